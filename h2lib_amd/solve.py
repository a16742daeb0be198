"""Krylov solves with device H2-matrices, and the interior Dirichlet problem
of the Laplace equation with the piecewise linear basis.

solve_gmres mirrors solve_gmres_avector (Library/krylovsolvers.c:183-227),
solve_cg solve_cg_avector (Library/krylovsolvers.c:18-52).
LaplaceDirichlet follows the reference's Dirichlet test
(Tests/test_laplacebem3d.c:100-160): f = L2 projection of the Dirichlet data
u(x) = 1/|x - x0|, b = (K + M/2) f with the double layer H2-matrix, V x = b
with the single layer H2-matrix, error against the Neumann data of u.
"""
import ctypes as C
import sys

import numpy as np

from ._lib import check, hip
from .bem3d import (BASIS_LINEAR_BEM3D, Surface3d, new_adlp_laplace_bem3d,
                    new_dlp_laplace_bem3d, new_slp_laplace_bem3d)
from .device import DeviceArray
from .geometry import BlockTree, ClusterTree, bem
from .h2matrix import H2Matrix

_done = False


def _sigs():
    global _done
    if _done:
        return
    h = hip()
    h.h2hip_projectL2_point_source.restype = C.c_int
    h.h2hip_projectL2_point_source.argtypes = [C.c_void_p, C.c_uint32,
                                               C.c_int, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
    h.h2hip_integrate_point_source.restype = C.c_int
    h.h2hip_integrate_point_source.argtypes = [C.c_void_p, C.c_uint32,
                                               C.c_int, C.c_void_p, C.c_int,
                                               C.c_void_p, C.c_void_p]
    h.h2hip_solve_gmres.restype = C.c_int
    h.h2hip_addeval_combined.restype = C.c_int
    h.h2hip_addeval_combined.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    h.h2hip_solve_cg.restype = C.c_int
    h.h2hip_solve_cg.argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_uint64, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_uint32, C.c_int,
                                 C.c_void_p, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_double)]
    h.h2hip_solve_gmres.argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_uint32, C.c_uint32,
                                    C.c_void_p, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_double)]
    b = bem()
    dp = C.c_void_p
    h.h2hip_gmres_batch_create.restype = C.c_int
    h.h2hip_gmres_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p,
                                           C.c_int, C.c_uint64, C.c_uint32,
                                           C.c_uint32, C.c_void_p]
    h.h2hip_gmres_batch_destroy.argtypes = [C.c_void_p]
    h.h2hip_gmres_batch_set_method.restype = C.c_int
    h.h2hip_gmres_batch_set_method.argtypes = [C.c_void_p, C.c_int]
    h.h2hip_gmres_batch_combine.restype = C.c_int
    h.h2hip_gmres_batch_combine.argtypes = [C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    h.h2hip_gmres_batch_free_slots.restype = C.c_int
    h.h2hip_gmres_batch_free_slots.argtypes = [C.c_void_p]
    h.h2hip_gmres_batch_submit.restype = C.c_int
    h.h2hip_gmres_batch_submit.argtypes = [C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_double,
                                           C.c_uint32, C.c_uint64]
    h.h2hip_gmres_batch_run.restype = C.c_int
    h.h2hip_gmres_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_uint32),
                                        C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_uint32]
    h.h2hip_gmres_batch_enqueue.restype = C.c_int
    h.h2hip_gmres_batch_enqueue.argtypes = [C.c_void_p]
    h.h2hip_gmres_batch_poll.restype = C.c_int
    h.h2hip_gmres_batch_poll.argtypes = [C.c_void_p, C.POINTER(C.c_uint32),
                                         C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint32]
    h.h2hip_gmres_batch_busy.restype = C.c_int
    h.h2hip_gmres_batch_busy.argtypes = [C.c_void_p]
    h.h2hip_gmres_multi_create.restype = C.c_int
    h.h2hip_gmres_multi_create.argtypes = [C.POINTER(C.c_void_p), C.c_uint32,
                                           dp, dp, C.c_int, C.c_uint32, dp]
    h.h2hip_gmres_multi_destroy.argtypes = [dp]
    h.h2hip_gmres_multi_solve.restype = C.c_int
    h.h2hip_gmres_multi_solve.argtypes = [dp, dp, dp, C.c_double, C.c_uint32,
                                          C.c_int, dp, dp]
    h.h2hip_gmres_multi_graph_builds.restype = C.c_uint64
    h.h2hip_gmres_multi_graph_builds.argtypes = [dp]
    b.h2b_integrate_l.argtypes = [dp, C.c_uint, C.c_int, dp, dp]
    b.h2b_projectL2_l.restype = C.c_int
    b.h2b_projectL2_l.argtypes = [dp, C.c_uint, C.c_int, dp, dp]
    b.h2b_normL2diff_l.restype = C.c_double
    b.h2b_normL2diff_l.argtypes = [dp, C.c_uint, dp, C.c_int, dp]
    b.h2b_normL2.restype = C.c_double
    b.h2b_normL2.argtypes = [dp, C.c_uint, C.c_int, dp]
    b.h2b_projectL2_c_helmholtz.argtypes = [dp, C.c_uint, C.c_int,
                                            C.c_double, dp, dp]
    b.h2b_integrate_c_helmholtz.argtypes = [dp, C.c_uint, C.c_int,
                                            C.c_double, dp, dp]
    b.h2b_integrate_l_helmholtz.argtypes = [dp, C.c_uint, C.c_int,
                                            C.c_double, dp, dp]
    h.h2hip_helmholtz_point_data.restype = C.c_int
    h.h2hip_helmholtz_point_data.argtypes = [dp, C.c_uint32, C.c_int, C.c_int,
                                             C.c_double, dp, C.c_int, dp, dp]
    h.h2hip_helmholtz_point_errors_c.restype = C.c_int
    h.h2hip_helmholtz_point_errors_c.argtypes = [dp, C.c_uint32, C.c_int,
                                                 C.c_double, dp, dp, dp, dp]
    h.h2hip_eval_brakhage_werner_l.restype = C.c_int
    h.h2hip_eval_brakhage_werner_l.argtypes = [dp, C.c_uint32, dp, dp, dp,
                                               C.c_uint32, dp, dp, dp]
    h.h2hip_combined_create.restype = C.c_int
    h.h2hip_combined_create.argtypes = [C.POINTER(C.c_void_p), dp, dp, dp,
                                        dp, dp, dp]
    h.h2hip_combined_destroy.argtypes = [dp]
    h.h2hip_eval_brakhage_werner_c.restype = C.c_int
    h.h2hip_eval_brakhage_werner_c.argtypes = [dp, C.c_uint32, dp, dp, dp,
                                               C.c_uint32, dp, dp, dp]
    b.h2b_normL2_c_helmholtz.argtypes = [dp, C.c_uint, dp, C.c_int,
                                         C.c_double, dp, dp]
    _done = True


def helmholtz_point_data(mesh, kind, k, x0, q=2):
    """Point-source data at x0 (kind 0: Dirichlet e^{ikr}/r, 1: Neumann)
    projected onto the piecewise constants: the mean over each triangle
    (projectL2_bem3d_c_avector, Library/bem3d.c:8827-8876)."""
    _sigs()
    out = np.zeros(mesh.triangles, dtype=np.complex128)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    bem().h2b_projectL2_c_helmholtz(mesh._sp, q, kind, float(k),
                                    x0.ctypes.data_as(C.c_void_p),
                                    out.ctypes.data_as(C.c_void_p))
    return out


def helmholtz_point_errors(mesh, x, kind, k, x0, q=2):
    """(|x - data|_L2, |data|_L2) for piecewise constant x
    (normL2diff_c_bem3d, normL2_bem3d)."""
    _sigs()
    out = np.zeros(2)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    xx = None if x is None else np.ascontiguousarray(x, dtype=np.complex128)
    bem().h2b_normL2_c_helmholtz(
        mesh._sp, q, None if xx is None else xx.ctypes.data_as(C.c_void_p),
        kind, float(k), x0.ctypes.data_as(C.c_void_p),
        out.ctypes.data_as(C.c_void_p))
    return out


PROJECT_C, INTEGRATE_C, INTEGRATE_L = 0, 1, 2


def helmholtz_point_data_device(gr, q, kind, mode, k, x0, d_out,
                                accumulate=False, stream=None):
    """Point-source data on the device (h2hip_helmholtz_point_data):
    mode PROJECT_C projection onto the constants, INTEGRATE_C integrals
    against the constants, INTEGRATE_L integrals against the linear basis,
    into the device array d_out (added with accumulate), ordered on the
    stream.  gr: the Surface3d of the operators."""
    _sigs()
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    check(hip().h2hip_helmholtz_point_data(
        gr.handle, q, kind, mode, float(k), x0.ctypes.data_as(C.c_void_p),
        int(bool(accumulate)), C.c_void_p(d_out.ptr),
        C.c_void_p(stream.handle if stream else 0)),
        "h2hip_helmholtz_point_data")


def helmholtz_point_errors_device(gr, q, kind, k, x0, d_x, stream=None):
    """(|x - data|_L2, |data|_L2) for the constant-basis device vector d_x
    (h2hip_helmholtz_point_errors_c)."""
    _sigs()
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    out = np.zeros(2)
    check(hip().h2hip_helmholtz_point_errors_c(
        gr.handle, q, kind, float(k), x0.ctypes.data_as(C.c_void_p),
        C.c_void_p(d_x.ptr), out.ctypes.data_as(C.c_void_p),
        C.c_void_p(stream.handle if stream else 0)),
        "h2hip_helmholtz_point_errors_c")
    return out


def _addeval_fn(H):
    """The device addeval of an operator object: an H2Matrix (self-adjoint
    product for symmetric storage) or a Combined."""
    if isinstance(H, Combined):
        return C.cast(hip().h2hip_addeval_combined, C.c_void_p)
    return C.cast(hip().h2hip_addevalsymm_h2matrix
                  if getattr(H, "symmetric", False)
                  else hip().h2hip_addeval_h2matrix, C.c_void_p)


class Combined:
    """c1 A1 + c2 A2 of two device H2-matrices as one operator (the
    Brakhage-Werner operator KM - i eta V of the reference's Helmholtz
    test, Tests/test_helmholtzbem3d.c:24-65)."""

    def __init__(self, A1, c1, A2=None, c2=0.0):
        _sigs()
        self.A1, self.A2 = A1, A2
        self.c1, self.c2 = c1, c2
        self.dtype = np.complex128 if (np.iscomplexobj(c1) or
                                       np.iscomplexobj(c2) or
                                       A1.dtype == np.complex128) \
            else np.float64
        self.n = A1.n
        a = np.array([np.real(c1), np.imag(c1)], dtype=np.float64)
        b = np.array([np.real(c2), np.imag(c2)], dtype=np.float64)
        h = C.c_void_p()
        check(hip().h2hip_combined_create(
            C.byref(h), _addeval_fn(A1), A1._h,
            a.ctypes.data_as(C.c_void_p),
            _addeval_fn(A2) if A2 is not None else None,
            A2._h if A2 is not None else None,
            b.ctypes.data_as(C.c_void_p)), "h2hip_combined_create")
        self._h = h

    def addeval(self, alpha, d_x, d_y, stream=None):
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        check(hip().h2hip_addeval_combined(
            a.ctypes.data_as(C.c_void_p), self._h, C.c_void_p(d_x.ptr),
            C.c_void_p(d_y.ptr), C.c_void_p(stream.handle if stream else 0)),
            "h2hip_addeval_combined")

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_combined_destroy(h)
            self._h = None


def solve_gmres(H, d_b, d_x, eps, maxiter, kmax, stream=None):
    """GMRES(kmax) for H x = b on device arrays; x holds the start vector.
    Returns (steps, residual estimate).  H: an H2Matrix (a matrix in
    symmetric storage is applied with the self-adjoint product) or a
    Combined operator."""
    _sigs()
    fn = _addeval_fn(H)
    it = C.c_uint32()
    res = C.c_double()
    check(hip().h2hip_solve_gmres(fn, H._h, int(H.dtype == np.complex128),
                                  d_b.n,
                                  C.c_void_p(d_b.ptr), C.c_void_p(d_x.ptr),
                                  eps, maxiter, kmax,
                                  C.c_void_p(stream.handle if stream else 0),
                                  C.byref(it), C.byref(res)),
          "h2hip_solve_gmres")
    return it.value, res.value


def solve_cg(H, d_b, d_x, eps, maxiter, use_graph=False, stream=None):
    """Conjugate gradients for H x = b (H self-adjoint positive definite)
    on device arrays; x holds the start vector.  Mirrors solve_cg_avector
    (Library/krylovsolvers.c:18-52).  use_graph replays 8 steps at a time as
    one captured graph.  Returns (steps, final |r|)."""
    _sigs()
    fn = _addeval_fn(H)
    it = C.c_uint32()
    res = C.c_double()
    check(hip().h2hip_solve_cg(fn, H._h, int(H.dtype == np.complex128),
                               d_b.n, C.c_void_p(d_b.ptr),
                               C.c_void_p(d_x.ptr), eps, maxiter,
                               int(bool(use_graph)),
                               C.c_void_p(stream.handle if stream else 0),
                               C.byref(it), C.byref(res)),
          "h2hip_solve_cg")
    return it.value, res.value


class GmresBatch:
    """Up to 8 independent GMRES(kmax) solves with one H2-matrix, advanced
    together so that every step reads the matrix once (continuous batching:
    solves enter free slots and leave when done, each with the semantics of
    solve_gmres); with method="cg" conjugate gradients instead, each solve
    with the semantics of solve_cg (for Hermitian positive definite H)."""

    def __init__(self, H, slots=8, kmax=30, stream=None, method="gmres"):
        """H: an H2Matrix, or a Combined of two H2-matrices (c1 A1 + c2 A2,
        two batched products per step)."""
        _sigs()
        self.H = H
        self.stream = stream
        A1 = H.A1 if isinstance(H, Combined) else H
        h = C.c_void_p()
        check(hip().h2hip_gmres_batch_create(
            C.byref(h), A1._h, int(H.dtype == np.complex128),
            len(A1.interp.ridx), slots, kmax,
            C.c_void_p(stream.handle if stream else 0)),
            "h2hip_gmres_batch_create")
        self._h = h
        self._keep = {}
        if isinstance(H, Combined):
            if H.A2 is None:
                raise ValueError("GmresBatch: a Combined needs two operators")
            c1 = np.array([np.real(H.c1), np.imag(H.c1)])
            c2 = np.array([np.real(H.c2), np.imag(H.c2)])
            check(hip().h2hip_gmres_batch_combine(
                h, c1.ctypes.data_as(C.c_void_p), H.A2._h,
                c2.ctypes.data_as(C.c_void_p)), "h2hip_gmres_batch_combine")
            self._coef = (c1, c2)
        self.method = "gmres"
        self.set_method(method)

    def set_method(self, method):
        """Switches between "gmres" and "cg" while no solve is active."""
        if method not in ("gmres", "cg"):
            raise ValueError("GmresBatch: method is 'gmres' or 'cg'")
        check(hip().h2hip_gmres_batch_set_method(self._h,
                                                  int(method == "cg")),
              "h2hip_gmres_batch_set_method")
        self.method = method

    def free_slots(self):
        return hip().h2hip_gmres_batch_free_slots(self._h)

    def submit(self, d_b, d_x, eps, maxiter, tag):
        """Starts a solve of H x = b (x holds the start vector); returns the
        slot."""
        s = hip().h2hip_gmres_batch_submit(self._h, C.c_void_p(d_b.ptr),
                                            C.c_void_p(d_x.ptr), eps,
                                            maxiter, tag)
        if s < 0:
            check(s, "h2hip_gmres_batch_submit")
        self._keep[tag] = (d_b, d_x)
        return s

    def _finished(self, fn):
        cap = 8
        nd = C.c_uint32()
        tags = np.zeros(cap, dtype=np.uint64)
        its = np.zeros(cap, dtype=np.uint32)
        res = np.zeros(cap)
        rc = fn(self._h, C.byref(nd), tags.ctypes.data_as(C.c_void_p),
                its.ctypes.data_as(C.c_void_p),
                res.ctypes.data_as(C.c_void_p), cap)
        if rc < 0:
            check(rc, "h2hip_gmres_batch")
        out = []
        for i in range(nd.value):
            out.append((int(tags[i]), int(its[i]), float(res[i])))
            self._keep.pop(int(tags[i]), None)
        return rc, out

    def run(self):
        """Advances all solves by one chunk of steps; returns the solves
        that finished as (tag, steps, residual)."""
        return self._finished(hip().h2hip_gmres_batch_run)[1]

    def enqueue(self):
        """Launches the next chunk without waiting for it."""
        check(hip().h2hip_gmres_batch_enqueue(self._h),
              "h2hip_gmres_batch_enqueue")

    def poll(self):
        """None while the chunk in flight runs; otherwise the solves it
        finished as (tag, steps, residual)."""
        rc, out = self._finished(hip().h2hip_gmres_batch_poll)
        return None if rc == 1 else out

    @property
    def active(self):
        return bool(hip().h2hip_gmres_batch_busy(self._h) & 1)

    @property
    def inflight(self):
        return bool(hip().h2hip_gmres_batch_busy(self._h) & 2)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_gmres_batch_destroy(h)
            self._h = None


class GmresMulti:
    """GMRES(kmax) on many independent systems A_s x_s = b_s (one device
    H2-matrix each), stepping in lock step; with use_graph one Arnoldi step
    of all active systems is one replay of a captured HIP graph.  Each
    system follows solve_gmres_avector (Library/krylovsolvers.c:183-227)."""

    def __init__(self, mats, kmax=30, stream=None):
        _sigs()
        if stream is None:
            raise ValueError("GmresMulti needs a stream (it is captured)")
        self.mats = list(mats)
        self.dtype = self.mats[0].dtype
        self.stream = stream
        hs = (C.c_void_p * len(self.mats))(*[H._h.value for H in self.mats])
        ns = np.array([H.n for H in self.mats], dtype=np.uint64)
        h = C.c_void_p()
        check(hip().h2hip_gmres_multi_create(
            C.byref(h), len(self.mats), C.cast(hs, C.c_void_p),
            ns.ctypes.data_as(C.c_void_p),
            int(self.dtype == np.complex128), kmax, stream.handle),
            "h2hip_gmres_multi_create")
        self._h = h

    def solve(self, d_bs, d_xs, eps, maxiter, use_graph=True):
        """Solves all systems (x_s: start vectors, overwritten); returns
        (iterations, residual estimates) per system."""
        c = len(self.mats)
        bs = (C.c_void_p * c)(*[d.ptr for d in d_bs])
        xs = (C.c_void_p * c)(*[d.ptr for d in d_xs])
        it = np.zeros(c, dtype=np.uint32)
        res = np.zeros(c)
        check(hip().h2hip_gmres_multi_solve(
            self._h, C.cast(bs, C.c_void_p), C.cast(xs, C.c_void_p), eps,
            maxiter, int(bool(use_graph)), it.ctypes.data_as(C.c_void_p),
            res.ctypes.data_as(C.c_void_p)), "h2hip_gmres_multi_solve")
        return it, res

    @property
    def graph_builds(self):
        return int(hip().h2hip_gmres_multi_graph_builds(self._h))

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_gmres_multi_destroy(h)
            self._h = None


class LaplaceDirichlet:
    """Interior Dirichlet problem on a mesh (piecewise linear basis)."""

    def __init__(self, mesh, clf=64, eta=2.0, m=4, q=2, q2=4, schedule=2):
        """schedule: transforms of the H2 products (H2Matrix.set_schedule;
        2, the flat transforms, is the fastest and holds the ancestors'
        bases restricted to the leaves on the device)."""
        _sigs()
        self.mesh = mesh
        self.q = q
        self.gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
        lin = BASIS_LINEAR_BEM3D
        self.slp = new_slp_laplace_bem3d(self.gr, q, q2, lin, lin)
        self.dlp = new_dlp_laplace_bem3d(self.gr, q, q2, lin, lin, 0.5)
        self.tree = ClusterTree(mesh, lin, clf)
        self.blocks = BlockTree(self.tree, self.tree, eta, 0)
        self.V = H2Matrix(self.slp, mesh, self.tree, self.tree, self.blocks,
                          m)
        self.KM = H2Matrix(self.dlp, mesh, self.tree, self.tree, self.blocks,
                           m, nearplan=self.V.nearplan)
        for H in (self.V, self.KM):
            H.set_schedule(schedule)
        self.n = mesh.vertices

    def project(self, x0, kind=0):
        """L2 projection of the Dirichlet (kind 0) or Neumann (kind 1) data
        of a point source at x0 (host array)."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        out = np.zeros(self.n)
        bem().h2b_projectL2_l(self.mesh._sp, self.q, kind,
                              x0.ctypes.data_as(C.c_void_p),
                              out.ctypes.data_as(C.c_void_p))
        return out

    def project_device(self, x0, d_out, kind=0, stream=None):
        """The same projection on the device (h2hip_projectL2_point_source),
        into the device array d_out, ordered on the stream."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        check(hip().h2hip_projectL2_point_source(
            self.gr.handle, self.q, kind, x0.ctypes.data_as(C.c_void_p),
            C.c_void_p(d_out.ptr), C.c_void_p(stream.handle if stream else 0)),
            "h2hip_projectL2_point_source")

    def rhs(self, x0, stream=None):
        """Device vectors f (projected Dirichlet data) and b = (K + M/2) f,
        computed on the device."""
        d_f = DeviceArray(self.n)
        self.project_device(x0, d_f, 0, stream)
        d_b = DeviceArray(self.n)
        d_b.zero(stream)
        self.KM.addeval(1.0, d_f, d_b, stream)
        return d_f, d_b

    def rhs_sources(self, x0s, stream=None):
        """(f, b) with b = (K + M/2) f for the data of several unit point
        sources, f = the projected Dirichlet data of the last one and b the
        sum over all (the problem is linear in the data).  Keep f alive
        until the stream has passed the product."""
        d_f = DeviceArray(self.n)
        d_b = DeviceArray(self.n)
        d_b.zero(stream)
        for x0 in x0s:
            self.project_device(x0, d_f, 0, stream)
            self.KM.addeval(1.0, d_f, d_b, stream)
        return d_f, d_b

    def solve(self, d_b, eps=1e-8, maxiter=1000, kmax=30, stream=None,
              d_x=None, solver="gmres", use_graph=False):
        """V x = b from x = d_x (zero when None) by GMRES(kmax) or, with
        solver="cg", by conjugate gradients (V is symmetric positive
        definite; the reference's Laplace test solves with CG,
        Tests/test_laplacebem3d.c:16-46)."""
        if d_x is None:
            d_x = DeviceArray(self.n)
            d_x.zero(stream)
        if solver == "cg":
            it, res = solve_cg(self.V, d_b, d_x, eps, maxiter, use_graph,
                               stream)
        elif solver == "gmres":
            it, res = solve_gmres(self.V, d_b, d_x, eps, maxiter, kmax,
                                  stream)
        else:
            raise ValueError("solver: 'gmres' or 'cg'")
        return d_x, it, res

    def neumann_error(self, x, x0):
        """Relative L2 error of the Neumann data of a point source at x0."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        p0 = x0.ctypes.data_as(C.c_void_p)
        e = bem().h2b_normL2diff_l(self.mesh._sp, self.q,
                                   x.ctypes.data_as(C.c_void_p), 1, p0)
        return e / bem().h2b_normL2(self.mesh._sp, self.q, 1, p0)


def laplace_point_integrals(mesh, kind, x0, q=2):
    """Integrals of the Laplace point-source data at x0 (kind 0 Dirichlet
    1/|x - x0|, 1 Neumann) against the linear basis on the host
    (integrate_bem3d_l_avector, Library/bem3d.c:8631-8702)."""
    _sigs()
    out = np.zeros(mesh.vertices)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    bem().h2b_integrate_l(mesh._sp, q, kind, x0.ctypes.data_as(C.c_void_p),
                          out.ctypes.data_as(C.c_void_p))
    return out


def laplace_point_integrals_device(gr, q, kind, x0, d_out, accumulate=False,
                                   stream=None):
    """The same integrals on the device into d_out (added with accumulate),
    ordered on the stream (h2hip_integrate_point_source)."""
    _sigs()
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    check(hip().h2hip_integrate_point_source(
        gr.handle, q, kind, x0.ctypes.data_as(C.c_void_p),
        int(bool(accumulate)), C.c_void_p(d_out.ptr),
        C.c_void_p(stream.handle if stream else 0)),
        "h2hip_integrate_point_source")


class LaplaceNeumannExterior:
    """Exterior Neumann problem of the Laplace equation, piecewise linear
    basis, indirect formulation with the adjoint double layer: the field is
    the single layer potential u = V sigma of a density sigma, whose
    exterior normal derivative is (K' - 1/2) sigma, so sigma solves

        (K' - M/2) sigma = b,   b_i = int phi_i g

    for the Neumann data g (Galerkin; K' the adjoint double layer operator,
    new_adlp_laplace_bem3d with alpha = -1/2, Library/laplacebem3d.c:
    1180-1346).  The operator is not symmetric: GMRES.  The data are those
    of unit point sources x0 inside the surface, u(x) = sum 1/|x - x0|
    (harmonic outside, decaying), g = du/dn.  The Dirichlet trace of the
    solution is V sigma; dirichlet_error compares its integrals against the
    basis with those of u."""

    def __init__(self, mesh, clf=64, eta=2.0, m=4, q=2, q2=4, schedule=2,
                 stream=None):
        _sigs()
        self.mesh = mesh
        self.q = q
        self.gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
        lin = BASIS_LINEAR_BEM3D
        self.adlp = new_adlp_laplace_bem3d(self.gr, q, q2, lin, lin, -0.5)
        self.slp = new_slp_laplace_bem3d(self.gr, q, q2, lin, lin)
        self.tree = ClusterTree(mesh, lin, clf)
        self.blocks = BlockTree(self.tree, self.tree, eta, 0)
        self.A = H2Matrix(self.adlp, mesh, self.tree, self.tree, self.blocks,
                          m, stream=stream)
        self.V = H2Matrix(self.slp, mesh, self.tree, self.tree, self.blocks,
                          m, stream=stream, nearplan=self.A.nearplan)
        if stream is not None:
            stream.sync()
        for H in (self.A, self.V):
            H.set_schedule(schedule)
        self.n = mesh.vertices

    def rhs_sources(self, x0s, stream=None):
        """(b, b): b = the integrals of the Neumann data of unit point
        sources at x0s against the basis, formed on the device (the pair
        mirrors the other problems' (f, b))."""
        d_b = DeviceArray(self.n)
        for j, x0 in enumerate(x0s):
            laplace_point_integrals_device(self.gr, self.q, 1, x0, d_b,
                                           accumulate=j > 0, stream=stream)
        return d_b, d_b

    def rhs(self, x0, stream=None):
        return self.rhs_sources([x0], stream)[1]

    def solve(self, d_b, eps=1e-10, maxiter=1000, kmax=30, stream=None):
        d_x = DeviceArray(self.n)
        d_x.zero(stream)
        it, res = solve_gmres(self.A, d_b, d_x, eps, maxiter, kmax, stream)
        return d_x, it, res

    def trace_integrals(self, d_sigma, stream=None):
        """Integrals of the Dirichlet trace V sigma against the basis (the
        Galerkin product V sigma), on the host."""
        d_y = DeviceArray(self.n)
        d_y.zero(stream)
        self.V.addeval(1.0, d_sigma, d_y, stream)
        if stream is not None:
            stream.sync()
        return d_y.download()

    def dirichlet_error(self, d_sigma, x0s, stream=None):
        """|V sigma - D| / |D| with D_i = int phi_i u, u the field of the
        sources: the discretisation error of the Dirichlet trace, measured
        on its Galerkin moments."""
        got = self.trace_integrals(d_sigma, stream)
        want = sum(laplace_point_integrals(self.mesh, 0, x0, self.q)
                   for x0 in x0s)
        return float(np.linalg.norm(got - want) / np.linalg.norm(want))


class HelmholtzExterior:
    """Exterior Dirichlet problem of the Helmholtz equation with the data of
    a point source x0 inside the surface, piecewise constant basis: f = the
    projected Dirichlet data e^{ik|x - x0|}/|x - x0|, b = (K - M/2) f with
    the double layer H2-matrix (alpha = -1/2), V x = b with the single layer
    H2-matrix; x approximates the Neumann data of the radiating field.
    precision 1 assembles both near fields in mixed precision."""

    def __init__(self, mesh, k, clf=64, eta=2.0, m=4, q=2, q2=4,
                 precision=0, stream=None, schedule=2):
        from .bem3d import new_dlp_helmholtz_bem3d, new_slp_helmholtz_bem3d
        _sigs()
        self.mesh, self.k, self.q = mesh, k, q
        self.gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
        self.slp = new_slp_helmholtz_bem3d(k, self.gr, q, q2, 0, 0)
        self.dlp = new_dlp_helmholtz_bem3d(k, self.gr, q, q2, 0, 0, -0.5)
        if precision:
            self.slp.set_precision(precision)
            self.dlp.set_precision(precision)
        self.rt = ClusterTree(mesh, 0, clf)
        self.bt = BlockTree(self.rt, self.rt, eta, 0)
        self.V = H2Matrix(self.slp, mesh, self.rt, self.rt, self.bt, m,
                          stream=stream)
        self.K = H2Matrix(self.dlp, mesh, self.rt, self.rt, self.bt, m,
                          stream=stream, nearplan=self.V.nearplan)
        if stream is not None:
            stream.sync()
        for H in (self.V, self.K):
            H.set_schedule(schedule)
        self.n = mesh.triangles

    def rhs(self, x0, stream=None):
        """(f, b) on the device for the point source at x0, formed on the
        device and ordered on the stream."""
        return self.rhs_sources([x0], stream)

    def rhs_sources(self, x0s, stream=None):
        """(f, b) with b = (K - M/2) f for several unit point sources (f =
        sum of their projected Dirichlet data), formed on the device and
        ordered on the stream.  Keep f alive until the stream has passed
        the product."""
        d_f = DeviceArray(self.n, np.complex128)
        for j, x0 in enumerate(x0s):
            helmholtz_point_data_device(self.gr, self.q, 0, PROJECT_C, self.k,
                                        x0, d_f, accumulate=j > 0,
                                        stream=stream)
        d_b = DeviceArray(self.n, np.complex128)
        d_b.zero(stream)
        self.K.addeval(1.0, d_f, d_b, stream)
        return d_f, d_b

    def solve(self, d_b, eps=1e-8, maxiter=1000, kmax=30, stream=None):
        d_x = DeviceArray(self.n, np.complex128)
        d_x.zero(stream)
        it, res = solve_gmres(self.V, d_b, d_x, eps, maxiter, kmax, stream)
        return d_x, it, res

    def neumann_error(self, x, x0):
        """|x - Neumann data|_L2 / |Neumann data|_L2 for x on the host or
        the device."""
        if isinstance(x, DeviceArray):
            d, nrm = helmholtz_point_errors_device(self.gr, self.q, 1, self.k,
                                                   x0, x)
        else:
            d, nrm = helmholtz_point_errors(self.mesh, x, 1, self.k, x0,
                                            self.q)
        return d / nrm


def helmholtz_point_integrals(mesh, kind, k, x0, q=2, basis=0):
    """Integrals of the point-source data against the constant (basis 0,
    integrate_bem3d_c_avector, Library/bem3d.c:8579-8629) or the linear
    basis (basis 1, integrate_bem3d_l_avector, Library/bem3d.c:8631-8702)."""
    _sigs()
    out = np.zeros(mesh.vertices if basis else mesh.triangles,
                   dtype=np.complex128)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    fn = bem().h2b_integrate_l_helmholtz if basis else \
        bem().h2b_integrate_c_helmholtz
    fn(mesh._sp, q, kind, float(k), x0.ctypes.data_as(C.c_void_p),
       out.ctypes.data_as(C.c_void_p))
    return out


class HelmholtzBrakhageWerner:
    """The reference's Helmholtz test problem (Tests/test_helmholtzbem3d.c:
    353-450, exterior): b = integrals of the Dirichlet data of a point
    source inside, (KM - i eta V) x = b with KM the double layer operator
    with alpha = 1/2, eta = k; the solution's exterior potential
    (eval_brakhage_werner_c / _l) is compared with the point-source field.
    basis 0 piecewise constant, 1 piecewise linear (both operators).
    precision 1 assembles both near fields in mixed precision."""

    def __init__(self, mesh, k, clf=32, eta_adm=1.0, m=5, q=2, q2=4,
                 precision=0, stream=None, basis=0, schedule=2):
        from .bem3d import new_dlp_helmholtz_bem3d, new_slp_helmholtz_bem3d
        _sigs()
        self.mesh, self.k, self.q, self.basis = mesh, k, q, basis
        self.eta = k
        self.gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
        self.slp = new_slp_helmholtz_bem3d(k, self.gr, q, q2, basis, basis)
        self.dlp = new_dlp_helmholtz_bem3d(k, self.gr, q, q2, basis, basis,
                                           0.5)
        if precision:
            self.slp.set_precision(precision)
            self.dlp.set_precision(precision)
        self.rt = ClusterTree(mesh, basis, clf)
        self.bt = BlockTree(self.rt, self.rt, eta_adm, 0)
        self.V = H2Matrix(self.slp, mesh, self.rt, self.rt, self.bt, m,
                          stream=stream)
        self.KM = H2Matrix(self.dlp, mesh, self.rt, self.rt, self.bt, m,
                           stream=stream, nearplan=self.V.nearplan)
        if stream is not None:
            stream.sync()
        for H in (self.V, self.KM):
            H.set_schedule(schedule)
        self.A = Combined(self.KM, 1.0, self.V, -1j * self.eta)
        self.n = mesh.vertices if basis else mesh.triangles

    def rhs(self, x0, stream=None):
        """b = integrals of the Dirichlet data against the basis, formed on
        the device."""
        d_b = DeviceArray(self.n, np.complex128)
        helmholtz_point_data_device(self.gr, self.q, 0,
                                    INTEGRATE_L if self.basis else
                                    INTEGRATE_C, self.k, x0, d_b,
                                    stream=stream)
        return d_b

    def rhs_sources(self, x0s, stream=None):
        """(b, b) for the sum of the data of several unit point sources
        (the pair mirrors the other problems' (f, b))."""
        d_b = DeviceArray(self.n, np.complex128)
        for j, x0 in enumerate(x0s):
            helmholtz_point_data_device(self.gr, self.q, 0,
                                        INTEGRATE_L if self.basis else
                                        INTEGRATE_C, self.k, x0, d_b,
                                        accumulate=j > 0, stream=stream)
        return d_b, d_b

    def solve(self, d_b, eps=1e-12, maxiter=1000, kmax=50, stream=None):
        d_x = DeviceArray(self.n, np.complex128)
        d_x.zero(stream)
        it, res = solve_gmres(self.A, d_b, d_x, eps, maxiter, kmax, stream)
        return d_x, it, res

    def potential(self, d_w, pts, stream=None):
        """The Brakhage-Werner potential of the density at the points."""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pts), dtype=np.complex128)
        k = np.array([np.real(self.k), np.imag(self.k)], dtype=np.float64)
        e = np.array([np.real(self.eta), np.imag(self.eta)], dtype=np.float64)
        fn = hip().h2hip_eval_brakhage_werner_l if self.basis else \
            hip().h2hip_eval_brakhage_werner_c
        check(fn(self.gr.handle, self.q, k.ctypes.data_as(C.c_void_p),
                 e.ctypes.data_as(C.c_void_p), C.c_void_p(d_w.ptr), len(pts),
                 pts.ctypes.data_as(C.c_void_p),
                 out.ctypes.data_as(C.c_void_p),
                 C.c_void_p(stream.handle if stream else 0)),
              "h2hip_eval_brakhage_werner")
        return out

    @staticmethod
    def outer_grid():
        """The test's evaluation points: a 20 x 20 grid of [-10, 10]^2 in
        the xz-plane outside the unit disc (max_rel_outer_error,
        Tests/test_helmholtzbem3d.c:224-287)."""
        pts = []
        for j in range(20):
            for i in range(20):
                x = -10.0 + (20.0 / 19) * i
                z = -10.0 + (20.0 / 19) * j
                if x * x + z * z > 1:
                    pts.append((x, 0.0, z))
        return np.array(pts)

    def outer_error(self, d_w, x0):
        """The test's error: max relative error of the potential against
        the point-source field e^{ik|x - x0|}/|x - x0| on the outer grid."""
        pts = self.outer_grid()
        u = self.potential(d_w, pts)
        r = np.linalg.norm(pts - np.asarray(x0)[None, :], axis=1)
        g = np.exp(1j * self.k * r) / r
        return float(np.max(np.abs(u - g) / np.abs(g)))

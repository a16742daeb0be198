"""ctypes access to the test oracles.

* ``Oracle``   -- the CPU restatement (oracle/liboracle.so).
* ``Reference`` -- the reference library compiled from its sources together
  with oracle/ref_driver.c (oracle/_ref/libh2ref_{real,complex}.so).

Both are checkers only: the product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = {False: os.path.join(ROOT, "oracle", "_ref", "libh2ref_real.so"),
          True: os.path.join(ROOT, "oracle", "_ref", "libh2ref_complex.so")}
# the reference plus the H2Lib-side binding of the GPU library
REF_HIP_SO = {
    False: os.path.join(ROOT, "oracle", "_ref", "libh2ref_hip_real.so"),
    True: os.path.join(ROOT, "oracle", "_ref", "libh2ref_hip_complex.so")}
# the reference built for speed (-O3, AVX2/FMA), timed as the CPU baseline
REF_FAST_SO = os.path.join(ROOT, "oracle", "_ref", "libh2ref_fast_real.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _idx(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.uint32)


def have_reference(complex_=False, hip=False, fast=False):
    if fast:
        return not complex_ and not hip and os.path.exists(REF_FAST_SO)
    return os.path.exists((REF_HIP_SO if hip else REF_SO)[complex_])


class Oracle:
    """The CPU restatement of the reference loops."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = C.CDLL(ORACLE_SO)
            lib.orc_bem_new.restype = C.c_void_p
            lib.orc_bem_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_uint, C.c_void_p, C.c_uint,
                                        C.c_void_p, C.c_uint, C.c_uint,
                                        C.c_void_p, C.c_void_p]
            lib.orc_bem_del.argtypes = [C.c_void_p]
            lib.orc_nearfield.argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_uint, C.c_void_p, C.c_uint,
                                          C.c_int, C.c_void_p, C.c_size_t]
            lib.orc_nearfield_evals.restype = C.c_double
            lib.orc_nearfield_evals.argtypes = [C.c_void_p, C.c_int,
                                                C.c_void_p, C.c_uint,
                                                C.c_void_p, C.c_uint]
            lib.orc_sq_new.restype = C.c_void_p
            lib.orc_sq_new.argtypes = [C.c_uint, C.c_uint]
            lib.orc_sq_del.argtypes = [C.c_void_p]
            lib.orc_sq_export.argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_uint),
                                          C.POINTER(C.c_double)]
            lib.orc_gauss.argtypes = [C.c_uint, C.c_void_p, C.c_void_p]
            lib.orc_bem_gram.restype = C.c_void_p
            lib.orc_bem_gram.argtypes = [C.c_void_p]
            lib.orc_bem_normals.restype = C.c_void_p
            lib.orc_bem_normals.argtypes = [C.c_void_p]
            cls._lib = lib
        return cls._lib

    @classmethod
    def gauss(cls, m):
        x = np.zeros(m)
        w = np.zeros(m)
        cls.lib().orc_gauss(m, _ptr(x), _ptr(w))
        return x, w

    @classmethod
    def rule(cls, q, q2, cls_id):
        lib = cls.lib()
        sq = lib.orc_sq_new(q, q2)
        n = C.c_uint()
        base = C.c_double()
        lib.orc_sq_export(sq, cls_id, None, None, None, C.byref(n),
                          C.byref(base))
        xc = 1 if cls_id == 4 else 2
        x = np.zeros(xc * n.value)
        y = np.zeros(xc * n.value)
        w = np.zeros(n.value)
        lib.orc_sq_export(sq, cls_id, _ptr(x), _ptr(y), _ptr(w), C.byref(n),
                          C.byref(base))
        lib.orc_sq_del(sq)
        return x, y, w, base.value

    @classmethod
    def h2_addeval(cls, interp, N, x, y, alpha=1.0, trans=False, S=None,
                   symm=False):
        """y += alpha H x (H^* x) for the host arrays of an
        h2lib_amd.h2matrix.Interpolation and packed near-field blocks N
        (in the order of interp.near); symm: the self-adjoint product from
        the lower block triangle (orc_h2_addevalsymm)."""
        lib = cls.lib()
        if not hasattr(lib, "_h2sig"):
            lib.orc_h2_addeval.restype = C.c_int
            lib.orc_h2_addeval.argtypes = [C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_uint32,
                                           C.c_void_p, C.c_void_p,
                                           C.c_uint64, C.c_uint32, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
            lib._h2sig = True
        I = interp
        cplx = np.iscomplexobj(N) or np.iscomplexobj(x) or \
            np.iscomplexobj(alpha) or np.iscomplexobj(S)
        dt = np.complex128 if cplx else np.float64
        N = np.ascontiguousarray(N, dtype=dt)
        xv = np.ascontiguousarray(x, dtype=dt)
        if y.dtype != dt or not y.flags.c_contiguous:
            raise TypeError("y must be a contiguous %s array" % dt)
        S = np.ascontiguousarray(I.S if S is None else S)
        S = S.astype(np.complex128 if cplx else np.float64)
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)

        class CB(C.Structure):
            _fields_ = [("nnodes", C.c_uint32), ("tree", C.c_void_p),
                        ("rank", C.c_void_p), ("idx", C.c_void_p),
                        ("V", C.c_void_p), ("voff", C.c_void_p),
                        ("E", C.c_void_p), ("eld", C.c_uint64)]

        rb = CB(len(I.rtree), _ptr(I.rtree), _ptr(I.rrank), _ptr(I.ridx),
                _ptr(I.rV), _ptr(I.rvoff), _ptr(I.rE), I.eld)
        cb = CB(len(I.ctree), _ptr(I.ctree), _ptr(I.crank), _ptr(I.cidx),
                _ptr(I.cV), _ptr(I.cvoff), _ptr(I.cE), I.eld)
        if symm:
            if not hasattr(lib, "_h2ssig"):
                lib.orc_h2_addevalsymm.restype = C.c_int
                lib.orc_h2_addevalsymm.argtypes = [
                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                    C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
                lib._h2ssig = True
            rc = lib.orc_h2_addevalsymm(int(cplx), _ptr(a), C.byref(rb),
                                        C.byref(cb), len(I.far), _ptr(I.far),
                                        _ptr(S), I.eld, len(I.near),
                                        _ptr(I.near), _ptr(N), _ptr(xv),
                                        _ptr(y))
        else:
            rc = lib.orc_h2_addeval(int(cplx), int(trans), _ptr(a),
                                    C.byref(rb), C.byref(cb), len(I.far),
                                    _ptr(I.far), _ptr(S), I.eld, len(I.near),
                                    _ptr(I.near), _ptr(N), _ptr(xv), _ptr(y))
        if rc != 0:
            raise ValueError("orc_h2_addeval rejected its input")
        return y


ADDEVAL_T = C.CFUNCTYPE(None, C.c_double, C.c_void_p,
                        C.POINTER(C.c_double), C.POINTER(C.c_double))


def oracle_gmres(matvec, n, b, x0, eps, maxiter, kmax):
    """The restated GMRES with y += alpha A x computed by the Python
    callable matvec(x) -> A x (float64 arrays); returns x, steps, residual."""
    lib = Oracle.lib()
    if not hasattr(lib, "_gmsig"):
        lib.orc_solve_gmres.restype = C.c_uint
        lib.orc_solve_gmres.argtypes = [ADDEVAL_T, C.c_void_p, C.c_size_t,
                                        C.c_void_p, C.c_void_p, C.c_double,
                                        C.c_uint, C.c_uint,
                                        C.POINTER(C.c_double)]
        lib._gmsig = True

    def cb(alpha, _A, xp, yp):
        xv = np.ctypeslib.as_array(xp, shape=(n,))
        yv = np.ctypeslib.as_array(yp, shape=(n,))
        yv += alpha * matvec(xv.copy())

    f = ADDEVAL_T(cb)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    xx = np.ascontiguousarray(x0, dtype=np.float64).copy()
    res = C.c_double()
    it = lib.orc_solve_gmres(f, None, n, _ptr(bb), _ptr(xx), eps, maxiter,
                             kmax, C.byref(res))
    return xx, it, res.value


ADDEVAL_C_T = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_void_p,
                          C.POINTER(C.c_double), C.POINTER(C.c_double))


def oracle_cg(matvec, n, b, x0, eps, maxiter, cplx=False):
    """The restated CG with y += alpha A x computed by the Python callable
    matvec(x) -> A x (float64 or complex128 arrays); returns x, steps,
    final |r|."""
    lib = Oracle.lib()
    if not hasattr(lib, "_cgsig"):
        lib.orc_solve_cg.restype = C.c_uint
        lib.orc_solve_cg.argtypes = [ADDEVAL_C_T, C.c_void_p, C.c_int,
                                     C.c_size_t, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_uint,
                                     C.POINTER(C.c_double)]
        lib._cgsig = True
    dt = np.complex128 if cplx else np.float64
    m = (2 if cplx else 1) * n

    def cb(alpha, _A, xp, yp):
        xv = np.ctypeslib.as_array(xp, shape=(m,)).view(dt)
        yv = np.ctypeslib.as_array(yp, shape=(m,)).view(dt)
        a = complex(alpha[0], alpha[1]) if cplx else alpha[0]
        yv += a * matvec(xv.copy())

    f = ADDEVAL_C_T(cb)
    bb = np.ascontiguousarray(b, dtype=dt)
    xx = np.ascontiguousarray(x0, dtype=dt).copy()
    res = C.c_double()
    it = lib.orc_solve_cg(f, None, int(cplx), n, _ptr(bb), _ptr(xx), eps,
                          maxiter, C.byref(res))
    return xx, it, res.value


def oracle_nearfield_packed(ob, blocks, ridx_all, cidx_all):
    """Near-field blocks (roff, rows, coff, cols) of an OracleBem computed
    one by one and packed column-major one after the other."""
    out = []
    for ro, nr, co, nc in np.asarray(blocks, dtype=np.int64):
        N = ob.nearfield(ridx_all[ro:ro + nr], cidx_all[co:co + nc])
        out.append(np.asarray(N).reshape(nr, nc).ravel(order="F"))
    return np.concatenate(out) if out else np.zeros(0)


class _BemBase:
    """Common interface of the restatement and the reference operators."""

    def __init__(self, kernel, op, basis_r, basis_c, x, t, q=2, q2=4,
                 alpha=0.0, k=0.0):
        self.kernel, self.op = kernel, op
        self.basis_r, self.basis_c = basis_r, basis_c
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.t = np.ascontiguousarray(t, dtype=np.uint32).reshape(-1, 3)
        self.complex = kernel == 1
        self.dtype = np.complex128 if self.complex else np.float64
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        kk = np.array([np.real(k), np.imag(k)], dtype=np.float64)
        self._keep = (a, kk)
        self.h = self._new(kernel, op, basis_r, basis_c, len(self.x),
                           _ptr(self.x), len(self.t), _ptr(self.t), q, q2,
                           _ptr(a), _ptr(kk))
        if not self.h:
            raise RuntimeError("operator construction failed")

    def ndof(self, basis):
        return len(self.x) if basis else len(self.t)

    def nearfield(self, ridx=None, cidx=None, far=False, ntrans=False,
                  rows=None, cols=None):
        """Dense block from the nearfield callback (column-major -> ndarray
        of shape (rows, cols), or (cols, rows) with ntrans)."""
        ridx = _idx(ridx)
        cidx = _idx(cidx)
        rows = len(ridx) if ridx is not None else (
            rows or self.ndof(self.basis_r))
        cols = len(cidx) if cidx is not None else (
            cols or self.ndof(self.basis_c))
        shp = (cols, rows) if ntrans else (rows, cols)
        # The reference adds the cc identity term at (t, t), which can lie
        # outside the window when rows > cols: allocate max(rows, cols)
        # columns so that such writes stay inside the buffer, and return the
        # window only.
        ncol = max(rows, cols)
        N = np.full(shp[0] * ncol, np.nan, dtype=self.dtype)
        self._nearfield(self.h, int(far), _ptr(ridx), rows, _ptr(cidx), cols,
                        int(ntrans), _ptr(N), shp[0])
        return N.reshape(ncol, shp[0]).T[:, :shp[1]].copy()


class OracleBem(_BemBase):
    def __init__(self, *a, **kw):
        lib = Oracle.lib()
        self._lib = lib
        self._new = lib.orc_bem_new
        self._nearfield = lib.orc_nearfield
        super().__init__(*a, **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.orc_bem_del(self.h)
            self.h = None

    def evals(self, ridx=None, cidx=None, far=False):
        ridx = _idx(ridx)
        cidx = _idx(cidx)
        rows = len(ridx) if ridx is not None else self.ndof(self.basis_r)
        cols = len(cidx) if cidx is not None else self.ndof(self.basis_c)
        return self._lib.orc_nearfield_evals(self.h, int(far), _ptr(ridx),
                                             rows, _ptr(cidx), cols)

    def gram(self):
        p = self._lib.orc_bem_gram(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)),
                                     shape=(len(self.t),)).copy()


class Reference:
    """The reference library built from its sources (real or complex)."""

    _libs = {}

    @classmethod
    def lib(cls, complex_=False, hip=False, fast=False):
        key = (bool(complex_), bool(hip), bool(fast))
        if key not in cls._libs:
            if fast and (complex_ or hip):
                raise ValueError("the fast reference build is real only")
            path = REF_FAST_SO if fast else \
                (REF_HIP_SO if hip else REF_SO)[complex_]
            lib = C.CDLL(path, mode=C.RTLD_LOCAL)
            lib.ref_mesh_size.argtypes = [C.c_int, C.c_uint,
                                          C.POINTER(C.c_uint),
                                          C.POINTER(C.c_uint)]
            lib.ref_mesh_fill.argtypes = [C.c_int, C.c_uint, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p]
            lib.ref_singquad.argtypes = [C.c_uint, C.c_uint, C.c_int,
                                         C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.POINTER(C.c_uint),
                                         C.POINTER(C.c_double)]
            lib.ref_bem_new.restype = C.c_void_p
            lib.ref_bem_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_uint, C.c_void_p, C.c_uint,
                                        C.c_void_p, C.c_uint, C.c_uint,
                                        C.c_void_p, C.c_void_p]
            lib.ref_bem_del.argtypes = [C.c_void_p]
            if not complex_:
                lib.ref_laplace_point_integrate_l.argtypes = [
                    C.c_void_p, C.c_int, C.c_void_p]
            if complex_:
                lib.ref_helmholtz_point_integrate.argtypes = [
                    C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
                lib.ref_helmholtz_bw_solve.restype = C.c_uint
                lib.ref_helmholtz_bw_solve.argtypes = [
                    C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                    C.c_void_p, C.c_double, C.c_uint, C.c_uint]
                lib.ref_eval_bw_c.argtypes = [C.c_void_p, C.c_void_p,
                                              C.c_double, C.c_void_p,
                                              C.c_uint, C.c_void_p]
                lib.ref_helmholtz_outer_error.restype = C.c_double
                lib.ref_helmholtz_outer_error.argtypes = [
                    C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
                lib.ref_helmholtz_point_project.argtypes = [
                    C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
                lib.ref_helmholtz_point_errors.argtypes = [
                    C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                    C.c_void_p]
            lib.ref_nearfield.argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                          C.c_uint, C.c_void_p, C.c_uint,
                                          C.c_int, C.c_void_p, C.c_uint]
            lib.ref_nearfield_blocks.argtypes = [
                C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p,
                C.c_void_p, C.c_void_p, C.c_int]
            lib.ref_h2_assemble_seconds.restype = C.c_double
            lib.ref_h2_assemble_seconds.argtypes = [C.c_void_p]
            lib.ref_h2_new.restype = C.c_void_p
            lib.ref_h2_new.argtypes = [C.c_void_p, C.c_uint, C.c_double,
                                       C.c_uint, C.c_int]
            lib.ref_h2_del.argtypes = [C.c_void_p]
            lib.ref_h2_idx.restype = C.c_uint
            lib.ref_h2_idx.argtypes = [C.c_void_p, C.c_void_p]
            lib.ref_h2_near_blocks.restype = C.c_uint
            lib.ref_h2_near_blocks.argtypes = [C.c_void_p, C.c_void_p,
                                               C.c_uint]
            lib.ref_h2_near_entries.argtypes = [C.c_void_p, C.c_uint,
                                                C.c_void_p]
            lib.ref_h2_sizes.argtypes = [C.c_void_p, C.c_void_p]
            lib.ref_h2_addevalsymm.argtypes = [C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
            lib.ref_h2_addeval.argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
            lib.ref_h2_gmres.restype = C.c_uint
            lib.ref_h2_gmres.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_double, C.c_uint, C.c_uint]
            lib.ref_h2_cg.restype = C.c_uint
            lib.ref_h2_cg.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_uint]
            if hip:
                lib.ref_bem_use_hip.argtypes = [C.c_void_p]
                lib.ref_bem_use_hip_deferred.argtypes = [C.c_void_p]
                lib.ref_bem_hip_precision.argtypes = [C.c_void_p, C.c_int]
                lib.ref_h2_hip_pack.argtypes = [C.c_void_p]
                lib.ref_h2_hip_addeval.argtypes = [C.c_void_p, C.c_int,
                                                   C.c_void_p, C.c_void_p,
                                                   C.c_void_p]
                lib.ref_h2_hip_gmres.restype = C.c_uint
                lib.ref_h2_hip_gmres.argtypes = [C.c_void_p, C.c_int,
                                                 C.c_void_p, C.c_void_p,
                                                 C.c_double, C.c_uint,
                                                 C.c_uint]
            cls._libs[key] = lib
        return cls._libs[key]

    @classmethod
    def mesh(cls, kind, split):
        """kind: 0 sphere, 1 cube, 2 cylinder.  Returns x, t, n, g."""
        lib = cls.lib(False)
        nv, nt = C.c_uint(), C.c_uint()
        lib.ref_mesh_size(kind, split, C.byref(nv), C.byref(nt))
        x = np.zeros((nv.value, 3))
        t = np.zeros((nt.value, 3), dtype=np.uint32)
        n = np.zeros((nt.value, 3))
        g = np.zeros(nt.value)
        lib.ref_mesh_fill(kind, split, _ptr(x), _ptr(t), _ptr(n), _ptr(g))
        return x, t, n, g

    @classmethod
    def rule(cls, q, q2, cls_id):
        lib = cls.lib(False)
        n = C.c_uint()
        base = C.c_double()
        lib.ref_singquad(q, q2, cls_id, None, None, None, C.byref(n),
                         C.byref(base))
        xc = 1 if cls_id == 4 else 2
        x = np.zeros(xc * n.value)
        y = np.zeros(xc * n.value)
        w = np.zeros(n.value)
        lib.ref_singquad(q, q2, cls_id, _ptr(x), _ptr(y), _ptr(w),
                         C.byref(n), C.byref(base))
        return x, y, w, base.value


class ReferenceBem(_BemBase):
    """A reference operator.  hip=True loads the reference build that also
    holds the H2Lib-side binding of the GPU library; use_hip() then installs
    the GPU callbacks into the reference's own bem3d object."""

    def __init__(self, kernel, op, basis_r, basis_c, x, t, hip=False,
                 fast=False, **kw):
        lib = Reference.lib(kernel == 1, hip, fast)
        self._lib = lib
        self.hip = hip
        self._new = lib.ref_bem_new
        self._nearfield = lib.ref_nearfield
        super().__init__(kernel, op, basis_r, basis_c, x, t, **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.ref_bem_del(self.h)
            self.h = None

    def laplace_point_integrate_l(self, kind):
        """Integrals of the reference's Laplace point-source data at
        (1.2, 1.2, 1.2) (0 Dirichlet, 1 Neumann) against the linear basis
        (real build only)."""
        out = np.zeros(len(self.x), dtype=np.float64)
        self._lib.ref_laplace_point_integrate_l(self.h, kind, _ptr(out))
        return out

    def helmholtz_point_project(self, kind, k, x0):
        """Point-source data (0 Dirichlet, 1 Neumann) projected onto the
        constant basis (complex build only)."""
        out = np.zeros(len(self.t), dtype=np.complex128)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self._lib.ref_helmholtz_point_project(self.h, kind, k, _ptr(x0),
                                              _ptr(out))
        return out

    def helmholtz_point_integrate(self, k, x0):
        out = np.zeros(self.ndof(self.basis_c), dtype=np.complex128)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self._lib.ref_helmholtz_point_integrate(self.h, k, _ptr(x0),
                                                _ptr(out))
        return out

    def eval_bw(self, w, eta, pts):
        """eval_brakhage_werner_c of the reference's Helmholtz test."""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        ww = np.ascontiguousarray(w, dtype=np.complex128)
        out = np.zeros(len(pts), dtype=np.complex128)
        self._lib.ref_eval_bw_c(self.h, _ptr(ww), eta, _ptr(pts), len(pts),
                                _ptr(out))
        return out

    def outer_error(self, x, k, x0):
        """max_rel_outer_error of the reference's Helmholtz test."""
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        return self._lib.ref_helmholtz_outer_error(self.h, _ptr(xx), k,
                                                   _ptr(x0))

    def helmholtz_point_errors(self, kind, k, x0, x):
        """(|x - data|_L2, |data|_L2) for the point-source data."""
        out = np.zeros(2)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        self._lib.ref_helmholtz_point_errors(self.h, kind, k, _ptr(x0),
                                             _ptr(xx), _ptr(out))
        return out

    def use_hip(self, deferred=False):
        """bem->nearfield / nearfield_far := the GPU callbacks
        (setup_hip_bem3d; deferred: setup_hip_deferred_bem3d, the driver
        flushes after its loops)."""
        if not self.hip:
            raise RuntimeError("load the reference with hip=True")
        if deferred:
            self._lib.ref_bem_use_hip_deferred(self.h)
        else:
            self._lib.ref_bem_use_hip(self.h)

    def hip_precision(self, mode):
        """Precision of the installed GPU callbacks
        (set_precision_hip_bem3d)."""
        self._lib.ref_bem_hip_precision(self.h, int(mode))

    def nearfield_blocks(self, ridx, cidx, blocks, threads=1):
        """All blocks (roff, rows, coff, cols) packed as the device plan
        packs them; returns the flat array."""
        ridx = _idx(ridx)
        cidx = _idx(cidx)
        blocks = np.ascontiguousarray(blocks, dtype=np.uint32).reshape(-1, 4)
        sizes = blocks[:, 1].astype(np.uint64) * blocks[:, 3]
        off = np.zeros(len(blocks), dtype=np.uint64)
        off[1:] = np.cumsum(sizes)[:-1]
        out = np.zeros(int(sizes.sum()), dtype=self.dtype)
        self._lib.ref_nearfield_blocks(self.h, _ptr(ridx), _ptr(cidx),
                                       len(blocks), _ptr(blocks), _ptr(off),
                                       _ptr(out), int(threads))
        return out


class ReferenceH2:
    """H2-matrix of a reference operator, built by interpolation."""

    def __init__(self, bem, clf, eta, m, admis=0):
        self.bem = bem
        self._lib = bem._lib
        self.h = self._lib.ref_h2_new(bem.h, clf, eta, m, admis)
        # wall seconds of the reference's assemble_bem3d_h2matrix
        self.assemble_s = self._lib.ref_h2_assemble_seconds(self.h)
        self.n = self._lib.ref_h2_idx(self.h, None)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.ref_h2_del(self.h)
            self.h = None

    def idx(self):
        out = np.zeros(self.n, dtype=np.uint32)
        self._lib.ref_h2_idx(self.h, _ptr(out))
        return out

    def near_blocks(self):
        nb = self._lib.ref_h2_near_blocks(self.h, None, 0)
        out = np.zeros((nb, 4), dtype=np.uint32)
        self._lib.ref_h2_near_blocks(self.h, _ptr(out), nb)
        return out

    def near_entries(self, b, rows, cols):
        out = np.zeros(rows * cols, dtype=self.bem.dtype)
        self._lib.ref_h2_near_entries(self.h, b, _ptr(out))
        return out.reshape(cols, rows).T.copy()

    def sizes(self):
        out = np.zeros(6)
        self._lib.ref_h2_sizes(self.h, _ptr(out))
        return out

    def addeval(self, x, y, alpha=1.0, trans=False):
        a = np.array([np.real(alpha), np.imag(alpha)])
        xx = np.ascontiguousarray(x, dtype=self.bem.dtype)
        yy = np.ascontiguousarray(y, dtype=self.bem.dtype).copy()
        self._lib.ref_h2_addeval(self.h, int(trans), _ptr(a), _ptr(xx),
                                 _ptr(yy))
        return yy

    def hip_addeval(self, x, y, alpha=1.0, trans=False):
        """The same product through new_hip_h2matrix +
        addeval_hip_h2matrix_avector (GPU)."""
        a = np.array([np.real(alpha), np.imag(alpha)])
        xx = np.ascontiguousarray(x, dtype=self.bem.dtype)
        yy = np.ascontiguousarray(y, dtype=self.bem.dtype).copy()
        self._lib.ref_h2_hip_addeval(self.h, int(trans), _ptr(a), _ptr(xx),
                                     _ptr(yy))
        return yy

    def hip_gmres(self, b, x0, eps, maxiter, kmax, device):
        """GMRES through the binding: device=False runs the reference's
        solve_gmres_avector with the GPU product as callback, device=True the
        device-resident solver."""
        bb = np.ascontiguousarray(b, dtype=self.bem.dtype)
        xx = np.ascontiguousarray(x0, dtype=self.bem.dtype).copy()
        it = self._lib.ref_h2_hip_gmres(self.h, int(bool(device)), _ptr(bb),
                                        _ptr(xx), eps, maxiter, kmax)
        return xx, it

    def addevalsymm(self, x, y, alpha=1.0):
        a = np.array([np.real(alpha), np.imag(alpha)])
        xx = np.ascontiguousarray(x, dtype=self.bem.dtype)
        yy = np.ascontiguousarray(y, dtype=self.bem.dtype).copy()
        self._lib.ref_h2_addevalsymm(self.h, _ptr(a), _ptr(xx), _ptr(yy))
        return yy

    def bw_solve(self, dlp, eta, b, eps, maxiter, kmax):
        """(KM - i eta V) x = b with this (single layer) matrix and the
        double layer H2-matrix dlp, the reference's GMRES."""
        bb = np.ascontiguousarray(b, dtype=np.complex128)
        xx = np.zeros_like(bb)
        it = self._lib.ref_helmholtz_bw_solve(self.h, dlp.h, eta, _ptr(bb),
                                              _ptr(xx), eps, maxiter, kmax)
        return xx, it

    def cg(self, b, x0, eps, maxiter):
        """The reference's solve_cg_avector on this H2-matrix."""
        bb = np.ascontiguousarray(b, dtype=self.bem.dtype)
        xx = np.ascontiguousarray(x0, dtype=self.bem.dtype).copy()
        it = self._lib.ref_h2_cg(self.h, _ptr(bb), _ptr(xx), eps, maxiter)
        return xx, it

    def gmres(self, b, x0, eps, maxiter, kmax):
        bb = np.ascontiguousarray(b, dtype=self.bem.dtype)
        xx = np.ascontiguousarray(x0, dtype=self.bem.dtype).copy()
        it = self._lib.ref_h2_gmres(self.h, _ptr(bb), _ptr(xx), eps, maxiter,
                                    kmax)
        return xx, it


def reference_dirichlet(mesh_x, mesh_t, clf, eta, m, eps, maxiter, kmax):
    """The reference's interior Dirichlet problem with GMRES(kmax): returns
    f, b, x, steps, relative Neumann error."""
    lib = Reference.lib(False)
    lib.ref_dirichlet_laplace.restype = C.c_uint
    lib.ref_dirichlet_laplace.argtypes = [C.c_void_p, C.c_void_p, C.c_double,
                                          C.c_uint, C.c_uint, C.c_void_p,
                                          C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_double)]
    slp = ReferenceBem(0, 0, 1, 1, mesh_x, mesh_t, q=2, q2=4)
    dlp = ReferenceBem(0, 1, 1, 1, mesh_x, mesh_t, q=2, q2=4, alpha=0.5)
    hs = ReferenceH2(slp, clf, eta, m, 0)
    hd = ReferenceH2(dlp, clf, eta, m, 0)
    n = hs.n
    f, b, x = np.zeros(n), np.zeros(n), np.zeros(n)
    err = C.c_double()
    it = lib.ref_dirichlet_laplace(hs.h, hd.h, eps, maxiter, kmax, _ptr(f),
                                   _ptr(b), _ptr(x), C.byref(err))
    return f, b, x, it, err.value

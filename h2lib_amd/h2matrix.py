"""H2-matrices on the device: construction by interpolation (C host code in
libh2bem.so, restating setup_h2matrix_aprx_inter_bem3d and
assemble_bem3d_h2matrix, Library/bem3d.c:9658-9686, 10313-10446) with the
near field from the HIP quadrature, and the products addeval / addevaltrans
(Library/h2matrix.c:1268-1286, 1345-1364) in libh2hip.so.
"""
import ctypes as C
import sys

import numpy as np

from ._lib import check, hip
from .bem3d import ADLP, DLP, LAPLACE, SLP, NearfieldPlan
from .geometry import bem


def dn_side(layer):
    """Side whose cluster basis takes the normal derivative of the
    Lagrange polynomials: 1 the columns (double layer, n_y), 2 the rows
    (adjoint double layer, n_x), 0 none (single layer)."""
    return {SLP: 0, DLP: 1, ADLP: 2}[layer]


class _H2Interp(C.Structure):
    _fields_ = [("m", C.c_uint32), ("k", C.c_uint32),
                ("rpts", C.POINTER(C.c_double)),
                ("cpts", C.POINTER(C.c_double)),
                ("rvoff", C.POINTER(C.c_uint64)),
                ("cvoff", C.POINTER(C.c_uint64)),
                ("rV", C.POINTER(C.c_double)), ("cV", C.POINTER(C.c_double)),
                ("rvlen", C.c_uint64), ("cvlen", C.c_uint64),
                ("rE", C.POINTER(C.c_double)), ("cE", C.POINTER(C.c_double)),
                ("nfar", C.c_uint32), ("S", C.POINTER(C.c_double))]


_sigs_done = False


def _sigs():
    global _sigs_done
    if _sigs_done:
        return
    b = bem()
    ip = C.POINTER(_H2Interp)
    b.h2b_build_h2interp.restype = ip
    b.h2b_build_h2interp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_uint, C.c_uint]
    b.h2b_h2interp_coupling_laplace.argtypes = [ip, C.c_void_p]
    b.h2b_h2interp_coupling_helmholtz.argtypes = [ip, C.c_void_p,
                                                  C.c_double, C.c_double,
                                                  C.c_void_p]
    b.h2b_h2interp_free.argtypes = [ip]
    h = hip()
    vp, u32, dp = C.c_void_p, C.c_uint32, C.c_void_p
    h.h2hip_h2matrix_create.restype = C.c_int
    h.h2hip_h2matrix_create.argtypes = [C.POINTER(vp), C.c_int, u32, dp, dp,
                                        u32, dp, u32, dp, dp, u32, dp, u32,
                                        dp, u32, dp]
    h.h2hip_h2matrix_destroy.argtypes = [vp]
    h.h2hip_h2matrix_set_basis.argtypes = [vp, C.c_int, dp, dp, dp,
                                           C.c_uint64]
    h.h2hip_h2matrix_set_couplings.argtypes = [vp, dp, C.c_uint64]
    h.h2hip_h2matrix_set_nearfield.argtypes = [vp, dp]
    h.h2hip_addevalsymm_h2matrix.restype = C.c_int
    h.h2hip_addevalsymm_h2matrix.argtypes = [dp, vp, vp, vp, vp]
    h.h2hip_h2matrix_set_schedule.restype = C.c_int
    h.h2hip_h2matrix_set_schedule.argtypes = [vp, C.c_int]
    h.h2hip_h2matrix_set_coupling_mode.restype = C.c_int
    h.h2hip_h2matrix_set_coupling_mode.argtypes = [vp, C.c_int]
    h.h2hip_h2matrix_set_adjoint_storage.restype = C.c_int
    h.h2hip_h2matrix_set_adjoint_storage.argtypes = [vp, C.c_int]
    h.h2hip_h2matrix_set_mfma.restype = C.c_int
    h.h2hip_h2matrix_set_mfma.argtypes = [vp, C.c_int]
    h.h2hip_h2matrix_set_coupling_pairs.restype = C.c_int
    h.h2hip_h2matrix_set_coupling_pairs.argtypes = [vp, C.c_int]
    h.h2hip_h2matrix_coupling_pairs.restype = C.c_int
    h.h2hip_h2matrix_coupling_pairs.argtypes = [vp]
    h.h2hip_h2matrix_interp_bases.restype = C.c_int
    h.h2hip_h2matrix_interp_bases.argtypes = [vp, C.c_int, vp, C.c_int,
                                              C.c_int, C.c_uint32, C.c_uint32,
                                              vp, vp]
    b.h2b_interp_points.argtypes = [C.c_void_p, C.c_uint, C.c_void_p]
    h.h2hip_h2matrix_interp_couplings.restype = C.c_int
    h.h2hip_h2matrix_interp_couplings.argtypes = [vp, C.c_int, dp, u32, dp,
                                                  dp, vp]
    h.h2hip_h2matrix_nearfield_device.restype = C.c_void_p
    h.h2hip_h2matrix_nearfield_device.argtypes = [vp]
    h.h2hip_h2matrix_bytes.argtypes = [vp, dp]
    h.h2hip_h2matrix_storage.argtypes = [vp, dp]
    h.h2hip_addeval_h2matrix.argtypes = [dp, vp, vp, vp, vp]
    h.h2hip_addevaltrans_h2matrix.argtypes = [dp, vp, vp, vp, vp]
    h.h2hip_addeval_h2matrix_batch.restype = C.c_int
    h.h2hip_addeval_h2matrix_batch.argtypes = [vp, C.c_uint32, dp, dp, dp,
                                               C.c_int, vp]
    _sigs_done = True


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def tree_arrays(t):
    """(off, size, father) per node and the ranks of a ClusterTree."""
    father = np.full(t.nnodes, 0xffffffff, dtype=np.uint32)
    for i in range(t.nnodes):
        for s in range(int(t.sons[i])):
            father[t.son[i][s]] = i
    tree = np.stack([t.off, t.size, father], axis=1).astype(np.uint32)
    return np.ascontiguousarray(tree)


class Interpolation:
    """Host arrays of an H2-matrix by interpolation (cluster bases of both
    trees, transfer and coupling matrices), built by libh2bem.so without
    touching the device: setup_h2matrix_aprx_inter_bem3d and the basis and
    coupling fills of assemble_bem3d_h2matrix (Library/bem3d.c:9658-9686,
    10313-10446).  V of node i: size x k column-major at voff[i]; E of node
    i: k x k column-major at i k^2; S of far block b at b k^2."""

    def __init__(self, mesh, rt, ct, btree, row_basis, col_basis, dlp, m, q,
                 couplings=True, bases=True):
        """dlp: normal derivative of the Lagrange polynomials on the
        column basis (True or 1: double layer), on the row basis (2:
        adjoint double layer) or on neither (False or 0).  bases=False:
        only the trees, blocks and interpolation points (the bases are then
        built on the device, H2Matrix device_bases)."""
        _sigs()
        if not bases:
            self._ip = None
            self.m, self.k = m, m ** 3
            self.eld = self.k * self.k
            self.rtree, self.ctree = tree_arrays(rt), tree_arrays(ct)
            self.rrank = np.full(rt.nnodes, self.k, dtype=np.uint32)
            self.crank = np.full(ct.nnodes, self.k, dtype=np.uint32)
            self.ridx, self.cidx = rt.idx, ct.idx
            self.near = np.ascontiguousarray(btree.near, dtype=np.uint32)
            self.far = np.ascontiguousarray(btree.far, dtype=np.uint32)
            self.rpts = np.zeros(3 * m * rt.nnodes)
            self.cpts = np.zeros(3 * m * ct.nnodes)
            bem().h2b_interp_points(rt._tp, m, self.rpts.ctypes.data_as(
                C.c_void_p))
            bem().h2b_interp_points(ct._tp, m, self.cpts.ctypes.data_as(
                C.c_void_p))
            self.S = None
            self._bp = btree._bp
            self.nfar = len(self.far)
            return
        ip = bem().h2b_build_h2interp(mesh._sp, rt._tp, ct._tp, btree._bp,
                                      row_basis, col_basis, int(dlp), m, q)
        if not ip:
            raise RuntimeError("interpolation build failed")
        if couplings:
            bem().h2b_h2interp_coupling_laplace(ip, btree._bp)
        self._ip = ip
        I = ip.contents
        self.m, self.k = I.m, I.k
        kk = I.k * I.k
        self.eld = kk
        self.rtree, self.ctree = tree_arrays(rt), tree_arrays(ct)
        self.rrank = np.full(rt.nnodes, I.k, dtype=np.uint32)
        self.crank = np.full(ct.nnodes, I.k, dtype=np.uint32)
        self.ridx, self.cidx = rt.idx, ct.idx
        self.near = np.ascontiguousarray(btree.near, dtype=np.uint32)
        self.far = np.ascontiguousarray(btree.far, dtype=np.uint32)
        as_np = np.ctypeslib.as_array
        self.rvoff = as_np(I.rvoff, (rt.nnodes,))
        self.cvoff = as_np(I.cvoff, (ct.nnodes,))
        self.rV = as_np(I.rV, (max(1, I.rvlen),))
        self.cV = as_np(I.cV, (max(1, I.cvlen),))
        self.rE = as_np(I.rE, (rt.nnodes * kk,))
        self.cE = as_np(I.cE, (ct.nnodes * kk,))
        self.S = as_np(I.S, (max(1, I.nfar * kk),)) if couplings else None
        self.rpts = as_np(I.rpts, (3 * self.m * rt.nnodes,))
        self.cpts = as_np(I.cpts, (3 * self.m * ct.nnodes,))
        self._bp = btree._bp
        self.nfar = I.nfar

    def helmholtz_couplings(self, k):
        """Complex couplings of the Helmholtz kernel with wave number k
        (same layout as S)."""
        out = np.zeros(2 * max(1, self.nfar) * self.eld)
        bem().h2b_h2interp_coupling_helmholtz(
            self._ip, self._bp, float(np.real(k)), float(np.imag(k)),
            out.ctypes.data_as(C.c_void_p))
        return out.view(np.complex128)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        ip = getattr(self, "_ip", None)
        if ip:
            bem().h2b_h2interp_free(ip)
            self._ip = None


class H2Matrix:
    """Device H2-matrix of one operator (mirror of h2matrix,
    Library/h2matrix.h:48-71)."""

    def __init__(self, op, mesh, rt, ct, btree, m, build_nearfield=True,
                 stream=None, device_couplings=True, symmetric=False,
                 nearplan=None, device_bases=True):
        """symmetric: keep only the blocks on and below the block diagonal
        (half the storage) for addevalsymm; the operator must be
        self-adjoint with equal row and column bases (a Laplace single
        layer operator).  nearplan: the near-field plan of another
        H2Matrix on the same surface, trees, blocks and bases (the plan
        does not depend on the kernel), reused instead of planned again."""
        _sigs()
        if symmetric and (rt is not ct or op.row_basis != op.col_basis or
                          op.layer != SLP or op.kernel != LAPLACE):
            raise ValueError("the symmetric storage needs a self-adjoint "
                             "operator on one tree with equal bases")
        self.symmetric = symmetric
        self.op, self.rt, self.ct, self.btree = op, rt, ct, btree
        self.dtype = op.dtype
        self.m = m
        device_bases = device_bases and device_couplings
        I = Interpolation(mesh, rt, ct, btree, op.row_basis, op.col_basis,
                          dn_side(op.layer), m, op.q_regular,
                          couplings=not device_couplings
                          and op.kernel == LAPLACE,
                          bases=not device_bases)
        self.interp = I
        self.k = I.k
        self.n = rt.n
        near, far = I.near, I.far
        nbl = btree.near_block_list()
        if symmetric:
            # blocks on or below the block diagonal: row node >= column
            # node (sons are numbered in order, as the reference walks them)
            keep = near[:, 0] >= near[:, 1]
            near = np.ascontiguousarray(near[keep])
            nbl = np.ascontiguousarray(np.asarray(nbl)[keep])
            far = np.ascontiguousarray(far[far[:, 0] > far[:, 1]])
        self.near_nodes, self.far_nodes = near, far
        h = C.c_void_p()
        check(hip().h2hip_h2matrix_create(
            C.byref(h), int(op.is_complex), rt.nnodes, _ptr(I.rtree),
            _ptr(I.rrank), rt.n, _ptr(rt.idx), ct.nnodes, _ptr(I.ctree),
            _ptr(I.crank), ct.n, _ptr(ct.idx), len(near), _ptr(near),
            len(far), _ptr(far)), "h2hip_h2matrix_create")
        self._h = h
        if device_bases:
            dside = dn_side(op.layer)
            for sd, basis, dn, pts in ((0, op.row_basis, int(dside == 2),
                                        I.rpts),
                                       (1, op.col_basis, int(dside == 1),
                                        I.cpts)):
                check(hip().h2hip_h2matrix_interp_bases(
                    h, sd, op.gr.handle, basis, dn, m, op.q_regular,
                    _ptr(pts), C.c_void_p(stream.handle if stream else 0)),
                    "h2hip_h2matrix_interp_bases")
        else:
            check(hip().h2hip_h2matrix_set_basis(
                h, 0, _ptr(I.rV), _ptr(I.rvoff), _ptr(I.rE), I.eld),
                "h2hip_h2matrix_set_basis")
            check(hip().h2hip_h2matrix_set_basis(
                h, 1, _ptr(I.cV), _ptr(I.cvoff), _ptr(I.cE), I.eld),
                "h2hip_h2matrix_set_basis")
        if device_couplings:
            kk = np.array([np.real(op.k), np.imag(op.k)], dtype=np.float64)
            check(hip().h2hip_h2matrix_interp_couplings(
                h, op.kernel, _ptr(kk), m, _ptr(I.rpts), _ptr(I.cpts),
                C.c_void_p(stream.handle if stream else 0)),
                "h2hip_h2matrix_interp_couplings")
        else:
            if symmetric:
                raise ValueError("symmetric storage computes its couplings "
                                 "on the device")
            S = I.S if op.kernel == LAPLACE else I.helmholtz_couplings(op.k)
            check(hip().h2hip_h2matrix_set_couplings(h, _ptr(S), I.eld),
                  "h2hip_h2matrix_set_couplings")
        self.nearplan = None
        if build_nearfield:
            if nearplan is not None and not symmetric:
                self.nearplan = nearplan
            else:
                self.nearplan = NearfieldPlan(op.gr, op.row_basis,
                                              op.col_basis, rt.idx, ct.idx,
                                              nbl)
            self.assemble_nearfield(stream)

    @classmethod
    def from_arrays(cls, interp, near_values, couplings=None, rt=None,
                    is_complex=False):
        """A device H2-matrix from host arrays: the cluster bases of an
        Interpolation, near-field blocks packed in the order of interp.near
        (None: left unset, e.g. for copy_nearfield_from_device) and
        couplings (default: interp.S), real or complex."""
        _sigs()
        self = cls.__new__(cls)
        self.op = None
        self.rt = rt
        self.nearplan = None
        self.interp = I = interp
        self.k = I.k
        self.n = len(I.ridx)
        self.dtype = np.complex128 if is_complex else np.float64
        h = C.c_void_p()
        check(hip().h2hip_h2matrix_create(
            C.byref(h), int(is_complex), len(I.rtree), _ptr(I.rtree),
            _ptr(I.rrank), len(I.ridx), _ptr(I.ridx), len(I.ctree),
            _ptr(I.ctree), _ptr(I.crank), len(I.cidx), _ptr(I.cidx),
            len(I.near), _ptr(I.near), len(I.far), _ptr(I.far)),
            "h2hip_h2matrix_create")
        self._h = h
        check(hip().h2hip_h2matrix_set_basis(
            h, 0, _ptr(I.rV), _ptr(I.rvoff), _ptr(I.rE), I.eld),
            "h2hip_h2matrix_set_basis")
        check(hip().h2hip_h2matrix_set_basis(
            h, 1, _ptr(I.cV), _ptr(I.cvoff), _ptr(I.cE), I.eld),
            "h2hip_h2matrix_set_basis")
        S = I.S if couplings is None else couplings
        S = np.ascontiguousarray(S, dtype=self.dtype)
        check(hip().h2hip_h2matrix_set_couplings(h, _ptr(S), I.eld),
              "h2hip_h2matrix_set_couplings")
        if near_values is not None:
            N = np.ascontiguousarray(near_values, dtype=self.dtype)
            check(hip().h2hip_h2matrix_set_nearfield(h, _ptr(N)),
                  "h2hip_h2matrix_set_nearfield")
        return self

    def copy_nearfield_from_device(self, d_src, entries):
        """Near-field blocks (packed in the order of the matrix's near
        blocks) from device memory at address d_src, entries elements."""
        es = 16 if self.dtype == np.complex128 else 8
        dst = hip().h2hip_h2matrix_nearfield_device(self._h)
        check(hip().h2hip_memcpy(C.c_void_p(dst), C.c_void_p(d_src),
                                 es * entries, 2), "h2hip_memcpy")

    def assemble_nearfield(self, stream=None):
        """Near-field blocks by the HIP quadrature, straight into the
        matrix's device storage."""
        d = hip().h2hip_h2matrix_nearfield_device(self._h)
        self.nearplan.execute(self.op, d, stream.handle if stream else None)
        if stream is None:
            check(hip().h2hip_stream_sync(None), "h2hip_stream_sync")

    def set_schedule(self, level_sync):
        """Transforms: upward sweep in one launch with counters (0), one
        launch per tree height (1), or flat transforms over the ancestors'
        bases restricted to the leaves (2, products y += alpha H x of one or
        several vectors; the adjoint uses 1)."""
        check(hip().h2hip_h2matrix_set_schedule(self._h, int(level_sync)),
              "h2hip_h2matrix_set_schedule")
        self.level_sync = int(level_sync)

    def set_coupling_mode(self, recompute):
        """Products y += alpha H x read the stored couplings (0) or compute
        them from the interpolation points (1; device couplings only)."""
        check(hip().h2hip_h2matrix_set_coupling_mode(self._h,
                                                     int(bool(recompute))),
              "h2hip_h2matrix_set_coupling_mode")
        self.recompute = int(bool(recompute))

    def set_adjoint_storage(self, keep):
        """Keep conjugate-transposed copies of the couplings and near blocks
        for the adjoint product (made by its first call after the blocks
        were set), which then runs the forward kernels."""
        check(hip().h2hip_h2matrix_set_adjoint_storage(self._h,
                                                       int(bool(keep))),
              "h2hip_h2matrix_set_adjoint_storage")
        self.adjoint_storage = int(bool(keep))

    def set_mfma(self, on):
        """Products of several vectors of a real matrix of rank 64 multiply
        the stored couplings on FP64 matrix cores."""
        check(hip().h2hip_h2matrix_set_mfma(self._h, int(bool(on))),
              "h2hip_h2matrix_set_mfma")
        self.mfma = int(bool(on))

    def set_coupling_pairs(self, on):
        """Computed couplings of a symmetric order-4 or -5 structure (Laplace
        or Helmholtz): each value evaluated once for a block and its mirror
        (the default)."""
        check(hip().h2hip_h2matrix_set_coupling_pairs(self._h, int(bool(on))),
              "h2hip_h2matrix_set_coupling_pairs")

    def coupling_pairs(self):
        """Mirrored pairs of far blocks found when the couplings were
        interpolated (0 when the structure is not symmetric)."""
        return int(hip().h2hip_h2matrix_coupling_pairs(self._h))

    def bytes(self):
        """Bytes one product y += alpha H x reads in the current schedule
        and coupling mode (couplings 0 when computed in the product)."""
        out = np.zeros(3, dtype=np.uint64)
        check(hip().h2hip_h2matrix_bytes(self._h, _ptr(out)),
              "h2hip_h2matrix_bytes")
        return {"near": int(out[0]), "coupling": int(out[1]),
                "bases": int(out[2])}

    def storage(self):
        """Storage in the reference's terms (near field, couplings of all
        admissible blocks, leaf and transfer matrices of both sides, one
        copy each) and the device bytes the matrix holds."""
        out = np.zeros(4, dtype=np.uint64)
        check(hip().h2hip_h2matrix_storage(self._h, _ptr(out)),
              "h2hip_h2matrix_storage")
        return {"near": int(out[0]), "coupling": int(out[1]),
                "bases": int(out[2]), "device": int(out[3])}

    def algorithmic_bytes(self):
        """Bytes of one product by the reference's data: near field,
        couplings, leaf and transfer matrices of both sides, and the vector
        traffic (x read, y read and written, coefficient vectors written
        and read once)."""
        st = self.storage()
        es = 16 if self.dtype == np.complex128 else 8
        kt = sum(int(x) for x in self.interp.rrank) + \
            sum(int(x) for x in self.interp.crank)
        return (st["near"] + st["coupling"] + st["bases"] +
                es * (3 * self.n + 2 * kt))

    def addeval(self, alpha, d_x, d_y, stream=None, trans=False):
        """y += alpha H x (trans: y += alpha H^* x) on device arrays."""
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        f = (hip().h2hip_addevaltrans_h2matrix if trans
             else hip().h2hip_addeval_h2matrix)
        check(f(_ptr(a), self._h, C.c_void_p(d_x.ptr), C.c_void_p(d_y.ptr),
                C.c_void_p(stream.handle if stream else 0)),
              "h2hip_addeval_h2matrix")

    def addevalsymm(self, alpha, d_x, d_y, stream=None):
        """y += alpha H x from the blocks on and below the block diagonal
        (addevalsymm_h2matrix_avector)."""
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        check(hip().h2hip_addevalsymm_h2matrix(
            _ptr(a), self._h, C.c_void_p(d_x.ptr), C.c_void_p(d_y.ptr),
            C.c_void_p(stream.handle if stream else 0)),
            "h2hip_addevalsymm_h2matrix")

    def addeval_batch(self, alphas, d_xs, d_ys, stream=None, trans=False):
        """y_v += alpha_v H x_v (trans: H^*) for up to 8 device arrays at
        once; the matrix is read once for all of them."""
        nv = len(d_xs)
        a = np.zeros(2 * nv)
        a[0::2] = np.real(alphas)
        a[1::2] = np.imag(alphas)
        xs = (C.c_void_p * nv)(*[d.ptr for d in d_xs])
        ys = (C.c_void_p * nv)(*[d.ptr for d in d_ys])
        check(hip().h2hip_addeval_h2matrix_batch(
            self._h, nv, _ptr(a), C.cast(xs, C.c_void_p),
            C.cast(ys, C.c_void_p), int(bool(trans)),
            C.c_void_p(stream.handle if stream else 0)),
            "h2hip_addeval_h2matrix_batch")

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_h2matrix_destroy(h)
            self._h = None

"""Host-side mirror of the reference's BEM3D operator interface.

The names follow the reference (H2Lib): ``new_slp_laplace_bem3d``
(Library/laplacebem3d.c:843), ``new_dlp_laplace_bem3d``
(Library/laplacebem3d.c:1021), ``new_slp_helmholtz_bem3d``
(Library/helmholtzbem3d.c:1008), ``new_dlp_helmholtz_bem3d``
(Library/helmholtzbem3d.c:1177), and the callbacks ``bem.nearfield`` /
``bem.nearfield_far`` (Library/bem3d.h:380-381, 405-406) with the same
argument meaning: row indices, column indices (None: identity), ``ntrans``
and the caller-owned column-major matrix ``N`` whose window is written.
All work is done by libh2hip.so on the GPU.
"""
import ctypes as C
import sys

import numpy as np

from ._lib import check, hip

BASIS_CONSTANT_BEM3D = 0
BASIS_LINEAR_BEM3D = 1

LAPLACE = 0
HELMHOLTZ = 1
SLP = 0
DLP = 1
ADLP = 2


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a


class Surface3d:
    """Triangulated surface on the device (mirror of surface3d,
    Library/surface3d.h:45-71).  Normals and Gram determinants are computed
    as prepare_surface3d does (Library/surface3d.c:51-108) unless given."""

    def __init__(self, x, t, n=None, g=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.t = np.ascontiguousarray(t, dtype=np.uint32).reshape(-1, 3)
        self.vertices = len(self.x)
        self.triangles = len(self.t)
        nn = None if n is None else np.ascontiguousarray(n, np.float64)
        gg = None if g is None else np.ascontiguousarray(g, np.float64)
        h = C.c_void_p()
        check(hip().h2hip_surface_create(
            C.byref(h), self.vertices, _ptr(self.x), self.triangles,
            _ptr(self.t), _ptr(nn), _ptr(gg)), "h2hip_surface_create")
        self._h = h
        self.n = np.zeros((self.triangles, 3))
        self.g = np.zeros(self.triangles)
        check(hip().h2hip_surface_geometry(h, _ptr(self.n), _ptr(self.g)),
              "h2hip_surface_geometry")

    @property
    def handle(self):
        return self._h

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_surface_destroy(h)
            self._h = None


class Bem3d:
    """One boundary integral operator with its near-field quadrature
    (mirror of bem3d, Library/bem3d.h:290-749)."""

    def __init__(self, gr, kernel, layer, q_regular, q_singular, row_basis,
                 col_basis, alpha=0.0, k=0.0):
        self.gr = gr
        self.kernel = kernel
        self.layer = layer
        self.row_basis = row_basis
        self.col_basis = col_basis
        self.q_regular = q_regular
        self.q_singular = q_singular
        self.alpha = alpha if layer != SLP else 0.0
        self.k = k
        self.is_complex = kernel == HELMHOLTZ
        self.dtype = np.complex128 if self.is_complex else np.float64
        a = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        kk = np.array([np.real(k), np.imag(k)], dtype=np.float64)
        h = C.c_void_p()
        check(hip().h2hip_op_create(C.byref(h), gr.handle, kernel, layer,
                                    row_basis, col_basis, q_regular,
                                    q_singular, _ptr(a), _ptr(kk)),
              "h2hip_op_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_op_destroy(h)
            self._h = None

    def set_precision(self, mode):
        """0: FP64 kernel evaluations (default); 1: mixed precision (FP32
        kernel values, FP64 geometry differences and outer sums); 2: as 1
        with FP32 accumulation of the regular pairs throughout, the outer
        sums compensated (Kahan)."""
        check(hip().h2hip_op_set_precision(self._h, int(mode)),
              "h2hip_op_set_precision")
        self.precision = int(mode)

    def ndof(self, basis):
        return self.gr.vertices if basis else self.gr.triangles

    def _fill(self, far, ridx, cidx, ntrans, N):
        ridx = _u32(ridx)
        cidx = _u32(cidx)
        rows = len(ridx) if ridx is not None else self.ndof(self.row_basis)
        cols = len(cidx) if cidx is not None else self.ndof(self.col_basis)
        if N is None:
            shp = (cols, rows) if ntrans else (rows, cols)
            N = np.zeros(shp, dtype=self.dtype, order="F")
        if N.dtype != self.dtype:
            raise TypeError("N must have dtype %s" % self.dtype)
        wr, wc = (cols, rows) if ntrans else (rows, cols)
        if N.ndim != 2 or N.shape[0] < wr or N.shape[1] < wc or \
                N.strides[0] != N.itemsize:
            raise ValueError("N must be a column-major %d x %d window" %
                             (wr, wc))
        ld = N.strides[1] // N.itemsize
        check(hip().h2hip_nearfield(self._h, int(far), _ptr(ridx), rows,
                                    _ptr(cidx), cols, int(bool(ntrans)),
                                    N.ctypes.data_as(C.c_void_p), ld),
              "h2hip_nearfield")
        return N

    def nearfield(self, ridx, cidx, ntrans=False, N=None):
        """bem->nearfield(ridx, cidx, bem, ntrans, N); returns N."""
        return self._fill(False, ridx, cidx, ntrans, N)

    def nearfield_far(self, ridx, cidx, ntrans=False, N=None):
        """bem->nearfield_far(ridx, cidx, bem, ntrans, N); returns N."""
        return self._fill(True, ridx, cidx, ntrans, N)


def new_slp_laplace_bem3d(gr, q_regular, q_singular, row_basis, col_basis):
    return Bem3d(gr, LAPLACE, SLP, q_regular, q_singular, row_basis,
                 col_basis)


def new_dlp_laplace_bem3d(gr, q_regular, q_singular, row_basis, col_basis,
                          alpha):
    return Bem3d(gr, LAPLACE, DLP, q_regular, q_singular, row_basis,
                 col_basis, alpha=alpha)


def new_slp_helmholtz_bem3d(k, gr, q_regular, q_singular, row_basis,
                            col_basis):
    return Bem3d(gr, HELMHOLTZ, SLP, q_regular, q_singular, row_basis,
                 col_basis, k=k)


def new_dlp_helmholtz_bem3d(k, gr, q_regular, q_singular, row_basis,
                            col_basis, alpha):
    return Bem3d(gr, HELMHOLTZ, DLP, q_regular, q_singular, row_basis,
                 col_basis, alpha=alpha, k=k)


def new_adlp_laplace_bem3d(gr, q_regular, q_singular, row_basis, col_basis,
                           alpha):
    """new_adlp_laplace_bem3d (Library/laplacebem3d.c:1180-1346): the
    adjoint double layer <y - x, n_x>/(4 pi |x - y|^3) plus alpha times the
    mass matrix."""
    return Bem3d(gr, LAPLACE, ADLP, q_regular, q_singular, row_basis,
                 col_basis, alpha=alpha)


def new_adlp_helmholtz_bem3d(k, gr, q_regular, q_singular, row_basis,
                             col_basis, alpha):
    """new_adlp_helmholtz_bem3d (Library/helmholtzbem3d.c:1346-1514)."""
    return Bem3d(gr, HELMHOLTZ, ADLP, q_regular, q_singular, row_basis,
                 col_basis, alpha=alpha, k=k)


class NearfieldPlan:
    """Many near-field blocks computed by one launch into a packed device
    buffer (the near-field part of assemble_bem3d_h2matrix,
    Library/bem3d.c:10313-10346)."""

    def __init__(self, gr, row_basis, col_basis, ridx_all, cidx_all, blocks):
        self.gr = gr
        self.row_basis = row_basis
        self.col_basis = col_basis
        self.ridx_all = _u32(ridx_all)
        self.cidx_all = _u32(cidx_all)
        self.blocks = np.ascontiguousarray(blocks, dtype=np.uint32).reshape(
            -1, 4)
        h = C.c_void_p()
        check(hip().h2hip_nfplan_create(
            C.byref(h), gr.handle, row_basis, col_basis, _ptr(self.ridx_all),
            len(self.ridx_all), _ptr(self.cidx_all), len(self.cidx_all),
            len(self.blocks), _ptr(self.blocks)), "h2hip_nfplan_create")
        self._h = h
        self.entries = int(hip().h2hip_nfplan_entries(h))
        self.offsets = np.zeros(len(self.blocks), dtype=np.uint64)
        check(hip().h2hip_nfplan_offsets(h, _ptr(self.offsets)),
              "h2hip_nfplan_offsets")

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        h = getattr(self, "_h", None)
        if h:
            hip().h2hip_nfplan_destroy(h)
            self._h = None

    def work(self, op, far=False):
        """(evaluations per class [4], pairs per class [4])."""
        out = np.zeros(8)
        check(hip().h2hip_nfplan_work(self._h, op.handle, int(far),
                                      _ptr(out)), "h2hip_nfplan_work")
        return out[:4], out[4:]

    def set_schedule(self, sorted_tiles, persistent, fused=False):
        """Launch schedule (entries are the same up to the order of atomic
        sums of the linear basis; constant basis bit for bit): tiles in
        decreasing order of work, and/or the persistent regular kernel;
        fused: each tile's workgroup also evaluates the tile's pairs with
        common vertices (one launch for all blocks; regular order 2,
        singular order 4)."""
        check(hip().h2hip_nfplan_set_schedule(self._h, int(bool(sorted_tiles)),
                                              int(bool(persistent))),
              "h2hip_nfplan_set_schedule")
        check(hip().h2hip_nfplan_set_fused(self._h, int(bool(fused))),
              "h2hip_nfplan_set_fused")
        self.schedule = (bool(sorted_tiles), bool(persistent), bool(fused))

    def set_tile(self, tile):
        """Tile edge (8..64 DOFs): the blocks are cut again into tiles of at
        most tile x tile DOFs, the schedule kept; the packed output does not
        change."""
        check(hip().h2hip_nfplan_set_tile(self._h, int(tile)),
              "h2hip_nfplan_set_tile")
        self.tile = int(tile)

    def execute(self, op, d_out, stream=None, far=False):
        """Computes all blocks into the device buffer at address d_out."""
        check(hip().h2hip_nearfield_execute(op.handle, self._h, int(far),
                                            C.c_void_p(d_out),
                                            C.c_void_p(stream or 0)),
              "h2hip_nearfield_execute")

    def block(self, flat, b):
        """Block b of a packed host copy of the output, as (rows, cols)."""
        rows, cols = int(self.blocks[b, 1]), int(self.blocks[b, 3])
        o = int(self.offsets[b])
        return flat[o:o + rows * cols].reshape(cols, rows).T

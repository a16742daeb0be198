"""Meshes, cluster trees and block trees from libh2bem.so (C host code).

Mirrors new_sphere_macrosurface3d / new_cube_macrosurface3d +
build_from_macrosurface3d_surface3d (Library/macrosurface3d.c:81, 390,
707), build_bem3d_cluster (Library/bem3d.c:932) and build_strict_block
(Library/block.c:258) of the reference.
"""
import ctypes as C
import os
import sys

import numpy as np

from ._lib import H2HipError, hip

_HERE = os.path.dirname(os.path.abspath(__file__))
BEM_SO = os.path.join(_HERE, "libh2bem.so")


class _Surface(C.Structure):
    _fields_ = [("vertices", C.c_uint32), ("triangles", C.c_uint32),
                ("x", C.POINTER(C.c_double)), ("t", C.POINTER(C.c_uint32)),
                ("n", C.POINTER(C.c_double)), ("g", C.POINTER(C.c_double)),
                ("hmin", C.c_double), ("hmax", C.c_double)]


class _CNode(C.Structure):
    _fields_ = [("off", C.c_uint32), ("size", C.c_uint32),
                ("sons", C.c_uint32), ("son", C.c_uint32 * 2),
                ("desc", C.c_uint32), ("level", C.c_uint32),
                ("bmin", C.c_double * 3), ("bmax", C.c_double * 3)]


_CNODE_DTYPE = np.dtype([("off", np.uint32), ("size", np.uint32),
                         ("sons", np.uint32), ("son", np.uint32, (2,)),
                         ("desc", np.uint32), ("level", np.uint32),
                         ("bmin", np.float64, (3,)),
                         ("bmax", np.float64, (3,))], align=True)
assert _CNODE_DTYPE.itemsize == C.sizeof(_CNode)


class _CTree(C.Structure):
    _fields_ = [("n", C.c_uint32), ("idx", C.POINTER(C.c_uint32)),
                ("nnodes", C.c_uint32), ("nodes", C.POINTER(_CNode))]


class _Blocks(C.Structure):
    _fields_ = [("nnear", C.c_uint32), ("nfar", C.c_uint32),
                ("near", C.POINTER(C.c_uint32)),
                ("far", C.POINTER(C.c_uint32))]


_bem = None


def bem():
    global _bem
    if _bem is None:
        hip()  # libh2hip.so first: libh2bem.so resolves its symbols there
        if not os.path.exists(BEM_SO):
            raise H2HipError("libh2bem.so is not built")
        lib = C.CDLL(BEM_SO, mode=C.RTLD_LOCAL)
        sp = C.POINTER(_Surface)
        lib.h2b_sphere_surface.restype = sp
        lib.h2b_sphere_surface.argtypes = [C.c_uint32]
        lib.h2b_cube_surface.restype = sp
        lib.h2b_cube_surface.argtypes = [C.c_uint32]
        lib.h2b_cylinder_surface.restype = sp
        lib.h2b_cylinder_surface.argtypes = [C.c_uint32]
        lib.h2b_surface_from_arrays.restype = sp
        lib.h2b_surface_from_arrays.argtypes = [C.c_uint32, C.c_void_p,
                                                C.c_uint32, C.c_void_p]
        lib.h2b_surface_transform.argtypes = [sp, C.c_void_p, C.c_double,
                                              C.c_void_p]
        lib.h2b_surface_free.argtypes = [sp]
        tp = C.POINTER(_CTree)
        lib.h2b_build_cluster.restype = tp
        lib.h2b_build_cluster.argtypes = [sp, C.c_int, C.c_uint32]
        lib.h2b_ctree_free.argtypes = [tp]
        bp = C.POINTER(_Blocks)
        lib.h2b_build_strict_block.restype = bp
        lib.h2b_build_strict_block.argtypes = [tp, tp, C.c_double, C.c_int]
        lib.h2b_blocks_free.argtypes = [bp]
        lib.h2b_near_block_list.argtypes = [bp, tp, tp, C.c_void_p]
        _bem = lib
    return _bem


def _copy(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Host mesh: x (V, 3), t (T, 3), n (T, 3), g (T,), hmin, hmax."""

    def __init__(self, sp):
        if not sp:
            raise H2HipError("mesh construction failed")
        s = sp.contents
        self.x = _copy(s.x, 3 * s.vertices, np.float64).reshape(-1, 3)
        self.t = _copy(s.t, 3 * s.triangles, np.uint32).reshape(-1, 3)
        self.n = _copy(s.n, 3 * s.triangles, np.float64).reshape(-1, 3)
        self.g = _copy(s.g, s.triangles, np.float64)
        self.hmin, self.hmax = s.hmin, s.hmax
        self._sp = sp

    @property
    def vertices(self):
        return len(self.x)

    @property
    def triangles(self):
        return len(self.t)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        sp = getattr(self, "_sp", None)
        if sp:
            bem().h2b_surface_free(sp)
            self._sp = None

    def transform(self, R, scale=1.0, shift=None):
        R = np.ascontiguousarray(R, dtype=np.float64)
        sh = None if shift is None else np.ascontiguousarray(shift, np.float64)
        bem().h2b_surface_transform(self._sp, R.ctypes.data_as(C.c_void_p),
                                    scale, None if sh is None else
                                    sh.ctypes.data_as(C.c_void_p))
        self.__init__(self._sp)
        return self


def sphere(split):
    """Octahedral unit sphere refined to 8 split^2 triangles."""
    return Mesh(bem().h2b_sphere_surface(split))


def cube(split):
    """Cube [-1, 1]^3 refined to 12 split^2 triangles."""
    return Mesh(bem().h2b_cube_surface(split))


def cylinder(split):
    """Closed cylinder of radius 1, |x| <= 4, refined to 40 split^2
    triangles."""
    return Mesh(bem().h2b_cylinder_surface(split))


def mesh_from_arrays(x, t):
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    return Mesh(bem().h2b_surface_from_arrays(
        len(x), x.ctypes.data_as(C.c_void_p), len(t),
        t.ctypes.data_as(C.c_void_p)))


class ClusterTree:
    """Adaptive cluster tree; nodes in depth-first order, node 0 the root."""

    def __init__(self, mesh, basis, clf):
        tp = bem().h2b_build_cluster(mesh._sp, basis, clf)
        T = tp.contents
        self.n = T.n
        self.idx = _copy(T.idx, T.n, np.uint32)
        nodes = np.frombuffer(
            C.string_at(T.nodes, T.nnodes * C.sizeof(_CNode)),
            dtype=_CNODE_DTYPE)
        self.off = nodes["off"].copy()
        self.size = nodes["size"].copy()
        self.sons = nodes["sons"].copy()
        self.son = nodes["son"].copy()
        self.level = nodes["level"].copy()
        self.bmin = nodes["bmin"].copy()
        self.bmax = nodes["bmax"].copy()
        self._tp = tp

    @property
    def nnodes(self):
        return len(self.off)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        tp = getattr(self, "_tp", None)
        if tp:
            bem().h2b_ctree_free(tp)
            self._tp = None


class BlockTree:
    """Leaves of the strict block tree: near (inadmissible) and far
    (admissible) pairs of cluster nodes, in depth-first order."""

    def __init__(self, rt, ct, eta, admis=0):
        bp = bem().h2b_build_strict_block(rt._tp, ct._tp, eta, admis)
        b = bp.contents
        self.near = (np.ctypeslib.as_array(b.near, shape=(2 * b.nnear,))
                     .reshape(-1, 2).copy() if b.nnear else
                     np.zeros((0, 2), np.uint32))
        self.far = (np.ctypeslib.as_array(b.far, shape=(2 * b.nfar,))
                    .reshape(-1, 2).copy() if b.nfar else
                    np.zeros((0, 2), np.uint32))
        self.rt, self.ct = rt, ct
        self._bp = bp

    def near_block_list(self):
        """(row offset, rows, column offset, cols) per near leaf."""
        r, c = self.near[:, 0], self.near[:, 1]
        return np.stack([self.rt.off[r], self.rt.size[r], self.ct.off[c],
                         self.ct.size[c]], axis=1).astype(np.uint32)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        bp = getattr(self, "_bp", None)
        if bp:
            bem().h2b_blocks_free(bp)
            self._bp = None

"""Many meshes in one launch set.

random_meshes draws meshes from the reference's macrosurfaces (sphere,
cube, cylinder: Library/macrosurface3d.c:81, 390, 432) with a triangle count
log-uniform in a range, a rotation uniform on SO(3) and a scale uniform in
[0.5, 2].  MultiMesh joins their surfaces into one (vertex numbers shifted,
so no two meshes share a vertex) and builds one near-field plan over the
near-field blocks of all of them: one regular launch and one launch per
singular rule class assemble every block of every mesh.  Each block is the
block its mesh alone would give (same tiles, same arithmetic), so the
entries are those of per-mesh plans (bit for bit for the constant basis;
the linear basis sums element matrices with atomic adds in varying order).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import geometry as G
from .bem3d import NearfieldPlan, Surface3d
from .h2matrix import H2Matrix, Interpolation

MACRO_TRIANGLES = {"sphere": 8, "cube": 12, "cylinder": 40}
KINDS = ("sphere", "cube", "cylinder")


def random_rotation(rng):
    """A rotation matrix uniform on SO(3): the unit quaternion of four
    independent normal variates (Shoemake's method)."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


class MeshSpec:
    """One drawn mesh: macrosurface kind, split, scale, rotation and the
    transformed mesh."""

    def __init__(self, kind, split, scale, R, mesh):
        self.kind, self.split, self.scale, self.R = kind, split, scale, R
        self.mesh = mesh

    @property
    def triangles(self):
        return self.mesh.triangles


def random_meshes(count, seed, tmin=512, tmax=32768):
    """count meshes drawn by seed: kind uniform over sphere / cube /
    cylinder, triangle count T log-uniform in [tmin, tmax] realised by the
    split s = round(sqrt(T / macro triangles)) (at least 1), rotation
    uniform on SO(3), scale uniform in [0.5, 2]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        kind = KINDS[int(rng.integers(0, 3))]
        T = np.exp(rng.uniform(np.log(tmin), np.log(tmax)))
        split = max(1, int(round(np.sqrt(T / MACRO_TRIANGLES[kind]))))
        R = random_rotation(rng)
        scale = float(rng.uniform(0.5, 2.0))
        mesh = getattr(G, kind)(split)
        mesh.transform(R, scale)
        out.append(MeshSpec(kind, split, scale, R, mesh))
    return out


def _host_threads():
    """Host threads for per-mesh work: the CPUs this process may use, at
    most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class MultiMesh:
    """The near-field blocks of many meshes as one plan on one joined
    surface.  basis 0 (constant) or 1 (linear) on both sides; clfs: leaf
    size per mesh; eta: admissibility of the strict block trees."""

    def __init__(self, meshes, clfs, basis=0, eta=2.0):
        self.meshes = [m.mesh if isinstance(m, MeshSpec) else m
                       for m in meshes]
        self.basis = basis
        self.clfs = [int(c) for c in clfs]
        xs, ts, ns, gs, idx, blocks = [], [], [], [], [], []
        self.trees, self.btrees, self.dof_off, self.block_range = \
            [], [], [], []
        voff = doff = nblk = 0

        # cluster and block trees of the meshes on host threads (the host
        # library's calls release the interpreter lock); joined in mesh order
        def trees(i):
            tree = G.ClusterTree(self.meshes[i], basis, self.clfs[i])
            bt = G.BlockTree(tree, tree, eta, 0)
            return tree, bt

        with ThreadPoolExecutor(_host_threads()) as ex:
            built = list(ex.map(trees, range(len(self.meshes))))
        for mesh, (tree, bt) in zip(self.meshes, built):
            nb = np.array(bt.near_block_list(), dtype=np.int64).reshape(-1, 4)
            xs.append(mesh.x)
            ts.append(mesh.t.astype(np.int64) + voff)
            ns.append(mesh.n)
            gs.append(mesh.g)
            idx.append(tree.idx.astype(np.int64) + doff)
            shifted = nb.copy()
            shifted[:, 0] += doff
            shifted[:, 2] += doff
            blocks.append(shifted)
            self.trees.append(tree)
            self.btrees.append(bt)
            self.dof_off.append(doff)
            self.block_range.append((nblk, nblk + len(nb)))
            voff += mesh.vertices
            doff += mesh.vertices if basis else mesh.triangles
            nblk += len(nb)
        self.ndof = doff
        self.x = np.ascontiguousarray(np.vstack(xs))
        self.t = np.ascontiguousarray(np.vstack(ts), dtype=np.uint32)
        self.gr = Surface3d(self.x, self.t, np.vstack(ns), np.concatenate(gs))
        self.idx = np.concatenate(idx).astype(np.uint32)
        self.blocks = np.vstack(blocks).astype(np.uint32)
        self.plan = NearfieldPlan(self.gr, basis, basis, self.idx, self.idx,
                                  self.blocks)
        self.entries = self.plan.entries
        off = np.append(self.plan.offsets.astype(np.int64), self.entries)
        # packed entries of mesh i: [entry_range[i][0], entry_range[i][1])
        self.entry_range = [(int(off[a]), int(off[b]))
                            for a, b in self.block_range]

    def execute(self, op, d_out, stream=None):
        """All near-field blocks of all meshes into the packed device
        buffer at d_out (op must live on self.gr)."""
        self.plan.execute(op, d_out, stream)

    def h2matrices(self, d_out, m, dlp=False, schedule=2):
        """Device H2-matrices of the meshes by interpolation of order m
        (couplings of the Laplace kernel), their near fields copied from the
        packed output d_out of execute (device to device); products with
        the transform schedule `schedule` (2 flat, measured fastest; 1
        where a tree does not allow it)."""
        from ._lib import H2HipError
        es = 8
        mats = []

        # the interpolation bases and couplings on a few host threads (each
        # fill runs OpenMP loops of its own), the device matrices in mesh
        # order
        def interp(i):
            return Interpolation(self.meshes[i], self.trees[i],
                                 self.trees[i], self.btrees[i], self.basis,
                                 self.basis, dlp, m, 2)

        with ThreadPoolExecutor(max(1, _host_threads() // 4)) as ex:
            interps = list(ex.map(interp, range(len(self.meshes))))
        for i, mesh in enumerate(self.meshes):
            tree, I = self.trees[i], interps[i]
            H = H2Matrix.from_arrays(I, None, rt=tree)
            a, b = self.entry_range[i]
            H.copy_nearfield_from_device(d_out.ptr + es * a, b - a)
            try:
                H.set_schedule(schedule)
            except H2HipError:
                H.set_schedule(1)
            mats.append(H)
        return mats

"""Serving boundary-element solves with continuous batching.

Requests (a mesh, Laplace or Helmholtz, the wave number, one to four unit
point sources) arrive over time.  Assembled operators stay on the device in
an operator cache; a request is admitted when its operators are cached, or
when they fit into the device-memory budget after evicting idle ones (least
recently used first), and when a slot of its operators' batched solver is
free.  Each cached operator pair owns a stream and a batched solver
(h2hip_gmres_batch, up to 8 solves sharing every product: CG for Laplace,
GMRES for Helmholtz); chunks of
different operators are launched without waiting, so solves on different
meshes overlap on the device, and a request enters between two chunks of
the solves already running.  New operators are built by worker threads
while the loop keeps serving, smallest first.

Problems (the reference's Dirichlet tests, Tests/test_laplacebem3d.c and
Tests/test_helmholtzbem3d.c):
* Laplace: interior Dirichlet problem, piecewise linear basis, data of unit
  point sources outside the surface, V x = (K + M/2) f;
* Helmholtz: exterior Dirichlet problem, piecewise constant basis, data of
  unit point sources inside the surface, in the Brakhage-Werner form of the
  reference's Helmholtz test: (KM - i k V) x = b with KM the double layer
  operator with alpha = 1/2 and b the integrals of the Dirichlet data
  (Tests/test_helmholtzbem3d.c:24-65, 353-450).  Unlike V x = (K - M/2) f
  it has no resonances (V is singular at the interior Dirichlet
  eigenvalues, of which the unit ball has several below 6);
* Neumann: exterior Neumann problem of the Laplace equation, piecewise
  linear basis, data of unit point sources inside the surface, in the
  indirect form with the adjoint double layer (K' - M/2) sigma = b, b the
  integrals of the Neumann data (GMRES; LaplaceNeumannExterior).
"""
import threading
import time
from collections import OrderedDict, deque

import numpy as np

from . import geometry as G
from .device import DeviceArray, Stream
from .solve import (GmresBatch, HelmholtzBrakhageWerner, LaplaceDirichlet,
                    LaplaceNeumannExterior)

LAPLACE, HELMHOLTZ, NEUMANN = "laplace", "helmholtz", "neumann"
KINDS = (LAPLACE, HELMHOLTZ, NEUMANN)


class Request:
    def __init__(self, rid, mesh_id, kind, k, sources, arrival):
        self.rid, self.mesh_id, self.kind, self.k = rid, mesh_id, kind, k
        self.sources = [np.asarray(s, dtype=np.float64) for s in sources]
        self.arrival = arrival
        self.admitted = self.done = None
        self.steps = 0
        self.residual = 0.0
        self.x = None

    @property
    def key(self):
        return (self.mesh_id, self.kind, float(self.k))


def footprint(mesh, kind, clf, m, slots, kmax, eta=2.0):
    """Device bytes of one operator pair (single and double layer
    H2-matrices) and its batched GMRES, from the trees and block lists
    (computed on the host before anything is built)."""
    basis = 0 if kind == HELMHOLTZ else 1
    es = 16 if kind == HELMHOLTZ else 8
    tree = G.ClusterTree(mesh, basis, clf)
    bt = G.BlockTree(tree, tree, eta, 0)
    nb = np.asarray(bt.near_block_list(), dtype=np.int64).reshape(-1, 4)
    near = int((nb[:, 1] * nb[:, 3]).sum())
    k = m ** 3
    n = tree.n
    leaf = tree.sons == 0
    # the flat transforms' blocks: each leaf's rows of all its ancestors'
    # bases, one orientation per side
    flat = 2 * 8 * k * int((tree.size[leaf].astype(np.int64) *
                            (tree.level[leaf].astype(np.int64) + 1)).sum())
    op = es * near + es * len(bt.far) * k * k + 8 * (4 * n * k +
                                                       4 * tree.nnodes * k * k)
    op += flat
    from ._lib import get_reproducible
    if get_reproducible():
        # partial sums of the fixed-order products for a full batch: a slot
        # per block (its larger side) and per flat upward item
        slots_len = len(bt.far) * k + int(np.maximum(nb[:, 1],
                                                     nb[:, 3]).sum())
        items = int((tree.level[leaf].astype(np.int64) + 1).sum())
        op += es * 8 * (slots_len + items * k)
    gm = slots * es * (n * (kmax + 2))
    return 2 * op + gm


class Entry:
    """One cached operator pair with its stream and batched GMRES; op is
    the operator the requests solve with (V, or KM - i k V)."""

    def __init__(self, key, prob, op, batch, stream, nbytes, build_s):
        self.key, self.prob, self.op, self.batch, self.stream = \
            key, prob, op, batch, stream
        self.bytes, self.build_s = nbytes, build_s
        self.running = {}  # rid -> (request, f, b, x)


class Server:
    """The request loop; meshes: {mesh_id: Mesh}."""

    def __init__(self, meshes, budget_bytes, slots=8, kmax=30, eps=1e-8,
                 maxiter=1000, clf=64, m_laplace=4, m_helmholtz=4,
                 laplace_method="cg", computed_couplings=True,
                 max_builds=1):
        """laplace_method: batched solver of the Laplace requests, "cg"
        (V is symmetric positive definite; the reference Laplace test's
        solver) or "gmres"; Helmholtz requests use GMRES(kmax).
        computed_couplings: the Laplace products compute their couplings
        from the interpolation points instead of reading them (measured
        faster for the Laplace kernel; Helmholtz operators read them: their
        kernel adds a sine and a cosine per entry)."""
        self.meshes = meshes
        self.laplace_method = laplace_method
        self.computed_couplings = computed_couplings
        self.budget = int(budget_bytes)
        self.slots, self.kmax, self.eps, self.maxiter = slots, kmax, eps, \
            maxiter
        self.clf = clf
        self.m = {LAPLACE: m_laplace, HELMHOLTZ: m_helmholtz,
                  NEUMANN: m_laplace}
        self.cache = OrderedDict()  # key -> Entry, least recently used first
        self.used = 0
        self.builds = self.evictions = 0
        self._fp = {}
        # operator builds in flight: key -> (thread, result dict); up to
        # max_builds at once (measured with two: the builds' near-field
        # launches slow the running solves, cold-cache p50 96 -> 172 ms with
        # the same p99, so one is the default)
        self._builds = {}
        self.max_builds = max_builds

    # ---- operators -------------------------------------------------------
    def _footprint(self, key):
        if key not in self._fp:
            mesh_id, kind, _ = key
            self._fp[key] = footprint(self.meshes[mesh_id], kind, self.clf,
                                      self.m[kind], self.slots, self.kmax)
        return self._fp[key]

    def _make(self, key, out):
        mesh_id, kind, k = key
        t0 = time.perf_counter()
        stream = Stream()
        mesh = self.meshes[mesh_id]
        if kind == LAPLACE:
            prob = LaplaceDirichlet(mesh, clf=self.clf, m=self.m[kind])
            op, method = prob.V, self.laplace_method
            if self.computed_couplings:
                prob.V.set_coupling_mode(1)
        elif kind == NEUMANN:
            prob = LaplaceNeumannExterior(mesh, clf=self.clf, m=self.m[kind],
                                          stream=stream)
            op, method = prob.A, "gmres"
            if self.computed_couplings:
                prob.A.set_coupling_mode(1)
        else:
            prob = HelmholtzBrakhageWerner(mesh, k, clf=self.clf,
                                           eta_adm=2.0, m=self.m[kind],
                                           stream=stream)
            stream.sync()
            op, method = prob.A, "gmres"
        batch = GmresBatch(op, slots=self.slots, kmax=self.kmax,
                           stream=stream, method=method)
        out["entry"] = Entry(key, prob, op, batch, stream,
                             self._footprint(key), time.perf_counter() - t0)

    def _evict_for(self, need):
        """Evicts idle entries (LRU first) until need more bytes fit; False
        if that is impossible now."""
        if self.used + need <= self.budget:
            return True
        idle = [k for k, e in self.cache.items()
                if not e.running and not e.batch.inflight]
        freeable = sum(self.cache[k].bytes for k in idle)
        if self.used - freeable + need > self.budget:
            return False
        for k in idle:
            if self.used + need <= self.budget:
                break
            e = self.cache.pop(k)
            self.used -= e.bytes
            self.evictions += 1
            del e
        return True

    def _start_build(self, key):
        need = self._footprint(key)
        if need > self.budget or not self._evict_for(need):
            return False
        self.used += need  # reserved while building
        out = {}
        th = threading.Thread(target=self._make, args=(key, out))
        th.start()
        self._builds[key] = (th, out)
        return True

    def _finish_builds(self, wait=False):
        for key, (th, out) in list(self._builds.items()):
            if th.is_alive() and not wait:
                continue
            th.join()
            del self._builds[key]
            if "entry" not in out:
                self.used -= self._footprint(key)
                raise RuntimeError("operator build failed for %r" % (key,))
            self.cache[key] = out["entry"]
            self.builds += 1

    # ---- the loop ----------------------------------------------------------
    def run(self, requests, timeout=600.0):
        """Serves the requests (sorted by arrival time, seconds from the
        start); returns them with admitted / done times, steps, residual
        estimates and the solutions on the host."""
        todo = deque(sorted(requests, key=lambda r: r.arrival))
        queue = []
        ndone = 0
        t0 = time.perf_counter()
        while ndone < len(requests):
            now = time.perf_counter() - t0
            if now > timeout:
                raise RuntimeError("serving did not finish in %g s" % timeout)
            while todo and todo[0].arrival <= now:
                queue.append(todo.popleft())
            self._finish_builds()
            progressed = False
            # completions
            for e in list(self.cache.values()):
                if not e.batch.inflight:
                    continue
                fin = e.batch.poll()
                if fin is None:
                    continue
                progressed = True
                t = time.perf_counter() - t0
                for rid, steps, res in fin:
                    req, f, b, x = e.running.pop(rid)
                    req.done, req.steps, req.residual = t, steps, res
                    req.x = x.download()
                    ndone += 1
            # the next operator to build: the smallest one that a queued
            # request waits for (shortest job first keeps small meshes
            # from waiting behind large ones)
            if len(self._builds) < self.max_builds:
                missing = {r.key for r in queue if r.key not in self.cache
                           and r.key not in self._builds}
                for key in sorted(missing, key=self._footprint):
                    if len(self._builds) >= self.max_builds:
                        break
                    self._start_build(key)
            # admissions, in arrival order
            for req in list(queue):
                e = self.cache.get(req.key)
                if e is None:
                    continue
                if e.batch.inflight or e.batch.free_slots() == 0:
                    continue
                self.cache.move_to_end(req.key)
                f, b = e.prob.rhs_sources(req.sources, e.stream)
                x = DeviceArray(e.prob.n, b.dtype)
                x.zero(e.stream)
                e.batch.submit(b, x, self.eps, self.maxiter, req.rid)
                e.running[req.rid] = (req, f, b, x)
                req.admitted = time.perf_counter() - t0
                queue.remove(req)
                progressed = True
            # next chunks
            for e in self.cache.values():
                if not e.batch.inflight and e.batch.active:
                    e.batch.enqueue()
                    progressed = True
            if not progressed:
                busy = any(e.batch.inflight for e in self.cache.values())
                if not busy and not self._builds and not queue and todo:
                    time.sleep(max(0.0, todo[0].arrival -
                                   (time.perf_counter() - t0)))
                elif not busy and self._builds:
                    next(iter(self._builds.values()))[0].join(0.002)
                else:
                    time.sleep(0.0002)
        self._finish_builds(wait=True)
        return requests


K_GRID = tuple(1.0 + 0.5 * i for i in range(11))  # 1, 1.5, ..., 6


def draw_requests(count, rate, mesh_ids, bounds, seed, ks=K_GRID,
                  kinds=(LAPLACE, HELMHOLTZ)):
    """Poisson arrivals at `rate` per second; each request picks a mesh,
    a problem from kinds (even odds; Laplace Dirichlet, Helmholtz
    Dirichlet, Laplace Neumann), a wave number uniformly from ks for
    Helmholtz (by default [1, 6] in steps of 1/2, so that requests can
    share operators), and one to four unit point sources: outside the
    mesh's bounding sphere for the interior Laplace Dirichlet problem,
    inside half its inscribed radius for the exterior problems.  bounds:
    {mesh_id: (inner, outer)}."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    t -= t[0]
    out = []
    for i in range(count):
        mid = mesh_ids[int(rng.integers(0, len(mesh_ids)))]
        if tuple(kinds) == (LAPLACE, HELMHOLTZ):
            # the draw of the two-problem stream as it always was
            kind = LAPLACE if rng.random() < 0.5 else HELMHOLTZ
        else:
            kind = kinds[int(rng.integers(0, len(kinds)))]
        k = float(ks[int(rng.integers(0, len(ks)))]) if kind == HELMHOLTZ \
            else 0.0
        inner, outer = bounds[mid]
        srcs = []
        for _ in range(int(rng.integers(1, 5))):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            if kind == LAPLACE:
                srcs.append(d * outer * rng.uniform(1.2, 2.5))
            else:
                srcs.append(d * inner * rng.uniform(0.0, 0.5))
        out.append(Request(i, mid, kind, k, srcs, float(t[i])))
    return out


def mesh_bounds(mesh):
    """(inscribed radius about the origin, bounding radius) of a mesh
    centred at the origin: the least distance of a triangle plane and the
    largest vertex norm."""
    dplane = np.abs(np.einsum("ij,ij->i", mesh.n,
                              mesh.x[mesh.t[:, 0]]))
    return float(dplane.min()), float(np.linalg.norm(mesh.x, axis=1).max())

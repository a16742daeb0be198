"""Autotuner of the batched many-mesh pipeline (measured on the device).

Two choices are made by measurement, never by a fixed rule:

* the leaf size (clf) of each mesh's cluster tree.  Small leaves shrink the
  near field (fewer quadrature entries) but deepen the trees and multiply
  the admissible blocks (more coupling matrices to build and to read in
  every H2 product).  For each size class of meshes (triangle count rounded
  to a power of two) the first mesh of the class is measured at every
  candidate leaf size: host time of the tree, the block tree and the
  interpolation bases, device time of its near-field assembly, and device
  time of one H2 product; the candidate with the least
  host + near field + matvecs x product wins for the class;

* the tile and quadrature schedule of a near-field plan: the tile edge
  (64 or 32 DOFs), tiles in plan order or by decreasing work, the regular
  pairs by one workgroup per tile or a persistent grid, and the singular
  (Sauter-Schwab) quadrature either in launches of its own per rule class
  or fused into the tiles' workgroups (the entries do not change beyond the
  order of the sums); the fastest combination is kept.

Results are kept per instance and can be saved to / loaded from JSON.
"""
import json
import math
import time

import numpy as np

from . import geometry as G
from .bem3d import NearfieldPlan, Surface3d, new_slp_laplace_bem3d
from .device import DeviceArray, Event
from .h2matrix import H2Matrix, Interpolation


def _time_ms(stream, fn, reps):
    fn()
    e0, e1 = Event(), Event()
    stream.sync()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    stream.sync()
    return e0.elapsed_ms(e1) / reps


class Autotuner:
    def __init__(self, stream, m=3, candidates=(32, 48, 64, 96, 128),
                 matvecs=50, eta=2.0, reps=3):
        self.stream = stream
        self.m = m
        self.candidates = tuple(int(c) for c in candidates)
        self.matvecs = matvecs
        self.eta = eta
        self.reps = reps
        self.leaf = {}       # size class -> {"clf": c, "costs": {...}}
        self.schedules = {}  # name -> {"best": [sorted, persistent], ...}

    @staticmethod
    def size_class(triangles):
        return int(round(math.log2(max(1, triangles))))

    def measure(self, mesh, clf):
        """Costs of one mesh at leaf size clf (seconds)."""
        t0 = time.perf_counter()
        tree = G.ClusterTree(mesh, 0, clf)
        bt = G.BlockTree(tree, tree, self.eta, 0)
        I = Interpolation(mesh, tree, tree, bt, 0, 0, False, self.m, 2)
        host = time.perf_counter() - t0
        nb = bt.near_block_list()
        gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
        op = new_slp_laplace_bem3d(gr, 2, 4, 0, 0)
        plan = NearfieldPlan(gr, 0, 0, tree.idx, tree.idx, nb)
        out = DeviceArray(max(1, plan.entries))
        near_ms = _time_ms(self.stream, lambda: plan.execute(
            op, out.ptr, self.stream.handle), self.reps)
        H = H2Matrix.from_arrays(I, np.zeros(plan.entries), rt=tree)
        x = DeviceArray.from_host(np.ones(tree.n))
        y = DeviceArray(tree.n)
        y.zero()
        mv_ms = _time_ms(self.stream, lambda: H.addeval(
            1.0, x, y, self.stream), 4 * self.reps)
        total = host + 1e-3 * (near_ms + self.matvecs * mv_ms)
        return {"host_s": host, "near_ms": near_ms, "matvec_ms": mv_ms,
                "near_entries": int(plan.entries), "far_blocks": len(I.far),
                "cost_s": total}

    def leaf_size(self, mesh):
        """The tuned leaf size for this mesh's size class (measured on this
        mesh the first time the class is seen)."""
        c = self.size_class(mesh.triangles)
        if c not in self.leaf:
            costs = {clf: self.measure(mesh, clf) for clf in self.candidates}
            best = min(costs, key=lambda k: costs[k]["cost_s"])
            self.leaf[c] = {"clf": best, "triangles": mesh.triangles,
                            "costs": costs}
        return self.leaf[c]["clf"]

    SCHEDULES = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1))
    TILES = (64, 32)

    def tune_schedule(self, name, plan, op, d_out, reps=None):
        """Times the tile edges and launch schedules of a plan, keeps the
        fastest on the plan and returns it as (tile, sorted, persistent,
        fused)."""
        reps = reps or self.reps
        times = {}
        for tile in self.TILES:
            plan.set_tile(tile)
            for sched in self.SCHEDULES:
                plan.set_schedule(*sched)
                times[(tile,) + sched] = _time_ms(
                    self.stream, lambda: plan.execute(
                        op, d_out, self.stream.handle), reps)
        best = min(times, key=times.get)
        plan.set_tile(best[0])
        plan.set_schedule(*best[1:])
        self.schedules[name] = {
            "best": list(best),
            "fields": ["tile", "sorted", "persistent", "fused"],
            "ms": {"%d-%d%d%d" % k: v for k, v in times.items()}}
        return best

    def to_json(self):
        return {"m": self.m, "candidates": list(self.candidates),
                "matvecs": self.matvecs,
                "leaf": {str(k): {"clf": v["clf"], "triangles":
                                  v["triangles"],
                                  "costs": {str(c): d for c, d in
                                            v["costs"].items()}}
                         for k, v in self.leaf.items()},
                "schedules": self.schedules}

    def save(self, path):
        with open(path, "w") as f:
            json.dump(self.to_json(), f, indent=1)

    def load(self, path):
        with open(path) as f:
            d = json.load(f)
        for k, v in d.get("leaf", {}).items():
            self.leaf[int(k)] = {"clf": int(v["clf"]),
                                 "triangles": v["triangles"],
                                 "costs": {int(c): x for c, x in
                                           v["costs"].items()}}
        self.schedules.update(d.get("schedules", {}))

#!/usr/bin/env python3
"""Benchmark of the near-field Galerkin quadrature on one MI355X.

Workload (the reference's H2 Laplace configuration): unit sphere refined
from the octahedron to 32768 flat triangles (split 64, 16386 vertices),
piecewise linear basis, single and double layer potential (alpha = 1/2),
Gauss orders q = 2 (regular) and q2 = 4 (Sauter-Schwab), FP64.  The blocks
are the near-field leaves of the strict block tree over the adaptive
cluster tree with leaf size 64 and admissibility max-norm, eta = 2.

One step assembles the near field of both operators: every near-field
block, into a packed device buffer, with all inputs already resident in
HBM.  value = near-field entries per second over the timed steps.

The second half of the metric is measured on the same mesh after the
timed steps: the H2 matrices of both operators (interpolation order 4) are
built once and kept on the device, then Dirichlet solves run one after the
other (boundary data -> right-hand side -> CG, the reference Laplace
test's solver -> solution on the host); the line reports solves/s, p50/p99 latency and the relative error
of the solution against the reference library's tightly solved solution
(tests/golden/dirichlet_sphere_s64.npz).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--solves R]
"""
import argparse
import json
import os
import sys
import time

# CPUs this process may run on, read before the OpenMP runtime starts (with
# OMP_PROC_BIND the runtime binds the calling thread to one place, and the
# affinity query would report that one CPU)
AVAIL_CPUS = len(os.sched_getaffinity(0))
# threads of the CPU baseline stay on neighbouring cores (read by the
# OpenMP runtime when it starts)
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("nearfield entries/s; Dirichlet solves/s + p99 latency at "
          "rel.err 1e-6, 1 GPU")
PEAK_FP64_TFLOPS = 78.6
# algorithmic flops per kernel evaluation, following the reference's loop
# body (Library/bem3d.c:1207-1225): Laplace SLP 46, DLP 53; the linear basis
# adds 2 * 9 flops per evaluation for the nine weighted sums
# (Library/bem3d.c:3169-3203) and every entry 2 for the final scaling.
FLOPS_PER_EVAL = {0: 46.0, 1: 53.0}
FLOPS_LL_PER_EVAL = 18.0


# near-field launch schedules (sorted tiles, persistent, fused)
SCHEDULES = {"per_class": (0, 0, 0), "fused": (0, 0, 1),
             "fused_persistent": (0, 1, 1)}
SECTIONS = ("mixed", "helmholtz", "batched", "helmholtz_solve",
            "serving_multi", "dense", "dirichlet")
# the summary line stays well inside the tail of stdout that is kept
LINE_MAX = 1900


def r4(v):
    """v to four significant digits (None stays None)."""
    return None if v is None else float("%.4g" % v)


def compact(d):
    """Section summaries for the last stdout line (the detail file keeps
    everything)."""
    out = {}
    dr = d.get("dirichlet")
    if dr:
        g, c = dr["gmres30"], dr["cg"]
        out.update({
            "dirichlet_solves_per_s": r4(g["solves_per_s"]),
            "p50_ms": r4(g["p50_ms"]), "p99_ms": r4(g["p99_ms"]),
            "rel_err": r4(g["rel_err"]),
            "dirichlet": "GMRES(30) to 1e-10|b|, %d steps" % g["steps"],
            "cg": {"solves_per_s": r4(c["solves_per_s"]),
                   "p99_ms": r4(c["p99_ms"]), "rel_err": r4(c["rel_err"]),
                   "steps": c["steps"]}})
        for key, name in (("serving", "served"),
                          ("serving_saturated", "served_capacity")):
            sv = dr.get(key)
            if sv:
                out[name] = {"solves_per_s": r4(sv["solves_per_s"]),
                             "p50_ms": r4(sv["p50_ms"]),
                             "p99_ms": r4(sv["p99_ms"]),
                             "rel_err": r4(sv.get("rel_err"))}
        mv = dr["matvec"]
        out["matvec"] = {"ms": r4(mv["stored"]["ms"]),
                         "GBps": r4(mv["stored"]["GB/s"]),
                         "frac": r4(mv["stored"]["frac"]),
                         "traffic": mv.get("traffic"),
                         "adjoint_ms": r4(mv["adjoint_ms"]),
                         "adjoint_copies_ms": r4(
                             mv["adjoint_ms_transposed_copies"]),
                         "computed_cpl_ms": r4(
                             (mv["computed_couplings"] or {}).get("ms")),
                         "computed_cpl_frac": r4(
                             (mv["computed_couplings"] or {}).get("frac")),
                         "computed_cpl_adjoint_ms": r4(
                             (mv["computed_couplings"] or {}).get(
                                 "adjoint_ms_V"))}
    mp = d.get("mixed_precision")
    if mp:
        out["mixed"] = {
            m: [r4(mp[m]["entries_per_s"]),
                r4(max(e["max_rel_to_max"] for e in
                       mp[m]["entry_err_vs_fp64"].values()))]
            for m in ("mode1", "mode2")}
    hm = d.get("helmholtz_mixed")
    if hm:
        out["helmholtz"] = {k: [r4(v["fp64_entries_per_s"]),
                                r4(v["mixed_entries_per_s"]),
                                r4(v["entry_err_vs_fp64"]["max_rel_to_max"])]
                            for k, v in hm["k"].items()}
    hs = d.get("helmholtz_solve")
    if hs:
        out["helmholtz_solve"] = {
            k: [v["fp64"]["steps"], r4(v["fp64"]["solve_s"]),
                r4(v["mixed_solution_err_vs_fp64"])]
            for k, v in hs["k"].items()}
    bt = d.get("batched")
    if bt:
        out["batched"] = {"entries_per_s": r4(bt["entries_per_s"]),
                          "graph_ms_per_step": r4(
                              bt["gmres"]["graph"]["ms_per_step_all_meshes"])}
    sm = d.get("serving_multi")
    if sm:
        out["serving_multi"] = {
            name: [r4(sm[name]["solves_per_s"]), r4(sm[name]["p50_ms"]),
                   r4(sm[name]["p99_ms"])] for name in ("cold", "warm")}
    dn = d.get("dense")
    if dn:
        out["dense"] = {"entries_per_s": r4(dn["entries_per_s"]),
                        "matvec_ms": r4(dn.get("matvec_ms"))}
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(split, clf, eta):
    from h2lib_amd import geometry as G
    mesh = G.sphere(split)
    tree = G.ClusterTree(mesh, 1, clf)
    blocks = G.BlockTree(tree, tree, eta, 0)
    return mesh, tree, blocks, blocks.near_block_list()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, tree, nb, entries, args):
    """The reference library, built from its sources for speed (gcc -O3
    -march=x86-64-v3, OpenMP), on the host cores: the single and double
    layer near fields of all blocks (the work of one GPU step), OpenMP over
    blocks with OMP_PROC_BIND=close, and a one-thread figure on every 16th
    block."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import ReferenceBem, have_reference
    fast = have_reference(fast=True)
    if not fast and not have_reference(False):
        return None
    # the box's CPU share for one GPU is 16 cores, whatever the host shows
    avail = AVAIL_CPUS
    threads = max(1, min(args.cpu_threads, avail))
    ops = [ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4, fast=fast),
           ReferenceBem(0, 1, 1, 1, mesh.x, mesh.t, q=2, q2=4, alpha=0.5,
                        fast=fast)]
    t0 = time.perf_counter()
    for op in ops:
        op.nearfield_blocks(tree.idx, tree.idx, nb, threads)
    dt = time.perf_counter() - t0
    dropin = dropin_assembly(mesh, args)
    nb = np.asarray(nb)
    sample = nb[::16]
    ents1 = int(sum(int(b[1]) * int(b[3]) for b in sample)) * len(ops)
    t0 = time.perf_counter()
    for op in ops:
        op.nearfield_blocks(tree.idx, tree.idx, sample, 1)
    dt1 = time.perf_counter() - t0
    return {"value": 2 * entries / dt, "unit": "entries/s", "cores": threads,
            "kind": "reference", "cpu_model": cpu_model(),
            "value_1thread": ents1 / dt1,
            "build": ("gcc -O3 -march=x86-64-v3 -fopenmp" if fast else
                      "gcc -O2 -fopenmp"),
            "sample": ("SLP+DLP near field, all %d blocks, %.1f s; 1 thread: "
                       "every 16th block, %.1f s" % (len(nb), dt, dt1)),
            "dropin_assembly_s": dropin,
            "_detail": {"host_cpus": os.cpu_count(), "affinity_cpus": avail,
                        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                        "entries": 2 * entries, "entries_1thread": ents1}}


def dropin_assembly(mesh, args):
    """The reference's own assemble_bem3d_h2matrix (single layer, linear
    basis, interpolation order 4, clf / eta of the workload) with its CPU
    callbacks, with the GPU callbacks installed by setup_hip_bem3d, and with
    the deferred GPU callbacks and one flush: wall seconds of each (the
    reference built with the binding, -O2, its OpenMP tasks on the host
    cores; the couplings are filled on the host in all three)."""
    from oracle_ffi import ReferenceBem, ReferenceH2, have_reference
    if not have_reference(False, True):
        return None
    out = {}
    for mode in ("cpu", "gpu", "gpu_deferred"):
        ref = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, hip=True, q=2, q2=4)
        if mode != "cpu":
            ref.use_hip(deferred=mode == "gpu_deferred")
        H = ReferenceH2(ref, args.clf, args.eta, 4, 0)
        out[mode] = round(H.assemble_s, 4)
        del H, ref
    return out


def bench_dense(h, stream, reps=20):
    """The dense workload: Laplace SLP, piecewise constant basis, unit sphere
    split 16 (2048 triangles), the full 2048 x 2048 matrix, q=2, q2=4.
    Timed device-resident (one block into a device buffer) and through the
    callback h2hip_nearfield into a host matrix (PCIe copy included)."""
    from h2lib_amd import geometry as G
    mesh = G.sphere(16)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = h.new_slp_laplace_bem3d(gr, 2, 4, 0, 0)
    T = mesh.triangles
    idx = np.arange(T, dtype=np.uint32)
    plan = h.NearfieldPlan(gr, 0, 0, idx, idx,
                           np.array([[0, T, 0, T]], dtype=np.uint32))
    out = h.DeviceArray(plan.entries)
    for _ in range(3):
        plan.execute(op, out.ptr, stream.handle)
    e0, e1 = h.Event(), h.Event()
    stream.sync()
    e0.record(stream)
    for _ in range(reps):
        plan.execute(op, out.ptr, stream.handle)
    e1.record(stream)
    stream.sync()
    ms = e0.elapsed_ms(e1) / reps
    # the product of the assembled matrix with a seeded uniform[-1, 1]
    # vector (addeval_amatrix_avector)
    from h2lib_amd.amatrix import addeval_amatrix
    d_x = h.DeviceArray.from_host(
        np.random.default_rng(16).uniform(-1.0, 1.0, T))
    d_y = h.DeviceArray(T)
    d_y.zero()
    for _ in range(3):
        addeval_amatrix(1.0, out, T, T, T, d_x, d_y, stream=stream)
    e0.record(stream)
    for _ in range(reps):
        addeval_amatrix(1.0, out, T, T, T, d_x, d_y, stream=stream)
    e1.record(stream)
    stream.sync()
    mv_ms = e0.elapsed_ms(e1) / reps
    N = np.zeros((T, T), order="F")
    op.nearfield(None, None, N=N)
    t0 = time.perf_counter()
    for _ in range(reps):
        op.nearfield(None, None, N=N)
    host_s = (time.perf_counter() - t0) / reps
    return {"workload": "Laplace SLP cc, sphere 2048 triangles, dense "
                        "2048 x 2048, q=2 q2=4, FP64",
            "entries_per_s": T * T / (ms * 1e-3), "ms": ms,
            "matvec_ms": mv_ms,
            "matvec_GBps": 8.0 * (T * T + 3 * T) / (mv_ms * 1e-3) / 1e9,
            "entries_per_s_host_matrix": T * T / host_s,
            "host_matrix_ms": host_s * 1e3}


def time_plan(h, ops, plan, outs, stream, steps):
    """Mean milliseconds of one pass of plan.execute over all operators."""
    for op, out in zip(ops, outs):
        plan.execute(op, out.ptr, stream.handle)
    e0, e1 = h.Event(), h.Event()
    stream.sync()
    e0.record(stream)
    for _ in range(steps):
        for op, out in zip(ops, outs):
            plan.execute(op, out.ptr, stream.handle)
    e1.record(stream)
    stream.sync()
    return e0.elapsed_ms(e1) / steps


def entry_errors(mp, f64):
    """Error of mixed-precision entries against the FP64 ones: the largest
    |difference| relative to the largest |entry|, and the Frobenius-norm
    ratio."""
    d = np.abs(mp - f64)
    return {"max_rel_to_max": float(d.max() / np.abs(f64).max()),
            "frobenius_rel": float(np.linalg.norm(mp - f64) /
                                   np.linalg.norm(f64))}


def bench_mixed(h, ops, plan, outs, stream, steps):
    """The same near-field blocks with mixed-precision kernel evaluations,
    timed like the FP64 steps, with the entry error against the FP64
    result: mode 1 (FP32 kernel values and inner sums, FP64 geometry
    differences and outer sums) and mode 2 (FP32 accumulation throughout
    for the regular pairs, outer sums compensated)."""
    res = {}
    for mode in (1, 2):
        mp_outs = [h.DeviceArray(plan.entries, o.dtype) for o in outs]
        for op in ops:
            op.set_precision(mode)
        try:
            ms = time_plan(h, ops, plan, mp_outs, stream, steps)
            errs = [entry_errors(m.download(), o.download())
                    for m, o in zip(mp_outs, outs)]
        finally:
            for op in ops:
                op.set_precision(0)
        res["mode%d" % mode] = {
            "entries_per_s": plan.entries * len(ops) / (ms * 1e-3),
            "ms_per_step": ms,
            "entry_err_vs_fp64": {"slp": errs[0], "dlp": errs[1]}}
    return res


def bench_helmholtz_mixed(h, args, stream):
    """Helmholtz single layer, piecewise constant basis, on the unit sphere
    split 128 (131072 triangles): the near field of the strict block tree
    (leaf size 64, eta 2) in FP64 and in mixed precision at k = 2, 4, 8,
    with entries/s of both and the entry error of mixed against FP64."""
    from h2lib_amd import geometry as G
    mesh = G.sphere(args.helm_split)
    tree = G.ClusterTree(mesh, 0, 64)
    nb = G.BlockTree(tree, tree, 2.0, 0).near_block_list()
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    plan = h.NearfieldPlan(gr, 0, 0, tree.idx, tree.idx, nb)
    out64 = h.DeviceArray(plan.entries, np.complex128)
    out32 = h.DeviceArray(plan.entries, np.complex128)
    res = {"workload": ("Helmholtz SLP cc near field, unit sphere %d "
                        "triangles, clf=64 eta=2, q=2 q2=4" %
                        mesh.triangles),
           "entries": int(plan.entries), "k": {}}
    for k in (2.0, 4.0, 8.0):
        op64 = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
        op32 = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
        op32.set_precision(1)
        ms64 = time_plan(h, [op64], plan, [out64], stream, args.helm_steps)
        ms32 = time_plan(h, [op32], plan, [out32], stream, args.helm_steps)
        e32 = entry_errors(out32.download(), out64.download())
        op32.set_precision(2)
        msc = time_plan(h, [op32], plan, [out32], stream, args.helm_steps)
        res["k"]["%g" % k] = {
            "fp64_entries_per_s": plan.entries / (ms64 * 1e-3),
            "mixed_entries_per_s": plan.entries / (ms32 * 1e-3),
            "compensated_entries_per_s": plan.entries / (msc * 1e-3),
            "fp64_ms": ms64, "mixed_ms": ms32, "compensated_ms": msc,
            "entry_err_vs_fp64": e32,
            "compensated_entry_err_vs_fp64": entry_errors(
                out32.download(), out64.download())}
        log("bench: helmholtz k=%g fp64 %.2f ms mixed %.2f ms compensated "
            "%.2f ms" % (k, ms64, ms32, msc))
    return res


def bench_batched(h, args, stream):
    """Many meshes at once: args.batch_meshes meshes drawn by seed from the
    sphere / cube / cylinder macrosurfaces (triangles log-uniform in
    [512, 32768], rotation uniform on SO(3), scale uniform in [0.5, 2]);
    Laplace SLP, piecewise constant basis, q=2 q2=4, FP64.  The autotuner
    picks each size class's leaf size (candidates 32..128, by measured host
    build + near field + 50 H2 products) and the launch schedule; all
    near-field blocks of all meshes are assembled by one launch set (one
    regular launch, one per singular class), timed over args.steps; then
    every mesh's H2-matrix (interpolation order 3) solves V s = 1 (the
    charge of a unit potential) by GMRES(30) to 1e-8 with one captured HIP
    graph per Arnoldi step of all meshes, and again with direct launches."""
    from h2lib_amd.autotune import Autotuner
    from h2lib_amd.batch import MultiMesh, random_meshes
    from h2lib_amd.solve import GmresMulti
    t0 = time.perf_counter()
    specs = random_meshes(args.batch_meshes, args.batch_seed)
    t_mesh = time.perf_counter() - t0
    tuner = Autotuner(stream, m=3, candidates=(32, 48, 64, 96, 128),
                      matvecs=50)
    t0 = time.perf_counter()
    clfs = [tuner.leaf_size(sp.mesh) for sp in specs]
    t_tune_leaf = time.perf_counter() - t0
    t0 = time.perf_counter()
    mm = MultiMesh(specs, clfs, basis=0)
    t_plan = time.perf_counter() - t0
    op = h.new_slp_laplace_bem3d(mm.gr, 2, 4, 0, 0)
    out = h.DeviceArray(mm.entries)
    best = tuner.tune_schedule("batched near field", mm.plan, op, out.ptr)
    ms = time_plan(h, [op], mm.plan, [out], stream, args.steps)
    evals, pairs = mm.plan.work(op)
    # the same blocks as one plan per mesh, launched one after another
    per_mesh = []
    for i, sp in enumerate(specs):
        m = sp.mesh
        gr = h.Surface3d(m.x, m.t, m.n, m.g)
        o1 = h.new_slp_laplace_bem3d(gr, 2, 4, 0, 0)
        p1 = h.NearfieldPlan(gr, 0, 0, mm.trees[i].idx, mm.trees[i].idx,
                             mm.btrees[i].near_block_list())
        per_mesh.append((gr, o1, p1, h.DeviceArray(p1.entries)))
    e0, e1 = h.Event(), h.Event()
    for _, o1, p1, d1 in per_mesh:
        p1.execute(o1, d1.ptr, stream.handle)
    stream.sync()
    e0.record(stream)
    for _ in range(args.steps):
        for _, o1, p1, d1 in per_mesh:
            p1.execute(o1, d1.ptr, stream.handle)
    e1.record(stream)
    stream.sync()
    ms_sep = e0.elapsed_ms(e1) / args.steps
    del per_mesh

    t0 = time.perf_counter()
    mats = mm.h2matrices(out, 3)
    # transforms of every mesh's products (flat where the tree allows it)
    sched = []
    for H in mats:
        try:
            H.set_schedule(args.batch_schedule)
        except h.H2HipError:
            H.set_schedule(1)
        sched.append(H.level_sync)
    t_h2 = time.perf_counter() - t0
    G = GmresMulti(mats, kmax=30, stream=stream)
    # b_i = integral of 1 over triangle i = g_i / 2 (original DOF order)
    bs = [h.DeviceArray.from_host(0.5 * sp.mesh.g) for sp in specs]
    gm = {}
    for use_graph in (True, False):
        xs = [h.DeviceArray(H.n) for H in mats]
        for x in xs:
            x.zero()
        stream.sync()
        t0 = time.perf_counter()
        it, res = G.solve(bs, xs, 1e-8, args.batch_maxiter,
                          use_graph=use_graph)
        dt = time.perf_counter() - t0
        gm["graph" if use_graph else "direct"] = {
            "s": dt, "steps_max": int(it.max()),
            "steps_median": int(np.median(it)),
            "ms_per_step_all_meshes": 1e3 * dt / max(1, int(it.max()))}
        if use_graph:
            # true relative residuals |b - V s| / |b| per mesh
            rel = []
            for H, b, x in zip(mats, bs, xs):
                r = h.DeviceArray(H.n)
                h.hip().h2hip_memcpy(r._p, b._p, b.nbytes, 2)
                H.addeval(-1.0, x, r, stream)
                stream.sync()
                rel.append(float(np.linalg.norm(r.download()) /
                                 np.linalg.norm(b.download())))
            gm["per_mesh"] = [[sp.kind, sp.triangles, round(sp.scale, 3),
                               c, int(k), float("%.3g" % e)]
                              for sp, c, k, e in zip(specs, clfs, it, rel)]
            gm["per_mesh_fields"] = ["kind", "triangles", "scale", "clf",
                                     "steps", "true_rel_residual"]
            gm["converged"] = int(sum(e <= 1.5e-8 for e in rel))
            gm["note"] = ("meshes that reach the step limit are cylinders: "
                          "their single layer matrices at q=2 q2=4 are "
                          "indefinite (symmetric part has a negative "
                          "eigenvalue; -1.4e-4 against 6.6e-2 at 1960 "
                          "triangles, the same with the reference's "
                          "quadrature), so restarted GMRES stagnates; the "
                          "timing per step is unaffected")
    gm["graph_captures"] = G.graph_builds
    gm["transform_schedules"] = {str(k): sched.count(k) for k in set(sched)}
    gm["speedup_graph_vs_direct"] = gm["direct"]["s"] / gm["graph"]["s"]
    tri = [sp.triangles for sp in specs]
    return {
        "workload": ("%d meshes (seed %d): sphere/cube/cylinder, %d..%d "
                     "triangles (total %d), random rotation and scale; "
                     "Laplace SLP cc, q=2 q2=4, FP64" %
                     (len(specs), args.batch_seed, min(tri), max(tri),
                      sum(tri))),
        "kinds": {k: sum(sp.kind == k for sp in specs)
                  for k in ("sphere", "cube", "cylinder")},
        "leaf_sizes": {str(c): clfs.count(c) for c in sorted(set(clfs))},
        "schedule": {"tile": int(best[0]), "sorted_tiles": int(best[1]),
                     "persistent": int(best[2]), "fused": int(best[3]),
                     "ms": tuner.schedules["batched near field"]["ms"]},
        "entries": int(mm.entries),
        "kernel_evaluations": float(evals.sum()),
        "entries_per_s": mm.entries / (ms * 1e-3), "ms": ms,
        "entries_per_s_one_plan_per_mesh": mm.entries / (ms_sep * 1e-3),
        "ms_one_plan_per_mesh": ms_sep,
        "host_s": {"meshes": t_mesh, "autotune_leaf": t_tune_leaf,
                   "trees_and_plan": t_plan, "h2_bases": t_h2},
        "gmres": gm,
        "autotuner": tuner.to_json(),
    }


def bench_helmholtz_solve(h, args, stream):
    """Exterior Helmholtz problem on the unit sphere split args.helm_split
    (131072 triangles by default), piecewise constant basis, point source
    (0.3, 0.2, 0.1) inside, near fields in FP64 and in mixed precision,
    H2-matrices by interpolation (leaf size 64), complex FP64 couplings.
    k below 6: f = projected Dirichlet data, b = (K - M/2) f, V x = b by
    GMRES(30) to 1e-8 (x: the Neumann data), order 4, eta 2; per k the
    Neumann error of each solution and the mixed solution's error against
    the FP64 one.  k = 8 lies 0.18 below an interior Dirichlet eigenvalue of
    the ball, where V is singular, and order 4 does not resolve e^{ikr}
    across the largest admissible blocks (17.8 % row-sum error on the
    8192-triangle sphere, the same for the reference's H2-matrix): the
    Brakhage-Werner form (KM - i k V) x = b of the reference's Helmholtz
    test with order 5 and eta 1, and the error of its exterior potential
    on the test's grid."""
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import (HelmholtzBrakhageWerner, HelmholtzExterior,
                                 solve_gmres)
    mesh = G.sphere(args.helm_split)
    x0 = np.array([0.3, 0.2, 0.1])
    out = {"workload": ("exterior Helmholtz, unit sphere %d triangles, cc, "
                        "point source (0.3, 0.2, 0.1), GMRES(30) to 1e-8, "
                        "clf 64; k < 6: V x = (K - M/2) f, order 4, eta 2; "
                        "k >= 6: Brakhage-Werner, order 5, eta 1" %
                        mesh.triangles), "k": {}}
    for k in args.helm_k:
        bw = k >= 6.0
        res = {"form": "brakhage-werner" if bw else "single layer"}
        sol = {}
        for prec, name in ((0, "fp64"), (1, "mixed")):
            t0 = time.perf_counter()
            if bw:
                prob = HelmholtzBrakhageWerner(mesh, k, clf=64, eta_adm=1.0,
                                               m=5, precision=prec,
                                               stream=stream)
            else:
                prob = HelmholtzExterior(mesh, k, clf=64, m=4,
                                         precision=prec, stream=stream)
            stream.sync()
            build = time.perf_counter() - t0
            # the products compute their couplings from the interpolation
            # points, each value once for a block and its mirror (0.93 ms
            # per single layer product at k = 2 against 1.30 ms once per
            # block; the Brakhage-Werner operator's double layer too)
            mode = {}
            for nm, H in (("V", prob.V),) + ((("KM", prob.KM),) if bw
                                              else ()):
                try:
                    H.set_coupling_mode(1)
                    mode[nm] = "computed"
                except h.H2HipError as exc:
                    log("bench: stored Helmholtz couplings: %s" % exc)
                    mode[nm] = "stored"
            t0 = time.perf_counter()
            if bw:
                d_b = prob.rhs(x0, stream)
                d_x = h.DeviceArray(prob.n, np.complex128)
                d_x.zero(stream)
                it, r = solve_gmres(prob.A, d_b, d_x, 1e-8,
                                    args.helm_maxiter, 30, stream)
            else:
                _, d_b = prob.rhs(x0, stream)
                d_x, it, r = prob.solve(d_b, eps=1e-8,
                                        maxiter=args.helm_maxiter, kmax=30,
                                        stream=stream)
            x = d_x.download()
            dt = time.perf_counter() - t0
            sol[name] = x
            res[name] = {"solve_s": dt, "steps": int(it),
                         "couplings": mode,
                         "operator_build_s": build}
            if bw:
                res[name]["outer_potential_err"] = prob.outer_error(d_x, x0)
            else:
                res[name]["neumann_rel_err"] = prob.neumann_error(x, x0)
            del prob
        res["mixed_solution_err_vs_fp64"] = float(
            np.linalg.norm(sol["mixed"] - sol["fp64"]) /
            np.linalg.norm(sol["fp64"]))
        out["k"]["%g" % k] = res
        log("bench: helmholtz solve k=%g (%s) fp64 %d steps %.2f s, mixed %d "
            "steps %.2f s, mixed vs fp64 %.2e" %
            (k, res["form"], res["fp64"]["steps"], res["fp64"]["solve_s"],
             res["mixed"]["steps"], res["mixed"]["solve_s"],
             res["mixed_solution_err_vs_fp64"]))
    return out


def bench_serving_multi(h, args):
    """Serving across meshes and problems: requests arrive as a Poisson
    stream (args.multi_rate per second, seeded); each picks a mesh from a
    catalog of sphere / cube / cylinder meshes with 2048 to 131072
    triangles, Laplace (interior Dirichlet, linear basis, sources outside)
    or Helmholtz (exterior Dirichlet, constant basis, sources inside, k
    uniform in [1, 6] in steps of 1/2, Brakhage-Werner form (KM - i k V) x = b,
    free of the resonances of the single layer operator), with one to four
    unit point sources.  Operators (H2, interpolation order 4, clf 64) are
    cached on the device within a budget of args.multi_budget_gb GB and
    built on demand by a worker thread; GMRES(30) to 1e-10 in batches of up
    to 8 per operator, chunks of different operators overlapping.  The same
    request stream is served twice: cold (empty cache) and warm.  Accuracy:
    the first request of each problem is solved again alone to 1e-13 and
    compared."""
    from h2lib_amd import geometry as G
    from h2lib_amd.serve import Server, draw_requests, mesh_bounds
    from h2lib_amd.solve import solve_gmres
    catalog = {"sphere2048": ("sphere", 16), "cube6912": ("cube", 24),
               "cylinder10240": ("cylinder", 16),
               "sphere18432": ("sphere", 48), "cube43200": ("cube", 60),
               "sphere131072": ("sphere", 128)}
    meshes = {mid: getattr(G, kind)(s) for mid, (kind, s) in catalog.items()}
    bounds = {mid: mesh_bounds(m) for mid, m in meshes.items()}
    reqs = draw_requests(args.multi_requests, args.multi_rate,
                         sorted(meshes), bounds, seed=77)
    srv = Server(meshes, args.multi_budget_gb * 1e9, eps=1e-10,
                 maxiter=3000)
    out = {"catalog": {mid: m.triangles for mid, m in meshes.items()},
           "requests": len(reqs), "arrival_rate_per_s": args.multi_rate,
           "budget_gb": args.multi_budget_gb, "gmres_eps": 1e-10}
    for name in ("cold", "warm"):
        for r in reqs:
            r.admitted = r.done = None
        t0 = time.perf_counter()
        srv.run(reqs)
        span = time.perf_counter() - t0
        lat = np.array([r.done - r.arrival for r in reqs])
        out[name] = {"solves_per_s": len(reqs) / span,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "builds": srv.builds, "evictions": srv.evictions,
                     "steps_median": int(np.median([r.steps for r in reqs]))}
        log("bench: serving across meshes (%s): %.1f solves/s, p50 %.0f ms, "
            "p99 %.0f ms" % (name, out[name]["solves_per_s"],
                             out[name]["p50_ms"], out[name]["p99_ms"]))
    # accuracy: first request of each problem, solved alone to 1e-13
    errs = {}
    seen = set()
    for r in reqs:
        if r.key in seen or len(errs) >= args.multi_check:
            continue
        seen.add(r.key)
        e = srv.cache.get(r.key)
        if e is None:
            continue
        f, b = e.prob.rhs_sources(r.sources, e.stream)
        x = h.DeviceArray(e.prob.n, b.dtype)
        x.zero()
        solve_gmres(e.op, b, x, 1e-13, 6000, 30, e.stream)
        xr = x.download()
        errs["%s/%s/k=%g" % r.key] = float(np.linalg.norm(r.x - xr) /
                                           np.linalg.norm(xr))
    out["rel_err_vs_tight_solve"] = errs
    out["per_request"] = [[r.mesh_id, r.kind, r.k, len(r.sources),
                           round(r.arrival, 4), round(r.done - r.arrival, 4),
                           r.steps] for r in reqs]
    out["per_request_fields"] = ["mesh", "kind", "k", "sources", "arrival_s",
                                 "latency_s_warm", "steps"]
    return out


def time_product(h, H, d_x, d_y, stream, reps=50, symm=False):
    """Mean milliseconds of one product y += H x (events on its stream)."""
    f = H.addevalsymm if symm else H.addeval
    for _ in range(3):
        f(1.0, d_x, d_y, stream)
    e0, e1 = h.Event(), h.Event()
    stream.sync()
    e0.record(stream)
    for _ in range(reps):
        f(1.0, d_x, d_y, stream)
    e1.record(stream)
    stream.sync()
    return e0.elapsed_ms(e1) / reps


def product_line(H, ms, n):
    """GB/s and roofline fraction of one product: algorithmic bytes (near
    field, couplings, leaf and transfer matrices of both sides, vector
    traffic) and the bytes the kernels read in the mode timed."""
    alg = H.algorithmic_bytes()
    rd = sum(H.bytes().values()) + 3 * 8 * n
    return {"ms": ms, "algorithmic_bytes": int(alg), "read_bytes": int(rd),
            "GB/s": alg / (ms * 1e-3) / 1e9,
            "frac": alg / (ms * 1e-3) / 8e12}


def bench_dirichlet(h, mesh, args, stream):
    """Dirichlet solves of the interior Laplace problem on the same mesh with
    the operators resident on the device: each solve projects the boundary
    data u = 1/|x - x0|, x0 = (1.2, 1.2, 1.2), forms b = (K + M/2) f and
    solves V x = b by GMRES(30) (the configuration's solver) and by CG (the
    reference Laplace test's solver, its steps replayed as captured graphs);
    the latency of a solve runs from the boundary data on the host to the
    solution on the host."""
    from h2lib_amd.solve import LaplaceDirichlet
    x0 = np.array([1.2, 1.2, 1.2])
    t0 = time.perf_counter()
    prob = LaplaceDirichlet(mesh, clf=args.clf, eta=args.eta, m=4)
    h.hip().h2hip_stream_sync(None)
    build_s = time.perf_counter() - t0

    # products of the single layer operator, events on its stream
    n = prob.n
    d_x = h.DeviceArray.from_host(
        np.random.default_rng(1).uniform(-1.0, 1.0, n))
    d_y = h.DeviceArray(n)
    d_y.zero()
    # the three schedules of the basis transforms, stored couplings; the
    # fastest serves the solves
    mv_by = {}
    for sched in args.schedules:
        for H in (prob.V, prob.KM):
            H.set_schedule(sched)
        mv_by[sched] = time_product(h, prob.V, d_x, d_y, stream)
    best = min(mv_by, key=mv_by.get)
    for H in (prob.V, prob.KM):
        H.set_schedule(best)
    stored = product_line(prob.V, mv_by[best], n)
    # the adjoint product y += V^* x on the same schedule
    for _ in range(3):
        prob.V.addeval(1.0, d_x, d_y, stream, trans=True)
    e0, e1 = h.Event(), h.Event()
    stream.sync()
    e0.record(stream)
    for _ in range(50):
        prob.V.addeval(1.0, d_x, d_y, stream, trans=True)
    e1.record(stream)
    stream.sync()
    trans_ms = e0.elapsed_ms(e1) / 50
    # the same with the conjugate-transposed copies of the blocks kept (the
    # adjoint then runs the forward kernels)
    prob.V.set_adjoint_storage(True)
    for _ in range(3):
        prob.V.addeval(1.0, d_x, d_y, stream, trans=True)
    e0, e1 = h.Event(), h.Event()
    stream.sync()
    e0.record(stream)
    for _ in range(50):
        prob.V.addeval(1.0, d_x, d_y, stream, trans=True)
    e1.record(stream)
    stream.sync()
    trans_copies_ms = e0.elapsed_ms(e1) / 50
    prob.V.set_adjoint_storage(False)
    # the same product with the couplings computed from the interpolation
    # points instead of read
    computed = None
    try:
        prob.V.set_coupling_mode(1)
    except h.H2HipError as exc:  # host-built couplings: no points kept
        log("bench: computed couplings unavailable: %s" % exc)
    else:
        computed = product_line(prob.V,
                                time_product(h, prob.V, d_x, d_y, stream),
                                n)
        # the adjoint with computed couplings (the double layer operator,
        # whose adjoint is not itself)
        for H in (prob.V, prob.KM):
            if H is prob.KM:
                H.set_coupling_mode(1)
            for _ in range(3):
                H.addeval(1.0, d_x, d_y, stream, trans=True)
            e0, e1 = h.Event(), h.Event()
            stream.sync()
            e0.record(stream)
            for _ in range(50):
                H.addeval(1.0, d_x, d_y, stream, trans=True)
            e1.record(stream)
            stream.sync()
            computed["adjoint_ms_" + ("V" if H is prob.V else "KM")] = \
                e0.elapsed_ms(e1) / 50
            if H is prob.KM:
                computed["ms_KM"] = time_product(h, prob.KM, d_x, d_y,
                                                 stream)
                H.set_coupling_mode(0)
        prob.V.set_coupling_mode(0)
    # the faster of the two serves the solves (both operators)
    use_computed = computed is not None and computed["ms"] < stored["ms"]
    modes = {"V": "stored", "KM": "stored"}
    if use_computed:
        for name, H in (("V", prob.V), ("KM", prob.KM)):
            try:
                H.set_coupling_mode(1)
                modes[name] = "computed"
            except h.H2HipError as exc:
                log("bench: %s keeps stored couplings: %s" % (name, exc))
    # the single layer matrix in symmetric storage (blocks on and below the
    # block diagonal) and its self-adjoint product
    from h2lib_amd.h2matrix import H2Matrix
    Vs = H2Matrix(prob.slp, mesh, prob.tree, prob.tree, prob.blocks, 4,
                  symmetric=True, stream=stream)
    Vs.set_schedule(best)
    stream.sync()
    symm = product_line(Vs, time_product(h, Vs, d_x, d_y, stream,
                                         symm=True), n)
    symm["device_bytes"] = Vs.storage()["device"]
    symm["device_bytes_full"] = prob.V.storage()["device"]

    # eight vectors per product (the serving path), stored couplings: the
    # coupling blocks on the vector units and on FP64 MFMA
    rng8 = np.random.default_rng(8)
    xs8 = [h.DeviceArray.from_host(rng8.uniform(-1.0, 1.0, n))
           for _ in range(8)]
    ys8 = [h.DeviceArray(n) for _ in range(8)]
    for y in ys8:
        y.zero()
    batch8 = {}
    prob.V.set_coupling_mode(0)
    for name, on in (("vector", False), ("mfma", True)):
        prob.V.set_mfma(on)
        for _ in range(3):
            prob.V.addeval_batch([1.0] * 8, xs8, ys8, stream)
        e0, e1 = h.Event(), h.Event()
        stream.sync()
        e0.record(stream)
        for _ in range(30):
            prob.V.addeval_batch([1.0] * 8, xs8, ys8, stream)
        e1.record(stream)
        stream.sync()
        batch8[name + "_ms"] = e0.elapsed_ms(e1) / 30
    prob.V.set_mfma(False)
    prob.V.set_coupling_mode(1 if modes["V"] == "computed" else 0)
    del xs8, ys8

    ref = None
    gpath = os.path.join(ROOT, "tests", "golden", "dirichlet_sphere_s64.npz")
    if (args.split, args.clf, args.eta) == (64, 64, 2.0) and \
            os.path.exists(gpath):
        ref = np.load(gpath)

    def rel(x):
        if ref is None:
            return None
        return float(np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"]))

    # solves: boundary data on the host -> solution on the host
    lines = {}
    x = None
    for solver in ("gmres", "cg"):
        lat, steps = [], []
        for r in range(args.solve_warmup + args.solves):
            t = time.perf_counter()
            d_f, d_b = prob.rhs(x0, stream)
            d_sol, it, res = prob.solve(d_b, eps=args.eps, maxiter=1000,
                                        kmax=30, stream=stream,
                                        solver=solver,
                                        use_graph=not args.cg_direct)
            x = d_sol.download()
            dt = time.perf_counter() - t
            if r >= args.solve_warmup:
                lat.append(dt)
                steps.append(it)
        lat = np.array(lat)
        lines[solver] = {
            "solves_per_s": len(lat) / lat.sum(),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "solves": len(lat), "steps": int(np.median(steps)),
            "rel_err": rel(x),
            "neumann_rel_err": prob.neumann_error(x, x0)}

    # solve only (b resident), the symmetric storage with CG as well
    from h2lib_amd.solve import solve_cg
    d_f, d_b = prob.rhs(x0, stream)
    ts, its = [], []
    for r in range(1 + max(1, args.solves // 4)):
        d_c = h.DeviceArray(n)
        d_c.zero(stream)
        t = time.perf_counter()
        it, _ = solve_cg(Vs, d_b, d_c, args.eps, 1000, not args.cg_direct,
                         stream)
        dt = time.perf_counter() - t
        if r:
            ts.append(dt)
            its.append(it)
    cg_symm = {"solve_ms": float(np.median(ts) * 1e3),
               "steps": int(np.median(its)), "rel_err": rel(d_c.download())}

    out = {
        "workload": ("interior Laplace Dirichlet problem, unit sphere %d "
                     "triangles, linear basis, H2 by interpolation m=4 "
                     "clf=%d eta=%g, u = 1/|x - (1.2,1.2,1.2)|, solves to "
                     "%g |b|" % (mesh.triangles, args.clf, args.eta,
                                 args.eps)),
        "gmres30": lines["gmres"], "cg": lines["cg"],
        "cg_symmetric_b_resident": cg_symm,
        "operator_build_s": build_s,
        "couplings_in_solves": modes,
        "matvec": {
            "kernels": ("fwd_flat_near + couple_near + back_flat (Laplace "
                        "SLP)" if best == 2 else
                        "fwd_up (+ fwd_level) + couple_near + down_back "
                        "(Laplace SLP)"),
            "schedule": ("counters", "level_sync", "flat")[best],
            "ms_by_schedule": {name: mv_by[i] for i, name in
                               enumerate(("counters", "level_sync", "flat"))
                               if i in mv_by},
            "stored": stored, "computed_couplings": computed,
            "adjoint_ms": trans_ms,
            "adjoint_ms_transposed_copies": trans_copies_ms,
            "eight_vectors_stored_couplings": batch8,
            "symmetric": symm},
    }
    if ref is not None:
        out["rel_err_reference_eps8"] = float(
            np.linalg.norm(ref["x_eps8"] - ref["x"]) /
            np.linalg.norm(ref["x"]))
        out["rel_err_against"] = ("the reference library's solution of the "
                                  "same discrete problem at residual "
                                  "1e-12 |b|")
    out["_prob"] = prob
    out["_ref"] = ref
    return out


def bench_serving(h, prob, ref, args, rate=None):
    """Continuous batching of Dirichlet solves with the operators resident
    on the device: requests arrive as a Poisson stream (seeded), each with
    the data of a point source x0 outside the unit sphere (the first one at
    x0 = (1.2, 1.2, 1.2), whose solution is checked against the reference);
    its right-hand side is formed on a second stream as soon as it arrives
    and it enters the batched solver (CG, the reference test's solver, or
    GMRES(30): --serve-method) as soon as one of the slots is free.
    Latency: arrival to solution complete on the device.  rate overrides
    the arrival rate (a rate far above the capacity measures the capacity:
    solves/s of a saturated server)."""
    from h2lib_amd.solve import GmresBatch
    rng = np.random.default_rng(2024)
    R = args.serve_requests
    rate = args.serve_rate if rate is None else rate
    arrivals = np.cumsum(rng.exponential(1.0 / rate, R))
    arrivals -= arrivals[0]
    x0s = [np.array([1.2, 1.2, 1.2])]
    for _ in range(R - 1):
        d = rng.standard_normal(3)
        x0s.append(d / np.linalg.norm(d) * rng.uniform(1.5, 3.0))
    run_st, rhs_st = h.Stream(), h.Stream()
    G = GmresBatch(prob.V, slots=args.serve_slots, kmax=30, stream=run_st,
                   method=args.serve_method)
    n = prob.n
    d_x = [h.DeviceArray(n) for _ in range(R)]
    for d in d_x:
        d.zero()
    rhs = [None] * R
    done_t = np.zeros(R)
    steps = np.zeros(R, dtype=int)
    formed, submitted, ndone = 0, 0, 0
    t0 = time.perf_counter()
    while ndone < R:
        now = time.perf_counter() - t0
        # right-hand sides of the requests that have arrived
        while formed < R and arrivals[formed] <= now:
            rhs[formed] = prob.rhs(x0s[formed], rhs_st)
            formed += 1
        # admissions into free slots (in arrival order)
        if submitted < formed and G.free_slots() > 0:
            rhs_st.sync()
            while submitted < formed and G.free_slots() > 0:
                G.submit(rhs[submitted][1], d_x[submitted], args.eps, 1000,
                         submitted)
                submitted += 1
        if submitted > ndone:
            for tag, it, res in G.run():
                done_t[tag] = time.perf_counter() - t0
                steps[tag] = it
                rhs[tag] = None
                ndone += 1
        elif formed < R:
            time.sleep(max(0.0, arrivals[formed] - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    lat = done_t - arrivals
    out = {
        "requests": R, "arrival_rate_per_s": rate,
        "slots": args.serve_slots,
        "solves_per_s": R / makespan,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "method": args.serve_method,
        "steps_median": int(np.median(steps)),
        "eps": args.eps,
    }
    if ref is not None:
        x = d_x[0].download()
        xr = ref["x"]
        out["rel_err"] = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--split", type=int, default=64)
    ap.add_argument("--clf", type=int, default=64)
    ap.add_argument("--eta", type=float, default=2.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--solves", type=int, default=20,
                    help="timed Dirichlet solves (0: skip them)")
    ap.add_argument("--solve-warmup", type=int, default=2)
    ap.add_argument("--eps", type=float, default=1e-10,
                    help="stopping tolerance of the Dirichlet solves "
                    "relative to |b|")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the serving run (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=60.0,
                    help="Poisson arrival rate of the serving run (1/s)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-method", choices=("cg", "gmres"), default="cg",
                    help="batched solver of the serving runs' Laplace "
                    "requests")
    ap.add_argument("--helm-split", type=int, default=128,
                    help="sphere split of the Helmholtz FP64 / mixed "
                    "near-field comparison (0: skip it)")
    ap.add_argument("--helm-steps", type=int, default=5)
    ap.add_argument("--helm-k", type=float, nargs="*", default=[2.0, 4.0, 8.0],
                    help="wave numbers of the Helmholtz solves (none: skip)")
    ap.add_argument("--helm-maxiter", type=int, default=1000)
    ap.add_argument("--batch-meshes", type=int, default=64,
                    help="meshes of the batched run (0: skip it)")
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-schedule", type=int, default=2, choices=(0, 1, 2),
                    help="transform schedule of the batched meshes' H2 "
                    "products (2 flat, 1 level-synchronous, 0 counters)")
    ap.add_argument("--batch-maxiter", type=int, default=300,
                    help="GMRES step limit of the batched solves (the "
                    "cylinder meshes give indefinite matrices at q=2 q2=4 "
                    "and stagnate under restarted GMRES)")
    ap.add_argument("--multi-requests", type=int, default=48,
                    help="requests of the serving run across meshes (0: "
                    "skip it)")
    ap.add_argument("--multi-rate", type=float, default=40.0)
    ap.add_argument("--multi-budget-gb", type=float, default=96.0,
                    help="device-memory budget of the operator cache")
    ap.add_argument("--multi-check", type=int, default=6,
                    help="requests whose solutions are checked against a "
                    "tight solve")
    ap.add_argument("--schedules", type=int, nargs="+", default=[0, 1, 2],
                    choices=(0, 1, 2),
                    help="transform schedules of the H2 product timed "
                    "(0 counter climb, 1 level-synchronous, 2 flat); the "
                    "fastest serves the solves")
    ap.add_argument("--cg-direct", action="store_true",
                    help="CG timings with direct launches instead of "
                         "graph-replayed chunks")
    ap.add_argument("--sections", nargs="*", default=list(SECTIONS),
                    choices=SECTIONS,
                    help="sections to run besides the headline near field "
                    "(one section per profiler run keeps kernel averages "
                    "apart)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the CPU baseline (at most the CPUs "
                    "this process may run on)")
    ap.add_argument("--detail", default=os.path.join("gpurun_out",
                                                     "bench_detail.json"),
                    help="file for the full results of every section (the "
                    "last stdout line is the compact summary)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run device-to-device copies of this size "
                    "(reference for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    sections = set(args.sections)

    import h2lib_amd as h

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        a, b = h.DeviceArray(n), h.DeviceArray(n)
        a.zero()
        for _ in range(3):
            h.hip().h2hip_memcpy(b._p, a._p, a.nbytes, 2)
        log("bench: copied %d MiB three times" % args.calibrate_copy_mib)
        return

    log("bench: building sphere split %d, cluster tree clf %d, eta %g" %
        (args.split, args.clf, args.eta))
    t0 = time.perf_counter()
    mesh, tree, btree, nb = build_workload(args.split, args.clf, args.eta)
    log("bench: %d triangles, %d vertices, %d near blocks (%.2f s host)" %
        (mesh.triangles, mesh.vertices, len(nb), time.perf_counter() - t0))

    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    ops = [h.new_slp_laplace_bem3d(gr, 2, 4, 1, 1),
           h.new_dlp_laplace_bem3d(gr, 2, 4, 1, 1, 0.5)]
    plan = h.NearfieldPlan(gr, 1, 1, tree.idx, tree.idx, nb)
    outs = [h.DeviceArray(plan.entries, np.float64) for _ in ops]
    stream = h.Stream()
    entries_per_step = plan.entries * len(ops)

    # algorithmic work per launch
    flops = []
    for k, op in enumerate(ops):
        evals, pairs = plan.work(op)
        flops.append(evals.sum() * (FLOPS_PER_EVAL[k] + FLOPS_LL_PER_EVAL) +
                     2.0 * plan.entries)
    work = plan.work(ops[0])

    # launch schedule of the plan, chosen by a short measurement before the
    # timed region: launches per rule class, or fused tiles (each tile's
    # workgroup also evaluates its pairs with common vertices) with one
    # workgroup per tile or a persistent grid
    # (the device is brought to its working clocks first: the first
    # launches of a run are slower and would penalise the first schedule)
    sched_ms = {}
    for _ in range(max(5, args.warmup)):
        for op, out in zip(ops, outs):
            plan.execute(op, out.ptr, stream.handle)
    for name, sched in SCHEDULES.items():
        plan.set_schedule(*sched)
        for op, out in zip(ops, outs):
            plan.execute(op, out.ptr, stream.handle)
        e0, e1 = h.Event(), h.Event()
        stream.sync()
        e0.record(stream)
        for _ in range(5):
            for op, out in zip(ops, outs):
                plan.execute(op, out.ptr, stream.handle)
        e1.record(stream)
        stream.sync()
        sched_ms[name] = e0.elapsed_ms(e1) / 5
    schedule = min(sched_ms, key=sched_ms.get)
    plan.set_schedule(*SCHEDULES[schedule])
    log("bench: near-field schedules (ms per step) %s, using %s" %
        (", ".join("%s %.3f" % kv for kv in sched_ms.items()), schedule))

    for _ in range(args.warmup):
        for op, out in zip(ops, outs):
            plan.execute(op, out.ptr, stream.handle)
    stream.sync()

    # timed region: K steps; events around every launch give the kernel
    # durations on the stream the kernels run on
    ev = [[h.Event(), h.Event()] for _ in range(args.steps * len(ops))]
    t_start, t_stop = h.Event(), h.Event()
    stream.sync()
    wall0 = time.perf_counter()
    t_start.record(stream)
    i = 0
    for _ in range(args.steps):
        for op, out in zip(ops, outs):
            ev[i][0].record(stream)
            plan.execute(op, out.ptr, stream.handle)
            ev[i][1].record(stream)
            i += 1
    t_stop.record(stream)
    stream.sync()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_ms(t_stop)
    kms = np.array([a.elapsed_ms(b) for a, b in ev]).reshape(args.steps,
                                                            len(ops))
    ms_per_step = total_ms / args.steps
    value = entries_per_step / (ms_per_step * 1e-3)
    k_avg = kms.mean(axis=0)
    achieved = sum(flops) / (k_avg.sum() * 1e-3) / 1e12
    log("bench: near field %.3e entries/s, %.3f ms per step, %.1f TFLOP/s" %
        (value, ms_per_step, achieved))

    # HBM bytes of one step (both operators' launches of the schedule in
    # use) from the committed counter summary (FETCH_SIZE and WRITE_SIZE per
    # dispatch, each scaled by what it reads for a plain copy); valid for
    # the default workload only
    traffic, counters = None, {}
    tpath = os.path.join(ROOT, "profiles",
                         "traffic_nearfield_h2_sphere32768.json")
    if os.path.exists(tpath) and (args.split, args.clf, args.eta) == \
            (64, 64, 2.0):
        with open(tpath) as f:
            counters = json.load(f)
        kb = counters.get("kernel_bytes", {})
        names = []
        for k in (0, 1):
            if schedule == "per_class":
                names.append("nf_regular_kernel<%d, 1, 1, 2, 0>" % k)
                names += ["nf_singular_kernel<%d, 1, 1, %d, 4, false>" %
                          (k, c) for c in (1, 2, 3)]
            elif schedule == "fused":
                names.append("nf_fused_kernel<%d, 1, 1, 2, 0>" % k)
            else:
                names.append("nf_fused_persistent<%d, 1, 1, 2, 0>" % k)
        if all(n in kb for n in names):
            traffic = sum(kb[n] for n in names)
    # FP64 flops the kernels execute per launch (counter-measured, committed
    # beside the traffic): the executed rate over the same kernel time
    executed = None
    vpath = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                         "profiles", "valu_nearfield_sphere32768.json")
    if os.path.exists(vpath) and (args.split, args.clf, args.eta) == \
            (64, 64, 2.0):
        with open(vpath) as f:
            fl = json.load(f).get("flops_fp64_per_launch", {})
        names = []
        for k in (0, 1):
            if schedule == "per_class":
                names.append("nf_regular_kernel<%d, 1, 1, 2, 0>" % k)
                names += ["nf_singular_kernel<%d, 1, 1, %d, 4, false>" %
                          (k, c) for c in (1, 2, 3)]
            elif schedule == "fused":
                names.append("nf_fused_kernel<%d, 1, 1, 2, 0>" % k)
            else:
                names.append("nf_fused_persistent<%d, 1, 1, 2, 0>" % k)
        if all(n in fl for n in names):
            executed = (sum(fl[n] for n in names) /
                        (k_avg.sum() * 1e-3) / 1e12)

    detail = {}
    if "mixed" in sections:
        log("bench: mixed precision, same blocks")
        detail["mixed_precision"] = bench_mixed(h, ops, plan, outs, stream,
                                                args.steps)
    if "helmholtz" in sections and args.helm_split > 0:
        log("bench: Helmholtz near field, FP64 and mixed")
        detail["helmholtz_mixed"] = bench_helmholtz_mixed(h, args, stream)
    if "batched" in sections and args.batch_meshes > 0:
        log("bench: %d meshes in one launch set" % args.batch_meshes)
        detail["batched"] = bench_batched(h, args, stream)
    if "helmholtz_solve" in sections and args.helm_split > 0 and \
            args.helm_k:
        log("bench: exterior Helmholtz solves, FP64 and mixed")
        detail["helmholtz_solve"] = bench_helmholtz_solve(h, args, stream)
    if "serving_multi" in sections and args.multi_requests > 0:
        log("bench: serving across meshes")
        detail["serving_multi"] = bench_serving_multi(h, args)
    if "dense" in sections:
        log("bench: dense 2048 x 2048 matrix")
        detail["dense"] = bench_dense(h, stream)
    if "dirichlet" in sections and args.solves > 0:
        log("bench: Dirichlet solves")
        dirichlet = bench_dirichlet(h, mesh, args, stream)
        if counters.get("h2_product_bytes"):
            dirichlet["matvec"]["traffic"] = counters["h2_product_bytes"]
        prob, ref = dirichlet.pop("_prob"), dirichlet.pop("_ref")
        if args.serve_requests > 0:
            log("bench: serving")
            dirichlet["serving"] = bench_serving(h, prob, ref, args)
            dirichlet["serving_saturated"] = bench_serving(h, prob, ref,
                                                           args, rate=1e5)
        detail["dirichlet"] = dirichlet
        del prob

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the reference on the host cores")
        cpu = cpu_baseline(mesh, tree, nb, plan.entries, args)
        if cpu:
            detail["cpu_baseline"] = cpu.pop("_detail")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic sphere mesh",
        "config": {
            "workload": ("Laplace SLP+DLP ll near field, sphere %d "
                         "triangles, clf=%d eta=%g q=2 q2=4" %
                         (mesh.triangles, args.clf, args.eta)),
            "entries_per_step": int(entries_per_step),
            "kernel_ms": [round(float(v), 4) for v in k_avg],
            "schedule": schedule,
        },
        "roofline": {
            "bound": "valu",
            "per": "step (SLP and DLP launches)",
            "kernel": ("nf_fused" if schedule.startswith("fused") else
                       "nf_regular+nf_singular"),
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "executed": executed,
            "executed_frac": (executed / PEAK_FP64_TFLOPS
                              if executed is not None else None),
        },
        "cpu_baseline": cpu,
    }
    detail["headline"] = {
        "schedule_ms": sched_ms,
        "near_blocks": int(len(nb)),
        "kernel_evaluations_per_launch": float(work[0].sum()),
        "pairs_per_launch_by_class": [int(v) for v in work[1]],
        "flops_per_launch": [float(f) for f in flops],
        "wall_s_timed": wall}
    line.update(compact(detail))
    full = dict(line)
    full["detail"] = detail
    if args.detail:
        try:
            d = os.path.dirname(args.detail)
            if d:
                os.makedirs(d, exist_ok=True)
            with open(args.detail, "w") as f:
                json.dump(full, f, indent=1, default=float)
            log("bench: full results in %s" % args.detail)
        except OSError as exc:
            log("bench: could not write %s: %s" % (args.detail, exc))
    out = json.dumps(line, separators=(",", ":"), default=float)
    if len(out) > LINE_MAX:
        log("bench: summary line %d characters, dropping the section "
            "summaries" % len(out))
        for key in ("serving_multi", "helmholtz", "batched", "dense",
                    "mixed", "matvec"):
            line.pop(key, None)
            out = json.dumps(line, separators=(",", ":"), default=float)
            if len(out) <= LINE_MAX:
                break
    print(out, flush=True)


if __name__ == "__main__":
    main()

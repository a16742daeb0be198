"""Many meshes in one launch set, the autotuner and the graph-captured
GMRES step (needs a GPU).

* the joined near-field plan of several drawn meshes gives, bit for bit,
  the entries of each mesh's own plan (the per-mesh plans are pinned to the
  oracle elsewhere);
* GMRES on many systems with one captured graph per step meets
  |b - A x| <= eps |b| (up to rounding) for every system and agrees with
  direct launches and with the one-system solver to rounding (the H2
  product sums with atomics, so runs are not bit-reproducible);
* the autotuner picks a candidate leaf size and a schedule without
  changing the entries.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_batch(count=5, seed=3):
    from h2lib_amd.batch import MultiMesh, random_meshes
    specs = random_meshes(count, seed, tmin=96, tmax=900)
    clfs = [16 + 16 * (i % 2) for i in range(count)]
    return specs, MultiMesh(specs, clfs, basis=0)


def test_joined_plan_matches_per_mesh_plans():
    import h2lib_amd as h
    specs, mm = small_batch()
    print("meshes", [s.triangles for s in specs], "entries", mm.entries,
          flush=True)
    op = h.new_slp_laplace_bem3d(mm.gr, 2, 4, 0, 0)
    out = h.DeviceArray(mm.entries)
    out.fill_nan()
    stream = h.Stream()
    mm.execute(op, out.ptr, stream.handle)
    stream.sync()
    flat = out.download()
    for i, s in enumerate(specs):
        m = s.mesh
        gr = h.Surface3d(m.x, m.t, m.n, m.g)
        op1 = h.new_slp_laplace_bem3d(gr, 2, 4, 0, 0)
        nb = mm.btrees[i].near_block_list()
        plan = h.NearfieldPlan(gr, 0, 0, mm.trees[i].idx, mm.trees[i].idx,
                               nb)
        o1 = h.DeviceArray(plan.entries)
        plan.execute(op1, o1.ptr, stream.handle)
        stream.sync()
        a, b = mm.entry_range[i]
        assert np.array_equal(flat[a:b], o1.download()), i


@pytest.mark.parametrize("schedule", [0, 2])
def test_graph_gmres_matches_direct_and_single(schedule):
    # Every solution meets the stopping rule on its true residual.  schedule 2: the flat transforms, whose coefficient
    # buffers are cleared by the products themselves, also across graph
    # replays.
    import h2lib_amd as h
    from h2lib_amd.solve import GmresMulti, solve_gmres
    specs, mm = small_batch(4, seed=11)
    op = h.new_slp_laplace_bem3d(mm.gr, 2, 4, 0, 0)
    out = h.DeviceArray(mm.entries)
    stream = h.Stream()
    mm.execute(op, out.ptr, stream.handle)
    stream.sync()
    mats = mm.h2matrices(out, 3)
    for H in mats:
        H.set_schedule(schedule)
    rng = np.random.default_rng(5)
    bs = [rng.standard_normal(H.n) for H in mats]
    eps = 1e-9

    def true_residual(H, b, d_x):
        r = h.DeviceArray.from_host(b)
        H.addeval(-1.0, d_x, r, stream)
        stream.sync()
        return np.linalg.norm(r.download()) / np.linalg.norm(b)

    G = GmresMulti(mats, kmax=10, stream=stream)
    sols = {}
    for use_graph in (True, False):
        d_b = [h.DeviceArray.from_host(b) for b in bs]
        d_x = [h.DeviceArray.from_host(np.zeros(H.n)) for H in mats]
        it, res = G.solve(d_b, d_x, eps, 500, use_graph=use_graph)
        sols[use_graph] = ([d.download() for d in d_x], it.copy())
        print("graph" if use_graph else "direct", "steps", it, flush=True)
        for s, H in enumerate(mats):
            assert true_residual(H, bs[s], d_x[s]) <= 1.5 * eps, (use_graph,
                                                                  s)
    assert G.graph_builds >= 1
    # reproducible mode: graph replays, direct launches and the one-system
    # solver run the same arithmetic, so the same steps and bits
    for s in range(len(mats)):
        assert np.array_equal(sols[True][0][s], sols[False][0][s]), s
    assert np.array_equal(sols[True][1], sols[False][1])
    for s, H in enumerate(mats):
        d_b = h.DeviceArray.from_host(bs[s])
        d_x = h.DeviceArray.from_host(np.zeros(H.n))
        it1, _ = solve_gmres(H, d_b, d_x, eps, 500, 10, stream)
        assert int(it1) == int(sols[True][1][s]), s
        assert np.array_equal(d_x.download(), sols[True][0][s]), s
    # a second solve with the same graphs and workspaces
    d_b = [h.DeviceArray.from_host(2.0 * b) for b in bs]
    d_x = [h.DeviceArray.from_host(np.zeros(H.n)) for H in mats]
    G.solve(d_b, d_x, eps, 500, use_graph=True)
    for s, H in enumerate(mats):
        assert true_residual(H, 2.0 * bs[s], d_x[s]) <= 1.5 * eps


def test_autotuner_picks_candidates_and_keeps_entries():
    import h2lib_amd as h
    from h2lib_amd.autotune import Autotuner
    specs, mm = small_batch(3, seed=5)
    stream = h.Stream()
    tuner = Autotuner(stream, m=2, candidates=(16, 32), matvecs=10, reps=1)
    for s in specs:
        assert tuner.leaf_size(s.mesh) in (16, 32)
    op = h.new_slp_laplace_bem3d(mm.gr, 2, 4, 0, 0)
    ref = h.DeviceArray(mm.entries)
    mm.execute(op, ref.ptr, stream.handle)
    out = h.DeviceArray(mm.entries)
    best = tuner.tune_schedule("test", mm.plan, op, out.ptr)
    assert best[0] in tuner.TILES and tuple(best[1:]) in tuner.SCHEDULES
    # constant basis: one store per entry, the same bits whatever the tile
    # edge and the schedule (the tuned one, and every other one)
    out.fill_nan()
    mm.execute(op, out.ptr, stream.handle)
    stream.sync()
    assert np.array_equal(out.download(), ref.download())
    assert "test" in tuner.to_json()["schedules"]
    for tile in tuner.TILES:
        mm.plan.set_tile(tile)
        for sched in tuner.SCHEDULES:
            mm.plan.set_schedule(*sched)
            out.fill_nan()
            mm.execute(op, out.ptr, stream.handle)
            stream.sync()
            assert np.array_equal(out.download(), ref.download()), (tile,
                                                                    sched)

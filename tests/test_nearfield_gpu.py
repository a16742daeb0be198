"""Parity of the HIP near-field quadrature with the oracle (needs a GPU).

Every kernel / layer / basis combination, near and far rules, plain and
transposed storage, random index subsets, full dense matrices, windows
inside larger matrices, and the batched device-resident plan.

Tolerance: FP64 throughout; entries agree with the oracle to
|gpu - oracle| <= 1e-12 * max|oracle| over each block (the two differ only
in summation order, fused multiply-adds and the Gauss nodes' last bit).
"""
import numpy as np
import pytest

from conftest import golden, rel_block_err, subsets
from oracle_ffi import OracleBem

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = {
    "lap_slp_cc": (0, 0, 0, 0, 0.0, 0.0),
    "lap_dlp_cc": (0, 1, 0, 0, 0.5, 0.0),
    "lap_slp_ll": (0, 0, 1, 1, 0.0, 0.0),
    "lap_dlp_ll": (0, 1, 1, 1, -0.5, 0.0),
    "lap_slp_cl": (0, 0, 0, 1, 0.0, 0.0),
    "lap_dlp_cl": (0, 1, 0, 1, 0.5, 0.0),
    "lap_slp_lc": (0, 0, 1, 0, 0.0, 0.0),
    "lap_dlp_lc": (0, 1, 1, 0, 0.5, 0.0),
    "helm_slp_cc": (1, 0, 0, 0, 0.0, 2.0),
    "helm_dlp_cc": (1, 1, 0, 0, 0.5 + 0.25j, 2.0),
    "helm_slp_ll": (1, 0, 1, 1, 0.0, 2.0 + 0.5j),
    "helm_dlp_ll": (1, 1, 1, 1, 0.5, 2.0),
    "helm_slp_cl": (1, 0, 0, 1, 0.0, 3.0),
    "helm_dlp_lc": (1, 1, 1, 0, -0.5, 3.0),
    # complex alpha with the linear basis (Re(alpha) only, as the reference)
    "helm_dlp_ll_calpha": (1, 1, 1, 1, 0.5 + 0.25j, 3.0),
    "helm_dlp_cl_calpha": (1, 1, 0, 1, -0.5 + 0.5j, 2.0),
    # the adjoint double layer (layer 2: n_x of the row triangle)
    "lap_adlp_cc": (0, 2, 0, 0, 0.5, 0.0),
    "lap_adlp_ll": (0, 2, 1, 1, -0.5, 0.0),
    "lap_adlp_cl": (0, 2, 0, 1, 0.5, 0.0),
    "lap_adlp_lc": (0, 2, 1, 0, -0.5, 0.0),
    "helm_adlp_cc": (1, 2, 0, 0, 0.5 + 0.25j, 2.0),
    "helm_adlp_ll": (1, 2, 1, 1, -0.5, 2.0 + 0.5j),
    "helm_adlp_cl": (1, 2, 0, 1, 0.5, 3.0),
    "helm_adlp_lc_calpha": (1, 2, 1, 0, 0.5 + 0.25j, 3.0),
}


def gpu_op(mesh, name, q=2, q2=4):
    import h2lib_amd as h
    kern, op, br, bc, alpha, k = CASES[name]
    gr = h.Surface3d(mesh["x"], mesh["t"])
    return h.Bem3d(gr, kern, op, q, q2, br, bc, alpha=alpha, k=k)


def oracle_op(mesh, name, q=2, q2=4):
    kern, op, br, bc, alpha, k = CASES[name]
    return OracleBem(kern, op, br, bc, mesh["x"], mesh["t"], q=q, q2=q2,
                     alpha=alpha, k=k)


@pytest.mark.parametrize("name", sorted(CASES))
def test_dense_matches_oracle(name, sphere8):
    g = gpu_op(sphere8, name)
    o = oracle_op(sphere8, name)
    print("dense", name, flush=True)
    for far in (False, True):
        a = g.nearfield_far(None, None) if far else g.nearfield(None, None)
        b = o.nearfield(far=far)
        assert rel_block_err(a, b) < TOL, far


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("ntrans", [False, True])
def test_subsets_match_oracle(name, ntrans, sphere8):
    g = gpu_op(sphere8, name)
    o = oracle_op(sphere8, name)
    rng = np.random.default_rng(5)
    kern, op, br, bc, alpha, k = CASES[name]
    nr = len(sphere8["x"]) if br else len(sphere8["t"])
    nc = len(sphere8["x"]) if bc else len(sphere8["t"])
    for rows, cols in [(37, 29), (64, 64), (65, 130), (5, 3), (1, 90),
                       (200, 1)]:
        ridx, cidx = subsets(rng, nr, nc, rows, cols, nr == nc)
        print("subset", name, ntrans, rows, cols, flush=True)
        for far in (False, True):
            f = g.nearfield_far if far else g.nearfield
            a = f(ridx, cidx, ntrans=ntrans)
            b = o.nearfield(ridx, cidx, far=far, ntrans=ntrans)
            assert rel_block_err(a, b) < TOL, (rows, cols, far)


@pytest.mark.parametrize("name", ["lap_dlp_cc", "helm_dlp_cc", "lap_adlp_cc",
                                  "lap_slp_ll", "lap_dlp_ll", "helm_slp_ll",
                                  "lap_dlp_cl", "lap_dlp_lc"])
@pytest.mark.parametrize("ntrans", [0, 1])
def test_repeated_indices_match_oracle(name, ntrans, sphere8):
    # Index arrays that repeat DOFs.  Constant basis: the reference computes
    # such columns again and adds the identity term at (t, t) for every
    # column s with ridx[t] == cidx[s], where a later column t overwrites
    # the terms of the columns before it (Library/bem3d.c:1150-1247).
    # Linear basis: a support triangle records the first three block
    # positions whose vertex it touches (Library/bem3d.c:3071-3116); a
    # vertex here appears three times, across the 64-row tile edge.
    g = gpu_op(sphere8, name)
    o = oracle_op(sphere8, name)
    kern, op, br, bc, alpha, k = CASES[name]
    rng = np.random.default_rng(13)
    nr = len(sphere8["x"]) if br else len(sphere8["t"])
    nc = len(sphere8["x"]) if bc else len(sphere8["t"])
    base = rng.permutation(nc)[:20]
    cidx = np.concatenate([base, base[[3, 0, 7, 3, 11]]]).astype(np.uint32)
    if not br and not bc:
        # rows from the other DOFs: a shared DOF at row position >= cols
        # would put the reference's identity term outside the window
        taken = set(int(v) for v in base)
        ridx = np.array([v for v in rng.permutation(nr)
                         if int(v) not in taken][:30])
        # shared DOFs at row positions below cols, before and after the
        # columns that hold them, some of them repeated among the rows too
        for pos, j in [(0, 0), (2, 3), (5, 21), (9, 7), (14, 3), (20, 11),
                       (24, 0)]:
            ridx[pos] = cidx[j]
    else:
        ridx = rng.permutation(nr)[:70]
        ridx[[66, 67]] = ridx[10]
        ridx[5] = ridx[2]
        ridx[40] = ridx[41]
    ridx = ridx.astype(np.uint32)
    for far in (False, True):
        f = g.nearfield_far if far else g.nearfield
        a = f(ridx, cidx, ntrans=ntrans)
        b = o.nearfield(ridx, cidx, far=far, ntrans=ntrans)
        assert rel_block_err(a, b) < TOL, far


def test_window_inside_larger_matrix(sphere8):
    # the callee writes only its window of a caller-owned matrix with ld
    g = gpu_op(sphere8, "lap_dlp_ll")
    o = oracle_op(sphere8, "lap_dlp_ll")
    rng = np.random.default_rng(9)
    ridx = rng.permutation(len(sphere8["x"]))[:40]
    cidx = rng.permutation(len(sphere8["x"]))[:33]
    big = np.full((70, 50), 7.0, order="F")
    win = big[10:50, 5:38]
    g.nearfield(ridx, cidx, N=win)
    assert rel_block_err(win, o.nearfield(ridx, cidx)) < TOL
    mask = np.ones_like(big, dtype=bool)
    mask[10:50, 5:38] = False
    assert np.all(big[mask] == 7.0)


@pytest.mark.parametrize("q,q2", [(1, 3), (3, 5), (4, 6)])
def test_other_orders(q, q2, sphere4):
    for name in ("lap_slp_cc", "helm_dlp_ll", "lap_dlp_cl", "lap_adlp_lc"):
        g = gpu_op(sphere4, name, q, q2)
        o = oracle_op(sphere4, name, q, q2)
        print("orders", name, q, q2, flush=True)
        assert rel_block_err(g.nearfield(None, None), o.nearfield()) < TOL


def test_sampled_entries_against_reference_fixture():
    # single entries of every rule class on the s = 16 sphere, values from
    # the reference library
    import h2lib_amd as h
    m = golden("mesh_sphere_s16.npz")
    s = golden("sampled_sphere_s16.npz")
    gr = h.Surface3d(m["x"], m["t"])
    for name, kern, op, alpha, k in [("lap_slp", 0, 0, 0.0, 0.0),
                                     ("lap_dlp", 0, 1, 0.5, 0.0),
                                     ("helm_slp", 1, 0, 0.0, 4.0)]:
        b = h.Bem3d(gr, kern, op, 2, 4, 0, 0, alpha=alpha, k=k)
        got = np.array([b.nearfield([i], [j])[0, 0]
                        for i, j in zip(s["rows"], s["cols"])])
        ref = s[name]
        tiny = np.abs(ref) < 1e-14 * np.abs(ref).max()
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        assert np.all(err[~tiny] < TOL), name


def test_batched_plan_matches_single_calls(sphere8):
    import h2lib_amd as h
    rng = np.random.default_rng(12)
    for name in ("lap_slp_ll", "helm_dlp_cc", "lap_dlp_lc", "lap_adlp_ll"):
        kern, op, br, bc, alpha, k = CASES[name]
        gr = h.Surface3d(sphere8["x"], sphere8["t"])
        b = h.Bem3d(gr, kern, op, 2, 4, br, bc, alpha=alpha, k=k)
        nr = gr.vertices if br else gr.triangles
        nc = gr.vertices if bc else gr.triangles
        ridx = rng.permutation(nr).astype(np.uint32)
        cidx = rng.permutation(nc).astype(np.uint32)
        blocks = []
        for _ in range(40):
            rows = int(rng.integers(1, 90))
            cols = int(rng.integers(1, 90))
            roff = int(rng.integers(0, nr - rows))
            coff = int(rng.integers(0, nc - cols))
            blocks.append((roff, rows, coff, cols))
        plan = h.NearfieldPlan(gr, br, bc, ridx, cidx, blocks)
        out = h.DeviceArray(plan.entries, b.dtype)
        out.fill_nan()
        stream = h.Stream()
        plan.execute(b, out.ptr, stream.handle)
        stream.sync()
        flat = out.download()
        for i, (roff, rows, coff, cols) in enumerate(blocks):
            want = b.nearfield(ridx[roff:roff + rows],
                               cidx[coff:coff + cols])
            got = plan.block(flat, i)
            assert rel_block_err(got, want) < 1e-14, (name, i)


@pytest.mark.parametrize("sorted_tiles,persistent,fused",
                         [(1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1),
                          (0, 1, 1)])
def test_schedules_give_identical_entries(sorted_tiles, persistent, fused,
                                          sphere8):
    # the launch schedule changes the order of the tiles only: the entries
    # are the same (bit for bit where no atomic sums are involved), in FP64
    # and in mixed precision; fused tiles evaluate their pairs with common
    # vertices in their own workgroup (one launch, no global atomics)
    import h2lib_amd as h
    rng = np.random.default_rng(21)
    for name in ("lap_slp_cc", "helm_dlp_cc", "helm_dlp_ll", "lap_dlp_lc",
                 "helm_slp_cl", "lap_adlp_ll"):
        kern, op, br, bc, alpha, k = CASES[name]
        gr = h.Surface3d(sphere8["x"], sphere8["t"])
        b = h.Bem3d(gr, kern, op, 2, 4, br, bc, alpha=alpha, k=k)
        nr = gr.vertices if br else gr.triangles
        nc = gr.vertices if bc else gr.triangles
        ridx = rng.permutation(nr).astype(np.uint32)
        cidx = rng.permutation(nc).astype(np.uint32)
        blocks = []
        for _ in range(60):
            rows = int(rng.integers(1, 140))
            cols = int(rng.integers(1, 140))
            blocks.append((int(rng.integers(0, nr - rows)), rows,
                           int(rng.integers(0, nc - cols)), cols))
        plan = h.NearfieldPlan(gr, br, bc, ridx, cidx, blocks)
        stream = h.Stream()
        for mode in (0, 1, 2):
            b.set_precision(mode)
            plan.set_schedule(0, 0)
            ref = h.DeviceArray(plan.entries, b.dtype)
            plan.execute(b, ref.ptr, stream.handle)
            plan.set_schedule(sorted_tiles, persistent, fused)
            out = h.DeviceArray(plan.entries, b.dtype)
            out.fill_nan()
            for _ in range(2):  # the tile counter is reset per launch
                plan.execute(b, out.ptr, stream.handle)
            stream.sync()
            got, want = out.download(), ref.download()
            if br == 0 and bc == 0:
                # one store or one atomic add per entry: deterministic
                assert np.array_equal(got, want), (name, mode)
            else:
                # the linear basis sums element matrices with atomic adds,
                # whose order varies from run to run
                err = np.abs(got - want).max() / np.abs(want).max()
                assert err < 1e-14, (name, mode, err)


@pytest.mark.parametrize("tile", [32, 24, 8])
def test_tile_edges_give_identical_entries(tile, sphere8):
    # a plan cut into smaller tiles gives the entries of the 64-DOF tiles
    # (constant basis bit for bit; the linear basis evaluates a triangle
    # pair once per tile its vertices fall into and sums in another order),
    # with the launches per rule class and with fused tiles
    import h2lib_amd as h
    rng = np.random.default_rng(5)
    for name in ("lap_dlp_cc", "lap_slp_ll", "helm_dlp_lc"):
        kern, op, br, bc, alpha, k = CASES[name]
        gr = h.Surface3d(sphere8["x"], sphere8["t"])
        b = h.Bem3d(gr, kern, op, 2, 4, br, bc, alpha=alpha, k=k)
        nr = gr.vertices if br else gr.triangles
        nc = gr.vertices if bc else gr.triangles
        ridx = rng.permutation(nr).astype(np.uint32)
        cidx = rng.permutation(nc).astype(np.uint32)
        blocks = []
        for _ in range(40):
            rows = int(rng.integers(1, 150))
            cols = int(rng.integers(1, 150))
            blocks.append((int(rng.integers(0, nr - rows)), rows,
                           int(rng.integers(0, nc - cols)), cols))
        plan = h.NearfieldPlan(gr, br, bc, ridx, cidx, blocks)
        stream = h.Stream()
        ref = h.DeviceArray(plan.entries, b.dtype)
        plan.execute(b, ref.ptr, stream.handle)
        stream.sync()
        want = ref.download()
        plan.set_tile(tile)
        for fused in (0, 1):
            plan.set_schedule(0, 0, fused)
            out = h.DeviceArray(plan.entries, b.dtype)
            out.fill_nan()
            plan.execute(b, out.ptr, stream.handle)
            stream.sync()
            got = out.download()
            if br == 0 and bc == 0:
                assert np.array_equal(got, want), (name, fused)
            else:
                err = np.abs(got - want).max() / np.abs(want).max()
                assert err < 1e-14, (name, fused, err)

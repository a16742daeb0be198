"""Dirichlet solves on the GPU against the reference library (needs a GPU).

The same problem as the reference's Laplace test (interior, linear basis,
u = 1/|x - (1.2, 1.2, 1.2)|) solved with GMRES(30): the projected data f and
the right-hand side b agree with the reference to 1e-12; the solutions agree
to 1e-6 relative (two GMRES runs stopped at residual 1e-10 |b|, rounding
differs); the Neumann error lies in the reference test's band.
"""
import numpy as np
import pytest

from oracle_ffi import have_reference, reference_dirichlet

pytestmark = pytest.mark.gpu

X0 = np.array([1.2, 1.2, 1.2])


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("split,clf,m", [(8, 16, 3), (16, 32, 3)])
def test_dirichlet_matches_reference(split, clf, m):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    mesh = G.sphere(split)
    prob = LaplaceDirichlet(mesh, clf=clf, eta=2.0, m=m)
    d_f, d_b = prob.rhs(X0)
    d_x, it, res = prob.solve(d_b, eps=1e-10, maxiter=2000, kmax=30)
    x = d_x.download()
    f_ref, b_ref, x_ref, it_ref, err_ref = reference_dirichlet(
        mesh.x, mesh.t, clf, 2.0, m, 1e-10, 2000, 30)
    print("dirichlet split", split, "steps", it, it_ref, flush=True)
    assert np.abs(d_f.download() - f_ref).max() < 1e-12 * np.abs(f_ref).max()
    assert np.abs(d_b.download() - b_ref).max() < 1e-12 * np.abs(b_ref).max()
    assert np.abs(x - x_ref).max() < 1e-6 * np.abs(x_ref).max()
    err = prob.neumann_error(x, X0)
    assert abs(err - err_ref) < 1e-6
    assert abs(it - it_ref) <= max(2, it_ref // 10)


def test_gmres_residual_on_random_rhs():
    # textbook check independent of the reference: |b - V x| <= eps |b|
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    mesh = G.sphere(10)
    prob = LaplaceDirichlet(mesh, clf=24, eta=2.0, m=3)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(prob.n)
    d_b = h.DeviceArray.from_host(b)
    for kmax in (5, 30):
        d_x, it, res = prob.solve(d_b, eps=1e-9, maxiter=5000, kmax=kmax)
        r = h.DeviceArray.from_host(b)
        prob.V.addeval(-1.0, d_x, r)
        assert np.linalg.norm(r.download()) <= 2e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("kind,q", [(0, 2), (1, 2), (0, 3)])
def test_device_projection_matches_host(kind, q):
    # the device projection against the host restatement of
    # projectL2_bem3d_l_avector (same rule, same stopping rule)
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    mesh = G.sphere(12)
    prob = LaplaceDirichlet(mesh, clf=24, eta=2.0, m=2, q=q)
    x0 = np.array([1.2, 1.2, 1.2]) if kind == 0 else np.array([0.3, 0.2, 0.1])
    want = prob.project(x0, kind)
    d = h.DeviceArray(prob.n)
    prob.project_device(x0, d, kind)
    got = d.download()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # a second call on the same surface reuses the cached data
    prob.project_device(2.0 * x0, d, kind)
    want2 = prob.project(2.0 * x0, kind)
    assert np.abs(d.download() - want2).max() < 1e-12 * np.abs(want2).max()


@pytest.mark.parametrize("kmax", [5, 30])
def test_device_gmres_matches_restatement(kmax):
    # device GMRES on an H2-matrix built from host arrays against the CPU
    # restatement of the reference's GMRES with the restated product, on
    # the same matrix: same steps (up to one), solutions to 1e-8
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix, Interpolation
    from h2lib_amd.solve import solve_gmres
    from oracle_ffi import (Oracle, OracleBem, oracle_gmres,
                            oracle_nearfield_packed)
    mesh = G.sphere(8)
    rt = G.ClusterTree(mesh, 1, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    I = Interpolation(mesh, rt, rt, bt, 1, 1, False, 3, 2)
    ob = OracleBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    N = oracle_nearfield_packed(ob, bt.near_block_list(), rt.idx, rt.idx)
    H = H2Matrix.from_arrays(I, N, rt=rt)
    n = rt.n
    b = np.random.default_rng(6).standard_normal(n)
    want, it_o, _ = oracle_gmres(
        lambda v: Oracle.h2_addeval(I, N, v, np.zeros(n)), n, b, np.zeros(n),
        1e-10, 2000, kmax)
    d_b = h.DeviceArray.from_host(b)
    d_x = h.DeviceArray.from_host(np.zeros(n))
    it, res = solve_gmres(H, d_b, d_x, 1e-10, 2000, kmax)
    print("gmres steps device", it, "restatement", it_o, flush=True)
    assert abs(it - it_o) <= 1
    assert np.abs(d_x.download() - want).max() < 1e-8 * np.abs(want).max()


@pytest.mark.parametrize("kmax,maxiter", [(10, 1), (10, 2), (30, 9),
                                          (30, 17), (10, 37), (30, 45)])
def test_device_gmres_stops_at_maxiter_like_restatement(kmax, maxiter):
    # a tolerance the solve does not reach: both stop after maxiter - 1
    # steps (the reference's loop), also where the limit falls inside a
    # chunk of steps queued ahead of the host's look at the estimates, or
    # right after a restart; the device's residual estimate is the true
    # residual of its iterate.  An iterate stopped early depends on the
    # rounding along the Krylov path (Gram-Schmidt here, Householder in the
    # reference), which this matrix amplifies (|x| is about 5e3 |b|;
    # measured after 36 steps of GMRES(10): residuals 4.6e-7 apart,
    # iterates 1.8e-8): both to 1e-5 of the restatement's
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix, Interpolation
    from h2lib_amd.solve import solve_gmres
    from oracle_ffi import (Oracle, OracleBem, oracle_gmres,
                            oracle_nearfield_packed)
    mesh = G.sphere(8)
    rt = G.ClusterTree(mesh, 1, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    I = Interpolation(mesh, rt, rt, bt, 1, 1, False, 3, 2)
    ob = OracleBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    N = oracle_nearfield_packed(ob, bt.near_block_list(), rt.idx, rt.idx)
    H = H2Matrix.from_arrays(I, N, rt=rt)
    n = rt.n
    b = np.random.default_rng(7).standard_normal(n)
    want, it_o, res_o = oracle_gmres(
        lambda v: Oracle.h2_addeval(I, N, v, np.zeros(n)), n, b, np.zeros(n),
        1e-15, maxiter, kmax)
    d_b = h.DeviceArray.from_host(b)
    d_x = h.DeviceArray.from_host(np.zeros(n))
    it, res = solve_gmres(H, d_b, d_x, 1e-15, maxiter, kmax)
    print("maxiter", maxiter, "kmax", kmax, "steps", it, it_o, flush=True)
    assert it == it_o == maxiter - 1
    got = d_x.download()
    true = np.linalg.norm(b - Oracle.h2_addeval(I, N, got, np.zeros(n)))
    assert abs(true - res) <= 1e-3 * res + 1e-10 * np.linalg.norm(b)
    assert abs(res - res_o) <= 1e-5 * res_o
    assert np.abs(got - want).max() <= 1e-5 * max(np.abs(want).max(), 1e-300)


def test_gmres_with_more_rows_than_reduction_blocks():
    # 73728 complex unknowns: the update kernels run 288 workgroups of 256
    # rows, more than the 240 of the multi-dot grid, and every one of them
    # leaves a partial |w|^2 for the last one to sum (regression: the
    # partial-sum buffer was sized for 240 workgroups)
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import HelmholtzExterior
    mesh = G.sphere(96)
    prob = HelmholtzExterior(mesh, 2.0, clf=64, m=3)
    _, d_b = prob.rhs(np.array([0.3, 0.2, 0.1]))
    d_x, it, res = prob.solve(d_b, eps=1e-6, maxiter=40, kmax=30)
    b = d_b.download()
    r = h.DeviceArray.from_host(b)
    prob.V.addeval(-1.0, d_x, r)
    true = np.linalg.norm(r.download())
    print("steps", it, "estimate %.3e true %.3e" % (res, true), flush=True)
    assert it > 0 and np.isfinite(true)
    # GMRES's least-squares residual is the true residual up to rounding
    assert abs(true - res) <= 1e-3 * res + 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_gmres_throughput_mode_matches_reproducible(kind):
    # GMRES with the throughput reductions (products with atomics, the
    # Gram-Schmidt sums of each step passed between its kernels without
    # atomics) against the fixed-order solve of the same system: true
    # residual at the requested level, steps within max(2, 3%), solutions
    # within 1e-7 (both stop at |r| <= 1e-9 |b|).  Constant basis: every
    # near-field entry is written once, so the operator is the same bits in
    # every run (the linear basis adds its pairs' contributions with
    # atomics, and GMRES(10) at 1e-9 turned a rounding-level change of the
    # operator into 254 instead of 262 steps in one run)
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import (Surface3d, new_slp_helmholtz_bem3d,
                                 new_slp_laplace_bem3d)
    from h2lib_amd.h2matrix import H2Matrix
    from h2lib_amd.solve import solve_gmres
    mesh = G.sphere(10)
    rng = np.random.default_rng(12)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = (new_slp_laplace_bem3d(gr, 2, 4, 0, 0) if kind == "real" else
          new_slp_helmholtz_bem3d(2.0, gr, 2, 4, 0, 0))
    rt = G.ClusterTree(mesh, 0, 24)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    b = rng.standard_normal(H.n)
    if kind != "real":
        b = b + 1j * rng.standard_normal(H.n)
    d_b = h.DeviceArray.from_host(b)
    out = {}
    for rep in (True, False):
        h.set_reproducible(rep)
        try:
            for kmax in (10, 30):
                d_x = h.DeviceArray(H.n, H.dtype)
                d_x.zero()
                it, res = solve_gmres(H, d_b, d_x, 1e-9, 4000, kmax)
                out[(rep, kmax)] = (it, d_x.download())
        finally:
            h.set_reproducible(True)
    for kmax in (10, 30):
        it_r, x_r = out[(True, kmax)]
        it_a, x_a = out[(False, kmax)]
        print(kind, "kmax", kmax, "steps", it_r, it_a, flush=True)
        assert it_r < 3999
        assert abs(it_a - it_r) <= max(2, 0.03 * it_r)
        assert np.abs(x_a - x_r).max() <= 1e-7 * np.abs(x_r).max()
        r = h.DeviceArray.from_host(b)
        H.addeval(-1.0, h.DeviceArray.from_host(x_a), r)
        assert np.linalg.norm(r.download()) <= 2e-9 * np.linalg.norm(b)

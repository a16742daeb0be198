"""The mixed-precision Helmholtz workload at full size: unit sphere of
131072 triangles (split 128), single layer, piecewise constant basis,
q = 2, q2 = 4, cluster leaf size 64, eta = 2, k = 2, 4, 8.

Parity and accuracy, each against a checker independent of the path:
* FP64 near-field blocks drawn by seed against the oracle (1e-12 of each
  block's largest entry, k = 2);
* mixed-precision entries (FP32 kernel values, FP64 geometry differences
  for singular pairs and anchor points, compensated FP32 inner sums)
  against the FP64 entries: within 2e-6 of the largest entry, for every
  wave number;
* the row sums of the single layer operator (near field plus the H2
  far field) against the closed form |T_i| e^{ik} sin(k)/k of the unit
  sphere (a single layer of unit density): within 2e-3 relative at k = 2
  (the polyhedral surface's error, 2.4e-2 / 6.0e-3 on the 512- / 2048-
  triangle spheres, falls as h^2), within 3e-2 at k = 8 with interpolation
  of order 5 and eta = 1, and mixed against FP64 within 1e-6;
* the exterior Dirichlet problem with a point source at (0.3, 0.2, 0.1):
  FP64 and mixed solutions meet their true residual bound, agree within
  1e-5, and the Neumann error lies in the band of the discretisation
  (k = 2); at k = 8, close to an interior Dirichlet eigenvalue of the ball
  where V is singular, the Brakhage-Werner form KM - i k V is solved
  instead (interpolation of order 5, eta = 1), FP64 against mixed within
  1e-5.
"""
import numpy as np
import pytest

from conftest import rel_block_err
from oracle_ffi import OracleBem

pytestmark = pytest.mark.gpu

SPLIT = 128


@pytest.fixture(scope="module")
def cfg4():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    mesh = G.sphere(SPLIT)
    tree = G.ClusterTree(mesh, 0, 64)
    bt = G.BlockTree(tree, tree, 2.0, 0)
    nb = np.asarray(bt.near_block_list())
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    plan = h.NearfieldPlan(gr, 0, 0, tree.idx, tree.idx, nb)
    return mesh, tree, bt, nb, gr, plan


@pytest.mark.parametrize("k", [2.0, 4.0, 8.0])
def test_mixed_entries_against_fp64(cfg4, k):
    import h2lib_amd as h
    mesh, tree, bt, nb, gr, plan = cfg4
    op64 = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
    op32 = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
    op32.set_precision(1)
    o64 = h.DeviceArray(plan.entries, np.complex128)
    o32 = h.DeviceArray(plan.entries, np.complex128)
    print("cfg4 near field k", k, plan.entries, "entries", flush=True)
    plan.execute(op64, o64.ptr)
    plan.execute(op32, o32.ptr)
    h.hip().h2hip_stream_sync(None)
    a64, a32 = o64.download(), o32.download()
    err = np.abs(a32 - a64).max() / np.abs(a64).max()
    print("  mixed vs fp64 %.2e" % err, flush=True)
    assert err < 2e-6
    if k == 2.0:
        orc = OracleBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=k)
        rng = np.random.default_rng(128)
        diag = [b for b in range(len(nb)) if nb[b, 0] == nb[b, 2]][:3]
        for b in list(diag) + list(rng.choice(len(nb), 9, replace=False)):
            r0, rows, c0, cols = (int(v) for v in nb[b])
            want = orc.nearfield(tree.idx[r0:r0 + rows],
                                 tree.idx[c0:c0 + cols])
            assert rel_block_err(plan.block(a64, b), want) < 1e-12, b


@pytest.mark.parametrize("k,m,eta,tol", [(2.0, 4, 2.0, 2e-3),
                                          (8.0, 5, 1.0, 3e-2)])
def test_row_sums_closed_form(cfg4, k, m, eta, tol):
    """At k = 8 interpolation of order 4 with eta = 2 cannot resolve
    e^{ikr} across the largest admissible blocks: the reference's own
    H2-matrix of that configuration misses the closed form by 17.8 % on the
    8192-triangle sphere (order 5, eta = 1: 1.7 %; the surface's own error
    there is 1.5e-3), so the check at k = 8 uses order 5 and eta = 1.  The
    mixed-precision product must give the FP64 one's row sums to 1e-6."""
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh, tree, bt0, nb0, gr, plan0 = cfg4
    bt = bt0 if eta == 2.0 else G.BlockTree(tree, tree, eta, 0)
    plan = plan0 if eta == 2.0 else h.NearfieldPlan(
        gr, 0, 0, tree.idx, tree.idx, bt.near_block_list())
    exact = np.exp(1j * k) * np.sin(k) / k
    rows = {}
    for precision in (0, 1):
        op = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
        op.set_precision(precision)
        H = H2Matrix(op, mesh, tree, tree, bt, m, nearplan=plan)
        H.set_schedule(2)
        ones = h.DeviceArray.from_host(np.ones(mesh.triangles,
                                               np.complex128))
        y = h.DeviceArray.from_host(np.zeros(mesh.triangles, np.complex128))
        H.addeval(1.0, ones, y)
        rs = y.download() / (0.5 * mesh.g)
        err = np.abs(rs - exact).max() / abs(exact)
        print("row sums k", k, "m", m, "eta", eta, "precision", precision,
              "rel err %.2e" % err, flush=True)
        assert err < tol
        rows[precision] = rs
        del H
    d = np.abs(rows[1] - rows[0]).max() / np.abs(rows[0]).max()
    print("mixed vs fp64 row sums %.2e" % d, flush=True)
    assert d < 1e-6


def test_exterior_problem_k2_fp64_and_mixed(cfg4):
    import h2lib_amd as h
    from h2lib_amd.solve import HelmholtzExterior
    mesh = cfg4[0]
    x0 = np.array([0.3, 0.2, 0.1])
    st = h.Stream()
    sols = {}
    for prec in (0, 1):
        prob = HelmholtzExterior(mesh, 2.0, clf=64, m=4, precision=prec,
                                 stream=st)
        _, d_b = prob.rhs(x0, st)
        d_x, it, res = prob.solve(d_b, eps=1e-8, maxiter=1000, kmax=30,
                                  stream=st)
        st.sync()
        r = h.DeviceArray.from_host(d_b.download())
        prob.V.addeval(-1.0, d_x, r, st)
        st.sync()
        true = np.linalg.norm(r.download()) / np.linalg.norm(d_b.download())
        x = d_x.download()
        err = prob.neumann_error(x, x0)
        print("k=2 precision", prec, "steps", it, "true residual %.2e" % true,
              "neumann %.4e" % err, flush=True)
        assert it < 999 and true <= 1.5e-8
        assert 0.015 < err < 0.035
        sols[prec] = x
        del prob
    d = np.linalg.norm(sols[1] - sols[0]) / np.linalg.norm(sols[0])
    print("mixed vs fp64 solution %.2e" % d, flush=True)
    assert d < 1e-5


def test_brakhage_werner_k8_fp64_and_mixed(cfg4):
    import h2lib_amd as h
    from h2lib_amd.solve import HelmholtzBrakhageWerner, solve_gmres
    mesh = cfg4[0]
    x0 = np.array([0.3, 0.2, 0.1])
    st = h.Stream()
    sols = {}
    for prec in (0, 1):
        prob = HelmholtzBrakhageWerner(mesh, 8.0, clf=64, eta_adm=1.0, m=5,
                                       precision=prec, stream=st)
        d_b = prob.rhs(x0, st)
        d_x = h.DeviceArray(prob.n, np.complex128)
        d_x.zero(st)
        it, res = solve_gmres(prob.A, d_b, d_x, 1e-8, 1000, 30, st)
        st.sync()
        r = h.DeviceArray.from_host(d_b.download())
        prob.A.addeval(-1.0, d_x, r, st)
        st.sync()
        true = np.linalg.norm(r.download()) / np.linalg.norm(d_b.download())
        print("BW k=8 precision", prec, "steps", it,
              "true residual %.2e" % true,
              "outer potential error %.2e" % prob.outer_error(d_x, x0),
              flush=True)
        assert it < 999 and true <= 1.5e-8
        sols[prec] = d_x.download()
        del prob
    d = np.linalg.norm(sols[1] - sols[0]) / np.linalg.norm(sols[0])
    print("mixed vs fp64 solution %.2e" % d, flush=True)
    assert d < 1e-5

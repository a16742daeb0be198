"""Brakhage-Werner formulation of the exterior Helmholtz problem on the GPU
(needs a GPU): the reference's Helmholtz test (Tests/test_helmholtzbem3d.c:
353-450, 695-725) with the piecewise constant basis, unit sphere with 512
triangles, k = 2, leaves of at most 32 triangles, max-cluster admissibility
with eta = 1, interpolation order 5, point source (0, 0, 0.2) inside:
b = integrals of the Dirichlet data, (KM - i k V) x = b with KM the double
layer operator with alpha = 1/2, GMRES(50) to 1e-12, then the potential
eval_brakhage_werner_c of x on a 20 x 20 grid of the xz-plane outside the
unit disc against the point-source field.

* the potential kernel against the reference's eval_brakhage_werner_c on a
  random density to 1e-12 (relative, max-norm);
* the combined operator's product against the reference's two products;
* the solve: the outer error lies in the reference test's accepted band
  [1e-3, 2e-3] and within 1e-6 of the reference library's own outer error
  on its own solution (the reference's complex Givens steps,
  Library/krylov.c:551-610, are not unitary, so step counts differ; both
  solutions are solved to a residual far below the discretisation error);
* mixed precision near fields keep the outer error within 1e-6;
* the piecewise linear variant (the test's second suite: leaves of at most
  16 vertices, eval_brakhage_werner_l): the potential against the
  reference's to 1e-12, the solve's outer error within 1e-7 of the
  reference library's on its own solution (2.33863e-4 for the reference's
  own test run, which prints "NOT okay" for this case because its accepted
  band [2.5e-4, 3.5e-4] starts above it; the band is not asserted).
"""
import numpy as np
import pytest

from oracle_ffi import ReferenceBem, ReferenceH2, have_reference

pytestmark = pytest.mark.gpu

X0 = np.array([0.0, 0.0, 0.2])
K = 2.0


def _problem(precision=0, basis=0):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import HelmholtzBrakhageWerner
    mesh = G.sphere(8)
    return mesh, HelmholtzBrakhageWerner(mesh, K, clf=16 if basis else 32,
                                         eta_adm=1.0, m=5,
                                         precision=precision, basis=basis)


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_potential_matches_reference():
    from h2lib_amd.device import DeviceArray
    mesh, prob = _problem()
    rng = np.random.default_rng(7)
    w = rng.standard_normal(mesh.triangles) + 1j * rng.standard_normal(
        mesh.triangles)
    pts = prob.outer_grid()
    assert len(pts) == 396
    got = prob.potential(DeviceArray.from_host(w), pts)
    ref = ReferenceBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=K)
    want = ref.eval_bw(w, K, pts)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("potential error %.2e" % err, flush=True)
    assert err < 1e-12


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_combined_product_matches_reference():
    from h2lib_amd.device import DeviceArray
    mesh, prob = _problem()
    rslp = ReferenceBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=K)
    rdlp = ReferenceBem(1, 1, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=K,
                        alpha=0.5)
    V = ReferenceH2(rslp, 32, 1.0, 5, 0)
    KM = ReferenceH2(rdlp, 32, 1.0, 5, 0)
    rng = np.random.default_rng(3)
    n = mesh.triangles
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    alpha = 0.7 - 0.4j
    want = KM.addeval(x, y0.copy(), alpha=alpha)
    want = V.addeval(x, want, alpha=-1j * K * alpha)
    d_x = DeviceArray.from_host(x)
    d_y = DeviceArray.from_host(y0)
    prob.A.addeval(alpha, d_x, d_y)
    got = d_y.download()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("combined product error %.2e" % err, flush=True)
    assert err < 1e-12


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_bw_solve_outer_error_matches_reference():
    mesh, prob = _problem()
    d_b = prob.rhs(X0)
    d_x, it, res = prob.solve(d_b, eps=1e-12, maxiter=1000, kmax=50)
    x = d_x.download()
    err = prob.outer_error(d_x, X0)
    rslp = ReferenceBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=K)
    rdlp = ReferenceBem(1, 1, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=K,
                        alpha=0.5)
    V = ReferenceH2(rslp, 32, 1.0, 5, 0)
    KM = ReferenceH2(rdlp, 32, 1.0, 5, 0)
    b_ref = rdlp.helmholtz_point_integrate(K, X0)
    x_ref, it_ref = V.bw_solve(KM, K, b_ref, 1e-12, 1000, 50)
    err_ref = rslp.outer_error(x_ref, K, X0)
    # the reference's error measure on this solution
    err_cross = rslp.outer_error(x, K, X0)
    b = d_b.download()
    r = KM.addeval(x, b.copy(), alpha=-1.0)
    r = V.addeval(x, r, alpha=1j * K)
    rel_res = np.linalg.norm(r) / np.linalg.norm(b)
    print("bw steps", it, it_ref, "true residual %.2e outer error %.6e "
          "(reference measure %.6e, reference solution %.6e)"
          % (rel_res, err, err_cross, err_ref), flush=True)
    assert np.abs(b - b_ref).max() < 1e-13 * np.abs(b_ref).max()
    assert rel_res < 1e-11
    assert 1e-3 <= err <= 2e-3
    assert abs(err - err_cross) < 1e-12
    assert abs(err - err_ref) < 1e-6


def test_bw_mixed_precision_outer_error():
    errs = []
    for prec in (0, 1):
        _, prob = _problem(prec)
        d_x, it, _ = prob.solve(prob.rhs(X0), eps=1e-12, maxiter=1000,
                                kmax=50)
        errs.append(prob.outer_error(d_x, X0))
        print("precision", prec, "steps", it, "outer error %.6e" % errs[-1],
              flush=True)
    assert 1e-3 <= errs[1] <= 2e-3
    assert abs(errs[1] - errs[0]) < 1e-6


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_potential_linear_matches_reference():
    from h2lib_amd.device import DeviceArray
    mesh, prob = _problem(basis=1)
    rng = np.random.default_rng(9)
    w = rng.standard_normal(mesh.vertices) + 1j * rng.standard_normal(
        mesh.vertices)
    pts = prob.outer_grid()
    got = prob.potential(DeviceArray.from_host(w), pts)
    ref = ReferenceBem(1, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4, k=K)
    want = ref.eval_bw(w, K, pts)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("linear potential error %.2e" % err, flush=True)
    assert err < 1e-12


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_bw_linear_solve_matches_reference():
    mesh, prob = _problem(basis=1)
    d_b = prob.rhs(X0)
    d_x, it, res = prob.solve(d_b, eps=1e-12, maxiter=1000, kmax=50)
    err = prob.outer_error(d_x, X0)
    rslp = ReferenceBem(1, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4, k=K)
    rdlp = ReferenceBem(1, 1, 1, 1, mesh.x, mesh.t, q=2, q2=4, k=K,
                        alpha=0.5)
    V = ReferenceH2(rslp, 16, 1.0, 5, 0)
    KM = ReferenceH2(rdlp, 16, 1.0, 5, 0)
    b_ref = rdlp.helmholtz_point_integrate(K, X0)
    x_ref, it_ref = V.bw_solve(KM, K, b_ref, 1e-12, 1000, 50)
    err_ref = rslp.outer_error(x_ref, K, X0)
    x = d_x.download()
    err_cross = rslp.outer_error(x, K, X0)
    b = d_b.download()
    r = KM.addeval(x, b.copy(), alpha=-1.0)
    r = V.addeval(x, r, alpha=1j * K)
    rel_res = np.linalg.norm(r) / np.linalg.norm(b)
    print("linear bw steps", it, it_ref, "true residual %.2e outer error "
          "%.6e (reference measure %.6e, reference solution %.6e)"
          % (rel_res, err, err_cross, err_ref), flush=True)
    assert np.abs(b - b_ref).max() < 1e-13 * np.abs(b_ref).max()
    assert rel_res < 1e-11
    assert abs(err - err_cross) < 1e-12
    assert abs(err - err_ref) < 1e-7
    assert abs(it - it_ref) <= 2


def test_bw_batched_gmres_equals_single():
    """The batched GMRES on the combined operator KM - i k V (two batched
    products per step, the server's Helmholtz path) gives each system the
    solution of the one-system GMRES on the same operator, to 1e-8.
    Interpolation order 4 (rank 64, the batched product's limit) and eta 2,
    as the server builds it."""
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import (GmresBatch, HelmholtzBrakhageWerner,
                                 solve_gmres)
    prob = HelmholtzBrakhageWerner(G.sphere(8), K, clf=32, eta_adm=2.0, m=4)
    srcs = [[X0], [np.array([0.1, -0.2, 0.0])],
            [X0, np.array([-0.3, 0.1, 0.1])]]
    bs = [prob.rhs_sources(s)[1] for s in srcs]
    alone = []
    for d_b in bs:
        d_x = h.DeviceArray(prob.n, np.complex128)
        d_x.zero()
        it, _ = solve_gmres(prob.A, d_b, d_x, 1e-10, 2000, 30)
        alone.append((d_x.download(), it))
    G = GmresBatch(prob.A, slots=2, kmax=30)
    xs = []
    for _ in bs:
        d_x = h.DeviceArray(prob.n, np.complex128)
        d_x.zero()
        xs.append(d_x)
    pending = list(range(len(bs)))
    done = {}
    rounds = 0
    while len(done) < len(bs):
        while pending and G.free_slots() > 0:
            t = pending.pop(0)
            G.submit(bs[t], xs[t], 1e-10, 2000, t)
        for tag, it, res in G.run():
            done[tag] = it
        rounds += 1
        assert rounds < 10000
    print("bw batch steps", [done[t] for t in range(len(bs))],
          [a[1] for a in alone], flush=True)
    for t in range(len(bs)):
        x = xs[t].download()
        assert np.linalg.norm(x - alone[t][0]) < \
            1e-8 * np.linalg.norm(alone[t][0]), t
        # FP64 atomics in the products: rounding, and near a restart the
        # step count, differ from run to run
        assert abs(done[t] - alone[t][1]) <= max(2, alone[t][1] // 20), t

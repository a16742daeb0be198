"""Conjugate gradients on the device (needs a GPU).

h2hip_solve_cg against the reference's solve_cg_avector
(Library/krylovsolvers.c:18-52) on the same H2 systems:

* the reference's Laplace test problem solved with CG as there
  (Tests/test_laplacebem3d.c:16-46, 129; interior problem, linear basis,
  u = 1/|x - (1.2, 1.2, 1.2)|): steps within 1 of the reference's, the
  solutions to 1e-8 (both stopped at |r| <= 1e-10 |b|; the device product
  sums with atomics, so rounding differs), true residual |b - V x| below
  1.01e-10 |b|, the Neumann error equal to 1e-8;
* the complex recurrence (conjugated inner products) on the Helmholtz
  single layer matrix: the iterates after 10 steps from a start vector
  against the reference's to 1e-9;
* the stopping rule: maxiter - 1 steps, no step from a converged start;
* graph replay against direct launches (same steps within max(2, 2%):
  the products' atomic sums round differently from run to run; solutions
  to 1e-8), also with the self-adjoint product of a matrix kept in symmetric
  storage (exactly symmetric, unlike the full matrix whose near field is
  symmetric only to quadrature accuracy);
* the batched solver's CG mode against the one-system CG (and the same
  batch switched back to GMRES against the one-system GMRES).
"""
import numpy as np
import pytest

from oracle_ffi import (ReferenceBem, ReferenceH2, have_reference,
                        reference_dirichlet)

pytestmark = pytest.mark.gpu

X0 = np.array([1.2, 1.2, 1.2])


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("use_graph", [False, True])
def test_laplace_dirichlet_cg_matches_reference(use_graph):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    mesh = G.sphere(8)
    prob = LaplaceDirichlet(mesh, clf=16, eta=2.0, m=3)
    d_f, d_b = prob.rhs(X0)
    d_x, it, res = prob.solve(d_b, eps=1e-10, maxiter=2000, solver="cg",
                              use_graph=use_graph)
    x = d_x.download()
    f_ref, b_ref, _, _, _ = reference_dirichlet(mesh.x, mesh.t, 16, 2.0, 3,
                                                1e-10, 2000, 30)
    slp = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    V = ReferenceH2(slp, 16, 2.0, 3, 0)
    x_ref, it_ref = V.cg(b_ref, np.zeros(prob.n), 1e-10, 2000)
    b = d_b.download()
    r = V.addeval(x, b.copy(), alpha=-1.0)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print("cg graph", use_graph, "steps", it, it_ref, "true residual %.2e"
          % rel, flush=True)
    assert np.abs(b - b_ref).max() < 1e-12 * np.abs(b_ref).max()
    assert abs(it - it_ref) <= 1
    assert rel < 1.01e-10
    assert abs(res - np.linalg.norm(r)) < 1e-3 * res + 1e-14
    assert np.abs(x - x_ref).max() < 1e-8 * np.abs(x_ref).max()
    assert abs(prob.neumann_error(x, X0) -
               prob.neumann_error(x_ref, X0)) < 1e-8


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_complex_cg_iterates_match_reference():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d, new_slp_helmholtz_bem3d
    from h2lib_amd.h2matrix import H2Matrix
    from h2lib_amd.solve import solve_cg
    k = 2.0
    mesh = G.sphere(8)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = new_slp_helmholtz_bem3d(k, gr, 2, 4, 0, 0)
    rt = G.ClusterTree(mesh, 0, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    n = mesh.triangles
    rng = np.random.default_rng(6)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    d_b = h.DeviceArray.from_host(b)
    d_x = h.DeviceArray.from_host(x0)
    it, _ = solve_cg(H, d_b, d_x, 1e-30, 11)
    ref = ReferenceBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=k)
    Hr = ReferenceH2(ref, 16, 2.0, 3, 0)
    want, it_ref = Hr.cg(b, x0, 1e-30, 11)
    got = d_x.download()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("complex cg iterate error %.2e" % err, flush=True)
    assert it == it_ref == 10
    assert err < 1e-9


def test_stopping_rule():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet, solve_cg
    mesh = G.sphere(8)
    prob = LaplaceDirichlet(mesh, clf=16, eta=2.0, m=3)
    b = np.random.default_rng(2).standard_normal(prob.n)
    d_b = h.DeviceArray.from_host(b)
    for maxiter, want in ((1, 0), (2, 1), (5, 4), (9, 8), (10, 9), (17, 16)):
        d_x = h.DeviceArray(prob.n)
        d_x.zero()
        it, _ = solve_cg(prob.V, d_b, d_x, 1e-30, maxiter)
        assert it == want
        if maxiter == 1:
            assert not d_x.download().any()
    # a converged start vector takes no step and stays as it is
    d_x = h.DeviceArray(prob.n)
    d_x.zero()
    solve_cg(prob.V, d_b, d_x, 1e-12, 1000)
    xs = d_x.download()
    it, res = solve_cg(prob.V, d_b, d_x, 1e-6, 1000)
    assert it == 0
    assert np.array_equal(d_x.download(), xs)


def test_symmetric_storage_and_graph():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import (BASIS_LINEAR_BEM3D, Surface3d,
                                 new_slp_laplace_bem3d)
    from h2lib_amd.h2matrix import H2Matrix
    from h2lib_amd.solve import solve_cg
    mesh = G.sphere(16)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    lin = BASIS_LINEAR_BEM3D
    op = new_slp_laplace_bem3d(gr, 2, 4, lin, lin)
    tree = G.ClusterTree(mesh, lin, 32)
    bt = G.BlockTree(tree, tree, 2.0, 0)
    n = mesh.vertices
    b = np.random.default_rng(8).standard_normal(n)
    d_b = h.DeviceArray.from_host(b)
    out = {}
    for symm in (False, True):
        H = H2Matrix(op, mesh, tree, tree, bt, 3, symmetric=symm)
        for g in (False, True):
            d_x = h.DeviceArray(n)
            d_x.zero()
            it, res = solve_cg(H, d_b, d_x, 1e-10, 3000, use_graph=g)
            r = h.DeviceArray.from_host(b)
            H.addeval(-1.0, d_x, r) if not symm else H.addevalsymm(-1.0,
                                                                   d_x, r)
            rel = np.linalg.norm(r.download()) / np.linalg.norm(b)
            print("symmetric", symm, "graph", g, "steps", it,
                  "true residual %.2e" % rel, flush=True)
            assert rel < 1.01e-10
            out[symm, g] = (it, d_x.download())
    for symm in (False, True):
        # the products sum with atomics, so the rounding and with it the
        # last steps before the stopping test vary between runs
        (it0, x0), (it1, x1) = out[symm, False], out[symm, True]
        assert abs(it1 - it0) <= max(2, it0 // 50)
        assert np.abs(x1 - x0).max() < 1e-8 * np.abs(x0).max()
    # the two storages hold different matrices: the full one's near field is
    # symmetric only to quadrature accuracy, the symmetric one exactly, and
    # CG (which assumes symmetry) needs fewer steps on the latter
    d = np.abs(out[True, False][1] - out[False, False][1]).max() / \
        np.abs(out[False, False][1]).max()
    print("symmetric vs full storage: steps %d / %d, solutions differ by "
          "%.2e" % (out[True, False][0], out[False, False][0], d), flush=True)
    assert d < 1e-3


def _run_batch(G, bs, xs, eps, maxiters):
    pending = list(range(len(bs)))
    done = {}
    rounds = 0
    while len(done) < len(bs):
        while pending and G.free_slots() > 0:
            t = pending.pop(0)
            G.submit(bs[t], xs[t], eps, maxiters[t], t)
        for tag, it, res in G.run():
            done[tag] = (it, res)
        rounds += 1
        assert rounds < 10000
    return done


def test_batched_cg_equals_single():
    """CG in the batched solver (products of up to four systems at once,
    systems entering as slots free up) against the one-system CG on the same
    operator: solutions to 1e-8 relative (both stop at |r| <= 1e-10 |b|;
    the batched product sums in another order), steps within max(2, 2%);
    the stopping rule per slot (maxiter - 1 steps, no step from a converged
    start); then the same batch switched back to GMRES against the
    one-system GMRES."""
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import (GmresBatch, LaplaceDirichlet, solve_cg,
                                 solve_gmres)
    mesh = G.sphere(8)
    prob = LaplaceDirichlet(mesh, clf=16, eta=2.0, m=3)
    n = prob.n
    rng = np.random.default_rng(12)
    rhs = [rng.standard_normal(n) for _ in range(6)]
    d_bs = [h.DeviceArray.from_host(b) for b in rhs]
    alone = []
    for d_b in d_bs:
        d_x = h.DeviceArray(n)
        d_x.zero()
        it, _ = solve_cg(prob.V, d_b, d_x, 1e-10, 3000)
        alone.append((d_x.download(), it))
    B = GmresBatch(prob.V, slots=4, kmax=30, method="cg")
    xs = [h.DeviceArray(n) for _ in rhs]
    for x in xs:
        x.zero()
    done = _run_batch(B, d_bs, xs, 1e-10, [3000] * len(rhs))
    print("batch cg steps", [done[t][0] for t in range(len(rhs))],
          [a[1] for a in alone], flush=True)
    for t in range(len(rhs)):
        x0, it0 = alone[t]
        assert np.abs(xs[t].download() - x0).max() < \
            1e-8 * np.abs(x0).max(), t
        assert abs(done[t][0] - it0) <= max(2, it0 // 50), t
    # stopping rule: maxiter 1, 2, 10, 17 -> 0, 1, 9, 16 steps
    xs2 = [h.DeviceArray(n) for _ in range(4)]
    for x in xs2:
        x.zero()
    done = _run_batch(B, d_bs[:4], xs2, 1e-30, [1, 2, 10, 17])
    assert [done[t][0] for t in range(4)] == [0, 1, 9, 16]
    assert not xs2[0].download().any()
    # a converged start takes no step and is left as it is
    before = xs[1].download()
    done = _run_batch(B, [d_bs[1]], [xs[1]], 1e-6, [3000])
    assert done[0][0] == 0
    assert np.array_equal(xs[1].download(), before)
    # back to GMRES on the same slots
    B.set_method("gmres")
    d_x = h.DeviceArray(n)
    d_x.zero()
    it_g, _ = solve_gmres(prob.V, d_bs[2], d_x, 1e-10, 3000, 30)
    xg = h.DeviceArray(n)
    xg.zero()
    done = _run_batch(B, [d_bs[2]], [xg], 1e-10, [3000])
    # reproducible mode: the slot repeats the one-system solver bit for bit
    assert done[0][0] == it_g
    assert np.array_equal(xg.download(), d_x.download())

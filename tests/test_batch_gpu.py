"""Batched H2 products and batched GMRES (needs a GPU).

The batched product y_v += alpha_v H x_v for several vectors at once must
give the single products (1e-12 relative) and, for complex data, the CPU
restatement of the reference's product (1e-12).  Batched GMRES: every
solve gives the same solution alone (h2hip_solve_gmres) and in the mix of
concurrent solves that enter and leave the slots at different times
(1e-8 relative at eps = 1e-10), with the same number of steps up to one.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle, OracleBem, oracle_nearfield_packed

pytestmark = pytest.mark.gpu


def real_h2(split=8, clf=16, m=3, layer=0):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(split)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = (h.new_slp_laplace_bem3d(gr, 2, 4, 1, 1) if layer == 0 else
          h.new_dlp_laplace_bem3d(gr, 2, 4, 1, 1, 0.5))
    rt = G.ClusterTree(mesh, 1, clf)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    return mesh, gr, op, rt, H2Matrix(op, mesh, rt, rt, bt, m)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nvec", [2, 3, 8])
@pytest.mark.parametrize("trans", [False, True])
def test_batch_product_matches_single(nvec, trans):
    import h2lib_amd as h
    mesh, gr, op, rt, H = real_h2()
    n = rt.n
    rng = np.random.default_rng(nvec)
    xs = [rng.uniform(-1, 1, n) for _ in range(nvec)]
    y0 = [rng.uniform(-1, 1, n) for _ in range(nvec)]
    alphas = rng.uniform(-2, 2, nvec)
    d_x = [h.DeviceArray.from_host(x) for x in xs]
    d_y = [h.DeviceArray.from_host(y) for y in y0]
    print("batch product", nvec, trans, flush=True)
    H.addeval_batch(alphas, d_x, d_y, trans=trans)
    for v in range(nvec):
        want = h.DeviceArray.from_host(y0[v])
        H.addeval(alphas[v], d_x[v], want, trans=trans)
        assert rel(d_y[v].download(), want.download()) < 1e-12, v


def test_complex_products_match_restatement():
    # a complex H2-matrix from real bases, complex near field and couplings:
    # single and batched GPU products against oracle/oracle_h2.c
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix, Interpolation
    mesh = G.sphere(8)
    rt = G.ClusterTree(mesh, 1, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    I = Interpolation(mesh, rt, rt, bt, 1, 1, False, 3, 2)
    ob = OracleBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    N = oracle_nearfield_packed(ob, bt.near_block_list(), rt.idx, rt.idx)
    rng = np.random.default_rng(3)
    Nc = N * (1.0 + 0.3j)
    Sc = I.S * (0.7 - 0.2j) + 1e-3j * rng.standard_normal(I.S.shape)
    Hc = H2Matrix.from_arrays(I, Nc, Sc, rt=rt, is_complex=True)
    n = rt.n
    xs = [rng.standard_normal(n) + 1j * rng.standard_normal(n)
          for _ in range(3)]
    alphas = np.array([0.5 - 1j, -0.25 + 0.1j, 2.0])
    for trans in (False, True):
        want = [Oracle.h2_addeval(I, Nc, x, np.zeros(n, complex), alpha=a,
                                  trans=trans, S=Sc)
                for x, a in zip(xs, alphas)]
        d_x = [h.DeviceArray.from_host(x) for x in xs]
        d_y = [h.DeviceArray.from_host(np.zeros(n, complex)) for _ in xs]
        Hc.addeval_batch(alphas, d_x, d_y, trans=trans)
        for v in range(3):
            assert rel(d_y[v].download(), want[v]) < 1e-12, (trans, v)
        d1 = h.DeviceArray.from_host(np.zeros(n, complex))
        Hc.addeval(alphas[0], d_x[0], d1, trans=trans)
        assert rel(d1.download(), want[0]) < 1e-12, trans


def test_gmres_batch_same_alone_and_in_mix():
    import h2lib_amd as h
    from h2lib_amd.solve import GmresBatch, solve_gmres
    mesh, gr, op, rt, H = real_h2(split=10, clf=24, m=3)
    n = rt.n
    rng = np.random.default_rng(9)
    rhs = [rng.standard_normal(n) for _ in range(6)]
    alone = []
    for b in rhs:
        d_b = h.DeviceArray.from_host(b)
        d_x = h.DeviceArray.from_host(np.zeros(n))
        it, res = solve_gmres(H, d_b, d_x, 1e-10, 2000, 30)
        alone.append((d_x.download(), it))
    G = GmresBatch(H, slots=4, kmax=30)
    d_bs = [h.DeviceArray.from_host(b) for b in rhs]
    d_xs = [h.DeviceArray.from_host(np.zeros(n)) for _ in rhs]
    pending = list(range(len(rhs)))
    done = {}
    # four start together, the others enter as slots free up
    rounds = 0
    while len(done) < len(rhs):
        while pending and G.free_slots() > 0:
            t = pending.pop(0)
            G.submit(d_bs[t], d_xs[t], 1e-10, 2000, t)
        for tag, it, res in G.run():
            done[tag] = it
        rounds += 1
        assert rounds < 10000
    print("batch gmres steps", [done[t] for t in range(len(rhs))],
          [a[1] for a in alone], flush=True)
    for t in range(len(rhs)):
        x_alone, it_alone = alone[t]
        # reproducible mode: every slot's arithmetic is the one-system
        # solver's, so the same steps and the same bits
        assert done[t] == it_alone, t
        assert np.array_equal(d_xs[t].download(), x_alone), t


@pytest.mark.parametrize("kern", [0, 1])
def test_batch_product_computed_couplings(kern):
    """Batched products with the couplings computed from the interpolation
    points (each value shared by all vectors) against the batched products
    reading the stored couplings: 1e-13, real (Laplace) and complex
    (Helmholtz, k = 1.5) kernels, both transform schedules."""
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(8)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = (h.new_slp_laplace_bem3d(gr, 2, 4, 1, 1) if kern == 0 else
          h.new_slp_helmholtz_bem3d(1.5, gr, 2, 4, 0, 0))
    rt = G.ClusterTree(mesh, 1 if kern == 0 else 0, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    n = rt.n
    rng = np.random.default_rng(21 + kern)
    for nvec in (2, 5, 8):
        xs = [rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n)
                                        if kern else 0) for _ in range(nvec)]
        alphas = rng.uniform(-2, 2, nvec) + (0.5j if kern else 0)
        d_x = [h.DeviceArray.from_host(x) for x in xs]
        for sched in (0, 2):
            H.set_schedule(sched)
            got = []
            for mode in (0, 1):
                H.set_coupling_mode(mode)
                d_y = [h.DeviceArray.from_host(np.zeros_like(x)) for x in xs]
                H.addeval_batch(alphas, d_x, d_y)
                got.append([d.download() for d in d_y])
            H.set_coupling_mode(0)
            print("computed couplings batch", kern, nvec, sched, flush=True)
            for v in range(nvec):
                assert rel(got[1][v], got[0][v]) < 1e-13, (nvec, sched, v)

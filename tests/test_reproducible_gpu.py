"""Reproducible products and the adjoint's flat transforms (needs a GPU).

* In reproducible mode (fixed-order sums) a product gives the same bits on
  every call, and each vector of a batched product the bits of the single
  product, for every transform schedule, both coupling modes, the adjoint,
  real and complex matrices.
* The reproducible products agree with the atomic ones to 1e-13 relative
  (the same sums in another order), and the flat adjoint (the other
  orientation of the blocks W_a|_l, shared between the sides when the row
  and column bases are equal) with the level-synchronous adjoint.
* Operators that are not Hermitian cannot be given to the batched CG.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def matrix(kind):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(10)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if kind == "lap_slp":
        op = h.new_slp_laplace_bem3d(gr, 2, 4, 1, 1)
        basis = 1
    elif kind == "lap_dlp":
        op = h.new_dlp_laplace_bem3d(gr, 2, 4, 1, 1, 0.5)
        basis = 1
    else:
        op = h.new_slp_helmholtz_bem3d(2.0 + 0.25j, gr, 2, 4, 0, 0)
        basis = 0
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    return H2Matrix(op, mesh, rt, rt, bt, 3), gr, op


def vectors(H, count, seed):
    rng = np.random.default_rng(seed)
    cplx = H.dtype == np.complex128
    out = []
    for _ in range(count):
        v = rng.uniform(-1, 1, H.n)
        if cplx:
            v = v + 1j * rng.uniform(-1, 1, H.n)
        out.append(v)
    return out


def products(H, xs, alphas, trans, batch):
    import h2lib_amd as h
    d_x = [h.DeviceArray.from_host(x) for x in xs]
    d_y = [h.DeviceArray.from_host(0.5 * x) for x in xs]
    if batch:
        H.addeval_batch(alphas, d_x, d_y, trans=trans)
    else:
        for a, x, y in zip(alphas, d_x, d_y):
            H.addeval(a, x, y, trans=trans)
    h.hip().h2hip_stream_sync(None)
    return [d.download() for d in d_y]


@pytest.mark.parametrize("kind", ["lap_slp", "lap_dlp", "helm_slp"])
@pytest.mark.parametrize("sched", [0, 1, 2])
def test_reproducible_bits_and_agreement(kind, sched):
    import h2lib_amd as h
    H, gr, op = matrix(kind)
    H.set_schedule(sched)
    xs = vectors(H, 5, 7)
    alphas = [1.0, -0.5, 2.0, 0.25, -1.5]
    if H.dtype == np.complex128:
        alphas = [a + 0.5j for a in alphas]
    for trans in (False, True):
        for mode in (0, 1):
            H.set_coupling_mode(mode)
            print(kind, "schedule", sched, "trans", trans, "couplings", mode,
                  flush=True)
            h.set_reproducible(True)
            try:
                one = products(H, xs, alphas, trans, False)
                again = products(H, xs, alphas, trans, False)
                many = products(H, xs, alphas, trans, True)
                two = products(H, xs[:2], alphas[:2], trans, True)
            finally:
                h.set_reproducible(True)
            for v in range(len(xs)):
                assert np.array_equal(one[v], again[v]), v
                assert np.array_equal(one[v], many[v]), v
            for v in range(2):
                assert np.array_equal(one[v], two[v]), v
            h.set_reproducible(False)
            try:
                atom = products(H, xs, alphas, trans, False)
                atom_many = products(H, xs, alphas, trans, True)
            finally:
                h.set_reproducible(True)
            for v in range(len(xs)):
                assert rel(atom[v], one[v]) < 1e-13, v
                assert rel(atom_many[v], one[v]) < 1e-13, v
    H.set_coupling_mode(0)


@pytest.mark.parametrize("kind", ["lap_slp", "lap_dlp", "helm_slp"])
def test_flat_adjoint_matches_level_sync(kind):
    H, gr, op = matrix(kind)
    xs = vectors(H, 3, 11)
    alphas = [1.0, -2.0, 0.5]
    H.set_schedule(1)
    want = products(H, xs, alphas, True, False)
    H.set_schedule(2)
    got = products(H, xs, alphas, True, False)
    got_b = products(H, xs, alphas, True, True)
    for v in range(len(xs)):
        assert rel(got[v], want[v]) < 1e-13, v
        assert rel(got_b[v], want[v]) < 1e-13, v
    # <H x, y> = <x, H^* y> with the flat transforms in both directions
    x, y = xs[0], xs[1]
    Hx = products(H, [x], [1.0], False, False)[0] - 0.5 * x
    Hty = products(H, [y], [1.0], True, False)[0] - 0.5 * y
    lhs, rhs = np.vdot(y, Hx), np.vdot(Hty, x)
    assert abs(lhs - rhs) <= 1e-12 * np.abs(Hx).max() * np.abs(y).sum()


def test_batched_cg_refuses_combined_operators():
    import h2lib_amd as h
    from h2lib_amd.solve import Combined, GmresBatch
    Hs, gr, op = matrix("helm_slp")
    G = GmresBatch(Combined(Hs, 1.0, Hs, -2j), slots=2, kmax=10)
    with pytest.raises(h.H2HipError):
        G.set_method("cg")
    B = GmresBatch(Hs, slots=2, kmax=10, method="cg")
    c = np.array([1.0, 0.0])
    with pytest.raises(h.H2HipError):
        import ctypes as C
        from h2lib_amd._lib import check
        check(h.hip().h2hip_gmres_batch_combine(
            B._h, c.ctypes.data_as(C.c_void_p), Hs._h,
            c.ctypes.data_as(C.c_void_p)), "h2hip_gmres_batch_combine")


@pytest.mark.parametrize("kind", ["lap_slp", "lap_dlp", "helm_slp"])
@pytest.mark.parametrize("sched", [0, 2])
def test_adjoint_from_transposed_copies(kind, sched):
    # the adjoint product from the kept conjugate transposes (forward
    # kernels) against the adjoint kernels: the same sums in another order,
    # one and three vectors, stored and computed couplings; the copies follow
    # a new near field
    import h2lib_amd as h
    H, gr, op = matrix(kind)
    H.set_schedule(sched)
    xs = vectors(H, 3, 11)
    alphas = [1.0, -0.75, 0.5]
    if H.dtype == np.complex128:
        alphas = [a - 0.25j for a in alphas]
    for mode in (0, 1):
        H.set_coupling_mode(mode)
        H.set_adjoint_storage(False)
        want = products(H, xs, alphas, True, False)
        H.set_adjoint_storage(True)
        for batch in (False, True):
            got = products(H, xs, alphas, True, batch)
            for v in range(len(xs)):
                assert rel(got[v], want[v]) < 1e-13, (mode, batch, v)
    H.set_coupling_mode(0)
    # a near field assembled again (mixed precision: entries change by
    # about 1e-8 relative) reaches the adjoint's copies
    before = products(H, xs[:1], alphas[:1], True, False)
    op.set_precision(1)
    H.assemble_nearfield()
    got = products(H, xs[:1], alphas[:1], True, False)
    H.set_adjoint_storage(False)
    want = products(H, xs[:1], alphas[:1], True, False)
    assert rel(got[0], want[0]) < 1e-13
    assert rel(before[0], want[0]) > 1e-11
    op.set_precision(0)


@pytest.mark.parametrize("schedule", [1, 2])
def test_batched_couplings_on_matrix_cores(schedule):
    # products of 2, 5 and 8 vectors with the couplings multiplied on FP64
    # MFMA (rank 64) against the vector kernels: the same sums in another
    # order
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    h.set_reproducible(False)
    try:
        prob = LaplaceDirichlet(G.sphere(16), clf=32, eta=2.0, m=4)
        H = prob.KM
        H.set_schedule(schedule)
        assert len(H.interp.far) > 0
        for nv in (2, 5, 8):
            xs = vectors(H, nv, 30 + nv)
            alphas = [1.0 - 0.25 * v for v in range(nv)]
            H.set_mfma(False)
            want = products(H, xs, alphas, False, True)
            H.set_mfma(True)
            got = products(H, xs, alphas, False, True)
            for v in range(nv):
                assert rel(got[v], want[v]) < 1e-13, (nv, v)
        H.set_mfma(False)
    finally:
        h.set_reproducible(True)

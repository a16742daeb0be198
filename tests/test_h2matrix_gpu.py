"""H2-matrix construction and products on the GPU against the reference
library (needs a GPU).

The reference builds its H2-matrix from the same mesh with the same cluster
tree, block tree and interpolation order; the products y += alpha H x and
y += alpha H^* x must agree to 1e-12 relative (max-norm).  Independent of
the reference: <H x, y> = <x, H^* y>, linearity, and the product of a
matrix without admissible blocks equals the dense near field.
"""
import numpy as np
import pytest

from conftest import golden
from oracle_ffi import ReferenceBem, ReferenceH2, have_reference

pytestmark = pytest.mark.gpu

TOL = 1e-12


def build(mesh, layer, basis, clf, eta, m, alpha=0.5):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if layer == 0:
        op = h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis)
    elif layer == 1:
        op = h.new_dlp_laplace_bem3d(gr, 2, 4, basis, basis, alpha)
    else:
        op = h.new_adlp_laplace_bem3d(gr, 2, 4, basis, basis, alpha)
    rt = G.ClusterTree(mesh, basis, clf)
    bt = G.BlockTree(rt, rt, eta, 0)
    return op, rt, bt, H2Matrix(op, mesh, rt, rt, bt, m)


def product(H, x, trans=False, alpha=1.0, y0=None):
    import h2lib_amd as h
    dx = h.DeviceArray.from_host(x)
    dy = h.DeviceArray.from_host(np.zeros_like(x) if y0 is None else y0)
    H.addeval(alpha, dx, dy, trans=trans)
    return dy.download()


def test_matches_reference_fixture():
    from h2lib_amd import geometry as G
    g = golden("h2_sphere_s8.npz")
    mesh = G.sphere(8)
    op, rt, bt, H = build(mesh, 0, 1, 16, 2.0, 3)
    assert np.array_equal(rt.idx, g["idx"])
    y = product(H, g["x"])
    assert np.abs(y - g["y"]).max() < TOL * np.abs(g["y"]).max()
    yt = product(H, g["x"], trans=True)
    assert np.abs(yt - g["yt"]).max() < TOL * np.abs(g["yt"]).max()


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("layer,basis", [(0, 1), (1, 1), (0, 0), (1, 0),
                                         (2, 1), (2, 0)])
def test_matches_live_reference(layer, basis):
    # layer 2, the adjoint double layer, takes the normal derivative of the
    # Lagrange polynomials on the row basis
    from h2lib_amd import geometry as G
    mesh = G.sphere(16)
    op, rt, bt, H = build(mesh, layer, basis, 32, 2.0, 3)
    ref = ReferenceBem(0, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                       alpha=0.5 if layer else 0.0)
    rh = ReferenceH2(ref, 32, 2.0, 3, 0)
    assert np.array_equal(rh.idx(), rt.idx)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, len(rt.idx))
    y0 = rng.uniform(-1, 1, len(rt.idx))
    print("h2 layer", layer, "basis", basis, flush=True)
    for trans in (False, True):
        want = rh.addeval(x, y0, alpha=-0.75, trans=trans)
        got = product(H, x, trans=trans, alpha=-0.75, y0=y0)
        assert np.abs(got - want).max() < TOL * np.abs(want).max(), trans


def test_adjoint_and_linearity():
    from h2lib_amd import geometry as G
    mesh = G.sphere(12)
    op, rt, bt, H = build(mesh, 1, 1, 24, 1.5, 3)
    rng = np.random.default_rng(8)
    n = len(rt.idx)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Hx = product(H, x)
    Hty = product(H, y, trans=True)
    assert abs(np.dot(Hx, y) - np.dot(x, Hty)) < 1e-12 * np.abs(Hx).max() * \
        np.abs(y).sum()
    z = rng.standard_normal(n)
    Hz = product(H, z)
    Hxz = product(H, 2.0 * x - 3.0 * z)
    assert np.abs(Hxz - (2.0 * Hx - 3.0 * Hz)).max() < 1e-12 * \
        np.abs(Hxz).max()


def test_all_near_equals_dense():
    # eta = 0: no admissible blocks, H is the near field of all leaf pairs,
    # i.e. the dense Galerkin matrix
    from h2lib_amd import geometry as G
    mesh = G.sphere(6)
    op, rt, bt, H = build(mesh, 0, 1, 20, 0.0, 2)
    assert len(bt.far) == 0
    dense = op.nearfield(None, None)
    x = np.random.default_rng(1).standard_normal(dense.shape[1])
    y = product(H, x)
    assert np.abs(y - dense @ x).max() < 1e-13 * np.abs(dense @ x).max()


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
@pytest.mark.parametrize("layer,basis,k", [(0, 0, 2.0), (1, 1, 4.0),
                                           (0, 1, 2.0 + 0.5j), (2, 1, 3.0),
                                           (2, 0, 2.0 + 0.5j)])
def test_helmholtz_matches_live_reference(layer, basis, k):
    # complex H2-matrix of a Helmholtz operator: real interpolation bases,
    # complex couplings (host) and near field (GPU quadrature)
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    alpha = 0.5 + 0.25j if layer else 0.0
    if layer == 0:
        op = h.new_slp_helmholtz_bem3d(k, gr, 2, 4, basis, basis)
    elif layer == 1:
        op = h.new_dlp_helmholtz_bem3d(k, gr, 2, 4, basis, basis, alpha)
    else:
        op = h.new_adlp_helmholtz_bem3d(k, gr, 2, 4, basis, basis, alpha)
    rt = G.ClusterTree(mesh, basis, 24)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    ref = ReferenceBem(1, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                       alpha=alpha, k=k)
    rh = ReferenceH2(ref, 24, 2.0, 3, 0)
    assert np.array_equal(rh.idx(), rt.idx)
    rng = np.random.default_rng(13)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    print("helmholtz h2 layer", layer, "basis", basis, "k", k, flush=True)
    for trans in (False, True):
        want = rh.addeval(x, y0, alpha=-0.75 + 0.5j, trans=trans)
        got = product(H, x, trans=trans, alpha=-0.75 + 0.5j, y0=y0)
        assert np.abs(got - want).max() < TOL * np.abs(want).max(), trans


@pytest.mark.parametrize("kern,layer,basis,k", [(0, 0, 0, 0.0),
                                                 (0, 1, 1, 0.0),
                                                 (0, 2, 1, 0.0),
                                                 (1, 0, 0, 3.0 + 0.5j),
                                                 (1, 1, 1, 2.0),
                                                 (1, 2, 0, 2.0)])
def test_device_couplings_equal_host_couplings(kern, layer, basis, k):
    # couplings computed on the device (h2hip_h2matrix_interp_couplings)
    # against the host restatement (h2b_h2interp_coupling_*), which is
    # pinned to the reference: products agree to 1e-13
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.cube(6)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if kern == 0:
        op = (h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis),
              h.new_dlp_laplace_bem3d(gr, 2, 4, basis, basis, 0.5),
              h.new_adlp_laplace_bem3d(gr, 2, 4, basis, basis, 0.5))[layer]
    else:
        op = (h.new_slp_helmholtz_bem3d(k, gr, 2, 4, basis, basis),
              h.new_dlp_helmholtz_bem3d(k, gr, 2, 4, basis, basis, 0.5),
              h.new_adlp_helmholtz_bem3d(k, gr, 2, 4, basis, basis,
                                         0.5))[layer]
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    Hd = H2Matrix(op, mesh, rt, rt, bt, 3, device_couplings=True)
    Hh = H2Matrix(op, mesh, rt, rt, bt, 3, device_couplings=False)
    rng = np.random.default_rng(4)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n)
    if kern == 1:
        x = x + 1j * rng.uniform(-1, 1, n)
    for trans in (False, True):
        a = product(Hd, x, trans=trans)
        b = product(Hh, x, trans=trans)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), trans


@pytest.mark.parametrize("kern,basis", [(0, 1), (0, 0), (1, 0)])
def test_level_synchronous_sweep_equals_counters(kern, basis):
    # the two schedules of the upward sweep compute the same coefficients;
    # products (single and batched) agree to the rounding of the atomic sums
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = (h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis) if kern == 0 else
          h.new_slp_helmholtz_bem3d(2.0, gr, 2, 4, basis, basis))
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    rng = np.random.default_rng(7)
    n = len(rt.idx)
    xs = [rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kern else 0)
          for _ in range(3)]
    res = {}
    for sched in (0, 1):
        H.set_schedule(sched)
        for trans in (False, True):
            res[sched, trans, "one"] = product(H, xs[0], trans=trans)
            d_x = [h.DeviceArray.from_host(x) for x in xs]
            d_y = [h.DeviceArray.from_host(np.zeros_like(x)) for x in xs]
            H.addeval_batch(np.ones(3), d_x, d_y, trans=trans)
            res[sched, trans, "batch"] = np.concatenate(
                [d.download() for d in d_y])
    for trans in (False, True):
        for kind in ("one", "batch"):
            a, b = res[1, trans, kind], res[0, trans, kind]
            assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), (trans,
                                                                    kind)


@pytest.mark.parametrize("kern,layer,basis,m", [(0, 0, 1, 3), (0, 1, 1, 4),
                                                (1, 0, 0, 4), (1, 1, 0, 3),
                                                (0, 0, 0, 5)])
def test_flat_transforms_equal_counters(kern, layer, basis, m):
    # schedule 2 (explicit bases of the ancestors restricted to the leaves)
    # computes the same products as the counter climb, to the rounding of
    # the atomic sums; the products it does not cover (adjoint, batched)
    # keep working through the level-synchronous sweep
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if kern == 0:
        op = (h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis) if layer == 0
              else h.new_dlp_laplace_bem3d(gr, 2, 4, basis, basis, 0.5))
    else:
        op = (h.new_slp_helmholtz_bem3d(2.0, gr, 2, 4, basis, basis)
              if layer == 0 else
              h.new_dlp_helmholtz_bem3d(2.0, gr, 2, 4, basis, basis, -0.5))
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, m)
    rng = np.random.default_rng(11)
    n = len(rt.idx)
    xs = [rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kern else 0)
          for _ in range(2)]
    res = {}
    for sched in (0, 2):
        H.set_schedule(sched)
        for trans in (False, True):
            res[sched, trans] = product(H, xs[0], trans=trans,
                                        alpha=0.75 - (0.5j if kern else 0))
        if m ** 3 <= 64:  # batched products: ranks of at most 64
            d_x = [h.DeviceArray.from_host(x) for x in xs]
            d_y = [h.DeviceArray.from_host(np.zeros_like(x)) for x in xs]
            H.addeval_batch(np.ones(2), d_x, d_y)
            res[sched, "batch"] = np.concatenate([d.download() for d in d_y])
    keys = ((False,), (True,)) + ((("batch",),) if m ** 3 <= 64 else ())
    for key in keys:
        a, b = res[(2,) + key], res[(0,) + key]
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), key


@pytest.mark.parametrize("kern,layer,basis,m,k", [(0, 0, 1, 4, 0.0),
                                                  (0, 1, 1, 3, 0.0),
                                                  (1, 0, 0, 4, 2.0),
                                                  (1, 1, 0, 3, 3.0 + 0.5j),
                                                  (0, 1, 1, 5, 0.0),
                                                  (1, 1, 0, 5, 3.0 + 0.5j)])
def test_computed_couplings_equal_stored(kern, layer, basis, m, k):
    # products that compute the couplings from the interpolation points
    # equal those that read them (same arithmetic for every entry), with
    # both transform schedules; the adjoint keeps reading them
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if kern == 0:
        op = (h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis) if layer == 0
              else h.new_dlp_laplace_bem3d(gr, 2, 4, basis, basis, 0.5))
    else:
        op = (h.new_slp_helmholtz_bem3d(k, gr, 2, 4, basis, basis)
              if layer == 0 else
              h.new_dlp_helmholtz_bem3d(k, gr, 2, 4, basis, basis, 0.5))
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, m)
    rng = np.random.default_rng(5)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kern else 0)
    alpha = 0.5 + (0.25j if kern else 0)
    want = product(H, x, alpha=alpha)
    want_t = product(H, x, trans=True)
    nb = H.bytes()
    H.set_coupling_mode(1)
    assert H.bytes()["coupling"] == 0 and nb["coupling"] > 0
    for sched in (0, 2):
        H.set_schedule(sched)
        got = product(H, x, alpha=alpha)
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max(), sched
        got_t = product(H, x, trans=True)
        assert np.abs(got_t - want_t).max() < 1e-13 * np.abs(want_t).max()
    # one tree on both sides: the far blocks come in mirrored pairs, and the
    # order-4 and -5 couplings are evaluated once for both blocks of a pair
    # (the flat schedule's forward product); the other products keep one
    # evaluation per block
    assert H.coupling_pairs() == (H.interp.nfar // 2 if m in (4, 5) else 0)
    H.set_schedule(2)
    H.set_coupling_pairs(False)
    off = product(H, x, alpha=alpha)
    H.set_coupling_pairs(True)
    on = product(H, x, alpha=alpha)
    assert np.abs(on - off).max() < 1e-14 * np.abs(off).max()
    H.set_coupling_mode(0)
    assert H.bytes()["coupling"] == nb["coupling"]


@pytest.mark.parametrize("kern,layer,basis,m", [(0, 0, 0, 3), (0, 1, 1, 4),
                                                (0, 0, 1, 5), (1, 1, 0, 3),
                                                (1, 0, 1, 2)])
def test_device_bases_equal_host_bases(kern, layer, basis, m):
    # cluster bases built on the device (leaf matrices with the single-
    # triangle rule, normal derivatives for the double layer's column
    # basis, transfer matrices) give the products of the host-built bases
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    if kern == 0:
        op = (h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis) if layer == 0
              else h.new_dlp_laplace_bem3d(gr, 2, 4, basis, basis, 0.5))
    else:
        op = (h.new_slp_helmholtz_bem3d(2.0, gr, 2, 4, basis, basis)
              if layer == 0 else
              h.new_dlp_helmholtz_bem3d(2.0, gr, 2, 4, basis, basis, -0.5))
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    Hd = H2Matrix(op, mesh, rt, rt, bt, m)
    Hh = H2Matrix(op, mesh, rt, rt, bt, m, device_bases=False)
    rng = np.random.default_rng(13)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kern else 0)
    for trans in (False, True):
        a = product(Hd, x, trans=trans)
        b = product(Hh, x, trans=trans)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max(), trans
    assert Hd.bytes() == Hh.bytes()


def test_flat_transforms_symmetric_storage():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = h.new_slp_laplace_bem3d(gr, 2, 4, 1, 1)
    rt = G.ClusterTree(mesh, 1, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 4, symmetric=True)
    x = np.random.default_rng(2).uniform(-1, 1, len(rt.idx))
    out = {}
    for sched in (0, 2):
        H.set_schedule(sched)
        d_x = h.DeviceArray.from_host(x)
        d_y = h.DeviceArray.from_host(np.zeros_like(x))
        H.addevalsymm(1.0, d_x, d_y)
        out[sched] = d_y.download()
    assert np.abs(out[2] - out[0]).max() < 1e-13 * np.abs(out[0]).max()


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("basis,clf,m", [(1, 16, 3), (0, 24, 3), (1, 32, 5)])
def test_symmetric_product_matches_reference(basis, clf, m):
    # addevalsymm from the blocks on and below the block diagonal only
    # (half the storage) against the reference's
    # addevalsymm_h2matrix_avector and the restatement, 1e-12; m = 5 gives
    # ranks of 125 (two rows per lane in the coupling kernel)
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    from oracle_ffi import Oracle
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = h.new_slp_laplace_bem3d(gr, 2, 4, basis, basis)
    rt = G.ClusterTree(mesh, basis, clf)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    Hs = H2Matrix(op, mesh, rt, rt, bt, m, symmetric=True)
    Hf = H2Matrix(op, mesh, rt, rt, bt, m)
    ref = ReferenceBem(0, 0, basis, basis, mesh.x, mesh.t, q=2, q2=4)
    rh = ReferenceH2(ref, clf, 2.0, m, 0)
    assert np.array_equal(rh.idx(), rt.idx)
    rng = np.random.default_rng(3)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    want = rh.addevalsymm(x, y0, alpha=-0.5)
    dx = h.DeviceArray.from_host(x)
    dy = h.DeviceArray.from_host(y0)
    Hs.addevalsymm(-0.5, dx, dy)
    got = dy.download()
    print("symm basis", basis, "clf", clf, "m", m,
          "err %.2e" % (np.abs(got - want).max() / np.abs(want).max()),
          flush=True)
    assert np.abs(got - want).max() < TOL * np.abs(want).max()
    # storage: about half of the full matrix's near field and couplings
    bs, bf = Hs.bytes(), Hf.bytes()
    assert bs["coupling"] < 0.55 * bf["coupling"]
    assert bs["near"] < 0.6 * bf["near"]
    # the full product of the same (nearly symmetric) matrix is close
    full = product(Hf, x, alpha=-0.5, y0=y0)
    assert np.abs(full - want).max() < 1e-2 * np.abs(want).max()


def test_symmetric_storage_rejects_unsymmetric_operators():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import H2Matrix
    mesh = G.sphere(4)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = h.new_dlp_laplace_bem3d(gr, 2, 4, 1, 1, 0.5)
    rt = G.ClusterTree(mesh, 1, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    with pytest.raises(ValueError):
        H2Matrix(op, mesh, rt, rt, bt, 3, symmetric=True)

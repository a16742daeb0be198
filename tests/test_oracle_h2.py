"""The CPU restatement of the H2-matrix products (oracle/oracle_h2.c),
checked without a GPU.

Pinned against (a) the reference library's addeval_h2matrix_avector /
addevaltrans_h2matrix_avector on the same mesh, trees and interpolation
order, live and through the committed fixture, both to 1e-12 relative;
(b) exact identities: with no admissible block (eta = 0) the product is the
dense near-field product; <H x, y> = <x, H^* y> for real and complex data
(complex entries exercise the conjugation of the adjoint).  A dropped
transfer matrix, a transposed coupling or a missing conjugate fails (a) or
(b).
"""
import numpy as np
import pytest

from conftest import golden
from oracle_ffi import (Oracle, OracleBem, ReferenceBem, ReferenceH2,
                        have_reference, oracle_nearfield_packed)

TOL = 1e-12


def host_h2(mesh, layer, basis, clf, eta, m, alpha=0.5):
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import Interpolation
    rt = G.ClusterTree(mesh, basis, clf)
    bt = G.BlockTree(rt, rt, eta, 0)
    interp = Interpolation(mesh, rt, rt, bt, basis, basis, layer == 1, m, 2)
    ob = OracleBem(0, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                   alpha=alpha if layer else 0.0)
    N = oracle_nearfield_packed(ob, bt.near_block_list(), rt.idx, rt.idx)
    return rt, bt, interp, N, ob


def test_matches_reference_fixture():
    from h2lib_amd import geometry as G
    g = golden("h2_sphere_s8.npz")
    mesh = G.sphere(8)
    rt, bt, interp, N, _ = host_h2(mesh, 0, 1, 16, 2.0, 3)
    assert np.array_equal(rt.idx, g["idx"])
    n = len(rt.idx)
    y = Oracle.h2_addeval(interp, N, g["x"], np.zeros(n))
    assert np.abs(y - g["y"]).max() < TOL * np.abs(g["y"]).max()
    yt = Oracle.h2_addeval(interp, N, g["x"], np.zeros(n), trans=True)
    assert np.abs(yt - g["yt"]).max() < TOL * np.abs(g["yt"]).max()


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("layer,basis", [(0, 1), (1, 1), (0, 0), (1, 0)])
def test_matches_live_reference(layer, basis):
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    rt, bt, interp, N, _ = host_h2(mesh, layer, basis, 16, 2.0, 3)
    ref = ReferenceBem(0, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                       alpha=0.5 if layer else 0.0)
    rh = ReferenceH2(ref, 16, 2.0, 3, 0)
    assert np.array_equal(rh.idx(), rt.idx)
    assert len(bt.far) > 0 and len(bt.near) > 0
    rng = np.random.default_rng(5)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    for trans in (False, True):
        want = rh.addeval(x, y0, alpha=-0.75, trans=trans)
        got = Oracle.h2_addeval(interp, N, x, y0.copy(), alpha=-0.75,
                                trans=trans)
        assert np.abs(got - want).max() < TOL * np.abs(want).max(), trans


def test_all_near_equals_dense():
    from h2lib_amd import geometry as G
    mesh = G.sphere(4)
    rt, bt, interp, N, ob = host_h2(mesh, 1, 1, 12, 0.0, 2)
    assert len(bt.far) == 0
    dense = ob.nearfield()
    x = np.random.default_rng(2).standard_normal(dense.shape[1])
    y = Oracle.h2_addeval(interp, N, x, np.zeros(len(x)))
    assert np.abs(y - dense @ x).max() < 1e-13 * np.abs(dense @ x).max()
    yt = Oracle.h2_addeval(interp, N, x, np.zeros(len(x)), trans=True)
    assert np.abs(yt - dense.T @ x).max() < 1e-13 * np.abs(dense.T @ x).max()


@pytest.mark.parametrize("cplx", [False, True])
def test_adjoint(cplx):
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    rt, bt, interp, N, _ = host_h2(mesh, 1, 1, 16, 2.0, 3)
    n = len(rt.idx)
    rng = np.random.default_rng(11)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    if cplx:
        N = N * (1.0 + 0.3j)
        x = x + 1j * rng.standard_normal(n)
        y = y + 1j * rng.standard_normal(n)
    dt = np.complex128 if cplx else np.float64
    Hx = Oracle.h2_addeval(interp, N, x, np.zeros(n, dtype=dt))
    Hty = Oracle.h2_addeval(interp, N, y, np.zeros(n, dtype=dt), trans=True)
    lhs = np.vdot(y, Hx)
    rhs = np.vdot(Hty, x)
    assert abs(lhs - rhs) < 1e-12 * np.abs(Hx).max() * np.abs(y).sum()
    # the complex near field must enter conjugated in the adjoint
    if cplx:
        Hty_wrong = Oracle.h2_addeval(interp, np.conj(N), y,
                                      np.zeros(n, dtype=dt), trans=True)
        assert abs(lhs - np.vdot(Hty_wrong, x)) > 1e-6 * abs(lhs)


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
@pytest.mark.parametrize("layer,basis,k", [(0, 0, 2.0), (1, 1, 3.0),
                                           (0, 1, 2.0 + 0.5j), (1, 0, 2.0)])
def test_helmholtz_couplings_match_live_reference(layer, basis, k):
    # the host Helmholtz couplings (h2b_h2interp_coupling_helmholtz) with
    # real bases and the complex near field: the restated product against
    # the reference's complex H2-matrix of the same operator
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import Interpolation
    mesh = G.sphere(8)
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    interp = Interpolation(mesh, rt, rt, bt, basis, basis, layer == 1, 3, 2)
    alpha = 0.5 if layer else 0.0
    ob = OracleBem(1, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                   alpha=alpha, k=k)
    N = oracle_nearfield_packed(ob, bt.near_block_list(), rt.idx, rt.idx)
    S = interp.helmholtz_couplings(k)
    ref = ReferenceBem(1, layer, basis, basis, mesh.x, mesh.t, q=2, q2=4,
                       alpha=alpha, k=k)
    rh = ReferenceH2(ref, 16, 2.0, 3, 0)
    assert np.array_equal(rh.idx(), rt.idx)
    rng = np.random.default_rng(8)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    for trans in (False, True):
        want = rh.addeval(x, y0, alpha=0.5 - 0.25j, trans=trans)
        got = Oracle.h2_addeval(interp, N, x, y0.copy(), alpha=0.5 - 0.25j,
                                trans=trans, S=S)
        assert np.abs(got - want).max() < TOL * np.abs(want).max(), trans


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("basis", [0, 1])
def test_symmetric_product_matches_reference(basis):
    # addevalsymm_h2matrix_avector on the reference's single layer matrix
    # (self-adjoint) against the restatement from the lower block triangle;
    # mathematics: for a symmetric matrix it equals the full product up to
    # the approximation's asymmetry (the interpolation H2-matrix of the
    # symmetric kernel is symmetric when row and column bases coincide,
    # the quadrature only nearly so)
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    rt, bt, interp, N, _ = host_h2(mesh, 0, basis, 16, 2.0, 3)
    ref = ReferenceBem(0, 0, basis, basis, mesh.x, mesh.t, q=2, q2=4)
    rh = ReferenceH2(ref, 16, 2.0, 3, 0)
    rng = np.random.default_rng(11)
    n = len(rt.idx)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    want = rh.addevalsymm(x, y0, alpha=0.75)
    got = Oracle.h2_addeval(interp, N, x, y0.copy(), alpha=0.75, symm=True)
    assert np.abs(got - want).max() < TOL * np.abs(want).max()
    full = Oracle.h2_addeval(interp, N, x, y0.copy(), alpha=0.75)
    assert np.abs(got - full).max() < 1e-2 * np.abs(full).max()


@pytest.mark.parametrize("basis,m", [(0, 3), (1, 4), (1, 5)])
def test_interp_points_equal_host_build(basis, m):
    # the points handed to the device basis construction are those of the
    # host construction, bit for bit, and lie in the (widened) boxes
    from h2lib_amd import geometry as G
    from h2lib_amd.h2matrix import Interpolation
    mesh = G.sphere(8)
    rt = G.ClusterTree(mesh, basis, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    full = Interpolation(mesh, rt, rt, bt, basis, basis, False, m, 2,
                         couplings=False)
    pts = Interpolation(mesh, rt, rt, bt, basis, basis, False, m, 2,
                        couplings=False, bases=False)
    assert np.array_equal(full.rpts, pts.rpts)
    assert np.array_equal(full.cpts, pts.cpts)
    p = pts.rpts.reshape(rt.nnodes, 3, m)
    lo = np.minimum(rt.bmin, rt.bmax - 5e-3)[:, :, None] - 5e-3
    hi = np.maximum(rt.bmax, rt.bmin + 5e-3)[:, :, None] + 5e-3
    assert np.all(p > lo) and np.all(p < hi)
    # Chebyshev nodes of the first kind are symmetric about the centre
    c = 0.5 * (p[:, :, 0] + p[:, :, -1])
    assert np.allclose(c, 0.5 * (p[:, :, 1] + p[:, :, -2]), atol=1e-14)

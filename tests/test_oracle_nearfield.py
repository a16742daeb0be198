"""Pins the oracle's near-field integrators (CPU only).

* against dense matrices the reference library wrote (golden fixtures,
  s = 4 sphere, all kernel / basis combinations the fixtures hold), and
  against single entries of every rule class on the s = 16 sphere;
* against the live reference library when it is present (random index
  subsets, transposed storage, far variant);
* against identities of potential theory on the unit sphere that do not
  depend on either implementation: the single layer of unit density equals 1
  on the sphere, the double layer of unit density equals -1/2 there, and the
  Helmholtz single layer of unit density equals e^{ik} sin(k)/k.
"""
import numpy as np
import pytest

from conftest import golden, rel_block_err, subsets
from oracle_ffi import OracleBem, ReferenceBem, have_reference

CASES = {
    # name: kernel, layer, row basis, column basis, alpha, k
    "lap_slp_cc": (0, 0, 0, 0, 0.0, 0.0),
    "lap_dlp_cc": (0, 1, 0, 0, 0.5, 0.0),
    "lap_slp_ll": (0, 0, 1, 1, 0.0, 0.0),
    "lap_dlp_ll": (0, 1, 1, 1, -0.5, 0.0),
    "lap_slp_cl": (0, 0, 0, 1, 0.0, 0.0),
    "lap_dlp_lc": (0, 1, 1, 0, 0.5, 0.0),
    "helm_slp_cc": (1, 0, 0, 0, 0.0, 2.0),
    "helm_dlp_cc": (1, 1, 0, 0, 0.5, 2.0),
    "helm_slp_ll": (1, 0, 1, 1, 0.0, 2.0),
    "helm_dlp_ll": (1, 1, 1, 1, 0.5, 2.0),
    # complex alpha: the reference keeps it for cc but holds alpha g in a
    # real variable for the linear bases (only Re(alpha) reaches the mass
    # term, Library/bem3d.c:1822-1823, 2418-2419, 3061-3062)
    "helm_dlp_cc_calpha": (1, 1, 0, 0, 0.5 + 0.25j, 3.0),
    "helm_dlp_ll_calpha": (1, 1, 1, 1, 0.5 + 0.25j, 3.0),
    "helm_dlp_cl_calpha": (1, 1, 0, 1, -0.5 + 0.5j, 2.0),
    "helm_dlp_lc_calpha": (1, 1, 1, 0, 0.25j, 2.0),
    # the adjoint double layer (layer 2); no dense fixture holds it, the
    # live reference and the transpose identity below pin it
    "lap_adlp_cc": (0, 2, 0, 0, 0.5, 0.0),
    "lap_adlp_ll": (0, 2, 1, 1, -0.5, 0.0),
    "lap_adlp_cl": (0, 2, 0, 1, 0.5, 0.0),
    "lap_adlp_lc": (0, 2, 1, 0, 0.5, 0.0),
    "helm_adlp_cc": (1, 2, 0, 0, 0.5, 2.0),
    "helm_adlp_ll": (1, 2, 1, 1, -0.5, 2.0),
    "helm_adlp_lc_calpha": (1, 2, 1, 0, 0.5 + 0.25j, 3.0),
}


def make(cls, mesh, name):
    kern, op, br, bc, alpha, k = CASES[name]
    return cls(kern, op, br, bc, mesh["x"], mesh["t"], q=2, q2=4,
               alpha=alpha, k=k)


@pytest.mark.parametrize("name", [n for n in sorted(CASES)
                                  if not n.endswith("_calpha")
                                  and "adlp" not in n])
def test_dense_against_reference_fixture(name, sphere4):
    g = golden("dense_sphere_s4.npz")
    o = make(OracleBem, sphere4, name)
    assert rel_block_err(o.nearfield(), g[name]) < 1e-13
    assert rel_block_err(o.nearfield(far=True), g[name + "_far"]) < 1e-13


def test_sampled_entries_all_classes():
    m = golden("mesh_sphere_s16.npz")
    s = golden("sampled_sphere_s16.npz")
    rows, cols = s["rows"], s["cols"]
    for name, kern, op, alpha, k in [("lap_slp", 0, 0, 0.0, 0.0),
                                     ("lap_dlp", 0, 1, 0.5, 0.0),
                                     ("helm_slp", 1, 0, 0.0, 4.0)]:
        o = OracleBem(kern, op, 0, 0, m["x"], m["t"], q=2, q2=4, alpha=alpha,
                      k=k)
        got = np.array([o.nearfield(np.array([i]), np.array([j]))[0, 0]
                        for i, j in zip(rows, cols)])
        ref = s[name]
        # per-entry relative error, every class (64 entries each)
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        # DLP entries of coincident flat triangles vanish up to rounding
        tiny = np.abs(ref) < 1e-14 * np.abs(ref).max()
        assert np.all(err[~tiny] < 1e-12), name


@pytest.mark.skipif(not (have_reference(False) and have_reference(True)),
                    reason="reference library not built")
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("far", [False, True])
@pytest.mark.parametrize("ntrans", [False, True])
def test_subsets_against_live_reference(name, far, ntrans, sphere8):
    o = make(OracleBem, sphere8, name)
    r = make(ReferenceBem, sphere8, name)
    rng = np.random.default_rng(11)
    nr = o.ndof(CASES[name][2])
    nc = o.ndof(CASES[name][3])
    for rows, cols in [(37, 29), (64, 64), (5, 3), (1, 90)]:
        ridx, cidx = subsets(rng, nr, nc, rows, cols, nr == nc)
        a = o.nearfield(ridx, cidx, far=far, ntrans=ntrans)
        b = r.nearfield(ridx, cidx, far=far, ntrans=ntrans)
        assert rel_block_err(a, b) < 1e-13, (rows, cols)
    # the diagonal block of a cluster: equal row and column index sets
    idx = rng.permutation(nr)[:48]
    if nr == nc:
        a = o.nearfield(idx, idx, far=far, ntrans=ntrans)
        b = r.nearfield(idx, idx, far=far, ntrans=ntrans)
        assert rel_block_err(a, b) < 1e-13


def test_single_layer_of_unit_density(sphere8):
    # int_Gamma 1/(4 pi |x-y|) dS_y = 1 on the unit sphere, so the rows of
    # the cc Galerkin matrix sum to the triangle areas (up to O(h^2) from
    # the flat triangulation)
    o = make(OracleBem, sphere8, "lap_slp_cc")
    V = o.nearfield()
    area = o.gram() / 2.0
    rel = V.sum(axis=1) / area - 1.0
    assert np.abs(rel).max() < 3e-2
    # a dropped or misplaced term would break the symmetry of V
    assert np.abs(V - V.T).max() < 1e-3 * np.abs(V).max()


@pytest.mark.parametrize("kern,k", [(0, 0.0), (1, 2.0)])
@pytest.mark.parametrize("br,bc", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_adjoint_double_layer_is_the_transposed_double_layer(kern, k, br, bc,
                                                             sphere8):
    # <y - x, n_x> / |x - y|^3 with x on the row triangle is the double
    # layer kernel with the roles of the triangles swapped, so the Galerkin
    # matrices satisfy K'_{ij} = K_{ji} (rows and columns of the bases
    # swapped, same alpha: the mass matrices transpose too).  The regular
    # Gauss rule is the same on both triangles, so pairs that share no
    # vertex agree to rounding; the Sauter-Schwab rules are not symmetric
    # in the two triangles, so touching pairs agree to quadrature accuracy.
    # Taking n_y for n_x, or the wrong sign, breaks the first bound.
    x, t = sphere8["x"], sphere8["t"]
    A = OracleBem(kern, 2, br, bc, x, t, q=2, q2=4, alpha=0.5,
                  k=k).nearfield()
    K = OracleBem(kern, 1, bc, br, x, t, q=2, q2=4, alpha=0.5,
                  k=k).nearfield()
    D = np.abs(A - K.T)
    scale = np.abs(A).max()
    assert D.max() < 2e-5 * scale
    if br == 0 and bc == 0:
        tt = t.astype(np.int64)
        touch = (tt[:, None, :, None] == tt[None, :, None, :]).any(axis=(2, 3))
        assert D[~touch].max() < 1e-14 * scale
        # the swap is not an identity: the two layers differ on this mesh
        assert np.abs(A - K)[~touch].max() > 1e-6 * scale


def test_double_layer_of_unit_density(sphere8):
    # int_Gamma <x-y, n_y>/(4 pi |x-y|^3) dS_y = -1/2 on the sphere, so
    # K + M/2 (alpha = 1/2) annihilates constants
    o = make(OracleBem, sphere8, "lap_dlp_cc")
    K = o.nearfield()
    area = o.gram() / 2.0
    assert np.abs(K.sum(axis=1) / area).max() < 3e-2
    # with alpha = -1/2 the rows sum to -area
    o2 = OracleBem(0, 1, 0, 0, sphere8["x"], sphere8["t"], q=2, q2=4,
                   alpha=-0.5)
    rs = o2.nearfield().sum(axis=1) / area
    assert np.abs(rs + 1.0).max() < 3e-2


@pytest.mark.parametrize("k", [1.0, 2.0])
def test_helmholtz_single_layer_of_unit_density(k, sphere8):
    # addition theorem, n = 0 term: int e^{ik|x-y|}/(4 pi |x-y|) dS_y
    # = e^{ik} sin(k)/k on the unit sphere
    o = OracleBem(1, 0, 0, 0, sphere8["x"], sphere8["t"], q=2, q2=4, k=k)
    V = o.nearfield()
    area = o.gram() / 2.0
    want = np.exp(1j * k) * np.sin(k) / k
    rel = V.sum(axis=1) / area / want - 1.0
    assert np.abs(rel).max() < 3e-2


def test_linear_basis_single_layer_of_unit_density(sphere8):
    # sum_j V_ij = int phi_i = (area of the support of vertex i) / 3
    o = make(OracleBem, sphere8, "lap_slp_ll")
    V = o.nearfield()
    t = sphere8["t"].astype(np.int64)
    area = o.gram() / 2.0
    supp = np.zeros(len(sphere8["x"]))
    for c in range(3):
        np.add.at(supp, t[:, c], area / 3.0)
    assert np.abs(V.sum(axis=1) / supp - 1.0).max() < 3e-2


@pytest.mark.skipif(not have_reference(fast=True),
                    reason="fast reference build absent")
@pytest.mark.parametrize("name", ["lap_slp_ll", "lap_dlp_ll", "lap_slp_cc"])
def test_fast_reference_build_matches(name, sphere8):
    """The reference built for speed (the benchmark's CPU baseline, -O3 with
    contracted multiply-adds) gives the entries of the -O2 build to 1e-13
    of each block's largest entry, so its timing is of the same work."""
    kern, op, br, bc, alpha, k = CASES[name]
    a = ReferenceBem(kern, op, br, bc, sphere8["x"], sphere8["t"], q=2, q2=4,
                     alpha=alpha, k=k)
    b = ReferenceBem(kern, op, br, bc, sphere8["x"], sphere8["t"], q=2, q2=4,
                     alpha=alpha, k=k, fast=True)
    assert rel_block_err(b.nearfield(), a.nearfield()) < 1e-13

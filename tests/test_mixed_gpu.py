"""Mixed-precision near-field quadrature against the FP64 kernels (needs a
GPU).

Mixed precision evaluates the kernel functions in FP32 (hardware 1/sqrt,
sine, cosine and 2^x after reduction of the phase to [-1/2, 1/2]
revolutions), forms x - y of singular pairs and the Sauter-Schwab points in
FP64 and sums the outer rule dimension in FP64.  The FP64 kernels are pinned
to the oracle at 1e-12 elsewhere; here every kernel / basis pair of the
mixed kernels stays within 2e-6 of them relative to the block maximum
(observed about 2e-7), for near and far rules, plain and transposed, on a
dense matrix and on index subsets.  The error is recorded per case.
"""
import numpy as np
import pytest

from conftest import rel_block_err, subsets

pytestmark = pytest.mark.gpu

TOL = 2e-6

CASES = {
    "lap_slp_cc": (0, 0, 0, 0, 0.0, 0.0),
    "lap_dlp_cc": (0, 1, 0, 0, 0.5, 0.0),
    "lap_slp_ll": (0, 0, 1, 1, 0.0, 0.0),
    "lap_dlp_ll": (0, 1, 1, 1, -0.5, 0.0),
    "lap_dlp_lc": (0, 1, 1, 0, 0.5, 0.0),
    "helm_slp_cc": (1, 0, 0, 0, 0.0, 2.0),
    "helm_slp_cc_k8": (1, 0, 0, 0, 0.0, 8.0),
    "helm_dlp_cc": (1, 1, 0, 0, 0.5 + 0.25j, 4.0),
    "helm_slp_ll": (1, 0, 1, 1, 0.0, 2.0 + 0.5j),
    "helm_dlp_ll": (1, 1, 1, 1, 0.5, 2.0),
    "helm_slp_cl": (1, 0, 0, 1, 0.0, 3.0),
}


def pair(mesh, name, q=2, q2=4, mode=1):
    import h2lib_amd as h
    kern, op, br, bc, alpha, k = CASES[name]
    gr = h.Surface3d(mesh["x"], mesh["t"])
    a = h.Bem3d(gr, kern, op, q, q2, br, bc, alpha=alpha, k=k)
    b = h.Bem3d(gr, kern, op, q, q2, br, bc, alpha=alpha, k=k)
    b.set_precision(mode)
    return a, b


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", [1, 2])
def test_mixed_dense_close_to_fp64(name, mode, sphere8):
    """mode 1: FP64 outer sums; mode 2: FP32 accumulation with compensated
    outer sums (the same bound; the kernel values' FP32 rounding
    dominates)."""
    a, b = pair(sphere8, name, mode=mode)
    for far in (False, True):
        f64 = a.nearfield_far(None, None) if far else a.nearfield(None, None)
        mp = b.nearfield_far(None, None) if far else b.nearfield(None, None)
        err = rel_block_err(mp, f64)
        print("mixed", mode, name, "far" if far else "near",
              "err %.3e" % err, flush=True)
        assert err < TOL, far
        assert err > 0.0  # the mixed kernels did run


def test_compensated_sums_differ_from_mode1(sphere8):
    """Mode 2 runs its own kernels: its entries differ from mode 1's in the
    last bits of the FP32 sums, and both stay within the bound."""
    a, b1 = pair(sphere8, "lap_slp_ll", mode=1)
    _, b2 = pair(sphere8, "lap_slp_ll", mode=2)
    f64 = a.nearfield(None, None)
    m1, m2 = b1.nearfield(None, None), b2.nearfield(None, None)
    assert not np.array_equal(m1, m2)
    assert rel_block_err(m2, f64) < TOL and rel_block_err(m1, f64) < TOL


@pytest.mark.parametrize("name", ["lap_slp_ll", "helm_dlp_cc", "lap_dlp_lc",
                                  "helm_slp_cl"])
@pytest.mark.parametrize("ntrans", [False, True])
def test_mixed_subsets_close_to_fp64(name, ntrans, sphere8):
    a, b = pair(sphere8, name)
    rng = np.random.default_rng(9)
    kern, op, br, bc, alpha, k = CASES[name]
    nr = len(sphere8["x"]) if br else len(sphere8["t"])
    nc = len(sphere8["x"]) if bc else len(sphere8["t"])
    for rows, cols in [(37, 29), (65, 130), (1, 90)]:
        ridx, cidx = subsets(rng, nr, nc, rows, cols, nr == nc)
        for far in (False, True):
            fa = a.nearfield_far if far else a.nearfield
            fb = b.nearfield_far if far else b.nearfield
            err = rel_block_err(fb(ridx, cidx, ntrans=ntrans),
                                fa(ridx, cidx, ntrans=ntrans))
            assert err < TOL, (rows, cols, far)


def test_precision_switch_back_to_fp64(sphere8):
    # switching back gives the FP64 entries bit for bit
    a, b = pair(sphere8, "helm_slp_cc")
    b.set_precision(0)
    assert np.array_equal(a.nearfield(None, None), b.nearfield(None, None))


def test_invalid_precision_mode_is_rejected(sphere8):
    a, _ = pair(sphere8, "lap_slp_cc")
    with pytest.raises(RuntimeError):
        a.set_precision(3)

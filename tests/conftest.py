import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(TESTS, "golden")
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(autouse=True, scope="session")
def reproducible_products():
    """GPU tests run the products and Krylov reductions in reproducible mode
    (fixed-order sums), so that results compare bit for bit between runs and
    between batched and single solves; tests of the atomic schedules switch
    it off themselves."""
    try:
        import h2lib_amd
        h2lib_amd.set_reproducible(True)
    except Exception:  # library not built: CPU tests need no GPU state
        pass
    yield


def golden(name):
    """Load a committed fixture (.npz, no pickles)."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def sphere4():
    return golden("mesh_sphere_s4.npz")


@pytest.fixture(scope="session")
def sphere8():
    return golden("mesh_sphere_s8.npz")


def gpu_available():
    try:
        import h2lib_amd
        return h2lib_amd.hip().h2hip_device_count() > 0
    except Exception:
        return False


def subsets(rng, nr, nc, rows, cols, same_dofs):
    """Random row and column index sets.  When rows and columns index the
    same DOFs, a few shared DOFs are placed at row positions below cols: the
    reference adds the constant-basis identity term at (t, t), which for
    t >= cols lies outside the window (undefined behaviour there)."""
    cidx = rng.permutation(nc)[:cols]
    if not same_dofs:
        return rng.permutation(nr)[:rows], cidx
    cset = set(int(v) for v in cidx)
    ridx = [int(v) for v in rng.permutation(nr) if int(v) not in cset]
    ridx = ridx[:rows]
    positions = list(range(0, min(rows, cols), 3))
    picks = rng.permutation(cols)[:len(positions)]
    for pos, j in zip(positions, picks):
        ridx[pos] = int(cidx[j])
    assert len(set(ridx)) == len(ridx)
    return np.array(ridx), cidx


def rel_block_err(a, b):
    """max |a - b| / max |b| over the finite entries; inf if the non-finite
    patterns differ."""
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb):
        return np.inf
    if not fb.any():
        return 0.0
    den = np.abs(b[fb]).max()
    return float(np.abs(a[fb] - b[fb]).max() / (den if den > 0 else 1.0))

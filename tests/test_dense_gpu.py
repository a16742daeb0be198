"""The dense workload: Laplace single layer, piecewise constant basis, unit
sphere of 2048 triangles (split 16), q = 2, q2 = 4, the full 2048 x 2048
matrix assembled on the device and multiplied by a seeded uniform[-1, 1]
vector (Tests/test_laplacebem3d.c:202-205, Library/amatrix.c:862-925).

Parity: the matrix against the oracle's (1e-12 of the largest entry), the
products y += A x and y += A^* x against numpy's products of the oracle's
matrix (1e-12 relative, max-norm), the product of a window inside a larger
leading dimension, and complex matrices (Helmholtz single layer, k = 2).
"""
import numpy as np
import pytest

from conftest import rel_block_err
from oracle_ffi import OracleBem

pytestmark = pytest.mark.gpu

TOL = 1e-12


def dense_device(h, mesh, kernel, k=0.0):
    """The full matrix through the device plan (one block, all DOFs)."""
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = h.Bem3d(gr, kernel, 0, 2, 4, 0, 0, k=k)
    T = mesh.triangles
    idx = np.arange(T, dtype=np.uint32)
    plan = h.NearfieldPlan(gr, 0, 0, idx, idx,
                           np.array([[0, T, 0, T]], dtype=np.uint32))
    out = h.DeviceArray(plan.entries, op.dtype)
    plan.execute(op, out.ptr)
    h.hip().h2hip_stream_sync(None)
    return gr, op, plan, out


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("kernel,k", [(0, 0.0), (1, 2.0)])
def test_dense_fill_and_product_match_oracle(kernel, k):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.amatrix import addeval_amatrix
    mesh = G.sphere(16)
    T = mesh.triangles
    print("dense 2048 kernel", kernel, flush=True)
    gr, op, plan, d_A = dense_device(h, mesh, kernel, k)
    A = d_A.download().reshape(T, T).T  # column-major
    want = OracleBem(kernel, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4,
                     k=k).nearfield()
    assert rel_block_err(A, want) < TOL
    rng = np.random.default_rng(16)
    x = rng.uniform(-1.0, 1.0, T)
    y0 = rng.uniform(-1.0, 1.0, T)
    alpha = 0.75
    if kernel:
        x = x + 1j * rng.uniform(-1.0, 1.0, T)
        y0 = y0 - 0.5j
        alpha = 0.75 - 0.25j
    for trans in (False, True):
        d_x = h.DeviceArray.from_host(x)
        d_y = h.DeviceArray.from_host(y0)
        addeval_amatrix(alpha, d_A, T, T, T, d_x, d_y, trans=trans)
        got = d_y.download()
        M = want.conj().T if trans else want
        ref = y0 + alpha * (M @ x)
        assert rel(got - y0, ref - y0) < TOL, trans


@pytest.mark.parametrize("rows,cols,ld", [(1, 1, 1), (37, 5, 40),
                                          (300, 1000, 311), (5000, 3, 5000),
                                          (64, 64, 64)])
def test_product_windows(rows, cols, ld):
    """Windows inside a larger leading dimension, one-column and one-row
    shapes, one and many column chunks; entries outside the window (NaN)
    must not be read."""
    import h2lib_amd as h
    from h2lib_amd.amatrix import addeval_amatrix
    rng = np.random.default_rng(rows * 7 + cols)
    full = np.full((ld, cols), np.nan)
    A = rng.standard_normal((rows, cols))
    full[:rows] = A
    d_A = h.DeviceArray.from_host(full.reshape(-1, order="F"))
    print("window", rows, cols, ld, flush=True)
    for trans in (False, True):
        n_in, n_out = (rows, cols) if trans else (cols, rows)
        x = rng.standard_normal(n_in)
        y = rng.standard_normal(n_out)
        d_x, d_y = h.DeviceArray.from_host(x), h.DeviceArray.from_host(y)
        addeval_amatrix(-1.5, d_A, rows, cols, ld, d_x, d_y, trans=trans)
        ref = y - 1.5 * ((A.T if trans else A) @ x)
        got = d_y.download()
        assert np.all(np.isfinite(got))
        assert rel(got, ref) < TOL


def test_product_is_deterministic():
    """Two products of the same data give the same bits (partial sums are
    added in chunk order)."""
    import h2lib_amd as h
    from h2lib_amd.amatrix import addeval_amatrix
    rng = np.random.default_rng(3)
    rows, cols = 2048, 2048
    A = rng.standard_normal(rows * cols)
    d_A = h.DeviceArray.from_host(A)
    x = h.DeviceArray.from_host(rng.standard_normal(cols))
    outs = []
    for _ in range(3):
        y = h.DeviceArray(rows)
        y.zero()
        addeval_amatrix(1.0, d_A, rows, cols, rows, x, y)
        outs.append(y.download())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])

"""The CPU restatement of the reference's GMRES (oracle/oracle_krylov.c),
checked without a GPU.

Pinned against (a) the reference's own solve_gmres_avector on the same H2
system (the reference's H2-matrix of the Laplace single layer potential,
its own product as the operator): same number of steps, solutions to
1e-9 relative; (b) textbook properties on random dense systems: the
residual bound |b - A x| <= eps |b| for several restart lengths and
agreement with a direct solve.  A wrong sign in a rotation or a missing
restart update fails (a) and (b).
"""
import numpy as np
import pytest

from oracle_ffi import ReferenceBem, ReferenceH2, have_reference, oracle_gmres


@pytest.mark.parametrize("kmax", [3, 10, 30])
def test_dense_residual_and_direct_solve(kmax):
    rng = np.random.default_rng(kmax)
    n = 60
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, it, res = oracle_gmres(lambda v: A @ v, n, b, np.zeros(n), 1e-12, 5000,
                              kmax)
    assert np.linalg.norm(b - A @ x) <= 2e-12 * np.linalg.norm(b)
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10
    assert 0 < it < 5000


def test_maxiter_stops():
    rng = np.random.default_rng(1)
    n = 40
    A = np.eye(n) + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, it, res = oracle_gmres(lambda v: A @ v, n, b, np.zeros(n), 1e-30, 7,
                              30)
    assert it == 6  # the loop stops when iter + 1 == maxiter


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("kmax", [5, 30])
def test_matches_reference_gmres(kmax):
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    ref = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    H = ReferenceH2(ref, 16, 2.0, 3, 0)
    n = H.n
    b = np.random.default_rng(5).standard_normal(n)
    want, it_ref = H.gmres(b, np.zeros(n), 1e-10, 2000, kmax)

    def mv(v):
        return H.addeval(v, np.zeros(n))

    got, it, res = oracle_gmres(mv, n, b, np.zeros(n), 1e-10, 2000, kmax)
    assert it == it_ref
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()

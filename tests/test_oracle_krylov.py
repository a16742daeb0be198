"""The CPU restatement of the reference's GMRES and CG
(oracle/oracle_krylov.c), checked without a GPU.

Pinned against (a) the reference's own solve_gmres_avector on the same H2
system (the reference's H2-matrix of the Laplace single layer potential,
its own product as the operator): same number of steps, solutions to
1e-9 relative; (b) textbook properties on random dense systems: the
residual bound |b - A x| <= eps |b| for several restart lengths and
agreement with a direct solve.  A wrong sign in a rotation or a missing
restart update fails (a) and (b).

CG is pinned the same way: (a) the reference's solve_cg_avector on its
H2-matrix of the Laplace single layer potential (symmetric positive
definite): same steps, solutions to 1e-9; and on the complex Helmholtz
single layer matrix the iterates after a fixed number of steps (the
recurrence with conjugated inner products, defined for any matrix) to
1e-10; (b) random SPD and Hermitian positive definite systems: residual
bound, direct solve, at most n steps on tiny well-separated spectra, and
the maxiter rule.  A dropped conjugate, a wrong sign of lambda or mu, or a
stale residual fails these.
"""
import numpy as np
import pytest

from oracle_ffi import (ReferenceBem, ReferenceH2, have_reference, oracle_cg,
                        oracle_gmres)


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


def _hpd(n, seed, cplx):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    if cplx:
        B = B + 1j * rng.standard_normal((n, n))
    return B @ B.conj().T / n + np.eye(n)


@pytest.mark.parametrize("cplx", [False, True])
def test_cg_dense_residual_and_direct_solve(cplx):
    n = 80
    A = _hpd(n, 11, cplx)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    x, it, res = oracle_cg(lambda v: A @ v, n, b, np.zeros(n), 1e-12, 1000,
                           cplx)
    r = np.linalg.norm(b - A @ x)
    assert r <= 2e-12 * np.linalg.norm(b)
    assert abs(res - r) <= 1e-13 * np.linalg.norm(b)
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10
    assert 0 < it < 1000


@pytest.mark.parametrize("cplx", [False, True])
def test_cg_terminates_within_n_steps(cplx):
    # eigenvalues 1..n well separated: exact CG ends after n steps
    n = 6
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) +
                        (1j * rng.standard_normal((n, n)) if cplx else 0))
    A = Q @ np.diag(np.arange(1.0, n + 1)) @ Q.conj().T
    b = rng.standard_normal(n) + 0j if cplx else rng.standard_normal(n)
    x, it, _ = oracle_cg(lambda v: A @ v, n, b, np.zeros(n), 1e-12, 100,
                         cplx)
    assert it <= n
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-11


def test_cg_maxiter_and_start_vector():
    n = 40
    A = _hpd(n, 2, False)
    b = np.random.default_rng(3).standard_normal(n)
    _, it, _ = oracle_cg(lambda v: A @ v, n, b, np.zeros(n), 1e-30, 5)
    assert it == 4  # the loop stops when iter + 1 == maxiter
    xs = np.linalg.solve(A, b)
    x, it, _ = oracle_cg(lambda v: A @ v, n, b, xs, 1e-10, 100)
    assert it == 0 and np.array_equal(x, xs)


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
def test_cg_matches_reference():
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    ref = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    H = ReferenceH2(ref, 16, 2.0, 3, 0)
    n = H.n
    b = np.random.default_rng(5).standard_normal(n)
    want, it_ref = H.cg(b, np.zeros(n), 1e-10, 2000)
    got, it, res = oracle_cg(lambda v: H.addeval(v, np.zeros(n)), n, b,
                             np.zeros(n), 1e-10, 2000)
    assert it == it_ref
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_cg_complex_iterates_match_reference():
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    ref = ReferenceBem(1, 0, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=2.0)
    H = ReferenceH2(ref, 16, 2.0, 3, 0)
    n = H.n
    rng = np.random.default_rng(6)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    want, it_ref = H.cg(b, x0, 1e-30, 11)
    got, it, _ = oracle_cg(lambda v: H.addeval(v, np.zeros(n, complex)), n,
                           b, x0, 1e-30, 11, cplx=True)
    assert it == it_ref == 10
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()

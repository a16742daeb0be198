"""Pins the oracle's quadrature rules (CPU only).

Checks against the reference's own tables (golden fixtures written from the
reference library built from its sources), and against what the mathematics
fixes independently: Gauss-Legendre identities, the measure of the pair of
reference triangles, exactness on polynomials, and a brute-force check of the
shared-vertex classification and permutation.
"""
import itertools

import numpy as np
import pytest

from conftest import golden
from oracle_ffi import Oracle

CLASSES = ["dist", "vert", "edge", "id", "single"]


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 8, 12])
def test_gauss_legendre_identities(m):
    x, w = Oracle.gauss(m)
    assert np.all(np.diff(x) > 0)
    # exact for polynomials of degree < 2m on [-1, 1]
    for d in range(2 * m):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(w, x ** d) - exact) < 1e-14
    # not exact for degree 2m (so a rule of lower order would be caught)
    d = 2 * m
    assert abs(np.dot(w, x ** d) - 2.0 / (d + 1)) > 1e-10


@pytest.mark.parametrize("q,q2", [(2, 4), (3, 5)])
@pytest.mark.parametrize("cls", range(5))
def test_rules_match_reference_tables(q, q2, cls):
    g = golden("singquad.npz")
    key = "%s_%d_%d_" % (CLASSES[cls], q, q2)
    x, y, w, base = Oracle.rule(q, q2, cls)
    assert base == 0.0
    np.testing.assert_allclose(x, g[key + "x"], rtol=0, atol=2e-15)
    np.testing.assert_allclose(y, g[key + "y"], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, g[key + "w"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("q,q2", [(2, 4), (3, 5), (1, 3)])
def test_rule_sizes_and_measure(q, q2):
    sizes = [q ** 4, 2 * q2 ** 4, 6 * q2 ** 4, 6 * q2 ** 4, q * q]
    for cls in range(5):
        x, y, w, _ = Oracle.rule(q, q2, cls)
        assert len(w) == sizes[cls]
        # area of the reference triangle is 1/2, of the pair 1/4
        assert abs(w.sum() - (0.5 if cls == 4 else 0.25)) < 1e-14


def _tri_moment(a, b):
    """Integral of x0^a x1^b over {0 <= x1 <= x0 <= 1}."""
    return 1.0 / ((b + 1) * (a + b + 2))


@pytest.mark.parametrize("cls", [0, 1, 2, 3])
def test_pair_rules_integrate_polynomials(cls):
    # the Duffy-transformed rules integrate smooth polynomials of low degree
    # on the triangle pair exactly
    q, q2 = 3, 5
    x, y, w, _ = Oracle.rule(q, q2, cls)
    n = len(w)
    x0, x1, y0, y1 = x[:n], x[n:], y[:n], y[n:]
    for a, b, c, d in itertools.product(range(3), repeat=4):
        if a + b + c + d > 3:
            continue
        exact = _tri_moment(a, b) * _tri_moment(c, d)
        got = np.dot(w, x0 ** a * x1 ** b * y0 ** c * y1 ** d)
        assert abs(got - exact) < 1e-13 * max(1.0, abs(exact)), (a, b, c, d)


def test_points_inside_reference_triangles():
    for cls in range(4):
        x, y, w, _ = Oracle.rule(3, 5, cls)
        n = len(w)
        for u, v in ((x[:n], x[n:]), (y[:n], y[n:])):
            assert np.all(v >= -1e-15) and np.all(u - v >= -1e-15)
            assert np.all(u <= 1 + 1e-15)
        assert np.all(w > 0)


def _select(tv, sv):
    """Shared-vertex count and permutation the reference's selection
    returns (Library/singquad2d.c:549-633), by brute force: among all
    permutation pairs with the shared vertices first in matching positions,
    the one that keeps the first triangle's order on the shared part and
    both triangles' order on the rest."""
    shared = [i for i in range(3) if tv[i] in sv]
    p = len(shared)
    if p in (0, 3):
        return p, [0, 1, 2], [0, 1, 2]
    tp = shared + [i for i in range(3) if i not in shared]
    sp = [sv.index(tv[i]) for i in shared]
    sp += [j for j in range(3) if j not in sp]
    return p, tp, sp


def test_classification_brute_force():
    import ctypes as C
    lib = Oracle.lib()
    lib.orc_select.restype = C.c_uint
    sq = lib.orc_sq_new(2, 3)
    rng = np.random.default_rng(3)
    seen = set()
    for _ in range(3000):
        tv = [int(v) for v in rng.choice(6, 3, replace=False)]
        sv = [int(v) for v in rng.choice(6, 3, replace=False)]
        a = (C.c_uint * 3)(*tv)
        b = (C.c_uint * 3)(*sv)
        tp = (C.c_uint * 3)()
        sp = (C.c_uint * 3)()
        rule = C.c_void_p()
        p = lib.orc_select(C.c_void_p(sq), a, b, tp, sp, C.byref(rule))
        ep, etp, esp = _select(tv, sv)
        assert p == ep
        if p in (1, 2):
            assert list(tp) == etp and list(sp) == esp
            for i in range(p):
                assert tv[tp[i]] == sv[sp[i]]
        seen.add(p)
    assert seen == {0, 1, 2, 3}
    lib.orc_sq_del(C.c_void_p(sq))

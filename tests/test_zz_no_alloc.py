"""h2hip_solve_cg_noalloc, the CG entry point that allocates nothing (needs
a GPU).

Its work space is a fixed device array of the library, shared by every call
on a GPU and keeping the last call's contents.  Checked through the C ABI
against h2hip_solve_cg (direct launches) on the same systems, two calls in a
row, the second smaller than the first, so that stale contents of the work
space from the larger solve would show: same step count, same final |r|,
the same solution bit for bit (the session runs the products and reductions
in reproducible mode, and both entry points launch the same kernels in the
same order).  Also a complex system, and the refusal of a system larger
than the work space (checked on the host before any device work).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

X0 = np.array([1.2, 1.2, 1.2])
MAXN2 = 1 << 20  # doubles per work vector of the fixed work space


def _noalloc(H, d_b, d_x, eps, maxiter):
    import h2lib_amd as h
    from h2lib_amd.solve import _addeval_fn, _sigs
    _sigs()
    f = h.hip().h2hip_solve_cg_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_void_p,
                  C.c_void_p, C.c_double, C.c_uint32, C.c_void_p,
                  C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    it, res = C.c_uint32(), C.c_double()
    rc = f(_addeval_fn(H), H._h, int(H.dtype == np.complex128), d_b.n,
           C.c_void_p(d_b.ptr), C.c_void_p(d_x.ptr), eps, maxiter, None,
           C.byref(it), C.byref(res))
    return rc, it.value, res.value


def _laplace(split, clf):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceDirichlet
    prob = LaplaceDirichlet(G.sphere(split), clf=clf, eta=2.0, m=3)
    _, d_b = prob.rhs(X0)
    return prob, d_b


def _both(H, d_b, n, dtype, eps, maxiter):
    import h2lib_amd as h
    from h2lib_amd.solve import solve_cg
    d_x = h.DeviceArray(n, dtype)
    d_x.zero()
    it_ref, res_ref = solve_cg(H, d_b, d_x, eps, maxiter)
    x_ref = d_x.download()
    d_y = h.DeviceArray(n, dtype)
    d_y.zero()
    rc, it, res = _noalloc(H, d_b, d_y, eps, maxiter)
    assert rc == 0, h.hip().h2hip_last_error()
    return (it, res, d_y.download()), (it_ref, res_ref, x_ref)


@pytest.mark.parametrize("maxiter", [2000, 13])
def test_noalloc_cg_equals_cg_two_calls_second_smaller(maxiter):
    print("noalloc CG: sphere 16 then sphere 8, maxiter %d" % maxiter,
          flush=True)
    for split, clf in ((16, 32), (8, 16)):
        prob, d_b = _laplace(split, clf)
        (it, res, x), (it_ref, res_ref, x_ref) = _both(
            prob.V, d_b, prob.n, np.float64, 1e-10, maxiter)
        print("  split %d: n %d, steps %d / %d, |r| %.3e / %.3e" %
              (split, prob.n, it, it_ref, res, res_ref), flush=True)
        assert it == it_ref
        assert res == res_ref
        assert np.array_equal(x, x_ref)
        if maxiter == 13:
            assert it == 12


def test_noalloc_cg_complex_after_real():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d, new_slp_helmholtz_bem3d
    from h2lib_amd.h2matrix import H2Matrix
    print("noalloc CG: complex Helmholtz SLP, sphere 8, 10 steps", flush=True)
    mesh = G.sphere(8)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    op = new_slp_helmholtz_bem3d(2.0, gr, 2, 4, 0, 0)
    rt = G.ClusterTree(mesh, 0, 16)
    bt = G.BlockTree(rt, rt, 2.0, 0)
    H = H2Matrix(op, mesh, rt, rt, bt, 3)
    n = mesh.triangles
    rng = np.random.default_rng(11)
    b = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    d_b = h.DeviceArray(n, np.complex128)
    d_b.upload(b)
    (it, res, x), (it_ref, res_ref, x_ref) = _both(H, d_b, n, np.complex128,
                                                   1e-14, 11)
    print("  steps %d / %d" % (it, it_ref), flush=True)
    assert it == it_ref == 10
    assert res == res_ref
    assert np.array_equal(x, x_ref)


def test_noalloc_cg_refuses_systems_beyond_the_work_space():
    import h2lib_amd as h
    print("noalloc CG: size limit", flush=True)
    prob, d_b = _laplace(4, 16)
    d_x = h.DeviceArray(prob.n, np.float64)
    d_x.zero()

    class Big:  # the same matrix, a length beyond the limit (never used:
        pass    # the size is refused before any device work)
    big = Big()
    big.n = MAXN2 + 1
    big.ptr = d_b.ptr
    rc, _, _ = _noalloc(prob.V, big, d_x, 1e-10, 100)
    assert rc != 0
    assert b"work space" in h.hip().h2hip_last_error()

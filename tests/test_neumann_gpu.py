"""The exterior Neumann problem of the Laplace equation on the GPU (needs a
GPU): the adjoint double layer H2-matrix with alpha = -1/2, right-hand side
integrals of the Neumann data formed on the device, GMRES(30).

Against the reference: the right-hand side agrees with the host
restatement (itself pinned to integrate_bem3d_l_avector) to 1e-14; the
solution agrees with the reference's solve_gmres_avector on the
reference's own adjoint double layer H2-matrix (new_adlp_laplace_bem3d,
interpolation, same trees) to 1e-8 relative at residual 1e-12 |b|.
Against mathematics: the Dirichlet trace V sigma of the solution
approaches the field of the sources under refinement.
"""
import numpy as np
import pytest

from oracle_ffi import ReferenceBem, ReferenceH2, have_reference

pytestmark = pytest.mark.gpu

SOURCES = [np.array([0.3, 0.2, 0.1]), np.array([-0.2, 0.1, -0.3])]


@pytest.mark.parametrize("kind", [0, 1])
def test_device_integrals_equal_host(kind):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import (laplace_point_integrals,
                                 laplace_point_integrals_device)
    mesh = G.sphere(12)
    gr = h.Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    d = h.DeviceArray(mesh.vertices)
    want = np.zeros(mesh.vertices)
    for j, x0 in enumerate([np.array([1.2, 1.2, 1.2])] + SOURCES):
        laplace_point_integrals_device(gr, 2, kind, x0, d, accumulate=j > 0)
        want += laplace_point_integrals(mesh, kind, x0)
        got = d.download()
        assert np.abs(got - want).max() < 1e-14 * np.abs(want).max(), j


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("split,clf,m", [(8, 16, 3), (16, 32, 3)])
def test_neumann_solution_matches_reference(split, clf, m):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceNeumannExterior
    mesh = G.sphere(split)
    prob = LaplaceNeumannExterior(mesh, clf=clf, eta=2.0, m=m)
    d_b, _ = prob.rhs_sources(SOURCES)
    d_x, it, res = prob.solve(d_b, eps=1e-12, maxiter=2000, kmax=30)
    x = d_x.download()
    b = d_b.download()
    ref = ReferenceBem(0, 2, 1, 1, mesh.x, mesh.t, q=2, q2=4, alpha=-0.5)
    rh = ReferenceH2(ref, clf, 2.0, m, 0)
    assert np.array_equal(rh.idx(), prob.tree.idx)
    x_ref, it_ref = rh.gmres(b, np.zeros_like(b), 1e-12, 2000, 30)
    print("neumann split", split, "steps", it, it_ref, flush=True)
    assert np.abs(x - x_ref).max() < 1e-8 * np.abs(x_ref).max()
    assert abs(it - it_ref) <= max(2, it_ref // 10)


def test_dirichlet_trace_converges():
    # the density's single layer potential reproduces the field of the
    # sources: its Galerkin moments approach those of u under refinement
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import LaplaceNeumannExterior
    errs = []
    for split in (4, 8, 16):
        mesh = G.sphere(split)
        prob = LaplaceNeumannExterior(mesh, clf=16, eta=2.0, m=4)
        d_b, _ = prob.rhs_sources(SOURCES)
        d_x, it, res = prob.solve(d_b, eps=1e-12, maxiter=2000, kmax=30)
        errs.append(prob.dirichlet_error(d_x, SOURCES))
        print("neumann trace split", split, "steps", it, "error %.3e" %
              errs[-1], flush=True)
    assert errs[2] < errs[1] < errs[0]
    assert errs[2] < 2e-2

"""Host restatement of the Laplace point-source integrals against the linear
basis (h2lib_amd/host/functionals.c, h2b_integrate_l), the right-hand side
of the exterior Neumann problem, against the reference's
integrate_bem3d_l_avector with eval_dirichlet_fundamental_laplacebem3d /
eval_neumann_fundamental_laplacebem3d (source fixed at (1.2, 1.2, 1.2) in
the reference) (no GPU).

Also pinned by mathematics: the integrals of the basis functions sum to
the integral of the data (the hat functions are a partition of unity), and
for a source inside a closed surface the flux of grad 1/|x - x0| through it
is -4 pi (Gauss), so the Neumann integrals sum to -4 pi up to the
quadrature error; a dropped sign or a normal of the wrong side fails it."""
import numpy as np
import pytest

from oracle_ffi import ReferenceBem, have_reference


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("kind", [0, 1])
def test_point_integrals_match_reference(kind):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import laplace_point_integrals
    mesh = G.sphere(8)
    ref = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    want = ref.laplace_point_integrate_l(kind)
    got = laplace_point_integrals(mesh, kind, (1.2, 1.2, 1.2))
    assert np.abs(got - want).max() < 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("x0", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.1)])
def test_neumann_integrals_carry_the_flux(x0):
    from h2lib_amd import geometry as G
    from h2lib_amd.solve import laplace_point_integrals
    errs = []
    for s in (8, 16):
        mesh = G.sphere(s)
        g = laplace_point_integrals(mesh, 1, x0, q=4)
        errs.append(abs(g.sum() + 4 * np.pi))
    assert errs[1] < errs[0] < 0.2
    # an exterior source carries no flux through the surface
    mesh = G.sphere(16)
    g = laplace_point_integrals(mesh, 1, (1.5, 0.3, 0.0), q=4)
    assert abs(g.sum()) < 0.05

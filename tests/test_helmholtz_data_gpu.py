"""Helmholtz point-source data and error norms on the device (needs a GPU).

h2hip_helmholtz_point_data and h2hip_helmholtz_point_errors_c against the
host restatements (h2b_projectL2_c_helmholtz, h2b_integrate_c_helmholtz,
h2b_integrate_l_helmholtz, h2b_normL2_c_helmholtz; these are checked
against the reference library on the CPU) and, where the reference is
built, against the reference's projectL2_bem3d_c_avector,
integrate_bem3d_c_avector / integrate_bem3d_l_avector, normL2diff_c_bem3d
and normL2_bem3d directly: to 1e-13 relative (max-norm), the rounding of
the device's FP64 sin/cos.  Accumulation over several sources equals the
sum of the single-source results.
"""
import numpy as np
import pytest

from oracle_ffi import ReferenceBem, have_reference

pytestmark = pytest.mark.gpu

SRC = [np.array([0.3, 0.2, 0.1]), np.array([-0.1, 0.25, -0.3])]


def _close(a, b, tol=1e-13):
    return np.abs(a - b).max() <= tol * np.abs(b).max()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("k", [2.0, 6.5])
def test_device_data_matches_host(kind, k):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d
    from h2lib_amd.solve import (INTEGRATE_C, INTEGRATE_L, PROJECT_C,
                                 helmholtz_point_data,
                                 helmholtz_point_data_device,
                                 helmholtz_point_integrals)
    mesh = G.cube(6)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    x0 = SRC[0]
    cases = ((PROJECT_C, mesh.triangles,
              lambda x: helmholtz_point_data(mesh, kind, k, x, 2)),
             (INTEGRATE_C, mesh.triangles,
              lambda x: helmholtz_point_integrals(mesh, kind, k, x, 2, 0)),
             (INTEGRATE_L, mesh.vertices,
              lambda x: helmholtz_point_integrals(mesh, kind, k, x, 2, 1)))
    for mode, n, host in cases:
        d = h.DeviceArray(n, np.complex128)
        helmholtz_point_data_device(gr, 2, kind, mode, k, x0, d)
        want = host(x0)
        assert _close(d.download(), want), mode
        # two sources accumulated
        helmholtz_point_data_device(gr, 2, kind, mode, k, SRC[1], d,
                                    accumulate=True)
        assert _close(d.download(), want + host(SRC[1])), mode


@pytest.mark.parametrize("kind", [0, 1])
def test_device_errors_match_host(kind):
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d
    from h2lib_amd.solve import (helmholtz_point_errors,
                                 helmholtz_point_errors_device)
    mesh = G.sphere(12)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    rng = np.random.default_rng(kind)
    x = rng.standard_normal(mesh.triangles) + 1j * rng.standard_normal(
        mesh.triangles)
    got = helmholtz_point_errors_device(gr, 3, kind, 3.0, SRC[0],
                                        h.DeviceArray.from_host(x))
    want = helmholtz_point_errors(mesh, x, kind, 3.0, SRC[0], 3)
    assert np.abs(got - want).max() <= 1e-13 * want.max()


@pytest.mark.skipif(not have_reference(True), reason="reference not built")
def test_device_data_matches_reference():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.bem3d import Surface3d
    from h2lib_amd.solve import (INTEGRATE_C, INTEGRATE_L, PROJECT_C,
                                 helmholtz_point_data_device,
                                 helmholtz_point_errors_device)
    k = 2.0
    mesh = G.sphere(8)
    gr = Surface3d(mesh.x, mesh.t, mesh.n, mesh.g)
    rc = ReferenceBem(1, 1, 0, 0, mesh.x, mesh.t, q=2, q2=4, k=k, alpha=0.5)
    rl = ReferenceBem(1, 1, 1, 1, mesh.x, mesh.t, q=2, q2=4, k=k, alpha=0.5)
    for kind in (0, 1):
        d = h.DeviceArray(mesh.triangles, np.complex128)
        helmholtz_point_data_device(gr, 2, kind, PROJECT_C, k, SRC[0], d)
        assert _close(d.download(),
                      rc.helmholtz_point_project(kind, k, SRC[0]))
        e = helmholtz_point_errors_device(gr, 2, kind, k, SRC[0], d)
        want = rc.helmholtz_point_errors(kind, k, SRC[0], d.download())
        assert np.abs(e - want).max() <= 1e-13 * want.max()
    d = h.DeviceArray(mesh.triangles, np.complex128)
    helmholtz_point_data_device(gr, 2, 0, INTEGRATE_C, k, SRC[0], d)
    assert _close(d.download(), rc.helmholtz_point_integrate(k, SRC[0]))
    d = h.DeviceArray(mesh.vertices, np.complex128)
    helmholtz_point_data_device(gr, 2, 0, INTEGRATE_L, k, SRC[0], d)
    assert _close(d.download(), rl.helmholtz_point_integrate(k, SRC[0]))

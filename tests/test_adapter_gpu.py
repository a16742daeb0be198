"""The H2Lib-side binding (h2lib_amd/adapter/h2lib_hip.c) inside the reference
library itself (needs a GPU).

The reference is built from its own sources together with the binding
(oracle/_ref/libh2ref_hip_*.so).  setup_hip_bem3d overwrites bem->nearfield
and bem->nearfield_far of the reference's own bem3d objects, so the
reference's own callers run on the GPU: the dense fill through the callback
and assemble_bem3d_h2matrix, whose OpenMP tasks call the callback
concurrently.  new_hip_h2matrix packs the reference's assembled h2matrix for
the GPU product, which also serves the reference's solve_gmres_avector as
its addeval_t callback.  Everything is compared with the unmodified
reference build: entries and products to 1e-12 relative, GMRES solutions to
1e-6 relative (rounding differs between two GMRES runs).
"""
import numpy as np
import pytest

from conftest import golden, rel_block_err, subsets
from oracle_ffi import ReferenceBem, ReferenceH2, have_reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (have_reference(False, True) and
                                      have_reference(True, True)),
                                 reason="reference with binding not built")]

TOL = 1e-12

CASES = [  # kernel, layer, row basis, col basis, alpha, k
    (0, 0, 0, 0, 0.0, 0.0),
    (0, 1, 0, 0, 0.5, 0.0),
    (0, 0, 1, 1, 0.0, 0.0),
    (0, 1, 1, 1, -0.5, 0.0),
    (0, 0, 0, 1, 0.0, 0.0),
    (0, 1, 1, 0, 0.5, 0.0),
    (1, 0, 0, 0, 0.0, 2.0),
    (1, 1, 1, 1, 0.5, 3.0),
]


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("case", CASES)
def test_callback_in_reference_bem3d(case):
    kern, layer, br, bc, alpha, k = case
    g = golden("mesh_sphere_s4.npz")
    x, t = g["x"], g["t"]
    plain = ReferenceBem(kern, layer, br, bc, x, t, q=2, q2=4, alpha=alpha,
                         k=k)
    hip = ReferenceBem(kern, layer, br, bc, x, t, hip=True, q=2, q2=4,
                       alpha=alpha, k=k)
    hip.use_hip()
    print("binding case", case, flush=True)
    for far in (False, True):
        # the regular rule on identical triangles (far field of the whole
        # matrix) divides by zero: inf in the reference, NaN here; both
        # non-finite at the same entries
        err = rel_block_err(hip.nearfield(far=far), plain.nearfield(far=far))
        assert err < TOL, far
    rng = np.random.default_rng(1)
    nr, nc = plain.ndof(br), plain.ndof(bc)
    ridx, cidx = subsets(rng, nr, nc, nr // 2, nc // 3, br == bc)
    for ntrans in (False, True):
        want = plain.nearfield(ridx, cidx, ntrans=ntrans)
        got = hip.nearfield(ridx, cidx, ntrans=ntrans)
        assert rel_block_err(got, want) < TOL, ntrans


@pytest.mark.parametrize("kern,layer,basis,k", [(0, 0, 1, 0.0),
                                                 (0, 1, 1, 0.0),
                                                 (1, 0, 0, 2.0)])
def test_reference_h2_assembly_and_product(kern, layer, basis, k):
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    alpha = 0.5 if layer else 0.0
    plain = ReferenceBem(kern, layer, basis, basis, mesh.x, mesh.t, q=2,
                         q2=4, alpha=alpha, k=k)
    hip = ReferenceBem(kern, layer, basis, basis, mesh.x, mesh.t, hip=True,
                       q=2, q2=4, alpha=alpha, k=k)
    hip.use_hip()
    print("binding h2", kern, layer, basis, flush=True)
    H0 = ReferenceH2(plain, 16, 2.0, 3, 0)
    # assemble_bem3d_h2matrix of the reference, near field from the GPU
    H1 = ReferenceH2(hip, 16, 2.0, 3, 0)
    nb = H0.near_blocks()
    assert np.array_equal(nb, H1.near_blocks())
    for b in range(0, len(nb), max(1, len(nb) // 17)):
        r, c = int(nb[b, 1]), int(nb[b, 3])
        assert rel(H1.near_entries(b, r, c), H0.near_entries(b, r, c)) < TOL
    rng = np.random.default_rng(2)
    n = H0.n
    dt = plain.dtype
    xv = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, n).astype(dt)
    if kern == 1:
        xv = xv + 1j * rng.uniform(-1, 1, n)
    a = 0.75 - 0.25j if kern == 1 else -0.75
    for trans in (False, True):
        want = H0.addeval(xv, y0, alpha=a, trans=trans)
        # the reference's own product over the GPU-assembled near field
        assert rel(H1.addeval(xv, y0, alpha=a, trans=trans), want) < TOL
        # the packed GPU product of the reference's h2matrix
        assert rel(H1.hip_addeval(xv, y0, alpha=a, trans=trans), want) < TOL


def test_gmres_through_binding():
    from h2lib_amd import geometry as G
    mesh = G.sphere(8)
    plain = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, q=2, q2=4)
    hip = ReferenceBem(0, 0, 1, 1, mesh.x, mesh.t, hip=True, q=2, q2=4)
    H0 = ReferenceH2(plain, 16, 2.0, 3, 0)
    H1 = ReferenceH2(hip, 16, 2.0, 3, 0)
    b = np.random.default_rng(5).standard_normal(H0.n)
    x0 = np.zeros(H0.n)
    want, it0 = H0.gmres(b, x0, 1e-10, 1000, 30)
    for device in (False, True):
        got, it = H1.hip_gmres(b, x0, 1e-10, 1000, 30, device)
        print("binding gmres device", device, "steps", it, it0, flush=True)
        assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()
        assert abs(int(it) - int(it0)) <= max(2, it0 // 10)


@pytest.mark.parametrize("case", [(0, 1, 1, 1, -0.5, 0.0),
                                  (1, 0, 0, 0, 0.0, 4.0)])
def test_mixed_precision_through_binding(case):
    # set_precision_hip_bem3d on the reference's own bem3d object: its
    # dense fill now gets mixed-precision entries (within 2e-6 of the
    # unmodified reference), and switching back restores 1e-12
    kern, layer, br, bc, alpha, k = case
    g = golden("mesh_sphere_s4.npz")
    x, t = g["x"], g["t"]
    plain = ReferenceBem(kern, layer, br, bc, x, t, q=2, q2=4, alpha=alpha,
                         k=k)
    hip = ReferenceBem(kern, layer, br, bc, x, t, hip=True, q=2, q2=4,
                       alpha=alpha, k=k)
    hip.use_hip()
    want = plain.nearfield()
    hip.hip_precision(1)
    err = rel_block_err(hip.nearfield(), want)
    print("binding mixed precision", case, "err %.2e" % err, flush=True)
    assert 0.0 < err < 2e-6
    hip.hip_precision(0)
    assert rel_block_err(hip.nearfield(), want) < TOL


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[5], CASES[7]])
def test_deferred_callbacks_and_flush(case):
    """setup_hip_deferred_bem3d: the reference's callbacks only record the
    blocks, flush_hip_bem3d writes them (the driver flushes after its
    loops): dense fills, transposed subsets and the far rule give the
    immediate callbacks' entries."""
    kern, layer, br, bc, alpha, k = case
    g = golden("mesh_sphere_s4.npz")
    x, t = g["x"], g["t"]
    plain = ReferenceBem(kern, layer, br, bc, x, t, q=2, q2=4, alpha=alpha,
                         k=k)
    hip = ReferenceBem(kern, layer, br, bc, x, t, hip=True, q=2, q2=4,
                       alpha=alpha, k=k)
    hip.use_hip(deferred=True)
    print("deferred binding case", case, flush=True)
    for far in (False, True):
        err = rel_block_err(hip.nearfield(far=far), plain.nearfield(far=far))
        assert err < TOL, far
    rng = np.random.default_rng(3)
    nr, nc = plain.ndof(br), plain.ndof(bc)
    ridx, cidx = subsets(rng, nr, nc, nr // 2, nc // 3, br == bc)
    for ntrans in (False, True):
        want = plain.nearfield(ridx, cidx, ntrans=ntrans)
        got = hip.nearfield(ridx, cidx, ntrans=ntrans)
        assert rel_block_err(got, want) < TOL, ntrans


@pytest.mark.parametrize("kern,layer,basis,k", [(0, 1, 1, 0.0),
                                                 (1, 0, 0, 2.0)])
def test_deferred_reference_h2_assembly(kern, layer, basis, k):
    """The reference's own assemble_bem3d_h2matrix with deferred callbacks
    (its OpenMP tasks record their blocks concurrently, one flush computes
    them) gives the near-field blocks and products of the CPU assembly."""
    from h2lib_amd import geometry as G
    mesh = G.sphere(12)
    alpha = 0.5 if layer else 0.0
    plain = ReferenceBem(kern, layer, basis, basis, mesh.x, mesh.t, q=2,
                         q2=4, alpha=alpha, k=k)
    hip = ReferenceBem(kern, layer, basis, basis, mesh.x, mesh.t, hip=True,
                       q=2, q2=4, alpha=alpha, k=k)
    hip.use_hip(deferred=True)
    H0 = ReferenceH2(plain, 16, 2.0, 3, 0)
    H1 = ReferenceH2(hip, 16, 2.0, 3, 0)
    nb = H0.near_blocks()
    print("deferred h2 assembly", kern, layer, basis, len(nb), "blocks",
          flush=True)
    for b in range(len(nb)):
        r, c = int(nb[b, 1]), int(nb[b, 3])
        assert rel(H1.near_entries(b, r, c), H0.near_entries(b, r, c)) < TOL
    rng = np.random.default_rng(4)
    xv = rng.uniform(-1, 1, H0.n).astype(plain.dtype)
    y0 = np.zeros(H0.n, dtype=plain.dtype)
    for trans in (False, True):
        want = H0.addeval(xv, y0, trans=trans)
        assert rel(H1.addeval(xv, y0, trans=trans), want) < TOL

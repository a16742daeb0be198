"""Host-side integer and geometry work against the reference (no GPU).

Meshes (macrosurface refinement of the sphere, the cube and the cylinder),
the adaptive cluster trees and the near-field block lists decide the DOF
order, the blocks and hence every entry: they must be identical to the
reference's (bit-exact coordinates and normals, equal index arrays),
checked against the committed fixtures and, where the reference is built,
live at further sizes.
"""
import numpy as np
import pytest

from conftest import golden
from oracle_ffi import Reference, ReferenceBem, ReferenceH2, have_reference


def mesh_arrays(m):
    return m.x, m.t, m.n, m.g


@pytest.mark.parametrize("name,kind,split", [("sphere", 0, 4),
                                             ("sphere", 0, 8),
                                             ("sphere", 0, 16),
                                             ("cube", 1, 4),
                                             ("cylinder", 2, 2)])
def test_meshes_match_fixtures(name, kind, split):
    from h2lib_amd import geometry as G
    g = golden("mesh_%s_s%d.npz" % (name, split))
    m = getattr(G, name)(split)
    x, t, n, gg = mesh_arrays(m)
    assert np.array_equal(t, g["t"])
    assert np.array_equal(x, g["x"])
    assert np.array_equal(n, g["n"])
    assert np.array_equal(gg, g["g"])


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("name,kind,split", [("sphere", 0, 5),
                                             ("cube", 1, 7),
                                             ("cylinder", 2, 3),
                                             ("cylinder", 2, 4),
                                             ("cylinder", 2, 6)])
def test_meshes_match_reference(name, kind, split):
    from h2lib_amd import geometry as G
    x, t, n, g = Reference.mesh(kind, split)
    m = getattr(G, name)(split)
    assert np.array_equal(m.t, t)
    assert np.array_equal(m.x, x)
    assert np.array_equal(m.n, n)
    assert np.array_equal(m.g, g)


@pytest.mark.skipif(not have_reference(False), reason="reference not built")
@pytest.mark.parametrize("name,split,basis,clf", [("sphere", 8, 1, 16),
                                                  ("sphere", 16, 0, 32),
                                                  ("cube", 6, 1, 20),
                                                  ("cylinder", 3, 0, 24)])
def test_clusters_and_blocks_match_reference(name, split, basis, clf):
    from h2lib_amd import geometry as G
    m = getattr(G, name)(split)
    tree = G.ClusterTree(m, basis, clf)
    blocks = G.BlockTree(tree, tree, 2.0, 0)
    ref = ReferenceBem(0, 0, basis, basis, m.x, m.t, q=2, q2=4)
    rh = ReferenceH2(ref, clf, 2.0, 2, 0)
    assert np.array_equal(tree.idx, rh.idx())
    assert np.array_equal(blocks.near_block_list(), rh.near_blocks())

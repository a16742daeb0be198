"""Host side of the many-mesh batch: the drawn meshes and the autotuner's
bookkeeping (no GPU)."""
import numpy as np
import pytest


def test_random_rotations_are_rotations():
    from h2lib_amd.batch import random_rotation
    rng = np.random.default_rng(0)
    Rs = [random_rotation(rng) for _ in range(200)]
    for R in Rs:
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
    # uniform on SO(3): E[R] = 0 and E[trace R] = 0
    assert np.abs(np.mean(Rs, axis=0)).max() < 0.15
    assert abs(np.mean([np.trace(R) for R in Rs])) < 0.2


def test_random_meshes_are_reproducible_and_in_range():
    from h2lib_amd import geometry as G
    from h2lib_amd.batch import MACRO_TRIANGLES, random_meshes
    a = random_meshes(12, seed=7, tmin=128, tmax=2048)
    b = random_meshes(12, seed=7, tmin=128, tmax=2048)
    kinds = set()
    for ma, mb in zip(a, b):
        assert (ma.kind, ma.split, ma.scale) == (mb.kind, mb.split, mb.scale)
        assert np.array_equal(ma.mesh.x, mb.mesh.x)
        kinds.add(ma.kind)
        T = ma.triangles
        assert T == MACRO_TRIANGLES[ma.kind] * ma.split ** 2
        # rounding the split keeps T within a factor of the range
        assert 128 / 4 <= T <= 2048 * 2
        assert 0.5 <= ma.scale <= 2.0
        # x -> s R x: the areas scale by s^2, the normals rotate with R
        ref = getattr(G, ma.kind)(ma.split)
        assert np.allclose(ma.mesh.g, ma.scale ** 2 * ref.g, rtol=1e-12)
        assert np.allclose(ma.mesh.n, ref.n @ ma.R.T, atol=1e-12)
        assert np.allclose(ma.mesh.x, ma.scale * ref.x @ ma.R.T,
                           atol=1e-14)
    assert len(kinds) >= 2


def test_autotuner_bookkeeping(tmp_path):
    from h2lib_amd.autotune import Autotuner
    t = Autotuner(stream=None, candidates=(32, 64))
    assert t.size_class(512) == 9 and t.size_class(700) == 9
    assert t.size_class(32768) == 15
    # a leaf-size choice recorded for a class is reused and survives a
    # save / load round trip
    t.leaf[9] = {"clf": 64, "triangles": 512,
                 "costs": {32: {"cost_s": 2.0}, 64: {"cost_s": 1.0}}}
    t.schedules["x"] = {"best": [1, 0], "ms": {"00": 2.0, "10": 1.0}}

    class FakeMesh:
        triangles = 600
    assert t.leaf_size(FakeMesh()) == 64
    p = tmp_path / "tune.json"
    t.save(str(p))
    u = Autotuner(stream=None)
    u.load(str(p))
    assert u.leaf[9]["clf"] == 64 and u.schedules["x"]["best"] == [1, 0]

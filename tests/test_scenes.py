"""Generators: determinism, shapes and the batched generator's per-mesh identity."""
import numpy as np

from gprt_amd import scenes


def test_blob_meshes_match_single_mesh_generator():
    sizes = [64, 100, 4097, 7, 333]
    verts, idx, offs = scenes.blob_meshes(sizes, seed0=1000)
    for i, n in enumerate(sizes):
        v1, i1 = scenes.blob_scene(n, seed=1000 + i)
        np.testing.assert_array_equal(verts[3 * offs[i]:3 * offs[i + 1]], v1)
        np.testing.assert_array_equal(idx[offs[i]:offs[i + 1]], i1)


def test_batch_sizes_log_uniform():
    s = scenes.batch_sizes(4096, seed=3)
    assert s.min() >= 64 and s.max() <= 262144
    # log-uniform: about half of the meshes below the geometric mean 4096
    assert 0.4 < np.mean(s < 4096) < 0.6


def test_generators_deterministic():
    a = scenes.uniform_scene(100, seed=5)[0]
    b = scenes.uniform_scene(100, seed=5)[0]
    np.testing.assert_array_equal(a, b)
    r = scenes.pinhole_rays(16, 8)
    assert len(r) == 128 and np.allclose(np.sqrt(r["dx"] ** 2 + r["dy"] ** 2 + r["dz"] ** 2), 1, atol=1e-6)

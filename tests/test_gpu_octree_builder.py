"""Device octree builder (create_sdf_SBS with the reference's adaptive frame
octree, utils/sparse_octree_builder.cpp:125-444) against the oracle's
restatement, and its scenes through the renderer.

- Distance functions (sphere, torus): the same bricks -- positions and mixed
  LODs -- with the same quantised words, bit for bit, as the oracle, for
  several depths, brick sizes, pads and value widths (brick order aside: the
  reference's order depends on its threads, the device sorts by LOD and
  Morton order).
- Meshes: the adaptive scene of a displaced icosphere has mixed LODs, covers
  the mesh's vertices with bricks of the finest LOD, and renders at parity
  with the oracle (the usual G-buffer bar), so mixed-LOD scenes quantised over
  the brick diagonal go through the whole path.
"""
import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import compare_gbuffer, preset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    yield r
    r.close()


def _bricks(nodes, values, n_words):
    return {(int(n[0]), int(n[1])): values[int(n[2]):int(n[2]) + n_words].tobytes() for n in nodes}


CASES = [  # shape, params, depth, brick_size, brick_pad, bytes_per_value
    (L.PRIMITIVE_SPHERE, (0.0, 0.0, 0.0, 0.7), 6, 4, 0, 1),
    (L.PRIMITIVE_SPHERE, (0.1, -0.05, 0.02, 0.55), 7, 2, 1, 2),
    (L.PRIMITIVE_SPHERE, (0.0, 0.0, 0.0, 0.8), 4, 3, 1, 4),
    (L.PRIMITIVE_TORUS, (0.6, 0.2), 6, 4, 1, 1),
    (L.PRIMITIVE_TORUS, (0.5, 0.25), 8, 2, 0, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[f"shape{c[0]}-d{c[2]}-bs{c[3]}-pad{c[4]}-bpv{c[5]}" for c in CASES])
def test_octree_builder_matches_oracle(gpu, case):
    shape, params, depth, bs, pad, bpv = case
    dev, ms = L.sbs_octree_from_primitive(gpu, shape, params, depth, bs, pad, bpv)
    o_nodes, o_values = oracle.sbs_octree(shape, params, depth, bs, pad, bpv)
    v = bs + 2 * pad + 1
    n_words = (v ** 3 + 4 // bpv - 1) // (4 // bpv)
    lods = np.unique(dev.nodes[:, 1] & 0xFFFF)
    print(case, dev.nodes.shape[0], "bricks, LODs", lods.tolist(), round(ms, 1), "ms")
    assert (dev.header.brick_size, dev.header.brick_pad, dev.header.bytes_per_value, dev.header.aux_data) == \
           (bs, pad, bpv, L.SDF_SBS_NODE_LAYOUT_DX)
    assert dev.nodes.shape[0] == o_nodes.shape[0] > 50
    assert np.array_equal(dev.nodes[:, 2], np.arange(dev.nodes.shape[0], dtype=np.uint32) * n_words)
    assert dev.values.size == o_values.size
    got, want = _bricks(dev.nodes, dev.values, n_words), _bricks(o_nodes, o_values, n_words)
    assert got.keys() == want.keys()
    assert all(got[k] == want[k] for k in want)
    if depth > 4:
        assert len(lods) > 1


LIMIT_CASES = [  # shape, params, depth, brick_size, brick_pad, nodes_limit
    (L.PRIMITIVE_SPHERE, (0.0, 0.0, 0.0, 0.7), 6, 4, 0, 12000),
    (L.PRIMITIVE_SPHERE, (0.0, 0.0, 0.0, 0.7), 6, 4, 1, 3000),
    (L.PRIMITIVE_TORUS, (0.6, 0.2), 7, 2, 0, 20000),
]


@pytest.mark.parametrize("case", LIMIT_CASES, ids=[f"shape{c[0]}-d{c[2]}-limit{c[5]}" for c in LIMIT_CASES])
def test_octree_node_limit_matches_oracle(gpu, case):
    """frame_octree_limit_nodes (:2229-2330): the device builder's frame,
    merged on the host, gives the oracle's bricks bit for bit."""
    shape, params, depth, bs, pad, lim = case
    dev, ms = L.sbs_octree_from_primitive(gpu, shape, params, depth, bs, pad, 1, nodes_limit=lim)
    full, _ = L.sbs_octree_from_primitive(gpu, shape, params, depth, bs, pad, 1)
    o_nodes, o_values = oracle.sbs_octree(shape, params, depth, bs, pad, 1, lim)
    v = bs + 2 * pad + 1
    n_words = (v ** 3 + 3) // 4
    print(case, dev.nodes.shape[0], "bricks of", full.nodes.shape[0], round(ms, 1), "ms")
    assert dev.nodes.shape[0] < full.nodes.shape[0]
    if pad == 0:  # the limit counts leaves by their 8 corners; a padded lattice can keep more
        assert dev.nodes.shape[0] <= lim
    got, want = _bricks(dev.nodes, dev.values, n_words), _bricks(o_nodes, o_values, n_words)
    assert got.keys() == want.keys()
    assert all(got[k] == want[k] for k in want)


@pytest.fixture(scope="module")
def ico_octree(gpu):
    v, t = L.mesh_displaced_icosphere(5, 0.8, 0.05, 4, 7)
    sbs, ms = L.sbs_octree_from_mesh(gpu, v, t, 7, 4, 0, 1)
    return v, t, sbs, ms


def test_mesh_octree_has_mixed_lods_and_covers_the_mesh(ico_octree):
    v, t, sbs, ms = ico_octree
    lod = sbs.nodes[:, 1] & 0xFFFF
    print(sbs.nodes.shape[0], "bricks, LODs", np.unique(lod, return_counts=True), round(ms, 1), "ms")
    assert 128 in lod and len(np.unique(lod)) > 1
    fine = sbs.nodes[lod == 128]
    cells = {(int(n[0] >> 16), int(n[0] & 0xFFFF), int(n[1] >> 16)) for n in fine}
    idx = np.floor((v.astype(np.float64) + 1.0) * 0.5 * 128).astype(int)
    covered = np.mean([(int(a), int(b), int(c)) in cells for a, b, c in idx])
    assert covered > 0.999, covered


def test_mesh_octree_scene_renders_at_parity(gpu, ico_octree):
    v, t, sbs, ms = ico_octree
    gpu.SetScene(sbs)
    sc = oracle.Scene(sbs, gpu.GetBVH()[0])
    wv, pr = L.orbit_camera(5, 1)
    res = compare_gbuffer(gpu, sc, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), wv, pr, 192, 160, 64)
    print(res)
    assert res["hits"] > 5000
    assert res["exact_t_fp32"] >= 0.999
    assert res["mask_mismatch_fp32"] <= 2
    assert res["depth_rel_max"] <= 1e-5

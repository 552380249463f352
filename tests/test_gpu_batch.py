"""Batched uneven workload (many scenes, one persistent launch): every item's
G-buffer equals the single-scene render of the same scene and camera, with
the default and with the autotuned work distribution."""
import numpy as np
import pytest

import litert_amd.api as L
from helpers import preset

pytestmark = pytest.mark.gpu

SPECS = [("ico", 8, 1, 37, 50), ("torus", 16, 4, 128, 128), ("ico", 32, 4, 200, 64), ("torus", 32, 1, 64, 256),
         ("ico", 16, 1, 1, 1), ("ico", 64, 1, 300, 17), ("torus", 8, 4, 96, 96)]


@pytest.fixture(scope="module")
def batch():
    rs, cams = [], []
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON)
    for k, (shape, lod, bpv, W, H) in enumerate(SPECS):
        if shape == "ico":
            s = L.displaced_icosphere(4, 0.7, 0.05, 3, seed=40 + k, lod=lod, brick_size=4, brick_pad=0, bytes_per_value=bpv)
        else:
            s = L.displaced_torus(0.6, 0.2, 0.02, 3, seed=40 + k, lod=lod, brick_size=4, brick_pad=0, bytes_per_value=bpv)
        r = L.CreateMultiRenderer(L.DEVICE_GPU)
        r.SetScene(s)
        r.SetPreset(p)
        rs.append(r)
        cams.append(L.orbit_camera(60 + k, 0, aspect=W / H))
    owner = L.CreateMultiRenderer(L.DEVICE_GPU)
    owner.SetPreset(p)
    yield owner, rs, cams
    for r in rs + [owner]:
        r.close()


def _check(owner, rs, cams):
    for i, (r, (wv, pr), spec) in enumerate(zip(rs, cams, SPECS)):
        W, H = spec[3], spec[4]
        t, n, prim = owner.BatchDownload(i)
        r.SetViewport(0, 0, W, H)
        r.UpdateCamera(wv, pr)
        t1, n1, p1 = r.RenderGBuffer(W, H)
        assert np.array_equal(prim, p1) and np.array_equal(t, t1) and np.array_equal(n, n1), i


def test_batch_matches_single_renders(batch):
    owner, rs, cams = batch
    wvs = np.stack([c[0] for c in cams])
    prs = np.stack([c[1] for c in cams])
    ms = owner.RenderBatch(rs, wvs, prs, [s[3] for s in SPECS], [s[4] for s in SPECS])
    print("batch kernel ms", ms)
    _check(owner, rs, cams)


def test_autotuned_batch_matches_single_renders(batch):
    owner, rs, cams = batch
    wvs = np.stack([c[0] for c in cams])
    prs = np.stack([c[1] for c in cams])
    owner.RenderBatch(rs, wvs, prs, [s[3] for s in SPECS], [s[4] for s in SPECS], autotune=True)
    chunk, blocks = owner.BatchConfig()
    print("tuned chunk", chunk, "blocks", blocks)
    assert chunk in (256, 512, 1024) and blocks > 0
    _check(owner, rs, cams)


def test_batch_rejects_bad_items(batch):
    owner, rs, cams = batch
    empty = L.CreateMultiRenderer(L.DEVICE_GPU)
    with pytest.raises(L.LrtError):
        owner.RenderBatch([empty], cams[0][0][None], cams[0][1][None], [8], [8])
    with pytest.raises(L.LrtError):
        owner.RenderBatch([rs[0]], cams[0][0][None], cams[0][1][None], [0], [8])
    empty.close()


@pytest.mark.parametrize("mode", [L.SDF_OCTREE_NODE_INTERSECT_NEWTON, L.SDF_OCTREE_NODE_INTERSECT_ST])
def test_seeded_batch_matches_single_seeded_renders(batch, mode):
    """With the owner's network and seeding on, each item equals the single
    scene's seeded render with the same network (random, always-confident
    weights, so every brick entry takes a seed)."""
    owner, rs, cams = batch
    rng = np.random.default_rng(17)
    params = rng.normal(0, 0.2, L.MLP_PARAMS).astype(np.float32)
    params[-3:] = [0.3, 0.6, 50.0]
    p = preset(mode)
    try:
        for r in rs + [owner]:
            r.SetPreset(p)
            r.MlpSetParams(params)
            r.SetSeeding(True, 0.9)
        wvs = np.stack([c[0] for c in cams])
        prs = np.stack([c[1] for c in cams])
        owner.RenderBatch(rs, wvs, prs, [s[3] for s in SPECS], [s[4] for s in SPECS], autotune=True)
        _check(owner, rs, cams)
    finally:
        for r in rs + [owner]:
            r.SetSeeding(False)
            r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON))

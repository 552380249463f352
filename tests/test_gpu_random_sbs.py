"""Randomised SBS scenes against the oracle's FP32 twin.

Bricks at random positions with random quantised corner values (a per-brick
offset plus noise, so rays both hit and miss and voxels hold several roots)
cover combinations of brick size, border padding and bytes per value that
the smooth generated scenes do not, under every root finder. The GPU and the
FP32 twin run the same arithmetic, so they must agree on which brick each ray
hits and, bit for bit, on its depth (within 1e-5 for the analytic cubic,
whose transcendental calls come from different math libraries).
"""
import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import MISS, preset

pytestmark = pytest.mark.gpu

MODES = [L.SDF_OCTREE_NODE_INTERSECT_ST, L.SDF_OCTREE_NODE_INTERSECT_ANALYTIC, L.SDF_OCTREE_NODE_INTERSECT_NEWTON,
         L.SDF_OCTREE_NODE_INTERSECT_IT]


def random_sbs(seed, bs, pad, bpv, lods, n_bricks):
    """DX-layout SBS with n_bricks bricks spread over the given LODs (they may
    overlap across LODs) and random packed values."""
    rng = np.random.default_rng(seed)
    v = bs + 2 * pad + 1
    per_word = 4 // bpv
    words = -(-v ** 3 // per_word)
    max_q = (1 << (8 * bpv)) - 1
    nodes = []
    for lod in lods:
        cells = rng.choice(lod ** 3, size=min(n_bricks // len(lods), lod ** 3), replace=False)
        for c in cells:
            x, y, z = c // (lod * lod), (c // lod) % lod, c % lod
            nodes.append((int(x), int(y), int(z), lod))
    values = np.zeros(len(nodes) * words, np.uint32)
    arr = []
    for i, (x, y, z, lod) in enumerate(nodes):
        offset = rng.uniform(-0.3, 0.8)
        q = np.clip(0.5 + 0.5 * (offset + rng.uniform(-0.5, 0.5, v ** 3)), 0.0, 1.0)
        qi = np.zeros(words * per_word, np.uint64)
        qi[:v ** 3] = np.round(q * max_q).astype(np.uint64)
        # element k in bits 8 * bpv * (k % per_word) of word k // per_word
        shifts = (8 * bpv * np.arange(per_word)).astype(np.uint64)
        packed = np.bitwise_or.reduce(qi.reshape(words, per_word) << shifts, axis=1)
        values[i * words:(i + 1) * words] = packed.astype(np.uint32)
        arr.append(((x << 16) | y, (z << 16) | lod, i * words, 0))
    hdr = L.SdfSBSHeader(bs, pad, bpv, L.SDF_SBS_NODE_LAYOUT_DX)
    return L.SdfSBS(hdr, np.array(arr, np.uint32), values)


# (brick_size, pad, bytes per value, LODs); the last four reach the largest
# accepted bricks (brick_size 16, 19^3 values with padding)
CASES = [(1, 0, 1, (16,)), (2, 0, 2, (8,)), (4, 0, 1, (8,)), (4, 1, 4, (4,)), (3, 1, 2, (6,)), (2, 1, 1, (8,)), (4, 0, 4, (6,)),
         (2, 0, 1, (4, 8)), (8, 1, 2, (4,)), (16, 1, 1, (2,)), (16, 0, 4, (3,)), (15, 1, 4, (2, 3))]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("mode", MODES)
def test_random_sbs_matches_fp32_twin(case, mode):
    bs, pad, bpv, lods = CASES[case]
    sbs = random_sbs(100 + case, bs, pad, bpv, lods, 160)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    p = preset(mode)
    W = H = 96
    n_hit = n_pix = n_prim_diff = n_exact = n_both = 0
    worst = 0.0
    for view in range(2):
        wv, pr = L.orbit_camera(500 + case, view)
        r.SetViewport(0, 0, W, H)
        r.SetPreset(p)
        t, n, prim = r.RenderGBuffer(W, H, wv, pr)
        o = sc.render(wv, pr, p, W, H, precision=32)
        n_pix += W * H
        n_hit += int((prim != MISS).sum())
        n_prim_diff += int((prim != o["prim"]).sum())
        both = (prim != MISS) & (prim == o["prim"])
        n_both += int(both.sum())
        ot = o["t"][both].astype(np.float32)
        n_exact += int((t[both] == ot).sum())
        if both.any():
            worst = max(worst, float(np.max(np.abs(t[both].astype(np.float64) - ot) / ot)))
    r.close()
    print(CASES[case], mode, "hits", n_hit, "prim differs", n_prim_diff, "bit-exact", n_exact, "/", n_both, "worst rel", worst)
    assert n_hit > 0.05 * n_pix and n_hit < 0.95 * n_pix
    assert n_prim_diff <= 1e-3 * n_pix
    if mode == L.SDF_OCTREE_NODE_INTERSECT_ANALYTIC:
        assert worst <= 1e-5
    else:
        assert n_exact >= 0.999 * n_both

"""The oracle's restatement of the reference's SBS builder for a distance
field (create_sdf_SBS(settings, header, sdf): construct_sdf_frame_octree +
frame_octree_to_SBS, utils/sparse_octree_builder.cpp:125-444) against what
the mathematics fixes for an exact distance field (a sphere, |grad f| = 1):

- Which cells the frame octree holds. A cell is refined while its centre lies
  within its half-diagonal of the surface or its corners change sign; for a
  1-Lipschitz field that condition holds for every ancestor of a cell where it
  holds, so a cell below the uniform top levels exists exactly when its parent
  satisfies it, and is a leaf when it does not satisfy it itself (or sits at
  the maximum depth). The test evaluates that characterisation level by level
  in FP64 and compares the set of kept bricks (position, LOD), allowing only
  cells whose decision lies within FP32 rounding of a threshold.
- What each brick stores: every lattice value (brick_size + 2 pad + 1 per
  axis, pad included) decodes to the FP64 distance at its world position
  within one quantisation step of the brick's diagonal range, and every kept
  brick satisfies the keep rule (min <= 0 < max + d_max).
- Coverage: every point of the sphere lies in a kept brick of the finest LOD.
"""
import numpy as np
import pytest

import oracle

R = 0.7
SQ3 = np.sqrt(3.0)


def _sphere(p):
    return np.linalg.norm(p, axis=-1) - R


def _corners(p, d):
    """World positions of the 8 corners (bit order x4 y2 z1) of cells p (n,3) of size d."""
    bits = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.float64)
    return 2.0 * (p[:, None, :] + bits[None, :, :]) * d - 1.0


def _cond(p, d):
    """Refinement test of add_node_rec and its distance to the decision
    boundary (for the FP32 tolerance)."""
    v = _sphere(2.0 * (p + 0.5) * d - 1.0)
    c = _sphere(_corners(p, d))
    mn, mx = c.min(axis=1), c.max(axis=1)
    near = np.abs(v) < SQ3 * d
    sign = mn * mx <= 0
    margin = np.minimum(np.abs(np.abs(v) - SQ3 * d), np.minimum(np.abs(mn), np.abs(mx)))
    return near | sign, near, margin


def _expected_leaves(depth):
    """(level, cells, fuzzy) of every frame-octree leaf at levels >= min(depth, 4)."""
    L4 = min(depth, 4)
    n = 1 << L4
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    out = []
    cur = g
    for level in range(L4, depth + 1):
        d = 2.0 ** -level
        cond, near, margin = _cond(cur, d)
        split = (near if level == L4 else cond) & (level < depth)
        fuzzy = margin < 1e-5
        out.append((level, cur[~split], fuzzy[~split]))
        kids = cur[split]
        bits = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.float64)
        cur = (2 * kids[:, None, :] + bits[None]).reshape(-1, 3)
    return out


def _lattice(p, d, bs, pad):
    r = np.arange(-pad, bs + pad + 1, dtype=np.float64)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    return 2.0 * d * p[:, None, :] - 1.0 + (2.0 * d / bs) * g[None]


def _decode(values, off, n_vals, bpv):
    per = 4 // bpv
    bits = 8 * bpv
    idx = np.arange(n_vals)
    w = values[off[:, None] + idx[None, :] // per].astype(np.uint64)
    return (w >> (bits * (idx % per)).astype(np.uint64)) & np.uint64((1 << bits) - 1)


@pytest.mark.parametrize("depth,bs,pad", [(5, 4, 0), (6, 2, 1), (6, 4, 1), (4, 3, 0)])
def test_octree_builder_matches_the_characterisation(depth, bs, pad):
    nodes, values = oracle.sbs_octree(0, [0, 0, 0, R], depth, bs, pad, 1)
    got = {(int(n[0] >> 16), int(n[0] & 0xFFFF), int(n[1] >> 16), int(n[1] & 0xFFFF)) for n in nodes}
    assert len(got) == len(nodes)  # no brick twice
    want, fuzzy_cells = set(), set()
    for level, cells, fuzzy in _expected_leaves(depth):
        d = 2.0 ** -level
        lat = _sphere(_lattice(cells, d, bs, pad))
        keep = (lat.min(axis=1) <= 0) & (lat.max(axis=1) > -2 * SQ3 * d)
        kfuzzy = (np.abs(lat.min(axis=1)) < 1e-5) | (np.abs(lat.max(axis=1) + 2 * SQ3 * d) < 1e-5)
        for c, k, f, kf in zip(cells.astype(int), keep, fuzzy, kfuzzy):
            key = (int(c[0]), int(c[1]), int(c[2]), 1 << level)
            if f or kf:
                fuzzy_cells.add(key)
            elif k:
                want.add(key)
    diff = (got ^ want) - fuzzy_cells
    assert not diff, sorted(diff)[:10]
    lods = {k[3] for k in got}
    if depth <= 4:
        assert lods == {1 << depth}
    else:
        assert (1 << depth) in lods and len(lods) > 1  # mixed LODs: coarse leaves inside the sphere


@pytest.mark.parametrize("bpv", [1, 2])
def test_octree_builder_values_decode_to_the_field(bpv):
    bs, pad, depth = 4, 1, 6
    nodes, values = oracle.sbs_octree(0, [0, 0, 0, R], depth, bs, pad, bpv)
    v = bs + 2 * pad + 1
    lod = (nodes[:, 1] & 0xFFFF).astype(np.float64)
    p = np.stack([nodes[:, 0] >> 16, nodes[:, 0] & 0xFFFF, nodes[:, 1] >> 16], -1).astype(np.float64)
    q = _decode(values, nodes[:, 2].astype(np.int64), v ** 3, bpv).astype(np.float64)
    d = 1.0 / lod
    d_max = 2 * SQ3 / lod
    maxv = float((1 << (8 * bpv)) - 1)
    f = np.stack([_sphere(_lattice(p[i:i + 1], d[i], bs, pad))[0] for i in range(len(p))])
    # the quantiser truncates maxv * (f + d_max) / (2 d_max), clamped to [0, maxv]
    expect = np.clip(np.floor(maxv * (f + d_max[:, None]) / (2 * d_max[:, None])), 0, maxv)
    step_err = np.abs(q - expect)
    # FP32 field values straddle a step boundary more often at 16 bits
    assert np.mean(step_err == 0) > (0.999 if bpv == 1 else 0.97) and step_err.max() <= 1
    lo = f.min(axis=1)
    assert np.all(lo <= 1e-6) and np.all(f.max(axis=1) > -d_max - 1e-6)


def test_octree_builder_covers_the_surface():
    depth, bs = 6, 4
    nodes, _ = oracle.sbs_octree(0, [0, 0, 0, R], depth, bs, 0, 1)
    fine = nodes[(nodes[:, 1] & 0xFFFF) == (1 << depth)]
    cells = {(int(n[0] >> 16), int(n[0] & 0xFFFF), int(n[1] >> 16)) for n in fine}
    rng = np.random.default_rng(3)
    u = rng.normal(size=(20000, 3))
    pts = R * u / np.linalg.norm(u, axis=1, keepdims=True)
    idx = np.floor((pts + 1.0) * 0.5 * (1 << depth)).astype(int)
    assert all((int(a), int(b), int(c)) in cells for a, b, c in idx)


def test_octree_builder_torus_has_mixed_lods_and_covers_it():
    depth = 6
    nodes, _ = oracle.sbs_octree(1, [0.6, 0.2], depth, 4, 0, 1)
    fine = nodes[(nodes[:, 1] & 0xFFFF) == (1 << depth)]
    cells = {(int(n[0] >> 16), int(n[0] & 0xFFFF), int(n[1] >> 16)) for n in fine}
    rng = np.random.default_rng(4)
    a, b = rng.uniform(0, 2 * np.pi, 20000), rng.uniform(0, 2 * np.pi, 20000)
    pts = np.stack([(0.6 + 0.2 * np.cos(b)) * np.cos(a), (0.6 + 0.2 * np.cos(b)) * np.sin(a), 0.2 * np.sin(b)], -1)
    idx = np.floor((pts + 1.0) * 0.5 * (1 << depth)).astype(int)
    assert all((int(x), int(y), int(z)) in cells for x, y, z in idx)


# ---- frame_octree_limit_nodes (:2229-2330) ----------------------------------
def _key_set(nodes):
    return {(int(n[0] >> 16), int(n[0] & 0xFFFF), int(n[1] >> 16), int(n[1] & 0xFFFF)) for n in nodes}


def test_node_limit_above_the_frame_size_changes_nothing():
    a, va = oracle.sbs_octree(0, [0, 0, 0, R], 6, 4, 0, 1)
    b, vb = oracle.sbs_octree(0, [0, 0, 0, R], 6, 4, 0, 1, 1 << 22)
    assert np.array_equal(a, b) and np.array_equal(va, vb)


def test_node_limit_is_monotone_and_bounds_the_bricks():
    counts = []
    for lim in (14000, 8000, 3000, 1000):
        nodes, _ = oracle.sbs_octree(0, [0, 0, 0, R], 6, 4, 0, 1, lim)
        counts.append(nodes.shape[0])
        # the limit counts leaves whose 8 corners pass the border test; the
        # kept bricks test their whole lattice, which can only keep fewer
        assert nodes.shape[0] <= lim
    assert counts == sorted(counts, reverse=True) and counts[-1] < counts[0]


def _weighted_diff(parent):
    """Mean |child corner - parent trilinear| over the 8 children x 8 corners,
    divided by (border children - 1), in FP64 (the merge order's key)."""
    p = np.asarray(parent, np.float64)
    pv = _sphere(_corners(p[None, :3], 2.0 ** -p[3]))[0]
    bits = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.float64)
    kids = 2 * p[None, :3] + bits
    kv = _sphere(_corners(kids, 2.0 ** -(p[3] + 1)))
    diff = 0.0
    for i in range(8):
        for j in range(8):
            q = 0.5 * (bits[i] + bits[j])
            w = np.array([(1 - q[0]) if not (c & 4) else q[0] for c in range(8)]) * \
                np.array([(1 - q[1]) if not (c & 2) else q[1] for c in range(8)]) * \
                np.array([(1 - q[2]) if not (c & 1) else q[2] for c in range(8)])
            diff += abs(kv[i, j] - float(w @ pv))
    d_max = 2 * SQ3 * 2.0 ** -(p[3] + 1)
    border = int(np.sum((kv.min(axis=1) <= 0) & (kv.max(axis=1) > -d_max)))
    return diff / (border - 1) if border > 1 else 1000.0


def test_node_limit_merges_the_best_approximated_parents_first():
    """In the layer it stops in, the merged parents (their children gone,
    which the unlimited build kept) approximate their children no worse than
    any parent left split: the merge order is the reference's mean difference
    between the children's corner values and the parent's trilinear
    interpolation."""
    depth, lim = 6, 8000
    full, _ = oracle.sbs_octree(0, [0, 0, 0, R], depth, 4, 0, 1)
    part, _ = oracle.sbs_octree(0, [0, 0, 0, R], depth, 4, 0, 1, lim)
    fine_full = {k for k in _key_set(full) if k[3] == 1 << depth}
    fine_part = {k for k in _key_set(part) if k[3] == 1 << depth}
    parents_full = {(x // 2, y // 2, z // 2) for x, y, z, _ in fine_full}
    parents_part = {(x // 2, y // 2, z // 2) for x, y, z, _ in fine_part}
    merged = parents_full - parents_part
    kept = parents_part
    assert len(merged) > 50 and len(kept) > 50
    wm = np.array([_weighted_diff((*c, depth - 1)) for c in merged])
    wk = np.array([_weighted_diff((*c, depth - 1)) for c in kept])
    assert wm.max() <= wk.min() + 1e-5, (wm.max(), wk.min())

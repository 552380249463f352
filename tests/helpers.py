"""Shared helpers for the parity tests: scene builders and error metrics."""
import functools

import numpy as np

import litert_amd.api as L
import oracle

MISS = 0xFFFFFFFF


@functools.lru_cache(maxsize=None)
def ico128_scene(seed=1):
    """Displaced icosphere(5) r 0.8, 3-octave noise amp 0.05, 128^3 effective
    (brick LOD 32 x 4^3 voxels), pad 1, 32-bit values."""
    return L.displaced_icosphere(5, 0.8, 0.05, 3, seed=seed, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)


@functools.lru_cache(maxsize=None)
def ico512_scene(seed=2):
    """Displaced icosphere(5), 4 octaves, 512^3 effective (LOD 128 x 4^3), pad 0, 8-bit
    values quantised over the voxel diagonal."""
    return L.displaced_icosphere(5, 0.8, 0.05, 4, seed=seed, lod=128, brick_size=4, brick_pad=0, bytes_per_value=1,
                                 voxel_diag_scale=True)


@functools.lru_cache(maxsize=None)
def oracle_scene(which):
    sbs = ico128_scene() if which == "ico128" else ico512_scene()
    pairs, _ = L.build_bvh(sbs)
    return oracle.Scene(sbs, pairs)


def preset(mode=L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_LAMBERT_NO_TEX, spp=1,
           ray_gen=L.RAY_GEN_MODE_REGULAR):
    p = L.getDefaultPreset()
    p.sdf_node_intersect = mode
    p.render_mode = render_mode
    p.spp = spp
    p.ray_gen_mode = ray_gen
    return p


def decode_oct(n):
    """decode_normal (Renderer/eye_ray.cpp:22-32), vectorised in float64."""
    x, y = n[..., 0].astype(np.float64), n[..., 1].astype(np.float64)
    z = 1.0 - np.abs(x) - np.abs(y)
    neg = z < 0
    sx = np.where(x >= 0, 1.0, -1.0)
    sy = np.where(y >= 0, 1.0, -1.0)
    x2 = np.where(neg, (1 - np.abs(y)) * sx, x)
    y2 = np.where(neg, (1 - np.abs(x)) * sy, y)
    v = np.stack([x2, y2, z], -1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def angle(a, b):
    return np.arccos(np.clip((decode_oct(a) * decode_oct(b)).sum(-1), -1.0, 1.0))


def grazing_or_seam(wv, pr, W, H, ys, xs, t, n, voxels_per_unit):
    """|cos(ray, n)| < 1e-2, or the hit lies within 1e-3 voxel of a voxel face."""
    out = []
    for y, x, tt, nn in zip(ys, xs, t, n):
        ray = oracle.eye_ray(wv, pr, int(x), int(y), 0.5, 0.5, W, H, 64)
        o, d = ray[:3], ray[4:7]
        c = abs(float(np.dot(d, _dec(nn))))
        p = (o + tt * d + 1.0) * voxels_per_unit
        seam = np.min(np.abs(p - np.round(p))) < 1e-3
        out.append(c < 1e-2 or seam)
    return np.array(out, bool)


def _dec(n):
    return decode_oct(np.asarray(n, np.float64)[None])[0]


def compare_gbuffer(r, sc, p, wv, pr, W, H, voxels_per_unit):
    """Renders one G-buffer on the GPU and scores it against the FP32 and FP64 oracle."""
    r.SetViewport(0, 0, W, H)
    r.SetPreset(p)
    r.UpdateCamera(wv, pr)
    t, n, prim = r.RenderGBuffer(W, H)
    o32 = sc.render(wv, pr, p, W, H, precision=32)
    o64 = sc.render(wv, pr, p, W, H, precision=64)
    g_hit, c_hit, d_hit = prim != MISS, o32["prim"] != MISS, o64["prim"] != MISS
    res = {}
    # bit-level agreement with the FP32 twin
    both32 = g_hit & c_hit
    res["mask_mismatch_fp32"] = int((g_hit != c_hit).sum())
    res["exact_t_fp32"] = float(np.mean(t[both32] == o32["t"][both32].astype(np.float32))) if both32.any() else 1.0
    # error against FP64
    mism = np.argwhere(g_hit != d_hit)
    res["mask_mismatch_fp64"] = len(mism)
    if len(mism):
        ys, xs = mism[:, 0], mism[:, 1]
        tt = np.where(g_hit[ys, xs], t[ys, xs], o64["t"][ys, xs])
        nn = np.where(g_hit[ys, xs][:, None], n[ys, xs], o64["n"][ys, xs])
        res["mismatch_all_grazing"] = bool(grazing_or_seam(wv, pr, W, H, ys, xs, tt, nn, voxels_per_unit).all())
    else:
        res["mismatch_all_grazing"] = True
    both = g_hit & d_hit
    rel = np.abs(t[both].astype(np.float64) - o64["t"][both]) / o64["t"][both]
    res["depth_rel_max"] = float(rel.max()) if rel.size else 0.0
    both3 = both & c_hit  # the FP32 twin's own distance from FP64 on the same pixels
    rel32 = np.abs(o32["t"][both3].astype(np.float32).astype(np.float64) - o64["t"][both3]) / o64["t"][both3]
    res["depth_rel_max_fp32_twin"] = float(rel32.max()) if rel32.size else 0.0
    ang = angle(n[both], o64["n"][both])
    ang32 = angle(o32["n"][both], o64["n"][both]) if (c_hit[both]).all() else np.zeros_like(ang)
    # seam pixels: the FP32 CPU path itself leaves the 1e-4 band there (normal
    # field of the trilinear interpolant is discontinuous across voxel faces)
    bad = (ang > 1e-4) & ~(ang32 > 1e-4)
    res["normal_bad"] = int(bad.sum())
    res["normal_p999"] = float(np.quantile(ang, 0.999)) if ang.size else 0.0
    res["n_pix"] = W * H
    res["hits"] = int(g_hit.sum())
    return res

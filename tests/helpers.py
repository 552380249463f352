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


# ---- position derivatives of the DR border term ----------------------------
# The reference's test_position_derivatives (diff_render/tests_diff_render.cpp:
# 359-562): one brick of size 2 holding the sphere r 0.8 (circle_smallest_scene,
# diff_render/DR_common.cpp:176-181), a 64 x 64 view from (0.3, 0, 2.5) with a
# 60 degree pinhole, an empty reference image, spp 64, border_spp 1024 and
# finite differences with delta 0.005, each mode averaged over 25 seeds.
POS_DER_W = POS_DER_H = 64
POS_DER_SEEDS = 25


@functools.lru_cache(maxsize=None)
def pos_der_scene():
    return L.grid_sbs("sphere", 0.8, 0.0, 0.0, 1, seed=1, brick_count=1, brick_size=2)


def pos_der_camera():
    return L.look_at_camera((0.3, 0.0, 2.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov=60.0, aspect=1.0, znear=0.01, zfar=100.0)


def pos_der_preset(mode, relax_eps, diff_mode):
    """tests_diff_render.cpp:404-425 for render mode `mode` (2 MASK, 3 LINEAR_DEPTH)."""
    dr = L.getDefaultPresetDR()
    dr.dr_render_mode = mode
    dr.dr_input_type = 1 if mode == 3 else 0
    dr.dr_reconstruction_flags = 2  # geometry
    dr.dr_diff_mode = diff_mode
    dr.spp = 64
    dr.border_spp = 1024
    dr.border_relax_eps = relax_eps
    dr.finite_diff_delta = 0.005
    return dr


def pos_der_distance_count(sbs):
    """param_count of the reference test: the distance values (values_f minus 3 colours x 8 corners per brick)."""
    return sbs.values_f.size - 24 * sbs.nodes.shape[0]


def pos_der_stats(grads):
    """stat::mean and stat::confidence(0.95) (utils/stat_utils.h:79-87, 168-183) over the seeds (rows)."""
    g = np.asarray(grads, np.float64)
    z = 1.959963984540054  # the 0.975 quantile of N(0, 1), what z_score(0.95) integrates to
    return g.mean(0), z * np.sqrt(g.var(0, ddof=1) / g.shape[0])


def pos_der_criteria(mean_default, conf_default, mean_fd, conf_fd):
    """tests_diff_render.cpp:483-538: relative error, average error in confidence
    widths and relative bias of the default-mode mean against the finite-difference mean."""
    m0, m1 = mean_default.astype(np.float64).copy(), mean_fd.astype(np.float64).copy()
    c0, c1 = np.maximum(conf_default, 1e-6), np.maximum(conf_fd, 1e-6)
    small = np.abs(m0) < 1e-3
    m0[small] = 0.0
    m1[small] = 0.0
    err = np.abs(m0 - m1) / np.maximum(c0, c1)
    total = 1e-6 + np.abs(m0).sum() + np.abs(m1).sum()
    rel = np.abs(m0 - m1).sum() / (0.5 * total)
    bias = (m0 - m1).sum() / (0.5 * total)
    return float(rel), float(err.mean()), float(bias)


def _lattice_field(lat, bs, p):
    """Trilinear value and world-space gradient of a one-brick lattice (bs+1)^3 over [-1, 1]^3."""
    q = (np.clip(p, -1.0, 1.0) + 1.0) * bs / 2.0
    i = np.clip(np.floor(q), 0, bs - 1).astype(np.int64)
    f = q - i
    v = np.zeros(p.shape[:-1])
    g = np.zeros(p.shape)
    for c in range(8):
        a, b, e = (c >> 2) & 1, (c >> 1) & 1, c & 1
        val = lat[i[..., 0] + a, i[..., 1] + b, i[..., 2] + e]
        wx = f[..., 0] if a else 1 - f[..., 0]
        wy = f[..., 1] if b else 1 - f[..., 1]
        wz = f[..., 2] if e else 1 - f[..., 2]
        v += wx * wy * wz * val
        g[..., 0] += (1 if a else -1) * wy * wz * val
        g[..., 1] += (1 if b else -1) * wx * wz * val
        g[..., 2] += (1 if e else -1) * wx * wy * val
    return v, g * (bs / 2.0)


def silhouette_inverse_gradient_norm(sbs, eye, nphi=360, nt=1000):
    """Mean of 1 / |grad f| over the silhouette of a one-brick SBS seen from
    `eye` (looking at the origin), weighted by screen-space arc length.

    Computed here from the lattice alone, independently of the oracle: for
    every angle around the view axis the screen radius of the tangent ray is
    bisected (a ray hits iff the trilinear field, densely sampled along it,
    reaches <= 0), the minimum along the tangent ray is refined by golden
    section, and the trilinear gradient is taken there.

    The relaxed-boundary estimator counts rays whose minimum distance along the
    ray is below eps, a band |grad f| times narrower than eps in space, and
    weights each by 1/eps and by the surface displacement -w_j / |grad f|; its
    expectation is therefore the true silhouette derivative times 1 / |grad f|
    (CalculateBorderRayDerivatives assumes a unit gradient,
    MultiRendererDR.cpp:1356-1360). This is the factor returned."""
    bs = int(sbs.header.brick_size)
    vs = bs + 1
    off = int(sbs.nodes[0, 2])
    lat = sbs.values_f[sbs.values[off + np.arange(vs ** 3)].astype(np.int64)].astype(np.float64).reshape(vs, vs, vs)
    eye = np.asarray(eye, np.float64)
    fwd = -eye / np.linalg.norm(eye)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    phi = np.linspace(0.0, 2 * np.pi, nphi, endpoint=False)
    radial = np.cos(phi)[:, None] * right[None] + np.sin(phi)[:, None] * up[None]
    ts = np.linspace(0.5, np.linalg.norm(eye) + 2.0, nt)

    def rays(r):
        d = fwd[None] + r[:, None] * radial
        return d / np.linalg.norm(d, axis=1, keepdims=True)

    def min_along(d):
        p = eye + ts[None, :, None] * d[:, None, :]
        v, _ = _lattice_field(lat, bs, p)
        v = np.where(np.all(np.abs(p) <= 1.0, axis=-1), v, np.inf)
        j = np.argmin(v, axis=1)
        return v[np.arange(len(j)), j], ts[j]

    lo, hi = np.zeros(nphi), np.full(nphi, 0.6)
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        hit = min_along(rays(mid))[0] <= 0
        lo, hi = np.where(hit, mid, lo), np.where(hit, hi, mid)
    d = rays(hi)
    _, tb = min_along(d)
    a, b = tb - 2 * (ts[1] - ts[0]), tb + 2 * (ts[1] - ts[0])
    for _ in range(60):
        x1, x2 = a + 0.382 * (b - a), a + 0.618 * (b - a)
        f1, _ = _lattice_field(lat, bs, eye + x1[:, None] * d)
        f2, _ = _lattice_field(lat, bs, eye + x2[:, None] * d)
        a, b = np.where(f1 < f2, a, x1), np.where(f1 < f2, x2, b)
    _, g = _lattice_field(lat, bs, eye + (0.5 * (a + b))[:, None] * d)
    gn = np.linalg.norm(g, axis=1)
    dphi = 2 * np.pi / nphi
    dr = (np.roll(hi, -1) - np.roll(hi, 1)) / (2 * dphi)
    w = np.sqrt(hi ** 2 + dr ** 2)
    return float(np.sum(w / gn) / np.sum(w))

"""Pins for the oracle functions that the reference's formulas alone leave open.

- ANALYTIC root finder (BVH2Common.cpp:362-436): the reference's
  R = (2a^2 - 9ab + 27c)/54 differs from Vieta's 2a^3 except where a^2 = a^3,
  i.e. a = c2/c3 in {0, 1}. There the analytic roots must be the true ones
  (FP64 bracketing to 1e-6); elsewhere they are not, which shows the check
  can tell the two formulas apart.
- Smoothed normals of _IN layouts (BVH2Common.cpp:838-906): on a linear field
  every voxel gradient is the plane normal, so the blend must return it; on a
  curved field the blend is continuous across voxel and brick faces while the
  plain trilinear gradient jumps there. A wrong neighbour slot or offset breaks
  the continuity at brick faces.
- Camera (eye_ray.cpp:146-160, render_common.h:25-55, with the GL
  perspective / lookAt conventions): every pixel's ray of a non-square image
  against the closed-form pinhole direction, which fixes the horizontal field
  of view and the aspect ratio.
- Reference-image resampling (PreprocessRefImages, MultiRendererDR.cpp:187-226)
  by what a supersampled bilinear filter fixes: the same size returns the
  image, an integer downsampling factor k returns the exact k x k block mean
  (sample count, positions and weights), a constant image stays constant at
  any size, and a linear ramp is reproduced in the interior (texel centres at
  (i + 0.5) / size and the bilinear weights).
"""
import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import decode_oct, preset

O = oracle.lib()
P = oracle._p


# ---- ANALYTIC --------------------------------------------------------------
def _cubic_basis(q0, d):
    """Column i: coefficients (t^0..t^3) of the trilinear weight of corner i along q0 + t d."""
    M = np.zeros((4, 8))
    for i in range(8):
        poly = np.array([1.0])
        for ax, bit in ((0, 4), (1, 2), (2, 1)):
            fac = np.array([q0[ax], d[ax]]) if i & bit else np.array([1.0 - q0[ax], -d[ax]])
            poly = np.convolve(poly, fac)
        M[:, i] = poly
    return M


def _first_root_fp64(coef, qfar):
    """First t in [0, qfar] with f(t) <= 0 for f = sum coef[k] t^k, by dense sampling and bisection."""
    f = lambda t: ((coef[3] * t + coef[2]) * t + coef[1]) * t + coef[0]
    ts = np.linspace(0.0, qfar, 20001)
    fv = f(ts)
    if fv[0] <= 1e-6:
        return 0.0, fv
    neg = np.nonzero(fv <= 0)[0]
    if len(neg) == 0:
        return None, fv
    lo, hi = ts[neg[0] - 1], ts[neg[0]]
    for _ in range(80):
        m = 0.5 * (lo + hi)
        lo, hi = (m, hi) if f(m) > 0 else (lo, m)
    return 0.5 * (lo + hi), fv


def _analytic_agreement(kind, n=1500, seed=11):
    """kind 'a0' / 'a1': distance values projected so that c2/c3 is 0 / 1; 'free': unconstrained."""
    rng = np.random.default_rng(seed)
    d_vox = 2.0 / 128
    agree = tested = 0
    while tested < n:
        q0 = rng.uniform(0, 1, 3)
        q0[rng.integers(3)] = 0.0
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if np.min(np.abs(d)) < 0.2:
            continue
        tx = [((1.0 if d[a] > 0 else 0.0) - q0[a]) / d[a] for a in range(3)]
        qfar = min(t for t in tx if t >= 0)
        if qfar < 0.05:
            continue
        M = _cubic_basis(q0, d)
        v = rng.uniform(-1, 1, 8) * d_vox
        if kind != "free":
            row = M[2] if kind == "a0" else M[2] - M[3]
            v = v - (row @ v) / (row @ row) * row
        coef = M @ v
        c3 = -coef[3] / d_vox  # the oracle's c3 is of -f / d
        if abs(c3) < 0.05 or (kind != "free" and abs(coef[2] - (0 if kind == "a0" else coef[3])) > 1e-12 * d_vox):
            continue
        t_bf, fv = _first_root_fp64(coef, qfar)
        if t_bf is not None and np.abs(fv).min() < 1e-9 * d_vox:
            continue  # tangential: the sign change is within noise
        t = np.zeros(1)
        it = np.zeros(1, np.uint32)
        hit = O.orc_voxel_root(L.SDF_OCTREE_NODE_INTERSECT_ANALYTIC, P(v), P(q0), P(d), d_vox, qfar, 64, P(t), P(it))
        tested += 1
        if t_bf is None:
            agree += hit == 0
        else:
            agree += hit == 1 and abs(t[0] - t_bf) < 1e-6
    return agree / tested


@pytest.mark.parametrize("kind", ["a0", "a1"])
def test_analytic_roots_exact_where_reference_R_equals_vieta(kind):
    assert _analytic_agreement(kind) >= 0.995


def test_analytic_roots_differ_from_vieta_elsewhere():
    """Away from a in {0, 1} the reference's R gives wrong roots on a visible
    fraction of voxels: the check above would catch a switch to either formula."""
    assert _analytic_agreement("free") < 0.9


# ---- smoothed normals --------------------------------------------------------
def _in_grid(field, brick_count=4, brick_size=4):
    """A grid SBS with the 27-neighbour table (ID32F_IRGB32F_IN) whose distance
    lattice holds field(p) at every lattice point p = 2 i / n - 1."""
    sbs = L.grid_sbs("sphere", 0.5, 0.0, 0.0, 1, seed=1, brick_count=brick_count, brick_size=brick_size, with_neighbors=True)
    pc = brick_count * brick_size + 1
    g = 2.0 * np.arange(pc) / (pc - 1) - 1.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    vf = sbs.values_f.copy()
    vf[:pc ** 3] = field(X, Y, Z).astype(np.float32).ravel()
    return L.SdfSBS(sbs.header, sbs.nodes, sbs.values, vf)


def _cast(sbs, rays, normal_mode):
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON)
    p.normal_mode = normal_mode
    sc = oracle.Scene(sbs, L.build_bvh(sbs)[0])
    return sc.cast_rays(p, rays, precision=64)


def _rays(orig, dirs):
    r = np.zeros((len(orig), 8), np.float32)
    r[:, :3], r[:, 4:7] = orig, dirs
    r[:, 3], r[:, 7] = 0.0, 1e9
    return r


def test_smoothed_normal_equals_plane_normal_on_linear_field():
    n = np.array([0.3, 0.5, 0.8])
    n /= np.linalg.norm(n)
    sbs = _in_grid(lambda x, y, z: n[0] * x + n[1] * y + n[2] * z - 0.1)
    rng = np.random.default_rng(3)
    tgt = rng.uniform(-0.7, 0.7, (400, 3))
    tgt -= ((tgt @ n) - 0.1)[:, None] * n  # points on the plane
    dirs = -(n[None] + 0.4 * rng.normal(size=(400, 3)))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    orig = tgt - 1.5 * dirs
    for mode in (L.NORMAL_MODE_GEOMETRY, L.NORMAL_MODE_SDF_SMOOTHED):
        out = _cast(sbs, _rays(orig, dirs), mode)
        ok = out[:, 1] != 0xFFFFFFFF
        assert ok.mean() > 0.95
        nn = decode_oct(out[ok][:, 3:5])
        assert np.arccos(np.clip(nn @ n, -1, 1)).max() < 1e-6


@pytest.mark.parametrize("face", ["voxel", "brick"])
def test_smoothed_normal_is_continuous_across_faces(face):
    """Parallel ray pairs 2e-5 apart across a voxel face (not a brick face) or a
    brick face, travelling inside the face plane (so each hit keeps its
    origin's side) toward a point of a sphere field on that plane: the
    smoothed normals of the two hits differ by O(1e-5), the plain trilinear
    gradients by the jump of the across-face derivative."""
    sbs = _in_grid(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.62)
    h = 2.0 / 16  # voxel size; bricks are 4 voxels
    planes = [-1 + k * h for k in range(1, 16) if (k % 4 != 0) == (face == "voxel")]
    rng = np.random.default_rng(5 if face == "voxel" else 6)
    orig, dirs = [], []
    for _ in range(600):
        ax = rng.integers(3)
        xf = planes[rng.integers(len(planes))]
        if abs(xf) > 0.4:
            continue
        # a point of the sphere r 0.62 on the plane p[ax] = xf, away from the other axes' faces
        rad = np.sqrt(0.62 ** 2 - xf ** 2)
        ang = rng.uniform(0, 2 * np.pi)
        p = np.zeros(3)
        o1, o2 = (ax + 1) % 3, (ax + 2) % 3
        p[ax], p[o1], p[o2] = xf, rad * np.cos(ang), rad * np.sin(ang)
        back = p.copy()
        back[ax] = 0.0
        back /= np.linalg.norm(back)  # outward, inside the plane
        e = np.zeros(3)
        e[ax] = 1e-5
        for s in (-1, 1):
            orig.append(p + s * e + 0.8 * back)
            dirs.append(-back)
    orig, dirs = np.array(orig), np.array(dirs)
    res = {}
    for mode in (L.NORMAL_MODE_GEOMETRY, L.NORMAL_MODE_SDF_SMOOTHED):
        out = _cast(sbs, _rays(orig, dirs), mode)
        a, b = out[0::2], out[1::2]
        ok = (a[:, 1] != 0xFFFFFFFF) & (b[:, 1] != 0xFFFFFFFF)
        na, nb = decode_oct(a[ok][:, 3:5]), decode_oct(b[ok][:, 3:5])
        res[mode] = np.arccos(np.clip((na * nb).sum(1), -1, 1))
    sm, geo = res[L.NORMAL_MODE_SDF_SMOOTHED], res[L.NORMAL_MODE_GEOMETRY]
    print(face, len(sm), "smoothed max", sm.max(), "geometric median", np.median(geo))
    assert len(sm) > 100
    assert sm.max() < 5e-4
    assert np.median(geo) > 20 * sm.max()


# ---- camera --------------------------------------------------------------------
def test_eye_rays_match_closed_form_pinhole_non_square():
    """Every pixel centre of a 96 x 64 image (aspect 1.5, vertical fov 50):
    dir = normalize(f + u r + v up), u = (2 (x + 0.5) / W - 1) aspect tan(fov/2),
    v = (1 - 2 (y + 0.5) / H) tan(fov/2), with (r, up, f) the lookAt frame."""
    eye, tgt, upv = np.array([0.4, -0.9, 2.2]), np.array([0.1, 0.2, -0.3]), np.array([0.0, 1.0, 0.0])
    W, H, fov = 96, 64, 50.0
    wv, pr = L.look_at_camera(eye, tgt, upv, fov=fov, aspect=W / H)
    f = (tgt - eye) / np.linalg.norm(tgt - eye)
    r = np.cross(f, upv)
    r /= np.linalg.norm(r)
    up = np.cross(r, f)
    th = np.tan(np.radians(fov) / 2)
    worst = 0.0
    for y in range(0, H, 3):
        for x in range(0, W, 5):
            ray = oracle.eye_ray(wv, pr, x, y, 0.5, 0.5, W, H, 64)
            u = (2 * (x + 0.5) / W - 1) * (W / H) * th
            v = (1 - 2 * (y + 0.5) / H) * th
            d = f + u * r + v * up
            d /= np.linalg.norm(d)
            assert np.allclose(ray[:3], eye, atol=1e-5)
            worst = max(worst, np.linalg.norm(ray[4:7] - d))
    assert worst < 1e-6, worst


# ---- reference-image resampling --------------------------------------------
def _rand_images(n, H, W, seed):
    return np.random.default_rng(seed).uniform(0, 1, (n, H, W, 4)).astype(np.float32)


def test_resample_same_size_returns_the_image():
    """Up to the rounding of u * size - 0.5 (weights of a few 1e-6 on the
    neighbouring texel), which the library skips by passing references of the
    render size through unchanged."""
    img = _rand_images(2, 37, 53, 1)
    out = oracle.resample_reference(img, 53, 37)
    assert np.max(np.abs(out - img)) <= 1e-5


@pytest.mark.parametrize("k", [2, 3])
def test_resample_integer_downsampling_is_the_block_mean(k):
    W, H = 24, 16
    img = _rand_images(2, H * k, W * k, 2 + k)
    out = oracle.resample_reference(img, W, H)
    block = img.astype(np.float64).reshape(2, H, k, W, k, 4).mean(axis=(2, 4))
    assert np.max(np.abs(out - block)) <= 2e-6


@pytest.mark.parametrize("size", [(40, 30), (64, 48), (100, 75)])
def test_resample_keeps_a_constant_image(size):
    img = np.full((1, 48, 64, 4), 0.3, np.float32)
    out = oracle.resample_reference(img, *size)
    assert out.shape == (1, size[1], size[0], 4)
    assert np.max(np.abs(out - 0.3)) <= 1e-6


def test_resample_reproduces_a_linear_ramp_inside():
    """Upsampling 40 -> 60 (one sample per pixel): at render pixel x the sample
    sits at texel coordinate (x + 0.5) * 40 / 60 - 0.5, and a bilinear filter
    returns the ramp's value there wherever both texels are inside."""
    W0, H0, W, H = 40, 20, 60, 30
    ramp = np.zeros((1, H0, W0, 4), np.float32)
    ramp[0, :, :, 0] = np.arange(W0, dtype=np.float32)[None, :] * 0.25
    ramp[0, :, :, 1] = np.arange(H0, dtype=np.float32)[:, None] * 0.5
    out = oracle.resample_reference(ramp, W, H)
    tx = (np.arange(W) + 0.5) * W0 / W - 0.5
    ty = (np.arange(H) + 0.5) * H0 / H - 0.5
    ix = (tx >= 0) & (tx <= W0 - 1)
    iy = (ty >= 0) & (ty <= H0 - 1)
    assert np.max(np.abs(out[0][:, ix, 0] - 0.25 * tx[ix][None, :])) <= 1e-5
    assert np.max(np.abs(out[0][iy, :, 1] - 0.5 * ty[iy][:, None])) <= 1e-5

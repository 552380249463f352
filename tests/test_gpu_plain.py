"""GPU parity of the plain primary-ray path against the CPU oracle.

Tolerances: hit mask may differ on at most 1e-4 of the pixels
and only on grazing/seam pixels; depth within 1e-5 relative and normals
within 1e-4 rad of the FP64 oracle. The FP32 oracle is the reference CPU
path restated operation for operation, so the GPU is also compared with it
directly (it is expected to agree bit for bit on almost every pixel).
"""
import os

import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import MISS, angle, decode_oct, ico128_scene, ico512_scene, compare_gbuffer as _compare, oracle_scene, preset

pytestmark = pytest.mark.gpu

MODES = [L.SDF_OCTREE_NODE_INTERSECT_ST, L.SDF_OCTREE_NODE_INTERSECT_ANALYTIC, L.SDF_OCTREE_NODE_INTERSECT_NEWTON,
         L.SDF_OCTREE_NODE_INTERSECT_IT]


@pytest.fixture(scope="module")
def r_ico128():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(ico128_scene())
    yield r
    r.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("view", [0, 1, 2])
def test_ico128_gbuffer_parity(r_ico128, mode, view):
    sc = oracle_scene("ico128")
    wv, pr = L.orbit_camera(7, view)
    res = _compare(r_ico128, sc, preset(mode), wv, pr, 256, 256, voxels_per_unit=32 * 4 / 2.0)
    print(mode, view, res)
    assert res["hits"] > 5000
    assert res["mask_mismatch_fp32"] <= 1e-4 * res["n_pix"]
    assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
    assert res["depth_rel_max"] <= 1e-5
    assert res["normal_bad"] == 0
    if mode != L.SDF_OCTREE_NODE_INTERSECT_ANALYTIC:  # transcendental calls differ between libm and the device
        assert res["exact_t_fp32"] > 0.999


def test_ico512_gbuffer_parity():
    sbs = ico512_scene()
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    sc = oracle_scene("ico512")
    wv, pr = L.orbit_camera(11, 0)
    res = _compare(r, sc, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), wv, pr, 512, 512, voxels_per_unit=128 * 4 / 2.0)
    print(res)
    assert res["mask_mismatch_fp32"] <= 1e-4 * res["n_pix"]
    assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
    assert res["depth_rel_max"] <= 1e-5
    assert res["normal_bad"] == 0
    assert res["exact_t_fp32"] > 0.999


def test_full_size_views_sampled_against_oracle():
    """The benchmark's own shape: 8 orbit views x 2048x2048 of the 512^3
    scene in one launch. 512 seeded pixels per view (image corners and
    edges included) against the FP32 oracle's RayQuery_NearestHit on the
    same eye rays: hit mask equal but for at most one grazing sample,
    t bit-identical on >= 99.9 % of the hits and within 1e-6 relative on
    all, normals within 1e-5 rad."""
    sbs = ico512_scene()
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    sc = oracle_scene("ico512")
    W = H = 2048
    cams = [L.orbit_camera(2, i) for i in range(8)]
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r.RenderViewsDevice(wvs, prs, W, H, 0)
    rng = np.random.default_rng(8)
    n_hits = n_exact = n_mask = 0
    for v in range(8):
        t, n, prim = r.DownloadView(v, W, H)
        xs = np.concatenate([[0, W - 1, 0, W - 1], rng.integers(0, W, 508)])
        ys = np.concatenate([[0, 0, H - 1, H - 1], rng.integers(0, H, 508)])
        rays = np.array([oracle.eye_ray(wvs[v], prs[v], int(x), int(y), 0.5, 0.5, W, H) for x, y in zip(xs, ys)], np.float32)
        ref = sc.cast_rays(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), rays, precision=32)
        g_prim, r_prim = prim[ys, xs], ref[:, 1].astype(np.uint32)
        n_mask += int(((g_prim != MISS) != (r_prim != MISS)).sum())
        hit = (g_prim != MISS) & (g_prim == r_prim)
        assert hit.sum() > 50, v
        gt, rt = t[ys, xs][hit], ref[hit, 0].astype(np.float32)
        n_hits += int(hit.sum())
        n_exact += int((gt == rt).sum())
        assert np.max(np.abs(gt.astype(np.float64) - rt) / rt) <= 1e-6, v
        assert np.max(angle(n[ys, xs][hit], ref[hit][:, 3:5].astype(np.float32))) <= 1e-5, v
    print("hits", n_hits, "bit-exact", n_exact, "mask mismatches", n_mask)
    assert n_mask <= 1
    assert n_exact >= 0.999 * n_hits
    r.close()


def _culled_and_traced(r, wvs, prs, W, H):
    """G-buffers of one multi-view launch with the tile cull on and off, and
    the culled-tile count."""
    out = []
    for on in (True, False):
        r.SetTileCulling(on)
        r.RenderViewsDevice(wvs, prs, W, H, 0)
        out.append([r.DownloadView(v, W, H) for v in range(wvs.shape[0])])
    r.SetTileCulling(True)
    culled = int(r.CountWorkEx(wvs, prs, W, H)[6])
    return out[0], out[1], culled


@pytest.mark.parametrize("case", ["orbit512", "random512", "orbit128_st", "spp4_random", "spp4_regular", "spp2_regular"])
def test_tile_culling_leaves_images_identical(case):
    """Tiles culled by their ray pyramid are written exactly as tracing writes
    them: G-buffers with the cull on and off are bit-identical (t, normal,
    primitive) over whole images, for orbit cameras, cameras around and
    inside the scene with any fov and aspect, an indexed FP32 scene with sphere
    tracing, and jittered / regular multi-sample presets. Orbit views cull most
    tiles; a regular pattern whose samples leave the pixel (spp 2) culls none."""
    rng = np.random.default_rng(31)
    sbs, mode, spp, gen = ico512_scene(), L.SDF_OCTREE_NODE_INTERSECT_NEWTON, 1, L.RAY_GEN_MODE_REGULAR
    if case == "orbit128_st":
        sbs, mode = ico128_scene(), L.SDF_OCTREE_NODE_INTERSECT_ST
    if case.startswith("spp"):
        spp = int(case[3])
        gen = L.RAY_GEN_MODE_RANDOM if case.endswith("random") else L.RAY_GEN_MODE_REGULAR
    if case == "random512":
        cams = []
        for _ in range(8):
            d = rng.normal(size=3)
            eye = d / np.linalg.norm(d) * rng.uniform(0.3, 4.0)
            cams.append(L.look_at_camera(eye, rng.uniform(-0.3, 0.3, 3), rng.normal(size=3), fov=float(rng.uniform(20, 110)),
                                         aspect=float(rng.uniform(0.5, 2.0))))
        W, H = 320, 240
    else:
        cams = [L.orbit_camera(11, i) for i in range(4)]
        W, H = 512, 384
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    r.SetPreset(preset(mode, render_mode=L.MULTI_RENDER_MODE_NORMAL, spp=spp, ray_gen=gen))
    on, off, culled = _culled_and_traced(r, wvs, prs, W, H)
    r.close()
    for v in range(len(cams)):
        for a, b in zip(on[v], off[v]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (case, v)
    tiles = len(cams) * ((W + 7) // 8) * ((H + 7) // 8)
    print(case, "culled", culled, "of", tiles)
    if case == "spp2_regular":
        assert culled == 0
    elif case != "random512":
        assert culled > 0.3 * tiles


@pytest.mark.parametrize("rmode", [L.MULTI_RENDER_MODE_PHONG_NO_TEX, L.MULTI_RENDER_MODE_LAMBERT_NO_TEX, L.MULTI_RENDER_MODE_LINEAR_DEPTH])
def test_tile_culling_leaves_colour_images_identical(rmode):
    """RenderFloat (the shading path, shadow rays included) with the tile cull
    on and off: bit-identical colour images; a culled tile shades as a miss."""
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(ico512_scene())
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=rmode)
    imgs = []
    for on in (True, False):
        r.SetTileCulling(on)
        imgs.append([r.RenderFloat(384, 256, *L.orbit_camera(13, i), p) for i in range(2)])
    r.SetTileCulling(True)
    r.SetPreset(p)
    culled = int(r.CountWorkEx(*[np.stack(x) for x in zip(*[L.orbit_camera(13, i) for i in range(2)])], 384, 256)[6])
    r.close()
    for a, b in zip(*imgs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert culled > 0


def test_random_cameras_sampled_against_oracle():
    """12 seeded look-at cameras around and inside the 512^3 scene (eye at
    radius 0.3-4, target jittered off the origin, random up vector, fov
    20-110 degrees, aspect 0.5-2, so waves mix direction octants), one
    multi-view launch at 320x240, 256 seeded pixels per view against the FP32
    oracle's RayQuery_NearestHit on the same eye rays: hit mask equal but for
    at most one grazing sample, t bit-identical on >= 99.9 % of the hits."""
    sbs = ico512_scene()
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    sc = oracle_scene("ico512")
    rng = np.random.default_rng(21)
    W, H = 320, 240
    cams = []
    for _ in range(12):
        d = rng.normal(size=3)
        eye = d / np.linalg.norm(d) * rng.uniform(0.3, 4.0)
        target = rng.uniform(-0.3, 0.3, 3)
        up = rng.normal(size=3)
        cams.append(L.look_at_camera(eye, target, up, fov=float(rng.uniform(20, 110)), aspect=float(rng.uniform(0.5, 2.0))))
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r.RenderViewsDevice(wvs, prs, W, H, 0)
    n_hits = n_exact = n_mask = 0
    for v in range(len(cams)):
        t, n, prim = r.DownloadView(v, W, H)
        xs, ys = rng.integers(0, W, 256), rng.integers(0, H, 256)
        rays = np.array([oracle.eye_ray(wvs[v], prs[v], int(x), int(y), 0.5, 0.5, W, H) for x, y in zip(xs, ys)], np.float32)
        ref = sc.cast_rays(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), rays, precision=32)
        g_prim, r_prim = prim[ys, xs], ref[:, 1].astype(np.uint32)
        n_mask += int(((g_prim != MISS) != (r_prim != MISS)).sum())
        hit = (g_prim != MISS) & (g_prim == r_prim)
        gt, rt = t[ys, xs][hit], ref[hit, 0].astype(np.float32)
        n_hits += int(hit.sum())
        n_exact += int((gt == rt).sum())
        if hit.any():
            assert np.max(np.abs(gt.astype(np.float64) - rt) / rt) <= 1e-6, v
            # an eye inside the surface sees hits in saturated voxels (all 8
            # values -d_max): a zero gradient, a NaN normal in both
            gn, rn = n[ys, xs][hit], ref[hit][:, 3:5].astype(np.float32)
            nan_g, nan_r = np.isnan(gn).any(-1), np.isnan(rn).any(-1)
            assert np.array_equal(nan_g, nan_r), v
            if (~nan_g).any():
                assert np.max(angle(gn[~nan_g], rn[~nan_g])) <= 1e-5, v
    print("hits", n_hits, "bit-exact", n_exact, "mask mismatches", n_mask)
    assert n_hits > 500
    assert n_mask <= 1
    assert n_exact >= 0.999 * n_hits
    r.close()


def test_scene_beyond_the_caches_sampled_against_oracle():
    """The 2048^3 displaced icosphere (brick LOD 512 x 4^3, 8-bit over the
    voxel diagonal: 1.7 M bricks, a 24-deep BVH, about 440 MB of scene arrays,
    past the 256 MB Infinity Cache): two orbit views at 1024x1024, 1,024
    seeded pixels per view against the FP32 oracle's RayQuery_NearestHit on
    the same eye rays, with the bars of the benchmark-shape test above."""
    sbs = L.displaced_icosphere(6, 0.8, 0.05, 4, seed=2, lod=512, brick_size=4, brick_pad=0, bytes_per_value=1, voxel_diag_scale=True)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, depth = r.GetBVH()
    assert sbs.nodes.shape[0] > 1_000_000 and depth > 20
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    sc = oracle.Scene(sbs, pairs)
    W = H = 1024
    cams = [L.orbit_camera(4, i) for i in range(2)]
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r.RenderViewsDevice(wvs, prs, W, H, 0)
    rng = np.random.default_rng(9)
    n_hits = n_exact = n_mask = 0
    for v in range(2):
        t, n, prim = r.DownloadView(v, W, H)
        xs = np.concatenate([[0, W - 1, 0, W - 1], rng.integers(0, W, 1020)])
        ys = np.concatenate([[0, 0, H - 1, H - 1], rng.integers(0, H, 1020)])
        rays = np.array([oracle.eye_ray(wvs[v], prs[v], int(x), int(y), 0.5, 0.5, W, H) for x, y in zip(xs, ys)], np.float32)
        ref = sc.cast_rays(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), rays, precision=32)
        g_prim, r_prim = prim[ys, xs], ref[:, 1].astype(np.uint32)
        n_mask += int(((g_prim != MISS) != (r_prim != MISS)).sum())
        hit = (g_prim != MISS) & (g_prim == r_prim)
        assert hit.sum() > 100, v
        gt, rt = t[ys, xs][hit], ref[hit, 0].astype(np.float32)
        n_hits += int(hit.sum())
        n_exact += int((gt == rt).sum())
        assert np.max(np.abs(gt.astype(np.float64) - rt) / rt) <= 1e-6, v
        assert np.max(angle(n[ys, xs][hit], ref[hit][:, 3:5].astype(np.float32))) <= 1e-5, v
    print("hits", n_hits, "bit-exact", n_exact, "mask mismatches", n_mask)
    assert n_mask <= 1
    assert n_exact >= 0.999 * n_hits
    r.close()


@pytest.mark.parametrize("rmode", [L.MULTI_RENDER_MODE_MASK, L.MULTI_RENDER_MODE_LINEAR_DEPTH, L.MULTI_RENDER_MODE_DEPTH,
                                   L.MULTI_RENDER_MODE_NORMAL, L.MULTI_RENDER_MODE_LAMBERT_NO_TEX,
                                   L.MULTI_RENDER_MODE_PRIMITIVE, L.MULTI_RENDER_MODE_ST_ITERATIONS,
                                   L.MULTI_RENDER_MODE_LAMBERT, L.MULTI_RENDER_MODE_PHONG, L.MULTI_RENDER_MODE_PHONG_NO_TEX,
                                   L.MULTI_RENDER_MODE_BARYCENTRIC, L.MULTI_RENDER_MODE_TEX_COORDS])
def test_render_float_modes(r_ico128, rmode):
    """RenderFloat colour per render mode vs the FP32 oracle (eye_ray.cpp:162-489)."""
    sc = oracle_scene("ico128")
    wv, pr = L.orbit_camera(3, 1)
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_ST, render_mode=rmode)
    img = r_ico128.RenderFloat(128, 128, wv, pr, p)
    ref = sc.render(wv, pr, p, 128, 128, precision=32, force_normal=False, rgba=True)["rgba"]
    diff = np.abs(img - ref).max(axis=-1)
    assert np.mean(diff <= 1e-5) > 0.9995, (rmode, np.sort(diff.ravel())[-10:])


@pytest.mark.parametrize("rmode", [L.MULTI_RENDER_MODE_LAMBERT, L.MULTI_RENDER_MODE_PHONG, L.MULTI_RENDER_MODE_PHONG_NO_TEX,
                                   L.MULTI_RENDER_MODE_DIFFUSE])
def test_shading_with_brick_colours_and_shadow_rays(rmode):
    """Lambert and Phong shading (eye_ray.cpp:376-483) of a coloured brick set:
    brick colours through the colour layout, and for Phong a shadow ray per
    hit toward the direct light (a secondary RayQuery_NearestHit from just
    before the hit). Colour vs the FP32 oracle."""
    sbs = L.grid_sbs("torus", 0.6, 0.2, 0.02, 3, seed=3, brick_count=16, brick_size=4)
    vs = sbs.header.brick_size + 2 * sbs.header.brick_pad + 1
    off = sbs.nodes[:, 2].astype(np.int64)
    ci = np.unique(sbs.values[(off[:, None] + vs ** 3 + np.arange(8)[None, :]).ravel()].astype(np.int64))
    vf = sbs.values_f.copy()
    rng = np.random.default_rng(5)
    for k in range(3):
        vf[ci + k] = rng.uniform(0.1, 0.9, ci.size).astype(np.float32)
    sbs = L.SdfSBS(sbs.header, sbs.nodes, sbs.values, vf)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    for view in range(2):
        wv, pr = L.orbit_camera(7, view)
        p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=rmode)
        img = r.RenderFloat(160, 160, wv, pr, p)
        ref = sc.render(wv, pr, p, 160, 160, precision=32, force_normal=False, rgba=True)["rgba"]
        diff = np.abs(img - ref).max(axis=-1)
        assert np.mean(diff <= 1e-5) > 0.9995, (rmode, view, np.sort(diff.ravel())[-10:])
        assert np.count_nonzero(img[..., :3].sum(-1) > 0) > 2000
    r.close()


def test_phong_shadow_rays_find_occluders():
    """A sphere (r 0.25) inside a torus ring (R 0.6, r 0.15), both sparse
    8-bit brick sets merged into one scene: the sphere shadows the ring's far
    inner side. Phong colour vs the FP32 oracle, and pixels that face the
    light but equal their Lambert colour (no diffuse/specular Phong term:
    the shadow ray hit) must exist on the GPU image."""
    a = L.displaced_torus(R=0.6, r=0.15, amplitude=0.0, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1)
    b = L.displaced_icosphere(subdiv=4, radius=0.25, amplitude=0.0, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1)
    assert not set(map(tuple, a.nodes[:, :2].tolist())) & set(map(tuple, b.nodes[:, :2].tolist()))
    nb = b.nodes.copy()
    nb[:, 2] += a.values.size
    sbs = L.SdfSBS(a.header, np.concatenate([a.nodes, nb]), np.concatenate([a.values, b.values]))
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    wv, pr = L.orbit_camera(7, 0)
    ph_p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_PHONG)
    ph = r.RenderFloat(160, 160, wv, pr, ph_p)
    ref = sc.render(wv, pr, ph_p, 160, 160, precision=32, force_normal=False, rgba=True)["rgba"]
    diff = np.abs(ph - ref).max(axis=-1)
    assert np.mean(diff <= 1e-5) > 0.9995, np.sort(diff.ravel())[-10:]
    lam = r.RenderFloat(160, 160, wv, pr, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_LAMBERT))
    t, n, prim = r.RenderGBuffer(160, 160, wv, pr)
    hit = prim != MISS
    lit = (decode_oct(n) @ (np.ones(3) / np.sqrt(3.0))) > 0.2  # outward normals, surfaces seen from outside
    same = np.all(np.abs(ph - lam) < 1e-7, axis=-1)
    assert np.count_nonzero(hit & same & lit) > 50  # in shadow
    assert np.count_nonzero(hit & ~same) > 500  # lit: Phong adds diffuse + specular
    r.close()


def _custom_lights():
    return [L.make_light(L.LIGHT_TYPE_DIRECT, (0.0, 1.0, 0.0), (0.5, 0.4, 0.3)),
            L.make_light(L.LIGHT_TYPE_POINT, (0.0, 0.9, 0.9), (1.5, 1.2, 1.0)),
            L.make_light(L.LIGHT_TYPE_DIRECT, (0.6, 0.0, 0.8), (0.3, 0.3, 0.6)),
            L.make_light(L.LIGHT_TYPE_AMBIENT, (0, 0, 0), (0.1, 0.12, 0.15))]


@pytest.mark.parametrize("rmode", [L.MULTI_RENDER_MODE_LAMBERT, L.MULTI_RENDER_MODE_LAMBERT_NO_TEX, L.MULTI_RENDER_MODE_PHONG,
                                   L.MULTI_RENDER_MODE_PHONG_NO_TEX])
def test_shading_with_set_lights(rmode):
    """SetLights (eye_ray.h:158): two direct lights, a point light (Lambert's
    1/l^2 falloff; Phong's shadow ray toward its position) and an ambient
    light, on the sphere-in-torus scene; colour vs the FP32 oracle with the
    same lights, and different from the default lights."""
    a = L.displaced_torus(R=0.6, r=0.15, amplitude=0.0, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1)
    b = L.displaced_icosphere(subdiv=4, radius=0.25, amplitude=0.0, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1)
    nb = b.nodes.copy()
    nb[:, 2] += a.values.size
    sbs = L.SdfSBS(a.header, np.concatenate([a.nodes, nb]), np.concatenate([a.values, b.values]))
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    wv, pr = L.orbit_camera(7, 0)
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=rmode)
    default = r.RenderFloat(128, 128, wv, pr, p)
    r.SetLights(_custom_lights())
    sc.set_lights(_custom_lights())
    img = r.RenderFloat(128, 128, wv, pr, p)
    ref = sc.render(wv, pr, p, 128, 128, precision=32, force_normal=False, rgba=True)["rgba"]
    diff = np.abs(img - ref).max(axis=-1)
    assert np.mean(diff <= 1e-5) > 0.9995, np.sort(diff.ravel())[-10:]
    assert np.abs(img - default).max() > 0.1
    with pytest.raises(L.LrtError):
        r.SetLights(_custom_lights() * 3)  # more than LRT_MAX_LIGHTS
    r.close()


def test_render_rgba8_and_timer(r_ico128):
    wv, pr = L.orbit_camera(3, 2)
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_LAMBERT_NO_TEX)
    f = r_ico128.RenderFloat(64, 64, wv, pr, p)
    u = r_ico128.Render(64, 64, wv, pr, p)
    exp = (255 * np.clip(f, 0, 1)).astype(np.uint32)
    packed = exp[..., 0] | (exp[..., 1] << 8) | (exp[..., 2] << 16) | (exp[..., 3] << 24)
    assert np.array_equal(u, packed)
    assert r_ico128.GetExecutionTime("CastRaySingleBlock")[0] > 0


def test_spp_random_rays(r_ico128):
    """spp 4 with PCG3D jitter (RAY_GEN_MODE_RANDOM) matches the oracle."""
    sc = oracle_scene("ico128")
    wv, pr = L.orbit_camera(5, 0)
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_LINEAR_DEPTH, spp=4,
               ray_gen=L.RAY_GEN_MODE_RANDOM)
    r_ico128.setSeed(12345)
    img = r_ico128.RenderFloat(96, 96, wv, pr, p)
    ref = sc.render(wv, pr, p, 96, 96, seed=12345, precision=32, rgba=True)["rgba"]
    assert np.mean(np.abs(img - ref).max(-1) <= 1e-6) > 0.999


def test_cast_rays_random(r_ico128):
    """Batched RayQuery_NearestHit on rays from outside, inside and on axis."""
    sc = oracle_scene("ico128")
    rng = np.random.default_rng(0)
    n = 4096
    o = rng.uniform(-1.5, 1.5, (n, 3))
    d = rng.normal(size=(n, 3))
    d[:256, 1:] = 0.0  # axis-aligned directions exercise SafeInverse with zeros
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 1e9
    r_ico128.SetPreset(preset())  # the oracle's preset (the fixture's default is another root finder)
    hits = r_ico128.RayQuery_NearestHit(rays)
    g = np.array([(h.t, h.primId, h.coords[2], h.coords[3]) for h in hits])
    ref = sc.cast_rays(preset(), rays, precision=32)
    gh, rh = g[:, 1] != MISS, ref[:, 1] != MISS
    assert (gh != rh).sum() <= 2
    both = gh & rh
    assert np.mean(g[both, 0].astype(np.float32) == ref[both, 0].astype(np.float32)) > 0.999


def test_cast_rays_on_grid_planes(r_ico128):
    """Axis-aligned rays whose origins lie exactly on brick and voxel planes:
    a zero direction component makes the inverse infinite, and a slab whose
    plane holds the origin gives inf * 0 = NaN. The slab test keeps the
    reference's std::min / std::max (raytrace_common.h:68-81), whose NaN
    handling decides whether such a ray, lying in a brick face, meets the
    brick; hit mask and depth must equal the oracle's."""
    sc = oracle_scene("ico128")
    rng = np.random.default_rng(11)
    n = 2048
    step = 2.0 / 128  # voxel planes of the 128^3 grid (brick planes every 4th)
    o = -1.0 + step * rng.integers(0, 129, (n, 3)).astype(np.float64)
    d = np.zeros((n, 3))
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    d[np.arange(n), axis] = sign
    o[np.arange(n), axis] = -1.5 * sign  # start outside, travel through the grid
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 1e9
    r_ico128.SetPreset(preset())
    hits = r_ico128.RayQuery_NearestHit(rays)
    g = np.array([(h.t, h.primId) for h in hits])
    ref = sc.cast_rays(preset(), rays, precision=32)
    gh, rh = g[:, 1] != MISS, ref[:, 1] != MISS
    print("hits", gh.sum(), "oracle", rh.sum(), "mask mismatches", (gh != rh).sum())
    assert gh.sum() > 200
    assert (gh != rh).sum() == 0
    both = gh & rh
    assert np.array_equal(g[both, 0].astype(np.float32), ref[both, 0].astype(np.float32))


@pytest.mark.parametrize("tfar", [1e9, 2.0])
def test_cast_rays_any_hit(r_ico128, tfar):
    """RayQuery_NearestHit with dirAndFar.w <= 0 (BVH2Common.cpp:3492-3494):
    any-hit rays up to |w| -- boxes beyond the current hit are culled -- as
    shadow rays use them; the hit / miss decision and the hit equal the
    oracle's on the same rays."""
    sc = oracle_scene("ico128")
    rng = np.random.default_rng(5)
    n = 4096
    o = rng.uniform(-1.5, 1.5, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, -tfar
    r_ico128.SetPreset(preset())
    hits = r_ico128.RayQuery_NearestHit(rays)
    g = np.array([(h.t, h.primId) for h in hits])
    ref = sc.cast_rays(preset(), rays, precision=32)
    gh, rh = g[:, 1] != MISS, ref[:, 1] != MISS
    assert gh.sum() > 500 and (gh != rh).sum() <= 2
    both = gh & rh
    assert np.mean(g[both, 0].astype(np.float32) == ref[both, 0].astype(np.float32)) > 0.999
    assert np.all(g[gh, 0] <= tfar)


def test_work_counters_match_oracle(r_ico128):
    """GPU per-ray work counters equal the FP32 oracle's (same traversal order;
    tile culling off, so every ray is traced as the reference traces it)."""
    sc = oracle_scene("ico128")
    wv, pr = L.orbit_camera(7, 0)
    r_ico128.SetPreset(preset())
    r_ico128.SetTileCulling(False)
    try:
        g = r_ico128.CountWork(wv[None], pr[None], 256, 256)
    finally:
        r_ico128.SetTileCulling(True)
    o = sc.render(wv, pr, preset(), 256, 256, precision=32, counters=True)["counters"]
    assert abs(int(g[0]) - o["node_pairs"]) <= 1e-4 * o["node_pairs"]
    assert abs(int(g[1]) - o["bricks"]) <= 1e-4 * o["bricks"]
    assert abs(int(g[2]) - o["voxels"]) <= 1e-4 * o["voxels"]


def test_brute_force_topology_independent():
    """On a small scene the BVH result equals the nearest hit over all bricks
    except where stop-at-first-SDF-hit returns a farther brick."""
    sbs = L.displaced_icosphere(3, 0.7, 0.05, 3, seed=4, lod=8, brick_size=4, brick_pad=0, bytes_per_value=2)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    r.SetPreset(preset())
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    wv, pr = L.orbit_camera(9, 0)
    rays = np.array([oracle.eye_ray(wv, pr, x, y, 0.5, 0.5, 64, 64) for y in range(64) for x in range(64)], np.float32)
    hits = r.RayQuery_NearestHit(rays)
    gt = np.array([h.t for h in hits])
    bt, bp = sc.brute_nearest(preset(), rays)
    hit = bp != MISS
    assert np.array_equal(np.array([h.primId for h in hits]) != MISS, hit)
    assert np.mean(np.abs(gt[hit] - bt[hit]) <= 1e-6 * bt[hit]) > 0.99


HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("mode", [L.SDF_OCTREE_NODE_INTERSECT_ST, L.SDF_OCTREE_NODE_INTERSECT_NEWTON])
def test_reference_fixture_gbuffer_parity(mode):
    """The reference's own SBS (saves/test_10_sbs.bin: 49,487 one-voxel bricks
    at LODs 16..128, 8-bit DX) rendered from the reference test's camera
    (tests.cpp render helper: lookAt from (0,0,3), fov 60, near 0.1, far 100)
    and from two orbit views, against the FP32 and FP64 oracle."""
    sbs = L.load_sbs(os.path.join(HERE, "golden", "test_10_sbs.bin"))
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    cams = [L.look_at_camera([0, 0, 3], [0, 0, 0], [0, 1, 0], fov=60.0, znear=0.1, zfar=100.0)]
    cams += [L.orbit_camera(21, i) for i in range(2)]
    for wv, pr in cams:
        res = _compare(r, sc, preset(mode), wv, pr, 256, 256, voxels_per_unit=128 * 1 / 2.0)
        print(mode, res)
        assert res["hits"] > 1000
        assert res["mask_mismatch_fp32"] <= 1e-4 * res["n_pix"]
        assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
        assert res["depth_rel_max"] <= 1e-5
        assert res["normal_bad"] == 0
        assert res["exact_t_fp32"] > 0.999
    r.close()


@pytest.mark.parametrize("W,H", [(1, 1), (7, 13), (129, 65), (1000, 3), (3, 700), (65535, 5), (5, 65535)])
def test_ragged_image_sizes(r_ico128, W, H):
    """Image sides that are not multiples of the 8x8 tile (partial tiles,
    single rows / columns, the largest side the API takes) against the FP32
    and FP64 oracle."""
    sc = oracle_scene("ico128")
    wv, pr = L.orbit_camera(31, 0, aspect=W / H)
    res = _compare(r_ico128, sc, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), wv, pr, W, H, voxels_per_unit=32 * 4 / 2.0)
    print(W, H, res)
    assert res["mask_mismatch_fp32"] == 0
    assert res["mask_mismatch_fp64"] <= max(1, 1e-4 * res["n_pix"]) and res["mismatch_all_grazing"]
    assert res["depth_rel_max"] <= 1e-5
    assert res["normal_bad"] == 0


def test_many_views_in_one_launch(r_ico128):
    """37 views of an odd-sized image in one launch (a wave finds its view and
    8x8 super-block from its packed index): every view equals the same view
    rendered alone, bit for bit."""
    W, H = 43, 29
    cams = [L.orbit_camera(17, i, aspect=W / H) for i in range(37)]
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r_ico128.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON))
    r_ico128.RenderViewsDevice(wvs, prs, W, H, 0)
    multi = [r_ico128.DownloadView(v, W, H) for v in range(37)]
    for v in (0, 1, 17, 36):
        t, n, prim = r_ico128.RenderGBuffer(W, H, wvs[v], prs[v])
        tm, nm, pm = multi[v]
        assert (prim != MISS).sum() > 20, v
        assert np.array_equal(prim, pm) and np.array_equal(t.view(np.uint32), tm.view(np.uint32)), v
        assert np.array_equal(n.view(np.uint32), nm.view(np.uint32)), v


@pytest.mark.parametrize("eye", [[0.95, 0.1, 0.05], [0.0, 0.0, 0.0], [0.3, -0.85, 0.2]])
def test_camera_inside_the_scene_box(r_ico128, eye):
    """Eye inside the [-1, 1]^3 scene box: near a box face among the bricks
    (the 0.1 SDF bias on tNear decides which bricks count), at the centre
    (every ray starts inside the solid), and just outside the surface. With
    the eye ~0.1 from the surface the reference arithmetic in FP32 (the
    oracle's twin) is itself up to ~1.5e-5 from FP64 in relative depth, so
    the depth bar is 1e-5 or the twin's own error, whichever is larger."""
    sc = oracle_scene("ico128")
    wv, pr = L.look_at_camera(eye, [0.2, 0.3, -0.4], [0, 1, 0], fov=90.0, znear=0.01, zfar=100.0)
    res = _compare(r_ico128, sc, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), wv, pr, 128, 128, voxels_per_unit=32 * 4 / 2.0)
    print(eye, res)
    assert res["mask_mismatch_fp32"] <= 2
    assert res["mask_mismatch_fp64"] <= 2 and res["mismatch_all_grazing"]
    assert res["exact_t_fp32"] > 0.999
    assert res["depth_rel_max"] <= max(1e-5, res["depth_rel_max_fp32_twin"])
    assert res["normal_bad"] == 0


def test_indexed_neighbour_layout_smoothed_normals():
    """ID32F_IRGB32F_IN scene (27 neighbour-brick indices per brick) with the
    geometric and the smoothed normal mode (BVH2Common.cpp:838-906: the
    normals of the 8 voxels around the hit, across brick borders through the
    neighbour table, blended by the smoothstep weights), against the FP32
    and FP64 oracle; the smoothed normals must differ from the geometric ones
    on many pixels, which shows the blend ran."""
    sbs = L.grid_sbs("torus", 0.6, 0.2, 0.02, 3, seed=3, brick_count=16, brick_size=4, with_neighbors=True)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    sc = oracle.Scene(sbs, pairs)
    normals = {}
    for nm in (L.NORMAL_MODE_GEOMETRY, L.NORMAL_MODE_SDF_SMOOTHED):
        p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON)
        p.normal_mode = nm
        for view in range(2):
            wv, pr = L.orbit_camera(41, view)
            res = _compare(r, sc, p, wv, pr, 160, 160, voxels_per_unit=16 * 4 / 2.0)
            print(nm, view, res)
            assert res["hits"] > 2000
            assert res["mask_mismatch_fp32"] <= 1e-4 * res["n_pix"]
            assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
            assert res["depth_rel_max"] <= 1e-5
            assert res["normal_bad"] == 0
            normals[(nm, view)] = r.RenderGBuffer(160, 160, wv, pr)
    for view in range(2):
        tg, ng, pg = normals[(L.NORMAL_MODE_GEOMETRY, view)]
        ts, ns, ps = normals[(L.NORMAL_MODE_SDF_SMOOTHED, view)]
        assert np.array_equal(tg, ts) and np.array_equal(pg, ps)  # same hits
        hit = pg != MISS
        assert np.mean(angle(ng[hit], ns[hit]) > 1e-3) > 0.2
    r.close()


def test_single_brick_scene():
    """A one-brick scene (its BVH is one pair whose second slot repeats the brick)."""
    sbs = L.grid_sbs("sphere", 0.5, 0.0, 0.0, 1, seed=1, brick_count=1, brick_size=4)
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(sbs)
    pairs, _ = r.GetBVH()
    assert pairs.shape[0] == 1
    sc = oracle.Scene(sbs, pairs)
    wv, pr = L.orbit_camera(2, 0)
    res = _compare(r, sc, preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON), wv, pr, 64, 64, voxels_per_unit=1 * 4 / 2.0)
    print(res)
    assert res["hits"] > 100 and res["mask_mismatch_fp32"] == 0 and res["exact_t_fp32"] > 0.999
    r.close()


def test_bad_inputs_raise():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    with pytest.raises(L.LrtError):
        r.RenderFloat(16, 16)  # no scene
    sbs = ico128_scene()
    bad = L.SdfSBS(sbs.header, sbs.nodes.copy(), sbs.values)
    bad.nodes[0, 2] = sbs.values.size  # data_offset out of range
    with pytest.raises(L.LrtError):
        r.SetScene(bad)
    for bs, pad in ((17, 0), (16, 2), (0, 0)):  # brick sizes outside 1..16, padding outside 0..1
        with pytest.raises(L.LrtError):
            r.SetScene(L.SdfSBS(L.SdfSBSHeader(bs, pad, sbs.header.bytes_per_value, sbs.header.aux_data), sbs.nodes, sbs.values))
    r.SetScene(sbs)
    # sizes checked before any work; outputs NULL so no host buffers are needed
    for w, h in ((65535, 65535), (0, 16), (65536, 1)):  # 65535^2 rounds up to 65536^2 packed pixels, past 32 bits
        assert L.lib.lrt_render_gbuffer(r._h, None, None, None, w, h) != 0, (w, h)


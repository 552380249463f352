"""GPU checks of the learned root-find seeder through the C ABI.

- Training events: features equal the oracle's restatement, targets equal the
  plain render's hits, one event per brick the traversal enters.
- FP32 training kernels: loss and gradient within 1e-4 relative (L2) of the
  FP64 oracle; one Adam step equals its closed form.
- BF16 MFMA inference: agrees with the oracle's BF16-operand emulation.
- Seeded rendering: the same hit-mask / depth / normal tolerances against the
  FP64 oracle as the plain path, with adversarial (always-confident, random)
  predictions and with a trained network, in the sphere-tracing and Newton
  modes; and the trained network cuts root-finder iterations.
"""
import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import MISS, ico128_scene, compare_gbuffer, oracle_scene, preset

pytestmark = pytest.mark.gpu

ST, NEWTON = L.SDF_OCTREE_NODE_INTERSECT_ST, L.SDF_OCTREE_NODE_INTERSECT_NEWTON


@pytest.fixture(scope="module")
def r_c1():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(ico128_scene())
    r.SetPreset(preset(NEWTON))
    yield r
    r.close()


def _collect(r, view_seed, W, max_events=2_000_000):
    r.MlpDatasetClear()
    wv, pr = L.orbit_camera(view_seed, 0)
    kept, seen = r.MlpCollect(wv[None], pr[None], W, W, max_events)
    f, t, m = r.MlpDataset()
    return wv, pr, f, t, m, kept, seen


def test_collected_features_match_oracle(r_c1):
    wv, pr, f, t, m, kept, seen = _collect(r_c1, 21, 64)
    assert kept == seen == f.shape[0] > 500
    sc = oracle_scene("ico128")
    rng = np.random.default_rng(0)
    pick = rng.choice(f.shape[0], 300, replace=False)
    rays = np.zeros((len(pick), 6), np.float32)
    for k, e in enumerate(pick):
        ray = m[e, 0]
        x, y = ray % 64, ray // 64
        rr = oracle.eye_ray(wv, pr, int(x), int(y), 0.5, 0.5, 64, 64, 32)
        rays[k, :3], rays[k, 3:] = rr[:3], rr[4:7]
    of, _ = sc.mlp_features(rays, m[pick, 1])
    err = np.abs(of - f[pick])
    print("feature max err", err.max())
    assert err.max() <= 1e-5 * (1 + np.abs(of).max())


def test_collected_targets_match_render(r_c1):
    wv, pr, f, t, m, kept, seen = _collect(r_c1, 22, 96)
    r_c1.SetViewport(0, 0, 96, 96)
    r_c1.UpdateCamera(wv, pr)
    tg, _, prim = r_c1.RenderGBuffer(96, 96)
    counts = r_c1.CountWorkEx(wv[None], pr[None], 96, 96)
    assert seen == counts[1]  # one event per brick entry
    sc = oracle_scene("ico128")
    hit_rays = set()
    last_event = {}
    for e in range(f.shape[0]):
        last_event[int(m[e, 0])] = e
    for e in np.nonzero(t[:, 1] > 0.5)[0]:
        ray = int(m[e, 0])
        assert ray not in hit_rays
        hit_rays.add(ray)
        assert last_event[ray] == e  # the hit ends the ray
    hit_px = {int(i) for i in np.nonzero(prim.ravel() != MISS)[0]}
    assert hit_rays == hit_px
    # t_hit (brick units from the brick entry) maps back to the rendered depth
    sel = np.nonzero(t[:, 1] > 0.5)[0][:200]
    rays = np.zeros((len(sel), 6), np.float32)
    for k, e in enumerate(sel):
        ray = int(m[e, 0])
        rr = oracle.eye_ray(wv, pr, ray % 96, ray // 96, 0.5, 0.5, 96, 96, 32)
        rays[k, :3], rays[k, 3:] = rr[:3], rr[4:7]
    _, t0 = sc.mlp_features(rays, m[sel, 1])
    s = 0.5 * (ico128_scene().nodes[m[sel, 1], 1] & 0xFFFF).astype(np.float64)
    depth = t0 + t[sel, 0] / s
    ref = tg.ravel()[m[sel, 0]]
    assert np.max(np.abs(depth - ref) / ref) <= 1e-6


def test_loss_and_gradient_match_oracle(r_c1):
    _, _, f, t, _, _, _ = _collect(r_c1, 23, 64)
    n = min(2000, f.shape[0])
    r_c1.MlpInit(3)
    params = r_c1.MlpGetParams()
    loss, g = r_c1.MlpLossGrad(f[:n], t[:n])
    ol, og = oracle.mlp_loss_grad(params, f[:n], t[:n])
    rel = np.linalg.norm(g - og) / np.linalg.norm(og)
    print("loss", loss, ol, "grad rel", rel)
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert rel <= 1e-4


def test_adam_step_closed_form(r_c1):
    _, _, f, t, _, _, _ = _collect(r_c1, 24, 32)
    n = f.shape[0]
    assert 64 < n <= 4096
    r_c1.MlpInit(4)
    p0 = r_c1.MlpGetParams()
    _, g = r_c1.MlpLossGrad(f, t)
    lr = 1e-3
    batch = 64 * ((n + 63) // 64)
    r_c1.MlpTrain(1, batch=batch, lr=lr, seed=1)  # one full-batch step
    p1 = r_c1.MlpGetParams()
    # first Adam step: m_hat = g, v_hat = g^2 -> lr * g / (|g| + eps)
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    stable = np.abs(g) > 1e-5
    assert np.max(np.abs(p1 - expect)[stable]) <= 1e-3 * lr
    assert np.max(np.abs(p1 - p0)) <= lr * 1.0001


def test_mfma_inference_matches_bf16_oracle(r_c1):
    _, _, f, _, _, _, _ = _collect(r_c1, 25, 64)
    r_c1.MlpInit(5)
    params = r_c1.MlpGetParams()
    out, _ = r_c1.MlpInfer(f)
    ref = oracle.mlp_forward(params, f, bf16=True)
    err = np.abs(out - ref)
    scale = 1 + np.abs(ref)
    print("infer p99", np.quantile(err / scale, 0.99), "max", (err / scale).max())
    assert np.quantile(err / scale, 0.99) <= 1e-2
    assert (err / scale).max() <= 5e-2
    # a wrong operand layout would not even correlate
    assert np.corrcoef(out[:, 2], ref[:, 2])[0, 1] > 0.999
    # ragged batch sizes (partial 16-sample tiles and a partial wave)
    for n in (1, 15, 17, 63, 65, 1000):
        o2, _ = r_c1.MlpInfer(f[:n])
        assert np.array_equal(o2, out[:n])


def _adversarial_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, L.MLP_PARAMS).astype(np.float32)
    p[-3:] = [0.3, 0.6, 50.0]  # b4: always "confident", seeds anywhere in the brick
    return p


@pytest.mark.parametrize("mode", [ST, NEWTON])
def test_seeded_render_exact_with_adversarial_seeds(r_c1, mode):
    sc = oracle_scene("ico128")
    r_c1.MlpSetParams(_adversarial_params(7))
    r_c1.SetSeeding(True, 0.9)
    try:
        for view in range(2):
            wv, pr = L.orbit_camera(7, view)
            res = compare_gbuffer(r_c1, sc, preset(mode), wv, pr, 256, 256, voxels_per_unit=32 * 4 / 2.0)
            cnt = r_c1.CountWorkEx(wv[None], pr[None], 256, 256)
            print(mode, view, res, "seeds", cnt[5], "bricks", cnt[1])
            assert cnt[5] == cnt[1] > 0  # every brick entry was seeded
            assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
            assert res["depth_rel_max"] <= 1e-5
            assert res["normal_bad"] == 0
    finally:
        r_c1.SetSeeding(False)


def _train_on_other_scenes(r, n_scenes=4, W=192, epochs=12):
    r.MlpDatasetClear()
    for k in range(n_scenes):
        s = L.displaced_icosphere(5, 0.8, 0.05, 3, seed=100 + k, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)
        r.SetScene(s)
        cams = [L.orbit_camera(300 + k, i) for i in range(4)]
        r.MlpCollect(np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams]), W, W, 4_000_000)
    r.SetScene(ico128_scene())
    r.MlpInit(11)
    losses, ms = r.MlpTrain(epochs, batch=4096, lr=2e-3, seed=3)
    return losses, ms


def test_trained_seeder_keeps_tolerances_and_saves_iterations(r_c1):
    losses, ms = _train_on_other_scenes(r_c1)
    n = r_c1.MlpDatasetSize()
    print("train events", n, "losses", losses, "ms", ms)
    assert losses[-1] < 0.7 * losses[0]
    f, t, _ = r_c1.MlpDataset(0, min(n, 20000))
    out, _ = r_c1.MlpInfer(f)
    acc = np.mean((out[:, 2] > 0) == (t[:, 1] > 0.5))
    print("hit accuracy", acc)
    assert acc > 0.85
    sc = oracle_scene("ico128")
    for mode in (ST, NEWTON):
        wv, pr = L.orbit_camera(7, 2)
        r_c1.SetPreset(preset(mode))
        plain = r_c1.CountWorkEx(wv[None], pr[None], 256, 256)
        r_c1.SetSeeding(True, 0.9)
        try:
            seeded = r_c1.CountWorkEx(wv[None], pr[None], 256, 256)
            res = compare_gbuffer(r_c1, sc, preset(mode), wv, pr, 256, 256, voxels_per_unit=32 * 4 / 2.0)
        finally:
            r_c1.SetSeeding(False)
        print(mode, "plain", plain, "seeded", seeded, res)
        assert seeded[0] == plain[0] and seeded[1] == plain[1]  # same traversal
        assert seeded[6] == plain[6] > 0  # the same tiles culled
        assert res["mask_mismatch_fp64"] <= 1e-4 * res["n_pix"] and res["mismatch_all_grazing"]
        assert res["depth_rel_max"] <= 1e-5
        assert res["normal_bad"] == 0
        # sphere tracing does fewer root-finder iterations: it skips voxels it
        # provably cannot stop in and starts near the surface. Newton already
        # converges in about one step from the bracket midpoint, so a seed moves
        # its count by a few iterations either way (parity within 1 %).
        if mode == ST:
            assert seeded[3] < 0.9 * plain[3]
        else:
            assert seeded[3] <= 1.01 * plain[3]
    r_c1.SetPreset(preset(NEWTON))


@pytest.mark.parametrize("rmode", [None, L.MULTI_RENDER_MODE_PHONG_NO_TEX])
def test_seeded_tile_culling_leaves_images_identical(r_c1, rmode):
    """The seeded render skips the tiles the cull pass marks and writes them
    as misses: G-buffers (t, normal, primitive) of a two-view launch, or a
    shaded image with shadow rays, bit-identical with the cull on and off,
    with confident seeds everywhere (adversarial network) and most tiles culled."""
    r_c1.MlpSetParams(_adversarial_params(5))
    cams = [L.orbit_camera(17, i) for i in range(2)]
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    W, H = 320, 192
    p = preset(NEWTON) if rmode is None else preset(NEWTON, render_mode=rmode)
    r_c1.SetPreset(p)
    r_c1.SetSeeding(True, 0.9)
    out = []
    try:
        for on in (True, False):
            r_c1.SetTileCulling(on)
            if rmode is None:
                r_c1.RenderViewsDevice(wvs, prs, W, H, 0)
                out.append([a for v in range(2) for a in r_c1.DownloadView(v, W, H)])
            else:
                out.append([r_c1.RenderFloat(W, H, *c, p) for c in cams])
        r_c1.SetTileCulling(True)
        cnt = r_c1.CountWorkEx(wvs, prs, W, H)
    finally:
        r_c1.SetSeeding(False)
        r_c1.SetTileCulling(True)
        r_c1.SetPreset(preset(NEWTON))
    for a, b in zip(*out):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    tiles = 2 * (W // 8) * (H // 8)
    print("seeded culled", cnt[6], "of", tiles, "seeds", cnt[5])
    assert cnt[6] > 0.3 * tiles and cnt[5] > 0


def test_seeding_rejects_bad_arguments(r_c1):
    with pytest.raises(L.LrtError):
        r_c1.SetSeeding(True, 1.5)
    with pytest.raises(L.LrtError):
        r_c1.MlpSetParams(np.zeros(10, np.float32))
    with pytest.raises(L.LrtError):
        r_c1.MlpTrain(1, batch=100)

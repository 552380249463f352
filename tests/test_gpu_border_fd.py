"""The HIP path's DR border (silhouette) term against finite differences.

The same port of the reference's test_position_derivatives
(diff_render/tests_diff_render.cpp:359-562) as tests/test_cpu_border_fd.py,
run through RenderDR on the GPU: the default mode (k_dr_primary +
k_dr_border_rays, spp 64 so each pixel's depth histogram is streamed) and the
finite-difference mode (k_dr_fd_*) each over 25 seeds, with the reference's
criteria and the 1/|grad f| factor of the relaxed-boundary estimator computed
from the lattice alone. One seed of each mode is also compared with the
oracle's FP32 twin.
"""
import numpy as np
import pytest

import litert_amd.api as L
import oracle
from helpers import (POS_DER_H, POS_DER_SEEDS, POS_DER_W, pos_der_camera, pos_der_criteria, pos_der_distance_count,
                     pos_der_preset, pos_der_scene, pos_der_stats, silhouette_inverse_gradient_norm)

pytestmark = pytest.mark.gpu

_cache = {}


@pytest.fixture(scope="module")
def renderer():
    sbs = pos_der_scene()
    g = L.MultiRendererDR()
    g.SetScene(sbs)
    wv, pr = pos_der_camera()
    g.SetReference(np.zeros((1, POS_DER_H, POS_DER_W, 4), np.float32), wv[None], pr[None])
    yield g
    g.close()


def _gpu_grads(g, mode, eps, diff_mode, seeds=POS_DER_SEEDS):
    key = (mode, eps, diff_mode, seeds)
    if key not in _cache:
        sbs = pos_der_scene()
        n, nd = sbs.values_f.size, pos_der_distance_count(sbs)
        rows = []
        for i in range(seeds):
            _, grad, _, _ = g.RenderDR(pos_der_preset(mode, eps, diff_mode), 0, seed=1000 + i, n_params=n)
            rows.append(grad[:nd].astype(np.float64))
        _cache[key] = np.array(rows)
    return _cache[key]


@pytest.fixture(scope="module")
def inv_grad_norm():
    return silhouette_inverse_gradient_norm(pos_der_scene(), (0.3, 0.0, 2.5))


@pytest.mark.parametrize("diff_mode", [0, 1], ids=["default", "finite_diff"])
def test_gpu_position_derivative_seed_matches_oracle_twin(renderer, diff_mode):
    """Seed 1000, MASK, relax_eps 2e-4: image and dLoss/dS against the
    oracle's FP32 twin (float atomics reorder the sums: 1e-5 relative)."""
    sbs = pos_der_scene()
    n, nd = sbs.values_f.size, pos_der_distance_count(sbs)
    sc = oracle.Scene(sbs, renderer.GetBVH()[0])
    wv, pr = pos_der_camera()
    ref = np.zeros((POS_DER_H, POS_DER_W, 4), np.float32)
    dr = pos_der_preset(2, 2e-4, diff_mode)
    img, grad, loss, _ = renderer.RenderDR(dr, 0, seed=1000, n_params=n)
    o_img, o_g, st = sc.dr_render_grad(dr, wv, pr, POS_DER_W, POS_DER_H, ref, n, seed=1000, precision=32)
    bp, bh = renderer.LastStats()
    print("border", bp, bh, st)
    assert np.mean(np.all(img == o_img.astype(np.float32), axis=-1)) > 0.999
    if diff_mode == 0:
        assert bp == st["border_pixels"] and abs(bh - st["border_rays_hit"]) <= 2
    assert np.linalg.norm(grad[:nd] - o_g[:nd]) <= 1e-5 * np.linalg.norm(o_g[:nd])


def test_gpu_position_derivatives_mask_wide_band(renderer):
    """Case 4.1 (MASK, relax_eps 1e-2) with the reference's criteria."""
    m0, c0 = pos_der_stats(_gpu_grads(renderer, 2, 1e-2, 0))
    m1, c1 = pos_der_stats(_gpu_grads(renderer, 2, 1e-2, 1))
    rel, avg, bias = pos_der_criteria(m0, c0, m1, c1)
    print(f"rel {rel:.4f} avg {avg:.3f} bias {bias:.4f}")
    assert np.abs(m1).sum() > 1.0
    assert rel <= 0.15 or avg <= 1.0
    assert abs(bias) <= 0.15


@pytest.mark.parametrize("mode", [2, 3], ids=["mask", "linear_depth"])
def test_gpu_position_derivatives_narrow_band(renderer, mode, inv_grad_norm):
    """Cases 4.3 / 4.13 (relax_eps 2e-4): ratio of summed gradients equals the
    1/|grad f| factor within 6 %, and the corrected default mode meets the
    reference's criteria with |bias| <= 0.06."""
    m0, c0 = pos_der_stats(_gpu_grads(renderer, mode, 2e-4, 0))
    m1, c1 = pos_der_stats(_gpu_grads(renderer, mode, 2e-4, 1))
    ratio = m0.sum() / m1.sum()
    rel, avg, bias = pos_der_criteria(m0 / inv_grad_norm, c0 / inv_grad_norm, m1, c1)
    print(f"factor {inv_grad_norm:.4f} ratio {ratio:.4f} corrected rel {rel:.4f} bias {bias:.4f}")
    assert abs(ratio / inv_grad_norm - 1.0) <= 0.06
    assert rel <= 0.15
    assert abs(bias) <= 0.06


# ---- further cases of the reference's test 4 (tests_diff_render.cpp:660-716) --
def _close_camera():
    """The reference's close view (:373-379): from (0.55, 0.55, 2) at (0.55, 0.55, 0), 15 degrees."""
    return L.look_at_camera((0.55, 0.55, 2.0), (0.55, 0.55, 0.0), (0.0, 1.0, 0.0), fov=15.0, aspect=1.0, znear=0.01, zfar=100.0)


def _grads(camera, mode, eps, diff_mode, sampling=0, border_spp=None, seeds=POS_DER_SEEDS):
    key = ("cam", camera, mode, eps, diff_mode, sampling, border_spp, seeds)
    if key not in _cache:
        sbs = pos_der_scene()
        n, nd = sbs.values_f.size, pos_der_distance_count(sbs)
        wv, pr = pos_der_camera() if camera == "far" else _close_camera()
        g = L.MultiRendererDR()
        g.SetScene(sbs)
        g.SetReference(np.zeros((1, POS_DER_H, POS_DER_W, 4), np.float32), wv[None], pr[None])
        rows = []
        for i in range(seeds):
            p = pos_der_preset(mode, eps, diff_mode)
            p.dr_border_sampling = sampling
            if border_spp is not None:
                p.border_spp = border_spp
            _, grad, _, _ = g.RenderDR(p, 0, seed=1000 + i, n_params=n)
            rows.append(grad[:nd].astype(np.float64))
        g.close()
        _cache[key] = np.array(rows)
    return _cache[key]


@pytest.mark.parametrize("case", [("4.2", "far", 2, 1, 1e-2), ("4.4", "far", 2, 1, 2e-4), ("4.12", "close", 3, 0, 2e-4)],
                         ids=["4.2-mask-svm-wide", "4.4-mask-svm-narrow", "4.12-depth-close"])
def test_gpu_position_derivatives_reference_cases(case):
    """Cases 4.2 and 4.4 (MASK with SVM border sampling, relax_eps 1e-2 and
    2e-4) and 4.12 (LINEAR_DEPTH, the close 15-degree view) with the
    reference's own criteria."""
    name, cam, mode, sampling, eps = case
    m0, c0 = pos_der_stats(_grads(cam, mode, eps, 0, sampling))
    m1, c1 = pos_der_stats(_grads(cam, mode, eps, 1))
    rel, avg, bias = pos_der_criteria(m0, c0, m1, c1)
    print(f"{name}: rel {rel:.4f} avg {avg:.3f} bias {bias:.4f}")
    assert np.abs(m1).sum() > 1.0
    assert rel <= 0.15 or avg <= 1.0
    assert abs(bias) <= 0.15


@pytest.mark.parametrize("mode", [0, 1], ids=["4.7-diffuse", "4.9-lambert"])
def test_gpu_position_derivatives_colour_modes(mode, inv_grad_norm):
    """Cases 4.7 (DIFFUSE) and 4.9 (LAMBERT, default lights), relax_eps 2e-4,
    empty reference. The gradient splits into the interior part (the same
    RenderDR with border_spp 0) and the border part (the rest); only the
    border part carries the relaxed-boundary estimator's 1/|grad f| factor.
    With the border part divided by the factor computed from the lattice, the
    default mode meets the reference's criteria with |bias| <= 0.06; the raw
    relative bias sits at the reference's 0.15 edge (0.14) or past it (0.19)."""
    default = _grads("far", mode, 2e-4, 0)
    interior = _grads("far", mode, 2e-4, 0, border_spp=0)
    corrected = interior + (default - interior) / inv_grad_norm
    m0, c0 = pos_der_stats(corrected)
    m1, c1 = pos_der_stats(_grads("far", mode, 2e-4, 1))
    raw = pos_der_criteria(*pos_der_stats(default), m1, c1)
    rel, avg, bias = pos_der_criteria(m0, c0, m1, c1)
    border_share = np.abs((default - interior).mean(0)).sum() / np.abs(default.mean(0)).sum()
    print(f"raw rel/avg/bias {raw}; border share {border_share:.3f}; corrected rel {rel:.4f} bias {bias:.4f}")
    assert border_share > 0.5
    assert rel <= 0.15
    assert abs(bias) <= 0.06

"""CPU pins of the seeder oracle: the FP64 network forward/backward against
PyTorch (an independent implementation of linear + SiLU layers and of the
logistic loss) and against central differences of its own loss, and the
feature restatement against direct geometric recomputation."""
import numpy as np
import pytest
import torch

import litert_amd.api as L
import oracle

P = L.MLP_PARAMS


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    off = 0
    for rows, cols in ((64, 24), (64, 64), (64, 64), (3, 64)):
        lim = np.sqrt(6.0 / (rows + cols)) * scale
        p[off:off + rows * cols] = rng.uniform(-lim, lim, rows * cols)
        off += rows * cols
        p[off:off + rows] = rng.normal(0, 0.1, rows)
        off += rows
    assert off == P
    return p


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 24)).astype(np.float32)
    t = np.zeros((n, 3), np.float32)
    t[:, 0] = rng.uniform(0, 1.5, n)
    t[:, 1] = rng.random(n) < 0.5
    t[:, 2] = 0.125
    return x, t


def _torch_net(params, bf16):
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    sizes = ((64, 24), (64, 64), (64, 64), (3, 64))

    def q(v):
        return v.to(torch.float32).to(torch.bfloat16).to(torch.float64) if bf16 else v

    def fwd(x):
        a = q(torch.tensor(x, dtype=torch.float64))
        off = 0
        for k, (rows, cols) in enumerate(sizes):
            W = p[off:off + rows * cols].reshape(rows, cols)
            off += rows * cols
            b = p[off:off + rows]
            off += rows
            z = a @ q(W).T + b if not bf16 else a @ (W + (q(W) - W).detach()).T + b
            a = q(torch.nn.functional.silu(z)) if k < 3 else z
        return a
    return p, fwd


def test_forward_matches_torch():
    params = _params(1)
    x, _ = _data(2, 200)
    _, fwd = _torch_net(params, False)
    ref = fwd(x).detach().numpy()
    got = oracle.mlp_forward(params, x, bf16=False)
    assert np.max(np.abs(got - ref)) <= 1e-12 * (1 + np.max(np.abs(ref)))


def test_forward_bf16_emulation_matches_torch_rounding():
    params = _params(3)
    x, _ = _data(4, 500)
    _, fwd = _torch_net(params, True)
    ref = fwd(x).detach().numpy()
    got = oracle.mlp_forward(params, x, bf16=True)
    err = np.abs(got - ref)
    # both round operands to bf16 with round-to-nearest-even; SiLU is x*sigmoid(x)
    # in torch and x/(1+exp(-x)) in the oracle, which can flip a rounding tie
    assert np.quantile(err, 0.99) <= 1e-9
    assert err.max() <= 5e-2
    # and emulation really changes the numbers (operands are rounded)
    assert np.max(np.abs(got - oracle.mlp_forward(params, x, bf16=False))) > 1e-4


def test_loss_and_gradient_match_torch_autograd():
    params = _params(5)
    x, t = _data(6, 300)
    p, fwd = _torch_net(params, False)
    z = fwd(x)
    tt = torch.tensor(t, dtype=torch.float64)
    h, th, m = tt[:, 1], tt[:, 0], tt[:, 2]
    bce = torch.nn.functional.binary_cross_entropy_with_logits(z[:, 2], h, reduction="none")
    loss = (bce + h * ((z[:, 0] - (th - m)) ** 2 + (z[:, 1] - (th + m)) ** 2)).mean()
    loss.backward()
    ol, og = oracle.mlp_loss_grad(params, x, t)
    assert abs(ol - loss.item()) <= 1e-12 * abs(loss.item())
    g = p.grad.numpy()
    assert np.linalg.norm(og - g) <= 1e-10 * np.linalg.norm(g)


def test_gradient_matches_central_differences():
    params = _params(7).astype(np.float64)
    x, t = _data(8, 64)
    _, g = oracle.mlp_loss_grad(params.astype(np.float32), x, t)
    rng = np.random.default_rng(9)
    idx = rng.choice(P, 40, replace=False)
    # the oracle reads float32 parameters: step by a float32-representable h
    for k in idx:
        h = 1e-3
        pp = params.astype(np.float32).copy()
        base = pp[k]
        pp[k] = np.float32(base + h)
        lp, _ = oracle.mlp_loss_grad(pp, x, t)
        pp[k] = np.float32(base - h)
        lm, _ = oracle.mlp_loss_grad(pp, x, t)
        hh = float(np.float32(base + h)) - float(np.float32(base - h))
        fd = (lp - lm) / hh
        assert abs(fd - g[k]) <= 1e-4 * (1 + abs(g[k])), (k, fd, g[k])


@pytest.fixture(scope="module")
def small_scene():
    sbs = L.displaced_icosphere(3, 0.7, 0.05, 2, seed=5, lod=16, brick_size=4, brick_pad=1, bytes_per_value=4)
    pairs, _ = L.build_bvh(sbs)
    return sbs, oracle.Scene(sbs, pairs)


def _box(sbs, b):
    px, py = sbs.nodes[b, 0] >> 16, sbs.nodes[b, 0] & 0xFFFF
    pz, sz = sbs.nodes[b, 1] >> 16, sbs.nodes[b, 1] & 0xFFFF
    bmin = -1 + 2.0 / sz * np.array([px, py, pz], np.float64)
    return bmin, bmin + 2.0 / sz, 0.5 * sz


def test_features_match_geometry(small_scene):
    sbs, sc = small_scene
    rng = np.random.default_rng(11)
    n = 64
    bricks = rng.integers(0, sbs.nodes.shape[0], n)
    rays = np.zeros((n, 6), np.float32)
    for i, b in enumerate(bricks):
        lo, hi, _ = _box(sbs, b)
        target = lo + (hi - lo) * rng.uniform(0.2, 0.8, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        rays[i, :3] = target - 3.0 * d
        rays[i, 3:] = d
    f, t0 = sc.mlp_features(rays, bricks)
    for i, b in enumerate(bricks):
        lo, hi, s = _box(sbs, b)
        o, d = rays[i, :3].astype(np.float64), rays[i, 3:].astype(np.float64)
        with np.errstate(divide="ignore"):
            ta, tb = (lo - o) / d, (hi - o) / d
        tn, tf = np.max(np.minimum(ta, tb)), np.min(np.maximum(ta, tb))
        assert abs(t0[i] - tn) <= 1e-5 * (1 + abs(tn))
        assert np.allclose(f[i, :3], d, atol=1e-7)
        assert np.allclose(f[i, 3:6], (o + tn * d - lo) * s, atol=1e-4)
        assert np.all(f[i, 3:6] >= -1e-4) and np.all(f[i, 3:6] <= 1 + 1e-4)
        assert abs(f[i, 6] - (tf - tn) * s) <= 1e-4
        c = 0.5 * (lo + hi)
        assert abs(f[i, 7] - (tn - np.dot(c - o, d)) * s) <= 1e-4
        # corners and octant minima straight from the voxel loads
        bs = sbs.header.brick_size
        vox = {(x, y, z): sc.load_voxel(int(b), x, y, z) for x in range(bs) for y in range(bs) for z in range(bs)}
        for cidx in range(8):
            cx, cy, cz = (cidx >> 2) & 1, (cidx >> 1) & 1, cidx & 1
            vals, _ = vox[(cx * (bs - 1), cy * (bs - 1), cz * (bs - 1))]
            assert abs(f[i, 8 + cidx] - vals[cidx] * s) <= 1e-5 * (1 + abs(vals[cidx] * s))
        for o_ in range(8):
            hx, hy, hz = (o_ >> 2) & 1, (o_ >> 1) & 1, o_ & 1
            m = min(vmin for (x, y, z), (_, vmin) in vox.items()
                    if (2 * x >= bs) == hx and (2 * y >= bs) == hy and (2 * z >= bs) == hz)
            assert abs(f[i, 16 + o_] - m * s) <= 1e-5 * (1 + abs(m * s))

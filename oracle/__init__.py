"""CPU oracle: restatement of the reference LiteRT SBS path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package. It is the checker of the HIP library and the CPU
baseline; the product library never calls it.

The reference itself cannot be compiled here (its HydraCore3 submodule --
LiteMath, CrossRT -- is empty and the Embree runtime is missing), so there is
no oracle/_ref build; see lrt_oracle.h for the restated functions.

Parity unpinned (only an import of the reference's third-party code could
pin these; details in lrt_oracle.h): LiteMath's camera/matrix conventions
(the GL convention used here is pinned against the closed-form pinhole),
Embree's BVH topology, LAMBERT's light direction to the last bit, and
whether LiteImage's bilinear sampler wraps or clamps at the image border
(resample_reference).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "liblrt_oracle.so")
BASELINE_LIB_PATH = os.path.join(BUILD_DIR, "liblrt_oracle_baseline.so")
SOURCES = ["oracle_api.cpp", "oracle_dr.cpp", "oracle_mlp.cpp", "oracle_octree.cpp"]
# -ffp-contract=off matches the reference's ISO C++17 build (no FP contraction).
FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
         "-Wall", "-Wno-unused-function"]
# The CPU-baseline build that bench.py times: optimised for speed with the
# AVX2 / FMA instruction set every x86-64 host of this pipeline has, still
# without FP contraction or fast-math, so it computes the same FP32 results.
BASELINE_FLAGS = ["-O3", "-march=x86-64-v3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
                  "-Wall", "-Wno-unused-function"]


def _compile(flags, out):
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [os.path.join(_HERE, s) for s in SOURCES if os.path.exists(os.path.join(_HERE, s))]
    tmp = out + ".tmp"
    r = subprocess.run(["g++", *flags, *srcs, "-o", tmp], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


def build():
    """Builds the checker (liblrt_oracle.so) and the CPU-baseline build; returns the checker's path."""
    _compile(BASELINE_FLAGS, BASELINE_LIB_PATH)
    return _compile(FLAGS, LIB_PATH)


_libs = {}


def _load(path):
    L = C.CDLL(path)
    P = C.c_void_p
    u32, u64, i32, f32, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_float, C.c_double
    sig = {
        "orc_scene_create": (P, [P, P, u32, P, u64, P, u64, P, u32]),
        "orc_scene_destroy": (None, [P]),
        "orc_scene_set_values_f": (i32, [P, P, u64]),
        "orc_scene_set_lights": (i32, [P, P, u32]),
        "orc_render": (i32, [P, P, P, P, u32, u32, u32, i32, i32, i32, P, P, P, P, P]),
        "orc_cast_rays": (i32, [P, P, P, u32, i32, i32, i32, P]),
        "orc_brute_nearest": (i32, [P, P, P, u32, i32, i32, P, P]),
        "orc_eval_trilinear": (f64, [P, P, i32]),
        "orc_eval_trilinear_diff": (None, [P, P, i32, P]),
        "orc_voxel_root": (i32, [u32, P, P, P, f64, f64, i32, P, P]),
        "orc_encode_normal": (None, [P, i32, P]),
        "orc_decode_normal": (None, [P, i32, P]),
        "orc_pcg3d": (None, [P]),
        "orc_rand2": (None, [u32, u32, u32, u32, P]),
        "orc_pack_xy": (None, [u32, u32, P]),
        "orc_safe_inverse": (None, [P, i32, P]),
        "orc_ray_box": (None, [P, P, P, P, i32, P]),
        "orc_load_voxel": (f64, [P, u32, u32, u32, u32, i32, P]),
        "orc_eye_ray": (None, [P, P, u32, u32, f32, f32, u32, u32, i32, P]),
    }
    optional = {
        "orc_dr_render_grad": (i32, [P, P, P, P, u32, u32, u32, P, i32, i32, i32, P, P, P]),
        "orc_adam_step": (None, [u32, P, P, P, P, u32, f32, f32, f32, f32]),
        "orc_dr_regularize": (i32, [P, u32, f32, f32, P]),
        "orc_dr_redistance": (i32, [P, u32, f32, u32]),
        "orc_mlp_forward": (None, [P, P, u32, i32, P]),
        "orc_mlp_loss_grad": (f64, [P, P, P, u32, i32, P]),
        "orc_mlp_brick_features": (i32, [P, u32, P]),
        "orc_mlp_features": (i32, [P, P, u32, P, P]),
        "orc_sbs_octree": (i32, [i32, P, u32, u32, u32, u32, u32, P, P, P, P]),
    }
    for name, (res, args) in list(sig.items()) + list(optional.items()):
        if name in optional and not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def lib(baseline=False):
    """The checker build, or (baseline=True) the CPU-baseline build bench.py times."""
    path = BASELINE_LIB_PATH if baseline else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _libs[path] = _load(path)
    return _libs[path]


class Counters(C.Structure):
    _fields_ = [("node_pairs", C.c_uint64), ("bricks", C.c_uint64), ("voxels", C.c_uint64),
                ("root_finds", C.c_uint64), ("rays", C.c_uint64), ("root_iters", C.c_uint64)]


class DRStats(C.Structure):
    _fields_ = [("loss", C.c_double), ("border_pixels", C.c_uint32), ("border_rays_hit", C.c_uint32),
                ("rays", C.c_uint64)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Scene:
    """Oracle copy of an SBS scene plus the fat BVH2 pairs (uint32 (n,16))."""

    def __init__(self, sbs, pairs, baseline=False):
        self._L = lib(baseline)
        self._keep = (np.ascontiguousarray(sbs.nodes, np.uint32), np.ascontiguousarray(sbs.values, np.uint32),
                      np.ascontiguousarray(sbs.values_f, np.float32), np.ascontiguousarray(pairs, np.uint32))
        n, v, vf, pr = self._keep
        self.h = self._L.orc_scene_create(C.byref(sbs.header), _p(n), n.shape[0], _p(v), v.size,
                                          _p(vf) if vf.size else None, vf.size, _p(pr), pr.shape[0])
        if not self.h:
            raise ValueError("orc_scene_create rejected the scene")

    def __del__(self):
        try:
            if self.h:
                self._L.orc_scene_destroy(self.h)
        except Exception:
            pass

    def render(self, world_view, proj, preset, W, H, seed=0, precision=32, threads=0, force_normal=True, rgba=False,
               counters=False):
        t = np.zeros((H, W), np.float64)
        n = np.zeros((H, W, 2), np.float64)
        prim = np.zeros((H, W), np.uint32)
        col = np.zeros((H, W, 4), np.float64) if rgba else None
        cnt = Counters() if counters else None
        wv = np.ascontiguousarray(world_view, np.float32)
        pr = np.ascontiguousarray(proj, np.float32)
        rc = self._L.orc_render(self.h, _p(wv), _p(pr), C.byref(preset), W, H, seed, precision, threads, int(force_normal),
                              _p(col), _p(t), _p(prim), _p(n), C.byref(cnt) if cnt is not None else None)
        if rc != 0:
            raise ValueError(f"orc_render failed {rc}")
        out = {"t": t, "n": n, "prim": prim}
        if rgba:
            out["rgba"] = col
        if counters:
            out["counters"] = {k: getattr(cnt, k) for k, _ in Counters._fields_}
        return out

    def cast_rays(self, preset, rays, precision=32, threads=0, force_normal=True):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 6), np.float64)
        rc = self._L.orc_cast_rays(self.h, C.byref(preset), _p(rays), rays.shape[0], precision, threads, int(force_normal),
                                 _p(out))
        if rc != 0:
            raise ValueError("orc_cast_rays failed")
        return out

    def brute_nearest(self, preset, rays, precision=32, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(rays.shape[0], np.float64)
        prim = np.zeros(rays.shape[0], np.uint32)
        rc = self._L.orc_brute_nearest(self.h, C.byref(preset), _p(rays), rays.shape[0], precision, threads, _p(t), _p(prim))
        if rc != 0:
            raise ValueError("orc_brute_nearest failed")
        return t, prim

    def load_voxel(self, node, vx, vy, vz, precision=32):
        out = np.zeros(8, np.float64)
        vmin = self._L.orc_load_voxel(self.h, node, vx, vy, vz, precision, _p(out))
        return out, vmin

    def set_lights(self, lights):
        """lights: a list of litert_amd.api.Light (same layout as LrtLight)."""
        import ctypes
        arr = (type(lights[0]) * len(lights))(*lights) if lights else None
        if self._L.orc_scene_set_lights(self.h, arr, len(lights)) != 0:
            raise ValueError("bad lights")

    def set_values_f(self, vf):
        vf = np.ascontiguousarray(vf, np.float32)
        self._keep = self._keep[:2] + (vf,) + self._keep[3:]
        if self._L.orc_scene_set_values_f(self.h, _p(vf), vf.size) != 0:
            raise ValueError("size mismatch")

    def dr_regularize(self, n_nodes, lam, power, n_params):
        """Regularization gradient (MultiRendererDR.cpp:1933-2057) of bricks [0, n_nodes)."""
        g = np.zeros(n_params, np.float64)
        if self._L.orc_dr_regularize(self.h, n_nodes, lam, power, _p(g)) != 0:
            raise ValueError("orc_dr_regularize failed")
        return g

    def mlp_features(self, rays, bricks):
        """Seeder features for rays (n x 6: pos, dir) entering bricks (n): (n x 24 float32, t0 float32)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.zeros((rays.shape[0], 24), np.float32)
        t0 = np.zeros(rays.shape[0], np.float32)
        for i in range(rays.shape[0]):
            t = C.c_float()
            if self._L.orc_mlp_features(self.h, _p(rays[i]), int(bricks[i]), _p(out[i]), C.byref(t)) != 0:
                raise ValueError("orc_mlp_features failed")
            t0[i] = t.value
        return out, t0

    def dr_render_grad(self, dr_preset, world_view, proj, W, H, ref, n_params, seed=0, precision=32, threads=0,
                       dtdv_mode=0):
        img = np.zeros((H, W, 4), np.float64)
        grad = np.zeros(n_params, np.float64)
        st = DRStats()
        wv = np.ascontiguousarray(world_view, np.float32)
        pr = np.ascontiguousarray(proj, np.float32)
        ref = np.ascontiguousarray(ref, np.float32)
        rc = self._L.orc_dr_render_grad(self.h, C.byref(dr_preset), _p(wv), _p(pr), W, H, seed, _p(ref), precision, threads,
                                      dtdv_mode, _p(img), _p(grad), C.byref(st))
        if rc != 0:
            raise ValueError(f"orc_dr_render_grad failed {rc}")
        return img, grad, {k: getattr(st, k) for k, _ in DRStats._fields_}


def dr_redistance(lattice, h, iters):
    """Redistance (MultiRendererDR.cpp:2122-2382) of a p x p x p lattice; returns the new lattice."""
    d = np.ascontiguousarray(lattice, np.float32).copy()
    p = d.shape[0]
    if lib().orc_dr_redistance(_p(d), p, h, iters) != 0:
        raise ValueError("orc_dr_redistance failed")
    return d


def sbs_octree(shape, params, depth, brick_size, brick_pad, bytes_per_value, nodes_limit=0):
    """create_sdf_SBS(settings, header, sdf) (utils/sdf_converter.cpp:169-181)
    of a sphere (shape 0, params centre xyz + radius) or a torus around the z
    axis (shape 1, params R, r), with SparseOctreeSettings::nodes_limit (0: its
    default 2^24): (nodes (n, 4) uint32, values uint32) in the frame octree's
    order."""
    prm = np.zeros(4, np.float32)
    prm[:len(params)] = params
    nn, nv = C.c_uint32(), C.c_uint64()
    L = lib()
    rc = L.orc_sbs_octree(shape, _p(prm), depth, brick_size, brick_pad, bytes_per_value, nodes_limit, C.byref(nn), None, C.byref(nv),
                          None)
    if rc != 0:
        raise ValueError(f"orc_sbs_octree failed {rc}")
    nodes = np.zeros((nn.value, 4), np.uint32)
    values = np.zeros(nv.value, np.uint32)
    rc = L.orc_sbs_octree(shape, _p(prm), depth, brick_size, brick_pad, bytes_per_value, nodes_limit, C.byref(nn), _p(nodes),
                          C.byref(nv), _p(values))
    if rc != 0:
        raise ValueError(f"orc_sbs_octree failed {rc}")
    return nodes, values


def resample_reference(images, W, H):
    """PreprocessRefImages (diff_render/MultiRendererDR.cpp:187-226) in FP32:
    images (n, H0, W0, 4) -- colour with the mask in w -- to (n, H, W, 4). Each
    render pixel is the mean of ceil(W0 / W) x ceil(H0 / H) bilinear samples at
    uv = ((spp * x + d + 0.5) / (spp * size)), accumulated in the reference's
    order (dx outer, dy inner), each divided by the sample count. The bilinear
    sample restates LiteImage's Sampler with filter LINEAR and its default
    WRAP addressing (LiteImage is part of LiteMath in the absent HydraCore3
    submodule): the Vulkan specification's linear texel filter, texel centres at
    (i + 0.5) / size, weights (1-a)(1-b), a(1-b), (1-a)b, ab, summed in that
    order. Whether LiteImage wraps or clamps at the image border is parity
    unpinned; the rest is pinned in tests/test_cpu_oracle_pins.py."""
    imgs = np.ascontiguousarray(images, np.float32)
    n, H0, W0, _ = imgs.shape
    f = np.float32
    spp_x = int(np.ceil(f(W0) / f(W)))
    spp_y = int(np.ceil(f(H0) / f(H)))
    n_s = f(spp_x * spp_y)
    xs, ys = np.arange(W, dtype=np.int64), np.arange(H, dtype=np.int64)
    acc = np.zeros((n, H, W, 4), np.float32)
    for dx in range(spp_x):
        for dy in range(spp_y):
            u = ((spp_x * xs + dx).astype(f) + f(0.5)) / f(spp_x * W)
            v = ((spp_y * ys + dy).astype(f) + f(0.5)) / f(spp_y * H)
            fx, fy = u * f(W0) - f(0.5), v * f(H0) - f(0.5)
            x0f, y0f = np.floor(fx), np.floor(fy)
            a, b = fx - x0f, fy - y0f
            x0, y0 = x0f.astype(np.int64) % W0, y0f.astype(np.int64) % H0
            x1, y1 = (x0 + 1) % W0, (y0 + 1) % H0
            oa, ob = f(1.0) - a, f(1.0) - b
            w00 = (oa[None, :] * ob[:, None])[None, :, :, None]
            w10 = (a[None, :] * ob[:, None])[None, :, :, None]
            w01 = (oa[None, :] * b[:, None])[None, :, :, None]
            w11 = (a[None, :] * b[:, None])[None, :, :, None]
            t00, t10 = imgs[:, y0][:, :, x0], imgs[:, y0][:, :, x1]
            t01, t11 = imgs[:, y1][:, :, x0], imgs[:, y1][:, :, x1]
            c = w00 * t00 + w10 * t10 + w01 * t01 + w11 * t11
            acc += c / n_s
    return acc


def eye_ray(world_view, proj, x, y, dx, dy, W, H, precision=32):
    out = np.zeros(8, np.float64)
    lib().orc_eye_ray(_p(np.ascontiguousarray(world_view, np.float32)), _p(np.ascontiguousarray(proj, np.float32)),
                      x, y, dx, dy, W, H, precision, _p(out))
    return out


def mlp_forward(params, x, bf16=False):
    """FP64 forward of the seeder network (optionally emulating BF16 operands): (n x 3)."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 24)
    y = np.zeros((x.shape[0], 3), np.float64)
    lib().orc_mlp_forward(_p(params), _p(x), x.shape[0], int(bf16), _p(y))
    return y


def mlp_loss_grad(params, x, tgt, bf16=False):
    """FP64 mean loss and gradient of the seeder network over n samples."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 24)
    tgt = np.ascontiguousarray(tgt, np.float32).reshape(-1, 3)
    g = np.zeros(params.size, np.float64)
    loss = lib().orc_mlp_loss_grad(_p(params), _p(x), _p(tgt), x.shape[0], int(bf16), _p(g))
    return loss, g

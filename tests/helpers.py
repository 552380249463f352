"""Shared helpers for the parity tests: scene builders and error metrics."""
import functools

import numpy as np

import litert_amd.api as L
import oracle

MISS = 0xFFFFFFFF


@functools.lru_cache(maxsize=None)
def c1_scene(seed=1):
    """Config C1: icosphere(5) r 0.8, 3-octave noise amp 0.05, 128^3 effective
    (brick LOD 32 x 4^3 voxels), pad 1, 32-bit values."""
    return L.displaced_icosphere(5, 0.8, 0.05, 3, seed=seed, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)


@functools.lru_cache(maxsize=None)
def c2_scene(seed=2):
    """Config C2: 4 octaves, 512^3 effective (LOD 128 x 4^3), pad 0, 8-bit
    values quantised over the voxel diagonal."""
    return L.displaced_icosphere(5, 0.8, 0.05, 4, seed=seed, lod=128, brick_size=4, brick_pad=0, bytes_per_value=1,
                                 voxel_diag_scale=True)


@functools.lru_cache(maxsize=None)
def oracle_scene(which):
    sbs = c1_scene() if which == "c1" else c2_scene()
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

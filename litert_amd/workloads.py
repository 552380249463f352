"""Seeded synthetic workloads of the benchmarks (scene, camera and preset
generation only; rendering goes through the library's entry points).

- the uneven batch: scenes of octree depth d ~ U{5..9} (brick LOD 2^(d-2) x
  4^3 voxels), 8-bit distances over the voxel diagonal or FP32, displaced
  icosphere or torus, each with its own orbit camera and W x H ~ U[128, 2048]^2;
- the differentiable-rendering problem: a noisy torus target (R 0.6, r 0.2)
  and a sphere start (r 0.5) as dense indexed SBS grids, seeded orbit views
  and the linear-depth reconstruction preset.
"""
import numpy as np

from . import api as L

BATCH_SEED = 5
DR_SEED = 4


def batch_specs(n_scenes=64, seed=BATCH_SEED):
    """One dict per scene: depth, bytes_per_value, torus, w, h, scene_seed, camera_seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        depth = int(rng.integers(5, 10))
        bpv = int(rng.choice([1, 4]))
        torus = bool(rng.integers(0, 2))
        w, h = int(rng.integers(128, 2049)), int(rng.integers(128, 2049))
        out.append({"depth": depth, "bytes_per_value": bpv, "torus": torus, "w": w, "h": h, "scene_seed": seed * 1000 + k,
                    "camera_seed": seed * 1000 + k})
    return out


def batch_scene(spec):
    kw = dict(seed=spec["scene_seed"], lod=1 << (spec["depth"] - 2), brick_size=4, brick_pad=0,
              bytes_per_value=spec["bytes_per_value"], voxel_diag_scale=spec["bytes_per_value"] == 1)
    if spec["torus"]:
        return L.displaced_torus(0.6, 0.2, 0.02, 3, **kw)
    return L.displaced_icosphere(5, 0.8, 0.05, 3, **kw)


def batch_camera(spec):
    return L.orbit_camera(spec["camera_seed"], 0, aspect=spec["w"] / spec["h"])


def dr_scenes(brick_count=32, brick_size=4, seed=DR_SEED):
    """(target, start): noisy torus and sphere on a dense indexed grid of
    (brick_count * brick_size)^3 voxels."""
    tgt = L.grid_sbs("torus", 0.6, 0.2, 0.02, 3, seed=seed, brick_count=brick_count, brick_size=brick_size)
    ini = L.grid_sbs("sphere", 0.5, 0.0, 0.0, 3, seed=seed, brick_count=brick_count, brick_size=brick_size)
    return tgt, ini


def dr_cameras(views=16, seed=DR_SEED):
    return [L.orbit_camera(seed, i) for i in range(views)]


def dr_reference_preset():
    """Preset of the reference renders: Newton, linear depth."""
    p = L.getDefaultPreset()
    p.sdf_node_intersect = L.SDF_OCTREE_NODE_INTERSECT_NEWTON
    p.render_mode = L.MULTI_RENDER_MODE_LINEAR_DEPTH
    return p


def dr_preset():
    """Linear-depth geometry reconstruction, border_spp 256, lr 1e-3."""
    dr = L.getDefaultPresetDR()
    dr.dr_render_mode = 3  # DR_RENDER_MODE_LINEAR_DEPTH
    dr.dr_input_type = 1   # DR_INPUT_TYPE_LINEAR_DEPTH
    dr.dr_reconstruction_flags = 2  # geometry
    dr.border_spp = 256
    dr.opt_lr = 1e-3
    return dr

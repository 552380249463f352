"""Python binding of liblrt_hip.so (ctypes) for tests and benchmarks.

Mirrors the reference renderer API (Renderer/eye_ray.h:77-252):
``CreateMultiRenderer(DEVICE_GPU)`` returns a :class:`MultiRenderer` with
``SetScene``, ``SetPreset``, ``SetViewport``, ``UpdateCamera``, ``Render``,
``RenderFloat`` and batched ``RayQuery_NearestHit``. All computation runs in
the HIP library; importing this module fails loudly when the library is not
built, and no call falls back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))  # litert_amd/
# LRT_LIBRARY names another build of the same library (same C ABI), e.g. one
# compiled with different occupancy settings; default: the in-tree build.
LIB_PATH = os.environ.get("LRT_LIBRARY") or os.path.join(_HERE, "liblrt_hip.so")

# ---- constants (include/lrt_types.h) ---------------------------------------
DEVICE_CPU, DEVICE_GPU, DEVICE_GPU_RTX = 0, 1, 2
SDF_OCTREE_NODE_INTERSECT_ST, SDF_OCTREE_NODE_INTERSECT_ANALYTIC, SDF_OCTREE_NODE_INTERSECT_NEWTON, \
    SDF_OCTREE_NODE_INTERSECT_BBOX, SDF_OCTREE_NODE_INTERSECT_IT = range(5)
MULTI_RENDER_MODE_MASK = 0
MULTI_RENDER_MODE_LAMBERT_NO_TEX = 1
MULTI_RENDER_MODE_DEPTH = 2
MULTI_RENDER_MODE_LINEAR_DEPTH = 3
MULTI_RENDER_MODE_INVERSE_LINEAR_DEPTH = 4
MULTI_RENDER_MODE_PRIMITIVE = 5
MULTI_RENDER_MODE_TYPE = 6
MULTI_RENDER_MODE_GEOM = 7
MULTI_RENDER_MODE_NORMAL = 8
MULTI_RENDER_MODE_BARYCENTRIC = 9
MULTI_RENDER_MODE_ST_ITERATIONS = 10
MULTI_RENDER_MODE_PHONG_NO_TEX = 12
MULTI_RENDER_MODE_TEX_COORDS = 15
MULTI_RENDER_MODE_DIFFUSE = 16
MULTI_RENDER_MODE_LAMBERT = 17
MULTI_RENDER_MODE_PHONG = 18
DR_BORDER_SAMPLING_RANDOM = 0  # DR_common.h:123-124
DR_BORDER_SAMPLING_SVM = 1
NORMAL_MODE_GEOMETRY, NORMAL_MODE_VERTEX, NORMAL_MODE_SDF_SMOOTHED = 0, 1, 2
RAY_GEN_MODE_REGULAR, RAY_GEN_MODE_RANDOM = 0, 1
REPRESENTATION_MODE_SURFACE, REPRESENTATION_MODE_VOLUME = 0, 1
SDF_SBS_NODE_LAYOUT_DX = 1 << 24
SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F = 4 << 24
SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN = 5 << 24
INVALID_IDX = 1 << 31


class SdfSBSHeader(C.Structure):
    _fields_ = [("brick_size", C.c_uint32), ("brick_pad", C.c_uint32), ("bytes_per_value", C.c_uint32), ("aux_data", C.c_uint32)]


class MultiRenderPreset(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("render_mode", "sdf_node_intersect", "normal_mode", "ray_gen_mode",
                                           "interpolation_mode", "representation_mode", "spp")]


class CRT_Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("primId", C.c_uint32), ("instId", C.c_uint32), ("geomId", C.c_uint32),
                ("coords", C.c_float * 4)]


LIGHT_TYPE_DIRECT, LIGHT_TYPE_POINT, LIGHT_TYPE_AMBIENT = 0, 1, 2


class Light(C.Structure):
    """Renderer/eye_ray.h:48-54: direction (direct) or position (point), type, colour."""
    _fields_ = [("space", C.c_float * 3), ("type", C.c_uint32), ("color", C.c_float * 3), ("_pad", C.c_uint32)]


def make_light(kind, space, color):
    return Light((C.c_float * 3)(*space), kind, (C.c_float * 3)(*color), 0)


class MultiRendererDRPreset(C.Structure):
    _fields_ = [
        ("dr_loss_function", C.c_uint32), ("dr_render_mode", C.c_uint32), ("dr_diff_mode", C.c_uint32),
        ("dr_reconstruction_flags", C.c_uint32), ("dr_input_type", C.c_uint32), ("dr_border_sampling", C.c_uint32),
        ("dr_raycasting_mask", C.c_uint32), ("dr_atomic_ders", C.c_uint32), ("spp", C.c_uint32),
        ("render_width", C.c_uint32), ("render_height", C.c_uint32), ("border_spp", C.c_uint32),
        ("border_relax_eps", C.c_float), ("border_integral_mult", C.c_float), ("opt_lr", C.c_float),
        ("opt_beta_1", C.c_float), ("opt_beta_2", C.c_float), ("opt_eps", C.c_float), ("opt_iterations", C.c_uint32),
        ("image_batch_size", C.c_uint32), ("reg_function", C.c_uint32), ("reg_lambda", C.c_float),
        ("reg_power", C.c_float), ("redistancing_enable", C.c_bool), ("redistancing_interval", C.c_uint32),
        ("dr_raycasting_thickness", C.c_uint32), ("debug_print", C.c_bool), ("debug_print_interval", C.c_uint32),
        ("debug_progress_images", C.c_uint32), ("debug_progress_interval", C.c_uint32),
        ("debug_render_mode", C.c_uint32), ("finite_diff_delta", C.c_float), ("finite_diff_brightness", C.c_float),
        ("debug_forced_border", C.c_bool), ("debug_pd_images", C.c_bool), ("debug_pd_brightness", C.c_float),
        ("debug_border_samples", C.c_bool), ("debug_border_samples_mega_image", C.c_bool),
        ("debug_border_save_normals", C.c_bool),
    ]


MLP_PARAMS = 10115
MLP_FEATURES = 24


def getDefaultPreset():
    """render_settings.h:94-108."""
    return MultiRenderPreset(MULTI_RENDER_MODE_LAMBERT_NO_TEX, SDF_OCTREE_NODE_INTERSECT_ST, NORMAL_MODE_GEOMETRY,
                             RAY_GEN_MODE_REGULAR, 0, REPRESENTATION_MODE_VOLUME, 1)


def getDefaultPresetDR():
    """diff_render/MultiRendererDR.h:14-69."""
    p = MultiRendererDRPreset()
    p.spp = 1
    p.dr_reconstruction_flags = 1
    p.dr_atomic_ders = 0xFFFFFFFF
    p.dr_raycasting_thickness = 5
    p.border_spp = 256
    p.border_relax_eps = 1e-3
    p.border_integral_mult = 1.0
    p.opt_lr = 0.01
    p.opt_beta_1 = 0.9
    p.opt_beta_2 = 0.999
    p.opt_eps = 1e-8
    p.opt_iterations = 500
    p.image_batch_size = 1
    p.reg_lambda = 0.25
    p.reg_power = 2.0
    p.redistancing_interval = 1
    p.debug_print_interval = 10
    p.debug_progress_images = 1001
    p.debug_progress_interval = 100
    p.finite_diff_delta = 0.001
    p.finite_diff_brightness = 100.0
    p.debug_pd_brightness = 0.1
    return p


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (python -m litert_amd.build)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32, u64, f32, i32 = C.c_uint32, C.c_uint64, C.c_float, C.c_int
    sig = {
        "lrt_last_error": (C.c_char_p, []),
        "lrt_create": (i32, [u32, u32, C.POINTER(P)]),
        "lrt_destroy": (i32, [P]),
        "lrt_set_scene_sbs": (i32, [P, C.POINTER(SdfSBSHeader), P, u32, P, u64, P, u64]),
        "lrt_get_bvh": (i32, [P, P, u32, C.POINTER(u32), C.POINTER(u32)]),
        "lrt_build_bvh": (i32, [P, u32, P, u32, C.POINTER(u32), C.POINTER(u32)]),
        "lrt_set_preset": (i32, [P, C.POINTER(MultiRenderPreset)]),
        "lrt_get_preset": (i32, [P, C.POINTER(MultiRenderPreset)]),
        "lrt_set_viewport": (i32, [P, i32, i32, i32, i32]),
        "lrt_update_camera": (i32, [P, P, P]),
        "lrt_set_seed": (i32, [P, u32]),
        "lrt_render": (i32, [P, P, u32, u32, i32]),
        "lrt_render_float": (i32, [P, P, u32, u32, i32]),
        "lrt_render_gbuffer": (i32, [P, P, P, P, u32, u32]),
        "lrt_cast_rays": (i32, [P, P, u32, P]),
        "lrt_get_execution_time": (i32, [P, C.c_char_p, P]),
        "lrt_render_views_device": (i32, [P, P, P, u32, u32, u32, i32, C.POINTER(f32)]),
        "lrt_download_view": (i32, [P, u32, P, P, P, P]),
        "lrt_render_views_noalloc": (i32, [P, P, P, u32, u32, u32, P, P, P, P]),
        "lrt_count_work": (i32, [P, P, P, u32, u32, u32, P]),
        "lrt_sync": (i32, [P]),
        "lrt_set_launch_blocks": (i32, [P, u32]),
        "lrt_set_tile_culling": (i32, [P, i32]),
        "lrt_sbs_displaced_icosphere": (i32, [u32, f32, f32, u32, u32, u32, u32, u32, u32, i32, i32, C.POINTER(P)]),
        "lrt_sbs_grid": (i32, [u32, f32, f32, f32, u32, u32, u32, u32, i32, C.POINTER(P)]),
        "lrt_dr_optimize_grid": (i32, [P, u32, i32, P, u32, u32, i32, C.POINTER(P), C.POINTER(f32)]),
        "lrt_set_lights": (i32, [P, P, u32]),
        "lrt_sbs_from_mesh": (i32, [P, u32, P, u32, u32, u32, u32, u32, i32, i32, C.POINTER(P), C.POINTER(f32)]),
        "lrt_sbs_from_mesh_device": (i32, [P, P, u32, P, u32, u32, u32, u32, u32, i32, C.POINTER(P), C.POINTER(f32)]),
        "lrt_sbs_octree_from_mesh_device": (i32, [P, P, u32, P, u32, u32, u32, u32, u32, u32, C.POINTER(P), C.POINTER(f32)]),
        "lrt_sbs_octree_from_primitive_device": (i32, [P, u32, P, u32, u32, u32, u32, u32, C.POINTER(P), C.POINTER(f32)]),
        "lrt_mesh_displaced_icosphere": (i32, [u32, f32, f32, u32, u32, P, C.POINTER(u32), P, C.POINTER(u32)]),
        "lrt_sbs_info": (i32, [P, C.POINTER(SdfSBSHeader), C.POINTER(u32), C.POINTER(u64), C.POINTER(u64)]),
        "lrt_sbs_copy": (i32, [P, P, P, P]),
        "lrt_sbs_free": (i32, [P]),
        "lrt_orbit_camera": (i32, [u32, u32, f32, f32, f32, f32, f32, P, P]),
        "lrt_sbs_load_file": (i32, [C.c_char_p, C.POINTER(P)]),
        "lrt_sbs_save_file": (i32, [C.c_char_p, C.POINTER(SdfSBSHeader), P, u32, P, u32, P, u32]),
        "lrt_camera_lookat": (i32, [P, P, P, f32, f32, f32, f32, P, P]),
        "lrt_load_scene_hydra": (i32, [P, C.c_char_p]),
    }
    optional = {}
    sig.update({
        "lrt_dr_set_reference": (i32, [P, P, P, u32, u32, u32, P, P]),
        "lrt_dr_render_grad": (i32, [P, C.POINTER(MultiRendererDRPreset), u32, u32, i32, P, P, C.POINTER(f32), C.POINTER(f32)]),
        "lrt_dr_optimize": (i32, [P, C.POINTER(MultiRendererDRPreset), C.POINTER(SdfSBSHeader), P, u32, P, u64, P, u64, u32, i32, P,
                                  C.POINTER(f32)]),
        "lrt_dr_get_last_dloss_ds": (i32, [P, P, u64]),
        "lrt_dr_get_last_image": (i32, [P, u32, P]),
        "lrt_dr_last_stats": (i32, [P, C.POINTER(u32), C.POINTER(u32)]),
        "lrt_dr_render_size": (i32, [P, C.POINTER(u32), C.POINTER(u32)]),
        "lrt_mlp_init": (i32, [P, u32]),
        "lrt_mlp_set_params": (i32, [P, P, u32]),
        "lrt_mlp_get_params": (i32, [P, P, u32]),
        "lrt_mlp_collect": (i32, [P, P, P, u32, u32, u32, u64, C.POINTER(u64), C.POINTER(u64)]),
        "lrt_mlp_dataset_size": (i32, [P, C.POINTER(u64)]),
        "lrt_mlp_dataset_read": (i32, [P, u64, u64, P, P, P]),
        "lrt_mlp_dataset_clear": (i32, [P]),
        "lrt_mlp_train": (i32, [P, u32, u32, f32, u32, P, C.POINTER(f32)]),
        "lrt_mlp_loss_grad": (i32, [P, P, P, u32, P, C.POINTER(f32)]),
        "lrt_mlp_infer": (i32, [P, P, u64, P, C.POINTER(f32)]),
        "lrt_mlp_infer_dataset": (i32, [P, u32, C.POINTER(f32), P, u64]),
        "lrt_set_seeding": (i32, [P, i32, f32]),
        "lrt_render_batch": (i32, [P, u32, P, P, P, P, P, i32, C.POINTER(f32)]),
        "lrt_batch_download": (i32, [P, u32, P, P, P]),
        "lrt_batch_config": (i32, [P, C.POINTER(u32), C.POINTER(u32)]),
        "lrt_sbs_displaced_torus": (i32, [f32, f32, f32, u32, u32, u32, u32, u32, u32, i32, i32, C.POINTER(P)]),
        "lrt_count_work_ex": (i32, [P, P, P, u32, u32, u32, P, u32]),
    })
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    for name, (res, args) in optional.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
    return lib


lib = _load()


class LrtError(RuntimeError):
    pass


def _check(code, what=""):
    if code != 0:
        raise LrtError(f"{what} failed ({code}): {lib.lrt_last_error().decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class SdfSBS:
    """Reference SdfSBS (sdfScene/sdf_scene.h:226-232): header, nodes (n,4)
    uint32 (pos_xy, pos_z_lod_size, data_offset, pad), values, values_f."""

    def __init__(self, header, nodes, values, values_f=None):
        self.header = header
        self.nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 4)
        self.values = np.ascontiguousarray(values, dtype=np.uint32)
        self.values_f = np.zeros(0, np.float32) if values_f is None else _f32(values_f)

    @staticmethod
    def _from_handle(h):
        hdr = SdfSBSHeader()
        nn, nv, nf = C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib.lrt_sbs_info(h, C.byref(hdr), C.byref(nn), C.byref(nv), C.byref(nf)), "lrt_sbs_info")
        nodes = np.zeros((nn.value, 4), np.uint32)
        values = np.zeros(nv.value, np.uint32)
        vf = np.zeros(nf.value, np.float32)
        _check(lib.lrt_sbs_copy(h, _ptr(nodes), _ptr(values), _ptr(vf) if nf.value else None), "lrt_sbs_copy")
        lib.lrt_sbs_free(h)
        return SdfSBS(hdr, nodes, values, vf)

    def lod(self):
        return int(self.nodes[0, 1] & 0xFFFF)


def displaced_icosphere(subdiv=5, radius=0.8, amplitude=0.05, octaves=3, seed=1, lod=32, brick_size=4, brick_pad=1,
                        bytes_per_value=4, voxel_diag_scale=False, threads=0):
    h = C.c_void_p()
    _check(lib.lrt_sbs_displaced_icosphere(subdiv, radius, amplitude, octaves, seed, lod, brick_size, brick_pad,
                                           bytes_per_value, int(voxel_diag_scale), threads, C.byref(h)),
           "lrt_sbs_displaced_icosphere")
    return SdfSBS._from_handle(h)


def displaced_torus(R=0.6, r=0.2, amplitude=0.02, octaves=3, seed=1, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1,
                    voxel_diag_scale=False, threads=0):
    """Noise-displaced torus as a sparse DX-layout SBS (no reference counterpart)."""
    h = C.c_void_p()
    _check(lib.lrt_sbs_displaced_torus(R, r, amplitude, octaves, seed, lod, brick_size, brick_pad, bytes_per_value,
                                       int(voxel_diag_scale), threads, C.byref(h)), "displaced_torus")
    return SdfSBS._from_handle(h)


def mesh_displaced_icosphere(subdiv=5, radius=0.8, amplitude=0.05, octaves=3, seed=1):
    """The displaced icosphere mesh behind displaced_icosphere(): (vertices (n, 3) float32, triangles (m, 3) uint32)."""
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    _check(lib.lrt_mesh_displaced_icosphere(subdiv, radius, amplitude, octaves, seed, None, C.byref(nv), None, C.byref(nt)),
           "lrt_mesh_displaced_icosphere")
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    _check(lib.lrt_mesh_displaced_icosphere(subdiv, radius, amplitude, octaves, seed, _ptr(v), C.byref(nv), _ptr(t), C.byref(nt)),
           "lrt_mesh_displaced_icosphere")
    return v, t


def sbs_from_mesh(vertices, triangles, lod=32, brick_size=4, brick_pad=0, bytes_per_value=1, voxel_diag_scale=False, threads=0,
                  renderer=None):
    """SBS of a closed triangle mesh (sdf_converter::create_sdf_SBS, utils/sdf_converter.cpp:183-228); on the
    host (OpenMP) or, with `renderer`, on its GPU. Returns (SdfSBS, build_ms)."""
    v = _f32(vertices).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    h = C.c_void_p()
    ms = C.c_float(0.0)
    if renderer is None:
        _check(lib.lrt_sbs_from_mesh(_ptr(v), v.shape[0], _ptr(t), t.shape[0], lod, brick_size, brick_pad, bytes_per_value,
                                     int(voxel_diag_scale), threads, C.byref(h), C.byref(ms)), "lrt_sbs_from_mesh")
    else:
        _check(lib.lrt_sbs_from_mesh_device(renderer._h, _ptr(v), v.shape[0], _ptr(t), t.shape[0], lod, brick_size, brick_pad,
                                            bytes_per_value, int(voxel_diag_scale), C.byref(h), C.byref(ms)),
               "lrt_sbs_from_mesh_device")
    return SdfSBS._from_handle(h), ms.value


PRIMITIVE_SPHERE, PRIMITIVE_TORUS = 0, 1


def sbs_octree_from_mesh(renderer, vertices, triangles, depth, brick_size=4, brick_pad=0, bytes_per_value=1, nodes_limit=0):
    """create_sdf_SBS(settings, header, mesh) with the reference's adaptive frame octree
    (utils/sparse_octree_builder.cpp:125-444) on the renderer's GPU: mixed-LOD bricks down to
    octree depth `depth`, frame_octree_limit_nodes at nodes_limit (0: 2^24). Returns (SdfSBS, build_ms)."""
    v = _f32(vertices).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    h = C.c_void_p()
    ms = C.c_float(0.0)
    _check(lib.lrt_sbs_octree_from_mesh_device(renderer._h, _ptr(v), v.shape[0], _ptr(t), t.shape[0], depth, brick_size, brick_pad,
                                               bytes_per_value, nodes_limit, C.byref(h), C.byref(ms)), "lrt_sbs_octree_from_mesh_device")
    return SdfSBS._from_handle(h), ms.value


def sbs_octree_from_primitive(renderer, shape, params, depth, brick_size=4, brick_pad=0, bytes_per_value=1, nodes_limit=0):
    """create_sdf_SBS(settings, header, sdf) with the reference's adaptive frame octree for a sphere
    (PRIMITIVE_SPHERE, params centre xyz + radius) or a torus around the z axis (PRIMITIVE_TORUS,
    params R, r) on the renderer's GPU, frame_octree_limit_nodes at nodes_limit (0: 2^24).
    Returns (SdfSBS, build_ms)."""
    prm = np.zeros(4, np.float32)
    prm[:len(params)] = params
    h = C.c_void_p()
    ms = C.c_float(0.0)
    _check(lib.lrt_sbs_octree_from_primitive_device(renderer._h, shape, _ptr(prm), depth, brick_size, brick_pad, bytes_per_value,
                                                    nodes_limit, C.byref(h), C.byref(ms)), "lrt_sbs_octree_from_primitive_device")
    return SdfSBS._from_handle(h), ms.value


def grid_sbs(shape="torus", R=0.6, r=0.2, amplitude=0.0, octaves=3, seed=1, brick_count=32, brick_size=4,
             with_neighbors=False):
    h = C.c_void_p()
    _check(lib.lrt_sbs_grid(1 if shape == "torus" else 0, R, r, amplitude, octaves, seed, brick_count, brick_size,
                            int(with_neighbors), C.byref(h)), "lrt_sbs_grid")
    return SdfSBS._from_handle(h)


def build_bvh(sbs):
    """Host-side BVH2 build for the bricks of `sbs` (same tree as SetScene)."""
    n, depth = C.c_uint32(), C.c_uint32()
    _check(lib.lrt_build_bvh(_ptr(sbs.nodes), sbs.nodes.shape[0], None, 0, C.byref(n), C.byref(depth)), "lrt_build_bvh")
    pairs = np.zeros((n.value, 16), np.uint32)
    _check(lib.lrt_build_bvh(_ptr(sbs.nodes), sbs.nodes.shape[0], _ptr(pairs), n.value, None, None), "lrt_build_bvh")
    return pairs, depth.value


def load_sbs(path):
    """load_sdf_SBS (sdfScene/sdf_scene.cpp:163-175)."""
    h = C.c_void_p()
    _check(lib.lrt_sbs_load_file(os.fsencode(path), C.byref(h)), "lrt_sbs_load_file")
    return SdfSBS._from_handle(h)


def save_sbs(path, sbs):
    nf = sbs.values_f.size
    _check(lib.lrt_sbs_save_file(os.fsencode(path), C.byref(sbs.header), _ptr(sbs.nodes), sbs.nodes.shape[0], _ptr(sbs.values),
                                 sbs.values.size, _ptr(sbs.values_f) if nf else None, nf), "lrt_sbs_save_file")


def look_at_camera(eye, target, up, fov=60.0, aspect=1.0, znear=0.1, zfar=100.0):
    wv = np.zeros(16, np.float32)
    pr = np.zeros(16, np.float32)
    _check(lib.lrt_camera_lookat(_ptr(_f32(eye)), _ptr(_f32(target)), _ptr(_f32(up)), fov, aspect, znear, zfar, _ptr(wv), _ptr(pr)),
           "lrt_camera_lookat")
    return wv, pr


def orbit_camera(seed, index, distance=3.0, fov=60.0, aspect=1.0, znear=0.01, zfar=100.0):
    wv = np.zeros(16, np.float32)
    pr = np.zeros(16, np.float32)
    _check(lib.lrt_orbit_camera(seed, index, distance, fov, aspect, znear, zfar, _ptr(wv), _ptr(pr)), "lrt_orbit_camera")
    return wv, pr


class MultiRenderer:
    """GPU MultiRenderer (Renderer/eye_ray.h:77-250) backed by liblrt_hip.so."""

    def __init__(self, device=DEVICE_GPU, max_primitives=10_000_000):
        self._h = C.c_void_p()
        _check(lib.lrt_create(device, max_primitives, C.byref(self._h)), "lrt_create")
        self.width = self.height = 0

    def close(self):
        if self._h:
            lib.lrt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SetScene(self, sbs):
        nv = sbs.values.size
        nf = sbs.values_f.size
        _check(lib.lrt_set_scene_sbs(self._h, C.byref(sbs.header), _ptr(sbs.nodes), sbs.nodes.shape[0], _ptr(sbs.values),
                                     nv, _ptr(sbs.values_f) if nf else None, nf), "SetScene")

    def LoadScene(self, path):
        """MultiRenderer::LoadScene / LoadSceneHydra (Renderer/eye_ray_host.cpp:52-121): one identity
        instance of an sdf_sbs geometry node and the first camera of a Hydra XML scene."""
        _check(lib.lrt_load_scene_hydra(self._h, os.fsencode(path)), "LoadScene")

    def GetBVH(self):
        n, depth = C.c_uint32(), C.c_uint32()
        _check(lib.lrt_get_bvh(self._h, None, 0, C.byref(n), C.byref(depth)), "GetBVH")
        pairs = np.zeros((n.value, 16), np.uint32)  # 2 x (boxMin[3], leftOffset, boxMax[3], escapeIndex)
        _check(lib.lrt_get_bvh(self._h, _ptr(pairs), n.value, None, None), "GetBVH")
        return pairs, depth.value

    def SetLights(self, lights):
        """MultiRenderer::SetLights (eye_ray.h:158): a list of Light."""
        arr = (Light * max(1, len(lights)))(*lights)
        _check(lib.lrt_set_lights(self._h, arr, len(lights)), "SetLights")

    def SetPreset(self, preset):
        _check(lib.lrt_set_preset(self._h, C.byref(preset)), "SetPreset")

    def GetPreset(self):
        p = MultiRenderPreset()
        _check(lib.lrt_get_preset(self._h, C.byref(p)), "GetPreset")
        return p

    def SetViewport(self, x0, y0, w, h):
        _check(lib.lrt_set_viewport(self._h, x0, y0, w, h), "SetViewport")
        self.width, self.height = w, h

    def UpdateCamera(self, world_view, proj):
        _check(lib.lrt_update_camera(self._h, _ptr(_f32(world_view)), _ptr(_f32(proj))), "UpdateCamera")

    def setSeed(self, seed):
        _check(lib.lrt_set_seed(self._h, seed), "setSeed")

    def RenderFloat(self, width, height, world_view=None, proj=None, preset=None, passes=1):
        """MultiRenderer::RenderFloat (eye_ray_host.cpp:575-585); returns (H, W, 4) float32."""
        if world_view is not None:
            self.SetViewport(0, 0, width, height)
            self.UpdateCamera(world_view, proj)
        if preset is not None:
            self.SetPreset(preset)
        img = np.zeros((height, width, 4), np.float32)
        _check(lib.lrt_render_float(self._h, _ptr(img), width, height, passes), "RenderFloat")
        return img

    def Render(self, width, height, world_view=None, proj=None, preset=None, passes=1):
        """MultiRenderer::Render (eye_ray_host.cpp:563-573); returns (H, W) uint32 RGBA8."""
        if world_view is not None:
            self.SetViewport(0, 0, width, height)
            self.UpdateCamera(world_view, proj)
        if preset is not None:
            self.SetPreset(preset)
        img = np.zeros((height, width), np.uint32)
        _check(lib.lrt_render(self._h, _ptr(img), width, height, passes), "Render")
        return img

    def RenderGBuffer(self, width, height, world_view=None, proj=None):
        if world_view is not None:
            self.UpdateCamera(world_view, proj)
        t = np.zeros((height, width), np.float32)
        n = np.zeros((height, width, 2), np.float32)
        prim = np.zeros((height, width), np.uint32)
        _check(lib.lrt_render_gbuffer(self._h, _ptr(t), _ptr(n), _ptr(prim), width, height), "RenderGBuffer")
        return t, n, prim

    def RayQuery_NearestHit(self, rays):
        """Batched RayQuery_NearestHit; rays (n, 8) float32 -> CRT_Hit array."""
        rays = _f32(rays).reshape(-1, 8)
        out = (CRT_Hit * rays.shape[0])()
        _check(lib.lrt_cast_rays(self._h, _ptr(rays), rays.shape[0], C.cast(out, C.c_void_p)), "RayQuery_NearestHit")
        return out

    def GetExecutionTime(self, name):
        out = (C.c_float * 4)()
        _check(lib.lrt_get_execution_time(self._h, name.encode(), C.cast(out, C.c_void_p)), "GetExecutionTime")
        return list(out)

    def RenderViewsDevice(self, world_views, projs, width, height, out_mode=0, timed=True):
        wv = _f32(world_views).reshape(-1, 16)
        pr = _f32(projs).reshape(-1, 16)
        ms = C.c_float(0.0)
        _check(lib.lrt_render_views_device(self._h, _ptr(wv), _ptr(pr), wv.shape[0], width, height, out_mode,
                                           C.byref(ms) if timed else None), "RenderViewsDevice")
        return ms.value

    def RenderViewsNoAlloc(self, world_views, projs, width, height, d_t=None, d_normal=None, d_prim=None, d_rgba=None):
        """lrt_render_views_noalloc: renders into caller-owned device buffers
        (integer device addresses, or None) without allocating; asynchronous
        (Sync() waits)."""
        wv = _f32(world_views).reshape(-1, 16)
        pr = _f32(projs).reshape(-1, 16)
        _check(lib.lrt_render_views_noalloc(self._h, _ptr(wv), _ptr(pr), wv.shape[0], width, height, d_t, d_normal, d_prim, d_rgba),
               "RenderViewsNoAlloc")

    def DownloadView(self, view, width, height, rgba=False):
        t = np.zeros((height, width), np.float32)
        n = np.zeros((height, width, 2), np.float32)
        prim = np.zeros((height, width), np.uint32)
        col = np.zeros((height, width, 4), np.float32) if rgba else None
        _check(lib.lrt_download_view(self._h, view, None if rgba else _ptr(t), None if rgba else _ptr(n),
                                     None if rgba else _ptr(prim), _ptr(col)), "DownloadView")
        return col if rgba else (t, n, prim)

    def CountWork(self, world_views, projs, width, height):
        wv = _f32(world_views).reshape(-1, 16)
        pr = _f32(projs).reshape(-1, 16)
        out = np.zeros(3, np.uint64)
        _check(lib.lrt_count_work(self._h, _ptr(wv), _ptr(pr), wv.shape[0], width, height, _ptr(out)), "CountWork")
        return out

    def CountWorkEx(self, world_views, projs, width, height):
        """(pairs, bricks, voxels, root iterations, MLP tiles, gated seeds, culled 8x8 tiles) of one counted launch."""
        wv, pr = _f32(world_views).reshape(-1, 16), _f32(projs).reshape(-1, 16)
        out = np.zeros(7, np.uint64)
        _check(lib.lrt_count_work_ex(self._h, _ptr(wv), _ptr(pr), wv.shape[0], width, height, _ptr(out), 7), "CountWorkEx")
        return out

    # ---- learned root-find seeder (no reference counterpart) ----
    def MlpInit(self, seed=0):
        _check(lib.lrt_mlp_init(self._h, seed), "MlpInit")

    def MlpSetParams(self, params):
        p = _f32(params).ravel()
        _check(lib.lrt_mlp_set_params(self._h, _ptr(p), p.size), "MlpSetParams")

    def MlpGetParams(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        _check(lib.lrt_mlp_get_params(self._h, _ptr(out), out.size), "MlpGetParams")
        return out

    def MlpCollect(self, world_views, projs, width, height, max_events):
        """Appends training events of one plain render; returns (kept, produced)."""
        wv, pr = _f32(world_views).reshape(-1, 16), _f32(projs).reshape(-1, 16)
        a, b = C.c_uint64(), C.c_uint64()
        _check(lib.lrt_mlp_collect(self._h, _ptr(wv), _ptr(pr), wv.shape[0], width, height, max_events, C.byref(a), C.byref(b)),
               "MlpCollect")
        return a.value, b.value

    def MlpDatasetSize(self):
        n = C.c_uint64()
        _check(lib.lrt_mlp_dataset_size(self._h, C.byref(n)), "MlpDatasetSize")
        return n.value

    def MlpDataset(self, first=0, count=None):
        n = self.MlpDatasetSize()
        count = n - first if count is None else count
        f = np.zeros((count, MLP_FEATURES), np.float32)
        t = np.zeros((count, 3), np.float32)
        m = np.zeros((count, 2), np.uint32)
        _check(lib.lrt_mlp_dataset_read(self._h, first, count, _ptr(f), _ptr(t), _ptr(m)), "MlpDataset")
        return f, t, m

    def MlpDatasetClear(self):
        _check(lib.lrt_mlp_dataset_clear(self._h), "MlpDatasetClear")

    def MlpTrain(self, epochs, batch=16384, lr=1e-3, seed=0):
        losses = np.zeros(epochs, np.float32)
        ms = C.c_float()
        _check(lib.lrt_mlp_train(self._h, epochs, batch, lr, seed, _ptr(losses), C.byref(ms)), "MlpTrain")
        return losses, ms.value

    def MlpLossGrad(self, feats, tgts):
        f, t = _f32(feats).reshape(-1, MLP_FEATURES), _f32(tgts).reshape(-1, 3)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib.lrt_mlp_loss_grad(self._h, _ptr(f), _ptr(t), f.shape[0], _ptr(g), C.byref(loss)), "MlpLossGrad")
        return loss.value, g

    def MlpInfer(self, feats):
        f = _f32(feats).reshape(-1, MLP_FEATURES)
        out = np.zeros((f.shape[0], 3), np.float32)
        ms = C.c_float()
        _check(lib.lrt_mlp_infer(self._h, _ptr(f), f.shape[0], _ptr(out), C.byref(ms)), "MlpInfer")
        return out, ms.value

    def MlpInferDataset(self, repeats=10, n_out=0):
        out = np.zeros((n_out, 3), np.float32)
        ms = C.c_float()
        _check(lib.lrt_mlp_infer_dataset(self._h, repeats, C.byref(ms), _ptr(out) if n_out else None, n_out), "MlpInferDataset")
        return ms.value, out

    # ---- batched uneven workload (no reference counterpart) ----
    def RenderBatch(self, scenes, world_views, projs, widths, heights, autotune=False):
        """One launch over len(scenes) items (renderers holding scenes); returns kernel ms."""
        hs = (C.c_void_p * len(scenes))(*[s._h for s in scenes])
        wv, pr = _f32(world_views).reshape(-1, 16), _f32(projs).reshape(-1, 16)
        w = np.ascontiguousarray(widths, np.uint32)
        h = np.ascontiguousarray(heights, np.uint32)
        ms = C.c_float()
        _check(lib.lrt_render_batch(self._h, len(scenes), hs, _ptr(wv), _ptr(pr), _ptr(w), _ptr(h), int(bool(autotune)),
                                    C.byref(ms)), "RenderBatch")
        self._batch_dims = list(zip(w.tolist(), h.tolist()))
        return ms.value

    def BatchDownload(self, item):
        W, H = self._batch_dims[item]
        t = np.zeros((H, W), np.float32)
        n = np.zeros((H, W, 2), np.float32)
        prim = np.zeros((H, W), np.uint32)
        _check(lib.lrt_batch_download(self._h, item, _ptr(t), _ptr(n), _ptr(prim)), "BatchDownload")
        return t, n, prim

    def BatchConfig(self):
        a, b = C.c_uint32(), C.c_uint32()
        _check(lib.lrt_batch_config(self._h, C.byref(a), C.byref(b)), "BatchConfig")
        return a.value, b.value

    def SetSeeding(self, enabled, threshold=0.9):
        _check(lib.lrt_set_seeding(self._h, int(bool(enabled)), threshold), "SetSeeding")

    def Sync(self):
        _check(lib.lrt_sync(self._h), "Sync")

    def SetLaunchBlocks(self, blocks):
        _check(lib.lrt_set_launch_blocks(self._h, blocks), "SetLaunchBlocks")

    def SetTileCulling(self, on):
        """lrt_set_tile_culling: skip 8x8 tiles that provably miss (identical images)."""
        _check(lib.lrt_set_tile_culling(self._h, int(bool(on))), "SetTileCulling")


class MultiRendererDR(MultiRenderer):
    """dr::MultiRendererDR (diff_render/MultiRendererDR.h:71-155) on the GPU."""

    def SetReference(self, images, world_views, projs, masks=None):
        imgs = _f32(images)
        n, H, W = imgs.shape[0], imgs.shape[1], imgs.shape[2]
        m = None if masks is None else _f32(masks)
        _check(lib.lrt_dr_set_reference(self._h, _ptr(imgs), _ptr(m), n, W, H, _ptr(_f32(world_views)), _ptr(_f32(projs))),
               "SetReference")
        self.ref_shape = (n, H, W)

    def RenderSize(self, preset=None):
        """(W, H) a preset renders at (MultiRendererDR.cpp:332-349: render_width x
        render_height when both are set, else the reference images' size); with
        no preset, the size of the last render."""
        if preset is not None:
            n, H, W = self.ref_shape
            if preset.render_width and preset.render_height:
                return int(preset.render_width), int(preset.render_height)
            return W, H
        w, h = C.c_uint32(), C.c_uint32()
        _check(lib.lrt_dr_render_size(self._h, C.byref(w), C.byref(h)), "RenderSize")
        return w.value, h.value

    def RenderDR(self, preset, view, seed=0, dtdv_mode=0, n_params=None):
        """One RenderDR (MultiRendererDR.cpp:578-750): returns (image, dLoss/dS, loss, kernel_ms)."""
        W, H = self.RenderSize(preset)
        img = np.zeros((H, W, 4), np.float32)
        g = None if n_params is None else np.zeros(n_params, np.float32)
        loss, ms = C.c_float(), C.c_float()
        _check(lib.lrt_dr_render_grad(self._h, C.byref(preset), view, seed, dtdv_mode, _ptr(img), _ptr(g), C.byref(loss),
                                      C.byref(ms)), "RenderDR")
        return img, g, loss.value, ms.value

    def LastStats(self):
        a, b = C.c_uint32(), C.c_uint32()
        _check(lib.lrt_dr_last_stats(self._h, C.byref(a), C.byref(b)), "LastStats")
        return a.value, b.value

    def OptimizeFixedStructure(self, preset, sbs, seed=0, dtdv_mode=0):
        """Runs preset.opt_iterations Adam steps; updates sbs.values_f in place; returns (losses, total_ms)."""
        losses = np.zeros(preset.opt_iterations, np.float32)
        ms = C.c_float()
        _check(lib.lrt_dr_optimize(self._h, C.byref(preset), C.byref(sbs.header), _ptr(sbs.nodes), sbs.nodes.shape[0],
                                   _ptr(sbs.values), sbs.values.size, _ptr(sbs.values_f), sbs.values_f.size, seed, dtdv_mode,
                                   _ptr(losses), C.byref(ms)), "OptimizeFixedStructure")
        return losses, ms.value

    def OptimizeGrid(self, start_grid_size, no_last_step_resize, presets, seed=0, dtdv_mode=0):
        """MultiRendererDR::OptimizeGrid (MultiRendererDR.cpp:228-276); returns (final grid SdfSBS, total_ms)."""
        arr = (MultiRendererDRPreset * len(presets))(*presets)
        h = C.c_void_p()
        ms = C.c_float()
        _check(lib.lrt_dr_optimize_grid(self._h, start_grid_size, int(no_last_step_resize), arr, len(presets), seed, dtdv_mode,
                                        C.byref(h), C.byref(ms)), "OptimizeGrid")
        return SdfSBS._from_handle(h), ms.value

    def getLastdLoss_dS(self, n):
        out = np.zeros(n, np.float32)
        _check(lib.lrt_dr_get_last_dloss_ds(self._h, _ptr(out), n), "getLastdLoss_dS")
        return out

    def getLastImage(self, view):
        """MultiRendererDR::getLastImage (MultiRendererDR.h:86): (H, W, 4) at the last render size."""
        W, H = self.RenderSize()
        out = np.zeros((H, W, 4), np.float32)
        _check(lib.lrt_dr_get_last_image(self._h, view, _ptr(out)), "getLastImage")
        return out


def CreateMultiRenderer(device=DEVICE_GPU, max_primitives=10_000_000):
    """CreateMultiRenderer (Renderer/eye_ray.h:252)."""
    return MultiRenderer(device, max_primitives)

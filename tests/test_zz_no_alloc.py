"""lrt_render_views_noalloc against lrt_render_views_device.

The no-allocation entry point renders into caller-owned device buffers with
its scratch in a per-GPU __device__ array. Its output must equal the
allocating entry point's bit for bit (same kernel), on two calls in a row
(the second with fewer views and a smaller image, so the shared scratch holds
the first call's contents), for G-buffer and RGBA output; a call must leave
the device's free memory unchanged, and requests beyond its limits must fail
with an error.
"""
import ctypes as C
import os

import numpy as np
import pytest

import litert_amd.api as L
from helpers import ico128_scene, preset

pytestmark = pytest.mark.gpu

_HIP = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        self.nbytes = nbytes
        _check(_HIP.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)), "hipMalloc")

    def fill(self, byte):
        _check(_HIP.hipMemset(self.p, byte, C.c_size_t(self.nbytes)), "hipMemset")

    def download(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(_HIP.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def free(self):
        if self.p:
            _check(_HIP.hipFree(self.p), "hipFree")
            self.p = C.c_void_p()


def _free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    _check(_HIP.hipDeviceSynchronize(), "hipDeviceSynchronize")
    _check(_HIP.hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value


@pytest.fixture(scope="module")
def renderer():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    r.SetScene(ico128_scene())
    yield r
    r.close()


def _cams(n, w, h):
    return [L.orbit_camera(31, i, aspect=w / h) for i in range(n)]


def _reference(r, cams, W, H, rgba):
    wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
    r.RenderViewsDevice(wvs, prs, W, H, 1 if rgba else 0)
    return [r.DownloadView(v, W, H, rgba=rgba) for v in range(len(cams))]


def test_no_alloc_gbuffer_matches_allocating_entry_point_twice(renderer):
    r = renderer
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    n1, W1, H1 = 8, 320, 256
    bt, bn, bp = DeviceBuffer(n1 * W1 * H1 * 4), DeviceBuffer(n1 * W1 * H1 * 8), DeviceBuffer(n1 * W1 * H1 * 4)
    try:
        for n, W, H in ((n1, W1, H1), (3, 200, 72)):  # the second call is smaller
            for b in (bt, bn, bp):
                b.fill(0x7F)
            cams = _cams(n, W, H)
            wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
            r.RenderViewsNoAlloc(wvs, prs, W, H, bt.p.value, bn.p.value, bp.p.value, None)
            r.Sync()
            t = bt.download(np.float32, (n, H, W))
            nn = bn.download(np.float32, (n, H, W, 2))
            prim = bp.download(np.uint32, (n, H, W))
            ref = _reference(r, cams, W, H, False)
            for v in range(n):
                rt, rn, rp = ref[v]
                assert np.array_equal(prim[v], rp), (n, W, H, v)
                assert np.array_equal(t[v], rt), (n, W, H, v)
                assert np.array_equal(nn[v], rn), (n, W, H, v)
            assert (prim != 0xFFFFFFFF).sum() > 1000
    finally:
        for b in (bt, bn, bp):
            b.free()


def test_no_alloc_rgba_matches_allocating_entry_point(renderer):
    r = renderer
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_PHONG_NO_TEX))
    n, W, H = 2, 128, 96
    bc = DeviceBuffer(n * W * H * 16)
    try:
        cams = _cams(n, W, H)
        wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
        r.RenderViewsNoAlloc(wvs, prs, W, H, None, None, None, bc.p.value)
        r.Sync()
        col = bc.download(np.float32, (n, H, W, 4))
        ref = _reference(r, cams, W, H, True)
        for v in range(n):
            assert np.array_equal(col[v], ref[v]), v
    finally:
        bc.free()


def test_no_alloc_call_leaves_free_device_memory_unchanged(renderer):
    r = renderer
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    n, W, H = 4, 640, 480  # larger than any image this handle rendered before
    bt = DeviceBuffer(n * W * H * 4)
    try:
        cams = _cams(n, W, H)
        wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
        r.RenderViewsNoAlloc(wvs, prs, 64, 64, bt.p.value, None, None, None)  # first use on this device: module and scratch set up
        r.Sync()
        before = _free_bytes()
        r.RenderViewsNoAlloc(wvs, prs, W, H, bt.p.value, None, None, None)
        r.Sync()
        assert _free_bytes() == before
    finally:
        bt.free()


def test_no_alloc_rejects_what_it_cannot_do(renderer):
    r = renderer
    r.SetPreset(preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON, render_mode=L.MULTI_RENDER_MODE_NORMAL))
    bt = DeviceBuffer(17 * 16 * 16 * 4)
    try:
        cams = _cams(17, 16, 16)
        wvs, prs = np.stack([c[0] for c in cams]), np.stack([c[1] for c in cams])
        with pytest.raises(L.LrtError, match="views per call"):
            r.RenderViewsNoAlloc(wvs, prs, 16, 16, bt.p.value, None, None, None)
        with pytest.raises(L.LrtError, match="no output"):
            r.RenderViewsNoAlloc(wvs[:1], prs[:1], 16, 16, None, None, None, None)
        params = np.random.default_rng(1).normal(0, 0.1, L.MLP_PARAMS).astype(np.float32)
        r.MlpSetParams(params)
        r.SetSeeding(True, 0.9)
        try:
            with pytest.raises(L.LrtError, match="seeder"):
                r.RenderViewsNoAlloc(wvs[:1], prs[:1], 16, 16, bt.p.value, None, None, None)
        finally:
            r.SetSeeding(False)
    finally:
        bt.free()

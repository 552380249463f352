"""Device SBS builder (sdf_converter::create_sdf_SBS from a triangle mesh,
utils/sdf_converter.cpp:183-228) against the host builder: the same bricks,
positions and quantised words bit for bit, for several meshes, LODs, brick
pads, value widths and quantisation ranges. The host builder's border rule and
quantiser are checked against the reference's in the CPU tests."""
import numpy as np
import pytest

import litert_amd.api as L

pytestmark = pytest.mark.gpu


def torus_mesh(R=0.6, r=0.2, nu=96, nv=48):
    """Closed parametric torus mesh (quads split in two), axis z."""
    u = np.arange(nu) * 2 * np.pi / nu
    v = np.arange(nv) * 2 * np.pi / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    verts = np.stack([(R + r * np.cos(V)) * np.cos(U), (R + r * np.cos(V)) * np.sin(U), r * np.sin(V)], -1).reshape(-1, 3)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    tris = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris += [(a, b, c), (a, c, d)]
    return verts.astype(np.float32), np.array(tris, np.uint32)


@pytest.fixture(scope="module")
def gpu():
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    yield r
    r.close()


CASES = [  # mesh, lod, brick_size, brick_pad, bytes_per_value, voxel_diag_scale
    ("ico5", 128, 4, 0, 1, True),
    ("ico5", 64, 4, 1, 4, False),
    ("ico4", 32, 8, 0, 2, False),
    ("torus", 64, 4, 0, 1, True),
    ("torus", 32, 4, 1, 4, False),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-lod{c[1]}-bs{c[2]}-pad{c[3]}-bpv{c[4]}" for c in CASES])
def test_device_builder_matches_host_builder(gpu, case):
    mesh, lod, bs, pad, bpv, vds = case
    if mesh.startswith("ico"):
        v, t = L.mesh_displaced_icosphere(int(mesh[3:]), 0.8, 0.05, 4, 7)
    else:
        v, t = torus_mesh()
    host, host_ms = L.sbs_from_mesh(v, t, lod, bs, pad, bpv, vds)
    dev, dev_ms = L.sbs_from_mesh(v, t, lod, bs, pad, bpv, vds, renderer=gpu)
    print(case, host.nodes.shape[0], "bricks; host", round(host_ms, 1), "ms, device", round(dev_ms, 1), "ms")
    assert host.nodes.shape[0] > 100
    assert (host.header.brick_size, host.header.brick_pad, host.header.bytes_per_value, host.header.aux_data) == \
           (dev.header.brick_size, dev.header.brick_pad, dev.header.bytes_per_value, dev.header.aux_data)
    assert np.array_equal(host.nodes, dev.nodes)
    assert np.array_equal(host.values, dev.values)


def test_device_built_scene_renders(gpu):
    """The device-built scene goes straight into SetScene and renders like the
    host-built one."""
    v, t = L.mesh_displaced_icosphere(5, 0.8, 0.05, 4, 7)
    dev, _ = L.sbs_from_mesh(v, t, 64, 4, 0, 1, True, renderer=gpu)
    host, _ = L.sbs_from_mesh(v, t, 64, 4, 0, 1, True)
    wv, pr = L.orbit_camera(3, 0)
    out = []
    for s in (dev, host):
        gpu.SetScene(s)
        out.append(gpu.RenderGBuffer(128, 128, wv, pr))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert np.count_nonzero(out[0][2] != 0xFFFFFFFF) > 2000

"""MultiRenderer::LoadScene / LoadSceneHydra (Renderer/eye_ray_host.cpp:52-121)
for the sdf_sbs geometry node (BVH2Common_host.cpp:142-148): a Hydra XML scene
in the layout of the reference's own scene files (scenes/02_sdf_scenes/
test_10.xml: materials_lib, geometry_lib with a loc relative to the scene's
folder, cam_lib, scenes/scene/instance) around the reference's SBS fixture.
The loaded scene and camera render the same G-buffer as SetScene +
UpdateCamera with the same camera; what this path cannot render is refused."""
import os
import shutil

import numpy as np
import pytest

import litert_amd.api as L

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "test_10_sbs.bin")

SCENE = """<?xml version="1.0"?>
<!-- a Hydra scene with one SBS node -->
<materials_lib>
  <material id="0" name="mysimplemat" type="hydra_material">
    <diffuse brdf_type="lambert"><color val="0.5 0.5 0.5" /></diffuse>
  </material>
</materials_lib>
<geometry_lib total_chunks="1">
  {geoms}
</geometry_lib>
<cam_lib>
  <camera id="0" name="my camera" type="uvn">
    <fov>50</fov>
    <nearClipPlane>0.01</nearClipPlane>
    <farClipPlane>100.0</farClipPlane>
    <up>0 1 0</up>
    <position>0.4 0.3 2.8</position>
    <look_at>0 0 0</look_at>
  </camera>
</cam_lib>
<scenes>
  <scene id="0" name="my scene" discard="1" bbox="-1 1 -1 1 -1 1">
    <remap_lists><remap_list id="0" size="2" val="0 4 " /></remap_lists>
    {instances}
  </scene>
</scenes>
"""
SBS_NODE = '<sdf_sbs id="{id}" name="teapot" type="sdf_sbs" bytesize="1" loc="data/test_10_sbs.bin"><material id="0" /></sdf_sbs>'
IDENTITY = "1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1 "


def _write(tmp_path, geoms, instances):
    d = tmp_path / "scene"
    (d / "data").mkdir(parents=True, exist_ok=True)
    shutil.copy(FIXTURE, d / "data" / "test_10_sbs.bin")
    path = d / "scene.xml"
    path.write_text(SCENE.format(geoms=geoms, instances=instances))
    return str(path)


def _instance(geom=0, matrix=IDENTITY):
    return f'<instance id="0" mesh_id="{geom}" rmap_id="0" scn_id="0" scn_sid="0" matrix="{matrix}" />'


def test_load_scene_renders_like_set_scene(tmp_path):
    path = _write(tmp_path, SBS_NODE.format(id=0), _instance())
    W, H = 160, 120
    a = L.CreateMultiRenderer(L.DEVICE_GPU)
    a.SetViewport(0, 0, W, H)
    a.LoadScene(path)
    ga = a.RenderGBuffer(W, H)
    b = L.CreateMultiRenderer(L.DEVICE_GPU)
    b.SetViewport(0, 0, W, H)
    b.SetScene(L.load_sbs(FIXTURE))
    wv, pr = L.look_at_camera((0.4, 0.3, 2.8), (0, 0, 0), (0, 1, 0), 50.0, W / H, 0.01, 100.0)
    b.UpdateCamera(wv, pr)
    gb = b.RenderGBuffer(W, H)
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    assert np.count_nonzero(ga[2] != 0xFFFFFFFF) > 1000
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["svs_node", "two_instances", "moved", "missing_geometry", "malformed"])
def test_load_scene_refuses_what_this_path_does_not_render(tmp_path, case):
    if case == "svs_node":
        path = _write(tmp_path, '<sdf_svs id="0" name="t" type="sdf_svs" loc="data/test_10_sbs.bin" />', _instance())
    elif case == "two_instances":
        path = _write(tmp_path, SBS_NODE.format(id=0), _instance() + _instance())
    elif case == "moved":
        path = _write(tmp_path, SBS_NODE.format(id=0), _instance(matrix="1 0 0 0.5 0 1 0 0 0 0 1 0 0 0 0 1"))
    elif case == "missing_geometry":
        path = _write(tmp_path, SBS_NODE.format(id=0), _instance(geom=3))
    else:
        path = _write(tmp_path, SBS_NODE.format(id=0) + "<broken", _instance())
    r = L.CreateMultiRenderer(L.DEVICE_GPU)
    with pytest.raises(L.LrtError):
        r.LoadScene(path)
    r.close()

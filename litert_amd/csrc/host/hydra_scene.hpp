// hydra_scene.hpp - the parts of a Hydra XML scene that
// MultiRenderer::LoadSceneHydra reads (Renderer/eye_ray_host.cpp:67-121): the
// first camera of cam_lib, the geometry nodes of geometry_lib (type = tag
// name, file = loc relative to the scene's folder) and the geometry instances
// of the first scene. A small XML reader of its own (elements, attributes,
// text, comments, declarations, the five predefined entities).
#pragma once
#include <string>
#include <vector>

namespace lrt {

struct HydraCamera {
  float fov = 45.0f, znear = 0.01f, zfar = 100.0f;
  float pos[3] = {0, 0, 1}, look_at[3] = {0, 0, 0}, up[3] = {0, 1, 0};
};

struct HydraGeom {
  std::string type;  // tag name: mesh, sdf_sbs, sdf_svs, ...
  std::string path;  // loc, joined to the scene's folder
  int id = -1;
};

struct HydraInstance {
  int geom_id = -1;
  float matrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // as written (row-major)
};

struct HydraScene {
  bool has_camera = false;
  HydraCamera camera;
  std::vector<HydraGeom> geoms;
  std::vector<HydraInstance> instances;
};

// Parses the scene file at `path`; throws std::runtime_error with a message on
// unreadable files or malformed XML.
HydraScene load_hydra_scene(const std::string& path);

}  // namespace lrt

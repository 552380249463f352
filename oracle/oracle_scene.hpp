// oracle_scene.hpp - the opaque orc_scene of the C API (TEST INFRASTRUCTURE).
#pragma once
#include "oracle_render.hpp"
struct orc_scene {
  orc::Scene s;
};

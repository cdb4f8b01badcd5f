// Host-side scene ingestion: shb_scene_desc -> the flat arrays the kernels
// read (world transforms, shading records, lights, area-light triangles).
// Free of HIP runtime calls so CPU checks can link it.
#ifndef SHB_SCENE_HOST_H_
#define SHB_SCENE_HOST_H_

#include <string>
#include <vector>

#include "../../include/shb.h"
#include "shb_common.h"

namespace shb {

// Host copy of the parts of the scene the kernels need, kept so the BVH can
// be rebuilt with another width and the raster inputs re-made.
struct HostScene {
  uint32_t num_vertices = 0, num_tris = 0;
  std::vector<float> world_xyz;      // 9 per triangle (BVH input)
  std::vector<float> wpos, wnrm, wtan, luv;  // per vertex (raster)
  std::vector<uint32_t> tri_vidx;    // 3 global vertex ids per triangle
  std::vector<int32_t> tri_mat;
  std::vector<float4> shade;         // 4 per triangle
  std::vector<int4> tri_info;
  std::vector<MaterialDev> mats;
  std::vector<uint8_t> texels;
  std::vector<LightDev> lights;
  std::vector<AreaTri> area_tris;
  std::vector<uint32_t> area_tri_global;  // global triangle id of each area_tris entry
  bool raster_ready = true;          // every geometry has luv, normals, tangents
  bool has_transparency = false;     // some RGBA texel with alpha < 255
  bool bad_material = false;         // a geometry's material_id is not a material
};

// Validates and converts the scene; returns SHB_OK or an SHB_ERR_* code with
// a message in *err.
int build_host_scene(const shb_scene_desc* d, HostScene* out, std::string* err);

// SRGBToLinear (reference src/colorspace.h:10-14).
float srgb_to_linear(uint8_t x);

}  // namespace shb

#endif  // SHB_SCENE_HOST_H_

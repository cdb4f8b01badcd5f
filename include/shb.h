/*
 * shb.h — C ABI of the MI355X SH-lightmap baker (libshb_hip.so).
 *
 * Plain C: POD structs, plain pointers and sizes, integer status codes.
 * Every entry point returns SHB_OK (0) or a negative SHB_ERR_* code and never
 * aborts across the boundary; the message of the last failure on a context is
 * returned by shb_last_error().
 *
 * Threading: one context per host thread. A context is not safe to use from
 * several threads at once. All GPU work of a context runs on its own stream.
 * Results are deterministic bit for bit for a given (scene, surface, config,
 * seed, BVH width), whatever the batch size, launch shapes or job packing.
 * Another BVH width builds another tree, which can change the hit of a ray
 * whose t lies within rounding of a box bound (3 of 68.6 M extension and 2
 * of 80.2 M shadow rays on a 262k-triangle room, no lightmap bit changed).
 *
 * Each entry point names the reference interface it replaces
 * (DaviesX/sh-baker, paths relative to the repository root). Host buffers are
 * caller-allocated; the library copies inputs and keeps no host pointers.
 */
#ifndef SHB_H_
#define SHB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  SHB_OK = 0,
  SHB_ERR_INVALID_ARG = -1,  /* bad pointer, size or config value          */
  SHB_ERR_NO_SCENE = -2,     /* no scene uploaded on this context           */
  SHB_ERR_DEVICE = -3,       /* HIP runtime error                           */
  SHB_ERR_OOM = -4,          /* device or host allocation failed            */
  SHB_ERR_LIMIT = -5,        /* a documented size limit was exceeded        */
  SHB_ERR_NO_SURFACE = -6,   /* bake called before rasterize / set_surface  */
  SHB_ERR_UNSUPPORTED = -7   /* a reference feature this build rejects      */
};

/* ----------------------------------------------------------------- scene */
enum {
  SHB_LIGHT_POINT = 0,
  SHB_LIGHT_DIRECTIONAL = 1,
  SHB_LIGHT_SPOT = 2,
  SHB_LIGHT_AREA = 3
};

enum {
  SHB_GEOM_HAS_NORMALS = 1u << 0,
  SHB_GEOM_HAS_TEXTURE_UVS = 1u << 1,
  SHB_GEOM_HAS_LIGHTMAP_UVS = 1u << 2,
  SHB_GEOM_HAS_TANGENTS = 1u << 3
};

/* One reference `Geometry` (src/scene.h:56-67). Vertex attributes live in the
 * flat arrays of shb_scene_desc at [vertex_offset, vertex_offset+vertex_count);
 * indices at [index_offset, index_offset+index_count) are local to the
 * geometry, three per triangle. */
typedef struct shb_geometry {
  uint32_t vertex_offset;
  uint32_t vertex_count;
  uint32_t index_offset;
  uint32_t index_count;
  int32_t material_id;  /* index into materials (-1: none); the bake needs a
                           valid one, raster and ray queries pass it through */
  uint32_t flags;       /* SHB_GEOM_HAS_*                                    */
  float transform[12];  /* row-major 3x4 affine (object -> world)            */
} shb_geometry;

/* One reference `Material` (src/scene.h:43-53): the albedo texture (8-bit
 * sRGB, row-major, 3 or 4 channels) and the emission intensity. */
typedef struct shb_material {
  uint32_t albedo_width;
  uint32_t albedo_height;
  uint32_t albedo_channels; /* 0: no texture (albedo 0.8 grey), 3 or 4      */
  uint32_t reserved;
  uint64_t albedo_offset;   /* byte offset into texel_pool                  */
  float emission_intensity;
  float reserved2;
} shb_material;

/* One reference `Light` (src/scene.h:70-87). Area lights name the geometry
 * whose triangles emit and the material that gives their emission. */
typedef struct shb_light {
  int32_t type;            /* SHB_LIGHT_*                                    */
  float position[3];
  float direction[3];
  float color[3];
  float intensity;
  float cos_inner_cone;
  float cos_outer_cone;
  int32_t geometry_index;  /* area lights only, else -1                      */
  int32_t material_index;  /* area lights only, else -1                      */
} shb_light;

/* The whole reference `Scene` (src/scene.h:110-115) as flat arrays.
 * Optional per-vertex arrays may be NULL when no geometry sets the flag. */
typedef struct shb_scene_desc {
  uint32_t num_vertices;
  const float* positions;     /* 3 * num_vertices, object space */
  const float* normals;       /* 3 * num_vertices or NULL       */
  const float* texture_uvs;   /* 2 * num_vertices or NULL       */
  const float* lightmap_uvs;  /* 2 * num_vertices or NULL       */
  const float* tangents;      /* 4 * num_vertices or NULL       */
  uint32_t num_indices;
  const uint32_t* indices;
  uint32_t num_geometries;
  const shb_geometry* geometries;
  uint32_t num_materials;
  const shb_material* materials;
  uint64_t texel_pool_size;
  const uint8_t* texel_pool;
  uint32_t num_lights;
  const shb_light* lights;
} shb_scene_desc;

/* ------------------------------------------------------ raster / surface */
/* Reference `RasterConfig` (src/rasterizer.h:12-16). */
typedef struct shb_raster_cfg {
  int32_t width;
  int32_t height;
  int32_t supersample_scale;
} shb_raster_cfg;

/* Reference `SurfacePoint` (src/rasterizer.h:18-23), packed to 48 bytes. */
typedef struct shb_surface_point {
  float position[3];
  float normal[3];
  float tangent[4];
  int32_t material_id;  /* < 0: texel not covered */
  int32_t reserved;
} shb_surface_point;

/* ------------------------------------------------------------------ bake */
enum {
  SHB_BAKE_COUNT_VISITS = 1u << 0, /* also count BVH node visits / triangle
                                      tests (slower; for measurement only) */
  SHB_BAKE_USE_GRAPH = 1u << 1,    /* capture the launch sequence in a hipGraph
                                      and replay it while the plan, buffers and
                                      tuning are unchanged (per-launch trace
                                      timings are then not reported)          */
  SHB_BAKE_TRACE_ALL_SHADOW = 1u << 2 /* trace every shadow query, also those
                                      the emitter test would answer (see
                                      rays_shadow_resolved): the same results,
                                      for checks and measurement              */
};

/* Reference `BakeConfig` (src/baker.h:10-14) plus the seed of the
 * counter-based random stream. */
typedef struct shb_bake_cfg {
  int32_t samples;           /* rays per texel                              */
  int32_t bounces;           /* max path depth                              */
  int32_t num_light_samples; /* light samples per NEE, 1..16 (the reference
                                CLI always uses 1; the adjoint takes 1)      */
  uint32_t flags;            /* SHB_BAKE_*                                  */
  uint64_t seed;
} shb_bake_cfg;

typedef struct shb_bake_stats {
  uint64_t valid_texels;    /* texels with material_id >= 0                 */
  uint64_t paths;           /* valid_texels * samples                       */
  uint64_t rays_extension;  /* closest-hit queries traced                   */
  uint64_t rays_shadow;     /* any-hit queries (every shadow ray of the
                               reference's FindOcclusion calls), including
                               rays_shadow_resolved                         */
  uint64_t node_visits;     /* only with SHB_BAKE_COUNT_VISITS              */
  uint64_t tri_tests;       /* only with SHB_BAKE_COUNT_VISITS              */
  double seconds_total;     /* wall time of the call                        */
  double seconds_device;    /* GPU time between the first and last kernel   */
  double seconds_trace;     /* GPU time inside the traversal kernels        */
  /* the shadow (any-hit) part of node_visits, tri_tests and seconds_trace  */
  uint64_t node_visits_shadow;
  uint64_t tri_tests_shadow;
  double seconds_trace_shadow;
  uint32_t launches_extension;  /* traversal kernel launches of each kind   */
  uint32_t launches_shadow;
  /* alpha < 1 splits (baker.cpp:97-109) whose transmitted ray took a spare
   * path slot; bake passes run (a pass that runs out of spare slots is
   * redone with more; the result does not depend on the count)            */
  uint64_t branches;
  uint32_t passes;
  uint32_t reserved;
  /* shadow queries of area lights answered without traversal: the ray
   * reaches past the sampled emitter point (tfar = |d| - 0.001 with t in
   * units of |d|, light.h:176-188), and one watertight test accepts the
   * emitter's own triangle well inside the ray's range, so the any-hit query
   * is occluded; rays_shadow - rays_shadow_resolved rays were traced       */
  uint64_t rays_shadow_resolved;
} shb_bake_stats;

/* Launch parameters of the wavefront kernels (chosen by shb_autotune).
 * A new context starts with block_size 256, batch_paths 2^25, queue_chunk
 * 12 and bvh_width 8. Results do not depend on the first three.        */
typedef struct shb_tuning {
  int32_t block_size;       /* threads per block of the traversal kernels   */
  int32_t batch_paths;      /* path slots resident per wavefront batch      */
  int32_t queue_chunk;      /* 0: one ray per thread; 1..64: persistent
                               traversal, a wave refills once this many of
                               its lanes are idle                          */
  int32_t bvh_width;        /* children per BVH node, 2..8 (rebuilds)       */
} shb_tuning;

/* ------------------------------------------------------ fixed-ray queries */
typedef struct shb_ray {
  float origin[3];
  float tnear;
  float direction[3];  /* need not be unit length; t is in units of |d| */
  float tfar;
} shb_ray;

/* Closest hit in the reference's convention (src/occlusion.cpp:9-107):
 * prim_id < 0 means miss. u, v are Embree barycentrics (weights of vertices 1
 * and 2); position = o + d*t; normal is the interpolated object-space vertex
 * normal (the world geometric normal without vertex normals); uv the texture
 * uv. Equal-t candidates resolve to the lowest prim_id. */
typedef struct shb_hit {
  float t;
  float u;
  float v;
  int32_t prim_id;      /* global triangle index (geometry-major order)    */
  int32_t geom_id;
  int32_t material_id;
  float position[3];
  float normal[3];
  float uv[2];
} shb_hit;

/* ------------------------------------------------------------- lifecycle */
typedef struct shb_ctx shb_ctx;

/* Creates a context on HIP device `device` with its own stream. */
int shb_create(shb_ctx** out, int device);
void shb_destroy(shb_ctx* ctx);
/* Message of the last failure on ctx ("" if none); valid until the next call. */
const char* shb_last_error(const shb_ctx* ctx);
const char* shb_version(void);

/* Uploads the scene and builds the BVH on the host (binned SAH, collapsed to
 * the compressed 8-wide node layout). SHB_ERR_LIMIT when the BVH would be
 * deeper than the traversal stack or hold 2^24 nodes or more (about 160 M
 * triangles); a batched bake's scenes hold fewer than 2^24 nodes together.
 * Replaces BuildBVH / ReleaseBVH (src/scene.cpp:209-264, src/scene.h:129-132)
 * and the scene arguments of BakeSHLightMap (src/baker.h:25-28). */
int shb_upload_scene(shb_ctx* ctx, const shb_scene_desc* scene);

/* Rasterizes the lightmap uvs into (width*ss) x (height*ss) surface points,
 * kept on the device for shb_bake; copies them to `out` when not NULL.
 * Replaces RasterizeScene (src/rasterizer.cpp:68-211). */
int shb_rasterize(shb_ctx* ctx, const shb_raster_cfg* cfg,
                  shb_surface_point* out);

/* Uploads caller-made surface points (the `surface_points` argument of
 * BakeSHLightMap, src/baker.h:25-28); count = width*ss * height*ss. */
int shb_set_surface(shb_ctx* ctx, const shb_raster_cfg* cfg,
                    const shb_surface_point* points);

/* Validity mask of the current surface: 1 where material_id >= 0.
 * Replaces CreateValidityMask (src/rasterizer.cpp:213-220). */
int shb_validity_mask(shb_ctx* ctx, uint8_t* mask_out);

/* Bakes the SH9 lightmap at the supersampled resolution:
 * sh_out[texel*27 + coeff*3 + channel], vis_out[texel] (either may be NULL).
 * Replaces BakeSHLightMap (src/baker.cpp:186-304). */
int shb_bake(shb_ctx* ctx, const shb_bake_cfg* cfg, float* sh_out,
             float* vis_out, shb_bake_stats* stats);

/* Bake followed by the post-bake steps of the reference CLI, all on the
 * device: validity mask, ss x ss downsample of SH, visibility and mask,
 * then `dilation` passes on SH and on visibility with the mask the SH pass
 * grew. Outputs are at the base width x height.
 * Replaces main.cpp:131-163 (BakeSHLightMap, CreateValidityMask,
 * DownsampleSHTexture, DownsampleEnvironmentVisibilityTexture,
 * DownsampleValidityMask, Dilate x2). */
int shb_bake_lightmap(shb_ctx* ctx, const shb_bake_cfg* cfg,
                      int32_t dilation, float* sh_out, float* vis_out,
                      uint8_t* mask_out, shb_bake_stats* stats);

/* Copies the result of the last shb_bake_lightmap of this context to host
 * memory (any pointer may be NULL). With NULL outputs, shb_bake_lightmap
 * leaves its result in device memory only (no PCIe transfer); this call
 * fetches it later. Valid until the next bake, rasterize or set_surface.
 * No reference counterpart: the reference returns host vectors
 * (main.cpp:131-163). */
int shb_read_lightmap(shb_ctx* ctx, float* sh_out, float* vis_out,
                      uint8_t* mask_out);

/* ss x ss box average of a (width*scale) x (height*scale) image of
 * `channels` floats per texel into width x height. Replaces DownsampleTexture
 * / DownsampleSHTexture / DownsampleEnvironmentVisibilityTexture
 * (src/baker.cpp:151-182, 306-323). */
int shb_downsample(shb_ctx* ctx, int32_t width, int32_t height, int32_t scale,
                   int32_t channels, const float* in, float* out);

/* Any-valid downsample of a mask. Replaces DownsampleValidityMask
 * (src/rasterizer.cpp:222-241). */
int shb_downsample_mask(shb_ctx* ctx, int32_t width, int32_t height,
                        int32_t scale, const uint8_t* in, uint8_t* out);

/* In-place dilation of `channels` floats per texel and of the mask.
 * Replaces Dilate<T> (src/dilation.cpp:7-73). */
int shb_dilate(shb_ctx* ctx, int32_t width, int32_t height, int32_t channels,
               float* data, uint8_t* mask, int32_t passes);

/* Traces a list of rays against the uploaded scene (closest hit, or any hit
 * when any_hit != 0: then only prim_id >= 0 / < 0 is meaningful).
 * Replaces FindOcclusion (src/occlusion.cpp:9-107). node_visits / tri_tests,
 * when not NULL, receive per-ray traversal counts. */
int shb_trace_rays(shb_ctx* ctx, const shb_ray* rays, uint32_t n,
                   int32_t any_hit, shb_hit* hits, uint32_t* node_visits,
                   uint32_t* tri_tests);

/* The flattened BVH the traversal kernels run on (no reference counterpart;
 * Embree's tree is internal to the reference): 80-byte compressed 8-wide
 * nodes (root first) and 48-byte triangles in leaf order (3 x float4,
 * v0.w = global triangle id as a bit pattern). With NULL buffers only the
 * counts are written; otherwise *num_nodes / *num_tris give the buffers'
 * capacities on entry and the counts on return. For host recounts of the
 * per-ray node visits and triangle tests behind the traversal roofline. */
int shb_bvh_export(shb_ctx* ctx, void* nodes, uint32_t* num_nodes, void* tris,
                   uint32_t* num_tris);

/* ---------------------------------------------------------- glTF loader */
/* Environment kinds of a loaded scene (src/scene.h:90-107). */
enum {
  SHB_ENV_NONE = 0,      /* no skybox and no directional light           */
  SHB_ENV_TEXTURE = 1,   /* extras.skybox: an RGB float LatLong map       */
  SHB_ENV_PREETHAM = 2   /* Preetham sky towards the brightest sun        */
};

typedef struct shb_environment_desc {
  int32_t type;              /* SHB_ENV_*                                   */
  int32_t width, height;     /* texture size (SHB_ENV_TEXTURE)              */
  int32_t channels;          /* 3 for a texture, else 0                     */
  const float* pixels;       /* width * height * 3 floats, or NULL          */
  float sun_direction[3];    /* towards the sun (SHB_ENV_PREETHAM)          */
  float sun_intensity;
  const char* path;          /* absolute path of the skybox file, or NULL   */
} shb_environment_desc;

/* A glTF scene loaded on the host (owns every array its description
 * points to). */
typedef struct shb_gltf shb_gltf;

/* Loads a .gltf / .glb file. Mesh primitives need a material, POSITION,
 * NORMAL and TEXCOORD_0; TEXCOORD_1 becomes the lightmap UVs (else they are
 * zero and SHB_GEOM_HAS_LIGHTMAP_UVS is clear); emissive materials
 * (KHR_materials_emissive_strength) give one area light per primitive;
 * KHR_lights_punctual lights are placed by their nodes; albedo textures are
 * RGBA8 (PNG or baseline JPEG) or a 1x1 sRGB texture of baseColorFactor.
 * On failure *out is NULL and err_out (err_len bytes, may be NULL) holds
 * the reason. Replaces LoadScene (src/loader.cpp:113-844). */
int shb_load_gltf(const char* path, shb_gltf** out, char* err_out, size_t err_len);
/* The scene as a description for shb_upload_scene (valid until free). */
const shb_scene_desc* shb_gltf_scene(const shb_gltf* scene);
/* The environment (src/loader.cpp:656-789): skybox texture or sun. */
int shb_gltf_environment(const shb_gltf* scene, shb_environment_desc* out);
/* Material name and absolute albedo file path (NULL when none). */
const char* shb_gltf_material_name(const shb_gltf* scene, uint32_t material);
const char* shb_gltf_albedo_path(const shb_gltf* scene, uint32_t material);
void shb_gltf_free(shb_gltf* scene);

/* ---------------------------------------------------------------- atlas */
/* Lightmap UVs for a scene without them. */
typedef struct shb_atlas shb_atlas;

/* Per-geometry relative scale: density_multiplier * sqrt(albedo width *
 * height) * sqrt(texture-UV tile count), clamped to 5x the median; out
 * holds num_geometries floats. Replaces atlas_internal::
 * CalculateGeometryScales (src/atlas.cpp:22-95). */
int shb_atlas_geometry_scales(const shb_scene_desc* scene, float density_multiplier,
                              float* out);

/* Charts the scaled world-space meshes (faces grown while normals stay
 * within 25 degrees; orthographic projection, minimum-area rectangle),
 * packs them into one resolution^2 atlas with `padding` texels between
 * charts at the largest density that fits, and returns the scene with
 * vertices split per chart and SHB_GEOM_HAS_LIGHTMAP_UVS set (materials,
 * lights, transforms and face order kept). SHB_ERR_LIMIT when the charts
 * cannot fit. Replaces CreateAtlasGeometries (src/atlas.cpp:97-250), whose
 * charting and packing come from xatlas. */
int shb_atlas_create(const shb_scene_desc* scene, uint32_t resolution, uint32_t padding,
                     float density_multiplier, shb_atlas** out, char* err_out,
                     size_t err_len);
const shb_scene_desc* shb_atlas_scene(const shb_atlas* atlas);
/* Atlas size, chart count, texels per scaled world unit, and per output
 * vertex the geometry-local index of its source vertex and its chart
 * (any may be NULL). */
int shb_atlas_info(const shb_atlas* atlas, int32_t* width, int32_t* height, uint32_t* charts,
                   float* texels_per_unit, const uint32_t** xref, const uint32_t** chart);
void shb_atlas_free(shb_atlas* atlas);

/* ------------------------------------------------------------ EXR output */
/* Save layouts of the reference's SaveMode (src/saver.h:19-24). */
enum {
  SHB_SAVE_COMBINED = 0,         /* one file, channels L0.R .. L22.B + EnvVisibility */
  SHB_SAVE_SPLIT = 1,            /* path_<coeff>.exr x 9, RGB (L0: RGBA, A = vis)   */
  SHB_SAVE_LUMINANCE_PACKED = 2  /* path_packed_{0,1,2}.exr, RGBA: L0 + vis, luma of
                                    L1m1 L10 L11 L2m2, luma of L2m1 L20 L21 L22     */
};

/* Writes an SH lightmap (sh[texel*27 + coeff*3 + channel], vis[texel] or
 * NULL) of width x height texels as OpenEXR, HALF channels, ZIP
 * compression. Host only (needs no context). On failure a message goes to
 * err_out (err_len bytes, may be NULL). Replaces SaveSHLightMap
 * (src/saver.cpp:103-422). */
int shb_save_lightmap(const char* path, int32_t width, int32_t height,
                      const float* sh, const float* vis, int32_t mode,
                      char* err_out, size_t err_len);

/* ------------------------------------------------------- environment SH */
/* SH9 (sh_out[coeff*3 + channel]) of an equirectangular radiance map of
 * width x height texels with `channels` >= 3 floats per texel (the first
 * three are RGB), +Y up, phi from +Z towards +X, rows with sin(theta) <
 * 1e-4 skipped. Runs on the context's device; pixels == NULL reuses the map
 * uploaded by the previous call (same size). seconds_device (may be NULL)
 * receives the kernels' time. Replaces ProjectEnvironmentToSH for
 * Environment::Type::Texture, i.e. ProjectLatLongMap (src/scene.cpp:31-71,
 * 152-158); sums are FP64 in a fixed order instead of a serial float sum. */
int shb_env_sh_latlong(shb_ctx* ctx, int32_t width, int32_t height,
                       int32_t channels, const float* pixels, float* sh_out,
                       double* seconds_device);

/* shb_env_sh_latlong with no allocation: the map is already in device
 * memory (pixels_device, width * height * channels floats) and the 27 SH
 * floats go to device memory (sh_out_device); both must be device memory of
 * ctx's GPU. The kernels' per-block partial sums live in a fixed 2 MiB array
 * of the library, one per GPU, held until the process ends, so the call
 * makes no hipMalloc / hipFree: maps needing more than 9708 blocks of
 * partial sums (above about 16384 x 8192 texels) are rejected with
 * SHB_ERR_LIMIT. Same kernels and summation order as shb_env_sh_latlong,
 * so the result is the same bit for bit. Runs on ctx's stream and returns
 * once it has drained. Calls that overlap on one GPU (from several host
 * threads or contexts) share the array and race.
 * No reference counterpart: for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops). */
int shb_env_sh_latlong_noalloc(shb_ctx* ctx, int32_t width, int32_t height,
                               int32_t channels, const float* pixels_device,
                               float* sh_out_device);

/* SH9 of the clamped Preetham/Perez clear sky for a sun direction (any
 * length; normalised first): 64 x 128 midpoint quadrature of the upper
 * hemisphere, radiance (0.2, 0.5, 0.9) * min(F, 30 F_zenith) / F_zenith.
 * Host only (needs no context). Replaces ProjectEnvironmentToSH for
 * Environment::Type::Preetham, i.e. ProjectPreethamToSH
 * (src/scene.cpp:73-150). */
int shb_env_sh_preetham(const float sun_direction[3], float* sh_out);

/* ------------------------------------------------ batched (no reference) */
/* Bakes several contexts (each with its own scene and surface; all on the
 * same device) in one persistent launch per bounce over a shared ray queue,
 * with the bounce loop captured in a hipGraph. sh_outs[i] / vis_outs[i] as in
 * shb_bake. No reference counterpart: the reference bakes one lightmap per
 * process (src/main.cpp:40-193). */
int shb_bake_batched(shb_ctx* const* ctxs, int32_t num_jobs,
                     const shb_bake_cfg* cfgs, float* const* sh_outs,
                     float* const* vis_outs, shb_bake_stats* stats);

/* A tile of a tiled bake: supersampled texels [x0, x0 + w) x [y0, y0 + h). */
typedef struct shb_tile {
  int32_t x0, y0, w, h;
} shb_tile;

/* Bakes the current surface of ctx as num_tiles (at most 64) tile jobs in
 * one batched run: the tiles share the context's scene and BVH and the
 * wavefront batches, and the launch sequence is captured in a hipGraph.
 * Tiles must lie inside the supersampled surface and must not overlap;
 * texels outside every tile stay 0. A texel in a tile gets the same bits as
 * from shb_bake (same random stream). sh_out / vis_out as in shb_bake (NULL
 * keeps that output on the device only). No reference counterpart: the
 * reference bakes one whole lightmap per call (src/baker.cpp:186-304). */
int shb_bake_tiled(shb_ctx* ctx, const shb_bake_cfg* cfg, const shb_tile* tiles,
                   int32_t num_tiles, float* sh_out, float* vis_out,
                   shb_bake_stats* stats);

/* Chooses block size, batch size, refill threshold and BVH width by timing
 * short bakes of the current scene and surface; applies and returns them. */
int shb_autotune(shb_ctx* ctx, const shb_bake_cfg* cfg, shb_tuning* best);
int shb_set_tuning(shb_ctx* ctx, const shb_tuning* tuning);
int shb_get_tuning(const shb_ctx* ctx, shb_tuning* tuning);

/* ------------------------------------------- differentiation (no reference) */
/* Per-light RGB radiance scales and per-material RGB albedo scales (all 1 by
 * default, which reproduces the reference bake; finite and >= 0). Either
 * pointer may be NULL. */
int shb_set_params(shb_ctx* ctx, const float* light_scale,
                   const float* albedo_scale);

/* Path-replay adjoint: given dL/dSH at the supersampled bake resolution
 * (same layout as shb_bake's sh_out), writes dL/d(light_scale) (3 per light)
 * and dL/d(albedo_scale) (3 per material) of shb_bake with the same cfg.
 * The bake's light selection uses radiance without these scales, so the
 * selection law does not depend on them; every path is re-traced from its
 * counter-based random numbers (memory O(texels + paths in flight)) and the
 * result is the exact derivative of the estimator for those numbers. Every
 * scale must be > 0. No reference counterpart. */
int shb_bake_adjoint(shb_ctx* ctx, const shb_bake_cfg* cfg,
                     const float* dloss_dsh, float* dloss_dlight,
                     float* dloss_dalbedo, shb_bake_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SHB_H_ */

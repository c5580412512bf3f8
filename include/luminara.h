/*
 * luminara.h - C ABI of the MI355X path tracer (render hot path of luminara).
 *
 * The reference renderer has no FFI; its seam is the Rust function
 *   render::render(&RenderConfig, &Camera, &dyn Hittable, &[LightInfo]) -> RenderResult
 *   (src/render.rs:1557-1562, called from src/main.rs:744)
 * whose pixel loop (src/render.rs:1587-1747) produces `rows: Vec<Vec<Color>>`, a
 * row-major linear-HDR framebuffer (row 0 = top) that the CPU post chain consumes.
 * The scene reaching it comes from scene::load_scene (src/scene.rs:846-1645).
 *
 * Every entry point below returns an int status (LMN_OK = 0) and never throws;
 * the message of the last failure on the calling thread is in lmn_last_error().
 * A Rust host binds these 1:1 with an `extern "C"` block.
 *
 * Ownership: the library owns scene handles and all device memory; callers own
 * every host buffer they pass in.  Descriptor contents are copied at create time.
 * A scene handle is externally synchronised (one host thread at a time).
 */
#ifndef LUMINARA_H
#define LUMINARA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LMN_OK 0
#define LMN_ERR_INVALID 1     /* bad argument / inconsistent descriptor      */
#define LMN_ERR_PARSE 2       /* scene text does not parse (TOML, PLY, OBJ)  */
#define LMN_ERR_IO 3          /* a referenced file cannot be read            */
#define LMN_ERR_UNSUPPORTED 4 /* valid reference scene feature not on path   */
#define LMN_ERR_DEVICE 5      /* HIP runtime error                           */
#define LMN_ERR_NOMEM 6       /* host or device allocation failed            */

/* ---- primitives (reference world objects, flattened) -------------------- */
/* One lmn_prim per hit() leaf of the reference.  A [[box]] contributes six
 * rect prims (src/rect.rs:128-141); a [[mesh]] one triangle prim per face
 * (src/obj.rs:212-240).  `object` is the index of the top-level world-list
 * entry the prim belongs to, in the reference's load order
 * (src/scene.rs:1180-1642); prims of a transformed box share one object.
 * Parameter layout p[] per kind (all FP64, world or object space as given
 * in the scene file):
 *   SPHERE          center[3], radius                      src/sphere.rs:24
 *   QUAD            q[3], u[3], v[3]                        src/quad.rs:20
 *   RECT_XY         x0, x1, y0, y1, k                       src/rect.rs:24
 *   RECT_XZ         x0, x1, z0, z1, k                       src/rect.rs:64
 *   RECT_YZ         y0, y1, z0, z1, k                       src/rect.rs:104
 *   TRIANGLE        v0[3], v1[3], v2[3]   (Triangle and MeshTriangle without
 *                   per-vertex normals/uvs)   src/triangle.rs:23, src/obj.rs:260
 *   PLANE           point[3], normal[3] (as given)          src/plane.rs:18
 *   TORUS           center[3], major_r, minor_r             src/torus.rs:24
 *   SUPERELLIPSOID  center[3], scale[3], e1, e2             src/superellipsoid.rs:26
 *   SPRING          center[3], coil_r, tube_r, pitch, turns src/spring.rs:35
 *   DISK            center[3], normal[3] (as given), radius src/disk.rs:16
 *   CYLINDER        center[3], radius, y_min, y_max         src/cylinder.rs:17
 *   CONE            center[3], radius, y_min, y_max (apex)  src/cone.rs:18
 *   ELLIPSOID       center[3], radii[3]                     src/ellipsoid.rs:14
 *   HEMISPHERE      center[3], radius (the half y >= c.y)   src/hemisphere.rs:14
 *   ANNULUS         center[3], normal[3] (as given), inner_r, outer_r
 *                                                           src/annulus.rs:17
 *   FACET           v0[3], v1[3], v2[3], normal[3] (unit, outward as the
 *                   reference orients it): one triangle of a prism or pyramid
 *                                   src/prism.rs:40-126, src/pyramid.rs:29-97
 *   WEDGE_FACE      min[3], max[3], face (0 bottom, 1 back, 2 left, 3 right,
 *                   4 ramp)                                 src/wedge.rs:22-120
 *   ROUNDED_BOX     center[3], half_size[3], radius         src/rounded_box.rs:10
 *   MOBIUS          center[3], radius, width, thickness     src/mobius.rs:11
 *   MESH_TRIANGLE   v0[3], v1[3], v2[3], attr: a MeshTriangle with per-vertex
 *                   normals and/or texture coordinates (an OBJ face with vn /
 *                   vt indices); p[9] = index into lmn_scene_desc.vertex_attrs
 *                                                           src/obj.rs:250-313
 * A capsule (src/capsule.rs:15-53) is three prims of one object: CYLINDER
 * (y from center.y to center.y + height), then the bottom and top SPHERE.
 * A prism is eight FACETs of one object (bottom, top, then two per side), a
 * pyramid six (two for the base, four sides), a wedge five WEDGE_FACEs. */
enum {
  LMN_PRIM_SPHERE = 0,
  LMN_PRIM_QUAD = 1,
  LMN_PRIM_RECT_XY = 2,
  LMN_PRIM_RECT_XZ = 3,
  LMN_PRIM_RECT_YZ = 4,
  LMN_PRIM_TRIANGLE = 5,
  LMN_PRIM_PLANE = 6,
  LMN_PRIM_TORUS = 7,
  LMN_PRIM_SUPERELLIPSOID = 8,
  LMN_PRIM_SPRING = 9,
  LMN_PRIM_DISK = 10,
  LMN_PRIM_CYLINDER = 11,
  LMN_PRIM_CONE = 12,
  LMN_PRIM_ELLIPSOID = 13,
  LMN_PRIM_HEMISPHERE = 14,
  LMN_PRIM_ANNULUS = 15,
  LMN_PRIM_FACET = 16,
  LMN_PRIM_WEDGE_FACE = 17,
  LMN_PRIM_ROUNDED_BOX = 18,
  LMN_PRIM_MOBIUS = 19,
  LMN_PRIM_MESH_TRIANGLE = 20,
  LMN_PRIM_KIND_COUNT = 21
};

typedef struct lmn_prim {
  uint32_t kind;
  uint32_t material;  /* index into lmn_scene_desc.materials          */
  int32_t transform;  /* index into lmn_scene_desc.transforms, or -1  */
  uint32_t object;    /* top-level world-list index (load order)      */
  double p[12];
} lmn_prim;

/* ---- instance transforms (src/transform.rs) ----------------------------- */
/* Ops are listed innermost first, exactly as the loader wraps them:
 * boxes   Scale|ScaleXYZ -> RotX -> RotY -> RotZ -> Translate (src/scene.rs:1577-1593)
 * spheres RotX -> RotY -> RotZ -> Translate -> Scale        (src/scene.rs:1216-1230) */
enum {
  LMN_XF_ROTATE_X = 1,  /* v[0] = degrees       src/transform.rs:116 */
  LMN_XF_ROTATE_Y = 2,  /* v[0] = degrees       src/transform.rs:39  */
  LMN_XF_ROTATE_Z = 3,  /* v[0] = degrees       src/transform.rs:191 */
  LMN_XF_TRANSLATE = 4, /* v[0..2] = offset     src/transform.rs:12  */
  LMN_XF_SCALE = 5,     /* v[0] = factor        src/transform.rs:265 */
  LMN_XF_SCALE_XYZ = 6  /* v[0..2] = factors    src/transform.rs:293 */
};
#define LMN_MAX_XF_OPS 6
typedef struct lmn_xform_op {
  uint32_t kind;
  uint32_t pad;
  double v[3];
} lmn_xform_op;
typedef struct lmn_transform {
  uint32_t n_ops;
  uint32_t pad;
  lmn_xform_op ops[LMN_MAX_XF_OPS];
} lmn_transform;

/* ---- materials (src/material.rs) ---------------------------------------- */
enum {
  LMN_MAT_LAMBERTIAN = 0, /* color = albedo                src/material.rs:59  */
  LMN_MAT_METAL = 1,      /* color = albedo, fuzz (<=1)    src/material.rs:89  */
  LMN_MAT_DIELECTRIC = 2, /* ior, color = tint, roughness, dispersion  :144     */
  LMN_MAT_EMISSIVE = 3,   /* color, intensity              src/material.rs:236 */
  LMN_MAT_MICROFACET = 4  /* color, roughness, metallic    src/material.rs:322 */
};
typedef struct lmn_material {
  uint32_t kind;
  int32_t texture;   /* image texture index replacing `color`, or -1 */
  double color[3];
  double fuzz;
  double ior;
  double roughness;
  double dispersion;
  double intensity;
  double metallic;
} lmn_material;

/* Image texture: RGB texels, row-major, row 0 = top of the image, values as
 * the sampler returns them (k/255 for 8-bit files, src/texture.rs:463-479).
 * Nearest-neighbour lookup, clamp, v flipped. */
typedef struct lmn_texture {
  uint32_t width;
  uint32_t height;
  const float* rgb; /* width*height*3 */
} lmn_texture;

/* ---- NEE lights (src/render.rs:17-39) ------------------------------------ */
enum { LMN_LIGHT_SPHERE = 0, LMN_LIGHT_RECT = 1, LMN_LIGHT_DISK = 2 };
typedef struct lmn_light {
  uint32_t kind;
  uint32_t material; /* emissive material it was registered from, ~0u if none */
  double origin[3]; /* sphere/disk center, rect corner */
  double u[3];      /* rect edge                        */
  double v[3];      /* rect edge                        */
  double normal[3]; /* rect/disk normal                 */
  double radius;    /* sphere/disk radius               */
  double emission[3];
} lmn_light;

/* ---- background (src/render.rs:43-172) ----------------------------------- */
enum { LMN_BG_SKY = 0, LMN_BG_SOLID = 1, LMN_BG_GRADIENT = 2, LMN_BG_SUN = 3 };
typedef struct lmn_background {
  uint32_t kind;
  uint32_t pad;
  double color0[3]; /* solid color / gradient bottom / sun color */
  double color1[3]; /* gradient top / sky color                   */
  double direction[3];
  double intensity;
} lmn_background;

/* ---- camera (src/camera.rs:48-103) --------------------------------------- */
typedef struct lmn_camera {
  double origin[3];
  double lower_left[3];
  double horizontal[3];
  double vertical[3];
  double u[3];
  double v[3];
  double lens_radius;
} lmn_camera;

/* ---- render settings (src/render.rs:239-521, the hot-path subset) -------- */
enum { LMN_FILTER_BOX = 0, LMN_FILTER_TRIANGLE = 1, LMN_FILTER_GAUSSIAN = 2, LMN_FILTER_MITCHELL = 3 };
typedef struct lmn_render_config {
  uint32_t width, height, samples, max_depth;
  uint64_t seed;
  uint32_t filter;
  uint32_t has_crop;
  uint32_t crop[4]; /* x, y, w, h */
  double exposure;
  uint32_t tone_map;      /* LMN_TONE_*: 0 ACES, 1 Reinhard, 2 Filmic, 3 none */
  uint32_t auto_exposure; /* [render] auto_exposure (src/scene.rs:870-872)      */
  lmn_background background;
  uint32_t adaptive;      /* [render] adaptive (src/scene.rs:914-916)            */
  uint32_t pad;
  double adaptive_threshold; /* [render] adaptive_threshold, default 0.03        */
} lmn_render_config;

/* ---- scene descriptor ---------------------------------------------------- */
/* Per-vertex attributes of a MESH_TRIANGLE (src/obj.rs:250-313).  With
 * normals the outward normal at barycentrics (u, v) is
 * unit(n0 (1 - u - v) + n1 u + n2 v); with uvs the texture coordinates are
 * uv0 (1 - u - v) + uv1 u + uv2 v (otherwise the geometric normal and (u, v)).
 * The loader stores the OBJ normals already made unit, as the reference does. */
typedef struct lmn_vertex_attr {
  double normal[3][3];
  double uv[3][2];
  uint32_t has_normals;
  uint32_t has_uvs;
} lmn_vertex_attr;

typedef struct lmn_scene_desc {
  uint32_t n_prims;
  uint32_t n_transforms;
  uint32_t n_materials;
  uint32_t n_textures;
  uint32_t n_lights;
  uint32_t n_objects;
  const lmn_prim* prims;
  const lmn_transform* transforms;
  const lmn_material* materials;
  const lmn_texture* textures;
  const lmn_light* lights;
  lmn_camera camera;
  lmn_render_config config; /* [render] settings as loaded (CLI applies on top) */
  uint64_t n_vertex_attrs;
  const lmn_vertex_attr* vertex_attrs; /* MESH_TRIANGLE attributes */
} lmn_scene_desc;

/* ---- per-call render parameters ----------------------------------------- */
enum {
  LMN_PATH_WAVEFRONT_FP32 = 0, /* the product path                         */
  LMN_PATH_PLAIN_FP64 = 1,     /* one path per thread, FP64 megakernel     */
  LMN_PATH_PLAIN_FP32 = 2      /* one path per thread, FP32 megakernel     */
};
enum {
  LMN_ISECT_EXACT = 0,    /* FP64 quartic / bracketed roots for marched prims */
  LMN_ISECT_REFERENCE = 1 /* the reference's SDF march semantics, FP64        */
};
typedef struct lmn_render_params {
  uint32_t width, height, spp, max_depth;
  uint64_t seed;
  uint32_t filter;
  uint32_t has_crop;
  uint32_t crop[4];
  uint32_t path;           /* LMN_PATH_*                                   */
  uint32_t intersect_mode; /* LMN_ISECT_*                                  */
  uint32_t scene_key;      /* RNG domain separator (batch index)          */
  uint32_t pool_paths;     /* wavefront path-pool size, 0 = library pick  */
  uint32_t block_size;     /* threads per workgroup, 0 = library pick     */
  uint32_t stack_lds;      /* LDS short-stack entries, 0 = library pick   */
  uint32_t flags;          /* LMN_FLAG_*                                  */
  uint32_t grid_blocks;    /* workgroups of the wavefront kernels, 0 = library pick */
  void* stream;            /* hipStream_t, NULL = library stream          */
  double adaptive_threshold; /* LMN_FLAG_ADAPTIVE: stop a pixel once the standard error
                                of its luminance is below this (RenderConfig default 0.03) */
  double time_limit;       /* seconds, 0 = none (RenderConfig.time_limit): rows begun after
                              it take one sample through the pixel centre; one frame only,
                              wavefront path only                                          */
} lmn_render_params;
#define LMN_FLAG_STAGE_TIMING 1u /* time each wavefront stage with HIP events */
#define LMN_FLAG_NO_GRAPH 2u     /* launch wavefront iterations directly, not as a captured hipGraph */
/* Plain FP64 kernel (LMN_PATH_PLAIN_FP64): the waves per SIMD its build is
 * held to, in bits 8-10 of flags: 2 (its full register budget, no spills),
 * 3, or 4 (the default, also for 0 and other values).  Results are
 * identical; the fastest depends on the scene. */
#define LMN_FLAG_PLAIN_WAVES_SHIFT 8u
#define LMN_FLAG_PLAIN_WAVES(n) ((uint32_t)(n) << LMN_FLAG_PLAIN_WAVES_SHIFT)
#define LMN_FLAG_ADAPTIVE 4u     /* adaptive sampling (src/render.rs:1626-1677): each pixel takes
                                    stratified samples in order until n >= max(spp/4, 16), n > 1
                                    and sqrt(var(luminance) / n) < adaptive_threshold; its value is
                                    the unweighted mean of the samples taken (no pixel filter) */

typedef struct lmn_stats {
  uint64_t paths;          /* samples averaged into the image, the reference's total_rays:
                              W*H*ceil(sqrt(spp))^2 (src/render.rs:1589), fewer with
                              adaptive sampling or a time limit                       */
  uint64_t camera_rays;
  uint64_t extension_rays; /* closest-hit queries after the camera ray    */
  uint64_t shadow_rays;    /* NEE visibility queries                      */
  double render_ms;        /* device time of the pixel loop               */
  uint32_t iterations;     /* wavefront iterations executed               */
  uint32_t extend_launches;
  double extend_ms;        /* summed stage times (LMN_FLAG_STAGE_TIMING):  */
  double shade_ms;         /* extend and shadow are the main BVH's kernels */
  double shadow_ms;
  double regen_ms;         /* the bookkeeping between iterations (new work items take the
                              freed slots; extend and shade generate their camera rays)    */
  double march_ms;         /* the marched kinds' extend and shadow kernels */
  uint64_t samples_traced; /* camera samples traced (adaptive rounds trace a few past the stop) */
  uint32_t rounds;         /* wavefront runs: adaptive rounds, time-limit bands              */
  uint32_t expired_rows;   /* rows rendered with one centre sample after the time limit      */
} lmn_stats;

/* ---- ray-batch parity harness -------------------------------------------- */
typedef struct lmn_ray {
  double origin[3];
  double direction[3];
  double t_min, t_max;
} lmn_ray;
typedef struct lmn_hit {
  double t;          /* reference-reported t (unnormalised direction units) */
  int64_t prim;      /* index into lmn_scene_desc.prims, -1 = miss          */
  double u, v;
  double normal[3];  /* facing the ray (HitRecord::new, src/hit.rs:19-43)   */
  int32_t front_face;
  int32_t pad;
} lmn_hit;

/* ---- entry points --------------------------------------------------------- */

/* Last error message of the calling thread ("" if none). */
const char* lmn_last_error(void);

/* Parse scene TOML exactly as scene::load_scene (src/scene.rs:846-1645) does
 * for the hot-path subset; file references (meshes, textures) are resolved
 * relative to base_dir (NULL = current directory). The descriptor is owned
 * by the library and released with lmn_desc_free. */
int lmn_load_scene(const char* toml_text, const char* base_dir, lmn_scene_desc** out);
void lmn_desc_free(lmn_scene_desc* desc);

/* Replace texture `index` texels (configs with programmatic albedo grids). */
int lmn_desc_set_texture(lmn_scene_desc* desc, uint32_t index, uint32_t width, uint32_t height,
                         const float* rgb);
/* Append an image texture; returns its index in *out_index. */
int lmn_desc_add_texture(lmn_scene_desc* desc, uint32_t width, uint32_t height, const float* rgb,
                         uint32_t* out_index);
/* Replace material `index`. */
int lmn_desc_set_material(lmn_scene_desc* desc, uint32_t index, const lmn_material* m);
/* Append a material; returns its index in *out_index. */
int lmn_desc_add_material(lmn_scene_desc* desc, const lmn_material* m, uint32_t* out_index);
/* Append an indexed triangle mesh as the loader appends a [[mesh]]: one
 * triangle prim (and one world object) per face, in face order, vertices
 * already scaled/offset (src/obj.rs:205, :232-239).  verts: n_verts*3
 * doubles; tris: n_tris*3 vertex indices; materials: one material index per
 * triangle, or NULL to use `material` for all.  Binary counterpart of the
 * ASCII PLY path for large meshes. */
int lmn_desc_add_triangles(lmn_scene_desc* desc, const double* verts, uint64_t n_verts, const uint32_t* tris,
                           uint64_t n_tris, uint32_t material, const uint32_t* materials);
/* The same with per-corner attributes: normals n_tris*9 doubles (the three
 * corners' normals of each triangle, unit as obj.rs makes them) and/or uvs
 * n_tris*6 doubles (NULL = absent).  Triangles with an attribute become
 * MESH_TRIANGLE prims (src/obj.rs:250-313), the rest TRIANGLE prims. */
int lmn_desc_add_mesh(lmn_scene_desc* desc, const double* verts, uint64_t n_verts, const uint32_t* tris,
                      uint64_t n_tris, const double* normals, const double* uvs, uint32_t material,
                      const uint32_t* materials);

/* Flatten, build the SAH BVH and upload a scene to the current HIP device. */
typedef struct lmn_scene lmn_scene;
int lmn_scene_create(const lmn_scene_desc* desc, lmn_scene** out);
void lmn_scene_destroy(lmn_scene* scene);
/* BVH shape of the uploaded scene (node count, leaf count, max depth). */
int lmn_scene_info(const lmn_scene* scene, uint64_t* n_nodes, uint64_t* n_leaves, uint32_t* max_depth,
                   uint64_t* device_bytes);

/* Render one frame: replaces the pixel loop src/render.rs:1587-1747.
 * hdr_rgb: caller-owned host buffer of crop_h*crop_w*3 floats (or H*W*3),
 * row 0 = top, linear HDR before exposure. */
int lmn_render(const lmn_scene* scene, const lmn_render_params* params, float* hdr_rgb, lmn_stats* stats);
/* Same, with the framebuffer left in device memory (d_hdr_rgb is a device
 * pointer of the same size).  No host copies inside the timed region. */
int lmn_render_device(const lmn_scene* scene, const lmn_render_params* params, float* d_hdr_rgb,
                      lmn_stats* stats);

/* Render many scenes of uneven size in one wavefront: path pools are shared
 * and every scene's paths are issued from one work list. */
int lmn_render_batch(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                     float* const* hdr_rgb, lmn_stats* stats);
/* Same with device framebuffers (d_hdr_rgb[i] device pointers). */
int lmn_render_batch_device(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                            float* const* d_hdr_rgb, lmn_stats* stats);

/* Closest-hit queries on a ray batch (SceneWorld::hit, src/scene.rs:817-834).
 * path: LMN_PATH_WAVEFRONT_FP32 traces in FP32, LMN_PATH_PLAIN_FP64 in FP64. */
int lmn_trace_rays(const lmn_scene* scene, const lmn_ray* rays, size_t n, uint32_t path,
                   uint32_t intersect_mode, lmn_hit* hits);

/* Device / library information. */
int lmn_device_info(char* name, size_t name_len, int* n_cu, double* hbm_bytes);

/* Traversal work of the extend stage, counted by an instrumented run of the
 * same kernels on the same frame: the algorithmic bytes of the roofline are
 * node_fetches * node_bytes + prim_tests * prim_bytes.  No reference
 * counterpart: measurement support for the render entry. */
typedef struct lmn_trace_counts {
  uint64_t node_fetches;
  uint64_t prim_tests;
  uint64_t tri_tests;
  uint64_t rays;
  uint32_t node_bytes;
  uint32_t prim_bytes;
  uint64_t hits;        /* rays of the pass that found a hit */
} lmn_trace_counts;
int lmn_profile_extend(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                       float* const* d_hdr_rgb, lmn_trace_counts* counts);

/* ---- autotuner of the wavefront launch -------------------------------------
 * No reference counterpart (the reference's only knob is its thread count,
 * src/main.rs:562-567).  Times short probe renders of the batch -- the same
 * scenes at reduced spp, written into d_hdr_rgb, or into buffers of the
 * library's own when d_hdr_rgb is NULL -- over the path-pool size,
 * workgroup size, LDS short-stack depth and grid size (coordinate descent
 * from the defaults), stores the best under the workload's signature in the
 * tuning cache (file $LMN_TUNE_CACHE, else ~/.cache/luminara_amd/tuning.txt)
 * and returns it.  Wavefront renders that leave a knob at 0 in
 * lmn_render_params take it from the cache.  Results do not depend on these
 * knobs: images are bitwise identical under every configuration. */
typedef struct lmn_tuning {
  uint32_t pool_paths, block_size, stack_lds, grid_blocks;
  double probe_ms;    /* probe render time of the chosen configuration  */
  double default_ms;  /* probe render time of the default configuration */
  uint64_t probe_paths;
  uint32_t candidates; /* configurations timed */
  uint32_t pad;
} lmn_tuning;
int lmn_autotune(const lmn_scene* const* scenes, size_t n_scenes, const lmn_render_params* params,
                 float* const* d_hdr_rgb, lmn_tuning* out);

/* ---- path-replay adjoint ------------------------------------------------------
 * Gradient of an image loss L with respect to scene parameters, given the
 * loss's gradient dL/dI with respect to the linear HDR framebuffer that
 * lmn_render produces for the same params.  Each path is replayed from its
 * counter-based seed (nothing is stored between the forward render and the
 * adjoint), then swept backwards to scatter gradients.  The estimator is the
 * derivative of the expected image of the reference's ray_color
 * (src/render.rs:1472-1537): sampled directions are held fixed (detached
 * sampling), Russian roulette enters through its survival probability.
 * No reference counterpart: the reference has no differentiable path.
 * Parameters:
 *   - the texels of one image texture (albedo of a Lambertian, or of a
 *     metallic = 0 microfacet, material);
 *   - the roughness of every microfacet material;
 *   - the intensity of one emissive material (its surfaces and the NEE
 *     lights registered from it).
 * Supported for max_depth <= LMN_ADJOINT_MAX_DEPTH.  Gradient accumulation
 * uses floating-point atomics: results agree run to run to rounding. */
#define LMN_ADJOINT_MAX_DEPTH 50
typedef struct lmn_adjoint_spec {
  int32_t texture;           /* texture whose texels are parameters, -1 = none        */
  int32_t emission_material; /* emissive material whose intensity is one, -1 = none   */
  uint32_t flags;            /* reserved, 0                                           */
  uint32_t pad;
} lmn_adjoint_spec;
typedef struct lmn_grads {
  float* texels;       /* out: width*height*3 of the spec's texture (NULL = skip)  */
  float* roughness;    /* out: n_materials; 0 for non-microfacet (NULL = skip)     */
  double emission;     /* out: dL/d intensity                                      */
  uint64_t paths;      /* out: paths replayed                                      */
  double adjoint_ms;   /* out: kernel time                                         */
} lmn_grads;
/* Host buffers: dL/dI (crop_h*crop_w*3 floats, row 0 = top) in, grads out. */
int lmn_adjoint(const lmn_scene* scene, const lmn_render_params* params, const lmn_adjoint_spec* spec,
                const float* dl_dimage, lmn_grads* grads);
/* Device buffers for dL/dI, grads->texels and grads->roughness. */
int lmn_adjoint_device(const lmn_scene* scene, const lmn_render_params* params, const lmn_adjoint_spec* spec,
                       const float* d_dl_dimage, lmn_grads* grads);

/* ---- image outputs: post chain, HDR encoding, AOV passes -------------------
 * What render() does with the pixel loop's rows before returning its
 * RenderResult (src/render.rs:1540-1554), restricted to the default chain:
 *   exposure, optionally automatic          src/render.rs:2594-2608, :525-541
 *   hdr_data = rows * exposure as f32        src/render.rs:2610-2624
 *   depth fog from a pixel-centre ray        src/render.rs:2634-2665, :3117-3124
 *   tone curve, sRGB, 8-bit quantisation     src/render.rs:3099-3202, :3606-3658
 *   depth / normal / albedo passes           src/render.rs:3508-3584
 * and the Radiance RGBE pixels of the HDR writer (src/main.rs:948-989).
 * Per-pixel arithmetic is FP64 in the reference's order (contraction off).
 * The options that change pixels otherwise (denoise, bloom, vignette, grading,
 * dither, gamma, brightness, ...) are not on this path.  The host-buffer entry
 * points return when their outputs are written.  The _device entry points are
 * ordered on the caller's stream (params->stream or `stream`): they return once
 * the work is enqueued there, or once it is done when the stream is NULL (the
 * library's stream) or the automatic exposure has to be read back. */
enum { LMN_TONE_ACES = 0, LMN_TONE_REINHARD = 1, LMN_TONE_FILMIC = 2, LMN_TONE_NONE = 3 };
typedef struct lmn_post_params {
  double exposure;        /* RenderConfig.exposure                                 */
  uint32_t tone_map;      /* LMN_TONE_* (RenderConfig.tone_map)                    */
  uint32_t auto_exposure; /* 1: exposure *= 0.18 / log-average luminance, or 1 when
                             that average is <= 1e-6 (RenderConfig.auto_exposure) */
  double depth_fog;       /* fog density, 0 = off (RenderConfig.depth_fog)         */
  double fog_color[3];    /* RenderConfig.depth_fog_color, default 0.8, 0.85, 0.9  */
} lmn_post_params;

/* AOV passes and the depth-fog buffer of one frame (same scene and params as
 * lmn_render; spp and max_depth are not used).  One ray per crop pixel through
 * the pixel centre, u = (i + 0.5) / (W - 1), v = (y + 0.5) / (H - 1), with the
 * camera's lens sample from the pixel's own stream, closest hit over
 * [0.001, inf).  Outputs (crop_h*crop_w pixels, row 0 = top; NULL = skip):
 *   depth   FP32 hit t in units of the unnormalised camera direction, +inf where
 *           the ray escapes (the fog buffer of :2634-2665; the depth pass of
 *           :3532 is this with +inf replaced by 0);
 *   normal  3 bytes: (n * 0.5 + 0.5) * 255 of the facing normal, truncated;
 *   albedo  3 bytes: one scatter's attenuation clamped to [0, 1] * 255, or for a
 *           material that does not scatter, emission / max(max(e), 1) * 255.
 * Normal and albedo bytes are 0 where the ray escapes.  params->path
 * LMN_PATH_PLAIN_FP64 traces in FP64, any other path in FP32. */
int lmn_aov(const lmn_scene* scene, const lmn_render_params* params, float* depth, uint8_t* normal,
            uint8_t* albedo);
int lmn_aov_device(const lmn_scene* scene, const lmn_render_params* params, float* d_depth, uint8_t* d_normal,
                   uint8_t* d_albedo);

/* Post chain of one framebuffer: hdr (height*width*3 floats, row 0 = top, linear
 * HDR as lmn_render writes it) -> rgba (height*width*4 bytes, alpha 255).
 * fog_depth: the depth output of lmn_aov for the same frame, required when
 * post->depth_fog > 0, else ignored.  hdr_exposed (optional): hdr * exposure
 * rounded to FP32 (RenderResult.hdr_data).  exposure_used (optional): the
 * exposure applied.  stream: hipStream_t, NULL = the library's stream. */
int lmn_post(const float* hdr, uint32_t width, uint32_t height, const lmn_post_params* post,
             const float* fog_depth, uint8_t* rgba, float* hdr_exposed, double* exposure_used);
int lmn_post_device(const float* d_hdr, uint32_t width, uint32_t height, const lmn_post_params* post,
                    const float* d_fog_depth, uint8_t* d_rgba, float* d_hdr_exposed, double* exposure_used,
                    void* stream);

/* Allocation-free variant of lmn_post_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as lmn_post_device, bit for bit; the
 * automatic exposure's scratch (2 + 1024 doubles, 8.2 KB, independent of the
 * frame size, so no frame is too large for it) is a static device array of
 * the library, one per GPU, held until the process ends.  stream NULL = the
 * default (null) stream, not the library's stream.  Calls that overlap on one
 * GPU share that array and race: serialise them.  The first call on a device
 * looks the array's address up (hipGetSymbolAddress) and uploads the sRGB
 * tables into constant memory; no call allocates. */
int lmn_post_device_noalloc(const float* d_hdr, uint32_t width, uint32_t height, const lmn_post_params* post,
                            const float* d_fog_depth, uint8_t* d_rgba, float* d_hdr_exposed, double* exposure_used,
                            void* stream);

/* Radiance RGBE pixels (src/main.rs:964-986): n_pixels*3 floats in, n_pixels*4
 * bytes out (the writer prepends the header "#?RADIANCE\nFORMAT=32-bit_rle_rgbe
 * \n\n-Y H +X W\n"). */
int lmn_encode_rgbe(const float* rgb, uint64_t n_pixels, uint8_t* rgbe);
int lmn_encode_rgbe_device(const float* d_rgb, uint64_t n_pixels, uint8_t* d_rgbe, void* stream);

/* A whole RenderResult in one call: lmn_render, then on the device the AOV
 * ray (when fog or a pass is asked for), the post chain and the RGBE encoding;
 * only the requested outputs cross PCIe.  Host buffers, crop_h*crop_w pixels,
 * NULL = not wanted. */
typedef struct lmn_image {
  uint8_t* rgba;   /* *4  RenderResult.pixels                              */
  float* hdr;      /* *3  RenderResult.hdr_data (after exposure)           */
  uint8_t* rgbe;   /* *4  Radiance pixels of hdr_data                      */
  float* depth;    /* *1  RenderResult.depth_pass (0 where the ray escapes) */
  uint8_t* normal; /* *3  RenderResult.normal_pass                         */
  uint8_t* albedo; /* *3  RenderResult.albedo_pass                         */
  double exposure;    /* out: exposure applied                             */
  double log_average; /* out: log-average luminance (auto-exposure), else 0 */
  double post_ms;     /* out: device time of AOV ray + post chain + encoding */
} lmn_image;
int lmn_render_image(const lmn_scene* scene, const lmn_render_params* params, const lmn_post_params* post,
                     lmn_image* image, lmn_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* LUMINARA_H */

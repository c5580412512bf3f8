/*
 * oracle.h - CPU FP64 restatement of the reference's render hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load this library, and only as the checker or
 * as the timed CPU baseline; the product (luminara_amd/) never links or calls
 * it.
 *
 * The reference (Rust, src/) cannot be built in this environment (no Rust
 * toolchain, crates not vendored), so this is a line-by-line C++ restatement
 * of the functions on the path, structured like the reference: a virtual
 * Hittable tree with the reference's wrappers (Translate/RotateX/Y/Z/Scale/
 * NonUniformScale), HittableList, the reference's 12-bucket SAH BvhNode with
 * left-first recursion, recursive ray_color with NEE, the row-parallel pixel
 * loop.  Each function cites the file:line it follows (oracle.cpp).
 *
 * RNG modes:
 *   ORACLE_RNG_COUNTER   the GPU's counter-based SplitMix64 streams (common
 *                        random numbers with the HIP kernels);
 *   ORACLE_RNG_REFERENCE rand 0.9.2 SmallRng = xoshiro256++ seeded by
 *                        SplitMix64 per row (src/render.rs:1608), restated
 *                        from the published algorithm (Blackman & Vigna).
 *                        The generator is pinned by the published outputs of
 *                        the reference implementation and of rand's
 *                        seed_from_u64 (tests/test_oracle_pins.py); the per-row
 *                        seeding and draw order follow the cited call sites.
 */
#ifndef LMN_ORACLE_H
#define LMN_ORACLE_H
#include "../include/luminara.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_RNG_COUNTER = 0, ORACLE_RNG_REFERENCE = 1 };

typedef struct oracle_params {
  uint32_t width, height, spp, max_depth;
  uint64_t seed;
  uint32_t filter;
  uint32_t has_crop;
  uint32_t crop[4];
  uint32_t rng_mode;
  uint32_t scene_key;
  uint32_t threads;        /* 0 = hardware concurrency */
  uint32_t exact_marched;  /* 1 = exact roots for torus/superellipsoid/spring */
  uint32_t row_begin;      /* render only rows [row_begin, row_end) of the crop, */
  uint32_t row_end;        /* every row_step-th one; row_end = 0 means all rows, */
  uint32_t row_step;       /* row_step 0 means 1                                 */
  uint32_t adaptive;       /* adaptive sampling (render.rs:1626-1677)            */
  double adaptive_threshold;
  double time_limit;       /* seconds, 0 = none (render.rs:1600-1625)            */
} oracle_params;

typedef struct oracle_stats {
  uint64_t camera_rays, extension_rays, shadow_rays; /* camera_rays: the reference's total_rays */
  double render_secs; /* pixel loop only (src/render.rs:1596 -> :1746) */
  uint32_t threads;
  uint32_t pad;
} oracle_stats;

/* Closest hit per ray, prim = descriptor prim index (-1 = miss). */
int oracle_trace(const lmn_scene_desc* desc, const lmn_ray* rays, size_t n, uint32_t exact_marched, lmn_hit* hits);

/* Pixel loop.  mean_rgb: rows*crop_w*3 doubles (weighted mean as the
 * reference returns it); var_rgb (optional): per-pixel per-channel sample
 * variance of the clamped samples (unweighted). */
int oracle_render(const lmn_scene_desc* desc, const oracle_params* p, double* mean_rgb, double* var_rgb,
                  oracle_stats* stats);

/* Single material scatter / NEE / background evaluations for known-answer
 * tests (fixed hit records). */
int oracle_scatter(const lmn_scene_desc* desc, uint32_t material, const double ray_dir[3], const double point[3],
                   const double normal[3], int front_face, double t, double u, double v, uint64_t key,
                   uint32_t counter, double out_dir[3], double out_att[3], uint32_t* out_counter);

const char* oracle_last_error(void);

/* Single-function evaluations for the known-answer tests: the mean of n NEE
 * samples at a fixed surface point (src/render.rs:1346-1466; sample i uses the
 * counter stream key + i*golden), one camera ray (src/camera.rs:82-103),
 * emitted radiance (src/material.rs:241-243), the background colour
 * (src/render.rs:72-172), a pixel filter weight (src/render.rs:198-237) and
 * Schlick's reflectance (src/material.rs:138-141). */
int oracle_nee(const lmn_scene_desc* desc, const double point[3], const double normal[3], const double albedo[3],
               uint64_t key, uint32_t n, double out_mean[3], uint64_t* out_shadow_rays);
int oracle_camera_ray(const lmn_scene_desc* desc, double s, double t, uint64_t key, double origin[3], double dir[3]);
int oracle_emitted(const lmn_scene_desc* desc, uint32_t material, double u, double v, double out[3]);
int oracle_background(const lmn_scene_desc* desc, const double dir[3], double out[3]);
double oracle_filter_weight(uint32_t filter, double dx, double dy);

/* Path-replay adjoint in FP64 (counter RNG): dL/dI (crop_h*crop_w*3) in;
 * gradients with respect to the texels of `texture` (width*height*3), every
 * material's roughness (n_materials; microfacet only) and the intensity of
 * `emission_material` out.  Either index may be -1. */
int oracle_adjoint(const lmn_scene_desc* desc, const oracle_params* p, int32_t texture, int32_t emission_material,
                   const double* dldi, double* g_tex, double* g_rough, double* g_em);
double oracle_reflectance(double cosine, double ratio);

/* Image outputs.  oracle_aov: the AOV passes of render.rs:3508-3584 (params:
 * width, height, crop, seed, scene_key, rng_mode, exact_marched, threads);
 * depth is +inf where the pixel's ray escapes (the fog buffer of
 * :2634-2665).  oracle_post: exposure (automatic when asked, :2594-2608),
 * hdr_data (:2610-2624), depth fog and the default per-pixel chain
 * (:3099-3202, :3606-3658) on an FP32 framebuffer of n pixels.
 * oracle_rgbe: the pixel loop of write_radiance_hdr (main.rs:964-986). */
int oracle_aov(const lmn_scene_desc* desc, const oracle_params* p, float* depth, uint8_t* normal, uint8_t* albedo);
int oracle_post(const float* hdr, uint64_t n, double exposure, uint32_t tone_map, uint32_t auto_exposure,
                const float* fog_depth, double fog_density, const double fog_rgb[3], uint8_t* rgba,
                float* hdr_exposed, double* exposure_used);
double oracle_log_average_luminance(const float* hdr, uint64_t n);
double oracle_tone(uint32_t mode, double x);
uint32_t oracle_srgb8(double x);
void oracle_rgbe(const float* rgb, uint64_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif

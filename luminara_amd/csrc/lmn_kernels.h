// Kernel-side data structures of the wavefront and the launch entry points
// used by the C ABI (lmn_api.hip).
#pragma once
#include "../../include/luminara.h"
#include "lmn_render.h"

namespace lmn {

// Path pool, structure of arrays (16 B per path per field):
//   ob   = origin.xyz, scene index (bits)
//   db   = direction.xyz, origin record (bits: the primitive the ray leaves)
//   meta = work item, remaining depth | skip-emission flag << 16,
//          rng key lo, rng key hi
//   aux  = rng counter, beta.xyz (bit patterns)
// Extend reads a stored ray from ob and db alone.
struct PathSoA {
  float4* ob;
  float4* db;
  uint4* meta;
  uint4* aux;
};
// Shadow-ray queue: o = origin.xyz, distance (negated once occluded);
// d = unit dir.xyz, work item; c = unoccluded contribution.rgb, scene index
struct ShadowSoA {
  float4* o;
  float4* d;
  float4* c;
};

// The two queue tails that every shading workgroup bumps (survivors, shadow
// rays) sit on cache lines of their own: same-line atomics serialise in one
// L2 channel.
struct WaveCounters {
  uint32_t active;       // paths in the current pool
  uint32_t n_surv;       // of which survivors of the last iteration; slots n_surv.. are new camera rays
  uint64_t cam_item;     // work item of slot n_surv (the camera rays are items cam_item, cam_item + 1, ...)
  uint64_t next_item;    // next work item to issue (chunk relative)
  uint64_t n_items;      // work items in the chunk
  uint64_t stat_extend;  // closest-hit queries issued
  uint64_t stat_shadow;  // any-hit queries issued
  uint64_t stat_camera;  // camera rays generated
  uint64_t iterations;
  uint32_t n_cand[2][4];  // marched candidates per kind + screened superellipsoids, [extend, shadow]
  uint32_t n_ovf[2];      // rays left to the marched fallback, [extend, shadow]
  uint32_t n_brk;         // superellipsoid brackets queued for refinement (extend)
  uint32_t pad2;
  alignas(256) uint32_t next_active;  // survivors compacted into the next pool
  alignas(256) uint32_t n_shadow;     // shadow rays queued this iteration
};

// Candidate queues of the marched pass.  A ray lists the marched
// primitives whose boxes it enters (within the main pass's hit distance);
// each kind (torus, superellipsoid, spring) has `cap` entries (ray index,
// primitive record) and is tested by a kernel of its own with one candidate
// per lane; superellipsoid candidates pass an FP32 screen first, the
// survivors (region 3) get the exact walk, and the extend stage refines the
// brackets it finds in a kernel of their own.  `key` holds per ray the closest
// marched hit (t bits << 32 |
// record, MARCH_EMPTY: none) or, for shadow rays, a nonzero occluded flag.
// Rays with more than MARCH_MAX_CAND candidates, or whose candidates do not
// fit, are listed in `ovf` and traced whole by the fallback kernel.
// A superellipsoid candidate's bracketing leaf, queued between the exact
// walk and its refinement (extend stage).
struct MarchBracket {
  uint32_t ray, rec, pad0, pad1;
  double a, b, fa, fb;
};
struct MarchQ {
  uint2* cand;               // 4 * cap entries; region r at cand + r * cap
  unsigned long long* key;   // per path / shadow ray
  uint32_t* ovf;             // per path / shadow ray
  MarchBracket* brk;         // cap_brk entries
  uint32_t cap, cap_brk;
  uint32_t kinds;            // bit k: kind K_TORUS + k present in the batch
  uint32_t pad;
};
constexpr unsigned long long MARCH_EMPTY = ~0ull;
constexpr uint32_t MARCH_SKIP = 0xffffffffu;  // reserved slot left unused
constexpr uint32_t MARCH_MAX_CAND = 16;

// LDS short stack with spill to a global per-thread region.  The first SS
// words of each thread live in LDS, deeper ones in `spill` (coalesced across
// the threads of the grid).  A closest-hit walk pushes (entry, distance)
// pairs: with PAIRS the two words of a pair sit side by side per lane (word w
// of lane t at [w / 2][t][w % 2]), so an entry is one 8-byte access to LDS
// or to the spill area, with one address and one depth check; without PAIRS
// word w of lane t is at [w][t], conflict-free for the one-word entries of a
// visibility walk.  SS is even, so a pair never straddles LDS and spill.
template <int SS, int BS, bool PAIRS = false>
struct LdsStack {
  static_assert(SS % 2 == 0, "pairs must not straddle LDS and spill");
  uint32_t* lds;
  uint32_t* spill;
  uint32_t tid;
  uint32_t slot;
  uint32_t nthreads;
  int sp;
  // word w of this lane: its index in LDS (w < SS) or in the spill area
  __device__ uint32_t lds_index(int w) const {
    return PAIRS ? uint32_t(w >> 1) * 2u * BS + 2u * tid + uint32_t(w & 1) : uint32_t(w) * BS + tid;
  }
  __device__ size_t spill_index(int w) const {
    const size_t d = size_t(w - SS);
    return PAIRS ? (d >> 1) * 2 * nthreads + 2 * size_t(slot) + (d & 1) : d * nthreads + slot;
  }
  // the same for the pair at word w, which push2/pop2 keep even: one
  // multiply-add off a per-lane base
  __device__ uint32_t pair_lds_index(int w) const { return uint32_t(w) * BS + 2u * tid; }
  __device__ size_t pair_spill_index(int w) const { return size_t(w - SS) * nthreads + 2 * size_t(slot); }
  // Accesses through address-space-qualified pointers: the LDS and spill
  // branches then stay LDS and global accesses (from two generic pointers
  // the compiler forms one flat access to whichever the branch picks).
#if defined(__HIP_DEVICE_COMPILE__)
  template <class T>
  __device__ __attribute__((address_space(3))) T* L(uint32_t i) const {
    return (__attribute__((address_space(3))) T*)(lds + i);
  }
  template <class T>
  __device__ __attribute__((address_space(1))) T* G(size_t i) const {
    return (__attribute__((address_space(1))) T*)(spill + i);
  }
#else
  template <class T>
  __device__ T* L(uint32_t i) const { return reinterpret_cast<T*>(lds + i); }
  template <class T>
  __device__ T* G(size_t i) const { return reinterpret_cast<T*>(spill + i); }
#endif
  __device__ void push(uint32_t v) {
    if (sp < SS)
      *L<uint32_t>(lds_index(sp)) = v;
    else
      *G<uint32_t>(spill_index(sp)) = v;
    ++sp;
  }
  __device__ uint32_t pop() {
    --sp;
    uint32_t v;
    if (sp < SS)
      v = *L<uint32_t>(lds_index(sp));
    else
      v = *G<uint32_t>(spill_index(sp));
    return v;
  }
  // a pair (a first, then b): pop2 returns them as pushed
  __device__ void push2(uint32_t a, uint32_t b) {
    if constexpr (PAIRS) {
      if (sp < SS)
        *L<uint2>(pair_lds_index(sp)) = make_uint2(a, b);
      else
        *G<uint2>(pair_spill_index(sp)) = make_uint2(a, b);
      sp += 2;
    } else {
      push(a);
      push(b);
    }
  }
  __device__ void pop2(uint32_t& a, uint32_t& b) {
    if constexpr (PAIRS) {
      sp -= 2;
      uint2 v;
      if (sp < SS) {
#if defined(__HIP_DEVICE_COMPILE__)
        // one LDS read, waited on in the same block: written as a plain load
        // the compiler merges it with the spill branch's into one flat load
        uint64_t x;
        const uint32_t addr = uint32_t(size_t(L<uint2>(pair_lds_index(sp))));
        asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(x) : "v"(addr) : "memory");
        v = make_uint2(uint32_t(x), uint32_t(x >> 32));
#else
        v = *L<uint2>(pair_lds_index(sp));
#endif
      } else {
        v = *G<uint2>(pair_spill_index(sp));
      }
      a = v.x;
      b = v.y;
    } else {
      b = pop();
      a = pop();
    }
  }
  __device__ bool empty() const { return sp == 0; }
};

// Scenes of one wavefront launch: per-scene views (device array), the
// chunk's parts (FrameDev each) and their first work items (n_parts + 1
// prefix, chunk-relative).
struct BatchDev {
  const SceneView<float>* views;
  const FrameDev* parts;
  const uint64_t* part_begin;
  uint32_t n_parts;
  uint32_t pad;
};

// Closest hit of a path, reconstructed by the extend stage while the hit
// primitive's record is still in cache: a = (unit normal facing the ray,
// true t), b = (u, v, material | front << 31, reported t).  b.z = ~0u: miss.
struct HitSoA {
  float4* a;
  float4* b;
  uint32_t* rec;  // the hit's primitive record: the origin record of the path's next ray
};
constexpr uint32_t HIT_MISS = 0xffffffffu;

struct WaveLaunch {
  hipStream_t stream;
  uint32_t grid;       // workgroups of the grid-stride kernels
  uint32_t block;      // lanes per workgroup: 128 or 256
  uint32_t pool;       // path pool size
  uint32_t stack_lds;  // LDS short-stack entries per lane
  uint32_t feat;       // FEAT_* specialisation of the scene
  uint32_t feat_main;  // FEAT_* of the main BVH's records (FEAT_F64: FP64 islands there)
  HitSoA hit;          // per path: the closest hit's surface (extend -> shade)
  ShadowSoA sh;
  float4* samples;
  WaveCounters* ctr;
  uint32_t* spill;
  TraceCounters* tc;
  MarchQ mq;           // marched pass (scenes with torus/superellipsoid/spring)
};

// One wavefront iteration: extend, shade, shadow, bookkeeping (new work
// items into the freed slots).  With `ev` (7 events), ev[0..6] bracket
// extend, its marched pass, shade, shadow, its marched pass and bookkeeping
// for per-stage timing.
void launch_wave_iteration(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, const PathSoA& nxt,
                           hipEvent_t* ev = nullptr);
void launch_counters_init(WaveCounters* ctr, uint64_t n_items, hipStream_t s);
// tile_w: the frame's entries run through 8x8 tiles of a crop this wide (0:
// row order), as the work items were issued
void launch_resolve(const float4* samples, double4* acc, uint32_t npix, uint32_t n_samples, uint32_t tile_w,
                    hipStream_t s);
// Same over listed pixels: samples[s * n_list + e] belongs to crop pixel list[e].
void launch_resolve_list(const float4* samples, double4* acc, const uint32_t* list, uint32_t n_list,
                         uint32_t n_samples, hipStream_t s);
// Adaptive sampling: per-pixel running state of the sequential stopping rule.
struct AdaptState {
  double mx, my, mz;  // running mean of the clamped colour
  double mean_lum, m2;  // Welford state of the clamped luminance
  uint32_t n;          // samples taken
  uint32_t done;       // the standard error fell below the threshold
};
// One sample into the adaptive state, in the reference's arithmetic and
// order (src/render.rs:1641-1668): the luminance clamp at 100, Welford's
// update, the running mean, the test after n >= min_samples and n > 1.
LMN_HD void adapt_take(AdaptState& a, double x, double y, double z, uint32_t min_samples, double threshold) {
#pragma clang fp contract(off)
  double lum0 = 0.2126 * x + 0.7152 * y + 0.0722 * z;
  if (lum0 > 100.0) {
    double f = 100.0 / lum0;
    x *= f;
    y *= f;
    z *= f;
  }
  a.n += 1;
  const double n = double(a.n);
  double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
  double d1 = lum - a.mean_lum;
  a.mean_lum += d1 / n;
  double d2 = lum - a.mean_lum;
  a.m2 += d1 * d2;
  a.mx = a.mx + (x - a.mx) / n;
  a.my = a.my + (y - a.my) / n;
  a.mz = a.mz + (z - a.mz) / n;
  if (a.n >= min_samples && a.n > 1) {
    double var = a.m2 / double(a.n - 1);
    if (sqrt(var / n) < threshold) a.done = 1;
  }
}
// max(spp/4, 16) (src/render.rs:1592)
LMN_HD uint32_t adaptive_min_samples(uint32_t spp_a) { return spp_a / 4 > 16 ? spp_a / 4 : 16; }
void launch_resolve_adaptive(const float4* samples, AdaptState* state, const uint32_t* list, uint32_t n_list,
                             uint32_t n_samples, uint32_t min_samples, double threshold, hipStream_t s);
void launch_adapt_compact(const AdaptState* state, const uint32_t* list_in, uint32_t n_in, uint32_t spp,
                          uint32_t* list_out, uint32_t* count, hipStream_t s);
void launch_adapt_finalize(const AdaptState* state, float* out, uint32_t npix, unsigned long long* taken,
                           hipStream_t s);
void launch_adapt_init(AdaptState* state, uint32_t* list, uint32_t npix, hipStream_t s);
void launch_zero_acc(double4* acc, uint32_t npix, hipStream_t s);
void launch_finalize(const double4* acc, float* out, uint32_t npix, hipStream_t s);
// Plain kernels, one path per thread: work items [k0, k0 + n) of a frame
// (k = s * npix + p), radiance and filter weight per item into samples[0, n).
void launch_plain_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, double4* samples, uint64_t k0,
                       uint64_t n, unsigned long long* rays, uint32_t waves, hipStream_t s);
void launch_plain_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, double4* samples, uint64_t k0,
                       uint64_t n, unsigned long long* rays, hipStream_t s);
// Adds `passes` whole sample passes of samples (pass-major) into acc per pixel.
void launch_plain_resolve(const double4* samples, double4* acc, uint32_t npix, uint32_t passes, hipStream_t s);
struct AdaptState;
void launch_plain_adaptive_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, AdaptState* state,
                                uint32_t s0, uint32_t s1, uint32_t min_samples, double threshold,
                                unsigned long long* rays, hipStream_t s);
void launch_plain_adaptive_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, AdaptState* state,
                                uint32_t s0, uint32_t s1, uint32_t min_samples, double threshold,
                                unsigned long long* rays, hipStream_t s);
// Path-replay adjoint (lmn_adjoint.hip)
struct AdjDev {
  int32_t texture;            // parameter texture, -1 = none
  int32_t emission_material;  // parameter emissive material, -1 = none
  uint64_t texel_base;        // texel index of the parameter texture's first texel
  uint32_t n_slots;           // roughness slots (microfacet materials)
  uint32_t pad;
};
void launch_adj_wsum_fp32(const CameraT<float>& C, const FrameDev& F, double* wsum, hipStream_t s);
void launch_adj_wsum_fp64(const CameraT<double>& C, const FrameDev& F, double* wsum, hipStream_t s);
void launch_adjoint_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, const AdjDev& A,
                         const float* dldi, const double* wsum, float* g_tex, double* part, uint32_t n_blocks,
                         uint64_t k0, uint64_t k1, uint32_t* spill, hipStream_t s);
void launch_adjoint_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, const AdjDev& A,
                         const float* dldi, const double* wsum, double* g_tex, double* part, uint32_t n_blocks,
                         uint64_t k0, uint64_t k1, uint32_t* spill, hipStream_t s);
void launch_adj_reduce(const double* part, uint32_t n_blocks, uint32_t width, double* out, hipStream_t s);
void launch_adj_zero_f32(float* p, size_t n, hipStream_t s);
void launch_adj_zero_f64(double* p, size_t n, hipStream_t s);
void launch_adj_to_float(const double* in, float* out, size_t n, hipStream_t s);

void launch_trace_fp32(const SceneView<float>& S, uint32_t feat, const lmn_ray* rays, lmn_hit* hits, size_t n,
                       hipStream_t s);
void launch_trace_fp64(const SceneView<double>& S, uint32_t feat, const lmn_ray* rays, lmn_hit* hits, size_t n,
                       hipStream_t s);

// AOV passes / depth-fog buffer: one pixel-centre ray per crop pixel (lmn_plain.hip).
void launch_aov_fp32(const SceneView<float>& S, uint32_t feat, const FrameDev& F, float* depth, uint8_t* normal,
                     uint8_t* albedo, hipStream_t s);
void launch_aov_fp64(const SceneView<double>& S, uint32_t feat, const FrameDev& F, float* depth, uint8_t* normal,
                     uint8_t* albedo, hipStream_t s);

// Post chain and HDR encoding (lmn_post.hip).
struct PostArgs {
  const float* hdr;           // n * 3 linear HDR
  const float* fog;           // n hit distances (+inf = escaped), null = no fog
  const double* exposure_dev; // [1] = exposure from the automatic exposure, null = `exposure`
  double exposure;
  double fog_density;
  double fog_rgb[3];
  uint32_t tone;              // LMN_TONE_*
  uint32_t vec;               // set by launch_post: every buffer 16-B aligned
  uint64_t n;
  uint8_t* rgba;              // n * 4
  float* hdr_exposed;         // n * 3, or null
};
constexpr uint32_t kLogLumBlocks = 1024;
// Uploads the sRGB quantisation thresholds to the current device (once).
hipError_t ensure_post_tables();
// Device address of the allocation-free post chain's scratch on the current
// device (2 + kLogLumBlocks doubles, looked up once per device; no allocation).
hipError_t noalloc_post_scratch(double** out, size_t* doubles);
void launch_post(const PostArgs& A, hipStream_t s);
// scratch: 2 + kLogLumBlocks doubles; scratch[0] = log-average luminance,
// scratch[1] = exposure * key factor.
void launch_log_luminance(const float* hdr, uint64_t n, double exposure, double* scratch, hipStream_t s);
void launch_rgbe(const float* rgb, uint64_t n, uint8_t* out, hipStream_t s);
void launch_depth_pass(float* depth, uint64_t n, hipStream_t s);

}  // namespace lmn

// HIP kernels of the render hot path for gfx950 (CDNA4, wave64):
//  * wavefront FP32 path tracer: extend (closest hit, LDS short stack),
//    shade (materials, emission, Russian roulette, NEE sampling, wave-ballot
//    compaction of survivors and shadow rays), shadow (any hit),
//    bookkeeping (new work items take the freed pool slots; their camera
//    rays are generated where extend and shade read them);
//  * plain one-path-per-thread megakernel in FP64 (or FP32);
//  * resolve: per-pixel luminance clamp, filter weights and FP64 sums;
//  * ray-batch trace for parity.
// The estimator is the reference's ray_color / pixel loop
// (src/render.rs:1472-1537, :1587-1747), unrolled into an iteration.
#include <algorithm>

#include "lmn_kernels.h"

namespace lmn {

__device__ __forceinline__ V3<float> f3(float4 a) { return {a.x, a.y, a.z}; }

// Block-aggregated slot reservation in two queues at once: each wave's
// ballot counts go to LDS, one thread takes the block's total from each
// global counter (one atomic per queue per block and round instead of one
// per wave: the counters are device-wide, so their atomics are the
// contended resource), and every lane gets its slot.  All lanes of the block
// must call it.  The second queue's slots are ordered by a 3-bit key
// within the block's range (stable within a key): the paths that survive a
// bounce go to the next pool grouped by the octant of their new direction,
// so the next extend's waves hold rays of one octant from nearby pixels and
// walk the tree more alike.  64 lanes of wave 0 scan the 8 x 8 (key, wave)
// counts; `lds` holds 64 + BS / 64 + 2 words and is reused after the
// last barrier of the next call, so a caller alternates two.
template <int BS>
__device__ __forceinline__ void block_reserve2_keyed(uint32_t* c0, uint32_t* c1, bool w0, bool w1, uint32_t key1,
                                                     uint32_t& s0, uint32_t& s1, uint32_t* lds) {
  constexpr int NW = BS / 64;
  static_assert(NW * 8 == 64, "one scan lane per (key, wave) pair");
  uint32_t* cnt1 = lds;            // [key][wave] counts, then exclusive offsets
  uint32_t* cnt0 = lds + 64;       // [wave] counts of the first queue, then offsets
  uint32_t* base = lds + 64 + NW;  // the block's first slot in each queue
  const uint32_t wave = threadIdx.x / 64u;
  const uint64_t lt = __lanemask_lt();
  const unsigned long long m0 = __ballot(w0);
  // the lanes of each key from one ballot per key bit (wave-uniform masks)
  const unsigned long long mv = __ballot(w1), b0 = __ballot((key1 & 1u) != 0u), b1 = __ballot((key1 & 2u) != 0u),
                           b2 = __ballot((key1 & 4u) != 0u);
  if (__lane_id() < 8u) {
    const uint32_t k = __lane_id();
    const unsigned long long m =
        mv & ((k & 1u) ? b0 : ~b0) & ((k & 2u) ? b1 : ~b1) & ((k & 4u) ? b2 : ~b2);
    cnt1[k * NW + wave] = uint32_t(__popcll(m));
  }
  const unsigned long long mine =
      mv & ((key1 & 1u) ? b0 : ~b0) & ((key1 & 2u) ? b1 : ~b1) & ((key1 & 4u) ? b2 : ~b2);
  if (__lane_id() == 0) cnt0[wave] = uint32_t(__popcll(m0));
  __syncthreads();
  if (wave == 0) {
    const uint32_t lane = __lane_id();
    const uint32_t v = cnt1[lane];
    uint32_t x = v;  // inclusive scan over the 64 (key, wave) counts, key-major
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = uint32_t(__shfl_up(int(x), off));
      if (int(lane) >= off) x += y;
    }
    cnt1[lane] = x - v;
    if (lane == 63) base[1] = x ? atomicAdd(c1, x) : 0u;
    if (lane == 0) {
      uint32_t t0 = 0;
      for (int k = 0; k < NW; ++k) {
        const uint32_t a = cnt0[k];
        cnt0[k] = t0;
        t0 += a;
      }
      base[0] = t0 ? atomicAdd(c0, t0) : 0u;
    }
  }
  __syncthreads();
  s0 = base[0] + cnt0[wave] + uint32_t(__popcll(m0 & lt));
  s1 = base[1] + cnt1[(key1 & 7u) * NW + wave] + uint32_t(__popcll(mine & lt));
}

// Slot in a device-wide queue for every lane of the wave that wants one:
// one atomic per wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* ctr, bool want) {
  const unsigned long long m = __ballot(want);
  if (m == 0ull) return 0u;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0u;
  if (int(__lane_id()) == leader) base = atomicAdd(ctr, uint32_t(__popcll(m)));
  base = uint32_t(__shfl(int(base), leader));
  return base + uint32_t(__popcll(m & __lanemask_lt()));
}

// The launch's tables (scene views, frame parts, part boundaries) are written
// by the host before the launch and only read by kernels: they are read
// through the constant address space, so a wave-uniform index becomes a
// scalar load through the scalar cache instead of a vector load through L1
// (the camera-ray part search is a chain of such loads).
template <class T>
__device__ __forceinline__ T uniform_load(const T* table, uint64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)table)[i];
#else
  return table[i];
#endif
}

// The scene view of a traversal kernel: read through the constant address
// space, except in the kernels with FP64 islands or transforms, whose
// traversal loops have no registers to spare (held in scalar registers the
// view cost them spills).  Inside for_scene's body the index is the lane's
// own, so these compile to per-lane loads of the fields the walk reads.
template <uint32_t FEAT>
__device__ __forceinline__ SceneView<float> scene_view(const BatchDev& B, uint32_t s) {
  if constexpr (kPinRecords<FEAT>)
    return uniform_load(B.views, s);
  else
    return B.views[s];
}

// ---------------------------------------------------------------------------
// Scene selection.  Paths of one scene are issued together, so a wave
// almost always holds one scene; the loop peels one scene per pass
// (readfirstlane picks it) until every lane has been served.  The body runs
// on the loop's divergent exit, where the compiler passes each lane's own
// sid (equal to the pass's index there): the scene's view comes from
// per-lane loads, and the lanes of a wave that holds rays of several scenes
// run the body together.  (Kept in scalar registers, the index made a wave
// run its scenes one after another: batch extend 111.4 -> 123.2 ms and
// shadow 41.2 -> 60.8 ms per step, the persistent shadow walk's refilled
// waves mixing scenes the most.)
template <class F>
__device__ __forceinline__ void for_scene(uint32_t sid, F&& body) {
  for (;;) {
    uint32_t s0 = uint32_t(__builtin_amdgcn_readfirstlane(int(sid)));
    if (sid == s0) {
      body(s0);
      return;
    }
  }
}
// The same with the pass's index in a scalar register inside the body, so
// that the records it selects are read with scalar loads (for the camera
// records of a wave's camera rays, which come from one part but for the
// waves that straddle a part boundary).  Two things keep the index scalar:
// - the loop ends on a wave-uniform condition, so the body runs inside the
//   pass with the lanes of one index (a body on a divergent loop exit runs
//   after the loop, once per lane's own index);
// - inside the branch the compiler knows s0 == sid and would pass on the
//   per-lane sid; the body takes a copy made opaque by an empty asm.
template <class F>
__device__ __forceinline__ void for_scene_scalar(uint32_t sid, F&& body) {
  bool done = false;
  do {
    const uint64_t waiting = __builtin_amdgcn_ballot_w64(!done);
    const uint32_t s0 = uint32_t(__builtin_amdgcn_readlane(int(sid), int(__builtin_ctzll(waiting))));
    uint32_t su = s0;
    asm volatile("" : "+s"(su));
    if (!done && sid == s0) {
      body(su);
      done = true;
    }
  } while (__builtin_amdgcn_ballot_w64(!done) != 0);
}

// Surface record of a closest hit (see HitSoA).
template <uint32_t FEAT>
__device__ __forceinline__ void store_hit(const SceneView<float>& S, HitSoA hit, uint32_t i, V3<float> o, V3<float> d,
                                          float t, uint32_t idx) {
  Surf<float> s;
  surface<float, FEAT>(S, idx, o, d, t, s, S.textured != 0);
  hit.a[i] = make_float4(s.n.x, s.n.y, s.n.z, s.t_true);
  hit.b[i] = make_float4(s.u, s.v, __uint_as_float(s.mat | (s.front ? 0x80000000u : 0u)), s.t);
  hit.rec[i] = idx;
}
__device__ __forceinline__ void store_miss(HitSoA hit, uint32_t i) {
  hit.b[i] = make_float4(0.f, 0.f, __uint_as_float(HIT_MISS), INFINITY);
}

__device__ __forceinline__ void add_sample(float4* samples, uint32_t item, V3<float> c) {
  float4 s = samples[item];
  s.x += c.x;
  s.y += c.y;
  s.z += c.z;
  samples[item] = s;
}

// ---------------------------------------------------------------------------
// New camera rays are not stored: the slots after an iteration's survivors
// stand for work items cam_item, cam_item + 1, ... (src/render.rs:1683-1690),
// and the stages that read a path recompute its camera ray from the item
// (generate: the item's part by binary search, its pixel, jitter, lens and
// time draws).  That saves writing 80 bytes per path in a kernel of its own
// and reading them back in extend and shade.
struct CameraPath {
  V3<float> o, d;
  float w;           // filter weight of the sample
  Rng rng;           // after the camera draws
  uint32_t scene, max_depth;
};
// SCALAR: the part's frame and camera records through scalar loads (the
// extend kernel, whose load-return path the traversal keeps busy); the
// other callers read them per lane, which spares their scalar registers.
template <bool SCALAR = false>
__device__ __forceinline__ void camera_path(const BatchDev& B, uint64_t item, CameraPath& c) {
  // part holding the item: largest p with part_begin[p] <= item.  A wave's
  // camera slots hold consecutive items, so the binary search runs once, on
  // the first active lane's item (the smallest), with wave-uniform values;
  // each lane then steps forward past the part boundaries below its own item
  // (none, or a few where the wave straddles parts).
  const uint64_t it0 = (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(item >> 32))))) << 32) |
                       uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(item))));
  uint32_t lo = 0, hi = B.n_parts;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (uniform_load(B.part_begin, mid) <= it0)
      lo = mid;
    else
      hi = mid;
  }
  while (lo + 1 < B.n_parts && uniform_load(B.part_begin, lo + 1) <= item) ++lo;
  auto gen = [&](uint32_t part) {
    const FrameDev F = uniform_load(B.parts, part);
    const CameraT<float> cam = uniform_load(B.views, F.scene).cam;
    uint32_t pixel;
    generate_item<float>(F, cam, F.item_base + (item - uniform_load(B.part_begin, part)), c.rng, c.o, c.d, c.w,
                         pixel);
    c.scene = F.scene;
    c.max_depth = F.max_depth;
  };
  if constexpr (SCALAR)
    for_scene_scalar(lo, gen);
  else
    for_scene(lo, gen);
}

// The ray of pool slot i: a stored path (i < n_surv) or a new camera ray.
// org: the record the ray leaves from (kNoOrigin for camera rays).
template <bool SCALAR = false>
__device__ __forceinline__ void pool_ray(const BatchDev& B, const PathSoA& cur, const WaveCounters* ctr, uint32_t i,
                                         V3<float>& o, V3<float>& d, uint32_t& sid, uint32_t& org) {
  const uint32_t ns = ctr->n_surv;
  if (i < ns) {
    const float4 a = cur.ob[i], b = cur.db[i];
    o = {a.x, a.y, a.z};
    d = {b.x, b.y, b.z};
    sid = __float_as_uint(a.w);
    org = __float_as_uint(b.w);
  } else {
    CameraPath c;
    camera_path<SCALAR>(B, ctr->cam_item + (i - ns), c);
    o = c.o;
    d = c.d;
    sid = c.scene;
    org = kNoOrigin;
  }
}

// ---------------------------------------------------------------------------
// Wavefront stages

// Extend, main pass: closest hit over the main BVH and the planes.
template <int SS, int BS, uint32_t FEAT, bool COUNT>
__device__ __forceinline__ void extend_body(const BatchDev& B, const PathSoA& cur, HitSoA hit,
                                            const WaveCounters* ctr, uint32_t* spill, TraceCounters* tc) {
  __shared__ uint32_t lds[SS * BS];
  uint32_t n = ctr->active;
  uint32_t nthreads = gridDim.x * BS;
  uint32_t gid = blockIdx.x * BS + threadIdx.x;
  for (uint32_t i = gid; i < n; i += nthreads) {
    V3<float> o, d;
    uint32_t sid, org;  // org: the record the ray leaves from
    pool_ray<true>(B, cur, ctr, i, o, d, sid, org);
    auto walk = [&](uint32_t s) {
      const SceneView<float> S = scene_view<FEAT>(B, s);
      LdsStack<SS, BS, kPinRecords<FEAT>> st{lds, spill, threadIdx.x, gid, nthreads, 0};
      float t = INFINITY;
      uint32_t idx = 0;
      bool h = trace_pass<float, FEAT, false, PASS_MAIN>(S, o, d, 0.001f, t, idx, st, COUNT ? tc : nullptr, org);
      if (h)
        store_hit<FEAT>(S, hit, i, o, d, t, idx);
      else
        store_miss(hit, i);
    };
    // a wave whose rays share one scene (the usual case) walks with the
    // scene's index in a scalar register: the view and the node and record
    // base addresses are then scalar; a wave of several scenes walks them
    // together with per-lane views.  (The kernels with FP64 islands or
    // transforms keep one walk: random-sphere extend 43.3 -> 44.5 ms with
    // both.)
    const uint32_t s0 = uint32_t(__builtin_amdgcn_readfirstlane(int(sid)));
    if (kPinRecords<FEAT> && __builtin_amdgcn_ballot_w64(sid != s0) == 0) {
      uint32_t su = s0;
      asm volatile("" : "+s"(su));
      walk(su);
    } else {
      for_scene(sid, walk);
    }
  }
}
// The traversal kernels are latency-bound: the register budget is capped at
// 8 waves per SIMD (64 VGPRs), which outweighs the few spills it costs.  With
// the 32-word LDS stack, LDS holds a CU to 20 waves (5 per SIMD) whatever the
// registers: the budget is then that of 5 waves (102 VGPRs).
#define LMN_TRAVERSAL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SS == 16 ? 8 : 5, SS == 16 ? 8 : 5)))

template <int SS, int BS, uint32_t FEAT>
__global__ __launch_bounds__(BS) LMN_TRAVERSAL_OCCUPANCY void k_extend(BatchDev B, PathSoA cur, HitSoA hit,
                                               const WaveCounters* __restrict__ ctr, uint32_t* spill) {
  extend_body<SS, BS, FEAT, false>(B, cur, hit, ctr, spill, nullptr);
}
// Instrumented extend (counts node fetches and primitive tests): a kernel of
// its own name so that profiles keep it apart from the timed k_extend.
template <int SS, int BS, uint32_t FEAT>
__global__ __launch_bounds__(BS) LMN_TRAVERSAL_OCCUPANCY void k_extend_count(BatchDev B, PathSoA cur, HitSoA hit,
                                                     const WaveCounters* __restrict__ ctr, uint32_t* spill,
                                                     TraceCounters* tc) {
  extend_body<SS, BS, FEAT, true>(B, cur, hit, ctr, spill, tc);
}

// Shadow rays: any hit over [0.001, dist - 0.001] (src/render.rs:1457-1458),
// main BVH then planes.  Without marched kinds the contribution of a visible
// light sample is added here; with them an occluded ray's distance is
// negated and the marched pass finishes the query.
//
// Persistent: a grid of about one workgroup per resident slot walks all the
// stage's rays, and a lane whose ray has ended takes its wave's next ray, so
// a wave no longer waits for its longest ray before its lanes get new work.
// Visibility rays end at their first hit and carry no surface work, so the
// refill pays (mesh-heavy shadow stage 25.6 -> 19.8 ms); closest-hit rays
// end with a surface record, and running that per refill round cost more
// than the refill saved (extend kept grid-stride; a persistent extend
// measured 167 -> 253 ms per batch step).  The rays of wave w are
// blocks w, w + W, w + 2W, ... of 64 (W waves in the grid), as a
// grid-stride kernel deals them out; walk steps run in rounds, and a round
// ends once half of the lanes that began it have finished (a quarter:
// batch shadow stage 53.7 ms, half: 53.0, an eighth: 57.3).  Every
// ray's walk is the grid-stride walk, step for step.  Exit: a lane whose
// next ray lies past the stage's count is done; the wave leaves when all its
// lanes are.
// MAIN: the features of the main BVH's walk (FEAT without FEAT_MARCH, and
// without FEAT_F64 where no scene has FP64 islands in its main BVH).
template <int SS, int BS, uint32_t FEAT, uint32_t MAIN>
__global__ __launch_bounds__(BS) LMN_TRAVERSAL_OCCUPANCY void k_shadow(BatchDev B, ShadowSoA sh,
                                                                       float4* __restrict__ samples,
                                                                       const WaveCounters* __restrict__ ctr,
                                                                       uint32_t* spill) {
  __shared__ uint32_t lds[SS * BS];
  constexpr bool DEFER = (FEAT & FEAT_MARCH) != 0;
  constexpr uint32_t kIdle = 0xffffffffu, kDone = 0xfffffffeu;
  const uint32_t n = ctr->n_shadow;
  const uint32_t nthreads = gridDim.x * BS;
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  const uint64_t n_waves = nthreads / 64u, wave = gid / 64u;
  uint64_t taken = 0;  // the wave's rays handed out so far (wave-uniform)
  LdsStack<SS, BS> st{lds, spill, threadIdx.x, gid, nthreads, 0};
  uint32_t i = kIdle, sid = 0, ref = 0, best = 0;
  V3<float> o = {0.f, 0.f, 0.f}, d = {0.f, 0.f, 0.f}, inv = {0.f, 0.f, 0.f};
  float closest = 0.f, dist = 0.f;
  bool found = false;
  for (;;) {
    const bool need = i == kIdle;
    const unsigned long long want = __ballot(need);
    if (need) {
      const uint64_t k = taken + uint64_t(__popcll(want & __lanemask_lt()));
      const uint64_t r = (wave + (k >> 6) * n_waves) * 64u + (k & 63u);
      i = r < n ? uint32_t(r) : kDone;
      if (i != kDone) {
        const float4 o4 = sh.o[i], d4 = sh.d[i];
        o = f3(o4);
        d = f3(d4);
        dist = o4.w;
        closest = o4.w - 0.001f;
        sid = __float_as_uint(sh.c[i].w);
        inv = slab_inv(d);
        ref = 0;
        best = 0;
        found = false;
        st.sp = 0;
      }
    }
    taken += uint64_t(__popcll(want));
    if (__ballot(i != kDone) == 0ull) return;
    if (i == kDone) continue;
    bool fin = false;
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = scene_view<MAIN>(B, s);
      uint32_t nf = 0, pt = 0, tt = 0;
      const uint32_t began = uint32_t(__popcll(__ballot(true)));
      for (;;) {
        if (!fin)
          fin = S.n_nodes4 == 0 ||
                bvh4_step<MAIN, true>(S, o, d, inv, 0.001f, closest, best, found, ref, st, nf, pt, tt, false);
        const uint32_t walking = uint32_t(__popcll(__ballot(!fin)));
        if (walking == 0u || 2u * walking <= began) break;
      }
      if (!fin) return;
      if (!found)  // the unbounded planes after the BVH (src/scene.rs:817-834)
        for (uint32_t k = 0; k < S.n_planes; ++k)
          if (test_prim<float, MAIN>(S, S.n_bounded + k, o, d, 0.001f, closest)) {
            found = true;
            break;
          }
      if (DEFER) {
        if (found) sh.o[i].w = -dist;
      } else if (!found) {
        add_sample(samples, __float_as_uint(sh.d[i].w), f3(sh.c[i]));
      }
    });
    if (fin) i = kIdle;
  }
}

// ---------------------------------------------------------------------------
// Marched pass (scenes with torus / superellipsoid / spring records): the
// main pass leaves each ray's hit distance (extend) or visibility (shadow);
// the marched kinds' FP64 root finders run after it, compacted:
//   collect  - every ray walks the marched BVH (FP32 boxes up to that
//              distance) and queues (ray, record) per kind;
//   test     - one kernel per kind, one candidate per lane: closest hit per
//              ray by a 64-bit atomic min on (t bits, record), or the
//              occluded flag;
//   fallback - rays with too many candidates, traced whole as before;
//   finish   - surface record of the marched hit / the shadow ray's sample.
// Tested in place, the root finders ran on the few lanes of a wave whose
// rays had entered a marched box, with the rest idle (8 % lane use).

// Ray `i` of the extend (path pool) or shadow (queue) stage: origin,
// direction, scene, range end; `live` false for shadow rays the main pass
// found occluded.
template <bool SHADOW>
__device__ __forceinline__ bool march_ray(const BatchDev& B, const PathSoA& cur, const HitSoA& hit,
                                          const ShadowSoA& sh, const WaveCounters* ctr, uint32_t i, V3<float>& o,
                                          V3<float>& d, uint32_t& sid, float& tmax) {
  if (SHADOW) {
    float4 o4 = sh.o[i], d4 = sh.d[i];
    o = f3(o4);
    d = f3(d4);
    sid = __float_as_uint(sh.c[i].w);
    tmax = o4.w - 0.001f;
    return o4.w > 0.f;
  }
  uint32_t org;
  pool_ray(B, cur, ctr, i, o, d, sid, org);
  const float4 h = hit.b[i];
  tmax = __float_as_uint(h.z) == HIT_MISS ? INFINITY : h.w;
  return true;
}

template <int SS, int BS, uint32_t FEAT, bool SHADOW>
__global__ __launch_bounds__(BS) void k_march_collect(BatchDev B, PathSoA cur, HitSoA hit, ShadowSoA sh, MarchQ q,
                                                      WaveCounters* __restrict__ ctr, uint32_t* spill) {
  __shared__ uint32_t lds[SS * BS];
  __shared__ uint32_t cnt[3], base[3];
  const uint32_t n = SHADOW ? ctr->n_shadow : ctr->active;
  const uint32_t nthreads = gridDim.x * BS;
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  uint32_t* qn = ctr->n_cand[SHADOW ? 1 : 0];
  // block-uniform rounds: every lane reaches the barriers
  for (uint32_t r0 = blockIdx.x * BS; r0 < n; r0 += nthreads) {
    const uint32_t i = r0 + threadIdx.x;
    uint32_t c[MARCH_MAX_CAND];  // record | kind << 30
    uint32_t nc = 0, nk0 = 0, nk1 = 0, nk2 = 0;
    bool over = false;
    V3<float> o, d;
    uint32_t sid;
    float tmax;
    if (i < n && march_ray<SHADOW>(B, cur, hit, sh, ctr, i, o, d, sid, tmax)) {
      q.key[i] = SHADOW ? 0ull : MARCH_EMPTY;
      for_scene(sid, [&](uint32_t s) {
        const SceneView<float> S = uniform_load(B.views, s);
        if (S.n_mnodes == 0) return;
        LdsStack<SS, BS> st{lds, spill, threadIdx.x, gid, nthreads, 0};
        V3<float> inv = {1.f / d.x, 1.f / d.y, 1.f / d.z};
        bvh_leaves(S.mnodes, o, inv, 0.001f, tmax, st, [&](uint32_t rec) {
          if (nc == MARCH_MAX_CAND) {
            over = true;
            return;
          }
          const uint32_t kind = S.prims[rec].kind_mat & KIND_MASK;
          if (is_quadric_kind(kind)) {  // disk ... annulus: the fallback traces the ray whole
            over = true;
            return;
          }
          uint32_t k = kind - K_TORUS;
          c[nc++] = rec | (k << 30);
          nk0 += k == 0u;
          nk1 += k == 1u;
          nk2 += k == 2u;
        });
      });
      if (over) nc = nk0 = nk1 = nk2 = 0;
    }
    // slots: LDS counts per kind, then one device atomic per kind per block
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t off0 = nk0 ? atomicAdd(&cnt[0], nk0) : 0u;
    uint32_t off1 = nk1 ? atomicAdd(&cnt[1], nk1) : 0u;
    uint32_t off2 = nk2 ? atomicAdd(&cnt[2], nk2) : 0u;
    __syncthreads();
    if (threadIdx.x < 3) base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&qn[threadIdx.x], cnt[threadIdx.x]) : 0u;
    __syncthreads();
    off0 += base[0];
    off1 += base[1];
    off2 += base[2];
    const bool fits = off0 + nk0 <= q.cap && off1 + nk1 <= q.cap && off2 + nk2 <= q.cap;
    for (uint32_t j = 0; j < nc; ++j) {
      uint32_t k = c[j] >> 30;
      uint32_t slot = k == 0u ? off0++ : (k == 1u ? off1++ : off2++);
      if (slot < q.cap) q.cand[size_t(k) * q.cap + slot] = fits ? make_uint2(i, c[j] & 0x3fffffffu) : make_uint2(MARCH_SKIP, 0u);
    }
    // only real rays with candidates that did not fit (or too many) go to the fallback;
    // lanes past n and rays without candidates have nothing to lose
    if (i < n && (over || (nc != 0u && !fits))) q.ovf[atomicAdd(&ctr->n_ovf[SHADOW ? 1 : 0], 1u)] = i;
    __syncthreads();  // cnt / base are rewritten next round
  }
}

// The root finders would take every register a wave can have (one wave per
// SIMD); capped at 4 waves per SIMD they spill some and run faster.
#define LMN_MARCH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 8)))

// Superellipsoid candidates: the exact walk's FP32 screen alone (down to
// level 8); candidates it cannot rule out move to region 3.  The walk's
// cost spreads from one node (a chord missing the body) to hundreds (rays
// grazing it), and a wave takes as long as its slowest lane: screening
// first leaves the exact kernel the candidates of similar, higher cost.
template <int BS, uint32_t FEAT, bool SHADOW>
__global__ __launch_bounds__(BS) void k_march_screen(BatchDev B, PathSoA cur, HitSoA hit, ShadowSoA sh, MarchQ q,
                                                     WaveCounters* __restrict__ ctr) {
  __shared__ uint32_t cnt, base;
  uint32_t* qn = ctr->n_cand[SHADOW ? 1 : 0];
  const uint32_t nq = qn[1];
  const uint32_t n = nq < q.cap ? nq : q.cap;
  const uint2* in = q.cand + size_t(1) * q.cap;
  uint2* out = q.cand + size_t(3) * q.cap;
  const uint32_t nthreads = gridDim.x * BS;
  for (uint32_t r0 = blockIdx.x * BS; r0 < n; r0 += nthreads) {  // block-uniform rounds
    const uint32_t j = r0 + threadIdx.x;
    uint2 e = make_uint2(MARCH_SKIP, 0u);
    bool keep = false;
    if (j < n) e = in[j];
    if (e.x != MARCH_SKIP) {
      V3<float> o, d;
      uint32_t sid;
      float tmax;
      march_ray<SHADOW>(B, cur, hit, sh, ctr, e.x, o, d, sid, tmax);
      for_scene(sid, [&](uint32_t s) {
        const SceneView<float> S = uniform_load(B.views, s);
        const MarchedRay r = marched_ray<FEAT>(S, e.y, o, d);
        keep = S.exact_marched == 0 || se_may_cross(S.prims64[e.y].p, r.o64, r.d64, 0.001, double(tmax), 8);
      });
    }
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    uint32_t off = keep ? atomicAdd(&cnt, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0) base = cnt ? atomicAdd(&qn[3], cnt) : 0u;
    __syncthreads();
    if (keep) out[base + off] = e;
    __syncthreads();  // cnt / base are rewritten next round
  }
}

template <int BS, uint32_t FEAT, uint32_t KIND, bool SHADOW>
__global__ __launch_bounds__(BS) LMN_MARCH_OCCUPANCY void k_march_test(BatchDev B, PathSoA cur, HitSoA hit,
                                                                      ShadowSoA sh, MarchQ q,
                                                                      const WaveCounters* __restrict__ ctr) {
  // superellipsoids: the screened region
  constexpr uint32_t region = KIND == K_SUPERELLIPSOID ? 3u : KIND - K_TORUS;
  const uint32_t nq = ctr->n_cand[SHADOW ? 1 : 0][region];
  const uint32_t n = nq < q.cap ? nq : q.cap;
  const uint2* cand = q.cand + size_t(region) * q.cap;
  const uint32_t nthreads = gridDim.x * BS;
  for (uint32_t j = blockIdx.x * BS + threadIdx.x; j < n; j += nthreads) {
    const uint2 e = cand[j];
    if (e.x == MARCH_SKIP) continue;
    V3<float> o, d;
    uint32_t sid;
    float tmax;
    march_ray<SHADOW>(B, cur, hit, sh, ctr, e.x, o, d, sid, tmax);
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = uniform_load(B.views, s);
      float t = tmax;
      if (test_marched<KIND, FEAT, SHADOW>(S, e.y, o, d, 0.001f, t)) {
        if (SHADOW)
          q.key[e.x] = 1ull;
        else
          atomicMin(&q.key[e.x], (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | e.y);
      }
    });
  }
}


// Superellipsoid candidates of the extend stage, exact mode: the interval
// walk up to the bracketing leaf; brackets are queued for k_march_refine, so
// that the walk's lanes and the refinement's Newton iterations each run
// with like work across the wave.  Reference mode, and brackets beyond the
// queue, are finished in place.
template <int BS, uint32_t FEAT>
__global__ __launch_bounds__(BS) LMN_MARCH_OCCUPANCY void k_march_bracket(BatchDev B, PathSoA cur, HitSoA hit,
                                                                         ShadowSoA sh, MarchQ q,
                                                                         WaveCounters* __restrict__ ctr) {
  const uint32_t nq = ctr->n_cand[0][3];
  const uint32_t n = nq < q.cap ? nq : q.cap;
  const uint2* cand = q.cand + size_t(3) * q.cap;
  const uint32_t nthreads = gridDim.x * BS;
  for (uint32_t j = blockIdx.x * BS + threadIdx.x; j < n; j += nthreads) {
    const uint2 e = cand[j];
    if (e.x == MARCH_SKIP) continue;
    V3<float> o, d;
    uint32_t sid;
    float tmax;
    march_ray<false>(B, cur, hit, sh, ctr, e.x, o, d, sid, tmax);
    SeBracket br{};
    bool queue = false;
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = uniform_load(B.views, s);
      float t = tmax;
      if (S.exact_marched == 0) {
        if (test_marched<K_SUPERELLIPSOID, FEAT>(S, e.y, o, d, 0.001f, t))
          atomicMin(&q.key[e.x], (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | e.y);
        return;
      }
      const MarchedRay r = marched_ray<FEAT>(S, e.y, o, d);
      queue = se_exact_bracket(S.prims64[e.y].p, r.o64, r.d64, 0.001, double(tmax), br);
    });
    const uint32_t slot = wave_append(&ctr->n_brk, queue);
    if (!queue) continue;
    if (slot < q.cap_brk) {
      q.brk[slot] = MarchBracket{e.x, e.y, 0u, 0u, br.a, br.b, br.fa, br.fb};
      continue;
    }
    for_scene(sid, [&](uint32_t s) {  // queue full: refine here
      const SceneView<float> S = uniform_load(B.views, s);
      const MarchedRay r = marched_ray<FEAT>(S, e.y, o, d);
      double t64;
      float t = tmax;
      if (se_refine(S.prims64[e.y].p, r.o64, r.d64, br, 0.001, double(tmax), t64) &&
          marched_accept<FEAT>(S, e.y, r, d, t64, t))
        atomicMin(&q.key[e.x], (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | e.y);
    });
  }
}

template <int BS, uint32_t FEAT>
__global__ __launch_bounds__(BS) void k_march_refine(BatchDev B, PathSoA cur, HitSoA hit, ShadowSoA sh, MarchQ q,
                                                     const WaveCounters* __restrict__ ctr) {
  const uint32_t nb = ctr->n_brk;
  const uint32_t n = nb < q.cap_brk ? nb : q.cap_brk;
  const uint32_t nthreads = gridDim.x * BS;
  for (uint32_t j = blockIdx.x * BS + threadIdx.x; j < n; j += nthreads) {
    const MarchBracket m = q.brk[j];
    V3<float> o, d;
    uint32_t sid;
    float tmax;
    march_ray<false>(B, cur, hit, sh, ctr, m.ray, o, d, sid, tmax);
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = uniform_load(B.views, s);
      const MarchedRay r = marched_ray<FEAT>(S, m.rec, o, d);
      double t64;
      float t = tmax;
      if (se_refine(S.prims64[m.rec].p, r.o64, r.d64, SeBracket{m.a, m.b, m.fa, m.fb}, 0.001, double(tmax), t64) &&
          marched_accept<FEAT>(S, m.rec, r, d, t64, t))
        atomicMin(&q.key[m.ray], (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | m.rec);
    });
  }
}

template <int SS, int BS, uint32_t FEAT, bool SHADOW>
__global__ __launch_bounds__(BS) LMN_MARCH_OCCUPANCY void k_march_fallback(BatchDev B, PathSoA cur, HitSoA hit,
                                                                          ShadowSoA sh, MarchQ q,
                                                                          const WaveCounters* __restrict__ ctr,
                                                                          uint32_t* spill) {
  __shared__ uint32_t lds[SS * BS];
  const uint32_t n = ctr->n_ovf[SHADOW ? 1 : 0];
  const uint32_t nthreads = gridDim.x * BS;
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  for (uint32_t j = gid; j < n; j += nthreads) {
    const uint32_t i = q.ovf[j];
    V3<float> o, d;
    uint32_t sid;
    float tmax;
    march_ray<SHADOW>(B, cur, hit, sh, ctr, i, o, d, sid, tmax);
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = uniform_load(B.views, s);
      LdsStack<SS, BS> st{lds, spill, threadIdx.x, gid, nthreads, 0};
      float t = tmax;
      uint32_t idx = 0;
      if (trace_pass<float, FEAT, SHADOW, PASS_MARCH>(S, o, d, 0.001f, t, idx, st)) {
        if (SHADOW)
          q.key[i] = 1ull;
        else
          atomicMin(&q.key[i], (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | idx);
      }
    });
  }
}

template <int BS, uint32_t FEAT, bool SHADOW>
__global__ __launch_bounds__(BS) void k_march_finish(BatchDev B, PathSoA cur, HitSoA hit, ShadowSoA sh,
                                                     float4* __restrict__ samples, MarchQ q,
                                                     const WaveCounters* __restrict__ ctr) {
  const uint32_t n = SHADOW ? ctr->n_shadow : ctr->active;
  const uint32_t nthreads = gridDim.x * BS;
  for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < n; i += nthreads) {
    if (SHADOW) {
      if (sh.o[i].w > 0.f && q.key[i] == 0ull) {
        float4 c = sh.c[i];
        add_sample(samples, __float_as_uint(sh.d[i].w), f3(c));
      }
      continue;
    }
    const unsigned long long key = q.key[i];
    if (key == MARCH_EMPTY) continue;
    V3<float> o, d;
    uint32_t sid, org;
    pool_ray(B, cur, ctr, i, o, d, sid, org);
    for_scene(sid, [&](uint32_t s) {
      const SceneView<float> S = uniform_load(B.views, s);
      store_hit<FEAT>(S, hit, i, o, d, __uint_as_float(uint32_t(key >> 32)), uint32_t(key));
    });
  }
}

// Shade: one recursion level of ray_color (src/render.rs:1472-1537) per
// path -- emission, scatter, Russian roulette, NEE light sample -- then
// wave-ballot compaction of the shadow queue and of the surviving paths.
// Shading waits on memory (the path state, hit record, material, texels and
// the sample's running sum) far more than it computes: it is built for 6
// waves per SIMD (80 VGPRs) rather than the 5 its registers would give
// (batch shade stage 65.8 -> 56.2 ms; at 8 waves the spills cost more than
// the extra waves gain).
template <int BS, uint32_t FEAT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_shade(BatchDev B, PathSoA cur, HitSoA hit, PathSoA nxt,
                                              ShadowSoA sh, float4* __restrict__ samples,
                                              WaveCounters* __restrict__ ctr) {
  __shared__ uint32_t res_lds[2][64 + (BS / 64) + 2];
  const uint32_t n = ctr->active, ns = ctr->n_surv;
  const uint64_t cam0 = ctr->cam_item;
  uint32_t nthreads = gridDim.x * BS;
  uint32_t round = 0;
  // block-uniform trip count: every lane reaches the reservation barriers
  for (uint32_t base = blockIdx.x * BS; base < n; base += nthreads, ++round) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    uint32_t j = valid ? i : base;  // invalid lanes read a valid slot and do nothing with it
    // the path: stored (a survivor) or a new camera ray, recomputed
    const bool camera = valid && j >= ns;
    V3<float> o, d, beta = {1.f, 1.f, 1.f};
    Rng rng;
    uint32_t item, depth, sid;
    bool skip = false;
    float cam_w = 0.f;
    if (!camera) {
      const float4 A = cur.ob[j], Bd = cur.db[j];
      const uint4 M = cur.meta[j], X = cur.aux[j];
      o = f3(A);
      d = f3(Bd);
      beta = {__uint_as_float(X.y), __uint_as_float(X.z), __uint_as_float(X.w)};
      rng.key = (uint64_t(M.w) << 32) | M.z;
      rng.n = X.x;
      item = M.x;
      depth = M.y & 0xffffu;
      sid = __float_as_uint(A.w);
      skip = (M.y >> 16) & 1u;
    } else {
      CameraPath c;
      const uint64_t it = cam0 + (j - ns);
      camera_path(B, it, c);
      o = c.o;
      d = c.d;
      rng = c.rng;
      item = uint32_t(it);
      depth = c.max_depth;
      sid = c.scene;
      cam_w = c.w;
    }
    float4 HA = hit.a[j], HB = hit.b[j];
    uint32_t mword = __float_as_uint(HB.z);
    const uint32_t rec = mword == HIT_MISS ? kNoOrigin : hit.rec[j];
    V3<float> add = {0.f, 0.f, 0.f};
    bool cont = false, emit = false;
    V3<float> sp_o = {0.f, 0.f, 0.f}, sp_d = {0.f, 0.f, 0.f}, ndir = {0.f, 0.f, 0.f}, nbeta = beta;
    V3<float> cb = {0.f, 0.f, 0.f};
    float sp_dist = 0.f;
    bool use_nee = false;
    if (valid) for_scene(sid, [&](uint32_t sc) {
      const SceneView<float> S = uniform_load(B.views, sc);
      if (mword == HIT_MISS) {
        add = hadamard(beta, background(S.bg, d));
        return;
      }
      Surf<float> s;
      s.p = surface_point(o, d, HA.w);
      s.n = {HA.x, HA.y, HA.z};
      s.t = HB.w;
      s.u = HB.x;
      s.v = HB.y;
      s.front = (mword >> 31) != 0;
      s.mat = mword & 0x7fffffffu;
      s.prim_id = 0;
      MaterialT<float> mcopy;
      const MaterialT<float>& m = record_copy<true>(S.materials, s.mat, mcopy);
      V3<float> em = skip ? V3<float>{0.f, 0.f, 0.f} : emitted(S, m, s.u, s.v);
      V3<float> att;
      if (!scatter(S, m, d, s, rng, ndir, att)) {
        add = hadamard(beta, em);
        return;
      }
      // Russian roulette on remaining depth < 47 (src/render.rs:1494-1502)
      float surv = clamp_r(max3(att), 0.05f, 1.0f);
      bool rr = depth < 47;
      if (rr && rng.next<float>() > surv) {
        add = hadamard(beta, em);
        return;
      }
      float rrw = rr ? 1.0f / surv : 1.0f;
      use_nee = !is_specular(m) && S.n_lights > 0;
      V3<float> c = {0.f, 0.f, 0.f};
      bool nee_ok = false;
      if (use_nee) nee_ok = nee_sample<float, true>(S, s.p, s.n, att, rng, sp_d, sp_dist, c);
      V3<float> attw = att * rrw;
      // NaN guard of one recursion level (src/render.rs:1526-1531)
      if (has_nan(em) || (nee_ok && has_nan(c)) || has_nan(attw)) return;
      add = hadamard(beta, em);
      emit = nee_ok;
      cb = hadamard(beta, c);
      sp_o = s.p;
      nbeta = hadamard(beta, attw);
      cont = depth > 1;
    });
    // a camera path's first vertex starts its sample (radiance, filter
    // weight); later vertices add to it
    if (camera)
      samples[item] = make_float4(add.x, add.y, add.z, cam_w);
    else if (add.x != 0.f || add.y != 0.f || add.z != 0.f)
      add_sample(samples, item, add);
    uint32_t ps, pn;
    const uint32_t octant = (ndir.x < 0.f ? 1u : 0u) | (ndir.y < 0.f ? 2u : 0u) | (ndir.z < 0.f ? 4u : 0u);
    block_reserve2_keyed<BS>(&ctr->n_shadow, &ctr->next_active, emit, cont, octant, ps, pn, res_lds[round & 1u]);
    if (emit) {
      sh.o[ps] = make_float4(sp_o.x, sp_o.y, sp_o.z, sp_dist);
      sh.d[ps] = make_float4(sp_d.x, sp_d.y, sp_d.z, __uint_as_float(item));
      sh.c[ps] = make_float4(cb.x, cb.y, cb.z, __uint_as_float(sid));
    }
    if (cont) {
      nxt.ob[pn] = make_float4(sp_o.x, sp_o.y, sp_o.z, __uint_as_float(sid));
      nxt.db[pn] = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(rec));
      nxt.meta[pn] = make_uint4(item, (depth - 1) | (use_nee ? 0x10000u : 0u), uint32_t(rng.key),
                                uint32_t(rng.key >> 32));
      nxt.aux[pn] = make_uint4(rng.n, __float_as_uint(nbeta.x), __float_as_uint(nbeta.y), __float_as_uint(nbeta.z));
    }
  }
}

// Bookkeeping between iterations: the survivors stay in their compacted
// slots and new work items (camera rays) take the free slots after them, as
// many as fit in the pool; the stages recompute those rays from their items.
__global__ void k_book(WaveCounters* ctr, uint32_t pool) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t na = ctr->next_active;
  uint64_t left = ctr->n_items - ctr->next_item;
  uint32_t room = pool - na;
  uint32_t m = left < room ? uint32_t(left) : room;
  ctr->stat_extend += ctr->active;
  ctr->stat_shadow += ctr->n_shadow;
  ctr->stat_camera += m;
  ctr->active = na + m;
  ctr->n_surv = na;
  ctr->cam_item = ctr->next_item;
  ctr->next_item += m;
  ctr->next_active = 0;
  ctr->n_shadow = 0;
  ctr->iterations += 1;
  for (int k = 0; k < 4; ++k) ctr->n_cand[0][k] = ctr->n_cand[1][k] = 0;
  ctr->n_ovf[0] = ctr->n_ovf[1] = 0;
  ctr->n_brk = 0;
}

__global__ void k_counters_init(WaveCounters* ctr, uint64_t n_items) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctr->active = 0;
  ctr->next_active = 0;
  ctr->n_shadow = 0;
  ctr->n_surv = 0;
  ctr->cam_item = 0;
  ctr->next_item = 0;
  ctr->n_items = n_items;
  ctr->stat_extend = 0;
  ctr->stat_shadow = 0;
  ctr->stat_camera = 0;
  ctr->iterations = 0;
  for (int k = 0; k < 4; ++k) ctr->n_cand[0][k] = ctr->n_cand[1][k] = 0;
  ctr->n_ovf[0] = ctr->n_ovf[1] = 0;
  ctr->n_brk = 0;
}

// ---------------------------------------------------------------------------
// Resolve: per-sample luminance clamp (src/render.rs:1691-1697) and filter
// weighting, accumulated per pixel in FP64 in sample order.
__global__ void k_resolve(const float4* __restrict__ samples, double4* __restrict__ acc, uint32_t npix,
                          uint32_t n_samples, const uint32_t* __restrict__ list, uint32_t tile_w) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npix) return;
  const uint32_t p = list ? list[e] : tile_w ? tile_pixel(e, tile_w) : e;
  double4 a = acc[p];
  auto add = [&](float4 v) {
    double x = v.x, y = v.y, z = v.z, w = v.w;
    double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
    if (lum > 100.0) {
      double f = 100.0 / lum;
      x *= f;
      y *= f;
      z *= f;
    }
    a.x += x * w;
    a.y += y * w;
    a.z += z * w;
    a.w += w;
  };
  // a launch covers one part's pixels (often about one wave per SIMD), so
  // the sample loads go out eight at a time: one memory round trip per eight
  // samples instead of one per sample; the sums keep sample order
  uint32_t s = 0;
  for (; s + 8u <= n_samples; s += 8u) {
    float4 v[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) v[k] = samples[size_t(s + k) * npix + e];
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) add(v[k]);
  }
  for (; s < n_samples; ++s) add(samples[size_t(s) * npix + e]);
  acc[p] = a;
}

// Adaptive sampling (src/render.rs:1626-1677): each listed pixel takes its
// round's samples in stratum order, luminance-clamped, into a Welford update
// of the luminance and a running mean of the colour, and stops at the first
// n >= max(spp/4, 16), n > 1 whose standard error sqrt(var / n) is below the
// threshold.  Samples of the round after that point are left unused, so the
// pixel's value is the mean of exactly the samples the sequential loop takes.
__global__ void k_resolve_adaptive(const float4* __restrict__ samples, AdaptState* __restrict__ state,
                                   const uint32_t* __restrict__ list, uint32_t n_list, uint32_t n_samples,
                                   uint32_t min_samples, double threshold) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_list) return;
  const uint32_t p = list[e];
  AdaptState a = state[p];
  for (uint32_t s = 0; s < n_samples && !a.done; ++s) {
    float4 v = samples[size_t(s) * n_list + e];
    adapt_take(a, v.x, v.y, v.z, min_samples, threshold);
  }
  state[p] = a;
}
// Pixels still sampling after a round: listed again (any order: a pixel's
// samples depend on its index and sample number only).
__global__ void k_adapt_compact(const AdaptState* __restrict__ state, const uint32_t* __restrict__ list_in,
                                uint32_t n_in, uint32_t spp, uint32_t* __restrict__ list_out,
                                uint32_t* __restrict__ count) {
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_in) return;
  const uint32_t p = list_in[e];
  const AdaptState a = state[p];
  if (!a.done && a.n < spp) list_out[atomicAdd(count, 1u)] = p;
}
__global__ void k_adapt_finalize(const AdaptState* __restrict__ state, float* __restrict__ out, uint32_t npix,
                                 unsigned long long* __restrict__ taken) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const AdaptState a = state[p];
  out[size_t(p) * 3 + 0] = float(a.mx);
  out[size_t(p) * 3 + 1] = float(a.my);
  out[size_t(p) * 3 + 2] = float(a.mz);
  atomicAdd(taken, (unsigned long long)a.n);
}
__global__ void k_adapt_init(AdaptState* __restrict__ state, uint32_t* __restrict__ list, uint32_t npix) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  state[p] = AdaptState{0.0, 0.0, 0.0, 0.0, 0.0, 0u, 0u};
  list[p] = p;
}

__global__ void k_zero_acc(double4* acc, uint32_t npix) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npix) acc[p] = make_double4(0, 0, 0, 0);
}

__global__ void k_finalize(const double4* __restrict__ acc, float* __restrict__ out, uint32_t npix) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  double4 a = acc[p];
  double inv = a.w > 0.0 ? 1.0 / a.w : 1.0;
  out[size_t(p) * 3 + 0] = float(a.x * inv);
  out[size_t(p) * 3 + 1] = float(a.y * inv);
  out[size_t(p) * 3 + 2] = float(a.z * inv);
}

// ---------------------------------------------------------------------------
// Launchers

template <int BS, uint32_t FEAT, bool SHADOW>
static void launch_march_tests(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur) {
  if (L.mq.kinds & 1u)
    hipLaunchKernelGGL((k_march_test<BS, FEAT, K_TORUS, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit,
                       L.sh, L.mq, L.ctr);
  if (L.mq.kinds & 2u) {
    hipLaunchKernelGGL((k_march_screen<BS, FEAT, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.sh,
                       L.mq, L.ctr);
    if constexpr (SHADOW) {
      hipLaunchKernelGGL((k_march_test<BS, FEAT, K_SUPERELLIPSOID, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B,
                         cur, L.hit, L.sh, L.mq, L.ctr);
    } else {
      hipLaunchKernelGGL((k_march_bracket<BS, FEAT>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.sh, L.mq,
                         L.ctr);
      hipLaunchKernelGGL((k_march_refine<BS, FEAT>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.sh, L.mq,
                         L.ctr);
    }
  }
  if (L.mq.kinds & 4u)
    hipLaunchKernelGGL((k_march_test<BS, FEAT, K_SPRING, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit,
                       L.sh, L.mq, L.ctr);
}

template <int SS, int BS, uint32_t FEAT, bool SHADOW>
static void launch_march(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur) {
  hipLaunchKernelGGL((k_march_collect<SS, BS, FEAT, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.sh,
                     L.mq, L.ctr, L.spill);
  launch_march_tests<BS, FEAT, SHADOW>(L, B, cur);
  hipLaunchKernelGGL((k_march_fallback<SS, BS, FEAT, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit,
                     L.sh, L.mq, L.ctr, L.spill);
  hipLaunchKernelGGL((k_march_finish<BS, FEAT, SHADOW>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.sh,
                     L.samples, L.mq, L.ctr);
}

// Workgroups of the persistent shadow launch: as many as the device holds
// at once, at most the grid-stride grid (the spill area is sized for that).
template <int SS, int BS, uint32_t FEAT, uint32_t MAIN>
static uint32_t persistent_grid(uint32_t grid) {
  static const uint32_t resident = [] {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                       reinterpret_cast<const void*>(&k_shadow<SS, BS, FEAT, MAIN>), BS, 0);
    return uint32_t(std::max(1, cus) * std::max(1, per_cu));
  }();
  return std::max(1u, std::min(grid, resident));
}

// The main passes of one feature set: MAIN is FEAT without FEAT_MARCH.
template <int SS, int BS, uint32_t FEAT, uint32_t MAIN>
static void launch_main_pass(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, bool shadow) {
  if (shadow)
    hipLaunchKernelGGL((k_shadow<SS, BS, FEAT, MAIN>), dim3(persistent_grid<SS, BS, FEAT, MAIN>(L.grid)), dim3(BS), 0,
                       L.stream, B, L.sh, L.samples, L.ctr, L.spill);
  else if (L.tc)
    hipLaunchKernelGGL((k_extend_count<SS, BS, MAIN>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.ctr,
                       L.spill, L.tc);
  else
    hipLaunchKernelGGL((k_extend<SS, BS, MAIN>), dim3(L.grid), dim3(BS), 0, L.stream, B, cur, L.hit, L.ctr, L.spill);
}
template <int SS, int BS, uint32_t FEAT>
static void launch_stage_ss(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, bool shadow,
                            hipEvent_t mid) {
  constexpr uint32_t MAIN = FEAT & ~FEAT_MARCH;
  constexpr bool MARCH = (FEAT & FEAT_MARCH) != 0;
  // FEAT_MARCH brings FEAT_F64 for the marched pass; the main passes carry
  // the FP64 sphere test (and the registers it takes) only for scenes with
  // FP64 islands in their main BVH
  if (MARCH && (MAIN & FEAT_F64) && !(L.feat_main & FEAT_F64))
    launch_main_pass<SS, BS, FEAT, MAIN & ~FEAT_F64>(L, B, cur, shadow);
  else
    launch_main_pass<SS, BS, FEAT, MAIN>(L, B, cur, shadow);
  if (shadow) {
    if (mid) (void)hipEventRecord(mid, L.stream);
    if constexpr (MARCH) launch_march<SS, BS, FEAT, true>(L, B, cur);
    return;
  }
  if (mid) (void)hipEventRecord(mid, L.stream);
  if constexpr (MARCH) launch_march<SS, BS, FEAT, false>(L, B, cur);
}
template <uint32_t FEAT>
static void launch_stage(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, bool shadow, hipEvent_t mid) {
  if (L.block == 128) {
    if (L.stack_lds == 32)
      launch_stage_ss<32, 128, FEAT>(L, B, cur, shadow, mid);
    else
      launch_stage_ss<16, 128, FEAT>(L, B, cur, shadow, mid);
  } else {
    if (L.stack_lds == 32)
      launch_stage_ss<32, 256, FEAT>(L, B, cur, shadow, mid);
    else
      launch_stage_ss<16, 256, FEAT>(L, B, cur, shadow, mid);
  }
}
// The FP32 path's feature sets: marched primitives are FP64 records too, so
// FEAT_MARCH always comes with FEAT_F64.
#define LMN_FEAT32_SWITCH(feat, CALL) \
  switch (feat) {                      \
    case 0: CALL(0); break;            \
    case 1: CALL(1); break;            \
    case 2: CALL(2); break;            \
    case 3: CALL(3); break;            \
    case 6: CALL(6); break;            \
    default: CALL(7); break;           \
  }

static void launch_trace_stage(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, bool shadow,
                               hipEvent_t mid) {
#define CALL(f) launch_stage<f>(L, B, cur, shadow, mid)
  LMN_FEAT32_SWITCH(L.feat, CALL)
#undef CALL
}
// Shading workgroups reserve their queue slots with one device atomic per
// queue and round (block_reserve2_keyed); 512 lanes per workgroup halve those
// atomics against 256 (batch shade stage 113 -> 69 ms); 1024 was slower.
constexpr int kShadeBlock = 512;
static void launch_shade(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, const PathSoA& nxt) {
  const uint32_t g = std::max<uint32_t>(1, uint32_t(uint64_t(L.grid) * L.block / kShadeBlock));
#define CALL(f)                                                                                                    \
  hipLaunchKernelGGL((k_shade<kShadeBlock, f>), dim3(g), dim3(kShadeBlock), 0, L.stream, B, cur, L.hit, nxt, L.sh, \
                     L.samples, L.ctr)
  LMN_FEAT32_SWITCH(L.feat, CALL)
#undef CALL
}

void launch_wave_iteration(const WaveLaunch& L, const BatchDev& B, const PathSoA& cur, const PathSoA& nxt,
                           hipEvent_t* ev) {
  // events: 0 extend 1 extend (marched) 2 shade 3 shadow 4 shadow (marched) 5 bookkeeping 6
  if (ev) (void)hipEventRecord(ev[0], L.stream);
  launch_trace_stage(L, B, cur, false, ev ? ev[1] : nullptr);
  if (ev) (void)hipEventRecord(ev[2], L.stream);
  launch_shade(L, B, cur, nxt);
  if (ev) (void)hipEventRecord(ev[3], L.stream);
  launch_trace_stage(L, B, cur, true, ev ? ev[4] : nullptr);
  if (ev) (void)hipEventRecord(ev[5], L.stream);
  hipLaunchKernelGGL(k_book, dim3(1), dim3(64), 0, L.stream, L.ctr, L.pool);
  if (ev) (void)hipEventRecord(ev[6], L.stream);
}

void launch_counters_init(WaveCounters* ctr, uint64_t n_items, hipStream_t s) {
  hipLaunchKernelGGL(k_counters_init, dim3(1), dim3(64), 0, s, ctr, n_items);
}
void launch_resolve(const float4* samples, double4* acc, uint32_t npix, uint32_t n_samples, uint32_t tile_w,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_resolve, dim3((npix + 255) / 256), dim3(256), 0, s, samples, acc, npix, n_samples,
                     (const uint32_t*)nullptr, tile_w);
}
void launch_resolve_list(const float4* samples, double4* acc, const uint32_t* list, uint32_t n_list,
                         uint32_t n_samples, hipStream_t s) {
  if (n_list)
    hipLaunchKernelGGL(k_resolve, dim3((n_list + 255) / 256), dim3(256), 0, s, samples, acc, n_list, n_samples, list,
                       0u);
}
void launch_resolve_adaptive(const float4* samples, AdaptState* state, const uint32_t* list, uint32_t n_list,
                             uint32_t n_samples, uint32_t min_samples, double threshold, hipStream_t s) {
  if (n_list)
    hipLaunchKernelGGL(k_resolve_adaptive, dim3((n_list + 255) / 256), dim3(256), 0, s, samples, state, list, n_list,
                       n_samples, min_samples, threshold);
}
void launch_adapt_compact(const AdaptState* state, const uint32_t* list_in, uint32_t n_in, uint32_t spp,
                          uint32_t* list_out, uint32_t* count, hipStream_t s) {
  if (n_in)
    hipLaunchKernelGGL(k_adapt_compact, dim3((n_in + 255) / 256), dim3(256), 0, s, state, list_in, n_in, spp,
                       list_out, count);
}
void launch_adapt_finalize(const AdaptState* state, float* out, uint32_t npix, unsigned long long* taken,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_adapt_finalize, dim3((npix + 255) / 256), dim3(256), 0, s, state, out, npix, taken);
}
void launch_adapt_init(AdaptState* state, uint32_t* list, uint32_t npix, hipStream_t s) {
  hipLaunchKernelGGL(k_adapt_init, dim3((npix + 255) / 256), dim3(256), 0, s, state, list, npix);
}
void launch_zero_acc(double4* acc, uint32_t npix, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_acc, dim3((npix + 255) / 256), dim3(256), 0, s, acc, npix);
}
void launch_finalize(const double4* acc, float* out, uint32_t npix, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3((npix + 255) / 256), dim3(256), 0, s, acc, out, npix);
}
}  // namespace lmn

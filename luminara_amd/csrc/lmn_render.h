// Device-side scene view, BVH traversal, surface reconstruction, materials,
// NEE light sampling, background and camera -- the per-ray work of the
// reference's integrator (src/render.rs:72-172, :1333-1537), in FP32 or FP64.
#pragma once
#include "lmn_geom.h"

namespace lmn {

enum : uint32_t { M_LAMBERTIAN = 0, M_METAL = 1, M_DIELECTRIC = 2, M_EMISSIVE = 3, M_MICROFACET = 4 };
enum : uint32_t { L_SPHERE = 0, L_RECT = 1, L_DISK = 2 };
enum : uint32_t { BG_SKY = 0, BG_SOLID = 1, BG_GRADIENT = 2, BG_SUN = 3 };

template <class R>
struct alignas(16) MaterialT {
  uint32_t kind;
  int32_t tex;
  R color[3];
  R fuzz, ior, roughness, dispersion, intensity, metallic;
  uint32_t param_slot;  // roughness parameter index for the adjoint, or ~0
  uint32_t pad;
};

struct TextureDev {
  uint32_t w, h;
  uint64_t offset;  // float offset into the texel pool (RGB triples)
};

template <class R>
struct alignas(16) LightT {
  uint32_t kind, mat;  // mat: emissive material it was registered from, ~0u if none
  V3<R> origin, u, v, normal, emission;
  R radius, area;
};

template <class R>
struct BackgroundT {
  uint32_t kind, pad;
  V3<R> c0, c1, dir_unit;
  R intensity;
};

template <class R>
struct CameraT {
  V3<R> origin, lower_left, horizontal, vertical, u, v;
  V3<R> ll_rel;  // lower_left - origin, formed in FP64
  R lens_radius;
};

template <class R>
struct SceneView {
  const Node4Q* nodes4;                     // main BVH, 4-wide, quantised boxes (FP32 path)
  const typename NodeSel<R>::type* nodes;   // main BVH (no marched kinds)
  const typename NodeSel<R>::type* mnodes;  // BVH over the marched kinds
  const PrimT<R>* prims;          // bounded prims (leaf order) then planes
  const PrimT<double>* prims64;   // FP64 copies, same order (FP64 islands)
  const XformT<R>* xforms;
  const XformT<double>* xforms64;
  const MaterialT<R>* materials;
  const TextureDev* textures;
  const float* texels;
  const LightT<R>* lights;
  const AttrT<R>* attrs;          // vertex attributes of smooth mesh triangles
  uint32_t n_nodes, n_mnodes, n_bounded, n_planes, n_lights;
  uint32_t n_nodes4;
  BackgroundT<R> bg;
  CameraT<R> cam;
  uint32_t exact_marched;  // LMN_ISECT_EXACT for marched prims
  uint32_t textured;       // a material reads an image texture (surface (u, v) are used)
};

// Traversal statistics of an instrumented launch (bytes per unit of work).
struct TraceCounters {
  unsigned long long node_fetches;
  unsigned long long prim_tests;
  unsigned long long tri_tests;
  unsigned long long rays;
  unsigned long long hits;
};

// Traversal stack in per-thread private memory (plain kernels).  The FP64
// path walks the binary BVH (one word per entry, depth <= 48); the FP32 path
// walks the 4-wide BVH (a node index and its entry distance per entry, up to
// three entries per level).  Scene creation checks a scene's worst case
// against these capacities.
constexpr int kPrivateStackWords64 = 64;
constexpr int kPrivateStackWords32 = 192;
template <int N>
struct PrivateStackN {
  uint32_t s[N];
  int sp = 0;
  LMN_HD void push(uint32_t v) { s[sp++] = v; }
  LMN_HD uint32_t pop() { return s[--sp]; }
  LMN_HD void push2(uint32_t a, uint32_t b) {
    s[sp] = a;
    s[sp + 1] = b;
    sp += 2;
  }
  LMN_HD void pop2(uint32_t& a, uint32_t& b) {
    sp -= 2;
    a = s[sp];
    b = s[sp + 1];
  }
  LMN_HD bool empty() const { return sp == 0; }
};
template <class R>
using PrivateStack = PrivateStackN<sizeof(R) == 4 ? kPrivateStackWords32 : kPrivateStackWords64>;

template <class R>
LMN_HD bool slab(const float* lo, const float* hi, V3<R> o, V3<R> inv, R tmin, R tmax, R& tenter);
// Slab test of a node box (src/aabb.rs:18-56), inclusive of grazing contact,
// reciprocal hoisted per ray.  NaN slab bounds are ignored (f64::max/min).
template <>
LMN_HD bool slab<float>(const float* lo, const float* hi, V3<float> o, V3<float> inv, float tmin, float tmax,
                        float& tenter) {
  float tx0 = (lo[0] - o.x) * inv.x, tx1 = (hi[0] - o.x) * inv.x;
  float ty0 = (lo[1] - o.y) * inv.y, ty1 = (hi[1] - o.y) * inv.y;
  float tz0 = (lo[2] - o.z) * inv.z, tz1 = (hi[2] - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  tenter = tn;
  return tn <= tf;
}
LMN_HD bool slab_d(const double* lo, const double* hi, V3<double> o, V3<double> inv, double tmin, double tmax,
                   double& tenter) {
  double tx0 = (lo[0] - o.x) * inv.x, tx1 = (hi[0] - o.x) * inv.x;
  double ty0 = (lo[1] - o.y) * inv.y, ty1 = (hi[1] - o.y) * inv.y;
  double tz0 = (lo[2] - o.z) * inv.z, tz1 = (hi[2] - o.z) * inv.z;
  double tn = fmax(fmax(fmin(tx0, tx1), fmin(ty0, ty1)), fmax(fmin(tz0, tz1), tmin));
  double tf = fmin(fmin(fmax(tx0, tx1), fmax(ty0, ty1)), fmin(fmax(tz0, tz1), tmax));
  tenter = tn;
  return tn <= tf;
}
template <class R> struct NodeOps;
template <> struct NodeOps<float> {
  LMN_HD static bool test(const NodeF& n, V3<float> o, V3<float> inv, float tmin, float tmax, float& te) {
    return slab<float>(n.lo, n.hi, o, inv, tmin, tmax, te);
  }
};
template <> struct NodeOps<double> {
  LMN_HD static bool test(const NodeD& n, V3<double> o, V3<double> inv, double tmin, double tmax, double& te) {
    return slab_d(n.lo, n.hi, o, inv, tmin, tmax, te);
  }
};

// Scene features that select a kernel specialisation: instance transforms,
// FP64 spheres on the FP32 path (large spheres), marched primitives (torus,
// superellipsoid, spring; FP64, out of line).  A specialisation without a
// feature carries none of its code or registers.
enum : uint32_t { FEAT_XF = 1, FEAT_F64 = 2, FEAT_MARCH = 4 };

// Whether record `km` is evaluated in FP64 by the R-precision path.
template <class R, uint32_t FEAT>
LMN_HD bool prim_is64(uint32_t km) {
  if constexpr (sizeof(R) == 4) {
    if (!(FEAT & (FEAT_F64 | FEAT_MARCH))) return false;
    return (km & FLAG_FP64) != 0;
  } else {
    if (!(FEAT & FEAT_MARCH)) return false;
    return (km & KIND_MASK) >= K_TORUS;  // marched kinds and quadrics (second BVH)
  }
}

// The record a ray leaves from (FP32 extension rays), or none.
constexpr uint32_t kNoOrigin = 0xffffffffu;

// A ray that leaves the surface of record `org` cannot meet that surface
// again if it is planar, and meets a sphere it leaves only at the far root,
// and only if it heads inward.  That is what the reference's t >= 0.001
// decides in FP64; an FP32 hit point lies up to a few ulps off the surface,
// and with the reference's unnormalised scatter directions (n + a unit
// vector can be arbitrarily short) 0.001 in t can be less than that, so the
// FP32 path would meet its own surface again.  `sphere_far` is set for an
// origin sphere; false = skip the record.
template <class R>
LMN_HD bool origin_rule(uint32_t kind, const R* p, V3<R> o, V3<R> d, bool& sphere_far) {
  sphere_far = false;
  if (kind != K_SPHERE) return kind != K_QUAD && kind != K_RECT_XY && kind != K_RECT_XZ && kind != K_RECT_YZ &&
                               kind != K_TRIANGLE && kind != K_PLANE;
  sphere_far = dot(o - V3<R>{p[0], p[1], p[2]}, d) < R(0);
  return sphere_far;
}

// Record i of a device table of 16-byte-aligned records, copied whole into
// `local` on the device: its 16-byte quarters are global loads issued
// together and all taken before any field is read, so a record whose fields
// are read behind a test of its kind word costs one memory round trip, not
// one per field.  On the host (and with PIN false) the record is read in place.
template <bool PIN, class T>
LMN_HD const T& record_copy(const T* table, uint32_t i, T& local) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (PIN) {
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "records are whole 16-byte quarters");
    constexpr int N = int(sizeof(T) / 16);
    typedef __attribute__((address_space(1))) const uint4 G4;
    G4* q = (G4*)((__attribute__((address_space(1))) const char*)table + uint64_t(i) * sizeof(T));
    uint4 w[N];
    for (int k = 0; k < N; ++k) w[k] = q[k];
    for (int k = 0; k < N; ++k) asm volatile("" : "+v"(w[k].x), "+v"(w[k].y), "+v"(w[k].z), "+v"(w[k].w));
    __builtin_memcpy(&local, w, sizeof(T));
    return local;
  }
#endif
  (void)local;
  return table[i];
}

// A primitive record.  On the device, where PIN is set, an FP32 record is
// four 16-byte global loads issued together and all taken before the kind
// (in the last quarter) is decoded, so a record costs one memory round trip
// rather than a load of the kind followed by loads of the fields its test
// reads.  The copy lives in `local`.  Without PIN the record is read in place, field
// by field: the kernels with FP64 islands or transforms have no registers to
// spare for the copy (it cost them spills in the traversal loop).
template <class R, bool PIN>
LMN_HD const PrimT<R>& prim_record(const PrimT<R>* prims, uint32_t idx, PrimT<R>& local) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4 && PIN) {
    typedef __attribute__((address_space(1))) const uint4 G4;
    G4* q = (G4*)((__attribute__((address_space(1))) const char*)prims + (uint64_t(idx) << 6));
    uint4 w[4] = {q[0], q[1], q[2], q[3]};
    for (int k = 0; k < 4; ++k) asm volatile("" : "+v"(w[k].x), "+v"(w[k].y), "+v"(w[k].z), "+v"(w[k].w));
    __builtin_memcpy(&local, w, sizeof local);
    return local;
  }
#endif
  (void)local;
  return prims[idx];
}
// FP32 kernels whose records are copied whole (see prim_record).
template <uint32_t FEAT> constexpr bool kPinRecords = (FEAT & (FEAT_F64 | FEAT_XF)) == 0;

// Intersects primitive `idx` (transform applied); on a hit in
// [tmin, closest] updates closest to the reference-reported t.  `org`: the
// record the ray leaves from (FP32 extension rays, see origin_rule).
template <class R, uint32_t FEAT>
LMN_HD bool test_prim(const SceneView<R>& S, uint32_t idx, V3<R> o, V3<R> d, R tmin, R& closest,
                      uint32_t org = kNoOrigin) {
  PrimT<R> local;
  const PrimT<R>& pr = prim_record<R, kPinRecords<FEAT>>(S.prims, idx, local);
  uint32_t km = pr.kind_mat;
  uint32_t kind = km & KIND_MASK;
  bool has_xf = (FEAT & FEAT_XF) && pr.xform != XF_NONE;
  V3<R> oo = o, dd = d;
  if (has_xf) {
    const XformT<R>& X = S.xforms[pr.xform];
    oo = xf_point(X.inv, o);
    dd = xf_dir(X.inv, d);
  }
  R t;
  bool hit;
  bool far_only = false;
  if constexpr (sizeof(R) == 4) {
    if (idx == org && !origin_rule<R>(kind, pr.p, oo, dd, far_only)) return false;
  }
  if (prim_is64<R, FEAT>(km)) {
    V3<double> o64 = cvt<double>(o), d64 = cvt<double>(d);
    if (has_xf) {
      const XformT<double>& X64 = S.xforms64[pr.xform];
      o64 = xf_point(X64.inv, o64);
      d64 = xf_dir(X64.inv, d64);
    }
    double t64;
    if (far_only)
      hit = hit_sphere<double>(S.prims64[idx].p, o64, d64, double(tmin), double(closest), t64, true);
    else if constexpr ((FEAT & FEAT_MARCH) != 0)
      hit = hit_prim_fp64(kind, S.prims64[idx].p, o64, d64, double(tmin), double(closest), S.exact_marched != 0, t64);
    else  // the main BVH's FP64 records are spheres (the second BVH holds the rest)
      hit = hit_sphere<double>(S.prims64[idx].p, o64, d64, double(tmin), double(closest), t64);
    t = R(t64);
    if (hit && !(t <= closest)) hit = false;  // rounding back to R must not leave the range
  } else if (far_only) {
    hit = hit_sphere<R>(pr.p, oo, dd, tmin, closest, t, true);
  } else {
    hit = hit_prim_analytic<R>(kind, pr.p, oo, dd, tmin, closest, t);
  }
  if (!hit) return false;
  if (has_xf) {
    const XformT<R>& X = S.xforms[pr.xform];
    if (X.tmode != 0) {
      R tr = t * t_report_factor(X, dd, d);
      if (!(tr <= closest)) return false;
      t = tr;
    }
  }
  closest = t;
  return true;
}

// test_prim for an FP32-path record of the marched kinds, in pieces for
// kernels that test one kind per launch: the object-space FP64 ray, and the
// acceptance of an FP64 root (rounding to FP32, reported-t rule), exactly as
// test_prim does them.
struct MarchedRay {
  V3<double> o64, d64;
  V3<float> dd;  // object-space FP32 direction (reported-t factor)
  bool has_xf;
};
template <uint32_t FEAT>
LMN_HD MarchedRay marched_ray(const SceneView<float>& S, uint32_t idx, V3<float> o, V3<float> d) {
  const PrimT<float>& pr = S.prims[idx];
  MarchedRay r{cvt<double>(o), cvt<double>(d), d, (FEAT & FEAT_XF) && pr.xform != XF_NONE};
  if (r.has_xf) {
    const XformT<double>& X64 = S.xforms64[pr.xform];
    r.o64 = xf_point(X64.inv, r.o64);
    r.d64 = xf_dir(X64.inv, r.d64);
    r.dd = xf_dir(S.xforms[pr.xform].inv, d);
  }
  return r;
}
template <uint32_t FEAT>
LMN_HD bool marched_accept(const SceneView<float>& S, uint32_t idx, const MarchedRay& r, V3<float> d, double t64,
                           float& closest) {
  float t = float(t64);
  if (!(t <= closest)) return false;  // rounding back to FP32 must not leave the range
  if (r.has_xf) {
    const XformT<float>& X = S.xforms[S.prims[idx].xform];
    if (X.tmode != 0) {
      float tr = t * t_report_factor(X, r.dd, d);
      if (!(tr <= closest)) return false;
      t = tr;
    }
  }
  closest = t;
  return true;
}

// The whole test for kind KIND.  ANY: visibility only (closest is then left
// unchanged on a superellipsoid hit in exact mode, whose visibility query
// stops at the bracketing leaf).
template <uint32_t KIND, uint32_t FEAT, bool ANY = false>
LMN_HD bool test_marched(const SceneView<float>& S, uint32_t idx, V3<float> o, V3<float> d, float tmin,
                         float& closest) {
  const MarchedRay r = marched_ray<FEAT>(S, idx, o, d);
  const double* p = S.prims64[idx].p;
  const bool exact = S.exact_marched != 0;
  const double t0 = double(tmin), t1 = double(closest);
  double t64;
  bool hit;
  if constexpr (KIND == K_TORUS)
    hit = exact ? hit_torus_exact(p, r.o64, r.d64, t0, t1, t64) : hit_torus_march(p, r.o64, r.d64, t0, t1, t64);
  else if constexpr (KIND == K_SUPERELLIPSOID && ANY) {
    if (exact) return hit_superellipsoid_exact_any(p, r.o64, r.d64, t0, t1);
    hit = hit_superellipsoid_march(p, r.o64, r.d64, t0, t1, t64);
  } else if constexpr (KIND == K_SUPERELLIPSOID)
    hit = exact ? hit_superellipsoid_exact(p, r.o64, r.d64, t0, t1, t64)
                : hit_superellipsoid_march(p, r.o64, r.d64, t0, t1, t64);
  else
    hit = exact ? hit_spring_exact(p, r.o64, r.d64, t0, t1, t64) : hit_spring_march(p, r.o64, r.d64, t0, t1, t64);
  return hit && marched_accept<FEAT>(S, idx, r, d, t64, closest);
}

// Every primitive record in the leaves of a binary BVH whose boxes the ray
// enters within [tmin, tmax], in no particular order: `f(record)` per record.
template <class Stack, class F>
LMN_HD void bvh_leaves(const NodeF* nodes, V3<float> o, V3<float> inv, float tmin, float tmax, Stack& st, F&& f) {
  NodeF nd = nodes[0];
  float te;
  bool live = NodeOps<float>::test(nd, o, inv, tmin, tmax, te);
  while (live) {
    if (nd.b & LEAF_BIT) {
      for (uint32_t k = 0; k < (nd.b & ~LEAF_BIT); ++k) f(nd.a + k);
    } else {
      const uint32_t c = nd.a;  // children at c, c + 1
      NodeF L = nodes[c], Rn = nodes[c + 1];
      float tl, tr;
      bool hl = NodeOps<float>::test(L, o, inv, tmin, tmax, tl);
      bool hr = NodeOps<float>::test(Rn, o, inv, tmin, tmax, tr);
      if (hl && hr) {
        st.push(c + 1);
        nd = L;
        continue;
      }
      if (hl || hr) {
        nd = hl ? L : Rn;
        continue;
      }
    }
    live = !st.empty();
    if (live) nd = nodes[st.pop()];
  }
}

// One BVH walk.  Ordered BVH2 traversal: the two children of a node are
// stored side by side, so one contiguous 64-byte fetch (FP32) brings both
// boxes; both are tested, the nearer is visited first and the farther
// pushed; a popped node is re-tested against the current closest hit before
// it is expanded.
template <class R, uint32_t FEAT, bool ANY, class Stack>
LMN_HD void bvh_walk(const SceneView<R>& S, const typename NodeSel<R>::type* nodes, V3<R> o, V3<R> d, V3<R> inv,
                     R tmin, R& closest, uint32_t& best, bool& found, Stack& st, uint32_t& nf, uint32_t& pt,
                     uint32_t& tt, bool counting, uint32_t org = kNoOrigin) {
  using Node = typename NodeSel<R>::type;
  Node nd = nodes[0];
  ++nf;
  R te;
  bool live = NodeOps<R>::test(nd, o, inv, tmin, closest, te);
  while (live) {
    if (nd.b & LEAF_BIT) {
      uint32_t first = nd.a, count = nd.b & ~LEAF_BIT;
      for (uint32_t k = 0; k < count; ++k) {
        ++pt;
        if (counting && (S.prims[first + k].kind_mat & KIND_MASK) == K_TRIANGLE) ++tt;
        if (test_prim<R, FEAT>(S, first + k, o, d, tmin, closest, org)) {
          found = true;
          best = first + k;
          if (ANY) break;
        }
      }
      if (ANY && found) return;
    } else {
      uint32_t c = nd.a;  // children at c, c + 1
      Node L = nodes[c], Rn = nodes[c + 1];
      nf += 2;
      R tl, tr;
      bool hl = NodeOps<R>::test(L, o, inv, tmin, closest, tl);
      bool hr = NodeOps<R>::test(Rn, o, inv, tmin, closest, tr);
      if (hl && hr) {
        bool left_first = tl <= tr;
        st.push(left_first ? c + 1 : c);
        nd = left_first ? L : Rn;
        continue;
      }
      if (hl || hr) {
        nd = hl ? L : Rn;
        continue;
      }
    }
    live = false;
    while (!st.empty()) {
      nd = nodes[st.pop()];
      ++nf;
      if (NodeOps<R>::test(nd, o, inv, tmin, closest, te)) {
        live = true;
        break;
      }
    }
  }
}

// The four 16-byte quarters of 4-wide node `ref`.  On the device they are
// global-memory loads off the scene's node array (a scalar base and a 32-bit
// offset: the host builder caps a tree at 2^26 nodes), all four issued
// before any is used.
struct Words4 {
  uint32_t x, y, z, w;
};
LMN_HD void node4_words(const Node4Q* nodes, uint32_t ref, Words4 q[4]) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const uint4 G4;
  G4* p = (G4*)((__attribute__((address_space(1))) const char*)nodes + (ref << 6));
  for (int k = 0; k < 4; ++k) {
    const uint4 v = p[k];
    q[k] = {v.x, v.y, v.z, v.w};
  }
#else
  __builtin_memcpy(q, nodes + ref, sizeof(Node4Q));
#endif
}
// Inverse direction for the FP32 slab tests: components clamped to 2^80 in
// magnitude, so that an axis-parallel ray gets a finite factor (the 4-wide
// nodes form slab distances as fma(q, step / d, (org - o) / d), where an
// infinite factor would meet 0 * inf).  The clamped box test is at least as
// permissive as the unclamped one.
LMN_HD V3<float> slab_inv(V3<float> d) {
  constexpr float C = 1.2089258196146292e24f;  // 2^80
  return {fminf(fmaxf(1.f / d.x, -C), C), fminf(fmaxf(1.f / d.y, -C), C), fminf(fmaxf(1.f / d.z, -C), C)};
}

// Entry distances of a 4-wide node's children (INFINITY where the ray
// misses the child's box within [tmin, closest]) and their entries.
LMN_HD void node4_children(const Words4 w[4], V3<float> o, V3<float> inv, float tmin, float closest, float tk[4],
                           uint32_t ch[4]) {
  const uint32_t meta = w[3].y, nc = meta >> 24;
  // slab distance of grid plane q on axis a: fma(q, a_a, b_a), with
  // a_a = step_a / d_a (exact: step_a is a power of two) and
  // b_a = (org_a - o_a) / d_a
  const float ax = ldexpf(inv.x, int(int8_t(meta & 0xffu)));
  const float ay = ldexpf(inv.y, int(int8_t((meta >> 8) & 0xffu)));
  const float az = ldexpf(inv.z, int(int8_t((meta >> 16) & 0xffu)));
  const float bx = (float_of(w[0].x) - o.x) * inv.x;
  const float by = (float_of(w[0].y) - o.y) * inv.y;
  const float bz = (float_of(w[0].z) - o.z) * inv.z;
  // near and far plane bytes per axis: the lower planes are near where the
  // direction component is positive
  const uint32_t lox = w[0].w, loy = w[1].x, loz = w[1].y, hix = w[1].z, hiy = w[1].w, hiz = w[2].x;
  const bool sx = inv.x < 0.f, sy = inv.y < 0.f, sz = inv.z < 0.f;
  const uint32_t nx = sx ? hix : lox, fx = sx ? lox : hix;
  const uint32_t ny = sy ? hiy : loy, fy = sy ? loy : hiy;
  const uint32_t nz = sz ? hiz : loz, fz = sz ? loz : hiz;
  ch[0] = w[2].y;
  ch[1] = w[2].z;
  ch[2] = w[2].w;
  ch[3] = w[3].x;
  for (int k = 0; k < 4; ++k) {
    const uint32_t sh = 8u * uint32_t(k);
    const float tnx = fmaf(float((nx >> sh) & 0xffu), ax, bx), tfx = fmaf(float((fx >> sh) & 0xffu), ax, bx);
    const float tny = fmaf(float((ny >> sh) & 0xffu), ay, by), tfy = fmaf(float((fy >> sh) & 0xffu), ay, by);
    const float tnz = fmaf(float((nz >> sh) & 0xffu), az, bz), tfz = fmaf(float((fz >> sh) & 0xffu), az, bz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, closest));
    tk[k] = (uint32_t(k) < nc && tn <= tf) ? tn : INFINITY;
  }
}

// Ordered walk of the FP32 4-wide BVH.  A node's (up to) four child boxes
// are tested together; the hits are sorted by entry distance, the nearest is
// visited next and the others pushed farthest first, each with its entry
// distance so that a pop can drop it once a closer hit is known (a
// visibility query, ANY, pushes the child alone).
// One step of the 4-wide walk: the node or leaf `ref` names, then the next
// entry to visit (popped from the stack) into `ref`.  Returns true once the
// walk has ended (stack empty, or a visibility query's first hit).  A walk
// is `ref = 0` (the root) followed by steps until one returns true, so a
// caller may interleave other work (a persistent kernel's ray fetch)
// between steps.  `inv` is slab_inv(d).
template <uint32_t FEAT, bool ANY, class Stack>
LMN_HD bool bvh4_step(const SceneView<float>& S, V3<float> o, V3<float> d, V3<float> inv, float tmin,
                      float& closest, uint32_t& best, bool& found, uint32_t& ref, Stack& st, uint32_t& nf,
                      uint32_t& pt, uint32_t& tt, bool counting, uint32_t org = kNoOrigin) {
  if (!(ref & LEAF4)) {
    Words4 w[4];
    node4_words(S.nodes4, ref, w);
    ++nf;
    float tk[4];
    uint32_t ch[4];
    node4_children(w, o, inv, tmin, closest, tk, ch);
    // sorting network on (tk, ch): misses (INFINITY) sink to the end
#define LMN_CSWAP(a, b)                    \
  if (tk[b] < tk[a]) {                     \
    float tt_ = tk[a];                     \
    tk[a] = tk[b];                         \
    tk[b] = tt_;                           \
    uint32_t cc_ = ch[a];                  \
    ch[a] = ch[b];                         \
    ch[b] = cc_;                           \
  }
    LMN_CSWAP(0, 1)
    LMN_CSWAP(2, 3)
    LMN_CSWAP(0, 2)
    LMN_CSWAP(1, 3)
    LMN_CSWAP(1, 2)
#undef LMN_CSWAP
    if (tk[0] != INFINITY) {
      // a visibility query returns at its first hit, so its range never
      // shrinks and an entry needs no distance: one word instead of two
      for (int k = 3; k >= 1; --k)
        if (tk[k] != INFINITY) {
          if (ANY)
            st.push(ch[k]);
          else
            st.push2(ch[k], bits_of(tk[k]));
        }
      ref = ch[0];
      return false;
    }
  } else {
    uint32_t first = ref & LEAF4_FIRST, count = ((ref >> 28) & 7u) + 1u;
    for (uint32_t k = 0; k < count; ++k) {
      ++pt;
      if (counting && (S.prims[first + k].kind_mat & KIND_MASK) == K_TRIANGLE) ++tt;
      if (test_prim<float, FEAT>(S, first + k, o, d, tmin, closest, org)) {
        found = true;
        best = first + k;
        if (ANY) return true;
      }
    }
  }
  if (ANY) {
    if (!st.empty()) {
      ref = st.pop();
      return false;
    }
  } else {
    while (!st.empty()) {
      uint32_t r, tb;
      st.pop2(r, tb);
      const float te = float_of(tb);
      if (te <= closest) {
        ref = r;
        return false;
      }
    }
  }
  return true;
}

template <uint32_t FEAT, bool ANY, class Stack>
LMN_HD void bvh4_walk(const SceneView<float>& S, V3<float> o, V3<float> d, V3<float> inv, float tmin,
                      float& closest, uint32_t& best, bool& found, Stack& st, uint32_t& nf, uint32_t& pt,
                      uint32_t& tt, bool counting, uint32_t org = kNoOrigin) {
  uint32_t ref = 0;
  while (!bvh4_step<FEAT, ANY>(S, o, d, inv, tmin, closest, best, found, ref, st, nf, pt, tt, counting, org)) {
  }
}

// Which part of the scene a trace covers: everything; the main BVH plus the
// unbounded planes; or only the marched primitives' BVH (the FP32 wavefront
// runs that in a kernel of its own so the FP64 marching code does not set
// the register budget of the main traversal).
enum : uint32_t { PASS_ALL = 0, PASS_MAIN = 1, PASS_MARCH = 2 };

// Closest hit over the BVH then the unbounded list (SceneWorld::hit,
// src/scene.rs:817-834).  ANY = stop at the first hit in range (NEE
// visibility only asks is_some(), src/render.rs:1458).  `closest` enters as
// t_max (or a previous pass's hit) and leaves as the closest hit's t.
template <class R, uint32_t FEAT, bool ANY, uint32_t PASS = PASS_ALL, class Stack>
LMN_HD bool trace_pass(const SceneView<R>& S, V3<R> o, V3<R> d, R tmin, R& closest, uint32_t& best, Stack& st,
                       TraceCounters* cnt = nullptr, uint32_t org = kNoOrigin) {
  bool found = false;
  V3<R> inv;
  if constexpr (sizeof(R) == 4)
    inv = slab_inv(d);
  else
    inv = {R(1) / d.x, R(1) / d.y, R(1) / d.z};
  uint32_t nf = 0, pt = 0, tt = 0;
  constexpr uint32_t MAIN_FEAT = FEAT & ~FEAT_MARCH;
  if constexpr (sizeof(R) == 4) {
    if (PASS != PASS_MARCH && S.n_nodes4 > 0)
      bvh4_walk<MAIN_FEAT, ANY>(S, o, d, inv, tmin, closest, best, found, st, nf, pt, tt, cnt != nullptr, org);
  } else {
    if (PASS != PASS_MARCH && S.n_nodes > 0)
      bvh_walk<R, MAIN_FEAT, ANY>(S, S.nodes, o, d, inv, tmin, closest, best, found, st, nf, pt, tt, cnt != nullptr,
                                  org);
  }
  if (PASS != PASS_MARCH && !(ANY && found)) {
    for (uint32_t k = 0; k < S.n_planes; ++k) {
      uint32_t idx = S.n_bounded + k;
      ++pt;
      if (test_prim<R, MAIN_FEAT>(S, idx, o, d, tmin, closest, org)) {
        found = true;
        best = idx;
        if (ANY) break;
      }
    }
  }
  if constexpr ((FEAT & FEAT_MARCH) != 0) {
    if (PASS != PASS_MAIN && S.n_mnodes > 0 && !(ANY && found))
      bvh_walk<R, FEAT, ANY>(S, S.mnodes, o, d, inv, tmin, closest, best, found, st, nf, pt, tt, cnt != nullptr);
  }
#if defined(__HIPCC__)
  if (cnt) {
    atomicAdd(&cnt->node_fetches, (unsigned long long)nf);
    atomicAdd(&cnt->prim_tests, (unsigned long long)pt);
    atomicAdd(&cnt->tri_tests, (unsigned long long)tt);
    atomicAdd(&cnt->rays, 1ull);
    if (found) atomicAdd(&cnt->hits, 1ull);
  }
#else
  (void)cnt;
  (void)tt;
#endif
  return found;
}

template <class R, uint32_t FEAT, bool ANY, class Stack>
LMN_HD bool trace(const SceneView<R>& S, V3<R> o, V3<R> d, R tmin, R tmax, R& t_out, uint32_t& idx_out, Stack& st,
                  TraceCounters* cnt = nullptr, uint32_t org = kNoOrigin) {
  R closest = tmax;
  uint32_t best = 0;
  bool found = trace_pass<R, FEAT, ANY, PASS_ALL>(S, o, d, tmin, closest, best, st, cnt, org);
  t_out = closest;
  idx_out = best;
  return found;
}

// ---------------------------------------------------------------------------
// Hit record reconstruction (HitRecord::new, src/hit.rs:19-43): the outward
// normal and (u, v) are evaluated in object space, the normal is flipped
// against the object-space ray and mapped back through the transform's
// normal matrix.
template <class R>
struct Surf {
  V3<R> p;       // world point
  V3<R> n;       // unit normal facing the incoming ray
  R t;           // reference-reported t
  R t_true;      // ray parameter of the point (differs from t under Scale wrappers)
  R u, v;
  bool front;
  uint32_t mat;
  uint32_t prim_id;
};

// The hit point of ray (o, d) at parameter t (one definition for every kernel
// that forms it, so that they round alike).
template <class R> LMN_HD V3<R> surface_point(V3<R> o, V3<R> d, R t) { return o + d * t; }

// want_uv = false: (u, v) are left 0 unless the record needs them itself (a
// smooth triangle's barycentrics interpolate its normal) -- for callers whose
// scene reads no image texture, where nothing else looks at them.
template <class R, uint32_t FEAT>
LMN_HD void surface(const SceneView<R>& S, uint32_t idx, V3<R> o, V3<R> d, R t_rep, Surf<R>& s,
                    bool want_uv = true) {
  PrimT<R> local;
  const PrimT<R>& pr = prim_record<R, kPinRecords<FEAT>>(S.prims, idx, local);
  uint32_t km = pr.kind_mat, kind = km & KIND_MASK;
  bool has_xf = (FEAT & FEAT_XF) && pr.xform != XF_NONE;
  V3<R> oo = o, dd = d;
  R t_true = t_rep;
  if (has_xf) {
    const XformT<R>& X = S.xforms[pr.xform];
    oo = xf_point(X.inv, o);
    dd = xf_dir(X.inv, d);
    if (X.tmode != 0) t_true = t_rep / t_report_factor(X, dd, d);
  }
  V3<R> n;
  if (prim_is64<R, FEAT>(km)) {
    V3<double> oo64 = cvt<double>(o), dd64 = cvt<double>(d);
    if (has_xf) {
      const XformT<double>& X64 = S.xforms64[pr.xform];
      oo64 = xf_point(X64.inv, oo64);
      dd64 = xf_dir(X64.inv, dd64);
    }
    const double* p64 = S.prims64[idx].p;
    V3<double> q = oo64 + dd64 * double(t_true);
    double u64, v64;
    if ((FEAT & FEAT_MARCH) && is_marched_kind(kind)) {
      n = cvt<R>(marched_normal_fp64(kind, p64, q, S.exact_marched != 0));
      marched_uv(kind, p64, q, u64, v64);
    } else if ((FEAT & FEAT_MARCH) && is_quadric_kind(kind)) {
      // only the second BVH holds these, and only kernels with FEAT_MARCH
      // walk it: the main traversal's kernels carry none of this (out of line) code
      V3<double> n64;
      quadric_surface(kind, p64, oo64, dd64, q, n64, u64, v64);
      n = cvt<R>(n64);
    } else {  // FP64 sphere: only the centre offset needs FP64, the map does not
      V3<R> nn = cvt<R>(prim_normal_analytic<double>(kind, p64, q));
      n = nn;
      const R pi = R(3.14159265358979323846);
      u64 = want_uv ? double((m_atan2(-nn.z, nn.x) + pi) / (R(2) * pi)) : 0.0;
      v64 = want_uv ? double(m_acos(-nn.y) / pi) : 0.0;
    }
    s.u = R(u64);
    s.v = R(v64);
  } else {
    n = prim_normal_analytic<R>(kind, pr.p, oo + dd * t_true);
    if (want_uv || (kind == K_TRIANGLE && (km & FLAG_SMOOTH)))
      prim_uv_analytic<R>(kind, pr.p, oo, dd, t_true, s.u, s.v);
    else
      s.u = s.v = R(0);
    if (kind == K_TRIANGLE && (km & FLAG_SMOOTH)) {
      // MeshTriangle::hit (src/obj.rs:291-303): the normal and texture
      // coordinates interpolated at the barycentrics (u, v)
      const AttrT<R>& A = S.attrs[attr_index(pr)];
      const R bu = s.u, bv = s.v;
      if (A.flags & 1u) {
        const V3<R> n0 = {A.n[0], A.n[1], A.n[2]}, n1 = {A.n[3], A.n[4], A.n[5]}, n2 = {A.n[6], A.n[7], A.n[8]};
        n = unit(n0 * (R(1) - bu - bv) + n1 * bu + n2 * bv);
      }
      if (A.flags & 2u) {
        const R w = R(1) - bu - bv;
        s.u = A.uv[0] * w + A.uv[2] * bu + A.uv[4] * bv;
        s.v = A.uv[1] * w + A.uv[3] * bu + A.uv[5] * bv;
      }
    }
  }
  bool front = dot(dd, n) < R(0);
  if (!front) n = -n;
  if (has_xf) {
    const XformT<R>& X = S.xforms[pr.xform];
    n = xf_nrm(X.nrm, n);
    if (X.renorm) n = unit(n);
  }
  s.p = surface_point(o, d, t_true);
  s.n = n;
  s.t = t_rep;
  s.t_true = t_true;
  s.front = front;
  s.mat = km >> 8;
  s.prim_id = pr.prim_id;
}

// ---------------------------------------------------------------------------
// Textures (src/texture.rs:19-23, :463-479)
template <class R>
LMN_HD V3<R> texture_value(const SceneView<R>& S, const MaterialT<R>& m, R u, R v) {
  if (m.tex < 0) return {m.color[0], m.color[1], m.color[2]};
  const TextureDev& T = S.textures[m.tex];
  R uc = clamp_r(u, R(0), R(1));
  R vc = R(1) - clamp_r(v, R(0), R(1));
  uint32_t i = uint32_t(uc * R(T.w));
  uint32_t j = uint32_t(vc * R(T.h));
  i = i < T.w - 1 ? i : T.w - 1;
  j = j < T.h - 1 ? j : T.h - 1;
  const float* px = S.texels + T.offset + (uint64_t(j) * T.w + i) * 3;
  return {R(px[0]), R(px[1]), R(px[2])};
}
// texel index touched by a lookup (path-replay adjoint), or -1 for solid colour
template <class R>
LMN_HD int64_t texel_index(const SceneView<R>& S, const MaterialT<R>& m, R u, R v) {
  if (m.tex < 0) return -1;
  const TextureDev& T = S.textures[m.tex];
  R uc = clamp_r(u, R(0), R(1));
  R vc = R(1) - clamp_r(v, R(0), R(1));
  uint32_t i = uint32_t(uc * R(T.w));
  uint32_t j = uint32_t(vc * R(T.h));
  i = i < T.w - 1 ? i : T.w - 1;
  j = j < T.h - 1 ? j : T.h - 1;
  return int64_t(T.offset / 3 + uint64_t(j) * T.w + i);
}

template <class R> LMN_HD bool is_specular(const MaterialT<R>& m) {
  switch (m.kind) {
    case M_METAL:
    case M_DIELECTRIC: return true;
    case M_MICROFACET: return m.metallic > R(0.9) && m.roughness < R(0.1);  // src/material.rs:323-326
    default: return false;
  }
}

template <class R> LMN_HD V3<R> emitted(const SceneView<R>& S, const MaterialT<R>& m, R u, R v) {
  if (m.kind != M_EMISSIVE) return {R(0), R(0), R(0)};
  return texture_value(S, m, u, v) * m.intensity;  // src/material.rs:241-243
}

template <class R> LMN_HD R powi5(R x) {
  R x2 = x * x;
  return x2 * x2 * x;
}
// Dielectric::reflectance, Schlick (src/material.rs:138-141)
template <class R> LMN_HD R schlick(R cosine, R ref_idx) {
  R r0 = (R(1) - ref_idx) / (R(1) + ref_idx);
  r0 = r0 * r0;
  return r0 + (R(1) - r0) * powi5(R(1) - cosine);
}
// Microfacet::fresnel (src/material.rs:291-294)
template <class R> LMN_HD V3<R> fresnel(R cos_theta, V3<R> f0) {
  R t = powi5(fmax_r(R(1) - cos_theta, R(0)));
  return f0 * (R(1) - t) + V3<R>{t, t, t};
}
// Microfacet::geometry, Schlick-Smith (src/material.rs:280-288)
template <class R> LMN_HD R ggx_geometry(R n_dot_v, R n_dot_l, R rough) {
  R k = (rough + R(1)) * (rough + R(1)) / R(8);
  return (n_dot_v / (n_dot_v * (R(1) - k) + k)) * (n_dot_l / (n_dot_l * (R(1) - k) + k));
}
template <class R> LMN_HD R vatan(R x) { return m_atan(x); }
// Microfacet::sample_ggx_normal (src/material.rs:297-319)
template <class R> LMN_HD V3<R> sample_ggx_normal(V3<R> n, R alpha, Rng& rng) {
  R r1 = rng.next<R>(), r2 = rng.next<R>();
  const R pi = R(3.14159265358979323846);
  R theta = m_atan(vsqrt(alpha * alpha * r1 / (R(1) - r1 + R(1e-10))));
  R phi = R(2) * pi * r2;
  R st = m_sin(theta), ct = m_cos(theta);
  V3<R> hl = {st * m_cos(phi), st * m_sin(phi), ct};
  V3<R> up = m_abs(n.y) < R(0.999) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
  V3<R> tg = unit(cross(up, n));
  V3<R> bt = cross(n, tg);
  return unit(tg * hl.x + bt * hl.y + n * hl.z);
}

// What the path-replay adjoint needs to know about a scatter event: the
// lobe taken and the cosines its weight was built from.
enum : uint32_t { LOBE_OTHER = 0, LOBE_LAMBERT = 1, LOBE_GGX_SPEC = 2, LOBE_GGX_DIFFUSE = 3 };
template <class R>
struct ScatterInfo {
  uint32_t lobe;
  R ndv, ndl, ndh;  // GGX specular lobe (ndv clamped as in the weight)
  V3<R> datt;       // d att / d albedo per channel (textured albedo lobes)
};

// Material::scatter for the five hot-path materials (src/material.rs:59-391).
// Returns false when the material absorbs (scatter() == None).  `ray_d` is
// the incoming (unnormalised) direction.  `info`, when given, receives the
// lobe and its cosines (the forward kernels pass nothing).
template <class R>
LMN_HD bool scatter(const SceneView<R>& S, const MaterialT<R>& m, V3<R> ray_d, const Surf<R>& h, Rng& rng,
                    V3<R>& out_dir, V3<R>& att, ScatterInfo<R>* info = nullptr) {
  if (info) info->lobe = LOBE_OTHER;
  switch (m.kind) {
    case M_LAMBERTIAN: {  // src/material.rs:59-71
      V3<R> dir = h.n + random_unit_vector<R>(rng);
      if (near_zero(dir)) dir = h.n;
      out_dir = dir;
      att = texture_value(S, m, h.u, h.v);
      if (info) {
        info->lobe = LOBE_LAMBERT;
        info->datt = {R(1), R(1), R(1)};
      }
      return true;
    }
    case M_METAL: {  // src/material.rs:94-106
      V3<R> refl = reflect(unit(ray_d), h.n);
      V3<R> sc = refl + random_in_unit_sphere<R>(rng) * m.fuzz;
      if (!(dot(sc, h.n) > R(0))) return false;
      out_dir = sc;
      att = {m.color[0], m.color[1], m.color[2]};
      return true;
    }
    case M_DIELECTRIC: {  // src/material.rs:149-213
      R ior = m.ior;
      V3<R> chan = {R(1), R(1), R(1)};
      if (m.dispersion > R(0)) {
        uint32_t c = uint32_t(rng.next<R>() * R(3));
        ior = ior + (c == 0 ? -m.dispersion : (c == 1 ? R(0) : m.dispersion));
        chan = c == 0 ? V3<R>{R(3), R(0), R(0)} : (c == 1 ? V3<R>{R(0), R(3), R(0)} : V3<R>{R(0), R(0), R(3)});
      }
      R eta = h.front ? R(1) / ior : ior;
      V3<R> ud = unit(ray_d);
      R ct = dot(-ud, h.n);
      ct = ct < R(1) ? ct : R(1);
      R stt = vsqrt(R(1) - ct * ct);
      bool cannot = eta * stt > R(1);
      V3<R> dir;
      if (cannot || schlick(ct, eta) > rng.next<R>())
        dir = reflect(ud, h.n);
      else
        dir = refract(ud, h.n, eta);
      if (m.roughness > R(0)) dir = dir + random_in_unit_sphere<R>(rng) * m.roughness;
      V3<R> a;
      bool white = m.color[0] == R(1) && m.color[1] == R(1) && m.color[2] == R(1);
      if (!h.front && !white) {
        a = {m_exp(-(R(1) - m.color[0]) * h.t), m_exp(-(R(1) - m.color[1]) * h.t), m_exp(-(R(1) - m.color[2]) * h.t)};
      } else {
        a = {m.color[0], m.color[1], m.color[2]};
      }
      if (m.dispersion > R(0)) a = hadamard(a, chan);
      out_dir = dir;
      att = a;
      return true;
    }
    case M_MICROFACET: {  // src/material.rs:328-390
      R alpha = m.roughness * m.roughness;
      V3<R> v = unit(-ray_d);
      V3<R> n = h.n;
      V3<R> albedo = texture_value(S, m, h.u, h.v);
      R ndv = fmax_r(dot(n, v), R(0.001));
      V3<R> f0 = V3<R>{R(0.04), R(0.04), R(0.04)} * (R(1) - m.metallic) + albedo * m.metallic;
      V3<R> fw = fresnel(ndv, f0);
      R sp = clamp_r((fw.x + fw.y + fw.z) / R(3), R(0.1), R(0.9));
      if (rng.next<R>() < sp) {
        V3<R> hv = sample_ggx_normal(n, alpha, rng);
        V3<R> l = reflect_around(v, hv);
        R ndl = dot(n, l);
        if (ndl <= R(0)) return false;
        R ndh = fmax_r(dot(n, hv), R(0));
        R vdh = fmax_r(dot(v, hv), R(0));
        R g = ggx_geometry(ndv, ndl, m.roughness);
        V3<R> f = fresnel(vdh, f0);
        V3<R> w = ndh > R(0.001) ? f * (g * vdh / (ndv * ndh)) : V3<R>{R(0), R(0), R(0)};
        out_dir = l;
        att = w / sp;
        if (info) {
          info->lobe = LOBE_GGX_SPEC;
          info->ndv = ndv;
          info->ndl = ndl;
          info->ndh = ndh;
          info->datt = {R(0), R(0), R(0)};  // f0 does not depend on albedo at metallic = 0
        }
        return true;
      }
      V3<R> dir = n + random_unit_vector<R>(rng);
      if (near_zero(dir)) dir = n;
      V3<R> kd = (V3<R>{R(1), R(1), R(1)} - fw) * (R(1) - m.metallic);
      out_dir = dir;
      att = hadamard(kd, albedo) / (R(1) - sp);
      if (info) {
        info->lobe = LOBE_GGX_DIFFUSE;
        info->datt = kd / (R(1) - sp);
      }
      return true;
    }
    default: return false;  // emissive: no scatter (src/material.rs:237-239)
  }
}

// Frisvad orthonormal basis (build_onb, src/render.rs:1333-1343)
template <class R> LMN_HD void build_onb(V3<R> n, V3<R>& a, V3<R>& b) {
  if (n.z < R(-0.9999999)) {
    a = {R(0), R(-1), R(0)};
    b = {R(-1), R(0), R(0)};
    return;
  }
  R aa = R(1) / (R(1) + n.z);
  R bb = -n.x * n.y * aa;
  a = {R(1) - n.x * n.x * aa, bb, -n.x};
  b = {bb, R(1) - n.y * n.y * aa, -n.y};
}

// The sampling half of sample_direct_light (src/render.rs:1346-1466): picks a
// light, samples a point and returns the unoccluded contribution and the
// shadow ray (origin hit_point, unit dir, range [0.001, dist - 0.001]).
// Returns false where the reference returns zero before its shadow query.
// `geo` receives the contribution per unit albedo and emission (the scalar
// factor of the estimate) and `light` the light picked.
// PIN: read the light record whole (record_copy), for the wavefront shading
// kernel; the one-path-per-thread kernels, at their register limit, read it
// in place.
template <class R, bool PIN = false>
LMN_HD bool nee_sample(const SceneView<R>& S, V3<R> hp, V3<R> normal, V3<R> albedo, Rng& rng, V3<R>& dir_out,
                       R& dist_out, V3<R>& contrib, R* geo = nullptr, uint32_t* light = nullptr) {
  uint32_t nl = S.n_lights;
  uint32_t li = uint32_t(rng.next<R>() * R(nl));
  li = li < nl - 1 ? li : nl - 1;
  if (light) *light = li;
  LightT<R> lcopy;
  const LightT<R>& L = record_copy<PIN && sizeof(R) == 4>(S.lights, li, lcopy);
  const R pi = R(3.14159265358979323846);
  V3<R> lp, ln, em = L.emission;
  R area;
  if (L.kind == L_SPHERE) {
    V3<R> tc = L.origin - hp;
    R dc2 = length_squared(tc);
    if (dc2 < L.radius * L.radius) return false;
    R dc = vsqrt(dc2);
    R stm = L.radius / dc;
    R ctm = vsqrt(fmax_r(R(1) - stm * stm, R(0)));
    R r1 = rng.next<R>(), r2 = rng.next<R>();
    R ct = R(1) - r1 * (R(1) - ctm);
    R st = vsqrt(fmax_r(R(1) - ct * ct, R(0)));
    R phi = R(2) * pi * r2;
    V3<R> w = tc / dc;
    V3<R> a, b;
    build_onb(w, a, b);
    V3<R> dir = unit(a * (st * m_cos(phi)) + b * (st * m_sin(phi)) + w * ct);
    V3<R> oc = hp - L.origin;
    R bc = dot(oc, dir);
    R cc = length_squared(oc) - L.radius * L.radius;
    R disc = bc * bc - cc;
    R t;
    if (disc > R(0)) {
      R sd = vsqrt(disc);
      R t1 = -bc - sd;
      t = t1 > R(0.001) ? t1 : -bc + sd;
    } else {
      t = dc;
    }
    lp = hp + dir * t;
    ln = unit(lp - L.origin);
    R sa = R(2) * pi * (R(1) - ctm);
    area = sa * t * t / fmax_r(m_abs(dot(-dir, ln)), R(0.001));
  } else if (L.kind == L_RECT) {
    R s = rng.next<R>(), t = rng.next<R>();
    lp = L.origin + L.u * s + L.v * t;
    ln = L.normal;
    area = L.area;
  } else {  // disk (src/render.rs:1422-1436)
    R r1 = rng.next<R>(), r2 = rng.next<R>();
    R r = vsqrt(r1) * L.radius;
    R th = R(2) * pi * r2;
    V3<R> w = L.normal;
    V3<R> av = m_abs(w.x) > R(0.9) ? V3<R>{R(0), R(1), R(0)} : V3<R>{R(1), R(0), R(0)};
    V3<R> ut = unit(cross(w, av));
    V3<R> vt = cross(w, ut);
    lp = L.origin + ut * (r * m_cos(th)) + vt * (r * m_sin(th));
    ln = L.normal;
    area = pi * L.radius * L.radius;
  }
  V3<R> tl = lp - hp;
  R d2 = length_squared(tl);
  R dist = vsqrt(d2);
  V3<R> dir = tl / dist;
  R ct = dot(normal, dir);
  if (!(ct > R(0))) return false;
  R ctl = m_abs(dot(-dir, ln));
  if (!(ctl > R(0.001))) return false;
  dir_out = dir;
  dist_out = dist;
  R g = ct * ctl * area * R(nl) / (pi * d2);
  contrib = hadamard(em, albedo) * g;
  if (geo) *geo = g;
  return true;
}

// Background::color (src/render.rs:72-172) for the hot-path variants.
template <class R> LMN_HD V3<R> background(const BackgroundT<R>& B, V3<R> d) {
  if (B.kind == BG_SOLID) return B.c0;
  V3<R> ud = unit(d);
  R t = R(0.5) * (ud.y + R(1));
  if (B.kind == BG_SKY) return V3<R>{R(1), R(1), R(1)} * (R(1) - t) + V3<R>{R(0.5), R(0.7), R(1)} * t;
  if (B.kind == BG_GRADIENT) return B.c0 * (R(1) - t) + B.c1 * t;
  V3<R> sky = V3<R>{R(1), R(1), R(1)} * (R(1) - t) + B.c1 * t;
  R sd = dot(ud, B.dir_unit);
  if (sd > R(0.9995)) return B.c0 * B.intensity;
  if (sd > R(0.995)) {
    R glow = (sd - R(0.995)) / (R(0.9995) - R(0.995));
    return sky + B.c0 * (B.intensity * R(0.3) * glow);
  }
  return sky;
}

// Camera::get_ray, thin lens (src/camera.rs:82-103): two or more disk draws
// even at zero aperture, then one time draw (time is unused on this path).
template <class R> LMN_HD void camera_ray(const CameraT<R>& C, R s, R t, Rng& rng, V3<R>& o, V3<R>& d) {
  V3<R> rd = random_in_unit_disk<R>(rng) * C.lens_radius;
  V3<R> off = C.u * rd.x + C.v * rd.y;
  o = C.origin + off;
  if constexpr (sizeof(R) == 4)  // lower_left - origin cancels: |origin| / focus_dist ulps in FP32
    d = C.ll_rel + C.horizontal * s + C.vertical * t - off;
  else  // the reference's order (src/camera.rs:94-100)
    d = C.lower_left + C.horizontal * s + C.vertical * t - C.origin - off;
  (void)rng.next<R>();
}

// PixelFilter::weight (src/render.rs:198-237)
template <class R> LMN_HD R mitchell_1d(R x) {
  x = m_abs(x);
  const R b = R(1) / R(3), c = R(1) / R(3);
  if (x < R(1))
    return ((R(12) - R(9) * b - R(6) * c) * x * x * x + (R(-18) + R(12) * b + R(6) * c) * x * x + (R(6) - R(2) * b)) / R(6);
  if (x < R(2))
    return ((-b - R(6) * c) * x * x * x + (R(6) * b + R(30) * c) * x * x + (R(-12) * b - R(48) * c) * x +
            (R(8) * b + R(24) * c)) / R(6);
  return R(0);
}
template <class R> LMN_HD R filter_weight(uint32_t f, R dx, R dy) {
  switch (f) {
    case 1: return fmax_r(R(1) - m_abs(dx) * R(2), R(0)) * fmax_r(R(1) - m_abs(dy) * R(2), R(0));
    case 2: {
      R sg = R(0.35);
      return m_exp(-(dx * dx + dy * dy) / (R(2) * sg * sg));
    }
    case 3: return mitchell_1d(dx * R(2)) * mitchell_1d(dy * R(2));
    default: return R(1);
  }
}

// Per-frame constants of the pixel loop (src/render.rs:1571-1596).  In a
// wavefront chunk one FrameDev describes one part: a range of whole sample
// passes [item_base / n, item_base / n + n_items / n) of `scene` over its n
// pixels: the whole crop (pix_list null, n = crop_w * crop_h) or the crop
// pixels listed in pix_list (n = n_list; adaptive rounds, row bands).
enum : uint32_t { FRAME_STRATIFIED = 0, FRAME_CENTER = 1 };
constexpr uint32_t kCenterSample = 0xffffffffu;  // stream of a pixel's single centre sample
struct FrameDev {
  uint32_t full_w, full_h;
  uint32_t crop_x, crop_y, crop_w, crop_h;
  uint32_t sqrt_spp, spp_a;
  uint32_t max_depth, filter;
  uint64_t fkey;
  uint64_t n_items;    // work items of the part (whole frame: crop_w * crop_h * spp_a)
  uint64_t item_base;  // first frame-local work item of the part
  uint32_t scene;      // index into the batch's scene views
  uint32_t mode;       // FRAME_STRATIFIED, or FRAME_CENTER: one sample through the pixel centre
  const uint32_t* pix_list;  // crop pixel of each list entry, or null (every crop pixel in order)
  uint32_t n_list;
  uint32_t tile_w;  // without a list: 0 = crop pixels in row order; else the crop width, pixels in 8x8 tiles
};

// Crop pixel of entry e when the entries run through 8x8 tiles, the tiles
// in row order (w: crop width, a multiple of 8; the crop height too): the 64
// work items of a wave are then the pixels of one tile, whose camera rays
// traverse nearly the same nodes.
LMN_HD uint32_t tile_pixel(uint32_t e, uint32_t w) {
  const uint32_t tile = e >> 6, in = e & 63u, tiles_x = w >> 3;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  return ((ty << 3) | (in >> 3)) * w + ((tx << 3) | (in & 7u));
}

// Generates work item k = sample * n + entry (n pixels, see FrameDev): its
// camera ray, RNG stream and filter weight (src/render.rs:1683-1690).  In
// FRAME_CENTER mode the ray goes through the pixel centre, without jitter
// draws, with weight 1: the sample a row gets once the time limit has passed
// (src/render.rs:1614-1624).
template <class R>
LMN_HD void generate_item(const FrameDev& F, const CameraT<R>& C, uint64_t k, Rng& rng, V3<R>& o, V3<R>& d, R& w,
                          uint32_t& pixel) {
  uint64_t n = F.pix_list ? uint64_t(F.n_list) : uint64_t(F.crop_w) * F.crop_h;
  uint32_t s = uint32_t(k / n);
  uint32_t e = uint32_t(k - uint64_t(s) * n);
  uint32_t pl = F.pix_list ? F.pix_list[e] : F.tile_w ? tile_pixel(e, F.tile_w) : e;
  uint32_t ci = pl % F.crop_w, cj = pl / F.crop_w;
  uint32_t gi = ci + F.crop_x, gj = cj + F.crop_y;
  uint64_t gpix = uint64_t(gj) * F.full_w + gi;
  if (F.mode == FRAME_CENTER) {
    rng.key = path_key(F.fkey, gpix, kCenterSample);
    rng.n = 0;
    R uc = (R(gi) + R(0.5)) / R(F.full_w - 1);
    R vc = (R(F.full_h - 1 - gj) + R(0.5)) / R(F.full_h - 1);
    camera_ray(C, uc, vc, rng, o, d);
    w = R(1);
    pixel = pl;
    return;
  }
  rng.key = path_key(F.fkey, gpix, s);
  rng.n = 0;
  uint32_t sy = s / F.sqrt_spp, sx = s - sy * F.sqrt_spp;
  R y = R(F.full_h - 1 - gj);
  R jx = (R(sx) + rng.next<R>()) / R(F.sqrt_spp);
  R jy = (R(sy) + rng.next<R>()) / R(F.sqrt_spp);
  R uc = (R(gi) + jx) / R(F.full_w - 1);
  R vc = (y + jy) / R(F.full_h - 1);
  camera_ray(C, uc, vc, rng, o, d);
  w = filter_weight<R>(F.filter, jx - R(0.5), jy - R(0.5));
  pixel = pl;
}

}  // namespace lmn

// Device scene layout and the primitive intersection library.
//
// Layout in HBM (one scene):
//   nodes   : binary SAH BVH with sibling pairs: root at 0, index 1 unused,
//             the two children of a node side by side at an even index.
//             FP32 nodes are 32 bytes {lo.xyz, a, hi.xyz, b}, FP64 nodes 64.
//             internal: a = first child, b = split axis; leaf: a = first
//             prim, b = LEAF_BIT | prim count.  The FP64 path traverses it.
//   nodes4  : the FP32 path's 4-wide BVH, collapsed from the binary one:
//             64-byte nodes holding the boxes of up to four children,
//             quantised to 8 bits per plane on a per-node grid (Node4Q).
//             Two nodes map onto the same primitive records.
//   prims   : primitive records reordered into leaf order (64 B FP32 /
//             128 B FP64), so a leaf's primitives are one contiguous fetch.
//   planes  : unbounded primitives (infinite planes), tested after the BVH
//             exactly as SceneWorld::hit does (src/scene.rs:817-834).
//   xforms  : instance transforms (rotations/translation/scale).
// Intersection routines follow the reference's hit() arithmetic
// (file:line at each function) in either precision.
#pragma once
#include "lmn_math.h"

namespace lmn {

enum : uint32_t {
  K_SPHERE = 0, K_QUAD = 1, K_RECT_XY = 2, K_RECT_XZ = 3, K_RECT_YZ = 4, K_TRIANGLE = 5,
  K_PLANE = 6, K_TORUS = 7, K_SUPERELLIPSOID = 8, K_SPRING = 9, K_DISK = 10, K_CYLINDER = 11, K_CONE = 12,
  K_ELLIPSOID = 13, K_HEMISPHERE = 14, K_ANNULUS = 15, K_FACET = 16, K_WEDGE_FACE = 17,
  K_ROUNDED_BOX = 18, K_MOBIUS = 19
};
// Torus, superellipsoid, spring: the marched kinds (own BVH, FP64 roots).
LMN_HD bool is_marched_kind(uint32_t k) { return k >= K_TORUS && k <= K_SPRING; }
// Disk, cylinder, cone, ellipsoid, hemisphere, annulus, prism and pyramid
// facets, wedge faces, and the sphere-traced rounded box and Mobius strip:
// rare kinds kept in the marched kinds' BVH
// and evaluated there in FP64, out of line, so the main traversal carries
// none of their code.
LMN_HD bool is_quadric_kind(uint32_t k) { return k >= K_DISK; }
constexpr uint32_t KIND_MASK = 0x1f;
constexpr uint32_t FLAG_FP64 = 0x20;  // FP32 path evaluates this primitive in FP64
constexpr uint32_t FLAG_SMOOTH = 0x40;  // triangle with vertex attributes (p[9]: AttrT index)
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t XF_NONE = 0xffffffffu;

// Parameter slots p[]:
//   SPHERE c.xyz r | QUAD q.xyz u.xyz v.xyz w.xyz D (w = n/(n.n), D = unit(n).q)
//   RECT_* a0 a1 b0 b1 k | TRIANGLE v0.xyz e1.xyz e2.xyz
//   PLANE point.xyz normal.xyz uaxis.xyz vaxis.xyz | TORUS c.xyz R r
//   SUPERELLIPSOID c.xyz s.xyz e1 e2 | SPRING c.xyz coil tube pitch turns
//   DISK c.xyz unit-normal.xyz r | CYLINDER, CONE c.xyz r y_min y_max
//   ELLIPSOID c.xyz radii.xyz | HEMISPHERE c.xyz r | ANNULUS c.xyz unit-normal.xyz r_in r_out
//   FACET v0.xyz e1.xyz e2.xyz normal.xyz | WEDGE_FACE min.xyz max.xyz face
//   ROUNDED_BOX c.xyz half.xyz r | MOBIUS c.xyz radius half-width thickness
template <class R>
struct alignas(16) PrimT {
  R p[13];
  uint32_t kind_mat;  // kind | flags | material << 8
  uint32_t xform;     // transform index or XF_NONE
  uint32_t prim_id;   // index into the descriptor's prim list
};
static_assert(sizeof(PrimT<float>) == 64, "FP32 primitive record must be 64 bytes");
static_assert(sizeof(PrimT<double>) == 128, "FP64 primitive record must be 128 bytes");

// Vertex attributes of a smooth mesh triangle (MeshTriangle, src/obj.rs:250-313):
// the corners' normals and texture coordinates.
template <class R>
struct alignas(16) AttrT {
  R n[9];          // n0, n1, n2
  R uv[6];         // uv0, uv1, uv2
  uint32_t flags;  // 1: normals, 2: uvs
};
static_assert(sizeof(AttrT<float>) == 64, "FP32 vertex attributes must be 64 bytes");
static_assert(sizeof(AttrT<double>) == 128, "FP64 vertex attributes must be 128 bytes");
// The AttrT index of a FLAG_SMOOTH triangle record: float records keep its
// bits in p[9], double records its value.
LMN_HD uint32_t attr_index(const PrimT<float>& r) { return bits_of(r.p[9]); }
LMN_HD uint32_t attr_index(const PrimT<double>& r) { return uint32_t(r.p[9]); }
LMN_HD void set_attr_index(PrimT<float>& r, uint32_t i) { r.p[9] = float_of(i); }
LMN_HD void set_attr_index(PrimT<double>& r, uint32_t i) { r.p[9] = double(i); }

struct NodeF {
  float lo[3];
  uint32_t a;
  float hi[3];
  uint32_t b;
};
static_assert(sizeof(NodeF) == 32, "FP32 BVH node must be 32 bytes");
struct alignas(16) NodeD {
  double lo[3];
  double hi[3];
  uint32_t a, b;
};
static_assert(sizeof(NodeD) == 64, "FP64 BVH node must be 64 bytes");
// FP32 4-wide BVH node as the host builds it, before quantisation (Node4Q):
// the FP32 boxes of up to four children in SoA order and the child entries:
// a node index, or a leaf LEAF4 | (count - 1) << 28 | first primitive record.
struct alignas(128) Node4F {
  float lox[4], loy[4], loz[4];
  float hix[4], hiy[4], hiz[4];
  uint32_t child[4];
  uint32_t n;  // children in use, 1..4
  uint32_t pad[3];
};
static_assert(sizeof(Node4F) == 128, "FP32 BVH4 node must be 128 bytes");
// The FP32 path's 4-wide node as the device reads it: 64 bytes, four 16-byte
// loads.  Child boxes are quantised to 8 bits per plane on a per-node grid
// of step 2^e per axis: plane(q) = fma(q, 2^e, org), rounded outward on the
// host so that every quantised box contains its child's FP32 box, with at
// least one grid step to spare except where the plane is the grid origin
// itself (there the plane and its slab distance are the unquantised ones).
//   w0: org.x org.y org.z qlo.x      w1: qlo.y qlo.z qhi.x qhi.y
//   w2: qhi.z child0 child1 child2   w3: child3 meta  pad pad
// qlo.a / qhi.a hold child k's lower / upper plane index of axis a in byte
// k; meta holds the exponent e of axis a's grid step 2^e as a signed byte
// in byte a, and the number of children in use in byte 3.  Every word the box test
// needs sits in a different 16-byte quarter, so all four loads are issued
// together and the node costs one memory round trip.
struct alignas(64) Node4Q {
  float org[3];
  uint32_t qlo_x;
  uint32_t qlo_y, qlo_z, qhi_x, qhi_y;
  uint32_t qhi_z, child0, child1, child2;
  uint32_t child3, meta, pad0, pad1;
};
static_assert(sizeof(Node4Q) == 64, "quantised FP32 BVH4 node must be 64 bytes");
constexpr uint32_t LEAF4 = 0x80000000u;
constexpr uint32_t LEAF4_FIRST = 0x0fffffffu;  // records addressable by a leaf entry
constexpr uint32_t LEAF4_MAX = 8;              // primitives per leaf entry

template <class R> struct NodeSel;
template <> struct NodeSel<float> { using type = NodeF; };
template <> struct NodeSel<double> { using type = NodeD; };

// Instance transform: composed wrappers of src/transform.rs.  `inv` maps a
// world ray into object space, `nrm` maps an object normal back.  The
// reference's Scale / NonUniformScale rescale the reported t (a quirk,
// src/transform.rs:277, :336): tmode 1 multiplies by tscale, tmode 2 by
// |d_obj| / |d_world|.
template <class R>
struct XformT {
  R inv[12];
  R nrm[9];
  R tscale;
  uint32_t tmode;
  uint32_t renorm;  // normalise the mapped normal (non-uniform scale)
};

template <class R> LMN_HD V3<R> xf_point(const R* m, V3<R> p) {
  return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
          m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
template <class R> LMN_HD V3<R> xf_dir(const R* m, V3<R> d) {
  return {m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
          m[8] * d.x + m[9] * d.y + m[10] * d.z};
}
template <class R> LMN_HD V3<R> xf_nrm(const R* m, V3<R> n) {
  return {m[0] * n.x + m[1] * n.y + m[2] * n.z, m[3] * n.x + m[4] * n.y + m[5] * n.z,
          m[6] * n.x + m[7] * n.y + m[8] * n.z};
}

template <class R> LMN_HD R t_report_factor(const XformT<R>& x, V3<R> d_obj, V3<R> d_world) {
  if (x.tmode == 1) return x.tscale;
  if (x.tmode == 2) return length(d_obj) / length(d_world);
  return R(1);
}

// ---------------------------------------------------------------------------
// Analytic primitives.  The hit routines return only t (object-space t ==
// world t for the affine wrappers); the reference's (u, v) are recomputed from
// the hit point when a hit is shaded (prim_uv), which keeps the traversal
// loop's register footprint small.

// Sphere::hit (src/sphere.rs:24-62)
template <class R>
LMN_HD bool hit_sphere(const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t, bool far_only = false) {
  V3<R> c = {p[0], p[1], p[2]};
  R r = p[3];
  V3<R> oc = o - c;
  R a = length_squared(d);
  R half_b = dot(oc, d);
  R cc = length_squared(oc) - r * r;
  R root;
  if constexpr (sizeof(R) == 4) {
    // FP32: the discriminant from the ray's perpendicular distance to the
    // centre and the roots as c/q, q/a (q = -b -+ sqrt(disc)) avoid the
    // cancellation of half_b^2 - a c for small or distant spheres
    // (Haines et al., "Precision Improvements for Ray/Sphere Intersection",
    // Ray Tracing Gems 2019).  Same roots, same near-then-far selection.
    V3<R> l = oc - d * (half_b / a);
    R disc = a * (r * r - length_squared(l));
    if (disc < R(0)) return false;
    R sq = vsqrt(disc);
    R q = half_b >= R(0) ? -half_b - sq : -half_b + sq;
    R t0 = cc / q, t1 = q / a;
    R tn = t0 < t1 ? t0 : t1, tf = t0 < t1 ? t1 : t0;
    if (q == R(0)) tn = tf = -half_b / a;
    root = far_only ? tf : tn;
    if (root < tmin || root > tmax) {
      root = tf;
      if (root < tmin || root > tmax) return false;
    }
  } else {  // the reference's half-b quadratic, unchanged
    R disc = half_b * half_b - a * cc;
    if (disc < R(0)) return false;
    R sq = vsqrt(disc);
    root = far_only ? (-half_b + sq) / a : (-half_b - sq) / a;
    if (root < tmin || root > tmax) {
      root = (-half_b + sq) / a;
      if (root < tmin || root > tmax) return false;
    }
  }
  if (!(root == root)) return false;
  t = root;
  return true;
}

// Quad::hit (src/quad.rs:30-62); alpha/beta from the planar hit vector
template <class R>
LMN_HD void quad_ab(const R* p, V3<R> planar, R& alpha, R& beta) {
  V3<R> uu = {p[3], p[4], p[5]}, vv = {p[6], p[7], p[8]}, w = {p[9], p[10], p[11]};
  alpha = dot(w, cross(planar, vv));
  beta = dot(w, cross(uu, planar));
}
template <class R>
LMN_HD bool hit_quad(const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
  V3<R> q = {p[0], p[1], p[2]}, w = {p[9], p[10], p[11]};
  V3<R> n = unit(w);
  R denom = dot(n, d);
  if (m_abs(denom) < R(1e-8)) return false;
  R tt = (p[12] - dot(n, o)) / denom;
  if (tt < tmin || tt > tmax || !(tt == tt)) return false;
  R alpha, beta;
  quad_ab(p, (o + d * tt) - q, alpha, beta);
  if (!(alpha >= R(0) && alpha <= R(1)) || !(beta >= R(0) && beta <= R(1))) return false;
  t = tt;
  return true;
}

// XyRect / XzRect / YzRect::hit (src/rect.rs:24-37, :64-77, :104-117).  `ax`
// is the plane axis, (ia, ib) the in-plane axes.
template <class R>
LMN_HD bool hit_rect(const R* p, int ax, int ia, int ib, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
  const R* oo = &o.x;
  const R* dd = &d.x;
  R tt = (p[4] - oo[ax]) / dd[ax];
  if (tt < tmin || tt > tmax || !(tt == tt)) return false;
  R a = oo[ia] + tt * dd[ia];
  R b = oo[ib] + tt * dd[ib];
  if (a < p[0] || a > p[1] || b < p[2] || b > p[3]) return false;
  t = tt;
  return true;
}

// Triangle::hit / MeshTriangle::hit, Moller-Trumbore
// (src/triangle.rs:23-64, src/obj.rs:260-313); u, v are the barycentrics.
template <class R>
LMN_HD bool hit_triangle_uv(const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t, R& u, R& v) {
  V3<R> v0 = {p[0], p[1], p[2]}, e1 = {p[3], p[4], p[5]}, e2 = {p[6], p[7], p[8]};
  V3<R> h = cross(d, e2);
  R a = dot(e1, h);
  if (m_abs(a) < R(1e-8)) return false;
  R f = R(1) / a;
  V3<R> s = o - v0;
  R uu = f * dot(s, h);
  if (!(uu >= R(0) && uu <= R(1))) return false;
  V3<R> q = cross(s, e1);
  R vv = f * dot(d, q);
  if (vv < R(0) || uu + vv > R(1)) return false;
  R tt = f * dot(e2, q);
  if (tt < tmin || tt > tmax || !(tt == tt)) return false;
  t = tt;
  u = uu;
  v = vv;
  return true;
}
template <class R>
LMN_HD bool hit_triangle(const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
  R u, v;
  return hit_triangle_uv(p, o, d, tmin, tmax, t, u, v);
}

// Plane::hit (src/plane.rs:43-69)
template <class R>
LMN_HD bool hit_plane(const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
  V3<R> pt = {p[0], p[1], p[2]}, n = {p[3], p[4], p[5]};
  R denom = dot(n, d);
  if (m_abs(denom) < R(1e-8)) return false;
  R tt = dot(pt - o, n) / denom;
  if (tt < tmin || tt > tmax || !(tt == tt)) return false;
  t = tt;
  return true;
}

// (u, v) of an analytic hit from the object-space ray and its t, with the
// reference's formulas: spherical map (src/sphere.rs:48-52), quad alpha/beta
// (src/quad.rs:43-46), rect fractions (src/rect.rs:34-35), triangle
// barycentrics (src/triangle.rs:35-42), plane fract (src/plane.rs:57-59).
template <class R>
LMN_HD void prim_uv_analytic(uint32_t kind, const R* p, V3<R> o, V3<R> d, R t, R& u, R& v) {
  const R pi = R(3.14159265358979323846);
  switch (kind) {
    case K_SPHERE: {
      V3<R> n = ((o + d * t) - V3<R>{p[0], p[1], p[2]}) / p[3];
      u = (m_atan2(-n.z, n.x) + pi) / (R(2) * pi);
      v = m_acos(-n.y) / pi;
      break;
    }
    case K_QUAD: quad_ab(p, (o + d * t) - V3<R>{p[0], p[1], p[2]}, u, v); break;
    case K_RECT_XY:
    case K_RECT_XZ:
    case K_RECT_YZ: {
      int ia = kind == K_RECT_YZ ? 1 : 0, ib = kind == K_RECT_XY ? 1 : 2;
      const R* oo = &o.x;
      const R* dd = &d.x;
      u = (oo[ia] + t * dd[ia] - p[0]) / (p[1] - p[0]);
      v = (oo[ib] + t * dd[ib] - p[2]) / (p[3] - p[2]);
      break;
    }
    case K_TRIANGLE: {
      R tt;
      if (!hit_triangle_uv(p, o, d, R(-INFINITY), R(INFINITY), tt, u, v)) u = v = R(0);
      break;
    }
    case K_PLANE: {
      V3<R> local = (o + d * t) - V3<R>{p[0], p[1], p[2]};
      R a = dot(local, V3<R>{p[6], p[7], p[8]});
      R b = dot(local, V3<R>{p[9], p[10], p[11]});
      u = m_abs(a - m_trunc(a));  // f64::fract().abs()
      v = m_abs(b - m_trunc(b));
      break;
    }
    default: u = v = R(0);
  }
}

// ---------------------------------------------------------------------------
// Marched primitives, evaluated in FP64 on every path.

LMN_HD double torus_sdf(V3<double> p, double R_, double r_) {
  double q = m_sqrt(p.x * p.x + p.z * p.z) - R_;
  return m_sqrt(q * q + p.y * p.y) - r_;
}

// Torus::hit, reference semantics: fixed-step SDF march with bisection
// (src/torus.rs:24-103).
LMN_HD bool hit_torus_march(const double* p, V3<double> ro, V3<double> d, double tmin, double tmax, double& t) {
  V3<double> o = ro - V3<double>{p[0], p[1], p[2]};
  double R_ = p[3], r_ = p[4];
  double step = r_ * 0.5;
  double ms = fmin((tmax - tmin) / step, 500.0);
  // `as usize` saturates: NaN and negatives become 0
  long max_steps = (ms == ms && ms > 0.0) ? long(ms) : 0;
  double tc = tmin;
  double prev = torus_sdf(o + d * tc, R_, r_);
  for (long i = 0; i < max_steps; ++i) {
    tc += step;
    if (tc > tmax) break;
    double dist = torus_sdf(o + d * tc, R_, r_);
    if (prev > 0.0 && dist < 0.0) {
      double lo = tc - step, hi = tc;
      for (int k = 0; k < 20; ++k) {
        double mid = (lo + hi) * 0.5;
        if (torus_sdf(o + d * mid, R_, r_) > 0.0)
          lo = mid;
        else
          hi = mid;
      }
      double th = (lo + hi) * 0.5;
      if (th < tmin || th > tmax) {
        prev = dist;
        continue;
      }
      t = th;
      return true;
    }
    prev = dist;
  }
  return false;
}

// Real roots of the monic quartic x^4 + a x^3 + b x^2 + c x + e, found as the
// eigen-free way: Ferrari's resolvent cubic solved in closed form, then every
// candidate root polished by Newton on the original quartic.  Returns the
// count and the roots in ascending order.
LMN_HD int solve_cubic_depressed_largest(double p, double q, double& y) {
  // largest real root of y^3 + p y + q = 0
  double disc = q * q / 4.0 + p * p * p / 27.0;
  if (disc >= 0.0) {
    double s = m_sqrt(disc);
    y = cbrt(-q / 2.0 + s) + cbrt(-q / 2.0 - s);
  } else {
    double rr = m_sqrt(-p / 3.0);
    double ang = m_acos(fmax(-1.0, fmin(1.0, -q / (2.0 * rr * rr * rr))));
    y = 2.0 * rr * m_cos(ang / 3.0);
  }
  return 1;
}

LMN_HD int solve_quartic(double a, double b, double c, double e, double roots[4]) {
  // depressed quartic y^4 + P y^2 + Q y + S via x = y - a/4
  double a2 = a * a;
  double P = b - 3.0 * a2 / 8.0;
  double Q = c - a * b / 2.0 + a2 * a / 8.0;
  double S = e - a * c / 4.0 + a2 * b / 16.0 - 3.0 * a2 * a2 / 256.0;
  int n = 0;
  double ys[4];
  if (m_abs(Q) < 1e-14 * (1.0 + m_abs(P) + m_abs(S))) {
    // biquadratic
    double disc = P * P - 4.0 * S;
    if (disc >= 0.0) {
      double sd = m_sqrt(disc);
      double z1 = (-P - sd) / 2.0, z2 = (-P + sd) / 2.0;
      if (z1 >= 0.0) { ys[n++] = -m_sqrt(z1); ys[n++] = m_sqrt(z1); }
      if (z2 >= 0.0) { ys[n++] = -m_sqrt(z2); ys[n++] = m_sqrt(z2); }
    }
  } else {
    // resolvent: m^3 + P m^2 + (P^2/4 - S) m - Q^2/8 = 0, need m > 0
    double rb = P, rc = P * P / 4.0 - S, rd = -Q * Q / 8.0;
    double pp = rc - rb * rb / 3.0;
    double qq = 2.0 * rb * rb * rb / 27.0 - rb * rc / 3.0 + rd;
    double yv;
    solve_cubic_depressed_largest(pp, qq, yv);
    double m = yv - rb / 3.0;
    // polish m on the resolvent
    for (int it = 0; it < 4; ++it) {
      double f = ((m + rb) * m + rc) * m + rd;
      double fp = (3.0 * m + 2.0 * rb) * m + rc;
      if (fp == 0.0) break;
      double nm = m - f / fp;
      if (!(nm == nm)) break;
      m = nm;
    }
    if (m > 0.0) {
      double sm = m_sqrt(2.0 * m);
      double k1 = -(2.0 * P + 2.0 * m);
      double k2 = 2.0 * Q / sm;
      double d1 = k1 - k2, d2 = k1 + k2;
      if (d1 >= 0.0) { double s1 = m_sqrt(d1); ys[n++] = (sm - s1) / 2.0; ys[n++] = (sm + s1) / 2.0; }
      if (d2 >= 0.0) { double s2 = m_sqrt(d2); ys[n++] = (-sm - s2) / 2.0; ys[n++] = (-sm + s2) / 2.0; }
    }
  }
  for (int i = 0; i < n; ++i) {
    double x = ys[i] - a / 4.0;
    for (int it = 0; it < 3; ++it) {  // Newton polish on the original quartic
      double f = (((x + a) * x + b) * x + c) * x + e;
      double fp = ((4.0 * x + 3.0 * a) * x + 2.0 * b) * x + c;
      if (fp == 0.0) break;
      double nx = x - f / fp;
      if (!(nx == nx)) break;
      x = nx;
    }
    roots[i] = x;
  }
  for (int i = 1; i < n; ++i)  // insertion sort
    for (int j = i; j > 0 && roots[j] < roots[j - 1]; --j) {
      double tmp = roots[j];
      roots[j] = roots[j - 1];
      roots[j - 1] = tmp;
    }
  return n;
}

// Torus, exact intersection: smallest root of the quartic
// (|p|^2 + R^2 - r^2)^2 = 4 R^2 (x^2 + z^2) in [tmin, tmax] (FP64).
LMN_HD bool hit_torus_exact(const double* p, V3<double> ro, V3<double> d, double tmin, double tmax, double& t) {
  V3<double> o = ro - V3<double>{p[0], p[1], p[2]};
  double R_ = p[3], r_ = p[4];
  // normalise the direction so the quartic is well scaled, map t back after
  double dl = length(d);
  V3<double> dn = d / dl;
  double od = dot(o, dn), oo = length_squared(o);
  double k = oo + R_ * R_ - r_ * r_;
  double R2 = R_ * R_;
  // (s^2 + 2 od s + k)^2 - 4R^2((ox + s dx)^2 + (oz + s dz)^2)
  double c3 = 4.0 * od;
  double c2 = 4.0 * od * od + 2.0 * k - 4.0 * R2 * (dn.x * dn.x + dn.z * dn.z);
  double c1 = 4.0 * od * k - 8.0 * R2 * (o.x * dn.x + o.z * dn.z);
  double c0 = k * k - 4.0 * R2 * (o.x * o.x + o.z * o.z);
  double roots[4];
  int n = solve_quartic(c3, c2, c1, c0, roots);
  for (int i = 0; i < n; ++i) {
    double tt = roots[i] / dl;
    if (tt >= tmin && tt <= tmax) {
      t = tt;
      return true;
    }
  }
  return false;
}

LMN_HD double superellipsoid_f(V3<double> q, double e1, double e2) {
  double ax = m_abs(q.x), ay = m_abs(q.y), az = m_abs(q.z);
  double exp2 = 2.0 / e2;
  double xz = pow(ax, exp2) + pow(az, exp2);
  double exp1 = 2.0 / e1;
  double val = pow(xz, e2 / e1) + pow(ay, exp1);
  return pow(val, e1 / 2.0) - 1.0;
}

// Superellipsoid::hit, reference semantics (src/superellipsoid.rs:26-90)
LMN_HD bool hit_superellipsoid_march(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax,
                                     double& t) {
  V3<double> o = {(ro.x - p[0]) / p[3], (ro.y - p[1]) / p[4], (ro.z - p[2]) / p[5]};
  V3<double> d = {rd.x / p[3], rd.y / p[4], rd.z / p[5]};
  double tc = tmin;
  for (int i = 0; i < 256; ++i) {
    if (tc > tmax) break;
    double dist = superellipsoid_f(o + d * tc, p[6], p[7]);
    if (m_abs(dist) < 1e-5) {
      t = tc;
      return true;
    }
    tc += fmax(dist, 1e-4);
  }
  return false;
}

// Superellipsoid, exact mode: the first crossing of the implicit surface
// val(q) = (|x|^(2/e2) + |z|^(2/e2))^(e2/e1) + |y|^(2/e1) = 1 along the ray
// (val is the reference's f + 1 raised to 2/e1, a monotone map, so the
// surface and the inside are the same), FP64.  Roots are isolated by
// interval arithmetic on a binary subdivision of the chord through the unit
// box (Mitchell, "Robust ray intersection with interval arithmetic", 1990):
// every term is monotone in |x|, |y|, |z| or in the partial sums, so the
// range of val over a sub-interval of t follows from the ranges of the
// coordinates.  Sub-intervals whose range excludes 1 hold no crossing and are
// skipped whole; the first interval of at most 1/4096 of the chord with a
// sign change across it is refined by bracketed Newton iteration.
// Intervals that may hold a crossing but show no sign change keep splitting
// down to 2^-24 of the chord, so slivers thinner than any fixed sampling are
// found; a touch without a sign change at that scale is no hit.  The
// interval bounds are screened first in FP32 with the hardware log2/exp2:
// their error near val = 1 is below 1e-5 for exponents up to 10 and
// coordinates inside the padded box, and an interval counts as ruled out
// only with a margin of 5e-4 on top; where a screened bound falls inside
// that band the FP64 bounds decide.  Below the first levels, intervals the
// plain bounds cannot rule out get the mean-value form as well (rays that
// graze the surface).  All sign evaluations are FP64.
LMN_HD double se_val(V3<double> q, double px, double qa, double py) {
  double a = pow_pos(m_abs(q.x), px) + pow_pos(m_abs(q.z), px);
  return pow_pos(a, qa) + pow_pos(m_abs(q.y), py);
}
// val at o + d t and its derivative in t (the terms' derivatives follow from
// their values: d|x|^p/dt = p |x|^p / x * dx/dt)
LMN_HD double se_val_d(V3<double> o, V3<double> d, double t, double px, double qa, double py, double& dv) {
  V3<double> q = o + d * t;
  double tx = pow_pos(m_abs(q.x), px), tz = pow_pos(m_abs(q.z), px), ty = pow_pos(m_abs(q.y), py);
  double A = tx + tz, tA = pow_pos(A, qa);
  double dtx = q.x != 0.0 ? px * tx / q.x * d.x : 0.0;
  double dtz = q.z != 0.0 ? px * tz / q.z * d.z : 0.0;
  double dty = q.y != 0.0 ? py * ty / q.y * d.y : 0.0;
  dv = (A > 0.0 ? qa * tA / A * (dtx + dtz) : 0.0) + dty;
  return tA + ty;
}
LMN_HD float pow_screen(float x, float y) { return exp2f(y * log2f(x)); }

// Range of |o + d t| over [ta, tb].
LMN_HD void se_range1(double o, double d, double ta, double tb, double& lo, double& hi) {
  double a = o + d * ta, b = o + d * tb;
  double amin = a < b ? a : b, amax = a < b ? b : a;
  if (amin <= 0.0 && amax >= 0.0) {
    lo = 0.0;
    hi = m_abs(amin) > m_abs(amax) ? m_abs(amin) : m_abs(amax);
  } else {
    lo = m_abs(amin) < m_abs(amax) ? m_abs(amin) : m_abs(amax);
    hi = m_abs(amin) < m_abs(amax) ? m_abs(amax) : m_abs(amin);
  }
}

// FP32 terms of the interval bound over a sub-chord: the coordinate
// magnitudes' ranges [ul, uh], the terms |c|^p at both ends of each range,
// and A = X + Z, B = A^qa at the ends of A's range.
struct SeScreen {
  float xl, xh, yl, yh, zl, zh;
  float txl, txh, tyl, tyh, tzl, tzh;
  float Al, Ah, bl, bh;
};

// Sign of val - 1 over [a, b] by the mean-value form val([a, b]) in
// val(m) +- (b - a)/2 * max|val'([a, b])|, for sub-chords on which no
// coordinate changes sign: +1 / -1 where val stays above / below 1, else 0.
// The plain interval bound adds up the variation of every term and cannot
// rule out sub-chords of rays that graze the surface (terms that cancel
// along the ray, as on rays leaving the surface at a shallow angle); this
// form's width shrinks like (b - a)^2 plus the true slope.  The slope bound
// comes from the FP32 terms: val' = qa A^(qa-1) A' + Y', A' = X' + Z', each
// term's derivative p |c|^(p-1) d|c|/dt ranging between its values at the
// sub-chord's ends (|c|^(p-1) = |c|^p / |c|), inflated by 1 % for the FP32
// error; val(m) is screened in FP32 like the plain bounds, and where that
// cannot decide the walk evaluates it in FP64 (se_mean_value_decide).
LMN_HD int se_mean_value_screen(V3<double> o, V3<double> d, double a, double b, float pxf, float qaf, float pyf,
                                const SeScreen& S, float margin, double& m, double& r) {
  const float kTiny = 1e-30f;  // terms of smaller magnitudes may have flushed to zero
  if (!(S.xl > kTiny && S.yl > kTiny && S.zl > kTiny && S.Al > kTiny)) return 0;
  m = 0.5 * (a + b);
  const V3<double> q = o + d * m;
  auto dterm = [](float p, float tl, float th, float ul, float uh, float k, float& lo, float& hi) {
    float g1 = p * (tl / ul), g2 = p * (th / uh);
    float gl = g1 < g2 ? g1 : g2, gh = g1 < g2 ? g2 : g1;
    lo = k >= 0.0f ? gl * k : gh * k;
    hi = k >= 0.0f ? gh * k : gl * k;
  };
  float xdl, xdh, zdl, zdh, ydl, ydh;
  dterm(pxf, S.txl, S.txh, S.xl, S.xh, float(q.x > 0.0 ? d.x : -d.x), xdl, xdh);
  dterm(pxf, S.tzl, S.tzh, S.zl, S.zh, float(q.z > 0.0 ? d.z : -d.z), zdl, zdh);
  dterm(pyf, S.tyl, S.tyh, S.yl, S.yh, float(q.y > 0.0 ? d.y : -d.y), ydl, ydh);
  const float adl = xdl + zdl, adh = xdh + zdh;
  const float P1 = S.bl / S.Al, P2 = S.bh / S.Ah;  // A^(qa-1) at the ends of A's range
  const float Pl = P1 < P2 ? P1 : P2, Ph = P1 < P2 ? P2 : P1;
  float bdl, bdh;  // A^(qa-1) A', with A^(qa-1) >= 0
  if (adl >= 0.0f) {
    bdl = Pl * adl;
    bdh = Ph * adh;
  } else if (adh <= 0.0f) {
    bdl = Ph * adl;
    bdh = Pl * adh;
  } else {
    bdl = Ph * adl;
    bdh = Ph * adh;
  }
  const float vdl = qaf * bdl + ydl, vdh = qaf * bdh + ydh;
  const float slope = fabsf(vdl) > fabsf(vdh) ? fabsf(vdl) : fabsf(vdh);
  if (!(slope < 1e30f)) return 0;  // also NaN
  // the centre value in FP32 first, with the screen's margin: clear of the
  // band on both sides, the FP64 form would straddle 1 as well
  const float r32 = 0.5f * float(b - a) * slope * 1.01f;
  const float vm32 = pow_screen(pow_screen(fabsf(float(q.x)), pxf) + pow_screen(fabsf(float(q.z)), pxf), qaf) +
                     pow_screen(fabsf(float(q.y)), pyf);
  if (vm32 - r32 > 1.0f + margin) return 1;
  if (vm32 + r32 < 1.0f - margin) return -1;
  if (vm32 - r32 < 1.0f - margin && vm32 + r32 > 1.0f + margin) return 0;
  r = 0.5 * (b - a) * double(slope) * 1.01;
  return 2;  // the FP64 centre value decides: se_mean_value_decide
}
LMN_HD int se_mean_value_decide(double vm, double r) {
  r += 1e-12 * (1.0 + vm);
  if (vm - r > 1.0 + 1e-12) return 1;
  if (vm + r < 1.0 - 1e-12) return -1;
  return 0;
}

// Object-space ray of a superellipsoid and its chord [t0, t1] through the
// unit box |q|_inf <= 1 that contains the surface, padded so that a ray
// entering where the surface touches the box (the axis points) starts
// strictly outside and sees the sign change; false if the chord is empty.
LMN_HD bool se_chord(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax, V3<double>& o,
                     V3<double>& d, double& t0, double& t1) {
  o = {(ro.x - p[0]) / p[3], (ro.y - p[1]) / p[4], (ro.z - p[2]) / p[5]};
  d = {rd.x / p[3], rd.y / p[4], rd.z / p[5]};
  const double B = 1.0 + 1e-9;
  t0 = tmin;
  t1 = tmax;
  const double* oo = &o.x;
  const double* dd = &d.x;
  for (int k = 0; k < 3; ++k) {
    double inv = 1.0 / dd[k];
    double a = (-B - oo[k]) * inv, b = (B - oo[k]) * inv;
    if (a > b) { double tmp = a; a = b; b = tmp; }
    t0 = fmax(t0, a);
    t1 = fmin(t1, b);
  }
  return t0 <= t1;
}

constexpr float kSeScreenMargin = 5e-4f;

// FP32 screen of the interval walk below, on its own: false where the
// screened bounds rule out every sub-chord down to `max_level`, i.e. where
// hit_superellipsoid_exact finds no crossing either (the screen only rules
// out sub-chords whose range excludes 1 with the margin).  The marched
// pass runs it on every candidate and the exact walk only on the rest,
// whose costs are closer to each other.
LMN_HD bool se_may_cross(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax, int max_level) {
  V3<double> o, d;
  double t0, t1;
  if (!se_chord(p, ro, rd, tmin, tmax, o, d, t0, t1)) return false;
  const float pxf = float(2.0 / p[7]), qaf = float(p[7] / p[6]), pyf = float(2.0 / p[6]);
  const double w = t1 - t0;
  int level = 0;
  uint32_t idx = 0;
  for (;;) {
    double scale = w / double(1u << level);
    double a = t0 + double(idx) * scale, b = (idx + 1u == (1u << level)) ? t1 : a + scale;
    double xl, xh, yl, yh, zl, zh;
    se_range1(o.x, d.x, a, b, xl, xh);
    se_range1(o.y, d.y, a, b, yl, yh);
    se_range1(o.z, d.z, a, b, zl, zh);
    float slo = pow_screen(pow_screen(float(xl), pxf) + pow_screen(float(zl), pxf), qaf) + pow_screen(float(yl), pyf);
    float shi = pow_screen(pow_screen(float(xh), pxf) + pow_screen(float(zh), pxf), qaf) + pow_screen(float(yh), pyf);
    if (!(slo > 1.0f + kSeScreenMargin || shi < 1.0f - kSeScreenMargin)) {
      if (level == max_level) return true;
      ++level;
      idx *= 2u;
      continue;
    }
    while (idx & 1u) {
      idx >>= 1;
      --level;
    }
    if (level == 0) return false;
    ++idx;
  }
}

// The first sign-change leaf [a, b] of the walk (fa, fb: val - 1 at its
// ends), found by se_exact_walk and refined by se_refine.  The root lies
// inside the leaf and the leaf inside [tmin, tmax], so the bracket alone
// answers a visibility query.
struct SeBracket {
  double a, b, fa, fb;
};

LMN_HD bool se_exact_walk(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax, SeBracket& br) {
  V3<double> o, d;
  double t0, t1;
  if (!se_chord(p, ro, rd, tmin, tmax, o, d, t0, t1)) return false;
  const double px = 2.0 / p[7], qa = p[7] / p[6], py = 2.0 / p[6];
  const float pxf = float(px), qaf = float(qa), pyf = float(py);
  const float kScreenMargin = kSeScreenMargin;
  const int kSolveLevel = 12, kMaxLevel = 24, kMeanValueLevel = 4;
  const double w = t1 - t0;
  // Every FP64 evaluation of the walk is val at one point of the chord (the
  // chord's start, a sub-chord's centre for the mean-value form, a leaf's
  // end), made at one place per step: lanes of a wave that need one at
  // different stages of their walks share it.
  enum : int { NEED_NONE, NEED_START, NEED_CENTRE, NEED_LEAF };
  int need = NEED_START;
  double fa = 0.0;  // val - 1 at the left end of the current interval (a sign once ruled out)
  double r_mv = 0.0;
  bool leaf = false;  // the current interval is past its screens: its right end decides
  int level = 0;
  uint32_t idx = 0;
  for (;;) {
    double scale = w / double(1u << level);
    double a = t0 + double(idx) * scale, b = (idx + 1u == (1u << level)) ? t1 : a + scale;
    int sign = 0;
    double t_eval = t0;
    if (need == NEED_NONE && !leaf) {
      double xl, xh, yl, yh, zl, zh;
      se_range1(o.x, d.x, a, b, xl, xh);
      se_range1(o.y, d.y, a, b, yl, yh);
      se_range1(o.z, d.z, a, b, zl, zh);
      SeScreen S;
      S.xl = float(xl), S.xh = float(xh), S.yl = float(yl), S.yh = float(yh), S.zl = float(zl), S.zh = float(zh);
      S.txl = pow_screen(S.xl, pxf), S.txh = pow_screen(S.xh, pxf), S.tzl = pow_screen(S.zl, pxf);
      S.tzh = pow_screen(S.zh, pxf), S.tyl = pow_screen(S.yl, pyf), S.tyh = pow_screen(S.yh, pyf);
      S.Al = S.txl + S.tzl, S.Ah = S.txh + S.tzh;
      S.bl = pow_screen(S.Al, qaf), S.bh = pow_screen(S.Ah, qaf);
      const float slo = S.bl + S.tyl, shi = S.bh + S.tyh;
      sign = slo > 1.0f + kScreenMargin ? 1 : (shi < 1.0f - kScreenMargin ? -1 : 0);
      const bool in_band = sign == 0 && !(slo < 1.0f - kScreenMargin && shi > 1.0f + kScreenMargin);
      if (sign == 0 && (in_band || level >= kMeanValueLevel)) {
        double m;
        int mv = se_mean_value_screen(o, d, a, b, pxf, qaf, pyf, S, kScreenMargin, m, r_mv);
        if (mv == 2) {
          need = NEED_CENTRE;
          t_eval = m;
        } else {
          sign = mv;
        }
      }
      if (sign == 0 && need == NEED_NONE) {
        if (level < kSolveLevel) {  // subdivide: left child first
          ++level;
          idx *= 2u;
          continue;
        }
        leaf = true;
      }
    }
    if (need == NEED_NONE && leaf) {
      need = NEED_LEAF;
      t_eval = b;
    }
    if (need != NEED_NONE) {
      const double v = se_val(o + d * t_eval, px, qa, py);
      const int what = need;
      need = NEED_NONE;
      if (what == NEED_START) {
        fa = v - 1.0;
        continue;
      }
      if (what == NEED_CENTRE) {
        sign = se_mean_value_decide(v, r_mv);
        if (sign == 0) {
          if (level < kSolveLevel) {
            ++level;
            idx *= 2u;
          } else {
            leaf = true;  // its right end next step
          }
          continue;
        }
      } else {  // NEED_LEAF
        leaf = false;
        const double fb = v - 1.0;
        if ((fa > 0.0) == (fb > 0.0) && level < kMaxLevel) {
          // no sign change across the leaf: a sliver entered and left within
          // it, or a grazing touch; keep splitting (down to 2^-24 of the chord)
          ++level;
          idx *= 2u;
          continue;
        }
        if ((fa > 0.0) != (fb > 0.0)) {
          br = SeBracket{a, b, fa, fb};
          return true;
        }
        fa = fb;
      }
    }
    if (sign != 0) fa = double(sign);  // constant sign over [a, b]: the sign at b
    // next interval: climb while this is a right child, then step right
    while (idx & 1u) {
      idx >>= 1;
      --level;
    }
    if (level == 0) return false;
    ++idx;
  }
}
// Newton's method from the secant point of the walk's bracket, kept inside
// the bracket (which always holds the root) by bisection.
LMN_HD bool se_refine(const double* p, V3<double> ro, V3<double> rd, const SeBracket& br, double tmin, double tmax,
                      double& t) {
  const V3<double> o = {(ro.x - p[0]) / p[3], (ro.y - p[1]) / p[4], (ro.z - p[2]) / p[5]};
  const V3<double> d = {rd.x / p[3], rd.y / p[4], rd.z / p[5]};
  const double px = 2.0 / p[7], qa = p[7] / p[6], py = 2.0 / p[6];
  double lo = br.a, hi = br.b, flo = br.fa;
  double th = (br.a * br.fb - br.b * br.fa) / (br.fb - br.fa);
  if (!(th > lo && th < hi)) th = 0.5 * (lo + hi);
  for (int it = 0; it < 64; ++it) {
    double g;
    double fx = se_val_d(o, d, th, px, qa, py, g) - 1.0;
    if (fx == 0.0) break;
    if ((fx > 0.0) == (flo > 0.0)) {
      lo = th;
      flo = fx;
    } else {
      hi = th;
    }
    double tn = th - fx / g;
    if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
    bool done = m_abs(tn - th) <= 1e-13 * m_abs(tn) || hi - lo <= 1e-13 * m_abs(hi);
    th = tn;
    if (done) break;
  }
  if (!(th >= tmin && th <= tmax)) return false;
  t = th;
  return true;
}

LMN_COLD bool hit_superellipsoid_exact(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax,
                                       double& t) {
  SeBracket br;
  return se_exact_walk(p, ro, rd, tmin, tmax, br) && se_refine(p, ro, rd, br, tmin, tmax, t);
}
LMN_COLD bool hit_superellipsoid_exact_any(const double* p, V3<double> ro, V3<double> rd, double tmin,
                                           double tmax) {
  SeBracket br;
  return se_exact_walk(p, ro, rd, tmin, tmax, br);
}
LMN_COLD bool se_exact_bracket(const double* p, V3<double> ro, V3<double> rd, double tmin, double tmax,
                               SeBracket& br) {
  return se_exact_walk(p, ro, rd, tmin, tmax, br);
}

LMN_HD double spring_sdf(V3<double> q, const double* p) {
  // src/spring.rs:41-73
  double coil = p[3], tube = p[4], pitch = p[5], turns = p[6];
  double height = pitch * turns;
  const double two_pi = 2.0 * 3.14159265358979323846;
  double angle = m_atan2(q.z, q.x);
  double base_y = pitch * angle / two_pi;
  double best = 1.0 / 0.0;
  int k_min = int(floor((q.y - base_y) / pitch)) - 1;
  for (int k = k_min; k <= k_min + 3; ++k) {
    double hy = base_y + pitch * double(k);
    if (hy < -tube || hy > height + tube) continue;
    double hyc = fmin(fmax(hy, 0.0), height);
    double ha = (hyc / pitch) * two_pi;
    double hx = coil * m_cos(ha), hz = coil * m_sin(ha);
    double dx = q.x - hx, dy = q.y - hyc, dz = q.z - hz;
    double dist = m_sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < best) best = dist;
  }
  return best - tube;
}

// Spring::hit, reference semantics (src/spring.rs:35-105)
LMN_HD bool hit_spring_march(const double* p, V3<double> ro, V3<double> d, double tmin, double tmax, double& t) {
  V3<double> o = ro - V3<double>{p[0], p[1], p[2]};
  double step_min = p[4] * 0.3;
  double tc = tmin;
  for (int i = 0; i < 512; ++i) {
    if (tc > tmax) break;
    double dist = spring_sdf(o + d * tc, p);
    if (dist < 1e-4) {
      t = tc;
      return true;
    }
    tc += fmax(dist, step_min);
  }
  return false;
}

// Spring, exact mode: sign change of the helix distance along the ray,
// sampled at a tenth of the tube radius and refined by bisection (FP64).
LMN_HD bool hit_spring_exact(const double* p, V3<double> ro, V3<double> d, double tmin, double tmax, double& t) {
  V3<double> o = ro - V3<double>{p[0], p[1], p[2]};
  double r = p[3] + p[4], height = p[5] * p[6];
  double lo_b[3] = {-r, -p[4], -r}, hi_b[3] = {r, height + p[4], r};
  const double* oo = &o.x;
  const double* dd = &d.x;
  double t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double inv = 1.0 / dd[k];
    double a = (lo_b[k] - oo[k]) * inv, b = (hi_b[k] - oo[k]) * inv;
    if (a > b) { double tmp = a; a = b; b = tmp; }
    t0 = fmax(t0, a);
    t1 = fmin(t1, b);
  }
  if (!(t0 <= t1)) return false;
  double dl = length(d);
  double h = 0.1 * p[4] / dl;
  double ta = t0, fa = spring_sdf(o + d * ta, p);
  if (fa <= 0.0) { t = t0; return true; }
  for (int i = 0; i < 1 << 16 && ta < t1; ++i) {
    double tb = fmin(ta + fmax(h, fa / dl), t1);  // the SDF is a distance bound
    double fb = spring_sdf(o + d * tb, p);
    if (fb <= 0.0) {
      double lo = ta, hi = tb;
      for (int k = 0; k < 80 && (hi - lo) * dl > 1e-13 * (1.0 + m_abs(hi) * dl); ++k) {
        double mid = 0.5 * (lo + hi);
        if (spring_sdf(o + d * mid, p) > 0.0) lo = mid; else hi = mid;
      }
      t = 0.5 * (lo + hi);
      return true;
    }
    if (tb >= t1) break;
    ta = tb;
    fa = fb;
  }
  return false;
}

// uv of the marched primitives from the object-space point
// (src/torus.rs:86-92, src/superellipsoid.rs:80-82, src/spring.rs:95-97)
LMN_HD void marched_uv(uint32_t kind, const double* p, V3<double> wp, double& u, double& v) {
  const double pi = 3.14159265358979323846;
  if (kind == K_TORUS) {
    V3<double> l = wp - V3<double>{p[0], p[1], p[2]};
    double theta = m_atan2(l.z, l.x);
    double xz = m_sqrt(l.x * l.x + l.z * l.z);
    double phi = m_atan2(l.y, xz - p[3]);
    u = theta / (2.0 * pi) + 0.5;
    v = phi / (2.0 * pi) + 0.5;
  } else if (kind == K_SUPERELLIPSOID) {
    V3<double> l = {(wp.x - p[0]) / p[3], (wp.y - p[1]) / p[4], (wp.z - p[2]) / p[5]};
    u = 0.5 + m_atan2(l.z, l.x) / (2.0 * pi);
    v = 0.5 + clamp_r(asin(l.y), -1.0, 1.0) / pi;
  } else {
    V3<double> l = wp - V3<double>{p[0], p[1], p[2]};
    u = l.y / (p[5] * p[6]);
    v = m_atan2(l.z, l.x) / (2.0 * pi) + 0.5;
  }
}

// Outward normal of the marched primitives by central differences
// (src/torus.rs:79-84, src/superellipsoid.rs:69-78, src/spring.rs:88-93)
LMN_HD V3<double> marched_normal(uint32_t kind, const double* p, V3<double> wp) {
  if (kind == K_TORUS) {
    V3<double> l = wp - V3<double>{p[0], p[1], p[2]};
    const double e = 1e-5;
    double nx = torus_sdf(l + V3<double>{e, 0, 0}, p[3], p[4]) - torus_sdf(l - V3<double>{e, 0, 0}, p[3], p[4]);
    double ny = torus_sdf(l + V3<double>{0, e, 0}, p[3], p[4]) - torus_sdf(l - V3<double>{0, e, 0}, p[3], p[4]);
    double nz = torus_sdf(l + V3<double>{0, 0, e}, p[3], p[4]) - torus_sdf(l - V3<double>{0, 0, e}, p[3], p[4]);
    return unit(V3<double>{nx, ny, nz});
  } else if (kind == K_SUPERELLIPSOID) {
    V3<double> l = {(wp.x - p[0]) / p[3], (wp.y - p[1]) / p[4], (wp.z - p[2]) / p[5]};
    const double e = 1e-5;
    double nx = superellipsoid_f(l + V3<double>{e, 0, 0}, p[6], p[7]) - superellipsoid_f(l - V3<double>{e, 0, 0}, p[6], p[7]);
    double ny = superellipsoid_f(l + V3<double>{0, e, 0}, p[6], p[7]) - superellipsoid_f(l - V3<double>{0, e, 0}, p[6], p[7]);
    double nz = superellipsoid_f(l + V3<double>{0, 0, e}, p[6], p[7]) - superellipsoid_f(l - V3<double>{0, 0, e}, p[6], p[7]);
    return unit(V3<double>{nx / p[3], ny / p[4], nz / p[5]});
  } else {
    V3<double> l = wp - V3<double>{p[0], p[1], p[2]};
    const double e = 1e-4;
    double nx = spring_sdf(l + V3<double>{e, 0, 0}, p) - spring_sdf(l - V3<double>{e, 0, 0}, p);
    double ny = spring_sdf(l + V3<double>{0, e, 0}, p) - spring_sdf(l - V3<double>{0, e, 0}, p);
    double nz = spring_sdf(l + V3<double>{0, 0, e}, p) - spring_sdf(l - V3<double>{0, 0, e}, p);
    return unit(V3<double>{nx, ny, nz});
  }
}

// Object-space intersection of one primitive in precision R.  Marched kinds
// are always FP64 and go through hit_marched_fp64.
// ---- disk, cylinder, cone (FP64, out of line) ---------------------------------
// Disk::hit (src/disk.rs:31-59): plane test with a parallel cull, then the
// radius.  A NaN distance is a miss (the reference would accept it).
LMN_HD bool hit_disk(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const V3<double> c = {p[0], p[1], p[2]}, n = {p[3], p[4], p[5]};
  double denom = dot(n, d);
  if (m_abs(denom) < 1e-8) return false;
  double tt = dot(c - o, n) / denom;
  if (!(tt >= tmin && tt <= tmax)) return false;
  V3<double> off = (o + d * tt) - c;
  if (length_squared(off) > p[6] * p[6]) return false;
  t = tt;
  return true;
}
// Cylinder::hit / Cone::hit (src/cylinder.rs:23-73, src/cone.rs:23-78):
// quadratic in the XZ offset (the cone's radius shrinking to its apex at
// y_max), near root first, then the far one, each within [tmin, tmax] and
// [y_min, y_max].  A NaN root (a ray along the axis) is a miss.
LMN_HD bool hit_cylinder_cone(bool cone, const double* p, V3<double> o, V3<double> d, double tmin, double tmax,
                              double& t) {
  const double height = p[5] - p[4];
  double k = 0.0, oy = 0.0;
  if (cone) {
    if (m_abs(height) < 1e-10) return false;
    k = (p[3] / height) * (p[3] / height);
    oy = o.y - p[5];
  }
  const double ox = o.x - p[0], oz = o.z - p[2];
  double a, half_b, c;
  if (cone) {
    a = d.x * d.x + d.z * d.z - k * d.y * d.y;
    half_b = ox * d.x + oz * d.z - k * oy * d.y;
    c = ox * ox + oz * oz - k * oy * oy;
  } else {
    a = d.x * d.x + d.z * d.z;
    half_b = ox * d.x + oz * d.z;
    c = ox * ox + oz * oz - p[3] * p[3];
  }
  double disc = half_b * half_b - a * c;
  if (disc < 0.0) return false;
  double sq = m_sqrt(disc);
  for (int s = -1; s <= 1; s += 2) {
    double root = (-half_b + double(s) * sq) / a;
    if (!(root >= tmin && root <= tmax)) continue;
    double y = o.y + root * d.y;
    if (y < p[4] || y > p[5]) continue;
    t = root;
    return true;
  }
  return false;
}
// Ellipsoid::hit (src/ellipsoid.rs:21-67): the unit-sphere quadratic of the
// ray scaled by 1/radii, near root first, then the far one.
LMN_HD bool hit_ellipsoid(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const V3<double> inv = {1.0 / p[3], 1.0 / p[4], 1.0 / p[5]};
  const V3<double> oc = {(o.x - p[0]) * inv.x, (o.y - p[1]) * inv.y, (o.z - p[2]) * inv.z};
  const V3<double> dd = {d.x * inv.x, d.y * inv.y, d.z * inv.z};
  double a = length_squared(dd), half_b = dot(oc, dd), c = length_squared(oc) - 1.0;
  double disc = half_b * half_b - a * c;
  if (disc < 0.0) return false;
  double sq = m_sqrt(disc);
  double root = (-half_b - sq) / a;
  if (!(root >= tmin && root <= tmax)) {
    root = (-half_b + sq) / a;
    if (!(root >= tmin && root <= tmax)) return false;
  }
  t = root;
  return true;
}
// Hemisphere::hit (src/hemisphere.rs:21-54): the sphere's roots in order,
// each kept only in range and on the upper half (y >= c.y).
LMN_HD bool hit_hemisphere(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const V3<double> oc = o - V3<double>{p[0], p[1], p[2]};
  double a = length_squared(d), half_b = dot(oc, d), c = length_squared(oc) - p[3] * p[3];
  double disc = half_b * half_b - a * c;
  if (disc < 0.0) return false;
  double sq = m_sqrt(disc);
  for (int s = -1; s <= 1; s += 2) {
    double root = (-half_b + double(s) * sq) / a;
    if (!(root >= tmin && root <= tmax)) continue;
    if (o.y + root * d.y < p[1]) continue;
    t = root;
    return true;
  }
  return false;
}
// Annulus::hit (src/annulus.rs:39-76): plane test with a parallel cull,
// then inner_r^2 <= |offset|^2 <= outer_r^2.
LMN_HD bool hit_annulus(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const V3<double> c = {p[0], p[1], p[2]}, n = {p[3], p[4], p[5]};
  double denom = dot(n, d);
  if (m_abs(denom) < 1e-8) return false;
  double tt = dot(c - o, n) / denom;
  if (!(tt >= tmin && tt <= tmax)) return false;
  double dist_sq = length_squared((o + d * tt) - c);
  if (dist_sq < p[6] * p[6] || dist_sq > p[7] * p[7]) return false;
  t = tt;
  return true;
}
// One triangle of a prism or pyramid: the solids' own Moller-Trumbore
// (src/prism.rs:100-126, src/pyramid.rs:29-50: cull |a| < 1e-10, t > 1e-10)
// under their strict range t_min < t < closest (src/prism.rs:50-78,
// src/pyramid.rs:86-94), so that of two facets at one t the first kept wins.
// Returns the barycentric (u, v) too.
LMN_HD bool facet_mt(const double* p, V3<double> o, V3<double> d, double& t, double& u, double& v) {
  const V3<double> v0 = {p[0], p[1], p[2]}, e1 = {p[3], p[4], p[5]}, e2 = {p[6], p[7], p[8]};
  V3<double> h = cross(d, e2);
  double a = dot(e1, h);
  if (m_abs(a) < 1e-10) return false;
  double f = 1.0 / a;
  V3<double> s = o - v0;
  u = f * dot(s, h);
  if (!(u >= 0.0 && u <= 1.0)) return false;
  V3<double> q = cross(s, e1);
  v = f * dot(d, q);
  if (v < 0.0 || u + v > 1.0) return false;
  t = f * dot(e2, q);
  return t > 1e-10;
}
LMN_HD bool hit_facet(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  double tt, u, v;
  if (!facet_mt(p, o, d, tt, u, v) || !(tt > tmin && tt < tmax)) return false;
  t = tt;
  return true;
}
// One face of a wedge (src/wedge.rs:22-120): the box [min, max] whose top is
// a ramp rising along +z from min.y at min.z to max.y at max.z.  Faces in the
// reference's order: 0 bottom, 1 back (z = max.z), 2 left, 3 right (the
// triangles under the ramp), 4 the ramp; each under the strict range
// t_min < t < closest.
LMN_HD double wedge_ramp_y(const double* p, double z) { return (z - p[2]) / (p[5] - p[2]) * (p[4] - p[1]) + p[1]; }
LMN_HD V3<double> wedge_ramp_normal(const double* p) {
  return unit(V3<double>{0.0, 1.0, -((p[4] - p[1]) / (p[5] - p[2]))});
}
LMN_HD bool hit_wedge_face(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const int face = int(p[6]);
  const V3<double> lo = {p[0], p[1], p[2]}, hi = {p[3], p[4], p[5]};
  double tt;
  if (face == 4) {
    const V3<double> n = wedge_ramp_normal(p);
    double denom = dot(d, n);
    if (!(m_abs(denom) > 1e-10)) return false;
    tt = dot(lo - o, n) / denom;
  } else {
    const int axis = face == 0 ? 1 : face == 1 ? 2 : 0;
    const double dc = axis == 0 ? d.x : axis == 1 ? d.y : d.z;
    const double oc = axis == 0 ? o.x : axis == 1 ? o.y : o.z;
    const double plane = face == 0 ? lo.y : face == 1 ? hi.z : face == 2 ? lo.x : hi.x;
    if (!(m_abs(dc) > 1e-10)) return false;
    tt = (plane - oc) / dc;
  }
  if (!(tt > tmin && tt < tmax)) return false;
  const V3<double> q = o + d * tt;
  bool in;
  switch (face) {
    case 0: in = q.x >= lo.x && q.x <= hi.x && q.z >= lo.z && q.z <= hi.z; break;
    case 1: in = q.x >= lo.x && q.x <= hi.x && q.y >= lo.y && q.y <= hi.y; break;
    case 2:
    case 3: in = q.z >= lo.z && q.z <= hi.z && q.y >= lo.y && q.y <= wedge_ramp_y(p, q.z); break;
    default:
      in = q.x >= lo.x && q.x <= hi.x && q.z >= lo.z && q.z <= hi.z && m_abs(q.y - wedge_ramp_y(p, q.z)) < 0.01;
  }
  if (!in) return false;
  t = tt;
  return true;
}
// ---- sphere-traced kinds: rounded box, Mobius strip -------------------------
// RoundedBox's signed distance (src/rounded_box.rs:30-35), at p relative to
// the centre.
LMN_HD double sdf_rounded_box(const double* p, V3<double> q) {
  const double qx = m_abs(q.x) - p[3], qy = m_abs(q.y) - p[4], qz = m_abs(q.z) - p[5];
  const double ox = fmax(qx, 0.0), oy = fmax(qy, 0.0), oz = fmax(qz, 0.0);
  const double outside = m_sqrt(ox * ox + oy * oy + oz * oz);
  const double inside = fmin(fmax(qx, fmax(qy, qz)), 0.0);
  return outside + inside - p[6];
}
// Mobius's distance estimate (src/mobius.rs:37-80): the nearest of the
// strip's segments at the point's angle and the angles pi either side.
LMN_HD double sdf_mobius(const double* p, V3<double> q) {
  const double r = p[3], w = p[4], th = p[5], pi = 3.14159265358979323846;
  const double angle = m_atan2(q.z, q.x);
  double best = __builtin_inf();
  for (int k = -1; k <= 1; ++k) {
    const double u = angle + double(k) * pi;
    const double cos_u = m_cos(u), sin_u = m_sin(u), cos_half = m_cos(u / 2.0), sin_half = m_sin(u / 2.0);
    const double cx = r * cos_u, cz = r * sin_u;
    const V3<double> dir = unit(V3<double>{cos_u * cos_half, sin_half, sin_u * cos_half});
    const double vp = (q.x - cx) * dir.x + q.y * dir.y + (q.z - cz) * dir.z;
    const double vc = fmin(fmax(vp, -w), w);
    const double dx = q.x - (cx + vc * dir.x), dy = q.y - vc * dir.y, dz = q.z - (cz + vc * dir.z);
    const double dist = m_sqrt(dx * dx + dy * dy + dz * dz) - th;
    if (dist < best) best = dist;
  }
  return best;
}
LMN_HD double sdf_traced(uint32_t kind, const double* p, V3<double> q) {
  return kind == K_ROUNDED_BOX ? sdf_rounded_box(p, q) : sdf_mobius(p, q);
}
// The reference's sphere trace (src/rounded_box.rs:38-76, src/mobius.rs:84-108):
// from t_min, accept where the distance is below tol, else step max(dist,
// floor) in t (the direction unnormalised), up to n steps or past t_max.
// The Mobius strip is traced this way in both intersection modes.
LMN_HD bool hit_traced_reference(uint32_t kind, const double* p, V3<double> o, V3<double> d, double tmin,
                                 double tmax, double& t) {
  const V3<double> oc = o - V3<double>{p[0], p[1], p[2]};
  const bool box = kind == K_ROUNDED_BOX;
  const int steps = box ? 256 : 512;
  const double tol = box ? 1e-5 : 1e-4, floor = box ? 1e-4 : p[5] * 0.3;
  double tt = tmin;
  for (int i = 0; i < steps; ++i) {
    if (tt > tmax) break;
    double dist = sdf_traced(kind, p, oc + d * tt);
    if (dist < tol) {
      t = tt;
      return true;
    }
    tt += fmax(dist, floor);
  }
  return false;
}
// Gradient of the rounded box's distance at q (unit; the outward normal).
LMN_HD V3<double> grad_rounded_box(const double* p, V3<double> q) {
  const double qx = m_abs(q.x) - p[3], qy = m_abs(q.y) - p[4], qz = m_abs(q.z) - p[5];
  const double sx = q.x < 0.0 ? -1.0 : 1.0, sy = q.y < 0.0 ? -1.0 : 1.0, sz = q.z < 0.0 ? -1.0 : 1.0;
  if (qx > 0.0 || qy > 0.0 || qz > 0.0) {
    const V3<double> o = {fmax(qx, 0.0) * sx, fmax(qy, 0.0) * sy, fmax(qz, 0.0) * sz};
    return o / length(o);
  }
  if (qx >= qy && qx >= qz) return {sx, 0.0, 0.0};
  if (qy >= qz) return {0.0, sy, 0.0};
  return {0.0, 0.0, sz};
}
// Exact mode for the rounded box, whose distance is exact: trace in units of
// length (t step = |dist| / |d|, never across the surface), outward from
// t_min or, for an origin inside, to the exit, until the distance is below
// 1e-7 of the box's size; then Newton steps on dist(t) with the analytic
// gradient (the distance is C1 across the face, edge and corner pieces) to
// full precision.  A trace that runs out of steps (a ray grazing the
// surface) is a miss.
LMN_HD bool hit_rounded_box_exact(const double* p, V3<double> o, V3<double> d, double tmin, double tmax, double& t) {
  const V3<double> oc = o - V3<double>{p[0], p[1], p[2]};
  const double len = length(d), inv_len = 1.0 / len;
  const double tol = 1e-7 * (p[3] + p[4] + p[5] + p[6]);
  double tt = tmin;
  const double sgn = sdf_rounded_box(p, oc + d * tt) < 0.0 ? -1.0 : 1.0;
  int i = 0;
  for (; i < 2048; ++i) {
    if (!(tt <= tmax)) return false;
    double dist = sgn * sdf_rounded_box(p, oc + d * tt);
    if (dist < tol) break;
    tt += dist * inv_len;
  }
  if (i == 2048) return false;
  for (int k = 0; k < 4; ++k) {
    const V3<double> q = oc + d * tt;
    const double g = sdf_rounded_box(p, q), dg = dot(grad_rounded_box(p, q), d);
    if (!(m_abs(dg) > 1e-6 * len)) break;
    const double step = g / dg;
    if (!(m_abs(step) <= 2.0 * tol * inv_len + 1e-12 * m_abs(tt))) break;  // stay at this root
    tt -= step;
  }
  if (!(tt >= tmin && tt <= tmax)) return false;
  t = tt;
  return true;
}
// Central-difference normal (step 1e-5 for the box, 1e-4 for the strip) and
// the reference's (u, v) at q relative to the centre (src/rounded_box.rs:47-69,
// src/mobius.rs:92-103).
LMN_HD void traced_surface(uint32_t kind, const double* p, V3<double> q, V3<double>& n, double& u, double& v) {
  const double eps = kind == K_ROUNDED_BOX ? 1e-5 : 1e-4;
  const double nx = sdf_traced(kind, p, q + V3<double>{eps, 0.0, 0.0}) - sdf_traced(kind, p, q - V3<double>{eps, 0.0, 0.0});
  const double ny = sdf_traced(kind, p, q + V3<double>{0.0, eps, 0.0}) - sdf_traced(kind, p, q - V3<double>{0.0, eps, 0.0});
  const double nz = sdf_traced(kind, p, q + V3<double>{0.0, 0.0, eps}) - sdf_traced(kind, p, q - V3<double>{0.0, 0.0, eps});
  n = unit(V3<double>{nx, ny, nz});
  if (kind == K_MOBIUS) {
    u = m_atan2(q.z, q.x) / (2.0 * 3.14159265358979323846) + 0.5;
    v = q.y / (p[4] + p[5]) * 0.5 + 0.5;
    return;
  }
  const V3<double> total = {p[3] + p[6], p[4] + p[6], p[5] + p[6]};
  const double ax = m_abs(n.x), ay = m_abs(n.y), az = m_abs(n.z);
  if (ax > ay && ax > az) {
    u = (q.z / total.z + 1.0) * 0.5;
    v = (q.y / total.y + 1.0) * 0.5;
  } else if (ay > az) {
    u = (q.x / total.x + 1.0) * 0.5;
    v = (q.z / total.z + 1.0) * 0.5;
  } else {
    u = (q.x / total.x + 1.0) * 0.5;
    v = (q.y / total.y + 1.0) * 0.5;
  }
}

LMN_COLD bool hit_quadric_fp64(uint32_t kind, const double* p, V3<double> o, V3<double> d, double tmin, double tmax,
                               bool exact, double& t) {
  switch (kind) {
    case K_ROUNDED_BOX:
      return exact ? hit_rounded_box_exact(p, o, d, tmin, tmax, t) : hit_traced_reference(kind, p, o, d, tmin, tmax, t);
    case K_MOBIUS: return hit_traced_reference(kind, p, o, d, tmin, tmax, t);
    case K_DISK: return hit_disk(p, o, d, tmin, tmax, t);
    case K_ELLIPSOID: return hit_ellipsoid(p, o, d, tmin, tmax, t);
    case K_HEMISPHERE: return hit_hemisphere(p, o, d, tmin, tmax, t);
    case K_ANNULUS: return hit_annulus(p, o, d, tmin, tmax, t);
    case K_FACET: return hit_facet(p, o, d, tmin, tmax, t);
    case K_WEDGE_FACE: return hit_wedge_face(p, o, d, tmin, tmax, t);
    default: return hit_cylinder_cone(kind == K_CONE, p, o, d, tmin, tmax, t);
  }
}
// Outward normal and (u, v) at object-space point q = o + d t of the hit.
LMN_COLD void quadric_surface(uint32_t kind, const double* p, V3<double> o, V3<double> d, V3<double> q,
                              V3<double>& n, double& u, double& v) {
  const double pi = 3.14159265358979323846;
  const V3<double> c = {p[0], p[1], p[2]};
  if (kind == K_ROUNDED_BOX || kind == K_MOBIUS) {
    traced_surface(kind, p, q - c, n, u, v);
    return;
  }
  if (kind == K_FACET) {  // the stored normal; (u, v) from the same ray's Moller-Trumbore
    double tt;
    n = {p[9], p[10], p[11]};
    facet_mt(p, o, d, tt, u, v);
    return;
  }
  if (kind == K_WEDGE_FACE) {  // src/wedge.rs:37-116: per-face normal and (u, v)
    const double width = p[3] - p[0], height = p[4] - p[1], depth = p[5] - p[2];
    switch (int(p[6])) {
      case 0: n = {0.0, -1.0, 0.0}; u = (q.x - p[0]) / width; v = (q.z - p[2]) / depth; break;
      case 1: n = {0.0, 0.0, 1.0}; u = (q.x - p[0]) / width; v = (q.y - p[1]) / height; break;
      case 2:
      case 3:
        n = {int(p[6]) == 2 ? -1.0 : 1.0, 0.0, 0.0};
        u = (q.z - p[2]) / depth;
        v = (q.y - p[1]) / height;
        break;
      default: n = wedge_ramp_normal(p); u = (q.x - p[0]) / width; v = (q.z - p[2]) / depth;
    }
    return;
  }
  if (kind == K_DISK) {  // src/disk.rs:48-56
    V3<double> off = q - c;
    n = {p[3], p[4], p[5]};
    u = m_atan2(off.z, off.x) / (2.0 * pi) + 0.5;
    v = m_sqrt(length_squared(off)) / p[6];
    return;
  }
  if (kind == K_ANNULUS) {  // src/annulus.rs:61-74: u angular, v radial
    V3<double> off = q - c;
    n = {p[3], p[4], p[5]};
    u = m_atan2(off.z, off.x) / (2.0 * pi) + 0.5;
    v = (m_sqrt(length_squared(off)) - p[6]) / (p[7] - p[6]);
    return;
  }
  if (kind == K_ELLIPSOID || kind == K_HEMISPHERE) {  // src/ellipsoid.rs:56-64, src/hemisphere.rs:41-46
    if (kind == K_ELLIPSOID)
      n = unit(V3<double>{(q.x - c.x) / (p[3] * p[3]), (q.y - c.y) / (p[4] * p[4]), (q.z - c.z) / (p[5] * p[5])});
    else
      n = (q - c) / p[3];
    u = (m_atan2(-n.z, n.x) + pi) / (2.0 * pi);
    v = m_acos(-n.y) / pi;
    return;
  }
  const double height = p[5] - p[4];
  if (kind == K_CYLINDER) {  // src/cylinder.rs:51-64
    n = {(q.x - c.x) / p[3], 0.0, (q.z - c.z) / p[3]};
    v = m_abs(height) < 1e-10 ? 0.5 : (q.y - p[4]) / height;
  } else {  // src/cone.rs:60-72
    double r_at_y = p[3] * (p[5] - q.y) / height;
    double rr = fmax_r(r_at_y, 1e-10);
    n = unit(V3<double>{(q.x - c.x) / rr, p[3] / height, (q.z - c.z) / rr});
    v = (q.y - p[4]) / height;
  }
  u = (m_atan2(-(q.z - c.z), q.x - c.x) + pi) / (2.0 * pi);
}

template <class R>
LMN_HD bool hit_prim_analytic(uint32_t kind, const R* p, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
  switch (kind) {
    case K_SPHERE: return hit_sphere(p, o, d, tmin, tmax, t);
    case K_QUAD: return hit_quad(p, o, d, tmin, tmax, t);
    case K_RECT_XY: return hit_rect(p, 2, 0, 1, o, d, tmin, tmax, t);
    case K_RECT_XZ: return hit_rect(p, 1, 0, 2, o, d, tmin, tmax, t);
    case K_RECT_YZ: return hit_rect(p, 0, 1, 2, o, d, tmin, tmax, t);
    case K_TRIANGLE: return hit_triangle(p, o, d, tmin, tmax, t);
    case K_PLANE: return hit_plane(p, o, d, tmin, tmax, t);
    default: return false;
  }
}

// Marched kinds (torus, superellipsoid, spring), FP64, out of line.
LMN_COLD bool hit_marched_fp64(uint32_t kind, const double* p, V3<double> o, V3<double> d, double tmin, double tmax,
                               bool exact, double& t) {
  switch (kind) {
    case K_TORUS: return exact ? hit_torus_exact(p, o, d, tmin, tmax, t) : hit_torus_march(p, o, d, tmin, tmax, t);
    case K_SUPERELLIPSOID:
      return exact ? hit_superellipsoid_exact(p, o, d, tmin, tmax, t) : hit_superellipsoid_march(p, o, d, tmin, tmax, t);
    default: return exact ? hit_spring_exact(p, o, d, tmin, tmax, t) : hit_spring_march(p, o, d, tmin, tmax, t);
  }
}

// FP64 evaluation of one primitive: spheres inline (the FP32 path's large
// spheres), the marched kinds out of line.
LMN_HD bool hit_prim_fp64(uint32_t kind, const double* p, V3<double> o, V3<double> d, double tmin, double tmax,
                          bool exact, double& t) {
  if (kind == K_SPHERE) return hit_sphere<double>(p, o, d, tmin, tmax, t);
  if (is_marched_kind(kind)) return hit_marched_fp64(kind, p, o, d, tmin, tmax, exact, t);
  if (is_quadric_kind(kind)) return hit_quadric_fp64(kind, p, o, d, tmin, tmax, exact, t);
  return hit_prim_analytic<double>(kind, p, o, d, tmin, tmax, t);
}

// Exact torus normal: gradient of (|p|^2 + R^2 - r^2)^2 - 4R^2(x^2+z^2).
LMN_HD V3<double> torus_exact_normal(const double* p, V3<double> wp) {
  V3<double> l = wp - V3<double>{p[0], p[1], p[2]};
  double k = length_squared(l) + p[3] * p[3] - p[4] * p[4];
  double R2 = p[3] * p[3];
  return unit(V3<double>{4.0 * k * l.x - 8.0 * R2 * l.x, 4.0 * k * l.y, 4.0 * k * l.z - 8.0 * R2 * l.z});
}

// Outward normal (object space) at object-space point q.
template <class R>
LMN_HD V3<R> prim_normal_analytic(uint32_t kind, const R* p, V3<R> q) {
  switch (kind) {
    case K_SPHERE: return (q - V3<R>{p[0], p[1], p[2]}) / p[3];
    case K_QUAD: return unit(V3<R>{p[9], p[10], p[11]});
    case K_RECT_XY: return {R(0), R(0), R(1)};
    case K_RECT_XZ: return {R(0), R(1), R(0)};
    case K_RECT_YZ: return {R(1), R(0), R(0)};
    case K_TRIANGLE: return unit(cross(V3<R>{p[3], p[4], p[5]}, V3<R>{p[6], p[7], p[8]}));
    case K_PLANE: return {p[3], p[4], p[5]};
    default: return {R(0), R(1), R(0)};
  }
}

// Superellipsoid normal in exact mode: the reference's central differences
// of f (src/superellipsoid.rs:69-78) with f evaluated through pow_pos, a few
// times cheaper than the library power and equal to it within a few ulp.
LMN_HD double superellipsoid_f_pos(V3<double> q, double e1, double e2) {
  const double px = 2.0 / e2;
  double xz = pow_pos(m_abs(q.x), px) + pow_pos(m_abs(q.z), px);
  double val = pow_pos(xz, e2 / e1) + pow_pos(m_abs(q.y), 2.0 / e1);
  return pow_pos(val, e1 / 2.0) - 1.0;
}
LMN_HD V3<double> superellipsoid_exact_normal(const double* p, V3<double> wp) {
  V3<double> l = {(wp.x - p[0]) / p[3], (wp.y - p[1]) / p[4], (wp.z - p[2]) / p[5]};
  const double e = 1e-5;
  double nx = superellipsoid_f_pos(l + V3<double>{e, 0, 0}, p[6], p[7]) -
              superellipsoid_f_pos(l - V3<double>{e, 0, 0}, p[6], p[7]);
  double ny = superellipsoid_f_pos(l + V3<double>{0, e, 0}, p[6], p[7]) -
              superellipsoid_f_pos(l - V3<double>{0, e, 0}, p[6], p[7]);
  double nz = superellipsoid_f_pos(l + V3<double>{0, 0, e}, p[6], p[7]) -
              superellipsoid_f_pos(l - V3<double>{0, 0, e}, p[6], p[7]);
  return unit(V3<double>{nx / p[3], ny / p[4], nz / p[5]});
}

LMN_COLD V3<double> marched_normal_fp64(uint32_t kind, const double* p, V3<double> q, bool exact) {
  if (exact && kind == K_TORUS) return torus_exact_normal(p, q);
  if (exact && kind == K_SUPERELLIPSOID) return superellipsoid_exact_normal(p, q);
  return marched_normal(kind, p, q);
}

}  // namespace lmn

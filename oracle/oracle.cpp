// CPU FP64 restatement of the reference's render hot path (see oracle.h).
// TEST INFRASTRUCTURE ONLY -- never linked into the product library.
//
// Every function follows the cited reference source line by line, in f64,
// with the reference's object structure (virtual Hittable tree, wrappers,
// BvhNode) so that traversal order, tie breaking and RNG draw order match.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------- vec3.rs
struct Vec3 {
  double x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
  double length_squared() const { return x * x + y * y + z * z; }        // :30-32
  double length() const { return std::sqrt(length_squared()); }          // :35-37
  double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }  // :40-42
  Vec3 cross(const Vec3& o) const {                                      // :44-50
    return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
  }
  Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  Vec3 operator-() const { return {-x, -y, -z}; }
  Vec3 operator*(double t) const { return {x * t, y * t, z * t}; }
  Vec3 operator/(double t) const { return {x / t, y / t, z / t}; }
  Vec3& operator+=(const Vec3& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  Vec3 unit() const { return *this / length(); }  // :53-57
  bool near_zero() const {                       // :59-62
    const double e = 1e-8;
    return std::fabs(x) < e && std::fabs(y) < e && std::fabs(z) < e;
  }
  Vec3 reflect(const Vec3& n) const { return *this - n * 2.0 * dot(n); }              // :64-66
  Vec3 reflect_around(const Vec3& h) const { return h * 2.0 * dot(h) - *this; }      // :69-71
  Vec3 refract(const Vec3& n, double eta) const {                                    // :73-78
    double ct = std::min((-*this).dot(n), 1.0);
    Vec3 perp = (*this + n * ct) * eta;
    Vec3 par = n * -std::sqrt(std::fabs(1.0 - perp.length_squared()));
    return perp + par;
  }
  Vec3 hadamard(const Vec3& o) const { return {x * o.x, y * o.y, z * o.z}; }  // :144-146
  Vec3 vmin(const Vec3& o) const { return {std::fmin(x, o.x), std::fmin(y, o.y), std::fmin(z, o.z)}; }
  Vec3 vmax(const Vec3& o) const { return {std::fmax(x, o.x), std::fmax(y, o.y), std::fmax(z, o.z)}; }
  double axis(int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
Vec3 v3(const double* a) { return {a[0], a[1], a[2]}; }

// ---------------------------------------------------------------- RNG
struct Rng {
  virtual double next_f64() = 0;  // RngCore::next_f64 (src/material.rs:31-39)
  virtual ~Rng() = default;
};
uint64_t splitmix_out(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// The HIP kernels' counter-based stream: draw n = mix(key + n * gamma).
struct CounterRng : Rng {
  uint64_t key;
  uint32_t n = 0;
  explicit CounterRng(uint64_t k, uint32_t c = 0) : key(k), n(c) {}
  double next_f64() override {
    uint64_t x = splitmix_out(key + uint64_t(++n) * 0x9E3779B97F4A7C15ULL);
    return double(x >> 11) * 0x1.0p-53;
  }
};
uint64_t frame_key(uint64_t seed, uint32_t scene_key) {
  return splitmix_out(seed * 0xD1B54A32D192ED03ULL + uint64_t(scene_key) * 0xA0761D6478BD642FULL +
                      0x632BE59BD9B4E019ULL);
}
uint64_t path_key(uint64_t fkey, uint64_t pixel, uint32_t sample) {
  return splitmix_out(fkey ^ splitmix_out((pixel << 24) ^ uint64_t(sample) ^ 0x8CB92BA72F3D8DD7ULL));
}
// rand 0.9.2 SmallRng: xoshiro256++ seeded by SplitMix64 (seed_from_u64);
// f64 = (next_u64 >> 11) * 2^-53.  Restated from the published algorithm.
struct XoshiroRng : Rng {
  uint64_t s[4];
  explicit XoshiroRng(uint64_t state) {
    for (int i = 0; i < 4; ++i) {
      state += 0x9e3779b97f4a7c15ULL;
      s[i] = splitmix_out(state);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next_u64() {
    uint64_t r = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double next_f64() override { return double(next_u64() >> 11) * 0x1.0p-53; }
};

// vec3.rs:86-122
Vec3 random_in_unit_sphere(Rng& r) {
  for (;;) {
    double a = r.next_f64(), b = r.next_f64(), c = r.next_f64();
    Vec3 v(-1.0 + a * 2.0, -1.0 + b * 2.0, -1.0 + c * 2.0);
    if (v.length_squared() < 1.0) return v;
  }
}
Vec3 random_unit_vector(Rng& r) { return random_in_unit_sphere(r).unit(); }
Vec3 random_in_unit_disk(Rng& r) {
  for (;;) {
    double a = r.next_f64(), b = r.next_f64();
    Vec3 v(a * 2.0 - 1.0, b * 2.0 - 1.0, 0.0);
    if (v.length_squared() < 1.0) return v;
  }
}

// ---------------------------------------------------------------- ray / aabb
struct Ray {
  Vec3 o, d;
  double time = 0;
  Vec3 at(double t) const { return o + d * t; }  // ray.rs:21-23
};

struct Aabb {
  Vec3 mn, mx;
  // aabb.rs:18-56
  bool hit(const Ray& r, double tmin, double tmax) const {
    for (int a = 0; a < 3; ++a) {
      double inv = 1.0 / r.d.axis(a);
      double t0 = (mn.axis(a) - r.o.axis(a)) * inv;
      double t1 = (mx.axis(a) - r.o.axis(a)) * inv;
      if (inv < 0.0) std::swap(t0, t1);
      tmin = std::fmax(t0, tmin);
      tmax = std::fmin(t1, tmax);
      if (a < 2 && tmax <= tmin) return false;
    }
    return tmax > tmin;
  }
  double surface_area() const {  // :59-62
    Vec3 d = mx - mn;
    return 2.0 * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
  static Aabb surrounding(const Aabb& a, const Aabb& b) { return {a.mn.vmin(b.mn), a.mx.vmax(b.mx)}; }
};

// ---------------------------------------------------------------- materials
struct Material;
struct HitRecord {  // hit.rs:7-15
  Vec3 p, n;
  double t = 0, u = 0, v = 0;
  bool front = false;
  const Material* mat = nullptr;
  int64_t prim = -1;
};
void set_face(HitRecord& h, const Ray& r, const Vec3& outward) {  // hit.rs:19-43
  h.front = r.d.dot(outward) < 0.0;
  h.n = h.front ? outward : -outward;
}

struct Texture {
  Vec3 color;
  const float* texels = nullptr;
  uint32_t w = 0, h = 0;
  int32_t id = -1;  // descriptor texture index
  int64_t texel(double u, double v) const {  // the texel a lookup reads (x + y * w)
    u = std::min(std::max(u, 0.0), 1.0);
    v = 1.0 - std::min(std::max(v, 0.0), 1.0);
    uint32_t i = std::min(uint32_t(u * double(w)), w - 1);
    uint32_t j = std::min(uint32_t(v * double(h)), h - 1);
    return int64_t(j) * w + i;
  }
  Vec3 value(double u, double v) const {  // texture.rs:19-23, :463-479
    if (!texels) return color;
    u = std::min(std::max(u, 0.0), 1.0);
    v = 1.0 - std::min(std::max(v, 0.0), 1.0);
    uint32_t i = std::min(uint32_t(u * double(w)), w - 1);
    uint32_t j = std::min(uint32_t(v * double(h)), h - 1);
    const float* px = texels + (size_t(j) * w + i) * 3;
    return {double(px[0]), double(px[1]), double(px[2])};
  }
};

// Lobe taken by a scatter and the cosines of its weight (for the adjoint).
enum { LOBE_OTHER = 0, LOBE_LAMBERT = 1, LOBE_GGX_SPEC = 2, LOBE_GGX_DIFFUSE = 3 };
struct ScatterInfo {
  int lobe = LOBE_OTHER;
  double ndv = 0, ndl = 0, ndh = 0;
  Vec3 datt;  // d att / d albedo per channel
};

struct Material {
  uint32_t kind = 0;
  uint32_t id = 0;  // descriptor material index
  Texture tex;
  Vec3 albedo;  // metal albedo / dielectric tint
  double fuzz = 0, ior = 1.5, roughness = 0, dispersion = 0, intensity = 1, metallic = 0;

  bool is_specular() const {
    if (kind == LMN_MAT_METAL || kind == LMN_MAT_DIELECTRIC) return true;  // material.rs:90-92, :145-147
    if (kind == LMN_MAT_MICROFACET) return metallic > 0.9 && roughness < 0.1;  // :323-326
    return false;
  }
  Vec3 emitted(double u, double v) const {  // :241-243
    if (kind != LMN_MAT_EMISSIVE) return {};
    return tex.value(u, v) * intensity;
  }
  static double reflectance(double c, double ri) {  // :138-141
    double r0 = (1.0 - ri) / (1.0 + ri);
    r0 = r0 * r0;
    double x = 1.0 - c;
    return r0 + (1.0 - r0) * (x * x * x * x * x);
  }
  static Vec3 fresnel(double c, const Vec3& f0) {  // :291-294
    double x = std::fmax(1.0 - c, 0.0);
    double t = x * x * x * x * x;
    return f0 * (1.0 - t) + Vec3(1, 1, 1) * t;
  }
  static double g1(double ndv, double k) { return ndv / (ndv * (1.0 - k) + k); }  // :280-282
  static double geometry(double ndv, double ndl, double r) {                       // :285-288
    double k = (r + 1.0) * (r + 1.0) / 8.0;
    return g1(ndv, k) * g1(ndl, k);
  }
  static Vec3 sample_ggx(const Vec3& n, double alpha, Rng& rng) {  // :297-319
    double r1 = rng.next_f64(), r2 = rng.next_f64();
    double theta = std::atan(std::sqrt(alpha * alpha * r1 / (1.0 - r1 + 1e-10)));
    double phi = 2.0 * M_PI * r2;
    double st = std::sin(theta), ct = std::cos(theta);
    Vec3 hl(st * std::cos(phi), st * std::sin(phi), ct);
    Vec3 up = std::fabs(n.y) < 0.999 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 tg = up.cross(n).unit();
    Vec3 bt = n.cross(tg);
    return (tg * hl.x + bt * hl.y + n * hl.z).unit();
  }

  bool scatter(const Ray& r, const HitRecord& h, Rng& rng, Ray& out, Vec3& att, ScatterInfo* info = nullptr) const {
    switch (kind) {
      case LMN_MAT_LAMBERTIAN: {  // :59-71
        Vec3 d = h.n + random_unit_vector(rng);
        if (d.near_zero()) d = h.n;
        out = Ray{h.p, d, r.time};
        att = tex.value(h.u, h.v);
        if (info) {
          info->lobe = LOBE_LAMBERT;
          info->datt = Vec3(1, 1, 1);
        }
        return true;
      }
      case LMN_MAT_METAL: {  // :94-106
        Vec3 refl = r.d.unit().reflect(h.n);
        Vec3 sc = refl + random_in_unit_sphere(rng) * fuzz;
        if (sc.dot(h.n) > 0.0) {
          out = Ray{h.p, sc, r.time};
          att = albedo;
          return true;
        }
        return false;
      }
      case LMN_MAT_DIELECTRIC: {  // :149-213
        double ior_ = ior;
        Vec3 chan(1, 1, 1);
        if (dispersion > 0.0) {
          uint32_t c = uint32_t(rng.next_f64() * 3.0);
          ior_ = ior + (c == 0 ? -dispersion : (c == 1 ? 0.0 : dispersion));
          chan = c == 0 ? Vec3(3, 0, 0) : (c == 1 ? Vec3(0, 3, 0) : Vec3(0, 0, 3));
        }
        double eta = h.front ? 1.0 / ior_ : ior_;
        Vec3 ud = r.d.unit();
        double ct = std::min((-ud).dot(h.n), 1.0);
        double st = std::sqrt(1.0 - ct * ct);
        bool cannot = eta * st > 1.0;
        Vec3 dir = (cannot || reflectance(ct, eta) > rng.next_f64()) ? ud.reflect(h.n) : ud.refract(h.n, eta);
        if (roughness > 0.0) dir += random_in_unit_sphere(rng) * roughness;
        Vec3 a;
        bool white = albedo.x == 1.0 && albedo.y == 1.0 && albedo.z == 1.0;
        if (!h.front && !white) {
          double dist = h.t;
          a = Vec3(std::exp(-(1.0 - albedo.x) * dist), std::exp(-(1.0 - albedo.y) * dist),
                   std::exp(-(1.0 - albedo.z) * dist));
        } else {
          a = albedo;
        }
        if (dispersion > 0.0) a = a.hadamard(chan);
        out = Ray{h.p, dir, r.time};
        att = a;
        return true;
      }
      case LMN_MAT_MICROFACET: {  // :328-390
        double alpha = roughness * roughness;
        Vec3 v = (-r.d).unit();
        Vec3 n = h.n;
        Vec3 alb = tex.value(h.u, h.v);
        double ndv = std::fmax(n.dot(v), 0.001);
        Vec3 f0 = Vec3(0.04, 0.04, 0.04) * (1.0 - metallic) + alb * metallic;
        Vec3 fw = fresnel(ndv, f0);
        double sp = (fw.x + fw.y + fw.z) / 3.0;
        sp = std::min(std::max(sp, 0.1), 0.9);
        if (rng.next_f64() < sp) {
          Vec3 hv = sample_ggx(n, alpha, rng);
          Vec3 l = v.reflect_around(hv);
          double ndl = n.dot(l);
          if (ndl <= 0.0) return false;
          double ndh = std::fmax(n.dot(hv), 0.0);
          double vdh = std::fmax(v.dot(hv), 0.0);
          double g = geometry(ndv, ndl, roughness);
          Vec3 f = fresnel(vdh, f0);
          Vec3 w = ndh > 0.001 ? f * (g * vdh / (ndv * ndh)) : Vec3();
          out = Ray{h.p, l, r.time};
          att = w / sp;
          if (info) {
            info->lobe = LOBE_GGX_SPEC;
            info->ndv = ndv;
            info->ndl = ndl;
            info->ndh = ndh;
          }
          return true;
        }
        Vec3 d = n + random_unit_vector(rng);
        if (d.near_zero()) d = n;
        Vec3 kd = (Vec3(1, 1, 1) - fw) * (1.0 - metallic);
        out = Ray{h.p, d, r.time};
        att = kd.hadamard(alb) / (1.0 - sp);
        if (info) {
          info->lobe = LOBE_GGX_DIFFUSE;
          info->datt = kd / (1.0 - sp);
        }
        return true;
      }
      default: return false;  // Emissive (:237-239)
    }
  }
};

// ---------------------------------------------------------------- hittables
struct Hittable {
  virtual bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const = 0;
  virtual bool bbox(Aabb& b) const = 0;
  virtual ~Hittable() = default;
};
using HPtr = std::unique_ptr<Hittable>;

struct Sphere : Hittable {  // sphere.rs:24-67
  Vec3 c;
  double r;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 oc = ray.o - c;
    double a = ray.d.length_squared();
    double hb = oc.dot(ray.d);
    double cc = oc.length_squared() - r * r;
    double disc = hb * hb - a * cc;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    double root = (-hb - sq) / a;
    if (root < tmin || root > tmax) {
      root = (-hb + sq) / a;
      if (root < tmin || root > tmax) return false;
    }
    Vec3 p = ray.at(root);
    Vec3 on = (p - c) / r;
    double theta = std::acos(-on.y);
    double phi = std::atan2(-on.z, on.x) + M_PI;
    h.p = p;
    h.t = root;
    h.u = phi / (2.0 * M_PI);
    h.v = theta / M_PI;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, on);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 rr(r, r, r);
    b = {c - rr, c + rr};
    return true;
  }
};

struct Quad : Hittable {  // quad.rs:20-72
  Vec3 q, u, v, normal, w;
  double D;
  const Material* m;
  int64_t id;
  Quad(Vec3 q_, Vec3 u_, Vec3 v_, const Material* m_, int64_t id_) : q(q_), u(u_), v(v_), m(m_), id(id_) {
    Vec3 n = u.cross(v);
    normal = n.unit();
    D = normal.dot(q);
    w = n / n.dot(n);
  }
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    double denom = normal.dot(r.d);
    if (std::fabs(denom) < 1e-8) return false;
    double t = (D - normal.dot(r.o)) / denom;
    if (t < tmin || t > tmax) return false;
    Vec3 p = r.at(t);
    Vec3 ph = p - q;
    double al = w.dot(ph.cross(v)), be = w.dot(u.cross(ph));
    if (!(al >= 0.0 && al <= 1.0) || !(be >= 0.0 && be <= 1.0)) return false;
    h.p = p;
    h.t = t;
    h.u = al;
    h.v = be;
    h.mat = m;
    h.prim = id;
    set_face(h, r, normal);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 p0 = q, p1 = q + u, p2 = q + v, p3 = q + u + v, e(1e-4, 1e-4, 1e-4);
    b = {p0.vmin(p1).vmin(p2).vmin(p3) - e, p0.vmax(p1).vmax(p2).vmax(p3) + e};
    return true;
  }
};

struct AxisRect : Hittable {  // rect.rs:24-125
  int ax, ia, ib;           // plane axis and in-plane axes
  double a0, a1, b0, b1, k;
  const Material* m;
  int64_t id;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    double t = (k - r.o.axis(ax)) / r.d.axis(ax);
    if (t < tmin || t > tmax) return false;
    double a = r.o.axis(ia) + t * r.d.axis(ia);
    double b = r.o.axis(ib) + t * r.d.axis(ib);
    if (a < a0 || a > a1 || b < b0 || b > b1) return false;
    h.u = (a - a0) / (a1 - a0);
    h.v = (b - b0) / (b1 - b0);
    h.t = t;
    h.p = r.at(t);
    h.mat = m;
    h.prim = id;
    Vec3 n(ax == 0 ? 1 : 0, ax == 1 ? 1 : 0, ax == 2 ? 1 : 0);
    set_face(h, r, n);
    return true;
  }
  bool bbox(Aabb& b) const override {
    double lo[3], hi[3];
    lo[ia] = a0;
    hi[ia] = a1;
    lo[ib] = b0;
    hi[ib] = b1;
    lo[ax] = k - 0.0001;
    hi[ax] = k + 0.0001;
    b = {v3(lo), v3(hi)};
    return true;
  }
};

struct Disk : Hittable {  // disk.rs:8-66
  Vec3 c, n;  // n: unit, (0, 1, 0) when near zero (Disk::new)
  double r;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    double denom = n.dot(ray.d);
    if (std::fabs(denom) < 1e-8) return false;
    double t = (c - ray.o).dot(n) / denom;
    if (t < tmin || t > tmax) return false;
    Vec3 point = ray.at(t);
    Vec3 offset = point - c;
    if (offset.length_squared() > r * r) return false;
    double rr = offset.length() / r;
    double theta = std::atan2(offset.z, offset.x);
    h.u = theta / (2.0 * M_PI) + 0.5;
    h.v = rr;
    h.t = t;
    h.p = point;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, n);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 e(r, r, r);
    b = {c - e, c + e};
    return true;
  }
};

// Cylinder (cylinder.rs:8-81) and Cone (cone.rs:8-86): Y-axis aligned,
// y in [y_min, y_max]; the cone's apex at y_max.
struct AxialQuadric : Hittable {
  bool cone;
  Vec3 c;
  double radius, y_min, y_max;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    double height = y_max - y_min;
    double k = 0.0, oy = 0.0;
    if (cone) {
      if (std::fabs(height) < 1e-10) return false;
      k = (radius / height) * (radius / height);
      oy = ray.o.y - y_max;
    }
    double dx = ray.d.x, dz = ray.d.z, dy = ray.d.y;
    double ox = ray.o.x - c.x, oz = ray.o.z - c.z;
    double a, half_b, cc;
    if (cone) {
      a = dx * dx + dz * dz - k * dy * dy;
      half_b = ox * dx + oz * dz - k * oy * dy;
      cc = ox * ox + oz * oz - k * oy * oy;
    } else {
      a = dx * dx + dz * dz;
      half_b = ox * dx + oz * dz;
      cc = ox * ox + oz * oz - radius * radius;
    }
    double disc = half_b * half_b - a * cc;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    for (double sign : {-1.0, 1.0}) {
      double root = (-half_b + sign * sq) / a;
      if (root < tmin || root > tmax) continue;
      double y = ray.o.y + root * ray.d.y;
      if (y < y_min || y > y_max) continue;
      Vec3 point = ray.at(root);
      Vec3 outward;
      if (cone) {
        double r_at_y = radius * (y_max - y) / height;
        double nx = (point.x - c.x) / std::fmax(r_at_y, 1e-10);
        double nz = (point.z - c.z) / std::fmax(r_at_y, 1e-10);
        outward = Vec3(nx, radius / height, nz).unit();
        h.v = (y - y_min) / height;
      } else {
        outward = Vec3((point.x - c.x) / radius, 0.0, (point.z - c.z) / radius);
        h.v = std::fabs(height) < 1e-10 ? 0.5 : (y - y_min) / height;
      }
      double theta = std::atan2(-(point.z - c.z), point.x - c.x) + M_PI;
      h.u = theta / (2.0 * M_PI);
      h.t = root;
      h.p = point;
      h.mat = m;
      h.prim = id;
      set_face(h, ray, outward);
      return true;
    }
    return false;
  }
  bool bbox(Aabb& b) const override {
    b = {Vec3(c.x - radius, y_min, c.z - radius), Vec3(c.x + radius, y_max, c.z + radius)};
    return true;
  }
};

// Ellipsoid (ellipsoid.rs:8-72): the unit-sphere quadratic after dividing the
// ray by the radii; the normal is the unit gradient.
struct Ellipsoid : Hittable {
  Vec3 c, radii;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 inv(1.0 / radii.x, 1.0 / radii.y, 1.0 / radii.z);
    Vec3 oc((ray.o.x - c.x) * inv.x, (ray.o.y - c.y) * inv.y, (ray.o.z - c.z) * inv.z);
    Vec3 dir(ray.d.x * inv.x, ray.d.y * inv.y, ray.d.z * inv.z);
    double a = dir.length_squared();
    double hb = oc.dot(dir);
    double cc = oc.length_squared() - 1.0;
    double disc = hb * hb - a * cc;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    double root = (-hb - sq) / a;
    if (root < tmin || root > tmax) {
      root = (-hb + sq) / a;
      if (root < tmin || root > tmax) return false;
    }
    Vec3 p = ray.at(root);
    Vec3 on = Vec3((p.x - c.x) / (radii.x * radii.x), (p.y - c.y) / (radii.y * radii.y),
                   (p.z - c.z) / (radii.z * radii.z)).unit();
    h.u = (std::atan2(-on.z, on.x) + M_PI) / (2.0 * M_PI);
    h.v = std::acos(-on.y) / M_PI;
    h.p = p;
    h.t = root;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, on);
    return true;
  }
  bool bbox(Aabb& b) const override {
    b = {c - radii, c + radii};
    return true;
  }
};

// Hemisphere (hemisphere.rs:8-62): the sphere's roots, nearest first, each
// kept only on the upper half y >= c.y.
struct Hemisphere : Hittable {
  Vec3 c;
  double r;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 oc = ray.o - c;
    double a = ray.d.length_squared();
    double hb = oc.dot(ray.d);
    double cc = oc.length_squared() - r * r;
    double disc = hb * hb - a * cc;
    if (disc < 0.0) return false;
    double sq = std::sqrt(disc);
    for (double sign : {-1.0, 1.0}) {
      double root = (-hb + sign * sq) / a;
      if (root < tmin || root > tmax) continue;
      Vec3 p = ray.at(root);
      if (p.y < c.y) continue;
      Vec3 on = (p - c) / r;
      h.u = (std::atan2(-on.z, on.x) + M_PI) / (2.0 * M_PI);
      h.v = std::acos(-on.y) / M_PI;
      h.p = p;
      h.t = root;
      h.mat = m;
      h.prim = id;
      set_face(h, ray, on);
      return true;
    }
    return false;
  }
  bool bbox(Aabb& b) const override {
    Vec3 rr(r, r, r);
    Vec3 lo = c - rr;
    lo.y = c.y;
    b = {lo, c + rr};
    return true;
  }
};

// Annulus (annulus.rs:8-82): plane test with a parallel cull, then the ring
// inner_r^2 <= |offset|^2 <= outer_r^2; u angular, v radial.
struct Annulus : Hittable {
  Vec3 c, n;  // n: unit, (0, 1, 0) when near zero (Annulus::new)
  double r_in, r_out;
  const Material* m;
  int64_t id;
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    double denom = n.dot(ray.d);
    if (std::fabs(denom) < 1e-8) return false;
    double t = (c - ray.o).dot(n) / denom;
    if (t < tmin || t > tmax) return false;
    Vec3 point = ray.at(t);
    Vec3 offset = point - c;
    double dist_sq = offset.length_squared();
    if (dist_sq < r_in * r_in || dist_sq > r_out * r_out) return false;
    double rr = std::sqrt(dist_sq);
    h.u = std::atan2(offset.z, offset.x) / (2.0 * M_PI) + 0.5;
    h.v = (rr - r_in) / (r_out - r_in);
    h.t = t;
    h.p = point;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, n);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 e(r_out, r_out, r_out);
    b = {c - e, c + e};
    return true;
  }
};

// RoundedBox (rounded_box.rs:10-89) and Mobius (mobius.rs:11-121): signed
// distance (estimate) sphere-traced from t_min as the reference does it; in
// exact mode the rounded box, whose distance is exact, is traced in units of
// length to 1e-10 of its size instead (inward to the exit from an origin
// inside it); the Mobius strip keeps the reference's trace in both modes.
struct SdfTraced : Hittable {
  bool mobius = false;
  bool exact = false;
  Vec3 c, half;      // rounded box: centre, half size
  double r = 0;      // rounded box radius / Mobius radius
  double w = 0, th = 0;  // Mobius half-width, thickness
  const Material* m;
  int64_t id;
  double sdf(const Vec3& p) const {
    if (!mobius) {  // rounded_box.rs:30-35
      Vec3 q(std::fabs(p.x) - half.x, std::fabs(p.y) - half.y, std::fabs(p.z) - half.z);
      double outside = Vec3(std::fmax(q.x, 0.0), std::fmax(q.y, 0.0), std::fmax(q.z, 0.0)).length();
      double inside = std::fmin(std::fmax(q.x, std::fmax(q.y, q.z)), 0.0);
      return outside + inside - r;
    }
    double angle = std::atan2(p.z, p.x);  // mobius.rs:37-80
    double best = INFINITY;
    for (double offset : {-M_PI, 0.0, M_PI}) {
      double u = angle + offset;
      double cos_u = std::cos(u), sin_u = std::sin(u), cos_half = std::cos(u / 2.0), sin_half = std::sin(u / 2.0);
      double cx = r * cos_u, cz = r * sin_u;
      Vec3 strip_dir = Vec3(cos_u * cos_half, sin_half, sin_u * cos_half).unit();
      Vec3 dp(p.x - cx, p.y, p.z - cz);
      double v_param = dp.x * strip_dir.x + dp.y * strip_dir.y + dp.z * strip_dir.z;
      double v_clamped = std::fmin(std::fmax(v_param, -w), w);
      double dx = p.x - (cx + v_clamped * strip_dir.x);
      double dy = p.y - v_clamped * strip_dir.y;
      double dz = p.z - (cz + v_clamped * strip_dir.z);
      double dist = std::sqrt(dx * dx + dy * dy + dz * dz) - th;
      if (dist < best) best = dist;
    }
    return best;
  }
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 o = ray.o - c;
    double t = tmin;
    bool found = false;
    if (exact && !mobius) {
      // trace in units of length to 1e-7 of the size, then Newton on the
      // distance along the ray with its analytic gradient
      double len = ray.d.length(), inv_len = 1.0 / len;
      double tol = 1e-7 * (half.x + half.y + half.z + r);
      double sgn = sdf(o + ray.d * t) < 0.0 ? -1.0 : 1.0;
      for (int i = 0; i < 2048 && t <= tmax; ++i) {
        double dist = sgn * sdf(o + ray.d * t);
        if (dist < tol) {
          found = true;
          break;
        }
        t += dist * inv_len;
      }
      for (int k = 0; found && k < 4; ++k) {
        Vec3 q = o + ray.d * t;
        Vec3 g;  // gradient: outward from the inner box, or the nearest face's axis inside it
        Vec3 a(std::fabs(q.x) - half.x, std::fabs(q.y) - half.y, std::fabs(q.z) - half.z);
        Vec3 sg(q.x < 0 ? -1.0 : 1.0, q.y < 0 ? -1.0 : 1.0, q.z < 0 ? -1.0 : 1.0);
        if (a.x > 0 || a.y > 0 || a.z > 0) {
          Vec3 out(std::fmax(a.x, 0.0) * sg.x, std::fmax(a.y, 0.0) * sg.y, std::fmax(a.z, 0.0) * sg.z);
          g = out / out.length();
        } else if (a.x >= a.y && a.x >= a.z) {
          g = Vec3(sg.x, 0, 0);
        } else if (a.y >= a.z) {
          g = Vec3(0, sg.y, 0);
        } else {
          g = Vec3(0, 0, sg.z);
        }
        double dg = g.dot(ray.d);
        if (!(std::fabs(dg) > 1e-6 * len)) break;
        double step = sdf(q) / dg;
        if (!(std::fabs(step) <= 2.0 * tol * inv_len + 1e-12 * std::fabs(t))) break;
        t -= step;
      }
      if (found && !(t >= tmin && t <= tmax)) found = false;
    } else {
      int steps = mobius ? 512 : 256;
      double tol = mobius ? 1e-4 : 1e-5, floor = mobius ? th * 0.3 : 1e-4;
      for (int i = 0; i < steps; ++i) {
        if (t > tmax) break;
        double dist = sdf(o + ray.d * t);
        if (dist < tol) {
          found = true;
          break;
        }
        t += std::fmax(dist, floor);
      }
    }
    if (!found) return false;
    Vec3 point = ray.at(t);
    Vec3 lp = point - c;
    double eps = mobius ? 1e-4 : 1e-5;
    double nx = sdf(lp + Vec3(eps, 0.0, 0.0)) - sdf(lp - Vec3(eps, 0.0, 0.0));
    double ny = sdf(lp + Vec3(0.0, eps, 0.0)) - sdf(lp - Vec3(0.0, eps, 0.0));
    double nz = sdf(lp + Vec3(0.0, 0.0, eps)) - sdf(lp - Vec3(0.0, 0.0, eps));
    Vec3 normal = Vec3(nx, ny, nz).unit();
    if (mobius) {
      h.u = std::atan2(lp.z, lp.x) / (2.0 * M_PI) + 0.5;
      h.v = lp.y / (w + th) * 0.5 + 0.5;
    } else {
      Vec3 total = half + Vec3(r, r, r);
      Vec3 an(std::fabs(normal.x), std::fabs(normal.y), std::fabs(normal.z));
      if (an.x > an.y && an.x > an.z) {
        h.u = (lp.z / total.z + 1.0) * 0.5;
        h.v = (lp.y / total.y + 1.0) * 0.5;
      } else if (an.y > an.z) {
        h.u = (lp.x / total.x + 1.0) * 0.5;
        h.v = (lp.z / total.z + 1.0) * 0.5;
      } else {
        h.u = (lp.x / total.x + 1.0) * 0.5;
        h.v = (lp.y / total.y + 1.0) * 0.5;
      }
    }
    h.p = point;
    h.t = t;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, normal);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 e = mobius ? Vec3(r + w + th, w + th, r + w + th) : half + Vec3(r, r, r);
    b = {c - e, c + e};
    return true;
  }
};

// ray_triangle / ray_tri of the faceted solids (prism.rs:100-126,
// pyramid.rs:29-50): Moller-Trumbore with cull |a| < 1e-10 and t > 1e-10.
static bool solid_ray_triangle(const Ray& ray, const Vec3& v0, const Vec3& v1, const Vec3& v2, double& t, double& u,
                               double& v) {
  Vec3 edge1 = v1 - v0;
  Vec3 edge2 = v2 - v0;
  Vec3 h = ray.d.cross(edge2);
  double a = edge1.dot(h);
  if (std::fabs(a) < 1e-10) return false;
  double f = 1.0 / a;
  Vec3 s = ray.o - v0;
  u = f * s.dot(h);
  if (!(u >= 0.0 && u <= 1.0)) return false;
  Vec3 q = s.cross(edge1);
  v = f * ray.d.dot(q);
  if (v < 0.0 || u + v > 1.0) return false;
  t = f * edge2.dot(q);
  return t > 1e-10;
}

// Prism (prism.rs:39-98) and Pyramid (pyramid.rs:52-111): a fixed list of
// triangles, each with the normal the reference gives its face, tested in
// order under the strict range t_min < t < closest; the box is the solid's.
struct FacetSolid : Hittable {
  struct Facet {
    Vec3 a, b, c, n;
  };
  std::vector<Facet> facets;
  Aabb box;
  const Material* m;
  int64_t id;  // id of the first facet; facet k reports id + k
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    double closest = tmax;
    int best = -1;
    double bu = 0, bv = 0;
    for (size_t k = 0; k < facets.size(); ++k) {
      double t, u, v;
      if (solid_ray_triangle(ray, facets[k].a, facets[k].b, facets[k].c, t, u, v) && t > tmin && t < closest) {
        closest = t;
        best = int(k);
        bu = u;
        bv = v;
      }
    }
    if (best < 0) return false;
    h.p = ray.at(closest);
    h.t = closest;
    h.u = bu;
    h.v = bv;
    h.mat = m;
    h.prim = id + best;
    set_face(h, ray, facets[best].n);
    return true;
  }
  bool bbox(Aabb& b) const override {
    b = box;
    return true;
  }
};

// Wedge (wedge.rs:8-127): bottom, back, left, right and ramp faces in that
// order under the strict range t_min < t < closest.
struct Wedge : Hittable {
  Vec3 mn, mx;
  const Material* m;
  int64_t id;  // id of the bottom face; face k reports id + k
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    double closest = tmax;
    int best = -1;
    Vec3 bn;
    double bu = 0, bv = 0;
    const double depth = mx.z - mn.z, height = mx.y - mn.y, width = mx.x - mn.x;
    auto take = [&](int face, double t, const Vec3& n, double u, double v) {
      closest = t;
      best = face;
      bn = n;
      bu = u;
      bv = v;
    };
    if (std::fabs(ray.d.y) > 1e-10) {  // bottom
      double t = (mn.y - ray.o.y) / ray.d.y;
      if (t > tmin && t < closest) {
        Vec3 p = ray.at(t);
        if (p.x >= mn.x && p.x <= mx.x && p.z >= mn.z && p.z <= mx.z)
          take(0, t, Vec3(0.0, -1.0, 0.0), (p.x - mn.x) / width, (p.z - mn.z) / depth);
      }
    }
    if (std::fabs(ray.d.z) > 1e-10) {  // back
      double t = (mx.z - ray.o.z) / ray.d.z;
      if (t > tmin && t < closest) {
        Vec3 p = ray.at(t);
        if (p.x >= mn.x && p.x <= mx.x && p.y >= mn.y && p.y <= mx.y)
          take(1, t, Vec3(0.0, 0.0, 1.0), (p.x - mn.x) / width, (p.y - mn.y) / height);
      }
    }
    for (int side = 0; side < 2; ++side) {  // left, right
      if (!(std::fabs(ray.d.x) > 1e-10)) break;
      double xs = side ? mx.x : mn.x;
      double t = (xs - ray.o.x) / ray.d.x;
      if (t > tmin && t < closest) {
        Vec3 p = ray.at(t);
        double ramp_y = (p.z - mn.z) / depth * height + mn.y;
        if (p.z >= mn.z && p.z <= mx.z && p.y >= mn.y && p.y <= ramp_y)
          take(2 + side, t, Vec3(side ? 1.0 : -1.0, 0.0, 0.0), (p.z - mn.z) / depth, (p.y - mn.y) / height);
      }
    }
    double slope = height / depth;  // ramp
    Vec3 normal = Vec3(0.0, 1.0, -slope).unit();
    double denom = ray.d.dot(normal);
    if (std::fabs(denom) > 1e-10) {
      double t = (mn - ray.o).dot(normal) / denom;
      if (t > tmin && t < closest) {
        Vec3 p = ray.at(t);
        if (p.x >= mn.x && p.x <= mx.x && p.z >= mn.z && p.z <= mx.z) {
          double expected_y = (p.z - mn.z) / depth * height + mn.y;
          if (std::fabs(p.y - expected_y) < 0.01) take(4, t, normal, (p.x - mn.x) / width, (p.z - mn.z) / depth);
        }
      }
    }
    if (best < 0) return false;
    h.p = ray.at(closest);
    h.t = closest;
    h.u = bu;
    h.v = bv;
    h.mat = m;
    h.prim = id + best;
    set_face(h, ray, bn);
    return true;
  }
  bool bbox(Aabb& b) const override {
    b = {mn, mx};
    return true;
  }
};

struct Triangle : Hittable {  // triangle.rs:23-70 / obj.rs:260-319
  Vec3 v0, v1, v2;
  const Material* m;
  int64_t id;
  const lmn_vertex_attr* attr = nullptr;  // MeshTriangle's normals / uvs (obj.rs:291-303)
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    Vec3 e1 = v1 - v0, e2 = v2 - v0;
    Vec3 hh = r.d.cross(e2);
    double a = e1.dot(hh);
    if (std::fabs(a) < 1e-8) return false;
    double f = 1.0 / a;
    Vec3 s = r.o - v0;
    double u = f * s.dot(hh);
    if (!(u >= 0.0 && u <= 1.0)) return false;
    Vec3 q = s.cross(e1);
    double v = f * r.d.dot(q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * e2.dot(q);
    if (t < tmin || t > tmax) return false;
    h.p = r.at(t);
    h.t = t;
    h.u = u;
    h.v = v;
    h.mat = m;
    h.prim = id;
    Vec3 outward = e1.cross(e2).unit();
    if (attr && attr->has_normals)
      outward = (v3(attr->normal[0]) * (1.0 - u - v) + v3(attr->normal[1]) * u + v3(attr->normal[2]) * v).unit();
    if (attr && attr->has_uvs) {
      double w = 1.0 - u - v;
      h.u = attr->uv[0][0] * w + attr->uv[1][0] * u + attr->uv[2][0] * v;
      h.v = attr->uv[0][1] * w + attr->uv[1][1] * u + attr->uv[2][1] * v;
    }
    set_face(h, r, outward);
    return true;
  }
  bool bbox(Aabb& b) const override {
    Vec3 e(1e-4, 1e-4, 1e-4);
    b = {v0.vmin(v1).vmin(v2) - e, v0.vmax(v1).vmax(v2) + e};
    return true;
  }
};

struct Plane : Hittable {  // plane.rs:18-73
  Vec3 pt, n, ua, va;
  const Material* m;
  int64_t id;
  Plane(Vec3 p, Vec3 nn, const Material* m_, int64_t id_) : pt(p), m(m_), id(id_) {
    n = nn.near_zero() ? Vec3(0, 1, 0) : nn.unit();
    Vec3 up = std::fabs(n.y) > 0.999 ? Vec3(1, 0, 0) : Vec3(0, 1, 0);
    ua = up.cross(n).unit();
    va = n.cross(ua);
  }
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    double denom = n.dot(r.d);
    if (std::fabs(denom) < 1e-8) return false;
    double t = (pt - r.o).dot(n) / denom;
    if (t < tmin || t > tmax) return false;
    Vec3 p = r.at(t);
    Vec3 l = p - pt;
    double a = l.dot(ua), b = l.dot(va);
    h.p = p;
    h.t = t;
    h.u = std::fabs(a - std::trunc(a));
    h.v = std::fabs(b - std::trunc(b));
    h.mat = m;
    h.prim = id;
    set_face(h, r, n);
    return true;
  }
  bool bbox(Aabb&) const override { return false; }
};

// ------- marched primitives: reference semantics, plus an independent exact
// mode (dense bracketing of the implicit function's sign changes, bisection
// in long double) used to pin the GPU's closed-form/bracketed roots.
struct Torus : Hittable {  // torus.rs:24-111
  Vec3 c;
  double R, r;
  const Material* m;
  int64_t id;
  bool exact;
  double sdf(const Vec3& p) const {
    double q = std::sqrt(p.x * p.x + p.z * p.z) - R;
    return std::sqrt(q * q + p.y * p.y) - r;
  }
  bool finish(const Ray& ray, double th, HitRecord& h, bool analytic_normal) const {
    Vec3 point = ray.at(th);
    Vec3 lp = point - c;
    Vec3 nn;
    if (analytic_normal) {
      double k = lp.length_squared() + R * R - r * r;
      nn = Vec3(4.0 * k * lp.x - 8.0 * R * R * lp.x, 4.0 * k * lp.y, 4.0 * k * lp.z - 8.0 * R * R * lp.z).unit();
    } else {
      const double e = 1e-5;
      double nx = sdf(lp + Vec3(e, 0, 0)) - sdf(lp - Vec3(e, 0, 0));
      double ny = sdf(lp + Vec3(0, e, 0)) - sdf(lp - Vec3(0, e, 0));
      double nz = sdf(lp + Vec3(0, 0, e)) - sdf(lp - Vec3(0, 0, e));
      nn = Vec3(nx, ny, nz).unit();
    }
    double theta = std::atan2(lp.z, lp.x);
    double xz = std::sqrt(lp.x * lp.x + lp.z * lp.z);
    double phi = std::atan2(lp.y, xz - R);
    h.p = point;
    h.t = th;
    h.u = theta / (2.0 * M_PI) + 0.5;
    h.v = phi / (2.0 * M_PI) + 0.5;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, nn);
    return true;
  }
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 o = ray.o - c, d = ray.d;
    if (exact) {
      // f(t) = (|p|^2 + R^2 - r^2)^2 - 4R^2(px^2 + pz^2); roots bracketed on
      // a grid fine enough to separate the tube crossings
      auto f = [&](long double t) {
        long double px = o.x + t * d.x, py = o.y + t * d.y, pz = o.z + t * d.z;
        long double k = px * px + py * py + pz * pz + (long double)R * R - (long double)r * r;
        return k * k - 4.0L * R * R * (px * px + pz * pz);
      };
      // clip to the bounding box
      double lo[3] = {-(R + r), -r, -(R + r)}, hi[3] = {R + r, r, R + r};
      double t0 = tmin, t1 = tmax;
      for (int a = 0; a < 3; ++a) {
        double inv = 1.0 / d.axis(a);
        double ta = (lo[a] - o.axis(a)) * inv, tb = (hi[a] - o.axis(a)) * inv;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::fmax(t0, ta);
        t1 = std::fmin(t1, tb);
      }
      if (!(t0 <= t1)) return false;
      int steps = 4096;
      long double h0 = ((long double)t1 - t0) / steps;
      long double ta = t0, fa = f(ta);
      if (fa == 0.0L) return finish(ray, double(ta), h, true);
      for (int i = 1; i <= steps; ++i) {
        long double tb = (i == steps) ? (long double)t1 : t0 + h0 * i;
        long double fb = f(tb);
        if ((fa > 0) != (fb > 0) || fb == 0.0L) {
          long double a = ta, b = tb, fl = fa;
          for (int it = 0; it < 200; ++it) {
            long double mid = 0.5L * (a + b);
            if (mid == a || mid == b) break;
            long double fm = f(mid);
            if ((fm > 0) == (fl > 0)) {
              a = mid;
              fl = fm;
            } else {
              b = mid;
            }
          }
          double th = double(0.5L * (a + b));
          if (th >= tmin && th <= tmax) return finish(ray, th, h, true);
        }
        ta = tb;
        fa = fb;
      }
      return false;
    }
    double step = r * 0.5;
    double ms = std::fmin((tmax - tmin) / step, 500.0);
    long max_steps = (ms == ms && ms > 0.0) ? long(ms) : 0;  // `as usize` saturates
    double t = tmin;
    double prev = sdf(o + d * t);
    for (long i = 0; i < max_steps; ++i) {
      t += step;
      if (t > tmax) break;
      double dist = sdf(o + d * t);
      if (prev > 0.0 && dist < 0.0) {
        double lo = t - step, hi = t;
        for (int k = 0; k < 20; ++k) {
          double mid = (lo + hi) * 0.5;
          if (sdf(o + d * mid) > 0.0)
            lo = mid;
          else
            hi = mid;
        }
        double th = (lo + hi) * 0.5;
        if (th < tmin || th > tmax) {
          prev = dist;
          continue;
        }
        return finish(ray, th, h, false);
      }
      prev = dist;
    }
    return false;
  }
  bool bbox(Aabb& b) const override {
    double rr = R + r;
    b = {c - Vec3(rr, r, rr), c + Vec3(rr, r, rr)};
    return true;
  }
};

struct Superellipsoid : Hittable {  // superellipsoid.rs:26-97
  Vec3 c, s;
  double e1, e2;
  const Material* m;
  int64_t id;
  bool exact;
  double f(const Vec3& p) const {
    double ax = std::fabs(p.x), ay = std::fabs(p.y), az = std::fabs(p.z);
    double ex2 = 2.0 / e2;
    double xz = std::pow(ax, ex2) + std::pow(az, ex2);
    double ex1 = 2.0 / e1;
    double val = std::pow(xz, e2 / e1) + std::pow(ay, ex1);
    return std::pow(val, e1 / 2.0) - 1.0;
  }
  bool finish(const Ray& ray, double t, HitRecord& h) const {
    Vec3 point = ray.at(t);
    Vec3 lp((point.x - c.x) / s.x, (point.y - c.y) / s.y, (point.z - c.z) / s.z);
    const double e = 1e-5;
    double nx = f(lp + Vec3(e, 0, 0)) - f(lp - Vec3(e, 0, 0));
    double ny = f(lp + Vec3(0, e, 0)) - f(lp - Vec3(0, e, 0));
    double nz = f(lp + Vec3(0, 0, e)) - f(lp - Vec3(0, 0, e));
    Vec3 nn = Vec3(nx / s.x, ny / s.y, nz / s.z).unit();
    h.p = point;
    h.t = t;
    h.u = 0.5 + std::atan2(lp.z, lp.x) / (2.0 * M_PI);
    h.v = 0.5 + std::min(std::max(std::asin(lp.y), -1.0), 1.0) / M_PI;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, nn);
    return true;
  }
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 o((ray.o.x - c.x) / s.x, (ray.o.y - c.y) / s.y, (ray.o.z - c.z) / s.z);
    Vec3 d(ray.d.x / s.x, ray.d.y / s.y, ray.d.z / s.z);
    if (exact) {
      // unit box, padded so that entries at the axis points (where the
      // surface touches the box) start strictly outside
      const double B = 1.0 + 1e-9;
      double t0 = tmin, t1 = tmax;
      for (int a = 0; a < 3; ++a) {
        double inv = 1.0 / d.axis(a);
        double ta = (-B - o.axis(a)) * inv, tb = (B - o.axis(a)) * inv;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::fmax(t0, ta);
        t1 = std::fmin(t1, tb);
      }
      if (!(t0 <= t1)) return false;
      int steps = 8192;
      double hs = (t1 - t0) / steps;
      double ta = t0, fa = f(o + d * ta);
      for (int i = 1; i <= steps; ++i) {
        double tb = (i == steps) ? t1 : t0 + hs * i;
        double fb = f(o + d * tb);
        if ((fa > 0.0) != (fb > 0.0)) {
          double a = ta, b = tb, fl = fa;
          for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (a + b);
            if (mid == a || mid == b) break;
            double fm = f(o + d * mid);
            if ((fm > 0.0) == (fl > 0.0)) {
              a = mid;
              fl = fm;
            } else {
              b = mid;
            }
          }
          double th = 0.5 * (a + b);
          if (th >= tmin && th <= tmax) return finish(ray, th, h);
        }
        ta = tb;
        fa = fb;
      }
      return false;
    }
    double t = tmin;
    for (int i = 0; i < 256; ++i) {
      if (t > tmax) break;
      double dist = f(o + d * t);
      if (std::fabs(dist) < 1e-5) return finish(ray, t, h);
      t += std::fmax(dist, 1e-4);
    }
    return false;
  }
  bool bbox(Aabb& b) const override {
    b = {c - s, c + s};
    return true;
  }
};

struct Spring : Hittable {  // spring.rs:35-114
  Vec3 c;
  double coil, tube, pitch, turns;
  const Material* m;
  int64_t id;
  bool exact;
  double sdf(const Vec3& p) const {
    double height = pitch * turns, two_pi = 2.0 * M_PI;
    double angle = std::atan2(p.z, p.x);
    double base_y = pitch * angle / two_pi;
    double best = INFINITY;
    int kmin = int(std::floor((p.y - base_y) / pitch)) - 1;
    for (int k = kmin; k <= kmin + 3; ++k) {
      double hy = base_y + pitch * double(k);
      if (hy < -tube || hy > height + tube) continue;
      double hyc = std::min(std::max(hy, 0.0), height);
      double ha = (hyc / pitch) * two_pi;
      double hx = coil * std::cos(ha), hz = coil * std::sin(ha);
      double dx = p.x - hx, dy = p.y - hyc, dz = p.z - hz;
      double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (dist < best) best = dist;
    }
    return best - tube;
  }
  bool finish(const Ray& ray, double t, HitRecord& h) const {
    Vec3 point = ray.at(t);
    Vec3 lp = point - c;
    const double e = 1e-4;
    double nx = sdf(lp + Vec3(e, 0, 0)) - sdf(lp - Vec3(e, 0, 0));
    double ny = sdf(lp + Vec3(0, e, 0)) - sdf(lp - Vec3(0, e, 0));
    double nz = sdf(lp + Vec3(0, 0, e)) - sdf(lp - Vec3(0, 0, e));
    h.p = point;
    h.t = t;
    h.u = lp.y / (pitch * turns);
    h.v = std::atan2(lp.z, lp.x) / (2.0 * M_PI) + 0.5;
    h.mat = m;
    h.prim = id;
    set_face(h, ray, Vec3(nx, ny, nz).unit());
    return true;
  }
  bool hit(const Ray& ray, double tmin, double tmax, HitRecord& h) const override {
    Vec3 o = ray.o - c, d = ray.d;
    if (exact) {
      double rr = coil + tube, hgt = pitch * turns;
      double lo[3] = {-rr, -tube, -rr}, hi[3] = {rr, hgt + tube, rr};
      double t0 = tmin, t1 = tmax;
      for (int a = 0; a < 3; ++a) {
        double inv = 1.0 / d.axis(a);
        double ta = (lo[a] - o.axis(a)) * inv, tb = (hi[a] - o.axis(a)) * inv;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::fmax(t0, ta);
        t1 = std::fmin(t1, tb);
      }
      if (!(t0 <= t1)) return false;
      double dl = d.length();
      double hs = 0.02 * tube / dl;
      double ta = t0;
      if (sdf(o + d * ta) <= 0.0) return finish(ray, t0, h);
      while (ta < t1) {
        double tb = std::fmin(ta + hs, t1);
        if (sdf(o + d * tb) <= 0.0) {
          double a = ta, b = tb;
          for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (a + b);
            if (mid == a || mid == b) break;
            if (sdf(o + d * mid) > 0.0)
              a = mid;
            else
              b = mid;
          }
          return finish(ray, 0.5 * (a + b), h);
        }
        ta = tb;
      }
      return false;
    }
    double step_min = tube * 0.3;
    double t = tmin;
    for (int i = 0; i < 512; ++i) {
      if (t > tmax) break;
      double dist = sdf(o + d * t);
      if (dist < 1e-4) return finish(ray, t, h);
      t += std::fmax(dist, step_min);
    }
    return false;
  }
  bool bbox(Aabb& b) const override {
    double rr = coil + tube, height = pitch * turns;
    b = {c + Vec3(-rr, -tube, -rr), c + Vec3(rr, height + tube, rr)};
    return true;
  }
};

// ---------------------------------------------------------------- wrappers
struct HittableList : Hittable {  // hit.rs:72-102
  std::vector<HPtr> objs;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    double closest = tmax;
    bool any = false;
    HitRecord tmp;
    for (auto& o : objs)
      if (o->hit(r, tmin, closest, tmp)) {
        closest = tmp.t;
        h = tmp;
        any = true;
      }
    return any;
  }
  bool bbox(Aabb& b) const override {
    bool have = false;
    for (auto& o : objs) {
      Aabb bb;
      if (o->bbox(bb)) {
        b = have ? Aabb::surrounding(b, bb) : bb;
        have = true;
      }
    }
    return have;
  }
};

struct Translate : Hittable {  // transform.rs:7-29
  HPtr in;
  Vec3 off;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    Ray m{r.o - off, r.d, 0.0};  // Ray::new drops the time
    if (!in->hit(m, tmin, tmax, h)) return false;
    h.p += off;
    return true;
  }
  bool bbox(Aabb& b) const override {
    Aabb bb;
    if (!in->bbox(bb)) return false;
    b = {bb.mn + off, bb.mx + off};
    return true;
  }
};

// RotateX / RotateY / RotateZ (transform.rs:32-256), axis = 0, 1, 2
struct Rotate : Hittable {
  HPtr in;
  int axis;
  double s, c;
  bool has_box;
  Aabb box;
  Rotate(HPtr inner, int ax, double deg) : in(std::move(inner)), axis(ax) {
    double rad = deg * (M_PI / 180.0);
    s = std::sin(rad);
    c = std::cos(rad);
    Aabb bb;
    has_box = in->bbox(bb);
    if (has_box) {
      Vec3 mn(INFINITY, INFINITY, INFINITY), mx(-INFINITY, -INFINITY, -INFINITY);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int k = 0; k < 2; ++k) {
            double x = i * bb.mx.x + (1 - i) * bb.mn.x;
            double y = j * bb.mx.y + (1 - j) * bb.mn.y;
            double z = k * bb.mx.z + (1 - k) * bb.mn.z;
            Vec3 nv = fwd(Vec3(x, y, z));
            mn = mn.vmin(nv);
            mx = mx.vmax(nv);
          }
      box = {mn, mx};
    }
  }
  Vec3 fwd(const Vec3& p) const {  // object -> world
    if (axis == 0) return {p.x, c * p.y - s * p.z, s * p.y + c * p.z};
    if (axis == 1) return {c * p.x + s * p.z, p.y, -s * p.x + c * p.z};
    return {c * p.x - s * p.y, s * p.x + c * p.y, p.z};
  }
  Vec3 inv(const Vec3& p) const {  // world -> object
    if (axis == 0) return {p.x, c * p.y + s * p.z, -s * p.y + c * p.z};
    if (axis == 1) return {c * p.x - s * p.z, p.y, s * p.x + c * p.z};
    return {c * p.x + s * p.y, -s * p.x + c * p.y, p.z};
  }
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    // RotateY re-casts with Ray::new (time 0); X and Z keep the time
    Ray m{inv(r.o), inv(r.d), axis == 1 ? 0.0 : r.time};
    if (!in->hit(m, tmin, tmax, h)) return false;
    h.p = fwd(h.p);
    h.n = fwd(h.n);
    return true;
  }
  bool bbox(Aabb& b) const override {
    if (has_box) b = box;
    return has_box;
  }
};

struct Scale : Hittable {  // transform.rs:259-284
  HPtr in;
  double f, invf;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    Ray m{r.o * invf, r.d * invf, 0.0};
    if (!in->hit(m, tmin, tmax, h)) return false;
    h.p = h.p * f;
    h.t *= f;
    return true;
  }
  bool bbox(Aabb& b) const override {
    Aabb bb;
    if (!in->bbox(bb)) return false;
    b = {bb.mn * f, bb.mx * f};
    return true;
  }
};

struct NonUniformScale : Hittable {  // transform.rs:287-356
  HPtr in;
  Vec3 sc, inv;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    Vec3 so(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
    Vec3 sd(r.d.x * inv.x, r.d.y * inv.y, r.d.z * inv.z);
    Ray m{so, sd, 0.0};
    if (!in->hit(m, tmin, tmax, h)) return false;
    h.p = Vec3(h.p.x * sc.x, h.p.y * sc.y, h.p.z * sc.z);
    h.n = Vec3(h.n.x * inv.x, h.n.y * inv.y, h.n.z * inv.z).unit();
    h.t *= sd.length() / r.d.length();
    return true;
  }
  bool bbox(Aabb& b) const override {
    Aabb bb;
    if (!in->bbox(bb)) return false;
    Vec3 mn(bb.mn.x * sc.x, bb.mn.y * sc.y, bb.mn.z * sc.z), mx(bb.mx.x * sc.x, bb.mx.y * sc.y, bb.mx.z * sc.z);
    b = {mn.vmin(mx), mn.vmax(mx)};
    return true;
  }
};

// ---------------------------------------------------------------- BVH
struct BvhNode : Hittable {  // bvh.rs:6-68
  HPtr left, right;
  Aabb box;
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const override {
    if (!box.hit(r, tmin, tmax)) return false;
    HitRecord lh;
    bool l = left->hit(r, tmin, tmax, lh);
    double tr = l ? lh.t : tmax;
    HitRecord rh;
    bool rr = right->hit(r, tmin, tr, rh);
    if (rr) {
      h = rh;  // right_hit.or(left_hit)
      return true;
    }
    if (l) h = lh;
    return l;
  }
  bool bbox(Aabb& b) const override {
    b = box;
    return true;
  }
};

double centroid_axis(const Aabb& b, int a) { return (b.mn.axis(a) + b.mx.axis(a)) * 0.5; }

// find_best_split (bvh.rs:71-163)
void find_best_split(const std::vector<HPtr>& objs, int& best_axis, size_t& best_mid) {
  size_t n = objs.size();
  const int NB = 12;
  Aabb overall;
  objs[0]->bbox(overall);
  Vec3 cmin(INFINITY, INFINITY, INFINITY), cmax(-INFINITY, -INFINITY, -INFINITY);
  for (auto& o : objs) {
    Aabb bb;
    if (o->bbox(bb)) {
      overall = Aabb::surrounding(overall, bb);
      Vec3 c = (bb.mn + bb.mx) * 0.5;
      cmin = cmin.vmin(c);
      cmax = cmax.vmax(c);
    }
  }
  double osa = overall.surface_area();
  best_axis = 0;
  best_mid = n / 2;
  if (osa < 1e-10) return;
  double best_cost = INFINITY;
  for (int ax = 0; ax < 3; ++ax) {
    double ext = cmax.axis(ax) - cmin.axis(ax);
    if (ext < 1e-10) continue;
    size_t cnt[NB] = {0};
    bool has[NB] = {false};
    Aabb bb_b[NB];
    for (auto& o : objs) {
      Aabb bb;
      if (!o->bbox(bb)) continue;
      double c = (bb.mn.axis(ax) + bb.mx.axis(ax)) * 0.5;
      double bf = (c - cmin.axis(ax)) / ext * double(NB);
      size_t b = (bf == bf && bf > 0.0) ? size_t(bf) : 0;  // `as usize` saturates
      b = std::min<size_t>(b, NB - 1);
      cnt[b]++;
      bb_b[b] = has[b] ? Aabb::surrounding(bb_b[b], bb) : bb;
      has[b] = true;
    }
    for (int split = 1; split < NB; ++split) {
      size_t lc = 0, rc = 0;
      bool lh = false, rh = false;
      Aabb lb, rb;
      for (int i = 0; i < split; ++i) {
        lc += cnt[i];
        if (has[i]) {
          lb = lh ? Aabb::surrounding(lb, bb_b[i]) : bb_b[i];
          lh = true;
        }
      }
      for (int i = split; i < NB; ++i) {
        rc += cnt[i];
        if (has[i]) {
          rb = rh ? Aabb::surrounding(rb, bb_b[i]) : bb_b[i];
          rh = true;
        }
      }
      if (lc == 0 || rc == 0) continue;
      double ls = lh ? lb.surface_area() : 0.0, rs = rh ? rb.surface_area() : 0.0;
      double cost = 1.0 + (double(lc) * ls + double(rc) * rs) / osa;
      if (cost < best_cost) {
        best_cost = cost;
        best_axis = ax;
        best_mid = lc;
      }
    }
  }
}

// BvhNode::build (bvh.rs:16-49)
HPtr bvh_build(std::vector<HPtr> objs) {
  size_t n = objs.size();
  if (n == 1) return std::move(objs[0]);
  auto node = std::make_unique<BvhNode>();
  if (n == 2) {
    Aabb a, b;
    objs[0]->bbox(a);
    objs[1]->bbox(b);
    node->box = Aabb::surrounding(a, b);
    node->left = std::move(objs[0]);
    node->right = std::move(objs[1]);
    return node;
  }
  int ax;
  size_t mid;
  find_best_split(objs, ax, mid);
  auto key = [ax](const HPtr& o) {
    Aabb bb;
    return o->bbox(bb) ? centroid_axis(bb, ax) : 0.0;
  };
  std::stable_sort(objs.begin(), objs.end(), [&](const HPtr& a, const HPtr& b) { return key(a) < key(b); });
  mid = std::min(std::max<size_t>(mid, 1), n - 1);
  std::vector<HPtr> rhalf;
  for (size_t i = mid; i < n; ++i) rhalf.push_back(std::move(objs[i]));
  objs.resize(mid);
  node->left = bvh_build(std::move(objs));
  node->right = bvh_build(std::move(rhalf));
  Aabb a, b;
  node->left->bbox(a);
  node->right->bbox(b);
  node->box = Aabb::surrounding(a, b);
  return node;
}

// ---------------------------------------------------------------- world
struct Light {
  uint32_t kind;
  Vec3 o, u, v, n, e;
  double radius;
  uint32_t mat;  // emissive material it was registered from, ~0u if none
};

struct World {
  std::vector<std::unique_ptr<Material>> mats;
  HPtr bvh;
  std::vector<HPtr> unbounded;
  std::vector<Light> lights;
  uint32_t bg_kind = 0;
  Vec3 bg0, bg1, bg_dir;
  double bg_int = 0;
  // camera
  Vec3 origin, lower_left, horizontal, vertical, cu, cv;
  double lens_radius = 0;

  // SceneWorld::hit (scene.rs:817-834)
  bool hit(const Ray& r, double tmin, double tmax, HitRecord& h) const {
    double closest = tmax;
    bool any = false;
    HitRecord tmp;
    if (bvh && bvh->hit(r, tmin, closest, tmp)) {
      closest = tmp.t;
      h = tmp;
      any = true;
    }
    for (auto& o : unbounded)
      if (o->hit(r, tmin, closest, tmp)) {
        closest = tmp.t;
        h = tmp;
        any = true;
      }
    return any;
  }

  // Background::color (render.rs:72-172)
  Vec3 background(const Ray& r) const {
    if (bg_kind == LMN_BG_SOLID) return bg0;
    Vec3 ud = r.d.unit();
    double t = 0.5 * (ud.y + 1.0);
    if (bg_kind == LMN_BG_SKY) return Vec3(1, 1, 1) * (1.0 - t) + Vec3(0.5, 0.7, 1.0) * t;
    if (bg_kind == LMN_BG_GRADIENT) return bg0 * (1.0 - t) + bg1 * t;
    Vec3 sky = Vec3(1, 1, 1) * (1.0 - t) + bg1 * t;
    double sd = ud.dot(bg_dir.unit());
    if (sd > 0.9995) return bg0 * bg_int;
    if (sd > 0.995) {
      double glow = (sd - 0.995) / (0.9995 - 0.995);
      return sky + bg0 * (bg_int * 0.3 * glow);
    }
    return sky;
  }

  // Camera::get_ray (camera.rs:82-103)
  Ray get_ray(double s, double t, Rng& rng) const {
    Vec3 rd = random_in_unit_disk(rng) * lens_radius;
    Vec3 off = cu * rd.x + cv * rd.y;
    Ray r;
    r.o = origin + off;
    r.d = lower_left + horizontal * s + vertical * t - origin - off;
    r.time = rng.next_f64();
    return r;
  }
};

struct Counters {
  uint64_t ext = 0, shadow = 0;
};

// build_onb (render.rs:1333-1343)
void build_onb(const Vec3& n, Vec3& a, Vec3& b) {
  if (n.z < -0.9999999) {
    a = Vec3(0, -1, 0);
    b = Vec3(-1, 0, 0);
    return;
  }
  double aa = 1.0 / (1.0 + n.z), bb = -n.x * n.y * aa;
  a = Vec3(1.0 - n.x * n.x * aa, bb, -n.x);
  b = Vec3(bb, 1.0 - n.y * n.y * aa, -n.y);
}

// sample_direct_light (render.rs:1346-1466)
// D_out (when given) receives the unoccluded contribution per unit albedo and
// light_mat the picked light's material (both left zero / ~0u otherwise).
Vec3 sample_direct_light(const World& w, const Vec3& hp, const Vec3& normal, const Vec3& albedo, Rng& rng,
                         Counters& cnt, Vec3* D_out = nullptr, uint32_t* light_mat = nullptr) {
  if (D_out) *D_out = Vec3();
  if (light_mat) *light_mat = ~0u;
  if (w.lights.empty()) return {};
  size_t nl = w.lights.size();
  double lf = rng.next_f64() * double(nl);
  size_t li = (lf == lf && lf > 0.0) ? size_t(lf) : 0;
  li = std::min(li, nl - 1);
  const Light& L = w.lights[li];
  Vec3 lp, ln, em = L.e;
  double area;
  if (L.kind == LMN_LIGHT_SPHERE) {
    Vec3 tc = L.o - hp;
    double dc2 = tc.length_squared();
    if (dc2 < L.radius * L.radius) return {};
    double dc = std::sqrt(dc2);
    double stm = L.radius / dc;
    double ctm = std::sqrt(std::fmax(1.0 - stm * stm, 0.0));
    double r1 = rng.next_f64(), r2 = rng.next_f64();
    double ct = 1.0 - r1 * (1.0 - ctm);
    double st = std::sqrt(std::fmax(1.0 - ct * ct, 0.0));
    double phi = 2.0 * M_PI * r2;
    Vec3 ww = tc / dc;
    Vec3 a, b;
    build_onb(ww, a, b);
    Vec3 dir = (a * (st * std::cos(phi)) + b * (st * std::sin(phi)) + ww * ct).unit();
    Vec3 oc = hp - L.o;
    double bc = oc.dot(dir);
    double cc = oc.length_squared() - L.radius * L.radius;
    double disc = bc * bc - cc;
    double t;
    if (disc > 0.0) {
      double sd = std::sqrt(disc);
      double t1 = -bc - sd;
      t = t1 > 0.001 ? t1 : -bc + sd;
    } else {
      t = dc;
    }
    lp = hp + dir * t;
    ln = (lp - L.o).unit();
    double sa = 2.0 * M_PI * (1.0 - ctm);
    area = sa * t * t / std::fmax(std::fabs((-dir).dot(ln)), 0.001);
  } else if (L.kind == LMN_LIGHT_RECT) {
    double s = rng.next_f64(), t = rng.next_f64();
    lp = L.o + L.u * s + L.v * t;
    ln = L.n;
    area = L.u.cross(L.v).length();
  } else {
    double r1 = rng.next_f64(), r2 = rng.next_f64();
    double r = std::sqrt(r1) * L.radius;
    double th = 2.0 * M_PI * r2;
    Vec3 ww = L.n;
    Vec3 av = std::fabs(ww.x) > 0.9 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 ut = ww.cross(av).unit();
    Vec3 vt = ww.cross(ut);
    lp = L.o + ut * (r * std::cos(th)) + vt * (r * std::sin(th));
    ln = L.n;
    area = M_PI * L.radius * L.radius;
  }
  Vec3 tl = lp - hp;
  double d2 = tl.length_squared();
  double dist = std::sqrt(d2);
  Vec3 dir = tl / dist;
  double ct = normal.dot(dir);
  if (ct <= 0.0) return {};
  double ctl = std::fabs((-dir).dot(ln));
  if (ctl <= 0.001) return {};
  Ray sr{hp, dir, 0.0};
  HitRecord hh;
  cnt.shadow++;
  if (w.hit(sr, 0.001, dist - 0.001, hh)) return {};
  double g = ct * ctl * area * double(nl) / (M_PI * d2);
  if (D_out) *D_out = em * g;
  if (light_mat) *light_mat = L.mat;
  return em.hadamard(albedo) * g;
}

// ray_color (render.rs:1472-1537)
Vec3 ray_color(const World& w, const Ray& r, Rng& rng, uint32_t depth, bool skip, Counters& cnt) {
  if (depth == 0) return {};
  HitRecord h;
  cnt.ext++;
  if (w.hit(r, 0.001, INFINITY, h)) {
    Vec3 em = skip ? Vec3() : h.mat->emitted(h.u, h.v);
    Ray sc;
    Vec3 att;
    if (h.mat->scatter(r, h, rng, sc, att)) {
      double mc = std::fmax(std::fmax(att.x, att.y), att.z);
      double surv = std::min(std::max(mc, 0.05), 1.0);
      if (depth < 47 && rng.next_f64() > surv) return em;
      double rrw = depth < 47 ? 1.0 / surv : 1.0;
      bool use_nee = !h.mat->is_specular() && !w.lights.empty();
      Vec3 direct = use_nee ? sample_direct_light(w, h.p, h.n, att, rng, cnt) : Vec3();
      Vec3 indirect = att.hadamard(ray_color(w, sc, rng, depth - 1, use_nee, cnt)) * rrw;
      Vec3 res = em + direct + indirect;
      if (std::isnan(res.x) || std::isnan(res.y) || std::isnan(res.z)) return {};
      return res;
    }
    return em;
  }
  return w.background(r);
}

// ---- path-replay adjoint (FP64): ray_color again, with the derivative of its
// expected value (directions held fixed, Russian roulette through its
// survival probability q; the NEE term is att (.) D and not divided by q):
//   d/dtheta = dlog f (att/q (.) Lnext + direct) + [q = max att] (dlog f - dlog p) direct
// with dlog f = datt/att for albedo, dlog D + dlog G for GGX roughness (and
// dlog p = dlog D), and emission linear in the intensity.
struct AdjCtx {
  int32_t tex_id = -1, em_mat = -1;
  std::vector<double> g_tex;    // texels * 3 of the parameter texture
  std::vector<double> g_rough;  // per material
  double g_em = 0;              // sum of the adjoint-weighted emission terms (to divide by I)
};

void ggx_dlog(double r, double ndv, double ndl, double ndh, double& dlogd, double& dlogg) {
  double a = r * r, c2 = ndh * ndh;
  double den = c2 * (a * a - 1.0) + 1.0;
  dlogd = (2.0 / a - 4.0 * c2 * a / den) * 2.0 * r;  // D = a^2 / (pi den^2)
  double k = (r + 1.0) * (r + 1.0) / 8.0, dk = (r + 1.0) / 4.0;
  dlogg = -((1.0 - ndv) / (ndv * (1.0 - k) + k) + (1.0 - ndl) / (ndl * (1.0 - k) + k)) * dk;
}

Vec3 ray_color_adj(const World& w, const Ray& r, Rng& rng, uint32_t depth, bool skip, Counters& cnt,
                   const Vec3& Aw, AdjCtx& ctx) {
  if (depth == 0) return {};
  HitRecord h;
  cnt.ext++;
  if (!w.hit(r, 0.001, INFINITY, h)) return w.background(r);
  Vec3 em = skip ? Vec3() : h.mat->emitted(h.u, h.v);
  if (!skip && int32_t(h.mat->id) == ctx.em_mat) ctx.g_em += Aw.dot(em);
  Ray sc;
  Vec3 att;
  ScatterInfo si;
  if (!h.mat->scatter(r, h, rng, sc, att, &si)) return em;
  double mc = std::fmax(std::fmax(att.x, att.y), att.z);
  double surv = std::min(std::max(mc, 0.05), 1.0);
  bool rr = depth < 47;
  if (rr && rng.next_f64() > surv) return em;
  double rrw = rr ? 1.0 / surv : 1.0;
  bool use_nee = !h.mat->is_specular() && !w.lights.empty();
  Vec3 D;
  uint32_t lmat = ~0u;
  Vec3 direct = use_nee ? sample_direct_light(w, h.p, h.n, att, rng, cnt, &D, &lmat) : Vec3();
  Vec3 Lnext = ray_color_adj(w, sc, rng, depth - 1, use_nee, cnt, Aw.hadamard(att) * rrw, ctx);
  Vec3 res = em + direct + att.hadamard(Lnext) * rrw;
  if (std::isnan(res.x) || std::isnan(res.y) || std::isnan(res.z)) return {};
  double Ad = Aw.dot(direct);
  if (lmat != ~0u && int32_t(lmat) == ctx.em_mat) ctx.g_em += Ad;
  bool rru = rr && mc > 0.05 && mc < 1.0;
  int am = att.x >= att.y ? (att.x >= att.z ? 0 : 2) : (att.y >= att.z ? 1 : 2);
  if ((si.lobe == LOBE_LAMBERT || si.lobe == LOBE_GGX_DIFFUSE) && h.mat->tex.texels && h.mat->tex.id == ctx.tex_id) {
    int64_t t = h.mat->tex.texel(h.u, h.v);
    double aw[3] = {Aw.x, Aw.y, Aw.z}, da[3] = {si.datt.x, si.datt.y, si.datt.z};
    double ln[3] = {Lnext.x, Lnext.y, Lnext.z}, dd[3] = {D.x, D.y, D.z};
    for (int c = 0; c < 3; ++c) {
      double g = aw[c] * da[c] * (rrw * ln[c] + dd[c]);
      if (rru && c == am) g += da[c] * Ad / surv;
      ctx.g_tex[size_t(t) * 3 + c] += g;
    }
  }
  if (si.lobe == LOBE_GGX_SPEC && h.mat->kind == LMN_MAT_MICROFACET && mc > 0.0) {
    double dd, dg;
    ggx_dlog(h.mat->roughness, si.ndv, si.ndl, si.ndh, dd, dg);
    double g = (dd + dg) * (Aw.dot(att.hadamard(Lnext) * rrw) + Ad);
    if (rru) g += dg * Ad;
    ctx.g_rough[h.mat->id] += g;
  }
  return res;
}

// PixelFilter::weight (render.rs:198-237)
double mitchell_1d(double x) {
  x = std::fabs(x);
  const double b = 1.0 / 3.0, c = 1.0 / 3.0;
  if (x < 1.0)
    return ((12.0 - 9.0 * b - 6.0 * c) * x * x * x + (-18.0 + 12.0 * b + 6.0 * c) * x * x + (6.0 - 2.0 * b)) / 6.0;
  if (x < 2.0)
    return ((-b - 6.0 * c) * x * x * x + (6.0 * b + 30.0 * c) * x * x + (-12.0 * b - 48.0 * c) * x + (8.0 * b + 24.0 * c)) /
           6.0;
  return 0.0;
}
double filter_weight(uint32_t f, double dx, double dy) {
  switch (f) {
    case LMN_FILTER_TRIANGLE: return std::fmax(1.0 - std::fabs(dx) * 2.0, 0.0) * std::fmax(1.0 - std::fabs(dy) * 2.0, 0.0);
    case LMN_FILTER_GAUSSIAN: {
      double sg = 0.35;
      return std::exp(-(dx * dx + dy * dy) / (2.0 * sg * sg));
    }
    case LMN_FILTER_MITCHELL: return mitchell_1d(dx * 2.0) * mitchell_1d(dy * 2.0);
    default: return 1.0;
  }
}

// ---------------------------------------------------------------- world build
HPtr make_leaf(const lmn_scene_desc& d, uint32_t i, const World& w, bool exact) {
  const lmn_prim& p = d.prims[i];
  const Material* m = w.mats[p.material].get();
  const double* q = p.p;
  int64_t id = int64_t(i);
  switch (p.kind) {
    case LMN_PRIM_SPHERE: {
      auto s = std::make_unique<Sphere>();
      s->c = v3(q);
      s->r = q[3];
      s->m = m;
      s->id = id;
      return s;
    }
    case LMN_PRIM_QUAD: return std::make_unique<Quad>(v3(q), v3(q + 3), v3(q + 6), m, id);
    case LMN_PRIM_RECT_XY:
    case LMN_PRIM_RECT_XZ:
    case LMN_PRIM_RECT_YZ: {
      auto r = std::make_unique<AxisRect>();
      if (p.kind == LMN_PRIM_RECT_XY) { r->ax = 2; r->ia = 0; r->ib = 1; }
      if (p.kind == LMN_PRIM_RECT_XZ) { r->ax = 1; r->ia = 0; r->ib = 2; }
      if (p.kind == LMN_PRIM_RECT_YZ) { r->ax = 0; r->ia = 1; r->ib = 2; }
      r->a0 = q[0]; r->a1 = q[1]; r->b0 = q[2]; r->b1 = q[3]; r->k = q[4];
      r->m = m;
      r->id = id;
      return r;
    }
    case LMN_PRIM_TRIANGLE:
    case LMN_PRIM_MESH_TRIANGLE: {
      auto t = std::make_unique<Triangle>();
      t->v0 = v3(q);
      t->v1 = v3(q + 3);
      t->v2 = v3(q + 6);
      t->m = m;
      t->id = id;
      if (p.kind == LMN_PRIM_MESH_TRIANGLE) t->attr = &d.vertex_attrs[size_t(q[9])];  // checked by build_world
      return t;
    }
    case LMN_PRIM_PLANE: return std::make_unique<Plane>(v3(q), v3(q + 3), m, id);
    case LMN_PRIM_TORUS: {
      auto t = std::make_unique<Torus>();
      t->c = v3(q);
      t->R = q[3];
      t->r = q[4];
      t->m = m;
      t->id = id;
      t->exact = exact;
      return t;
    }
    case LMN_PRIM_SUPERELLIPSOID: {
      auto s = std::make_unique<Superellipsoid>();
      s->c = v3(q);
      s->s = v3(q + 3);
      s->e1 = q[6];
      s->e2 = q[7];
      s->m = m;
      s->id = id;
      s->exact = exact;
      return s;
    }
    case LMN_PRIM_DISK: {
      auto dk = std::make_unique<Disk>();
      dk->c = v3(q);
      Vec3 n = v3(q + 3);
      dk->n = (std::fabs(n.x) < 1e-8 && std::fabs(n.y) < 1e-8 && std::fabs(n.z) < 1e-8) ? Vec3(0, 1, 0) : n.unit();
      dk->r = q[6];
      dk->m = m;
      dk->id = id;
      return dk;
    }
    case LMN_PRIM_ROUNDED_BOX:
    case LMN_PRIM_MOBIUS: {
      auto sd = std::make_unique<SdfTraced>();
      sd->mobius = p.kind == LMN_PRIM_MOBIUS;
      sd->exact = exact;
      sd->c = v3(q);
      if (sd->mobius) {
        sd->r = q[3];
        sd->w = q[4];
        sd->th = q[5];
      } else {
        sd->half = v3(q + 3);
        sd->r = q[6];
      }
      sd->m = m;
      sd->id = id;
      return sd;
    }
    case LMN_PRIM_ELLIPSOID: {
      auto e = std::make_unique<Ellipsoid>();
      e->c = v3(q);
      e->radii = v3(q + 3);
      e->m = m;
      e->id = id;
      return e;
    }
    case LMN_PRIM_HEMISPHERE: {
      auto hs = std::make_unique<Hemisphere>();
      hs->c = v3(q);
      hs->r = q[3];
      hs->m = m;
      hs->id = id;
      return hs;
    }
    case LMN_PRIM_ANNULUS: {
      auto a = std::make_unique<Annulus>();
      a->c = v3(q);
      Vec3 n = v3(q + 3);
      a->n = (std::fabs(n.x) < 1e-8 && std::fabs(n.y) < 1e-8 && std::fabs(n.z) < 1e-8) ? Vec3(0, 1, 0) : n.unit();
      a->r_in = q[6];
      a->r_out = q[7];
      a->m = m;
      a->id = id;
      return a;
    }
    case LMN_PRIM_CYLINDER:
    case LMN_PRIM_CONE: {
      auto a = std::make_unique<AxialQuadric>();
      a->cone = p.kind == LMN_PRIM_CONE;
      a->c = v3(q);
      a->radius = q[3];
      a->y_min = q[4];
      a->y_max = q[5];
      a->m = m;
      a->id = id;
      return a;
    }
    default: {
      auto s = std::make_unique<Spring>();
      s->c = v3(q);
      s->coil = q[3];
      s->tube = q[4];
      s->pitch = q[5];
      s->turns = q[6];
      s->m = m;
      s->id = id;
      s->exact = exact;
      return s;
    }
  }
}

HPtr wrap(HPtr obj, const lmn_transform& t) {
  for (uint32_t k = 0; k < t.n_ops && k < LMN_MAX_XF_OPS; ++k) {
    const lmn_xform_op& op = t.ops[k];
    switch (op.kind) {
      case LMN_XF_ROTATE_X: obj = std::make_unique<Rotate>(std::move(obj), 0, op.v[0]); break;
      case LMN_XF_ROTATE_Y: obj = std::make_unique<Rotate>(std::move(obj), 1, op.v[0]); break;
      case LMN_XF_ROTATE_Z: obj = std::make_unique<Rotate>(std::move(obj), 2, op.v[0]); break;
      case LMN_XF_TRANSLATE: {
        auto tr = std::make_unique<Translate>();
        tr->in = std::move(obj);
        tr->off = v3(op.v);
        obj = std::move(tr);
        break;
      }
      case LMN_XF_SCALE: {
        auto s = std::make_unique<Scale>();
        s->in = std::move(obj);
        s->f = std::fabs(op.v[0]) < 1e-10 ? 1.0 : op.v[0];
        s->invf = 1.0 / s->f;
        obj = std::move(s);
        break;
      }
      case LMN_XF_SCALE_XYZ: {
        auto s = std::make_unique<NonUniformScale>();
        s->in = std::move(obj);
        double f[3];
        for (int a = 0; a < 3; ++a) f[a] = std::fabs(op.v[a]) < 1e-10 ? 1.0 : op.v[a];
        s->sc = v3(f);
        s->inv = Vec3(1.0 / f[0], 1.0 / f[1], 1.0 / f[2]);
        obj = std::move(s);
        break;
      }
      default: break;
    }
  }
  return obj;
}

bool build_world(const lmn_scene_desc& d, bool exact, World& w) {
  for (uint32_t i = 0; i < d.n_prims; ++i)
    if (d.prims[i].kind == LMN_PRIM_MESH_TRIANGLE &&
        !(d.prims[i].p[9] >= 0.0 && d.prims[i].p[9] < double(d.n_vertex_attrs) && d.vertex_attrs)) {
      g_err = "vertex attribute index out of range";
      return false;
    }
  for (uint32_t i = 0; i < d.n_materials; ++i) {
    const lmn_material& m = d.materials[i];
    auto mm = std::make_unique<Material>();
    mm->kind = m.kind;
    mm->id = i;
    mm->tex.color = v3(m.color);
    if (m.texture >= 0) {
      if (uint32_t(m.texture) >= d.n_textures) {
        g_err = "texture index out of range";
        return false;
      }
      const lmn_texture& t = d.textures[m.texture];
      mm->tex.texels = t.rgb;
      mm->tex.w = t.width;
      mm->tex.h = t.height;
      mm->tex.id = m.texture;
    }
    mm->albedo = v3(m.color);
    mm->fuzz = m.fuzz;
    mm->ior = m.ior;
    mm->roughness = m.roughness;
    mm->dispersion = m.dispersion;
    mm->intensity = m.intensity;
    mm->metallic = m.metallic;
    w.mats.push_back(std::move(mm));
  }
  // world list in load order: prims sharing `object` form one entry
  std::vector<HPtr> bounded;
  uint32_t i = 0;
  while (i < d.n_prims) {
    uint32_t j = i + 1;
    while (j < d.n_prims && d.prims[j].object == d.prims[i].object) ++j;
    if (d.prims[i].material >= d.n_materials) {
      g_err = "material index out of range";
      return false;
    }
    HPtr obj;
    if (d.prims[i].kind == LMN_PRIM_FACET) {  // a prism or pyramid
      auto fs = std::make_unique<FacetSolid>();
      for (uint32_t k = i; k < j; ++k) {
        const double* q = d.prims[k].p;
        fs->facets.push_back({v3(q), v3(q + 3), v3(q + 6), v3(q + 9)});
      }
      // the solid's box: prism.rs:87-97 / pyramid.rs:99-110 span exactly its vertices
      Vec3 lo = fs->facets[0].a, hi = lo;
      for (auto& f : fs->facets)
        for (const Vec3* v : {&f.a, &f.b, &f.c}) {
          lo = Vec3(std::fmin(lo.x, v->x), std::fmin(lo.y, v->y), std::fmin(lo.z, v->z));
          hi = Vec3(std::fmax(hi.x, v->x), std::fmax(hi.y, v->y), std::fmax(hi.z, v->z));
        }
      fs->box = {lo, hi};
      fs->m = w.mats[d.prims[i].material].get();
      fs->id = int64_t(i);
      obj = std::move(fs);
    } else if (d.prims[i].kind == LMN_PRIM_WEDGE_FACE) {
      auto wg = std::make_unique<Wedge>();
      wg->mn = v3(d.prims[i].p);
      wg->mx = v3(d.prims[i].p + 3);
      wg->m = w.mats[d.prims[i].material].get();
      wg->id = int64_t(i);
      obj = std::move(wg);
    } else if (j - i == 1) {
      obj = make_leaf(d, i, w, exact);
    } else {
      auto list = std::make_unique<HittableList>();
      for (uint32_t k = i; k < j; ++k) list->objs.push_back(make_leaf(d, k, w, exact));
      obj = std::move(list);
    }
    if (d.prims[i].transform >= 0) obj = wrap(std::move(obj), d.transforms[d.prims[i].transform]);
    Aabb bb;
    if (obj->bbox(bb))
      bounded.push_back(std::move(obj));
    else
      w.unbounded.push_back(std::move(obj));
    i = j;
  }
  if (!bounded.empty()) w.bvh = bvh_build(std::move(bounded));  // SceneWorld::from_list (scene.rs:786-808)
  for (uint32_t k = 0; k < d.n_lights; ++k) {
    const lmn_light& l = d.lights[k];
    w.lights.push_back(
        Light{l.kind, v3(l.origin), v3(l.u), v3(l.v), v3(l.normal), v3(l.emission), l.radius, l.material});
  }
  const lmn_background& b = d.config.background;
  w.bg_kind = b.kind;
  w.bg0 = v3(b.color0);
  w.bg1 = v3(b.color1);
  w.bg_dir = v3(b.direction);
  w.bg_int = b.intensity;
  const lmn_camera& c = d.camera;
  w.origin = v3(c.origin);
  w.lower_left = v3(c.lower_left);
  w.horizontal = v3(c.horizontal);
  w.vertical = v3(c.vertical);
  w.cu = v3(c.u);
  w.cv = v3(c.v);
  w.lens_radius = c.lens_radius;
  return true;
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_trace(const lmn_scene_desc* desc, const lmn_ray* rays, size_t n, uint32_t exact_marched, lmn_hit* hits) {
  if (!desc || (n && (!rays || !hits))) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, exact_marched != 0, w)) return LMN_ERR_INVALID;
  for (size_t i = 0; i < n; ++i) {
    Ray r{v3(rays[i].origin), v3(rays[i].direction), 0.0};
    HitRecord h;
    lmn_hit& o = hits[i];
    std::memset(&o, 0, sizeof o);
    if (w.hit(r, rays[i].t_min, rays[i].t_max, h)) {
      o.t = h.t;
      o.prim = h.prim;
      o.u = h.u;
      o.v = h.v;
      o.normal[0] = h.n.x;
      o.normal[1] = h.n.y;
      o.normal[2] = h.n.z;
      o.front_face = h.front ? 1 : 0;
    } else {
      o.prim = -1;
    }
  }
  return LMN_OK;
}

// Pixel loop (render.rs:1557-1747), non-adaptive, rows scheduled dynamically
// over threads like Rayon's work stealing over rows.
int oracle_render(const lmn_scene_desc* desc, const oracle_params* p, double* mean, double* var, oracle_stats* st) {
  if (!desc || !p || !mean) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, p->exact_marched != 0, w)) return LMN_ERR_INVALID;
  uint32_t fw = std::max<uint32_t>(p->width, 1), fh = std::max<uint32_t>(p->height, 1);
  uint32_t cx = 0, cy = 0, cw = fw, ch = fh;
  if (p->has_crop) {
    cx = std::min(p->crop[0], fw);
    cy = std::min(p->crop[1], fh);
    cw = std::min(p->crop[2], fw);
    ch = std::min(p->crop[3], fh);
  }
  uint32_t rb = p->row_begin, re = p->row_end ? std::min(p->row_end, ch) : ch;
  uint32_t rs = p->row_step ? p->row_step : 1;
  uint32_t n_rows = re > rb ? (re - rb + rs - 1) / rs : 0;
  uint32_t sq = uint32_t(std::ceil(std::sqrt(double(p->spp))));
  uint64_t fkey = frame_key(p->seed, p->scene_key);
  std::atomic<uint32_t> next_row{0};
  std::atomic<uint64_t> n_ext{0}, n_sh{0}, n_cam{0};
  std::atomic<bool> time_expired{false};
  const uint32_t adaptive_min_samples = p->adaptive ? std::max(sq * sq / 4, 16u) : 0u;  // render.rs:1592
  unsigned nt = p->threads ? p->threads : std::max(1u, std::thread::hardware_concurrency());
  auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    Counters cnt;
    for (;;) {
      uint32_t r = next_row.fetch_add(1);
      if (r >= n_rows) break;
      // render.rs:1600-1605: checked once per row
      if (p->time_limit > 0.0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > p->time_limit)
        time_expired.store(true);
      const bool expired = time_expired.load();
      uint32_t j = rb + r * rs;
      uint32_t gj = j + cy;
      double y = double(fh - 1 - gj);
      std::unique_ptr<XoshiroRng> row_rng;
      if (p->rng_mode == ORACLE_RNG_REFERENCE) row_rng = std::make_unique<XoshiroRng>(uint64_t(gj) * p->seed);
      uint64_t row_rays = 0;
      for (uint32_t i = 0; i < cw; ++i) {
        uint32_t gi = i + cx;
        size_t idx = (size_t(r) * cw + i) * 3;
        if (expired) {  // render.rs:1614-1624: one sample through the pixel centre
          CounterRng crng(path_key(fkey, uint64_t(gj) * fw + gi, 0xffffffffu));
          Rng& rng = row_rng ? static_cast<Rng&>(*row_rng) : static_cast<Rng&>(crng);
          double u = (double(gi) + 0.5) / double(fw - 1);
          double v = (y + 0.5) / double(fh - 1);
          Ray ray = w.get_ray(u, v, rng);
          Vec3 smp = ray_color(w, ray, rng, p->max_depth, false, cnt);
          ++row_rays;
          double lum = 0.2126 * smp.x + 0.7152 * smp.y + 0.0722 * smp.z;
          Vec3 c = lum > 100.0 ? smp * (100.0 / lum) : smp;
          mean[idx] = c.x;
          mean[idx + 1] = c.y;
          mean[idx + 2] = c.z;
          if (var)
            for (int k = 0; k < 3; ++k) var[idx + k] = 0.0;
          continue;
        }
        if (p->adaptive) {  // render.rs:1626-1677
          Vec3 mn;
          double m2_lum = 0.0, mean_lum = 0.0;
          uint32_t nn = 0;
          double m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
          bool stop = false;
          for (uint32_t sy = 0; sy < sq && !stop; ++sy) {
            for (uint32_t sx = 0; sx < sq; ++sx) {
              CounterRng crng(path_key(fkey, uint64_t(gj) * fw + gi, sy * sq + sx));
              Rng& rng = row_rng ? static_cast<Rng&>(*row_rng) : static_cast<Rng&>(crng);
              double u = (double(gi) + (double(sx) + rng.next_f64()) / double(sq)) / double(fw - 1);
              double v = (y + (double(sy) + rng.next_f64()) / double(sq)) / double(fh - 1);
              Ray ray = w.get_ray(u, v, rng);
              Vec3 smp = ray_color(w, ray, rng, p->max_depth, false, cnt);
              double luminance = 0.2126 * smp.x + 0.7152 * smp.y + 0.0722 * smp.z;
              Vec3 clamped = luminance > 100.0 ? smp * (100.0 / luminance) : smp;
              nn += 1;
              double lum = 0.2126 * clamped.x + 0.7152 * clamped.y + 0.0722 * clamped.z;
              double delta_lum = lum - mean_lum;
              mean_lum += delta_lum / double(nn);
              double delta_lum2 = lum - mean_lum;
              m2_lum += delta_lum * delta_lum2;
              mn = Vec3(mn.x + (clamped.x - mn.x) / double(nn), mn.y + (clamped.y - mn.y) / double(nn),
                        mn.z + (clamped.z - mn.z) / double(nn));
              double cv[3] = {clamped.x, clamped.y, clamped.z};
              for (int k = 0; k < 3; ++k) {  // Welford per channel, for the statistical tests
                double dlt = cv[k] - m1[k];
                m1[k] += dlt / nn;
                m2[k] += dlt * (cv[k] - m1[k]);
              }
              if (nn >= adaptive_min_samples && nn > 1) {
                double variance = m2_lum / double(nn - 1);
                double std_error = std::sqrt(variance / double(nn));
                if (std_error < p->adaptive_threshold) {
                  stop = true;
                  break;
                }
              }
            }
          }
          row_rays += nn;
          mean[idx] = mn.x;
          mean[idx + 1] = mn.y;
          mean[idx + 2] = mn.z;
          if (var)
            for (int k = 0; k < 3; ++k) var[idx + k] = nn > 1 ? m2[k] / (nn - 1) : 0.0;
          continue;
        }
        Vec3 color;
        double tw = 0.0;
        double m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
        uint32_t ns = 0;
        for (uint32_t sy = 0; sy < sq; ++sy)
          for (uint32_t sx = 0; sx < sq; ++sx) {
            CounterRng crng(path_key(fkey, uint64_t(gj) * fw + gi, sy * sq + sx));
            Rng& rng = row_rng ? static_cast<Rng&>(*row_rng) : static_cast<Rng&>(crng);
            double jx = (double(sx) + rng.next_f64()) / double(sq);
            double jy = (double(sy) + rng.next_f64()) / double(sq);
            double u = (double(gi) + jx) / double(fw - 1);
            double v = (y + jy) / double(fh - 1);
            Ray r = w.get_ray(u, v, rng);
            Vec3 s = ray_color(w, r, rng, p->max_depth, false, cnt);
            double lum = 0.2126 * s.x + 0.7152 * s.y + 0.0722 * s.z;
            Vec3 c = lum > 100.0 ? s * (100.0 / lum) : s;
            double wt = filter_weight(p->filter, jx - 0.5, jy - 0.5);
            color += c * wt;
            tw += wt;
            ++ns;
            double cv[3] = {c.x, c.y, c.z};
            for (int k = 0; k < 3; ++k) {  // Welford
              double dlt = cv[k] - m1[k];
              m1[k] += dlt / ns;
              m2[k] += dlt * (cv[k] - m1[k]);
            }
          }
        Vec3 out = tw > 0.0 ? color / tw : color;
        row_rays += uint64_t(sq) * sq;
        mean[idx] = out.x;
        mean[idx + 1] = out.y;
        mean[idx + 2] = out.z;
        if (var)
          for (int k = 0; k < 3; ++k) var[idx + k] = ns > 1 ? m2[k] / (ns - 1) : 0.0;
      }
      n_cam += row_rays;
    }
    n_ext += cnt.ext;
    n_sh += cnt.shadow;
  };
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; ++k) th.emplace_back(worker);
  for (auto& t : th) t.join();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st) {
    st->camera_rays = n_cam.load();
    st->extension_rays = n_ext.load() - st->camera_rays;
    st->shadow_rays = n_sh.load();
    st->render_secs = secs;
    st->threads = nt;
  }
  return LMN_OK;
}

int oracle_scatter(const lmn_scene_desc* desc, uint32_t material, const double ray_dir[3], const double point[3],
                   const double normal[3], int front_face, double t, double u, double v, uint64_t key,
                   uint32_t counter, double out_dir[3], double out_att[3], uint32_t* out_counter) {
  World w;
  if (!desc || material >= desc->n_materials) {
    g_err = "bad material";
    return -LMN_ERR_INVALID;
  }
  if (!build_world(*desc, false, w)) return -LMN_ERR_INVALID;
  Ray r{Vec3(0, 0, 0), v3(ray_dir), 0.0};
  HitRecord h;
  h.p = v3(point);
  h.n = v3(normal);
  h.front = front_face != 0;
  h.t = t;
  h.u = u;
  h.v = v;
  h.mat = w.mats[material].get();
  CounterRng rng(key, counter);
  Ray out;
  Vec3 att;
  bool ok = h.mat->scatter(r, h, rng, out, att);
  out_dir[0] = out.d.x;
  out_dir[1] = out.d.y;
  out_dir[2] = out.d.z;
  out_att[0] = att.x;
  out_att[1] = att.y;
  out_att[2] = att.z;
  if (out_counter) *out_counter = rng.n;
  return ok ? 1 : 0;
}

// n scatters at one surface point, sample k drawn from key0 + k (the
// material tests' Monte Carlo estimates).
int oracle_scatter_batch(const lmn_scene_desc* desc, uint32_t material, const double ray_dir[3],
                         const double point[3], const double normal[3], int front_face, double t, double u, double v,
                         uint64_t key0, uint32_t n, double* out_dir, double* out_att, uint8_t* out_ok) {
  World w;
  if (!desc || material >= desc->n_materials || !out_dir || !out_att || !out_ok) {
    g_err = "bad argument";
    return LMN_ERR_INVALID;
  }
  if (!build_world(*desc, false, w)) return LMN_ERR_INVALID;
  Ray r{Vec3(0, 0, 0), v3(ray_dir), 0.0};
  HitRecord h;
  h.p = v3(point);
  h.n = v3(normal);
  h.front = front_face != 0;
  h.t = t;
  h.u = u;
  h.v = v;
  h.mat = w.mats[material].get();
  for (uint32_t k = 0; k < n; ++k) {
    CounterRng rng(key0 + k, 0);
    Ray out;
    Vec3 att;
    out_ok[k] = h.mat->scatter(r, h, rng, out, att) ? 1 : 0;
    out_dir[3 * k] = out.d.x;
    out_dir[3 * k + 1] = out.d.y;
    out_dir[3 * k + 2] = out.d.z;
    out_att[3 * k] = att.x;
    out_att[3 * k + 1] = att.y;
    out_att[3 * k + 2] = att.z;
  }
  return LMN_OK;
}

// The reference-stream generator (rand 0.9.2 SmallRng = xoshiro256++):
// n u64 outputs from the raw state s[4] (state != NULL), or from
// seed_from_u64(seed); f64 draws = (u64 >> 11) * 2^-53 into out_f64 when given.
int oracle_xoshiro(const uint64_t* state, uint64_t seed, uint32_t n, uint64_t* out_u64, double* out_f64) {
  if (!out_u64) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  XoshiroRng g(seed);
  if (state)
    for (int i = 0; i < 4; ++i) g.s[i] = state[i];
  for (uint32_t k = 0; k < n; ++k) {
    XoshiroRng c = g;
    out_u64[k] = g.next_u64();
    if (out_f64) out_f64[k] = c.next_f64();
  }
  return LMN_OK;
}

// ---- single-function evaluations for the known-answer tests ---------------

int oracle_nee(const lmn_scene_desc* desc, const double point[3], const double normal[3], const double albedo[3],
               uint64_t key, uint32_t n, double out_mean[3], uint64_t* out_shadow_rays) {
  if (!desc || !point || !normal || !albedo || !out_mean || n == 0) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, true, w)) return LMN_ERR_INVALID;
  Counters cnt;
  Vec3 sum;
  for (uint32_t i = 0; i < n; ++i) {
    CounterRng rng(key + uint64_t(i) * 0x9E3779B97F4A7C15ull);
    sum += sample_direct_light(w, v3(point), v3(normal), v3(albedo), rng, cnt);
  }
  out_mean[0] = sum.x / n;
  out_mean[1] = sum.y / n;
  out_mean[2] = sum.z / n;
  if (out_shadow_rays) *out_shadow_rays = cnt.shadow;
  return LMN_OK;
}

int oracle_camera_ray(const lmn_scene_desc* desc, double s, double t, uint64_t key, double origin[3],
                      double dir[3]) {
  if (!desc || !origin || !dir) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, true, w)) return LMN_ERR_INVALID;
  CounterRng rng(key);
  Ray r = w.get_ray(s, t, rng);
  origin[0] = r.o.x;
  origin[1] = r.o.y;
  origin[2] = r.o.z;
  dir[0] = r.d.x;
  dir[1] = r.d.y;
  dir[2] = r.d.z;
  return LMN_OK;
}

int oracle_emitted(const lmn_scene_desc* desc, uint32_t material, double u, double v, double out[3]) {
  if (!desc || !out || material >= desc->n_materials) {
    g_err = "bad material";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, true, w)) return LMN_ERR_INVALID;
  Vec3 e = w.mats[material]->emitted(u, v);
  out[0] = e.x;
  out[1] = e.y;
  out[2] = e.z;
  return LMN_OK;
}

int oracle_background(const lmn_scene_desc* desc, const double dir[3], double out[3]) {
  if (!desc || !dir || !out) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, true, w)) return LMN_ERR_INVALID;
  Vec3 c = w.background(Ray{Vec3(), v3(dir), 0.0});
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
  return LMN_OK;
}

double oracle_filter_weight(uint32_t filter, double dx, double dy) { return filter_weight(filter, dx, dy); }

double oracle_reflectance(double cosine, double ratio) { return Material::reflectance(cosine, ratio); }

// Path-replay adjoint on the CPU in FP64 (counter RNG): each sample is traced
// once for its radiance (the luminance clamp's Jacobian needs it) and replayed
// from the same seed by ray_color_adj.  Rows are split statically over the
// threads and their sums merged in thread order, so the result is
// deterministic.
int oracle_adjoint(const lmn_scene_desc* desc, const oracle_params* p, int32_t texture, int32_t emission_material,
                   const double* dldi, double* g_tex, double* g_rough, double* g_em) {
  if (!desc || !p || !dldi) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, p->exact_marched != 0, w)) return LMN_ERR_INVALID;
  uint32_t fw = std::max<uint32_t>(p->width, 1), fh = std::max<uint32_t>(p->height, 1);
  uint32_t cx = 0, cy = 0, cw = fw, ch = fh;
  if (p->has_crop) {
    cx = std::min(p->crop[0], fw);
    cy = std::min(p->crop[1], fh);
    cw = std::min(p->crop[2], fw);
    ch = std::min(p->crop[3], fh);
  }
  uint32_t sq = uint32_t(std::ceil(std::sqrt(double(p->spp))));
  uint64_t fkey = frame_key(p->seed, p->scene_key);
  size_t n_tex = 0;
  if (texture >= 0) {
    if (uint32_t(texture) >= desc->n_textures) {
      g_err = "texture out of range";
      return LMN_ERR_INVALID;
    }
    n_tex = size_t(desc->textures[texture].width) * desc->textures[texture].height * 3;
  }
  unsigned nt = p->threads ? p->threads : std::max(1u, std::thread::hardware_concurrency());
  std::vector<AdjCtx> ctx(nt);
  for (auto& c : ctx) {
    c.tex_id = texture;
    c.em_mat = emission_material;
    c.g_tex.assign(n_tex, 0.0);
    c.g_rough.assign(desc->n_materials, 0.0);
  }
  auto worker = [&](unsigned ti) {
    Counters cnt;
    AdjCtx& C = ctx[ti];
    for (uint32_t j = ti; j < ch && p->max_depth > 0; j += nt) {
      uint32_t gj = j + cy;
      double y = double(fh - 1 - gj);
      for (uint32_t i = 0; i < cw; ++i) {
        uint32_t gi = i + cx;
        uint64_t gpix = uint64_t(gj) * fw + gi;
        double tw = 0.0;
        for (uint32_t s = 0; s < sq * sq; ++s) {
          CounterRng rng(path_key(fkey, gpix, s));
          uint32_t sy = s / sq, sx = s % sq;
          double jx = (double(sx) + rng.next_f64()) / double(sq);
          double jy = (double(sy) + rng.next_f64()) / double(sq);
          tw += filter_weight(p->filter, jx - 0.5, jy - 0.5);
        }
        const double* g = dldi + (size_t(j) * cw + i) * 3;
        for (uint32_t s = 0; s < sq * sq; ++s) {
          uint32_t sy = s / sq, sx = s % sq;
          Vec3 L;
          double wt = 0;
          for (int pass = 0; pass < 2; ++pass) {
            CounterRng rng(path_key(fkey, gpix, s));
            double jx = (double(sx) + rng.next_f64()) / double(sq);
            double jy = (double(sy) + rng.next_f64()) / double(sq);
            double u = (double(gi) + jx) / double(fw - 1);
            double v = (y + jy) / double(fh - 1);
            Ray r = w.get_ray(u, v, rng);
            if (pass == 0) {
              L = ray_color(w, r, rng, p->max_depth, false, cnt);
              wt = filter_weight(p->filter, jx - 0.5, jy - 0.5);
              continue;
            }
            double gx = g[0], gy = g[1], gz = g[2];
            double lum = 0.2126 * L.x + 0.7152 * L.y + 0.0722 * L.z;
            if (lum > 100.0) {
              double f = 100.0 / lum, sg = (L.x * gx + L.y * gy + L.z * gz) / lum;
              gx = f * (gx - 0.2126 * sg);
              gy = f * (gy - 0.7152 * sg);
              gz = f * (gz - 0.0722 * sg);
            }
            double sc = tw > 0.0 ? wt / tw : 1.0;
            ray_color_adj(w, r, rng, p->max_depth, false, cnt, Vec3(gx * sc, gy * sc, gz * sc), C);
          }
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; ++k) th.emplace_back(worker, k);
  for (auto& t : th) t.join();
  double em = 0.0;
  for (unsigned k = 0; k < nt; ++k) {
    for (size_t q = 0; q < n_tex; ++q)
      if (g_tex) g_tex[q] = (k ? g_tex[q] : 0.0) + ctx[k].g_tex[q];
    for (uint32_t m = 0; m < desc->n_materials; ++m)
      if (g_rough) g_rough[m] = (k ? g_rough[m] : 0.0) + ctx[k].g_rough[m];
    em += ctx[k].g_em;
  }
  if (g_em) {
    double inten = emission_material >= 0 ? desc->materials[emission_material].intensity : 0.0;
    *g_em = inten > 0.0 ? em / inten : 0.0;
  }
  return LMN_OK;
}


// ---- image outputs: AOV passes, post chain, Radiance pixels ----------------

namespace {

// Rust float semantics: f64::clamp keeps a NaN; `as u8` saturates, truncates
// toward zero and maps NaN to 0.
double clamp_nan(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
uint8_t as_u8(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 255.0) return 255;
  return uint8_t(x);
}
// The GPU's per-pixel AOV stream (counter mode).
uint64_t aov_key(uint64_t fkey, uint64_t pixel) {
  return splitmix_out(fkey ^ splitmix_out(pixel ^ 0x3C6EF372FE94F82BULL) ^ 0xA4093822299F31D0ULL);
}

double aces_tonemap(double x) {  // render.rs:3608-3616
  x = std::fmax(x, 0.0);
  double a = 2.51, b = 0.03, c = 2.43, d = 0.59, e = 0.14;
  return clamp_nan((x * (a * x + b)) / (x * (c * x + d) + e), 0.0, 1.0);
}
double reinhard_tonemap(double x) {  // render.rs:3619-3622
  x = std::fmax(x, 0.0);
  return x / (1.0 + x);
}
double uc2(double v) {  // render.rs:3629-3637
  double a = 0.15, b = 0.50, c = 0.10, d = 0.20, e = 0.02, f = 0.30;
  return ((v * (a * v + c * b) + d * e) / (v * (a * v + b) + d * f)) - e / f;
}
double filmic_tonemap(double x) {  // render.rs:3626-3641
  x = std::fmax(x, 0.0);
  double w = 11.2;
  return clamp_nan(uc2(x) / uc2(w), 0.0, 1.0);
}
double tone_fn(uint32_t mode, double x) {  // render.rs:2626-2632
  switch (mode) {
    case 0: return aces_tonemap(x);
    case 1: return reinhard_tonemap(x);
    case 2: return filmic_tonemap(x);
    default: return clamp_nan(x, 0.0, 1.0);
  }
}
uint8_t gamma_correct(double x) {  // render.rs:3649-3658, gamma 0, no dither
  x = clamp_nan(x, 0.0, 1.0);
  double s = x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
  return as_u8(clamp_nan(s + 0.0, 0.0, 0.999) * 256.0);
}

}  // namespace

// compute_log_average_luminance (render.rs:525-541), one sequential sum.
double oracle_log_average_luminance(const float* hdr, uint64_t n) {
  double sum = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    double x = hdr[3 * i], y = hdr[3 * i + 1], z = hdr[3 * i + 2];
    double lum = 0.2126 * x + 0.7152 * y + 0.0722 * z;
    sum += std::log(1e-4 + std::fmax(lum, 0.0));
  }
  if (n == 0) return 0.0;
  return std::exp(sum / double(n));
}

double oracle_tone(uint32_t mode, double x) { return tone_fn(mode, x); }
uint32_t oracle_srgb8(double x) { return gamma_correct(x); }

// Exposure (render.rs:2594-2608), hdr_data (:2610-2624), and the per-pixel
// loop of :3099-3202 with vignette 1, depth fog, no white balance / hue /
// saturation / grading / dither / brightness: tone curve, gamma_correct.
int oracle_post(const float* hdr, uint64_t n, double exposure, uint32_t tone_map, uint32_t auto_exposure,
                const float* fog_depth, double fog_density, const double fog_rgb[3], uint8_t* rgba,
                float* hdr_exposed, double* exposure_used) {
  if (!hdr || (fog_density > 0.0 && (!fog_depth || !fog_rgb))) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  double ex = exposure;
  if (auto_exposure) {
    double log_avg = oracle_log_average_luminance(hdr, n);
    double auto_exp = log_avg > 1e-6 ? 0.18 / log_avg : 1.0;
    ex = auto_exp * exposure;
  }
  if (exposure_used) *exposure_used = ex;
  for (uint64_t i = 0; i < n; ++i) {
    double vignette_factor = 1.0;
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = double(hdr[3 * i + k]) * ex * vignette_factor;
    if (hdr_exposed)
      for (int k = 0; k < 3; ++k) hdr_exposed[3 * i + k] = float(double(hdr[3 * i + k]) * ex);
    if (fog_density > 0.0) {
      double depth = double(fog_depth[i]);
      double fog_factor = 1.0 - std::exp(-fog_density * depth);
      for (int k = 0; k < 3; ++k) c[k] = c[k] * (1.0 - fog_factor) + fog_rgb[k] * fog_factor;
    }
    if (rgba) {
      for (int k = 0; k < 3; ++k) rgba[4 * i + k] = gamma_correct(tone_fn(tone_map, c[k]));
      rgba[4 * i + 3] = 255;
    }
  }
  return LMN_OK;
}

// write_radiance_hdr's pixel loop (main.rs:964-986): f32 arithmetic, log2f.
void oracle_rgbe(const float* rgb, uint64_t n, uint8_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    float r = std::fmax(rgb[i * 3], 0.0f), g = std::fmax(rgb[i * 3 + 1], 0.0f), b = std::fmax(rgb[i * 3 + 2], 0.0f);
    float max_val = std::fmax(std::fmax(r, g), b);
    uint8_t* o = out + 4 * i;
    if (max_val < 1e-32f) {
      o[0] = o[1] = o[2] = o[3] = 0;
      continue;
    }
    float l = std::ceil(std::log2(max_val));  // f32 log2, then ceil
    int32_t e = l >= 2147483647.0f ? 2147483647 : int32_t(l);  // `as i32` saturates
    float p2 = std::ldexp(1.0f, e > 1024 ? 1024 : e);  // 2.0_f32.powi(e): exact, inf above 2^127
    float scale = 256.0f / p2;
    o[0] = as_u8(double(std::fmin(r * scale, 255.0f)));
    o[1] = as_u8(double(std::fmin(g * scale, 255.0f)));
    o[2] = as_u8(double(std::fmin(b * scale, 255.0f)));
    o[3] = uint8_t(uint32_t(e) + 128u);  // `(exp + 128) as u8` wraps
  }
}

// AOV passes (render.rs:3508-3584); the same ray gives the fog depth of
// :2634-2665.  Rows in parallel.  depth is +inf where the ray escapes (the fog
// buffer); the reference's depth_pass holds 0 there.
int oracle_aov(const lmn_scene_desc* desc, const oracle_params* p, float* depth, uint8_t* normal, uint8_t* albedo) {
  if (!desc || !p) {
    g_err = "null argument";
    return LMN_ERR_INVALID;
  }
  World w;
  if (!build_world(*desc, p->exact_marched != 0, w)) return LMN_ERR_INVALID;
  uint32_t fw = std::max<uint32_t>(p->width, 1), fh = std::max<uint32_t>(p->height, 1);
  uint32_t cx = 0, cy = 0, cw = fw, ch = fh;
  if (p->has_crop) {
    cx = std::min(p->crop[0], fw);
    cy = std::min(p->crop[1], fh);
    cw = std::min(p->crop[2], fw);
    ch = std::min(p->crop[3], fh);
  }
  uint64_t fkey = frame_key(p->seed, p->scene_key);
  std::atomic<uint32_t> next_row{0};
  unsigned nt = p->threads ? p->threads : std::max(1u, std::thread::hardware_concurrency());
  auto worker = [&]() {
    for (;;) {
      uint32_t j = next_row.fetch_add(1);
      if (j >= ch) break;
      uint32_t gj = j + cy;
      double y = double(fh - 1 - gj);
      for (uint32_t i = 0; i < cw; ++i) {
        uint32_t gi = i + cx;
        double u = (double(gi) + 0.5) / double(fw - 1);
        double v = (y + 0.5) / double(fh - 1);
        uint64_t pix = uint64_t(gj) * fw + gi;
        CounterRng crng(aov_key(fkey, pix));
        XoshiroRng xrng(pix * p->seed);
        Rng& rng = p->rng_mode == ORACLE_RNG_REFERENCE ? static_cast<Rng&>(xrng) : static_cast<Rng&>(crng);
        Ray r = w.get_ray(u, v, rng);
        size_t o = size_t(j) * cw + i;
        float dep = INFINITY;
        uint8_t nb[3] = {0, 0, 0}, ab[3] = {0, 0, 0};
        HitRecord h;
        if (w.hit(r, 0.001, INFINITY, h)) {
          dep = float(h.t);
          double n3[3] = {h.n.x, h.n.y, h.n.z};
          for (int k = 0; k < 3; ++k) nb[k] = as_u8(clamp_nan((n3[k] * 0.5 + 0.5) * 255.0, 0.0, 255.0));
          if (albedo) {
            Ray sc;
            Vec3 att;
            if (h.mat->scatter(r, h, rng, sc, att)) {
              double a3[3] = {att.x, att.y, att.z};
              for (int k = 0; k < 3; ++k) ab[k] = as_u8(clamp_nan(a3[k], 0.0, 1.0) * 255.0);
            } else {
              Vec3 e = h.mat->emitted(h.u, h.v);
              double max_e = std::fmax(std::fmax(e.x, std::fmax(e.y, e.z)), 1.0);
              double e3[3] = {e.x, e.y, e.z};
              for (int k = 0; k < 3; ++k) ab[k] = as_u8(clamp_nan(e3[k] / max_e * 255.0, 0.0, 255.0));
            }
          }
        }
        if (depth) depth[o] = dep;
        for (int k = 0; k < 3; ++k) {
          if (normal) normal[3 * o + k] = nb[k];
          if (albedo) albedo[3 * o + k] = ab[k];
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; ++k) th.emplace_back(worker);
  for (auto& t : th) t.join();
  return LMN_OK;
}

}  // extern "C"

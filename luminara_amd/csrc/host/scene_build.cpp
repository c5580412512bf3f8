// Host-side scene flattening and SAH BVH build (see scene_build.h).
#include "scene_build.h"

#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>
#include <functional>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>

namespace lmn {

namespace {

void mat_identity(double m[12]) {
  std::memset(m, 0, 12 * sizeof(double));
  m[0] = m[5] = m[10] = 1.0;
}
// c = a * b for 3x4 affine matrices (b applied first)
void mat_mul(const double a[12], const double b[12], double c[12]) {
  double r[12];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 4; ++j) {
      double s = a[i * 4 + 0] * b[0 * 4 + j] + a[i * 4 + 1] * b[1 * 4 + j] + a[i * 4 + 2] * b[2 * 4 + j];
      if (j == 3) s += a[i * 4 + 3];
      r[i * 4 + j] = s;
    }
  }
  std::memcpy(c, r, sizeof r);
}
void mat3_mul(const double a[9], const double b[9], double c[9]) {
  double r[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] + a[i * 3 + 2] * b[2 * 3 + j];
  std::memcpy(c, r, sizeof r);
}

// Rust's f64::to_radians multiplies by PI/180.
double to_radians(double deg) { return deg * (M_PI / 180.0); }

// Scale::new / NonUniformScale::new replace a near-zero factor by 1
// (src/transform.rs:266-268, :294-297)
double fix_scale(double f) { return std::fabs(f) < 1e-10 ? 1.0 : f; }

}  // namespace

void compose_transform(const lmn_transform& t, XformT<double>& x, double fwd[12]) {
  double inv[12], nrm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  mat_identity(inv);
  mat_identity(fwd);
  x.tmode = 0;
  x.tscale = 1.0;
  x.renorm = 0;
  for (uint32_t k = 0; k < t.n_ops && k < LMN_MAX_XF_OPS; ++k) {
    const lmn_xform_op& op = t.ops[k];
    double F[12], I[12], N[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    mat_identity(F);
    mat_identity(I);
    switch (op.kind) {
      case LMN_XF_ROTATE_X: {  // src/transform.rs:148-176
        double r = to_radians(op.v[0]), s = std::sin(r), c = std::cos(r);
        F[5] = c; F[6] = -s; F[9] = s; F[10] = c;
        I[5] = c; I[6] = s; I[9] = -s; I[10] = c;
        N[4] = c; N[5] = -s; N[7] = s; N[8] = c;
        break;
      }
      case LMN_XF_ROTATE_Y: {  // src/transform.rs:72-101
        double r = to_radians(op.v[0]), s = std::sin(r), c = std::cos(r);
        F[0] = c; F[2] = s; F[8] = -s; F[10] = c;
        I[0] = c; I[2] = -s; I[8] = s; I[10] = c;
        N[0] = c; N[2] = s; N[6] = -s; N[8] = c;
        break;
      }
      case LMN_XF_ROTATE_Z: {  // src/transform.rs:224-251
        double r = to_radians(op.v[0]), s = std::sin(r), c = std::cos(r);
        F[0] = c; F[1] = -s; F[4] = s; F[5] = c;
        I[0] = c; I[1] = s; I[4] = -s; I[5] = c;
        N[0] = c; N[1] = -s; N[3] = s; N[4] = c;
        break;
      }
      case LMN_XF_TRANSLATE:  // src/transform.rs:18-24
        F[3] = op.v[0]; F[7] = op.v[1]; F[11] = op.v[2];
        I[3] = -op.v[0]; I[7] = -op.v[1]; I[11] = -op.v[2];
        break;
      case LMN_XF_SCALE: {  // src/transform.rs:272-279
        double f = fix_scale(op.v[0]), inv_f = 1.0 / f;
        F[0] = F[5] = F[10] = f;
        I[0] = I[5] = I[10] = inv_f;
        x.tmode = 1;
        x.tscale = f;
        break;
      }
      case LMN_XF_SCALE_XYZ: {  // src/transform.rs:306-338
        double s[3] = {fix_scale(op.v[0]), fix_scale(op.v[1]), fix_scale(op.v[2])};
        for (int a = 0; a < 3; ++a) {
          F[a * 5] = s[a];
          I[a * 5] = 1.0 / s[a];
          N[a * 4] = 1.0 / s[a];
        }
        x.tmode = 2;
        x.renorm = 1;
        break;
      }
      default: break;
    }
    mat_mul(F, fwd, fwd);   // forward: new op applied after (outside) the inner ones
    mat_mul(inv, I, inv);   // inverse: outer op undone first
    mat3_mul(N, nrm, nrm);
  }
  std::memcpy(x.inv, inv, sizeof inv);
  std::memcpy(x.nrm, nrm, sizeof nrm);
}

// Object-space bounds of a primitive (the reference's bounding_box()).
static void object_bounds(const lmn_prim& p, double lo[3], double hi[3]) {
  const double* q = p.p;
  switch (p.kind) {
    case LMN_PRIM_SPHERE:  // src/sphere.rs:64-67
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[3]; hi[k] = q[k] + q[3]; }
      break;
    case LMN_PRIM_QUAD: {  // src/quad.rs:64-72
      for (int k = 0; k < 3; ++k) {
        double c[4] = {q[k], q[k] + q[3 + k], q[k] + q[6 + k], q[k] + q[3 + k] + q[6 + k]};
        lo[k] = std::min(std::min(c[0], c[1]), std::min(c[2], c[3])) - 1e-4;
        hi[k] = std::max(std::max(c[0], c[1]), std::max(c[2], c[3])) + 1e-4;
      }
      break;
    }
    case LMN_PRIM_RECT_XY:  // src/rect.rs:39-44
      lo[0] = q[0]; hi[0] = q[1]; lo[1] = q[2]; hi[1] = q[3]; lo[2] = q[4] - 0.0001; hi[2] = q[4] + 0.0001;
      break;
    case LMN_PRIM_RECT_XZ:  // src/rect.rs:79-84
      lo[0] = q[0]; hi[0] = q[1]; lo[2] = q[2]; hi[2] = q[3]; lo[1] = q[4] - 0.0001; hi[1] = q[4] + 0.0001;
      break;
    case LMN_PRIM_RECT_YZ:  // src/rect.rs:119-124
      lo[1] = q[0]; hi[1] = q[1]; lo[2] = q[2]; hi[2] = q[3]; lo[0] = q[4] - 0.0001; hi[0] = q[4] + 0.0001;
      break;
    case LMN_PRIM_TRIANGLE:       // src/triangle.rs:66-70
    case LMN_PRIM_MESH_TRIANGLE:  // src/obj.rs:315-319
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(std::min(q[k], q[3 + k]), q[6 + k]) - 1e-4;
        hi[k] = std::max(std::max(q[k], q[3 + k]), q[6 + k]) + 1e-4;
      }
      break;
    case LMN_PRIM_TORUS: {  // src/torus.rs:105-111
      double r = q[3] + q[4];
      double e[3] = {r, q[4], r};
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - e[k]; hi[k] = q[k] + e[k]; }
      break;
    }
    case LMN_PRIM_SUPERELLIPSOID:  // src/superellipsoid.rs:92-97
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[3 + k]; hi[k] = q[k] + q[3 + k]; }
      break;
    case LMN_PRIM_SPRING: {  // src/spring.rs:107-114
      double r = q[3] + q[4], h = q[5] * q[6];
      lo[0] = q[0] - r; hi[0] = q[0] + r;
      lo[1] = q[1] - q[4]; hi[1] = q[1] + h + q[4];
      lo[2] = q[2] - r; hi[2] = q[2] + r;
      break;
    }
    case LMN_PRIM_DISK:  // src/disk.rs:61-65: centre +/- radius on every axis
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[6]; hi[k] = q[k] + q[6]; }
      break;
    case LMN_PRIM_CYLINDER:  // src/cylinder.rs:75-80
    case LMN_PRIM_CONE:      // src/cone.rs:80-85
      lo[0] = q[0] - q[3]; hi[0] = q[0] + q[3];
      lo[1] = q[4]; hi[1] = q[5];
      lo[2] = q[2] - q[3]; hi[2] = q[2] + q[3];
      break;
    case LMN_PRIM_ELLIPSOID:  // src/ellipsoid.rs:69-71: centre +/- radii
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[3 + k]; hi[k] = q[k] + q[3 + k]; }
      break;
    case LMN_PRIM_HEMISPHERE:  // src/hemisphere.rs:56-61: the upper half only
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[3]; hi[k] = q[k] + q[3]; }
      lo[1] = q[1];
      break;
    case LMN_PRIM_ANNULUS:  // src/annulus.rs:78-81: centre +/- outer radius
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - q[7]; hi[k] = q[k] + q[7]; }
      break;
    case LMN_PRIM_FACET:  // one triangle of a prism or pyramid: its vertices, padded
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(std::min(q[k], q[3 + k]), q[6 + k]) - 1e-4;
        hi[k] = std::max(std::max(q[k], q[3 + k]), q[6 + k]) + 1e-4;
      }
      break;
    case LMN_PRIM_ROUNDED_BOX:  // src/rounded_box.rs:81-87: centre +/- (half size + radius)
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - (q[3 + k] + q[6]); hi[k] = q[k] + (q[3 + k] + q[6]); }
      break;
    case LMN_PRIM_MOBIUS: {  // src/mobius.rs:113-120
      double e[3] = {q[3] + q[4] + q[5], q[4] + q[5], q[3] + q[4] + q[5]};
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - e[k]; hi[k] = q[k] + e[k]; }
      break;
    }
    case LMN_PRIM_WEDGE_FACE:  // src/wedge.rs:123-125: the wedge's box, padded
      for (int k = 0; k < 3; ++k) { lo[k] = q[k] - 1e-4; hi[k] = q[3 + k] + 1e-4; }
      break;
    default:
      for (int k = 0; k < 3; ++k) { lo[k] = -INFINITY; hi[k] = INFINITY; }
  }
}

void prim_world_bounds(const lmn_scene_desc& desc, const lmn_prim& p, double lo[3], double hi[3]) {
  object_bounds(p, lo, hi);
  if (p.transform < 0) return;
  XformT<double> x;
  double fwd[12];
  compose_transform(desc.transforms[p.transform], x, fwd);
  double nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    double v[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
    for (int k = 0; k < 3; ++k) {
      double w = fwd[k * 4 + 0] * v[0] + fwd[k * 4 + 1] * v[1] + fwd[k * 4 + 2] * v[2] + fwd[k * 4 + 3];
      nlo[k] = std::min(nlo[k], w);
      nhi[k] = std::max(nhi[k], w);
    }
  }
  // pad by a relative epsilon: transformed surfaces are evaluated in the
  // object frame, so the box must cover their rounding
  for (int k = 0; k < 3; ++k) {
    double pad = 1e-9 * (std::fabs(nlo[k]) + std::fabs(nhi[k]) + 1.0);
    lo[k] = nlo[k] - pad;
    hi[k] = nhi[k] + pad;
  }
}

namespace {

struct BRef {
  float c[3];
  uint32_t idx;
};

struct Builder {
  const std::vector<double>& blo;  // 3 per prim
  const std::vector<double>& bhi;
  std::vector<BRef> refs;
  std::vector<NodeD> nodes;
  std::vector<uint32_t> order;
  BvhOptions opt;
  uint32_t depth_max = 0, leaves = 0;

  // Subtrees reached at depth `split_depth` with at least `par_min`
  // primitives are deferred to `tasks` and built on worker threads.
  struct Task {
    uint32_t id;
    size_t b, e;
    uint32_t depth;
  };
  std::vector<Task>* tasks = nullptr;
  uint32_t split_depth = 0;
  size_t par_min = 0;

  Builder(const std::vector<double>& lo, const std::vector<double>& hi) : blo(lo), bhi(hi) {}

  void bounds_of(size_t b, size_t e, double lo[3], double hi[3], double clo[3], double chi[3]) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = clo[k] = INFINITY;
      hi[k] = chi[k] = -INFINITY;
    }
    for (size_t i = b; i < e; ++i) {
      uint32_t p = refs[i].idx;
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], blo[p * 3 + k]);
        hi[k] = std::max(hi[k], bhi[p * 3 + k]);
        clo[k] = std::min(clo[k], double(refs[i].c[k]));
        chi[k] = std::max(chi[k], double(refs[i].c[k]));
      }
    }
  }
  static double area(const double lo[3], const double hi[3]) {
    double d[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    for (double& x : d)
      if (!(x > 0)) x = 0;
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }

  // Builds the subtree of refs[b, e) into the already allocated node `id`.
  // Children are allocated as adjacent pairs at even indices (root at 0,
  // index 1 unused), so a pair of FP32 nodes is one aligned 64-byte block.
  void build(uint32_t id, size_t b, size_t e, uint32_t depth) {
    if (tasks && depth == split_depth && e - b >= par_min) {
      tasks->push_back({id, b, e, depth});
      return;
    }
    double lo[3], hi[3], clo[3], chi[3];
    bounds_of(b, e, lo, hi, clo, chi);
    for (int k = 0; k < 3; ++k) {
      nodes[id].lo[k] = lo[k];
      nodes[id].hi[k] = hi[k];
    }
    depth_max = std::max(depth_max, depth);
    size_t n = e - b;
    bool leaf = n <= 1 || depth >= uint32_t(opt.max_depth);
    int best_axis = -1;
    size_t mid = b;
    if (!leaf) {
      // binned SAH over centroid bounds (the reference bins 12 buckets per
      // axis, src/bvh.rs:71-163)
      const int NB = opt.bins;
      double sa = area(lo, hi);
      double best_cost = INFINITY;
      int best_split = -1;
      std::vector<double> blo_b(NB * 3), bhi_b(NB * 3);
      std::vector<size_t> cnt(NB);
      std::vector<double> rarea(NB);
      std::vector<size_t> rcount(NB);
      for (int ax = 0; ax < 3; ++ax) {
        double ext = chi[ax] - clo[ax];
        if (!(ext > 1e-12 * (std::fabs(chi[ax]) + std::fabs(clo[ax]) + 1e-30))) continue;
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int i = 0; i < NB * 3; ++i) {
          blo_b[i] = INFINITY;
          bhi_b[i] = -INFINITY;
        }
        double scale = NB / ext;
        for (size_t i = b; i < e; ++i) {
          int bi = int((refs[i].c[ax] - clo[ax]) * scale);
          bi = std::min(std::max(bi, 0), NB - 1);
          cnt[bi]++;
          uint32_t p = refs[i].idx;
          for (int k = 0; k < 3; ++k) {
            blo_b[bi * 3 + k] = std::min(blo_b[bi * 3 + k], blo[p * 3 + k]);
            bhi_b[bi * 3 + k] = std::max(bhi_b[bi * 3 + k], bhi[p * 3 + k]);
          }
        }
        double rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t rc = 0;
        for (int i = NB - 1; i > 0; --i) {
          rc += cnt[i];
          for (int k = 0; k < 3; ++k) {
            rl[k] = std::min(rl[k], blo_b[i * 3 + k]);
            rh[k] = std::max(rh[k], bhi_b[i * 3 + k]);
          }
          rarea[i] = area(rl, rh);
          rcount[i] = rc;
        }
        double ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t lc = 0;
        for (int i = 0; i < NB - 1; ++i) {
          lc += cnt[i];
          for (int k = 0; k < 3; ++k) {
            ll[k] = std::min(ll[k], blo_b[i * 3 + k]);
            lh[k] = std::max(lh[k], bhi_b[i * 3 + k]);
          }
          if (lc == 0 || rcount[i + 1] == 0) continue;
          double cost = 1.0 + (double(lc) * area(ll, lh) + double(rcount[i + 1]) * rarea[i + 1]) / std::max(sa, 1e-300);
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = ax;
            best_split = i;
          }
        }
      }
      if (best_axis >= 0 && n <= size_t(opt.max_leaf) && double(n) <= best_cost) leaf = true;
      if (!leaf) {
        if (best_axis >= 0) {
          double ext = chi[best_axis] - clo[best_axis];
          double scale = NB / ext;
          auto it = std::partition(refs.begin() + b, refs.begin() + e, [&](const BRef& r) {
            int bi = int((r.c[best_axis] - clo[best_axis]) * scale);
            bi = std::min(std::max(bi, 0), NB - 1);
            return bi <= best_split;
          });
          mid = size_t(it - refs.begin());
        }
        if (best_axis < 0 || mid == b || mid == e) {
          if (n <= size_t(opt.max_leaf)) {
            leaf = true;
          } else {  // coincident centroids: median split on the widest axis
            int ax = 0;
            for (int k = 1; k < 3; ++k)
              if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
            mid = b + n / 2;
            std::nth_element(refs.begin() + b, refs.begin() + mid, refs.begin() + e,
                             [&](const BRef& x, const BRef& y) { return x.c[ax] < y.c[ax]; });
            best_axis = ax;
          }
        }
      }
    }
    if (leaf) {
      nodes[id].a = uint32_t(order.size());
      nodes[id].b = LEAF_BIT | uint32_t(n);
      for (size_t i = b; i < e; ++i) order.push_back(refs[i].idx);
      leaves++;
      return;
    }
    uint32_t c = uint32_t(nodes.size());
    nodes.resize(c + 2);
    nodes[id].a = c;
    nodes[id].b = uint32_t(best_axis < 0 ? 0 : best_axis);
    build(c, b, mid, depth + 1);
    build(c + 1, mid, e, depth + 1);
  }

  // The same tree as build(0, 0, n, 0): the top levels on this thread, the
  // subtrees below them on worker threads (each in a builder of its own over
  // a copy of its references, splitting exactly as here), then merged in
  // depth-first order.  Node numbering differs from the one-thread build;
  // the tree does not.
  void build_parallel(size_t n) {
    unsigned nt = std::min(64u, std::max(1u, std::thread::hardware_concurrency()));
    if (const char* env = std::getenv("LMN_BUILD_THREADS")) nt = unsigned(std::max(1, std::atoi(env)));
    if (n < 65536 || nt < 2) {
      build(0, 0, n, 0);
      return;
    }
    std::vector<Task> T;
    tasks = &T;
    split_depth = 7;
    par_min = 1024;
    build(0, 0, n, 0);
    tasks = nullptr;
    std::vector<std::unique_ptr<Builder>> sub(T.size());
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t k = next.fetch_add(1);
        if (k >= T.size()) break;
        sub[k] = std::make_unique<Builder>(blo, bhi);
        Builder& S = *sub[k];
        S.opt = opt;
        S.refs.assign(refs.begin() + T[k].b, refs.begin() + T[k].e);
        S.nodes.resize(2);
        S.nodes[1].b = LEAF_BIT;
        S.build(0, 0, S.refs.size(), T[k].depth);
      }
    };
    std::vector<std::thread> th;
    for (unsigned k = 0; k < std::min<size_t>(nt, T.size()); ++k) th.emplace_back(worker);
    for (auto& t : th) t.join();
    for (size_t k = 0; k < T.size(); ++k) {
      const Builder& S = *sub[k];
      const uint32_t G = uint32_t(nodes.size()), O = uint32_t(order.size());
      auto moved = [&](NodeD x) {
        if (x.b & LEAF_BIT)
          x.a += O;
        else
          x.a = G + (x.a - 2);
        return x;
      };
      nodes[T[k].id] = moved(S.nodes[0]);
      for (size_t i = 2; i < S.nodes.size(); ++i) nodes.push_back(moved(S.nodes[i]));
      order.insert(order.end(), S.order.begin(), S.order.end());
      depth_max = std::max(depth_max, S.depth_max);
      leaves += S.leaves;
    }
  }
};

}  // namespace

// Collapses the FP32 binary BVH (sibling-pair layout) into a 4-wide one: a
// 4-wide node starts from a binary node's two children and keeps opening the
// child of largest surface area until it holds four, the usual top-down
// collapse of a binary SAH tree.  Leaves keep their primitive ranges; a leaf
// with more than LEAF4_MAX primitives (possible only at the depth cap) is
// split into entries of LEAF4_MAX under extra nodes with the leaf's box.
// Returns the worst-case traversal stack in 32-bit words: along any
// root-to-leaf path, every node can push all but one of its children, two
// words each.
uint32_t collapse_bvh4(const std::vector<NodeF>& n2, std::vector<Node4F>& n4) {
  n4.clear();
  if (n2.empty()) return 0;
  auto area = [&](uint32_t i) {
    const NodeF& n = n2[i];
    double dx = double(n.hi[0]) - n.lo[0], dy = double(n.hi[1]) - n.lo[1], dz = double(n.hi[2]) - n.lo[2];
    return dx * dy + dy * dz + dz * dx;
  };
  struct Entry {
    float lo[3], hi[3];
    uint32_t node2;   // binary node, or ~0u for a leaf chunk
    uint32_t first, count;
  };
  uint32_t worst = 0;
  // returns the index of the new 4-wide node; `pushed` = stack words above it
  std::function<uint32_t(std::vector<Entry>, uint32_t)> make = [&](std::vector<Entry> cand, uint32_t pushed) {
    uint32_t id = uint32_t(n4.size());
    n4.emplace_back();
    Node4F node{};
    node.n = uint32_t(cand.size());
    uint32_t here = pushed + 2 * (uint32_t(cand.size()) - 1);
    worst = std::max(worst, here);
    for (size_t k = 0; k < cand.size(); ++k) {
      const Entry& e = cand[k];
      node.lox[k] = e.lo[0];
      node.loy[k] = e.lo[1];
      node.loz[k] = e.lo[2];
      node.hix[k] = e.hi[0];
      node.hiy[k] = e.hi[1];
      node.hiz[k] = e.hi[2];
    }
    for (size_t k = 0; k < cand.size(); ++k) {
      const Entry& e = cand[k];
      if (e.node2 == ~0u || (n2[e.node2].b & LEAF_BIT)) {
        uint32_t first = e.node2 == ~0u ? e.first : n2[e.node2].a;
        uint32_t count = e.node2 == ~0u ? e.count : (n2[e.node2].b & ~LEAF_BIT);
        if (count <= LEAF4_MAX) {
          node.child[k] = LEAF4 | ((count - 1) << 28) | first;
          continue;
        }
        std::vector<Entry> parts;  // oversized leaf: chunks under extra nodes
        uint32_t per = (count + 3) / 4;
        for (uint32_t c = 0; c < count; c += per) {
          Entry p = e;
          p.node2 = ~0u;
          p.first = first + c;
          p.count = std::min(per, count - c);
          parts.push_back(p);
        }
        node.child[k] = make(parts, here);
        continue;
      }
      // internal: open the largest-area internal entry until four
      std::vector<Entry> sub;
      uint32_t c = n2[e.node2].a;
      for (uint32_t j : {c, c + 1}) {
        Entry x{};
        for (int a = 0; a < 3; ++a) {
          x.lo[a] = n2[j].lo[a];
          x.hi[a] = n2[j].hi[a];
        }
        x.node2 = j;
        sub.push_back(x);
      }
      while (sub.size() < 4) {
        int bi = -1;
        double ba = -1.0;
        for (size_t q = 0; q < sub.size(); ++q)
          if (sub[q].node2 != ~0u && !(n2[sub[q].node2].b & LEAF_BIT) && area(sub[q].node2) > ba) {
            ba = area(sub[q].node2);
            bi = int(q);
          }
        if (bi < 0) break;
        uint32_t open = sub[bi].node2, oc = n2[open].a;
        sub.erase(sub.begin() + bi);
        for (uint32_t j : {oc, oc + 1}) {
          Entry x{};
          for (int a = 0; a < 3; ++a) {
            x.lo[a] = n2[j].lo[a];
            x.hi[a] = n2[j].hi[a];
          }
          x.node2 = j;
          sub.push_back(x);
        }
      }
      node.child[k] = make(sub, here);
    }
    n4[id] = node;
    return id;
  };
  Entry root{};
  for (int a = 0; a < 3; ++a) {
    root.lo[a] = n2[0].lo[a];
    root.hi[a] = n2[0].hi[a];
  }
  root.node2 = 0;
  if (n2[0].b & LEAF_BIT) {
    make({root}, 0);
  } else {
    // the root's own box is not stored: start from its opened children
    std::vector<Entry> top;
    uint32_t c = n2[0].a;
    for (uint32_t j : {c, c + 1}) {
      Entry x{};
      for (int a = 0; a < 3; ++a) {
        x.lo[a] = n2[j].lo[a];
        x.hi[a] = n2[j].hi[a];
      }
      x.node2 = j;
      top.push_back(x);
    }
    while (top.size() < 4) {
      int bi = -1;
      double ba = -1.0;
      for (size_t q = 0; q < top.size(); ++q)
        if (!(n2[top[q].node2].b & LEAF_BIT) && area(top[q].node2) > ba) {
          ba = area(top[q].node2);
          bi = int(q);
        }
      if (bi < 0) break;
      uint32_t open = top[bi].node2, oc = n2[open].a;
      top.erase(top.begin() + bi);
      for (uint32_t j : {oc, oc + 1}) {
        Entry x{};
        for (int a = 0; a < 3; ++a) {
          x.lo[a] = n2[j].lo[a];
          x.hi[a] = n2[j].hi[a];
        }
        x.node2 = j;
        top.push_back(x);
      }
    }
    make(top, 0);
  }
  return worst;
}

// Quantises the 4-wide nodes' child boxes onto a per-node grid (Node4Q).
// Axis a of a node gets the origin org = the lowest child plane and the
// smallest power-of-two step s with 253 s >= the node's extent; a child's
// lower plane index is the largest q with org + q s below its FP32 plane,
// the upper one the smallest q with org + q s above it (strictly: an index
// landing exactly on the plane moves one step out, so the slab distance
// the kernel forms as fma(q, s / d, (org - o) / d) has a whole step of slack
// against its own rounding; only q = 0, the origin plane itself, is exact).
// The comparisons run in long double, exact for the FP32 differences here.
void quantize_bvh4(const std::vector<Node4F>& in, std::vector<Node4Q>& out) {
  out.assign(in.size(), Node4Q{});
  for (size_t i = 0; i < in.size(); ++i) {
    const Node4F& f = in[i];
    Node4Q& q = out[i];
    const uint32_t n = f.n;
    const float* lo[3] = {f.lox, f.loy, f.loz};
    const float* hi[3] = {f.hix, f.hiy, f.hiz};
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0}, meta = n << 24;
    for (int a = 0; a < 3; ++a) {
      float org = lo[a][0], top = hi[a][0];
      for (uint32_t k = 1; k < n; ++k) {
        org = std::min(org, lo[a][k]);
        top = std::max(top, hi[a][k]);
      }
      const long double ext = (long double)top - (long double)org;
      int e = -126;  // smallest normal step: a flat axis has every index 0
      if (ext > 0) {
        int x;
        std::frexp(ext / 253, &x);  // 2^(x-1) <= ext / 253 < 2^x
        e = std::min(std::max(x - 1, -126), 127);
        while (e > -126 && std::ldexp((long double)253, e - 1) >= ext) --e;
        while (e < 127 && std::ldexp((long double)253, e) < ext) ++e;
      }
      const long double step = std::ldexp((long double)1, e);
      q.org[a] = org;
      meta |= uint32_t(uint8_t(int8_t(e))) << (8 * a);
      for (uint32_t k = 0; k < n; ++k) {
        const long double dl = (long double)lo[a][k] - (long double)org;
        const long double dh = (long double)hi[a][k] - (long double)org;
        long double ql = std::floor(dl / step), qh = std::ceil(dh / step);
        ql = std::min(std::max(ql, (long double)0), (long double)255);
        qh = std::min(std::max(qh, (long double)0), (long double)255);
        while (ql > 0 && ql * step > dl) ql -= 1;
        if (ql > 0 && ql * step == dl) ql -= 1;
        while (qh < 255 && qh * step < dh) qh += 1;
        if (qh < 255 && qh * step == dh) qh += 1;
        qlo[a] |= uint32_t(ql) << (8 * k);
        qhi[a] |= uint32_t(qh) << (8 * k);
      }
    }
    q.qlo_x = qlo[0];
    q.qlo_y = qlo[1];
    q.qlo_z = qlo[2];
    q.qhi_x = qhi[0];
    q.qhi_y = qhi[1];
    q.qhi_z = qhi[2];
    q.child0 = f.child[0];
    q.child1 = f.child[1];
    q.child2 = f.child[2];
    q.child3 = f.child[3];
    q.meta = meta;
  }
}

namespace {

float round_down(double x) {
  float f = float(x);
  if (double(f) > x) f = std::nextafter(f, -INFINITY);
  return f;
}
float round_up(double x) {
  float f = float(x);
  if (double(f) < x) f = std::nextafter(f, INFINITY);
  return f;
}

template <class R>
void fill_record(const lmn_prim& p, const XformT<double>* x, PrimT<R>& r, uint32_t prim_id, bool fp64_island) {
  (void)x;
  std::memset(&r, 0, sizeof r);
  const double* q = p.p;
  double v[13] = {0};
  switch (p.kind) {
    case LMN_PRIM_QUAD: {  // Quad::new (src/quad.rs:20-26)
      double u[3] = {q[3], q[4], q[5]}, w[3] = {q[6], q[7], q[8]};
      double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
      double nl = std::sqrt(nn);
      double D = (n[0] / nl) * q[0] + (n[1] / nl) * q[1] + (n[2] / nl) * q[2];
      for (int k = 0; k < 9; ++k) v[k] = q[k];
      for (int k = 0; k < 3; ++k) v[9 + k] = n[k] / nn;
      v[12] = D;
      break;
    }
    case LMN_PRIM_TRIANGLE:
      for (int k = 0; k < 3; ++k) {
        v[k] = q[k];
        v[3 + k] = q[3 + k] - q[k];
        v[6 + k] = q[6 + k] - q[k];
      }
      break;
    case LMN_PRIM_PLANE: {  // Plane::new (src/plane.rs:18-39)
      double n[3] = {q[3], q[4], q[5]};
      bool nz = std::fabs(n[0]) < 1e-8 && std::fabs(n[1]) < 1e-8 && std::fabs(n[2]) < 1e-8;
      if (nz) {
        n[0] = 0; n[1] = 1; n[2] = 0;
      } else {
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (double& c : n) c = c / l;
      }
      double up[3] = {0, 1, 0};
      if (std::fabs(n[1]) > 0.999) { up[0] = 1; up[1] = 0; }
      double ua[3] = {up[1] * n[2] - up[2] * n[1], up[2] * n[0] - up[0] * n[2], up[0] * n[1] - up[1] * n[0]};
      double ul = std::sqrt(ua[0] * ua[0] + ua[1] * ua[1] + ua[2] * ua[2]);
      for (double& c : ua) c = c / ul;
      double va[3] = {n[1] * ua[2] - n[2] * ua[1], n[2] * ua[0] - n[0] * ua[2], n[0] * ua[1] - n[1] * ua[0]};
      for (int k = 0; k < 3; ++k) {
        v[k] = q[k];
        v[3 + k] = n[k];
        v[6 + k] = ua[k];
        v[9 + k] = va[k];
      }
      break;
    }
    case LMN_PRIM_DISK:       // Disk::new (src/disk.rs:16-28) and Annulus::new
    case LMN_PRIM_ANNULUS: {  // (src/annulus.rs:17-36): unit normal, (0, 1, 0) when near zero
      double n[3] = {q[3], q[4], q[5]};
      if (std::fabs(n[0]) < 1e-8 && std::fabs(n[1]) < 1e-8 && std::fabs(n[2]) < 1e-8) {
        n[0] = 0; n[1] = 1; n[2] = 0;
      } else {
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (double& c : n) c = c / l;
      }
      for (int k = 0; k < 3; ++k) {
        v[k] = q[k];
        v[3 + k] = n[k];
      }
      v[6] = q[6];
      v[7] = q[7];
      break;
    }
    case LMN_PRIM_FACET:  // v0, e1 = v1 - v0, e2 = v2 - v0 (src/prism.rs:104-105), normal
      for (int k = 0; k < 3; ++k) {
        v[k] = q[k];
        v[3 + k] = q[3 + k] - q[k];
        v[6 + k] = q[6 + k] - q[k];
        v[9 + k] = q[9 + k];
      }
      break;
    default:
      for (int k = 0; k < 12; ++k) v[k] = q[k];
  }
  for (int k = 0; k < 13; ++k) r.p[k] = R(v[k]);
  uint32_t kind = p.kind;
  uint32_t flags = fp64_island ? FLAG_FP64 : 0;
  r.kind_mat = kind | flags | (p.material << 8);
  r.xform = p.transform < 0 ? XF_NONE : uint32_t(p.transform);
  r.prim_id = prim_id;
}

// Which primitives the FP32 path evaluates in FP64: the marched kinds, and
// spheres large enough that FP32 cancellation in |o-c|^2 - r^2 exceeds the
// hit-distance tolerance (e.g. a r = 1000 ground sphere).
bool fp64_island(const lmn_prim& p) {
  if (p.kind == LMN_PRIM_TORUS || p.kind == LMN_PRIM_SUPERELLIPSOID || p.kind == LMN_PRIM_SPRING) return true;
  if (p.kind >= LMN_PRIM_DISK) return true;  // disk ... annulus: second BVH, FP64
  if (p.kind == LMN_PRIM_SPHERE) return std::fabs(p.p[3]) >= 128.0;
  return false;
}

template <class R>
void fill_material(const lmn_material& m, MaterialT<R>& o) {
  std::memset(&o, 0, sizeof o);
  o.kind = m.kind;
  o.tex = m.texture;
  for (int k = 0; k < 3; ++k) o.color[k] = R(m.color[k]);
  o.fuzz = R(m.fuzz);
  o.ior = R(m.ior);
  o.roughness = R(m.roughness);
  o.dispersion = R(m.dispersion);
  o.intensity = R(m.intensity);
  o.metallic = R(m.metallic);
  o.param_slot = ~0u;
}

template <class R>
void fill_light(const lmn_light& l, LightT<R>& o) {
  std::memset(&o, 0, sizeof o);
  o.kind = l.kind;
  o.mat = l.material;
  o.origin = {R(l.origin[0]), R(l.origin[1]), R(l.origin[2])};
  o.u = {R(l.u[0]), R(l.u[1]), R(l.u[2])};
  o.v = {R(l.v[0]), R(l.v[1]), R(l.v[2])};
  o.normal = {R(l.normal[0]), R(l.normal[1]), R(l.normal[2])};
  o.emission = {R(l.emission[0]), R(l.emission[1]), R(l.emission[2])};
  o.radius = R(l.radius);
  // u.cross(v).length() (src/render.rs:1419), evaluated in FP64
  double c[3] = {l.u[1] * l.v[2] - l.u[2] * l.v[1], l.u[2] * l.v[0] - l.u[0] * l.v[2], l.u[0] * l.v[1] - l.u[1] * l.v[0]};
  o.area = R(std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
}

template <class R>
void fill_bg(const lmn_background& b, BackgroundT<R>& o) {
  std::memset(&o, 0, sizeof o);
  o.kind = b.kind;
  o.c0 = {R(b.color0[0]), R(b.color0[1]), R(b.color0[2])};
  o.c1 = {R(b.color1[0]), R(b.color1[1]), R(b.color1[2])};
  double l = std::sqrt(b.direction[0] * b.direction[0] + b.direction[1] * b.direction[1] + b.direction[2] * b.direction[2]);
  o.dir_unit = {R(b.direction[0] / l), R(b.direction[1] / l), R(b.direction[2] / l)};
  o.intensity = R(b.intensity);
}

template <class R>
void fill_camera(const lmn_camera& c, CameraT<R>& o) {
  auto v = [](const double* a) { return V3<R>{R(a[0]), R(a[1]), R(a[2])}; };
  o.origin = v(c.origin);
  o.lower_left = v(c.lower_left);
  o.horizontal = v(c.horizontal);
  o.vertical = v(c.vertical);
  o.u = v(c.u);
  o.v = v(c.v);
  double rel[3] = {c.lower_left[0] - c.origin[0], c.lower_left[1] - c.origin[1], c.lower_left[2] - c.origin[2]};
  o.ll_rel = v(rel);
  o.lens_radius = R(c.lens_radius);
}

}  // namespace

namespace {
int build_scene_tables(const lmn_scene_desc& desc, const BvhOptions& opt, HostScene& out, std::string& err);
}  // namespace

// A MESH_TRIANGLE is a triangle to everything but the surface: the tables
// are built from a copy of the prims with those marked TRIANGLE, and their
// records then get FLAG_SMOOTH and the index of their attributes.
int build_host_scene(const lmn_scene_desc& desc, const BvhOptions& opt, HostScene& out, std::string& err) {
  std::vector<lmn_prim> prims(desc.prims, desc.prims + desc.n_prims);
  std::vector<int64_t> attr_of(desc.n_prims, -1);
  for (uint32_t i = 0; i < desc.n_prims; ++i) {
    if (prims[i].kind != LMN_PRIM_MESH_TRIANGLE) continue;
    const double a = prims[i].p[9];
    if (!(a >= 0.0) || a != std::floor(a) || a >= double(desc.n_vertex_attrs) || !desc.vertex_attrs ||
        a >= 4294967295.0) {
      err = "prim " + std::to_string(i) + ": vertex attribute index out of range";
      return LMN_ERR_INVALID;
    }
    attr_of[i] = int64_t(a);
    prims[i].kind = LMN_PRIM_TRIANGLE;
    prims[i].p[9] = 0.0;
  }
  lmn_scene_desc d = desc;
  d.prims = prims.data();
  int rc = build_scene_tables(d, opt, out, err);
  if (rc != LMN_OK) return rc;
  out.attrF.assign(std::max<uint64_t>(desc.n_vertex_attrs, 1), AttrT<float>{});
  out.attrD.assign(std::max<uint64_t>(desc.n_vertex_attrs, 1), AttrT<double>{});
  for (uint64_t i = 0; i < desc.n_vertex_attrs; ++i) {
    const lmn_vertex_attr& a = desc.vertex_attrs[i];
    const uint32_t flags = (a.has_normals ? 1u : 0u) | (a.has_uvs ? 2u : 0u);
    for (int k = 0; k < 9; ++k) {
      out.attrF[i].n[k] = float(a.normal[k / 3][k % 3]);
      out.attrD[i].n[k] = a.normal[k / 3][k % 3];
    }
    for (int k = 0; k < 6; ++k) {
      out.attrF[i].uv[k] = float(a.uv[k / 2][k % 2]);
      out.attrD[i].uv[k] = a.uv[k / 2][k % 2];
    }
    out.attrF[i].flags = out.attrD[i].flags = flags;
  }
  for (size_t r = 0; r < out.order.size(); ++r) {
    const int64_t a = attr_of[out.order[r]];
    if (a < 0) continue;
    out.primsF[r].kind_mat |= FLAG_SMOOTH;
    out.primsD[r].kind_mat |= FLAG_SMOOTH;
    set_attr_index(out.primsF[r], uint32_t(a));
    set_attr_index(out.primsD[r], uint32_t(a));
  }
  return LMN_OK;
}

namespace {
int build_scene_tables(const lmn_scene_desc& desc, const BvhOptions& opt, HostScene& out, std::string& err) {
  // validate references
  for (uint32_t i = 0; i < desc.n_prims; ++i) {
    const lmn_prim& p = desc.prims[i];
    if (p.kind >= LMN_PRIM_KIND_COUNT) {
      err = "prim " + std::to_string(i) + ": unknown kind";
      return LMN_ERR_INVALID;
    }
    if (p.material >= desc.n_materials) {
      err = "prim " + std::to_string(i) + ": material index out of range";
      return LMN_ERR_INVALID;
    }
    if (p.transform >= 0 && uint32_t(p.transform) >= desc.n_transforms) {
      err = "prim " + std::to_string(i) + ": transform index out of range";
      return LMN_ERR_INVALID;
    }
    if (p.material >= (1u << 24)) {
      err = "too many materials";
      return LMN_ERR_INVALID;
    }
  }
  for (uint32_t i = 0; i < desc.n_materials; ++i) {
    const lmn_material& m = desc.materials[i];
    if (m.kind > LMN_MAT_MICROFACET) {
      err = "material " + std::to_string(i) + ": unknown kind";
      return LMN_ERR_INVALID;
    }
    if (m.texture >= 0 && uint32_t(m.texture) >= desc.n_textures) {
      err = "material " + std::to_string(i) + ": texture index out of range";
      return LMN_ERR_INVALID;
    }
  }
  for (uint32_t i = 0; i < desc.n_textures; ++i) {
    if (desc.textures[i].width == 0 || desc.textures[i].height == 0 || !desc.textures[i].rgb) {
      err = "texture " + std::to_string(i) + ": empty";
      return LMN_ERR_INVALID;
    }
  }
  for (uint32_t i = 0; i < desc.n_lights; ++i)
    if (desc.lights[i].kind > LMN_LIGHT_DISK) {
      err = "light " + std::to_string(i) + ": unknown kind";
      return LMN_ERR_INVALID;
    }

  // main BVH over the analytic kinds, a second BVH over the marched kinds
  // (torus, superellipsoid, spring) and the FP64 quadrics (disk, cylinder,
  // cone), the unbounded planes last
  std::vector<uint32_t> bounded, marched, planes;
  for (uint32_t i = 0; i < desc.n_prims; ++i) {
    uint32_t k = desc.prims[i].kind;
    if (k == LMN_PRIM_PLANE)
      planes.push_back(i);
    else if (k >= LMN_PRIM_TORUS)  // marched kinds and the quadrics
      marched.push_back(i);
    else
      bounded.push_back(i);
  }
  if (desc.n_prims >= LEAF_BIT) {
    err = "too many primitives";
    return LMN_ERR_INVALID;
  }
  auto build_tree = [&](const std::vector<uint32_t>& ids, uint32_t rec_base, std::vector<NodeD>& nodesD,
                        std::vector<NodeF>& nodesF, int max_leaf) {
    std::vector<double> blo(ids.size() * 3), bhi(ids.size() * 3);
    Builder B(blo, bhi);
    B.opt = opt;
    B.opt.max_leaf = std::min(opt.max_leaf, max_leaf);
    B.refs.resize(ids.size());
    for (size_t i = 0; i < ids.size(); ++i) {
      prim_world_bounds(desc, desc.prims[ids[i]], &blo[i * 3], &bhi[i * 3]);
      for (int k = 0; k < 3; ++k) B.refs[i].c[k] = float(0.5 * (blo[i * 3 + k] + bhi[i * 3 + k]));
      B.refs[i].idx = uint32_t(i);
    }
    if (!ids.empty()) {
      B.nodes.resize(2);  // root + unused slot keeps child pairs 64-byte aligned
      B.nodes[1].b = LEAF_BIT;  // empty leaf, never referenced
      B.build_parallel(ids.size());
    }
    for (NodeD& n : B.nodes)
      if (n.b & LEAF_BIT) n.a += rec_base;
    out.tree_depth = std::max(out.tree_depth, B.depth_max);
    out.n_leaves += B.leaves;
    nodesD = B.nodes;
    nodesF.resize(B.nodes.size());
    for (size_t i = 0; i < B.nodes.size(); ++i) {
      const NodeD& d = B.nodes[i];
      NodeF& f = nodesF[i];
      for (int k = 0; k < 3; ++k) {
        // conservative FP32 boxes: outward rounding plus a few ulps for the
        // FP32 rounding of the primitive parameters themselves
        double pad = 4.0 * FLT_EPSILON * (std::fabs(d.lo[k]) + std::fabs(d.hi[k]));
        f.lo[k] = round_down(d.lo[k] - pad);
        f.hi[k] = round_up(d.hi[k] + pad);
      }
      f.a = d.a;
      f.b = d.b;
    }
    for (uint32_t r : B.order) out.order.push_back(ids[r]);
  };
  out.order.clear();
  out.tree_depth = 0;
  out.n_leaves = 0;
  build_tree(bounded, 0, out.nodesD, out.nodesF, opt.max_leaf);
  // the marched kinds' root finders cost far more than a box test: one
  // primitive per leaf, so a ray reaches only those whose own box it enters
  build_tree(marched, uint32_t(bounded.size()), out.mnodesD, out.mnodesF, 1);
  if (bounded.size() + marched.size() > LEAF4_FIRST) {
    err = "too many primitives for the 4-wide BVH";
    return LMN_ERR_INVALID;
  }
  {
    std::vector<Node4F> wide;
    out.stack4_words = collapse_bvh4(out.nodesF, wide);
    if (wide.size() > (size_t(1) << 26)) {  // the kernels form a node's byte offset in 32 bits
      err = "too many BVH nodes for the 4-wide tree (at most 2^26)";
      return LMN_ERR_INVALID;
    }
    quantize_bvh4(wide, out.nodes4);
  }
  out.n_bounded = uint32_t(bounded.size() + marched.size());
  out.n_planes = uint32_t(planes.size());
  for (uint32_t p : planes) out.order.push_back(p);
  size_t nrec = out.order.size();
  out.primsF.resize(nrec);
  out.primsD.resize(nrec);
  out.featF = out.featD = 0;
  out.islands_main = false;
  out.march_kinds = 0;
  for (size_t i = 0; i < nrec; ++i) {
    const lmn_prim& p = desc.prims[out.order[i]];
    bool isl = fp64_island(p);
    fill_record<float>(p, nullptr, out.primsF[i], out.order[i], isl);
    fill_record<double>(p, nullptr, out.primsD[i], out.order[i], false);
    if (p.transform >= 0) out.featF |= FEAT_XF, out.featD |= FEAT_XF;
    if (isl) out.featF |= FEAT_F64;
    if (isl && p.kind < LMN_PRIM_TORUS) out.islands_main = true;  // an FP64 record in the main BVH
    if (p.kind >= LMN_PRIM_TORUS) out.featF |= FEAT_MARCH, out.featD |= FEAT_MARCH;
    if (p.kind == LMN_PRIM_TORUS) out.march_kinds |= 1u;
    if (p.kind == LMN_PRIM_SUPERELLIPSOID) out.march_kinds |= 2u;
    if (p.kind == LMN_PRIM_SPRING) out.march_kinds |= 4u;
  }
  out.xfD.resize(desc.n_transforms);
  out.xfF.resize(desc.n_transforms);
  for (uint32_t i = 0; i < desc.n_transforms; ++i) {
    double fwd[12];
    compose_transform(desc.transforms[i], out.xfD[i], fwd);
    XformT<float>& f = out.xfF[i];
    for (int k = 0; k < 12; ++k) f.inv[k] = float(out.xfD[i].inv[k]);
    for (int k = 0; k < 9; ++k) f.nrm[k] = float(out.xfD[i].nrm[k]);
    f.tscale = float(out.xfD[i].tscale);
    f.tmode = out.xfD[i].tmode;
    f.renorm = out.xfD[i].renorm;
  }
  out.matF.resize(desc.n_materials);
  out.matD.resize(desc.n_materials);
  out.slot_mat.clear();
  for (uint32_t i = 0; i < desc.n_materials; ++i) {
    fill_material(desc.materials[i], out.matF[i]);
    fill_material(desc.materials[i], out.matD[i]);
    if (desc.materials[i].kind == LMN_MAT_MICROFACET) {  // roughness parameters of the adjoint
      out.matF[i].param_slot = out.matD[i].param_slot = uint32_t(out.slot_mat.size());
      out.slot_mat.push_back(i);
    }
  }
  out.tex.resize(desc.n_textures);
  out.texels.clear();
  for (uint32_t i = 0; i < desc.n_textures; ++i) {
    const lmn_texture& t = desc.textures[i];
    out.tex[i].w = t.width;
    out.tex[i].h = t.height;
    out.tex[i].offset = out.texels.size();
    out.texels.insert(out.texels.end(), t.rgb, t.rgb + size_t(t.width) * t.height * 3);
  }
  if (out.texels.empty()) out.texels.push_back(0.f);
  out.lightF.resize(desc.n_lights);
  out.lightD.resize(desc.n_lights);
  for (uint32_t i = 0; i < desc.n_lights; ++i) {
    fill_light(desc.lights[i], out.lightF[i]);
    fill_light(desc.lights[i], out.lightD[i]);
  }
  fill_bg(desc.config.background, out.bgF);
  fill_bg(desc.config.background, out.bgD);
  fill_camera(desc.camera, out.camF);
  fill_camera(desc.camera, out.camD);
  return LMN_OK;
}
}  // namespace

}  // namespace lmn

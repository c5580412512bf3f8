// TOML scene -> lmn_scene_desc.  Mirrors scene::load_scene
// (src/scene.rs:846-1645): [render] settings over RenderConfig::default()
// (src/render.rs:422-521), the camera built from the TOML width/height aspect
// (src/scene.rs:1146-1174, src/camera.rs:48-76), objects added grouped by type
// in the reference's fixed order with NEE lights registered alongside
// (src/scene.rs:1180-1642), materials from build_material (src/scene.rs:1709).
// Scene features that are valid in the reference but not part of the render
// hot path (procedural textures, exotic materials and primitives) are
// rejected with LMN_ERR_UNSUPPORTED naming the feature.
#include "scene_load.h"

#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

#include "mesh_io.h"
#include "toml.h"

namespace lmn {

void OwnedDesc::sync() {
  textures.resize(texels.size());
  for (size_t i = 0; i < texels.size(); ++i) textures[i].rgb = texels[i].data();
  desc.n_prims = uint32_t(prims.size());
  desc.n_transforms = uint32_t(transforms.size());
  desc.n_materials = uint32_t(materials.size());
  desc.n_textures = uint32_t(textures.size());
  desc.n_lights = uint32_t(lights.size());
  desc.n_objects = n_objects;
  desc.prims = prims.data();
  desc.transforms = transforms.data();
  desc.materials = materials.data();
  desc.textures = textures.data();
  desc.lights = lights.data();
  desc.n_vertex_attrs = vertex_attrs.size();
  desc.vertex_attrs = vertex_attrs.data();
}

namespace {

struct LoadError {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw LoadError{code, msg}; }

const char* kind_name(TomlValue::Kind k) {
  switch (k) {
    case TomlValue::STRING: return "string";
    case TomlValue::INTEGER: return "integer";
    case TomlValue::FLOAT: return "float";
    case TomlValue::BOOLEAN: return "boolean";
    case TomlValue::ARRAY: return "array";
    default: return "table";
  }
}

// serde-style typed accessors: a missing required field or a wrong type is a
// parse error, as with the reference's derive(Deserialize) structs.
double num(const TomlValue& v, const std::string& ctx) {
  if (!v.is_number()) fail(LMN_ERR_PARSE, ctx + ": invalid type " + kind_name(v.kind) + ", expected f64");
  return v.as_double();
}
uint64_t uint_field(const TomlValue& v, const std::string& ctx, uint64_t max) {
  if (v.kind != TomlValue::INTEGER) fail(LMN_ERR_PARSE, ctx + ": invalid type " + kind_name(v.kind) + ", expected integer");
  if (v.integer < 0 || uint64_t(v.integer) > max) fail(LMN_ERR_PARSE, ctx + ": integer out of range");
  return uint64_t(v.integer);
}
void vec3(const TomlValue& v, const std::string& ctx, double out[3]) {
  if (v.kind != TomlValue::ARRAY) fail(LMN_ERR_PARSE, ctx + ": expected an array of 3 numbers");
  if (v.array.size() != 3) fail(LMN_ERR_PARSE, ctx + ": invalid length " + std::to_string(v.array.size()) + ", expected 3");
  for (int k = 0; k < 3; ++k) out[k] = num(v.array[k], ctx);
}
void vec2(const TomlValue& v, const std::string& ctx, double out[2]) {
  if (v.kind != TomlValue::ARRAY || v.array.size() != 2) fail(LMN_ERR_PARSE, ctx + ": expected an array of 2 numbers");
  for (int k = 0; k < 2; ++k) out[k] = num(v.array[k], ctx);
}
const TomlValue& req(const TomlValue& t, const std::string& key, const std::string& ctx) {
  const TomlValue* v = t.get(key);
  if (!v) fail(LMN_ERR_PARSE, ctx + ": missing field `" + key + "`");
  return *v;
}
bool opt_num(const TomlValue& t, const std::string& key, const std::string& ctx, double& out) {
  const TomlValue* v = t.get(key);
  if (!v) return false;
  out = num(*v, ctx + "." + key);
  return true;
}
bool opt_vec3(const TomlValue& t, const std::string& key, const std::string& ctx, double out[3]) {
  const TomlValue* v = t.get(key);
  if (!v) return false;
  vec3(*v, ctx + "." + key, out);
  return true;
}
double req_num(const TomlValue& t, const std::string& key, const std::string& ctx) {
  return num(req(t, key, ctx), ctx + "." + key);
}
void req_vec3(const TomlValue& t, const std::string& key, const std::string& ctx, double out[3]) {
  vec3(req(t, key, ctx), ctx + "." + key, out);
}
std::string opt_str(const TomlValue& t, const std::string& key, const std::string& ctx, bool& present) {
  const TomlValue* v = t.get(key);
  present = v != nullptr;
  if (!v) return "";
  if (v->kind != TomlValue::STRING) fail(LMN_ERR_PARSE, ctx + "." + key + ": expected a string");
  return v->str;
}

const std::vector<TomlValue>& table_array(const TomlValue& root, const std::string& key) {
  static const std::vector<TomlValue> empty;
  const TomlValue* v = root.get(key);
  if (!v) return empty;
  if (v->kind != TomlValue::ARRAY) fail(LMN_ERR_PARSE, "`" + key + "`: expected an array of tables");
  for (auto& e : v->array)
    if (e.kind != TomlValue::TABLE) fail(LMN_ERR_PARSE, "`" + key + "`: expected an array of tables");
  return v->array;
}

// Tanner Helland's blackbody approximation, as the reference's
// blackbody_to_rgb (src/scene.rs:1677-1707) evaluates it.
void blackbody_rgb(double temp_k, double out[3]) {
  double temp = std::fmin(std::fmax(temp_k, 1000.0), 40000.0) / 100.0;
  auto clamp = [](double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); };
  double r, g, b;
  if (temp <= 66.0) {
    r = 1.0;
  } else {
    double t = temp - 60.0;
    r = clamp(329.698727446 * std::pow(t, -0.1332047592) / 255.0, 0.0, 1.0);
  }
  if (temp <= 66.0) {
    g = clamp(99.4708025861 * std::log(temp) - 161.1195681661, 0.0, 255.0) / 255.0;
  } else {
    double t = temp - 60.0;
    g = clamp(288.1221695283 * std::pow(t, -0.0755148492) / 255.0, 0.0, 1.0);
  }
  if (temp >= 66.0) {
    b = 1.0;
  } else if (temp <= 19.0) {
    b = 0.0;
  } else {
    double t = temp - 10.0;
    b = clamp(138.5177312231 * std::log(t) - 305.0447927307, 0.0, 255.0) / 255.0;
  }
  out[0] = r;
  out[1] = g;
  out[2] = b;
}

// Material tag names: serde accepts the variant identifier and its aliases
// (src/scene.rs:512-772).
std::string material_tag(const TomlValue& m, const std::string& ctx) {
  if (m.kind != TomlValue::TABLE) fail(LMN_ERR_PARSE, ctx + ": expected a material table");
  const TomlValue* t = m.get("type");
  if (!t) fail(LMN_ERR_PARSE, ctx + ": missing field `type`");
  if (t->kind != TomlValue::STRING) fail(LMN_ERR_PARSE, ctx + ".type: expected a string");
  return t->str;
}

struct MatInfo {
  bool emissive = false;  // Emissive or Blackbody: a light for NEE where registered
  double emission[3] = {0, 0, 0};
};

class Loader {
 public:
  Loader(const std::string& base_dir, OwnedDesc& out) : base_(base_dir), out_(out) {}

  void run(const std::string& text) {
    TomlValue root;
    std::string err;
    if (!toml_parse(text, root, err)) fail(LMN_ERR_PARSE, "TOML parse error: " + err);
    if (root.kind != TomlValue::TABLE) fail(LMN_ERR_PARSE, "TOML parse error: expected a table");
    render_settings(root);
    camera(root);
    objects(root);
    out_.sync();
  }

 private:
  std::string base_;
  OwnedDesc& out_;

  void render_settings(const TomlValue& root) {
    lmn_render_config& c = out_.desc.config;
    // RenderConfig::default() (src/render.rs:422-521)
    c.width = 800;
    c.height = 450;
    c.samples = 100;
    c.max_depth = 50;
    c.seed = 31337;
    c.filter = LMN_FILTER_BOX;
    c.has_crop = 0;
    c.exposure = 1.0;
    c.tone_map = 0;
    c.auto_exposure = 0;
    c.adaptive = 0;
    c.pad = 0;
    c.adaptive_threshold = 0.03;
    c.background = lmn_background{};
    c.background.kind = LMN_BG_SKY;
    const TomlValue* r = root.get("render");
    if (!r) return;
    if (r->kind != TomlValue::TABLE) fail(LMN_ERR_PARSE, "render: expected a table");
    const std::string ctx = "render";
    if (auto v = r->get("width")) c.width = uint32_t(uint_field(*v, ctx + ".width", 0xffffffffu));
    if (auto v = r->get("height")) c.height = uint32_t(uint_field(*v, ctx + ".height", 0xffffffffu));
    if (auto v = r->get("samples")) c.samples = uint32_t(uint_field(*v, ctx + ".samples", 0xffffffffu));
    if (auto v = r->get("max_depth")) c.max_depth = uint32_t(uint_field(*v, ctx + ".max_depth", 0xffffffffu));
    if (auto v = r->get("seed")) c.seed = uint_field(*v, ctx + ".seed", ~0ull);
    opt_num(*r, "exposure", ctx, c.exposure);
    if (const TomlValue* ae = r->get("auto_exposure")) {  // src/scene.rs:870-872
      if (ae->kind != TomlValue::BOOLEAN) fail(LMN_ERR_PARSE, "render.auto_exposure: expected a boolean");
      c.auto_exposure = ae->boolean ? 1u : 0u;
    }
    bool present = false;
    std::string pf = opt_str(*r, "pixel_filter", ctx, present);
    if (present) {  // src/scene.rs:906-913
      if (pf == "triangle" || pf == "tent")
        c.filter = LMN_FILTER_TRIANGLE;
      else if (pf == "gaussian" || pf == "gauss")
        c.filter = LMN_FILTER_GAUSSIAN;
      else if (pf == "mitchell")
        c.filter = LMN_FILTER_MITCHELL;
      else
        c.filter = LMN_FILTER_BOX;
    }
    std::string tm = opt_str(*r, "tone_map", ctx, present);
    if (present) {  // src/scene.rs:1100-1107
      if (tm == "reinhard")
        c.tone_map = 1;
      else if (tm == "filmic" || tm == "uncharted2")
        c.tone_map = 2;
      else if (tm == "none")
        c.tone_map = 3;
      else
        c.tone_map = 0;
    }
    if (const TomlValue* ad = r->get("adaptive")) {  // src/scene.rs:914-916
      if (ad->kind != TomlValue::BOOLEAN) fail(LMN_ERR_PARSE, "render.adaptive: expected a boolean");
      c.adaptive = ad->boolean ? 1u : 0u;
    }
    opt_num(*r, "adaptive_threshold", ctx, c.adaptive_threshold);  // :917-919
    if (const TomlValue* bg = r->get("background")) background(*bg, c.background);
  }

  void background(const TomlValue& bg, lmn_background& b) {
    // src/scene.rs:1108-1143
    const std::string ctx = "render.background";
    std::string tag = material_tag(bg, ctx);
    b = lmn_background{};
    if (tag == "Sky" || tag == "sky") {
      b.kind = LMN_BG_SKY;
    } else if (tag == "Solid" || tag == "solid") {
      b.kind = LMN_BG_SOLID;
      req_vec3(bg, "color", ctx, b.color0);
    } else if (tag == "Gradient" || tag == "gradient") {
      b.kind = LMN_BG_GRADIENT;
      req_vec3(bg, "bottom", ctx, b.color0);
      req_vec3(bg, "top", ctx, b.color1);
    } else if (tag == "Black" || tag == "black") {
      b.kind = LMN_BG_SOLID;
    } else if (tag == "Sunset" || tag == "sunset") {
      b.kind = LMN_BG_SUN;
      const double d[3] = {0.3, 0.15, -1.0}, sc[3] = {1.0, 0.6, 0.2}, sky[3] = {0.8, 0.4, 0.2};
      std::memcpy(b.direction, d, sizeof d);
      std::memcpy(b.color0, sc, sizeof sc);
      std::memcpy(b.color1, sky, sizeof sky);
      b.intensity = 30.0;
    } else if (tag == "Sun" || tag == "sun") {
      b.kind = LMN_BG_SUN;
      req_vec3(bg, "direction", ctx, b.direction);
      const double sc[3] = {1.0, 0.95, 0.85}, sky[3] = {0.5, 0.7, 1.0};
      if (!opt_vec3(bg, "sun_color", ctx, b.color0)) std::memcpy(b.color0, sc, sizeof sc);
      b.intensity = 20.0;
      opt_num(bg, "intensity", ctx, b.intensity);
      if (!opt_vec3(bg, "sky_color", ctx, b.color1)) std::memcpy(b.color1, sky, sizeof sky);
    } else if (tag == "Starfield" || tag == "starfield" || tag == "EnvMap" || tag == "env_map") {
      fail(LMN_ERR_UNSUPPORTED, "background type `" + tag + "` is not on the GPU path");
    } else {
      fail(LMN_ERR_PARSE, ctx + ": unknown variant `" + tag + "`");
    }
  }

  void camera(const TomlValue& root) {
    // CameraConfig::default() with aspect from the TOML width/height
    // (src/camera.rs:32-45, src/scene.rs:1147-1150)
    double look_from[3] = {0, 1, 3}, look_at[3] = {0, 0, 0}, vup[3] = {0, 1, 0};
    double vfov = 40.0, aperture = 0.0, focus = 1.0;
    double aspect = double(out_.desc.config.width) / double(out_.desc.config.height);
    if (const TomlValue* c = root.get("camera")) {
      if (c->kind != TomlValue::TABLE) fail(LMN_ERR_PARSE, "camera: expected a table");
      opt_vec3(*c, "look_from", "camera", look_from);
      opt_vec3(*c, "look_at", "camera", look_at);
      opt_vec3(*c, "vup", "camera", vup);
      opt_num(*c, "vfov", "camera", vfov);
      opt_num(*c, "aperture", "camera", aperture);
      opt_num(*c, "focus_dist", "camera", focus);
      if (const TomlValue* p = c->get("panorama")) {
        if (p->kind != TomlValue::BOOLEAN) fail(LMN_ERR_PARSE, "camera.panorama: expected a boolean");
        if (p->boolean) fail(LMN_ERR_UNSUPPORTED, "camera.panorama is not on the GPU path");
      }
    }
    // Camera::new (src/camera.rs:48-76)
    const double rads_per_deg = M_PI / 180.0;
    double theta = vfov * rads_per_deg;
    double h = std::tan(theta / 2.0);
    double vh = 2.0 * h, vw = aspect * vh;
    double w[3], u[3], v[3];
    for (int k = 0; k < 3; ++k) w[k] = look_from[k] - look_at[k];
    double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int k = 0; k < 3; ++k) w[k] = w[k] / wl;
    double cx = vup[1] * w[2] - vup[2] * w[1], cy = vup[2] * w[0] - vup[0] * w[2], cz = vup[0] * w[1] - vup[1] * w[0];
    double cl = std::sqrt(cx * cx + cy * cy + cz * cz);
    u[0] = cx / cl;
    u[1] = cy / cl;
    u[2] = cz / cl;
    v[0] = w[1] * u[2] - w[2] * u[1];
    v[1] = w[2] * u[0] - w[0] * u[2];
    v[2] = w[0] * u[1] - w[1] * u[0];
    lmn_camera& cam = out_.desc.camera;
    for (int k = 0; k < 3; ++k) {
      cam.origin[k] = look_from[k];
      cam.horizontal[k] = u[k] * vw * focus;
      cam.vertical[k] = v[k] * vh * focus;
      cam.lower_left[k] = look_from[k] - cam.horizontal[k] / 2.0 - cam.vertical[k] / 2.0 - w[k] * focus;
      cam.u[k] = u[k];
      cam.v[k] = v[k];
    }
    cam.lens_radius = aperture / 2.0;
  }

  // build_material (src/scene.rs:1709-2048) for the hot-path subset.
  uint32_t material(const TomlValue& m, const std::string& ctx, MatInfo* info = nullptr) {
    std::string tag = material_tag(m, ctx);
    lmn_material mat{};
    mat.texture = -1;
    mat.intensity = 1.0;
    mat.ior = 1.5;
    MatInfo mi;
    if (tag == "Lambertian" || tag == "lambertian" || tag == "matte") {
      mat.kind = LMN_MAT_LAMBERTIAN;
      req_vec3(m, "color", ctx, mat.color);
    } else if (tag == "Metal" || tag == "metal") {
      mat.kind = LMN_MAT_METAL;
      req_vec3(m, "color", ctx, mat.color);
      double f = 0.0;
      opt_num(m, "fuzz", ctx, f);
      mat.fuzz = std::fmin(f, 1.0);  // Metal::new clamps fuzz to <= 1 (src/material.rs:81-86)
    } else if (tag == "Dielectric" || tag == "dielectric") {
      mat.kind = LMN_MAT_DIELECTRIC;
      mat.ior = req_num(m, "refraction_index", ctx);
      mat.color[0] = mat.color[1] = mat.color[2] = 1.0;
      opt_vec3(m, "tint", ctx, mat.color);
      opt_num(m, "roughness", ctx, mat.roughness);
      opt_num(m, "dispersion", ctx, mat.dispersion);
      if (!(mat.dispersion > 0.0)) mat.dispersion = 0.0;  // only d > 0 builds a dispersive glass
    } else if (tag == "Emissive" || tag == "emissive") {
      mat.kind = LMN_MAT_EMISSIVE;
      req_vec3(m, "color", ctx, mat.color);
      opt_num(m, "intensity", ctx, mat.intensity);
      bool has_tex = false;
      std::string file = opt_str(m, "texture", ctx, has_tex);
      if (has_tex) texture_file(file, ctx, mat);
      mi.emissive = true;
      for (int k = 0; k < 3; ++k) mi.emission[k] = mat.color[k] * mat.intensity;
    } else if (tag == "Blackbody" || tag == "blackbody") {
      mat.kind = LMN_MAT_EMISSIVE;
      blackbody_rgb(req_num(m, "temperature", ctx), mat.color);
      opt_num(m, "intensity", ctx, mat.intensity);
      mi.emissive = true;
      for (int k = 0; k < 3; ++k) mi.emission[k] = mat.color[k] * mat.intensity;
    } else if (tag == "Microfacet" || tag == "microfacet" || tag == "pbr") {
      mat.kind = LMN_MAT_MICROFACET;
      req_vec3(m, "color", ctx, mat.color);
      double r = 0.5, mt = 0.0;
      opt_num(m, "roughness", ctx, r);
      opt_num(m, "metallic", ctx, mt);
      // Microfacet::new clamps (src/material.rs:255-261)
      mat.roughness = std::fmin(std::fmax(r, 0.01), 1.0);
      mat.metallic = std::fmin(std::fmax(mt, 0.0), 1.0);
      bool has_tex = false;
      std::string file = opt_str(m, "texture", ctx, has_tex);
      if (has_tex) texture_file(file, ctx, mat);
    } else if (tag == "Mirror" || tag == "mirror") {
      mat.kind = LMN_MAT_METAL;
      mat.color[0] = mat.color[1] = mat.color[2] = 0.95;
      mat.fuzz = 0.0;
    } else if (tag == "Glass" || tag == "glass") {
      mat.kind = LMN_MAT_DIELECTRIC;
      mat.ior = 1.5;
      mat.color[0] = mat.color[1] = mat.color[2] = 1.0;
    } else {
      static const char* known[] = {
          "Checker", "checker", "Marble", "marble", "Turbulence", "turbulence", "Image", "image", "Wood", "wood",
          "Rings", "rings", "Toon", "toon", "cel", "Anisotropic", "anisotropic", "brushed", "Clearcoat",
          "clearcoat", "Velvet", "velvet", "Iridescent", "iridescent", "Translucent", "translucent",
          "Subsurface", "subsurface", "sss", "BlendMat", "blend", "Opacity", "opacity", "alpha", "Dots", "dots",
          "Grid", "grid", "UvChecker", "uv_checker", "GradientTex", "gradient_tex", "Stripe", "stripe",
          "Voronoi", "voronoi", "Noise", "noise", "Spiral", "spiral", "Hexgrid", "hexgrid", "Plasma", "plasma",
          "Terrain", "terrain", "earth", "Rust", "rust", "patina", "oxidized", "Brick", "brick", "bricks", "Camo",
          "camo", "camouflage", "Lava", "lava", "magma", "Cloud", "cloud", "clouds", "TriPlanar", "tri_planar",
          "triplanar", "MixColor", "mix_color", "Wavy", "wavy", "wave", "Fbm", "fbm", "fractal", "ColorRamp",
          "color_ramp", "ramp"};
      for (const char* k : known)
        if (tag == k) fail(LMN_ERR_UNSUPPORTED, ctx + ": material `" + tag + "` is not on the GPU path");
      fail(LMN_ERR_PARSE, ctx + ": unknown variant `" + tag + "`");
    }
    if (info) *info = mi;
    out_.materials.push_back(mat);
    return uint32_t(out_.materials.size() - 1);
  }

  // Emissive/PBR `texture = "file"`: the reference falls back to the solid
  // colour with a warning when the image cannot be loaded
  // (src/scene.rs:1732-1739, :1750-1757); decoding image files is not part of
  // this library, so only the fallback case is accepted.
  void texture_file(const std::string& file, const std::string& ctx, lmn_material&) {
    std::ifstream f(resolve(file));
    if (f.good()) fail(LMN_ERR_UNSUPPORTED, ctx + ": image texture files are not decoded by this library ('" + file + "')");
    fprintf(stderr, "Warning: failed to load texture '%s', using solid color\n", file.c_str());
  }

  std::string resolve(const std::string& path) const {
    if (base_.empty() || (!path.empty() && path[0] == '/')) return path;
    return base_ + "/" + path;
  }

  uint32_t add_transform(const std::vector<lmn_xform_op>& ops) {
    lmn_transform t{};
    t.n_ops = uint32_t(ops.size());
    for (size_t k = 0; k < ops.size(); ++k) t.ops[k] = ops[k];
    out_.transforms.push_back(t);
    return uint32_t(out_.transforms.size() - 1);
  }

  lmn_prim& add_prim(uint32_t kind, uint32_t mat, int32_t xf, uint32_t object) {
    lmn_prim p{};
    p.kind = kind;
    p.material = mat;
    p.transform = xf;
    p.object = object;
    out_.prims.push_back(p);
    return out_.prims.back();
  }
  uint32_t new_object() { return out_.n_objects++; }

  static lmn_xform_op op(uint32_t kind, double a, double b = 0, double c = 0) {
    lmn_xform_op o{};
    o.kind = kind;
    o.v[0] = a;
    o.v[1] = b;
    o.v[2] = c;
    return o;
  }

  void push_light_sphere(const double c[3], double r, const double e[3], uint32_t mat) {
    lmn_light l{};
    l.kind = LMN_LIGHT_SPHERE;
    l.material = mat;
    for (int k = 0; k < 3; ++k) {
      l.origin[k] = c[k];
      l.emission[k] = e[k];
    }
    l.radius = r;
    out_.lights.push_back(l);
  }

  void reject_if_present(const TomlValue& root, const char* key, const char* why) {
    if (!table_array(root, key).empty())
      fail(LMN_ERR_UNSUPPORTED, std::string("[[") + key + "]] " + why + " is not on the GPU path");
  }

  void objects(const TomlValue& root) {
    // Fixed type order of load_scene (src/scene.rs:1180-1642).
    for (auto& s : table_array(root, "sphere")) {
      const std::string ctx = "sphere";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double r = req_num(s, "radius", ctx);
      MatInfo mi;
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material", &mi);
      if (mi.emissive) push_light_sphere(c, r, mi.emission, mat);  // src/scene.rs:1182-1202
      if (s.get("bump_strength") || s.get("normal_map"))
        fail(LMN_ERR_UNSUPPORTED, "sphere bump/normal maps are not on the GPU path");
      std::vector<lmn_xform_op> ops;  // src/scene.rs:1216-1230
      double a;
      if (opt_num(s, "rotate_x", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_X, a));
      if (opt_num(s, "rotate_y", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Y, a));
      if (opt_num(s, "rotate_z", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Z, a));
      double t[3];
      if (opt_vec3(s, "translate", ctx, t)) ops.push_back(op(LMN_XF_TRANSLATE, t[0], t[1], t[2]));
      if (opt_num(s, "scale", ctx, a)) ops.push_back(op(LMN_XF_SCALE, a));
      int32_t xf = ops.empty() ? -1 : int32_t(add_transform(ops));
      lmn_prim& p = add_prim(LMN_PRIM_SPHERE, mat, xf, new_object());
      p.p[0] = c[0];
      p.p[1] = c[1];
      p.p[2] = c[2];
      p.p[3] = r;
    }
    reject_if_present(root, "moving_sphere", "(motion blur)");
    for (auto& s : table_array(root, "plane")) {
      const std::string ctx = "plane";
      double pt[3], n[3];
      req_vec3(s, "point", ctx, pt);
      req_vec3(s, "normal", ctx, n);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_PLANE, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = pt[k];
        p.p[3 + k] = n[k];
      }
    }
    for (auto& s : table_array(root, "torus")) {
      const std::string ctx = "torus";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double R = req_num(s, "major_radius", ctx), r = req_num(s, "minor_radius", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      if (s.get("bump_strength")) fail(LMN_ERR_UNSUPPORTED, "torus bump maps are not on the GPU path");
      lmn_prim& p = add_prim(LMN_PRIM_TORUS, mat, -1, new_object());
      p.p[0] = c[0];
      p.p[1] = c[1];
      p.p[2] = c[2];
      p.p[3] = R;
      p.p[4] = r;
    }
    for (auto& s : table_array(root, "rounded_box")) {  // src/scene.rs:1264-1272
      const std::string ctx = "rounded_box";
      double c[3], hs[3];
      req_vec3(s, "center", ctx, c);
      req_vec3(s, "half_size", ctx, hs);
      double r = req_num(s, "radius", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_ROUNDED_BOX, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = c[k];
        p.p[3 + k] = hs[k];
      }
      p.p[6] = r;
    }
    for (auto& s : table_array(root, "mobius")) {  // src/scene.rs:1274-1283 (thickness default 0.05)
      const std::string ctx = "mobius";
      double c[3], th = 0.05;
      req_vec3(s, "center", ctx, c);
      double r = req_num(s, "radius", ctx), w = req_num(s, "width", ctx);
      opt_num(s, "thickness", ctx, th);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_MOBIUS, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) p.p[k] = c[k];
      p.p[3] = r;
      p.p[4] = w;
      p.p[5] = th;
    }
    for (auto& s : table_array(root, "spring")) {
      const std::string ctx = "spring";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double cr = req_num(s, "coil_radius", ctx), tr = req_num(s, "tube_radius", ctx);
      double pitch = req_num(s, "pitch", ctx), turns = req_num(s, "turns", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_SPRING, mat, -1, new_object());
      p.p[0] = c[0];
      p.p[1] = c[1];
      p.p[2] = c[2];
      p.p[3] = cr;
      p.p[4] = tr;
      p.p[5] = pitch;
      p.p[6] = turns;
    }
    for (auto& s : table_array(root, "superellipsoid")) {
      const std::string ctx = "superellipsoid";
      double c[3], sc[3];
      req_vec3(s, "center", ctx, c);
      req_vec3(s, "scale", ctx, sc);
      double e1 = req_num(s, "e1", ctx), e2 = req_num(s, "e2", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_SUPERELLIPSOID, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = c[k];
        p.p[3 + k] = sc[k];
      }
      p.p[6] = e1;
      p.p[7] = e2;
    }
    for (auto& s : table_array(root, "capsule")) {  // src/scene.rs:1308-1316, src/capsule.rs:15-53
      const std::string ctx = "capsule";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double r = req_num(s, "radius", ctx), h = req_num(s, "height", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      // one world-list entry: the cylinder body, then the bottom and top spheres
      uint32_t obj = new_object();
      lmn_prim& body = add_prim(LMN_PRIM_CYLINDER, mat, -1, obj);
      for (int k = 0; k < 3; ++k) body.p[k] = c[k];
      body.p[3] = r;
      body.p[4] = c[1];
      body.p[5] = c[1] + h;
      for (int cap = 0; cap < 2; ++cap) {
        lmn_prim& sp = add_prim(LMN_PRIM_SPHERE, mat, -1, obj);
        sp.p[0] = c[0];
        sp.p[1] = cap ? c[1] + h : c[1];
        sp.p[2] = c[2];
        sp.p[3] = r;
      }
    }
    for (auto& s : table_array(root, "ellipsoid")) {  // src/scene.rs:1318-1327
      const std::string ctx = "ellipsoid";
      double c[3], radii[3];
      req_vec3(s, "center", ctx, c);
      req_vec3(s, "radii", ctx, radii);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      if (s.get("bump_strength")) fail(LMN_ERR_UNSUPPORTED, "ellipsoid bump maps are not on the GPU path");
      lmn_prim& p = add_prim(LMN_PRIM_ELLIPSOID, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = c[k];
        p.p[3 + k] = radii[k];
      }
    }
    for (auto& s : table_array(root, "light")) {  // src/scene.rs:1329-1344
      const std::string ctx = "light";
      double pos[3], col[3] = {1, 1, 1};
      req_vec3(s, "position", ctx, pos);
      opt_vec3(s, "color", ctx, col);
      double intensity = 10.0, radius = 0.1;
      opt_num(s, "intensity", ctx, intensity);
      opt_num(s, "radius", ctx, radius);
      lmn_material mat{};
      mat.kind = LMN_MAT_EMISSIVE;
      mat.texture = -1;
      mat.intensity = intensity;
      for (int k = 0; k < 3; ++k) mat.color[k] = col[k];
      out_.materials.push_back(mat);
      lmn_prim& p = add_prim(LMN_PRIM_SPHERE, uint32_t(out_.materials.size() - 1), -1, new_object());
      for (int k = 0; k < 3; ++k) p.p[k] = pos[k];
      p.p[3] = radius;
      double e[3] = {col[0] * intensity, col[1] * intensity, col[2] * intensity};
      push_light_sphere(pos, radius, e, uint32_t(out_.materials.size() - 1));
    }
    for (auto& s : table_array(root, "hemisphere")) {  // src/scene.rs:1346-1349
      const std::string ctx = "hemisphere";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double r = req_num(s, "radius", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_HEMISPHERE, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) p.p[k] = c[k];
      p.p[3] = r;
    }
    reject_if_present(root, "fog", "(constant medium)");
    // cones and cylinders: y from center.y to center.y + height, then
    // RotateX/Y/Z and Translate wrappers (src/scene.rs:1364-1375, :1422-1456)
    auto axial = [&](const TomlValue& s, const std::string& ctx, uint32_t kind) {
      double c[3];
      req_vec3(s, "center", ctx, c);
      double r = req_num(s, "radius", ctx), h = req_num(s, "height", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      if (s.get("bump_strength") || s.get("normal_map"))
        fail(LMN_ERR_UNSUPPORTED, ctx + " bump/normal maps are not on the GPU path");
      std::vector<lmn_xform_op> ops;
      double a;
      if (opt_num(s, "rotate_x", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_X, a));
      if (opt_num(s, "rotate_y", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Y, a));
      if (opt_num(s, "rotate_z", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Z, a));
      double t[3];
      if (opt_vec3(s, "translate", ctx, t)) ops.push_back(op(LMN_XF_TRANSLATE, t[0], t[1], t[2]));
      int32_t xf = ops.empty() ? -1 : int32_t(add_transform(ops));
      lmn_prim& p = add_prim(kind, mat, xf, new_object());
      for (int k = 0; k < 3; ++k) p.p[k] = c[k];
      p.p[3] = r;
      p.p[4] = c[1];
      p.p[5] = c[1] + h;
    };
    for (auto& s : table_array(root, "cone")) axial(s, "cone", LMN_PRIM_CONE);
    for (auto& s : table_array(root, "annulus")) {  // src/scene.rs:1377-1386
      const std::string ctx = "annulus";
      double c[3], n[3];
      req_vec3(s, "center", ctx, c);
      req_vec3(s, "normal", ctx, n);
      double ri = req_num(s, "inner_radius", ctx), ro = req_num(s, "outer_radius", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_ANNULUS, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = c[k];
        p.p[3 + k] = n[k];
      }
      p.p[6] = ri;
      p.p[7] = ro;
    }
    for (auto& s : table_array(root, "disk")) {  // src/scene.rs:1388-1420
      const std::string ctx = "disk";
      double c[3], n[3];
      req_vec3(s, "center", ctx, c);
      req_vec3(s, "normal", ctx, n);
      double r = req_num(s, "radius", ctx);
      MatInfo mi;
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material", &mi);
      if (mi.emissive) {  // an emissive disk is an NEE light (normal.unit(), as given)
        lmn_light l{};
        l.kind = LMN_LIGHT_DISK;
        l.material = mat;
        double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
          l.origin[k] = c[k];
          l.normal[k] = n[k] / len;
          l.emission[k] = mi.emission[k];
        }
        l.radius = r;
        out_.lights.push_back(l);
      }
      lmn_prim& p = add_prim(LMN_PRIM_DISK, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = c[k];
        p.p[3 + k] = n[k];
      }
      p.p[6] = r;
    }
    for (auto& s : table_array(root, "cylinder")) axial(s, "cylinder", LMN_PRIM_CYLINDER);
    // RotateX/Y/Z then Translate wrappers of a prism, wedge or pyramid
    // (src/scene.rs:1465-1468, :1475-1478, :1486-1489)
    auto rot_translate = [&](const TomlValue& s, const std::string& ctx) -> int32_t {
      std::vector<lmn_xform_op> ops;
      double a, t[3];
      if (opt_num(s, "rotate_x", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_X, a));
      if (opt_num(s, "rotate_y", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Y, a));
      if (opt_num(s, "rotate_z", ctx, a)) ops.push_back(op(LMN_XF_ROTATE_Z, a));
      if (opt_vec3(s, "translate", ctx, t)) ops.push_back(op(LMN_XF_TRANSLATE, t[0], t[1], t[2]));
      return ops.empty() ? -1 : int32_t(add_transform(ops));
    };
    typedef std::array<double, 3> P3;
    auto facet = [&](uint32_t mat, int32_t xf, uint32_t obj, const P3& a, const P3& b, const P3& c, const P3& n) {
      lmn_prim& p = add_prim(LMN_PRIM_FACET, mat, xf, obj);
      for (int k = 0; k < 3; ++k) {
        p.p[k] = a[k];
        p.p[3 + k] = b[k];
        p.p[6 + k] = c[k];
        p.p[9 + k] = n[k];
      }
    };
    auto unit3 = [](const P3& v) {
      double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      return P3{v[0] / l, v[1] / l, v[2] / l};
    };
    auto cross3 = [](const P3& a, const P3& b) {
      return P3{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    };
    auto sub3 = [](const P3& a, const P3& b) { return P3{a[0] - b[0], a[1] - b[1], a[2] - b[2]}; };
    for (auto& s : table_array(root, "prism")) {  // src/scene.rs:1450-1470, src/prism.rs:18-85
      const std::string ctx = "prism";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double side = req_num(s, "side", ctx), h = req_num(s, "height", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      if (s.get("bump_strength") || s.get("normal_map"))
        fail(LMN_ERR_UNSUPPORTED, "prism bump/normal maps are not on the GPU path");
      int32_t xf = rot_translate(s, ctx);
      uint32_t obj = new_object();
      double r = side / std::sqrt(3.0);  // circumradius of the equilateral base
      const P3 bot[3] = {P3{c[0], c[1], c[2] + r}, P3{c[0] - side / 2.0, c[1], c[2] - r / 2.0},
                         P3{c[0] + side / 2.0, c[1], c[2] - r / 2.0}};
      P3 top[3];
      for (int i = 0; i < 3; ++i) top[i] = P3{bot[i][0], bot[i][1] + h, bot[i][2]};
      facet(mat, xf, obj, bot[0], bot[1], bot[2], P3{0.0, -1.0, 0.0});
      facet(mat, xf, obj, top[0], top[2], top[1], P3{0.0, 1.0, 0.0});
      for (int i = 0; i < 3; ++i) {  // side normal edge x up, as the reference forms it
        int j = (i + 1) % 3;
        P3 n = unit3(cross3(sub3(bot[j], bot[i]), P3{0.0, 1.0, 0.0}));
        facet(mat, xf, obj, bot[i], bot[j], top[j], n);
        facet(mat, xf, obj, bot[i], top[j], top[i], n);
      }
    }
    for (auto& s : table_array(root, "wedge")) {  // src/scene.rs:1472-1480, src/wedge.rs:22-120
      const std::string ctx = "wedge";
      double mn[3], mx[3];
      req_vec3(s, "min", ctx, mn);
      req_vec3(s, "max", ctx, mx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      int32_t xf = rot_translate(s, ctx);
      uint32_t obj = new_object();
      for (int face = 0; face < 5; ++face) {
        lmn_prim& p = add_prim(LMN_PRIM_WEDGE_FACE, mat, xf, obj);
        for (int k = 0; k < 3; ++k) {
          p.p[k] = mn[k];
          p.p[3 + k] = mx[k];
        }
        p.p[6] = face;
      }
    }
    for (auto& s : table_array(root, "pyramid")) {  // src/scene.rs:1482-1491, src/pyramid.rs:52-97
      const std::string ctx = "pyramid";
      double c[3];
      req_vec3(s, "center", ctx, c);
      double hb = req_num(s, "base_size", ctx) / 2.0, h = req_num(s, "height", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      int32_t xf = rot_translate(s, ctx);
      uint32_t obj = new_object();
      const P3 b0{c[0] - hb, c[1], c[2] - hb}, b1{c[0] + hb, c[1], c[2] - hb}, b2{c[0] + hb, c[1], c[2] + hb},
          b3{c[0] - hb, c[1], c[2] + hb}, apex{c[0], c[1] + h, c[2]};
      const P3 tri[6][3] = {{b0, b2, b1}, {b0, b3, b2}, {b0, b1, apex}, {b1, b2, apex}, {b2, b3, apex}, {b3, b0, apex}};
      for (auto& t : tri)  // normal unit(e1 x e2), as the reference forms it at each hit
        facet(mat, xf, obj, t[0], t[1], t[2], unit3(cross3(sub3(t[1], t[0]), sub3(t[2], t[0]))));
    }
    for (auto& s : table_array(root, "triangle")) {
      const std::string ctx = "triangle";
      double v[9];
      req_vec3(s, "v0", ctx, v);
      req_vec3(s, "v1", ctx, v + 3);
      req_vec3(s, "v2", ctx, v + 6);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_TRIANGLE, mat, -1, new_object());
      for (int k = 0; k < 9; ++k) p.p[k] = v[k];
    }
    for (auto& s : table_array(root, "quad")) {  // src/scene.rs:1503-1543
      const std::string ctx = "quad";
      double q[3], u[3], v[3];
      req_vec3(s, "q", ctx, q);
      req_vec3(s, "u", ctx, u);
      req_vec3(s, "v", ctx, v);
      MatInfo mi;
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material", &mi);
      if (mi.emissive) {
        lmn_light l{};
        l.kind = LMN_LIGHT_RECT;
        l.material = mat;
        double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
          l.origin[k] = q[k];
          l.u[k] = u[k];
          l.v[k] = v[k];
          l.normal[k] = n[k] / nl;
          l.emission[k] = mi.emission[k];
        }
        out_.lights.push_back(l);
      }
      if (s.get("normal_map")) fail(LMN_ERR_UNSUPPORTED, "quad normal maps are not on the GPU path");
      lmn_prim& p = add_prim(LMN_PRIM_QUAD, mat, -1, new_object());
      for (int k = 0; k < 3; ++k) {
        p.p[k] = q[k];
        p.p[3 + k] = u[k];
        p.p[6 + k] = v[k];
      }
    }
    for (auto& s : table_array(root, "mesh")) {  // src/scene.rs:1545-1559
      const std::string ctx = "mesh";
      const TomlValue& f = req(s, "file", ctx);
      if (f.kind != TomlValue::STRING) fail(LMN_ERR_PARSE, "mesh.file: expected a string");
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      double scale = 1.0, off[3] = {0, 0, 0};
      opt_num(s, "scale", ctx, scale);
      opt_vec3(s, "offset", ctx, off);
      std::vector<double> tv;
      std::vector<lmn_vertex_attr> ta;
      std::string err;
      const bool ply = f.str.size() >= 4 && f.str.compare(f.str.size() - 4, 4, ".ply") == 0;
      if (ply) {
        // the binary sidecar when it is current, else the text
        MeshData mesh;
        bool from_sidecar = false;
        int rc = load_ply_file(resolve(f.str), f.str, mesh, from_sidecar, err);
        if (rc != LMN_OK) fail(rc, err);
        fprintf(stderr, "Loaded PLY: %zu vertices, %llu faces, %zu triangles%s\n", mesh.n_vertices(),
                (unsigned long long)mesh.n_faces, mesh.n_triangles(), from_sidecar ? " (binary sidecar)" : "");
        emit_mesh(mesh, scale, off, tv);
      } else {
        std::ifstream in(resolve(f.str), std::ios::binary);
        if (!in.good()) fail(LMN_ERR_IO, "Failed to read mesh '" + f.str + "'");
        std::stringstream ss;
        ss << in.rdbuf();
        int rc = load_obj_text(ss.str(), scale, off, tv, ta, err);
        if (rc != LMN_OK) fail(rc, err);
      }
      size_t nt = tv.size() / 9;
      out_.prims.reserve(out_.prims.size() + nt);
      for (size_t t = 0; t < nt; ++t) {
        // a face with normals or uvs is a MeshTriangle with attributes; the
        // rest shade exactly as plain triangles (geometric normal, (u, v))
        const bool smooth = !ta.empty() && (ta[t].has_normals || ta[t].has_uvs);
        lmn_prim& p = add_prim(smooth ? LMN_PRIM_MESH_TRIANGLE : LMN_PRIM_TRIANGLE, mat, -1, new_object());
        for (int k = 0; k < 9; ++k) p.p[k] = tv[t * 9 + k];
        if (smooth) {
          p.p[9] = double(out_.vertex_attrs.size());
          out_.vertex_attrs.push_back(ta[t]);
        }
      }
    }
    for (auto& s : table_array(root, "box")) {  // src/scene.rs:1561-1600
      const std::string ctx = "box";
      double mn[3], mx[3];
      req_vec3(s, "min", ctx, mn);
      req_vec3(s, "max", ctx, mx);
      // make_box calls the material factory once per side
      // (src/rect.rs:128-141); all six materials are identical.
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      double rx, ry, rz, sc, t[3], sxyz[3];
      bool has_rx = opt_num(s, "rotate_x", ctx, rx), has_ry = opt_num(s, "rotate_y", ctx, ry);
      bool has_rz = opt_num(s, "rotate_z", ctx, rz), has_t = opt_vec3(s, "translate", ctx, t);
      bool has_s = opt_num(s, "scale", ctx, sc), has_sxyz = opt_vec3(s, "scale_xyz", ctx, sxyz);
      bool has_xf = has_rx || has_ry || has_rz || has_t || has_s || has_sxyz;
      int32_t xf = -1;
      if (has_xf) {
        std::vector<lmn_xform_op> ops;
        if (has_sxyz)
          ops.push_back(op(LMN_XF_SCALE_XYZ, sxyz[0], sxyz[1], sxyz[2]));
        else if (has_s)
          ops.push_back(op(LMN_XF_SCALE, sc));
        if (has_rx) ops.push_back(op(LMN_XF_ROTATE_X, rx));
        if (has_ry) ops.push_back(op(LMN_XF_ROTATE_Y, ry));
        if (has_rz) ops.push_back(op(LMN_XF_ROTATE_Z, rz));
        if (has_t) ops.push_back(op(LMN_XF_TRANSLATE, t[0], t[1], t[2]));
        xf = int32_t(add_transform(ops));
      }
      uint32_t obj = has_xf ? new_object() : 0;
      // side order +z, -z, +y, -y, +x, -x (src/rect.rs:133-140)
      const uint32_t kinds[6] = {LMN_PRIM_RECT_XY, LMN_PRIM_RECT_XY, LMN_PRIM_RECT_XZ,
                                 LMN_PRIM_RECT_XZ, LMN_PRIM_RECT_YZ, LMN_PRIM_RECT_YZ};
      for (int side = 0; side < 6; ++side) {
        lmn_prim& p = add_prim(kinds[side], mat, xf, has_xf ? obj : new_object());
        bool hi = (side % 2) == 0;
        if (kinds[side] == LMN_PRIM_RECT_XY) {
          p.p[0] = mn[0]; p.p[1] = mx[0]; p.p[2] = mn[1]; p.p[3] = mx[1]; p.p[4] = hi ? mx[2] : mn[2];
        } else if (kinds[side] == LMN_PRIM_RECT_XZ) {
          p.p[0] = mn[0]; p.p[1] = mx[0]; p.p[2] = mn[2]; p.p[3] = mx[2]; p.p[4] = hi ? mx[1] : mn[1];
        } else {
          p.p[0] = mn[1]; p.p[1] = mx[1]; p.p[2] = mn[2]; p.p[3] = mx[2]; p.p[4] = hi ? mx[0] : mn[0];
        }
      }
    }
    for (auto& s : table_array(root, "rect_xy")) {
      const std::string ctx = "rect_xy";
      double x[2], y[2];
      vec2(req(s, "x", ctx), ctx + ".x", x);
      vec2(req(s, "y", ctx), ctx + ".y", y);
      double k = req_num(s, "k", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");  // emissive XY rects are not NEE lights
      lmn_prim& p = add_prim(LMN_PRIM_RECT_XY, mat, -1, new_object());
      p.p[0] = x[0]; p.p[1] = x[1]; p.p[2] = y[0]; p.p[3] = y[1]; p.p[4] = k;
    }
    for (auto& s : table_array(root, "rect_xz")) {  // src/scene.rs:1607-1622
      const std::string ctx = "rect_xz";
      double x[2], z[2];
      vec2(req(s, "x", ctx), ctx + ".x", x);
      vec2(req(s, "z", ctx), ctx + ".z", z);
      double k = req_num(s, "k", ctx);
      const TomlValue& mdesc = req(s, "material", ctx);
      MatInfo mi;
      uint32_t mat = material(mdesc, ctx + ".material", &mi);
      // only `Emissive` (not Blackbody) XZ rects become NEE lights
      std::string tag = material_tag(mdesc, ctx + ".material");
      if (mi.emissive && (tag == "Emissive" || tag == "emissive")) {
        lmn_light l{};
        l.kind = LMN_LIGHT_RECT;
        l.material = mat;
        l.origin[0] = x[0]; l.origin[1] = k; l.origin[2] = z[0];
        l.u[0] = x[1] - x[0];
        l.v[2] = z[1] - z[0];
        l.normal[1] = -1.0;
        for (int c = 0; c < 3; ++c) l.emission[c] = mi.emission[c];
        out_.lights.push_back(l);
      }
      lmn_prim& p = add_prim(LMN_PRIM_RECT_XZ, mat, -1, new_object());
      p.p[0] = x[0]; p.p[1] = x[1]; p.p[2] = z[0]; p.p[3] = z[1]; p.p[4] = k;
    }
    for (auto& s : table_array(root, "rect_yz")) {
      const std::string ctx = "rect_yz";
      double y[2], z[2];
      vec2(req(s, "y", ctx), ctx + ".y", y);
      vec2(req(s, "z", ctx), ctx + ".z", z);
      double k = req_num(s, "k", ctx);
      uint32_t mat = material(req(s, "material", ctx), ctx + ".material");
      lmn_prim& p = add_prim(LMN_PRIM_RECT_YZ, mat, -1, new_object());
      p.p[0] = y[0]; p.p[1] = y[1]; p.p[2] = z[0]; p.p[3] = z[1]; p.p[4] = k;
    }
    reject_if_present(root, "csg", "(CSG)");
  }
};

}  // namespace

int load_obj_text(const std::string& content, double scale, const double offset[3], std::vector<double>& tv,
                  std::vector<lmn_vertex_attr>& ta, std::string& err) {
  // src/obj.rs:18-153.  Numbers as Rust parses them: an index is an unsigned
  // decimal, a coordinate anything strtod takes whole.
  struct Corner {
    uint64_t pos, uv, nrm;  // 1-based; 0 = absent
  };
  std::vector<double> verts, uvs, nrms;
  std::vector<std::array<Corner, 3>> faces;
  std::istringstream in(content);
  std::string line;
  size_t line_num = 0;
  auto fail_line = [&](const std::string& what) {
    err = "Line " + std::to_string(line_num) + ": " + what;
    return LMN_ERR_PARSE;
  };
  auto num = [](const std::string& w, double& x) {
    char* end = nullptr;
    x = std::strtod(w.c_str(), &end);
    return !w.empty() && *end == 0;
  };
  auto index = [](const std::string& w, uint64_t& v) {
    if (w.empty() || w[0] == '-' || w[0] == '+') return false;
    char* end = nullptr;
    v = std::strtoull(w.c_str(), &end, 10);
    return *end == 0;
  };
  while (std::getline(in, line)) {
    ++line_num;
    std::istringstream ls(line);
    std::vector<std::string> parts;
    std::string w;
    while (ls >> w) parts.push_back(w);
    if (parts.empty() || parts[0][0] == '#') continue;
    const std::string& tag = parts[0];
    if (tag == "v" || tag == "vn") {
      if (parts.size() < 4) return fail_line(tag == "v" ? "vertex needs 3 coordinates" : "normal needs 3 components");
      double x[3];
      for (int k = 0; k < 3; ++k)
        if (!num(parts[1 + k], x[k])) return fail_line("invalid float literal");
      if (tag == "v") {
        for (int k = 0; k < 3; ++k) verts.push_back(x[k] * scale + offset[k]);
      } else {  // Vec3::unit: divided by its length
        double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        for (int k = 0; k < 3; ++k) nrms.push_back(x[k] / l);
      }
    } else if (tag == "vt") {
      if (parts.size() < 3) return fail_line("texture coord needs 2 values");
      double u, v;
      if (!num(parts[1], u) || !num(parts[2], v)) return fail_line("invalid float literal");
      uvs.push_back(u);
      uvs.push_back(v);
    } else if (tag == "f") {
      std::vector<Corner> fv;
      for (size_t i = 1; i < parts.size(); ++i) {
        std::vector<std::string> seg;
        size_t b = 0;
        for (;;) {
          size_t e = parts[i].find('/', b);
          seg.push_back(parts[i].substr(b, e == std::string::npos ? std::string::npos : e - b));
          if (e == std::string::npos) break;
          b = e + 1;
        }
        Corner c{0, 0, 0};
        if (!index(seg[0], c.pos)) return fail_line("invalid digit found in string");
        if (seg.size() >= 2 && !seg[1].empty() && !index(seg[1], c.uv)) return fail_line("invalid digit found in string");
        if (seg.size() >= 3 && !seg[2].empty() && !index(seg[2], c.nrm)) return fail_line("invalid digit found in string");
        fv.push_back(c);
      }
      if (fv.size() < 3) return fail_line("face needs at least 3 vertices");
      for (size_t i = 1; i + 1 < fv.size(); ++i) faces.push_back({fv[0], fv[i], fv[i + 1]});  // fan
    }
  }
  const uint64_t nv = verts.size() / 3, nt = uvs.size() / 2, nn = nrms.size() / 3;
  tv.reserve(tv.size() + faces.size() * 9);
  ta.reserve(ta.size() + faces.size());
  for (const auto& f : faces) {
    for (const Corner& c : f)
      if (c.pos == 0 || c.pos > nv) {
        err = "Vertex index " + std::to_string(c.pos) + " out of range";
        return LMN_ERR_PARSE;
      }
    for (const Corner& c : f)
      for (int k = 0; k < 3; ++k) tv.push_back(verts[(c.pos - 1) * 3 + k]);
    lmn_vertex_attr a{};
    // attributes only when every corner names one that exists (a 0 index
    // wraps in the reference's `n - 1` and finds nothing)
    bool hn = true, hu = true;
    for (const Corner& c : f) {
      hn = hn && c.nrm >= 1 && c.nrm <= nn;
      hu = hu && c.uv >= 1 && c.uv <= nt;
    }
    for (int j = 0; j < 3; ++j) {
      for (int k = 0; k < 3 && hn; ++k) a.normal[j][k] = nrms[(f[j].nrm - 1) * 3 + k];
      for (int k = 0; k < 2 && hu; ++k) a.uv[j][k] = uvs[(f[j].uv - 1) * 2 + k];
    }
    a.has_normals = hn ? 1u : 0u;
    a.has_uvs = hu ? 1u : 0u;
    ta.push_back(a);
  }
  fprintf(stderr, "Loaded OBJ: %llu vertices, %llu tex coords, %llu normals, %llu triangles\n",
          (unsigned long long)nv, (unsigned long long)nt, (unsigned long long)nn, (unsigned long long)faces.size());
  return LMN_OK;
}

void emit_mesh(const MeshData& mesh, double scale, const double offset[3], std::vector<double>& tv) {
  // vertices scaled then offset (src/obj.rs:206), each triangle's corners
  // copied out (src/obj.rs:226-240)
  std::vector<double> v(mesh.verts.size());
  for (size_t i = 0; i < mesh.n_vertices(); ++i)
    for (int k = 0; k < 3; ++k) v[i * 3 + k] = mesh.verts[i * 3 + k] * scale + offset[k];
  tv.reserve(tv.size() + mesh.tris.size() * 3);
  for (uint32_t c : mesh.tris)
    for (int k = 0; k < 3; ++k) tv.push_back(v[size_t(c) * 3 + k]);
}

int load_ply_text(const std::string& content, double scale, const double offset[3], std::vector<double>& tv,
                  std::string& err) {
  // src/obj.rs:157-246
  MeshData mesh;
  int rc = parse_ply(content.c_str(), content.size(), mesh, err);
  if (rc == LMN_OK) emit_mesh(mesh, scale, offset, tv);
  return rc;
}

int load_scene_toml(const std::string& text, const std::string& base_dir, OwnedDesc& out, std::string& err) {
  try {
    Loader(base_dir, out).run(text);
    return LMN_OK;
  } catch (const LoadError& e) {
    err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    err = "out of host memory while loading the scene";
    return LMN_ERR_NOMEM;
  }
}

}  // namespace lmn

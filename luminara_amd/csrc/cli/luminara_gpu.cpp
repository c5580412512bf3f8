// luminara-gpu: the reference's command line (src/main.rs:154-945) for the
// flags that reach the render hot path, rendering on the MI355X through the
// library's C ABI.  Scene loading follows scene::load_scene
// (lmn_load_scene); -w/-H/-s/-d/--seed override the [render] table like
// src/main.rs:204-230 and -p/--preview like :231-237 (the camera keeps the
// TOML aspect, :205-210).  lmn_render_image runs the frame and, on the
// device, the reference's default post chain -- exposure (--auto-exposure),
// depth fog, tone map, sRGB, 8-bit (src/render.rs:2594-2665, :3099-3202,
// :3606-3658) -- the AOV passes (:3508-3584) and the Radiance RGBE pixels;
// this front end writes PNG or PPM (src/main.rs:774-810), --save-hdr
// (:915-944, :948-989), --save-depth/--save-normals/--save-albedo
// (:855-913) and --save-json (:840-857).  Effects beyond the default chain
// are not part of this path.
#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/luminara.h"

namespace {

struct Cli {
  std::string scene, output = "output.png", save_hdr, save_json, save_depth, save_normals, save_albedo, tone_map;
  long width = -1, height = -1, samples = -1, depth = -1, threads = -1;
  long long seed = -1;
  double exposure = -1.0, depth_fog = -1.0, time_limit = -1.0, adaptive_threshold = -1.0;
  double fog_color[3] = {-1.0, -1.0, -1.0};
  bool quiet = false, preview = false, benchmark = false, help = false, auto_exposure = false, adaptive = false;
};

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "Error: %s\n", msg.c_str());
  std::exit(1);
}

// Numeric option values: a value that does not parse leaves the setting
// unchanged, as `args[i].parse().ok()` does.
template <class T>
void parse_num(const char* s, T& out) {
  char* end = nullptr;
  errno = 0;
  if constexpr (std::is_floating_point<T>::value) {
    double v = std::strtod(s, &end);
    if (end != s && *end == '\0' && errno == 0) out = T(v);
  } else {
    long long v = std::strtoll(s, &end, 10);
    if (end != s && *end == '\0' && errno == 0 && v >= 0) out = T(v);
  }
}

// --fog-color R,G,B: three comma-separated numbers, else ignored (:1396-1404).
void parse_rgb(const char* s, double out[3]) {
  double v[3];
  int n = 0;
  std::string str(s);
  size_t pos = 0;
  while (n < 4) {
    size_t comma = str.find(',', pos);
    std::string tok = str.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
    size_t a = tok.find_first_not_of(" \t"), b = tok.find_last_not_of(" \t");
    tok = a == std::string::npos ? "" : tok.substr(a, b - a + 1);
    char* end = nullptr;
    double x = std::strtod(tok.c_str(), &end);
    if (!tok.empty() && *end == '\0') {  // filter_map keeps the values that parse
      if (n < 3) v[n] = x;
      ++n;
    }
    if (comma == std::string::npos) break;
    pos = comma + 1;
  }
  if (n == 3)
    for (int k = 0; k < 3; ++k) out[k] = v[k];
}

Cli parse(int argc, char** argv) {
  Cli c;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* { return (i + 1 < argc) ? argv[++i] : nullptr; };
    if (a == "-o" || a == "--output") {
      if (const char* v = next()) c.output = v;
    } else if (a == "-w" || a == "--width") {
      if (const char* v = next()) parse_num(v, c.width);
    } else if (a == "-H" || a == "--height") {
      if (const char* v = next()) parse_num(v, c.height);
    } else if (a == "-s" || a == "--samples") {
      if (const char* v = next()) parse_num(v, c.samples);
    } else if (a == "-d" || a == "--depth") {
      if (const char* v = next()) parse_num(v, c.depth);
    } else if (a == "--seed") {
      if (const char* v = next()) parse_num(v, c.seed);
    } else if (a == "-t" || a == "--threads") {
      if (const char* v = next()) parse_num(v, c.threads);
    } else if (a == "-e" || a == "--exposure") {
      if (const char* v = next()) parse_num(v, c.exposure);
    } else if (a == "--adaptive") {
      c.adaptive = true;
    } else if (a == "--adaptive-threshold") {
      if (const char* v = next()) parse_num(v, c.adaptive_threshold);
    } else if (a == "--time-limit" || a == "--max-time") {
      if (const char* v = next()) parse_num(v, c.time_limit);
    } else if (a == "--auto-exposure") {
      c.auto_exposure = true;
    } else if (a == "--tone-map") {
      if (const char* v = next()) c.tone_map = v;
    } else if (a == "--depth-fog") {
      if (const char* v = next()) parse_num(v, c.depth_fog);
    } else if (a == "--fog-color") {
      if (const char* v = next()) parse_rgb(v, c.fog_color);
    } else if (a == "-q" || a == "--quiet") {
      c.quiet = true;
    } else if (a == "-p" || a == "--preview") {
      c.preview = true;
    } else if (a == "--save-hdr") {
      if (const char* v = next()) c.save_hdr = v;
    } else if (a == "--save-json") {
      if (const char* v = next()) c.save_json = v;
    } else if (a == "--save-depth") {
      if (const char* v = next()) c.save_depth = v;
    } else if (a == "--save-normals") {
      if (const char* v = next()) c.save_normals = v;
    } else if (a == "--save-albedo") {
      if (const char* v = next()) c.save_albedo = v;
    } else if (a == "--benchmark" || a == "--bench") {
      c.benchmark = true;
    } else if (a == "-h" || a == "--help") {
      c.help = true;
    } else if (!a.empty() && a[0] != '-' && c.scene.empty()) {
      c.scene = a;
    } else {
      std::fprintf(stderr, "Warning: option '%s' is not supported by luminara-gpu and is ignored\n", a.c_str());
    }
  }
  return c;
}

// ---- writers --------------------------------------------------------------
void put_be32(std::vector<uint8_t>& b, uint32_t v) {
  for (int k = 3; k >= 0; --k) b.push_back(uint8_t(v >> (8 * k)));
}
void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
  put_be32(out, uint32_t(data.size()));
  size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  uLong crc = crc32(0L, out.data() + start, uInt(out.size() - start));
  put_be32(out, uint32_t(crc));
}
bool write_png(const std::string& path, uint32_t w, uint32_t h, const std::vector<uint8_t>& rgba) {
  std::vector<uint8_t> raw;
  raw.reserve(size_t(h) * (size_t(w) * 4 + 1));
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);  // filter: none
    raw.insert(raw.end(), rgba.begin() + size_t(y) * w * 4, rgba.begin() + size_t(y + 1) * w * 4);
  }
  uLongf zlen = compressBound(uLong(raw.size()));
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) return false;
  z.resize(zlen);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, w);
  put_be32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA
  png_chunk(out, "IHDR", ihdr);
  png_chunk(out, "IDAT", z);
  png_chunk(out, "IEND", {});
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(out.data()), std::streamsize(out.size()));
  return bool(f);
}
std::string ppm_text(uint32_t w, uint32_t h, const std::vector<uint8_t>& rgba) {
  std::ostringstream s;
  s << "P3\n" << w << " " << h << "\n255\n";
  for (size_t i = 0; i < size_t(w) * h; ++i)
    s << int(rgba[i * 4]) << " " << int(rgba[i * 4 + 1]) << " " << int(rgba[i * 4 + 2]) << "\n";
  return s.str();
}
// Radiance RGBE, flat (no run-length), as src/main.rs:948-989 writes it; the
// pixels come encoded from the device.
bool write_rgbe(const std::string& path, uint32_t w, uint32_t h, const std::vector<uint8_t>& rgbe) {
  std::ostringstream head;
  head << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << h << " +X " << w << "\n";
  std::string out = head.str();
  out.append(reinterpret_cast<const char*>(rgbe.data()), rgbe.size());
  std::ofstream f(path, std::ios::binary);
  f << out;
  return bool(f);
}
// Depth pass image (src/main.rs:856-876): grey = depth * 255 / max depth.
std::vector<uint8_t> depth_rgba(const std::vector<float>& d) {
  float mx = 0.0f;
  for (float v : d) mx = std::fmax(mx, v);
  float norm = mx > 0.0f ? 255.0f / mx : 1.0f;
  std::vector<uint8_t> px(d.size() * 4);
  for (size_t i = 0; i < d.size(); ++i) {
    float x = d[i] * norm;
    x = x < 0.0f ? 0.0f : (x > 255.0f ? 255.0f : x);
    uint8_t v = x > 0.0f ? uint8_t(x) : 0;  // NaN -> 0
    px[4 * i] = px[4 * i + 1] = px[4 * i + 2] = v;
    px[4 * i + 3] = 255;
  }
  return px;
}
std::vector<uint8_t> rgb_to_rgba(const std::vector<uint8_t>& rgb) {
  std::vector<uint8_t> px(rgb.size() / 3 * 4);
  for (size_t i = 0; i < rgb.size() / 3; ++i) {
    for (int k = 0; k < 3; ++k) px[4 * i + k] = rgb[3 * i + k];
    px[4 * i + 3] = 255;
  }
  return px;
}
void save_pass(const std::string& path, const char* what, uint32_t w, uint32_t h, const std::vector<uint8_t>& px) {
  if (path.empty()) return;
  if (!write_png(path, w, h, px))
    std::fprintf(stderr, "Error: failed to save %s pass '%s'\n", what, path.c_str());
  else
    std::fprintf(stderr, "Saved %s pass to %s\n", what, path.c_str());
}

std::string read_file(const std::string& path, bool& ok) {
  std::ifstream f(path, std::ios::binary);
  ok = bool(f);
  std::ostringstream s;
  s << f.rdbuf();
  return s.str();
}

void usage() {
  std::fprintf(stderr,
               "Usage: luminara-gpu scene.toml [options]\n\n"
               "  -o, --output FILE       PNG output, or PPM for '.ppm' / '-' (stdout) [output.png]\n"
               "  -w, --width N           image width (the camera keeps the scene's aspect)\n"
               "  -H, --height N          image height\n"
               "  -s, --samples N         samples per pixel (rounded up to a square)\n"
               "  -d, --depth N           maximum bounce depth\n"
               "      --seed N            random seed\n"
               "  -e, --exposure X        exposure multiplier\n"
               "      --auto-exposure     exposure from the log-average luminance\n"
               "      --adaptive          adaptive sampling (stop a pixel once converged)\n"
               "      --adaptive-threshold X  standard-error threshold [0.03]\n"
               "      --time-limit SECS   rows begun after SECS take one sample\n"
               "      --tone-map NAME     aces, reinhard, filmic (uncharted2) or none\n"
               "      --depth-fog N       atmospheric depth fog density (e.g. 0.1)\n"
               "      --fog-color R,G,B   fog colour (0-1 each, default: 0.8,0.85,0.9)\n"
               "  -p, --preview           1/4 resolution, <= 4 spp, depth <= 10\n"
               "  -q, --quiet             no progress output\n"
               "  -t, --threads N         accepted; the render runs on the GPU\n"
               "      --save-hdr FILE     Radiance HDR of the exposed framebuffer\n"
               "      --save-depth FILE   depth pass image\n"
               "      --save-normals FILE normal pass image\n"
               "      --save-albedo FILE  albedo pass image\n"
               "      --save-json FILE    render statistics\n"
               "      --benchmark         the demo scene at 400x225, 64 spp; prints Mrays/s\n");
}

// scene::demo_scene (src/scene.rs): ground sphere r = 1000, the glass,
// Lambertian and metal r = 1 spheres, a 10 x 10 grid of r = 0.2 spheres (80 %
// Lambertian albedo U*U, 15 % metal U[0.5, 1] fuzz U[0, 0.5], 5 % glass 1.5)
// clear of (4, 0.2, 0); camera (13, 2, 3) -> origin, vfov 20, aperture 0.1,
// focus 10; 800 x 450, 64 spp, depth 50, seed 31337, default sky, no NEE
// lights.  The reference draws the grid from an OS-seeded generator, so each
// of its runs renders another instance; this one draws it from a fixed
// SplitMix64 stream, in the reference's draw order.
std::string demo_scene_toml() {
  uint64_t state = 0x9e3779b97f4a7c15ull;
  auto uniform = [&state]() {
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  std::ostringstream t;
  t.precision(17);
  t << "[render]\nwidth = 800\nheight = 450\nsamples = 64\nmax_depth = 50\nseed = 31337\n\n"
    << "[camera]\nlook_from = [13.0, 2.0, 3.0]\nlook_at = [0.0, 0.0, 0.0]\nvup = [0.0, 1.0, 0.0]\n"
    << "vfov = 20.0\naperture = 0.1\nfocus_dist = 10.0\n\n";
  auto sphere = [&t](double x, double y, double z, double r, const std::string& m) {
    t << "[[sphere]]\ncenter = [" << x << ", " << y << ", " << z << "]\nradius = " << r << "\nmaterial = " << m
      << "\n\n";
  };
  auto color = [](double r, double g, double b) {
    std::ostringstream c;
    c.precision(17);
    c << "[" << r << ", " << g << ", " << b << "]";
    return c.str();
  };
  sphere(0.0, -1000.0, 0.0, 1000.0, "{ type = \"lambertian\", color = [0.5, 0.5, 0.5] }");
  sphere(0.0, 1.0, 0.0, 1.0, "{ type = \"dielectric\", refraction_index = 1.5 }");
  sphere(-4.0, 1.0, 0.0, 1.0, "{ type = \"lambertian\", color = [0.4, 0.2, 0.1] }");
  sphere(4.0, 1.0, 0.0, 1.0, "{ type = \"metal\", color = [0.7, 0.6, 0.5], fuzz = 0.0 }");
  for (int a = -5; a < 5; ++a)
    for (int b = -5; b < 5; ++b) {
      double choose = uniform();
      double cx = a + 0.9 * uniform(), cz = b + 0.9 * uniform();
      double dx = cx - 4.0, dz = cz;
      if (std::sqrt(dx * dx + dz * dz) <= 0.9) continue;
      std::string m;
      if (choose < 0.8) {
        double c1[3], c2[3];
        for (double& c : c1) c = uniform();
        for (double& c : c2) c = uniform();
        m = "{ type = \"lambertian\", color = " + color(c1[0] * c2[0], c1[1] * c2[1], c1[2] * c2[2]) + " }";
      } else if (choose < 0.95) {
        double c[3];
        for (double& v : c) v = 0.5 + 0.5 * uniform();
        std::ostringstream f;
        f.precision(17);
        f << uniform() * 0.5;
        m = "{ type = \"metal\", color = " + color(c[0], c[1], c[2]) + ", fuzz = " + f.str() + " }";
      } else {
        m = "{ type = \"dielectric\", refraction_index = 1.5 }";
      }
      sphere(cx, 0.2, cz, 0.2, m);
    }
  return t.str();
}

}  // namespace

int main(int argc, char** argv) {
  Cli cli = parse(argc, argv);
  if (cli.help) {
    usage();
    return 0;
  }
  const auto t_start = std::chrono::steady_clock::now();
  std::string text, base = ".";
  if (cli.scene.empty()) {  // src/main.rs:198-201
    std::fprintf(stderr, "No scene file provided, rendering demo scene...\n");
    text = demo_scene_toml();
  } else {
    bool ok = false;
    text = read_file(cli.scene, ok);
    if (!ok) die("cannot read scene file '" + cli.scene + "'");
    base = cli.scene.find('/') == std::string::npos ? "." : cli.scene.substr(0, cli.scene.rfind('/'));
  }
  lmn_scene_desc* desc = nullptr;
  if (lmn_load_scene(text.c_str(), base.c_str(), &desc) != 0)
    die("invalid scene '" + (cli.scene.empty() ? std::string("demo") : cli.scene) + "': " + lmn_last_error());
  if (cli.benchmark) {  // src/main.rs:688-706: the demo scene, whatever scene was given
    lmn_desc_free(desc);
    desc = nullptr;
    if (lmn_load_scene(demo_scene_toml().c_str(), ".", &desc) != 0) die(std::string("demo scene: ") + lmn_last_error());
    std::fprintf(stderr, "Benchmark: rendering demo scene at 400x225 @ 64 spp...\n");
  }
  lmn_render_config rc = desc->config;
  lmn_scene* scene = nullptr;
  int rcode = lmn_scene_create(desc, &scene);
  lmn_desc_free(desc);
  if (rcode != 0) die(std::string("scene upload failed: ") + lmn_last_error());

  if (cli.benchmark) {  // the benchmark takes none of the overrides (src/main.rs:690-694)
    Cli b;
    b.benchmark = true;
    b.quiet = true;
    cli = b;
  }
  if (cli.width >= 0) rc.width = uint32_t(cli.width);
  if (cli.height >= 0) rc.height = uint32_t(cli.height);
  if (cli.samples >= 0) rc.samples = uint32_t(cli.samples);
  if (cli.depth >= 0) rc.max_depth = uint32_t(cli.depth);
  if (cli.seed >= 0) rc.seed = uint64_t(cli.seed);
  if (cli.exposure >= 0.0) rc.exposure = cli.exposure;
  if (cli.preview) {
    rc.width = std::max<uint32_t>(rc.width / 4, 1);
    rc.height = std::max<uint32_t>(rc.height / 4, 1);
    rc.samples = std::min<uint32_t>(rc.samples, 4);
    rc.max_depth = std::min<uint32_t>(rc.max_depth, 10);
  }
  if (cli.auto_exposure) rc.auto_exposure = 1;
  if (cli.adaptive) rc.adaptive = 1;  // src/main.rs:526-534
  if (cli.adaptive_threshold >= 0.0) rc.adaptive_threshold = cli.adaptive_threshold;
  lmn_post_params post;
  std::memset(&post, 0, sizeof post);
  post.exposure = rc.exposure;
  post.tone_map = rc.tone_map;
  post.auto_exposure = rc.auto_exposure;
  post.depth_fog = cli.depth_fog >= 0.0 ? cli.depth_fog : 0.0;
  const double fog_default[3] = {0.8, 0.85, 0.9};
  for (int k = 0; k < 3; ++k) post.fog_color[k] = cli.fog_color[0] >= 0.0 || cli.fog_color[1] >= 0.0 ||
                                                          cli.fog_color[2] >= 0.0
                                                      ? cli.fog_color[k]
                                                      : fog_default[k];
  if (!cli.tone_map.empty()) {  // src/main.rs:550-561
    const std::string& t = cli.tone_map;
    if (t == "aces")
      post.tone_map = LMN_TONE_ACES;
    else if (t == "reinhard")
      post.tone_map = LMN_TONE_REINHARD;
    else if (t == "filmic" || t == "uncharted2")
      post.tone_map = LMN_TONE_FILMIC;
    else if (t == "none")
      post.tone_map = LMN_TONE_NONE;
    else {
      std::fprintf(stderr, "Warning: unknown tone map '%s', using ACES\n", t.c_str());
      post.tone_map = LMN_TONE_ACES;
    }
  }
  char dev[256] = {0};
  int n_cu = 0;
  double hbm = 0;
  lmn_device_info(dev, sizeof dev, &n_cu, &hbm);

  lmn_render_params p;
  std::memset(&p, 0, sizeof p);
  if (cli.benchmark) {
    rc.width = 400;
    rc.height = 225;
    rc.samples = 64;
    rc.has_crop = 0;
  }
  if (rc.width == 0 || rc.height == 0) die("width and height must be > 0");
  if (rc.samples == 0) die("samples must be > 0");
  p.width = rc.width;
  p.height = rc.height;
  p.spp = rc.samples;
  p.max_depth = rc.max_depth;
  p.seed = rc.seed;
  p.filter = rc.filter;
  p.has_crop = rc.has_crop;
  std::memcpy(p.crop, rc.crop, sizeof p.crop);
  p.path = LMN_PATH_WAVEFRONT_FP32;
  p.intersect_mode = LMN_ISECT_EXACT;
  if (rc.adaptive) p.flags |= LMN_FLAG_ADAPTIVE;
  p.adaptive_threshold = rc.adaptive_threshold;
  p.time_limit = cli.time_limit > 0.0 ? cli.time_limit : 0.0;
  uint32_t out_w = rc.width, out_h = rc.height;
  if (rc.has_crop) {
    out_w = std::min(rc.crop[2], rc.width > rc.crop[0] ? rc.width - rc.crop[0] : 0u);
    out_h = std::min(rc.crop[3], rc.height > rc.crop[1] ? rc.height - rc.crop[1] : 0u);
  }
  if (!cli.quiet && !cli.benchmark)
    std::fprintf(stderr, "Rendering %ux%u (%.1fMP) @ %u spp, max depth %u, on %s (%d CUs)\n", rc.width, rc.height,
                 double(out_w) * out_h / 1e6, rc.samples, rc.max_depth, dev, n_cu);
  const size_t npix = size_t(std::max(out_w, 1u)) * std::max(out_h, 1u);
  std::vector<uint8_t> rgba(npix * 4), rgbe, normal, albedo;
  std::vector<float> hdr, depth;
  lmn_image img;
  std::memset(&img, 0, sizeof img);
  img.rgba = rgba.data();
  if (!cli.save_hdr.empty() && !cli.benchmark) {
    hdr.resize(npix * 3);
    rgbe.resize(npix * 4);
    img.hdr = hdr.data();
    img.rgbe = rgbe.data();
  }
  if (!cli.save_depth.empty() && !cli.benchmark) {
    depth.resize(npix);
    img.depth = depth.data();
  }
  if (!cli.save_normals.empty() && !cli.benchmark) {
    normal.resize(npix * 3);
    img.normal = normal.data();
  }
  if (!cli.save_albedo.empty() && !cli.benchmark) {
    albedo.resize(npix * 3);
    img.albedo = albedo.data();
  }
  lmn_stats st;
  std::memset(&st, 0, sizeof st);
  if (lmn_render_image(scene, &p, &post, &img, &st) != 0) die(std::string("render failed: ") + lmn_last_error());
  lmn_scene_destroy(scene);
  const double trace_s = st.render_ms / 1e3;
  if (cli.benchmark) {
    double mrays = double(st.paths) / trace_s / 1e6;
    uint32_t sq = uint32_t(std::ceil(std::sqrt(double(rc.samples))));
    std::fprintf(stderr, "Benchmark result: %.1f Mrays/s (%.3fs, %s)\n", mrays, trace_s, dev);
    std::fprintf(stderr, "  %llu primary rays, %u spp (stratified %ux%u)\n", (unsigned long long)st.paths, sq * sq,
                 sq, sq);
    return 0;
  }
  if (!cli.quiet) {  // src/render.rs:1762-1773
    const uint64_t npx = uint64_t(out_w) * out_h;
    const uint32_t sq = uint32_t(std::ceil(std::sqrt(double(rc.samples))));
    if (rc.adaptive) {
      const double max_rays = double(npx) * sq * sq;
      std::fprintf(stderr, "Adaptive: %llu rays (%.1f avg spp, max %u, %.1f%% saved)\n",
                   (unsigned long long)st.paths, double(st.paths) / double(npx), sq * sq,
                   100.0 * (1.0 - double(st.paths) / max_rays));
    } else {
      std::fprintf(stderr, "Primary rays: %llu (%u spp, %ux%u stratified)\n", (unsigned long long)st.paths, sq * sq,
                   sq, sq);
    }
  }
  if (post.auto_exposure && !cli.quiet)  // src/render.rs:2595-2605
    std::fprintf(stderr, "Auto-exposure: %.3f (scene avg luminance: %.4f)\n", img.exposure / post.exposure,
                 img.log_average);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  const double mrays = double(st.paths) / trace_s / 1e6;
  if (!cli.quiet)
    std::fprintf(stderr, "Rendered in %.2fs (trace: %.2fs, %.1f Mrays/s, %llu total rays)\n", secs, trace_s, mrays,
                 (unsigned long long)st.paths);

  const std::string& out = cli.output;
  bool ppm = out == "-" || (out.size() >= 4 && out.compare(out.size() - 4, 4, ".ppm") == 0);
  if (ppm) {
    std::string s = ppm_text(out_w, out_h, rgba);
    if (out == "-") {
      std::fwrite(s.data(), 1, s.size(), stdout);
      if (!cli.quiet) std::fprintf(stderr, "Written to stdout (PPM)\n");
    } else {
      std::ofstream f(out, std::ios::binary);
      f << s;
      if (!f) die("failed to write '" + out + "'");
      if (!cli.quiet) std::fprintf(stderr, "Saved to %s\n", out.c_str());
    }
  } else {
    if (!write_png(out, out_w, out_h, rgba)) die("failed to save image '" + out + "'");
    if (!cli.quiet) std::fprintf(stderr, "Saved to %s\n", out.c_str());
  }
  if (!cli.save_json.empty()) {
    std::ostringstream j;
    j.setf(std::ios::fixed);
    j << "{\n  \"width\": " << out_w << ",\n  \"height\": " << out_h << ",\n  \"samples\": " << rc.samples
      << ",\n  \"max_depth\": " << rc.max_depth << ",\n  \"total_rays\": " << st.paths;
    j.precision(4);
    j << ",\n  \"render_time_secs\": " << trace_s << ",\n  \"total_time_secs\": " << secs;
    j.precision(2);
    j << ",\n  \"mrays_per_sec\": " << mrays << ",\n  \"device\": \"" << dev << "\",\n  \"output\": \"" << out
      << "\"\n}";
    std::ofstream f(cli.save_json);
    f << j.str();
    if (!f)
      std::fprintf(stderr, "Warning: failed to write JSON stats '%s'\n", cli.save_json.c_str());
    else if (!cli.quiet)
      std::fprintf(stderr, "Stats saved to %s\n", cli.save_json.c_str());
  }
  if (!cli.save_depth.empty()) save_pass(cli.save_depth, "depth", out_w, out_h, depth_rgba(depth));
  if (!cli.save_normals.empty()) save_pass(cli.save_normals, "normal", out_w, out_h, rgb_to_rgba(normal));
  if (!cli.save_albedo.empty()) save_pass(cli.save_albedo, "albedo", out_w, out_h, rgb_to_rgba(albedo));
  if (!cli.save_hdr.empty()) {
    if (!write_rgbe(cli.save_hdr, out_w, out_h, rgbe)) die("failed to save HDR '" + cli.save_hdr + "'");
    std::fprintf(stderr, "Saved HDR to %s\n", cli.save_hdr.c_str());
  }
  return 0;
}

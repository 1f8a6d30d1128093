/*
 * gt_util.cpp - host helpers of the C-ABI: error strings, the uniform block,
 * the WASD camera matrix, seeded synthetic scenes and the .ply reader.
 */
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gt.h"
#include "gt_internal.h"

static thread_local std::string g_last_error;

extern "C" void gt_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

extern "C" const char* gt_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* gt_version(void) { return "gt 0.1.0 (gfx950)"; }
extern "C" void gt_free(void* p) { std::free(p); }

/* src/main.ts:520-534 */
extern "C" void gt_uniforms_init(gt_uniforms* u, const float to_world[16], double hfov_deg, uint32_t width,
                                 uint32_t height, float t_min, float t_max, float T_min, float alpha_min,
                                 uint32_t k, uint32_t sh_degree_max, int32_t early_stop, uint32_t debug_mode) {
  std::memset(u, 0, sizeof(*u));
  std::memcpy(u->to_world, to_world, 16 * sizeof(float));
  const float ax = (float)std::tan(0.5 * hfov_deg * M_PI / 180.0);
  u->camera_aspect[0] = ax;
  u->camera_aspect[1] = (float)((double)ax * ((double)height / (double)width));
  u->t_min = t_min;
  u->t_max = t_max;
  u->T_min = T_min;
  u->alpha_min = alpha_min;
  u->k = k;
  u->sh_degree_max = sh_degree_max;
  u->accumulated_frames = 0;
  u->early_stop = early_stop <= 0 ? 0xffffffffu : (uint32_t)early_stop;
  u->debug_mode = debug_mode;
}

/* src/camera.ts:126-140 (constructor: recalculateAngles) and 153-170
 * (update: matrix = rotateX(rotationY(yaw), pitch), position restored).
 * wgpu-matrix evaluates in double and stores to Float32Array. */
extern "C" void gt_camera_look_at(const float pos[3], const float target[3], float m[16]) {
  float b[3];
  for (int i = 0; i < 3; ++i) b[i] = (float)((double)pos[i] - (double)target[i]);
  const double len = std::sqrt((double)b[0] * b[0] + (double)b[1] * b[1] + (double)b[2] * b[2]);
  if (len > 1e-5)
    for (int i = 0; i < 3; ++i) b[i] = (float)((double)b[i] / len);
  const double yaw = std::atan2((double)b[0], (double)b[2]);
  const double pitch = -std::asin((double)b[1]);
  // rotationY(yaw)
  float ry[16] = {0};
  const double cy = std::cos(yaw), sy = std::sin(yaw);
  ry[0] = (float)cy; ry[2] = (float)-sy; ry[5] = 1.0f; ry[8] = (float)sy; ry[10] = (float)cy; ry[15] = 1.0f;
  // rotateX(ry, pitch)
  const double c = std::cos(pitch), s = std::sin(pitch);
  for (int i = 0; i < 16; ++i) m[i] = ry[i];
  for (int j = 0; j < 4; ++j) {
    const double m1 = ry[4 + j], m2 = ry[8 + j];
    m[4 + j] = (float)(c * m1 + s * m2);
    m[8 + j] = (float)(c * m2 - s * m1);
  }
  m[12] = pos[0]; m[13] = pos[1]; m[14] = pos[2]; m[15] = 1.0f;
}

/* ---------------- seeded synthetic scenes ----------------
 * SplitMix64 (Steele, Lea, Flood 2014) streams; normals by the Box-Muller
 * transform.  Gaussian i draws from its own stream (seed, i), in this order:
 *   mean offsets (3 normals, or 3 uniforms when there are no clusters),
 *   log-scales (3), quaternion (4 normals), opacity logit (1),
 *   f_dc (3), f_rest (45).
 * Cluster c's centre draws 3 uniforms from stream (seed, 2^40 + c);
 * Gaussian i belongs to cluster i mod n_clusters. */
namespace {
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double normal() {
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};
SplitMix stream(uint64_t seed, uint64_t i) {
  SplitMix m{seed ^ 0x6a09e667f3bcc909ull};
  m.s += i * 0xd1b54a32d192ed03ull;
  (void)m.next();
  return m;
}
}  // namespace

extern "C" int gt_generate_scene(const gt_gen_params* g, float* raw) {
  if (!g || (!raw && g->n)) { gt_set_error("gt_generate_scene: NULL argument"); return GT_E_INVALID; }
  std::vector<double> centres(3 * (size_t)(g->n_clusters ? g->n_clusters : 1), 0.0);
  for (uint32_t c = 0; c < g->n_clusters; ++c) {
    SplitMix m = stream(g->seed, (1ull << 40) + c);
    for (int j = 0; j < 3; ++j) centres[3 * c + j] = (2.0 * m.uniform() - 1.0) * g->box_half;
  }
  for (uint32_t i = 0; i < g->n; ++i) {
    SplitMix m = stream(g->seed, i);
    float* r = raw + (size_t)i * GT_RAW_FLOATS;
    if (g->n_clusters == 0) {
      for (int j = 0; j < 3; ++j) r[j] = (float)((2.0 * m.uniform() - 1.0) * g->box_half);
    } else {
      const uint32_t c = i % g->n_clusters;
      for (int j = 0; j < 3; ++j) r[j] = (float)(centres[3 * c + j] + g->cluster_sigma * m.normal());
    }
    for (int j = 0; j < 3; ++j) r[52 + j] = (float)(g->log_scale_mean + g->log_scale_std * m.normal());
    for (int j = 0; j < 4; ++j) r[55 + j] = (float)m.normal();
    r[51] = (float)(g->logit_mean + g->logit_std * m.normal());
    for (int j = 0; j < 3; ++j) r[3 + j] = (float)(g->dc_std * m.normal());
    for (int j = 0; j < 45; ++j) r[6 + j] = (float)(g->rest_std * m.normal());
  }
  return GT_OK;
}

/* ---------------- .ply reader (src/GS_ply.ts:58-132, 171-196) ----------------
 * Binary little-endian vertex element; float properties are read as f32 and
 * uchar properties as value/255 (GS_ply.ts:122-128).  Unlike the reference,
 * other property types are sized correctly instead of being skipped without
 * advancing (GS_ply.ts:120-129 mis-parses them), and SH degree < 3 leaves the
 * missing coefficients at zero instead of breaking the 320-B layout.
 * The file is untrusted input: the header's vertex count is checked against
 * the bytes the file holds before anything is allocated; every property
 * index must lie in its field (f_dc 0..2, scale 0..2, rot 0..3, f_rest below
 * the f_rest count, which must be a multiple of 3); the 14 fields besides
 * f_rest that the reference reads (GS_ply.ts:225-235) must be present; and
 * no exception leaves the call. */
namespace {

/* decimal index after a property-name prefix; -1 unless the rest is 1-6 digits */
int prop_index(const std::string& name, size_t prefix) {
  if (name.size() <= prefix || name.size() - prefix > 6) return -1;
  int v = 0;
  for (size_t i = prefix; i < name.size(); ++i) {
    const char c = name[i];
    if (c < '0' || c > '9') return -1;
    v = v * 10 + (c - '0');
  }
  return v;
}

int ply_read(const char* path, float** raw_out, uint32_t* n_out, uint32_t* deg_out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { gt_set_error("cannot open %s", path); return GT_E_IO; }
  std::string line;
  std::getline(f, line);
  if (line.rfind("ply", 0) != 0) { gt_set_error("%s: not a ply file", path); return GT_E_IO; }
  struct Prop { std::string name, type; int size; };
  std::vector<Prop> props;
  uint64_t n = 0;
  bool in_vertex = false, binary_le = false, header_done = false, seen_vertex = false;
  int header_lines = 0;
  while (std::getline(f, line)) {
    if (++header_lines > 100000) { gt_set_error("%s: header too long", path); return GT_E_IO; }
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string w;
    ls >> w;
    if (w == "format") {
      std::string fmt;
      ls >> fmt;
      binary_le = fmt == "binary_little_endian";
    } else if (w == "element") {
      std::string name;
      ls >> name;
      in_vertex = name == "vertex";
      if (in_vertex) {
        if (seen_vertex) { gt_set_error("%s: two vertex elements", path); return GT_E_IO; }
        seen_vertex = true;
        std::string cnt;
        ls >> cnt;
        if (cnt.empty() || cnt.size() > 19 || cnt.find_first_not_of("0123456789") != std::string::npos) {
          gt_set_error("%s: bad vertex count '%s'", path, cnt.c_str());
          return GT_E_IO;
        }
        n = std::strtoull(cnt.c_str(), nullptr, 10);
      } else if (!seen_vertex) {
        /* its records would precede the vertex records in the body */
        std::string cnt;
        ls >> cnt;
        if (cnt != "0") {
          gt_set_error("%s: element %s before the vertex element is not supported", path, name.c_str());
          return GT_E_IO;
        }
      }
    } else if (w == "property" && in_vertex) {
      std::string type, name;
      ls >> type >> name;
      int size = 0;
      if (type == "float" || type == "float32" || type == "int" || type == "int32" || type == "uint" ||
          type == "uint32")
        size = 4;
      else if (type == "double" || type == "float64")
        size = 8;
      else if (type == "uchar" || type == "uint8" || type == "char" || type == "int8")
        size = 1;
      else if (type == "short" || type == "ushort" || type == "int16" || type == "uint16")
        size = 2;
      else {
        gt_set_error("%s: unsupported property type %s", path, type.c_str());
        return GT_E_IO;
      }
      for (const Prop& q : props)
        if (q.name == name) { gt_set_error("%s: property %s given twice", path, name.c_str()); return GT_E_IO; }
      props.push_back({name, type, size});
    } else if (w == "end_header") {
      header_done = true;
      break;
    }
  }
  if (!header_done || !binary_le) { gt_set_error("%s: need a binary_little_endian ply", path); return GT_E_IO; }
  if (n > 0xffffffffull) { gt_set_error("%s: too many vertices", path); return GT_E_LIMIT; }
  size_t stride = 0;
  for (auto& p : props) stride += p.size;
  int n_rest = 0;
  for (auto& p : props)
    if (p.name.rfind("f_rest_", 0) == 0) n_rest++;
  if (n_rest % 3 != 0) {
    gt_set_error("%s: %d f_rest coefficients is not a multiple of 3", path, n_rest);
    return GT_E_IO;
  }
  const int per_color = n_rest / 3;
  int deg = 0;
  while ((deg + 1) * (deg + 1) - 1 < per_color) deg++;
  if ((deg + 1) * (deg + 1) - 1 != per_color || deg > 3) {
    gt_set_error("%s: %d f_rest coefficients is not SH degree 0..3", path, n_rest);
    return GT_E_IO;
  }
  // column of each raw field (-1 = a property the reader ignores)
  std::vector<int> col(props.size(), -1);
  uint32_t present = 0; /* bit b: raw column b (0..5) or 45 + b (51..58) was found */
  for (size_t i = 0; i < props.size(); ++i) {
    const std::string& nm = props[i].name;
    int idx = -1;
    const bool dc = nm.rfind("f_dc_", 0) == 0, rest = nm.rfind("f_rest_", 0) == 0;
    const bool scale = nm.rfind("scale_", 0) == 0, rot = nm.rfind("rot_", 0) == 0;
    if (nm == "x") idx = 0; else if (nm == "y") idx = 1; else if (nm == "z") idx = 2;
    else if (nm == "opacity") idx = 51;
    else if (dc || rest || scale || rot) {
      const size_t pre = dc ? 5 : rest ? 7 : scale ? 6 : 4;
      const int lim = dc ? 3 : rest ? n_rest : scale ? 3 : 4;
      const int r = prop_index(nm, pre);
      if (r < 0 || r >= lim) {
        gt_set_error("%s: property %s names no field (index out of range)", path, nm.c_str());
        return GT_E_IO;
      }
      if (dc) idx = 3 + r;
      else if (rest) idx = 6 + (r / per_color) * 15 + (r % per_color); /* f_rest_{rgb*per_color+j} -> rgb*15+j */
      else if (scale) idx = 52 + r;
      else idx = 55 + r;
    }
    col[i] = idx;
    if (idx >= 0 && idx < 6) present |= 1u << idx;
    if (idx >= 51) present |= 1u << (idx - 45);
  }
  if (present != (1u << 14) - 1u) {
    gt_set_error("%s: x y z, f_dc_0..2, opacity, scale_0..2 and rot_0..3 are required", path);
    return GT_E_IO;
  }
  /* the records must be in the file before anything is sized by n */
  const std::streamoff data_pos = f.tellg();
  f.seekg(0, std::ios::end);
  const std::streamoff end_pos = f.tellg();
  if (data_pos < 0 || end_pos < data_pos) { gt_set_error("%s: cannot size the vertex data", path); return GT_E_IO; }
  const uint64_t avail = (uint64_t)(end_pos - data_pos);
  uint64_t need = 0;
  if (__builtin_mul_overflow((uint64_t)stride, n, &need) || need > avail) {
    gt_set_error("%s: truncated vertex data (%llu vertices of %zu B, %llu B in the file)", path,
                 (unsigned long long)n, stride, (unsigned long long)avail);
    return GT_E_IO;
  }
  f.seekg(data_pos);
  std::vector<char> buf((size_t)need);
  f.read(buf.data(), (std::streamsize)buf.size());
  if ((size_t)f.gcount() != buf.size()) { gt_set_error("%s: truncated vertex data", path); return GT_E_IO; }
  float* raw = (float*)std::calloc((size_t)n * GT_RAW_FLOATS + 1, sizeof(float));
  if (!raw) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  for (uint64_t v = 0; v < n; ++v) {
    const char* p = buf.data() + v * stride;
    float* r = raw + v * GT_RAW_FLOATS;
    for (size_t i = 0; i < props.size(); ++i) {
      const Prop& pr = props[i];
      double val = 0.0;
      if (pr.type == "float" || pr.type == "float32") { float x; std::memcpy(&x, p, 4); val = x; }
      else if (pr.type == "double" || pr.type == "float64") { double x; std::memcpy(&x, p, 8); val = x; }
      else if (pr.type == "uchar" || pr.type == "uint8") { val = (unsigned char)*p / 255.0; }
      else if (pr.type == "char" || pr.type == "int8") { val = (signed char)*p; }
      else if (pr.type == "int" || pr.type == "int32") { int32_t x; std::memcpy(&x, p, 4); val = x; }
      else if (pr.type == "uint" || pr.type == "uint32") { uint32_t x; std::memcpy(&x, p, 4); val = x; }
      else if (pr.type == "short" || pr.type == "int16") { int16_t x; std::memcpy(&x, p, 2); val = x; }
      else { uint16_t x; std::memcpy(&x, p, 2); val = x; }
      if (col[i] >= 0) r[col[i]] = (float)val;
      p += pr.size;
    }
  }
  *raw_out = raw;
  *n_out = (uint32_t)n;
  if (deg_out) *deg_out = (uint32_t)deg;
  return GT_OK;
}

}  // namespace

extern "C" int gt_ply_read(const char* path, float** raw_out, uint32_t* n_out, uint32_t* deg_out) {
  if (!path || !raw_out || !n_out) { gt_set_error("gt_ply_read: NULL argument"); return GT_E_INVALID; }
  *raw_out = nullptr;
  *n_out = 0;
  try {
    return ply_read(path, raw_out, n_out, deg_out);
  } catch (const std::bad_alloc&) {
    gt_set_error("%s: out of host memory", path);
    return GT_E_NOMEM;
  } catch (const std::exception& e) {
    gt_set_error("%s: %s", path, e.what());
    return GT_E_IO;
  }
}

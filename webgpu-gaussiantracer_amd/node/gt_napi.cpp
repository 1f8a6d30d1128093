/*
 * gt_napi.cpp - Node-API addon over libgt.so for the JavaScript host.
 *
 * The reference's host is TypeScript driving WebGPU + WebRTX (src/main.ts,
 * src/camera.ts, src/GS_ply.ts); here the same host-side roles call the
 * C-ABI of include/gt.h.  Exported functions (all synchronous, host
 * buffers; errors are thrown as JS Error objects carrying gt_last_error()):
 *   version()                                  -> string
 *   plyRead(path)                              -> {raw: Float32Array(n*59), n, shDegree}
 *   generateScene({n, nClusters, boxHalf, ...}) -> Float32Array(n*59)
 *   cameraLookAt(pos[3], target[3])            -> Float32Array(16)
 *   sceneCreate(raw, n, alphaMin)              -> scene handle (external)
 *   sceneSetAlphaMin(scene, alphaMin)
 *   sceneInfo(scene)                           -> {nGaussians, nProxies, ...}
 *   sceneDestroy(scene)
 *   render(scene, toWorld, opts)               -> Float32Array(W*H*4)
 * opts: {width, height, hFov, t_min, t_max, T_min, alpha_min,
 *        hit_array_size, sh_degree_max, earlyStop, debugMode} as the
 * uniforms object and upadateUniformData of src/main.ts:448-534.
 */
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/gt.h"

namespace {

#define NAPI_TRY(env, call)                                           \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                                 \
    }                                                                 \
  } while (0)

napi_value throw_gt(napi_env env, int rc) {
  std::string msg = "gt error " + std::to_string(rc) + ": " + gt_last_error();
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

bool get_f32_array(napi_env env, napi_value v, float** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off = 0;
  void* p = nullptr;
  if (napi_get_typedarray_info(env, v, &t, len, &p, &ab, &off) != napi_ok || t != napi_float32_array) return false;
  *data = static_cast<float*>(p);
  return true;
}

napi_value new_f32_array(napi_env env, size_t n, float** data) {
  napi_value ab, arr;
  void* p = nullptr;
  if (napi_create_arraybuffer(env, n * sizeof(float), &p, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr) != napi_ok) return nullptr;
  *data = static_cast<float*>(p);
  return arr;
}

double num_prop(napi_env env, napi_value obj, const char* name, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  double d = dflt;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) return dflt;
  napi_valuetype ty;
  napi_typeof(env, v, &ty);
  if (ty != napi_number) return dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

bool bool_prop(napi_env env, napi_value obj, const char* name, bool dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) return dflt;
  bool b = dflt;
  napi_valuetype ty;
  napi_typeof(env, v, &ty);
  if (ty == napi_boolean) napi_get_value_bool(env, v, &b);
  return b;
}

void set_num(napi_env env, napi_value obj, const char* name, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  napi_set_named_property(env, obj, name, v);
}

bool get_vec3(napi_env env, napi_value v, float out[3]) {
  bool is_arr = false;
  napi_is_array(env, v, &is_arr);
  float* p = nullptr;
  size_t len = 0;
  if (get_f32_array(env, v, &p, &len)) {
    if (len < 3) return false;
    std::memcpy(out, p, 3 * sizeof(float));
    return true;
  }
  if (!is_arr) return false;
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double d = 0;
    if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) return false;
    out[i] = (float)d;
  }
  return true;
}

/* scene handles are externals; sceneDestroy clears the pointer slot */
struct Handle {
  gt_scene* s;
};

void handle_finalize(napi_env, void* data, void*) {
  Handle* h = static_cast<Handle*>(data);
  if (h->s) gt_scene_destroy(h->s);
  delete h;
}

gt_scene* get_scene(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<Handle*>(p)->s) {
    napi_throw_type_error(env, nullptr, "not a live scene handle");
    return nullptr;
  }
  return static_cast<Handle*>(p)->s;
}

napi_value js_version(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_TRY(env, napi_create_string_utf8(env, gt_version(), NAPI_AUTO_LENGTH, &v));
  return v;
}

napi_value js_ply_read(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  char path[4096];
  size_t len = 0;
  if (napi_get_value_string_utf8(env, argv[0], path, sizeof(path), &len) != napi_ok) {
    napi_throw_type_error(env, nullptr, "plyRead(path: string)");
    return nullptr;
  }
  float* raw = nullptr;
  uint32_t n = 0, deg = 0;
  const int rc = gt_ply_read(path, &raw, &n, &deg);
  if (rc != GT_OK) return throw_gt(env, rc);
  float* dst = nullptr;
  napi_value arr = new_f32_array(env, (size_t)n * GT_RAW_FLOATS, &dst);
  if (!arr) {
    gt_free(raw);
    napi_throw_error(env, nullptr, "allocation failed");
    return nullptr;
  }
  if (n) std::memcpy(dst, raw, (size_t)n * GT_RAW_FLOATS * sizeof(float));
  gt_free(raw);
  napi_value obj;
  NAPI_TRY(env, napi_create_object(env, &obj));
  NAPI_TRY(env, napi_set_named_property(env, obj, "raw", arr));
  set_num(env, obj, "n", n);
  set_num(env, obj, "shDegree", deg);
  return obj;
}

napi_value js_generate(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  gt_gen_params g{};
  g.n = (uint32_t)num_prop(env, argv[0], "n", 0);
  g.n_clusters = (uint32_t)num_prop(env, argv[0], "nClusters", 0);
  g.box_half = (float)num_prop(env, argv[0], "boxHalf", 2.0);
  g.cluster_sigma = (float)num_prop(env, argv[0], "clusterSigma", 0.8);
  g.log_scale_mean = (float)num_prop(env, argv[0], "logScaleMean", -3.0);
  g.log_scale_std = (float)num_prop(env, argv[0], "logScaleStd", 0.5);
  g.logit_mean = (float)num_prop(env, argv[0], "logitMean", 0.0);
  g.logit_std = (float)num_prop(env, argv[0], "logitStd", 1.5);
  g.dc_std = (float)num_prop(env, argv[0], "dcStd", 0.5);
  g.rest_std = (float)num_prop(env, argv[0], "restStd", 0.05);
  g.seed = (uint64_t)num_prop(env, argv[0], "seed", 0);
  float* dst = nullptr;
  napi_value arr = new_f32_array(env, (size_t)g.n * GT_RAW_FLOATS, &dst);
  if (!arr) {
    napi_throw_error(env, nullptr, "allocation failed");
    return nullptr;
  }
  const int rc = gt_generate_scene(&g, dst);
  if (rc != GT_OK) return throw_gt(env, rc);
  return arr;
}

napi_value js_camera_look_at(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  float pos[3], tgt[3];
  if (!get_vec3(env, argv[0], pos) || !get_vec3(env, argv[1], tgt)) {
    napi_throw_type_error(env, nullptr, "cameraLookAt(pos: number[3], target: number[3])");
    return nullptr;
  }
  float* m = nullptr;
  napi_value arr = new_f32_array(env, 16, &m);
  if (!arr) return nullptr;
  gt_camera_look_at(pos, tgt, m);
  return arr;
}

napi_value js_scene_create(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  size_t argc = 4;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < 3) {
    napi_throw_type_error(env, nullptr, "sceneCreate(raw: Float32Array, n: number, alphaMin: number, gpuBuild?: boolean)");
    return nullptr;
  }
  bool gpu = false;
  if (argc >= 4) {
    napi_valuetype t;
    if (napi_typeof(env, argv[3], &t) == napi_ok && t == napi_boolean) napi_get_value_bool(env, argv[3], &gpu);
  }
  float* raw = nullptr;
  size_t len = 0;
  uint32_t n = 0;
  double amin = 0;
  if (!get_f32_array(env, argv[0], &raw, &len) || napi_get_value_uint32(env, argv[1], &n) != napi_ok ||
      napi_get_value_double(env, argv[2], &amin) != napi_ok) {
    napi_throw_type_error(env, nullptr, "sceneCreate(raw: Float32Array, n: number, alphaMin: number)");
    return nullptr;
  }
  if (len < (size_t)n * GT_RAW_FLOATS) {
    napi_throw_range_error(env, nullptr, "raw holds fewer than n * 59 floats");
    return nullptr;
  }
  gt_scene* s = nullptr;
  /* gpuBuild: upload once and build on the GPU (gt_scene_create_raw_gpu) */
  const int rc = gpu ? gt_scene_create_raw_gpu(raw, n, amin, &s) : gt_scene_create_raw(raw, n, amin, &s);
  if (rc != GT_OK) return throw_gt(env, rc);
  Handle* h = new Handle{s};
  napi_value ext;
  if (napi_create_external(env, h, handle_finalize, nullptr, &ext) != napi_ok) {
    gt_scene_destroy(s);
    delete h;
    napi_throw_error(env, nullptr, "napi_create_external failed");
    return nullptr;
  }
  return ext;
}

napi_value js_scene_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    Handle* h = static_cast<Handle*>(p);
    if (h->s) gt_scene_destroy(h->s);
    h->s = nullptr;
  }
  return nullptr;
}

napi_value js_scene_set_alpha_min(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  gt_scene* s = get_scene(env, argv[0]);
  if (!s) return nullptr;
  double a = 0;
  NAPI_TRY(env, napi_get_value_double(env, argv[1], &a));
  const int rc = gt_scene_set_alpha_min(s, a);
  if (rc != GT_OK) return throw_gt(env, rc);
  return nullptr;
}

napi_value js_scene_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  gt_scene* s = get_scene(env, argv[0]);
  if (!s) return nullptr;
  gt_scene_info si{};
  const int rc = gt_scene_get_info(s, &si);
  if (rc != GT_OK) return throw_gt(env, rc);
  napi_value obj;
  NAPI_TRY(env, napi_create_object(env, &obj));
  set_num(env, obj, "nGaussians", si.n_gaussians);
  set_num(env, obj, "nProxies", si.n_proxies);
  set_num(env, obj, "nNodes", si.n_nodes);
  set_num(env, obj, "maxStack", si.max_stack);
  set_num(env, obj, "bvhDepth", si.bvh_depth);
  set_num(env, obj, "deviceBytes", (double)si.device_bytes);
  set_num(env, obj, "buildMs", si.build_ms);
  return obj;
}

napi_value js_render(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  gt_scene* s = get_scene(env, argv[0]);
  if (!s) return nullptr;
  float* tw = nullptr;
  size_t len = 0;
  if (!get_f32_array(env, argv[1], &tw, &len) || len < 16) {
    napi_throw_type_error(env, nullptr, "render(scene, toWorld: Float32Array(16), opts)");
    return nullptr;
  }
  const napi_value o = argv[2];
  const double w = num_prop(env, o, "width", 0), hgt = num_prop(env, o, "height", 0);
  if (!(w >= 1 && hgt >= 1 && w <= 16384 && hgt <= 16384)) {
    napi_throw_range_error(env, nullptr, "opts.width and opts.height must be in 1..16384");
    return nullptr;
  }
  const uint32_t W = (uint32_t)w, H = (uint32_t)hgt;
  gt_uniforms u;
  gt_uniforms_init(&u, tw, num_prop(env, o, "hFov", 60.0), W, H, (float)num_prop(env, o, "t_min", 1e-3),
                   (float)num_prop(env, o, "t_max", 1e5), (float)num_prop(env, o, "T_min", 0.03),
                   (float)num_prop(env, o, "alpha_min", 0.2), (uint32_t)num_prop(env, o, "hit_array_size", 6),
                   (uint32_t)num_prop(env, o, "sh_degree_max", 3), (int32_t)num_prop(env, o, "earlyStop", 0),
                   (uint32_t)num_prop(env, o, "debugMode", 0));
  /* antialiasing (GS_shader.ts:282-307, compiled out in the reference):
   * jittered rays; with accumulatedFrames > 0 the frame is averaged into
   * opts.prev, the previous average */
  const bool aa = bool_prop(env, o, "antialiasing", false);
  u.accumulated_frames = (uint32_t)num_prop(env, o, "accumulatedFrames", 0);
  /* opts.out: a Float32Array(W*H*4) to render into and return (a render
   * loop reuses it); otherwise a new array */
  float* rgba = nullptr;
  napi_value arr = nullptr;
  bool has_out = false;
  if (napi_has_named_property(env, o, "out", &has_out) == napi_ok && has_out) {
    napi_value ov;
    size_t olen = 0;
    napi_valuetype ty = napi_undefined;
    if (napi_get_named_property(env, o, "out", &ov) == napi_ok) napi_typeof(env, ov, &ty);
    if (ty != napi_undefined && ty != napi_null) {
      if (!get_f32_array(env, ov, &rgba, &olen) || olen != (size_t)W * H * 4) {
        napi_throw_type_error(env, nullptr, "opts.out must be a Float32Array of width * height * 4");
        return nullptr;
      }
      arr = ov;
    }
  }
  if (!arr) {
    arr = new_f32_array(env, (size_t)W * H * 4, &rgba);
    if (!arr) {
      napi_throw_error(env, nullptr, "allocation failed");
      return nullptr;
    }
  }
  if (aa && u.accumulated_frames) {
    bool has = false;
    napi_value pv;
    float* prev = nullptr;
    size_t plen = 0;
    if (napi_has_named_property(env, o, "prev", &has) != napi_ok || !has ||
        napi_get_named_property(env, o, "prev", &pv) != napi_ok || !get_f32_array(env, pv, &prev, &plen) ||
        plen < (size_t)W * H * 4) {
      napi_throw_type_error(env, nullptr, "antialiasing with accumulatedFrames > 0 needs opts.prev: Float32Array(W*H*4)");
      return nullptr;
    }
    if (prev != rgba) std::memmove(rgba, prev, (size_t)W * H * 4 * sizeof(float));
  }
  gt_frame f{};
  f.rgba = rgba;
  f.antialiasing = aa ? 1u : 0u;
  const int rc = gt_render(s, &u, W, H, &f, GT_BUF_HOST, nullptr);
  if (rc != GT_OK) return throw_gt(env, rc);
  return arr;
}

bool get_u8_array(napi_env env, napi_value v, uint8_t** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off = 0;
  void* p = nullptr;
  if (napi_get_typedarray_info(env, v, &t, len, &p, &ab, &off) != napi_ok) return false;
  if (t != napi_uint8_array && t != napi_uint8_clamped_array) return false;
  *data = static_cast<uint8_t*>(p);
  return true;
}

bool get_u32(napi_env env, napi_value v, uint32_t* out) {
  double d = 0;
  if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0.0) || d > 4294967295.0) return false;
  *out = (uint32_t)d;
  return true;
}

/* toneMap(rgba: Float32Array, width, height, mode = 0) -> Uint8Array (8-bit
 * RGBA as the reference's canvas shows the frame, src/toneMapper.ts:11-26) */
napi_value js_tone_map(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  size_t argc = 4;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < 3) {
    napi_throw_type_error(env, nullptr, "toneMap(rgba: Float32Array, width, height, mode?)");
    return nullptr;
  }
  float* rgba = nullptr;
  size_t len = 0;
  uint32_t w = 0, h = 0, mode = 0;
  if (!get_f32_array(env, argv[0], &rgba, &len) || !get_u32(env, argv[1], &w) || !get_u32(env, argv[2], &h) ||
      (argc > 3 && !get_u32(env, argv[3], &mode))) {
    napi_throw_type_error(env, nullptr, "toneMap(rgba: Float32Array, width, height, mode?)");
    return nullptr;
  }
  if (len != (size_t)w * h * 4) {
    napi_throw_range_error(env, nullptr, "toneMap: rgba must hold width * height * 4 floats");
    return nullptr;
  }
  napi_value ab, arr;
  void* p = nullptr;
  NAPI_TRY(env, napi_create_arraybuffer(env, len, &p, &ab));
  NAPI_TRY(env, napi_create_typedarray(env, napi_uint8_array, len, ab, 0, &arr));
  const int rc = gt_tonemap(rgba, w, h, mode, static_cast<uint8_t*>(p), GT_BUF_HOST, nullptr);
  if (rc != GT_OK) return throw_gt(env, rc);
  return arr;
}

/* writePng(path, rgba8: Uint8Array, width, height): the Save button
 * (src/main.ts:606-611) as a file */
napi_value js_write_png(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  char path[4096];
  size_t plen = 0;
  uint8_t* px = nullptr;
  size_t len = 0;
  uint32_t w = 0, h = 0;
  if (napi_get_value_string_utf8(env, argv[0], path, sizeof(path), &plen) != napi_ok ||
      !get_u8_array(env, argv[1], &px, &len) || !get_u32(env, argv[2], &w) || !get_u32(env, argv[3], &h)) {
    napi_throw_type_error(env, nullptr, "writePng(path: string, rgba8: Uint8Array, width, height)");
    return nullptr;
  }
  if (len != (size_t)w * h * 4) {
    napi_throw_range_error(env, nullptr, "writePng: rgba8 must hold width * height * 4 bytes");
    return nullptr;
  }
  const int rc = gt_write_png(path, px, w, h);
  if (rc != GT_OK) return throw_gt(env, rc);
  napi_value undef;
  napi_get_undefined(env, &undef);
  return undef;
}

napi_value init(napi_env env, napi_value exports) {
  const napi_property_descriptor props[] = {
      {"version", nullptr, js_version, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"plyRead", nullptr, js_ply_read, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"generateScene", nullptr, js_generate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"cameraLookAt", nullptr, js_camera_look_at, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"sceneCreate", nullptr, js_scene_create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"sceneDestroy", nullptr, js_scene_destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"sceneSetAlphaMin", nullptr, js_scene_set_alpha_min, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"sceneInfo", nullptr, js_scene_info, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"render", nullptr, js_render, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"toneMap", nullptr, js_tone_map, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"writePng", nullptr, js_write_png, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(gtrace, init)

// Node N-API addon over libxpbi.so: each export wraps one C-ABI entry point of
// include/xpbi.h 1:1 and throws a JavaScript Error carrying xpbi_last_error()
// when the call fails.  The JavaScript Simulation class (../sim.js) keeps the
// reference's web/js/sim.js surface on top of these.
#define NAPI_VERSION 8
#include <node_api.h>

#include <cstring>
#include <string>
#include <vector>

#include "xpbi.h"

namespace {

#define NAPI_CALL(env, call)                                  \
  do {                                                        \
    if ((call) != napi_ok) {                                  \
      napi_throw_error((env), nullptr, "N-API call failed");  \
      return nullptr;                                         \
    }                                                         \
  } while (0)

napi_value throw_xpbi(napi_env env, xpbi_sim* s, int rc, const char* what) {
  std::string m = std::string(what) + " failed (" + std::to_string(rc) + "): " + xpbi_last_error(s);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

xpbi_sim* get_sim(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a simulation handle");
    return nullptr;
  }
  return (xpbi_sim*)p;
}

bool get_bytes(napi_env env, napi_value v, void** data, size_t* len) {
  bool is_ta = false, is_ab = false;
  napi_is_typedarray(env, v, &is_ta);
  if (is_ta) {
    napi_typedarray_type t;
    size_t n = 0, off = 0;
    napi_value ab;
    void* base = nullptr;
    if (napi_get_typedarray_info(env, v, &t, &n, &base, &ab, &off) != napi_ok) return false;
    size_t esz = 1;
    switch (t) {
      case napi_int16_array: case napi_uint16_array: esz = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: esz = 4; break;
      case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: esz = 8; break;
      default: esz = 1;
    }
    *data = base;
    *len = n * esz;
    return true;
  }
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  return false;
}

int32_t arg_i32(napi_env env, napi_value v) {
  int32_t x = 0;
  napi_get_value_int32(env, v, &x);
  return x;
}
double arg_f64(napi_env env, napi_value v) {
  double x = 0;
  napi_get_value_double(env, v, &x);
  return x;
}
napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value r = nullptr;
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return nullptr;
  napi_get_named_property(env, obj, name, &r);
  napi_valuetype t;
  napi_typeof(env, r, &t);
  return (t == napi_undefined || t == napi_null) ? nullptr : r;
}

// a scene descriptor {N, positions:Float32Array, tags?, matIds?:Uint32Array,
// states?, velocities?, h, domainMin:[3], domainMax:[3], device?}; the typed
// arrays stay owned by JavaScript and are read during the create call only
bool parse_desc(napi_env env, napi_value obj, xpbi_scene_desc* d) {
  std::memset(d, 0, sizeof(*d));
  napi_value v;
  if ((v = prop(env, obj, "N"))) d->N = (uint32_t)arg_f64(env, v);
  void* p;
  size_t n;
  struct { const char* name; size_t per; const void** dst; } arrs[] = {
      {"positions", 12, (const void**)&d->positions}, {"tags", 4, (const void**)&d->tags},
      {"matIds", 4, (const void**)&d->matIds},       {"states", 4, (const void**)&d->states},
      {"velocities", 12, (const void**)&d->velocities},
  };
  for (auto& a : arrs) {
    if ((v = prop(env, obj, a.name)) && get_bytes(env, v, &p, &n)) {
      if (n < a.per * d->N) {
        std::string m = std::string(a.name) + " too short";
        napi_throw_range_error(env, nullptr, m.c_str());
        return false;
      }
      *a.dst = p;
    }
  }
  if ((v = prop(env, obj, "h"))) d->h = arg_f64(env, v);
  for (int k = 0; k < 3; k++) {
    napi_value e;
    if ((v = prop(env, obj, "domainMin"))) { napi_get_element(env, v, k, &e); d->domainMin[k] = arg_f64(env, e); }
    if ((v = prop(env, obj, "domainMax"))) { napi_get_element(env, v, k, &e); d->domainMax[k] = arg_f64(env, e); }
  }
  if ((v = prop(env, obj, "device"))) d->device = arg_i32(env, v);
  return true;
}

napi_value wrap_sim(napi_env env, xpbi_sim* s) {
  napi_value out;
  NAPI_CALL(env, napi_create_external(env, s, nullptr, nullptr, &out));
  return out;
}

// create(desc) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 1) { napi_throw_type_error(env, nullptr, "create(desc)"); return nullptr; }
  xpbi_scene_desc d;
  if (!parse_desc(env, a[0], &d)) return nullptr;
  xpbi_sim* s = nullptr;
  int rc = xpbi_create(&d, &s);
  if (rc) return throw_xpbi(env, nullptr, rc, "xpbi_create");
  return wrap_sim(env, s);
}

// createBatch([desc, ...]) -> handle of one simulation holding every scene
napi_value CreateBatch(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  bool isArr = false;
  if (argc < 1 || napi_is_array(env, a[0], &isArr) != napi_ok || !isArr) {
    napi_throw_type_error(env, nullptr, "createBatch([desc, ...])");
    return nullptr;
  }
  uint32_t n = 0;
  napi_get_array_length(env, a[0], &n);
  std::vector<xpbi_scene_desc> d(n);
  for (uint32_t q = 0; q < n; q++) {
    napi_value e;
    napi_get_element(env, a[0], q, &e);
    if (!parse_desc(env, e, &d[q])) return nullptr;
  }
  xpbi_sim* s = nullptr;
  int rc = xpbi_create_batch(d.data(), (int)n, &s);
  if (rc) return throw_xpbi(env, nullptr, rc, "xpbi_create_batch");
  return wrap_sim(env, s);
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (s) xpbi_destroy(s);
  return nullptr;
}

napi_value GridInfo(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  float o[3], cs;
  int32_t dims[3];
  uint32_t nc, np;
  int rc = xpbi_grid_info(s, o, &cs, dims, &nc, &np);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_grid_info");
  napi_value obj, arr, arr2, x;
  napi_create_object(env, &obj);
  napi_create_array_with_length(env, 3, &arr);
  napi_create_array_with_length(env, 3, &arr2);
  for (int k = 0; k < 3; k++) {
    napi_create_double(env, o[k], &x);
    napi_set_element(env, arr, k, x);
    napi_create_int32(env, dims[k], &x);
    napi_set_element(env, arr2, k, x);
  }
  napi_set_named_property(env, obj, "gridOrigin", arr);
  napi_set_named_property(env, obj, "gridDims", arr2);
  napi_create_double(env, cs, &x);
  napi_set_named_property(env, obj, "cellSize", x);
  napi_create_double(env, nc, &x);
  napi_set_named_property(env, obj, "numCells", x);
  napi_create_double(env, np, &x);
  napi_set_named_property(env, obj, "N", x);
  return obj;
}

// writeMaterials(handle, ArrayBuffer|TypedArray of n*64 bytes)
napi_value WriteMaterials(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p;
  size_t n;
  if (argc < 2 || !get_bytes(env, a[1], &p, &n) || n % 64 != 0 || n > 512) {
    napi_throw_type_error(env, nullptr, "materials must be up to 8 records of 64 bytes");
    return nullptr;
  }
  int rc = xpbi_write_materials(s, (const xpbi_material*)p, (int)(n / 64));
  if (rc) return throw_xpbi(env, s, rc, "xpbi_write_materials");
  return nullptr;
}

// writeParams(handle, 256-byte buffer)
napi_value WriteParams(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p;
  size_t n;
  if (argc < 2 || !get_bytes(env, a[1], &p, &n) || n != sizeof(xpbi_params)) {
    napi_throw_type_error(env, nullptr, "params must be a 256-byte buffer");
    return nullptr;
  }
  int rc = xpbi_write_params(s, (const xpbi_params*)p);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_write_params");
  return nullptr;
}

// step(handle, substeps, iterations, hasWater)
napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  bool water = false;
  if (argc > 3) napi_get_value_bool(env, a[3], &water);
  int rc = xpbi_step(s, arg_i32(env, a[1]), arg_i32(env, a[2]), water ? 1 : 0);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_step");
  return nullptr;
}

// read(handle, field) -> ArrayBuffer
napi_value Read(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int field = arg_i32(env, a[1]);
  float o[3], cs;
  int32_t dims[3];
  uint32_t nc = 0, np = 0;
  xpbi_grid_info(s, o, &cs, dims, &nc, &np);
  int scenes = 1;
  xpbi_scene_count(s, &scenes);
  if (scenes > 1) {  // a batch: every scene's particles and cells
    uint32_t ncq, npq, fp, fc;
    nc = np = 0;
    for (int q = 0; q < scenes; q++) {
      xpbi_scene_info(s, q, o, &cs, dims, &ncq, &npq, &fp, &fc);
      nc += ncq;
      np += npq;
    }
  }
  size_t bytes;
  switch (field) {
    case XPBI_FIELD_POS: case XPBI_FIELD_VEL: case XPBI_FIELD_AUX: bytes = 16ull * np; break;
    case XPBI_FIELD_F: case XPBI_FIELD_L: case XPBI_FIELD_SP: bytes = 48ull * np; break;
    case XPBI_FIELD_CELLSTART: bytes = 4ull * nc; break;
    case XPBI_FIELD_PERM: case XPBI_FIELD_ID: bytes = 4ull * np; break;
    case XPBI_FIELD_MATS: bytes = 512ull * (size_t)scenes; break;
    default: napi_throw_range_error(env, nullptr, "unknown field"); return nullptr;
  }
  void* data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_create_arraybuffer(env, bytes, &data, &ab));
  int rc = xpbi_read(s, field, data, bytes);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_read");
  return ab;
}

// write(handle, field, TypedArray)
napi_value Write(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p;
  size_t n;
  if (argc < 3 || !get_bytes(env, a[2], &p, &n)) {
    napi_throw_type_error(env, nullptr, "write(handle, field, typedArray)");
    return nullptr;
  }
  int rc = xpbi_write(s, arg_i32(env, a[1]), p, n);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_write");
  return nullptr;
}

// profileSubstep(handle, iterations, hasWater) -> [[label, ms], ...]
napi_value ProfileSubstep(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  bool water = false;
  if (argc > 2) napi_get_value_bool(env, a[2], &water);
  std::vector<float> ms(4096);
  int cnt = 0;
  int rc = xpbi_profile_substep(s, arg_i32(env, a[1]), water ? 1 : 0, ms.data(), (int)ms.size(), &cnt);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_profile_substep");
  if (cnt > (int)ms.size()) cnt = (int)ms.size();
  napi_value arr;
  napi_create_array_with_length(env, cnt, &arr);
  for (int k = 0; k < cnt; k++) {
    napi_value pair, lab, t;
    napi_create_array_with_length(env, 2, &pair);
    napi_create_string_utf8(env, xpbi_profile_label(s, k), NAPI_AUTO_LENGTH, &lab);
    napi_create_double(env, ms[k], &t);
    napi_set_element(env, pair, 0, lab);
    napi_set_element(env, pair, 1, t);
    napi_set_element(env, arr, k, pair);
  }
  return arr;
}

// timeSubsteps(handle, substeps, iterations, hasWater) -> ms
napi_value TimeSubsteps(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  bool water = false;
  if (argc > 3) napi_get_value_bool(env, a[3], &water);
  float ms = 0;
  int rc = xpbi_time_substeps(s, arg_i32(env, a[1]), arg_i32(env, a[2]), water ? 1 : 0, &ms);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_time_substeps");
  napi_value out;
  napi_create_double(env, ms, &out);
  return out;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int rc = xpbi_sync(s);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_sync");
  return nullptr;
}

// writeSceneMaterials(handle, scene, bytes) / writeSceneParams(handle, scene, buf256)
napi_value WriteSceneMaterials(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p;
  size_t n;
  if (argc < 3 || !get_bytes(env, a[2], &p, &n) || n % 64 != 0 || n > 512) {
    napi_throw_type_error(env, nullptr, "materials must be up to 8 records of 64 bytes");
    return nullptr;
  }
  int rc = xpbi_write_scene_materials(s, arg_i32(env, a[1]), (const xpbi_material*)p, (int)(n / 64));
  if (rc) return throw_xpbi(env, s, rc, "xpbi_write_scene_materials");
  return nullptr;
}
napi_value WriteSceneParams(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p;
  size_t n;
  if (argc < 3 || !get_bytes(env, a[2], &p, &n) || n != sizeof(xpbi_params)) {
    napi_throw_type_error(env, nullptr, "params must be a 256-byte buffer");
    return nullptr;
  }
  int rc = xpbi_write_scene_params(s, arg_i32(env, a[1]), (const xpbi_params*)p);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_write_scene_params");
  return nullptr;
}

// sceneInfo(handle, scene) -> {gridOrigin, cellSize, gridDims, numCells, N, firstParticle, firstCell}
napi_value SceneInfo(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  float o[3], cs;
  int32_t dims[3];
  uint32_t nc, np, fp, fc;
  int rc = xpbi_scene_info(s, arg_i32(env, a[1]), o, &cs, dims, &nc, &np, &fp, &fc);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_scene_info");
  napi_value obj, arr, arr2, x;
  napi_create_object(env, &obj);
  napi_create_array_with_length(env, 3, &arr);
  napi_create_array_with_length(env, 3, &arr2);
  for (int k = 0; k < 3; k++) {
    napi_create_double(env, o[k], &x);
    napi_set_element(env, arr, k, x);
    napi_create_int32(env, dims[k], &x);
    napi_set_element(env, arr2, k, x);
  }
  napi_set_named_property(env, obj, "gridOrigin", arr);
  napi_set_named_property(env, obj, "gridDims", arr2);
  const struct { const char* k; double v; } nums[] = {
      {"cellSize", cs}, {"numCells", (double)nc}, {"N", (double)np}, {"firstParticle", (double)fp},
      {"firstCell", (double)fc}};
  for (auto& e : nums) {
    napi_create_double(env, e.v, &x);
    napi_set_named_property(env, obj, e.k, x);
  }
  return obj;
}

// setTolerance(handle, tol, checkEvery)
napi_value SetTolerance(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int rc = xpbi_set_tolerance(s, (float)arg_f64(env, a[1]), argc > 2 ? arg_i32(env, a[2]) : 4);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_set_tolerance");
  return nullptr;
}

// lastIterations(handle) -> {iterations, residual}
napi_value LastIterations(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int it = 0;
  double r = 0.0;
  int rc = xpbi_last_iterations(s, &it, &r);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_last_iterations");
  napi_value obj, x;
  napi_create_object(env, &obj);
  napi_create_int32(env, it, &x);
  napi_set_named_property(env, obj, "iterations", x);
  napi_create_double(env, r, &x);
  napi_set_named_property(env, obj, "residual", x);
  return obj;
}

// setWarmstart(handle, Float32Array(5122) | null)
napi_value SetWarmstart(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (argc > 1) {
    napi_valuetype t;
    napi_typeof(env, a[1], &t);
    if (t != napi_null && t != napi_undefined && !get_bytes(env, a[1], &p, &n)) {
      napi_throw_type_error(env, nullptr, "setWarmstart(handle, Float32Array | null)");
      return nullptr;
    }
  }
  int rc = xpbi_set_warmstart(s, (const float*)p, n / 4);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_set_warmstart");
  return nullptr;
}

// setLambda0(handle, Float32Array(N) | null): the given start per particle id
napi_value SetLambda0(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (argc > 1) {
    napi_valuetype t;
    napi_typeof(env, a[1], &t);
    if (t != napi_null && t != napi_undefined && !get_bytes(env, a[1], &p, &n)) {
      napi_throw_type_error(env, nullptr, "setLambda0(handle, Float32Array | null)");
      return nullptr;
    }
  }
  int rc = xpbi_set_lambda0(s, (const float*)p, n / 4);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_set_lambda0");
  return nullptr;
}

// graphStep(handle, substeps, iterations, hasWater)
napi_value GraphStep(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  bool water = false;
  if (argc > 3) napi_get_value_bool(env, a[3], &water);
  int rc = xpbi_graph_step(s, arg_i32(env, a[1]), arg_i32(env, a[2]), water ? 1 : 0);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_graph_step");
  return nullptr;
}

// setLaunch(handle, blocks) ; getLaunch(handle) -> {blocks, workEntries}
napi_value SetLaunch(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int rc = xpbi_set_launch(s, arg_i32(env, a[1]));
  if (rc) return throw_xpbi(env, s, rc, "xpbi_set_launch");
  return nullptr;
}
// setNeighbourSlots(handle, slots)
napi_value SetNeighbourSlots(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int rc = xpbi_set_neighbour_slots(s, arg_i32(env, a[1]));
  if (rc) return throw_xpbi(env, s, rc, "xpbi_set_neighbour_slots");
  return nullptr;
}
napi_value GetLaunch(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  int b = 0, w = 0;
  int rc = xpbi_get_launch(s, &b, &w);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_get_launch");
  napi_value obj, x;
  napi_create_object(env, &obj);
  napi_create_int32(env, b, &x);
  napi_set_named_property(env, obj, "blocks", x);
  napi_create_int32(env, w, &x);
  napi_set_named_property(env, obj, "workEntries", x);
  return obj;
}

// autotune(handle, iterations, hasWater, substeps) -> {blocks, ms}
napi_value Autotune(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  xpbi_sim* s = get_sim(env, a[0]);
  if (!s) return nullptr;
  bool water = false;
  if (argc > 2) napi_get_value_bool(env, a[2], &water);
  int blocks = 0;
  float ms = 0.0f;
  int rc = xpbi_autotune(s, arg_i32(env, a[1]), water ? 1 : 0, argc > 3 ? arg_i32(env, a[3]) : 2, &blocks, &ms);
  if (rc) return throw_xpbi(env, s, rc, "xpbi_autotune");
  napi_value obj, x;
  napi_create_object(env, &obj);
  napi_create_int32(env, blocks, &x);
  napi_set_named_property(env, obj, "blocks", x);
  napi_create_double(env, ms, &x);
  napi_set_named_property(env, obj, "ms", x);
  return obj;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
      {"create", Create},         {"destroy", Destroy},          {"gridInfo", GridInfo},
      {"writeMaterials", WriteMaterials}, {"writeParams", WriteParams}, {"step", Step},
      {"read", Read},             {"write", Write},              {"profileSubstep", ProfileSubstep},
      {"timeSubsteps", TimeSubsteps}, {"sync", Sync},
      {"createBatch", CreateBatch}, {"sceneInfo", SceneInfo}, {"writeSceneMaterials", WriteSceneMaterials},
      {"writeSceneParams", WriteSceneParams}, {"setTolerance", SetTolerance}, {"lastIterations", LastIterations},
      {"setWarmstart", SetWarmstart}, {"setLambda0", SetLambda0}, {"graphStep", GraphStep}, {"setLaunch", SetLaunch},
      {"getLaunch", GetLaunch}, {"autotune", Autotune}, {"setNeighbourSlots", SetNeighbourSlots},
  };
  for (auto& f : fns) {
    napi_value fn;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn);
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(xpbi_napi, Init)

// gs_napi.cc — Node.js N-API binding of the trainer's C ABI
// (include/splat_gs.h). One JS function per C entry point; the JavaScript
// shim (splat.js_amd/js/trainer.mjs) builds the GSTrainer surface on top.
//
// The library is opened with dlopen from ../libsplatgs.so, so loading the
// addon does not touch the GPU and a missing library is reported as a JS
// Error naming the path.
//
// Readbacks (readLossAsync, readGaussiansAsync, evalCamPsnrAsync) return
// Promises: the JS thread enqueues the copy (gs_*_async) and a
// napi_async_work waits for it on a worker thread, so the event loop keeps
// running and stepOnce keeps enqueueing meanwhile. Errors carry a `code`
// ("GS_E_HIP" for device errors, which the shim maps to device loss).
#include <dlfcn.h>
#include <node_api.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "splat_gs.h"

namespace {

struct Api {
  void* so = nullptr;
  std::string err;
  int (*default_opts)(gs_opts*);
  int (*create)(const gs_opts*, gs_trainer_t**);
  void (*destroy)(gs_trainer_t*);
  int (*setup)(gs_trainer_t*, const float*, int64_t, const float*, const gs_camera*, int32_t,
               const gs_image*, int32_t, int32_t, int32_t, float);
  int (*step)(gs_trainer_t*, int32_t);
  int (*step_views)(gs_trainer_t*, const int32_t*, int32_t);
  int (*read_loss)(gs_trainer_t*, double*, int32_t*);
  int (*render)(gs_trainer_t*, const gs_camera*, int32_t, int64_t, float*);
  int (*read_gaussians)(gs_trainer_t*, float*, float*);
  int (*eval_cam_psnr)(gs_trainer_t*, int32_t, double*);
  int (*eval_cam_psnr_refined)(gs_trainer_t*, int32_t, int32_t, double*);
  int (*get_camera)(gs_trainer_t*, int32_t, gs_cam_state*);
  int (*read_cam_grads)(gs_trainer_t*, int32_t*, int32_t);
  int (*set_holdout)(gs_trainer_t*, int32_t);
  int (*set_excluded)(gs_trainer_t*, const int32_t*, int32_t);
  int (*set_iter)(gs_trainer_t*, int64_t);
  int (*get_info)(gs_trainer_t*, gs_info*);
  int (*read_step_stats)(gs_trainer_t*, gs_step_stats*);
  int (*read_loss_async)(gs_trainer_t*, gs_pending_t**);
  int (*eval_cam_psnr_async)(gs_trainer_t*, int32_t, gs_pending_t**);
  int (*read_gaussians_async)(gs_trainer_t*, float*, float*, gs_pending_t**);
  int (*pending_wait)(gs_pending_t*, gs_readback_result*);
  const char* (*pending_error)(gs_pending_t*);
  void (*pending_free)(gs_pending_t*);
  int (*sync)(gs_trainer_t*);
  int (*refine)(gs_trainer_t*, const gs_refine_opts*, gs_refine_result*);
  int (*export_ply)(gs_trainer_t*, int32_t, void*, int64_t, int64_t*);
  const char* (*last_error)(gs_trainer_t*);
  int (*set_raster_shape)(gs_trainer_t*, int32_t, int32_t);
  int (*autotune)(gs_trainer_t*, const int32_t*, int32_t, int32_t, gs_tune_entry*);
};

Api g_api;

std::string addon_dir() {
  Dl_info info;
  if (dladdr((void*)&addon_dir, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t s = p.rfind('/');
    return s == std::string::npos ? "." : p.substr(0, s);
  }
  return ".";
}

bool load_api() {
  if (g_api.so) return true;
  const std::string path = addon_dir() + "/../libsplatgs.so";
  void* so = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!so) {
    g_api.err = std::string("cannot load ") + path + ": " + dlerror();
    return false;
  }
#define SYM(field, name)                                              \
  g_api.field = reinterpret_cast<decltype(g_api.field)>(dlsym(so, name)); \
  if (!g_api.field) {                                                 \
    g_api.err = std::string("missing symbol ") + name;                \
    return false;                                                     \
  }
  SYM(default_opts, "gs_default_opts");
  SYM(create, "gs_create");
  SYM(destroy, "gs_destroy");
  SYM(setup, "gs_setup");
  SYM(step, "gs_step");
  SYM(step_views, "gs_step_views");
  SYM(set_raster_shape, "gs_set_raster_shape");
  SYM(autotune, "gs_autotune");
  SYM(read_loss, "gs_read_loss");
  SYM(render, "gs_render");
  SYM(read_gaussians, "gs_read_gaussians");
  SYM(eval_cam_psnr, "gs_eval_cam_psnr");
  SYM(eval_cam_psnr_refined, "gs_eval_cam_psnr_refined");
  SYM(get_camera, "gs_get_camera");
  SYM(read_cam_grads, "gs_read_cam_grads");
  SYM(set_holdout, "gs_set_holdout");
  SYM(set_excluded, "gs_set_excluded");
  SYM(set_iter, "gs_set_iter");
  SYM(get_info, "gs_get_info");
  SYM(read_step_stats, "gs_read_step_stats");
  SYM(read_loss_async, "gs_read_loss_async");
  SYM(eval_cam_psnr_async, "gs_eval_cam_psnr_async");
  SYM(read_gaussians_async, "gs_read_gaussians_async");
  SYM(pending_wait, "gs_pending_wait");
  SYM(pending_error, "gs_pending_error");
  SYM(pending_free, "gs_pending_free");
  SYM(sync, "gs_sync");
  SYM(refine, "gs_refine");
  SYM(export_ply, "gs_export_ply");
  SYM(last_error, "gs_last_error");
#undef SYM
  g_api.so = so;
  return true;
}

napi_value throw_err(napi_env env, const std::string& msg) {
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

const char* code_name(int rc) {
  switch (rc) {
    case GS_E_ARG: return "GS_E_ARG";
    case GS_E_HIP: return "GS_E_HIP";
    case GS_E_NOMEM: return "GS_E_NOMEM";
    case GS_E_LIMIT: return "GS_E_LIMIT";
    default: return "GS_E_UNKNOWN";
  }
}

bool check(napi_env env, gs_trainer_t* h, int rc) {
  if (rc == 0) return true;
  std::string msg = h ? g_api.last_error(h) : "trainer destroyed";
  if (msg.empty()) msg = "splat_gs error " + std::to_string(rc);
  napi_throw_error(env, code_name(rc), msg.c_str());
  return false;
}

napi_value make_error(napi_env env, int rc, const std::string& msg) {
  napi_value code, text, err;
  napi_create_string_utf8(env, code_name(rc), NAPI_AUTO_LENGTH, &code);
  napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &text);
  napi_create_error(env, code, text, &err);
  return err;
}

#define NAPI_CALL(env, call)                          \
  do {                                                \
    if ((call) != napi_ok) {                          \
      throw_err(env, "N-API call failed: " #call);    \
      return nullptr;                                 \
    }                                                 \
  } while (0)

double get_num(napi_env env, napi_value obj, const char* key, double dflt, bool* present = nullptr) {
  bool has = false;
  if (present) *present = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    if (present) *present = true;
    return b ? 1.0 : 0.0;
  }
  if (t != napi_number) return dflt;
  double d;
  napi_get_value_double(env, v, &d);
  if (present) *present = true;
  return d;
}

void set_num(napi_env env, napi_value obj, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}

bool get_f32(napi_env env, napi_value v, float** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  void* p;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, len, &p, &ab, &off);
  if (t != napi_float32_array) return false;
  *data = (float*)p;
  return true;
}

// The JS handle is an external around a box, so destroy() can free the
// trainer at once and the finalizer then finds nothing left to free.
struct Box {
  gs_trainer_t* h;
};

gs_trainer_t* get_handle(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) return nullptr;
  return ((Box*)p)->h;
}

bool read_camera(napi_env env, napi_value c, gs_camera* out) {
  memset(out, 0, sizeof(*out));
  napi_value R, t;
  if (napi_get_named_property(env, c, "R", &R) != napi_ok) return false;
  if (napi_get_named_property(env, c, "t", &t) != napi_ok) return false;
  for (uint32_t i = 0; i < 9; i++) {
    napi_value e;
    double d = 0;
    if (napi_get_element(env, R, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) return false;
    out->R[i] = (float)d;
  }
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    double d = 0;
    if (napi_get_element(env, t, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) return false;
    out->t[i] = (float)d;
  }
  out->f = (float)get_num(env, c, "f", 0);
  out->cx = (float)get_num(env, c, "cx", 0);
  out->cy = (float)get_num(env, c, "cy", 0);
  out->w = (int32_t)get_num(env, c, "w", 0);
  out->h = (int32_t)get_num(env, c, "h", 0);
  out->imgIdx = (int32_t)get_num(env, c, "imgIdx", 0);
  out->g = (float)get_num(env, c, "g", 0);
  out->b = (float)get_num(env, c, "b", 0);
  return true;
}

void finalize_trainer(napi_env, void* data, void*) {
  Box* b = (Box*)data;
  if (!b) return;
  if (b->h && g_api.destroy) g_api.destroy(b->h);
  delete b;
}

// create(opts) -> external
napi_value Create(napi_env env, napi_callback_info info) {
  if (!load_api()) return throw_err(env, g_api.err);
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_opts o;
  g_api.default_opts(&o);
  if (argc >= 1) {
    napi_valuetype t;
    napi_typeof(env, argv[0], &t);
    if (t == napi_object) {
      napi_value a = argv[0];
      o.eCut = (float)get_num(env, a, "eCut", o.eCut);
      o.aMin = (float)get_num(env, a, "aMin", o.aMin);
      o.radClamp = (float)get_num(env, a, "radClamp", o.radClamp);
      o.shDeg = (int32_t)get_num(env, a, "shDeg", o.shDeg);
      o.anisoReg = (float)get_num(env, a, "anisoReg", o.anisoReg);
      o.opacityReg = (float)get_num(env, a, "opacityReg", o.opacityReg);
      o.shLr = (float)get_num(env, a, "shLr", o.shLr);
      o.minScale = (float)get_num(env, a, "minScale", o.minScale);
      o.maxIters = (int32_t)get_num(env, a, "maxIters", o.maxIters);
      o.maxSplats = (int64_t)get_num(env, a, "maxSplats", (double)o.maxSplats);
      o.capMult = (float)get_num(env, a, "capMult", o.capMult);
      o.entriesCap = (int64_t)get_num(env, a, "entriesCap", (double)o.entriesCap);
      o.tileGrad = (int32_t)get_num(env, a, "tileGrad", o.tileGrad);
      o.posLr = (float)get_num(env, a, "posLr", o.posLr);
      o.loss = (int32_t)get_num(env, a, "loss", o.loss);
      o.ssimLambda = (float)get_num(env, a, "ssimLambda", o.ssimLambda);
      o.gradAccum = (int32_t)get_num(env, a, "gradAccum", o.gradAccum);
      o.sortMode = (int32_t)get_num(env, a, "sortMode", o.sortMode);
      o.seed = (uint32_t)get_num(env, a, "seed", o.seed);
      o.device = (int32_t)get_num(env, a, "device", o.device);
      o.precision = (int32_t)get_num(env, a, "precision", o.precision);
      o.camOpt = get_num(env, a, "camOpt", o.camOpt) != 0.0 ? 1 : 0;
      o.expComp = get_num(env, a, "expComp", o.expComp) != 0.0 ? 1 : 0;
      o.viewsPerStep = (int32_t)get_num(env, a, "viewsPerStep", o.viewsPerStep);
      o.graph = get_num(env, a, "graph", o.graph) != 0.0 ? 1 : 0;
      o.viewsLaunch = (int32_t)get_num(env, a, "viewsLaunch", o.viewsLaunch);
      o.tile = (int32_t)get_num(env, a, "tile", o.tile);
      o.batch = (int32_t)get_num(env, a, "batch", o.batch);
    }
  }
  gs_trainer_t* h = nullptr;
  const int rc = g_api.create(&o, &h);
  if (rc != 0) return throw_err(env, "GSTrainer.create: no usable HIP device (gs_create returned " + std::to_string(rc) + ")");
  napi_value ext;
  Box* box = new Box{h};
  NAPI_CALL(env, napi_create_external(env, box, finalize_trainer, nullptr, &ext));
  return ext;
}

// destroy(h): frees the device state now (trainer.js's device.destroy())
napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* p = nullptr;
  if (argc < 1 || napi_get_value_external(env, argv[0], &p) != napi_ok || !p) return nullptr;
  Box* b = (Box*)p;
  if (b->h) g_api.destroy(b->h);
  b->h = nullptr;
  return nullptr;
}

// setup(h, params, n, sh|null, cams[], images[], maxW, maxH, radius)
napi_value Setup(napi_env env, napi_callback_info info) {
  size_t argc = 9;
  napi_value argv[9];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 9) return throw_err(env, "setup: expected 9 arguments");
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return throw_err(env, "setup: trainer destroyed");
  float* params = nullptr;
  size_t plen = 0;
  if (!get_f32(env, argv[1], &params, &plen)) return throw_err(env, "setup: gaussians.data must be a Float32Array");
  int64_t n = 0;
  NAPI_CALL(env, napi_get_value_int64(env, argv[2], &n));
  if ((int64_t)plen < n * 16) return throw_err(env, "setup: gaussians.data shorter than n * 16");
  float* sh = nullptr;
  size_t shlen = 0;
  napi_valuetype st;
  napi_typeof(env, argv[3], &st);
  if (st != napi_null && st != napi_undefined && !get_f32(env, argv[3], &sh, &shlen))
    return throw_err(env, "setup: sh must be a Float32Array or null");
  if (sh) {
    // the trainer reads n * 3K coefficients of its own SH degree
    gs_info in;
    if (!check(env, h, g_api.get_info(h, &in))) return nullptr;
    const int64_t want = n * 3 * (int64_t)in.shK;
    if ((int64_t)shlen != want)
      return throw_err(env, "setup: gaussians.sh holds " + std::to_string(shlen) + " floats; SH degree " +
                                std::to_string(in.shDeg) + " needs n * 3 * " + std::to_string(in.shK) + " = " +
                                std::to_string(want));
    if (want == 0) sh = nullptr;
  }
  uint32_t ncam = 0, nimg = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[4], &ncam));
  NAPI_CALL(env, napi_get_array_length(env, argv[5], &nimg));
  std::vector<gs_camera> cams(ncam);
  for (uint32_t i = 0; i < ncam; i++) {
    napi_value c;
    NAPI_CALL(env, napi_get_element(env, argv[4], i, &c));
    if (!read_camera(env, c, &cams[i])) return throw_err(env, "setup: bad camera " + std::to_string(i));
  }
  std::vector<gs_image> ims(nimg);
  for (uint32_t i = 0; i < nimg; i++) {
    napi_value im, rgb;
    NAPI_CALL(env, napi_get_element(env, argv[5], i, &im));
    ims[i].tw = (int32_t)get_num(env, im, "tw", 0);
    ims[i].th = (int32_t)get_num(env, im, "th", 0);
    NAPI_CALL(env, napi_get_named_property(env, im, "rgb", &rgb));
    float* p = nullptr;
    size_t len = 0;
    if (!get_f32(env, rgb, &p, &len) || (int64_t)len < (int64_t)ims[i].tw * ims[i].th * 3)
      return throw_err(env, "setup: images[" + std::to_string(i) + "].rgb must be a Float32Array of tw*th*3");
    ims[i].rgb = p;
  }
  int32_t mw, mh;
  double rad;
  NAPI_CALL(env, napi_get_value_int32(env, argv[6], &mw));
  NAPI_CALL(env, napi_get_value_int32(env, argv[7], &mh));
  NAPI_CALL(env, napi_get_value_double(env, argv[8], &rad));
  if (!check(env, h, g_api.setup(h, params, n, sh, cams.data(), (int32_t)ncam, ims.data(), (int32_t)nimg, mw, mh,
                                 (float)rad)))
    return nullptr;
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  int32_t cam = -1;
  if (argc >= 2) napi_get_value_int32(env, argv[1], &cam);
  check(env, h, g_api.step(h, cam));
  return nullptr;
}

// stepViews(h, cams[]): one batched step over those views (-1 picks)
napi_value StepViews(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return nullptr;
  uint32_t len = 0;
  if (argc < 2 || napi_get_array_length(env, argv[1], &len) != napi_ok) return throw_err(env, "stepViews: cams array");
  std::vector<int32_t> cams(len);
  for (uint32_t k = 0; k < len; k++) {
    napi_value e;
    napi_get_element(env, argv[1], k, &e);
    napi_get_value_int32(env, e, &cams[k]);
  }
  check(env, h, g_api.step_views(h, cams.data(), (int32_t)len));
  return nullptr;
}

// setRasterShape(h, tile, batch)
napi_value SetRasterShape(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return nullptr;
  int32_t tile = 0, batch = 0;
  if (argc < 3 || napi_get_value_int32(env, argv[1], &tile) != napi_ok ||
      napi_get_value_int32(env, argv[2], &batch) != napi_ok)
    return throw_err(env, "setRasterShape: tile, batch");
  check(env, h, g_api.set_raster_shape(h, tile, batch));
  return nullptr;
}

// autotune(h, cams[], reps) -> [{tile, batch, status, ms, keys, dropped}] (9 shapes)
napi_value Autotune(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return nullptr;
  uint32_t len = 0;
  int32_t reps = 5;
  if (argc < 2 || napi_get_array_length(env, argv[1], &len) != napi_ok) return throw_err(env, "autotune: cams array");
  if (argc >= 3) napi_get_value_int32(env, argv[2], &reps);
  std::vector<int32_t> cams(len);
  for (uint32_t k = 0; k < len; k++) {
    napi_value e;
    napi_get_element(env, argv[1], k, &e);
    napi_get_value_int32(env, e, &cams[k]);
  }
  gs_tune_entry out[9];
  if (!check(env, h, g_api.autotune(h, cams.data(), (int32_t)len, reps, out))) return nullptr;
  napi_value arr;
  napi_create_array_with_length(env, 9, &arr);
  for (uint32_t i = 0; i < 9; i++) {
    napi_value o, v;
    napi_create_object(env, &o);
    napi_create_int32(env, out[i].tile, &v);
    napi_set_named_property(env, o, "tile", v);
    napi_create_int32(env, out[i].batch, &v);
    napi_set_named_property(env, o, "batch", v);
    napi_create_int32(env, out[i].status, &v);
    napi_set_named_property(env, o, "status", v);
    napi_create_double(env, out[i].ms, &v);
    napi_set_named_property(env, o, "ms", v);
    napi_create_double(env, (double)out[i].keys, &v);
    napi_set_named_property(env, o, "keys", v);
    napi_create_double(env, (double)out[i].dropped, &v);
    napi_set_named_property(env, o, "dropped", v);
    napi_set_element(env, arr, i, o);
  }
  return arr;
}

// readLoss(h) -> number | null
napi_value ReadLoss(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  double mse = 0;
  int32_t has = 0;
  if (!check(env, h, g_api.read_loss(h, &mse, &has))) return nullptr;
  napi_value r;
  if (!has) NAPI_CALL(env, napi_get_null(env, &r));
  else NAPI_CALL(env, napi_create_double(env, mse, &r));
  return r;
}

// render(h, cam, trainMode, offset, Float32Array out)
napi_value Render(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  gs_camera c;
  if (!read_camera(env, argv[1], &c)) return throw_err(env, "renderView: bad camera");
  int32_t tm = 0;
  int64_t off = 0;
  napi_get_value_int32(env, argv[2], &tm);
  napi_get_value_int64(env, argv[3], &off);
  float* out = nullptr;
  size_t len = 0;
  if (!get_f32(env, argv[4], &out, &len) || (int64_t)len < (int64_t)c.w * c.h * 4)
    return throw_err(env, "renderView: output must be a Float32Array of w*h*4");
  check(env, h, g_api.render(h, &c, tm, off, out));
  return nullptr;
}

// readGaussians(h, Float32Array params, Float32Array|null sh)
napi_value ReadGaussians(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  gs_info in;
  if (!check(env, h, g_api.get_info(h, &in))) return nullptr;
  float *p = nullptr, *sh = nullptr;
  size_t lp = 0, ls = 0;
  if (!get_f32(env, argv[1], &p, &lp) || (int64_t)lp < in.n * 16)
    return throw_err(env, "readGaussians: params must be a Float32Array of n*16");
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  if (t != napi_null && t != napi_undefined) {
    if (!get_f32(env, argv[2], &sh, &ls) || (int64_t)ls < in.n * 3 * in.shK)
      return throw_err(env, "readGaussians: sh must be a Float32Array of n*3*shK");
  }
  check(env, h, g_api.read_gaussians(h, p, sh));
  return nullptr;
}

napi_value EvalCamPsnr(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  int32_t ci = 0;
  napi_get_value_int32(env, argv[1], &ci);
  double v = 0;
  if (!check(env, h, g_api.eval_cam_psnr(h, ci, &v))) return nullptr;
  napi_value r;
  NAPI_CALL(env, napi_create_double(env, v, &r));
  return r;
}

// evalCamPsnrRefined(h, ci, steps) -> psnr
napi_value EvalCamPsnrRefined(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  int32_t ci = 0, steps = 200;
  napi_get_value_int32(env, argv[1], &ci);
  if (argc > 2) napi_get_value_int32(env, argv[2], &steps);
  double v = 0;
  if (!check(env, h, g_api.eval_cam_psnr_refined(h, ci, steps, &v))) return nullptr;
  napi_value r;
  NAPI_CALL(env, napi_create_double(env, v, &r));
  return r;
}

// getCamera(h, ci) -> {R: number[9], t: number[3], f, g, b}
napi_value GetCamera(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  int32_t ci = 0;
  napi_get_value_int32(env, argv[1], &ci);
  gs_cam_state st;
  if (!check(env, h, g_api.get_camera(h, ci, &st))) return nullptr;
  napi_value o, R, t;
  NAPI_CALL(env, napi_create_object(env, &o));
  NAPI_CALL(env, napi_create_array_with_length(env, 9, &R));
  NAPI_CALL(env, napi_create_array_with_length(env, 3, &t));
  for (uint32_t k = 0; k < 9; k++) {
    napi_value v;
    napi_create_double(env, st.R[k], &v);
    napi_set_element(env, R, k, v);
  }
  for (uint32_t k = 0; k < 3; k++) {
    napi_value v;
    napi_create_double(env, st.t[k], &v);
    napi_set_element(env, t, k, v);
  }
  napi_set_named_property(env, o, "R", R);
  napi_set_named_property(env, o, "t", t);
  set_num(env, o, "f", st.f);
  set_num(env, o, "g", st.g);
  set_num(env, o, "b", st.b);
  return o;
}

// readCamGrads(h, Int32Array (numCams + 1) * 8)
napi_value ReadCamGrads(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  napi_typedarray_type ty;
  size_t len = 0;
  void* data = nullptr;
  if (napi_get_typedarray_info(env, argv[1], &ty, &len, &data, nullptr, nullptr) != napi_ok || ty != napi_int32_array)
    return throw_err(env, "readCamGrads: expected an Int32Array");
  check(env, h, g_api.read_cam_grads(h, (int32_t*)data, (int32_t)len));
  return nullptr;
}

napi_value SetHoldout(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t ci = -1;
  napi_get_value_int32(env, argv[1], &ci);
  gs_trainer_t* h = get_handle(env, argv[0]);
  check(env, h, g_api.set_holdout(h, ci));
  return nullptr;
}

napi_value SetExcluded(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  uint32_t n = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[1], &n));
  std::vector<int32_t> v(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    NAPI_CALL(env, napi_get_element(env, argv[1], i, &e));
    napi_get_value_int32(env, e, &v[i]);
  }
  check(env, h, g_api.set_excluded(h, v.data(), (int32_t)n));
  return nullptr;
}

napi_value SetIter(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  int64_t it = 0;
  napi_get_value_int64(env, argv[1], &it);
  check(env, h, g_api.set_iter(h, it));
  return nullptr;
}


napi_value Info(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  gs_info in;
  if (!check(env, h, g_api.get_info(h, &in))) return nullptr;
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "n", (double)in.n);
  set_num(env, o, "cap", (double)in.cap);
  set_num(env, o, "iter", (double)in.iter);
  set_num(env, o, "shDeg", in.shDeg);
  set_num(env, o, "shK", in.shK);
  set_num(env, o, "stride", in.stride);
  set_num(env, o, "numCams", in.numCams);
  set_num(env, o, "lastCam", in.lastCam);
  set_num(env, o, "holdout", in.holdout);
  set_num(env, o, "entriesCap", (double)in.entriesCap);
  set_num(env, o, "maxTiles", in.maxTiles);
  set_num(env, o, "horizon", in.horizon);
  return o;
}

// stepStats(h) -> {lastKeys, lastVisible, overflowKeys} (waits for the GPU)
napi_value StepStats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return throw_err(env, "stepStats: trainer destroyed");
  gs_step_stats s;
  if (!check(env, h, g_api.read_step_stats(h, &s))) return nullptr;
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "lastKeys", (double)s.lastKeys);
  set_num(env, o, "lastVisible", (double)s.lastVisible);
  set_num(env, o, "overflowKeys", (double)s.overflowKeys);
  return o;
}

// ---- Promise readbacks ----
struct Job {
  enum Kind { kLoss, kPsnr, kGaussians } kind;
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  gs_pending_t* pending = nullptr;
  napi_ref keep[2] = {nullptr, nullptr};   // destination arrays of readGaussians
  int rc = 0;
  gs_readback_result r{};
  std::string err;
};

void job_execute(napi_env, void* data) {
  Job* j = (Job*)data;
  j->rc = g_api.pending_wait(j->pending, &j->r);
  if (j->rc != 0) j->err = g_api.pending_error(j->pending);
}

void job_complete(napi_env env, napi_status status, void* data) {
  Job* j = (Job*)data;
  if (status != napi_ok && j->rc == 0) {
    j->rc = GS_E_ARG;
    j->err = "readback cancelled";
  }
  if (j->rc != 0) {
    napi_reject_deferred(env, j->deferred, make_error(env, j->rc, j->err.empty() ? "readback failed" : j->err));
  } else {
    napi_value v;
    if (j->kind == Job::kLoss) {
      // {mse: number | null, overflow: number}
      napi_create_object(env, &v);
      napi_value m;
      if (j->r.has) napi_create_double(env, j->r.value, &m);
      else napi_get_null(env, &m);
      napi_set_named_property(env, v, "mse", m);
      set_num(env, v, "overflow", (double)j->r.overflow);
    } else if (j->kind == Job::kPsnr) {
      napi_create_double(env, j->r.value, &v);
    } else {
      napi_get_undefined(env, &v);
    }
    napi_resolve_deferred(env, j->deferred, v);
  }
  g_api.pending_free(j->pending);
  for (napi_ref& r : j->keep)
    if (r) napi_delete_reference(env, r);
  napi_delete_async_work(env, j->work);
  delete j;
}

// Starts the wait of an enqueued readback; returns the Promise.
napi_value start_job(napi_env env, Job* j) {
  napi_value promise, name;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, "splat_gs_readback", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, nullptr, name, job_execute, job_complete, j, &j->work) != napi_ok ||
      napi_queue_async_work(env, j->work) != napi_ok) {
    g_api.pending_free(j->pending);
    for (napi_ref& r : j->keep)
      if (r) napi_delete_reference(env, r);
    delete j;
    return throw_err(env, "readback: cannot queue the async work");
  }
  return promise;
}

// readLossAsync(h) -> Promise<{mse, overflow}>
napi_value ReadLossAsync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return throw_err(env, "readLoss: trainer destroyed");
  Job* j = new Job();
  j->kind = Job::kLoss;
  if (!check(env, h, g_api.read_loss_async(h, &j->pending))) { delete j; return nullptr; }
  return start_job(env, j);
}

// evalCamPsnrAsync(h, ci) -> Promise<number>
napi_value EvalCamPsnrAsync(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return throw_err(env, "evalCamPsnr: trainer destroyed");
  int32_t ci = 0;
  napi_get_value_int32(env, argv[1], &ci);
  Job* j = new Job();
  j->kind = Job::kPsnr;
  if (!check(env, h, g_api.eval_cam_psnr_async(h, ci, &j->pending))) { delete j; return nullptr; }
  return start_job(env, j);
}

// readGaussiansAsync(h, Float32Array params, Float32Array|null sh) -> Promise<undefined>
napi_value ReadGaussiansAsync(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  if (!h) return throw_err(env, "readGaussians: trainer destroyed");
  gs_info in;
  if (!check(env, h, g_api.get_info(h, &in))) return nullptr;
  float *p = nullptr, *sh = nullptr;
  size_t lp = 0, ls = 0;
  if (!get_f32(env, argv[1], &p, &lp) || (int64_t)lp < in.n * 16)
    return throw_err(env, "readGaussians: params must be a Float32Array of n*16");
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  const bool hasSh = t != napi_null && t != napi_undefined;
  if (hasSh && (!get_f32(env, argv[2], &sh, &ls) || (int64_t)ls < in.n * 3 * in.shK))
    return throw_err(env, "readGaussians: sh must be a Float32Array of n*3*shK");
  Job* j = new Job();
  j->kind = Job::kGaussians;
  if (!check(env, h, g_api.read_gaussians_async(h, p, sh, &j->pending))) { delete j; return nullptr; }
  // the wait writes into both arrays: keep them from collection until then
  napi_create_reference(env, argv[1], 1, &j->keep[0]);
  if (hasSh) napi_create_reference(env, argv[2], 1, &j->keep[1]);
  return start_job(env, j);
}

napi_value Sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  check(env, h, g_api.sync(h));
  return nullptr;
}

// refine(h, {seed, growRate, growUntil, growLimit}) -> {moved, grown, n, maxTile, overflow}
napi_value Refine(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  gs_refine_opts ro{};
  ro.size = (int32_t)sizeof(ro);
  ro.seed = (uint32_t)get_num(env, argv[1], "seed", 0);
  ro.growRate = (float)get_num(env, argv[1], "growRate", -1);
  ro.growUntil = get_num(env, argv[1], "growUntil", -1);
  ro.growLimit = (int64_t)get_num(env, argv[1], "growLimit", 0);
  gs_refine_result r;
  if (!check(env, h, g_api.refine(h, &ro, &r))) return nullptr;
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  set_num(env, o, "moved", (double)r.moved);
  set_num(env, o, "grown", (double)r.grown);
  set_num(env, o, "n", (double)r.n);
  set_num(env, o, "maxTile", (double)r.maxTile);
  set_num(env, o, "overflow", (double)r.overflow);
  return o;
}

// exportPly(h, bake) -> Buffer holding the binary PLY file
napi_value ExportPly(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  gs_trainer_t* h = get_handle(env, argv[0]);
  bool bake = true;
  if (argc > 1) napi_get_value_bool(env, argv[1], &bake);
  int64_t size = 0;
  if (!check(env, h, g_api.export_ply(h, bake ? 1 : 0, nullptr, 0, &size))) return nullptr;
  void* data = nullptr;
  napi_value out;
  NAPI_CALL(env, napi_create_buffer(env, (size_t)size, &data, &out));
  if (!check(env, h, g_api.export_ply(h, bake ? 1 : 0, data, size, &size))) return nullptr;
  return out;
}

// loaded() -> true when libsplatgs.so opened and every symbol resolved
napi_value Loaded(napi_env env, napi_callback_info) {
  napi_value r;
  napi_get_boolean(env, load_api(), &r);
  return r;
}

napi_value LoadError(napi_env env, napi_callback_info) {
  napi_value r;
  napi_create_string_utf8(env, g_api.err.c_str(), NAPI_AUTO_LENGTH, &r);
  return r;
}

napi_value Init(napi_env env, napi_value exports) {
  const struct {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"create", Create},           {"setup", Setup},         {"step", Step},
      {"stepViews", StepViews},     {"setRasterShape", SetRasterShape}, {"autotune", Autotune},
      {"readLoss", ReadLoss},       {"render", Render},       {"readGaussians", ReadGaussians},
      {"evalCamPsnr", EvalCamPsnr}, {"setHoldout", SetHoldout}, {"setExcluded", SetExcluded},
      {"setIter", SetIter},         {"info", Info},           {"sync", Sync},
      {"refine", Refine},           {"exportPly", ExportPly},
      {"evalCamPsnrRefined", EvalCamPsnrRefined}, {"getCamera", GetCamera},
      {"readCamGrads", ReadCamGrads},
      {"destroy", Destroy},         {"stepStats", StepStats},
      {"readLossAsync", ReadLossAsync}, {"evalCamPsnrAsync", EvalCamPsnrAsync},
      {"readGaussiansAsync", ReadGaussiansAsync},
      {"loaded", Loaded},           {"loadError", LoadError},
  };
  for (const auto& f : fns) {
    napi_value fn;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn);
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)

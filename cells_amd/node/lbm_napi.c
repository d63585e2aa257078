/*
 * lbm_napi.c - thin N-API addon exposing the C ABI of include/lbm.h to Node.js.
 *
 * Every export maps one-to-one onto an lbm_* entry point; the JS classes in index.js rebuild the
 * reference's Solver / Solver3D surface (src/solver.ts, src/solver3d.ts) on top. Calls that only
 * enqueue GPU work run on the JS thread; the waits (run's completion, readMacro, readForces'
 * copy) run on the libuv pool through napi_async_work and settle a Promise, so the event loop
 * never blocks on the GPU. A failing status throws (or rejects with) an Error carrying
 * lbm_last_error().
 */
#define NAPI_VERSION 6
#include <node_api.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm.h"

#define CHECK_NAPI(env, call)                                             \
  do {                                                                    \
    if ((call) != napi_ok) {                                              \
      napi_throw_error((env), NULL, "N-API call failed: " #call);         \
      return NULL;                                                        \
    }                                                                     \
  } while (0)

static napi_value throw_status(napi_env env, int rc) {
  (void)rc;
  napi_throw_error(env, NULL, lbm_last_error());
  return NULL;
}

/* A handle is a JS object { ptr: External(Box) }; the box lets destroy() free the solver early and
 * the external's finalizer free whatever is left when the object is collected. */
typedef struct {
  lbm_solver* s;
} Box;

static void finalize_box(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  Box* b = (Box*)data;
  if (b->s) lbm_destroy(b->s);
  free(b);
}

static Box* get_box(napi_env env, napi_value obj) {
  napi_value ext;
  void* p = NULL;
  napi_valuetype t;
  if (napi_typeof(env, obj, &t) != napi_ok || t != napi_object) return NULL;
  if (napi_get_named_property(env, obj, "ptr", &ext) != napi_ok) return NULL;
  if (napi_typeof(env, ext, &t) != napi_ok || t != napi_external) return NULL;
  if (napi_get_value_external(env, ext, &p) != napi_ok) return NULL;
  return (Box*)p;
}

static lbm_solver* get_handle(napi_env env, napi_value obj) {
  Box* b = get_box(env, obj);
  return b ? b->s : NULL;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t k = argc; k < want; k++) napi_get_undefined(env, &argv[k]);
  return 1;
}

static double num_prop(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  napi_value v;
  napi_valuetype t;
  double d = dflt;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return dflt;
  if (napi_typeof(env, v, &t) != napi_ok) return dflt;
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1.0 : 0.0;
  }
  if (t != napi_number) return dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

static lbm_solver* solver_arg(napi_env env, napi_value v) {
  lbm_solver* s = get_handle(env, v);
  if (!s) napi_throw_type_error(env, NULL, "solver handle expected (destroyed?)");
  return s;
}

static napi_value make_handle(napi_env env, lbm_solver* s) {
  napi_value obj, ext;
  Box* b = (Box*)malloc(sizeof(Box));
  b->s = s;
  napi_create_object(env, &obj);
  napi_create_external(env, b, finalize_box, NULL, &ext);
  napi_set_named_property(env, obj, "ptr", ext);
  return obj;
}

/* Fills an lbm_config2d from the flat object the JS wrapper builds (all keys present). */
static void read_cfg2(napi_env env, napi_value o, lbm_config2d* c) {
  lbm_config2d_defaults(c);
  c->width = (int32_t)num_prop(env, o, "width", 0);
  c->height = (int32_t)num_prop(env, o, "height", 0);
  c->left = (int32_t)num_prop(env, o, "left", 0);
  c->right = (int32_t)num_prop(env, o, "right", 0);
  c->bottom = (int32_t)num_prop(env, o, "bottom", 0);
  c->top = (int32_t)num_prop(env, o, "top", 0);
  c->parabolic = (int32_t)num_prop(env, o, "parabolic", 0);
  c->workgroup_size = (int32_t)num_prop(env, o, "workgroupSize", 128);
  c->force_every = (int32_t)num_prop(env, o, "forceEvery", 1);
  c->tau = num_prop(env, o, "tau", 0.6);
  c->lambda = num_prop(env, o, "lambda", 3.0 / 16.0);
  c->smagorinsky = num_prop(env, o, "smagorinsky", 0);
  c->u_in = num_prop(env, o, "uIn", 0);
  c->u_lid = num_prop(env, o, "uLid", 0);
  c->sponge_fraction = num_prop(env, o, "spongeFraction", 0);
  c->sponge_tau = num_prop(env, o, "spongeTau", 1);
  c->body_force[0] = num_prop(env, o, "bodyForceX", 0);
  c->body_force[1] = num_prop(env, o, "bodyForceY", 0);
  c->absorb = num_prop(env, o, "absorb", 0);
  c->inlet_layer_fraction = num_prop(env, o, "inletLayerFraction", 0);
}

static void read_cfg3(napi_env env, napi_value o, lbm_config3d* c) {
  lbm_config3d_defaults(c);
  c->width = (int32_t)num_prop(env, o, "width", 0);
  c->height = (int32_t)num_prop(env, o, "height", 0);
  c->depth = (int32_t)num_prop(env, o, "depth", 0);
  c->precision = (int32_t)num_prop(env, o, "precisionCode", 0);
  c->workgroup_size = (int32_t)num_prop(env, o, "workgroupSize", 128);
  c->force_every = (int32_t)num_prop(env, o, "forceEvery", 1);
  c->tau = num_prop(env, o, "tau", 0.6);
  c->lambda = num_prop(env, o, "lambda", c->lambda); /* absent or NaN: trtLambda3(tau) */
  c->smagorinsky = num_prop(env, o, "smagorinsky", 0);
  c->u_in = num_prop(env, o, "uIn", 0);
  c->sponge_fraction = num_prop(env, o, "spongeFraction", 0);
  c->sponge_tau = num_prop(env, o, "spongeTau", 1);
  c->absorb = num_prop(env, o, "absorb", 0);
  c->velocity_set = (int32_t)num_prop(env, o, "velocitySet", 19);
}

static napi_value js_create2d(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_config2d c;
  read_cfg2(env, argv[0], &c);
  int32_t dev = 0;
  napi_get_value_int32(env, argv[1], &dev);
  lbm_solver* s = NULL;
  int rc = lbm_create_2d(&c, dev, &s);
  if (rc != LBM_OK) return throw_status(env, rc);
  return make_handle(env, s);
}

static napi_value js_create3d(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_config3d c;
  read_cfg3(env, argv[0], &c);
  int32_t dev = 0;
  napi_get_value_int32(env, argv[1], &dev);
  lbm_solver* s = NULL;
  int rc = lbm_create_3d(&c, dev, &s);
  if (rc != LBM_OK) return throw_status(env, rc);
  return make_handle(env, s);
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Box* b = get_box(env, argv[0]);
  if (b && b->s) {
    lbm_destroy(b->s);
    b->s = NULL;
  }
  return NULL;
}

static napi_value js_update2d(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  lbm_config2d c;
  read_cfg2(env, argv[1], &c);
  int rc = lbm_update_2d(s, &c);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_update3d(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  lbm_config3d c;
  read_cfg3(env, argv[1], &c);
  int rc = lbm_update_3d(s, &c);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_set_inlet_ramp(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t kind = 0;
  double u = 0, steps = 0;
  napi_get_value_int32(env, argv[1], &kind);
  napi_get_value_double(env, argv[2], &u);
  napi_get_value_double(env, argv[3], &steps);
  int rc = lbm_set_inlet_ramp(s, kind, u, steps);
  return rc ? throw_status(env, rc) : NULL;
}

/* Float32Array argument -> pointer and element count. */
static int f32_arg(napi_env env, napi_value v, float** data, size_t* n) {
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return 0;
  napi_typedarray_type type;
  void* p;
  size_t len, off;
  napi_value ab;
  napi_get_typedarray_info(env, v, &type, &len, &p, &ab, &off);
  if (type != napi_float32_array) return 0;
  *data = (float*)p;
  *n = len;
  return 1;
}

static napi_value js_set_sdf(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  float* d;
  size_t n;
  if (!f32_arg(env, argv[1], &d, &n)) {
    napi_throw_type_error(env, NULL, "setSdf expects a Float32Array");
    return NULL;
  }
  int rc = lbm_set_sdf(s, d, n);
  return rc ? throw_status(env, rc) : NULL;
}

/* deviceName(ordinal) -> string */
static napi_value js_device_name(napi_env env, napi_callback_info info) {
  napi_value argv[1], out;
  if (!get_args(env, info, 1, argv)) return NULL;
  int32_t dev = 0;
  napi_get_value_int32(env, argv[0], &dev);
  char name[256];
  int rc = lbm_device_name(dev, name, sizeof(name));
  if (rc) return throw_status(env, rc);
  CHECK_NAPI(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &out));
  return out;
}

/* copyRoofline(ordinal, nCells, q, elemBytes, workgroup, iters) -> total milliseconds */
static napi_value js_copy_roofline(napi_env env, napi_callback_info info) {
  napi_value argv[6], out;
  if (!get_args(env, info, 6, argv)) return NULL;
  double v[6];
  for (int k = 0; k < 6; k++)
    if (napi_get_value_double(env, argv[k], &v[k]) != napi_ok) {
      napi_throw_type_error(env, NULL, "copyRoofline expects six numbers");
      return NULL;
    }
  double ms = 0;
  int rc = lbm_copy_roofline((int)v[0], (int64_t)v[1], (int32_t)v[2], (int32_t)v[3], (int32_t)v[4], (int32_t)v[5], &ms);
  if (rc) return throw_status(env, rc);
  CHECK_NAPI(env, napi_create_double(env, ms, &out));
  return out;
}

/* autotune(handle, trialSteps) -> { block, sweep, tiled, stride }: lbm_autotune, then lbm_launch_info
 * and lbm_population_layout */
static napi_value js_autotune(napi_env env, napi_callback_info info) {
  napi_value argv[2], out, v;
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t trial = 200, block = 0, sweep = 0;
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  if (t == napi_number) napi_get_value_int32(env, argv[1], &trial);
  int32_t tiled = 0;
  int64_t stride = 0;
  int rc = lbm_autotune(s, trial, NULL);
  if (!rc) rc = lbm_launch_info(s, &block, &sweep);
  if (!rc) rc = lbm_population_layout(s, &tiled, &stride);
  if (rc) return throw_status(env, rc);
  CHECK_NAPI(env, napi_create_object(env, &out));
  napi_create_int32(env, block, &v);
  napi_set_named_property(env, out, "block", v);
  napi_create_int32(env, sweep, &v);
  napi_set_named_property(env, out, "sweep", v);
  napi_get_boolean(env, tiled != 0, &v);
  napi_set_named_property(env, out, "tiled", v);
  napi_create_double(env, (double)stride, &v);
  napi_set_named_property(env, out, "stride", v);
  return out;
}

/* setPopulationLayout(handle, tiled, stride): lbm_set_population_layout (device layout only; every
 * array the caller reads keeps the reference's layout and bits) */
static napi_value js_set_population_layout(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  bool tiled = false;
  double stride = 0;
  napi_get_value_bool(env, argv[1], &tiled);
  napi_get_value_double(env, argv[2], &stride);
  int rc = lbm_set_population_layout(s, tiled ? 1 : 0, (int64_t)stride);
  if (rc) return throw_status(env, rc);
  return NULL;
}

/* divergedAt(handle) -> number: lbm_diverged (waits for the work enqueued before it) */
static napi_value js_diverged_at(napi_env env, napi_callback_info info) {
  napi_value argv[1], out;
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int64_t first = 0;
  int rc = lbm_diverged(s, &first);
  if (rc) return throw_status(env, rc);
  CHECK_NAPI(env, napi_create_int64(env, first, &out));
  return out;
}

static napi_value js_rebuild_flags(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int rc = lbm_rebuild_flags(s);
  return rc ? throw_status(env, rc) : NULL;
}

/* setOpenFaces(handle, kind) */
static napi_value js_set_open_faces(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t kind = -1;
  napi_get_value_int32(env, argv[1], &kind);
  int rc = lbm_set_open_faces(s, kind);
  return rc ? throw_status(env, rc) : NULL;
}

/* brush(handle, cx, cy, r, erase, rebuild) */
static napi_value js_brush(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  double v[3];
  bool erase = false, rebuild = true;
  for (int k = 0; k < 3; k++)
    if (napi_get_value_double(env, argv[1 + k], &v[k]) != napi_ok) {
      napi_throw_type_error(env, NULL, "brush expects numbers cx, cy, r");
      return NULL;
    }
  napi_get_value_bool(env, argv[4], &erase);
  napi_get_value_bool(env, argv[5], &rebuild);
  int rc = lbm_brush(s, v[0], v[1], v[2], erase, rebuild);
  return rc ? throw_status(env, rc) : NULL;
}

/* brushStroke(handle, xy: Float64Array of (x, y) pairs, r, erase) */
static napi_value js_brush_stroke(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  bool is_ta = false;
  napi_is_typedarray(env, argv[1], &is_ta);
  napi_typedarray_type type = napi_int8_array;
  void* p = NULL;
  size_t len = 0, off;
  napi_value ab;
  if (is_ta) napi_get_typedarray_info(env, argv[1], &type, &len, &p, &ab, &off);
  if (!is_ta || type != napi_float64_array || len % 2) {
    napi_throw_type_error(env, NULL, "brushStroke expects a Float64Array of (x, y) pairs");
    return NULL;
  }
  double r = 0;
  bool erase = false;
  if (napi_get_value_double(env, argv[2], &r) != napi_ok) {
    napi_throw_type_error(env, NULL, "brushStroke expects a number r");
    return NULL;
  }
  napi_get_value_bool(env, argv[3], &erase);
  int rc = lbm_brush_stroke(s, (const double*)p, len / 2, r, erase);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_init_field(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  float* d = NULL;
  size_t n = 0;
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  if (t != napi_undefined && t != napi_null && !f32_arg(env, argv[1], &d, &n)) {
    napi_throw_type_error(env, NULL, "initField expects a Float32Array or nothing");
    return NULL;
  }
  int rc = lbm_init_field(s, d, n);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_reset_forces(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int rc = lbm_reset_forces(s);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_enqueue_steps(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int64_t n = 0;
  napi_get_value_int64(env, argv[1], &n);
  int rc = lbm_enqueue_steps(s, n);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_enqueue_macro(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int rc = lbm_enqueue_macro(s);
  return rc ? throw_status(env, rc) : NULL;
}

static napi_value js_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t dims[3];
  int64_t n, step;
  int32_t q, prec;
  int rc = lbm_info(s, dims, &n, &step, &q, &prec);
  if (rc) return throw_status(env, rc);
  napi_value o, v;
  napi_create_object(env, &o);
  const char* names[3] = {"W", "H", "D"};
  for (int k = 0; k < 3; k++) {
    napi_create_int32(env, dims[k], &v);
    napi_set_named_property(env, o, names[k], v);
  }
  napi_create_double(env, (double)n, &v);
  napi_set_named_property(env, o, "N", v);
  napi_create_double(env, (double)step, &v);
  napi_set_named_property(env, o, "step", v);
  napi_create_int32(env, q, &v);
  napi_set_named_property(env, o, "Q", v);
  napi_create_int32(env, prec, &v);
  napi_set_named_property(env, o, "precision", v);
  return o;
}

static napi_value js_time_steps(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int64_t n = 0;
  napi_get_value_int64(env, argv[1], &n);
  double ms = 0;
  int rc = lbm_time_steps(s, n, &ms);
  if (rc) return throw_status(env, rc);
  napi_value v;
  napi_create_double(env, ms, &v);
  return v;
}

/* ---- asynchronous waits ---- */
enum { JOB_SYNC = 0, JOB_MACRO = 1, JOB_FORCES = 2, JOB_ADJ_FWD = 3, JOB_ADJ_BWD = 4 };

typedef struct {
  int kind;
  lbm_solver* s;
  napi_deferred deferred;
  napi_async_work work;
  napi_ref keep;   /* keeps the handle object (and output array) alive */
  napi_ref out_ref;
  float* out;
  size_t n;
  int64_t ticket;
  lbm_adjoint_config acfg; /* JOB_ADJ_FWD */
  double value;            /* J (JOB_ADJ_FWD), seed (JOB_ADJ_BWD) */
  double* out64;           /* dJ/dsdf (JOB_ADJ_BWD) */
  int rc;
  char err[512];
} Job;

static void job_execute(napi_env env, void* data) {
  (void)env;
  Job* j = (Job*)data;
  if (j->kind == JOB_SYNC) j->rc = lbm_sync(j->s);
  else if (j->kind == JOB_MACRO) j->rc = lbm_read_macro(j->s, j->out, j->n);
  else if (j->kind == JOB_ADJ_FWD) j->rc = lbm_adjoint_forward(j->s, &j->acfg, &j->value);
  else if (j->kind == JOB_ADJ_BWD) j->rc = lbm_adjoint_backward(j->s, j->value, j->out64, j->n);
  else j->rc = lbm_read_forces_wait(j->s, j->ticket);
  if (j->rc) {
    strncpy(j->err, lbm_last_error(), sizeof(j->err) - 1);
    j->err[sizeof(j->err) - 1] = 0;
  }
}

static void job_complete(napi_env env, napi_status status, void* data) {
  Job* j = (Job*)data;
  napi_value result = NULL, err, msg;
  int rc = j->rc;
  if (status != napi_ok && rc == 0) {
    rc = -3;
    strcpy(j->err, "async work cancelled");
  }
  if (rc == 0) {
    if (j->kind == JOB_SYNC) {
      napi_get_undefined(env, &result);
    } else if (j->kind == JOB_MACRO || j->kind == JOB_ADJ_BWD) {
      napi_get_reference_value(env, j->out_ref, &result);
    } else if (j->kind == JOB_ADJ_FWD) {
      napi_create_double(env, j->value, &result);
    } else {
      /* Bookkeeping back on the JS thread (src/solver.ts:322-334). */
      int32_t dims[3];
      int64_t n, step;
      int32_t q, prec;
      lbm_info(j->s, dims, &n, &step, &q, &prec);
      const int dim = q == 9 ? 2 : 3;
      int64_t* steps = (int64_t*)malloc(sizeof(int64_t) * LBM_HISTORY_LEN);
      double* vals = (double*)malloc(sizeof(double) * LBM_HISTORY_LEN * 3);
      int64_t count = 0;
      rc = lbm_read_forces_end(j->s, j->ticket, steps, vals, LBM_HISTORY_LEN, &count);
      if (rc == 0) {
        napi_create_array_with_length(env, (size_t)count, &result);
        const char* names[3] = {"fx", "fy", "fz"};
        for (int64_t k = 0; k < count; k++) {
          napi_value o, v;
          napi_create_object(env, &o);
          napi_create_double(env, (double)steps[k], &v);
          napi_set_named_property(env, o, "step", v);
          for (int d = 0; d < dim; d++) {
            napi_create_double(env, vals[k * dim + d], &v);
            napi_set_named_property(env, o, names[d], v);
          }
          napi_set_element(env, result, (uint32_t)k, o);
        }
      } else {
        strncpy(j->err, lbm_last_error(), sizeof(j->err) - 1);
        j->err[sizeof(j->err) - 1] = 0;
      }
      free(steps);
      free(vals);
    }
  }
  if (rc == 0) {
    napi_resolve_deferred(env, j->deferred, result);
  } else {
    napi_create_string_utf8(env, j->err, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  if (j->out_ref) napi_delete_reference(env, j->out_ref);
  napi_delete_reference(env, j->keep);
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value start_job(napi_env env, napi_value handle, Job* j, const char* name) {
  napi_value promise, res_name;
  napi_create_promise(env, &j->deferred, &promise);
  napi_create_reference(env, handle, 1, &j->keep);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &res_name);
  napi_create_async_work(env, NULL, res_name, job_execute, job_complete, j, &j->work);
  napi_queue_async_work(env, j->work);
  return promise;
}

/* Promise that settles when the GPU has finished everything enqueued so far. */
static napi_value js_sync_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  Job* j = (Job*)calloc(1, sizeof(Job));
  j->kind = JOB_SYNC;
  j->s = s;
  return start_job(env, argv[0], j, "lbm_sync");
}

static napi_value js_read_macro_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t dims[3];
  int64_t n, step;
  int32_t q, prec;
  lbm_info(s, dims, &n, &step, &q, &prec);
  napi_value ab, ta;
  void* data;
  CHECK_NAPI(env, napi_create_arraybuffer(env, (size_t)n * 16, &data, &ab));
  CHECK_NAPI(env, napi_create_typedarray(env, napi_float32_array, (size_t)n * 4, ab, 0, &ta));
  Job* j = (Job*)calloc(1, sizeof(Job));
  j->kind = JOB_MACRO;
  j->s = s;
  j->out = (float*)data;
  j->n = (size_t)n * 4;
  napi_create_reference(env, ta, 1, &j->out_ref);
  return start_job(env, argv[0], j, "lbm_read_macro");
}

static napi_value js_read_forces_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int64_t ticket = 0;
  int rc = lbm_read_forces_begin(s, &ticket);
  if (rc) return throw_status(env, rc);
  Job* j = (Job*)calloc(1, sizeof(Job));
  j->kind = JOB_FORCES;
  j->s = s;
  j->ticket = ticket;
  return start_job(env, argv[0], j, "lbm_read_forces");
}

static napi_value js_read_populations(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t dims[3];
  int64_t n, step;
  int32_t q, prec;
  lbm_info(s, dims, &n, &step, &q, &prec);
  const size_t elem = prec == LBM_FP32 ? 4 : 2;
  const size_t count = (size_t)q * (size_t)n;
  napi_value ab, ta;
  void* data;
  CHECK_NAPI(env, napi_create_arraybuffer(env, count * elem, &data, &ab));
  CHECK_NAPI(env, napi_create_typedarray(env, elem == 4 ? napi_float32_array : napi_uint16_array, count, ab, 0, &ta));
  int rc = lbm_read_populations(s, data, count * elem);
  return rc ? throw_status(env, rc) : ta;
}

static napi_value js_read_flags(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t dims[3];
  int64_t n, step;
  int32_t q, prec;
  lbm_info(s, dims, &n, &step, &q, &prec);
  napi_value ab, ta;
  void* data;
  CHECK_NAPI(env, napi_create_arraybuffer(env, (size_t)n * 4, &data, &ab));
  CHECK_NAPI(env, napi_create_typedarray(env, napi_uint32_array, (size_t)n, ab, 0, &ta));
  int rc = lbm_read_flags(s, (uint32_t*)data, (size_t)n);
  return rc ? throw_status(env, rc) : ta;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value v;
  napi_create_int32(env, lbm_abi_version(), &v);
  return v;
}

/* adjointForwardAsync(handle, opts): lbm_adjoint_forward on the libuv pool, resolves to J. opts holds
 * steps, windowStart, scale, checkpointEvery, direction ([dx, dy, dz]), rampKind, rampU,
 * rampSteps and recordBudgetBytes (the fields of lbm_adjoint_config). */
static napi_value js_adjoint_forward_async(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  Job* j = (Job*)calloc(1, sizeof(Job));
  j->kind = JOB_ADJ_FWD;
  j->s = s;
  napi_value o = argv[1];
  j->acfg.steps = (int64_t)num_prop(env, o, "steps", 0);
  j->acfg.window_start = (int64_t)num_prop(env, o, "windowStart", 0);
  j->acfg.scale = num_prop(env, o, "scale", 1.0);
  j->acfg.checkpoint_every = (int32_t)num_prop(env, o, "checkpointEvery", 200);
  j->acfg.ramp_kind = (int32_t)num_prop(env, o, "rampKind", 0);
  j->acfg.ramp_u = num_prop(env, o, "rampU", 0);
  j->acfg.ramp_steps = num_prop(env, o, "rampSteps", 0);
  j->acfg.record_budget_bytes = (int64_t)num_prop(env, o, "recordBudgetBytes", 0);
  j->acfg.direction[0] = 1.0;
  napi_value dir;
  bool is_array = false;
  if (napi_get_named_property(env, argv[1], "direction", &dir) == napi_ok &&
      napi_is_array(env, dir, &is_array) == napi_ok && is_array) {
    for (uint32_t k = 0; k < 3; k++) {
      napi_value e;
      double d = 0;
      if (napi_get_element(env, dir, k, &e) == napi_ok && napi_get_value_double(env, e, &d) == napi_ok)
        j->acfg.direction[k] = d;
      else
        j->acfg.direction[k] = 0;
    }
  }
  return start_job(env, argv[0], j, "lbm_adjoint_forward");
}

/* adjointBackwardAsync(handle, seed): lbm_adjoint_backward on the libuv pool, resolves to a
 * Float64Array of N values (seed * dJ/dsdf). */
static napi_value js_adjoint_backward_async(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  lbm_solver* s = solver_arg(env, argv[0]);
  if (!s) return NULL;
  int32_t dims[3];
  int64_t n, step;
  int32_t q, prec;
  if (lbm_info(s, dims, &n, &step, &q, &prec)) return throw_status(env, -1);
  double seed = 1.0;
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  if (t == napi_number) napi_get_value_double(env, argv[1], &seed);
  Job* j = (Job*)calloc(1, sizeof(Job));
  j->kind = JOB_ADJ_BWD;
  j->s = s;
  j->value = seed;
  j->n = (size_t)n;
  napi_value ab, arr;
  void* data = NULL;
  CHECK_NAPI(env, napi_create_arraybuffer(env, sizeof(double) * (size_t)n, &data, &ab));
  CHECK_NAPI(env, napi_create_typedarray(env, napi_float64_array, (size_t)n, ab, 0, &arr));
  j->out64 = (double*)data;
  napi_create_reference(env, arr, 1, &j->out_ref);
  return start_job(env, argv[0], j, "lbm_adjoint_backward");
}

#define EXPORT(name, fn) {name, NULL, fn, NULL, NULL, NULL, napi_enumerable, NULL}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      EXPORT("abiVersion", js_abi_version),
      EXPORT("create2d", js_create2d),
      EXPORT("create3d", js_create3d),
      EXPORT("destroy", js_destroy),
      EXPORT("update2d", js_update2d),
      EXPORT("update3d", js_update3d),
      EXPORT("setInletRamp", js_set_inlet_ramp),
      EXPORT("setSdf", js_set_sdf),
      EXPORT("rebuildFlags", js_rebuild_flags),
      EXPORT("setOpenFaces", js_set_open_faces),
      EXPORT("deviceName", js_device_name),
      EXPORT("copyRoofline", js_copy_roofline),
      EXPORT("autotune", js_autotune),
      EXPORT("divergedAt", js_diverged_at),
      EXPORT("setPopulationLayout", js_set_population_layout),
      EXPORT("brush", js_brush),
      EXPORT("brushStroke", js_brush_stroke),
      EXPORT("initField", js_init_field),
      EXPORT("resetForces", js_reset_forces),
      EXPORT("enqueueSteps", js_enqueue_steps),
      EXPORT("enqueueMacro", js_enqueue_macro),
      EXPORT("info", js_info),
      EXPORT("timeSteps", js_time_steps),
      EXPORT("syncAsync", js_sync_async),
      EXPORT("readMacroAsync", js_read_macro_async),
      EXPORT("readForcesAsync", js_read_forces_async),
      EXPORT("readPopulations", js_read_populations),
      EXPORT("readFlags", js_read_flags),
      EXPORT("adjointForwardAsync", js_adjoint_forward_async),
      EXPORT("adjointBackwardAsync", js_adjoint_backward_async),
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

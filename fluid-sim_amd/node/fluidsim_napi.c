/*
 * fluidsim_napi.c -- Node-API addon over the C ABI of include/fluidsim.h.
 *
 * A thin marshalling layer: every function converts JS values to plain C
 * arguments, calls one fs_* entry point and throws a JS Error carrying
 * fs_last_error() on failure. The handle is a JS external whose finalizer
 * destroys the solver if the caller never called destroy().
 *
 * All calls are synchronous on the calling thread, so calls on one handle are
 * serialised by construction (the C handle is not thread-safe); solver.mjs wraps
 * the readbacks in Promises to keep the reference's async signatures.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fluidsim.h"

#define NAPI_CALL(env, call)                                  \
  do {                                                        \
    if ((call) != napi_ok) {                                  \
      napi_throw_error((env), NULL, "N-API call failed: " #call); \
      return NULL;                                            \
    }                                                         \
  } while (0)

typedef struct {
  void* solver;
} Handle;

static napi_value throw_fs(napi_env env, int rc) {
  char msg[640];
  snprintf(msg, sizeof msg, "fluidsim: %s (status %d)", fs_last_error(), rc);
  napi_throw_error(env, NULL, msg);
  return NULL;
}

static void finalize_handle(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  Handle* h = (Handle*)data;
  if (h->solver) fs_solver_destroy(h->solver);
  free(h);
}

static Handle* get_handle(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a solver handle");
    return NULL;
  }
  Handle* h = (Handle*)p;
  if (!h->solver) {
    napi_throw_error(env, NULL, "solver has been destroyed");
    return NULL;
  }
  return h;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return 0;
  }
  return 1;
}

static int prop_double(napi_env env, napi_value obj, const char* key, double* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok) return 0;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_number) return 0;
  return napi_get_value_double(env, v, out) == napi_ok;
}

static napi_value make_number(napi_env env, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  return v;
}

/* create({nx, ny, lx, ly, nu, advection, pressureSolver, pressureIterations, tolerance,
 *         mgMinSize, mgPre, mgPost, mgCoarse, guess, device}) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  fs_config c;
  fs_config_default(&c);
  double d;
  if (prop_double(env, argv[0], "nx", &d)) c.nx = (int)d;
  if (prop_double(env, argv[0], "ny", &d)) c.ny = (int)d;
  if (prop_double(env, argv[0], "lx", &d)) c.lx = d;
  if (prop_double(env, argv[0], "ly", &d)) c.ly = d;
  if (prop_double(env, argv[0], "nu", &d)) c.nu = d;
  if (prop_double(env, argv[0], "advection", &d)) c.advection = (int)d;
  if (prop_double(env, argv[0], "pressureSolver", &d)) c.pressure_solver = (int)d;
  if (prop_double(env, argv[0], "pressureIterations", &d)) c.pressure_iterations = (int)d;
  if (prop_double(env, argv[0], "tolerance", &d)) c.tolerance = d;
  if (prop_double(env, argv[0], "mgMinSize", &d)) c.mg_min_size = (int)d;
  if (prop_double(env, argv[0], "mgPre", &d)) c.mg_pre = (int)d;
  if (prop_double(env, argv[0], "mgPost", &d)) c.mg_post = (int)d;
  if (prop_double(env, argv[0], "mgCoarse", &d)) c.mg_coarse = (int)d;
  if (prop_double(env, argv[0], "guess", &d)) c.guess = (int)d;
  if (prop_double(env, argv[0], "solidMaskWidth", &d)) c.solid_mask_width = d;
  if (prop_double(env, argv[0], "device", &d)) c.device = (int)d;
  void* s = NULL;
  int rc = fs_solver_create(&c, &s);
  if (rc) return throw_fs(env, rc);
  Handle* h = (Handle*)malloc(sizeof(Handle));
  if (!h) {
    fs_solver_destroy(s);
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  h->solver = s;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, finalize_handle, NULL, &ext));
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* p = NULL;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    Handle* h = (Handle*)p;
    if (h->solver) fs_solver_destroy(h->solver);
    h->solver = NULL;
  }
  return NULL;
}

/* uploadVelocity(h, u, v): Float64Array or Float32Array of nx*ny each */
static napi_value js_upload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_typedarray_type tu, tv;
  size_t nu_, nv_;
  void *pu, *pv;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, argv[1], &tu, &nu_, &pu, &ab, &off) != napi_ok ||
      napi_get_typedarray_info(env, argv[2], &tv, &nv_, &pv, &ab, &off) != napi_ok) {
    napi_throw_type_error(env, NULL, "u and v must be Float64Array or Float32Array");
    return NULL;
  }
  if (tu != tv || nu_ != nv_) {
    napi_throw_type_error(env, NULL, "u and v must have the same type and length");
    return NULL;
  }
  int rc;
  if (tu == napi_float64_array) rc = fs_upload_velocity_f64(h->solver, (const double*)pu, (const double*)pv, nu_);
  else if (tu == napi_float32_array) rc = fs_upload_velocity_f32(h->solver, (const float*)pu, (const float*)pv, nu_);
  else {
    napi_throw_type_error(env, NULL, "u and v must be Float64Array or Float32Array");
    return NULL;
  }
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* loadGuess(h, bytes): the learned guess's weights blob (Uint8Array / Buffer) */
static napi_value js_load_guess(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  if (napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off) != napi_ok || t != napi_uint8_array) {
    napi_throw_type_error(env, NULL, "the weights must be a Uint8Array or Buffer");
    return NULL;
  }
  int rc = fs_load_guess(h->solver, data, len);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value new_f32(napi_env env, size_t n, float** data) {
  napi_value ab, arr;
  void* p;
  if (napi_create_arraybuffer(env, n * sizeof(float), &p, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr) != napi_ok) return NULL;
  *data = (float*)p;
  return arr;
}

/* readVelocity(h, n) -> {u: Float32Array, v: Float32Array} */
static napi_value js_read(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double dn;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &dn));
  size_t n = (size_t)dn;
  float *pu, *pv;
  napi_value u = new_f32(env, n, &pu), v = new_f32(env, n, &pv);
  if (!u || !v) {
    napi_throw_error(env, NULL, "cannot allocate the readback arrays");
    return NULL;
  }
  int rc = fs_read_velocity(h->solver, pu, pv, n);
  if (rc) return throw_fs(env, rc);
  napi_value obj;
  NAPI_CALL(env, napi_create_object(env, &obj));
  NAPI_CALL(env, napi_set_named_property(env, obj, "u", u));
  NAPI_CALL(env, napi_set_named_property(env, obj, "v", v));
  return obj;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double dt;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &dt));
  int rc = fs_step(h->solver, dt);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value js_step_cfl(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double cfl;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &cfl));
  int rc = fs_step_cfl(h->solver, cfl);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int rc = fs_sync(h->solver);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value js_max_div(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double out = 0;
  int rc = fs_max_divergence(h->solver, &out);
  if (rc) return throw_fs(env, rc);
  return make_number(env, out);
}

static napi_value js_residual_after(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double it;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &it));
  double out = 0;
  int rc = fs_poisson_residual_after(h->solver, (int)it, &out);
  if (rc) return throw_fs(env, rc);
  return make_number(env, out);
}

/* setParam(h, name, value) for the mutable fields */
static napi_value js_set(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  char key[64];
  size_t len;
  NAPI_CALL(env, napi_get_value_string_utf8(env, argv[1], key, sizeof key, &len));
  double val;
  NAPI_CALL(env, napi_get_value_double(env, argv[2], &val));
  int rc;
  if (!strcmp(key, "pressureIterations")) rc = fs_set_pressure_iterations(h->solver, (int)val);
  else if (!strcmp(key, "tolerance")) rc = fs_set_tolerance(h->solver, val);
  else if (!strcmp(key, "nu")) rc = fs_set_nu(h->solver, val);
  else if (!strcmp(key, "guess")) rc = fs_set_guess(h->solver, (int)val);
  else {
    napi_throw_error(env, NULL, "unknown parameter");
    return NULL;
  }
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* setForcing(h, {kf, width, amp, drag, seed} | null) */
static napi_value js_forcing(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_valuetype t;
  NAPI_CALL(env, napi_typeof(env, argv[1], &t));
  int rc;
  if (t == napi_null || t == napi_undefined) {
    rc = fs_set_forcing(h->solver, NULL);
  } else {
    fs_forcing f = {28, 2, 0.4, 0, 1};
    double d;
    if (prop_double(env, argv[1], "kf", &d)) f.kf = d;
    if (prop_double(env, argv[1], "width", &d)) f.width = d;
    if (prop_double(env, argv[1], "amp", &d)) f.amp = d;
    if (prop_double(env, argv[1], "drag", &d)) f.drag = d;
    if (prop_double(env, argv[1], "seed", &d)) f.seed = (uint32_t)d;
    rc = fs_set_forcing(h->solver, &f);
  }
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* setSolid(h, {x, y, r, vx, vy, omega, passes} | null) */
static napi_value js_solid(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_valuetype t;
  NAPI_CALL(env, napi_typeof(env, argv[1], &t));
  int rc;
  if (t == napi_null || t == napi_undefined) {
    rc = fs_set_solid(h->solver, NULL);
  } else {
    fs_disk d = {0, 0, 0, 0, 0, 0, 1};
    double x;
    if (prop_double(env, argv[1], "x", &x)) d.x = x;
    if (prop_double(env, argv[1], "y", &x)) d.y = x;
    if (prop_double(env, argv[1], "r", &x)) d.r = x;
    if (prop_double(env, argv[1], "vx", &x)) d.vx = x;
    if (prop_double(env, argv[1], "vy", &x)) d.vy = x;
    if (prop_double(env, argv[1], "omega", &x)) d.omega = x;
    if (prop_double(env, argv[1], "passes", &x)) d.passes = (int)x;
    rc = fs_set_solid(h->solver, &d);
  }
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* solidImpulse(h) -> [fx, fy, torque] of the last coupling, dA included */
static napi_value js_solid_impulse(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double imp[3];
  int rc = fs_solid_impulse(h->solver, imp);
  if (rc) return throw_fs(env, rc);
  napi_value arr;
  NAPI_CALL(env, napi_create_array_with_length(env, 3, &arr));
  for (uint32_t k = 0; k < 3; k++) NAPI_CALL(env, napi_set_element(env, arr, k, make_number(env, imp[k])));
  return arr;
}

static napi_value js_stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  fs_stats st;
  int rc = fs_get_stats(h->solver, &st);
  if (rc) return throw_fs(env, rc);
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "t", make_number(env, st.t));
  napi_set_named_property(env, o, "steps", make_number(env, (double)st.steps));
  napi_set_named_property(env, o, "lastVCycles", make_number(env, st.last_vcycles));
  napi_set_named_property(env, o, "totalVCycles", make_number(env, (double)st.total_vcycles));
  napi_set_named_property(env, o, "lastResidual", make_number(env, st.last_residual));
  napi_set_named_property(env, o, "lastDt", make_number(env, st.last_dt));
  napi_set_named_property(env, o, "gatedSteps", make_number(env, (double)st.gated_steps));
  napi_set_named_property(env, o, "gateMisses", make_number(env, (double)st.gate_misses));
  napi_set_named_property(env, o, "graphLaunches", make_number(env, (double)st.graph_launches));
  return o;
}

static napi_value js_diag(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  fs_diag d;
  int rc = fs_diagnostics(h->solver, &d);
  if (rc) return throw_fs(env, rc);
  napi_value o;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "kineticEnergy", make_number(env, d.kinetic_energy));
  napi_set_named_property(env, o, "maxDivergence", make_number(env, d.max_divergence));
  napi_set_named_property(env, o, "l2Divergence", make_number(env, d.l2_divergence));
  napi_set_named_property(env, o, "momentumX", make_number(env, d.momentum_x));
  napi_set_named_property(env, o, "momentumY", make_number(env, d.momentum_y));
  napi_set_named_property(env, o, "enstrophy", make_number(env, d.enstrophy));
  return o;
}

static napi_value js_device(napi_env env, napi_callback_info info) {
  (void)info;
  int n = 0;
  int rc = fs_device_count(&n);
  if (rc) return throw_fs(env, rc);
  char name[256] = "";
  if (n > 0 && (rc = fs_device_name(0, name, sizeof name))) return throw_fs(env, rc);
  napi_value o, s;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_set_named_property(env, o, "count", make_number(env, n));
  NAPI_CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s));
  napi_set_named_property(env, o, "description", s);
  return o;
}

static napi_value init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"create", js_create},   {"destroy", js_destroy},         {"uploadVelocity", js_upload},
      {"readVelocity", js_read}, {"step", js_step},             {"stepCFL", js_step_cfl},
      {"sync", js_sync},       {"maxDivergence", js_max_div},   {"poissonResidualAfter", js_residual_after},
      {"setParam", js_set},    {"setForcing", js_forcing},      {"stats", js_stats},
      {"diagnostics", js_diag}, {"device", js_device},         {"setSolid", js_solid},
      {"solidImpulse", js_solid_impulse}, {"loadGuess", js_load_guess},
  };
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; k++) {
    napi_value f;
    NAPI_CALL(env, napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].cb, NULL, &f));
    NAPI_CALL(env, napi_set_named_property(env, exports, fns[k].name, f));
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

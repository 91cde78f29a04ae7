/*
 * fluidsim_napi.c -- Node-API addon over the C ABI of include/fluidsim.h.
 *
 * A thin marshalling layer: every function converts JS values to plain C
 * arguments, calls fs_* entry points and throws (or rejects with) a JS Error
 * carrying fs_last_error() on failure. The handle is a JS external whose finalizer
 * destroys the solver if the caller never called destroy().
 *
 * Readbacks (velocity, fields as rgba texels, max divergence, the Poisson residual
 * curve, diagnostics, the energy spectrum, waiting for submitted work) return
 * Promises resolved from napi_create_async_work: the copy runs on a worker thread,
 * as the reference's mapAsync readbacks do (src/gpu/solver.js:336-362). The C handle
 * is not thread-safe, so each handle carries a lock that every call on it takes,
 * synchronous ones on the JS thread included; a pending readback holds a reference
 * to the handle, so the handle outlives it, and one that runs after destroy()
 * rejects.
 */
#include <node_api.h>
#include <pthread.h>
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

/* Calls on a handle run one at a time in the order they were made on the JS thread: each
 * takes a ticket when it is made and runs when its ticket is served, so a readback asked
 * for before a step() reads the state before that step, whichever thread runs it. */
typedef struct {
  void* solver;
  pthread_mutex_t mu;
  pthread_cond_t cv;
  unsigned long next_ticket, serving;
} Handle;

static unsigned long take_ticket(Handle* h) {
  pthread_mutex_lock(&h->mu);
  unsigned long t = h->next_ticket++;
  pthread_mutex_unlock(&h->mu);
  return t;
}
static void acquire(Handle* h, unsigned long t) {
  pthread_mutex_lock(&h->mu);
  while (h->serving != t) pthread_cond_wait(&h->cv, &h->mu);
  pthread_mutex_unlock(&h->mu);
}
static void release_turn(Handle* h) {
  pthread_mutex_lock(&h->mu);
  h->serving++;
  pthread_cond_broadcast(&h->cv);
  pthread_mutex_unlock(&h->mu);
}
/* a synchronous call on the JS thread */
static void lock(Handle* h) { acquire(h, take_ticket(h)); }
static void unlock(Handle* h) { release_turn(h); }

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
  pthread_cond_destroy(&h->cv);
  pthread_mutex_destroy(&h->mu);
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
  if (prop_double(env, argv[0], "dyeFade", &d)) c.dye_fade = d;
  if (prop_double(env, argv[0], "confinement", &d)) c.confinement = d;
  if (prop_double(env, argv[0], "mgCoarseExact", &d)) c.mg_coarse_exact = (int)d;
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
  pthread_mutex_init(&h->mu, NULL);
  pthread_cond_init(&h->cv, NULL);
  h->next_ticket = h->serving = 0;
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
    lock(h);
    if (h->solver) fs_solver_destroy(h->solver);
    h->solver = NULL;
    unlock(h);
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
  if (tu != napi_float64_array && tu != napi_float32_array) {
    napi_throw_type_error(env, NULL, "u and v must be Float64Array or Float32Array");
    return NULL;
  }
  lock(h);
  int rc = tu == napi_float64_array
               ? fs_upload_velocity_f64(h->solver, (const double*)pu, (const double*)pv, nu_)
               : fs_upload_velocity_f32(h->solver, (const float*)pu, (const float*)pv, nu_);
  unlock(h);
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
  lock(h);
  int rc = fs_load_guess(h->solver, data, len);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* ------------------------------------------------------------ async readbacks */

enum { J_READ_VELOCITY, J_READ_FIELD, J_MAX_DIV, J_RESIDUAL_AFTER, J_SPECTRUM, J_SYNC };

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref href; /* keeps the handle's external (and so the Handle) alive */
  Handle* h;
  unsigned long ticket;
  int kind, field, iterations, len;
  size_t n, bytes;
  void *a, *b;
  double scalar;
  int rc;
  char err[640];
} Job;

static void job_execute(napi_env env, void* data) {
  (void)env;
  Job* j = (Job*)data;
  acquire(j->h, j->ticket);
  void* s = j->h->solver;
  if (!s) {
    j->rc = FS_E_INVALID_ARG;
    snprintf(j->err, sizeof j->err, "fluidsim: solver has been destroyed");
  } else {
    switch (j->kind) {
      case J_READ_VELOCITY: j->rc = fs_read_velocity(s, (float*)j->a, (float*)j->b, j->n); break;
      case J_READ_FIELD: j->rc = fs_read_field(s, j->field, j->a, j->bytes); break;
      case J_MAX_DIV: j->rc = fs_max_divergence(s, &j->scalar); break;
      case J_RESIDUAL_AFTER: j->rc = fs_poisson_residual_after(s, j->iterations, &j->scalar); break;
      case J_SPECTRUM: j->rc = fs_energy_spectrum(s, (double*)j->a, (int64_t*)j->b, j->len); break;
      default: j->rc = fs_sync(s); break;
    }
    if (j->rc) snprintf(j->err, sizeof j->err, "fluidsim: %s (status %d)", fs_last_error(), j->rc);
  }
  release_turn(j->h);
}

/* the render targets: rgba half floats, returned as their uint16 bits */
static int is_half_field(int field) { return field == FS_F_OUTPUT || field == FS_F_BLOOM; }

static napi_value copy_typed(napi_env env, napi_typedarray_type t, size_t count, size_t esize, const void* src) {
  napi_value ab, arr;
  void* p;
  if (napi_create_arraybuffer(env, count * esize, &p, &ab) != napi_ok) return NULL;
  if (count) memcpy(p, src, count * esize);
  if (napi_create_typedarray(env, t, count, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static void job_complete(napi_env env, napi_status status, void* data) {
  Job* j = (Job*)data;
  napi_value result = NULL;
  if (status == napi_ok && j->rc == 0) {
    switch (j->kind) {
      case J_READ_VELOCITY: {
        napi_value u = copy_typed(env, napi_float32_array, j->n, sizeof(float), j->a);
        napi_value v = copy_typed(env, napi_float32_array, j->n, sizeof(float), j->b);
        napi_create_object(env, &result);
        if (u && v) {
          napi_set_named_property(env, result, "u", u);
          napi_set_named_property(env, result, "v", v);
        }
        break;
      }
      case J_READ_FIELD:
        result = j->field == FS_F_PRESSURE ? copy_typed(env, napi_float64_array, j->n, sizeof(double), j->a)
                 : is_half_field(j->field) ? copy_typed(env, napi_uint16_array, j->n, sizeof(uint16_t), j->a)
                                           : copy_typed(env, napi_float32_array, j->n, sizeof(float), j->a);
        break;
      case J_MAX_DIV:
      case J_RESIDUAL_AFTER: napi_create_double(env, j->scalar, &result); break;
      case J_SPECTRUM: {
        napi_value E = copy_typed(env, napi_float64_array, (size_t)j->len, sizeof(double), j->a);
        double* c = (double*)malloc(sizeof(double) * (size_t)j->len);
        napi_value counts = NULL;
        if (c) {
          for (int k = 0; k < j->len; k++) c[k] = (double)((int64_t*)j->b)[k];
          counts = copy_typed(env, napi_float64_array, (size_t)j->len, sizeof(double), c);
          free(c);
        }
        napi_create_object(env, &result);
        if (E) napi_set_named_property(env, result, "E", E);
        if (counts) napi_set_named_property(env, result, "counts", counts);
        break;
      }
      default: napi_get_undefined(env, &result); break;
    }
    napi_resolve_deferred(env, j->deferred, result);
  } else {
    napi_value msg, errv;
    if (status != napi_ok) snprintf(j->err, sizeof j->err, "fluidsim: the readback was cancelled");
    napi_create_string_utf8(env, j->err, NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &errv);
    napi_reject_deferred(env, j->deferred, errv);
  }
  napi_delete_reference(env, j->href);
  napi_delete_async_work(env, j->work);
  free(j->a);
  free(j->b);
  free(j);
}

/* a Job for `kind` on the handle argv0, its ticket taken now (call order) */
static Job* new_job(napi_env env, napi_value hval, int kind) {
  Handle* h = get_handle(env, hval);
  if (!h) return NULL;
  Job* j = (Job*)calloc(1, sizeof(Job));
  if (!j) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  j->h = h;
  j->kind = kind;
  return j;
}

static napi_value start_job(napi_env env, napi_value hval, Job* j) {
  napi_value promise, name;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_reference(env, hval, 1, &j->href) != napi_ok ||
      napi_create_string_utf8(env, "fluidsim readback", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, job_execute, job_complete, j, &j->work) != napi_ok) {
    free(j->a);
    free(j->b);
    free(j);
    napi_throw_error(env, NULL, "cannot start the readback");
    return NULL;
  }
  j->ticket = take_ticket(j->h);
  if (napi_queue_async_work(env, j->work) != napi_ok) {
    napi_throw_error(env, NULL, "cannot queue the readback");
    return NULL;
  }
  return promise;
}

/* readVelocity(h, n) -> Promise<{u: Float32Array, v: Float32Array}> */
static napi_value js_read(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  double dn;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &dn));
  Job* j = new_job(env, argv[0], J_READ_VELOCITY);
  if (!j) return NULL;
  j->n = (size_t)dn;
  j->a = malloc(j->n * sizeof(float) + 1);
  j->b = malloc(j->n * sizeof(float) + 1);
  return start_job(env, argv[0], j);
}

/* readField(h, field, n) -> Promise<Float32Array | Float64Array (pressure) | Uint16Array
 * (output, bloom: n is their count of half floats)> */
static napi_value js_read_field(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  double f, dn;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &f));
  NAPI_CALL(env, napi_get_value_double(env, argv[2], &dn));
  Job* j = new_job(env, argv[0], J_READ_FIELD);
  if (!j) return NULL;
  j->field = (int)f;
  j->n = (size_t)dn;
  if (j->field == FS_F_DYE) j->n *= 4;
  j->bytes = j->n * (j->field == FS_F_PRESSURE ? sizeof(double) : is_half_field(j->field) ? sizeof(uint16_t)
                                                                                            : sizeof(float));
  j->a = malloc(j->bytes + 1);
  return start_job(env, argv[0], j);
}

/* maxDivergence(h) -> Promise<number> */
static napi_value js_max_div(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Job* j = new_job(env, argv[0], J_MAX_DIV);
  return j ? start_job(env, argv[0], j) : NULL;
}

/* poissonResidualAfter(h, it) -> Promise<number> */
static napi_value js_residual_after(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  double it;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &it));
  Job* j = new_job(env, argv[0], J_RESIDUAL_AFTER);
  if (!j) return NULL;
  j->iterations = (int)it;
  return start_job(env, argv[0], j);
}

/* energySpectrum(h, len) -> Promise<{E: Float64Array, counts: Float64Array}> */
static napi_value js_spectrum(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  double len;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &len));
  Job* j = new_job(env, argv[0], J_SPECTRUM);
  if (!j) return NULL;
  j->len = (int)len;
  j->a = calloc((size_t)j->len + 1, sizeof(double));
  j->b = calloc((size_t)j->len + 1, sizeof(int64_t));
  return start_job(env, argv[0], j);
}

/* syncAsync(h) -> Promise<undefined> once the submitted work is done */
static napi_value js_sync_async(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Job* j = new_job(env, argv[0], J_SYNC);
  return j ? start_job(env, argv[0], j) : NULL;
}

/* runKernel(h, kernel, dt): one kernel for per-kernel parity checks */
static napi_value js_run_kernel(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double k, dt;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &k));
  NAPI_CALL(env, napi_get_value_double(env, argv[2], &dt));
  lock(h);
  int rc = fs_run_kernel(h->solver, (int)k, dt);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* setForcingModes(h, kx[], ky[], phase[], amplitude, drag, phaseDrift) */
static napi_value js_forcing_modes(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  uint32_t count = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[1], &count));
  double* buf = (double*)malloc(sizeof(double) * 3 * (count + 1));
  if (!buf) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  for (int a = 0; a < 3; a++)
    for (uint32_t m = 0; m < count; m++) {
      napi_value e;
      if (napi_get_element(env, argv[1 + a], m, &e) != napi_ok ||
          napi_get_value_double(env, e, &buf[a * count + m]) != napi_ok) {
        free(buf);
        napi_throw_type_error(env, NULL, "kx, ky and phase must be arrays of numbers of one length");
        return NULL;
      }
    }
  double amp, drag, drift;
  if (napi_get_value_double(env, argv[4], &amp) != napi_ok || napi_get_value_double(env, argv[5], &drag) != napi_ok ||
      napi_get_value_double(env, argv[6], &drift) != napi_ok) {
    free(buf);
    napi_throw_type_error(env, NULL, "amplitude, drag and phaseDrift must be numbers");
    return NULL;
  }
  lock(h);
  int rc = fs_set_forcing_modes(h->solver, buf, buf + count, buf + 2 * count, (int)count, amp, drag, drift);
  unlock(h);
  free(buf);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* stepSplats(h, dt, [{x, y, dx, dy, r, g, b, radius, isDye}]) */
static napi_value js_step_splats(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double dt;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &dt));
  uint32_t count = 0;
  NAPI_CALL(env, napi_get_array_length(env, argv[2], &count));
  fs_splat* sp = (fs_splat*)calloc(count + 1, sizeof(fs_splat));
  if (!sp) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  for (uint32_t m = 0; m < count; m++) {
    napi_value o;
    if (napi_get_element(env, argv[2], m, &o) != napi_ok) {
      free(sp);
      napi_throw_type_error(env, NULL, "splats must be an array of objects");
      return NULL;
    }
    fs_splat* p = &sp[m];
    p->radius = 0.2;
    prop_double(env, o, "x", &p->x);
    prop_double(env, o, "y", &p->y);
    prop_double(env, o, "dx", &p->dx);
    prop_double(env, o, "dy", &p->dy);
    prop_double(env, o, "r", &p->r);
    prop_double(env, o, "g", &p->g);
    prop_double(env, o, "b", &p->b);
    double rad;
    if (prop_double(env, o, "radius", &rad) && rad) p->radius = rad;
    bool has = false;
    napi_value d;
    if (napi_has_named_property(env, o, "isDye", &has) == napi_ok && has &&
        napi_get_named_property(env, o, "isDye", &d) == napi_ok) {
      bool b = false;
      napi_value bv;
      if (napi_coerce_to_bool(env, d, &bv) == napi_ok && napi_get_value_bool(env, bv, &b) == napi_ok) p->is_dye = b;
    }
  }
  lock(h);
  int rc = fs_step_splats(h->solver, dt, sp, (int)count);
  unlock(h);
  free(sp);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* uploadDye(h, Float32Array(nx*ny*4)) */
static napi_value js_upload_dye(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  if (napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off) != napi_ok || t != napi_float32_array) {
    napi_throw_type_error(env, NULL, "the dye must be a Float32Array of nx*ny*4 rgba values");
    return NULL;
  }
  lock(h);
  int rc = fs_upload_dye(h->solver, (const float*)data, len);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double dt;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &dt));
  lock(h);
  int rc = fs_step(h->solver, dt);
  unlock(h);
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
  lock(h);
  int rc = fs_step_cfl(h->solver, cfl);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

/* render(h, mode, scale, bloomThreshold, bloomPasses) */
static napi_value js_render(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double mode, scale, thr, passes;
  NAPI_CALL(env, napi_get_value_double(env, argv[1], &mode));
  NAPI_CALL(env, napi_get_value_double(env, argv[2], &scale));
  NAPI_CALL(env, napi_get_value_double(env, argv[3], &thr));
  NAPI_CALL(env, napi_get_value_double(env, argv[4], &passes));
  lock(h);
  int rc = fs_render(h->solver, mode, scale, thr, (int)passes);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Handle* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  lock(h);
  int rc = fs_sync(h->solver);
  unlock(h);
  if (rc) return throw_fs(env, rc);
  return NULL;
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
  int which = !strcmp(key, "pressureIterations") ? 0 : !strcmp(key, "tolerance") ? 1 : !strcmp(key, "nu") ? 2
              : !strcmp(key, "guess") ? 3 : !strcmp(key, "confinement") ? 4 : -1;
  if (which < 0) {
    napi_throw_error(env, NULL, "unknown parameter");
    return NULL;
  }
  lock(h);
  int rc = which == 0 ? fs_set_pressure_iterations(h->solver, (int)val)
           : which == 1 ? fs_set_tolerance(h->solver, val)
           : which == 2 ? fs_set_nu(h->solver, val)
           : which == 3 ? fs_set_guess(h->solver, (int)val)
                        : fs_set_confinement(h->solver, val);
  unlock(h);
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
    lock(h);
    rc = fs_set_forcing(h->solver, NULL);
    unlock(h);
  } else {
    fs_forcing f = {28, 2, 0.4, 0, 1};
    double d;
    if (prop_double(env, argv[1], "kf", &d)) f.kf = d;
    if (prop_double(env, argv[1], "width", &d)) f.width = d;
    if (prop_double(env, argv[1], "amp", &d)) f.amp = d;
    if (prop_double(env, argv[1], "drag", &d)) f.drag = d;
    if (prop_double(env, argv[1], "seed", &d)) f.seed = (uint32_t)d;
    lock(h);
    rc = fs_set_forcing(h->solver, &f);
    unlock(h);
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
    lock(h);
    rc = fs_set_solid(h->solver, NULL);
    unlock(h);
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
    lock(h);
    rc = fs_set_solid(h->solver, &d);
    unlock(h);
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
  lock(h);
  int rc = fs_solid_impulse(h->solver, imp);
  unlock(h);
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
  lock(h);
  int rc = fs_get_stats(h->solver, &st);
  unlock(h);
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
  lock(h);
  int rc = fs_diagnostics(h->solver, &d);
  unlock(h);
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
      {"solidImpulse", js_solid_impulse}, {"loadGuess", js_load_guess}, {"readField", js_read_field},
      {"energySpectrum", js_spectrum}, {"syncAsync", js_sync_async}, {"runKernel", js_run_kernel},
      {"setForcingModes", js_forcing_modes}, {"stepSplats", js_step_splats}, {"uploadDye", js_upload_dye},
      {"render", js_render},
  };
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; k++) {
    napi_value f;
    NAPI_CALL(env, napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].cb, NULL, &f));
    NAPI_CALL(env, napi_set_named_property(env, exports, fns[k].name, f));
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

// Thin N-API binding of liblbm.so for Node (N-API version 8, Node >= 12).
// Every function maps one-to-one onto an include/lbm.h entry point; errors
// become JavaScript exceptions carrying lbm_last_error().  lbm_step runs on
// the libuv thread pool and resolves a Promise, so a host loop can await it
// the way the reference awaits queue.onSubmittedWorkDone().
#define NAPI_VERSION 8
#include <node_api.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <string>

#include "lbm.h"

namespace {

#define NAPI_OK(call)                                       \
  do {                                                      \
    if ((call) != napi_ok) {                                \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                       \
    }                                                       \
  } while (0)

napi_value throw_lbm(napi_env env, int rc) {
  std::string msg = "lbm error " + std::to_string(rc) + ": " + lbm_last_error();
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool get_prop(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

double num(napi_env env, napi_value obj, const char* name, double dflt) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

void num_array(napi_env env, napi_value obj, const char* name, double* out, int n) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return;
  for (int i = 0; i < n; i++) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) == napi_ok) {
      double d;
      if (napi_get_value_double(env, e, &d) == napi_ok) out[i] = d;
    }
  }
}

// Every engine object reaches JavaScript as an external holding a Handle:
// the object's kind, the element counts its buffers must hold (checked
// before native code touches a typed array) and the number of asynchronous
// calls in flight on it.  While one is in flight, destroy and the
// synchronous calls throw instead of racing with the libuv thread.  The
// Handle outlives destroy (marked dead) and is freed by the GC finalizer.
enum Kind { kSolve = 1, kAmr = 2, kServe = 3 };
struct Handle {
  void* p = nullptr;
  Kind kind = kSolve;
  size_t cells = 0;       // solve: lattice cells; amr: L0 cells
  int q = 0;
  size_t nblocks = 0;     // amr: L0 blocks
  size_t slots = 0;       // amr: pool slots
  int levels = 0;         // amr: levels of the hierarchy (L0 included)
  std::atomic<int> busy{0};
  bool dead = false;
};

void handle_finalize(napi_env, void* data, void*) {
  Handle* h = static_cast<Handle*>(data);
  if (h->busy.load() == 0) delete h;  // else in-flight work still reads it (a few bytes leak)
}

napi_value wrap(napi_env env, Handle* h) {
  napi_value ext;
  if (napi_create_external(env, h, handle_finalize, nullptr, &ext) != napi_ok) {
    napi_throw_error(env, nullptr, "N-API call failed: napi_create_external");
    return nullptr;
  }
  return ext;
}

// The Handle of argument v, or null with a pending exception: not a handle
// of this kind, destroyed, or (idle) an asynchronous call still in flight.
Handle* get_handle(napi_env env, napi_value v, Kind kind, bool idle) {
  static const char* names[] = {"", "solve", "AMR solve", "server"};
  void* p = nullptr;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external || napi_get_value_external(env, v, &p) != napi_ok ||
      !p || static_cast<Handle*>(p)->kind != kind) {
    napi_throw_type_error(env, nullptr, (std::string("expected a ") + names[kind] + " handle").c_str());
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (h->dead) {
    napi_throw_error(env, nullptr, (std::string("the ") + names[kind] + " was destroyed").c_str());
    return nullptr;
  }
  if (idle && h->busy.load() != 0) {
    napi_throw_error(env, nullptr,
                     (std::string("the ") + names[kind] + " has an asynchronous call in flight; await it first").c_str());
    return nullptr;
  }
  return h;
}

// Throws a RangeError unless a typed array of `len` elements holds `need`.
bool long_enough(napi_env env, size_t len, size_t need, const char* what) {
  if (len >= need) return true;
  std::string m = std::string(what) + ": the array holds " + std::to_string(len) + " elements, " +
                  std::to_string(need) + " needed";
  napi_throw_range_error(env, nullptr, m.c_str());
  return false;
}

#define GET(var, arg, kind, idle)                      \
  Handle* var##_h = get_handle(env, arg, kind, idle); \
  if (!var##_h) return nullptr

// createSolve({q, nx, ny, nz, collision: 'bgk'|'trt', omega, omegaMinus,
//              force, walls, wallU, sponge: {width, axes, u},
//              body: {a, centre} | null, solidEq, diffuse, chiEps, storage: 'f32'|'f16'|'bf16',
//              autotune})
// Fills an lbm_desc from a JS descriptor object; false (with a pending JS
// exception) on a bad field.
bool parse_desc(napi_env env, napi_value o, lbm_desc& d) {
  memset(&d, 0, sizeof d);
  d.abi_version = LBM_ABI_VERSION;
  d.q = (int)num(env, o, "q", 19);
  d.nx = (int)num(env, o, "nx", 1);
  d.ny = (int)num(env, o, "ny", 1);
  d.nz = (int)num(env, o, "nz", 1);
  d.omega = (float)num(env, o, "omega", 1.0);
  napi_value v;
  d.collision = LBM_BGK;
  if (get_prop(env, o, "omegaMinus", &v)) {
    d.collision = LBM_TRT;
    d.omega_minus = (float)num(env, o, "omegaMinus", d.omega);
  }
  double t3[3] = {0, 0, 0};
  num_array(env, o, "force", t3, 3);
  for (int k = 0; k < 3; k++) d.force[k] = (float)t3[k];
  double w3[3] = {0, 0, 0};
  num_array(env, o, "walls", w3, 3);
  for (int k = 0; k < 3; k++) d.walls[k] = w3[k] != 0;
  double wu[2] = {0, 0};
  num_array(env, o, "wallU", wu, 2);
  d.wall_u[0] = (float)wu[0];
  d.wall_u[1] = (float)wu[1];
  d.sponge_axes[0] = d.sponge_axes[1] = 1;
  d.sponge_axes[2] = d.q == 9 ? 0 : 1;
  if (get_prop(env, o, "sponge", &v)) {
    d.sponge_width = (float)num(env, v, "width", 0);
    double ax[3] = {1, 1, d.q == 9 ? 0.0 : 1.0}, su[3] = {0, 0, 0};
    num_array(env, v, "axes", ax, 3);
    num_array(env, v, "u", su, 3);
    for (int k = 0; k < 3; k++) {
      d.sponge_axes[k] = ax[k] != 0;
      d.sponge_u[k] = (float)su[k];
    }
  }
  if (get_prop(env, o, "body", &v)) {
    d.body.shape = d.q == 9 ? LBM_BODY_CIRCLE : LBM_BODY_SPHERE;
    d.body.radius = (float)num(env, v, "a", 0);
    double c[3] = {0, 0, 0};
    num_array(env, v, "centre", c, 3);
    for (int k = 0; k < 3; k++) d.body.centre[k] = (float)c[k];
  }
  d.solid_eq = (int)num(env, o, "solidEq", 1);
  d.autotune = (int)num(env, o, "autotune", 0);
  d.diffuse = (int)num(env, o, "diffuse", 0);
  d.chi_eps = (float)num(env, o, "chiEps", 0);
  d.storage = LBM_STORE_F32;
  if (get_prop(env, o, "storage", &v)) {
    char buf[16] = {0};
    size_t len = 0;
    napi_get_value_string_utf8(env, v, buf, sizeof buf, &len);
    if (!strcmp(buf, "f16")) d.storage = LBM_STORE_F16;
    else if (!strcmp(buf, "bf16")) d.storage = LBM_STORE_BF16;
    else if (strcmp(buf, "f32")) {
      napi_throw_type_error(env, nullptr, "storage must be 'f32', 'f16' or 'bf16'");
      return false;
    }
  }
  return true;
}

napi_value CreateSolve(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) {
    napi_throw_type_error(env, nullptr, "createSolve(desc)");
    return nullptr;
  }
  lbm_desc d;
  if (!parse_desc(env, argv[0], d)) return nullptr;
  lbm_solve* s = nullptr;
  int rc = lbm_solve_create(&d, &s);
  if (rc) return throw_lbm(env, rc);
  Handle* h = new Handle();
  h->p = s;
  h->kind = kSolve;
  h->q = d.q;
  h->cells = (size_t)d.nx * d.ny * d.nz;
  return wrap(env, h);
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(s, argv[0], kSolve, true);
  lbm_solve_destroy(static_cast<lbm_solve*>(s_h->p));
  s_h->p = nullptr;
  s_h->dead = true;
  return nullptr;
}

lbm_solve* solve_of(Handle* h) { return static_cast<lbm_solve*>(h->p); }

// Typed-array argument as (data, length in elements).
bool typed(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, len, data, &ab, &off);
  return t == want;
}

napi_value UploadMacro(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "uploadMacro(solve, Float32Array)");
    return nullptr;
  }
  GET(s, argv[0], kSolve, true);
  if (!long_enough(env, len, 4 * s_h->cells, "uploadMacro")) return nullptr;
  int rc = lbm_upload_macro(solve_of(s_h), static_cast<float*>(data));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

napi_value UploadF(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "uploadF(solve, Float32Array)");
    return nullptr;
  }
  GET(s, argv[0], kSolve, true);
  if (!long_enough(env, len, (size_t)s_h->q * s_h->cells, "uploadF")) return nullptr;
  int rc = lbm_upload_f(solve_of(s_h), static_cast<float*>(data));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

// Bytes of a Uint8Array or an ArrayBuffer.
bool bytes_of(napi_env env, napi_value v, void** data, size_t* len) {
  bool is = false;
  napi_is_arraybuffer(env, v, &is);
  if (is) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  return typed(env, v, napi_uint8_array, data, len);
}

// saveState(solve) -> ArrayBuffer (lbm_state_header + populations)
napi_value SaveState(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(sh, argv[0], kSolve, true);
  lbm_solve* s = solve_of(sh_h);
  size_t n = 0;
  int rc = lbm_state_size(s, &n);
  if (rc) return throw_lbm(env, rc);
  void* data = nullptr;
  napi_value ab;
  NAPI_OK(napi_create_arraybuffer(env, n, &data, &ab));
  rc = lbm_save_state(s, data, n);
  if (rc) return throw_lbm(env, rc);
  return ab;
}

// loadState(solve, ArrayBuffer | Uint8Array)
napi_value LoadState(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !bytes_of(env, argv[1], &data, &len)) {
    napi_throw_type_error(env, nullptr, "loadState(solve, ArrayBuffer | Uint8Array)");
    return nullptr;
  }
  GET(s, argv[0], kSolve, true);
  int rc = lbm_load_state(solve_of(s_h), data, len);
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

// readInto(solve, TypedArray): fills a caller-sized array of PerCell floats
// per cell (0: q, the populations).
template <int (*Fn)(lbm_solve*, float*), int PerCell>
napi_value ReadFloat(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "expected (solve, Float32Array)");
    return nullptr;
  }
  GET(s, argv[0], kSolve, true);
  if (!long_enough(env, len, (size_t)(PerCell ? PerCell : s_h->q) * s_h->cells, "read")) return nullptr;
  int rc = Fn(solve_of(s_h), static_cast<float*>(data));
  if (rc) return throw_lbm(env, rc);
  return argv[1];
}

napi_value ReadTotals(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  napi_value ab, arr;
  void* data;
  GET(s, argv[0], kSolve, true);
  NAPI_OK(napi_create_arraybuffer(env, 32, &data, &ab));
  int rc = lbm_read_totals(solve_of(s_h), static_cast<double*>(data));
  if (rc) return throw_lbm(env, rc);
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, 4, ab, 0, &arr));
  return arr;
}

napi_value ReadForces(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t maxr = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &maxr);
  if (maxr < 0) maxr = 0;
  napi_value ab, arr;
  void* data;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)maxr * 48 + 8, &data, &ab));
  GET(s, argv[0], kSolve, true);
  int32_t count = 0;
  int rc = lbm_read_forces(solve_of(s_h), static_cast<double*>(data), maxr, &count);
  if (rc) return throw_lbm(env, rc);
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)count * 6, ab, 0, &arr));
  return arr;
}

napi_value ReadClasses(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_uint8_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "readClasses(solve, Uint8Array)");
    return nullptr;
  }
  GET(s, argv[0], kSolve, true);
  if (!long_enough(env, len, s_h->cells, "readClasses")) return nullptr;
  int rc = lbm_read_classes(solve_of(s_h), static_cast<uint8_t*>(data));
  if (rc) return throw_lbm(env, rc);
  return argv[1];
}

struct StepWork {
  napi_async_work work;
  napi_deferred deferred;
  Handle* h;
  int n;
  int rc;
  std::string err;
};

void StepExecute(napi_env, void* p) {
  StepWork* w = static_cast<StepWork*>(p);
  lbm_solve* s = solve_of(w->h);
  w->rc = lbm_step(s, w->n);
  if (!w->rc) w->rc = lbm_sync(s);
  if (w->rc) w->err = lbm_last_error();
}

void StepComplete(napi_env env, napi_status, void* p) {
  StepWork* w = static_cast<StepWork*>(p);
  w->h->busy.fetch_sub(1);
  if (w->rc) {
    napi_value msg, err;
    std::string m = "lbm error " + std::to_string(w->rc) + ": " + w->err;
    napi_create_string_utf8(env, m.c_str(), m.size(), &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, w->deferred, undef);
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

// Queues async work that holds `h` busy until its complete callback.
bool queue_busy(napi_env env, napi_async_work work, Handle* h) {
  h->busy.fetch_add(1);
  if (napi_queue_async_work(env, work) == napi_ok) return true;
  h->busy.fetch_sub(1);
  napi_throw_error(env, nullptr, "N-API call failed: napi_queue_async_work");
  return false;
}

// step(solve, n) -> Promise resolved once the n steps have completed.
napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(s, argv[0], kSolve, true);
  StepWork* w = new StepWork();
  w->h = s_h;
  w->n = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &w->n);
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "lbm_step", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, StepExecute, StepComplete, w, &w->work));
  if (!queue_busy(env, w->work, s_h)) return nullptr;
  return promise;
}

napi_value Info(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int64_t step = 0, bytes = 0;
  int32_t bx = 0, by = 0, flags = 0;
  GET(s, argv[0], kSolve, true);
  int rc = lbm_solve_info(solve_of(s_h), &step, &bx, &by, &flags, &bytes);
  if (rc) return throw_lbm(env, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  napi_create_double(env, (double)step, &v);
  napi_set_named_property(env, o, "step", v);
  napi_create_int32(env, bx, &v);
  napi_set_named_property(env, o, "blockX", v);
  napi_create_int32(env, by, &v);
  napi_set_named_property(env, o, "blockY", v);
  napi_create_int32(env, flags, &v);
  napi_set_named_property(env, o, "flags", v);
  napi_create_double(env, (double)bytes, &v);
  napi_set_named_property(env, o, "deviceBytes", v);
  return o;
}

// rigidAttach(solve, {shape: {kind, a, b, c, r}, x, q, v, omega, density | mass,
//                     pinned, gravity, vMax, oMax, planar, noFluidForce, swept,
//                     rhoWLocal, window, anchor})
// The body object of d3-body.mjs makeBodyState (mass, not density) is taken as
// it is; the flags are the physics pipeline's overrides (main-3d.js:1060-1070).
napi_value RigidAttach(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  lbm_rigid_body b;
  memset(&b, 0, sizeof b);
  napi_value sh;
  if (!get_prop(env, argv[1], "shape", &sh)) {
    napi_throw_type_error(env, nullptr, "rigid body needs a shape");
    return nullptr;
  }
  b.shape = (int)num(env, sh, "kind", 0);
  b.a = num(env, sh, "a", 0);
  b.b = num(env, sh, "b", b.a);
  b.c = num(env, sh, "c", b.a);
  b.r = num(env, sh, "r", 0);
  const double pi = 3.14159265358979323846;
  const double vol = b.shape == LBM_SHAPE_SPHERE ? (4.0 / 3.0) * pi * b.a * b.a * b.a
                   : b.shape == LBM_SHAPE_SPHEROID ? (4.0 / 3.0) * pi * b.a * b.a * b.c : 8.0 * b.a * b.b * b.c;
  napi_value v;
  b.density = get_prop(env, argv[1], "mass", &v) && !get_prop(env, argv[1], "density", &v)
                  ? num(env, argv[1], "mass", 0) / vol
                  : num(env, argv[1], "density", 1.0);
  num_array(env, argv[1], "x", b.x, 3);
  b.q[0] = 1.0;
  num_array(env, argv[1], "q", b.q, 4);
  num_array(env, argv[1], "v", b.v, 3);
  num_array(env, argv[1], "omega", b.omega, 3);
  num_array(env, argv[1], "gravity", b.gravity, 3);
  b.pinned = num(env, argv[1], "pinned", 0) != 0;
  b.v_max = num(env, argv[1], "vMax", 0.2);
  b.o_max = num(env, argv[1], "oMax", 0.5);
  b.planar = num(env, argv[1], "planar", 0) != 0;
  b.no_fluid_force = num(env, argv[1], "noFluidForce", 0) != 0;
  b.swept_force = num(env, argv[1], "swept", 1) != 0;
  b.rho_w_local = num(env, argv[1], "rhoWLocal", 1) != 0;
  double w3[3] = {0, 0, 0};
  num_array(env, argv[1], "window", w3, 3);
  for (int k = 0; k < 3; k++) b.window[k] = w3[k] != 0;
  num_array(env, argv[1], "anchor", b.window_anchor, 3);
  GET(s, argv[0], kSolve, true);
  int rc = lbm_rigid_attach(solve_of(s_h), &b);
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

napi_value RigidRead(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  lbm_rigid_state st;
  GET(s, argv[0], kSolve, true);
  int rc = lbm_rigid_read(solve_of(s_h), &st);
  if (rc) return throw_lbm(env, rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  auto arr = [&](const char* name, const double* p, int n) {
    napi_value a, e;
    napi_create_array_with_length(env, n, &a);
    for (int k = 0; k < n; k++) {
      napi_create_double(env, p[k], &e);
      napi_set_element(env, a, k, e);
    }
    napi_set_named_property(env, o, name, a);
  };
  arr("x", st.x, 3);
  arr("q", st.q, 4);
  arr("v", st.v, 3);
  arr("L", st.L, 3);
  arr("omega", st.omega, 3);
  arr("d", st.d, 3);
  arr("force", st.force, 3);
  arr("torque", st.torque, 3);
  arr("inertia", st.inertia, 3);
  napi_value m;
  napi_create_double(env, st.mass, &m);
  napi_set_named_property(env, o, "mass", m);
  return o;
}

napi_value DeviceInfo(napi_env env, napi_callback_info) {
  char name[256] = {0};
  double hbm = 0;
  int32_t cus = 0;
  int rc = lbm_device_info(name, sizeof name, &hbm, &cus);
  if (rc) return throw_lbm(env, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &v);
  napi_set_named_property(env, o, "name", v);
  napi_create_double(env, hbm, &v);
  napi_set_named_property(env, o, "hbmBytes", v);
  napi_create_int32(env, cus, &v);
  napi_set_named_property(env, o, "computeUnits", v);
  return o;
}

// ---------------------------------------------------------------------------
// Two-level AMR solve (lbm_amr_*)
// ---------------------------------------------------------------------------
lbm_amr* amr_of(Handle* h) { return static_cast<lbm_amr*>(h->p); }

// createAmr({q, nx, ny, nz, rb, maxSlots, omega, omegaMinus, force,
//            manageEvery, refineAt, keepAt, lead, explodeLinear, orphans,
//            interface: 'explode' | 'interp', reflux, explodeDirs,
//            criterion: 'vorticity' | 'q', sponge: {width, axes, u},
//            body: {a, centre}, levels, ladderOctaves})
napi_value CreateAmr(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) {
    napi_throw_type_error(env, nullptr, "createAmr(desc)");
    return nullptr;
  }
  napi_value o = argv[0], v;
  lbm_amr_desc d;
  memset(&d, 0, sizeof d);
  d.abi_version = LBM_ABI_VERSION;
  d.q = (int)num(env, o, "q", 27);
  d.nx = (int)num(env, o, "nx", 1);
  d.ny = (int)num(env, o, "ny", 1);
  d.nz = (int)num(env, o, "nz", 1);
  d.rb = (int)num(env, o, "rb", 8);
  d.max_slots = (int)num(env, o, "maxSlots", 1);
  d.omega = (float)num(env, o, "omega", 1.0);
  d.collision = LBM_BGK;
  if (get_prop(env, o, "omegaMinus", &v)) {
    d.collision = LBM_TRT;
    d.omega_minus = (float)num(env, o, "omegaMinus", d.omega);
  }
  double t3[3] = {0, 0, 0};
  num_array(env, o, "force", t3, 3);
  for (int k = 0; k < 3; k++) d.force[k] = (float)t3[k];
  d.manage_every = (int)num(env, o, "manageEvery", 0);
  d.refine_at = (float)num(env, o, "refineAt", 1e30);
  d.keep_at = (float)num(env, o, "keepAt", 1e30);
  d.lead = (int)num(env, o, "lead", 0);
  d.explode_linear = (int)num(env, o, "explodeLinear", 1);
  d.orphans = (int)num(env, o, "orphans", 1);
  napi_value iv;
  if (get_prop(env, o, "interface", &iv)) {
    char buf[16] = {0};
    size_t len = 0;
    napi_get_value_string_utf8(env, iv, buf, sizeof buf, &len);
    if (!strcmp(buf, "interp")) d.interface_interp = 1;
    else if (strcmp(buf, "explode")) {
      napi_throw_type_error(env, nullptr, "interface must be 'explode' or 'interp'");
      return nullptr;
    }
  }
  d.reflux = (int)num(env, o, "reflux", 0);
  d.explode_dirs = (int)num(env, o, "explodeDirs", 0);
  d.levels = (int)num(env, o, "levels", 2);
  d.ladder_octaves = (float)num(env, o, "ladderOctaves", 0);
  napi_value cv;
  if (get_prop(env, o, "criterion", &cv)) {
    char buf[16] = {0};
    size_t len = 0;
    napi_get_value_string_utf8(env, cv, buf, sizeof buf, &len);
    if (!strcmp(buf, "q")) d.criterion = 1;
    else if (strcmp(buf, "vorticity")) {
      napi_throw_type_error(env, nullptr, "criterion must be 'vorticity' or 'q'");
      return nullptr;
    }
  }
  if (get_prop(env, o, "sponge", &v)) {
    d.sponge_width = (float)num(env, v, "width", 0);
    double ax[3] = {1, 1, 1}, su[3] = {0, 0, 0};
    num_array(env, v, "axes", ax, 3);
    num_array(env, v, "u", su, 3);
    for (int k = 0; k < 3; k++) {
      d.sponge_axes[k] = ax[k] != 0;
      d.sponge_u[k] = (float)su[k];
    }
  }
  if (get_prop(env, o, "body", &v)) {
    d.body.shape = LBM_BODY_SPHERE;
    d.body.radius = (float)num(env, v, "a", 0);
    double c[3] = {0, 0, 0};
    num_array(env, v, "centre", c, 3);
    for (int k = 0; k < 3; k++) d.body.centre[k] = (float)c[k];
  }
  lbm_amr* a = nullptr;
  int rc = lbm_amr_create(&d, &a);
  if (rc) return throw_lbm(env, rc);
  Handle* h = new Handle();
  h->p = a;
  h->kind = kAmr;
  h->q = d.q;
  h->cells = (size_t)d.nx * d.ny * d.nz;
  // the D2Q9 quadtree (q 9, nz 1) tiles the xy plane only
  h->nblocks = d.rb > 0 ? (size_t)(d.nx / d.rb) * (d.ny / d.rb) * (d.q == 9 ? 1 : d.nz / d.rb) : 0;
  h->slots = (size_t)d.max_slots;
  h->levels = d.levels;
  return wrap(env, h);
}

napi_value AmrDestroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(a, argv[0], kAmr, true);
  lbm_amr_destroy(amr_of(a_h));
  a_h->p = nullptr;
  a_h->dead = true;
  return nullptr;
}

// amrUploadTopology(amr, Int32Array blockSlot, Int32Array slotToBlock) and
// amrReadTopology(amr, Int32Array, Int32Array)
template <bool Upload>
napi_value AmrTopology(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void *bs, *sb;
  size_t nb, ns;
  if (argc < 3 || !typed(env, argv[1], napi_int32_array, &bs, &nb) || !typed(env, argv[2], napi_int32_array, &sb, &ns)) {
    napi_throw_type_error(env, nullptr, "expected (amr, Int32Array blockSlot, Int32Array slotToBlock)");
    return nullptr;
  }
  GET(ah, argv[0], kAmr, true);
  if (!long_enough(env, nb, ah_h->nblocks, "blockSlot") || !long_enough(env, ns, ah_h->slots, "slotToBlock"))
    return nullptr;
  lbm_amr* a = amr_of(ah_h);
  const int rc = Upload ? lbm_amr_upload_topology(a, static_cast<int32_t*>(bs), static_cast<int32_t*>(sb))
                        : lbm_amr_read_topology(a, static_cast<int32_t*>(bs), static_cast<int32_t*>(sb));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

// amrUploadLevelTopology(amr, level, Int32Array blockSlot, Int32Array
// slotToBlock): level m's maps over level m - 1's global grid (L0 dims times
// 2^(m-1), blocks of rb of its cells), the reference's per-level
// uploadLevelTopology (main-3d.js:822) fed from refineHierarchy's byLevel.
napi_value AmrUploadLevelTopology(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void *bs, *sb;
  size_t nb, ns;
  int32_t level = 0;
  if (argc < 4 || napi_get_value_int32(env, argv[1], &level) != napi_ok ||
      !typed(env, argv[2], napi_int32_array, &bs, &nb) || !typed(env, argv[3], napi_int32_array, &sb, &ns)) {
    napi_throw_type_error(env, nullptr, "expected (amr, level, Int32Array blockSlot, Int32Array slotToBlock)");
    return nullptr;
  }
  GET(ah, argv[0], kAmr, true);
  if (level < 1 || level >= ah_h->levels) {
    napi_throw_range_error(env, nullptr, "level outside 1 .. levels - 1");
    return nullptr;
  }
  // blocks at level m: the L0 block count times 8^(m-1) (4^(m-1) in 2D)
  const size_t want_blocks = ah_h->nblocks << ((ah_h->q == 9 ? 2 : 3) * (level - 1));
  if (!long_enough(env, nb, want_blocks, "blockSlot") || !long_enough(env, ns, ah_h->slots, "slotToBlock"))
    return nullptr;
  const int rc = lbm_amr_upload_level_topology(amr_of(ah_h), level, static_cast<int32_t*>(bs),
                                               static_cast<int32_t*>(sb));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

// amrUploadMacro / amrReadMacro (amr, Float32Array)
template <int (*Fn)(lbm_amr*, float*)>
napi_value AmrFloat(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "expected (amr, Float32Array)");
    return nullptr;
  }
  GET(a, argv[0], kAmr, true);
  if (!long_enough(env, len, 4 * a_h->cells, "AMR macroscopic field")) return nullptr;
  int rc = Fn(amr_of(a_h), static_cast<float*>(data));
  if (rc) return throw_lbm(env, rc);
  return argv[1];
}

int amr_upload_macro(lbm_amr* a, float* m) { return lbm_amr_upload_macro(a, m); }

napi_value AmrReadCriterion(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void *crit, *want;
  size_t n1, n2;
  if (argc < 3 || !typed(env, argv[1], napi_float32_array, &crit, &n1) || !typed(env, argv[2], napi_uint8_array, &want, &n2)) {
    napi_throw_type_error(env, nullptr, "amrReadCriterion(amr, Float32Array, Uint8Array)");
    return nullptr;
  }
  GET(a, argv[0], kAmr, true);
  if (!long_enough(env, n1, a_h->nblocks, "criterion") || !long_enough(env, n2, a_h->nblocks, "want"))
    return nullptr;
  int rc = lbm_amr_read_criterion(amr_of(a_h), static_cast<float*>(crit), static_cast<uint8_t*>(want));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

napi_value AmrReadActive(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 2 || !typed(env, argv[1], napi_int32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "amrReadActive(amr, Int32Array(maxSlots + 1))");
    return nullptr;
  }
  GET(a, argv[0], kAmr, true);
  if (!long_enough(env, len, a_h->slots + 1, "amrReadActive")) return nullptr;
  int rc = lbm_amr_read_active(amr_of(a_h), static_cast<int32_t*>(data));
  if (rc) return throw_lbm(env, rc);
  return argv[1];
}

// amrReadForces(amr, maxRecords) -> Float64Array [fx, fy, fz, tx, ty, tz] * n
napi_value AmrReadForces(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t maxr = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &maxr);
  if (maxr < 0) maxr = 0;
  napi_value ab, arr;
  void* data;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)maxr * 48 + 8, &data, &ab));
  int32_t count = 0;
  GET(a, argv[0], kAmr, true);
  int rc = lbm_amr_read_forces(amr_of(a_h), static_cast<double*>(data), maxr, &count);
  if (rc) return throw_lbm(env, rc);
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)count * 6, ab, 0, &arr));
  return arr;
}

napi_value AmrReadTotals(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  napi_value ab, arr;
  void* data;
  GET(a, argv[0], kAmr, true);
  NAPI_OK(napi_create_arraybuffer(env, 32, &data, &ab));
  int rc = lbm_amr_read_totals(amr_of(a_h), static_cast<double*>(data));
  if (rc) return throw_lbm(env, rc);
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, 4, ab, 0, &arr));
  return arr;
}

napi_value AmrManageNow(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(a, argv[0], kAmr, true);
  int rc = lbm_amr_manage_now(amr_of(a_h));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

napi_value AmrInfo(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int64_t step = 0;
  int32_t act = 0, exh = 0;
  GET(a, argv[0], kAmr, true);
  int rc = lbm_amr_info(amr_of(a_h), &step, &act, &exh);
  if (rc) return throw_lbm(env, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  napi_create_double(env, (double)step, &v);
  napi_set_named_property(env, o, "step", v);
  napi_create_int32(env, act, &v);
  napi_set_named_property(env, o, "activeTiles", v);
  napi_create_int32(env, exh, &v);
  napi_set_named_property(env, o, "exhausted", v);
  return o;
}

struct AmrStepWork {
  napi_async_work work;
  napi_deferred deferred;
  Handle* h;
  int n;
  int rc;
  std::string err;
};

void AmrStepExecute(napi_env, void* p) {
  AmrStepWork* w = static_cast<AmrStepWork*>(p);
  w->rc = lbm_amr_step(amr_of(w->h), w->n);
  if (w->rc) w->err = lbm_last_error();
}

void AmrStepComplete(napi_env env, napi_status, void* p) {
  AmrStepWork* w = static_cast<AmrStepWork*>(p);
  w->h->busy.fetch_sub(1);
  if (w->rc) {
    napi_value msg, err;
    std::string m = "lbm error " + std::to_string(w->rc) + ": " + w->err;
    napi_create_string_utf8(env, m.c_str(), m.size(), &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, w->deferred, undef);
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

// amrStep(amr, n) -> Promise resolved once the n macro-steps have completed.
napi_value AmrStep(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(a, argv[0], kAmr, true);
  AmrStepWork* w = new AmrStepWork();
  w->h = a_h;
  w->n = 1;
  if (argc > 1) napi_get_value_int32(env, argv[1], &w->n);
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "lbm_amr_step", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, AmrStepExecute, AmrStepComplete, w, &w->work));
  if (!queue_busy(env, w->work, a_h)) return nullptr;
  return promise;
}

// ---------------------------------------------------------------------------
// Batched serving (lbm_serve_*)
// ---------------------------------------------------------------------------
lbm_serve* serve_of(Handle* h) { return static_cast<lbm_serve*>(h->p); }

int64_t ticket_of(napi_env env, napi_value v) {
  int64_t t = 0;
  napi_get_value_int64(env, v, &t);
  return t;
}

// serveCreate({budgetBytes, maxResident, chunkSteps, autotune, targetCells, persistent})
napi_value ServeCreate(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) {
    napi_throw_type_error(env, nullptr, "serveCreate(desc)");
    return nullptr;
  }
  lbm_serve_desc d;
  memset(&d, 0, sizeof d);
  d.abi_version = LBM_ABI_VERSION;
  d.budget_bytes = num(env, argv[0], "budgetBytes", 0);
  d.max_resident = (int)num(env, argv[0], "maxResident", 0);
  d.chunk_steps = (int)num(env, argv[0], "chunkSteps", 32);
  d.autotune = (int)num(env, argv[0], "autotune", 0);
  d.target_cells = num(env, argv[0], "targetCells", 0);
  d.persistent = (int)num(env, argv[0], "persistent", 0);
  lbm_serve* sv = nullptr;
  int rc = lbm_serve_create(&d, &sv);
  if (rc) return throw_lbm(env, rc);
  Handle* h = new Handle();
  h->p = sv;
  h->kind = kServe;
  return wrap(env, h);
}

napi_value ServeDestroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(sv, argv[0], kServe, true);  // no serveWait in flight
  lbm_serve_destroy(serve_of(sv_h));
  sv_h->p = nullptr;
  sv_h->dead = true;
  return nullptr;
}

// serveSubmit(server, desc, Float32Array mac, steps) -> ticket
napi_value ServeSubmit(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 4 || !typed(env, argv[2], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "serveSubmit(server, desc, Float32Array, steps)");
    return nullptr;
  }
  lbm_desc d;
  if (!parse_desc(env, argv[1], d)) return nullptr;
  const size_t cells = (size_t)d.nx * d.ny * (d.q == 9 ? 1 : d.nz);
  if (len < 4 * cells) {
    napi_throw_range_error(env, nullptr, "mac must hold 4 floats per cell");
    return nullptr;
  }
  int32_t steps = 0;
  napi_get_value_int32(env, argv[3], &steps);
  int64_t t = 0;
  GET(sv, argv[0], kServe, false);
  int rc = lbm_serve_submit(serve_of(sv_h), &d, static_cast<float*>(data), steps, &t);
  if (rc) return throw_lbm(env, rc);
  napi_value v;
  NAPI_OK(napi_create_int64(env, t, &v));
  return v;
}

// serveSubmitState(server, desc, ArrayBuffer | Uint8Array state, steps, keepState) -> ticket
// serveSubmitState(server, desc, Float32Array mac, steps, keepState) -> ticket
napi_value ServeSubmitState(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 5) {
    napi_throw_type_error(env, nullptr, "serveSubmitState(server, desc, state | mac, steps, keepState)");
    return nullptr;
  }
  lbm_desc d;
  if (!parse_desc(env, argv[1], d)) return nullptr;
  void* data;
  size_t len;
  const float* mac = nullptr;
  const void* state = nullptr;
  size_t state_bytes = 0;
  if (typed(env, argv[2], napi_float32_array, &data, &len)) {
    const size_t cells = (size_t)d.nx * d.ny * (d.q == 9 ? 1 : d.nz);
    if (len < 4 * cells) {
      napi_throw_range_error(env, nullptr, "mac must hold 4 floats per cell");
      return nullptr;
    }
    mac = static_cast<const float*>(data);
  } else if (bytes_of(env, argv[2], &data, &len)) {
    state = data;
    state_bytes = len;
  } else {
    napi_throw_type_error(env, nullptr, "state must be an ArrayBuffer or Uint8Array, mac a Float32Array");
    return nullptr;
  }
  int32_t steps = 0;
  napi_get_value_int32(env, argv[3], &steps);
  bool keep = false;
  napi_get_value_bool(env, argv[4], &keep);
  int64_t t = 0;
  GET(sv, argv[0], kServe, false);
  int rc = lbm_serve_submit_state(serve_of(sv_h), &d, mac, state, state_bytes, steps, keep ? 1 : 0, &t);
  if (rc) return throw_lbm(env, rc);
  napi_value v;
  NAPI_OK(napi_create_int64(env, t, &v));
  return v;
}

// serveReadState(server, ticket) -> ArrayBuffer
napi_value ServeReadState(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(svh, argv[0], kServe, false);
  lbm_serve* sv = serve_of(svh_h);
  const int64_t t = ticket_of(env, argv[1]);
  size_t n = 0;
  int rc = lbm_serve_read_state(sv, t, nullptr, &n);
  if (rc) return throw_lbm(env, rc);
  void* data = nullptr;
  napi_value ab;
  NAPI_OK(napi_create_arraybuffer(env, n, &data, &ab));
  rc = lbm_serve_read_state(sv, t, data, &n);
  if (rc) return throw_lbm(env, rc);
  return ab;
}

struct WaitWork {
  napi_async_work work;
  napi_deferred deferred;
  Handle* h;
  int64_t ticket;
  double timeout;
  lbm_serve_result r;
  int rc;
  std::string err;
};

void WaitExecute(napi_env, void* p) {
  WaitWork* w = static_cast<WaitWork*>(p);
  w->rc = lbm_serve_wait(serve_of(w->h), w->ticket, w->timeout, &w->r);
  if (w->rc) w->err = lbm_last_error();
}

void WaitComplete(napi_env env, napi_status, void* p) {
  WaitWork* w = static_cast<WaitWork*>(p);
  w->h->busy.fetch_sub(1);
  if (w->rc) {
    napi_value msg, err;
    std::string m = "lbm error " + std::to_string(w->rc) + ": " + w->err;
    napi_create_string_utf8(env, m.c_str(), m.size(), &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_value o, v, arr;
    napi_create_object(env, &o);
    napi_create_int64(env, w->r.ticket, &v);
    napi_set_named_property(env, o, "ticket", v);
    napi_create_int32(env, w->r.steps, &v);
    napi_set_named_property(env, o, "steps", v);
    napi_create_double(env, w->r.cells, &v);
    napi_set_named_property(env, o, "cells", v);
    napi_create_double(env, w->r.t_submit, &v);
    napi_set_named_property(env, o, "tSubmit", v);
    napi_create_double(env, w->r.t_admit, &v);
    napi_set_named_property(env, o, "tAdmit", v);
    napi_create_double(env, w->r.t_finish, &v);
    napi_set_named_property(env, o, "tFinish", v);
    napi_create_array_with_length(env, 6, &arr);
    for (uint32_t k = 0; k < 6; k++) {
      napi_create_double(env, w->r.force[k], &v);
      napi_set_element(env, arr, k, v);
    }
    napi_set_named_property(env, o, "force", arr);
    napi_resolve_deferred(env, w->deferred, o);
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

// serveWait(server, ticket, timeoutSeconds) -> Promise<result>
napi_value ServeWait(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(sv, argv[0], kServe, false);
  WaitWork* w = new WaitWork();
  w->h = sv_h;
  w->ticket = ticket_of(env, argv[1]);
  w->timeout = -1;
  if (argc > 2) napi_get_value_double(env, argv[2], &w->timeout);
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "lbm_serve_wait", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, WaitExecute, WaitComplete, w, &w->work));
  if (!queue_busy(env, w->work, sv_h)) return nullptr;
  return promise;
}

// serveReadMacro(server, ticket, Float32Array)
napi_value ServeReadMacro(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data;
  size_t len;
  if (argc < 3 || !typed(env, argv[2], napi_float32_array, &data, &len)) {
    napi_throw_type_error(env, nullptr, "serveReadMacro(server, ticket, Float32Array)");
    return nullptr;
  }
  GET(sv, argv[0], kServe, false);
  // the request's cell count, from its result (finished requests return at once)
  lbm_serve_result res;
  int rc = lbm_serve_wait(serve_of(sv_h), ticket_of(env, argv[1]), 0.0, &res);
  if (rc) return throw_lbm(env, rc);
  if (!long_enough(env, len, 4 * (size_t)res.cells, "serveReadMacro")) return nullptr;
  rc = lbm_serve_read_macro(serve_of(sv_h), ticket_of(env, argv[1]), static_cast<float*>(data));
  if (rc) return throw_lbm(env, rc);
  return argv[2];
}

// serveReadForces(server, ticket, maxRecords) -> Float64Array(6 * count)
napi_value ServeReadForces(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t maxr = 1;
  if (argc > 2) napi_get_value_int32(env, argv[2], &maxr);
  if (maxr < 0) maxr = 0;
  napi_value ab, arr;
  void* data;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)maxr * 48 + 8, &data, &ab));
  int32_t count = 0;
  GET(sv, argv[0], kServe, false);
  int rc = lbm_serve_read_forces(serve_of(sv_h), ticket_of(env, argv[1]), static_cast<double*>(data), maxr,
                                 &count);
  if (rc) return throw_lbm(env, rc);
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)count * 6, ab, 0, &arr));
  return arr;
}

napi_value ServeRelease(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  GET(sv, argv[0], kServe, false);
  int rc = lbm_serve_release(serve_of(sv_h), ticket_of(env, argv[1]));
  if (rc) return throw_lbm(env, rc);
  return nullptr;
}

napi_value ServeStats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  lbm_serve_stats st;
  GET(sv, argv[0], kServe, false);
  int rc = lbm_serve_read_stats(serve_of(sv_h), &st);
  if (rc) return throw_lbm(env, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  const struct { const char* k; double v; } f[] = {
      {"submitted", (double)st.submitted}, {"completed", (double)st.completed}, {"failed", (double)st.failed},
      {"resident", (double)st.resident}, {"queued", (double)st.queued}, {"cellUpdates", st.cell_updates},
      {"p50Latency", st.p50_latency}, {"p99Latency", st.p99_latency}, {"mlups", st.mlups},
      {"poolBytes", st.pool_bytes}, {"peakResidentBytes", st.peak_resident_bytes},
      {"peakResident", (double)st.peak_resident}, {"chunks", (double)st.chunks},
      {"deviceSeconds", st.device_seconds}, {"deviceBytes", st.device_bytes},
      {"graphInstantiations", (double)st.graph_instantiations}, {"graphUpdates", (double)st.graph_updates}};
  for (const auto& e : f) {
    napi_create_double(env, e.v, &v);
    napi_set_named_property(env, o, e.k, v);
  }
  return o;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn { const char* name; napi_callback cb; };
  const Fn fns[] = {
      {"createSolve", CreateSolve},
      {"destroy", Destroy},
      {"uploadMacro", UploadMacro},
      {"uploadF", UploadF},
      {"readF", ReadFloat<lbm_read_f, 0>},
      {"readMacro", ReadFloat<lbm_read_macro, 4>},
      {"readTotals", ReadTotals},
      {"readForces", ReadForces},
      {"readClasses", ReadClasses},
      {"step", Step},
      {"info", Info},
      {"deviceInfo", DeviceInfo},
      {"createAmr", CreateAmr},
      {"amrDestroy", AmrDestroy},
      {"amrUploadTopology", AmrTopology<true>},
      {"amrReadTopology", AmrTopology<false>},
      {"amrUploadLevelTopology", AmrUploadLevelTopology},
      {"amrUploadMacro", AmrFloat<amr_upload_macro>},
      {"amrReadMacro", AmrFloat<lbm_amr_read_macro>},
      {"amrReadCriterion", AmrReadCriterion},
      {"amrReadActive", AmrReadActive},
      {"amrReadTotals", AmrReadTotals},
      {"amrReadForces", AmrReadForces},
      {"amrManageNow", AmrManageNow},
      {"amrInfo", AmrInfo},
      {"amrStep", AmrStep},
      {"serveCreate", ServeCreate},
      {"serveDestroy", ServeDestroy},
      {"serveSubmit", ServeSubmit},
      {"serveSubmitState", ServeSubmitState},
      {"serveReadState", ServeReadState},
      {"saveState", SaveState},
      {"rigidAttach", RigidAttach},
      {"rigidRead", RigidRead},
      {"loadState", LoadState},
      {"serveWait", ServeWait},
      {"serveReadMacro", ServeReadMacro},
      {"serveReadForces", ServeReadForces},
      {"serveRelease", ServeRelease},
      {"serveStats", ServeStats},
  };
  for (const Fn& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.name, fn) != napi_ok) return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)

/*
 * Node.js N-API addon over the C ABI of librq.so (include/rq.h).
 *
 * JavaScript hosts (the molecule.js compute-pass contract, molecule_host.js)
 * drive the MI355X solver through these functions.  Every call maps to one
 * rq_* entry point; failures throw an Error carrying rq_last_error().
 * Handles are napi externals whose finaliser calls rq_destroy.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/rq.h"

#define NAPI_CALL(env, call)                                          \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error((env), NULL, "N-API call failed: " #call);     \
      return NULL;                                                    \
    }                                                                 \
  } while (0)

static napi_value throw_rq(napi_env env, rq_solver* h, int code) {
  char buf[640];
  snprintf(buf, sizeof buf, "rq error %d: %s", code, rq_last_error(h));
  napi_throw_error(env, NULL, buf);
  return NULL;
}

#define RQ_CALL(env, h, call)                 \
  do {                                        \
    int rc_ = (call);                         \
    if (rc_ != RQ_OK) return throw_rq(env, h, rc_); \
  } while (0)

static void finalize_solver(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  rq_solver** box = (rq_solver**)data;
  if (box && *box) rq_destroy(*box);
  free(box);
}

static rq_solver* get_handle(napi_env env, napi_value v) {
  void* data = NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) {
    napi_throw_type_error(env, NULL, "expected a solver handle");
    return NULL;
  }
  rq_solver* h = *(rq_solver**)data;
  if (!h) napi_throw_error(env, NULL, "solver handle already destroyed");
  return h;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return 1;
}

static int has_prop(napi_env env, napi_value obj, const char* key, napi_value* out) {
  napi_valuetype t;
  if (napi_typeof(env, obj, &t) != napi_ok || t != napi_object) return 0;
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  if (napi_get_named_property(env, obj, key, out) != napi_ok) return 0;
  napi_valuetype vt;
  napi_typeof(env, *out, &vt);
  return vt != napi_undefined && vt != napi_null;
}

static double prop_num(napi_env env, napi_value obj, const char* key, double dflt) {
  napi_value v;
  double d = dflt;
  if (has_prop(env, obj, key, &v)) napi_get_value_double(env, v, &d);
  return d;
}

static int prop_str_eq(napi_env env, napi_value obj, const char* key, const char* val) {
  napi_value v;
  if (!has_prop(env, obj, key, &v)) return 0;
  char buf[64];
  size_t n = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof buf, &n) != napi_ok) return 0;
  return strcmp(buf, val) == 0;
}

static napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

static napi_value f64array(napi_env env, const double* src, size_t n) {
  void* data = NULL;
  napi_value ab, arr;
  napi_create_arraybuffer(env, n * sizeof(double), &data, &ab);
  if (n) memcpy(data, src, n * sizeof(double));
  napi_create_typedarray(env, napi_float64_array, n, ab, 0, &arr);
  return arr;
}

/* atoms: array of [i, j, k, Z, rc] arrays or {i, j, k, Z, rc} objects */
static rq_atom* read_atoms(napi_env env, napi_value arr, uint32_t* n_out) {
  bool is_arr = false;
  napi_is_array(env, arr, &is_arr);
  if (!is_arr) {
    napi_throw_type_error(env, NULL, "atoms must be an array");
    return NULL;
  }
  uint32_t n = 0;
  napi_get_array_length(env, arr, &n);
  rq_atom* atoms = (rq_atom*)calloc(n ? n : 1, sizeof(rq_atom));
  for (uint32_t a = 0; a < n; ++a) {
    napi_value e;
    napi_get_element(env, arr, a, &e);
    bool sub = false;
    napi_is_array(env, e, &sub);
    double v[5] = {0, 0, 0, 0, 0};
    if (sub) {
      for (uint32_t c = 0; c < 5; ++c) {
        napi_value x;
        if (napi_get_element(env, e, c, &x) == napi_ok) napi_get_value_double(env, x, &v[c]);
      }
    } else {
      v[0] = prop_num(env, e, "i", 0);
      v[1] = prop_num(env, e, "j", 0);
      v[2] = prop_num(env, e, "k", 0);
      v[3] = prop_num(env, e, "Z", 0);
      v[4] = prop_num(env, e, "rc", 0);
    }
    atoms[a].i = (float)v[0];
    atoms[a].j = (float)v[1];
    atoms[a].k = (float)v[2];
    atoms[a].Z = (float)v[3];
    atoms[a].rc = (float)v[4];
  }
  *n_out = n;
  return atoms;
}

/* create(params, atoms) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_params p;
  rq_default_params(&p);
  napi_value o = argv[0];
  p.NN = (int32_t)prop_num(env, o, "NN", p.NN);
  p.screen = prop_num(env, o, "screen", p.screen);
  p.dv = prop_num(env, o, "dv", 0.0);
  if (prop_str_eq(env, o, "poisson", "multigrid")) p.poisson_mode = RQ_POISSON_MULTIGRID;
  if (prop_str_eq(env, o, "poisson", "twogrid")) p.poisson_mode = RQ_POISSON_TWOGRID;
  p.jacobi_sweeps = (int32_t)prop_num(env, o, "jacobiSweeps", p.jacobi_sweeps);
  p.smooth_a = (float)prop_num(env, o, "smoothA", p.smooth_a);
  p.smooth_c = (float)prop_num(env, o, "smoothC", p.smooth_c);
  p.mg_levels = (int32_t)prop_num(env, o, "mgLevels", p.mg_levels);
  p.mg_nu1 = (int32_t)prop_num(env, o, "mgNu1", p.mg_nu1);
  p.mg_nu2 = (int32_t)prop_num(env, o, "mgNu2", p.mg_nu2);
  p.twogrid_interval = (int32_t)prop_num(env, o, "twogridInterval", p.twogrid_interval);
  p.twogrid_coarse_sweeps = (int32_t)prop_num(env, o, "twogridCoarseSweeps", p.twogrid_coarse_sweeps);
  if (prop_str_eq(env, o, "boundary", "levelset")) p.boundary_mode = RQ_BOUNDARY_LEVELSET;
  if (prop_str_eq(env, o, "boundary", "off")) p.boundary_mode = RQ_BOUNDARY_OFF;
  if (prop_str_eq(env, o, "init", "s2")) p.init_mode = RQ_INIT_S2;
  if (prop_str_eq(env, o, "energy", "s2")) p.energy_mode = RQ_ENERGY_S2;
  if (prop_str_eq(env, o, "pother", "frame")) p.pother_mode = RQ_POTHER_FRAME;
  if (prop_str_eq(env, o, "pother", "directSweeps")) p.pother_mode = RQ_POTHER_DIRECT_SWEEPS;
  p.sic_interval = (int32_t)prop_num(env, o, "sicInterval", 0);
  p.rc_dirichlet = (int32_t)prop_num(env, o, "rcDirichlet", 0);
  p.ls_dt_w = (float)prop_num(env, o, "dtW", p.ls_dt_w);
  p.ls_curv_reg = (float)prop_num(env, o, "curvReg", p.ls_curv_reg);
  p.flip_ratio = (float)prop_num(env, o, "flipRatio", p.flip_ratio);
  uint32_t n = 0;
  rq_atom* atoms = read_atoms(env, argv[1], &n);
  if (!atoms) return NULL;
  rq_solver* h = NULL;
  int rc = rq_create(&p, atoms, (int32_t)n, &h);
  free(atoms);
  if (rc != RQ_OK) return throw_rq(env, NULL, rc);
  rq_solver** box = (rq_solver**)malloc(sizeof(rq_solver*));
  *box = h;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, box, finalize_solver, NULL, &ext));
  return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  void* data = NULL;
  if (napi_get_value_external(env, argv[0], &data) == napi_ok && data) {
    rq_solver** box = (rq_solver**)data;
    if (*box) rq_destroy(*box);
    *box = NULL;
  }
  return NULL;
}

#define SIMPLE_PASS(name, fn)                                         \
  static napi_value name(napi_env env, napi_callback_info info) {     \
    napi_value argv[1];                                               \
    if (!get_args(env, info, 1, argv)) return NULL;                   \
    rq_solver* h = get_handle(env, argv[0]);                          \
    if (!h) return NULL;                                              \
    RQ_CALL(env, h, fn(h));                                           \
    return NULL;                                                      \
  }

SIMPLE_PASS(js_init_trial, rq_init_trial)
SIMPLE_PASS(js_update, rq_update)
SIMPLE_PASS(js_normalize, rq_normalize)
SIMPLE_PASS(js_boundary, rq_boundary)
SIMPLE_PASS(js_recompute_k, rq_recompute_k)
SIMPLE_PASS(js_twogrid, rq_twogrid_correct)
SIMPLE_PASS(js_sync, rq_sync)

static napi_value js_initialize(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double b = 100;
  napi_get_value_double(env, argv[1], &b);
  RQ_CALL(env, h, rq_initialize(h, (int32_t)b));
  return NULL;
}

static napi_value js_poisson(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double n = 1;
  napi_get_value_double(env, argv[1], &n);
  RQ_CALL(env, h, rq_poisson(h, (int32_t)n));
  return NULL;
}

/* step(h, firstStep, nSteps, {poissonInterval, normInterval, boundaryInterval}) */
static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double first = 0, n = 1;
  napi_get_value_double(env, argv[1], &first);
  napi_get_value_double(env, argv[2], &n);
  rq_schedule s;
  rq_default_schedule(&s);
  s.poisson_interval = (int32_t)prop_num(env, argv[3], "poissonInterval", s.poisson_interval);
  s.norm_interval = (int32_t)prop_num(env, argv[3], "normInterval", s.norm_interval);
  s.boundary_interval = (int32_t)prop_num(env, argv[3], "boundaryInterval", s.boundary_interval);
  RQ_CALL(env, h, rq_step(h, (int64_t)first, (int32_t)n, &s));
  return NULL;
}

static napi_value js_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t NN, S, na;
  double hg, dv;
  RQ_CALL(env, h, rq_get_info(h, &NN, &S, &na, &hg, &dv));
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "NN", num(env, NN));
  napi_set_named_property(env, o, "S", num(env, S));
  napi_set_named_property(env, o, "nAtoms", num(env, na));
  napi_set_named_property(env, o, "h", num(env, hg));
  napi_set_named_property(env, o, "dv", num(env, dv));
  return o;
}

/* energy(h) -> {T, VeK, Vee, dipole: [x, y, z], perDomain: Float64Array(3n)} */
static napi_value js_energy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t na = 0;
  rq_get_info(h, NULL, NULL, &na, NULL, NULL);
  double out[6];
  double* per = (double*)malloc(sizeof(double) * 3 * (size_t)na);
  int rc = rq_energy(h, out, per);
  if (rc != RQ_OK) {
    free(per);
    return throw_rq(env, h, rc);
  }
  napi_value o, dip;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "T", num(env, out[0]));
  napi_set_named_property(env, o, "VeK", num(env, out[1]));
  napi_set_named_property(env, o, "Vee", num(env, out[2]));
  napi_create_array_with_length(env, 3, &dip);
  for (uint32_t c = 0; c < 3; ++c) napi_set_element(env, dip, c, num(env, out[3 + c]));
  napi_set_named_property(env, o, "dipole", dip);
  napi_set_named_property(env, o, "perDomain", f64array(env, per, 3 * (size_t)na));
  free(per);
  return o;
}

static napi_value js_forces_hf(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double first = 0, count = 0;
  napi_get_value_double(env, argv[1], &first);
  napi_get_value_double(env, argv[2], &count);
  double* out = (double*)calloc(3 * (size_t)count + 1, sizeof(double));
  int rc = rq_forces_hf(h, (int32_t)first, (int32_t)count, out);
  if (rc != RQ_OK) {
    free(out);
    return throw_rq(env, h, rc);
  }
  napi_value arr = f64array(env, out, 3 * (size_t)count);
  free(out);
  return arr;
}

static napi_value js_forces_gradp(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  double radius = 0;
  napi_get_value_double(env, argv[1], &radius);
  int32_t na = 0;
  rq_get_info(h, NULL, NULL, &na, NULL, NULL);
  double* out = (double*)calloc(3 * (size_t)na + 1, sizeof(double));
  int rc = rq_forces_gradp(h, (int32_t)radius, out);
  if (rc != RQ_OK) {
    free(out);
    return throw_rq(env, h, rc);
  }
  napi_value arr = f64array(env, out, 3 * (size_t)na);
  free(out);
  return arr;
}

static int field_id(napi_env env, napi_value v) {
  char buf[32];
  size_t n = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof buf, &n) != napi_ok) return -1;
  static const char* names[] = {"U", "U2", "K", "P", "P2", "rho", "label", "bestU", "W", "Pother"};
  for (int f = 0; f < 10; ++f)
    if (strcmp(buf, names[f]) == 0) return f;
  return -1;
}

/* readField(h, name) -> Float32Array (Int32Array for "label"), length S^3 */
static napi_value js_read_field(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int f = field_id(env, argv[1]);
  if (f < 0) {
    napi_throw_type_error(env, NULL, "unknown field name");
    return NULL;
  }
  int32_t S = 0;
  rq_get_info(h, NULL, &S, NULL, NULL, NULL);
  size_t n = (size_t)S * S * S;
  void* data = NULL;
  napi_value ab, arr;
  NAPI_CALL(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
  RQ_CALL(env, h, rq_read_field(h, f, data, n * 4));
  NAPI_CALL(env, napi_create_typedarray(env, f == RQ_FIELD_LABEL ? napi_int32_array : napi_float32_array, n, ab, 0,
                                        &arr));
  return arr;
}

/* displaySlice(h, k) -> Float32Array of 3 S^2 + 2 S floats: the molecule.js
 * extract buffer (density, label + 1000 Z, boundary planes; K line; W line) */
static napi_value js_display_slice(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int32_t k = 0;
  NAPI_CALL(env, napi_get_value_int32(env, argv[1], &k));
  int32_t S = 0;
  rq_get_info(h, NULL, &S, NULL, NULL, NULL);
  size_t n = 3 * (size_t)S * S + 2 * (size_t)S;
  void* data = NULL;
  napi_value ab, arr;
  NAPI_CALL(env, napi_create_arraybuffer(env, n * 4, &data, &ab));
  RQ_CALL(env, h, rq_display_slice(h, k, (float*)data, n));
  NAPI_CALL(env, napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
  return arr;
}

/* potherFrame(h, frame): end-of-frame potential (molecule.js:3009-3076) */
static napi_value js_pother_frame(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  int64_t frame = 0;
  NAPI_CALL(env, napi_get_value_int64(env, argv[1], &frame));
  RQ_CALL(env, h, rq_pother_frame(h, frame));
  return NULL;
}

/* writeField(h, name, array): a Float32Array (an Int32Array for "label") of
 * S^3 elements.  The array type is checked before the handle, so a wrong
 * type is a TypeError whatever the handle. */
static napi_value js_write_field(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  int f = field_id(env, argv[1]);
  if (f < 0) {
    napi_throw_type_error(env, NULL, "unknown field name");
    return NULL;
  }
  napi_typedarray_type t = napi_int8_array;
  size_t len = 0, off = 0;
  void* data = NULL;
  napi_value ab;
  bool is_ta = false;
  NAPI_CALL(env, napi_is_typedarray(env, argv[2], &is_ta));
  if (is_ta) NAPI_CALL(env, napi_get_typedarray_info(env, argv[2], &t, &len, &data, &ab, &off));
  /* the label map is int32 on the host, every other field float32: the
   * byte count passed on is that of the array's own elements */
  const napi_typedarray_type want = f == RQ_FIELD_LABEL ? napi_int32_array : napi_float32_array;
  if (!is_ta || t != want) {
    napi_throw_type_error(env, NULL,
                          f == RQ_FIELD_LABEL ? "writeField: label needs an Int32Array"
                                              : "writeField: field needs a Float32Array");
    return NULL;
  }
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  RQ_CALL(env, h, rq_write_field(h, f, data, len * 4));
  return NULL;
}

static napi_value js_set_atoms(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  uint32_t n = 0;
  rq_atom* atoms = read_atoms(env, argv[1], &n);
  if (!atoms) return NULL;
  int rc = rq_set_atoms(h, atoms, (int32_t)n);
  free(atoms);
  if (rc != RQ_OK) return throw_rq(env, h, rc);
  return NULL;
}

/* setLockGroups(h, Int32Array(nAtoms)) (molecule.js:487-500) */
static napi_value js_set_lock_groups(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  napi_typedarray_type t = napi_int8_array;
  size_t len = 0, off = 0;
  void* data = NULL;
  napi_value ab;
  bool is_ta = false;
  NAPI_CALL(env, napi_is_typedarray(env, argv[1], &is_ta));
  if (is_ta) NAPI_CALL(env, napi_get_typedarray_info(env, argv[1], &t, &len, &data, &ab, &off));
  if (!is_ta || t != napi_int32_array) {
    napi_throw_type_error(env, NULL, "setLockGroups: expected an Int32Array");
    return NULL;
  }
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  RQ_CALL(env, h, rq_set_lock_groups(h, (const int32_t*)data, (int32_t)len));
  return NULL;
}

/* setAtomsS2(h, [{Z_nuc, initZeff, initRcut, splitType, splitIdx, splitAxis: [x, y, z],
 * splitRot}]): the molecule.js per-atom fields (molecule.js:1875-1903); a
 * missing key keeps the library default (0: Z_nuc = Z, initZeff = Z,
 * initRcut = 1e6; sphere; axis [0, 0, 1]).  splitType 'sphere' | 'hemi' |
 * 'third' or its code. */
static napi_value js_set_atoms_s2(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  bool is_arr = false;
  napi_is_array(env, argv[1], &is_arr);
  if (!is_arr) {
    napi_throw_type_error(env, NULL, "setAtomsS2: expected an array of per-atom objects");
    return NULL;
  }
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  rq_atom_s2* ext = (rq_atom_s2*)calloc(n ? n : 1, sizeof(rq_atom_s2));
  for (uint32_t a = 0; a < n; ++a) {
    napi_value e, v;
    napi_get_element(env, argv[1], a, &e);
    ext[a].Z_nuc = (float)prop_num(env, e, "Z_nuc", 0);
    ext[a].init_zeff = (float)prop_num(env, e, "initZeff", 0);
    ext[a].init_rcut = (float)prop_num(env, e, "initRcut", 0);
    if (prop_str_eq(env, e, "splitType", "hemi")) ext[a].split_type = RQ_SPLIT_HEMI;
    else if (prop_str_eq(env, e, "splitType", "third")) ext[a].split_type = RQ_SPLIT_THIRD;
    else if (has_prop(env, e, "splitType", &v)) {
      napi_valuetype vt;
      napi_typeof(env, v, &vt);
      if (vt == napi_number) napi_get_value_int32(env, v, &ext[a].split_type);
    }
    ext[a].split_idx = (int32_t)prop_num(env, e, "splitIdx", 0);
    ext[a].split_axis[0] = 0.0f;
    ext[a].split_axis[1] = 0.0f;
    ext[a].split_axis[2] = 1.0f;
    if (has_prop(env, e, "splitAxis", &v)) {
      for (uint32_t c = 0; c < 3; ++c) {
        napi_value x;
        double d = ext[a].split_axis[c];
        if (napi_get_element(env, v, c, &x) == napi_ok) napi_get_value_double(env, x, &d);
        ext[a].split_axis[c] = (float)d;
      }
    }
    ext[a].split_rot = (float)prop_num(env, e, "splitRot", 0);
  }
  const int rc = rq_set_atoms_s2(h, ext, (int32_t)n);
  free(ext);
  if (rc != RQ_OK) return throw_rq(env, h, rc);
  return NULL;
}

/* adjointGradient(h, {nSteps, bootstrap, checkpoint, poissonInterval, normInterval, start: 'trial' | 'saved'})
 *   -> {energy: [T, VeK, Vee], dEdR: Float64Array(3n)} */
static napi_value js_adjoint(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  rq_adjoint_spec sp;
  memset(&sp, 0, sizeof sp);
  sp.n_steps = (int64_t)prop_num(env, argv[1], "nSteps", 100);
  sp.bootstrap = (int32_t)prop_num(env, argv[1], "bootstrap", 100);
  sp.checkpoint_interval = (int32_t)prop_num(env, argv[1], "checkpoint", 64);
  sp.schedule.poisson_interval = (int32_t)prop_num(env, argv[1], "poissonInterval", 2);
  sp.schedule.norm_interval = (int32_t)prop_num(env, argv[1], "normInterval", 20);
  sp.schedule.boundary_interval = 1 << 30;
  if (prop_str_eq(env, argv[1], "start", "saved")) sp.start = RQ_ADJ_START_SAVED;
  int32_t na = 0;
  rq_get_info(h, NULL, NULL, &na, NULL, NULL);
  double e3[3];
  double* g = (double*)calloc(3 * (size_t)na + 1, sizeof(double));
  int rc = rq_adjoint_gradient(h, &sp, e3, g);
  if (rc != RQ_OK) {
    free(g);
    return throw_rq(env, h, rc);
  }
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "energy", f64array(env, e3, 3));
  napi_set_named_property(env, o, "dEdR", f64array(env, g, 3 * (size_t)na));
  free(g);
  return o;
}

/* adjointSaveStart(h): the current U and P become the start of
 * adjointGradient(h, {start: 'saved'}) */
static napi_value js_adjoint_save_start(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  rq_solver* h = get_handle(env, argv[0]);
  if (!h) return NULL;
  RQ_CALL(env, h, rq_adjoint_save_start(h));
  return NULL;
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  return num(env, rq_abi_version());
}

static napi_value js_dv_for_electrons(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  double n = 0;
  napi_get_value_double(env, argv[0], &n);
  return num(env, rq_dv_for_electrons((int32_t)n));
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"abiVersion", 0, js_abi_version, 0, 0, 0, napi_enumerable, 0},
      {"dvForElectrons", 0, js_dv_for_electrons, 0, 0, 0, napi_enumerable, 0},
      {"create", 0, js_create, 0, 0, 0, napi_enumerable, 0},
      {"destroy", 0, js_destroy, 0, 0, 0, napi_enumerable, 0},
      {"info", 0, js_info, 0, 0, 0, napi_enumerable, 0},
      {"initTrial", 0, js_init_trial, 0, 0, 0, napi_enumerable, 0},
      {"initialize", 0, js_initialize, 0, 0, 0, napi_enumerable, 0},
      {"poisson", 0, js_poisson, 0, 0, 0, napi_enumerable, 0},
      {"twogridCorrect", 0, js_twogrid, 0, 0, 0, napi_enumerable, 0},
      {"update", 0, js_update, 0, 0, 0, napi_enumerable, 0},
      {"normalize", 0, js_normalize, 0, 0, 0, napi_enumerable, 0},
      {"boundary", 0, js_boundary, 0, 0, 0, napi_enumerable, 0},
      {"recomputeK", 0, js_recompute_k, 0, 0, 0, napi_enumerable, 0},
      {"step", 0, js_step, 0, 0, 0, napi_enumerable, 0},
      {"sync", 0, js_sync, 0, 0, 0, napi_enumerable, 0},
      {"energy", 0, js_energy, 0, 0, 0, napi_enumerable, 0},
      {"forcesHF", 0, js_forces_hf, 0, 0, 0, napi_enumerable, 0},
      {"forcesGradP", 0, js_forces_gradp, 0, 0, 0, napi_enumerable, 0},
      {"readField", 0, js_read_field, 0, 0, 0, napi_enumerable, 0},
      {"displaySlice", 0, js_display_slice, 0, 0, 0, napi_enumerable, 0},
      {"potherFrame", 0, js_pother_frame, 0, 0, 0, napi_enumerable, 0},
      {"writeField", 0, js_write_field, 0, 0, 0, napi_enumerable, 0},
      {"setAtoms", 0, js_set_atoms, 0, 0, 0, napi_enumerable, 0},
      {"setLockGroups", 0, js_set_lock_groups, 0, 0, 0, napi_enumerable, 0},
      {"setAtomsS2", 0, js_set_atoms_s2, 0, 0, 0, napi_enumerable, 0},
      {"adjointGradient", 0, js_adjoint, 0, 0, 0, napi_enumerable, 0},
      {"adjointSaveStart", 0, js_adjoint_save_start, 0, 0, 0, napi_enumerable, 0},
  };
  NAPI_CALL(env, napi_define_properties(env, exports, sizeof props / sizeof props[0], props));
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

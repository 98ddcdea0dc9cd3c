/*
 * avbd_node.c — N-API 8 addon over the C ABI of libavbd_mi355x.so (include/avbd.h).
 *
 * The binding the reference's TypeScript host loads in place of its WebGPU solver class
 * (GpuSolver3D, src/avbd3d/gpu/solver.ts:504; World constructs it at src/lib/world.ts:847 and
 * drives it from World.flush/step, world.ts:1736-1995). Each export is one C ABI call on typed
 * arrays; gpu-solver-3d.js wraps them in a class with the reference's method names. A negative
 * status becomes a thrown Error carrying avbd_last_error (code "StorageCapacityError" for
 * AVBD_ERR_CAPACITY, as the reference throws StorageCapacityError).
 *
 * Build (done by the package's build step):
 *   cc -O2 -fPIC -shared -I/usr/include/node -Iinclude avbd_node.c -L../avbd_amd -lavbd_mi355x \
 *      -Wl,-rpath,'$ORIGIN/../avbd_amd' -o avbd_node.node
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "avbd.h"

#define CHECK(call)                                        \
  do {                                                     \
    if ((call) != napi_ok) {                               \
      napi_throw_error(env, NULL, "N-API call failed");    \
      return NULL;                                         \
    }                                                      \
  } while (0)
#define ARGS(n)                                                      \
  size_t argc = (n);                                                 \
  napi_value argv[(n) > 0 ? (n) : 1];                                \
  CHECK(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));       \
  if (argc < (size_t)(n)) {                                          \
    napi_throw_type_error(env, NULL, "too few arguments");           \
    return NULL;                                                     \
  }

static napi_value status(napi_env env, avbd_solver *s, int rc) {
  if (rc < 0) {
    const char *msg = s ? avbd_last_error(s) : avbd_create_error();
    const char *code = rc == AVBD_ERR_CAPACITY ? "StorageCapacityError" : rc == AVBD_ERR_DESTROYED ? "Destroyed" : NULL;
    napi_throw_error(env, code, msg && msg[0] ? msg : "avbd error");
    return NULL;
  }
  napi_value r;
  napi_create_int32(env, rc, &r);
  return r;
}
static avbd_solver *handle(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "not a solver handle");
    return NULL;
  }
  return (avbd_solver *)p;
}
/* The data of a typed array (or NULL for null/undefined) and its element count. */
static void *data(napi_env env, napi_value v, size_t *count) {
  napi_valuetype t;
  *count = 0;
  if (napi_typeof(env, v, &t) != napi_ok || t == napi_null || t == napi_undefined) return NULL;
  bool is;
  if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
    napi_typedarray_type tt;
    void *p;
    napi_value ab;
    size_t off;
    if (napi_get_typedarray_info(env, v, &tt, count, &p, &ab, &off) != napi_ok) return NULL;
    return p;
  }
  if (napi_is_arraybuffer(env, v, &is) == napi_ok && is) {
    void *p;
    if (napi_get_arraybuffer_info(env, v, &p, count) != napi_ok) return NULL;
    return p;
  }
  return NULL;
}
static uint32_t u32(napi_env env, napi_value v) {
  uint32_t x = 0;
  napi_get_value_uint32(env, v, &x);
  return x;
}
static int32_t i32(napi_env env, napi_value v) {
  int32_t x = 0;
  napi_get_value_int32(env, v, &x);
  return x;
}
static double f64(napi_env env, napi_value v) {
  double x = 0;
  napi_get_value_double(env, v, &x);
  return x;
}
static bool flag(napi_env env, napi_value v) {
  bool b = false;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_boolean) napi_get_value_bool(env, v, &b);
  return b;
}
static napi_value js_uint(napi_env env, uint32_t x) {
  napi_value r;
  napi_create_uint32(env, x, &r);
  return r;
}

/* create(params: ArrayBuffer(avbd_params), bodies: Float32Array, joints: Float32Array | null,
 *        info: Int32Array | null, capacity: Uint32Array(6) | null) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  ARGS(5);
  size_t np, nb, nj, ni, nc;
  const avbd_params *p = (const avbd_params *)data(env, argv[0], &np);
  const float *bodies = (const float *)data(env, argv[1], &nb);
  const float *joints = (const float *)data(env, argv[2], &nj);
  const int32_t *jinfo = (const int32_t *)data(env, argv[3], &ni);
  const avbd_capacity *cap = (const avbd_capacity *)data(env, argv[4], &nc);
  if (!p || np < sizeof(avbd_params)) {
    napi_throw_type_error(env, NULL, "params: an ArrayBuffer holding avbd_params");
    return NULL;
  }
  if (nb % AVBD_BODY_FLOATS || nj % AVBD_JOINT_FLOATS || ni != 4 * (nj / AVBD_JOINT_FLOATS) || (cap && nc != 6)) {
    napi_throw_type_error(env, NULL, "records: 40 floats per body and per joint, 4 info words per joint, 6 capacities");
    return NULL;
  }
  avbd_solver *s = NULL;
  int rc = avbd_create(p, bodies, (uint32_t)(nb / AVBD_BODY_FLOATS), joints, jinfo, (uint32_t)(nj / AVBD_JOINT_FLOATS), cap, &s);
  if (rc < 0) return status(env, NULL, rc);
  napi_value ext;
  CHECK(napi_create_external(env, s, NULL, NULL, &ext));
  return ext;
}
static napi_value js_destroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  avbd_destroy(s);
  return NULL;
}
static napi_value js_step(napi_env env, napi_callback_info info) {  /* (h, reuseManifolds) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  avbd_step_opts o = {flag(env, argv[1]) ? 1 : 0};
  return status(env, s, avbd_step(s, &o));
}
static napi_value js_steps(napi_env env, napi_callback_info info) {  /* (h, steps, substeps) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  return status(env, s, avbd_steps(s, u32(env, argv[1]), u32(env, argv[2])));
}
static napi_value js_set_params(napi_env env, napi_callback_info info) {  /* (h, ArrayBuffer) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const avbd_params *p = (const avbd_params *)data(env, argv[1], &n);
  if (!s) return NULL;
  if (!p || n < sizeof(avbd_params)) {
    napi_throw_type_error(env, NULL, "params: an ArrayBuffer holding avbd_params");
    return NULL;
  }
  return status(env, s, avbd_set_params(s, p));
}
static napi_value js_get_params(napi_env env, napi_callback_info info) {  /* (h, ArrayBuffer out) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  avbd_params *p = (avbd_params *)data(env, argv[1], &n);
  if (!s) return NULL;
  if (!p || n < sizeof(avbd_params)) {
    napi_throw_type_error(env, NULL, "params: an ArrayBuffer of avbd_params' size");
    return NULL;
  }
  return status(env, s, avbd_get_params(s, p));
}
/* (h, Float32Array out, first): records [first, first + out.length / 40) */
static napi_value js_read_bodies(napi_env env, napi_callback_info info) {
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  float *out = (float *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_read_bodies(s, out, u32(env, argv[2]), (uint32_t)(n / AVBD_BODY_FLOATS)));
}
static napi_value js_write_bodies(napi_env env, napi_callback_info info) {  /* rewriteBodies: (h, Float32Array, first) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const float *in = (const float *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_write_bodies(s, in, u32(env, argv[2]), (uint32_t)(n / AVBD_BODY_FLOATS)));
}
static napi_value js_read_joints(napi_env env, napi_callback_info info) {  /* (h, Float32Array out, first) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  float *out = (float *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_read_joints(s, out, u32(env, argv[2]), (uint32_t)(n / AVBD_JOINT_FLOATS)));
}
static napi_value js_read_joint_info(napi_env env, napi_callback_info info) {  /* (h, Int32Array out) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  int32_t *out = (int32_t *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_read_joint_info(s, out, 0, (uint32_t)(n / 4)));
}
/* (h, Uint32Array(10) out): pairs, contacts, manifolds, overflow, clashes, colors, colorCap, step, clashing[2] */
static napi_value js_read_counters(napi_env env, napi_callback_info info) {
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  avbd_counters *out = (avbd_counters *)data(env, argv[1], &n);
  if (!s) return NULL;
  if (!out || n < sizeof(avbd_counters) / 4) {
    napi_throw_type_error(env, NULL, "counters: a Uint32Array(10)");
    return NULL;
  }
  return status(env, s, avbd_read_counters(s, out));
}
/* (h, Uint32Array manifolds, Uint32Array contacts) -> manifolds written */
static napi_value js_read_contacts(napi_env env, napi_callback_info info) {
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t nm, nc;
  uint32_t *man = (uint32_t *)data(env, argv[1], &nm);
  uint32_t *con = (uint32_t *)data(env, argv[2], &nc);
  if (!s) return NULL;
  return status(env, s, avbd_read_contacts(s, man, (uint32_t)(nm / AVBD_MANIFOLD_WORDS), con, (uint32_t)(nc / AVBD_CONTACT_WORDS)));
}
static napi_value js_read_pairs(napi_env env, napi_callback_info info) {  /* (h, Uint32Array out) -> pairs written */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  uint32_t *out = (uint32_t *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_read_pairs(s, out, (uint32_t)(n / 2)));
}
static napi_value js_get_sizes(napi_env env, napi_callback_info info) {  /* (h) -> [bodies, joints] */
  ARGS(1);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  uint32_t b = 0, j = 0;
  if (avbd_get_sizes(s, &b, &j) < 0) return status(env, s, AVBD_ERR_INVALID);
  napi_value arr;
  CHECK(napi_create_array_with_length(env, 2, &arr));
  napi_set_element(env, arr, 0, js_uint(env, b));
  napi_set_element(env, arr, 1, js_uint(env, j));
  return arr;
}
static napi_value js_add_bodies(napi_env env, napi_callback_info info) {  /* (h, Float32Array) -> first | -1 */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const float *in = (const float *)data(env, argv[1], &n);
  if (!s) return NULL;
  int32_t first = -1;
  int rc = avbd_add_bodies(s, in, (uint32_t)(n / AVBD_BODY_FLOATS), &first);
  if (rc < 0) return status(env, s, rc);
  napi_value r;
  napi_create_int32(env, first, &r);
  return r;
}
/* (h, index, fields, Float32Array(40), fixed, marker) */
static napi_value js_patch_body(napi_env env, napi_callback_info info) {
  ARGS(6);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const float *rec = (const float *)data(env, argv[3], &n);
  if (!s) return NULL;
  if (!rec || n < AVBD_BODY_FLOATS) {
    napi_throw_type_error(env, NULL, "patchBody: a 40-float record");
    return NULL;
  }
  return status(env, s, avbd_patch_body(s, u32(env, argv[1]), u32(env, argv[2]), rec, flag(env, argv[4]), flag(env, argv[5])));
}
/* (h, Uint32Array indices, Float32Array positions, Float32Array size(3), friction, Float32Array | null rotations) */
static napi_value js_rewrite_fixed(napi_env env, napi_callback_info info) {
  ARGS(6);
  avbd_solver *s = handle(env, argv[0]);
  size_t ni, np, ns, nr;
  const uint32_t *idx = (const uint32_t *)data(env, argv[1], &ni);
  const float *pos = (const float *)data(env, argv[2], &np);
  const float *size = (const float *)data(env, argv[3], &ns);
  const float *rot = (const float *)data(env, argv[5], &nr);
  if (!s) return NULL;
  if (np < 3 * ni || ns < 3 || (rot && nr < 4 * ni)) {
    napi_throw_type_error(env, NULL, "rewriteFixed: 3 floats per position, 4 per rotation, a 3-float size");
    return NULL;
  }
  return status(env, s, avbd_rewrite_fixed(s, idx, (uint32_t)ni, pos, size, (float)f64(env, argv[4]), rot));
}
/* (h, ArrayBuffer of avbd_joint_spec, count, fracture, linear, Uint32Array slots out) */
static napi_value js_append_joints(napi_env env, napi_callback_info info) {
  ARGS(6);
  avbd_solver *s = handle(env, argv[0]);
  size_t nb, ns;
  const avbd_joint_spec *specs = (const avbd_joint_spec *)data(env, argv[1], &nb);
  uint32_t *slots = (uint32_t *)data(env, argv[5], &ns);
  uint32_t count = u32(env, argv[2]);
  if (!s) return NULL;
  if (nb < count * sizeof(avbd_joint_spec) || ns < count) {
    napi_throw_type_error(env, NULL, "appendJoints: specs and slots sized for count");
    return NULL;
  }
  return status(env, s, avbd_append_joints(s, specs, count, (float)f64(env, argv[3]), flag(env, argv[4]), slots));
}
/* (h, ArrayBuffer of avbd_spring_spec, count, Uint32Array slots out) */
static napi_value js_append_springs(napi_env env, napi_callback_info info) {
  ARGS(4);
  avbd_solver *s = handle(env, argv[0]);
  size_t nb, ns;
  const avbd_spring_spec *springs = (const avbd_spring_spec *)data(env, argv[1], &nb);
  uint32_t *slots = (uint32_t *)data(env, argv[3], &ns);
  uint32_t count = u32(env, argv[2]);
  if (!s) return NULL;
  if (nb < count * sizeof(avbd_spring_spec) || ns < count) {
    napi_throw_type_error(env, NULL, "appendSprings: specs and slots sized for count");
    return NULL;
  }
  return status(env, s, avbd_append_springs(s, springs, count, slots));
}
/* (h, a, b, Float32Array rA, Float32Array rB, stiffnessLin, stiffnessAng) -> slot */
static napi_value js_append_joint(napi_env env, napi_callback_info info) {
  ARGS(7);
  avbd_solver *s = handle(env, argv[0]);
  size_t na, nb;
  const float *rA = (const float *)data(env, argv[3], &na);
  const float *rB = (const float *)data(env, argv[4], &nb);
  if (!s) return NULL;
  if (na < 3 || nb < 3) {
    napi_throw_type_error(env, NULL, "appendJoint: 3-float anchors");
    return NULL;
  }
  uint32_t slot = 0;
  int rc = avbd_append_joint(s, i32(env, argv[1]), i32(env, argv[2]), rA, rB, (float)f64(env, argv[5]), (float)f64(env, argv[6]), &slot);
  if (rc < 0) return status(env, s, rc);
  return js_uint(env, slot);
}
/* (h, b, Float32Array anchorWorld, Float32Array anchorBody) -> slot */
static napi_value js_append_world_joint(napi_env env, napi_callback_info info) {
  ARGS(4);
  avbd_solver *s = handle(env, argv[0]);
  size_t na, nb;
  const float *aw = (const float *)data(env, argv[2], &na);
  const float *ab = (const float *)data(env, argv[3], &nb);
  if (!s) return NULL;
  if (na < 3 || nb < 3) {
    napi_throw_type_error(env, NULL, "appendWorldJoint: anchors need three finite f32 numbers");
    return NULL;
  }
  uint32_t slot = 0;
  int rc = avbd_append_world_joint(s, u32(env, argv[1]), aw, ab, &slot);
  if (rc < 0) return status(env, s, rc);
  return js_uint(env, slot);
}
static napi_value js_release_joints(napi_env env, napi_callback_info info) {  /* (h, Uint32Array slots) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const uint32_t *slots = (const uint32_t *)data(env, argv[1], &n);
  if (!s) return NULL;
  return status(env, s, avbd_release_joints(s, slots, (uint32_t)n));
}
static napi_value js_set_world_anchor(napi_env env, napi_callback_info info) {  /* (h, slot, Float32Array p) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t n;
  const float *p = (const float *)data(env, argv[2], &n);
  if (!s) return NULL;
  if (n < 3) {
    napi_throw_type_error(env, NULL, "setWorldAnchor: a 3-float point");
    return NULL;
  }
  return status(env, s, avbd_set_world_anchor(s, u32(env, argv[1]), p));
}
static napi_value js_disable_constraint(napi_env env, napi_callback_info info) {  /* (h, slot) */
  ARGS(2);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  return status(env, s, avbd_disable_constraint(s, u32(env, argv[1])));
}
/* (h, Uint32Array indices, Float32Array scales, Float32Array | null linear, Float32Array | null angular) */
static napi_value js_set_body_aux(napi_env env, napi_callback_info info) {
  ARGS(5);
  avbd_solver *s = handle(env, argv[0]);
  size_t ni, ns, nl, na;
  const uint32_t *idx = (const uint32_t *)data(env, argv[1], &ni);
  const float *scales = (const float *)data(env, argv[2], &ns);
  const float *lin = (const float *)data(env, argv[3], &nl);
  const float *ang = (const float *)data(env, argv[4], &na);
  if (!s) return NULL;
  if (ns < ni || (lin && nl < ni) || (ang && na < ni)) {
    napi_throw_error(env, NULL, "GpuSolver3D.setBodyAux: one scale per index");
    return NULL;
  }
  return status(env, s, avbd_set_body_aux(s, idx, (uint32_t)ni, scales, lin, ang));
}
static napi_value js_set_lifecycle(napi_env env, napi_callback_info info) {  /* (h, index, bits) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  return status(env, s, avbd_set_lifecycle(s, u32(env, argv[1]), u32(env, argv[2])));
}
static napi_value js_set_filters(napi_env env, napi_callback_info info) {  /* (h, Uint32Array groups, Uint32Array collidesWith) */
  ARGS(3);
  avbd_solver *s = handle(env, argv[0]);
  size_t ng, nw;
  const uint32_t *g = (const uint32_t *)data(env, argv[1], &ng);
  const uint32_t *w = (const uint32_t *)data(env, argv[2], &nw);
  if (!s) return NULL;
  if (ng != nw) {
    napi_throw_type_error(env, NULL, "setFilters: one collidesWith per group");
    return NULL;
  }
  return status(env, s, avbd_set_filters(s, g, w, (uint32_t)ng));
}
static napi_value js_adapt(napi_env env, napi_callback_info info) {  /* (h): reads the counters, adapts */
  ARGS(1);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  avbd_counters c;
  int rc = avbd_read_counters(s, &c);
  if (rc < 0) return status(env, s, rc);
  return status(env, s, avbd_adapt(s, &c));
}
static napi_value js_sync(napi_env env, napi_callback_info info) {
  ARGS(1);
  avbd_solver *s = handle(env, argv[0]);
  if (!s) return NULL;
  return status(env, s, avbd_sync(s));
}
static napi_value js_sizes(napi_env env, napi_callback_info info) {  /* struct sizes for the JS packers */
  napi_value o;
  CHECK(napi_create_object(env, &o));
  napi_set_named_property(env, o, "params", js_uint(env, (uint32_t)sizeof(avbd_params)));
  napi_set_named_property(env, o, "counters", js_uint(env, (uint32_t)sizeof(avbd_counters)));
  napi_set_named_property(env, o, "jointSpec", js_uint(env, (uint32_t)sizeof(avbd_joint_spec)));
  napi_set_named_property(env, o, "springSpec", js_uint(env, (uint32_t)sizeof(avbd_spring_spec)));
  return o;
}

static napi_value init(napi_env env, napi_value exports) {
  const struct {
    const char *name;
    napi_callback fn;
  } fns[] = {
      {"create", js_create},
      {"destroy", js_destroy},
      {"step", js_step},
      {"steps", js_steps},
      {"setParams", js_set_params},
      {"getParams", js_get_params},
      {"readBodies", js_read_bodies},
      {"rewriteBodies", js_write_bodies},
      {"readJoints", js_read_joints},
      {"readJointInfo", js_read_joint_info},
      {"readCounters", js_read_counters},
      {"readContacts", js_read_contacts},
      {"readPairs", js_read_pairs},
      {"getSizes", js_get_sizes},
      {"addBodies", js_add_bodies},
      {"patchBody", js_patch_body},
      {"rewriteFixed", js_rewrite_fixed},
      {"appendJoints", js_append_joints},
      {"appendSprings", js_append_springs},
      {"appendJoint", js_append_joint},
      {"appendWorldJoint", js_append_world_joint},
      {"releaseJoints", js_release_joints},
      {"setWorldAnchor", js_set_world_anchor},
      {"disableConstraint", js_disable_constraint},
      {"setBodyAux", js_set_body_aux},
      {"setLifecycle", js_set_lifecycle},
      {"setFilters", js_set_filters},
      {"adapt", js_adapt},
      {"sync", js_sync},
      {"structSizes", js_sizes},
  };
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; k++) {
    napi_value f;
    CHECK(napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].fn, NULL, &f));
    CHECK(napi_set_named_property(env, exports, fns[k].name, f));
  }
  return exports;
}
NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

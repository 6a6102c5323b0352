/*
 * ketsv_napi.c: the Node-API addon through which ket's TypeScript host reaches
 * libketsv (include/ketsv.h).  It replaces the seam of src/circuit.ts:143-207
 * (simulatePure / applyOps over the apply* functions of src/statevector.ts):
 * the host packs its FlatOps into the ksv_op tape (16 int32 per op: kind, nq,
 * q[12], mat as a little-endian int64) and a parallel Float64Array of matrix
 * entries, and reads results back into typed arrays.  No HIP code lives here.
 *
 * Errors keep ket's classes: KSV_ERR_RANGE -> RangeError, KSV_ERR_TYPE (and a
 * malformed argument) -> TypeError, anything else -> Error, each with the
 * library's message (ksv_last_error).
 *
 * Exports:
 *   create(n, f32?) -> state          free(state)
 *   initBasis(state, index)           apply(state, ops: Int32Array, mats: Float64Array)
 *   amplitudes(state) -> Float64Array (re, im interleaved)
 *   amplitude(state, index) -> [re, im]
 *   probs(state) -> [BigUint64Array index, Float64Array p]   (p > 1e-14, ascending index)
 *   sample(state, r: Float64Array) -> BigUint64Array          (run()'s draw)
 *   expect(state, x: BigUint64Array, z: BigUint64Array, c: Float64Array, mode) -> number
 *   plan(n, ops, mats, f32?) -> {passes, deviceOps, blocks}    (schedule only, no device)
 *   xorshift(seed, count) -> Float64Array                      (makePrng's stream)
 */
#include <node_api.h>
#include <stdlib.h>
#include <string.h>

#include "ketsv.h"

typedef struct {
    ksv_state* s;
    int n;
} handle;

static napi_value throw_ksv(napi_env env, int rc) {
    char msg[512];
    ksv_last_error(msg, sizeof msg);
    if (rc == KSV_ERR_RANGE) napi_throw_range_error(env, NULL, msg);
    else if (rc == KSV_ERR_TYPE) napi_throw_type_error(env, NULL, msg);
    else napi_throw_error(env, NULL, msg);
    return NULL;
}

static napi_value type_error(napi_env env, const char* msg) {
    napi_throw_type_error(env, NULL, msg);
    return NULL;
}

static int args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
    for (size_t i = argc; i < want; i++) napi_get_undefined(env, &argv[i]);
    return 1;
}

static void finalize(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    handle* h = (handle*)data;
    if (h->s) ksv_destroy(h->s);
    free(h);
}

static handle* handle_of(napi_env env, napi_value v) {
    void* p = NULL;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        type_error(env, "state must be a handle from create()");
        return NULL;
    }
    handle* h = (handle*)p;
    if (!h->s) {
        napi_throw_error(env, NULL, "state has been freed");
        return NULL;
    }
    return h;
}

static ksv_state* state_of(napi_env env, napi_value v) {
    handle* h = handle_of(env, v);
    return h ? h->s : NULL;
}

static int typed(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* len, const char* what) {
    bool is = false;
    napi_typedarray_type t;
    napi_value ab;
    size_t off;
    if (napi_is_typedarray(env, v, &is) != napi_ok || !is ||
        napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok || t != want) {
        type_error(env, what);
        return 0;
    }
    return 1;
}

/* an index given as a Number or a BigInt */
static int index_of(napi_env env, napi_value v, uint64_t* out) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_bigint) {
        bool lossless = false;
        napi_get_value_bigint_uint64(env, v, out, &lossless);
        if (!lossless) return 0;
        return 1;
    }
    if (t == napi_number) {
        double d;
        napi_get_value_double(env, v, &d);
        if (!(d >= 0) || d != (double)(uint64_t)d) return 0;
        *out = (uint64_t)d;
        return 1;
    }
    return 0;
}

static napi_value create(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args(env, info, 2, argv)) return NULL;
    int32_t n = 0;
    bool f32 = false;
    if (napi_get_value_int32(env, argv[0], &n) != napi_ok) return type_error(env, "qubit count must be a number");
    napi_valuetype t;
    napi_typeof(env, argv[1], &t);
    if (t == napi_boolean) napi_get_value_bool(env, argv[1], &f32);
    ksv_state* s = NULL;
    int rc = ksv_create(n, f32 ? KSV_F32 : KSV_F64, &s);
    if (rc) return throw_ksv(env, rc);
    handle* h = (handle*)malloc(sizeof *h);
    if (!h) {
        ksv_destroy(s);
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    h->s = s;
    h->n = n;
    napi_value ext;
    napi_create_external(env, h, finalize, NULL, &ext);
    return ext;
}

static napi_value free_state(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args(env, info, 1, argv)) return NULL;
    void* p = NULL;
    if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p) return type_error(env, "state must be a handle from create()");
    handle* h = (handle*)p;
    if (h->s) ksv_destroy(h->s);
    h->s = NULL;
    return NULL;
}

static napi_value init_basis(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args(env, info, 2, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    uint64_t idx;
    if (!index_of(env, argv[1], &idx)) return type_error(env, "index must be a non-negative integer or BigInt");
    int rc = ksv_init_basis(s, idx);
    if (rc) return throw_ksv(env, rc);
    return NULL;
}

static int tape_of(napi_env env, napi_value ops_v, napi_value mats_v, const ksv_op** ops, size_t* nops,
                   const double** mats, size_t* nmats) {
    void *o, *m;
    size_t no, nm;
    if (!typed(env, ops_v, napi_int32_array, &o, &no, "ops must be an Int32Array (16 int32 per op)")) return 0;
    if (!typed(env, mats_v, napi_float64_array, &m, &nm, "mats must be a Float64Array")) return 0;
    if (no % 16) {
        type_error(env, "ops must hold 16 int32 per op");
        return 0;
    }
    *ops = (const ksv_op*)o;
    *nops = no / 16;
    *mats = (const double*)m;
    *nmats = nm;
    return 1;
}

static napi_value apply(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!args(env, info, 3, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    const ksv_op* ops;
    const double* mats;
    size_t nops, nmats;
    if (!tape_of(env, argv[1], argv[2], &ops, &nops, &mats, &nmats)) return NULL;
    int rc = ksv_apply(s, ops, nops, mats, nmats);
    if (rc) return throw_ksv(env, rc);
    return NULL;
}

static napi_value plan(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!args(env, info, 4, argv)) return NULL;
    int32_t n = 0;
    bool f32 = false;
    if (napi_get_value_int32(env, argv[0], &n) != napi_ok) return type_error(env, "qubit count must be a number");
    napi_valuetype t;
    napi_typeof(env, argv[3], &t);
    if (t == napi_boolean) napi_get_value_bool(env, argv[3], &f32);
    const ksv_op* ops;
    const double* mats;
    size_t nops, nmats;
    if (!tape_of(env, argv[1], argv[2], &ops, &nops, &mats, &nmats)) return NULL;
    ksv_program_summary sm;
    int rc = ksv_plan(n, f32 ? KSV_F32 : KSV_F64, ops, nops, mats, nmats, 0, 0, 0, 0, 0, &sm);
    if (rc) return throw_ksv(env, rc);
    napi_value out, v;
    napi_create_object(env, &out);
    napi_create_double(env, (double)sm.passes, &v);
    napi_set_named_property(env, out, "passes", v);
    napi_create_double(env, (double)sm.device_ops, &v);
    napi_set_named_property(env, out, "deviceOps", v);
    napi_create_double(env, (double)sm.blocks, &v);
    napi_set_named_property(env, out, "blocks", v);
    return out;
}

static napi_value new_array(napi_env env, napi_typedarray_type t, size_t count, void** data) {
    napi_value ab, arr;
    napi_create_arraybuffer(env, count * 8 > 0 ? count * 8 : 8, data, &ab);
    napi_create_typedarray(env, t, count, ab, 0, &arr);
    return arr;
}

static napi_value amplitudes(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args(env, info, 1, argv)) return NULL;
    handle* h = handle_of(env, argv[0]);
    if (!h) return NULL;
    const size_t count = (size_t)1 << h->n;
    void* d;
    napi_value arr = new_array(env, napi_float64_array, 2 * count, &d);
    int rc = ksv_get_amplitudes(h->s, 0, count, (double*)d);
    if (rc) return throw_ksv(env, rc);
    return arr;
}

static napi_value amplitude(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args(env, info, 2, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    uint64_t idx;
    if (!index_of(env, argv[1], &idx)) return type_error(env, "index must be a non-negative integer or BigInt");
    double a[2];
    int rc = ksv_amplitude(s, idx, a);
    if (rc) return throw_ksv(env, rc);
    napi_value out, v;
    napi_create_array_with_length(env, 2, &out);
    napi_create_double(env, a[0], &v);
    napi_set_element(env, out, 0, v);
    napi_create_double(env, a[1], &v);
    napi_set_element(env, out, 1, v);
    return out;
}

static napi_value probs(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!args(env, info, 1, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    size_t count = 0;
    int rc = ksv_probs_filtered(s, 1e-14, NULL, NULL, 0, &count);
    if (rc) return throw_ksv(env, rc);
    void *ib, *pb;
    napi_value ia = new_array(env, napi_biguint64_array, count, &ib);
    napi_value pa = new_array(env, napi_float64_array, count, &pb);
    rc = ksv_probs_filtered(s, 1e-14, (uint64_t*)ib, (double*)pb, count, &count);
    if (rc) return throw_ksv(env, rc);
    napi_value out;
    napi_create_array_with_length(env, 2, &out);
    napi_set_element(env, out, 0, ia);
    napi_set_element(env, out, 1, pa);
    return out;
}

static napi_value sample(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args(env, info, 2, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    void* r;
    size_t shots;
    if (!typed(env, argv[1], napi_float64_array, &r, &shots, "draws must be a Float64Array")) return NULL;
    void* o;
    napi_value out = new_array(env, napi_biguint64_array, shots, &o);
    int rc = ksv_sample(s, (const double*)r, shots, (uint64_t*)o);
    if (rc) return throw_ksv(env, rc);
    return out;
}

static napi_value expect(napi_env env, napi_callback_info info) {
    napi_value argv[5];
    if (!args(env, info, 5, argv)) return NULL;
    ksv_state* s = state_of(env, argv[0]);
    if (!s) return NULL;
    void *x, *z, *c;
    size_t nx, nz, nc;
    if (!typed(env, argv[1], napi_biguint64_array, &x, &nx, "x masks must be a BigUint64Array")) return NULL;
    if (!typed(env, argv[2], napi_biguint64_array, &z, &nz, "z masks must be a BigUint64Array")) return NULL;
    if (!typed(env, argv[3], napi_float64_array, &c, &nc, "coefficients must be a Float64Array")) return NULL;
    if (nx != nz || nx != nc) return type_error(env, "x, z and coefficients differ in length");
    int32_t mode = 0;
    napi_get_value_int32(env, argv[4], &mode);
    double* terms = (double*)malloc(8 * (nc ? nc : 1));
    if (!terms) {
        napi_throw_error(env, NULL, "out of host memory");
        return NULL;
    }
    double total = 0;
    int rc = ksv_expect_pauli(s, (const uint64_t*)x, (const uint64_t*)z, (const double*)c, nc, mode, terms, &total);
    free(terms);
    if (rc) return throw_ksv(env, rc);
    napi_value out;
    napi_create_double(env, total, &out);
    return out;
}

static napi_value xorshift(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!args(env, info, 2, argv)) return NULL;
    double seed_d;
    uint32_t count;
    if (napi_get_value_double(env, argv[0], &seed_d) != napi_ok) return type_error(env, "seed must be a number");
    if (napi_get_value_uint32(env, argv[1], &count) != napi_ok) return type_error(env, "count must be a number");
    /* (seed >>> 0) || 1, src/circuit.ts:814 */
    uint32_t seed = 0;
    napi_get_value_uint32(env, argv[0], &seed);
    if (!seed) seed = 1;
    void* d;
    napi_value arr = new_array(env, napi_float64_array, count, &d);
    int rc = ksv_xorshift32(&seed, count, (double*)d);
    if (rc) return throw_ksv(env, rc);
    return arr;
}

NAPI_MODULE_INIT() {
    napi_property_descriptor d[] = {
        {"create", 0, create, 0, 0, 0, napi_default, 0},
        {"free", 0, free_state, 0, 0, 0, napi_default, 0},
        {"initBasis", 0, init_basis, 0, 0, 0, napi_default, 0},
        {"apply", 0, apply, 0, 0, 0, napi_default, 0},
        {"plan", 0, plan, 0, 0, 0, napi_default, 0},
        {"amplitudes", 0, amplitudes, 0, 0, 0, napi_default, 0},
        {"amplitude", 0, amplitude, 0, 0, 0, napi_default, 0},
        {"probs", 0, probs, 0, 0, 0, napi_default, 0},
        {"sample", 0, sample, 0, 0, 0, napi_default, 0},
        {"expect", 0, expect, 0, 0, 0, napi_default, 0},
        {"xorshift", 0, xorshift, 0, 0, 0, napi_default, 0},
    };
    napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
    return exports;
}

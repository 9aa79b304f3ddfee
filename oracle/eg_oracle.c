/*
 * eg_oracle.c -- CPU restatement of the reference's single-key group-by path.
 *
 * TEST INFRASTRUCTURE ONLY (see eg_oracle.h): the checker for the HIP
 * implementation and the timed CPU baseline.  The product never links it.
 *
 * Reference files cited are relative to src/cuda/ of the reference repo.
 * The reference has no CPU group-by: this file restates the semantics of its
 * GPU code paths (types, identities, output typing, request rewriting) in
 * plain C.  Two independent group-by implementations live here:
 *   orc_groupby           -- brute force: stable sort of row ids by key, then
 *                            one sequential reduction per run (the oracle);
 *   orc_groupby_parallel  -- OpenMP hash-partition + per-partition hash table
 *                            (the timed CPU baseline), cross-checked against
 *                            the brute force in the tests.
 */
#include "eg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[256];

static int fail(const char* msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

const char* orc_last_error(void) { return g_err; }

/* ======================================================================== */
/* Counter-based generator                                                   */
/* ======================================================================== */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Multipliers and Weyl increments are the
 * published ones. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_row(uint64_t row, uint32_t stream_id, uint64_t seed, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)row, (uint32_t)(row >> 32), stream_id, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* Stream ids of the generator (one independent Philox stream per column). */
#define ORC_STREAM_PERM 0x7065726du /* constants of the row permutation */
#define ORC_STREAM_ZIPF 1u

/* Keyed bijection on [0, 2^k): xor, odd multiply mod 2^k and xor-shift-right
 * are each bijective on k-bit words, so their composition is too. */
typedef struct perm_ctx {
    uint64_t n, mask, a1, a2, x0, x1;
    int s;
} perm_ctx;

static void perm_init(perm_ctx* p, uint64_t n, uint64_t seed)
{
    int k = 1;
    while (k < 64 && (((uint64_t)1) << k) < n) ++k;
    uint32_t o0[4], o1[4];
    uint32_t ctr0[4] = {0u, 0u, ORC_STREAM_PERM, 0u};
    uint32_t ctr1[4] = {1u, 0u, ORC_STREAM_PERM, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr0, key, o0);
    orc_philox4x32_10(ctr1, key, o1);
    p->n = n;
    p->mask = (k == 64) ? ~(uint64_t)0 : ((((uint64_t)1) << k) - 1);
    p->a1 = (((uint64_t)o0[1] << 32) | o0[0]) | 1u;
    p->a2 = (((uint64_t)o0[3] << 32) | o0[2]) | 1u;
    p->x0 = ((uint64_t)o1[1] << 32) | o1[0];
    p->x1 = ((uint64_t)o1[3] << 32) | o1[2];
    p->s = (k + 1) / 2;
}

static uint64_t perm_round(const perm_ctx* p, uint64_t x)
{
    x ^= p->x0 & p->mask;
    x = (x * p->a1) & p->mask;
    x ^= x >> p->s;
    x = (x * p->a2) & p->mask;
    x ^= x >> p->s;
    x ^= p->x1 & p->mask;
    x = (x * p->a1) & p->mask;
    x ^= x >> p->s;
    return x;
}

static uint64_t perm_apply(const perm_ctx* p, uint64_t i)
{
    /* cycle walking: i < n lies on a cycle of the bijection that contains a
     * point < n (itself), so the loop ends. */
    uint64_t y = perm_round(p, i);
    while (y >= p->n) y = perm_round(p, y);
    return y;
}

uint64_t orc_perm(uint64_t i, uint64_t n, uint64_t seed)
{
    perm_ctx p;
    perm_init(&p, n, seed);
    return perm_apply(&p, i);
}

static void store_int(void* out, int32_t type, int64_t i, int64_t v)
{
    switch (type) {
    case ORC_INT8: ((int8_t*)out)[i] = (int8_t)v; break;
    case ORC_INT16: ((int16_t*)out)[i] = (int16_t)v; break;
    case ORC_INT32: ((int32_t*)out)[i] = (int32_t)v; break;
    case ORC_INT64: ((int64_t*)out)[i] = v; break;
    case ORC_UINT8: ((uint8_t*)out)[i] = (uint8_t)v; break;
    case ORC_UINT16: ((uint16_t*)out)[i] = (uint16_t)v; break;
    case ORC_UINT32: ((uint32_t*)out)[i] = (uint32_t)v; break;
    case ORC_UINT64: ((uint64_t*)out)[i] = (uint64_t)v; break;
    case ORC_FLOAT32: ((float*)out)[i] = (float)v; break;
    case ORC_FLOAT64: ((double*)out)[i] = (double)v; break;
    default: break;
    }
}

/* Uniform exact-count keys: the reference writes key = i % G and shuffles
 * the rows (benchmarks/groupby/groupby_util.cuh:99-109); here row i holds
 * perm(i) mod D for a seeded bijection perm of [0, n), so every key in
 * [0, D) occurs floor(n/D) or ceil(n/D) times.  Rows [begin, begin+count). */
void orc_gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count,
                          int64_t n, int64_t distinct, uint64_t seed)
{
    perm_ctx p;
    perm_init(&p, (uint64_t)n, seed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i) {
        uint64_t y = perm_apply(&p, (uint64_t)(begin + i));
        store_int(out, type, i, (int64_t)(y % (uint64_t)distinct));
    }
}

/* Zipf CDF over ranks [0, D): w_r = 1/(r+1)^s, cdf = inclusive prefix / total,
 * last entry forced to 1.0 (data_gen/generator.cuh:391-484). */
void orc_zipf_cdf(double* cdf, int64_t distinct, double s)
{
    double acc = 0.0;
    for (int64_t r = 0; r < distinct; ++r) {
        acc += 1.0 / pow((double)(r + 1), s);
        cdf[r] = acc;
    }
    for (int64_t r = 0; r < distinct; ++r) cdf[r] = cdf[r] / acc;
    cdf[distinct - 1] = 1.0;
}

static double u53(const uint32_t o[4])
{
    uint64_t hi = o[0] >> 5, lo = o[1] >> 6; /* 27 + 26 bits */
    return (double)((hi << 26) | lo) * (1.0 / 9007199254740992.0);
}

/* i.i.d. Zipf draws: smallest rank r with cdf[r] > u
 * (data_gen/generator.cuh:360-383 inverse-CDF binary search). */
void orc_gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count,
                       const double* cdf, int64_t distinct, uint64_t seed)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i) {
        uint32_t o[4];
        philox_row((uint64_t)(begin + i), ORC_STREAM_ZIPF, seed, o);
        double u = u53(o);
        int64_t lo = 0, hi = distinct - 1;
        while (lo < hi) {
            int64_t mid = lo + ((hi - lo) >> 1);
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        store_int(out, type, i, lo);
    }
}

void orc_gen_values(void* out, int32_t type, int32_t dist, int64_t begin,
                    int64_t count, uint64_t seed, uint32_t stream_id)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i) {
        int64_t row = begin + i;
        uint32_t o[4];
        switch (dist) {
        case ORC_VAL_UNIFORM_I64:
            philox_row((uint64_t)row, stream_id, seed, o);
            store_int(out, type, i, (int64_t)o[0] - 2147483648LL);
            break;
        case ORC_VAL_UNIFORM_F32:
            philox_row((uint64_t)row, stream_id, seed, o);
            if (type == ORC_FLOAT32)
                ((float*)out)[i] = (float)(o[0] >> 8) * (1.0f / 16777216.0f);
            else if (type == ORC_FLOAT64)
                ((double*)out)[i] = (double)(o[0] >> 8) * (1.0 / 16777216.0);
            break;
        case ORC_VAL_UNIFORM_I32_1000:
            philox_row((uint64_t)row, stream_id, seed, o);
            store_int(out, type, i, (int64_t)(((uint64_t)o[0] * 1000u) >> 32));
            break;
        case ORC_VAL_ALL_ONES:
            store_int(out, type, i, 1);
            break;
        case ORC_VAL_SEQ:
            store_int(out, type, i, row);
            break;
        default:
            break;
        }
    }
}

/* ======================================================================== */
/* Hashing (operators/groupby/groupby_partition.cu:64-104)                    */
/* ======================================================================== */

/* MurmurHash3 finalizers (Austin Appleby, public domain). */
uint32_t orc_fmix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

uint64_t orc_fmix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static int type_size(int32_t t)
{
    switch (t) {
    case ORC_INT8: case ORC_UINT8: case ORC_BOOL8: return 1;
    case ORC_INT16: case ORC_UINT16: return 2;
    case ORC_INT32: case ORC_UINT32: case ORC_FLOAT32: return 4;
    case ORC_INT64: case ORC_UINT64: case ORC_FLOAT64: return 8;
    default: return 0;
    }
}

static int is_signed_int(int32_t t)
{
    return t == ORC_INT8 || t == ORC_INT16 || t == ORC_INT32 || t == ORC_INT64;
}

static int is_unsigned_int(int32_t t)
{
    return t == ORC_UINT8 || t == ORC_UINT16 || t == ORC_UINT32 || t == ORC_UINT64;
}

static int is_float(int32_t t) { return t == ORC_FLOAT32 || t == ORC_FLOAT64; }

/* raw bits of element i, zero-extended */
static uint64_t raw_bits(const void* p, int32_t t, int64_t i)
{
    switch (type_size(t)) {
    case 1: return ((const uint8_t*)p)[i];
    case 2: return ((const uint16_t*)p)[i];
    case 4: return ((const uint32_t*)p)[i];
    default: return ((const uint64_t*)p)[i];
    }
}

static int64_t as_i64(const void* p, int32_t t, int64_t i)
{
    switch (t) {
    case ORC_INT8: return ((const int8_t*)p)[i];
    case ORC_INT16: return ((const int16_t*)p)[i];
    case ORC_INT32: return ((const int32_t*)p)[i];
    case ORC_INT64: return ((const int64_t*)p)[i];
    default: return (int64_t)raw_bits(p, t, i);
    }
}

static double as_f64(const void* p, int32_t t, int64_t i)
{
    if (t == ORC_FLOAT32) return (double)((const float*)p)[i];
    if (t == ORC_FLOAT64) return ((const double*)p)[i];
    if (is_signed_int(t)) return (double)as_i64(p, t, i);
    return (double)raw_bits(p, t, i);
}

/* bit_mix (groupby_partition.cu:83-90): fmix32 of the low 32 bits for keys of
 * at most 4 bytes, fmix64 otherwise. */
static uint64_t bit_mix(const void* keys, int32_t t, int64_t i)
{
    if (type_size(t) <= 4) return orc_fmix32((uint32_t)as_i64(keys, t, i));
    return orc_fmix64((uint64_t)as_i64(keys, t, i));
}

/* compute_partition_ids_kernel (groupby_partition.cu:94-104) */
void orc_partition_ids(const void* keys, int32_t type, int64_t n, uint32_t mask,
                       int32_t* pids)
{
    for (int64_t i = 0; i < n; ++i) pids[i] = (int32_t)(bit_mix(keys, type, i) & mask);
}

/* ======================================================================== */
/* Output typing (operators/groupby/aggregate.cu:20-53, agg_func.cuh:18-36)    */
/* ======================================================================== */

int32_t orc_agg_output_type(int32_t agg_kind, int32_t type)
{
    if (agg_kind == ORC_AGG_COUNT_ALL) return ORC_INT64;
    if (agg_kind == ORC_AGG_AVG) return ORC_FLOAT64;
    if (agg_kind == ORC_AGG_SUM) {
        switch (type) {
        case ORC_INT8: case ORC_INT16: case ORC_INT32: case ORC_INT64: return ORC_INT64;
        case ORC_UINT8: case ORC_UINT16: case ORC_UINT32: case ORC_UINT64: return ORC_UINT64;
        case ORC_FLOAT32: return ORC_FLOAT32;
        case ORC_FLOAT64: return ORC_FLOAT64;
        default: return ORC_EMPTY;
        }
    }
    if (agg_kind == ORC_AGG_MAX || agg_kind == ORC_AGG_MIN || agg_kind == ORC_AGG_ANY)
        return type_size(type) ? type : ORC_EMPTY;
    return ORC_EMPTY;
}

/* ======================================================================== */
/* Aggregation states                                                         */
/* ======================================================================== */

/* Internal ops after the request rewrite of groupby.cu:121-393: AVG becomes a
 * SUM shared with any SUM of the same column plus the shared count;
 * COUNT_ALL and hint.agg_count read the shared count (the reference's SUM of a
 * materialised ones column, groupby.cu:202-231). */
enum { OP_COUNT = 0, OP_SUM = 1, OP_MAX = 2, OP_MIN = 3, OP_ANY = 4 };

typedef struct op_t {
    int kind;
    int64_t col; /* value column, -1 for OP_COUNT */
    int32_t vtype;
} op_t;

typedef union state_t {
    uint64_t u;   /* integer SUM (wrapping), COUNT, unsigned MIN/MAX */
    int64_t i;    /* signed MIN/MAX */
    double d;     /* float SUM (fp64 accumulation), float MIN/MAX */
} state_t;

typedef struct plan_t {
    op_t ops[ORC_MAX_COLS];
    int nops;
    int count_op;               /* index of OP_COUNT, -1 if none */
    /* output columns after the key: source op and final kind */
    int out_op[ORC_MAX_COLS];
    int out_kind[ORC_MAX_COLS]; /* original agg kind, or ORC_AGG_COUNT_ALL for agg_count */
    int32_t out_type[ORC_MAX_COLS];
    int nout;
} plan_t;

static int find_op(const plan_t* p, int kind, int64_t col)
{
    for (int k = 0; k < p->nops; ++k)
        if (p->ops[k].kind == kind && p->ops[k].col == col) return k;
    return -1;
}

static int add_op(plan_t* p, int kind, int64_t col, int32_t vtype)
{
    int k = find_op(p, kind, col);
    if (k >= 0 && kind != OP_ANY) return k;
    if (p->nops >= ORC_MAX_COLS) return -1;
    p->ops[p->nops].kind = kind;
    p->ops[p->nops].col = col;
    p->ops[p->nops].vtype = vtype;
    return p->nops++;
}

static int value_type_ok(int32_t t)
{
    return is_signed_int(t) || is_unsigned_int(t) || is_float(t);
}

static int build_plan(const orc_col* vals, int32_t nvals, const orc_req* reqs,
                      int32_t nreqs, int32_t agg_count, plan_t* p)
{
    memset(p, 0, sizeof *p);
    p->count_op = -1;
    if (nreqs <= 0) return fail("No aggregation function is specified");
    if (nreqs + 2 > ORC_MAX_COLS) return fail("too many aggregation requests");
    int needs_count = agg_count ? 1 : 0;
    for (int r = 0; r < nreqs; ++r) {
        int kind = reqs[r].agg_kind;
        if (kind == ORC_AGG_AVG || kind == ORC_AGG_COUNT_ALL) needs_count = 1;
    }
    if (needs_count) p->count_op = add_op(p, OP_COUNT, -1, ORC_INT64);
    for (int r = 0; r < nreqs; ++r) {
        int kind = reqs[r].agg_kind;
        int64_t c = reqs[r].column_index;
        if (kind == ORC_AGG_COUNT_ALL) {
            p->out_op[p->nout] = p->count_op;
            p->out_kind[p->nout] = ORC_AGG_COUNT_ALL;
            p->out_type[p->nout] = ORC_INT64;
            p->nout++;
            continue;
        }
        if (c < 0 || c >= nvals) return fail("aggregation column index out of range");
        int32_t vt = vals[c].type;
        if (!value_type_ok(vt)) return fail("unsupported value column type");
        int op;
        switch (kind) {
        case ORC_AGG_SUM: case ORC_AGG_AVG: op = add_op(p, OP_SUM, c, vt); break;
        case ORC_AGG_MAX: op = add_op(p, OP_MAX, c, vt); break;
        case ORC_AGG_MIN: op = add_op(p, OP_MIN, c, vt); break;
        case ORC_AGG_ANY: op = add_op(p, OP_ANY, c, vt); break;
        default: return fail("Unsupported aggregation function");
        }
        if (op < 0) return fail("too many aggregation ops");
        p->out_op[p->nout] = op;
        p->out_kind[p->nout] = kind;
        p->out_type[p->nout] = orc_agg_output_type(kind, vt);
        p->nout++;
    }
    if (agg_count) {
        p->out_op[p->nout] = p->count_op;
        p->out_kind[p->nout] = ORC_AGG_COUNT_ALL;
        p->out_type[p->nout] = ORC_INT64;
        p->nout++;
    }
    return 0;
}

static void state_init(const op_t* o, state_t* s, int first_row_pending)
{
    (void)first_row_pending;
    s->u = 0;
    if (o->kind == OP_SUM && is_float(o->vtype)) s->d = 0.0;
    else if (o->kind == OP_MAX) {
        if (is_float(o->vtype)) s->d = -INFINITY;
        else if (is_signed_int(o->vtype)) s->i = INT64_MIN;
        else s->u = 0;
    } else if (o->kind == OP_MIN) {
        if (is_float(o->vtype)) s->d = INFINITY;
        else if (is_signed_int(o->vtype)) s->i = INT64_MAX;
        else s->u = UINT64_MAX;
    }
}

/* One row into one state.  Semantics of agg_func.cuh:37-207:
 *   SUM over integers wraps mod 2^64 (atomics/atomics.cuh:42-43);
 *   SUM over float32 accumulates in fp64 here (the FP64 oracle);
 *   MAX/MIN exact; ANY keeps the first row's value. */
static void state_update(const op_t* o, state_t* s, const orc_col* vals, int64_t row, int first)
{
    const void* v = o->col >= 0 ? vals[o->col].data : NULL;
    switch (o->kind) {
    case OP_COUNT: s->u += 1; break;
    case OP_SUM:
        if (is_float(o->vtype)) s->d += as_f64(v, o->vtype, row);
        else s->u += (uint64_t)as_i64(v, o->vtype, row);
        break;
    case OP_MAX:
        if (is_float(o->vtype)) { double x = as_f64(v, o->vtype, row); if (x > s->d) s->d = x; }
        else if (is_signed_int(o->vtype)) { int64_t x = as_i64(v, o->vtype, row); if (x > s->i) s->i = x; }
        else { uint64_t x = raw_bits(v, o->vtype, row); if (x > s->u) s->u = x; }
        break;
    case OP_MIN:
        if (is_float(o->vtype)) { double x = as_f64(v, o->vtype, row); if (x < s->d) s->d = x; }
        else if (is_signed_int(o->vtype)) { int64_t x = as_i64(v, o->vtype, row); if (x < s->i) s->i = x; }
        else { uint64_t x = raw_bits(v, o->vtype, row); if (x < s->u) s->u = x; }
        break;
    case OP_ANY:
        if (first) s->u = raw_bits(v, o->vtype, row);
        break;
    }
}

/* Write output cell g of column j (groupby.cu:326-388 assembly; AVG =
 * double(sum)/double(count), groupby.cu:29-51). */
static void emit(const plan_t* p, int j, const state_t* st, void* out, int64_t g)
{
    const op_t* o = &p->ops[p->out_op[j]];
    const state_t* s = &st[p->out_op[j]];
    int kind = p->out_kind[j];
    int32_t ot = p->out_type[j];
    if (kind == ORC_AGG_COUNT_ALL) { ((int64_t*)out)[g] = (int64_t)s->u; return; }
    if (kind == ORC_AGG_AVG) {
        double cnt = (double)st[p->count_op].u;
        double sum;
        if (is_float(o->vtype)) sum = s->d;
        else if (is_signed_int(o->vtype)) sum = (double)(int64_t)s->u;
        else sum = (double)s->u;
        ((double*)out)[g] = sum / cnt;
        return;
    }
    if (ot == ORC_FLOAT32) {
        ((float*)out)[g] = (float)s->d;
        return;
    }
    if (ot == ORC_FLOAT64) {
        ((double*)out)[g] = s->d;
        return;
    }
    if (kind == ORC_AGG_ANY) {
        memcpy((char*)out + g * type_size(ot), &s->u, (size_t)type_size(ot));
        return;
    }
    /* integer outputs: SUM -> 64-bit, MAX/MIN -> value type */
    if (is_signed_int(ot)) store_int(out, ot, g, (kind == ORC_AGG_SUM) ? (int64_t)s->u : s->i);
    else store_int(out, ot, g, (int64_t)s->u);
}

static int key_type_ok(int32_t t)
{
    return is_signed_int(t) || is_unsigned_int(t) || t == ORC_BOOL8 || t == ORC_FLOAT32 || t == ORC_FLOAT64;
}

/* Float keys compare as numbers with operator== (memory/row_utils.cuh:16-21:
 * -0.0 equals +0.0), except that every NaN is one key, after +inf.  The
 * value (a float32 widened exactly) is ranked by its IEEE-754 bits: the bit
 * order of non-negative doubles is their numeric order, and reversed for
 * negative ones. */
static uint64_t float_key_ord(double v)
{
    if (isnan(v)) return UINT64_MAX;
    if (v == 0.0) v = 0.0;
    uint64_t b;
    memcpy(&b, &v, sizeof b);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

/* key comparison value: signed keys as int64, unsigned as uint64 mapped into
 * the same order by flipping the sign bit; floats by float_key_ord */
static uint64_t key_ord(const void* k, int32_t t, int64_t i)
{
    if (t == ORC_FLOAT32) return float_key_ord((double)((const float*)k)[i]);
    if (t == ORC_FLOAT64) return float_key_ord(((const double*)k)[i]);
    if (is_signed_int(t)) return (uint64_t)as_i64(k, t, i) ^ 0x8000000000000000ULL;
    return raw_bits(k, t, i);
}

static int alloc_result(orc_result* res, const plan_t* p, int32_t key_type, int64_t rows)
{
    memset(res, 0, sizeof *res);
    res->num_rows = rows;
    res->num_cols = 1 + p->nout;
    res->types[0] = key_type;
    size_t n = (size_t)(rows > 0 ? rows : 1);
    res->cols[0] = calloc(n, (size_t)type_size(key_type));
    if (!res->cols[0]) return fail("out of memory");
    for (int j = 0; j < p->nout; ++j) {
        res->types[1 + j] = p->out_type[j];
        res->cols[1 + j] = calloc(n, (size_t)type_size(p->out_type[j]));
        if (!res->cols[1 + j]) return fail("out of memory");
    }
    return 0;
}

void orc_result_free(orc_result* res)
{
    for (int j = 0; j < res->num_cols && j < ORC_MAX_COLS; ++j) {
        free(res->cols[j]);
        res->cols[j] = NULL;
    }
    res->num_cols = 0;
    res->num_rows = 0;
}

/* stable merge sort of row ids by key_ord */
typedef struct { uint64_t k; int64_t r; } kr_t;

static void merge_sort_kr(kr_t* a, kr_t* tmp, int64_t n)
{
    for (int64_t w = 1; w < n; w <<= 1) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) tmp[o++] = (a[j].k < a[i].k) ? a[j++] : a[i++];
            while (i < mid) tmp[o++] = a[i++];
            while (j < hi) tmp[o++] = a[j++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(kr_t));
    }
}

static int check_inputs(const orc_col* key, const orc_col* vals, int32_t nvals)
{
    if (!key) return fail("Aggregation only is not supported yet");
    if (!key_type_ok(key->type)) return fail("unsupported key type");
    for (int c = 0; c < nvals; ++c)
        if (vals[c].size != key->size) return fail("value column size differs from key column");
    return 0;
}

/* The brute-force oracle. Rows of the result are in ascending key order. */
int orc_groupby(const orc_col* key, const orc_col* vals, int32_t nvals,
                const orc_req* reqs, int32_t nreqs, int32_t agg_count,
                orc_result* res)
{
    plan_t p;
    if (check_inputs(key, vals, nvals)) return -1;
    if (build_plan(vals, nvals, reqs, nreqs, agg_count, &p)) return -1;
    int64_t n = key->size;
    kr_t* a = (kr_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kr_t));
    kr_t* t = (kr_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kr_t));
    if (!a || !t) { free(a); free(t); return fail("out of memory"); }
    for (int64_t i = 0; i < n; ++i) { a[i].k = key_ord(key->data, key->type, i); a[i].r = i; }
    merge_sort_kr(a, t, n);
    int64_t groups = 0;
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || a[i].k != a[i - 1].k) ++groups;
    if (alloc_result(res, &p, key->type, groups)) { free(a); free(t); return -1; }
    state_t st[ORC_MAX_COLS];
    int64_t g = -1;
    for (int64_t i = 0; i < n; ++i) {
        int64_t row = a[i].r;
        int first = (i == 0 || a[i].k != a[i - 1].k);
        if (first) {
            if (g >= 0)
                for (int j = 0; j < p.nout; ++j) emit(&p, j, st, res->cols[1 + j], g);
            ++g;
            memcpy((char*)res->cols[0] + g * type_size(key->type),
                   (const char*)key->data + row * type_size(key->type), (size_t)type_size(key->type));
            for (int k = 0; k < p.nops; ++k) state_init(&p.ops[k], &st[k], 1);
        }
        for (int k = 0; k < p.nops; ++k) state_update(&p.ops[k], &st[k], vals, row, first);
    }
    if (g >= 0)
        for (int j = 0; j < p.nout; ++j) emit(&p, j, st, res->cols[1 + j], g);
    free(a);
    free(t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Parallel CPU group-by: the timed CPU baseline.                             */
/* Pass 1 scatters row ids into 2^PB partitions by the top bits of bit_mix    */
/* (per-thread histograms + prefix, so the scatter is stable); pass 2 runs    */
/* one open-addressing table per partition.  Output rows are in partition   */
/* order, then first-encounter order.                                         */
/* ------------------------------------------------------------------------ */
#define PAR_PB 10

int orc_groupby_parallel(const orc_col* key, const orc_col* vals, int32_t nvals,
                         const orc_req* reqs, int32_t nreqs, int32_t agg_count,
                         int32_t nthreads, orc_result* res)
{
    plan_t p;
    if (check_inputs(key, vals, nvals)) return -1;
    if (key->type == ORC_FLOAT32 || key->type == ORC_FLOAT64) return fail("float keys: brute-force oracle only");
    if (build_plan(vals, nvals, reqs, nreqs, agg_count, &p)) return -1;
    const int64_t n = key->size;
    const int P = 1 << PAR_PB;
    int T = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
    T = 1;
#endif
    int64_t* hist = (int64_t*)calloc((size_t)T * P, sizeof(int64_t));
    int64_t* rows = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t* pstart = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    int64_t* gcount = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    if (!hist || !rows || !pstart || !gcount) {
        free(hist); free(rows); free(pstart); free(gcount);
        return fail("out of memory");
    }
    const int shift = 64 - PAR_PB;
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t lo = n * t / T, hi = n * (t + 1) / T;
        int64_t* h = hist + (size_t)t * P;
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t m = bit_mix(key->data, key->type, i);
            if (type_size(key->type) <= 4) m <<= 32;
            h[m >> shift]++;
        }
#pragma omp barrier
#pragma omp single
        {
            int64_t acc = 0;
            for (int b = 0; b < P; ++b) {
                pstart[b] = acc;
                for (int u = 0; u < T; ++u) {
                    int64_t c = hist[(size_t)u * P + b];
                    hist[(size_t)u * P + b] = acc;
                    acc += c;
                }
            }
            pstart[P] = acc;
        }
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t m = bit_mix(key->data, key->type, i);
            if (type_size(key->type) <= 4) m <<= 32;
            rows[h[m >> shift]++] = i;
        }
    }
    /* pass 2: per-partition hash aggregation into per-partition buffers */
    const int nops = p.nops;
    state_t** pst = (state_t**)calloc((size_t)P, sizeof(state_t*));
    int64_t** pfirst = (int64_t**)calloc((size_t)P, sizeof(int64_t*));
    int err = (!pst || !pfirst);
    if (!err) {
#pragma omp parallel for num_threads(T) schedule(dynamic, 4)
        for (int b = 0; b < P; ++b) {
            int64_t cnt = pstart[b + 1] - pstart[b];
            if (cnt == 0) continue;
            int64_t cap = 16;
            while (cap < 2 * cnt) cap <<= 1;
            int64_t* slot_row = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
            int64_t* slot_gid = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
            state_t* st = (state_t*)malloc((size_t)cnt * nops * sizeof(state_t));
            int64_t* first = (int64_t*)malloc((size_t)cnt * sizeof(int64_t));
            if (!slot_row || !slot_gid || !st || !first) {
                free(slot_row); free(slot_gid); free(st); free(first);
                continue; /* reported below as a missing partition */
            }
            for (int64_t s = 0; s < cap; ++s) slot_row[s] = -1;
            int64_t ng = 0;
            for (int64_t q = pstart[b]; q < pstart[b + 1]; ++q) {
                int64_t r = rows[q];
                uint64_t kr = raw_bits(key->data, key->type, r);
                uint64_t h = orc_fmix64(kr ^ 0x9e3779b97f4a7c15ULL);
                int64_t s = (int64_t)(h & (uint64_t)(cap - 1));
                int64_t gid;
                for (;;) {
                    if (slot_row[s] < 0) {
                        slot_row[s] = r;
                        gid = ng++;
                        slot_gid[s] = gid;
                        first[gid] = r;
                        for (int k = 0; k < nops; ++k) state_init(&p.ops[k], &st[gid * nops + k], 1);
                        for (int k = 0; k < nops; ++k) state_update(&p.ops[k], &st[gid * nops + k], vals, r, 1);
                        break;
                    }
                    if (raw_bits(key->data, key->type, slot_row[s]) == kr) {
                        gid = slot_gid[s];
                        for (int k = 0; k < nops; ++k) state_update(&p.ops[k], &st[gid * nops + k], vals, r, 0);
                        break;
                    }
                    s = (s + 1) & (cap - 1);
                }
            }
            free(slot_row);
            free(slot_gid);
            pst[b] = st;
            pfirst[b] = first;
            gcount[b] = ng;
        }
        for (int b = 0; b < P; ++b)
            if (pstart[b + 1] > pstart[b] && !pst[b]) err = 1;
    }
    int64_t total = 0;
    if (!err) {
        for (int b = 0; b < P; ++b) { int64_t c = gcount[b]; gcount[b] = total; total += c; }
        gcount[P] = total;
        if (alloc_result(res, &p, key->type, total)) err = 2;
    }
    if (!err) {
        const int ks = type_size(key->type);
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
        for (int b = 0; b < P; ++b) {
            int64_t ng = gcount[b + 1] - gcount[b];
            for (int64_t g = 0; g < ng; ++g) {
                int64_t o = gcount[b] + g;
                memcpy((char*)res->cols[0] + o * ks, (const char*)key->data + pfirst[b][g] * ks, (size_t)ks);
                for (int j = 0; j < p.nout; ++j) emit(&p, j, &pst[b][g * nops], res->cols[1 + j], o);
            }
        }
    }
    if (pst) for (int b = 0; b < P; ++b) free(pst[b]);
    if (pfirst) for (int b = 0; b < P; ++b) free(pfirst[b]);
    free(pst); free(pfirst); free(hist); free(rows); free(pstart); free(gcount);
    if (err == 1) return fail("out of memory");
    return err ? -1 : 0;
}

/* Scalar aggregate (operators/groupby/aggregate.cu:131-165): one row per
 * request; COUNT_ALL = number of rows; AVG = double(SUM)/rows. */
int orc_aggregate(const orc_col* cols, int32_t ncols, const orc_req* reqs,
                  int32_t nreqs, orc_result* res)
{
    memset(res, 0, sizeof *res);
    if (nreqs > ORC_MAX_COLS) return fail("too many aggregation requests");
    int64_t n = ncols > 0 ? cols[0].size : 0;
    res->num_rows = 1;
    res->num_cols = nreqs;
    for (int r = 0; r < nreqs; ++r) {
        int kind = reqs[r].agg_kind;
        if (kind == ORC_AGG_COUNT_ALL) {
            res->types[r] = ORC_INT64;
            res->cols[r] = malloc(8);
            ((int64_t*)res->cols[r])[0] = n;
            continue;
        }
        int64_t c = reqs[r].column_index;
        if (c < 0 || c >= ncols) { orc_result_free(res); return fail("column index out of range"); }
        int32_t vt = cols[c].type;
        if (!value_type_ok(vt)) { orc_result_free(res); return fail("unsupported value column type"); }
        op_t o;
        o.col = c;
        o.vtype = vt;
        switch (kind) {
        case ORC_AGG_SUM: case ORC_AGG_AVG: o.kind = OP_SUM; break;
        case ORC_AGG_MAX: o.kind = OP_MAX; break;
        case ORC_AGG_MIN: o.kind = OP_MIN; break;
        case ORC_AGG_ANY: o.kind = OP_ANY; break;
        default: orc_result_free(res); return fail("Unsupported aggregation function");
        }
        state_t s;
        state_init(&o, &s, 1);
        for (int64_t i = 0; i < cols[c].size; ++i) state_update(&o, &s, cols, i, i == 0);
        int32_t ot = orc_agg_output_type(kind, vt);
        res->types[r] = ot;
        res->cols[r] = calloc(1, 8);
        if (kind == ORC_AGG_AVG) {
            double sum = is_float(vt) ? s.d : (is_signed_int(vt) ? (double)(int64_t)s.u : (double)s.u);
            ((double*)res->cols[r])[0] = sum / (double)cols[c].size;
        } else if (ot == ORC_FLOAT32) {
            ((float*)res->cols[r])[0] = (float)s.d;
        } else if (ot == ORC_FLOAT64) {
            ((double*)res->cols[r])[0] = s.d;
        } else if (kind == ORC_AGG_ANY) {
            memcpy(res->cols[r], &s.u, (size_t)type_size(ot));
        } else if (is_signed_int(ot)) {
            store_int(res->cols[r], ot, 0, kind == ORC_AGG_SUM ? (int64_t)s.u : s.i);
        } else {
            store_int(res->cols[r], ot, 0, (int64_t)s.u);
        }
    }
    return 0;
}

/* ======================================================================== */
/* Sort / partition sub-operators                                             */
/* ======================================================================== */

/* Radix-sortable bits of element i: the key with its sign bit flipped for
 * signed integers, and the usual order-preserving transform for floats (the
 * ordering cub::DeviceRadixSort sorts by, operators/sort/sort.cu:15-66).
 * cub documents that -0.0 and +0.0 sort as equal keys: -0.0 takes the bits
 * of +0.0 before the transform, so a stable sort keeps zeros in input order. */
static uint64_t twiddle(const void* k, int32_t t, int64_t i)
{
    uint64_t b = raw_bits(k, t, i);
    int w = 8 * type_size(t);
    uint64_t sign = (uint64_t)1 << (w - 1);
    if (is_signed_int(t)) return b ^ sign;
    if (is_float(t)) {
        uint64_t all = (w == 64) ? ~(uint64_t)0 : ((((uint64_t)1) << w) - 1);
        if (b == sign) b = 0;
        return (b & sign) ? (~b & all) : (b ^ sign);
    }
    return b;
}

/* Stable sort of (key, value) pairs by bits [begin_bit, end_bit) of the
 * twiddled key, ascending or descending (sort.cu:115-153 / 203-215). */
int orc_sort_pairs(const void* keys, int32_t key_type, const void* vals,
                   int32_t val_bytes, int64_t n, int32_t begin_bit,
                   int32_t end_bit, int32_t descending, void* out_keys,
                   void* out_vals)
{
    int w = 8 * type_size(key_type);
    if (!w) return fail("unsupported key type");
    if (end_bit <= 0) end_bit = w;
    if (begin_bit < 0 || begin_bit > end_bit || end_bit > w) return fail("bad bit range");
    int nb = end_bit - begin_bit;
    uint64_t mask = nb >= 64 ? ~(uint64_t)0 : ((((uint64_t)1) << nb) - 1);
    kr_t* a = (kr_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kr_t));
    kr_t* t = (kr_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kr_t));
    if (!a || !t) { free(a); free(t); return fail("out of memory"); }
    for (int64_t i = 0; i < n; ++i) {
        uint64_t d = (twiddle(keys, key_type, i) >> begin_bit) & mask;
        a[i].k = descending ? (~d & mask) : d;
        a[i].r = i;
    }
    merge_sort_kr(a, t, n);
    int ks = type_size(key_type);
    for (int64_t i = 0; i < n; ++i) {
        if (out_keys)
            memcpy((char*)out_keys + i * ks, (const char*)keys + a[i].r * ks, (size_t)ks);
        if (out_vals && vals)
            memcpy((char*)out_vals + i * val_bytes, (const char*)vals + a[i].r * val_bytes, (size_t)val_bytes);
    }
    free(a);
    free(t);
    return 0;
}

/* compute_partition_offsets (operators/partition/partition.cu:122-148):
 * histogram of (key >> begin_bit) & mask over 2^(end-begin) bins, then an
 * inclusive prefix sum.  offsets must hold 2^(end-begin)+1 entries; the last
 * one (left unwritten by the reference) is set to n here. */
int orc_partition_offsets(const void* keys, int32_t key_type, int64_t n,
                          int32_t begin_bit, int32_t end_bit, int64_t* offsets)
{
    if (!key_type_ok(key_type)) return fail("Cannot extract radix from non-integral type");
    int nb = end_bit - begin_bit;
    if (nb <= 0 || nb > 24) return fail("bad bit range");
    int64_t P = (int64_t)1 << nb;
    memset(offsets, 0, (size_t)(P + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        uint64_t d = ((uint64_t)as_i64(keys, key_type, i) >> begin_bit) & (uint64_t)(P - 1);
        offsets[d]++;
    }
    int64_t acc = 0;
    for (int64_t b = 0; b < P; ++b) { acc += offsets[b]; offsets[b] = acc; }
    offsets[P] = n;
    return 0;
}

/* Hash partition by pid = (bit_mix(key) >> shift) & (2^bits - 1): stable
 * permutation perm (row ids in partition order) and exclusive offsets
 * (2^bits + 1 entries).  The reference builds the same grouping with a 16-bit
 * pid and a stable radix sort (groupby_partition.cu:580-663). */
int orc_hash_partition(const void* keys, int32_t key_type, int64_t n,
                       int32_t bits, int32_t shift, int64_t* perm,
                       int64_t* offsets)
{
    if (!key_type_ok(key_type)) return fail("unsupported key type");
    if (bits <= 0 || bits > 24) return fail("bad partition bits");
    int64_t P = (int64_t)1 << bits;
    memset(offsets, 0, (size_t)(P + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        uint64_t d = (bit_mix(keys, key_type, i) >> shift) & (uint64_t)(P - 1);
        offsets[d + 1]++;
    }
    for (int64_t b = 0; b < P; ++b) offsets[b + 1] += offsets[b];
    int64_t* cur = (int64_t*)malloc((size_t)P * sizeof(int64_t));
    if (!cur) return fail("out of memory");
    memcpy(cur, offsets, (size_t)P * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        uint64_t d = (bit_mix(keys, key_type, i) >> shift) & (uint64_t)(P - 1);
        perm[cur[d]++] = i;
    }
    free(cur);
    return 0;
}

/* ======================================================================== */
/* Dispatch                                                                   */
/* ======================================================================== */

/* auto_select_groupby_algo thresholds (operators/groupby/groupby.cu:104-118):
 * 4-byte keys: NDV <= 2^14 -> HASH else PARTITION;
 * 8-byte keys: NDV <= 2^18 -> HASH else PARTITION. */
int32_t orc_rule_select(int32_t key_bytes, int64_t ndv)
{
    if (key_bytes == 4) return ndv <= ((int64_t)1 << 14) ? ORC_ALGO_HASH : ORC_ALGO_PARTITION;
    return ndv <= ((int64_t)1 << 18) ? ORC_ALGO_HASH : ORC_ALGO_PARTITION;
}

int64_t orc_count_distinct(const void* keys, int32_t type, int64_t n)
{
    if (n <= 0) return 0;
    kr_t* a = (kr_t*)malloc((size_t)n * sizeof(kr_t));
    kr_t* t = (kr_t*)malloc((size_t)n * sizeof(kr_t));
    if (!a || !t) { free(a); free(t); return -1; }
    for (int64_t i = 0; i < n; ++i) { a[i].k = key_ord(keys, type, i); a[i].r = i; }
    merge_sort_kr(a, t, n);
    int64_t d = 1;
    for (int64_t i = 1; i < n; ++i) d += (a[i].k != a[i - 1].k);
    free(a);
    free(t);
    return d;
}

/* compress/dictionary_encode_dict_ready.cu:15-80 (cuco static_map insert of
 * (dict[i], i), then find): out[r] = the position of keys[r] in dict, `miss`
 * (the reference's sentinel<int32_t>() = INT32_MIN) when absent.  A key held
 * more than once in dict maps to its first position (the reference's map keeps
 * whichever insert wins; dictionaries there hold unique keys).  Stable sort of
 * (key, position) + lower-bound search. */
int orc_dict_encode(const void* keys, const void* dict, int32_t type, int64_t n, int64_t m, int64_t miss,
                    int64_t* out)
{
    kr_t* a = (kr_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(kr_t));
    kr_t* t = (kr_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(kr_t));
    if (!a || !t) { free(a); free(t); return fail("out of memory"); }
    for (int64_t i = 0; i < m; ++i) { a[i].k = key_ord(dict, type, i); a[i].r = i; }
    merge_sort_kr(a, t, m);
    for (int64_t r = 0; r < n; ++r) {
        const uint64_t k = key_ord(keys, type, r);
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (a[mid].k < k) lo = mid + 1; else hi = mid;
        }
        out[r] = (lo < m && a[lo].k == k) ? a[lo].r : miss;
    }
    free(a);
    free(t);
    return 0;
}

/* compress/dictionary_encode.cu (dictionary_encode_with_order): out[r] = the
 * rank of keys[r] among the distinct keys in ascending order; returns the
 * number of distinct keys, -1 on allocation failure. */
int64_t orc_dict_encode_ordered(const void* keys, int32_t type, int64_t n, int64_t* out)
{
    if (n <= 0) return 0;
    kr_t* a = (kr_t*)malloc((size_t)n * sizeof(kr_t));
    kr_t* t = (kr_t*)malloc((size_t)n * sizeof(kr_t));
    if (!a || !t) { free(a); free(t); return -1; }
    for (int64_t i = 0; i < n; ++i) { a[i].k = key_ord(keys, type, i); a[i].r = i; }
    merge_sort_kr(a, t, n);
    int64_t d = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && a[i].k != a[i - 1].k) ++d;
        out[a[i].r] = d;
    }
    free(a);
    free(t);
    return d + 1;
}

/* ------------------------------------------------------------------------ */
/* Checker helpers for full-size results (no reference counterpart: the     */
/* reference's tests compare group-by results after sorting them by key,    */
/* floating_point_test.cu:16).                                              */
/* ------------------------------------------------------------------------ */

/* perm = the row order that sorts keys ascending (signed keys by value,
 * unsigned by value), stable: an OpenMP LSD radix sort on 8-bit digits of the
 * order-preserving 64-bit key, passes whose digit is the same for every row
 * skipped.  Returns 0, or -1 when memory runs out. */
int orc_argsort_keys(const void* keys, int32_t type, int64_t n, int32_t nthreads, int64_t* perm)
{
    if (n <= 0) return 0;
    if (!key_type_ok(type)) return fail("argsort: unsupported key type");
    int T = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
    T = 1;
#endif
    uint64_t* ka = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* kb = (uint64_t*)malloc((size_t)n * 8);
    int64_t* pb = (int64_t*)malloc((size_t)n * 8);
    int64_t* hist = (int64_t*)malloc((size_t)T * 256 * 8);
    if (!ka || !kb || !pb || !hist) {
        free(ka); free(kb); free(pb); free(hist);
        return fail("out of memory");
    }
    int64_t* pa = perm;
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        ka[i] = key_ord(keys, type, i);
        pa[i] = i;
    }
    for (int shift = 0; shift < 64; shift += 8) {
        int skip = 0;
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const int64_t lo = n * t / T, hi = n * (t + 1) / T;
            int64_t* h = hist + (size_t)t * 256;
            for (int d = 0; d < 256; ++d) h[d] = 0;
            for (int64_t i = lo; i < hi; ++i) h[(ka[i] >> shift) & 255]++;
#pragma omp barrier
#pragma omp single
            {
                int64_t acc = 0;
                for (int d = 0; d < 256; ++d) {
                    int64_t tot = 0;
                    for (int u = 0; u < T; ++u) tot += hist[(size_t)u * 256 + d];
                    if (tot == n) skip = 1;
                    for (int u = 0; u < T; ++u) {
                        const int64_t c = hist[(size_t)u * 256 + d];
                        hist[(size_t)u * 256 + d] = acc;
                        acc += c;
                    }
                }
            }
            if (!skip)
                for (int64_t i = lo; i < hi; ++i) {
                    const int64_t o = h[(ka[i] >> shift) & 255]++;
                    kb[o] = ka[i];
                    pb[o] = pa[i];
                }
        }
        if (skip) continue;
        uint64_t* tk = ka; ka = kb; kb = tk;
        int64_t* tp = pa; pa = pb; pb = tp;
    }
    if (pa != perm) {
        memcpy(perm, pa, (size_t)n * 8);
        pb = pa;
    }
    free(ka); free(kb); free(pb); free(hist);
    return 0;
}

/* dst[i] = src[idx[i]] for elements of `bytes` bytes (1, 2, 4 or 8). */
int orc_gather(const void* src, int32_t bytes, const int64_t* idx, int64_t n, void* dst, int32_t nthreads)
{
    int T = nthreads > 0 ? nthreads : 1;
    if (bytes != 1 && bytes != 2 && bytes != 4 && bytes != 8) return fail("gather: element size");
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < n; ++i)
        memcpy((char*)dst + i * bytes, (const char*)src + idx[i] * bytes, (size_t)bytes);
    return 0;
}

// eg_device.hpp -- device-side building blocks shared by every group-by kernel:
// hashing, key and order encodings, typed loads, operand construction, and the
// emission of one finished group into the output columns.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "eg_spec.hpp"

namespace eg {
namespace dev {

constexpr uint64_t kSign64 = 0x8000000000000000ull;
constexpr uint32_t kSign32 = 0x80000000u;

// MurmurHash3 finalizers (Austin Appleby, public domain); the reference uses
// the same mixers for partition ids (operators/groupby/groupby_partition.cu:64-90).
__device__ __forceinline__ uint32_t fmix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint64_t fmix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// bit_mix of groupby_partition.cu:83-90: fmix32 for keys of <= 4 bytes,
// fmix64 otherwise.
template <int KB>
__device__ __forceinline__ uint64_t bit_mix(uint64_t raw)
{
    if constexpr (KB <= 4) return (uint64_t)fmix32((uint32_t)raw);
    else return fmix64(raw);
}

// ---- key encodings -------------------------------------------------------
// Table keys are stored encoded so that the all-zero word means "empty":
//   enc64(k) = widen(k) ^ 2^63   (widen = sign- or zero-extension)
// For 4-byte keys enc64 is never zero.  For 8-byte keys exactly one key value
// (INT64_MIN, or 2^63 for uint64) encodes to zero; it lives in a side slot.
__device__ __forceinline__ uint64_t widen_key(uint64_t raw, int kb, int is_signed)
{
    if (kb == 4) return is_signed ? (uint64_t)(int64_t)(int32_t)(uint32_t)raw : (uint64_t)(uint32_t)raw;
    return raw;
}

__device__ __forceinline__ uint64_t enc64(uint64_t wide) { return wide ^ kSign64; }
__device__ __forceinline__ uint64_t dec64(uint64_t enc) { return enc ^ kSign64; }

// ---- order-preserving encodings -----------------------------------------
__device__ __forceinline__ uint64_t ord_i64(int64_t v) { return (uint64_t)v ^ kSign64; }

__device__ __forceinline__ uint64_t ord_f64(double d)
{
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b & kSign64) ? ~b : (b | kSign64);
}

__device__ __forceinline__ double unord_f64(uint64_t o)
{
    uint64_t b = (o & kSign64) ? (o ^ kSign64) : ~o;
    return __longlong_as_double((long long)b);
}

// ---- typed loads -----------------------------------------------------------
// Raw element bits, zero-extended to 64 bits.
__device__ __forceinline__ uint64_t load_raw(const void* p, int bytes, int64_t i)
{
    switch (bytes) {
    case 1: return ((const uint8_t*)p)[i];
    case 2: return ((const uint16_t*)p)[i];
    case 4: return ((const uint32_t*)p)[i];
    default: return ((const uint64_t*)p)[i];
    }
}

// type ids (gamax::type_id numbering)
enum : int32_t {
    T_INT8 = 1, T_INT16 = 2, T_INT32 = 3, T_INT64 = 4, T_BOOL8 = 5, T_UINT8 = 6,
    T_UINT16 = 7, T_UINT32 = 8, T_UINT64 = 9, T_FLOAT32 = 10, T_FLOAT64 = 11
};

__host__ __device__ __forceinline__ int type_bytes(int32_t type)
{
    switch (type) {
    case T_INT8: case T_UINT8: case T_BOOL8: return 1;
    case T_INT16: case T_UINT16: return 2;
    case T_INT32: case T_UINT32: case T_FLOAT32: return 4;
    default: return 8;
    }
}

__device__ __forceinline__ int64_t raw_as_i64(uint64_t raw, int32_t type)
{
    switch (type) {
    case T_INT8: return (int64_t)(int8_t)(uint8_t)raw;
    case T_INT16: return (int64_t)(int16_t)(uint16_t)raw;
    case T_INT32: return (int64_t)(int32_t)(uint32_t)raw;
    default: return (int64_t)raw;
    }
}

__device__ __forceinline__ double raw_as_f64(uint64_t raw, int32_t type)
{
    switch (type) {
    case T_FLOAT32: return (double)__uint_as_float((uint32_t)raw);
    case T_FLOAT64: return __longlong_as_double((long long)raw);
    case T_INT8: case T_INT16: case T_INT32: case T_INT64: return (double)raw_as_i64(raw, type);
    default: return (double)raw;
    }
}

// Canonical 64-bit value of a column element (ConvCode): integers sign- or
// zero-extended, float32 widened to double bits, 8-byte types unchanged.
__device__ __forceinline__ uint64_t canon(uint64_t raw, int32_t conv)
{
    switch (conv) {
    case CONV_SEXT32: return (uint64_t)(int64_t)(int32_t)(uint32_t)raw;
    case CONV_F32: return (uint64_t)__double_as_longlong((double)__uint_as_float((uint32_t)raw));
    case CONV_SEXT16: return (uint64_t)(int64_t)(int16_t)(uint16_t)raw;
    case CONV_SEXT8: return (uint64_t)(int64_t)(int8_t)(uint8_t)raw;
    default: return raw;
    }
}

// Same with the conversion known at compile time.
template <int CONV>
__device__ __forceinline__ uint64_t canon_c(uint64_t raw)
{
    if constexpr (CONV == CONV_SEXT32) return (uint64_t)(int64_t)(int32_t)(uint32_t)raw;
    else if constexpr (CONV == CONV_F32) return (uint64_t)__double_as_longlong((double)__uint_as_float((uint32_t)raw));
    else if constexpr (CONV == CONV_SEXT16) return (uint64_t)(int64_t)(int16_t)(uint16_t)raw;
    else if constexpr (CONV == CONV_SEXT8) return (uint64_t)(int64_t)(int8_t)(uint8_t)raw;
    else return raw;
}

// ---- op lists ----------------------------------------------------------------
// Kernels read their aggregation ops through an op-list policy:
//   RtOps      the AggSpec's ops, read from the kernel arguments: one
//              compiled kernel serves every request list;
//   FixOps<W>  the same ops as template constants (W = code | (col + 1) << 4,
//              in class order), so a row's work folds to straight-line code:
//              exactly these operand transforms and atomics.
// Both read the value columns from the AggSpec and give identical results.
struct RtOps {
    static constexpr bool kFixed = false;
    __device__ static int n(const AggSpec& s) { return s.nops; }
    __device__ static int col(const AggSpec& s, int a) { return s.ops[a].col; }
    __device__ static uint64_t xmask(const AggSpec& s, int a) { return s.ops[a].xmask; }
    __device__ static uint64_t fmask(const AggSpec& s, int a) { return s.ops[a].fmask; }
    // class of op a from the class-ordered list
    __device__ static int cls(const AggSpec& s, int a)
    {
        const int e0 = s.ncls[0], e1 = e0 + s.ncls[1], e2 = e1 + s.ncls[2];
        return a < e0 ? CLS_ADD_U64 : a < e1 ? CLS_ADD_F64 : a < e2 ? CLS_MAX_U64 : CLS_XCHG;
    }
};

template <uint32_t... W>
struct FixOps {
    static constexpr bool kFixed = true;
    static constexpr int kN = sizeof...(W);
    static_assert(kN >= 1 && kN <= kMaxOps, "op count");
    __device__ static constexpr int word(int a)
    {
        constexpr uint32_t w[kN] = {W...};
        return (int)w[a];
    }
    __device__ static constexpr int n(const AggSpec&) { return kN; }
    __device__ static constexpr int col(const AggSpec&, int a) { return ((word(a) >> 4) & 15) - 1; }
    __device__ static constexpr uint64_t xmask(const AggSpec&, int a) { return op_xmask(word(a) & 15); }
    __device__ static constexpr uint64_t fmask(const AggSpec&, int a) { return op_fmask(word(a) & 15); }
    __device__ static constexpr int cls(const AggSpec&, int a) { return op_class(word(a) & 15); }
};

template <int... I, class F>
__device__ __forceinline__ void static_for_seq(std::integer_sequence<int, I...>, F&& f)
{
    (f(std::integral_constant<int, I>{}), ...);
}

// f(a) for every op a (unrolled with constant a for a fixed op list).
template <class OPS, class F>
__device__ __forceinline__ void for_each_op(const AggSpec& s, F&& f)
{
    if constexpr (OPS::kFixed) static_for_seq(std::make_integer_sequence<int, OPS::kN>{}, f);
    else
        for (int a = 0; a < s.nops; ++a) f(a);
}

// Operand of op a for a row with canonical column values cv (see OpDesc).
template <class OPS, int NV>
__device__ __forceinline__ uint64_t op_operand(const AggSpec& s, int a, const uint64_t (&cv)[NV])
{
    const int c = OPS::col(s, a);
    uint64_t v = 1ull;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        if (k == c) v = cv[k];
    return v ^ OPS::xmask(s, a) ^ ((uint64_t)((int64_t)v >> 63) & OPS::fmask(s, a));
}

// Operand of op `od` for canonical value v (see OpDesc).
__device__ __forceinline__ uint64_t operand(const OpDesc& od, uint64_t v)
{
    return v ^ od.xmask ^ ((uint64_t)((int64_t)v >> 63) & od.fmask);
}

__device__ __forceinline__ int32_t class_of(const AggSpec& spec, int a) { return RtOps::cls(spec, a); }

// ---- state updates ---------------------------------------------------------
// Merge operand `v` into the 64-bit state at `p` (LDS or global; the compiler
// picks ds_* or global_* atomics from the address space of the pointer).
template <typename P>
__device__ __forceinline__ void apply_atomic(int32_t cls, P* p, uint64_t v)
{
    switch (cls) {
    case CLS_ADD_U64: atomicAdd((unsigned long long*)p, (unsigned long long)v); break;
    case CLS_ADD_F64: atomicAdd((double*)p, __longlong_as_double((long long)v)); break;
    case CLS_MAX_U64: atomicMax((unsigned long long*)p, (unsigned long long)v); break;
    default: atomicExch((unsigned long long*)p, (unsigned long long)v); break;
    }
}

// Non-atomic merge (single owner of the state).
__device__ __forceinline__ uint64_t combine(int32_t cls, uint64_t a, uint64_t b)
{
    switch (cls) {
    case CLS_ADD_U64: return a + b;
    case CLS_ADD_F64:
        return (uint64_t)__double_as_longlong(__longlong_as_double((long long)a) +
                                              __longlong_as_double((long long)b));
    case CLS_MAX_U64: return a > b ? a : b;
    default: return b;
    }
}

// Merges the row (canonical column values cv) into the states at ptr(a) for
// every op: one tight loop per atomic class for the runtime list, straight
// code for a fixed one.
template <class OPS, int NV, class PtrFn>
__device__ __forceinline__ void apply_ops(const AggSpec& spec, const uint64_t (&cv)[NV], PtrFn ptr)
{
    if constexpr (OPS::kFixed) {
        for_each_op<OPS>(spec, [&](auto a) { apply_atomic(OPS::cls(spec, a), ptr(a), op_operand<OPS, NV>(spec, a, cv)); });
    } else {
        auto val = [&](int a) { return op_operand<OPS, NV>(spec, a, cv); };
        const int e0 = spec.ncls[0], e1 = e0 + spec.ncls[1], e2 = e1 + spec.ncls[2];
        int a = 0;
        for (; a < e0; ++a) atomicAdd((unsigned long long*)ptr(a), (unsigned long long)val(a));
        for (; a < e1; ++a) atomicAdd((double*)ptr(a), __longlong_as_double((long long)val(a)));
        for (; a < e2; ++a) atomicMax((unsigned long long*)ptr(a), (unsigned long long)val(a));
        for (; a < spec.nops; ++a) atomicExch((unsigned long long*)ptr(a), (unsigned long long)val(a));
    }
}

// Same, merging 64-bit partial states st(a) (already operands).
template <class OPS, class PtrFn, class StFn>
__device__ __forceinline__ void merge_states(const AggSpec& spec, PtrFn ptr, StFn st)
{
    if constexpr (OPS::kFixed) {
        for_each_op<OPS>(spec, [&](auto a) { apply_atomic(OPS::cls(spec, a), ptr(a), (uint64_t)st(a)); });
    } else {
        const int e0 = spec.ncls[0], e1 = e0 + spec.ncls[1], e2 = e1 + spec.ncls[2];
        int a = 0;
        for (; a < e0; ++a) atomicAdd((unsigned long long*)ptr(a), (unsigned long long)st(a));
        for (; a < e1; ++a) atomicAdd((double*)ptr(a), __longlong_as_double((long long)st(a)));
        for (; a < e2; ++a) atomicMax((unsigned long long*)ptr(a), (unsigned long long)st(a));
        for (; a < spec.nops; ++a) atomicExch((unsigned long long*)ptr(a), (unsigned long long)st(a));
    }
}

// ---- output ----------------------------------------------------------------
__device__ __forceinline__ void store_bytes(void* p, int bytes, int64_t i, uint64_t bits)
{
    switch (bytes) {
    case 1: ((uint8_t*)p)[i] = (uint8_t)bits; break;
    case 2: ((uint16_t*)p)[i] = (uint16_t)bits; break;
    case 4: ((uint32_t*)p)[i] = (uint32_t)bits; break;
    default: ((uint64_t*)p)[i] = bits; break;
    }
}

__device__ __forceinline__ void store_double_as(void* p, int bytes, int64_t i, double d)
{
    if (bytes == 4) ((float*)p)[i] = (float)d;
    else ((double*)p)[i] = d;
}

// One output value of group g from its state s (and the count state cnt for
// AVG); with constant code/bytes/is_float the switch folds away.
__device__ __forceinline__ void emit_one(int32_t code, int32_t bytes, int32_t is_float, void* data, int64_t g,
                                         uint64_t s, uint64_t cnt)
{
    switch (code) {
    case OUT_U64: store_bytes(data, 8, g, s); break;
    case OUT_F32: ((float*)data)[g] = (float)__longlong_as_double((long long)s); break;
    case OUT_F64: ((double*)data)[g] = __longlong_as_double((long long)s); break;
    case OUT_MAX_I: store_bytes(data, bytes, g, s ^ kSign64); break;
    case OUT_MIN_I: store_bytes(data, bytes, g, (~s) ^ kSign64); break;
    case OUT_MAX_U: store_bytes(data, bytes, g, s); break;
    case OUT_MIN_U: store_bytes(data, bytes, g, ~s); break;
    case OUT_MAX_F: store_double_as(data, bytes, g, unord_f64(s)); break;
    case OUT_MIN_F: store_double_as(data, bytes, g, unord_f64(~s)); break;
    case OUT_ANY:
        // canonical value: truncation restores integers, floats come back from double
        if (is_float && bytes == 4) ((float*)data)[g] = (float)__longlong_as_double((long long)s);
        else store_bytes(data, bytes, g, s);
        break;
    default: {
        const double sum = (code == OUT_AVG_I) ? (double)(int64_t)s
                         : (code == OUT_AVG_U) ? (double)s
                         : __longlong_as_double((long long)s);
        ((double*)data)[g] = sum / (double)cnt;
        break;
    }
    }
}

// Writes group `g`: its key (from the 64-bit encoding) and every output
// column.  States of op a are read at st[a * stride].
template <typename W>
__device__ __forceinline__ void emit_group(const OutSpec& os, int64_t g, uint64_t key_enc,
                                           const W* st, int64_t stride)
{
    store_bytes(os.key_data, os.key_bytes, g, dec64(key_enc));
    for (int j = 0; j < os.nouts; ++j) {
        const OutDesc& o = os.outs[j];
        const uint64_t s = (uint64_t)st[(int64_t)o.op * stride];
        const uint64_t cnt = o.op2 >= 0 ? (uint64_t)st[(int64_t)o.op2 * stride] : 0ull;
        emit_one(o.code, o.bytes, o.is_float, o.data, g, s, cnt);
    }
}

// Output-list policies of the partition unit kernel: RtOuts reads the output
// descriptors at run time (emit_group); FixOuts<W...> has them compiled in,
// one word per output column: code | bytes << 4 | op << 8 | (op2 + 1) << 16 |
// is_float << 24 (the run-time compiler instantiates it per request list).
struct RtOuts {
    static constexpr bool kFixed = false;
};

template <uint32_t... W>
struct FixOuts {
    static constexpr bool kFixed = true;
    static constexpr int kN = sizeof...(W);
    __device__ static constexpr uint32_t word(int j)
    {
        constexpr uint32_t w[kN > 0 ? kN : 1] = {W...};
        return w[j];
    }
};

template <class OUTS, typename W>
__device__ __forceinline__ void emit_outs(const OutSpec& os, int key_bytes, int64_t g, uint64_t key_enc, const W* st,
                                          int64_t stride)
{
    if constexpr (!OUTS::kFixed) {
        emit_group(os, g, key_enc, st, stride);
    } else {
        store_bytes(os.key_data, key_bytes, g, dec64(key_enc));
        static_for_seq(std::make_integer_sequence<int, OUTS::kN>{}, [&](auto jc) {
            constexpr int j = decltype(jc)::value;
            constexpr uint32_t w = OUTS::word(j);
            constexpr int32_t op = (int32_t)((w >> 8) & 0xff), op2 = (int32_t)((w >> 16) & 0xff) - 1;
            const uint64_t s = (uint64_t)st[(int64_t)op * stride];
            const uint64_t cnt = op2 >= 0 ? (uint64_t)st[(int64_t)op2 * stride] : 0ull;
            emit_one((int32_t)(w & 15), (int32_t)((w >> 4) & 15), (int32_t)((w >> 24) & 1), os.outs[j].data, g, s, cnt);
        });
    }
}

// ---- XCD-aware block order ------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b mod 8), and each XCD has its own L2.  xcd_remap(b, n) is a bijection on
// [0, n) giving XCD x the contiguous range of work items
// [x*n/8, (x+1)*n/8) (rounded), in order, so neighbouring items -- whose
// output cache lines overlap -- run on one XCD and their partial lines merge
// in that L2 instead of leaving as partial writes from two XCDs.
constexpr uint32_t kXcds = 8;

__host__ __device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n)
{
    const uint32_t x = b % kXcds, k = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// ---- wave helpers ----------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt()
{
    int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Wave-aggregated increment of a device counter: one atomic per wave; returns
// this lane's slot.  Every lane of the wave must call it.
__device__ __forceinline__ int64_t wave_claim(unsigned long long* counter, bool take)
{
    uint64_t m = __ballot(take);
    int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if (m != 0 && lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader < 0 ? 0 : leader);
    return (int64_t)(base + __popcll(m & lanemask_lt()));
}

// Workgroup-aggregated increment of a device counter: one atomic per
// workgroup (same-address device atomics serialise, so a wave-level claim
// over millions of waves becomes the bottleneck).  Every thread of the
// workgroup must call it; it synchronises.  `sh`: 17 words of LDS, at most
// 16 waves per workgroup.
__device__ __forceinline__ int64_t block_claim(unsigned long long* counter, bool take, unsigned long long* sh)
{
    const uint64_t m = __ballot(take);
    const int w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    if (lane_id() == 0) sh[w] = (unsigned long long)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < nw; ++i) {
            const unsigned long long c = sh[i];
            sh[i] = tot;
            tot += c;
        }
        sh[16] = tot ? atomicAdd(counter, tot) : 0ull;
    }
    __syncthreads();
    const int64_t r = (int64_t)(sh[16] + sh[w] + (unsigned long long)__popcll(m & lanemask_lt()));
    __syncthreads();  // sh is reused by the next call
    return r;
}

// Like block_claim for n (0..) rows per thread: this thread's first index,
// the block's rows contiguous in thread order.
__device__ __forceinline__ int64_t block_claim_n(unsigned long long* counter, uint32_t n, unsigned long long* sh)
{
    uint32_t x = n;  // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane_id() >= off) x += y;
    }
    const int w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    if (lane_id() == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < nw; ++i) {
            const unsigned long long c = sh[i];
            sh[i] = tot;
            tot += c;
        }
        sh[16] = tot ? atomicAdd(counter, tot) : 0ull;
    }
    __syncthreads();
    const int64_t r = (int64_t)(sh[16] + sh[w] + (unsigned long long)(x - n));
    __syncthreads();  // sh is reused by the next call
    return r;
}

// Exclusive scan of NT values held one per thread (NT threads; wsum: NT/64
// words of LDS).
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t add = 0;
    for (int i = 0; i < w; ++i) add += wsum[i];
    __syncthreads();
    return x - v + add;
}

}  // namespace dev
}  // namespace eg

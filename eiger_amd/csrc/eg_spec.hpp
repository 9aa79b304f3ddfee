// eg_spec.hpp -- plain-old-data descriptors shared by host code and kernels.
//
// A group-by request list is lowered on the host into an AggSpec: the distinct
// value columns it reads, and one "op" per aggregation state.  Every state is
// one 64-bit word whose identity is the all-zero word, so tables are cleared
// with zero fills:
//   add-u64   COUNT, integer SUM (wraps mod 2^64 like the reference's
//             unsigned long long atomics, atomics/atomics.cuh:32-45)
//   add-f64   float SUM (fp64 accumulation, rounded once at output)
//   max-u64   MAX of an order-preserving encoding, MIN of its complement
//   xchg      ANY (some row's raw value bits)
// The final columns are described by OutSpec entries that read one or two
// states (AVG = sum / count, groupby.cu:29-51).
#pragma once

#include <stdint.h>

namespace eg {

constexpr int kMaxOps = 16;     // aggregation states per group
constexpr int kMaxVCols = 8;    // distinct value columns per query
constexpr int kMaxOuts = 32;    // output columns after the key

// Operand codes: how a row's value becomes the 64-bit operand of its state.
enum OpCode : int32_t {
    OPC_COUNT = 0,    // operand 1, add-u64
    OPC_SUM_I = 1,    // sign-extended integer, add-u64
    OPC_SUM_U = 2,    // zero-extended integer, add-u64
    OPC_SUM_F = 3,    // value as double, add-f64
    OPC_MAX_I = 4,    // ord(int64), max-u64
    OPC_MIN_I = 5,    // ~ord(int64), max-u64
    OPC_MAX_U = 6,    // uint64, max-u64
    OPC_MIN_U = 7,    // ~uint64, max-u64
    OPC_MAX_F = 8,    // ord(double), max-u64
    OPC_MIN_F = 9,    // ~ord(double), max-u64
    OPC_ANY = 10      // raw bits, xchg
};

// Update classes (the atomic that merges an operand into a state).
enum OpClass : int32_t { CLS_ADD_U64 = 0, CLS_ADD_F64 = 1, CLS_MAX_U64 = 2, CLS_XCHG = 3 };

__host__ __device__ constexpr int32_t op_class(int32_t code)
{
    return (code == OPC_COUNT || code == OPC_SUM_I || code == OPC_SUM_U) ? CLS_ADD_U64
         : code == OPC_SUM_F ? CLS_ADD_F64
         : code == OPC_ANY ? CLS_XCHG
         : CLS_MAX_U64;
}

// How a column's raw bits become its canonical 64-bit value: integers
// sign- or zero-extended to 64 bits, floats converted to double bits.
enum ConvCode : int32_t { CONV_RAW = 0, CONV_SEXT32 = 1, CONV_SEXT16 = 2, CONV_SEXT8 = 3, CONV_F32 = 4 };

struct ColDesc {
    const void* data;   // device pointer (offset already applied)
    int32_t type;       // gamax type id
    int32_t bytes;      // element size
    int32_t conv;       // ConvCode
    int32_t pad;
};

// Operand of an op from the canonical value v of its column (v = 1 for
// COUNT):  operand = v ^ xmask ^ ((v >> 63, arithmetic) & fmask).
//   SUM, COUNT, ANY, MAX(uint)  xmask 0,      fmask 0
//   MIN(uint)                   xmask ~0,     fmask 0
//   MAX(int)                    xmask 2^63,   fmask 0
//   MIN(int)                    xmask ~2^63,  fmask 0
//   MAX(float)                  xmask 2^63,   fmask ~2^63   (order-preserving bits)
//   MIN(float)                  xmask ~2^63,  fmask ~2^63
__host__ __device__ constexpr uint64_t op_xmask(int32_t code)
{
    return code == OPC_MIN_U ? ~0ull
         : (code == OPC_MAX_I || code == OPC_MAX_F) ? 0x8000000000000000ull
         : (code == OPC_MIN_I || code == OPC_MIN_F) ? 0x7fffffffffffffffull
         : 0ull;
}

__host__ __device__ constexpr uint64_t op_fmask(int32_t code)
{
    return (code == OPC_MAX_F || code == OPC_MIN_F) ? 0x7fffffffffffffffull : 0ull;
}

struct OpDesc {
    int32_t code;       // OpCode
    int32_t col;        // index into AggSpec::cols, -1 for COUNT
    uint64_t xmask;     // op_xmask(code)
    uint64_t fmask;     // op_fmask(code)
};

// Ops are ordered by class: ncls[0] add-u64, then ncls[1] add-f64, then
// ncls[2] max-u64, then ncls[3] exchange.
struct AggSpec {
    int32_t nvcols;
    int32_t nops;
    int32_t ncls[4];
    ColDesc cols[kMaxVCols];
    OpDesc ops[kMaxOps];
};

// Output finalisation codes.
enum OutCode : int32_t {
    OUT_U64 = 0,      // state bits as a 64-bit integer (COUNT, integer SUM)
    OUT_F32 = 1,      // double state rounded to float (float SUM)
    OUT_F64 = 2,      // double state (double SUM)
    OUT_MAX_I = 3,    // decode ord -> int64 -> out width
    OUT_MIN_I = 4,
    OUT_MAX_U = 5,
    OUT_MIN_U = 6,
    OUT_MAX_F = 7,    // decode ord -> double -> out float/double
    OUT_MIN_F = 8,
    OUT_ANY = 9,      // raw bits, out width
    OUT_AVG_I = 10,   // (double)(int64)sum / (double)count
    OUT_AVG_U = 11,   // (double)(uint64)sum / (double)count
    OUT_AVG_F = 12    // fp64 sum / (double)count
};

struct OutDesc {
    void* data;         // device column, num_groups rows
    int32_t code;       // OutCode
    int32_t bytes;      // output element size
    int32_t op;         // state read
    int32_t op2;        // count state (AVG only)
    int32_t is_float;   // output column is FLOAT32/FLOAT64
    int32_t pad;
};

struct OutSpec {
    void* key_data;     // output key column
    int32_t key_bytes;
    int32_t nouts;
    OutDesc outs[kMaxOuts];
};

// Key column descriptor.
struct KeyDesc {
    const void* data;
    int32_t bytes;      // 4 or 8
    int32_t is_signed;
};

}  // namespace eg

// eg_radix.hpp -- the radix engine shared by the sort and partition strategies
// and by the sort/partition sub-operators.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "eg_internal.hpp"

namespace eg {

constexpr int kMaxPay = 9;  // payload columns carried by one radix pass

// Where a pass takes its digit from.
enum DigitMode : int32_t {
    DIG_KEY = 0,   // radix-ordered key bits: twiddle(key) - kmin (sign flip / float order)
    DIG_HASH = 1   // bit_mix(key) bits (fmix32 for <= 4-byte keys, fmix64 otherwise)
};

struct RadixKeys {
    const void* data;
    int32_t bytes;      // 1, 2, 4 or 8
    int32_t type;       // gamax type id (twiddle rule)
    int32_t mode;       // DigitMode
    int32_t desc;       // descending digits
    uint64_t kmin;      // subtracted from the twiddled key (DIG_KEY)
};

struct Payload {
    const void* in;     // nullptr + rowid: generate the row index
    void* out;
    int32_t bytes;
    int32_t rowid;      // 1: generated int64 row ids (argsort / GFUR)
};

// One stable counting pass on digit (key >> shift) & (2^bits - 1): keys and
// payloads move from in to out.  bits <= 8.  n < 2^32.
void radix_pass(const RadixKeys& k, void* keys_out, const std::vector<Payload>& pay, int64_t n, int shift,
                int bits, hipStream_t s);

// Stable LSD sort over digit bits [begin_bit, end_bit) in balanced passes of
// at most 6 bits.
// Returns the buffers holding the result (the last pass writes the *_out
// buffers when the pass count is odd, the *_alt buffers otherwise; the
// first pass reads the inputs).
struct SortBuffers {
    const void* keys_in;
    void* keys_a;
    void* keys_b;
    std::vector<Payload> pay_in;   // out fields unused
    std::vector<void*> pay_a, pay_b;
};
// After the call, *keys_res / pay_res point at the sorted data (a or b).
void radix_sort(const RadixKeys& k, SortBuffers& b, int64_t n, int begin_bit, int end_bit, hipStream_t s,
                const void** keys_res, std::vector<const void*>& pay_res);

// Exclusive prefix sum of n uint32 values; returns the total.
uint64_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, bool want_total);

// Radix-ordered (twiddled) bits of a key, device and host.  -0.0 sorts as
// +0.0 (cub::DeviceRadixSort treats the two zeros as equal keys).
__host__ __device__ inline uint64_t twiddle_bits(uint64_t raw, int32_t type, int bytes)
{
    const int w = 8 * bytes;
    const uint64_t sign = 1ull << (w - 1);
    const uint64_t all = (w == 64) ? ~0ull : ((1ull << w) - 1);
    if (type == EG_INT8 || type == EG_INT16 || type == EG_INT32 || type == EG_INT64) return raw ^ sign;
    if (type == EG_FLOAT32 || type == EG_FLOAT64) {
        if (raw == sign) raw = 0;
        return (raw & sign) ? (~raw & all) : (raw ^ sign);
    }
    return raw;
}

// Min and max of the twiddled keys (device reduction, synchronises s).
void key_range(const void* keys, int32_t type, int bytes, int64_t n, hipStream_t s, uint64_t* mn, uint64_t* mx);

}  // namespace eg

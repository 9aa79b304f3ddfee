// gen.hip -- counter-based synthetic data on the device.
//
// Semantics follow the reference's generators: uniform keys are "i % G then
// shuffle" (benchmarks/groupby/groupby_util.cuh:99-109) and Zipf keys are
// inverse-CDF draws with weights 1/(r+1)^s (data_gen/generator.cuh:360-484).
// Randomness comes from Philox4x32-10 (Salmon et al., SC'11) keyed by the seed,
// with the row index as counter, so any row range can be generated on its own
// and a host restatement produces the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "eg_device.hpp"
#include "eg_internal.hpp"

namespace eg {
namespace {

constexpr uint32_t kStreamPerm = 0x7065726du;
constexpr uint32_t kStreamZipf = 1u;

__host__ __device__ inline void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
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

// Keyed bijection on [0, 2^k) (xor, odd multiply, xor-shift-right; each is a
// bijection of k-bit words), cycle-walked into [0, n).
struct Perm {
    uint64_t n, mask, a1, a2, x0, x1;
    int s;
};

Perm make_perm(uint64_t n, uint64_t seed)
{
    int k = 1;
    while (k < 64 && (1ull << k) < n) ++k;
    uint32_t o0[4], o1[4];
    philox(0u, 0u, kStreamPerm, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o0);
    philox(1u, 0u, kStreamPerm, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o1);
    Perm p;
    p.n = n;
    p.mask = (k == 64) ? ~0ull : ((1ull << k) - 1);
    p.a1 = (((uint64_t)o0[1] << 32) | o0[0]) | 1u;
    p.a2 = (((uint64_t)o0[3] << 32) | o0[2]) | 1u;
    p.x0 = ((uint64_t)o1[1] << 32) | o1[0];
    p.x1 = ((uint64_t)o1[3] << 32) | o1[2];
    p.s = (k + 1) / 2;
    return p;
}

__device__ inline uint64_t perm_round(const Perm& p, uint64_t x)
{
    x ^= p.x0 & p.mask;
    x = (x * p.a1) & p.mask;
    x ^= x >> p.s;
    x = (x * p.a2) & p.mask;
    x ^= x >> p.s;
    x ^= p.x1 & p.mask;
    x = (x * p.a1) & p.mask;
    x ^= x >> p.s;
    return x;
}

__device__ inline void store_num(void* out, int32_t type, int64_t i, int64_t v)
{
    if (type == dev::T_FLOAT32) ((float*)out)[i] = (float)v;
    else if (type == dev::T_FLOAT64) ((double*)out)[i] = (double)v;
    else dev::store_bytes(out, dev::type_bytes(type), i, (uint64_t)v);
}

__global__ void gen_uniform_kernel(void* out, int32_t type, int64_t begin, int64_t count,
                                   uint64_t distinct, Perm p)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        // Cycle walking ends: begin+i < n lies on a cycle of the bijection
        // that contains a point below n (itself).
        uint64_t y = perm_round(p, (uint64_t)(begin + i));
        while (y >= p.n) y = perm_round(p, y);
        store_num(out, type, i, (int64_t)(y % distinct));
    }
}

__device__ inline double u53(const uint32_t o[4])
{
    uint64_t hi = o[0] >> 5, lo = o[1] >> 6;  // 27 + 26 bits
    return (double)((hi << 26) | lo) * (1.0 / 9007199254740992.0);
}

__global__ void gen_zipf_kernel(void* out, int32_t type, int64_t begin, int64_t count,
                                const double* __restrict__ cdf, int64_t distinct, uint32_t k0,
                                uint32_t k1)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint64_t row = (uint64_t)(begin + i);
        uint32_t o[4];
        philox((uint32_t)row, (uint32_t)(row >> 32), kStreamZipf, 0u, k0, k1, o);
        double u = u53(o);
        int64_t lo = 0, hi = distinct - 1;
        while (lo < hi) {
            int64_t mid = lo + ((hi - lo) >> 1);
            if (cdf[mid] > u) hi = mid;
            else lo = mid + 1;
        }
        store_num(out, type, i, lo);
    }
}

__global__ void gen_values_kernel(void* out, int32_t type, int32_t dist, int64_t begin,
                                  int64_t count, uint32_t k0, uint32_t k1, uint32_t stream_id)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint64_t row = (uint64_t)(begin + i);
        uint32_t o[4] = {0u, 0u, 0u, 0u};
        if (dist <= EG_VAL_UNIFORM_I32_1000)
            philox((uint32_t)row, (uint32_t)(row >> 32), stream_id, 0u, k0, k1, o);
        switch (dist) {
        case EG_VAL_UNIFORM_I64: store_num(out, type, i, (int64_t)o[0] - 2147483648LL); break;
        case EG_VAL_UNIFORM_F32:
            if (type == dev::T_FLOAT32) ((float*)out)[i] = (float)(o[0] >> 8) * (1.0f / 16777216.0f);
            else ((double*)out)[i] = (double)(o[0] >> 8) * (1.0 / 16777216.0);
            break;
        case EG_VAL_UNIFORM_I32_1000:
            store_num(out, type, i, (int64_t)(((uint64_t)o[0] * 1000u) >> 32));
            break;
        case EG_VAL_ALL_ONES: store_num(out, type, i, 1); break;
        default: store_num(out, type, i, (int64_t)row); break;
        }
    }
}

unsigned grid_for(int64_t count)
{
    int64_t cus = device_info().num_cus;
    int64_t want = (count + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, cus * 16));
}

void check_type(int32_t type, bool allow_float)
{
    if (!(is_int(type) || (allow_float && is_float(type))))
        raise(EG_ERR_UNSUPPORTED, "generator: unsupported output type");
}

}  // namespace

void gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count, int64_t n,
                      int64_t distinct, uint64_t seed, hipStream_t s)
{
    check_type(type, true);
    if (n <= 0 || distinct <= 0 || begin < 0 || count < 0 || begin + count > n)
        raise(EG_ERR_INVALID, "gen_keys_uniform: bad row range or distinct count");
    if (count == 0) return;
    Perm p = make_perm((uint64_t)n, seed);
    hipLaunchKernelGGL(gen_uniform_kernel, dim3(grid_for(count)), dim3(256), 0, s, out, type, begin,
                       count, (uint64_t)distinct, p);
    EG_LAUNCH_CHECK();
}

void gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count, const double* d_cdf,
                   int64_t distinct, uint64_t seed, hipStream_t s)
{
    check_type(type, true);
    if (distinct <= 0 || begin < 0 || count < 0 || !d_cdf)
        raise(EG_ERR_INVALID, "gen_keys_zipf: bad arguments");
    if (count == 0) return;
    hipLaunchKernelGGL(gen_zipf_kernel, dim3(grid_for(count)), dim3(256), 0, s, out, type, begin,
                       count, d_cdf, distinct, (uint32_t)seed, (uint32_t)(seed >> 32));
    EG_LAUNCH_CHECK();
}

void gen_values(void* out, int32_t type, int32_t dist, int64_t begin, int64_t count,
                uint64_t seed, uint32_t stream_id, hipStream_t s)
{
    if (dist == EG_VAL_UNIFORM_F32) {
        if (!is_float(type)) raise(EG_ERR_UNSUPPORTED, "gen_values: float distribution needs a float type");
    } else {
        check_type(type, true);
    }
    if (dist < EG_VAL_UNIFORM_I64 || dist > EG_VAL_SEQ) raise(EG_ERR_INVALID, "gen_values: unknown distribution");
    if (begin < 0 || count < 0) raise(EG_ERR_INVALID, "gen_values: bad row range");
    if (count == 0) return;
    hipLaunchKernelGGL(gen_values_kernel, dim3(grid_for(count)), dim3(256), 0, s, out, type, dist,
                       begin, count, (uint32_t)seed, (uint32_t)(seed >> 32), stream_id);
    EG_LAUNCH_CHECK();
}

// Host Zipf CDF: sequential fp64 prefix of 1/(r+1)^s, normalised, last = 1.0
// (data_gen/generator.cuh:391-484).
void zipf_cdf(double* cdf, int64_t distinct, double s)
{
    if (!cdf || distinct <= 0) raise(EG_ERR_INVALID, "zipf_cdf: bad arguments");
    double acc = 0.0;
    for (int64_t r = 0; r < distinct; ++r) {
        acc += 1.0 / std::pow((double)(r + 1), s);
        cdf[r] = acc;
    }
    for (int64_t r = 0; r < distinct; ++r) cdf[r] = cdf[r] / acc;
    cdf[distinct - 1] = 1.0;
}

}  // namespace eg

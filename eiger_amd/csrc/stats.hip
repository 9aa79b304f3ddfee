// stats.hip -- distinct-count estimate and the rule-based strategy choice.
//
// compute_ndv_approx (stats/stats_compute.cu:103-115,186-196) is a cuco
// HyperLogLog; this is a HyperLogLog of our own (Flajolet et al. 2007, with
// the linear-counting small-range correction), 2^14 registers, one read of the
// column.  Each workgroup keeps its registers in LDS (32-bit, so LDS atomicMax
// applies), writes them out, and a second kernel takes the per-register max.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "eg_device.hpp"
#include "eg_internal.hpp"

namespace eg {
namespace {

constexpr int kP = 14;
constexpr int kM = 1 << kP;

// One register update: the register is read first, and the LDS atomic is
// issued only when it would raise it (after the first rows most do not;
// registers only grow, so a stale read never skips a needed update).
__device__ __forceinline__ void hll_add(uint32_t* reg, uint64_t raw)
{
    const uint64_t h = dev::fmix64(raw ^ 0x9e3779b97f4a7c15ull);
    const uint32_t idx = (uint32_t)(h >> (64 - kP));
    const uint64_t w = (h << kP) | (1ull << (kP - 1));
    const uint32_t rho = (uint32_t)__clzll((long long)w) + 1u;
    if (__hip_atomic_load(&reg[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < rho) atomicMax(&reg[idx], rho);
}

// KB = 4 or 8: 16-byte vector loads when the column is 16-byte aligned;
// KB = 0: any width, one typed load per row.
template <int KB>
__global__ __launch_bounds__(1024) void hll_kernel(const void* data, int bytes, int64_t n,
                                                   uint8_t* __restrict__ partial)
{
    __shared__ uint32_t reg[kM];
    for (int i = threadIdx.x; i < kM; i += blockDim.x) reg[i] = 0u;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t done = 0;
    if constexpr (KB == 4 || KB == 8) {
        constexpr int kPer = 16 / KB;
        const int64_t nvec = (((uintptr_t)data & 15u) == 0u) ? n / kPer : 0;
        const uint4* v = static_cast<const uint4*>(data);
        for (int64_t i = gid; i < nvec; i += stride) {
            const uint4 x = v[i];
            if constexpr (KB == 4) {
                hll_add(reg, x.x);
                hll_add(reg, x.y);
                hll_add(reg, x.z);
                hll_add(reg, x.w);
            } else {
                hll_add(reg, (uint64_t)x.x | ((uint64_t)x.y << 32));
                hll_add(reg, (uint64_t)x.z | ((uint64_t)x.w << 32));
            }
        }
        done = nvec * kPer;
    }
    for (int64_t i = done + gid; i < n; i += stride) hll_add(reg, dev::load_raw(data, bytes, i));
    __syncthreads();
    uint8_t* out = partial + (size_t)blockIdx.x * kM;
    for (int i = threadIdx.x; i < kM; i += blockDim.x) out[i] = (uint8_t)reg[i];
}

void launch_hll(const void* data, int bytes, int64_t n, int parts, uint8_t* partial, hipStream_t s)
{
    if (bytes == 4) hipLaunchKernelGGL(hll_kernel<4>, dim3(parts), dim3(1024), 0, s, data, bytes, n, partial);
    else if (bytes == 8) hipLaunchKernelGGL(hll_kernel<8>, dim3(parts), dim3(1024), 0, s, data, bytes, n, partial);
    else hipLaunchKernelGGL(hll_kernel<0>, dim3(parts), dim3(1024), 0, s, data, bytes, n, partial);
    EG_LAUNCH_CHECK();
}

// Workgroups of the register pass: two per CU (64 KB of registers each),
// fewer for short columns.
int hll_parts(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 16383) / 16384, (int64_t)device_info().num_cus * 2));
}

// Per-register maximum over the partial register sets.  With acc set, also
// the estimator's inputs: acc[0] += sum of 2^(40 - r) (the harmonic sum in
// units of 2^-40: exact, as the host's double sum is, while r <= 39, i.e.
// for any column below ~2^50 distinct keys) and acc[1] += the zero registers.
__global__ __launch_bounds__(256) void hll_merge_kernel(const uint8_t* __restrict__ partial, int parts,
                                                        uint8_t* __restrict__ out, unsigned long long* acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // the grid covers kM exactly
    uint8_t m = 0;
    for (int p = 0; p < parts; ++p) m = std::max(m, partial[(size_t)p * kM + i]);
    out[i] = m;
    if (!acc) return;
    unsigned long long sum = m <= 40 ? 1ull << (40 - m) : 0ull, zeros = m == 0 ? 1ull : 0ull;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum += __shfl_xor(sum, off);
        zeros += __shfl_xor(zeros, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[0], sum);
        atomicAdd(&acc[1], zeros);
    }
}

}  // namespace

namespace {
double hll_estimate_from(double sum, int64_t zeros)
{
    const double alpha = 0.7213 / (1.0 + 1.079 / kM);
    double e = alpha * (double)kM * (double)kM / sum;
    if (e <= 2.5 * kM && zeros > 0) e = (double)kM * std::log((double)kM / (double)zeros);
    return e;
}

double hll_estimate(const uint8_t* reg)
{
    double sum = 0.0;
    int zeros = 0;
    for (int i = 0; i < kM; ++i) {
        sum += std::ldexp(1.0, -(int)reg[i]);
        zeros += reg[i] == 0;
    }
    const double alpha = 0.7213 / (1.0 + 1.079 / kM);
    double e = alpha * (double)kM * (double)kM / sum;
    if (e <= 2.5 * kM && zeros > 0) e = (double)kM * std::log((double)kM / (double)zeros);
    return e;
}

// Registers of one column into regs (kM bytes, device); asynchronous.
void hll_registers(const void* data, int bytes, int64_t n, uint8_t* regs, DevBuf& partial, hipStream_t s)
{
    const int parts = hll_parts(n);
    partial = DevBuf((size_t)parts * kM, s);
    KernelTimer timer("hll", s);
    launch_hll(data, bytes, n, parts, partial.get<uint8_t>(), s);
    hipLaunchKernelGGL(hll_merge_kernel, dim3(kM / 256), dim3(256), 0, s, (const uint8_t*)partial.get<uint8_t>(), parts,
                       regs, (unsigned long long*)nullptr);
    EG_LAUNCH_CHECK();
}
}  // namespace

int64_t ndv_approx(const void* data, int32_t type, int64_t n, hipStream_t s)
{
    const int bytes = type_size(type);
    if (!bytes) raise(EG_ERR_UNSUPPORTED, "compute_ndv_approx: unsupported column type");
    if (n <= 0) return 0;
    const int parts = hll_parts(n);
    // partial registers, the merged registers, then the two estimator sums
    DevBuf partial((size_t)parts * kM + kM + 16, s);
    uint8_t* pp = partial.get<uint8_t>();
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(pp + (size_t)parts * kM + kM);
    EG_HIP(hipMemsetAsync(acc, 0, 16, s));
    {
        KernelTimer timer("hll", s);
        launch_hll(data, bytes, n, parts, pp, s);
        hipLaunchKernelGGL(hll_merge_kernel, dim3(kM / 256), dim3(256), 0, s, (const uint8_t*)pp, parts,
                           pp + (size_t)parts * kM, acc);
        EG_LAUNCH_CHECK();
    }
    // 16 bytes back instead of the 16 KB of registers
    int64_t* h = pinned_i64(2);
    EG_HIP(hipMemcpyAsync(h, acc, 16, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    const double sum = std::ldexp((double)(uint64_t)h[0], -40);
    int64_t r = (int64_t)std::llround(hll_estimate_from(sum, h[1]));
    return std::max<int64_t>(1, std::min<int64_t>(r, n));
}

}  // namespace eg

extern "C" {

// The rule for many key columns with one synchronisation: every column's
// HyperLogLog runs asynchronously, the registers come back in one copy.
int eg_auto_select_algo_batch(const eg_column_view* keys, int64_t num, int32_t* algos, int64_t* ndv, void* stream)
{
    return eg::guard([&] {
        if (num < 0 || (num > 0 && (!keys || !algos))) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (num == 0) return;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        eg::DevBuf regs((size_t)num * eg::kM, s);
        std::vector<eg::DevBuf> partials((size_t)num);
        for (int64_t i = 0; i < num; ++i) {
            const eg_column_view& c = keys[i];
            const int bytes = eg::type_size(c.type);
            if (!bytes || c.type == EG_STRING) eg::raise(EG_ERR_UNSUPPORTED, "unsupported key type");
            if (c.size > 0 && !c.data) eg::raise(EG_ERR_INVALID, "null column data");
            uint8_t* r = regs.get<uint8_t>() + (size_t)i * eg::kM;
            if (c.size <= 0) {
                eg::fill_zero(r, eg::kM, s);
                continue;
            }
            eg::hll_registers((const char*)c.data + c.offset * bytes, bytes, c.size, r, partials[(size_t)i], s);
        }
        std::vector<uint8_t> h((size_t)num * eg::kM);
        EG_HIP(hipMemcpyAsync(h.data(), regs.get(), h.size(), hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < num; ++i) {
            const int64_t n = keys[i].size;
            const int64_t est = n <= 0 ? 0
                : std::max<int64_t>(1, std::min<int64_t>((int64_t)std::llround(eg::hll_estimate(h.data() + (size_t)i * eg::kM)), n));
            if (ndv) ndv[i] = est;
            const int64_t thr = eg::type_size(keys[i].type) == 4 ? ((int64_t)1 << 14) : ((int64_t)1 << 18);
            algos[i] = est <= thr ? EG_ALGO_HASH : EG_ALGO_PARTITION;
        }
    });
}

int eg_compute_ndv_approx(eg_column_view col, void* stream, int64_t* out_ndv)
{
    return eg::guard([&] {
        if (!out_ndv) eg::raise(EG_ERR_INVALID, "out_ndv is null");
        const int bytes = eg::type_size(col.type);
        if (!bytes) eg::raise(EG_ERR_UNSUPPORTED, "compute_ndv_approx: unsupported column type");
        if (col.size > 0 && !col.data) eg::raise(EG_ERR_INVALID, "null column data");
        const void* d = col.data ? (const char*)col.data + col.offset * bytes : nullptr;
        *out_ndv = eg::ndv_approx(d, col.type, col.size, reinterpret_cast<hipStream_t>(stream));
    });
}

// auto_select_groupby_algo (operators/groupby/groupby.cu:68-119).
int eg_auto_select_algo(const eg_column_view* keys, int64_t num_keys, eg_groupby_hint* hint, void* stream)
{
    return eg::guard([&] {
        if (!hint) eg::raise(EG_ERR_INVALID, "hint is null");
        if (hint->groupby_algo != EG_ALGO_AUTO) return;
        if (num_keys <= 0 || !keys) eg::raise(EG_ERR_INVALID, "no key columns");
        for (int64_t i = 0; i < num_keys; ++i)
            if (keys[i].type == EG_STRING) {
                hint->groupby_algo = EG_ALGO_HASH;
                return;
            }
        if (num_keys != 1) {
            hint->groupby_algo = EG_ALGO_HASH;
            return;
        }
        int64_t est = 0;
        int rc = eg_compute_ndv_approx(keys[0], stream, &est);
        if (rc != EG_OK) eg::raise(rc, eg_last_error());
        const bool four = eg::type_size(keys[0].type) == 4;
        const int64_t thr = four ? ((int64_t)1 << 14) : ((int64_t)1 << 18);
        hint->groupby_algo = est <= thr ? EG_ALGO_HASH : EG_ALGO_PARTITION;
    });
}

}  // extern "C"

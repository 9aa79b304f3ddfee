// radix.hip -- stable radix passes, LSD sort and the sort/partition
// sub-operators (gamax::sort, gamax::partition).
//
// The reference delegates these to cub::DeviceRadixSort / DeviceHistogram /
// DeviceScan (operators/sort/sort.cu:15-66, operators/partition/partition.cu:
// 24-171).  Here one pass over up to 8 digit bits is:
//   radix_hist_kernel    per 4096-row tile, an LDS histogram of the digit;
//   offsets              digit-major exclusive prefix of the tile-major
//                        [tile][digit] counts (coalesced rows, three kernels);
//   radix_scatter_kernel per tile, a stable rank of every row among the
//                        tile's rows of the same digit (each wave ranks its own
//                        1024 consecutive rows with wave64 ballots on the digit
//                        bits and running per-wave counters, then one prefix
//                        over waves and digits), the tile rearranged in LDS so
//                        rows of one digit are contiguous, then written out in
//                        runs (coalesced stores).  Payload columns follow the
//                        same permutation, one column at a time through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "eg_device.hpp"
#include "eg_internal.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

constexpr int kRB = 256;        // threads per block
// Digit width of a sort pass: fan-out 64 keeps each digit's rows in runs long
// enough for full-line stores (a 6-bit pass on 2^28 rows takes about half the
// time of an 8-bit one), so a 32-bit key takes six passes instead of four.
constexpr int kSortDigitBits = 6;
constexpr int kRI = 16;         // rows per thread
constexpr int kRT = kRB * kRI;  // rows per tile

struct PayArgs {
    int32_t n;
    const void* in[kMaxPay];
    void* out[kMaxPay];
    int32_t bytes[kMaxPay];
    int32_t rowid[kMaxPay];
};

__device__ __forceinline__ uint32_t digit_of(const RadixKeys& k, uint64_t raw, int shift, uint32_t mask)
{
    uint64_t v;
    if (k.mode == DIG_HASH) v = (k.bytes <= 4) ? (uint64_t)dev::fmix32((uint32_t)raw) : dev::fmix64(raw);
    else v = twiddle_bits(raw, k.type, k.bytes) - k.kmin;
    const uint32_t d = (uint32_t)(v >> shift) & mask;
    return k.desc ? mask - d : d;
}

template <int KB>
__device__ __forceinline__ uint64_t load_key(const void* p, int64_t i)
{
    if constexpr (KB == 1) return ((const uint8_t*)p)[i];
    else if constexpr (KB == 2) return ((const uint16_t*)p)[i];
    else if constexpr (KB == 4) return ((const uint32_t*)p)[i];
    else return ((const uint64_t*)p)[i];
}

template <int KB>
__device__ __forceinline__ void store_key(void* p, int64_t i, uint64_t v)
{
    if constexpr (KB == 1) ((uint8_t*)p)[i] = (uint8_t)v;
    else if constexpr (KB == 2) ((uint16_t*)p)[i] = (uint16_t)v;
    else if constexpr (KB == 4) ((uint32_t*)p)[i] = (uint32_t)v;
    else ((uint64_t*)p)[i] = v;
}

// Key widths are template parameters so the unrolled loads of a thread are
// independent (a runtime width switch around each load serialises them).
template <int KB>
__global__ __launch_bounds__(kRB) void radix_hist_kernel(RadixKeys k, int64_t n, int shift, uint32_t mask,
                                                          uint32_t* __restrict__ hist, int ntiles)
{
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRT;
    uint64_t raw[kRI];
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int64_t e = base + (int64_t)i * kRB + t;
        raw[i] = e < n ? load_key<KB>(k.data, e) : 0ull;
    }
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int64_t e = base + (int64_t)i * kRB + t;
        if (e < n) atomicAdd(&h[digit_of(k, raw[i], shift, mask)], 1u);
    }
    __syncthreads();
    // [tile][digit], the 2^bits digits in use: one contiguous row per tile
    if ((uint32_t)t <= mask) hist[(int64_t)blockIdx.x * (mask + 1) + t] = h[t];
}

using dev::block_excl_scan;

__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* wsum) { return block_excl_scan<256>(v, wsum); }

// A column of the tile leaves in digit order in three steps: its values are
// loaded into registers (load_column), put at their tile position in LDS
// (stage_column), and written out in runs of one digit (store_column,
// coalesced stores).  The next column's loads are issued before the current
// column's stores, so their latency overlaps.
template <int B>
__device__ __forceinline__ void load_column_w(const void* src, bool rid, const uint32_t (&pos)[kRI], int64_t wbase,
                                              int lane, uint64_t (&v)[kRI])
{
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int64_t e = wbase + (int64_t)i * 64 + lane;
        v[i] = rid ? (uint64_t)e : (pos[i] != 0xffffffffu ? load_key<B>(src, e) : 0ull);
    }
}

__device__ __forceinline__ void load_column(const void* src, int bytes, bool rid, const uint32_t (&pos)[kRI],
                                            int64_t wbase, int lane, uint64_t (&v)[kRI])
{
    switch (rid ? 8 : bytes) {  // uniform; the loads of each case are independent
    case 1: load_column_w<1>(src, false, pos, wbase, lane, v); break;
    case 2: load_column_w<2>(src, false, pos, wbase, lane, v); break;
    case 4: load_column_w<4>(src, false, pos, wbase, lane, v); break;
    default: load_column_w<8>(src, rid, pos, wbase, lane, v); break;
    }
}

__device__ __forceinline__ void stage_column(const uint64_t (&v)[kRI], const uint32_t (&pos)[kRI],
                                             unsigned long long* stage)
{
#pragma unroll
    for (int i = 0; i < kRI; ++i)
        if (pos[i] != 0xffffffffu) stage[pos[i]] = v[i];
}

template <int B>
__device__ __forceinline__ void store_column_w(void* dst, int t, int32_t tcount, const unsigned long long* stage,
                                               const uint8_t* sdig, const uint32_t* gofs, const uint32_t* lstart)
{
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int p = i * kRB + t;
        if (p < tcount) {
            const uint32_t d = sdig[p];
            store_key<B>(dst, (int64_t)gofs[d] + (p - (int64_t)lstart[d]), stage[p]);
        }
    }
}

__device__ __forceinline__ void store_column(void* dst, int bytes, int t, int32_t tcount,
                                             const unsigned long long* stage, const uint8_t* sdig,
                                             const uint32_t* gofs, const uint32_t* lstart)
{
    switch (bytes) {
    case 1: store_column_w<1>(dst, t, tcount, stage, sdig, gofs, lstart); break;
    case 2: store_column_w<2>(dst, t, tcount, stage, sdig, gofs, lstart); break;
    case 4: store_column_w<4>(dst, t, tcount, stage, sdig, gofs, lstart); break;
    default: store_column_w<8>(dst, t, tcount, stage, sdig, gofs, lstart); break;
    }
}

// Stable multisplit of one 4096-row tile.  Wave w owns rows
// [w*1024, (w+1)*1024) of the tile and walks them 64 at a time: the lanes
// with the same digit are found with one ballot per digit bit, the lowest of
// them advances the wave's running count of that digit in LDS.  No barrier is
// needed inside the walk (each wave updates only its own counters); one scan
// over (wave, digit) then gives every row its place in the tile.
template <int KB>
__global__ __launch_bounds__(kRB) void radix_scatter_kernel(RadixKeys k, void* __restrict__ kout, PayArgs pay,
                                                             int64_t n, int shift, int bits,
                                                             const uint32_t* __restrict__ offs, int ntiles)
{
    constexpr int kW = kRB / 64;   // waves
    constexpr int kWR = kRT / kW;  // rows per wave
    static_assert(kRB == 256 && kWR == kRI * 64, "one digit per thread; kRI rows per lane");
    __shared__ unsigned long long stage[kRT];
    __shared__ uint8_t sdig[kRT];
    __shared__ uint32_t whist[kW][256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t wsum[kW];

    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint32_t mask = (1u << bits) - 1u;
    // tile order per XCD: the runs of one digit from neighbouring tiles are
    // adjacent in the output and share their boundary cache lines
    const uint32_t tile = dev::xcd_remap(blockIdx.x, (uint32_t)ntiles);
    const int64_t tbase = (int64_t)tile * kRT;
#pragma unroll
    for (int w = 0; w < kW; ++w) whist[w][t] = 0u;
    gofs[t] = t < (1 << bits) ? offs[(int64_t)tile * (1 << bits) + t] : 0u;

    const int64_t wbase = tbase + (int64_t)wave * kWR;
    uint64_t kr[kRI];
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int64_t e = wbase + (int64_t)i * 64 + lane;
        kr[i] = e < n ? load_key<KB>(k.data, e) : 0ull;
    }
    __syncthreads();

    uint32_t dg[kRI], rk[kRI];
    const uint64_t lt = dev::lanemask_lt();
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        const int64_t e = wbase + (int64_t)i * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = valid ? digit_of(k, kr[i], shift, mask) : 0u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before =
            valid ? __hip_atomic_load(&whist[wave][d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0u;
        const uint32_t r = (uint32_t)__popcll(peers & lt);
        // every lane of the wave has read the counter before the leader writes it
        __builtin_amdgcn_wave_barrier();
        if (valid && r == 0u)
            __hip_atomic_store(&whist[wave][d], before + (uint32_t)__popcll(peers), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        __builtin_amdgcn_wave_barrier();
        dg[i] = d;
        rk[i] = valid ? before + r : 0xffffffffu;
    }
    __syncthreads();
    // per digit: exclusive prefix over the waves, and the digit's tile count
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kW; ++w) {
        const uint32_t c = whist[w][t];
        whist[w][t] = run;
        run += c;
    }
    const uint32_t ls = block_excl_scan256(run, wsum);
    lstart[t] = ls;
    __syncthreads();
    const int32_t tcount = (int32_t)min<int64_t>((int64_t)kRT, n - tbase);
    uint32_t pos[kRI];
#pragma unroll
    for (int i = 0; i < kRI; ++i) {
        pos[i] = 0xffffffffu;
        if (rk[i] != 0xffffffffu) {
            pos[i] = lstart[dg[i]] + whist[wave][dg[i]] + rk[i];
            sdig[pos[i]] = (uint8_t)dg[i];
        }
    }

    // keys (from registers), then every payload column, through the LDS stage;
    // column c+1 is being loaded while column c is stored
    uint64_t v[kRI];
    stage_column(kr, pos, stage);
    __syncthreads();
    if (pay.n > 0) load_column(pay.in[0], pay.bytes[0], pay.rowid[0] != 0, pos, wbase, lane, v);
    store_column(kout, KB, t, tcount, stage, sdig, gofs, lstart);
    for (int c = 0; c < pay.n; ++c) {
        __syncthreads();  // the stage is free again
        stage_column(v, pos, stage);
        __syncthreads();
        if (c + 1 < pay.n) load_column(pay.in[c + 1], pay.bytes[c + 1], pay.rowid[c + 1] != 0, pos, wbase, lane, v);
        store_column(pay.out[c], pay.rowid[c] != 0 ? 8 : pay.bytes[c], t, tcount, stage, sdig, gofs, lstart);
    }
}

// ---- tile offsets of a radix pass -----------------------------------------------
// counts[t][d] (tile-major, nd digits) -> offs[t][d] = rows of smaller digits
// + rows of digit d in tiles before t.  Tiles are taken in chunks of `per`;
// every access is a contiguous row of nd words (a digit per thread):
//   1. chunk sums csum[c][d];
//   2. one workgroup per digit scans its chunk sums (coff[c][d]) and writes
//      the digit's total;
//   3. one workgroup per chunk adds the digit's base (a scan of the totals)
//      and walks its tiles.
constexpr int kOB = 256;  // threads of the offset kernels (>= digits)

__global__ __launch_bounds__(kOB) void tile_chunk_sums_kernel(const uint32_t* __restrict__ counts, int ntiles, int nd,
                                                              int per, uint32_t* __restrict__ csum)
{
    const int d = threadIdx.x;
    if (d >= nd) return;
    const int t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
    uint32_t acc = 0;
#pragma unroll 8
    for (int t = t0; t < t1; ++t) acc += counts[(int64_t)t * nd + d];
    csum[(int64_t)blockIdx.x * nd + d] = acc;
}

// one workgroup (1024 threads) per digit; nchunks <= 4 * 1024
__global__ __launch_bounds__(1024) void tile_chunk_scan_kernel(const uint32_t* __restrict__ csum, int nchunks, int nd,
                                                               uint32_t* __restrict__ coff, uint32_t* __restrict__ total)
{
    __shared__ uint32_t wsum[16];
    const int d = blockIdx.x;
    uint32_t v[4], run = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = threadIdx.x * 4 + i;
        v[i] = c < nchunks ? csum[(int64_t)c * nd + d] : 0u;
        run += v[i];
    }
    uint32_t ex = block_excl_scan<1024>(run, wsum);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = threadIdx.x * 4 + i;
        if (c < nchunks) coff[(int64_t)c * nd + d] = ex;
        ex += v[i];
    }
    if (threadIdx.x == 1023) total[d] = ex;
}

__global__ __launch_bounds__(kOB) void tile_offsets_kernel(const uint32_t* __restrict__ counts, int ntiles, int nd,
                                                           int per, const uint32_t* __restrict__ coff,
                                                           const uint32_t* __restrict__ total, uint32_t* __restrict__ offs)
{
    __shared__ uint32_t wsum[kOB / 64];
    const int d = threadIdx.x;
    const uint32_t base = block_excl_scan<kOB>(d < nd ? total[d] : 0u, wsum);
    if (d >= nd) return;
    const int t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
    uint32_t run = base + coff[(int64_t)blockIdx.x * nd + d];
    for (int t = t0; t < t1; ++t) {
        const uint32_t c = counts[(int64_t)t * nd + d];
        offs[(int64_t)t * nd + d] = run;
        run += c;
    }
}

// ---- device-wide exclusive scan of uint32 (reduce, scan the block sums, add) ----
constexpr int kSB = 256, kSI = 16, kST = kSB * kSI;

__global__ __launch_bounds__(kSB) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                          uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[4];
    const int64_t base = (int64_t)blockIdx.x * kST;
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < kSI; ++i) {
        const int64_t e = base + (int64_t)i * kSB + threadIdx.x;
        if (e < n) acc += in[e];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Scans one kST chunk per block (blocked arrangement: thread t owns kSI
// consecutive values), adding the chunk's offset from `chunk_offs`.
__global__ __launch_bounds__(kSB) void scan_chunks_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                          int64_t n, const uint32_t* __restrict__ chunk_offs)
{
    __shared__ uint32_t wsum[4];
    const int64_t base = (int64_t)blockIdx.x * kST + (int64_t)threadIdx.x * kSI;
    uint32_t v[kSI];
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < kSI; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        tot += v[i];
    }
    uint32_t ex = block_excl_scan256(tot, wsum) + (chunk_offs ? chunk_offs[blockIdx.x] : 0u);
#pragma unroll
    for (int i = 0; i < kSI; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

__global__ void read_last_kernel(const uint32_t* in, const uint32_t* ex, int64_t n, unsigned long long* total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = (unsigned long long)ex[n - 1] + in[n - 1];
}

void scan_rec(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s)
{
    const int64_t chunks = (n + kST - 1) / kST;
    if (chunks == 1) {
        hipLaunchKernelGGL(scan_chunks_kernel, dim3(1), dim3(kSB), 0, s, in, out, n, (const uint32_t*)nullptr);
        EG_LAUNCH_CHECK();
        return;
    }
    DevBuf sums((size_t)chunks * 4, s), offs((size_t)chunks * 4, s);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)chunks), dim3(kSB), 0, s, in, n, sums.get<uint32_t>());
    EG_LAUNCH_CHECK();
    scan_rec(sums.get<uint32_t>(), offs.get<uint32_t>(), chunks, s);
    hipLaunchKernelGGL(scan_chunks_kernel, dim3((unsigned)chunks), dim3(kSB), 0, s, in, out, n,
                       (const uint32_t*)offs.get<uint32_t>());
    EG_LAUNCH_CHECK();
}

// ---- key range ----------------------------------------------------------------------
__global__ __launch_bounds__(256) void key_range_kernel(const void* keys, int32_t type, int bytes, int64_t n,
                                                        unsigned long long* mn, unsigned long long* mx)
{
    uint64_t lo = ~0ull, hi = 0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t v = twiddle_bits(dev::load_raw(keys, bytes, i), type, bytes);
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t a = __shfl_xor((unsigned long long)lo, off), b = __shfl_xor((unsigned long long)hi, off);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(mn, (unsigned long long)lo);
        atomicMax(mx, (unsigned long long)hi);
    }
}

__global__ void init_range_kernel(unsigned long long* p)
{
    if (threadIdx.x == 0) {
        p[0] = ~0ull;
        p[1] = 0ull;
    }
}

}  // namespace

uint64_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, bool want_total)
{
    if (n <= 0) return 0;
    scan_rec(in, out, n, s);
    if (!want_total) return 0;
    DevBuf tot(8, s);
    hipLaunchKernelGGL(read_last_kernel, dim3(1), dim3(64), 0, s, in, (const uint32_t*)out, n,
                       tot.get<unsigned long long>());
    EG_LAUNCH_CHECK();
    int64_t* h = pinned_i64(1);
    EG_HIP(hipMemcpyAsync(h, tot.get(), 8, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    return (uint64_t)h[0];
}

void key_range(const void* keys, int32_t type, int bytes, int64_t n, hipStream_t s, uint64_t* mn, uint64_t* mx)
{
    DevBuf r(16, s);
    hipLaunchKernelGGL(init_range_kernel, dim3(1), dim3(64), 0, s, r.get<unsigned long long>());
    EG_LAUNCH_CHECK();
    if (n > 0) {
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)device_info().num_cus * 8));
        KernelTimer timer("key_range", s);
        hipLaunchKernelGGL(key_range_kernel, dim3((unsigned)grid), dim3(256), 0, s, keys, type, bytes, n,
                           r.get<unsigned long long>(), r.get<unsigned long long>() + 1);
        EG_LAUNCH_CHECK();
    }
    int64_t* h = pinned_i64(2);
    EG_HIP(hipMemcpyAsync(h, r.get(), 16, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    *mn = (uint64_t)h[0];
    *mx = (uint64_t)h[1];
}

void radix_pass(const RadixKeys& k, void* keys_out, const std::vector<Payload>& pay, int64_t n, int shift,
                int bits, hipStream_t s)
{
    if (n <= 0) return;
    if (n >= (int64_t)1 << 32) raise(EG_ERR_UNSUPPORTED, "radix pass: at most 2^32 - 1 rows");
    if (bits < 1 || bits > 8) raise(EG_ERR_INVALID, "radix pass: 1..8 digit bits");
    if ((int)pay.size() > kMaxPay) raise(EG_ERR_UNSUPPORTED, "radix pass: too many payload columns");
    const int ntiles = (int)((n + kRT - 1) / kRT);
    const uint32_t mask = (1u << bits) - 1u;
    const int nd = 1 << bits;
    const int64_t nh = (int64_t)nd * ntiles;  // [tile][digit] counts
    DevBuf hist((size_t)nh * 4, s), offs((size_t)nh * 4, s);
    {
        KernelTimer timer("radix_hist", s);
        auto hk = k.bytes == 1 ? radix_hist_kernel<1> : k.bytes == 2 ? radix_hist_kernel<2>
                : k.bytes == 4 ? radix_hist_kernel<4> : radix_hist_kernel<8>;
        hipLaunchKernelGGL(hk, dim3(ntiles), dim3(kRB), 0, s, k, n, shift, mask, hist.get<uint32_t>(), ntiles);
        EG_LAUNCH_CHECK();
    }
    {
        KernelTimer timer("radix_scan", s);
        const int per = std::max(16, (ntiles + 4095) / 4096);  // tiles per chunk: at most 4096 chunks
        const int nchunks = (ntiles + per - 1) / per;
        DevBuf cs((size_t)nchunks * nd * 4 * 2 + (size_t)nd * 4, s);
        uint32_t* csum = cs.get<uint32_t>();
        uint32_t* coff = csum + (size_t)nchunks * nd;
        uint32_t* total = coff + (size_t)nchunks * nd;
        hipLaunchKernelGGL(tile_chunk_sums_kernel, dim3(nchunks), dim3(kOB), 0, s, (const uint32_t*)hist.get<uint32_t>(),
                           ntiles, nd, per, csum);
        EG_LAUNCH_CHECK();
        hipLaunchKernelGGL(tile_chunk_scan_kernel, dim3(nd), dim3(1024), 0, s, (const uint32_t*)csum, nchunks, nd, coff,
                           total);
        EG_LAUNCH_CHECK();
        hipLaunchKernelGGL(tile_offsets_kernel, dim3(nchunks), dim3(kOB), 0, s, (const uint32_t*)hist.get<uint32_t>(),
                           ntiles, nd, per, (const uint32_t*)coff, (const uint32_t*)total, offs.get<uint32_t>());
        EG_LAUNCH_CHECK();
    }
    PayArgs pa{};
    pa.n = (int32_t)pay.size();
    for (size_t c = 0; c < pay.size(); ++c) {
        pa.in[c] = pay[c].in;
        pa.out[c] = pay[c].out;
        pa.bytes[c] = pay[c].bytes;
        pa.rowid[c] = pay[c].rowid;
    }
    KernelTimer timer("radix_scatter", s);
    auto sk = k.bytes == 1 ? radix_scatter_kernel<1> : k.bytes == 2 ? radix_scatter_kernel<2>
            : k.bytes == 4 ? radix_scatter_kernel<4> : radix_scatter_kernel<8>;
    hipLaunchKernelGGL(sk, dim3(ntiles), dim3(kRB), 0, s, k, keys_out, pa, n, shift, bits,
                       (const uint32_t*)offs.get<uint32_t>(), ntiles);
    EG_LAUNCH_CHECK();
}

void radix_sort(const RadixKeys& k0, SortBuffers& b, int64_t n, int begin_bit, int end_bit, hipStream_t s,
                const void** keys_res, std::vector<const void*>& pay_res)
{
    const size_t np = b.pay_in.size();
    pay_res.assign(np, nullptr);
    if (end_bit <= begin_bit || n <= 0) {
        // no digit bits: the input order is already sorted; copy it to A
        if (n > 0) {
            EG_HIP(hipMemcpyAsync(b.keys_a, b.keys_in, (size_t)n * k0.bytes, hipMemcpyDeviceToDevice, s));
            for (size_t c = 0; c < np; ++c) {
                if (b.pay_in[c].rowid) {
                    // generated row ids: one pass of zero digit bits is not possible, write iota
                    std::vector<int64_t> iota((size_t)n);
                    for (int64_t i = 0; i < n; ++i) iota[(size_t)i] = i;
                    EG_HIP(hipMemcpyAsync(b.pay_a[c], iota.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
                    EG_HIP(hipStreamSynchronize(s));
                } else {
                    EG_HIP(hipMemcpyAsync(b.pay_a[c], b.pay_in[c].in, (size_t)n * b.pay_in[c].bytes,
                                          hipMemcpyDeviceToDevice, s));
                }
            }
        }
        *keys_res = b.keys_a;
        for (size_t c = 0; c < np; ++c) pay_res[c] = b.pay_a[c];
        return;
    }
    RadixKeys k = k0;
    const void* kin = b.keys_in;
    std::vector<Payload> pay(np);
    for (size_t c = 0; c < np; ++c) pay[c] = b.pay_in[c];
    bool to_a = true;
    // balanced passes of at most kSortDigitBits bits
    const int total = end_bit - begin_bit;
    const int passes = (total + kSortDigitBits - 1) / kSortDigitBits;
    const int width = (total + passes - 1) / passes;
    for (int shift = begin_bit; shift < end_bit; shift += width) {
        const int bits = std::min(width, end_bit - shift);
        void* kout = to_a ? b.keys_a : b.keys_b;
        for (size_t c = 0; c < np; ++c) pay[c].out = to_a ? b.pay_a[c] : b.pay_b[c];
        k.data = kin;
        radix_pass(k, kout, pay, n, shift, bits, s);
        kin = kout;
        for (size_t c = 0; c < np; ++c) {
            pay[c].in = pay[c].out;
            if (pay[c].rowid) {
                pay[c].rowid = 0;
                pay[c].bytes = 8;
            }
        }
        to_a = !to_a;
    }
    *keys_res = kin;
    for (size_t c = 0; c < np; ++c) pay_res[c] = pay[c].in;
}

}  // namespace eg

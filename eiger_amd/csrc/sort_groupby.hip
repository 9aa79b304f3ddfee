// sort_groupby.hip -- sort-based group-by (gamax::groupby::sort::groupby and
// groupby_gfur, operators/groupby/groupby_sort.cu:346-479).
//
// The reference sorts (key, value) pairs once per aggregation request and runs
// cub::DeviceReduce::ReduceByKey per request (groupby_sort.cu:32-78,
// 363-419).  Here:
//   1. key_range: min/max of the radix-ordered keys; only the bits of
//      (key - min) are sorted, so dense keys take few passes (the reference
//      only narrows the bit range under hint.zero_based_ids,
//      groupby_sort.cu:222-246; the range here is exact and always applies).
//   2. one LSD radix sort carrying every value column (SORT) or the row ids
//      (SORT_GFUR: values are gathered through the permutation, like the
//      reference's permutation_iterator, groupby_sort.cu:429-479);
//   3. a segmented reduction that computes every aggregation in one pass:
//      seg_heads_kernel counts group heads per 4096-row tile, a scan gives each
//      tile its first group index, and seg_reduce_kernel folds each lane's
//      consecutive rows in registers, merges runs that cross lanes through LDS
//      states, writes the tile's complete groups and hands the group that
//      crosses into the next tile to seg_tail_kernel through a per-tile
//      boundary state.
// Output rows are in ascending key order, as in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "eg_device.hpp"
#include "eg_internal.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

constexpr int kSegB = 256;

__device__ __forceinline__ bool is_head(const void* keys, int kb, int64_t i)
{
    return i == 0 || dev::load_raw(keys, kb, i) != dev::load_raw(keys, kb, i - 1);
}

template <int ITEMS>
__global__ __launch_bounds__(kSegB) void seg_heads_kernel(const void* keys, int kb, int64_t n, uint32_t* heads)
{
    const int64_t r0 = (int64_t)blockIdx.x * (kSegB * ITEMS) + (int64_t)threadIdx.x * ITEMS;
    uint32_t c = 0;
    uint64_t prev = r0 > 0 && r0 < n ? dev::load_raw(keys, kb, r0 - 1) : 0ull;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t r = r0 + i;
        if (r < n) {
            const uint64_t k = dev::load_raw(keys, kb, r);
            c += (r == 0 || k != prev) ? 1u : 0u;
            prev = k;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    __shared__ uint32_t ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) heads[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total)
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
    uint32_t add = 0, tot = 0;
    for (int i = 0; i < kSegB / 64; ++i) {
        if (i < w) add += wsum[i];
        tot += wsum[i];
    }
    *total = tot;
    __syncthreads();
    return x - v + add;
}

// Largest tile u with head_off[u] <= g (the tile holding group g's head).
__device__ __forceinline__ uint32_t owner_tile(const uint32_t* head_off, uint32_t ntiles, uint32_t g)
{
    uint32_t lo = 0, hi = ntiles;  // head_off has ntiles + 1 entries
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (head_off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// States of local group lg, op a at st[a * T + lg].
template <int ITEMS, int NV>
__global__ __launch_bounds__(kSegB) void seg_reduce_kernel(const void* keys, int kb, int is_signed, AggSpec spec,
                                                           const int64_t* perm, int64_t n,
                                                           const uint32_t* __restrict__ head_off, uint32_t ntiles,
                                                           OutSpec os, unsigned long long* bstate,
                                                           unsigned long long* bkey, uint32_t* btrail)
{
    constexpr int T = kSegB * ITEMS;
    extern __shared__ unsigned long long smem[];
    unsigned long long* st = smem;                                  // nops * T
    unsigned long long* gkey = smem + (size_t)spec.nops * T;       // T
    __shared__ uint32_t wsum[kSegB / 64];
    __shared__ uint32_t s_first_cont, s_last_cont;

    const int t = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const int64_t tstart = (int64_t)tile * T;
    const int64_t r0 = tstart + (int64_t)t * ITEMS;
    const int nops = spec.nops;

    uint64_t k[ITEMS];
    bool real_head[ITEMS];
    uint32_t nforced = 0;
    uint64_t prev = (r0 > 0 && r0 < n) ? dev::load_raw(keys, kb, r0 - 1) : 0ull;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t r = r0 + i;
        k[i] = r < n ? dev::load_raw(keys, kb, r) : 0ull;
        real_head[i] = r < n && (r == 0 || k[i] != prev);
        const bool forced = r < n && (real_head[i] || r == tstart);
        nforced += forced ? 1u : 0u;
        prev = k[i];
    }
    if (t == 0) s_first_cont = (tstart > 0 && !real_head[0]) ? 1u : 0u;
    if (t == kSegB - 1) {
        const int64_t nx = tstart + T;
        s_last_cont = (nx < n && dev::load_raw(keys, kb, nx) == dev::load_raw(keys, kb, nx - 1)) ? 1u : 0u;
    }
    uint32_t G = 0;
    const uint32_t lg0 = block_excl_scan(nforced, wsum, &G);
    for (int i = t; i < nops * (int)G; i += kSegB) {
        const int a = i / (int)G, g = i - a * (int)G;
        st[(size_t)a * T + g] = 0ull;
    }
    __syncthreads();

    // fold this lane's rows; flush each run into its local group's LDS states
    uint64_t acc[kMaxOps];
    int lg = (int)lg0 - 1;
    bool have = false;
    auto flush = [&]() {
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a)
            if (a < nops) dev::apply_atomic(dev::class_of(spec, a), &st[(size_t)a * T + lg], acc[a]);
    };
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t r = r0 + i;
        if (r >= n) break;
        const bool forced = real_head[i] || r == tstart;
        if (forced) {
            if (have) flush();
            ++lg;
            if (real_head[i] || r == tstart) gkey[lg] = k[i];
        }
        const int64_t src = perm ? perm[r] : r;
        uint64_t cv[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c)
            cv[c] = c < spec.nvcols ? dev::canon(dev::load_raw(spec.cols[c].data, spec.cols[c].bytes, src),
                                                 spec.cols[c].conv)
                                    : 0ull;
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a) {
            if (a < nops) {
                const OpDesc od = spec.ops[a];
                uint64_t v = 1ull;
#pragma unroll
                for (int c = 0; c < NV; ++c)
                    if (c == od.col) v = cv[c];
                const uint64_t o = dev::operand(od, v);
                acc[a] = (forced || !have) ? o : dev::combine(dev::class_of(spec, a), acc[a], o);
            }
        }
        have = true;
    }
    if (have) flush();
    __syncthreads();

    const bool first_cont = s_first_cont != 0u;
    const bool last_cont = s_last_cont != 0u;
    const uint32_t hbase = head_off[tile];
    for (int g = t; g < (int)G; g += kSegB) {
        const uint32_t gidx = hbase + (uint32_t)g - (first_cont ? 1u : 0u);
        if (first_cont && g == 0) {
            // continuation of a group whose head is in an earlier tile
            const uint32_t owner = owner_tile(head_off, ntiles, gidx);
            unsigned long long* dst = bstate + (size_t)owner * nops;
            for (int a = 0; a < nops; ++a) dev::apply_atomic(dev::class_of(spec, a), dst + a, st[(size_t)a * T + g]);
        } else if (last_cont && g == (int)G - 1) {
            unsigned long long* dst = bstate + (size_t)tile * nops;
            for (int a = 0; a < nops; ++a) dev::apply_atomic(dev::class_of(spec, a), dst + a, st[(size_t)a * T + g]);
            bkey[tile] = gkey[g];
            btrail[tile] = 1u;
        } else {
            dev::emit_group(os, (int64_t)gidx, dev::enc64(dev::widen_key(gkey[g], kb, is_signed)), st + g, (int64_t)T);
        }
    }
}

__global__ void seg_tail_kernel(uint32_t ntiles, const uint32_t* head_off, const unsigned long long* bstate,
                                const unsigned long long* bkey, const uint32_t* btrail, int nops, int kb,
                                int is_signed, OutSpec os)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles || btrail[t] == 0u) return;
    const uint32_t gidx = head_off[t + 1] - 1u;
    dev::emit_group(os, (int64_t)gidx, dev::enc64(dev::widen_key(bkey[t], kb, is_signed)), bstate + (size_t)t * nops,
                    (int64_t)1);
}

template <int ITEMS, int NV>
void launch_seg(const void* keys, int kb, int is_signed, const AggSpec& spec, const int64_t* perm, int64_t n,
                const uint32_t* head_off, uint32_t ntiles, const OutSpec& os, unsigned long long* bstate,
                unsigned long long* bkey, uint32_t* btrail, hipStream_t s)
{
    constexpr int T = kSegB * ITEMS;
    const size_t lds = (size_t)(spec.nops + 1) * T * 8;
    auto kern = seg_reduce_kernel<ITEMS, NV>;
    if (enable_dynamic_lds((const void*)kern) < lds) raise(EG_ERR_UNSUPPORTED, "segmented reduction tile too large");
    KernelTimer timer("segreduce", s);
    hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kSegB), lds, s, keys, kb, is_signed, spec, perm, n, head_off, ntiles,
                       os, bstate, bkey, btrail);
    EG_LAUNCH_CHECK();
}

template <int ITEMS>
void launch_seg_nv(const void* keys, int kb, int is_signed, const AggSpec& spec, const int64_t* perm, int64_t n,
                   const uint32_t* head_off, uint32_t ntiles, const OutSpec& os, unsigned long long* bstate,
                   unsigned long long* bkey, uint32_t* btrail, hipStream_t s)
{
    const int nv = spec.nvcols;
    if (nv <= 1) launch_seg<ITEMS, 1>(keys, kb, is_signed, spec, perm, n, head_off, ntiles, os, bstate, bkey, btrail, s);
    else if (nv <= 2) launch_seg<ITEMS, 2>(keys, kb, is_signed, spec, perm, n, head_off, ntiles, os, bstate, bkey, btrail, s);
    else if (nv <= 4) launch_seg<ITEMS, 4>(keys, kb, is_signed, spec, perm, n, head_off, ntiles, os, bstate, bkey, btrail, s);
    else launch_seg<ITEMS, 8>(keys, kb, is_signed, spec, perm, n, head_off, ntiles, os, bstate, bkey, btrail, s);
}

}  // namespace

// Segmented reduction of rows already ordered by key (equal keys adjacent):
// writes every group into freshly allocated outputs.  perm: optional row
// gather for the value columns (GFUR).
Result reduce_sorted(const Plan& plan, const void* sorted_keys, const AggSpec& spec, const int64_t* perm,
                     hipStream_t s)
{
    const int64_t n = plan.rows;
    const int kb = plan.key.bytes;
    // tile size from the LDS the per-tile states need: (nops + 1) * T * 8 bytes
    const size_t max_lds = (size_t)device_info().max_lds - 256;  // static LDS of the kernel
    int items = 16;
    while (items > 4 && (size_t)(spec.nops + 1) * kSegB * items * 8 > max_lds) items >>= 1;
    const int T = kSegB * items;
    const int64_t ntiles64 = (n + T - 1) / T;
    if (ntiles64 >= ((int64_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "too many rows for the segmented reduction");
    const uint32_t ntiles = (uint32_t)ntiles64;

    DevBuf heads((size_t)(ntiles + 1) * 4, s), head_off((size_t)(ntiles + 1) * 4, s);
    fill_zero(heads.get(), (size_t)(ntiles + 1) * 4, s);
    {
        KernelTimer timer("seg_heads", s);
        if (items == 16) hipLaunchKernelGGL(seg_heads_kernel<16>, dim3(ntiles), dim3(kSegB), 0, s, sorted_keys, kb, n, heads.get<uint32_t>());
        else if (items == 8) hipLaunchKernelGGL(seg_heads_kernel<8>, dim3(ntiles), dim3(kSegB), 0, s, sorted_keys, kb, n, heads.get<uint32_t>());
        else hipLaunchKernelGGL(seg_heads_kernel<4>, dim3(ntiles), dim3(kSegB), 0, s, sorted_keys, kb, n, heads.get<uint32_t>());
        EG_LAUNCH_CHECK();
    }
    // head_off[t] = groups whose head lies before tile t; head_off[ntiles] = G
    const uint64_t groups = scan_exclusive_u32(heads.get<uint32_t>(), head_off.get<uint32_t>(), (int64_t)ntiles + 1, s, true);

    Result res;
    OutSpec os;
    alloc_outputs(plan, (int64_t)groups, s, res, os);
    DevBuf bstate((size_t)ntiles * spec.nops * 8 + 8, s), bkey((size_t)ntiles * 8, s), btrail((size_t)ntiles * 4, s);
    fill_zero(bstate.get(), (size_t)ntiles * spec.nops * 8 + 8, s);
    fill_zero(btrail.get(), (size_t)ntiles * 4, s);
    const int sg = plan.key.is_signed;
    if (items == 16) launch_seg_nv<16>(sorted_keys, kb, sg, spec, perm, n, head_off.get<uint32_t>(), ntiles, os, bstate.get<unsigned long long>(), bkey.get<unsigned long long>(), btrail.get<uint32_t>(), s);
    else if (items == 8) launch_seg_nv<8>(sorted_keys, kb, sg, spec, perm, n, head_off.get<uint32_t>(), ntiles, os, bstate.get<unsigned long long>(), bkey.get<unsigned long long>(), btrail.get<uint32_t>(), s);
    else launch_seg_nv<4>(sorted_keys, kb, sg, spec, perm, n, head_off.get<uint32_t>(), ntiles, os, bstate.get<unsigned long long>(), bkey.get<unsigned long long>(), btrail.get<uint32_t>(), s);
    hipLaunchKernelGGL(seg_tail_kernel, dim3((ntiles + 255) / 256), dim3(256), 0, s, ntiles,
                       (const uint32_t*)head_off.get<uint32_t>(), (const unsigned long long*)bstate.get<unsigned long long>(),
                       (const unsigned long long*)bkey.get<unsigned long long>(), (const uint32_t*)btrail.get<uint32_t>(),
                       spec.nops, kb, sg, os);
    EG_LAUNCH_CHECK();
    EG_HIP(hipStreamSynchronize(s));
    return res;
}

Result run_sort(const Plan& plan, const eg_groupby_hint& hint, bool gfur, hipStream_t s)
{
    (void)hint;
    const int64_t n = plan.rows;
    const int kb = plan.key.bytes;
    uint64_t mn = 0, mx = 0;
    key_range(plan.key.data, plan.key_type, kb, n, s, &mn, &mx);
    const uint64_t span = mx - mn;
    const int bits = span == 0 ? 0 : 64 - __builtin_clzll(span);

    RadixKeys rk{};
    rk.data = plan.key.data;
    rk.bytes = kb;
    rk.type = plan.key_type;
    rk.mode = DIG_KEY;
    rk.kmin = mn;
    SortBuffers b;
    b.keys_in = plan.key.data;
    DevBuf ka((size_t)n * kb, s), kbuf((size_t)n * kb, s);
    b.keys_a = ka.get();
    b.keys_b = kbuf.get();
    std::vector<DevBuf> bufs;
    AggSpec spec = plan.spec;
    if (gfur) {
        b.pay_in.push_back(Payload{nullptr, nullptr, 8, 1});
        bufs.emplace_back((size_t)n * 8, s);
        bufs.emplace_back((size_t)n * 8, s);
        b.pay_a.push_back(bufs[0].get());
        b.pay_b.push_back(bufs[1].get());
    } else {
        for (int c = 0; c < spec.nvcols; ++c) {
            const int by = spec.cols[c].bytes;
            b.pay_in.push_back(Payload{spec.cols[c].data, nullptr, by, 0});
            bufs.emplace_back((size_t)n * by, s);
            bufs.emplace_back((size_t)n * by, s);
            b.pay_a.push_back(bufs[bufs.size() - 2].get());
            b.pay_b.push_back(bufs[bufs.size() - 1].get());
        }
    }
    const void* sorted_keys = nullptr;
    std::vector<const void*> sorted_pay;
    radix_sort(rk, b, n, 0, bits, s, &sorted_keys, sorted_pay);
    const int64_t* perm = nullptr;
    if (gfur) perm = (const int64_t*)sorted_pay[0];
    else
        for (int c = 0; c < spec.nvcols; ++c) spec.cols[c].data = sorted_pay[(size_t)c];
    return reduce_sorted(plan, sorted_keys, spec, perm, s);
}

}  // namespace eg

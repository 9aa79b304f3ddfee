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
#include "eg_lds_table.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

using namespace lds;

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

// Writes group g from states held in registers.
__device__ __forceinline__ void emit_regs(const OutSpec& os, int64_t g, uint64_t key_enc, const uint64_t (&acc)[kMaxOps])
{
    dev::store_bytes(os.key_data, os.key_bytes, g, dev::dec64(key_enc));
    for (int j = 0; j < os.nouts; ++j) {
        const OutDesc& o = os.outs[j];
        uint64_t st[2] = {0ull, 0ull};
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a) {
            if (a == o.op) st[0] = acc[a];
            if (a == o.op2) st[1] = acc[a];
        }
        const uint64_t s = st[0];
        switch (o.code) {
        case OUT_U64: dev::store_bytes(o.data, 8, g, s); break;
        case OUT_F32: ((float*)o.data)[g] = (float)__longlong_as_double((long long)s); break;
        case OUT_F64: ((double*)o.data)[g] = __longlong_as_double((long long)s); break;
        case OUT_MAX_I: dev::store_bytes(o.data, o.bytes, g, s ^ dev::kSign64); break;
        case OUT_MIN_I: dev::store_bytes(o.data, o.bytes, g, (~s) ^ dev::kSign64); break;
        case OUT_MAX_U: dev::store_bytes(o.data, o.bytes, g, s); break;
        case OUT_MIN_U: dev::store_bytes(o.data, o.bytes, g, ~s); break;
        case OUT_MAX_F: dev::store_double_as(o.data, o.bytes, g, dev::unord_f64(s)); break;
        case OUT_MIN_F: dev::store_double_as(o.data, o.bytes, g, dev::unord_f64(~s)); break;
        case OUT_ANY:
            if (o.is_float && o.bytes == 4) ((float*)o.data)[g] = (float)__longlong_as_double((long long)s);
            else dev::store_bytes(o.data, o.bytes, g, s);
            break;
        default: {
            const double cnt = (double)st[1];
            const double sum = (o.code == OUT_AVG_I) ? (double)(int64_t)s
                             : (o.code == OUT_AVG_U) ? (double)s
                             : __longlong_as_double((long long)s);
            ((double*)o.data)[g] = sum / cnt;
            break;
        }
        }
    }
}

// One 4096-row tile of sorted rows (ITEMS consecutive rows per lane).  A run
// (rows of one key) that starts and ends inside a lane is folded in registers
// and written straight out.  Runs that cross lane boundaries meet in LDS:
// slot L (nops states) collects the group whose head is in lane L, slot
// kSegB the group whose head lies before the tile.  A continuation run finds
// its slot through a max-scan of "last lane with a head".  The tile's last
// group, if it continues into the next tile, goes to the tile's boundary
// state instead; the continuation group of an earlier tile's head is added
// to that tile's boundary state.
template <int ITEMS, int KB, class VL>
__global__ __launch_bounds__(kSegB) void seg_reduce_kernel(const void* keys, int kb_unused, int is_signed, AggSpec spec,
                                                           const int64_t* perm, int64_t n,
                                                           const uint32_t* __restrict__ head_off, uint32_t ntiles,
                                                           OutSpec os, unsigned long long* bstate,
                                                           unsigned long long* bkey, uint32_t* btrail)
{
    constexpr int T = kSegB * ITEMS;
    constexpr int kSlots = kSegB + 1;
    constexpr int NV = VL::NV > 0 ? VL::NV : 1;
    constexpr int kb = KB;
    (void)kb_unused;
    extern __shared__ unsigned long long slot[];  // nops * kSlots
    __shared__ uint32_t wsum[kSegB / 64];
    __shared__ int wmax[kSegB / 64];
    __shared__ uint32_t s_tile_cont;

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t tile = blockIdx.x;
    const int64_t tstart = (int64_t)tile * T;
    const int64_t r0 = tstart + (int64_t)t * ITEMS;
    const int nops = spec.nops;
    for (int i = t; i < nops * kSlots; i += kSegB) slot[i] = 0ull;
    if (t == 0) s_tile_cont = 0u;

    uint64_t k[ITEMS];
    bool head[ITEMS];
    uint32_t nheads = 0;
    auto key_at = [&](int64_t r) -> uint64_t {
        return KB == 4 ? (uint64_t)((const uint32_t*)keys)[r] : ((const uint64_t*)keys)[r];
    };
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) k[i] = r0 + i < n ? key_at(r0 + i) : 0ull;
    uint64_t prev = (r0 > 0 && r0 < n) ? key_at(r0 - 1) : 0ull;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t r = r0 + i;
        head[i] = r < n && (r == 0 || k[i] != prev);
        nheads += head[i] ? 1u : 0u;
        prev = k[i];
    }
    const int64_t r_end = r0 + ITEMS;  // first row after this lane
    const bool next_head = r_end >= n || (r0 < n && key_at(r_end) != k[ITEMS - 1]);
    // groups whose head precedes this lane within the tile, and the tile total
    uint32_t tile_heads = 0;
    const uint32_t hbefore = block_excl_scan(nheads, wsum, &tile_heads);
    // owner of a continuation: the last lane before this one with a head (-1: before the tile)
    int mine = nheads > 0 ? t : -1;
    int incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off);
        if (lane >= off) incl = y > incl ? y : incl;
    }
    if (lane == 63) wmax[w] = incl;
    __syncthreads();  // also orders the slot clearing above
    int prev_max = -1;
    for (int i = 0; i < w; ++i) prev_max = wmax[i] > prev_max ? wmax[i] : prev_max;
    int excl = __shfl_up(incl, 1);
    if (lane == 0) excl = -1;
    excl = excl > prev_max ? excl : prev_max;
    const int owner = excl < 0 ? kSegB : excl;  // slot of a continuation run

    const uint32_t hbase = head_off[tile];
    const bool last_cont_tile = [&] {
        const int64_t nx = tstart + T;
        return nx < n && key_at(nx) == key_at(nx - 1);
    }();

    uint64_t acc[kMaxOps];
#pragma unroll
    for (int a = 0; a < kMaxOps; ++a) acc[a] = 0ull;
    bool have = false, cont = r0 < n && r0 > 0 && !head[0];
    uint32_t hseen = 0;          // heads of this lane seen so far
    uint64_t run_key = 0;
    auto to_slot = [&](int sidx) {
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a)
            if (a < nops) dev::apply_atomic(dev::class_of(spec, a), &slot[(size_t)a * kSlots + sidx], acc[a]);
    };
    // closes the current run; `closed`: the run ends inside this lane
    auto finish = [&](bool closed) {
        if (!have) return;
        if (cont) {
            to_slot(owner);
            if (owner == kSegB) s_tile_cont = 1u;
        } else if (closed) {
            emit_regs(os, (int64_t)hbase + hbefore + hseen - 1, dev::enc64(dev::widen_key(run_key, kb, is_signed)), acc);
        } else {
            to_slot(t);
        }
    };
    // values of the lane's rows: typed loads issued eight rows at a time
    // (fixed layouts), then folded
    constexpr int H = 8;
    static_assert(ITEMS % H == 0, "rows per lane in halves");
#pragma unroll
    for (int h0 = 0; h0 < ITEMS; h0 += H) {
    uint64_t rawv[VL::GEN ? 1 : H][NV];
    if constexpr (!VL::GEN) {
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const int64_t r = r0 + h0 + j;
            const bool valid = r < n;
            load_row<VL>(spec, valid ? (perm ? perm[r] : r) : 0, valid, rawv[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const int i = h0 + j;
        const int64_t r = r0 + i;
        if (r >= n) break;
        if (head[i]) {
            finish(true);
            have = false;
            cont = false;
            ++hseen;
            run_key = k[i];
        }
        uint64_t cv[NV];
        if constexpr (VL::GEN) {
            uint64_t raw[NV];
            load_row<VL>(spec, perm ? perm[r] : r, true, raw);
            canon_row<VL, NV>(spec, raw, cv);
        } else {
            canon_row<VL, NV>(spec, rawv[j], cv);
        }
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a) {
            if (a < nops) {
                const OpDesc od = spec.ops[a];
                uint64_t v = 1ull;
#pragma unroll
                for (int c = 0; c < NV; ++c)
                    if (c == od.col) v = cv[c];
                const uint64_t o = dev::operand(od, v);
                acc[a] = have ? dev::combine(dev::class_of(spec, a), acc[a], o) : o;
            }
        }
        have = true;
    }
    }
    // the lane's last run: complete when the next row starts a new group
    finish(next_head);
    __syncthreads();

    // groups collected in slots: the one with its head in this lane
    const bool owns_open = have && !cont && !next_head;
    if (owns_open) {
        const uint32_t gidx = hbase + hbefore + hseen - 1;
        if (last_cont_tile && hbefore + hseen == tile_heads) {
            // the tile's last group continues into the next tile
            unsigned long long* dst = bstate + (size_t)tile * nops;
            for (int a = 0; a < nops; ++a)
                dev::apply_atomic(dev::class_of(spec, a), dst + a, slot[(size_t)a * kSlots + t]);
            bkey[tile] = run_key;
            btrail[tile] = 1u;
        } else {
            dev::emit_group(os, (int64_t)gidx, dev::enc64(dev::widen_key(run_key, kb, is_signed)), slot + t,
                            (int64_t)kSlots);
        }
    }
    if (t == 0 && s_tile_cont) {
        // continuation of a group whose head is in an earlier tile
        const uint32_t own = owner_tile(head_off, ntiles, hbase - 1u);
        unsigned long long* dst = bstate + (size_t)own * nops;
        for (int a = 0; a < nops; ++a)
            dev::apply_atomic(dev::class_of(spec, a), dst + a, slot[(size_t)a * kSlots + kSegB]);
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

template <int ITEMS>
void launch_seg(const void* keys, int kb, int is_signed, const AggSpec& spec, const int64_t* perm, int64_t n,
                const uint32_t* head_off, uint32_t ntiles, const OutSpec& os, unsigned long long* bstate,
                unsigned long long* bkey, uint32_t* btrail, hipStream_t s)
{
    const size_t lds = (size_t)spec.nops * (kSegB + 1) * 8;
    auto go = [&](auto kern) {
        if (enable_dynamic_lds((const void*)kern) < lds) raise(EG_ERR_UNSUPPORTED, "segmented reduction slots too large");
        KernelTimer timer("segreduce", s);
        hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kSegB), lds, s, keys, kb, is_signed, spec, perm, n, head_off, ntiles,
                           os, bstate, bkey, btrail);
        EG_LAUNCH_CHECK();
    };
    with_vlayout(spec, true, [&](auto vl) {
        using VL = decltype(vl);
        if (kb == 4) go(seg_reduce_kernel<ITEMS, 4, VL>);
        else go(seg_reduce_kernel<ITEMS, 8, VL>);
    });
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
    constexpr int items = 16;
    const int T = kSegB * items;
    const int64_t ntiles64 = (n + T - 1) / T;
    if (ntiles64 >= ((int64_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "too many rows for the segmented reduction");
    const uint32_t ntiles = (uint32_t)ntiles64;

    DevBuf heads((size_t)(ntiles + 1) * 4, s), head_off((size_t)(ntiles + 1) * 4, s);
    fill_zero(heads.get(), (size_t)(ntiles + 1) * 4, s);
    {
        KernelTimer timer("seg_heads", s);
        hipLaunchKernelGGL(seg_heads_kernel<items>, dim3(ntiles), dim3(kSegB), 0, s, sorted_keys, kb, n,
                           heads.get<uint32_t>());
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
    launch_seg<items>(sorted_keys, kb, sg, spec, perm, n, head_off.get<uint32_t>(), ntiles, os,
                      bstate.get<unsigned long long>(), bkey.get<unsigned long long>(), btrail.get<uint32_t>(), s);
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

// eg_agg_kernels.hpp -- bodies of the two row-streaming aggregation kernels,
// written against an op-list policy (dev::RtOps or dev::FixOps):
//   hash_agg   the persistent hash kernel (hash strategy, hash_groupby.hip);
//   part_agg   one partition unit per workgroup (partition strategy,
//              partition_groupby.hip).
// The library instantiates them with RtOps; the run-time compiler (jit.hip)
// instantiates the same bodies with the request list's ops compiled in, so
// both kernels compute identical results.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "eg_atomics.hpp"
#include "eg_device.hpp"
#include "eg_hash_table.hpp"
#include "eg_lds_table.hpp"

namespace eg {
namespace agg {

using namespace lds;

// Wider value sets need more registers per lane: cap the workgroup at 512.
template <int NV>
constexpr int max_block() { return NV <= 2 ? 1024 : 512; }

// Merges the replicas of the block's LDS table and adds every group to the
// HBM table.
template <class OPS, int KB>
__device__ __forceinline__ void flush_lds(LdsTable<KB>& tab, const GTable& gt, const KeyDesc& key, const AggSpec& spec,
                          int rep_log2, int r_mask, uint32_t& ins)
{
    using KT = typename LdsKey<KB>::T;
    __syncthreads();
    // merge replicas 1..R-1 into replica 0 (LDS atomics)
    const int R = 1 << rep_log2;
    if (R > 1) {
        for (int idx = threadIdx.x; idx < tab.T; idx += blockDim.x) {
            const int r = idx & r_mask;
            if (r == 0) continue;
            const int s = idx >> rep_log2;
            int dst;
            KT kw = 0;
            if (s == tab.cr) {
                if (tab.side[r] == 0ull) continue;
                dst = tab.find((KT)0, 0, 1);
            } else {
                kw = tab.keys[idx];
                if (kw == 0) continue;
                dst = tab.find(kw, 0, tab.cr);
            }
            auto src = [&](int a) { return tab.st[(size_t)a * tab.T + idx]; };
            if (dst >= 0) {
                dev::merge_states<OPS>(spec, [&](int a) { return &tab.st[(size_t)a * tab.T + dst]; }, src);
            } else {
                // replica 0 is full: merge this slot straight into HBM
                unsigned long long* g = gt_slot(gt, lds_to_enc64<KB>(kw, key.is_signed), ins);
                if (g) dev::merge_states<OPS>(spec, [&](int a) { return g + 1 + a; }, src);
            }
        }
        __syncthreads();
    }
    // flush replica 0 into the HBM table
    for (int s = threadIdx.x; s <= tab.cr; s += blockDim.x) {
        const int idx = s << rep_log2;
        KT kw;
        if (s == tab.cr) {
            if (tab.side[0] == 0ull) continue;
            kw = 0;
        } else {
            kw = tab.keys[idx];
            if (kw == 0) continue;
        }
        unsigned long long* g = gt_slot(gt, lds_to_enc64<KB>(kw, key.is_signed), ins);
        if (!g) continue;
        dev::merge_states<OPS>(spec, [&](int a) { return g + 1 + a; }, [&](int a) { return tab.st[(size_t)a * tab.T + idx]; });
    }
}

template <int KB, class VL, class OPS>
__device__ __forceinline__ void hash_agg(const KeyDesc& key, int64_t n, const AggSpec& spec, const GTable& gt,
                                         int cr_log2, int rep_log2, int probe_limit, int peel_min)
{
    constexpr int NV = VL::NV > 0 ? VL::NV : 1;
    extern __shared__ __align__(16) unsigned long long smem[];
    const bool use_lds = cr_log2 >= 0;
    // Rows reach the HBM table as spills of cold keys (hot keys take LDS
    // slots first), where combining seldom finds partners; only without an
    // LDS stage do hot keys meet there.
    const int hbm_peel = use_lds ? 65 : peel_min;
    const int nops = OPS::n(spec);
    const int lane = dev::lane_id();
    LdsTable<KB> tab(smem, nops, use_lds ? cr_log2 : 0, use_lds ? rep_log2 : 0);
    if (use_lds) {
        const int words = LdsTable<KB>::words(nops, cr_log2, rep_log2);
        for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
        __syncthreads();
    }
    const int r_mask = (1 << (use_lds ? rep_log2 : 0)) - 1;
    uint32_t ins = 0;  // HBM-table insertions by this lane
    dev::HotTags hot;  // groups the wave combined last

    // ---- full tiles: vector loads, next tile in flight while this one runs ----
    const int64_t tile = (int64_t)blockDim.x * kRows;
    const int64_t full = n / tile;
    int64_t t = blockIdx.x;
    uint64_t kcur[kRows], vcur[NV][kRows];
    if (t < full) load_tile<VL>(key, KB, spec, t * tile + (int64_t)threadIdx.x * kRows, kcur, vcur);
    for (; t < full; t += gridDim.x) {
        uint64_t knxt[kRows], vnxt[NV][kRows];
        const int64_t tn = t + gridDim.x;
        if (tn < full) load_tile<VL>(key, KB, spec, tn * tile + (int64_t)threadIdx.x * kRows, knxt, vnxt);
        uint32_t spill = 0;
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            uint64_t raw[NV], cv[NV];
#pragma unroll
            for (int c = 0; c < NV; ++c) raw[c] = vcur[c][i];
            canon_row<VL, NV>(spec, raw, cv);
            if (lds_row<OPS, KB, NV>(tab, use_lds, r_mask, lane, key, spec, kcur[i], cv, true, probe_limit, peel_min, hot))
                spill |= 1u << i;
        }
        if (__ballot(spill != 0) != 0ull) {
#pragma unroll 1
            for (int i = 0; i < kRows; ++i) {
                const bool pend = (spill >> i) & 1u;
                if (__ballot(pend) == 0ull) continue;
                uint64_t raw[NV], cv[NV];
#pragma unroll
                for (int c = 0; c < NV; ++c) raw[c] = pick4(vcur[c], i);
                canon_row<VL, NV>(spec, raw, cv);
                hbm_row<OPS, KB, NV>(gt, key, spec, pend, pick4(kcur, i), cv, hbm_peel, ins, hot);
            }
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            kcur[i] = knxt[i];
#pragma unroll
            for (int c = 0; c < NV; ++c) vcur[c][i] = vnxt[c][i];
        }
    }
    // ---- tail rows: one row per lane, typed loads ----
    for (int64_t b = full * tile + (int64_t)blockIdx.x * blockDim.x; b < n; b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = b + threadIdx.x;
        const bool valid = r < n;
        const uint64_t kraw = valid ? dev::load_raw(key.data, KB, r) : 0ull;
        uint64_t raw[NV], cv[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c)
            raw[c] = (valid && c < spec.nvcols) ? dev::load_raw(spec.cols[c].data, spec.cols[c].bytes, r) : 0ull;
        canon_row<VL, NV>(spec, raw, cv);
        const bool sp = lds_row<OPS, KB, NV>(tab, use_lds, r_mask, lane, key, spec, kraw, cv, valid, probe_limit, peel_min, hot);
        if (__ballot(sp) != 0ull) hbm_row<OPS, KB, NV>(gt, key, spec, sp, kraw, cv, hbm_peel, ins, hot);
    }
    if (use_lds) flush_lds<OPS>(tab, gt, key, spec, rep_log2, r_mask, ins);
    gt_count_inserts(gt, ins);
}

constexpr int kPBMax = 1024;  // largest workgroup of the unit kernel

// A unit of the partition kernel: rows [start, end) of the partitioned input.
struct Unit {
    int64_t start;
    int64_t end;
    int32_t sliced;  // the partition spans several units
    int32_t pad;
};

// GATHER: values are read through the row-id permutation `perm` (the
// PARTITION_GFUR layout), else straight from the partitioned columns.
template <int KB, class VL, class OPS, bool GATHER>
__device__ __forceinline__ void part_agg(const KeyDesc& key, const AggSpec& spec, const int64_t* perm,
                                         const Unit* units, const GTable& gt, const OutSpec& os,
                                         unsigned long long* counter, int64_t capacity, int cr_log2,
                                         int probe_limit, int peel_min)
{
    constexpr int NV = VL::NV > 0 ? VL::NV : 1;
    constexpr int J = 4;  // rows per lane per step: independent loads in flight
    using KT = typename LdsKey<KB>::T;
    extern __shared__ __align__(16) unsigned long long smem[];
    __shared__ unsigned long long sh[17];
    const int nops = OPS::n(spec);
    const int lane = dev::lane_id();
    LdsTable<KB, true> tab(smem, nops, cr_log2, 0);  // keys of a unit share hash bits: bucketed probing
    const int words = LdsTable<KB>::words(nops, cr_log2, 0);
    for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
    __syncthreads();

    const Unit u = units[blockIdx.x];  // never empty
    uint32_t ins = 0;  // HBM-table insertions by this lane
    dev::HotTags hot;  // groups the wave combined last
    // Loads of one step (J rows per lane): all issued before any is used;
    // rows past the unit's end read its first row instead (in bounds, no
    // branch) and are masked by `valid`.
    auto load_step = [&](int64_t b, uint64_t (&kr)[J], uint64_t (&raw)[J][NV], bool (&valid)[J]) {
        int64_t src[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int64_t r = b + (int64_t)j * blockDim.x + threadIdx.x;
            valid[j] = r < u.end;
            src[j] = valid[j] ? r : u.start;
            kr[j] = KB == 4 ? (uint64_t)((const uint32_t*)key.data)[src[j]] : ((const uint64_t*)key.data)[src[j]];
        }
        if constexpr (GATHER) {
#pragma unroll
            for (int j = 0; j < J; ++j) src[j] = perm[src[j]];
        }
#pragma unroll
        for (int j = 0; j < J; ++j) load_row<VL>(spec, src[j], true, raw[j]);
    };
    const int64_t step = (int64_t)J * blockDim.x;
    uint64_t kr[J], raw[J][NV];
    bool valid[J];
    load_step(u.start, kr, raw, valid);
    for (int64_t b = u.start; b < u.end; b += step) {
        // the next step's rows are in flight while this step aggregates
        uint64_t kn[J], rn[J][NV];
        bool vn[J];
        if (b + step < u.end) load_step(b + step, kn, rn, vn);
#pragma unroll
        for (int j = 0; j < J; ++j) {
            uint64_t cv[NV];
            canon_row<VL, NV>(spec, raw[j], cv);
            const bool sp = lds_row<OPS, KB, NV>(tab, true, 0, lane, key, spec, kr[j], cv, valid[j], probe_limit, peel_min, hot);
            if (__ballot(sp) != 0ull) hbm_row<OPS, KB, NV>(gt, key, spec, sp, kr[j], cv, 65, ins, hot);  // spills: cold keys
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            kr[j] = kn[j];
            valid[j] = vn[j];
#pragma unroll
            for (int c = 0; c < NV; ++c) raw[j][c] = rn[j][c];
        }
    }
    __syncthreads();
    // finished groups of this unit
    auto slot_key = [&](int sl, KT& kw) {
        kw = 0;
        if (sl < tab.cr) {
            kw = tab.keys[sl];
            return kw != 0;
        }
        return sl == tab.cr && tab.side[0] != 0ull;
    };
    if (u.sliced) {
        // the partition continues in other units: merge into the HBM table
        for (int sl = threadIdx.x; sl <= tab.cr; sl += blockDim.x) {
            KT kw;
            if (!slot_key(sl, kw)) continue;
            unsigned long long* g = gt_slot(gt, lds_to_enc64<KB>(kw, key.is_signed), ins);
            if (g)
                dev::merge_states<OPS>(spec, [&](int a) { return g + 1 + a; },
                                       [&](int a) { return tab.st[(size_t)a * tab.T + sl]; });
        }
    } else {
        // Written straight out.  One device atomic per unit claims the
        // unit's output range (a same-address atomic per chunk of slots
        // serialises across the ~10^5 units); each wave then owns a
        // contiguous run of it.
        const int w = threadIdx.x >> 6;
        const int nw = (blockDim.x + 63) >> 6;
        uint32_t wave_groups = 0;
        for (int base = 0; base <= tab.cr; base += blockDim.x) {
            KT kw;
            wave_groups += (uint32_t)__popcll(__ballot(slot_key(base + threadIdx.x, kw)));
        }
        if (lane == 0) sh[w] = wave_groups;
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
        int64_t pos = (int64_t)(sh[16] + sh[w]);
        for (int base = 0; base <= tab.cr; base += blockDim.x) {
            const int sl = base + threadIdx.x;
            KT kw;
            const bool occ = slot_key(sl, kw);
            const uint64_t m = __ballot(occ);
            if (occ) {
                const int64_t idx = pos + __popcll(m & dev::lanemask_lt());
                if (idx < capacity) dev::emit_group(os, idx, lds_to_enc64<KB>(kw, key.is_signed), tab.st + sl, (int64_t)tab.T);
                else atomicExch(&gt.meta[2], 1ull);  // output capacity exceeded: retry larger
            }
            pos += __popcll(m);
        }
    }
    gt_count_inserts(gt, ins);
}

}  // namespace agg
}  // namespace eg

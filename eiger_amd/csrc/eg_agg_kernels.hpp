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

// Whether a fixed op list can keep per-lane partial states of one group in
// registers: every op must merge associatively (no exchange, i.e. ANY).
template <class OPS>
__device__ constexpr bool privatisable()
{
    if constexpr (!OPS::kFixed) {
        return false;
    } else {
        bool ok = true;
        for (int a = 0; a < OPS::kN; ++a) ok = ok && op_class(OPS::word(a) & 15) != CLS_XCHG;
        return ok;
    }
}

// Ops held in registers by a privatising kernel (1 otherwise: unused).
template <class OPS>
__device__ constexpr int priv_ops()
{
    if constexpr (OPS::kFixed) return OPS::kN;
    else return 1;
}

// Fewest of a step's first rows (one per lane) holding lane 0's key for the
// hash kernel to keep that key in registers.
constexpr int kPrivLanes = 8;

template <int KB, class VL, class OPS>
__device__ __forceinline__ void hash_agg(const KeyDesc& key, int64_t n, const AggSpec& spec, const GTable& gt,
                                         int cr, int rep_log2, int probe_limit, int peel_min)
{
    constexpr int NV = VL::NV > 0 ? VL::NV : 1;
    extern __shared__ __align__(16) unsigned long long smem[];
    const bool use_lds = cr > 0;
    // Rows reach the HBM table as spills of cold keys (hot keys take LDS
    // slots first), where combining seldom finds partners; only without an
    // LDS stage do hot keys meet there.
    const int hbm_peel = use_lds ? 65 : peel_min;
    const int nops = OPS::n(spec);
    const int lane = dev::lane_id();
    LdsTable<KB> tab(smem, nops, use_lds ? cr : 1, use_lds ? rep_log2 : 0);
    if (use_lds) {
        const int words = LdsTable<KB>::words(nops, cr, rep_log2);
        for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
        __syncthreads();
    }
    const int r_mask = (1 << (use_lds ? rep_log2 : 0)) - 1;
    uint32_t ins = 0;  // HBM-table insertions by this lane
    dev::HotTags hot;  // groups the wave combined last
    // A key held by many lanes of a tile's first rows (lane 0's, see
    // kPrivLanes) accumulates in per-lane registers from then on; the wave's
    // partials reach its slot (LDS replica 0, else the HBM table) with one
    // wave reduction and one atomic per state when another key turns hot or
    // the full tiles end.  Fixed op lists without ANY only.
    using KT = typename LdsKey<KB>::T;
    constexpr bool kPriv = privatisable<OPS>();
    constexpr int kNP = kPriv ? priv_ops<OPS>() : 1;
    uint64_t acc[kNP];
#pragma unroll
    for (int a = 0; a < kNP; ++a) acc[a] = 0ull;
    uint64_t hot_key = ~0ull;  // wave-uniform
    bool priv_on = false;      // wave-uniform: acc holds rows of hot_key
    auto flush_priv = [&]() {
        if constexpr (kPriv) {
            if (!priv_on) return;
            uint64_t tot[kNP];
            dev::for_each_op<OPS>(spec, [&](auto a) { tot[a] = dev::wave_reduce(OPS::cls(spec, a), acc[a]); });
            if (lane == 0) {
                int sl = -1;
                if (use_lds) {
                    const KT e = (KB == 4) ? (KT)((uint32_t)hot_key ^ dev::kSign32)
                                           : (KT)dev::enc64(dev::widen_key(hot_key, KB, key.is_signed));
                    sl = tab.find(e, 0, probe_limit);
                }
                if (sl >= 0) {
                    dev::for_each_op<OPS>(spec, [&](auto a) {
                        dev::apply_atomic(OPS::cls(spec, a), &tab.st[(size_t)a * tab.T + sl], tot[a]);
                    });
                } else {
                    unsigned long long* g = gt_slot(gt, dev::enc64(dev::widen_key(hot_key, KB, key.is_signed)), ins);
                    if (g) dev::for_each_op<OPS>(spec, [&](auto a) { dev::apply_atomic(OPS::cls(spec, a), g + 1 + a, tot[a]); });
                }
            }
#pragma unroll
            for (int a = 0; a < kNP; ++a) acc[a] = 0ull;
            priv_on = false;
        }
    };

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
        if constexpr (kPriv) {
            if (peel_min <= 64) {
                // lane 0's key takes the registers when at least kPrivLanes
                // first rows of the step hold it and more than twice as many
                // as hold the current hot key (so a hot key stays put)
                const uint64_t k0 = dev::readlane64(kcur[0], 0);
                if (!(priv_on && k0 == hot_key)) {
                    const uint32_t c0 = (uint32_t)__popcll(__ballot(kcur[0] == k0));
                    const uint32_t ch = priv_on ? (uint32_t)__popcll(__ballot(kcur[0] == hot_key)) : 0u;
                    if (c0 >= (uint32_t)kPrivLanes && c0 > 2u * ch) {
                        flush_priv();
                        hot_key = k0;
                        priv_on = true;
                    }
                }
            }
        }
        uint32_t spill = 0;
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            uint64_t raw[NV], cv[NV];
#pragma unroll
            for (int c = 0; c < NV; ++c) raw[c] = vcur[c][i];
            canon_row<VL, NV>(spec, raw, cv);
            bool mine = false;  // a row of the register-held hot key
            if constexpr (kPriv) {
                if (priv_on && kcur[i] == hot_key) {
                    dev::for_each_op<OPS>(spec, [&](auto a) {
                        acc[a] = dev::combine(OPS::cls(spec, a), acc[a], dev::op_operand<OPS, NV>(spec, a, cv));
                    });
                    mine = true;
                }
            }
            if (lds_row<OPS, KB, NV>(tab, use_lds, r_mask, lane, key, spec, kcur[i], cv, !mine, probe_limit, peel_min, hot))
                spill |= 1u << i;
        }
        if (__ballot(spill != 0) != 0ull) {
            if (hbm_peel <= 64) {
#pragma unroll 1
                for (int i = 0; i < kRows; ++i) {
                    const bool pend = (spill >> i) & 1u;
                    if (__ballot(pend) == 0ull) continue;
                    uint64_t raw[NV], cv[NV];
#pragma unroll
                    for (int c = 0; c < NV; ++c) raw[c] = pick4(vcur[c], i);
                    canon_row<VL, NV>(spec, raw, cv);
                    hbm_rows<OPS, KB, NV>(gt, key, spec, pend, pick4(kcur, i), cv, hbm_peel, ins, hot);
                }
            } else {
                // every spilled row of the step finds its slot, then all of
                // them share the lane-group rounds (one request per op class
                // and row); a few rows go one per lane
                unsigned long long* g[kRows];
                uint64_t cvs[kRows][NV], m[kRows];
                int np = 0;
#pragma unroll
                for (int i = 0; i < kRows; ++i) {
                    uint64_t raw[NV];
#pragma unroll
                    for (int c = 0; c < NV; ++c) raw[c] = vcur[c][i];
                    canon_row<VL, NV>(spec, raw, cvs[i]);
                    g[i] = nullptr;
                    if ((spill >> i) & 1u) g[i] = gt_slot(gt, dev::enc64(dev::widen_key(kcur[i], KB, key.is_signed)), ins);
                    m[i] = __ballot(g[i] != nullptr);
                    np += __popcll(m[i]);
                }
                if (np >= kGroupedMinStep) {
                    hbm_grouped_multi<OPS, NV, kRows>(spec, m, g, cvs);
                } else {
#pragma unroll
                    for (int i = 0; i < kRows; ++i)
                        if (g[i]) dev::apply_ops<OPS, NV>(spec, cvs[i], [&](int a) { return g[i] + 1 + a; });
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            kcur[i] = knxt[i];
#pragma unroll
            for (int c = 0; c < NV; ++c) vcur[c][i] = vnxt[c][i];
        }
    }
    flush_priv();
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
        if (__ballot(sp) != 0ull) hbm_rows<OPS, KB, NV>(gt, key, spec, sp, kraw, cv, hbm_peel, ins, hot);
    }
    if (use_lds) flush_lds<OPS>(tab, gt, key, spec, rep_log2, r_mask, ins);
    gt_count_inserts(gt, ins);
}

constexpr int kPBMax = 1024;  // largest workgroup of the unit kernel

// Output ranges for the groups that partition units write straight out.  One
// same-address device atomic per unit serialises: about 80 claims per
// microsecond on MI355X, so 2^17 units cost ~1.6 ms.  Units therefore claim
// from one of kClaimLanes lane counters instead (lane = workgroup id mod 8,
// which is also the workgroup's XCD; each counter on its own 256-byte line).
// A lane's positions are logical: they map to chunks of `chunk` output rows,
// each claimed from the global counter by the unit whose range holds the
// chunk's first row.  A lane's last chunk is left partly used; the host then
// moves the rows stored past the group count into those holes.
constexpr int kClaimLanes = 8;
constexpr int kClaimHeadWords = 32 * (1 + kClaimLanes);
struct OutClaims {
    // [0] global position; [1 + x] lane x's last chunk (index << 40 | base);
    // [32 (1 + x)] lane x's logical position; from kClaimHeadWords on, lane
    // x's chunk bases + 1 (0 = not claimed yet), maxc per lane
    unsigned long long* words;
    int64_t chunk;  // rows per chunk: a power of two above any unit's group count
    int64_t maxc;
};
__host__ __device__ constexpr int64_t claim_words(int64_t maxc) { return kClaimHeadWords + kClaimLanes * maxc; }

// Claims m rows on lane x for the calling thread: out[0] = logical start o,
// out[1] = base of chunk o / chunk, out[2] = base of the next chunk (a unit
// spans at most two).  The unit holding a chunk's first row claims it before
// it waits for any other chunk, and that unit's lane position was taken
// before any later one's, so every wait ends.
__device__ __forceinline__ void claim_rows(const OutClaims& c, int x, unsigned long long m, unsigned long long* out)
{
    const unsigned long long o = atomicAdd(c.words + 32 * (1 + x), m);
    const uint64_t ch = (uint64_t)c.chunk;
    const uint64_t c0 = o / ch, c1 = (o + m - 1) / ch;
    unsigned long long* bases = c.words + kClaimHeadWords + (int64_t)x * c.maxc;
    unsigned long long b[2] = {0ull, 0ull};
    for (uint64_t k = c0; k <= c1; ++k) {
        if (k * ch < o) continue;
        const unsigned long long base = atomicAdd(c.words, (unsigned long long)ch);
        __hip_atomic_store(bases + k, base + 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicMax(c.words + 1 + x, (unsigned long long)((k << 40) | base));
        b[k - c0] = base;
    }
    for (uint64_t k = c0; k <= c1; ++k) {
        if (k * ch >= o) continue;
        unsigned long long v;
        while ((v = __hip_atomic_load(bases + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0ull)
            __builtin_amdgcn_s_sleep(2);
        b[k - c0] = v - 1ull;
    }
    out[0] = o;
    out[1] = b[0];
    out[2] = b[1];
}

// A unit of the partition kernel: rows [start, end) of the partitioned input.
struct Unit {
    int64_t start;
    int64_t end;
    int32_t sliced;  // the partition spans several units
    int32_t pad;
};

// LDS table of one partition unit: `cr` slots, any multiple of kBucket, so
// the table can be sized to the LDS budget rather than to a power of two (a
// 1536-slot table of five states leaves room for two 1024-thread workgroups
// per CU, a 2048-slot one for only one).  Probing reads a bucket of four key
// slots per 16-byte LDS load, from the key's home bucket on; the keys of a
// unit share their partition hash bits, so the home bucket comes from an
// independent multiplicative hash.  Layout in 8-byte words: nops state arrays
// of cr + 1 words (index cr: the side slot of the key whose LDS encoding is
// 0), then the cr keys (16-byte aligned), then the side flag.
__host__ __device__ constexpr int unit_keys_word(int nops, int cr) { return (nops * (cr + 1) + 1) & ~1; }
__host__ __device__ constexpr int unit_table_words(int nops, int cr, int kb)
{
    return unit_keys_word(nops, cr) + (cr * kb + 7) / 8 + 1;
}

template <int KB>
struct UnitTable {
    using KT = typename LdsKey<KB>::T;
    unsigned long long* st;
    KT* keys;
    unsigned long long* side;
    int cr, T;
    uint32_t nb;  // buckets

    __device__ UnitTable(unsigned long long* smem, int nops, int cr_)
        : st(smem), keys(reinterpret_cast<KT*>(smem + unit_keys_word(nops, cr_))),
          side(smem + unit_keys_word(nops, cr_) + (cr_ * KB + 7) / 8), cr(cr_), T(cr_ + 1),
          nb((uint32_t)cr_ / (uint32_t)kBucket)
    {
    }
    __device__ __forceinline__ uint32_t home(KT e) const
    {
        const uint32_t x = (KB == 4) ? (uint32_t)e : (uint32_t)e ^ (uint32_t)((uint64_t)e >> 32);
        return __umulhi(x * 0x9e3779b1u, nb);
    }
    __device__ __forceinline__ void read(uint32_t bucket, KT (&k)[kBucket]) const
    {
        const KT* p = keys + (size_t)bucket * kBucket;
        if constexpr (KB == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(p);
            k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
        } else {
            const uint4 v0 = reinterpret_cast<const uint4*>(p)[0];
            const uint4 v1 = reinterpret_cast<const uint4*>(p)[1];
            k[0] = v0.x | ((unsigned long long)v0.y << 32); k[1] = v0.z | ((unsigned long long)v0.w << 32);
            k[2] = v1.x | ((unsigned long long)v1.y << 32); k[3] = v1.z | ((unsigned long long)v1.w << 32);
        }
    }
    // Slot of key e (not 0) from bucket b on, inserted if absent; -1 when
    // `probe` buckets are full of other keys.  A key is only ever written
    // into an empty slot of its own probe sequence, by one CAS, so it lands
    // in exactly one slot.
    __device__ __forceinline__ int find(KT e, uint32_t b, int probe)
    {
        for (int p = 0; p < probe; ++p) {
            KT k[kBucket];
            read(b, k);
            int at = -1;
#pragma unroll
            for (int i = kBucket - 1; i >= 0; --i)
                if (k[i] == e) at = i;
            if (at >= 0) return (int)b * kBucket + at;
#pragma unroll
            for (int i = 0; i < kBucket; ++i) {
                if (k[i] != 0) continue;
                const KT prev = atomicCAS(&keys[b * kBucket + i], (KT)0, e);
                if (prev == 0 || prev == e) return (int)b * kBucket + i;
            }
            b = (b + 1 == nb) ? 0u : b + 1;
        }
        return -1;
    }
};

// GATHER: values are read through the row-id permutation `perm` (the
// PARTITION_GFUR layout), else straight from the partitioned columns.
// cr: LDS table slots (a multiple of kBucket); probe: buckets probed before a
// key spills to the HBM table.  OUTS: the output-list policy (dev::RtOuts or
// dev::FixOuts).  A step holds J rows per lane and no loads of the next step:
// the registers that would take stay free for a second 1024-thread
// workgroup per CU, whose loads overlap this one's aggregation.
template <int KB, class VL, class OPS, bool GATHER, class OUTS = dev::RtOuts, bool PF = false>
__device__ __forceinline__ void part_agg(const KeyDesc& key, const AggSpec& spec, const int64_t* perm,
                                         const Unit* units, const GTable& gt, const OutSpec& os,
                                         const OutClaims& claims, int64_t capacity, int cr,
                                         int probe, int peel_min)
{
    constexpr int NV = VL::NV > 0 ? VL::NV : 1;
    // rows per lane per step (independent loads in flight); PF: two rows, and
    // the next step's loads issued before this step aggregates
    constexpr int J = PF ? 2 : 4;
    using KT = typename LdsKey<KB>::T;
    extern __shared__ __align__(16) unsigned long long smem[];
    __shared__ unsigned long long sh[20];
    const int nops = OPS::n(spec);
    const int lane = dev::lane_id();
    UnitTable<KB> tab(smem, nops, cr);
    const Unit u = units[blockIdx.x];  // never empty
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
    load_step(u.start, kr, raw, valid);  // in flight while the table is cleared
    const int words = unit_table_words(nops, cr, KB);
    for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
    __syncthreads();

    uint32_t ins = 0;  // HBM-table insertions by this lane
    dev::HotTags hot;  // groups the wave combined last
    uint64_t hot_key = ~0ull;  // wave-uniform: the raw key of the last hot step
    // A hot key's rows are merged in registers instead (per-lane partial
    // states, identity 0 for every class), and the wave's partials go to the
    // key's slot with one wave reduction and one atomic per state when the
    // hot key changes or the unit ends: no LDS traffic and no same-address
    // contention for its rows.  Fixed op lists without ANY only.
    constexpr bool kPriv = privatisable<OPS>();
    constexpr int kNP = kPriv ? priv_ops<OPS>() : 1;
    uint64_t acc[kNP];
#pragma unroll
    for (int a = 0; a < kNP; ++a) acc[a] = 0ull;
    bool priv_on = false;  // wave-uniform: acc holds rows of hot_key
    auto flush_priv = [&]() {
        if constexpr (kPriv) {
            if (!priv_on) return;
            const KT e = (KB == 4) ? (KT)((uint32_t)hot_key ^ dev::kSign32)
                                   : (KT)dev::enc64(dev::widen_key(hot_key, KB, key.is_signed));
            uint64_t tot[kNP];
            dev::for_each_op<OPS>(spec, [&](auto a) { tot[a] = dev::wave_reduce(OPS::cls(spec, a), acc[a]); });
            if (lane == 0) {
                int sl = -1;
                if (e == 0) {
                    sl = cr;
                    if (__hip_atomic_load(tab.side, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0ull)
                        __hip_atomic_store(tab.side, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    sl = tab.find(e, tab.home(e), probe);
                }
                if (sl >= 0) {
                    dev::for_each_op<OPS>(spec, [&](auto a) {
                        dev::apply_atomic(OPS::cls(spec, a), &tab.st[(size_t)a * tab.T + sl], tot[a]);
                    });
                } else {
                    // the table is full: the partials go to the HBM table
                    unsigned long long* g = gt_slot(gt, dev::enc64(dev::widen_key(hot_key, KB, key.is_signed)), ins);
                    if (g) dev::for_each_op<OPS>(spec, [&](auto a) { dev::apply_atomic(OPS::cls(spec, a), g + 1 + a, tot[a]); });
                }
            }
#pragma unroll
            for (int a = 0; a < kNP; ++a) acc[a] = 0ull;
            priv_on = false;
        }
    };
    uint64_t kn[J], rn[J][NV];
    bool vn[J];
    for (int64_t b = u.start; b < u.end; b += step) {
        if constexpr (PF) {
            if (b != u.start) {
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    kr[j] = kn[j];
                    valid[j] = vn[j];
#pragma unroll
                    for (int c = 0; c < NV; ++c) raw[j][c] = rn[j][c];
                }
            }
            if (b + step < u.end) load_step(b + step, kn, rn, vn);
        } else if (b != u.start) {
            load_step(b, kr, raw, valid);
        }
        // Hot-key test, once per step: do at least peel_min lanes hold the
        // key of lane 0's first row?  Only then does the step take the
        // wave-level path (combining a group's rows before its atomics);
        // otherwise every lane probes and updates on its own, with no
        // wave-wide operation (their scalar work bounds the kernel).
        bool hot_step = false;
        if (peel_min <= 64) {
            // lane 0's key of the first row, and the key found hot last (a
            // hot key stays hot), against every row of the step
            const uint64_t k0 = dev::readlane64(kr[0], 0);
            if (__popcll(__ballot(valid[0] && kr[0] == k0)) >= (uint32_t)peel_min && !(priv_on && k0 == hot_key)) {
                flush_priv();
                hot_step = true;
                hot_key = k0;
                priv_on = kPriv;
            }
            if constexpr (!kPriv) {
#pragma unroll
                for (int j = 0; j < J; ++j)
                    hot_step = hot_step || __popcll(__ballot(valid[j] && kr[j] == hot_key)) >= (uint32_t)peel_min;
            } else {
                hot_step = false;  // the hot key's rows take the register path below
            }
        }
        if (!hot_step) {
#pragma unroll
            for (int j = 0; j < J; ++j) {
                if (!valid[j]) continue;
                if constexpr (kPriv) {
                    if (priv_on && kr[j] == hot_key) {
                        uint64_t cv[NV];
                        canon_row<VL, NV>(spec, raw[j], cv);
                        dev::for_each_op<OPS>(spec, [&](auto a) {
                            acc[a] = dev::combine(OPS::cls(spec, a), acc[a], dev::op_operand<OPS, NV>(spec, a, cv));
                        });
                        continue;
                    }
                }
                const KT e = (KB == 4) ? (KT)((uint32_t)kr[j] ^ dev::kSign32)
                                       : (KT)dev::enc64(dev::widen_key(kr[j], KB, key.is_signed));
                int slot;
                if (e == 0) {
                    slot = cr;  // the side slot
                    if (__hip_atomic_load(tab.side, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0ull)
                        __hip_atomic_store(tab.side, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    slot = tab.find(e, tab.home(e), probe);
                }
                uint64_t cv[NV];
                canon_row<VL, NV>(spec, raw[j], cv);
                if (slot >= 0) dev::apply_ops<OPS, NV>(spec, cv, [&](int a) { return &tab.st[(size_t)a * tab.T + slot]; });
                else hbm_row<OPS, KB, NV>(gt, key, spec, true, kr[j], cv, 65, ins, hot);  // the table is full: a cold key spills
            }
            continue;
        }
        // wave-level path: slots of every row, then combining and atomics
        KT e[J];
        int slot[J];
        bool side = false, miss = false;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            e[j] = (KB == 4) ? (KT)((uint32_t)kr[j] ^ dev::kSign32)
                             : (KT)dev::enc64(dev::widen_key(kr[j], KB, key.is_signed));
            slot[j] = -1;
            if (valid[j]) {
                if (e[j] == 0) {
                    slot[j] = cr;
                    side = true;
                } else {
                    KT k[kBucket];
                    const uint32_t hb = tab.home(e[j]);
                    tab.read(hb, k);
#pragma unroll
                    for (int i = kBucket - 1; i >= 0; --i)
                        if (k[i] == e[j]) slot[j] = (int)hb * kBucket + i;
                }
            }
            miss = miss || (valid[j] && slot[j] < 0);
        }
        if (miss) {
#pragma unroll
            for (int j = 0; j < J; ++j)
                if (valid[j] && slot[j] < 0) slot[j] = tab.find(e[j], tab.home(e[j]), probe);
        }
        if (side && __hip_atomic_load(tab.side, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0ull)
            __hip_atomic_store(tab.side, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        bool spill = false;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            uint64_t cv[NV];
            canon_row<VL, NV>(spec, raw[j], cv);
            const bool done = dev::wave_peel<OPS, 1>(slot[j] >= 0, (uint32_t)slot[j], hot.lds, spec, peel_min,
                                                     [&](int a) { return dev::op_operand<OPS, NV>(spec, a, cv); },
                                                     [&](uint64_t sl) { return &tab.st[(int64_t)sl]; }, (int64_t)tab.T);
            if (slot[j] >= 0 && !done)
                dev::apply_ops<OPS, NV>(spec, cv, [&](int a) { return &tab.st[(size_t)a * tab.T + slot[j]]; });
            spill = spill || (valid[j] && slot[j] < 0);
        }
        if (__ballot(spill) != 0ull) {
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const bool sp = valid[j] && slot[j] < 0;
                if (!sp) continue;
                uint64_t cv[NV];
                canon_row<VL, NV>(spec, raw[j], cv);
                hbm_row<OPS, KB, NV>(gt, key, spec, true, kr[j], cv, 65, ins, hot);
            }
        }
    }
    flush_priv();
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
        // the partition continues in other units: merge into the HBM table.
        // Each lane finds its slot's HBM slot, then groups of 8 lanes (16
        // above 8 ops) merge one slot each, lane j its state j: one request
        // per op class and slot (see hbm_grouped).
        const int nops = OPS::n(spec);
        const int lg = nops <= 8 ? 3 : 4;
        const int grp = lane >> lg, j = lane & ((1 << lg) - 1);
        for (int base = 0; base <= tab.cr; base += blockDim.x) {
            const int sl = base + threadIdx.x;
            KT kw;
            unsigned long long* g = nullptr;
            if (sl <= tab.cr && slot_key(sl, kw)) g = gt_slot(gt, lds_to_enc64<KB>(kw, key.is_signed), ins);
            const uint64_t m = __ballot(g != nullptr);
            const int np = __popcll(m);
            for (int r = 0; r < np; r += 64 >> lg) {
                const int q = r + grp;
                const int src = q < np ? nth_set_bit(m, q) : 0;
                unsigned long long* d = (unsigned long long*)__shfl((unsigned long long)(uintptr_t)g, src);
                const int ssl = __shfl(sl, src);
                if (q < np && j < nops)
                    dev::apply_atomic(OPS::cls(spec, j), d + 1 + j, (uint64_t)tab.st[(size_t)j * tab.T + ssl]);
            }
        }
    } else {
        // Written straight out: one claim per unit on its lane (see
        // OutClaims); each wave then owns a contiguous run of the unit's
        // logical range, mapped to the (at most two) chunks it falls in.
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
            if (tot) claim_rows(claims, (int)(blockIdx.x % kClaimLanes), tot, sh + 16);
            else sh[16] = sh[17] = sh[18] = 0ull;
        }
        __syncthreads();
        int64_t pos = (int64_t)(sh[16] + sh[w]);
        const int64_t cend = (int64_t)((sh[16] / (uint64_t)claims.chunk + 1) * (uint64_t)claims.chunk);
        const int64_t b0 = (int64_t)sh[17] - (cend - claims.chunk), b1 = (int64_t)sh[18] - cend;
        for (int base = 0; base <= tab.cr; base += blockDim.x) {
            const int sl = base + threadIdx.x;
            KT kw;
            const bool occ = slot_key(sl, kw);
            const uint64_t m = __ballot(occ);
            if (occ) {
                const int64_t li = pos + __popcll(m & dev::lanemask_lt());
                const int64_t idx = li + (li < cend ? b0 : b1);
                // past the capacity: the claimed total tells the host to retry larger
                if (idx < capacity)
                    dev::emit_outs<OUTS>(os, KB, idx, lds_to_enc64<KB>(kw, key.is_signed), tab.st + sl, (int64_t)tab.T);
            }
            pos += __popcll(m);
        }
    }
    gt_count_inserts(gt, ins);
}

}  // namespace agg
}  // namespace eg

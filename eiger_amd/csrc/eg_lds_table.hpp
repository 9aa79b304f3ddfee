// eg_lds_table.hpp -- building blocks of the LDS-privatised aggregation used
// by the hash and partition strategies: compiled column layouts with vector
// loads, the replicated open-addressing LDS table, canonical value
// conversion, and the per-row LDS / HBM update paths (with wave-level
// combining of hot groups).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include <type_traits>

#include "eg_atomics.hpp"
#include "eg_device.hpp"
#include "eg_hash_table.hpp"

namespace eg {
namespace lds {

constexpr int kRows = 4;      // consecutive rows per lane per tile

template <int KB>
struct LdsKey {
    using T = typename std::conditional<KB == 4, uint32_t, unsigned long long>::type;
};

// Four consecutive elements (row r multiple of 4, 16-byte aligned column).
template <int W>
__device__ __forceinline__ void vload4(const void* p, int64_t r, uint64_t (&o)[kRows])
{
    if constexpr (W == 4) {
        uint4 v = *(const uint4*)((const uint32_t*)p + r);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else {
        const uint4* q = (const uint4*)((const uint64_t*)p + r);
        uint4 a = q[0], b = q[1];
        o[0] = a.x | ((uint64_t)a.y << 32); o[1] = a.z | ((uint64_t)a.w << 32);
        o[2] = b.x | ((uint64_t)b.y << 32); o[3] = b.z | ((uint64_t)b.w << 32);
    }
}

__device__ __forceinline__ uint64_t pick4(const uint64_t (&a)[kRows], int i)
{
    return i == 0 ? a[0] : i == 1 ? a[1] : i == 2 ? a[2] : a[3];
}

// Value-column layouts compiled into the kernel.  GEN: any widths, read with
// per-element typed loads; otherwise NV columns of W0/W1 bytes read with
// 16-byte vector loads.
template <int NV_, int W0_, int W1_, bool GEN_, int C0_ = CONV_RAW, int C1_ = CONV_RAW>
struct VLayout {
    static constexpr int NV = NV_;
    static constexpr bool GEN = GEN_;
    template <int C>
    static constexpr int width() { return C == 0 ? W0_ : W1_; }
    template <int C>
    static constexpr int conv() { return C == 0 ? C0_ : C1_; }
};

// Column kinds of the compiled layouts: 0 = 8-byte, 1 = int32, 2 = uint32,
// 3 = float32.
constexpr int kind_width(int k) { return k == 0 ? 8 : 4; }
constexpr int kind_conv(int k) { return k == 1 ? CONV_SEXT32 : k == 3 ? CONV_F32 : CONV_RAW; }

template <class F>
void with_kind(int k, F&& f)
{
    switch (k) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    default: f(std::integral_constant<int, 3>{}); break;
    }
}

// Calls f(VLayout{}) with the compiled layout of spec's value columns: fixed
// widths and conversions for up to two 4/8-byte columns (when `fixed_ok`),
// the generic typed-load layout otherwise.
template <class F>
void with_vlayout(const AggSpec& spec, bool fixed_ok, F&& f)
{
    const int nv = spec.nvcols;
    bool fast = fixed_ok && nv <= 2;
    for (int c = 0; c < nv; ++c) fast = fast && (spec.cols[c].bytes == 4 || spec.cols[c].bytes == 8);
    if (fast) {
        auto kind = [&](int c) {
            const ColDesc& d = spec.cols[c];
            return d.bytes == 8 ? 0 : d.conv == CONV_SEXT32 ? 1 : d.conv == CONV_F32 ? 3 : 2;
        };
        if (nv == 0) return f(VLayout<0, 4, 4, false>{});
        if (nv == 1)
            return with_kind(kind(0), [&](auto k0) {
                constexpr int K0 = decltype(k0)::value;
                f(VLayout<1, kind_width(K0), 4, false, kind_conv(K0)>{});
            });
        return with_kind(kind(0), [&](auto k0) {
            with_kind(kind(1), [&](auto k1) {
                constexpr int K0 = decltype(k0)::value, K1 = decltype(k1)::value;
                f(VLayout<2, kind_width(K0), kind_width(K1), false, kind_conv(K0), kind_conv(K1)>{});
            });
        });
    }
    if (nv <= 2) return f(VLayout<2, 0, 0, true>{});
    if (nv <= 4) return f(VLayout<4, 0, 0, true>{});
    return f(VLayout<8, 0, 0, true>{});
}

// Values of row r (typed loads; one coalesced element per lane).
template <class VL>
__device__ __forceinline__ void load_row(const AggSpec& spec, int64_t r, bool valid,
                                         uint64_t (&v)[VL::NV > 0 ? VL::NV : 1])
{
    if constexpr (VL::GEN) {
#pragma unroll
        for (int c = 0; c < VL::NV; ++c)
            v[c] = (valid && c < spec.nvcols) ? dev::load_raw(spec.cols[c].data, spec.cols[c].bytes, r) : 0ull;
    } else {
        if constexpr (VL::NV == 0) v[0] = 0ull;
        if constexpr (VL::NV >= 1) {
            constexpr int W = VL::template width<0>();
            v[0] = !valid ? 0ull : W == 4 ? (uint64_t)((const uint32_t*)spec.cols[0].data)[r]
                                          : (uint64_t)((const uint64_t*)spec.cols[0].data)[r];
        }
        if constexpr (VL::NV >= 2) {
            constexpr int W = VL::template width<1>();
            v[1] = !valid ? 0ull : W == 4 ? (uint64_t)((const uint32_t*)spec.cols[1].data)[r]
                                          : (uint64_t)((const uint64_t*)spec.cols[1].data)[r];
        }
    }
}

template <class VL>
__device__ __forceinline__ void load_tile(const KeyDesc& key, int kb, const AggSpec& spec, int64_t r0,
                                          uint64_t (&k)[kRows], uint64_t (&v)[VL::NV > 0 ? VL::NV : 1][kRows])
{
    if constexpr (VL::GEN) {
#pragma unroll
        for (int i = 0; i < kRows; ++i) k[i] = dev::load_raw(key.data, kb, r0 + i);
#pragma unroll
        for (int c = 0; c < VL::NV; ++c)
            if (c < spec.nvcols)
#pragma unroll
                for (int i = 0; i < kRows; ++i) v[c][i] = dev::load_raw(spec.cols[c].data, spec.cols[c].bytes, r0 + i);
    } else {
        if (kb == 4) vload4<4>(key.data, r0, k);
        else vload4<8>(key.data, r0, k);
        if constexpr (VL::NV >= 1) vload4<VL::template width<0>()>(spec.cols[0].data, r0, v[0]);
        if constexpr (VL::NV >= 2) vload4<VL::template width<1>()>(spec.cols[1].data, r0, v[1]);
    }
}

// Layout of the dynamic LDS (8-byte words), T = (Cr + 1) * R slots:
//   states: nops * T words; op a of slot idx at [a * T + idx]
//   keys:   T entries of KB bytes (encoded; 0 = empty), 16-byte aligned
//   side:   R words, side slot of replica r used
// Slot (s, r) has index (s << rlog) | r; s = Cr is the side slot holding the
// key whose LDS encoding is zero.  The dynamic LDS base must be 16-byte
// aligned.
__host__ __device__ constexpr int lds_keys_word(int nops, int T) { return (nops * T + 1) & ~1; }

__host__ __device__ constexpr int lds_table_words(int nops, int cr, int rep_log2, int kb)
{
    return lds_keys_word(nops, (cr + 1) << rep_log2) + (((cr + 1) << rep_log2) * kb + 7) / 8 + (1 << rep_log2);
}

// Probing of the unreplicated table (R = 1) reads keys kBucket at a time
// from aligned 16-byte groups of slots, so a wave's probe loop -- as long as
// its worst lane's displacement -- takes a few iterations at load 1/2 instead
// of tens.  Two probe orders, fixed per table:
//   bucketed: b, b+1, .., b+3, then the next group (b = h rounded down); a
//             key is in its home group ~97% of the time at load 1/2, and the
//             first probe reads that whole group.  For tables of hash-scattered
//             keys (partition units).
//   linear:   h, h+1, ...; the first probe reads slot h alone.  Dense key
//             ranges sit exactly at home under multiplicative hashing, so the
//             single-slot read always hits.  For the hash kernel.
constexpr int kBucket = 4;

// BUCKETED: the bucketed probe order (the table is then never replicated).
template <int KB, bool BUCKETED = false>
struct LdsTable {
    using KT = typename LdsKey<KB>::T;
    static constexpr bool bucketed = BUCKETED;
    unsigned long long* st;
    KT* keys;
    unsigned long long* side;
    int T, cr, rlog;

    // cr slots per replica: a power of two when replicated, else any
    // multiple of kBucket (a 1-slot table is never probed)
    __device__ LdsTable(unsigned long long* smem, int nops, int cr_slots, int rep_log2)
    {
        cr = cr_slots;
        rlog = rep_log2;
        T = (cr + 1) << rlog;
        st = smem;
        keys = reinterpret_cast<KT*>(st + lds_keys_word(nops, T));
        side = st + lds_keys_word(nops, T) + ((size_t)T * KB + 7) / 8;
    }
    __device__ static int words(int nops, int cr, int rep_log2)
    {
        return lds_table_words(nops, cr, rep_log2, KB);
    }
    // Home slot of encoded key e: multiplicative (Fibonacci) hashing scaled to
    // the table, the high word of (e * 2^32/phi) * cr -- for cr = 2^k the top
    // k bits of the product; 8-byte keys fold their halves first.  Only probe
    // lengths depend on it.
    __device__ __forceinline__ uint32_t home(KT e) const
    {
        const uint32_t x = (KB == 4) ? (uint32_t)e : (uint32_t)e ^ (uint32_t)((uint64_t)e >> 32);
        return __umulhi(x * 0x9e3779b1u, (uint32_t)cr);
    }
    // The kBucket keys from slot b on (R = 1, b a multiple of kBucket).
    __device__ __forceinline__ void read_bucket(uint32_t b, KT (&k)[kBucket]) const
    {
        if constexpr (KB == 4) {
            const uint4 v = *reinterpret_cast<const uint4*>(keys + b);
            k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
        } else {
            const uint4 v0 = reinterpret_cast<const uint4*>(keys + b)[0];
            const uint4 v1 = reinterpret_cast<const uint4*>(keys + b)[1];
            k[0] = v0.x | ((unsigned long long)v0.y << 32); k[1] = v0.z | ((unsigned long long)v0.w << 32);
            k[2] = v1.x | ((unsigned long long)v1.y << 32); k[3] = v1.z | ((unsigned long long)v1.w << 32);
        }
    }
    // Finds or inserts encoded key e in replica r; -1 when probe_limit slots
    // are taken by other keys.  Every inserter of a key walks the same slot
    // sequence and keys never leave a slot, so a key lands in exactly one
    // slot: the first one of its sequence that was empty.
    __device__ __forceinline__ int find(KT e, int r, int probe_limit)
    {
        if (e == 0) {
            if (__hip_atomic_load(&side[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0ull)
                __hip_atomic_store(&side[r], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return (cr << rlog) | r;
        }
        uint32_t h = home(e);
        if (BUCKETED || (rlog == 0 && cr >= kBucket)) {
            if constexpr (BUCKETED) h &= ~(uint32_t)(kBucket - 1);
            for (int left = probe_limit; left > 0;) {
                const uint32_t b = h & ~(uint32_t)(kBucket - 1);
                const int i0 = (int)(h - b);
                KT k[kBucket];
                read_bucket(b, k);
                int at = -1;
#pragma unroll
                for (int i = kBucket - 1; i >= 0; --i)
                    if ((BUCKETED || i >= i0) && k[i] == e) at = i;
                if (at >= 0) return (int)b + at;
#pragma unroll
                for (int i = 0; i < kBucket; ++i) {
                    if ((!BUCKETED && i < i0) || k[i] != 0) continue;
                    const KT prev = atomicCAS(&keys[b + i], (KT)0, e);
                    if (prev == 0 || prev == e) return (int)b + i;
                }
                left -= kBucket - i0;
                h = b + kBucket >= (uint32_t)cr ? 0u : b + kBucket;
            }
            return -1;
        }
        for (int p = 0; p < probe_limit; ++p) {
            const int idx = (int)(h << rlog) | r;
            KT cur = __hip_atomic_load(&keys[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (cur == e) return idx;
            if (cur == 0) {
                KT prev = atomicCAS(&keys[idx], (KT)0, e);
                if (prev == 0 || prev == e) return idx;
            }
            h = h + 1 == (uint32_t)cr ? 0u : h + 1;
        }
        return -1;
    }
};

template <int KB>
__device__ __forceinline__ uint64_t lds_to_enc64(typename LdsKey<KB>::T kw, int is_signed)
{
    if (KB == 4) return dev::enc64(dev::widen_key((uint32_t)kw ^ dev::kSign32, 4, is_signed));
    return (uint64_t)kw;
}

// Canonical values of a row's columns (raw -> 64-bit per ConvCode); the
// compiled layouts know each column's conversion.
template <class VL, int NV>
__device__ __forceinline__ void canon_row(const AggSpec& spec, const uint64_t (&raw)[NV], uint64_t (&cv)[NV])
{
    if constexpr (VL::GEN) {
#pragma unroll
        for (int c = 0; c < NV; ++c) cv[c] = c < spec.nvcols ? dev::canon(raw[c], spec.cols[c].conv) : 0ull;
    } else {
        if constexpr (VL::NV >= 1) cv[0] = dev::canon_c<VL::template conv<0>()>(raw[0]);
        if constexpr (VL::NV >= 2) cv[1] = dev::canon_c<VL::template conv<1>()>(raw[1]);
        if constexpr (VL::NV == 0) cv[0] = 0ull;
    }
}

// One row into the LDS table; returns true when it needs the HBM table.
template <class OPS, int KB, int NV, bool BK>
__device__ __forceinline__ bool lds_row(LdsTable<KB, BK>& tab, bool use_lds, int r_mask, int lane, const KeyDesc& key,
                                        const AggSpec& spec, uint64_t kraw, const uint64_t (&cv)[NV],
                                        bool valid, int probe_limit, int peel_min, dev::HotTags& hot)
{
    using KT = typename LdsKey<KB>::T;
    int slot = -1;
    if (use_lds) {
        // first probe inline: once the table holds the groups it almost always hits
        KT e = 0;
        uint32_t h = 0;
        bool hit = false;
        if (valid) {
            e = (KB == 4) ? (KT)((uint32_t)kraw ^ dev::kSign32) : (KT)dev::enc64(dev::widen_key(kraw, KB, key.is_signed));
            h = tab.home(e);
            if constexpr (BK) {
                // the whole home group in one read
                const uint32_t b = h & ~(uint32_t)(kBucket - 1);
                KT k[kBucket];
                tab.read_bucket(b, k);
                int at = -1;
#pragma unroll
                for (int i = kBucket - 1; i >= 0; --i)
                    if (k[i] == e) at = i;
                hit = e != 0 && at >= 0;
                if (hit) slot = (int)b + at;
            } else {
                const int idx = (int)(h << tab.rlog) | (lane & r_mask);
                const KT cur = __hip_atomic_load(&tab.keys[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                hit = e != 0 && cur == e;
                if (hit) slot = idx;
            }
        }
        if (__ballot(valid && !hit) != 0ull) {
            if (valid && !hit) slot = tab.find(e, lane & r_mask, probe_limit);
        }
    }
    bool done = false;
    if (use_lds && r_mask == 0 && peel_min <= 64) {
        done = dev::wave_peel<OPS>(slot >= 0, (uint32_t)slot, hot.lds, spec, peel_min,
                                   [&](int a) { return dev::op_operand<OPS, NV>(spec, a, cv); },
                              [&](uint64_t s) { return &tab.st[(int64_t)s]; }, (int64_t)tab.T);
    }
    if (slot >= 0 && !done) dev::apply_ops<OPS, NV>(spec, cv, [&](int a) { return &tab.st[(size_t)a * tab.T + slot]; });
    return valid && slot < 0;
}

// One row into the HBM table (called by every lane of the wave).
template <class OPS, int KB, int NV>
__device__ __forceinline__ void hbm_row(const GTable& gt, const KeyDesc& key, const AggSpec& spec, bool pending,
                                        uint64_t kraw, const uint64_t (&cv)[NV], int peel_min, uint32_t& ins,
                                        dev::HotTags& hot)
{
    const uint64_t e64 = dev::enc64(dev::widen_key(kraw, KB, key.is_signed));
    bool done = false;
    if (peel_min <= 64)
        done = dev::wave_peel<OPS>(pending, e64, hot.hbm, spec, peel_min,
                                   [&](int a) { return dev::op_operand<OPS, NV>(spec, a, cv); },
                              [&](uint64_t t) {
                                  unsigned long long* g = gt_slot(gt, t, ins);
                                  return g ? g + 1 : (unsigned long long*)nullptr;
                              },
                              (int64_t)1);
    if (pending && !done) {
        unsigned long long* g = gt_slot(gt, e64, ins);
        if (g) dev::apply_ops<OPS, NV>(spec, cv, [&](int a) { return g + 1 + a; });
    }
}

// Fewest pending rows of a wave that hbm_rows redistributes over lane groups
// (one row per lane), and that a hash-kernel step does (four rows per lane,
// whose rows share the rounds).
constexpr int kGroupedMin = 16;
constexpr int kGroupedMinStep = 6;

// Index of the q-th set bit of m (0 <= q < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int q)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll((m >> pos) & ((1ull << w) - 1ull));
        if (q >= c) {
            q -= c;
            pos += w;
        }
    }
    return pos;
}

// Operand of op j (a per-lane index; 0 past the last op) for row values cv.
template <class OPS, int NV>
__device__ __forceinline__ uint64_t lane_operand(const AggSpec& spec, int j, const uint64_t (&cv)[NV])
{
    uint64_t val = 0ull;
    dev::for_each_op<OPS>(spec, [&](auto a) {
        const uint64_t x = dev::op_operand<OPS, NV>(spec, a, cv);
        if ((int)a == j) val = x;
    });
    return val;
}

// The rows of lanes m (slot g, values cv) applied by groups of lanes, one row
// a group and one op a lane.
template <class OPS, int NV>
__device__ __forceinline__ void hbm_grouped(const AggSpec& spec, uint64_t m, unsigned long long* g,
                                                      const uint64_t (&cv)[NV])
{
    const int nops = OPS::n(spec);
    const int lg = nops <= 8 ? 3 : 4;  // lanes per row
    const int lane = dev::lane_id();
    const int grp = lane >> lg, j = lane & ((1 << lg) - 1);
    const int np = __popcll(m);
    for (int base = 0; base < np; base += 64 >> lg) {
        const int q = base + grp;
        const int src = q < np ? nth_set_bit(m, q) : 0;
        unsigned long long* d = (unsigned long long*)__shfl((unsigned long long)(uintptr_t)g, src);
        uint64_t rv[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) rv[c] = __shfl((unsigned long long)cv[c], src);
        const uint64_t val = lane_operand<OPS, NV>(spec, j, rv);
        if (q < np && j < nops) dev::apply_atomic(OPS::cls(spec, j), d + 1 + j, val);
    }
}

// Like hbm_grouped for rows spread over R registers per lane: row k of lane l
// is pending when bit l of m[k] is set, with slot g[k] and values cv[k].
// The rows of all R sets share the rounds, so a wave whose lanes each spill
// a row or two still fills its groups.
template <class OPS, int NV, int R>
__device__ __forceinline__ void hbm_grouped_multi(const AggSpec& spec, const uint64_t (&m)[R],
                                                  unsigned long long* const (&g)[R], const uint64_t (&cv)[R][NV])
{
    const int nops = OPS::n(spec);
    const int lg = nops <= 8 ? 3 : 4;  // lanes per row
    const int lane = dev::lane_id();
    const int grp = lane >> lg, j = lane & ((1 << lg) - 1);
    int pre[R + 1];  // wave-uniform prefix counts
    pre[0] = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) pre[k + 1] = pre[k] + __popcll(m[k]);
    const int np = pre[R];
    for (int base = 0; base < np; base += 64 >> lg) {
        const int q = base + grp;
        int r = 0, qr = q;
        uint64_t mr = m[0];
#pragma unroll
        for (int k = 1; k < R; ++k)
            if (q >= pre[k]) {
                r = k;
                mr = m[k];
                qr = q - pre[k];
            }
        const int src = q < np ? nth_set_bit(mr, qr) : 0;
        unsigned long long* d = nullptr;
        uint64_t rv[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c) rv[c] = 0ull;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            unsigned long long* dk = (unsigned long long*)__shfl((unsigned long long)(uintptr_t)g[k], src);
            if (r == k) d = dk;
#pragma unroll
            for (int c = 0; c < NV; ++c) {
                const uint64_t x = __shfl((unsigned long long)cv[k][c], src);
                if (r == k) rv[c] = x;
            }
        }
        const uint64_t val = lane_operand<OPS, NV>(spec, j, rv);
        if (q < np && j < nops) dev::apply_atomic(OPS::cls(spec, j), d + 1 + j, val);
    }
}

// The wave's pending rows into the HBM table; every lane of the wave must be
// at the call.  Each lane finds its own row's slot, then groups of 8 lanes
// (16 for more than 8 ops) take one row each, lane j of a group applying op
// j, so all of a row's states of one op class reach its slot's cache line as
// a single request.  Device atomics on MI355X resolve at a fixed ~23 G
// requests/s -- the same at agent or workgroup scope, with one table per XCD,
// or with a table that fits L2 -- and a request may carry every lane that
// hits its line: a row of five states costs three requests (one per class)
// instead of five.  With wave combining on (peel_min <= 64) the rows of keys
// shared by many lanes are combined first, as in hbm_row.
template <class OPS, int KB, int NV>
__device__ __forceinline__ void hbm_rows(const GTable& gt, const KeyDesc& key, const AggSpec& spec, bool pending,
                                         uint64_t kraw, const uint64_t (&cv)[NV], int peel_min, uint32_t& ins,
                                         dev::HotTags& hot)
{
    const uint64_t e64 = dev::enc64(dev::widen_key(kraw, KB, key.is_signed));
    if (peel_min <= 64 && dev::wave_peel<OPS>(pending, e64, hot.hbm, spec, peel_min,
                                              [&](int a) { return dev::op_operand<OPS, NV>(spec, a, cv); },
                                              [&](uint64_t t) {
                                                  unsigned long long* g = gt_slot(gt, t, ins);
                                                  return g ? g + 1 : (unsigned long long*)nullptr;
                                              },
                                              (int64_t)1))
        pending = false;
    unsigned long long* g = nullptr;
    if (pending) g = gt_slot(gt, e64, ins);
    const uint64_t m = __ballot(g != nullptr);
    if (__popcll(m) < kGroupedMin) {
        // a few rows: the redistribution costs more instructions than the
        // requests it saves
        if (g) dev::apply_ops<OPS, NV>(spec, cv, [&](int a) { return g + 1 + a; });
        return;
    }
    hbm_grouped<OPS, NV>(spec, m, g, cv);
}

}  // namespace lds
}  // namespace eg

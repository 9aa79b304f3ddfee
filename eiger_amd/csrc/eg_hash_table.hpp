// eg_hash_table.hpp -- the HBM aggregation table shared by the hash and
// partition strategies: open addressing with linear probing over slots of
// W = 2^k 64-bit words [encoded key | states...], one cache line or less per
// slot.  Keys move 0 -> key exactly once (compare-and-swap), so a stale read
// of an empty slot only costs one extra CAS.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "eg_device.hpp"
#ifndef __HIPCC_RTC__
#include "eg_internal.hpp"
#endif

namespace eg {

struct GTable {
    unsigned long long* base;   // (mask+1) * W words
    unsigned long long* side;   // W words: states of the key whose encoding is 0
    unsigned long long* meta;   // [0] occupied slots, [1] side used, [2] overflow
    uint64_t mask;
    int W;
};

// Longest probe sequence before the table counts as full.  Tables are sized
// for a load of at most 1/2, where linear probing stays far below this; a
// table that reaches it is (nearly) full and is rebuilt larger.
constexpr uint64_t kMaxProbes = 2048;

// Slot of encoded key `enc` (inserted if absent); nullptr when the table is
// full (meta[2] is then set and the caller retries with a larger table).
// `ins` counts this lane's insertions; gt_count_inserts publishes the count.
__device__ __forceinline__ unsigned long long* gt_slot(const GTable& gt, uint64_t enc, uint32_t& ins)
{
    if (enc == 0) {
        if (__hip_atomic_load(&gt.meta[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            atomicExch(&gt.meta[1], 1ull);
        return gt.side;
    }
    uint64_t h = dev::fmix64(enc) & gt.mask;
    const uint64_t limit = gt.mask < kMaxProbes ? gt.mask + 1 : kMaxProbes;
    for (uint64_t p = 0; p < limit; ++p) {
        unsigned long long* slot = gt.base + h * (uint64_t)gt.W;
        uint64_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == enc) return slot;
        if (cur == 0) {
            uint64_t prev = atomicCAS(slot, 0ull, (unsigned long long)enc);
            if (prev == 0) {
                ++ins;
                return slot;
            }
            if (prev == enc) return slot;
        }
        h = (h + 1) & gt.mask;
    }
    atomicExch(&gt.meta[2], 1ull);
    return nullptr;
}

// Every group of the table (slots, then the side slot) handed to
// emit(index, enc, states), indices dense from *counter on.  A thread looks
// at R slots per round and a block claims its round's groups with one device
// atomic: same-address atomics serialise (~80 per microsecond), so one claim
// per 256 slots cost a 3M-slot table ~0.2 ms.  Large tables take
// kCompactSlots per thread; small ones one (more workgroups, shorter chains).
constexpr int kCompactSlots = 8;
constexpr uint64_t kCompactWideSlots = 1u << 16;  // tables from this size take kCompactSlots
template <int R, class EMIT>
__device__ __forceinline__ void compact_groups(const GTable& gt, unsigned long long* counter, unsigned long long* sh,
                                               EMIT emit)
{
    const uint64_t total = gt.mask + 2;  // slots + side slot
    const uint64_t chunk = (uint64_t)blockDim.x * R;
    for (uint64_t base = (uint64_t)blockIdx.x * chunk; base < total; base += (uint64_t)gridDim.x * chunk) {
        uint64_t enc[R];
        uint32_t occ = 0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const uint64_t s = base + (uint64_t)i * blockDim.x + threadIdx.x;
            enc[i] = 0;
            if (s <= gt.mask) {
                enc[i] = gt.base[s * (uint64_t)gt.W];
                if (enc[i] != 0) occ |= 1u << i;
            } else if (s == gt.mask + 1 && gt.meta[1] != 0) {
                occ |= 1u << i;
            }
        }
        int64_t idx = dev::block_claim_n(counter, (uint32_t)__popc(occ), sh);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (!((occ >> i) & 1u)) continue;
            const uint64_t s = base + (uint64_t)i * blockDim.x + threadIdx.x;
            const unsigned long long* slot = s <= gt.mask ? gt.base + s * (uint64_t)gt.W : gt.side;
            emit(idx++, enc[i], slot + 1);
        }
    }
}

// Adds the wave's insertion count to meta[0] with one atomic.  Every lane of
// the wave calls it (at the end of a kernel).
__device__ __forceinline__ void gt_count_inserts(const GTable& gt, uint32_t ins)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) ins += __shfl_xor(ins, off);
    if (dev::lane_id() == 0 && ins != 0u) atomicAdd(&gt.meta[0], (unsigned long long)ins);
}

#ifndef __HIPCC_RTC__  // host side
int table_words_per_slot(int nops);

// Owns a zeroed HBM table of `slots` (power of two) slots.
struct HbmTable {
    DevBuf buf;
    GTable gt;
    HbmTable(int64_t slots, int W, hipStream_t s);
    void read_meta(int64_t* out3, hipStream_t s);   // synchronises s
    // Appends every occupied slot to the outputs at positions from *counter.
    void compact(const OutSpec& os, unsigned long long* counter, hipStream_t s);
};

// LDS table geometry of the hash kernel.
struct HashLayout {
    int cr = -1;          // slots per replica, -1 = no LDS stage
    int rep_log2 = 0;     // replicas (log2)
    size_t lds_bytes = 0;
    int probe_limit = 16;
    int peel_min = 8;
};

HashLayout plan_hash_layout(int nops, int key_bytes, int64_t group_bound, const HashTuning& t);
#endif

}  // namespace eg

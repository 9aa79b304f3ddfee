// eg_atomics.hpp -- the wave64 atomics layer.
//
// The reference merges every row into its group with one device atomic per
// aggregation (atomics/atomics.cuh:10-126, groupby_hash.cu:77-99), which
// serialises on hot keys.  Here a wavefront first looks for lanes that update
// the same target: a ballot over the first pending lanes' tags finds them, a
// cross-lane DPP reduction combines their operands, and one lane issues a single
// atomic per aggregation state.  Lanes left over (distinct targets) go
// straight to the memory atomic.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "eg_device.hpp"

namespace eg {
namespace dev {

template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf, false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

__device__ __forceinline__ uint32_t readlane32(uint32_t v, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane)
{
    return (uint64_t)readlane32((uint32_t)v, lane) | ((uint64_t)readlane32((uint32_t)(v >> 32), lane) << 32);
}

// Reduction of a 64-bit state over all 64 lanes (every lane of the wave must
// execute it with the full wave active; lanes outside the group pass the
// class identity 0).  Within each 16-lane row: DPP quad_perm [1,0,3,2] and
// [2,3,0,1] (xor 1, xor 2), row_half_mirror (other quad of the 8) and
// row_mirror (other half of the row) -- every lane then holds its row's
// total; the four row totals are read out and combined, so the result is
// wave-uniform.  No LDS traffic, unlike a shuffle butterfly.
__device__ __forceinline__ uint64_t wave_reduce(int32_t cls, uint64_t v)
{
    v = combine(cls, v, dpp64<0xb1>(v));
    v = combine(cls, v, dpp64<0x4e>(v));
    v = combine(cls, v, dpp64<0x141>(v));
    v = combine(cls, v, dpp64<0x140>(v));
    const uint64_t a = combine(cls, readlane64(v, 0), readlane64(v, 16));
    const uint64_t b = combine(cls, readlane64(v, 32), readlane64(v, 48));
    return combine(cls, a, b);
}

constexpr int kPeelTries = 2;

__device__ __forceinline__ uint32_t read_leader(uint32_t tag, int leader) { return readlane32(tag, leader); }
__device__ __forceinline__ uint64_t read_leader(uint64_t tag, int leader) { return readlane64(tag, leader); }

// Wave-aggregated update of the states of one group.
//   pending   : this lane still has a row to merge
//   tag       : equal tags <=> same target group (32- or 64-bit)
//   hot       : wave-uniform tag of the group last combined (in/out); tried
//               first, since a hot group stays hot -- a hint only, any value
//               is correct
//   operand(a): 64-bit operand of op a for this lane's row (all lanes)
//   base(tag) : state pointer of op 0 of group `tag` (leader lane only);
//               op a lives at base + a * stride
// Returns true when this lane's row was merged here.  Every lane of the wave
// must call it, with the full wave active and the same min_lanes.  Otherwise
// the leader is the lowest pending lane; its tag is read with a scalar lane
// read.
// TRIES: leaders tried after the hot tag (each costs a lane read, a ballot
// and a population count of scalar work per call).
template <class OPS, int TRIES = kPeelTries, class Tag, class OperandFn, class BaseFn>
__device__ __forceinline__ bool wave_peel(bool pending, Tag tag, Tag& hot, const AggSpec& spec, int min_lanes,
                                          OperandFn operand, BaseFn base, int64_t stride)
{
    uint64_t pend = __ballot(pending);
    bool done = false;
    const int lane = lane_id();
    // merges the rows of lanes m (all with tag ltag) through lane `leader`
    auto combine_group = [&](uint64_t m, int leader, Tag ltag) {
        const bool mine = (m >> lane) & 1ull;
        decltype(base(ltag)) p = nullptr;
        if (lane == leader) p = base(ltag);
        for_each_op<OPS>(spec, [&](auto a) {
            const int32_t cls = OPS::cls(spec, a);
            uint64_t v = operand(a);
            if (cls != CLS_XCHG) v = wave_reduce(cls, mine ? v : 0ull);
            if (lane == leader && p) apply_atomic(cls, p + (int)a * stride, v);
        });
        done = done || mine;
    };
    if (pend == 0) return false;
    {
        const uint64_t m = __ballot(tag == hot) & pend;
        if (__popcll(m) >= min_lanes) {
            pend &= ~m;
            combine_group(m, __ffsll((unsigned long long)m) - 1, hot);
        }
    }
#pragma unroll
    for (int it = 0; it < TRIES; ++it) {
        if (pend == 0) break;
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const Tag ltag = read_leader(tag, leader);
        const uint64_t m = __ballot(tag == ltag) & pend;
        pend &= ~m;
        if (__popcll(m) < min_lanes) continue;
        combine_group(m, leader, ltag);
        hot = ltag;
    }
    return done;
}

// Wave-uniform hints of the groups last combined (see wave_peel).
struct HotTags {
    uint32_t lds = 0xffffffffu;
    uint64_t hbm = ~0ull;
};

}  // namespace dev
}  // namespace eg

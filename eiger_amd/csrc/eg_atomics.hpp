// eg_atomics.hpp -- the wave64 atomics layer.
//
// The reference merges every row into its group with one device atomic per
// aggregation (atomics/atomics.cuh:10-126, groupby_hash.cu:77-99), which
// serialises on hot keys.  Here a wavefront first looks for lanes that update
// the same target: a ballot over the first pending lanes' tags finds them, a
// cross-lane butterfly reduces their operands, and one lane issues a single
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

// Butterfly reduction of a 64-bit state over all 64 lanes (every lane of the
// wave must execute it; lanes outside the group pass the class identity 0).
__device__ __forceinline__ uint64_t wave_reduce(int32_t cls, uint64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint64_t o = __shfl_xor((unsigned long long)v, off);
        v = combine(cls, v, o);
    }
    return v;
}

constexpr int kPeelTries = 2;

// Wave-aggregated update of the states of one group.
//   pending   : this lane still has a row to merge
//   tag       : equal tags <=> same target group
//   operand(a): 64-bit operand of op a for this lane's row (all lanes)
//   base(tag) : state pointer of op 0 of group `tag` (leader lane only);
//               op a lives at base + a * stride
// Returns true when this lane's row was merged here.  Every lane of the wave
// must call it with the same nops and min_lanes.
template <class OPS, class OperandFn, class BaseFn>
__device__ __forceinline__ bool wave_peel(bool pending, uint64_t tag, const AggSpec& spec, int min_lanes,
                                          OperandFn operand, BaseFn base, int64_t stride)
{
    uint64_t pend = __ballot(pending);
    bool done = false;
    const int lane = lane_id();
#pragma unroll
    for (int it = 0; it < kPeelTries; ++it) {
        if (pend == 0) break;
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const uint64_t ltag = __shfl((unsigned long long)tag, leader);
        const uint64_t m = __ballot(pending && tag == ltag) & pend;
        pend &= ~m;
        if (__popcll(m) < min_lanes) continue;
        const bool mine = (m >> lane) & 1ull;
        decltype(base(0ull)) p = nullptr;
        if (lane == leader) p = base(ltag);
        for_each_op<OPS>(spec, [&](auto a) {
            const int32_t cls = OPS::cls(spec, a);
            uint64_t v = operand(a);
            if (cls != CLS_XCHG) v = wave_reduce(cls, mine ? v : 0ull);
            if (lane == leader && p) apply_atomic(cls, p + (int)a * stride, v);
        });
        done = done || mine;
    }
    return done;
}

}  // namespace dev
}  // namespace eg

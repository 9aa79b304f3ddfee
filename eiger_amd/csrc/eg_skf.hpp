// eg_skf.hpp -- fusion of multi-column integer group keys into one 64-bit key
// (skf.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "eg_internal.hpp"

namespace eg {

constexpr int kMaxFusedCols = 8;

struct FusedKeys {
    DevBuf fused;                  // rows x uint64 fused keys
    int ncols = 0;
    int64_t rows = 0;
    int total_bits = 0;
    int32_t type[kMaxFusedCols] = {};
    int bytes[kMaxFusedCols] = {};
    int bits[kMaxFusedCols] = {};
    int shift[kMaxFusedCols] = {};
    uint64_t min[kMaxFusedCols] = {};  // of the order-preserving (twiddled) values
};

// Fuses the key columns; false when they need more than 64 bits (nothing is
// allocated then).  Synchronises s (one range reduction per column).
bool fuse_keys(const ViewIn* keys, int64_t ncols, hipStream_t s, FusedKeys& f);

// The original key columns of `groups` fused keys (device), one buffer per
// column of the original widths.
std::vector<DevBuf> unfuse_keys(const FusedKeys& f, const void* fused, int64_t groups, hipStream_t s);

}  // namespace eg

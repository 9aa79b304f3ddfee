// eg_skf.hpp -- fusion of multi-column integer group keys into one 64-bit key
// (skf.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "eg_internal.hpp"

namespace eg {

// Bit fields one fusion pass writes; a stage that holds this many is folded
// into its dense rank before the next column.
constexpr int kMaxFuseItems = 9;

// One bit field of a fused key.
struct FuseItem {
    int src = 0;      // key column, or -1: the dense rank of the previous stage's fused key
    int ranked = 0;   // key column replaced by the dense rank of its range-reduced value
    int bits = 0;
    int shift = 0;
};

// Stage s fuses its items into one 64-bit value per row; the last stage's
// value is the fused key.  Every earlier stage's value is replaced by its
// dense rank in the next stage (first item), and `table` maps that rank back.
struct FuseStage {
    int nitems = 0;
    FuseItem items[kMaxFuseItems];
    DevBuf table;  // distinct fused values of this stage, ascending (all but the last stage)
};

struct FusedKeys {
    DevBuf fused;                  // rows x uint64 fused keys
    int ncols = 0;
    int64_t rows = 0;
    int total_bits = 0;            // bits of the fused key
    // per key column
    std::vector<int32_t> orig_type;  // the caller's type (float columns are fused in mapped form)
    std::vector<int32_t> type;
    std::vector<int> bytes;
    std::vector<int> bits;          // of the range-reduced column
    std::vector<uint64_t> min;      // of the order-preserving (twiddled) values
    std::vector<DevBuf> col_table;  // ranked columns: distinct range-reduced values, ascending
    std::vector<FuseStage> stages;
};

// Fuses the key columns.  Up to nine columns whose range-reduced widths sum
// to at most 64 bits take one stage; wider or longer key sets are folded
// through dense ranks (inputs below 2^32 rows).  Synchronises s (one range
// reduction per column, one distinct count per rank).
void fuse_keys(const ViewIn* keys, int64_t ncols, hipStream_t s, FusedKeys& f);

// The original key columns of `groups` fused keys (device), one buffer per
// column of the original widths.
std::vector<DevBuf> unfuse_keys(const FusedKeys& f, const void* fused, int64_t groups, hipStream_t s);

}  // namespace eg

// pending.hip -- entry points whose kernels are not built yet; they fail
// loudly with EG_ERR_UNSUPPORTED instead of falling back to anything.
#include <hip/hip_runtime.h>

#include "eg_internal.hpp"

namespace eg {
Result run_sort(const Plan&, const eg_groupby_hint&, bool, hipStream_t)
{
    raise(EG_ERR_UNSUPPORTED, "sort group-by not built yet");
}
Result run_partition(const Plan&, int64_t, bool, hipStream_t)
{
    raise(EG_ERR_UNSUPPORTED, "partition group-by not built yet");
}
}  // namespace eg

extern "C" {
#define EG_PENDING return eg::guard([&] { eg::raise(EG_ERR_UNSUPPORTED, "not built yet"); })
int eg_sort_column_pair(eg_column_view, eg_column_view, int32_t, int32_t, int32_t, void*, size_t, void*,
                        eg_column_view*, eg_column_view*) { EG_PENDING; }
int eg_sort_column(eg_column_view, int32_t, int32_t, int32_t, void*, size_t, void*, eg_column_view*) { EG_PENDING; }
int eg_argsort(eg_column_view, int32_t, int32_t, int32_t, void*, size_t, void*, eg_column_view*,
               eg_column_view*) { EG_PENDING; }
int eg_partition_column_pair(eg_column_view, eg_column_view, int32_t, int32_t, void*, size_t, void*,
                             eg_column_view*, eg_column_view*) { EG_PENDING; }
int eg_partition_column(eg_column_view, int32_t, int32_t, void*, size_t, void*, eg_column_view*) { EG_PENDING; }
int eg_arg_partition(eg_column_view, int32_t, int32_t, void*, size_t, void*, eg_column_view*,
                     eg_column_view*) { EG_PENDING; }
int eg_compute_partition_offsets(eg_column_view, int32_t, int32_t, void*, eg_column_view*) { EG_PENDING; }
}

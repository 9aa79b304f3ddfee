// sort_api.hip -- the sort and partition sub-operators at the C ABI
// (gamax::sort::{sort_column_pair, sort_column, argsort},
//  gamax::partition::{partition_column_pair, partition_column, arg_partition,
//  compute_partition_offsets}; operators/sort/sort.cu:115-215,
//  operators/partition/partition.cu:11-171).
//
// Sorting is stable on bits [begin_bit, end_bit) of the radix-ordered key
// (sign bit flipped for signed integers, the usual transform for floats), as
// cub::DeviceRadixSort orders them.  A partition is the same stable sort,
// ascending (partition.cu:24-43).  The temp/temp_bytes arguments are accepted
// for signature compatibility; scratch comes from the library's allocator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "eg_device.hpp"
#include "eg_internal.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

const void* col_ptr(const eg_column_view& c, int bytes)
{
    if (c.offset < 0 || c.size < 0) raise(EG_ERR_INVALID, "bad column view");
    if (c.size > 0 && !c.data) raise(EG_ERR_INVALID, "null column data");
    return c.data ? (const char*)c.data + c.offset * bytes : nullptr;
}

int key_bytes(int32_t type)
{
    const int b = type_size(type);
    if (!b || type == EG_BOOL8) raise(EG_ERR_UNSUPPORTED, "unsupported sort key type");
    return b;
}

void set_out(eg_column_view* out, void* data, int64_t n, int32_t type)
{
    out->data = data;
    out->size = n;
    out->type = type;
    out->reserved = 0;
    out->offset = 0;
}

// Sorts key (+ optional payload) and hands the results to the caller.
void sort_impl(const eg_column_view& key, const eg_column_view* val, bool rowid, int begin_bit, int end_bit,
               bool desc, hipStream_t s, eg_column_view* out_key, eg_column_view* out_val)
{
    const int kb = key_bytes(key.type);
    const int64_t n = key.size;
    if (end_bit <= 0) end_bit = 8 * kb;
    if (begin_bit < 0 || begin_bit > end_bit || end_bit > 8 * kb) raise(EG_ERR_INVALID, "bad bit range");
    RadixKeys rk{};
    rk.data = col_ptr(key, kb);
    rk.bytes = kb;
    rk.type = key.type;
    rk.mode = DIG_KEY;
    rk.desc = desc ? 1 : 0;
    SortBuffers b;
    b.keys_in = rk.data;
    DevBuf ka((size_t)n * kb, s), kbb((size_t)n * kb, s);
    b.keys_a = ka.get();
    b.keys_b = kbb.get();
    DevBuf pa, pb;
    int32_t ptype = EG_INT64;
    if (val || rowid) {
        int pbytes = 8;
        const void* pin = nullptr;
        if (val) {
            ptype = val->type;
            pbytes = type_size(val->type);
            if (!pbytes) raise(EG_ERR_UNSUPPORTED, "unsupported value type");
            if (val->size != n) raise(EG_ERR_INVALID, "key and value columns differ in size");
            pin = col_ptr(*val, pbytes);
        }
        pa = DevBuf((size_t)n * pbytes, s);
        pb = DevBuf((size_t)n * pbytes, s);
        b.pay_in.push_back(Payload{pin, nullptr, pbytes, rowid ? 1 : 0});
        b.pay_a.push_back(pa.get());
        b.pay_b.push_back(pb.get());
    }
    const void* kres = nullptr;
    std::vector<const void*> pres;
    radix_sort(rk, b, n, begin_bit, end_bit, s, &kres, pres);
    EG_HIP(hipStreamSynchronize(s));
    set_out(out_key, kres == ka.get() ? ka.release() : kbb.release(), n, key.type);
    if (out_val) {
        if (!pres.empty()) set_out(out_val, pres[0] == pa.get() ? pa.release() : pb.release(), n, ptype);
        else set_out(out_val, nullptr, 0, ptype);
    }
}

// Bucket counts of (key >> begin_bit) & mask.  Up to kLdsBins buckets are
// counted in LDS per block first (one device atomic per bucket per block), so
// few buckets never serialise on device atomics.
constexpr int kLdsBins = 8192;

__global__ void radix_count_kernel(const void* keys, int bytes, int64_t n, int begin_bit, uint64_t mask,
                                   unsigned long long* counts)
{
    __shared__ uint32_t h[kLdsBins];
    const bool lds = mask < (uint64_t)kLdsBins;
    if (lds) {
        for (int i = threadIdx.x; i <= (int)mask; i += blockDim.x) h[i] = 0u;
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t d = (dev::load_raw(keys, bytes, i) >> begin_bit) & mask;
        if (lds) atomicAdd(&h[d], 1u);
        else atomicAdd(&counts[d], 1ull);
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i <= (int)mask; i += blockDim.x)
            if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
    }
}

__global__ void inclusive_offsets_kernel(const unsigned long long* counts, int64_t P, int64_t n, int64_t* offs)
{
    // single block: serial prefix per chunk of 1024 threads with a carry
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < P; base += blockDim.x) {
        const int64_t i = base + threadIdx.x;
        long long v = i < P ? (long long)counts[i] : 0;
        // block inclusive scan (Hillis-Steele on 64-bit through shuffles + LDS)
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        long long x = v;
        for (int off = 1; off < 64; off <<= 1) {
            long long y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        __shared__ long long ws[16];
        if (lane == 63) ws[w] = x;
        __syncthreads();
        long long add = carry;
        for (int k = 0; k < w; ++k) add += ws[k];
        if (i < P) offs[i] = x + add;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = add + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) offs[P] = n;
}

}  // namespace
}  // namespace eg

extern "C" {

int eg_sort_column_pair(eg_column_view key, eg_column_view val, int32_t begin_bit, int32_t end_bit, int32_t order,
                        void* temp, size_t temp_bytes, void* stream, eg_column_view* out_key, eg_column_view* out_val)
{
    (void)temp;
    (void)temp_bytes;
    return eg::guard([&] {
        if (!out_key || !out_val) eg::raise(EG_ERR_INVALID, "null output");
        eg::sort_impl(key, &val, false, begin_bit, end_bit, order == EG_DESCENDING, (hipStream_t)stream, out_key, out_val);
    });
}

int eg_sort_column(eg_column_view col, int32_t begin_bit, int32_t end_bit, int32_t order, void* temp,
                   size_t temp_bytes, void* stream, eg_column_view* out)
{
    (void)temp;
    (void)temp_bytes;
    return eg::guard([&] {
        if (!out) eg::raise(EG_ERR_INVALID, "null output");
        eg::sort_impl(col, nullptr, false, begin_bit, end_bit, order == EG_DESCENDING, (hipStream_t)stream, out, nullptr);
    });
}

int eg_argsort(eg_column_view col, int32_t begin_bit, int32_t end_bit, int32_t order, void* temp, size_t temp_bytes,
               void* stream, eg_column_view* out_sorted, eg_column_view* out_idx)
{
    (void)temp;
    (void)temp_bytes;
    return eg::guard([&] {
        if (!out_sorted || !out_idx) eg::raise(EG_ERR_INVALID, "null output");
        eg::sort_impl(col, nullptr, true, begin_bit, end_bit, order == EG_DESCENDING, (hipStream_t)stream, out_sorted,
                      out_idx);
    });
}

int eg_partition_column_pair(eg_column_view key, eg_column_view val, int32_t begin_bit, int32_t end_bit, void* temp,
                             size_t temp_bytes, void* stream, eg_column_view* out_key, eg_column_view* out_val)
{
    return eg_sort_column_pair(key, val, begin_bit, end_bit, EG_ASCENDING, temp, temp_bytes, stream, out_key, out_val);
}

int eg_partition_column(eg_column_view col, int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes,
                        void* stream, eg_column_view* out)
{
    return eg_sort_column(col, begin_bit, end_bit, EG_ASCENDING, temp, temp_bytes, stream, out);
}

int eg_arg_partition(eg_column_view col, int32_t begin_bit, int32_t end_bit, void* temp, size_t temp_bytes,
                     void* stream, eg_column_view* out_sorted, eg_column_view* out_idx)
{
    return eg_argsort(col, begin_bit, end_bit, EG_ASCENDING, temp, temp_bytes, stream, out_sorted, out_idx);
}

int eg_compute_partition_offsets(eg_column_view partitioned_key, int32_t begin_bit, int32_t end_bit, void* stream,
                                 eg_column_view* out_offsets)
{
    return eg::guard([&] {
        if (!out_offsets) eg::raise(EG_ERR_INVALID, "null output");
        const int32_t t = partitioned_key.type;
        if (!eg::is_int(t)) eg::raise(EG_ERR_UNSUPPORTED, "Cannot extract radix from non-integral type");
        const int kb = eg::type_size(t);
        const int nb = end_bit - begin_bit;
        if (begin_bit < 0 || nb <= 0 || nb > 24 || end_bit > 8 * kb) eg::raise(EG_ERR_INVALID, "bad bit range");
        hipStream_t s = (hipStream_t)stream;
        const int64_t n = partitioned_key.size;
        const int64_t P = (int64_t)1 << nb;
        const void* keys = eg::col_ptr(partitioned_key, kb);
        eg::DevBuf counts((size_t)P * 8, s);
        eg::fill_zero(counts.get(), (size_t)P * 8, s);
        if (n > 0) {
            const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)eg::device_info().num_cus * 8));
            hipLaunchKernelGGL(eg::radix_count_kernel, dim3((unsigned)grid), dim3(256), 0, s, keys, kb, n, begin_bit,
                               (uint64_t)(P - 1), counts.get<unsigned long long>());
            EG_LAUNCH_CHECK();
        }
        eg::DevBuf offs((size_t)(P + 1) * 8, s);
        hipLaunchKernelGGL(eg::inclusive_offsets_kernel, dim3(1), dim3(1024), 0, s,
                           (const unsigned long long*)counts.get<unsigned long long>(), P, n, offs.get<int64_t>());
        EG_LAUNCH_CHECK();
        EG_HIP(hipStreamSynchronize(s));
        eg::set_out(out_offsets, offs.release(), P + 1, EG_INT64);
    });
}

}  // extern "C"

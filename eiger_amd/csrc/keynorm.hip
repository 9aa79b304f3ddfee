// keynorm.hip -- single key columns of the types the strategies do not take
// directly (1- and 2-byte integers, BOOL8, FLOAT32, FLOAT64) are mapped to a
// 4- or 8-byte integer key column that keeps equality and order, and the
// group keys are mapped back:
//   INT8, INT16          -> INT32  (sign extension)
//   UINT8, UINT16, BOOL8 -> UINT32 (zero extension)
//   FLOAT32 -> UINT32, FLOAT64 -> UINT64: the order-preserving bits (sign
//   bit set for non-negatives, every bit flipped for negatives) of the value
//   with -0.0 taken as +0.0 and every NaN as the positive quiet NaN.
// So two keys fall in one group exactly when the reference's key equality
// (operator==, memory/row_utils.cuh:16-21) holds, except that all NaNs form
// one group (== makes every NaN row a group of its own), and the mapped
// order is the numeric order (NaN last), which the sort strategy's ascending
// output keeps.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eg_device.hpp"
#include "eg_internal.hpp"

namespace eg {
namespace {

__device__ __forceinline__ uint64_t map_key(uint64_t raw, int32_t type)
{
    switch (type) {
    case EG_INT8: return (uint64_t)(uint32_t)(int32_t)(int8_t)(uint8_t)raw;
    case EG_INT16: return (uint64_t)(uint32_t)(int32_t)(int16_t)(uint16_t)raw;
    case EG_FLOAT32: {
        uint32_t b = (uint32_t)raw;
        if ((b & 0x7fffffffu) == 0u) b = 0u;                          // -0.0 -> +0.0
        else if ((b & 0x7fffffffu) > 0x7f800000u) b = 0x7fc00000u;    // NaN -> the positive quiet NaN
        return (b & 0x80000000u) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
    }
    case EG_FLOAT64: {
        uint64_t b = raw;
        if ((b & 0x7fffffffffffffffull) == 0ull) b = 0ull;
        else if ((b & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) b = 0x7ff8000000000000ull;
        return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
    }
    default: return raw;  // UINT8, UINT16, BOOL8: zero extension (load_raw's)
    }
}

__device__ __forceinline__ uint64_t unmap_key(uint64_t m, int32_t type)
{
    switch (type) {
    case EG_FLOAT32: {
        const uint32_t o = (uint32_t)m;
        return (o & 0x80000000u) ? (uint64_t)(o & 0x7fffffffu) : (uint64_t)(~o);
    }
    case EG_FLOAT64: return (m & 0x8000000000000000ull) ? (m & 0x7fffffffffffffffull) : ~m;
    default: return m;  // integers: the low bytes (store_bytes truncates)
    }
}

__global__ __launch_bounds__(256) void map_keys_kernel(const void* in, int32_t type, int in_bytes, int64_t n,
                                                       void* out, int out_bytes)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dev::store_bytes(out, out_bytes, i, map_key(dev::load_raw(in, in_bytes, i), type));
}

__global__ __launch_bounds__(256) void unmap_keys_kernel(const void* in, int in_bytes, int64_t n, int32_t type,
                                                         void* out, int out_bytes)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dev::store_bytes(out, out_bytes, i, unmap_key(dev::load_raw(in, in_bytes, i), type));
}

int64_t grid_for(int64_t n)
{
    return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)device_info().num_cus * 16));
}

}  // namespace

bool key_needs_mapping(int32_t type)
{
    switch (type) {
    case EG_INT8: case EG_INT16: case EG_UINT8: case EG_UINT16: case EG_BOOL8: case EG_FLOAT32: case EG_FLOAT64:
        return true;
    default: return false;
    }
}

int32_t mapped_key_type(int32_t type)
{
    switch (type) {
    case EG_INT8: case EG_INT16: return EG_INT32;
    case EG_UINT8: case EG_UINT16: case EG_BOOL8: case EG_FLOAT32: return EG_UINT32;
    case EG_FLOAT64: return EG_UINT64;
    default: return type;
    }
}

DevBuf map_keys(const void* keys, int32_t type, int64_t n, hipStream_t s)
{
    const int32_t mt = mapped_key_type(type);
    DevBuf out((size_t)std::max<int64_t>(n, 1) * type_size(mt), s);
    if (n > 0) {
        hipLaunchKernelGGL(map_keys_kernel, dim3((unsigned)grid_for(n)), dim3(256), 0, s, keys, type, type_size(type),
                           n, out.get(), type_size(mt));
        EG_LAUNCH_CHECK();
    }
    return out;
}

DevBuf unmap_keys(const void* mapped, int64_t n, int32_t type, hipStream_t s)
{
    DevBuf out((size_t)std::max<int64_t>(n, 1) * type_size(type), s);
    if (n > 0) {
        hipLaunchKernelGGL(unmap_keys_kernel, dim3((unsigned)grid_for(n)), dim3(256), 0, s, mapped,
                           type_size(mapped_key_type(type)), n, type, out.get(), type_size(type));
        EG_LAUNCH_CHECK();
    }
    return out;
}

}  // namespace eg

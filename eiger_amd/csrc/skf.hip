// skf.hip -- group keys of several columns fused into one 64-bit key (the
// reference's SKF path: compress::multikey_compress,
// compress/multikey_compress.hpp:11-43, used by groupby.cu:396-465).
//
// Each integer key column c is mapped to its order-preserving unsigned form
// t_c (sign bit flipped for signed types), its range [min_c, max_c] is read
// with one reduction, and it takes bits_c = bit width of (max_c - min_c) bits
// of the fused key, the first column in the most significant bits:
//     fused = sum_c (t_c - min_c) << shift_c.
// Equal fused keys <=> equal key tuples, and the fused order is the
// lexicographic order of the tuples, so the single-key strategies (including
// the key-ordered sort output) apply unchanged.  Output keys are decoded from
// the fused key (unfuse).  Key sets needing more than 64 bits are not fused
// (the reference then falls back to its multi-column hash table; this
// library reports them as unsupported).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "eg_device.hpp"
#include "eg_internal.hpp"
#include "eg_radix.hpp"
#include "eg_skf.hpp"

namespace eg {
namespace {

struct FuseArgs {
    int32_t ncols;
    const void* data[kMaxFusedCols];
    void* out[kMaxFusedCols];
    int32_t bytes[kMaxFusedCols];
    int32_t type[kMaxFusedCols];
    int32_t shift[kMaxFusedCols];
    uint64_t min[kMaxFusedCols];
    uint64_t mask[kMaxFusedCols];
};

__global__ __launch_bounds__(256) void fuse_keys_kernel(FuseArgs a, int64_t n, uint64_t* __restrict__ fused)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t f = 0;
        for (int c = 0; c < a.ncols; ++c) {
            const uint64_t t = twiddle_bits(dev::load_raw(a.data[c], a.bytes[c], i), a.type[c], a.bytes[c]);
            f |= (t - a.min[c]) << a.shift[c];
        }
        fused[i] = f;
    }
}

__global__ __launch_bounds__(256) void unfuse_keys_kernel(FuseArgs a, int64_t n, const uint64_t* __restrict__ fused)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t f = fused[i];
        for (int c = 0; c < a.ncols; ++c) {
            const uint64_t t = ((f >> a.shift[c]) & a.mask[c]) + a.min[c];
            // the twiddle of an integer type is its own inverse
            dev::store_bytes(a.out[c], a.bytes[c], i, twiddle_bits(t, a.type[c], a.bytes[c]));
        }
    }
}

int bit_width(uint64_t v)
{
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

int64_t grid_of(int64_t n)
{
    return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)device_info().num_cus * 8));
}

}  // namespace

bool fuse_keys(const ViewIn* keys, int64_t ncols, hipStream_t s, FusedKeys& f)
{
    if (ncols < 1 || ncols > kMaxFusedCols) raise(EG_ERR_UNSUPPORTED, "at most 8 key columns can be fused");
    const int64_t n = keys[0].size;
    f.ncols = (int)ncols;
    f.rows = n;
    for (int64_t c = 0; c < ncols; ++c) {  // all checks before any device work
        if (keys[c].size != n) raise(EG_ERR_INVALID, "key columns have different row counts");
        if (!is_int(keys[c].type) && keys[c].type != EG_BOOL8)
            raise(EG_ERR_UNSUPPORTED, "multi-column keys must be integer columns");
    }
    int total = 0;
    for (int64_t c = 0; c < ncols; ++c) {
        const ViewIn& k = keys[c];
        f.type[c] = k.type;
        f.bytes[c] = type_size(k.type);
        uint64_t mn = 0, mx = 0;
        if (n > 0) key_range(k.data, k.type, f.bytes[c], n, s, &mn, &mx);
        f.min[c] = mn;
        f.bits[c] = bit_width(mx - mn);
        total += f.bits[c];
    }
    f.total_bits = total;
    if (total > 64) return false;
    int shift = total;
    for (int64_t c = 0; c < ncols; ++c) {
        shift -= f.bits[c];
        f.shift[c] = f.bits[c] > 0 ? shift : 0;  // first column in the most significant bits; constant columns take none
    }
    f.fused = DevBuf((size_t)std::max<int64_t>(n, 1) * 8, s);
    if (n > 0) {
        FuseArgs a{};
        a.ncols = (int32_t)ncols;
        for (int64_t c = 0; c < ncols; ++c) {
            a.data[c] = keys[c].data;
            a.bytes[c] = f.bytes[c];
            a.type[c] = f.type[c];
            a.shift[c] = f.shift[c];
            a.min[c] = f.min[c];
        }
        hipLaunchKernelGGL(fuse_keys_kernel, dim3((unsigned)grid_of(n)), dim3(256), 0, s, a, n, f.fused.get<uint64_t>());
        EG_LAUNCH_CHECK();
    }
    return true;
}

std::vector<DevBuf> unfuse_keys(const FusedKeys& f, const void* fused, int64_t groups, hipStream_t s)
{
    std::vector<DevBuf> cols;
    FuseArgs a{};
    a.ncols = f.ncols;
    for (int c = 0; c < f.ncols; ++c) {
        cols.emplace_back((size_t)std::max<int64_t>(groups, 1) * f.bytes[c], s);
        a.out[c] = cols.back().get();
        a.bytes[c] = f.bytes[c];
        a.type[c] = f.type[c];
        a.shift[c] = f.shift[c];
        a.min[c] = f.min[c];
        a.mask[c] = f.bits[c] >= 64 ? ~0ull : ((1ull << f.bits[c]) - 1ull);
    }
    if (groups > 0) {
        hipLaunchKernelGGL(unfuse_keys_kernel, dim3((unsigned)grid_of(groups)), dim3(256), 0, s, a, groups,
                           (const uint64_t*)fused);
        EG_LAUNCH_CHECK();
    }
    return cols;
}

}  // namespace eg

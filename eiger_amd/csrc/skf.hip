// skf.hip -- group keys of several columns fused into one 64-bit key (the
// reference's SKF path: compress::multikey_compress,
// compress/multikey_compress.hpp:11-43, used by groupby.cu:396-465).
//
// Each integer key column c is mapped to its order-preserving unsigned form
// t_c (sign bit flipped for signed types; float columns first take their
// mapped integer form, keynorm.hip), its range [min_c, max_c] is read
// with one reduction, and it takes bits_c = bit width of (max_c - min_c) bits
// of the fused key, the first column in the most significant bits:
//     fused = sum_c (t_c - min_c) << shift_c.
// Equal fused keys <=> equal key tuples, and the fused order is the
// lexicographic order of the tuples, so the single-key strategies (including
// the key-ordered sort output) apply unchanged.  Output keys are decoded from
// the fused key (unfuse).
//
// Key sets needing more than 64 bits or more than nine fields (the
// reference's multi-column hash table, groupby_hash.cu:212-270) are folded:
// when the next column does not fit beside the columns fused so far (bits,
// or the nine fields of one pass), their fused value is replaced by its
// dense rank (the number of distinct smaller values: a radix sort with row
// ids, head flags and a scan), which needs at most 32 bits below 2^32 rows;
// a column that still does not fit is replaced by its own dense rank.  Ranks
// preserve order, so the fused order stays lexicographic, and each rank's
// table of distinct values (ascending) decodes it.
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
    int32_t nitems;
    const void* data[kMaxFuseItems];       // fuse: item sources
    void* out[kMaxFuseItems];              // unfuse: item destinations
    const uint64_t* table[kMaxFuseItems];  // unfuse: rank -> value, nullptr for plain fields
    int32_t bytes[kMaxFuseItems];
    int32_t type[kMaxFuseItems];
    int32_t shift[kMaxFuseItems];
    uint64_t min[kMaxFuseItems];
    uint64_t mask[kMaxFuseItems];
};

__global__ __launch_bounds__(256) void fuse_keys_kernel(FuseArgs a, int64_t n, uint64_t* __restrict__ fused)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t f = 0;
        for (int c = 0; c < a.nitems; ++c) {
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
        for (int c = 0; c < a.nitems; ++c) {
            uint64_t x = (f >> a.shift[c]) & a.mask[c];
            if (a.table[c]) x = a.table[c][x];
            // the twiddle of an integer type is its own inverse
            dev::store_bytes(a.out[c], a.bytes[c], i, twiddle_bits(x + a.min[c], a.type[c], a.bytes[c]));
        }
    }
}

// Dense ranks of sorted values, in tiles of kRankTile rows (kRankItems
// consecutive rows a thread).
constexpr int kRankB = 256;
constexpr int kRankItems = 16;
constexpr int kRankTile = kRankB * kRankItems;

__device__ __forceinline__ uint32_t tile_heads(const uint64_t* v, int64_t n, int64_t r0)
{
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < kRankItems; ++i) {
        const int64_t r = r0 + i;
        if (r < n && (r == 0 || v[r] != v[r - 1])) ++c;
    }
    return c;
}

// heads[tile] = rows of the tile whose value differs from the row before
__global__ __launch_bounds__(kRankB) void rank_heads_kernel(const uint64_t* __restrict__ v, int64_t n,
                                                            uint32_t* __restrict__ heads)
{
    __shared__ uint32_t wsum[kRankB / 64];
    uint32_t c = tile_heads(v, n, (int64_t)blockIdx.x * kRankTile + (int64_t)threadIdx.x * kRankItems);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kRankB / 64; ++w) t += wsum[w];
        heads[blockIdx.x] = t;
    }
}

// ranks[perm[r]] = distinct values before sorted row r; table[rank] = value.
// head_off[tile] = heads in the tiles before.
__global__ __launch_bounds__(kRankB) void rank_write_kernel(const uint64_t* __restrict__ v,
                                                            const int64_t* __restrict__ perm, int64_t n,
                                                            const uint32_t* __restrict__ head_off,
                                                            uint32_t* __restrict__ ranks, uint64_t* __restrict__ table)
{
    __shared__ uint32_t wsum[kRankB / 64];
    const int64_t r0 = (int64_t)blockIdx.x * kRankTile + (int64_t)threadIdx.x * kRankItems;
    const uint32_t mine = tile_heads(v, n, r0);
    uint32_t run = dev::block_excl_scan<kRankB>(mine, wsum) + head_off[blockIdx.x];
    uint64_t prev = r0 > 0 && r0 < n ? v[r0 - 1] : 0ull;
    for (int i = 0; i < kRankItems; ++i) {
        const int64_t r = r0 + i;
        if (r >= n) break;
        const uint64_t x = v[r];
        if (r == 0 || x != prev) table[run++] = x;
        prev = x;
        ranks[perm[r]] = run - 1u;
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

// Dense order-preserving ranks of n (< 2^32) values below 2^width: ranks
// (rows x uint32) and the distinct values in ascending order (table).
// Returns the distinct count.  Synchronises s.
int64_t dense_ranks(const uint64_t* v, int64_t n, int width, hipStream_t s, DevBuf& ranks, DevBuf& table)
{
    if (width == 0) {  // every value is 0: one rank
        ranks = DevBuf((size_t)n * 4, s);
        table = DevBuf(8, s);
        fill_zero(ranks.get(), (size_t)n * 4, s);
        fill_zero(table.get(), 8, s);
        return 1;
    }
    RadixKeys rk{};
    rk.data = v;
    rk.bytes = 8;
    rk.type = EG_UINT64;
    rk.mode = DIG_KEY;
    rk.kmin = 0;
    DevBuf ka((size_t)n * 8, s), kb((size_t)n * 8, s), pa((size_t)n * 8, s), pb((size_t)n * 8, s);
    SortBuffers b;
    b.keys_in = v;
    b.keys_a = ka.get();
    b.keys_b = kb.get();
    b.pay_in.push_back(Payload{nullptr, nullptr, 8, 1});  // row ids
    b.pay_a.push_back(pa.get());
    b.pay_b.push_back(pb.get());
    const void* sorted = nullptr;
    std::vector<const void*> pay;
    radix_sort(rk, b, n, 0, width, s, &sorted, pay);

    const int64_t ntiles = (n + kRankTile - 1) / kRankTile;
    DevBuf heads((size_t)(ntiles + 1) * 4, s), off((size_t)(ntiles + 1) * 4, s);
    fill_zero(heads.get(), (size_t)(ntiles + 1) * 4, s);
    hipLaunchKernelGGL(rank_heads_kernel, dim3((unsigned)ntiles), dim3(kRankB), 0, s, (const uint64_t*)sorted, n,
                       heads.get<uint32_t>());
    EG_LAUNCH_CHECK();
    const int64_t nd =
        (int64_t)scan_exclusive_u32(heads.get<uint32_t>(), off.get<uint32_t>(), ntiles + 1, s, true);
    ranks = DevBuf((size_t)n * 4, s);
    table = DevBuf((size_t)std::max<int64_t>(nd, 1) * 8, s);
    hipLaunchKernelGGL(rank_write_kernel, dim3((unsigned)ntiles), dim3(kRankB), 0, s, (const uint64_t*)sorted,
                       (const int64_t*)pay[0], n, (const uint32_t*)off.get<uint32_t>(), ranks.get<uint32_t>(),
                       table.get<uint64_t>());
    EG_LAUNCH_CHECK();
    return nd;
}

uint64_t field_mask(int bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull); }

}  // namespace

void fuse_keys(const ViewIn* keys, int64_t ncols, hipStream_t s, FusedKeys& f)
{
    if (ncols < 1 || ncols > (int64_t)1 << 20) raise(EG_ERR_INVALID, "bad key column count");
    const int64_t n = keys[0].size;
    f.ncols = (int)ncols;
    f.rows = n;
    f.type.assign((size_t)ncols, 0);
    f.bytes.assign((size_t)ncols, 0);
    f.bits.assign((size_t)ncols, 0);
    f.min.assign((size_t)ncols, 0);
    f.col_table.clear();
    f.col_table.resize((size_t)ncols);
    f.stages.clear();
    f.orig_type.assign((size_t)ncols, 0);
    for (int64_t c = 0; c < ncols; ++c) {  // all checks before any device work
        if (keys[c].size != n) raise(EG_ERR_INVALID, "key columns have different row counts");
        if (!is_int(keys[c].type) && keys[c].type != EG_BOOL8 && !is_float(keys[c].type))
            raise(EG_ERR_UNSUPPORTED, "multi-column keys must be integer, BOOL8 or float columns");
    }
    // float columns are fused through their mapped integer form (-0.0 as
    // +0.0, every NaN as one value, numeric order; keynorm.hip)
    std::vector<DevBuf> mapped((size_t)ncols);
    std::vector<const void*> src((size_t)ncols);
    int total = 0;
    for (int64_t c = 0; c < ncols; ++c) {
        const ViewIn& k = keys[c];
        f.orig_type[c] = k.type;
        src[c] = k.data;
        int32_t t = k.type;
        if (is_float(t)) {
            if (n > 0) {
                mapped[c] = map_keys(k.data, t, n, s);
                src[c] = mapped[c].get();
            }
            t = mapped_key_type(t);
        }
        f.type[c] = t;
        f.bytes[c] = type_size(t);
        uint64_t mn = 0, mx = 0;
        if (n > 0) key_range(src[c], t, f.bytes[c], n, s, &mn, &mx);
        f.min[c] = mn;
        f.bits[c] = bit_width(mx - mn);
        total += f.bits[c];
    }
    if ((total > 64 || ncols > kMaxFuseItems) && n >= ((int64_t)1 << 32))
        raise(EG_ERR_UNSUPPORTED, "multi-column keys wider than 64 bits take fewer than 2^32 rows");

    DevBuf prev_ranks;                 // dense ranks of the previous stage's fused values
    std::vector<DevBuf> col_ranks((size_t)ncols);  // dense ranks of ranked columns
    // Assigns the stage's shifts (first item in the most significant bits,
    // zero-width items none) and fuses its items for every row.
    auto run_stage = [&](FuseStage& st, int width) {
        int shift = width;
        for (int j = 0; j < st.nitems; ++j) {
            shift -= st.items[j].bits;
            st.items[j].shift = st.items[j].bits > 0 ? shift : 0;
        }
        DevBuf out((size_t)std::max<int64_t>(n, 1) * 8, s);
        if (n > 0) {
            FuseArgs a{};
            a.nitems = st.nitems;
            for (int j = 0; j < st.nitems; ++j) {
                const FuseItem& it = st.items[j];
                a.shift[j] = it.shift;
                if (it.src < 0 || it.ranked) {
                    a.data[j] = it.src < 0 ? prev_ranks.get() : col_ranks[(size_t)it.src].get();
                    a.bytes[j] = 4;
                    a.type[j] = EG_UINT32;
                    a.min[j] = 0;
                } else {
                    a.data[j] = src[(size_t)it.src];
                    a.bytes[j] = f.bytes[it.src];
                    a.type[j] = f.type[it.src];
                    a.min[j] = f.min[it.src];
                }
            }
            hipLaunchKernelGGL(fuse_keys_kernel, dim3((unsigned)grid_of(n)), dim3(256), 0, s, a, n,
                               out.get<uint64_t>());
            EG_LAUNCH_CHECK();
        }
        return out;
    };

    FuseStage cur;
    int width = 0;
    for (int c = 0; c < (int)ncols; ++c) {
        FuseItem it;
        it.src = c;
        it.bits = f.bits[c];
        if (width + it.bits > 64 || cur.nitems == kMaxFuseItems) {
            // the columns so far become the dense rank of their fused value
            DevBuf v = run_stage(cur, width);
            const int64_t nd = dense_ranks(v.get<uint64_t>(), n, width, s, prev_ranks, cur.table);
            f.stages.push_back(std::move(cur));
            cur = FuseStage{};
            width = bit_width((uint64_t)(nd - 1));
            FuseItem p;
            p.src = -1;
            p.bits = width;
            cur.items[cur.nitems++] = p;
            if (width + it.bits > 64) {  // and this column its own dense rank
                FuseStage one;
                one.items[one.nitems++] = it;
                DevBuf w = run_stage(one, it.bits);
                const int64_t ndc =
                    dense_ranks(w.get<uint64_t>(), n, it.bits, s, col_ranks[(size_t)c], f.col_table[(size_t)c]);
                it.ranked = 1;
                it.bits = bit_width((uint64_t)(ndc - 1));
            }
        }
        cur.items[cur.nitems++] = it;
        width += it.bits;
    }
    f.total_bits = width;
    f.fused = run_stage(cur, width);
    f.stages.push_back(std::move(cur));
}

std::vector<DevBuf> unfuse_keys(const FusedKeys& f, const void* fused, int64_t groups, hipStream_t s)
{
    std::vector<DevBuf> cols;
    for (int c = 0; c < f.ncols; ++c) cols.emplace_back((size_t)std::max<int64_t>(groups, 1) * f.bytes[c], s);
    if (groups <= 0) return cols;
    // last stage first: each earlier stage's fused values come out of the
    // rank table of the stage before
    DevBuf cur_buf;
    const void* cur = fused;
    for (int st = (int)f.stages.size() - 1; st >= 0; --st) {
        const FuseStage& S = f.stages[(size_t)st];
        DevBuf prev;
        FuseArgs a{};
        a.nitems = S.nitems;
        for (int j = 0; j < S.nitems; ++j) {
            const FuseItem& it = S.items[j];
            a.shift[j] = it.shift;
            a.mask[j] = field_mask(it.bits);
            if (it.src < 0) {
                prev = DevBuf((size_t)groups * 8, s);
                a.out[j] = prev.get();
                a.bytes[j] = 8;
                a.type[j] = EG_UINT64;
                a.min[j] = 0;
                a.table[j] = f.stages[(size_t)st - 1].table.get<uint64_t>();
            } else {
                a.out[j] = cols[(size_t)it.src].get();
                a.bytes[j] = f.bytes[it.src];
                a.type[j] = f.type[it.src];
                a.min[j] = f.min[it.src];
                a.table[j] = it.ranked ? f.col_table[(size_t)it.src].get<uint64_t>() : nullptr;
            }
        }
        hipLaunchKernelGGL(unfuse_keys_kernel, dim3((unsigned)grid_of(groups)), dim3(256), 0, s, a, groups,
                           (const uint64_t*)cur);
        EG_LAUNCH_CHECK();
        cur_buf = std::move(prev);
        cur = cur_buf.get();
    }
    for (int c = 0; c < f.ncols; ++c)
        if (is_float(f.orig_type[(size_t)c])) cols[(size_t)c] = unmap_keys(cols[(size_t)c].get(), groups, f.orig_type[(size_t)c], s);
    return cols;
}

}  // namespace eg

// dict.hip -- dictionary encoding of an integer column
// (gamax::compress::dictionary_encode, compress/dictionary_encode.hpp:12-37):
//   * with a given dictionary (dictionary_encode_dict_ready.cu:15-80): each
//     row gets the position of its key in the dictionary column, INT32, or
//     INT32_MIN (the reference's empty-value sentinel, types/sentinel.hpp:10-16)
//     when the key is not in it;
//   * self-built (dictionary_encode.cu): each row gets a dense id in
//     [0, number of distinct keys), any integral output type; with_order
//     numbers the distinct keys in ascending key order.  SORT_OPT's
//     re-encoding of a key column against its sorted unique keys
//     (groupby_sort.cu:394-413) is this operation.
//
// The dictionary goes into an open-addressing table of 2^c >= 2 m slots
// (m dictionary entries): slot value = dictionary position + 1 (0 = empty,
// so a zero-fill clears it), claimed by compare-and-swap; a duplicate key
// keeps its smallest position (atomicMin), so results are deterministic.  A
// second pass stores each slot's key beside it.  The lookup streams the column
// once (read key, write id): HBM-bound, 4 + 4 bytes per int32 row.  Tables of
// up to 8192 slots are copied into each workgroup's LDS; larger ones are probed
// in HBM as packed 16-byte {key, position} slots, one 64-byte line per bucket.  Unlike the reference's cuco
// map (INT_MIN reserved as the empty key), every key value is valid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "eg_device.hpp"
#include "eg_internal.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

constexpr int kDictBlock = 1024;
constexpr int kLdsSlotsMax = 8192;  // 32 KB of positions + up to 64 KB of keys
constexpr int64_t kMaxDict = (int64_t)1 << 30;

// First slot of the key's bucket: buckets are aligned groups of 4 slots, and
// probing runs linearly from there, so one 16-byte read covers a bucket.
__device__ __forceinline__ uint32_t dict_home(uint64_t raw, int cap_log2)
{
    return (uint32_t)(dev::fmix64(raw) >> (64 - (cap_log2 - 2))) << 2;
}

__global__ __launch_bounds__(256) void dict_build_kernel(const void* dict, int kb, int64_t m, uint32_t* slots,
                                                         int cap_log2)
{
    const uint32_t mask = (1u << cap_log2) - 1u;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint64_t k = dev::load_raw(dict, kb, i);
        const uint32_t v = (uint32_t)i + 1u;
        uint32_t h = dict_home(k, cap_log2);
        // at most m of the 2^cap_log2 >= 2m slots are ever taken, so an empty
        // slot is met within the bound.  Slots are taken in probe order and
        // never freed, so after the build the occupied slots of every probe
        // sequence form a prefix: a lookup stops at the first empty slot.
        for (uint32_t p = 0; p <= mask; ++p) {
            const uint32_t old = atomicCAS(&slots[h], 0u, v);
            if (old == 0u) break;
            if (dev::load_raw(dict, kb, (int64_t)old - 1) == k) {
                atomicMin(&slots[h], v);
                break;
            }
            h = (h + 1u) & mask;
        }
    }
}

__global__ __launch_bounds__(256) void dict_slot_keys_kernel(const void* dict, int kb, const uint32_t* slots,
                                                             int64_t cap, uint64_t* skeys)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += stride) {
        const uint32_t v = slots[s];
        skeys[s] = v ? dev::load_raw(dict, kb, (int64_t)v - 1) : 0ull;
    }
}

__global__ __launch_bounds__(256) void dict_pack_kernel(const void* dict, int kb, const uint32_t* slots, int64_t cap,
                                                        uint4* packed)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += stride) {
        const uint32_t v = slots[s];
        const uint64_t k = v ? dev::load_raw(dict, kb, (int64_t)v - 1) : 0ull;
        packed[s] = make_uint4((uint32_t)k, (uint32_t)(k >> 32), v, 0u);
    }
}

// Keys of the LDS copy: 4 bytes for keys of up to 4 bytes, else 8.  The HBM
// table keeps 8-byte keys.
template <int KB, bool LDS>
using SlotKey = typename std::conditional<(KB > 4 || !LDS), uint64_t, uint32_t>::type;

template <int KB>
__device__ __forceinline__ uint64_t load_key(const void* p, int64_t i)
{
    if constexpr (KB == 1) return ((const uint8_t*)p)[i];
    else if constexpr (KB == 2) return ((const uint16_t*)p)[i];
    else if constexpr (KB == 4) return ((const uint32_t*)p)[i];
    else return ((const uint64_t*)p)[i];
}

template <int OB>
__device__ __forceinline__ void store_id(void* p, int64_t i, uint64_t bits)
{
    if constexpr (OB == 1) ((uint8_t*)p)[i] = (uint8_t)bits;
    else if constexpr (OB == 2) ((uint16_t*)p)[i] = (uint16_t)bits;
    else if constexpr (OB == 4) ((uint32_t*)p)[i] = (uint32_t)bits;
    else ((uint64_t*)p)[i] = bits;
}

// One bucket (4 slots from h, h a multiple of 4): 16-byte reads.
template <typename KT>
struct DictBucket {
    uint32_t v[4];
    KT k[4];
};

template <typename KT>
__device__ __forceinline__ DictBucket<KT> read_bucket(const uint32_t* sidx, const KT* skey, uint32_t h)
{
    DictBucket<KT> b;
    const uint4 v = *(const uint4*)(sidx + h);
    b.v[0] = v.x, b.v[1] = v.y, b.v[2] = v.z, b.v[3] = v.w;
    if constexpr (sizeof(KT) == 4) {
        const uint4 k = *(const uint4*)(skey + h);
        b.k[0] = k.x, b.k[1] = k.y, b.k[2] = k.z, b.k[3] = k.w;
    } else {
        const ulonglong2 k0 = *(const ulonglong2*)(skey + h), k1 = *(const ulonglong2*)(skey + h + 2);
        b.k[0] = k0.x, b.k[1] = k0.y, b.k[2] = k1.x, b.k[3] = k1.y;
    }
    return b;
}

// true when the bucket decides the lookup (res = position or miss).
template <typename KT>
__device__ __forceinline__ bool match_bucket(const DictBucket<KT>& b, uint64_t k, uint64_t miss, uint64_t& res)
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (b.v[j] == 0u) {
            res = miss;
            return true;
        }
        if ((uint64_t)b.k[j] == k) {
            res = (uint64_t)(b.v[j] - 1u);
            return true;
        }
    }
    return false;
}

// Where a lookup reads its buckets: the LDS copy (positions and keys in two
// arrays) or the packed HBM table (one 16-byte {key, position} entry per
// slot, so a bucket is one 64-byte line: one random line per lookup instead
// of two).
template <typename KT>
struct SplitTable {
    const uint32_t* sidx;
    const KT* skey;
    __device__ __forceinline__ DictBucket<KT> bucket(uint32_t h) const { return read_bucket(sidx, skey, h); }
};

struct PackedTable {
    const uint4* slot;
    __device__ __forceinline__ DictBucket<uint64_t> bucket(uint32_t h) const
    {
        DictBucket<uint64_t> b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint4 t = slot[h + j];
            b.v[j] = t.z;
            b.k[j] = (uint64_t)t.x | ((uint64_t)t.y << 32);
        }
        return b;
    }
};

// Lookup continued from bucket h (the key's earlier buckets were full).
template <class TAB>
__device__ __noinline__ uint64_t probe_from(const TAB& t, uint64_t k, uint32_t h, int cap_log2, uint64_t miss)
{
    const uint32_t mask = (1u << cap_log2) - 1u;
    for (uint32_t p = 0; p <= mask; p += 4) {
        uint64_t r;
        if (match_bucket(t.bucket(h), k, miss, r)) return r;
        h = (h + 4u) & mask;
    }
    return miss;
}

template <int KB, int OB, class TAB>
__device__ __forceinline__ void find_rows(const TAB& t, const void* __restrict__ keys, int64_t n, int cap_log2,
                                          uint64_t miss, void* __restrict__ out)
{
    const uint32_t mask = (1u << cap_log2) - 1u;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // four rows per step: their first buckets are read together
    for (; i + 3 * stride < n; i += 4 * stride) {
        uint64_t k[4];
        uint32_t h[4];
        decltype(t.bucket(0u)) b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) k[u] = load_key<KB>(keys, i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            h[u] = dict_home(k[u], cap_log2);
            b[u] = t.bucket(h[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            uint64_t r;
            if (!match_bucket(b[u], k[u], miss, r)) r = probe_from(t, k[u], (h[u] + 4u) & mask, cap_log2, miss);
            store_id<OB>(out, i + u * stride, r);
        }
    }
    for (; i < n; i += stride) {
        const uint64_t k = load_key<KB>(keys, i);
        store_id<OB>(out, i, probe_from(t, k, dict_home(k, cap_log2), cap_log2, miss));
    }
}

// LDS: `slots` positions and `aux` 64-bit keys, copied into LDS per
// workgroup; HBM: `aux` is the packed table.
template <int KB, int OB, bool LDS>
__global__ __launch_bounds__(kDictBlock) void dict_find_kernel(const void* __restrict__ keys, int64_t n,
                                                               const uint32_t* __restrict__ slots,
                                                               const void* __restrict__ aux, int cap_log2,
                                                               uint64_t miss, void* __restrict__ out)
{
    if constexpr (LDS) {
        using KT = SlotKey<KB, true>;
        extern __shared__ __align__(16) unsigned char dict_smem[];
        const int cap = 1 << cap_log2;
        uint32_t* li = (uint32_t*)dict_smem;
        KT* lk = (KT*)(dict_smem + (size_t)cap * 4);
        const uint64_t* skeys = (const uint64_t*)aux;
        for (int s = threadIdx.x; s < cap; s += blockDim.x) {
            li[s] = slots[s];
            lk[s] = (KT)skeys[s];
        }
        __syncthreads();
        find_rows<KB, OB>(SplitTable<KT>{li, lk}, keys, n, cap_log2, miss, out);
    } else {
        find_rows<KB, OB>(PackedTable{(const uint4*)aux}, keys, n, cap_log2, miss, out);
    }
}

// Direct-indexed table for dictionaries whose key span (max - min + 1) is
// at most kDirectMax: slot d = key - min holds ~position (0 = absent;
// atomicMax keeps the first position of a repeated key), and a row costs one
// 4-byte LDS read.  key - min is taken modulo 2^(8 KB), so keys outside
// [min, max] land above the span without a signed compare.
constexpr int64_t kDirectMax = 32768;  // 128 KB of LDS: one 1024-thread workgroup per CU

__global__ __launch_bounds__(256) void dict_direct_build_kernel(const void* dict, int kb, int64_t m, uint64_t raw_min,
                                                                uint64_t wmask, uint32_t* table)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
        atomicMax(&table[(dev::load_raw(dict, kb, i) - raw_min) & wmask], ~(uint32_t)i);
}

template <int KB, int OB>
__global__ __launch_bounds__(kDictBlock) void dict_direct_kernel(const void* __restrict__ keys, int64_t n,
                                                                 const uint32_t* __restrict__ table, uint64_t span,
                                                                 uint64_t raw_min, uint64_t miss, void* __restrict__ out)
{
    extern __shared__ __align__(16) unsigned char dict_smem[];
    uint32_t* dt = (uint32_t*)dict_smem;
    for (int64_t s = threadIdx.x; s <= (int64_t)span; s += blockDim.x) dt[s] = table[s];
    __syncthreads();
    constexpr uint64_t wmask = KB == 8 ? ~0ull : ((1ull << (8 * KB)) - 1ull);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto find = [&](uint64_t k) -> uint64_t {
        const uint64_t d = (k - raw_min) & wmask;
        const uint32_t v = d <= span ? dt[d] : 0u;
        return v ? (uint64_t)(uint32_t)~v : miss;
    };
    for (; i + 3 * stride < n; i += 4 * stride) {
        uint64_t k[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) k[u] = load_key<KB>(keys, i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) store_id<OB>(out, i + u * stride, find(k[u]));
    }
    for (; i < n; i += stride) store_id<OB>(out, i, find(load_key<KB>(keys, i)));
}

template <int KB, int OB>
void launch_direct(const void* keys, int64_t n, const uint32_t* table, uint64_t span, uint64_t raw_min, uint64_t miss,
                   void* out, hipStream_t s)
{
    const void* kern = (const void*)dict_direct_kernel<KB, OB>;
    const size_t lds = (size_t)(span + 1) * 4;
    if (lds > enable_dynamic_lds(kern)) raise(EG_ERR_DEVICE, "dictionary table exceeds LDS");
    int per_cu = 0;
    EG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kDictBlock, lds));
    per_cu = std::max(per_cu, 1);
    const int64_t grid =
        std::max<int64_t>(1, std::min<int64_t>((n + kDictBlock - 1) / kDictBlock, (int64_t)device_info().num_cus * per_cu));
    hipLaunchKernelGGL((dict_direct_kernel<KB, OB>), dim3((unsigned)grid), dim3(kDictBlock), lds, s, keys, n, table,
                       span, raw_min, miss, out);
    EG_LAUNCH_CHECK();
}

template <int KB>
void launch_direct_ob(int ob, const void* keys, int64_t n, const uint32_t* table, uint64_t span, uint64_t raw_min,
                      uint64_t miss, void* out, hipStream_t s)
{
    switch (ob) {
    case 1: launch_direct<KB, 1>(keys, n, table, span, raw_min, miss, out, s); break;
    case 2: launch_direct<KB, 2>(keys, n, table, span, raw_min, miss, out, s); break;
    case 4: launch_direct<KB, 4>(keys, n, table, span, raw_min, miss, out, s); break;
    default: launch_direct<KB, 8>(keys, n, table, span, raw_min, miss, out, s); break;
    }
}

template <int KB, int OB, bool LDS>
void launch_find(const void* keys, int64_t n, const uint32_t* slots, const void* skeys, int cap_log2,
                 uint64_t miss, void* out, hipStream_t s)
{
    const void* kern = (const void*)dict_find_kernel<KB, OB, LDS>;
    size_t lds = 0;
    if (LDS) {
        lds = ((size_t)4 + sizeof(SlotKey<KB, true>)) << cap_log2;
        if (lds > enable_dynamic_lds(kern)) raise(EG_ERR_DEVICE, "dictionary table exceeds LDS");
    }
    int per_cu = 0;
    EG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kDictBlock, lds));
    per_cu = std::max(per_cu, 1);
    const int64_t grid =
        std::max<int64_t>(1, std::min<int64_t>((n + kDictBlock - 1) / kDictBlock, (int64_t)device_info().num_cus * per_cu));
    hipLaunchKernelGGL((dict_find_kernel<KB, OB, LDS>), dim3((unsigned)grid), dim3(kDictBlock), lds, s, keys, n,
                       slots, skeys, cap_log2, miss, out);
    EG_LAUNCH_CHECK();
}

template <int KB, bool LDS>
void launch_find_ob(int ob, const void* keys, int64_t n, const uint32_t* slots, const void* skeys,
                    int cap_log2, uint64_t miss, void* out, hipStream_t s)
{
    switch (ob) {
    case 1: launch_find<KB, 1, LDS>(keys, n, slots, skeys, cap_log2, miss, out, s); break;
    case 2: launch_find<KB, 2, LDS>(keys, n, slots, skeys, cap_log2, miss, out, s); break;
    case 4: launch_find<KB, 4, LDS>(keys, n, slots, skeys, cap_log2, miss, out, s); break;
    default: launch_find<KB, 8, LDS>(keys, n, slots, skeys, cap_log2, miss, out, s); break;
    }
}

template <bool LDS>
void launch_find_kb(int kb, int ob, const void* keys, int64_t n, const uint32_t* slots, const void* skeys,
                    int cap_log2, uint64_t miss, void* out, hipStream_t s)
{
    switch (kb) {
    case 1: launch_find_ob<1, LDS>(ob, keys, n, slots, skeys, cap_log2, miss, out, s); break;
    case 2: launch_find_ob<2, LDS>(ob, keys, n, slots, skeys, cap_log2, miss, out, s); break;
    case 4: launch_find_ob<4, LDS>(ob, keys, n, slots, skeys, cap_log2, miss, out, s); break;
    default: launch_find_ob<8, LDS>(ob, keys, n, slots, skeys, cap_log2, miss, out, s); break;
    }
}

int64_t small_grid(int64_t n)
{
    return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)device_info().num_cus * 8));
}

const void* view_ptr(const eg_column_view& c, int bytes)
{
    if (c.offset < 0 || c.size < 0) raise(EG_ERR_INVALID, "bad column view");
    if (c.size > 0 && !c.data) raise(EG_ERR_INVALID, "null column data");
    return c.data ? (const char*)c.data + c.offset * bytes : nullptr;
}

// The dict-ready form takes integral keys (dictionary_encode_dict_ready.cu);
// the self-built forms also take floats (dictionary_encode.cu type-dispatches
// every non-string type).
void check_key_type(int32_t t, bool floats_ok)
{
    if (!is_int(t) && t != EG_BOOL8 && !(floats_ok && is_float(t)))
        raise(EG_ERR_UNSUPPORTED, "Dictionary encoding is not supported for non-integral type.");
}

}  // namespace

// Positions of column's keys in dict (n rows of `out_type`, `miss` where a key
// is absent).  Both columns have the same integer type.
DevBuf dict_encode(const void* keys, int64_t n, const void* dict, int64_t m, int32_t type, int32_t out_type,
                   uint64_t miss, hipStream_t s)
{
    const int kb = type_size(type);
    const int ob = type_size(out_type);
    DevBuf out((size_t)std::max<int64_t>(n, 1) * ob, s);
    if (n == 0) return out;
    if (m > kMaxDict) raise(EG_ERR_UNSUPPORTED, "dictionary larger than 2^30 entries");
    if (m > 0) {
        uint64_t tmin = 0, tmax = 0;
        key_range(dict, type, kb, m, s, &tmin, &tmax);  // order-preserving form
        const uint64_t span = tmax - tmin;
        if (span < (uint64_t)kDirectMax) {
            const uint64_t raw_min = twiddle_bits(tmin, type, kb);  // the twiddle of an integer is its own inverse
            const uint64_t wmask = kb == 8 ? ~0ull : ((1ull << (8 * kb)) - 1ull);
            DevBuf table((size_t)(span + 1) * 4, s);
            fill_zero(table.get(), (size_t)(span + 1) * 4, s);
            hipLaunchKernelGGL(dict_direct_build_kernel, dim3((unsigned)small_grid(m)), dim3(256), 0, s, dict, kb, m,
                               raw_min, wmask, table.get<uint32_t>());
            EG_LAUNCH_CHECK();
            const uint32_t* t = table.get<uint32_t>();
            switch (kb) {
            case 1: launch_direct_ob<1>(ob, keys, n, t, span, raw_min, miss, out.get(), s); break;
            case 2: launch_direct_ob<2>(ob, keys, n, t, span, raw_min, miss, out.get(), s); break;
            case 4: launch_direct_ob<4>(ob, keys, n, t, span, raw_min, miss, out.get(), s); break;
            default: launch_direct_ob<8>(ob, keys, n, t, span, raw_min, miss, out.get(), s); break;
            }
            EG_HIP(hipStreamSynchronize(s));  // the table is freed on return
            return out;
        }
    }
    // hashed table; an empty dictionary gives an empty 64-slot table (every key absent)
    int cap_log2 = 6;
    while (((int64_t)1 << cap_log2) < 2 * m) ++cap_log2;
    const int64_t cap = (int64_t)1 << cap_log2;
    DevBuf slots((size_t)cap * 4, s);
    fill_zero(slots.get(), (size_t)cap * 4, s);
    const bool in_lds = cap <= kLdsSlotsMax;
    DevBuf aux((size_t)cap * (in_lds ? 8 : 16), s);
    if (m > 0) {
        hipLaunchKernelGGL(dict_build_kernel, dim3((unsigned)small_grid(m)), dim3(256), 0, s, dict, kb, m,
                           slots.get<uint32_t>(), cap_log2);
        EG_LAUNCH_CHECK();
    }
    if (in_lds) {
        hipLaunchKernelGGL(dict_slot_keys_kernel, dim3((unsigned)small_grid(cap)), dim3(256), 0, s, dict, kb,
                           (const uint32_t*)slots.get<uint32_t>(), cap, aux.get<uint64_t>());
        EG_LAUNCH_CHECK();
        launch_find_kb<true>(kb, ob, keys, n, slots.get<uint32_t>(), aux.get(), cap_log2, miss, out.get(), s);
    } else {
        hipLaunchKernelGGL(dict_pack_kernel, dim3((unsigned)small_grid(cap)), dim3(256), 0, s, dict, kb,
                           (const uint32_t*)slots.get<uint32_t>(), cap, aux.get<uint4>());
        EG_LAUNCH_CHECK();
        launch_find_kb<false>(kb, ob, keys, n, slots.get<uint32_t>(), aux.get(), cap_log2, miss, out.get(), s);
    }
    EG_HIP(hipStreamSynchronize(s));  // the tables are freed on return
    return out;
}

}  // namespace eg

using namespace eg;

int eg_dictionary_encode_with_dict(eg_column_view column, eg_column_view dict, int32_t encode_type, void* stream,
                                   eg_column_view* out)
{
    return guard([&] {
        if (!out) raise(EG_ERR_INVALID, "null output");
        if (encode_type != EG_INT32) raise(EG_ERR_UNSUPPORTED, "Unsupported encoded type");
        check_key_type(column.type, false);
        if (dict.type != column.type) raise(EG_ERR_INVALID, "dictionary and column types differ");
        const int kb = type_size(column.type);
        const void* keys = view_ptr(column, kb);
        const void* d = view_ptr(dict, kb);
        hipStream_t s = (hipStream_t)stream;
        DevBuf enc = dict_encode(keys, column.size, d, dict.size, column.type, EG_INT32, (uint64_t)(uint32_t)INT32_MIN, s);
        out->data = enc.release();
        out->size = column.size;
        out->type = EG_INT32;
        out->reserved = 0;
        out->offset = 0;
    });
}

int eg_dictionary_encode(eg_column_view column, int32_t encode_type, int64_t dictionary_size_hint, int32_t with_order,
                         void* stream, eg_column_view* out, int64_t* out_dict_size, eg_column_view* out_dict)
{
    eg_table t = nullptr;
    const int rc = guard([&] {
        if (!out || !out_dict_size) raise(EG_ERR_INVALID, "null output");
        if (!is_int(encode_type)) raise(EG_ERR_UNSUPPORTED, "Encoded type must be integral.");
        check_key_type(column.type, true);
        const int kb = type_size(column.type);
        const void* keys = view_ptr(column, kb);
        hipStream_t s = (hipStream_t)stream;
        // 1- and 2-byte integer, BOOL8 and float keys: the distinct-key pass
        // and the lookup run on the mapped 4- or 8-byte integer keys (equal
        // keys <=> equal mapped keys, order kept; floats: -0.0 = +0.0, one
        // NaN); the dictionary is mapped back
        const bool mapped_key = key_needs_mapping(column.type);
        eg_column_view src = column;
        DevBuf mapped;
        if (mapped_key) {
            mapped = map_keys(keys, column.type, column.size, s);
            keys = mapped.get();
            src.data = mapped.get();
            src.offset = 0;
            src.type = mapped_key_type(column.type);
        }
        // the distinct keys: ascending (sort) for with_order, any order (hash) otherwise
        eg_groupby_hint h;
        eg_hint_init(&h);
        if (dictionary_size_hint > 0) h.groups_upper_bound = dictionary_size_hint;
        const int r = with_order ? eg_distinct_sort(&src, 1, &h, stream, &t) : eg_distinct_hash(&src, 1, &h, stream, &t);
        if (r != EG_OK) raise(r, eg_last_error());
        eg_column_view dv{};
        if (eg_table_release_column(t, 0, &dv) != EG_OK) raise(EG_ERR_DEVICE, "dictionary_encode: no key column");
        eg_table_free(t);
        t = nullptr;
        struct FreeOnExit {
            void* p;
            ~FreeOnExit() { if (p) eg_free(p); }
        } dict_owner{const_cast<void*>(dv.data)};
        const int ob = type_size(encode_type);
        if (ob < 8) {
            const int vb = 8 * ob - (is_signed_int(encode_type) ? 1 : 0);
            if (dv.size > ((int64_t)1 << vb)) raise(EG_ERR_INVALID, "dictionary does not fit the encoded type");
        }
        DevBuf enc = dict_encode(keys, column.size, dv.data, dv.size, src.type, encode_type, 0ull, s);
        out->data = enc.release();
        out->size = column.size;
        out->type = encode_type;
        out->reserved = 0;
        out->offset = 0;
        *out_dict_size = dv.size;
        if (out_dict && mapped_key) {
            DevBuf back = unmap_keys(dv.data, dv.size, column.type, s);
            EG_HIP(hipStreamSynchronize(s));
            out_dict->data = back.release();
            out_dict->size = dv.size;
            out_dict->type = column.type;
            out_dict->reserved = 0;
            out_dict->offset = 0;
        } else if (out_dict) {
            out_dict->data = dv.data;
            dict_owner.p = nullptr;
            out_dict->size = dv.size;
            out_dict->type = column.type;
            out_dict->reserved = 0;
            out_dict->offset = 0;
        }
    });
    if (t) eg_table_free(t);
    return rc;
}

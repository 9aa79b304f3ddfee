// batch.hip -- many independent single-key group-bys in one persistent launch.
//
// The reference runs one group-by per call with host synchronisation inside
// (groupby_hash.cu:256-262,398; groupby.cu:390) and has no batched entry point.
// Here a batch of queries is planned once on the host and then executed with
// no host synchronisation and no allocation, so the whole execution can be
// captured in a hipGraph and replayed:
//
//   batch_fill_kernel    zeroes every query's HBM table and the counters
//                        (a fill kernel, not hipMemsetAsync: graph-safe);
//   batch_agg_kernel     a persistent grid; each workgroup pulls work items
//                        (query, row range) from an atomic queue until it is
//                        drained.  A query with a small group bound is
//                        aggregated in the workgroup's LDS table (hot keys
//                        combined by the wave first) and flushed into the
//                        query's HBM table; a large one goes straight to its
//                        HBM table (wave-combined atomics);
//   batch_compact_kernel one workgroup per (query, slot range) writes the
//                        query's groups into the caller's output columns at
//                        wave-claimed positions, counting groups per query.
//
// Queries whose group bound is above the LDS bound are partitioned first
// (tuning.radix_bits > 0), with no host synchronisation in between:
//
//   bpart_hist_kernel    per 2^14-row tile of such a query, an LDS histogram of
//                        the low hash bits of its keys (the partition id);
//   bpart_scan_kernel    one workgroup per query scans its [partition][tile]
//                        counts into each tile's first row per partition;
//   bpart_scatter_kernel moves every row (key and value) to its partition in
//                        a scratch buffer;
//   batch_agg_kernel     then takes one work item per partition, aggregates it
//                        in the workgroup's LDS table (a partition holds about
//                        half a table of groups) and writes its groups straight
//                        to the query's output; keys the table cannot hold
//                        spill to a small HBM table, appended by the
//                        compaction.
//
// Per-query failures cannot be retried asynchronously: a table that fills
// (group_bound below the true group count) or an output that exceeds its
// capacity sets the query's status bit, read back by eg_batch_results.
//
// The tuning (rows per work item, LDS table slots, the group bound up to which
// the LDS stage is used, workgroups per CU) comes from eg_batch_autotune,
// which times candidate tunings on the caller's own batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "eg_atomics.hpp"
#include "eg_device.hpp"
#include "eg_hash_table.hpp"
#include "eg_internal.hpp"
#include "eg_lds_table.hpp"

struct eg_batch_plan_s;

namespace eg {
namespace {

using namespace lds;

constexpr int kBatchBlock = 1024;
constexpr uint32_t kStTableFull = 1u, kStCapacity = 2u;

struct BQuery {
    const void* keys;
    const void* vals;
    unsigned long long* table;  // (mask + 1) * W words, then the side slot (W words)
    uint64_t mask;
    int64_t rows;
    int64_t capacity;
    void* out_keys;
    void* out_aggs[EG_BATCH_MAX_AGGS];
    int32_t use_lds;
    int32_t pbits;        // partition bits (0: not partitioned)
    int64_t ntiles;       // tiles of the partition pass
    int64_t hist_off;     // its [partition][tile] counts / row offsets
    int64_t scratch_off;  // its rows in the partitioned scratch columns
};

// A work item: rows [begin, end) of query q (kind 0), or partition `begin`
// of a partitioned query (kind 1).
struct BItem {
    int32_t q;
    int32_t kind;
    int64_t begin;
    int64_t end;
};

// A tile of the partition pass: rows [t * kPTile, (t + 1) * kPTile) of query q.
struct BPTile {
    int32_t q;
    int32_t t;
};
constexpr int64_t kPTile = (int64_t)1 << 14;
constexpr int kMaxPartBits = 12;
constexpr int kPB = 256;  // threads of the partition-pass kernels

struct BCItem {
    int32_t q;
    int32_t pad;
    uint64_t slot_begin;
    uint64_t slot_end;  // inclusive end is mask + 1 (the side slot) for the last range
};

// Per-batch constant layout of the outputs (identical for every query).
struct BOut {
    int32_t nouts;
    int32_t key_bytes;
    int32_t code[EG_BATCH_MAX_AGGS];
    int32_t bytes[EG_BATCH_MAX_AGGS];
    int32_t op[EG_BATCH_MAX_AGGS];
    int32_t op2[EG_BATCH_MAX_AGGS];
    int32_t is_float[EG_BATCH_MAX_AGGS];
};

__device__ __forceinline__ GTable query_table(const BQuery& q, int W, unsigned long long* meta)
{
    GTable gt;
    gt.base = q.table;
    gt.side = q.table + (q.mask + 1) * (uint64_t)W;
    gt.meta = meta;
    gt.mask = q.mask;
    gt.W = W;
    return gt;
}

// meta words per query: [0] inserts (unused), [1] side slot used, [2] full
constexpr int kMetaWords = 4;

__global__ void batch_fill_kernel(unsigned long long* p, uint64_t words)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 2;
    for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < words; i += stride) {
        if (i + 1 < words) {
            *reinterpret_cast<ulonglong2*>(p + i) = make_ulonglong2(0ull, 0ull);
        } else {
            p[i] = 0ull;
        }
    }
}

template <int KB>
__device__ __forceinline__ uint32_t part_of(uint64_t raw, uint32_t mask)
{
    return (uint32_t)dev::bit_mix<KB>(raw) & mask;
}

template <int KB>
__global__ __launch_bounds__(kPB) void bpart_hist_kernel(const BQuery* __restrict__ qs, const BPTile* __restrict__ tiles,
                                                         uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[1 << kMaxPartBits];
    const BPTile tt = tiles[blockIdx.x];
    const BQuery& q = qs[tt.q];
    const uint32_t P = 1u << q.pbits, mask = P - 1u;
    for (uint32_t i = threadIdx.x; i < P; i += kPB) h[i] = 0u;
    __syncthreads();
    const int64_t b = (int64_t)tt.t * kPTile, e = min(q.rows, b + kPTile);
    for (int64_t r = b + threadIdx.x; r < e; r += kPB) atomicAdd(&h[part_of<KB>(dev::load_raw(q.keys, KB, r), mask)], 1u);
    __syncthreads();
    uint32_t* out = hist + q.hist_off + tt.t;
    for (uint32_t d = threadIdx.x; d < P; d += kPB) out[(int64_t)d * q.ntiles] = h[d];
}

// Exclusive scan of one partitioned query's counts in [partition][tile]
// order (in place): each entry becomes the query-local row where that tile's
// rows of that partition start.  One 1024-thread workgroup per query, 8
// entries per thread per round.
__global__ __launch_bounds__(1024) void bpart_scan_kernel(const BQuery* __restrict__ qs, const int32_t* __restrict__ pqs,
                                                          uint32_t* __restrict__ hist)
{
    constexpr int K = 8;
    __shared__ uint32_t wsum[16];
    const BQuery& q = qs[pqs[blockIdx.x]];
    const int64_t n = ((int64_t)1 << q.pbits) * q.ntiles;
    uint32_t* a = hist + q.hist_off;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int64_t base = 0; base < n; base += (int64_t)K * blockDim.x) {
        const int64_t i0 = base + (int64_t)threadIdx.x * K;
        uint32_t v[K], tot = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v[k] = i0 + k < n ? a[i0 + k] : 0u;
            tot += v[k];
        }
        uint32_t x = tot;  // inclusive scan over the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t before = carry, round = 0;
        for (int i = 0; i < 16; ++i) {
            if (i < w) before += wsum[i];
            round += wsum[i];
        }
        uint32_t ex = before + x - tot;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (i0 + k < n) a[i0 + k] = ex;
            ex += v[k];
        }
        carry += round;
        __syncthreads();  // wsum is reused
    }
}

// Moves every row of a tile to its partition in the scratch columns (the
// order inside a partition is the order the LDS cursors hand out).
template <int KB>
__global__ __launch_bounds__(kPB) void bpart_scatter_kernel(const BQuery* __restrict__ qs, const BPTile* __restrict__ tiles,
                                                            const uint32_t* __restrict__ offs, void* __restrict__ skeys,
                                                            void* __restrict__ svals, int vbytes)
{
    __shared__ uint32_t cur[1 << kMaxPartBits];
    const BPTile tt = tiles[blockIdx.x];
    const BQuery& q = qs[tt.q];
    const uint32_t P = 1u << q.pbits, mask = P - 1u;
    const uint32_t* o = offs + q.hist_off + tt.t;
    for (uint32_t d = threadIdx.x; d < P; d += kPB) cur[d] = o[(int64_t)d * q.ntiles];
    __syncthreads();
    const int64_t b = (int64_t)tt.t * kPTile, e = min(q.rows, b + kPTile);
    for (int64_t r = b + threadIdx.x; r < e; r += kPB) {
        const uint64_t raw = dev::load_raw(q.keys, KB, r);
        const int64_t dst = q.scratch_off + atomicAdd(&cur[part_of<KB>(raw, mask)], 1u);
        dev::store_bytes(skeys, KB, dst, raw);
        if (svals) dev::store_bytes(svals, vbytes, dst, dev::load_raw(q.vals, vbytes, r));
    }
}

// Group g of query q from its states: op a at st[a * stride].
__device__ __forceinline__ void emit_batch(const BOut& bo, const BQuery& q, int64_t g, uint64_t key_enc,
                                           const unsigned long long* st, int64_t stride)
{
    dev::store_bytes(q.out_keys, bo.key_bytes, g, dev::dec64(key_enc));
#pragma unroll
    for (int j = 0; j < EG_BATCH_MAX_AGGS; ++j) {
        if (j >= bo.nouts) break;
        void* d = q.out_aggs[j];
        const uint64_t s = st[(int64_t)bo.op[j] * stride];
        switch (bo.code[j]) {
        case OUT_U64: dev::store_bytes(d, 8, g, s); break;
        case OUT_F32: ((float*)d)[g] = (float)__longlong_as_double((long long)s); break;
        case OUT_F64: ((double*)d)[g] = __longlong_as_double((long long)s); break;
        case OUT_MAX_I: dev::store_bytes(d, bo.bytes[j], g, s ^ dev::kSign64); break;
        case OUT_MIN_I: dev::store_bytes(d, bo.bytes[j], g, (~s) ^ dev::kSign64); break;
        case OUT_MAX_U: dev::store_bytes(d, bo.bytes[j], g, s); break;
        case OUT_MIN_U: dev::store_bytes(d, bo.bytes[j], g, ~s); break;
        case OUT_MAX_F: dev::store_double_as(d, bo.bytes[j], g, dev::unord_f64(s)); break;
        case OUT_MIN_F: dev::store_double_as(d, bo.bytes[j], g, dev::unord_f64(~s)); break;
        case OUT_ANY:
            if (bo.is_float[j] && bo.bytes[j] == 4) ((float*)d)[g] = (float)__longlong_as_double((long long)s);
            else dev::store_bytes(d, bo.bytes[j], g, s);
            break;
        default: {
            const double cnt = (double)(uint64_t)st[(int64_t)bo.op2[j] * stride];
            const double sum = (bo.code[j] == OUT_AVG_I) ? (double)(int64_t)s
                             : (bo.code[j] == OUT_AVG_U) ? (double)s
                             : __longlong_as_double((long long)s);
            ((double*)d)[g] = sum / cnt;
            break;
        }
        }
    }
}

// Eight waves per SIMD, so two workgroups fit a CU (blocks_per_cu = 2).
template <int KB, int NV>
__global__ __launch_bounds__(kBatchBlock, 8) void batch_agg_kernel(const BQuery* __restrict__ qs,
                                                                const BItem* __restrict__ items, int n_items,
                                                                unsigned int* next_item, unsigned long long* metas,
                                                                AggSpec spec, int key_signed, int vbytes, int W,
                                                                int cr_log2, int probe_limit, int peel_min,
                                                                const uint32_t* __restrict__ offs, const void* skeys,
                                                                const void* svals, unsigned long long* counts,
                                                                uint32_t* qstatus, BOut bo)
{
    extern __shared__ __align__(16) unsigned long long smem[];
    __shared__ int s_item;
    __shared__ unsigned long long sh[17];
    const int nops = spec.nops;
    const int lane = dev::lane_id();
    LdsTable<KB, true> tab(smem, nops, 1 << cr_log2, 0);
    const int words = LdsTable<KB>::words(nops, 1 << cr_log2, 0);
    uint32_t ins = 0;
    dev::HotTags hot;  // groups the wave combined last
    KeyDesc key;
    key.bytes = KB;
    key.is_signed = key_signed;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(next_item, 1u);
        __syncthreads();
        const int it = s_item;
        if (it >= n_items) break;  // the same value in every thread: the whole block leaves
        const BItem w = items[it];
        const BQuery q = qs[w.q];
        const GTable gt = query_table(q, W, metas + (size_t)w.q * kMetaWords);
        if (w.kind == 1) {
            // One partition of a partitioned query: its rows from the scratch
            // columns into the LDS table, its groups straight to the output.
            const uint32_t d = (uint32_t)w.begin;
            const uint32_t* o = offs + q.hist_off;
            const int64_t pb = q.scratch_off + (int64_t)o[(int64_t)d * q.ntiles];
            const int64_t pe = q.scratch_off + (d + 1 < (1u << q.pbits) ? (int64_t)o[(int64_t)(d + 1) * q.ntiles] : q.rows);
            key.data = skeys;
            for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
            __syncthreads();
            for (int64_t b = pb; b < pe; b += blockDim.x) {
                const int64_t r = b + threadIdx.x;
                const bool valid = r < pe;
                const uint64_t kraw = valid ? dev::load_raw(skeys, KB, r) : 0ull;
                uint64_t cv[NV];
#pragma unroll
                for (int c = 0; c < NV; ++c)
                    cv[c] = (valid && spec.nvcols > 0) ? dev::canon(dev::load_raw(svals, vbytes, r), spec.cols[0].conv) : 0ull;
                const bool sp = lds_row<dev::RtOps, KB, NV>(tab, true, 0, lane, key, spec, kraw, cv, valid, probe_limit, peel_min, hot);
                if (__ballot(sp) != 0ull) hbm_rows<dev::RtOps, KB, NV>(gt, key, spec, sp, kraw, cv, 65, ins, hot);
            }
            __syncthreads();
            using KT = typename LdsKey<KB>::T;
            auto slot_key = [&](int sl, KT& kw) {
                kw = 0;
                if (sl < tab.cr) {
                    kw = tab.keys[sl];
                    return kw != 0;
                }
                return sl == tab.cr && tab.side[0] != 0ull;
            };
            const int wv = threadIdx.x >> 6;
            const int nw = (blockDim.x + 63) >> 6;
            uint32_t wave_groups = 0;
            for (int base = 0; base <= tab.cr; base += blockDim.x) {
                KT kw;
                wave_groups += (uint32_t)__popcll(__ballot(slot_key(base + threadIdx.x, kw)));
            }
            if (lane == 0) sh[wv] = wave_groups;
            __syncthreads();
            if (threadIdx.x == 0) {
                unsigned long long tot = 0;
                for (int i = 0; i < nw; ++i) {
                    const unsigned long long c = sh[i];
                    sh[i] = tot;
                    tot += c;
                }
                sh[16] = tot ? atomicAdd(&counts[w.q], tot) : 0ull;
            }
            __syncthreads();
            int64_t pos = (int64_t)(sh[16] + sh[wv]);
            for (int base = 0; base <= tab.cr; base += blockDim.x) {
                const int sl = base + threadIdx.x;
                KT kw;
                const bool occ = slot_key(sl, kw);
                const uint64_t m = __ballot(occ);
                if (occ) {
                    const int64_t idx = pos + __popcll(m & dev::lanemask_lt());
                    if (idx < q.capacity) emit_batch(bo, q, idx, lds_to_enc64<KB>(kw, key_signed), tab.st + sl, (int64_t)tab.T);
                    else atomicOr(&qstatus[w.q], kStCapacity);
                }
                pos += __popcll(m);
            }
            __syncthreads();  // LDS, sh and s_item are reused by the next item
            continue;
        }
        key.data = q.keys;
        const bool use_lds = q.use_lds != 0;
        if (use_lds) {
            for (int i = threadIdx.x; i < words; i += blockDim.x) smem[i] = 0ull;
            __syncthreads();
        }
        for (int64_t b = w.begin; b < w.end; b += blockDim.x) {
            const int64_t r = b + threadIdx.x;
            const bool valid = r < w.end;
            const uint64_t kraw = valid ? dev::load_raw(q.keys, KB, r) : 0ull;
            uint64_t cv[NV];
#pragma unroll
            for (int c = 0; c < NV; ++c)
                cv[c] = (valid && spec.nvcols > 0) ? dev::canon(dev::load_raw(q.vals, vbytes, r), spec.cols[0].conv)
                                                   : 0ull;
            bool sp = valid;
            if (use_lds) sp = lds_row<dev::RtOps, KB, NV>(tab, true, 0, lane, key, spec, kraw, cv, valid, probe_limit, peel_min, hot);
            if (__ballot(sp) != 0ull) hbm_rows<dev::RtOps, KB, NV>(gt, key, spec, sp, kraw, cv, use_lds ? 65 : peel_min, ins, hot);
        }
        if (use_lds) {
            __syncthreads();
            for (int s = threadIdx.x; s <= tab.cr; s += blockDim.x) {
                typename LdsKey<KB>::T kw = 0;
                if (s == tab.cr) {
                    if (tab.side[0] == 0ull) continue;
                } else {
                    kw = tab.keys[s];
                    if (kw == 0) continue;
                }
                unsigned long long* g = gt_slot(gt, lds_to_enc64<KB>(kw, key_signed), ins);
                if (!g) continue;
                dev::merge_states<dev::RtOps>(spec, [&](int a) { return g + 1 + a; },
                                  [&](int a) { return tab.st[(size_t)a * tab.T + s]; });
            }
        }
        __syncthreads();  // LDS and s_item are reused by the next item
    }
}

__global__ void batch_compact_kernel(const BQuery* __restrict__ qs, const BCItem* __restrict__ citems,
                                     unsigned long long* metas, unsigned long long* counts, uint32_t* qstatus,
                                     BOut bo, int W)
{
    __shared__ unsigned long long sh[17];
    const BCItem c = citems[blockIdx.x];
    const BQuery q = qs[c.q];
    unsigned long long* meta = metas + (size_t)c.q * kMetaWords;
    if (c.slot_begin == 0 && threadIdx.x == 0 && meta[2] != 0ull) atomicOr(&qstatus[c.q], kStTableFull);
    for (uint64_t base = c.slot_begin; base < c.slot_end; base += blockDim.x) {
        const uint64_t s = base + threadIdx.x;
        bool occ = false;
        const unsigned long long* slot = nullptr;
        uint64_t enc = 0;
        if (s < c.slot_end) {
            if (s <= q.mask) {
                slot = q.table + s * (uint64_t)W;
                enc = slot[0];
                occ = enc != 0;
            } else {
                slot = q.table + (q.mask + 1) * (uint64_t)W;  // side slot: key 0 encoding
                occ = meta[1] != 0;
            }
        }
        const int64_t idx = dev::block_claim(&counts[c.q], occ, sh);
        if (occ) {
            if (idx < q.capacity) emit_batch(bo, q, idx, enc, slot + 1, 1);
            else atomicOr(&qstatus[c.q], kStCapacity);
        }
    }
}

int ceil_pow2_log2(int64_t x)
{
    int b = 0;
    while (((int64_t)1 << b) < x) ++b;
    return b;
}

}  // namespace

// Host-side plan: everything one execution needs, in device memory.
struct BatchPlan {
    int64_t nq = 0;
    int32_t key_type = 0, val_type = 0;
    int kb = 4, vbytes = 8, nv = 0;
    AggSpec spec{};
    BOut bo{};
    int W = 0;
    eg_batch_tuning tuning{};
    DevBuf qs, items, citems, tables, metas;  // metas: per-query meta + counts + status + queue
    DevBuf ptiles, pqs, hist, skeys, svals;     // the partition pass (radix_bits > 0)
    int n_items = 0, n_citems = 0, n_ptiles = 0, n_pq = 0;
    uint64_t table_words = 0;
    unsigned long long* counts = nullptr;
    uint32_t* qstatus = nullptr;
    unsigned int* next_item = nullptr;
    uint64_t meta_words = 0;  // words of `metas` cleared per run
    size_t lds = 0;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    double last_ms = 0.0;

    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
    }
};

void batch_default_tuning(eg_batch_tuning* t)
{
    t->tile_rows_log2 = 16;
    t->lds_slots_log2 = 11;
    t->lds_max_bound_log2 = 11;
    t->blocks_per_cu = 1;
    t->radix_bits = 12;
}

BatchPlan* batch_plan_create(const eg_batch_query* queries, int64_t nq, int32_t key_type, int32_t val_type,
                             const eg_agg_request* reqs, int64_t nreqs, const eg_batch_tuning* tuning, hipStream_t s)
{
    if (nq <= 0 || !queries) raise(EG_ERR_INVALID, "empty batch");
    if (nq >= ((int64_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "too many queries in one batch");
    auto p = std::make_unique<BatchPlan>();
    p->nq = nq;
    p->key_type = key_type;
    p->val_type = val_type;
    if (tuning) p->tuning = *tuning;
    else batch_default_tuning(&p->tuning);
    const eg_batch_tuning& T = p->tuning;
    if (T.tile_rows_log2 < 10 || T.tile_rows_log2 > 26 || T.lds_slots_log2 < 6 || T.lds_slots_log2 > 13 ||
        T.blocks_per_cu < 1 || T.blocks_per_cu > 8 || T.radix_bits < 0 || T.radix_bits > kMaxPartBits)
        raise(EG_ERR_INVALID, "batch tuning out of range");

    // Lower the requests once: every query shares the key/value types.
    static const char dummy[16] = {0};
    ViewIn kv{dummy, 1, key_type};
    std::vector<ViewIn> vv;
    bool uses_vals = false;
    for (int64_t r = 0; r < nreqs; ++r)
        if (reqs[r].agg_kind != EG_AGG_COUNT_ALL) {
            if (reqs[r].column_index != 0) raise(EG_ERR_INVALID, "batched queries have one value column (index 0)");
            uses_vals = true;
        }
    if (uses_vals) vv.push_back(ViewIn{dummy, 1, val_type});
    Plan plan = make_plan(&kv, vv, reqs, nreqs, false, false);
    if ((int)plan.outs.size() > EG_BATCH_MAX_AGGS) raise(EG_ERR_UNSUPPORTED, "too many requests for a batch");
    p->spec = plan.spec;
    p->kb = plan.key.bytes;
    p->nv = plan.spec.nvcols;
    p->vbytes = uses_vals ? type_size(val_type) : 8;
    p->W = table_words_per_slot(plan.spec.nops);
    p->bo.nouts = (int32_t)plan.outs.size();
    p->bo.key_bytes = plan.key.bytes;
    for (size_t j = 0; j < plan.outs.size(); ++j) {
        p->bo.code[j] = plan.outs[j].code;
        p->bo.bytes[j] = type_size(plan.outs[j].type);
        p->bo.op[j] = plan.outs[j].op;
        p->bo.op2[j] = plan.outs[j].op2;
        p->bo.is_float[j] = is_float(plan.outs[j].type) ? 1 : 0;
    }

    // LDS table of the aggregation kernel (slots clamped to what fits)
    const DeviceInfo& di = device_info();
    int cr = T.lds_slots_log2;
    auto lds_of = [&](int c) { return (size_t)lds::lds_table_words(p->spec.nops, 1 << c, 0, p->kb) * 8; };
    while (cr > 6 && lds_of(cr) + 256 > (size_t)di.max_lds) --cr;
    p->tuning.lds_slots_log2 = cr;
    p->lds = lds_of(cr);
    const int64_t per_part = (int64_t)1 << (cr - 1);  // groups per partition: half a table

    // Per-query tables, work items and compaction ranges.
    const int W = p->W;
    const int64_t tile = (int64_t)1 << T.tile_rows_log2;
    const int64_t lds_bound = (int64_t)1 << T.lds_max_bound_log2;
    std::vector<BQuery> hq((size_t)nq);
    std::vector<BItem> hitems;
    std::vector<int64_t> item_rows;  // estimated rows per item (longest first)
    std::vector<BCItem> hc;
    std::vector<BPTile> htiles;
    std::vector<int32_t> hpq;
    std::vector<uint64_t> offs((size_t)nq);
    uint64_t words = 0;
    int64_t hist_entries = 0, scratch_rows = 0;
    constexpr uint64_t kCompactRange = 1u << 16;
    for (int64_t i = 0; i < nq; ++i) {
        const eg_batch_query& u = queries[i];
        if (u.rows < 0 || u.capacity < 0) raise(EG_ERR_INVALID, "negative batch query size");
        if (u.rows > 0 && (!u.keys || (uses_vals && !u.vals))) raise(EG_ERR_INVALID, "null batch query column");
        if (u.capacity > 0 && !u.out_keys) raise(EG_ERR_INVALID, "null batch query output");
        for (int j = 0; j < p->bo.nouts; ++j)
            if (u.capacity > 0 && !u.out_aggs[j]) raise(EG_ERR_INVALID, "null batch query output");
        const int64_t bound = std::max<int64_t>(1, std::min<int64_t>(u.group_bound > 0 ? u.group_bound : u.rows, std::max<int64_t>(u.rows, 1)));
        // above the LDS bound: partitioned (its HBM table only takes the keys
        // a partition's LDS table cannot hold), or per-row HBM-table atomics
        const bool part = T.radix_bits > 0 && bound > lds_bound && u.rows > 0 && u.rows < ((int64_t)1 << 31);
        const int pbits = part ? std::min(T.radix_bits, std::max(1, ceil_pow2_log2((bound + per_part - 1) / per_part))) : 0;
        // a partitioned query's HBM table takes what its partitions' LDS
        // tables cannot: the groups beyond half a table per partition (when
        // radix_bits caps the partition count), plus an eighth of the bound
        const int64_t excess = part ? std::max<int64_t>(0, bound - (per_part << pbits)) : 0;
        const int64_t slots = part ? (int64_t)1 << std::max(6, ceil_pow2_log2(2 * (excess + bound / 8)))
                                   : (int64_t)1 << std::max(4, ceil_pow2_log2(2 * bound));
        BQuery& q = hq[(size_t)i];
        std::memset(&q, 0, sizeof q);
        q.keys = u.keys;
        q.vals = u.vals;
        q.mask = (uint64_t)slots - 1;
        q.rows = u.rows;
        q.capacity = u.capacity;
        q.out_keys = u.out_keys;
        for (int j = 0; j < EG_BATCH_MAX_AGGS; ++j) q.out_aggs[j] = u.out_aggs[j];
        q.use_lds = bound <= lds_bound ? 1 : 0;
        offs[(size_t)i] = words;
        words += (uint64_t)(slots + 1) * W;
        if (part) {
            q.pbits = pbits;
            q.ntiles = (u.rows + kPTile - 1) / kPTile;
            q.hist_off = hist_entries;
            q.scratch_off = scratch_rows;
            hist_entries += ((int64_t)1 << q.pbits) * q.ntiles;
            scratch_rows += u.rows;
            hpq.push_back((int32_t)i);
            for (int64_t t = 0; t < q.ntiles; ++t) htiles.push_back(BPTile{(int32_t)i, (int32_t)t});
            for (int64_t d = 0; d < ((int64_t)1 << q.pbits); ++d) {
                hitems.push_back(BItem{(int32_t)i, 1, d, d + 1});
                item_rows.push_back(u.rows >> q.pbits);
            }
        } else {
            for (int64_t b = 0; b < u.rows; b += tile) {
                hitems.push_back(BItem{(int32_t)i, 0, b, std::min(u.rows, b + tile)});
                item_rows.push_back(std::min(u.rows, b + tile) - b);
            }
        }
        for (uint64_t b = 0; b < (uint64_t)slots + 1; b += kCompactRange)
            hc.push_back(BCItem{(int32_t)i, 0, b, std::min<uint64_t>((uint64_t)slots + 1, b + kCompactRange)});
    }
    // Longest items first keeps the tail of the persistent grid short.
    {
        std::vector<size_t> order(hitems.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return item_rows[a] > item_rows[b]; });
        std::vector<BItem> sorted(hitems.size());
        for (size_t k = 0; k < order.size(); ++k) sorted[k] = hitems[order[k]];
        hitems.swap(sorted);
    }
    if (hist_entries >= ((int64_t)1 << 40) || htiles.size() >= ((size_t)1 << 31))
        raise(EG_ERR_UNSUPPORTED, "batch partition pass too large");
    if (hitems.size() >= ((size_t)1 << 31) || hc.size() >= ((size_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "batch too large");
    p->n_items = (int)hitems.size();
    p->n_citems = (int)hc.size();
    p->table_words = words;
    p->tables = DevBuf((size_t)std::max<uint64_t>(words, 1) * 8, s);
    for (int64_t i = 0; i < nq; ++i) hq[(size_t)i].table = p->tables.get<unsigned long long>() + offs[(size_t)i];

    // metas: per-query meta words, then counts, then status, then the item queue
    const uint64_t meta_q = (uint64_t)nq * kMetaWords;
    const uint64_t mw = meta_q + (uint64_t)nq + ((uint64_t)nq + 1) / 2 + 1;
    p->meta_words = mw;
    p->metas = DevBuf((size_t)mw * 8, s);
    unsigned long long* m = p->metas.get<unsigned long long>();
    p->counts = m + meta_q;
    p->qstatus = reinterpret_cast<uint32_t*>(m + meta_q + nq);
    p->next_item = reinterpret_cast<unsigned int*>(m + mw - 1);

    p->qs = DevBuf(hq.size() * sizeof(BQuery), s);
    p->items = DevBuf(std::max<size_t>(hitems.size(), 1) * sizeof(BItem), s);
    p->citems = DevBuf(std::max<size_t>(hc.size(), 1) * sizeof(BCItem), s);
    EG_HIP(hipMemcpyAsync(p->qs.get(), hq.data(), hq.size() * sizeof(BQuery), hipMemcpyHostToDevice, s));
    if (!hitems.empty())
        EG_HIP(hipMemcpyAsync(p->items.get(), hitems.data(), hitems.size() * sizeof(BItem), hipMemcpyHostToDevice, s));
    EG_HIP(hipMemcpyAsync(p->citems.get(), hc.data(), hc.size() * sizeof(BCItem), hipMemcpyHostToDevice, s));
    p->n_ptiles = (int)htiles.size();
    p->n_pq = (int)hpq.size();
    if (!htiles.empty()) {
        p->ptiles = DevBuf(htiles.size() * sizeof(BPTile), s);
        p->pqs = DevBuf(hpq.size() * sizeof(int32_t), s);
        p->hist = DevBuf((size_t)hist_entries * 4, s);
        p->skeys = DevBuf((size_t)scratch_rows * p->kb, s);
        if (uses_vals) p->svals = DevBuf((size_t)scratch_rows * p->vbytes, s);
        EG_HIP(hipMemcpyAsync(p->ptiles.get(), htiles.data(), htiles.size() * sizeof(BPTile), hipMemcpyHostToDevice, s));
        EG_HIP(hipMemcpyAsync(p->pqs.get(), hpq.data(), hpq.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    EG_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope

    p->grid = di.num_cus * T.blocks_per_cu;
    // raise the kernel's LDS limit now: no attribute calls during graph capture
    const void* kern = p->kb == 4 ? (const void*)batch_agg_kernel<4, 1> : (const void*)batch_agg_kernel<8, 1>;
    if (enable_dynamic_lds(kern) < p->lds) raise(EG_ERR_UNSUPPORTED, "batch LDS table too large");
    return p.release();
}

template <int KB, int NV>
void launch_batch_agg(BatchPlan& p, hipStream_t s)
{
    auto kern = batch_agg_kernel<KB, NV>;
    hipLaunchKernelGGL(kern, dim3(p.grid), dim3(kBatchBlock), p.lds, s, p.qs.get<BQuery>(), p.items.get<BItem>(),
                       p.n_items, p.next_item, p.metas.get<unsigned long long>(), p.spec, is_signed_int(p.key_type) ? 1 : 0,
                       p.vbytes, p.W, p.tuning.lds_slots_log2, 16, 8, (const uint32_t*)p.hist.get<uint32_t>(),
                       (const void*)p.skeys.get(), (const void*)p.svals.get(), p.counts, p.qstatus, p.bo);
    EG_LAUNCH_CHECK();
}

template <int KB>
void launch_partition_pass(BatchPlan& p, hipStream_t s)
{
    hipLaunchKernelGGL(bpart_hist_kernel<KB>, dim3((unsigned)p.n_ptiles), dim3(kPB), 0, s, p.qs.get<BQuery>(),
                       p.ptiles.get<BPTile>(), p.hist.get<uint32_t>());
    EG_LAUNCH_CHECK();
    hipLaunchKernelGGL(bpart_scan_kernel, dim3((unsigned)p.n_pq), dim3(1024), 0, s, p.qs.get<BQuery>(),
                       p.pqs.get<int32_t>(), p.hist.get<uint32_t>());
    EG_LAUNCH_CHECK();
    hipLaunchKernelGGL(bpart_scatter_kernel<KB>, dim3((unsigned)p.n_ptiles), dim3(kPB), 0, s, p.qs.get<BQuery>(),
                       p.ptiles.get<BPTile>(), (const uint32_t*)p.hist.get<uint32_t>(), p.skeys.get(), p.svals.get(),
                       p.vbytes);
    EG_LAUNCH_CHECK();
}

// One execution: no host synchronisation, no allocation.
void batch_run(BatchPlan& p, hipStream_t s)
{
    const DeviceInfo& di = device_info();
    auto fill = [&](unsigned long long* ptr, uint64_t words) {
        if (!words) return;
        const uint64_t g = std::min<uint64_t>((words / 2 + 255) / 256 + 1, (uint64_t)di.num_cus * 16);
        hipLaunchKernelGGL(batch_fill_kernel, dim3((unsigned)g), dim3(256), 0, s, ptr, words);
        EG_LAUNCH_CHECK();
    };
    fill(p.tables.get<unsigned long long>(), p.table_words);
    fill(p.metas.get<unsigned long long>(), p.meta_words);
    if (p.n_ptiles > 0) {
        KernelTimer timer("batch_partition", s);
        if (p.kb == 4) launch_partition_pass<4>(p, s);
        else launch_partition_pass<8>(p, s);
    }
    {
        KernelTimer timer("batch_agg", s);
        if (p.kb == 4) launch_batch_agg<4, 1>(p, s);
        else launch_batch_agg<8, 1>(p, s);
    }
    KernelTimer timer("batch_compact", s);
    hipLaunchKernelGGL(batch_compact_kernel, dim3(p.n_citems), dim3(256), 0, s, p.qs.get<BQuery>(), p.citems.get<BCItem>(),
                       p.metas.get<unsigned long long>(), p.counts, p.qstatus, p.bo, p.W);
    EG_LAUNCH_CHECK();
}

void batch_graph_launch(BatchPlan& p, hipStream_t s)
{
    if (!p.exec) {
        if (timing_enabled()) raise(EG_ERR_INVALID, "kernel timing records events: disable it before graph capture");
        if (!p.capture_stream) EG_HIP(hipStreamCreateWithFlags(&p.capture_stream, hipStreamNonBlocking));
        EG_HIP(hipStreamBeginCapture(p.capture_stream, hipStreamCaptureModeThreadLocal));
        try {
            batch_run(p, p.capture_stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(p.capture_stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        EG_HIP(hipStreamEndCapture(p.capture_stream, &p.graph));
        EG_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
    }
    EG_HIP(hipGraphLaunch(p.exec, s));
}

void batch_results(BatchPlan& p, int64_t* num_groups, int32_t* status, hipStream_t s)
{
    std::vector<unsigned long long> c((size_t)p.nq);
    std::vector<uint32_t> st((size_t)p.nq);
    EG_HIP(hipMemcpyAsync(c.data(), p.counts, (size_t)p.nq * 8, hipMemcpyDeviceToHost, s));
    EG_HIP(hipMemcpyAsync(st.data(), p.qstatus, (size_t)p.nq * 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < p.nq; ++i) {
        if (num_groups) num_groups[i] = (int64_t)c[(size_t)i];
        if (status) status[i] = (int32_t)st[(size_t)i];
    }
}

}  // namespace eg

struct eg_batch_plan_s {
    eg::BatchPlan* p;
};

extern "C" {

void eg_batch_tuning_default(eg_batch_tuning* t)
{
    if (t) eg::batch_default_tuning(t);
}

int eg_batch_plan_create(const eg_batch_query* queries, int64_t num_queries, int32_t key_type, int32_t val_type,
                         const eg_agg_request* reqs, int64_t num_reqs, const eg_batch_tuning* tuning, void* stream,
                         eg_batch_plan* out)
{
    return eg::guard([&] {
        if (!out) eg::raise(EG_ERR_INVALID, "null output");
        *out = nullptr;
        eg::BatchPlan* p = eg::batch_plan_create(queries, num_queries, key_type, val_type, reqs, num_reqs, tuning,
                                                 (hipStream_t)stream);
        *out = new eg_batch_plan_s{p};
    });
}

int eg_batch_plan_tuning(eg_batch_plan plan, eg_batch_tuning* out)
{
    return eg::guard([&] {
        if (!plan || !out) eg::raise(EG_ERR_INVALID, "null argument");
        *out = plan->p->tuning;
    });
}

int eg_batch_run(eg_batch_plan plan, void* stream)
{
    return eg::guard([&] {
        if (!plan) eg::raise(EG_ERR_INVALID, "null plan");
        eg::batch_run(*plan->p, (hipStream_t)stream);
    });
}

int eg_batch_graph_launch(eg_batch_plan plan, void* stream)
{
    return eg::guard([&] {
        if (!plan) eg::raise(EG_ERR_INVALID, "null plan");
        eg::batch_graph_launch(*plan->p, (hipStream_t)stream);
    });
}

int eg_batch_results(eg_batch_plan plan, int64_t* num_groups, int32_t* status, void* stream)
{
    return eg::guard([&] {
        if (!plan) eg::raise(EG_ERR_INVALID, "null plan");
        eg::batch_results(*plan->p, num_groups, status, (hipStream_t)stream);
    });
}

void eg_batch_plan_free(eg_batch_plan plan)
{
    if (!plan) return;
    delete plan->p;
    delete plan;
}

int eg_batch_autotune(const eg_batch_query* queries, int64_t num_queries, int32_t key_type, int32_t val_type,
                      const eg_agg_request* reqs, int64_t num_reqs, const eg_batch_tuning* candidates,
                      int64_t num_candidates, int32_t reps, void* stream, eg_batch_tuning* best, double* times_ms)
{
    return eg::guard([&] {
        if (!best || !candidates || num_candidates <= 0) eg::raise(EG_ERR_INVALID, "no candidate tunings");
        hipStream_t s = (hipStream_t)stream;
        hipEvent_t e0, e1;
        EG_HIP(hipEventCreate(&e0));
        EG_HIP(hipEventCreate(&e1));
        double best_ms = 1e300;
        int64_t best_i = -1;
        for (int64_t c = 0; c < num_candidates; ++c) {
            std::unique_ptr<eg::BatchPlan> p(
                eg::batch_plan_create(queries, num_queries, key_type, val_type, reqs, num_reqs, &candidates[c], s));
            eg::batch_run(*p, s);  // warm-up
            EG_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < std::max(1, reps); ++r) eg::batch_run(*p, s);
            EG_HIP(hipEventRecord(e1, s));
            EG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            EG_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double per = (double)ms / std::max(1, reps);
            if (times_ms) times_ms[c] = per;
            if (per < best_ms) {
                best_ms = per;
                best_i = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *best = candidates[best_i];
    });
}

}  // extern "C"

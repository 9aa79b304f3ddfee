// hash_groupby.hip -- hash-based group-by for MI355X (gfx950).
//
// Replaces gamax::groupby::hash::groupby (operators/groupby/groupby_hash.cu:302-426).
// The reference builds dense group ids with a cuco static_set of row indices,
// then runs a second pass of per-row global atomics (compute_groupid_kernel
// :48-75, compute_aggregation :77-99).  Here one streaming pass does both:
//
//   1. hash_agg_kernel: a persistent grid, one workgroup per CU slot.  Each
//      workgroup keeps private open-addressing tables in LDS holding the
//      encoded key and every aggregation state inline.  For small group counts
//      the table is replicated R times (R up to 64) and lane l uses replica
//      l mod R, so lanes of a wave never collide on one LDS address; the
//      replicas are merged in LDS before the flush.  Rows are read with
//      16-byte vector loads, four consecutive rows per lane, and the next
//      tile is loaded while the current one is aggregated.  A row that
//      cannot find a slot within `probe_limit` probes spills to the HBM
//      table.  Hot keys (many lanes of a wave on one group) are combined by
//      the wave first (eg_atomics.hpp) so one atomic per state is issued.
//   2. hash_compact_kernel: scans the HBM table and writes the final key and
//      output columns at wave-aggregated positions.
//
// The HBM table stores slots of W = 2^k 64-bit words [enc key | states...]
// (AoS, so a random update touches one cache line).  The all-zero word is
// "empty" for keys and the identity of every state, so the table is cleared
// with a plain zero fill.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "eg_agg_kernels.hpp"
#include "eg_atomics.hpp"
#include "eg_device.hpp"
#include "eg_hash_table.hpp"
#include "eg_internal.hpp"
#include "eg_jit.hpp"
#include "eg_lds_table.hpp"

namespace eg {
namespace {

using namespace lds;
using agg::max_block;

template <int KB, class VL>
__global__ __launch_bounds__(max_block<VL::NV>()) void hash_agg_kernel(KeyDesc key, int64_t n, AggSpec spec,
                                                                        GTable gt, int cr, int rep_log2,
                                                                        int probe_limit, int peel_min)
{
    agg::hash_agg<KB, VL, dev::RtOps>(key, n, spec, gt, cr, rep_log2, probe_limit, peel_min);
}

// Groups at indices from `capacity` on are not written (the caller sees the
// count and re-runs the compaction with room for them).
template <int R>
__global__ __launch_bounds__(256) void hash_compact_kernel(GTable gt, OutSpec os, unsigned long long* counter,
                                                           int64_t capacity)
{
    __shared__ unsigned long long sh[17];
    compact_groups<R>(gt, counter, sh, [&](int64_t idx, uint64_t enc, const unsigned long long* st) {
        if (idx < capacity) dev::emit_group(os, idx, enc, st, 1);
    });
}

// Compaction of gt into os from *counter on (grid and slots per thread by size).
void launch_compact(const GTable& gt, const OutSpec& os, unsigned long long* counter, hipStream_t s,
                    int64_t capacity = INT64_MAX)
{
    const DeviceInfo& di = device_info();
    const uint64_t total = gt.mask + 2;
    const bool wide = total >= kCompactWideSlots;
    const int64_t per = 256 * (wide ? kCompactSlots : 1);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)((total + per - 1) / per), (int64_t)di.num_cus * 8));
    KernelTimer timer("hash_compact", s);
    if (wide)
        hipLaunchKernelGGL(hash_compact_kernel<kCompactSlots>, dim3((unsigned)grid), dim3(256), 0, s, gt, os, counter,
                           capacity);
    else hipLaunchKernelGGL(hash_compact_kernel<1>, dim3((unsigned)grid), dim3(256), 0, s, gt, os, counter, capacity);
    EG_LAUNCH_CHECK();
}

// use_jit: take the specialised kernel once it is loaded (loading a module
// allocates device memory for its code, so callers that must not allocate
// pass false and always run the generic kernel).
template <int KB, class VL>
void launch_agg(const DeviceInfo& di, const HashTuning& t, const KeyDesc& key, int64_t n, const AggSpec& spec,
                const GTable& gt, const HashLayout& L, hipStream_t s, bool use_jit)
{
    auto kern = hash_agg_kernel<KB, VL>;
    enable_dynamic_lds((const void*)kern);
    const int block = std::min(t.block, max_block<VL::NV>());
    const int64_t tile = (int64_t)block * kRows;
    const int64_t want = (n + tile - 1) / tile;
    auto grid_for = [&](int per_cu) {
        return std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)di.num_cus * std::max(per_cu, 1)));
    };
    KernelTimer timer("hash_agg", s);
    // the kernel with this request's ops compiled in, once it is available
    if (hipFunction_t f = use_jit ? jit::lookup(jit::Kind::HASH_AGG, KB, jit::vlayout_name<VL>(), spec, n) : nullptr) {
        const int64_t grid = grid_for(t.blocks_per_cu > 0 ? t.blocks_per_cu : jit::occupancy(f, block, L.lds_bytes));
        KeyDesc a_key = key;
        int64_t a_n = n;
        AggSpec a_spec = spec;
        GTable a_gt = gt;
        int a_cr = L.cr, a_rep = L.rep_log2, a_probe = L.probe_limit, a_peel = L.peel_min;
        void* args[] = {&a_key, &a_n, &a_spec, &a_gt, &a_cr, &a_rep, &a_probe, &a_peel};
        if (jit::launch(f, (unsigned)grid, (unsigned)block, L.lds_bytes, s, args)) return;
    }
    int per_cu = t.blocks_per_cu;
    if (per_cu <= 0) EG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, L.lds_bytes));
    const int64_t grid = grid_for(per_cu);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), L.lds_bytes, s, key, n, spec, gt, L.cr,
                       L.rep_log2, L.probe_limit, L.peel_min);
    EG_LAUNCH_CHECK();
}

// Picks the compiled value layout: vector loads for up to two 4/8-byte
// columns on 16-byte aligned data, the generic typed-load kernel otherwise.
template <int KB>
void dispatch_layout(bool vec, const DeviceInfo& di, const HashTuning& t, const KeyDesc& key, int64_t n,
                     const AggSpec& spec, const GTable& gt, const HashLayout& L, hipStream_t s, bool use_jit = true)
{
    with_vlayout(spec, vec, [&](auto vl) { launch_agg<KB, decltype(vl)>(di, t, key, n, spec, gt, L, s, use_jit); });
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

HashLayout plan_hash_layout(int nops, int key_bytes, int64_t group_bound, const HashTuning& t)
{
    const DeviceInfo& di = device_info();
    HashLayout L;
    auto bytes_of = [&](int cr_log2, int rep_log2) {
        return (size_t)lds::lds_table_words(nops, 1 << cr_log2, rep_log2, key_bytes) * 8;
    };
    const size_t budget = (size_t)di.max_lds;
    // replicas: the most (<= 64) that still give every replica >= 2x the group bound
    int best_c = -1, best_r = 0;
    for (int rl = 6; rl >= 0; --rl) {
        if (t.replicas > 0 && (1 << rl) > t.replicas) continue;
        int cl = -1;
        for (int c = 14; c >= 4; --c)
            if (bytes_of(c, rl) <= budget && (t.lds_slots == 0 || ((1 << c) << rl) <= t.lds_slots)) {
                cl = c;
                break;
            }
        if (cl < 0) continue;
        if (rl == 0 || (int64_t)(1 << cl) >= 2 * group_bound) {
            best_c = cl;
            best_r = rl;
            break;
        }
    }
    if (best_c >= 0 && group_bound <= (int64_t)t.lds_max_ratio * ((int64_t)1 << best_c)) {
        L.cr = 1 << best_c;
        L.rep_log2 = best_r;
        if (best_r == 0) {
            // one table: every multiple of four slots the LDS holds (a larger
            // table keeps more of the keys that arrive, and fewer rows spill)
            auto bytes_cr = [&](int cr) { return (size_t)lds::lds_table_words(nops, cr, 0, key_bytes) * 8; };
            const int cap = t.lds_slots > 0 ? t.lds_slots : (1 << 14);
            while (L.cr + 4 <= cap && bytes_cr(L.cr + 4) + 1024 <= budget) L.cr += 4;  // 1 KB for static LDS
        }
        L.lds_bytes = (size_t)lds::lds_table_words(nops, L.cr, L.rep_log2, key_bytes) * 8;
    }
    // A table much smaller than the group count still catches the hot keys
    // (they arrive early); keep the probing short so cold rows spill cheaply.
    const int64_t slots = std::max(L.cr, 1);
    L.probe_limit = group_bound > 4 * slots ? 2 : group_bound > slots ? 4 : 16;
    L.peel_min = t.peel_min;
    return L;
}

// Group bounds up to which run_hash allocates the outputs before it reads
// the group count (at most 40 bytes a group for the benchmark's requests).
constexpr int64_t kEarlyOutputRows = (int64_t)1 << 22;

Result run_hash(const Plan& plan, int64_t group_bound, hipStream_t s, const HashTuning& t)
{
    const DeviceInfo& di = device_info();
    const int64_t n = plan.rows;
    const int nops = plan.spec.nops;
    const int W = table_words_per_slot(nops);
    const int KB = plan.key.bytes;
    const HashLayout L = plan_hash_layout(nops, KB, group_bound, t);

    bool vec = aligned16(plan.key.data);
    for (int c = 0; c < plan.spec.nvcols; ++c) vec = vec && aligned16(plan.spec.cols[c].data);

    int64_t bound = std::max<int64_t>(1, std::min<int64_t>(group_bound, n));
    int64_t slots = t.table_slots > 0 ? t.table_slots : 1024;
    while (slots < 2 * bound) slots <<= 1;

    // With a modest bound the outputs are allocated for it before the table is
    // read, and the compaction runs right after the aggregation: one
    // synchronisation per call instead of two (a bound the groups exceed
    // costs one more compaction into exact outputs).
    const bool early = bound <= kEarlyOutputRows;
    for (;;) {
        HbmTable ht(slots, W, s);
        if (KB == 4) dispatch_layout<4>(vec, di, t, plan.key, n, plan.spec, ht.gt, L, s);
        else dispatch_layout<8>(vec, di, t, plan.key, n, plan.spec, ht.gt, L, s);
        Result eres;
        OutSpec eos;
        if (early) {
            alloc_outputs(plan, bound, s, eres, eos);
            launch_compact(ht.gt, eos, ht.gt.meta + 3, s, bound);  // meta[3]: zeroed with the table
        }
        int64_t meta[3];
        ht.read_meta(meta, s);
        if (meta[2] != 0) {
            // The table filled up: the group bound was too small.  Retry with a
            // table sized for every row being its own group.
            if (slots >= 2 * n) raise(EG_ERR_CAPACITY, "hash table overflow");
            timing_count("hash_retry");
            int64_t ns = slots;
            while (ns < 2 * n) ns <<= 1;
            slots = ns;
            continue;
        }
        const int64_t groups = meta[0] + (meta[1] != 0 ? 1 : 0);
        if (early && groups <= bound) {
            eres.num_groups = groups;
            return eres;  // read_meta synchronised after the compaction
        }
        Result res;
        OutSpec os;
        alloc_outputs(plan, groups, s, res, os);
        if (groups > 0) {
            DevBuf counter(8, s);
            fill_zero(counter.get(), 8, s);
            ht.compact(os, counter.get<unsigned long long>(), s);
        }
        EG_HIP(hipStreamSynchronize(s));
        return res;
    }
}

// ---- no-allocation variant ---------------------------------------------------------
// The HBM table of run_hash_noalloc lives in this array: one copy per device,
// loaded with the module and held until the process ends.  Host code never
// passes the array's name to a kernel (on the host it only stands for the
// symbol); it uses the device address that hipGetSymbolAddress returns for
// the current device, looked up once per device.
namespace {
__device__ __align__(256) unsigned long long scratch[kNoAllocScratchBytes / 8];

unsigned long long* scratch_ptr()
{
    static std::mutex mu;
    static std::vector<void*> ptrs;
    int dev = 0;
    EG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((size_t)dev >= ptrs.size()) ptrs.resize((size_t)dev + 1, nullptr);
    if (!ptrs[(size_t)dev]) {
        void* p = nullptr;
        EG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) raise(EG_ERR_DEVICE, "no device address for the group-by scratch");
        ptrs[(size_t)dev] = p;
    }
    return static_cast<unsigned long long*>(ptrs[(size_t)dev]);
}
}  // namespace

int64_t run_hash_noalloc(const Plan& plan, int64_t group_bound, const OutSpec& os, int64_t capacity, hipStream_t s)
{
    const DeviceInfo& di = device_info();
    const int64_t n = plan.rows;
    const int nops = plan.spec.nops;
    const int W = table_words_per_slot(nops);
    const int KB = plan.key.bytes;
    const HashTuning t;
    const HashLayout L = plan_hash_layout(nops, KB, group_bound, t);
    bool vec = aligned16(plan.key.data);
    for (int c = 0; c < plan.spec.nvcols; ++c) vec = vec && aligned16(plan.spec.cols[c].data);

    // table: slots * W words, then the side slot (W words) and 4 meta words;
    // meta[3] is the compaction's output counter
    const size_t max_words = kNoAllocScratchBytes / 8;
    auto words_of = [&](int64_t slots) { return (size_t)slots * (size_t)W + (size_t)W + 4; };
    const int64_t bound = std::max<int64_t>(1, std::min<int64_t>(group_bound, n));
    int64_t slots = 1024;
    while (slots < 2 * bound) slots <<= 1;
    if (words_of(slots) > max_words)
        raise(EG_ERR_CAPACITY, "groups_upper_bound " + std::to_string(group_bound) + " needs a table of " +
                                   std::to_string(words_of(slots) * 8) + " bytes, above the " +
                                   std::to_string(kNoAllocScratchBytes) + "-byte no-allocation scratch");
    int64_t max_slots = slots;
    while (words_of(max_slots * 2) <= max_words) max_slots <<= 1;

    unsigned long long* base = scratch_ptr();
    for (;;) {
        GTable gt;
        gt.base = base;
        gt.side = base + (size_t)slots * W;
        gt.meta = gt.side + W;
        gt.mask = (uint64_t)slots - 1;
        gt.W = W;
        fill_zero(base, words_of(slots) * 8, s);  // a kernel: every call starts from an empty table
        if (KB == 4) dispatch_layout<4>(vec, di, t, plan.key, n, plan.spec, gt, L, s, false);
        else dispatch_layout<8>(vec, di, t, plan.key, n, plan.spec, gt, L, s, false);
        int64_t meta[3] = {0, 0, 0};
        EG_HIP(hipMemcpyAsync(meta, gt.meta, sizeof meta, hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        if (meta[2] != 0) {
            // more groups than the bound said: once more with the largest
            // table the scratch holds
            if (slots < max_slots) {
                slots = max_slots;
                continue;
            }
            raise(EG_ERR_CAPACITY, "more groups than the no-allocation scratch's table holds (" +
                                       std::to_string(slots / 2) + " at load 1/2)");
        }
        const int64_t groups = meta[0] + (meta[1] != 0 ? 1 : 0);
        if (groups > capacity)
            raise(EG_ERR_CAPACITY, std::to_string(groups) + " groups exceed the output capacity of " +
                                       std::to_string(capacity) + " rows");
        if (groups > 0) {
            launch_compact(gt, os, gt.meta + 3, s);
        }
        EG_HIP(hipStreamSynchronize(s));
        return groups;
    }
}

// ---- HbmTable ---------------------------------------------------------------------
int table_words_per_slot(int nops)
{
    int p = 1;
    while (p < 1 + nops) p <<= 1;
    return p;
}

HbmTable::HbmTable(int64_t slots, int W, hipStream_t s)
    : buf((size_t)(slots * W + W + 4) * 8, s)
{
    const size_t words = (size_t)(slots * W + W + 4);
    fill_zero(buf.get(), words * 8, s);
    gt.base = buf.get<unsigned long long>();
    gt.side = gt.base + (size_t)slots * W;
    gt.meta = gt.side + W;
    gt.mask = (uint64_t)slots - 1;
    gt.W = W;
}

void HbmTable::read_meta(int64_t* out3, hipStream_t s)
{
    int64_t* h = pinned_i64(4);
    EG_HIP(hipMemcpyAsync(h, gt.meta, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    out3[0] = h[0];
    out3[1] = h[1];
    out3[2] = h[2];
}

void HbmTable::compact(const OutSpec& os, unsigned long long* counter, hipStream_t s)
{
    launch_compact(gt, os, counter, s);
}

}  // namespace eg

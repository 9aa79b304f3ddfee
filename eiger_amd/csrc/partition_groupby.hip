// partition_groupby.hip -- partition-based group-by
// (gamax::groupby::partition::groupby / groupby_gfur,
// operators/groupby/groupby_partition.cu:556-707).
//
// The reference computes pid = fmix(key) & 0xFFFF, sorts (pid, key) and every
// value column with a 16-bit CUB radix sort each (groupby_partition.cu:580-663),
// then aggregates one partition per block in a shared-memory cuco set of 8,209
// row indices and throws when a partition has more distinct keys
// (:116-297, 543-549).  Here:
//   1. the number of partitions P = 2^b is chosen so that a partition holds
//      about a thousand groups (from hint.groups_upper_bound or the HyperLogLog
//      estimate);
//   2. LSD radix passes on the fmix bits move the key and every value column
//      together (PARTITION) or the key and row ids (PARTITION_GFUR, values
//      gathered at aggregation time);
//   3. part_bounds_kernel finds each partition's first row (binary search);
//   4. part_agg_kernel aggregates a unit (a partition, or a slice of a large
//      one) in an LDS table and writes the finished groups of single-unit
//      partitions straight to the output.  Groups of sliced partitions, and
//      keys that do not fit the LDS table, go through the HBM table and are
//      appended by its compaction -- a spill, where the reference throws.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <cstdio>
#include <memory>
#include <vector>

#include "eg_agg_kernels.hpp"
#include "eg_atomics.hpp"
#include "eg_device.hpp"
#include "eg_hash_table.hpp"
#include "eg_internal.hpp"
#include "eg_jit.hpp"
#include "eg_lds_table.hpp"
#include "eg_radix.hpp"

namespace eg {
namespace {

using namespace lds;

// Widest digit of a partitioning pass (fan-out 256); EG_PARTITION_DIGIT_BITS
// (4..8) lowers it for tuning.  Passes over the same bits are balanced.
constexpr int kMaxDigitBits = 8;

int max_digit_bits()
{
    static const int v = [] {
        const char* e = std::getenv("EG_PARTITION_DIGIT_BITS");
        const int b = e ? std::atoi(e) : 0;
        return (b >= 4 && b <= 8) ? b : kMaxDigitBits;
    }();
    return v;
}

// Unit kernel geometry: 1024-thread workgroups whose LDS table takes what
// leaves room for two workgroups per CU (160 KB of LDS: about 79 KB each,
// e.g. 1788 slots for the five 8-byte states and 4-byte keys of the
// benchmark), filled to 2/3 by design.  EG_PARTITION_TUNE="slots,threads"
// fixes both for tuning (slots a multiple of 4).
struct PartTune {
    int cr = 0;  // 0: from the LDS budget
    int block = 1024;
};
PartTune part_tune()
{
    PartTune t;
    if (const char* e = std::getenv("EG_PARTITION_TUNE")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%d,%d", &a, &b) == 2 && a >= 64 && a <= 16384 && a % 4 == 0 &&
            (b == 256 || b == 512 || b == 1024)) {
            t.cr = a;
            t.block = b;
        }
    }
    return t;
}

using agg::kPBMax;
using agg::Unit;

using agg::OutClaims;

// Eight waves per SIMD: two 1024-thread workgroups per CU, which the LDS
// table is sized for (without the bound the compiler's scalar registers
// alone leave room for only one).
template <int KB, class VL, bool GATHER>
__global__ __launch_bounds__(kPBMax, 8) void part_agg_kernel(KeyDesc key, AggSpec spec, const int64_t* perm,
                                                       const Unit* units, GTable gt, OutSpec os,
                                                       OutClaims claims, int64_t capacity,
                                                       int cr, int probe, int peel_min)
{
    agg::part_agg<KB, VL, dev::RtOps, GATHER>(key, spec, perm, units, gt, os, claims, capacity, cr, probe,
                                              peel_min);
}

template <int KB>
__device__ __forceinline__ uint32_t pid_of(uint64_t raw, uint32_t mask)
{
    return (uint32_t)dev::bit_mix<KB>(raw) & mask;
}

// First row of every partition: after the radix passes the rows are ordered
// by partition id, so thread p binary-searches the first row whose id is
// >= p (offs[P] = n).  About log2(n) dependent loads per thread instead of a
// pass over every key.
template <int KB>
__global__ void part_bounds_kernel(const void* keys, int64_t n, uint32_t mask, int64_t* offs)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > (int64_t)mask + 1) return;
    int64_t lo = 0, hi = n;  // invariant: rows < lo have id < p, rows >= hi have id >= p
    while (lo < hi) {        // at most 1 + log2(n) iterations
        const int64_t mid = lo + ((hi - lo) >> 1);
        if ((int64_t)pid_of<KB>(dev::load_raw(keys, KB, mid), mask) < p) lo = mid + 1;
        else hi = mid;
    }
    offs[p] = lo;
}

// Appends the HBM table's groups after the directly written ones.
__global__ __launch_bounds__(256) void part_compact_kernel(GTable gt, OutSpec os, unsigned long long* counter,
                                                           int64_t capacity)
{
    __shared__ unsigned long long sh[17];
    compact_groups<kCompactSlots>(gt, counter, sh, [&](int64_t idx, uint64_t enc, const unsigned long long* st) {
        if (idx < capacity) dev::emit_group(os, idx, enc, st, 1);
    });
}

// Units from the partition boundaries, on the device: a partition of at most
// `slice` rows is one unit, a larger one ceil(rows / slice) sliced units.
// units_count_kernel counts each partition's units (cnt) and the sliced
// partitions (meta[0]); after an exclusive scan of cnt (pre),
// units_fill_kernel writes them and the unit total (meta[1]).
__global__ __launch_bounds__(256) void units_count_kernel(const int64_t* __restrict__ offs, int64_t P, int64_t slice,
                                                          uint32_t* __restrict__ cnt, unsigned long long* meta)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int64_t rows = offs[p + 1] - offs[p];
    cnt[p] = rows <= 0 ? 0u : rows <= slice ? 1u : (uint32_t)((rows + slice - 1) / slice);
    if (rows > slice) atomicAdd(&meta[0], 1ull);
}

__global__ __launch_bounds__(256) void units_fill_kernel(const int64_t* __restrict__ offs, int64_t P, int64_t slice,
                                                         const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ pre, agg::Unit* __restrict__ units,
                                                         unsigned long long* meta)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const uint32_t k = cnt[p];
    const int64_t a = offs[p], e = offs[p + 1];
    for (uint32_t i = 0; i < k; ++i) {
        const int64_t x = a + (int64_t)i * slice;
        units[pre[p] + i] = agg::Unit{x, k == 1 ? e : (x + slice < e ? x + slice : e), k > 1 ? 1 : 0, 0};
    }
    if (p == P - 1) meta[1] = (unsigned long long)pre[p] + k;
}

// Output rows past the group count moved into the holes below it (the
// unused tails of the lanes' last chunks): op i copies len[i] rows from
// src[i] to dst[i] in every output column; sources and destinations are
// disjoint (every source is >= the group count, every destination below it).
constexpr int kMaxMoves = 32;
struct MoveList {
    int64_t src[kMaxMoves], dst[kMaxMoves], len[kMaxMoves];
};

__global__ void move_rows_kernel(OutSpec os, MoveList ml)
{
    const int op = blockIdx.y;
    const int64_t len = ml.len[op];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = ml.src[op] + i, d = ml.dst[op] + i;
        dev::store_bytes(os.key_data, os.key_bytes, d, dev::load_raw(os.key_data, os.key_bytes, s));
        for (int j = 0; j < os.nouts; ++j)
            dev::store_bytes(os.outs[j].data, os.outs[j].bytes, d, dev::load_raw(os.outs[j].data, os.outs[j].bytes, s));
    }
}

int ceil_log2(int64_t x)
{
    int b = 0;
    while (((int64_t)1 << b) < x) ++b;
    return b;
}

template <int KB, bool GATHER>
void launch_part_agg(const KeyDesc& key, const AggSpec& spec, const int64_t* perm, const Unit* units, int nunits,
                     const GTable& gt, const OutSpec& os, const OutClaims& claims, int64_t capacity, int cr,
                     size_t lds, int block, int64_t rows, hipStream_t s)
{
    with_vlayout(spec, true, [&](auto vl) {
        using VL = decltype(vl);
        auto kern = part_agg_kernel<KB, VL, GATHER>;
        if (enable_dynamic_lds((const void*)kern) < lds) raise(EG_ERR_UNSUPPORTED, "partition LDS table too large");
        KernelTimer timer("partition_agg", s);
        int probe = 16, peel_min = default_peel_min();  // probe: buckets of four slots
        // the kernel with this request's ops compiled in, once it is available
        const jit::Kind kind = GATHER ? jit::Kind::PART_AGG_GATHER : jit::Kind::PART_AGG;
        if (hipFunction_t f = jit::lookup(kind, KB, jit::vlayout_name<VL>(), spec, rows, &os)) {
            KeyDesc a_key = key;
            AggSpec a_spec = spec;
            const int64_t* a_perm = perm;
            const Unit* a_units = units;
            GTable a_gt = gt;
            OutSpec a_os = os;
            OutClaims a_claims = claims;
            int64_t a_cap = capacity;
            int a_cr = cr;
            void* args[] = {&a_key, &a_spec, &a_perm, &a_units, &a_gt, &a_os, &a_claims, &a_cap, &a_cr,
                            &probe, &peel_min};
            if (jit::launch(f, (unsigned)nunits, (unsigned)block, lds, s, args)) return;
        }
        hipLaunchKernelGGL(kern, dim3(nunits), dim3(block), lds, s, key, spec, perm, units, gt, os, claims, capacity,
                           cr, probe, peel_min);
        EG_LAUNCH_CHECK();
    });
}

template <int KB>
void launch_part(bool gather, const KeyDesc& key, const AggSpec& spec, const int64_t* perm, const Unit* units,
                 int nunits, const GTable& gt, const OutSpec& os, const OutClaims& claims, int64_t capacity,
                 int cr, size_t lds, int block, int64_t rows, hipStream_t s)
{
    if (gather)
        launch_part_agg<KB, true>(key, spec, perm, units, nunits, gt, os, claims, capacity, cr, lds, block,
                                  rows, s);
    else
        launch_part_agg<KB, false>(key, spec, perm, units, nunits, gt, os, claims, capacity, cr, lds, block,
                                   rows, s);
}

}  // namespace

Result run_partition(const Plan& plan, int64_t group_bound, bool gfur, hipStream_t s)
{
    const DeviceInfo& di = device_info();
    const int64_t n = plan.rows;
    const int kb = plan.key.bytes;
    const int nops = plan.spec.nops;
    const int64_t bound = std::max<int64_t>(1, std::min<int64_t>(group_bound, n));

    // LDS table and workgroup (see PartTune); the partition count is chosen
    // so a partition holds about 2/3 of the table.
    const PartTune tune = part_tune();
    const size_t static_lds = 256;
    auto lds_of = [&](int c) { return (size_t)agg::unit_table_words(nops, c, kb) * 8; };
    int cr = tune.cr;
    if (cr == 0) {
        const size_t budget = std::min<size_t>((size_t)di.max_lds, 160 * 1024) / 2 - static_lds;
        cr = 4;
        while (cr < 16384 && lds_of(cr + 4) <= budget) cr += 4;
    }
    while (cr > 64 && lds_of(cr) + static_lds > (size_t)di.max_lds) cr -= 4;
    const size_t lds = lds_of(cr);
    const int64_t per_part = std::max<int64_t>(1, (int64_t)cr * 2 / 3);  // target groups per partition
    const int bits = std::min(24, std::max(1, ceil_log2((bound + per_part - 1) / per_part)));
    const int digit_bits = max_digit_bits();
    const uint32_t pmask = (uint32_t)((1u << bits) - 1u);

    std::unique_ptr<HostPhase> phase_prep(new HostPhase("host_partition_prep", s));
    // ---- radix passes on the hash bits -----------------------------------------
    RadixKeys rk{};
    rk.data = plan.key.data;
    rk.bytes = kb;
    rk.type = plan.key_type;
    rk.mode = DIG_HASH;
    SortBuffers b;
    b.keys_in = plan.key.data;
    DevBuf ka((size_t)n * kb, s), kbuf((size_t)n * kb, s);
    b.keys_a = ka.get();
    b.keys_b = kbuf.get();
    std::vector<DevBuf> bufs;
    AggSpec spec = plan.spec;
    if (gfur) {
        b.pay_in.push_back(Payload{nullptr, nullptr, 8, 1});
        bufs.emplace_back((size_t)n * 8, s);
        bufs.emplace_back((size_t)n * 8, s);
        b.pay_a.push_back(bufs[0].get());
        b.pay_b.push_back(bufs[1].get());
    } else {
        for (int c = 0; c < spec.nvcols; ++c) {
            const int by = spec.cols[c].bytes;
            b.pay_in.push_back(Payload{spec.cols[c].data, nullptr, by, 0});
            bufs.emplace_back((size_t)n * by, s);
            bufs.emplace_back((size_t)n * by, s);
            b.pay_a.push_back(bufs[bufs.size() - 2].get());
            b.pay_b.push_back(bufs[bufs.size() - 1].get());
        }
    }
    // balanced digits: ceil(bits / digit_bits) passes of near-equal width
    phase_prep.reset();
    const int passes = (bits + digit_bits - 1) / digit_bits;
    const void* pkeys = nullptr;
    std::vector<const void*> ppay;
    {
        HostPhase phase("host_partition_radix", s);
        // LSD passes over [0, bits) of near-equal width
        const int w = (bits + passes - 1) / passes;
        const void* kin = b.keys_in;
        std::vector<Payload> pay = b.pay_in;
        bool to_a = true;
        for (int shift = 0; shift < bits; shift += w) {
            const int bw = std::min(w, bits - shift);
            void* kout = to_a ? b.keys_a : b.keys_b;
            for (size_t c = 0; c < pay.size(); ++c) pay[c].out = to_a ? b.pay_a[c] : b.pay_b[c];
            RadixKeys k = rk;
            k.data = kin;
            radix_pass(k, kout, pay, n, shift, bw, s);
            kin = kout;
            for (auto& p : pay) {
                p.in = p.out;
                if (p.rowid) {
                    p.rowid = 0;
                    p.bytes = 8;
                }
            }
            to_a = !to_a;
        }
        pkeys = kin;
        for (auto& p : pay) ppay.push_back(p.in);
    }
    const int64_t* perm = nullptr;
    if (gfur) perm = (const int64_t*)ppay[0];
    else
        for (int c = 0; c < spec.nvcols; ++c) spec.cols[c].data = ppay[(size_t)c];
    KeyDesc pkey = plan.key;
    pkey.data = pkeys;

    // ---- partition boundaries and units -------------------------------------------
    const int64_t P = (int64_t)pmask + 1;
    std::unique_ptr<HostPhase> phase_units(new HostPhase("host_partition_units", s));
    DevBuf offs((size_t)(P + 1) * 8, s);
    {
        const int64_t grid = (P + 1 + 255) / 256;  // one thread per boundary
        KernelTimer timer("partition_bounds", s);
        if (kb == 4) hipLaunchKernelGGL(part_bounds_kernel<4>, dim3((unsigned)grid), dim3(256), 0, s, pkeys, n, pmask, offs.get<int64_t>());
        else hipLaunchKernelGGL(part_bounds_kernel<8>, dim3((unsigned)grid), dim3(256), 0, s, pkeys, n, pmask, offs.get<int64_t>());
        EG_LAUNCH_CHECK();
    }
    // Rows per unit: a partition is split when it is much larger than the
    // average (skew), or when there are too few partitions to occupy every CU,
    // since a split partition's groups merge through the HBM table with
    // atomics instead of being written straight out.
    const int64_t avg_rows = (n + P - 1) / P;
    const int64_t balance = std::max<int64_t>(1, n / (4 * (int64_t)di.num_cus));
    const int64_t slice = std::max<int64_t>({(int64_t)tune.block * 16, 8 * per_part, std::min(4 * avg_rows, balance)});
    // units built on the device; 16 bytes (unit and sliced-partition counts)
    // come back instead of the boundaries, and no unit list goes up
    const int64_t max_units = P + n / slice + 1;
    if (max_units >= ((int64_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "too many partition units");
    DevBuf dunits((size_t)max_units * sizeof(Unit), s);
    int64_t nunits = 0, sliced_parts = 0;
    {
        DevBuf ub((size_t)P * 8 + 16, s);
        uint32_t* cnt = ub.get<uint32_t>();
        uint32_t* pre = cnt + P;
        unsigned long long* umeta = reinterpret_cast<unsigned long long*>(pre + P);
        EG_HIP(hipMemsetAsync(umeta, 0, 16, s));
        const unsigned ugrid = (unsigned)((P + 255) / 256);
        hipLaunchKernelGGL(units_count_kernel, dim3(ugrid), dim3(256), 0, s, (const int64_t*)offs.get<int64_t>(), P,
                           slice, cnt, umeta);
        EG_LAUNCH_CHECK();
        scan_exclusive_u32(cnt, pre, P, s, false);
        hipLaunchKernelGGL(units_fill_kernel, dim3(ugrid), dim3(256), 0, s, (const int64_t*)offs.get<int64_t>(), P,
                           slice, (const uint32_t*)cnt, (const uint32_t*)pre, dunits.get<Unit>(), umeta);
        EG_LAUNCH_CHECK();
        int64_t* hm = pinned_i64(2);
        EG_HIP(hipMemcpyAsync(hm, umeta, 16, hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        sliced_parts = hm[0];
        nunits = hm[1];
    }

    phase_units.reset();
    HostPhase phase_agg("host_partition_aggregate", s);
    // ---- aggregation ------------------------------------------------------------------
    // HBM table: groups of sliced partitions plus LDS spills
    int64_t capacity = bound;
    int64_t slots = 1 << 14;
    // sliced partitions' groups, plus keys the LDS tables may not hold (a
    // partition above its expected size, long probe runs): 1/32 of the bound
    const int64_t spill_groups = std::min<int64_t>(bound, sliced_parts * per_part * 2) + bound / 32;
    while (slots < 2 * spill_groups) slots <<= 1;
    const int W = table_words_per_slot(nops);
    // Output chunk of the lane claims (see agg::OutClaims): a power of two
    // above a unit's largest group count (its slots + the side slot), and at
    // least 1/256 of the bound (up to 2^16 rows) so few chunks are claimed.
    int64_t chunk = 4096;
    while (chunk < (int64_t)cr + 1) chunk <<= 1;
    while (chunk < 65536 && chunk * 256 < bound) chunk <<= 1;
    for (;;) {
        // room for the unfilled tails of the lanes' last chunks
        const int64_t cap_rows = capacity + (int64_t)agg::kClaimLanes * chunk;
        Result res;
        OutSpec os;
        alloc_outputs(plan, cap_rows, s, res, os);
        HbmTable ht(slots, W, s);
        OutClaims claims;
        claims.chunk = chunk;
        claims.maxc = cap_rows / chunk + 2;
        const size_t claim_bytes = (size_t)agg::claim_words(claims.maxc) * 8;
        DevBuf cbuf(claim_bytes, s);
        fill_zero(cbuf.get(), claim_bytes, s);
        claims.words = cbuf.get<unsigned long long>();
        if (kb == 4)
            launch_part<4>(gfur, pkey, spec, perm, dunits.get<Unit>(), (int)nunits, ht.gt, os, claims,
                           cap_rows, cr, lds, tune.block, n, s);
        else
            launch_part<8>(gfur, pkey, spec, perm, dunits.get<Unit>(), (int)nunits, ht.gt, os, claims,
                           cap_rows, cr, lds, tune.block, n, s);
        {
            // append HBM-table groups after every claimed chunk (also when
            // the table filled up: the outputs are then dropped for the
            // retry below, and the capacity bounds every write)
            const uint64_t total = ht.gt.mask + 2;
            const int64_t per = 256 * kCompactSlots;
            const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)((total + per - 1) / per), (int64_t)di.num_cus * 8));
            KernelTimer timer("partition_compact", s);
            hipLaunchKernelGGL(part_compact_kernel, dim3((unsigned)grid), dim3(256), 0, s, ht.gt, os,
                               claims.words, cap_rows);
            EG_LAUNCH_CHECK();
        }
        // the claims' head and the table's meta words in one synchronisation
        int64_t* h = (int64_t*)pinned_staging((size_t)agg::kClaimHeadWords * 8, 2);
        int64_t* hm = pinned_i64(4);
        EG_HIP(hipMemcpyAsync(h, claims.words, (size_t)agg::kClaimHeadWords * 8, hipMemcpyDeviceToHost, s));
        EG_HIP(hipMemcpyAsync(hm, ht.gt.meta, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        const int64_t meta[3] = {hm[0], hm[1], hm[2]};
        const bool table_full = meta[2] != 0;
        const int64_t total = h[0];  // output rows claimed: chunks, then the HBM table's groups
        if (table_full || total > cap_rows) {
            // the HBM table or the output ran out: retry with room for every row
            if (capacity >= n && slots >= 2 * n) raise(EG_ERR_CAPACITY, "partition group-by overflow");
            timing_count(table_full ? "partition_retry_table" : "partition_retry_capacity");
            if (total > cap_rows) capacity = n;
            if (table_full) {
                int64_t ns = slots;
                while (ns < 2 * std::max<int64_t>(bound, meta[0] * 2)) ns <<= 1;
                slots = std::max<int64_t>(ns, slots * 4);
            }
            continue;
        }
        // Rows written: every lane's logical rows (its chunks, the last one
        // partly used) plus the HBM table's groups after the chunks.  Rows at
        // or past the group count move into the holes below it.
        int64_t fills = 0, chunk_rows = 0;
        std::vector<std::pair<int64_t, int64_t>> holes;
        for (int x = 0; x < agg::kClaimLanes; ++x) {
            const int64_t f = h[32 * (1 + x)];
            const int64_t nch = (f + chunk - 1) / chunk;
            fills += f;
            chunk_rows += nch * chunk;
            if (f % chunk != 0) {
                const uint64_t last = (uint64_t)h[1 + x];
                if ((int64_t)(last >> 40) != nch - 1) raise(EG_ERR_DEVICE, "partition output claims inconsistent");
                const int64_t base = (int64_t)(last & ((1ull << 40) - 1));
                holes.emplace_back(base + f % chunk, base + chunk);
            }
        }
        const int64_t groups = fills + (total - chunk_rows);
        if (total < chunk_rows || groups > capacity + (int64_t)agg::kClaimLanes * chunk)
            raise(EG_ERR_DEVICE, "partition output claims inconsistent");
        std::sort(holes.begin(), holes.end());
        std::vector<std::pair<int64_t, int64_t>> dst, src;  // [begin, end) runs
        int64_t at = groups;
        for (const auto& hl : holes) {
            if (hl.first < groups) dst.emplace_back(hl.first, std::min(hl.second, groups));
            if (hl.second <= at) continue;
            if (hl.first > at) src.emplace_back(at, std::min(hl.first, total));
            at = std::max(at, hl.second);
        }
        if (at < total) src.emplace_back(at, total);
        MoveList ml{};
        int nm = 0;
        int64_t longest = 0;
        size_t si = 0, di2 = 0;
        int64_t so = 0, doff = 0;
        while (di2 < dst.size() && si < src.size()) {
            const int64_t len = std::min(dst[di2].second - dst[di2].first - doff, src[si].second - src[si].first - so);
            if (nm == kMaxMoves) raise(EG_ERR_DEVICE, "partition output holes: too many moves");
            ml.src[nm] = src[si].first + so;
            ml.dst[nm] = dst[di2].first + doff;
            ml.len[nm] = len;
            longest = std::max(longest, len);
            ++nm;
            doff += len;
            so += len;
            if (doff == dst[di2].second - dst[di2].first) { ++di2; doff = 0; }
            if (so == src[si].second - src[si].first) { ++si; so = 0; }
        }
        if (nm > 0) {
            const unsigned gx = (unsigned)std::min<int64_t>((longest + 255) / 256, 1024);
            hipLaunchKernelGGL(move_rows_kernel, dim3(gx, (unsigned)nm), dim3(256), 0, s, os, ml);
            EG_LAUNCH_CHECK();
        }
        res.num_groups = groups;
        return res;
    }
}

}  // namespace eg

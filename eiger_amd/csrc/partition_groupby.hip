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

// Digit width of one partitioning pass.  Fan-out 64 keeps the rows of one
// digit in runs long enough for full-line stores (measured: a 6-bit pass on
// 2^28 rows takes about half the time of an 8-bit one).
constexpr int kMaxDigitBits = 6;

// Widest digit of a partitioning pass; EG_PARTITION_DIGIT_BITS (4..8)
// overrides it for tuning.
int max_digit_bits()
{
    static const int v = [] {
        const char* e = std::getenv("EG_PARTITION_DIGIT_BITS");
        const int b = e ? std::atoi(e) : 0;
        return (b >= 4 && b <= 8) ? b : kMaxDigitBits;
    }();
    return v;
}

// Workgroup size and LDS table slots (log2) of the unit kernel; overridable
// for tuning with EG_PARTITION_TUNE="slots_log2,threads".
struct PartTune {
    int cr_log2 = 11;
    int block = 1024;
    bool fixed = false;  // set by the override
};
PartTune part_tune()
{
    PartTune t;
    if (const char* e = std::getenv("EG_PARTITION_TUNE")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%d,%d", &a, &b) == 2 && a >= 6 && a <= 12 && (b == 256 || b == 512 || b == 1024)) {
            t.cr_log2 = a;
            t.block = b;
            t.fixed = true;
        }
    }
    return t;
}

using agg::kPBMax;
using agg::Unit;

template <int KB, class VL, bool GATHER>
__global__ __launch_bounds__(kPBMax) void part_agg_kernel(KeyDesc key, AggSpec spec, const int64_t* perm,
                                                       const Unit* units, GTable gt, OutSpec os,
                                                       unsigned long long* counter, int64_t capacity,
                                                       int cr_log2, int probe_limit, int peel_min)
{
    agg::part_agg<KB, VL, dev::RtOps, GATHER>(key, spec, perm, units, gt, os, counter, capacity, cr_log2,
                                              probe_limit, peel_min);
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
__global__ void part_compact_kernel(GTable gt, OutSpec os, unsigned long long* counter, int64_t capacity)
{
    __shared__ unsigned long long sh[17];
    const uint64_t total = gt.mask + 2;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < total; base += stride) {
        const uint64_t s = base + threadIdx.x;
        bool occ = false;
        const unsigned long long* slot = nullptr;
        uint64_t enc = 0;
        if (s <= gt.mask) {
            slot = gt.base + s * (uint64_t)gt.W;
            enc = slot[0];
            occ = enc != 0;
        } else if (s == gt.mask + 1) {
            slot = gt.side;
            occ = gt.meta[1] != 0;
        }
        const int64_t idx = dev::block_claim(counter, occ, sh);
        if (occ && idx < capacity) dev::emit_group(os, idx, enc, slot + 1, 1);
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
                     const GTable& gt, const OutSpec& os, unsigned long long* counter, int64_t capacity, int cr_log2,
                     size_t lds, int block, int64_t rows, hipStream_t s)
{
    with_vlayout(spec, true, [&](auto vl) {
        using VL = decltype(vl);
        auto kern = part_agg_kernel<KB, VL, GATHER>;
        if (enable_dynamic_lds((const void*)kern) < lds) raise(EG_ERR_UNSUPPORTED, "partition LDS table too large");
        KernelTimer timer("partition_agg", s);
        int probe_limit = 64, peel_min = default_peel_min();
        // the kernel with this request's ops compiled in, once it is available
        const jit::Kind kind = GATHER ? jit::Kind::PART_AGG_GATHER : jit::Kind::PART_AGG;
        if (hipFunction_t f = jit::lookup(kind, KB, jit::vlayout_name<VL>(), spec, rows)) {
            KeyDesc a_key = key;
            AggSpec a_spec = spec;
            const int64_t* a_perm = perm;
            const Unit* a_units = units;
            GTable a_gt = gt;
            OutSpec a_os = os;
            unsigned long long* a_counter = counter;
            int64_t a_cap = capacity;
            int a_cr = cr_log2;
            void* args[] = {&a_key, &a_spec, &a_perm, &a_units, &a_gt, &a_os, &a_counter, &a_cap, &a_cr,
                            &probe_limit, &peel_min};
            if (jit::launch(f, (unsigned)nunits, (unsigned)block, lds, s, args)) return;
        }
        hipLaunchKernelGGL(kern, dim3(nunits), dim3(block), lds, s, key, spec, perm, units, gt, os, counter, capacity,
                           cr_log2, probe_limit, peel_min);
        EG_LAUNCH_CHECK();
    });
}

template <int KB>
void launch_part(bool gather, const KeyDesc& key, const AggSpec& spec, const int64_t* perm, const Unit* units,
                 int nunits, const GTable& gt, const OutSpec& os, unsigned long long* counter, int64_t capacity,
                 int cr_log2, size_t lds, int block, int64_t rows, hipStream_t s)
{
    if (gather)
        launch_part_agg<KB, true>(key, spec, perm, units, nunits, gt, os, counter, capacity, cr_log2, lds, block,
                                  rows, s);
    else
        launch_part_agg<KB, false>(key, spec, perm, units, nunits, gt, os, counter, capacity, cr_log2, lds, block,
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

    // LDS table and workgroup: 1024 slots with 512 threads runs the unit
    // kernel about 1.5x faster than 2048 slots with 1024 threads (three
    // resident workgroups per CU instead of one), but doubles the partition
    // count; take it when that costs no extra radix pass, widening the digits
    // from 6 to 7 bits if that is what it takes (a 7-bit pass costs ~7% more
    // than a 6-bit one; 2^28 rows, D = 2^22: 6.4 ms against 6.7 ms).
    PartTune tune = part_tune();
    auto bits_for = [&](int c) {
        const int64_t pp = (int64_t)1 << (c - 1);
        return std::min(24, std::max(1, ceil_log2((bound + pp - 1) / pp)));
    };
    int digit_bits = max_digit_bits();
    if (!tune.fixed) {
        auto passes = [&](int c, int mb) { return (bits_for(c) + mb - 1) / mb; };
        const int p11 = passes(11, digit_bits);
        if (bound > ((int64_t)1 << 12)) {
            if (passes(10, digit_bits) == p11) {
                tune.cr_log2 = 10;
                tune.block = 512;
            } else if (digit_bits < 7 && passes(10, 7) == p11) {
                tune.cr_log2 = 10;
                tune.block = 512;
                digit_bits = 7;
            }
        }
    }
    int cr_log2 = tune.cr_log2;
    auto lds_of = [&](int c) { return (size_t)lds_table_words(nops, c, 0, kb) * 8; };
    while (cr_log2 > 6 && lds_of(cr_log2) + 256 > (size_t)di.max_lds) --cr_log2;  // 256: static LDS
    const size_t lds = lds_of(cr_log2);
    const int64_t per_part = (int64_t)1 << (cr_log2 - 1);  // target groups per partition (load 1/2)
    const int bits = std::min(24, std::max(1, ceil_log2((bound + per_part - 1) / per_part)));
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
    int64_t* ho = (int64_t*)pinned_staging((size_t)(P + 1) * 8, 0);
    EG_HIP(hipMemcpyAsync(ho, offs.get(), (size_t)(P + 1) * 8, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    // Rows per unit: a partition is split when it is much larger than the
    // average (skew), or when there are too few partitions to occupy every CU,
    // since a split partition's groups merge through the HBM table with
    // atomics instead of being written straight out.
    const int64_t avg_rows = (n + P - 1) / P;
    const int64_t balance = std::max<int64_t>(1, n / (4 * (int64_t)di.num_cus));
    const int64_t slice = std::max<int64_t>({(int64_t)tune.block * 16, 8 * per_part, std::min(4 * avg_rows, balance)});
    std::vector<Unit> units;
    for (int64_t p = 0; p < P; ++p) {
        const int64_t a = ho[p], e = ho[p + 1];
        if (e <= a) continue;
        if (e - a <= slice) {
            units.push_back(Unit{a, e, 0, 0});
        } else {
            for (int64_t x = a; x < e; x += slice) units.push_back(Unit{x, std::min(e, x + slice), 1, 0});
        }
    }
    if (units.size() >= ((size_t)1 << 31)) raise(EG_ERR_UNSUPPORTED, "too many partition units");
    DevBuf dunits(units.size() * sizeof(Unit), s);
    {
        // staged through pinned memory: the copy is a DMA, not a pageable bounce
        void* hu = pinned_staging(units.size() * sizeof(Unit), 1);
        std::memcpy(hu, units.data(), units.size() * sizeof(Unit));
        EG_HIP(hipMemcpyAsync(dunits.get(), hu, units.size() * sizeof(Unit), hipMemcpyHostToDevice, s));
    }

    phase_units.reset();
    HostPhase phase_agg("host_partition_aggregate", s);
    // ---- aggregation ------------------------------------------------------------------
    // HBM table: groups of sliced partitions plus LDS spills
    int64_t sliced_parts = 0;
    for (int64_t p = 0; p < P; ++p) sliced_parts += (ho[p + 1] - ho[p]) > slice ? 1 : 0;
    int64_t capacity = bound;
    int64_t slots = 1 << 14;
    // sliced partitions' groups, plus keys the LDS tables may not hold (a
    // partition above its expected size, long probe runs): 1/32 of the bound
    const int64_t spill_groups = std::min<int64_t>(bound, sliced_parts * per_part * 2) + bound / 32;
    while (slots < 2 * spill_groups) slots <<= 1;
    const int W = table_words_per_slot(nops);
    for (;;) {
        Result res;
        OutSpec os;
        alloc_outputs(plan, capacity, s, res, os);
        HbmTable ht(slots, W, s);
        DevBuf counter(8, s);
        fill_zero(counter.get(), 8, s);
        if (kb == 4)
            launch_part<4>(gfur, pkey, spec, perm, dunits.get<Unit>(), (int)units.size(), ht.gt, os,
                           counter.get<unsigned long long>(), capacity, cr_log2, lds, tune.block, n, s);
        else
            launch_part<8>(gfur, pkey, spec, perm, dunits.get<Unit>(), (int)units.size(), ht.gt, os,
                           counter.get<unsigned long long>(), capacity, cr_log2, lds, tune.block, n, s);
        int64_t meta[3];
        ht.read_meta(meta, s);
        const bool table_full = meta[2] != 0;
        if (!table_full) {
            // append HBM-table groups
            const uint64_t total = ht.gt.mask + 2;
            const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)((total + 255) / 256), (int64_t)di.num_cus * 8));
            KernelTimer timer("partition_compact", s);
            hipLaunchKernelGGL(part_compact_kernel, dim3((unsigned)grid), dim3(256), 0, s, ht.gt, os,
                               counter.get<unsigned long long>(), capacity);
            EG_LAUNCH_CHECK();
        }
        int64_t* h = pinned_i64(1);
        EG_HIP(hipMemcpyAsync(h, counter.get(), 8, hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        const int64_t groups = h[0];
        if (table_full || groups > capacity) {
            // the HBM table or the output ran out: retry with room for every row
            if (capacity >= n && slots >= 2 * n) raise(EG_ERR_CAPACITY, "partition group-by overflow");
            timing_count(table_full ? "partition_retry_table" : "partition_retry_capacity");
            if (groups > capacity) capacity = n;
            if (table_full) {
                int64_t ns = slots;
                while (ns < 2 * std::max<int64_t>(bound, meta[0] * 2)) ns <<= 1;
                slots = std::max<int64_t>(ns, slots * 4);
            }
            continue;
        }
        res.num_groups = groups;
        return res;
    }
}

}  // namespace eg

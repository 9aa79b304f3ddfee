// api.hip -- the extern "C" boundary of libeiger_groupby.so.
//
// Mirrors gamax::groupby::groupby (operators/groupby/groupby.cu:121-465):
// request rewriting (AVG -> SUM + shared count, COUNT_ALL -> the shared count,
// hint.agg_count -> trailing count), strategy selection, and result assembly
// [key, one column per request in order, count].  The reference materialises
// an int64 "ones" column and sums it for counts (groupby.cu:202-231); here the
// count is a state of its own, so no extra column is written or read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "eg_internal.hpp"
#include "eg_skf.hpp"

namespace eg {
const char* last_error_cstr();
int timing_enable(int on);
int timing_reset();
void timing_query(const char* name, int64_t* launches, double* total_ms);
void zipf_cdf(double* cdf, int64_t distinct, double s);
Result run_sort(const Plan& plan, const eg_groupby_hint& hint, bool gfur, hipStream_t s);
Result run_partition(const Plan& plan, int64_t group_bound, bool gfur, hipStream_t s);
bool mlp_select(const void* keys, int32_t key_type, int64_t rows, int64_t group_bound, int nops, int value_bytes,
                hipStream_t s, int32_t* algo, int64_t* card_est, float* singletons);
int64_t ndv_approx(const void* data, int32_t type, int64_t n, hipStream_t s);
void scalar_aggregate(const std::vector<ViewIn>& cols, const eg_agg_request* reqs, int64_t nreqs,
                      hipStream_t s, Result& res);

// ---- plan ------------------------------------------------------------------------
namespace {
struct PlanBuilder {
    Plan p;
    std::map<int64_t, int> vcol_of;  // request column index -> AggSpec col

    int vcol(const ViewIn& v, int64_t idx)
    {
        auto it = vcol_of.find(idx);
        if (it != vcol_of.end()) return it->second;
        if (p.spec.nvcols >= kMaxVCols) throw WideError("too many distinct value columns");
        int c = p.spec.nvcols++;
        p.spec.cols[c].data = v.data;
        p.spec.cols[c].type = v.type;
        p.spec.cols[c].bytes = type_size(v.type);
        p.spec.cols[c].conv = v.type == EG_INT32 ? CONV_SEXT32 : v.type == EG_INT16 ? CONV_SEXT16
                            : v.type == EG_INT8 ? CONV_SEXT8 : v.type == EG_FLOAT32 ? CONV_F32 : CONV_RAW;
        vcol_of[idx] = c;
        return c;
    }

    int op(int32_t code, int col)
    {
        for (int a = 0; a < p.spec.nops; ++a)
            if (p.spec.ops[a].code == code && p.spec.ops[a].col == col) return a;
        if (p.spec.nops >= kMaxOps) throw WideError("too many aggregation states");
        int a = p.spec.nops++;
        p.spec.ops[a].code = code;
        p.spec.ops[a].col = col;
        return a;
    }
};
}  // namespace

// Operand masks of each op (see OpDesc) and the class order the kernels
// expect; output columns are re-pointed at the moved ops.
void order_ops_by_class(Plan& p)
{
    AggSpec& sp = p.spec;
    std::vector<int> order;
    for (int cls = 0; cls < 4; ++cls) {
        sp.ncls[cls] = 0;
        for (int a = 0; a < sp.nops; ++a)
            if (op_class(sp.ops[a].code) == cls) {
                order.push_back(a);
                sp.ncls[cls]++;
            }
    }
    OpDesc old[kMaxOps];
    std::vector<int> where(sp.nops);
    for (int a = 0; a < sp.nops; ++a) old[a] = sp.ops[a];
    for (int i = 0; i < sp.nops; ++i) {
        OpDesc od = old[order[i]];
        od.xmask = op_xmask(od.code);
        od.fmask = op_fmask(od.code);
        sp.ops[i] = od;
        where[order[i]] = i;
    }
    for (auto& o : p.outs) {
        if (o.op >= 0) o.op = where[o.op];
        if (o.op2 >= 0) o.op2 = where[o.op2];
    }
}

Plan make_plan(const ViewIn* key, const std::vector<ViewIn>& vals, const eg_agg_request* reqs,
               int64_t nreqs, bool agg_count, bool rewrite_avg)
{
    if (nreqs <= 0 || !reqs) raise(EG_ERR_INVALID, "No aggregation function is specified");
    PlanBuilder b;
    if (key) {
        if (!(key->type == EG_INT32 || key->type == EG_INT64 || key->type == EG_UINT32 || key->type == EG_UINT64))
            raise(EG_ERR_UNSUPPORTED, "group-by keys must be 4- or 8-byte integers");
        if (key->size < 0) raise(EG_ERR_INVALID, "negative key column size");
        if (key->size > 0 && !key->data) raise(EG_ERR_INVALID, "null key column data");
        b.p.key.data = key->data;
        b.p.key.bytes = type_size(key->type);
        b.p.key.is_signed = is_signed_int(key->type) ? 1 : 0;
        b.p.key_type = key->type;
        b.p.rows = key->size;
    } else {
        // scalar aggregate: no key; rows = the first column's (table_view::num_rows)
        b.p.key.bytes = 8;
        b.p.key_type = EG_INT64;
        b.p.rows = vals.empty() ? 0 : vals[0].size;
    }

    bool needs_count = agg_count;
    for (int64_t r = 0; r < nreqs; ++r) {
        int k = reqs[r].agg_kind;
        if (k == EG_AGG_AVG && !rewrite_avg)
            raise(EG_ERR_UNSUPPORTED, "AVG is only rewritten by the top-level groupby");
        if (k == EG_AGG_AVG || k == EG_AGG_COUNT_ALL) needs_count = true;
    }
    int count_op = needs_count ? b.op(OPC_COUNT, -1) : -1;

    for (int64_t r = 0; r < nreqs; ++r) {
        const int k = reqs[r].agg_kind;
        if (k == EG_AGG_COUNT_ALL) {
            b.p.outs.push_back(OutCol{OUT_U64, EG_INT64, count_op, -1});
            continue;
        }
        if (k < EG_AGG_SUM || k > EG_AGG_ANY) raise(EG_ERR_UNSUPPORTED, "Unsupported aggregation function");
        const int64_t ci = reqs[r].column_index;
        if (ci < 0 || ci >= (int64_t)vals.size()) raise(EG_ERR_INVALID, "aggregation column index out of range");
        const ViewIn& v = vals[(size_t)ci];
        if (!(is_int(v.type) || is_float(v.type))) raise(EG_ERR_UNSUPPORTED, "unsupported value column type");
        if (v.size != b.p.rows) raise(EG_ERR_INVALID, "value column size differs from the key column");
        if (v.size > 0 && !v.data) raise(EG_ERR_INVALID, "null value column data");
        const int col = b.vcol(v, ci);
        const bool sg = is_signed_int(v.type), fl = is_float(v.type);
        const int32_t ot = agg_output_type(k, v.type);
        switch (k) {
        case EG_AGG_SUM: {
            int a = b.op(fl ? OPC_SUM_F : (sg ? OPC_SUM_I : OPC_SUM_U), col);
            b.p.outs.push_back(OutCol{fl ? (v.type == EG_FLOAT32 ? OUT_F32 : OUT_F64) : OUT_U64, ot, a, -1});
            break;
        }
        case EG_AGG_AVG: {
            int a = b.op(fl ? OPC_SUM_F : (sg ? OPC_SUM_I : OPC_SUM_U), col);
            b.p.outs.push_back(OutCol{fl ? OUT_AVG_F : (sg ? OUT_AVG_I : OUT_AVG_U), EG_FLOAT64, a, count_op});
            break;
        }
        case EG_AGG_MAX: {
            int a = b.op(fl ? OPC_MAX_F : (sg ? OPC_MAX_I : OPC_MAX_U), col);
            b.p.outs.push_back(OutCol{fl ? OUT_MAX_F : (sg ? OUT_MAX_I : OUT_MAX_U), ot, a, -1});
            break;
        }
        case EG_AGG_MIN: {
            int a = b.op(fl ? OPC_MIN_F : (sg ? OPC_MIN_I : OPC_MIN_U), col);
            b.p.outs.push_back(OutCol{fl ? OUT_MIN_F : (sg ? OUT_MIN_I : OUT_MIN_U), ot, a, -1});
            break;
        }
        default: {
            int a = b.op(OPC_ANY, col);
            b.p.outs.push_back(OutCol{OUT_ANY, ot, a, -1});
            break;
        }
        }
    }
    if (agg_count) b.p.outs.push_back(OutCol{OUT_U64, EG_INT64, count_op, -1});
    if ((int)b.p.outs.size() > kMaxOuts) raise(EG_ERR_UNSUPPORTED, "too many output columns");
    order_ops_by_class(b.p);
    return b.p;
}

void set_outspec(const Plan& plan, void* key_data, void* const* out_data, OutSpec& os)
{
    std::memset(&os, 0, sizeof os);
    os.key_data = key_data;
    os.key_bytes = plan.key.bytes;
    os.nouts = (int32_t)plan.outs.size();
    for (size_t j = 0; j < plan.outs.size(); ++j) {
        const OutCol& o = plan.outs[j];
        OutDesc& d = os.outs[j];
        d.data = out_data[j];
        d.code = o.code;
        d.bytes = type_size(o.type);
        d.op = o.op;
        d.op2 = o.op2;
        d.is_float = is_float(o.type) ? 1 : 0;
    }
}

void alloc_outputs(const Plan& plan, int64_t groups, hipStream_t s, Result& res, OutSpec& os)
{
    res.num_groups = groups;
    res.cols.clear();
    res.types.clear();
    const size_t rows = (size_t)std::max<int64_t>(groups, 0);
    res.cols.emplace_back(rows * plan.key.bytes, s);
    res.types.push_back(plan.key_type);
    void* outs[kMaxOuts] = {};
    for (size_t j = 0; j < plan.outs.size(); ++j) {
        res.cols.emplace_back(rows * type_size(plan.outs[j].type), s);
        res.types.push_back(plan.outs[j].type);
        outs[j] = res.cols.back().get();
    }
    set_outspec(plan, res.cols[0].get(), outs, os);
}

}  // namespace eg

// ---- extern "C" -----------------------------------------------------------------------
struct eg_table_s {
    struct Col {
        void* data;
        int64_t size;
        int32_t type;
    };
    std::vector<Col> cols;
};

namespace {
using namespace eg;

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

ViewIn view_in(const eg_column_view& v)
{
    if (v.offset < 0) raise(EG_ERR_INVALID, "negative column offset");
    const int bytes = type_size(v.type);
    if (bytes == 0) raise(EG_ERR_UNSUPPORTED, "unsupported column type");
    ViewIn r;
    r.data = v.data ? (const char*)v.data + v.offset * bytes : nullptr;
    r.size = v.size;
    r.type = v.type;
    return r;
}

eg_table to_table(Result& r)
{
    auto* t = new eg_table_s;
    for (size_t j = 0; j < r.cols.size(); ++j)
        t->cols.push_back(eg_table_s::Col{r.cols[j].release(), r.num_groups, r.types[j]});
    return t;
}

int64_t group_bound_of(const Plan& p, const eg_groupby_hint& h, hipStream_t s)
{
    if (h.groups_upper_bound > 0 && h.groups_upper_bound < p.rows) return h.groups_upper_bound;
    if (p.rows <= 4096) return p.rows;
    // No usable bound from the caller: HyperLogLog estimate with headroom.
    int64_t est = ndv_approx(p.key.data, p.key_type, p.rows, s);
    int64_t b = est + est / 4 + 1024;
    return std::min<int64_t>(b, p.rows);
}

enum class Entry { TOP, HASH, SORT, SORT_GFUR, PARTITION, PARTITION_GFUR };

// Requests that need more states or value columns than one pass holds: the
// requests are split into chunks that fit, each chunk is aggregated by the
// sort strategy (rows in ascending key order, so every chunk's rows line up),
// and the result is [key, every request's column in order, trailing count].
Result wide_groupby(const ViewIn& k, const std::vector<ViewIn>& vs, const eg_agg_request* reqs, int64_t nreqs,
                    bool top, bool agg_count, const eg_groupby_hint& h, hipStream_t s)
{
    Result all;
    int64_t groups = -1;
    int64_t r0 = 0;
    while (r0 < nreqs) {
        // the longest prefix of the remaining requests that fits one plan
        int64_t take = 0;
        Plan plan;
        for (int64_t m = 1; r0 + m <= nreqs; ++m) {
            const bool last = r0 + m == nreqs;
            try {
                plan = make_plan(&k, vs, reqs + r0, m, top && agg_count && last, top);
                take = m;
            } catch (const WideError&) {
                break;
            }
        }
        if (take == 0) raise(EG_ERR_UNSUPPORTED, "one aggregation request needs more states than a pass holds");
        plan = make_plan(&k, vs, reqs + r0, take, top && agg_count && r0 + take == nreqs, top);
        Result part;
        if (k.size == 0) {
            OutSpec os;
            alloc_outputs(plan, 0, s, part, os);
        } else {
            part = run_sort(plan, h, false, s);
        }
        if (groups >= 0 && part.num_groups != groups) raise(EG_ERR_DEVICE, "chunked group-by: group counts differ");
        groups = part.num_groups;
        for (size_t j = 0; j < part.cols.size(); ++j) {
            if (j == 0 && !all.cols.empty()) continue;  // the key column once
            all.cols.push_back(std::move(part.cols[j]));
            all.types.push_back(part.types[j]);
        }
        r0 += take;
    }
    all.num_groups = groups;
    return all;
}

int groupby_entry(Entry entry, const eg_column_view* keys, int64_t num_keys,
                  const eg_column_view* vals, int64_t num_vals, const eg_agg_request* reqs,
                  int64_t num_reqs, eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return guard([&] {
        if (!out) raise(EG_ERR_INVALID, "out is null");
        *out = nullptr;
        if (num_keys <= 0 || !keys) raise(EG_ERR_INVALID, "Aggregation only is not supported yet");
        if (num_vals < 0 || (num_vals > 0 && !vals)) raise(EG_ERR_INVALID, "bad value columns");
        eg_groupby_hint local;
        eg_hint_init(&local);
        eg_groupby_hint& h = hint ? *hint : local;
        hipStream_t s = as_stream(stream);

        if (num_keys > 1) {
            // Multi-column keys: the sort and partition strategies take one key
            // column (groupby_sort.cu:209-211, groupby.cu:287-294) unless the
            // keys are fused (SKF, groupby.cu:396-465); hash and AUTO take them.
            // Here every multi-column request runs on the fused key.
            const bool strategy_entry = entry != Entry::TOP && entry != Entry::HASH;
            const bool forced = entry == Entry::TOP && h.groupby_algo != EG_ALGO_AUTO && h.groupby_algo != EG_ALGO_HASH;
            if (strategy_entry || (forced && !h.enable_skf))
                raise(EG_ERR_UNSUPPORTED, "multi-column keys need the hash strategy or key fusion (enable_skf)");
            std::vector<ViewIn> kv;
            for (int64_t i = 0; i < num_keys; ++i) kv.push_back(view_in(keys[i]));
            FusedKeys f;
            fuse_keys(kv.data(), num_keys, s, f);
            eg_column_view fk{f.fused.get(), f.rows, EG_UINT64, 0, 0};
            eg_groupby_hint h2 = h;
            if (f.total_bits < 62 && h2.groups_upper_bound > ((int64_t)1 << f.total_bits))
                h2.groups_upper_bound = (int64_t)1 << f.total_bits;  // SKF tightens the bound (groupby.cu:416)
            eg_table t = nullptr;
            const int rc = groupby_entry(entry, &fk, 1, vals, num_vals, reqs, num_reqs, &h2, stream, &t);
            if (rc != EG_OK) raise(rc, last_error_cstr());
            h.groupby_algo = h2.groupby_algo;
            h.optimize_sort = h2.optimize_sort;
            // replace the fused key column by the decoded key columns
            std::unique_ptr<eg_table_s> tt(t);
            const int64_t groups = tt->cols.empty() ? 0 : tt->cols[0].size;
            std::vector<DevBuf> kc = unfuse_keys(f, tt->cols[0].data, groups, s);
            EG_HIP(hipStreamSynchronize(s));
            if (tt->cols[0].data) mm_free(tt->cols[0].data);
            std::vector<eg_table_s::Col> cols;
            for (int64_t i = 0; i < num_keys; ++i) cols.push_back(eg_table_s::Col{kc[(size_t)i].release(), groups, kv[(size_t)i].type});
            for (size_t j = 1; j < tt->cols.size(); ++j) cols.push_back(tt->cols[j]);
            tt->cols = std::move(cols);
            *out = tt.release();
            return;
        }

        ViewIn k = view_in(keys[0]);
        std::vector<ViewIn> vs;
        for (int64_t i = 0; i < num_vals; ++i) vs.push_back(view_in(vals[i]));
        const bool top = entry == Entry::TOP;
        // 1- and 2-byte integer, BOOL8 and float keys run on a mapped 4- or
        // 8-byte integer key (keynorm.hip); the group keys are mapped back.
        // Partition takes integral keys only (groupby_partition.cu:556-567).
        const int32_t orig_key_type = k.type;
        const bool mapped_key = key_needs_mapping(orig_key_type);
        if (mapped_key) {
            const bool part = entry == Entry::PARTITION || entry == Entry::PARTITION_GFUR ||
                              (top && (h.groupby_algo == EG_ALGO_PARTITION || h.groupby_algo == EG_ALGO_PARTITION_GFUR));
            if (part && is_float(orig_key_type))
                raise(EG_ERR_UNSUPPORTED, "Non-integral key type is not supported for partition-based groupby.");
            k.type = mapped_key_type(orig_key_type);
        }
        DevBuf mapped;
        eg_column_view mapped_view{};
        const eg_column_view* kviews = keys;
        auto map_key_column = [&]() {
            if (!mapped_key) return;
            if (k.size < 0) raise(EG_ERR_INVALID, "negative key column size");
            if (k.size > 0 && !k.data) raise(EG_ERR_INVALID, "null key column data");
            mapped = map_keys(k.data, orig_key_type, k.size, s);
            k.data = mapped.get();
            mapped_view = eg_column_view{mapped.get(), k.size, k.type, 0, 0};
            kviews = &mapped_view;
        };
        auto finish = [&](Result& res) {
            if (mapped_key && !res.cols.empty()) {
                res.cols[0] = unmap_keys(res.cols[0].get(), res.num_groups, orig_key_type, s);
                res.types[0] = orig_key_type;
            }
            EG_HIP(hipStreamSynchronize(s));
            *out = to_table(res);
        };
        Plan plan;
        try {
            plan = make_plan(&k, vs, reqs, num_reqs, top && h.agg_count != 0, top);
        } catch (const WideError&) {
            map_key_column();
            Result res = wide_groupby(k, vs, reqs, num_reqs, top, h.agg_count != 0, h, s);
            finish(res);
            return;
        }
        map_key_column();
        plan.key.data = k.data;

        int algo;
        int64_t learned_bound = 0;  // group bound from the dispatch model's estimate
        switch (entry) {
        case Entry::HASH: algo = EG_ALGO_HASH; break;
        case Entry::SORT: algo = EG_ALGO_SORT; break;
        case Entry::SORT_GFUR: algo = EG_ALGO_SORT_GFUR; break;
        case Entry::PARTITION: algo = EG_ALGO_PARTITION; break;
        case Entry::PARTITION_GFUR: algo = EG_ALGO_PARTITION_GFUR; break;
        default:
            if (h.groupby_algo == EG_ALGO_AUTO && plan.rows > 0) {
                // learned dispatch when a model is loaded, else the reference's rule
                int vbytes = 0;
                for (int c = 0; c < plan.spec.nvcols; ++c) vbytes += plan.spec.cols[c].bytes;
                int32_t a = 0;
                int64_t card = 0;
                float singletons = 1.f;
                if (mlp_select(plan.key.data, plan.key_type, plan.rows, h.groups_upper_bound, plan.spec.nops, vbytes, s,
                               &a, &card, &singletons)) {
                    h.groupby_algo = a;
                    // The tables are sized from a group-count bound.  A caller's
                    // bound is an upper bound, often loose (skewed keys use few
                    // of the possible values), so a smaller estimate replaces
                    // it: the model's estimate with 2.5x headroom.  On inputs
                    // of 2^24 rows or more whose sample is mostly singletons
                    // (more than 1/4 of its keys seen once) the full-column
                    // HyperLogLog estimate (standard error 0.8%, 25% headroom)
                    // sizes the tables instead: one pass over the keys, cheap
                    // against such an input, and tighter tables.  Below that
                    // the pass and its synchronisation would cost more than
                    // the tighter sizing saves; the model's estimate stays
                    // within 1.34x of the true count on every held-out query
                    // of that kind (2,913), far inside the headroom.
                    const int64_t caller = (h.groups_upper_bound > 0 && h.groups_upper_bound < plan.rows)
                                               ? h.groups_upper_bound : plan.rows;
                    learned_bound = std::min<int64_t>(caller, card + card + card / 2 + 1024);
                    if (learned_bound < caller && singletons > 0.25f && plan.rows >= ((int64_t)1 << 24)) {
                        const int64_t est = ndv_approx(plan.key.data, plan.key_type, plan.rows, s);
                        learned_bound = std::min<int64_t>(caller, est + est / 4 + 1024);
                    }
                } else {
                    int rc = eg_auto_select_algo(kviews, num_keys, &h, stream);
                    if (rc != EG_OK) raise(rc, last_error_cstr());
                }
            }
            algo = h.groupby_algo;
            if (algo == EG_ALGO_AUTO) algo = EG_ALGO_HASH;  // empty input
            break;
        }
        if (algo == EG_ALGO_SORT_OPT) h.optimize_sort = 1;

        HostPhase phase_all("host_groupby_call", s);
        Result res;
        if (plan.rows == 0) {
            OutSpec os;
            alloc_outputs(plan, 0, s, res, os);
        } else {
            switch (algo) {
            case EG_ALGO_HASH: res = run_hash(plan, learned_bound > 0 ? learned_bound : group_bound_of(plan, h, s), s); break;
            case EG_ALGO_SORT: case EG_ALGO_SORT_OPT: res = run_sort(plan, h, false, s); break;
            case EG_ALGO_SORT_GFUR: res = run_sort(plan, h, true, s); break;
            case EG_ALGO_PARTITION:
                res = run_partition(plan, learned_bound > 0 ? learned_bound : group_bound_of(plan, h, s), false, s);
                break;
            case EG_ALGO_PARTITION_GFUR: res = run_partition(plan, group_bound_of(plan, h, s), true, s); break;
            default: raise(EG_ERR_UNSUPPORTED, "Unsupported groupby algorithm");
            }
        }
        finish(res);
    });
}
}  // namespace

extern "C" {

const char* eg_last_error(void) { return eg::last_error_cstr(); }

const char* eg_version(void) { return "eiger-mi355x 0.1 (gfx950)"; }

void eg_hint_init(eg_groupby_hint* h)
{
    if (!h) return;
    h->groups_upper_bound = INT64_MAX;
    h->zero_based_ids = 0;
    h->optimize_sort = 0;
    h->groupby_algo = EG_ALGO_AUTO;
    h->agg_count = 0;
    h->enable_skf = 0;
    h->allow_hash_fallback = 0;
}

int eg_groupby(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
               int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
               eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::TOP, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_hash(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
                    int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
                    eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::HASH, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_sort(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
                    int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
                    eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::SORT, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_sort_gfur(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
                         int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
                         eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::SORT_GFUR, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_partition(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
                         int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
                         eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::PARTITION, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_partition_gfur(const eg_column_view* keys, int64_t num_keys,
                              const eg_column_view* vals, int64_t num_vals,
                              const eg_agg_request* reqs, int64_t num_reqs,
                              eg_groupby_hint* hint, void* stream, eg_table* out)
{
    return groupby_entry(Entry::PARTITION_GFUR, keys, num_keys, vals, num_vals, reqs, num_reqs, hint, stream, out);
}

int eg_groupby_hash_noalloc(const eg_column_view* keys, int64_t num_keys, const eg_column_view* vals,
                            int64_t num_vals, const eg_agg_request* reqs, int64_t num_reqs,
                            int64_t groups_upper_bound, const eg_column_view* out_cols, int64_t num_out_cols,
                            int64_t* num_groups, void* stream)
{
    return guard([&] {
        if (!num_groups) raise(EG_ERR_INVALID, "num_groups is null");
        *num_groups = 0;
        if (num_keys != 1 || !keys)
            raise(num_keys > 1 ? EG_ERR_UNSUPPORTED : EG_ERR_INVALID,
                  "the no-allocation group-by takes exactly one key column");
        if (num_vals < 0 || (num_vals > 0 && !vals)) raise(EG_ERR_INVALID, "bad value columns");
        ViewIn k = view_in(keys[0]);
        std::vector<ViewIn> vs;
        for (int64_t i = 0; i < num_vals; ++i) vs.push_back(view_in(vals[i]));
        Plan plan;
        try {
            plan = make_plan(&k, vs, reqs, num_reqs, false, false);
        } catch (const WideError& e) {
            raise(EG_ERR_UNSUPPORTED, std::string("no-allocation group-by: ") + e.what());
        }
        // caller-owned outputs: [key, one column per request], each typed as the result
        if (!out_cols || num_out_cols != 1 + (int64_t)plan.outs.size())
            raise(EG_ERR_INVALID, "expected " + std::to_string(1 + plan.outs.size()) + " output columns");
        int64_t capacity = INT64_MAX;
        void* outs[kMaxOuts] = {};
        void* key_out = nullptr;
        for (int64_t j = 0; j < num_out_cols; ++j) {
            const eg_column_view& o = out_cols[j];
            const int32_t want = j == 0 ? plan.key_type : plan.outs[(size_t)j - 1].type;
            if (o.type != want) raise(EG_ERR_INVALID, "output column " + std::to_string(j) + " has the wrong type");
            if (o.size < 0 || o.offset < 0) raise(EG_ERR_INVALID, "bad output column view");
            if (o.size > 0 && !o.data) raise(EG_ERR_INVALID, "null output column data");
            capacity = std::min(capacity, o.size);
            void* p = o.data ? (char*)o.data + o.offset * type_size(o.type) : nullptr;
            if (j == 0) key_out = p;
            else outs[j - 1] = p;
        }
        if (plan.rows == 0) return;
        const int64_t bound = groups_upper_bound > 0 ? std::min(groups_upper_bound, plan.rows) : plan.rows;
        OutSpec os;
        set_outspec(plan, key_out, outs, os);
        *num_groups = run_hash_noalloc(plan, bound, os, capacity, as_stream(stream));
    });
}

// Distinct key rows: the group-by with one COUNT_ALL request, count dropped.
static int distinct_entry(bool sort, const eg_column_view* keys, int64_t num_keys, eg_groupby_hint* hint,
                          void* stream, eg_table* out)
{
    if (!out) return eg::guard([] { eg::raise(EG_ERR_INVALID, "out is null"); });
    *out = nullptr;
    eg_groupby_hint local;
    eg_hint_init(&local);
    eg_groupby_hint& h = hint ? *hint : local;
    const eg_agg_request cnt{-1, EG_AGG_COUNT_ALL, 0};
    const bool agg_count = h.agg_count != 0;
    h.agg_count = 0;
    int rc;
    if (sort && num_keys > 1) {
        // multi-column keys only through key fusion (groupby_sort.hpp:39-44)
        if (!h.enable_skf)
            return eg::guard([] { eg::raise(EG_ERR_UNSUPPORTED, "sort::distinct: multi-column keys need enable_skf"); });
        const int32_t algo = h.groupby_algo;
        h.groupby_algo = EG_ALGO_SORT;
        rc = groupby_entry(Entry::TOP, keys, num_keys, nullptr, 0, &cnt, 1, &h, stream, out);
        h.groupby_algo = algo;
    } else {
        rc = groupby_entry(sort ? Entry::SORT : Entry::HASH, keys, num_keys, nullptr, 0, &cnt, 1, &h, stream, out);
    }
    h.agg_count = agg_count ? 1 : 0;
    if (rc != EG_OK) return rc;
    return eg::guard([&] {
        eg_table t = *out;
        if (t->cols.size() != (size_t)num_keys + 1) eg::raise(EG_ERR_DEVICE, "distinct: unexpected result shape");
        if (t->cols.back().data) eg::mm_free(t->cols.back().data);
        t->cols.pop_back();
    });
}

int eg_distinct_hash(const eg_column_view* keys, int64_t num_keys, eg_groupby_hint* hint, void* stream,
                     eg_table* out)
{
    return distinct_entry(false, keys, num_keys, hint, stream, out);
}

int eg_distinct_sort(const eg_column_view* keys, int64_t num_keys, eg_groupby_hint* hint, void* stream,
                     eg_table* out)
{
    return distinct_entry(true, keys, num_keys, hint, stream, out);
}

int eg_aggregate(const eg_column_view* cols, int64_t num_cols, const eg_agg_request* reqs,
                 int64_t num_reqs, void* stream, eg_table* out)
{
    return guard([&] {
        if (!out) raise(EG_ERR_INVALID, "out is null");
        *out = nullptr;
        if (num_cols < 0 || (num_cols > 0 && !cols)) raise(EG_ERR_INVALID, "bad input columns");
        if (num_reqs <= 0 || !reqs) raise(EG_ERR_INVALID, "No aggregation function is specified");
        std::vector<ViewIn> vs;
        for (int64_t i = 0; i < num_cols; ++i) vs.push_back(view_in(cols[i]));
        Result res;
        scalar_aggregate(vs, reqs, num_reqs, as_stream(stream), res);
        EG_HIP(hipStreamSynchronize(as_stream(stream)));
        *out = to_table(res);
    });
}

int eg_agg_output_type(int32_t agg_kind, int32_t type, int32_t* out_type)
{
    return guard([&] {
        if (!out_type) raise(EG_ERR_INVALID, "out_type is null");
        int32_t t = eg::agg_output_type(agg_kind, type);
        if (t == EG_TYPE_EMPTY) raise(EG_ERR_UNSUPPORTED, "Unsupported aggregation function or type");
        *out_type = t;
    });
}

int64_t eg_table_num_columns(eg_table t) { return t ? (int64_t)t->cols.size() : 0; }

int64_t eg_table_num_rows(eg_table t)
{
    // table::num_rows (memory/table.hpp:160-162): rows of the first column.
    return (t && !t->cols.empty()) ? t->cols[0].size : 0;
}

int eg_table_get_column(eg_table t, int64_t i, eg_column_view* out)
{
    return guard([&] {
        if (!t || !out) raise(EG_ERR_INVALID, "null argument");
        if (i < 0 || i >= (int64_t)t->cols.size()) raise(EG_ERR_INVALID, "column out of range");
        const auto& c = t->cols[(size_t)i];
        out->data = c.data;
        out->size = c.size;
        out->type = c.type;
        out->reserved = 0;
        out->offset = 0;
    });
}

int eg_table_release_column(eg_table t, int64_t i, eg_column_view* out)
{
    return guard([&] {
        if (!t || !out) raise(EG_ERR_INVALID, "null argument");
        if (i < 0 || i >= (int64_t)t->cols.size()) raise(EG_ERR_INVALID, "column out of range");
        auto& c = t->cols[(size_t)i];
        out->data = c.data;
        out->size = c.size;
        out->type = c.type;
        out->reserved = 0;
        out->offset = 0;
        c.data = nullptr;
    });
}

void eg_table_free(eg_table t)
{
    if (!t) return;
    for (auto& c : t->cols)
        if (c.data) {
            try {
                eg::mm_free(c.data);
            } catch (...) {
            }
        }
    delete t;
}

int eg_malloc(void** ptr, int64_t bytes, int32_t clear, void* stream)
{
    return guard([&] {
        if (!ptr || bytes < 0) raise(EG_ERR_INVALID, "bad arguments");
        *ptr = eg::mm_alloc((size_t)bytes, as_stream(stream));
        if (clear) {
            eg::fill_zero(*ptr, (size_t)bytes, as_stream(stream));
            EG_HIP(hipStreamSynchronize(as_stream(stream)));
        }
    });
}

int eg_free(void* ptr)
{
    return guard([&] { eg::mm_free(ptr); });
}

int eg_memory_usage(int64_t* current_bytes, int64_t* peak_bytes)
{
    return guard([&] { eg::mm_usage(current_bytes, peak_bytes); });
}

int eg_set_memory_mode(int32_t mode, int64_t pool_bytes)
{
    return guard([&] { eg::mm_set_mode(mode, pool_bytes); });
}

int eg_make_column_from_host(const void* host_data, int64_t num_items, int32_t type, void* stream,
                             eg_column_view* out)
{
    return guard([&] {
        if (!out || num_items < 0 || (num_items > 0 && !host_data)) raise(EG_ERR_INVALID, "bad arguments");
        const int bytes = type_size(type);
        if (!bytes) raise(EG_ERR_UNSUPPORTED, "unsupported column type");
        void* p = eg::mm_alloc((size_t)num_items * bytes, as_stream(stream));
        if (num_items > 0) {
            hipError_t e = hipMemcpyAsync(p, host_data, (size_t)num_items * bytes, hipMemcpyHostToDevice,
                                          as_stream(stream));
            if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));
            if (e != hipSuccess) {
                eg::mm_free(p);
                raise(EG_ERR_DEVICE, std::string("make_column_from_host: ") + hipGetErrorString(e));
            }
        }
        out->data = p;
        out->size = num_items;
        out->type = type;
        out->reserved = 0;
        out->offset = 0;
    });
}

int eg_copy_column_to_host(eg_column_view col, void* host_data, void* stream)
{
    return guard([&] {
        const int bytes = type_size(col.type);
        if (!bytes) raise(EG_ERR_UNSUPPORTED, "unsupported column type");
        if (col.size < 0 || col.offset < 0) raise(EG_ERR_INVALID, "bad column view");
        if (col.size == 0) return;
        if (!host_data || !col.data) raise(EG_ERR_INVALID, "null pointer");
        EG_HIP(hipMemcpyAsync(host_data, (const char*)col.data + col.offset * bytes, (size_t)col.size * bytes,
                              hipMemcpyDeviceToHost, as_stream(stream)));
        EG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int eg_kernel_timing_enable(int32_t on) { return eg::timing_enable(on); }

int eg_kernel_timing_reset(void) { return eg::timing_reset(); }

int eg_kernel_timing_query(const char* name, int64_t* launches, double* total_ms)
{
    return guard([&] {
        if (!name) raise(EG_ERR_INVALID, "name is null");
        eg::timing_query(name, launches, total_ms);
    });
}

int eg_gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count, int64_t n,
                        int64_t distinct, uint64_t seed, void* stream)
{
    return guard([&] {
        eg::gen_keys_uniform(out, type, begin, count, n, distinct, seed, as_stream(stream));
        EG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int eg_zipf_cdf(double* cdf, int64_t distinct, double s)
{
    return guard([&] { eg::zipf_cdf(cdf, distinct, s); });
}

int eg_gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count, const double* d_cdf,
                     int64_t distinct, uint64_t seed, void* stream)
{
    return guard([&] {
        eg::gen_keys_zipf(out, type, begin, count, d_cdf, distinct, seed, as_stream(stream));
        EG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int eg_gen_values(void* out, int32_t type, int32_t dist, int64_t begin, int64_t count,
                  uint64_t seed, uint32_t stream_id, void* stream)
{
    return guard([&] {
        eg::gen_values(out, type, dist, begin, count, seed, stream_id, as_stream(stream));
        EG_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

}  // extern "C"

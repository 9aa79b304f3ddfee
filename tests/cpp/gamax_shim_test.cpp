// gamax_shim_test.cpp -- the reference's own group-by / sort / partition test
// cases, written against the header-only gamax C++ mirror (gamax/gamax.hpp,
// gamax/phys.hpp, gamax/pullops.hpp)
// and run on the GPU through libeiger_groupby.so.  Each case follows a
// reference test (file:line relative to src/cuda/ of the reference repo); the
// expected values are the closed forms and fixtures those tests hold.
//
// Prints "PASS <case>" per case; exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "gamax/gamax.hpp"
#include "gamax/phys.hpp"
#include "gamax/pullops.hpp"

using namespace gamax;
using groupby::aggregation_id;
using groupby::aggregation_request;
using groupby::groupby_algo;
using groupby::groupby_hint;

static int g_failed = 0;

#define CHECK(cond)                                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

static void pass(const std::string& name) { std::printf("PASS %s\n", name.c_str()), std::fflush(stdout); }

static const std::vector<std::pair<const char*, groupby_algo>> kAlgos = {
    {"HASH", groupby::HASH},           {"SORT", groupby::SORT},
    {"SORT_GFUR", groupby::SORT_GFUR}, {"PARTITION", groupby::PARTITION},
    {"PARTITION_GFUR", groupby::PARTITION_GFUR}, {"AUTO", groupby::AUTO_ALGO}};

template <typename T>
static std::unique_ptr<table> one_col_table(const std::vector<T>& h, data_type t)
{
    auto tab = std::make_unique<table>();
    tab->add_column(make_column(h, t));
    return tab;
}

// Key -> value map of a two-column result (order-free comparison, as the
// reference's check_cols_eq_after_sort, tests/test_helper.hpp).
template <typename K, typename V>
static std::map<K, V> as_map(const table& t, int vcol)
{
    auto k = copy_column_to_host<K>(t.get_column(0).view());
    auto v = copy_column_to_host<V>(t.get_column(vcol).view());
    CHECK(k.size() == v.size());
    std::map<K, V> m;
    for (size_t i = 0; i < k.size(); ++i) {
        CHECK(m.count(k[i]) == 0);
        m[k[i]] = v[i];
    }
    return m;
}

// tests/groupby/groupby_test.hpp:64-92: keys [0, G) repeated N/G times and
// shuffled, values all ones, two SUM requests; every SUM is N/G and the keys
// sum to G(G-1)/2.
static void test_uniform_closed_form(groupby_algo algo, const char* name, int64_t n, int64_t g)
{
    std::vector<int32_t> keys((size_t)n);
    for (int64_t i = 0; i < n; ++i) keys[(size_t)i] = (int32_t)(i % g);
    std::mt19937_64 rng(42);
    std::shuffle(keys.begin(), keys.end(), rng);
    std::vector<int32_t> ones((size_t)n, 1);
    auto kt = one_col_table(keys, data_type(type_id::INT32));
    auto vt = std::make_unique<table>();
    vt->add_column(make_column(ones, data_type(type_id::INT32)));
    vt->add_column(make_column(ones, data_type(type_id::INT32)));
    std::vector<aggregation_request> req{{0, aggregation_id::SUM}, {1, aggregation_id::SUM}};
    groupby_hint hint;
    hint.groups_upper_bound = g;
    hint.groupby_algo = algo;
    auto out = groupby::groupby(kt->view(), vt->view(), req, hint);
    CHECK(out->num_columns() == 3);
    CHECK(out->num_rows() == g);
    CHECK(out->get_column(1).type() == data_type(type_id::INT64));
    auto k = copy_column_to_host<int32_t>(out->get_column(0).view());
    auto s0 = copy_column_to_host<int64_t>(out->get_column(1).view());
    auto s1 = copy_column_to_host<int64_t>(out->get_column(2).view());
    unsigned long long ksum = 0;
    for (int64_t i = 0; i < g; ++i) {
        ksum += (unsigned long long)k[(size_t)i];
        CHECK(s0[(size_t)i] == n / g);
        CHECK(s1[(size_t)i] == n / g);
    }
    CHECK(ksum == (unsigned long long)(g - 1) * (unsigned long long)g / 2);
    if (algo == groupby::AUTO_ALGO) CHECK(hint.groupby_algo != groupby::AUTO_ALGO);  // resolved in place
    pass(std::string("uniform_closed_form ") + name + " G=" + std::to_string(g));
}

// tests/types/floating_point_test.cu:19-60: MAX over float / double values.
template <typename T>
static void test_floating_point_max(groupby_algo algo, const char* name, type_id vt)
{
    std::vector<int32_t> hk{1, 2, 2, 3, 3, 3, 4, 4, 4, 4};
    std::vector<T> hv{(T)1.1, (T)2.1, (T)2.2, (T)3.1, (T)3.2, (T)3.3, (T)4.1, (T)4.2, (T)4.3, (T)4.4};
    auto kt = one_col_table(hk, data_type(type_id::INT32));
    auto vtab = one_col_table(hv, data_type(vt));
    std::vector<aggregation_request> req{{0, aggregation_id::MAX}};
    groupby_hint hint;
    hint.groups_upper_bound = 4;
    hint.groupby_algo = algo;
    auto out = groupby::groupby(kt->view(), vtab->view(), req, hint);
    CHECK(out->num_columns() == 2);
    auto m = as_map<int32_t, T>(*out, 1);
    CHECK(m.size() == 4);
    CHECK(m[1] == (T)1.1 && m[2] == (T)2.2 && m[3] == (T)3.3 && m[4] == (T)4.4);
    pass(std::string("floating_point_max ") + name + (vt == type_id::FLOAT32 ? " f32" : " f64"));
}

// phys/phys_test.cu:533-548: keys {1,1,2,2,3}, values {10..50}: SUM {30,70,50}.
static void test_phys_sum(groupby_algo algo, const char* name)
{
    std::vector<int64_t> hk{1, 1, 2, 2, 3};
    std::vector<int64_t> hv{10, 20, 30, 40, 50};
    auto kt = one_col_table(hk, data_type(type_id::INT64));
    auto vtab = one_col_table(hv, data_type(type_id::INT64));
    std::vector<aggregation_request> req{{0, aggregation_id::SUM}};
    groupby_hint hint;
    hint.groupby_algo = algo;
    auto out = groupby::groupby(kt->view(), vtab->view(), req, hint);
    auto m = as_map<int64_t, int64_t>(*out, 1);
    CHECK(m.size() == 3 && m[1] == 30 && m[2] == 70 && m[3] == 50);
    pass(std::string("phys_sum ") + name);
}

// groupby.cu:29-51,326-388: AVG = double(SUM)/double(COUNT); agg_count adds a
// trailing int64 count column.
static void test_avg_and_agg_count()
{
    std::vector<int32_t> hk{5, 5, 5, 7, 7, 9};
    std::vector<int32_t> hv{1, 2, 4, 10, 11, -3};
    auto kt = one_col_table(hk, data_type(type_id::INT32));
    auto vtab = one_col_table(hv, data_type(type_id::INT32));
    std::vector<aggregation_request> req{{0, aggregation_id::AVG}, {-1, aggregation_id::COUNT_ALL}};
    groupby_hint hint;
    hint.agg_count = true;
    auto out = groupby::groupby(kt->view(), vtab->view(), req, hint);
    CHECK(out->num_columns() == 4);
    CHECK(out->get_column(1).type() == data_type(type_id::FLOAT64));
    auto avg = as_map<int32_t, double>(*out, 1);
    auto cnt = as_map<int32_t, int64_t>(*out, 2);
    auto cnt2 = as_map<int32_t, int64_t>(*out, 3);
    CHECK(std::fabs(avg[5] - 7.0 / 3.0) < 1e-12 && avg[7] == 10.5 && avg[9] == -3.0);
    CHECK(cnt[5] == 3 && cnt[7] == 2 && cnt[9] == 1 && cnt2 == cnt);
    pass("avg_and_agg_count");
}

// aggregate/aggregate_test.cu:11-111: scalar SUM over 0..n-1.
static void test_scalar_aggregate()
{
    const int64_t n = 100000;
    std::vector<int64_t> h((size_t)n);
    std::iota(h.begin(), h.end(), 0);
    auto t = one_col_table(h, data_type(type_id::INT64));
    std::vector<aggregation_request> req{{0, aggregation_id::SUM}, {0, aggregation_id::MAX},
                                         {-1, aggregation_id::COUNT_ALL}};
    auto out = groupby::aggregate::aggregate(t->view(), req);
    CHECK(out->num_rows() == 1 && out->num_columns() == 3);
    CHECK(copy_column_to_host<int64_t>(out->get_column(0).view())[0] == n * (n - 1) / 2);
    CHECK(copy_column_to_host<int64_t>(out->get_column(1).view())[0] == n - 1);
    CHECK(copy_column_to_host<int64_t>(out->get_column(2).view())[0] == n);
    pass("scalar_aggregate");
}

// tests/sort/sort_partition_test.cu:16-59: partitioned pairs are a stable
// permutation of the input grouped by the bits; offsets are the inclusive
// prefix of the bucket counts.
static void test_partition_and_sort()
{
    const int64_t n = 200000;
    std::mt19937 rng(7);
    std::vector<int32_t> hk((size_t)n);
    std::vector<int64_t> hv((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        hk[(size_t)i] = (int32_t)(rng() & 0xFFFF);
        hv[(size_t)i] = i;
    }
    auto kc = make_column(hk, data_type(type_id::INT32));
    auto vc = make_column(hv, data_type(type_id::INT64));
    auto [pk, pv] = partition::partition_column_pair(kc->view(), vc->view(), 0, 8);
    auto k = copy_column_to_host<int32_t>(pk->view());
    auto v = copy_column_to_host<int64_t>(pv->view());
    for (int64_t i = 0; i < n; ++i) CHECK(hk[(size_t)v[(size_t)i]] == k[(size_t)i]);
    for (int64_t i = 1; i < n; ++i) {
        const int a = k[(size_t)i - 1] & 0xFF, b = k[(size_t)i] & 0xFF;
        CHECK(a < b || (a == b && v[(size_t)i - 1] < v[(size_t)i]));
    }
    auto off = copy_column_to_host<int64_t>(partition::compute_partition_offsets(pk->view(), 0, 8)->view());
    CHECK(off.size() == 257 && off[255] == n && off[256] == n);
    std::vector<int64_t> cnt(256, 0);
    for (auto x : hk) cnt[(size_t)(x & 0xFF)]++;
    int64_t acc = 0;
    for (int b = 0; b < 256; ++b) CHECK(off[(size_t)b] == (acc += cnt[(size_t)b]));

    auto sorted = copy_column_to_host<int32_t>(sort::sort_column(kc->view(), 0, -1, sort_order::DESCENDING)->view());
    std::vector<int32_t> exp = hk;
    std::sort(exp.begin(), exp.end(), std::greater<int32_t>());
    CHECK(sorted == exp);
    auto [sk, idx] = sort::argsort(kc->view());
    auto ix = copy_column_to_host<int64_t>(idx->view());
    auto sv = copy_column_to_host<int32_t>(sk->view());
    for (int64_t i = 0; i < n; ++i) CHECK(hk[(size_t)ix[(size_t)i]] == sv[(size_t)i]);
    CHECK(std::is_sorted(sv.begin(), sv.end()));
    pass("partition_and_sort");
}

// compress/dictionary_encode_dict_ready_test.cu:59-92 through the C++ mirror:
// dict = 0..D-1, keys (0..n-1) mod D shuffled, so each key's position is the
// key; an absent key gives INT32_MIN; the self-built ordered ids of the same
// keys are the keys again (their distinct values are exactly 0..D-1).
static void test_dictionary_encode()
{
    const int32_t n = 4096, d = 16;
    std::vector<int32_t> hk((size_t)n), hd((size_t)d);
    for (int32_t i = 0; i < n; ++i) hk[(size_t)i] = i % d;
    std::shuffle(hk.begin(), hk.end(), std::mt19937(3));
    for (int32_t i = 0; i < d; ++i) hd[(size_t)i] = i;
    hk[5] = 1000;  // absent from the dictionary
    auto kc = make_column(hk, data_type(type_id::INT32));
    auto dc = make_column(hd, data_type(type_id::INT32));
    auto e = copy_column_to_host<int32_t>(compress::dictionary_encode(kc->view(), dc->view())->view());
    for (int32_t i = 0; i < n; ++i) CHECK(e[(size_t)i] == (i == 5 ? INT32_MIN : hk[(size_t)i]));
    hk[5] = 0;
    auto kc2 = make_column(hk, data_type(type_id::INT32));
    auto [ids, size] = compress::dictionary_encode_with_order(kc2->view(), data_type(type_id::INT64));
    CHECK(size == d);
    auto hid = copy_column_to_host<int64_t>(ids->view());
    for (int32_t i = 0; i < n; ++i) CHECK(hid[(size_t)i] == hk[(size_t)i]);
    auto [ids2, size2] = compress::dictionary_encode(kc2->view());
    CHECK(size2 == d && ids2->size() == n);
    pass("dictionary_encode");
}

// phys/phys_test.cu:455-476,520-548: the engine-facing wrappers.  Empty
// inputs: group_by gives 0 rows with schema [key, SUM output]; the scalar
// aggregate gives one row of zeros.  Keys {1,1,2,2,3}, values {10..50}:
// three groups, SUM {30,70,50}; values {10,20,30}: SUM 60 (SUM(int32) is
// INT64, aggregate.cu:20-53).  The caller's hint is not resolved in place.
static void test_phys_wrappers()
{
    const phys::Context ctx{};
    auto ek = one_col_table(std::vector<int32_t>{}, data_type(type_id::INT32));
    auto ev = one_col_table(std::vector<double>{}, data_type(type_id::FLOAT64));
    auto g0 = phys::group_by(ek->view(), ev->view(), {{0, aggregation_id::SUM}}, ctx);
    CHECK(g0->num_rows() == 0 && g0->num_columns() == 2);
    CHECK(g0->get_column(0).type() == data_type(type_id::INT32));
    CHECK(g0->get_column(1).type() == data_type(type_id::FLOAT64));
    auto s0 = phys::scalar_aggregate(ek->view(), {{0, aggregation_id::SUM}, {-1, aggregation_id::COUNT_ALL}}, ctx);
    CHECK(s0->num_rows() == 1 && s0->num_columns() == 2);
    CHECK(s0->get_column(0).type() == data_type(type_id::INT64));
    CHECK(copy_column_to_host<int64_t>(s0->get_column(0).view())[0] == 0);
    CHECK(copy_column_to_host<int64_t>(s0->get_column(1).view())[0] == 0);

    auto kt = one_col_table(std::vector<int32_t>{1, 1, 2, 2, 3}, data_type(type_id::INT32));
    auto vt = one_col_table(std::vector<int32_t>{10, 20, 30, 40, 50}, data_type(type_id::INT32));
    groupby_hint hint;  // AUTO
    auto g = phys::group_by(kt->view(), vt->view(), {{0, aggregation_id::SUM}}, ctx, hint);
    CHECK(g->num_rows() == 3 && g->num_columns() == 2);
    auto m = as_map<int32_t, int64_t>(*g, 1);
    CHECK(m.size() == 3 && m[1] == 30 && m[2] == 70 && m[3] == 50);
    CHECK(hint.groupby_algo == groupby::AUTO_ALGO);

    auto st = one_col_table(std::vector<int32_t>{10, 20, 30}, data_type(type_id::INT32));
    auto s = phys::scalar_aggregate(st->view(), {{0, aggregation_id::SUM}}, ctx);
    CHECK(s->num_rows() == 1 && copy_column_to_host<int64_t>(s->get_column(0).view())[0] == 60);

    // phys::distinct (phys.cu:622-663): keys {1,1,2,2,3} -> {1,2,3} by hash and
    // by sort (ascending); PARTITION throws; empty input keeps the schema.
    groupby_hint hh;
    auto dh = phys::distinct(kt->view(), ctx, hh);
    auto vh = copy_column_to_host<int32_t>(dh->get_column(0).view());
    std::sort(vh.begin(), vh.end());
    CHECK(dh->num_columns() == 1 && vh == (std::vector<int32_t>{1, 2, 3}));
    groupby_hint hs;
    hs.groupby_algo = groupby::SORT;
    auto ds = phys::distinct(kt->view(), ctx, hs);
    CHECK(copy_column_to_host<int32_t>(ds->get_column(0).view()) == (std::vector<int32_t>{1, 2, 3}));
    groupby_hint hp;
    hp.groupby_algo = groupby::PARTITION;
    bool threw = false;
    try {
        phys::distinct(kt->view(), ctx, hp);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
    auto de = phys::distinct(ek->view(), ctx);
    CHECK(de->num_rows() == 0 && de->num_columns() == 1 && de->get_column(0).type() == data_type(type_id::INT32));
    pass("phys_wrappers");
}

// Output directory for arrays the Python side checks against the oracle
// (argv[1]; none: nothing is written).
static const char* g_dump_dir = nullptr;

template <class T>
static void dump(const char* name, const std::vector<T>& v)
{
    if (!g_dump_dir) return;
    const std::string path = std::string(g_dump_dir) + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    CHECK(f != nullptr);
    if (!f) return;
    CHECK(std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size());
    std::fclose(f);
}

// The pull-based GroupBy operator (designed here; the reference snapshot has
// no GroupBy pullop): a scan in 100000-row chunks under a GroupByOperator
// emitting 1000-row chunks gives the same groups as one groupby call; an
// empty input gives no chunk.  The input and the emitted rows are also
// written out (g_dump_dir) and compared with the oracle by the Python test.
static void test_pullops_group_by()
{
    const int64_t n = 1000000;
    std::mt19937_64 rng(5);
    std::vector<int32_t> hk((size_t)n);
    std::vector<int64_t> hv((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        hk[(size_t)i] = (int32_t)(rng() % 5000) - 2500;
        hv[(size_t)i] = (int64_t)(rng() % 2000001) - 1000000;
    }
    std::vector<std::unique_ptr<column>> cols;
    cols.push_back(make_column(hk, data_type(type_id::INT32)));
    cols.push_back(make_column(hv, data_type(type_id::INT64)));
    table input(std::move(cols));
    std::vector<aggregation_request> reqs{{0, aggregation_id::SUM}, {-1, aggregation_id::COUNT_ALL},
                                          {0, aggregation_id::MAX}};

    pullops::TableScan scan(input.view(), 100000);
    pullops::GroupByOperator gb(&scan, 1, reqs, groupby_hint{}, 1000);
    gb.open();
    std::map<int32_t, std::tuple<int64_t, int64_t, int64_t>> got;
    std::vector<int32_t> ok;
    std::vector<int64_t> os, oc, om;
    int chunks = 0;
    for (auto c = gb.next(); !c.is_end(); c = gb.next()) {
        CHECK(c.num_items <= 1000 && c.data->num_columns() == 4);
        auto k = copy_column_to_host<int32_t>(c.data->get_column(0).view());
        auto s = copy_column_to_host<int64_t>(c.data->get_column(1).view());
        auto n1 = copy_column_to_host<int64_t>(c.data->get_column(2).view());
        auto mx = copy_column_to_host<int64_t>(c.data->get_column(3).view());
        for (size_t i = 0; i < k.size(); ++i) {
            CHECK(got.count(k[i]) == 0);
            got[k[i]] = std::make_tuple(s[i], n1[i], mx[i]);
        }
        ok.insert(ok.end(), k.begin(), k.end());
        os.insert(os.end(), s.begin(), s.end());
        oc.insert(oc.end(), n1.begin(), n1.end());
        om.insert(om.end(), mx.begin(), mx.end());
        c.free_mem();
        ++chunks;
    }
    gb.close();
    dump("pullops_in_keys", hk);
    dump("pullops_in_vals", hv);
    dump("pullops_out_keys", ok);
    dump("pullops_out_sum", os);
    dump("pullops_out_count", oc);
    dump("pullops_out_max", om);
    CHECK(gb.hint_used().groupby_algo != groupby::AUTO_ALGO);

    std::vector<size_type> k0{0}, v0{1};
    groupby_hint h;
    auto direct = groupby::groupby(input.view().select_columns(k0), input.view().select_columns(v0), reqs, h);
    auto ds = as_map<int32_t, int64_t>(*direct, 1);
    auto dc = as_map<int32_t, int64_t>(*direct, 2);
    auto dm = as_map<int32_t, int64_t>(*direct, 3);
    CHECK(got.size() == ds.size() && chunks == (int)((ds.size() + 999) / 1000));
    for (auto const& [k, t] : got)
        CHECK(std::get<0>(t) == ds[k] && std::get<1>(t) == dc[k] && std::get<2>(t) == dm[k]);

    std::vector<std::unique_ptr<column>> ecols;
    ecols.push_back(make_column(std::vector<int32_t>{}, data_type(type_id::INT32)));
    ecols.push_back(make_column(std::vector<int64_t>{}, data_type(type_id::INT64)));
    table empty(std::move(ecols));
    pullops::TableScan escan(empty.view(), 100000);
    pullops::GroupByOperator egb(&escan, 1, reqs);
    egb.open();
    CHECK(egb.next().is_end());
    egb.close();
    pass("pullops_group_by");
}

// groupby_hash.cu:32-40 / groupby.cu:322-323: bad calls throw std::runtime_error.
static void test_errors()
{
    std::vector<int32_t> hk{1, 2, 3};
    auto kt = one_col_table(hk, data_type(type_id::INT32));
    std::vector<aggregation_request> none;
    groupby_hint hint;
    bool threw = false;
    try {
        groupby::groupby(kt->view(), kt->view(), none, hint);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    std::vector<aggregation_request> bad_col{{3, aggregation_id::SUM}};
    try {
        groupby::groupby(kt->view(), kt->view(), bad_col, hint);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
    pass("errors");
}

int main(int argc, char** argv)
{
    if (argc > 1) g_dump_dir = argv[1];
    for (auto [name, algo] : kAlgos) {
        test_uniform_closed_form(algo, name, 1 << 20, 1024);
        test_uniform_closed_form(algo, name, 1 << 20, 4);
        test_floating_point_max<float>(algo, name, type_id::FLOAT32);
        test_floating_point_max<double>(algo, name, type_id::FLOAT64);
        test_phys_sum(algo, name);
    }
    test_avg_and_agg_count();
    test_scalar_aggregate();
    test_partition_and_sort();
    test_phys_wrappers();
    test_pullops_group_by();
    test_dictionary_encode();
    test_errors();
    std::printf("ALL PASSED\n");
    return g_failed;
}

// gamax.hpp -- header-only C++ mirror of the reference's gamax operator API
// over the C ABI of libeiger_groupby.so (include/eiger_groupby.h).
//
// A caller of gamax::groupby::groupby & co. (operators/groupby/groupby.hpp:44-49)
// switches by including this header instead of the reference's headers and
// linking libeiger_groupby.so.  Names, argument order, defaults and error
// behaviour follow the reference:
//   types        types/data_type.hpp:12-58 (type_id, data_type, size_type, sort_order)
//   columns      memory/column.hpp:13-172 (column_view, column: owning, move-only)
//   tables       memory/table.hpp:78-222 (table_view, table)
//   group-by     operators/groupby/{groupby,aggregation,groupby_hash,groupby_sort,
//                groupby_partition}.hpp
//   sort         operators/sort/sort.hpp:14-75
//   partition    operators/partition/partition.hpp:11-79
// Errors: EG_ERR_OOM becomes std::bad_alloc, every other failure
// std::runtime_error carrying the library's message (the reference throws
// std::runtime_error, groupby_hash.cu:32-40, groupby.cu:287-324).
// The stream type is hipStream_t where the reference takes cudaStream_t.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "eiger_groupby.h"

namespace gamax {

using size_type = int64_t;
constexpr size_type MAXIMUM_SIZE = std::numeric_limits<size_type>::max();

enum class type_id {
    EMPTY, INT8, INT16, INT32, INT64, BOOL8, UINT8, UINT16, UINT32, UINT64, FLOAT32, FLOAT64, STRING,
    NUM_TYPE_IDS
};

class data_type {
public:
    data_type() = default;
    explicit constexpr data_type(type_id id) : _id{id} {}
    constexpr type_id id() const noexcept { return _id; }

private:
    type_id _id{type_id::EMPTY};
};
constexpr bool operator==(data_type const& a, data_type const& b) { return a.id() == b.id(); }
inline bool operator!=(data_type const& a, data_type const& b) { return !(a == b); }

enum class sort_order { ASCENDING, DESCENDING };

inline size_type size_of(data_type t)
{
    switch (t.id()) {
    case type_id::INT8: case type_id::UINT8: case type_id::BOOL8: return 1;
    case type_id::INT16: case type_id::UINT16: return 2;
    case type_id::INT32: case type_id::UINT32: case type_id::FLOAT32: return 4;
    case type_id::INT64: case type_id::UINT64: case type_id::FLOAT64: return 8;
    default: throw std::runtime_error("Invalid type id");
    }
}

namespace detail {
inline void check(int rc)
{
    if (rc == EG_OK) return;
    if (rc == EG_ERR_OOM) throw std::bad_alloc();
    throw std::runtime_error(eg_last_error());
}
}  // namespace detail

class column_view {
public:
    column_view() = default;
    column_view(void* data, size_type size, data_type type, size_type offset = 0)
        : _data(data), _size(size - offset), _type(type), _offset(offset)
    {
        if (offset > size) throw std::runtime_error("Offset is greater than the size of the column");
    }
    size_type size() const noexcept { return _size; }
    data_type type() const noexcept { return _type; }
    template <typename T>
    const T* data() const noexcept { return static_cast<const T*>(_data) + _offset; }
    const void* data() const noexcept
    {
        return _data ? static_cast<const char*>(_data) + _offset * size_of(_type) : nullptr;
    }
    eg_column_view abi() const noexcept
    {
        return eg_column_view{_data, _size, static_cast<int32_t>(_type.id()), 0, _offset};
    }

private:
    void* _data{nullptr};
    size_type _size{0};
    data_type _type;
    size_type _offset{0};
};

// Owning device column (memory/column.hpp:96-172); storage comes from the
// library's allocator (eg_malloc), as the reference's comes from its
// MemoryManager.
class column {
public:
    column() = delete;
    column(column const&) = delete;
    column& operator=(column const&) = delete;
    column& operator=(column&&) = delete;
    column(column&& o) noexcept : _data(std::exchange(o._data, nullptr)), _size(o._size), _type(o._type) {}
    explicit column(data_type type, size_type num_rows, bool clear = false, hipStream_t stream = nullptr)
        : _size(num_rows), _type(type)
    {
        if (num_rows > 0) detail::check(eg_malloc(&_data, num_rows * size_of(type), clear ? 1 : 0, stream));
    }
    // Adopts a device allocation made by the library (eg_malloc).
    column(void* adopted, size_type size, data_type type) : _data(adopted), _size(size), _type(type) {}
    ~column()
    {
        if (_data) eg_free(_data);
    }
    size_type size() const noexcept { return _size; }
    data_type type() const noexcept { return _type; }
    void set_size(size_type size) noexcept { _size = size; }
    template <typename T = void>
    T* data() const noexcept { return static_cast<T*>(_data); }
    column_view view(size_type offset = 0) const { return column_view(_data, _size, _type, offset); }
    operator column_view() const { return view(); }

private:
    void* _data{nullptr};
    size_type _size{0};
    data_type _type;
};
using column_ptr = std::unique_ptr<column>;

// Host array -> device column (memory/column.hpp:174-197 make_column).
template <typename T>
std::unique_ptr<column> make_column(const std::vector<T>& host, data_type type, hipStream_t stream = nullptr)
{
    if ((size_type)sizeof(T) != size_of(type)) throw std::runtime_error("host element size differs from the type");
    eg_column_view v{};
    detail::check(eg_make_column_from_host(host.data(), (int64_t)host.size(), (int32_t)type.id(), stream, &v));
    return std::make_unique<column>(const_cast<void*>(v.data), v.size, type);
}

// Device column -> host vector (memory/column.hpp:199-202 copy_column_to_host).
template <typename T>
std::vector<T> copy_column_to_host(column_view col, hipStream_t stream = nullptr)
{
    if ((size_type)sizeof(T) != size_of(col.type())) throw std::runtime_error("host element size differs from the type");
    std::vector<T> out((size_t)col.size());
    if (col.size() > 0) detail::check(eg_copy_column_to_host(col.abi(), out.data(), stream));
    return out;
}

class table_view {
public:
    table_view() = default;
    table_view(std::vector<column_view> const& cols) : _columns(cols) {}
    table_view(std::vector<std::unique_ptr<column>> const& cols)
    {
        for (auto const& c : cols) _columns.push_back(c->view());
    }
    size_type num_columns() const noexcept { return (size_type)_columns.size(); }
    size_type num_rows() const noexcept { return _columns.empty() ? 0 : _columns[0].size(); }
    column_view get_column(size_type i) const { return _columns.at((size_t)i); }
    column_view operator[](size_type i) const { return get_column(i); }
    table_view select_columns(std::vector<size_type> const& idx) const
    {
        std::vector<column_view> s;
        for (auto i : idx) s.push_back(_columns.at((size_t)i));
        return table_view(s);
    }
    void add_column(column_view c)
    {
        if (c.size() != num_rows() && !_columns.empty())
            throw std::runtime_error("New column must have the same number of rows as the table");
        _columns.push_back(c);
    }
    std::vector<eg_column_view> abi() const
    {
        std::vector<eg_column_view> v;
        for (auto const& c : _columns) v.push_back(c.abi());
        return v;
    }

private:
    std::vector<column_view> _columns;
};

class table {
public:
    table() = default;
    table(table&&) = default;
    table& operator=(table&&) = default;
    table(std::vector<std::unique_ptr<column>>&& columns) : _columns(std::move(columns)) {}
    size_type num_columns() const noexcept { return (size_type)_columns.size(); }
    size_type num_rows() const noexcept { return _columns.empty() || !_columns[0] ? 0 : _columns[0]->size(); }
    column& get_column(size_type i) const { return *_columns.at((size_t)i); }
    table_view view() const { return table_view(_columns); }
    void add_column(std::unique_ptr<column>&& c)
    {
        if (num_rows() != 0 && c->size() != num_rows())
            throw std::runtime_error("Column has different number of rows than the table");
        _columns.push_back(std::move(c));
    }
    std::unique_ptr<column> remove_column(size_type i)
    {
        if (i < 0 || i >= num_columns()) throw std::invalid_argument("column out of range");
        auto c = std::move(_columns[(size_t)i]);
        _columns.erase(_columns.begin() + i);
        return c;
    }
    std::unique_ptr<column> release_column(size_type i)
    {
        if (i < 0 || i >= num_columns()) throw std::invalid_argument("column out of range");
        return std::move(_columns[(size_t)i]);
    }

private:
    std::vector<std::unique_ptr<column>> _columns;
};

namespace detail {
// Moves every column of a library result table into a gamax::table.
inline std::unique_ptr<table> adopt(eg_table t)
{
    std::vector<std::unique_ptr<column>> cols;
    const int64_t nc = eg_table_num_columns(t);
    for (int64_t i = 0; i < nc; ++i) {
        eg_column_view v{};
        const int rc = eg_table_release_column(t, i, &v);
        if (rc != EG_OK) {
            eg_table_free(t);
            check(rc);
        }
        cols.push_back(std::make_unique<column>(const_cast<void*>(v.data), v.size, data_type((type_id)v.type)));
    }
    eg_table_free(t);
    return std::make_unique<table>(std::move(cols));
}
}  // namespace detail

namespace groupby {

enum groupby_algo { AUTO_ALGO, HASH, SORT, SORT_OPT, SORT_GFUR, PARTITION, PARTITION_GFUR, UNSUPPORTED_GROUPBY_ALGO };

struct groupby_hint {
    size_type groups_upper_bound{MAXIMUM_SIZE};
    bool zero_based_ids{false};
    bool optimize_sort{false};
    gamax::groupby::groupby_algo groupby_algo{AUTO_ALGO};
    bool agg_count{false};
    bool enable_skf{false};
    bool allow_hash_fallback{false};
};

enum class aggregation_id { SUM, MAX, MIN, AVG, ANY, COUNT_ALL, NOT_SUPPORTED_AGGREGATION };

struct aggregation_request {
    size_type column_index;
    aggregation_id agg_kind;
    aggregation_request(size_type column_index, aggregation_id agg_kind)
        : column_index(column_index), agg_kind(agg_kind) {}
};

inline data_type agg_output_data_type(aggregation_id const& agg_kind, data_type const& type)
{
    int32_t out = 0;
    detail::check(eg_agg_output_type((int32_t)agg_kind, (int32_t)type.id(), &out));
    return data_type((type_id)out);
}

namespace detail {
using entry_fn = int (*)(const eg_column_view*, int64_t, const eg_column_view*, int64_t, const eg_agg_request*,
                         int64_t, eg_groupby_hint*, void*, eg_table*);

inline eg_groupby_hint to_abi(groupby_hint const& h)
{
    eg_groupby_hint a{};
    eg_hint_init(&a);
    a.groups_upper_bound = h.groups_upper_bound;
    a.zero_based_ids = h.zero_based_ids;
    a.optimize_sort = h.optimize_sort;
    a.groupby_algo = (int32_t)h.groupby_algo;
    a.agg_count = h.agg_count;
    a.enable_skf = h.enable_skf;
    a.allow_hash_fallback = h.allow_hash_fallback;
    return a;
}

// The hint is in/out (groupby.cu:73-118): AUTO is resolved in place.
inline std::unique_ptr<table> call(entry_fn fn, table_view key, table_view val,
                                   std::vector<aggregation_request>& reqs, groupby_hint& hint, hipStream_t stream)
{
    auto k = key.abi();
    auto v = val.abi();
    std::vector<eg_agg_request> r;
    for (auto const& q : reqs) r.push_back(eg_agg_request{q.column_index, (int32_t)q.agg_kind, 0});
    eg_groupby_hint h = to_abi(hint);
    eg_table out = nullptr;
    gamax::detail::check(fn(k.data(), (int64_t)k.size(), v.data(), (int64_t)v.size(), r.data(), (int64_t)r.size(), &h,
                            stream, &out));
    hint.groupby_algo = (gamax::groupby::groupby_algo)h.groupby_algo;
    hint.optimize_sort = h.optimize_sort != 0;
    hint.groups_upper_bound = h.groups_upper_bound;
    return gamax::detail::adopt(out);
}

using distinct_fn = int (*)(const eg_column_view*, int64_t, eg_groupby_hint*, void*, eg_table*);

inline std::unique_ptr<table> call_distinct(distinct_fn fn, table_view key, groupby_hint& hint, hipStream_t stream)
{
    auto k = key.abi();
    eg_groupby_hint h = to_abi(hint);
    eg_table out = nullptr;
    gamax::detail::check(fn(k.data(), (int64_t)k.size(), &h, stream, &out));
    hint.groupby_algo = (gamax::groupby::groupby_algo)h.groupby_algo;
    hint.groups_upper_bound = h.groups_upper_bound;
    return gamax::detail::adopt(out);
}
}  // namespace detail

// operators/groupby/groupby.hpp:44-49
inline std::unique_ptr<table> groupby(table_view key, table_view val, std::vector<aggregation_request>& agg_requests,
                                      groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby, key, val, agg_requests, hint, stream);
}

namespace hash {
// operators/groupby/groupby_hash.hpp:17-22
inline std::unique_ptr<table> groupby(table_view key, table_view val, std::vector<aggregation_request>& agg_requests,
                                      groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby_hash, key, val, agg_requests, hint, stream);
}
// operators/groupby/groupby_hash.hpp:24-32
inline std::unique_ptr<table> distinct(table_view key, groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call_distinct(eg_distinct_hash, key, hint, stream);
}
}  // namespace hash

namespace sort {
// operators/groupby/groupby_sort.hpp:17-29
inline std::unique_ptr<table> groupby(table_view key, table_view val, std::vector<aggregation_request>& agg_requests,
                                      groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby_sort, key, val, agg_requests, hint, stream);
}
inline std::unique_ptr<table> groupby_gfur(table_view key, table_view val,
                                           std::vector<aggregation_request>& agg_requests, groupby_hint& hint,
                                           hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby_sort_gfur, key, val, agg_requests, hint, stream);
}
// operators/groupby/groupby_sort.hpp:31-48
inline std::unique_ptr<table> distinct(table_view key, groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call_distinct(eg_distinct_sort, key, hint, stream);
}
}  // namespace sort

namespace partition {
// operators/groupby/groupby_partition.hpp:52-64
inline std::unique_ptr<table> groupby(table_view key, table_view val, std::vector<aggregation_request>& agg_requests,
                                      groupby_hint& hint, hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby_partition, key, val, agg_requests, hint, stream);
}
inline std::unique_ptr<table> groupby_gfur(table_view key, table_view val,
                                           std::vector<aggregation_request>& agg_requests, groupby_hint& hint,
                                           hipStream_t stream = nullptr)
{
    return detail::call(eg_groupby_partition_gfur, key, val, agg_requests, hint, stream);
}
}  // namespace partition

namespace aggregate {
// operators/groupby/aggregation.hpp:34-39
inline std::unique_ptr<table> aggregate(table_view input, std::vector<aggregation_request>& agg_requests,
                                        hipStream_t stream = nullptr)
{
    auto c = input.abi();
    std::vector<eg_agg_request> r;
    for (auto const& q : agg_requests) r.push_back(eg_agg_request{q.column_index, (int32_t)q.agg_kind, 0});
    eg_table out = nullptr;
    gamax::detail::check(eg_aggregate(c.data(), (int64_t)c.size(), r.data(), (int64_t)r.size(), stream, &out));
    return gamax::detail::adopt(out);
}
}  // namespace aggregate
}  // namespace groupby

namespace detail {
inline std::unique_ptr<column> own(eg_column_view v) { return std::make_unique<column>(const_cast<void*>(v.data), v.size, data_type((type_id)v.type)); }
}  // namespace detail

namespace sort {
// operators/sort/sort.hpp:14-75
inline std::pair<std::unique_ptr<column>, std::unique_ptr<column>> sort_column_pair(
    column_view key_col, column_view val_col, int begin_bit = 0, int end_bit = -1,
    sort_order order = sort_order::ASCENDING, void* d_temp_storage = nullptr, size_t temp_storage_bytes = 0,
    hipStream_t stream = nullptr)
{
    eg_column_view k{}, v{};
    detail::check(eg_sort_column_pair(key_col.abi(), val_col.abi(), begin_bit, end_bit, (int32_t)order, d_temp_storage,
                                      temp_storage_bytes, stream, &k, &v));
    return {detail::own(k), detail::own(v)};
}
inline std::unique_ptr<column> sort_column(column_view col, int begin_bit = 0, int end_bit = -1,
                                           sort_order order = sort_order::ASCENDING, void* d_temp_storage = nullptr,
                                           size_t temp_storage_bytes = 0, hipStream_t stream = nullptr)
{
    eg_column_view k{};
    detail::check(eg_sort_column(col.abi(), begin_bit, end_bit, (int32_t)order, d_temp_storage, temp_storage_bytes,
                                 stream, &k));
    return detail::own(k);
}
inline std::pair<std::unique_ptr<column>, std::unique_ptr<column>> argsort(
    column_view col, int begin_bit = 0, int end_bit = -1, sort_order order = sort_order::ASCENDING,
    void* d_temp_storage = nullptr, size_t temp_storage_bytes = 0, hipStream_t stream = nullptr)
{
    eg_column_view k{}, i{};
    detail::check(eg_argsort(col.abi(), begin_bit, end_bit, (int32_t)order, d_temp_storage, temp_storage_bytes, stream,
                             &k, &i));
    return {detail::own(k), detail::own(i)};
}
}  // namespace sort

namespace partition {
// operators/partition/partition.hpp:11-79
inline std::pair<std::unique_ptr<column>, std::unique_ptr<column>> partition_column_pair(
    column_view key_col, column_view val_col, int begin_bit, int end_bit, void* d_temp_storage = nullptr,
    size_t temp_storage_bytes = 0, hipStream_t stream = nullptr)
{
    eg_column_view k{}, v{};
    detail::check(eg_partition_column_pair(key_col.abi(), val_col.abi(), begin_bit, end_bit, d_temp_storage,
                                           temp_storage_bytes, stream, &k, &v));
    return {detail::own(k), detail::own(v)};
}
inline std::unique_ptr<column> partition_column(column_view col, int begin_bit, int end_bit,
                                                void* d_temp_storage = nullptr, size_t temp_storage_bytes = 0,
                                                hipStream_t stream = nullptr)
{
    eg_column_view k{};
    detail::check(eg_partition_column(col.abi(), begin_bit, end_bit, d_temp_storage, temp_storage_bytes, stream, &k));
    return detail::own(k);
}
inline std::pair<std::unique_ptr<column>, std::unique_ptr<column>> arg_partition(
    column_view col, int begin_bit, int end_bit, void* d_temp_storage = nullptr, size_t temp_storage_bytes = 0,
    hipStream_t stream = nullptr)
{
    eg_column_view k{}, i{};
    detail::check(eg_arg_partition(col.abi(), begin_bit, end_bit, d_temp_storage, temp_storage_bytes, stream, &k, &i));
    return {detail::own(k), detail::own(i)};
}
inline std::unique_ptr<column> compute_partition_offsets(column_view partitioned_key_col, int begin_bit, int end_bit,
                                                         hipStream_t stream = nullptr)
{
    eg_column_view o{};
    detail::check(eg_compute_partition_offsets(partitioned_key_col.abi(), begin_bit, end_bit, stream, &o));
    return detail::own(o);
}
}  // namespace partition

namespace compress {
// compress/dictionary_encode.hpp:12-37 (integer key columns).  The
// self-built forms return (encoded, dictionary size); load_factor is
// accepted for signature parity and not used (the table is sized here).
inline std::pair<std::unique_ptr<column>, size_type> dictionary_encode(
    column_view col, data_type encode_type = data_type{type_id::INT32}, size_type dictionary_size_hint = 0,
    double /*load_factor*/ = 0.25, hipStream_t stream = nullptr)
{
    eg_column_view e{};
    int64_t n = 0;
    detail::check(eg_dictionary_encode(col.abi(), (int32_t)encode_type.id(), dictionary_size_hint, 0, stream, &e, &n,
                                       nullptr));
    return {detail::own(e), n};
}
inline std::pair<std::unique_ptr<column>, size_type> dictionary_encode_with_order(
    column_view col, data_type encode_type = data_type{type_id::INT32}, size_type dictionary_size_hint = 0,
    double /*load_factor*/ = 0.25, hipStream_t stream = nullptr)
{
    eg_column_view e{};
    int64_t n = 0;
    detail::check(eg_dictionary_encode(col.abi(), (int32_t)encode_type.id(), dictionary_size_hint, 1, stream, &e, &n,
                                       nullptr));
    return {detail::own(e), n};
}
// dictionary_encode_dict_ready.cu:15-80: positions in `dict`, INT32_MIN if absent.
inline std::unique_ptr<column> dictionary_encode(column_view col, column_view dict,
                                                 data_type encode_type = data_type{type_id::INT32},
                                                 hipStream_t stream = nullptr)
{
    eg_column_view e{};
    detail::check(eg_dictionary_encode_with_dict(col.abi(), dict.abi(), (int32_t)encode_type.id(), stream, &e));
    return detail::own(e);
}
}  // namespace compress

namespace stats {
// stats/stats_compute.hpp:72
inline size_type compute_ndv_approx(column_view col, hipStream_t stream = nullptr)
{
    int64_t ndv = 0;
    detail::check(eg_compute_ndv_approx(col.abi(), stream, &ndv));
    return ndv;
}
}  // namespace stats
}  // namespace gamax

// pullops.hpp -- pull-based ("volcano") operators around the group-by:
// Operator::open / next / close as the reference's pipelines drive them
// (benchmarks/tpc/tpch_q18.cu:58-75: open the root, call next until a chunk
// is_end, free each chunk, close).  The reference snapshot has no GroupBy
// operator (its pullops headers are absent, benchmarks/join/join_exp.cu:17
// includes one that does not exist), so this layer is designed here, not
// mirrored:
//   TableScan       emits an existing device table in chunks of vec_size rows;
//   GroupByOperator blocking: open() drains the child, concatenates its chunks
//                   on the device and runs gamax::groupby::groupby once;
//                   next() emits the result in chunks of vec_size rows.
// Chunks own their device columns.  Header-only; needs gamax.hpp, the HIP
// runtime (device copies) and libeiger_groupby.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gamax.hpp"

namespace gamax {
namespace pullops {

// A batch of rows; a chunk without data is the end of the stream.
struct Chunk {
    std::unique_ptr<table> data;
    size_type num_items = 0;
    bool is_end() const noexcept { return !data; }
    void free_mem() noexcept
    {
        data.reset();
        num_items = 0;
    }
};

class Operator {
public:
    virtual ~Operator() = default;
    virtual void open() = 0;
    virtual Chunk next() = 0;
    virtual void close() = 0;
};

namespace detail {
inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Rows [begin, begin + count) of `src` as a new owning table.
inline std::unique_ptr<table> slice_copy(table_view src, size_type begin, size_type count, hipStream_t s)
{
    std::vector<std::unique_ptr<column>> cols;
    for (size_type c = 0; c < src.num_columns(); ++c) {
        const column_view v = src.get_column(c);
        auto out = std::make_unique<column>(v.type(), count, false, s);
        if (count > 0) {
            const size_t w = (size_t)size_of(v.type());
            hip_check(hipMemcpyAsync(out->data(), static_cast<const char*>(v.data()) + (size_t)begin * w,
                                     (size_t)count * w, hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync");
        }
        cols.push_back(std::move(out));
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return std::make_unique<table>(std::move(cols));
}

// The chunks' columns end to end (same schema), as a new owning table.
inline std::unique_ptr<table> concatenate(std::vector<std::unique_ptr<table>> const& parts, hipStream_t s)
{
    if (parts.empty()) throw std::runtime_error("concatenate: no chunks");
    const size_type nc = parts[0]->num_columns();
    size_type rows = 0;
    for (auto const& p : parts) {
        if (p->num_columns() != nc) throw std::runtime_error("chunks with different column counts");
        rows += p->num_rows();
    }
    std::vector<std::unique_ptr<column>> cols;
    for (size_type c = 0; c < nc; ++c) {
        const data_type t = parts[0]->get_column(c).type();
        auto out = std::make_unique<column>(t, rows, false, s);
        const size_t w = (size_t)size_of(t);
        size_type at = 0;
        for (auto const& p : parts) {
            column const& in = p->get_column(c);
            if (in.type() != t) throw std::runtime_error("chunks with different column types");
            if (in.size() > 0)
                hip_check(hipMemcpyAsync(out->data<char>() + (size_t)at * w, in.data(), (size_t)in.size() * w,
                                         hipMemcpyDeviceToDevice, s),
                          "hipMemcpyAsync");
            at += in.size();
        }
        cols.push_back(std::move(out));
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return std::make_unique<table>(std::move(cols));
}
}  // namespace detail

// Emits the rows of a device table (not owned) in chunks of vec_size rows.
class TableScan : public Operator {
public:
    TableScan(table_view input, size_type vec_size, hipStream_t stream = nullptr)
        : _in(std::move(input)), _vec(vec_size), _stream(stream)
    {
        if (vec_size <= 0) throw std::runtime_error("vec_size must be positive");
    }
    void open() override { _pos = 0; }
    Chunk next() override
    {
        const size_type n = _in.num_rows();
        if (_pos >= n) return Chunk{};
        const size_type cnt = std::min(_vec, n - _pos);
        Chunk c{detail::slice_copy(_in, _pos, cnt, _stream), cnt};
        _pos += cnt;
        return c;
    }
    void close() override {}

private:
    table_view _in;
    size_type _vec;
    hipStream_t _stream;
    size_type _pos = 0;
};

// Blocking grouped aggregation over the child's rows.  The child's chunks
// hold the key columns first (key_columns of them), then the value columns;
// requests index the value columns, as in gamax::groupby::groupby.  The
// result chunks hold [keys, one column per request (, count)].
class GroupByOperator : public Operator {
public:
    GroupByOperator(Operator* child, size_type key_columns, std::vector<groupby::aggregation_request> requests,
                    groupby::groupby_hint hint = {}, size_type vec_size = size_type(1) << 24,
                    hipStream_t stream = nullptr)
        : _child(child), _nkeys(key_columns), _reqs(std::move(requests)), _hint(hint), _vec(vec_size),
          _stream(stream)
    {
        if (!child) throw std::runtime_error("GroupByOperator: no child");
        if (key_columns <= 0) throw std::runtime_error("GroupByOperator: no key columns");
        if (vec_size <= 0) throw std::runtime_error("vec_size must be positive");
    }
    void open() override
    {
        _child->open();
        std::vector<std::unique_ptr<table>> parts;
        for (Chunk c = _child->next(); !c.is_end(); c = _child->next())
            if (c.data->num_rows() > 0) parts.push_back(std::move(c.data));
        _pos = 0;
        _result.reset();
        if (parts.empty()) return;  // no input rows: no groups
        auto all = parts.size() == 1 ? std::move(parts[0]) : detail::concatenate(parts, _stream);
        parts.clear();
        const table_view v = all->view();
        std::vector<size_type> kidx, vidx;
        for (size_type c = 0; c < v.num_columns(); ++c) (c < _nkeys ? kidx : vidx).push_back(c);
        groupby::groupby_hint h = _hint;
        _result = groupby::groupby(v.select_columns(kidx), v.select_columns(vidx), _reqs, h, _stream);
        _hint_used = h;
    }
    Chunk next() override
    {
        if (!_result || _pos >= _result->num_rows()) return Chunk{};
        const size_type cnt = std::min(_vec, _result->num_rows() - _pos);
        Chunk c;
        if (_pos == 0 && cnt == _result->num_rows()) {
            c.data = std::move(_result);  // the whole result in one chunk
        } else {
            c.data = detail::slice_copy(_result->view(), _pos, cnt, _stream);
        }
        c.num_items = cnt;
        _pos += cnt;
        return c;
    }
    void close() override
    {
        _result.reset();
        _child->close();
    }
    // The hint after the call (AUTO resolved to the strategy that ran).
    groupby::groupby_hint const& hint_used() const noexcept { return _hint_used; }

private:
    Operator* _child;
    size_type _nkeys;
    std::vector<groupby::aggregation_request> _reqs;
    groupby::groupby_hint _hint;
    groupby::groupby_hint _hint_used;
    size_type _vec;
    hipStream_t _stream;
    std::unique_ptr<table> _result;
    size_type _pos = 0;
};

}  // namespace pullops
}  // namespace gamax

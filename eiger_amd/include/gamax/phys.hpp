// phys.hpp -- the engine-facing aggregation layer over the gamax mirror:
// gamax::phys::group_by, scalar_aggregate and distinct, the functions the
// reference's executor calls for a GroupBy / Aggregate / Distinct step
// (phys/phys.hpp:59-69,213-240, phys/phys.cu:535-663).  Header-only; needs
// gamax.hpp and libeiger_groupby.so.
//
// Behaviour kept from the reference:
//   * AggSpec = (input column, aggregation id); the request list is built
//     from it in order;
//   * group_by over 0 rows returns 0 rows with the schema
//     [key types..., agg output types...] without calling the operator
//     (phys.cu:585-606);
//   * scalar_aggregate over 0 rows returns one row of zeros typed per
//     AggSpec (phys.cu:541-566), where the operator itself would return the
//     aggregation identities;
//   * COUNT_ALL never reads its column index (INT64 output);
//   * the caller's hint is copied, so the caller's own hint is not resolved
//     in place (phys.cu:613-616).
#pragma once

#include <memory>
#include <vector>

#include "gamax.hpp"

namespace gamax {
namespace phys {

// phys/phys.hpp:47-56 (the profiling sink is not carried over).
struct Context {
    hipStream_t stream{nullptr};
};

using table_ptr = std::unique_ptr<gamax::table>;

// phys/phys.hpp:65-68
struct AggSpec {
    gamax::size_type column;
    gamax::groupby::aggregation_id agg;
};

namespace detail {
inline data_type output_type(AggSpec const& s, table_view values)
{
    if (s.agg == gamax::groupby::aggregation_id::COUNT_ALL) return data_type{type_id::INT64};
    return gamax::groupby::agg_output_data_type(s.agg, values.get_column(s.column).type());
}

inline std::vector<gamax::groupby::aggregation_request> requests_of(std::vector<AggSpec> const& specs)
{
    std::vector<gamax::groupby::aggregation_request> r;
    r.reserve(specs.size());
    for (auto const& s : specs) r.emplace_back(s.column, s.agg);
    return r;
}
}  // namespace detail

// Scalar aggregate: one row, one column per AggSpec.
inline table_ptr scalar_aggregate(table_view values, std::vector<AggSpec> specs, Context const& ctx)
{
    if (values.num_rows() == 0) {
        std::vector<column_ptr> cols;
        cols.reserve(specs.size());
        for (auto const& s : specs)
            cols.push_back(std::make_unique<column>(detail::output_type(s, values), 1, /*clear=*/true, ctx.stream));
        return std::make_unique<table>(std::move(cols));
    }
    auto reqs = detail::requests_of(specs);
    return gamax::groupby::aggregate::aggregate(values, reqs, ctx.stream);
}

// Grouped aggregate.  `keys` and `values` have equal row counts; the result
// is [keys..., one column per AggSpec], one row per group.
inline table_ptr group_by(table_view keys, table_view values, std::vector<AggSpec> specs, Context const& ctx,
                          gamax::groupby::groupby_hint const& hint_in = {})
{
    if (keys.num_rows() == 0) {
        std::vector<column_ptr> cols;
        cols.reserve((size_t)keys.num_columns() + specs.size());
        for (size_type i = 0; i < keys.num_columns(); ++i)
            cols.push_back(std::make_unique<column>(keys.get_column(i).type(), 0, false, ctx.stream));
        for (auto const& s : specs)
            cols.push_back(std::make_unique<column>(detail::output_type(s, values), 0, false, ctx.stream));
        return std::make_unique<table>(std::move(cols));
    }
    auto reqs = detail::requests_of(specs);
    auto hint = hint_in;
    return gamax::groupby::groupby(keys, values, reqs, hint, ctx.stream);
}

// Distinct rows of `keys` (phys.cu:622-663): HASH / AUTO go to
// hash::distinct, the sort family to sort::distinct, PARTITION variants
// throw; 0 rows give 0 rows with the key schema.
inline table_ptr distinct(table_view keys, Context const& ctx, gamax::groupby::groupby_hint const& hint_in = {})
{
    if (keys.num_rows() == 0) {
        std::vector<column_ptr> cols;
        for (size_type i = 0; i < keys.num_columns(); ++i)
            cols.push_back(std::make_unique<column>(keys.get_column(i).type(), 0, false, ctx.stream));
        return std::make_unique<table>(std::move(cols));
    }
    auto hint = hint_in;
    switch (hint.groupby_algo) {
    case gamax::groupby::HASH:
    case gamax::groupby::AUTO_ALGO: return gamax::groupby::hash::distinct(keys, hint, ctx.stream);
    case gamax::groupby::SORT:
    case gamax::groupby::SORT_OPT:
    case gamax::groupby::SORT_GFUR: return gamax::groupby::sort::distinct(keys, hint, ctx.stream);
    case gamax::groupby::PARTITION:
    case gamax::groupby::PARTITION_GFUR:
        throw std::runtime_error("phys::distinct: PARTITION variants are not supported; use HASH or SORT");
    default: throw std::runtime_error("phys::distinct: unknown groupby_algo");
    }
}

}  // namespace phys
}  // namespace gamax

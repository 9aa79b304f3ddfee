/*
 * eiger_groupby.h -- C ABI of the MI355X single-key grouped-aggregation library
 * (libeiger_groupby.so).
 *
 * This is the drop-in boundary.  Each entry point replaces one function of the
 * reference's C++ operator API (paths relative to src/cuda/ of the reference
 * repository, cited per declaration).  The reference has no C ABI: a binding
 * (ctypes, cgo, JNI, or the header-only gamax C++ shim under
 * eiger_amd/include/gamax/) calls these functions with plain pointers and sizes.
 *
 * Conventions
 *   - Every function returns an eg_status (0 = EG_OK).  On failure the
 *     thread-local message behind eg_last_error() says why.  The C++ shim turns
 *     EG_ERR_INVALID / EG_ERR_UNSUPPORTED / EG_ERR_CAPACITY into
 *     std::runtime_error and EG_ERR_OOM into std::bad_alloc, as the reference
 *     throws (operators/groupby/groupby_hash.cu:32-40, groupby.cu:287-324,
 *     memory/mem_manager.cu:152-167).
 *   - Column data are device pointers.  Inputs are never modified or freed.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Single-query entry points are synchronous at return, as the reference's
 *     (groupby.cu:390).  eg_batch_groupby is asynchronous and graph-capturable.
 *   - Result tables (eg_table) own their device columns; free them with
 *     eg_table_free, or take a column with eg_table_release_column and free its
 *     data with eg_free.
 */
#ifndef EIGER_GROUPBY_H
#define EIGER_GROUPBY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
typedef enum eg_status {
    EG_OK = 0,
    EG_ERR_INVALID = 1,      /* bad argument (reference: std::runtime_error) */
    EG_ERR_UNSUPPORTED = 2,  /* shape/type/agg not supported (std::runtime_error) */
    EG_ERR_CAPACITY = 3,     /* a caller-sized buffer or fixed table is too small */
    EG_ERR_OOM = 4,          /* device allocation failed (std::bad_alloc) */
    EG_ERR_DEVICE = 5        /* a HIP runtime call failed */
} eg_status;

/* ---- type ids: numbering of gamax::type_id (types/data_type.hpp:12-28) ---- */
enum {
    EG_TYPE_EMPTY = 0, EG_INT8 = 1, EG_INT16 = 2, EG_INT32 = 3, EG_INT64 = 4,
    EG_BOOL8 = 5, EG_UINT8 = 6, EG_UINT16 = 7, EG_UINT32 = 8, EG_UINT64 = 9,
    EG_FLOAT32 = 10, EG_FLOAT64 = 11, EG_STRING = 12
};

/* ---- aggregation ids: gamax::groupby::aggregation_id (aggregation.hpp:9-22) */
enum {
    EG_AGG_SUM = 0, EG_AGG_MAX = 1, EG_AGG_MIN = 2, EG_AGG_AVG = 3,
    EG_AGG_ANY = 4, EG_AGG_COUNT_ALL = 5
};

/* ---- strategies: gamax::groupby::groupby_algo (groupby.hpp:16-25) --------- */
enum {
    EG_ALGO_AUTO = 0, EG_ALGO_HASH = 1, EG_ALGO_SORT = 2, EG_ALGO_SORT_OPT = 3,
    EG_ALGO_SORT_GFUR = 4, EG_ALGO_PARTITION = 5, EG_ALGO_PARTITION_GFUR = 6
};

/* ---- sort order: gamax::sort_order (types/data_type.hpp:55-58) ------------ */
enum { EG_ASCENDING = 0, EG_DESCENDING = 1 };

/* gamax::column_view (memory/column.hpp:13-58): data + size + type + offset
 * in elements.  `size` counts the rows after `offset`. */
typedef struct eg_column_view {
    const void* data;
    int64_t size;
    int32_t type;
    int32_t reserved;
    int64_t offset;
} eg_column_view;

/* gamax::groupby::aggregation_request (aggregation.hpp:24-31).  COUNT_ALL
 * ignores column_index (-1 by convention). */
typedef struct eg_agg_request {
    int64_t column_index;
    int32_t agg_kind;
    int32_t reserved;
} eg_agg_request;

/* gamax::groupby::groupby_hint (groupby.hpp:27-42).  In/out, as in the
 * reference: EG_ALGO_AUTO is resolved in place (groupby.cu:68-119) and
 * EG_ALGO_SORT_OPT sets optimize_sort (groupby.cu:304-306). */
typedef struct eg_groupby_hint {
    int64_t groups_upper_bound;   /* INT64_MAX = unknown */
    int32_t zero_based_ids;
    int32_t optimize_sort;
    int32_t groupby_algo;
    int32_t agg_count;
    int32_t enable_skf;
    int32_t allow_hash_fallback;
} eg_groupby_hint;

/* Owning result table (gamax::table, memory/table.hpp:143-222). */
typedef struct eg_table_s* eg_table;

/* ---- library ------------------------------------------------------------- */
const char* eg_last_error(void);
const char* eg_version(void);
/* Default-initialised hint (groupby.hpp:27-42 defaults). */
void eg_hint_init(eg_groupby_hint* hint);

/* ---- group-by -------------------------------------------------------------
 * Replaces gamax::groupby::groupby (operators/groupby/groupby.hpp:44-49,
 * groupby.cu:396-465).  Result columns: [key, one column per request in
 * request order, trailing COUNT if hint->agg_count] (groupby.cu:326-388), typed
 * as agg_output_data_type (aggregate.cu:20-53); num_rows = number of groups.
 * Row order is strategy-dependent and unspecified.  Single-column integer
 * keys (4 or 8 bytes); an empty input gives a 0-row table with that schema
 * (phys/phys.cu:589-611). */
int eg_groupby(const eg_column_view* keys, int64_t num_keys,
               const eg_column_view* vals, int64_t num_vals,
               const eg_agg_request* reqs, int64_t num_reqs,
               eg_groupby_hint* hint, void* stream, eg_table* out);

/* Strategy entry points with the same contract as the reference's
 * (requests SUM/MAX/MIN/ANY/COUNT_ALL; AVG is rewritten only by eg_groupby):
 *   hash::groupby            operators/groupby/groupby_hash.hpp:17-22
 *   sort::groupby            operators/groupby/groupby_sort.hpp:17-22
 *   sort::groupby_gfur       operators/groupby/groupby_sort.hpp:24-29
 *   partition::groupby       operators/groupby/groupby_partition.hpp:52-57
 *   partition::groupby_gfur  operators/groupby/groupby_partition.hpp:59-64 */
int eg_groupby_hash(const eg_column_view* keys, int64_t num_keys,
                    const eg_column_view* vals, int64_t num_vals,
                    const eg_agg_request* reqs, int64_t num_reqs,
                    eg_groupby_hint* hint, void* stream, eg_table* out);
int eg_groupby_sort(const eg_column_view* keys, int64_t num_keys,
                    const eg_column_view* vals, int64_t num_vals,
                    const eg_agg_request* reqs, int64_t num_reqs,
                    eg_groupby_hint* hint, void* stream, eg_table* out);
int eg_groupby_sort_gfur(const eg_column_view* keys, int64_t num_keys,
                         const eg_column_view* vals, int64_t num_vals,
                         const eg_agg_request* reqs, int64_t num_reqs,
                         eg_groupby_hint* hint, void* stream, eg_table* out);
int eg_groupby_partition(const eg_column_view* keys, int64_t num_keys,
                         const eg_column_view* vals, int64_t num_vals,
                         const eg_agg_request* reqs, int64_t num_reqs,
                         eg_groupby_hint* hint, void* stream, eg_table* out);
int eg_groupby_partition_gfur(const eg_column_view* keys, int64_t num_keys,
                              const eg_column_view* vals, int64_t num_vals,
                              const eg_agg_request* reqs, int64_t num_reqs,
                              eg_groupby_hint* hint, void* stream, eg_table* out);

/* gamax::groupby::aggregate::aggregate (aggregation.hpp:34-39,
 * aggregate.cu:131-165): one row, one column per request. */
int eg_aggregate(const eg_column_view* cols, int64_t num_cols,
                 const eg_agg_request* reqs, int64_t num_reqs, void* stream,
                 eg_table* out);

/* gamax::groupby::agg_output_data_type (aggregation.hpp:33, aggregate.cu:42-53). */
int eg_agg_output_type(int32_t agg_kind, int32_t type, int32_t* out_type);

/* ---- result tables ------------------------------------------------------- */
int64_t eg_table_num_columns(eg_table t);
int64_t eg_table_num_rows(eg_table t);
int eg_table_get_column(eg_table t, int64_t i, eg_column_view* out);
/* Moves column i out (table::release_column, memory/table.hpp:194-210); the
 * slot stays, empty.  The caller frees out->data with eg_free. */
int eg_table_release_column(eg_table t, int64_t i, eg_column_view* out);
void eg_table_free(eg_table t);

/* ---- device memory (gamax::MemoryManager, memory/mem_manager.hpp:12-158) -- */
enum { EG_MM_ASYNC = 0, EG_MM_POOL = 1 };
int eg_malloc(void** ptr, int64_t bytes, int32_t clear, void* stream);
int eg_free(void* ptr);
int eg_memory_usage(int64_t* current_bytes, int64_t* peak_bytes);
/* EG_MM_ASYNC: stream-ordered allocations (the reference's CudaMM);
 * EG_MM_POOL: one slab of pool_bytes with a first-fit free list, 256-B
 * aligned (the reference's UserSpaceMM).  Only while nothing is allocated. */
int eg_set_memory_mode(int32_t mode, int64_t pool_bytes);
/* gamax::make_column_from_host (memory/column.hpp:185-191): a new device
 * column (free its data with eg_free) holding a copy of num_items host
 * elements of `type`. */
int eg_make_column_from_host(const void* host_data, int64_t num_items, int32_t type,
                             void* stream, eg_column_view* out);
/* gamax::copy_column_to_host (memory/column.hpp:199-202); synchronous. */
int eg_copy_column_to_host(eg_column_view col, void* host_data, void* stream);

/* ---- sort / partition sub-operators ----------------------------------------
 * Outputs are new device allocations owned by the caller (free with eg_free).
 * temp/temp_bytes: optional caller scratch (NULL = allocate internally), as
 * the reference's d_temp_storage/temp_storage_bytes.  end_bit <= 0 means all
 * bits of the key type. */
/* gamax::sort::sort_column_pair (operators/sort/sort.hpp:25-34, sort.cu:115-153) */
int eg_sort_column_pair(eg_column_view key, eg_column_view val, int32_t begin_bit,
                        int32_t end_bit, int32_t order, void* temp,
                        size_t temp_bytes, void* stream, eg_column_view* out_key,
                        eg_column_view* out_val);
/* gamax::sort::sort_column (sort.hpp:46-54, sort.cu:167-197) */
int eg_sort_column(eg_column_view col, int32_t begin_bit, int32_t end_bit,
                   int32_t order, void* temp, size_t temp_bytes, void* stream,
                   eg_column_view* out);
/* gamax::sort::argsort (sort.hpp:66-75, sort.cu:199-215): sorted column and
 * int64 row ids. */
int eg_argsort(eg_column_view col, int32_t begin_bit, int32_t end_bit,
               int32_t order, void* temp, size_t temp_bytes, void* stream,
               eg_column_view* out_sorted, eg_column_view* out_idx);
/* gamax::partition::partition_column_pair (operators/partition/partition.hpp:46-54,
 * partition.cu:24-43): stable order by bits [begin_bit, end_bit). */
int eg_partition_column_pair(eg_column_view key, eg_column_view val,
                             int32_t begin_bit, int32_t end_bit, void* temp,
                             size_t temp_bytes, void* stream,
                             eg_column_view* out_key, eg_column_view* out_val);
/* gamax::partition::partition_column (partition.hpp:56-63) */
int eg_partition_column(eg_column_view col, int32_t begin_bit, int32_t end_bit,
                        void* temp, size_t temp_bytes, void* stream,
                        eg_column_view* out);
/* gamax::partition::arg_partition (partition.hpp:65-72, partition.cu:88-99) */
int eg_arg_partition(eg_column_view col, int32_t begin_bit, int32_t end_bit,
                     void* temp, size_t temp_bytes, void* stream,
                     eg_column_view* out_sorted, eg_column_view* out_idx);
/* gamax::partition::compute_partition_offsets (partition.hpp:74-79,
 * partition.cu:122-171): 2^(end-begin)+1 int64 entries, the first 2^(end-begin)
 * the inclusive prefix of the bucket counts, the last = rows. */
int eg_compute_partition_offsets(eg_column_view partitioned_key, int32_t begin_bit,
                                 int32_t end_bit, void* stream,
                                 eg_column_view* out_offsets);

/* ---- statistics and dispatch ----------------------------------------------- */
/* gamax::stats::compute_ndv_approx (stats/stats_compute.hpp:72, stats_compute.cu:103-196):
 * HyperLogLog over the whole column, 2^14 registers. */
int eg_compute_ndv_approx(eg_column_view col, void* stream, int64_t* out_ndv);
/* auto_select_groupby_algo (operators/groupby/groupby.cu:68-119): resolves
 * hint->groupby_algo in place when it is EG_ALGO_AUTO. */
int eg_auto_select_algo(const eg_column_view* keys, int64_t num_keys,
                        eg_groupby_hint* hint, void* stream);

/* ---- kernel timing (measurement support; no reference counterpart) -------- */
int eg_kernel_timing_enable(int32_t on);
int eg_kernel_timing_reset(void);
/* Sums the hipEvent-measured durations of every launch of kernel `name`
 * since the last reset (synchronises the recorded events). */
int eg_kernel_timing_query(const char* name, int64_t* launches, double* total_ms);

/* Reference copy kernel for bandwidth calibration: copies `bytes` between two
 * scratch buffers `iters` times with 16-byte vector loads/stores (kernel name
 * "probe_copy"); *ms = mean time per copy.  Mirrors the reference's
 * read/write bandwidth probe (benchmarks/misc/read_write_nvbench.cu:29-80). */
int eg_probe_copy(int64_t bytes, int32_t iters, void* stream, double* ms);

/* ---- synthetic data (data_gen/generator.cuh semantics; counter-based) ------ */
enum {
    EG_VAL_UNIFORM_I64 = 0, EG_VAL_UNIFORM_F32 = 1, EG_VAL_UNIFORM_I32_1000 = 2,
    EG_VAL_ALL_ONES = 3, EG_VAL_SEQ = 4
};
/* key[i] = perm(begin+i) mod distinct, perm a seeded bijection of [0, n). */
int eg_gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count,
                        int64_t n, int64_t distinct, uint64_t seed, void* stream);
/* Host-side Zipf CDF (w_r = 1/(r+1)^s); cdf holds `distinct` doubles (host). */
int eg_zipf_cdf(double* cdf, int64_t distinct, double s);
/* key[i] = smallest r with d_cdf[r] > u(begin+i); d_cdf on the device. */
int eg_gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count,
                     const double* d_cdf, int64_t distinct, uint64_t seed,
                     void* stream);
int eg_gen_values(void* out, int32_t type, int32_t dist, int64_t begin,
                  int64_t count, uint64_t seed, uint32_t stream_id, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EIGER_GROUPBY_H */

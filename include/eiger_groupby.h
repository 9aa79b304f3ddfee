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
 *     (groupby.cu:390).  eg_batch_run / eg_batch_graph_launch are asynchronous
 *     and graph-capturable.
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
 * Row order is strategy-dependent and unspecified.  A key column may be any
 * integer type, BOOL8, FLOAT32 or FLOAT64 (partition: integers only).
 * Several key columns (integer, BOOL8 or float) are fused into one 64-bit
 * key after range reduction (the reference's SKF path, groupby.cu:396-465;
 * compress/multikey_compress.hpp:11-43) -- with AUTO or HASH, or any
 * strategy when enable_skf is set.  Key sets wider than 64 bits or of more
 * than nine columns (the reference's multi-column hash table,
 * groupby_hash.cu:212-270) are folded through order-preserving dense ranks
 * first (inputs below 2^32 rows).  An empty input gives a 0-row table with
 * that schema (phys/phys.cu:589-611). */
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

/* The hash strategy with nothing allocated: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * requests and results as eg_groupby_hash with one key column, written into
 * caller-owned columns out_cols = [key, one per request], typed as that
 * result; each column's size is its capacity in rows (offset honoured).
 * *num_groups receives the group count; synchronous at return.  The HBM
 * table lives in a static device array of 32 MiB per GPU (held until the
 * process ends), so groups_upper_bound (<= 0: the row count) is limited to
 * what that table holds at load 1/2: 2^20 / W groups, W the power of two
 * >= 1 + aggregation states (262144 groups for up to 3 states, 131072 for
 * 4 to 7).  A larger
 * bound, more groups than the table holds, or more groups than the output
 * capacity give EG_ERR_CAPACITY (outputs untouched).  Calls that overlap on
 * one GPU share that array and race: serialise them per device.  The
 * generic kernels run (no run-time specialisation). */
int eg_groupby_hash_noalloc(const eg_column_view* keys, int64_t num_keys,
                            const eg_column_view* vals, int64_t num_vals,
                            const eg_agg_request* reqs, int64_t num_reqs,
                            int64_t groups_upper_bound,
                            const eg_column_view* out_cols, int64_t num_out_cols,
                            int64_t* num_groups, void* stream);

/* Distinct rows of the key columns, one per unique key (or key tuple):
 *   hash::distinct  operators/groupby/groupby_hash.hpp:24-32 (any key
 *                   shape the group-by takes; first-encounter order is not
 *                   kept: row order is unspecified)
 *   sort::distinct  operators/groupby/groupby_sort.hpp:31-48 (one key
 *                   column, or several with hint->enable_skf; ascending)
 * The result has exactly the key columns. */
int eg_distinct_hash(const eg_column_view* keys, int64_t num_keys,
                     eg_groupby_hint* hint, void* stream, eg_table* out);
int eg_distinct_sort(const eg_column_view* keys, int64_t num_keys,
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
enum { EG_MM_ASYNC = 0, EG_MM_POOL = 1, EG_MM_CACHE = 2 };
int eg_malloc(void** ptr, int64_t bytes, int32_t clear, void* stream);
int eg_free(void* ptr);
int eg_memory_usage(int64_t* current_bytes, int64_t* peak_bytes);
/* EG_MM_CACHE (default): blocks from hipMalloc in size classes (four per
 * power of two), kept after eg_free and reused by later allocations on the
 * same stream (stream order makes the reuse safe); cached blocks are returned
 * to HIP only when an allocation would otherwise fail.
 * EG_MM_ASYNC: stream-ordered hipMallocAsync (the reference's CudaMM);
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

/* ---- dictionary encoding (gamax::compress, compress/dictionary_encode.hpp:12-37) --
 * Integer (and BOOL8) key columns only.  The encoded column is a new device
 * allocation owned by the caller (free with eg_free). */
/* dictionary_encode(column, dict, encode_type) (dictionary_encode_dict_ready.cu:15-80):
 * each row gets the position of its key in `dict` (same type as column), or
 * INT32_MIN when absent; a key present more than once in dict gets its first
 * position.  encode_type must be EG_INT32. */
int eg_dictionary_encode_with_dict(eg_column_view column, eg_column_view dict, int32_t encode_type,
                                   void* stream, eg_column_view* out);
/* dictionary_encode / dictionary_encode_with_order (dictionary_encode.hpp:12-24):
 * each row gets a dense id in [0, *out_dict_size); with_order != 0 numbers the
 * distinct keys in ascending key order.  encode_type: any integer type wide
 * enough for the number of distinct keys (else EG_ERR_INVALID).
 * dictionary_size_hint (0 = none) sizes the distinct-key pass.  out_dict
 * (optional, NULL = not returned) receives the distinct keys, id i at row i. */
int eg_dictionary_encode(eg_column_view column, int32_t encode_type, int64_t dictionary_size_hint,
                         int32_t with_order, void* stream, eg_column_view* out, int64_t* out_dict_size,
                         eg_column_view* out_dict);

/* ---- statistics and dispatch ----------------------------------------------- */
/* gamax::stats::compute_ndv_approx (stats/stats_compute.hpp:72, stats_compute.cu:103-196):
 * HyperLogLog over the whole column, 2^14 registers. */
int eg_compute_ndv_approx(eg_column_view col, void* stream, int64_t* out_ndv);
/* auto_select_groupby_algo (operators/groupby/groupby.cu:68-119): resolves
 * hint->groupby_algo in place when it is EG_ALGO_AUTO. */
int eg_auto_select_algo(const eg_column_view* keys, int64_t num_keys,
                        eg_groupby_hint* hint, void* stream);
/* The same rule for `num` independent single-column keys with one
 * synchronisation (no reference counterpart: batched callers); algos[i] gets
 * EG_ALGO_HASH or EG_ALGO_PARTITION, ndv[i] (optional) the estimate. */
int eg_auto_select_algo_batch(const eg_column_view* keys, int64_t num, int32_t* algos, int64_t* ndv,
                              void* stream);

/* ---- batched group-by (no reference counterpart) -----------------------------
 * Many independent single-key group-bys executed by one persistent launch.
 * The reference runs one group-by per call, each synchronous
 * (operators/groupby/groupby.hpp:44-49, groupby.cu:390); this entry point is
 * for callers that hold thousands of small and uneven queries.  Every query
 * has one key column (key_type) and at most one value column (val_type,
 * column_index 0 in every request); the requests (SUM/MAX/MIN/ANY/COUNT_ALL)
 * are the same for all queries.  Outputs are caller-allocated columns of
 * `capacity` rows: [out_keys, out_aggs[j] per request], typed as
 * eg_agg_output_type.  Row order is unspecified.
 *
 * eg_batch_plan_create allocates the plan's device tables and uploads the
 * work list (synchronous).  eg_batch_run and eg_batch_graph_launch are
 * asynchronous, make no allocation and no host synchronisation, so the
 * execution can be captured in a hipGraph (eg_batch_graph_launch captures it
 * once on its first call and replays the graph afterwards).  Queries whose
 * group_bound is below their true group count, or whose groups exceed
 * `capacity`, come back with a non-zero status from eg_batch_results. */
#define EG_BATCH_MAX_AGGS 4
enum { EG_BATCH_TABLE_FULL = 1, EG_BATCH_CAPACITY = 2 };
typedef struct eg_batch_query {
    const void* keys;        /* device, rows elements of key_type */
    const void* vals;        /* device, rows elements of val_type (NULL if only COUNT_ALL) */
    int64_t rows;
    int64_t group_bound;     /* upper bound of the distinct keys (<= 0: rows) */
    int64_t capacity;        /* rows available in each output column */
    void* out_keys;
    void* out_aggs[EG_BATCH_MAX_AGGS];
} eg_batch_query;
/* Execution parameters picked by eg_batch_autotune. */
typedef struct eg_batch_tuning {
    int32_t tile_rows_log2;      /* rows per work item (10..26) */
    int32_t lds_slots_log2;      /* LDS table slots per workgroup (6..13) */
    int32_t lds_max_bound_log2;  /* queries with group_bound <= 2^this use the LDS table */
    int32_t blocks_per_cu;       /* persistent workgroups per CU (1..8) */
    int32_t radix_bits;          /* queries above the LDS bound: one radix pass on up to this many
                                    hash bits (0..12) into partitions of about half an LDS table,
                                    each aggregated in LDS; 0: per-row atomics on an HBM table */
} eg_batch_tuning;
typedef struct eg_batch_plan_s* eg_batch_plan;
void eg_batch_tuning_default(eg_batch_tuning* t);
int eg_batch_plan_create(const eg_batch_query* queries, int64_t num_queries, int32_t key_type,
                         int32_t val_type, const eg_agg_request* reqs, int64_t num_reqs,
                         const eg_batch_tuning* tuning, void* stream, eg_batch_plan* out);
/* The tuning the plan runs with (LDS slots clamped to what fits). */
int eg_batch_plan_tuning(eg_batch_plan plan, eg_batch_tuning* out);
int eg_batch_run(eg_batch_plan plan, void* stream);
int eg_batch_graph_launch(eg_batch_plan plan, void* stream);
/* Synchronises `stream`; num_groups / status: num_queries entries (host). */
int eg_batch_results(eg_batch_plan plan, int64_t* num_groups, int32_t* status, void* stream);
void eg_batch_plan_free(eg_batch_plan plan);
/* Times every candidate tuning on this batch (reps timed runs after a
 * warm-up, hipEvents on `stream`) and returns the fastest; times_ms
 * (optional, num_candidates entries) receives each candidate's mean time. */
int eg_batch_autotune(const eg_batch_query* queries, int64_t num_queries, int32_t key_type,
                      int32_t val_type, const eg_agg_request* reqs, int64_t num_reqs,
                      const eg_batch_tuning* candidates, int64_t num_candidates, int32_t reps,
                      void* stream, eg_batch_tuning* best, double* times_ms);

/* ---- learned dispatch (no reference counterpart) ----------------------------
 * A 16 -> 64 -> 64 -> 4 ReLU MLP (fp32) over 16 features of a 4096-row key
 * sample.  Outputs 0..2: logits of {HASH, SORT, PARTITION}; output 3: predicted
 * log2(group count) / 32.  Parameters: one flat array of EG_MLP_PARAMS floats,
 * [W1 64x16 | b1 64 | W2 64x64 | b2 64 | W3 4x64 | b3 4], row-major (out, in).
 * The rule of auto_select_groupby_algo (groupby.cu:68-119) stays the fallback:
 * eg_groupby with EG_ALGO_AUTO uses the MLP only when a model is loaded
 * (eg_mlp_load) and its outputs are finite.  The choice never changes a
 * result, only the strategy that computes it. */
#define EG_MLP_FEATURES 16
#define EG_MLP_OUTPUTS 4
#define EG_MLP_PARAMS 5508
typedef struct eg_mlp_query {
    const void* keys;        /* device */
    int64_t rows;
    int64_t group_bound;     /* hint (<= 0 or >= rows: unknown) */
    int32_t key_type;
    int32_t nops;            /* aggregation states of the query */
    int32_t value_bytes;     /* bytes per row of the distinct value columns */
    int32_t reserved;
} eg_mlp_query;
typedef struct eg_mlp_train_config {
    int32_t epochs;
    int32_t batch;           /* multiple of 64 */
    float lr;                /* Adam step size */
    float lambda;            /* weight of the log-cardinality loss */
    float weight_decay;
    uint64_t seed;           /* minibatch shuffling */
} eg_mlp_train_config;
/* out_features: num_queries x EG_MLP_FEATURES floats, device. */
int eg_mlp_features(const eg_mlp_query* queries, int64_t num_queries, float* out_features, void* stream);
/* params, x (n x 16), y (n x 4): device. */
int eg_mlp_forward(const float* params, const float* x, int64_t n, float* y, void* stream);
/* Mean loss over n samples and its gradient (EG_MLP_PARAMS floats, device):
 * cross-entropy of softmax(y[0..2]) against target (n x 3, rows summing to 1)
 * plus lambda * (y[3] - card)^2. */
int eg_mlp_loss_grad(const float* params, const float* x, const float* target, const float* card, int64_t n,
                     float lambda, float* grad, float* loss, void* stream);
/* Adam training of params in place (device) over shuffled minibatches;
 * loss_per_epoch (host, optional) receives each epoch's mean loss. */
int eg_mlp_train(float* params, const float* x, const float* target, const float* card, int64_t n,
                 const eg_mlp_train_config* cfg, float* loss_per_epoch, void* stream);
/* Loads a model for the dispatcher (host array; NULL unloads). */
int eg_mlp_load(const float* host_params, int64_t n);
int eg_mlp_loaded(void);
/* The loaded model's strategy (EG_ALGO_HASH/SORT/PARTITION) and group-count
 * estimate for one key column. */
int eg_mlp_select(eg_column_view key, int64_t group_bound, int32_t nops, int32_t value_bytes, void* stream,
                  int32_t* algo, int64_t* card_est);
/* The loaded model's choice for `num` queries: one feature launch, one
 * forward launch, one synchronisation. */
int eg_mlp_select_batch(const eg_mlp_query* queries, int64_t num, int32_t* algos, int64_t* card_est,
                        void* stream);

/* ---- kernel timing (measurement support; no reference counterpart) -------- */
int eg_kernel_timing_enable(int32_t on);
int eg_kernel_timing_reset(void);
/* Sums the hipEvent-measured durations of every launch of kernel `name`
 * since the last reset (synchronises the recorded events). */
int eg_kernel_timing_query(const char* name, int64_t* launches, double* total_ms);

/* ---- run-time kernel specialisation (no counterpart in the reference: its
 * CUDA build instantiates one kernel per aggregation functor at compile time,
 * groupby_hash.cu:77-99; here the op list is a kernel argument and hipRTC
 * compiles the hash / partition aggregation kernels with the request's ops as
 * constants) ------------------------------------------------------------- */
enum { EG_JIT_OFF = 0, EG_JIT_ASYNC = 1, EG_JIT_SYNC = 2 };
/* mode: EG_JIT_ASYNC (default; compile on a background thread, use the
 * generic kernel until the specialised one is ready), EG_JIT_SYNC (compile on
 * first use), EG_JIT_OFF, or -1 to keep the mode.  min_rows: inputs with fewer
 * rows use the generic kernel (default 2^20; -1 keeps the value).  The
 * environment variables EG_JIT (off|sync|async) and EG_JIT_MIN_ROWS set the
 * initial values. */
int eg_jit_configure(int32_t mode, int64_t min_rows);
/* Blocks until no compilation is queued or running. */
int eg_jit_wait(void);
/* out6 = {kernels compiled, compiles or loads failed, compiles pending,
 *         specialised launches, mode, min_rows}. */
int eg_jit_stats(int64_t* out6);
/* Queues, on the background compiler thread, the hash and partition
 * aggregation kernels specialised for this request list (key type, value
 * column types, requests as eg_groupby lowers them), for `arch` (e.g.
 * "gfx950:sramecc+:xnack-", the device's gcnArchName; NULL = the current
 * device's).  Needs no device when `arch` is given.  Lets a caller warm the
 * specialisation cache before its data arrives; *queued (may be NULL) = the
 * number of new compiles.  A process may exit with compiles queued or
 * running. */
int eg_jit_prefetch(int32_t key_type, const int32_t* val_types, int64_t num_vals,
                    const eg_agg_request* reqs, int64_t num_reqs, const char* arch,
                    int32_t* queued);
/* Diagnostic: compiles a fixed set of specialised kernels (hash and partition,
 * 4- and 8-byte keys, vector and typed-load value layouts) for `arch` (e.g.
 * "gfx950") without a device; the compiler log of failures goes to log. */
int eg_jit_selftest(const char* arch, char* log, int64_t log_len);

/* Reference copy kernel for bandwidth calibration: copies `bytes` between two
 * scratch buffers `iters` times with 16-byte vector loads/stores (kernel name
 * "probe_copy", four workgroups per CU); *ms = the fastest copy.  Mirrors the reference's
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

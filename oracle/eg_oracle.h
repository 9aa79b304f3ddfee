/*
 * eg_oracle.h -- CPU restatement of the reference's single-key group-by path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libeiger_groupby.so,
 * the gamax C++ headers, the eiger_amd Python package) may link, load or call
 * this code.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, and only as the checker / the timed CPU baseline.
 *
 * Every function names the reference file:line whose semantics it restates
 * (paths relative to src/cuda/ of the reference repository).
 */
#ifndef EG_ORACLE_H
#define EG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Type ids: numbering of gamax::type_id (types/data_type.hpp:12-28). */
enum {
    ORC_EMPTY = 0, ORC_INT8 = 1, ORC_INT16 = 2, ORC_INT32 = 3, ORC_INT64 = 4,
    ORC_BOOL8 = 5, ORC_UINT8 = 6, ORC_UINT16 = 7, ORC_UINT32 = 8,
    ORC_UINT64 = 9, ORC_FLOAT32 = 10, ORC_FLOAT64 = 11, ORC_STRING = 12
};

/* Aggregation ids: gamax::groupby::aggregation_id (aggregation.hpp:9-22). */
enum {
    ORC_AGG_SUM = 0, ORC_AGG_MAX = 1, ORC_AGG_MIN = 2, ORC_AGG_AVG = 3,
    ORC_AGG_ANY = 4, ORC_AGG_COUNT_ALL = 5
};

/* Group-by strategies: gamax::groupby::groupby_algo (groupby.hpp:16-25). */
enum {
    ORC_ALGO_AUTO = 0, ORC_ALGO_HASH = 1, ORC_ALGO_SORT = 2,
    ORC_ALGO_SORT_OPT = 3, ORC_ALGO_SORT_GFUR = 4, ORC_ALGO_PARTITION = 5,
    ORC_ALGO_PARTITION_GFUR = 6
};

/* Value distributions of the synthetic generator. */
enum {
    ORC_VAL_UNIFORM_I64 = 0,  /* int64 uniform in [-2^31, 2^31) */
    ORC_VAL_UNIFORM_F32 = 1,  /* float32 uniform in [0, 1), 24-bit grid */
    ORC_VAL_UNIFORM_I32_1000 = 2, /* int32 uniform in [0, 1000) */
    ORC_VAL_ALL_ONES = 3,     /* 1 (groupby_util.cuh ALL_ONES) */
    ORC_VAL_SEQ = 4           /* row index (groupby_util.cuh SEQ) */
};

#define ORC_MAX_COLS 64

typedef struct orc_col {
    const void* data;
    int64_t size;
    int32_t type;
    int32_t pad_;
} orc_col;

typedef struct orc_req {
    int64_t column_index;
    int32_t agg_kind;
    int32_t pad_;
} orc_req;

/* Result table: num_cols columns of num_rows rows, malloc-owned. */
typedef struct orc_result {
    int64_t num_rows;
    int32_t num_cols;
    int32_t types[ORC_MAX_COLS];
    void* cols[ORC_MAX_COLS];
} orc_result;

/* ---- counter-based generator ------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t orc_perm(uint64_t i, uint64_t n, uint64_t seed);
void orc_gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count,
                          int64_t n, int64_t distinct, uint64_t seed);
void orc_zipf_cdf(double* cdf, int64_t distinct, double s);
void orc_gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count,
                       const double* cdf, int64_t distinct, uint64_t seed);
void orc_gen_values(void* out, int32_t type, int32_t dist, int64_t begin,
                    int64_t count, uint64_t seed, uint32_t stream_id);

/* ---- hashing ----------------------------------------------------------- */
uint32_t orc_fmix32(uint32_t x);
uint64_t orc_fmix64(uint64_t x);
void orc_partition_ids(const void* keys, int32_t type, int64_t n, uint32_t mask,
                       int32_t* pids);

/* ---- output typing ----------------------------------------------------- */
int32_t orc_agg_output_type(int32_t agg_kind, int32_t type);

/* ---- group-by ----------------------------------------------------------- */
int orc_groupby(const orc_col* key, const orc_col* vals, int32_t nvals,
                const orc_req* reqs, int32_t nreqs, int32_t agg_count,
                orc_result* res);
int orc_groupby_parallel(const orc_col* key, const orc_col* vals, int32_t nvals,
                         const orc_req* reqs, int32_t nreqs, int32_t agg_count,
                         int32_t nthreads, orc_result* res);
int orc_aggregate(const orc_col* cols, int32_t ncols, const orc_req* reqs,
                  int32_t nreqs, orc_result* res);
void orc_result_free(orc_result* res);
const char* orc_last_error(void);

/* ---- sort / partition sub-operators ------------------------------------ */
int orc_sort_pairs(const void* keys, int32_t key_type, const void* vals,
                   int32_t val_bytes, int64_t n, int32_t begin_bit,
                   int32_t end_bit, int32_t descending, void* out_keys,
                   void* out_vals);
int orc_partition_offsets(const void* keys, int32_t key_type, int64_t n,
                          int32_t begin_bit, int32_t end_bit, int64_t* offsets);
int orc_hash_partition(const void* keys, int32_t key_type, int64_t n,
                       int32_t bits, int32_t shift, int64_t* perm,
                       int64_t* offsets);

/* ---- dispatch ----------------------------------------------------------- */
int32_t orc_rule_select(int32_t key_bytes, int64_t ndv);
int64_t orc_count_distinct(const void* keys, int32_t type, int64_t n);

/* ---- dictionary encoding ------------------------------------------------- */
int orc_dict_encode(const void* keys, const void* dict, int32_t type, int64_t n, int64_t m, int64_t miss,
                    int64_t* out);
int64_t orc_dict_encode_ordered(const void* keys, int32_t type, int64_t n, int64_t* out);

/* ---- checker helpers (OpenMP) --------------------------------------------- */
int orc_argsort_keys(const void* keys, int32_t type, int64_t n, int32_t nthreads, int64_t* perm);
int orc_gather(const void* src, int32_t bytes, const int64_t* idx, int64_t n, void* dst, int32_t nthreads);

/* ---- learned dispatch (eg_oracle_mlp.c): fp64 restatement ---------------- */
#define ORC_MLP_PARAMS 5508
int orc_mlp_forward(const double* params, const float* x, int64_t n, double* y);
int orc_mlp_loss_grad(const double* params, const float* x, const float* target, const float* card, int64_t n,
                      double lambda, double* grad, double* loss);
int orc_sample_features(const void* keys, int32_t key_type, int64_t rows, int64_t group_bound, int32_t nops,
                        int32_t value_bytes, float* out);

#ifdef __cplusplus
}
#endif
#endif

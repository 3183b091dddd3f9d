/* qe_gpu.h — C ABI of the MI355X (gfx950) execution path for the engine's
 * hash table, inner hash join, hash group-by aggregation and backward as-of
 * join.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes, no C++ or
 * torch types. Every call runs on the context's single HIP stream. No
 * exception crosses this boundary: failures return a status and the message
 * is kept per context (qe_gpu_last_error). The C++ operator wrappers
 * (query-engine_amd/host) turn statuses back into the reference's exception
 * types (std::invalid_argument, std::overflow_error, std::bad_alloc).
 *
 * The reference engine has no FFI of its own; each entry point names the
 * reference interface it replaces (file:line in the reference checkout).
 */
#ifndef QE_GPU_H
#define QE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status / types -------------------------------------------------------*/
typedef enum qe_gpu_status {
    QE_GPU_OK = 0,
    QE_GPU_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference   */
    QE_GPU_OUT_OF_MEMORY = 2,    /* std::bad_alloc (core/buffer.cpp:23)       */
    QE_GPU_OVERFLOW = 3,         /* std::overflow_error (ops/hashtable.cpp:64) */
    QE_GPU_HIP_ERROR = 4,
    QE_GPU_UNSUPPORTED = 5
} qe_gpu_status;

/* Same enumerators, same order as qe::Type (core/types.h:18). Widths per
 * byte_width (core/types.h:32-48): I32/STR 4, I64/F64/TS 8, BOOL 1. */
typedef enum qe_type {
    QE_I32 = 0, QE_I64 = 1, QE_F64 = 2, QE_BOOL = 3, QE_TS = 4, QE_STR = 5
} qe_type;

/* A host column view, field-for-field the reference qe::Column
 * (core/column.h:55-67): LSB-first u64 validity words; all_valid is
 * authoritative (core/column.h:42-54) and validity may be NULL when set. */
typedef struct qe_col_desc {
    int32_t type;
    int32_t all_valid;
    uint64_t len;
    const void* data;
    const uint64_t* validity;
} qe_col_desc;

/* Aggregate functions, same order as qe::AggFunc (ops/aggregate.h:61). */
typedef enum qe_agg_func {
    QE_AGG_COUNT_STAR = 0, QE_AGG_COUNT = 1, QE_AGG_SUM = 2,
    QE_AGG_MIN = 3, QE_AGG_MAX = 4, QE_AGG_AVG = 5
} qe_agg_func;

/* qe::AggSpec without the name (ops/aggregate.h:66-94). */
typedef struct qe_agg_spec {
    int32_t func;
    uint32_t input_col; /* ignored for COUNT(*) */
} qe_agg_spec;

typedef enum qe_join_type { QE_JOIN_INNER = 0, QE_JOIN_LEFT = 1 } qe_join_type;
typedef enum qe_null_policy { QE_NULL_EQUAL = 0, QE_NULL_NEVER_MATCH = 1 } qe_null_policy;

/* As-of search method: plain per-probe binary search (the exact answer the
 * learned index must reproduce) or the learned per-key index (spline pieces +
 * monotone MLP router over buckets holding their knots inline, a search
 * outward from the piece's predicted row, a verified knot-range search with a
 * whole-partition fallback for buckets of more than seven knots). */
typedef enum qe_asof_mode { QE_ASOF_BINARY = 0, QE_ASOF_LEARNED = 1, QE_ASOF_LEARNED_NO_MLP = 2 } qe_asof_mode;
/* Or'ed into the mode: report each match as its position in the index's
 * sorted build order (qe_gpu_asof_sorted, qe_gpu_asof_take_sorted) instead of
 * the original build row, which spares the probe one dependent random read.
 * Enough for a caller that only needs which probes matched (an inner as-of
 * feeding probe-side columns only), or that reads build columns it took into
 * sorted order once with qe_gpu_asof_take_sorted. No reference counterpart:
 * the reference's AsofJoin emits build rows. */
#define QE_ASOF_OUT_SORTED 0x100

/* Sentinel for "no match" in match-index outputs: kNoGroup / kNullBuildRow
 * (ops/hashtable.h:58, ops/join_internal.h:31). */
#define QE_GPU_NO_ROW 0xFFFFFFFFu

typedef struct qe_gpu_ctx qe_gpu_ctx;
typedef struct qe_dev_col qe_dev_col;     /* device column (HBM)            */
typedef struct qe_dev_table qe_dev_table; /* device result: rows + columns  */
typedef struct qe_dev_pairs qe_dev_pairs; /* device (probe_row, build_row)  */
typedef struct qe_dev_ht qe_dev_ht;       /* device hash table              */
typedef struct qe_dev_asof qe_dev_asof;   /* sorted as-of build side        */

/* ---- context ----------------------------------------------------------------*/
/* One context per GPU; owns one HIP stream and a caching device allocator. */
qe_gpu_status qe_gpu_ctx_create(int device, qe_gpu_ctx** out);
void qe_gpu_ctx_destroy(qe_gpu_ctx* ctx);
const char* qe_gpu_last_error(const qe_gpu_ctx* ctx);
/* The context's device as the HIP runtime reports it, for result tags
 * (bench/bench_report.h:22-46 records the host the same way). */
typedef struct {
    char name[256];
    char arch[64];               /* e.g. "gfx950:sramecc+:xnack-"            */
    int32_t device;
    int32_t compute_units;
    uint64_t hbm_bytes;
    uint64_t lds_bytes_per_block;
    uint64_t l2_bytes;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_bits;
    int32_t hip_runtime_version;
    int32_t hip_driver_version;
} qe_gpu_device_info_t;
qe_gpu_status qe_gpu_device_info(qe_gpu_ctx* ctx, qe_gpu_device_info_t* out);
qe_gpu_status qe_gpu_sync(qe_gpu_ctx* ctx);
/* Device memory limit for this context (0 = no limit): bytes in live device
 * buffers may not exceed it (cached free blocks are released first when a new
 * allocation would pass it). Lets a test force the OUT_OF_MEMORY path. */
qe_gpu_status qe_gpu_set_mem_limit(qe_gpu_ctx* ctx, uint64_t bytes);
/* Bytes in the context's live device buffers and bytes it holds from the
 * device (live plus cached free blocks). */
qe_gpu_status qe_gpu_mem_usage(qe_gpu_ctx* ctx, uint64_t* live_bytes, uint64_t* held_bytes);
/* Kernel timing: when enabled, the dominant kernel of each call is bracketed by
 * HIP events on the context stream; qe_gpu_timing reads the accumulated time
 * (ms) and launch count per kernel name, then clears it. */
qe_gpu_status qe_gpu_set_timing(qe_gpu_ctx* ctx, int enabled);
qe_gpu_status qe_gpu_timing(qe_gpu_ctx* ctx, const char* kernel, double* ms, uint64_t* launches);
/* Execution-shape knobs (results never depend on them). No counterpart in the
 * reference (its kernels have no launch shape); used by the autotuner and the
 * benchmark. QE_TUNE_JOIN_PAIRS: row pairs per thread per tile of the
 * streaming join probe, 2 or 4 (default 4). QE_TUNE_ASOF_BUCKET_KNOTS:
 * spline knots per learned-router bucket, 1..64 (default 1), read by
 * qe_gpu_asof_train: larger values shrink the radix table.
 * QE_TUNE_JOIN_SPILL_CAP: records per block in the spill lists of a fused
 * join -> group-by whose groups exceed LDS (0 = a quarter of the block's rows;
 * rows past the cap fold with HBM atomics). QE_TUNE_JOIN_BLOCKS_PER_CU:
 * 1024-thread blocks per CU of the fused join's streaming probe, 1..8
 * (default 1), capped by what the LDS partials leave room for.
 * QE_TUNE_ASOF_DATA_LINES: 1 (default) stores the sorted times of a router
 * bucket of at most 28 rows spanning less than 2^32 time units, or 56 rows
 * spanning less than 2^16, inline in the bucket's line, so a learned probe
 * there reads that one line; 0 keeps knot records in every line. Read by
 * qe_gpu_asof_train. */
typedef enum qe_tune_key {
    QE_TUNE_JOIN_PAIRS = 0,
    QE_TUNE_ASOF_BUCKET_KNOTS = 1,
    QE_TUNE_JOIN_SPILL_CAP = 2,
    QE_TUNE_JOIN_BLOCKS_PER_CU = 3,
    QE_TUNE_ASOF_DATA_LINES = 4
} qe_tune_key;
qe_gpu_status qe_gpu_set_tuning(qe_gpu_ctx* ctx, int32_t key, int64_t value);
/* Inspection: the XCD (die) id that ran each of nblocks one-wave blocks of a
 * plain launch (block b lands on XCD b % 8 on MI355X). No counterpart in the
 * reference (it has no device placement). */
qe_gpu_status qe_gpu_xcd_of_blocks(qe_gpu_ctx* ctx, uint32_t nblocks, uint32_t* out);

/* ---- columns ----------------------------------------------------------------*/
/* Host -> HBM once, Batch-compatible layout (same widths, u64 validity words,
 * all_valid rule). Replaces the zero-copy Scan source (ops/scan.cpp:27-64,
 * ops/table.h:24-45): device operators read the uploaded copy. */
qe_gpu_status qe_gpu_column_upload(qe_gpu_ctx* ctx, const qe_col_desc* desc, qe_dev_col** out);
/* Uninitialised device column (validity allocated iff nullable). */
qe_gpu_status qe_gpu_column_alloc(qe_gpu_ctx* ctx, int32_t type, uint64_t len, int32_t nullable, qe_dev_col** out);
/* HBM -> host: data (len*width bytes), validity (ceil(len/64) words, may be
 * NULL), all_valid flag (may be NULL). */
qe_gpu_status qe_gpu_column_download(qe_gpu_ctx* ctx, const qe_dev_col* col, void* data, uint64_t* validity, int32_t* all_valid);
/* Download rows [offset, offset+count) only (the <=2048-row output batches). */
qe_gpu_status qe_gpu_column_download_range(qe_gpu_ctx* ctx, const qe_dev_col* col, uint64_t offset, uint64_t count, void* data, uint8_t* valid_bytes);
void qe_gpu_column_free(qe_gpu_ctx* ctx, qe_dev_col* col);
uint64_t qe_gpu_column_len(const qe_dev_col* col);
int32_t qe_gpu_column_type(const qe_dev_col* col);
int32_t qe_gpu_column_all_valid(const qe_dev_col* col);
/* A view of rows [offset, offset+len) of `col` (no copy; offset % 64 == 0 so
 * the validity sub-view is word aligned, as ops/scan.cpp:15-18 requires). */
qe_gpu_status qe_gpu_column_slice(qe_gpu_ctx* ctx, const qe_dev_col* col, uint64_t offset, uint64_t len, qe_dev_col** out);

/* ---- dictionary-encoded VARCHAR (STR) ------------------------------------------*/
/* A STR column's int32 codes index a string dictionary, which replaces the
 * reference's StringDict (core/string_dict.h:34-72) on the device path: `n`
 * strings concatenated in `bytes`, string i = bytes[offsets[i], offsets[i+1])
 * (its bytes_/offsets_ layout, core/string_dict.h:21-23). The library copies
 * it. STR values compare by their bytes, never by code (core/string_dict.h:
 * 9-14): group keys and join / as-of keys match by value (ops/join.cpp:48-71,
 * ops/aggregate.cpp:164-187), MIN/MAX and sort keys order by byte value
 * (ops/aggregate.cpp:189-205, ops/sort.cpp:82-86). STR outputs (group keys,
 * MIN/MAX, gathered / filtered / sliced columns) carry the dictionary of the
 * column they come from; a dictionary must outlive every column and as-of
 * index that refers to it. */
typedef struct qe_gpu_dict qe_gpu_dict;
qe_gpu_status qe_gpu_dict_create(qe_gpu_ctx* ctx, const char* bytes, const uint32_t* offsets, uint32_t n,
                                 qe_gpu_dict** out);
void qe_gpu_dict_free(qe_gpu_ctx* ctx, qe_gpu_dict* d);
uint32_t qe_gpu_dict_size(const qe_gpu_dict* d);
/* Attach (or, with NULL, detach) the dictionary of a STR column. */
qe_gpu_status qe_gpu_column_set_dict(qe_gpu_ctx* ctx, qe_dev_col* col, const qe_gpu_dict* d);
const qe_gpu_dict* qe_gpu_column_dict(const qe_dev_col* col);
/* out[r] = table[col[r]] as out_type (I32 or STR, no dictionary attached);
 * NULL rows stay NULL; a valid code outside [0, table_len) is
 * QE_GPU_INVALID_ARGUMENT. For callers re-encoding codes between dictionaries
 * (the value-id canonicalisation of ops/join.cpp:48-71). */
qe_gpu_status qe_gpu_column_remap(qe_gpu_ctx* ctx, const qe_dev_col* col, const int32_t* table, uint32_t table_len,
                                  int32_t out_type, qe_dev_col** out);

/* ---- result tables ------------------------------------------------------------*/
uint64_t qe_gpu_table_rows(const qe_dev_table* t);
uint32_t qe_gpu_table_cols(const qe_dev_table* t);
const qe_dev_col* qe_gpu_table_col(const qe_dev_table* t, uint32_t i);
void qe_gpu_table_free(qe_gpu_ctx* ctx, qe_dev_table* t);

/* ---- hash group-by ------------------------------------------------------------*/
/* Replaces qe::Aggregate(child, key_cols, aggs) (ops/aggregate.h:110-111;
 * drain_and_build ops/aggregate.cpp:243-437, next :439-493). Output columns =
 * key columns (input types) ++ one column per agg with the reference result
 * type (ops/aggregate.cpp:78-97). One row per group, unordered; zero keys ->
 * exactly one row even over empty input; keyed over empty -> zero rows.
 * F64 SUM/AVG accumulate in FP64 (summation order is the only difference
 * from the reference's sequential fold). */
qe_gpu_status qe_gpu_aggregate(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                               const uint32_t* key_cols, uint32_t nkeys,
                               const qe_agg_spec* aggs, uint32_t naggs, qe_dev_table** out);

/* ---- filter ---------------------------------------------------------------------*/
/* Comparison operators, same order as expr::CmpOp (expr/expr.h:79). */
typedef enum qe_cmp_op {
    QE_CMP_LT = 0, QE_CMP_LE = 1, QE_CMP_GT = 2, QE_CMP_GE = 3, QE_CMP_EQ = 4, QE_CMP_NE = 5
} qe_cmp_op;
/* One term `column op literal` (expr::cmp(op, col(t, i), lit(s))). The literal
 * is typed like expr::Scalar (expr/expr.h:84-104): lit_i carries I32/I64/TS/
 * BOOL, lit_f carries F64, lit_is_null marks a typed NULL. */
typedef struct qe_pred_term {
    uint32_t col;
    int32_t op;
    int32_t lit_type;
    int32_t lit_is_null;
    int64_t lit_i;
    double lit_f;
} qe_pred_term;
/* Replaces qe::Filter(child, predicate) (ops/filter.h, ops/filter.cpp:28-66)
 * for a predicate that is a conjunction (expr::logic_and) of such terms: rows
 * where every term is TRUE pass (a NULL comparison does not), operands are
 * promoted as expr::cmp promotes them, and every column is compacted with its
 * validity, passing rows in input order (compact_column,
 * core/owned_batch.cpp:101-166). */
qe_gpu_status qe_gpu_filter(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                            const qe_pred_term* terms, uint32_t nterms, qe_dev_table** out);
/* The same predicate, returning the passing row indices (a device selection
 * vector, ascending) in sel_out (device u32[len]) and their count. */
qe_gpu_status qe_gpu_filter_select(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                   const qe_pred_term* terms, uint32_t nterms, uint32_t* sel_out,
                                   uint64_t* count_out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc): the same
 * selection as qe_gpu_filter_select, identical output, with no device
 * allocation, no free and no allocating library call. Its working memory is a
 * static 16.4 MiB device array per GPU (held until the process ends), so it
 * takes at most 2^27 rows (QE_GPU_UNSUPPORTED beyond). Every call on one GPU
 * shares that array: calls that overlap on one GPU, from any context or
 * thread, race. It synchronises the context stream (the count is read back),
 * so it cannot be captured into a HIP graph. */
qe_gpu_status qe_gpu_filter_select_noalloc(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols,
                                           const qe_pred_term* terms, uint32_t nterms, uint32_t* sel_out,
                                           uint64_t* count_out);

/* ---- hash join ----------------------------------------------------------------*/
/* Replaces qe::HashJoin(probe, build, probe_keys, build_keys, type)
 * (ops/join.h:79-81; build_side ops/join.cpp:133-197, build_pairs_for_probe
 * :199-247). Produces every (probe_row, build_row) pair; LEFT unmatched rows
 * carry build_row = QE_GPU_NO_ROW. NULL keys never match; F64 keys use the
 * canonical -0/+0, NaN==NaN equality. Key types must match positionally. */
qe_gpu_status qe_gpu_hash_join(qe_gpu_ctx* ctx,
                               const qe_dev_col* const* probe_keys, const qe_dev_col* const* build_keys,
                               uint32_t nkeys, int32_t join_type, qe_dev_pairs** out);
uint64_t qe_gpu_pairs_count(const qe_dev_pairs* p);
/* Copy pairs [offset, offset+count) to host arrays (either may be NULL). */
qe_gpu_status qe_gpu_pairs_download(qe_gpu_ctx* ctx, const qe_dev_pairs* p, uint64_t offset, uint64_t count,
                                    uint32_t* probe_rows, uint32_t* build_rows);
/* Materialise output rows [offset, offset+count) of a join: probe columns
 * gathered by probe_row, then build columns by build_row (NULL where
 * build_row == QE_GPU_NO_ROW). Mirrors emit_probe_column/emit_build_column
 * (ops/join_internal.h:136-179). */
qe_gpu_status qe_gpu_pairs_gather(qe_gpu_ctx* ctx, const qe_dev_pairs* p, uint64_t offset, uint64_t count,
                                  const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                  const qe_dev_col* const* build_cols, uint32_t nbuild, qe_dev_table** out);
void qe_gpu_pairs_free(qe_gpu_ctx* ctx, qe_dev_pairs* p);

/* Fused Aggregate(HashJoin(probe, build, ..., type), group_keys, aggs) — the
 * plan plan/plan.cpp:581-590 lowers for scan(P).join(scan(B)).aggregate(...).
 * group_keys / agg input_col index the join output schema (probe columns then
 * build columns, ops/join.h:22-28). Output as qe_gpu_aggregate. */
qe_gpu_status qe_gpu_join_aggregate(qe_gpu_ctx* ctx,
                                    const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                    const qe_dev_col* const* build_cols, uint32_t nbuild,
                                    const uint32_t* probe_keys, const uint32_t* build_keys, uint32_t nkeys,
                                    int32_t join_type, const uint32_t* group_keys, uint32_t ngroup_keys,
                                    const qe_agg_spec* aggs, uint32_t naggs, qe_dev_table** out);
/* Autotuner for the fused join -> group-by: runs this query once per launch
 * shape (row pairs per thread 2/4 x blocks per CU 1/2/4), times the streaming
 * probe with HIP events (best of two runs), leaves the fastest shape set on
 * the context (QE_TUNE_JOIN_PAIRS, QE_TUNE_JOIN_BLOCKS_PER_CU) and reports it.
 * The caller persists the choice (the benchmark keeps it in a tuning file).
 * No counterpart in the reference. */
typedef struct qe_join_tuning {
    int32_t pairs;
    int32_t blocks_per_cu;
    double probe_ms;       /* probe kernel time of the chosen shape       */
    double default_ms;     /* the same with the default shape (4 pairs, 1 block per CU) */
    int32_t candidates;    /* shapes timed                                */
} qe_join_tuning;
qe_gpu_status qe_gpu_join_autotune(qe_gpu_ctx* ctx, const qe_dev_col* const* probe_cols, uint32_t nprobe,
                                   const qe_dev_col* const* build_cols, uint32_t nbuild, const uint32_t* probe_keys,
                                   const uint32_t* build_keys, uint32_t nkeys, int32_t join_type,
                                   const uint32_t* group_keys, uint32_t ngk, const qe_agg_spec* aggs, uint32_t naggs,
                                   qe_join_tuning* out);

/* ---- hash table -----------------------------------------------------------------*/
/* Replaces qe::HashTable (ops/hashtable.h:108-190): key -> dense group id in
 * first-seen row order (ids stable across calls and growth), kEqual /
 * kNeverMatch NULL policies (kNeverMatch NULL rows get fresh dead ids), key
 * store read-back. initial_capacity / max_load_factor follow HashTableConfig
 * (ops/hashtable.h:83-94); the seed is accepted for interface parity. */
qe_gpu_status qe_gpu_ht_create(qe_gpu_ctx* ctx, const int32_t* key_types, uint32_t nkeys, int32_t null_policy,
                               uint64_t initial_capacity, double max_load_factor, uint64_t seed, qe_dev_ht** out);
/* insert_or_find (ops/hashtable.cpp:93-133) over rows sel[0..n) (or 0..n when
 * sel is NULL; sel is a device u32 array). out_groups: device u32[n]. */
qe_gpu_status qe_gpu_ht_insert_or_find(qe_gpu_ctx* ctx, qe_dev_ht* ht, const qe_dev_col* const* keys,
                                       const uint32_t* sel, uint64_t n, uint32_t* out_groups);
/* find (ops/hashtable.cpp:135-163): QE_GPU_NO_ROW on a miss. Never mutates. */
qe_gpu_status qe_gpu_ht_find(qe_gpu_ctx* ctx, const qe_dev_ht* ht, const qe_dev_col* const* keys,
                             const uint32_t* sel, uint64_t n, uint32_t* out_groups);
uint64_t qe_gpu_ht_num_groups(const qe_dev_ht* ht);
uint64_t qe_gpu_ht_capacity(const qe_dev_ht* ht);
/* group_key_word / group_is_null (ops/hashtable.h:143-155) for groups
 * [0, num_groups): host arrays words[num_groups], nulls[num_groups]. */
qe_gpu_status qe_gpu_ht_key_words(qe_gpu_ctx* ctx, const qe_dev_ht* ht, uint32_t col, uint64_t* words, uint8_t* nulls);
void qe_gpu_ht_free(qe_gpu_ctx* ctx, qe_dev_ht* ht);

/* ---- sort ---------------------------------------------------------------------------*/
/* Stable sort permutation, keys[0] most significant, per-key desc /
 * nulls_first flags (either array may be NULL = ASC / NULLS LAST). Replaces
 * Sort (ops/sort.h:84): integer-family keys as its radix path
 * (ops/sort.cpp:125-188); F64 keys in the NaN-greatest total order and STR keys
 * by value, as its comparison path (ops/sort.cpp:60-100). perm_out: device
 * u32[n]. */
qe_gpu_status qe_gpu_sort_perm(qe_gpu_ctx* ctx, const qe_dev_col* const* keys, uint32_t nkeys, const int32_t* desc,
                               const int32_t* nulls_first, uint32_t* perm_out);

/* Rows idx[0..m) of `cols` (device u32 indices, each < the columns' length)
 * as a new table, validity and dictionaries kept: Sort's whole-row gather by
 * its permutation (gather_rows, ops/sort_internal.cpp:95-151). */
qe_gpu_status qe_gpu_take(qe_gpu_ctx* ctx, const qe_dev_col* const* cols, uint32_t ncols, const uint32_t* idx,
                          uint64_t m, qe_dev_table** out);

/* ---- as-of join -------------------------------------------------------------------*/
/* Build side of tsx::AsofJoin (tsx/asof.h:101-105; build_side
 * tsx/asof.cpp:177-232): partition by key tuple (NULL keys never match),
 * drop NULL timestamps, sort each partition by timestamp on the device
 * (the Sort wrap tsx/asof.cpp:163-166), CSR offsets per partition. Time
 * column I32/I64/TS, key types as the probe's (checked at probe). Zero keys =
 * one global partition. */
qe_gpu_status qe_gpu_asof_build(qe_gpu_ctx* ctx, const qe_dev_col* const* build_keys, uint32_t nkeys,
                                const qe_dev_col* build_time, qe_dev_asof** out);
/* Train the learned index on the sorted build side: per partition, spline
 * pieces of `piece_rows` rows with recorded max error, and a monotone
 * 1-16-16-1 MLP router trained on the device (forward/backward kernels,
 * Adam) for `steps` steps. */
qe_gpu_status qe_gpu_asof_train(qe_gpu_ctx* ctx, qe_dev_asof* idx, uint32_t piece_rows, uint32_t steps, float lr);
/* Probe (build_pairs_for_probe, tsx/asof.cpp:234-308): for each probe row the
 * build row with equal key and the greatest tb <= t (original build row
 * index), or QE_GPU_NO_ROW. has_tolerance: drop if (u64)t-(u64)tb > tol.
 * match_out: device u32[n]. */
qe_gpu_status qe_gpu_asof_probe(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* probe_keys,
                                uint32_t nkeys, const qe_dev_col* probe_time, int32_t has_tolerance,
                                int64_t tolerance, int32_t mode, uint32_t* match_out);
/* Turn per-probe matches into join pairs: INNER keeps matches, LEFT keeps
 * every probe row (build_row QE_GPU_NO_ROW when unmatched). */
qe_gpu_status qe_gpu_asof_pairs(qe_gpu_ctx* ctx, const uint32_t* match, uint64_t n, int32_t join_type,
                                qe_dev_pairs** out);
typedef struct qe_asof_stats {
    uint64_t build_rows;       /* matchable rows (non-NULL key and ts)           */
    uint64_t partitions;
    uint64_t pieces;           /* spline pieces                                  */
    uint64_t max_piece_error;  /* max recorded |predicted - actual| position     */
    uint64_t router_buckets;   /* radix buckets addressed by the router          */
    double mean_bucket_knots;  /* mean knots searched per probe bucket (learned router as probed) */
    double mean_bucket_knots_linear; /* same with the linear (no-MLP) router everywhere */
    uint64_t fallbacks;        /* probes whose piece was not in its bucket's knot range and was
                                  searched over the whole partition (last probe) */
    double final_loss;         /* mean router training loss over trained partitions */
    uint64_t mlp_candidates;   /* partitions with a trained router               */
    uint64_t mlp_partitions;   /* partitions whose learned probe uses the MLP    */
    double mean_bucket_knots_mlp;  /* same with the trained MLP router in every trained partition */
    /* The candidate router (MLP where it searches fewer knots) and the linear
     * one timed on a sample of up to 2^21 build rows, in ms; the MLP is kept
     * only if at least 2% faster (0 when no partition was a candidate). */
    double sample_probe_ms_mlp;
    double sample_probe_ms_linear;
    uint64_t data_line_buckets;  /* router buckets whose rows sit inline in their line (learned router as probed) */
} qe_asof_stats;
qe_gpu_status qe_gpu_asof_stats(qe_gpu_ctx* ctx, const qe_dev_asof* idx, qe_asof_stats* out);
/* Build-side columns (of the build side's length) gathered into the index's
 * sorted order: row i of the result is the build row at sorted position i,
 * the rows a QE_ASOF_OUT_SORTED match addresses. */
qe_gpu_status qe_gpu_asof_take_sorted(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* cols,
                                      uint32_t ncols, qe_dev_table** out);
/* Probe-side ordering before the probe. It replaces the Sort that
 * tsx::AsofJoin wraps around its probe child (tsx/asof.cpp:161-164), which
 * orders the probes by (keys..., time) so that neighbouring probes read
 * neighbouring build rows.
 * Here one partitioning pass reorders the rows of `cols` (each of
 * probe_time's length; validity and dictionaries kept) into 256 groups by the
 * index region each row's probe will read:
 *   - the learned router's bucket line, for a trained index;
 *   - else the partition's first sorted row.
 * Within a group, rows keep their input order, so the result is
 * deterministic. Rows that probe nothing (NULL key or time, no partition)
 * join the first group.
 * Probing the returned key and time columns gives the same match for each
 * row as probing the input. Each group's random reads stay within 1/256 of
 * the index. */
qe_gpu_status qe_gpu_asof_cluster(qe_gpu_ctx* ctx, const qe_dev_asof* idx, const qe_dev_col* const* probe_keys,
                                  uint32_t nkeys, const qe_dev_col* probe_time, const qe_dev_col* const* cols,
                                  uint32_t ncols, qe_dev_table** out);
/* Sorted build side for inspection: sorted timestamps, original row of each
 * sorted position, and CSR partition offsets (partitions+1). Any may be NULL. */
qe_gpu_status qe_gpu_asof_sorted(qe_gpu_ctx* ctx, const qe_dev_asof* idx, int64_t* ts, uint32_t* rows, uint64_t* offsets);
void qe_gpu_asof_free(qe_gpu_ctx* ctx, qe_dev_asof* idx);

/* Learned-index router internals, exposed so tests can check the device
 * forward pass and the MFMA gradient kernel against a CPU implementation.
 * Parameter layout (QE_MLP_PARAMS floats): w1[16] b1[16] W2[16][16] (row =
 * output unit) b2[16] w3[16] b3. No counterpart in the reference (the learned
 * index is new). */
#define QE_MLP_PARAMS 321
qe_gpu_status qe_gpu_mlp_forward(qe_gpu_ctx* ctx, const float* params, const float* x, uint64_t n, float* y);
/* Gradient of the batch loss (1/n) sum (y(x_i) - c_i)^2 w.r.t. every parameter,
 * and that loss; host arrays, n in [1, 8192]. */
qe_gpu_status qe_gpu_mlp_grad(qe_gpu_ctx* ctx, const float* params, const float* x, const float* c, uint32_t n,
                              float* grads, float* loss);

/* ---- batched group-by (many independent queries over uneven morsels) -------------*/
/* One GROUP BY key query of the batch: its key and value columns, the row
 * offsets of its morsels (host array, nmorsels+1 entries, the morsels of
 * exec/morsel_scan.cpp:39-76 laid end to end), its aggregates (input_col 0 =
 * key, 1 = value) and an optional upper bound on distinct groups (0 = rows).
 * Replaces running one Aggregate per query through the morsel-parallel driver
 * (exec/parallel.cpp:431-599). */
typedef struct qe_batch_query {
    const qe_dev_col* key;
    const qe_dev_col* value;
    uint32_t nmorsels;
    const uint64_t* morsel_offsets;
    uint32_t naggs;
    const qe_agg_spec* aggs;
    uint64_t group_bound;
} qe_batch_query;
/* Execution shape: rows per work tile, persistent blocks per CU, LDS slot
 * cache entries per block (0 or a power of two, at most 512: 80 bytes each
 * within a 64 KiB block budget). */
typedef struct qe_batch_tuning {
    uint32_t tile_rows;
    uint32_t blocks_per_cu;
    uint32_t lds_slots;
} qe_batch_tuning;
typedef enum qe_batch_mode {
    QE_BATCH_PERSISTENT = 0, /* one persistent launch over a device tile queue      */
    QE_BATCH_PER_BATCH = 1,  /* one launch per 2048-row batch (the baseline shape)   */
    QE_BATCH_GRAPH = 2       /* the persistent pipeline captured once in a hipGraph  */
} qe_batch_mode;
typedef struct qe_batch_plan qe_batch_plan;
qe_gpu_status qe_gpu_batch_prepare(qe_gpu_ctx* ctx, const qe_batch_query* queries, uint32_t nq, qe_batch_plan** out);
/* Runs every query; outs (nq tables, or NULL to only execute) receive key ++
 * aggregate columns per query, as qe_gpu_aggregate. tuning NULL = defaults. */
qe_gpu_status qe_gpu_batch_run(qe_gpu_ctx* ctx, qe_batch_plan* plan, const qe_batch_query* queries, int32_t mode,
                               const qe_batch_tuning* tuning, qe_dev_table** outs);
/* Times tile size x blocks per CU x LDS cache on this batch; returns the fastest. */
qe_gpu_status qe_gpu_batch_autotune(qe_gpu_ctx* ctx, qe_batch_plan* plan, qe_batch_tuning* out);
/* The strategy prepare chose for each query (kinds[nq]: QE_BATCH_SMALL, _PART,
 * _PART_PASS or _HASH) and its partition count (nparts[nq], 0 unless
 * partitioned; may be NULL). Host-side only, no device work. No reference
 * counterpart: the reference's per-query driver has one strategy (a hash map
 * per thread, ops/aggregate.cpp); this reports which of the device's the
 * batch uses. */
typedef enum qe_batch_kind {
    QE_BATCH_SMALL = 0,     /* dense LDS table per tile                                  */
    QE_BATCH_PART = 1,      /* dense range: histogram, scatter by partition, LDS per part */
    QE_BATCH_HASH = 2,      /* device hash table                                          */
    QE_BATCH_PART_PASS = 3  /* dense range of <= 2 partitions, folded from the columns    */
} qe_batch_kind;
qe_gpu_status qe_gpu_batch_strategies(qe_gpu_ctx* ctx, const qe_batch_plan* plan, int32_t* kinds, uint32_t* nparts);
void qe_gpu_batch_free(qe_gpu_ctx* ctx, qe_batch_plan* plan);
/* Uniform values in [lo, hi) (integers floored) with Bernoulli(null_frac)
 * NULLs; a pure function of (seed, stream, row). */
qe_gpu_status qe_gpu_gen_uniform(qe_gpu_ctx* ctx, uint64_t seed, uint64_t stream, qe_dev_col* col, double lo, double hi,
                                 double null_frac);

/* ---- device memory helpers for callers without their own allocator ---------------*/
qe_gpu_status qe_gpu_malloc(qe_gpu_ctx* ctx, uint64_t bytes, void** out);
qe_gpu_status qe_gpu_free(qe_gpu_ctx* ctx, void* p);
qe_gpu_status qe_gpu_memcpy_d2h(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
qe_gpu_status qe_gpu_memcpy_h2d(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Plain streaming copy kernel (16-byte loads/stores, bytes % 16 == 0): the
 * measured HBM ceiling and the calibration for PMC byte counters. Returns the
 * kernel time in ms (HIP events on the context stream). No counterpart in the
 * reference (measurement support). */
qe_gpu_status qe_gpu_stream_copy(qe_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes, double* ms);
/* Read-only streaming kernel over src (bytes % 16 == 0): the measured ceiling
 * for kernels that mostly read. sink is a 1 KiB device buffer it practically
 * never writes. Returns the kernel time in ms. No counterpart in the reference
 * (measurement support). */
qe_gpu_status qe_gpu_stream_read(qe_gpu_ctx* ctx, const void* src, uint64_t bytes, void* sink, double* ms);
/* The fused join probe's map lookups alone: reads an I64 key column (data
 * 16-byte aligned; validity ignored) and gathers map[clamp(key - kmin, 0, range - 1)] from a 16-bit map,
 * with no payload and no aggregation. Over the probe's keys and a map of the
 * probe's size its time is the floor of that probe. sink: a 1 KiB device
 * buffer it practically never writes. No counterpart in the reference
 * (measurement support). */
qe_gpu_status qe_gpu_map_gather(qe_gpu_ctx* ctx, const qe_dev_col* key, int64_t kmin, uint64_t range,
                                const uint16_t* map, void* sink, double* ms);

/* ---- synthetic workload generators (bench / tests; seeded, counter-based) ---------*/
/* Config shapes from the benchmark definitions; every value is a pure
 * function of (seed, row), generated directly in HBM. */
qe_gpu_status qe_gpu_gen_groupby(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint64_t key_range,
                                 double null_frac, qe_dev_col* key_i64, qe_dev_col* val_f64);
qe_gpu_status qe_gpu_gen_join_build(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint32_t groups,
                                    qe_dev_col* key_i64, qe_dev_col* grp_i32);
qe_gpu_status qe_gpu_gen_join_probe(qe_gpu_ctx* ctx, uint64_t seed, uint64_t n, uint64_t build_n, double zipf_s,
                                    double miss_frac, qe_dev_col* key_i64, qe_dev_col* pay_f64);
/* As-of tick shape: sym Zipf(zipf_s) over nsym, per-sym strictly increasing
 * microsecond times. sessions == 0: exponential gaps (mean 1 ms) with 10x
 * bursts in 512-row regimes; sessions > 0: that many trading days with a
 * U-shaped intraday intensity and overnight gaps. Left: Zipf syms, times
 * uniform over the sym's right-side span +-5%. */
qe_gpu_status qe_gpu_gen_asof(qe_gpu_ctx* ctx, uint64_t seed, uint64_t right_n, uint64_t left_n, uint32_t nsym,
                              double zipf_s, uint32_t sessions, qe_dev_col* r_sym, qe_dev_col* r_ts, qe_dev_col* r_px,
                              qe_dev_col* l_sym, qe_dev_col* l_ts);

#ifdef __cplusplus
}
#endif
#endif /* QE_GPU_H */

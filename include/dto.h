/*
 * dto.h — C ABI of the MI355X (gfx950) Dual Threshold Optimization library.
 *
 * The reference crate (BrentLab/Dual_Threshold_Optimization v2.0.1) has no FFI;
 * its boundary is the Rust library API plus the CLI.  Each entry point below
 * names the reference item it replaces (file:line in the reference tree).
 * Conventions: plain pointers and sizes, caller-owned output buffers, int status
 * codes (DTO_OK = 0) instead of panics, a thread-local message behind
 * dto_last_error().  Every compute entry point runs on the GPU (HIP, gfx950)
 * and fails with DTO_ERR_NO_DEVICE when none is present; there is no CPU
 * fallback.
 *
 * Permutation convention (src/collections/permuted.rs:56-60, 90-101): a
 * permutation of a list of length L is an index array idx[0..L) over the
 * list's rank-sorted positions; permuted position k holds the feature that sits
 * at sorted position idx[k], while ranks stay at their positions.
 */
#ifndef DTO_H
#define DTO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DTO_OK 0
#define DTO_ERR_INVALID_ARGUMENT 1 /* null pointer, bad size, bad flag */
#define DTO_ERR_POPULATION 2       /* compute_population_size would panic */
#define DTO_ERR_HYPERGEOMETRIC 3   /* Hypergeometric::new fails (K > N or n > N) */
#define DTO_ERR_EMPTY_GRID 4       /* a list has no thresholds: optimize would panic */
#define DTO_ERR_NO_DEVICE 5        /* no HIP device */
#define DTO_ERR_DEVICE 6           /* HIP runtime error */
#define DTO_ERR_BAD_PERMUTATION 7  /* an explicit permutation index is out of range */
#define DTO_ERR_TOO_LARGE 8        /* list or threshold count beyond the kernels' limits */
#define DTO_ERR_IO 9               /* file cannot be opened / read */
#define DTO_ERR_PARSE 10           /* malformed CSV line or rank */
#define DTO_ERR_BUSY 11            /* serving queue full */

/* job flags */
#define DTO_DROP_OUTSIDE_BACKGROUND 0x1u /* drop list features absent from the
                                            background instead of failing */
#define DTO_DEFER_UPLOAD 0x2u /* build the job's tables on the host only: a
                                 server copies them into its HBM pool when it
                                 admits the job; any other run uploads them
                                 on first use */

/* run modes */
#define DTO_MODE_SCREENED 0 /* FP32 log-space screen + exact FP64 recheck */
#define DTO_MODE_PLAIN 1    /* exact FP64 p-value on every grid cell */

/* Limits of the device kernels (checked at job creation). */
#define DTO_MAX_LIST_LEN 65535u /* and the LDS check: lists up to about 61,000 entries fit */
#define DTO_MAX_THRESHOLDS 4096u

/* One optimisation record: OptimizationResultRecord without the debug
 * feature sets (src/dto/results_objects.rs:21-32). rank1/rank2 are threshold
 * values; idx1/idx2 their positions in the threshold lists.  ties (records
 * from optimize only; 0 in grid records): bits 0-15 count the grid cells at
 * the minimum p, bits 16-31 those of them at the largest intersection, each
 * saturated at 0xFFFF; optimize warns on stderr when either is above 1
 * (src/dto/optimize_main.rs:86-108, see DTO_TIES_AT_MIN_P). */
typedef struct {
  uint64_t rank1, rank2;
  uint64_t set1_len, set2_len;
  uint64_t population_size;
  uint64_t intersection_size;
  double pvalue;
  int32_t permuted;
  uint32_t idx1, idx2;
  uint32_t ties;
} dto_record;
#define DTO_TIES_AT_MIN_P(r) ((r)->ties & 0xFFFFu)
#define DTO_TIES_AT_MAX_K(r) ((r)->ties >> 16)

/* Result of a whole DTO run: the fields of the JSON object built by
 * empirical_pvalue (src/stat_operations/empirical_pvalue.rs:136-186). */
typedef struct {
  dto_record unpermuted;
  double empirical_pvalue;
  double fdr;
  uint64_t n_permutations;
  uint64_t n_le; /* #{permuted p <= unpermuted p} */
} dto_result;

typedef struct {
  uint64_t population_size; /* N */
  uint64_t len1, len2;      /* list lengths after any background drop */
  uint32_t n_thresholds1, n_thresholds2;
  uint64_t dropped1, dropped2; /* features dropped (DTO_DROP_OUTSIDE_BACKGROUND) */
  uint64_t device_bytes;       /* resident device footprint of the job */
} dto_job_info;

typedef struct dto_job dto_job;

/* Message for the last failing call on this thread ("" if none). */
const char *dto_last_error(void);
/* Number of visible HIP devices (0 when none). */
int dto_device_count(void);

/* ---- scalar statistics ------------------------------------------------- */
/* hypergeometric_pvalue (src/stat_operations/hypergeometric_pvalue.rs:33-50)
 * for `count` quadruples, evaluated on the GPU. */
int dto_hypergeometric_pvalue(const uint64_t *N, const uint64_t *K, const uint64_t *n,
                              const uint64_t *k, uint64_t count, double *out);
/* fdr (src/stat_operations/fdr.rs:29-60); closed form, host.  Where the
 * reference panics (sensitivity <= 0, fdr.rs:36-39) it returns NaN and sets
 * dto_last_error(); dto_fdr_checked reports that case as
 * DTO_ERR_INVALID_ARGUMENT instead. */
double dto_fdr(uint64_t set1_len, uint64_t set2_len, uint64_t overlap, uint64_t population,
               double sensitivity);
int dto_fdr_checked(uint64_t set1_len, uint64_t set2_len, uint64_t overlap, uint64_t population,
                    double sensitivity, double *out);
/* generate_thresholds (src/collections/ranked.rs:359-375): writes up to cap
 * values into out (may be NULL) and returns how many the reference makes. */
uint32_t dto_generate_thresholds(uint32_t max_rank, uint32_t *out, uint32_t cap);
/* The ln n! table the kernels use: statrs factorial::ln_factorial for
 * x = 0..n_max (replaces the per-call evaluation inside statrs). */
int dto_ln_factorial_table(uint64_t n_max, double *out);

/* ---- jobs (upload once, run many) --------------------------------------- */
/* Builds a job from two ranked lists and an optional background (bg = NULL:
 * none).  Performs RankedFeatureList::from (stable sort + thresholds,
 * src/collections/ranked.rs:176-191), compute_population_size
 * (src/dto/compute_population_size.rs:66-104), the list1/list2 feature join,
 * threshold bins, set sizes and the ln n! tables, and uploads them to the
 * current device. */
int dto_job_create(const char *const *ids1, const uint32_t *ranks1, uint64_t n1,
                   const char *const *ids2, const uint32_t *ranks2, uint64_t n2,
                   const char *const *bg, uint64_t nbg, uint32_t flags, dto_job **out);
/* The library-level form used by optimize / process_threshold_pairs
 * (src/dto/optimize_main.rs:53-59): the caller passes the population size
 * and no population check runs; lists are still stable-sorted and get
 * reference thresholds. */
int dto_job_create_with_population(const char *const *ids1, const uint32_t *ranks1, uint64_t n1,
                                   const char *const *ids2, const uint32_t *ranks2, uint64_t n2,
                                   uint64_t population_size, dto_job **out);
/* Same, reading the reference's file formats: CSV "feature,rank" without
 * header (src/read/read_ranked_feature_list_from_csv.rs:50-70) and one feature
 * per line (src/read/read_feature_list_from_file.rs:45-53). bg_path may be NULL. */
int dto_job_create_from_files(const char *path1, const char *path2, const char *bg_path,
                              uint32_t flags, dto_job **out);
/* Job on explicit threshold lists (increasing values), replacing
 * generate_thresholds; used for dense custom grids. */
int dto_job_set_thresholds(dto_job *job, const uint32_t *thr1, uint32_t t1, const uint32_t *thr2,
                           uint32_t t2);
int dto_job_info_get(const dto_job *job, dto_job_info *info);
/* Threshold values of list 1 or 2 (RankedFeatureList::thresholds,
 * ranked.rs:261-263): writes up to cap values into out (may be NULL) and
 * returns how many the list has, or minus a status code on error. */
int dto_job_thresholds(const dto_job *job, int list, uint32_t *out, uint32_t cap);
void dto_job_destroy(dto_job *job);

/* ---- hot path ------------------------------------------------------------ */
/* optimize(l1, l2, permute, N, debug=false) (src/dto/optimize_main.rs:53-118)
 * for one permutation; idx1/idx2 = NULL means that list is unpermuted.  Host
 * index arrays of len1/len2 entries. */
int dto_optimize(dto_job *job, const uint32_t *idx1, const uint32_t *idx2, uint32_t mode,
                 dto_record *out);
/* process_threshold_pairs (src/dto/process_threshold_pairs.rs:71-131) for one
 * permutation: all T1*T2 records, row-major over (threshold1, threshold2). */
int dto_process_threshold_pairs(dto_job *job, const uint32_t *idx1, const uint32_t *idx2,
                                dto_record *out);
/* No counterpart in the reference: dto_process_threshold_pairs for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing:
 * d_idx1/d_idx2 (device arrays of len1/len2 entries, or NULL for unpermuted)
 * and d_out (a device array of T1*T2 records) are the caller's, and the
 * overlap counts go to a fixed device array of DTO_NOALLOC_MAX_CELLS words
 * that each GPU holds for the life of the process.  Grids larger than that
 * fail with DTO_ERR_TOO_LARGE.  Asynchronous on stream (a hipStream_t, NULL =
 * default); d_status (device int, may be NULL) is cleared and then set
 * nonzero if a permutation index is out of range.  Every call on one GPU
 * shares that array: calls that overlap in time on one GPU race. */
#define DTO_NOALLOC_MAX_CELLS (1u << 22) /* 16 MiB: a 2048 x 2048 threshold grid */
int dto_process_threshold_pairs_noalloc(dto_job *job, const uint32_t *d_idx1, const uint32_t *d_idx2,
                                        dto_record *d_out, int *d_status, void *stream);
/* The CLI's run (src/main.rs:122-126, 156-165): task 0 unpermuted plus
 * n_perm permutations, run_single_node (src/run/single_node.rs:83-137) on the
 * GPU, then empirical_pvalue.  idx1/idx2: host arrays of n_perm*len1 and
 * n_perm*len2 explicit permutations, or NULL for on-device Philox shuffles
 * keyed by seed.  perm_out (n_perm records) may be NULL. */
int dto_run(dto_job *job, uint64_t n_perm, uint64_t seed, const uint32_t *idx1,
            const uint32_t *idx2, uint32_t mode, dto_result *result, dto_record *perm_out);
/* Device-resident form of dto_run's permutation loop for callers that keep
 * inputs in HBM: d_idx1/d_idx2 are device pointers (or NULL for Philox),
 * d_out a device array of n_perm records, stream a hipStream_t (NULL =
 * default).  Asynchronous; perm_base offsets the Philox permutation ids. */
int dto_run_device(dto_job *job, uint64_t n_perm, uint64_t perm_base, uint64_t seed,
                   const uint32_t *d_idx1, const uint32_t *d_idx2, uint32_t mode,
                   dto_record *d_out, void *stream);
/* Philox permutation the device draws for (seed, job_id, perm, list): writes
 * len entries (list 1 or 2); used to feed the same shuffles to a checker. */
int dto_philox_permutation(uint64_t seed, uint32_t job_id, uint32_t perm, uint32_t list,
                           uint32_t len, uint32_t *out);
/* empirical_pvalue + fdr over records (src/stat_operations/empirical_pvalue.rs:109-187). */
int dto_empirical_pvalue(const dto_record *unpermuted, const dto_record *perm, uint64_t n_perm,
                         dto_result *result);
/* serde_json::to_string_pretty of the result object (src/main.rs:165):
 * sorted keys, 2-space indent, shortest round-trip floats.  Returns the
 * length written (excluding the NUL) or -1 if cap is too small. */
int dto_result_json(const dto_result *result, char *buf, size_t cap);

/* ---- batched jobs (one persistent launch) -------------------------------- */
typedef struct dto_batch dto_batch;
/* A batch of jobs, each with its own permutation count, run by one persistent
 * kernel over a device work queue.  Philox permutations keyed by (seed, job
 * position in the batch, permutation). */
int dto_batch_create(dto_job *const *jobs, const uint64_t *n_perm, uint64_t n_jobs, uint64_t seed,
                     dto_batch **out);
/* Launch (captured in a hipGraph on first use, replayed afterwards). */
int dto_batch_launch(dto_batch *b, void *stream);
/* Copy results: one dto_result per job. */
int dto_batch_results(dto_batch *b, dto_result *results);
void dto_batch_destroy(dto_batch *b);
/* Autotuner: times the candidate launch shapes on this batch and keeps the
 * fastest; returns the chosen waves-per-workgroup through *waves. */
int dto_batch_autotune(dto_batch *b, int *waves);

/* ---- serving (continuous batching) --------------------------------------- */
typedef struct dto_server dto_server;
typedef struct {
  uint64_t hbm_budget_bytes; /* admission limit on resident job state (0: 8 GiB) */
  uint32_t max_active_jobs;  /* requests with tasks in flight at once (0: 256) */
  uint32_t perms_per_slice;  /* permutations per job per launch slice (0: 1024) */
  uint64_t seed;
} dto_server_config;
int dto_server_create(const dto_server_config *cfg, dto_server **out);
/* Submits a job (the server takes ownership) with n_perm permutations;
 * returns a request id.  Never blocks: DTO_ERR_BUSY when the queue is full. */
int dto_server_submit(dto_server *s, dto_job *job, uint64_t n_perm, uint64_t *request_id);
/* Runs scheduling steps until at least one request completes or timeout_ms
 * passes; writes up to cap completed (request_id, result, latency_s). */
int dto_server_poll(dto_server *s, uint32_t timeout_ms, uint64_t *ids, dto_result *results,
                    double *latency_s, uint32_t cap, uint32_t *n_done);
int dto_server_pending(const dto_server *s, uint64_t *queued, uint64_t *active);
void dto_server_destroy(dto_server *s);

#ifdef __cplusplus
}
#endif
#endif

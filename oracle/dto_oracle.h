/*
 * CPU oracle for the DTO threshold-pair p-value grid and its permutation null.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (dual_threshold_optimization_amd/) never links or calls it.
 *
 * It restates, in plain C, the reference crate's hot path and the third-party
 * arithmetic it depends on (statrs 0.17.1, Cargo.lock:734-737):
 *   - statrs factorial::ln_factorial / ln_binomial, gamma::ln_gamma and
 *     Hypergeometric::sf (call site src/stat_operations/hypergeometric_pvalue.rs:33-50)
 *   - RankedFeatureList::from: stable sort + generate_thresholds
 *     (src/collections/ranked.rs:176-191, 359-375, 527-542)
 *   - feature sets by threshold, plain and permuted
 *     (src/collections/ranked.rs:561-568, src/collections/permuted.rs:90-101)
 *   - intersect_genes (src/stat_operations/intersect_genes.rs:38-56)
 *   - process_threshold_pairs (src/dto/process_threshold_pairs.rs:71-131)
 *   - optimize tie-break (src/dto/optimize_main.rs:53-118)
 *   - fdr (src/stat_operations/fdr.rs:29-60), empirical p-value
 *     (src/stat_operations/empirical_pvalue.rs:109-187)
 *   - compute_population_size (src/dto/compute_population_size.rs:66-104)
 *   - the std::thread chunked permutation loop (src/run/single_node.rs:83-137)
 *
 * Randomness: the reference draws unseeded thread_rng shuffles
 * (src/collections/permuted.rs:56-60).  The oracle instead takes explicit
 * permutation index arrays with the same meaning: position k of the permuted
 * list holds the feature that sits at sorted position idx[k].
 *
 * All floating-point code is compiled with -ffp-contract=off and uses the
 * host C library's log/exp, which is what Rust's f64::ln / f64::exp lower to
 * on x86_64-linux-gnu.
 */
#ifndef DTO_ORACLE_H
#define DTO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- statrs 0.17.1 arithmetic -------------------------------------- */
double or_ln_gamma(double x);
double or_ln_factorial(uint64_t x);
double or_ln_binomial(uint64_t n, uint64_t k);
/* Hypergeometric::new(N,K,n) then sf(x).  Returns 0, or -1 where statrs
 * returns Err (K > N or n > N). */
int or_hypergeom_sf(uint64_t N, uint64_t K, uint64_t n, uint64_t x, double *out);
/* hypergeometric_pvalue(N, K, n, k): -1 where the reference panics. */
int or_hypergeometric_pvalue(uint64_t N, uint64_t K, uint64_t n, uint64_t k, double *out);
double or_fdr(uint64_t l1, uint64_t l2, uint64_t overlap, uint64_t N, double sensitivity);

/* ---- thresholds ------------------------------------------------------ */
/* Writes at most cap thresholds, returns how many the reference makes. */
uint32_t or_generate_thresholds(uint32_t max_rank, uint32_t *out, uint32_t cap);

/* ---- ranked lists ---------------------------------------------------- */
typedef struct or_list or_list;
/* RankedFeatureList::from(genes, ranks): copies ids, stable-sorts by rank and
 * builds thresholds. */
or_list *or_list_new(const char *const *ids, const uint32_t *ranks, uint64_t n);
/* Replace the list's thresholds with increasing custom values (0 ok, -1 not increasing). */
int or_list_set_thresholds(or_list *l, const uint32_t *thr, uint32_t nt);
void or_list_free(or_list *l);
uint64_t or_list_len(const or_list *l);
uint32_t or_list_n_thresholds(const or_list *l);
const uint32_t *or_list_thresholds(const or_list *l);
const uint32_t *or_list_ranks(const or_list *l);
const char *or_list_id(const or_list *l, uint64_t pos);

/* compute_population_size: returns 0 and N, or -1 (no background and the id
 * sets differ) / -2 (list1 id missing from background) / -3 (list2 id missing). */
int or_population_size(const or_list *l1, const or_list *l2,
                       const char *const *bg, uint64_t nbg, int has_bg,
                       uint64_t *N);

/* One record of the threshold grid (OptimizationResultRecord without the
 * debug feature sets; src/dto/results_objects.rs:21-32). */
typedef struct {
  uint64_t rank1, rank2;
  uint64_t set1_len, set2_len;
  uint64_t population_size;
  uint64_t intersection_size;
  double pvalue;
  int32_t permuted;
  uint32_t i, j; /* threshold indices */
  /* optimize only (0 in grid records): how many cells share the minimum p
   * (low 16 bits) and how many of those share the largest intersection
   * (high 16 bits), each saturated at 0xFFFF; optimize_main.rs:86-108 warns
   * when either is above 1 */
  uint32_t ties;
} or_record;

/* process_threshold_pairs with the reference's per-cell hash-set
 * intersection.  idx1/idx2 = NULL means unpermuted.  out has T1*T2 slots,
 * row-major over (threshold1, threshold2).  Returns 0, or -1 if a cell's
 * Hypergeometric::new fails. */
int or_process_threshold_pairs(const or_list *l1, const or_list *l2,
                               const uint64_t *idx1, const uint64_t *idx2,
                               uint64_t N, or_record *out);
/* optimize (non-debug): min p, then max intersection, then min (rank1, rank2).
 * Returns 0, -1 (hypergeometric error) or -2 (empty grid). */
int or_optimize(const or_list *l1, const or_list *l2,
                const uint64_t *idx1, const uint64_t *idx2, uint64_t N,
                or_record *best);

/* Same results as or_optimize, but overlap counts come from a (bin1, bin2)
 * histogram + 2D prefix sum and ln_factorial is memoised in a table.  The
 * p-value arithmetic (term order, operations) is the one above, so results
 * are bit-identical; this is only a faster checker for larger lists.
 * Cross-checked against or_optimize by the tests. */
int or_fast_grid(const or_list *l1, const or_list *l2,
                 const uint64_t *idx1, const uint64_t *idx2, uint64_t N,
                 int32_t *k_out, double *p_out);
int or_fast_optimize(const or_list *l1, const or_list *l2,
                     const uint64_t *idx1, const uint64_t *idx2, uint64_t N,
                     or_record *best);
/* Many permutations on nthreads host threads: perm p uses idx1 + p*N1 and
 * idx2 + p*N2.  fast != 0 selects or_fast_optimize. */
int or_optimize_many(const or_list *l1, const or_list *l2,
                     const uint64_t *idx1, const uint64_t *idx2, uint64_t P,
                     uint64_t N, int fast, int nthreads, or_record *out);

/* The reference's CPU cost structure, for the baseline timing: the
 * std::thread chunked loop of src/run/single_node.rs:95-133 over P
 * permutations, each evaluated only on every row_stride-th threshold1 row
 * (row_stride = 1 is the full grid).  Returns the number of cells evaluated. */
uint64_t or_reference_sample(const or_list *l1, const or_list *l2,
                             const uint64_t *idx1, const uint64_t *idx2,
                             uint64_t P, uint64_t N, uint32_t row_stride,
                             int nthreads, double *checksum);

/* The host C library's exp over an array (the reference's f64::exp). */
void or_exp_many(const double *x, uint64_t n, double *out);

/* empirical p-value over permuted records (src/stat_operations/empirical_pvalue.rs:159-173). */
double or_empirical_pvalue(double p0, const double *perm_p, uint64_t P);

#ifdef __cplusplus
}
#endif
#endif

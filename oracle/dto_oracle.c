/*
 * CPU oracle for the DTO hot path.  TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product.  Each function cites the reference file:line it restates.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, host libm).
 *
 * Parity pins (see tests/test_oracle.py):
 *   - statrs arithmetic: the exact-equality known answers
 *     hypergeometric_pvalue(1000,50,60,10) == 0.00044068070222441115 and
 *     (6060,5808,154,153) == 0.010413637619010246
 *     (src/stat_operations/hypergeometric_pvalue.rs:27-31), the 10-gene
 *     optimize answer p == 0.33333333333333337 (src/dto/optimize_main.rs:170),
 *     the README end-to-end answer p == 0.15632183908046102 (README.md:174-186),
 *     the 1e-10 cases (hypergeometric_pvalue.rs:57-86), and exact rational
 *     tails (Python fractions) on small grids.
 *   - thresholds: closed-form checks and T[0] == 1 (ranked.rs:258-259).
 *   - overlaps / tie-break: brute force on tiny inputs; the README and
 *     10-gene optimum pairs.
 */
#include "dto_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* statrs 0.17.1 arithmetic                                               */
/* ===================================================================== */

/* statrs factorial.rs: FCACHE[0] = 1, FCACHE[i] = FCACHE[i-1] * i, i <= 170. */
#define OR_MAX_FACTORIAL 170
static double or_fcache[OR_MAX_FACTORIAL + 1];
static pthread_once_t or_fcache_once = PTHREAD_ONCE_INIT;
static void or_fcache_init(void) {
  or_fcache[0] = 1.0;
  for (int i = 1; i <= OR_MAX_FACTORIAL; i++) or_fcache[i] = or_fcache[i - 1] * (double)i;
}

/* statrs gamma.rs: Lanczos approximation, g = 10.900511, 11 coefficients
 * (Godfrey's set as used by Math.NET / statrs). */
static const double OR_GAMMA_R = 10.900511;
static const double OR_GAMMA_DK[11] = {
    2.48574089138753565546e-5,  1.05142378581721974210,    -3.45687097222016235469,
    4.51227709466894823700,     -2.98285225323576655721,   1.05639711577126713077,
    -1.95428773191645869583e-1, 1.70970543404441224307e-2, -5.71926117404305781283e-4,
    4.63399473359905636708e-6,  -2.71994908488607703910e-9};
static const double OR_LN_2_SQRT_E_OVER_PI = 0.6207822376352452223455184457816472122518527279025978;
static const double OR_LN_PI = 1.1447298858494001741434273513530587116472948129153;

double or_ln_gamma(double x) {
  if (x < 0.5) {
    double s = OR_GAMMA_DK[0];
    for (int i = 1; i <= 10; i++) s = s + OR_GAMMA_DK[i] / ((double)i - x);
    return OR_LN_PI - log(sin(M_PI * x)) - log(s) - OR_LN_2_SQRT_E_OVER_PI -
           (0.5 - x) * log((0.5 - x + OR_GAMMA_R) / M_E);
  }
  double s = OR_GAMMA_DK[0];
  for (int i = 1; i <= 10; i++) s = s + OR_GAMMA_DK[i] / ((x + (double)i) - 1.0);
  return (log(s) + OR_LN_2_SQRT_E_OVER_PI) + (x - 0.5) * log(((x - 0.5) + OR_GAMMA_R) / M_E);
}

/* statrs factorial::ln_factorial: cache hit -> ln(FCACHE[x]), else ln_gamma(x+1). */
double or_ln_factorial(uint64_t x) {
  pthread_once(&or_fcache_once, or_fcache_init);
  if (x <= OR_MAX_FACTORIAL) return log(or_fcache[x]);
  return or_ln_gamma((double)x + 1.0);
}

/* statrs factorial::ln_binomial: k > n -> -inf, else lf(n) - lf(k) - lf(n-k). */
double or_ln_binomial(uint64_t n, uint64_t k) {
  if (k > n) return -INFINITY;
  return (or_ln_factorial(n) - or_ln_factorial(k)) - or_ln_factorial(n - k);
}

/* statrs Hypergeometric::new + DiscreteCDF::sf.  min = (n+K) saturating_sub N,
 * max = min(K, n); x < min -> 1.0; x >= max -> 0.0; otherwise the ascending
 * fold acc + exp((ln_binomial(K,i) + ln_binomial(N-K,n-i)) - ln_binomial(N,n)). */
int or_hypergeom_sf(uint64_t N, uint64_t K, uint64_t n, uint64_t x, double *out) {
  if (K > N || n > N) return -1;
  uint64_t mn = (n + K > N) ? n + K - N : 0;
  uint64_t mx = K < n ? K : n;
  if (x < mn) { *out = 1.0; return 0; }
  if (x >= mx) { *out = 0.0; return 0; }
  double ln_denom = or_ln_binomial(N, n);
  double acc = 0.0;
  for (uint64_t i = x + 1; i <= mx; i++) {
    double e = (or_ln_binomial(K, i) + or_ln_binomial(N - K, n - i)) - ln_denom;
    acc = acc + exp(e);
  }
  *out = acc;
  return 0;
}

/* src/stat_operations/hypergeometric_pvalue.rs:33-50: new(...).expect(), then
 * k == 0 -> 1.0, else sf(k - 1). */
int or_hypergeometric_pvalue(uint64_t N, uint64_t K, uint64_t n, uint64_t k, double *out) {
  if (K > N || n > N) return -1;
  if (k == 0) { *out = 1.0; return 0; }
  return or_hypergeom_sf(N, K, n, k - 1, out);
}

/* src/stat_operations/fdr.rs:29-60. */
double or_fdr(uint64_t l1, uint64_t l2, uint64_t overlap, uint64_t N, double sensitivity) {
  double df = fmax((double)overlap / sensitivity, 0.0);
  double b = fmax((double)l1 - df, 0.0);
  double r = fmax((double)l2 - df, 0.0);
  double num = b * r;
  double den = (double)N * (double)overlap;
  return den > 0.0 ? num / den : 0.0;
}

/* ===================================================================== */
/* thresholds: src/collections/ranked.rs:359-375                          */
/* ===================================================================== */
/* T1 = 1; T_{n+1} = floor(T_n * 1.01 + 1.0) in f64 (no FMA) while <= max
 * rank.  The "set last threshold to max rank" fix-up at :369-372 edits the
 * old vector before the assignment at :374, so it changes nothing. */
uint32_t or_generate_thresholds(uint32_t max_rank, uint32_t *out, uint32_t cap) {
  uint32_t n = 0;
  uint64_t cur = 1;
  while (cur <= max_rank) {
    if (n < cap && out) out[n] = (uint32_t)cur;
    n++;
    double v = floor((double)cur * 1.01 + 1.0);
    if (v >= 4294967295.0) break; /* `as u32` saturates; stop instead of looping */
    cur = (uint64_t)v;
  }
  return n;
}

/* ===================================================================== */
/* ranked lists                                                            */
/* ===================================================================== */
struct or_list {
  uint64_t n;
  char **ids;      /* sorted order */
  uint32_t *ranks; /* sorted ascending, stable */
  uint32_t nt;
  uint32_t *thr;
  uint32_t *bin;   /* bin[p] = first t with thr[t] >= ranks[p], nt if none */
};

typedef struct { uint32_t rank; uint64_t pos; } or_rp;
static int or_rp_cmp(const void *a, const void *b) {
  const or_rp *x = a, *y = b;
  if (x->rank != y->rank) return x->rank < y->rank ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* RankedFeatureList::from: sort_genes_and_ranks (stable sort_by_key on rank,
 * ranked.rs:527-542) then generate_thresholds with max rank = last rank. */
or_list *or_list_new(const char *const *ids, const uint32_t *ranks, uint64_t n) {
  or_list *l = calloc(1, sizeof(or_list));
  l->n = n;
  l->ids = calloc(n ? n : 1, sizeof(char *));
  l->ranks = calloc(n ? n : 1, sizeof(uint32_t));
  or_rp *rp = malloc((n ? n : 1) * sizeof(or_rp));
  for (uint64_t i = 0; i < n; i++) { rp[i].rank = ranks[i]; rp[i].pos = i; }
  qsort(rp, n, sizeof(or_rp), or_rp_cmp);
  for (uint64_t i = 0; i < n; i++) {
    l->ids[i] = strdup(ids[rp[i].pos]);
    l->ranks[i] = rp[i].rank;
  }
  free(rp);
  uint32_t max_rank = n ? l->ranks[n - 1] : 0;
  l->nt = or_generate_thresholds(max_rank, NULL, 0);
  l->thr = calloc(l->nt ? l->nt : 1, sizeof(uint32_t));
  or_generate_thresholds(max_rank, l->thr, l->nt);
  l->bin = calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t t = 0;
  for (uint64_t p = 0; p < n; p++) {
    while (t < l->nt && l->thr[t] < l->ranks[p]) t++;
    l->bin[p] = t;
  }
  return l;
}

/* Custom threshold values (increasing), replacing generate_thresholds: the
 * dense-grid configuration has no reference API (thresholds are private,
 * ranked.rs:132), so the oracle takes them as given and keeps the reference's
 * per-threshold set definition, ranks <= t (ranked.rs:561-568). */
int or_list_set_thresholds(or_list *l, const uint32_t *thr, uint32_t nt) {
  for (uint32_t t = 1; t < nt; t++)
    if (thr[t] <= thr[t - 1]) return -1;
  free(l->thr);
  l->nt = nt;
  l->thr = calloc(nt ? nt : 1, sizeof(uint32_t));
  memcpy(l->thr, thr, (size_t)nt * sizeof(uint32_t));
  uint32_t t = 0;
  for (uint64_t p = 0; p < l->n; p++) {
    while (t < l->nt && l->thr[t] < l->ranks[p]) t++;
    l->bin[p] = t;
  }
  return 0;
}

void or_list_free(or_list *l) {
  if (!l) return;
  for (uint64_t i = 0; i < l->n; i++) free(l->ids[i]);
  free(l->ids); free(l->ranks); free(l->thr); free(l->bin); free(l);
}
uint64_t or_list_len(const or_list *l) { return l->n; }
uint32_t or_list_n_thresholds(const or_list *l) { return l->nt; }
const uint32_t *or_list_thresholds(const or_list *l) { return l->thr; }
const uint32_t *or_list_ranks(const or_list *l) { return l->ranks; }
const char *or_list_id(const or_list *l, uint64_t pos) { return pos < l->n ? l->ids[pos] : NULL; }

/* ---- a plain string hash set (stands in for Rust's HashSet<&str>) ---- */
typedef struct { const char **slot; uint64_t mask; } or_sset;
static uint64_t or_hash(const char *s) {
  uint64_t h = 1469598103934665603ULL;
  for (; *s; s++) { h ^= (unsigned char)*s; h *= 1099511628211ULL; }
  return h ^ (h >> 29);
}
static void or_sset_init(or_sset *h, uint64_t n) {
  uint64_t cap = 16;
  while (cap < 2 * n + 2) cap <<= 1;
  h->slot = calloc(cap, sizeof(char *));
  h->mask = cap - 1;
}
static void or_sset_free(or_sset *h) { free(h->slot); }
static void or_sset_insert(or_sset *h, const char *s) {
  uint64_t i = or_hash(s) & h->mask;
  while (h->slot[i]) {
    if (strcmp(h->slot[i], s) == 0) return;
    i = (i + 1) & h->mask;
  }
  h->slot[i] = s;
}
static int or_sset_has(const or_sset *h, const char *s) {
  uint64_t i = or_hash(s) & h->mask;
  while (h->slot[i]) {
    if (strcmp(h->slot[i], s) == 0) return 1;
    i = (i + 1) & h->mask;
  }
  return 0;
}

/* src/dto/compute_population_size.rs:66-104.  No background: N =
 * |list1.intersect(list2)| (list1 entries whose id is in both) and it must
 * equal both lengths.  Background: every list id must be in it; N = number of
 * background entries (duplicates and blank lines included). */
int or_population_size(const or_list *l1, const or_list *l2, const char *const *bg,
                       uint64_t nbg, int has_bg, uint64_t *N) {
  if (!has_bg) {
    or_sset s2; or_sset_init(&s2, l2->n);
    for (uint64_t i = 0; i < l2->n; i++) or_sset_insert(&s2, l2->ids[i]);
    uint64_t inter = 0;
    for (uint64_t i = 0; i < l1->n; i++) inter += or_sset_has(&s2, l1->ids[i]);
    or_sset_free(&s2);
    if (inter != l1->n || inter != l2->n) return -1;
    *N = inter;
    return 0;
  }
  or_sset sb; or_sset_init(&sb, nbg);
  for (uint64_t i = 0; i < nbg; i++) or_sset_insert(&sb, bg[i]);
  int rc = 0;
  for (uint64_t i = 0; i < l1->n && !rc; i++) if (!or_sset_has(&sb, l1->ids[i])) rc = -2;
  for (uint64_t i = 0; i < l2->n && !rc; i++) if (!or_sset_has(&sb, l2->ids[i])) rc = -3;
  or_sset_free(&sb);
  if (!rc) *N = nbg;
  return rc;
}

/* ===================================================================== */
/* process_threshold_pairs / optimize (reference structure)               */
/* ===================================================================== */

/* Feature set for threshold t (ranked.rs:561-568 unpermuted;
 * permuted.rs:90-101 permuted: position k contributes ids[idx[k]]). */
static uint64_t or_feature_set(const or_list *l, const uint64_t *idx, uint32_t t,
                               const char **out) {
  uint64_t m = 0;
  for (uint64_t k = 0; k < l->n; k++)
    if (l->ranks[k] <= t) out[m++] = l->ids[idx ? idx[k] : k];
  return m;
}

/* intersect_genes.rs:38-56: HashSet of genes1 ids, count genes2 entries in it. */
static uint64_t or_intersect(const char **g1, uint64_t n1, const char **g2, uint64_t n2) {
  or_sset h; or_sset_init(&h, n1);
  for (uint64_t i = 0; i < n1; i++) or_sset_insert(&h, g1[i]);
  uint64_t c = 0;
  for (uint64_t i = 0; i < n2; i++) c += or_sset_has(&h, g2[i]);
  or_sset_free(&h);
  return c;
}

/* Rows [row0, T1) step row_stride of process_threshold_pairs.rs:71-131.
 * The list2 sets are cached per threshold2 as the HashMap at :78,92-98 does. */
static int or_grid_rows(const or_list *l1, const or_list *l2, const uint64_t *idx1,
                        const uint64_t *idx2, uint64_t N, uint32_t row_stride,
                        or_record *out, uint64_t *cells) {
  const char **g1 = malloc((l1->n ? l1->n : 1) * sizeof(char *));
  const char ***g2 = calloc(l2->nt ? l2->nt : 1, sizeof(char **));
  uint64_t *g2n = calloc(l2->nt ? l2->nt : 1, sizeof(uint64_t));
  int rc = 0;
  uint64_t nc = 0;
  int permuted = (idx1 != NULL) || (idx2 != NULL);
  for (uint32_t i = 0; i < l1->nt && !rc; i += row_stride) {
    uint64_t n1 = or_feature_set(l1, idx1, l1->thr[i], g1);
    for (uint32_t j = 0; j < l2->nt; j++) {
      if (!g2[j]) {
        g2[j] = malloc((l2->n ? l2->n : 1) * sizeof(char *));
        g2n[j] = or_feature_set(l2, idx2, l2->thr[j], g2[j]);
      }
      uint64_t k = or_intersect(g1, n1, g2[j], g2n[j]);
      double p;
      if (or_hypergeometric_pvalue(N, n1, g2n[j], k, &p)) { rc = -1; break; }
      if (out) {
        or_record *r = &out[(uint64_t)i * l2->nt + j];
        r->rank1 = l1->thr[i]; r->rank2 = l2->thr[j];
        r->set1_len = n1; r->set2_len = g2n[j];
        r->population_size = N; r->intersection_size = k;
        r->pvalue = p; r->permuted = permuted; r->i = i; r->j = j; r->ties = 0;
      } else {
        /* keep the work observable for the timing sample */
        static __thread volatile double sink;
        sink = sink + p;
      }
      nc++;
    }
  }
  for (uint32_t j = 0; j < l2->nt; j++) free(g2[j]);
  free(g2); free(g2n); free(g1);
  if (cells) *cells = nc;
  return rc;
}

int or_process_threshold_pairs(const or_list *l1, const or_list *l2, const uint64_t *idx1,
                               const uint64_t *idx2, uint64_t N, or_record *out) {
  return or_grid_rows(l1, l2, idx1, idx2, N, 1, out, NULL);
}

/* src/dto/optimize_main.rs:73-116 over a record array. */
static int or_select_best(const or_record *r, uint64_t n, or_record *best) {
  if (n == 0) return -2;
  double minp = INFINITY;
  for (uint64_t c = 0; c < n; c++) minp = fmin(minp, r[c].pvalue); /* f64::min fold */
  uint64_t maxk = 0; int found = 0;
  for (uint64_t c = 0; c < n; c++)
    if (r[c].pvalue == minp) {
      if (!found || r[c].intersection_size > maxk) maxk = r[c].intersection_size;
      found = 1;
    }
  const or_record *b = NULL;
  uint64_t n_p = 0, n_pk = 0; /* best_results.len() at optimize_main.rs:86 and :102 */
  for (uint64_t c = 0; c < n; c++) {
    if (r[c].pvalue != minp) continue;
    n_p++;
    if (r[c].intersection_size != maxk) continue;
    n_pk++;
    if (!b || r[c].rank1 < b->rank1 || (r[c].rank1 == b->rank1 && r[c].rank2 < b->rank2)) b = &r[c];
  }
  *best = *b;
  best->ties = (uint32_t)(n_p < 0xFFFF ? n_p : 0xFFFF) | ((uint32_t)(n_pk < 0xFFFF ? n_pk : 0xFFFF) << 16);
  return 0;
}

int or_optimize(const or_list *l1, const or_list *l2, const uint64_t *idx1, const uint64_t *idx2,
                uint64_t N, or_record *best) {
  uint64_t n = (uint64_t)l1->nt * l2->nt;
  if (n == 0) return -2;
  or_record *r = malloc(n * sizeof(or_record));
  int rc = or_process_threshold_pairs(l1, l2, idx1, idx2, N, r);
  if (!rc) rc = or_select_best(r, n, best);
  free(r);
  return rc;
}

/* ===================================================================== */
/* faster checker: histogram overlaps + memoised ln_factorial             */
/* ===================================================================== */

/* id -> list1 positions (CSR), built with a string hash map. */
typedef struct {
  uint64_t *off; /* n2 + 1 */
  uint64_t *pos;
} or_join;

static void or_build_join(const or_list *l1, const or_list *l2, or_join *j) {
  /* map each list1 id to its positions */
  uint64_t cap = 16;
  while (cap < 2 * l1->n + 2) cap <<= 1;
  int64_t *head = malloc(cap * sizeof(int64_t));
  const char **key = calloc(cap, sizeof(char *));
  int64_t *next = malloc((l1->n ? l1->n : 1) * sizeof(int64_t));
  for (uint64_t i = 0; i < cap; i++) head[i] = -1;
  for (uint64_t m = 0; m < l1->n; m++) {
    uint64_t s = or_hash(l1->ids[m]) & (cap - 1);
    while (key[s] && strcmp(key[s], l1->ids[m])) s = (s + 1) & (cap - 1);
    key[s] = l1->ids[m];
    next[m] = head[s];
    head[s] = (int64_t)m;
  }
  j->off = calloc(l2->n + 1, sizeof(uint64_t));
  uint64_t total = 0;
  for (int pass = 0; pass < 2; pass++) {
    total = 0;
    for (uint64_t q = 0; q < l2->n; q++) {
      uint64_t s = or_hash(l2->ids[q]) & (cap - 1);
      while (key[s] && strcmp(key[s], l2->ids[q])) s = (s + 1) & (cap - 1);
      if (pass == 0) j->off[q] = total;
      if (key[s])
        for (int64_t m = head[s]; m >= 0; m = next[m]) {
          if (pass == 1) j->pos[total] = (uint64_t)m;
          total++;
        }
    }
    if (pass == 0) { j->off[l2->n] = total; j->pos = malloc((total ? total : 1) * sizeof(uint64_t)); }
  }
  free(head); free(key); free(next);
}

int or_fast_grid(const or_list *l1, const or_list *l2, const uint64_t *idx1, const uint64_t *idx2,
                 uint64_t N, int32_t *k_out, double *p_out) {
  uint32_t T1 = l1->nt, T2 = l2->nt;
  or_join J; or_build_join(l1, l2, &J);
  /* nb1[m]: bin of list1 entry m after the permutation (position k holds idx1[k]) */
  uint32_t *nb1 = malloc((l1->n ? l1->n : 1) * sizeof(uint32_t));
  for (uint64_t k = 0; k < l1->n; k++) nb1[idx1 ? idx1[k] : k] = l1->bin[k];
  /* histogram of (a, b) */
  uint64_t cells = (uint64_t)T1 * T2;
  int64_t *g = calloc(cells ? cells : 1, sizeof(int64_t));
  for (uint64_t k = 0; k < l2->n; k++) {
    uint64_t q = idx2 ? idx2[k] : k;
    uint32_t a = T1;
    for (uint64_t t = J.off[q]; t < J.off[q + 1]; t++) if (nb1[J.pos[t]] < a) a = nb1[J.pos[t]];
    uint32_t b = l2->bin[k];
    if (a < T1 && b < T2) g[(uint64_t)a * T2 + b]++;
  }
  for (uint32_t i = 0; i < T1; i++)
    for (uint32_t j = 0; j < T2; j++) {
      int64_t v = g[(uint64_t)i * T2 + j];
      if (i) v += g[(uint64_t)(i - 1) * T2 + j];
      if (j) v += g[(uint64_t)i * T2 + j - 1];
      if (i && j) v -= g[(uint64_t)(i - 1) * T2 + j - 1];
      g[(uint64_t)i * T2 + j] = v;
    }
  /* set sizes */
  uint64_t *K = calloc(T1 ? T1 : 1, sizeof(uint64_t)), *n2 = calloc(T2 ? T2 : 1, sizeof(uint64_t));
  for (uint64_t p = 0; p < l1->n; p++) if (l1->bin[p] < T1) K[l1->bin[p]]++;
  for (uint32_t i = 1; i < T1; i++) K[i] += K[i - 1];
  for (uint64_t p = 0; p < l2->n; p++) if (l2->bin[p] < T2) n2[l2->bin[p]]++;
  for (uint32_t j = 1; j < T2; j++) n2[j] += n2[j - 1];
  /* memoised ln_factorial (same values statrs computes per call) */
  double *LF = malloc((N + 1) * sizeof(double));
  for (uint64_t x = 0; x <= N; x++) LF[x] = or_ln_factorial(x);
  int rc = 0;
  for (uint32_t i = 0; i < T1 && !rc; i++)
    for (uint32_t j = 0; j < T2; j++) {
      uint64_t k = (uint64_t)g[(uint64_t)i * T2 + j];
      uint64_t Kc = K[i], nc = n2[j];
      if (k_out) k_out[(uint64_t)i * T2 + j] = (int32_t)k;
      if (Kc > N || nc > N) { rc = -1; break; }
      double p;
      if (k == 0) p = 1.0;
      else {
        uint64_t x = k - 1;
        uint64_t mn = (nc + Kc > N) ? nc + Kc - N : 0;
        uint64_t mx = Kc < nc ? Kc : nc;
        if (x < mn) p = 1.0;
        else if (x >= mx) p = 0.0;
        else {
          double d = (LF[N] - LF[nc]) - LF[N - nc];
          double acc = 0.0;
          for (uint64_t t = x + 1; t <= mx; t++) {
            double lb1 = (LF[Kc] - LF[t]) - LF[Kc - t];
            double lb2 = (LF[N - Kc] - LF[nc - t]) - LF[(N - Kc) - (nc - t)];
            acc = acc + exp((lb1 + lb2) - d);
          }
          p = acc;
        }
      }
      if (p_out) p_out[(uint64_t)i * T2 + j] = p;
    }
  free(LF); free(K); free(n2); free(g); free(nb1); free(J.off); free(J.pos);
  return rc;
}

int or_fast_optimize(const or_list *l1, const or_list *l2, const uint64_t *idx1,
                     const uint64_t *idx2, uint64_t N, or_record *best) {
  uint32_t T1 = l1->nt, T2 = l2->nt;
  uint64_t n = (uint64_t)T1 * T2;
  if (n == 0) return -2;
  int32_t *k = malloc(n * sizeof(int32_t));
  double *p = malloc(n * sizeof(double));
  int rc = or_fast_grid(l1, l2, idx1, idx2, N, k, p);
  if (!rc) {
    uint64_t *K = calloc(T1, sizeof(uint64_t)), *n2 = calloc(T2, sizeof(uint64_t));
    for (uint64_t q = 0; q < l1->n; q++) if (l1->bin[q] < T1) K[l1->bin[q]]++;
    for (uint32_t i = 1; i < T1; i++) K[i] += K[i - 1];
    for (uint64_t q = 0; q < l2->n; q++) if (l2->bin[q] < T2) n2[l2->bin[q]]++;
    for (uint32_t j = 1; j < T2; j++) n2[j] += n2[j - 1];
    or_record *r = malloc(n * sizeof(or_record));
    int permuted = (idx1 != NULL) || (idx2 != NULL);
    for (uint32_t i = 0; i < T1; i++)
      for (uint32_t j = 0; j < T2; j++) {
        or_record *c = &r[(uint64_t)i * T2 + j];
        c->rank1 = l1->thr[i]; c->rank2 = l2->thr[j];
        c->set1_len = K[i]; c->set2_len = n2[j]; c->population_size = N;
        c->intersection_size = (uint64_t)k[(uint64_t)i * T2 + j];
        c->pvalue = p[(uint64_t)i * T2 + j]; c->permuted = permuted; c->i = i; c->j = j; c->ties = 0;
      }
    rc = or_select_best(r, n, best);
    free(r); free(K); free(n2);
  }
  free(k); free(p);
  return rc;
}

/* ===================================================================== */
/* threaded permutation loop: src/run/single_node.rs:95-133               */
/* ===================================================================== */
typedef struct {
  const or_list *l1, *l2;
  const uint64_t *idx1, *idx2;
  uint64_t p0, p1, N;
  int fast;
  uint32_t row_stride;
  or_record *out;
  int rc;
  uint64_t cells;
  double checksum;
} or_task;

static void *or_worker(void *arg) {
  or_task *t = arg;
  for (uint64_t p = t->p0; p < t->p1 && !t->rc; p++) {
    const uint64_t *a = t->idx1 ? t->idx1 + p * t->l1->n : NULL;
    const uint64_t *b = t->idx2 ? t->idx2 + p * t->l2->n : NULL;
    if (t->out) {
      t->rc = t->fast ? or_fast_optimize(t->l1, t->l2, a, b, t->N, &t->out[p])
                      : or_optimize(t->l1, t->l2, a, b, t->N, &t->out[p]);
    } else {
      uint64_t c = 0;
      t->rc = or_grid_rows(t->l1, t->l2, a, b, t->N, t->row_stride, NULL, &c);
      t->cells += c;
    }
  }
  return NULL;
}

static int or_run_threads(or_task *proto, uint64_t P, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  uint64_t chunk = (P + (uint64_t)nthreads - 1) / (uint64_t)nthreads; /* tasks.len().div_ceil(n) */
  if (chunk == 0) chunk = 1;
  int nt = (int)((P + chunk - 1) / chunk);
  pthread_t *th = malloc((nt ? nt : 1) * sizeof(pthread_t));
  or_task *ts = malloc((nt ? nt : 1) * sizeof(or_task));
  for (int w = 0; w < nt; w++) {
    ts[w] = *proto;
    ts[w].p0 = (uint64_t)w * chunk;
    ts[w].p1 = ts[w].p0 + chunk < P ? ts[w].p0 + chunk : P;
    pthread_create(&th[w], NULL, or_worker, &ts[w]);
  }
  int rc = 0;
  proto->cells = 0;
  for (int w = 0; w < nt; w++) {
    pthread_join(th[w], NULL);
    if (ts[w].rc) rc = ts[w].rc;
    proto->cells += ts[w].cells;
  }
  free(th); free(ts);
  return rc;
}

int or_optimize_many(const or_list *l1, const or_list *l2, const uint64_t *idx1,
                     const uint64_t *idx2, uint64_t P, uint64_t N, int fast, int nthreads,
                     or_record *out) {
  or_task t = {l1, l2, idx1, idx2, 0, 0, N, fast, 1, out, 0, 0, 0.0};
  return or_run_threads(&t, P, nthreads);
}

uint64_t or_reference_sample(const or_list *l1, const or_list *l2, const uint64_t *idx1,
                             const uint64_t *idx2, uint64_t P, uint64_t N, uint32_t row_stride,
                             int nthreads, double *checksum) {
  or_task t = {l1, l2, idx1, idx2, 0, 0, N, 0, row_stride ? row_stride : 1, NULL, 0, 0, 0.0};
  int rc = or_run_threads(&t, P, nthreads);
  if (checksum) *checksum = rc ? -1.0 : 0.0;
  return rc ? 0 : t.cells;
}

double or_empirical_pvalue(double p0, const double *perm_p, uint64_t P) {
  if (P == 0) return 1.0;
  uint64_t c = 0;
  for (uint64_t i = 0; i < P; i++) c += perm_p[i] <= p0;
  return (double)c / (double)P;
}

void or_exp_many(const double *x, uint64_t n, double *out) {
  for (uint64_t i = 0; i < n; i++) out[i] = exp(x[i]);
}

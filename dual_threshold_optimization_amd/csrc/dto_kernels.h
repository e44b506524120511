/*
 * gfx950 kernels for the DTO threshold-pair grid and its permutation null.
 *
 * Replaces, per permutation, process_threshold_pairs + optimize
 * (src/dto/process_threshold_pairs.rs:71-131, src/dto/optimize_main.rs:53-118),
 * and the std::thread permutation loop (src/run/single_node.rs:83-137).
 *
 * One workgroup owns one permutation at a time (grid-stride over
 * permutations; the per-job tables stay in LDS across them):
 *   A  scatter: list1 feature m takes the bin of the position it lands on,
 *      nb[idx1[k]] = bin1[k]                                  (LDS, u16)
 *   B  pairs: list2 position e holds entry q = idx2[e]; its list1 partner's
 *      bin a = nb[join[q]], its own bin b = bin2[e]; counting sort of the
 *      pairs by a (row) into nb                                 (LDS)
 *   C  per-wave column histograms of row blocks, prefix over waves
 *   D  each wave sweeps its rows: add the row's pairs into its running
 *      column counts, wave prefix scan -> exact overlap k(i, j) for the row,
 *      then per cell either
 *        PLAIN:    the exact FP64 statrs tail (dto_tail_exact), or
 *        SCREENED: an FP32 log-space bound of the tail from the LDS table
 *                  H(x) = ln x! - x (ln N - 1); cells that cannot reach the
 *                  minimum are dropped, p == 0 and p == 1 classes are exact.
 *   E  SCREENED only: reverse sweep over rows that can hold the minimum;
 *      surviving cells are evaluated exactly in FP64 (bit-identical to the
 *      reference) and reduced with the reference tie-break.
 * No MFMA: nothing here is a dense contraction.
 */
#pragma once
#include <hip/hip_runtime.h>

#include "dto_device.h"
#include "dto_math.h"

#pragma clang fp contract(off)

namespace dto {

static __constant__ uint64_t kExpTabC[256] = DTO_EXP_TAB_INIT;

constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mask_lt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* optimize's order (src/dto/optimize_main.rs:73-116): smaller p, then larger
 * intersection, then smaller (rank1, rank2); thresholds increase with their
 * index, so (i, j) order is (rank1, rank2) order. */
__device__ __forceinline__ bool better(double p1, uint32_t k1, uint32_t ij1, double p2, uint32_t k2,
                                       uint32_t ij2) {
  if (ij2 == kNone) return ij1 != kNone;
  if (ij1 == kNone) return false;
  if (p1 != p2) return p1 < p2;
  if (k1 != k2) return k1 > k2;
  return ij1 < ij2;
}
__device__ __forceinline__ void take_better(BestKey &b, double p, uint32_t k, uint32_t ij) {
  if (better(p, k, ij, b.p, b.k, b.ij)) {
    b.p = p;
    b.k = k;
    b.ij = ij;
  }
}

__device__ BestKey wave_best(BestKey b) {
  for (int off = 32; off >= 1; off >>= 1) {
    double p = __shfl_xor(b.p, off);
    uint32_t k = __shfl_xor(b.k, off);
    uint32_t ij = __shfl_xor(b.ij, off);
    take_better(b, p, k, ij);
  }
  return b;
}

__device__ __forceinline__ float wave_minf(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
  uint32_t x = v;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  return x - v;
}

/* Exclusive scan of data[0..len) in LDS, total written to data[len]. */
template <int NT>
__device__ void block_excl_scan(uint32_t *data, uint32_t len, uint32_t *wave_sums) {
  const uint32_t tid = threadIdx.x, w = tid >> 6;
  const uint32_t per = (len + NT - 1) / NT;
  const uint32_t beg = min(tid * per, len), end = min(beg + per, len);
  uint32_t s = 0;
  for (uint32_t t = beg; t < end; ++t) s += data[t];
  uint32_t ex = wave_excl_scan(s);
  if (lane_id() == 63) wave_sums[w] = ex + s;
  __syncthreads();
  uint32_t run = ex;
  for (uint32_t v = 0; v < w; ++v) run += wave_sums[v];
  for (uint32_t t = beg; t < end; ++t) {
    uint32_t x = data[t];
    data[t] = run;
    run += x;
  }
  if (tid == NT - 1) data[len] = run;
  __syncthreads();
}

struct Lds {
  float *h;
  uint16_t *nb;
  uint32_t *rowstart, *cursor, *cw;
  float *rowmin;
  uint32_t *coln;
  float *colb;
  uint2 *cand;
  uint64_t *tab;
  float *redU;
  BestKey *redB; /* [3][W]: zero class, one class, evaluated */
  uint32_t *redS;
};

__device__ __forceinline__ Lds lds_view(unsigned char *smem, const LdsLayout &Lo, int W) {
  Lds s;
  s.h = (float *)(smem + Lo.h);
  s.nb = (uint16_t *)(smem + Lo.nb);
  s.rowstart = (uint32_t *)(smem + Lo.rowstart);
  s.cursor = (uint32_t *)(smem + Lo.cursor);
  s.cw = (uint32_t *)(smem + Lo.cw);
  s.rowmin = (float *)(smem + Lo.rowmin);
  s.coln = (uint32_t *)(smem + Lo.coln);
  s.colb = (float *)(smem + Lo.colb);
  s.cand = (uint2 *)(smem + Lo.cand);
  s.tab = (uint64_t *)(smem + Lo.exptab);
  s.redU = (float *)(smem + Lo.red);
  s.redB = (BestKey *)(smem + Lo.red + 16 * W);
  s.redS = (uint32_t *)(smem + Lo.red + 16 * W + 3 * W * sizeof(BestKey));
  return s;
}

/* Per-job tables that stay in LDS while a workgroup works on that job. */
template <int NT, bool HLDS>
__device__ void load_job_tables(const DevJob &J, const Lds &S) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t t = tid; t < 256; t += NT) S.tab[t] = kExpTabC[t];
  if (HLDS)
    for (uint32_t t = tid; t <= J.N; t += NT) S.h[t] = J.H32[t];
  for (uint32_t t = tid; t < J.T2; t += NT) {
    S.coln[t] = J.n[t];
    S.colb[t] = J.B32[t];
  }
}

/* Phases A-C for one permutation: leaves the row-sorted list2 bins in S.nb,
 * row offsets in S.rowstart, and in each wave's S.cw block the column counts
 * of all pairs in rows before the wave's first row. */
template <int NT, int RMAX, bool EXPLICIT>
__device__ void build_rows(const DevJob &J, const LdsLayout &Lo, const Lds &S, const uint32_t *i1,
                           const uint32_t *i2, const Shuffle &s1, const Shuffle &s2, int *err) {
  constexpr int W = NT / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const uint32_t T1 = J.T1, T2 = J.T2, T2wp = Lo.T2w_pad;
  for (uint32_t t = tid; t <= T1; t += NT) S.rowstart[t] = 0;
  for (uint32_t t = tid; t < W * T2wp; t += NT) S.cw[t] = 0;
  /* A: list1 feature m sits at the position k with idx1[k] == m */
  for (uint32_t k = tid; k < J.N1; k += NT) {
    uint32_t m = EXPLICIT ? (i1 ? i1[k] : k) : shuffle_index(s1, k);
    if (m < J.N1)
      S.nb[m] = J.bin1[k];
    else
      atomicOr(err, 1);
  }
  __syncthreads();
  /* B1: (a, b) per list2 position, counted per row a */
  uint32_t pr[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const uint32_t e = tid + (uint32_t)r * NT;
    pr[r] = kNone;
    if (e < J.N2) {
      uint32_t q = EXPLICIT ? (i2 ? i2[e] : e) : shuffle_index(s2, e);
      uint32_t a = kBinNone;
      if (q >= J.N2) {
        atomicOr(err, 1);
      } else if (!J.dup) {
        int32_t m = J.join[q];
        if (m >= 0) a = S.nb[m];
      } else {
        /* intersect_genes collapses duplicate ids of genes1: the entry is in
         * S1'(t) as soon as any list1 entry with its id is */
        for (uint32_t t = J.join_off[q]; t < J.join_off[q + 1]; ++t) {
          uint32_t v = S.nb[J.join_pos[t]];
          a = v < a ? v : a;
        }
      }
      uint32_t b = J.bin2[e];
      if (a < T1 && b < T2) {
        pr[r] = (a << 16) | b;
        atomicAdd(&S.rowstart[a], 1u);
      }
    }
  }
  __syncthreads();
  block_excl_scan<NT>(S.rowstart, T1, S.redS);
  for (uint32_t t = tid; t < T1; t += NT) S.cursor[t] = S.rowstart[t];
  __syncthreads();
  /* B2: scatter b into row order (overwrites the list1 bins, all read above) */
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    if (pr[r] != kNone) {
      uint32_t pos = atomicAdd(&S.cursor[pr[r] >> 16], 1u);
      S.nb[pos] = (uint16_t)(pr[r] & 0xFFFFu);
    }
  }
  __syncthreads();
  /* C: per-wave column counts of its own row block, then exclusive prefix over waves */
  const uint32_t r0 = (uint32_t)(((uint64_t)w * T1) / W), r1 = (uint32_t)(((uint64_t)(w + 1) * T1) / W);
  uint32_t *C = S.cw + w * T2wp;
  for (uint32_t e = S.rowstart[r0] + lane; e < S.rowstart[r1]; e += 64) {
    uint32_t b = S.nb[e];
    atomicAdd(&C[b >> 1], 1u << ((b & 1u) * 16u));
  }
  __syncthreads();
  for (uint32_t t = tid; t < J.T2w; t += NT) {
    uint32_t run = 0;
    for (int v = 0; v < W; ++v) {
      uint32_t x = S.cw[v * T2wp + t];
      S.cw[v * T2wp + t] = run;
      run += x;
    }
  }
  __syncthreads();
}

/* Adds (sign +1) or removes (-1) row i's pairs into a wave's column counts. */
__device__ __forceinline__ void row_update(const Lds &S, uint32_t *C, uint32_t i, bool add) {
  const uint32_t lane = lane_id();
  for (uint32_t e = S.rowstart[i] + lane; e < S.rowstart[i + 1]; e += 64) {
    uint32_t b = S.nb[e];
    uint32_t inc = 1u << ((b & 1u) * 16u);
    if (add)
      atomicAdd(&C[b >> 1], inc);
    else
      atomicSub(&C[b >> 1], inc);
  }
  wave_sync();
}

/* Start of this lane's run of columns: exclusive wave scan of lane totals. */
__device__ __forceinline__ uint32_t row_lane_offset(const uint32_t *C, uint32_t wbeg, uint32_t c2,
                                                    uint32_t T2w) {
  uint32_t tot = 0;
  for (uint32_t t = 0; t < c2; ++t) {
    uint32_t wi = wbeg + t;
    if (wi < T2w) {
      uint32_t x = C[wi];
      tot += (x & 0xFFFFu) + (x >> 16);
    }
  }
  return wave_excl_scan(tot);
}

struct ScreenCell {
  int cls; /* 0 sum, 1 p == 1 exactly, 2 p == 0 exactly */
  float x; /* FP32 estimate of the first term's log (sum class) */
};

/* Classify one cell and form the FP32 log of its first tail term. */
template <bool HLDS>
__device__ __forceinline__ ScreenCell screen_cell(const DevJob &J, const float *H, uint32_t K, float A,
                                                  uint32_t n, float B, uint32_t k) {
  ScreenCell c;
  const uint32_t N = J.N;
  const uint32_t mn = (K + n > N) ? K + n - N : 0u;
  const uint32_t mx = K < n ? K : n;
  c.x = 0.f;
  if (k == 0 || k <= mn) {
    c.cls = 1; /* k == 0 -> 1.0; k - 1 < min -> sf = 1.0 */
    return c;
  }
  if (k > mx) {
    c.cls = 2; /* k - 1 >= max -> sf = 0.0 (duplicate list2 ids only) */
    return c;
  }
  float s = (H[k] + H[K - k]) + (H[n - k] + H[(N - K) - (n - k)]);
  c.x = (A + B) - s;
  c.cls = 0;
  if (c.x + J.margin < -745.2f) {
    /* every term from k on is smaller (ratio < 1) and its exponent is below
     * the exp underflow point -745.1332...: the reference's sum is exactly 0 */
    uint64_t num = (uint64_t)(K - k) * (uint64_t)(n - k);
    uint64_t den = (uint64_t)(k + 1) * (uint64_t)((N - K) - (n - k) + 1);
    if (num < den) c.cls = 2;
  }
  return c;
}

/* Upper bound of the log of the reference's computed sum for a sum-class cell. */
__device__ __forceinline__ float screen_upper(const DevJob &J, uint32_t K, uint32_t n, uint32_t k, float x) {
  const uint32_t N = J.N;
  const uint32_t mx = K < n ? K : n;
  uint64_t num = (uint64_t)(K - k) * (uint64_t)(n - k);
  uint64_t den = (uint64_t)(k + 1) * (uint64_t)((N - K) - (n - k) + 1);
  if (num >= den) return 1e-6f; /* p <= 1 (+ rounding) */
  float lm = __logf((float)(mx - k + 1));
  float g = fminf(lm, __logf((float)den) - __logf((float)(den - num)));
  float U = x + J.margin + g + 1e-4f;
  if (U < -690.f) U = fmaxf(U, lm - 744.44f) + 0.7f; /* subnormal rounding of the terms */
  return U;
}

template <int NT>
__device__ void flush_candidates(const DevJob &J, const Lds &S, uint2 *cand, uint32_t &nc, BestKey &xb,
                                 float &Uw) {
  const uint32_t lane = lane_id();
  wave_sync();
  BestKey lb = {__builtin_inf(), 0u, kNone};
  for (uint32_t c = lane; c < nc; c += 64) {
    uint2 cd = cand[c];
    uint32_t i = cd.x >> 16, j = cd.x & 0xFFFFu, k = cd.y;
    double p = dto_tail_exact(J.LF, J.N, J.K[i], S.coln[j], k, J.D[j], S.tab, nullptr);
    take_better(lb, p, k, cd.x);
  }
  lb = wave_best(lb);
  take_better(xb, lb.p, lb.k, lb.ij);
  if (xb.ij != kNone) {
    float lp = xb.p > 0.0 ? (float)(log(xb.p) + 1e-3) : -__builtin_inff();
    Uw = fminf(Uw, lp);
  }
  nc = 0;
  wave_sync();
}

/* One permutation: overlap grid -> best cell.  Returns the best key in
 * thread 0 (others: undefined). */
template <int NT, int RMAX, bool EXPLICIT, bool HLDS, bool PLAIN>
__device__ BestKey optimize_perm(const DevJob &J, const LdsLayout &Lo, const Lds &S, const uint32_t *i1,
                                 const uint32_t *i2, const Shuffle &s1, const Shuffle &s2, int *err) {
  constexpr int W = NT / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  build_rows<NT, RMAX, EXPLICIT>(J, Lo, S, i1, i2, s1, s2, err);

  const uint32_t T1 = J.T1, T2 = J.T2, T2w = J.T2w;
  const uint32_t r0 = (uint32_t)(((uint64_t)w * T1) / W), r1 = (uint32_t)(((uint64_t)(w + 1) * T1) / W);
  uint32_t *C = S.cw + w * Lo.T2w_pad;
  const uint32_t c2 = (T2w + 63) / 64;
  const uint32_t wbeg = lane * c2;
  const float *H = HLDS ? S.h : J.H32;

  BestKey zb = {0.0, 0u, kNone};            /* p == 0 class */
  BestKey ob = {1.0, 0u, kNone};            /* p == 1 class */
  BestKey xb = {__builtin_inf(), 0u, kNone}; /* exactly evaluated cells */
  float Urun = __builtin_inff();

  /* ---- D: forward sweep ---- */
  for (uint32_t i = r0; i < r1; ++i) {
    row_update(S, C, i, true);
    uint32_t run = row_lane_offset(C, wbeg, c2, T2w);
    const uint32_t K = J.K[i];
    const float A = J.A32[i];
    float lmin = __builtin_inff();
    for (uint32_t t = 0; t < c2; ++t) {
      const uint32_t wi = wbeg + t;
      const uint32_t x = wi < T2w ? C[wi] : 0u;
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t j = 2 * wi + h;
        run += h ? (x >> 16) : (x & 0xFFFFu);
        if (j >= T2) continue;
        const uint32_t k = run, n = S.coln[j], ij = (i << 16) | j;
        if (PLAIN) {
          double p = dto_tail_exact(J.LF, J.N, K, n, k, J.D[j], S.tab, nullptr);
          take_better(xb, p, k, ij);
          continue;
        }
        ScreenCell c = screen_cell<HLDS>(J, H, K, A, n, S.colb[j], k);
        if (c.cls == 1) {
          take_better(ob, 1.0, k, ij);
          Urun = fminf(Urun, 0.f);
        } else if (c.cls == 2) {
          take_better(zb, 0.0, k, ij);
          Urun = -__builtin_inff();
        } else {
          float L = c.x - J.margin;
          lmin = fminf(lmin, L < -708.f ? -__builtin_inff() : L);
          if (c.x + J.margin < Urun) Urun = fminf(Urun, screen_upper(J, K, n, k, c.x));
        }
      }
    }
    if (!PLAIN) {
      lmin = wave_minf(lmin);
      if (lane == 0) S.rowmin[i] = lmin;
    }
  }

  if (PLAIN) {
    xb = wave_best(xb);
    if (lane == 0) S.redB[2 * W + w] = xb;
    __syncthreads();
    BestKey b = {__builtin_inf(), 0u, kNone};
    if (tid == 0)
      for (int v = 0; v < W; ++v) take_better(b, S.redB[2 * W + v].p, S.redB[2 * W + v].k, S.redB[2 * W + v].ij);
    __syncthreads();
    return b;
  }

  Urun = wave_minf(Urun);
  zb = wave_best(zb);
  ob = wave_best(ob);
  if (lane == 0) {
    S.redU[w] = Urun;
    S.redB[w] = zb;
    S.redB[W + w] = ob;
  }
  __syncthreads();
  float Uw = __builtin_inff();
  for (int v = 0; v < W; ++v) Uw = fminf(Uw, S.redU[v]);

  /* ---- E: reverse sweep over rows that can hold the minimum ---- */
  uint2 *cand = S.cand + w * kCandCap;
  uint32_t nc = 0;
  for (int32_t ii = (int32_t)r1 - 1; ii >= (int32_t)r0; --ii) {
    const uint32_t i = (uint32_t)ii;
    if (S.rowmin[i] <= Uw) {
      uint32_t run = row_lane_offset(C, wbeg, c2, T2w);
      const uint32_t K = J.K[i];
      const float A = J.A32[i];
      for (uint32_t t = 0; t < c2; ++t) {
        const uint32_t wi = wbeg + t;
        const uint32_t x = wi < T2w ? C[wi] : 0u;
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
          const uint32_t j = 2 * wi + h;
          run += h ? (x >> 16) : (x & 0xFFFFu);
          bool f = false;
          const uint32_t k = run;
          if (j < T2) {
            ScreenCell c = screen_cell<HLDS>(J, H, K, A, S.coln[j], S.colb[j], k);
            if (c.cls == 0) {
              float L = c.x - J.margin;
              L = L < -708.f ? -__builtin_inff() : L;
              f = L <= Uw;
            }
          }
          const uint64_t m = __ballot(f);
          const uint32_t cnt = (uint32_t)__popcll(m);
          if (cnt) {
            if (nc + cnt > (uint32_t)kCandCap) flush_candidates<NT>(J, S, cand, nc, xb, Uw);
            if (f) cand[nc + mask_lt(m)] = make_uint2((i << 16) | j, k);
            nc += cnt;
          }
        }
      }
    }
    row_update(S, C, i, false);
  }
  if (nc) flush_candidates<NT>(J, S, cand, nc, xb, Uw);
  if (lane == 0) S.redB[2 * W + w] = xb;
  __syncthreads();
  BestKey b = {__builtin_inf(), 0u, kNone};
  if (tid == 0)
    for (int v = 0; v < 3 * W; ++v) take_better(b, S.redB[v].p, S.redB[v].k, S.redB[v].ij);
  __syncthreads();
  return b;
}

__device__ __forceinline__ void write_record(const DevJob &J, const BestKey &b, int32_t permuted,
                                             dto_record *out, int *err) {
  dto_record r;
  if (b.ij == kNone || (b.ij >> 16) >= J.T1 || (b.ij & 0xFFFFu) >= J.T2) {
    /* no cell selected: cannot happen for a non-empty grid; flag it */
    memset(&r, 0, sizeof r);
    r.pvalue = __builtin_nan("");
    *out = r;
    atomicOr(err, 2);
    return;
  }
  const uint32_t i = b.ij >> 16, j = b.ij & 0xFFFFu;
  r.rank1 = J.thr1[i];
  r.rank2 = J.thr2[j];
  r.set1_len = J.K[i];
  r.set2_len = J.n[j];
  r.population_size = J.N;
  r.intersection_size = b.k;
  r.pvalue = b.p;
  r.permuted = permuted;
  r.idx1 = i;
  r.idx2 = j;
  r.reserved = 0;
  *out = r;
}

/* Permutations [perm_base, perm_base + n_perm) of one job.  EXPLICIT: idx1 /
 * idx2 hold n_perm permutations back to back (NULL: identity); otherwise the
 * Philox shuffle keyed by (seed, job_id, permutation id, list). */
template <int NT, int RMAX, bool EXPLICIT, bool HLDS, bool PLAIN>
__global__ __launch_bounds__(NT) void k_optimize(DevJob J, LdsLayout Lo, const uint32_t *__restrict__ idx1,
                                                 const uint32_t *__restrict__ idx2, uint64_t n_perm,
                                                 uint64_t perm_base, uint64_t seed, uint32_t job_id,
                                                 int32_t permuted, dto_record *__restrict__ out,
                                                 int *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int W = NT / 64;
  const Lds S = lds_view(smem, Lo, W);
  load_job_tables<NT, HLDS>(J, S);
  __syncthreads();
  for (uint64_t pp = blockIdx.x; pp < n_perm; pp += gridDim.x) {
    const uint64_t pid = perm_base + pp;
    const uint32_t *i1 = (EXPLICIT && idx1) ? idx1 + pp * J.N1 : nullptr;
    const uint32_t *i2 = (EXPLICIT && idx2) ? idx2 + pp * J.N2 : nullptr;
    Shuffle s1, s2;
    if (!EXPLICIT) {
      s1 = shuffle_make(seed, job_id, (uint32_t)pid, 1u, J.N1);
      s2 = shuffle_make(seed, job_id, (uint32_t)pid, 2u, J.N2);
    }
    BestKey b = optimize_perm<NT, RMAX, EXPLICIT, HLDS, PLAIN>(J, Lo, S, i1, i2, s1, s2, err);
    if (threadIdx.x == 0) write_record(J, b, permuted, out + pp, err);
  }
}

/* ------------------------------------------------------------------ */
/* batched persistent kernel                                             */
/* ------------------------------------------------------------------ */
/* One job of a batch: its tables, LDS layout, permutation count and the
 * (P+1)-record output (record 0: unpermuted, record p: permutation p-1). */
struct BatchJob {
  DevJob J;
  LdsLayout lo;
  dto_record *out;
  uint32_t n_perm;
  uint32_t job_id;
};

/* A slice of one job's tasks: tasks [t0, t0 + count), task 0 = unpermuted. */
struct WorkItem {
  uint32_t job;
  uint32_t t0;
  uint32_t count;
  uint32_t pad;
};

static __global__ void k_reset_counter(uint32_t *counter, int *err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    counter[0] = 0;
    err[0] = 0;
  }
}

/* Persistent kernel: every workgroup pulls work items from a device queue
 * (one atomic counter) until it is drained; a workgroup reloads the per-job
 * LDS tables only when it moves to another job.  Every workgroup exits once
 * the counter passes n_items. */
template <int NT, int RMAX, bool HLDS>
__global__ __launch_bounds__(NT) void k_batch(const BatchJob *__restrict__ jobs, const WorkItem *__restrict__ items,
                                              uint32_t n_items, uint64_t seed, uint32_t *__restrict__ counter,
                                              int *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int W = NT / 64;
  __shared__ uint32_t s_item;
  uint32_t cur = kNone;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const WorkItem wi = items[it];
    const BatchJob &B = jobs[wi.job];
    const Lds S = lds_view(smem, B.lo, W);
    if (wi.job != cur) {
      load_job_tables<NT, HLDS>(B.J, S);
      __syncthreads();
      cur = wi.job;
    }
    for (uint32_t t = wi.t0; t < wi.t0 + wi.count; ++t) {
      BestKey b;
      Shuffle s1, s2;
      if (t == 0) {
        b = optimize_perm<NT, RMAX, true, HLDS, false>(B.J, B.lo, S, nullptr, nullptr, s1, s2, err);
      } else {
        s1 = shuffle_make(seed, B.job_id, t - 1, 1u, B.J.N1);
        s2 = shuffle_make(seed, B.job_id, t - 1, 2u, B.J.N2);
        b = optimize_perm<NT, RMAX, false, HLDS, false>(B.J, B.lo, S, nullptr, nullptr, s1, s2, err);
      }
      if (threadIdx.x == 0) write_record(B.J, b, t == 0 ? 0 : 1, B.out + t, err);
    }
  }
}

/* Overlap counts of every cell for one permutation (debug grid export). */
template <int NT, int RMAX>
__global__ __launch_bounds__(NT) void k_grid_counts(DevJob J, LdsLayout Lo, const uint32_t *__restrict__ idx1,
                                                    const uint32_t *__restrict__ idx2,
                                                    uint32_t *__restrict__ kout, int *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int W = NT / 64;
  const Lds S = lds_view(smem, Lo, W);
  Shuffle s1, s2;
  build_rows<NT, RMAX, true>(J, Lo, S, idx1, idx2, s1, s2, err);
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  const uint32_t T1 = J.T1, T2 = J.T2, T2w = J.T2w;
  const uint32_t r0 = (uint32_t)(((uint64_t)w * T1) / W), r1 = (uint32_t)(((uint64_t)(w + 1) * T1) / W);
  uint32_t *C = S.cw + w * Lo.T2w_pad;
  const uint32_t c2 = (T2w + 63) / 64, wbeg = lane * c2;
  for (uint32_t i = r0; i < r1; ++i) {
    row_update(S, C, i, true);
    uint32_t run = row_lane_offset(C, wbeg, c2, T2w);
    for (uint32_t t = 0; t < c2; ++t) {
      const uint32_t wi = wbeg + t;
      const uint32_t x = wi < T2w ? C[wi] : 0u;
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t j = 2 * wi + h;
        run += h ? (x >> 16) : (x & 0xFFFFu);
        if (j < T2) kout[(uint64_t)i * T2 + j] = run;
      }
    }
  }
}

/* Exact p-value of every cell from the count grid (debug grid export). */
static __global__ __launch_bounds__(256) void k_grid_pvalues(DevJob J, const uint32_t *__restrict__ kin,
                                                      int32_t permuted, dto_record *__restrict__ out) {
  __shared__ uint64_t tab[256];
  tab[threadIdx.x] = kExpTabC[threadIdx.x];
  __syncthreads();
  const uint64_t cells = (uint64_t)J.T1 * J.T2;
  for (uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (uint64_t)gridDim.x * 256) {
    const uint32_t i = (uint32_t)(c / J.T2), j = (uint32_t)(c % J.T2);
    const uint32_t k = kin[c];
    double p = dto_tail_exact(J.LF, J.N, J.K[i], J.n[j], k, J.D[j], tab, nullptr);
    BestKey b = {p, k, (i << 16) | j};
    write_record(J, b, permuted, out + c, nullptr);
  }
}

/* hypergeometric_pvalue for arrays (Hypergeometric::new already validated on
 * the host).  LF covers 0..max N. */
static __global__ __launch_bounds__(256) void k_hypergeom(const double *__restrict__ LF, const uint64_t *__restrict__ N,
                                                   const uint64_t *__restrict__ K, const uint64_t *__restrict__ n,
                                                   const uint64_t *__restrict__ k, uint64_t count,
                                                   double *__restrict__ out) {
  __shared__ uint64_t tab[256];
  tab[threadIdx.x] = kExpTabC[threadIdx.x];
  __syncthreads();
  for (uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x; c < count; c += (uint64_t)gridDim.x * 256) {
    const uint32_t NN = (uint32_t)N[c], nn = (uint32_t)n[c];
    const double d = (LF[NN] - LF[nn]) - LF[NN - nn];
    out[c] = dto_tail_exact(LF, NN, (uint32_t)K[c], nn, (uint32_t)k[c], d, tab, nullptr);
  }
}

/* exp on the device, for the bit-exactness check against the host library. */
static __global__ __launch_bounds__(256) void k_exp(const double *__restrict__ x, uint64_t count, double *__restrict__ out) {
  __shared__ uint64_t tab[256];
  tab[threadIdx.x] = kExpTabC[threadIdx.x];
  __syncthreads();
  for (uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x; c < count; c += (uint64_t)gridDim.x * 256)
    out[c] = dto_exp(x[c], tab);
}

/* Philox shuffle on the device (checks that host and device draw the same permutation). */
static __global__ __launch_bounds__(256) void k_shuffle(uint64_t seed, uint32_t job, uint32_t perm, uint32_t list,
                                                 uint32_t n, uint32_t *__restrict__ out) {
  Shuffle s = shuffle_make(seed, job, perm, list, n);
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) out[k] = shuffle_index(s, k);
}

}  // namespace dto

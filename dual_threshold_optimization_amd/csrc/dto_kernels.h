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

/* Work counters and cycle stamps of the screened kernels (profiling builds
 * only: make STATS=1); off, they compile to nothing. */
#ifndef DTO_STATS
#define DTO_STATS 0
#endif

namespace dto {

static __constant__ uint64_t kExpTabC[256] = DTO_EXP_TAB_INIT;

constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
/* The wave's index in the workgroup, as a wave-uniform (scalar) value. */
__device__ __forceinline__ uint32_t wave_id() { return (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

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
/* a <- the better of a and b, with the cells of both counted: cells at the
 * winning p add up, and so do, of those, the cells at the winning k.  Every
 * cell enters exactly one key once, so the counts are exact. */
__device__ __forceinline__ void merge_best(BestKey &a, const BestKey &b) {
  if (b.ij == kNone) return;
  if (a.ij == kNone) {
    a = b;
    return;
  }
  if (b.p != a.p) {
    if (b.p < a.p) a = b;
    return;
  }
  const uint32_t np = a.np + b.np;
  if (b.k > a.k) {
    a = b;
  } else if (b.k == a.k) {
    a.npk += b.npk;
    a.ij = b.ij < a.ij ? b.ij : a.ij;
  }
  a.np = np;
}
__device__ __forceinline__ void take_better(BestKey &b, double p, uint32_t k, uint32_t ij) {
  merge_best(b, BestKey{p, k, ij, 1u, 1u});
}
__device__ __forceinline__ BestKey no_best() { return BestKey{__builtin_inf(), 0u, kNone, 0u, 0u}; }

__device__ __forceinline__ BestKey wave_best(BestKey b) {
  for (int off = 32; off >= 1; off >>= 1) {
    BestKey o;
    o.p = __shfl_xor(b.p, off);
    o.k = __shfl_xor(b.k, off);
    o.ij = __shfl_xor(b.ij, off);
    o.np = __shfl_xor(b.np, off);
    o.npk = __shfl_xor(b.npk, off);
    merge_best(b, o);
  }
  return b;
}

/* Wave-wide scans with DPP (gfx9 family): row_shr 1, 2, 4, 8 within each
 * 16-lane row, then row_bcast:15 and row_bcast:31 carry the row totals. */
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false); /* row_shr:1 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false); /* row_shr:2 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false); /* row_shr:4 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false); /* row_shr:8 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
  return x;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) { return wave_incl_scan(v) - v; }

/* Order-preserving map between float bit patterns and ints (an involution,
 * exact for every non-NaN value), so that wave minima are integer minima:
 * v_min_i32 takes its DPP operand directly and needs no NaN quieting. */
__device__ __forceinline__ int ord_key(int b) { return b ^ ((b >> 31) & 0x7FFFFFFF); }

/* Minimum over the wave of each of R values (never NaN), returned in every
 * lane; the R reductions are interleaved so that their DPP latencies
 * overlap. */
template <int R>
__device__ __forceinline__ void wave_minf_n(float *v) {
  constexpr int idn = 0x7FFFFFFF; /* identity of the integer minimum */
  int x[R];
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = ord_key(__float_as_int(v[r]));
#define DTO_MIN_STEP(ctrl, rmask)                                                                                  \
  _Pragma("unroll") for (int r = 0; r < R; ++r) x[r] =                                                            \
      min(x[r], __builtin_amdgcn_update_dpp(idn, x[r], ctrl, rmask, 0xf, false));
  DTO_MIN_STEP(0x111, 0xf) /* row_shr:1 */
  DTO_MIN_STEP(0x112, 0xf) /* row_shr:2 */
  DTO_MIN_STEP(0x114, 0xf) /* row_shr:4 */
  DTO_MIN_STEP(0x118, 0xf) /* row_shr:8 */
  DTO_MIN_STEP(0x142, 0xa) /* row_bcast:15 */
  DTO_MIN_STEP(0x143, 0xc) /* row_bcast:31 */
#undef DTO_MIN_STEP
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = __int_as_float(ord_key(__builtin_amdgcn_readlane(x[r], 63)));
}

/* Minimum over the wave, returned in every lane (fminf: a NaN lane is
 * ignored). */
__device__ __forceinline__ float wave_minf(float v) {
  const int inf = __float_as_int(__builtin_inff());
  int x = __float_as_int(v);
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x111, 0xf, 0xf, false))));
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x112, 0xf, 0xf, false))));
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x114, 0xf, 0xf, false))));
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x118, 0xf, 0xf, false))));
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x142, 0xa, 0xf, false))));
  x = __float_as_int(fminf(__int_as_float(x), __int_as_float(__builtin_amdgcn_update_dpp(inf, x, 0x143, 0xc, 0xf, false))));
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

/* Exclusive scan of data[0..len) in LDS, total written to data[len]. */
template <int NT>
__device__ __forceinline__ void block_excl_scan(uint32_t *data, uint32_t len, uint32_t *wave_sums) {
  const uint32_t tid = threadIdx.x, w = wave_id();
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
  uint16_t *pairs;
  uint16_t *nb;
  uint32_t *rowstart, *cursor, *cw;
  float *rowmin;
  uint32_t *coln;
  float *colb;
  uint2 *cand0;
  uint2 *cand;
  uint4 *cand3;
  uint64_t *tab;
  float *redU;
  BestKey *redB; /* [3][W]: p == 0 class, p == 1 class, evaluated cells */
  uint32_t *redS;
  int *ushare; /* the reverse sweep's best upper bound of the minimum log p, as an ordered key */
};

/* NBL false: the bins array is the workgroup's slice of a global buffer
 * (gnb), which leaves the LDS to the screen table on larger jobs. */
template <bool NBL = true>
__device__ __forceinline__ Lds lds_view(unsigned char *smem, const LdsLayout &Lo, int W, uint16_t *gnb = nullptr) {
  Lds s;
  /* the screen table and the pair staging start the dynamic region (their
   * layout offsets are always 0), so table addresses are the scaled indices */
  s.h = (float *)smem;
  s.pairs = (uint16_t *)smem;
  s.nb = NBL ? (uint16_t *)(smem + Lo.nb) : gnb;
  s.rowstart = (uint32_t *)(smem + Lo.rowstart);
  s.cursor = (uint32_t *)(smem + Lo.cursor);
  s.cw = (uint32_t *)(smem + Lo.cw);
  s.rowmin = (float *)(smem + Lo.rowmin);
  s.coln = (uint32_t *)(smem + Lo.coln);
  s.colb = (float *)(smem + Lo.colb);
  s.cand0 = (uint2 *)(smem + Lo.cand0);
  s.cand = (uint2 *)(smem + Lo.cand);
  s.cand3 = (uint4 *)(smem + Lo.cand3);
  s.tab = (uint64_t *)(smem + Lo.exptab);
  s.redU = (float *)(smem + Lo.red);
  s.redB = (BestKey *)(smem + Lo.red + 16 * W);
  s.redS = (uint32_t *)(smem + Lo.red + 16 * W + 3 * W * sizeof(BestKey));
  s.ushare = (int *)(smem + Lo.misc) + 2;
  return s;
}

/* Bins array access: LDS, or (NBL false) the workgroup's global slice.
 * The global slice is written and read only by the workgroup's own waves,
 * and every write is followed by a barrier (a workgroup-scope release and
 * acquire) before another wave reads it, so plain loads and stores are
 * ordered by the memory model; volatile ones would wait for each load. */
template <bool NBL>
__device__ __forceinline__ uint32_t nb_ld(const Lds &S, uint32_t i) {
  return S.nb[i];
}
template <bool NBL>
__device__ __forceinline__ void nb_st(const Lds &S, uint32_t i, uint16_t v) {
  S.nb[i] = v;
}

/* Order-preserving map between floats and ints (an involution), so that an
 * integer atomic minimum keeps the smallest float. */
__device__ __forceinline__ int fkey(float f) {
  const int b = __float_as_int(f);
  return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float fkey_value(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF)); }

/* Shares a wave's upper bound of the minimum log p with the other waves. */
__device__ __forceinline__ void share_bound(const Lds &S, float U) {
  if (lane_id() == 0) atomicMin(S.ushare, fkey(U));
}

/* One screen table entry by byte offset.  In LDS the table starts the
 * dynamic region, which is LDS address 0 in every kernel that reads it (none
 * of them declares static LDS), so the offset is the LDS address itself and
 * needs no base register or add. */
template <bool HLDS>
__device__ __forceinline__ float htab(const char *Hb, uint32_t off) {
  if (HLDS) return *(const __attribute__((address_space(3))) float *)off;
  return *(const float *)(Hb + off);
}

/* Copies n floats global -> LDS (both 16-byte aligned) with LDS DMA
 * (global_load_lds_dwordx4): lane l of a wave writes 16 bytes at the wave's
 * LDS base + 16 l, and the data never passes through VGPRs.  Lanes past the
 * end are masked off, so nothing beyond dst[n) is written.  Ends with the
 * wave's loads complete; the caller's barrier publishes them. */
template <int NT>
__device__ __forceinline__ void copy_table(float *dst, const float *src, uint32_t n) {
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const uint32_t n4 = n / 4;
  const float4 *s4 = reinterpret_cast<const float4 *>(src);
  float4 *d4 = reinterpret_cast<float4 *>(dst);
  for (uint32_t c0 = w * 64u; c0 < n4; c0 += NT) {
    const uint32_t t = c0 + lane;
    if (t < n4)
      __builtin_amdgcn_global_load_lds((const void *)(s4 + t), (__attribute__((address_space(3))) void *)(d4 + c0),
                                       16, 0, 0);
  }
  for (uint32_t t = n4 * 4 + tid; t < n; t += NT) dst[t] = src[t];
  __builtin_amdgcn_s_waitcnt(0);
}

/* Per-job tables that stay in LDS while a workgroup works on that job. */
template <int NT, bool HLDS>
__device__ __forceinline__ void load_job_tables(const DevJob &J, const Lds &S) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t t = tid; t < 256; t += NT) S.tab[t] = kExpTabC[t];
  if (HLDS) copy_table<NT>(S.h, J.H32, J.N + 1);
  for (uint32_t t = tid; t < J.T2; t += NT) {
    S.coln[t] = J.n[t];
    S.colb[t] = J.B32[t];
  }
}

/* Visits the Philox permutation of one list, idx[k] for every k < n, spread
 * over the workgroup.  The domain goes in chunks of NT * kShufU consecutive
 * elements, each thread evaluating kShufU consecutive ones; a workgroup
 * prefix count (wave scan + per-wave totals in wsum[W]) ranks those that
 * fall below n.  Every thread runs every chunk and hands fn its kShufU
 * (rank, value) pairs at once, fn(rank[], value[]), value >= n marking an
 * element outside [0, n), so that fn can issue all its loads together. */
constexpr uint32_t kShufU = 16;
template <int NT, typename Fn>
__device__ __forceinline__ void shuffle_visit(const Shuffle &s, uint32_t *wsum, Fn fn) {
  constexpr uint32_t W = NT / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  uint32_t carried = 0;
  for (uint32_t base = 0; base < s.dom; base += NT * kShufU) {
    uint32_t v[kShufU], rk[kShufU], cnt = 0;
#pragma unroll
    for (uint32_t u = 0; u < kShufU; ++u) {
      const uint32_t p = base + tid * kShufU + u;
      v[u] = p < s.dom ? feistel_once(s, p) : s.n;
      rk[u] = cnt; /* rank among this thread's elements below n */
      cnt += v[u] < s.n ? 1u : 0u;
    }
    const uint32_t ex = wave_excl_scan(cnt);
    if (lane == 63) wsum[w] = ex + cnt;
    __syncthreads();
    uint32_t r = carried + ex, total = 0;
    for (uint32_t x = 0; x < W; ++x) {
      const uint32_t t = wsum[x];
      r += x < w ? t : 0u;
      total += t;
    }
    __syncthreads();
    carried += total;
#pragma unroll
    for (uint32_t u = 0; u < kShufU; ++u) rk[u] += r;
    fn(rk, v);
  }
}

/* Phases A-B for one permutation: leaves the row-sorted list2 bins in S.nb,
 * row offsets in S.rowstart and zeroed per-wave column counts.  Each list2
 * position's pair is staged in `pairs`: with HLDS in the screen-table region
 * (reloaded afterwards), which holds the whole pair (a << 16 | b) as a u32;
 * otherwise (LDS next to the bins, or a per-workgroup global buffer) only
 * its row a as a u16, and the column b, the position's own bin, is read
 * again from bin2 when the pairs are sorted. */
template <int NT, bool EXPLICIT, bool HLDS, bool PL, bool NBL = true>
__device__ __forceinline__ void build_rows(const DevJob &J, const LdsLayout &Lo, const Lds &S, const uint32_t *i1,
                                           const uint32_t *i2, const Shuffle &s1, const Shuffle &s2, int *err,
                                           uint16_t *gpairs) {
  constexpr int W = NT / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  const uint32_t T1 = J.T1, T2 = J.T2, T2wp = Lo.T2w_pad;
  uint16_t *lpairs = S.pairs;
  uint16_t *vpairs = gpairs; /* global staging, read back by other waves after a barrier */
  for (uint32_t t = tid; t <= T1; t += NT) S.rowstart[t] = 0;
  for (uint32_t t = tid; t < W * T2wp; t += NT) S.cw[t] = 0;
  if (tid == 0) *S.ushare = fkey(__builtin_inff()); /* published after the build's barriers */
  /* the tables this build reads, each held as its own scalar pair: read
   * straight from the job they are one spilled tuple with the job's other
   * pointers, reloaded whole (16 lane reads) at every use */
  const uint16_t *bin1 = J.bin1, *bin2 = J.bin2;
  const int32_t *join = J.join;
  asm("" : "+s"(bin1), "+s"(bin2), "+s"(join));
  /* A: list1 feature m sits at the position k with idx1[k] == m (loads of
   * kU positions in flight per thread) */
  constexpr uint32_t kU = 16;
  if (!EXPLICIT) {
    shuffle_visit<NT>(s1, S.redS, [&](const uint32_t *k, const uint32_t *m) {
      uint32_t bn[kShufU];
#pragma unroll
      for (uint32_t u = 0; u < kShufU; ++u) bn[u] = m[u] < J.N1 ? bin1[k[u]] : 0u;
#pragma unroll
      for (uint32_t u = 0; u < kShufU; ++u)
        if (m[u] < J.N1) nb_st<NBL>(S, m[u], (uint16_t)bn[u]);
    });
  } else {
    for (uint32_t k0 = tid; k0 < J.N1; k0 += NT * kU) {
      uint32_t m[kU], bn[kU];
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) {
        const uint32_t k = k0 + u * NT;
        m[u] = 0u;
        bn[u] = 0u;
        if (k < J.N1) {
          m[u] = i1 ? i1[k] : k;
          bn[u] = bin1[k];
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) {
        if (k0 + u * NT < J.N1) {
          if (m[u] < J.N1)
            nb_st<NBL>(S, m[u], (uint16_t)bn[u]);
          else
            atomicOr(err, 1);
        }
      }
    }
  }
  if (!NBL) __threadfence_block();
  __syncthreads();
  /* B1: (a, b) per list2 position, counted per row a */
  auto pair_of = [&](uint32_t e, uint32_t q, uint32_t b, int32_t jm) {
    uint32_t a = kBinNone;
    if (!J.dup) {
      if (jm >= 0) a = nb_ld<NBL>(S, (uint32_t)jm);
    } else {
      /* intersect_genes collapses duplicate ids of genes1: the entry is in
       * S1'(t) as soon as any list1 entry with its id is */
      for (uint32_t t = J.join_off[q]; t < J.join_off[q + 1]; ++t) {
        uint32_t v = nb_ld<NBL>(S, J.join_pos[t]);
        a = v < a ? v : a;
      }
    }
    uint32_t pr = HLDS ? kNone : kBinNone;
    if (a < T1 && b < T2) {
      pr = HLDS ? (a << 16) | b : a;
      atomicAdd(&S.rowstart[a], 1u);
    }
    if (HLDS)
      ((uint32_t *)lpairs)[e] = pr;
    else if (PL)
      lpairs[e] = (uint16_t)pr;
    else
      vpairs[e] = (uint16_t)pr;
  };
  if (!EXPLICIT) {
    shuffle_visit<NT>(s2, S.redS, [&](const uint32_t *e, const uint32_t *q) {
      uint32_t b[kShufU];
      int32_t jm[kShufU];
#pragma unroll
      for (uint32_t u = 0; u < kShufU; ++u) {
        const bool in = q[u] < J.N2;
        b[u] = in ? bin2[e[u]] : kBinNone;
        jm[u] = (in && !J.dup) ? join[q[u]] : -1;
      }
#pragma unroll
      for (uint32_t u = 0; u < kShufU; ++u)
        if (q[u] < J.N2) pair_of(e[u], q[u], b[u], jm[u]);
    });
  } else {
    for (uint32_t e0 = tid; e0 < J.N2; e0 += NT * kU) {
      uint32_t q[kU], b[kU];
      int32_t jm[kU];
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) {
        const uint32_t e = e0 + u * NT;
        q[u] = kNone;
        b[u] = kBinNone;
        if (e < J.N2) {
          q[u] = i2 ? i2[e] : e;
          b[u] = bin2[e];
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) jm[u] = (q[u] < J.N2 && !J.dup) ? join[q[u]] : -1;
#pragma unroll
      for (uint32_t u = 0; u < kU; ++u) {
        const uint32_t e = e0 + u * NT;
        if (e >= J.N2) continue;
        if (q[u] >= J.N2)
          atomicOr(err, 1);
        else
          pair_of(e, q[u], b[u], jm[u]);
      }
    }
  }
  if (!PL) __threadfence_block();
  __syncthreads();
  block_excl_scan<NT>(S.rowstart, T1, S.redS);
  for (uint32_t t = tid; t < T1; t += NT) S.cursor[t] = S.rowstart[t];
  __syncthreads();
  /* B2: scatter b into row order (overwrites the list1 bins, all read above);
   * kU rows and bins are loaded per thread before the first atomic, so the
   * global loads of bin2 overlap */
  for (uint32_t e0 = tid; e0 < J.N2; e0 += NT * kU) {
    uint32_t a[kU];
    uint16_t b[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      const uint32_t e = e0 + u * NT;
      a[u] = kBinNone;
      b[u] = 0;
      if (e < J.N2) {
        if (HLDS) {
          const uint32_t pr = ((const uint32_t *)lpairs)[e];
          a[u] = pr == kNone ? kBinNone : pr >> 16;
          b[u] = (uint16_t)(pr & 0xFFFFu);
        } else {
          a[u] = PL ? lpairs[e] : vpairs[e];
          b[u] = bin2[e];
        }
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; ++u) {
      if (a[u] != kBinNone) {
        const uint32_t pos = atomicAdd(&S.cursor[a[u]], 1u);
        nb_st<NBL>(S, pos, b[u]);
      }
    }
  }
  if (!NBL) __threadfence_block();
  __syncthreads();
  if (HLDS && PL) {
    /* the pairs used the screen-table region: put the table back */
    copy_table<NT>(S.h, J.H32, J.N + 1);
    __syncthreads();
  }
}

/* Adds or removes the pairs of rows [ia, ib) (contiguous in S.nb) into a
 * wave's column counts (four bins in flight per lane).  No wait at the end:
 * the wave's later LDS reads of C are processed after these updates (the
 * LDS executes one wave's operations in order). */
template <bool NBL = true>
__device__ __forceinline__ void rows_update(const Lds &S, uint32_t *C, uint32_t ia, uint32_t ib, bool add) {
  const uint32_t lane = lane_id();
  const uint32_t e1 = S.rowstart[ib];
  for (uint32_t e = S.rowstart[ia] + lane; e < e1; e += 256) {
    uint32_t b[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) b[u] = nb_ld<NBL>(S, min(e + 64u * u, e1 - 1u));
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      if (e + 64u * u < e1) {
        const uint32_t inc = 1u << ((b[u] & 1u) * 16u);
        if (add)
          atomicAdd(&C[b[u] >> 1], inc);
        else
          atomicSub(&C[b[u] >> 1], inc);
      }
    }
  }
}

/* Software-pipelined form of rows_update for the forward sweep: the bins of
 * the first kPreBins * 64 pairs of the next update are read into registers
 * while the current row is scanned, so that the update itself issues its
 * adds without waiting for a load. */
constexpr uint32_t kPreBins = 5;

template <bool NBL = true>
__device__ __forceinline__ void prefetch_bins(const Lds &S, uint32_t e0, uint32_t e1, uint32_t *pb) {
  const uint32_t lane = lane_id(), last = e1 > e0 ? e1 - 1u : 0u;
#pragma unroll
  for (uint32_t u = 0; u < kPreBins; ++u) pb[u] = nb_ld<NBL>(S, min(e0 + lane + 64u * u, last));
}

template <bool NBL = true>
__device__ __forceinline__ void add_bins(const Lds &S, uint32_t *C, uint32_t e0, uint32_t e1, const uint32_t *pb) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t u = 0; u < kPreBins; ++u)
    if (e0 + lane + 64u * u < e1) atomicAdd(&C[pb[u] >> 1], 1u << ((pb[u] & 1u) * 16u));
  /* pairs past the prefetched ones (rows with many pairs) */
  for (uint32_t e = e0 + 64u * kPreBins + lane; e < e1; e += 64u) {
    const uint32_t b = nb_ld<NBL>(S, e);
    atomicAdd(&C[b >> 1], 1u << ((b & 1u) * 16u));
  }
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

/* t_{i+1} / t_i for consecutive terms of the tail (exact mathematics):
 * (K-i)(n-i) / ((i+1)(N-K-n+i+1)), in FP32.  Relative error <= 7 * 2^-24
 * (two products, one 1-ulp reciprocal, one product, and the int->float
 * conversions, exact below 2^24). */
__device__ __forceinline__ float term_ratio(uint32_t N, uint32_t K, uint32_t n, uint32_t i) {
  float a = (float)(K - i) * (float)(n - i);
  float b = (float)(i + 1) * (float)((N - K) - (n - i) + 1);
  return a * __builtin_amdgcn_rcpf(b);
}

/* Lower bound of log S, S = sum_{i >= k} t_i / t_k, from one ratio: the
 * ratios fall with i, so each of the first M+1 terms is at least r^l with
 * r = ratio at k+M-1, and S >= (1 - r^(M+1)) / (1 - r).  FP32 rounding is
 * covered by taking 1e-3 off the result. */
__device__ __forceinline__ float tail_lower_log(uint32_t N, uint32_t K, uint32_t n, uint32_t k) {
  const uint32_t mx = K < n ? K : n;
  const uint32_t M = mx - k < 16u ? mx - k : 16u;
  if (M == 0u) return 0.f;
  const float r = term_ratio(N, K, n, k + M - 1u) * (1.f - 16.f * 0x1p-24f);
  float S;
  if (r >= 0.9999f)
    S = 0.99f * (float)(M + 1u); /* each term >= 0.9999^16 > 0.99 */
  else
    S = (1.f - exp2f((float)(M + 1u) * __log2f(r))) / (1.f - r);
  return fmaxf(0.f, __logf(S) - 1e-3f);
}

/* Bracket of S = sum_{i >= k} t_i / t_k from the ratio recurrence: up to mmax
 * terms in FP32 (fewer once the rest is below 2^-16 S), then a geometric
 * bound of the rest (the ratio only falls with i).  Accumulated relative
 * error of the products and the sum is at most 9 (m+1) 2^-24; hi is +inf
 * when the terms are still growing at the cut.  The sum also stops once it
 * passes s_stop (both ends stay valid bounds). */
struct TailBracket {
  float lo, hi;
};
__device__ __forceinline__ TailBracket tail_bracket(uint32_t N, uint32_t K, uint32_t n, uint32_t k,
                                                    uint32_t mmax, float s_stop) {
  const uint32_t mx = K < n ? K : n;
  float S = 1.f, P = 1.f;
  uint32_t i = k, m = 0;
  while (i < mx && m < mmax) {
    const float r = term_ratio(N, K, n, i);
    P *= r;
    S += P;
    ++i;
    ++m;
    /* past s_stop the caller needs no more than the lower end */
    if (P > 1e30f || S > s_stop) break;
    /* the geometric bound of the rest (ratios only fall) is below 2^-16 S */
    if (r < 0.999f && P * r <= S * (1.f - r) * 0x1p-16f) break;
  }
  const float eps = (float)(9u * (m + 1u)) * 0x1p-24f + 1e-7f;
  float rem = 0.f;
  if (i < mx) {
    const float r = term_ratio(N, K, n, i) * (1.f + 8.f * 0x1p-24f);
    rem = (r < 0.999f && P <= 1e30f) ? P * (1.f + eps) * r / (1.f - r) * 1.001f : __builtin_inff();
  }
  TailBracket t;
  t.lo = S * (1.f - eps);
  t.hi = (S + rem) * (1.f + eps);
  return t;
}

/* Upper bound of the log of the reference's computed sum for a sum-class
 * cell, from its FP32 first-term log x. */
__device__ __forceinline__ float screen_upper(const DevJob &J, uint32_t K, uint32_t n, uint32_t k, float x,
                                              uint32_t mmax = 64, float s_stop = __builtin_inff()) {
  const uint32_t mx = K < n ? K : n;
  const TailBracket tb = tail_bracket(J.N, K, n, k, mmax, s_stop);
  if (!(tb.hi < __builtin_inff())) return 1e-6f; /* p <= 1 (+ rounding) */
  float U = x + J.margin + __logf(tb.hi) + 1e-4f;
  if (U < -690.f) U = fmaxf(U, __logf((float)(mx - k + 1)) - 744.44f) + 0.7f; /* subnormal term rounding */
  return U;
}

/* The reference's first tail exponent for (K, n, k), FP64 in statrs order. */
__device__ __forceinline__ double first_exponent(const DevJob &J, uint32_t K, uint32_t n, uint32_t k, double d) {
  const double *LF = J.LF;
  const uint32_t N = J.N;
  double lb1 = (LF[K] - LF[k]) - LF[K - k];
  double lb2 = (LF[N - K] - LF[n - k]) - LF[(N - K) - (n - k)];
  return (lb1 + lb2) - d;
}

struct WaveStats {
  uint32_t cells, terms, rows;
  uint64_t cyc[5]; /* build, forward sweep, reverse sweep, tier-2 brackets (flush_candidates), exact */
  uint32_t push1, push2, flushes; /* tier-1 survivors, tier-2 survivors, flush_candidates calls */
  uint32_t gap;                    /* sum of (first upper bound - log of the minimum), milli-nats */
  uint32_t cheap;                  /* cells passing the first-term test of the reverse sweep (per lane) */
};

__device__ __forceinline__ uint64_t cycles() {
#if DTO_STATS
  return clock64();
#else
  return 0;
#endif
}

/* dto_tail_exact for one cell with the whole wave (wave-uniform arguments,
 * result in every lane): lane l evaluates term k + 64 b + l of block b, so a
 * block's table loads and exps run side by side; the terms are then added
 * in ascending order, one at a time, with the same early exit, so the sum
 * is bit-identical to dto_tail_exact's.  Used when a wave has only a few
 * exact cells, where one lane per cell would leave the wave (and the
 * workgroup, at the next barrier) waiting on that lane's serial terms. */
__device__ __forceinline__ double wave_tail_exact(const DevJob &J, const Lds &S, uint32_t K, uint32_t n, uint32_t k,
                                                  double d, uint32_t *terms) {
  *terms = 0;
  const uint32_t N = J.N;
  if (k == 0) return 1.0;
  const uint32_t x = k - 1;
  const uint32_t mn = (K + n > N) ? K + n - N : 0;
  const uint32_t mx = K < n ? K : n;
  if (x < mn) return 1.0;
  if (x >= mx) return 0.0;
  const double *LF = J.LF;
  const double lfK = LF[K], lfNK = LF[N - K];
  const uint32_t NK = N - K, lane = lane_id();
  double acc = 0.0;
  for (uint32_t i0 = k; i0 <= mx; i0 += 64) {
    const uint32_t i = i0 + lane;
    double t = 0.0;
    bool fall = false; /* rho_i <= 1 - 2^-20, as num <= den - ceil(den / 2^20) */
    if (i <= mx) {
      const double lb1 = (lfK - LF[i]) - LF[K - i];
      const double lb2 = (lfNK - LF[n - i]) - LF[NK - (n - i)];
      t = dto_exp((lb1 + lb2) - d, S.tab);
      const uint64_t num = (uint64_t)(K - i) * (uint64_t)(n - i);
      const uint64_t den = (uint64_t)(i + 1) * (uint64_t)(NK - n + i + 1);
      fall = num <= den - ((den + 0xFFFFFull) >> 20);
    }
    const uint64_t fm = __ballot(fall);
    const uint32_t cnt = min(64u, mx - i0 + 1u);
    const int tlo = __double2loint(t), thi = __double2hiint(t);
    for (uint32_t l = 0; l < cnt; ++l) {
      const double tl =
          __hiloint2double(__builtin_amdgcn_readlane(thi, (int)l), __builtin_amdgcn_readlane(tlo, (int)l));
      acc = acc + tl;
      if (tl * 0x1p56 <= acc && ((fm >> l) & 1ull)) {
        *terms = i0 + l + 1 - k;
        return acc;
      }
    }
  }
  *terms = mx + 1 - k;
  return acc;
}

/* Exact FP64 evaluation (bit-identical to the reference) of the tier-3 cells
 * whose lower bound is <= thr. */
__device__ __forceinline__ void flush_exact(const DevJob &J, const Lds &S, uint4 *c3, uint32_t &n3, float thr, BestKey &xb,
                            WaveStats &st) {
  const uint32_t lane = lane_id();
  wave_sync();
  BestKey lb = no_best();
  uint32_t cells = 0;
  static_assert(kCand3Cap <= 64, "one lane per listed cell");
  const uint32_t nex = (uint32_t)__popcll(__ballot(lane < n3 && __uint_as_float(c3[lane].z) <= thr));
  if (nex <= 8u) {
    /* few cells: each with the whole wave */
    for (uint32_t c = 0; c < n3; ++c) {
      const uint4 cd = c3[c];
      if (__uint_as_float(cd.z) > thr) continue;
      const uint32_t i = cd.x >> 16, j = cd.x & 0xFFFFu, k = cd.y;
      uint32_t nt = 0;
      const double p = wave_tail_exact(J, S, J.K[i], S.coln[j], k, J.D[j], &nt);
      if (lane == 0) {
        /* every lane holds p: one lane enters the cell, so it counts once */
        st.terms += nt;
        ++cells;
        take_better(lb, p, k, cd.x);
      }
    }
  } else {
    for (uint32_t c = lane; c < n3; c += 64) {
      const uint4 cd = c3[c];
      if (__uint_as_float(cd.z) > thr) continue;
      const uint32_t i = cd.x >> 16, j = cd.x & 0xFFFFu, k = cd.y;
      uint32_t nt = 0;
      double p = dto_tail_exact(J.LF, J.N, J.K[i], S.coln[j], k, J.D[j], S.tab, &nt);
      st.terms += nt;
      ++cells;
      take_better(lb, p, k, cd.x);
    }
  }
  if (DTO_STATS) {
    for (int off = 32; off >= 1; off >>= 1) cells += __shfl_xor(cells, off);
    st.cells += cells;
  }
  lb = wave_best(lb);
  merge_best(xb, lb);
  n3 = 0;
  wave_sync();
}

__device__ __forceinline__ float log_bound_of(const BestKey &xb) {
  if (xb.ij == kNone) return __builtin_inff();
  return xb.p > 0.0 ? (float)(log(xb.p) + 1e-3) : -__builtin_inff();
}

/* Tier 2 for a full tier-1 list: every lane brackets one cell's log p with
 * the exact FP64 first exponent and the FP32 ratio recurrence; cells whose
 * lower end is within the wave's best upper end go to the tier-3 list. */
template <int NT>
__device__ __forceinline__ void flush_candidates(const DevJob &J, const Lds &S, uint2 *cand, uint32_t &nc, uint4 *c3,
                                 uint32_t &n3, BestKey &xb, float &Uw, float &U2w, WaveStats &st) {
  constexpr int PER = (kCandCap + 63) / 64;
  const uint32_t lane = lane_id();
  wave_sync();
  float L2[PER];
  float U2min = __builtin_inff();
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const uint32_t c = lane + 64u * r;
    L2[r] = __builtin_inff();
    if (c < nc) {
      const uint2 cd = cand[c];
      const uint32_t i = cd.x >> 16, j = cd.x & 0xFFFFu, k = cd.y;
      const uint32_t K = J.K[i], n = S.coln[j];
      const uint32_t mx = K < n ? K : n;
      const double x64 = first_exponent(J, K, n, k, J.D[j]);
      /* a sum above s_stop already puts the lower end above U2w (the
       * 2e-3 covers the FP32 rounding of the exponent and eps <= 5.5e-4) */
      const float s_stop = __expf(fminf(U2w - (float)x64 + 2e-3f, 80.f));
      const TailBracket tb = tail_bracket(J.N, K, n, k, 1024, s_stop);
      const double lo = x64 + (double)__logf(tb.lo) - 1e-4;
      L2[r] = lo < -708.0 ? -__builtin_inff() : (float)(lo - 1e-4);
      if (tb.hi < __builtin_inff()) {
        double hi = x64 + (double)__logf(tb.hi) + 1e-4;
        if (hi < -690.0) hi = fmax(hi, log((double)(mx - k + 1)) - 744.44) + 0.7;
        U2min = fminf(U2min, (float)(hi + 1e-4));
      } else {
        U2min = fminf(U2min, 1e-6f);
      }
    }
  }
  U2w = fminf(U2w, wave_minf(U2min));
  Uw = fminf(Uw, U2w);
  share_bound(S, U2w);
  /* survivors move to the tier-3 list half a wave at a time: at most 32
   * entries per push, so a push after a flush always fits the list */
  static_assert(kCand3Cap >= 32, "a half-wave push must fit an empty tier-3 list");
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const uint32_t c = lane + 64u * r;
    const bool f = c < nc && L2[r] <= U2w;
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
      const bool fh = f && (lane >> 5) == half;
      const uint64_t m = __ballot(fh);
      const uint32_t cnt = (uint32_t)__popcll(m);
      if (cnt) {
        if (n3 + cnt > (uint32_t)kCand3Cap) {
          flush_exact(J, S, c3, n3, U2w, xb, st);
          U2w = fminf(U2w, log_bound_of(xb));
          Uw = fminf(Uw, U2w);
          share_bound(S, U2w);
        }
        if (fh) {
          const uint2 cd = cand[c];
          c3[n3 + mask_lt(m)] = make_uint4(cd.x, cd.y, __float_as_uint(L2[r]), 0u);
        }
        n3 += cnt;
        st.push2 += cnt;
      }
    }
  }
  nc = 0;
  st.flushes++;
  wave_sync();
}

/* True when a cell's p is exactly 0.0 in the reference: k > min(K, n) (the
 * tail is empty), or every term from k on is below the exp underflow point
 * (-745.1332...): x + margin < -745.2 bounds the first exponent and a ratio
 * below 1 makes every later term smaller. */
__device__ __forceinline__ bool exact_zero(const DevJob &J, uint32_t K, uint32_t n, uint32_t k, float x) {
  const uint32_t N = J.N;
  if (k > min(K, n)) return true;
  if (!(x + J.margin < -745.2f)) return false;
  const uint64_t num = (uint64_t)(K - k) * (uint64_t)(n - k);
  const uint64_t den = (uint64_t)(k + 1) * (uint64_t)((N - K) - (n - k) + 1);
  return num < den;
}

/* The reverse sweep's first-term survivors, in full-lane batches: each lane
 * takes one cell, rebuilds its screen value, drops it if its p is exactly 0
 * (already in the zero class), and applies the tier-1 bound (first term
 * minus the screen margin plus a lower bound of the tail, -inf below -708
 * where sums may be subnormal); cells within Uw move to the tier-1 list,
 * which is flushed through tier 2 whenever it would overflow. */
template <int NT, bool HLDS>
__device__ __forceinline__ void flush_first(const DevJob &J, const Lds &S, const char *Hb, uint2 *cand0, uint32_t &n0,
                                            uint2 *cand, uint32_t &nc, uint4 *c3, uint32_t &n3, BestKey &xb,
                                            float &Uw, float &U2w, WaveStats &st, uint32_t top, uint32_t Kl,
                                            float Al) {
  constexpr uint32_t W = NT / 64;
  const uint32_t lane = lane_id(), N = J.N;
  wave_sync();
#pragma unroll 1
  for (uint32_t r0 = 0; r0 < n0; r0 += 64) {
    bool f = false;
    const bool act = r0 + lane < n0;
    uint2 cd = act ? cand0[r0 + lane] : make_uint2(top << 16, 0u);
    /* every listed cell is in a row of the current batch (row top - l W is
     * lane l's): the row constants come from that lane's registers */
    const int src = (int)((top - (cd.x >> 16)) / W);
    const uint32_t K = (uint32_t)__shfl((int)Kl, src);
    const float A = __shfl(Al, src);
    if (act) {
      const uint32_t j = cd.x & 0xFFFFu, k = cd.y;
      const uint32_t n = S.coln[j];
      if (k <= min(K, n)) {
        const float x = (A + S.colb[j]) -
                        ((htab<HLDS>(Hb, 4u * k) + htab<HLDS>(Hb, 4u * (K - k))) +
                         (htab<HLDS>(Hb, 4u * (n - k)) + htab<HLDS>(Hb, 4u * ((N - K) - (n - k)))));
        if (!(x + J.margin < -745.2f && exact_zero(J, K, n, k, x))) {
          float L = x - J.margin;
          if (L >= -708.f) {
            L += tail_lower_log(N, K, n, k);
            f = L <= Uw;
          } else {
            f = true; /* possibly subnormal: no usable lower bound */
          }
        }
      }
    }
    const uint64_t m = __ballot(f);
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (cnt) {
      if (nc + cnt > (uint32_t)kCandCap) {
        uint64_t tf = cycles();
        flush_candidates<NT>(J, S, cand, nc, c3, n3, xb, Uw, U2w, st);
        st.cyc[3] += cycles() - tf;
      }
      if (f) cand[nc + mask_lt(m)] = cd;
      nc += cnt;
      st.push1 += cnt;
    }
  }
  n0 = 0;
  wave_sync();
}

/* A lane's row segment: kSegWords count words (2 columns each).  With the
 * reference thresholds T2 <= 640 (N <= ~32k), so a row is one segment and the
 * lane's column constants live in registers (REGCOLS); wider grids loop over
 * segments and read the column constants from LDS.  The loads of a segment
 * are issued together, so a segment costs two LDS round trips. */
constexpr uint32_t kSegWords = 5;
constexpr uint32_t kSegCells = 2 * kSegWords;

/* Column constant of a cell outside the sum class or past the last column:
 * its screen value becomes ~3e38, above every threshold, and stays finite so
 * that the packed minimum below never meets a NaN. */
constexpr float kNoCell = 3.0e38f;

/* The low kPackBits mantissa bits of a lane's running minimum carry the
 * index of the cell that holds it; this moves the value by at most
 * 2^kPackBits ulp, i.e. |x| 2^(kPackBits - 23), which every bound derived
 * from a packed value adds back (pack_slack). */
constexpr uint32_t kPackBits = 4;
static_assert(kSegCells <= (1u << kPackBits), "a segment's cell index must fit the packed bits");
__device__ __forceinline__ float pack_cell(float x, uint32_t idx) {
  return __int_as_float((__float_as_int(x) & ~((1 << kPackBits) - 1)) | (int)idx);
}
__device__ __forceinline__ uint32_t packed_cell(float v) {
  return (uint32_t)__float_as_int(v) & ((1u << kPackBits) - 1u);
}
__device__ __forceinline__ float pack_slack(float v) { return fabsf(v) * 0x1p-18f + 1e-30f; }

struct ColRegs {
  uint32_t n4[kSegCells]; /* 4 n_j */
  float b[kSegCells];
  uint32_t wa[kSegWords]; /* LDS byte addresses of the lane's count words */
};

/* Lower tail.  The first term is a useless lower bound of p for a cell below
 * its mean, where p is close to 1 but P(X = k) can be as small as the upper
 * tail's terms.  Cantelli's inequality, P(X <= mu - lam) <= var / (var +
 * lam^2) for lam > 0, gives P(X >= k) >= 1/2 whenever lam = mu - k + 1 >=
 * max(sigma, 1).  With sigma' >= max(sigma_ij, 1) over a lane's cells of a
 * row and D >= N (sigma' - 1), the test k N + D <= K n implies it (k <= mu +
 * 1 - sigma').  Rows held in registers test 4 k N + 4 D <= K (4 n), valid
 * while 4 N^2 + 4 D < 2^32 (N <= kLiftMaxN); wider rows test k N + D <= K n,
 * valid while N^2 + D < 2^32 (N <= kLiftMaxN1).  Larger jobs get no lift
 * (c = -inf). */
constexpr uint32_t kLiftMaxN = 32000;
constexpr uint32_t kLiftMaxN1 = 65000;
constexpr float kLogHalfBelow = -0.69315f; /* below log(1/2) */

/* D for row constant K: var_ij = K (N - K) / N^2 * n (N - n) / (N - 1) and
 * V bounds the last factor over the lane's columns. */
__device__ __forceinline__ uint32_t lift_offset(uint32_t K, uint32_t N, float Nf, float invN2, float V) {
  const float s = (float)K * (float)(N - K) * invN2;
  const float sg = fmaxf(__builtin_sqrtf(s * V) * 1.001f + 1e-3f, 1.f);
  return (uint32_t)ceilf(fminf((sg - 1.f) * Nf + 1.f, 1e9f));
}

/* A cell's screen value raised to log(1/2) (c) when the test above holds.
 * p == 0 cells (k > K >= mu) and p == 1 cells (x ~ +3e38) are unchanged.
 * X4: the scaled test (N4 = 4 N, D4 = 4 D), else the plain one (N4 = N, D4 = D). */
template <bool X4>
__device__ __forceinline__ float lift_lower(float x, uint32_t k, uint32_t n4, uint32_t K, uint32_t N4, uint32_t D4,
                                            float c) {
  const bool below = __umul24(k, N4) + D4 <= __umul24(X4 ? n4 : n4 >> 2, K);
  return fmaxf(x, below ? c : -__builtin_inff());
}

/* LDS byte address of a pointer into the dynamic region. */
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

/* read_segment for a row held in registers (REGCOLS): the lane's word
 * addresses are fixed for the job, and words past the lane's columns point
 * at the wave's zero padding word, so the loads need neither index
 * arithmetic nor masks. */
__device__ __forceinline__ uint32_t read_segment_regs(const ColRegs &cr, uint32_t *wv) {
#pragma unroll
  for (uint32_t u = 0; u < kSegWords; ++u) wv[u] = *(const __attribute__((address_space(3))) uint32_t *)cr.wa[u];
  uint32_t tot = 0;
#pragma unroll
  for (uint32_t u = 0; u < kSegWords; ++u) tot += (wv[u] & 0xFFFFu) + (wv[u] >> 16);
  return tot;
}

/* Reads a segment's count words (all loads issued together) and returns
 * the lane's count total.  Indices are clamped to the wave's last padding
 * word, which is always zero (the padding is at least one word and no pair
 * reaches it), so words past the row need no per-lane mask; only words past
 * the segment (u >= nseg, the same for every lane) are masked. */
__device__ __forceinline__ uint32_t read_segment(const uint32_t *C, uint32_t w0, uint32_t nseg, uint32_t last,
                                                 uint32_t *wv) {
  uint32_t raw[kSegWords];
#pragma unroll
  for (uint32_t u = 0; u < kSegWords; ++u) raw[u] = C[min(w0 + u, last)];
  uint32_t tot = 0;
#pragma unroll
  for (uint32_t u = 0; u < kSegWords; ++u) {
    wv[u] = u < nseg ? raw[u] : 0u;
    tot += (wv[u] & 0xFFFFu) + (wv[u] >> 16);
  }
  return tot;
}

/* The cells of one segment, straight-line: overlap k (prefix of the column
 * counts), column size n (kept as 4n), and the FP32 log x of the first tail
 * term,
 *   x = (A_i + B_j) - (H[k] + H[K-k] + H[n-k] + H[N-K-n+k]),
 * for sum-class cells (max(0, K+n-N) < k <= min(K, n)).  Cells with p == 1
 * (k == 0, or k <= K+n-N, i.e. n-k >= N-K) or no column get x ~ +3e38;
 * cells with k > K (p == 0: an empty tail; k <= n always, since k counts
 * entries of the column's set) get x ~ -3e38, so that a segment minimum
 * below the underflow bound flags every cell that may have p == 0 (ZC
 * false: the job has no such cells, see DevJob.kle, and the test is left
 * out).  Hb is
 * the byte base of the H table; indices are byte offsets (4x), and the two
 * that can wrap for cells of the other classes saturate at 0. */
template <bool HLDS, bool REGCOLS, bool ZC = true>
__device__ __forceinline__ void segment_fast(const DevJob &J, const Lds &S, const char *Hb, const uint32_t *wv,
                                             uint32_t nseg, uint32_t w0, uint32_t run, uint32_t K, uint32_t NK4,
                                             float A, const ColRegs &cr, uint32_t *kk, uint32_t *nn4, float *xx) {
  const uint32_t T2 = J.T2, K4 = K << 2;
  float bc[kSegCells], h0[kSegCells], h1[kSegCells], h2[kSegCells], h3[kSegCells];
  uint32_t d2[kSegCells];
  /* indices and table loads of every cell first (one wait for all) */
#pragma unroll
  for (uint32_t c = 0; c < kSegCells; ++c) {
    run += (c & 1u) ? (wv[c >> 1] >> 16) : (wv[c >> 1] & 0xFFFFu);
    const uint32_t k = run, k4 = k << 2;
    uint32_t n4;
    if (REGCOLS) {
      n4 = cr.n4[c];
      bc[c] = cr.b[c];
    } else {
      const uint32_t j = 2u * w0 + c;
      const bool valid = (c >> 1) < nseg && j < T2;
      const uint32_t jc = valid ? j : T2 - 1u;
      n4 = S.coln[jc] << 2;
      bc[c] = valid ? S.colb[jc] : kNoCell;
    }
    const uint32_t i1 = __builtin_elementwise_sub_sat(K4, k4); /* K - k */
    const uint32_t i2 = n4 - k4;                                /* n - k */
    const uint32_t i3 = __builtin_elementwise_sub_sat(NK4, i2); /* N - K - n + k */
    kk[c] = k;
    nn4[c] = n4;
    d2[c] = i2;
    h0[c] = htab<HLDS>(Hb, k4);
    h1[c] = htab<HLDS>(Hb, i1);
    h2[c] = htab<HLDS>(Hb, i2);
    h3[c] = htab<HLDS>(Hb, i3);
  }
  /* then each cell's class and screen value, one cell at a time so that
   * only one cell's condition masks are live */
#pragma unroll
  for (uint32_t c = 0; c < kSegCells; ++c) {
    __builtin_amdgcn_sched_barrier(0);
    const bool one = kk[c] == 0u || d2[c] >= NK4;
    const bool zero = ZC && kk[c] > K; /* ZC false: the job has no k > K cells */
    const float b = zero ? -kNoCell : (one ? kNoCell : bc[c]);
    xx[c] = (A + b) - ((h0[c] + h1[c]) + (h2[c] + h3[c]));
  }
  __builtin_amdgcn_sched_barrier(0);
}

/* flush_first's tier-1 test, run on a segment's first-term survivors where
 * they sit in registers: a segment with more survivors than the first-term
 * list holds is tested here instead of through the list.  The cells are row
 * i's (constants K, and x as the screen computed it, which is the value
 * flush_first rebuilds); survivors flagged in surv, column of cell c is
 * j0 + c.  Tier-1 survivors go to the tier-2 list. */
template <int NT>
__device__ __forceinline__ void tier1_segment(const DevJob &J, const Lds &S, uint32_t surv, const uint32_t *kk,
                                              const uint32_t *nn4, const float *xx, uint32_t K, uint32_t i,
                                              uint32_t j0, uint2 *cand, uint32_t &nc, uint4 *c3, uint32_t &n3,
                                              BestKey &xb, float &Uw, float &U2w, WaveStats &st) {
  const uint32_t N = J.N;
#pragma unroll
  for (uint32_t c = 0; c < kSegCells; ++c) {
    bool f = false;
    const uint32_t k = kk[c], n = nn4[c] >> 2;
    if (((surv >> c) & 1u) != 0u && k <= min(K, n)) {
      const float x = xx[c];
      if (!(x + J.margin < -745.2f && exact_zero(J, K, n, k, x))) {
        float L = x - J.margin;
        if (L >= -708.f) {
          L += tail_lower_log(N, K, n, k);
          f = L <= Uw;
        } else {
          f = true; /* possibly subnormal: no usable lower bound */
        }
      }
    }
    const uint64_t m = __ballot(f);
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (cnt) {
      if (nc + cnt > (uint32_t)kCandCap) {
        uint64_t tf = cycles();
        flush_candidates<NT>(J, S, cand, nc, c3, n3, xb, Uw, U2w, st);
        st.cyc[3] += cycles() - tf;
      }
      if (f) cand[nc + mask_lt(m)] = make_uint2((i << 16) | (j0 + c), k);
      nc += cnt;
      st.push1 += cnt;
    }
  }
}

/* The best of a class of cells that share one p (p == 0, or p == 1), in
 * optimize's order at equal p (optimize_main.rs:93-116): larger k, then
 * smaller (i, j); n counts the class's cells, nk those at the best k. */
struct ZeroKey {
  uint32_t k, ij, n, nk;
};
__device__ __forceinline__ void merge_zero(ZeroKey &a, const ZeroKey &b) {
  if (b.ij == kNone) return;
  if (a.ij == kNone) {
    a = b;
    return;
  }
  const uint32_t n = a.n + b.n;
  if (b.k > a.k) {
    a = b;
  } else if (b.k == a.k) {
    a.nk += b.nk;
    a.ij = b.ij < a.ij ? b.ij : a.ij;
  }
  a.n = n;
}
__device__ __forceinline__ void take_zero(ZeroKey &z, uint32_t k, uint32_t ij) { merge_zero(z, ZeroKey{k, ij, 1u, 1u}); }
__device__ __forceinline__ ZeroKey wave_zero(ZeroKey z) {
  for (int off = 32; off >= 1; off >>= 1) {
    ZeroKey o;
    o.k = __shfl_xor(z.k, off);
    o.ij = __shfl_xor(z.ij, off);
    o.n = __shfl_xor(z.n, off);
    o.nk = __shfl_xor(z.nk, off);
    merge_zero(z, o);
  }
  return z;
}
__device__ __forceinline__ BestKey class_key(double p, const ZeroKey &z) { return BestKey{p, z.k, z.ij, z.n, z.nk}; }

/* Exact best among the p == 1 cells (k <= max(0, K+n-N)): larger k, then
 * smaller (i, j).  Only needed when no cell has p < 1; clears the wave's
 * column counts and rebuilds them row by row. */
template <int NT, bool NBL = true>
__device__ __forceinline__ ZeroKey class1_best(const DevJob &J, const LdsLayout &Lo, const Lds &S, uint32_t *C) {
  constexpr uint32_t W = NT / 64;
  const uint32_t w = wave_id(), lane = lane_id();
  const uint32_t T1 = J.T1, T2 = J.T2, T2w = J.T2w, N = J.N;
  const uint32_t c2 = (T2w + 63) / 64, wbeg = lane * c2;
  ZeroKey o = {0u, kNone, 0u, 0u};
  for (uint32_t t = lane; t < Lo.T2w_pad; t += 64) C[t] = 0u;
  wave_sync();
  uint32_t added = 0;
  for (uint32_t i = w; i < T1; i += W) {
    rows_update<NBL>(S, C, added, i + 1, true);
    added = i + 1;
    const uint32_t K = J.K[i];
    uint32_t run = row_lane_offset(C, wbeg, c2, T2w);
    for (uint32_t t = 0; t < c2; ++t) {
      const uint32_t wi = wbeg + t;
      const uint32_t x = wi < T2w ? C[wi] : 0u;
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t j = 2 * wi + h;
        run += h ? (x >> 16) : (x & 0xFFFFu);
        if (j >= T2) continue;
        const uint32_t n = S.coln[j];
        const uint32_t mn = (K + n > N) ? K + n - N : 0u;
        if (run <= mn) take_zero(o, run, (i << 16) | j);
      }
    }
  }
  return wave_zero(o);
}

/* The launch's task counter and where a workgroup puts its next task; with
 * explicit permutations, their index arrays (for the prefetch). */
struct NextTask {
  uint32_t *counter;
  uint32_t *slot;
  const uint32_t *idx1, *idx2;
  uint64_t n_perm;
};

/* One permutation: overlap grid -> best cell.  Returns the best key in
 * thread 0 (others: undefined).  Rows are interleaved over the waves: every
 * wave adds every row's pairs to its own running column counts, and scans
 * the rows i with i % W == w.  ZCM, the screen's k > K test (cells with an
 * empty tail): 1 in, 0 left out (the job cannot have such cells,
 * DevJob.kle), 2 chosen by J.kle at run time (the batch kernel, whose jobs
 * differ). */
template <int NT, bool EXPLICIT, bool HLDS, bool PLAIN, bool PL, bool REGCOLS, bool NBL = true, int ZCM = 2>
__device__ __forceinline__ BestKey optimize_perm(const DevJob &J, const LdsLayout &Lo, const Lds &S, const uint32_t *i1,
                                 const uint32_t *i2, const Shuffle &s1, const Shuffle &s2, int *err,
                                 WaveStats &st, uint16_t *gpairs, const NextTask *nt = nullptr) {
  constexpr uint32_t W = NT / 64;
  const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
  uint64_t tc0 = cycles();
  build_rows<NT, EXPLICIT, HLDS, PL, NBL>(J, Lo, S, i1, i2, s1, s2, err, gpairs);
  uint64_t tc1 = cycles();
  st.cyc[0] += tc1 - tc0;
  /* wave 0 claims the workgroup's next task once the build is done: by then
   * every workgroup of the launch has started, so none takes a second task
   * while another still has none.  The claim does not depend on the grid's
   * contents, and the slot is read after this task's last barrier, so every
   * task moves the workgroup on to a new one. */
  uint32_t pnext = kNone;
  if (nt && w == 0u) {
    if (lane == 0) {
      pnext = atomicAdd(nt->counter, 1u);
      *nt->slot = pnext;
    }
    pnext = (uint32_t)__builtin_amdgcn_readfirstlane((int)pnext);
  }

  const uint32_t T1 = J.T1, T2 = J.T2, T2w = J.T2w;
  uint32_t *C = S.cw + w * Lo.T2w_pad;
  /* count words per lane; wide grids (several segments per row) take an odd
   * number, so that the lanes' words, and their columns' constants, fall in
   * different LDS banks (an even stride puts 64 lanes on 64 / 2^t banks).
   * Lanes past the grid read the zero padding word and hold no cell. */
  const uint32_t c2 = REGCOLS ? (T2w + 63) / 64 : ((T2w + 63) / 64) | 1u;
  const uint32_t wbeg = lane * c2;
  const char *Hb = HLDS ? (const char *)S.h : (const char *)J.H32;

  if (PLAIN) {
    BestKey xb = no_best();
    uint32_t added = 0;
    for (uint32_t i = w; i < T1; i += W) {
      rows_update<NBL>(S, C, added, i + 1, true);
      added = i + 1;
      const uint32_t K = J.K[i];
      uint32_t run = row_lane_offset(C, wbeg, c2, T2w);
#pragma unroll 1
      for (uint32_t t = 0; t < c2; ++t) {
        const uint32_t wi = wbeg + t;
        const uint32_t x = wi < T2w ? C[wi] : 0u;
#pragma unroll 1
        for (uint32_t h = 0; h < 2; ++h) {
          const uint32_t j = 2 * wi + h;
          run += h ? (x >> 16) : (x & 0xFFFFu);
          if (j >= T2) continue;
          const uint32_t k = run;
          double p = dto_tail_exact(J.LF, J.N, K, S.coln[j], k, J.D[j], S.tab, nullptr);
          take_better(xb, p, k, (i << 16) | j);
        }
      }
    }
    xb = wave_best(xb);
    if (lane == 0) S.redB[2 * W + w] = xb;
    __syncthreads();
    BestKey b = no_best();
    if (tid == 0)
      for (uint32_t v = 0; v < W; ++v) merge_best(b, S.redB[2 * W + v]);
    __syncthreads();
    return b;
  }

  /* this lane's column constants, kept in registers across rows; columns
   * past T2 repeat the last column's n (their k equals its k) and get no cell */
  ColRegs cr;
  /* the lower-tail lift: the bound of n (N - n) / (N - 1) over the lane's
   * columns (REGCOLS), or over all n (N^2 / 4 / (N - 1), wider rows) */
  const float liftNf = (float)J.N, liftInvN2 = 1.f / (liftNf * liftNf);
  const uint32_t liftN4 = REGCOLS ? J.N << 2 : J.N;
  const float liftc = J.N <= (REGCOLS ? kLiftMaxN : kLiftMaxN1) ? kLogHalfBelow : -__builtin_inff();
  float liftV = !REGCOLS && J.N > 1u ? liftNf * liftNf / (4.f * (liftNf - 1.f)) * 1.001f : 0.f;
  if (REGCOLS) {
#pragma unroll
    for (uint32_t c = 0; c < kSegCells; ++c) {
      const uint32_t j = 2u * wbeg + c;
      const bool valid = (c >> 1) < c2 && j < T2;
      const uint32_t jc = valid ? j : T2 - 1u;
      cr.n4[c] = S.coln[jc] << 2;
      cr.b[c] = valid ? S.colb[jc] : kNoCell;
      const float nf = (float)S.coln[jc];
      liftV = fmaxf(liftV, nf * (liftNf - nf));
    }
    liftV = J.N > 1u ? liftV / (liftNf - 1.f) : 0.f;
#pragma unroll
    for (uint32_t u = 0; u < kSegWords; ++u) {
      const uint32_t last = Lo.T2w_pad - 1u;
      cr.wa[u] = lds_addr(C + (u < c2 ? min(wbeg + u, last) : last));
    }
  }
  const float margin = J.margin;
  const float uthr = -745.2f - margin; /* below: the first term may underflow */

  ZeroKey zb = {0u, kNone, 0u, 0u}; /* p == 0 cells */
  BestKey xb = no_best();           /* exactly evaluated cells */
  /* this lane's cell with the smallest screen value: its tail bracket, taken
   * after the sweep, bounds the minimum log p from above */
  float bx = __builtin_inff();
  uint32_t bK = 0, bn = 0, bk = 0;

  /* ---- D: forward sweep over this wave's rows w, w+W, ... ---- */
  uint32_t added = 0; /* rows [0, added) are in C */
  /* pairs [pe0, pe1) are the next update's (rows [added, i + 1)); the bins
   * of its first pairs are in pb */
  uint32_t pe0 = S.rowstart[0], pe1 = w < T1 ? S.rowstart[w + 1] : pe0;
  uint32_t pb[kPreBins];
  prefetch_bins<NBL>(S, pe0, pe1, pb);
  float lq[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
  /* the row constants of 64 steps at a time, lane l holding step 64 b + l's:
   * one load per 64 rows instead of a global load latency per row */
  uint32_t Kl = 0u, NKl = 0u; /* K_i and 4 (N - K_i) */
  float Al = 0.f;
  uint32_t pfacc = 0u; /* keeps the prefetch loads */
  for (uint32_t i = w, q = 0; i < T1; i += W, ++q) {
    if ((q & 63u) == 0u) {
      const uint32_t ir = i + lane * W;
      Kl = ir < T1 ? J.K[ir] : 0u;
      NKl = (J.N - Kl) << 2;
      Al = ir < T1 ? J.A32[ir] : 0.f;
    }
    add_bins<NBL>(S, C, pe0, pe1, pb);
    added = i + 1;
    const uint32_t inext = i + W;
    uint32_t ne1 = 0u;
    if (inext < T1) ne1 = S.rowstart[inext + 1];
    const uint32_t K = (uint32_t)__builtin_amdgcn_readlane((int)Kl, (int)(q & 63u));
    const uint32_t NK4 = (uint32_t)__builtin_amdgcn_readlane((int)NKl, (int)(q & 63u));
    const float A = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Al), (int)(q & 63u)));
    float lmin = __builtin_inff(); /* packed (value, cell) */
    const uint32_t liftD = lift_offset(K, J.N, liftNf, liftInvN2, liftV) << (REGCOLS ? 2 : 0);
    uint32_t run = REGCOLS ? 0u : row_lane_offset(C, wbeg, c2, T2w);
    /* REGCOLS: the row is one segment */
#pragma unroll 1
    for (uint32_t t0 = 0; REGCOLS ? t0 == 0u : t0 < c2; t0 += kSegWords) {
      const uint32_t nseg = REGCOLS ? c2 : (c2 - t0 < kSegWords ? c2 - t0 : kSegWords);
      uint32_t wv[kSegWords];
      const uint32_t tot = REGCOLS ? read_segment_regs(cr, wv) : read_segment(C, wbeg + t0, nseg, Lo.T2w_pad - 1u, wv);
      if (REGCOLS) run = wave_excl_scan(tot);
      uint32_t kk[kSegCells], nn4[kSegCells];
      float xx[kSegCells];
      if (ZCM == 2 ? J.kle != 0u : ZCM == 0)
        segment_fast<HLDS, REGCOLS, false>(J, S, Hb, wv, nseg, wbeg + t0, run, K, NK4, A, cr, kk, nn4, xx);
      else
        segment_fast<HLDS, REGCOLS, true>(J, S, Hb, wv, nseg, wbeg + t0, run, K, NK4, A, cr, kk, nn4, xx);
      {
        /* row minima and the lane's best cell from the lifted values (a
         * lifted best cell only loosens its own upper bound: x rises) */
#pragma unroll
        for (uint32_t c = 0; c < kSegCells; ++c)
          xx[c] = lift_lower<REGCOLS>(xx[c], kk[c], nn4[c], K, liftN4, liftD, liftc);
      }
      if (t0 == 0u && inext < T1) {
        /* the next update's pairs: rows [i + 1, inext + 1) */
        pe0 = pe1;
        pe1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)ne1);
        prefetch_bins<NBL>(S, pe0, pe1, pb);
      }
      run += tot;
      float smin = __builtin_inff();
#pragma unroll
      for (uint32_t c = 0; c < kSegCells; ++c) smin = fminf(smin, pack_cell(xx[c], c));
      const bool rare = smin < uthr + pack_slack(smin);
      if (smin < bx) {
        /* this lane's best so far is in this segment */
        const uint32_t c = packed_cell(smin);
        uint32_t k = kk[0], n4 = nn4[0];
#pragma unroll
        for (uint32_t u = 1; u < kSegCells; ++u) {
          k = c == u ? kk[u] : k;
          n4 = c == u ? nn4[u] : n4;
        }
        bx = smin;
        bK = K;
        bn = n4 >> 2;
        bk = k;
      }
      lmin = fminf(lmin, smin);
      if (__ballot(rare)) {
        /* exact p == 0 cells of this segment */
#pragma unroll
        for (uint32_t c = 0; c < kSegCells; ++c) {
          const uint32_t j = 2u * (wbeg + t0) + c;
          if ((c >> 1) < nseg && j < T2 && exact_zero(J, K, nn4[c] >> 2, kk[c], xx[c]))
            take_zero(zb, kk[c], (i << 16) | j);
        }
      }
    }
    /* row minima are reduced four rows at a time (rows i - 3W .. i) */
    lq[0] = lq[1];
    lq[1] = lq[2];
    lq[2] = lq[3];
    lq[3] = lmin;
    if ((q & 3u) == 3u || inext >= T1) {
      wave_minf_n<4>(lq);
      if (lane == 0) {
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
          if (q + r < 3u) continue; /* before the first row */
          /* lower bound of the row's sum-class log p (subnormal sums: -inf) */
          const float L = lq[r] - margin - pack_slack(lq[r]);
          S.rowmin[i - (3u - r) * W] = L < -708.f ? -__builtin_inff() : L;
        }
      }
    }
  }

  /* U*: every p <= 1 (+ rounding), so log p <= 1e-6 always.  Each lane
   * brackets its best cell's tail with 16 ratio steps, which settles the
   * fast-falling tails of small-p cells; the wave minimum U1 of those bounds
   * is valid.  Only lanes still unsettled and with a screen value below U1
   * (a lane's bound is never below its screen value) go on to 64 steps, and
   * stop as soon as their sum shows they cannot get below U1: a lane whose
   * terms keep growing (a cell below the mode) no longer sets the length of
   * the wave's loop. */
  float Urun = 1e-6f;
  {
    const float xs = bx + pack_slack(bx);
    float r1 = 1e-6f;
    if (bx < 1e30f) r1 = screen_upper(J, bK, bn, bk, xs, 16);
    const bool open = bx < 1e30f && r1 == 1e-6f; /* 1e-6: not settled in 16 steps */
    const float U1 = wave_minf(fminf(r1, 1e-6f));
    Urun = U1;
    if (open && xs < U1)
      Urun = fminf(Urun, screen_upper(J, bK, bn, bk, xs, 64, __expf(fminf(U1 - xs, 80.f))));
  }
  Urun = wave_minf(Urun);
  zb = wave_zero(zb);
  if (zb.ij != kNone) Urun = -__builtin_inff();
  uint64_t tc2 = cycles();
  st.cyc[1] += tc2 - tc1;
  if (lane == 0) {
    S.redU[w] = Urun;
    S.redB[w] = class_key(0.0, zb);
  }
  /* no barrier here: each wave goes on to its own rows' reverse sweep as
   * soon as its forward sweep is done, starting from its own bound, and
   * takes up the other waves' bounds through the shared minimum as they
   * arrive (the reverse sweep reads only this wave's counts, lists and row
   * bounds) */
  share_bound(S, Urun);
  const float Ustar = fminf(Urun, fkey_value(*(volatile int *)S.ushare));
  float Uw = Ustar;  /* tier-1 threshold (wave, tightens as tier 2 runs) */
  float U2w = Ustar; /* best tier-2 upper end seen by this wave */

  /* ---- E: reverse sweep over this wave's rows that can hold the minimum ---- */
  static_assert(kCandCap >= 64 && kCand0Cap >= 64, "one ballot (64 cells) must fit an empty list");
  uint2 *cand0 = S.cand0 + w * kCand0Cap;
  uint2 *cand = S.cand + w * kCandCap;
  uint4 *c3 = S.cand3 + w * kCand3Cap;
  uint32_t n0 = 0, nc = 0, n3 = 0;
  /* C holds rows [0, ctop).  The owned rows go from the top down, 64 per
   * batch: lane l holds the bound of the batch's l-th row, a ballot finds the
   * next row that can still hold the minimum (Uw only tightens), and the
   * rows above it are removed in one pass.  Rows below the last rescanned
   * row are never removed (the next permutation clears the counts). */
  uint32_t ctop = added;
  const uint32_t nown = added > w ? (added - 1u - w) / W + 1u : 0u; /* owned rows */
#pragma unroll 1
  for (uint32_t b0 = 0; b0 < nown; b0 += 64) {
    const uint32_t top = added - 1u - b0 * W; /* this batch's first (highest) row */
    const bool have = b0 + lane < nown;
    const float rm = have ? S.rowmin[top - lane * W] : __builtin_inff();
    const uint32_t Kl = have ? J.K[top - lane * W] : 0u;
    const uint32_t NKl = (J.N - Kl) << 2;
    const float Al = have ? J.A32[top - lane * W] : 0.f;
    uint64_t done = 0;
#pragma unroll 1
    for (;;) {
      /* bounds found by the other waves apply here too */
      const float us = fkey_value(*(volatile int *)S.ushare);
      Uw = fminf(Uw, us);
      U2w = fminf(U2w, us);
      const uint64_t m = __ballot(rm <= Uw) & ~done;
      if (!m) break;
      const uint32_t l = (uint32_t)__builtin_ctzll(m);
      done |= (2ull << l) - 1ull; /* rows at or above this one are passed */
      const uint32_t i = top - l * W;
      if (S.rowstart[i + 1u] - S.rowstart[0] < S.rowstart[ctop] - S.rowstart[i + 1u]) {
        /* fewer pairs below than between: rebuild the counts from zero */
        for (uint32_t t = lane; t < Lo.T2w_pad; t += 64u) C[t] = 0u;
        rows_update<NBL>(S, C, 0u, i + 1u, true);
      } else {
        rows_update<NBL>(S, C, i + 1u, ctop, false);
      }
      ctop = i + 1u;
      st.rows++;
      const uint32_t K = (uint32_t)__builtin_amdgcn_readlane((int)Kl, (int)l);
      const uint32_t NK4 = (uint32_t)__builtin_amdgcn_readlane((int)NKl, (int)l);
      const float A = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(Al), (int)l));
      const uint32_t liftD = lift_offset(K, J.N, liftNf, liftInvN2, liftV) << (REGCOLS ? 2 : 0);
      uint32_t run = REGCOLS ? 0u : row_lane_offset(C, wbeg, c2, T2w);
#pragma unroll 1
      for (uint32_t t0 = 0; REGCOLS ? t0 == 0u : t0 < c2; t0 += kSegWords) {
        const uint32_t nseg = REGCOLS ? c2 : (c2 - t0 < kSegWords ? c2 - t0 : kSegWords);
        uint32_t wv[kSegWords];
        const uint32_t tot =
            REGCOLS ? read_segment_regs(cr, wv) : read_segment(C, wbeg + t0, nseg, Lo.T2w_pad - 1u, wv);
        if (REGCOLS) run = wave_excl_scan(tot);
        uint32_t kk[kSegCells], nn4[kSegCells];
        float xx[kSegCells];
        if (ZCM == 2 ? J.kle != 0u : ZCM == 0)
          segment_fast<HLDS, REGCOLS, false>(J, S, Hb, wv, nseg, wbeg + t0, run, K, NK4, A, cr, kk, nn4, xx);
        else
          segment_fast<HLDS, REGCOLS, true>(J, S, Hb, wv, nseg, wbeg + t0, run, K, NK4, A, cr, kk, nn4, xx);
        run += tot;
        /* cheap first test: the first-term bound alone; cells whose bound
         * is below -708 (possibly subnormal or zero sums) always pass */
        const float thr = fmaxf(Uw + margin, -708.f + margin);
        uint32_t surv = 0;
        /* lifted lower-tail cells do not pass; xx keeps the first-term values */
#pragma unroll
        for (uint32_t c = 0; c < kSegCells; ++c)
          surv |= lift_lower<REGCOLS>(xx[c], kk[c], nn4[c], K, liftN4, liftD, liftc) <= thr ? 1u << c : 0u;
        if (!__ballot(surv != 0u)) continue;
        st.cheap += (uint32_t)__popc(surv);
        uint32_t stot = 0;
#pragma unroll
        for (uint32_t c = 0; c < kSegCells; ++c) stot += (uint32_t)__popcll(__ballot(((surv >> c) & 1u) != 0u));
        if (stot > (uint32_t)kCand0Cap) {
          /* more survivors than the list holds: tier 1 where they are */
          tier1_segment<NT>(J, S, surv, kk, nn4, xx, K, i, 2u * (wbeg + t0), cand, nc, c3, n3, xb, Uw, U2w, st);
          continue;
        }
        /* make room first: then the segment's survivors fit the list */
        if (n0 + stot > (uint32_t)kCand0Cap)
          flush_first<NT, HLDS>(J, S, Hb, cand0, n0, cand, nc, c3, n3, xb, Uw, U2w, st, top, Kl, Al);
#pragma unroll
        for (uint32_t c = 0; c < kSegCells; ++c) {
          const bool f = ((surv >> c) & 1u) != 0u;
          const uint64_t m = __ballot(f);
          if (f) cand0[n0 + mask_lt(m)] = make_uint2((i << 16) | (2u * (wbeg + t0) + c), kk[c]);
          n0 += (uint32_t)__popcll(m);
        }
      }
    }
    /* the list holds rows of this batch only */
    if (n0) flush_first<NT, HLDS>(J, S, Hb, cand0, n0, cand, nc, c3, n3, xb, Uw, U2w, st, top, Kl, Al);
  }
  uint64_t tc3 = cycles();
  st.cyc[2] += tc3 - tc2;
  if (nc) flush_candidates<NT>(J, S, cand, nc, c3, n3, xb, Uw, U2w, st);
  uint64_t tc4 = cycles();
  st.cyc[3] += tc4 - tc3;
  if (nt && w == 0u && nt->idx1 && nt->idx2 && pnext < nt->n_perm) {
    /* the next task's index arrays, late in this task so that they are
     * still in the cache when its build reads them: one load per 128-byte
     * line, kPf lines per lane issued together before a single wait (the
     * sum keeps the loads) */
    constexpr uint32_t kPf = 16;
    const uint32_t *p1 = nt->idx1 + (uint64_t)pnext * J.N1, *p2 = nt->idx2 + (uint64_t)pnext * J.N2;
    const uint32_t l1 = (J.N1 + 31u) / 32u, l2 = (J.N2 + 31u) / 32u;
    for (uint32_t t0 = 0; t0 < l1 + l2; t0 += 64u * kPf) {
      uint32_t v[kPf];
#pragma unroll
      for (uint32_t u = 0; u < kPf; ++u) {
        const uint32_t t = t0 + u * 64u + lane;
        v[u] = t < l1 ? p1[32u * t] : (t < l1 + l2 ? p2[32u * (t - l1)] : 0u);
      }
#pragma unroll
      for (uint32_t u = 0; u < kPf; ++u) pfacc += v[u];
    }
  }
  /* tier 3: exact FP64 for cells whose bracket reaches the best bound this
   * wave knows of (its own, and what the other waves have shared so far: no
   * barrier, a looser bound only evaluates more cells) */
  uint64_t tc5 = cycles();
  {
    const float U2o = fminf(U2w, log_bound_of(xb));
    share_bound(S, U2o);
    const float U2s = fminf(U2o, fkey_value(*(volatile int *)S.ushare));
    if (n3) flush_exact(J, S, c3, n3, U2s, xb, st);
  }
  st.cyc[4] += cycles() - tc5;
  if (lane == 0) S.redB[2 * W + w] = xb;
  __syncthreads();
  if (pfacc == 0xFFFFFFFFu && J.N == 0u) S.redS[0] = pfacc; /* never true: N >= 1 */
  /* p == 1 cells matter only when no cell has p < 1 */
  bool need1 = true;
  for (uint32_t v = 0; v < W; ++v) {
    if (S.redB[v].ij != kNone) need1 = false;                                    /* a p == 0 cell */
    if (S.redB[2 * W + v].ij != kNone && S.redB[2 * W + v].p < 1.0) need1 = false; /* a p < 1 cell */
  }
  if (need1) {
    const ZeroKey o = class1_best<NT, NBL>(J, Lo, S, C);
    if (lane == 0) S.redB[W + w] = class_key(1.0, o);
  } else if (lane == 0) {
    S.redB[W + w] = no_best();
  }
  __syncthreads();
  /* the three classes hold disjoint cells: p == 0 cells flagged by the
   * forward sweep (never evaluated), p == 1 cells (k <= max(0, K + n - N)),
   * and the evaluated sum-class cells; the tie counts add up across them */
  BestKey b = no_best();
  if (tid == 0) {
    for (uint32_t v = 0; v < 3 * W; ++v) merge_best(b, S.redB[v]);
    if (DTO_STATS) {
      float U0 = __builtin_inff(); /* the forward sweeps' bound */
      for (uint32_t v = 0; v < W; ++v) U0 = fminf(U0, S.redU[v]);
      if (b.p > 0.0 && b.p < 1.0 && U0 < 1e30f) st.gap += (uint32_t)fmaxf(0.f, (U0 - (float)log(b.p)) * 1000.f);
    }
  }
  __syncthreads();
  return b;
}

/* Adds one wave's work counters to the job's counters (one atomic each per
 * wave); nothing without DTO_STATS. */
__device__ __forceinline__ void flush_stats(const DevJob &J, const WaveStats &st) {
  if (!DTO_STATS) return;
  uint32_t t = st.terms, ch = st.cheap;
  for (int off = 32; off >= 1; off >>= 1) {
    t += __shfl_xor(t, off);
    ch += __shfl_xor(ch, off);
  }
  if (lane_id() == 0 && J.stats) {
    if (st.cells) atomicAdd(&J.stats[0], (unsigned long long)st.cells);
    if (t) atomicAdd(&J.stats[1], (unsigned long long)t);
    if (st.rows) atomicAdd(&J.stats[2], (unsigned long long)st.rows);
    for (int c = 0; c < 5; ++c)
      if (st.cyc[c]) atomicAdd(&J.stats[3 + c], (unsigned long long)st.cyc[c]);
    if (st.push1) atomicAdd(&J.stats[8], (unsigned long long)st.push1);
    if (st.push2) atomicAdd(&J.stats[9], (unsigned long long)st.push2);
    if (st.flushes) atomicAdd(&J.stats[10], (unsigned long long)st.flushes);
    if (st.gap) atomicAdd(&J.stats[11], (unsigned long long)st.gap);
    if (ch) atomicAdd(&J.stats[12], (unsigned long long)ch);
  }
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
  r.ties = (b.np < 0xFFFFu ? b.np : 0xFFFFu) | ((b.npk < 0xFFFFu ? b.npk : 0xFFFFu) << 16);
  *out = r;
}

/* Permutations [perm_base, perm_base + n_perm) of one job.  EXPLICIT: idx1 /
 * idx2 hold n_perm permutations back to back (NULL: identity); otherwise the
 * Philox shuffle keyed by (seed, job_id, permutation id, list).  ws: per
 * workgroup pair staging (N2 u16 each) when !PL.  Workgroups take
 * permutations from a device counter (zeroed before the launch), so that
 * permutations of uneven cost spread evenly; every workgroup exits once the
 * counter passes n_perm.  The screen keeps its k > K test for every job:
 * one code path in the sweeps costs less than leaving the test out. */
template <int NT, bool EXPLICIT, bool HLDS, bool PLAIN, bool PL, bool REGCOLS, bool NBL>
__global__ __launch_bounds__(NT) void k_optimize(DevJob J, LdsLayout Lo, const uint32_t *__restrict__ idx1,
                                                 const uint32_t *__restrict__ idx2, uint64_t n_perm,
                                                 uint64_t perm_base, uint64_t seed, uint32_t job_id,
                                                 int32_t permuted, dto_record *__restrict__ out,
                                                 int *__restrict__ err, uint16_t *__restrict__ ws,
                                                 uint16_t *__restrict__ gnb, uint32_t *__restrict__ counter) {
  /* no static LDS here: the dynamic region then starts at address 0 */
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t &s_next = *(uint32_t *)(smem + Lo.misc);
  constexpr int W = NT / 64;
  const Lds S = lds_view<NBL>(smem, Lo, W, NBL ? nullptr : gnb + (uint64_t)blockIdx.x * max(J.N1, J.N2));
  load_job_tables<NT, HLDS>(J, S);
  __syncthreads();
  uint16_t *gpairs = PL ? nullptr : ws + (uint64_t)blockIdx.x * J.N2;
  WaveStats st = {0u, 0u, 0u, {0, 0, 0, 0, 0}, 0u, 0u, 0u, 0u, 0u};
  /* the first task; each task claims the next one (optimize_perm) */
  if (threadIdx.x == 0) s_next = atomicAdd(counter, 1u);
  __syncthreads();
  uint64_t pp = s_next;
  __syncthreads();
  const NextTask nt = {counter, &s_next, EXPLICIT && !PLAIN ? idx1 : nullptr, EXPLICIT && !PLAIN ? idx2 : nullptr,
                       n_perm};
  for (;;) {
    if (pp >= n_perm) break;
    const uint64_t pid = perm_base + pp;
    const uint32_t *i1 = (EXPLICIT && idx1) ? idx1 + pp * J.N1 : nullptr;
    const uint32_t *i2 = (EXPLICIT && idx2) ? idx2 + pp * J.N2 : nullptr;
    Shuffle s1, s2;
    if (!EXPLICIT) {
      s1 = shuffle_make(seed, job_id, (uint32_t)pid, 1u, J.N1);
      s2 = shuffle_make(seed, job_id, (uint32_t)pid, 2u, J.N2);
    }
    BestKey b =
        optimize_perm<NT, EXPLICIT, HLDS, PLAIN, PL, REGCOLS, NBL, 1>(J, Lo, S, i1, i2, s1, s2, err, st, gpairs, &nt);
    /* a counter value claimed during this task (optimize_perm claims one
     * every time, after the build): published by its last barrier, and new,
     * so the loop always moves on */
    const uint64_t pn = s_next;
    if (threadIdx.x == 0) write_record(J, b, permuted, out + pp, err);
    pp = pn;
  }
  flush_stats(J, st);
}

/* Zeroes one work counter (a kernel, not a memset, so that callers may
 * capture the launch sequence in a graph). */
static __global__ void k_zero_counter(uint32_t *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
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
template <int NT, bool HLDS, bool PL, bool REGCOLS, bool NBL>
__global__ __launch_bounds__(NT) void k_batch(const BatchJob *__restrict__ jobs, const WorkItem *__restrict__ items,
                                              uint32_t n_items, uint64_t seed, uint32_t *__restrict__ counter,
                                              int *__restrict__ err, uint16_t *__restrict__ ws, uint32_t ws_stride,
                                              uint16_t *__restrict__ gnb, uint32_t gnb_stride, uint32_t item_slot) {
  /* no static LDS here: the dynamic region then starts at address 0.  The
   * work item slot is the misc slot of the batch's largest layout, which
   * every other job's layout ends before. */
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int W = NT / 64;
  uint32_t &s_item = *(uint32_t *)(smem + item_slot);
  uint16_t *gpairs = PL ? nullptr : ws + (uint64_t)blockIdx.x * ws_stride;
  uint32_t cur = kNone;
  WaveStats st = {0u, 0u, 0u, {0, 0, 0, 0, 0}, 0u, 0u, 0u, 0u, 0u};
  const DevJob *last = nullptr;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t it = s_item;
    __syncthreads();
    if (it >= n_items) {
      if (last) flush_stats(*last, st);
      break;
    }
    const WorkItem wi = items[it];
    const BatchJob &B = jobs[wi.job];
    const Lds S = lds_view<NBL>(smem, B.lo, W, NBL ? nullptr : gnb + (uint64_t)blockIdx.x * gnb_stride);
    if (wi.job != cur) {
      if (last) flush_stats(*last, st);
      st = {0u, 0u, 0u, {0, 0, 0, 0, 0}, 0u, 0u, 0u, 0u, 0u};
      last = &B.J;
      load_job_tables<NT, HLDS>(B.J, S);
      __syncthreads();
      cur = wi.job;
    }
    for (uint32_t t = wi.t0; t < wi.t0 + wi.count; ++t) {
      BestKey b;
      Shuffle s1, s2;
      if (t == 0) {
        b = optimize_perm<NT, true, HLDS, false, PL, REGCOLS, NBL>(B.J, B.lo, S, nullptr, nullptr, s1, s2, err, st,
                                                                  gpairs);
      } else {
        s1 = shuffle_make(seed, B.job_id, t - 1, 1u, B.J.N1);
        s2 = shuffle_make(seed, B.job_id, t - 1, 2u, B.J.N2);
        b = optimize_perm<NT, false, HLDS, false, PL, REGCOLS, NBL>(B.J, B.lo, S, nullptr, nullptr, s1, s2, err, st,
                                                                   gpairs);
      }
      if (threadIdx.x == 0) write_record(B.J, b, t == 0 ? 0 : 1, B.out + t, err);
    }
  }
}

/* Overlap counts of every cell for one permutation (debug grid export). */
template <int NT, bool PL>
__global__ __launch_bounds__(NT) void k_grid_counts(DevJob J, LdsLayout Lo, const uint32_t *__restrict__ idx1,
                                                    const uint32_t *__restrict__ idx2,
                                                    uint32_t *__restrict__ kout, int *__restrict__ err,
                                                    uint16_t *__restrict__ ws) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int W = NT / 64;
  const Lds S = lds_view(smem, Lo, W);
  Shuffle s1, s2;
  build_rows<NT, true, false, PL>(J, Lo, S, idx1, idx2, s1, s2, err, ws);
  const uint32_t w = wave_id(), lane = lane_id();
  const uint32_t T1 = J.T1, T2 = J.T2, T2w = J.T2w;
  uint32_t *C = S.cw + w * Lo.T2w_pad;
  const uint32_t c2 = (T2w + 63) / 64, wbeg = lane * c2;
  for (uint32_t i = 0; i < T1; ++i) {
    row_update(S, C, i, true);
    if (i % W != w) continue;
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

/* Philox shuffle on the device (checks that host and device draw the same
 * permutation); one workgroup of 256 threads. */
static __global__ __launch_bounds__(256) void k_shuffle(uint64_t seed, uint32_t job, uint32_t perm, uint32_t list,
                                                 uint32_t n, uint32_t *__restrict__ out) {
  __shared__ uint32_t wsum[4];
  const Shuffle s = shuffle_make(seed, job, perm, list, n);
  shuffle_visit<256>(s, wsum, [&](const uint32_t *k, const uint32_t *v) {
#pragma unroll
    for (uint32_t u = 0; u < kShufU; ++u)
      if (v[u] < n) out[k[u]] = v[u];
  });
}

}  // namespace dto

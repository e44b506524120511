/*
 * Device-side job layout shared by the host library and the kernels.
 *
 * One job = two rank-sorted lists joined on feature id, their threshold lists
 * and the tables the kernels read.  All of it is uploaded once
 * (dto_job_create) and stays resident in HBM for every permutation.
 */
#pragma once
#include <stdint.h>

#include "../../include/dto.h"

namespace dto {

constexpr uint32_t kBinNone = 0xFFFFu; /* (a,b) pair outside the grid */

struct DevJob {
  uint32_t N;        /* population size */
  uint32_t N1, N2;   /* list lengths */
  uint32_t T1, T2;   /* threshold counts */
  uint32_t T2w;      /* ceil(T2 / 2): column counts are packed two per u32 */
  uint32_t dup;      /* 1: some list2 entry joins several list1 entries (CSR) */
  uint32_t kle;      /* 1: every overlap k_ij <= K_i (no cell has an empty tail) */
  float margin;      /* bound on |screen log-term - exact log-term| (nats) */
  const uint16_t *bin1;     /* [N1] threshold bin of sorted list1 position (T1 = none) */
  const uint16_t *bin2;     /* [N2] */
  const int32_t *join;      /* [N2] list1 sorted position of list2 entry q, -1 none */
  const uint32_t *join_off; /* [N2+1] CSR over join_pos when dup */
  const uint32_t *join_pos;
  const uint32_t *K;        /* [T1] |S1(T1[i])| */
  const uint32_t *n;        /* [T2] |S2(T2[j])| */
  const uint32_t *thr1;     /* [T1] threshold values */
  const uint32_t *thr2;     /* [T2] */
  const double *LF;         /* [N+1] statrs ln_factorial */
  const double *D;          /* [T2] ln_binomial(N, n_j), statrs order */
  const float *H32;         /* [N+1] LF[x] - x*(ln N - 1), rounded to f32 */
  const float *A32;         /* [T1] H(K_i) + H(N-K_i) */
  const float *B32;         /* [T2] H(n_j) + H(N-n_j) - H(N) */
  unsigned long long *stats; /* [16] work counters (dto_internal_stats) */
};

/* Byte offsets of the dynamic LDS regions of one workgroup. */
struct LdsLayout {
  uint32_t h;        /* f32[N+1] screen table (only when h_in_lds); always 0 */
  uint32_t pairs;    /* u16[N2] row of each list2 position's pair (staging): aliases h; always 0 */
  uint32_t pairs_in_lds;
  uint32_t nb;       /* u16[max(N1,N2)]: list1 bins after the shuffle, then row-sorted list2 bins (nb_in_lds) */
  uint32_t rowstart; /* u32[T1+1] */
  uint32_t cursor;   /* u32[T1] */
  uint32_t cw;       /* u32[W][T2w_pad] per-wave running column counts (2 x u16 per word) */
  uint32_t rowmin;   /* f32[T1] per-row minimum screen lower bound */
  uint32_t coln;     /* u32[T2] */
  uint32_t colb;     /* f32[T2] */
  uint32_t cand0;    /* uint2[W][kCand0Cap] cells past the first-term test (aliases cursor) */
  uint32_t cand;     /* uint2[W][kCandCap] tier-1 survivors */
  uint32_t cand3;    /* uint4[W][kCand3Cap] tier-2 survivors */
  uint32_t exptab;   /* u64[256] */
  uint32_t red;      /* per-wave reduction slots */
  uint32_t misc;     /* u32[4]: [0] the workgroup's current work item, [2] the shared bound key */
  uint32_t total;
  uint32_t T2w_pad;
  uint32_t h_in_lds;
  uint32_t nb_in_lds; /* 0: the bins array is a per-workgroup global slice (leaves the LDS to h) */
};

constexpr int kCand0Cap = 128;
constexpr int kCandCap = 64;
constexpr int kCand3Cap = 32;

/* The best cell found so far: key (p, -k, i, j), lexicographic minimum,
 * with the number of cells seen at its p (np) and, of those, at its k (npk):
 * optimize warns when either is above one. */
struct BestKey {
  double p;
  uint32_t k;
  uint32_t ij; /* i << 16 | j; 0xFFFFFFFF = none */
  uint32_t np, npk;
};

}  // namespace dto

/*
 * Arithmetic shared by the host library and the gfx950 kernels.
 *
 *  - dto_exp:        the host C library's exp (glibc >= 2.28 table-driven
 *                    design), restated so device results are bit-identical
 *                    to what the reference gets from f64::exp on
 *                    x86_64-linux-gnu.
 *  - dto_tail_exact: statrs 0.17.1 Hypergeometric::sf term fold
 *                    (called from src/stat_operations/hypergeometric_pvalue.rs:40-49),
 *                    replayed in the same operation order from an ln n! table,
 *                    with an early exit that provably leaves the sum unchanged.
 *  - Philox4x32-10 and a keyed Feistel bijection with cycle walking: the
 *    counter-based shuffle that replaces thread_rng() + SliceRandom::shuffle
 *    (src/collections/permuted.rs:56-60).
 *
 * Floating-point contraction is off for everything here: the reference (Rust)
 * never fuses a multiply into an add, so only the fma() calls written out in
 * dto_exp (which restate the fused operations of the library exp) may fuse.
 */
#pragma once
#include <stdint.h>
#include <string.h>

#include "dto_exp_table.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DTO_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define DTO_HD static inline
#endif

#pragma clang fp contract(off)

namespace dto {

DTO_HD uint64_t as_u64(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}
DTO_HD double as_f64(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
}

/* ------------------------------------------------------------------ */
/* exp                                                                  */
/* ------------------------------------------------------------------ */
/* exp(x) = 2^(k/128) * exp(r), |r| <= ln2/256, exp(r) - 1 by a degree-5
 * polynomial.  Constants: 128/ln2, ln2/128 split hi/lo, the 1.5*2^52 rounding
 * shift, and the minimax coefficients C2..C5 of the published design.  The
 * fused operations are where the x86-64 FMA build of that library fuses
 * them; on the underflow side of the special case the scaled sum and its
 * error term are evaluated unfused.  Validated
 * bit-for-bit against the host library on 4e8 random arguments
 * (tests/test_exp_replica.py keeps a smaller sweep). */
struct ExpConst {
  static constexpr double InvLn2N = 0x1.71547652b82fep0 * 128.0;
  static constexpr double NegLn2hiN = -0x1.62e42fefa0000p-8;
  static constexpr double NegLn2loN = -0x1.cf79abc9e3b3ap-47;
  static constexpr double Shift = 0x1.8p52;
  static constexpr double C2 = 0x1.ffffffffffdbdp-2;
  static constexpr double C3 = 0x1.555555555543cp-3;
  static constexpr double C4 = 0x1.55555cf172b91p-5;
  static constexpr double C5 = 0x1.1111167a4d017p-7;
};

DTO_HD double dto_fma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return fma(a, b, c);
#endif
}

/* Subnormal / large-|x| tail of the exp (|x| >= 512 and finite). */
DTO_HD double dto_exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  double scale, y;
  if ((ki & 0x80000000u) == 0) {
    sbits -= 1009ull << 52;
    scale = as_f64(sbits);
    y = 0x1p1009 * dto_fma(scale, tmp, scale);
    return y;
  }
  sbits += 1022ull << 52;
  scale = as_f64(sbits);
  double st = scale * tmp;
  y = scale + st;
  if (y < 1.0) {
    double hi, lo;
    lo = (scale - y) + st;
    hi = 1.0 + y;
    lo = ((1.0 - hi) + y) + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

DTO_HD double dto_exp(double x, const uint64_t *tab) {
  uint32_t abstop = (uint32_t)(as_u64(x) >> 52) & 0x7ffu;
  if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
    if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x; /* |x| < 2^-54 */
    if (abstop >= 0x409u) {                              /* |x| >= 1024, inf, nan */
      if (as_u64(x) == 0xfff0000000000000ull) return 0.0;
      if (abstop >= 0x7ffu) return 1.0 + x;
      return (as_u64(x) >> 63) ? 0.0 : as_f64(0x7ff0000000000000ull);
    }
    abstop = 0; /* 512 <= |x| < 1024: special-cased below */
  }
  double kd = dto_fma(ExpConst::InvLn2N, x, ExpConst::Shift);
  uint64_t ki = as_u64(kd);
  kd -= ExpConst::Shift;
  double r = dto_fma(kd, ExpConst::NegLn2loN, dto_fma(kd, ExpConst::NegLn2hiN, x));
  uint64_t idx = 2 * (ki % 128);
  uint64_t top = ki << 45;
  double tail = as_f64(tab[idx]);
  uint64_t sbits = tab[idx + 1] + top;
  double r2 = r * r;
  double r4 = r2 * r2;
  double p1 = dto_fma(r2, dto_fma(r, ExpConst::C3, ExpConst::C2), tail + r);
  double tmp = dto_fma(r4, dto_fma(r, ExpConst::C5, ExpConst::C4), p1);
  if (abstop == 0) return dto_exp_special(tmp, sbits, ki);
  double scale = as_f64(sbits);
  return dto_fma(scale, tmp, scale);
}

/* ------------------------------------------------------------------ */
/* hypergeometric upper tail                                            */
/* ------------------------------------------------------------------ */
/* statrs Hypergeometric(N, K, n).sf(k - 1) for k >= 1, i.e. the
 * hypergeometric_pvalue of src/stat_operations/hypergeometric_pvalue.rs:33-50
 * once Hypergeometric::new has accepted (K <= N, n <= N):
 *   min = (n + K) saturating_sub N, max = min(K, n)
 *   k - 1 < min -> 1.0;  k - 1 >= max -> 0.0
 *   else acc = 0; for i = k ..= max:
 *     acc = acc + exp((ln_binomial(K,i) + ln_binomial(N-K,n-i)) - ln_binomial(N,n))
 * with ln_binomial(a,b) = (LF[a] - LF[b]) - LF[a-b] and d = ln_binomial(N,n)
 * precomputed in the same order.
 *
 * Early exit: after adding term i, stop when term_i * 2^56 <= acc and the
 * exact term ratio rho_i = (K-i)(n-i) / ((i+1)(N-K-n+i+1)) is <= 1 - 2^-20.
 * rho decreases with i, so every later exact exponent is below exponent i by
 * at least 2^-20 nats, far more than the ~1e-8 error of the computed
 * exponents; the computed later terms are then no larger than term_i (up to
 * one ulp of exp's own rounding, or one subnormal step), hence below half an
 * ulp of acc, and leave acc unchanged under round-to-nearest exactly as if
 * they had been added.  When term_i underflowed to 0 every later term
 * underflows too.  Returns the number of terms evaluated through *terms. */
DTO_HD double dto_tail_exact(const double *LF, uint32_t N, uint32_t K, uint32_t n, uint32_t k,
                             double d, const uint64_t *tab, uint32_t *terms) {
  if (terms) *terms = 0;
  if (k == 0) return 1.0;
  uint32_t x = k - 1;
  uint32_t mn = (K + n > N) ? K + n - N : 0;
  uint32_t mx = K < n ? K : n;
  if (x < mn) return 1.0;
  if (x >= mx) return 0.0;
  const double lfK = LF[K];
  const double lfNK = LF[N - K];
  const uint32_t NK = N - K;
  double acc = 0.0;
  uint32_t i = k;
  for (; i <= mx; ++i) {
    double lb1 = (lfK - LF[i]) - LF[K - i];
    double lb2 = (lfNK - LF[n - i]) - LF[NK - (n - i)];
    double t = dto_exp((lb1 + lb2) - d, tab);
    acc = acc + t;
    if (t * 0x1p56 <= acc) {
      /* rho_i <= 1 - 2^-20, checked as num <= den - ceil(den / 2^20) */
      uint64_t num = (uint64_t)(K - i) * (uint64_t)(n - i);
      uint64_t den = (uint64_t)(i + 1) * (uint64_t)(NK - n + i + 1);
      if (num <= den - ((den + 0xFFFFFull) >> 20)) {
        ++i;
        break;
      }
    }
  }
  if (terms) *terms = i - k;
  return acc;
}

/* Full statrs term fold with no early exit (reference cost, used by tests). */
DTO_HD double dto_tail_full(const double *LF, uint32_t N, uint32_t K, uint32_t n, uint32_t k, double d,
                            const uint64_t *tab) {
  if (k == 0) return 1.0;
  uint32_t x = k - 1;
  uint32_t mn = (K + n > N) ? K + n - N : 0;
  uint32_t mx = K < n ? K : n;
  if (x < mn) return 1.0;
  if (x >= mx) return 0.0;
  double acc = 0.0;
  for (uint32_t i = k; i <= mx; ++i) {
    double lb1 = (LF[K] - LF[i]) - LF[K - i];
    double lb2 = (LF[N - K] - LF[n - i]) - LF[(N - K) - (n - i)];
    acc = acc + dto_exp((lb1 + lb2) - d, tab);
  }
  return acc;
}

/* ------------------------------------------------------------------ */
/* counter-based shuffle                                                */
/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw; SC'11). */
struct U32x4 {
  uint32_t x, y, z, w;
};
DTO_HD U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

/* A keyed permutation of [0, n).  F is an alternating Feistel network over
 * the mixed-radix domain Z_a x Z_b, a = 2^p, of dom = a b elements: for
 * n <= 4096, a = b = 64 (dom = 4096); above, p = floor(ceil(log2 n) / 2) and
 * b = ceil(n / a), so dom exceeds n by less than a (under 1 % at n = 20,000,
 * where a power-of-two domain wastes 64 %).  Element v is the pair
 * (L = v mod a, R = v div a); even rounds add a keyed mix of R to L mod a,
 * odd rounds a keyed mix of L, scaled into [0, b), to R mod b.  Each round,
 * and F, is a bijection of the domain.  The permutation is F's image of the
 * domain in order, keeping the values < n:
 *   idx = [F(v) for v = 0, 1, ..., dom - 1 if F(v) < n],
 * which holds every value of [0, n) exactly once; the 4096 floor keeps short
 * lists well mixed (their order is that of their values' preimages in a
 * 4096-element permutation).  (Every domain element is evaluated once, so a
 * GPU workgroup computes it without divergence, with a prefix count for the
 * positions.)  The 8 round keys come from two Philox4x32-10 blocks keyed by
 * the seed, with (permutation, job, list) as the counter.  The mix and the
 * scaling use 24-bit multiplies (full rate on the GPU). */
#define DTO_FEISTEL_ROUNDS 8
struct Shuffle {
  uint32_t n, pa, amask, b, dom;
  uint32_t rk[DTO_FEISTEL_ROUNDS];
};

DTO_HD Shuffle shuffle_make(uint64_t seed, uint32_t job, uint32_t perm, uint32_t list, uint32_t n) {
  Shuffle s;
  s.n = n;
  uint32_t bl = 0; /* ceil(log2 n) */
  while (bl < 31 && (1u << bl) < n) ++bl;
  if (bl <= 12) {
    s.pa = 6;
    s.b = 64;
  } else {
    s.pa = bl / 2;
    s.b = (n + (1u << s.pa) - 1) >> s.pa;
  }
  s.amask = (1u << s.pa) - 1;
  s.dom = s.b << s.pa;
  U32x4 c = {perm, job, list, 0x44544F32u};
  U32x4 a = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  c.w ^= 0x5A5A5A5Au;
  U32x4 bb = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  s.rk[0] = a.x;
  s.rk[1] = a.y;
  s.rk[2] = a.z;
  s.rk[3] = a.w;
  s.rk[4] = bb.x;
  s.rk[5] = bb.y;
  s.rk[6] = bb.z;
  s.rk[7] = bb.w;
  return s;
}

/* Round function: two multiply-xorshift steps on 24-bit operands. */
DTO_HD uint32_t feistel_mix(uint32_t v, uint32_t key) {
  uint32_t h = ((v ^ key) & 0xFFFFFFu) * 0x9E3779u;
  h ^= h >> 15;
  h = (h & 0xFFFFFFu) * 0x85EBCBu;
  h ^= h >> 13;
  return h ^ (key >> 24);
}

/* F(v): the Feistel bijection of [0, dom). */
DTO_HD uint32_t feistel_once(const Shuffle &s, uint32_t v) {
  uint32_t L = v & s.amask, R = v >> s.pa;
#pragma unroll
  for (int r = 0; r < DTO_FEISTEL_ROUNDS; ++r) {
    if ((r & 1) == 0) {
      L = (L + feistel_mix(R, s.rk[r])) & s.amask;
    } else {
      /* the mix's low 16 bits scaled into [0, b) (b <= 2^16, so the product
       * fits; the mask on b, a no-op, lets the GPU use a 24-bit multiply) */
      R += ((feistel_mix(L, s.rk[r]) & 0xFFFFu) * (s.b & 0xFFFFFFu)) >> 16;
      R = R >= s.b ? R - s.b : R;
    }
  }
  return (R << s.pa) | L;
}

/* The whole permutation (host): out[k] = idx[k], k < n. */
static inline void shuffle_fill(const Shuffle &s, uint32_t *out) {
  uint32_t k = 0;
  for (uint32_t p = 0; p < s.dom; ++p) {
    const uint32_t v = feistel_once(s, p);
    if (v < s.n) out[k++] = v;
  }
}

}  // namespace dto

/*
 * C ABI of the DTO library (include/dto.h): device residency of jobs, kernel
 * launches, and the reference's run / empirical p-value / JSON steps.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dto.h"
#include "dto_host.hpp"
#include "dto_kernels.h"
#include "dto_internal.h"

namespace dto {
const char *last_error_cstr();
}

using namespace dto;



/* ------------------------------------------------------------------ */
/* devices                                                               */
/* ------------------------------------------------------------------ */
static std::mutex g_dev_mu;
static DevProps g_props[64];

int dto::current_device(int *dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(DTO_ERR_NO_DEVICE, "no HIP device available: the DTO kernels need an MI355X (gfx950)");
  }
  HIP_TRY(hipGetDevice(dev));
  return DTO_OK;
}

int dto::device_props(int dev, DevProps *out) {
  std::lock_guard<std::mutex> g(g_dev_mu);
  if (dev < 0 || dev >= 64) return fail(DTO_ERR_DEVICE, "device id out of range");
  if (!g_props[dev].ok) {
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    g_props[dev].cus = p.multiProcessorCount;
    g_props[dev].lds_per_block = p.sharedMemPerBlock;
    g_props[dev].ok = 1;
  }
  *out = g_props[dev];
  return DTO_OK;
}

/* ------------------------------------------------------------------ */
/* job                                                                   */
/* ------------------------------------------------------------------ */
static uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

LdsLayout dto::make_layout(const HostJob &h, int W, bool with_h, size_t lds_limit, bool nb_lds) {
  LdsLayout L{};
  const uint32_t N1 = (uint32_t)h.l1.ids.size(), N2 = (uint32_t)h.l2.ids.size();
  const uint32_t T1 = (uint32_t)h.thr1.size(), T2 = (uint32_t)h.thr2.size();
  const uint32_t T2w = (T2 + 1) / 2;
  L.T2w_pad = (T2w + 4u) & ~3u; /* at least one zero padding word */
  /* the screen table (or, without it, the pair staging) sits at offset 0 so
   * that table reads address it with the scaled index alone */
  /* pairs: u32 (row << 16 | column) where they alias the screen table, else the row alone as a u16 */
  const uint32_t hb = with_h ? 4u * (uint32_t)(h.N + 1) : 0u, pb = (with_h ? 4u : 2u) * std::max(N2, 1u);
  uint32_t fixed = 0;
  auto region = [&](uint32_t &field, uint32_t bytes) {
    field = fixed;
    fixed = align16(fixed + bytes);
  };
  region(L.nb, nb_lds ? 2u * std::max(std::max(N1, N2), 1u) : 0u);
  L.nb_in_lds = nb_lds ? 1u : 0u;
  region(L.rowstart, 4u * (T1 + 1));
  /* cursor is used while building the rows, cand0 in the reverse sweep */
  region(L.cursor, std::max(4u * T1, 8u * (uint32_t)W * (uint32_t)kCand0Cap));
  L.cand0 = L.cursor;
  region(L.cw, 4u * (uint32_t)W * L.T2w_pad);
  region(L.rowmin, 4u * T1);
  region(L.coln, 4u * T2);
  region(L.colb, 4u * T2);
  region(L.cand, 8u * (uint32_t)W * (uint32_t)kCandCap);
  region(L.cand3, 16u * (uint32_t)W * (uint32_t)kCand3Cap);
  region(L.exptab, 8u * 256u);
  region(L.red, (16u + 3u * (uint32_t)sizeof(BestKey) + 16u) * (uint32_t)W);
  region(L.misc, 16u);
  uint32_t base = 0;
  L.h = L.pairs = 0;
  if (with_h) {
    L.h_in_lds = 1;
    L.pairs_in_lds = 1;
    base = align16(std::max(hb, pb));
  } else if (align16(pb) + fixed <= lds_limit) {
    L.h_in_lds = 0;
    L.pairs_in_lds = 1;
    base = align16(pb);
  } else {
    L.h_in_lds = 0;
    L.pairs_in_lds = 0;
  }
  for (uint32_t *f : {&L.nb, &L.rowstart, &L.cursor, &L.cw, &L.rowmin, &L.coln, &L.colb, &L.cand, &L.cand3,
                      &L.exptab, &L.red, &L.misc})
    *f += base;
  L.cand0 = L.cursor;
  L.total = base + fixed;
  /* the batch kernel keeps its work item in the largest layout's misc slot,
   * which must therefore end the layout */
  if (L.misc + 16u != L.total) abort();
  return L;
}

template <typename T>
static size_t put(std::vector<unsigned char> &buf, const std::vector<T> &v) {
  size_t off = (buf.size() + 255) & ~(size_t)255;
  buf.resize(off + std::max<size_t>(v.size() * sizeof(T), 4));
  if (!v.empty()) memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

LdsLayout dto::screen_layout(const HostJob &h, int W, size_t lim) {
  LdsLayout L = make_layout(h, W, true, lim, true);
  if (L.total <= lim) return L;
  L = make_layout(h, W, true, lim, false);
  if (L.total <= lim) return L;
  return make_layout(h, W, false, lim, true);
}

int prepare_job(dto_job *j) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  DevProps pr;
  if ((rc = device_props(dev, &pr))) return rc;
  const HostJob &h = j->h;
  /* LDS: keep the screen table resident when it fits */
  const size_t lim = pr.lds_per_block;
  LdsLayout without_h = make_layout(h, kNT / 64, false, lim);
  if (without_h.total > lim)
    return fail(DTO_ERR_TOO_LARGE, "job needs " + std::to_string(without_h.total) + " B of LDS per workgroup, device has " +
                                       std::to_string(lim));
  j->lo_plain = without_h;
  j->lo_screen = screen_layout(h, kNT / 64, lim);
  std::vector<unsigned char> &buf = j->image;
  buf.clear();
  ArenaOffsets &o = j->off;
  o.bin1 = put(buf, h.bin1);
  o.bin2 = put(buf, h.bin2);
  o.join = put(buf, h.join);
  o.joff = put(buf, h.join_off);
  o.jpos = put(buf, h.join_pos);
  o.K = put(buf, h.K);
  o.n = put(buf, h.n);
  o.thr1 = put(buf, h.thr1);
  o.thr2 = put(buf, h.thr2);
  o.LF = put(buf, h.LF);
  o.D = put(buf, h.D);
  o.H = put(buf, h.H32);
  o.A = put(buf, h.A32);
  o.B = put(buf, h.B32);
  o.err = put(buf, std::vector<int>(4, 0));
  o.ctr = put(buf, std::vector<uint32_t>(kCounterRing, 0));
  o.stats = put(buf, std::vector<unsigned long long>(16, 0));
  buf.resize((buf.size() + 255) & ~(size_t)255);
  j->arena_bytes = buf.size();
  j->device = dev;
  DevJob &d = j->dev;
  d = DevJob{};
  d.N = (uint32_t)h.N;
  d.N1 = (uint32_t)h.l1.ids.size();
  d.N2 = (uint32_t)h.l2.ids.size();
  d.T1 = (uint32_t)h.thr1.size();
  d.T2 = (uint32_t)h.thr2.size();
  d.T2w = (d.T2 + 1) / 2;
  d.dup = h.dup ? 1u : 0u;
  d.kle = h.kle ? 1u : 0u;
  d.margin = h.margin;
  return DTO_OK;
}

void bind_job(dto_job *j, unsigned char *b) {
  const ArenaOffsets &o = j->off;
  DevJob &d = j->dev;
  d.bin1 = (const uint16_t *)(b + o.bin1);
  d.bin2 = (const uint16_t *)(b + o.bin2);
  d.join = (const int32_t *)(b + o.join);
  d.join_off = (const uint32_t *)(b + o.joff);
  d.join_pos = (const uint32_t *)(b + o.jpos);
  d.K = (const uint32_t *)(b + o.K);
  d.n = (const uint32_t *)(b + o.n);
  d.thr1 = (const uint32_t *)(b + o.thr1);
  d.thr2 = (const uint32_t *)(b + o.thr2);
  d.LF = (const double *)(b + o.LF);
  d.D = (const double *)(b + o.D);
  d.H32 = (const float *)(b + o.H);
  d.A32 = (const float *)(b + o.A);
  d.B32 = (const float *)(b + o.B);
  j->d_err = (int *)(b + o.err);
  j->d_ctr = (uint32_t *)(b + o.ctr);
  d.stats = (unsigned long long *)(b + o.stats);
  j->resident = true;
}

int upload_job(dto_job *j) {
  int rc = prepare_job(j);
  if (rc) return rc;
  DevProps pr;
  if ((rc = device_props(j->device, &pr))) return rc;
  if (j->arena) {
    (void)hipFree(j->arena);
    j->arena = nullptr;
  }
  j->resident = false;
  HIP_TRY(hipMalloc(&j->arena, j->image.size()));
  HIP_TRY(hipMemcpy(j->arena, j->image.data(), j->image.size(), hipMemcpyHostToDevice));
  bind_job(j, (unsigned char *)j->arena);
  std::vector<unsigned char>().swap(j->image); /* the device copy is the job's now */
  /* global pair staging for layouts whose pairs do not fit in LDS */
  if (j->ws) {
    (void)hipFree(j->ws);
    j->ws = nullptr;
  }
  if (j->gnb) {
    (void)hipFree(j->gnb);
    j->gnb = nullptr;
  }
  j->ws_slots = (uint32_t)pr.cus * 4u;
  if (!j->lo_plain.pairs_in_lds || !j->lo_screen.pairs_in_lds) {
    HIP_TRY(hipMalloc(&j->ws, (size_t)j->ws_slots * std::max<uint32_t>(j->dev.N2, 1) * sizeof(uint16_t)));
    if (!j->ws_ev) HIP_TRY(hipEventCreateWithFlags(&j->ws_ev, hipEventDisableTiming));
  }
  if (!j->lo_screen.nb_in_lds) {
    HIP_TRY(hipMalloc(&j->gnb, (size_t)j->ws_slots * std::max(std::max(j->dev.N1, j->dev.N2), 1u) * sizeof(uint16_t)));
    if (!j->ws_ev) HIP_TRY(hipEventCreateWithFlags(&j->ws_ev, hipEventDisableTiming));
  }
  return DTO_OK;
}

int ensure_resident(dto_job *j) {
  if (j->resident) return DTO_OK;
  return upload_job(j);
}

/* ------------------------------------------------------------------ */
/* launches                                                              */
/* ------------------------------------------------------------------ */
/* Launches that stage their pairs or keep their bins in the job's global
 * buffers (lists too long for the LDS) run one after another, whatever their
 * streams: each waits for the previous one's completion event.  Inside a stream capture the wait is
 * left out (a capture cannot wait on work recorded outside it), so captured
 * launches of one such job must not overlap. */
struct StagingOrder {
  dto_job *j;
  hipStream_t st;
  bool on = false;
  StagingOrder(const dto_job *job, const LdsLayout &lo, hipStream_t s) : j(const_cast<dto_job *>(job)), st(s) {
    if ((lo.pairs_in_lds && lo.nb_in_lds) || !j->ws_ev) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) (void)hipGetLastError();
    if (cs != hipStreamCaptureStatusNone) return;
    j->ws_mu.lock();
    on = true;
    if (j->ws_used) (void)hipStreamWaitEvent(st, j->ws_ev, 0);
  }
  ~StagingOrder() {
    if (!on) return;
    if (hipEventRecord(j->ws_ev, st) == hipSuccess) j->ws_used = true;
    j->ws_mu.unlock();
  }
};

template <bool E, bool HL, bool PL, bool PA, bool RC, bool NBL = true>
static hipError_t launch_opt(const dto_job *j, const LdsLayout &lo, uint32_t grid, hipStream_t st,
                             const uint32_t *i1, const uint32_t *i2, uint64_t n_perm, uint64_t base, uint64_t seed,
                             uint32_t job_id, int32_t permuted, dto_record *out) {
  auto kern = k_optimize<kNT, E, HL, PL, PA, RC, NBL>;
  (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lo.total);
  /* a slot of the job's counter ring for this launch, zeroed on its stream */
  uint32_t *ctr = j->d_ctr + (const_cast<dto_job *>(j)->ctr_next.fetch_add(1u) % kCounterRing);
  StagingOrder order(j, lo, st);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, ctr);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kNT), lo.total, st, j->dev, lo, i1, i2, n_perm, base, seed, job_id,
                     permuted, out, j->d_err, j->ws, j->gnb, ctr);
  return hipGetLastError();
}

/* Lane segments fit in registers when a row is at most kSegWords words per lane. */
static bool regcols(const dto_job *j) { return (j->dev.T2w + 63) / 64 <= kSegWords; }

template <bool E, bool HL, bool PL>
static hipError_t launch_opt_r(const dto_job *j, const LdsLayout &lo, uint32_t grid, hipStream_t st,
                               const uint32_t *i1, const uint32_t *i2, uint64_t n_perm, uint64_t base,
                               uint64_t seed, uint32_t job_id, int32_t permuted, dto_record *out) {
  if (PL) /* the plain kernel reads column constants per cell; one variant */
    return lo.pairs_in_lds
               ? launch_opt<E, false, true, true, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out)
               : launch_opt<E, false, true, false, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out);
  const bool rc = regcols(j);
  if (HL && !lo.nb_in_lds) /* the table in LDS, the bins global (pairs alias the table) */
    return rc ? launch_opt<E, true, false, true, true, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id,
                                                              permuted, out)
              : launch_opt<E, true, false, true, false, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id,
                                                               permuted, out);
  if (lo.pairs_in_lds) {
    if (rc) return launch_opt<E, HL, false, true, true>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out);
    return launch_opt<E, HL, false, true, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out);
  }
  if (rc) return launch_opt<E, false, false, false, true>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out);
  return launch_opt<E, false, false, false, false>(j, lo, grid, st, i1, i2, n_perm, base, seed, job_id, permuted, out);
}

/* Number of workgroups for a permutation launch: enough to fill every CU at
 * the occupancy the LDS footprint allows, never more than the work. */
static uint32_t perm_grid(const dto_job *j, const LdsLayout &lo, uint64_t n_perm) {
  DevProps pr;
  if (device_props(j->device, &pr)) return 1;
  uint64_t per_cu = std::max<uint64_t>(1, pr.lds_per_block ? (160u * 1024u) / std::max<uint32_t>(lo.total, 1) : 1);
  per_cu = std::min<uint64_t>(per_cu, 4);
  uint64_t g = std::min<uint64_t>(n_perm, (uint64_t)pr.cus * per_cu);
  if (!lo.pairs_in_lds || !lo.nb_in_lds) g = std::min<uint64_t>(g, j->ws_slots); /* one global slot per workgroup */
  return (uint32_t)std::max<uint64_t>(g, 1);
}

static int launch_perms(dto_job *j, uint32_t mode, bool expl, const uint32_t *d1, const uint32_t *d2, uint64_t n_perm,
                        uint64_t base, uint64_t seed, uint32_t job_id, int32_t permuted, dto_record *d_out,
                        hipStream_t st) {
  if (n_perm == 0) return DTO_OK;
  if (n_perm >= (1ull << 32) - 65536ull) return fail(DTO_ERR_TOO_LARGE, "too many permutations in one launch");
  hipError_t e;
  if (mode == DTO_MODE_PLAIN) {
    const LdsLayout &lo = j->lo_plain;
    uint32_t g = perm_grid(j, lo, n_perm);
    e = expl ? launch_opt_r<true, false, true>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out)
             : launch_opt_r<false, false, true>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out);
  } else if (mode == DTO_MODE_SCREENED) {
    const LdsLayout &lo = j->lo_screen;
    uint32_t g = perm_grid(j, lo, n_perm);
    if (lo.h_in_lds)
      e = expl ? launch_opt_r<true, true, false>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out)
               : launch_opt_r<false, true, false>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out);
    else
      e = expl ? launch_opt_r<true, false, false>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out)
               : launch_opt_r<false, false, false>(j, lo, g, st, d1, d2, n_perm, base, seed, job_id, permuted, d_out);
  } else {
    return fail(DTO_ERR_INVALID_ARGUMENT, "unknown mode");
  }
  if (e != hipSuccess) return fail(DTO_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return DTO_OK;
}

static int check_err_flag(dto_job *j) {
  int h[1] = {0};
  HIP_TRY(hipMemcpy(h, j->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (h[0]) {
    int z = 0;
    HIP_TRY(hipMemcpy(j->d_err, &z, sizeof(int), hipMemcpyHostToDevice));
    if (h[0] & 2) return fail(DTO_ERR_DEVICE, "internal error: a permutation selected no grid cell");
    return fail(DTO_ERR_BAD_PERMUTATION, "a permutation index is out of range");
  }
  return DTO_OK;
}

/* explicit permutations must be permutations of [0, len) */
static int validate_perms(const uint32_t *idx, uint64_t P, uint64_t len) {
  if (!idx) return DTO_OK;
  std::vector<uint64_t> stamp(len, 0);
  for (uint64_t p = 0; p < P; ++p) {
    const uint32_t *row = idx + p * len;
    for (uint64_t k = 0; k < len; ++k) {
      uint32_t v = row[k];
      if (v >= len || stamp[v] == p + 1)
        return fail(DTO_ERR_BAD_PERMUTATION, "permutation " + std::to_string(p) + " is not a permutation of [0, " +
                                                 std::to_string(len) + ")");
      stamp[v] = p + 1;
    }
  }
  return DTO_OK;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

/* The debug grid of one permutation on a stream: overlap counts into kcnt
 * (T1 * T2 words), then every cell's exact record into out; an out-of-range
 * permutation index sets *err. */
hipError_t dto::launch_grid(dto_job *job, const uint32_t *d1, const uint32_t *d2, uint32_t *kcnt, dto_record *out,
                              int *err, hipStream_t st) {
  const size_t cells = (size_t)job->dev.T1 * job->dev.T2;
  const LdsLayout &lo = job->lo_plain;
  if (lo.pairs_in_lds) {
    (void)hipFuncSetAttribute((const void *)k_grid_counts<kNT, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lo.total);
    hipLaunchKernelGGL((k_grid_counts<kNT, true>), dim3(1), dim3(kNT), lo.total, st, job->dev, lo, d1, d2, kcnt, err,
                       job->ws);
  } else {
    (void)hipFuncSetAttribute((const void *)k_grid_counts<kNT, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lo.total);
    StagingOrder order(job, lo, st);
    hipLaunchKernelGGL((k_grid_counts<kNT, false>), dim3(1), dim3(kNT), lo.total, st, job->dev, lo, d1, d2, kcnt,
                       err, job->ws);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int32_t permuted = (d1 || d2) ? 1 : 0;
  const uint32_t g = (uint32_t)std::min<size_t>((cells + 255) / 256, 8192);
  hipLaunchKernelGGL(k_grid_pvalues, dim3(g), dim3(256), 0, st, job->dev, (const uint32_t *)kcnt, permuted, out);
  return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* C ABI                                                                  */
/* ------------------------------------------------------------------ */
extern "C" {

const char *dto_last_error(void) { return dto::last_error_cstr(); }

int dto_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

double dto_fdr(uint64_t l1, uint64_t l2, uint64_t overlap, uint64_t population, double sensitivity) {
  /* fdr.rs:29-60; where the reference panics (sensitivity <= 0, fdr.rs:36-39)
   * the result is NaN and dto_last_error() holds the panic message */
  if (sensitivity <= 0.0) {
    fail(DTO_ERR_INVALID_ARGUMENT, "Sensitivity must be greater than 0.");
    return __builtin_nan("");
  }
  double df = fmax((double)overlap / sensitivity, 0.0);
  double b = fmax((double)l1 - df, 0.0);
  double r = fmax((double)l2 - df, 0.0);
  double num = b * r;
  double den = (double)population * (double)overlap;
  return den > 0.0 ? num / den : 0.0;
}

int dto_fdr_checked(uint64_t l1, uint64_t l2, uint64_t overlap, uint64_t population, double sensitivity,
                    double *out) {
  if (!out) return fail(DTO_ERR_INVALID_ARGUMENT, "null output");
  if (sensitivity <= 0.0) return fail(DTO_ERR_INVALID_ARGUMENT, "Sensitivity must be greater than 0.");
  *out = dto_fdr(l1, l2, overlap, population, sensitivity);
  return DTO_OK;
}

uint32_t dto_generate_thresholds(uint32_t max_rank, uint32_t *out, uint32_t cap) {
  std::vector<uint32_t> t = generate_thresholds(max_rank);
  if (out)
    for (uint32_t i = 0; i < cap && i < t.size(); ++i) out[i] = t[i];
  return (uint32_t)t.size();
}

int dto_ln_factorial_table(uint64_t n_max, double *out) {
  if (!out) return fail(DTO_ERR_INVALID_ARGUMENT, "null output");
  for (uint64_t x = 0; x <= n_max; ++x) out[x] = ln_factorial(x);
  return DTO_OK;
}

int dto_hypergeometric_pvalue(const uint64_t *N, const uint64_t *K, const uint64_t *n, const uint64_t *k,
                              uint64_t count, double *out) {
  if (count == 0) return DTO_OK;
  if (!N || !K || !n || !k || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  uint64_t maxN = 0;
  for (uint64_t c = 0; c < count; ++c) {
    /* Hypergeometric::new(N, K, n).expect(...) */
    if (K[c] > N[c] || n[c] > N[c]) return fail(DTO_ERR_HYPERGEOMETRIC, "Failed to create hypergeometric distribution");
    if (N[c] >= (1ull << 31)) return fail(DTO_ERR_TOO_LARGE, "population size too large");
    maxN = std::max(maxN, N[c]);
  }
  int dev, rc;
  if ((rc = current_device(&dev))) return rc;
  std::vector<double> LF(maxN + 1);
  for (uint64_t x = 0; x <= maxN; ++x) LF[x] = ln_factorial(x);
  DevBuf dLF, dq, dout;
  HIP_TRY(hipMalloc(&dLF.p, LF.size() * 8));
  HIP_TRY(hipMalloc(&dq.p, count * 8 * 4));
  HIP_TRY(hipMalloc(&dout.p, count * 8));
  uint64_t *q = (uint64_t *)dq.p;
  HIP_TRY(hipMemcpy(dLF.p, LF.data(), LF.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q, N, count * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q + count, K, count * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q + 2 * count, n, count * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(q + 3 * count, k, count * 8, hipMemcpyHostToDevice));
  uint32_t g = (uint32_t)std::min<uint64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_hypergeom, dim3(g), dim3(256), 0, 0, (const double *)dLF.p, q, q + count, q + 2 * count,
                     q + 3 * count, count, (double *)dout.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dout.p, count * 8, hipMemcpyDeviceToHost));
  return DTO_OK;
}

/* internal test hooks (not in dto.h): device exp and device shuffle */
int dto_internal_exp(const double *x, uint64_t count, double *out) {
  int dev, rc;
  if ((rc = current_device(&dev))) return rc;
  DevBuf dx, dy;
  HIP_TRY(hipMalloc(&dx.p, count * 8));
  HIP_TRY(hipMalloc(&dy.p, count * 8));
  HIP_TRY(hipMemcpy(dx.p, x, count * 8, hipMemcpyHostToDevice));
  uint32_t g = (uint32_t)std::min<uint64_t>((count + 255) / 256, 8192);
  hipLaunchKernelGGL(k_exp, dim3(g), dim3(256), 0, 0, (const double *)dx.p, count, (double *)dy.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dy.p, count * 8, hipMemcpyDeviceToHost));
  return DTO_OK;
}

int dto_internal_exp_host(const double *x, uint64_t count, double *out) {
  static const uint64_t tab[256] = DTO_EXP_TAB_INIT;
  for (uint64_t c = 0; c < count; ++c) out[c] = dto_exp(x[c], tab);
  return DTO_OK;
}

/* one Philox4x32-10 block (host), for the published known-answer vectors */
int dto_internal_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  if (!ctr || !key || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  U32x4 c;
  c.x = ctr[0];
  c.y = ctr[1];
  c.z = ctr[2];
  c.w = ctr[3];
  const U32x4 r = philox4x32_10(c, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
  return DTO_OK;
}

int dto_internal_shuffle_device(uint64_t seed, uint32_t job, uint32_t perm, uint32_t list, uint32_t n, uint32_t *out) {
  int dev, rc;
  if ((rc = current_device(&dev))) return rc;
  DevBuf d;
  HIP_TRY(hipMalloc(&d.p, std::max<uint32_t>(n, 1) * 4));
  hipLaunchKernelGGL(k_shuffle, dim3(1), dim3(256), 0, 0, seed, job, perm, list, n, (uint32_t *)d.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, d.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return DTO_OK;
}

/* work counters of a job's screened launches (test / profiling hook, not in
 * dto.h), 16 values: exact cells, exact terms, rescanned rows, cycles of the
 * build / forward / reverse / tier-1 flush / exact phases, tier-1 pushes,
 * tier-2 survivors, tier-1 flushes, first-bound gap (milli-nats), cells past
 * the reverse sweep's first-term test */
int dto_internal_stats(dto_job *job, uint64_t *out, int reset) {
  if (!job || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  if (!DTO_STATS) return fail(DTO_ERR_INVALID_ARGUMENT, "the library was built without DTO_STATS (make STATS=1)");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(out, job->dev.stats, 16 * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (reset) HIP_TRY(hipMemset(job->dev.stats, 0, 16 * sizeof(uint64_t)));
  return DTO_OK;
}

int dto_philox_permutation(uint64_t seed, uint32_t job_id, uint32_t perm, uint32_t list, uint32_t len, uint32_t *out) {
  if (!out && len) return fail(DTO_ERR_INVALID_ARGUMENT, "null output");
  Shuffle s = shuffle_make(seed, job_id, perm, list, len);
  shuffle_fill(s, out);
  return DTO_OK;
}

int dto_job_create(const char *const *ids1, const uint32_t *ranks1, uint64_t n1, const char *const *ids2,
                   const uint32_t *ranks2, uint64_t n2, const char *const *bg, uint64_t nbg, uint32_t flags,
                   dto_job **out) {
  if (!out) return fail(DTO_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  dto_job *j = new dto_job();
  int rc = build_host_job(ids1, ranks1, n1, ids2, ranks2, n2, bg, nbg, flags & ~DTO_DEFER_UPLOAD, j->h);
  if (!rc) rc = (flags & DTO_DEFER_UPLOAD) ? prepare_job(j) : upload_job(j);
  if (rc) {
    dto_job_destroy(j);
    return rc;
  }
  *out = j;
  return DTO_OK;
}

int dto_job_create_with_population(const char *const *ids1, const uint32_t *ranks1, uint64_t n1,
                                   const char *const *ids2, const uint32_t *ranks2, uint64_t n2,
                                   uint64_t population_size, dto_job **out) {
  if (!out) return fail(DTO_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  dto_job *j = new dto_job();
  int rc = build_host_job_with_population(ids1, ranks1, n1, ids2, ranks2, n2, population_size, j->h);
  if (!rc) rc = upload_job(j);
  if (rc) {
    dto_job_destroy(j);
    return rc;
  }
  *out = j;
  return DTO_OK;
}

int dto_job_create_from_files(const char *path1, const char *path2, const char *bg_path, uint32_t flags,
                              dto_job **out) {
  std::vector<std::string> i1, i2, ib;
  std::vector<uint32_t> r1, r2;
  int rc;
  if ((rc = read_ranked_csv(path1, i1, r1))) return rc;
  if ((rc = read_ranked_csv(path2, i2, r2))) return rc;
  if (bg_path && (rc = read_feature_lines(bg_path, ib))) return rc;
  std::vector<const char *> p1, p2, pb;
  for (auto &s : i1) p1.push_back(s.c_str());
  for (auto &s : i2) p2.push_back(s.c_str());
  for (auto &s : ib) pb.push_back(s.c_str());
  static const char *empty_bg[1] = {""};
  const char *const *bgp = bg_path ? (pb.empty() ? empty_bg : pb.data()) : nullptr;
  return dto_job_create(p1.data(), r1.data(), i1.size(), p2.data(), r2.data(), i2.size(), bgp, pb.size(), flags, out);
}

int dto_job_set_thresholds(dto_job *job, const uint32_t *thr1, uint32_t t1, const uint32_t *thr2, uint32_t t2) {
  if (!job || (t1 && !thr1) || (t2 && !thr2)) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  if (job->resident && !job->arena) return fail(DTO_ERR_INVALID_ARGUMENT, "the job belongs to a server");
  const bool own = job->resident;
  HostJob saved = job->h;
  job->h.thr1.assign(thr1, thr1 + t1);
  job->h.thr2.assign(thr2, thr2 + t2);
  int rc = job->h.derive();
  if (!rc) rc = own ? upload_job(job) : prepare_job(job);
  if (rc) {
    std::string msg = dto_last_error();
    job->h = saved;
    (void)(own ? upload_job(job) : prepare_job(job));
    return fail(rc, msg);
  }
  return DTO_OK;
}

int dto_job_info_get(const dto_job *job, dto_job_info *info) {
  if (!job || !info) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  info->population_size = job->h.N;
  info->len1 = job->h.l1.ids.size();
  info->len2 = job->h.l2.ids.size();
  info->n_thresholds1 = (uint32_t)job->h.thr1.size();
  info->n_thresholds2 = (uint32_t)job->h.thr2.size();
  info->dropped1 = job->h.dropped1;
  info->dropped2 = job->h.dropped2;
  info->device_bytes = job->arena_bytes;
  return DTO_OK;
}

int dto_job_thresholds(const dto_job *job, int list, uint32_t *out, uint32_t cap) {
  if (!job || (list != 1 && list != 2)) return -fail(DTO_ERR_INVALID_ARGUMENT, "bad job or list");
  const std::vector<uint32_t> &t = list == 1 ? job->h.thr1 : job->h.thr2;
  if (out)
    for (uint32_t i = 0; i < cap && i < t.size(); ++i) out[i] = t[i];
  return (int)t.size();
}

void dto_job_destroy(dto_job *job) {
  if (!job) return;
  if (job->arena) (void)hipFree(job->arena);
  if (job->ws) (void)hipFree(job->ws);
  if (job->gnb) (void)hipFree(job->gnb);
  if (job->ws_ev) (void)hipEventDestroy(job->ws_ev);
  delete job;
}

int dto_run_device(dto_job *job, uint64_t n_perm, uint64_t perm_base, uint64_t seed, const uint32_t *d_idx1,
                   const uint32_t *d_idx2, uint32_t mode, dto_record *d_out, void *stream) {
  if (!job || (!d_out && n_perm)) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  int rc;
  if ((rc = ensure_resident(job))) return rc;
  bool expl = d_idx1 || d_idx2;
  return launch_perms(job, mode, expl, d_idx1, d_idx2, n_perm, perm_base, seed, 0, 1, d_out, (hipStream_t)stream);
}

int dto_optimize(dto_job *job, const uint32_t *idx1, const uint32_t *idx2, uint32_t mode, dto_record *out) {
  if (!job || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  int rc;
  if ((rc = ensure_resident(job))) return rc;
  if ((rc = validate_perms(idx1, 1, job->dev.N1)) || (rc = validate_perms(idx2, 1, job->dev.N2))) return rc;
  DevBuf di, dr;
  const size_t n1 = job->dev.N1, n2 = job->dev.N2;
  HIP_TRY(hipMalloc(&di.p, (n1 + n2 + 2) * 4));
  HIP_TRY(hipMalloc(&dr.p, sizeof(dto_record)));
  uint32_t *d1 = nullptr, *d2 = nullptr;
  if (idx1) {
    d1 = (uint32_t *)di.p;
    HIP_TRY(hipMemcpy(d1, idx1, n1 * 4, hipMemcpyHostToDevice));
  }
  if (idx2) {
    d2 = (uint32_t *)di.p + n1 + 1;
    HIP_TRY(hipMemcpy(d2, idx2, n2 * 4, hipMemcpyHostToDevice));
  }
  int32_t permuted = (idx1 || idx2) ? 1 : 0;
  if ((rc = launch_perms(job, mode, true, d1, d2, 1, 0, 0, 0, permuted, (dto_record *)dr.p, 0))) return rc;
  HIP_TRY(hipMemcpy(out, dr.p, sizeof(dto_record), hipMemcpyDeviceToHost));
  return check_err_flag(job);
}

int dto_process_threshold_pairs(dto_job *job, const uint32_t *idx1, const uint32_t *idx2, dto_record *out) {
  if (!job || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  int rc;
  if ((rc = ensure_resident(job))) return rc;
  if ((rc = validate_perms(idx1, 1, job->dev.N1)) || (rc = validate_perms(idx2, 1, job->dev.N2))) return rc;
  const size_t n1 = job->dev.N1, n2 = job->dev.N2, cells = (size_t)job->dev.T1 * job->dev.T2;
  DevBuf di, dk, dr;
  HIP_TRY(hipMalloc(&di.p, (n1 + n2 + 2) * 4));
  HIP_TRY(hipMalloc(&dk.p, cells * 4));
  HIP_TRY(hipMalloc(&dr.p, cells * sizeof(dto_record)));
  uint32_t *d1 = nullptr, *d2 = nullptr;
  if (idx1) {
    d1 = (uint32_t *)di.p;
    HIP_TRY(hipMemcpy(d1, idx1, n1 * 4, hipMemcpyHostToDevice));
  }
  if (idx2) {
    d2 = (uint32_t *)di.p + n1 + 1;
    HIP_TRY(hipMemcpy(d2, idx2, n2 * 4, hipMemcpyHostToDevice));
  }
  HIP_TRY(launch_grid(job, d1, d2, (uint32_t *)dk.p, (dto_record *)dr.p, job->d_err, 0));
  HIP_TRY(hipMemcpy(out, dr.p, cells * sizeof(dto_record), hipMemcpyDeviceToHost));
  return check_err_flag(job);
}


int dto_empirical_pvalue(const dto_record *unpermuted, const dto_record *perm, uint64_t n_perm, dto_result *result) {
  if (!unpermuted || !result || (n_perm && !perm)) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  /* empirical_pvalue.rs:136-173 */
  result->unpermuted = *unpermuted;
  result->fdr = dto_fdr(unpermuted->set1_len, unpermuted->set2_len, unpermuted->intersection_size,
                        unpermuted->population_size, 0.8);
  result->n_permutations = n_perm;
  uint64_t le = 0;
  for (uint64_t p = 0; p < n_perm; ++p) le += perm[p].pvalue <= unpermuted->pvalue;
  result->n_le = le;
  result->empirical_pvalue = n_perm ? (double)le / (double)n_perm : 1.0;
  return DTO_OK;
}

int dto_run(dto_job *job, uint64_t n_perm, uint64_t seed, const uint32_t *idx1, const uint32_t *idx2, uint32_t mode,
            dto_result *result, dto_record *perm_out) {
  if (!job || !result) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  if ((idx1 == nullptr) != (idx2 == nullptr))
    return fail(DTO_ERR_INVALID_ARGUMENT, "give explicit permutations for both lists or neither");
  int rc;
  if ((rc = ensure_resident(job))) return rc;
  const uint64_t n1 = job->dev.N1, n2 = job->dev.N2;
  if ((rc = validate_perms(idx1, n_perm, n1)) || (rc = validate_perms(idx2, n_perm, n2))) return rc;
  DevBuf dr;
  HIP_TRY(hipMalloc(&dr.p, (n_perm + 1) * sizeof(dto_record)));
  dto_record *d_rec = (dto_record *)dr.p;
  /* task 0: unpermuted (main.rs:122-125) */
  if ((rc = launch_perms(job, mode, true, nullptr, nullptr, 1, 0, 0, 0, 0, d_rec, 0))) return rc;
  if (n_perm) {
    if (idx1) {
      /* explicit permutations, streamed in chunks of at most ~512 MiB */
      uint64_t chunk = std::max<uint64_t>(1, (512ull << 20) / ((n1 + n2) * 4));
      chunk = std::min(chunk, n_perm);
      DevBuf di;
      HIP_TRY(hipMalloc(&di.p, chunk * (n1 + n2) * 4));
      uint32_t *d1 = (uint32_t *)di.p, *d2 = d1 + chunk * n1;
      for (uint64_t p0 = 0; p0 < n_perm; p0 += chunk) {
        uint64_t c = std::min(chunk, n_perm - p0);
        HIP_TRY(hipMemcpy(d1, idx1 + p0 * n1, c * n1 * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d2, idx2 + p0 * n2, c * n2 * 4, hipMemcpyHostToDevice));
        if ((rc = launch_perms(job, mode, true, d1, d2, c, p0, seed, 0, 1, d_rec + 1 + p0, 0))) return rc;
        HIP_TRY(hipDeviceSynchronize());
      }
    } else {
      if ((rc = launch_perms(job, mode, false, nullptr, nullptr, n_perm, 0, seed, 0, 1, d_rec + 1, 0))) return rc;
    }
  }
  std::vector<dto_record> rec(n_perm + 1);
  HIP_TRY(hipMemcpy(rec.data(), d_rec, (n_perm + 1) * sizeof(dto_record), hipMemcpyDeviceToHost));
  if ((rc = check_err_flag(job))) return rc;
  if (perm_out && n_perm) memcpy(perm_out, rec.data() + 1, n_perm * sizeof(dto_record));
  return dto_empirical_pvalue(&rec[0], rec.data() + 1, n_perm, result);
}

int dto_result_json(const dto_result *result, char *buf, size_t cap) {
  if (!result) return -1;
  std::string s = result_json(*result);
  if (!buf || cap < s.size() + 1) return -1;
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // extern "C"

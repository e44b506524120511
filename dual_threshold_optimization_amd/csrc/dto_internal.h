/*
 * Library-internal state shared by the translation units of libdto_amd.so.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "dto_device.h"
#include "dto_host.hpp"

#define HIP_TRY(expr)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) return dto::fail(DTO_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

namespace dto {
constexpr int kNT = 512; /* threads per workgroup of the grid kernels (8 waves) */

struct DevProps {
  int ok = 0;
  int cus = 0;
  size_t lds_per_block = 0;
};
int current_device(int *dev);
int device_props(int dev, DevProps *out);
/* with_h: the screen table in LDS; nb_lds: the bins array in LDS (else a
 * per-workgroup global slice of max(N1, N2) u16). */
LdsLayout make_layout(const HostJob &h, int W, bool with_h, size_t lds_limit, bool nb_lds = true);
/* The screened kernels' layout for a job: table and bins in LDS when they
 * fit, else the table in LDS and the bins global, else the table global. */
LdsLayout screen_layout(const HostJob &h, int W, size_t lds_limit);
}  // namespace dto

constexpr uint32_t kCounterRing = 256;

/* Byte offsets of a job's tables inside its device arena. */
struct ArenaOffsets {
  size_t bin1, bin2, join, joff, jpos, K, n, thr1, thr2, LF, D, H, A, B, err, ctr, stats;
};

/* A job: host tables plus their device copy (one HBM arena).  The arena is
 * the job's own allocation (`arena`), or, for a job created with
 * DTO_DEFER_UPLOAD and admitted by a server, a region of the server's pool
 * (`arena` stays NULL; the server owns the memory). */
struct dto_job {
  dto::HostJob h;
  dto::DevJob dev{};
  void *arena = nullptr;
  size_t arena_bytes = 0; /* size of the arena image */
  bool resident = false;  /* dev points at an arena (own or pooled) */
  std::vector<unsigned char> image; /* the arena's contents (kept for deferred jobs) */
  ArenaOffsets off{};
  int *d_err = nullptr;
  /* work counters of permutation launches: a ring of kCounterRing slots, one
   * per launch, so that launches on different streams do not share one (a
   * job with more launches than that in flight at once would) */
  uint32_t *d_ctr = nullptr;
  std::atomic<uint32_t> ctr_next{0};
  int device = -1;
  dto::LdsLayout lo_screen{}, lo_plain{};
  uint16_t *ws = nullptr; /* global pair staging (only when the LDS cannot hold the pairs) */
  uint16_t *gnb = nullptr; /* global bins arrays, one per workgroup (only when lo_screen.nb_in_lds == 0) */
  uint32_t ws_slots = 0;
  /* launches using ws run in order: each waits for the previous one's event */
  hipEvent_t ws_ev = nullptr;
  bool ws_used = false;
  std::mutex ws_mu;
};
/* Tables, layouts and arena image of a job on the current device (no device
 * memory); then either its own arena (upload_job) or a caller's region
 * (bind_job, after the caller has copied j->image there). */
int prepare_job(dto_job *j);
int upload_job(dto_job *j);
void bind_job(dto_job *j, unsigned char *base);
/* Uploads a deferred job into its own arena on first use; fails when the
 * job sits in a server's pool (the server owns it). */
int ensure_resident(dto_job *j);
namespace dto {
/* The debug grid of one permutation on a stream (k_grid_counts then
 * k_grid_pvalues): overlap counts into kcnt (T1 * T2 words), every cell's
 * record into out; an out-of-range permutation index sets *err. */
hipError_t launch_grid(dto_job *job, const uint32_t *d1, const uint32_t *d2, uint32_t *kcnt, dto_record *out, int *err,
                       hipStream_t st);
}  // namespace dto

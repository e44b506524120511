/*
 * Library-internal state shared by the translation units of libdto_amd.so.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>

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
LdsLayout make_layout(const HostJob &h, int W, bool with_h, size_t lds_limit);
}  // namespace dto

constexpr uint32_t kCounterRing = 256;

/* A job: host tables plus their device copy (one HBM arena). */
struct dto_job {
  dto::HostJob h;
  dto::DevJob dev{};
  void *arena = nullptr;
  size_t arena_bytes = 0;
  int *d_err = nullptr;
  /* work counters of permutation launches: a ring of kCounterRing slots, one
   * per launch, so that launches on different streams do not share one (a
   * job with more launches than that in flight at once would) */
  uint32_t *d_ctr = nullptr;
  std::atomic<uint32_t> ctr_next{0};
  int device = -1;
  dto::LdsLayout lo_screen{}, lo_plain{};
  uint32_t *ws = nullptr; /* global pair staging (only when the LDS cannot hold the pairs) */
  uint32_t ws_slots = 0;
};
int upload_job(dto_job *j);

/*
 * dto_process_threshold_pairs_noalloc: the debug grid of one permutation for
 * callers that cannot allocate device memory.  Its only scratch (the overlap
 * counts and a status word) is a file-scope device array of fixed size.
 */
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/dto.h"
#include "dto_internal.h"
#include "dto_kernels.h"

using namespace dto;

/* Device scratch of dto_process_threshold_pairs_noalloc: a status word
 * (used when the caller passes none) and the overlap counts, in one
 * file-scope array that every GPU the library runs on holds until the
 * process ends.  In host code the array's name is only a handle for the
 * symbol: the device address comes from hipGetSymbolAddress, once per device
 * with that device current, and only that address reaches a kernel. */
constexpr uint64_t kScratchHead = 64; /* words before the counts (keeps them 256-byte aligned) */
static __device__ __align__(256) uint32_t scratch[kScratchHead + DTO_NOALLOC_MAX_CELLS];
static std::mutex g_scratch_mu;
static uint32_t *g_scratch[64];

static int scratch_of_current_device(int dev, uint32_t **out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (dev < 0 || dev >= 64) return fail(DTO_ERR_DEVICE, "device id out of range");
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DTO_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = (uint32_t *)p;
  }
  *out = g_scratch[dev];
  return DTO_OK;
}

extern "C" int dto_process_threshold_pairs_noalloc(dto_job *job, const uint32_t *d_idx1, const uint32_t *d_idx2,
                                                   dto_record *d_out, int *d_status, void *stream) {
  if (!job || !d_out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  const uint64_t cells = (uint64_t)job->dev.T1 * job->dev.T2;
  if (cells > DTO_NOALLOC_MAX_CELLS)
    return fail(DTO_ERR_TOO_LARGE, "dto_process_threshold_pairs_noalloc: the grid has " + std::to_string(cells) +
                                       " cells, more than DTO_NOALLOC_MAX_CELLS (" +
                                       std::to_string((unsigned long long)DTO_NOALLOC_MAX_CELLS) + ")");
  if (!job->resident)
    return fail(DTO_ERR_INVALID_ARGUMENT, "the job's tables are not on the device yet (DTO_DEFER_UPLOAD): "
                                          "run it once with another entry point first");
  int dev, rc;
  if ((rc = current_device(&dev))) return rc;
  if (dev != job->device)
    return fail(DTO_ERR_INVALID_ARGUMENT, "the job lives on device " + std::to_string(job->device) +
                                              ", the current device is " + std::to_string(dev));
  uint32_t *scr = nullptr;
  if ((rc = scratch_of_current_device(dev, &scr))) return rc;
  hipStream_t st = (hipStream_t)stream;
  int *status = d_status ? d_status : (int *)scr;
  /* the status word must start at zero (the kernel only sets bits); a
   * kernel clears it, not a memset, so that the call can be captured */
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, (uint32_t *)status);
  HIP_TRY(hipGetLastError());
  HIP_TRY(launch_grid(job, d_idx1, d_idx2, scr + kScratchHead, d_out, status, st));
  return DTO_OK;
}

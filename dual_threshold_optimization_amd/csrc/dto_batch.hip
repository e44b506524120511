/*
 * Batched jobs and the serving loop.
 *
 * The reference runs one job per process and spreads its permutations over
 * std::thread chunks (src/run/single_node.rs:83-137) or MPI ranks
 * (src/run/multi_node.rs:47-162).  Here many jobs of uneven size share one
 * persistent kernel (k_batch): every workgroup pulls (job, task range) items
 * from a device queue until it is empty.  A batch's launch sequence (queue
 * reset + persistent kernel) is captured once in a hipGraph and replayed.
 *
 * The server does continuous batching: requests are admitted, oldest first,
 * while their tables and records fit a device pool of hbm_budget_bytes (one
 * allocation at creation; admission and completion only move regions of it,
 * so the poll loop never calls hipMalloc or hipFree).  Requests are grouped
 * by the launch shape their job needs alone; a slice takes up to
 * perms_per_slice tasks from every active request of one shape class, and up
 * to four slices are in flight on four streams, so requests that arrive while
 * others run join the next slice and overlap with them on the device.
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dto.h"
#include "dto_internal.h"
#include "dto_kernels.h"

using namespace dto;

namespace {

/* Launch shapes the batch kernels are built for: threads per workgroup, the
 * screen table in LDS or not, the pair staging in LDS or global memory. */
/* Global buffers of a launch: pair staging (ws, ws_stride u16 per
 * workgroup) and bins arrays (gnb, gnb_stride u16 per workgroup). */
struct Globals {
  uint16_t *ws = nullptr, *gnb = nullptr;
  uint32_t ws_stride = 0, gnb_stride = 0;
};

template <int NT, bool HL, bool PL, bool RC, bool NBL>
hipError_t launch_batch_t(uint32_t grid, uint32_t lds, hipStream_t st, const BatchJob *jobs, const WorkItem *items,
                          uint32_t n_items, uint64_t seed, uint32_t *counter, int *err, const Globals &g) {
  auto kern = k_batch<NT, HL, PL, RC, NBL>;
  (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  /* every layout ends with its 16-byte misc slot: the largest one's is free in all */
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, jobs, items, n_items, seed, counter, err, g.ws, g.ws_stride,
                     g.gnb, g.gnb_stride, lds - 16u);
  return hipGetLastError();
}

template <int NT>
hipError_t launch_batch_n(bool hl, bool nbl, bool pl, bool rc, uint32_t grid, uint32_t lds, hipStream_t st,
                          const BatchJob *jobs, const WorkItem *items, uint32_t n_items, uint64_t seed,
                          uint32_t *counter, int *err, const Globals &g) {
  if (hl && nbl)
    return rc ? launch_batch_t<NT, true, true, true, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g)
              : launch_batch_t<NT, true, true, false, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g);
  if (hl)
    return rc ? launch_batch_t<NT, true, true, true, false>(grid, lds, st, jobs, items, n_items, seed, counter, err, g)
              : launch_batch_t<NT, true, true, false, false>(grid, lds, st, jobs, items, n_items, seed, counter, err, g);
  if (pl)
    return rc ? launch_batch_t<NT, false, true, true, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g)
              : launch_batch_t<NT, false, true, false, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g);
  return rc ? launch_batch_t<NT, false, false, true, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g)
            : launch_batch_t<NT, false, false, false, true>(grid, lds, st, jobs, items, n_items, seed, counter, err, g);
}

/* Batch-wide launch parameters for a set of jobs at nt threads. */
struct Shape {
  int nt = 512;
  bool hl = true;  /* screen table in LDS */
  bool nbl = true; /* bins arrays in LDS */
  bool pl = true;  /* pair staging in LDS */
  bool rc = true;  /* every job's lane segment fits in registers */
  uint32_t lds = 0;
  uint32_t max_n2 = 0, max_n = 0; /* largest N2, largest max(N1, N2) */
  std::vector<LdsLayout> lo;
};

hipError_t launch_batch(const Shape &sh, uint32_t grid, hipStream_t st, const BatchJob *jobs, const WorkItem *items,
                        uint32_t n_items, uint64_t seed, uint32_t *counter, int *err, const Globals &g) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, counter, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (sh.nt == 256)
    return launch_batch_n<256>(sh.hl, sh.nbl, sh.pl, sh.rc, grid, sh.lds, st, jobs, items, n_items, seed, counter,
                               err, g);
  return launch_batch_n<512>(sh.hl, sh.nbl, sh.pl, sh.rc, grid, sh.lds, st, jobs, items, n_items, seed, counter, err,
                             g);
}

/* One shape for all jobs of a launch: the first of (table and bins in
 * LDS), (table in LDS, bins global), (table global) that every job fits. */
int plan_shape(const std::vector<dto_job *> &jobs, int nt, size_t lds_limit, Shape &s) {
  s.nt = nt;
  const int W = nt / 64;
  bool fit_hb = true, fit_h = true, fit_p = true;
  s.max_n2 = 1;
  s.max_n = 1;
  s.rc = true;
  for (dto_job *j : jobs) {
    if ((j->dev.T2w + 63) / 64 > kSegWords) s.rc = false;
    if (make_layout(j->h, W, true, lds_limit, true).total > lds_limit) fit_hb = false;
    if (make_layout(j->h, W, true, lds_limit, false).total > lds_limit) fit_h = false;
    s.max_n2 = std::max(s.max_n2, j->dev.N2);
    s.max_n = std::max(s.max_n, std::max(j->dev.N1, j->dev.N2));
  }
  s.hl = fit_h;
  s.nbl = fit_hb || !fit_h;
  s.lo.clear();
  s.lds = 0;
  for (dto_job *j : jobs) {
    LdsLayout L = make_layout(j->h, W, s.hl, lds_limit, s.nbl);
    if (L.total > lds_limit) return fail(DTO_ERR_TOO_LARGE, "job does not fit in LDS");
    if (!L.pairs_in_lds) fit_p = false;
    s.lo.push_back(L);
    s.lds = std::max(s.lds, L.total);
  }
  s.pl = fit_p;
  return DTO_OK;
}

uint32_t grid_for(int cus, uint32_t lds, int nt, uint32_t n_items) {
  uint32_t by_lds = std::max<uint32_t>(1, (160u * 1024u) / std::max<uint32_t>(lds, 1));
  uint32_t by_threads = std::max<uint32_t>(1, 2048u / (uint32_t)nt);
  uint32_t per_cu = std::min(by_lds, by_threads);
  return std::max<uint32_t>(1, std::min<uint32_t>(n_items, (uint32_t)cus * per_cu));
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

/* ------------------------------------------------------------------ */
/* batch                                                                 */
/* ------------------------------------------------------------------ */
struct dto_batch {
  std::vector<dto_job *> jobs;
  std::vector<uint64_t> n_perm;
  uint64_t seed = 0;
  int device = -1;
  int cus = 0;
  size_t lds_limit = 0;
  Shape shape;
  uint32_t chunk = 8; /* tasks per work item */
  uint32_t grid = 0;
  std::vector<size_t> out_off;
  size_t total_tasks = 0;
  BatchJob *d_jobs = nullptr;
  WorkItem *d_items = nullptr;
  uint32_t n_items = 0;
  uint32_t *d_counter = nullptr;
  int *d_err = nullptr;
  dto_record *d_out = nullptr;
  uint16_t *d_ws = nullptr;  /* global pair staging when the LDS cannot hold the pairs */
  uint16_t *d_gnb = nullptr; /* global bins arrays when the LDS holds the table but not the bins */
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t own_stream = nullptr;
};

static void batch_free_plan(dto_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_ws) (void)hipFree(b->d_ws);
  if (b->d_gnb) (void)hipFree(b->d_gnb);
  b->d_jobs = nullptr;
  b->d_items = nullptr;
  b->d_ws = nullptr;
  b->d_gnb = nullptr;
}

static Globals batch_globals(const dto_batch *b) {
  Globals g;
  g.ws = b->d_ws;
  g.ws_stride = b->shape.max_n2;
  g.gnb = b->d_gnb;
  g.gnb_stride = b->shape.max_n;
  return g;
}

/* Builds the job table and the work queue for the current shape and chunk. */
static int batch_plan(dto_batch *b) {
  batch_free_plan(b);
  int rc = plan_shape(b->jobs, b->shape.nt, b->lds_limit, b->shape);
  if (rc) return rc;
  std::vector<BatchJob> hj(b->jobs.size());
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    hj[i].J = b->jobs[i]->dev;
    hj[i].lo = b->shape.lo[i];
    hj[i].out = b->d_out + b->out_off[i];
    hj[i].n_perm = (uint32_t)b->n_perm[i];
    hj[i].job_id = (uint32_t)i;
  }
  /* largest jobs first (cost ~ cells x tasks) so the queue drains evenly */
  std::vector<size_t> ord(b->jobs.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    double cx = (double)b->jobs[x]->dev.T1 * b->jobs[x]->dev.T2, cy = (double)b->jobs[y]->dev.T1 * b->jobs[y]->dev.T2;
    return cx > cy;
  });
  std::vector<WorkItem> items;
  for (size_t i : ord) {
    uint64_t tasks = b->n_perm[i] + 1;
    for (uint64_t t0 = 0; t0 < tasks; t0 += b->chunk) {
      WorkItem w;
      w.job = (uint32_t)i;
      w.t0 = (uint32_t)t0;
      w.count = (uint32_t)std::min<uint64_t>(b->chunk, tasks - t0);
      w.pad = 0;
      items.push_back(w);
    }
  }
  b->n_items = (uint32_t)items.size();
  HIP_TRY(hipMalloc(&b->d_jobs, hj.size() * sizeof(BatchJob)));
  HIP_TRY(hipMalloc(&b->d_items, std::max<size_t>(1, items.size()) * sizeof(WorkItem)));
  HIP_TRY(hipMemcpy(b->d_jobs, hj.data(), hj.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
  if (!items.empty()) HIP_TRY(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  b->grid = grid_for(b->cus, b->shape.lds, b->shape.nt, b->n_items);
  if (!b->shape.pl) HIP_TRY(hipMalloc(&b->d_ws, (size_t)b->grid * b->shape.max_n2 * sizeof(uint16_t)));
  if (!b->shape.nbl) HIP_TRY(hipMalloc(&b->d_gnb, (size_t)b->grid * b->shape.max_n * sizeof(uint16_t)));
  return DTO_OK;
}

extern "C" {

int dto_batch_create(dto_job *const *jobs, const uint64_t *n_perm, uint64_t n_jobs, uint64_t seed, dto_batch **out) {
  if (!out || (n_jobs && (!jobs || !n_perm))) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  *out = nullptr;
  dto_batch *b = new dto_batch();
  int rc = current_device(&b->device);
  DevProps pr;
  if (!rc) rc = device_props(b->device, &pr);
  if (rc) {
    delete b;
    return rc;
  }
  b->cus = pr.cus;
  b->lds_limit = pr.lds_per_block;
  b->seed = seed;
  for (uint64_t i = 0; i < n_jobs; ++i) {
    if (!jobs[i] || n_perm[i] >= 0xFFFFFFFFull) {
      delete b;
      return fail(DTO_ERR_INVALID_ARGUMENT, "bad job or permutation count");
    }
    if ((rc = ensure_resident(jobs[i]))) {
      delete b;
      return rc;
    }
    b->jobs.push_back(jobs[i]);
    b->n_perm.push_back(n_perm[i]);
    b->out_off.push_back(b->total_tasks);
    b->total_tasks += n_perm[i] + 1;
  }
  if (hipMalloc(&b->d_out, std::max<size_t>(1, b->total_tasks) * sizeof(dto_record)) != hipSuccess ||
      hipMalloc(&b->d_counter, 64) != hipSuccess || hipMalloc(&b->d_err, 64) != hipSuccess) {
    dto_batch_destroy(b);
    return fail(DTO_ERR_DEVICE, "hipMalloc failed for batch buffers");
  }
  b->shape.nt = 512;
  if ((rc = batch_plan(b))) {
    dto_batch_destroy(b);
    return rc;
  }
  *out = b;
  return DTO_OK;
}

int dto_batch_launch(dto_batch *b, void *stream) {
  if (!b) return fail(DTO_ERR_INVALID_ARGUMENT, "null batch");
  if (b->n_items == 0) return DTO_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!st) {
    if (!b->own_stream) HIP_TRY(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking));
    st = b->own_stream;
  }
  if (!b->exec) {
    /* capture queue reset + persistent kernel once, replay afterwards */
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_batch(b->shape, b->grid, st, b->d_jobs, b->d_items, b->n_items, b->seed, b->d_counter,
                                b->d_err, batch_globals(b));
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(DTO_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(b->exec, st));
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  return DTO_OK;
}

int dto_batch_results(dto_batch *b, dto_result *results) {
  if (!b || !results) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  HIP_TRY(hipDeviceSynchronize());
  int herr = 0;
  HIP_TRY(hipMemcpy(&herr, b->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr) return fail(DTO_ERR_BAD_PERMUTATION, "permutation index out of range in batch");
  std::vector<dto_record> rec(b->total_tasks);
  HIP_TRY(hipMemcpy(rec.data(), b->d_out, b->total_tasks * sizeof(dto_record), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < b->jobs.size(); ++i) {
    const dto_record *r = rec.data() + b->out_off[i];
    int rc = dto_empirical_pvalue(&r[0], r + 1, b->n_perm[i], &results[i]);
    if (rc) return rc;
  }
  return DTO_OK;
}

/* Times each candidate launch shape (threads per workgroup x tasks per work
 * item) on this batch and keeps the fastest. */
int dto_batch_autotune(dto_batch *b, int *waves) {
  if (!b) return fail(DTO_ERR_INVALID_ARGUMENT, "null batch");
  const int nts[2] = {256, 512};
  const uint32_t chunks[3] = {1, 4, 16};
  float best = 1e30f;
  int best_nt = b->shape.nt;
  uint32_t best_chunk = b->chunk;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  for (int nt : nts) {
    for (uint32_t c : chunks) {
      b->shape.nt = nt;
      b->chunk = c;
      if (batch_plan(b)) continue;
      float ms = 0.f;
      bool good = true;
      for (int rep = 0; rep < 2 && good; ++rep) {
        HIP_TRY(hipEventRecord(e0, 0));
        if (launch_batch(b->shape, b->grid, 0, b->d_jobs, b->d_items, b->n_items, b->seed, b->d_counter, b->d_err,
                         batch_globals(b)) != hipSuccess)
          good = false;
        HIP_TRY(hipEventRecord(e1, 0));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      }
      if (good && ms < best) {
        best = ms;
        best_nt = nt;
        best_chunk = c;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->shape.nt = best_nt;
  b->chunk = best_chunk;
  int rc = batch_plan(b);
  if (waves) *waves = best_nt / 64;
  return rc;
}

void dto_batch_destroy(dto_batch *b) {
  if (!b) return;
  batch_free_plan(b);
  if (b->d_out) (void)hipFree(b->d_out);
  if (b->d_counter) (void)hipFree(b->d_counter);
  if (b->d_err) (void)hipFree(b->d_err);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
}

}  // extern "C"

/* ------------------------------------------------------------------ */
/* server                                                                */
/* ------------------------------------------------------------------ */
namespace {

constexpr int kSlots = 4;         /* slices in flight, one stream each */
constexpr uint32_t kSliceChunk = 4; /* tasks per work item in a slice */
constexpr size_t kPoolAlign = 256;

/* First-fit allocator over the server's one device block: admission,
 * completion and slice staging take and return regions without any device
 * allocation (hipMalloc / hipFree would synchronise the device). */
struct Pool {
  unsigned char *base = nullptr;
  size_t size = 0, used = 0;
  std::map<size_t, size_t> free_; /* offset -> bytes, coalesced */

  void init(void *b, size_t n) {
    base = (unsigned char *)b;
    size = n;
    used = 0;
    free_.clear();
    if (n) free_[0] = n;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (std::max<size_t>(bytes, 1) + kPoolAlign - 1) & ~(kPoolAlign - 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used += bytes;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + kPoolAlign - 1) & ~(kPoolAlign - 1);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

/* Launch class of a job on its own: screen table in LDS, pairs in LDS,
 * column constants in registers, and workgroups per CU its LDS allows.
 * Requests of one class share slices; a large request never moves small
 * ones onto its slower shape or its lower occupancy. */
struct ShapeClass {
  bool hl, nbl, pl, rc;
  uint32_t per_cu;
  bool operator==(const ShapeClass &o) const {
    return hl == o.hl && nbl == o.nbl && pl == o.pl && rc == o.rc && per_cu == o.per_cu;
  }
};

ShapeClass class_of(const dto_job *j) {
  const LdsLayout &L = j->lo_screen;
  ShapeClass c;
  c.hl = L.h_in_lds != 0;
  c.nbl = L.nb_in_lds != 0;
  c.pl = L.pairs_in_lds != 0;
  c.rc = (j->dev.T2w + 63) / 64 <= kSegWords;
  c.per_cu = std::min<uint32_t>(4, std::max<uint32_t>(1, (160u * 1024u) / std::max<uint32_t>(L.total, 1)));
  return c;
}

constexpr size_t kNoRegion = ~(size_t)0;

struct Request {
  uint64_t id = 0;
  dto_job *job = nullptr;
  uint64_t n_perm = 0;
  uint64_t next_task = 0; /* next task to schedule (0 = unpermuted) */
  uint64_t done_tasks = 0;
  uint32_t inflight = 0; /* slices in flight that hold some of its tasks */
  bool failed = false;
  bool admitted = false;
  double t_submit = 0.0;
  ShapeClass cls{};
  dto_record *d_out = nullptr;
  size_t out_off = kNoRegion, out_bytes = 0;
  size_t job_off = kNoRegion, job_bytes = 0; /* the job's tables in the pool (deferred jobs) */
};

struct Part {
  Request *req;
  uint32_t t0, count;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  BatchJob *d_jobs = nullptr;
  WorkItem *d_items = nullptr;
  BatchJob *h_jobs = nullptr; /* pinned staging */
  WorkItem *h_items = nullptr;
  int *h_err = nullptr; /* pinned: the launch's error word, copied back on the stream */
  uint32_t *d_counter = nullptr;
  int *d_err = nullptr;
  size_t ws_off = kNoRegion, ws_bytes = 0;
  size_t gnb_off = kNoRegion, gnb_bytes = 0;
  std::vector<Part> parts;
  bool inflight = false;
};

}  // namespace

struct dto_server {
  dto_server_config cfg{};
  int device = -1;
  int cus = 0;
  size_t lds_limit = 0;
  uint64_t next_id = 1;
  std::deque<Request *> queue;
  std::vector<Request *> active;
  Pool pool;
  void *pool_mem = nullptr;
  hipStream_t up = nullptr; /* uploads of admitted jobs */
  hipEvent_t up_ev = nullptr;
  bool up_any = false;
  Slot slots[kSlots];
  uint32_t cap_items = 0;
  std::vector<dto_job *> graveyard; /* own-arena jobs, freed when no slice is in flight */
  uint32_t max_queue = 1u << 20;
};

static void server_free_request(dto_server *sv, Request *r) {
  if (r->out_off != kNoRegion) sv->pool.release(r->out_off, r->out_bytes);
  if (r->job_off != kNoRegion) sv->pool.release(r->job_off, r->job_bytes);
  if (r->job) {
    if (r->job->arena || r->job->ws || r->job->gnb)
      sv->graveyard.push_back(r->job); /* its hipFree would wait for the device */
    else
      dto_job_destroy(r->job);
  }
  delete r;
}

static bool server_busy(const dto_server *sv) {
  for (const Slot &s : sv->slots)
    if (s.inflight) return true;
  return false;
}

/* Admits queued requests, oldest first, while the pool holds their tables
 * and records; copies a deferred job's tables into its region. */
static int server_admit(dto_server *sv, std::string *dropped) {
  bool any = false;
  while (!sv->queue.empty() && sv->active.size() < sv->cfg.max_active_jobs) {
    Request *r = sv->queue.front();
    dto_job *j = r->job;
    const size_t out_b = (r->n_perm + 1) * sizeof(dto_record);
    const size_t job_b = j->resident ? 0 : j->image.size();
    const size_t need = ((out_b + kPoolAlign - 1) & ~(kPoolAlign - 1)) + ((job_b + kPoolAlign - 1) & ~(kPoolAlign - 1));
    if (need > sv->pool.size) {
      /* can never fit: drop it rather than block the queue */
      sv->queue.pop_front();
      *dropped += (dropped->empty() ? "" : ", ") + std::to_string(r->id);
      server_free_request(sv, r);
      continue;
    }
    size_t oo, jo = kNoRegion;
    if (!sv->pool.alloc(out_b, &oo)) break;
    if (job_b && !sv->pool.alloc(job_b, &jo)) {
      sv->pool.release(oo, out_b);
      break;
    }
    r->out_off = oo;
    r->out_bytes = out_b;
    r->d_out = (dto_record *)(sv->pool.base + oo);
    if (job_b) {
      r->job_off = jo;
      r->job_bytes = job_b;
      HIP_TRY(hipMemcpyAsync(sv->pool.base + jo, j->image.data(), job_b, hipMemcpyHostToDevice, sv->up));
      bind_job(j, sv->pool.base + jo);
      any = true;
    }
    r->admitted = true;
    sv->active.push_back(r);
    sv->queue.pop_front();
  }
  if (any) {
    HIP_TRY(hipEventRecord(sv->up_ev, sv->up));
    sv->up_any = true;
  }
  return DTO_OK;
}

/* Builds and launches one slice on slot s: every active request of the
 * class of the oldest request with unscheduled tasks gives up to
 * perms_per_slice of its tasks. */
static int server_launch_slice(dto_server *sv, Slot &s) {
  s.parts.clear();
  const Request *oldest = nullptr;
  for (Request *r : sv->active)
    if (!r->failed && r->next_task < r->n_perm + 1 && (!oldest || r->t_submit < oldest->t_submit)) oldest = r;
  if (!oldest) return DTO_OK;
  const ShapeClass cls = oldest->cls;
  uint32_t n_items = 0, lds = 0, max_n2 = 1, max_n = 1;
  for (Request *r : sv->active) {
    if (r->failed || !(r->cls == cls) || r->next_task >= r->n_perm + 1) continue;
    const uint32_t c = (uint32_t)std::min<uint64_t>(sv->cfg.perms_per_slice, r->n_perm + 1 - r->next_task);
    const uint32_t its = (c + kSliceChunk - 1) / kSliceChunk;
    if (n_items + its > sv->cap_items || s.parts.size() >= sv->cfg.max_active_jobs) break;
    s.parts.push_back(Part{r, (uint32_t)r->next_task, c});
    n_items += its;
    lds = std::max(lds, r->job->lo_screen.total);
    max_n2 = std::max(max_n2, r->job->dev.N2);
    max_n = std::max(max_n, std::max(r->job->dev.N1, r->job->dev.N2));
  }
  if (s.parts.empty()) return DTO_OK;
  const uint32_t grid = grid_for(sv->cus, lds, kNT, n_items);
  if (!cls.pl) {
    s.ws_bytes = (size_t)grid * max_n2 * sizeof(uint16_t);
    if (!sv->pool.alloc(s.ws_bytes, &s.ws_off)) {
      s.parts.clear(); /* no room for the staging now: try again later */
      s.ws_off = kNoRegion;
      return DTO_OK;
    }
  }
  if (!cls.nbl) {
    s.gnb_bytes = (size_t)grid * max_n * sizeof(uint16_t);
    if (!sv->pool.alloc(s.gnb_bytes, &s.gnb_off)) {
      s.parts.clear();
      s.gnb_off = kNoRegion;
      if (s.ws_off != kNoRegion) sv->pool.release(s.ws_off, s.ws_bytes);
      s.ws_off = kNoRegion;
      return DTO_OK;
    }
  }
  uint32_t it = 0;
  for (size_t i = 0; i < s.parts.size(); ++i) {
    Part &p = s.parts[i];
    BatchJob &bj = s.h_jobs[i];
    bj.J = p.req->job->dev;
    bj.lo = p.req->job->lo_screen;
    bj.out = p.req->d_out;
    bj.n_perm = (uint32_t)p.req->n_perm;
    bj.job_id = (uint32_t)(p.req->id & 0xFFFFFFFFu);
    for (uint32_t t0 = p.t0; t0 < p.t0 + p.count; t0 += kSliceChunk) {
      WorkItem &w = s.h_items[it++];
      w.job = (uint32_t)i;
      w.t0 = t0;
      w.count = std::min(kSliceChunk, p.t0 + p.count - t0);
      w.pad = 0;
    }
    p.req->next_task += p.count;
    p.req->inflight++;
  }
  /* jobs admitted since the last slice: their tables must have arrived */
  if (sv->up_any) HIP_TRY(hipStreamWaitEvent(s.st, sv->up_ev, 0));
  HIP_TRY(hipMemcpyAsync(s.d_jobs, s.h_jobs, s.parts.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s.st));
  HIP_TRY(hipMemcpyAsync(s.d_items, s.h_items, n_items * sizeof(WorkItem), hipMemcpyHostToDevice, s.st));
  Shape shape;
  shape.nt = kNT;
  shape.hl = cls.hl;
  shape.pl = cls.pl;
  shape.nbl = cls.nbl;
  shape.rc = cls.rc;
  shape.lds = lds;
  shape.max_n2 = max_n2;
  shape.max_n = max_n;
  Globals g;
  g.ws = cls.pl ? nullptr : (uint16_t *)(sv->pool.base + s.ws_off);
  g.ws_stride = max_n2;
  g.gnb = cls.nbl ? nullptr : (uint16_t *)(sv->pool.base + s.gnb_off);
  g.gnb_stride = max_n;
  hipError_t e = launch_batch(shape, grid, s.st, s.d_jobs, s.d_items, n_items, sv->cfg.seed, s.d_counter, s.d_err, g);
  if (e != hipSuccess) return fail(DTO_ERR_DEVICE, std::string("slice launch: ") + hipGetErrorString(e));
  HIP_TRY(hipMemcpyAsync(s.h_err, s.d_err, sizeof(int), hipMemcpyDeviceToHost, s.st));
  HIP_TRY(hipEventRecord(s.ev, s.st));
  s.inflight = true;
  return DTO_OK;
}

extern "C" {

int dto_server_create(const dto_server_config *cfg, dto_server **out) {
  if (!cfg || !out) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  *out = nullptr;
  dto_server *sv = new dto_server();
  sv->cfg = *cfg;
  if (!sv->cfg.max_active_jobs) sv->cfg.max_active_jobs = 256;
  if (!sv->cfg.perms_per_slice) sv->cfg.perms_per_slice = 1024;
  if (!sv->cfg.hbm_budget_bytes) sv->cfg.hbm_budget_bytes = 8ull << 30;
  DevProps pr;
  int rc = current_device(&sv->device);
  if (!rc) rc = device_props(sv->device, &pr);
  if (rc) {
    delete sv;
    return rc;
  }
  sv->cus = pr.cus;
  sv->lds_limit = pr.lds_per_block;
  sv->cap_items = sv->cfg.max_active_jobs * ((sv->cfg.perms_per_slice + kSliceChunk - 1) / kSliceChunk);
  bool ok = hipMalloc(&sv->pool_mem, sv->cfg.hbm_budget_bytes) == hipSuccess &&
            hipStreamCreateWithFlags(&sv->up, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&sv->up_ev, hipEventDisableTiming) == hipSuccess;
  for (Slot &s : sv->slots) {
    ok = ok && hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) == hipSuccess &&
         hipMalloc(&s.d_counter, 64) == hipSuccess && hipMalloc(&s.d_err, 64) == hipSuccess &&
         hipMalloc(&s.d_jobs, (size_t)sv->cfg.max_active_jobs * sizeof(BatchJob)) == hipSuccess &&
         hipMalloc(&s.d_items, (size_t)sv->cap_items * sizeof(WorkItem)) == hipSuccess &&
         hipHostMalloc(&s.h_jobs, (size_t)sv->cfg.max_active_jobs * sizeof(BatchJob), hipHostMallocDefault) ==
             hipSuccess &&
         hipHostMalloc(&s.h_items, (size_t)sv->cap_items * sizeof(WorkItem), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&s.h_err, 64, hipHostMallocDefault) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    dto_server_destroy(sv);
    return fail(DTO_ERR_DEVICE, "server setup failed (device pool of " + std::to_string(cfg->hbm_budget_bytes) +
                                    " B, streams, slice buffers)");
  }
  sv->pool.init(sv->pool_mem, sv->cfg.hbm_budget_bytes);
  *out = sv;
  return DTO_OK;
}

int dto_server_submit(dto_server *sv, dto_job *job, uint64_t n_perm, uint64_t *request_id) {
  if (!sv || !job || !request_id) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  if (sv->queue.size() >= sv->max_queue) return fail(DTO_ERR_BUSY, "request queue full");
  if (n_perm >= 0xFFFFFFFFull) return fail(DTO_ERR_TOO_LARGE, "too many permutations in one request");
  if (job->resident && !job->arena) return fail(DTO_ERR_INVALID_ARGUMENT, "the job already belongs to a server");
  if (job->device != sv->device) return fail(DTO_ERR_INVALID_ARGUMENT, "the job was built for another device");
  Request *r = new Request();
  r->id = sv->next_id++;
  r->job = job;
  r->n_perm = n_perm;
  r->t_submit = now_s();
  r->cls = class_of(job);
  sv->queue.push_back(r);
  *request_id = r->id;
  return DTO_OK;
}

int dto_server_pending(const dto_server *sv, uint64_t *queued, uint64_t *active) {
  if (!sv) return fail(DTO_ERR_INVALID_ARGUMENT, "null server");
  if (queued) *queued = sv->queue.size();
  if (active) *active = sv->active.size();
  return DTO_OK;
}

int dto_server_poll(dto_server *sv, uint32_t timeout_ms, uint64_t *ids, dto_result *results, double *latency_s,
                    uint32_t cap, uint32_t *n_done) {
  if (!sv || !n_done || (cap && !results)) return fail(DTO_ERR_INVALID_ARGUMENT, "null pointer");
  *n_done = 0;
  const double t_end = now_s() + timeout_ms * 1e-3;
  std::string failed, dropped;
  std::vector<dto_record> rec;
  for (;;) {
    /* retire finished slices; a slice that reports an error fails only the
     * requests it carried */
    for (Slot &s : sv->slots) {
      if (!s.inflight) continue;
      hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) return fail(DTO_ERR_DEVICE, std::string("slice: ") + hipGetErrorString(q));
      const int herr = *(volatile int *)s.h_err;
      for (Part &p : s.parts) {
        p.req->inflight--;
        if (herr)
          p.req->failed = true;
        else
          p.req->done_tasks += p.count;
      }
      s.parts.clear();
      if (s.ws_off != kNoRegion) sv->pool.release(s.ws_off, s.ws_bytes);
      if (s.gnb_off != kNoRegion) sv->pool.release(s.gnb_off, s.gnb_bytes);
      s.ws_off = s.gnb_off = kNoRegion;
      s.inflight = false; /* the next launch on this slot clears its error word */
    }
    /* complete (or drop failed) requests */
    for (size_t i = 0; i < sv->active.size() && *n_done < cap;) {
      Request *r = sv->active[i];
      if (r->inflight) {
        ++i;
        continue;
      }
      if (r->failed) {
        failed += (failed.empty() ? "" : ", ") + std::to_string(r->id);
      } else if (r->done_tasks == r->n_perm + 1) {
        rec.resize(r->n_perm + 1);
        HIP_TRY(hipMemcpy(rec.data(), r->d_out, rec.size() * sizeof(dto_record), hipMemcpyDeviceToHost));
        int rc = dto_empirical_pvalue(&rec[0], rec.data() + 1, r->n_perm, &results[*n_done]);
        if (rc) return rc;
        if (ids) ids[*n_done] = r->id;
        if (latency_s) latency_s[*n_done] = now_s() - r->t_submit;
        ++*n_done;
      } else {
        ++i;
        continue;
      }
      sv->active.erase(sv->active.begin() + (long)i);
      server_free_request(sv, r);
    }
    if (!server_busy(sv) && !sv->graveyard.empty()) {
      for (dto_job *j : sv->graveyard) dto_job_destroy(j);
      sv->graveyard.clear();
    }
    int rc = server_admit(sv, &dropped);
    if (rc) return rc;
    for (Slot &s : sv->slots)
      if (!s.inflight && (rc = server_launch_slice(sv, s))) return rc;
    if (!failed.empty() || !dropped.empty()) {
      std::string msg;
      if (!failed.empty()) msg += "a device error (out-of-range permutation index or no grid cell) failed requests " + failed;
      if (!dropped.empty()) msg += (msg.empty() ? "" : "; ") + std::string("requests larger than the HBM budget were dropped: ") + dropped;
      return fail(failed.empty() ? DTO_ERR_TOO_LARGE : DTO_ERR_DEVICE, msg);
    }
    if (*n_done > 0 || now_s() >= t_end) return DTO_OK;
    if (!server_busy(sv) && sv->queue.empty() && sv->active.empty()) return DTO_OK;
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

void dto_server_destroy(dto_server *sv) {
  if (!sv) return;
  (void)hipDeviceSynchronize();
  for (Slot &s : sv->slots) {
    if (s.st) (void)hipStreamDestroy(s.st);
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.d_jobs) (void)hipFree(s.d_jobs);
    if (s.d_items) (void)hipFree(s.d_items);
    if (s.h_jobs) (void)hipHostFree(s.h_jobs);
    if (s.h_items) (void)hipHostFree(s.h_items);
    if (s.h_err) (void)hipHostFree(s.h_err);
    if (s.d_counter) (void)hipFree(s.d_counter);
    if (s.d_err) (void)hipFree(s.d_err);
  }
  for (Request *r : sv->active) server_free_request(sv, r);
  for (Request *r : sv->queue) server_free_request(sv, r);
  for (dto_job *j : sv->graveyard) dto_job_destroy(j);
  if (sv->up) (void)hipStreamDestroy(sv->up);
  if (sv->up_ev) (void)hipEventDestroy(sv->up_ev);
  if (sv->pool_mem) (void)hipFree(sv->pool_mem);
  delete sv;
}

}  // extern "C"

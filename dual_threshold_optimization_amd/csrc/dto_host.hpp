/*
 * Host-side job state (no HIP types): the upload step of the DTO library.
 */
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dto.h"
#include "dto_device.h"

namespace dto {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

/* statrs 0.17.1 factorial::ln_factorial (host C library log). */
double ln_factorial(uint64_t x);
double ln_gamma(double x);
/* generate_thresholds (src/collections/ranked.rs:359-375). */
std::vector<uint32_t> generate_thresholds(uint32_t max_rank);

struct RankedList {
  std::vector<std::string> ids; /* rank-sorted (stable) */
  std::vector<uint32_t> ranks;
};

/* Host state of a job.  derive() fills the tables from the lists and
 * thresholds; the device copy lives in DeviceJob (dto_api.hip). */
struct HostJob {
  RankedList l1, l2;
  std::vector<uint32_t> thr1, thr2;
  uint64_t N = 0;
  uint64_t dropped1 = 0, dropped2 = 0;

  std::vector<uint16_t> bin1, bin2;
  std::vector<int32_t> join;
  std::vector<uint32_t> join_off, join_pos;
  bool dup = false;
  bool kle = false; /* every overlap k_ij <= K_i: no list1 entry pairs with two list2 entries */
  std::vector<uint32_t> K, n;
  std::vector<double> LF, D;
  std::vector<float> H32, A32, B32;
  float margin = 0.f;
  double h_max = 0.0;

  int derive(); /* DTO_OK or an error code (message set) */
};

/* RankedFeatureList::from + compute_population_size (+ optional background
 * drop) into a HostJob with reference thresholds. */
int build_host_job(const char *const *ids1, const uint32_t *ranks1, uint64_t n1, const char *const *ids2,
                   const uint32_t *ranks2, uint64_t n2, const char *const *bg, uint64_t nbg, uint32_t flags,
                   HostJob &out);

/* Lists + thresholds with a caller-given population size (no checks). */
int build_host_job_with_population(const char *const *ids1, const uint32_t *ranks1, uint64_t n1,
                                   const char *const *ids2, const uint32_t *ranks2, uint64_t n2, uint64_t N,
                                   HostJob &out);

/* Readers with the reference's semantics. */
int read_ranked_csv(const char *path, std::vector<std::string> &ids, std::vector<uint32_t> &ranks);
int read_feature_lines(const char *path, std::vector<std::string> &ids);

/* serde_json / ryu formatting of one f64 (src/main.rs:165 output). */
std::string format_f64(double v);
std::string result_json(const dto_result &r);

}  // namespace dto

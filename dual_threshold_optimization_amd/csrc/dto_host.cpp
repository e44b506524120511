/*
 * Upload step of the DTO library (host C++, no device calls): ranked lists,
 * thresholds, population size, the list1/list2 join, threshold bins, set
 * sizes and the ln n! tables the kernels read; the reference's file readers
 * and its JSON output format.
 *
 * Compiled with -ffp-contract=off: every floating-point expression here must
 * round exactly as the reference's (Rust never fuses multiply-add).
 */
#include "dto_host.hpp"
#include "dto_text.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <charconv>
#include <fstream>
#include <numeric>
#include <sstream>
#include <unordered_map>
#include <unordered_set>

namespace dto {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
const char *last_error_cstr() { return g_err.c_str(); }

/* ------------------------------------------------------------------ */
/* statrs 0.17.1 ln_factorial                                           */
/* ------------------------------------------------------------------ */
/* Lanczos g = 10.900511 with the 11-term coefficient set statrs ships
 * (gamma.rs); valid for x >= 0.5, the only branch ln_factorial reaches. */
static const double kGammaR = 10.900511;
static const double kGammaDk[11] = {
    2.48574089138753565546e-5,  1.05142378581721974210,    -3.45687097222016235469,
    4.51227709466894823700,     -2.98285225323576655721,   1.05639711577126713077,
    -1.95428773191645869583e-1, 1.70970543404441224307e-2, -5.71926117404305781283e-4,
    4.63399473359905636708e-6,  -2.71994908488607703910e-9};
static const double kLn2SqrtEOverPi = 0.6207822376352452223455184457816472122518527279025978;

double ln_gamma(double x) {
  /* x >= 0.5 branch: s = d0 + sum d_i / (x + i - 1);
   * ln s + ln(2 sqrt(e/pi)) + (x - 0.5) ln((x - 0.5 + g) / e) */
  double s = kGammaDk[0];
  for (int i = 1; i <= 10; ++i) s = s + kGammaDk[i] / ((x + (double)i) - 1.0);
  return (log(s) + kLn2SqrtEOverPi) + (x - 0.5) * log(((x - 0.5) + kGammaR) / M_E);
}

static const double *factorial_cache() {
  static double fc[171];
  static bool init = [] {
    fc[0] = 1.0;
    for (int i = 1; i <= 170; ++i) fc[i] = fc[i - 1] * (double)i;
    return true;
  }();
  (void)init;
  return fc;
}

double ln_factorial(uint64_t x) {
  if (x <= 170) return log(factorial_cache()[x]);
  return ln_gamma((double)x + 1.0);
}

std::vector<uint32_t> generate_thresholds(uint32_t max_rank) {
  /* T1 = 1, T_{n+1} = floor(T_n * 1.01 + 1.0) while <= max rank.  The
   * reference's last-threshold fix-up (ranked.rs:369-372) writes to the vector
   * it is about to replace, so the last threshold stays below the max rank. */
  std::vector<uint32_t> t;
  uint64_t cur = 1;
  while (cur <= max_rank) {
    t.push_back((uint32_t)cur);
    double v = floor((double)cur * 1.01 + 1.0);
    if (v >= 4294967295.0) break;
    cur = (uint64_t)v;
  }
  return t;
}

/* ------------------------------------------------------------------ */
/* job construction                                                      */
/* ------------------------------------------------------------------ */
static void stable_sort_list(RankedList &l) {
  std::vector<size_t> ord(l.ids.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return l.ranks[a] < l.ranks[b]; });
  RankedList s;
  s.ids.reserve(ord.size());
  s.ranks.reserve(ord.size());
  for (size_t o : ord) {
    s.ids.push_back(std::move(l.ids[o]));
    s.ranks.push_back(l.ranks[o]);
  }
  l = std::move(s);
}

static std::string debug_list(const std::vector<std::string> &v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) s += ", ";
    s += "\"" + v[i] + "\"";
  }
  return s + "]";
}

int build_host_job(const char *const *ids1, const uint32_t *ranks1, uint64_t n1, const char *const *ids2,
                   const uint32_t *ranks2, uint64_t n2, const char *const *bg, uint64_t nbg, uint32_t flags,
                   HostJob &J) {
  if ((n1 && (!ids1 || !ranks1)) || (n2 && (!ids2 || !ranks2)) || (nbg && !bg))
    return fail(DTO_ERR_INVALID_ARGUMENT, "null list pointer");
  if (flags & ~(DTO_DROP_OUTSIDE_BACKGROUND | DTO_DEFER_UPLOAD)) return fail(DTO_ERR_INVALID_ARGUMENT, "unknown flag");
  RankedList a, b;
  for (uint64_t i = 0; i < n1; ++i) {
    if (!ids1[i]) return fail(DTO_ERR_INVALID_ARGUMENT, "null feature id");
    a.ids.emplace_back(ids1[i]);
    a.ranks.push_back(ranks1[i]);
  }
  for (uint64_t i = 0; i < n2; ++i) {
    if (!ids2[i]) return fail(DTO_ERR_INVALID_ARGUMENT, "null feature id");
    b.ids.emplace_back(ids2[i]);
    b.ranks.push_back(ranks2[i]);
  }
  J.dropped1 = J.dropped2 = 0;
  if (bg) {
    /* compute_population_size with a background (compute_population_size.rs:87-102) */
    std::unordered_set<std::string> bgs;
    for (uint64_t i = 0; i < nbg; ++i) bgs.insert(bg[i] ? bg[i] : "");
    RankedList *lists[2] = {&a, &b};
    const char *names[2] = {"first", "second"};
    for (int L = 0; L < 2; ++L) {
      RankedList &l = *lists[L];
      std::vector<std::string> missing;
      RankedList kept;
      for (size_t i = 0; i < l.ids.size(); ++i) {
        if (bgs.count(l.ids[i])) {
          kept.ids.push_back(l.ids[i]);
          kept.ranks.push_back(l.ranks[i]);
        } else {
          missing.push_back(l.ids[i]);
        }
      }
      if (!missing.empty()) {
        if (!(flags & DTO_DROP_OUTSIDE_BACKGROUND))
          return fail(DTO_ERR_POPULATION, std::string("The following genes in the ") + names[L] +
                                              " ranked feature list are not in the background: " +
                                              debug_list(missing));
        (L == 0 ? J.dropped1 : J.dropped2) = missing.size();
        l = std::move(kept);
      }
    }
    J.N = nbg;
  } else {
    /* no background (compute_population_size.rs:72-86): N = list1 entries whose
     * id is in both lists; it must equal both lengths */
    std::unordered_set<std::string> s2(b.ids.begin(), b.ids.end());
    uint64_t inter = 0;
    for (auto &id : a.ids) inter += s2.count(id);
    if (inter != a.ids.size() || inter != b.ids.size())
      return fail(DTO_ERR_POPULATION, "If no background is provided, the feature lists must have identical genes.");
    J.N = inter;
  }
  stable_sort_list(a);
  stable_sort_list(b);
  J.l1 = std::move(a);
  J.l2 = std::move(b);
  J.thr1 = generate_thresholds(J.l1.ranks.empty() ? 0 : J.l1.ranks.back());
  J.thr2 = generate_thresholds(J.l2.ranks.empty() ? 0 : J.l2.ranks.back());
  return J.derive();
}

int build_host_job_with_population(const char *const *ids1, const uint32_t *ranks1, uint64_t n1,
                                   const char *const *ids2, const uint32_t *ranks2, uint64_t n2, uint64_t N,
                                   HostJob &J) {
  if ((n1 && (!ids1 || !ranks1)) || (n2 && (!ids2 || !ranks2)))
    return fail(DTO_ERR_INVALID_ARGUMENT, "null list pointer");
  RankedList a, b;
  for (uint64_t i = 0; i < n1; ++i) {
    if (!ids1[i]) return fail(DTO_ERR_INVALID_ARGUMENT, "null feature id");
    a.ids.emplace_back(ids1[i]);
    a.ranks.push_back(ranks1[i]);
  }
  for (uint64_t i = 0; i < n2; ++i) {
    if (!ids2[i]) return fail(DTO_ERR_INVALID_ARGUMENT, "null feature id");
    b.ids.emplace_back(ids2[i]);
    b.ranks.push_back(ranks2[i]);
  }
  J.dropped1 = J.dropped2 = 0;
  J.N = N;
  stable_sort_list(a);
  stable_sort_list(b);
  J.l1 = std::move(a);
  J.l2 = std::move(b);
  J.thr1 = generate_thresholds(J.l1.ranks.empty() ? 0 : J.l1.ranks.back());
  J.thr2 = generate_thresholds(J.l2.ranks.empty() ? 0 : J.l2.ranks.back());
  return J.derive();
}

static void bins_of(const std::vector<uint32_t> &ranks, const std::vector<uint32_t> &thr, std::vector<uint16_t> &bin,
                    std::vector<uint32_t> &setsz) {
  bin.resize(ranks.size());
  setsz.assign(thr.size(), 0);
  for (size_t p = 0; p < ranks.size(); ++p) {
    size_t t = std::lower_bound(thr.begin(), thr.end(), ranks[p]) - thr.begin();
    bin[p] = (uint16_t)t; /* == thr.size(): above the last threshold, in no set */
    if (t < thr.size()) setsz[t]++;
  }
  for (size_t t = 1; t < setsz.size(); ++t) setsz[t] += setsz[t - 1];
}

int HostJob::derive() {
  const size_t T1 = thr1.size(), T2 = thr2.size();
  if (T1 == 0 || T2 == 0)
    return fail(DTO_ERR_EMPTY_GRID, "a ranked list has no thresholds; optimize() has no result to return");
  if (T1 > DTO_MAX_THRESHOLDS || T2 > DTO_MAX_THRESHOLDS)
    return fail(DTO_ERR_TOO_LARGE, "more than DTO_MAX_THRESHOLDS thresholds");
  if (l1.ids.size() > DTO_MAX_LIST_LEN || l2.ids.size() > DTO_MAX_LIST_LEN)
    return fail(DTO_ERR_TOO_LARGE, "list longer than DTO_MAX_LIST_LEN");
  if (N >= (1ull << 31)) return fail(DTO_ERR_TOO_LARGE, "population size too large");
  for (size_t t = 1; t < T1; ++t)
    if (thr1[t] <= thr1[t - 1]) return fail(DTO_ERR_INVALID_ARGUMENT, "thresholds must increase");
  for (size_t t = 1; t < T2; ++t)
    if (thr2[t] <= thr2[t - 1]) return fail(DTO_ERR_INVALID_ARGUMENT, "thresholds must increase");
  bins_of(l1.ranks, thr1, bin1, K);
  bins_of(l2.ranks, thr2, bin2, n);
  /* Hypergeometric::new(N, |S1|, |S2|) must accept every cell; set sizes
   * grow with the threshold index, so the last ones decide */
  if (K.back() > N || n.back() > N)
    return fail(DTO_ERR_HYPERGEOMETRIC, "Failed to create hypergeometric distribution");

  /* join: list2 entry -> list1 entries with the same id */
  std::unordered_map<std::string, std::vector<uint32_t>> pos1;
  pos1.reserve(l1.ids.size() * 2);
  for (size_t m = 0; m < l1.ids.size(); ++m) pos1[l1.ids[m]].push_back((uint32_t)m);
  const size_t N2 = l2.ids.size();
  join.assign(N2, -1);
  join_off.assign(N2 + 1, 0);
  join_pos.clear();
  dup = false;
  for (size_t q = 0; q < N2; ++q) {
    join_off[q] = (uint32_t)join_pos.size();
    auto it = pos1.find(l2.ids[q]);
    if (it != pos1.end()) {
      join[q] = (int32_t)it->second[0];
      if (it->second.size() > 1) dup = true;
      for (uint32_t m : it->second) join_pos.push_back(m);
    }
  }
  join_off[N2] = (uint32_t)join_pos.size();
  /* with one partner per list2 entry and no list1 entry shared by two list2
   * entries, the pairs with a <= i are distinct list1 entries: k_ij <= K_i */
  kle = !dup;
  if (kle) {
    std::vector<uint8_t> used(l1.ids.size(), 0);
    for (size_t q = 0; q < N2 && kle; ++q)
      if (join[q] >= 0) {
        if (used[(size_t)join[q]]) kle = false;
        used[(size_t)join[q]] = 1;
      }
  }

  /* ln n! table, exactly as statrs evaluates each call */
  LF.resize(N + 1);
  for (uint64_t x = 0; x <= N; ++x) LF[x] = ln_factorial(x);
  D.resize(T2);
  for (size_t j = 0; j < T2; ++j) D[j] = (LF[N] - LF[n[j]]) - LF[N - n[j]];

  /* Screen table H(x) = ln x! - x (ln N - 1).  The log of the first tail
   * term is a signed sum of nine ln-factorials whose arguments cancel
   * (2N - 2N), so it equals the same signed sum of H; H stays small
   * (|H| <= ~N/e) where ln x! reaches N ln N, which keeps FP32 accurate. */
  const double c = N > 1 ? log((double)N) - 1.0 : 0.0;
  std::vector<double> H(N + 1);
  h_max = 0.0;
  for (uint64_t x = 0; x <= N; ++x) {
    H[x] = LF[x] - (double)x * c;
    h_max = std::max(h_max, fabs(H[x]));
  }
  H32.resize(N + 1);
  for (uint64_t x = 0; x <= N; ++x) H32[x] = (float)H[x];
  A32.resize(T1);
  B32.resize(T2);
  for (size_t i = 0; i < T1; ++i) A32[i] = (float)(H[K[i]] + H[N - K[i]]);
  for (size_t j = 0; j < T2; ++j) B32[j] = (float)((H[n[j]] + H[N - n[j]]) - H[N]);
  /* FP32 error of x = (A + B) - (H[k] + H[K-k] + H[n-k] + H[N-K-n+k]):
   * 4 table roundings, 2 roundings of A/B (|.| <= 3M), 5 additions
   * (|partials| <= 10M): <= 38 * 2^-24 * M; plus slack for the FP64 rounding
   * of the reference's own exponent (~1e-9). */
  margin = (float)(40.0 * ldexp(1.0, -24) * h_max + 1e-3);
  return DTO_OK;
}

/* ------------------------------------------------------------------ */
/* readers                                                               */
/* ------------------------------------------------------------------ */
/* BufRead::lines() hands out each line as a String and fails on a line that
 * is not valid UTF-8; the readers' .expect("Could not read line") panics
 * with the io::Error's Debug form. */
static const char *kBadUtf8 =
    "Could not read line: Error { kind: InvalidData, message: \"stream did not contain valid UTF-8\" }";

/* BufRead::lines(): split on '\n', drop one trailing '\r', no final empty line. */
static bool read_lines(const char *path, std::vector<std::string> &lines) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  std::string all = ss.str();
  size_t p = 0;
  while (p < all.size()) {
    size_t e = all.find('\n', p);
    std::string line = all.substr(p, e == std::string::npos ? std::string::npos : e - p);
    if (e != std::string::npos && !line.empty() && line.back() == '\r') line.pop_back();
    lines.push_back(line);
    if (e == std::string::npos) break;
    p = e + 1;
  }
  return true;
}

int read_ranked_csv(const char *path, std::vector<std::string> &ids, std::vector<uint32_t> &ranks) {
  std::vector<std::string> lines;
  if (!path || !read_lines(path, lines)) return fail(DTO_ERR_IO, std::string("Could not open file: ") + (path ? path : ""));
  for (size_t i = 0; i < lines.size(); ++i) {
    const std::string &line = lines[i];
    if (!utf8_valid(line)) return fail(DTO_ERR_PARSE, kBadUtf8);
    size_t c = line.find(',');
    if (c == std::string::npos || line.find(',', c + 1) != std::string::npos)
      return fail(DTO_ERR_PARSE, std::string("Invalid format in file ") + path + " at line " + std::to_string(i + 1));
    std::string feat = rust_trim(line.substr(0, c));
    std::string rs = rust_trim(line.substr(c + 1));
    /* usize::from_str: optional '+', then one or more decimal digits, no overflow */
    size_t s0 = (!rs.empty() && rs[0] == '+') ? 1 : 0;
    if (s0 == rs.size()) return fail(DTO_ERR_PARSE, "Invalid rank value");
    uint64_t v = 0;
    for (size_t t = s0; t < rs.size(); ++t) {
      if (rs[t] < '0' || rs[t] > '9') return fail(DTO_ERR_PARSE, "Invalid rank value");
      uint64_t d = (uint64_t)(rs[t] - '0');
      if (v > (UINT64_MAX - d) / 10) return fail(DTO_ERR_PARSE, "Invalid rank value");
      v = v * 10 + d;
    }
    ids.push_back(feat);
    ranks.push_back((uint32_t)v); /* `rank as u32` */
  }
  return DTO_OK;
}

int read_feature_lines(const char *path, std::vector<std::string> &ids) {
  std::vector<std::string> lines;
  if (!path || !read_lines(path, lines)) return fail(DTO_ERR_IO, std::string("Could not open file: ") + (path ? path : ""));
  for (auto &l : lines) {
    if (!utf8_valid(l)) return fail(DTO_ERR_PARSE, kBadUtf8);
    ids.push_back(rust_trim(l));
  }
  return DTO_OK;
}

/* ------------------------------------------------------------------ */
/* JSON                                                                   */
/* ------------------------------------------------------------------ */
/* serde_json writes finite f64 with ryu's shortest round-trip digits:
 * integer-valued -> "d.0"; 1e-5 <= |v| < 1e16 plain decimal; otherwise
 * d.ddde[-]X. */
std::string format_f64(double v) {
  if (!std::isfinite(v)) return "null";
  std::string out;
  if (std::signbit(v)) {
    out = "-";
    v = -v;
  }
  if (v == 0.0) return out + "0.0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  size_t e = s.find('e');
  std::string mant = s.substr(0, e);
  int exp10 = atoi(s.c_str() + e + 1);
  std::string digits;
  for (char ch : mant)
    if (ch != '.') digits += ch;
  const int length = (int)digits.size();
  const int kk = exp10 + 1; /* 10^(kk-1) <= v < 10^kk */
  const int k = kk - length;
  if (0 <= k && kk <= 16) {
    out += digits + std::string((size_t)(kk - length), '0') + ".0";
  } else if (0 < kk && kk <= 16) {
    out += digits.substr(0, (size_t)kk) + "." + digits.substr((size_t)kk);
  } else if (-5 < kk && kk <= 0) {
    out += "0." + std::string((size_t)(-kk), '0') + digits;
  } else if (length == 1) {
    out += digits + "e" + std::to_string(kk - 1);
  } else {
    out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
  }
  return out;
}

/* empirical_pvalue's json! object printed by serde_json::to_string_pretty
 * (keys sorted, two-space indent). */
std::string result_json(const dto_result &r) {
  const dto_record &u = r.unpermuted;
  std::string s = "{\n";
  s += "  \"empirical_pvalue\": " + format_f64(r.empirical_pvalue) + ",\n";
  s += "  \"fdr\": " + format_f64(r.fdr) + ",\n";
  s += "  \"population_size\": " + std::to_string(u.population_size) + ",\n";
  s += "  \"rank1\": " + std::to_string(u.rank1) + ",\n";
  s += "  \"rank2\": " + std::to_string(u.rank2) + ",\n";
  s += "  \"set1_len\": " + std::to_string(u.set1_len) + ",\n";
  s += "  \"set2_len\": " + std::to_string(u.set2_len) + ",\n";
  s += "  \"unpermuted_intersection_size\": " + std::to_string(u.intersection_size) + ",\n";
  s += "  \"unpermuted_pvalue\": " + format_f64(u.pvalue) + "\n";
  s += "}";
  return s;
}

}  // namespace dto

/* Test hooks (not in dto.h): the readers alone, without building a job.
 * ids are written NUL-terminated one after another into ids_buf. */
static int pack_ids(const std::vector<std::string> &ids, char *buf, uint64_t cap, uint64_t *used) {
  uint64_t need = 0;
  for (auto &s : ids) need += s.size() + 1;
  if (used) *used = need;
  if (!buf) return DTO_OK;
  if (need > cap) return dto::fail(DTO_ERR_INVALID_ARGUMENT, "id buffer too small");
  for (auto &s : ids) {
    memcpy(buf, s.c_str(), s.size() + 1);
    buf += s.size() + 1;
  }
  return DTO_OK;
}

extern "C" int dto_internal_read_ranked_csv(const char *path, char *ids_buf, uint64_t ids_cap, uint64_t *ids_used,
                                            uint32_t *ranks, uint64_t cap, uint64_t *n) {
  std::vector<std::string> ids;
  std::vector<uint32_t> r;
  int rc = dto::read_ranked_csv(path, ids, r);
  if (rc) return rc;
  if (n) *n = ids.size();
  if (ranks)
    for (uint64_t i = 0; i < cap && i < r.size(); ++i) ranks[i] = r[i];
  return pack_ids(ids, ids_buf, ids_cap, ids_used);
}

extern "C" int dto_internal_read_feature_lines(const char *path, char *ids_buf, uint64_t ids_cap, uint64_t *ids_used,
                                               uint64_t *n) {
  std::vector<std::string> ids;
  int rc = dto::read_feature_lines(path, ids);
  if (rc) return rc;
  if (n) *n = ids.size();
  return pack_ids(ids, ids_buf, ids_cap, ids_used);
}

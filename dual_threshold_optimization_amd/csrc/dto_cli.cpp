/*
 * dto_amd: drop-in for the reference CLI (src/main.rs:20-166).
 *
 *   dto_amd -1 <FILE> -2 <FILE> [-b <FILE>] [-p <N>] [-t <N>] [-m]
 *
 * Same flags, defaults (-p 1000, -t 1), stderr messages and pretty JSON on
 * stdout.  The permutation loop runs on the GPU; -t is accepted for
 * compatibility (threads == 0 warns and becomes 1, as in main.rs:73-76).
 * Extensions: --seed <u64> (Philox key, default: time-based like the
 * reference's unseeded thread_rng), --mode screened|plain,
 * --drop-outside-background.  Errors exit with status 101, as a Rust panic does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/dto.h"
#include "dto_text.h"

static void usage(FILE *f) {
  fprintf(f,
          "Dual Threshold Optimization CLI (MI355X)\n\n"
          "Usage: dto_amd [OPTIONS] --ranked-list1 <FILE> --ranked-list2 <FILE>\n\n"
          "Options:\n"
          "  -1, --ranked-list1 <FILE>   First ranked feature list (CSV feature,rank; no header)\n"
          "  -2, --ranked-list2 <FILE>   Second ranked feature list (CSV feature,rank; no header)\n"
          "  -b, --background <FILE>     Background feature list (one feature per line, optional)\n"
          "  -p, --permutations <N>      Number of permutations to perform [default: 1000]\n"
          "  -t, --threads <N>           Number of threads [default: 1]\n"
          "  -m, --multi-node            Enable multi-node mode (MPI)\n"
          "      --seed <U64>            Philox seed for the permutations\n"
          "      --mode <screened|plain> p-value grid evaluation [default: screened]\n"
          "      --drop-outside-background  Drop list features absent from the background\n"
          "  -h, --help                  Print help\n"
          "  -V, --version               Print version\n");
}

[[noreturn]] static void panic(const std::string &msg) {
  fprintf(stderr, "thread 'main' panicked:\n%s\n", msg.c_str());
  exit(101);
}

static bool parse_u64(const std::string &s, uint64_t *v) {
  if (s.empty()) return false;
  size_t i = s[0] == '+' ? 1 : 0;
  if (i == s.size()) return false;
  uint64_t x = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    uint64_t d = (uint64_t)(s[i] - '0');
    if (x > (UINT64_MAX - d) / 10) return false;
    x = x * 10 + d;
  }
  *v = x;
  return true;
}

/* Largest rank of a CSV the library has already parsed: last field of each
 * line, the same `usize` parse and `as u32` truncation. */
static uint32_t max_rank(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return 0;
  uint32_t mx = 0;
  std::string line;
  int ch;
  auto take = [&]() {
    size_t comma = line.find(',');
    if (comma != std::string::npos) {
      std::string r = line.substr(comma + 1);
      if (!r.empty() && r.back() == '\r') r.pop_back();
      uint64_t v = 0;
      if (dto::utf8_valid(r) && parse_u64(dto::rust_trim(r), &v) && (uint32_t)v > mx) mx = (uint32_t)v;
    }
    line.clear();
  };
  while ((ch = fgetc(f)) != EOF) {
    if (ch == '\n')
      take();
    else
      line.push_back((char)ch);
  }
  if (!line.empty()) take();
  fclose(f);
  return mx;
}

/* optimize's stderr lines for a task whose minimum p is shared
 * (src/dto/optimize_main.rs:86-108); {:.15} is fixed notation with 15
 * decimals, as %.15f prints it. */
static void tie_warnings(const dto_record &r) {
  if (DTO_TIES_AT_MIN_P(&r) <= 1) return;
  fprintf(stderr,
          "Multiple results with the same minimum p-value (%.15f). Choosing the result with the largest intersection "
          "size.\n",
          r.pvalue);
  if (DTO_TIES_AT_MAX_K(&r) > 1)
    fprintf(stderr,
            "Multiple results with the same maximum intersection size (%llu). Choosing an arbitrary result based on "
            "the order of thresholds.\n",
            (unsigned long long)r.intersection_size);
}

int main(int argc, char **argv) {
  std::string l1, l2, bg;
  bool have_bg = false, multi = false, drop = false, have_seed = false;
  uint64_t perms = 1000, threads = 1, seed = 0;
  uint32_t mode = DTO_MODE_SCREENED;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i], val;
    auto need = [&](const char *name) -> std::string {
      size_t eq = a.find('=');
      if (a.rfind("--", 0) == 0 && eq != std::string::npos) return a.substr(eq + 1);
      if (a.size() > 2 && a[0] == '-' && a[1] != '-') return a.substr(2);
      if (i + 1 >= argc) {
        fprintf(stderr, "error: a value is required for '%s' but none was supplied\n", name);
        exit(2);
      }
      return argv[++i];
    };
    auto is = [&](const char *s, const char *l) {
      if (s && (a == s || (a.rfind(s, 0) == 0 && a.size() > 2 && a[1] != '-'))) return true;
      return a == l || a.rfind(std::string(l) + "=", 0) == 0;
    };
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return 0;
    } else if (a == "-V" || a == "--version") {
      printf("dto_amd 2.0.1\n");
      return 0;
    } else if (a == "-m" || a == "--multi-node") {
      multi = true;
    } else if (a == "--drop-outside-background") {
      drop = true;
    } else if (is("-1", "--ranked-list1")) {
      l1 = need("--ranked-list1 <FILE>");
    } else if (is("-2", "--ranked-list2")) {
      l2 = need("--ranked-list2 <FILE>");
    } else if (is("-b", "--background")) {
      bg = need("--background <FILE>");
      have_bg = true;
    } else if (is("-p", "--permutations")) {
      if (!parse_u64(need("--permutations <PERMUTATIONS>"), &perms)) {
        fprintf(stderr, "error: invalid value for '--permutations <PERMUTATIONS>'\n");
        return 2;
      }
    } else if (is("-t", "--threads")) {
      if (!parse_u64(need("--threads <THREADS>"), &threads)) {
        fprintf(stderr, "error: invalid value for '--threads <THREADS>'\n");
        return 2;
      }
    } else if (is(nullptr, "--seed")) {
      if (!parse_u64(need("--seed <SEED>"), &seed)) {
        fprintf(stderr, "error: invalid value for '--seed <SEED>'\n");
        return 2;
      }
      have_seed = true;
    } else if (is(nullptr, "--mode")) {
      std::string m = need("--mode <MODE>");
      if (m == "screened")
        mode = DTO_MODE_SCREENED;
      else if (m == "plain")
        mode = DTO_MODE_PLAIN;
      else {
        fprintf(stderr, "error: invalid value '%s' for '--mode <MODE>'\n", m.c_str());
        return 2;
      }
    } else {
      fprintf(stderr, "error: unexpected argument '%s' found\n\n", a.c_str());
      usage(stderr);
      return 2;
    }
  }
  if (l1.empty() || l2.empty()) {
    fprintf(stderr, "error: the following required arguments were not provided:\n");
    if (l1.empty()) fprintf(stderr, "  --ranked-list1 <FILE>\n");
    if (l2.empty()) fprintf(stderr, "  --ranked-list2 <FILE>\n");
    return 2;
  }
  if (threads == 0) {
    fprintf(stderr, "Warning: Number of threads cannot be 0. Setting threads to 1.\n");
    threads = 1;
  }
  fprintf(stderr, "Ranked list 1: %s\n", l1.c_str());
  fprintf(stderr, "Ranked list 2: %s\n", l2.c_str());
  fprintf(stderr, "Permutations: %llu\n", (unsigned long long)perms);
  fprintf(stderr, "Threads: %llu\n", (unsigned long long)threads);
  fprintf(stderr, "Multi-node mode: %s\n", multi ? "enabled" : "disabled");

  dto_job *job = nullptr;
  int rc = dto_job_create_from_files(l1.c_str(), l2.c_str(), have_bg ? bg.c_str() : nullptr,
                                     drop ? DTO_DROP_OUTSIDE_BACKGROUND : 0u, &job);
  /* main.rs prints the threshold product after reading both lists and
   * before the background and population checks: errors past the readers
   * come after that line */
  const bool lists_read = rc == DTO_OK || (rc != DTO_ERR_IO && rc != DTO_ERR_PARSE);
  if (lists_read) {
    uint64_t t1 = 0, t2 = 0;
    if (rc == DTO_OK) {
      dto_job_info info;
      dto_job_info_get(job, &info);
      t1 = info.n_thresholds1;
      t2 = info.n_thresholds2;
    } else {
      t1 = dto_generate_thresholds(max_rank(l1), nullptr, 0);
      t2 = dto_generate_thresholds(max_rank(l2), nullptr, 0);
    }
    fprintf(stderr,
            "The product of the lengths of the threshold lists (this describes the asymptotic runtime of a single "
            "job): %llu\n",
            (unsigned long long)(t1 * t2));
  }
  if (rc) panic(dto_last_error());
  if (multi) panic("Multi-node mode requires MPI. Rebuild with the 'mpi' feature enabled.");
  if (!have_seed) seed = (uint64_t)time(nullptr) ^ ((uint64_t)clock() << 32);
  dto_result res;
  std::vector<dto_record> perm_rec(perms ? perms : 1);
  rc = dto_run(job, perms, seed, nullptr, nullptr, mode, &res, perm_rec.data());
  if (rc) panic(dto_last_error());
  /* every task's optimize warns about its own ties (optimize_main.rs:86-108);
   * the reference's threads print them in completion order, here task order */
  tie_warnings(res.unpermuted);
  for (uint64_t p = 0; p < perms; ++p) tie_warnings(perm_rec[p]);
  char buf[4096];
  if (dto_result_json(&res, buf, sizeof buf) < 0) panic("json formatting failed");
  printf("%s\n", buf);
  dto_job_destroy(job);
  return 0;
}

// Random-input check of the host library's text handling, built by
// tests/test_host_sanitizers.py with g++ -fsanitize=address,undefined
// together with cudalucas_amd/csrc/ll_host.cpp (no HIP code): work lines,
// worktodo files (next assignment, clear), report formats and the file lock.
// Usage: fuzz_host_text <work dir> <lines> <files>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>

#include "ll_api.h"

// The length policy in ll_host.cpp asks the engine which lengths have
// kernels; this check builds the host file alone.
extern "C" int ll_fft_supported(uint32_t) { return 1; }

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const std::string dir = argv[1];
  const long n_lines = atol(argv[2]), n_files = atol(argv[3]);
  std::mt19937_64 rng(12345);
  const char *alphabet = "TestDoubleCheck=,/#\\ 0123456789abcdefABCDEFKkMmN/A\r\t%xqcQdTCrRpPfFiozyaNhmsD";
  const size_t na = strlen(alphabet);
  ll_assignment a;
  char out[64];
  for (long it = 0; it < n_lines; it++) {
    std::string s;
    const int len = (int)(rng() % 260);
    if (rng() % 3 == 0) s = (rng() % 2) ? "Test=" : "DoubleCheck=";
    for (int i = 0; i < len; i++) s += alphabet[rng() % na];
    ll_work_parse_line(s.c_str(), &a);
    ll_report r;
    memset(&r, 0, sizeof r);
    r.p = (uint32_t)rng();
    r.n = (uint32_t)rng();
    r.iteration = (uint32_t)rng();
    r.max_err = (float)(rng() % 1000) / 1000.0f;
    r.interval_us = rng() % 3 ? rng() : 0;
    r.interval_iterations = (uint32_t)(rng() % 5);
    r.eta_s = rng();
    r.now = (int64_t)(rng() % 4000000000ull);
    r.program = "x";
    ll_format_report(s.c_str(), &r, out, rng() % 2 ? sizeof out : (size_t)(rng() % 4));
  }
  const std::string work = dir + "/worktodo.txt";
  for (long it = 0; it < n_files; it++) {
    FILE *f = fopen(work.c_str(), "w");
    if (!f) return 3;
    const int lines = (int)(rng() % 12);
    for (int l = 0; l < lines; l++) {
      std::string s = (rng() % 2) ? "Test=N/A,44497" : "";
      const int len = (int)(rng() % 300);
      for (int i = 0; i < len; i++) s += alphabet[rng() % na];
      fprintf(f, "%s\n", s.c_str());
    }
    fclose(f);
    ll_work_next(work.c_str(), &a);
    ll_work_clear(work.c_str(), 44497);
  }
  const std::string results = dir + "/results.txt";
  FILE *g = ll_fopen_locked(results.c_str(), "a", 10);
  if (!g) return 4;
  FILE *h = ll_fopen_locked(results.c_str(), "a", 10);  // held: times out
  if (h) return 5;
  ll_fclose_locked(g);
  printf("fuzz done\n");
  return 0;
}

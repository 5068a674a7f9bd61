// Random-input check of the host library's text handling, built by
// tests/test_host_sanitizers.py with g++ -fsanitize=address,undefined
// together with cudalucas_amd/csrc/ll_host.cpp (no HIP code): work lines,
// worktodo files (next assignment, clear), report formats, the file lock,
// and the digit codec and checkpoint files on random small states.
// Usage: fuzz_host_text <work dir> <lines> <files>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

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
  // Digit codec and checkpoint files on random states (p, n small), each
  // round trip checked: unpack -> balance -> standardize -> pack gives the
  // words back, and a written checkpoint reads back as the same words.
  for (long it = 0; it < n_files; it++) {
    const uint32_t n = 64u << (rng() % 5);
    const uint32_t p = n * (2 + (uint32_t)(rng() % 18)) + 1 + (uint32_t)(rng() % (n - 1));
    const size_t nw = (p + 31) / 32;
    std::vector<uint32_t> w(nw), back(nw);
    for (auto &x : w) x = (uint32_t)rng();
    w[nw - 1] &= (p % 32) ? ((1u << (p % 32)) - 1) : 0xffffffffu;
    if (w[nw - 1] == ((p % 32) ? ((1u << (p % 32)) - 1) : 0xffffffffu)) w[nw - 1] = 0;  // not 2^p - 1
    std::vector<int32_t> d(n);
    ll_unpack_bits(p, n, w.data(), d.data());
    ll_balance_digits(p, n, d.data());
    const uint32_t offset = (uint32_t)(rng() % p);
    ll_res64(p, n, d.data(), offset);
    ll_checkpoint meta = {p, n, 1000, offset, rng() % 100000000, 0, 0};
    std::vector<uint32_t> ck(ll_checkpoint_words(p)), rd(ll_checkpoint_words(p));
    if (ll_checkpoint_encode(&meta, d.data(), ck.data()) != LL_OK) return 6;
    if (ll_checkpoint_write(dir.c_str(), p, ck.data(), nullptr, 0) != LL_OK) return 7;
    if (ll_checkpoint_read(dir.c_str(), p, rd.data()) != 0 || rd != ck) return 8;
    ll_checkpoint dm;
    if (ll_checkpoint_decode(p, rd.data(), &dm) != LL_OK || dm.offset != offset) return 9;
    std::vector<int32_t> d2(n);
    ll_checkpoint_digits(p, n, rd.data(), d2.data());
    ll_standardize_digits(p, n, d.data());
    ll_pack_bits(p, n, d.data(), back.data());
    if (back != w) return 10;
    ll_checkpoint_remove(dir.c_str(), p);
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

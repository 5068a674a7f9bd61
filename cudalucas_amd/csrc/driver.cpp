// CUDALucasAMD: the CUDALucas command line and ini surface on top of the
// MI355X LL engine (include/ll_api.h).
//
// Mirrors tdulcet/CUDALucas v2.06: parse_args (CUDALucas.cu:3581-3818),
// process_options and its ini keys (CUDALucas.cu:3388-3522, parse.c:683-770),
// check() with its report/checkpoint cadence and roundoff/device-error
// recovery (CUDALucas.cu:2862-3186), round_off_test (CUDALucas.cu:2589-2648),
// the -r self-test (CUDALucas.cu:2650-2805) and the worktodo/results workflow
// (parse.c:241-678).  Device work goes through the C-ABI only.
#include <math.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "ll_api.h"

static const char kProgram[] = "CUDALucasAMD v1.0";

struct Config {
  int device = -1;
  int threads[2] = {0, 0};
  std::string ini = "CUDALucas.ini";
  int cpi = -1, ri = -1, ei = -1, el = -1, er = -1;
  int fftlen = -1;  // words; 0 = automatic
  int save_all = -1;
  std::string save_folder;
  int selftest = -1;
  int roundoff = -1;
  std::string workfile, resultsfile;
  std::string output_string, output_header;  // OutputString, OutputHeader
  int output_interval = -1;                  // OutputHInterval; 0 = the one-line report
  // options without a reference counterpart (tests and tooling)
  std::string ckpt_dir;
  int64_t fixed_offset = -1;
  uint64_t inject_after = 0;
  float inject_value = 0.f;
  int inject = 0;  // injections left (re-armed after each restart)
  int64_t stop_after = -1;  // iterations before a checkpoint-and-exit
  uint32_t force_width = 0, force_height = 0, force_lines = 0;
  bool force_fft = false;  // take -f as given, without the lb/ub clamp
  bool info = false;       // -info
  int cufftbench[3] = {0, 0, 0};   // -cufftbench start end passes
  int threadbench[4] = {0, 0, 0, 0};  // -threadbench start end passes mode
  std::string selector_model, selector_times;  // learned FFT-length selection
};

// The learned length selector (--selector), when loaded.
static ll_selector *g_sel = nullptr;

// A length-selection step: the learned selector when one is loaded, else
// the reference policy's list index (choose_fft_length, CUDALucas.cu:1027).
static int next_length(uint32_t q, int &idx, int n_failed) {
  if (g_sel) {
    const int n = ll_selector_choose(g_sel, q, (uint32_t)n_failed, nullptr);
    if (n > 0) return n;
  }
  idx++;
  return ll_choose_fft_length(q, &idx);
}

// The length to try after a roundoff failure at n: the next one larger than
// n, or 0 if there is none (the policy's upper bound, CUDALucas.cu:3047,
// where the list index stops growing).  idx moves only when there is one.
static int larger_length(uint32_t q, int &idx, int n) {
  int i = idx;
  const int next = next_length(q, i, n);
  if (next <= n) return 0;
  idx = i;
  return next;
}

static volatile sig_atomic_t g_quit = 0;

static void on_signal(int sig) {
  g_quit = 1;
  const char *m = sig == SIGINT ? "\tSIGINT caught, writing checkpoint.\n" : "\tSIGTERM caught, writing checkpoint.\n";
  ssize_t r = write(2, m, strlen(m));
  (void)r;
}

static double now_s() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

static void print_time(unsigned sec, int mode) {
  unsigned day = 0;
  if (mode >= 2) {
    day = sec / 86400;
    sec %= 86400;
  }
  const unsigned h = sec / 3600, m = (sec % 3600) / 60, s = sec % 60;
  if (day) printf("%u:%02u:%02u:%02u", day, h, m, s);
  else if (h) printf("%u:%02u:%02u", h, m, s);
  else printf("%u:%02u", m, s);
}

// check_interval (CUDALucas.cu:3374-3386): k * 10^m with k in {1, 2, 5}
static int check_interval(int v) {
  int k = v, j = 1;
  while (j <= k) j *= 10;
  j /= 10;
  k = (int)floor(k / (double)j + 0.5);
  if (k == 3) k = 2;
  else if (k > 3 && k < 8) k = 5;
  else if (k > 7) k = 10;
  return k * j;
}

// ------------------------------------------------------------------- ini
// IniGetInt / IniGetInt2 / IniGetStr (parse.c:683-771): the value of the
// first line "<key>=<value>" that `accept` takes (a number, two numbers, a
// non-empty string); a line whose value it refuses is passed over for a
// later one, and with none the caller's default holds.  The newline and one
// carriage return before it are not part of the value.
static bool ini_get(const std::string &file, const char *key, std::string &out,
                    bool (*accept)(const std::string &)) {
  FILE *f = fopen(file.c_str(), "r");
  if (!f) return false;
  char buf[1024];
  const size_t kl = strlen(key);
  bool found = false;
  while (!found && fgets(buf, sizeof buf, f)) {
    if (!strncmp(buf, key, kl) && buf[kl] == '=') {
      std::string v(buf + kl + 1);
      if (!v.empty() && v.back() == '\n') v.pop_back();
      if (!v.empty() && v.back() == '\r') v.pop_back();
      if (accept(v)) {
        out = v;
        found = true;
      }
    }
  }
  fclose(f);
  return found;
}

static bool ini_is_int(const std::string &v) {
  int x;
  return sscanf(v.c_str(), "%d", &x) == 1;
}
static bool ini_is_int_pair(const std::string &v) {
  int x, y;
  return sscanf(v.c_str(), "%d %d", &x, &y) == 2;
}
static bool ini_is_text(const std::string &v) { return !v.empty(); }

static void ini_int(const std::string &file, const char *key, int &dst, int dflt) {
  if (dst >= 0) return;  // command line wins
  std::string v;
  dst = dflt;
  if (ini_get(file, key, v, ini_is_int)) sscanf(v.c_str(), "%d", &dst);
}

static void ini_str(const std::string &file, const char *key, std::string &dst, const char *dflt) {
  if (!dst.empty()) return;
  std::string v;
  dst = ini_get(file, key, v, ini_is_text) ? v : std::string(dflt);
}

// process_options (CUDALucas.cu:3388-3522)
static void process_options(Config &c) {
  ini_int(c.ini, "SaveAllCheckpoints", c.save_all, 0);
  ini_int(c.ini, "RoundoffTest", c.roundoff, 0);
  ini_int(c.ini, "DeviceNumber", c.device, 0);
  ini_int(c.ini, "CheckpointIterations", c.cpi, 100000);
  ini_int(c.ini, "ReportIterations", c.ri, 10000);
  ini_int(c.ini, "ErrorIterations", c.ei, 100);
  ini_int(c.ini, "ErrorLimit", c.el, 40);
  ini_int(c.ini, "ErrorReset", c.er, 85);
  if (c.threads[0] <= 0) {
    std::string v;
    if (ini_get(c.ini, "Threads", v, ini_is_int_pair)) sscanf(v.c_str(), "%d %d", &c.threads[0], &c.threads[1]);
  }
  if (c.fftlen < 0) {
    std::string v;
    c.fftlen = ini_get(c.ini, "FFTLength", v, ini_is_text) ? ll_fft_from_str(v.c_str()) : 0;
    if (c.fftlen < 0) c.fftlen = 0;
  }
  ini_str(c.ini, "WorkFile", c.workfile, "worktodo.txt");
  ini_str(c.ini, "ResultsFile", c.resultsfile, "results.txt");
  ini_str(c.ini, "SaveFolder", c.save_folder, "savefiles");
  ini_str(c.ini, "OutputString", c.output_string, "0");
  ini_str(c.ini, "OutputHeader", c.output_header, "0");
  ini_int(c.ini, "OutputHInterval", c.output_interval, 0);
  if (c.output_interval < 0) c.output_interval = 15;  // CUDALucas.cu:3488
  const int k1 = check_interval(c.cpi), k2 = check_interval(c.ri), k3 = check_interval(c.ei);
  if (k1 != c.cpi) printf("CheckpointIterations = %d must have the form k*10^m for k = 1, 2, or 5.\nChanging to %d.\n", c.cpi, k1);
  if (k2 != c.ri) printf("ReportIterations = %d must have the form k*10^m for k = 1, 2, or 5.\nChanging to %d.\n", c.ri, k2);
  if (k3 != c.ei) printf("ErrorIterations = %d must have the form k*10^m for k = 1, 2, or 5.\nChanging to %d.\n", c.ei, k3);
  c.cpi = k1;
  c.ri = k2;
  c.ei = k3;
  if (c.el > 47) c.el = 47;
  if (c.el < 1) c.el = 1;
}

// ------------------------------------------------------------- engine glue
struct Engine {
  ll_ctx *ctx = nullptr;
  uint32_t p = 0, n = 0;
  ~Engine() { close(); }
  void close() {
    if (ctx) ll_destroy(ctx);
    ctx = nullptr;
  }
  int open(const Config &c, uint32_t p_, uint32_t n_) {
    close();
    p = p_;
    n = n_;
    ll_options o;
    memset(&o, 0, sizeof o);
    o.device = c.device;
    o.threads[0] = c.threads[0];
    o.threads[1] = c.threads[1];
    o.width = c.force_width;
    o.height = c.force_height;
    o.lines = c.force_lines;
    const int rc = ll_create(p, n, &o, &ctx);
    if (rc) fprintf(stderr, "ll_create(%u, %u): %s\n", p, n, ctx ? ll_last_error(ctx) : "failed");
    return rc;
  }
};

// The state a rollback returns to: packed shifted residue + position.
struct Snapshot {
  std::vector<uint32_t> words;  // ll_checkpoint layout
  ll_checkpoint meta{};
};

static void snapshot_from_digits(Snapshot &s, uint32_t p, uint32_t n, const int32_t *d, uint32_t next,
                                 uint32_t offset, uint64_t total_us) {
  s.words.assign(ll_checkpoint_words(p), 0);
  s.meta.p = p;
  s.meta.n = n;
  s.meta.next_iter = next;
  s.meta.offset = offset;
  s.meta.total_us = total_us;
  s.meta.time_adj_us = 0;
  s.meta.iter_adj = 0;
  ll_checkpoint_encode(&s.meta, d, s.words.data());
}

static int load_snapshot(Engine &e, const Snapshot &s) {
  std::vector<int32_t> d(e.n);
  ll_checkpoint_digits(e.p, e.n, s.words.data(), d.data());
  return ll_load_digits(e.ctx, d.data(), s.meta.offset);
}

// round_off_test (CUDALucas.cu:2589-2648): 1000 iterations, each checked.
static int round_off_test(Engine &e, uint32_t iters) {
  printf("Running careful round off test for %u iterations.\n"
         "If average error > 0.25, or maximum error > 0.35,\n"
         "the test will restart with a longer FFT.\n", iters);
  std::vector<float> errs(iters);
  float m = 0.f;
  if (ll_iterate(e.ctx, iters, errs.data(), &m)) return -1;
  double total = 0;
  for (uint32_t k = 0; k < iters; k++) {
    total += errs[k];
    if (errs[k] > 0.35f) {
      printf("Iteration = %u < %u && err = %5.5f > 0.35, increasing n from %uK\n", k, iters, errs[k], e.n / 1024);
      return 1;
    }
  }
  const double avg = total / iters;
  if (avg > 0.25) {
    printf("Iteration %u, average error = %5.5f > 0.25 (max error = %5.5f), increasing FFT length and restarting\n", iters, avg, m);
    return 1;
  }
  printf("Iteration %u, average error = %5.5f <= 0.25 (max error = %5.5f), continuing test.\n", iters, avg, m);
  return 0;
}

// One LL test of 2^q - 1 (check, CUDALucas.cu:2862-3186).  Returns 0 on a
// finished or cleanly interrupted test, 2 on an unrecoverable error.
// process_output (CUDALucas.cu:1633-1701): the report as a table row by
// OutputString, with OutputHeader above the first row and then every
// OutputHInterval - 1 rows (the reference's count, kept across tests).
static int g_header_countdown = 0;

static void report_table(const Config &cfg, uint32_t q, uint32_t n, uint32_t iteration, uint32_t offset,
                         uint64_t res, float maxerr, uint32_t iters, double seconds, uint64_t eta_s) {
  ll_report r;
  memset(&r, 0, sizeof r);
  r.p = q;
  r.n = n;
  r.iteration = iteration;
  r.offset = offset;
  r.residue = res;
  r.max_err = maxerr;
  r.interval_iterations = iters;
  r.interval_us = (uint64_t)(seconds * 1e6);
  r.eta_s = eta_s;
  r.now = (int64_t)time(nullptr);
  r.program = kProgram;
  std::string line(256, '\0');
  const int len = ll_format_report(cfg.output_string.c_str(), &r, &line[0], line.size());
  if (len >= (int)line.size()) {
    line.assign((size_t)len + 1, '\0');
    ll_format_report(cfg.output_string.c_str(), &r, &line[0], line.size());
  }
  line.resize(len > 0 ? (size_t)len : 0);
  if (g_header_countdown == 0 && cfg.output_interval > 0) {
    g_header_countdown = cfg.output_interval - 1;
    printf("%s\n", cfg.output_header.c_str());
  }
  g_header_countdown--;
  printf("%s\n", line.c_str());
}

static int check(uint32_t q, const Config &cfg, const char *aid) {
  const uint32_t last = q - 2;
  const double error_limit = (cfg.el - cfg.el / 8.0 * log((double)cfg.ei) / log(100.0)) / 100.0;
  std::vector<uint32_t> words(ll_checkpoint_words(q));
  Snapshot good;
  int which = ll_checkpoint_read(cfg.ckpt_dir.c_str(), q, words.data());
  uint64_t total_us = 0;
  if (which >= 0) {
    good.words = words;
    ll_checkpoint_decode(q, words.data(), &good.meta);
    total_us = good.meta.total_us;
    if (which == 1) fprintf(stderr, "\nThe checkpoint appears to be corrupt. Using the backup file.\n");
  }
  // a manual length (words) is any index >= the list size (choose_fft_length)
  int idx = cfg.fftlen > 0 ? cfg.fftlen : (which >= 0 ? (int)good.meta.n : -1);
  int n = cfg.force_fft && cfg.fftlen > 0 && ll_fft_supported((uint32_t)cfg.fftlen) ? cfg.fftlen
                                                                                  : ll_choose_fft_length(q, &idx);
  if (g_sel && cfg.fftlen <= 0) {
    if (which >= 0 && ll_fft_supported(good.meta.n)) {
      n = (int)good.meta.n;  // keep the length the checkpoint was written at
    } else if (which < 0) {
      float pred = 0.f;
      const int ns = ll_selector_choose(g_sel, q, 0, &pred);
      if (ns > 0) {
        printf("Learned length selection: %dK (predicted max roundoff %.3f; length policy %dK)\n", ns / 1024,
               pred, n / 1024);
        n = ns;
      } else {
        fprintf(stderr, "selector: %s\n", ll_selector_last_error(g_sel));
      }
    }
  }
  if (n <= 0) {
    fprintf(stderr, "No FFT length available for exponent %u\n", q);
    return 2;
  }
  const int base_idx = idx, base_n = n;
  Engine eng;
  if (eng.open(cfg, q, (uint32_t)n)) return 2;
  std::vector<int32_t> digits(n);
  uint32_t j;
  if (which < 0) {
    uint32_t offset;
    if (cfg.fixed_offset >= 0) offset = (uint32_t)(cfg.fixed_offset % q);
    else {
      srand((unsigned)time(nullptr));
      offset = (uint32_t)(rand() % q);
    }
    ll_initial_digits(q, (uint32_t)n, offset, digits.data());
    // snapshot of s_0 = 4 at "next iteration 1"
    snapshot_from_digits(good, q, (uint32_t)n, digits.data(), 1, offset, 0);
    printf("Starting M%u fft length = %dK\n", q, n / 1024);
    j = 1;
  } else {
    j = good.meta.next_iter;
    printf("\nContinuing M%u @ iteration %u with fft length %dK, %5.2f%% done\n\n", q, j, n / 1024,
           (float)j / q * 100);
  }
  if (load_snapshot(eng, good)) {
    fprintf(stderr, "%s\n", ll_last_error(eng.ctx));
    return 2;
  }
  if (cfg.roundoff > 0 && j == 1) {
    for (;;) {
      const int r = round_off_test(eng, last < 1000 ? last : 1000);
      if (r < 0) return 2;
      if (r == 0) break;
      n = larger_length(q, idx, n);
      if (n <= 0) {
        printf("No larger FFT length for M%u. Quitting.\n\n", q);
        return 2;
      }
      if (eng.open(cfg, q, (uint32_t)n) || load_snapshot(eng, good)) return 2;
    }
    // the test ran real iterations: continue from where they left off
    digits.assign(n, 0);
    uint32_t off = 0;
    ll_read_digits(eng.ctx, digits.data(), &off);
    j = (last < 1000 ? last : 1000) + 1;
    snapshot_from_digits(good, q, (uint32_t)n, digits.data(), j, off, 0);
  }
  // --inject-error ITER:VALUE:COUNT makes iteration ITER report VALUE the
  // first COUNT times it runs (re-armed after every restart)
  int inject_left = cfg.inject;
  auto arm = [&](uint32_t next) {
    if (inject_left > 0 && cfg.inject_after >= next) {
      ll_debug_inject(eng.ctx, 1, cfg.inject_after - next, cfg.inject_value);
      inject_left--;
    }
  };
  arm(j);
  bool retry = true, fft_reset = false;
  uint32_t last_chk = j - 1;
  double t_mark = now_s();
  float maxerr = 0.f;
  int64_t budget = cfg.stop_after;
  std::vector<float> errs;
  while (j <= last) {
    uint32_t stop = ((j + cfg.ri - 1) / cfg.ri) * cfg.ri;  // next report iteration
    if (stop < j) stop = j;
    if (stop > last) stop = last;
    if (budget >= 0 && (int64_t)(stop - j + 1) > budget) stop = j + (uint32_t)budget - 1;
    const uint32_t count = stop - j + 1;
    errs.resize(count);
    float m = 0.f;
    int rc = count ? ll_iterate(eng.ctx, count, errs.data(), &m) : LL_OK;
    if (rc == LL_ERR_HIP) {
      // device error (CUDALucas.cu:2938-2950): restart from the last good state
      printf("%s\nResetting device and restarting from last checkpoint.\n\n", ll_last_error(eng.ctx));
      if (eng.open(cfg, q, (uint32_t)n) || load_snapshot(eng, good)) return 2;
      j = good.meta.next_iter;
      arm(j);
      continue;
    }
    if (rc) {
      fprintf(stderr, "%s\n", ll_last_error(eng.ctx));
      return 2;
    }
    if (m > error_limit) {
      uint32_t at = j;
      for (uint32_t k = 0; k < count; k++)
        if (errs[k] > error_limit) {
          at = j + k;
          break;
        }
      printf("Round off error at iteration = %u, err = %0.5g > %0.5g, fft = %dK.\n", at, m, error_limit, n / 1024);
      {
        int probe = idx;
        if (larger_length(q, probe, n) <= 0) {  // at the upper bound (CUDALucas.cu:3047-3051)
          printf("Something is wrong! Quitting.\n\n");
          return 2;
        }
      }
      if (retry) {
        printf("Restarting from last checkpoint to see if the error is repeatable.\n\n");
        retry = false;
      } else if (fft_reset) {
        printf("The error won't go away. I give up.\n\n");
        return 2;
      } else {
        printf("The error persists.\nTrying a larger fft until the next checkpoint.\n\n");
        n = larger_length(q, idx, n);
        if (n <= 0) return 2;
        fft_reset = true;
      }
      if (eng.open(cfg, q, (uint32_t)n) || load_snapshot(eng, good)) return 2;
      j = good.meta.next_iter;
      arm(j);
      continue;
    }
    maxerr = m > maxerr ? m : maxerr;
    j = stop + 1;
    if (budget >= 0) budget -= count;
    const bool at_report = (stop % cfg.ri == 0) || stop == last;
    const bool quitting = g_quit || budget == 0;
    if (!at_report && !quitting) continue;
    uint32_t offset = 0;
    digits.assign(n, 0);
    if ((rc = ll_read_digits(eng.ctx, digits.data(), &offset))) {
      fprintf(stderr, "%s\n", ll_last_error(eng.ctx));
      return 2;
    }
    const uint64_t res = ll_res64(q, (uint32_t)n, digits.data(), offset);
    const double t = now_s();
    const double dt = t - t_mark;
    t_mark = t;
    total_us += (uint64_t)(dt * 1e6);
    snapshot_from_digits(good, q, (uint32_t)n, digits.data(), j, offset, total_us);
    if (stop % cfg.cpi == 0 || quitting) {
      if (ll_checkpoint_write(cfg.ckpt_dir.c_str(), q, good.words.data(),
                              cfg.save_all > 0 ? cfg.save_folder.c_str() : nullptr, res))
        fprintf(stderr, "Couldn't write checkpoint.\n");
    }
    if (quitting) {
      printf(" Estimated time spent so far: ");
      print_time((unsigned)(total_us / 1000000), 0);
      printf("\n\n");
      return 0;
    }
    if (stop < last) {
      const double ms = dt * 1000.0 / (stop - last_chk);
      if (cfg.output_interval) {
        report_table(cfg, q, (uint32_t)n, stop, offset, res, maxerr, stop - last_chk, dt,
                     (uint64_t)(ms * (last - stop) / 1000.0));
      } else {
        printf("Iteration %u / %u, 0x%016llx, %dK, %s, error = %5.5f, real: ", stop, q,
               (unsigned long long)res, n / 1024, kProgram, maxerr);
        print_time((unsigned)dt, 0);
        printf(", %4.4f ms/iter, ETA: ", ms);
        print_time((unsigned)(ms * (last - stop) / 1000.0), 2);
        printf(", %5.2f%%\n", (float)stop / q * 100);
      }
      fflush(stdout);
      // illegal residues (CUDALucas.cu:3006-3016)
      if (res == 0) {
        printf("Illegal residue: 0x0000000000000000. See mersenneforum.org for help.\n\n");
        return 2;
      }
      if (res == 2) {
        std::vector<uint32_t> w((q + 31) / 32);
        ll_residue_words(q, (uint32_t)n, digits.data(), offset, w.data());
        bool only2 = w[0] == 2;
        for (size_t i = 1; i < w.size() && only2; i++) only2 = w[i] == 0;
        if (only2) {
          printf("Illegal residue: 0x0000000000000002. See mersenneforum.org for help.\n\n");
          return 2;
        }
      }
      last_chk = stop;
      maxerr *= cfg.er / 100.0f;  // ErrorReset: decay the reported maximum
      if (fft_reset) {
        // a larger FFT fixed the error: back to the original length
        printf("Resetting fft.\n\n");
        fft_reset = false;
        retry = true;
        idx = base_idx;
        n = base_n;
        if (eng.open(cfg, q, (uint32_t)n) || load_snapshot(eng, good)) return 2;
      } else if (!retry) {
        printf("Looks like the error went away, continuing.\n\n");
        retry = true;
      }
    }
  }
  // final result (CUDALucas.cu:3146-3178)
  uint32_t offset = 0;
  digits.assign(n, 0);
  if (ll_read_digits(eng.ctx, digits.data(), &offset)) return 2;
  std::vector<uint32_t> w((q + 31) / 32 + 2, 0);
  ll_residue_words(q, (uint32_t)n, digits.data(), offset, w.data());
  bool zero = true;
  for (size_t i = 0; i < (q + 31) / 32; i++) zero = zero && w[i] == 0;
  const uint64_t res = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  FILE *fp = ll_fopen_locked(cfg.resultsfile.c_str(), "a", 0);  // fopen_and_lock (CUDALucas.cu:3149)
  if (zero) {
    printf("M( %u )P, n = %dK, %s", q, n / 1024, kProgram);
    if (fp) fprintf(fp, "M( %u )P, offset = %u, n = %dK, %s", q, offset, n / 1024, kProgram);
  } else {
    printf("M( %u )C, 0x%016llx, %dK, %s", q, (unsigned long long)res, n / 1024, kProgram);
    if (fp) fprintf(fp, "M( %u )C, 0x%016llx, offset = %u, n = %dK, %s", q, (unsigned long long)res, offset, n / 1024, kProgram);
  }
  total_us += (uint64_t)((now_s() - t_mark) * 1e6);
  printf(", estimated total time = ");
  print_time((unsigned)(total_us / 1000000), 0);
  printf("\n\n");
  if (fp) {
    if (aid && *aid && strncasecmp(aid, "N/A", 3)) fprintf(fp, ", AID: %s\n", aid);
    else fprintf(fp, "\n");
    ll_fclose_locked(fp);
  } else {
    fprintf(stderr, "Cannot write results to %s\n\n", cfg.resultsfile.c_str());
    return 1;
  }
  if (!zero) ll_checkpoint_remove(cfg.ckpt_dir.c_str(), q);
  return 0;
}

// check_residue (CUDALucas.cu:2650-2805): Res64 after 10,000 iterations
// against the reference's table.
struct SelfTest {
  uint32_t p;
  uint64_t res64;
};
static const SelfTest k_short[] = {
    {86243, 0x23992ccd735a03d9ull},   {132049, 0x4c52a92b54635f9eull},  {216091, 0x30247786758b8792ull},
    {756839, 0x5d2cbe7cb24a109aull},  {859433, 0x3c4ad525c2d0aed0ull},  {1257787, 0x3f45bf9bea7213eaull},
    {1398269, 0xa4a6d2f0e34629dbull}, {2976221, 0x2a7111b7f70fea2full}, {3021377, 0x6387a70a85d46bafull},
    {6972593, 0x88f1d2640adb89e1ull}, {13466917, 0x9fdc1f4092b15d69ull}, {20996011, 0x5fc58920a821da11ull},
    {24036583, 0xcbdef38a0bdc4f00ull}, {25964951, 0x62eb3ff0a5f6237cull}, {30402457, 0x0b8600ef47e69d27ull},
    {32582657, 0x02751b7fcec76bb1ull}, {37156667, 0x67ad7646a1fad514ull}, {42643801, 0x8f90d78d5007bba7ull},
    {43112609, 0xe86891ebf6cd70c4ull}, {57885161, 0x76c27556683cd84dull}};
static const SelfTest k_long[] = {
    {22133, 0xb3ed67a5d4be2a9full},    {43633, 0xa3f3e669e061bfbeull},    {85933, 0x474e01b06d9698f8ull},
    {169409, 0x3d32414583ac7d9aull},   {190097, 0xf681a2a2ad0397fbull},   {210739, 0x7e3b5a8b4e4c43a0ull},
    {292921, 0xade86fd9a4ad99f1ull},   {333803, 0xe7311b6b953a609full},   {374587, 0xdb4ca1fad1c3e276ull},
    {435571, 0xb6711834a5dd0ad1ull},   {657719, 0xc56715170d6819efull},   {738083, 0x2d6f637091a3ce6dull},
    {818239, 0xe63f1cc63af264cbull},   {978041, 0xc049101d2a0f945bull},   {1137271, 0x8e006ecd6719158full},
    {1216693, 0x847bbf5d0a077cadull},  {1296011, 0xc5002ae321cc3055ull},  {1414741, 0x7c6237b4f067d8caull},
    {1612249, 0xfd4c832d1dbf746full},  {1927129, 0x155eb42707e90cd9ull},  {2240863, 0xf5e629fcc1247e4aull},
    {2397383, 0xf58df7be5d73b369ull},  {2553659, 0x0a9357c955a13ca5ull},  {2865601, 0xdac4555e41809026ull},
    {3176779, 0x1a8ed18fa285f151ull},  {3215629, 0x9e0c8eb95b15ae66ull},  {3332107, 0x7a539a9b8060b0d8ull},
    {3564823, 0x820e428ad972cd86ull},  {3797219, 0xa7223e4ecad04860ull},  {3874583, 0xf60c0a837ec80a50ull},
    {3951977, 0xb5da72b662eb3652ull},  {4415431, 0x31d58de3e1d79a85ull},  {4723783, 0x89fac94192cc3becull},
    {5031737, 0x6371a0f4aef2af32ull},  {5646379, 0xf03d3bbaa51fbe63ull},  {6259537, 0x026805e6b4f1b0feull},
    {6336103, 0x0f0718d2da4aaad3ull},  {6565633, 0x98d84929de4e6f63ull},  {7024163, 0xc030210791a14e9cull},
    {7482053, 0xf54842c06ba7e3c4ull},  {7634537, 0x24872a305faae4f8ull},  {7786967, 0x060dfca872385f9bull},
    {8700169, 0x4f95b9afd65b3c7eull},  {9914521, 0xc6ff39fd699d624cull},  {11125619, 0x1484b22a17312da0ull},
    {12333809, 0xbd8435b83f062d6aull}, {12484649, 0xd56aa903718b2353ull}, {12936919, 0x1527f970ad3020bfull},
    {13840423, 0x002a2aa0b97306b7ull}, {14742617, 0xf63addc70cde0addull}, {15043099, 0xe52f5e2045a2db30ull},
    {15343429, 0x5eda3ab1816505b9ull}, {16543493, 0x85ea99f94dc20a8bull}, {17142793, 0x94edb737270e4ca7ull},
    {17217653, 0x3ec0aa4728b97ef3ull}, {19535569, 0x708d03d1fe4c06f2ull}, {21921901, 0x38651890258c09bbull},
    {22368691, 0xcc264d5bc61bc957ull}, {24599717, 0x76e7ab92aca768abull}, {25490893, 0x135f7e63e7357efbull},
    {27271147, 0xb1a47be0366530efull}, {29640913, 0x4d51f44ffae1c913ull}, {30232693, 0xcab02c7178561218ull},
    {32597297, 0xec15163c817fcfddull}, {33778141, 0x9145db2f2080172eull}, {38492887, 0x36efe22824adb35cull},
    {43194913, 0xf13bdac634316fa2ull}, {47885689, 0x51a71b010cacfd92ull}, {48471289, 0x5f504f82abefdd1cull},
    {50227213, 0x643739d94fe6789cull}, {53735041, 0x653267115ab57688ull}, {58404433, 0x6e4b3224349080c8ull},
    {59570449, 0x8d7c6550107a2d0aull}, {64229677, 0x796701d5562dce22ull}, {66556463, 0xe418b5d4fb5c7e45ull},
    {66847171, 0xc01ffc681dbb8966ull}, {75846319, 0x626ab6b09b607ba9ull}, {85111207, 0x14bb68062d100ebdull},
    {86845813, 0x88220ac98093b65cull}, {95507747, 0xa92c24ad9b6655f1ull}, {97454309, 0xd6d4b3b14c1b73d8ull},
    {105879517, 0x3310a4b03a18f4a8ull}, {111056879, 0x59e95ddb73c59624ull}, {115080019, 0x293d0b202c9e3c63ull},
    {117377567, 0xdf57cd5d866f1065ull}, {118813021, 0x1cd6a8caef092277ull}, {126558077, 0x40b1cf42af7f99a7ull},
    {131142761, 0x046ab21b10a27189ull}, {131715607, 0x8fcfb36b868738ecull}, {142017539, 0x801794367f35ccb6ull},
    {147162241, 0x72c8773764a260b0ull}, {149447533, 0x24702427a0f5d673ull}};

static int check_residue(const Config &cfg, int longtest) {
  const SelfTest *tab = longtest ? k_long : k_short;
  const int count = longtest ? (int)(sizeof k_long / sizeof k_long[0]) : (int)(sizeof k_short / sizeof k_short[0]);
  int bad = 0;
  for (int i = 0; i < count; i++) {
    const uint32_t q = tab[i].p;
    int idx = -1;
    int n = ll_choose_fft_length(q, &idx);
    float m = 0.f;
    uint64_t res = 0;
    bool failed = false;
    for (;;) {
      if (n <= 0) {
        printf("No FFT length available for M%u, skipped\n", q);
        break;
      }
      Engine e;
      if (e.open(cfg, q, (uint32_t)n)) return 2;
      printf("Starting self test M%u fft length = %dK\n", q, n / 1024);
      std::vector<int32_t> d(n);
      const uint32_t off = (uint32_t)(cfg.fixed_offset >= 0 ? cfg.fixed_offset % q : (uint32_t)rand() % q);
      ll_initial_digits(q, (uint32_t)n, off, d.data());
      const double t0 = now_s();
      if (ll_load_digits(e.ctx, d.data(), off) || ll_iterate(e.ctx, 10000, nullptr, &m)) {
        fprintf(stderr, "%s\n", ll_last_error(e.ctx));
        return 2;
      }
      if (m > 0.4f) {  // CUDALucas.cu:2767-2771, bounded by the list's upper bound
        const int bigger = larger_length(q, idx, n);
        if (bigger <= 0) {
          printf("Round off error %5.5f > 0.4 at the largest allowed fft length %dK for M%u\n", m, n / 1024, q);
          failed = true;
          break;
        }
        n = bigger;
        continue;
      }
      uint32_t offset = 0;
      if (ll_read_digits(e.ctx, d.data(), &offset)) return 2;
      res = ll_res64(q, (uint32_t)n, d.data(), offset);
      const double dt = now_s() - t0;
      printf("Iteration 10000 / %u, 0x%016llx, %dK, %s, error = %5.5f, real: ", q, (unsigned long long)res,
             n / 1024, kProgram, m);
      print_time((unsigned)dt, 0);
      printf(", %4.4f ms/iter\n", dt * 1000.0 / 10000);
      break;
    }
    if (n <= 0) continue;
    if (failed) {
      bad++;
      continue;
    }
    if (res != tab[i].res64) {
      printf("Expected residue [%016llx] does not match actual residue [%016llx]\n\n",
             (unsigned long long)tab[i].res64, (unsigned long long)res);
      bad++;
    } else {
      printf("This residue is correct.\n\n");
    }
    fflush(stdout);
  }
  if (bad) fprintf(stderr, "Error: There %s %d bad selftest%s!\n", bad > 1 ? "were" : "was", bad, bad > 1 ? "s" : "");
  return bad ? 1 : 0;
}

// ------------------------------------------------------- tuning files
// "<device name> fft.txt" and "<device name> schedules.txt" in the working
// directory, as the reference keeps "<device name> fft.txt" and
// "<device name> threads.txt" (CUDALucas.cu:830, 975).
static std::string device_file(int device, const char *suffix) {
  ll_device_info d;
  if (ll_get_device_info(device < 0 ? 0 : device, &d) != LL_OK) return "";
  // Some ROCm installs report an empty device name: the architecture
  // ("gfx950") names the files then.
  std::string name(d.name);
  if (name.empty()) name = std::string(d.arch).substr(0, std::string(d.arch).find(':'));
  return name + " " + suffix;
}

// A data file shipped in data/ beside the binary.
static std::string packaged_file(const char *name) {
  char exe[4096];
  const ssize_t k = readlink("/proc/self/exe", exe, sizeof exe - 1);
  if (k <= 0) return "";
  exe[k] = 0;
  std::string dir(exe);
  const size_t slash = dir.rfind('/');
  return slash == std::string::npos ? "" : dir.substr(0, slash) + "/data/" + name;
}

// init_device with show_prop (CUDALucas.cu:737-810)
static void print_device_info(int device) {
  ll_device_info d;
  if (ll_get_device_info(device, &d) != LL_OK) {
    printf("device %d not available\n\n", device);
    return;
  }
  printf("%s %d-bit build\n\n", kProgram, (int)(sizeof(void *) * 8));
  printf("HIP runtime version       %d\n", d.runtime_version);
  printf("HIP driver version        %d\n\n", d.driver_version);
  printf("---------------- DEVICE %d ----------------\n", device);
  printf("Device Name               %s\n", d.name);
  printf("Architecture              %s\n", d.arch);
  printf("clockRate (MHz)           %d\n", d.clock_mhz);
  printf("memClockRate (MHz)        %d\n", d.mem_clock_mhz);
  printf("memBusWidth (bits)        %d\n", d.mem_bus_bits);
  printf("totalGlobalMem            %llu\n", (unsigned long long)d.global_mem);
  printf("l2CacheSize               %d\n", d.l2_bytes);
  printf("sharedMemPerBlock         %u\n", d.lds_per_block);
  printf("wavefrontSize             %d\n", d.wavefront);
  printf("maxThreadsPerBlock        %d\n", d.max_threads_per_block);
  printf("multiProcessorCount       %d\n", d.compute_units);
  printf("pciDeviceID               %d\n", d.pci_device);
  printf("pciBusID                  %d\n\n", d.pci_bus);
}

// init_ffts (CUDALucas.cu:947-1025): the lengths listed in fft.txt, with
// the defaults below its first and above its last entry.
static void load_fft_file(const std::string &path) {
  FILE *f = path.empty() ? nullptr : ll_fopen_locked(path.c_str(), "r", 0);
  if (!f) return;
  std::vector<uint32_t> file;
  char buf[256];
  while (fgets(buf, sizeof buf, f)) {
    const int k = atoi(buf);
    if (k > 0 && (file.empty() || (uint32_t)k * 1024u > file.back())) file.push_back((uint32_t)k * 1024u);
  }
  ll_fclose_locked(f);
  if (file.empty()) return;
  ll_fft_length_list_set(0, nullptr);
  std::vector<uint32_t> merged;
  const int count = ll_fft_length_count();
  for (int i = 0; i < count; i++)
    if ((uint32_t)ll_fft_length_at(i) < file.front()) merged.push_back((uint32_t)ll_fft_length_at(i));
  merged.insert(merged.end(), file.begin(), file.end());
  for (int i = 0; i < count; i++)
    if ((uint32_t)ll_fft_length_at(i) > file.back()) merged.push_back((uint32_t)ll_fft_length_at(i));
  ll_fft_length_list_set((uint32_t)merged.size(), merged.data());
}

// init_threads reading threads.txt (CUDALucas.cu:812-945), for schedules:
// lines "K path W H R ms", a later line for the same length wins.
static void load_schedule_file(const std::string &path) {
  FILE *f = path.empty() ? nullptr : ll_fopen_locked(path.c_str(), "r", 0);
  if (!f) return;
  std::vector<ll_schedule> t;
  char buf[256];
  while (fgets(buf, sizeof buf, f)) {
    unsigned k, path_, w, h, r;
    float ms;
    if (sscanf(buf, "%u %u %u %u %u %f", &k, &path_, &w, &h, &r, &ms) != 6) continue;
    ll_schedule e{k * 1024u, path_, w, h, r, ms};
    bool replaced = false;
    for (auto &x : t)
      if (x.n == e.n) {
        x = e;
        replaced = true;
      }
    if (!replaced) t.push_back(e);
  }
  ll_fclose_locked(f);
  ll_schedule_table_set((uint32_t)t.size(), t.data());
}

static bool is_prime_u32(uint32_t n) {
  if (n < 2) return false;
  for (uint32_t d = 2; (uint64_t)d * d <= n; d++)
    if (n % d == 0) return false;
  return true;
}

// Largest prime exponent the length-policy bounds give length n (the "max
// exp" column of fft.txt).
static uint32_t max_exponent(uint32_t n) {
  uint32_t lo = n, hi = 40u * n;
  while (lo + 1 < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    int32_t lb;
    ll_fft_bounds(mid, &lb, nullptr);
    if ((uint32_t)lb <= n) lo = mid;
    else hi = mid;
  }
  uint32_t p = lo | 1u;
  while (p > 3 && !is_prime_u32(p)) p -= 2;
  return p;
}

// Average ms per iteration of length n over `passes` passes of 50
// iterations; *spread receives max - average.
static double time_length(const Config &cfg, uint32_t n, int passes, double *spread) {
  ll_options o;
  memset(&o, 0, sizeof o);
  o.device = cfg.device < 0 ? 0 : cfg.device;
  ll_ctx *c = nullptr;
  const uint32_t p = 17u * n + 1u;
  std::vector<int32_t> d(n);
  ll_initial_digits(p, n, 1, d.data());
  double total = 0, worst = 0;
  int rc = ll_create(p, n, &o, &c);
  if (!rc) rc = ll_load_digits(c, d.data(), 1);
  if (!rc) rc = ll_iterate(c, 64, nullptr, nullptr);
  for (int k = 0; k < passes && !rc; k++) {
    rc = ll_iterate(c, 50, nullptr, nullptr);
    const double t = ll_last_iterate_ms(c) / 50.0;
    total += t;
    if (t > worst) worst = t;
  }
  ll_destroy(c);
  if (rc) return -1.0;
  if (spread) *spread = worst - total / passes;
  return total / passes;
}

// cufftbench (CUDALucas.cu:2413-2587): time every supported length in
// [s, e] K, keep each length that is faster than every longer one (powers of
// two are always kept), and write them to "<device> fft.txt".
static int cufftbench(const Config &cfg, int s, int e, int passes) {
  if (s < 1 || e < s || passes < 1) {
    fprintf(stderr, "can't parse -cufftbench options\n\n");
    return 2;
  }
  printf("FFT bench, testing supported fft sizes %dK to %dK, doing %d passes.\n", s, e, passes);
  std::vector<std::pair<int, double>> t;
  for (int k = s; k <= e; k++) {
    const uint32_t n = (uint32_t)k * 1024u;
    if (!ll_fft_supported(n)) continue;
    double spread = 0;
    const double ms = time_length(cfg, n, passes, &spread);
    if (ms < 0) {
      fprintf(stderr, "fft size = %dK failed\n", k);
      continue;
    }
    printf("fft size = %dK, ave time = %6.4f msec, max-ave = %0.5f\n", k, ms, spread);
    fflush(stdout);
    t.emplace_back(k, ms);
  }
  std::vector<bool> keep(t.size(), false);
  double best = 1e30;
  for (size_t i = t.size(); i-- > 0;) {
    const bool pow2 = (t[i].first & (t[i].first - 1)) == 0;
    if (t[i].second < best) {
      best = t[i].second;
      keep[i] = true;
    } else if (pow2) {
      keep[i] = true;
    }
  }
  const std::string path = device_file(cfg.device, "fft.txt");
  ll_device_info d;
  ll_get_device_info(cfg.device < 0 ? 0 : cfg.device, &d);
  FILE *f = path.empty() ? nullptr : ll_fopen_locked(path.c_str(), "w", 0);
  FILE *out = f ? f : stdout;
  if (!f) printf("Cannot open %s.\n", path.c_str());
  fprintf(out, "Device              %s\n", d.name);
  fprintf(out, "Architecture        %s\n", d.arch);
  fprintf(out, "clockRate (MHz)     %d\n", d.clock_mhz);
  fprintf(out, "memClockRate (MHz)  %d\n", d.mem_clock_mhz);
  fprintf(out, "\n  fft    max exp  ms/iter\n");
  for (size_t i = 0; i < t.size(); i++)
    if (keep[i])
      fprintf(out, "%5d %10u %8.4f\n", t[i].first, max_exponent((uint32_t)t[i].first * 1024u), t[i].second);
  if (f) {
    ll_fclose_locked(f);
    printf("Optimal fft lengths saved in %s.\n", path.c_str());
  }
  return 0;
}

// threadbench (CUDALucas.cu:2143-2276) for this engine's schedules: every
// W x H x R (and the LDS-resident kernel) of each length is timed and the
// fastest appended to "<device> schedules.txt".  mode & 1: only lengths in
// the FFT-length list; mode & 8: print only the best of each length.
// kernel_percentages (CUDALucas.cu:2002-2141), for this engine's kernels:
// "-threadbench 0 K passes" times, in two rounds of `passes` x 50
// iterations at length K, a whole iteration (hipGraph replay) and each
// kernel launched on its own, and prints each kernel's share of their sum.
static int kernel_percentages(const Config &cfg, int k, int passes) {
  const uint32_t n = (uint32_t)k * 1024u;
  if (k < 1 || passes < 1 || !ll_fft_supported(n)) {
    fprintf(stderr, "can't parse -threadbench options (fft %dK, %d passes)\n\n", k, passes);
    return 2;
  }
  const uint32_t p = max_exponent(n) | 1u;
  Engine eng;
  if (eng.open(cfg, p, n)) return 2;
  ll_info info;
  ll_get_info(eng.ctx, &info);
  std::vector<int32_t> d(n);
  ll_initial_digits(p, n, 0, d.data());
  if (ll_load_digits(eng.ctx, d.data(), 0) || ll_iterate(eng.ctx, 100, nullptr, nullptr)) return 2;
  double total[2] = {0, 0}, first[2] = {0, 0}, second[2] = {0, 0};
  for (int round = 0; round < 2; round++)
    for (int pass = 0; pass < passes; pass++) {
      float km[2] = {0, 0};
      if (ll_profile_kernels(eng.ctx, 50, km) || ll_iterate(eng.ctx, 50, nullptr, nullptr)) return 2;
      first[round] += km[0];
      second[round] += km[1];
      total[round] += ll_last_iterate_ms(eng.ctx) / 50.0;
    }
  const double sum = first[0] + second[0];
  printf("Kernel bench, fft = %dK, p = %u, %s %u x %u x %u, %d passes of 50 iterations, two rounds (ms).\n", k, p,
         info.path ? "sweeps" : "resident", info.width, info.height, info.lines, passes);
  printf("Total time: %11.5f, %8.5f, %6.2f\n", total[0] / passes, total[1] / passes, 100.0);
  if (info.path) {
    printf("Square sweep: %9.5f, %8.5f, %6.2f\n", first[0] / passes, first[1] / passes, first[0] / sum * 100.0);
    printf("Carry sweep: %10.5f, %8.5f, %6.2f\n", second[0] / passes, second[1] / passes, second[0] / sum * 100.0);
  } else {
    printf("Resident kernel: %6.5f, %8.5f, %6.2f\n", first[0] / passes, first[1] / passes, 100.0);
  }
  return 0;
}

static int threadbench(const Config &cfg, int s, int e, int passes, int mode) {
  if (s == 0) return kernel_percentages(cfg, e, passes);
  if (s < 1 || e < s || passes < 1) {
    fprintf(stderr, "can't parse -threadbench options\n\n");
    return 2;
  }
  printf("Schedule bench, testing every schedule for ffts %dK to %dK, doing %d passes.\n", s, e, passes);
  const std::string path = device_file(cfg.device, "schedules.txt");
  bool appended = false;
  for (int k = s; k <= e; k++) {
    const uint32_t n = (uint32_t)k * 1024u;
    if (mode & 1) {
      bool listed = false;
      for (int i = 0; i < ll_fft_length_count(); i++) listed |= (uint32_t)ll_fft_length_at(i) == n;
      if (!listed) continue;
    }
    const int cnt = ll_fft_schedules(n, nullptr, 0);
    if (cnt <= 0) continue;
    std::vector<ll_schedule> v(cnt);
    const int got = ll_tune_length(cfg.device < 0 ? 0 : cfg.device, 0, n, 50u * (uint32_t)passes, v.data(), cnt);
    if (got <= 0 || v[0].ms < 0) {
      fprintf(stderr, "fft = %dK: tuning failed\n", k);
      continue;
    }
    if (!(mode & 8))
      for (const auto &x : v)
        printf("fft = %dK, %s %u x %u x %u, ave time = %7.4f ms\n", k, x.path ? "sweeps" : "resident", x.width,
               x.height, x.lines, x.ms);
    printf("fft = %dK, min time = %7.4f ms, %s %u x %u x %u\n", k, v[0].ms, v[0].path ? "sweeps" : "resident",
           v[0].width, v[0].height, v[0].lines);
    if (!(mode & 8)) printf("\n");
    fflush(stdout);
    // the file is locked only while a line is appended (CUDALucas.cu:2269-2274)
    FILE *f = path.empty() ? nullptr : ll_fopen_locked(path.c_str(), "a", 0);
    if (f) {
      fprintf(f, "%5d %u %u %u %u %10.5f\n", k, v[0].path, v[0].width, v[0].height, v[0].lines, v[0].ms);
      ll_fclose_locked(f);
      appended = true;
    }
  }
  if (appended) printf("Best schedules appended to %s.\n", path.c_str());
  return 0;
}


// ------------------------------------------------------------- arguments
static void usage() {
  fprintf(stderr,
          "$ CUDALucasAMD -h|-v\n\n"
          "$ CUDALucasAMD [-d device_number] [-i inifile] [-threads t1 t2] [-c checkpoint_iteration]\n"
          "               [-x report_iteration] [-e error_limit] [-f fft_length] [-s folder] exponent|input_filename\n\n"
          "$ CUDALucasAMD [-d device_number] [-i inifile] -r [0|1]\n\n"
          "  -f    set fft length (e.g. 4096K, 4M)\n"
          "  -r    run the residue self-test (0: 20 Mersenne primes, 1: 92 exponents)\n"
          "  -s    save all checkpoint files into folder\n"
          "  -info print device information\n"
          "  -cufftbench start end passes: time every supported fft length in [start, end] K and\n"
          "        save the fastest ones in \"<device name> fft.txt\"\n"
          "  -threadbench start end passes [mode]: time every kernel schedule of each length and\n"
          "        append the fastest to \"<device name> schedules.txt\"\n"
          "  -threadbench 0 fft passes: time an iteration of length fft K and each kernel's share\n"
          "  --selector model [--selector-times file]: learned fft length selection\n"
          "  -polite and -k are accepted for compatibility and have no effect.\n");
}

static bool need(int argc, char **argv, int i, int k) {
  for (int a = 1; a <= k; a++)
    if (i + a >= argc || argv[i + a][0] == '-') return false;
  return true;
}

int main(int argc, char **argv) {
  Config cfg;
  uint32_t q = 0;
  bool print_config = false;
  printf("\n");
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    if (!strcmp(a, "-h")) { usage(); return 2; }
    if (!strcmp(a, "-v")) { printf("%s\n\n", kProgram); return 2; }
    if (!strcmp(a, "-d") || !strcmp(a, "-i") || !strcmp(a, "-e") || !strcmp(a, "-c") || !strcmp(a, "-x") ||
        !strcmp(a, "-f") || !strcmp(a, "-s") || !strcmp(a, "-polite")) {
      if (!need(argc, argv, i, 1)) { fprintf(stderr, "can't parse %s option\n\n", a); return 2; }
      const char *v = argv[++i];
      if (!strcmp(a, "-d")) cfg.device = atoi(v);
      else if (!strcmp(a, "-i")) cfg.ini = v;
      else if (!strcmp(a, "-e")) cfg.el = atoi(v);
      else if (!strcmp(a, "-c")) { cfg.cpi = atoi(v); if (!cfg.cpi) { fprintf(stderr, "can't parse -c option\n\n"); return 2; } }
      else if (!strcmp(a, "-x")) { cfg.ri = atoi(v); if (!cfg.ri) { fprintf(stderr, "can't parse -x option\n\n"); return 2; } }
      else if (!strcmp(a, "-f")) { cfg.fftlen = ll_fft_from_str(v); if (cfg.fftlen < 0) { fprintf(stderr, "can't parse fft length \"%s\"\n\n", v); return 2; } }
      else if (!strcmp(a, "-s")) { cfg.save_all = 1; cfg.save_folder = v; }
      continue;
    }
    if (!strcmp(a, "-threads")) {
      if (!need(argc, argv, i, 2)) { fprintf(stderr, "can't parse -threads option\n\n"); return 2; }
      cfg.threads[0] = atoi(argv[++i]);
      cfg.threads[1] = atoi(argv[++i]);
      continue;
    }
    if (!strcmp(a, "-r")) {
      cfg.selftest = (i + 1 < argc && argv[i + 1][0] != '-') ? atoi(argv[++i]) : 0;
      continue;
    }
    if (!strcmp(a, "-info")) { cfg.info = true; continue; }
    if (!strcmp(a, "-k")) continue;
    if (!strcmp(a, "-memtest")) {  // CUDALucas.cu:2295-2411, a VRAM burn-in
      fprintf(stderr, "-memtest is not supported: it tests the card's memory, not LL arithmetic\n\n");
      return 2;
    }
    if (!strcmp(a, "-cufftbench")) {
      if (!need(argc, argv, i, 3)) { fprintf(stderr, "can't parse -cufftbench option\n\n"); return 2; }
      for (int k = 0; k < 3; k++) cfg.cufftbench[k] = atoi(argv[++i]);
      continue;
    }
    if (!strcmp(a, "-threadbench")) {
      if (!need(argc, argv, i, 3)) { fprintf(stderr, "can't parse -threadbench option\n\n"); return 2; }
      for (int k = 0; k < 3; k++) cfg.threadbench[k] = atoi(argv[++i]);
      if (i + 1 < argc && argv[i + 1][0] != '-') cfg.threadbench[3] = atoi(argv[++i]);
      continue;
    }
    if (!strcmp(a, "--selector") && i + 1 < argc) { cfg.selector_model = argv[++i]; continue; }
    if (!strcmp(a, "--selector-times") && i + 1 < argc) { cfg.selector_times = argv[++i]; continue; }
    // options without a reference counterpart
    if (!strcmp(a, "--checkpoint-dir") && i + 1 < argc) { cfg.ckpt_dir = argv[++i]; continue; }
    if (!strcmp(a, "--offset") && i + 1 < argc) { cfg.fixed_offset = atoll(argv[++i]); continue; }
    if (!strcmp(a, "--stop-after") && i + 1 < argc) { cfg.stop_after = atoll(argv[++i]); continue; }
    if (!strcmp(a, "--roundoff-test")) { cfg.roundoff = 1; continue; }
    if (!strcmp(a, "--print-config")) { print_config = true; continue; }
    if (!strcmp(a, "--force-fft")) { cfg.force_fft = true; continue; }
    if (!strcmp(a, "--schedule") && i + 1 < argc) {
      sscanf(argv[++i], "%ux%ux%u", &cfg.force_width, &cfg.force_height, &cfg.force_lines);
      continue;
    }
    if (!strcmp(a, "--inject-error") && i + 1 < argc) {
      unsigned long long after = 0;
      float v = 0.f;
      int times = 1;
      if (sscanf(argv[++i], "%llu:%f:%d", &after, &v, &times) < 2) { fprintf(stderr, "can't parse --inject-error\n"); return 2; }
      cfg.inject = times;
      cfg.inject_after = after;
      cfg.inject_value = v;
      continue;
    }
    if (a[0] == '-') { fprintf(stderr, "can't parse options\n\n"); return 2; }
    if (q || !cfg.workfile.empty()) { fprintf(stderr, "can't parse options\n\n"); return 2; }
    const long v = atol(a);
    if (v > 0) q = (uint32_t)v;
    else cfg.workfile = a;
  }
  process_options(cfg);
  if (print_config) {
    printf("device=%d threads=%d,%d checkpoint=%d report=%d error_iter=%d error_limit=%d error_reset=%d "
           "fft=%d save_all=%d save_folder=%s roundoff=%d workfile=%s results=%s output_interval=%d\n",
           cfg.device, cfg.threads[0], cfg.threads[1], cfg.cpi, cfg.ri, cfg.ei, cfg.el, cfg.er, cfg.fftlen,
           cfg.save_all, cfg.save_folder.c_str(), cfg.roundoff, cfg.workfile.c_str(), cfg.resultsfile.c_str(),
           cfg.output_interval);
    return 0;
  }
  // handle SIGINT/SIGTERM even if the parent started us with them blocked
  sigset_t quit_set;
  sigemptyset(&quit_set);
  sigaddset(&quit_set, SIGINT);
  sigaddset(&quit_set, SIGTERM);
  sigprocmask(SIG_UNBLOCK, &quit_set, nullptr);
  signal(SIGINT, on_signal);
  signal(SIGTERM, on_signal);
  const int dev = cfg.device < 0 ? 0 : cfg.device;
  if (cfg.info) print_device_info(dev);
  if (cfg.cufftbench[2] > 0) return cufftbench(cfg, cfg.cufftbench[0], cfg.cufftbench[1], cfg.cufftbench[2]);
  // "<device> schedules.txt" from -threadbench; without one, the table
  // tuned for this GPU that ships beside the binary
  {
    const std::string own = device_file(dev, "schedules.txt");
    FILE *f = own.empty() ? nullptr : fopen(own.c_str(), "r");
    if (f) {
      fclose(f);
      load_schedule_file(own);
    } else {
      load_schedule_file(packaged_file("mi355x_schedules_single.txt"));
    }
  }
  if (cfg.threadbench[2] > 0)
    return threadbench(cfg, cfg.threadbench[0], cfg.threadbench[1], cfg.threadbench[2], cfg.threadbench[3]);
  load_fft_file(device_file(dev, "fft.txt"));
  if (!cfg.selector_model.empty()) {
    if (ll_selector_load(dev, cfg.selector_model.c_str(), &g_sel) != LL_OK) {
      fprintf(stderr, "selector: %s\n", ll_selector_last_error(g_sel));
      return 2;
    }
    if (!cfg.selector_times.empty()) {
      std::vector<uint32_t> ns;
      std::vector<float> ms;
      FILE *f = fopen(cfg.selector_times.c_str(), "r");
      char buf[256];
      while (f && fgets(buf, sizeof buf, f)) {
        unsigned k;
        float t;
        if (buf[0] != '#' && sscanf(buf, "%u %f", &k, &t) == 2) {
          ns.push_back(k * 1024u);
          ms.push_back(t);
        }
      }
      if (f) fclose(f);
      ll_selector_set_times(g_sel, (uint32_t)ns.size(), ns.data(), ms.data());
    }
  }
  if (cfg.selftest >= 0) return check_residue(cfg, cfg.selftest);
  if (q) {
    // valid_assignment; a composite exponent moves on to the next prime
    // (CUDALucas.cu:3560-3575)
    const int valid = ll_valid_assignment(q, cfg.fftlen > 0 ? (uint32_t)cfg.fftlen : 0);
    if (valid == 0) return 2;
    if (valid < 0) {
      uint32_t next = q | 1;
      while (next <= 0x7fffffffu && !ll_is_prime(next)) next += 2;
      if (next > 0x7fffffffu) {
        fprintf(stderr, "Warning: exponent is too big besides!\n");
        return 2;
      }
      fprintf(stderr, "Using %u instead.\n", next);
      q = next;
    }
    return check(q, cfg, nullptr);
  }
  // The worktodo loop (CUDALucas.cu:3549-3558, parse.c:486-678): a length
  // given on the work line holds for that test only.
  for (;;) {
    ll_assignment work;
    if (ll_work_next(cfg.workfile.c_str(), &work)) return 2;
    Config this_test = cfg;
    if (work.fft_length) this_test.fftlen = (int)work.fft_length;
    const int rc = check(work.exponent, this_test, work.key);
    if (rc) return rc;
    if (g_quit) return 0;
    if (ll_work_clear(cfg.workfile.c_str(), work.exponent)) return 2;
  }
}

// Engine contexts behind the C-ABI (include/ll_api.h): schedule choice,
// device tables and buffers, the iteration loop with hipGraph replay,
// load/unload of digits, per-iteration error readback.
#include <errno.h>
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/file.h>
#include <unistd.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "ll_api.h"
#include "ll_host.h"
#include "ll_kernels.h"

using namespace llamd;

namespace {

// A carry sweep waits on other workgroups of its own launch, which each XCD
// dispatches in order; two such launches side by side on one GPU can hold
// the workgroup slots each other's awaited workgroups need.  So every call
// that runs carry sweeps holds its GPU's lock from its first launch to its
// synchronisation: a mutex for the threads of this process, and an
// exclusive flock on "<tmp>/cudalucas_amd-gpu-<PCI bus id>.lock" (TMPDIR,
// else /tmp) for other processes using this library on the same GPU.
// Where the lock file cannot be opened only the in-process mutex applies.
class GpuLock {
 public:
  explicit GpuLock(int device) : dev_(device >= 0 && device < kDevices ? device : 0) {
    mutex_[dev_].lock();
    const int fd = lock_fd(dev_);
    if (fd >= 0) {
      while (flock(fd, LOCK_EX) != 0 && errno == EINTR) {
      }
      locked_ = true;
    }
  }
  ~GpuLock() {
    if (locked_) (void)flock(fds_[dev_], LOCK_UN);
    mutex_[dev_].unlock();
  }
  GpuLock(const GpuLock &) = delete;
  GpuLock &operator=(const GpuLock &) = delete;

 private:
  static constexpr int kDevices = 64;
  static std::mutex mutex_[kDevices];
  static int fds_[kDevices];
  // Opened once per device (under the device's mutex) and kept open.
  static int lock_fd(int dev) {
    if (fds_[dev] != -1) return fds_[dev];
    char bus[64] = {0};
    if (hipDeviceGetPCIBusId(bus, sizeof bus, dev) != hipSuccess) snprintf(bus, sizeof bus, "dev%d", dev);
    for (char *q = bus; *q; q++)
      if (*q == ':' || *q == '.' || *q == '/') *q = '_';
    const char *tmp = getenv("TMPDIR");
    char path[512];
    snprintf(path, sizeof path, "%s/cudalucas_amd-gpu-%s.lock", tmp && *tmp ? tmp : "/tmp", bus);
    const int fd = open(path, O_RDWR | O_CREAT | O_CLOEXEC, 0666);
    fds_[dev] = fd >= 0 ? fd : -2;
    return fds_[dev];
  }
  int dev_;
  bool locked_ = false;
};
std::mutex GpuLock::mutex_[GpuLock::kDevices];
int GpuLock::fds_[GpuLock::kDevices] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                        -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                        -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1,
                                        -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};

static_assert(LL_NOALLOC_SCRATCH_BYTES == kNoAllocScratchBytes, "the header states the scratch size");

constexpr uint32_t kRingSize = 1u << 16;   // per-iteration error slots
constexpr uint32_t kGraphIters = 32;       // iterations per captured graph
constexpr uint32_t kTwoLevelShift = 11;    // 2048-entry fine root table

struct Schedule {
  uint32_t path;  // 0 resident, 1 sweeps
  uint32_t W, H, R;
  const ResidentKernel *rk;
  const SquareKernel *sk;
  const CarryKernel *ck;
  bool generic;   // the runtime-plan kernels (ll_generic.hip)
  uint32_t sq_grid;  // square-sweep workgroups along x (square_grid)
  CarryLaunch ca;     // carry-sweep kernel, workgroup size and grid (carry_launch)
};

bool seven_smooth(uint32_t x) {
  if (!x) return false;
  for (uint32_t f : {2u, 3u, 5u, 7u})
    while (x % f == 0) x /= f;
  return x == 1;
}

// Radix passes of the generic plan of L: the odd prime factors, then the
// power of two in radix-8 passes with a 4 or 2 remainder first.
// Empty if L is not 7-smooth or needs more than kGenMaxPasses passes.
std::vector<uint32_t> generic_radices(uint32_t L) {
  std::vector<uint32_t> r;
  uint32_t x = L;
  for (uint32_t f : {7u, 5u, 3u})
    while (x % f == 0) {
      r.push_back(f);
      x /= f;
    }
  int k = 0;
  while (x % 2 == 0) {
    x /= 2;
    k++;
  }
  if (x != 1) return {};
  if (k % 3) r.push_back(1u << (k % 3));
  for (int i = 0; i < k / 3; i++) r.push_back(8);
  if (r.size() > (size_t)kGenMaxPasses) return {};
  return r;
}

// The generic schedule of M complex points, if any: H | M columns of 64 to
// 4800 points (two fit in LDS), W = M / H even lines, R a power of two
// lines per carry block whose LDS fits; o may force W, H or R.  Preferred:
// H near 2048, R = 8 (128-byte column runs) or 4, fewer line passes.  With
// max_lds below kGenMaxLds (the batched engine: several workgroups per CU)
// both sweeps' LDS stay under it and the split with the fewest workgroups
// (largest items) is preferred.
bool pick_generic(uint32_t M, uint32_t want_r, const ll_options *o, Schedule *s,
                  uint32_t max_lds = kGenMaxLds) {
  double best = 1e30;
  bool found = false;
  const bool balanced = max_lds < kGenMaxLds;
  for (uint32_t H = 64; H <= 4800; H += 2) {
    if (M % H || !seven_smooth(H) || square_generic_lds(H) > max_lds) continue;
    const uint32_t W = M / H;
    if (W % 2 || W < 8 || !seven_smooth(W)) continue;
    if (o && o->width && o->width != W) continue;
    if (o && o->height && o->height != H) continue;
    if (generic_radices(W).empty() || generic_radices(H).empty()) continue;
    for (uint32_t R = 16; R >= 1; R /= 2) {
      if (H % R || carry_generic_lds(W, R) > max_lds) continue;
      if (want_r && want_r != R) continue;
      // Under an LDS bound (the batched engine's work items): the fewest
      // items, W / 2 column pairs plus H / R line blocks, so each item has
      // the most work for its fixed latency (measured on the 256-exponent
      // workload: 0.624 ms per batch iteration at H near 2W and R = 8,
      // 0.511 ms with the fewest items); R below 8 reads columns in short
      // runs.  Otherwise columns near 2048 points and R = 8.
      double cost;
      if (balanced) {
        cost = log2((double)(W / 2 + H / R)) * 2.0;
        cost += R >= 8 ? 0.0 : R == 4 ? 0.6 : R == 2 ? 1.5 : 4.0;
      } else {
        cost = fabs(log2((double)H) - 11.0) * 2.0;
        cost += R == 8 ? 0.0 : R == 4 ? 0.5 : R == 16 ? 0.7 : R == 2 ? 1.5 : 4.0;
      }
      cost += 0.1 * (double)generic_radices(W).size();
      if (cost < best) {
        best = cost;
        found = true;
        s->path = 1;
        s->W = W;
        s->H = H;
        s->R = R;
        s->generic = true;
        s->sk = generic_square();
        s->ck = generic_carry();
      }
    }
  }
  return found;
}

// Preferred schedules installed by ll_schedule_table_set (a tuning file).
std::vector<ll_schedule> &schedule_table() {
  static std::vector<ll_schedule> t;
  return t;
}

bool pick_schedule_auto(uint32_t n, const ll_options *o, Schedule *s);

bool pick_schedule(uint32_t n, const ll_options *o, Schedule *s) {
  const bool explicit_choice = o && (o->width || o->height || o->lines || o->threads[1] > 0 ||
                                     (o->flags & (LL_OPT_FORCE_SWEEPS | LL_OPT_GENERIC)));
  if (!explicit_choice) {
    for (const ll_schedule &e : schedule_table()) {
      if (e.n != n) continue;
      ll_options t;
      memset(&t, 0, sizeof t);
      if (o) t = *o;
      if (e.path == 1) {
        t.flags |= LL_OPT_FORCE_SWEEPS;
        t.width = e.width;
        t.height = e.height;
        t.lines = e.lines;
      }
      if (pick_schedule_auto(n, &t, s)) return true;
      break;  // stale entry: fall back to the built-in preference
    }
  }
  return pick_schedule_auto(n, o, s);
}

bool pick_schedule_auto(uint32_t n, const ll_options *o, Schedule *s) {
  memset(s, 0, sizeof(*s));
  if (n < 512 || n % 2) return false;
  const uint32_t M = n / 2;
  const uint32_t flags = o ? o->flags : 0;
  if (!(flags & (LL_OPT_FORCE_SWEEPS | LL_OPT_GENERIC)) && !(o && (o->width || o->height))) {
    if ((s->rk = find_resident(M))) {
      s->path = 0;
      return true;
    }
  }
  uint32_t want_r = 0;
  if (o && o->lines) want_r = o->lines;
  else if (o && o->threads[1] > 0) {
    want_r = (uint32_t)o->threads[1] / 32;
    want_r = want_r <= 2 ? 2 : want_r >= 8 ? 8 : 4;
  }
  if (flags & LL_OPT_GENERIC) return pick_generic(M, want_r, o, s);
  size_t count = 0;
  const CarryKernel *tab = carry_table(&count);
  double best = 1e30;
  for (size_t i = 0; i < count; i++) {
    const CarryKernel *ck = &tab[i];
    if (M % ck->w) continue;
    const uint32_t H = M / ck->w;
    const SquareKernel *sk = find_square(H);
    if (!sk || H % ck->r) continue;
    if (o && o->width && o->width != ck->w) continue;
    if (o && o->height && o->height != H) continue;
    if (want_r && want_r != ck->r) continue;
    // prefer H = 2048, then R = 8 (128-byte column runs; R = 1 reads each
    // column element on its own, which the 16M+ lengths measured 1.3-1.7x
    // slower than R = 2 beside 4096-point columns), then W close to H / 2
    double cost = fabs(log2((double)H) - 11.0) * 4.0;
    cost += ck->r == 8 ? 0.0 : ck->r == 4 ? 1.0 : ck->r == 2 ? 2.0 : 5.0;
    cost += fabs(log2((double)ck->w) - log2((double)H / 2.0));
    if (cost < best) {
      best = cost;
      s->path = 1;
      s->W = ck->w;
      s->H = H;
      s->R = ck->r;
      s->ck = ck;
      s->sk = sk;
    }
  }
  if (s->ck) return true;
  return pick_generic(M, want_r, o, s);
}

}  // namespace

// Small host values that cross to and from the device on every call.
struct HostWords {
  uint32_t offset;
  uint32_t unload_err;
  ResidentDesc rdesc;
};

struct ll_ctx {
  uint32_t p = 0, n = 0, m = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  Schedule sch{};
  bool loaded = false;
  bool use_graph = true;
  uint32_t offset = 0;       // host copy of the shift offset
  uint64_t epoch = 0;        // carry-sweep launches so far
  uint64_t iterations = 0;   // iterations since the last load
  float max_err = 0.f;
  char msg[256] = {0};
  // fault injection
  uint32_t inject_kind = 0;
  uint64_t inject_after = 0;
  float inject_value = 0.f;
  // device memory (both paths)
  std::vector<void *> allocs;
  cplx *x = nullptr;
  int32_t *digits = nullptr;
  uint32_t *unload_err = nullptr;
  // sweeps
  PlanArgs args{};
  PlanArgs *d_args = nullptr;         // device copy of args
  PlanArgs *d_args_unload = nullptr;  // args with err -> unload_err
  // the no-allocation path: args whose shuttle and error slots are in the
  // device's static scratch array (null if the shuttle does not fit)
  PlanArgs *d_args_static = nullptr;
  uint8_t *scratch = nullptr;
  size_t shuttle_bytes = 0;
  uint32_t ring_size = kRingSize;
  uint32_t *ring = nullptr;
  DevState *st = nullptr;
  hipGraphExec_t graph = nullptr;
  // resident
  ResidentDesc rdesc{};
  ResidentDesc *d_rdesc = nullptr;
  float *r_err = nullptr;
  uint32_t *r_offset = nullptr;  // device shift offset of the resident path
  // pinned host memory: the error slots of one call, the offset, the
  // resident descriptor (so readbacks and uploads need no staging)
  uint32_t *h_ring = nullptr;
  HostWords *h_words = nullptr;
  // timing (hipEvents on `stream`)
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_gpu_ms = 0.0;
};

static int fail(ll_ctx *c, int code, const char *fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->msg, sizeof(c->msg), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(c, expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((c), LL_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

template <typename T>
static int dalloc(ll_ctx *c, T **ptr, size_t count) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 16);
  if (e != hipSuccess) return fail(c, LL_ERR_ALLOC, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
  c->allocs.push_back(p);
  *ptr = (T *)p;
  return LL_OK;
}

template <typename T>
static int upload(ll_ctx *c, T **ptr, const std::vector<T> &v) {
  int rc = dalloc(c, ptr, v.size());
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LL_OK;
}

// Tables of one generic plan: radices, roots exp(-2 pi i e/L), and the
// digit reversal of its in-place forward transform (position q = sum k_p
// s_p, s_p = L / (r_0 ... r_p), holds frequency sum k_p r_0 ... r_(p-1)).
static int build_generic_line(ll_ctx *c, uint32_t L, GenLine *g) {
  const std::vector<uint32_t> rad = generic_radices(L);
  if (rad.empty()) return fail(c, LL_ERR_UNSUPPORTED, "no generic plan for length %u", L);
  memset(g, 0, sizeof(*g));
  g->L = L;
  g->np = (uint32_t)rad.size();
  for (size_t i = 0; i < rad.size(); i++) g->radix[i] = (uint8_t)rad[i];
  std::vector<double> roots(2 * (size_t)L);
  for (uint32_t e = 0; e < L; e++) root_of_unity(e, L, &roots[2 * e], &roots[2 * e + 1]);
  std::vector<uint32_t> perm(L), iperm(L);
  for (uint32_t q = 0; q < L; q++) {
    uint32_t span = L, place = 1, f = 0;
    for (uint32_t r : rad) {
      span /= r;
      f += ((q / span) % r) * place;
      place *= r;
    }
    perm[q] = f;
    iperm[f] = q;
  }
  int rc;
  if ((rc = upload(c, (double **)&g->roots, roots))) return rc;
  if ((rc = upload(c, (uint32_t **)&g->perm, perm))) return rc;
  if ((rc = upload(c, (uint32_t **)&g->iperm, iperm))) return rc;
  return LL_OK;
}

static int build_sweeps(ll_ctx *c) {
  const uint32_t p = c->p, n = c->n, M = c->m, W = c->sch.W, H = c->sch.H, R = c->sch.R;
  PlanArgs &A = c->args;
  std::vector<double> qc;
  int rc;
  memset(&A.gW, 0, sizeof(A.gW));
  memset(&A.gH, 0, sizeof(A.gH));
  if (c->sch.generic) {
    if ((rc = build_generic_line(c, W, &A.gW))) return rc;
    if ((rc = build_generic_line(c, H, &A.gH))) return rc;
  }
  quarter_cosines(W, qc);
  if ((rc = upload(c, (double **)&A.qcW, qc))) return rc;
  quarter_cosines(H, qc);
  if ((rc = upload(c, (double **)&A.qcH, qc))) return rc;
  cplx *ptH = nullptr;
  if ((rc = dalloc(c, &ptH, H))) return rc;
  if (c->sch.sk->table) HIPCHK(c, launch_square_table(c->sch.sk, A.qcH, ptH, c->stream));
  A.ptH = ptH;
  A.ptW = nullptr;
  if (c->sch.ck->table) {
    cplx *ptW = nullptr;
    if ((rc = dalloc(c, &ptW, W))) return rc;
    HIPCHK(c, launch_carry_table(c->sch.ck, A.qcW, ptW, c->stream));
    A.ptW = ptW;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // two-level roots of order M
  const uint32_t shift = std::min<uint32_t>(kTwoLevelShift, (uint32_t)log2((double)M));
  const uint32_t S = 1u << shift;
  const uint32_t nhi = (M + S - 1) / S;
  // coarse roots as (rounded value, rounding remainder) pairs
  std::vector<double> hi(4 * (size_t)nhi), lo(2 * (size_t)S);
  for (uint32_t i = 0; i < nhi; i++) {
    long double re, im;
    root_of_unity_ld((uint64_t)i * S, M, &re, &im);
    hi[4 * i] = (double)re;
    hi[4 * i + 1] = (double)im;
    hi[4 * i + 2] = (double)(re - (long double)hi[4 * i]);
    hi[4 * i + 3] = (double)(im - (long double)hi[4 * i + 1]);
  }
  for (uint32_t i = 0; i < S; i++) root_minus_one(i, M, &lo[2 * i], &lo[2 * i + 1]);
  if ((rc = upload(c, (double **)&A.twM.hi, hi))) return rc;
  if ((rc = upload(c, (double **)&A.twM.lo, lo))) return rc;
  A.twM.shift = shift;
  A.twM.mask = S - 1;
  // factored weights: word 2H a + i has f = fA[a] + fB[i] (mod n)
  std::vector<double> wA(W), iwA(W), iwAs(W), wB(2 * H), iwB(2 * H);
  std::vector<int32_t> fA(W), fB(2 * H);
  for (uint32_t a = 0; a < W; a++) {
    fA[a] = (int32_t)weight_exponent(p, n, (uint64_t)2 * H * a);
    weight_pair((uint32_t)fA[a], n, &wA[a], &iwA[a]);
    iwAs[a] = inverse_weight_scaled((uint32_t)fA[a], n, 4ull * M);
  }
  for (uint32_t i = 0; i < 2 * H; i++) {
    fB[i] = (int32_t)weight_exponent(p, n, i);
    weight_pair((uint32_t)fB[i], n, &wB[i], &iwB[i]);
  }
  if ((rc = upload(c, (double **)&A.wA, wA))) return rc;
  if ((rc = upload(c, (double **)&A.iwA, iwA))) return rc;
  if ((rc = upload(c, (double **)&A.iwAs, iwAs))) return rc;
  if ((rc = upload(c, (int32_t **)&A.fA, fA))) return rc;
  if ((rc = upload(c, (double **)&A.wB, wB))) return rc;
  if ((rc = upload(c, (double **)&A.iwB, iwB))) return rc;
  if ((rc = upload(c, (int32_t **)&A.fB, fB))) return rc;
  const size_t shuttle = (size_t)(H / R) * W;
  c->shuttle_bytes = shuttle * sizeof(uint64_t);
  if ((rc = dalloc(c, &A.shuttle, shuttle))) return rc;
  HIPCHK(c, hipMemsetAsync(A.shuttle, 0xff, shuttle * sizeof(uint64_t), c->stream));
  if ((rc = dalloc(c, &c->st, 1))) return rc;
  HIPCHK(c, hipMemsetAsync(c->st, 0, sizeof(DevState), c->stream));
  if ((rc = dalloc(c, &c->ring, c->ring_size))) return rc;
  A.x = c->x;
  A.st = c->st;
  A.err = c->ring;
  A.err_mask = c->ring_size - 1;
  A.digits = c->digits;
  A.scale = 0.25 / (double)M;
  A.base_lo = ldexp(1.0, (int)(p / n));
  A.rshift = (uint32_t)log2((double)R);
  A.p = p;
  A.n = n;
  A.W = W;
  A.H = H;
  A.q = p / n;
  A.b = p % n;
  A.remaining = nullptr;
  A.stagger_sq = A.stagger_ca = 0;
  if (c->sch.sk->h) A.stagger_sq = sweep_stagger(c->sch.sk->fn, c->sch.sk->threads, 0, W / 2, 32 * H, c->device);
  c->sch.sq_grid = square_grid(c->sch.sk, W, 1);
  c->sch.ca = carry_launch(c->sch.ck, H, R, 1);
  if (c->sch.ck->w) A.stagger_ca = sweep_stagger(c->sch.ca.fn, c->sch.ca.threads, 0, H / R, 16 * R * W, c->device);
  if ((rc = dalloc(c, &c->d_args, 1))) return rc;
  if ((rc = dalloc(c, &c->d_args_unload, 1))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_args, &A, sizeof(PlanArgs), hipMemcpyHostToDevice, c->stream));
  PlanArgs U = A;
  U.err = c->unload_err;
  HIPCHK(c, hipMemcpyAsync(c->d_args_unload, &U, sizeof(PlanArgs), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->shuttle_bytes <= kNoAllocShuttleBytes) {
    PlanArgs S = A;
    S.shuttle = (uint64_t *)c->scratch;
    S.err = (uint32_t *)(c->scratch + kNoAllocShuttleBytes);
    S.err_mask = kNoAllocRing - 1;
    if ((rc = dalloc(c, &c->d_args_static, 1))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_args_static, &S, sizeof(PlanArgs), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return LL_OK;
}

static int build_resident(ll_ctx *c) {
  const uint32_t p = c->p, n = c->n, M = c->m;
  std::vector<double> w(n), iw(n), qc;
  for (uint32_t j = 0; j < n; j++) weight_pair(weight_exponent(p, n, j), n, &w[j], &iw[j]);
  quarter_cosines(M, qc);
  ResidentDesc &D = c->rdesc;
  int rc;
  if ((rc = upload(c, (double **)&D.w, w))) return rc;
  if ((rc = upload(c, (double **)&D.iw, iw))) return rc;
  // factored weights for the iteration kernel's T threads of C words
  const uint32_t T = c->sch.rk->threads, C = n / T;
  std::vector<double> wT(T), iwT(T), iwTs(T), wC(C), iwC(C);
  std::vector<int32_t> fT(T), fC(C);
  for (uint32_t t = 0; t < T; t++) {
    fT[t] = (int32_t)weight_exponent(p, n, (uint64_t)t * C);
    weight_pair((uint32_t)fT[t], n, &wT[t], &iwT[t]);
    iwTs[t] = inverse_weight_scaled((uint32_t)fT[t], n, 4ull * M);
  }
  for (uint32_t i = 0; i < C; i++) {
    fC[i] = (int32_t)weight_exponent(p, n, i);
    weight_pair((uint32_t)fC[i], n, &wC[i], &iwC[i]);
  }
  if ((rc = upload(c, (double **)&D.wT, wT))) return rc;
  if ((rc = upload(c, (double **)&D.iwTs, iwTs))) return rc;
  if ((rc = upload(c, (double **)&D.wC, wC))) return rc;
  if ((rc = upload(c, (double **)&D.iwC, iwC))) return rc;
  if ((rc = upload(c, (int32_t **)&D.fT, fT))) return rc;
  if ((rc = upload(c, (int32_t **)&D.fC, fC))) return rc;
  if ((rc = upload(c, (double **)&D.qc, qc))) return rc;
  cplx *pt = nullptr;
  if ((rc = dalloc(c, &pt, M))) return rc;
  HIPCHK(c, launch_resident_table(c->sch.rk, D.qc, pt, c->stream));
  D.pt = pt;
  if ((rc = dalloc(c, &c->r_err, c->ring_size))) return rc;
  if ((rc = dalloc(c, &c->r_offset, 1))) return rc;
  if ((rc = dalloc(c, &c->d_rdesc, 1))) return rc;
  D.x = c->x;
  D.err = c->r_err;
  D.offset = c->r_offset;
  D.scale = 0.25 / (double)M;
  D.remaining = nullptr;
  D.p = p;
  D.m = M;
  D.iters = 0;
  return LL_OK;
}

static void release(ll_ctx *c) {
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  if (c->h_words) (void)hipHostFree(c->h_words);
  c->h_ring = nullptr;
  c->h_words = nullptr;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

static int build_graph(ll_ctx *c) {
  if (c->graph) return LL_OK;
  hipGraph_t g = nullptr;
  // relaxed: other threads may create contexts (and copy) while this captures
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  for (uint32_t i = 0; i < kGraphIters; i++) {
    (void)launch_square(c->sch.sk, c->d_args, 1, c->sch.W, c->sch.H, c->sch.sq_grid, c->stream);
    (void)launch_carry(c->sch.ck, c->sch.ca, c->d_args, 1, c->sch.W, c->sch.H, c->sch.R, c->stream);
  }
  HIPCHK(c, hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return fail(c, LL_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return LL_OK;
}

extern "C" {

int ll_fft_supported(uint32_t n) {
  Schedule s;
  return pick_schedule(n, nullptr, &s) ? 1 : 0;
}

static int ctx_create(uint32_t p, uint32_t n, const ll_options *opts, uint32_t ring_size, bool graph,
                      ll_ctx **out);

int ll_create(uint32_t p, uint32_t n, const ll_options *opts, ll_ctx **out) {
  return ctx_create(p, n, opts, kRingSize, true, out);
}

}  // extern "C"

// graph: instantiate the iteration graph now (two-sweep path), so that no
// ll_iterate call pays for it.
static int ctx_create(uint32_t p, uint32_t n, const ll_options *opts, uint32_t ring_size, bool graph,
                      ll_ctx **out) {
  if (!out) return LL_ERR_ARG;
  *out = nullptr;
  if (p < 3 || n < 512 || n % 2 || (uint64_t)p < 4ull * n / 3) return LL_ERR_ARG;
  if ((double)p / n > 31.0) return LL_ERR_ARG;  // digits must fit int32 with headroom
  if (p / n < 2) return LL_ERR_ARG;              // digits of at least 2 bits
  ll_ctx *c = new ll_ctx();
  c->p = p;
  c->n = n;
  c->m = n / 2;
  c->ring_size = ring_size;
  c->device = opts ? opts->device : 0;
  c->use_graph = !(opts && (opts->flags & LL_OPT_NO_GRAPH));
  if (!pick_schedule(n, opts, &c->sch)) {
    fail(c, LL_ERR_UNSUPPORTED, "no kernel schedule for FFT length %u", n);
    *out = c;  // caller can read the message, then destroy
    return LL_ERR_UNSUPPORTED;
  }
  int rc = LL_OK;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) rc = fail(c, LL_ERR_HIP, "hipSetDevice(%d): %s", c->device, hipGetErrorString(e));
  if (!rc) {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(c, LL_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  for (int i = 0; i < 4 && !rc; i++) {
    e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) rc = fail(c, LL_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  if (!rc && (hipHostMalloc((void **)&c->h_ring, sizeof(uint32_t) * ring_size, 0) != hipSuccess ||
              hipHostMalloc((void **)&c->h_words, sizeof(HostWords), 0) != hipSuccess))
    rc = fail(c, LL_ERR_ALLOC, "hipHostMalloc of the readback buffers failed");
  if (!rc) {
    void *sp = nullptr;
    e = noalloc_scratch(c->device, &sp);
    if (e != hipSuccess) rc = fail(c, LL_ERR_HIP, "no-allocation scratch lookup: %s", hipGetErrorString(e));
    c->scratch = (uint8_t *)sp;
  }
  if (!rc) rc = dalloc(c, &c->x, c->m);
  if (!rc) rc = dalloc(c, &c->digits, n);
  if (!rc) rc = dalloc(c, &c->unload_err, 1);
  if (!rc) rc = c->sch.path == 0 ? build_resident(c) : build_sweeps(c);
  if (!rc && graph && c->sch.path == 1 && c->use_graph) rc = build_graph(c);
  *out = c;
  return rc;
}

extern "C" {

void ll_destroy(ll_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
}

int ll_load_digits(ll_ctx *c, const int32_t *digits, uint32_t offset) {
  if (!c || !digits) return LL_ERR_ARG;
  if (!c->x) return fail(c, LL_ERR_STATE, "context not built");
  HIPCHK(c, hipSetDevice(c->device));
  offset %= c->p;
  HIPCHK(c, hipMemcpyAsync(c->digits, digits, sizeof(int32_t) * c->n, hipMemcpyHostToDevice, c->stream));
  c->h_words->offset = offset;
  if (c->sch.path == 0) {
    c->h_words->rdesc = c->rdesc;
    c->h_words->rdesc.iters = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &c->h_words->rdesc, sizeof(ResidentDesc), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->r_offset, &c->h_words->offset, sizeof(uint32_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, launch_resident_load(c->d_rdesc, 1, c->digits, c->stream));
  } else {
    HIPCHK(c, launch_load(c->sch.ck, c->d_args, 1, c->sch.W, c->sch.H, c->sch.R, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->st->offset, &c->h_words->offset, sizeof(uint32_t), hipMemcpyHostToDevice,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->offset = offset;
  c->loaded = true;
  c->iterations = 0;
  return LL_OK;
}

// Both iterate forms run `count` iterations (no more than the error slots
// they are given), then bring back only the error slots those iterations
// wrote, and the offset, through pinned host memory (c->h_ring, c->h_words):
// a call costs its kernels plus one synchronisation.
static int iterate_resident(ll_ctx *c, uint32_t count, float *dev_err) {
  HostWords &hw = *c->h_words;
  hw.rdesc = c->rdesc;
  hw.rdesc.iters = count;
  hw.rdesc.err = dev_err;
  HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &hw.rdesc, sizeof(ResidentDesc), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, launch_resident(c->sch.rk, c->d_rdesc, 1, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_ring, dev_err, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&hw.offset, c->r_offset, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->last_gpu_ms += ms;
  c->offset = hw.offset;
  return LL_OK;
}

// dargs: the PlanArgs the kernels read; ring/ring_size: the error slots
// those args name; graph: replay the context's captured graph (built on
// c->d_args) for whole groups of kGraphIters iterations.
static int iterate_sweeps(ll_ctx *c, uint32_t count, const PlanArgs *dargs, const uint32_t *ring,
                          uint32_t ring_size, bool graph) {
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  uint32_t left = count;
  if (graph && c->use_graph && count >= kGraphIters) {
    int rc = build_graph(c);
    if (rc) return rc;
    while (left >= kGraphIters) {
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      left -= kGraphIters;
    }
  }
  for (; left; left--) {
    HIPCHK(c, launch_square(c->sch.sk, dargs, 1, c->sch.W, c->sch.H, c->sch.sq_grid, c->stream));
    HIPCHK(c, launch_carry(c->sch.ck, c->sch.ca, dargs, 1, c->sch.W, c->sch.H, c->sch.R, c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  // the slots of epochs epoch .. epoch + count - 1, which may wrap the ring
  const uint32_t first = (uint32_t)(c->epoch & (ring_size - 1));
  const uint32_t n1 = std::min(count, ring_size - first);
  HIPCHK(c, hipMemcpyAsync(c->h_ring, ring + first, sizeof(uint32_t) * n1, hipMemcpyDeviceToHost, c->stream));
  if (n1 < count)
    HIPCHK(c, hipMemcpyAsync(c->h_ring + n1, ring, sizeof(uint32_t) * (count - n1), hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(&c->h_words->offset, &c->st->offset, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->last_gpu_ms += ms;
  c->epoch += count;
  c->offset = c->h_words->offset;
  return LL_OK;
}

// ll_iterate and ll_iterate_noalloc: the same kernels, with the carry
// shuttle and error slots either the context's own or the device's static
// scratch array.  Neither allocates; the no-allocation form also replays no
// graph (direct launches only).
static int iterate_common(ll_ctx *c, uint32_t count, float *per_iter, float *max_err, bool noalloc) {
  if (!c) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_iterate before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->inject_kind == 2) {
    c->inject_kind = 0;
    return fail(c, LL_ERR_HIP, "injected device error");
  }
  uint32_t slots = c->ring_size;
  if (noalloc) {
    slots = std::min(slots, kNoAllocRing);
    if (c->sch.path == 1 && !c->d_args_static)
      return fail(c, LL_ERR_ARG, "carry shuttle of %zu bytes exceeds the %zu-byte no-allocation scratch",
                  c->shuttle_bytes, kNoAllocShuttleBytes);
  }
  c->last_gpu_ms = 0.0;
  float m = 0.f;
  uint32_t done = 0;
  std::unique_ptr<GpuLock> lock;
  if (c->sch.path == 1) lock.reset(new GpuLock(c->device));
  if (noalloc && c->sch.path == 1)  // stale carries of other contexts must not carry this epoch's tag
    HIPCHK(c, launch_fill((uint32_t *)c->scratch, 0xffffffffu, c->shuttle_bytes / 4, c->stream));
  while (done < count) {
    const uint32_t chunk = std::min<uint32_t>(count - done, slots);
    int rc;
    if (c->sch.path == 0) {
      float *dev_err = noalloc ? (float *)(c->scratch + kNoAllocShuttleBytes) : c->r_err;
      rc = iterate_resident(c, chunk, dev_err);
    } else if (noalloc) {
      rc = iterate_sweeps(c, chunk, c->d_args_static, (const uint32_t *)(c->scratch + kNoAllocShuttleBytes),
                          kNoAllocRing, false);
    } else {
      rc = iterate_sweeps(c, chunk, c->d_args, c->ring, c->ring_size, true);
    }
    if (rc) return rc;
    float *errs = (float *)c->h_ring;
    if (c->inject_kind == 1) {
      if (c->inject_after < chunk) {
        errs[c->inject_after] = c->inject_value;
        c->inject_kind = 0;
      } else {
        c->inject_after -= chunk;
      }
    }
    for (uint32_t i = 0; i < chunk; i++) m = std::max(m, errs[i]);
    if (per_iter) memcpy(per_iter + done, errs, sizeof(float) * chunk);
    done += chunk;
    c->iterations += chunk;
  }
  c->max_err = std::max(c->max_err, m);
  if (max_err) *max_err = m;
  return LL_OK;
}

int ll_iterate(ll_ctx *c, uint32_t count, float *per_iter, float *max_err) {
  return iterate_common(c, count, per_iter, max_err, false);
}

int ll_iterate_noalloc(ll_ctx *c, uint32_t count, float *per_iter, float *max_err) {
  return iterate_common(c, count, per_iter, max_err, true);
}

int ll_read_digits(ll_ctx *c, int32_t *digits, uint32_t *offset) {
  if (!c || !digits) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_read_digits before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_fill(c->unload_err, 0u, 1, c->stream));
  HostWords &hw = *c->h_words;
  if (c->sch.path == 0) {
    hw.rdesc = c->rdesc;
    hw.rdesc.iters = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &hw.rdesc, sizeof(ResidentDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_resident_unload(c->d_rdesc, 1, c->digits, c->unload_err, c->stream));
    HIPCHK(c, hipMemcpyAsync(&hw.offset, c->r_offset, 4, hipMemcpyDeviceToHost, c->stream));
  } else {
    HIPCHK(c, launch_unload(c->sch.ck, c->d_args_unload, 1, c->sch.W, c->sch.H, c->sch.R, c->stream));
    HIPCHK(c, hipMemcpyAsync(&hw.offset, &c->st->offset, 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(digits, c->digits, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&hw.unload_err, c->unload_err, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t dev_offset = hw.offset;
  float e;
  memcpy(&e, &hw.unload_err, 4);
  if (!(e < 0.05f)) return fail(c, LL_ERR_ROUNDOFF, "digit readback rounding error %g", (double)e);
  c->offset = dev_offset;
  if (offset) *offset = dev_offset;
  return LL_OK;
}

double ll_last_iterate_ms(const ll_ctx *c) { return c ? c->last_gpu_ms : -1.0; }

int ll_profile_kernels(ll_ctx *c, uint32_t count, float *kernel_ms) {
  if (!c || !kernel_ms || count == 0) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_profile_kernels before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  kernel_ms[0] = kernel_ms[1] = 0.f;
  if (c->sch.path == 0) {
    for (uint32_t done = 0; done < count;) {
      const uint32_t chunk = std::min<uint32_t>(count - done, c->ring_size);
      c->last_gpu_ms = 0.0;
      int rc = iterate_resident(c, chunk, c->r_err);
      if (rc) return rc;
      kernel_ms[0] += (float)c->last_gpu_ms;
      done += chunk;
    }
    kernel_ms[0] /= (float)count;  // per iteration inside the resident kernel
    c->iterations += count;
    return LL_OK;
  }
  double sq = 0.0, ca = 0.0;
  GpuLock lock(c->device);
  for (uint32_t i = 0; i < count; i++) {
    // the events take each kernel's own start and end (hipExtLaunchKernel),
    // not the dispatch gaps between the launches
    HIPCHK(c, launch_square(c->sch.sk, c->d_args, 1, c->sch.W, c->sch.H, c->sch.sq_grid, c->stream, c->ev[0],
                            c->ev[1]));
    HIPCHK(c, launch_carry(c->sch.ck, c->sch.ca, c->d_args, 1, c->sch.W, c->sch.H, c->sch.R, c->stream, c->ev[2],
                           c->ev[3]));
    HIPCHK(c, hipEventSynchronize(c->ev[3]));
    float a = 0.f, b = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
    sq += a;
    ca += b;
  }
  uint32_t off = 0;
  HIPCHK(c, hipMemcpyAsync(&c->h_words->offset, &c->st->offset, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  off = c->h_words->offset;
  c->offset = off;
  c->epoch += count;
  c->iterations += count;
  kernel_ms[0] = (float)(sq / count);
  kernel_ms[1] = (float)(ca / count);
  return LL_OK;
}

int ll_bench_copy(int32_t device, uint64_t bytes, uint32_t reps, float *ms) {
  if (!ms || bytes < 16 || reps == 0) return LL_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return LL_ERR_HIP;
  const size_t n = bytes / 16;
  cplx *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LL_OK;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) rc = LL_ERR_ALLOC;
  if (!rc && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = LL_ERR_HIP;
  if (!rc && hipMemsetAsync(a, 0, n * 16, s) != hipSuccess) rc = LL_ERR_HIP;
  if (!rc && launch_copy(a, b, n, s) != hipSuccess) rc = LL_ERR_HIP;  // warm-up
  if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = LL_ERR_HIP;
  for (uint32_t i = 0; !rc && i < reps; i++)
    if (launch_copy(i & 1 ? b : a, i & 1 ? a : b, n, s) != hipSuccess) rc = LL_ERR_HIP;
  if (!rc && hipEventRecord(e1, s) != hipSuccess) rc = LL_ERR_HIP;
  if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = LL_ERR_HIP;
  float t = 0.f;
  if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = LL_ERR_HIP;
  *ms = t / (float)reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int ll_reset_error(ll_ctx *c) {
  if (!c) return LL_ERR_ARG;
  c->max_err = 0.f;
  return LL_OK;
}

float ll_error_max(const ll_ctx *c) { return c ? c->max_err : -1.f; }

int ll_get_info(const ll_ctx *c, ll_info *info) {
  if (!c || !info) return LL_ERR_ARG;
  info->p = c->p;
  info->n = c->n;
  info->path = c->sch.generic ? 2 : c->sch.path;
  info->width = c->sch.path ? c->sch.W : c->m;
  info->height = c->sch.path ? c->sch.H : 1;
  info->lines = c->sch.path ? c->sch.R : 1;
  info->stream = (void *)c->stream;
  info->iterations = c->iterations;
  return LL_OK;
}

const char *ll_last_error(const ll_ctx *c) { return c ? c->msg : "null context"; }

int ll_debug_inject(ll_ctx *c, uint32_t kind, uint64_t after, float value) {
  if (!c || kind > 2) return LL_ERR_ARG;
  c->inject_kind = kind;
  c->inject_after = after;
  c->inject_value = value;
  return LL_OK;
}

}  // extern "C"

// ===================================================================== batch
// Many LL tests of uneven exponents and lengths on one GPU.  Tests of one
// FFT length and schedule form a group; a group of two-sweep tests runs as
// one square and one carry launch per iteration (blockIdx.y = test), a group
// of LDS-resident tests as one launch per step (one workgroup per test),
// on a second stream so it overlaps the HBM-bound sweeps.  Per-test
// device-side "remaining" counters retire finished tests inside replayed
// hipGraphs without host involvement.

namespace {

constexpr uint32_t kBatchRing = 1024;  // iterations per batch step at most
constexpr uint32_t kBatchLds = 40 * 1024;  // LDS per batched workgroup: 4 per CU
// Two-sweep tests up to this length run in the single-launch iteration on
// the generic bodies; longer ones, where the specialised kernels are
// 1.7x faster, keep a square and a carry launch per schedule group.
constexpr uint32_t kBatchFusedMaxN = 262144;

struct BatchGroup {
  Schedule sch{};
  uint32_t n = 0;
  std::vector<uint32_t> members;  // test indices
  PlanArgs *d_args = nullptr;     // sweeps: per-member args
  ResidentDesc *d_desc = nullptr; // resident: per-member descriptors
  std::vector<ResidentDesc> desc;
};

}  // namespace

struct ll_batch {
  int device = 0;
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<ll_ctx *> tests;
  std::vector<BatchGroup> groups;
  std::vector<void *> allocs;
  uint32_t *d_remaining = nullptr;
  uint32_t *d_ring = nullptr;        // per test kBatchRing error slots
  uint32_t *h_ring = nullptr;        // pinned copy of d_ring
  uint32_t *h_remaining = nullptr;   // pinned copy of d_remaining
  std::vector<uint32_t> remaining;   // host copy after the last step
  std::vector<uint64_t> done;        // iterations run per test
  std::vector<float> max_err;        // running maximum per test
  std::vector<std::vector<float>> step_errs;  // per-iteration errors of the last step
  // the single-launch iteration of the two-sweep tests
  std::vector<uint32_t> sweep_tests;   // test indices of the two-sweep tests
  PlanArgs *d_fargs = nullptr;         // their PlanArgs, in sweep_tests order
  unsigned long long *d_sq_done = nullptr, *d_ca_done = nullptr;
  BatchItem *d_items = nullptr;        // capacity: every item of every test
  uint32_t item_count = 0;             // items of the tests still running
  std::vector<bool> in_items;          // per sweep test: has items in d_items
  uint32_t lds = 0;
  hipGraphExec_t graph = nullptr;
  double last_ms = 0.0;
  char msg[256] = {0};
};

static int bfail(ll_batch *b, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b->msg, sizeof(b->msg), fmt, ap);
  va_end(ap);
  return code;
}

#define BCHK(b, expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return bfail((b), LL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static void batch_release(ll_batch *b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (auto *c : b->tests) ll_destroy(c);
  b->tests.clear();
  for (void *p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  if (b->h_ring) (void)hipHostFree(b->h_ring);
  if (b->h_remaining) (void)hipHostFree(b->h_remaining);
  b->h_ring = b->h_remaining = nullptr;
  for (auto &e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
}

template <typename T>
static int balloc(ll_batch *b, T **ptr, size_t count) {
  void *p = nullptr;
  if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) return bfail(b, LL_ERR_ALLOC, "hipMalloc batch");
  b->allocs.push_back(p);
  *ptr = (T *)p;
  return LL_OK;
}

// `iters` iterations of every running two-sweep test on the batch stream:
// one launch per iteration (batch_iteration) whose workgroups are every
// test's square items, then every test's carry items.
static int enqueue_sweeps(ll_batch *b, uint32_t iters) {
  for (uint32_t i = 0; i < iters; i++) {
    if (b->item_count)
      BCHK(b, launch_batch_iteration(b->d_items, b->item_count, b->d_fargs, b->d_sq_done, b->d_ca_done,
                                     b->lds, b->stream));
    // the longer tests: one square and one carry launch per schedule group,
    // on the same stream, so no two carry launches run side by side
    for (auto &grp : b->groups)
      if (grp.sch.path == 1 && !grp.sch.generic) {
        const uint32_t k = (uint32_t)grp.members.size();
        const uint32_t grid = square_grid(grp.sch.sk, grp.sch.W, k);
        BCHK(b, launch_square(grp.sch.sk, grp.d_args, k, grp.sch.W, grp.sch.H, grid, b->stream));
        const CarryLaunch cl = carry_launch(grp.sch.ck, grp.sch.H, grp.sch.R, k);
        BCHK(b, launch_carry(grp.sch.ck, cl, grp.d_args, k, grp.sch.W, grp.sch.H, grp.sch.R, b->stream));
      }
  }
  return LL_OK;
}

// The work items of the two-sweep tests that still run (all square items
// first); rebuilt, with the captured graph dropped, when tests finish.
static int build_items(ll_batch *b) {
  std::vector<BatchItem> items;
  std::vector<bool> in(b->sweep_tests.size(), false);
  for (int carry = 0; carry < 2; carry++)
    for (uint32_t k = 0; k < b->sweep_tests.size(); k++) {
      const uint32_t t = b->sweep_tests[k];
      if (b->remaining[t] == 0) continue;
      in[k] = true;
      const ll_ctx *c = b->tests[t];
      const uint32_t n = carry ? c->sch.H / c->sch.R : c->sch.W / 2;
      for (uint32_t blk = 0; blk < n; blk++) {
        BatchItem it;
        it.test = k;
        it.carry = (uint32_t)carry;
        it.block = blk;
        items.push_back(it);
      }
    }
  if (in == b->in_items && b->item_count == items.size()) return LL_OK;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->in_items = in;
  b->item_count = (uint32_t)items.size();
  if (!items.empty())
    BCHK(b, hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                           b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return LL_OK;
}

static int batch_graph(ll_batch *b) {
  if (b->graph) return LL_OK;
  hipGraph_t g = nullptr;
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_sweeps(b, kGraphIters);
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return bfail(b, LL_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
  hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return bfail(b, LL_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return LL_OK;
}

extern "C" {

int ll_batch_create(int32_t device, uint32_t count, const ll_batch_test *tests, uint32_t flags,
                    ll_batch **out) {
  if (!out || !tests || count == 0) return LL_ERR_ARG;
  ll_batch *b = new ll_batch();
  *out = b;
  b->device = device;
  if (hipSetDevice(device) != hipSuccess) return bfail(b, LL_ERR_HIP, "hipSetDevice(%d)", device);
  BCHK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BCHK(b, hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking));
  for (auto &e : b->ev) BCHK(b, hipEventCreate(&e));
  int rc = balloc(b, &b->d_remaining, count);
  if (!rc) rc = balloc(b, &b->d_ring, (size_t)count * kBatchRing);
  if (rc) return rc;
  if (hipHostMalloc((void **)&b->h_ring, sizeof(uint32_t) * count * kBatchRing, 0) != hipSuccess ||
      hipHostMalloc((void **)&b->h_remaining, sizeof(uint32_t) * count, 0) != hipSuccess)
    return bfail(b, LL_ERR_ALLOC, "hipHostMalloc of the batch readback buffers failed");
  b->remaining.resize(count);
  b->done.assign(count, 0);
  b->max_err.assign(count, 0.f);
  ll_options o;
  memset(&o, 0, sizeof o);
  o.device = device;
  o.flags = flags & LL_OPT_FORCE_SWEEPS;
  generic_kernels_init();
  for (uint32_t i = 0; i < count; i++) {
    ll_ctx *c = nullptr;
    // two-sweep tests run the generic kernels' bodies in the single-launch
    // iteration, at a schedule whose LDS lets several workgroups share a CU
    Schedule pick, gpick;
    memset(&gpick, 0, sizeof gpick);
    ll_options oi = o;
    if (pick_schedule(tests[i].n, &o, &pick) && pick.path == 1 &&
        (pick.generic || tests[i].n <= kBatchFusedMaxN) &&
        pick_generic(tests[i].n / 2, 0, nullptr, &gpick, kBatchLds)) {
      oi.flags |= LL_OPT_GENERIC;
      oi.width = gpick.W;
      oi.height = gpick.H;
      oi.lines = gpick.R;
    }
    rc = ctx_create(tests[i].p, tests[i].n, &oi, kBatchRing, false, &c);
    if (c) b->tests.push_back(c);
    if (rc) return bfail(b, rc, "test %u (p=%u n=%u): %s", i, tests[i].p, tests[i].n, c ? c->msg : "");
    std::vector<int32_t> d(tests[i].n);
    ll_initial_digits(tests[i].p, tests[i].n, tests[i].offset % tests[i].p, d.data());
    if ((rc = ll_load_digits(c, d.data(), tests[i].offset))) return bfail(b, rc, "load %u: %s", i, c->msg);
    b->remaining[i] = tests[i].iterations;
    // group by schedule
    BatchGroup *grp = nullptr;
    for (auto &g : b->groups)
      if (g.n == c->n && g.sch.path == c->sch.path && g.sch.W == c->sch.W && g.sch.H == c->sch.H &&
          g.sch.R == c->sch.R && g.sch.rk == c->sch.rk && g.sch.generic == c->sch.generic)
        grp = &g;
    if (!grp) {
      b->groups.emplace_back();
      grp = &b->groups.back();
      grp->sch = c->sch;
      grp->n = c->n;
    }
    grp->members.push_back(i);
  }
  BCHK(b, hipMemcpyAsync(b->d_remaining, b->remaining.data(), 4 * count, hipMemcpyHostToDevice, b->stream));
  for (auto &g : b->groups) {
    const size_t k = g.members.size();
    if (g.sch.path == 1) {
      std::vector<PlanArgs> a(k);
      for (size_t t = 0; t < k; t++) {
        a[t] = b->tests[g.members[t]]->args;
        a[t].remaining = b->d_remaining + g.members[t];
        a[t].err = b->d_ring + (size_t)g.members[t] * kBatchRing;
        a[t].err_mask = kBatchRing - 1;
        a[t].stagger_sq = a[t].stagger_ca = 0;  // several tests share the launch
      }
      if ((rc = balloc(b, &g.d_args, k))) return rc;
      BCHK(b, hipMemcpyAsync(g.d_args, a.data(), k * sizeof(PlanArgs), hipMemcpyHostToDevice, b->stream));
      BCHK(b, hipStreamSynchronize(b->stream));
    } else {
      g.desc.resize(k);
      for (size_t t = 0; t < k; t++) {
        g.desc[t] = b->tests[g.members[t]]->rdesc;
        g.desc[t].remaining = b->d_remaining + g.members[t];
        g.desc[t].err = (float *)(b->d_ring + (size_t)g.members[t] * kBatchRing);
      }
      if ((rc = balloc(b, &g.d_desc, k))) return rc;
    }
  }
  // the single-launch iteration: PlanArgs of every two-sweep test, its
  // counters, room for all items
  std::vector<PlanArgs> fa;
  size_t items = 0;
  for (uint32_t i = 0; i < count; i++) {
    const ll_ctx *c = b->tests[i];
    if (c->sch.path != 1 || !c->sch.generic) continue;
    PlanArgs a = c->args;
    a.remaining = b->d_remaining + i;
    a.err = b->d_ring + (size_t)i * kBatchRing;
    a.err_mask = kBatchRing - 1;
    fa.push_back(a);
    b->sweep_tests.push_back(i);
    items += c->sch.W / 2 + c->sch.H / c->sch.R;
    b->lds = std::max(b->lds, std::max(square_generic_lds(c->sch.H), carry_generic_lds(c->sch.W, c->sch.R)));
  }
  if (!fa.empty()) {
    if ((rc = balloc(b, &b->d_fargs, fa.size())) || (rc = balloc(b, &b->d_sq_done, fa.size())) ||
        (rc = balloc(b, &b->d_ca_done, fa.size())) || (rc = balloc(b, &b->d_items, items)))
      return rc;
    BCHK(b, hipMemcpyAsync(b->d_fargs, fa.data(), fa.size() * sizeof(PlanArgs), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemsetAsync(b->d_sq_done, 0, fa.size() * sizeof(unsigned long long), b->stream));
    BCHK(b, hipMemsetAsync(b->d_ca_done, 0, fa.size() * sizeof(unsigned long long), b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  if ((rc = build_items(b))) return rc;
  return LL_OK;
}

void ll_batch_destroy(ll_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  batch_release(b);
  delete b;
}

const char *ll_batch_last_error(const ll_batch *b) { return b ? b->msg : "null batch"; }

int ll_batch_step(ll_batch *b, uint32_t iterations, float *step_max_err) {
  if (!b) return LL_ERR_ARG;
  if (iterations > kBatchRing) return bfail(b, LL_ERR_ARG, "at most %u iterations per step", kBatchRing);
  BCHK(b, hipSetDevice(b->device));
  const uint32_t count = (uint32_t)b->tests.size();
  std::vector<uint32_t> before = b->remaining;
  GpuLock lock(b->device);
  // (error slots need no clearing: the resident kernel writes each of its
  // iterations' slots, and the square sweep clears its iteration's slot)
  BCHK(b, hipEventRecord(b->ev[0], b->stream));
  BCHK(b, hipStreamWaitEvent(b->stream2, b->ev[0], 0));
  // LDS-resident groups: one launch for the whole step, on the second stream
  for (auto &g : b->groups) {
    if (g.sch.path != 0) continue;
    for (auto &d : g.desc) d.iters = iterations;
    BCHK(b, hipMemcpyAsync(g.d_desc, g.desc.data(), g.desc.size() * sizeof(ResidentDesc),
                           hipMemcpyHostToDevice, b->stream2));
    BCHK(b, launch_resident(g.sch.rk, g.d_desc, (uint32_t)g.desc.size(), b->stream2));
  }
  // two-sweep tests: one launch per iteration (and the longer tests' group
  // launches), graph-replayed
  {
    int rc = build_items(b);
    if (rc) return rc;
  }
  bool any_sweeps = b->item_count > 0;
  for (auto &g : b->groups) any_sweeps |= g.sch.path == 1 && !g.sch.generic;
  if (any_sweeps) {
    uint32_t left = iterations;
    if (left >= kGraphIters) {
      int rc = batch_graph(b);
      if (rc) return rc;
      for (; left >= kGraphIters; left -= kGraphIters) BCHK(b, hipGraphLaunch(b->graph, b->stream));
    }
    if (left) {
      int rc = enqueue_sweeps(b, left);
      if (rc) return rc;
    }
  }
  BCHK(b, hipEventRecord(b->ev[1], b->stream2));
  BCHK(b, hipStreamWaitEvent(b->stream, b->ev[1], 0));
  BCHK(b, hipEventRecord(b->ev[2], b->stream));
  BCHK(b, hipMemcpyAsync(b->h_remaining, b->d_remaining, 4 * count, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipMemcpyAsync(b->h_ring, b->d_ring, sizeof(uint32_t) * count * kBatchRing, hipMemcpyDeviceToHost,
                         b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  memcpy(b->remaining.data(), b->h_remaining, 4 * count);
  float ms = 0.f;
  BCHK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[2]));
  b->last_ms = ms;
  b->step_errs.resize(count);
  for (uint32_t i = 0; i < count; i++) {
    ll_ctx *c = b->tests[i];
    const uint32_t ran = before[i] - b->remaining[i];
    float m = 0.f;
    b->step_errs[i].resize(ran);
    const uint32_t *ring = b->h_ring + (size_t)i * kBatchRing;
    for (uint32_t k = 0; k < ran; k++) {
      const uint32_t bits = c->sch.path == 0 ? ring[k] : ring[(c->epoch + k) & (kBatchRing - 1)];
      float e;
      memcpy(&e, &bits, 4);
      b->step_errs[i][k] = e;
      m = std::max(m, e);
    }
    if (c->sch.path == 1) c->epoch += ran;
    c->iterations += ran;
    b->done[i] += ran;
    b->max_err[i] = std::max(b->max_err[i], m);
    if (step_max_err) step_max_err[i] = m;
  }
  return LL_OK;
}

double ll_batch_last_step_ms(const ll_batch *b) { return b ? b->last_ms : -1.0; }

int ll_batch_status(const ll_batch *b, uint32_t i, uint64_t *iterations_done, uint32_t *remaining,
                    float *max_err) {
  if (!b || i >= b->tests.size()) return LL_ERR_ARG;
  if (iterations_done) *iterations_done = b->done[i];
  if (remaining) *remaining = b->remaining[i];
  if (max_err) *max_err = b->max_err[i];
  return LL_OK;
}

int ll_batch_read_digits(ll_batch *b, uint32_t i, int32_t *digits, uint32_t *offset) {
  if (!b || i >= b->tests.size() || !digits) return LL_ERR_ARG;
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipStreamSynchronize(b->stream2));
  const int rc = ll_read_digits(b->tests[i], digits, offset);
  if (rc) return bfail(b, rc, "read %u: %s", i, b->tests[i]->msg);
  return LL_OK;
}

int ll_batch_groups(const ll_batch *b) { return b ? (int)b->groups.size() : -1; }

int ll_batch_step_errors(const ll_batch *b, uint32_t i, float *errs, uint32_t capacity) {
  if (!b || i >= b->tests.size()) return LL_ERR_ARG;
  if (i >= b->step_errs.size()) return 0;
  const auto &e = b->step_errs[i];
  if (errs && capacity < e.size()) return LL_ERR_ARG;
  if (errs) memcpy(errs, e.data(), e.size() * sizeof(float));
  return (int)e.size();
}

int ll_fft_schedules(uint32_t n, ll_schedule *out, uint32_t capacity) {
  if (n < 512 || n % 2) return 0;
  const uint32_t M = n / 2;
  std::vector<ll_schedule> v;
  if (find_resident(M)) v.push_back({n, 0, M, 1, 1, 0.f});
  size_t count = 0;
  const CarryKernel *tab = carry_table(&count);
  for (size_t i = 0; i < count; i++) {
    const CarryKernel *ck = &tab[i];
    if (M % ck->w) continue;
    const uint32_t H = M / ck->w;
    if (!find_square(H) || H % ck->r) continue;
    bool dup = false;
    for (const auto &e : v) dup |= e.path == 1 && e.width == ck->w && e.lines == ck->r;
    if (!dup) v.push_back({n, 1, ck->w, H, ck->r, 0.f});
  }
  if (v.empty()) {  // lengths only the generic kernels run
    Schedule g;
    memset(&g, 0, sizeof g);
    if (pick_generic(M, 0, nullptr, &g)) v.push_back({n, 1, g.W, g.H, g.R, 0.f});
  }
  if (out) {
    if (capacity < v.size()) return LL_ERR_ARG;
    std::copy(v.begin(), v.end(), out);
  }
  return (int)v.size();
}

int ll_tune_length(int32_t device, uint32_t p, uint32_t n, uint32_t iters, ll_schedule *out, uint32_t capacity) {
  const int k = ll_fft_schedules(n, nullptr, 0);
  if (k <= 0) return LL_ERR_UNSUPPORTED;
  if (!out || capacity < (uint32_t)k || iters == 0) return LL_ERR_ARG;
  ll_fft_schedules(n, out, capacity);
  if (p == 0) p = 17u * n + 1u;  // ~17 bits per word: timing does not depend on p
  std::vector<int32_t> digits(n);
  ll_initial_digits(p, n, 1, digits.data());
  for (int i = 0; i < k; i++) {
    ll_options o;
    memset(&o, 0, sizeof o);
    o.device = device;
    if (out[i].path == 1) {
      o.flags = LL_OPT_FORCE_SWEEPS;
      o.width = out[i].width;
      o.height = out[i].height;
      o.lines = out[i].lines;
    }
    ll_ctx *c = nullptr;
    int rc = ll_create(p, n, &o, &c);
    if (!rc) rc = ll_load_digits(c, digits.data(), 1);
    if (!rc) rc = ll_iterate(c, std::min(iters, 64u), nullptr, nullptr);  // warm-up, graph build
    if (!rc) rc = ll_iterate(c, iters, nullptr, nullptr);
    out[i].ms = rc ? -1.f : (float)(ll_last_iterate_ms(c) / iters);
    ll_destroy(c);
    if (rc == LL_ERR_HIP) return rc;
  }
  std::stable_sort(out, out + k, [](const ll_schedule &a, const ll_schedule &b) {
    return (a.ms >= 0 ? a.ms : 1e30f) < (b.ms >= 0 ? b.ms : 1e30f);
  });
  return k;
}

int ll_schedule_table_set(uint32_t count, const ll_schedule *table) {
  if (count && !table) return LL_ERR_ARG;
  schedule_table().assign(table, table + count);
  return LL_OK;
}

int ll_get_device_info(int32_t device, ll_device_info *out) {
  if (!out) return LL_ERR_ARG;
  memset(out, 0, sizeof(*out));
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return LL_ERR_ARG;
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, device) != hipSuccess) return LL_ERR_HIP;
  snprintf(out->name, sizeof out->name, "%s", pr.name);
  snprintf(out->arch, sizeof out->arch, "%.63s", pr.gcnArchName);
  out->compute_units = pr.multiProcessorCount;
  out->clock_mhz = pr.clockRate / 1000;
  out->mem_clock_mhz = pr.memoryClockRate / 1000;
  out->mem_bus_bits = pr.memoryBusWidth;
  out->global_mem = pr.totalGlobalMem;
  out->l2_bytes = pr.l2CacheSize;
  out->lds_per_block = (uint32_t)pr.sharedMemPerBlock;
  out->max_threads_per_block = pr.maxThreadsPerBlock;
  out->wavefront = pr.warpSize;
  out->pci_bus = pr.pciBusID;
  out->pci_device = pr.pciDeviceID;
  int v = 0;
  if (hipRuntimeGetVersion(&v) == hipSuccess) out->runtime_version = v;
  if (hipDriverGetVersion(&v) == hipSuccess) out->driver_version = v;
  return LL_OK;
}

int ll_fft_schedule(uint32_t n, const ll_options *opts, ll_info *info) {
  Schedule s;
  if (!info) return LL_ERR_ARG;
  if (!pick_schedule(n, opts, &s)) return LL_ERR_UNSUPPORTED;
  memset(info, 0, sizeof(*info));
  info->n = n;
  info->path = s.generic ? 2 : s.path;
  if (s.path == 0) {
    info->width = n / 2;
    info->height = 1;
    info->lines = 1;
  } else {
    info->width = s.W;
    info->height = s.H;
    info->lines = s.R;
  }
  return LL_OK;
}

}  // extern "C"

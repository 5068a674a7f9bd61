// Engine contexts behind the C-ABI (include/ll_api.h): schedule choice,
// device tables and buffers, the iteration loop with hipGraph replay,
// load/unload of digits, per-iteration error readback.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ll_api.h"
#include "ll_host.h"
#include "ll_kernels.h"

using namespace llamd;

namespace {

constexpr uint32_t kRingSize = 1u << 16;   // per-iteration error slots
constexpr uint32_t kGraphIters = 32;       // iterations per captured graph
constexpr uint32_t kTwoLevelShift = 11;    // 2048-entry fine root table

struct Schedule {
  uint32_t path;  // 0 resident, 1 sweeps
  uint32_t W, H, R;
  const ResidentKernel *rk;
  const SquareKernel *sk;
  const CarryKernel *ck;
};

bool pick_schedule(uint32_t n, const ll_options *o, Schedule *s) {
  memset(s, 0, sizeof(*s));
  if (n < 512 || n % 2) return false;
  const uint32_t M = n / 2;
  const uint32_t flags = o ? o->flags : 0;
  if (!(flags & LL_OPT_FORCE_SWEEPS) && !(o && (o->width || o->height))) {
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
    // prefer H = 2048, then R = 8 (128-byte column runs), then W close to H / 2
    double cost = fabs(log2((double)H) - 11.0) * 4.0;
    cost += ck->r == 8 ? 0.0 : ck->r == 4 ? 1.0 : 2.0;
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
  return s->ck != nullptr;
}

}  // namespace

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
  uint32_t *ring = nullptr;
  DevState *st = nullptr;
  hipGraphExec_t graph = nullptr;
  // resident
  ResidentDesc rdesc{};
  ResidentDesc *d_rdesc = nullptr;
  float *r_err = nullptr;
  uint32_t *r_offset = nullptr;
  std::vector<uint32_t> host_ring;
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
  HIPCHK(c, hipMemcpy(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return LL_OK;
}

static int build_sweeps(ll_ctx *c) {
  const uint32_t p = c->p, n = c->n, M = c->m, W = c->sch.W, H = c->sch.H, R = c->sch.R;
  PlanArgs &A = c->args;
  std::vector<double> qc;
  int rc;
  quarter_cosines(W, qc);
  if ((rc = upload(c, (double **)&A.qcW, qc))) return rc;
  quarter_cosines(H, qc);
  if ((rc = upload(c, (double **)&A.qcH, qc))) return rc;
  // two-level roots of order M
  const uint32_t shift = std::min<uint32_t>(kTwoLevelShift, (uint32_t)log2((double)M));
  const uint32_t S = 1u << shift;
  const uint32_t nhi = (M + S - 1) / S;
  std::vector<double> hi(2 * (size_t)nhi), lo(2 * (size_t)S);
  for (uint32_t i = 0; i < nhi; i++) root_of_unity((uint64_t)i * S, M, &hi[2 * i], &hi[2 * i + 1]);
  for (uint32_t i = 0; i < S; i++) root_minus_one(i, M, &lo[2 * i], &lo[2 * i + 1]);
  if ((rc = upload(c, (double **)&A.twM.hi, hi))) return rc;
  if ((rc = upload(c, (double **)&A.twM.lo, lo))) return rc;
  A.twM.shift = shift;
  A.twM.mask = S - 1;
  // factored weights: word 2H a + i has f = fA[a] + fB[i] (mod n)
  std::vector<double> wA(W), iwA(W), wB(2 * H), iwB(2 * H);
  std::vector<int32_t> fA(W), fB(2 * H);
  for (uint32_t a = 0; a < W; a++) {
    fA[a] = (int32_t)weight_exponent(p, n, (uint64_t)2 * H * a);
    weight_pair((uint32_t)fA[a], n, &wA[a], &iwA[a]);
  }
  for (uint32_t i = 0; i < 2 * H; i++) {
    fB[i] = (int32_t)weight_exponent(p, n, i);
    weight_pair((uint32_t)fB[i], n, &wB[i], &iwB[i]);
  }
  if ((rc = upload(c, (double **)&A.wA, wA))) return rc;
  if ((rc = upload(c, (double **)&A.iwA, iwA))) return rc;
  if ((rc = upload(c, (int32_t **)&A.fA, fA))) return rc;
  if ((rc = upload(c, (double **)&A.wB, wB))) return rc;
  if ((rc = upload(c, (double **)&A.iwB, iwB))) return rc;
  if ((rc = upload(c, (int32_t **)&A.fB, fB))) return rc;
  const size_t shuttle = (size_t)(H / R) * W;
  if ((rc = dalloc(c, &A.shuttle, shuttle))) return rc;
  HIPCHK(c, hipMemset(A.shuttle, 0xff, shuttle * sizeof(uint64_t)));
  if ((rc = dalloc(c, &c->st, 1))) return rc;
  HIPCHK(c, hipMemset(c->st, 0, sizeof(DevState)));
  if ((rc = dalloc(c, &c->ring, kRingSize))) return rc;
  A.x = c->x;
  A.st = c->st;
  A.err = c->ring;
  A.err_mask = kRingSize - 1;
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
  c->host_ring.resize(kRingSize);
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
  if ((rc = upload(c, (double **)&D.qc, qc))) return rc;
  if ((rc = dalloc(c, &c->r_err, kRingSize))) return rc;
  if ((rc = dalloc(c, &c->r_offset, 1))) return rc;
  if ((rc = dalloc(c, &c->d_rdesc, 1))) return rc;
  D.x = c->x;
  D.err = c->r_err;
  D.offset_out = c->r_offset;
  D.scale = 0.25 / (double)M;
  D.p = p;
  D.m = M;
  D.offset = 0;
  D.iters = 0;
  c->host_ring.resize(kRingSize);
  return LL_OK;
}

static void release(ll_ctx *c) {
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  for (void *p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

static int build_graph(ll_ctx *c) {
  if (c->graph) return LL_OK;
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (uint32_t i = 0; i < kGraphIters; i++) {
    (void)launch_square(c->sch.sk, c->args, c->stream);
    (void)launch_carry(c->sch.ck, c->args, c->stream);
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

int ll_create(uint32_t p, uint32_t n, const ll_options *opts, ll_ctx **out) {
  if (!out) return LL_ERR_ARG;
  *out = nullptr;
  if (p < 3 || n < 512 || n % 2 || (uint64_t)p < 4ull * n / 3) return LL_ERR_ARG;
  if ((double)p / n > 31.0) return LL_ERR_ARG;  // digits must fit int32 with headroom
  if (p / n < 2) return LL_ERR_ARG;              // digits of at least 2 bits
  ll_ctx *c = new ll_ctx();
  c->p = p;
  c->n = n;
  c->m = n / 2;
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
  if (!rc) rc = dalloc(c, &c->x, c->m);
  if (!rc) rc = dalloc(c, &c->digits, n);
  if (!rc) rc = dalloc(c, &c->unload_err, 1);
  if (!rc) rc = c->sch.path == 0 ? build_resident(c) : build_sweeps(c);
  *out = c;
  return rc;
}

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
  if (c->sch.path == 0) {
    c->rdesc.offset = offset;
    c->rdesc.iters = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &c->rdesc, sizeof(ResidentDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_resident_load(c->d_rdesc, 1, c->digits, c->stream));
  } else {
    HIPCHK(c, launch_load(c->sch.ck, c->args, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->st->offset, &offset, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->offset = offset;
  c->loaded = true;
  c->iterations = 0;
  return LL_OK;
}

static int iterate_resident(ll_ctx *c, uint32_t count, float *errs) {
  ResidentDesc D = c->rdesc;
  D.offset = c->offset;
  D.iters = count;
  HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &D, sizeof(ResidentDesc), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, launch_resident(c->sch.rk, c->d_rdesc, 1, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipMemcpyAsync(errs, c->r_err, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream));
  uint32_t off = 0;
  HIPCHK(c, hipMemcpyAsync(&off, c->r_offset, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->last_gpu_ms += ms;
  c->offset = off;
  return LL_OK;
}

static int iterate_sweeps(ll_ctx *c, uint32_t count, float *errs) {
  HIPCHK(c, launch_fill(c->ring, 0u, kRingSize, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  uint32_t left = count;
  if (c->use_graph && count >= kGraphIters) {
    int rc = build_graph(c);
    if (rc) return rc;
    while (left >= kGraphIters) {
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      left -= kGraphIters;
    }
  }
  for (; left; left--) {
    HIPCHK(c, launch_square(c->sch.sk, c->args, c->stream));
    HIPCHK(c, launch_carry(c->sch.ck, c->args, c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipMemcpyAsync(c->host_ring.data(), c->ring, sizeof(uint32_t) * kRingSize,
                           hipMemcpyDeviceToHost, c->stream));
  uint32_t off = 0;
  HIPCHK(c, hipMemcpyAsync(&off, &c->st->offset, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  c->last_gpu_ms += ms;
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t bits = c->host_ring[(c->epoch + i) & (kRingSize - 1)];
    memcpy(&errs[i], &bits, 4);
  }
  c->epoch += count;
  c->offset = off;
  return LL_OK;
}

int ll_iterate(ll_ctx *c, uint32_t count, float *per_iter, float *max_err) {
  if (!c) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_iterate before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->inject_kind == 2) {
    c->inject_kind = 0;
    return fail(c, LL_ERR_HIP, "injected device error");
  }
  std::vector<float> errs(std::min<uint32_t>(count, kRingSize));
  c->last_gpu_ms = 0.0;
  float m = 0.f;
  uint32_t done = 0;
  while (done < count) {
    const uint32_t chunk = std::min<uint32_t>(count - done, kRingSize);
    int rc = c->sch.path == 0 ? iterate_resident(c, chunk, errs.data()) : iterate_sweeps(c, chunk, errs.data());
    if (rc) return rc;
    if (c->inject_kind == 1) {
      if (c->inject_after < chunk) {
        errs[c->inject_after] = c->inject_value;
        c->inject_kind = 0;
      } else {
        c->inject_after -= chunk;
      }
    }
    for (uint32_t i = 0; i < chunk; i++) m = std::max(m, errs[i]);
    if (per_iter) memcpy(per_iter + done, errs.data(), sizeof(float) * chunk);
    done += chunk;
    c->iterations += chunk;
  }
  c->max_err = std::max(c->max_err, m);
  if (max_err) *max_err = m;
  return LL_OK;
}

int ll_read_digits(ll_ctx *c, int32_t *digits, uint32_t *offset) {
  if (!c || !digits) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_read_digits before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_fill(c->unload_err, 0u, 1, c->stream));
  if (c->sch.path == 0) {
    ResidentDesc D = c->rdesc;
    D.offset = c->offset;
    D.iters = 0;
    HIPCHK(c, hipMemcpyAsync(c->d_rdesc, &D, sizeof(ResidentDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_resident_unload(c->d_rdesc, 1, c->digits, c->unload_err, c->stream));
  } else {
    PlanArgs A = c->args;
    A.err = c->unload_err;
    HIPCHK(c, launch_unload(c->sch.ck, A, c->stream));
  }
  uint32_t ebits = 0;
  HIPCHK(c, hipMemcpyAsync(digits, c->digits, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&ebits, c->unload_err, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float e;
  memcpy(&e, &ebits, 4);
  if (!(e < 0.05f)) return fail(c, LL_ERR_ROUNDOFF, "digit readback rounding error %g", (double)e);
  if (offset) *offset = c->offset;
  return LL_OK;
}

double ll_last_iterate_ms(const ll_ctx *c) { return c ? c->last_gpu_ms : -1.0; }

int ll_profile_kernels(ll_ctx *c, uint32_t count, float *kernel_ms) {
  if (!c || !kernel_ms || count == 0) return LL_ERR_ARG;
  if (!c->loaded) return fail(c, LL_ERR_STATE, "ll_profile_kernels before ll_load_digits");
  HIPCHK(c, hipSetDevice(c->device));
  kernel_ms[0] = kernel_ms[1] = 0.f;
  if (c->sch.path == 0) {
    std::vector<float> errs(std::min<uint32_t>(count, kRingSize));
    for (uint32_t done = 0; done < count;) {
      const uint32_t chunk = std::min<uint32_t>(count - done, kRingSize);
      c->last_gpu_ms = 0.0;
      int rc = iterate_resident(c, chunk, errs.data());
      if (rc) return rc;
      kernel_ms[0] += (float)c->last_gpu_ms;
      done += chunk;
    }
    kernel_ms[0] /= (float)count;  // per iteration inside the resident kernel
    c->iterations += count;
    return LL_OK;
  }
  HIPCHK(c, launch_fill(c->ring, 0u, kRingSize, c->stream));
  double sq = 0.0, ca = 0.0;
  for (uint32_t i = 0; i < count; i++) {
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, launch_square(c->sch.sk, c->args, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, launch_carry(c->sch.ck, c->args, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[2]));
    float a = 0.f, b = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    sq += a;
    ca += b;
  }
  uint32_t off = 0;
  HIPCHK(c, hipMemcpy(&off, &c->st->offset, sizeof(uint32_t), hipMemcpyDeviceToHost));
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
  info->path = c->sch.path;
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

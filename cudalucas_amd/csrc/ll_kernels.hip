// HIP kernels of the Lucas-Lehmer iteration s <- s^2 - 2 mod 2^p - 1 on gfx950.
//
// One iteration of the reference (CUDALucas.cu:1869-1986: cuFFT INVERSE,
// square, cuFFT INVERSE, rcb, splice) is rebuilt in two shapes:
//
//  * resident_kernel: for N <= 16K words the whole M = N/2-point complex
//    signal lives in one workgroup's LDS and the kernel loops over many
//    iterations without touching HBM (batched: one workgroup per exponent).
//
//  * two sweeps over HBM for large N (ll_sweeps.h): square_sweep over
//    column pairs, carry_sweep over blocks of lines, carries handed between
//    blocks through an epoch-tagged shuttle.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <tuple>
#include <vector>

#include "ll_sweeps.h"

namespace llamd {

// ----------------------------------------------------------- resident path
// One workgroup per exponent; the signal (natural order) stays in LDS for
// `iters` iterations.  Thread t owns words [t*C, (t+1)*C), C = N/T.
template <int M, int T>
__global__ __launch_bounds__(T) void resident_kernel(const ResidentDesc *descs) {
  constexpr int LP = lpitch(M);
  constexpr int N = 2 * M;
  constexpr int C = N / T;
  static_assert(C % 2 == 0 && C >= 2, "words per thread");
  __shared__ cplx lds[LP];
  __shared__ double qc[M / 4 + 1];
  __shared__ int64_t s_carry[T];
  __shared__ float red[T / 64];
  __shared__ Minus2 s_m2;
  // per-word weight factors shared by every thread (word t C + i)
  __shared__ double s_wC[C], s_iwC[C];
  __shared__ int s_fC[C];
  const ResidentDesc D = descs[blockIdx.x];
  const uint32_t left = D.remaining ? *D.remaining : D.iters;
  const uint32_t iters = D.iters < left ? D.iters : left;
  if (iters == 0) return;
  const int tid = threadIdx.x;
  load_quarter<M, T>(qc, D.qc);
  const QuarterRoots<M> tw{qc};
  for (int i = tid; i < M; i += T) lds[lpad(i)] = D.x[i];
  for (int i = tid; i < C; i += T) {
    s_wC[i] = D.wC[i];
    s_iwC[i] = D.iwC[i];
    s_fC[i] = D.fC[i];
  }
  // this thread's factor, kept in registers for every iteration (the
  // weights are never read from global memory inside the loop)
  const double wT = D.wT[tid], iwTs = D.iwTs[tid];
  const int fT = D.fT[tid];
  const uint32_t p = D.p, b = p % N, q = p / N;
  const uint32_t j0 = (uint32_t)tid * C;
  uint32_t offset = *D.offset;
  // Up to 4096 points (registers to spare) each thread reads its Stockham
  // twiddles once from the plan's per-pass table and keeps them in
  // registers for every transform of the launch (the inverse uses their
  // conjugates); longer transforms take them from the quarter-wave table
  // in LDS, which costs them no registers.
  constexpr bool CT = T <= 256 || M <= 4096;
  PassTwiddles<M, 1, T> ptw;
  if constexpr (CT) ptw.init_table(D.pt);
  __syncthreads();
  for (uint32_t it = 0; it < iters; it++) {
    if (tid == 0) s_m2 = minus2_position(offset, p, N);
    if constexpr (CT) fft_lds_cached<M, 1, T, false>(lds, ptw);
    else fft_lds<M, 1, T, false>(lds, tw);
    for (int k = tid; k <= M / 2; k += T) {
      const int kb = (M - k) % M;
      cplx a = lds[lpad(k)], bb = lds[lpad(kb)];
      pair_square(a, bb, tw(k));
      lds[lpad(k)] = a;
      if (kb != k) lds[lpad(kb)] = bb;
    }
    __syncthreads();
    if constexpr (CT) fft_lds_cached<M, 1, T, true>(lds, ptw);
    else fft_lds<M, 1, T, true>(lds, tw);
    const Minus2 m2 = s_m2;
    // bit i: word i's weight exponent sum wrapped / word i is a big digit
    uint64_t wraps = 0, bigs = 0;
    static_assert(C <= 64, "per-word flags in one word");
    float emax = 0.f;
    int64_t carry = 0;
    // round, -2 and carry along this thread's C words; the balanced digits
    // go back into the thread's own LDS slots (as exact doubles), so no
    // per-word register array lives across the barrier below
#pragma unroll
    for (int w = 0; w < C; w += 2) {
      const cplx v = lds[lpad((j0 + w) / 2)];
      double dg[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const uint32_t j = j0 + w + e;
        const int fs = fT + s_fC[w + e];
        const int wrap = fs >= (int)N;
        const int f = fs - (wrap ? (int)N : 0);
        wraps |= (uint64_t)wrap << (w + e);
        const double t = ldexp(mul3(e ? v.y : v.x, iwTs, s_iwC[w + e]), wrap);
        const double rt = rint(t);
        emax = fmaxf(emax, (float)fabs(t - rt));
        int64_t iv = (int64_t)rt;
        if (j == m2.digit) iv -= m2.value;
        const int big = f < (int)b;
        bigs |= (uint64_t)big << (w + e);
        dg[e] = (double)balance(iv + carry, (int)q + big, carry);
      }
      lds[lpad((j0 + w) / 2)] = {dg[0], dg[1]};
    }
    s_carry[tid] = carry;
    __syncthreads();
    {
      // carry-in from the previous thread's words, propagated through this
      // thread's whole chain (see carry_sweep); the remainder stays in the
      // last word; then reweight
      int64_t cc = s_carry[(tid + T - 1) % T];
#pragma unroll
      for (int w = 0; w < C; w += 2) {
        const cplx dv = lds[lpad((j0 + w) / 2)];
        const int w0 = (int)q + (int)((bigs >> w) & 1), w1 = (int)q + (int)((bigs >> (w + 1)) & 1);
        const int64_t d0 = balance((int64_t)dv.x + cc, w0, cc);
        int64_t d1 = balance((int64_t)dv.y + cc, w1, cc);
        if (w + 2 == C) d1 += cc << w1;
        const double x0 = ldexp(mul3((double)d0, wT, s_wC[w]), -(int)((wraps >> w) & 1));
        const double x1 = ldexp(mul3((double)d1, wT, s_wC[w + 1]), -(int)((wraps >> (w + 1)) & 1));
        lds[lpad((j0 + w) / 2)] = {x0, x1};
      }
    }
    const float m = block_max<T>(emax, red);
    if (tid == 0) D.err[it] = m;
    offset = m2.new_offset;
    __syncthreads();
  }
  for (int i = tid; i < M; i += T) D.x[i] = lds[lpad(i)];
  if (tid == 0) {
    *D.offset = offset;
    if (D.remaining) *D.remaining = left - iters;
  }
}

// digits <- rint(x / w) for the resident layout
__global__ void resident_unload(const ResidentDesc *descs, int32_t *digits, uint32_t *err) {
  const ResidentDesc D = descs[blockIdx.y];
  const uint32_t N = 2 * D.m;
  float emax = 0.f;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
    const cplx v = D.x[j / 2];
    const double t = (j & 1 ? v.y : v.x) * D.iw[j];
    const double r = rint(t);
    emax = fmaxf(emax, (float)fabs(t - r));
    digits[(size_t)blockIdx.y * N + j] = (int32_t)r;
  }
  emax = wave_max(emax);
  if ((threadIdx.x & 63) == 0) atomicMax(err, __float_as_uint(emax));
}

__global__ void resident_load(const ResidentDesc *descs, const int32_t *digits) {
  const ResidentDesc D = descs[blockIdx.y];
  const uint32_t M = D.m;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
    const size_t base = (size_t)blockIdx.y * 2 * M;
    D.x[i] = {(double)digits[base + 2 * i] * D.w[2 * i],
              (double)digits[base + 2 * i + 1] * D.w[2 * i + 1]};
  }
}

// Plain streaming copy (one read + one write per 16 bytes): the calibration
// kernel for HBM byte counters and achievable bandwidth.
__global__ __launch_bounds__(256) void copy_stream(const cplx *__restrict__ in,
                                                   cplx *__restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

__global__ void fill_u32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Per-pass twiddles of the H-point column plan (the square sweep's
// PassTwiddles::init_table layout), built once per plan from the
// quarter-wave table, so every twiddle equals the one the quarter table
// gives.
template <int H>
__global__ __launch_bounds__(256) void column_pass_table(const double *qc, cplx *out) {
  build_wave_table<H, 256, false>(out, QuarterRoots<H>{qc});
}

// ------------------------------------------------------- static scratch
// The no-allocation iteration path's shuttle and error slots.  In host code
// the name `scratch` is only a stand-in: kernels are given the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void **ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {nullptr};
  if (!ptr || device < 0 || device >= kMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

// ------------------------------------------------------------- dispatch

#define LL_SQUARE(H) {H, (const void *)square_sweep<H, 256>, 256, (const void *)column_pass_table<H>, nullptr}
#define LL_SQUARE_T(H, T) {H, (const void *)square_sweep<H, T>, T, (const void *)column_pass_table<H>, nullptr}
#define LL_SQUARE_PIPE(H, T)                                                                           \
  {H, (const void *)square_sweep<H, T>, T, (const void *)column_pass_table<H>,                         \
   (const void *)square_sweep_pipe<H, T>}
#define LL_RESIDENT(M, T) {M, (const void *)resident_kernel<M, T>, T, (const void *)column_pass_table<M>}

static const SquareKernel k_square[] = {
    LL_SQUARE_T(128, 64), LL_SQUARE(256), LL_SQUARE(512), LL_SQUARE(1024), LL_SQUARE(2048), LL_SQUARE_PIPE(4096, 512),
};

// Carry-sweep instantiations live in ll_carry_*.hip (compiled in parallel).
extern const CarryKernel k_carry_large[];
extern const size_t k_carry_large_count;
extern const CarryKernel k_carry_small[];
extern const size_t k_carry_small_count;
extern const CarryKernel k_carry_mixed[];
extern const size_t k_carry_mixed_count;
extern const CarryKernel k_carry_wide[];
extern const size_t k_carry_wide_count;
extern const CarryKernel k_carry_huge[];
extern const size_t k_carry_huge_count;
extern const CarryKernel k_carry_huge1[];
extern const size_t k_carry_huge1_count;

static const ResidentKernel k_resident[] = {
    LL_RESIDENT(256, 64),   LL_RESIDENT(512, 128),  LL_RESIDENT(1024, 512),
    LL_RESIDENT(2048, 512), LL_RESIDENT(4096, 512), LL_RESIDENT(8192, 512),
    LL_RESIDENT(1536, 512), LL_RESIDENT(2560, 256), LL_RESIDENT(3584, 256),
    LL_RESIDENT(3072, 256), LL_RESIDENT(5120, 512), LL_RESIDENT(6144, 512),
    LL_RESIDENT(7168, 512),
};

const SquareKernel *find_square(uint32_t H) {
  for (const auto &k : k_square)
    if (k.h == H) return &k;
  return nullptr;
}

static const std::vector<CarryKernel> &all_carry() {
  static const std::vector<CarryKernel> v = [] {
    std::vector<CarryKernel> t(k_carry_large, k_carry_large + k_carry_large_count);
    t.insert(t.end(), k_carry_small, k_carry_small + k_carry_small_count);
    t.insert(t.end(), k_carry_mixed, k_carry_mixed + k_carry_mixed_count);
    t.insert(t.end(), k_carry_wide, k_carry_wide + k_carry_wide_count);
    t.insert(t.end(), k_carry_huge, k_carry_huge + k_carry_huge_count);
    t.insert(t.end(), k_carry_huge1, k_carry_huge1 + k_carry_huge1_count);
    return t;
  }();
  return v;
}

const CarryKernel *find_carry(uint32_t W, uint32_t R) {
  for (const auto &k : all_carry())
    if (k.w == W && k.r == R) return &k;
  return nullptr;
}

const CarryKernel *carry_table(size_t *count) {
  *count = all_carry().size();
  return all_carry().data();
}

const ResidentKernel *find_resident(uint32_t M) {
  for (const auto &k : k_resident)
    if (k.m == M) return &k;
  return nullptr;
}

uint32_t sweep_stagger(const void *fn, uint32_t threads, uint32_t dyn_lds, uint32_t groups,
                       uint32_t bytes_per_group, int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, dyn_lds) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  if (per_cu <= 0 || (uint64_t)groups > (uint64_t)per_cu * (uint64_t)cus) return 0;
  return bytes_per_group / 6144;
}

// Workgroups of one round of a kernel on the current device (its occupancy
// times the CUs; fn = null: the CUs), looked up once per kernel and device.
static uint32_t one_round(const void *fn, uint32_t threads) {
  static std::mutex mu;
  static std::vector<std::tuple<const void *, int, uint32_t>> known;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  for (const auto &e : known)
    if (std::get<0>(e) == fn && std::get<1>(e) == dev) return std::get<2>(e);
  int per_cu = 1, cus = 0;
  if ((fn && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)threads, 0) != hipSuccess) ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu <= 0 ||
      cus <= 0)
    return 0;
  const uint32_t r = (uint32_t)per_cu * (uint32_t)cus;
  known.emplace_back(fn, dev, r);
  return r;
}

uint32_t square_grid(const SquareKernel *k, uint32_t W, uint32_t tests) {
  if (!k->pipe || tests == 0) return W / 2;
  const uint32_t plain = one_round(k->fn, k->threads);
  if (plain == 0 || (uint64_t)(W / 2) * tests <= plain) return W / 2;
  const uint32_t r = one_round(k->pipe, k->threads) / tests;
  return r > 0 && r < W / 2 ? r : W / 2;
}

// A launch whose start and stop events (if set) take the kernel's own
// execution times.
static hipError_t launch_timed(const void *fn, dim3 grid, dim3 block, void **args, size_t lds, hipStream_t s,
                               hipEvent_t start, hipEvent_t stop) {
  if (start || stop) return hipExtLaunchKernel(fn, grid, block, args, lds, s, start, stop, 0);
  return hipLaunchKernel(fn, grid, block, args, lds, s);
}

hipError_t launch_square(const SquareKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                         uint32_t H, uint32_t grid, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const PlanArgs *a = d_args;
  void *args[] = {&a};
  const uint32_t lds = k->h ? 0 : square_generic_lds(H);
  if (grid == 0 || grid > W / 2) return hipErrorInvalidValue;
  const void *fn = k->pipe && grid < W / 2 ? k->pipe : k->fn;
  return launch_timed(fn, dim3(grid, tests), dim3(k->threads), args, lds, s, start, stop);
}

CarryLaunch carry_launch(const CarryKernel *k, uint32_t H, uint32_t R, uint32_t tests) {
  const uint32_t G = R ? H / R : 0;
  CarryLaunch c{k->carry, k->threads, G};
  if (tests == 0) return c;
  const uint64_t need = (uint64_t)G * tests;
  const uint32_t plain = one_round(k->carry, k->threads);
  if (k->pipe && plain && need > plain) {
    // (the persistent forms run 512 threads whatever the plain kernel's)
    const uint32_t r = one_round(k->pipe, 512) / tests;
    if (r > 0 && r < G) {
      c.fn = k->pipe;
      c.threads = 512;
      c.grid = r;
    }
    return c;
  }
  // (measured: with at most one block per two CUs the 256-thread form is
  // as fast or faster, 512K words 0.0220 against 0.0228 ms per iteration)
  if (k->wide) {
    const uint32_t wr = one_round(k->wide, 512), cus = one_round(nullptr, 0);
    if (wr && need <= wr && 2 * need > cus) {
      c.fn = k->wide;
      c.threads = 512;
    }
  }
  return c;
}

hipError_t launch_carry(const CarryKernel *k, const CarryLaunch &cl, const PlanArgs *d_args, uint32_t tests,
                        uint32_t W, uint32_t H, uint32_t R, hipStream_t s, hipEvent_t start, hipEvent_t stop) {
  const PlanArgs *a = d_args;
  void *args[] = {&a};
  const uint32_t lds = k->w ? 0 : carry_generic_lds(W, R);
  if (!cl.fn || cl.grid == 0 || cl.grid > H / R || (cl.fn != k->carry && cl.fn != k->pipe && cl.fn != k->wide))
    return hipErrorInvalidValue;
  return launch_timed(cl.fn, dim3(cl.grid, tests), dim3(cl.threads), args, lds, s, start, stop);
}

hipError_t launch_load(const CarryKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                       uint32_t H, uint32_t R, hipStream_t s) {
  const PlanArgs *a = d_args;
  void *args[] = {&a};
  const uint32_t lds = k->w ? 0 : carry_generic_lds(W, R);
  return hipLaunchKernel(k->load, dim3(H / R, tests), dim3(k->threads), args, lds, s);
}

hipError_t launch_unload(const CarryKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                         uint32_t H, uint32_t R, hipStream_t s) {
  const PlanArgs *a = d_args;
  void *args[] = {&a};
  const uint32_t lds = k->w ? 0 : carry_generic_lds(W, R);
  return hipLaunchKernel(k->unload, dim3(H / R, tests), dim3(k->threads), args, lds, s);
}

hipError_t launch_resident(const ResidentKernel *k, const ResidentDesc *d_descs, uint32_t count,
                           hipStream_t s) {
  const ResidentDesc *d = d_descs;
  void *args[] = {&d};
  return hipLaunchKernel(k->fn, dim3(count), dim3(k->threads), args, 0, s);
}

hipError_t launch_resident_unload(const ResidentDesc *d_descs, uint32_t count, int32_t *digits,
                                  uint32_t *err, hipStream_t s) {
  hipLaunchKernelGGL(resident_unload, dim3(16, count), dim3(256), 0, s, d_descs, digits, err);
  return hipGetLastError();
}

hipError_t launch_resident_load(const ResidentDesc *d_descs, uint32_t count, const int32_t *digits,
                                hipStream_t s) {
  hipLaunchKernelGGL(resident_load, dim3(16, count), dim3(256), 0, s, d_descs, digits);
  return hipGetLastError();
}

hipError_t launch_square_table(const SquareKernel *k, const double *qcH, cplx *ptH, hipStream_t s) {
  void *args[] = {&qcH, &ptH};
  return hipLaunchKernel(k->table, dim3(1), dim3(256), args, 0, s);
}

hipError_t launch_resident_table(const ResidentKernel *k, const double *qc, cplx *pt, hipStream_t s) {
  void *args[] = {&qc, &pt};
  return hipLaunchKernel(k->table, dim3(1), dim3(256), args, 0, s);
}

hipError_t launch_carry_table(const CarryKernel *k, const double *qcW, cplx *ptW, hipStream_t s) {
  void *args[] = {&qcW, &ptW};
  return hipLaunchKernel(k->table, dim3(1), dim3(256), args, 0, s);
}

hipError_t launch_copy(const cplx *in, cplx *out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(copy_stream, dim3(8192), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_fill(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  hipLaunchKernelGGL(fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace llamd

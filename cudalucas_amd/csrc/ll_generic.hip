// Generic two-sweep kernels: the LL iteration for any W x H x R schedule
// whose sizes are known only at run time, one instantiation for all of
// them.  They cover the FFT lengths of the reference's list
// (CUDALucas.cu:955-971) that the specialised tables (ll_kernels.hip,
// ll_carry_*.hip) have no template for: lines and columns with repeated or
// several odd factors (3136K = 2^15 * 49 complex points: W = 784) and
// lines beyond 2048 points (lengths above 16M words).
//
// The transforms run in place in dynamic LDS with a runtime radix plan
// (GenLine): forward passes decimate in frequency (natural order in,
// digit-reversed out), inverse passes decimate in time (digit-reversed in,
// natural out), so no pass needs a second buffer; the digit reversal is
// undone where data enters or leaves LDS (perm / iperm tables).  Twiddles
// come from one table of L roots exp(-2 pi i e/L) per plan.  The HBM layout
// of the signal, the four-step twiddle, the pairing squaring, the exact
// FP64 carries and the epoch-tagged carry hand-off are those of the
// specialised kernels (ll_sweeps.h), so both produce the same integers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ll_sweeps.h"

namespace llamd {

namespace {

constexpr int kGenThreads = 256;
// at least 4 waves per SIMD (4 workgroups per CU where the LDS allows):
// 128 registers per lane, so the largest radix is 8
constexpr int kGenWavesPerEU = 4;

// One butterfly of a pass: forward (DIF) reads R values at stride s, takes
// their length-R DFT and multiplies output k by w_L^(k j tstep); inverse
// (DIT) multiplies input k by the conjugate, then takes the inverse DFT.
// Roots of the plan: staged in LDS (LR) or read from global memory.
template <bool LR>
__device__ __forceinline__ cplx gen_root(const cplx *roots, int e) {
  if constexpr (LR) return roots[e];
  else return gload(roots + e);
}

template <int R, bool INV, bool LR>
__device__ __forceinline__ void gen_bfly(cplx *buf, int base, int s, int jt, const cplx *__restrict__ roots) {
  cplx v[R];
#pragma unroll
  for (int i = 0; i < R; i++) v[i] = buf[lpad(base + i * s)];
  if constexpr (INV) {
    if (jt) {
#pragma unroll
      for (int k = 1; k < R; k++) v[k] = cmulc(v[k], gen_root<LR>(roots, k * jt));
    }
    dft<R, true>(v);
  } else {
    dft<R, false>(v);
    if (jt) {
#pragma unroll
      for (int k = 1; k < R; k++) v[k] = cmul(v[k], gen_root<LR>(roots, k * jt));
    }
  }
#pragma unroll
  for (int i = 0; i < R; i++) buf[lpad(base + i * s)] = v[i];
}

// One pass of radix R and stride s over NB transforms at lds + b * LP.
template <bool INV, bool LR>
__device__ __forceinline__ void gen_pass(cplx *lds, int LP, int NB, const GenLine &g, const cplx *roots, int R,
                                         int s) {
  const int L = (int)g.L;
  const int Q = L / R;
  const int tstep = L / (R * s);
  for (int u = threadIdx.x; u < NB * Q; u += blockDim.x) {
    const int b = u / Q, v = u - b * Q;
    const int blk = v / s, j = v - blk * s;
    const int base = blk * R * s + j;
    cplx *buf = lds + b * LP;
    const int jt = j * tstep;
    switch (R) {
      case 2: gen_bfly<2, INV, LR>(buf, base, s, jt, roots); break;
      case 3: gen_bfly<3, INV, LR>(buf, base, s, jt, roots); break;
      case 4: gen_bfly<4, INV, LR>(buf, base, s, jt, roots); break;
      case 5: gen_bfly<5, INV, LR>(buf, base, s, jt, roots); break;
      case 7: gen_bfly<7, INV, LR>(buf, base, s, jt, roots); break;
      default: gen_bfly<8, INV, LR>(buf, base, s, jt, roots); break;
    }
  }
}

// Whole transforms of the NB lines; callers synchronise before.
// Pass p has stride s_p = L / (r_0 ... r_p): forward passes run p = 0, 1,
// ...; inverse passes run backwards from s = 1, s_(p-1) = s_p r_p.
template <bool INV, bool LR = false>
__device__ __forceinline__ void gen_fft(cplx *lds, int LP, int NB, const GenLine &g, const cplx *roots) {
  const int np = (int)g.np;
  if constexpr (!INV) {
    int s = (int)g.L;
    for (int p = 0; p < np; p++) {
      const int r = g.radix[p];
      s /= r;
      gen_pass<false, LR>(lds, LP, NB, g, roots, r, s);
      __syncthreads();
    }
  } else {
    int s = 1;
    for (int p = np - 1; p >= 0; p--) {
      const int r = g.radix[p];
      gen_pass<true, LR>(lds, LP, NB, g, roots, r, s);
      __syncthreads();
      s *= r;
    }
  }
}

// Loops over i = threadIdx.x + k * blockDim.x < n in chunks of kChunk, the
// chunk's reads (ld) all issued before any result is used (st), so a
// thread waits for memory once per chunk, not once per element.
constexpr int kChunk = 4;
template <typename LD, typename ST>
__device__ __forceinline__ void chunked(int n, const LD &ld, const ST &st) {
  const int T = blockDim.x;
  for (int base = threadIdx.x; base < n; base += kChunk * T) {
    decltype(ld(0)) v[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (base + k * T < n) v[k] = ld(base + k * T);
#pragma unroll
    for (int k = 0; k < kChunk; k++)
      if (base + k * T < n) st(base + k * T, v[k]);
  }
}

struct CplxPair {
  cplx a, b;
};
struct CplxAt {
  cplx v;
  uint32_t pos;
};

__device__ __forceinline__ float gen_block_max(float v, float *red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float m = 0.f;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmaxf(m, red[w]);
  return m;  // valid in thread 0
}

}  // namespace

// Plans of at most kStage points keep their roots and digit-reversal table
// in LDS beside the data (16 L + 4 L bytes), and those workgroups issue the
// four-step twiddles of their pairing and store phases one transform ahead,
// so that no pass or phase waits on a global read: the batched engine runs
// thousands of such small workgroups, each bound by latency.
constexpr int kStage = 1024;
constexpr int kStagePer = kStage / kGenThreads;  // table entries per thread
__host__ __device__ constexpr uint32_t stage_bytes(uint32_t L) { return L <= (uint32_t)kStage ? 20u * L : 0u; }

// A plan's roots and inverse digit reversal on their way into LDS: read()
// issues the global loads, write() stores them once they have arrived, so
// that the signal loads go out in between.
struct StagedPlan {
  cplx r[kStagePer];
  uint32_t q[kStagePer];
  __device__ __forceinline__ void read(const GenLine &g) {
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int i = (int)threadIdx.x + k * kGenThreads;
      if (i < (int)g.L) {
        r[k] = gload(g.roots + i);
        q[k] = gload(g.iperm + i);
      }
    }
  }
  __device__ __forceinline__ void write(const GenLine &g, cplx *rts, uint32_t *ipl) const {
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int i = (int)threadIdx.x + k * kGenThreads;
      if (i < (int)g.L) {
        rts[i] = r[k];
        ipl[i] = q[k];
      }
    }
  }
};

// --------------------------------------------------------- square sweep
// W / 2 workgroups; workgroup g owns columns g and W - g (0 and W / 2 for
// g = 0), as square_sweep.
template <bool ST>
__device__ __forceinline__ void square_generic_body(const PlanArgs &A, const uint32_t g, cplx *dyn) {
  const bool finished = A.remaining && *A.remaining == 0;
  if (g == 0 && threadIdx.x == 0) {  // column pair 0 opens the iteration
    DevState &st = *A.st;
    if (finished) {
      st.run = 0;
    } else {
      const Minus2 m2 = minus2_position(st.offset, A.p, A.n);
      st.m2_digit = m2.digit;
      st.m2_value = m2.value;
      st.offset = m2.new_offset;
      A.err[st.epoch & A.err_mask] = 0u;
      st.epoch += 1;
      st.run = 1;
    }
  }
  if (finished) return;
  const int H = (int)A.H, W = (int)A.W, LP = lpitch(H);
  const uint32_t cols[2] = {g, g == 0 ? (uint32_t)W / 2 : (uint32_t)W - g};
  const uint32_t rs = A.rshift, rmask = (1u << rs) - 1;
  const int tid = threadIdx.x;
  cplx *const rts = dyn + 2 * LP;                 // staged roots (ST)
  uint32_t *const ipl = (uint32_t *)(rts + H);    // staged inverse digit reversal (ST)
  StagedPlan sp;
  if constexpr (ST) sp.read(A.gH);
  chunked(
      2 * H,
      [&](int i) {
        const int which = i >= H, b = i - which * H;
        const uint32_t c = cols[which];
        const size_t off = (((size_t)((uint32_t)b >> rs) * W + c) << rs) | ((uint32_t)b & rmask);
        return CplxPair{gload_stream(A.x + off), twiddle_m(A.twM, c * (uint32_t)b)};
      },
      [&](int i, const CplxPair &v) {
        const int which = i >= H, b = i - which * H;
        dyn[which * LP + lpad(b)] = cmul(v.a, v.b);
      });
  if constexpr (ST) sp.write(A.gH, rts, ipl);
  // the pairing twiddles of this thread, read while the transform runs
  cplx tp[ST ? kStagePer : 1];
  if constexpr (ST) {
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int k2 = tid + k * kGenThreads;
      if (g != 0 && k2 < H) tp[k] = twiddle_m(A.twM, g + (uint32_t)W * (uint32_t)k2);
    }
  }
  __syncthreads();
  if constexpr (ST) gen_fft<false, true>(dyn, LP, 2, A.gH, rts);
  else gen_fft<false>(dyn, LP, 2, A.gH, A.gH.roots);
  const uint32_t *ip = ST ? ipl : A.gH.iperm;
  if (g != 0) {
    cplx *c0 = dyn, *c1 = dyn + LP;
    auto pair = [&](int k2, cplx t) {
      const int pa = lpad((int)ip[k2]), pb = lpad((int)ip[H - 1 - k2]);
      cplx a = c0[pa], b = c1[pb];
      pair_square(a, b, t);
      c0[pa] = a;
      c1[pb] = b;
    };
    if constexpr (ST) {
#pragma unroll
      for (int k = 0; k < kStagePer; k++) {
        const int k2 = tid + k * kGenThreads;
        if (k2 < H) pair(k2, tp[k]);
      }
    } else {
      for (int k2 = tid; k2 < H; k2 += blockDim.x) pair(k2, twiddle_m(A.twM, g + (uint32_t)W * (uint32_t)k2));
    }
  } else {
    // column 0: k2 <-> H - k2 ; column W/2: k2 <-> H - 1 - k2
    for (int k2 = tid; k2 <= H / 2; k2 += blockDim.x) {
      const int k2b = (H - k2) % H;
      const int pa = lpad((int)ip[k2]), pb = lpad((int)ip[k2b]);
      cplx a = dyn[pa], b = dyn[pb];
      pair_square(a, b, twiddle_m(A.twM, (uint32_t)W * (uint32_t)k2));
      dyn[pa] = a;
      if (k2b != k2) dyn[pb] = b;
    }
    cplx *c1 = dyn + LP;
    for (int k2 = tid; k2 < (H + 1) / 2; k2 += blockDim.x) {
      const int k2b = H - 1 - k2;
      const int pa = lpad((int)ip[k2]), pb = lpad((int)ip[k2b]);
      cplx a = c1[pa], b = c1[pb];
      pair_square(a, b, twiddle_m(A.twM, (uint32_t)W / 2 + (uint32_t)W * (uint32_t)k2));
      c1[pa] = a;
      if (k2b != k2) c1[pb] = b;
    }
  }
  const StreamOut out(A.x, A.n / 2);
  auto store = [&](int i, const cplx &w) {
    const int which = i >= H, b = i - which * H;
    const uint32_t c = cols[which];
    const uint32_t off = (((uint32_t)b >> rs) * (uint32_t)W + c) << rs | ((uint32_t)b & rmask);
    out.store(off, cmulc(dyn[which * LP + lpad(b)], w));
  };
  if constexpr (ST) {
    // the store twiddles, read while the inverse transform runs
    cplx ts[2 * kStagePer];
#pragma unroll
    for (int k = 0; k < 2 * kStagePer; k++) {
      const int i = tid + k * kGenThreads;
      if (i < 2 * H) {
        const int which = i >= H, b = i - which * H;
        ts[k] = twiddle_m(A.twM, cols[which] * (uint32_t)b);
      }
    }
    __syncthreads();
    gen_fft<true, true>(dyn, LP, 2, A.gH, rts);
#pragma unroll
    for (int k = 0; k < 2 * kStagePer; k++) {
      const int i = tid + k * kGenThreads;
      if (i < 2 * H) store(i, ts[k]);
    }
  } else {
    __syncthreads();
    gen_fft<true>(dyn, LP, 2, A.gH, A.gH.roots);
    chunked(
        2 * H,
        [&](int i) {
          const int which = i >= H, b = i - which * H;
          return twiddle_m(A.twM, cols[which] * (uint32_t)b);
        },
        store);
  }
}

__global__ __launch_bounds__(kGenThreads, kGenWavesPerEU) void square_generic(const PlanArgs *__restrict__ args) {
  extern __shared__ cplx dyn[];
  const PlanArgs &A = args[blockIdx.y];
  // neighbouring column pairs on one XCD, as square_sweep
  const uint32_t g = pair_of_workgroup(blockIdx.x, gridDim.x, A.rshift);
  if (A.H <= (uint32_t)kStage) square_generic_body<true>(A, g, dyn);
  else square_generic_body<false>(A, g, dyn);
}

// ---------------------------------------------------------- carry sweep
// H / R workgroups, workgroup k owning block k of R lines, as carry_sweep:
// inverse line transforms, unweight, round, exact carries along each
// column's 2R words, hand-off of the chain's carry-out to the next block
// (epoch-tagged shuttle), carry-in through the whole chain, -2, reweight,
// forward line transforms.  Balanced digits wait in LDS for the carry-in.
template <bool ST>
__device__ __forceinline__ void carry_generic_body(const PlanArgs &A, const uint32_t blk, cplx *dyn) {
  if (!A.st->run) return;  // a finished batched test
  const int W = (int)A.W, R = 1 << A.rshift, LP = lpitch(W);
  const uint32_t H = A.H, G = H / (uint32_t)R, b0 = blk * (uint32_t)R;
  cplx *lds = dyn;
  double *wb = (double *)(lds + R * LP);
  double *iwb = wb + 2 * R;
  int *fb = (int *)(iwb + 2 * R);
  float *red = (float *)(fb + 2 * R);
  // staged roots and inverse digit reversal of the line plan (ST), after
  // the block's own arrays, 16-byte aligned
  cplx *const rts = lds + R * LP + (2 * R * (8 + 8 + 4) + (kGenThreads / 64) * 4 + 15) / 16;
  uint32_t *const ipl = (uint32_t *)(rts + W);
  const int tid = threadIdx.x;
  const int rs = (int)A.rshift;
  StagedPlan sp;
  // per-column weights of this thread's columns a = tid + k T (ST)
  int fa_c[ST ? kStagePer : 1];
  double iwa_c[ST ? kStagePer : 1], wa_c[ST ? kStagePer : 1];
  if constexpr (ST) {
    sp.read(A.gW);
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int a = tid + k * kGenThreads;
      if (a < W) {
        fa_c[k] = gload(A.fA + a);
        iwa_c[k] = gload(A.iwAs + a);
        wa_c[k] = gload(A.wA + a);
      }
    }
  }
  const uint32_t *gip = A.gW.iperm;
  chunked(
      R * W,
      [&](int e) { return CplxAt{gload_stream(A.x + (size_t)b0 * W + e), gload(gip + (e >> rs))}; },
      [&](int e, const CplxAt &v) { lds[(e & (R - 1)) * LP + lpad((int)v.pos)] = v.v; });
  if (tid < 2 * R) {
    const uint32_t i = 2 * b0 + tid;
    fb[tid] = gload(A.fB + i);
    wb[tid] = gload(A.wB + i);
    iwb[tid] = gload(A.iwB + i);
  }
  if constexpr (ST) sp.write(A.gW, rts, ipl);
  const uint32_t epoch = A.st->epoch;
  const uint32_t m2_digit = A.st->m2_digit;
  const double dmod = (double)A.st->m2_value;
  __syncthreads();
  if constexpr (ST) gen_fft<true, true>(lds, LP, R, A.gW, rts);
  else gen_fft<true>(lds, LP, R, A.gW, A.gW.roots);

  const int q = (int)A.q;
  const int n = (int)A.n, bb = (int)A.b;
  double emax = 0.0;
  // round, subtract, carry along each column's chain; publish the carry-out
  auto round_column = [&](const int a, const int fa, const double iwa) {
    const uint32_t m2i = m2_digit - 2u * ((uint32_t)a * H + b0);
    double carry = 0.0;
    for (int r = 0; r < R; r++) {
      const cplx v = lds[r * LP + lpad(a)];
      double d[2];
      for (int e = 0; e < 2; e++) {
        const int i = 2 * r + e;
        const int f0 = fa + fb[i];
        const int wrap = f0 >= n;
        const int big = f0 - (wrap ? n : 0) < bb;
        const double t = ldexp(mul3(e ? v.y : v.x, iwa, iwb[i]), wrap);
        double rt = rint(t);
        emax = fmax(emax, fabs(t - rt));
        if ((uint32_t)i == m2i) rt -= dmod;
        d[e] = balance_f64(rt + carry, q + big, carry);
      }
      lds[r * LP + lpad(a)] = {d[0], d[1]};
    }
    const bool fits = fabs(carry) < 2147483648.0;
    if (!fits) emax = 2.0;
    const int32_t c32 = fits ? (int32_t)carry : 0;
    const uint64_t word = ((uint64_t)epoch << 32) | (uint32_t)c32;
    __hip_atomic_store(&A.shuttle[(size_t)blk * W + a], word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  // carry-in from the previous block (block 0: the last block, one column
  // to the left), through the whole chain, then reweight
  auto carry_in_column = [&](const int a, const int fa, const double wa) {
    const size_t src = blk == 0 ? (size_t)(G - 1) * W + (a + W - 1) % W : (size_t)(blk - 1) * W + a;
    uint64_t word;
    uint32_t spins = 0;
    for (;;) {
      word = __hip_atomic_load(&A.shuttle[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(word >> 32) == epoch) break;
      if (++spins > kSpinLimit) {
        emax = 3.0;
        word = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    double cc = (double)(int32_t)(uint32_t)word;
    for (int r = 0; r < R; r++) {
      const cplx dv = lds[r * LP + lpad(a)];
      double x[2];
      for (int e = 0; e < 2; e++) {
        const int i = 2 * r + e;
        const int f0 = fa + fb[i];
        const int wrap = f0 >= n;
        const int width = q + (f0 - (wrap ? n : 0) < bb);
        double d = balance_f64((e ? dv.y : dv.x) + cc, width, cc);
        if (i == 2 * R - 1) d += ldexp(cc, width);
        x[e] = ldexp(mul3(d, wa, wb[i]), -wrap);
      }
      lds[r * LP + lpad(a)] = {x[0], x[1]};
    }
  };
  if constexpr (ST) {
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int a = tid + k * kGenThreads;
      if (a < W) round_column(a, fa_c[k], iwa_c[k]);
    }
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
      const int a = tid + k * kGenThreads;
      if (a < W) carry_in_column(a, fa_c[k], wa_c[k]);
    }
  } else {
    for (int a = tid; a < W; a += blockDim.x) round_column(a, gload(A.fA + a), gload(A.iwAs + a));
    for (int a = tid; a < W; a += blockDim.x) carry_in_column(a, gload(A.fA + a), gload(A.wA + a));
  }
  __syncthreads();
  if constexpr (ST) gen_fft<false, true>(lds, LP, R, A.gW, rts);
  else gen_fft<false>(lds, LP, R, A.gW, A.gW.roots);
  const StreamOut out(A.x, A.n / 2);
  const uint32_t *ip = ST ? ipl : gip;
  chunked(
      R * W, [&](int e) { return ip[e >> rs]; },
      [&](int e, const uint32_t &q) { out.store(b0 * (uint32_t)W + e, lds[(e & (R - 1)) * LP + lpad((int)q)]); });
  const float m = gen_block_max((float)emax, red);
  if (tid == 0) {
    atomicMax(&A.err[(epoch - 1) & A.err_mask], __float_as_uint(m));
    if (blk == 0 && A.remaining) *A.remaining -= 1;
  }
}

__global__ __launch_bounds__(kGenThreads, kGenWavesPerEU) void carry_generic(const PlanArgs *__restrict__ args) {
  extern __shared__ cplx dyn[];
  const PlanArgs &A = args[blockIdx.y];
  if (A.W <= (uint32_t)kStage) carry_generic_body<true>(A, blockIdx.x, dyn);
  else carry_generic_body<false>(A, blockIdx.x, dyn);
}

// ------------------------------------------------- batched: one launch
// One LL iteration of every two-sweep test of a batch in a single launch.
// Workgroup i runs item items[i]: first every test's square items, then
// every test's carry items.  A test's carry items start once all its square
// items of this iteration have finished (per-test counters: sq_done counts
// square items ever finished, ca_done carry items; the carry items of
// iteration e of a test wait for sq_done >= (e + 1) W / 2, e read from
// ca_done, which this item has not yet raised).  Every wait is on items of
// lower index (or on the last carry block's carries, published before it
// waits), which each XCD dispatches first, as in the single-test carry
// sweep; it holds while the tests are fewer than the workgroups the GPU
// holds at once.
__global__ __launch_bounds__(kGenThreads, kGenWavesPerEU) void batch_iteration(const BatchItem *__restrict__ items,
                                                               const PlanArgs *__restrict__ args,
                                                               unsigned long long *sq_done,
                                                               unsigned long long *ca_done) {
  extern __shared__ cplx dyn[];
  __shared__ int s_go;
  const BatchItem it = items[blockIdx.x];
  const PlanArgs &A = args[it.test];
  if (!it.carry) {
    if (A.H <= (uint32_t)kStage) square_generic_body<true>(A, it.block, dyn);
    else square_generic_body<false>(A, it.block, dyn);
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(&sq_done[it.test], 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (threadIdx.x == 0) {
    const unsigned long long G = A.H >> A.rshift;
    const unsigned long long e =
        __hip_atomic_load(&ca_done[it.test], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) / G;
    const unsigned long long want = (e + 1) * (A.W / 2);
    uint32_t spins = 0;
    int go = 1;
    // relaxed polls: an acquire on every poll would invalidate this XCD's
    // caches each time (buffer_inv) under every other workgroup; the one
    // acquire fence below orders the item's reads after the wait
    while (__hip_atomic_load(&sq_done[it.test], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
      if (++spins > kSpinLimit) {  // never hang: the iteration reports failure
        go = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    s_go = go;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (s_go) {
    if (A.W <= (uint32_t)kStage) carry_generic_body<true>(A, it.block, dyn);
    else carry_generic_body<false>(A, it.block, dyn);
  } else if (threadIdx.x == 0) {
    atomicMax(&A.err[(A.st->epoch - 1) & A.err_mask], __float_as_uint(3.0f));
  }
  __syncthreads();
  // relaxed: ca_done only counts iterations for the waits above; what the
  // carry items wrote is read in the next launch, after the kernel boundary
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add(&ca_done[it.test], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

hipError_t launch_batch_iteration(const BatchItem *items, uint32_t count, const PlanArgs *args,
                                  unsigned long long *sq_done, unsigned long long *ca_done, uint32_t lds,
                                  hipStream_t s) {
  void *a[] = {&items, &args, &sq_done, &ca_done};
  return hipLaunchKernel((const void *)batch_iteration, dim3(count), dim3(kGenThreads), a, lds, s);
}

// ------------------------------------------------------ load / unload (HBM)
// load: balanced digits (natural order) -> weighted, line-transformed
// signal, the state the square sweep expects.
__global__ __launch_bounds__(kGenThreads, kGenWavesPerEU) void load_generic(const PlanArgs *__restrict__ args) {
  extern __shared__ cplx dyn[];
  const PlanArgs &A = args[blockIdx.y];
  const int W = (int)A.W, R = 1 << A.rshift, LP = lpitch(W), rs = (int)A.rshift;
  const uint32_t H = A.H, b0 = blockIdx.x * (uint32_t)R;
  const int tid = threadIdx.x;
  for (int a = tid; a < W; a += blockDim.x) {
    const int fa = A.fA[a];
    const double wa = A.wA[a], iwa = A.iwA[a];
    for (int r = 0; r < R; r++) {
      const uint32_t j = 2u * ((uint32_t)a * H + b0 + r);
      const WordWeight we = word_weight(A, fa, wa, iwa, 2 * (b0 + r));
      const WordWeight wo = word_weight(A, fa, wa, iwa, 2 * (b0 + r) + 1);
      dyn[r * LP + lpad(a)] = {(double)A.digits[j] * we.w, (double)A.digits[j + 1] * wo.w};
    }
  }
  __syncthreads();
  gen_fft<false>(dyn, LP, R, A.gW, A.gW.roots);
  const uint32_t *ip = A.gW.iperm;
  for (int e = tid; e < R * W; e += blockDim.x) {
    const int r = e & (R - 1), a = e >> rs;
    A.x[(size_t)b0 * W + e] = dyn[r * LP + lpad((int)ip[a])];
  }
}

// unload: inverse of load; digits rounded, rounding error max into err[0].
__global__ __launch_bounds__(kGenThreads, kGenWavesPerEU) void unload_generic(const PlanArgs *__restrict__ args) {
  extern __shared__ cplx dyn[];
  const PlanArgs &A = args[blockIdx.y];
  const int W = (int)A.W, R = 1 << A.rshift, LP = lpitch(W), rs = (int)A.rshift;
  const uint32_t H = A.H, b0 = blockIdx.x * (uint32_t)R;
  float *red = (float *)(dyn + R * LP);
  const int tid = threadIdx.x;
  const uint32_t *ip = A.gW.iperm;
  for (int e = tid; e < R * W; e += blockDim.x) {
    const int r = e & (R - 1), a = e >> rs;
    dyn[r * LP + lpad((int)ip[a])] = A.x[(size_t)b0 * W + e];
  }
  __syncthreads();
  gen_fft<true>(dyn, LP, R, A.gW, A.gW.roots);
  float emax = 0.f;
  const double scale = 1.0 / (double)W;
  for (int a = tid; a < W; a += blockDim.x) {
    const int fa = A.fA[a];
    const double wa = A.wA[a], iwa = A.iwA[a];
    for (int r = 0; r < R; r++) {
      const uint32_t j = 2u * ((uint32_t)a * H + b0 + r);
      const cplx v = dyn[r * LP + lpad(a)];
      const WordWeight we = word_weight(A, fa, wa, iwa, 2 * (b0 + r));
      const WordWeight wo = word_weight(A, fa, wa, iwa, 2 * (b0 + r) + 1);
      const double t0 = v.x * we.iw * scale, t1 = v.y * wo.iw * scale;
      const double r0 = rint(t0), r1 = rint(t1);
      emax = fmaxf(emax, fmaxf((float)fabs(t0 - r0), (float)fabs(t1 - r1)));
      A.digits[j] = (int32_t)r0;
      A.digits[j + 1] = (int32_t)r1;
    }
  }
  const float m = gen_block_max(emax, red);
  if (tid == 0) atomicMax(&A.err[0], __float_as_uint(m));
}

// ------------------------------------------------------------- dispatch

uint32_t square_generic_lds(uint32_t H) {
  return 2u * (uint32_t)lpitch((int)H) * (uint32_t)sizeof(cplx) + stage_bytes(H);
}

uint32_t carry_generic_lds(uint32_t W, uint32_t R) {
  const uint32_t own = R * (uint32_t)lpitch((int)W) * (uint32_t)sizeof(cplx) + 2 * R * (8 + 8 + 4) + (kGenThreads / 64) * 4;
  return stage_bytes(W) ? (own + 15) / 16 * 16 + stage_bytes(W) : own;
}

static const SquareKernel k_square_generic = {0, (const void *)square_generic, kGenThreads, nullptr, nullptr};
static const CarryKernel k_carry_generic = {0, 0, (const void *)carry_generic, (const void *)load_generic,
                                            (const void *)unload_generic, kGenThreads, nullptr};

// The generic kernels take up to kGenMaxLds bytes of dynamic LDS, above
// the default launch limit: raised once per process (and device set).
static void allow_generic_lds() {
  static bool done = false;
  if (done) return;
  const void *fns[] = {(const void *)square_generic, (const void *)carry_generic, (const void *)load_generic,
                       (const void *)unload_generic, (const void *)batch_iteration};
  for (const void *f : fns) {
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenMaxLds);
    (void)hipGetLastError();  // a refusal must not linger as the next launch's error
  }
  done = true;
}

void generic_kernels_init() { allow_generic_lds(); }

const SquareKernel *generic_square() {
  allow_generic_lds();
  return &k_square_generic;
}

const CarryKernel *generic_carry() {
  allow_generic_lds();
  return &k_carry_generic;
}

}  // namespace llamd

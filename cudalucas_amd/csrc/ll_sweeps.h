// Device code of the Lucas-Lehmer iteration s <- s^2 - 2 mod 2^p - 1 on gfx950:
// the two-sweep kernels and their helpers, instantiated by the ll_*.hip files.
//
// One iteration of the reference (CUDALucas.cu:1869-1986: cuFFT INVERSE,
// square, cuFFT INVERSE, rcb, splice) is rebuilt in two shapes:
//
//  * resident_kernel: for N <= 16K words the whole M = N/2-point complex
//    signal lives in one workgroup's LDS and the kernel loops over many
//    iterations without touching HBM (batched: one workgroup per exponent).
//
//  * two sweeps over HBM for large N, M = W x H.  Complex element
//    m = a*H + b holds words (x[2m], x[2m+1]); lines b are tiled by R, so the
//    element lives at ((b / R) W + a) R + (b mod R): a carry workgroup's R
//    lines are one contiguous block, a column is read in runs of R.
//      square_sweep : per column pair (k1, W-k1): four-step twiddle,
//                     forward H-point DFT, real-pairing squaring, inverse
//                     H-point DFT, inverse twiddle.
//      carry_sweep  : per block of R lines b: inverse W-point DFT,
//                     unweight/round/-2/balanced carry, carry look-back
//                     from the previous block, reweight, forward W-point
//                     DFT.
//    Each sweep reads and writes the signal once: 32 N bytes per iteration.
//
// Carries between line blocks travel through an epoch-tagged shuttle
// (64-bit words: iteration epoch in the high half, int32 carry in the low
// half) written and polled with agent-scope atomics, so no fence or flag is
// needed.  Block k waits only on block k-1, dispatched before it; the
// cyclic wrap (the last block's carries feed block 0, one column to the
// left) is the one wait on a later block, and it cannot deadlock because
// every block publishes its own carries before it waits, so the last block
// never waits on block 0.  That argument holds for one carry launch at a
// time on the GPU (each XCD dispatches its workgroups in order); two carry
// launches running side by side can hold each other's workgroup slots, so
// callers run them one after another.  The square sweep's workgroup 0 opens
// each iteration in a small device state (epoch, offset, -2 position) that
// the carry sweep reads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ll_fft.h"
#include "ll_kernels.h"

namespace llamd {

// Upper bound on carry look-back polls (a few hundred ms); reaching it marks
// the iteration failed (error 3.0) instead of hanging the device.
constexpr uint32_t kSpinLimit = 1u << 22;

// ------------------------------------------------------------------ helpers

// Staggered start: workgroup i waits i * sixteenths / 16 ticks of the
// 100 MHz real-time counter (measured from its own start) before its first
// load, so that the sweep's loads arrive in workgroup order and early
// workgroups compute and store while later ones load.
__device__ __forceinline__ void stagger_start(uint32_t sixteenths) {
  if (sixteenths) {
    const uint64_t until = __builtin_amdgcn_s_memrealtime() + (((uint64_t)blockIdx.x * sixteenths) >> 4);
    while (__builtin_amdgcn_s_memrealtime() < until) __builtin_amdgcn_s_sleep(1);
  }
}

// Column pair of square-sweep workgroup i of n.  With lines tiled by R < 8
// a column is read in runs of 16 R bytes, so S = 8 / R neighbouring columns
// share each 128-byte line; workgroups i, i + 8, ..., i + 8 (S - 1), which
// the dispatcher sends one after another to the same XCD, take S
// neighbouring pairs, so that a line's parts meet in that XCD's L2 instead
// of each being fetched from HBM.  A bijection of [0, n); the identity for
// R = 8 and for the last n mod 8S workgroups.
__device__ __forceinline__ uint32_t pair_of_workgroup(uint32_t i, uint32_t n, uint32_t rshift) {
  if (rshift >= 3) return i;
  const uint32_t S = 8u >> rshift, span = 8 * S;
  if (i >= n / span * span) return i;
  const uint32_t x = i % 8, q = i / 8;
  return span * (q / S) + S * x + q % S;
}

__device__ __forceinline__ cplx twiddle_m(const TwoLevel &t, uint32_t e) {
  const uint32_t k = 2 * (e >> t.shift);
  const cplx h = gload(t.hi + k), hr = gload(t.hi + k + 1);
  const cplx l = gload(t.lo + (e & t.mask));  // (cos - 1, sin) of a small angle
  // h + (remainder + h l): the small terms first, one rounding at the end
  return {h.x + (hr.x + fma(h.x, l.x, -h.y * l.y)), h.y + (hr.y + fma(h.x, l.y, h.y * l.x))};
}

// x * (a * b) with the product a * b carried exactly (its rounding error
// recovered by an fma): one rounding instead of two.
__device__ __forceinline__ double mul3(double x, double a, double b) {
  const double p = a * b;
  const double pe = fma(a, b, -p);
  return fma(x, p, x * pe);
}

// Balanced digit of value v in base 2^width and the carry out.
__device__ __forceinline__ int64_t balance(int64_t v, int width, int64_t &carry) {
  const int64_t half = (int64_t)1 << (width - 1);
  const int64_t d = ((v + half) & ((half << 1) - 1)) - half;
  carry = (v - d) >> width;
  return d;
}

// The same split in FP64, exact for |v| < 2^53: every value here is an
// integer and the base 2^width a power of two, so v * 2^-width, its rint and
// the fma below are all exact.  Ties round to even, so the digit lies in
// [-2^(width-1), 2^(width-1)].
// Balanced digit of the integer-valued v in base 2^w and the carry out,
// exactly: v 2^-w and carry 2^w are exact scalings, and v - carry 2^w is
// an integer of magnitude <= 2^(w-1).
__device__ __forceinline__ double balance_f64(double v, int w, double &carry) {
  carry = rint(ldexp(v, -w));
  return v - ldexp(carry, w);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

template <int T>
__device__ __forceinline__ float block_max(float v, float *scratch) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  float m = 0.f;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < T / 64; w++) m = fmaxf(m, scratch[w]);
  }
  return m;  // valid in thread 0
}

// Position of the -2 for the iteration that maps offset -> 2*offset mod p:
// subtract 2^(pos) from digit `digit` (CUDALucas.cu:1882-1886, exact).
struct Minus2 {
  uint32_t new_offset;
  uint32_t digit;
  int64_t value;
};

__device__ __forceinline__ Minus2 minus2_position(uint32_t offset, uint32_t p, uint32_t n) {
  Minus2 m;
  m.new_offset = (uint32_t)((2ull * offset) % p);
  uint64_t bit = (uint64_t)m.new_offset + 1;
  if (bit == p) bit = 0;
  m.digit = (uint32_t)((bit * n) / p);
  const uint64_t start = ((uint64_t)m.digit * p + n - 1) / n;
  m.value = (int64_t)1 << (bit - start);
  return m;
}

// Words of the pair (2(aH+b), 2(aH+b)+1) use f = fA[a] + fB[2b+e] (mod n).
struct WordWeight {
  double w, iw;
  int width;
};

__device__ __forceinline__ WordWeight word_weight(const PlanArgs &A, int fa, double wa, double iwa,
                                                  int i) {
  int f = fa + A.fB[i];
  const bool wrap = f >= (int)A.n;
  if (wrap) f -= (int)A.n;
  WordWeight r;
  r.w = wa * A.wB[i] * (wrap ? 0.5 : 1.0);
  r.iw = iwa * A.iwB[i] * (wrap ? 2.0 : 1.0);
  r.width = (int)A.q + (f < (int)A.b ? 1 : 0);
  return r;
}

// Real-pairing squaring of the half-length spectrum: with A = Z[k],
// B = Z[M-k], t = exp(-2 pi i k/M), E2 = A + conj(B), O2 = -i (A - conj(B)),
//   Z'[k]   = E2^2 + t O2^2 + 2i E2 O2
//   Z'[M-k] = conj(E2^2 + t O2^2 - 2i E2 O2)
// which is 4x the half-length spectrum of the squared real signal; the
// factor 1/4 is applied with the unweighting.
__device__ __forceinline__ void pair_square(cplx &A, cplx &B, cplx t) {
  const cplx e2 = {A.x + B.x, A.y - B.y};
  const cplx o2 = {A.y + B.y, B.x - A.x};
  const cplx P = csqr(e2);
  const cplx Q = cmul(t, csqr(o2));
  const cplx eo = cmul(e2, o2);
  const cplx S = {-2.0 * eo.y, 2.0 * eo.x};
  const cplx PQ = cadd(P, Q);
  A = cadd(PQ, S);
  B = cconj(csub(PQ, S));
}

// ------------------------------------------------------------ square sweep
// W/2 workgroups: workgroup g in [1, W/2) owns the column pair (g, W - g);
// workgroup 0 owns the two self-paired columns 0 and W/2.  After the forward
// H-point DFT, element k2 of column k1 holds Z[k1 + W k2]; its partner
// M - k is (W - k1, H - 1 - k2) for k1 != 0 and (0, H - k2) for k1 = 0.
template <int H, int T>
__global__ __launch_bounds__(T, 2) void square_sweep(const PlanArgs *__restrict__ args) {
  const PlanArgs &A = args[blockIdx.y];
  stagger_start(A.stagger_sq);
  // Opens the iteration for the carry sweep that follows this launch
  // (workgroup 0) and tells whether this is a batched test that is done.
  // Called once the column loads are in flight, so that the state reads do
  // not hold them back.
  auto open_iteration = [&]() -> bool {
    const bool finished = A.remaining && *A.remaining == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      DevState &st = *A.st;
      if (finished) {
        st.run = 0;
      } else {
        const Minus2 m2 = minus2_position(st.offset, A.p, A.n);
        st.m2_digit = m2.digit;
        st.m2_value = m2.value;
        st.offset = m2.new_offset;
        // the error slot of this iteration, which the carry sweep's
        // workgroups raise with atomicMax after this launch has ended
        A.err[st.epoch & A.err_mask] = 0u;
        st.epoch += 1;
        st.run = 1;
      }
    }
    return finished;
  };
  constexpr int LP = lpitch(H);
  __shared__ cplx lds[2 * LP];
  const int tid = threadIdx.x;
  const uint32_t W = A.W;
  const uint32_t g = pair_of_workgroup(blockIdx.x, gridDim.x, A.rshift);
  const uint32_t c0 = g;
  const uint32_t c1 = g == 0 ? W / 2 : W - g;
  // column c, element b lives at ((b >> rs) W + c) << rs | (b & rmask)
  const uint32_t rs = A.rshift, rmask = (1u << rs) - 1;
  const cplx *col0 = A.x + ((size_t)c0 << rs);
  const cplx *col1 = A.x + ((size_t)c1 << rs);
  // The four-step twiddle w_M^(c b) between the line and the column DFTs
  // is applied on the way in and undone on the way out.
  //
  // Fused form (when the first and last passes give every thread whole
  // butterflies): each thread loads the R0 inputs of its first-pass
  // butterfly straight from HBM -- rows j + i Q0 of its column, twiddled by
  // w^(c j) * w^(c Q0 i) -- and runs that pass in registers; the last pass
  // of the inverse transform likewise writes straight to HBM.  That saves
  // two LDS round trips and their barriers.
  constexpr int R0 = plan_radix(H, 0), Q0 = H / R0;
  constexpr int NP = plan_passes(H), RL = plan_radix(H, NP - 1), QL = H / RL;
  // (measured: no gain for 4096-point columns, whose 512 threads run short
  // of registers when fused)
  constexpr bool FUSED = (2 * Q0) % T == 0 && (2 * QL) % T == 0 && NP >= 2 && 2 * (R0 + RL) <= T && H < 4096;
  const uint32_t cols[2] = {c0, c1};
  // w^(c i Q0) (i < R0) and w^(c i QL) (i < RL) of both columns, shared by
  // every thread of the column
  __shared__ cplx colf0[FUSED ? 2 * R0 : 1], colfL[FUSED ? 2 * RL : 1];
  if constexpr (FUSED) {
    constexpr int PER = 2 * Q0 / T;
    cplx v[PER][R0];
#pragma unroll
    for (int s = 0; s < PER; s++) {
      const int u = tid + s * T, b = u / Q0, j = u % Q0;
      const cplx *col = b ? col1 : col0;
#pragma unroll
      for (int i = 0; i < R0; i++) {
        const int row = j + i * Q0;
        v[s][i] = gload_stream(col + (((size_t)(row >> rs) * W << rs) | (row & rmask)));
      }
    }
    if (open_iteration()) return;
    if (tid < 2 * R0) {
      colf0[tid] = twiddle_m(A.twM, cols[tid / R0] * (uint32_t)((tid % R0) * Q0));
    } else if (tid < 2 * (R0 + RL)) {
      const int t = tid - 2 * R0;
      colfL[t] = twiddle_m(A.twM, cols[t / RL] * (uint32_t)((t % RL) * QL));
    }
    lds_barrier();  // the column loads stay in flight
#pragma unroll
    for (int s = 0; s < PER; s++) {
      const int u = tid + s * T, b = u / Q0, j = u % Q0;
      const cplx tj = twiddle_m(A.twM, cols[b] * (uint32_t)j);
#pragma unroll
      for (int i = 0; i < R0; i++) v[s][i] = cmul(v[s][i], i ? cmul(tj, colf0[b * R0 + i]) : tj);
      dft<R0, false>(v[s]);
#pragma unroll
      for (int i = 0; i < R0; i++) lds[b * LP + lpad(j * R0 + i)] = v[s][i];
    }
  } else {
    // Element b = tid + s T of column c takes w^(c tid) * w^(c s T).  All
    // of a thread's loads are issued before the first is used.
    static_assert(H % T == 0, "square sweep: H must be a multiple of T");
    constexpr int NE = H / T;
    cplx v0[NE], v1[NE];
#pragma unroll
    for (int st = 0; st < NE; st++) {
      const int i = tid + st * T;
      const size_t off = ((size_t)(i >> rs) * W << rs) | (i & rmask);
      v0[st] = gload_stream(col0 + off);
      v1[st] = gload_stream(col1 + off);
    }
    if (open_iteration()) return;
    const cplx t0 = twiddle_m(A.twM, c0 * (uint32_t)tid), t1 = twiddle_m(A.twM, c1 * (uint32_t)tid);
#pragma unroll
    for (int st = 0; st < NE; st++) {
      const int i = tid + st * T;
      const cplx w0 = cmul(t0, twiddle_m(A.twM, c0 * (uint32_t)(st * T)));
      const cplx w1 = cmul(t1, twiddle_m(A.twM, c1 * (uint32_t)(st * T)));
      lds[lpad(i)] = cmul(v0[st], w0);
      lds[LP + lpad(i)] = cmul(v1[st], w1);
    }
  }
  __syncthreads();
  // Each thread's Stockham twiddles, cached in registers for both transforms
  // (the inverse uses their conjugates), read from the plan's per-pass table.
  PassTwiddles<H, 2, T> ptw;
  ptw.init_table(A.ptH);
  if constexpr (FUSED) fft_lds_cached<H, 2, T, false, 1, R0>(lds, ptw.next);
  else fft_lds_cached<H, 2, T, false>(lds, ptw);
  if (g != 0) {
    // unrolled by 4 so that four iterations' table loads are in flight at once
    static_assert(H % T == 0, "square sweep: H must be a multiple of T");
#pragma unroll 4
    for (int st = 0; st < H / T; st++) {
      const int k2 = tid + st * T, k2b = H - 1 - k2;
      cplx a = lds[lpad(k2)], b = lds[LP + lpad(k2b)];
      pair_square(a, b, twiddle_m(A.twM, c0 + W * (uint32_t)k2));
      lds[lpad(k2)] = a;
      lds[LP + lpad(k2b)] = b;
    }
  } else {
    // column 0: k2 <-> H - k2 ; column W/2: k2 <-> H - 1 - k2
    for (int k2 = tid; k2 <= H / 2; k2 += T) {
      const int k2b = (H - k2) % H;
      cplx a = lds[lpad(k2)], b = lds[lpad(k2b)];
      pair_square(a, b, twiddle_m(A.twM, W * (uint32_t)k2));
      lds[lpad(k2)] = a;
      if (k2b != k2) lds[lpad(k2b)] = b;
    }
    for (int k2 = tid; k2 < (H + 1) / 2; k2 += T) {
      const int k2b = H - 1 - k2;
      cplx a = lds[LP + lpad(k2)], b = lds[LP + lpad(k2b)];
      pair_square(a, b, twiddle_m(A.twM, W / 2 + W * (uint32_t)k2));
      lds[LP + lpad(k2)] = a;
      if (k2b != k2) lds[LP + lpad(k2b)] = b;
    }
  }
  __syncthreads();
  if constexpr (FUSED) {
    constexpr int PER = 2 * QL / T;
    // the last pass's row twiddles w^(c j), loaded while the other passes run
    cplx tjL[PER];
#pragma unroll
    for (int s = 0; s < PER; s++) {
      const int u = tid + s * T;
      tjL[s] = twiddle_m(A.twM, cols[u / QL] * (uint32_t)(u % QL));
    }
    fft_lds_cached<H, 2, T, true, 0, 1, NP - 1>(lds, ptw);
    // last inverse pass: rows j + i QL of each column, straight to HBM
    constexpr int NSL = plan_span(H, NP - 1);
    static_assert(NSL == QL, "last pass spans the whole column");
    cplx v[PER][RL];
#pragma unroll
    for (int s = 0; s < PER; s++) {
      const int u = tid + s * T, b = u / QL, j = u % QL;
#pragma unroll
      for (int i = 0; i < RL; i++) {
        cplx a = lds[b * LP + lpad(j + i * QL)];
        if (i > 0) a = cmulc(a, twiddle_level<NP - 1>(ptw).w[s][i - 1]);
        v[s][i] = a;
      }
    }
    const StreamOut out(A.x, A.n / 2);
#pragma unroll
    for (int s = 0; s < PER; s++) {
      const int u = tid + s * T, b = u / QL, j = u % QL;
      const uint32_t col = cols[b] << rs;
      dft<RL, true>(v[s]);
      const cplx tj = tjL[s];
#pragma unroll
      for (int i = 0; i < RL; i++) {
        const int row = j + i * QL;
        const cplx w = i ? cmul(tj, colfL[b * RL + i]) : tj;
        out.store(col + ((((uint32_t)row >> rs) * W << rs) | (row & rmask)), cmulc(v[s][i], w));
      }
    }
  } else {
    fft_lds_cached<H, 2, T, true>(lds, ptw);
    const cplx t0 = twiddle_m(A.twM, c0 * (uint32_t)tid), t1 = twiddle_m(A.twM, c1 * (uint32_t)tid);
    const StreamOut out(A.x, A.n / 2);
#pragma unroll 2
    for (int st = 0; st < H / T; st++) {
      const int i = tid + st * T;
      const uint32_t off = (((uint32_t)i >> rs) * W << rs) | (i & rmask);
      const cplx w0 = cmul(t0, twiddle_m(A.twM, c0 * (uint32_t)(st * T)));
      const cplx w1 = cmul(t1, twiddle_m(A.twM, c1 * (uint32_t)(st * T)));
      out.store((c0 << rs) + off, cmulc(lds[lpad(i)], w0));
      out.store((c1 << rs) + off, cmulc(lds[LP + lpad(i)], w1));
    }
  }
}

// ------------------------------------------------- persistent square sweep
// For columns too long for two workgroups per CU (H = 4096: 139,776 bytes
// of LDS), where a grid of W/2 column pairs runs in several rounds: a grid
// of at most one round, each workgroup taking pairs g, g + gridDim.x, ...
// and loading its next pair into registers while it transforms the
// current one, so that the loads of the next pair and the stores of the
// current one overlap the transforms instead of following them.  To leave
// registers for the pair in flight, the second pass of the column
// transforms reads its twiddles from a copy of its block of the per-pass
// table in LDS instead of holding them in registers; otherwise the
// arithmetic is square_sweep's (non-fused form) term for term, so the
// results are the same integers.
template <int H, int T>
__global__ __launch_bounds__(T, 1) void square_sweep_pipe(const PlanArgs *__restrict__ args) {
  const PlanArgs &A = args[blockIdx.y];
  static_assert(H % T == 0, "square sweep: H must be a multiple of T");
  constexpr int NE = H / T;
  constexpr int LP = lpitch(H);
  __shared__ cplx lds[2 * LP];
  constexpr int S1 = plan_radix(H, 0), S2 = S1 * plan_radix(H, 1);
  constexpr int NT1 = (plan_radix(H, 1) - 1) * S1;  // twiddles of the second pass
  __shared__ cplx tw1[NT1];
  const int tid = threadIdx.x;
  const uint32_t W = A.W, npairs = W / 2;
  const uint32_t rs = A.rshift, rmask = (1u << rs) - 1;
  uint32_t g = pair_of_workgroup(blockIdx.x, gridDim.x, A.rshift);
  if (g >= npairs) return;
  auto partner = [&](uint32_t gg) { return gg == 0 ? W / 2 : W - gg; };
  cplx n0[NE], n1[NE];  // the pair being loaded
  auto fetch = [&](uint32_t gg, int tid) {
    const cplx *col0 = A.x + ((size_t)gg << rs), *col1 = A.x + ((size_t)partner(gg) << rs);
#pragma unroll
    for (int st = 0; st < NE; st++) {
      const int i = tid + st * T;
      const size_t off = ((size_t)(i >> rs) * W << rs) | (i & rmask);
      n0[st] = gload_stream(col0 + off);
      n1[st] = gload_stream(col1 + off);
    }
  };
  fetch(g, tid);
  // opens the iteration for the carry sweep that follows (workgroup 0, which
  // takes pair 0 first), or stops a finished batched test
  const bool finished = A.remaining && *A.remaining == 0;
  if (blockIdx.x == 0 && tid == 0) {
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
  for (int i = tid; i < NT1; i += T) tw1[i] = gload(A.ptH + (S1 - 1) + i);
  PassTwiddles<H, 2, T, 2, S2> ptw;  // the third pass's, in registers
  ptw.init_table(A.ptH);
  const StreamOut out(A.x, A.n / 2);
  for (;;) {
    // the thread index, opaque to the compiler in each pass of the loop
    // (as in fft_lds_3pass_in_loop)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t c0 = g, c1 = partner(g);
    const uint32_t gn = g + gridDim.x;
    {
      const cplx t0 = twiddle_m(A.twM, c0 * (uint32_t)tid), t1 = twiddle_m(A.twM, c1 * (uint32_t)tid);
#pragma unroll
      for (int st = 0; st < NE; st++) {
        const int i = tid + st * T;
        const cplx w0 = cmul(t0, twiddle_m(A.twM, c0 * (uint32_t)(st * T)));
        const cplx w1 = cmul(t1, twiddle_m(A.twM, c1 * (uint32_t)(st * T)));
        lds[lpad(i)] = cmul(n0[st], w0);
        lds[LP + lpad(i)] = cmul(n1[st], w1);
      }
    }
    if (gn < npairs) fetch(gn, tid);
    __syncthreads();
    fft_lds_3pass_in_loop<H, 2, T, false>(lds, tw1, ptw);
    if (g != 0) {
#pragma unroll 4
      for (int st = 0; st < H / T; st++) {
        const int k2 = tid + st * T, k2b = H - 1 - k2;
        cplx a = lds[lpad(k2)], b = lds[LP + lpad(k2b)];
        pair_square(a, b, twiddle_m(A.twM, c0 + W * (uint32_t)k2));
        lds[lpad(k2)] = a;
        lds[LP + lpad(k2b)] = b;
      }
    } else {
      // column 0: k2 <-> H - k2 ; column W/2: k2 <-> H - 1 - k2
      for (int k2 = tid; k2 <= H / 2; k2 += T) {
        const int k2b = (H - k2) % H;
        cplx a = lds[lpad(k2)], b = lds[lpad(k2b)];
        pair_square(a, b, twiddle_m(A.twM, W * (uint32_t)k2));
        lds[lpad(k2)] = a;
        if (k2b != k2) lds[lpad(k2b)] = b;
      }
      for (int k2 = tid; k2 < (H + 1) / 2; k2 += T) {
        const int k2b = H - 1 - k2;
        cplx a = lds[LP + lpad(k2)], b = lds[LP + lpad(k2b)];
        pair_square(a, b, twiddle_m(A.twM, W / 2 + W * (uint32_t)k2));
        lds[LP + lpad(k2)] = a;
        if (k2b != k2) lds[LP + lpad(k2b)] = b;
      }
    }
    __syncthreads();
    fft_lds_3pass_in_loop<H, 2, T, true>(lds, tw1, ptw);
    {
      const cplx t0 = twiddle_m(A.twM, c0 * (uint32_t)tid), t1 = twiddle_m(A.twM, c1 * (uint32_t)tid);
#pragma unroll 2
      for (int st = 0; st < H / T; st++) {
        const int i = tid + st * T;
        const uint32_t off = (((uint32_t)i >> rs) * W << rs) | (i & rmask);
        const cplx w0 = cmul(t0, twiddle_m(A.twM, c0 * (uint32_t)(st * T)));
        const cplx w1 = cmul(t1, twiddle_m(A.twM, c1 * (uint32_t)(st * T)));
        out.store((c0 << rs) + off, cmulc(lds[lpad(i)], w0));
        out.store((c1 << rs) + off, cmulc(lds[LP + lpad(i)], w1));
      }
    }
    if (gn >= npairs) break;
    g = gn;
    __syncthreads();  // every thread has read this pair from LDS
  }
}

// ------------------------------------------------------------- carry sweep
// H/R workgroups, workgroup k owning block k of R lines.  Each publishes the
// carries out of its block before it waits for those into it, so a wait is
// on a lower-numbered block, which the hardware dispatches first -- except
// block 0, which waits for the last block's carries (the cyclic wrap
// 2^p = 1) after publishing its own; no block waits on block 0 finishing.
// Should a wait ever outlast kSpinLimit polls, the iteration is reported as
// failed (error 3.0) rather than hanging the device.
// The carry sweep keeps a copy of its line plan's per-pass twiddles (16 W
// bytes of LDS, in place of the quarter-wave table) for its wave-local FFT
// passes when that costs no workgroup per CU (160 KiB of LDS and 2048
// threads per CU); the LDS bytes below are the kernel's own arrays.
template <int W, int R, int T>
constexpr bool carry_pass_table() {
  if constexpr (!wave_local_ok<W, R, T>()) {
    return false;
  } else {
    constexpr int kLds = 163840, cap = 2048 / T;
    constexpr int base = R * lpitch(W) * 16 + (W / 4 + 1) * 8 + 2 * R * 4 + 4 * R * 8 + (T / 64) * 4;
    constexpr int with = base - (W / 4 + 1) * 8 + 16 * W;  // the table replaces qc
    constexpr int a = kLds / base < cap ? kLds / base : cap;
    constexpr int b = kLds / with < cap ? kLds / with : cap;
    return with <= kLds && b >= a;
  }
}

// The table carry_sweep<W, R, T> reads (W - 1 roots of its line plan, from
// the quarter-wave table of order W), built once per plan.
template <int W, int R, int T>
__global__ __launch_bounds__(256) void line_pass_table(const double *qc, cplx *out) {
  build_wave_table<W, 256, auto_plan_uniform<W, R, T>()>(out, QuarterRoots<W>{qc});
}

// Workgroups of carry_sweep<W, R, T> that the LDS of a CU holds, at most 2.
// HIP passes the second __launch_bounds__ argument on as the minimum waves
// per SIMD (amdgpu_waves_per_eu), which bounds the registers per lane; with
// T = 256 (one wave per SIMD per workgroup) the value is also workgroups per
// CU, so a sweep whose lines fill over half the LDS is not held to half the
// register file.  For T = 64 and 128 the value asks for more waves per SIMD
// than one workgroup brings (a register bound only; the LDS still limits
// how many workgroups share a CU).
template <int W, int R, int T>
constexpr int carry_groups_per_cu() {
  constexpr bool tab = carry_pass_table<W, R, T>();
  constexpr int bytes = R * lpitch(W) * 16 + (tab ? 8 : (W / 4 + 1) * 8) + 2 * R * 4 + 4 * R * 8 + (T / 64) * 4 +
                        (tab ? 16 * W : 16);
  return 163840 / bytes >= 2 ? 2 : 1;
}

// Shape of carry_sweep's fused end passes: the wave-local plan (UNI, NP
// passes, first radix R0, last radix RL) and whether both end passes map
// onto the threads' load and store positions (M0 first-pass butterflies
// per thread).
template <int W, int R, int T>
struct CarryEnds {
  static constexpr bool wl = wave_local_ok<W, R, T>();
  static constexpr bool UNI = auto_plan_uniform<W, R, T>();
  static constexpr int NP = wl ? wave_passes<UNI>(W) : 0;
  static constexpr int R0 = wl ? wave_radix<UNI>(W, 0) : 1;
  static constexpr int RL = wl ? wave_radix<UNI>(W, NP - 1) : 1;
  static constexpr int TA = T / R;
  static constexpr bool ok = wl && NP >= 2 && T % R == 0 && (R * W) % T == 0 && (W / R0) % TA == 0 &&
                             (W / RL) % TA == 0;
  static constexpr int M0 = ok ? W / R0 / TA : 1;
};

template <int W, int R, int T>
__global__ __launch_bounds__(T, (T >= 512 ? 1 : carry_groups_per_cu<W, R, T>())) void carry_sweep(
    const PlanArgs *__restrict__ args) {
  const PlanArgs &A = args[blockIdx.y];
  stagger_start(A.stagger_ca);
  constexpr int LP = lpitch(W);
  constexpr int CPT = (W + T - 1) / T;  // columns per thread
  __shared__ cplx lds[R * LP];
  __shared__ double qc[carry_pass_table<W, R, T>() ? 1 : W / 4 + 1];
  __shared__ int fb[2 * R];  // weight factors shared by every column
  __shared__ double wb[2 * R], iwb[2 * R];
  __shared__ float red[T / 64];
  constexpr bool TAB = carry_pass_table<W, R, T>();
  __shared__ cplx ptab[TAB ? W : 1];  // wave-local pass twiddles (copy of ptW)
  const int tid = threadIdx.x;
  // Fused end passes (wave-local plans whose first and last passes map onto
  // the threads' load and store positions): thread t holds elements
  // e = t + s T, i.e. line r = t mod R and columns a = t / R + TA s
  // (TA = T / R).  The inverse transform's first pass (radix R0, inputs
  // a = j + Q0 i) then runs in registers straight from the loads, and the
  // forward transform's last pass (radix RL, outputs a = j + QL i) goes
  // straight to HBM, each saving an LDS round trip of the whole block.
  using Fused = CarryEnds<W, R, T>;
  constexpr bool FUSE = Fused::ok;
  const uint32_t H = A.H;
  const uint32_t G = H / R;
  const uint32_t blk = blockIdx.x;

  const uint32_t b0 = blk * R;
  const QuarterRoots<W> tw{qc};

  // load R lines (the square sweep has already undone the column twiddle):
  // every load of the thread is issued first, then the small tables, so
  // their latencies overlap.  Workgroups of 512 or more threads (one per
  // CU, registers to spare) also fetch here the per-column weights the
  // carry step uses after the inverse transform.
  constexpr bool PF = T >= 512;
  int fa_c[PF ? CPT : 1];
  double iwa_c[PF ? CPT : 1], wa_c[PF ? CPT : 1];
  {
    constexpr int NE = (R * W + T - 1) / T;
    cplx v[NE];
#pragma unroll
    for (int s = 0; s < NE; s++) {
      const int e = tid + s * T;
      if ((R * W) % T == 0 || e < R * W) v[s] = gload_stream(A.x + (size_t)b0 * W + e);
    }
    // checked once the line loads are in flight (reads only so far)
    if (!A.st->run) return;  // a finished batched test
    if constexpr (TAB) {
      for (int i = tid; i < W - 1; i += T) ptab[i] = gload(A.ptW + i);
    } else {
      load_quarter<W, T>(qc, A.qcW);
    }
#pragma unroll
    for (int c = 0; c < (PF ? CPT : 0); c++) {
      const int a = tid + c * T;
      if (W % T == 0 || a < W) {
        fa_c[c] = gload(A.fA + a);
        iwa_c[c] = gload(A.iwAs + a);
        wa_c[c] = gload(A.wA + a);
      }
    }
    if (tid < 2 * R) {
      const uint32_t i = 2 * blk * R + tid;
      fb[tid] = gload(A.fB + i);
      wb[tid] = gload(A.wB + i);
      iwb[tid] = gload(A.iwB + i);
    }
    if constexpr (FUSE) {
      // first inverse pass (span 1, no twiddles): butterfly j of line r
      // takes columns j + Q0 i, which this thread loaded as v[s0 + M0 i]
      constexpr int R0 = Fused::R0, M0 = Fused::M0, TA = T / R;
      const int r = tid % R, j0 = tid / R;
#pragma unroll
      for (int s0 = 0; s0 < M0; s0++) {
        cplx x[R0];
#pragma unroll
        for (int i = 0; i < R0; i++) x[i] = v[s0 + M0 * i];
        dft<R0, true>(x);
        const int j = j0 + TA * s0;
#pragma unroll
        for (int i = 0; i < R0; i++) lds[r * LP + lpad(j * R0 + i)] = x[i];
      }
    } else {
#pragma unroll
      for (int s = 0; s < NE; s++) {
        const int e = tid + s * T;
        if ((R * W) % T == 0 || e < R * W) lds[(e % R) * LP + lpad(e / R)] = v[s];
      }
    }
  }
  const uint32_t epoch = A.st->epoch;
  const uint32_t m2_digit = A.st->m2_digit;
  const double dmod = (double)A.st->m2_value;
  __syncthreads();
  if constexpr (FUSE) {
    if constexpr (TAB) fft_lds_wave<W, R, T, true, Fused::UNI, 1, Fused::R0>(lds, WaveTable{ptab});
    else fft_lds_wave<W, R, T, true, Fused::UNI, 1, Fused::R0>(lds, tw);
    __syncthreads();
  } else {
    fft_lds_auto<W, R, T, true, TAB>(lds, tw, ptab);
  }

  // Round, subtract, carry along the 2R consecutive words of each column,
  // in exact FP64 (balance_f64).  The balanced digits go back into the
  // column's own LDS slots until the carry-in arrives.
  // Word 2 H a + i of the column has weight exponent f = (fa + fb[i]) mod n
  // (wrap: the sum reached n, a factor 2 in the weight) and width q + 1
  // below b, q above; the factors of 2 are exact ldexp scalings.
  const int q = (int)A.q;
  double emax = 0.0;
  // 512-thread workgroups keep the balanced digits of their columns in
  // registers until the carry-in arrives; the others park them in LDS
  double dg[PF ? CPT : 1][PF ? 2 * R : 1];
  // and the wrap and big flags of each word (bits 2i and 2i + 1)
  static_assert(2 * R <= 32, "two flag bits per word of the chain");
  uint64_t flags[PF ? CPT : 1];
  auto round_column = [&](const int a, const int c, const int fa, const double iwa) {
    // m2i < 2R: this chain takes the -2, at word m2i (one chain of all)
    const uint32_t m2i = m2_digit - 2u * ((uint32_t)a * H + b0);
    double carry = 0.0;
    uint64_t fl = 0;
    auto chain = [&](auto with_m2) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const cplx v = lds[r * LP + lpad(a)];
        double d[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int i = 2 * r + e;
          const int f0 = fa + fb[i];
          const int wrap = f0 >= (int)A.n;
          const int big = f0 - (wrap ? (int)A.n : 0) < (int)A.b;
          if constexpr (PF) fl |= ((uint64_t)wrap | ((uint64_t)big << 1)) << (2 * i);
          const double t = ldexp(mul3(e ? v.y : v.x, iwa, iwb[i]), wrap);
          double rt = rint(t);
          emax = fmax(emax, fabs(t - rt));
          if constexpr (decltype(with_m2)::value) rt -= (uint32_t)i == m2i ? dmod : 0.0;
          d[e] = balance_f64(rt + carry, q + big, carry);
        }
        if constexpr (PF) {
          dg[c][2 * r] = d[0];
          dg[c][2 * r + 1] = d[1];
        } else {
          lds[r * LP + lpad(a)] = {d[0], d[1]};
        }
      }
    };
    if (m2i < 2u * R) chain(std::true_type{});
    else chain(std::false_type{});
    if constexpr (PF) flags[c] = fl;
    const bool fits = fabs(carry) < 2147483648.0;
    if (!fits) emax = 2.0;  // carry overflow: flagged as a failed iteration
    const int32_t c32 = fits ? (int32_t)carry : 0;
    const uint64_t word = ((uint64_t)epoch << 32) | (uint32_t)c32;
    __hip_atomic_store(&A.shuttle[(size_t)blk * W + a], word, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  };
  if constexpr (PF) {
#pragma unroll
    for (int c = 0; c < CPT; c++)
      if (W % T == 0 || tid + c * T < W) round_column(tid + c * T, c, fa_c[c], iwa_c[c]);
  } else {
    for (int a = tid; a < W; a += T) round_column(a, 0, gload(A.fA + a), gload(A.iwAs + a));
  }

  // carry-in from the previous block (block 0: the last block, one column
  // to the left), propagated through the whole chain, then reweight
  auto carry_in_column = [&](const int a, const int c, const int fa, const double wa) {
    const size_t src = blk == 0 ? (size_t)(G - 1) * W + (a + W - 1) % W : (size_t)(blk - 1) * W + a;
    uint64_t word;
    uint32_t spins = 0;
    for (;;) {
      word = __hip_atomic_load(&A.shuttle[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(word >> 32) == epoch) break;
      if (++spins > kSpinLimit) {  // never hang: flag the iteration as failed
        emax = 3.0;
        word = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    // What is left after one word is about sqrt(N)/12 whatever the digit
    // size, which exceeds half a digit at low bits/word, so the carry runs
    // through the whole chain; anything left after 2R words (0 in
    // practice) stays in the last word.
    double cc = (double)(int32_t)(uint32_t)word;
#pragma unroll
    for (int r = 0; r < R; r++) {
      cplx dv;
      if constexpr (PF) dv = {dg[c][2 * r], dg[c][2 * r + 1]};
      else dv = lds[r * LP + lpad(a)];
      double x[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int i = 2 * r + e;
        int wrap, width;
        if constexpr (PF) {
          wrap = (int)(flags[c] >> (2 * i)) & 1;
          width = q + ((int)(flags[c] >> (2 * i + 1)) & 1);
        } else {
          const int f0 = fa + fb[i];
          wrap = f0 >= (int)A.n;
          width = q + (f0 - (wrap ? (int)A.n : 0) < (int)A.b);
        }
        double d = balance_f64((e ? dv.y : dv.x) + cc, width, cc);
        if (i == 2 * R - 1) d += ldexp(cc, width);
        x[e] = ldexp(mul3(d, wa, wb[i]), -wrap);
      }
      lds[r * LP + lpad(a)] = {x[0], x[1]};
    }
  };
  if constexpr (PF) {
#pragma unroll
    for (int c = 0; c < CPT; c++)
      if (W % T == 0 || tid + c * T < W) carry_in_column(tid + c * T, c, fa_c[c], wa_c[c]);
  } else {
    for (int a = tid; a < W; a += T) carry_in_column(a, 0, gload(A.fA + a), gload(A.wA + a));
  }
  __syncthreads();
  if constexpr (FUSE) {
    constexpr int NPW = Fused::NP;
    if constexpr (TAB) fft_lds_wave<W, R, T, false, Fused::UNI, 0, 1, NPW - 1>(lds, WaveTable{ptab});
    else fft_lds_wave<W, R, T, false, Fused::UNI, 0, 1, NPW - 1>(lds, tw);
    __syncthreads();
    // last forward pass (span QL): butterfly j of line r reads columns
    // j + QL i and writes the same columns, element (a R + r) of the block:
    // HBM index b0 W + t + T (s + PL i)
    constexpr int RL = Fused::RL, QL = W / RL, TA = T / R, PL = QL / TA;
    // an opaque copy of tid keeps the compiler from holding the load
    // phase's LDS addresses in registers across the whole kernel
    int t2 = tid;
    asm volatile("" : "+v"(t2));
    const int r = t2 % R, j0 = t2 / R;
    const StreamOut out(A.x, A.n / 2);
#pragma unroll
    for (int s = 0; s < PL; s++) {
      const int j = j0 + TA * s;
      cplx x[RL];
#pragma unroll
      for (int i = 0; i < RL; i++) {
        cplx a = lds[r * LP + lpad(j + QL * i)];
        if (i > 0) {
          cplx w;
          if constexpr (TAB) w = ptab[QL - 1 + (i - 1) * QL + j];
          else w = tw(i * j);
          a = cmul(a, w);
        }
        x[i] = a;
      }
      dft<RL, false>(x);
#pragma unroll
      for (int i = 0; i < RL; i++) out.store(b0 * W + tid + T * (s + PL * i), x[i]);
    }
  } else {
    fft_lds_auto<W, R, T, false, TAB>(lds, tw, ptab);
    constexpr int NE = (R * W + T - 1) / T;
    cplx v[NE];
    // an opaque copy of tid keeps the compiler from holding the load
    // phase's LDS addresses in registers across the whole kernel
    int t2 = tid;
    asm volatile("" : "+v"(t2));
#pragma unroll
    for (int s = 0; s < NE; s++) {
      const int e = t2 + s * T;
      if ((R * W) % T == 0 || e < R * W) v[s] = lds[(e % R) * LP + lpad(e / R)];
    }
    const StreamOut out(A.x, A.n / 2);
#pragma unroll
    for (int s = 0; s < NE; s++) {
      const int e = tid + s * T;
      if ((R * W) % T == 0 || e < R * W) out.store(b0 * W + e, v[s]);
    }
  }
  const float m = block_max<T>((float)emax, red);
  if (tid == 0) {
    atomicMax(&A.err[(epoch - 1) & A.err_mask], __float_as_uint(m));
    // no carry workgroup reads `remaining`: the next square sweep does
    if (blk == 0 && A.remaining) *A.remaining -= 1;
  }
}

// The 512-thread form of carry_sweep<W, R, 256> (more waves per CU where a
// schedule has no more line blocks than one round of it holds).  It reads
// the same line table, or none, so a context built for one form can launch
// either.
template <int W, int R>
const void *carry_wide_form() {
  static_assert(carry_pass_table<W, R, 256>() == carry_pass_table<W, R, 512>() &&
                    (!carry_pass_table<W, R, 256>() || auto_plan_uniform<W, R, 256>() == auto_plan_uniform<W, R, 512>()),
                "the 512-thread carry sweep must read the 256-thread form's line table");
  return (const void *)carry_sweep<W, R, 512>;
}

// ------------------------------------------------- persistent carry sweep
// For lengths whose H / R line blocks take more than one round of
// carry_sweep at T >= 512 (one workgroup per CU): one round of workgroups,
// workgroup w taking blocks w, w + gridDim.x, ... and loading its next
// block into registers while it works on the current one, so that the
// loads of the next block and the stores of the current one overlap the
// transforms and the carries.  The arithmetic is carry_sweep's term for
// term.  Every wait ends: a block publishes its own carries before it
// waits for its predecessor's, so the chain of waits behind any block
// steps back one block and one pass over the grid at a time and ends at a
// block that waits on nothing.  The exception is block 0, whose carry-in
// comes from the last block: workgroup 0 would hold it (and so its later
// blocks, and the blocks after those) for the whole sweep.  So workgroup 0
// publishes block 0's carries first, parks its balanced digits in block 0's
// own words of the signal (which no other workgroup reads or writes in
// this launch), and finishes the block -- carry-in, reweight, forward
// transform, store -- after its last other block.
template <int W, int R, int T>
__global__ __launch_bounds__(T, 1) void carry_sweep_pipe(const PlanArgs *__restrict__ args) {
  static_assert(T >= 512, "carry_sweep_pipe keeps the per-column weights in registers");
  const PlanArgs &A = args[blockIdx.y];
  constexpr int LP = lpitch(W);
  constexpr int CPT = (W + T - 1) / T;     // columns per thread
  constexpr int NE = (R * W + T - 1) / T;  // block elements per thread
  constexpr bool TAB = carry_pass_table<W, R, T>();
  using Fused = CarryEnds<W, R, T>;
  constexpr bool FUSE = Fused::ok;
  __shared__ cplx lds[R * LP];
  __shared__ double qc[TAB ? 1 : W / 4 + 1];
  __shared__ int fb[2 * R];
  __shared__ double wb[2 * R], iwb[2 * R];
  __shared__ float red[T / 64];
  __shared__ cplx ptab[TAB ? W : 1];
  const uint32_t H = A.H;
  const uint32_t G = H / R;
  const uint32_t first = blockIdx.x, stride = gridDim.x;
  if (first >= G) return;
  const QuarterRoots<W> tw{qc};
  // thread t holds elements e = t + s T of a block: line e mod R, column e / R
  cplx v[NE];
  auto fetch = [&](uint32_t blk, int t) {
    const cplx *src = A.x + (size_t)blk * R * W;
#pragma unroll
    for (int s = 0; s < NE; s++) {
      const int e = t + s * T;
      if ((R * W) % T == 0 || e < R * W) v[s] = gload_stream(src + e);
    }
  };
  fetch(first, threadIdx.x);
  if (!A.st->run) return;  // a finished batched test
  const int tid0 = threadIdx.x;
  if constexpr (TAB) {
    for (int i = tid0; i < W - 1; i += T) ptab[i] = gload(A.ptW + i);
  } else {
    load_quarter<W, T>(qc, A.qcW);
  }
  int fa_c[CPT];
  double iwa_c[CPT], wa_c[CPT];
#pragma unroll
  for (int c = 0; c < CPT; c++) {
    const int a = tid0 + c * T;
    if (W % T == 0 || a < W) {
      fa_c[c] = gload(A.fA + a);
      iwa_c[c] = gload(A.iwAs + a);
      wa_c[c] = gload(A.wA + a);
    }
  }
  const uint32_t epoch = A.st->epoch;
  const uint32_t m2_digit = A.st->m2_digit;
  const double dmod = (double)A.st->m2_value;
  const int q = (int)A.q;
  double emax = 0.0;
  // the balanced digits go back into their column's own LDS slots until the
  // carry-in arrives (registers hold the next block's loads)
  uint64_t flags[CPT];  // wrap and big flags of each word (bits 2i, 2i + 1)
  static_assert(2 * R <= 32, "two flag bits per word of the chain");
  const StreamOut out(A.x, A.n / 2);

  // the weight factors of block blk's rows, shared by every column
  auto row_weights = [&](uint32_t blk, int t) {
    if (t < 2 * R) {
      const uint32_t i = 2 * blk * R + t;
      fb[t] = gload(A.fB + i);
      wb[t] = gload(A.wB + i);
      iwb[t] = gload(A.iwB + i);
    }
  };
  // round, subtract, carry along column a's 2R words; publish the carry-out
  auto round_column = [&](const uint32_t blk, const int a, const int c) {
    const uint32_t b0 = blk * R;
    const uint32_t m2i = m2_digit - 2u * ((uint32_t)a * H + b0);
    double carry = 0.0;
    uint64_t fl = 0;
    auto chain = [&](auto with_m2) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const cplx x = lds[r * LP + lpad(a)];
        double d[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int i = 2 * r + e;
          const int f0 = fa_c[c] + fb[i];
          const int wrap = f0 >= (int)A.n;
          const int big = f0 - (wrap ? (int)A.n : 0) < (int)A.b;
          fl |= ((uint64_t)wrap | ((uint64_t)big << 1)) << (2 * i);
          const double t = ldexp(mul3(e ? x.y : x.x, iwa_c[c], iwb[i]), wrap);
          double rt = rint(t);
          emax = fmax(emax, fabs(t - rt));
          if constexpr (decltype(with_m2)::value) rt -= (uint32_t)i == m2i ? dmod : 0.0;
          d[e] = balance_f64(rt + carry, q + big, carry);
        }
        lds[r * LP + lpad(a)] = {d[0], d[1]};
      }
    };
    if (m2i < 2u * R) chain(std::true_type{});
    else chain(std::false_type{});
    flags[c] = fl;
    const bool fits = fabs(carry) < 2147483648.0;
    if (!fits) emax = fmax(emax, 2.0);  // carry overflow: flagged as a failed iteration
    const int32_t c32 = fits ? (int32_t)carry : 0;
    const uint64_t word = ((uint64_t)epoch << 32) | (uint32_t)c32;
    __hip_atomic_store(&A.shuttle[(size_t)blk * W + a], word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  // carry-in from the previous block (block 0: the last block, one column
  // to the left) through the whole chain, then reweight into LDS
  auto carry_in_column = [&](const uint32_t blk, const int a, const int c) {
    const size_t src = blk == 0 ? (size_t)(G - 1) * W + (a + W - 1) % W : (size_t)(blk - 1) * W + a;
    uint64_t word;
    uint32_t spins = 0;
    for (;;) {
      word = __hip_atomic_load(&A.shuttle[src], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((uint32_t)(word >> 32) == epoch) break;
      if (++spins > kSpinLimit) {  // never hang: flag the iteration as failed
        emax = fmax(emax, 3.0);
        word = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    double cc = (double)(int32_t)(uint32_t)word;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const cplx dv = lds[r * LP + lpad(a)];
      double x[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int i = 2 * r + e;
        const int wrap = (int)(flags[c] >> (2 * i)) & 1;
        const int width = q + ((int)(flags[c] >> (2 * i + 1)) & 1);
        double d = balance_f64((e ? dv.y : dv.x) + cc, width, cc);
        if (i == 2 * R - 1) d += ldexp(cc, width);
        x[e] = ldexp(mul3(d, wa_c[c], wb[i]), -wrap);
      }
      lds[r * LP + lpad(a)] = {x[0], x[1]};
    }
  };
  // forward transform of the block in LDS and its store to HBM
  auto forward_store = [&](const uint32_t blk) {
    const uint32_t b0 = blk * R;
    if constexpr (FUSE) {
      constexpr int NPW = Fused::NP;
      if constexpr (TAB) fft_lds_wave<W, R, T, false, Fused::UNI, 0, 1, NPW - 1, true>(lds, WaveTable{ptab});
      else fft_lds_wave<W, R, T, false, Fused::UNI, 0, 1, NPW - 1, true>(lds, tw);
      __syncthreads();
      constexpr int RL = Fused::RL, QL = W / RL, TA = T / R, PL = QL / TA;
      int t2 = threadIdx.x;
      asm volatile("" : "+v"(t2));
      const int r = t2 % R, j0 = t2 / R;
#pragma unroll
      for (int s = 0; s < PL; s++) {
        const int j = j0 + TA * s;
        cplx x[RL];
#pragma unroll
        for (int i = 0; i < RL; i++) {
          cplx a = lds[r * LP + lpad(j + QL * i)];
          if (i > 0) {
            cplx w;
            if constexpr (TAB) w = ptab[QL - 1 + (i - 1) * QL + j];
            else w = tw(i * j);
            a = cmul(a, w);
          }
          x[i] = a;
        }
        dft<RL, false>(x);
#pragma unroll
        for (int i = 0; i < RL; i++) out.store(b0 * W + t2 + T * (s + PL * i), x[i]);
      }
    } else {
      fft_lds_auto<W, R, T, false, TAB, true>(lds, tw, ptab);
      int t2 = threadIdx.x;
      asm volatile("" : "+v"(t2));
#pragma unroll
      for (int s = 0; s < NE; s++) {
        const int e = t2 + s * T;
        if ((R * W) % T == 0 || e < R * W) out.store(b0 * W + e, lds[(e % R) * LP + lpad(e / R)]);
      }
    }
  };

  for (uint32_t blk = first;;) {
    // the thread index, opaque to the compiler in each pass of the loop, so
    // that the passes' LDS offsets are not all hoisted out of it
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const uint32_t next = blk + stride;
    __syncthreads();  // the previous block's LDS reads are done
    row_weights(blk, t);
    // the block into LDS; when fused, the first inverse pass runs in
    // registers on the way (butterfly j of line r takes columns j + Q0 i,
    // which this thread holds as v[s0 + M0 i])
    if constexpr (FUSE) {
      constexpr int R0 = Fused::R0, M0 = Fused::M0, TA = T / R;
      const int r = t % R, j0 = t / R;
#pragma unroll
      for (int s0 = 0; s0 < M0; s0++) {
        cplx x[R0];
#pragma unroll
        for (int i = 0; i < R0; i++) x[i] = v[s0 + M0 * i];
        dft<R0, true>(x);
        const int j = j0 + TA * s0;
#pragma unroll
        for (int i = 0; i < R0; i++) lds[r * LP + lpad(j * R0 + i)] = x[i];
      }
    } else {
#pragma unroll
      for (int s = 0; s < NE; s++) {
        const int e = t + s * T;
        if ((R * W) % T == 0 || e < R * W) lds[(e % R) * LP + lpad(e / R)] = v[s];
      }
    }
    if (next < G) fetch(next, t);
    __syncthreads();
    if constexpr (FUSE) {
      if constexpr (TAB) fft_lds_wave<W, R, T, true, Fused::UNI, 1, Fused::R0, -1, true>(lds, WaveTable{ptab});
      else fft_lds_wave<W, R, T, true, Fused::UNI, 1, Fused::R0, -1, true>(lds, tw);
      __syncthreads();
    } else {
      fft_lds_auto<W, R, T, true, TAB, true>(lds, tw, ptab);
    }
    {
      int t3 = threadIdx.x;
      asm volatile("" : "+v"(t3));
#pragma unroll
      for (int c = 0; c < CPT; c++)
        if (W % T == 0 || t3 + c * T < W) round_column(blk, t3 + c * T, c);
    }
    if (blk == 0) {
      // park block 0's digits in its own words: thread t's column a, line r
      // at element a R + r, read back by the same thread after the loop
#pragma unroll
      for (int c = 0; c < CPT; c++) {
        const int a = t + c * T;
        if (W % T == 0 || a < W) {
#pragma unroll
          for (int r = 0; r < R; r++) gstore(A.x + (size_t)a * R + r, lds[r * LP + lpad(a)]);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < CPT; c++)
        if (W % T == 0 || t + c * T < W) carry_in_column(blk, t + c * T, c);
      __syncthreads();
      forward_store(blk);
    }
    if (next >= G) break;
    blk = next;
  }
  if (first == 0) {
    // block 0: its digits back (the words this thread parked), its rows'
    // weights and flags, the last block's carries, then the forward half
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    __syncthreads();
    row_weights(0, t);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const int a = t + c * T;
      if (W % T == 0 || a < W) {
        uint64_t fl = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
          lds[r * LP + lpad(a)] = gload(A.x + (size_t)a * R + r);
#pragma unroll
          for (int e = 0; e < 2; e++) {
            const int i = 2 * r + e;
            const int f0 = fa_c[c] + fb[i];
            const int wrap = f0 >= (int)A.n;
            const int big = f0 - (wrap ? (int)A.n : 0) < (int)A.b;
            fl |= ((uint64_t)wrap | ((uint64_t)big << 1)) << (2 * i);
          }
        }
        flags[c] = fl;
        carry_in_column(0, a, c);
      }
    }
    __syncthreads();
    forward_store(0);
  }
  const float m = block_max<T>((float)emax, red);
  if (threadIdx.x == 0) {
    atomicMax(&A.err[(epoch - 1) & A.err_mask], __float_as_uint(m));
    // no carry workgroup reads `remaining`: the next square sweep does
    if (first == 0 && A.remaining) *A.remaining -= 1;
  }
}

// The persistent form (512 threads) of carry_sweep<W, R, 256>, for
// schedules whose line blocks take several rounds; same line table.
template <int W, int R>
const void *carry_pipe_form() {
  static_assert(carry_pass_table<W, R, 256>() == carry_pass_table<W, R, 512>() &&
                    (!carry_pass_table<W, R, 256>() || auto_plan_uniform<W, R, 256>() == auto_plan_uniform<W, R, 512>()),
                "the persistent carry sweep must read the 256-thread form's line table");
  return (const void *)carry_sweep_pipe<W, R, 512>;
}

// ------------------------------------------------------ load / unload (HBM)
// load: balanced digits (natural order) -> weighted, line-transformed,
// twiddled signal, the state the square sweep expects.
template <int W, int R, int T>
__global__ __launch_bounds__(T) void load_lines(const PlanArgs *__restrict__ args) {
  const PlanArgs &A = args[blockIdx.y];
  constexpr int LP = lpitch(W);
  __shared__ cplx lds[R * LP];
  __shared__ double qc[W / 4 + 1];
  load_quarter<W, T>(qc, A.qcW);
  const QuarterRoots<W> tw{qc};
  const int tid = threadIdx.x;
  const uint32_t H = A.H;
  const uint32_t b0 = blockIdx.x * R;
  for (int a = tid; a < W; a += T) {
    const int fa = A.fA[a];
    const double wa = A.wA[a], iwa = A.iwA[a];
    for (int r = 0; r < R; r++) {
      const uint32_t j = 2u * ((uint32_t)a * H + b0 + r);
      const WordWeight we = word_weight(A, fa, wa, iwa, 2 * (b0 + r));
      const WordWeight wo = word_weight(A, fa, wa, iwa, 2 * (b0 + r) + 1);
      lds[r * LP + lpad(a)] = {(double)A.digits[j] * we.w, (double)A.digits[j + 1] * wo.w};
    }
  }
  __syncthreads();
  fft_lds<W, R, T, false>(lds, tw);
  for (int e = tid; e < R * W; e += T) {
    const int r = e % R, k1 = e / R;
    A.x[(size_t)b0 * W + e] = lds[r * LP + lpad(k1)];
  }
}

// unload: inverse of load; digits rounded, rounding error max into err[0].
template <int W, int R, int T>
__global__ __launch_bounds__(T) void unload_lines(const PlanArgs *__restrict__ args) {
  const PlanArgs &A = args[blockIdx.y];
  constexpr int LP = lpitch(W);
  __shared__ cplx lds[R * LP];
  __shared__ double qc[W / 4 + 1];
  __shared__ float red[T / 64];
  load_quarter<W, T>(qc, A.qcW);
  const QuarterRoots<W> tw{qc};
  const int tid = threadIdx.x;
  const uint32_t H = A.H;
  const uint32_t b0 = blockIdx.x * R;
  for (int e = tid; e < R * W; e += T) {
    const int r = e % R, k1 = e / R;
    lds[r * LP + lpad(k1)] = A.x[(size_t)b0 * W + e];
  }
  __syncthreads();
  fft_lds<W, R, T, true>(lds, tw);
  float emax = 0.f;
  const double scale = 1.0 / (double)W;
  for (int a = tid; a < W; a += T) {
    const int fa = A.fA[a];
    const double wa = A.wA[a], iwa = A.iwA[a];
    for (int r = 0; r < R; r++) {
      const uint32_t j = 2u * ((uint32_t)a * H + b0 + r);
      const cplx v = lds[r * LP + lpad(a)];
      const WordWeight we = word_weight(A, fa, wa, iwa, 2 * (b0 + r));
      const WordWeight wo = word_weight(A, fa, wa, iwa, 2 * (b0 + r) + 1);
      const double t0 = v.x * we.iw * scale, t1 = v.y * wo.iw * scale;
      const double r0 = rint(t0), r1 = rint(t1);
      emax = fmaxf(emax, fmaxf((float)fabs(t0 - r0), (float)fabs(t1 - r1)));
      A.digits[j] = (int32_t)r0;
      A.digits[j + 1] = (int32_t)r1;
    }
  }
  const float m = block_max<T>(emax, red);
  if (tid == 0) atomicMax(&A.err[0], __float_as_uint(m));
}

}  // namespace llamd

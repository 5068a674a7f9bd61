// Device-side building blocks for the Lucas-Lehmer iteration on gfx950:
// complex helpers, radix-R DFT butterflies, and a Stockham autosort FFT that
// runs NB independent length-L transforms held in LDS by one workgroup.
//
// Conventions: forward transforms use exp(-2 pi i jk/L), inverse transforms
// exp(+2 pi i jk/L); neither is normalised.  Twiddle tables hold the forward
// roots exp(-2 pi i e/L); inverse passes use their conjugates.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

#include "ll_types.h"

namespace llamd {

// Loads and stores through global-address-space pointers: the signal and
// table pointers arrive in structs as generic pointers, which would
// otherwise compile to flat memory instructions (slower to issue, and
// counted against the LDS wait counter too).
typedef double gd2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cplx gload(const cplx *p) {
  const gd2 v = *(const __attribute__((address_space(1))) gd2 *)p;
  return {v.x, v.y};
}
__device__ __forceinline__ void gstore(cplx *p, cplx c) {
  const gd2 v = {c.x, c.y};
  *(__attribute__((address_space(1))) gd2 *)p = v;
}
template <typename V>
__device__ __forceinline__ V gload(const V *p) {
  return *(const __attribute__((address_space(1))) V *)p;
}
// Streaming forms for the signal, which each sweep reads and writes exactly
// once: non-temporal loads, and stores written through to memory (cache
// policy sc0 sc1: a 32 MiB store phase takes about 8.0 us instead of 9.2
// with the default or the non-temporal policy, measured on MI355X).  The
// stores go through a buffer resource over the whole signal (raw buffer
// stores are the form whose cache policy the compiler takes as an
// argument); word 3 of the resource is the gfx9 untyped-buffer value.
__device__ __forceinline__ cplx gload_stream(const cplx *p) {
  const gd2 v = __builtin_nontemporal_load((const __attribute__((address_space(1))) gd2 *)p);
  return {v.x, v.y};
}
struct StreamOut {
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ __forceinline__ StreamOut(cplx *base, uint32_t count)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(count * sizeof(cplx)), 0x00020000)) {}
  // element `index` of the signal <- c
  __device__ __forceinline__ void store(uint32_t index, cplx c) const {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 v = __builtin_bit_cast(u4, gd2{c.x, c.y});
    __builtin_amdgcn_raw_buffer_store_b128(v, rsrc, (int)(index * sizeof(cplx)), 0, 1 | 16);  // sc0 | sc1
  }
};

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.x, -a.y}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) {
  return {fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y)};
}
__device__ __forceinline__ cplx csqr(cplx a) {
  return {(a.x + a.y) * (a.x - a.y), 2.0 * a.x * a.y};
}
// multiply by -i (forward) or +i (inverse): the primitive 4th root used by
// the transform direction.
template <bool INV>
__device__ __forceinline__ cplx mul_w4(cplx a) {
  return INV ? cplx{-a.y, a.x} : cplx{a.y, -a.x};
}

// ---------------------------------------------------------------- radices
constexpr double kC8 = 0.7071067811865476;   // cos(pi/4)
constexpr double kC16 = 0.9238795325112867;  // cos(pi/8)
constexpr double kS16 = 0.3826834323650898;  // sin(pi/8)
constexpr double kS3 = 0.8660254037844386;   // sin(2pi/3)
constexpr double kC51 = 0.30901699437494745, kS51 = 0.9510565162951535;
constexpr double kC52 = -0.8090169943749475, kS52 = 0.5877852522924731;
constexpr double kC71 = 0.6234898018587335, kS71 = 0.7818314824680298;
constexpr double kC72 = -0.2225209339563144, kS72 = 0.9749279121818236;
constexpr double kC73 = -0.9009688679024191, kS73 = 0.4338837391175581;

template <bool INV>
__device__ __forceinline__ void dft2(cplx &a, cplx &b) {
  cplx t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(cplx &a, cplx &b, cplx &c, cplx &d) {
  cplx s0 = cadd(a, c), d0 = csub(a, c);
  cplx s1 = cadd(b, d), d1 = mul_w4<INV>(csub(b, d));
  a = cadd(s0, s1);
  c = csub(s0, s1);
  b = cadd(d0, d1);
  d = csub(d0, d1);
}

// multiply by w8^k, w8 = exp(-+2 pi i/8)
template <bool INV, int K>
__device__ __forceinline__ cplx mul_w8(cplx a) {
  if constexpr (K == 0) return a;
  else if constexpr (K == 2) return mul_w4<INV>(a);
  else if constexpr (K == 1) {
    // (1 -+ i)/sqrt2
    return INV ? cplx{kC8 * (a.x - a.y), kC8 * (a.x + a.y)}
               : cplx{kC8 * (a.x + a.y), kC8 * (a.y - a.x)};
  } else {  // K == 3: (-1 -+ i)/sqrt2
    return INV ? cplx{-kC8 * (a.x + a.y), kC8 * (a.x - a.y)}
               : cplx{kC8 * (a.y - a.x), -kC8 * (a.x + a.y)};
  }
}

template <bool INV>
__device__ __forceinline__ void dft8(cplx *v) {
  dft4<INV>(v[0], v[2], v[4], v[6]);
  dft4<INV>(v[1], v[3], v[5], v[7]);
  v[3] = mul_w8<INV, 1>(v[3]);
  v[5] = mul_w8<INV, 2>(v[5]);
  v[7] = mul_w8<INV, 3>(v[7]);
  // outputs: X[k + 4m] from even part E[k] (= v[2k]) and odd O[k] (= v[2k+1])
  cplx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cplx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// multiply by w16^k for k in {1,...,6, 9}, w16 = exp(-+2 pi i/16)
template <bool INV, int K>
__device__ __forceinline__ cplx mul_w16(cplx a) {
  if constexpr (K % 2 == 0) {
    return mul_w8<INV, K / 2>(a);
  } else {
    // cos and sin of K*pi/8 for the odd K used by dft16
    constexpr double c = K == 1 ? kC16 : K == 3 ? kS16 : -kC16;  // K == 9
    constexpr double s = K == 1 ? kS16 : K == 3 ? kC16 : -kS16;
    static_assert(K == 1 || K == 3 || K == 9, "unused twiddle");
    // forward w = c - i s, inverse w = c + i s
    constexpr double wi = INV ? s : -s;
    return {fma(a.x, c, -a.y * wi), fma(a.x, wi, a.y * c)};
  }
}

template <bool INV>
__device__ __forceinline__ void dft16(cplx *v) {
  // 4 x 4: columns over stride-4 inputs, twiddle w16^(i*u), then rows.
  dft4<INV>(v[0], v[4], v[8], v[12]);
  dft4<INV>(v[1], v[5], v[9], v[13]);
  dft4<INV>(v[2], v[6], v[10], v[14]);
  dft4<INV>(v[3], v[7], v[11], v[15]);
  v[5] = mul_w16<INV, 1>(v[5]);
  v[9] = mul_w16<INV, 2>(v[9]);
  v[13] = mul_w16<INV, 3>(v[13]);
  v[6] = mul_w16<INV, 2>(v[6]);
  v[10] = mul_w16<INV, 4>(v[10]);
  v[14] = mul_w16<INV, 6>(v[14]);
  v[7] = mul_w16<INV, 3>(v[7]);
  v[11] = mul_w16<INV, 6>(v[11]);
  v[15] = mul_w16<INV, 9>(v[15]);
  dft4<INV>(v[0], v[1], v[2], v[3]);
  dft4<INV>(v[4], v[5], v[6], v[7]);
  dft4<INV>(v[8], v[9], v[10], v[11]);
  dft4<INV>(v[12], v[13], v[14], v[15]);
  // element (u, i) = v[4u + i] holds X[u + 4 i]; transpose to natural order
  cplx t[16];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int i = 0; i < 4; i++) t[u + 4 * i] = v[4 * u + i];
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = t[k];
}

template <bool INV>
__device__ __forceinline__ void dft3(cplx *v) {
  cplx s = cadd(v[1], v[2]);
  cplx d = csub(v[1], v[2]);
  cplx m = {fma(-0.5, s.x, v[0].x), fma(-0.5, s.y, v[0].y)};
  // sigma * i * sin(2pi/3) * d, sigma = -1 forward, +1 inverse
  cplx r = INV ? cplx{-kS3 * d.y, kS3 * d.x} : cplx{kS3 * d.y, -kS3 * d.x};
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <bool INV>
__device__ __forceinline__ void dft5(cplx *v) {
  cplx s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
  cplx s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
  cplx a1 = {v[0].x + kC51 * s1.x + kC52 * s2.x, v[0].y + kC51 * s1.y + kC52 * s2.y};
  cplx a2 = {v[0].x + kC52 * s1.x + kC51 * s2.x, v[0].y + kC52 * s1.y + kC51 * s2.y};
  // b_u = sum_i sin(2 pi i u / 5) d_i
  cplx b1 = {kS51 * d1.x + kS52 * d2.x, kS51 * d1.y + kS52 * d2.y};
  cplx b2 = {kS52 * d1.x - kS51 * d2.x, kS52 * d1.y - kS51 * d2.y};
  // y_u = a_u + sigma i b_u ; y_{5-u} = a_u - sigma i b_u
  cplx ib1 = INV ? cplx{-b1.y, b1.x} : cplx{b1.y, -b1.x};
  cplx ib2 = INV ? cplx{-b2.y, b2.x} : cplx{b2.y, -b2.x};
  v[0] = cadd(v[0], cadd(s1, s2));
  v[1] = cadd(a1, ib1);
  v[4] = csub(a1, ib1);
  v[2] = cadd(a2, ib2);
  v[3] = csub(a2, ib2);
}

template <bool INV>
__device__ __forceinline__ void dft7(cplx *v) {
  cplx s1 = cadd(v[1], v[6]), d1 = csub(v[1], v[6]);
  cplx s2 = cadd(v[2], v[5]), d2 = csub(v[2], v[5]);
  cplx s3 = cadd(v[3], v[4]), d3 = csub(v[3], v[4]);
  // cos(2 pi i u / 7) for (u, i): u=1: c1 c2 c3 ; u=2: c2 c3 c1 ; u=3: c3 c1 c2
  cplx a1 = {v[0].x + kC71 * s1.x + kC72 * s2.x + kC73 * s3.x,
             v[0].y + kC71 * s1.y + kC72 * s2.y + kC73 * s3.y};
  cplx a2 = {v[0].x + kC72 * s1.x + kC73 * s2.x + kC71 * s3.x,
             v[0].y + kC72 * s1.y + kC73 * s2.y + kC71 * s3.y};
  cplx a3 = {v[0].x + kC73 * s1.x + kC71 * s2.x + kC72 * s3.x,
             v[0].y + kC73 * s1.y + kC71 * s2.y + kC72 * s3.y};
  // sin(2 pi i u / 7): u=1: s1 s2 s3 ; u=2: s2 -s3 -s1 ; u=3: s3 -s1 s2
  cplx b1 = {kS71 * d1.x + kS72 * d2.x + kS73 * d3.x, kS71 * d1.y + kS72 * d2.y + kS73 * d3.y};
  cplx b2 = {kS72 * d1.x - kS73 * d2.x - kS71 * d3.x, kS72 * d1.y - kS73 * d2.y - kS71 * d3.y};
  cplx b3 = {kS73 * d1.x - kS71 * d2.x + kS72 * d3.x, kS73 * d1.y - kS71 * d2.y + kS72 * d3.y};
  cplx ib1 = INV ? cplx{-b1.y, b1.x} : cplx{b1.y, -b1.x};
  cplx ib2 = INV ? cplx{-b2.y, b2.x} : cplx{b2.y, -b2.x};
  cplx ib3 = INV ? cplx{-b3.y, b3.x} : cplx{b3.y, -b3.x};
  v[0] = cadd(v[0], cadd(s1, cadd(s2, s3)));
  v[1] = cadd(a1, ib1);
  v[6] = csub(a1, ib1);
  v[2] = cadd(a2, ib2);
  v[5] = csub(a2, ib2);
  v[3] = cadd(a3, ib3);
  v[4] = csub(a3, ib3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(cplx *v) {
  if constexpr (R == 2) dft2<INV>(v[0], v[1]);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<INV>(v);
  else if constexpr (R == 7) dft7<INV>(v);
  else if constexpr (R == 8) dft8<INV>(v);
  else if constexpr (R == 16) dft16<INV>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------------ radix plans
constexpr int odd_part(int L) { while (L % 2 == 0) L /= 2; return L; }
constexpr int log2_pow2(int L) { int k = 0; while ((1 << k) < L) k++; return k; }

// Radix of pass `pass` for a length-L transform: odd prime factors first
// (7, 5, 3), then the power-of-two part split into 16s with an 8, 4 or
// 8 x 8 remainder chosen to avoid radix 2.
constexpr int plan_radix(int L, int pass) {
  int rads[32] = {0};
  int n = 0;
  int o = odd_part(L);
  while (o % 7 == 0) { rads[n++] = 7; o /= 7; }
  while (o % 5 == 0) { rads[n++] = 5; o /= 5; }
  while (o % 3 == 0) { rads[n++] = 3; o /= 3; }
  int k = log2_pow2(L / odd_part(L));
  // k = 4a + r
  int r = k % 4;
  int a = k / 4;
  if (r == 1 && a >= 1) { a -= 1; rads[n++] = 8; rads[n++] = 4; }      // 16 = 8*2 -> 8*4*... : 2^5 = 8*4
  else if (r == 2) { rads[n++] = 4; }
  else if (r == 3) { rads[n++] = 8; }
  else if (r == 1) { rads[n++] = 2; }
  for (int i = 0; i < a; i++) rads[n++] = 16;
  return pass < n ? rads[pass] : 0;
}

// LDS index with one pad slot every 16 complex values.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lpitch(int L) { return L + L / 16 + 1; }

// Roots exp(-2 pi i e/L), e in [0, L), from a quarter-wave cosine table
// qc[e] = cos(2 pi e/L), e in [0, L/4], held in LDS (L divisible by 4).
template <int L>
struct QuarterRoots {
  const double *qc;
  __device__ __forceinline__ cplx operator()(int e) const {
    constexpr int Q = L / 4;
    const int q = e / Q, r = e - q * Q;
    const double c = qc[r], s = qc[Q - r];  // cos, sin of 2 pi r/L
    // angle q*pi/2 + theta: (cos, sin) rotated by q quarter turns
    const double cr = q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
    const double si = q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
    return {cr, -si};
  }
};

// Copies the quarter-wave table of order L from global memory into LDS.
template <int L, int T>
__device__ __forceinline__ void load_quarter(double *dst, const double *__restrict__ src) {
  for (int i = threadIdx.x; i <= L / 4; i += T) dst[i] = gload(src + i);
}

// One Stockham pass over NB transforms of length L held at buf[b*LP + lpad(i)].
// Pass with radix R and span NS reads x[j + i*L/R], multiplies by
// w^(i*(j mod NS)) with w = exp(-+2 pi i/(NS*R)), applies a length-R DFT and
// writes to (j/NS)*NS*R + (j mod NS) + u*NS.
template <int L, int R, int NS, int NB, int T, bool INV, typename TW>
__device__ __forceinline__ void stockham_pass(cplx *buf, const TW &tw, const int tid = threadIdx.x) {
  constexpr int Q = L / R;
  constexpr int NBF = NB * Q;
  constexpr int PER = (NBF + T - 1) / T;
  constexpr int LP = lpitch(L);
  cplx v[PER][R];
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    if ((NBF % T == 0) || u < NBF) {
      const int b = u / Q, j = u % Q;
      const int k = j % NS;
#pragma unroll
      for (int i = 0; i < R; i++) {
        cplx a = buf[b * LP + lpad(j + i * Q)];
        if (NS > 1 && i > 0) {
          const cplx w = tw((i * k) * (L / (NS * R)));
          a = INV ? cmulc(a, w) : cmul(a, w);
        }
        v[s][i] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    if ((NBF % T == 0) || u < NBF) {
      const int b = u / Q, j = u % Q;
      const int k = j % NS;
      dft<R, INV>(v[s]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int i = 0; i < R; i++) buf[b * LP + lpad(base + i * NS)] = v[s][i];
    }
  }
  __syncthreads();
}

// OPQ: as in fft_lds_wave.
template <int L, int NB, int T, bool INV, int PASS = 0, int NS = 1, int END = -1, bool OPQ = false, typename TW>
__device__ __forceinline__ void fft_lds(cplx *buf, const TW &tw) {
  if constexpr (NS < L && PASS != END) {
    constexpr int R = plan_radix(L, PASS);
    static_assert(R > 0, "bad plan");
    if constexpr (OPQ) {
      int t = threadIdx.x;
      asm volatile("" : "+v"(t));
      stockham_pass<L, R, NS, NB, T, INV>(buf, tw, t);
    } else {
      stockham_pass<L, R, NS, NB, T, INV>(buf, tw);
    }
    fft_lds<L, NB, T, INV, PASS + 1, NS * R, END, OPQ>(buf, tw);
  }
}

// Workgroup barrier for LDS only: waits for this wave's LDS operations, not
// for its outstanding global loads, which stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------ wave-local passes
// Each wave of the workgroup owns LW = NB / (T / 64) consecutive lines and
// runs every pass of their transforms by itself: in a pass with Q
// butterflies per line, lane l handles butterflies v = l + 64 s of its
// lines (line v / Q, index v mod Q).  When LW Q is a multiple of 64 in
// every pass, no workgroup barrier is needed between passes: LDS
// instructions of one wave execute in order, and a wavefront fence keeps
// the compiler from moving them across a pass boundary.
// Plans for the wave-local form: UNI = true takes a single radix 8 or 16
// for every pass when L is a power of it (measured faster for L = 512 than
// the mixed default plan), UNI = false the default plan.
constexpr int uniform_radix(int L) {
  for (int r = 16; r >= 8; r /= 2) {
    int x = L;
    while (x % r == 0) x /= r;
    if (x == 1) return r;
  }
  return 0;
}
template <bool UNI>
constexpr int wave_radix(int L, int pass) {
  return UNI ? uniform_radix(L) : plan_radix(L, pass);
}

template <int L, int NB, int T, bool UNI, int PASS = 0, int NS = 1>
constexpr bool wave_passes_ok() {
  if constexpr (NS >= L) {
    return true;
  } else {
    constexpr int R = wave_radix<UNI>(L, PASS);
    constexpr int LW = NB / (T / 64);
    if constexpr (R == 0) return false;
    else return (LW * (L / R)) % 64 == 0 && wave_passes_ok<L, NB, T, UNI, PASS + 1, NS * R>();
  }
}

template <int L, int NB, int T, bool UNI>
constexpr bool wave_plan_ok() {
  return T % 64 == 0 && NB % (T / 64) == 0 && (!UNI || uniform_radix(L) != 0) &&
         wave_passes_ok<L, NB, T, UNI>();
}

template <int L, int NB, int T>
constexpr bool wave_local_ok() {
  return wave_plan_ok<L, NB, T, true>() || wave_plan_ok<L, NB, T, false>();
}

// Twiddles of every pass of one plan, laid out pass by pass: the pass of
// span NS and radix R keeps w^(i k), i in [1, R), k in [0, NS), at
// pt[NS - 1 + (i - 1) NS + k] (the earlier passes fill the first NS - 1
// entries, so a length-L plan needs L - 1).  Lanes of a pass read
// consecutive k, so the lookups are contiguous LDS reads.
struct WaveTable {
  const cplx *pt;
};

template <int L, int T, bool UNI, int PASS = 0, int NS = 1, typename TW>
__device__ __forceinline__ void build_wave_table(cplx *pt, const TW &tw) {
  if constexpr (NS < L) {
    constexpr int R = wave_radix<UNI>(L, PASS);
    if constexpr (NS > 1) {
      for (int idx = threadIdx.x; idx < (R - 1) * NS; idx += T) {
        const int i = idx / NS + 1, k = idx % NS;
        pt[NS - 1 + idx] = tw((i * k) * (L / (NS * R)));
      }
    }
    build_wave_table<L, T, UNI, PASS + 1, NS * R>(pt, tw);
  }
}

template <int L, int R, int NS, int NB, int T, bool INV, typename TW>
__device__ __forceinline__ void stockham_pass_wave(cplx *buf, const TW &tw, const int tid = threadIdx.x) {
  constexpr int Q = L / R;
  constexpr int LW = NB / (T / 64);
  constexpr int PER = LW * Q / 64;
  constexpr int LP = lpitch(L);
  cplx v[PER][R];
  const int lane = tid & 63, line0 = (tid >> 6) * LW;
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = lane + 64 * s;
    const int b = line0 + u / Q, j = u % Q;
    const int k = j % NS;
#pragma unroll
    for (int i = 0; i < R; i++) {
      cplx a = buf[b * LP + lpad(j + i * Q)];
      if (NS > 1 && i > 0) {
        cplx w;
        if constexpr (std::is_same<TW, WaveTable>::value) w = tw.pt[NS - 1 + (i - 1) * NS + k];
        else w = tw((i * k) * (L / (NS * R)));
        a = INV ? cmulc(a, w) : cmul(a, w);
      }
      v[s][i] = a;
    }
  }
  // every lane of this wave has read its inputs before any lane writes
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = lane + 64 * s;
    const int b = line0 + u / Q, j = u % Q;
    const int k = j % NS;
    dft<R, INV>(v[s]);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int i = 0; i < R; i++) buf[b * LP + lpad(base + i * NS)] = v[s][i];
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// NB transforms of length L, each wave transforming its own lines; callers
// synchronise the workgroup before and after.  Passes PASS .. END - 1 are
// run (END = -1: to the last pass).  OPQ: each pass takes the thread index
// through an empty asm statement, for kernels that run the transform in a
// loop (the compiler cannot then hoist every pass's LDS offsets out of the
// loop and keep them all live across it).
template <int L, int NB, int T, bool INV, bool UNI = wave_plan_ok<L, NB, T, true>(), int PASS = 0,
          int NS = 1, int END = -1, bool OPQ = false, typename TW>
__device__ __forceinline__ void fft_lds_wave(cplx *buf, const TW &tw) {
  static_assert(wave_plan_ok<L, NB, T, UNI>(), "fft_lds_wave: plan is not wave-local");
  if constexpr (NS < L && PASS != END) {
    constexpr int R = wave_radix<UNI>(L, PASS);
    if constexpr (OPQ) {
      int t = threadIdx.x;
      asm volatile("" : "+v"(t));
      stockham_pass_wave<L, R, NS, NB, T, INV>(buf, tw, t);
    } else {
      stockham_pass_wave<L, R, NS, NB, T, INV>(buf, tw);
    }
    fft_lds_wave<L, NB, T, INV, UNI, PASS + 1, NS * R, END, OPQ>(buf, tw);
  }
}

// Number of passes of a wave-local plan.
template <bool UNI>
constexpr int wave_passes(int L) {
  int n = 0, ns = 1;
  while (ns < L) ns *= wave_radix<UNI>(L, n++);
  return n;
}

// Either form, chosen at compile time.  With TAB the wave-local form takes
// its twiddles from the per-pass table `pt` (build_wave_table of the plan
// auto_plan_uniform selects, L - 1 entries), else from `tw`.
template <int L, int NB, int T, bool INV, bool TAB, bool OPQ = false, typename TW>
__device__ __forceinline__ void fft_lds_auto(cplx *buf, const TW &tw, const cplx *pt) {
  if constexpr (wave_local_ok<L, NB, T>()) {
    constexpr bool UNI = wave_plan_ok<L, NB, T, true>();
    __syncthreads();
    if constexpr (TAB) fft_lds_wave<L, NB, T, INV, UNI, 0, 1, -1, OPQ>(buf, WaveTable{pt});
    else fft_lds_wave<L, NB, T, INV, UNI, 0, 1, -1, OPQ>(buf, tw);
    __syncthreads();
  } else {
    fft_lds<L, NB, T, INV, 0, 1, -1, OPQ>(buf, tw);
  }
}

// The plan fft_lds_auto's wave-local form uses (uniform radix when valid).
template <int L, int NB, int T>
constexpr bool auto_plan_uniform() {
  return wave_plan_ok<L, NB, T, true>();
}

// ------------------------------------------------ per-thread twiddle cache
// In pass PASS of an NB x L Stockham transform run by T threads, thread t
// always handles butterflies u = t + s*T, so its twiddles w^(i*(j mod NS))
// are fixed: they are loaded once per kernel from the plan's per-pass table
// (build_wave_table of the default plan, in global memory) and shared by
// the forward and the inverse transform (the inverse uses their
// conjugates), which takes the table lookups out of the passes.
template <int L, int NB, int T, int PASS = 0, int NS = 1, bool END = (NS >= L)>
struct PassTwiddles;

template <int L, int NB, int T, int PASS, int NS>
struct PassTwiddles<L, NB, T, PASS, NS, true> {
  __device__ __forceinline__ void init_table(const cplx *) {}
};

template <int L, int NB, int T, int PASS, int NS>
struct PassTwiddles<L, NB, T, PASS, NS, false> {
  static constexpr int R = plan_radix(L, PASS);
  static constexpr int Q = L / R;
  static constexpr int NBF = NB * Q;
  static constexpr int PER = (NBF + T - 1) / T;
  static constexpr int NW = NS > 1 ? R - 1 : 1;
  cplx w[PER][NW];
  PassTwiddles<L, NB, T, PASS + 1, NS * R> next;

  __device__ __forceinline__ void init_table(const cplx *__restrict__ pt) {
    if constexpr (NS > 1) {
#pragma unroll
      for (int s = 0; s < PER; s++) {
        const int u = (int)threadIdx.x + s * T;
        const int k = (u % Q) % NS;
#pragma unroll
        for (int i = 1; i < R; i++) w[s][i - 1] = gload(pt + (NS - 1 + (i - 1) * NS + k));
      }
    }
    next.init_table(pt);
  }
};

// stockham_pass with the thread's twiddles taken from a PassTwiddles level.
template <int L, int R, int NS, int NB, int T, bool INV, typename PT>
__device__ __forceinline__ void stockham_pass_cached(cplx *buf, const PT &pt, const int tid = threadIdx.x) {
  constexpr int Q = L / R;
  constexpr int NBF = NB * Q;
  constexpr int PER = (NBF + T - 1) / T;
  constexpr int LP = lpitch(L);
  cplx v[PER][R];
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    if ((NBF % T == 0) || u < NBF) {
      const int b = u / Q, j = u % Q;
#pragma unroll
      for (int i = 0; i < R; i++) {
        cplx a = buf[b * LP + lpad(j + i * Q)];
        if (NS > 1 && i > 0) a = INV ? cmulc(a, pt.w[s][i - 1]) : cmul(a, pt.w[s][i - 1]);
        v[s][i] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    if ((NBF % T == 0) || u < NBF) {
      const int b = u / Q, j = u % Q;
      const int k = j % NS;
      dft<R, INV>(v[s]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int i = 0; i < R; i++) buf[b * LP + lpad(base + i * NS)] = v[s][i];
    }
  }
  __syncthreads();
}

// fft_lds with cached twiddles (pt: the PassTwiddles of the same L, NB, T).
// Passes PASS .. END-1 are run (END = -1: to the last pass).
template <int L, int NB, int T, bool INV, int PASS = 0, int NS = 1, int END = -1, typename PT>
__device__ __forceinline__ void fft_lds_cached(cplx *buf, const PT &pt) {
  if constexpr (NS < L && PASS != END) {
    constexpr int R = plan_radix(L, PASS);
    stockham_pass_cached<L, R, NS, NB, T, INV>(buf, pt);
    fft_lds_cached<L, NB, T, INV, PASS + 1, NS * R, END>(buf, pt.next);
  }
}

// Number of passes of the plan of L, and the span NS before pass P.
constexpr int plan_passes(int L) {
  int n = 0, ns = 1;
  while (ns < L) ns *= plan_radix(L, n++);
  return n;
}
constexpr int plan_span(int L, int P) {
  int ns = 1;
  for (int i = 0; i < P; i++) ns *= plan_radix(L, i);
  return ns;
}

// The PassTwiddles level of pass K.
template <int K, typename PT>
__device__ __forceinline__ const auto &twiddle_level(const PT &pt) {
  if constexpr (K == 0) return pt;
  else return twiddle_level<K - 1>(pt.next);
}

// Stockham pass of an NB x L transform whose twiddles w^(i k) are read from
// LDS at tab[(i - 1) NS + k] (the layout of the per-pass table).
template <int L, int R, int NS, int NB, int T, bool INV>
__device__ __forceinline__ void stockham_pass_tab(cplx *buf, const cplx *tab, const int tid) {
  constexpr int Q = L / R;
  constexpr int NBF = NB * Q;
  static_assert(NBF % T == 0, "stockham_pass_tab: whole butterflies per thread");
  constexpr int PER = NBF / T;
  constexpr int LP = lpitch(L);
  cplx v[PER][R];
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    const int b = u / Q, j = u % Q, k = j % NS;
#pragma unroll
    for (int i = 0; i < R; i++) {
      cplx a = buf[b * LP + lpad(j + i * Q)];
      if (i > 0) {
        const cplx w = tab[(i - 1) * NS + k];
        a = INV ? cmulc(a, w) : cmul(a, w);
      }
      v[s][i] = a;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < PER; s++) {
    const int u = tid + s * T;
    const int b = u / Q, j = u % Q, k = j % NS;
    dft<R, INV>(v[s]);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int i = 0; i < R; i++) buf[b * LP + lpad(base + i * NS)] = v[s][i];
  }
  __syncthreads();
}

struct UnitRoots {
  __device__ __forceinline__ cplx operator()(int) const { return {1.0, 0.0}; }
};

// A three-pass transform for a kernel that runs it in a loop: the first
// pass has no twiddles, the second reads them from LDS (tab, its block of
// the per-pass table), the third holds them in registers (pt2: the
// PassTwiddles level of pass 2).  Each pass takes the thread index through
// an empty asm statement, so that the compiler cannot hoist the passes'
// LDS offsets out of the loop and keep them all live across it.
template <int L, int NB, int T, bool INV, typename PT>
__device__ __forceinline__ void fft_lds_3pass_in_loop(cplx *buf, const cplx *tab, const PT &pt2) {
  constexpr int R0 = plan_radix(L, 0), R1 = plan_radix(L, 1), R2 = plan_radix(L, 2);
  static_assert(R0 * R1 * R2 == L, "fft_lds_3pass_in_loop: a three-pass plan");
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  stockham_pass<L, R0, 1, NB, T, INV>(buf, UnitRoots{}, tid);
  tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  stockham_pass_tab<L, R1, R0, NB, T, INV>(buf, tab, tid);
  tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  stockham_pass_cached<L, R2, R0 * R1, NB, T, INV>(buf, pt2, tid);
}

}  // namespace llamd

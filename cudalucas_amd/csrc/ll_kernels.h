// Launch-side view of the LL kernels (ll_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ll_types.h"

namespace llamd {

// exp(-2 pi i e/M) = h * (1 + lo[e & mask]), lo = (cos-1, sin), where the
// coarse root h = hi[2k] + hi[2k + 1] (k = e >> shift) is kept as a rounded
// value and its rounding remainder, so the product is within about half an
// ulp of the exact root.
struct TwoLevel {
  const cplx *hi;
  const cplx *lo;
  uint32_t shift;
  uint32_t mask;
};

// Device-resident iteration state of the two-sweep path.  Workgroup 0 of
// the square sweep opens each iteration (epoch, offset, position of the -2,
// whether a batched test still runs); the carry sweep of the same iteration
// only reads it, so no atomics are needed.
struct DevState {
  uint32_t epoch;      // iterations opened so far: the carry shuttle's tag
  uint32_t offset;     // shift offset of the value once this iteration ends
  uint32_t run;        // 0: a finished batched test (the carry sweep idles)
  uint32_t m2_digit;   // digit that takes the -2 in this iteration
  long long m2_value;  // 2^(bit position inside that digit)
};

// A runtime mixed-radix plan of one transform length L (the generic
// kernels): passes of radix radix[0], radix[1], ... (each 2, 3, 4, 5, 7, 8
// or 16) run in place, decimation in frequency forward (natural order in,
// digit-reversed out) and in time inverse (digit-reversed in, natural out).
// Position q of a forward result holds frequency perm[q]; frequency k sits
// at position iperm[k].
constexpr int kGenMaxPasses = 16;
struct GenLine {
  const cplx *roots;      // exp(-2 pi i e/L), e < L
  const uint32_t *perm;   // position -> frequency
  const uint32_t *iperm;  // frequency -> position
  uint32_t L, np;
  uint8_t radix[kGenMaxPasses];
};

struct PlanArgs {
  cplx *x;                 // M complex; element a*H + b at ((b/R) W + a) R + b mod R
  const double *qcW;       // cos(2 pi e/W), e <= W/4
  const double *qcH;       // cos(2 pi e/H), e <= H/4
  const cplx *ptH;         // column plan's per-pass twiddles, H - 1 roots
  const cplx *ptW;         // line plan's per-pass twiddles (wave-local carry sweeps)
  TwoLevel twM;            // exp(-2 pi i e/M), e < M
  const double *wA, *iwA;  // 2^(+-fA[a]/n), a < W
  const double *iwAs;      // 2^(-fA[a]/n) / (4 M), rounded once: the carry sweep's unweight
  const int32_t *fA;       // (-b * 2H a) mod n
  const double *wB, *iwB;  // 2^(+-fB[i]/n), i < 2H
  const int32_t *fB;       // (-b * i) mod n
  uint64_t *shuttle;       // (H/R) x W epoch-tagged carries
  DevState *st;
  uint32_t *err;           // ring of per-iteration max errors (float bits)
  uint32_t err_mask;
  int32_t *digits;         // N words, natural order (load/unload)
  double scale;            // 1 / (4 M) after the squaring round trip
  double base_lo;          // 2^q, q = floor(p / n): the small digit base
  uint32_t rshift;         // log2(R): lines are tiled by R in memory
  uint32_t *remaining;     // iterations left (batched tests), or null
  uint32_t p, n, W, H, q, b;
  // staggered start of the two-sweep kernels: workgroup i of the square
  // (carry) sweep waits i * stagger_sq (stagger_ca) / 16 ticks of the
  // 100 MHz real-time counter before its first load; 0: no wait
  uint32_t stagger_sq, stagger_ca;
  GenLine gW, gH;          // plans of the generic kernels (unused otherwise)
};

struct ResidentDesc {
  cplx *x;            // M complex, natural order (weighted words)
  const double *w;    // 2^(f(j)/n)  (load / unload)
  const double *iw;   // 2^(-f(j)/n) (load / unload)
  // factored weights of the iteration kernel: word j = t C + i (thread t of
  // T, C = n / T words each) has f(j) = fT[t] + fC[i] (mod n), weight
  // wT[t] wC[i] (halved when the sum wraps) and unweight iwTs[t] iwC[i]
  // (doubled when it wraps; iwTs includes the 1 / (4 M) of the round trip)
  const double *wT, *iwTs, *wC, *iwC;
  const int32_t *fT, *fC;
  const double *qc;   // cos(2 pi e/M), e <= M/4
  const cplx *pt;     // the M-point plan's per-pass twiddles, M - 1 roots
  float *err;         // iters floats
  uint32_t *offset;   // shift offset, read at start and written at the end
  uint32_t *remaining; // iterations left (batched tests), or null
  double scale;       // 1 / (4 M)
  uint32_t p, m, iters;
};

struct SquareKernel {
  uint32_t h;         // 0: the generic kernel (any column length, runtime plan)
  const void *fn;
  uint32_t threads;
  const void *table;  // builds the per-pass twiddles of the column plan
  const void *pipe;   // persistent form of the same sweep (square_sweep_pipe), or null
};

struct CarryKernel {
  uint32_t w, r;      // 0, 0: the generic kernel (any line length and R)
  const void *carry, *load, *unload;
  uint32_t threads;
  const void *table;  // builds the line pass table ptW (null: not used)
  const void *pipe;   // persistent form of the carry sweep (carry_sweep_pipe), or null
  const void *wide;   // 512-thread form of a 256-thread sweep (same line table), or null
};

// The carry-sweep launch of a schedule: kernel, workgroup size and
// workgroups along x (carry_launch).
struct CarryLaunch {
  const void *fn;
  uint32_t threads, grid;
};

// The generic square / carry kernels (ll_generic.hip): one instantiation
// for every W x H x R the specialised tables lack.  Dynamic LDS bytes of a
// launch are square_generic_lds(H) and carry_generic_lds(W, R); lengths
// whose LDS exceeds kGenMaxLds have no generic schedule.
constexpr uint32_t kGenMaxLds = 160 * 1024 - 256;  // room for the kernels' own few bytes
const SquareKernel *generic_square();
const CarryKernel *generic_carry();
uint32_t square_generic_lds(uint32_t H);
uint32_t carry_generic_lds(uint32_t W, uint32_t R);
void generic_kernels_init();

// One work item of the batched single-launch iteration: square column pair
// or carry block `block` of test `test` (an index into the launch's
// PlanArgs array).
struct BatchItem {
  uint32_t test : 31;
  uint32_t carry : 1;
  uint32_t block;
};
// count items (all square items first), args per test, per-test counters
// of finished square and carry items (zero at batch creation); lds: the
// largest square_generic_lds / carry_generic_lds of the tests.
hipError_t launch_batch_iteration(const BatchItem *items, uint32_t count, const PlanArgs *args,
                                  unsigned long long *sq_done, unsigned long long *ca_done, uint32_t lds,
                                  hipStream_t s);

struct ResidentKernel {
  uint32_t m;
  const void *fn;
  uint32_t threads;
  const void *table;  // builds the per-pass twiddles of the M-point plan (ResidentDesc::pt)
};

// The no-allocation iteration path (ll_iterate_noalloc) keeps the carry
// shuttle and the per-iteration error slots in one file-scope device array
// of this many bytes per GPU: the last kNoAllocRing * 4 bytes are the error
// slots, the rest the shuttle.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
constexpr uint32_t kNoAllocRing = 16384;
constexpr size_t kNoAllocShuttleBytes = kNoAllocScratchBytes - (size_t)kNoAllocRing * 4;
// Device address of that array on `device`, which must be the current
// device; looked up once per device (hipGetSymbolAddress) and kept.
hipError_t noalloc_scratch(int device, void **ptr);

const SquareKernel *find_square(uint32_t H);
const CarryKernel *find_carry(uint32_t W, uint32_t R);
const CarryKernel *carry_table(size_t *count);
const ResidentKernel *find_resident(uint32_t M);

// Staggered-start step of a sweep kernel (PlanArgs::stagger_sq / _ca):
// spreads the launch's loads over about the time the memory system needs
// for them (bytes_per_group / 6144 sixteenths of a 10 ns tick per
// workgroup, measured best at 1.5M-4M words) when all `groups` workgroups
// run at once (one round); 0 when the grid runs in several rounds, where
// later workgroups start as earlier ones finish anyway.
uint32_t sweep_stagger(const void *fn, uint32_t threads, uint32_t dyn_lds, uint32_t groups,
                       uint32_t bytes_per_group, int device);

// Sweep launches over `tests` tests of one shape: d_args is a device array of
// their PlanArgs, blockIdx.y selects the test.
// W, H, R: the schedule's line length, column length and lines per carry
// block (the generic kernels take them at run time).
// Workgroups along x of a square-sweep launch of `tests` tests with W
// columns on the current device: W / 2 (one per column pair), or one round
// of the persistent form when k has one and W / 2 x tests pairs would take
// more than one round of the plain kernel.
uint32_t square_grid(const SquareKernel *k, uint32_t W, uint32_t tests);
// grid: square_grid's value (the persistent form runs when grid < W / 2).
// start / stop (launch_square, launch_carry): if set, events that take the
// kernel's own start and end times (hipExtLaunchKernel), for timing one
// launch without the dispatch gaps around it.
hipError_t launch_square(const SquareKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                         uint32_t H, uint32_t grid, hipStream_t s, hipEvent_t start = nullptr,
                         hipEvent_t stop = nullptr);
// The carry-sweep launch of `tests` tests with H / R line blocks each on
// the current device: the persistent form (one round of workgroups) when
// k has one and the blocks would take more than one round of the plain
// kernel; else the 512-thread form when k has one, the blocks fit in one
// round of it and outnumber half the CUs (twice the waves per CU where
// there are about as many blocks as CUs); else the plain kernel, H / R
// workgroups.
CarryLaunch carry_launch(const CarryKernel *k, uint32_t H, uint32_t R, uint32_t tests);
hipError_t launch_carry(const CarryKernel *k, const CarryLaunch &cl, const PlanArgs *d_args, uint32_t tests,
                        uint32_t W, uint32_t H, uint32_t R, hipStream_t s, hipEvent_t start = nullptr,
                        hipEvent_t stop = nullptr);
hipError_t launch_load(const CarryKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                       uint32_t H, uint32_t R, hipStream_t s);
hipError_t launch_unload(const CarryKernel *k, const PlanArgs *d_args, uint32_t tests, uint32_t W,
                         uint32_t H, uint32_t R, hipStream_t s);
hipError_t launch_resident(const ResidentKernel *k, const ResidentDesc *d_descs, uint32_t count,
                           hipStream_t s);
hipError_t launch_resident_unload(const ResidentDesc *d_descs, uint32_t count, int32_t *digits,
                                  uint32_t *err, hipStream_t s);
hipError_t launch_resident_load(const ResidentDesc *d_descs, uint32_t count, const int32_t *digits,
                                hipStream_t s);
hipError_t launch_fill(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_copy(const cplx *in, cplx *out, size_t n, hipStream_t s);
// ptH of a column plan from its quarter-wave table qcH (both device memory)
hipError_t launch_square_table(const SquareKernel *k, const double *qcH, cplx *ptH, hipStream_t s);
// ptW of a carry kernel's line plan from qcW (k->table must be set)
hipError_t launch_carry_table(const CarryKernel *k, const double *qcW, cplx *ptW, hipStream_t s);
hipError_t launch_resident_table(const ResidentKernel *k, const double *qc, cplx *pt, hipStream_t s);

}  // namespace llamd

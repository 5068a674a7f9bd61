// Carry-sweep / load / unload instantiations: lines with odd factors, 1080..2880 points (R = 4 and 2), for the
// 4320K..11520K-word lengths of the reference's list that have 2048-point columns.
#include "ll_sweeps.h"

namespace llamd {

#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}

// with the persistent form for lengths whose blocks take several rounds
#define LL_CARRY_PIPE(W, R, T)                                                           \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr,       \
   (const void *)carry_sweep_pipe<W, R, T>}

extern const CarryKernel k_carry_wide[] = {
    LL_CARRY_PIPE(1080, 4, 512), LL_CARRY_PIPE(1152, 4, 512), LL_CARRY_PIPE(1176, 4, 512),
    LL_CARRY_PIPE(1296, 4, 512), LL_CARRY_PIPE(1400, 4, 512), LL_CARRY_PIPE(1440, 4, 512),
    LL_CARRY_PIPE(1512, 4, 512), LL_CARRY_PIPE(1568, 4, 512), LL_CARRY_T(1600, 4, 512),
    LL_CARRY_PIPE(1620, 4, 512), LL_CARRY_T(1800, 2, 256), LL_CARRY_T(1944, 2, 256),
    LL_CARRY_T(2016, 2, 256), LL_CARRY_T(2160, 2, 512), LL_CARRY_PIPE(2304, 2, 512),
    LL_CARRY_T(2352, 2, 512), LL_CARRY_PIPE(2592, 2, 512), LL_CARRY_PIPE(2700, 2, 512),
    LL_CARRY_T(2800, 2, 512), LL_CARRY_T(2880, 2, 512),
};
extern const size_t k_carry_wide_count = sizeof(k_carry_wide) / sizeof(k_carry_wide[0]);

}  // namespace llamd

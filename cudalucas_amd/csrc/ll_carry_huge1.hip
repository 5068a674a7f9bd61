// Carry-sweep / load / unload instantiations: lines of 4500..8192 points
// (R = 1) beside 4096- or 2048-point columns, for the 19600K..65536K
// lengths of the reference's list whose line has a template.
#include "ll_sweeps.h"

namespace llamd {

#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}

extern const CarryKernel k_carry_huge1[] = {
    LL_CARRY_T(4500, 1, 256), LL_CARRY_T(4536, 1, 256), LL_CARRY_T(4900, 1, 256),
    LL_CARRY_T(5184, 1, 256), LL_CARRY_T(5292, 1, 256), LL_CARRY_T(5400, 1, 256),
    LL_CARRY_T(5488, 1, 256), LL_CARRY_T(6272, 1, 256), LL_CARRY_T(6400, 1, 256),
    LL_CARRY_T(6804, 1, 256), LL_CARRY_T(6912, 1, 256), LL_CARRY_T(7000, 1, 256),
    LL_CARRY_T(7168, 1, 256), LL_CARRY_T(8000, 1, 256), LL_CARRY_T(8100, 1, 256),
    LL_CARRY_T(8192, 1, 256),
};
extern const size_t k_carry_huge1_count = sizeof(k_carry_huge1) / sizeof(k_carry_huge1[0]);

}  // namespace llamd

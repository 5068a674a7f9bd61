// Carry-sweep / load / unload instantiations: lines of 4500..8192 points
// (R = 1) beside 4096- or 2048-point columns, for the 19600K..65536K
// lengths of the reference's list whose line has a template.
#include "ll_sweeps.h"

namespace llamd {

#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}

// with the persistent 512-thread form for schedules whose blocks take
// several rounds
#define LL_CARRY_HP(W, R)                                                                \
  {W, R, (const void *)carry_sweep<W, R, 256>, (const void *)load_lines<W, R, 256>,      \
   (const void *)unload_lines<W, R, 256>, 256,                                           \
   carry_pass_table<W, R, 256>() ? (const void *)line_pass_table<W, R, 256> : nullptr,   \
   carry_pipe_form<W, R>(), nullptr}

extern const CarryKernel k_carry_huge1[] = {
    LL_CARRY_HP(4500, 1), LL_CARRY_HP(4536, 1), LL_CARRY_HP(4900, 1),
    LL_CARRY_HP(5184, 1), LL_CARRY_HP(5292, 1), LL_CARRY_HP(5400, 1),
    LL_CARRY_HP(5488, 1), LL_CARRY_T(6272, 1, 256), LL_CARRY_T(6400, 1, 256),
    LL_CARRY_HP(6804, 1), LL_CARRY_T(6912, 1, 256), LL_CARRY_HP(7000, 1),
    LL_CARRY_HP(7168, 1), LL_CARRY_HP(8000, 1), LL_CARRY_HP(8100, 1),
    LL_CARRY_HP(8192, 1),
};
extern const size_t k_carry_huge1_count = sizeof(k_carry_huge1) / sizeof(k_carry_huge1[0]);

}  // namespace llamd

// Carry-sweep / load / unload instantiations: lines of 2268..4096 points
// (R = 2) beside 4096- and 2048-point columns, for the 10240K..32768K
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

extern const CarryKernel k_carry_huge[] = {
    LL_CARRY_T(2268, 2, 256), LL_CARRY_HP(2500, 2), LL_CARRY_HP(2744, 2),
    LL_CARRY_HP(2916, 2), LL_CARRY_HP(3136, 2), LL_CARRY_HP(3200, 2),
    LL_CARRY_HP(3500, 2), LL_CARRY_HP(3584, 2), LL_CARRY_HP(3888, 2),
    LL_CARRY_HP(4000, 2), LL_CARRY_HP(4096, 2), LL_CARRY_T(2560, 2, 256),
    LL_CARRY_T(3072, 2, 256), LL_CARRY_T(3456, 2, 256), LL_CARRY_T(3600, 2, 256),
};
extern const size_t k_carry_huge_count = sizeof(k_carry_huge) / sizeof(k_carry_huge[0]);

}  // namespace llamd

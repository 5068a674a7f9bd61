// Carry-sweep / load / unload instantiations: lines with a repeated or second odd factor, 324..900 points (R = 8), for the
// 1296K..3600K-word lengths of the reference's list that have 2048-point columns.
#include "ll_sweeps.h"

namespace llamd {

#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}

// with a 512-thread form for schedules whose blocks fit in one round of it
#define LL_CARRY_W(W, R)                                                                 \
  {W, R, (const void *)carry_sweep<W, R, 256>, (const void *)load_lines<W, R, 256>,      \
   (const void *)unload_lines<W, R, 256>, 256,                                           \
   carry_pass_table<W, R, 256>() ? (const void *)line_pass_table<W, R, 256> : nullptr,   \
   nullptr, carry_wide_form<W, R>()}

extern const CarryKernel k_carry_mixed[] = {
    LL_CARRY_W(324, 8), LL_CARRY_W(336, 8), LL_CARRY_W(392, 8),
    LL_CARRY_W(400, 8), LL_CARRY_W(432, 8), LL_CARRY_T(540, 8, 512),
    LL_CARRY_T(576, 8, 512), LL_CARRY_T(588, 8, 512), LL_CARRY_T(648, 8, 512),
    LL_CARRY_T(672, 8, 512), LL_CARRY_T(720, 8, 512), LL_CARRY_T(756, 8, 512),
    LL_CARRY_T(784, 8, 512), LL_CARRY_T(800, 8, 512), LL_CARRY_T(900, 8, 512),
};
extern const size_t k_carry_mixed_count = sizeof(k_carry_mixed) / sizeof(k_carry_mixed[0]);

}  // namespace llamd

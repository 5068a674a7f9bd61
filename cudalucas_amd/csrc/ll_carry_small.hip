// Carry-sweep / load / unload instantiations: short lines (72..448) for 18K..224K-word FFTs.
#include "ll_sweeps.h"

namespace llamd {

#define LL_CARRY_T(W, R, T)                                                              \
  {W, R, (const void *)carry_sweep<W, R, T>, (const void *)load_lines<W, R, T>,          \
   (const void *)unload_lines<W, R, T>, T,                                               \
   carry_pass_table<W, R, T>() ? (const void *)line_pass_table<W, R, T> : nullptr}
#define LL_CARRY(W, R) LL_CARRY_T(W, R, 256)

// with a 512-thread form for schedules whose blocks fit in one round of it
#define LL_CARRY_W(W, R)                                                                 \
  {W, R, (const void *)carry_sweep<W, R, 256>, (const void *)load_lines<W, R, 256>,      \
   (const void *)unload_lines<W, R, 256>, 256,                                           \
   carry_pass_table<W, R, 256>() ? (const void *)line_pass_table<W, R, 256> : nullptr,   \
   nullptr, carry_wide_form<W, R>()}

extern const CarryKernel k_carry_small[] = {
    LL_CARRY_T(72, 8, 64),   LL_CARRY_T(80, 8, 64),   LL_CARRY_T(128, 8, 64),
    LL_CARRY_T(140, 8, 128), LL_CARRY_T(144, 8, 128), LL_CARRY_T(160, 8, 128),
    LL_CARRY_T(168, 8, 128), LL_CARRY_T(192, 8, 128), LL_CARRY_T(200, 8, 128),
    LL_CARRY_T(224, 8, 128), LL_CARRY_T(240, 8, 128), LL_CARRY_T(252, 8, 128),
    LL_CARRY_W(288, 8),        LL_CARRY_W(320, 8),        LL_CARRY_W(360, 8),
    LL_CARRY_W(448, 8),
};
extern const size_t k_carry_small_count = sizeof(k_carry_small) / sizeof(k_carry_small[0]);

}  // namespace llamd

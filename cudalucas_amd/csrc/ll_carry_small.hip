// Carry-sweep / load / unload instantiations: short lines (72..448) for 18K..224K-word FFTs.
#include "ll_carry_table.h"

namespace llamd {

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

// Carry-sweep / load / unload instantiations: lines with a repeated or second odd factor, 324..900 points (R = 8), for the
// 1296K..3600K-word lengths of the reference's list that have 2048-point columns.
#include "ll_carry_table.h"

namespace llamd {

extern const CarryKernel k_carry_mixed[] = {
    LL_CARRY_W(324, 8), LL_CARRY_W(336, 8), LL_CARRY_W(392, 8),
    LL_CARRY_W(400, 8), LL_CARRY_W(432, 8), LL_CARRY_T(540, 8, 512),
    LL_CARRY_T(576, 8, 512), LL_CARRY_T(588, 8, 512), LL_CARRY_T(648, 8, 512),
    LL_CARRY_T(672, 8, 512), LL_CARRY_T(720, 8, 512), LL_CARRY_T(756, 8, 512),
    LL_CARRY_T(784, 8, 512), LL_CARRY_T(800, 8, 512), LL_CARRY_T(900, 8, 512),
};
extern const size_t k_carry_mixed_count = sizeof(k_carry_mixed) / sizeof(k_carry_mixed[0]);

}  // namespace llamd

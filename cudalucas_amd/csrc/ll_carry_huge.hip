// Carry-sweep / load / unload instantiations: lines of 2268..4096 points
// (R = 2) beside 4096- and 2048-point columns, for the 10240K..32768K
// lengths of the reference's list whose line has a template.
#include "ll_carry_table.h"

namespace llamd {

extern const CarryKernel k_carry_huge[] = {
    LL_CARRY_T(2268, 2, 256), LL_CARRY_HP(2500, 2), LL_CARRY_HP(2744, 2),
    LL_CARRY_HP(2916, 2), LL_CARRY_HP(3136, 2), LL_CARRY_HP(3200, 2),
    LL_CARRY_HP(3500, 2), LL_CARRY_HP(3584, 2), LL_CARRY_HP(3888, 2),
    LL_CARRY_HP(4000, 2), LL_CARRY_HP(4096, 2), LL_CARRY_T(2560, 2, 256),
    LL_CARRY_T(3072, 2, 256), LL_CARRY_T(3456, 2, 256), LL_CARRY_T(3600, 2, 256),
};
extern const size_t k_carry_huge_count = sizeof(k_carry_huge) / sizeof(k_carry_huge[0]);

}  // namespace llamd

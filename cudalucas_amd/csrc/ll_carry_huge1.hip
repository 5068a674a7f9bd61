// Carry-sweep / load / unload instantiations: lines of 4500..8192 points
// (R = 1) beside 4096- or 2048-point columns, for the 19600K..65536K
// lengths of the reference's list whose line has a template.
#include "ll_carry_table.h"

namespace llamd {

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

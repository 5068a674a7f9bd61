// Carry-sweep / load / unload instantiations: lines with odd factors, 1080..2880 points (R = 4 and 2), for the
// 4320K..11520K-word lengths of the reference's list that have 2048-point columns.
#include "ll_carry_table.h"

namespace llamd {

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

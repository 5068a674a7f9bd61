// Carry-sweep / load / unload instantiations: line lengths 256..2048.
#include "ll_carry_table.h"

namespace llamd {

extern const CarryKernel k_carry_large[] = {
    LL_CARRY_W(256, 4),  LL_CARRY_W(512, 4),  LL_CARRY_W(1024, 4), LL_CARRY(2048, 2),
    LL_CARRY_W(384, 4),  LL_CARRY_W(640, 4),  LL_CARRY_W(768, 4),  LL_CARRY_W(896, 4),
    LL_CARRY_PIPE(1280, 4, 512), LL_CARRY_PIPE(1536, 4, 512), LL_CARRY_T(1792, 4, 512), LL_CARRY(1024, 2),
    LL_CARRY_W(512, 8),  LL_CARRY_PIPE(1024, 8, 512), LL_CARRY_W(256, 8), LL_CARRY_W(384, 8),
    LL_CARRY_W(640, 8),  LL_CARRY_W(768, 8), LL_CARRY_PIPE(896, 8, 512),
    LL_CARRY_T(512, 16, 512), LL_CARRY_PIPE(2048, 4, 512),
};
extern const size_t k_carry_large_count = sizeof(k_carry_large) / sizeof(k_carry_large[0]);

}  // namespace llamd

// Internal host helpers shared by the engine (ll_host.cpp).
#pragma once
#include <stdint.h>

#include <vector>

namespace llamd {

// exp(-2 pi i e / L), from an exact rational reduction
void root_of_unity(uint64_t e, uint64_t L, double *re, double *im);
// the same in long double (about 64 bits), before rounding to double
void root_of_unity_ld(uint64_t e, uint64_t L, long double *re, long double *im);
// 2^(-f/n) / four_m, rounded once to double
double inverse_weight_scaled(uint32_t f, uint32_t n, uint64_t four_m);
// exp(-2 pi i e / L) - 1 with an accurate small real part
void root_minus_one(uint64_t e, uint64_t L, double *re, double *im);
// interleaved (re, im) of exp(-2 pi i e / L) for e < L
void roots_table(uint32_t L, std::vector<double> &out);
// cos(2 pi e / L) for e in [0, L/4]
void quarter_cosines(uint32_t L, std::vector<double> &out);
// f(j) = (-(p mod n) j) mod n, the IBDWT weight exponent of word j
uint32_t weight_exponent(uint32_t p, uint32_t n, uint64_t j);
// 2^(f/n) and 2^(-f/n)
void weight_pair(uint32_t f, uint32_t n, double *w, double *iw);

}  // namespace llamd

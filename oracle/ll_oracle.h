/* CPU oracle for the Lucas-Lehmer iteration path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * host driver, the Python package) may link, import or call this code.  It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and there only as the checker.
 *
 * Contents
 *   - exact LL arithmetic on GMP (s <- s^2 - 2 mod 2^p - 1);
 *   - a CPU restatement of the reference's IBDWT iteration
 *     (CUDALucas.cu:1125-1252 tables, 255-314 square, 322-534 rcb/splice,
 *      1869-1986 lucas_square) with its own mixed-radix FFT standing in for
 *     cuFFT's Z2Z INVERSE transform (CUDALucas.cu:1239,1888,1890);
 *   - restatements of the digit codec, Res64 extraction, checkpoint CRC and
 *     FFT-length policy (CUDALucas.cu:947-1068,1280-1342,1465-1614,1712-1733).
 */
#ifndef LL_ORACLE_H
#define LL_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded inputs ---------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
int orc_is_prime(uint64_t n);
/* next prime >= lo + (splitmix64(seed) mod (hi - lo)) */
uint32_t orc_seeded_prime(uint64_t seed, uint64_t lo, uint64_t hi);

/* ---- exact arithmetic (GMP) ------------------------------------------- */
/* words = ceil(p/32) little-endian uint32 words of a value in [0, 2^p-1). */
int orc_gmp_ll(uint32_t p, const uint32_t *start, uint64_t iters,
               uint32_t *out_words);
/* One modular squaring minus `sub` (sub may be 0). */
int orc_gmp_square_sub(uint32_t p, const uint32_t *in, uint32_t sub,
                       uint32_t *out);
/* Modular product a*b mod 2^p-1 (for tests of the convolution). */
int orc_gmp_mulmod(uint32_t p, const uint32_t *a, const uint32_t *b,
                   uint32_t *out);

/* ---- codec (CUDALucas.cu:1280-1342, 1465-1614) ----------------------- */
int orc_digit_width(uint32_t p, uint32_t n, uint32_t j);
void orc_unpack_bits(uint32_t p, uint32_t n, const uint32_t *packed,
                     int32_t *digits);
void orc_balance_digits(uint32_t p, uint32_t n, int32_t *digits);
int orc_standardize_digits(uint32_t p, uint32_t n, int32_t *digits);
void orc_pack_bits(uint32_t p, uint32_t n, const int32_t *digits,
                   uint32_t *packed);
uint64_t orc_find_residue(uint32_t p, uint32_t n, const int32_t *digits,
                          uint32_t offset);
/* Exact integer value of a digit vector (any digit size), reduced mod
 * 2^p-1 into [0, 2^p-1), then multiplied by 2^-offset (unshifted). */
int orc_digits_to_residue(uint32_t p, uint32_t n, const int32_t *digits,
                          uint32_t offset, uint32_t *out_words);
/* Initial state 4 * 2^offset as digits (CUDALucas.cu:1261-1267). */
void orc_initial_digits(uint32_t p, uint32_t n, uint32_t offset,
                        int32_t *digits);
uint64_t orc_res64_words(const uint32_t *words);

/* ---- checkpoint checksum (CUDALucas.cu:1712-1733) --------------------- */
uint32_t orc_checkpoint_crc(const void *bytes, size_t len);

/* ---- FFT-length policy (CUDALucas.cu:947-1068) ------------------------ */
int orc_default_fft_count(void);
int orc_default_fft_length(int index); /* words */
/* lb, ub power laws of choose_fft_length */
void orc_fft_bounds(uint32_t p, int32_t *lb, int32_t *ub);

/* ---- FFT used by the restated iteration ------------------------------ */
/* Unnormalised complex DFT with exponent sign +1 (cuFFT's INVERSE),
 * in place on m interleaved (re, im) doubles.  m must be 7-smooth. */
int orc_fft_inverse(uint32_t m, double *data);

/* ---- CPU IBDWT iteration (restatement of lucas_square) --------------- */
typedef struct orc_ibdwt orc_ibdwt;
orc_ibdwt *orc_ibdwt_create(uint32_t p, uint32_t n);
void orc_ibdwt_destroy(orc_ibdwt *c);
/* Digits are any int32 vector whose value is the shifted state. */
void orc_ibdwt_load(orc_ibdwt *c, const int32_t *digits, uint32_t offset);
/* Runs count iterations; per_iter (may be NULL) receives each iteration's
 * maximum |x - rint(x)|; returns the maximum over all of them. */
double orc_ibdwt_iterate(orc_ibdwt *c, uint64_t count, double *per_iter);
/* Threads for the IBDWT's data-parallel loops (1 by default); returns the
 * count in effect (1 without OpenMP).  Results do not depend on it. */
int orc_set_threads(int n);
void orc_ibdwt_read(orc_ibdwt *c, int32_t *digits, uint32_t *offset);

#ifdef __cplusplus
}
#endif
#endif

/* Hand-written declarations of the documented GMP 6.x mpz ABI.
 *
 * Test infrastructure only: the oracle links the system runtime
 * (libgmp.so.10) without gmp.h, which this image does not ship.  Every
 * name below is a public, documented entry point of GMP (the mpz_* macros
 * of gmp.h expand to these __gmpz_* symbols); the struct layout is the
 * documented __mpz_struct.
 */
#ifndef ORACLE_GMP_ABI_H
#define ORACLE_GMP_ABI_H
#include <stddef.h>

typedef struct {
  int _mp_alloc;
  int _mp_size;
  unsigned long *_mp_d;
} oracle_mpz_struct;
typedef oracle_mpz_struct oracle_mpz_t[1];

void __gmpz_init(oracle_mpz_struct *);
void __gmpz_clear(oracle_mpz_struct *);
void __gmpz_set_ui(oracle_mpz_struct *, unsigned long);
void __gmpz_set(oracle_mpz_struct *, const oracle_mpz_struct *);
void __gmpz_mul(oracle_mpz_struct *, const oracle_mpz_struct *, const oracle_mpz_struct *);
void __gmpz_add(oracle_mpz_struct *, const oracle_mpz_struct *, const oracle_mpz_struct *);
void __gmpz_sub(oracle_mpz_struct *, const oracle_mpz_struct *, const oracle_mpz_struct *);
void __gmpz_sub_ui(oracle_mpz_struct *, const oracle_mpz_struct *, unsigned long);
void __gmpz_add_ui(oracle_mpz_struct *, const oracle_mpz_struct *, unsigned long);
void __gmpz_tdiv_q_2exp(oracle_mpz_struct *, const oracle_mpz_struct *, unsigned long);
void __gmpz_tdiv_r_2exp(oracle_mpz_struct *, const oracle_mpz_struct *, unsigned long);
void __gmpz_mul_2exp(oracle_mpz_struct *, const oracle_mpz_struct *, unsigned long);
int __gmpz_cmp(const oracle_mpz_struct *, const oracle_mpz_struct *);
void __gmpz_setbit(oracle_mpz_struct *, unsigned long);
void *__gmpz_export(void *, size_t *, int, size_t, int, size_t, const oracle_mpz_struct *);
void __gmpz_import(oracle_mpz_struct *, size_t, int, size_t, int, size_t, const void *);

#endif

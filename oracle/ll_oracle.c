/* CPU oracle for the Lucas-Lehmer iteration path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product never links it.
 *
 * Every function names the reference code it restates (file:line in
 * tdulcet/CUDALucas v2.06).  Where the reference uses doubles for integer
 * bookkeeping (digit/bit positions, big-digit flags) this restatement uses
 * exact integer arithmetic; the values are the same wherever the reference's
 * doubles are exact.
 *
 * Pinning (see tests/test_oracle.py): the GMP routines against the
 * reference's Res64 self-test tables (CUDALucas.cu:2673-2729) and the
 * Mersenne-prime closed form (residue 0 after p-2 iterations); the IBDWT
 * restatement against GMP full residues; the FFT against a brute-force DFT;
 * the CRC against the standard CRC-32/MPEG-2 check value; the codec by
 * round trips against exact integers.
 */
#include "ll_oracle.h"
#include "gmp_abi.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* seeded inputs                                                          */
/* ===================================================================== */

uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Threads for the IBDWT's data-parallel loops (FFT passes, the pair
 * squaring, rounding and reweighting); the carry chain stays sequential.
 * Each element is computed by the same operations whichever thread runs it,
 * so the results do not depend on the count.  Default 1: no thread pool is
 * started unless a caller asks (the oracle is also used from forked
 * processes). */
static int g_threads = 1;

int orc_set_threads(int n) {
#ifdef _OPENMP
  g_threads = n < 1 ? 1 : n;
#else
  (void)n;
  g_threads = 1;
#endif
  return g_threads;
}

static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)((unsigned __int128)a * b % m);
}

static uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = mulmod64(r, a, m);
    a = mulmod64(a, a, m);
    e >>= 1;
  }
  return r;
}

/* Deterministic Miller-Rabin for 64-bit n. */
int orc_is_prime(uint64_t n) {
  static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return 0;
  for (int i = 0; i < 12; i++) {
    if (n == bases[i]) return 1;
    if (n % bases[i] == 0) return 0;
  }
  uint64_t d = n - 1;
  int s = 0;
  while (!(d & 1)) { d >>= 1; s++; }
  for (int i = 0; i < 12; i++) {
    uint64_t x = powmod64(bases[i], d, n);
    if (x == 1 || x == n - 1) continue;
    int comp = 1;
    for (int r = 1; r < s; r++) {
      x = mulmod64(x, x, n);
      if (x == n - 1) { comp = 0; break; }
    }
    if (comp) return 0;
  }
  return 1;
}

uint32_t orc_seeded_prime(uint64_t seed, uint64_t lo, uint64_t hi) {
  uint64_t st = seed;
  uint64_t x = lo + (hi > lo ? orc_splitmix64(&st) % (hi - lo) : 0);
  while (!orc_is_prime(x)) x++;
  return (uint32_t)x;
}

/* ===================================================================== */
/* exact arithmetic on GMP                                                */
/* ===================================================================== */

static void mpz_from_words(oracle_mpz_struct *z, const uint32_t *w, uint32_t p) {
  size_t nw = (p + 31) / 32;
  __gmpz_import(z, nw, -1, 4, 0, 0, w);
}

static void mpz_to_words(const oracle_mpz_struct *z, uint32_t *w, uint32_t p) {
  size_t nw = (p + 31) / 32, cnt = 0;
  memset(w, 0, nw * 4);
  if (z->_mp_size == 0) return;
  uint32_t *tmp = (uint32_t *)__gmpz_export(NULL, &cnt, -1, 4, 0, 0, z);
  memcpy(w, tmp, (cnt < nw ? cnt : nw) * 4);
  free(tmp);
}

/* z <- z mod (2^p - 1), z >= 0 on entry, result in [0, 2^p - 1). */
static void mpz_mod_mersenne(oracle_mpz_struct *z, uint32_t p,
                             oracle_mpz_struct *hi, const oracle_mpz_struct *mp) {
  for (;;) {
    __gmpz_tdiv_q_2exp(hi, z, p);
    if (hi->_mp_size == 0) break;
    __gmpz_tdiv_r_2exp(z, z, p);
    __gmpz_add(z, z, hi);
  }
  if (__gmpz_cmp(z, mp) >= 0) __gmpz_sub(z, z, mp);
}

int orc_gmp_ll(uint32_t p, const uint32_t *start, uint64_t iters,
               uint32_t *out_words) {
  oracle_mpz_t s, hi, mp;
  __gmpz_init(s); __gmpz_init(hi); __gmpz_init(mp);
  __gmpz_set_ui(mp, 0); __gmpz_setbit(mp, p); __gmpz_sub_ui(mp, mp, 1);
  if (start) mpz_from_words(s, start, p);
  else __gmpz_set_ui(s, 4);
  for (uint64_t i = 0; i < iters; i++) {
    __gmpz_mul(s, s, s);
    mpz_mod_mersenne(s, p, hi, mp);
    if (s->_mp_size == 0 || (s->_mp_size == 1 && s->_mp_d[0] < 2))
      __gmpz_add(s, s, mp);
    __gmpz_sub_ui(s, s, 2);
  }
  if (__gmpz_cmp(s, mp) == 0) __gmpz_set_ui(s, 0);
  mpz_to_words(s, out_words, p);
  __gmpz_clear(s); __gmpz_clear(hi); __gmpz_clear(mp);
  return 0;
}

int orc_gmp_square_sub(uint32_t p, const uint32_t *in, uint32_t sub,
                       uint32_t *out) {
  oracle_mpz_t s, hi, mp;
  __gmpz_init(s); __gmpz_init(hi); __gmpz_init(mp);
  __gmpz_set_ui(mp, 0); __gmpz_setbit(mp, p); __gmpz_sub_ui(mp, mp, 1);
  mpz_from_words(s, in, p);
  __gmpz_mul(s, s, s);
  mpz_mod_mersenne(s, p, hi, mp);
  __gmpz_add(s, s, mp);
  __gmpz_sub_ui(s, s, sub);
  mpz_mod_mersenne(s, p, hi, mp);
  mpz_to_words(s, out, p);
  __gmpz_clear(s); __gmpz_clear(hi); __gmpz_clear(mp);
  return 0;
}

int orc_gmp_mulmod(uint32_t p, const uint32_t *a, const uint32_t *b,
                   uint32_t *out) {
  oracle_mpz_t x, y, hi, mp;
  __gmpz_init(x); __gmpz_init(y); __gmpz_init(hi); __gmpz_init(mp);
  __gmpz_set_ui(mp, 0); __gmpz_setbit(mp, p); __gmpz_sub_ui(mp, mp, 1);
  mpz_from_words(x, a, p);
  mpz_from_words(y, b, p);
  __gmpz_mul(x, x, y);
  mpz_mod_mersenne(x, p, hi, mp);
  mpz_to_words(x, out, p);
  __gmpz_clear(x); __gmpz_clear(y); __gmpz_clear(hi); __gmpz_clear(mp);
  return 0;
}

/* ===================================================================== */
/* variable-base digits                                                   */
/* ===================================================================== */

/* Bit position where digit j starts: ceil(p*j/n) (CUDALucas.cu:1265). */
static uint64_t digit_start(uint32_t p, uint32_t n, uint64_t j) {
  return ((uint64_t)p * j + n - 1) / n;
}

/* Width of digit j: q+1 ("big") iff f(j) = (-b*j) mod n < b, b = p mod n;
 * the reference derives the same flag from weight comparisons
 * (CUDALucas.cu:1192-1213) and reads it back with size()
 * (CUDALucas.cu:1280-1283). */
int orc_digit_width(uint32_t p, uint32_t n, uint32_t j) {
  uint64_t b = p % n;
  uint64_t f = (n - (b * j) % n) % n;
  return (int)(p / n) + (f < b ? 1 : 0);
}

/* unpack_bits (CUDALucas.cu:1285-1311): standard digits from packed words. */
void orc_unpack_bits(uint32_t p, uint32_t n, const uint32_t *packed,
                     int32_t *digits) {
  uint64_t acc = 0;
  int have = 0;
  size_t w = 0, nw = (p + 31) / 32;
  for (uint32_t i = 0; i < n; i++) {
    int width = orc_digit_width(p, n, i);
    while (have < width) {
      uint64_t next = w < nw ? packed[w] : 0;
      acc |= next << have;
      have += 32;
      w++;
    }
    digits[i] = (int32_t)(acc & ((1ull << width) - 1));
    acc >>= width;
    have -= width;
  }
}

/* balance_digits (CUDALucas.cu:1313-1342): one upward pass mapping
 * [0, 2^w) digits into [-2^(w-1), 2^(w-1)), final carry into digit 0. */
void orc_balance_digits(uint32_t p, uint32_t n, int32_t *digits) {
  int carry = 0;
  for (uint32_t j = 0; j < n; j++) {
    int width = orc_digit_width(p, n, j);
    int32_t half = 1 << (width - 1);
    digits[j] += carry;
    carry = 0;
    if (digits[j] >= half) {
      digits[j] -= 2 * half;
      carry = 1;
    }
  }
  digits[0] += carry;
}

/* standardize_digits (CUDALucas.cu:1465-1489): balanced -> [0, 2^w),
 * with the sign of the top non-zero digit as the initial borrow.
 * Returns 1 if the value is zero. */
int orc_standardize_digits(uint32_t p, uint32_t n, int32_t *digits) {
  int32_t j = (int32_t)n - 1;
  int carry = 0;
  while (j > 0 && digits[j] == 0) j--;
  if (j == 0 && digits[0] == 0) return 1;
  if (digits[j] < 0) carry = -1;
  for (uint32_t i = 0; i < n; i++) {
    int width = orc_digit_width(p, n, i);
    digits[i] += carry;
    if (digits[i] < 0) {
      digits[i] += 1 << width;
      carry = -1;
    } else {
      carry = 0;
    }
  }
  return 0;
}

/* pack_bits (CUDALucas.cu:1594-1614): standard digits -> packed words. */
void orc_pack_bits(uint32_t p, uint32_t n, const int32_t *digits,
                   uint32_t *packed) {
  uint64_t acc = 0;
  int have = 0;
  size_t w = 0;
  for (uint32_t i = 0; i < n; i++) {
    acc += (uint64_t)(uint32_t)digits[i] << have;
    have += orc_digit_width(p, n, i);
    if (have >= 32) {
      packed[w++] = (uint32_t)acc;
      acc >>= 32;
      have -= 32;
    }
  }
  if (w < (p + 31) / 32) packed[w] = (uint32_t)acc;
}

/* find_residue (CUDALucas.cu:1491-1534): low 64 bits of the unshifted
 * residue, read from balanced digits starting at the digit that holds bit
 * `offset`, with the sign of the cyclically highest non-zero digit as the
 * initial borrow. */
uint64_t orc_find_residue(uint32_t p, uint32_t n, const int32_t *digits,
                          uint32_t offset) {
  uint32_t digit = (uint32_t)(((uint64_t)offset * n) / p);
  int bit = (int)(offset - digit_start(p, n, digit));
  int32_t j = (int32_t)((n + digit - 1) % n);
  int carry = 0, k = 0;
  uint64_t residue = 0;
  while (digits[j] == 0 && (uint32_t)j != digit) {
    j--;
    if (j < 0) j += n;
  }
  if ((uint32_t)j == digit && digits[digit] == 0) return 0;
  if (digits[j] < 0) carry = -1;
  for (int i = 0; i < 10; i++) {
    int width = orc_digit_width(p, n, digit);
    int64_t tx = (int64_t)digits[digit] + carry;
    if (tx < 0) {
      tx += (int64_t)1 << width;
      carry = -1;
    } else {
      carry = 0;
    }
    if (k < 64) residue += (uint64_t)tx << k;
    k += width;
    if (i == 0) {
      k -= bit;
      residue >>= bit;
    }
    if (k >= 64) break;
    digit++;
    if (digit == n) digit = 0;
  }
  return residue;
}

/* Bits [p, p + 64) of a two's-complement word array, as a signed value. */
static int64_t bits_above(const uint32_t *w, size_t L, uint32_t p) {
  uint64_t h = 0;
  for (int k = 0; k < 64; k++) {
    uint64_t b = (uint64_t)p + k;
    uint32_t bit = (b >> 5) < L ? (w[b >> 5] >> (b & 31)) & 1u
                                : (w[L - 1] >> 31) & 1u;
    h |= (uint64_t)bit << k;
  }
  return (int64_t)h;
}

/* Exact value of any digit vector mod 2^p-1, then rotated right by offset:
 * the unshifted residue (CUDALucas.cu:1491-1534 reads its low 64 bits). */
int orc_digits_to_residue(uint32_t p, uint32_t n, const int32_t *digits,
                          uint32_t offset, uint32_t *out_words) {
  size_t nw = (p + 31) / 32;
  size_t L = nw + 4;
  int64_t *acc = (int64_t *)calloc(L, sizeof(int64_t));
  uint32_t *w = (uint32_t *)calloc(L, sizeof(uint32_t));
  if (!acc || !w) { free(acc); free(w); return -1; }
  for (uint32_t j = 0; j < n; j++) {
    uint64_t s = digit_start(p, n, j);
    int64_t t = (int64_t)digits[j] * ((int64_t)1 << (s & 31));
    acc[s >> 5] += t & 0xffffffffll;
    acc[(s >> 5) + 1] += t >> 32;
  }
  int64_t carry = 0;
  for (size_t i = 0; i < L; i++) {
    int64_t x = acc[i] + carry;
    w[i] = (uint32_t)x;
    carry = x >> 32;
  }
  /* X = Lo + 2^p H  ==  Lo + H  (mod 2^p - 1); repeat until 0 <= X < 2^p */
  for (;;) {
    int64_t H = bits_above(w, L, p);
    if (H == 0) break;
    for (uint64_t b = p; b < 32ull * L; b++) w[b >> 5] &= ~(1u << (b & 31));
    uint64_t add = (uint64_t)H;
    uint32_t ext = H < 0 ? 0xffffffffu : 0u;
    uint64_t c = 0;
    for (size_t i = 0; i < L; i++) {
      uint64_t a = i == 0 ? (uint32_t)add : i == 1 ? (uint32_t)(add >> 32) : ext;
      uint64_t s = (uint64_t)w[i] + a + c;
      w[i] = (uint32_t)s;
      c = s >> 32;
    }
  }
  int all_ones = 1;
  for (uint32_t b = 0; b < p; b++)
    if (!((w[b >> 5] >> (b & 31)) & 1u)) { all_ones = 0; break; }
  if (all_ones) memset(w, 0, L * 4);
  memset(out_words, 0, nw * 4);
  offset %= p;
  for (uint32_t i = 0; i < p; i++) {
    uint32_t src = i + offset;
    if (src >= p) src -= p;
    if ((w[src >> 5] >> (src & 31)) & 1u) out_words[i >> 5] |= 1u << (i & 31);
  }
  free(acc); free(w);
  return 0;
}

/* init_x for a new test (CUDALucas.cu:1261-1267): one set bit, the value
 * 4 * 2^offset mod 2^p-1. */
void orc_initial_digits(uint32_t p, uint32_t n, uint32_t offset,
                        int32_t *digits) {
  uint64_t bit = ((uint64_t)offset + 2) % p;
  uint32_t digit = (uint32_t)((bit * n) / p);
  uint64_t pos = bit - digit_start(p, n, digit);
  memset(digits, 0, sizeof(int32_t) * n);
  digits[digit] = (int32_t)(1u << pos);
}

uint64_t orc_res64_words(const uint32_t *words) {
  return (uint64_t)words[0] | ((uint64_t)words[1] << 32);
}

/* ===================================================================== */
/* checkpoint checksum (CUDALucas.cu:1712-1733)                          */
/* ===================================================================== */

/* MSB-first CRC, polynomial 0x04C11DB7, initial value 0, no reflection, no
 * final xor, over the bytes in order. */
uint32_t orc_checkpoint_crc(const void *bytes, size_t len) {
  const uint8_t *s = (const uint8_t *)bytes;
  uint32_t crc = 0;
  for (size_t i = 0; i < len; i++) {
    for (int b = 7; b >= 0; b--) {
      uint32_t in = (s[i] >> b) & 1u;
      if ((crc >> 31) == in) crc <<= 1;
      else crc = (crc << 1) ^ 0x04C11DB7u;
    }
  }
  return crc;
}

/* ===================================================================== */
/* FFT-length policy (CUDALucas.cu:955-971, 1027-1068)                   */
/* ===================================================================== */

/* The reference's default length list in units of 1024 words. */
static const int k_default_mult[] = {
    1,     2,     4,     8,     10,    14,    16,    18,    20,    32,    36,
    42,    48,    50,    56,    60,    64,    70,    80,    84,    96,    112,
    120,   126,   128,   144,   160,   162,   168,   180,   192,   224,   256,
    288,   320,   324,   336,   360,   384,   392,   400,   448,   512,   576,
    640,   648,   672,   720,   768,   784,   800,   864,   896,   900,   1024,
    1152,  1176,  1280,  1296,  1344,  1440,  1568,  1600,  1728,  1792,  2048,
    2160,  2304,  2352,  2592,  2688,  2880,  3024,  3136,  3200,  3584,  3600,
    4096,  4320,  4608,  4704,  5120,  5184,  5600,  5760,  6048,  6144,  6272,
    6400,  6480,  7168,  7200,  7776,  8064,  8192,  8640,  9216,  9408,  10240,
    10368, 10584, 10800, 11200, 11520, 12096, 12288, 12544, 12960, 13824, 14336,
    14400, 16384, 17496, 18144, 19208, 19600, 20000, 20250, 21952, 23328, 23814,
    24300, 24500, 25088, 25600, 26244, 27000, 27216, 28000, 28672, 31104, 31250,
    32000, 32400, 32768, 33614, 34992, 36000, 36288, 38416, 39200, 39366, 40500,
    41472, 42336, 43200, 43904, 47628, 49000, 50000, 50176, 51200, 52488, 54432,
    55296, 56000, 57344, 60750, 62500, 64000, 64800, 65536};

int orc_default_fft_count(void) {
  return (int)(sizeof(k_default_mult) / sizeof(k_default_mult[0]));
}

int orc_default_fft_length(int index) {
  if (index < 0 || index >= orc_default_fft_count()) return -1;
  return k_default_mult[index] * 1024;
}

void orc_fft_bounds(uint32_t p, int32_t *lb, int32_t *ub) {
  const double ff1 = 0.999 * 1024.0 * 0.0000357822505975293;
  const double ff2 = 0.99 * 1024.0 * 0.0002670641830112380;
  const double e1 = 1.022179977969700;
  const double e2 = 0.929905288591965;
  *lb = (int32_t)ceil(ff1 * exp(e1 * log((double)p)));
  *ub = (int32_t)floor(ff2 * exp(e2 * log((double)p)));
}

/* ===================================================================== */
/* accurate trigonometry and weights                                     */
/* ===================================================================== */

/* cos and sin of 2*pi*e/L from an exact octant reduction of the rational
 * e/L followed by long double evaluation on [0, pi/4]. */
static void cis_2pi_frac(uint64_t e, uint64_t L, double *c, double *s) {
  e %= L;
  uint64_t o = (8 * e) / L;        /* octant */
  uint64_t r = 8 * e - o * L;      /* 0 <= r < L */
  long double q = 0.78539816339744830961566084581987572L; /* pi/4 */
  long double t = q * (long double)r / (long double)L;
  long double tp = q * (long double)(L - r) / (long double)L;
  long double ct = cosl(t), st = sinl(t), ctp = cosl(tp), stp = sinl(tp);
  long double cc, ss;
  switch (o) {
    case 0: cc = ct; ss = st; break;
    case 1: cc = stp; ss = ctp; break;
    case 2: cc = -st; ss = ct; break;
    case 3: cc = -ctp; ss = stp; break;
    case 4: cc = -ct; ss = -st; break;
    case 5: cc = -stp; ss = -ctp; break;
    case 6: cc = st; ss = -ct; break;
    default: cc = ctp; ss = -stp; break;
  }
  *c = (double)cc;
  *s = (double)ss;
}

/* ===================================================================== */
/* mixed-radix complex FFT, exponent sign +1, unnormalised               */
/* (stands in for cufftExecZ2Z(..., CUFFT_INVERSE), CUDALucas.cu:1888)   */
/* ===================================================================== */

typedef struct {
  uint32_t m;
  int nrad;
  int rad[64];
  double *tw;      /* m complex: exp(+2 pi i e / m) */
  double *scratch; /* m complex */
} fft_plan;

static int fft_plan_init(fft_plan *pl, uint32_t m) {
  uint32_t r = m;
  pl->m = m;
  pl->nrad = 0;
  pl->tw = NULL;
  pl->scratch = NULL;
  while (r % 4 == 0) { pl->rad[pl->nrad++] = 4; r /= 4; }
  while (r % 2 == 0) { pl->rad[pl->nrad++] = 2; r /= 2; }
  while (r % 3 == 0) { pl->rad[pl->nrad++] = 3; r /= 3; }
  while (r % 5 == 0) { pl->rad[pl->nrad++] = 5; r /= 5; }
  while (r % 7 == 0) { pl->rad[pl->nrad++] = 7; r /= 7; }
  if (r != 1) return -1;
  pl->tw = (double *)malloc(sizeof(double) * 2 * m);
  pl->scratch = (double *)malloc(sizeof(double) * 2 * m);
  if (!pl->tw || !pl->scratch) return -1;
  for (uint32_t e = 0; e < m; e++)
    cis_2pi_frac(e, m, &pl->tw[2 * e], &pl->tw[2 * e + 1]);
  return 0;
}

static void fft_plan_free(fft_plan *pl) {
  free(pl->tw);
  free(pl->scratch);
}

/* Stockham autosort passes: pass with radix R and span ns reads
 * x[j + i*m/R], twiddles by w^(i*(j mod ns)) with w = exp(+2 pi i/(ns R)),
 * takes a length-R DFT and writes to (j/ns)*ns*R + j mod ns + u*ns. */
static void fft_run(fft_plan *pl, double *data) {
  uint32_t m = pl->m, ns = 1;
  double *in = data, *out = pl->scratch;
  for (int pass = 0; pass < pl->nrad; pass++) {
    uint32_t R = (uint32_t)pl->rad[pass];
    uint32_t stride = m / R;
    uint32_t twstep = m / (ns * R);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (uint32_t j = 0; j < stride; j++) {
      uint32_t k = j % ns;
      double vr[7], vi[7];
      for (uint32_t i = 0; i < R; i++) {
        double ar = in[2 * (j + i * stride)], ai = in[2 * (j + i * stride) + 1];
        uint64_t e = (uint64_t)i * k * twstep;
        double wr = pl->tw[2 * e], wi = pl->tw[2 * e + 1];
        vr[i] = ar * wr - ai * wi;
        vi[i] = ar * wi + ai * wr;
      }
      uint32_t base = (j / ns) * ns * R + k;
      if (R == 2) {
        out[2 * base] = vr[0] + vr[1];
        out[2 * base + 1] = vi[0] + vi[1];
        out[2 * (base + ns)] = vr[0] - vr[1];
        out[2 * (base + ns) + 1] = vi[0] - vi[1];
      } else if (R == 4) {
        double s0r = vr[0] + vr[2], s0i = vi[0] + vi[2];
        double d0r = vr[0] - vr[2], d0i = vi[0] - vi[2];
        double s1r = vr[1] + vr[3], s1i = vi[1] + vi[3];
        double d1r = vr[1] - vr[3], d1i = vi[1] - vi[3];
        out[2 * base] = s0r + s1r;
        out[2 * base + 1] = s0i + s1i;
        /* u=1: d0 + i d1 */
        out[2 * (base + ns)] = d0r - d1i;
        out[2 * (base + ns) + 1] = d0i + d1r;
        out[2 * (base + 2 * ns)] = s0r - s1r;
        out[2 * (base + 2 * ns) + 1] = s0i - s1i;
        /* u=3: d0 - i d1 */
        out[2 * (base + 3 * ns)] = d0r + d1i;
        out[2 * (base + 3 * ns) + 1] = d0i - d1r;
      } else {
        /* direct length-R DFT with exp(+2 pi i iu/R) from the table */
        uint32_t rstep = m / R;
        for (uint32_t u = 0; u < R; u++) {
          double sr = 0, si = 0;
          for (uint32_t i = 0; i < R; i++) {
            uint64_t e = ((uint64_t)i * u % R) * rstep;
            double wr = pl->tw[2 * e], wi = pl->tw[2 * e + 1];
            sr += vr[i] * wr - vi[i] * wi;
            si += vr[i] * wi + vi[i] * wr;
          }
          out[2 * (base + u * ns)] = sr;
          out[2 * (base + u * ns) + 1] = si;
        }
      }
    }
    double *t = in; in = out; out = t;
    ns *= R;
  }
  if (in != data) memcpy(data, in, sizeof(double) * 2 * m);
}

int orc_fft_inverse(uint32_t m, double *data) {
  fft_plan pl;
  if (fft_plan_init(&pl, m)) { fft_plan_free(&pl); return -1; }
  fft_run(&pl, data);
  fft_plan_free(&pl);
  return 0;
}

/* ===================================================================== */
/* CPU IBDWT iteration                                                   */
/* ===================================================================== */

struct orc_ibdwt {
  uint32_t p, n, offset;
  double *x;     /* n weighted words */
  double *ttp;   /* 2^(f(j)/n) */
  double *ttmp;  /* 2^(-f(j)/n) * 2/n * (-1)^j */
  double *ct;    /* n/4: 0.5 cos(2 pi j / n) */
  uint8_t *width;
  int64_t *v;
  fft_plan plan;
};

/* write_gpu_data (CUDALucas.cu:1125-1252), one weight per word computed
 * directly instead of the group x column factorisation. */
orc_ibdwt *orc_ibdwt_create(uint32_t p, uint32_t n) {
  if (n % 4 || n < 8) return NULL;
  orc_ibdwt *c = (orc_ibdwt *)calloc(1, sizeof(orc_ibdwt));
  if (!c) return NULL;
  c->p = p;
  c->n = n;
  c->x = (double *)calloc(n, sizeof(double));
  c->ttp = (double *)malloc(sizeof(double) * n);
  c->ttmp = (double *)malloc(sizeof(double) * n);
  c->ct = (double *)malloc(sizeof(double) * (n / 4));
  c->width = (uint8_t *)malloc(n);
  c->v = (int64_t *)malloc(sizeof(int64_t) * n);
  if (!c->x || !c->ttp || !c->ttmp || !c->ct || !c->width || !c->v ||
      fft_plan_init(&c->plan, n / 2)) {
    orc_ibdwt_destroy(c);
    return NULL;
  }
  uint64_t b = p % n;
  for (uint32_t j = 0; j < n; j++) {
    uint64_t f = (n - (b * j) % n) % n;
    long double e = (long double)f / (long double)n;
    c->ttp[j] = (double)exp2l(e);
    long double im = exp2l(-e) * 2.0L / (long double)n;
    c->ttmp[j] = (double)((j & 1) ? -im : im);
    c->width[j] = (uint8_t)(p / n + (f < b ? 1 : 0));
  }
  c->ct[0] = 0.0;
  for (uint32_t j = 1; j < n / 4; j++) {
    double cs, sn;
    cis_2pi_frac(j, n, &cs, &sn);
    c->ct[j] = 0.5 * cs;
  }
  return c;
}

void orc_ibdwt_destroy(orc_ibdwt *c) {
  if (!c) return;
  free(c->x); free(c->ttp); free(c->ttmp); free(c->ct); free(c->width);
  free(c->v);
  if (c->plan.tw || c->plan.scratch) fft_plan_free(&c->plan);
  free(c);
}

/* init_x / apply_weights (CUDALucas.cu:216-249, 1254-1278). */
void orc_ibdwt_load(orc_ibdwt *c, const int32_t *digits, uint32_t offset) {
  c->offset = offset % c->p;
  for (uint32_t j = 0; j < c->n; j++) {
    c->v[j] = digits[j];
    c->x[j] = (double)digits[j] * c->ttp[j];
  }
}

/* square (CUDALucas.cu:255-314) in complex notation.  a holds m = n/2
 * complex values from the first transform; j pairs with m - j and the
 * output is conjugated so that the second transform is also INVERSE.
 *   D = A - conj(B),  A' = A - W D,  B' = B + conj(W D),
 *   SA = A'^2, SB = B'^2, E = SA - conj(SB), F = conj(W) E,
 *   A <- conj(SA - F),  B <- conj(SB) + F,
 * with W = (0.5 - ct[n/4 - j]) + i ct[j]. */
static void square_split(double *a, uint32_t n, const double *ct) {
  uint32_t m = n / 2;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (uint32_t j = 1; j < n / 4; j++) {
    double wr = 0.5 - ct[n / 4 - j], wi = ct[j];
    double ar = a[2 * j], ai = a[2 * j + 1];
    double br = a[2 * (m - j)], bi = a[2 * (m - j) + 1];
    double dr = ar - br, di = ai + bi;
    double tr = wr * dr - wi * di, ti = wr * di + wi * dr;
    ar -= tr; ai -= ti;
    br += tr; bi -= ti;
    double sar = (ar - ai) * (ar + ai), sai = 2.0 * ar * ai;
    double sbr = (br - bi) * (br + bi), sbi = 2.0 * br * bi;
    double er = sar - sbr, ei = sai + sbi;
    double fr = wr * er + wi * ei, fi = wr * ei - wi * er;
    a[2 * j] = sar - fr;
    a[2 * j + 1] = fi - sai;
    a[2 * (m - j)] = sbr + fr;
    a[2 * (m - j) + 1] = fi - sbi;
  }
  /* DC / Nyquist packed in element 0 */
  {
    double s = a[0] + a[1], d = a[0] - a[1];
    double s2 = s * s, d2 = d * d;
    double h = 0.5 * (d2 - s2);
    a[0] = s2 + h;
    a[1] = h;
  }
  /* middle element m/2 */
  {
    double xr = a[m], xi = a[m + 1];
    a[m] = (xr + xi) * (xr - xi);
    a[m + 1] = -2.0 * xr * xi;
  }
}

/* lucas_square (CUDALucas.cu:1869-1986) + rcb/splice (CUDALucas.cu:322-534)
 * with a sequential, fully balanced carry (exact) in place of the
 * block-parallel carry + splice. */
double orc_ibdwt_iterate(orc_ibdwt *c, uint64_t count, double *per_iter) {
  uint32_t p = c->p, n = c->n;
  double maxall = 0.0;
  for (uint64_t it = 0; it < count; it++) {
    c->offset = (uint32_t)((2ull * c->offset) % p);
    uint64_t bit = ((uint64_t)c->offset + 1) % p;
    uint32_t digit = (uint32_t)((bit * n) / p);
    uint32_t pos = (uint32_t)(bit - digit_start(p, n, digit));

    fft_run(&c->plan, c->x);
    square_split(c->x, n, c->ct);
    fft_run(&c->plan, c->x);

    double maxerr = 0.0;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1) reduction(max : maxerr)
    for (uint32_t j = 0; j < n; j++) {
      double t = c->x[j] * c->ttmp[j];
      double r = rint(t);
      double e = fabs(t - r);
      if (e > maxerr) maxerr = e;
      c->v[j] = (int64_t)r;
    }
    c->v[digit] -= (int64_t)1 << pos;
    int64_t carry = 0;
    int wrapped = 0;
    for (uint32_t j = 0;; j++) {
      if (j == n) { j = 0; wrapped = 1; }
      if (wrapped && carry == 0) break;
      int w = c->width[j];
      int64_t t = c->v[j] + carry;
      int64_t half = (int64_t)1 << (w - 1);
      int64_t d = ((t + half) & (((int64_t)1 << w) - 1)) - half;
      carry = (t - d) >> w;
      c->v[j] = d;
    }
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (uint32_t j = 0; j < n; j++) c->x[j] = (double)c->v[j] * c->ttp[j];
    if (per_iter) per_iter[it] = maxerr;
    if (maxerr > maxall) maxall = maxerr;
  }
  return maxall;
}

void orc_ibdwt_read(orc_ibdwt *c, int32_t *digits, uint32_t *offset) {
  for (uint32_t j = 0; j < c->n; j++) digits[j] = (int32_t)c->v[j];
  if (offset) *offset = c->offset;
}

// Host-side arithmetic of the LL engine: variable-base digit codec, Res64,
// checkpoint checksum, FFT-length policy, the worktodo file and the
// extended-precision tables (roots of unity, IBDWT weights) the kernels read.
#include <ctype.h>
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>
#include <unistd.h>

#include <mutex>
#include <string>
#include <vector>

#include "ll_api.h"
#include "ll_host.h"

namespace llamd {

// ------------------------------------------------------------------ tables

// cos and sin of -2 pi e / L (the forward root), from an exact octant
// reduction of the rational e/L and long double evaluation on [0, pi/4].
void root_of_unity(uint64_t e, uint64_t L, double *re, double *im) {
  long double r, i;
  root_of_unity_ld(e, L, &r, &i);
  *re = (double)r;
  *im = (double)i;
}

void root_of_unity_ld(uint64_t e, uint64_t L, long double *re, long double *im) {
  e %= L;
  const uint64_t o = (8 * e) / L;
  const uint64_t r = 8 * e - o * L;
  const long double quarter_pi = 0.78539816339744830961566084581987572L;
  const long double t = quarter_pi * (long double)r / (long double)L;
  const long double u = quarter_pi * (long double)(L - r) / (long double)L;
  const long double ct = cosl(t), st = sinl(t), cu = cosl(u), su = sinl(u);
  long double c, s;  // cos, sin of +2 pi e / L
  switch (o) {
    case 0: c = ct; s = st; break;
    case 1: c = su; s = cu; break;
    case 2: c = -st; s = ct; break;
    case 3: c = -cu; s = su; break;
    case 4: c = -ct; s = -st; break;
    case 5: c = -su; s = -cu; break;
    case 6: c = st; s = -ct; break;
    default: c = cu; s = -su; break;
  }
  *re = c;
  *im = -s;
}

// exp(-2 pi i e / L) - 1 for small angles, computed as
// (-2 sin^2(theta/2), -sin(theta)) to keep the small real part accurate.
void root_minus_one(uint64_t e, uint64_t L, double *re, double *im) {
  const long double two_pi = 6.28318530717958647692528676655900577L;
  const long double th = two_pi * (long double)e / (long double)L;
  const long double h = sinl(th / 2);
  *re = (double)(-2.0L * h * h);
  *im = (double)(-sinl(th));
}

void roots_table(uint32_t L, std::vector<double> &out) {
  out.resize(2 * (size_t)L);
  for (uint32_t e = 0; e < L; e++) root_of_unity(e, L, &out[2 * e], &out[2 * e + 1]);
}

void quarter_cosines(uint32_t L, std::vector<double> &out) {
  out.resize(L / 4 + 1);
  for (uint32_t e = 0; e <= L / 4; e++) {
    double re, im;
    root_of_unity(e, L, &re, &im);
    out[e] = re;
  }
}

uint32_t weight_exponent(uint32_t p, uint32_t n, uint64_t j) {
  const uint64_t b = p % n;
  return (uint32_t)((n - (b * j) % n) % n);
}

double inverse_weight_scaled(uint32_t f, uint32_t n, uint64_t four_m) {
  return (double)(exp2l(-(long double)f / (long double)n) / (long double)four_m);
}

void weight_pair(uint32_t f, uint32_t n, double *w, double *iw) {
  const long double e = (long double)f / (long double)n;
  *w = (double)exp2l(e);
  *iw = (double)exp2l(-e);
}

// --------------------------------------------------------------- codec

static uint64_t digit_start(uint32_t p, uint32_t n, uint64_t j) {
  return ((uint64_t)p * j + n - 1) / n;
}

}  // namespace llamd

using namespace llamd;

extern "C" {

int ll_digit_width(uint32_t p, uint32_t n, uint32_t j) {
  const uint32_t b = p % n;
  return (int)(p / n) + (weight_exponent(p, n, j) < b ? 1 : 0);
}

void ll_unpack_bits(uint32_t p, uint32_t n, const uint32_t *packed, int32_t *digits) {
  const size_t nw = (p + 31) / 32;
  uint64_t acc = 0;
  int have = 0;
  size_t w = 0;
  for (uint32_t j = 0; j < n; j++) {
    const int width = ll_digit_width(p, n, j);
    while (have < width) {
      acc |= (uint64_t)(w < nw ? packed[w] : 0u) << have;
      have += 32;
      w++;
    }
    digits[j] = (int32_t)(acc & ((1ull << width) - 1));
    acc >>= width;
    have -= width;
  }
}

void ll_balance_digits(uint32_t p, uint32_t n, int32_t *digits) {
  int32_t carry = 0;
  for (uint32_t j = 0; j < n; j++) {
    const int width = ll_digit_width(p, n, j);
    const int32_t half = 1 << (width - 1);
    digits[j] += carry;
    carry = 0;
    if (digits[j] >= half) {
      digits[j] -= 2 * half;
      carry = 1;
    }
  }
  digits[0] += carry;
}

int ll_standardize_digits(uint32_t p, uint32_t n, int32_t *digits) {
  std::vector<int64_t> d(digits, digits + n);
  std::vector<uint8_t> width(n);
  for (uint32_t j = 0; j < n; j++) width[j] = (uint8_t)ll_digit_width(p, n, j);
  for (int lap = 0; lap < 64; lap++) {
    int64_t carry = 0;
    for (uint32_t j = 0; j < n; j++) {
      const int64_t t = d[j] + carry;
      d[j] = t & (((int64_t)1 << width[j]) - 1);
      carry = t >> width[j];  // floor division
    }
    if (carry == 0) break;
    d[0] += carry;  // 2^p == 1
  }
  bool all_ones = true, zero = true;
  for (uint32_t j = 0; j < n; j++) {
    if (d[j] != (((int64_t)1 << width[j]) - 1)) all_ones = false;
    if (d[j] != 0) zero = false;
  }
  if (all_ones) {
    std::fill(d.begin(), d.end(), 0);
    zero = true;
  }
  for (uint32_t j = 0; j < n; j++) digits[j] = (int32_t)d[j];
  return zero ? 1 : 0;
}

void ll_pack_bits(uint32_t p, uint32_t n, const int32_t *digits, uint32_t *packed) {
  const size_t nw = (p + 31) / 32;
  uint64_t acc = 0;
  int have = 0;
  size_t w = 0;
  for (uint32_t j = 0; j < n; j++) {
    acc |= (uint64_t)(uint32_t)digits[j] << have;
    have += ll_digit_width(p, n, j);
    while (have >= 32) {
      if (w < nw) packed[w] = (uint32_t)acc;
      w++;
      acc >>= 32;
      have -= 32;
    }
  }
  if (w < nw) packed[w++] = (uint32_t)acc;
  while (w < nw) packed[w++] = 0;
}

int ll_residue_words(uint32_t p, uint32_t n, const int32_t *digits, uint32_t offset,
                     uint32_t *words) {
  if (p < 3 || n == 0) return LL_ERR_ARG;
  std::vector<int32_t> d(digits, digits + n);
  ll_standardize_digits(p, n, d.data());
  const size_t nw = (p + 31) / 32;
  std::vector<uint32_t> v(nw + 1, 0);
  ll_pack_bits(p, n, d.data(), v.data());
  // rotate right by offset within p bits: out bit i = v bit (i + offset) mod p
  offset %= p;
  memset(words, 0, nw * 4);
  auto bit = [&](uint32_t k) { return (v[k >> 5] >> (k & 31)) & 1u; };
  for (uint32_t i = 0; i < p; i++) {
    uint32_t s = i + offset;
    if (s >= p) s -= p;
    if (bit(s)) words[i >> 5] |= 1u << (i & 31);
  }
  return LL_OK;
}

uint64_t ll_res64(uint32_t p, uint32_t n, const int32_t *digits, uint32_t offset) {
  std::vector<uint32_t> w((p + 31) / 32 + 2, 0);
  if (ll_residue_words(p, n, digits, offset, w.data())) return 0;
  return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

void ll_initial_digits(uint32_t p, uint32_t n, uint32_t offset, int32_t *digits) {
  const uint64_t bit = ((uint64_t)offset + 2) % p;
  const uint32_t digit = (uint32_t)((bit * n) / p);
  memset(digits, 0, sizeof(int32_t) * n);
  digits[digit] = (int32_t)(1u << (bit - digit_start(p, n, digit)));
}

uint32_t ll_checkpoint_crc(const void *bytes, size_t len) {
  const uint8_t *s = (const uint8_t *)bytes;
  uint32_t crc = 0;
  for (size_t i = 0; i < len; i++) {
    for (int b = 7; b >= 0; b--) {
      const uint32_t top = crc >> 31;
      crc <<= 1;
      if (top != ((s[i] >> b) & 1u)) crc ^= 0x04C11DB7u;
    }
  }
  return crc;
}

// ------------------------------------------------------ checkpoint files

size_t ll_checkpoint_words(uint32_t p) { return (p + 31) / 32 + 10; }

int ll_checkpoint_encode(const ll_checkpoint *m, const int32_t *digits, uint32_t *words) {
  if (!m || !digits || !words || m->p < 3 || m->n == 0) return LL_ERR_ARG;
  const size_t end = (m->p + 31) / 32;
  std::vector<int32_t> d(digits, digits + m->n);
  ll_standardize_digits(m->p, m->n, d.data());
  ll_pack_bits(m->p, m->n, d.data(), words);
  words[end + 0] = m->p;
  words[end + 1] = m->n;
  words[end + 2] = m->next_iter;
  words[end + 3] = m->offset;
  words[end + 4] = (uint32_t)(m->total_us >> 15);
  words[end + 5] = (uint32_t)(m->time_adj_us >> 15);
  words[end + 6] = m->iter_adj;
  words[end + 7] = 0;
  words[end + 8] = 0;  // magic number: unused by the format
  words[end + 9] = ll_checkpoint_crc(words, 4 * (end + 9));
  return LL_OK;
}

int ll_checkpoint_decode(uint32_t p, const uint32_t *words, ll_checkpoint *m) {
  if (!words || !m || p < 3) return LL_ERR_ARG;
  const size_t end = (p + 31) / 32;
  if (words[end + 9] != ll_checkpoint_crc(words, 4 * (end + 9))) return LL_ERR_ARG;
  if (words[end] != p) return LL_ERR_ARG;
  m->p = p;
  m->n = words[end + 1];
  m->next_iter = words[end + 2];
  m->offset = words[end + 3];
  m->total_us = (uint64_t)words[end + 4] << 15;
  m->time_adj_us = (uint64_t)words[end + 5] << 15;
  m->iter_adj = words[end + 6];
  if (m->offset >= p) return LL_ERR_ARG;
  return LL_OK;
}

int ll_checkpoint_digits(uint32_t p, uint32_t n, const uint32_t *words, int32_t *digits) {
  if (!words || !digits || n == 0) return LL_ERR_ARG;
  ll_unpack_bits(p, n, words, digits);
  ll_balance_digits(p, n, digits);
  return LL_OK;
}

static void ckpt_path(char *buf, size_t len, const char *dir, char kind, uint32_t p) {
  if (dir && *dir) snprintf(buf, len, "%s/%c%u", dir, kind, p);
  else snprintf(buf, len, "%c%u", kind, p);
}

int ll_checkpoint_write(const char *dir, uint32_t p, const uint32_t *words, const char *save_dir,
                        uint64_t res64) {
  if (!words) return LL_ERR_ARG;
  const size_t nbytes = 4 * ll_checkpoint_words(p);
  char c[1024], t[1024], tmp[1040];
  ckpt_path(c, sizeof c, dir, 'c', p);
  ckpt_path(t, sizeof t, dir, 't', p);
  snprintf(tmp, sizeof tmp, "%s.new", c);
  FILE *f = fopen(tmp, "wb");
  if (!f) return LL_ERR_ARG;
  const bool ok = fwrite(words, 1, nbytes, f) == nbytes;
  if (fclose(f) != 0 || !ok) {
    remove(tmp);
    return LL_ERR_ARG;
  }
  remove(t);
  rename(c, t);
  if (rename(tmp, c) != 0) return LL_ERR_ARG;
  if (save_dir && *save_dir) {
    const size_t end = (p + 31) / 32;
    mkdir(save_dir, 0755);
    char sname[1100];
    snprintf(sname, sizeof sname, "%s/s%u.%u.%016llx", save_dir, p, words[end + 2] - 1,
             (unsigned long long)res64);
    FILE *g = fopen(sname, "wb");
    if (!g) return LL_ERR_ARG;
    const bool ok2 = fwrite(words, 1, nbytes, g) == nbytes;
    if (fclose(g) != 0 || !ok2) return LL_ERR_ARG;
  }
  return LL_OK;
}

int ll_checkpoint_read(const char *dir, uint32_t p, uint32_t *words) {
  if (!words) return LL_ERR_ARG;
  const size_t nbytes = 4 * ll_checkpoint_words(p);
  for (int which = 0; which < 2; which++) {
    char path[1024];
    ckpt_path(path, sizeof path, dir, which ? 't' : 'c', p);
    FILE *f = fopen(path, "rb");
    if (!f) continue;
    const size_t got = fread(words, 1, nbytes, f);
    fclose(f);
    ll_checkpoint meta;
    if (got == nbytes && ll_checkpoint_decode(p, words, &meta) == LL_OK) return which;
  }
  return LL_ERR_ARG;
}

void ll_checkpoint_remove(const char *dir, uint32_t p) {
  char c[1024], t[1024];
  ckpt_path(c, sizeof c, dir, 'c', p);
  ckpt_path(t, sizeof t, dir, 't', p);
  remove(c);
  remove(t);
}

// ------------------------------------------------------ FFT-length policy

// The reference's 162 default lengths in units of 1024 words
// (CUDALucas.cu:955-971).
static const int k_lengths_k[] = {
    1,     2,     4,     8,     10,    14,    16,    18,    20,    32,    36,    42,    48,
    50,    56,    60,    64,    70,    80,    84,    96,    112,   120,   126,   128,   144,
    160,   162,   168,   180,   192,   224,   256,   288,   320,   324,   336,   360,   384,
    392,   400,   448,   512,   576,   640,   648,   672,   720,   768,   784,   800,   864,
    896,   900,   1024,  1152,  1176,  1280,  1296,  1344,  1440,  1568,  1600,  1728,  1792,
    2048,  2160,  2304,  2352,  2592,  2688,  2880,  3024,  3136,  3200,  3584,  3600,  4096,
    4320,  4608,  4704,  5120,  5184,  5600,  5760,  6048,  6144,  6272,  6400,  6480,  7168,
    7200,  7776,  8064,  8192,  8640,  9216,  9408,  10240, 10368, 10584, 10800, 11200, 11520,
    12096, 12288, 12544, 12960, 13824, 14336, 14400, 16384, 17496, 18144, 19208, 19600, 20000,
    20250, 21952, 23328, 23814, 24300, 24500, 25088, 25600, 26244, 27000, 27216, 28000, 28672,
    31104, 31250, 32000, 32400, 32768, 33614, 34992, 36000, 36288, 38416, 39200, 39366, 40500,
    41472, 42336, 43200, 43904, 47628, 49000, 50000, 50176, 51200, 52488, 54432, 55296, 56000,
    57344, 60750, 62500, 64000, 64800, 65536};

// The active list: the defaults, or one installed by ll_fft_length_list_set
// (the analog of init_ffts merging "<device> fft.txt", CUDALucas.cu:947-1025).
static std::vector<int> &length_list() {
  static std::vector<int> v = [] {
    std::vector<int> d;
    for (int k : k_lengths_k) d.push_back(k * 1024);
    return d;
  }();
  return v;
}

int ll_fft_length_list_set(uint32_t count, const uint32_t *lengths) {
  std::vector<int> &v = length_list();
  if (count == 0) {
    v.clear();
    for (int k : k_lengths_k) v.push_back(k * 1024);
    return LL_OK;
  }
  if (!lengths) return LL_ERR_ARG;
  for (uint32_t i = 0; i < count; i++)
    if (lengths[i] == 0 || lengths[i] > 0x7fffffffu || (i && lengths[i] <= lengths[i - 1])) return LL_ERR_ARG;
  v.assign(lengths, lengths + count);
  return LL_OK;
}

int ll_fft_length_count(void) { return (int)length_list().size(); }

int ll_fft_length_at(int index) {
  if (index < 0 || index >= ll_fft_length_count()) return -1;
  return length_list()[index];
}

void ll_fft_bounds(uint32_t p, int32_t *lb, int32_t *ub) {
  // power laws fitted by the reference (CUDALucas.cu:1034-1050)
  const double ff1 = 0.999 * 1024.0 * 0.0000357822505975293;
  const double ff2 = 0.99 * 1024.0 * 0.0002670641830112380;
  const double e1 = 1.022179977969700;
  const double e2 = 0.929905288591965;
  if (lb) *lb = (int32_t)ceil(ff1 * exp(e1 * log((double)p)));
  if (ub) *ub = (int32_t)floor(ff2 * exp(e2 * log((double)p)));
}

int ll_choose_fft_length(uint32_t p, int *index) {
  const int count = ll_fft_length_count();
  int32_t lb, ub;
  ll_fft_bounds(p, &lb, &ub);
  int lbi = 0;
  while (lbi < count && ll_fft_length_at(lbi) < lb) lbi++;
  int ubi = lbi;
  while (ubi < count && ll_fft_length_at(ubi) <= ub) ubi++;
  ubi--;
  int i;
  if (*index >= count) {  // manual length
    i = 0;
    while (i < count && ll_fft_length_at(i) < *index) i++;
  } else {
    i = *index < 0 ? lbi : *index;
  }
  if (i < lbi) i = lbi;
  if (ubi >= lbi && i > ubi) i = ubi;
  while (i < count && !ll_fft_supported((uint32_t)ll_fft_length_at(i))) i++;
  if (i >= count) return LL_ERR_UNSUPPORTED;
  *index = i;
  return ll_fft_length_at(i);
}

int ll_fft_from_str(const char *s) {
  if (!s || !*s) return -1;
  char *end = nullptr;
  const unsigned long v = strtoul(s, &end, 10);
  if (end == s) return -1;
  long mult = 1;
  if (*end == 'k' || *end == 'K') { mult = 1024; end++; }
  else if (*end == 'm' || *end == 'M') { mult = 1024 * 1024; end++; }
  if (*end != '\0') return -1;
  const unsigned long long n = (unsigned long long)v * mult;
  if (n > 0x7fffffffull) return -1;
  return (int)n;
}

// ---------------------------------------------------------------- worktodo

// A work line holds at most 190 characters before its newline: the
// reference reads lines into a 192-byte buffer and rejects a line that fills
// it (MAX_LINE_LENGTH, parse.c:67, 280-283).
static const size_t kWorkLineMax = 190;

static void copy_text(char *dst, const char *b, const char *e) {
  while (e > b && isspace((unsigned char)e[-1])) e--;
  size_t k = (size_t)(e - b);
  if (k >= LL_WORK_TEXT) k = LL_WORK_TEXT - 1;
  memcpy(dst, b, k);
  dst[k] = 0;
}

// A number field: digits from b, then nothing but an optional unit letter at
// u (K or M, u == nullptr for a plain number) and ignored trailing text.
// Returns LL_WORK_OK and the value, or the status that rejects the field
// (parse.c:372-430).
static int work_number(const char *b, const char *u, unsigned long long scale, unsigned long long &v) {
  char *stop = nullptr;
  errno = 0;
  const unsigned long x = strtoul(b, &stop, 10);
  if (stop == b) return LL_WORK_INVALID_FORMAT;
  if (u && stop < u) return LL_WORK_INVALID_FORMAT;  // "40 96K", "4096 K"
  if (errno || x > 0x7fffffffull / scale) return LL_WORK_INVALID_DATA;
  v = (unsigned long long)x * scale;
  return LL_WORK_OK;
}

int ll_work_parse_line(const char *line, ll_assignment *a) {
  if (!a) return LL_ERR_ARG;
  memset(a, 0, sizeof *a);
  if (!line) return LL_WORK_EOF;
  const size_t len = strcspn(line, "\n");
  if (len > kWorkLineMax) return LL_WORK_LONG_LINE;
  char text[kWorkLineMax + 1];  // the line alone, so no scan runs past it
  memcpy(text, line, len);
  text[len] = 0;
  const char *end = text + len;
  const char *s = text;
  while (s < end && isspace((unsigned char)*s)) s++;
  if (s == end) return LL_WORK_BLANK;
  if (*s == '#' || (end - s >= 2 && (s[0] == '/' || s[0] == '\\') && s[1] == s[0])) return LL_WORK_COMMENT;
  if (strncasecmp(s, "Test=", 5) && strncasecmp(s, "DoubleCheck=", 12)) return LL_WORK_NO_TEST_EQUAL;
  s = strchr(s, '=') + 1;
  // The assignment ends at a comment delimiter (#, // or \\).
  const char *stop = end, *note = nullptr;
  for (const char *c = s; c < end && !note; c++) {
    if (*c == '#') stop = c, note = c + 1;
    else if (c + 1 < end && (*c == '/' || *c == '\\') && c[1] == *c) stop = c, note = c + 2;
  }
  int commas = 0;
  for (const char *c = s; c < stop; c++) commas += *c == ',';
  if (commas > 4) return LL_WORK_INVALID_FORMAT;  // key, exponent, length, TF bits, P-1 done
  // Fields (parse.c:344-437): the first letter in a field decides its kind;
  // a-f marks the assignment key (hex digits or N/A), K or M after digits an
  // FFT length; a field without one is a number, and the largest number is
  // the exponent (trial-factoring depth and the P-1 flag are smaller).
  for (const char *f = s; f <= stop;) {
    while (f < stop && isspace((unsigned char)*f)) f++;
    const char *fe = f;
    while (fe < stop && *fe != ',') fe++;
    const char *kind = f;
    while (kind < fe && !strchr("abcdefABCDEFkKmM", *kind)) kind++;
    // A leading M is a length without digits, so "M<p>" is rejected as in
    // the reference (its M<p> exponent branch, parse.c:418-419, is never
    // reached).
    int st = LL_WORK_OK;
    unsigned long long v = 0;
    if (kind < fe && (*kind == 'k' || *kind == 'K')) {
      st = work_number(f, kind, 1024, v);
      a->fft_length = (uint32_t)v;
    } else if (kind < fe && (*kind == 'm' || *kind == 'M')) {
      st = work_number(f, kind, 1024 * 1024, v);
      a->fft_length = (uint32_t)v;
    } else if (kind < fe) {
      copy_text(a->key, f, fe);
    } else {
      st = work_number(f, nullptr, 1, v);
      if (st == LL_WORK_OK && v > a->exponent) a->exponent = (uint32_t)v;
    }
    if (st != LL_WORK_OK) return st;
    f = fe + 1;
  }
  if (note) {
    while (note < end && isspace((unsigned char)*note)) note++;
    copy_text(a->comment, note, end);
  }
  if (a->exponent < 100 || (a->fft_length > 0 && a->fft_length < 4096)) return LL_WORK_INVALID_DATA;
  return LL_WORK_OK;
}

int ll_is_prime(uint32_t n) {
  if (n < 2) return 0;
  if (n % 2 == 0) return n == 2;
  for (uint64_t d = 3; d * d <= n; d += 2)
    if (n % d == 0) return 0;
  return 1;
}

int ll_valid_assignment(uint32_t p, uint32_t fft_length) {
  if (p < 7500) {
    fprintf(stderr, "Warning: exponents < 7500 are not supported!\n");
    return 0;
  }
  if (fft_length % 1024) {
    fprintf(stderr, "Warning: FFT length %u is invalid, it must be a multiple of 1024.\n", fft_length);
    return 0;
  }
  if (!ll_is_prime(p)) {
    fprintf(stderr, "Warning: exponent is not prime!\n");
    return -1;
  }
  return 1;
}

// One line of f without its newline; false at the end of the file.  Lines
// of any length are read whole, so a long line is one rejected line (the
// reference's fixed buffer reads its remainder as a further line).
static bool work_read_line(FILE *f, std::string &line) {
  line.clear();
  char buf[256];
  while (fgets(buf, sizeof buf, f)) {
    line += buf;
    if (!line.empty() && line.back() == '\n') {
      line.pop_back();
      return true;
    }
  }
  return !line.empty();
}

static const char *work_reason(int st) {
  switch (st) {
    case LL_WORK_LONG_LINE: return "line is too long.";
    case LL_WORK_NO_TEST_EQUAL: return "doesn't begin with Test= or DoubleCheck=.";
    case LL_WORK_INVALID_FORMAT: return "invalid format.";
    case LL_WORK_INVALID_DATA: return "invalid data.";
    default: return "unknown error.";
  }
}

int ll_work_next(const char *file, ll_assignment *out) {
  if (!file || !out) return LL_ERR_ARG;
  FILE *f = fopen(file, "r");
  if (!f) {
    fprintf(stderr, "Can't open workfile %s\n\n", file);
    return 1;
  }
  std::string line;
  unsigned lineno = 0;
  while (work_read_line(f, line)) {
    lineno++;
    int st = ll_work_parse_line(line.c_str(), out);
    if (st == LL_WORK_BLANK || st == LL_WORK_COMMENT) continue;
    if (st == LL_WORK_OK) {
      if (ll_valid_assignment(out->exponent, out->fft_length) > 0) {
        fclose(f);
        return 0;
      }
      st = LL_WORK_INVALID_DATA;
    }
    fprintf(stderr, "Warning: ignoring line %u: \"%s\" in \"%s\". Reason: %s\n", lineno, line.c_str(), file,
            work_reason(st));
  }
  fclose(f);
  memset(out, 0, sizeof *out);
  fprintf(stderr, "No valid assignment found.\n\n");
  return 2;
}

int ll_work_clear(const char *file, uint32_t p) {
  if (!file) return LL_ERR_ARG;
  FILE *f = fopen(file, "r");
  if (!f) {
    fprintf(stderr, "Can't open workfile %s\n\n", file);
    return 3;
  }
  std::vector<std::string> lines;
  std::vector<int> status;
  std::vector<uint32_t> expo;
  std::string line;
  ll_assignment a;
  while (work_read_line(f, line)) {
    status.push_back(ll_work_parse_line(line.c_str(), &a));
    expo.push_back(a.exponent);
    lines.push_back(line);
  }
  fclose(f);
  // The target and the first line that goes with it (parse.c:592-615).
  size_t target = lines.size(), first = lines.size();
  bool blank_seen = false;
  for (size_t i = 0; i < lines.size(); i++) {
    if (status[i] == LL_WORK_OK) {
      if (expo[i] == p) {
        target = i;
        if (first > i) first = i;
        break;
      }
      first = i + 1;
      blank_seen = true;  // an assignment above: blank lines no longer matter
    } else if (status[i] == LL_WORK_BLANK && !blank_seen) {
      first = i + 1;
      blank_seen = true;
    }
  }
  if (target == lines.size()) {
    fprintf(stderr, "Assignment M%u completed but not found in workfile\n\n", p);
    return 5;
  }
  const std::string tmp = std::string(file) + ".tmp";
  FILE *o = fopen(tmp.c_str(), "w");
  if (!o) {
    fprintf(stderr, "Can't open tmp workfile\n\n");
    return 4;
  }
  bool ok = true;
  for (size_t i = 0; i < lines.size(); i++) {
    const bool keep = i > target || (i < target && (status[i] == LL_WORK_OK || i < first));
    if (keep) ok = ok && fprintf(o, "%s\n", lines[i].c_str()) >= 0;
  }
  ok = (fclose(o) == 0) && ok;
  if (!ok || rename(tmp.c_str(), file) != 0) {
    remove(tmp.c_str());
    fprintf(stderr, "Cannot move tmp workfile to regular workfile\n\n");
    return 6;
  }
  return 0;
}

// ------------------------------------------------------------ report line

// print_time_from_seconds (CUDALucas.cu:1429-1463): m:ss, h:mm:ss, or with
// days d:hh:mm:ss, right-aligned in 12 characters.
static std::string report_clock(uint64_t sec, bool days) {
  char t[48], out[64];
  const uint64_t d = days ? sec / 86400 : 0;
  if (days) sec %= 86400;
  const unsigned h = (unsigned)(sec / 3600), m = (unsigned)(sec % 3600 / 60), s = (unsigned)(sec % 60);
  if (d) snprintf(t, sizeof t, "%llu:%02u:%02u:%02u", (unsigned long long)d, h, m, s);
  else if (h) snprintf(t, sizeof t, "%u:%02u:%02u", h, m, s);
  else snprintf(t, sizeof t, "%u:%02u", m, s);
  snprintf(out, sizeof out, "%12s", t);
  return out;
}

int ll_format_report(const char *format, const ll_report *r, char *out, size_t cap) {
  if (!format || !r) return LL_ERR_ARG;
  const time_t now = (time_t)r->now;
  struct tm tm_now;
  localtime_r(&now, &tm_now);
  const double iters = r->interval_iterations ? (double)r->interval_iterations : 1.0;
  std::string line;
  char f[96];
  for (const char *c = format; *c;) {
    if (*c != '%') {
      line += *c++;
      continue;
    }
    const char code = c[1];
    if (!code) break;
    c += 2;
    f[0] = 0;
    double v = 0;
    const char *fixed = nullptr;  // a %.5f field: its printf format
    switch (code) {
      case 'a': strftime(f, sizeof f, "%a", &tm_now); break;
      case 'M': strftime(f, sizeof f, "%b", &tm_now); break;
      case 'D': strftime(f, sizeof f, "%d", &tm_now); break;
      case 'h': strftime(f, sizeof f, "%H", &tm_now); break;
      case 'm': strftime(f, sizeof f, ":%M", &tm_now); break;
      case 's': strftime(f, sizeof f, ":%S", &tm_now); break;
      case 'y': strftime(f, sizeof f, "%y", &tm_now); break;
      case 'z': strftime(f, sizeof f, "%Z", &tm_now); break;
      case 'p': snprintf(f, sizeof f, "%9u", r->p); break;
      case 'P': {
        char m[24];
        snprintf(m, sizeof m, "M%u", r->p);
        snprintf(f, sizeof f, "%10s", m);
        break;
      }
      case 'f': snprintf(f, sizeof f, "%5uK", r->n / 1024); break;
      case 'F': snprintf(f, sizeof f, "%8u", r->n); break;
      case 'i': snprintf(f, sizeof f, "%9u", r->iteration); break;
      case 'o': snprintf(f, sizeof f, "%9u", r->offset); break;
      case 'r': snprintf(f, sizeof f, "0x%016llx", (unsigned long long)r->residue); break;
      case 'R': snprintf(f, sizeof f, "0x%016llX", (unsigned long long)r->residue); break;
      case 'C': snprintf(f, sizeof f, "%s", report_clock(r->interval_us / 1000000, false).c_str()); break;
      case 'T': snprintf(f, sizeof f, "%s", report_clock(r->eta_s, true).c_str()); break;
      case 'N': snprintf(f, sizeof f, "%s", r->program ? r->program : ""); break;
      case '%': snprintf(f, sizeof f, "%%"); break;
      case 'x': fixed = "%7.5f", v = r->max_err; break;
      case 'q': fixed = "%9.5f", v = r->interval_us / 1000.0 / iters; break;
      case 'c': fixed = "%10.5f", v = r->interval_us / 1000000.0; break;
      case 'Q': fixed = "%10.5f", v = iters * 1000000.0 / (double)r->interval_us; break;
      case 'd': fixed = "%10.5f", v = r->iteration / (float)(r->p - 2) * 100; break;
      default: break;  // an unknown code prints nothing
    }
    if (fixed) {
      // The digit after the code keeps that many decimals (0 drops the
      // point too); anything else keeps all five.  The character is taken
      // either way (CUDALucas.cu:3286-3333).
      int keep = 5;
      if (*c) {
        const int k = *c++ - '0';
        if (k >= 0 && k <= 5) keep = k;
      }
      snprintf(f, sizeof f, fixed, v);
      const size_t cut = (size_t)(5 - keep + (keep == 0));
      const size_t len = strlen(f);
      f[len > cut ? len - cut : 0] = 0;
    }
    line += f;
  }
  if (out && cap) {
    const size_t k = line.size() < cap - 1 ? line.size() : cap - 1;
    memcpy(out, line.data(), k);
    out[k] = 0;
  }
  return (int)line.size();
}

// ----------------------------------------------------------- locked files

namespace {
struct HeldLock {
  FILE *file;
  int fd;  // -1: the file was opened without a lock
  std::string lock_path;
};
std::mutex g_lock_mutex;
std::vector<HeldLock> g_held;
}  // namespace

FILE *ll_fopen_locked(const char *path, const char *mode, uint32_t wait_ms) {
  if (!path || !mode) {
    errno = EINVAL;
    return nullptr;
  }
  const std::string lock = std::string(path) + ".lck";
  int fd;
  unsigned sleep_ms = 0;
  uint64_t waited_ms = 0;
  for (;;) {
    fd = open(lock.c_str(), O_CREAT | O_EXCL | O_WRONLY, 0644);
    if (fd >= 0) break;
    if (errno != EEXIST) {
      perror("Cannot open lockfile");
      break;
    }
    if (wait_ms && waited_ms >= wait_ms) {
      errno = ETIMEDOUT;
      return nullptr;
    }
    if (sleep_ms == 0) fprintf(stderr, "%s is locked, waiting ...\n", path);
    if (sleep_ms < 1000) sleep_ms++;
    const struct timespec ts = {(time_t)(sleep_ms / 1000), (long)(sleep_ms % 1000) * 1000000L};
    nanosleep(&ts, nullptr);
    waited_ms += sleep_ms;
  }
  if (fd >= 0 && sleep_ms > 0) printf("Locked %s\n", path);
  FILE *f = fopen(path, mode);
  if (!f) {
    const int err = errno;
    if (fd >= 0) {
      close(fd);
      remove(lock.c_str());
    }
    errno = err;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_lock_mutex);
  g_held.push_back({f, fd, lock});
  return f;
}

int ll_fclose_locked(FILE *f) {
  if (!f) return EOF;
  HeldLock held{nullptr, -1, ""};
  {
    std::lock_guard<std::mutex> g(g_lock_mutex);
    for (size_t i = 0; i < g_held.size(); i++)
      if (g_held[i].file == f) {
        held = g_held[i];
        g_held.erase(g_held.begin() + (long)i);
        break;
      }
  }
  if (!held.file) fprintf(stderr, "File was not locked!\n");
  const int rc = fclose(f);
  if (held.fd >= 0) {
    if (close(held.fd) != 0) perror("Failed to close lockfile");
    if (remove(held.lock_path.c_str()) != 0) perror("Failed to delete lockfile");
  }
  return rc;
}

}  // extern "C"

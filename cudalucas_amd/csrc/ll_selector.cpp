// Learned FFT-length selection: an MLP predicts the maximum roundoff of
// (p, n) from a handful of features, and the selector takes the fastest
// length whose extrapolated full-test roundoff stays under the limit.
//
// The reference picks lengths from fixed power-law bounds
// (choose_fft_length, CUDALucas.cu:1027-1068) and only learns that a length
// was too short from the round-off test (CUDALucas.cu:2589-2648); the
// selector keeps that test as its fallback check (ll_selector_choose with
// `above` set returns the next candidate).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ll_api.h"

namespace {

constexpr uint32_t kMaxLengthK = 16384;  // candidates: multiples of 1K up to 16M words
constexpr double kDomainLow = 0.62;      // candidates from 0.62 n_ref ...
constexpr double kDomainHigh = 1.25;     // ... (training also covers up to 1.25 n_ref)
constexpr double kMaxBitsPerWord = 28.0; // and never past 28 bits per word
constexpr double kLongerLimit = 2.0;     // timed choices may go up to 2 n_ref

// The selector's lengths: multiples of 1K words with a specialised
// schedule (the lengths its roundoff traces cover), to which each exponent's
// reference-policy length is added even where only the generic kernels run
// it.
const std::vector<uint32_t> &supported_lengths() {
  static std::vector<uint32_t> v = [] {
    std::vector<uint32_t> out;
    for (uint32_t k = 1; k <= kMaxLengthK; k++) {
      ll_info info;
      if (ll_fft_schedule(k * 1024u, nullptr, &info) == LL_OK && info.path != 2) out.push_back(k * 1024u);
    }
    return out;
  }();
  return v;
}

// lengths plus nref, ascending
std::vector<uint32_t> with_reference(const std::vector<uint32_t> &lengths, uint32_t nref) {
  std::vector<uint32_t> v = lengths;
  if (!std::binary_search(v.begin(), v.end(), nref)) v.insert(std::lower_bound(v.begin(), v.end(), nref), nref);
  return v;
}

int exponent_of(uint32_t n, uint32_t f) {
  int e = 0;
  while (n % f == 0) {
    n /= f;
    e++;
  }
  return e;
}

}  // namespace

struct ll_selector {
  ll_mlp *mlp = nullptr;
  std::vector<uint32_t> sizes;
  std::vector<float> mean, scale;
  float target_mean = 0.f, target_scale = 1.f;
  float window = 128.f;  // iterations the training target's maximum covers
  float margin = 0.f;    // log2 units added to every prediction
  float limit = 0.35f;   // acceptable full-test maximum roundoff
  float growth_scale = 1.f;  // measured / Gaussian-model growth of the maximum
  std::vector<std::pair<uint32_t, float>> times;  // (n, ms/iter), sorted by n
  char msg[256] = {0};
};

static int sfail(ll_selector *s, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->msg, sizeof s->msg, fmt, ap);
  va_end(ap);
  return code;
}

extern "C" {

int ll_selector_features(uint32_t p, uint32_t n, float *f) {
  ll_info info;
  if (!f || p < 3 || n == 0) return LL_ERR_ARG;
  if (ll_fft_schedule(n, nullptr, &info) != LL_OK) return LL_ERR_UNSUPPORTED;
  f[0] = (float)log2((double)p);
  f[1] = (float)log2((double)n);
  f[2] = (float)((double)p / (double)n);
  f[3] = (float)exponent_of(n, 3);
  f[4] = (float)exponent_of(n, 5);
  f[5] = (float)exponent_of(n, 7);
  f[6] = info.path == 0 ? 1.0f : 0.0f;
  return LL_SELECTOR_FEATURES;
}

int ll_selector_candidates(uint32_t p, uint32_t *ns, uint32_t capacity) {
  int idx = -1;
  const int nref = ll_choose_fft_length(p, &idx);
  if (nref <= 0) return LL_ERR_UNSUPPORTED;
  const double lo = std::max(kDomainLow * nref, p / kMaxBitsPerWord), hi = kDomainHigh * nref;
  uint32_t k = 0;
  for (uint32_t n : with_reference(supported_lengths(), (uint32_t)nref)) {
    if (n < lo || n > hi) continue;
    if (ns) {
      if (k >= capacity) return LL_ERR_ARG;
      ns[k] = n;
    }
    k++;
  }
  return (int)k;
}

int ll_selector_load(int32_t device, const char *path, ll_selector **out) {
  if (!out || !path) return LL_ERR_ARG;
  ll_selector *s = new ll_selector();
  *out = s;
  FILE *f = fopen(path, "r");
  if (!f) return sfail(s, LL_ERR_ARG, "cannot open %s", path);
  char magic[64] = {0}, kind[64] = {0};
  int version = 0;
  std::vector<float> params;
  int rc = LL_OK;
  if (fscanf(f, "%63s %63s %d", magic, kind, &version) != 3 || strcmp(magic, "cudalucas_amd") ||
      strcmp(kind, "roundoff-mlp") || version != 1) {
    rc = sfail(s, LL_ERR_ARG, "%s: not a roundoff-mlp v1 file", path);
  }
  char key[64];
  while (!rc && fscanf(f, "%63s", key) == 1) {
    auto read_floats = [&](std::vector<float> &v, size_t count) {
      v.resize(count);
      for (size_t i = 0; i < count; i++)
        if (fscanf(f, "%f", &v[i]) != 1) return false;
      return true;
    };
    if (!strcmp(key, "sizes")) {
      unsigned cnt = 0;
      if (fscanf(f, "%u", &cnt) != 1 || cnt < 2 || cnt > 16) {
        rc = sfail(s, LL_ERR_ARG, "bad sizes");
        break;
      }
      s->sizes.resize(cnt);
      for (auto &x : s->sizes)
        if (fscanf(f, "%u", &x) != 1) rc = sfail(s, LL_ERR_ARG, "bad sizes");
    } else if (!strcmp(key, "feature_mean") || !strcmp(key, "feature_scale")) {
      if (!read_floats(!strcmp(key, "feature_mean") ? s->mean : s->scale, LL_SELECTOR_FEATURES))
        rc = sfail(s, LL_ERR_ARG, "bad %s", key);
    } else if (!strcmp(key, "target")) {
      if (fscanf(f, "%f %f", &s->target_mean, &s->target_scale) != 2) rc = sfail(s, LL_ERR_ARG, "bad target");
    } else if (!strcmp(key, "window")) {
      if (fscanf(f, "%f", &s->window) != 1 || s->window < 1) rc = sfail(s, LL_ERR_ARG, "bad window");
    } else if (!strcmp(key, "growth_scale")) {
      if (fscanf(f, "%f", &s->growth_scale) != 1 || !(s->growth_scale >= 1.f && s->growth_scale < 4.f))
        rc = sfail(s, LL_ERR_ARG, "bad growth_scale");
    } else if (!strcmp(key, "margin")) {
      if (fscanf(f, "%f", &s->margin) != 1) rc = sfail(s, LL_ERR_ARG, "bad margin");
    } else if (!strcmp(key, "limit")) {
      if (fscanf(f, "%f", &s->limit) != 1 || !(s->limit > 0 && s->limit < 0.5f))
        rc = sfail(s, LL_ERR_ARG, "bad limit");
    } else if (!strcmp(key, "params")) {
      unsigned cnt = 0;
      if (fscanf(f, "%u", &cnt) != 1 || !read_floats(params, cnt)) rc = sfail(s, LL_ERR_ARG, "bad params");
    } else if (key[0] == '#') {
      int c;
      while ((c = fgetc(f)) != EOF && c != '\n') {
      }
    } else {
      rc = sfail(s, LL_ERR_ARG, "unknown key %s", key);
    }
  }
  fclose(f);
  if (rc) return rc;
  if (s->sizes.empty() || s->sizes[0] != LL_SELECTOR_FEATURES || s->mean.size() != LL_SELECTOR_FEATURES ||
      s->scale.size() != LL_SELECTOR_FEATURES)
    return sfail(s, LL_ERR_ARG, "%s: incomplete model", path);
  rc = ll_mlp_create(device, (uint32_t)s->sizes.size() - 1, s->sizes.data(), &s->mlp);
  if (rc) return sfail(s, rc, "mlp: %s", s->mlp ? ll_mlp_last_error(s->mlp) : "");
  if (params.size() != ll_mlp_param_count(s->mlp))
    return sfail(s, LL_ERR_ARG, "%s: %zu params, network has %u", path, params.size(), ll_mlp_param_count(s->mlp));
  rc = ll_mlp_set_params(s->mlp, params.data());
  if (rc) return sfail(s, rc, "mlp: %s", ll_mlp_last_error(s->mlp));
  return LL_OK;
}

void ll_selector_destroy(ll_selector *s) {
  if (!s) return;
  ll_mlp_destroy(s->mlp);
  delete s;
}

const char *ll_selector_last_error(const ll_selector *s) { return s ? s->msg : "null selector"; }

int ll_selector_set_times(ll_selector *s, uint32_t count, const uint32_t *ns, const float *ms) {
  if (!s || (count && (!ns || !ms))) return LL_ERR_ARG;
  s->times.clear();
  for (uint32_t i = 0; i < count; i++)
    if (ms[i] > 0) s->times.emplace_back(ns[i], ms[i]);
  std::sort(s->times.begin(), s->times.end());
  return LL_OK;
}

int ll_selector_predict(ll_selector *s, uint32_t p, uint32_t count, const uint32_t *ns, float *log2_err) {
  if (!s || !s->mlp || !ns || !log2_err || count == 0) return LL_ERR_ARG;
  std::vector<float> x((size_t)count * LL_SELECTOR_FEATURES), y(count);
  for (uint32_t i = 0; i < count; i++) {
    float *f = &x[(size_t)i * LL_SELECTOR_FEATURES];
    if (ll_selector_features(p, ns[i], f) < 0) return sfail(s, LL_ERR_UNSUPPORTED, "length %u", ns[i]);
    for (int k = 0; k < LL_SELECTOR_FEATURES; k++) f[k] = (f[k] - s->mean[k]) / s->scale[k];
  }
  int rc = ll_mlp_forward(s->mlp, x.data(), count, y.data());
  if (rc) return sfail(s, rc, "mlp: %s", ll_mlp_last_error(s->mlp));
  for (uint32_t i = 0; i < count; i++) {
    // maximum of ~Gaussian errors over N*I outputs grows as sqrt(ln(N I)):
    // extrapolate from the training window to the p - 2 iterations of a test
    const double n = ns[i];
    const double growth = 0.5 * log2(log(n * (double)(p - 2)) / log(n * (double)s->window));
    log2_err[i] = y[i] * s->target_scale + s->target_mean + s->growth_scale * (float)std::max(0.0, growth) + s->margin;
  }
  return LL_OK;
}

int ll_selector_choose(ll_selector *s, uint32_t p, uint32_t above, float *pred_err) {
  if (!s || p < 3) return LL_ERR_ARG;
  int idx = -1;
  const int nref = ll_choose_fft_length(p, &idx);
  if (nref <= 0) return sfail(s, LL_ERR_UNSUPPORTED, "no length for p=%u", p);
  // With a timing table every length of the trained domain competes on
  // measured speed, and so do the longer lengths up to twice the reference's
  // (fewer bits per word than anything the reference would use: safe without
  // a prediction); without one, the first choice is the shortest safe length
  // no longer than the reference's.
  const bool timed = !s->times.empty();
  std::vector<uint32_t> cand, longer;
  for (uint32_t n : with_reference(supported_lengths(), (uint32_t)nref)) {
    if (n < kDomainLow * nref || n < p / kMaxBitsPerWord || n <= above) continue;
    if (!above && n > (timed ? kLongerLimit * nref : (double)nref)) break;
    if (n > kDomainHigh * nref && !above) longer.push_back(n);
    else cand.push_back(n);
  }
  if (cand.empty()) return sfail(s, LL_ERR_UNSUPPORTED, "no supported length above %u", above);
  std::vector<float> pred(cand.size());
  int rc = ll_selector_predict(s, p, (uint32_t)cand.size(), cand.data(), pred.data());
  if (rc) return rc;
  const float lim = log2f(s->limit);
  for (uint32_t n : longer) {  // accepted without a prediction
    cand.push_back(n);
    pred.push_back(-INFINITY);
  }
  int best = -1;
  double best_t = 1e30;
  for (size_t i = 0; i < cand.size(); i++) {
    if (pred[i] > lim) continue;
    double t = (double)cand[i];  // without a timing table: the shortest length
    if (!s->times.empty()) {
      auto it = std::lower_bound(s->times.begin(), s->times.end(), std::make_pair(cand[i], 0.0f));
      if (it == s->times.end() || it->first != cand[i]) continue;
      t = it->second;
    }
    if (t < best_t) {
      best_t = t;
      best = (int)i;
    }
  }
  if (best < 0) {
    // nothing predicted safe: the reference's length (or the next one up)
    const uint32_t n = above ? cand.front() : (uint32_t)nref;
    if (pred_err) {
      float e;
      *pred_err = ll_selector_predict(s, p, 1, &n, &e) == LL_OK ? exp2f(e) : -1.f;
    }
    return (int)n;
  }
  if (pred_err) *pred_err = pred[best] == -INFINITY ? 0.f : exp2f(pred[best]);
  return (int)cand[best];
}

}  // extern "C"

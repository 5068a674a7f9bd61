// FP32 multilayer perceptron on gfx950: forward, backward and Adam kernels
// for the learned FFT-length selector (a roundoff-error surrogate).  Hidden
// layers use tanh, the output layer is linear; the loss is the mean squared
// error.  Parameters are one contiguous vector: for each layer l the weights
// W_l [out x in] row-major, then the biases b_l [out].
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "ll_api.h"

namespace {

constexpr int kMaxWidth = 128;  // widest layer the kernels accept

// y[s][o] = act(b[o] + sum_i W[o][i] x[s][i]); one thread per (s, o).
__global__ __launch_bounds__(256) void layer_forward(const float *__restrict__ W, const float *__restrict__ b,
                                                     const float *__restrict__ X, float *__restrict__ Z,
                                                     float *__restrict__ A, int B, int in, int out, int act) {
  __shared__ float w[kMaxWidth * kMaxWidth];
  __shared__ float bb[kMaxWidth];
  for (int i = threadIdx.x; i < in * out; i += blockDim.x) w[i] = W[i];
  for (int i = threadIdx.x; i < out; i += blockDim.x) bb[i] = b[i];
  __syncthreads();
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * out) return;
  const int s = idx / out, o = idx - s * out;
  const float *x = X + (size_t)s * in;
  float z = bb[o];
  for (int i = 0; i < in; i++) z = fmaf(w[o * in + i], x[i], z);
  Z[idx] = z;
  A[idx] = act ? tanhf(z) : z;
}

// dZ = dA * act'(Z) in place (tanh' = 1 - tanh^2).
__global__ void activation_backward(float *__restrict__ dA, const float *__restrict__ A, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dA[i] *= 1.0f - A[i] * A[i];
}

// dW[o][i] = sum_s dZ[s][o] X[s][i], db[o] = sum_s dZ[s][o]: one block per
// output neuron o, threads over the inputs (and one for the bias), each
// summing its column over the batch.
__global__ __launch_bounds__(256) void layer_param_grad(const float *__restrict__ dZ, const float *__restrict__ X,
                                                        float *__restrict__ dW, float *__restrict__ db, int B,
                                                        int in, int out) {
  const int o = blockIdx.x;
  __shared__ float part[256];
  for (int i = 0; i <= in; i++) {  // i == in: the bias
    float acc = 0.f;
    for (int s = threadIdx.x; s < B; s += blockDim.x)
      acc = fmaf(dZ[(size_t)s * out + o], i < in ? X[(size_t)s * in + i] : 1.0f, acc);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (i < in) dW[o * in + i] = part[0];
      else db[o] = part[0];
    }
    __syncthreads();
  }
}

// dX[s][i] = sum_o W[o][i] dZ[s][o]; one thread per (s, i).
__global__ void layer_input_grad(const float *__restrict__ W, const float *__restrict__ dZ, float *__restrict__ dX,
                                 int B, int in, int out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * in) return;
  const int s = idx / in, i = idx - s * in;
  float acc = 0.f;
  for (int o = 0; o < out; o++) acc = fmaf(W[o * in + i], dZ[(size_t)s * out + o], acc);
  dX[idx] = acc;
}

// Mean squared error of a single-output layer: loss += sum (y - t)^2 / B,
// dY = 2 (y - t) / B.
__global__ __launch_bounds__(256) void mse_loss(const float *__restrict__ Y, const float *__restrict__ T,
                                                float *__restrict__ dY, float *__restrict__ loss, int B) {
  __shared__ float part[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float d = 0.f;
  if (i < B) {
    d = Y[i] - T[i];
    dY[i] = 2.0f * d / (float)B;
  }
  part[threadIdx.x] = d * d / (float)B;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// Adam (Kingma & Ba) with bias correction folded into the step size.
__global__ void adam_update(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr_t, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g;
  const float vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

__global__ void zero_f32(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

}  // namespace

struct ll_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<uint32_t> sizes;  // layers + 1
  std::vector<size_t> woff, boff;
  size_t nparam = 0;
  float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
  // activations for a batch: A[l] (B x sizes[l]), Z[l] for l >= 1, dA buffers
  uint32_t cap = 0;
  std::vector<float *> A, Z, dA;
  float *T = nullptr;
  uint64_t adam_t = 0;
  std::vector<void *> allocs;
  char msg[256] = {0};
};

static int mfail(ll_mlp *m, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(m->msg, sizeof m->msg, fmt, ap);
  va_end(ap);
  return code;
}

#define MCHK(m, expr)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return mfail((m), LL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static int malloc_f(ll_mlp *m, float **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, (n + 4) * sizeof(float)) != hipSuccess) return mfail(m, LL_ERR_ALLOC, "hipMalloc mlp");
  m->allocs.push_back(q);
  *p = (float *)q;
  return LL_OK;
}

static int ensure_batch(ll_mlp *m, uint32_t B) {
  if (B <= m->cap) return LL_OK;
  const size_t L = m->sizes.size() - 1;
  m->A.assign(L + 1, nullptr);
  m->Z.assign(L + 1, nullptr);
  m->dA.assign(L + 1, nullptr);
  int rc;
  for (size_t l = 0; l <= L; l++) {
    if ((rc = malloc_f(m, &m->A[l], (size_t)B * m->sizes[l]))) return rc;
    if ((rc = malloc_f(m, &m->dA[l], (size_t)B * m->sizes[l]))) return rc;
    if (l && (rc = malloc_f(m, &m->Z[l], (size_t)B * m->sizes[l]))) return rc;
  }
  if ((rc = malloc_f(m, &m->T, B))) return rc;
  m->cap = B;
  return LL_OK;
}

static int forward(ll_mlp *m, uint32_t B) {
  const size_t L = m->sizes.size() - 1;
  for (size_t l = 1; l <= L; l++) {
    const int in = (int)m->sizes[l - 1], out = (int)m->sizes[l];
    const int total = (int)B * out;
    hipLaunchKernelGGL(layer_forward, dim3((total + 255) / 256), dim3(256), 0, m->stream, m->params + m->woff[l - 1],
                       m->params + m->boff[l - 1], m->A[l - 1], m->Z[l], m->A[l], (int)B, in, out, l < L ? 1 : 0);
  }
  MCHK(m, hipGetLastError());
  return LL_OK;
}

// loss and parameter gradient of the batch currently in A[0] / T
static int backward(ll_mlp *m, uint32_t B) {
  const size_t L = m->sizes.size() - 1;
  hipLaunchKernelGGL(zero_f32, dim3(1), dim3(64), 0, m->stream, m->loss, 1);
  hipLaunchKernelGGL(mse_loss, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->A[L], m->T, m->dA[L], m->loss,
                     (int)B);
  for (size_t l = L; l >= 1; l--) {
    const int in = (int)m->sizes[l - 1], out = (int)m->sizes[l];
    if (l < L)
      hipLaunchKernelGGL(activation_backward, dim3((B * out + 255) / 256), dim3(256), 0, m->stream, m->dA[l], m->A[l],
                         (int)(B * out));
    hipLaunchKernelGGL(layer_param_grad, dim3(out), dim3(256), 0, m->stream, m->dA[l], m->A[l - 1],
                       m->grad + m->woff[l - 1], m->grad + m->boff[l - 1], (int)B, in, out);
    if (l > 1)
      hipLaunchKernelGGL(layer_input_grad, dim3((B * in + 255) / 256), dim3(256), 0, m->stream,
                         m->params + m->woff[l - 1], m->dA[l], m->dA[l - 1], (int)B, in, out);
  }
  MCHK(m, hipGetLastError());
  return LL_OK;
}

extern "C" {

int ll_mlp_create(int32_t device, uint32_t layers, const uint32_t *sizes, ll_mlp **out) {
  if (!out || !sizes || layers == 0) return LL_ERR_ARG;
  for (uint32_t l = 0; l <= layers; l++)
    if (sizes[l] == 0 || sizes[l] > (uint32_t)kMaxWidth) return LL_ERR_ARG;
  if (sizes[layers] != 1) return LL_ERR_ARG;  // the loss assumes one output
  ll_mlp *m = new ll_mlp();
  *out = m;
  m->device = device;
  m->sizes.assign(sizes, sizes + layers + 1);
  size_t off = 0;
  for (uint32_t l = 0; l < layers; l++) {
    m->woff.push_back(off);
    off += (size_t)sizes[l] * sizes[l + 1];
    m->boff.push_back(off);
    off += sizes[l + 1];
  }
  m->nparam = off;
  MCHK(m, hipSetDevice(device));
  MCHK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  int rc;
  if ((rc = malloc_f(m, &m->params, off)) || (rc = malloc_f(m, &m->grad, off)) || (rc = malloc_f(m, &m->m, off)) ||
      (rc = malloc_f(m, &m->v, off)) || (rc = malloc_f(m, &m->loss, 1)))
    return rc;
  // Glorot-uniform initial weights, zero biases (seeded, reproducible)
  std::vector<float> p(off, 0.f);
  std::mt19937_64 rng(12345);
  for (uint32_t l = 0; l < layers; l++) {
    const float a = sqrtf(6.0f / (float)(sizes[l] + sizes[l + 1]));
    std::uniform_real_distribution<float> U(-a, a);
    for (size_t i = 0; i < (size_t)sizes[l] * sizes[l + 1]; i++) p[m->woff[l] + i] = U(rng);
  }
  MCHK(m, hipMemcpy(m->params, p.data(), off * sizeof(float), hipMemcpyHostToDevice));
  MCHK(m, hipMemset(m->m, 0, off * sizeof(float)));
  MCHK(m, hipMemset(m->v, 0, off * sizeof(float)));
  return LL_OK;
}

void ll_mlp_destroy(ll_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *p : m->allocs) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char *ll_mlp_last_error(const ll_mlp *m) { return m ? m->msg : "null mlp"; }

uint32_t ll_mlp_param_count(const ll_mlp *m) { return m ? (uint32_t)m->nparam : 0; }

int ll_mlp_set_params(ll_mlp *m, const float *params) {
  if (!m || !params) return LL_ERR_ARG;
  MCHK(m, hipMemcpy(m->params, params, m->nparam * sizeof(float), hipMemcpyHostToDevice));
  MCHK(m, hipMemset(m->m, 0, m->nparam * sizeof(float)));
  MCHK(m, hipMemset(m->v, 0, m->nparam * sizeof(float)));
  m->adam_t = 0;
  return LL_OK;
}

int ll_mlp_get_params(ll_mlp *m, float *params) {
  if (!m || !params) return LL_ERR_ARG;
  MCHK(m, hipStreamSynchronize(m->stream));
  MCHK(m, hipMemcpy(params, m->params, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return LL_OK;
}

int ll_mlp_forward(ll_mlp *m, const float *x, uint32_t batch, float *y) {
  if (!m || !x || !y || batch == 0) return LL_ERR_ARG;
  int rc = ensure_batch(m, batch);
  if (rc) return rc;
  MCHK(m, hipMemcpyAsync(m->A[0], x, (size_t)batch * m->sizes[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if ((rc = forward(m, batch))) return rc;
  MCHK(m, hipMemcpyAsync(y, m->A.back(), batch * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return LL_OK;
}

int ll_mlp_gradient(ll_mlp *m, const float *x, const float *t, uint32_t batch, float *grad, float *loss) {
  if (!m || !x || !t || batch == 0) return LL_ERR_ARG;
  int rc = ensure_batch(m, batch);
  if (rc) return rc;
  MCHK(m, hipMemcpyAsync(m->A[0], x, (size_t)batch * m->sizes[0] * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->T, t, batch * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if ((rc = forward(m, batch)) || (rc = backward(m, batch))) return rc;
  if (grad) MCHK(m, hipMemcpyAsync(grad, m->grad, m->nparam * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss) MCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return LL_OK;
}

int ll_mlp_train(ll_mlp *m, const float *x, const float *t, uint32_t count, uint32_t batch, uint32_t epochs,
                 float lr, uint64_t seed, float *epoch_loss) {
  if (!m || !x || !t || count == 0 || batch == 0 || batch > count) return LL_ERR_ARG;
  const uint32_t in = m->sizes[0];
  int rc = ensure_batch(m, batch);
  if (rc) return rc;
  float *dX = nullptr, *dT = nullptr;
  if ((rc = malloc_f(m, &dX, (size_t)count * in)) || (rc = malloc_f(m, &dT, count))) return rc;
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::mt19937_64 rng(seed);
  std::vector<float> xs((size_t)count * in), ts(count);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t ep = 0; ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), rng);
    for (uint32_t i = 0; i < count; i++) {
      memcpy(&xs[(size_t)i * in], &x[(size_t)order[i] * in], in * sizeof(float));
      ts[i] = t[order[i]];
    }
    MCHK(m, hipMemcpyAsync(dX, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(dT, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    double sum = 0.0;
    uint32_t nb = 0;
    for (uint32_t s = 0; s + batch <= count; s += batch) {
      MCHK(m, hipMemcpyAsync(m->A[0], dX + (size_t)s * in, (size_t)batch * in * sizeof(float),
                             hipMemcpyDeviceToDevice, m->stream));
      MCHK(m, hipMemcpyAsync(m->T, dT + s, batch * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
      if ((rc = forward(m, batch)) || (rc = backward(m, batch))) return rc;
      m->adam_t++;
      const float lr_t = lr * sqrtf(1.0f - powf(b2, (float)m->adam_t)) / (1.0f - powf(b1, (float)m->adam_t));
      hipLaunchKernelGGL(adam_update, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0, m->stream, m->params,
                         m->grad, m->m, m->v, (int)m->nparam, lr_t, b1, b2, eps);
      float l = 0.f;
      MCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
      MCHK(m, hipStreamSynchronize(m->stream));
      sum += l;
      nb++;
    }
    if (epoch_loss) epoch_loss[ep] = (float)(sum / std::max(nb, 1u));
  }
  return LL_OK;
}

}  // extern "C"

// mlp.hip — the tile-cost model: a 3-layer perceptron (in -> H -> H -> 1,
// ReLU, FP32) trained on the GPU with the library's own forward, backward
// and Adam kernels.  It predicts log(kernel time) of a BLAS call from its
// shape and configuration; the tuner (xkblas_amd/tuner.py) uses it to pick
// the tile shape and split count.
//
// Layout: parameters in one FP32 vector [W1 (in x H, row-major by input),
// b1 (H), W2 (H x H), b2 (H), w3 (H), b3 (1)].  Samples are row-major
// (n x in).
//
// Kernels:
//   mlp_forward_kernel   one thread per sample, activations in registers;
//   mlp_backward_kernel  a workgroup handles 128 samples: forward again,
//                        then the per-sample deltas go to LDS and the weight
//                        gradients are reduced over the 128 samples inside
//                        the workgroup (each thread owns 16 entries of W2's
//                        gradient) before one atomicAdd per entry;
//   mlp_adam_kernel      Adam step on the parameter vector.
// The loss is the mean squared error; gradients are exact (no sampling).
#include "common.h"
#include "runtime.h"
#include "../../include/xkblas.h"
#include <cmath>
#include <cstring>
#include <vector>

namespace xk {

constexpr int MLP_MAX_IN = 16;
constexpr int MLP_H = 64;

struct MlpShape {
  int in;
  long oW1, ob1, oW2, ob2, ow3, ob3, total;
};

static MlpShape mlp_shape(int in)
{
  MlpShape s;
  s.in = in;
  s.oW1 = 0;
  s.ob1 = s.oW1 + (long)in * MLP_H;
  s.oW2 = s.ob1 + MLP_H;
  s.ob2 = s.oW2 + (long)MLP_H * MLP_H;
  s.ow3 = s.ob2 + MLP_H;
  s.ob3 = s.ow3 + MLP_H;
  s.total = s.ob3 + 1;
  return s;
}

__device__ __forceinline__ void mlp_fwd_sample(const MlpShape& s, const float* __restrict__ P, const float* x,
                                               float* h1, float* h2, float& y)
{
#pragma unroll 4
  for (int j = 0; j < MLP_H; ++j) {
    float z = P[s.ob1 + j];
    for (int i = 0; i < s.in; ++i) z = fmaf(x[i], P[s.oW1 + i * MLP_H + j], z);
    h1[j] = z > 0.0f ? z : 0.0f;
  }
  for (int j = 0; j < MLP_H; ++j) {
    float z = P[s.ob2 + j];
#pragma unroll 8
    for (int i = 0; i < MLP_H; ++i) z = fmaf(h1[i], P[s.oW2 + i * MLP_H + j], z);
    h2[j] = z > 0.0f ? z : 0.0f;
  }
  float o = P[s.ob3];
#pragma unroll 8
  for (int i = 0; i < MLP_H; ++i) o = fmaf(h2[i], P[s.ow3 + i], o);
  y = o;
}

__global__ __launch_bounds__(128) void mlp_forward_kernel(MlpShape s, const float* __restrict__ P,
                                                          const float* __restrict__ X, int n, float* __restrict__ Y)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float x[MLP_MAX_IN], h1[MLP_H], h2[MLP_H], y;
  for (int i = 0; i < s.in; ++i) x[i] = X[(long)e * s.in + i];
  mlp_fwd_sample(s, P, x, h1, h2, y);
  Y[e] = y;
}

// Gradient of L = (1/n) sum (y - t)^2 accumulated into G; loss sum into *L.
__global__ __launch_bounds__(128) void mlp_backward_kernel(MlpShape s, const float* __restrict__ P,
                                                           const float* __restrict__ X, const float* __restrict__ T,
                                                           int n, float inv_n, float* __restrict__ G,
                                                           float* __restrict__ loss)
{
  __shared__ float sh1[128][MLP_H + 1];   // h1 per sample
  __shared__ float sd2[128][MLP_H + 1];   // delta at layer-2 pre-activation
  __shared__ float sd1[128][MLP_H + 1];   // delta at layer-1 pre-activation
  __shared__ float sx[128][MLP_MAX_IN];
  const int t = threadIdx.x;
  const int e = blockIdx.x * 128 + t;
  const bool valid = e < n;
  float x[MLP_MAX_IN], h1[MLP_H], h2[MLP_H], y = 0.0f;
  for (int i = 0; i < MLP_MAX_IN; ++i) x[i] = (valid && i < s.in) ? X[(long)e * s.in + i] : 0.0f;
  float dy = 0.0f;
  if (valid) {
    mlp_fwd_sample(s, P, x, h1, h2, y);
    const float r = y - T[e];
    dy = 2.0f * r * inv_n;
    atomicAdd(loss, r * r * inv_n);
  } else {
    for (int j = 0; j < MLP_H; ++j) h1[j] = h2[j] = 0.0f;
  }
  // layer 3
  float d2[MLP_H];
  for (int j = 0; j < MLP_H; ++j) d2[j] = (h2[j] > 0.0f) ? dy * P[s.ow3 + j] : 0.0f;
  // layer 2 -> layer 1 delta
  for (int i = 0; i < MLP_H; ++i) {
    float a = 0.0f;
#pragma unroll 8
    for (int j = 0; j < MLP_H; ++j) a = fmaf(P[s.oW2 + i * MLP_H + j], d2[j], a);
    sd1[t][i] = (h1[i] > 0.0f) ? a : 0.0f;
    sh1[t][i] = h1[i];
    sd2[t][i] = d2[i];
  }
  for (int i = 0; i < MLP_MAX_IN; ++i) sx[t][i] = x[i];
  // gradients of w3, b3 need h2 and dy: reduce through LDS reuse of sd1? do
  // them directly with one atomic per entry per workgroup below
  __shared__ float sdy[128];
  __shared__ float sh2[128][MLP_H + 1];
  sdy[t] = dy;
  for (int j = 0; j < MLP_H; ++j) sh2[t][j] = h2[j];
  __syncthreads();
  // W2 gradient: G_W2[i][j] = sum_s h1[s][i] d2[s][j]; 4096 entries, 32 per thread
  for (int q = t; q < MLP_H * MLP_H; q += 128) {
    const int i = q / MLP_H, j = q % MLP_H;
    float a = 0.0f;
    for (int r = 0; r < 128; ++r) a = fmaf(sh1[r][i], sd2[r][j], a);
    atomicAdd(&G[s.oW2 + q], a);
  }
  // W1 gradient: G_W1[i][j] = sum_s x[s][i] d1[s][j]
  for (int q = t; q < s.in * MLP_H; q += 128) {
    const int i = q / MLP_H, j = q % MLP_H;
    float a = 0.0f;
    for (int r = 0; r < 128; ++r) a = fmaf(sx[r][i], sd1[r][j], a);
    atomicAdd(&G[s.oW1 + q], a);
  }
  if (t < MLP_H) {
    float a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (int r = 0; r < 128; ++r) {
      a1 += sd1[r][t];
      a2 += sd2[r][t];
      a3 = fmaf(sh2[r][t], sdy[r], a3);
    }
    atomicAdd(&G[s.ob1 + t], a1);
    atomicAdd(&G[s.ob2 + t], a2);
    atomicAdd(&G[s.ow3 + t], a3);
  }
  if (t == 0) {
    float a = 0.0f;
    for (int r = 0; r < 128; ++r) a += sdy[r];
    atomicAdd(&G[s.ob3], a);
  }
}

__global__ void mlp_adam_kernel(float* P, const float* G, float* M, float* V, long n, float lr, float b1, float b2,
                                float eps, float c1, float c2)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.0f - b1) * g;
  const float v = b2 * V[i] + (1.0f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void mlp_zero_kernel(float* p, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

struct Mlp {
  MlpShape s;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *loss = nullptr;
  float *X = nullptr, *T = nullptr, *Y = nullptr;
  long cap = 0;  // samples the X/T/Y buffers hold
  long step = 0;
  hipStream_t st = nullptr;
};

static void mlp_reserve(Mlp* m, long n)
{
  if (n <= m->cap) return;
  if (m->X) {
    XK_HIP(hipFree(m->X));
    XK_HIP(hipFree(m->T));
    XK_HIP(hipFree(m->Y));
  }
  XK_HIP(hipMalloc(&m->X, sizeof(float) * n * m->s.in));
  XK_HIP(hipMalloc(&m->T, sizeof(float) * n));
  XK_HIP(hipMalloc(&m->Y, sizeof(float) * n));
  m->cap = n;
}

static void mlp_grad_dev(Mlp* m, const float* X, const float* T, int n)
{
  const long P = m->s.total;
  hipLaunchKernelGGL(mlp_zero_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, m->st, m->G, P);
  hipLaunchKernelGGL(mlp_zero_kernel, dim3(1), dim3(256), 0, m->st, m->loss, 1L);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((n + 127) / 128), dim3(128), 0, m->st, m->s, m->P, X, T, n,
                     1.0f / (float)n, m->G, m->loss);
  XK_HIP(hipGetLastError());
}

}  // namespace xk

using namespace xk;

extern "C" {

void* xkblas_mlp_create(int in_features, unsigned seed)
{
  if (in_features < 1 || in_features > MLP_MAX_IN) return nullptr;
  xkblas_init();
  Mlp* m = new Mlp();
  m->s = mlp_shape(in_features);
  const long P = m->s.total;
  XK_HIP(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
  XK_HIP(hipMalloc(&m->P, sizeof(float) * P));
  XK_HIP(hipMalloc(&m->G, sizeof(float) * P));
  XK_HIP(hipMalloc(&m->M, sizeof(float) * P));
  XK_HIP(hipMalloc(&m->V, sizeof(float) * P));
  XK_HIP(hipMalloc(&m->loss, sizeof(float)));
  // He-uniform initialisation from a splitmix64 stream
  std::vector<float> h(P, 0.0f);
  unsigned long long z = seed * 0x9E3779B97F4A7C15ull + 1;
  auto rnd = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    unsigned long long x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) / 9007199254740992.0;
  };
  const double a1 = std::sqrt(6.0 / in_features), a2 = std::sqrt(6.0 / MLP_H);
  for (long i = m->s.oW1; i < m->s.ob1; ++i) h[i] = (float)((2 * rnd() - 1) * a1);
  for (long i = m->s.oW2; i < m->s.ob2; ++i) h[i] = (float)((2 * rnd() - 1) * a2);
  for (long i = m->s.ow3; i < m->s.ob3; ++i) h[i] = (float)((2 * rnd() - 1) * a2);
  XK_HIP(hipMemcpy(m->P, h.data(), sizeof(float) * P, hipMemcpyHostToDevice));
  std::vector<float> zeros(P, 0.0f);
  XK_HIP(hipMemcpy(m->M, zeros.data(), sizeof(float) * P, hipMemcpyHostToDevice));
  XK_HIP(hipMemcpy(m->V, zeros.data(), sizeof(float) * P, hipMemcpyHostToDevice));
  return m;
}

long xkblas_mlp_num_params(void* h) { return h ? static_cast<Mlp*>(h)->s.total : -1; }

int xkblas_mlp_set_params(void* h, const float* params)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !params) return -1;
  XK_HIP(hipMemcpy(m->P, params, sizeof(float) * m->s.total, hipMemcpyHostToDevice));
  return 0;
}

int xkblas_mlp_get_params(void* h, float* params)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || !params) return -1;
  XK_HIP(hipMemcpy(params, m->P, sizeof(float) * m->s.total, hipMemcpyDeviceToHost));
  return 0;
}

int xkblas_mlp_forward(void* h, const float* X, int n, float* Y)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || n < 0) return -1;
  if (n == 0) return 0;
  mlp_reserve(m, n);
  XK_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * (size_t)n * m->s.in, hipMemcpyHostToDevice, m->st));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 127) / 128), dim3(128), 0, m->st, m->s, m->P, m->X, n, m->Y);
  XK_HIP(hipGetLastError());
  XK_HIP(hipMemcpyAsync(Y, m->Y, sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
  XK_HIP(hipStreamSynchronize(m->st));
  return 0;
}

// Loss and full-batch gradient (for checks against the CPU restatement).
int xkblas_mlp_gradient(void* h, const float* X, const float* T, int n, float* grad, float* loss)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || n <= 0) return -1;
  mlp_reserve(m, n);
  XK_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * (size_t)n * m->s.in, hipMemcpyHostToDevice, m->st));
  XK_HIP(hipMemcpyAsync(m->T, T, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
  mlp_grad_dev(m, m->X, m->T, n);
  XK_HIP(hipMemcpyAsync(grad, m->G, sizeof(float) * m->s.total, hipMemcpyDeviceToHost, m->st));
  XK_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
  XK_HIP(hipStreamSynchronize(m->st));
  return 0;
}

// Adam on mini-batches of `batch` consecutive samples (the caller shuffles);
// returns the mean loss of the last epoch.
float xkblas_mlp_train(void* h, const float* X, const float* T, int n, int epochs, int batch, float lr)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m || n <= 0 || epochs <= 0) return -1.0f;
  if (batch <= 0 || batch > n) batch = n;
  mlp_reserve(m, n);
  XK_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * (size_t)n * m->s.in, hipMemcpyHostToDevice, m->st));
  XK_HIP(hipMemcpyAsync(m->T, T, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
  const long P = m->s.total;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float last = 0.0f;
  for (int ep = 0; ep < epochs; ++ep) {
    double sum = 0.0;
    int nb = 0;
    for (int off = 0; off < n; off += batch) {
      const int cnt = std::min(batch, n - off);
      mlp_grad_dev(m, m->X + (long)off * m->s.in, m->T + off, cnt);
      ++m->step;
      const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, m->st, m->P, m->G, m->M,
                         m->V, P, lr, b1, b2, eps, c1, c2);
      if (ep == epochs - 1) {
        float l = 0.0f;
        XK_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
        XK_HIP(hipStreamSynchronize(m->st));
        sum += l;
        ++nb;
      }
    }
    if (ep == epochs - 1) last = (float)(sum / std::max(nb, 1));
  }
  XK_HIP(hipStreamSynchronize(m->st));
  return last;
}

void xkblas_mlp_destroy(void* h)
{
  Mlp* m = static_cast<Mlp*>(h);
  if (!m) return;
  (void)hipStreamSynchronize(m->st);
  (void)hipFree(m->P);
  (void)hipFree(m->G);
  (void)hipFree(m->M);
  (void)hipFree(m->V);
  (void)hipFree(m->loss);
  if (m->X) {
    (void)hipFree(m->X);
    (void)hipFree(m->T);
    (void)hipFree(m->Y);
  }
  (void)hipStreamDestroy(m->st);
  delete m;
}

}  // extern "C"

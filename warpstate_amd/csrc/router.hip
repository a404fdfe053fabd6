// Learned kernel router: a small FP32 MLP cost model with hand-written HIP
// training and inference kernels, and the exhaustive autotuner it is trained
// from (it stands in for AutoMatcher's size threshold, core/src/router.rs:205-225,
// and for an exhaustive tuning sweep on every request).
//
// Network: 8 inputs (7 request features + 1 zero pad) -> 64 -> 64 -> 64 -> 16
// outputs (predicted standardized log time of each kernel configuration),
// ReLU between layers, masked mean-squared-error loss, Adam.
//
// Kernels:
//   k_mlp_forward   one wave per sample, lane j computes neuron j of every
//                   layer; weights sit transposed in LDS so that lanes read
//                   consecutive words (no bank conflicts).
//   k_mlp_delta     per-sample backward pass of the layer errors.
//   k_mlp_wgrad     weight gradients dW[j][k] = sum_b D[b][j] * A[b][k] and
//                   bias gradients, one lane per weight, reduced over the batch.
//   k_adam          parameter update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/warpstate.h"

namespace ws {

constexpr int kIn = 8, kHid = 64, kOut = WS_ROUTER_CONFIGS;
static_assert(kOut <= 64, "one lane per output");
// Parameter layout (row-major W[out][in]).
constexpr int kW1 = 0, kB1 = kW1 + kHid * kIn;
constexpr int kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid;
constexpr int kW3 = kB2 + kHid, kB3 = kW3 + kHid * kHid;
constexpr int kW4 = kB3 + kHid, kB4 = kW4 + kOut * kHid;
constexpr int kParams = kB4 + kOut;

struct MlpBuffers {
  const float* x;    // [B][kIn]
  const float* y;    // [B][kOut]
  const float* m;    // [B][kOut] mask (1 = valid target)
  float* h1;         // [B][kHid] post-ReLU activations
  float* h2;
  float* h3;
  float* out;        // [B][kOut]
  float* d1;         // [B][kHid] layer errors
  float* d2;
  float* d3;
  float* d4;         // [B][kOut]
  uint32_t n;
};

// Stages all parameters into LDS; weight matrices transposed (in-major).
__device__ void stage_params(const float* __restrict__ p, float* s) {
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
    float v = p[i];
    int dst = i;
    if (i < kB1) { int o = i / kIn, k = i % kIn; dst = kW1 + k * kHid + o; }
    else if (i >= kW2 && i < kB2) { int r = i - kW2, o = r / kHid, k = r % kHid; dst = kW2 + k * kHid + o; }
    else if (i >= kW3 && i < kB3) { int r = i - kW3, o = r / kHid, k = r % kHid; dst = kW3 + k * kHid + o; }
    else if (i >= kW4 && i < kB4) { int r = i - kW4, o = r / kHid, k = r % kHid; dst = kW4 + k * kOut + o; }
    s[dst] = v;
  }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params, MlpBuffers b) {
  __shared__ float sp[kParams];
  __shared__ float act[4][kHid];
  stage_params(params, sp);
  __syncthreads();
  const int wave = threadIdx.x >> 6, j = threadIdx.x & 63;
  for (uint32_t s = blockIdx.x * 4 + wave; s < b.n; s += gridDim.x * 4) {
    const float* x = b.x + (size_t)s * kIn;
    float a = sp[kB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a += sp[kW1 + k * kHid + j] * x[k];
    a = fmaxf(a, 0.f);
    if (b.h1) b.h1[(size_t)s * kHid + j] = a;
    act[wave][j] = a;
    __builtin_amdgcn_wave_barrier();
    float a2 = sp[kB2 + j];
    for (int k = 0; k < kHid; ++k) a2 += sp[kW2 + k * kHid + j] * act[wave][k];
    a2 = fmaxf(a2, 0.f);
    if (b.h2) b.h2[(size_t)s * kHid + j] = a2;
    __builtin_amdgcn_wave_barrier();
    act[wave][j] = a2;
    __builtin_amdgcn_wave_barrier();
    float a3 = sp[kB3 + j];
    for (int k = 0; k < kHid; ++k) a3 += sp[kW3 + k * kHid + j] * act[wave][k];
    a3 = fmaxf(a3, 0.f);
    if (b.h3) b.h3[(size_t)s * kHid + j] = a3;
    __builtin_amdgcn_wave_barrier();
    act[wave][j] = a3;
    __builtin_amdgcn_wave_barrier();
    if (j < kOut) {
      float o = sp[kB4 + j];
      for (int k = 0; k < kHid; ++k) o += sp[kW4 + k * kOut + j] * act[wave][k];
      b.out[(size_t)s * kOut + j] = o;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Layer errors of the masked MSE loss L = scale * sum m * (out - y)^2.
__global__ __launch_bounds__(256) void k_mlp_delta(const float* __restrict__ params, MlpBuffers b, float scale) {
  __shared__ float sp[kParams];
  __shared__ float err[4][kHid];
  stage_params(params, sp);
  __syncthreads();
  const int wave = threadIdx.x >> 6, j = threadIdx.x & 63;
  for (uint32_t s = blockIdx.x * 4 + wave; s < b.n; s += gridDim.x * 4) {
    if (j < kOut) {
      const size_t q = (size_t)s * kOut + j;
      const float d = 2.f * scale * b.m[q] * (b.out[q] - b.y[q]);
      b.d4[q] = d;
      err[wave][j] = d;
    }
    __builtin_amdgcn_wave_barrier();
    // d3_j = relu'(h3_j) * sum_o W4[o][j] d4_o; W4 staged as W4T[j][o].
    float g = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) g += sp[kW4 + j * kOut + o] * err[wave][o];
    const float d3 = b.h3[(size_t)s * kHid + j] > 0.f ? g : 0.f;
    b.d3[(size_t)s * kHid + j] = d3;
    __builtin_amdgcn_wave_barrier();
    err[wave][j] = d3;
    __builtin_amdgcn_wave_barrier();
    // d2_j = relu'(h2_j) * sum_o W3[o][j] d3_o; W3 staged as W3T[j][o].
    g = 0.f;
    for (int o = 0; o < kHid; ++o) g += sp[kW3 + j * kHid + o] * err[wave][o];
    const float d2 = b.h2[(size_t)s * kHid + j] > 0.f ? g : 0.f;
    b.d2[(size_t)s * kHid + j] = d2;
    __builtin_amdgcn_wave_barrier();
    err[wave][j] = d2;
    __builtin_amdgcn_wave_barrier();
    g = 0.f;
    for (int o = 0; o < kHid; ++o) g += sp[kW2 + j * kHid + o] * err[wave][o];
    const float d1 = b.h1[(size_t)s * kHid + j] > 0.f ? g : 0.f;
    b.d1[(size_t)s * kHid + j] = d1;
    __builtin_amdgcn_wave_barrier();
  }
}

// dW[o][k] = sum_s D[s][o] * A[s][k]; db[o] = sum_s D[s][o].
// One lane per (o, k) plus `rows` lanes for the biases.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ D, const float* __restrict__ A,
                                                    uint32_t n, int rows, int cols, int dstride, int astride,
                                                    float* __restrict__ dW, float* __restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < rows * cols) {
    const int o = t / cols, k = t % cols;
    float acc = 0.f;
    for (uint32_t s = 0; s < n; ++s) acc += D[(size_t)s * dstride + o] * A[(size_t)s * astride + k];
    dW[t] = acc;
  } else if (t < rows * cols + rows) {
    const int o = t - rows * cols;
    float acc = 0.f;
    for (uint32_t s = 0; s < n; ++s) acc += D[(size_t)s * dstride + o];
    db[o] = acc;
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace ws

using namespace ws;

struct ws_router {
  int dev = 0;
  hipStream_t st = nullptr;
  std::mutex mu;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  float *x = nullptr, *y = nullptr, *mask = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *out = nullptr,
        *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *d4 = nullptr;
  // Pinned host staging for small inference batches: the forward kernel reads
  // the features and writes the predictions in host memory directly, so a
  // request-path prediction is one launch and one synchronization.
  static constexpr uint32_t kPinnedRows = 256;
  float *hx = nullptr, *hout = nullptr;
  uint32_t cap = 0;
  uint32_t step = 0;
  float xmean[kIn] = {}, xstd[kIn] = {1, 1, 1, 1, 1, 1, 1, 1};
  float ymean = 0, ystd = 1;
};

namespace {
thread_local std::string r_err;
ws_status rfail(ws_status s, const std::string& m) {
  r_err = m;
  return s;
}
#define R_HIP(expr)                                                                       \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) return rfail(WS_ERR_GPU_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

ws_status ensure_batch(ws_router* r, uint32_t n) {
  if (n <= r->cap) return WS_OK;
  float** bufs[] = {&r->x, &r->y, &r->mask, &r->h1, &r->h2, &r->h3, &r->out, &r->d1, &r->d2, &r->d3, &r->d4};
  const size_t widths[] = {kIn, kOut, kOut, kHid, kHid, kHid, kOut, kHid, kHid, kHid, kOut};
  for (int i = 0; i < 11; ++i) {
    if (*bufs[i]) (void)hipFree(*bufs[i]);
    *bufs[i] = nullptr;
    R_HIP(hipMalloc(reinterpret_cast<void**>(bufs[i]), (size_t)n * widths[i] * sizeof(float)));
  }
  r->cap = n;
  return WS_OK;
}

MlpBuffers make_buffers(ws_router* r, uint32_t n) {
  MlpBuffers b{};
  b.x = r->x; b.y = r->y; b.m = r->mask;
  b.h1 = r->h1; b.h2 = r->h2; b.h3 = r->h3; b.out = r->out;
  b.d1 = r->d1; b.d2 = r->d2; b.d3 = r->d3; b.d4 = r->d4;
  b.n = n;
  return b;
}

uint32_t blocks_for(uint32_t n) { return std::max(1u, std::min(1024u, (n + 3) / 4)); }

// Full-batch gradient into r->g for the (already normalized, uploaded) batch.
ws_status gradient(ws_router* r, uint32_t n, float scale) {
  MlpBuffers b = make_buffers(r, n);
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n)), dim3(256), 0, r->st, r->p, b);
  R_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_delta, dim3(blocks_for(n)), dim3(256), 0, r->st, r->p, b, scale);
  R_HIP(hipGetLastError());
  struct L { int w, bb, rows, cols, dstride, astride; const float* D; const float* A; };
  const L layers[] = {
      {kW1, kB1, kHid, kIn, kHid, kIn, r->d1, r->x},
      {kW2, kB2, kHid, kHid, kHid, kHid, r->d2, r->h1},
      {kW3, kB3, kHid, kHid, kHid, kHid, r->d3, r->h2},
      {kW4, kB4, kOut, kHid, kOut, kHid, r->d4, r->h3},
  };
  for (const L& l : layers) {
    const int threads = l.rows * l.cols + l.rows;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((threads + 255) / 256), dim3(256), 0, r->st, l.D, l.A, n, l.rows, l.cols,
                       l.dstride, l.astride, r->g + l.w, r->g + l.bb);
    R_HIP(hipGetLastError());
  }
  return WS_OK;
}

// Standardizes features with the stored statistics (pad column stays 0).
void normalize(const ws_router* r, const float* x, uint32_t n, std::vector<float>& out) {
  out.resize((size_t)n * kIn);
  for (uint32_t s = 0; s < n; ++s)
    for (int k = 0; k < kIn; ++k)
      out[(size_t)s * kIn + k] = k < WS_ROUTER_FEATURES ? (x[(size_t)s * kIn + k] - r->xmean[k]) / r->xstd[k] : 0.f;
}

}  // namespace

extern "C" {

const char* ws_router_last_error(void) { return r_err.c_str(); }

uint32_t ws_router_param_count(void) { return (uint32_t)kParams; }

ws_status ws_router_create(int device, uint32_t seed, ws_router** out) {
  if (!out) return rfail(WS_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return rfail(WS_ERR_NO_GPU_ADAPTER, "no GPU adapter available");
  }
  ws_router* r = new ws_router();
  r->dev = device;
  auto bail = [&](hipError_t e) {
    std::string msg = hipGetErrorString(e);
    ws_router_destroy(r);
    return rfail(WS_ERR_GPU_DEVICE, msg);
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bail(e);
  if ((e = hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking)) != hipSuccess) return bail(e);
  for (float** q : {&r->p, &r->g, &r->m, &r->v})
    if ((e = hipMalloc(reinterpret_cast<void**>(q), kParams * sizeof(float))) != hipSuccess) return bail(e);
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&r->hx), ws_router::kPinnedRows * kIn * sizeof(float),
                         hipHostMallocDefault)) != hipSuccess)
    return bail(e);
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&r->hout), ws_router::kPinnedRows * kOut * sizeof(float),
                         hipHostMallocDefault)) != hipSuccess)
    return bail(e);
  // He-uniform initialization, zero biases.
  std::vector<float> init(kParams, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int rows, int cols) {
    const float lim = std::sqrt(6.f / (float)cols);
    std::uniform_real_distribution<float> u(-lim, lim);
    for (int i = 0; i < rows * cols; ++i) init[off + i] = u(rng);
  };
  fill(kW1, kHid, kIn);
  fill(kW2, kHid, kHid);
  fill(kW3, kHid, kHid);
  fill(kW4, kOut, kHid);
  if ((e = hipMemcpy(r->p, init.data(), kParams * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
  std::vector<float> zero(kParams, 0.f);
  if ((e = hipMemcpy(r->m, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
  if ((e = hipMemcpy(r->v, zero.data(), kParams * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) return bail(e);
  *out = r;
  return WS_OK;
}

void ws_router_destroy(ws_router* r) {
  if (!r) return;
  (void)hipSetDevice(r->dev);
  if (r->st) (void)hipStreamSynchronize(r->st);
  for (float* q : {r->p, r->g, r->m, r->v, r->x, r->y, r->mask, r->h1, r->h2, r->h3, r->out, r->d1, r->d2, r->d3,
                   r->d4})
    if (q) (void)hipFree(q);
  if (r->hx) (void)hipHostFree(r->hx);
  if (r->hout) (void)hipHostFree(r->hout);
  if (r->st) (void)hipStreamDestroy(r->st);
  delete r;
}

ws_status ws_router_get_params(ws_router* r, float* out, uint32_t n) {
  if (!r || !out || n != (uint32_t)kParams) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(r->mu);
  R_HIP(hipSetDevice(r->dev));
  R_HIP(hipMemcpy(out, r->p, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return WS_OK;
}

ws_status ws_router_set_params(ws_router* r, const float* in, uint32_t n) {
  if (!r || !in || n != (uint32_t)kParams) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(r->mu);
  R_HIP(hipSetDevice(r->dev));
  R_HIP(hipMemcpy(r->p, in, kParams * sizeof(float), hipMemcpyHostToDevice));
  return WS_OK;
}

ws_status ws_router_set_normalization(ws_router* r, const float* xmean, const float* xstd, float ymean, float ystd) {
  if (!r || !xmean || !xstd || !(ystd > 0)) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int k = 0; k < kIn; ++k) {
    r->xmean[k] = xmean[k];
    r->xstd[k] = xstd[k] > 1e-6f ? xstd[k] : 1.f;
  }
  r->ymean = ymean;
  r->ystd = ystd;
  return WS_OK;
}

ws_status ws_router_get_normalization(ws_router* r, float* xmean, float* xstd, float* ymean, float* ystd) {
  if (!r || !xmean || !xstd || !ymean || !ystd) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int k = 0; k < kIn; ++k) { xmean[k] = r->xmean[k]; xstd[k] = r->xstd[k]; }
  *ymean = r->ymean;
  *ystd = r->ystd;
  return WS_OK;
}

// Gradient of scale * sum m * (f(x) - y)^2 with respect to the parameters,
// on raw (already standardized) inputs: the check against the CPU oracle.
ws_status ws_router_gradient(ws_router* r, const float* x, const float* y, const float* mask, uint32_t n,
                             float scale, float* grad_out) {
  if (!r || !x || !y || !mask || !grad_out || n == 0) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(r->mu);
  R_HIP(hipSetDevice(r->dev));
  ws_status s = ensure_batch(r, n);
  if (s != WS_OK) return s;
  R_HIP(hipMemcpyAsync(r->x, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice, r->st));
  R_HIP(hipMemcpyAsync(r->y, y, (size_t)n * kOut * 4, hipMemcpyHostToDevice, r->st));
  R_HIP(hipMemcpyAsync(r->mask, mask, (size_t)n * kOut * 4, hipMemcpyHostToDevice, r->st));
  s = gradient(r, n, scale);
  if (s != WS_OK) return s;
  R_HIP(hipMemcpyAsync(grad_out, r->g, kParams * 4, hipMemcpyDeviceToHost, r->st));
  R_HIP(hipStreamSynchronize(r->st));
  return WS_OK;
}

// Raw network outputs for standardized inputs x[n][8].
ws_status ws_router_forward(ws_router* r, const float* x, uint32_t n, float* out) {
  if (!r || !x || !out || n == 0) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(r->mu);
  R_HIP(hipSetDevice(r->dev));
  if (n <= ws_router::kPinnedRows) {
    std::memcpy(r->hx, x, (size_t)n * kIn * 4);
    MlpBuffers b{};
    b.x = r->hx;
    b.out = r->hout;
    b.n = n;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n)), dim3(256), 0, r->st, r->p, b);
    R_HIP(hipGetLastError());
    R_HIP(hipStreamSynchronize(r->st));
    std::memcpy(out, r->hout, (size_t)n * kOut * 4);
    return WS_OK;
  }
  ws_status s = ensure_batch(r, n);
  if (s != WS_OK) return s;
  R_HIP(hipMemcpyAsync(r->x, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice, r->st));
  MlpBuffers b = make_buffers(r, n);
  b.h1 = b.h2 = b.h3 = nullptr;
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n)), dim3(256), 0, r->st, r->p, b);
  R_HIP(hipGetLastError());
  R_HIP(hipMemcpyAsync(out, r->out, (size_t)n * kOut * 4, hipMemcpyDeviceToHost, r->st));
  R_HIP(hipStreamSynchronize(r->st));
  return WS_OK;
}

// Trains on raw features x[n][8] (7 used) and log-time targets y[n][kOut] with
// mask[n][kOut] (1 = measured). Standardization statistics are computed here.
// Mini-batch Adam for `epochs` passes; *final_loss receives the last epoch's
// masked MSE in standardized units.
ws_status ws_router_train(ws_router* r, const float* x, const float* y, const float* mask, uint32_t n,
                          uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float* final_loss) {
  if (!r || !x || !y || !mask || n == 0 || batch == 0) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::lock_guard<std::mutex> lk(r->mu);
  R_HIP(hipSetDevice(r->dev));
  // Feature and target statistics.
  for (int k = 0; k < kIn; ++k) {
    double s1 = 0, s2 = 0;
    for (uint32_t i = 0; i < n; ++i) { s1 += x[(size_t)i * kIn + k]; s2 += (double)x[(size_t)i * kIn + k] * x[(size_t)i * kIn + k]; }
    const double mu = s1 / n, var = std::max(0.0, s2 / n - mu * mu);
    r->xmean[k] = (float)mu;
    r->xstd[k] = var > 1e-12 ? (float)std::sqrt(var) : 1.f;
  }
  double ys = 0, ys2 = 0, yc = 0;
  for (size_t i = 0; i < (size_t)n * kOut; ++i)
    if (mask[i] > 0) { ys += y[i]; ys2 += (double)y[i] * y[i]; yc += 1; }
  if (yc == 0) return rfail(WS_ERR_INVALID_ARGUMENT, "no measured targets");
  r->ymean = (float)(ys / yc);
  r->ystd = (float)std::max(1e-3, std::sqrt(std::max(0.0, ys2 / yc - (ys / yc) * (ys / yc))));
  std::vector<float> xn;
  normalize(r, x, n, xn);
  std::vector<float> yn((size_t)n * kOut);
  for (size_t i = 0; i < yn.size(); ++i) yn[i] = mask[i] > 0 ? (y[i] - r->ymean) / r->ystd : 0.f;
  const uint32_t bs = std::min(batch, n);
  ws_status s = ensure_batch(r, bs);
  if (s != WS_OK) return s;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::mt19937 rng(seed);
  std::vector<float> bx((size_t)bs * kIn), by((size_t)bs * kOut), bm((size_t)bs * kOut), bout((size_t)bs * kOut);
  float loss = 0;
  const float b1 = 0.9f, b2 = 0.999f;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    double eloss = 0, ecount = 0;
    for (uint32_t o = 0; o + bs <= n || (o < n && o == 0); o += bs) {
      const uint32_t m = std::min(bs, n - o);
      float cnt = 0;
      for (uint32_t i = 0; i < m; ++i) {
        const uint32_t src = order[o + i];
        std::memcpy(&bx[(size_t)i * kIn], &xn[(size_t)src * kIn], kIn * 4);
        std::memcpy(&by[(size_t)i * kOut], &yn[(size_t)src * kOut], kOut * 4);
        std::memcpy(&bm[(size_t)i * kOut], &mask[(size_t)src * kOut], kOut * 4);
        for (int k = 0; k < kOut; ++k) cnt += bm[(size_t)i * kOut + k] > 0 ? 1.f : 0.f;
      }
      if (cnt == 0) continue;
      R_HIP(hipMemcpyAsync(r->x, bx.data(), (size_t)m * kIn * 4, hipMemcpyHostToDevice, r->st));
      R_HIP(hipMemcpyAsync(r->y, by.data(), (size_t)m * kOut * 4, hipMemcpyHostToDevice, r->st));
      R_HIP(hipMemcpyAsync(r->mask, bm.data(), (size_t)m * kOut * 4, hipMemcpyHostToDevice, r->st));
      s = gradient(r, m, 1.f / cnt);
      if (s != WS_OK) return s;
      ++r->step;
      const float c1 = 1.f - std::pow(b1, (float)r->step), c2 = 1.f - std::pow(b2, (float)r->step);
      hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, r->st, r->p, r->g, r->m, r->v, kParams, lr,
                         b1, b2, 1e-8f, c1, c2);
      R_HIP(hipGetLastError());
      if (ep + 1 == epochs) {
        R_HIP(hipMemcpyAsync(bout.data(), r->out, (size_t)m * kOut * 4, hipMemcpyDeviceToHost, r->st));
        R_HIP(hipStreamSynchronize(r->st));
        for (size_t i = 0; i < (size_t)m * kOut; ++i)
          if (bm[i] > 0) { const double d = bout[i] - by[i]; eloss += d * d; ecount += 1; }
      }
    }
    if (ep + 1 == epochs) loss = ecount > 0 ? (float)(eloss / ecount) : 0.f;
  }
  R_HIP(hipStreamSynchronize(r->st));
  if (final_loss) *final_loss = loss;
  return WS_OK;
}

// Predicted time (ms) of every configuration for raw features x[n][8]; the
// chosen configuration is the fastest one whose mask entry is set.
ws_status ws_router_predict(ws_router* r, const float* x, uint32_t n, const float* mask, float* pred_ms,
                            uint32_t* choice) {
  if (!r || !x || n == 0) return rfail(WS_ERR_INVALID_ARGUMENT, "bad arguments");
  std::vector<float> xn, out((size_t)n * kOut);
  normalize(r, x, n, xn);
  ws_status s = ws_router_forward(r, xn.data(), n, out.data());
  if (s != WS_OK) return s;
  for (uint32_t i = 0; i < n; ++i) {
    int best = -1;
    float bv = 0;
    for (int k = 0; k < kOut; ++k) {
      const float t = std::exp(out[(size_t)i * kOut + k] * r->ystd + r->ymean);
      if (pred_ms) pred_ms[(size_t)i * kOut + k] = t;
      if (mask && !(mask[(size_t)i * kOut + k] > 0)) continue;
      if (best < 0 || t < bv) { best = k; bv = t; }
    }
    if (choice) choice[i] = best < 0 ? 0u : (uint32_t)best;
  }
  return WS_OK;
}

}  // extern "C"

// Frontier-size predictor: a 264 -> 64 -> 64 -> 1 ReLU MLP in FP32, trained
// and evaluated with these kernels.  It predicts log(1 + items) of one
// (group, 4 KiB window) work item from the window's 256-bin byte histogram and
// 8 statistics of the group's automaton; the batched scheduler uses the
// prediction only to size and order work items.  No counterpart in the
// reference.
//
// Parameter layout (floats): W1[264][64], b1[64], W2[64][64], b2[64], W3[64],
// b3[1].  Features of sample n are either a dense row X[n][264] or, for
// inference over many (group, window) pairs, the concatenation of
// hist[window][256] and stats[group][8].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ngap_amd.h"
#include "c_guard.h"

void ngap_set_last_error_internal(const std::string &m) noexcept;

namespace {

constexpr int kIn = 264, kHid = 64;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kHid, kParams = kB3 + 1;
constexpr int kTile = 16;      // samples per forward/backward sub-tile
constexpr int kThreads = 256;

int fail(int code, const std::string &m) {
  ngap_set_last_error_internal(m);
  return code;
}
#define MLP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess)                                                        \
      return fail(_e == hipErrorOutOfMemory ? NGAP_ERR_OOM : NGAP_ERR_DEVICE,    \
                  std::string(#expr) + ": " + hipGetErrorString(_e));            \
  } while (0)

// Loads the features of samples [s0, s0 + kTile) into x[kTile][kIn].
__device__ void load_tile(float *x, const float *X, const float *hist, const float *stats,
                          const uint32_t *pairs, uint64_t n, uint64_t s0) {
  for (int k = threadIdx.x; k < kTile * kIn; k += blockDim.x) {
    const int s = k / kIn, i = k % kIn;
    const uint64_t ns = s0 + s;
    float v = 0.f;
    if (ns < n) {
      if (X) {
        v = X[ns * kIn + i];
      } else {
        const uint32_t g = pairs[2 * ns], w = pairs[2 * ns + 1];
        v = i < 256 ? hist[(uint64_t)w * 256 + i] : stats[(uint64_t)g * 8 + (i - 256)];
      }
    }
    x[k] = v;
  }
}

// h[kTile][64] = relu(x[kTile][in] W[in][64] + b); 4 outputs per thread.
__device__ void dense64(float *h, const float *x, int in, const float *W, const float *b) {
  const int o = threadIdx.x % kHid, s0 = threadIdx.x / kHid;
  float acc[kTile / 4];
#pragma unroll
  for (int k = 0; k < kTile / 4; k++) acc[k] = b[o];
  for (int i = 0; i < in; i++) {
    const float w = W[i * kHid + o];
#pragma unroll
    for (int k = 0; k < kTile / 4; k++) acc[k] += x[(s0 + 4 * k) * in + i] * w;
  }
#pragma unroll
  for (int k = 0; k < kTile / 4; k++) h[(s0 + 4 * k) * kHid + o] = fmaxf(acc[k], 0.f);
}

// Forward over n samples; optionally keeps H1/H2 for the backward pass.
__global__ void __launch_bounds__(kThreads)
    k_forward(const float *P, const float *X, const float *hist, const float *stats,
              const uint32_t *pairs, uint64_t n, float *Y, float *H1, float *H2) {
  __shared__ float x[kTile * kIn];
  __shared__ float h1[kTile * kHid];
  __shared__ float h2[kTile * kHid];
  for (uint64_t s0 = (uint64_t)blockIdx.x * kTile; s0 < n; s0 += (uint64_t)gridDim.x * kTile) {
    load_tile(x, X, hist, stats, pairs, n, s0);
    __syncthreads();
    dense64(h1, x, kIn, P + kW1, P + kB1);
    __syncthreads();
    dense64(h2, h1, kHid, P + kW2, P + kB2);
    __syncthreads();
    if (threadIdx.x < kTile) {
      const int s = threadIdx.x;
      float y = P[kB3];
      for (int i = 0; i < kHid; i++) y += h2[s * kHid + i] * P[kW3 + i];
      if (s0 + s < n) Y[s0 + s] = y;
    }
    if (H1)
      for (int k = threadIdx.x; k < kTile * kHid; k += blockDim.x)
        if (s0 + k / kHid < n) {
          H1[(s0 + k / kHid) * kHid + k % kHid] = h1[k];
          H2[(s0 + k / kHid) * kHid + k % kHid] = h2[k];
        }
    __syncthreads();
  }
}

// Backward of the mean squared error over n samples (targets T): gradients
// accumulated into G (zeroed by the caller), loss sum into *loss.  Each block
// walks its samples in tiles; the block's W1 partial sums live in LDS (each
// entry owned by one thread), the rest in registers, and are added to G once.
constexpr size_t kBackLds = (size_t)(kTile * kIn + 4 * kTile * kHid + kTile + kIn * kHid) * 4;

__global__ void __launch_bounds__(kThreads)
    k_backward(const float *P, const float *X, const float *T, const float *Y,
               const float *H1, const float *H2, uint64_t n, float *G, float *loss) {
  extern __shared__ float lds[];
  float *x = lds;                      // [kTile][kIn]
  float *h1 = x + kTile * kIn;         // [kTile][kHid]
  float *h2 = h1 + kTile * kHid;
  float *d2 = h2 + kTile * kHid;
  float *d1 = d2 + kTile * kHid;
  float *dy = d1 + kTile * kHid;       // [kTile]
  float *gw1 = dy + kTile;             // [kIn][kHid]
  const int t = threadIdx.x;
  for (int e = t; e < kIn * kHid; e += kThreads) gw1[e] = 0.f;
  float gw2[kHid * kHid / kThreads];
  float gw3 = 0.f, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
#pragma unroll
  for (int k = 0; k < kHid * kHid / kThreads; k++) gw2[k] = 0.f;
  const float scale = 2.0f / (float)n;
  for (uint64_t s0 = (uint64_t)blockIdx.x * kTile; s0 < n; s0 += (uint64_t)gridDim.x * kTile) {
    load_tile(x, X, nullptr, nullptr, nullptr, n, s0);
    for (int k = t; k < kTile * kHid; k += blockDim.x) {
      const uint64_t ns = s0 + k / kHid;
      h1[k] = ns < n ? H1[ns * kHid + k % kHid] : 0.f;
      h2[k] = ns < n ? H2[ns * kHid + k % kHid] : 0.f;
    }
    if (t < kTile) {
      const uint64_t ns = s0 + t;
      const float e = ns < n ? Y[ns] - T[ns] : 0.f;
      dy[t] = e * scale;
      lsum += e * e;
    }
    __syncthreads();
    // d2 = dy W3^T * relu'(h2);   d1 = d2 W2^T * relu'(h1)
    for (int k = t; k < kTile * kHid; k += blockDim.x) {
      const int s = k / kHid, o = k % kHid;
      d2[k] = h2[k] > 0.f ? dy[s] * P[kW3 + o] : 0.f;
    }
    __syncthreads();
    for (int k = t; k < kTile * kHid; k += blockDim.x) {
      const int s = k / kHid, i = k % kHid;
      float a = 0.f;
      for (int o = 0; o < kHid; o++) a += d2[s * kHid + o] * P[kW2 + i * kHid + o];
      d1[k] = h1[k] > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int e = t; e < kIn * kHid; e += kThreads) {
      const int i = e / kHid, o = e % kHid;
      float a = 0.f;
#pragma unroll 4
      for (int s = 0; s < kTile; s++) a += x[s * kIn + i] * d1[s * kHid + o];
      gw1[e] += a;
    }
#pragma unroll
    for (int k = 0; k < kHid * kHid / kThreads; k++) {
      const int e = t + k * kThreads, i = e / kHid, o = e % kHid;
      float a = 0.f;
      for (int s = 0; s < kTile; s++) a += h1[s * kHid + i] * d2[s * kHid + o];
      gw2[k] += a;
    }
    if (t < kHid) {
      float a = 0.f, c = 0.f, d = 0.f;
      for (int s = 0; s < kTile; s++) {
        a += h2[s * kHid + t] * dy[s];
        c += d1[s * kHid + t];
        d += d2[s * kHid + t];
      }
      gw3 += a;
      gb1 += c;
      gb2 += d;
    }
    if (t == 0)
      for (int s = 0; s < kTile; s++) gb3 += dy[s];
    __syncthreads();
  }
  for (int e = t; e < kIn * kHid; e += kThreads) atomicAdd(&G[kW1 + e], gw1[e]);
#pragma unroll
  for (int k = 0; k < kHid * kHid / kThreads; k++) atomicAdd(&G[kW2 + t + k * kThreads], gw2[k]);
  if (t < kHid) {
    atomicAdd(&G[kW3 + t], gw3);
    atomicAdd(&G[kB1 + t], gb1);
    atomicAdd(&G[kB2 + t], gb2);
  }
  if (t == 0) atomicAdd(&G[kB3], gb3);
  if (t < kTile) atomicAdd(loss, lsum);
}

// Adam step over all parameters; clears the gradient.
__global__ void k_adam(float *P, float *G, float *M, float *V, int count, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  G[i] = 0.f;
}

__global__ void k_zero(float *p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.f;
}

struct Buf {
  void *p = nullptr;
  size_t n = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  hipError_t reserve(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
    if (e == hipSuccess) n = bytes ? bytes : 8;
    return e;
  }
  float *f() const { return reinterpret_cast<float *>(p); }
};

}  // namespace

struct ngap_mlp {
  int device = 0;
  int blocks = 1024;
  Buf P, G, M, V, loss;
  Buf X, T, Y, H1, H2, hist, stats, pairs;
  int step = 0;
};

namespace {

int forward(ngap_mlp *m, uint64_t n, bool keep) {
  hipLaunchKernelGGL(k_forward, dim3(m->blocks), dim3(kThreads), 0, 0, m->P.f(), m->X.f(),
                     nullptr, nullptr, nullptr, n, m->Y.f(), keep ? m->H1.f() : nullptr,
                     keep ? m->H2.f() : nullptr);
  MLP_TRY(hipGetLastError());
  return NGAP_OK;
}

int load_batch(ngap_mlp *m, const float *x, const float *t, uint64_t n) {
  MLP_TRY(m->X.reserve(n * kIn * 4));
  MLP_TRY(m->Y.reserve(n * 4));
  MLP_TRY(m->H1.reserve(n * kHid * 4));
  MLP_TRY(m->H2.reserve(n * kHid * 4));
  MLP_TRY(hipMemcpy(m->X.p, x, n * kIn * 4, hipMemcpyHostToDevice));
  if (t) {
    MLP_TRY(m->T.reserve(n * 4));
    MLP_TRY(hipMemcpy(m->T.p, t, n * 4, hipMemcpyHostToDevice));
  }
  return NGAP_OK;
}

int grad_on_device(ngap_mlp *m, uint64_t n) {
  int rc = forward(m, n, true);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(1), 0, 0, m->loss.f(), 1);
  MLP_TRY(hipGetLastError());
  uint32_t blocks = (uint32_t)std::min<uint64_t>((n + kTile * 8 - 1) / (kTile * 8), 512);
  if (!blocks) blocks = 1;
  hipLaunchKernelGGL(k_backward, dim3(blocks), dim3(kThreads), kBackLds, 0, m->P.f(), m->X.f(),
                     m->T.f(), m->Y.f(), m->H1.f(), m->H2.f(), n, m->G.f(), m->loss.f());
  MLP_TRY(hipGetLastError());
  return NGAP_OK;
}

}  // namespace

extern "C" {

int ngap_mlp_create(int32_t device, uint64_t seed, ngap_mlp **out) {
  return ngapamd::c_guard([&]() -> int {
    if (!out) return fail(NGAP_ERR_INVALID, "NULL argument");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      return fail(NGAP_ERR_DEVICE, "no HIP device available");
    MLP_TRY(hipSetDevice(device));
    MLP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k_backward),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBackLds));
    ngap_mlp *m = new ngap_mlp();
    m->device = device;
    std::vector<float> p(kParams, 0.f);
    // He-uniform initialisation from a splitmix64 stream
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      return (double)(x >> 11) / 9007199254740992.0;
    };
    auto fill = [&](int off, int cnt, int fan_in) {
      const double a = std::sqrt(6.0 / fan_in);
      for (int i = 0; i < cnt; i++) p[off + i] = (float)((2 * next() - 1) * a);
    };
    fill(kW1, kIn * kHid, kIn);
    fill(kW2, kHid * kHid, kHid);
    fill(kW3, kHid, kHid);
    int rc = NGAP_OK;
    if (m->P.reserve(kParams * 4) || m->G.reserve(kParams * 4) || m->M.reserve(kParams * 4) ||
        m->V.reserve(kParams * 4) || m->loss.reserve(4)) {
      delete m;
      return fail(NGAP_ERR_OOM, "MLP allocation failed");
    }
    if (hipMemcpy(m->P.p, p.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess) rc = 1;
    std::vector<float> zero(kParams, 0.f);
    if (hipMemcpy(m->G.p, zero.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->M.p, zero.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->V.p, zero.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = 1;
    if (rc) {
      delete m;
      return fail(NGAP_ERR_DEVICE, "MLP initialisation failed");
    }
    *out = m;
    return NGAP_OK;
  });
}

int ngap_mlp_destroy(ngap_mlp *m) {
  return ngapamd::c_guard([&]() -> int {
    delete m;
    return NGAP_OK;
  });
}

int32_t ngap_mlp_num_params(void) { return kParams; }

int ngap_mlp_get_params(ngap_mlp *m, float *out) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || !out) return fail(NGAP_ERR_INVALID, "NULL argument");
    MLP_TRY(hipMemcpy(out, m->P.p, kParams * 4, hipMemcpyDeviceToHost));
    return NGAP_OK;
  });
}

int ngap_mlp_set_params(ngap_mlp *m, const float *in) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || !in) return fail(NGAP_ERR_INVALID, "NULL argument");
    MLP_TRY(hipMemcpy(m->P.p, in, kParams * 4, hipMemcpyHostToDevice));
    return NGAP_OK;
  });
}

int ngap_mlp_predict(ngap_mlp *m, const float *x, uint64_t n, float *y) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || (!x && n) || (!y && n)) return fail(NGAP_ERR_INVALID, "NULL argument");
    if (!n) return NGAP_OK;
    MLP_TRY(hipSetDevice(m->device));
    int rc = load_batch(m, x, nullptr, n);
    if (rc) return rc;
    rc = forward(m, n, false);
    if (rc) return rc;
    MLP_TRY(hipMemcpy(y, m->Y.p, n * 4, hipMemcpyDeviceToHost));
    return NGAP_OK;
  });
}

int ngap_mlp_predict_pairs(ngap_mlp *m, const float *hist, uint32_t num_windows,
                           const float *stats, uint32_t num_groups, const uint32_t *pairs,
                           uint64_t n, float *y) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || !hist || !stats || !pairs || !y) return fail(NGAP_ERR_INVALID, "NULL argument");
    for (uint64_t i = 0; i < n; i++)
      if (pairs[2 * i] >= num_groups || pairs[2 * i + 1] >= num_windows)
        return fail(NGAP_ERR_INVALID, "pair out of range");
    if (!n) return NGAP_OK;
    MLP_TRY(hipSetDevice(m->device));
    MLP_TRY(m->hist.reserve((size_t)num_windows * 256 * 4));
    MLP_TRY(m->stats.reserve((size_t)num_groups * 8 * 4));
    MLP_TRY(m->pairs.reserve(n * 8));
    MLP_TRY(m->Y.reserve(n * 4));
    MLP_TRY(hipMemcpy(m->hist.p, hist, (size_t)num_windows * 256 * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->stats.p, stats, (size_t)num_groups * 8 * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->pairs.p, pairs, n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_forward, dim3(m->blocks), dim3(kThreads), 0, 0, m->P.f(), nullptr,
                       m->hist.f(), m->stats.f(), reinterpret_cast<uint32_t *>(m->pairs.p), n,
                       m->Y.f(), nullptr, nullptr);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpy(y, m->Y.p, n * 4, hipMemcpyDeviceToHost));
    return NGAP_OK;
  });
}

int ngap_mlp_gradient(ngap_mlp *m, const float *x, const float *t, uint64_t n, float *grad,
                      float *loss) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || !x || !t || !grad || !loss || !n) return fail(NGAP_ERR_INVALID, "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    int rc = load_batch(m, x, t, n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero, dim3((kParams + 255) / 256), dim3(256), 0, 0, m->G.f(), kParams);
    MLP_TRY(hipGetLastError());
    rc = grad_on_device(m, n);
    if (rc) return rc;
    MLP_TRY(hipMemcpy(grad, m->G.p, kParams * 4, hipMemcpyDeviceToHost));
    float l = 0;
    MLP_TRY(hipMemcpy(&l, m->loss.p, 4, hipMemcpyDeviceToHost));
    *loss = l / (float)n;
    hipLaunchKernelGGL(k_zero, dim3((kParams + 255) / 256), dim3(256), 0, 0, m->G.f(), kParams);
    MLP_TRY(hipGetLastError());
    return NGAP_OK;
  });
}

int ngap_mlp_train(ngap_mlp *m, const float *x, const float *t, uint64_t n, int32_t epochs,
                   uint32_t batch, float lr, uint64_t seed, float *final_loss) {
  return ngapamd::c_guard([&]() -> int {
    if (!m || !x || !t || !n || epochs < 1 || batch < 1)
      return fail(NGAP_ERR_INVALID, "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    // shuffled copy per epoch on the host; each mini-batch is uploaded once
    std::vector<uint64_t> idx(n);
    for (uint64_t i = 0; i < n; i++) idx[i] = i;
    std::vector<float> bx((size_t)batch * kIn), bt(batch);
    uint64_t z = seed;
    double last = 0;
    for (int e = 0; e < epochs; e++) {
      for (uint64_t i = n - 1; i > 0; i--) {
        z = z * 6364136223846793005ull + 1442695040888963407ull;
        std::swap(idx[i], idx[(z >> 33) % (i + 1)]);
      }
      double lsum = 0;
      uint64_t cnt = 0;
      for (uint64_t b0 = 0; b0 < n; b0 += batch) {
        const uint64_t nb = std::min<uint64_t>(batch, n - b0);
        for (uint64_t k = 0; k < nb; k++) {
          memcpy(&bx[k * kIn], x + idx[b0 + k] * kIn, kIn * 4);
          bt[k] = t[idx[b0 + k]];
        }
        int rc = load_batch(m, bx.data(), bt.data(), nb);
        if (rc) return rc;
        rc = grad_on_device(m, nb);
        if (rc) return rc;
        m->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, 0, m->P.f(), m->G.f(),
                           m->M.f(), m->V.f(), kParams, lr, b1, b2, 1e-8f, c1, c2);
        MLP_TRY(hipGetLastError());
        float l = 0;
        MLP_TRY(hipMemcpy(&l, m->loss.p, 4, hipMemcpyDeviceToHost));
        lsum += l;
        cnt += nb;
      }
      last = lsum / (double)cnt;
    }
    if (final_loss) *final_loss = (float)last;
    return NGAP_OK;
  });
}

}  // extern "C"

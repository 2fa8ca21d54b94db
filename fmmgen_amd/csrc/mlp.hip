// Kernels of the order-selection network (layout and semantics in mlp.h).
#include "mlp.h"

#include <mutex>

namespace fmm {
namespace {

constexpr int F = kMlpF, H = kMlpH, O = kMlpO;

__device__ __forceinline__ void stage_params(const float* __restrict__ P, float* sP) {
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
}

// Sum of the mask (the number of valid outputs of the batch) into cnt[0];
// loss[0] cleared for the forward pass that adds into it.
__global__ void k_mask_count(long long n, const float* __restrict__ mask, float* __restrict__ cnt,
                             float* __restrict__ loss) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += mask[i] != 0.f ? 1.f : 0.f;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[0] = red[0];
    loss[0] = 0.f;
  }
}

// Forward pass, one thread per sample.  With T: the masked mean-square loss
// (added into loss[0]) and dL/dY.
__global__ __launch_bounds__(64) void k_mlp_fwd(int B, const float* __restrict__ X, const float* __restrict__ P,
                                                float* __restrict__ H1, float* __restrict__ H2,
                                                float* __restrict__ Y, const float* __restrict__ T,
                                                const float* __restrict__ mask, const float* __restrict__ cnt,
                                                float* __restrict__ dY, float* __restrict__ loss) {
  __shared__ float sP[kMlpParams];
  __shared__ float red[64];
  stage_params(P, sP);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float lsum = 0.f;
  if (b < B) {
    float x[F], h1[H], h2[H];
#pragma unroll
    for (int f = 0; f < F; ++f) x[f] = X[(size_t)b * F + f];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float a = sP[kMlpB1 + j];
#pragma unroll
      for (int f = 0; f < F; ++f) a += sP[kMlpW1 + j * F + f] * x[f];
      h1[j] = a > 0.f ? a : 0.f;
    }
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float a = sP[kMlpB2 + j];
#pragma unroll
      for (int i = 0; i < H; ++i) a += sP[kMlpW2 + j * H + i] * h1[i];
      h2[j] = a > 0.f ? a : 0.f;
    }
    if (H1) {
#pragma unroll
      for (int j = 0; j < H; ++j) {
        H1[(size_t)b * H + j] = h1[j];
        H2[(size_t)b * H + j] = h2[j];
      }
    }
    const float inv = (T && cnt[0] > 0.f) ? 1.f / cnt[0] : 0.f;
#pragma unroll
    for (int o = 0; o < O; ++o) {
      float a = sP[kMlpB3 + o];
#pragma unroll
      for (int i = 0; i < H; ++i) a += sP[kMlpW3 + o * H + i] * h2[i];
      Y[(size_t)b * O + o] = a;
      if (T) {
        const bool valid = mask[(size_t)b * O + o] != 0.f;
        const float e = valid ? a - T[(size_t)b * O + o] : 0.f;
        dY[(size_t)b * O + o] = 2.f * e * inv;
        lsum += e * e * inv;
      }
    }
  }
  if (T) {
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
  }
}

// Backward through the hidden layers, one thread per sample:
//   D2 = (W3^T dY) * [h2 > 0],  D1 = (W2^T D2) * [h1 > 0].
__global__ __launch_bounds__(64) void k_mlp_bwd(int B, const float* __restrict__ P, const float* __restrict__ H1,
                                                const float* __restrict__ H2, const float* __restrict__ dY,
                                                float* __restrict__ D2, float* __restrict__ D1) {
  __shared__ float sP[kMlpParams];
  stage_params(P, sP);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float dy[O], d2[H];
#pragma unroll
  for (int o = 0; o < O; ++o) dy[o] = dY[(size_t)b * O + o];
#pragma unroll
  for (int i = 0; i < H; ++i) {
    float a = 0.f;
#pragma unroll
    for (int o = 0; o < O; ++o) a += sP[kMlpW3 + o * H + i] * dy[o];
    d2[i] = H2[(size_t)b * H + i] > 0.f ? a : 0.f;
    D2[(size_t)b * H + i] = d2[i];
  }
  for (int k = 0; k < H; ++k) {
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < H; ++i) a += sP[kMlpW2 + i * H + k] * d2[i];
    D1[(size_t)b * H + k] = H1[(size_t)b * H + k] > 0.f ? a : 0.f;
  }
}

// dW[o][i] = sum_b Dout[b][o] Ain[b][i] and db[o] = sum_b Dout[b][o]; one
// thread per (o, i), i == NI for the bias.
__global__ void k_mlp_wgrad(int B, int NO, int NI, const float* __restrict__ Dout, const float* __restrict__ Ain,
                            float* __restrict__ dW, float* __restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NO * (NI + 1)) return;
  const int o = t / (NI + 1), i = t % (NI + 1);
  float a = 0.f;
  if (i < NI) {
    for (int b = 0; b < B; ++b) a += Dout[(size_t)b * NO + o] * Ain[(size_t)b * NI + i];
    dW[o * NI + i] = a;
  } else {
    for (int b = 0; b < B; ++b) a += Dout[(size_t)b * NO + o];
    db[o] = a;
  }
}

__global__ void k_adam(int n, const float* __restrict__ g, float* __restrict__ p, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_to_half(int n, const float* __restrict__ a, _Float16* __restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = (_Float16)a[i];
}

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float dot2(h2_t a, h2_t b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }

// FP16 inference, one thread per sample: activations and weights in FP16,
// products accumulated in FP32 by v_dot2_f32_f16.
__global__ __launch_bounds__(64) void k_mlp_infer(int B, const float* __restrict__ X,
                                                  const _Float16* __restrict__ HP, float* __restrict__ Y,
                                                  int* __restrict__ order, float log10_tol, float margin,
                                                  int min_order, int max_order) {
  __shared__ h2_t sW[(kMlpParams + 1) / 2];
  const _Float16* sP = reinterpret_cast<const _Float16*>(sW);
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) reinterpret_cast<_Float16*>(sW)[i] = HP[i];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  h2_t x[F / 2], h1[H / 2], h2[H / 2];
#pragma unroll
  for (int f = 0; f < F / 2; ++f) x[f] = h2_t{(_Float16)X[(size_t)b * F + 2 * f], (_Float16)X[(size_t)b * F + 2 * f + 1]};
  auto w2 = [&](int off) { return h2_t{sP[off], sP[off + 1]}; };
#pragma unroll
  for (int j = 0; j < H; j += 2) {
    float a0 = (float)sP[kMlpB1 + j], a1 = (float)sP[kMlpB1 + j + 1];
#pragma unroll
    for (int f = 0; f < F / 2; ++f) {
      a0 = dot2(w2(kMlpW1 + j * F + 2 * f), x[f], a0);
      a1 = dot2(w2(kMlpW1 + (j + 1) * F + 2 * f), x[f], a1);
    }
    h1[j / 2] = h2_t{(_Float16)(a0 > 0.f ? a0 : 0.f), (_Float16)(a1 > 0.f ? a1 : 0.f)};
  }
#pragma unroll
  for (int j = 0; j < H; j += 2) {
    float a0 = (float)sP[kMlpB2 + j], a1 = (float)sP[kMlpB2 + j + 1];
#pragma unroll
    for (int i = 0; i < H / 2; ++i) {
      a0 = dot2(w2(kMlpW2 + j * H + 2 * i), h1[i], a0);
      a1 = dot2(w2(kMlpW2 + (j + 1) * H + 2 * i), h1[i], a1);
    }
    h2[j / 2] = h2_t{(_Float16)(a0 > 0.f ? a0 : 0.f), (_Float16)(a1 > 0.f ? a1 : 0.f)};
  }
  int q = -1;
#pragma unroll
  for (int o = 0; o < O; ++o) {
    float a = (float)sP[kMlpB3 + o];
#pragma unroll
    for (int i = 0; i < H / 2; ++i) a = dot2(w2(kMlpW3 + o * H + 2 * i), h2[i], a);
    if (Y) Y[(size_t)b * O + o] = a;
    if (q < 0 && a + margin <= log10_tol) q = o + kMlpOrder0;
  }
  if (order) {
    if (q < 0) q = max_order;
    order[b] = q < min_order ? min_order : (q > max_order ? max_order : q);
  }
}

inline int nb(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// Workspace of the allocation-free gradient.  In host code `scratch` names
// a host-side stand-in, never the device array: host code uses only the
// per-device address from hipGetSymbolAddress.  Nothing in it has to start
// at zero (every kernel writes what it later reads).
static __device__ __align__(256) float scratch[mlp_work_floats(kMlpScratchBatch)];

constexpr int kMaxDevices = 64;

}  // namespace

float* mlp_scratch() {
  static std::mutex mu;
  static float* addr[kMaxDevices] = {};
  int dev = 0;
  FMM_HIP(hipGetDevice(&dev));
  FMM_CHECK(dev >= 0 && dev < kMaxDevices, kUnsupported, "device index beyond the workspace table");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    FMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FMM_CHECK(p != nullptr, kHipError, "no device address for the gradient workspace");
    addr[dev] = static_cast<float*>(p);
  }
  return addr[dev];
}

void mlp_forward(int B, const float* X, const float* params, float* Y, hipStream_t st) {
  if (B <= 0) return;
  k_mlp_fwd<<<nb(B, 64), 64, 0, st>>>(B, X, params, nullptr, nullptr, Y, nullptr, nullptr, nullptr, nullptr,
                                      nullptr);
}

void mlp_grad(int B, const float* X, const float* T, const float* mask, const float* params, float* grad,
              float* loss, float* work, hipStream_t st) {
  FMM_CHECK(B > 0, kInvalidArgument, "empty batch");
  float* H1 = work;
  float* H2 = H1 + (size_t)B * H;
  float* Y = H2 + (size_t)B * H;
  float* dY = Y + (size_t)B * O;
  float* D2 = dY + (size_t)B * O;
  float* D1 = D2 + (size_t)B * H;
  float* cnt = D1 + (size_t)B * H;
  k_mask_count<<<1, 256, 0, st>>>((long long)B * O, mask, cnt, loss);
  k_mlp_fwd<<<nb(B, 64), 64, 0, st>>>(B, X, params, H1, H2, Y, T, mask, cnt, dY, loss);
  k_mlp_bwd<<<nb(B, 64), 64, 0, st>>>(B, params, H1, H2, dY, D2, D1);
  k_mlp_wgrad<<<nb(O * (H + 1), 128), 128, 0, st>>>(B, O, H, dY, H2, grad + kMlpW3, grad + kMlpB3);
  k_mlp_wgrad<<<nb(H * (H + 1), 128), 128, 0, st>>>(B, H, H, D2, H1, grad + kMlpW2, grad + kMlpB2);
  k_mlp_wgrad<<<nb(H * (F + 1), 128), 128, 0, st>>>(B, H, F, D1, X, grad + kMlpW1, grad + kMlpB1);
}

void mlp_adam(const float* grad, float* params, float* m, float* v, float lr, float beta1, float beta2, float eps,
              int step, hipStream_t st) {
  FMM_CHECK(step >= 1, kInvalidArgument, "Adam step must be >= 1");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  k_adam<<<nb(kMlpParams, 256), 256, 0, st>>>(kMlpParams, grad, params, m, v, lr, beta1, beta2, eps, c1, c2);
}

void mlp_to_half(const float* params, _Float16* hparams, hipStream_t st) {
  k_to_half<<<nb(kMlpParams, 256), 256, 0, st>>>(kMlpParams, params, hparams);
}

void mlp_infer_f16(int B, const float* X, const _Float16* hparams, float* Y, int* order, float log10_tol,
                   float margin, int min_order, int max_order, hipStream_t st) {
  if (B <= 0) return;
  k_mlp_infer<<<nb(B, 64), 64, 0, st>>>(B, X, hparams, Y, order, log10_tol, margin, min_order, max_order);
}

}  // namespace fmm

// Order-selection network: a 3-layer perceptron (F inputs, two hidden ReLU
// layers of H units, O outputs) with FP32 weights.  Training kernels
// (forward with saved activations, masked mean-square loss, backward,
// weight gradients, Adam) run in FP32; inference runs in FP16 with FP32
// accumulation.  The network has no counterpart in the reference: it picks
// per leaf cell the lowest expansion order whose predicted error meets a
// field tolerance.
//
// Parameter layout (FP32, flat): W1[H][F], b1[H], W2[H][H], b2[H], W3[O][H], b3[O].
// Output o is the predicted log10 of a leaf's far-field error, relative to
// the rms field, at expansion order kMlpOrder0 + o.
#pragma once
#include "fmm_common.h"

namespace fmm {

constexpr int kMlpF = 12;
constexpr int kMlpH = 64;
constexpr int kMlpO = 11;
constexpr int kMlpOrder0 = 2;  // order of output 0 (orders 2 .. 12)
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + kMlpH * kMlpF, kMlpW2 = kMlpB1 + kMlpH, kMlpB2 = kMlpW2 + kMlpH * kMlpH,
              kMlpW3 = kMlpB2 + kMlpH, kMlpB3 = kMlpW3 + kMlpO * kMlpH, kMlpParams = kMlpB3 + kMlpO;

// Workspace of one training batch of B samples (device floats).
constexpr size_t mlp_work_floats(int B) { return (size_t)B * (2 * kMlpH + kMlpO) * 2 + 1; }

// Largest batch of the allocation-free gradient (its workspace is a fixed
// device array of mlp_work_floats(kMlpScratchBatch) floats, 9.1 MB).
constexpr int kMlpScratchBatch = 8192;

// FP32 forward: Y[B][O].
void mlp_forward(int B, const float* X, const float* params, float* Y, hipStream_t st);
// Masked mean-square loss over valid (mask != 0) outputs and its gradient
// with respect to all parameters: grad[kMlpParams], loss[1] (device).
void mlp_grad(int B, const float* X, const float* T, const float* mask, const float* params, float* grad,
              float* loss, float* work, hipStream_t st);
// The fixed workspace on the current device (its address is looked up on
// the first call per device; nothing is allocated).  Every call on one
// device shares it: gradients that overlap on one device race.
float* mlp_scratch();
// One Adam step (bias-corrected, step >= 1).
void mlp_adam(const float* grad, float* params, float* m, float* v, float lr, float beta1, float beta2, float eps,
              int step, hipStream_t st);
// FP16 parameters for inference.
void mlp_to_half(const float* params, _Float16* hparams, hipStream_t st);
// FP16 inference: Y[B][O] (optional) and, when order != nullptr, the lowest
// order o + kMlpOrder0 with Y[o] + margin <= log10(tol), else max_order;
// orders are clamped to [min_order, max_order].
void mlp_infer_f16(int B, const float* X, const _Float16* hparams, float* Y, int* order, float log10_tol,
                   float margin, int min_order, int max_order, hipStream_t st);

}  // namespace fmm

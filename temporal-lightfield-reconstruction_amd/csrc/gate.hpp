// Learned gate: a 24-64-64-1 MLP with BF16 operands and FP32 accumulation on
// MFMA (v_mfma_f32_16x16x32_bf16), with hand-written forward, backward and
// Adam kernels.  It has no counterpart in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlr {

constexpr int kGateIn = 24;
constexpr int kGateHid = 64;
// Parameter vector layout (fp32 master copy):
//   W1[24][64] (k-major), b1[64], W2[64][64] (k-major), b2[64], w3[64], b3
constexpr int kGateOffW1 = 0;
constexpr int kGateOffB1 = kGateOffW1 + kGateIn * kGateHid;
constexpr int kGateOffW2 = kGateOffB1 + kGateHid;
constexpr int kGateOffB2 = kGateOffW2 + kGateHid * kGateHid;
constexpr int kGateOffW3 = kGateOffB2 + kGateHid;
constexpr int kGateOffB3 = kGateOffW3 + kGateHid;
constexpr int kGateParams = kGateOffB3 + 1;  // 5825

// Logits z for n feature rows (24 floats each).
hipError_t launch_gate_forward(const float* X, int n, const float* params, float* logits, hipStream_t s);
// Same for *n_dev rows (a device-side count, at most max_n).
hipError_t launch_gate_forward_dev(const float* X, const unsigned* n_dev, size_t max_n, const float* params,
                                   float* logits, hipStream_t s);
// Accumulates d(loss)/d(params) over rows idx[0..n) into grad (zeroed by the
// caller) and the summed loss into *loss.  Loss: weighted binary cross
// entropy on logits, weight pos_w for label 1.
hipError_t launch_gate_grad(const float* X, const float* y, const int* idx, int n, const float* params,
                            float* grad, float* loss, float pos_w, hipStream_t s);
// Adam step on the fp32 master parameters; grad is scaled by gscale.
hipError_t launch_gate_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1,
                            float beta2, float eps, int step, float gscale, hipStream_t s);

}  // namespace tlr

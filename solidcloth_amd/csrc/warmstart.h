// Learned warm-start for the Gauss-Seidel sweeps: a per-vertex MLP over
// one-ring features predicts the correction the sweeps will apply, and the
// prediction is added to the inertial positions before the sweeps.
//
// Network: 12 inputs -> 64 -> 64 -> 3 outputs, GELU (tanh form) after both
// hidden layers. Parameters are FP32 masters (Adam), the MFMA operands are
// BF16 copies, activations enter each product as BF16, products accumulate in
// FP32 (v_mfma_f32_16x16x16_bf16), biases and GELU are FP32.
//
// Parameter vector (kMlpParams floats): W1 [12][64] (input-major), b1 [64],
// W2 [64][64], b2 [64], W3 [64][3], b3 [3].
//
// Features of vertex v (all FP32, rounded to BF16 once; scales fixed):
//   0-2  (xhat - prev) / ell              inertial displacement of the substep
//   3-5  10 c / ell                       c = last substep's correction (pos - xhat)
//   6-8  10 (mean of c over the one-ring neighbours - c) / ell
//   9-11 100 x mean over incident triangles of (warp, weft, shear) strain at
//        xhat: 0.5(|wu|^2 - lu2), 0.5(|wv|^2 - lv2), wu.wv (ClothPBD.cpp:78,89,101)
// ell = the mean rest edge length. Output y: the correction is ell y / 10.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace sc {

enum {
    kMlpIn = 12,
    kMlpHidden = 64,
    kMlpOut = 3,
    kOffW1 = 0,
    kOffB1 = 768,
    kOffW2 = 832,
    kOffB2 = 4928,
    kOffW3 = 4992,
    kOffB3 = 5184,
    kMlpParams = 5187,
    kMlpStride = 5188,  // per-block gradient partial: the gradient, then the loss sum
    kMlpOperands = 44,  // packed BF16 MFMA operands, 64 lanes x 4 values each
    kMlpMaxBlocks = 256,
};

constexpr float kWsScaleCorr = 10.f;
constexpr float kWsScaleStrain = 100.f;
constexpr float kWsScaleOut = 10.f;

struct MlpWeights {
    const float* master;     // kMlpParams FP32 (biases are read from here)
    const uint16_t* packed;  // kMlpOperands x 256 BF16
};

// Per-model buffers of the warm start.
struct WarmBufs {
    float4* xhat;       // positions after prediction
    float4* corr;       // correction of the last substep (pos - xhat)
    uint16_t* X;        // features, 16 BF16 per vertex (12 used)
    float4* T;          // training targets (output units) and weight (0 = pinned)
    float4* triStrain;  // per triangle stream position: warp, weft, shear
    const int* adjOff;  // one-ring neighbours, CSR
    const int* adj;
    const int* vtOff;   // incident triangle stream positions, CSR
    const int* vt;
};

void launch_ws_tri_strain(hipStream_t s, const float4* pos, TriStreams tr, int nt, float4* out);
void launch_ws_features(hipStream_t s, const float4* pos, const float4* prev, WarmBufs b, int nv, float invEll);
// pos += outScale * y for vertices with non-zero inverse mass
void launch_mlp_apply(hipStream_t s, MlpWeights w, const uint16_t* X, int n, float4* pos, float outScale);
void launch_mlp_forward(hipStream_t s, MlpWeights w, const uint16_t* X, int n, float* Y);
// corr = pos - xhat; with train, T = (corr targetScale, weight)
void launch_ws_after(hipStream_t s, const float4* pos, WarmBufs b, int nv, int train, float targetScale);
// loss = sum_v w_v |y_v - t_v|^2 * invCount and its gradient: per-block
// partials (kMlpStride floats each) reduced in block order into grad
int mlp_backward_blocks(int n);
void launch_mlp_backward(hipStream_t s, MlpWeights w, const uint16_t* X, const float4* T, int n, float invCount,
                         float* partial, int blocks);
void launch_mlp_reduce(hipStream_t s, const float* partial, int blocks, float* grad);
// the step size is read from device memory, so a captured graph follows changes
void launch_mlp_adam(hipStream_t s, float* master, float* m, float* v, const float* grad, const int* step,
                     const float* lr);
// BF16 operands from the masters; increments *step when step is non-null
void launch_mlp_pack(hipStream_t s, const float* master, uint16_t* packed, int* step);
// FP32 features (n x 12) to the BF16 layout
void launch_pack_features(hipStream_t s, const float* Xf, int n, uint16_t* X);
void launch_unpack_features(hipStream_t s, const uint16_t* X, int n, float* Xf);
// Forward pass over device features Xf (n x 12 FP32) into device Y (n x 3)
// with the BF16 rows staged in the library's static per-GPU scratch array
// instead of allocated memory: n <= kNoAllocRows, enqueued on s. Calls that
// overlap on one GPU share the array and race.
enum { kNoAllocRows = 1 << 18 };
void launch_mlp_forward_noalloc(hipStream_t s, MlpWeights w, const float* Xf, int n, float* Y);
// sums of squared constraint errors in FP64: out[0] = sum of warp^2 + weft^2,
// out[1] = sum of shear^2, out[2] = sum of (sStiff theta)^2 over the bends;
// partial holds 3 doubles per block
int residual_blocks(int items);
void launch_residual(hipStream_t s, const float4* pos, TriStreams tr, int nt, const int4* bidx, int nb,
                     double sStiff, double* partial, int blocks, double* out);

}  // namespace sc

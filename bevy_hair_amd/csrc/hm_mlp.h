// Shapes, argument blocks and launchers of the learned shading term.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hm_kernels.h"

namespace hm {

// Features (all 32 used): tangent 3, view 3, key-light direction 3, the 16
// DOM layer transmittances, sigma_a * MLP_SIGMA_SCALE 3, roughness, then the
// key light's theta_h / pi, cos(theta_o) and optical depth * MLP_DEPTH_SCALE.
constexpr int MLP_IN = 32;
constexpr int MLP_NFEAT = 32;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 16;
constexpr int MLP_NOUT = 16;    // all outputs used
constexpr int MLP_TILES = 22;   // 16x32 bf16 operand tiles over the four layers
constexpr int MLP_NW = 64 * 32 + 64 * 64 + 64 * 64 + 16 * 64;  // 11264 weights
constexpr int MLP_NBIAS = 64 + 64 + 64 + 16;                    // 208 biases
constexpr float MLP_SIGMA_SCALE = 0.25f;
constexpr float MLP_DEPTH_SCALE = 0.125f;

// Outputs: the dual-scattering terms of dual_scatter (shade.wesl:153-190),
//   0       forward R lobe sum (the same in every channel)
//   1-3     forward TT lobe sum, 4-6 forward TRT lobe sum
//   7-9     a_back, 10-12 delta_b, 13-15 sigma_b^2.
// The non-negative ones (all but delta_b, a signed angle) span orders of
// magnitude: the network predicts log(value + MLP_LOG_EPS), so that its error
// is relative.
constexpr float MLP_LOG_EPS = 1e-6f;
__host__ __device__ constexpr bool mlp_log_output(int k) { return k < 10 || k >= 13; }

// Canonical float32 layout: W_L is (out x in) row-major, layers back to back.
__host__ __device__ constexpr int mlp_w_offset(int L) {
    return L == 0 ? 0 : (L == 1 ? 64 * 32 : (L == 2 ? 64 * 32 + 64 * 64 : 64 * 32 + 2 * 64 * 64));
}
__host__ __device__ constexpr int mlp_b_offset(int L) { return 64 * L; }

struct TrainArgs {
    const float* x;        // [n][32] features
    const float* t;        // [n][MLP_NOUT] targets
    uint32_t n;            // samples held
    uint32_t first;        // index of the batch's first sample in the epoch order
    uint32_t batch;        // samples in this step
    uint64_t perm_a, perm_b;
    const __bf16* packed;  // current weights, operand tiles
    const float* w;        // float32 master weights
    const float* bias;     // float32 biases
    float* gw;             // weight gradients (accumulated)
    float* gb;             // bias gradients
    __bf16* act0;          // [batch][32] layer-0 inputs
    __bf16* act[3];        // [batch][64] inputs of layers 1..3
    float* dz[4];          // [batch][64 or 16] dL/dz per layer
    float scale[MLP_NOUT]; // loss weight per output
    float inv_count;       // 1 / (batch * MLP_NOUT)
    float* loss;
};

struct SampleArgs {
    float* x;
    float* t;
    uint32_t* count;
    uint32_t capacity;
    uint32_t seed;
    float keep;            // fraction of hit pixels kept
};

struct MlpShadeArgs {
    const __bf16* packed;
    const float* bias;
    uint32_t* group_flags; // per 64-pixel group of the shade list: 0 network, 1 out of range (plain),
                           // 2 probe error (plain)
    uint32_t groups_cap;   // group_flags length
    float lo[32], hi[32];  // trained feature range
    float tol;             // probe relative error bound (<= 0: no probe)
    float floor;           // absolute floor of the probe's relative error
};

hipError_t launch_mlp_pack(hipStream_t s, const float* w, __bf16* packed);
hipError_t launch_mlp_forward(hipStream_t s, const float* x, uint32_t n, const __bf16* packed, const float* bias,
                              float* y);
hipError_t launch_mlp_train_step(hipStream_t s, const TrainArgs& a, float* w, float* b, float* m, float* v,
                                 float lr, int step);
hipError_t launch_mlp_grads(hipStream_t s, const TrainArgs& a);
hipError_t launch_mlp_samples(hipStream_t s, const ShadeArgs& a, const SampleArgs& sa, uint32_t width,
                              uint32_t height);
hipError_t launch_shade_mlp(hipStream_t s, const ShadeArgs& a, const MlpShadeArgs& m, const PixelList& L, int blocks);

}  // namespace hm

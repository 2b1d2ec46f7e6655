// Shapes, argument blocks and launchers of the learned shading term.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hm_kernels.h"

namespace hm {

constexpr int MLP_IN = 32;      // 29 features + 3 zero pads
constexpr int MLP_NFEAT = 29;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 16;     // 12 outputs + 4 zero pads
constexpr int MLP_NOUT = 12;
constexpr int MLP_TILES = 22;   // 16x32 bf16 operand tiles over the four layers
constexpr int MLP_NW = 64 * 32 + 64 * 64 + 64 * 64 + 16 * 64;  // 11264 weights
constexpr int MLP_NBIAS = 64 + 64 + 64 + 16;                    // 208 biases
constexpr float MLP_SIGMA_SCALE = 0.25f;

// Canonical float32 layout: W_L is (out x in) row-major, layers back to back.
__host__ __device__ constexpr int mlp_w_offset(int L) {
    return L == 0 ? 0 : (L == 1 ? 64 * 32 : (L == 2 ? 64 * 32 + 64 * 64 : 64 * 32 + 2 * 64 * 64));
}
__host__ __device__ constexpr int mlp_b_offset(int L) { return 64 * L; }

struct TrainArgs {
    const float* x;        // [n][32] features
    const float* t;        // [n][12] targets
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
    float inv_count;       // 1 / (batch * 12)
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
    uint32_t* tile_flags;  // per 16x4 tile: 0 network, 1 out of range, 2 probe error; bit 2 once re-shaded plain
    uint32_t tiles_x, tiles_y;
    float lo[32], hi[32];  // trained feature range
    float tol;             // probe relative error bound
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
hipError_t launch_shade_mlp(hipStream_t s, const ShadeArgs& a, const MlpShadeArgs& m, uint32_t width, uint32_t height);

}  // namespace hm

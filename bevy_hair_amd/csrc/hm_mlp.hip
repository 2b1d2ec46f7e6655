// Learned dual scattering: a 32 -> 64 -> 64 -> 64 -> 16 MLP on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16: bf16 operands, f32 accumulation)
// stands in, per pixel, for dual_scatter (shade.wesl:153-190), the 16-point
// azimuthal quadrature behind every light's back lobe and the key light's
// global multiple scattering and DOM forward-scattering spread
// (shade.wesl:220-242), which are then evaluated from the predicted terms.
// Single scattering stays analytic.
//
// Operand layout (one wave, 16 pixels per column block):
//   A (weights, 16 rows x 32 k): lane l holds row l%16, k = 8*(l/16) + j
//   B (activations, 32 k x 16 px): lane l holds pixel l%16, the same k
//   D (16 rows x 16 px): lane l holds rows 4*(l/16) + i, pixel l%16
// A layer's D feeds the next layer's B without moving data: lane group g
// packs rows {4g..4g+3} of two adjacent 16-row blocks, and the next layer's
// weights are stored with their k index permuted the same way.
#include <cstring>

#include "hm_device.h"
#include "hm_kernels.h"
#include "hm_lists.h"
#include "hm_mlp.h"
#include "hm_shade.h"

namespace hm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Tile t of layer L: tile_base(L) + r * kchunks(L) + c, 64 lanes x 8 bf16.
__device__ __host__ constexpr int tile_base(int L) { return L == 0 ? 0 : (L == 1 ? 4 : (L == 2 ? 12 : 20)); }
__device__ __host__ constexpr int kchunks(int L) { return L == 0 ? 1 : 2; }

// Logical input index of operand element (lane group g, slot j) in k-chunk c.
__device__ __host__ inline int k_of(int L, int c, int g, int j) {
    if (L == 0) return 32 * c + 8 * g + j;
    return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ w, __bf16* __restrict__ packed) {
    int t = blockIdx.x;  // tile
    int lane = threadIdx.x;
    if (t >= MLP_TILES || lane >= 64) return;
    int L = t < 4 ? 0 : (t < 12 ? 1 : (t < 20 ? 2 : 3));
    int local = t - tile_base(L);
    int r = local / kchunks(L), c = local % kchunks(L);
    int in = L == 0 ? MLP_IN : MLP_H;
    int row = 16 * r + (lane & 15), g = lane >> 4;
    for (int j = 0; j < 8; ++j) {
        int k = k_of(L, c, g, j);
        packed[((size_t)t * 64 + lane) * 8 + j] = (__bf16)w[mlp_w_offset(L) + row * in + k];
    }
}

// Forward pass of one 16-pixel column block. `T` points at the packed tiles
// (LDS or global), `bias` at the 208 biases. When `hid` is given, the packed
// hidden activations are kept for the backward pass.
__device__ __forceinline__ f32x4 mlp_fwd16(const bf16x8* T, const float* bias, bf16x8 b0, int lane,
                                           bf16x8* hid /* [6] or null */) {
    const int g = lane >> 4;
    f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 acc[4];
    bf16x8 B[2];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = mfma(T[(tile_base(0) + r) * 64 + lane], b0, z);
#pragma unroll
    for (int L = 1; L <= 3; ++L) {
        const float* bl = bias + 64 * (L - 1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            bf16x8 v;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float lo = acc[2 * c][i] + bl[32 * c + 4 * g + i];
                float hi = acc[2 * c + 1][i] + bl[32 * c + 16 + 4 * g + i];
                v[i] = (__bf16)(lo > 0.0f ? lo : 0.0f);
                v[4 + i] = (__bf16)(hi > 0.0f ? hi : 0.0f);
            }
            B[c] = v;
            if (hid) hid[2 * (L - 1) + c] = v;
        }
        if (L < 3) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[r] = mfma(T[(tile_base(L) + r * 2) * 64 + lane], B[0], z);
                acc[r] = mfma(T[(tile_base(L) + r * 2 + 1) * 64 + lane], B[1], acc[r]);
            }
        }
    }
    f32x4 y = mfma(T[(tile_base(3)) * 64 + lane], B[0], z);
    y = mfma(T[(tile_base(3) + 1) * 64 + lane], B[1], y);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] += bias[192 + 4 * g + i];
    return y;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

// Two activations rounded to bf16 together (one v_cvt_pk_bf16_f32), then
// ReLU on their bit patterns: a negative value or -0 has the sign bit set,
// so it is below zero as a 16-bit integer, and max(x, 0) over packed 16-bit
// integers gives +0 there and x elsewhere (one v_pk_max_i16). This equals
// rounding max(x, 0) to bf16 for every non-NaN x, since rounding keeps the sign.
__device__ __forceinline__ uint32_t relu_pk(float a, float b) {
    const bf16x2 v = __builtin_convertvector(f32x2{a, b}, bf16x2);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), i16x2{}));
}

// The next layer's operand from two 16-row output tiles (lane rows 4 g .. 4 g + 3
// of each): elements 0-3 from `lo`, 4-7 from `hi`.
__device__ __forceinline__ bf16x8 relu_operand(f32x4 lo, f32x4 hi) {
    const uint4 w = {relu_pk(lo[0], lo[1]), relu_pk(lo[2], lo[3]), relu_pk(hi[0], hi[1]), relu_pk(hi[2], hi[3])};
    return __builtin_bit_cast(bf16x8, w);
}

// Forward pass of NB 16-pixel column blocks at once, layer by layer: each
// weight tile is read from `T` once per layer and feeds NB independent
// matrix-core chains; the outputs of block cb are returned in y[cb].
template <int NB>
__device__ __forceinline__ void mlp_fwd_blocks(const bf16x8* T, const float* bias, const bf16x8 b0[NB], int lane,
                                               f32x4 y[NB]) {
    const int g = lane >> 4;
    // The biases enter as the accumulator input of each layer's first
    // product: lane rows 16 r + 4 g .. + 3 of output tile r.
    auto bias_tiles = [&](int layer, f32x4 bt[4]) {
        const float4* bl = reinterpret_cast<const float4*>(bias + 64 * layer);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4 v = bl[4 * r + g];
            bt[r] = f32x4{v.x, v.y, v.z, v.w};
        }
    };
    f32x4 acc[NB][4], bt[4];
    bias_tiles(0, bt);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const bf16x8 w = T[(tile_base(0) + r) * 64 + lane];
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) acc[cb][r] = mfma(w, b0[cb], bt[r]);
    }
    bf16x8 B[NB][2];
#pragma unroll
    for (int L = 1; L <= 3; ++L) {
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
#pragma unroll
            for (int c = 0; c < 2; ++c) B[cb][c] = relu_operand(acc[cb][2 * c], acc[cb][2 * c + 1]);
        }
        if (L < 3) {
            bias_tiles(L, bt);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bf16x8 w0 = T[(tile_base(L) + r * 2) * 64 + lane];
                const bf16x8 w1 = T[(tile_base(L) + r * 2 + 1) * 64 + lane];
#pragma unroll
                for (int cb = 0; cb < NB; ++cb) acc[cb][r] = mfma(w0, B[cb][0], bt[r]);
#pragma unroll
                for (int cb = 0; cb < NB; ++cb) acc[cb][r] = mfma(w1, B[cb][1], acc[cb][r]);
            }
        }
    }
    const bf16x8 w0 = T[tile_base(3) * 64 + lane];
    const bf16x8 w1 = T[(tile_base(3) + 1) * 64 + lane];
    const float4 bo4 = reinterpret_cast<const float4*>(bias + 192)[g];
    const f32x4 bo = {bo4.x, bo4.y, bo4.z, bo4.w};
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
        f32x4 v = mfma(w0, B[cb][0], bo);
        y[cb] = mfma(w1, B[cb][1], v);
    }
}

__device__ __forceinline__ void mlp_fwd64(const bf16x8* T, const float* bias, const bf16x8 b0[4], int lane,
                                          f32x4 y[4]) {
    mlp_fwd_blocks<4>(T, bias, b0, lane, y);
}

__device__ __forceinline__ bf16x8 load_x8(const float* row8) {
    float4 a = *reinterpret_cast<const float4*>(row8);
    float4 b = *reinterpret_cast<const float4*>(row8 + 4);
    bf16x8 v;
    v[0] = (__bf16)a.x; v[1] = (__bf16)a.y; v[2] = (__bf16)a.z; v[3] = (__bf16)a.w;
    v[4] = (__bf16)b.x; v[5] = (__bf16)b.y; v[6] = (__bf16)b.z; v[7] = (__bf16)b.w;
    return v;
}

// Batch inference: Y[n][16] = mlp(X[n][32]), 64 samples per wave step
// through the same four-block pass as the shading kernel.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint32_t n, const __bf16* packed,
                                                     const float* __restrict__ bias, float* __restrict__ Y) {
    const int lane = threadIdx.x & 63, g = lane >> 4, q = lane & 15;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    const bf16x8* T = reinterpret_cast<const bf16x8*>(packed);
    for (uint32_t base = wave * 64u; base < n; base += nw * 64u) {
        bf16x8 b0[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const uint32_t s = base + 16u * cb + (uint32_t)q;
            b0[cb] = bf16x8{};
            if (s < n) b0[cb] = load_x8(X + (size_t)s * 32 + 8 * g);
        }
        f32x4 y[4];
        mlp_fwd64(T, bias, b0, lane, y);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const uint32_t s = base + 16u * cb + (uint32_t)q;
            if (s < n) *reinterpret_cast<float4*>(Y + (size_t)s * 16 + 4 * g) = make_float4(y[cb][0], y[cb][1], y[cb][2], y[cb][3]);
        }
    }
}

// ---- training ---------------------------------------------------------------------

// Sample index k of the batch starting at `base` under this epoch's
// permutation i -> (a * i + b) mod n (a coprime to n).
__device__ __forceinline__ uint32_t sample_index(uint64_t i, const TrainArgs& a) {
    return (uint32_t)((a.perm_a * i + a.perm_b) % a.n);
}

// Forward over the batch, storing each layer's input activations (bf16,
// row-major) and the output error dL/dy; accumulates the loss.
__global__ __launch_bounds__(256) void k_mlp_train_fwd(TrainArgs a) {
    const int lane = threadIdx.x & 63, g = lane >> 4, q = lane & 15;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    const bf16x8* T = reinterpret_cast<const bf16x8*>(a.packed);
    float loss = 0.0f;
    for (uint32_t base = wave * 16u; base < a.batch; base += nw * 16u) {
        uint32_t k = base + (uint32_t)q;
        bool ok = k < a.batch;
        uint32_t s = ok ? sample_index((uint64_t)a.first + k, a) : 0u;
        bf16x8 b0 = {};
        if (ok) b0 = load_x8(a.x + (size_t)s * 32 + 8 * g);
        bf16x8 hid[6];
        f32x4 y = mlp_fwd16(T, a.bias, b0, lane, hid);
        if (!ok) continue;
        // layer inputs in natural neuron order: X0 [k][32], H1..H3 [k][64]
        *reinterpret_cast<bf16x8*>(a.act0 + (size_t)k * 32 + 8 * g) = b0;
#pragma unroll
        for (int L = 0; L < 3; ++L) {
            __bf16* row = a.act[L] + (size_t)k * 64;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                bf16x8 v = hid[2 * L + c];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    row[32 * c + 4 * g + i] = v[i];
                    row[32 * c + 16 + 4 * g + i] = v[4 + i];
                }
            }
        }
        float dz[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int o = 4 * g + i;
            if (o < MLP_NOUT) {
                float sc = a.scale[o];
                float d = (y[i] - a.t[(size_t)s * MLP_NOUT + o]) * sc;
                loss += d * d;
                dz[i] = 2.0f * d * sc * a.inv_count;
            } else {
                dz[i] = 0.0f;
            }
        }
        *reinterpret_cast<float4*>(a.dz[3] + (size_t)k * 16 + 4 * g) = make_float4(dz[0], dz[1], dz[2], dz[3]);
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0 && loss != 0.0f) atomicAdd(a.loss, loss * a.inv_count);
}

// Backward through layer L (3 .. 0) for 32-sample chunks:
//   dW_L += dZ_L^T . H_L, db_L += sum dZ_L, and for L > 0
//   dZ_{L-1} = (dZ_L . W_L) * [H_L > 0]  (H_L is layer L's input activation).
template <int L>
__global__ __launch_bounds__(256) void k_mlp_train_bwd(TrainArgs a) {
    constexpr int O = L == 3 ? 16 : 64;       // layer outputs
    constexpr int K = L == 0 ? 32 : 64;       // layer inputs
    constexpr int RB = O / 16, NB = K / 16;   // dW tiles
    const int lane = threadIdx.x & 63, g = lane >> 4, q = lane & 15;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    const float* dZ = a.dz[L];
    const __bf16* H = L == 0 ? a.act0 : a.act[L - 1];
    const float* W = a.w + mlp_w_offset(L);
    const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 acc[RB][NB];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[r][n] = z;
    float db = 0.0f;
    for (uint32_t base = wave * 32u; base < a.batch; base += nw * 32u) {
        // dW: A = dZ^T (row = output, k = sample), B = H (k = sample, col = input)
        bf16x8 Ad[RB], Bh[NB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t s = base + 8u * g + j;
                Ad[r][j] = (__bf16)(s < a.batch ? dZ[(size_t)s * O + 16 * r + q] : 0.0f);
            }
        }
#pragma unroll
        for (int n = 0; n < NB; ++n) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t s = base + 8u * g + j;
                Bh[n][j] = s < a.batch ? H[(size_t)s * K + 16 * n + q] : (__bf16)0.0f;
            }
        }
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int n = 0; n < NB; ++n) acc[r][n] = mfma(Ad[r], Bh[n], acc[r][n]);
        if (lane < O) {
            for (int s = 0; s < 32; ++s)
                if (base + s < a.batch) db += dZ[(size_t)(base + s) * O + lane];
        }
        if (L > 0) {
            // dH^T (input x sample) = W^T (input x output) . dZ^T (output x sample)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                uint32_t s = base + 16u * cb + (uint32_t)q;
                bool ok = s < a.batch;
#pragma unroll
                for (int r = 0; r < K / 16; ++r) {
                    f32x4 d = z;
#pragma unroll
                    for (int c = 0; c < (O + 31) / 32; ++c) {
                        bf16x8 Aw, Bd;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            int o = 32 * c + 8 * g + j;
                            bool in = o < O;
                            Aw[j] = (__bf16)(in ? W[o * K + 16 * r + q] : 0.0f);
                            Bd[j] = (__bf16)(in && ok ? dZ[(size_t)s * O + o] : 0.0f);
                        }
                        d = mfma(Aw, Bd, d);
                    }
                    if (ok) {
                        float out[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            int in_n = 16 * r + 4 * g + i;
                            float h = (float)H[(size_t)s * K + in_n];
                            out[i] = h > 0.0f ? d[i] : 0.0f;
                        }
                        *reinterpret_cast<float4*>(a.dz[L - 1] + (size_t)s * 64 + 16 * r + 4 * g) =
                            make_float4(out[0], out[1], out[2], out[3]);
                    }
                }
            }
        }
    }
    float* gW = a.gw + mlp_w_offset(L);
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[r][n][i];
                if (v != 0.0f) atomicAdd(gW + (16 * r + 4 * g + i) * K + 16 * n + q, v);
            }
    if (lane < O && db != 0.0f) atomicAdd(a.gb + mlp_b_offset(L) + lane, db);
}

// Adam on the float32 master copy (weights then biases), zeroing the gradient.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, float* __restrict__ grad,
                                                  float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr,
                                                  float b1, float b2, float eps, float c1, float c2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gr = grad[i];
    grad[i] = 0.0f;
    float mi = b1 * m[i] + (1.0f - b1) * gr;
    float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Keeps the zero padding of the weights at zero (unused inputs of layer 0,
// padded outputs of layer 3) after an update.
__global__ void k_mlp_mask(float* __restrict__ w, float* __restrict__ b) {
    uint32_t i = threadIdx.x;
    if constexpr (MLP_IN > MLP_NFEAT) {
        if (i < 64 * (MLP_IN - MLP_NFEAT)) {
            uint32_t row = i / (MLP_IN - MLP_NFEAT), col = MLP_NFEAT + i % (MLP_IN - MLP_NFEAT);
            w[mlp_w_offset(0) + row * MLP_IN + col] = 0.0f;
        }
    }
    if (i < (MLP_OUT - MLP_NOUT) * MLP_H) w[mlp_w_offset(3) + MLP_NOUT * MLP_H + i] = 0.0f;
    if (i < MLP_OUT - MLP_NOUT) b[mlp_b_offset(3) + MLP_NOUT + i] = 0.0f;
}

// ---- features and targets from the plain path ----------------------------------------

// The DOM inputs of one pixel from dom_visibility's four taps
// (shade.wesl:279-310): returns that visibility, bit for bit, and fills
// `layers` with the transmittance through layers 0..k of the bilinearly
// weighted layer counts (k = 0..15). The counts stream in four groups of four
// slices (one 16-byte load per tap and group), each group continuing every
// tap's optical depth and the weighted layer sums, so that no more than two
// groups' counts are held at once.
__device__ float dom_features(const View& p, const uint32_t* __restrict__ dom, F3 world, float beta,
                              float layers[16]) {
    for (int k = 0; k < 16; ++k) layers[k] = 1.0f;
    DomPos P;
    if (!dom_pos(p, world, beta, P)) return 1.0f;
    const uint4* counts = reinterpret_cast<const uint4*>(dom + (size_t)to_u32(p.dom_info.x) * to_u32(p.dom_info.y));
    bool on[4];
    float t[4], occ[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t f = dom[P.pix[c]];
        on[c] = P.inside[c] && f != HM_DOM_EMPTY;
        t[c] = on[c] ? dom_tap_t(p, f, P.dist) : 0.0f;
        occ[c] = 0.0f;
    }
    // bilinear weights, zero for a tap outside the map or with an empty front
    float wts[4] = {(1.0f - P.tx) * (1.0f - P.ty), P.tx * (1.0f - P.ty), (1.0f - P.tx) * P.ty, P.tx * P.ty};
#pragma unroll
    for (int c = 0; c < 4; ++c) wts[c] = on[c] ? wts[c] : 0.0f;
    float pre = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint4 v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = counts[(size_t)P.pix[c] * (HM_DOM_LAYERS / 4) + q];
#pragma unroll
        for (int c = 0; c < 4; ++c) occ[c] = dom_occ4(occ[c], v[c], t[c], 4 * q);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float cw = 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t cs[4] = {v[c].x, v[c].y, v[c].z, v[c].w};
                cw += wts[c] * u2f(cs[k]);
            }
            pre += cw;
            layers[4 * q + k] = hexp(-pre * p.filter_params.w);
        }
    }
    float vis[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) vis[c] = on[c] ? hexp(-occ[c] * p.filter_params.w) : 1.0f;
    return mixs(mixs(vis[0], vis[1], P.tx), mixs(vis[2], vis[3], P.tx), P.ty);
}

// Features of one pixel (order and scaling in hm_mlp.h); returns the key
// light's optical depth -log(visibility), as the plain method forms it.
__device__ float mlp_features(const View& p, const PixelShade& px, F3 key_dir, const uint32_t* dom, float f[32]) {
    f[0] = px.tangent.x; f[1] = px.tangent.y; f[2] = px.tangent.z;
    f[3] = px.view.x; f[4] = px.view.y; f[5] = px.view.z;
    F3 l = safe_normalize(key_dir);
    f[6] = l.x; f[7] = l.y; f[8] = l.z;
    const float vis = dom_features(p, dom, px.world, px.beta, f + 9);
    const float depth = -hlog(mx(vis, 1e-6f));
    f[25] = px.sigma_a.x * MLP_SIGMA_SCALE;
    f[26] = px.sigma_a.y * MLP_SIGMA_SCALE;
    f[27] = px.sigma_a.z * MLP_SIGMA_SCALE;
    f[28] = px.roughness;
    const float sin_i = clampf(dot3(px.t_n, l), -1.0f, 1.0f);
    const float sin_o = clampf(dot3(px.t_n, px.v_n), -1.0f, 1.0f);
    f[29] = (fasin(sin_i) + fasin(sin_o)) * (1.0f / kPi);  // theta_h / pi
    f[30] = px.cos_theta_o;
    f[31] = depth * MLP_DEPTH_SCALE;
    return depth;
}

// The plain method's values of the outputs (hm_mlp.h) for one pixel.
__device__ void mlp_targets(const DualScatter& ds, float t[MLP_NOUT]) {
    t[0] = ds.fwd.r.x;
    t[1] = ds.fwd.tt.x; t[2] = ds.fwd.tt.y; t[3] = ds.fwd.tt.z;
    t[4] = ds.fwd.trt.x; t[5] = ds.fwd.trt.y; t[6] = ds.fwd.trt.z;
    t[7] = ds.a_back.x; t[8] = ds.a_back.y; t[9] = ds.a_back.z;
    t[10] = ds.delta_b.x; t[11] = ds.delta_b.y; t[12] = ds.delta_b.z;
    t[13] = ds.sigma_b2.x; t[14] = ds.sigma_b2.y; t[15] = ds.sigma_b2.z;
    for (int k = 0; k < MLP_NOUT; ++k)
        if (mlp_log_output(k)) t[k] = logf(mx(t[k], 0.0f) + MLP_LOG_EPS);
}

// The dual-scattering terms from the decoded outputs; a_f and beta_f^2 follow
// from the forward lobes as in dual_scatter.
__device__ __forceinline__ DualScatter network_dual_scatter(const float o[MLP_NOUT], const Shape& shape) {
    DualScatter ds;
    ds.fwd = {splat3(o[0]), {o[1], o[2], o[3]}, {o[4], o[5], o[6]}};
    ds.a_f = vmin(ds.fwd.r + ds.fwd.tt + ds.fwd.trt, splat3(0.95f));
    ds.beta_f2 = lobe_mean(ds.fwd, ds.a_f, shape.beta * shape.beta);
    ds.a_back = {o[7], o[8], o[9]};
    ds.delta_b = {o[10], o[11], o[12]};
    ds.sigma_b2 = vmax(F3{o[13], o[14], o[15]}, splat3(1e-6f));
    return ds;
}

// Appends (features, targets) of the hit pixels kept by a hash test.
__global__ __launch_bounds__(256) void k_mlp_samples(ShadeArgs a, SampleArgs s) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    uint32_t width = to_u32(p.screen.x), height = to_u32(p.screen.y);
    if (x >= width || y >= height) return;
    size_t pix = (size_t)y * width + x;
    uint64_t word = a.center[pix];
    int key = key_light(a);
    if (gb_empty(word) || key < 0) return;
    if ((float)pcg((uint32_t)pix ^ s.seed) * (1.0f / 4294967296.0f) >= s.keep) return;
    PixelShade px = pixel_setup(p, word, a.beta[pix], x, y);
    DualScatter ds = dual_scatter(px.cos_theta_o, px.sigma_a, px.shape);
    F3 kd = dir_of(a.directionals[key]);
    float f[32], t[MLP_NOUT];
    (void)mlp_features(p, px, kd, a.dom, f);
    mlp_targets(ds, t);
    uint32_t slot = atomicAdd(s.count, 1u);
    if (slot >= s.capacity) return;
    for (int k = 0; k < 32; ++k) s.x[(size_t)slot * 32 + k] = f[k];
    for (int k = 0; k < MLP_NOUT; ++k) s.t[(size_t)slot * MLP_NOUT + k] = t[k];
}

// ---- MLP shading ---------------------------------------------------------------------

// Sum of v over the 64 lanes of the wave, in every lane.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sum of v over the aligned group of 8 lanes that holds this lane.
__device__ __forceinline__ float sum8(float v) {
#pragma unroll
    for (int m = 4; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ F3 sum8_3(F3 v) { return {sum8(v.x), sum8(v.y), sum8(v.z)}; }
__device__ __forceinline__ F3 lane_read3(F3 v, int lane) {
    return {__shfl(v.x, lane, 64), __shfl(v.y, lane, 64), __shfl(v.z, lane, 64)};
}

// pixel_color of one pixel (the same in every lane) with two sets of terms,
// `net` and `ref`, evaluated by the whole wave: lane i takes the point lights
// i, i + 64, ..., the directional and rect lights likewise; only the
// summation order differs from pixel_color. The geometry and single
// scattering are shared by the two colours.
template <bool EXTRAS>
__device__ void color_pair_by_wave(const ShadeArgs& a, const PixelShade& px, const LocalBack& net, F3 net_ms,
                                   const LocalBack& ref, F3 ref_ms, int key, float depth, int lane, F3& got,
                                   F3& want) {
    F3 cn = splat3(0.0f), cr = splat3(0.0f);
    for (uint32_t i = (uint32_t)lane; i < a.point_count; i += 64u) {
        const hm_point_light& L = a.points[i];
        F3 ltf = point_pos(L) - px.world;
        float att = distance_attenuation(dot3(ltf, ltf), L.color_inverse_square_range[3]);
        if (att == 0.0f) continue;
        BsdfGeom g = bsdf_geom_u(px.vg, safe_normalize_unit(safe_normalize_mid(ltf)));
        F3 single = single_lobes(g, px);
        F3 w = point_color(L) * att;
        if (L.kind != HM_LIGHT_POINT) w = w * spot_mask(L, px.world);
        cn = cn + (single + back_lobe(g.theta_h, net)) * w;
        cr = cr + (single + back_lobe(g.theta_h, ref)) * w;
    }
    for (uint32_t i = (uint32_t)lane; i < a.directional_count; i += 64u) {
        const hm_directional_light& L = a.directionals[i];
        BsdfGeom g = bsdf_geom_u(px.vg, safe_normalize_unit(dir_of(L)));
        F3 single = single_lobes(g, px);
        F3 w = F3{L.color[0], L.color[1], L.color[2]};
        if ((int)i == key) {
            single = hexp(-depth) * single;
            cn = cn + (single + net_ms) * w;
            cr = cr + (single + ref_ms) * w;
        } else {
            cn = cn + (single + back_lobe(g.theta_h, net)) * w;
            cr = cr + (single + back_lobe(g.theta_h, ref)) * w;
        }
    }
    F3 env = splat3(0.0f);
    if (EXTRAS) {
        for (uint32_t i = (uint32_t)lane; i < a.rect_count; i += 64u) {
            const hm_rect_light& L = a.rects[i];
            cn = cn + rect_term(L, px, net);
            cr = cr + rect_term(L, px, ref);
        }
        // the probe does not depend on the dual-scattering terms
        if (a.env.on) env = probe_radiance(a.env, px) * px.ao;
    }
    got = (F3{wave_sum(cn.x), wave_sum(cn.y), wave_sum(cn.z)} + env) * a.exposure;
    want = (F3{wave_sum(cr.x), wave_sum(cr.y), wave_sum(cr.z)} + env) * a.exposure;
}

// dual_scatter of one pixel (the same in every lane) by the whole wave:
// lanes 0-15 take its 16 azimuthal samples.
__device__ DualScatter dual_scatter_by_wave(const PixelShade& px, int lane) {
    Lobes mine = {splat3(0.0f), splat3(0.0f), splat3(0.0f)};
    if (lane < 16) mine = dual_sample((uint32_t)lane, px.cos_theta_o, px.sigma_a);
    mine = {sum8_3(mine.r), sum8_3(mine.tt), sum8_3(mine.trt)};
    const Lobes back = {lane_read3(mine.r, 0), lane_read3(mine.tt, 0), lane_read3(mine.trt, 0)};
    const Lobes fwd = {lane_read3(mine.r, 8), lane_read3(mine.tt, 8), lane_read3(mine.trt, 8)};
    return dual_finish(fwd, back, px.shape);
}

// Network shading, one wave per group of 64 consecutive covered pixels, one
// lane per pixel. Per group: the pixel set-up and features -> LDS (bf16) ->
// MFMA forward (two 16-pixel column blocks at a time, in place) -> LDS ->
// dual-scattering terms -> the key light's multiple scattering -> colour
// (pixel_color, the plain method's light loop, with the network's terms in
// place of dual_scatter's). A pixel with a feature outside the trained range
// takes the plain method's terms instead, formed by its own lane with
// dual_scatter (per pixel and by the same code, so its colour does not depend
// on the order of the covered-pixel list or on its group's other pixels). When
// `tol` > 0, every PROBE_STRIDE-th range-clean group also shades one probe
// pixel by the plain method (all 64 lanes on that pixel) and the group is
// re-shaded by the plain method when the network colour misses it by more
// than `tol`; which pixels share a group follows the list order, so that
// re-shading does not.
constexpr uint32_t PROBE_STRIDE = 4;

// F: SHADE_* variant bits (pixel_color); PROBE: the error-bound probe (tol > 0).
template <uint32_t F, bool PROBE>
__global__ __launch_bounds__(256, 3) void k_shade_mlp(ShadeArgs a, MlpShadeArgs m, PixelList L) {
    constexpr bool CULL = (F & SHADE_CULL) != 0u;
    __shared__ bf16x8 s_tiles[MLP_TILES * 64];
    __shared__ __align__(16) float s_bias[MLP_NBIAS];
    __shared__ float4 s_io[4][64 * 4];  // per wave: 64 x 32 bf16 features, then 64 x 16 outputs
    __shared__ uint32_t pre[LIST_SUBS + 1];
    list_prefix(L, pre);  // synchronizes the block
    const uint32_t n = pre[LIST_SUBS], groups = (n + 63u) / 64u;
    if (blockIdx.x * 4u >= groups) return;  // the grid is sized for the list's capacity
    {
        const uint4* src = reinterpret_cast<const uint4*>(m.packed);
        uint4* dst = reinterpret_cast<uint4*>(s_tiles);
        for (int i = threadIdx.x; i < MLP_TILES * 64; i += 256) dst[i] = src[i];
        for (int i = threadIdx.x; i < MLP_NBIAS; i += 256) s_bias[i] = m.bias[i];
    }
    __syncthreads();
    const View& p = *reinterpret_cast<const View*>(&a.view);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, q = lane & 15;
    const uint32_t width = to_u32(p.screen.x), pixels = width * to_u32(p.screen.y);
    const int key = key_light(a);
    const F3 kd = key >= 0 ? dir_of(a.directionals[key]) : F3{0.0f, 1.0f, 0.0f};
    float4* io = s_io[wv];
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (uint32_t grp = blockIdx.x * 4u + (uint32_t)wv; grp < groups && grp < m.groups_cap;
         grp += gridDim.x * 4u) {
        const uint32_t i = grp * 64u + (uint32_t)lane;
        const uint32_t pix = i < n ? list_entry(L, pre, i) : pixels;
        const bool valid = pix < pixels;
        float depth = 0.0f;
        bool outside = false;
        PixelShade px;  // kept across the network for the light loop
        px.world = splat3(0.0f);
        {
            float f[32];
            for (int k = 0; k < 32; ++k) f[k] = 0.0f;
            if (valid) {
                px = pixel_setup(p, a.center[pix], a.beta[pix], pix % width, pix / width);
                depth = mlp_features(p, px, kd, a.dom, f);
                // outside the trained box when some feature differs from
                // min(max(f, lo), hi) (always, for an empty box lo > hi; a NaN
                // feature differs too)
                float off_box = 0.0f;
                for (int k = 9; k < MLP_NFEAT; ++k) off_box += fabsf(f[k] - fminf(fmaxf(f[k], m.lo[k]), m.hi[k]));
                outside = !(off_box == 0.0f);
            }
            // row `lane`: 32 bf16 = four 16-byte vectors
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                bf16x8 v;
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = (__bf16)f[8 * c + k];
                io[lane * 4 + c] = __builtin_bit_cast(float4, v);
            }
        }
        // flag (per group, for the count): 1 when a pixel is outside the
        // trained range; that pixel alone takes the plain method
        const uint64_t out_lanes = __ballot(outside && valid);
        uint32_t flag = out_lanes ? 1u : 0u;
        float o[MLP_NOUT] = {};
        if (!__all(outside || !valid)) {
            wave_sync();
            // Block cb is pixels 16 cb .. 16 cb + 15: its features are read
            // from those rows and its outputs written back over them, two
            // blocks at a time (two blocks' network state in registers).
#pragma unroll 1
            for (int cb = 0; cb < 4; cb += 2) {
                const bf16x8 b0[2] = {__builtin_bit_cast(bf16x8, io[(cb * 16 + q) * 4 + g]),
                                      __builtin_bit_cast(bf16x8, io[((cb + 1) * 16 + q) * 4 + g])};
                // An offset the compiler cannot see through: the weight and
                // bias reads stay inside the loop instead of being hoisted
                // out of it into registers held across both passes.
                int off = 0;
                asm volatile("" : "+v"(off));
                f32x4 y[2];
                mlp_fwd_blocks<2>(s_tiles + off, s_bias + off, b0, lane, y);
                io[(cb * 16 + q) * 4 + g] = make_float4(y[0][0], y[0][1], y[0][2], y[0][3]);
                io[((cb + 1) * 16 + q) * 4 + g] = make_float4(y[1][0], y[1][1], y[1][2], y[1][3]);
            }
            wave_sync();
            const float* y = reinterpret_cast<const float*>(io + lane * 4);
#pragma unroll
            for (int k = 0; k < MLP_NOUT; ++k) o[k] = mlp_log_output(k) ? mx(hexp(y[k]) - MLP_LOG_EPS, 0.0f) : y[k];
        }
        if (PROBE && flag == 0u && grp % PROBE_STRIDE == 0u) {
            // probe: one pixel of the group, the same in every lane, shaded
            // with the network's terms and with the plain method's
            const uint32_t in_group = n - grp * 64u < 64u ? n - grp * 64u : 64u;
            const int r = (int)(pcg(grp) % in_group);
            const uint32_t ppix = __shfl(pix, r, 64);
            if (ppix < pixels) {
                const PixelShade pp = pixel_setup(p, a.center[ppix], a.beta[ppix], ppix % width, ppix / width);
                const float pd = __shfl(depth, r, 64);
                float po[MLP_NOUT];
                for (int k = 0; k < MLP_NOUT; ++k) po[k] = __shfl(o[k], r, 64);
                const DualScatter dn = network_dual_scatter(po, pp.shape);
                const DualScatter ds = dual_scatter_by_wave(pp, lane);
                const float th = key >= 0 ? key_theta_h(pp, kd) : 0.0f;
                const F3 net_ms = key >= 0 ? key_multiple_scattering(th, pd, pp.shape, dn) : splat3(0.0f);
                const F3 key_ms = key >= 0 ? key_multiple_scattering(th, pd, pp.shape, ds) : splat3(0.0f);
                F3 got, ref;
                color_pair_by_wave<(F & SHADE_EXTRAS) != 0u>(a, pp, local_terms(dn), net_ms, local_terms(ds), key_ms,
                                                             key, pd, lane, got, ref);
                const float err = fabsf(got.x - ref.x) + fabsf(got.y - ref.y) + fabsf(got.z - ref.z);
                const float mag = fabsf(ref.x) + fabsf(ref.y) + fabsf(ref.z);
                if (err > m.tol * mx(mag, m.floor)) flag = 2u;
            }
        }
        // The dual-scattering terms: the network's, or the plain method's for
        // the group (probe miss) or for the lanes outside the trained range.
        DualScatter ds;
        if (valid && (flag == 2u || outside)) ds = dual_scatter(px.cos_theta_o, px.sigma_a, px.shape);
        else if (valid) ds = network_dual_scatter(o, px.shape);
        LocalBack lb = {splat3(0.0f), splat3(0.0f), splat3(0.0f)};
        F3 key_ms = splat3(0.0f);
        if (valid) {
            lb = local_terms(ds);
            if (key >= 0) key_ms = key_multiple_scattering(key_theta_h(px, kd), depth, px.shape, ds);
        }
        F3 c = splat3(0.0f);
        if (CULL) c = pixel_color<F>(a, px, valid, lb, key, depth, key_ms);  // the whole wave calls
        else if (valid) c = pixel_color<F>(a, px, true, lb, key, depth, key_ms);
        if (valid) {
            a.shaded[pix] = make_float4(c.x, c.y, c.z, 1.0f);
            a.shaded_depth[pix] = px.view_dist;
        }
        if (lane == 0) m.group_flags[grp] = flag;
        wave_sync();  // the LDS rows are rewritten by the next group
    }
}

// ---- launchers --------------------------------------------------------------------------

hipError_t launch_mlp_pack(hipStream_t s, const float* w, __bf16* packed) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_TILES), dim3(64), 0, s, w, packed);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(hipStream_t s, const float* x, uint32_t n, const __bf16* packed, const float* bias,
                              float* y) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, s, x, n, packed, bias, y);
    return hipGetLastError();
}

hipError_t launch_mlp_grads(hipStream_t s, const TrainArgs& a) {
    if (a.batch == 0) return hipSuccess;
    uint32_t fb = (a.batch + 63) / 64;
    hipLaunchKernelGGL(k_mlp_train_fwd, dim3(fb < 4096 ? fb : 4096), dim3(256), 0, s, a);
    uint32_t bb = (a.batch + 511) / 512;  // each wave takes >= 4 chunks of 32 samples
    if (bb > 256) bb = 256;
    if (bb == 0) bb = 1;
    hipLaunchKernelGGL(k_mlp_train_bwd<3>, dim3(bb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mlp_train_bwd<2>, dim3(bb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mlp_train_bwd<1>, dim3(bb), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mlp_train_bwd<0>, dim3(bb), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_train_step(hipStream_t s, const TrainArgs& a, float* w, float* b, float* m, float* v,
                                 float lr, int step) {
    hipError_t e = launch_mlp_grads(s, a);
    if (e != hipSuccess) return e;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NW + 255) / 256), dim3(256), 0, s, w, a.gw, m, v, (uint32_t)MLP_NW, lr,
                       b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_mlp_adam, dim3(1), dim3(256), 0, s, b, a.gb, m + MLP_NW, v + MLP_NW, (uint32_t)MLP_NBIAS,
                       lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_mlp_mask, dim3(1), dim3(256), 0, s, w, b);
    hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_TILES), dim3(64), 0, s, w, const_cast<__bf16*>(a.packed));
    return hipGetLastError();
}

hipError_t launch_mlp_samples(hipStream_t s, const ShadeArgs& a, const SampleArgs& sa, uint32_t width,
                              uint32_t height) {
    if (width == 0 || height == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_samples, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, a, sa);
    return hipGetLastError();
}

hipError_t launch_shade_mlp(hipStream_t s, const ShadeArgs& a, const MlpShadeArgs& m, const PixelList& L,
                            int blocks) {
    if (m.groups_cap == 0) return hipSuccess;
    const uint32_t want = (m.groups_cap + 3u) / 4u;
    const dim3 grid(want < (uint32_t)blocks ? want : (uint32_t)blocks);
    const bool probe = m.tol > 0.0f;
    switch (shade_variant(a)) {
        case 0u:
            if (probe) hipLaunchKernelGGL((k_shade_mlp<0u, true>), grid, dim3(256), 0, s, a, m, L);
            else hipLaunchKernelGGL((k_shade_mlp<0u, false>), grid, dim3(256), 0, s, a, m, L);
            break;
        case SHADE_CULL:
            if (probe) hipLaunchKernelGGL((k_shade_mlp<SHADE_CULL, true>), grid, dim3(256), 0, s, a, m, L);
            else hipLaunchKernelGGL((k_shade_mlp<SHADE_CULL, false>), grid, dim3(256), 0, s, a, m, L);
            break;
        case SHADE_EXTRAS:
            if (probe) hipLaunchKernelGGL((k_shade_mlp<SHADE_EXTRAS, true>), grid, dim3(256), 0, s, a, m, L);
            else hipLaunchKernelGGL((k_shade_mlp<SHADE_EXTRAS, false>), grid, dim3(256), 0, s, a, m, L);
            break;
        default:
            if (probe) hipLaunchKernelGGL((k_shade_mlp<SHADE_CULL | SHADE_EXTRAS, true>), grid, dim3(256), 0, s, a, m, L);
            else hipLaunchKernelGGL((k_shade_mlp<SHADE_CULL | SHADE_EXTRAS, false>), grid, dim3(256), 0, s, a, m, L);
            break;
    }
    return hipGetLastError();
}

}  // namespace hm

// Learned multiple-scattering term: a 29 -> 64 -> 64 -> 64 -> 12 MLP on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: bf16 operands, f32
// accumulation). It stands in, per pixel, for dual_scatter (shade.wesl:153-190)
// and for the key light's multiple-scattering part of hair_bsdf
// (shade.wesl:220-242): outputs 0-2 are that key-light term per unit light
// colour, 3-11 the dual-scattering local terms a_back, delta_b, sigma_b^2
// that the other lights' back lobe needs. Single scattering stays analytic.
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

__device__ __forceinline__ bf16x8 load_x8(const float* row8) {
    float4 a = *reinterpret_cast<const float4*>(row8);
    float4 b = *reinterpret_cast<const float4*>(row8 + 4);
    bf16x8 v;
    v[0] = (__bf16)a.x; v[1] = (__bf16)a.y; v[2] = (__bf16)a.z; v[3] = (__bf16)a.w;
    v[4] = (__bf16)b.x; v[5] = (__bf16)b.y; v[6] = (__bf16)b.z; v[7] = (__bf16)b.w;
    return v;
}

// Batch inference: Y[n][16] = mlp(X[n][32]).
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint32_t n, const __bf16* packed,
                                                     const float* __restrict__ bias, float* __restrict__ Y) {
    const int lane = threadIdx.x & 63, g = lane >> 4, q = lane & 15;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * 4u;
    const bf16x8* T = reinterpret_cast<const bf16x8*>(packed);
    for (uint32_t base = wave * 16u; base < n; base += nw * 16u) {
        uint32_t s = base + (uint32_t)q;
        bf16x8 b0 = {};
        if (s < n) b0 = load_x8(X + (size_t)s * 32 + 8 * g);
        f32x4 y = mlp_fwd16(T, bias, b0, lane, nullptr);
        if (s < n) {
            float4 o = make_float4(y[0], y[1], y[2], y[3]);
            *reinterpret_cast<float4*>(Y + (size_t)s * 16 + 4 * g) = o;
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

// Keeps the zero padding of the weights at zero (inputs 29..31 of layer 0,
// outputs 12..15 of layer 3) after an update.
__global__ void k_mlp_mask(float* __restrict__ w, float* __restrict__ b) {
    uint32_t i = threadIdx.x;
    if (i < 64 * (MLP_IN - MLP_NFEAT)) {
        uint32_t row = i / (MLP_IN - MLP_NFEAT), col = MLP_NFEAT + i % (MLP_IN - MLP_NFEAT);
        w[mlp_w_offset(0) + row * MLP_IN + col] = 0.0f;
    }
    if (i < (MLP_OUT - MLP_NOUT) * MLP_H) w[mlp_w_offset(3) + MLP_NOUT * MLP_H + i] = 0.0f;
    if (i < MLP_OUT - MLP_NOUT) b[mlp_b_offset(3) + MLP_NOUT + i] = 0.0f;
}

// ---- features and targets from the plain path ----------------------------------------

// The 16 DOM transmittance samples: visibility (shade.wesl:279-310) at the
// pixel's Eq.-8 light depth shifted by (k - 8) / 4 opacity layers. Sample 8
// is the plain method's own visibility.
__device__ void dom_samples(const View& p, const uint32_t* __restrict__ dom, F3 world, float beta, float out[16]) {
    if (p.flags[3] == 0u) {
        for (int k = 0; k < 16; ++k) out[k] = 1.0f;
        return;
    }
    float cx = mat_row(p.lclip, 0, world.x, world.y, world.z, 1.0f);
    float cy = mat_row(p.lclip, 1, world.x, world.y, world.z, 1.0f);
    float cw = mat_row(p.lclip, 3, world.x, world.y, world.z, 1.0f);
    float ux = cx / cw * 0.5f + 0.5f, uy = 1.0f - (cy / cw * 0.5f + 0.5f);
    if (cw <= 0.0f || ux < 0.0f || uy < 0.0f || ux > 1.0f || uy > 1.0f) {
        for (int k = 0; k < 16; ++k) out[k] = 1.0f;
        return;
    }
    uint32_t S = to_u32(p.dom_info.x);
    float near_ = p.dom_info.z, far_ = p.dom_info.w;
    float thickness = mx(far_ - near_, 1e-3f) / 16.0f;
    float dist = dot3(world - xyz(p.light_eye), xyz(p.light_forward)) - (-logf(mx(beta, 1e-4f)) * p.filter_params.z);
    float fx = ux * (float)S - 0.5f, fy = uy * (float)S - 0.5f;
    float flx = floorf(fx), fly = floorf(fy);
    int32_t x0 = to_i32(flx), y0 = to_i32(fly);
    float tx = fx - flx, ty = fy - fly;
    float wts[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    // per texel: prefix occupancy; visibility blended bilinearly like dom_visibility
    float vis[16];
    for (int k = 0; k < 16; ++k) vis[k] = 0.0f;
    for (int c = 0; c < 4; ++c) {
        int32_t x = x0 + (c & 1), y = y0 + (c >> 1);
        bool inside = x >= 0 && y >= 0 && x < (int32_t)S && y < (int32_t)S;
        uint32_t front = inside ? dom[(size_t)y * S + x] : HM_DOM_EMPTY;
        if (front == HM_DOM_EMPTY) {
            for (int k = 0; k < 16; ++k) vis[k] += wts[c];
            continue;
        }
        float z_front = dequantize_depth(front, near_, far_);
        float pre[17];
        pre[0] = 0.0f;
        const uint4* c4 = reinterpret_cast<const uint4*>(dom + (size_t)S * S + ((size_t)y * S + x) * HM_DOM_LAYERS);
        for (int qd = 0; qd < 4; ++qd) {
            uint4 v = c4[qd];
            pre[4 * qd + 1] = pre[4 * qd] + (float)v.x;
            pre[4 * qd + 2] = pre[4 * qd + 1] + (float)v.y;
            pre[4 * qd + 3] = pre[4 * qd + 2] + (float)v.z;
            pre[4 * qd + 4] = pre[4 * qd + 3] + (float)v.w;
        }
        float rel0 = (dist - z_front - thickness * 0.05f) / thickness;
        for (int k = 0; k < 16; ++k) {
            float t = mx(rel0 + 0.25f * (float)(k - 8), 0.0f);
            float occ;
            if (t >= 16.0f) {
                occ = pre[16];
            } else {
                int s = (int)t;
                occ = pre[s] + (pre[s + 1] - pre[s]) * (t - (float)s);
            }
            vis[k] += wts[c] * expf(-occ * p.filter_params.w);
        }
    }
    for (int k = 0; k < 16; ++k) out[k] = vis[k];
}

// Features of one pixel (see hm_mlp.h for the order and scaling).
__device__ void mlp_features(const View& p, const PixelShade& px, F3 key_dir, const uint32_t* dom, float f[32]) {
    f[0] = px.tangent.x; f[1] = px.tangent.y; f[2] = px.tangent.z;
    f[3] = px.view.x; f[4] = px.view.y; f[5] = px.view.z;
    F3 l = safe_normalize(key_dir);
    f[6] = l.x; f[7] = l.y; f[8] = l.z;
    dom_samples(p, dom, px.world, px.beta, f + 9);
    f[25] = px.sigma_a.x * MLP_SIGMA_SCALE;
    f[26] = px.sigma_a.y * MLP_SIGMA_SCALE;
    f[27] = px.sigma_a.z * MLP_SIGMA_SCALE;
    f[28] = px.roughness;
    f[29] = f[30] = f[31] = 0.0f;
}

// The plain method's values of the 12 outputs for one pixel and key light.
__device__ void mlp_targets(const PixelShade& px, const DualScatter& ds, F3 key_dir, float depth, float t[12]) {
    BsdfGeom gk = bsdf_geom_n(px.t_n, px.v_n, key_dir);
    float direct = expf(-depth);
    F3 back = local_back(gk.theta_h, ds.a_back, ds.delta_b, ds.sigma_b2);
    F3 ms = direct * (0.7f * back) + global_spread(gk.theta_h, depth, direct, px.shape, ds);
    t[0] = ms.x; t[1] = ms.y; t[2] = ms.z;
    t[3] = ds.a_back.x; t[4] = ds.a_back.y; t[5] = ds.a_back.z;
    t[6] = ds.delta_b.x; t[7] = ds.delta_b.y; t[8] = ds.delta_b.z;
    t[9] = ds.sigma_b2.x; t[10] = ds.sigma_b2.y; t[11] = ds.sigma_b2.z;
}

__device__ __forceinline__ int key_light(const ShadeArgs& a) {
    for (uint32_t i = 0; i < a.directional_count; ++i)
        if (a.directionals[i].flags & HM_LIGHT_FLAG_SHADOWS) return (int)i;
    return -1;
}

__device__ __forceinline__ F3 dir_of(const hm_directional_light& L) {
    return {L.direction_to_light[0], L.direction_to_light[1], L.direction_to_light[2]};
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
    float depth = -logf(mx(dom_visibility(p, a.dom, px.world, px.beta), 1e-6f));
    F3 kd = dir_of(a.directionals[key]);
    float f[32], t[12];
    mlp_features(p, px, kd, a.dom, f);
    mlp_targets(px, ds, kd, depth, t);
    uint32_t slot = atomicAdd(s.count, 1u);
    if (slot >= s.capacity) return;
    for (int k = 0; k < 32; ++k) s.x[(size_t)slot * 32 + k] = f[k];
    for (int k = 0; k < 12; ++k) s.t[(size_t)slot * 12 + k] = t[k];
}

// ---- MLP shading ---------------------------------------------------------------------

// Colour of one pixel from the network's 12 outputs: analytic single
// scattering for every light, the predicted local terms for the back lobe of
// the non-key lights, the predicted key-light multiple scattering.
__device__ F3 shade_from_mlp(const ShadeArgs& a, const PixelShade& px, const float o[12], int key, float depth) {
    F3 a_back = {o[3], o[4], o[5]}, delta_b = {o[6], o[7], o[8]};
    F3 sigma_b2 = vmax(F3{o[9], o[10], o[11]}, splat3(1e-6f));
    F3 color = splat3(0.0f);
    for (int pass = 0; pass < 2; ++pass) {
        for (uint32_t i = 0; i < a.point_count; ++i) {
            const hm_point_light& L = a.points[i];
            if ((pass == 0) != (L.kind == HM_LIGHT_POINT)) continue;
            F3 ltf = point_pos(L) - px.world;
            float att = distance_attenuation(dot3(ltf, ltf), L.color_inverse_square_range[3]);
            if (att == 0.0f) continue;
            BsdfGeom g = bsdf_geom_n(px.t_n, px.v_n, safe_normalize(ltf));
            F3 lobe = single_scatter(g, px.sigma_a, px.shape) + 0.7f * local_back(g.theta_h, a_back, delta_b, sigma_b2);
            F3 c = lobe * point_color(L) * att;
            if (pass == 1) c = c * spot_mask(L, px.world);
            color = color + c;
        }
    }
    for (uint32_t i = 0; i < a.directional_count; ++i) {
        const hm_directional_light& L = a.directionals[i];
        BsdfGeom g = bsdf_geom_n(px.t_n, px.v_n, dir_of(L));
        F3 single = single_scatter(g, px.sigma_a, px.shape);
        F3 lobe;
        if ((int)i == key) {
            lobe = expf(-depth) * single + F3{o[0], o[1], o[2]};
        } else {
            lobe = single + 0.7f * local_back(g.theta_h, a_back, delta_b, sigma_b2);
        }
        color = color + lobe * F3{L.color[0], L.color[1], L.color[2]};
    }
    return color * a.exposure;
}

// One wave shades a 16x4 tile: features -> LDS -> MFMA forward -> colour.
// Tiles whose inputs leave the trained range are flagged for the plain path.
__global__ __launch_bounds__(256) void k_shade_mlp(ShadeArgs a, MlpShadeArgs m) {
    __shared__ bf16x8 s_tiles[MLP_TILES * 64];
    __shared__ float s_bias[MLP_NBIAS];
    __shared__ __bf16 s_x[4][64 * 32];
    __shared__ float s_y[4][64 * 16];
    const View& p = *reinterpret_cast<const View*>(&a.view);
    {
        const uint4* src = reinterpret_cast<const uint4*>(m.packed);
        uint4* dst = reinterpret_cast<uint4*>(s_tiles);
        for (int i = threadIdx.x; i < MLP_TILES * 64; i += 256) dst[i] = src[i];
        for (int i = threadIdx.x; i < MLP_NBIAS; i += 256) s_bias[i] = m.bias[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, q = lane & 15;
    uint32_t x = blockIdx.x * 16u + (uint32_t)(lane & 15);
    uint32_t y = blockIdx.y * 16u + (uint32_t)(wv * 4 + (lane >> 4));
    uint32_t width = to_u32(p.screen.x), height = to_u32(p.screen.y);
    bool inb = x < width && y < height;
    size_t pix = inb ? (size_t)y * width + x : 0;
    uint64_t word = inb ? a.center[pix] : HM_GB_EMPTY;
    bool hit = !gb_empty(word);
    int key = key_light(a);
    PixelShade px;
    float depth = 0.0f;
    bool outside = false;
    float f[32];
    for (int k = 0; k < 32; ++k) f[k] = 0.0f;
    if (hit) {
        px = pixel_setup(p, word, a.beta[pix], x, y);
        F3 kd = key >= 0 ? dir_of(a.directionals[key]) : F3{0.0f, 1.0f, 0.0f};
        mlp_features(p, px, kd, a.dom, f);
        depth = -logf(mx(f[9 + 8], 1e-6f));
        for (int k = 9; k < MLP_NFEAT; ++k) outside |= f[k] < m.lo[k] || f[k] > m.hi[k];
    }
    // stage features for the matrix cores (row = pixel lane)
    for (int k = 0; k < 32; ++k) s_x[wv][lane * 32 + k] = (__bf16)f[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const bf16x8* X = reinterpret_cast<const bf16x8*>(s_x[wv]);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        bf16x8 b0 = X[(cb * 16 + q) * 4 + g];
        f32x4 yv = mlp_fwd16(s_tiles, s_bias, b0, lane, nullptr);
        *reinterpret_cast<float4*>(&s_y[wv][(cb * 16 + q) * 16 + 4 * g]) = make_float4(yv[0], yv[1], yv[2], yv[3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    bool flag = __any(hit && outside);
    if (lane == 0) {
        uint32_t ty = blockIdx.y * 4u + (uint32_t)wv;
        if (ty < m.tiles_y) m.tile_flags[ty * m.tiles_x + blockIdx.x] = flag ? 1u : 0u;
    }
    if (!inb) return;
    if (!hit) {
        a.shaded[pix] = make_float4(0, 0, 0, 0);
        a.shaded_depth[pix] = 0.0f;
        return;
    }
    if (flag) return;  // the plain path shades this tile
    float o[12];
    for (int k = 0; k < 12; ++k) o[k] = s_y[wv][lane * 16 + k];
    F3 c = shade_from_mlp(a, px, o, key, depth);
    a.shaded[pix] = make_float4(c.x, c.y, c.z, 1.0f);
    a.shaded_depth[pix] = px.view_dist;
}

// One probe pixel per network-shaded tile is shaded by the plain method; a
// tile whose probe misses by more than the bound is flagged for fallback.
__global__ __launch_bounds__(256) void k_mlp_check(ShadeArgs a, MlpShadeArgs m) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t tile = blockIdx.x * 256u + threadIdx.x;
    if (tile >= m.tiles_x * m.tiles_y || m.tile_flags[tile] != 0u) return;
    uint32_t width = to_u32(p.screen.x), height = to_u32(p.screen.y);
    uint32_t x0 = (tile % m.tiles_x) * 16u, y0 = (tile / m.tiles_x) * 4u;
    uint32_t start = pcg(tile) & 63u;
    for (uint32_t k = 0; k < 64u; ++k) {
        uint32_t idx = (start + k) & 63u;
        uint32_t x = x0 + (idx & 15u), y = y0 + (idx >> 4);
        if (x >= width || y >= height) continue;
        size_t pix = (size_t)y * width + x;
        uint64_t word = a.center[pix];
        if (gb_empty(word)) continue;
        PixelShade px = pixel_setup(p, word, a.beta[pix], x, y);
        F3 ref = plain_pixel_color(a, p, px);
        float4 got = a.shaded[pix];
        float err = fabsf(got.x - ref.x) + fabsf(got.y - ref.y) + fabsf(got.z - ref.z);
        float mag = fabsf(ref.x) + fabsf(ref.y) + fabsf(ref.z);
        if (err > m.tol * mx(mag, m.floor)) m.tile_flags[tile] = 2u;
        return;
    }
}

// Plain shading of every flagged tile (one wave per 16x4 tile).
__global__ __launch_bounds__(256) void k_shade_tiles(ShadeArgs a, MlpShadeArgs m) {
    const View& p = *reinterpret_cast<const View*>(&a.view);
    uint32_t tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (tile >= m.tiles_x * m.tiles_y) return;
    uint32_t f = m.tile_flags[tile];
    if (f == 0u) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t width = to_u32(p.screen.x), height = to_u32(p.screen.y);
    uint32_t x = (tile % m.tiles_x) * 16u + (lane & 15u), y = (tile / m.tiles_x) * 4u + (lane >> 4);
    if (lane == 0) m.tile_flags[tile] = f | 4u;  // remembers that the plain path ran
    if (x >= width || y >= height) return;
    size_t pix = (size_t)y * width + x;
    uint64_t word = a.center[pix];
    if (gb_empty(word)) return;
    PixelShade px = pixel_setup(p, word, a.beta[pix], x, y);
    F3 c = plain_pixel_color(a, p, px);
    a.shaded[pix] = make_float4(c.x, c.y, c.z, 1.0f);
    a.shaded_depth[pix] = px.view_dist;
}

// ---- launchers --------------------------------------------------------------------------

hipError_t launch_mlp_pack(hipStream_t s, const float* w, __bf16* packed) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_TILES), dim3(64), 0, s, w, packed);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(hipStream_t s, const float* x, uint32_t n, const __bf16* packed, const float* bias,
                              float* y) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + 63) / 64;
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

hipError_t launch_shade_mlp(hipStream_t s, const ShadeArgs& a, const MlpShadeArgs& m, uint32_t width,
                            uint32_t height) {
    if (width == 0 || height == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_mlp, dim3((width + 15) / 16, (height + 15) / 16), dim3(256), 0, s, a, m);
    uint32_t tiles = m.tiles_x * m.tiles_y;
    hipLaunchKernelGGL(k_mlp_check, dim3((tiles + 255) / 256), dim3(256), 0, s, a, m);
    hipLaunchKernelGGL(k_shade_tiles, dim3((tiles + 3) / 4), dim3(256), 0, s, a, m);
    return hipGetLastError();
}

}  // namespace hm

// mlp.hpp -- learned CG warm start: a 32-128-128-128-5 MLP on BF16 MFMA (gfx950),
// FP32 accumulation, trained on the GPU by the kernels below.  Included at the end of
// yarnstep.hip (it reads the context's device buffers).
//
// The network corrects the quadratic extrapolation x_q = 3 v_n - 3 v_{n-1} + v_{n-2}
// of the implicit step's velocity: per node it sees 32 features (the Jacobi correction
// D^-1 (b - A x_q) at the node and the L part of it at the four neighbours, the last two
// steps' velocity changes, the sliding offset and three flags) and predicts x - x_q;
// CG then starts from x0 = x_q + out_std * y and refines to the same residual as from a
// zero guess.
#pragma once

namespace {

constexpr int MLP_IN = 32, MLP_H = 128, MLP_OUT = 5, MLP_OUTP = 16;
constexpr int MLP_TILE = 64;   // rows per block (4 waves x 16)
// flat FP32 parameter layout: W1[128x32] b1 W2[128x128] b2 W3[128x128] b3 W4[5x128] b4
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_H * MLP_IN, OFF_W2 = OFF_B1 + MLP_H, OFF_B2 = OFF_W2 + MLP_H * MLP_H,
              OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H * MLP_H, OFF_W4 = OFF_B3 + MLP_H,
              OFF_B4 = OFF_W4 + MLP_OUT * MLP_H, MLP_NPARAM = OFF_B4 + MLP_OUT;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (bit pattern); finite inputs
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }
__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

struct MlpW {            // BF16 weights as stored for the MFMA B operand (row-major [out][in])
    const unsigned short* W1;  // 128 x 32
    const unsigned short* W2;  // 128 x 128
    const unsigned short* W3;  // 128 x 128
    const unsigned short* W4;  // 16 x 128 (rows 5..15 zero)
    const float* b;            // FP32 master params (biases read from here)
};

// Row strides of the bf16 activation tiles in LDS: 4 elements of padding move
// consecutive rows by two banks, so the 16 rows an MFMA fragment load touches fall in
// different banks.
constexpr int MLP_LDX = MLP_IN + 4, MLP_LDH = MLP_H + 4;

// C = A W^T over a 64-row tile, A in LDS (bf16, [64][LDA]), W in global memory (bf16,
// row-major [N][K]).  Wave w computes columns [w N/4, (w+1) N/4) of all 64 rows, so each
// wave reads a quarter of W (one 8-byte fragment per lane and k-slab, reused for the
// four row tiles).  Every output element accumulates its k-slabs in ascending order.
// epi(row, col, value + bias).
template <int K, int N, int LDA, class Epi>
__device__ __forceinline__ void mfma_layer(const unsigned short* As, const unsigned short* Wg, const float* bias, Epi epi) {
    constexpr int NW = N / 4, NC = NW / 16;
    static_assert(NC >= 1 && NW % 16 == 0, "columns per wave");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kb = lane >> 4;
    floatx4 acc[4][NC];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < NC; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 16; ++kt) {
        bf16x4 bv[NC];
#pragma unroll
        for (int ct = 0; ct < NC; ++ct)
            bv[ct] = *(const bf16x4*)(Wg + (size_t)(wave * NW + ct * 16 + i) * K + kt * 16 + 4 * kb);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            bf16x4 a = *(const bf16x4*)(As + (rt * 16 + i) * LDA + kt * 16 + 4 * kb);
#pragma unroll
            for (int ct = 0; ct < NC; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[ct], acc[rt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) {
        const int col = wave * NW + ct * 16 + i;
        const float bb = bias ? bias[col] : 0.0f;
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) epi(rt * 16 + 4 * kb + r, col, acc[rt][ct][r] + bb);
    }
}

// The 16-column output layer: wave w computes rows [16 w, 16 w + 16).  epi(row, col, value).
template <int K, int LDA, class Epi>
__device__ __forceinline__ void mfma_out_layer(const unsigned short* As, const unsigned short* Wg, Epi epi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kb = lane >> 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 16; ++kt) {
        bf16x4 a = *(const bf16x4*)(As + (wave * 16 + i) * LDA + kt * 16 + 4 * kb);
        bf16x4 bv = *(const bf16x4*)(Wg + (size_t)i * K + kt * 16 + 4 * kb);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) epi(wave * 16 + 4 * kb + r, i, acc[r] + 0.0f);
}

// The three hidden layers and the output layer of one 64-row tile whose standardised
// bf16 inputs are in xs; hid(layer, row, col, z) sees every hidden pre-activation after
// its bf16(silu(z)) is stored, out(row, col, y) every output (y = z + b4, col < 5).
template <class Hid, class Out>
__device__ __forceinline__ void mlp_tile(const unsigned short* xs, unsigned short* h0, unsigned short* h1, const MlpW& w,
                                         Hid hid, Out out) {
    auto store = [&](unsigned short* dst, int layer) {
        return [=](int r, int c, float z) {
            float hval = z * sigmoidf_(z);
            unsigned short hb = f2bf(hval);
            dst[r * MLP_LDH + c] = hb;
            hid(layer, r, c, z, hb);
        };
    };
    mfma_layer<MLP_IN, MLP_H, MLP_LDX>(xs, w.W1, w.b + OFF_B1, store(h0, 0));
    __syncthreads();
    mfma_layer<MLP_H, MLP_H, MLP_LDH>(h0, w.W2, w.b + OFF_B2, store(h1, 1));
    __syncthreads();
    mfma_layer<MLP_H, MLP_H, MLP_LDH>(h1, w.W3, w.b + OFF_B3, store(h0, 2));
    __syncthreads();
    const float* b4 = w.b + OFF_B4;
    mfma_out_layer<MLP_H, MLP_LDH>(h0, w.W4, [=](int r, int c, float z) {
        if (c < MLP_OUT) out(r, c, z + b4[c]);
    });
}

struct MlpFwdArgs {
    const float* X;      // rows x 32 raw features
    const int* idx;      // optional row gather (training batches)
    int rows;
    const float* in_mean; const float* in_istd;   // standardisation
    MlpW w;
    float* Y;            // rows x 5 outputs (normalised units), may be null
    // training stores (per row of the batch): X0 bf16, Z fp32, H bf16
    unsigned short* sX0; float* sZ; unsigned short* sH;
};

// Forward over 64-row tiles.  Z_l and H_l = bf16(silu(Z_l)) are kept for the backward.
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpFwdArgs a) {
    __shared__ unsigned short xs[MLP_TILE * MLP_LDX];
    __shared__ unsigned short h0[MLP_TILE * MLP_LDH];
    __shared__ unsigned short h1[MLP_TILE * MLP_LDH];
    int base = blockIdx.x * MLP_TILE;
    for (int t = threadIdx.x; t < MLP_TILE * MLP_IN; t += blockDim.x) {
        int r = t / MLP_IN, c = t % MLP_IN;
        int gr = base + r;
        float v = 0.0f;
        if (gr < a.rows) {
            int src = a.idx ? a.idx[gr] : gr;
            v = (a.X[(size_t)src * MLP_IN + c] - a.in_mean[c]) * a.in_istd[c];
        }
        unsigned short hb = f2bf(v);
        xs[r * MLP_LDX + c] = hb;
        if (a.sX0 && gr < a.rows) a.sX0[(size_t)gr * MLP_IN + c] = hb;
    }
    __syncthreads();
    const int R = a.rows;
    float* Y = a.Y;
    mlp_tile(xs, h0, h1, a.w,
             [=](int layer, int r, int c, float z, unsigned short hb) {
                 int gr = base + r;
                 if (a.sZ && gr < R) {
                     a.sZ[((size_t)layer * R + gr) * MLP_H + c] = z;
                     a.sH[((size_t)layer * R + gr) * MLP_H + c] = hb;
                 }
             },
             [=](int r, int c, float y) {
                 int gr = base + r;
                 if (gr < R && Y) Y[(size_t)gr * MLP_OUT + c] = y;
             });
}

struct MlpBwdArgs {
    int rows;            // batch rows
    const unsigned short* sX0; const float* sZ; const unsigned short* sH;
    const float* Y; const float* T; const float* M;   // outputs, targets, masks (rows x 5)
    const int* idx;      // gather for T, M (same as the forward)
    MlpW w;
    float* part;         // [nblocks][MLP_NPARAM + 1] partial gradients (+ loss)
    float inv_count;     // 1 / (rows * 5)
};

// Backward of one 64-row tile; writes the tile's partial gradient of every parameter
// and its partial loss.  Large products on MFMA (BF16 operands, FP32 accumulation).
__global__ void __launch_bounds__(256) k_mlp_bwd(MlpBwdArgs a) {
    __shared__ unsigned short dz[MLP_TILE * MLP_H];    // bf16 dZ of the current layer
    __shared__ unsigned short hs[MLP_TILE * MLP_H];    // bf16 input activations of the layer
    __shared__ float dzf[MLP_TILE * MLP_H];            // fp32 dZ / dH scratch
    __shared__ float dy[MLP_TILE * MLP_OUTP];
    __shared__ float red[256];
    const int R = a.rows;
    int base = blockIdx.x * MLP_TILE;
    int nrows = min(MLP_TILE, R - base);
    float* P = a.part + (size_t)blockIdx.x * (MLP_NPARAM + 1);
    int tid = threadIdx.x;
    // dY and loss
    float lsum = 0.0f;
    for (int t = tid; t < MLP_TILE * MLP_OUTP; t += blockDim.x) {
        int r = t / MLP_OUTP, o = t % MLP_OUTP;
        float g = 0.0f;
        if (r < nrows && o < MLP_OUT) {
            int gr = base + r;
            int src = a.idx ? a.idx[gr] : gr;
            float m = a.M[(size_t)src * MLP_OUT + o];
            float d = (a.Y[(size_t)gr * MLP_OUT + o] - a.T[(size_t)src * MLP_OUT + o]) * m;
            lsum += d * d;
            g = 2.0f * d * a.inv_count;
        }
        dy[t] = g;
    }
    red[tid] = lsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) { if (tid < s) red[tid] += red[tid + s]; __syncthreads(); }
    if (tid == 0) P[MLP_NPARAM] = red[0] * a.inv_count;
    // layer 4 (K = 5): dW4, db4 scalar; dH3 = dY W4
    const unsigned short* H3 = a.sH + (size_t)2 * R * MLP_H;
    for (int t = tid; t < MLP_OUT * MLP_H; t += blockDim.x) {
        int o = t / MLP_H, k = t % MLP_H;
        float s = 0.0f;
        for (int r = 0; r < nrows; ++r) s += dy[r * MLP_OUTP + o] * bf2f(H3[(size_t)(base + r) * MLP_H + k]);
        P[OFF_W4 + t] = s;
    }
    if (tid < MLP_OUT) {
        float s = 0.0f;
        for (int r = 0; r < nrows; ++r) s += dy[r * MLP_OUTP + tid];
        P[OFF_B4 + tid] = s;
    }
    for (int t = tid; t < MLP_TILE * MLP_H; t += blockDim.x) {
        int r = t / MLP_H, k = t % MLP_H;
        float s = 0.0f;
        for (int o = 0; o < MLP_OUT; ++o) s += dy[r * MLP_OUTP + o] * bf2f(a.w.W4[o * MLP_H + k]);
        dzf[t] = s;   // dH3
    }
    __syncthreads();
    const unsigned short* Wl[3] = {a.w.W1, a.w.W2, a.w.W3};
    const int offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
    int lane = tid & 63, wave = tid >> 6, i = lane & 15, kb = lane >> 4;
    for (int l = 2; l >= 0; --l) {
        // dZ = dH * silu'(Z); fp32 for the bias, bf16 for the products
        const float* Z = a.sZ + (size_t)l * R * MLP_H;
        for (int t = tid; t < MLP_TILE * MLP_H; t += blockDim.x) {
            int r = t / MLP_H, k = t % MLP_H;
            float g = 0.0f;
            if (r < nrows) {
                float z = Z[(size_t)(base + r) * MLP_H + k];
                float s = sigmoidf_(z);
                g = dzf[t] * (s * (1.0f + z * (1.0f - s)));
            }
            dzf[t] = g;
            dz[t] = f2bf(g);
        }
        // layer input activations (bf16) into LDS
        int K = l == 0 ? MLP_IN : MLP_H;
        for (int t = tid; t < MLP_TILE * K; t += blockDim.x) {
            int r = t / K, k = t % K;
            unsigned short v = 0;
            if (r < nrows) v = l == 0 ? a.sX0[(size_t)(base + r) * MLP_IN + k] : a.sH[((size_t)(l - 1) * R + base + r) * MLP_H + k];
            hs[t] = v;
        }
        __syncthreads();
        // db
        if (tid < MLP_H) {
            float s = 0.0f;
            for (int r = 0; r < nrows; ++r) s += dzf[r * MLP_H + tid];
            P[offB[l] + tid] = s;
        }
        // dW[o][k] = sum_r dz[r][o] * h[r][k]: A = dz^T (M = out, K = rows), B = h (K = rows, N = in)
        int ntiles_n = K / 16;
        for (int tile = wave; tile < 8 * ntiles_n; tile += 4) {
            int mt = tile / ntiles_n, nt = tile % ntiles_n;
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < MLP_TILE / 16; ++kt) {
                int r0 = kt * 16 + 4 * kb;
                bf16x4 av = {(short)dz[(r0 + 0) * MLP_H + mt * 16 + i], (short)dz[(r0 + 1) * MLP_H + mt * 16 + i],
                             (short)dz[(r0 + 2) * MLP_H + mt * 16 + i], (short)dz[(r0 + 3) * MLP_H + mt * 16 + i]};
                bf16x4 bv = {(short)hs[(r0 + 0) * K + nt * 16 + i], (short)hs[(r0 + 1) * K + nt * 16 + i],
                             (short)hs[(r0 + 2) * K + nt * 16 + i], (short)hs[(r0 + 3) * K + nt * 16 + i]};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) P[offW[l] + (mt * 16 + 4 * kb + r) * K + nt * 16 + i] = acc[r];
        }
        __syncthreads();
        if (l == 0) break;
        // dH_{l-1}[r][k] = sum_o dz[r][o] W_l[o][k]: A = dz (M = rows, K = out), B = W_l (K = out, N = in)
        const unsigned short* Wg = Wl[l];
        for (int nt = 0; nt < MLP_H / 16; ++nt) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < MLP_H / 16; ++kt) {
                const unsigned short* ap = dz + (wave * 16 + i) * MLP_H + kt * 16 + 4 * kb;
                bf16x4 av = {(short)ap[0], (short)ap[1], (short)ap[2], (short)ap[3]};
                int o0 = kt * 16 + 4 * kb;
                bf16x4 bv = {(short)Wg[(o0 + 0) * MLP_H + nt * 16 + i], (short)Wg[(o0 + 1) * MLP_H + nt * 16 + i],
                             (short)Wg[(o0 + 2) * MLP_H + nt * 16 + i], (short)Wg[(o0 + 3) * MLP_H + nt * 16 + i]};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) dzf[(wave * 16 + 4 * kb + r) * MLP_H + nt * 16 + i] = acc[r];
        }
        __syncthreads();
    }
}

// grad[p] = sum over blocks of the partials (fixed order); Adam step on the FP32
// master parameters when lr > 0; the last slot carries the mean loss.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(const float* part, int nblk, float* grad, float* param,
                                                          float* m1, float* m2, float lr, float b1t, float b2t) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += part[(size_t)b * (MLP_NPARAM + 1) + p];
    grad[p] = g;
    if (p == MLP_NPARAM || lr <= 0.0f) return;
    const float be1 = 0.9f, be2 = 0.999f, eps = 1e-8f;
    float mm = be1 * m1[p] + (1.0f - be1) * g;
    float vv = be2 * m2[p] + (1.0f - be2) * g * g;
    m1[p] = mm; m2[p] = vv;
    float mh = mm / (1.0f - b1t), vh = vv / (1.0f - b2t);
    param[p] -= lr * mh / (sqrtf(vh) + eps);
}

// BF16 copies of the weight matrices (W4 padded to 16 rows)
__global__ void __launch_bounds__(256) k_mlp_to_bf16(const float* param, unsigned short* W1, unsigned short* W2,
                                                      unsigned short* W3, unsigned short* W4) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_H * MLP_IN) W1[t] = f2bf(param[OFF_W1 + t]);
    if (t < MLP_H * MLP_H) { W2[t] = f2bf(param[OFF_W2 + t]); W3[t] = f2bf(param[OFF_W3 + t]); }
    if (t < MLP_OUTP * MLP_H) W4[t] = t < MLP_OUT * MLP_H ? f2bf(param[OFF_W4 + t]) : (unsigned short)0;
}

// ---------------------------------------------------------------------------
// Features and warm start on a cloth context
// ---------------------------------------------------------------------------
struct FeatArgs {
    const double* uv; const double* uvbar;
    const unsigned char* flags;
    const double* b; const double* invd;   // padded
    const double* Ae;                       // padded A x_q, x_q the extrapolated guess
    const double* dprev; const double* dprev2;   // padded velocity changes of the last two steps
    int W, Lg, N, Wp, P;
    double L;
};

// The 32 features of node k (u = k mod W, v = k / W), in four groups of straight-line
// loads (one group per wavefront of k_mlp_warm; k_features runs all four):
//   0..4    D^-1 (b - A x_q) at the node: the Jacobi correction of the extrapolated
//           guess x_q;  29..31 interior crossing, handle, flipped           (group 0)
//   5..9    dv of the last step, 10..14 dv of the step before,
//   15, 16  sliding offset (uv - uvbar) / L                                 (group 1)
//   17..22  the Jacobi correction's x, y, z at (u-1, v) and (u+1, v)        (group 2)
//   23..28  the same at (u, v-1) and (u, v+1)                               (group 3)
// Neighbours outside the grid and the crossing DOFs (components 3, 4) of boundary
// nodes read zero.
template <int G>
__device__ __forceinline__ void node_features(const FeatArgs& a, int k, float* f) {
    const int u = k % a.W, v = k / a.W;
    const bool inter = u > 0 && u < a.W - 1 && v > 0 && v < a.Lg - 1;
    auto jac = [&](int uu, int vv, int c) {
        size_t i = (size_t)c * a.P + padidx(a.Wp, uu, vv);
        return (float)(a.invd[i] * (a.b[i] - a.Ae[i]));
    };
    if constexpr (G == 0) {
#pragma unroll
        for (int c = 0; c < 5; ++c) f[c] = (c < 3 || inter) ? jac(u, v, c) : 0.0f;
        f[29] = inter ? 1.0f : 0.0f;
        f[30] = (a.flags[k] & 2) ? 1.0f : 0.0f;
        f[31] = (a.flags[k] & 1) ? 1.0f : 0.0f;
    } else if constexpr (G == 1) {
        const size_t pk = padidx(a.Wp, u, v);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            f[5 + c] = (c < 3 || inter) ? (float)a.dprev[(size_t)c * a.P + pk] : 0.0f;
            f[10 + c] = (c < 3 || inter) ? (float)a.dprev2[(size_t)c * a.P + pk] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < 2; ++c)
            f[15 + c] = inter ? (float)((a.uv[(size_t)c * a.N + k] - a.uvbar[(size_t)c * a.N + k]) / a.L) : 0.0f;
    } else {
        // G 2: (u-1, v), (u+1, v); G 3: (u, v-1), (u, v+1)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int uu = G == 2 ? u + 2 * q - 1 : u, vv = G == 3 ? v + 2 * q - 1 : v;
            const bool in = uu >= 0 && uu < a.W && vv >= 0 && vv < a.Lg;
#pragma unroll
            for (int c = 0; c < 3; ++c) f[17 + 6 * (G - 2) + 3 * q + c] = in ? jac(uu, vv, c) : 0.0f;
        }
    }
}

// the feature indices of group G: 0 -> {0..4, 29..31}, 1 -> 5..16, 2 -> 17..22, 3 -> 23..28
__device__ __forceinline__ bool in_group(int G, int j) {
    return G == 0 ? (j < 5 || j >= 29) : G == 1 ? (j >= 5 && j < 17) : G == 2 ? (j >= 17 && j < 23) : (j >= 23 && j < 29);
}

__global__ void __launch_bounds__(256) k_features(FeatArgs a, float* F) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    float f[MLP_IN];
    node_features<0>(a, k, f);
    node_features<1>(a, k, f);
    node_features<2>(a, k, f);
    node_features<3>(a, k, f);
    for (int j = 0; j < MLP_IN; ++j) F[(size_t)k * MLP_IN + j] = f[j];
}

template <int G>
__device__ __forceinline__ void stage_features(const FeatArgs& a, int k, int r, const float* in_mean, const float* in_istd,
                                               unsigned short* xs) {
    float f[MLP_IN];
    if (k < a.N) node_features<G>(a, k, f);
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j)
        if (in_group(G, j)) xs[r * MLP_LDX + j] = f2bf(k < a.N ? (f[j] - in_mean[j]) * in_istd[j] : 0.0f);
}

// The learned warm start of nodes [64 tile, 64 tile + 64) by a 256-thread block: features
// (wavefront G gathers group G of every row), standardisation, the network (the same tile
// forward as k_mlp_fwd, so the same bits), and x0 += out_std * y on the node's DOFs (x0
// holds x_q on entry).
__device__ __forceinline__ void mlp_warm_tile(const FeatArgs& a, const MlpW& w, const float* in_mean, const float* in_istd,
                                              const float* out_std, double* x0, int tile) {
    __shared__ unsigned short xs[MLP_TILE * MLP_LDX];
    __shared__ unsigned short h0[MLP_TILE * MLP_LDH];
    __shared__ unsigned short h1[MLP_TILE * MLP_LDH];
    const int base = tile * MLP_TILE;
    {
        const int r = threadIdx.x & (MLP_TILE - 1), k = base + r;
        switch (threadIdx.x / MLP_TILE) {
            case 0: stage_features<0>(a, k, r, in_mean, in_istd, xs); break;
            case 1: stage_features<1>(a, k, r, in_mean, in_istd, xs); break;
            case 2: stage_features<2>(a, k, r, in_mean, in_istd, xs); break;
            default: stage_features<3>(a, k, r, in_mean, in_istd, xs); break;
        }
    }
    __syncthreads();
    mlp_tile(xs, h0, h1, w, [](int, int, int, float, unsigned short) {},
             [=](int r, int c, float y) {
                 const int k = base + r;
                 if (k >= a.N) return;
                 const int u = k % a.W, v = k / a.W;
                 const bool inter = u > 0 && u < a.W - 1 && v > 0 && v < a.Lg - 1;
                 if (c < 3 || inter) x0[(size_t)c * a.P + padidx(a.Wp, u, v)] += (double)y * (double)out_std[c];
             });
}

// The learned warm start in one launch per 64 nodes.
__global__ void __launch_bounds__(256) k_mlp_warm(FeatArgs a, MlpW w, const float* in_mean, const float* in_istd,
                                                   const float* out_std, double* x0) {
    mlp_warm_tile(a, w, in_mean, in_istd, out_std, x0, blockIdx.x);
}

// padded v_old (x, y, z, u, v; boundary u, v zero)
__global__ void __launch_bounds__(256) k_pack_vel(const double* v, const double* uvd, int W, int Lg, int N, int Wp, int P, double* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int u = k % W, vv = k / W;
    int pk = padidx(Wp, u, vv);
    bool inter = u > 0 && u < W - 1 && vv > 0 && vv < Lg - 1;
    for (int c = 0; c < 3; ++c) out[(size_t)c * P + pk] = v[(size_t)c * N + k];
    for (int c = 0; c < 2; ++c) out[(size_t)(3 + c) * P + pk] = inter ? uvd[(size_t)c * N + k] : 0.0;
}

// x0 from the padded previous velocity v_n and the last two steps' changes:
// v_n (kind 1), the linear extrapolation v_n + dprev = 2 v_n - v_{n-1} (kind 2), or the
// quadratic one v_n + 2 dprev - dprev2 = 3 v_n - 3 v_{n-1} + v_{n-2} (kind 3)
__global__ void __launch_bounds__(256) k_guess_x0(const double* vpad, const double* dprev, const double* dprev2, int kind,
                                                   int N5, double* x0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N5) return;
    double x = vpad[i];
    if (kind == 2) x = vpad[i] + dprev[i];
    if (kind == 3) x = vpad[i] + (2.0 * dprev[i] - dprev2[i]);
    x0[i] = x;
}

}  // namespace

struct yc_mlp {
    float* param;        // FP32 master (MLP_NPARAM + 1)
    float* m1; float* m2; float* grad;
    unsigned short *W1, *W2, *W3, *W4;
    float* in_mean; float* in_istd; float* out_std;
    // training buffers
    float* X; float* T; float* M; int nrows;
    int* idx; int cap_batch;
    unsigned short* sX0; float* sZ; unsigned short* sH; float* Y; float* part;
    int adam_t;
    hipStream_t stream;
    // this device's copy of the staging array of yc_mlp_forward_noalloc (not owned)
    float* scratch;
};

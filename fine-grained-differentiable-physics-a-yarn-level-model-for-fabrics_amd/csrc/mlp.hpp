// mlp.hpp -- learned CG warm start: a 32-128-128-128-5 MLP on BF16 MFMA (gfx950),
// FP32 accumulation, trained on the GPU by the kernels below.  Included at the end of
// yarnstep.hip (it reads the context's device buffers).
//
// Per node the network sees 32 features (velocity, Eulerian velocity, the Jacobi
// estimate D^-1 (b - A v) of the velocity change, the previous step's velocity change,
// the sliding offset uv - uvbar, the four neighbour offsets and three flags) and
// predicts the velocity change dv of the implicit step; CG then starts from
// x0 = v + dv and refines to the same residual as from a zero guess.
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

// One 16-row slab of C = A (LDS, bf16, row-major [64][K]) x W^T for a wave, all N
// columns; epilogue callback gets (row, col, value-with-bias).
template <int K, int N, class Epi>
__device__ __forceinline__ void mfma_layer(const unsigned short* As, const unsigned short* Wg, const float* bias, Epi epi) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int i = lane & 15, kb = lane >> 4;
    int row0 = wave * 16;
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < K / 16; ++kt) {
            const unsigned short* ap = As + (row0 + i) * K + kt * 16 + 4 * kb;
            bf16x4 a = {(short)ap[0], (short)ap[1], (short)ap[2], (short)ap[3]};
            const unsigned short* bp = Wg + (nt * 16 + i) * K + kt * 16 + 4 * kb;
            bf16x4 bv = {(short)bp[0], (short)bp[1], (short)bp[2], (short)bp[3]};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc, 0, 0, 0);
        }
        int col = nt * 16 + i;
        float bb = bias ? bias[col] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) epi(row0 + 4 * kb + r, col, acc[r] + bb);
    }
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
    __shared__ unsigned short xs[MLP_TILE * MLP_IN];
    __shared__ unsigned short h0[MLP_TILE * MLP_H];
    __shared__ unsigned short h1[MLP_TILE * MLP_H];
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
        xs[t] = hb;
        if (a.sX0 && gr < a.rows) a.sX0[(size_t)gr * MLP_IN + c] = hb;
    }
    __syncthreads();
    const int R = a.rows;
    auto store = [&](unsigned short* dst, int layer) {
        return [=](int r, int c, float z) {
            float hval = z * sigmoidf_(z);
            unsigned short hb = f2bf(hval);
            dst[r * MLP_H + c] = hb;
            int gr = base + r;
            if (a.sZ && gr < R) {
                a.sZ[((size_t)layer * R + gr) * MLP_H + c] = z;
                a.sH[((size_t)layer * R + gr) * MLP_H + c] = hb;
            }
        };
    };
    mfma_layer<MLP_IN, MLP_H>(xs, a.w.W1, a.w.b + OFF_B1, store(h0, 0));
    __syncthreads();
    mfma_layer<MLP_H, MLP_H>(h0, a.w.W2, a.w.b + OFF_B2, store(h1, 1));
    __syncthreads();
    mfma_layer<MLP_H, MLP_H>(h1, a.w.W3, a.w.b + OFF_B3, store(h0, 2));
    __syncthreads();
    float* Y = a.Y;
    const float* b4 = a.w.b + OFF_B4;
    mfma_layer<MLP_H, MLP_OUTP>(h0, a.w.W4, nullptr, [=](int r, int c, float z) {
        int gr = base + r;
        if (c < MLP_OUT && gr < R && Y) Y[(size_t)gr * MLP_OUT + c] = z + b4[c];
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
    const double* x; const double* v; const double* uvd; const double* uv; const double* uvbar;
    const unsigned char* flags;
    const double* b; const double* invd; const double* Av;   // padded vectors; Av = A v_old
    const double* dprev;                                      // padded previous dv
    int W, Lg, N, Wp, P;
    double L;
    float* F;    // N x 32
};

__global__ void __launch_bounds__(256) k_features(FeatArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    int u = k % a.W, v = k / a.W;
    int pk = padidx(a.Wp, u, v);
    bool inter = u > 0 && u < a.W - 1 && v > 0 && v < a.Lg - 1;
    float f[MLP_IN];
    const size_t N = a.N;
    for (int c = 0; c < 3; ++c) f[c] = (float)a.v[c * N + k];
    for (int c = 0; c < 2; ++c) f[3 + c] = inter ? (float)a.uvd[c * N + k] : 0.0f;
    for (int c = 0; c < 5; ++c) {
        size_t i = (size_t)c * a.P + pk;
        f[5 + c] = (float)(a.invd[i] * (a.b[i] - a.Av[i]));
        f[10 + c] = (float)a.dprev[i];
    }
    for (int c = 0; c < 2; ++c) f[15 + c] = inter ? (float)((a.uv[c * N + k] - a.uvbar[c * N + k]) / a.L) : 0.0f;
    const int du[4] = {-1, 1, 0, 0}, dv[4] = {0, 0, -1, 1};
    for (int q = 0; q < 4; ++q) {
        int uu = u + du[q], vv = v + dv[q];
        bool in = uu >= 0 && uu < a.W && vv >= 0 && vv < a.Lg;
        int kk = uu + vv * a.W;
        for (int c = 0; c < 3; ++c) f[17 + 3 * q + c] = in ? (float)((a.x[c * N + kk] - a.x[c * N + k]) / a.L) : 0.0f;
    }
    f[29] = inter ? 1.0f : 0.0f;
    f[30] = (a.flags[k] & 2) ? 1.0f : 0.0f;
    f[31] = (a.flags[k] & 1) ? 1.0f : 0.0f;
    for (int c = 0; c < MLP_IN; ++c) a.F[(size_t)k * MLP_IN + c] = f[c];
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

// x0 = v_old + out_std * y (the network's dv in normalised units)
__global__ void __launch_bounds__(256) k_warm_x0(const double* vpad, const float* Y, const float* out_std, int W, int Lg, int N,
                                                  int Wp, int P, double* x0) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int u = k % W, vv = k / W;
    int pk = padidx(Wp, u, vv);
    bool inter = u > 0 && u < W - 1 && vv > 0 && vv < Lg - 1;
    for (int c = 0; c < 5; ++c) {
        size_t i = (size_t)c * P + pk;
        double d = (c < 3 || inter) ? (double)Y[(size_t)k * MLP_OUT + c] * (double)out_std[c] : 0.0;
        x0[i] = vpad[i] + d;
    }
}

// dv of the last solve relative to the velocity it started from
// x0 = v_old (kind 1) or the linear extrapolation v_old + (v_old - v_older) (kind 2) of
// the velocities, from the padded previous velocity and the last step's change
__global__ void __launch_bounds__(256) k_guess_x0(const double* vpad, const double* dprev, int kind, int N5, double* x0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N5) x0[i] = kind == 2 ? vpad[i] + dprev[i] : vpad[i];
}

__global__ void __launch_bounds__(256) k_dprev(const double* x, const double* vpad, int N5, double* dprev) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N5) dprev[i] = x[i] - vpad[i];
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
    // inference scratch for one context at a time
    float* F; float* Yi; int capN;
    // this device's copy of the staging array of yc_mlp_forward_noalloc (not owned)
    float* scratch;
};

// yarnstep.hip -- MI355X (gfx950) implicit yarn-level time step behind the C ABI of
// include/yarnstep.h.
//
// Device layout (FP64 throughout, the reference's dtype, src/Constants.h:9):
//   node state      structure of arrays, planes of N = W*Lg doubles (x, v, xfix: 3
//                   planes; uv, uvd, uvbar: 2 planes; crossing normals: 3 planes)
//   system matrix   159 slot planes of N doubles (ys::NSLOT): the fixed 13-node
//                   stencil of the yarn grid, columns implicit, no index arrays
//   CG vectors      5 planes (x, y, z, u, v) over the grid padded by 2 nodes on each
//                   side, so the stencil product needs no bounds checks; the padding
//                   and the boundary nodes' u, v entries stay exactly zero
// Kernels: k_assemble1 (edges, bends, crimp, faces, handles, friction; one thread
// per node), k_assemble2 (shear, diagonal, finiteness), CG (3 kernels per
// iteration, replayed from a hipGraph), k_update + k_normals.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <type_traits>

#include "../../include/yarnstep.h"
#include "elements.hpp"

using namespace ys;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(YC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int BLK = 256;
constexpr int MAX_DEVICES = 64;   // device indices of the per-device tables

// Mean device time per launch of `reps` back-to-back launches issued by `launch` on
// `stream` (after one untimed launch), from one event pair around the whole run.  An
// event pair around each single launch would add the event packets' own gaps (about
// 2 us per launch on MI355X) to a kernel of a few microseconds.
template <class F>
int time_launches(hipStream_t stream, int reps, double* avg_ms, F launch) {
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(YC_ERR_HIP, "hipEventCreate"); }
    launch();
    hipError_t err = hipEventRecord(e0, stream);
    for (int i = 0; i < reps && err == hipSuccess; ++i) launch();
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return fail(YC_ERR_HIP, std::string("timing: ") + hipGetErrorString(err));
    *avg_ms = (double)ms / reps;
    return YC_OK;
}
// per-node kernels (assembly, update, normals) are heavy per thread and have only one
// thread per node: small blocks spread them over all 256 CUs
constexpr int NBLK = 64;

struct CGScal {
    double rr, bb, thr;
    int done, maxit, zero_rhs;
    int itc[2];   // completed iterations, written by parity so readers never race the writer
    // single-reduction CG: alpha and gamma = r.u of the last iteration, by parity, and
    // the iteration count reported at the stop
    double alpha[2], gamma[2];
    int fin_iters;
    // persistent CG over many steps without host synchronisation: totals since the last
    // reset (iterations, solves stopped at maxit, solves with a non-finite residual)
    long long total_iters;
    int nonconv, nonfinite;
};

// Partial sums of the CG reductions, one slot per block of the producing kernel.  A
// consumer kernel reduces them in every block (fixed order, identical result in all
// blocks), so no device-scope fence or atomic is needed inside a kernel: the kernel
// boundary makes the partials visible.  r.z alternates between two buffers by
// iteration parity so that the old and the new value are both readable.
struct CGPart {
    double* pAp;     // per SpMV block
    double* rr;      // per vector block
    double* rz[2];   // per vector block, by parity
    double* bb;      // per vector block (init)
    int nS, nV;      // number of SpMV / vector blocks
};

// ---------------------------------------------------------------------------
// Block reduction helpers (wave64 shuffles + LDS), deterministic order
// ---------------------------------------------------------------------------
__device__ double block_sum(double v) {
    __shared__ double sh[BLK / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BLK / 64; ++i) s += sh[i];
    return s;  // valid in thread 0
}

// DPP move of a double (both 32-bit halves) with the given control and masks; lanes that
// are masked off or read outside their row get 0.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_mov(double v) {
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, BANK_MASK, true);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, BANK_MASK, true);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront into lane 63, in a fixed order: row_shr 1, 2, 4, 8
// build each 16-lane row's sum in its lane 15, row_bcast 15 and 31 carry the rows on.
// Register-to-register (DPP), without the LDS round trips of shuffles.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v = v + dpp_mov<0x111, 0xf, 0xf>(v);   // row_shr:1
    v = v + dpp_mov<0x112, 0xf, 0xf>(v);   // row_shr:2
    v = v + dpp_mov<0x114, 0xf, 0xf>(v);   // row_shr:4
    v = v + dpp_mov<0x118, 0xf, 0xf>(v);   // row_shr:8
    v = v + dpp_mov<0x142, 0xa, 0xf>(v);   // row_bcast:15 into rows 1 and 3
    v = v + dpp_mov<0x143, 0xc, 0xf>(v);   // row_bcast:31 into rows 2 and 3
    return v;                               // the total is in lane 63
}

// Several block sums in one pass (fixed order); results valid in thread 0.
template <int M>
__device__ void block_sum_n(double (&v)[M]) {
    __shared__ double sh[M][BLK / 64];
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = wave_sum_dpp(v[m]);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 63)
#pragma unroll
        for (int m = 0; m < M; ++m) sh[m][wid] = v[m];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double t = 0.0;
            for (int i = 0; i < BLK / 64; ++i) t += sh[m][i];
            v[m] = t;
        }
}

// Sum of partials[0..n) in a fixed order by one block (all threads participate).
__device__ double block_sum_partials(const double* partials, int n) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += BLK) s += partials[i];
    return block_sum(s);
}

// Block-uniform read of the stop flag (thread 0 reads, the block shares it), so that a
// block never diverges around the barriers of the reductions.
__device__ __forceinline__ int ld_done(const CGScal* s) {
    __shared__ int d;
    if (threadIdx.x == 0) d = __atomic_load_n(&s->done, __ATOMIC_RELAXED);
    __syncthreads();
    return d;
}

}  // namespace

#include "assembly.hpp"
#include "assemble1.hpp"

namespace {

// T = double: the step's rows of A, F, b (src/PhysicsCloth.cpp:46-156,
// src/TrainHelper.cpp:19-20), one block per tile (assemble1_block).
template <class T>
__global__ void __launch_bounds__(ASM1_THREADS, 1) k_assemble1(ClothDev C, AsmOut O, double h) {
    assemble1_block<T, false>(C, O, h, blockIdx.x);
}

// The tangent run of every direction in one launch (blockIdx.y: the direction, with its
// own descriptors): one direction's tiles use a quarter of the SIMDs at one wavefront
// each, so the directions go together.  REF: the reference's gradient (every direction
// detached, no dA), so that the Jacobian blocks' derivative parts are not computed.
template <bool REF>
__global__ void __launch_bounds__(NBLK, 1) k_assemble1_dirs(const ClothDev* __restrict__ Cs, const AsmOut* __restrict__ Os,
                                                           double h) {
    assemble1_block<dd, REF>(Cs[blockIdx.y], Os[blockIdx.y], h, blockIdx.x);
}

// Shear segments touching the node (4 own, 2 each as warp neighbour of the left and
// right crossings, 2 each as weft neighbour of the top and bottom crossings), replayed from
// the element-major records, then the Jacobi diagonal and the NaN/Inf screen of the node's
// rows.
template <class T, class Sink>
__device__ void emit_shear_role(Sink& S, const int (&k)[3], int r, const v3<T>& F, const mat3<T> (&B)[3]);
__device__ __forceinline__ int shear_seg(const ClothDev& C, int e, int q);
__device__ __forceinline__ void shear_load(const double* rec, int nseg, int s, int r, d3& F, m33 (&B)[3]);

// The up to 12 segments touching a node, in the order of the crossings (u-1, v), (u, v-1),
// (u, v), (u, v+1), (u+1, v) and, per crossing, of the reference's shear_segs ((left,bot),
// (left,top), (right,bot), (right,top); src/Cloth.cpp:170-181): the crossing's offset from
// the node, the segment q, and the node's role in it (0: centre, 1: warp neighbour, 2: weft
// neighbour).
constexpr int SEG_DU[12] = {-1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1};
constexpr int SEG_DV[12] = {0, 0, -1, -1, 0, 0, 0, 0, 1, 1, 0, 0};
constexpr int SEG_Q[12] = {2, 3, 0, 2, 0, 1, 2, 3, 1, 3, 0, 1};
constexpr int SEG_R[12] = {1, 1, 2, 2, 0, 0, 0, 0, 2, 2, 1, 1};
// Index in SHEAR_O (the 3 x 3 neighbourhood's stencil offsets 0-4, 9-12) of the node at
// (du, dv) from the row's node.
constexpr int shear_oi(int du, int dv) {
    return dv == 0 ? (du == 0 ? 0 : du < 0 ? 1 : 2)
                   : du == 0 ? (dv < 0 ? 3 : 4) : (dv < 0 ? (du < 0 ? 5 : 6) : (du < 0 ? 7 : 8));
}
// The node column of block jj (0: crossing, 1: warp neighbour, 2: weft neighbour) of the
// node's segment j, as a SHEAR_O index.
constexpr int seg_block_oi(int j, int jj) {
    return shear_oi(SEG_DU[j] + (jj == 1 ? ((SEG_Q[j] >> 1) == 0 ? -1 : 1) : 0),
                    SEG_DV[j] + (jj == 2 ? ((SEG_Q[j] & 1) == 0 ? 1 : -1) : 0));
}
constexpr int SHEAR_O[9] = {0, 1, 2, 3, 4, 9, 10, 11, 12};
constexpr int ASM2_THREADS = 3 * TILE_N;

// The value run of the second assembly kernel for one tile: three threads per node, thread
// rc taking the node's L row rc.  The shear writes only the L-L blocks of the 3 x 3 node
// neighbourhood, so row rc receives 27 slots (9 column nodes x 3 components); they start
// from the first kernel's values and collect the segments' contributions in registers, in
// the segments' order (the same sums, in the same order, as one update per contribution),
// with each segment's record loaded while the one before is added.  Thread rc also adds
// the shear force to F and b of its component and writes its Jacobi diagonal entry; thread
// 0 the E rows' entries.  A[3F+v, 3D+u] = B[u][v] (src/MatrixMp.cpp:65-76): row component
// rc takes B[u][rc], u = 0..2.
// ROWS: the block takes 64 consecutive nodes of one grid row (block b: row b / ceil(W / 64),
// nodes from 64 (b mod ceil(W / 64))) instead of an 8 x 8 tile, so that a wavefront's reads of
// a record field (crossing-major) are one contiguous run; the slots of A (tile-major) are the
// smaller stream and take the strided side.  The arithmetic of each node is the same.
template <bool ROWS = false>
__device__ __forceinline__ void assemble2_value(const ClothDev& C, const AsmOut& O, double h, int tile) {
    const int nl = threadIdx.x & (TILE_N - 1), rc = threadIdx.x / TILE_N;
    int k;
    if constexpr (ROWS) {
        const int rpb = (C.W + TILE_N - 1) / TILE_N, row = tile / rpb, uu = (tile - row * rpb) * TILE_N + nl;
        k = uu < C.W ? uu + row * C.W : -1;
    } else {
        k = tile_node(C, tile, nl);
    }
    if (k < 0) return;
    const int u = k % C.W, v = k / C.W;
    const bool inter = interior(C, u, v);
    // -h^2: K_x's coefficient in A (or 1 / 0 for the rows of one matrix alone)
    const double coef = O.only_kind < 0 ? -(h * h) : (O.only_kind == KIND_KX ? 1.0 : 0.0);
    double* gA = O.A + a_base(C.ntx, u, v);
    double acc[27];
#pragma unroll
    for (int oi = 0; oi < 9; ++oi)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) acc[oi * 3 + cc] = gA[(size_t)(rc * 45 + SHEAR_O[oi] * 3 + cc) * TILE_N];
    bool ok[12];
    int sidx[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        const int cu = u + SEG_DU[j], cv = v + SEG_DV[j];
        ok[j] = interior(C, cu, cv);
        sidx[j] = ok[j] ? shear_seg(C, crossing(C, cu, cv), SEG_Q[j]) : 0;
    }
    // a record's part for row rc: F[rc] and B[jj][cc][rc] (planes rec[(r 30 + f) nseg + s])
    double Fb[2], Bb[2][9];
    auto load = [&](int j, int buf) {
        const double* base = O.shear_rec + (size_t)SEG_R[j] * 30 * O.nseg + sidx[j];
        Fb[buf] = base[(size_t)rc * O.nseg];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) Bb[buf][jj * 3 + cc] = base[(size_t)(3 + jj * 9 + cc * 3 + rc) * O.nseg];
    };
    if (ok[0]) load(0, 0);
    double fsum = 0.0;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        if (j + 1 < 12 && ok[j + 1]) load(j + 1, (j + 1) & 1);
        if (!ok[j]) continue;
        fsum += Fb[j & 1];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) acc[seg_block_oi(j, jj) * 3 + cc] += coef * Bb[j & 1][jj * 3 + cc];
    }
    const int pk = padidx(O.Wp, u, v);
    if (O.vpad) {   // the padded starting velocity (the solve's guesses and residuals read it)
        O.vpad[(size_t)rc * O.P + pk] = C.v[(size_t)rc * C.N + k];
        if (rc == 0)
            for (int c = 0; c < 2; ++c) O.vpad[(size_t)(3 + c) * O.P + pk] = inter ? C.uvd[(size_t)c * C.N + k] : 0.0;
    }
    O.F[rc * C.N + k] += fsum;
    const double bv = O.b[rc * O.P + pk] + h * fsum;
    O.b[rc * O.P + pk] = bv;
    bool bad = !isfinite(bv);
    // Jacobi diagonal (the reference solver's DiagonalPreconditioner: 1/a_ii, or 1 if zero)
    // self column (SHEAR_O index 0), component rc (selected: a run-time index would put
    // the accumulators in scratch memory)
    const double d = rc == 0 ? acc[0] : (rc == 1 ? acc[1] : acc[2]);
    O.invdiag[rc * O.P + pk] = d != 0.0 ? 1.0 / d : 1.0;
    if (rc == 0)
        for (int c = 3; c < 5; ++c) {
            double inv = 0.0;
            if (inter) {   // E-E self entries, untouched by the shear
                const double de = gA[(size_t)((c == 3 ? SLOT_EU : SLOT_EV) + 10) * TILE_N];
                inv = de != 0.0 ? 1.0 / de : 1.0;
            }
            O.invdiag[c * O.P + pk] = inv;
            if (!isfinite(O.b[c * O.P + pk])) bad = true;
        }
#pragma unroll
    for (int oi = 0; oi < 9; ++oi)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const double a = acc[oi * 3 + cc];
            if (!isfinite(a)) bad = true;
            gA[(size_t)(rc * 45 + SHEAR_O[oi] * 3 + cc) * TILE_N] = a;
        }
    if (bad) atomicOr(&O.flag[1], 1);
}

// The tangent run (T = dd): one thread per node replays the segments' tangent records
// (k_shear_elem_tan) through the row sink into dF, db and, in exact mode, dA.
template <class T, bool REF = false>
__device__ __forceinline__ void assemble2_node(const ClothDev& C, const AsmOut& O, double h, int k) {
    static_assert(!std::is_same<T, double>::value, "the value run is assemble2_value");
    RowSinkT<T, REF> S;
    S.init(&C, k, nullptr, h, O.dA, O.detach);
    int u = S.mu_, v = S.mv_;
    auto K = [&](int uu, int vv) { return uu + vv * C.W; };
    // shear_segs order per crossing: (left,bot), (left,top), (right,bot), (right,top)
#pragma unroll 1
    for (int ci = 0; ci < 9; ++ci) {
        int cu = u - 1 + ci / 3, cv = v - 1 + ci % 3;
        if (!interior(C, cu, cv)) continue;
        if (cu != u && cv != v) continue;
        int kc = K(cu, cv);
#pragma unroll 1
        for (int sq = 0; sq < 4; ++sq) {
            int a = sq >> 1, bq = sq & 1;
            int k1 = a == 0 ? K(cu - 1, cv) : K(cu + 1, cv);
            int k3 = bq == 0 ? K(cu, cv + 1) : K(cu, cv - 1);
            if (kc != k && k1 != k && k3 != k) continue;
            // replay the segment's tangent record for this node's role (the tangent outputs
            // take only the tangent parts of the shear's forces and K_x blocks, so the
            // values are 0)
            int sidx = shear_seg(C, crossing(C, cu, cv), 2 * a + bq);
            int r = kc == k ? 0 : (k1 == k ? 1 : 2);
            const int kk[3] = {kc, k1, k3};
            d3 F;
            m33 B[3];
            shear_load(O.shear_rec, O.nseg, sidx, r, F, B);
            v3<T> Ft = V3<T>(T(0.0, F.x), T(0.0, F.y), T(0.0, F.z));
            mat3<T> Bt[3];
            for (int j = 0; j < 3; ++j)
                for (int aa = 0; aa < 3; ++aa)
                    for (int bb = 0; bb < 3; ++bb) Bt[j].m[aa][bb] = T(0.0, B[j].m[aa][bb]);
            emit_shear_role<T>(S, kk, r, Ft, Bt);
        }
    }
    int pk = padidx(O.Wp, u, v);
    for (int c = 0; c < 3; ++c) {
        O.dF[c * C.N + k] += tan_of(S.F[c]);
        O.db[c * O.P + pk] += h * tan_of(S.F[c]);
    }
    if (S.bad) atomicOr(&O.flag[0], 1);
}

template <class T, bool REF = false>
__device__ __forceinline__ void assemble2_tile(const ClothDev& C, const AsmOut& O, double h, int tile) {
    int k = tile_node(C, tile);
    if (k >= 0) assemble2_node<T, REF>(C, O, h, k);
}

__global__ void __launch_bounds__(ASM2_THREADS) k_assemble2(ClothDev C, AsmOut O, double h) {
    assemble2_value<true>(C, O, h, blockIdx.x);
}

// The second kernel's tangent run for every direction in one launch (blockIdx.y); REF as
// for k_assemble1_dirs (the shear records then hold no block tangents, and none are read).
template <bool REF>
__global__ void __launch_bounds__(NBLK, 1) k_assemble2_dirs(const ClothDev* __restrict__ Cs, const AsmOut* __restrict__ Os,
                                                           double h) {
    const ClothDev& C = Cs[blockIdx.y];
    const AsmOut& O = Os[blockIdx.y];
    assemble2_tile<dd, REF>(C, O, h, blockIdx.x);
}

}  // namespace

#include "shear.hpp"
#include "mlp.hpp"

namespace {

// ---------------------------------------------------------------------------
// Stencil SpMV.  One block per 8 x 8 tile, 4 wavefronts: wavefront g < 3 computes the
// L row g of the tile's 64 nodes, wavefront 3 both E rows.  The matrix slots of a tile
// are contiguous planes (a_base), so every slot load of a wavefront is 512 contiguous
// bytes.
// ---------------------------------------------------------------------------
struct SpmvArgs {
    const double* A;
    const double* x;   // padded 5-plane
    double* y;         // padded 5-plane
    int W, Lg, N, Wp, P, ntx;
};

// Rows of one node of a tile: L row g (g < 3) into *y0, or both E rows (g == 3) into
// *y0, *y1.  X(c, du, dv) returns entry c of the input vector at the node offset by
// (du, dv).  S(i) returns slot i of the rows: slot g*45 + i for an L row, slot
// SLOT_EU + i (i < 24: the E_u row, then the E_v row) for the E rows.
// FUSED: fused multiply-adds in three (L) or two (E) independent chains, for the latency of
// the persistent solve's product; otherwise one chain of separate products and sums in slot
// order (the reference form, whose iteration counts match the oracle's exactly).
template <bool FUSED = false, class S, class X>
__device__ __forceinline__ void stencil_rows(int g, S slot, X xin, double* y0, double* y1) {
    if constexpr (FUSED) {
        if (g < 3) {
            const int odu[13] = {0, -1, 1, 0, 0, -2, 2, 0, 0, -1, 1, -1, 1};
            const int odv[13] = {0, 0, 0, -1, 1, 0, 0, -2, 2, -1, -1, 1, 1};
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
            for (int o = 0; o < 13; ++o) {
                a0 = __builtin_fma(slot(o * 3 + 0), xin(0, odu[o], odv[o]), a0);
                a1 = __builtin_fma(slot(o * 3 + 1), xin(1, odu[o], odv[o]), a1);
                a2 = __builtin_fma(slot(o * 3 + 2), xin(2, odu[o], odv[o]), a2);
            }
            a0 = __builtin_fma(slot(SLOT_LE + 0), xin(3, -1, 0), a0);
            a1 = __builtin_fma(slot(SLOT_LE + 1), xin(3, 0, 0), a1);
            a2 = __builtin_fma(slot(SLOT_LE + 2), xin(3, 1, 0), a2);
            a0 = __builtin_fma(slot(SLOT_LE + 3), xin(4, 0, -1), a0);
            a1 = __builtin_fma(slot(SLOT_LE + 4), xin(4, 0, 0), a1);
            a2 = __builtin_fma(slot(SLOT_LE + 5), xin(4, 0, 1), a2);
            *y0 = (a0 + a1) + a2;
        } else {
            double au0 = 0.0, au1 = 0.0, av0 = 0.0, av1 = 0.0;
#pragma unroll
            for (int d = -1; d <= 1; ++d) {
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    if (cc & 1) {
                        au1 = __builtin_fma(slot((d + 1) * 3 + cc), xin(cc, d, 0), au1);
                        av1 = __builtin_fma(slot(12 + (d + 1) * 3 + cc), xin(cc, 0, d), av1);
                    } else {
                        au0 = __builtin_fma(slot((d + 1) * 3 + cc), xin(cc, d, 0), au0);
                        av0 = __builtin_fma(slot(12 + (d + 1) * 3 + cc), xin(cc, 0, d), av0);
                    }
                }
                au1 = __builtin_fma(slot(9 + d + 1), xin(3, d, 0), au1);
                av1 = __builtin_fma(slot(12 + 9 + d + 1), xin(4, 0, d), av1);
            }
            *y0 = au0 + au1;
            *y1 = av0 + av1;
        }
        return;
    }
    if (g < 3) {
        const int odu[13] = {0, -1, 1, 0, 0, -2, 2, 0, 0, -1, 1, -1, 1};
        const int odv[13] = {0, 0, 0, -1, 1, 0, 0, -2, 2, -1, -1, 1, 1};
        double acc = 0.0;
#pragma unroll
        for (int o = 0; o < 13; ++o) {
            acc += slot(o * 3 + 0) * xin(0, odu[o], odv[o]);
            acc += slot(o * 3 + 1) * xin(1, odu[o], odv[o]);
            acc += slot(o * 3 + 2) * xin(2, odu[o], odv[o]);
        }
        acc += slot(SLOT_LE + 0) * xin(3, -1, 0);
        acc += slot(SLOT_LE + 1) * xin(3, 0, 0);
        acc += slot(SLOT_LE + 2) * xin(3, 1, 0);
        acc += slot(SLOT_LE + 3) * xin(4, 0, -1);
        acc += slot(SLOT_LE + 4) * xin(4, 0, 0);
        acc += slot(SLOT_LE + 5) * xin(4, 0, 1);
        *y0 = acc;
    } else {
        double au = 0.0, av = 0.0;
#pragma unroll
        for (int d = -1; d <= 1; ++d) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                au += slot((d + 1) * 3 + cc) * xin(cc, d, 0);
                av += slot(12 + (d + 1) * 3 + cc) * xin(cc, 0, d);
            }
            au += slot(9 + d + 1) * xin(3, d, 0);
            av += slot(12 + 9 + d + 1) * xin(4, 0, d);
        }
        *y0 = au;
        *y1 = av;
    }
}

__device__ __forceinline__ int row_slot0(int g) { return g < 3 ? g * 45 : SLOT_EU; }

// y = A x on one tile; optionally partial sums of x . y per block (one block per tile)
__device__ __forceinline__ void spmv_tile(const SpmvArgs& a, double* partials, int tile) {
    int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    int u = (tile % a.ntx) * TILE + (lane & 7), v = (tile / a.ntx) * TILE + (lane >> 3);
    double dotp = 0.0;
    if (u < a.W && v < a.Lg) {
        const double* At = a.A + a_base(a.ntx, u, v);
        const int p0 = padidx(a.Wp, u, v);
        const double* x = a.x;
        const int Wp = a.Wp, P = a.P;
        auto xin = [&](int c, int du, int dv) { return x[c * P + p0 + du + dv * Wp]; };
        const double* Ag = At + (size_t)row_slot0(g) * TILE_N;
        auto slot = [&](int i) { return Ag[(size_t)i * TILE_N]; };
        double y0 = 0.0, y1 = 0.0;
        stencil_rows(g, slot, xin, &y0, &y1);
        if (g < 3) {
            a.y[g * P + p0] = y0;
            dotp = y0 * x[g * P + p0];
        } else {
            a.y[3 * P + p0] = y0;
            a.y[4 * P + p0] = y1;
            dotp = y0 * x[3 * P + p0] + y1 * x[4 * P + p0];
        }
    }
    if (partials) {
        double s = block_sum(dotp);
        if (threadIdx.x == 0) partials[tile] = s;
    }
}

__global__ void __launch_bounds__(BLK) k_spmv(SpmvArgs a, double* partials, const CGScal* sc, int check_done) {
    if (check_done && ld_done(sc)) return;
    spmv_tile(a, partials, blockIdx.x);
}

// ---------------------------------------------------------------------------
// CG vector kernels (Jacobi-PCG in the reference solver's order, src/Solvers.cpp:22-24)
// ---------------------------------------------------------------------------
struct VecArgs {
    double* x; double* r; double* pb[2]; double* y; const double* b; const double* invd;
    int W, Lg, N, Wp, P;
};

// thread -> (plane, node) over the 5N real entries; returns the padded index
__device__ __forceinline__ int vec_index(const VecArgs& a, int t, int& c) {
    c = t / a.N;
    int k = t - c * a.N;
    int u = k % a.W, v = k / a.W;
    return c * a.P + padidx(a.Wp, u, v);
}

// init: r = b - y (y = A x0 when has_x0, else x = 0, r = b); p = D^-1 r; partials of
// r.r, r.z (into parity buffer 1, read by iteration 0) and b.b
__device__ __forceinline__ void cg_init_block(const VecArgs& a, const CGPart& P, int has_x0, int blk) {
    int t = blk * BLK + threadIdx.x;
    double rr = 0.0, rz = 0.0, bb = 0.0;
    if (t < 5 * a.N) {
        int c;
        int i = vec_index(a, t, c);
        double bi = a.b[i];
        double ri = has_x0 ? bi - a.y[i] : bi;
        if (!has_x0) a.x[i] = 0.0;
        double zi = a.invd[i] * ri;
        a.r[i] = ri;
        a.pb[0][i] = zi;
        rr = ri * ri; rz = ri * zi; bb = bi * bi;
    }
    double s1 = block_sum(rr);
    double s2 = block_sum(rz);
    double s3 = block_sum(bb);
    if (threadIdx.x == 0) { P.rr[blk] = s1; P.rz[1][blk] = s2; P.bb[blk] = s3; }
}

__global__ void __launch_bounds__(BLK) k_cg_init(VecArgs a, CGPart P, int has_x0) { cg_init_block(a, P, has_x0, blockIdx.x); }

// one block: stopping threshold and the initial decision (src/Solvers.cpp:22-24 semantics)
__global__ void __launch_bounds__(BLK) k_cg_init2(CGPart P, CGScal* sc, double tol, int maxit) {
    double rr = block_sum_partials(P.rr, P.nV);
    double bb = block_sum_partials(P.bb, P.nV);
    if (threadIdx.x == 0) {
        sc->rr = rr; sc->bb = bb;
        double thr = tol * tol * bb;
        sc->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
        sc->itc[0] = 0; sc->itc[1] = 0; sc->maxit = maxit; sc->zero_rhs = bb == 0.0;
        sc->done = (bb == 0.0 || rr < sc->thr || maxit <= 0) ? 1 : 0;
    }
}

__device__ __forceinline__ double bcast(double v) {
    __shared__ double s;
    if (threadIdx.x == 0) s = v;
    __syncthreads();
    double r = s;
    __syncthreads();
    return r;
}

// alpha = r.z / p.Ap; x += alpha p; r -= alpha y; partials of |r|^2 and r.z
__global__ void __launch_bounds__(BLK) k_cg_xr(VecArgs a, CGPart P, const CGScal* sc, int q) {
    if (ld_done(sc)) return;
    double pAp = bcast(block_sum_partials(P.pAp, P.nS));
    double rz = bcast(block_sum_partials(P.rz[1 - q], P.nV));
    double alpha = rz / pAp;
    int t = blockIdx.x * BLK + threadIdx.x;
    double rr = 0.0, rzn = 0.0;
    if (t < 5 * a.N) {
        int c;
        int i = vec_index(a, t, c);
        a.x[i] += alpha * a.pb[0][i];
        double ri = a.r[i] - alpha * a.y[i];
        a.r[i] = ri;
        rr = ri * ri;
        rzn = ri * (a.invd[i] * ri);
    }
    double s1 = block_sum(rr);
    double s2 = block_sum(rzn);
    if (threadIdx.x == 0) { P.rr[blockIdx.x] = s1; P.rz[q][blockIdx.x] = s2; }
}

// Stopping test on the |r|^2 of k_cg_xr, beta, p = D^-1 r + beta p.  Block 0 keeps the
// iteration count in the parity slot q (readers use slot 1-q, so no block ever reads
// a slot being written in the same launch).
__global__ void __launch_bounds__(BLK) k_cg_p(VecArgs a, CGPart Pt, CGScal* sc, int q) {
    if (ld_done(sc)) return;
    double rr = bcast(block_sum_partials(Pt.rr, Pt.nV));
    double rzn = bcast(block_sum_partials(Pt.rz[q], Pt.nV));
    double rzo = bcast(block_sum_partials(Pt.rz[1 - q], Pt.nV));
    bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr < sc->thr) {
        if (lead) { sc->rr = rr; sc->done = 1; }
        return;
    }
    int itn = sc->itc[1 - q] + 1;
    if (lead) { sc->rr = rr; sc->itc[q] = itn; if (itn >= sc->maxit) sc->done = 1; }
    double beta = rzn / rzo;
    int t = blockIdx.x * BLK + threadIdx.x;
    if (t < 5 * a.N) {
        int c;
        int i = vec_index(a, t, c);
        a.pb[0][i] = a.invd[i] * a.r[i] + beta * a.pb[0][i];
    }
}

// ---------------------------------------------------------------------------
// Single-reduction Jacobi-PCG: the reference solver's recurrences (src/Solvers.cpp:22-24,
// Eigen's ConjugateGradient with DiagonalPreconditioner) with one kernel and one grid-wide
// reduction per iteration.  With u = D^-1 r, w = A u, s = A p kept as vectors,
//   p_j = u_j + beta_j p_{j-1}      s_j = w_j + beta_j s_{j-1}      (= A p_j by linearity)
//   x_{j+1} = x_j + alpha_j p_j     r_{j+1} = r_j - alpha_j s_j
//   u_{j+1} = D^-1 r_{j+1}          w_{j+1} = A u_{j+1}
//   beta_j = gamma_j / gamma_{j-1},  gamma = r.u,
//   alpha_j = gamma_j / p_j.A p_j, with p_j.A p_j expanded as
//     u_j.w_j + beta_j (u_j.s_{j-1} + p_{j-1}.w_j) + beta_j^2 p_{j-1}.s_{j-1},
// whose five inner products are all known at the end of the previous iteration.  Nothing
// assumes A symmetric (it is not: the reference writes the L-L blocks transposed), so the
// iterates are the reference's up to rounding, with the same stopping rule on the
// recursive |r|^2 < tol^2 |b|^2 and the same iteration count convention.
// A block owns one 8 x 8 tile.  It recomputes the element-wise recurrences on the
// tile's 2-node halo itself (bit-identical to the owner's values) and keeps
// u_{j+1} of tile + halo in LDS for the stencil product, so no second grid-wide
// synchronisation is needed.  Vectors read on the halo are double-buffered by
// iteration parity; u is never stored.
// ---------------------------------------------------------------------------
constexpr int HALO_W = TILE + 4;
constexpr int HALO_N = HALO_W * HALO_W;

struct Cg1Args {
    const double* A; const double* invd;
    double* x;
    double* r[2]; double* w[2]; double* p[2]; double* s[2];
    double* part[2];        // per tile, 6 x (KP * BLK) by parity: r.u, u.w, u.s, p.w, p.s, r.r
    const double* bbpart;   // per vector block of k_cg_init: b.b
    const double* rrpart;   // per vector block of k_cg_init: r_0.r_0
    int nbb;
    CGScal* sc;
    int W, Lg, N, Wp, P, ntx, nT;
};

// init = 1: w_0 = A D^-1 r_0 and the partials of the next iteration (r_0 in r[0]; p and
// s count as zero); block 0 also sets the stopping threshold max(tol^2 |b|^2, DBL_MIN)
// and maxit.  init = 0: iteration j with parity q = j & 1 (reads r, w from [q] and p, s
// from [1-q]; writes the opposite); tol and maxit are not used.
// Every load that does not depend on alpha and beta -- the tile's matrix rows, the
// partials, the scalars and the vector entries of tile + halo -- is issued before the
// first barrier, so one memory latency covers them all.
constexpr int CG1_NE = (5 * HALO_N + BLK - 1) / BLK;   // vector entries per thread
constexpr int CG1_MAXP = 4;                              // partial rounds: nT <= 4 * BLK

// KP = ceil(nT / BLK): partial sums per thread.  The partial arrays are KP * BLK long
// per quantity and zero past nT, so every load below is unconditional (a conditional
// load would make the compiler wait for it inside its branch).
template <int KP, int q>
__global__ void __launch_bounds__(BLK) k_cg1(Cg1Args a, int init, double tol, int maxit) {
    __shared__ double su[5][HALO_N];
    __shared__ double sp[5][TILE_N];
    __shared__ double sred[8];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int tu0 = (tile % a.ntx) * TILE, tv0 = (tile / a.ntx) * TILE;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int rq = init ? 0 : q;           // parity of r, w read
    const int PS = KP * BLK;               // stride of the partial arrays
    // (a) scalars of the previous iteration and the partials (first, so that waiting
    // for them does not wait for the larger loads issued after them).  They are read
    // in the init launch too and ignored there, so that no load sits in a branch.
    const CGScal* sc = a.sc;
    const int done = sc->done, j = sc->itc[1 - q], s_maxit = sc->maxit;
    const double s_thr = sc->thr, gprev = sc->gamma[1 - q];
    const double* pp = a.part[q];
    double pv[KP][6];
#pragma unroll
    for (int k = 0; k < KP; ++k)
#pragma unroll
        for (int m = 0; m < 6; ++m) pv[k][m] = pp[m * PS + threadIdx.x + k * BLK];
    // (b) this thread's matrix rows into registers (a node outside the grid reads the
    // tile's first node and discards the result)
    const int nu = tu0 + (lane & 7), nv = tv0 + (lane >> 3);
    const bool node_in = nu < a.W && nv < a.Lg;
    double Ar[45];
    {
        const double* Ag = a.A + a_base(a.ntx, node_in ? nu : tu0, node_in ? nv : tv0) + (size_t)row_slot0(g) * TILE_N;
        if (g < 3) {
#pragma unroll
            for (int i = 0; i < 45; ++i) Ar[i] = Ag[(size_t)i * TILE_N];
        } else {
#pragma unroll
            for (int i = 0; i < 24; ++i) Ar[i] = Ag[(size_t)i * TILE_N];
#pragma unroll
            for (int i = 24; i < 45; ++i) Ar[i] = 0.0;
        }
    }
    // (c) vector entries of tile + halo.  Entries outside the grid read padded index 0
    // (a padding corner, always zero).
    const double* r_in = a.r[rq];
    const double* w_in = a.w[rq];
    const double* s_in = a.s[1 - q];
    const double* p_in = a.p[1 - q];
    double e_id[CG1_NE], e_r[CG1_NE], e_w[CG1_NE], e_s[CG1_NE], e_p[CG1_NE], e_x[CG1_NE];
    size_t e_i[CG1_NE];
    bool e_in[CG1_NE], e_own[CG1_NE];
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t) {
        int e = threadIdx.x + t * BLK;
        int c = e / HALO_N, ln = e - c * HALO_N;
        int lx = ln % HALO_W, ly = ln / HALO_W;
        int u = tu0 - 2 + lx, v = tv0 - 2 + ly;
        e_own[t] = e < 5 * HALO_N && lx >= 2 && lx < TILE + 2 && ly >= 2 && ly < TILE + 2;
        e_in[t] = e < 5 * HALO_N && u >= 0 && u < a.W && v >= 0 && v < a.Lg;
        e_i[t] = e_in[t] ? (size_t)c * a.P + padidx(a.Wp, u, v) : 0;
        const size_t io = e_own[t] ? e_i[t] : 0;
        e_id[t] = a.invd[e_i[t]];
        e_r[t] = r_in[e_i[t]];
        e_w[t] = w_in[e_i[t]];
        e_s[t] = s_in[e_i[t]];
        e_p[t] = p_in[io];
        e_x[t] = a.x[io];
    }
    // (d) reduce the partials; every thread gets the sums
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (init) {
        for (int i = threadIdx.x; i < a.nbb; i += BLK) { acc[0] += a.bbpart[i]; acc[1] += a.rrpart[i]; }
    } else {
#pragma unroll
        for (int k = 0; k < KP; ++k)
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[m] += pv[k][m];
    }
    block_sum_n(acc);
    if (threadIdx.x == 0) {
        for (int m = 0; m < 6; ++m) sred[m] = acc[m];
        sred[6] = (double)done;
    }
    __syncthreads();
    double alpha = 0.0, beta = 0.0;
    bool first = false;
    if (init) {
        if (lead) {
            CGScal* sc = a.sc;
            double bb = sred[0];
            sc->bb = bb;
            sc->rr = sred[1];
            double thr = tol * tol * bb;
            sc->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
            sc->maxit = maxit;
            sc->itc[0] = 0; sc->itc[1] = 0; sc->fin_iters = 0;
            sc->zero_rhs = bb == 0.0;
            sc->done = (bb == 0.0 || maxit <= 0) ? 1 : 0;
        }
    } else {
        if (sred[6] != 0.0) return;   // block-uniform
        const double gam = sred[0], d1 = sred[1], d2 = sred[2], d3 = sred[3], d4 = sred[4], rr = sred[5];
        if (rr < s_thr || j >= s_maxit) {
            // Stop: r_j met the threshold (the reference loop broke in iteration j-1,
            // which it does not count) or maxit iterations are done.
            if (lead) {
                a.sc->rr = rr;
                a.sc->fin_iters = rr < s_thr ? (j > 0 ? j - 1 : 0) : j;
                a.sc->done = 1;
            }
            return;
        }
        first = j == 0;
        beta = first ? 0.0 : gam / gprev;
        const double pAp = first ? d1 : d1 + beta * (d2 + d3) + beta * beta * d4;
        alpha = gam / pAp;
        if (lead) { a.sc->gamma[q] = gam; a.sc->alpha[q] = alpha; a.sc->itc[q] = j + 1; a.sc->rr = rr; }
    }
    // (e) element-wise recurrences on tile + halo; u_{j+1} and p_j into LDS.  The
    // global stores wait until after the LDS barrier, which then does not wait for them.
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] = 0.0;   // r.u, u.w, u.s, p.w, p.s, r.r
    double o_x[CG1_NE], o_p[CG1_NE], o_s[CG1_NE], o_r[CG1_NE];
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t) {
        o_x[t] = o_p[t] = o_s[t] = o_r[t] = 0.0;
        int e = threadIdx.x + t * BLK;
        if (e >= 5 * HALO_N) continue;
        int c = e / HALO_N, ln = e - c * HALO_N;
        double sj = init ? 0.0 : (first ? e_w[t] : e_w[t] + beta * e_s[t]);
        double rn = e_r[t] - alpha * sj;
        double un = e_id[t] * rn;
        double pj = 0.0;
        if (e_in[t] && e_own[t]) {
            if (!init) {
                pj = e_id[t] * e_r[t];
                if (!first) pj += beta * e_p[t];
                o_x[t] = e_x[t] + alpha * pj;
                o_p[t] = pj;
                o_s[t] = sj;
                o_r[t] = rn;
            }
            acc[0] += rn * un;
            acc[2] += un * sj;
            acc[4] += pj * sj;
            acc[5] += rn * rn;
        }
        su[c][ln] = un;
        if (e_own[t]) {
            int lx = ln % HALO_W, ly = ln / HALO_W;
            sp[c][(lx - 2) + (ly - 2) * TILE] = pj;
        }
    }
    __syncthreads();
    // (f) w_{j+1} = A u_{j+1} on the tile's own nodes, from the registers
    double y0 = 0.0, y1 = 0.0;
    if (node_in) {
        const int l0 = (2 + (lane & 7)) + (2 + (lane >> 3)) * HALO_W;
        auto xin = [&](int c, int du, int dv) { return su[c][l0 + du + dv * HALO_W]; };
        auto slot = [&](int i) { return Ar[i]; };
        stencil_rows(g, slot, xin, &y0, &y1);
        if (g < 3) {
            acc[1] += y0 * su[g][l0];
            acc[3] += y0 * sp[g][lane];
        } else {
            acc[1] += y0 * su[3][l0] + y1 * su[4][l0];
            acc[3] += y0 * sp[3][lane] + y1 * sp[4][lane];
        }
    }
    block_sum_n(acc);
    // (g) all global stores last, so that no barrier above waits for them
    double* part_out = a.part[init ? 0 : 1 - q];
    if (threadIdx.x == 0)
        for (int m = 0; m < 6; ++m) part_out[m * PS + blockIdx.x] = acc[m];
    if (node_in) {
        double* w_out = a.w[init ? 0 : 1 - q];
        const int p0 = padidx(a.Wp, nu, nv);
        if (g < 3) {
            w_out[g * a.P + p0] = y0;
        } else {
            w_out[3 * a.P + p0] = y0;
            w_out[4 * a.P + p0] = y1;
        }
    }
    if (!init) {
        double* r_out = a.r[1 - q];
        double* s_out = a.s[q];
        double* p_out = a.p[q];
#pragma unroll
        for (int t = 0; t < CG1_NE; ++t)
            if (e_in[t] && e_own[t]) {
                a.x[e_i[t]] = o_x[t];
                p_out[e_i[t]] = o_p[t];
                s_out[e_i[t]] = o_s[t];
                r_out[e_i[t]] = o_r[t];
            }
    }
}

// ---------------------------------------------------------------------------
// Persistent form of the single-reduction CG: one launch runs the whole solve, its
// blocks all resident at once (launch_cgp).  Each block owns one 8 x 8 tile for all iterations and keeps the tile's rows
// of A (the 20 MB that every per-iteration kernel would otherwise stream from memory
// again), D^-1, x and p in registers.  Blocks exchange only r, w, s on the tile
// halos and their six partial sums per iteration.
// Cross-block data goes through relaxed device-scope loads and stores (sc1: coherent
// across the XCDs without L2 write-back or invalidation).  A block publishes an
// iteration by storing its partials, waiting until all of its stores have completed
// (s_waitcnt 0 in every thread, then a barrier), and only then storing its tag
// (solve epoch << 32 | stage) in the same 64-byte slot.  A block starts iteration j
// when every slot of parity j & 1 carries stage j; since no block can then be more than
// one stage ahead, parity double buffers suffice.  Every wait is bounded: a block that
// spins longer than spin_limit sets err and all blocks leave at their next wait.
// ---------------------------------------------------------------------------
struct CgSlot {
    double v[6];
    unsigned long long tag;
    unsigned long long pad;
};

struct CgpArgs {
    const double* A; const double* invd; const double* bbpart; const double* rrpart;
    const double* brow;     // block-Jacobi rows (20 per entry, k_bj_factor) or null: Jacobi
    double* x;
    const double* r0;       // r_0 = b - A x_0 (k_cg_init)
    double* w[2];           // w_j = A M^-1 r_j of every tile, by iteration parity
    CgSlot* slot[2];
    CGScal* sc;
    int* err;
    unsigned long long epoch;
    double tol;
    int maxit, nbb, spin_limit;
    int W, Lg, Wp, P, ntx, nT;
};

__device__ __forceinline__ double ld_dev(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until every slot of `sl` carries `want` (or the spin limit / a peer's error).
// Returns false, block-uniformly, when the block must leave.
template <int KP>
__device__ bool cgp_wait(const CgpArgs& a, const CgSlot* sl, unsigned long long want) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        int i = threadIdx.x + k * BLK;
        if (i >= a.nT) continue;
        int spins = 0;
        while (__hip_atomic_load(&sl[i].tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
            if (++spins > a.spin_limit || __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    int bad = __syncthreads_or(!ok);
    if (bad && threadIdx.x == 0) __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return !bad;
}

// Block sums of six values that every thread receives: wavefront sums (DPP) into
// sh[m][wave], one barrier, then every thread adds the four wavefront sums in the same
// order (the order thread 0 of block_sum_n uses, so the sums are the same bits).  sh
// must not be read by a slower thread from an earlier call: the caller separates two
// uses of one sh by at least one other barrier.
__device__ __forceinline__ void block_allsum6(double (&v)[6], double (*sh)[BLK / 64]) {
#pragma unroll
    for (int m = 0; m < 6; ++m) v[m] = wave_sum_dpp(v[m]);
    if ((threadIdx.x & 63) == 63)
#pragma unroll
        for (int m = 0; m < 6; ++m) sh[m][threadIdx.x >> 6] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < BLK / 64; ++i) t += sh[m][i];
        v[m] = t;
    }
}

// Block-sums this block's six partials and has thread 0 store them into the block's
// slot (the tag follows in cgp_tag, once every store of the stage has completed).
__device__ __forceinline__ void cgp_partials(CgSlot* sl, double (&acc)[6], double (*sh)[BLK / 64], int tile) {
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] = wave_sum_dpp(acc[m]);
    if ((threadIdx.x & 63) == 63)
#pragma unroll
        for (int m = 0; m < 6; ++m) sh[m][threadIdx.x >> 6] = acc[m];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < BLK / 64; ++i) t += sh[m][i];
            st_dev(&sl[tile].v[m], t);
        }
}

// Publishes the stage: every thread waits for its own stores (partials and vectors), one
// barrier, then thread 0 stores the tag in the block's slot.
__device__ __forceinline__ void cgp_tag(CgSlot* sl, unsigned long long tag, int tile) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(&sl[tile].tag, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The solve of one tile's block; `tile` numbers the block within its system.
// BJ: the 2x2-node block-Jacobi preconditioner (rows in a.brow) in place of D^-1.  Its
// groups sit at even grid coordinates, and a tile's halo starts at an even coordinate
// (tile origin - 2) and spans whole groups, so u = M^-1 r on the halo needs only r on the
// halo: the block computes it locally, like D^-1 r, after one more barrier.
template <int KP, bool BJ = false>
__device__ __forceinline__ void cgp_tile(const CgpArgs& a, int tile, unsigned long long epoch) {
    __shared__ double su[5][HALO_N];
    __shared__ double sr[BJ ? 5 : 1][BJ ? HALO_N : 1];
    __shared__ double sp[5][TILE_N];
    __shared__ double sred[8];
    __shared__ double shA[6][BLK / 64], shB[6][BLK / 64];   // the two block sums of a stage
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int tu0 = (tile % a.ntx) * TILE, tv0 = (tile / a.ntx) * TILE;
    const bool lead = tile == 0 && threadIdx.x == 0;
    // resident data: matrix rows, entry indices, D^-1, x and p of the owned entries
    const int nu = tu0 + (lane & 7), nv = tv0 + (lane >> 3);
    const bool node_in = nu < a.W && nv < a.Lg;
    const int p0 = node_in ? padidx(a.Wp, nu, nv) : 0;
    const int l0 = (2 + (lane & 7)) + (2 + (lane >> 3)) * HALO_W;
    double Ar[45];
    {
        const double* Ag = a.A + a_base(a.ntx, node_in ? nu : tu0, node_in ? nv : tv0) + (size_t)row_slot0(g) * TILE_N;
        if (g < 3) {
#pragma unroll
            for (int i = 0; i < 45; ++i) Ar[i] = Ag[(size_t)i * TILE_N];
        } else {
#pragma unroll
            for (int i = 0; i < 24; ++i) Ar[i] = Ag[(size_t)i * TILE_N];
#pragma unroll
            for (int i = 24; i < 45; ++i) Ar[i] = 0.0;
        }
    }
    int e_i[CG1_NE], e_ln[CG1_NE], e_c[CG1_NE];
    bool e_in[CG1_NE], e_own[CG1_NE];
    double e_id[CG1_NE], x_reg[CG1_NE], p_reg[CG1_NE];
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t) {
        int e = threadIdx.x + t * BLK;
        int c = e / HALO_N, ln = e - c * HALO_N;
        int lx = ln % HALO_W, ly = ln / HALO_W;
        int u = tu0 - 2 + lx, v = tv0 - 2 + ly;
        e_c[t] = c; e_ln[t] = ln;
        e_own[t] = e < 5 * HALO_N && lx >= 2 && lx < TILE + 2 && ly >= 2 && ly < TILE + 2;
        e_in[t] = e < 5 * HALO_N && u >= 0 && u < a.W && v >= 0 && v < a.Lg;
        e_i[t] = e_in[t] ? c * a.P + padidx(a.Wp, u, v) : 0;
        e_id[t] = a.invd[e_i[t]];
        x_reg[t] = a.x[e_own[t] ? e_i[t] : 0];
        p_reg[t] = 0.0;
    }
    // block-Jacobi apply, by group: the halo's 6 x 6 groups of 20 DOFs, 7 threads per
    // group (3 rows each, 2 for the last), so that a thread reads its group's 20 inputs
    // once.  brow[o]: the 20 coefficients of the thread's o-th row (over node q = du + 2 dv
    // from the group's origin and component c: coefficient q * 5 + c), zero for nodes
    // outside the grid; bj_c / bj_ln: where its result goes in su.
    constexpr int BJ_T = 7, BJ_GROUPS = (HALO_W / 2) * (HALO_W / 2);
    static_assert(BJ_T * BJ_GROUPS <= BLK && 3 * BJ_T >= 20 && HALO_W % 2 == 0, "block-Jacobi thread map");
    double brow[3][BJ ? 20 : 1];
    int bj_c[3], bj_ln[3], bj_ln0 = 0;
    bool bj_ok[3];
    double u_reg[CG1_NE];   // u_j = M^-1 r_j of the owned entries, from the previous stage
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t) u_reg[t] = 0.0;
    {
        const int gi = threadIdx.x / BJ_T, jt = threadIdx.x - BJ_T * gi;
        const int gx0 = 2 * (gi % (HALO_W / 2)), gy0 = 2 * (gi / (HALO_W / 2));
        bj_ln0 = gi < BJ_GROUPS ? gx0 + gy0 * HALO_W : 0;   // spare threads read group 0
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const int oi = 3 * jt + o, qn = oi / 5, c = oi - 5 * qn;
            const int lx = gx0 + (qn & 1), ly = gy0 + (qn >> 1);
            const int u = tu0 - 2 + lx, v = tv0 - 2 + ly;
            bj_ok[o] = gi < BJ_GROUPS && oi < 20;
            bj_c[o] = c;
            bj_ln[o] = lx + ly * HALO_W;
            const bool in = bj_ok[o] && u >= 0 && u < a.W && v >= 0 && v < a.Lg;
            if constexpr (BJ) {
                const double* src = a.brow + (size_t)(in ? c * a.P + padidx(a.Wp, u, v) : 0) * 20;
#pragma unroll
                for (int m = 0; m < 20; ++m) brow[o][m] = in ? src[m] : 0.0;
            }
        }
    }
    // stopping threshold from the b.b and r_0.r_0 partials of k_cg_init (every block
    // reduces them in the same order, so all blocks hold the same values)
    double bb, rr0;
    {
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = threadIdx.x; i < a.nbb; i += BLK) { acc[0] += a.bbpart[i]; acc[1] += a.rrpart[i]; }
        block_sum_n(acc);
        if (threadIdx.x == 0) { sred[0] = acc[0]; sred[1] = acc[1]; }
        __syncthreads();
        bb = sred[0]; rr0 = sred[1];
    }
    double thr = a.tol * a.tol * bb;
    thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || a.maxit <= 0 || !isfinite(bb) || !isfinite(rr0)) {
        if (lead) {
            a.sc->bb = bb; a.sc->rr = rr0; a.sc->thr = thr; a.sc->maxit = a.maxit;
            a.sc->zero_rhs = bb == 0.0; a.sc->fin_iters = 0; a.sc->itc[0] = a.sc->itc[1] = 0; a.sc->done = 1;
            if (!isfinite(bb) || !isfinite(rr0)) a.sc->nonfinite = 1;
            else if (bb != 0.0 && !(rr0 < thr)) a.sc->nonconv += 1;
        }
        if (bb == 0.0)   // the reference solver returns x = 0 for a zero right-hand side
#pragma unroll
            for (int t = 0; t < CG1_NE; ++t)
                if (e_in[t] && e_own[t]) a.x[e_i[t]] = 0.0;
        return;
    }
    // r_j and s_{j-1} on tile + halo stay in registers from stage to stage: the block
    // computes the recurrences of its halo entries itself, from the same inputs and in
    // the same operations as their owners, so only w_j (a stencil product) crosses
    // tiles.  Entries outside the grid are held at zero.  r_0 comes from k_cg_init.
    double e_r[CG1_NE], e_s[CG1_NE];
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t) {
        e_r[t] = e_in[t] ? a.r0[e_i[t]] : 0.0;
        e_s[t] = 0.0;
    }
    // stage -1: w_0 = A D^-1 r_0 and the partials of iteration 0
    double alpha = 0.0, beta = 0.0, gprev = 1.0;
    for (int j = -1;; ++j) {
        const int q = j & 1;
        const bool init = j < 0;
        const bool first = j == 0;
        if (!init && !cgp_wait<KP>(a, a.slot[q], (epoch << 32) | (unsigned long long)j)) return;   // error
        // one round trip for the partials of every tile and w_j on tile + halo; loads
        // whose values an init stage ignores are still issued (their buffers exist), so
        // that none of them sits in a branch
        const CgSlot* sl = a.slot[q];
        double pv[KP][6];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            int i = threadIdx.x + k * BLK;
            int ic = i < a.nT ? i : a.nT - 1;
#pragma unroll
            for (int m = 0; m < 6; ++m) pv[k][m] = ld_dev(&sl[ic].v[m]);
        }
        const double* w_in = a.w[q];
        double e_w[CG1_NE];
#pragma unroll
        for (int t = 0; t < CG1_NE; ++t) {
            double wv = ld_dev(w_in + e_i[t]);
            e_w[t] = e_in[t] ? wv : 0.0;
        }
        if (!init) {
            double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < KP; ++k)
                if (threadIdx.x + k * BLK < a.nT)
#pragma unroll
                    for (int m = 0; m < 6; ++m) acc[m] += pv[k][m];
            block_allsum6(acc, shA);
            const double gam = acc[0], d1 = acc[1], d2 = acc[2], d3 = acc[3], d4 = acc[4], rr = acc[5];
            const bool finite = isfinite(rr);
            if (rr < thr || j >= a.maxit || !finite) {
                if (lead) {
                    a.sc->bb = bb; a.sc->rr = rr; a.sc->thr = thr; a.sc->maxit = a.maxit; a.sc->zero_rhs = 0;
                    a.sc->fin_iters = rr < thr ? (j > 0 ? j - 1 : 0) : j;
                    a.sc->itc[0] = a.sc->itc[1] = a.sc->fin_iters;
                    a.sc->done = 1;
                    a.sc->total_iters += a.sc->fin_iters;
                    if (!finite) a.sc->nonfinite = 1;
                    else if (!(rr < thr)) a.sc->nonconv += 1;
                }
                break;
            }
            beta = first ? 0.0 : gam / gprev;
            const double pAp = first ? d1 : d1 + beta * (d2 + d3) + beta * beta * d4;
            alpha = gam / pAp;
            gprev = gam;
        }
        if (init || first) {
#pragma unroll
            for (int t = 0; t < CG1_NE; ++t) {
                e_s[t] = 0.0;
                if (init) e_w[t] = 0.0;
            }
        }
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // r.u, u.w, u.s, p.w, p.s, r.r
        double k_rn[CG1_NE], k_sj[CG1_NE];
        if constexpr (BJ) {
            // r_{j+1} on tile + halo into LDS, then u_{j+1} = M^-1 r_{j+1} into su by group
            // (every su entry of tile + halo is written: 36 groups x 20 rows); the owned
            // entries' dot products with u follow the barrier after the main pass
#pragma unroll
            for (int t = 0; t < CG1_NE; ++t)
                if (threadIdx.x + t * BLK < 5 * HALO_N) {
                    double sj = e_w[t] + beta * e_s[t];
                    sr[e_c[t]][e_ln[t]] = e_r[t] - alpha * sj;
                }
            __syncthreads();
            double in[20];
#pragma unroll
            for (int qn = 0; qn < 4; ++qn)
#pragma unroll
                for (int cc = 0; cc < 5; ++cc)
                    in[qn * 5 + cc] = sr[cc][bj_ln0 + (qn & 1) + (qn >> 1) * HALO_W];
            // three independent fused multiply-add chains (the build contracts nothing by
            // itself; a preconditioner needs no particular rounding, only the same one for
            // every block that computes an entry, which one code path gives)
            double uu[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 20; ++m)
#pragma unroll
                for (int o = 0; o < 3; ++o) uu[o] = __builtin_fma(brow[o][m], in[m], uu[o]);
#pragma unroll
            for (int o = 0; o < 3; ++o)
                if (bj_ok[o]) su[bj_c[o]][bj_ln[o]] = uu[o];
        }
#pragma unroll
        for (int t = 0; t < CG1_NE; ++t) {
            k_rn[t] = k_sj[t] = 0.0;
            if (threadIdx.x + t * BLK >= 5 * HALO_N) continue;
            double sj = e_w[t] + beta * e_s[t];
            double rn = e_r[t] - alpha * sj;
            double un = BJ ? 0.0 : e_id[t] * rn;
            double pj = 0.0;
            k_rn[t] = rn;
            k_sj[t] = sj;
            if (e_in[t] && e_own[t]) {
                if (!init) {
                    pj = (BJ ? u_reg[t] : e_id[t] * e_r[t]) + beta * p_reg[t];
                    x_reg[t] += alpha * pj;
                    p_reg[t] = pj;
                }
                if (!BJ) {
                    acc[0] += rn * un;
                    acc[2] += un * sj;
                }
                acc[4] += pj * sj;
                acc[5] += rn * rn;
            }
            if (!BJ) su[e_c[t]][e_ln[t]] = un;
            if (e_own[t]) {
                int lx = e_ln[t] % HALO_W, ly = e_ln[t] / HALO_W;
                sp[e_c[t]][(lx - 2) + (ly - 2) * TILE] = pj;
            }
        }
        __syncthreads();
        if constexpr (BJ) {
#pragma unroll
            for (int t = 0; t < CG1_NE; ++t)
                if (threadIdx.x + t * BLK < 5 * HALO_N && e_in[t] && e_own[t]) {
                    double un = su[e_c[t]][e_ln[t]];
                    acc[0] += k_rn[t] * un;
                    acc[2] += un * k_sj[t];
                    u_reg[t] = un;
                }
        }
        // r_{j+1} and s_j of tile + halo for the next stage
#pragma unroll
        for (int t = 0; t < CG1_NE; ++t) {
            e_r[t] = k_rn[t];
            e_s[t] = k_sj[t];
        }
        double y0 = 0.0, y1 = 0.0;
        if (node_in) {
            auto xin = [&](int c, int du, int dv) { return su[c][l0 + du + dv * HALO_W]; };
            auto slotv = [&](int i) { return Ar[i]; };
            stencil_rows<true>(g, slotv, xin, &y0, &y1);
            if (g < 3) {
                acc[1] += y0 * su[g][l0];
                acc[3] += y0 * sp[g][lane];
            } else {
                acc[1] += y0 * su[3][l0] + y1 * su[4][l0];
                acc[3] += y0 * sp[3][lane] + y1 * sp[4][lane];
            }
        }
        CgSlot* slot_out = a.slot[init ? 0 : 1 - q];
        cgp_partials(slot_out, acc, shB, tile);
        // output: w_{j+1} of the tile
        double* w_out = a.w[init ? 0 : 1 - q];
        if (node_in) {
            if (g < 3) {
                st_dev(w_out + g * a.P + p0, y0);
            } else {
                st_dev(w_out + 3 * a.P + p0, y0);
                st_dev(w_out + 4 * a.P + p0, y1);
            }
        }
        cgp_tag(slot_out, (epoch << 32) | (unsigned long long)(j + 1), tile);
    }
    // the solution of the owned entries
#pragma unroll
    for (int t = 0; t < CG1_NE; ++t)
        if (e_in[t] && e_own[t]) a.x[e_i[t]] = x_reg[t];
}

template <int KP, bool BJ>
__global__ void __launch_bounds__(BLK) k_cgp(CgpArgs a) { cgp_tile<KP, BJ>(a, blockIdx.x, a.epoch); }

// 2x2-node block-Jacobi preconditioner: for the group of nodes (2gu + i, 2gv + j), i, j in
// {0, 1}, the inverse of A restricted to the group's DOFs (up to 20: 3 L and, at interior
// crossings, 2 E per node), by Gauss-Jordan without pivoting (a principal block of the
// step's symmetric positive definite system).  One wavefront inverts three groups: lane
// 20 g + j holds column j of group g in registers; at pivot k lane k's column goes through
// LDS to the group's other lanes.  Row (component c, node) goes to
// brow[(c P + padded node) 20 + q' 5 + c'], zero for DOFs that do not exist.
constexpr int BJF_GROUPS = 3;   // groups per 64-lane block

__device__ __forceinline__ bool bj_exists(int W, int Lg, int gu, int gv, int q, int c) {
    int u = gu + (q & 1), v = gv + (q >> 1);
    if (u >= W || v >= Lg) return false;
    return c < 3 || (u > 0 && u < W - 1 && v > 0 && v < Lg - 1);
}

// Inverts the group at (gu, gv) held by lanes 20 gl + j of a 64-thread block and stores its
// rows; every lane calls it (it passes 40 barriers).  load(qi, ci, s): the entry of node qi's
// row component ci in slot s.
template <class Load>
__device__ __forceinline__ void bj_invert_store(int W, int Lg, int Wp, int P, double* brow, bool live, int gu, int gv,
                                                int gl, int j, double (*colk)[20], Load load) {
    const int qj = j / 5, cj = j - 5 * qj;
    const bool jex = live && bj_exists(W, Lg, gu, gv, qj, cj);
    double col[20];
#pragma unroll
    for (int i = 0; i < 20; ++i) {
        const int qi = i / 5, ci = i - 5 * qi;
        double val = i == j ? 1.0 : 0.0;
        if (jex && bj_exists(W, Lg, gu, gv, qi, ci)) {
            int s = slot_of(ci, (qj & 1) - (qi & 1), (qj >> 1) - (qi >> 1), cj);
            val = s >= 0 ? load(qi, ci, s) : 0.0;
        }
        col[i] = val;
    }
#pragma unroll
    for (int k = 0; k < 20; ++k) {
        if (gl < BJF_GROUPS && j == k)
#pragma unroll
            for (int i = 0; i < 20; ++i) colk[gl][i] = col[i];
        __syncthreads();
        if (gl < BJF_GROUPS) {
            double ck[20];
#pragma unroll
            for (int i = 0; i < 20; ++i) ck[i] = colk[gl][i];
            const double ip = 1.0 / ck[k];
            if (j == k) {
#pragma unroll
                for (int i = 0; i < 20; ++i) col[i] = i == k ? ip : -ck[i] * ip;
            } else {
                const double f = col[k] * ip;
#pragma unroll
                for (int i = 0; i < 20; ++i) col[i] = i == k ? f : col[i] - ck[i] * f;
            }
        }
        __syncthreads();
    }
    if (!live) return;
#pragma unroll
    for (int i = 0; i < 20; ++i) {
        const int qi = i / 5, ci = i - 5 * qi;
        const int u = gu + (qi & 1), v = gv + (qi >> 1);
        if (u >= W || v >= Lg) continue;
        brow[((size_t)ci * P + padidx(Wp, u, v)) * 20 + j] = jex && bj_exists(W, Lg, gu, gv, qi, ci) ? col[i] : 0.0;
    }
}

// Block b of a grid of nb as the dispatcher places it: on XCD b % 8 of the eight, in turn.
// Returns the block index that gives each XCD a contiguous range of [0, nb) (a bijection),
// so that blocks whose data share cache lines run on the XCD whose L2 holds them.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// From A (or A^T) in HBM: three groups per wavefront; `unit` numbers the wavefront's
// three groups, `lane` is the thread's lane, colk the wavefront's own pivot columns.  Every
// wavefront of the block calls it (it passes the block's barriers).
__device__ __forceinline__ void bj_factor_unit(const double* A, double* brow, int W, int Lg, int Wp, int P, int ntx, int ngx,
                                               int ngroups, int unit, int lane, double (*colk)[20]) {
    const int gl = lane / 20, j = lane - 20 * gl;
    const int grp = unit * BJF_GROUPS + gl;
    const bool live = gl < BJF_GROUPS && grp < ngroups;
    const int gu = live ? (grp % ngx) * 2 : 0, gv = live ? (grp / ngx) * 2 : 0;
    bj_invert_store(W, Lg, Wp, P, brow, live, gu, gv, gl, j, colk, [&](int qi, int, int s) {
        return A[a_base(ntx, gu + (qi & 1), gv + (qi >> 1)) + (size_t)s * TILE_N];
    });
}

// One wavefront per block.  The groups of a tile read the same 64-byte lines of its slot
// planes (neighbouring nodes of one slot are adjacent), so consecutive groups go to one XCD.
__global__ void __launch_bounds__(64) k_bj_factor(const double* A, double* brow, int W, int Lg, int Wp, int P, int ntx,
                                                   int ngx, int ngroups) {
    __shared__ double colk[BJF_GROUPS][20];
    bj_factor_unit(A, brow, W, Lg, Wp, P, ntx, ngx, ngroups, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, colk);
}

// The block-Jacobi rows of A and the learned warm start in one launch, so that they run side
// by side without a second stream: blocks [0, nbj) factor twelve groups each (four
// wavefronts of three, consecutive blocks of an XCD on consecutive groups), the others run
// the network on 64 nodes each.  Both write what their own kernels write, bit for bit.
struct BjArgs {
    const double* A; double* brow;
    int W, Lg, Wp, P, ntx, ngx, ngroups, nbj;
};
__global__ void __launch_bounds__(256) k_mlp_warm_bj(FeatArgs a, MlpW w, const float* in_mean, const float* in_istd,
                                                      const float* out_std, double* x0, BjArgs bj) {
    if ((int)blockIdx.x < bj.nbj) {
        __shared__ double colk[4][BJF_GROUPS][20];
        const int wv = threadIdx.x >> 6;
        bj_factor_unit(bj.A, bj.brow, bj.W, bj.Lg, bj.Wp, bj.P, bj.ntx, bj.ngx, bj.ngroups,
                       4 * xcd_block(blockIdx.x, bj.nbj) + wv, threadIdx.x & 63, colk[wv]);
    } else {
        mlp_warm_tile(a, w, in_mean, in_istd, out_std, x0, blockIdx.x - bj.nbj);
    }
}


// ---------------------------------------------------------------------------
// Update (src/Cloth.cpp:111-148) and geometry refresh (src/Cloth.cpp:307-348)
// ---------------------------------------------------------------------------
struct StateDev {
    double* x; double* v; double* xfix; double* uv; double* uvd; double* uvbar; double* nrm;
    const unsigned char* flags;
    int W, Lg, N;
};

__device__ __forceinline__ void update_node(const StateDev& s, const double* sol, const double* is_slide, int Wp, int P,
                                            double h, int k) {
    if (k >= s.N) return;
    int u = k % s.W, v = k / s.W;
    int pk = padidx(Wp, u, v);
    bool fixed = (s.flags[k] & 2) != 0;
    for (int c = 0; c < 3; ++c) {
        double nv = sol[c * P + pk];
        s.v[c * s.N + k] = nv;
        double nx = s.x[c * s.N + k] + nv * h;
        s.x[c * s.N + k] = nx;
        if (!fixed) s.xfix[c * s.N + k] = nx;
    }
    if (u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1) {
        for (int c = 0; c < 2; ++c) {
            double nv = sol[(3 + c) * P + pk];
            s.uvd[c * s.N + k] = nv;
            s.uv[c * s.N + k] = s.uv[c * s.N + k] + nv * h;
            s.uvbar[c * s.N + k] = s.uvbar[c * s.N + k] + is_slide[c * s.N + k] * nv * h;
        }
    }
}

// Component c of node k's update, in update_node's arithmetic.
__device__ __forceinline__ void update_comp(const StateDev& s, const double* sol, const double* is_slide, int Wp, int P,
                                            double h, int k, int c) {
    const int u = k % s.W, v = k / s.W;
    const int pk = padidx(Wp, u, v);
    if (c < 3) {
        const bool fixed = (s.flags[k] & 2) != 0;
        double nv = sol[c * P + pk];
        s.v[c * s.N + k] = nv;
        double nx = s.x[c * s.N + k] + nv * h;
        s.x[c * s.N + k] = nx;
        if (!fixed) s.xfix[c * s.N + k] = nx;
    } else if (u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1) {
        const int e = c - 3;
        double nv = sol[(3 + e) * P + pk];
        s.uvd[e * s.N + k] = nv;
        s.uv[e * s.N + k] = s.uv[e * s.N + k] + nv * h;
        s.uvbar[e * s.N + k] = s.uvbar[e * s.N + k] + is_slide[e * s.N + k] * nv * h;
    }
}

// One thread per (component, node), component-major (5 N threads, so that the loads of a
// wavefront are consecutive entries of one plane).  With dprev set, also the entry's
// velocity change of this solve relative to the velocity it started from (dprev = sol -
// vpad, padded layout); the previous change moves to dprev2.  Padding entries are never
// written: sol and vpad are zero there, so are dprev and dprev2.
__global__ void __launch_bounds__(NBLK) k_update(StateDev s, const double* sol, const double* is_slide, int Wp, int P, double h,
                                                 const double* vpad = nullptr, double* dprev = nullptr, double* dprev2 = nullptr) {
    const int t = blockIdx.x * NBLK + threadIdx.x;
    if (t >= 5 * s.N) return;
    const int c = t / s.N, k = t - c * s.N;
    if (dprev) {
        const size_t i = (size_t)c * P + padidx(Wp, k % s.W, k / s.W);
        dprev2[i] = dprev[i];
        dprev[i] = sol[i] - vpad[i];
    }
    update_comp(s, sol, is_slide, Wp, P, h, k, c);
}

// Cloth::NodeNorm (src/Cloth.cpp:440-463): smallest-eigenvalue eigenvector of the 3x3
// scatter matrix of the 5-point cross (cyclic Jacobi), aligned with
// (x_right - x_left) x (x_bot - x_top), flipped by the weave pattern.
// Eigen-decomposition of the 3x3 scatter matrix of the 5-point cross at interior node
// k: eigenvalues lam, eigenvectors as the columns of V, the index im of the smallest,
// and the sign sg that the reference applies to that eigenvector.
struct NodeFrame {
    double lam[3], V[3][3], d[5][3], sg;
    int im;
};

// M[r][j] for a column j known only at run time, by selects
__device__ __forceinline__ double frame_col(const double (&M)[3][3], int r, int j) {
    return j == 0 ? M[r][0] : (j == 1 ? M[r][1] : M[r][2]);
}

// Inlined, with the smallest eigenpair picked by selects rather than a run-time index, so
// that the frame stays in registers.
__device__ __forceinline__ void node_frame(const StateDev& s, int k, NodeFrame& f) {
    int ids[5] = {k, k - 1, k + 1, k - s.W, k + s.W};
    double p[5][3];
    for (int i = 0; i < 5; ++i) for (int c = 0; c < 3; ++c) p[i][c] = s.x[c * s.N + ids[i]];
    double m[3] = {0, 0, 0};
    for (int i = 0; i < 5; ++i) for (int c = 0; c < 3; ++c) m[c] += p[i][c];
    for (int c = 0; c < 3; ++c) m[c] = m[c] * (1.0 / 5.0);
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < 5; ++i) {
        for (int c = 0; c < 3; ++c) f.d[i][c] = p[i][c] - m[c];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) A[a][b] += f.d[i][a] * f.d[i][b];
    }
    double (&V)[3][3] = f.V;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double tr = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off == 0.0 || off <= 1e-34 * tr) break;
        for (int pp = 0; pp < 2; ++pp)
            for (int q = pp + 1; q < 3; ++q) {
                if (A[pp][q] == 0.0) continue;
                double th = (A[q][q] - A[pp][pp]) / (2.0 * A[pp][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < 3; ++r) { double a0 = A[r][pp], a1 = A[r][q]; A[r][pp] = cs * a0 - sn * a1; A[r][q] = sn * a0 + cs * a1; }
                for (int r = 0; r < 3; ++r) { double a0 = A[pp][r], a1 = A[q][r]; A[pp][r] = cs * a0 - sn * a1; A[q][r] = sn * a0 + cs * a1; }
                for (int r = 0; r < 3; ++r) { double a0 = V[r][pp], a1 = V[r][q]; V[r][pp] = cs * a0 - sn * a1; V[r][q] = sn * a0 + cs * a1; }
            }
    }
    for (int i = 0; i < 3; ++i) f.lam[i] = A[i][i];
    int im = 0;
    double lmin = A[0][0];
    if (A[1][1] < lmin) { im = 1; lmin = A[1][1]; }
    if (A[2][2] < lmin) im = 2;
    f.im = im;
    double n[3];
    for (int c = 0; c < 3; ++c) n[c] = frame_col(V, c, im);
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; ++c) n[c] = n[c] / nn;
    double a[3], b[3];
    for (int c = 0; c < 3; ++c) { a[c] = p[2][c] - p[1][c]; b[c] = p[4][c] - p[3][c]; }
    double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    double nn2 = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double cosv = (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2]) / (cn * nn2);
    double sg = cosv < 0 ? -1.0 : 1.0;
    if (s.flags[k] & 1) sg = -sg;
    f.sg = sg;
}

__device__ __forceinline__ void normals_node(const StateDev& s, int k) {
    if (k >= s.N) return;
    int u = k % s.W, v = k / s.W;
    if (!(u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1)) {
        for (int c = 0; c < 3; ++c) s.nrm[c * s.N + k] = 0.0;
        return;
    }
    NodeFrame f;
    node_frame(s, k, f);
    // the eigenvector normalised, then signed
    double n[3] = {frame_col(f.V, 0, f.im), frame_col(f.V, 1, f.im), frame_col(f.V, 2, f.im)};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; ++c) s.nrm[c * s.N + k] = f.sg * (n[c] / nn);
}

__global__ void __launch_bounds__(NBLK) k_normals(StateDev s) { normals_node(s, blockIdx.x * NBLK + threadIdx.x); }

// Tangent of the crossing normals along the position tangent dx: first-order
// perturbation of the smallest eigenvector of the scatter matrix,
// dn = sum_{j != min} e_j (e_j^T dC e_min) / (lam_min - lam_j), signed as n (the sign
// is locally constant).  This is the derivative the reference's autograd takes through
// torch::svd (src/Cloth.cpp:446) when the singular values are distinct.
__global__ void __launch_bounds__(NBLK) k_normals_tan(StateDev s, const double* dx, double* dnrm, size_t dstride) {
    // direction blockIdx.y: its tangents dstride doubles after the previous direction's
    dx += (size_t)blockIdx.y * dstride;
    dnrm += (size_t)blockIdx.y * dstride;
    int k = blockIdx.x * NBLK + threadIdx.x;
    if (k >= s.N) return;
    int u = k % s.W, v = k / s.W;
    if (!(u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1)) {
        for (int c = 0; c < 3; ++c) dnrm[c * s.N + k] = 0.0;
        return;
    }
    NodeFrame f;
    node_frame(s, k, f);
    int ids[5] = {k, k - 1, k + 1, k - s.W, k + s.W};
    double dp[5][3], dm[3] = {0, 0, 0};
    for (int i = 0; i < 5; ++i) for (int c = 0; c < 3; ++c) { dp[i][c] = dx[c * s.N + ids[i]]; dm[c] += dp[i][c]; }
    for (int c = 0; c < 3; ++c) dm[c] = dm[c] * (1.0 / 5.0);
    double dC[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < 5; ++i) {
        double dd_[3] = {dp[i][0] - dm[0], dp[i][1] - dm[1], dp[i][2] - dm[2]};
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) dC[a][b] += dd_[a] * f.d[i][b] + f.d[i][a] * dd_[b];
    }
    // unit eigenvectors
    double E[3][3];
    for (int j = 0; j < 3; ++j) {
        double nn = sqrt(f.V[0][j] * f.V[0][j] + f.V[1][j] * f.V[1][j] + f.V[2][j] * f.V[2][j]);
        for (int r = 0; r < 3; ++r) E[r][j] = f.V[r][j] / nn;
    }
    int im = f.im;
    const double e0 = frame_col(E, 0, im), e1 = frame_col(E, 1, im), e2 = frame_col(E, 2, im);
    const double lam_min = im == 0 ? f.lam[0] : (im == 1 ? f.lam[1] : f.lam[2]);
    double Ce[3];
    for (int a = 0; a < 3; ++a) Ce[a] = dC[a][0] * e0 + dC[a][1] * e1 + dC[a][2] * e2;
    double dn[3] = {0, 0, 0};
    for (int j = 0; j < 3; ++j) {
        if (j == im) continue;
        double cj = (E[0][j] * Ce[0] + E[1][j] * Ce[1] + E[2][j] * Ce[2]) / (lam_min - f.lam[j]);
        for (int r = 0; r < 3; ++r) dn[r] += cj * E[r][j];
    }
    for (int c = 0; c < 3; ++c) dnrm[c * s.N + k] = f.sg * dn[c];
}

// Tangent of Cloth::update (src/Cloth.cpp:111-148): with the solution x_sol and its
// tangent dsol, v' = x_sol[L], x' = x + h v', x_fix' = x' off the handles,
// u' = x_sol[E], uv' = uv + h u', uvbar' = uvbar + h is_slide u'.
struct TanState {
    double *dx, *dv, *dxfix, *duv, *duvd, *duvbar;
};

__global__ void __launch_bounds__(NBLK) k_update_tan(StateDev s, TanState t, const double* sol, const double* dsol,
                                                    const double* is_slide, const double* dis_slide, int Wp, int P, double h) {
    int k = blockIdx.x * NBLK + threadIdx.x;
    if (k >= s.N) return;
    int u = k % s.W, v = k / s.W;
    int pk = padidx(Wp, u, v);
    bool fixed = (s.flags[k] & 2) != 0;
    for (int c = 0; c < 3; ++c) {
        double dnv = dsol[c * P + pk];
        t.dv[c * s.N + k] = dnv;
        double dnx = t.dx[c * s.N + k] + dnv * h;
        t.dx[c * s.N + k] = dnx;
        if (!fixed) t.dxfix[c * s.N + k] = dnx;
    }
    if (u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1) {
        for (int c = 0; c < 2; ++c) {
            double nv = sol[(3 + c) * P + pk], dnv = dsol[(3 + c) * P + pk];
            t.duvd[c * s.N + k] = dnv;
            t.duv[c * s.N + k] = t.duv[c * s.N + k] + dnv * h;
            t.duvbar[c * s.N + k] = t.duvbar[c * s.N + k] + (dis_slide[c * s.N + k] * nv + is_slide[c * s.N + k] * dnv) * h;
        }
    }
}

// At = A^T in the same stencil layout (the stencil is symmetric in structure): the
// entry of row (n, rr), column (m, cc) of At is A's entry of row (m, cc), column (n, rr).
__device__ __forceinline__ void transpose_tile(const double* A, double* At, int W, int Lg, int ntx, int tile) {
    int u = (tile % ntx) * TILE + (threadIdx.x & (TILE - 1));
    int v = (tile / ntx) * TILE + (threadIdx.x / TILE);
    if (u >= W || v >= Lg) return;
    double* out = At + a_base(ntx, u, v);
    for (int sidx = 0; sidx < NSLOT; ++sidx) {
        int rr, du, dv, cc;
        slot_decode(sidx, rr, du, dv, cc);
        int mu = u + du, mv = v + dv;
        double x = 0.0;
        if (mu >= 0 && mu < W && mv >= 0 && mv < Lg) {
            int s2 = slot_of(cc, -du, -dv, rr);
            if (s2 >= 0) x = A[a_base(ntx, mu, mv) + (size_t)s2 * TILE_N];
        }
        out[(size_t)sidx * TILE_N] = x;
    }
}

__global__ void __launch_bounds__(NBLK) k_transpose(const double* A, double* At, int W, int Lg, int ntx) {
    transpose_tile(A, At, W, Lg, ntx, blockIdx.x);
}

}  // namespace

#include "adjoint.hpp"

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct yc_ctx {
    yc_desc d;
    int device;
    int W, Lg, N, Nc, n, Wp, Hp, P;
    hipStream_t stream;
    // device buffers
    double *x, *v, *xfix, *uv, *uvd, *uvbar, *nrm;
    unsigned char* flags;
    double *A, *F, *is_slide, *comp, *fepart;
    double *b, *invd, *cx, *cr, *cp, *cy;  // padded 5-plane vectors
    double* ext;
    double *vpad, *dprev, *dprev2, *Ae;   // padded v_old, the last two steps' dv, A x_q (learned warm start)
    double* partials;
    CGPart part;
    CGScal* sc;
    int* flag;
    CGScal* h_sc;  // pinned
    int* h_flag;   // pinned
    ClothDev cd;
    bool has_ext;
    bool assembled, solved;
    hipGraphExec_t graph;
    int graph_chunk;
    int last_iters;
    int ks_cg_blocks;
    // matrix tiles and the single-reduction CG
    int ntx, nty, nT;
    size_t a_size;
    int cg_variant;                         // YC_CG_REFERENCE or YC_CG_FUSED
    double *cr1, *cw1, *cp1, *cs0, *cs1;    // second parity buffers (r, w, p) and s
    double* part1;                          // 2 parities x 6 x (cg1_kp * BLK) partials
    int cg1_kp;                             // partial rounds of the fused CG kernel
    CgSlot* cgslot;                         // 2 parities x nT slots of the persistent CG
    int* cg_err;
    unsigned long long epoch;               // solves run by the persistent CG
    unsigned long long* bepoch;             // device counter of the batched solves' epochs
    int coop_ok;                            // the persistent CG fits on the device at once
    int coop_bj_ok;                         // ... with the block-Jacobi preconditioner
    int precond;                            // YC_PRECOND_*
    double* bj;                             // block-Jacobi rows, 5 P x 20 (when selected)
    const double* bj_of;                    // the matrix (A or At) bj was computed from, or null
    hipEvent_t io_event;                    // orders device state transfers with a caller's stream
    hipGraphExec_t graph1;
    // forward-mode tangents of the step (directional derivatives)
    int ndir;
    double* tstate;                         // ndir x TAN_PLANES planes of N
    std::vector<double> tpar;               // ndir x YC_NPARAM
    double *dA, *dF, *db, *dis_slide, *dcomp, *At, *xsol, *bsave, *dsol, *trhs;
    double* tdext;                          // ndir x n tangents of the external force, or null
    ClothDev* tdev;                         // ndir tangent descriptors (device), for k_assemble1_dirs
    double* tshear;                         // ndir x 90 x 4Nc shear-tangent records
    AsmOut* tout;                           // ndir tangent outputs (device)
    std::vector<ClothDev> h_tdev;           // their host copies (kept alive for the upload)
    std::vector<AsmOut> h_tout;
    double* shear_rec;                      // 3 roles x 30 x (4 Nc) shear records
    double* bend_rec;                       // 3 roles x BEND_REC x n_bends bend records
    double last_h;                          // step of the last assembly
    double* Aexp;                           // M, K_x or K_v alone, for yc_export_matrix (on first use)
    double* dexp;                           // the diagonal scratch of that assembly (5 P)
    double* fric_out;                       // when set: the assembly writes the friction inputs here
    // reverse mode (yc_adjoint_*): checkpoints of the recorded rollout and the backward
    // pass's buffers
    int adj_max, adj_K, adj_steps;          // capacity, checkpoint interval, steps recorded
    double* adj_ckpt;                       // blocks of G_PLANES N: s_0, s_K, s_2K, ...
    double* adj_final;                      // the state after the last recorded step
    double* adj_seg;                        // K blocks: a segment's states, recomputed
    double* adj_ext;                        // per-step external force (checkpoint interval > 1), or null
    std::vector<char> adj_has_ext;          // whether step t had one
    double *adj_g0, *adj_g1;                // state adjoints (G_PLANES N each)
    double *adj_gpar, *adj_gcomp, *adj_ybar, *adj_pbar, *adj_gslide, *adj_fric_in, *adj_lam, *adj_xs;
    double *adj_gst, *adj_lpart, *adj_sums;
    double *adj_tx, *adj_tu;                // target positions / Eulerian coordinates (steps x N x 3 / 2), or null
    double adj_tw;                          // their loss weight
};

namespace {
void adjoint_free(yc_ctx* c);
}

namespace {

ClothDev make_dev(const yc_ctx* c) {
    ClothDev C = cloth_params(c->d);
    C.x = c->x; C.v = c->v; C.xfix = c->xfix; C.uv = c->uv; C.uvd = c->uvd; C.uvbar = c->uvbar;
    C.dx = C.dv = C.dxfix = C.duv = C.duvd = C.duvbar = C.dnrm = nullptr;
    C.nrm = c->nrm; C.flags = c->flags;
    return C;
}

AsmOut make_out(yc_ctx* c) {
    AsmOut O;
    O.A = c->A; O.F = c->F; O.b = c->b; O.is_slide = c->is_slide; O.comp = c->comp; O.fepart = c->fepart;
    O.dA = O.dF = O.db = O.dis_slide = O.dcomp = nullptr;
    O.fric_in = c->fric_out;
    O.vpad = c->vpad;
    O.dext = nullptr;
    O.detach = 0;
    O.shear_rec = c->shear_rec; O.nseg = 4 * c->Nc;
    O.bend_rec = c->bend_rec;
    O.only_kind = -1;
    O.invdiag = c->invd; O.ext = c->has_ext ? c->ext : nullptr; O.flag = c->flag; O.Wp = c->Wp; O.P = c->P;
    return O;
}

StateDev make_state(yc_ctx* c) {
    StateDev s;
    s.x = c->x; s.v = c->v; s.xfix = c->xfix; s.uv = c->uv; s.uvd = c->uvd; s.uvbar = c->uvbar; s.nrm = c->nrm;
    s.flags = c->flags; s.W = c->W; s.Lg = c->Lg; s.N = c->N;
    return s;
}

VecArgs make_vec(yc_ctx* c) {
    VecArgs a;
    a.x = c->cx; a.r = c->cr; a.pb[0] = c->cp; a.pb[1] = c->cp; a.y = c->cy; a.b = c->b; a.invd = c->invd;
    a.W = c->W; a.Lg = c->Lg; a.N = c->N; a.Wp = c->Wp; a.P = c->P;
    return a;
}

SpmvArgs make_spmv(yc_ctx* c, const double* in, double* out, const double* Aptr = nullptr) {
    SpmvArgs a;
    a.A = Aptr ? Aptr : c->A; a.x = in; a.y = out; a.W = c->W; a.Lg = c->Lg; a.N = c->N; a.Wp = c->Wp; a.P = c->P; a.ntx = c->ntx;
    return a;
}

int grid_for(long long threads) { return (int)((threads + BLK - 1) / BLK); }

// flip_norm by weave pattern (src/Cloth.cpp:66-92)

// State between a caller's node-major device arrays ((N, 3) and (N, 2), state index
// u + v * width) and the context's component planes, one thread per node.
struct StateIO {
    const double* in[6];   // lpos, lposfix, lvel, epos, eposbar, evel (set) or null
    double* out[6];        // the same (get) or null
};

__global__ void __launch_bounds__(NBLK) k_state_in(StateDev s, StateIO io) {
    int k = blockIdx.x * NBLK + threadIdx.x;
    if (k >= s.N) return;
    for (int q = 0; q < 3; ++q) {
        s.x[q * s.N + k] = io.in[0][3 * k + q];
        s.xfix[q * s.N + k] = io.in[1][3 * k + q];
        s.v[q * s.N + k] = io.in[2][3 * k + q];
    }
    for (int q = 0; q < 2; ++q) {
        s.uv[q * s.N + k] = io.in[3][2 * k + q];
        s.uvbar[q * s.N + k] = io.in[4][2 * k + q];
        s.uvd[q * s.N + k] = io.in[5][2 * k + q];
    }
}

__global__ void __launch_bounds__(NBLK) k_state_out(StateDev s, StateIO io) {
    int k = blockIdx.x * NBLK + threadIdx.x;
    if (k >= s.N) return;
    const double* src3[3] = {s.x, s.xfix, s.v};
    const double* src2[3] = {s.uv, s.uvbar, s.uvd};
    for (int a = 0; a < 3; ++a)
        if (io.out[a])
            for (int q = 0; q < 3; ++q) io.out[a][3 * k + q] = src3[a][q * s.N + k];
    for (int a = 0; a < 3; ++a)
        if (io.out[3 + a])
            for (int q = 0; q < 2; ++q) io.out[3 + a][2 * k + q] = src2[a][q * s.N + k];
}

int refresh_geometry(yc_ctx* c) {
    hipLaunchKernelGGL(k_normals, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, make_state(c));
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// CG iteration with parity q: y = A p and p.y partials; x, r update with |r|^2 and r.z
// partials (r.z into slot q); stopping test and p update
int cg_launch_iteration(yc_ctx* c, int q) {
    hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cp, c->cy), c->part.pAp,
                       (const CGScal*)c->sc, 1);
    hipLaunchKernelGGL(k_cg_xr, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, (const CGScal*)c->sc, q);
    hipLaunchKernelGGL(k_cg_p, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, c->sc, q);
    return YC_OK;
}

int build_graph(yc_ctx* c) {
    if (c->graph) return YC_OK;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < c->graph_chunk; ++i) cg_launch_iteration(c, i & 1);   // chunk is even
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    HIPCHK(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return YC_OK;
}

Cg1Args make_cg1(yc_ctx* c) {
    Cg1Args a;
    a.A = c->A; a.invd = c->invd; a.x = c->cx;
    a.r[0] = c->cr; a.r[1] = c->cr1; a.w[0] = c->cy; a.w[1] = c->cw1;
    a.p[0] = c->cp; a.p[1] = c->cp1; a.s[0] = c->cs0; a.s[1] = c->cs1;
    a.part[0] = c->part1; a.part[1] = c->part1 + 6 * (size_t)c->cg1_kp * BLK;
    a.bbpart = c->part.bb; a.rrpart = c->part.rr; a.nbb = c->part.nV;
    a.sc = c->sc;
    a.W = c->W; a.Lg = c->Lg; a.N = c->N; a.Wp = c->Wp; a.P = c->P; a.ntx = c->ntx; a.nT = c->nT;
    return a;
}

// The persistent CG needs every tile's block resident at once: the grid must fit the
// device's CUs at the kernel's occupancy.
template <int KP, bool BJ>
bool cgp_fits_kp(const yc_ctx* c) {
    int dev = c->device, ncu = 0, per_cu = 0;
    if (dev < 0 || dev >= MAX_DEVICES) return false;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cgp<KP, BJ>, BLK, 0) != hipSuccess) return false;
    return (long long)per_cu * ncu >= c->nT;
}

template <bool BJ>
bool cgp_fits(const yc_ctx* c) {
    switch (c->cg1_kp) {
        case 1: return cgp_fits_kp<1, BJ>(c);
        case 2: return cgp_fits_kp<2, BJ>(c);
        case 3: return cgp_fits_kp<3, BJ>(c);
        default: return cgp_fits_kp<4, BJ>(c);
    }
}

template <bool BJ>
const void* cgp_kernel(int kp) {
    switch (kp) {
        case 1: return (const void*)k_cgp<1, BJ>;
        case 2: return (const void*)k_cgp<2, BJ>;
        case 3: return (const void*)k_cgp<3, BJ>;
        default: return (const void*)k_cgp<4, BJ>;
    }
}

// The block-Jacobi rows of the matrix Aptr (k_bj_factor), unless they are already those
// of Aptr: c->bj_of names the matrix the rows were computed from (the assembly computes
// them for the new A, or the warm guess beside its network; everything else that rewrites
// A or At clears it).
int bj_factor(yc_ctx* c, const double* Aptr) {
    if (c->bj_of == Aptr) return YC_OK;
    const int ngx = (c->W + 1) / 2, ngy = (c->Lg + 1) / 2, ng = ngx * ngy;
    hipLaunchKernelGGL(k_bj_factor, dim3((ng + BJF_GROUPS - 1) / BJF_GROUPS), dim3(64), 0, c->stream, Aptr, c->bj, c->W,
                       c->Lg, c->Wp, c->P, c->ntx, ngx, ng);
    HIPCHK(hipGetLastError());
    c->bj_of = Aptr;
    return YC_OK;
}

// Persistent solves (k_cgp here, k_b_cgp in batches) launch as ordinary kernels.  Their
// blocks wait on one another, so all of a solve's blocks must be resident at once: the grid
// fits the device at the kernel's occupancy (cgp_fits), and nothing else of this library
// that waits across blocks may hold CUs meanwhile.  (A cooperative launch would also guarantee residency, but the runtime
// runs it on a separate hardware queue behind a cross-queue wait on each side, about 30 us
// per step at 128 x 128 crossings.)  Two persistent solves in flight together could each
// hold part of the device and wait for the other until the spin limit, so the solves of all
// contexts and batches of the process run one at a time per device: a solve queued on a
// stream other than the previous solve's first waits for everything queued on that stream
// so far.
// Kernels of other libraries or processes do not wait on these blocks and finish by
// themselves; they can only delay a solve.
struct PersistOrder {
    std::mutex mu;
    hipStream_t last[MAX_DEVICES] = {};
    hipEvent_t ev[MAX_DEVICES] = {};
};
PersistOrder& persist_order() {
    static PersistOrder p;
    return p;
}

// Queues `launch` (a persistent solve, or a graph that holds one) on stream st of device
// dev after the device's previous persistent solve.  The launch runs under the lock, so
// the order of the records is the order of the launches.  Not for a capturing stream.
template <class F>
int persist_launch(int dev, hipStream_t st, F&& launch) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(YC_ERR_ARG, "device index beyond the persistent-solve table");
    PersistOrder& po = persist_order();
    std::lock_guard<std::mutex> lock(po.mu);
    if (po.last[dev] && po.last[dev] != st) {
        if (!po.ev[dev]) HIPCHK(hipEventCreateWithFlags(&po.ev[dev], hipEventDisableTiming));
        HIPCHK(hipEventRecord(po.ev[dev], po.last[dev]));
        HIPCHK(hipStreamWaitEvent(st, po.ev[dev], 0));
    }
    po.last[dev] = st;
    return launch();
}

// A stream about to be destroyed stops being its device's last persistent-solve stream.
void persist_forget(int dev, hipStream_t st) {
    if (dev < 0 || dev >= MAX_DEVICES || !st) return;
    PersistOrder& po = persist_order();
    std::lock_guard<std::mutex> lock(po.mu);
    if (po.last[dev] == st) po.last[dev] = nullptr;
}

// With the block-Jacobi preconditioner the rows of the matrix solved are made current first.
int launch_cgp(yc_ctx* c, double tol, int maxit, const double* Aptr = nullptr) {
    CgpArgs a;
    a.A = Aptr ? Aptr : c->A; a.invd = c->invd; a.bbpart = c->part.bb; a.rrpart = c->part.rr;
    const bool bj = c->precond == YC_PRECOND_BLOCK2;
    a.brow = bj ? c->bj : nullptr;
    if (bj) {
        int rc = bj_factor(c, a.A);
        if (rc) return rc;
    }
    a.x = c->cx;
    a.r0 = c->cr; a.w[0] = c->cy; a.w[1] = c->cw1;
    a.slot[0] = c->cgslot; a.slot[1] = c->cgslot + c->nT;
    a.sc = c->sc; a.err = c->cg_err;
    a.epoch = ++c->epoch;
    a.tol = tol; a.maxit = maxit; a.nbb = c->part.nV;
    a.spin_limit = 1 << 18;   // polls of about 1 us each: a quarter second
    a.W = c->W; a.Lg = c->Lg; a.Wp = c->Wp; a.P = c->P; a.ntx = c->ntx; a.nT = c->nT;
    void* args[] = {&a};
    const void* fn = bj ? cgp_kernel<true>(c->cg1_kp) : cgp_kernel<false>(c->cg1_kp);
    return persist_launch(c->device, c->stream, [&]() -> int {
        HIPCHK(hipLaunchKernel(fn, dim3(c->nT), dim3(BLK), args, 0, c->stream));
        return YC_OK;
    });
}

template <int Q>
void launch_cg1_q(yc_ctx* c, const Cg1Args& a, int init, double tol, int maxit) {
    dim3 grid(c->nT), blk(BLK);
    switch (c->cg1_kp) {
        case 1: hipLaunchKernelGGL((k_cg1<1, Q>), grid, blk, 0, c->stream, a, init, tol, maxit); break;
        case 2: hipLaunchKernelGGL((k_cg1<2, Q>), grid, blk, 0, c->stream, a, init, tol, maxit); break;
        case 3: hipLaunchKernelGGL((k_cg1<3, Q>), grid, blk, 0, c->stream, a, init, tol, maxit); break;
        default: hipLaunchKernelGGL((k_cg1<4, Q>), grid, blk, 0, c->stream, a, init, tol, maxit); break;
    }
}

void launch_cg1(yc_ctx* c, const Cg1Args& a, int q, int init, double tol, int maxit) {
    if (q == 0) launch_cg1_q<0>(c, a, init, tol, maxit);
    else launch_cg1_q<1>(c, a, init, tol, maxit);
}

int build_graph1(yc_ctx* c) {
    if (c->graph1) return YC_OK;
    hipGraph_t g;
    Cg1Args a = make_cg1(c);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < c->graph_chunk; ++i)   // chunk is even: parity continues across chunks
        launch_cg1(c, a, i & 1, 0, 0.0, 0);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    HIPCHK(hipGraphInstantiate(&c->graph1, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return YC_OK;
}

void to_padded(const yc_ctx* c, const double* glob, std::vector<double>& pad) {
    pad.assign((size_t)5 * c->P, 0.0);
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            int k = u + v * c->W;
            int pk = (u + 2) + (v + 2) * c->Wp;
            for (int a = 0; a < 3; ++a) pad[(size_t)a * c->P + pk] = glob[3 * k + a];
            if (u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1) {
                int e = (u - 1) + (v - 1) * (c->W - 2);
                for (int a = 0; a < 2; ++a) pad[(size_t)(3 + a) * c->P + pk] = glob[3 * c->N + 2 * e + a];
            }
        }
}

void from_padded(const yc_ctx* c, const std::vector<double>& pad, double* glob) {
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            int k = u + v * c->W;
            int pk = (u + 2) + (v + 2) * c->Wp;
            for (int a = 0; a < 3; ++a) glob[3 * k + a] = pad[(size_t)a * c->P + pk];
            if (u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1) {
                int e = (u - 1) + (v - 1) * (c->W - 2);
                for (int a = 0; a < 2; ++a) glob[3 * c->N + 2 * e + a] = pad[(size_t)(3 + a) * c->P + pk];
            }
        }
}

// async: no host synchronisation; the flags are sticky and read by the caller later.
// With the block-Jacobi preconditioner the rows of the new A follow the assembly, unless
// defer_bj: then their first user computes them (the warm guess, beside its network).
int do_assemble(yc_ctx* c, double h, bool async = false, bool defer_bj = false) {
    if (!async) HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
    c->cd = make_dev(c);
    c->last_h = h;
    static_assert(NBLK == TILE_N, "assembly blocks are matrix tiles");
    if constexpr (USE_BEND_REC) if (c->cd.n_bends > 0)
        hipLaunchKernelGGL(k_bend_elem, dim3((c->cd.n_bends + 255) / 256), dim3(256), 0, c->stream, c->cd, c->bend_rec,
                           c->cd.n_bends);
    hipLaunchKernelGGL(k_assemble1<double>, dim3(c->nT), dim3(ASM1_THREADS), ASM1_LDS, c->stream, c->cd, make_out(c), h);
    HIPCHK(hipGetLastError());
    if (c->Nc > 0)
        hipLaunchKernelGGL(k_shear_elem, dim3((4 * c->Nc + 255) / 256), dim3(256), 0, c->stream, c->cd, (const double*)c->comp,
                           c->shear_rec, 4 * c->Nc);
    hipLaunchKernelGGL(k_assemble2, dim3(c->Lg * ((c->W + TILE_N - 1) / TILE_N)), dim3(ASM2_THREADS), 0, c->stream, c->cd, make_out(c), h);
    HIPCHK(hipGetLastError());
    c->bj_of = nullptr;   // A is new
    if (c->precond == YC_PRECOND_BLOCK2 && !defer_bj) {
        int rc = bj_factor(c, c->A);
        if (rc) return rc;
    }
    if (async) {
        HIPCHK(hipGetLastError());
        c->assembled = true;
        return YC_OK;
    }
    HIPCHK(hipMemcpyAsync(c->h_flag, c->flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->assembled = true;
    if (c->h_flag[0]) return fail(YC_ERR_HIP, "assembly: stencil slot out of range");
    if (c->h_flag[1]) return fail(YC_ERR_NONFINITE, "NaN/Inf in A or b");
    return YC_OK;
}

// Results of a solve from the scalars copied to h_sc.
int finish_solve(yc_ctx* c, bool fused, int* iters, double* relres) {
    if (c->h_sc->zero_rhs) HIPCHK(hipMemsetAsync(c->cx, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    c->solved = true;
    double rel = c->h_sc->bb > 0 ? std::sqrt(c->h_sc->rr / c->h_sc->bb) : 0.0;
    c->last_iters = fused ? c->h_sc->fin_iters : std::max(c->h_sc->itc[0], c->h_sc->itc[1]);
    if (iters) *iters = c->last_iters;
    if (relres) *relres = rel;
    if (!std::isfinite(rel)) return fail(YC_ERR_NONFINITE, "NaN/Inf in the solution");
    if (c->h_sc->rr >= c->h_sc->thr && !c->h_sc->zero_rhs) return fail(YC_ERR_NOT_CONVERGED, "CG stopped at maxit above tolerance");
    return YC_OK;
}

int do_solve(yc_ctx* c, const double* x0_global, double tol, int maxit, int* iters, double* relres, bool x0_in_cx = false,
             bool async = false) {
    if (!c->assembled) return fail(YC_ERR_ARG, "yc_solve before yc_assemble");
    int has_x0 = 0;
    if (x0_in_cx) {
        hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cx, c->cy), (double*)nullptr,
                           (const CGScal*)c->sc, 0);
        has_x0 = 1;
    } else if (x0_global) {
        std::vector<double> pad;
        to_padded(c, x0_global, pad);
        HIPCHK(hipMemcpyAsync(c->cx, pad.data(), pad.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cx, c->cy), (double*)nullptr,
                           (const CGScal*)c->sc, 0);
        has_x0 = 1;
    }
    const bool persistent = c->cg_variant == YC_CG_PERSISTENT;
    const bool fused = c->cg_variant == YC_CG_FUSED || persistent;
    if (c->precond != YC_PRECOND_JACOBI && !persistent)
        return fail(YC_ERR_ARG, "the block-Jacobi preconditioner needs the persistent CG variant");
    hipLaunchKernelGGL(k_cg_init, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, has_x0);
    if (persistent) {
        int rc = launch_cgp(c, tol, maxit);
        if (rc) return rc;
        if (async) {   // results accumulate in sc; the caller synchronises once
            c->solved = true;
            return YC_OK;
        }
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->h_flag[2]) {
            (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
            return fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
        }
        return finish_solve(c, true, iters, relres);
    }
    if (fused)
        launch_cg1(c, make_cg1(c), 0, 1, tol, maxit);
    else
        hipLaunchKernelGGL(k_cg_init2, dim3(1), dim3(BLK), 0, c->stream, c->part, c->sc, tol, maxit);
    HIPCHK(hipGetLastError());
    int rc = fused ? build_graph1(c) : build_graph(c);
    if (rc) return rc;
    hipGraphExec_t graph = fused ? c->graph1 : c->graph;
    HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    long long launched = 0;
    // launch ahead as many chunks as the previous solve needed, then check once per chunk
    int ahead = std::max(1, (int)(0.9 * c->last_iters) / c->graph_chunk);
    while (!c->h_sc->done && launched < (long long)maxit + c->graph_chunk) {
        int n = launched == 0 ? ahead : 1;
        for (int i = 0; i < n; ++i) HIPCHK(hipGraphLaunch(graph, c->stream));
        launched += (long long)n * c->graph_chunk;
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return finish_solve(c, fused, iters, relres);
}

int do_update(yc_ctx* c, double h) {
    if (!c->solved) return fail(YC_ERR_ARG, "yc_update before yc_solve");
    hipLaunchKernelGGL(k_update, dim3((5 * c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, make_state(c), (const double*)c->cx,
                       (const double*)c->is_slide, c->Wp, c->P, h, (const double*)c->vpad, c->dprev, c->dprev2);
    HIPCHK(hipGetLastError());
    int rc = refresh_geometry(c);
    c->assembled = false;
    c->solved = false;
    return rc;
}

template <class T>
int dalloc(T** p, size_t count) {
    HIPCHK(hipMalloc((void**)p, count * sizeof(T) + 64));
    // hipMemset of device memory returns before the fill is done, on the null stream,
    // which does not order against the contexts' non-blocking streams: wait for it, or a
    // following upload on a context stream can land first and be zeroed
    HIPCHK(hipMemset(*p, 0, count * sizeof(T) + 64));
    HIPCHK(hipStreamSynchronize(nullptr));
    return YC_OK;
}

// ---------------------------------------------------------------------------
// Forward-mode tangents of the step.  Direction d holds the tangents of the state
// (TAN_PLANES planes of N: dx, dv, dxfix (3 each), duv, duvd, duvbar (2 each), dnrm (3))
// and of the YC_NPARAM material parameters.
// ---------------------------------------------------------------------------
constexpr int TAN_PLANES = 18;

TanState tan_state(yc_ctx* c, int d) {
    double* b = c->tstate + (size_t)d * TAN_PLANES * c->N;
    size_t N = c->N;
    TanState t;
    t.dx = b; t.dv = b + 3 * N; t.dxfix = b + 6 * N; t.duv = b + 9 * N; t.duvd = b + 11 * N; t.duvbar = b + 13 * N;
    return t;
}
double* tan_nrm(yc_ctx* c, int d) { return c->tstate + (size_t)d * TAN_PLANES * c->N + 15 * (size_t)c->N; }

ClothDev make_tan_dev(yc_ctx* c, int d, int ref) {
    ClothDev C = make_dev(c);
    TanState t = tan_state(c, d);
    C.dx = t.dx; C.dv = t.dv; C.dxfix = t.dxfix; C.duv = t.duv; C.duvd = t.duvd; C.duvbar = t.duvbar;
    C.dnrm = tan_nrm(c, d);
    const double* p = c->tpar.data() + (size_t)d * YC_NPARAM;
    const double area = PI_REF * std::pow(c->d.R, 2);
    for (int y = 0; y < 2; ++y) {
        int yy = c->d.n_yarns == 1 ? 0 : y;
        C.drho[y] = p[YC_P_RHO0 + yy];
        C.dks[y] = p[YC_P_Y0 + yy] * area;    // ks = Y pi R^2 (src/Yarn.h:14)
        C.dkb[y] = p[YC_P_B0 + yy] * area;    // kb = B pi R^2 (src/Yarn.h:15)
    }
    C.dS = p[YC_P_S]; C.dmu = p[YC_P_MU]; C.dkc = p[YC_P_KC]; C.dkf = p[YC_P_KF]; C.ddf = p[YC_P_DF];
    C.dhandle_k = p[YC_P_HANDLE_K];
    C.ref_mode = ref;
    return C;
}

// Solve A_use x = c->b into c->cx without hipGraphs (the graphs hold c->A), through the
// persistent CG when it fits, else the reference-order kernels.
// x0: start from c->cx, whose product A x0 is in c->cy (else from zero).
int solve_on(yc_ctx* c, const double* Aptr, double tol, int maxit, int* iters, bool x0 = false) {
    if (c->precond != YC_PRECOND_JACOBI && !c->coop_ok)
        return fail(YC_ERR_ARG, "the block-Jacobi preconditioner needs the persistent CG");
    hipLaunchKernelGGL(k_cg_init, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, x0 ? 1 : 0);
    if (c->coop_ok) {
        int rc = launch_cgp(c, tol, maxit, Aptr);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->h_flag[2]) {
            (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
            return fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
        }
        if (iters) *iters = c->h_sc->fin_iters;
    } else {
        hipLaunchKernelGGL(k_cg_init2, dim3(1), dim3(BLK), 0, c->stream, c->part, c->sc, tol, maxit);
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        long long it = 0;
        while (!c->h_sc->done && it < (long long)maxit + 64) {
            for (int i = 0; i < 32; ++i, ++it) {
                int q = (int)(it & 1);
                hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cp, c->cy, Aptr), c->part.pAp,
                                   (const CGScal*)c->sc, 1);
                hipLaunchKernelGGL(k_cg_xr, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, (const CGScal*)c->sc, q);
                hipLaunchKernelGGL(k_cg_p, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, c->sc, q);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (iters) *iters = std::max(c->h_sc->itc[0], c->h_sc->itc[1]);
    }
    if (c->h_sc->zero_rhs) HIPCHK(hipMemsetAsync(c->cx, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    if (!std::isfinite(c->h_sc->bb) || !std::isfinite(c->h_sc->rr))
        return fail(YC_ERR_NONFINITE, "tangent: NaN/Inf in the right-hand side or the solution");
    if (c->h_sc->rr >= c->h_sc->thr && !c->h_sc->zero_rhs) return fail(YC_ERR_NOT_CONVERGED, "tangent CG stopped at maxit above tolerance");
    return YC_OK;
}

__global__ void k_axpy_neg(const double* a, const double* y, double* out, int n) {
    int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = a[i] - y[i];
}

// The outputs of direction d's tangent assembly (per-direction buffers, so that the first
// kernel runs for every direction at once).
AsmOut make_tan_out(yc_ctx* c, int d, int mode) {
    const size_t N = c->N, PV = (size_t)5 * c->P;
    AsmOut O = make_out(c);
    O.dA = mode == YC_GRAD_EXACT ? c->dA + (size_t)d * c->a_size : nullptr;
    O.dF = c->dF + (size_t)d * 5 * N; O.db = c->db + (size_t)d * PV;
    O.dis_slide = c->dis_slide + (size_t)d * 2 * N; O.dcomp = c->dcomp + (size_t)d * N;
    O.detach = mode == YC_GRAD_REF;
    O.b = c->bsave;   // the tangent kernels write db only; b is not touched
    O.dext = c->tdext ? c->tdext + (size_t)d * c->n : nullptr;
    O.shear_rec = c->tshear + (size_t)d * 90 * O.nseg;   // the direction's shear-tangent records
    return O;
}

// First part of every direction's tangent: normal tangents, then the first assembly
// kernel of all directions in one launch.
int tangent_assemble_all(yc_ctx* c, int mode, double h) {
    StateDev sd = make_state(c);
    hipLaunchKernelGGL(k_normals_tan, dim3((c->N + NBLK - 1) / NBLK, c->ndir), dim3(NBLK), 0, c->stream, sd,
                       (const double*)tan_state(c, 0).dx, tan_nrm(c, 0), (size_t)TAN_PLANES * c->N);
    c->h_tdev.resize(c->ndir);
    c->h_tout.resize(c->ndir);
    for (int d = 0; d < c->ndir; ++d) {
        c->h_tdev[d] = make_tan_dev(c, d, mode == YC_GRAD_REF);
        c->h_tout[d] = make_tan_out(c, d, mode);
    }
    HIPCHK(hipMemcpyAsync(c->tdev, c->h_tdev.data(), sizeof(ClothDev) * c->ndir, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->tout, c->h_tout.data(), sizeof(AsmOut) * c->ndir, hipMemcpyHostToDevice, c->stream));
    // make_tan_out gives every direction of the reference mode detach = 1 and no dA
    if (mode == YC_GRAD_REF)
        hipLaunchKernelGGL(k_assemble1_dirs<true>, dim3(c->nT, c->ndir), dim3(NBLK), 0, c->stream, (const ClothDev*)c->tdev,
                           (const AsmOut*)c->tout, h);
    else
        hipLaunchKernelGGL(k_assemble1_dirs<false>, dim3(c->nT, c->ndir), dim3(NBLK), 0, c->stream, (const ClothDev*)c->tdev,
                           (const AsmOut*)c->tout, h);
    // the shear segments' tangents, once per segment and direction, then the second kernel
    if (mode == YC_GRAD_REF) {
        hipLaunchKernelGGL(k_shear_elem_tan<true>, dim3((4 * c->Nc + 255) / 256, c->ndir), dim3(256), 0, c->stream,
                           (const ClothDev*)c->tdev, (const AsmOut*)c->tout, (const double*)c->comp);
        hipLaunchKernelGGL(k_assemble2_dirs<true>, dim3(c->nT, c->ndir), dim3(NBLK), 0, c->stream,
                           (const ClothDev*)c->tdev, (const AsmOut*)c->tout, h);
    } else {
        hipLaunchKernelGGL(k_shear_elem_tan<false>, dim3((4 * c->Nc + 255) / 256, c->ndir), dim3(256), 0, c->stream,
                           (const ClothDev*)c->tdev, (const AsmOut*)c->tout, (const double*)c->comp);
        hipLaunchKernelGGL(k_assemble2_dirs<false>, dim3(c->nT, c->ndir), dim3(NBLK), 0, c->stream,
                           (const ClothDev*)c->tdev, (const AsmOut*)c->tout, h);
    }
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// Rest of direction d's tangent, after tangent_assemble_all (the solution is in xsol, the
// right-hand side in bsave).  exact (mode 1): the derivative of the step as computed,
// A x = b differentiated as A dx = db - dA x.  ref (mode 0): the reference's gradient: A,
// M and K_v held constant, the non-jamming shear angle held constant, and the linear solve
// transposed (A^T dx = db), so that the transpose of this tangent map is the reference's
// backward solve with A (src/Solvers.cpp:51-53).
// queued: with the persistent CG the solve is only launched; its outcome accumulates in
// the context's scalars and the caller reads it once for all directions.
int tangent_dir(yc_ctx* c, int d, int mode, double h, double tol, int maxit, int* iters, bool queued = false) {
    const size_t PV = (size_t)5 * c->P;
    StateDev sd = make_state(c);
    TanState ts = tan_state(c, d);
    const AsmOut& O = c->h_tout[d];
    if (mode == YC_GRAD_EXACT) {
        hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->xsol, c->trhs, O.dA), (double*)nullptr,
                           (const CGScal*)c->sc, 0);
        hipLaunchKernelGGL(k_axpy_neg, dim3(grid_for((long long)PV)), dim3(BLK), 0, c->stream, (const double*)O.db,
                           (const double*)c->trhs, c->b, (int)PV);
    } else {
        HIPCHK(hipMemcpyAsync(c->b, O.db, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    const double* Au = mode == YC_GRAD_EXACT ? c->A : c->At;
    // the direction's velocity tangent of the previous step as the initial guess (the
    // primal step's warm start from its previous velocity, for the tangent)
    hipLaunchKernelGGL(k_pack_vel, dim3(grid_for(c->N)), dim3(BLK), 0, c->stream, (const double*)ts.dv, (const double*)ts.duvd,
                       c->W, c->Lg, c->N, c->Wp, c->P, c->cx);
    hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cx, c->cy, Au), (double*)nullptr,
                       (const CGScal*)c->sc, 0);
    int rc = YC_OK;
    if (queued) {
        hipLaunchKernelGGL(k_cg_init, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, 1);
        rc = launch_cgp(c, tol, maxit, Au);
    } else {
        rc = solve_on(c, Au, tol, maxit, iters, true);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_update_tan, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, sd, ts, (const double*)c->xsol,
                       (const double*)c->cx, (const double*)c->is_slide, (const double*)O.dis_slide, c->Wp, c->P, h);
    HIPCHK(hipGetLastError());
    return YC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* yc_last_error(void) { return g_err.c_str(); }

int yc_create(const yc_desc* desc, yc_ctx** out) {
    g_err.clear();
    if (!desc || !out) return fail(YC_ERR_ARG, "null argument");
    *out = nullptr;
    if (desc->width < 3 || desc->length < 3) return fail(YC_ERR_ARG, "cloth must be at least 3x3 crossings");
    if (desc->n_yarns < 1 || desc->n_yarns > 2) return fail(YC_ERR_ARG, "n_yarns must be 1 or 2");
    if (desc->n_handles < 0 || desc->n_handles > YC_MAX_HANDLES) return fail(YC_ERR_ARG, "too many handles");
    if ((long long)desc->width * desc->length > (1LL << 26)) return fail(YC_ERR_ARG, "cloth too large");
    yc_ctx* c = new yc_ctx();   // value-initialised: every pointer and count starts at zero
    c->d = *desc;
    c->W = desc->width; c->Lg = desc->length; c->N = c->W * c->Lg; c->Nc = (c->W - 2) * (c->Lg - 2);
    c->n = 3 * c->N + 2 * c->Nc;
    c->Wp = c->W + 4; c->Hp = c->Lg + 4;
    c->P = ((c->Wp * c->Hp + 63) / 64) * 64;
    c->graph_chunk = 32;
    c->ntx = (c->W + TILE - 1) / TILE; c->nty = (c->Lg + TILE - 1) / TILE; c->nT = c->ntx * c->nty;
    c->a_size = (size_t)c->nT * NSLOT * TILE_N;
    c->cg_variant = YC_CG_REFERENCE;
    int rc = YC_OK;
    auto chk = [&](int r) { if (r && !rc) rc = r; };
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return fail(YC_ERR_HIP, "no HIP device"); }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return fail(YC_ERR_HIP, "stream"); }
    size_t N = c->N;
    chk(dalloc(&c->x, 3 * N)); chk(dalloc(&c->v, 3 * N)); chk(dalloc(&c->xfix, 3 * N));
    chk(dalloc(&c->uv, 2 * N)); chk(dalloc(&c->uvd, 2 * N)); chk(dalloc(&c->uvbar, 2 * N)); chk(dalloc(&c->nrm, 3 * N));
    chk(dalloc(&c->flags, N));
    chk(dalloc(&c->A, c->a_size)); chk(dalloc(&c->F, 5 * N));
    chk(dalloc(&c->is_slide, 2 * N)); chk(dalloc(&c->comp, N)); chk(dalloc(&c->fepart, 2 * N));
    size_t PV = (size_t)5 * c->P;
    chk(dalloc(&c->b, PV)); chk(dalloc(&c->invd, PV)); chk(dalloc(&c->cx, PV)); chk(dalloc(&c->cr, PV));
    chk(dalloc(&c->cp, PV)); chk(dalloc(&c->cy, PV));
    chk(dalloc(&c->ext, (size_t)c->n));
    chk(dalloc(&c->vpad, PV)); chk(dalloc(&c->dprev, PV)); chk(dalloc(&c->dprev2, PV)); chk(dalloc(&c->Ae, PV));
    chk(dalloc(&c->cr1, PV)); chk(dalloc(&c->cw1, PV)); chk(dalloc(&c->cp1, PV));
    chk(dalloc(&c->cs0, PV)); chk(dalloc(&c->cs1, PV));
    c->cg1_kp = std::min((c->nT + BLK - 1) / BLK, CG1_MAXP);
    chk(dalloc(&c->cgslot, (size_t)2 * c->nT));
    chk(dalloc(&c->shear_rec, (size_t)90 * 4 * c->Nc + 1));
    {
        size_t nb = (size_t)c->Lg * (c->W - 2) + (size_t)c->W * (c->Lg - 2);
        chk(dalloc(&c->bend_rec, (size_t)3 * BEND_REC * nb + 1));
    }
    chk(dalloc(&c->cg_err, 1));
    chk(dalloc(&c->bepoch, 1));
    if (!rc) {   // batched solves tag in their own range: 2^30 + n, host-launched solves count from 1
        unsigned long long e0 = 1ULL << 30;
        if (hipMemcpy(c->bepoch, &e0, sizeof(e0), hipMemcpyHostToDevice) != hipSuccess) rc = fail(YC_ERR_HIP, "upload");
    }
    chk(dalloc(&c->part1, (size_t)12 * c->cg1_kp * BLK));
    c->part.nS = c->nT;
    c->part.nV = grid_for(5LL * c->N);
    chk(dalloc(&c->partials, (size_t)c->part.nS + 4 * (size_t)c->part.nV + 64));
    if (!rc) {
        c->part.pAp = c->partials;
        c->part.rr = c->part.pAp + c->part.nS;
        c->part.rz[0] = c->part.rr + c->part.nV;
        c->part.rz[1] = c->part.rz[0] + c->part.nV;
        c->part.bb = c->part.rz[1] + c->part.nV;
    }
    chk(dalloc(&c->sc, 1));
    chk(dalloc(&c->flag, 4));
    if (!rc && hipHostMalloc((void**)&c->h_sc, sizeof(CGScal)) != hipSuccess) rc = fail(YC_ERR_HIP, "pinned alloc");
    if (!rc && hipHostMalloc((void**)&c->h_flag, 4 * sizeof(int)) != hipSuccess) rc = fail(YC_ERR_HIP, "pinned alloc");
    if (rc) { yc_destroy(c); return rc; }
    c->coop_ok = cgp_fits<false>(c) ? 1 : 0;
    c->coop_bj_ok = cgp_fits<true>(c) ? 1 : 0;
    if (hipFuncSetAttribute((const void*)k_assemble1<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ASM1_LDS) != hipSuccess) {
        yc_destroy(c);
        return fail(YC_ERR_HIP, "assembly LDS size not supported");
    }
    // constructor rest state (src/Cloth.cpp:52-99) and flags
    std::vector<double> x(3 * N), uv(2 * N);
    std::vector<unsigned char> fl(N, 0);
    for (int j = 0; j < c->Lg; ++j)
        for (int i = 0; i < c->W; ++i) {
            int k = i + j * c->W;
            double px = i * desc->L, py = j * desc->L, pz = desc->hight;
            if (desc->pose == 0) { px = i * desc->L; py = 0.0; pz = desc->hight - j * desc->L; }
            x[k] = px; x[N + k] = py; x[2 * N + k] = pz;
            uv[k] = i * desc->L; uv[N + k] = j * desc->L;
            bool interior_ = i > 0 && i < c->W - 1 && j > 0 && j < c->Lg - 1;
            if (interior_ && pattern_flip(desc->pattern, i, j)) fl[k] |= 1;
        }
    for (int hnd = 0; hnd < desc->n_handles; ++hnd) {
        int hu = desc->handles[2 * hnd], hv = desc->handles[2 * hnd + 1];
        if (hu < 0 || hu >= c->W || hv < 0 || hv >= c->Lg) { yc_destroy(c); return fail(YC_ERR_ARG, "handle outside the cloth"); }
        fl[hu + hv * c->W] |= 2;
    }
    bool ok = hipMemcpy(c->x, x.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->xfix, x.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->uv, uv.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->uvbar, uv.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(c->flags, fl.data(), N, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { yc_destroy(c); return fail(YC_ERR_HIP, "upload"); }
    c->cd = make_dev(c);
    rc = refresh_geometry(c);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(YC_ERR_HIP, "sync");
    if (rc) { yc_destroy(c); return rc; }
    *out = c;
    return YC_OK;
}

void yc_destroy(yc_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    persist_forget(c->device, c->stream);
    adjoint_free(c);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->graph1) (void)hipGraphExecDestroy(c->graph1);
    void* bufs[] = {c->x, c->v, c->xfix, c->uv, c->uvd, c->uvbar, c->nrm, c->flags, c->A, c->F, c->is_slide, c->comp,
                    c->fepart, c->b, c->invd, c->cx, c->cr, c->cp, c->cy, c->ext, c->vpad, c->dprev, c->dprev2, c->Ae, c->partials, c->sc, c->flag,
                    c->cr1, c->cw1, c->cp1, c->cs0, c->cs1, c->part1, c->cgslot, c->cg_err,
                    c->tstate, c->dA, c->dF, c->db, c->dis_slide, c->dcomp, c->At, c->xsol, c->bsave, c->dsol, c->trhs, c->tdext, c->tshear,
                    c->shear_rec, c->bend_rec, c->bepoch, c->bj, c->Aexp, c->dexp};
    for (void* p : bufs) if (p) (void)hipFree(p);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    if (c->tdev) (void)hipFree(c->tdev);
    if (c->tout) (void)hipFree(c->tout);
    if (c->io_event) (void)hipEventDestroy(c->io_event);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int yc_sizes(const yc_ctx* c, int* N, int* Nc, int* n, int* nslots) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (N) *N = c->N;
    if (Nc) *Nc = c->Nc;
    if (n) *n = c->n;
    if (nslots) *nslots = NSLOT;
    return YC_OK;
}

int yc_set_state(yc_ctx* c, const double* lpos, const double* lposfix, const double* lvel,
                 const double* epos, const double* eposbar, const double* evel) {
    if (!c || !lpos || !lposfix || !lvel || !epos || !eposbar || !evel) return fail(YC_ERR_ARG, "null argument");
    size_t N = c->N;
    std::vector<double> a(3 * N), b2(3 * N), cv(3 * N), d(2 * N), e(2 * N), f(2 * N);
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            // state index u + v*length as the reference (src/Cloth.cpp:426), which addresses the
            // nodes one to one only when width == length; u + v*width equals it there and is
            // the trajectory loader's order (src/Loader.cpp:160) for non-square cloths
            int k = u + v * c->W, s = k;
            for (int q = 0; q < 3; ++q) { a[q * N + k] = lpos[3 * s + q]; b2[q * N + k] = lposfix[3 * s + q]; cv[q * N + k] = lvel[3 * s + q]; }
            for (int q = 0; q < 2; ++q) { d[q * N + k] = epos[2 * s + q]; e[q * N + k] = eposbar[2 * s + q]; f[q * N + k] = evel[2 * s + q]; }
        }
    HIPCHK(hipMemcpyAsync(c->x, a.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->xfix, b2.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->v, cv.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->uv, d.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->uvbar, e.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->uvd, f.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // a new state has no previous step: the extrapolated guesses start from v alone
    HIPCHK(hipMemsetAsync(c->dprev, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->dprev2, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    int rc = refresh_geometry(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->assembled = c->solved = false;
    return YC_OK;
}

// Orders the context's stream after the caller's (before == true) or the caller's after
// the context's.
static int join_streams(yc_ctx* c, hipStream_t caller, bool before) {
    if (!c->io_event && hipEventCreateWithFlags(&c->io_event, hipEventDisableTiming) != hipSuccess)
        return fail(YC_ERR_HIP, "hipEventCreateWithFlags");
    hipStream_t from = before ? caller : c->stream, to = before ? c->stream : caller;
    HIPCHK(hipEventRecord(c->io_event, from));
    HIPCHK(hipStreamWaitEvent(to, c->io_event, 0));
    return YC_OK;
}

int yc_set_state_device(yc_ctx* c, const double* lpos, const double* lposfix, const double* lvel, const double* epos,
                        const double* eposbar, const double* evel, void* stream) {
    if (!c || !lpos || !lposfix || !lvel || !epos || !eposbar || !evel) return fail(YC_ERR_ARG, "null argument");
    hipStream_t cs = (hipStream_t)stream;
    int rc = join_streams(c, cs, true);
    if (rc) return rc;
    StateIO io = {{lpos, lposfix, lvel, epos, eposbar, evel}, {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}};
    hipLaunchKernelGGL(k_state_in, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, make_state(c), io);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(c->dprev, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->dprev2, 0, (size_t)5 * c->P * sizeof(double), c->stream));
    rc = refresh_geometry(c);
    if (rc) return rc;
    c->assembled = c->solved = false;
    return join_streams(c, cs, false);
}

int yc_get_state_device(yc_ctx* c, double* lpos, double* lposfix, double* lvel, double* epos, double* eposbar,
                        double* evel, void* stream) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    hipStream_t cs = (hipStream_t)stream;
    int rc = join_streams(c, cs, true);
    if (rc) return rc;
    StateIO io = {{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}, {lpos, lposfix, lvel, epos, eposbar, evel}};
    hipLaunchKernelGGL(k_state_out, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, make_state(c), io);
    HIPCHK(hipGetLastError());
    return join_streams(c, cs, false);
}

int yc_get_state(yc_ctx* c, double* lpos, double* lposfix, double* lvel, double* epos, double* eposbar, double* evel) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    size_t N = c->N;
    std::vector<double> a(3 * N), b2(3 * N), cv(3 * N), d(2 * N), e(2 * N), f(2 * N);
    HIPCHK(hipMemcpyAsync(a.data(), c->x, 3 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(b2.data(), c->xfix, 3 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(cv.data(), c->v, 3 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(d.data(), c->uv, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(e.data(), c->uvbar, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(f.data(), c->uvd, 2 * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            int k = u + v * c->W, s = k;   // state index: see yc_set_state
            for (int q = 0; q < 3; ++q) {
                if (lpos) lpos[3 * s + q] = a[q * N + k];
                if (lposfix) lposfix[3 * s + q] = b2[q * N + k];
                if (lvel) lvel[3 * s + q] = cv[q * N + k];
            }
            for (int q = 0; q < 2; ++q) {
                if (epos) epos[2 * s + q] = d[q * N + k];
                if (eposbar) eposbar[2 * s + q] = e[q * N + k];
                if (evel) evel[2 * s + q] = f[q * N + k];
            }
        }
    return YC_OK;
}

int yc_set_ext_force(yc_ctx* c, const double* ext) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (!ext) { c->has_ext = false; return YC_OK; }
    HIPCHK(hipMemcpyAsync(c->ext, ext, (size_t)c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_ext = true;
    return YC_OK;
}

int yc_assemble(yc_ctx* c, double h) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    g_err.clear();
    return do_assemble(c, h);
}

}  // extern "C"

namespace {

// The structural pattern of a tiled stencil matrix (A, or M, K_x, K_v alone) as CSR of the
// global n x n matrix, rows in global order and columns sorted.
int export_tiles(yc_ctx* c, const double* Adev, int* rowptr, int* col, double* val, int* nnz) {
    std::vector<double> Am(c->a_size);
    HIPCHK(hipMemcpyAsync(Am.data(), Adev, Am.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // rows in global order; every structurally possible slot inside the grid
    std::vector<std::vector<std::pair<int, double>>> rows(c->n);
    const int offu[13] = {0, -1, 1, 0, 0, -2, 2, 0, 0, -1, 1, -1, 1};
    const int offv[13] = {0, 0, 0, -1, 1, 0, 0, -2, 2, -1, -1, 1, 1};
    auto inter = [&](int u, int v) { return u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1; };
    auto gL = [&](int u, int v, int a) { return 3 * (u + v * c->W) + a; };
    auto gE = [&](int u, int v, int a) { return 3 * c->N + 2 * ((u - 1) + (v - 1) * (c->W - 2)) + a; };
    auto inside = [&](int u, int v) { return u >= 0 && u < c->W && v >= 0 && v < c->Lg; };
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            const double* At = Am.data() + a_base(c->ntx, u, v);
            auto A = [&](int slot) { return At[(size_t)slot * TILE_N]; };
            for (int rc = 0; rc < 3; ++rc) {
                int row = gL(u, v, rc);
                for (int o = 0; o < 13; ++o) {
                    int uu = u + offu[o], vv = v + offv[o];
                    if (!inside(uu, vv)) continue;
                    for (int cc = 0; cc < 3; ++cc) rows[row].push_back({gL(uu, vv, cc), A(rc * 45 + o * 3 + cc)});
                }
                for (int d = -1; d <= 1; ++d) {
                    if (inside(u + d, v) && inter(u + d, v)) rows[row].push_back({gE(u + d, v, 0), A(rc * 45 + SLOT_LE + d + 1)});
                    if (inside(u, v + d) && inter(u, v + d)) rows[row].push_back({gE(u, v + d, 1), A(rc * 45 + SLOT_LE + 3 + d + 1)});
                }
            }
            if (!inter(u, v)) continue;
            for (int a = 0; a < 2; ++a) {
                int row = gE(u, v, a);
                int base = a == 0 ? SLOT_EU : SLOT_EV;
                for (int d = -1; d <= 1; ++d) {
                    int uu = u + (a == 0 ? d : 0), vv = v + (a == 1 ? d : 0);
                    for (int cc = 0; cc < 3; ++cc) rows[row].push_back({gL(uu, vv, cc), A(base + (d + 1) * 3 + cc)});
                    if (inter(uu, vv)) rows[row].push_back({gE(uu, vv, a), A(base + 9 + d + 1)});
                }
            }
        }
    int total = 0;
    for (auto& r : rows) { std::sort(r.begin(), r.end()); total += (int)r.size(); }
    if (nnz) *nnz = total;
    if (rowptr) {
        rowptr[0] = 0;
        for (int i = 0; i < c->n; ++i) rowptr[i + 1] = rowptr[i] + (int)rows[i].size();
    }
    if (col || val) {
        int j = 0;
        for (auto& r : rows)
            for (auto& e : r) { if (col) col[j] = e.first; if (val) val[j] = e.second; ++j; }
    }
    return YC_OK;
}

// The value assembly with the rows of one matrix alone (kind 0: M, 1: K_x, 2: K_v) into
// c->Aexp.  Everything else it writes (F, b, the friction outputs, the shear records) gets
// the values the last assembly wrote, from the same state and step; the Jacobi diagonal goes
// to scratch, so the context's system stays as assembled.
int assemble_kind(yc_ctx* c, int kind) {
    if (!c->Aexp) {
        int rc = dalloc(&c->Aexp, c->a_size);
        if (!rc) rc = dalloc(&c->dexp, (size_t)5 * c->P);
        if (rc) return rc;
    }
    AsmOut O = make_out(c);
    O.A = c->Aexp;
    O.invdiag = c->dexp;
    O.only_kind = kind;
    HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_assemble1<double>, dim3(c->nT), dim3(ASM1_THREADS), ASM1_LDS, c->stream, c->cd, O, c->last_h);
    if (c->Nc > 0)
        hipLaunchKernelGGL(k_shear_elem, dim3((4 * c->Nc + 255) / 256), dim3(256), 0, c->stream, c->cd, (const double*)c->comp,
                           c->shear_rec, 4 * c->Nc);
    hipLaunchKernelGGL(k_assemble2, dim3(c->Lg * ((c->W + TILE_N - 1) / TILE_N)), dim3(ASM2_THREADS), 0, c->stream, c->cd, O, c->last_h);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_flag, c->flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->h_flag[0]) return fail(YC_ERR_HIP, "assembly: stencil slot out of range");
    return YC_OK;
}

// Global row or column index -> (node u, v, component 0..4); false for an index past n.
bool global_dof(const yc_ctx* c, int g, int& u, int& v, int& comp) {
    if (g < 0 || g >= c->n) return false;
    if (g < 3 * c->N) {
        int k = g / 3;
        comp = g - 3 * k; u = k % c->W; v = k / c->W;
        return true;
    }
    int r = g - 3 * c->N, e = r / 2;
    comp = 3 + (r - 2 * e);
    u = e % (c->W - 2) + 1; v = e / (c->W - 2) + 1;
    return true;
}

}  // namespace

extern "C" {

int yc_export_csr(yc_ctx* c, int* rowptr, int* col, double* val, int* nnz) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (!c->assembled) return fail(YC_ERR_ARG, "export before assemble");
    return export_tiles(c, c->A, rowptr, col, val, nnz);
}

int yc_export_matrix(yc_ctx* c, int which, int* rowptr, int* col, double* val, int* nnz) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (!c->assembled) return fail(YC_ERR_ARG, "export before assemble");
    if (which < YC_MAT_A || which > YC_MAT_KV) return fail(YC_ERR_ARG, "unknown matrix");
    // the structural pattern (every slot inside the grid) is the same for all four
    if (which == YC_MAT_A || (!rowptr && !col && !val)) return export_tiles(c, c->A, rowptr, col, val, nnz);
    int rc = assemble_kind(c, which - 1);
    if (rc) return rc;
    return export_tiles(c, c->Aexp, rowptr, col, val, nnz);
}

int yc_load_system(yc_ctx* c, const int* rowptr, const int* col, const double* val, const double* b) {
    if (!c || !rowptr || !col || !val || !b) return fail(YC_ERR_ARG, "null argument");
    g_err.clear();
    std::vector<double> At(c->a_size, 0.0);
    std::vector<double> diag(c->n, 0.0);
    for (int r = 0; r < c->n; ++r)
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            int u, v, rc_, uc, vc, cc;
            if (!global_dof(c, col[j], uc, vc, cc)) return fail(YC_ERR_ARG, "column index out of range");
            global_dof(c, r, u, v, rc_);
            if (col[j] == r) diag[r] = val[j];
            const int s = slot_of(rc_, uc - u, vc - v, cc);
            if (s < 0) {
                if (val[j] != 0.0) return fail(YC_ERR_ARG, "entry outside the yarn stencil: row " + std::to_string(r) +
                                                               ", column " + std::to_string(col[j]));
                continue;
            }
            At[a_base(c->ntx, u, v) + (size_t)s * TILE_N] += val[j];
        }
    for (int i = 0; i < c->n; ++i)
        if (!std::isfinite(diag[i]) || !std::isfinite(b[i])) return fail(YC_ERR_NONFINITE, "NaN/Inf in A or b");
    for (double x : At)
        if (!std::isfinite(x)) return fail(YC_ERR_NONFINITE, "NaN/Inf in A or b");
    std::vector<double> bp, dp, inv(c->n);
    // the reference solver's diagonal preconditioner: 1/a_ii, or 1 where a_ii is zero
    for (int i = 0; i < c->n; ++i) inv[i] = diag[i] != 0.0 ? 1.0 / diag[i] : 1.0;
    to_padded(c, b, bp);
    to_padded(c, inv.data(), dp);
    HIPCHK(hipMemcpyAsync(c->A, At.data(), At.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->b, bp.data(), bp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->invd, dp.data(), dp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->bj_of = nullptr;
    c->assembled = true;
    c->solved = false;
    return YC_OK;
}

int yc_update_vel(yc_ctx* c, const double* new_vel, const double* is_slide, double h) {
    if (!c || !new_vel) return fail(YC_ERR_ARG, "null argument");
    std::vector<double> pad;
    to_padded(c, new_vel, pad);
    HIPCHK(hipMemcpyAsync(c->cx, pad.data(), pad.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> sl;
    if (is_slide) {
        sl.assign((size_t)2 * c->N, 0.0);
        for (int v = 1; v < c->Lg - 1; ++v)
            for (int u = 1; u < c->W - 1; ++u) {
                const int k = u + v * c->W, e = (u - 1) + (v - 1) * (c->W - 2);
                sl[k] = is_slide[2 * e];
                sl[(size_t)c->N + k] = is_slide[2 * e + 1];
            }
        HIPCHK(hipMemcpyAsync(c->is_slide, sl.data(), sl.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    if (!c->assembled)   // the warm-start history reads v_old in padded form, as after an assembly
        hipLaunchKernelGGL(k_pack_vel, dim3(grid_for(c->N)), dim3(BLK), 0, c->stream, (const double*)c->v, (const double*)c->uvd,
                           c->W, c->Lg, c->N, c->Wp, c->P, c->vpad);
    c->solved = true;
    int rc = do_update(c, h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return YC_OK;
}

int yc_export_vec(yc_ctx* c, int which, double* out) {
    if (!c || !out) return fail(YC_ERR_ARG, "null argument");
    size_t N = c->N;
    if (which == 0) {
        std::vector<double> F(5 * N);
        HIPCHK(hipMemcpyAsync(F.data(), c->F, F.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int v = 0; v < c->Lg; ++v)
            for (int u = 0; u < c->W; ++u) {
                int k = u + v * c->W;
                for (int a = 0; a < 3; ++a) out[3 * k + a] = F[a * N + k];
                if (u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1) {
                    int e = (u - 1) + (v - 1) * (c->W - 2);
                    for (int a = 0; a < 2; ++a) out[3 * N + 2 * e + a] = F[(3 + a) * N + k];
                }
            }
        return YC_OK;
    }
    if (which == 2 || which == 4) {   // velocities (get_SpaVel) or positions (get_SpaPos)
        std::vector<double> v3(3 * N), v2(2 * N);
        HIPCHK(hipMemcpyAsync(v3.data(), which == 2 ? c->v : c->x, v3.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(v2.data(), which == 2 ? c->uvd : c->uv, v2.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int v = 0; v < c->Lg; ++v)
            for (int u = 0; u < c->W; ++u) {
                int k = u + v * c->W;
                for (int a = 0; a < 3; ++a) out[3 * k + a] = v3[a * N + k];
                if (u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1) {
                    int e = (u - 1) + (v - 1) * (c->W - 2);
                    for (int a = 0; a < 2; ++a) out[3 * N + 2 * e + a] = v2[a * N + k];
                }
            }
        return YC_OK;
    }
    if (which == 1 || which == 3) {
        std::vector<double> pad((size_t)5 * c->P);
        HIPCHK(hipMemcpyAsync(pad.data(), which == 1 ? c->b : c->cx, pad.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        from_padded(c, pad, out);
        return YC_OK;
    }
    return fail(YC_ERR_ARG, "unknown vector");
}

int yc_export_crossing(yc_ctx* c, double* is_slide, double* comp, double* fe_partial) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    size_t N = c->N;
    std::vector<double> sl(2 * N), cp(N), fp(2 * N);
    HIPCHK(hipMemcpyAsync(sl.data(), c->is_slide, sl.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(cp.data(), c->comp, cp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fp.data(), c->fepart, fp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 1; v < c->Lg - 1; ++v)
        for (int u = 1; u < c->W - 1; ++u) {
            int k = u + v * c->W, e = (u - 1) + (v - 1) * (c->W - 2);
            for (int a = 0; a < 2; ++a) {
                if (is_slide) is_slide[2 * e + a] = sl[a * N + k];
                if (fe_partial) fe_partial[2 * e + a] = fp[a * N + k];
            }
            if (comp) comp[e] = cp[k];
        }
    return YC_OK;
}

int yc_export_normals(yc_ctx* c, double* out) {
    if (!c || !out) return fail(YC_ERR_ARG, "null argument");
    size_t N = c->N;
    std::vector<double> nr(3 * N);
    HIPCHK(hipMemcpyAsync(nr.data(), c->nrm, nr.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < N; ++k) for (int a = 0; a < 3; ++a) out[3 * k + a] = nr[a * N + k];
    return YC_OK;
}

int yc_solve(yc_ctx* c, const double* x0, double tol, int maxit, int* iters, double* relres) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    g_err.clear();
    return do_solve(c, x0, tol, maxit, iters, relres);
}

int yc_update(yc_ctx* c, double h) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    int rc = do_update(c, h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return YC_OK;
}

int yc_step(yc_ctx* c, double h, const double* x0, double tol, int maxit, yc_step_info* info) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    g_err.clear();
    int it = 0;
    double rr = 0;
    int rc = do_assemble(c, h);
    if (!rc) rc = do_solve(c, x0, tol, maxit, &it, &rr);
    if (info) { info->cg_iters = it; info->cg_relres = rr; info->status = rc; }
    if (rc == YC_ERR_NONFINITE || rc == YC_ERR_HIP || rc == YC_ERR_ARG) return rc;
    int rc2 = do_update(c, h);
    if (rc2) return rc2;
    HIPCHK(hipStreamSynchronize(c->stream));
    return rc;
}

}  // extern "C"

namespace {

int warm_guess(yc_ctx* c, yc_mlp* m);

// Steps back to back with the persistent CG and no host synchronisation until the end:
// assembly flags and the solves' outcomes accumulate on the device.
// guess (without a network): YC_GUESS_ZERO, YC_GUESS_PREV or YC_GUESS_EXTRAP.
int run_async(yc_ctx* c, yc_mlp* m, double h, int steps, double tol, int maxit, long long* total_iters,
              int guess = YC_GUESS_ZERO) {
    g_err.clear();
    HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync((char*)c->sc + offsetof(CGScal, total_iters), 0, sizeof(long long) + 2 * sizeof(int), c->stream));
    const bool x0 = m != nullptr || guess != YC_GUESS_ZERO;
    // with the network and the block-Jacobi preconditioner, the rows (from A) and the
    // network's guess (from A, b and the state) are independent, and warm_guess computes
    // them in one launch side by side
    for (int s = 0; s < steps; ++s) {
        int rc = do_assemble(c, h, true, m != nullptr);
        if (!rc && m) rc = warm_guess(c, m);
        if (!rc && !m && guess != YC_GUESS_ZERO) {
            hipLaunchKernelGGL(k_guess_x0, dim3(grid_for(5LL * c->P)), dim3(BLK), 0, c->stream, (const double*)c->vpad,
                               (const double*)c->dprev, (const double*)c->dprev2, guess, 5 * c->P, c->cx);
            HIPCHK(hipGetLastError());
        }
        if (!rc) rc = do_solve(c, nullptr, tol, maxit, nullptr, nullptr, x0, true);
        if (!rc) rc = do_update(c, h);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_flag, c->flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (total_iters) *total_iters = c->h_sc->total_iters;
    if (c->h_flag[2]) {
        (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
        return fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
    }
    if (c->h_flag[0]) return fail(YC_ERR_HIP, "assembly: stencil slot out of range");
    if (c->h_flag[1] || c->h_sc->nonfinite) return fail(YC_ERR_NONFINITE, "NaN/Inf in A, b or a solution");
    if (c->h_sc->nonconv) return fail(YC_ERR_NOT_CONVERGED, "a CG solve stopped at maxit above tolerance");
    return YC_OK;
}

}  // namespace

extern "C" {

int yc_run(yc_ctx* c, double h, int steps, double tol, int maxit, long long* total_iters) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->cg_variant == YC_CG_PERSISTENT && c->coop_ok) return run_async(c, nullptr, h, steps, tol, maxit, total_iters);
    long long tot = 0;
    int worst = YC_OK;
    for (int s = 0; s < steps; ++s) {
        yc_step_info info;
        int rc = yc_step(c, h, nullptr, tol, maxit, &info);
        tot += info.cg_iters;
        if (rc == YC_ERR_NOT_CONVERGED) worst = rc;
        else if (rc) { if (total_iters) *total_iters = tot; return rc; }
    }
    if (total_iters) *total_iters = tot;
    return worst;
}

int yc_run_guess(yc_ctx* c, int guess, double h, int steps, double tol, int maxit, long long* total_iters) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (guess < YC_GUESS_ZERO || guess > YC_GUESS_QUADRATIC) return fail(YC_ERR_ARG, "unknown guess");
    if (c->cg_variant != YC_CG_PERSISTENT || !c->coop_ok) return fail(YC_ERR_ARG, "yc_run_guess runs the persistent CG");
    return run_async(c, nullptr, h, steps, tol, maxit, total_iters, guess);
}

int yc_tangent_end(yc_ctx* c) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    (void)hipStreamSynchronize(c->stream);
    double** bufs[] = {&c->tstate, &c->dA, &c->dF, &c->db, &c->dis_slide, &c->dcomp, &c->At, &c->xsol, &c->bsave, &c->dsol, &c->trhs,
                       &c->tdext, &c->tshear};
    for (double** p : bufs) { if (*p) (void)hipFree(*p); *p = nullptr; }
    if (c->tdev) (void)hipFree(c->tdev);
    if (c->tout) (void)hipFree(c->tout);
    c->tdev = nullptr;
    c->tout = nullptr;
    c->ndir = 0;
    c->tpar.clear();
    return YC_OK;
}

int yc_tangent_begin(yc_ctx* c, int ndir, const double* dparam, const double* dstate) {
    if (!c || ndir <= 0 || !dparam) return fail(YC_ERR_ARG, "bad argument");
    g_err.clear();
    yc_tangent_end(c);
    size_t N = c->N, PV = (size_t)5 * c->P;
    int rc = YC_OK;
    auto chk = [&](int r) { if (r && !rc) rc = r; };
    chk(dalloc(&c->tstate, (size_t)ndir * TAN_PLANES * N));
    // per-direction outputs of the tangent assembly (all directions assemble in one launch)
    chk(dalloc(&c->dA, (size_t)ndir * c->a_size)); chk(dalloc(&c->At, c->a_size));
    chk(dalloc(&c->dF, (size_t)ndir * 5 * N)); chk(dalloc(&c->db, (size_t)ndir * PV));
    chk(dalloc(&c->dis_slide, (size_t)ndir * 2 * N)); chk(dalloc(&c->dcomp, (size_t)ndir * N));
    if (!rc && hipMalloc((void**)&c->tdev, sizeof(ClothDev) * ndir) != hipSuccess) rc = fail(YC_ERR_HIP, "hipMalloc");
    chk(dalloc(&c->tshear, (size_t)ndir * 90 * 4 * c->Nc + 1));
    if (!rc && hipMalloc((void**)&c->tout, sizeof(AsmOut) * ndir) != hipSuccess) rc = fail(YC_ERR_HIP, "hipMalloc");
    chk(dalloc(&c->xsol, PV)); chk(dalloc(&c->bsave, PV)); chk(dalloc(&c->dsol, PV)); chk(dalloc(&c->trhs, PV));
    if (rc) { yc_tangent_end(c); return rc; }
    c->ndir = ndir;
    c->tpar.assign(dparam, dparam + (size_t)ndir * YC_NPARAM);
    if (dstate) {
        std::vector<double> t((size_t)TAN_PLANES * N, 0.0);
        for (int d = 0; d < ndir; ++d) {
            const double* src = dstate + (size_t)d * 15 * N;
            const double *lpos = src, *lposfix = src + 3 * N, *lvel = src + 6 * N;
            const double *epos = src + 9 * N, *eposbar = src + 11 * N, *evel = src + 13 * N;
            for (int v = 0; v < c->Lg; ++v)
                for (int u = 0; u < c->W; ++u) {
                    int k = u + v * c->W, st = k;   // state index: see yc_set_state
                    for (int q = 0; q < 3; ++q) {
                        t[0 * N + q * N + k] = lpos[3 * st + q];
                        t[3 * N + q * N + k] = lvel[3 * st + q];
                        t[6 * N + q * N + k] = lposfix[3 * st + q];
                    }
                    for (int q = 0; q < 2; ++q) {
                        t[9 * N + q * N + k] = epos[2 * st + q];
                        t[11 * N + q * N + k] = evel[2 * st + q];
                        t[13 * N + q * N + k] = eposbar[2 * st + q];
                    }
                }
            HIPCHK(hipMemcpyAsync(c->tstate + (size_t)d * TAN_PLANES * N, t.data(), t.size() * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    return YC_OK;
}

int yc_tangent_set_ext(yc_ctx* c, const double* dext) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->ndir <= 0) return fail(YC_ERR_ARG, "yc_tangent_set_ext before yc_tangent_begin");
    if (!dext) {
        if (c->tdext) { HIPCHK(hipStreamSynchronize(c->stream)); (void)hipFree(c->tdext); c->tdext = nullptr; }
        return YC_OK;
    }
    const size_t cnt = (size_t)c->ndir * c->n;
    if (!c->tdext) {
        int rc = dalloc(&c->tdext, cnt);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->tdext, dext, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return YC_OK;
}

int yc_tangent_get(yc_ctx* c, int dir, double* dlpos, double* dlposfix, double* dlvel, double* depos, double* deposbar,
                   double* devel) {
    if (!c || dir < 0 || dir >= c->ndir) return fail(YC_ERR_ARG, "bad direction");
    size_t N = c->N;
    std::vector<double> t((size_t)TAN_PLANES * N);
    HIPCHK(hipMemcpyAsync(t.data(), c->tstate + (size_t)dir * TAN_PLANES * N, t.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            int k = u + v * c->W, st = k;   // state index: see yc_set_state
            for (int q = 0; q < 3; ++q) {
                if (dlpos) dlpos[3 * st + q] = t[0 * N + q * N + k];
                if (dlvel) dlvel[3 * st + q] = t[3 * N + q * N + k];
                if (dlposfix) dlposfix[3 * st + q] = t[6 * N + q * N + k];
            }
            for (int q = 0; q < 2; ++q) {
                if (depos) depos[2 * st + q] = t[9 * N + q * N + k];
                if (devel) devel[2 * st + q] = t[11 * N + q * N + k];
                if (deposbar) deposbar[2 * st + q] = t[13 * N + q * N + k];
            }
        }
    return YC_OK;
}

int yc_step_tangent(yc_ctx* c, double h, int mode, double tol, int maxit, yc_step_info* info) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->ndir <= 0) return fail(YC_ERR_ARG, "yc_step_tangent before yc_tangent_begin");
    if (mode != YC_GRAD_REF && mode != YC_GRAD_EXACT) return fail(YC_ERR_ARG, "unknown gradient mode");
    g_err.clear();
    int it = 0;
    double rr = 0;
    int rc = do_assemble(c, h);
    if (!rc) rc = do_solve(c, nullptr, tol, maxit, &it, &rr);
    if (info) { info->cg_iters = it; info->cg_relres = rr; info->status = rc; }
    if (rc) return rc;   // a tangent needs the converged step
    const size_t PV = (size_t)5 * c->P;
    HIPCHK(hipMemcpyAsync(c->xsol, c->cx, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bsave, c->b, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (mode == YC_GRAD_REF) {
        if (c->bj_of == c->At) c->bj_of = nullptr;   // At is rewritten
        hipLaunchKernelGGL(k_transpose, dim3(c->nT), dim3(NBLK), 0, c->stream, (const double*)c->A, c->At, c->W, c->Lg, c->ntx);
    }
    HIPCHK(hipGetLastError());
    rc = tangent_assemble_all(c, mode, h);
    // with the persistent CG the directions' solves are queued back to back and their
    // outcomes read once (the solves' totals are reset first)
    const bool queued = c->coop_ok != 0;
    if (queued)
        HIPCHK(hipMemsetAsync((char*)c->sc + offsetof(CGScal, total_iters), 0, sizeof(long long) + 2 * sizeof(int), c->stream));
    for (int d = 0; d < c->ndir && !rc; ++d) rc = tangent_dir(c, d, mode, h, tol, maxit, nullptr, queued);
    if (!rc && queued) {
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->h_flag[2]) {
            (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
            rc = fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
        } else if (c->h_sc->nonfinite) {
            rc = fail(YC_ERR_NONFINITE, "tangent: NaN/Inf in the right-hand side or the solution");
        } else if (c->h_sc->nonconv) {
            rc = fail(YC_ERR_NOT_CONVERGED, "tangent CG stopped at maxit above tolerance");
        }
    }
    // restore the step's own right-hand side and solution, then update
    HIPCHK(hipMemcpyAsync(c->b, c->bsave, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->cx, c->xsol, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (rc) return rc;
    rc = do_update(c, h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return YC_OK;
}

int yc_set_cg_variant(yc_ctx* c, int variant) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (variant != YC_CG_REFERENCE && variant != YC_CG_FUSED && variant != YC_CG_PERSISTENT)
        return fail(YC_ERR_ARG, "unknown CG variant");
    if (variant != YC_CG_REFERENCE && c->nT > CG1_MAXP * BLK)
        return fail(YC_ERR_ARG, "fused CG supports up to 1024 matrix tiles (256 x 256 crossings)");
    if (variant == YC_CG_PERSISTENT && !c->coop_ok)
        return fail(YC_ERR_ARG, "persistent CG: the cloth's tiles do not fit on the device at once");
    c->cg_variant = variant;
    return YC_OK;
}

int yc_set_preconditioner(yc_ctx* c, int kind) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (kind != YC_PRECOND_JACOBI && kind != YC_PRECOND_BLOCK2) return fail(YC_ERR_ARG, "unknown preconditioner");
    if (kind == YC_PRECOND_BLOCK2) {
        if (!c->coop_bj_ok) return fail(YC_ERR_ARG, "block Jacobi: the cloth's persistent CG does not fit on the device");
        if (!c->bj) {
            int rc = dalloc(&c->bj, (size_t)5 * c->P * 20);
            if (rc) return rc;
        }
    }
    c->precond = kind;
    c->bj_of = nullptr;
    return YC_OK;
}

int yc_export_precond(yc_ctx* c, double* rows) {
    if (!c || !rows) return fail(YC_ERR_ARG, "null argument");
    if (c->precond != YC_PRECOND_BLOCK2) return fail(YC_ERR_ARG, "the context does not use the block-Jacobi preconditioner");
    if (!c->assembled) return fail(YC_ERR_ARG, "yc_export_precond before yc_assemble");
    int rc = bj_factor(c, c->A);
    if (rc) return rc;
    std::vector<double> pad((size_t)5 * c->P * 20);
    HIPCHK(hipMemcpyAsync(pad.data(), c->bj, pad.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->Lg; ++v)
        for (int u = 0; u < c->W; ++u) {
            const int k = u + v * c->W, pk = (u + 2) + (v + 2) * c->Wp;
            for (int q = 0; q < 3; ++q)
                std::memcpy(rows + (size_t)(3 * k + q) * 20, pad.data() + ((size_t)q * c->P + pk) * 20, 20 * sizeof(double));
            if (u > 0 && u < c->W - 1 && v > 0 && v < c->Lg - 1) {
                const int e = (u - 1) + (v - 1) * (c->W - 2);
                for (int q = 0; q < 2; ++q)
                    std::memcpy(rows + ((size_t)3 * c->N + 2 * e + q) * 20, pad.data() + ((size_t)(3 + q) * c->P + pk) * 20,
                                20 * sizeof(double));
            }
        }
    return YC_OK;
}

int yc_time_spmv(yc_ctx* c, int reps, double* avg_ms) {
    if (!c || reps <= 0 || !avg_ms) return fail(YC_ERR_ARG, "bad argument");
    // the CG loop's stencil product y = A p (partials of p.y included, as in the loop)
    SpmvArgs a = make_spmv(c, c->cp, c->cy);
    int st = time_launches(c->stream, reps, avg_ms, [&] {
        hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, a, c->part.pAp, (const CGScal*)c->sc, 0);
    });
    return st;
}

int yc_time_assemble(yc_ctx* c, double h, int reps, double* avg_ms) {
    if (!c || reps <= 0 || !avg_ms) return fail(YC_ERR_ARG, "bad argument");
    int rc = YC_OK;
    int st = time_launches(c->stream, reps, avg_ms, [&] {
        if (!rc) rc = do_assemble(c, h, true);
    });
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return st;
}

int yc_time_solve(yc_ctx* c, double tol, int maxit, int reps, double* avg_ms, int* iters) {
    if (!c || reps <= 0 || !avg_ms) return fail(YC_ERR_ARG, "bad argument");
    if (!c->assembled) return fail(YC_ERR_ARG, "yc_time_solve before yc_assemble");
    if (c->cg_variant != YC_CG_PERSISTENT || !c->coop_ok) return fail(YC_ERR_ARG, "yc_time_solve times the persistent CG");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(YC_ERR_HIP, "hipEventCreate"); }
    double tot = 0.0;
    // the block-Jacobi rows of the one matrix are made current before the first launch
    // (untimed): the event pairs time the solve kernel
    int rc = c->precond == YC_PRECOND_BLOCK2 ? bj_factor(c, c->A) : YC_OK;
    for (int i = 0; i <= reps && !rc; ++i) {   // the first launch is untimed
        hipLaunchKernelGGL(k_cg_init, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, 0);
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(YC_ERR_HIP, "hipEventRecord"); break; }
        rc = launch_cgp(c, tol, maxit);
        if (rc) break;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = fail(YC_ERR_HIP, "event"); break; }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = fail(YC_ERR_HIP, "hipEventElapsedTime"); break; }
        if (i > 0) tot += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIPCHK(hipMemcpy(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost));
    *avg_ms = tot / reps;
    if (iters) *iters = c->h_sc->fin_iters;
    return YC_OK;
}

double yc_cg_iteration_bytes(const yc_ctx* c) {
    if (!c) return 0.0;
    // one Jacobi-PCG iteration streamed from memory: A once, and 15 vector passes
    // (p, y, x, r, z, D^-1 reads and writes) of 8 bytes per DOF; with the block-Jacobi
    // preconditioner its 20-coefficient rows replace D^-1 (152 more bytes per DOF)
    double b = yc_spmv_bytes(c) - 16.0 * c->n + 120.0 * c->n;
    return c->precond == YC_PRECOND_BLOCK2 ? b + 152.0 * c->n : b;
}

double yc_spmv_bytes(const yc_ctx* c) {
    if (!c) return 0.0;
    // structural nonzeros of A (LL 13-point stencil + LE/EL/EE along the yarns)
    double W = c->W;
    double nnz = 9.0 * (13.0 * W * W - 20.0 * W - 4.0) + 14.0 * (W - 2.0) * (3.0 * W - 8.0);
    // values once, p read and A p written once per DOF
    return 8.0 * nnz + 16.0 * c->n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned warm start
// ---------------------------------------------------------------------------
namespace {

int mlp_refresh_bf16(yc_mlp* m) {
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, m->stream, (const float*)m->param,
                       m->W1, m->W2, m->W3, m->W4);
    HIPCHK(hipGetLastError());
    return YC_OK;
}

int noalloc_scratch_dev(float** out);

MlpW mlp_w(const yc_mlp* m) {
    MlpW w;
    w.W1 = m->W1; w.W2 = m->W2; w.W3 = m->W3; w.W4 = m->W4; w.b = m->param;
    return w;
}

int mlp_reserve_batch(yc_mlp* m, int rows) {
    if (rows <= m->cap_batch) return YC_OK;
    void* bufs[] = {m->idx, m->sX0, m->sZ, m->sH, m->Y, m->part};
    for (void* p : bufs) if (p) (void)hipFree(p);
    int nblk = (rows + MLP_TILE - 1) / MLP_TILE;
    HIPCHK(hipMalloc((void**)&m->idx, (size_t)rows * sizeof(int)));
    HIPCHK(hipMalloc((void**)&m->sX0, (size_t)rows * MLP_IN * 2));
    HIPCHK(hipMalloc((void**)&m->sZ, (size_t)3 * rows * MLP_H * 4));
    HIPCHK(hipMalloc((void**)&m->sH, (size_t)3 * rows * MLP_H * 2));
    HIPCHK(hipMalloc((void**)&m->Y, (size_t)rows * MLP_OUT * 4));
    HIPCHK(hipMalloc((void**)&m->part, (size_t)nblk * (MLP_NPARAM + 1) * 4));
    m->cap_batch = rows;
    return YC_OK;
}

// forward + backward (+ Adam when lr > 0) on rows idx_dev[0..rows) of the dataset
int mlp_batch(yc_mlp* m, const int* idx_dev, int rows, float lr) {
    int nblk = (rows + MLP_TILE - 1) / MLP_TILE;
    MlpFwdArgs f;
    f.X = m->X; f.idx = idx_dev; f.rows = rows; f.in_mean = m->in_mean; f.in_istd = m->in_istd; f.w = mlp_w(m);
    f.Y = m->Y; f.sX0 = m->sX0; f.sZ = m->sZ; f.sH = m->sH;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(nblk), dim3(256), 0, m->stream, f);
    MlpBwdArgs b;
    b.rows = rows; b.sX0 = m->sX0; b.sZ = m->sZ; b.sH = m->sH; b.Y = m->Y; b.T = m->T; b.M = m->M; b.idx = idx_dev;
    b.w = mlp_w(m); b.part = m->part; b.inv_count = 1.0f / (float)(rows * MLP_OUT);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nblk), dim3(256), 0, m->stream, b);
    float b1t = 1.0f, b2t = 1.0f;
    if (lr > 0.0f) {
        m->adam_t += 1;
        b1t = std::pow(0.9f, (float)m->adam_t);
        b2t = std::pow(0.999f, (float)m->adam_t);
    }
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, (const float*)m->part, nblk,
                       m->grad, m->param, m->m1, m->m2, lr, b1t, b2t);
    HIPCHK(hipGetLastError());
    if (lr > 0.0f) return mlp_refresh_bf16(m);
    return YC_OK;
}

FeatArgs feat_args(yc_ctx* c) {
    FeatArgs fa;
    fa.uv = c->uv; fa.uvbar = c->uvbar; fa.flags = c->flags;
    fa.b = c->b; fa.invd = c->invd; fa.Ae = c->Ae; fa.dprev = c->dprev; fa.dprev2 = c->dprev2;
    fa.W = c->W; fa.Lg = c->Lg; fa.N = c->N; fa.Wp = c->Wp; fa.P = c->P; fa.L = c->d.L;
    return fa;
}

// x_q = v + (2 dprev - dprev2) into xe, then A x_q into c->Ae (the features' residual)
int extrapolate_and_apply(yc_ctx* c, double* xe) {
    hipLaunchKernelGGL(k_guess_x0, dim3(grid_for(5LL * c->P)), dim3(BLK), 0, c->stream, (const double*)c->vpad,
                       (const double*)c->dprev, (const double*)c->dprev2, YC_GUESS_QUADRATIC, 5 * c->P, xe);
    hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, xe, c->Ae), (double*)nullptr,
                       (const CGScal*)c->sc, 0);
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// The network's initial guess in c->cx: x_q, its stencil product, then one launch of
// features + network + correction.
// With the block-Jacobi preconditioner and rows not yet those of A, the rows are computed
// in the network's launch (k_mlp_warm_bj), beside it.
int warm_guess(yc_ctx* c, yc_mlp* m) {
    int rc = extrapolate_and_apply(c, c->cx);
    if (rc) return rc;
    const int nmlp = (c->N + MLP_TILE - 1) / MLP_TILE;
    if (c->precond == YC_PRECOND_BLOCK2 && c->bj_of != c->A) {
        BjArgs bj;
        bj.A = c->A; bj.brow = c->bj;
        bj.W = c->W; bj.Lg = c->Lg; bj.Wp = c->Wp; bj.P = c->P; bj.ntx = c->ntx;
        bj.ngx = (c->W + 1) / 2;
        bj.ngroups = bj.ngx * ((c->Lg + 1) / 2);
        bj.nbj = (bj.ngroups + 4 * BJF_GROUPS - 1) / (4 * BJF_GROUPS);
        hipLaunchKernelGGL(k_mlp_warm_bj, dim3(bj.nbj + nmlp), dim3(256), 0, c->stream, feat_args(c), mlp_w(m),
                           (const float*)m->in_mean, (const float*)m->in_istd, (const float*)m->out_std, c->cx, bj);
        HIPCHK(hipGetLastError());
        c->bj_of = c->A;
        return YC_OK;
    }
    hipLaunchKernelGGL(k_mlp_warm, dim3(nmlp), dim3(256), 0, c->stream, feat_args(c), mlp_w(m),
                       (const float*)m->in_mean, (const float*)m->in_istd, (const float*)m->out_std, c->cx);
    HIPCHK(hipGetLastError());
    return YC_OK;
}

}  // namespace

extern "C" {

int yc_mlp_param_count(void) { return MLP_NPARAM; }

int yc_mlp_create(yc_mlp** out) {
    if (!out) return fail(YC_ERR_ARG, "null argument");
    yc_mlp* m = new yc_mlp();
    std::memset((void*)m, 0, sizeof(yc_mlp));
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) { delete m; return fail(YC_ERR_HIP, "stream"); }
    int rc = YC_OK;
    auto chk = [&](int r) { if (r && !rc) rc = r; };
    chk(dalloc(&m->param, MLP_NPARAM + 1)); chk(dalloc(&m->m1, MLP_NPARAM + 1)); chk(dalloc(&m->m2, MLP_NPARAM + 1));
    chk(dalloc(&m->grad, MLP_NPARAM + 1));
    chk(dalloc(&m->W1, MLP_H * MLP_IN)); chk(dalloc(&m->W2, MLP_H * MLP_H)); chk(dalloc(&m->W3, MLP_H * MLP_H));
    chk(dalloc(&m->W4, MLP_OUTP * MLP_H));
    chk(dalloc(&m->in_mean, MLP_IN)); chk(dalloc(&m->in_istd, MLP_IN)); chk(dalloc(&m->out_std, MLP_OUT));
    chk(noalloc_scratch_dev(&m->scratch));   // the device's staging of yc_mlp_forward_noalloc
    if (rc) { yc_mlp_destroy(m); return rc; }
    std::vector<float> one(MLP_IN, 1.0f);
    if (hipMemcpy(m->in_istd, one.data(), MLP_IN * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->out_std, one.data(), MLP_OUT * 4, hipMemcpyHostToDevice) != hipSuccess) {
        yc_mlp_destroy(m);
        return fail(YC_ERR_HIP, "upload");
    }
    *out = m;
    return YC_OK;
}

void yc_mlp_destroy(yc_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->param, m->m1, m->m2, m->grad, m->W1, m->W2, m->W3, m->W4, m->in_mean, m->in_istd, m->out_std,
                    m->X, m->T, m->M, m->idx, m->sX0, m->sZ, m->sH, m->Y, m->part};
    for (void* p : bufs) if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int yc_mlp_set_params(yc_mlp* m, const float* p) {
    if (!m || !p) return fail(YC_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(m->param, p, MLP_NPARAM * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemsetAsync(m->m1, 0, (MLP_NPARAM + 1) * 4, m->stream));
    HIPCHK(hipMemsetAsync(m->m2, 0, (MLP_NPARAM + 1) * 4, m->stream));
    m->adam_t = 0;
    int rc = mlp_refresh_bf16(m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return YC_OK;
}

int yc_mlp_get_params(yc_mlp* m, float* p) {
    if (!m || !p) return fail(YC_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(p, m->param, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return YC_OK;
}

int yc_mlp_set_norm(yc_mlp* m, const float* in_mean, const float* in_std, const float* out_std) {
    if (!m || !in_mean || !in_std || !out_std) return fail(YC_ERR_ARG, "null argument");
    float istd[MLP_IN];
    for (int i = 0; i < MLP_IN; ++i) istd[i] = in_std[i] > 0.0f ? 1.0f / in_std[i] : 0.0f;
    HIPCHK(hipMemcpyAsync(m->in_mean, in_mean, MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->in_istd, istd, MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->out_std, out_std, MLP_OUT * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return YC_OK;
}

int yc_mlp_set_data(yc_mlp* m, const float* X, const float* T, const float* M, int rows) {
    if (!m || !X || !T || !M || rows <= 0) return fail(YC_ERR_ARG, "bad argument");
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    if (m->M) (void)hipFree(m->M);
    m->X = m->T = m->M = nullptr;
    HIPCHK(hipMalloc((void**)&m->X, (size_t)rows * MLP_IN * 4));
    HIPCHK(hipMalloc((void**)&m->T, (size_t)rows * MLP_OUT * 4));
    HIPCHK(hipMalloc((void**)&m->M, (size_t)rows * MLP_OUT * 4));
    HIPCHK(hipMemcpyAsync(m->X, X, (size_t)rows * MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->T, T, (size_t)rows * MLP_OUT * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->M, M, (size_t)rows * MLP_OUT * 4, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->nrows = rows;
    return YC_OK;
}

int yc_mlp_forward(yc_mlp* m, const float* X, int rows, float* Y) {
    if (!m || !X || !Y || rows <= 0) return fail(YC_ERR_ARG, "bad argument");
    float *dX = nullptr, *dY = nullptr;
    HIPCHK(hipMalloc((void**)&dX, (size_t)rows * MLP_IN * 4));
    HIPCHK(hipMalloc((void**)&dY, (size_t)rows * MLP_OUT * 4));
    HIPCHK(hipMemcpyAsync(dX, X, (size_t)rows * MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    MlpFwdArgs f;
    f.X = dX; f.idx = nullptr; f.rows = rows; f.in_mean = m->in_mean; f.in_istd = m->in_istd; f.w = mlp_w(m);
    f.Y = dY; f.sX0 = nullptr; f.sZ = nullptr; f.sH = nullptr;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((rows + MLP_TILE - 1) / MLP_TILE), dim3(256), 0, m->stream, f);
    HIPCHK(hipMemcpyAsync(Y, dY, (size_t)rows * MLP_OUT * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    (void)hipFree(dX);
    (void)hipFree(dY);
    return YC_OK;
}

}  // extern "C"

namespace {

// Staging of yc_mlp_forward_noalloc: the features of up to MLP_NOALLOC_ROWS rows, then
// their outputs, in one file-scope device array.  Each device holds its own copy for the
// life of the process; host code reaches a device's copy only through the address
// hipGetSymbolAddress returns on that device (noalloc_scratch_dev), never through the
// array's name.
constexpr int MLP_NOALLOC_ROWS = 1 << 16;
constexpr size_t MLP_NOALLOC_FLOATS = (size_t)MLP_NOALLOC_ROWS * (MLP_IN + MLP_OUT);   // 9.7 MB

}  // namespace

static __device__ __align__(256) float scratch[MLP_NOALLOC_FLOATS];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_dev[MAX_DEVICES];

// The current device's copy of `scratch`, resolved once per device.
int noalloc_scratch_dev(float** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(YC_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (float*)p;
    }
    *out = g_scratch_dev[dev];
    return YC_OK;
}

}  // namespace

extern "C" {

int yc_mlp_forward_noalloc(yc_mlp* m, const float* X, int rows, float* Y) {
    if (!m || !X || !Y) return fail(YC_ERR_ARG, "null argument");
    if (rows <= 0 || rows > MLP_NOALLOC_ROWS)
        return fail(YC_ERR_ARG, "yc_mlp_forward_noalloc: rows must be 1.." + std::to_string(MLP_NOALLOC_ROWS));
    if (!m->scratch) return fail(YC_ERR_ARG, "yc_mlp_forward_noalloc: the network has no staging (created on another device?)");
    float* dX = m->scratch;
    float* dY = m->scratch + (size_t)MLP_NOALLOC_ROWS * MLP_IN;
    HIPCHK(hipMemcpyAsync(dX, X, (size_t)rows * MLP_IN * 4, hipMemcpyHostToDevice, m->stream));
    MlpFwdArgs f;
    f.X = dX; f.idx = nullptr; f.rows = rows; f.in_mean = m->in_mean; f.in_istd = m->in_istd; f.w = mlp_w(m);
    f.Y = dY; f.sX0 = nullptr; f.sZ = nullptr; f.sH = nullptr;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((rows + MLP_TILE - 1) / MLP_TILE), dim3(256), 0, m->stream, f);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(Y, dY, (size_t)rows * MLP_OUT * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return YC_OK;
}

int yc_mlp_grad(yc_mlp* m, const int* idx, int rows, float* grad, float* loss) {
    if (!m || !idx || rows <= 0 || !m->X) return fail(YC_ERR_ARG, "bad argument (dataset set?)");
    for (int i = 0; i < rows; ++i) if (idx[i] < 0 || idx[i] >= m->nrows) return fail(YC_ERR_ARG, "row index out of range");
    int rc = mlp_reserve_batch(m, rows);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->idx, idx, (size_t)rows * 4, hipMemcpyHostToDevice, m->stream));
    rc = mlp_batch(m, m->idx, rows, 0.0f);
    if (rc) return rc;
    std::vector<float> g(MLP_NPARAM + 1);
    HIPCHK(hipMemcpyAsync(g.data(), m->grad, (MLP_NPARAM + 1) * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (grad) std::memcpy(grad, g.data(), MLP_NPARAM * 4);
    if (loss) *loss = g[MLP_NPARAM];
    return YC_OK;
}

int yc_mlp_train(yc_mlp* m, const int* idx, int rows, int batch, float lr, float* mean_loss) {
    if (!m || !idx || rows <= 0 || batch <= 0 || !m->X) return fail(YC_ERR_ARG, "bad argument (dataset set?)");
    for (int i = 0; i < rows; ++i) if (idx[i] < 0 || idx[i] >= m->nrows) return fail(YC_ERR_ARG, "row index out of range");
    int rc = mlp_reserve_batch(m, std::max(batch, 1));
    if (rc) return rc;
    int* didx = nullptr;
    HIPCHK(hipMalloc((void**)&didx, (size_t)rows * 4));
    HIPCHK(hipMemcpyAsync(didx, idx, (size_t)rows * 4, hipMemcpyHostToDevice, m->stream));
    double lsum = 0.0;
    int nb = 0;
    for (int s = 0; s + batch <= rows; s += batch) {
        rc = mlp_batch(m, didx + s, batch, lr);
        if (rc) break;
        float l = 0.0f;
        HIPCHK(hipMemcpyAsync(&l, m->grad + MLP_NPARAM, 4, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        lsum += l;
        ++nb;
    }
    (void)hipFree(didx);
    if (mean_loss) *mean_loss = nb ? (float)(lsum / nb) : 0.0f;
    return rc;
}

int yc_mlp_time_forward(yc_mlp* m, int rows, int reps, double* avg_ms) {
    if (!m || rows <= 0 || reps <= 0 || !avg_ms) return fail(YC_ERR_ARG, "bad argument");
    float *dX = nullptr, *dY = nullptr;
    HIPCHK(hipMalloc((void**)&dX, (size_t)rows * MLP_IN * 4));
    HIPCHK(hipMalloc((void**)&dY, (size_t)rows * MLP_OUT * 4));
    HIPCHK(hipMemsetAsync(dX, 0, (size_t)rows * MLP_IN * 4, m->stream));
    MlpFwdArgs f;
    f.X = dX; f.idx = nullptr; f.rows = rows; f.in_mean = m->in_mean; f.in_istd = m->in_istd; f.w = mlp_w(m);
    f.Y = dY; f.sX0 = nullptr; f.sZ = nullptr; f.sH = nullptr;
    int nblk = (rows + MLP_TILE - 1) / MLP_TILE;
    int st = time_launches(m->stream, reps, avg_ms, [&] {
        hipLaunchKernelGGL(k_mlp_fwd, dim3(nblk), dim3(256), 0, m->stream, f);
    });
    (void)hipFree(dX);
    (void)hipFree(dY);
    return st;
}

int yc_time_warm_guess(yc_ctx* c, yc_mlp* m, int reps, double* avg_ms, double* avg_mlp_ms) {
    if (!c || !m || reps <= 0 || !avg_ms || !avg_mlp_ms) return fail(YC_ERR_ARG, "bad argument");
    if (!c->assembled) return fail(YC_ERR_ARG, "warm guess before assemble");
    int st = time_launches(c->stream, reps, avg_ms, [&] { (void)warm_guess(c, m); });
    if (st) return st;
    return time_launches(c->stream, reps, avg_mlp_ms, [&] {
        hipLaunchKernelGGL(k_mlp_warm, dim3((c->N + MLP_TILE - 1) / MLP_TILE), dim3(256), 0, c->stream, feat_args(c), mlp_w(m),
                           (const float*)m->in_mean, (const float*)m->in_istd, (const float*)m->out_std, c->cx);
    });
}

int yc_features(yc_ctx* c, float* out) {
    if (!c || !out) return fail(YC_ERR_ARG, "null argument");
    if (!c->assembled) return fail(YC_ERR_ARG, "features before assemble");
    float* F = nullptr;
    HIPCHK(hipMalloc((void**)&F, (size_t)c->N * MLP_IN * 4));
    int rc = extrapolate_and_apply(c, c->cy);   // x_q in the CG work vector
    if (!rc) {
        hipLaunchKernelGGL(k_features, dim3(grid_for(c->N)), dim3(BLK), 0, c->stream, feat_args(c), F);
        if (hipGetLastError() != hipSuccess) rc = fail(YC_ERR_HIP, "k_features launch");
    }
    if (!rc && (hipMemcpyAsync(out, F, (size_t)c->N * MLP_IN * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail(YC_ERR_HIP, "features download");
    (void)hipFree(F);
    return rc;
}

int yc_warm_guess(yc_ctx* c, yc_mlp* m) {
    if (!c || !m) return fail(YC_ERR_ARG, "null argument");
    if (!c->assembled) return fail(YC_ERR_ARG, "warm guess before assemble");
    int rc = warm_guess(c, m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return YC_OK;
}

int yc_targets(yc_ctx* c, float* out) {
    if (!c || !out) return fail(YC_ERR_ARG, "null argument");
    if (!c->solved) return fail(YC_ERR_ARG, "targets before solve");
    size_t PV = (size_t)5 * c->P;
    std::vector<double> x(PV), v(PV), d(PV), d2(PV);
    HIPCHK(hipMemcpyAsync(x.data(), c->cx, PV * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(v.data(), c->vpad, PV * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(d.data(), c->dprev, PV * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(d2.data(), c->dprev2, PV * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // x - x_q with x_q = v + (2 dprev - dprev2) as k_guess_x0 forms it
    for (int vv = 0; vv < c->Lg; ++vv)
        for (int u = 0; u < c->W; ++u) {
            int k = u + vv * c->W, pk = (u + 2) + (vv + 2) * c->Wp;
            for (int q = 0; q < 5; ++q) {
                size_t i = (size_t)q * c->P + pk;
                out[(size_t)k * 5 + q] = (float)(x[i] - (v[i] + (2.0 * d[i] - d2[i])));
            }
        }
    return YC_OK;
}

int yc_step_warm(yc_ctx* c, yc_mlp* m, double h, double tol, int maxit, yc_step_info* info) {
    if (!c || !m) return fail(YC_ERR_ARG, "null argument");
    g_err.clear();
    int it = 0;
    double rr = 0;
    int rc = do_assemble(c, h, false, true);
    if (!rc) rc = warm_guess(c, m);
    if (!rc) rc = do_solve(c, nullptr, tol, maxit, &it, &rr, true);
    if (info) { info->cg_iters = it; info->cg_relres = rr; info->status = rc; }
    if (rc == YC_ERR_NONFINITE || rc == YC_ERR_HIP || rc == YC_ERR_ARG) return rc;
    int rc2 = do_update(c, h);
    if (rc2) return rc2;
    HIPCHK(hipStreamSynchronize(c->stream));
    return rc;
}

int yc_run_warm(yc_ctx* c, yc_mlp* m, double h, int steps, double tol, int maxit, long long* total_iters) {
    if (!c || !m) return fail(YC_ERR_ARG, "null argument");
    if (c->cg_variant == YC_CG_PERSISTENT && c->coop_ok) return run_async(c, m, h, steps, tol, maxit, total_iters);
    long long tot = 0;
    int worst = YC_OK;
    for (int s = 0; s < steps; ++s) {
        yc_step_info info;
        int rc = yc_step_warm(c, m, h, tol, maxit, &info);
        tot += info.cg_iters;
        if (rc == YC_ERR_NOT_CONVERGED) worst = rc;
        else if (rc) { if (total_iters) *total_iters = tot; return rc; }
    }
    if (total_iters) *total_iters = tot;
    return worst;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reverse mode: a recorded rollout and its backward pass (adjoint.hpp)
// ---------------------------------------------------------------------------
namespace {

void adjoint_free(yc_ctx* c) {
    double** bufs[] = {&c->adj_ckpt, &c->adj_final, &c->adj_seg, &c->adj_ext, &c->adj_g0, &c->adj_g1, &c->adj_gpar,
                       &c->adj_gcomp, &c->adj_ybar, &c->adj_pbar, &c->adj_gslide, &c->adj_fric_in, &c->adj_lam, &c->adj_xs,
                       &c->adj_gst, &c->adj_lpart, &c->adj_sums, &c->adj_tx, &c->adj_tu};
    for (double** p : bufs) { if (*p) (void)hipFree(*p); *p = nullptr; }
    c->adj_max = c->adj_K = c->adj_steps = 0;
    c->adj_has_ext.clear();
    c->fric_out = nullptr;
}

// The descriptor of the adjoint kernels: the state's tangent pointers are seed codes
// (tan_at), the parameters' tangents zero unless a parameter pass sets them.
ClothDev make_adj_dev(yc_ctx* c, int mode) {
    ClothDev C = make_dev(c);
    C.dx = seed_ptr(SEED_X); C.dv = seed_ptr(SEED_V); C.dxfix = seed_ptr(SEED_XFIX); C.duv = seed_ptr(SEED_UV);
    C.duvd = seed_ptr(SEED_UVD); C.duvbar = seed_ptr(SEED_UVBAR); C.dnrm = seed_ptr(SEED_NRM);
    C.ref_mode = mode == YC_GRAD_REF;
    return C;
}

size_t adj_block(const yc_ctx* c) { return (size_t)G_PLANES * c->N; }

StateDev block_state(yc_ctx* c, double* blk) {
    StateDev s = make_state(c);
    const size_t N = c->N;
    s.x = blk + G_X * N; s.v = blk + G_V * N; s.xfix = blk + G_XFIX * N; s.uv = blk + G_UV * N;
    s.uvd = blk + G_UVD * N; s.uvbar = blk + G_UVBAR * N; s.nrm = blk + G_NRM * N;
    return s;
}

int ckpt_io(yc_ctx* c, double* blk, int save) {
    hipLaunchKernelGGL(k_ckpt, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, make_state(c), blk, save);
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// One recorded step, as yc_step with the zero CG guess (a recomputed step repeats it bit
// for bit); a step whose solve fails or stops above tolerance is not taken, as in
// yc_step_tangent.
int adj_forward_step(yc_ctx* c, double h, double tol, int maxit, int* iters, double* relres) {
    int rc = do_assemble(c, h);
    if (!rc) rc = do_solve(c, nullptr, tol, maxit, iters, relres);
    if (rc) return rc;
    return do_update(c, h);
}

template <bool REF>
int adj_launch(yc_ctx* c, const ClothDev& Cadj, double h) {
    const int nT = c->nT;
    AdjW w;
    w.lam = c->adj_lam; w.xs = c->adj_xs; w.ybar = c->adj_ybar; w.Wp = c->Wp; w.P = c->P; w.h = h;
    AdjOut o;
    o.g = c->adj_g1; o.gpar = c->adj_gpar; o.gcomp = c->adj_gcomp; o.ybar = c->adj_ybar;
    ClothDev Cs = Cadj;
    Cs.S = 1.0;   // k_adj_shear_par: the shear at S = 1 is its S derivative
    if (c->Nc > 0) {
        hipLaunchKernelGGL(k_adj_shear<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cadj, w, o, (const double*)c->comp);
        hipLaunchKernelGGL(k_adj_shear_par<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cs, w, o, (const double*)c->comp);
    } else {
        HIPCHK(hipMemsetAsync(c->adj_gcomp, 0, (size_t)c->N * sizeof(double), c->stream));
    }
    hipLaunchKernelGGL(k_adj_fric<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cadj, w, o, (const double*)c->adj_fric_in,
                       (const double*)c->adj_gslide, -1);
    for (int p = 0; p < 3; ++p) {
        ClothDev Cp = Cadj;
        if (p == 0) Cp.dmu = 1.0;
        else if (p == 1) Cp.dkf = 1.0;
        else Cp.ddf = 1.0;
        hipLaunchKernelGGL(k_adj_fric<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cp, w, o, (const double*)c->adj_fric_in,
                           (const double*)c->adj_gslide, p);
    }
    hipLaunchKernelGGL(k_adj_elems<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cadj, w, o);
    for (int p = 0; p < 3; ++p) {
        ClothDev Cp = Cadj;
        if (p == 0) { Cp.drho[0] = Cp.drho[1] = 1.0; Cp.dkb[0] = Cp.dkb[1] = 1.0; Cp.dhandle_k = 1.0; }
        else if (p == 1) { Cp.dks[0] = Cp.dks[1] = 1.0; }
        else Cp.dkc = 1.0;
        hipLaunchKernelGGL(k_adj_elems_par<REF>, dim3(nT), dim3(TILE_N), 0, c->stream, Cp, w, o, p);
    }
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// The loss gradient of the state after recorded step t (1-based), into adj_g1, from the
// caller's array (host) or the targets.
int adj_add_loss(yc_ctx* c, int t, const double* gstate, double* state_blk) {
    const size_t N = c->N;
    if (gstate) {
        HIPCHK(hipMemcpyAsync(c->adj_gst, gstate + (size_t)(t - 1) * 15 * N, 15 * N * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        hipLaunchKernelGGL(k_adj_add_gstate, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, c->N,
                           (const double*)c->adj_gst, c->adj_g1);
    } else if (c->adj_tx) {
        hipLaunchKernelGGL(k_adj_target_loss, dim3((c->N + NBLK - 1) / NBLK), dim3(NBLK), 0, c->stream, c->N,
                           (const double*)(state_blk + G_X * N), (const double*)(state_blk + G_UV * N),
                           (const double*)(c->adj_tx + (size_t)(t - 1) * 3 * N), (const double*)(c->adj_tu + (size_t)(t - 1) * 2 * N),
                           c->adj_tw, c->adj_g1, c->adj_lpart);
    }
    HIPCHK(hipGetLastError());
    return YC_OK;
}

// Copies the state adjoint block g (G_PLANES planes) to the yc_get_state layout.
void adj_to_state_layout(const yc_ctx* c, const std::vector<double>& g, double* out) {
    const size_t N = c->N;
    for (size_t k = 0; k < N; ++k) {
        for (int q = 0; q < 3; ++q) {
            out[3 * k + q] = g[(G_X + q) * N + k];
            out[3 * N + 3 * k + q] = g[(G_XFIX + q) * N + k];
            out[6 * N + 3 * k + q] = g[(G_V + q) * N + k];
        }
        for (int q = 0; q < 2; ++q) {
            out[9 * N + 2 * k + q] = g[(G_UV + q) * N + k];
            out[11 * N + 2 * k + q] = g[(G_UVBAR + q) * N + k];
            out[13 * N + 2 * k + q] = g[(G_UVD + q) * N + k];
        }
    }
}

template <bool REF>
int adjoint_backward(yc_ctx* c, double h, double tol, int maxit, const double* gstate, double* dparam, double* dstate0,
                     double* dext, double* loss, long long* cg_iters) {
    const int T = c->adj_steps, K = c->adj_K;
    const size_t N = c->N, B = adj_block(c), PV = (size_t)5 * c->P;
    const int mode = REF ? YC_GRAD_REF : YC_GRAD_EXACT;
    const ClothDev Cadj = make_adj_dev(c, mode);
    const int nblk_n = (c->N + NBLK - 1) / NBLK;
    long long iters_total = 0;
    int rc = ckpt_io(c, c->adj_final, 1);   // the state after the last step, restored at the end
    if (rc) return rc;
    // with the persistent CG the whole pass is queued without a host round trip: the
    // assemblies' flags and the solves' outcomes accumulate on the device and are read once
    const bool queued = c->coop_ok != 0;
    if (queued) {
        HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
        HIPCHK(hipMemsetAsync((char*)c->sc + offsetof(CGScal, total_iters), 0, sizeof(long long) + 2 * sizeof(int), c->stream));
    }
    double* lamlog = nullptr;   // every step's lambda, for dext (device, converted at the end)
    if (dext) {
        rc = dalloc(&lamlog, (size_t)T * PV);
        if (rc) return rc;
    }
    HIPCHK(hipMemsetAsync(c->adj_g1, 0, B * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->adj_gpar, 0, (size_t)YC_NPARAM * N * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->adj_lpart, 0, N * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->adj_lam, 0, PV * sizeof(double), c->stream));
    rc = adj_add_loss(c, T, gstate, c->adj_final);
    if (rc) return rc;
    std::swap(c->adj_g0, c->adj_g1);   // g0: the adjoint of the state after the step being reversed
    std::vector<double> hl(PV);
    std::vector<double> glob(c->n);
    c->fric_out = c->adj_fric_in;
    const bool had_ext = c->has_ext;
    double* ext_save = c->adj_ext ? c->adj_ext + (size_t)c->adj_max * c->n : nullptr;
    if (ext_save && had_ext)
        HIPCHK(hipMemcpyAsync(ext_save, c->ext, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    auto state_ptr = [&](int n, int seg0) -> double* {
        if (n == T) return c->adj_final;
        if (n % K == 0) return c->adj_ckpt + (size_t)(n / K) * B;
        return c->adj_seg + (size_t)(n - seg0) * B;
    };
    for (int q = (T - 1) / K; q >= 0 && !rc; --q) {
        const int s0 = q * K, s1 = std::min(s0 + K, T);
        // recompute the segment's states after its checkpoint
        if (s1 - s0 > 1) {
            rc = ckpt_io(c, c->adj_ckpt + (size_t)q * B, 0);
            for (int n = s0 + 1; n < s1 && !rc; ++n) {
                if (c->adj_ext && c->adj_has_ext[n - 1]) {
                    HIPCHK(hipMemcpyAsync(c->ext, c->adj_ext + (size_t)(n - 1) * c->n, (size_t)c->n * sizeof(double),
                                          hipMemcpyDeviceToDevice, c->stream));
                    c->has_ext = true;
                } else if (c->adj_ext) {
                    c->has_ext = false;
                }
                rc = adj_forward_step(c, h, tol, maxit, nullptr, nullptr);
                if (!rc) rc = ckpt_io(c, c->adj_seg + (size_t)(n - s0) * B, 1);
            }
        }
        for (int n = s1 - 1; n >= s0 && !rc; --n) {
            double* sn = state_ptr(n, s0);
            double* sn1 = state_ptr(n + 1, s0);
            // 1. normals' adjoint at the state after the step
            hipLaunchKernelGGL(k_adj_nrm, dim3(nblk_n), dim3(NBLK), 0, c->stream, block_state(c, sn1),
                               (const double*)(c->adj_g0 + G_NRM * N), c->adj_pbar);
            // 2. recompute the step's system and friction inputs
            rc = ckpt_io(c, sn, 0);
            if (!rc) rc = do_assemble(c, h, queued);
            if (rc) break;
            hipLaunchKernelGGL(k_pack_vel, dim3(grid_for(c->N)), dim3(BLK), 0, c->stream, (const double*)(sn1 + G_V * N),
                               (const double*)(sn1 + G_UVD * N), c->W, c->Lg, c->N, c->Wp, c->P, c->adj_xs);
            // 3. the update's adjoint: the right-hand side of the backward solve in b
            hipLaunchKernelGGL(k_adj_update, dim3(nblk_n), dim3(NBLK), 0, c->stream, (const unsigned char*)c->flags, c->W, c->Lg,
                               (const double*)c->adj_g0, (const double*)c->adj_pbar, (const double*)c->is_slide,
                               (const double*)c->adj_xs, c->Wp, c->P, h, c->b, c->adj_gslide, c->adj_g1);
            // 4. the backward solve, from the previous step's lambda
            const double* Au = c->A;
            if (!REF) {
                if (c->bj_of == c->At) c->bj_of = nullptr;   // At is rewritten
                hipLaunchKernelGGL(k_transpose, dim3(c->nT), dim3(NBLK), 0, c->stream, (const double*)c->A, c->At, c->W, c->Lg, c->ntx);
                Au = c->At;
            }
            HIPCHK(hipMemcpyAsync(c->cx, c->adj_lam, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_spmv, dim3(c->part.nS), dim3(BLK), 0, c->stream, make_spmv(c, c->cx, c->cy, Au), (double*)nullptr,
                               (const CGScal*)c->sc, 0);
            if (queued) {
                hipLaunchKernelGGL(k_cg_init, dim3(c->part.nV), dim3(BLK), 0, c->stream, make_vec(c), c->part, 1);
                rc = launch_cgp(c, tol, maxit, Au);
            } else {
                int it = 0;
                rc = solve_on(c, Au, tol, maxit, &it, true);
                iters_total += it;
            }
            if (rc) break;
            HIPCHK(hipMemcpyAsync(c->adj_lam, c->cx, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            if (lamlog)   // the adjoint of the step's external force is h lambda
                HIPCHK(hipMemcpyAsync(lamlog + (size_t)n * PV, c->cx, PV * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            // 5. the assembly's vector-Jacobian product, into g1 and the parameter partials
            rc = adj_launch<REF>(c, Cadj, h);
            if (rc) break;
            // 6. the loss gradient of the state before the step (the states after steps 1..T)
            if (n >= 1) rc = adj_add_loss(c, n, gstate, sn);
            if (rc) break;
            std::swap(c->adj_g0, c->adj_g1);
        }
    }
    c->fric_out = nullptr;
    if (!rc && queued) {   // the outcomes of the queued assemblies and solves
        HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_flag, c->flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        iters_total = c->h_sc->total_iters;
        if (c->h_flag[2]) {
            (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
            rc = fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
        } else if (c->h_flag[0]) {
            rc = fail(YC_ERR_HIP, "assembly: stencil slot out of range");
        } else if (c->h_flag[1] || c->h_sc->nonfinite) {
            rc = fail(YC_ERR_NONFINITE, "adjoint: NaN/Inf in a recomputed system or a backward solve");
        } else if (c->h_sc->nonconv) {
            rc = fail(YC_ERR_NOT_CONVERGED, "adjoint: a backward solve stopped at maxit above tolerance");
        }
    }
    if (!rc && dext) {
        for (int n = 0; n < T && !rc; ++n) {
            if (hipMemcpyAsync(hl.data(), lamlog + (size_t)n * PV, PV * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(YC_ERR_HIP, "dext download");
                break;
            }
            from_padded(c, hl, glob.data());
            for (int i = 0; i < c->n; ++i) dext[(size_t)n * c->n + i] = h * glob[i];
        }
    }
    if (lamlog) { (void)hipStreamSynchronize(c->stream); (void)hipFree(lamlog); }
    if (!rc) {   // the initial state's normals come from its x (yc_set_state)
        hipLaunchKernelGGL(k_adj_nrm, dim3(nblk_n), dim3(NBLK), 0, c->stream, block_state(c, c->adj_ckpt),
                           (const double*)(c->adj_g0 + G_NRM * N), c->adj_pbar);
        hipLaunchKernelGGL(k_adj_fold_nrm, dim3(nblk_n), dim3(NBLK), 0, c->stream, c->W, c->Lg, (const double*)c->adj_pbar, c->adj_g0);
        hipLaunchKernelGGL(k_plane_sums, dim3(YC_NPARAM), dim3(256), 0, c->stream, (const double*)c->adj_gpar, c->N, c->adj_sums);
        hipLaunchKernelGGL(k_plane_sums, dim3(1), dim3(256), 0, c->stream, (const double*)c->adj_lpart, c->N, c->adj_sums + YC_NPARAM);
        HIPCHK(hipGetLastError());
        double sums[YC_NPARAM + 1];
        HIPCHK(hipMemcpyAsync(sums, c->adj_sums, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
        std::vector<double> g0;
        if (dstate0) {
            g0.resize(B);
            HIPCHK(hipMemcpyAsync(g0.data(), c->adj_g0, B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        // ks = Y pi R^2 and kb = B pi R^2 (src/Yarn.h:14-15); a one-yarn cloth's second
        // yarn slot is its first yarn
        const double area = PI_REF * std::pow(c->d.R, 2);
        sums[YC_P_Y0] *= area; sums[YC_P_Y1] *= area; sums[YC_P_B0] *= area; sums[YC_P_B1] *= area;
        if (c->d.n_yarns == 1) {
            sums[YC_P_RHO0] += sums[YC_P_RHO1]; sums[YC_P_Y0] += sums[YC_P_Y1]; sums[YC_P_B0] += sums[YC_P_B1];
            sums[YC_P_RHO1] = sums[YC_P_Y1] = sums[YC_P_B1] = 0.0;
        }
        if (dparam) for (int i = 0; i < YC_NPARAM; ++i) dparam[i] = sums[i];
        if (loss) *loss = sums[YC_NPARAM];
        if (dstate0) adj_to_state_layout(c, g0, dstate0);
    }
    if (cg_iters) *cg_iters = iters_total;
    // back to the state after the last recorded step
    int rc2 = ckpt_io(c, c->adj_final, 0);
    if (ext_save) {   // the caller's external force as it was
        c->has_ext = had_ext;
        if (had_ext && hipMemcpyAsync(c->ext, ext_save, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            rc2 = fail(YC_ERR_HIP, "restoring the external force");
    }
    if (!rc) rc = rc2;
    (void)hipStreamSynchronize(c->stream);
    c->assembled = false;
    c->solved = false;
    return rc;
}

}  // namespace

extern "C" {

int yc_adjoint_end(yc_ctx* c) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    (void)hipStreamSynchronize(c->stream);
    adjoint_free(c);
    return YC_OK;
}

int yc_adjoint_begin(yc_ctx* c, int max_steps, int checkpoint_every) {
    if (!c || max_steps <= 0 || checkpoint_every <= 0) return fail(YC_ERR_ARG, "bad argument");
    g_err.clear();
    yc_adjoint_end(c);
    const size_t N = c->N, B = adj_block(c), PV = (size_t)5 * c->P;
    const int K = std::min(checkpoint_every, max_steps);
    int rc = YC_OK;
    auto chk = [&](int r) { if (r && !rc) rc = r; };
    chk(dalloc(&c->adj_ckpt, ((size_t)max_steps / K + 1) * B));
    chk(dalloc(&c->adj_final, B));
    if (K > 1) chk(dalloc(&c->adj_seg, (size_t)K * B));
    chk(dalloc(&c->adj_g0, B)); chk(dalloc(&c->adj_g1, B));
    chk(dalloc(&c->adj_gpar, (size_t)YC_NPARAM * N)); chk(dalloc(&c->adj_gcomp, N)); chk(dalloc(&c->adj_ybar, (size_t)FRIC_IN * N));
    chk(dalloc(&c->adj_pbar, 15 * N)); chk(dalloc(&c->adj_gslide, 2 * N)); chk(dalloc(&c->adj_fric_in, (size_t)FRIC_IN * N));
    chk(dalloc(&c->adj_lam, PV)); chk(dalloc(&c->adj_xs, PV));
    chk(dalloc(&c->adj_gst, 15 * N)); chk(dalloc(&c->adj_lpart, N)); chk(dalloc(&c->adj_sums, YC_NPARAM + 1));
    if (rc) { adjoint_free(c); return rc; }
    c->adj_max = max_steps;
    c->adj_K = K;
    c->adj_steps = 0;
    rc = ckpt_io(c, c->adj_ckpt, 1);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(YC_ERR_HIP, "sync");
    if (rc) adjoint_free(c);
    return rc;
}

int yc_adjoint_step(yc_ctx* c, double h, double tol, int maxit, yc_step_info* info) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->adj_max <= 0) return fail(YC_ERR_ARG, "yc_adjoint_step before yc_adjoint_begin");
    if (c->adj_steps >= c->adj_max) return fail(YC_ERR_ARG, "more steps than yc_adjoint_begin allowed");
    g_err.clear();
    const int t = c->adj_steps;
    if (c->adj_K > 1) {   // a recomputed segment needs the step's external force
        if (c->has_ext && !c->adj_ext) {   // one slot per step and one for the caller's force during a backward pass
            int r = dalloc(&c->adj_ext, (size_t)(c->adj_max + 1) * c->n);
            if (r) return r;
        }
        c->adj_has_ext.resize(c->adj_max, 0);
        c->adj_has_ext[t] = c->has_ext ? 1 : 0;
        if (c->has_ext)
            HIPCHK(hipMemcpyAsync(c->adj_ext + (size_t)t * c->n, c->ext, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    int it = 0;
    double rr = 0;
    int rc = adj_forward_step(c, h, tol, maxit, &it, &rr);
    if (info) { info->cg_iters = it; info->cg_relres = rr; info->status = rc; }
    if (rc) return rc;   // not taken, not recorded
    c->adj_steps = t + 1;
    if (c->adj_steps % c->adj_K == 0 && c->adj_steps < c->adj_max) {
        int r = ckpt_io(c, c->adj_ckpt + (size_t)(c->adj_steps / c->adj_K) * adj_block(c), 1);
        if (r) return r;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return rc;
}

int yc_adjoint_run(yc_ctx* c, double h, int steps, double tol, int maxit, long long* total_iters) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->adj_max <= 0) return fail(YC_ERR_ARG, "yc_adjoint_run before yc_adjoint_begin");
    if (steps < 0 || c->adj_steps + steps > c->adj_max) return fail(YC_ERR_ARG, "more steps than yc_adjoint_begin allowed");
    if (!(c->cg_variant == YC_CG_PERSISTENT && c->coop_ok)) {   // step by step, each checked
        long long tot = 0;
        for (int s = 0; s < steps; ++s) {
            yc_step_info info;
            int rc = yc_adjoint_step(c, h, tol, maxit, &info);
            tot += info.cg_iters;
            if (rc) { if (total_iters) *total_iters = tot; return rc; }
        }
        if (total_iters) *total_iters = tot;
        return YC_OK;
    }
    g_err.clear();
    if (c->adj_K > 1) {
        if (c->has_ext && !c->adj_ext) {
            int r = dalloc(&c->adj_ext, (size_t)(c->adj_max + 1) * c->n);
            if (r) return r;
        }
        c->adj_has_ext.resize(c->adj_max, 0);
    }
    HIPCHK(hipMemsetAsync(c->flag, 0, 2 * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync((char*)c->sc + offsetof(CGScal, total_iters), 0, sizeof(long long) + 2 * sizeof(int), c->stream));
    for (int s = 0; s < steps; ++s) {
        const int t = c->adj_steps + s;
        if (c->adj_K > 1) {
            c->adj_has_ext[t] = c->has_ext ? 1 : 0;
            if (c->has_ext)
                HIPCHK(hipMemcpyAsync(c->adj_ext + (size_t)t * c->n, c->ext, (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice,
                                      c->stream));
        }
        int rc = do_assemble(c, h, true);
        if (!rc) rc = do_solve(c, nullptr, tol, maxit, nullptr, nullptr, false, true);
        if (!rc) rc = do_update(c, h);
        if (!rc && (t + 1) % c->adj_K == 0 && t + 1 < c->adj_max) rc = ckpt_io(c, c->adj_ckpt + (size_t)((t + 1) / c->adj_K) * adj_block(c), 1);
        if (rc) { c->adj_steps = 0; return rc; }
    }
    HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, sizeof(CGScal), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_flag, c->flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_flag + 2, c->cg_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (total_iters) *total_iters = c->h_sc->total_iters;
    int rc = YC_OK;
    if (c->h_flag[2]) {
        (void)hipMemsetAsync(c->cg_err, 0, sizeof(int), c->stream);
        rc = fail(YC_ERR_HIP, "persistent CG: a block waited past its spin limit");
    } else if (c->h_flag[0]) {
        rc = fail(YC_ERR_HIP, "assembly: stencil slot out of range");
    } else if (c->h_flag[1] || c->h_sc->nonfinite) {
        rc = fail(YC_ERR_NONFINITE, "NaN/Inf in A, b or a solution");
    } else if (c->h_sc->nonconv) {
        rc = fail(YC_ERR_NOT_CONVERGED, "a CG solve stopped at maxit above tolerance");
    }
    if (rc) { c->adj_steps = 0; return rc; }   // the recording is void: steps after a failed one ran on its result
    c->adj_steps += steps;
    return YC_OK;
}

int yc_adjoint_set_targets(yc_ctx* c, const double* lpos, const double* epos, double weight) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->adj_max <= 0) return fail(YC_ERR_ARG, "yc_adjoint_set_targets before yc_adjoint_begin");
    if (c->adj_tx) { (void)hipStreamSynchronize(c->stream); (void)hipFree(c->adj_tx); (void)hipFree(c->adj_tu); }
    c->adj_tx = c->adj_tu = nullptr;
    if (!lpos && !epos) return YC_OK;
    if (!lpos || !epos) return fail(YC_ERR_ARG, "targets need both positions and Eulerian coordinates");
    const size_t N = c->N, T = c->adj_max;
    int rc = dalloc(&c->adj_tx, T * 3 * N);
    if (!rc) rc = dalloc(&c->adj_tu, T * 2 * N);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->adj_tx, lpos, T * 3 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->adj_tu, epos, T * 2 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->adj_tw = weight;
    return YC_OK;
}

int yc_adjoint_backward(yc_ctx* c, int mode, double h, double tol, int maxit, const double* gstate, double* dparam,
                        double* dstate0, double* dext, double* loss, long long* cg_iters) {
    if (!c) return fail(YC_ERR_ARG, "null context");
    if (c->adj_max <= 0) return fail(YC_ERR_ARG, "yc_adjoint_backward before yc_adjoint_begin");
    if (c->adj_steps <= 0) return fail(YC_ERR_ARG, "no recorded step");
    if (mode != YC_GRAD_REF && mode != YC_GRAD_EXACT) return fail(YC_ERR_ARG, "unknown gradient mode");
    if (c->precond != YC_PRECOND_JACOBI && !c->coop_ok) return fail(YC_ERR_ARG, "the block-Jacobi preconditioner needs the persistent CG");
    g_err.clear();
    if (mode == YC_GRAD_EXACT && !c->At) {   // A^T of the exact adjoint's solves (shared with the tangents)
        int r = dalloc(&c->At, c->a_size);
        if (r) return r;
    }
    return mode == YC_GRAD_REF ? adjoint_backward<true>(c, h, tol, maxit, gstate, dparam, dstate0, dext, loss, cg_iters)
                               : adjoint_backward<false>(c, h, tol, maxit, gstate, dparam, dstate0, dext, loss, cg_iters);
}

}  // extern "C"

#include "batch.hpp"
#include "batch_adjoint.hpp"

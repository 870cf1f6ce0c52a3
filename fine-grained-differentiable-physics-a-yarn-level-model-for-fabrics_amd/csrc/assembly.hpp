// assembly.hpp -- declarations shared by the assembly kernels of yarnstep.hip and the
// host build of their body (tests/asm_host.cpp): the outputs, the node-to-thread map of a
// tile, the LDS sizes, the weave's flip pattern and the cloth's scalars in the kernels'
// form.  Included after elements.hpp.
#pragma once
#include <cmath>
#include <cstddef>

namespace {
using namespace ys;

struct AsmOut {
    double* A;          // tiles of NSLOT planes of TILE_N (a_base)
    double* F;          // 5 planes of N
    double* b;          // padded 5-plane vector
    double* is_slide;   // 2 planes of N
    double* comp;       // N
    double* fepart;     // 2 planes of N
    double* invdiag;    // padded 5-plane vector
    const double* ext;  // global-layout n vector or null
    const double* dext; // tangent run: the direction's tangent of ext (global layout) or null
    int* flag;          // [0] structural error, [1] non-finite
    int Wp, P;          // padded width, plane stride
    // tangent run: tangents of A (exact mode, else null), F, b, is_slide and comp, and
    // the comp tangent read by the shear; detach = the reference's gradient
    double *dA, *dF, *db, *dis_slide, *dcomp;
    int detach;
    // value run: element-major shear and bend records (k_shear_elem, k_bend_elem)
    // replayed by the node kernels
    const double* shear_rec;
    int nseg;
    const double* bend_rec;
    // value run: -1 for A; 0, 1, 2 for M, K_x or K_v alone (RowSinkT::select_kind)
    int only_kind;
    // value run, when set: the friction's accumulated inputs of every node (FRIC_IN planes
    // of N, zero off the interior), which the adjoint reads back
    double* fric_in;
    // value run, when set: the step's starting velocity in the padded 5-plane layout (x, y,
    // z, u, v; zero u, v off the interior), written by the second kernel
    double* vpad;
};

// The friction's inputs accumulated over a node's elements, as planes of fric_in and of
// the adjoint weights: the compression sums SBC_L (3) and SBC_LE (2 x 3), the E force
// before friction (2), its thread partial (2), the K_x and K_v E-E self entries (2 + 2).
enum FricIn { FI_SBCL = 0, FI_SBCLE = 3, FI_FNS = 9, FI_FEP = 11, FI_KXEE = 13, FI_KVEE = 15, FRIC_IN = 17 };

__device__ __forceinline__ int padidx(int Wp, int u, int v) { return (u + 2) + (v + 2) * Wp; }

// Bends in the node kernel: computed in place (the default), or replayed from the
// element-major records of k_bend_elem.  Measured at 128 x 128: replaying is slower
// (455 us against 360 us for the node kernel), so the records are off.
constexpr bool USE_BEND_REC = false;

// LDS of the first assembly kernel's value run: one tile's slot planes (81,408 bytes), then
// the weft groups' L-L self block (9 planes) and the two edge groups' private planes
// (ASM1_PRV each: 3 node columns x 16 entries), which afterwards hold the accumulators of
// three groups (3 x RowSinkT::NACC planes) and group 0's compression sum (3 planes) --
// 135,168 bytes, above the 64 KiB default, allowed per kernel at context creation
constexpr int ASM1_PRV = 48;
static_assert(3 * RowSinkT<double>::NACC + 3 <= 2 * ASM1_PRV, "accumulator planes");
constexpr size_t ASM1_LDS = (size_t)(NSLOT + 9 + 2 * ASM1_PRV) * TILE_N * sizeof(double);
constexpr int ASM1_THREADS = 4 * TILE_N;

// Node of thread t in the assembly kernels: one block per 8 x 8 matrix tile, so that a
// wavefront's read-modify-writes of one slot are 512 contiguous bytes.
__device__ __forceinline__ int tile_node(const ClothDev& C, int tile, int lane) {
    int u = (tile % C.ntx) * TILE + (lane & (TILE - 1));
    int v = (tile / C.ntx) * TILE + (lane / TILE);
    return (u < C.W && v < C.Lg) ? u + v * C.W : -1;
}
__device__ __forceinline__ int tile_node(const ClothDev& C, int tile) { return tile_node(C, tile, threadIdx.x); }

// Flipped crossings of the weave (bit 0 of the node flags; plain, twill, satin), as
// src/Cloth.cpp:52-99 sets them.
inline bool pattern_flip(int pattern, int i, int j) {
    if (pattern == 0) return (i + j) % 2 != 0;
    if (pattern == 1) { int rv = j % 4; int ru = (i + rv) % 4; return ru < 2; }
    int rv = (j - 1) % 4; int q = rv / 2; int r = rv % 2; int iu = q * 3 + r * 2; int ru = (i - 1 + iu) % 4;
    return ru == 0;
}

// The cloth's scalars as the element kernels take them (the yarns' stretch and bending
// stiffness from Y, B and the radius, src/Yarn.h:14-15; the jamming threshold,
// src/Cloth.cpp:29); state pointers and tangents unset (null / zero).
inline ClothDev cloth_params(const yc_desc& d) {
    ClothDev C;
    C.W = d.width; C.Lg = d.length; C.N = d.width * d.length; C.Nc = (d.width - 2) * (d.length - 2);
    C.ntx = (d.width + TILE - 1) / TILE;
    C.n_edges = C.Lg * (C.W - 1) + C.W * (C.Lg - 1);
    C.n_bends = C.Lg * (C.W - 2) + C.W * (C.Lg - 2);
    C.L = d.L; C.R = d.R;
    for (int y = 0; y < 2; ++y) {
        int yy = d.n_yarns == 1 ? 0 : y;
        C.rho[y] = d.rho[yy];
        C.ks[y] = d.Y[yy] * PI_REF * std::pow(d.R, 2);   // src/Yarn.h:14-15
        C.kb[y] = d.B[yy] * PI_REF * std::pow(d.R, 2);
    }
    C.kc = d.kc; C.kf = d.kf; C.df = d.df; C.mu = d.mu; C.S = d.S; C.handle_k = d.handle_k;
    for (int i = 0; i < 3; ++i) { C.G[i] = d.G[i]; C.wind_vel[i] = d.wind_vel[i]; }
    C.wind_density = d.wind_density; C.wind_drag = d.wind_drag;
    C.jam_thrd = 2.0 * std::asin(d.R / d.L);   // src/Cloth.cpp:29
    C.x = C.v = C.xfix = C.uv = C.uvd = C.uvbar = C.nrm = nullptr;
    C.flags = nullptr;
    C.dx = C.dv = C.dxfix = C.duv = C.duvd = C.duvbar = C.dnrm = nullptr;
    C.drho[0] = C.drho[1] = C.dks[0] = C.dks[1] = C.dkb[0] = C.dkb[1] = 0.0;
    C.dkc = C.dkf = C.ddf = C.dmu = C.dS = C.dhandle_k = 0.0;
    C.ref_mode = 0;
    return C;
}

}  // namespace

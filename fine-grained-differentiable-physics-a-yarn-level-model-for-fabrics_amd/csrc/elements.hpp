// elements.hpp -- per-element forces and Jacobians of the yarn-level cloth, gfx950.
//
// Gather form: one thread owns one crossing node ("me") and evaluates every element
// that touches it; each element reports its contributions through a sink that keeps
// only the rows owned by the thread.  The element algebra is written in closed form
// (the edge inertia terms use the two-node structure of the edge mass matrix
// instead of explicit 8x8 products) but keeps the reference's values, including its
// non-derivative quirks; each function cites the reference lines it reproduces.
#pragma once
#include <type_traits>
#include "vec3.hpp"

#pragma clang fp contract(off)

namespace ys {

constexpr double PI_REF = 3.1415926535897;   // src/Constants.h:24 (truncated PI)
constexpr int NTHRD = 8;                     // src/Constants.h:26

// Stencil slots of one node row block.  L rows (x,y,z) have 45 slots: 13 neighbour
// offsets x 3 L columns, then E_u at du = -1,0,1 and E_v at dv = -1,0,1.  E_u and E_v
// rows have 12 slots: 3 offsets along their own yarn x 3 L columns, then 3 EE.
constexpr int NSLOT = 159;
constexpr int SLOT_LE = 39;
constexpr int SLOT_EU = 135;
constexpr int SLOT_EV = 147;

// Layout of the system matrix in HBM: the grid is cut into tiles of 8 x 8 nodes and a
// tile stores its slots one after another as planes of 64 doubles (node lane
// (v % 8) * 8 + u % 8).  A wavefront that works on one tile reads a slot as 512
// contiguous bytes.  Slot s of node (u, v) lives at a_base(ntx, u, v) + 64 s, ntx
// being the number of tiles along u.
constexpr int TILE = 8;
constexpr int TILE_N = TILE * TILE;
__host__ __device__ inline size_t a_base(int ntx, int u, int v) {
    return (size_t)((v / TILE) * ntx + (u / TILE)) * NSLOT * TILE_N + (size_t)((v % TILE) * TILE + (u % TILE));
}

struct ClothDev {
    int W, Lg, N, Nc;
    int ntx;              // tiles of the system matrix along u
    int n_edges, n_bends;
    double L, R;
    double rho[2], ks[2], kb[2];
    double kc, kf, df, mu, S, handle_k;
    double G[3], wind_vel[3], wind_density, wind_drag;
    double jam_thrd;
    const double* x;      // 3 planes of N
    const double* v;      // 3 planes
    const double* xfix;   // 3 planes
    const double* uv;     // 2 planes
    const double* uvd;    // 2 planes
    const double* uvbar;  // 2 planes
    const double* nrm;    // 3 planes (zero on the boundary)
    const unsigned char* flags;  // bit0 flip_norm, bit1 handle
    // Tangent run (dd kernels only): directional derivatives of the state arrays (same
    // layouts; null = zero) and of the material parameters, and the gradient mode.
    const double *dx, *dv, *dxfix, *duv, *duvd, *duvbar, *dnrm;
    double drho[2], dks[2], dkb[2], dkc, dkf, ddf, dmu, dS, dhandle_k;
    int ref_mode;   // 1: the reference's detached quantities (non-jamming shear angle)
};

// A parameter or state value as scalar type T (dd carries its tangent).
template <class T> YS_DEV T par(double v, double d);
template <> YS_DEV double par<double>(double v, double) { return v; }
template <> YS_DEV dd par<dd>(double v, double d) { return dd(v, d); }

// Seeded tangents of the adjoint kernels.  There a state field's tangent pointer in
// ClothDev is not an array but a small field code (SEED_X ... SEED_COMP, below
// SEED_MAX), and the tangent of that field is the unit seed of the calling thread: 1 at
// the node, field and components the thread has put into its LDS seed slot
// (ys_seed_node, ys_seed_fm = field << 4 | component mask), 0 everywhere else.  One
// kernel thus differentiates the element code along a different input in every thread.
enum SeedField { SEED_X = 1, SEED_V, SEED_XFIX, SEED_UV, SEED_UVD, SEED_UVBAR, SEED_NRM, SEED_COMP, SEED_MAX = 16 };
constexpr int SEED_THREADS = 256;
__shared__ int ys_seed_node[SEED_THREADS];
__shared__ int ys_seed_fm[SEED_THREADS];
YS_DEV void set_seed(int node, int field, int mask) {
    ys_seed_node[threadIdx.x] = node;
    ys_seed_fm[threadIdx.x] = field << 4 | mask;
}
__host__ __device__ inline const double* seed_ptr(int field) { return reinterpret_cast<const double*>((size_t)field); }
// tangent of component c of node k of the field whose tangent pointer is p (planes of N)
YS_DEV double tan_at(const double* p, int N, int k, int c) {
    const size_t f = reinterpret_cast<size_t>(p);
    if (f < SEED_MAX) {
        if (f == 0) return 0.0;
        const int fm = ys_seed_fm[threadIdx.x];
        return (k == ys_seed_node[threadIdx.x] && (fm >> 4) == (int)f && ((fm >> c) & 1)) ? 1.0 : 0.0;
    }
    return p[(size_t)c * N + k];
}

template <class T> YS_DEV v3<T> plane3(const double* a, const double* da, int N, int k) {
    return V3<T>(par<T>(a[k], tan_at(da, N, k, 0)), par<T>(a[N + k], tan_at(da, N, k, 1)),
                 par<T>(a[2 * N + k], tan_at(da, N, k, 2)));
}
template <class T = double> YS_DEV v3<T> pos(const ClothDev& C, int k) { return plane3<T>(C.x, C.dx, C.N, k); }
template <class T = double> YS_DEV v3<T> vel(const ClothDev& C, int k) { return plane3<T>(C.v, C.dv, C.N, k); }
template <class T = double> YS_DEV T uvc(const ClothDev& C, int k, int c) { return par<T>(C.uv[c * C.N + k], tan_at(C.duv, C.N, k, c)); }
template <class T = double> YS_DEV T uvdc(const ClothDev& C, int k, int c) { return par<T>(C.uvd[c * C.N + k], tan_at(C.duvd, C.N, k, c)); }
template <class T = double> YS_DEV T uvbarc(const ClothDev& C, int k, int c) { return par<T>(C.uvbar[c * C.N + k], tan_at(C.duvbar, C.N, k, c)); }
YS_DEV bool interior(const ClothDev& C, int u, int v) { return u > 0 && u < C.W - 1 && v > 0 && v < C.Lg - 1; }
YS_DEV bool flip(const ClothDev& C, int k) { return (C.flags[k] & 1) != 0; }
YS_DEV int crossing(const ClothDev& C, int u, int v) { return (u - 1) + (v - 1) * (C.W - 2); }

// libgomp static schedule owner of iteration i of n over NTHRD threads.
YS_DEV int omp_owner(int n, int i) {
    int q = n / NTHRD, r = n % NTHRD;
    int cut = r * (q + 1);
    return i < cut ? i / (q + 1) : r + (i - cut) / q;
}

// crimp copy of a node: x - R n on warp yarns, x + R n on weft yarns, x on the
// boundary (src/Cloth.cpp:321-325)
template <class T> YS_DEV v3<T> crimp_pos(const ClothDev& C, int u, int v, int c) {
    int k = u + v * C.W;
    v3<T> p = pos<T>(C, k);
    if (!interior(C, u, v)) return p;
    v3<T> n = plane3<T>(C.nrm, C.dnrm, C.N, k);
    return c == 0 ? p - C.R * n : p + C.R * n;
}

// Slot of the entry in row component rc of a node, column component cc of the node at
// offset (du, dv); -1 outside the stencil.
__host__ __device__ inline int slot_of(int rc, int du, int dv, int cc) {
    if (rc < 3) {
        if (cc < 3) {
            int o;
            if (dv == 0 && du == 0) o = 0;
            else if (dv == 0 && du == -1) o = 1;
            else if (dv == 0 && du == 1) o = 2;
            else if (du == 0 && dv == -1) o = 3;
            else if (du == 0 && dv == 1) o = 4;
            else if (dv == 0 && du == -2) o = 5;
            else if (dv == 0 && du == 2) o = 6;
            else if (du == 0 && dv == -2) o = 7;
            else if (du == 0 && dv == 2) o = 8;
            else if (du == -1 && dv == -1) o = 9;
            else if (du == 1 && dv == -1) o = 10;
            else if (du == -1 && dv == 1) o = 11;
            else if (du == 1 && dv == 1) o = 12;
            else return -1;
            return rc * 45 + o * 3 + cc;
        }
        if (cc == 3 && dv == 0 && du >= -1 && du <= 1) return rc * 45 + SLOT_LE + du + 1;
        if (cc == 4 && du == 0 && dv >= -1 && dv <= 1) return rc * 45 + SLOT_LE + 3 + dv + 1;
        return -1;
    }
    if (rc == 3) {
        if (dv != 0 || du < -1 || du > 1) return -1;
        if (cc < 3) return SLOT_EU + (du + 1) * 3 + cc;
        if (cc == 3) return SLOT_EU + 9 + du + 1;
        return -1;
    }
    if (du != 0 || dv < -1 || dv > 1) return -1;
    if (cc < 3) return SLOT_EV + (dv + 1) * 3 + cc;
    if (cc == 4) return SLOT_EV + 9 + dv + 1;
    return -1;
}

// Inverse of slot_of.
__host__ __device__ inline void slot_decode(int s, int& rc, int& du, int& dv, int& cc) {
    const int odu[13] = {0, -1, 1, 0, 0, -2, 2, 0, 0, -1, 1, -1, 1};
    const int odv[13] = {0, 0, 0, -1, 1, 0, 0, -2, 2, -1, -1, 1, 1};
    if (s < SLOT_EU) {
        rc = s / 45;
        int r = s - rc * 45;
        if (r < SLOT_LE) { du = odu[r / 3]; dv = odv[r / 3]; cc = r % 3; }
        else if (r < SLOT_LE + 3) { cc = 3; du = r - SLOT_LE - 1; dv = 0; }
        else { cc = 4; du = 0; dv = r - SLOT_LE - 4; }
    } else if (s < SLOT_EV) {
        rc = 3;
        int r = s - SLOT_EU;
        dv = 0;
        if (r < 9) { du = r / 3 - 1; cc = r % 3; } else { du = r - 10; cc = 3; }
    } else {
        rc = 4;
        int r = s - SLOT_EV;
        du = 0;
        if (r < 9) { dv = r / 3 - 1; cc = r % 3; } else { dv = r - 10; cc = 4; }
    }
}

enum Kind { KIND_M = 0, KIND_KX = 1, KIND_KV = 2 };

// ---------------------------------------------------------------------------
// Row sink: accumulates the rows of node "me" of A, F, M v, K_v^T v and the
// friction inputs.  A slots are read-modify-written in HBM (one owner per slot).
// T = double: the step.  T = dd: its tangent; A's slots are not touched and the
// tangents of the A entries go to dA when dA is set (exact mode).  With `detach`
// (the reference's gradient), M v and K_v^T v take tangents through v only, as the
// reference builds M and K_v outside the autograd graph (src/MatrixMp.cpp:46-98).
// REF = true fixes the reference's gradient at compile time (detached, no dA): the
// derivative parts of the Jacobian blocks are then unused and the compiler drops them.
// ---------------------------------------------------------------------------
template <class T, bool REF = false>
struct RowSinkT {
    const ClothDev* C;
    int me, mu_, mv_;
    bool inter;
    double* A;          // slot 0 of node "me"; slots TILE_N apart
    double* Aself;      // if set: where the L-L self block (9 planes) goes instead of A
    double* dA;         // tangent run: slot 0 of node "me" in the tangent matrix, or null
    bool detach;
    double coef[3];     // +1, -h^2, -h
    T F[5], Mv[5], KvTv[5];
    T KxEE[2], KvEE[2], FEpart[2];
    v3<T> SBC_L;
    v3<T> SBC_LE[2];
    int fric_owner;
    int bad;
    // While racc_on, the entries of A in this node's rows that an edge emits are summed in
    // registers, per column (0: the node itself, 1: the edge's other end) and entry (0-8 the
    // L-L block row-major, 9-11 L rows x E column, 12-14 E row x L columns, 15 E-E), and
    // edge_end_planes() adds each sum once, instead of one slot update per term.
    // (value run only: the tangent run has no registers to spare for the dual sums)
    static constexpr bool kRacc = std::is_same<T, double>::value;
    bool racc_on;
    double racc[kRacc ? 2 : 1][kRacc ? 16 : 1];

    YS_DEV void init(const ClothDev* c, int k, double* a, double h, double* da = nullptr, bool det = false) {
        C = c; me = k; mu_ = k % c->W; mv_ = k / c->W; inter = interior(*c, mu_, mv_);
        size_t base = a_base(c->ntx, mu_, mv_);
        A = a ? a + base : nullptr; dA = da ? da + base : nullptr; detach = det; Aself = nullptr;
        coef[0] = 1.0; coef[1] = -(h * h); coef[2] = -h;
        for (int i = 0; i < 5; ++i) { F[i] = T(0.0); Mv[i] = T(0.0); KvTv[i] = T(0.0); }
        KxEE[0] = KxEE[1] = KvEE[0] = KvEE[1] = FEpart[0] = FEpart[1] = T(0.0);
        SBC_L = V3<T>(T(0.0), T(0.0), T(0.0)); SBC_LE[0] = SBC_L; SBC_LE[1] = SBC_L;
        fric_owner = inter ? omp_owner(c->Nc, crossing(*c, mu_, mv_)) : -1;
        bad = 0;
        racc_on = false;
    }
    // Rows of one matrix of the system alone (kind 0: M, 1: K_x, 2: K_v, each with
    // coefficient 1): the reference's get_SpaM / get_SpaFPos / get_SpaFVel
    // (src/PhysicsCloth.h:72-77).  kind < 0 keeps A = M - h^2 K_x - h K_v.  The force, M v,
    // K_v^T v and the friction inputs do not depend on the coefficients.
    YS_DEV void select_kind(int kind) {
        if (kind < 0) return;
        for (int i = 0; i < 3; ++i) coef[i] = i == kind ? 1.0 : 0.0;
    }
    YS_DEV int slot(int rc, int kcol, int cc) {
        int s = slot_of(rc, kcol % C->W - mu_, kcol / C->W - mv_, cc);
        if (s < 0) bad = 1;
        return s;
    }
    // true when x went to the register sums (value run, inside an edge)
    YS_DEV bool accA(int kcol, int i, const T& x) {
        if constexpr (kRacc) {
            if (racc_on) {
                if (kcol == me) racc[0][i] += x;
                else racc[1][i] += x;
                return true;
            }
        }
        return false;
    }
    YS_DEV void edge_begin() {
        static_assert(kRacc, "value run only");
        racc_on = true;
        for (int i = 0; i < 16; ++i) { racc[0][i] = 0.0; racc[1][i] = 0.0; }
    }
    // the sums added to private planes P (`stride` apart): 16 planes per node column,
    // column 0 the node itself, other_col (1 or 2) the edge's other end
    YS_DEV void edge_end_planes(double* P, int stride, int other_col) {
        static_assert(kRacc, "value run only");
        racc_on = false;
        for (int i = 0; i < 16; ++i) {
            P[(size_t)i * stride] += racc[0][i];
            P[(size_t)(other_col * 16 + i) * stride] += racc[1][i];
        }
    }
    YS_DEV void addA(int rc, int kcol, int cc, T x) {
        int s = slot(rc, kcol, cc);
        if (s < 0) return;
        if (Aself && rc < 3 && cc < 3 && s == rc * 45 + cc) {   // the L-L self block
            Aself[(size_t)(rc * 3 + cc) * TILE_N] += val(x);
            return;
        }
        if (A) A[(size_t)s * TILE_N] += val(x);
        if (!REF && dA) dA[(size_t)s * TILE_N] += tan_of(x);
    }
    // accumulate into element i of a small array through a switch, so that an index known
    // only at run time (the yarn direction c) does not force the array into scratch memory
    YS_DEV static void add5(T (&a)[5], int i, T x) {
        switch (i) { case 0: a[0] += x; break; case 1: a[1] += x; break; case 2: a[2] += x; break;
                     case 3: a[3] += x; break; default: a[4] += x; break; }
    }
    YS_DEV static void add2(T (&a)[2], int i, T x) { if (i == 0) a[0] += x; else a[1] += x; }
    YS_DEV T velL(int k, int c) { return par<T>(C->v[c * C->N + k], tan_at(C->dv, C->N, k, c)); }
    YS_DEV T velE(int k, int c) { return uvdc<T>(*C, k, c); }
    // a block entry in a product with the velocity (value only when detached)
    YS_DEV T blk(T b) const { return (REF || detach) ? T(val(b)) : b; }

    // LL block B of (force node F, dof node D): A[3F+v, 3D+u] = B[u][v] (src/MatrixMp.cpp:65-76)
    YS_DEV void LL(int kind, int kF, int kD, const mat3<T>& B) {
        if (kF == me) {
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) {
                    if (!accA(kD, v * 3 + u, coef[kind] * B.m[u][v])) addA(v, kD, u, coef[kind] * B.m[u][v]);
                    if (kind == KIND_M) Mv[v] += blk(B.m[u][v]) * velL(kD, u);
                }
        }
        if (kind == KIND_KV && kD == me)
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) KvTv[u] += blk(B.m[u][v]) * velL(kF, v);
    }
    // LE: A[3kL+v, 3N+2e(kE)+c] = B[v]
    YS_DEV void LE(int kind, int kL, int kE, int c, v3<T> B) {
        T b[3] = {B.x, B.y, B.z};
        if (kL == me)
            for (int v = 0; v < 3; ++v) {
                if (!accA(kE, 9 + v, coef[kind] * b[v])) addA(v, kE, 3 + c, coef[kind] * b[v]);
                if (kind == KIND_M) add5(Mv, v, blk(b[v]) * velE(kE, c));
            }
        if (kind == KIND_KV && kE == me)
            for (int v = 0; v < 3; ++v) add5(KvTv, 3 + c, blk(b[v]) * velL(kL, v));
    }
    // EL: A[3N+2e(kE)+c, 3kL+u] = B[u]
    YS_DEV void EL(int kind, int kL, int kE, int c, v3<T> B) {
        T b[3] = {B.x, B.y, B.z};
        if (kE == me)
            for (int u = 0; u < 3; ++u) {
                if (!accA(kL, 12 + u, coef[kind] * b[u])) addA(3 + c, kL, u, coef[kind] * b[u]);
                if (kind == KIND_M) add5(Mv, 3 + c, blk(b[u]) * velL(kL, u));
            }
        if (kind == KIND_KV && kL == me)
            for (int u = 0; u < 3; ++u) add5(KvTv, u, blk(b[u]) * velE(kE, c));
    }
    // EE: A[3N+2e(kF)+c, 3N+2e(kD)+c] = val
    YS_DEV void EE(int kind, int kF, int kD, int c, T v) {
        if (kF == me) {
            if (!accA(kD, 15, coef[kind] * v)) addA(3 + c, kD, 3 + c, coef[kind] * v);
            if (kind == KIND_M) add5(Mv, 3 + c, blk(v) * velE(kD, c));
            if (kD == me) {
                if (kind == KIND_KX) add2(KxEE, c, v);
                if (kind == KIND_KV) add2(KvEE, c, v);
            }
        }
        if (kind == KIND_KV && kD == me) add5(KvTv, 3 + c, blk(v) * velE(kF, c));
    }
    // the accumulators as NACC planes `stride` apart (value run): store, and add in.  Plane
    // of F[c], M v[c], K_v^T v[c]: 3c, 3c + 1, 3c + 2; of KxEE[c], KvEE[c], FEpart[c]:
    // 15 + 3c, 16 + 3c, 17 + 3c; SBC_L: 21-23; SBC_LE[c]: 24 + 3c ..; bad: 30.
    static constexpr int NACC = 31;
    static constexpr int ACC_SBCL = 21;
    YS_DEV void store_acc(double* p, int stride) const {
        int i = 0;
        for (int c = 0; c < 5; ++c) { p[(size_t)(i++) * stride] = val(F[c]); p[(size_t)(i++) * stride] = val(Mv[c]); p[(size_t)(i++) * stride] = val(KvTv[c]); }
        for (int c = 0; c < 2; ++c) { p[(size_t)(i++) * stride] = val(KxEE[c]); p[(size_t)(i++) * stride] = val(KvEE[c]); p[(size_t)(i++) * stride] = val(FEpart[c]); }
        p[(size_t)(i++) * stride] = val(SBC_L.x); p[(size_t)(i++) * stride] = val(SBC_L.y); p[(size_t)(i++) * stride] = val(SBC_L.z);
        for (int c = 0; c < 2; ++c) { p[(size_t)(i++) * stride] = val(SBC_LE[c].x); p[(size_t)(i++) * stride] = val(SBC_LE[c].y); p[(size_t)(i++) * stride] = val(SBC_LE[c].z); }
        p[(size_t)i * stride] = (double)bad;
    }
    YS_DEV void add_acc(const double* p, int stride) {
        int i = 0;
        for (int c = 0; c < 5; ++c) { F[c] += p[(size_t)(i++) * stride]; Mv[c] += p[(size_t)(i++) * stride]; KvTv[c] += p[(size_t)(i++) * stride]; }
        for (int c = 0; c < 2; ++c) { KxEE[c] += p[(size_t)(i++) * stride]; KvEE[c] += p[(size_t)(i++) * stride]; FEpart[c] += p[(size_t)(i++) * stride]; }
        SBC_L.x += p[(size_t)(i++) * stride]; SBC_L.y += p[(size_t)(i++) * stride]; SBC_L.z += p[(size_t)(i++) * stride];
        for (int c = 0; c < 2; ++c) { SBC_LE[c].x += p[(size_t)(i++) * stride]; SBC_LE[c].y += p[(size_t)(i++) * stride]; SBC_LE[c].z += p[(size_t)(i++) * stride]; }
        if (p[(size_t)i * stride] != 0.0) bad = 1;
    }
    // the weft component's sums (F, M v, K_v^T v of E_v; KxEE, KvEE, FEpart, SBC_LE of the
    // weft) of another group's accumulator planes, for the weft friction (value run)
    YS_DEV void add_acc_weft(const double* p, int stride) {
        F[4] += p[(size_t)12 * stride]; Mv[4] += p[(size_t)13 * stride]; KvTv[4] += p[(size_t)14 * stride];
        KxEE[1] += p[(size_t)18 * stride]; KvEE[1] += p[(size_t)19 * stride]; FEpart[1] += p[(size_t)20 * stride];
        SBC_LE[1].x += p[(size_t)27 * stride]; SBC_LE[1].y += p[(size_t)28 * stride]; SBC_LE[1].z += p[(size_t)29 * stride];
    }
    YS_DEV void FL(int k, v3<T> f) { if (k == me) { F[0] += f.x; F[1] += f.y; F[2] += f.z; } }
    YS_DEV void addF(int i, T f) { add5(F, i, f); }   // a force row of the node itself (the friction)
    // E force on the crossing of node k; elem_owner is the static owner thread of the
    // emitting element in its own loop, for the partial sum read by the friction
    YS_DEV void FE(int k, int c, T f, int elem_owner) {
        if (k == me) {
            add5(F, 3 + c, f);
            if (elem_owner == fric_owner) add2(FEpart, c, f);
        }
    }
    YS_DEV void SBCL(int k, v3<T> f) { if (k == me) SBC_L = SBC_L + f; }
    YS_DEV void SBCLE(int k, int c, v3<T> f) {
        if (k != me) return;
        if (c == 0) SBC_LE[0] = SBC_LE[0] + f; else SBC_LE[1] = SBC_LE[1] + f;
    }
};
typedef RowSinkT<double> RowSink;

// The shared placement of an edge's blocks (src/PhysicsCloth.cpp:267-337, 1529-1599):
// LL four blocks; LE/EL/EE self blocks of each interior end; cross blocks when both
// ends are interior.  Names: LEab = dF_La/dE_b, ELab = dF_Ea/dL_b, EEab = dF_Ea/dE_b.
template <class T>
struct EdgeBlocks {
    mat3<T> LL00, LL01, LL10, LL11;
    v3<T> LE00, LE01, LE10, LE11, EL00, EL01, EL10, EL11;
    T EE00, EE01, EE10, EE11;
};

template <class T, class Sink>
YS_DEV void emit_edge(Sink& S, int kind, int k0, int k1, bool i0, bool i1, int c, const EdgeBlocks<T>& b, bool ll) {
    if (ll) {
        S.LL(kind, k0, k0, b.LL00);
        S.LL(kind, k0, k1, b.LL01);
        S.LL(kind, k1, k0, b.LL10);
        S.LL(kind, k1, k1, b.LL11);
    }
    if (i0) { S.EL(kind, k0, k0, c, b.EL00); S.LE(kind, k0, k0, c, b.LE00); S.EE(kind, k0, k0, c, b.EE00); }
    if (i1) { S.EL(kind, k1, k1, c, b.EL11); S.LE(kind, k1, k1, c, b.LE11); S.EE(kind, k1, k1, c, b.EE11); }
    if (i0 && i1) {
        S.LE(kind, k1, k0, c, b.LE10);
        S.LE(kind, k0, k1, c, b.LE01);
        S.EL(kind, k1, k0, c, b.EL01);   // row E(n0), col L(n1)
        S.EL(kind, k0, k1, c, b.EL10);   // row E(n1), col L(n0)
        S.EE(kind, k0, k1, c, b.EE01);
        S.EE(kind, k1, k0, c, b.EE10);
    }
}

// ---------------------------------------------------------------------------
// Edge: Mass, Gravity, Stretch, Inertia, MDotVDot, ParaColli
// (src/PhysicsCloth.cpp:250-577, 989-2577; geometry src/ClothUnit.h:134-154)
// ---------------------------------------------------------------------------
template <class T, class Sink>
YS_DEV void edge_element(const ClothDev& C, Sink& S, int u0, int v0, int c, int yarn, int owner) {
    typedef v3<T> V;
    typedef mat3<T> M;
    int u1 = u0 + (c == 0), v1 = v0 + (c == 1);
    int k0 = u0 + v0 * C.W, k1 = u1 + v1 * C.W;
    bool i0 = interior(C, u0, v0), i1 = interior(C, u1, v1);
    V x0 = pos<T>(C, k0), x1 = pos<T>(C, k1);
    V q0 = vel<T>(C, k0), q1 = vel<T>(C, k1);
    T s0 = uvdc<T>(C, k0, c), s1 = uvdc<T>(C, k1, c);
    T D = uvc<T>(C, k1, c) - uvc<T>(C, k0, c);
    V dx = x1 - x0;
    V w = V3<T>(dx.x / D, dx.y / D, dx.z / D);
    T l = nrm(dx);
    V d = V3<T>(dx.x / l, dx.y / l, dx.z / l);
    M P = projector(d);
    T rho = par<T>(C.rho[yarn], C.drho[yarn]), ks = par<T>(C.ks[yarn], C.dks[yarn]);
    T ww = dot(w, w);
    T nw = nrm(w);
    EdgeBlocks<T> b;

    // two-node mass pattern [[2,1],[1,2]]: mixed velocities
    V V0 = 2.0 * q0 + q1, V1 = q0 + 2.0 * q1;
    T S0 = 2.0 * s0 + s1, S1 = s0 + 2.0 * s1;
    V Pvs = S0 * q0 + S1 * q1;
    T Sss = s0 * S0 + s1 * S1;
    T Vvv = dot(q0, V0) + dot(q1, V1);

    // --- Mass (src/PhysicsCloth.cpp:250-338) ---
    {
        T cf = (1.0 / 6.0) * rho * D;
        b.LL00 = I33T<T>(2.0 * cf); b.LL11 = b.LL00; b.LL01 = I33T<T>(cf); b.LL10 = b.LL01;
        b.LE00 = (-2.0 * cf) * w; b.LE11 = b.LE00; b.LE01 = (-1.0 * cf) * w; b.LE10 = b.LE01;
        b.EL00 = b.LE00; b.EL11 = b.LE00; b.EL01 = b.LE01; b.EL10 = b.LE01;
        b.EE00 = 2.0 * cf * ww; b.EE11 = b.EE00; b.EE01 = 0.5 * b.EE00; b.EE10 = b.EE01;
        emit_edge(S, KIND_M, k0, k1, i0, i1, c, b, true);
    }
    // --- Gravity_F / Gravity_D (src/PhysicsCloth.cpp:340-425) ---
    {
        V G = V3<T>(C.G[0], C.G[1], C.G[2]);
        V fl = (-0.5 * rho * D) * G;
        S.FL(k0, fl); S.FL(k1, fl);
        T gx = dot(G, x0 + x1);
        if (i0) S.FE(k0, c, 0.5 * rho * gx, owner);
        if (i1) S.FE(k1, c, -0.5 * rho * gx, owner);
        V g = (0.5 * rho) * G;
        if (i0) { S.EL(KIND_KX, k0, k0, c, g); S.LE(KIND_KX, k0, k0, c, g); }
        if (i1) { S.EL(KIND_KX, k1, k1, c, -g); S.LE(KIND_KX, k1, k1, c, -g); }
        if (i0 && i1) {
            S.LE(KIND_KX, k1, k0, c, g);     // L1E0
            S.LE(KIND_KX, k0, k1, c, -g);    // L0E1
            S.EL(KIND_KX, k1, k0, c, g);     // E0L1
            S.EL(KIND_KX, k0, k1, c, -g);    // E1L0
        }
    }
    // --- Stretch_F / Stretch_D (src/PhysicsCloth.cpp:427-577) ---
    {
        V f = (ks * (nw - 1.0)) * d;
        S.FL(k0, f); S.FL(k1, -f);
        T sg = c == 0 ? 1.0 : -1.0;
        S.SBCL(k0, sg * f); S.SBCL(k1, -sg * f);
        T ew = 0.5 * ks * (pow(nw, 2.0) - 1.0);
        if (i0) S.FE(k0, c, -ew, owner);
        if (i1) S.FE(k1, c, ew, owner);
        M K = (ks / l) * P - I33T<T>(ks / D);
        b.LL00 = K; b.LL11 = K; b.LL01 = -K; b.LL10 = -K;
        V le = (ks * (nw / D)) * d;
        b.LE00 = le; b.LE11 = le; b.LE01 = -le; b.LE10 = -le;
        V el = (ks / D) * w;
        b.EL00 = el; b.EL11 = el; b.EL01 = -el; b.EL10 = -el;
        T ee = -ks * (pow(nw, 2.0) / D);
        b.EE00 = ee; b.EE11 = ee; b.EE01 = -ee; b.EE10 = -ee;
        emit_edge(S, KIND_KX, k0, k1, i0, i1, c, b, true);
        if (i0) S.SBCLE(k0, c, sg * le);
        if (i1) S.SBCLE(k1, c, sg * le);
    }
    // --- Inertia_F (src/PhysicsCloth.cpp:1071-1186): T_x = 1/2 qd^T dM/dx qd ---
    {
        V Tx0 = (rho / 6.0) * (Pvs - Sss * w);
        S.FL(k0, Tx0); S.FL(k1, -Tx0);
        T Tuv0 = (rho / 12.0) * (ww * Sss - Vvv);
        if (i0) S.FE(k0, c, Tuv0, owner);
        if (i1) S.FE(k1, c, -Tuv0, owner);
    }
    // --- Inertia_D_Pos (src/PhysicsCloth.cpp:1188-1600) ---
    {
        T Ti = (rho / 6.0) * Sss / (D * D);
        b.LL00 = I33T<T>(Ti); b.LL01 = I33T<T>(-Ti); b.LL10 = I33T<T>(-Ti); b.LL11 = I33T<T>(Ti);
        V le = (-rho / (6.0 * D)) * (Pvs + Sss * w);
        b.LE00 = le; b.LE01 = -le; b.LE10 = -le; b.LE11 = le;
        V el = (rho / (6.0 * D)) * ((3.0 * Sss) * w - Pvs);
        b.EL00 = el; b.EL01 = -el; b.EL10 = -el; b.EL11 = el;
        T ee = (rho / 6.0) * (ww / D) * Sss;
        b.EE00 = ee; b.EE01 = -ee; b.EE10 = -ee; b.EE11 = ee;
        emit_edge(S, KIND_KX, k0, k1, i0, i1, c, b, true);
    }
    // --- Inertia_D_Vel (src/PhysicsCloth.cpp:1602-1901): entry (i,j) = (dM/dq_i qd)_j ---
    {
        T r6 = rho / 6.0;
        b.LL00 = I33T<T>(r6 * S0); b.LL01 = I33T<T>(r6 * S1); b.LL10 = I33T<T>(-r6 * S0); b.LL11 = I33T<T>(-r6 * S1);
        b.LE00 = r6 * (V0 - (2.0 * S0) * w); b.LE01 = r6 * (V1 - (2.0 * S1) * w);
        b.LE10 = -b.LE00; b.LE11 = -b.LE01;
        b.EL00 = -r6 * V0; b.EL01 = -r6 * V1; b.EL10 = r6 * V0; b.EL11 = r6 * V1;
        b.EE00 = r6 * ww * S0; b.EE01 = r6 * ww * S1; b.EE10 = -b.EE00; b.EE11 = -b.EE01;
        emit_edge(S, KIND_KV, k0, k1, i0, i1, c, b, true);
    }
    // --- MDotVDot_F / _D_Pos / _D_Vel (src/PhysicsCloth.cpp:1903-2577) ---
    {
        T Dd = s1 - s0;
        V dq = q1 - q0;
        V wd = (1.0 / (D * D)) * (D * dq - Dd * dx);
        T wwd = dot(w, wd);
        T r6 = rho / 6.0;
        // (Mdot qd) per node part a: L and E components
        V MdL0 = (r6 * Dd) * (V0 - S0 * w) - (r6 * D * S0) * wd;
        V MdL1 = (r6 * Dd) * (V1 - S1 * w) - (r6 * D * S1) * wd;
        T MdE0 = r6 * Dd * (ww * S0 - dot(w, V0)) + r6 * D * (2.0 * wwd * S0 - dot(wd, V0));
        T MdE1 = r6 * Dd * (ww * S1 - dot(w, V1)) + r6 * D * (2.0 * wwd * S1 - dot(wd, V1));
        S.FL(k0, -MdL0); S.FL(k1, -MdL1);
        if (i0) S.FE(k0, c, -MdE0, owner);
        if (i1) S.FE(k1, c, -MdE1, owner);

        // D_Pos (added as written): x-columns only feed E rows; uv-columns feed L and E rows
        {
            T r3 = rho / 3.0;
            b.LL00 = Z33T<T>(); b.LL01 = Z33T<T>(); b.LL10 = Z33T<T>(); b.LL11 = Z33T<T>();
            b.EL00 = (-r3 * S0) * wd; b.EL01 = (r3 * S0) * wd;
            b.EL10 = (-r3 * S1) * wd; b.EL11 = (r3 * S1) * wd;
            V g = (1.0 / (D * D)) * dq - (4.0 * Dd / (D * D * D)) * dx;
            T gw = dot(g, w);
            // column uv_end, sgn = -1 (end 0) / +1 (end 1)
            for (int end = 0; end < 2; ++end) {
                T sg = end == 0 ? -1.0 : 1.0;
                T wTw_dot_uv = 2.0 * sg * wwd / D - 2.0 * sg * gw;
                V L0 = (r6 * Dd * sg * S0 / D) * w - r6 * MdL0 + (r6 * D * sg * S0) * g;
                V L1 = (r6 * Dd * sg * S1 / D) * w - r6 * MdL1 + (r6 * D * sg * S1) * g;
                T E0 = r6 * Dd * (sg * dot(w, V0) / D + 2.0 * sg * ww * S0 / D) - r6 * MdE0 +
                            r6 * D * (sg * dot(g, V0) + wTw_dot_uv * S0);
                T E1 = r6 * Dd * (sg * dot(w, V1) / D + 2.0 * sg * ww * S1 / D) - r6 * MdE1 +
                            r6 * D * (sg * dot(g, V1) + wTw_dot_uv * S1);
                if (end == 0) { b.LE00 = L0; b.LE10 = L1; b.EE00 = E0; b.EE10 = E1; }
                else { b.LE01 = L0; b.LE11 = L1; b.EE01 = E0; b.EE11 = E1; }
            }
            emit_edge(S, KIND_KX, k0, k1, i0, i1, c, b, false);
        }
        // D_Vel (subtracted as written)
        {
            V WW = Dd * w + D * wd;
            T EWq = Dd * ww + 2.0 * D * wwd;
            const double K2[2][2] = {{2.0, 1.0}, {1.0, 2.0}};
            const V Vb[2] = {V0, V1};
            const T Sb[2] = {S0, S1};
            M LLb[2][2];
            V ELb[2][2], LEb[2][2];
            T EEb[2][2];
            for (int ap = 0; ap < 2; ++ap)
                for (int end = 0; end < 2; ++end) {
                    T sg = end == 0 ? -1.0 : 1.0;
                    T k2 = K2[ap][end];
                    LLb[ap][end] = I33T<T>(-r6 * (-sg * Sb[ap] + k2 * Dd));
                    ELb[ap][end] = -r6 * ((-sg) * Vb[ap] + (2.0 * sg * Sb[ap]) * w - k2 * WW);
                    LEb[ap][end] = -r6 * (sg * Vb[ap] - k2 * WW);
                    EEb[ap][end] = -r6 * (-sg * ww * Sb[ap] + k2 * EWq);
                }
            b.LL00 = LLb[0][0]; b.LL01 = LLb[0][1]; b.LL10 = LLb[1][0]; b.LL11 = LLb[1][1];
            b.EL00 = ELb[0][0]; b.EL01 = ELb[0][1]; b.EL10 = ELb[1][0]; b.EL11 = ELb[1][1];
            b.LE00 = LEb[0][0]; b.LE01 = LEb[0][1]; b.LE10 = LEb[1][0]; b.LE11 = LEb[1][1];
            b.EE00 = EEb[0][0]; b.EE01 = EEb[0][1]; b.EE10 = EEb[1][0]; b.EE11 = EEb[1][1];
            emit_edge(S, KIND_KV, k0, k1, i0, i1, c, b, true);
        }
    }
    // --- ParaColli_F / _D (src/PhysicsCloth.cpp:989-1069): EE uses the value (Q6) ---
    {
        double dpc = flip(C, k0) != flip(C, k1) ? 4.0 * C.R : 2.0 * C.R;
        T rl = dpc - D > 0.0 ? dpc - D : T(0.0);
        T Fu0 = par<T>(C.kc, C.dkc) * C.L * rl;
        if (i0) S.FE(k0, c, Fu0, owner);
        if (i1) S.FE(k1, c, -Fu0, owner);
        T val = -par<T>(C.kc, C.dkc) * C.L * rl;
        if (i0) S.EE(KIND_KX, k0, k0, c, val);
        if (i1) S.EE(KIND_KX, k1, k1, c, val);
        if (i0 && i1) { S.EE(KIND_KX, k0, k1, c, val); S.EE(KIND_KX, k1, k0, c, val); }
    }
}

// ---------------------------------------------------------------------------
// Bending (src/PhysicsCloth.cpp:579-987; geometry src/ClothUnit.h:197-210)
// ---------------------------------------------------------------------------
template <class T>
struct BendVals {
    v3<T> Fx1, Fx2, Fx0;
    T Fu2, Fu1;
    mat3<T> L11, L12, L21, L22;
    v3<T> L1E1, L2E1, E1L1, E1L2;
    T E1E1;
};

template <class T>
YS_DEV BendVals<T> bend_values(v3<T> x2, v3<T> x0, v3<T> x1, T D, T kb) {
    typedef v3<T> V;
    typedef mat3<T> M;
    BendVals<T> o;
    V a = x1 - x0, bb = x2 - x0;
    T l1 = nrm(a), l2 = nrm(bb);
    V d1 = V3<T>(a.x / l1, a.y / l1, a.z / l1), d2 = V3<T>(bb.x / l2, bb.y / l2, bb.z / l2);
    M P1 = projector(d1), P2 = projector(d2);
    T ct = -dot(d1, d2);
    ct = fmin(fmax(ct, -1.0), 1.0);   // the reference clamps to [-1+DBL_MIN, 1-DBL_MIN]
    T th = acos(ct);
    T s = sin(th), c = cos(th);
    o.Fx1 = V3<T>(0, 0, 0); o.Fx2 = o.Fx1;
    o.L11 = Z33T<T>(); o.L12 = o.L11; o.L21 = o.L11; o.L22 = o.L11;
    o.L1E1 = o.Fx1; o.L2E1 = o.Fx1; o.E1L1 = o.Fx1; o.E1L2 = o.Fx1;
    if (s != 0.0) {
        V P1d2 = P1 * d2, P2d1 = P2 * d1;
        o.Fx1 = ((-2.0 * kb * th) / (l1 * D * s)) * P1d2;
        o.Fx2 = ((-2.0 * kb * th) / (l2 * D * s)) * P2d1;
        T cs2 = c / pow(s, 2.0);
        M A1 = outerP(P1d2, d2, P1);   // P1 d2 d2^T P1
        M A2 = outerP(P2d1, d1, P2);
        o.L11 = ((2.0 * kb) / (pow(l1, 2.0) * D * s)) *
                (th * (outer(P1d2, d1) + cs2 * A1 + c * P1 + outerP(d1, d2, P1)) - (1.0 / s) * A1);
        o.L12 = ((-2.0 * kb) / (l1 * l2 * D * s)) *
                ((th * (P1 - cs2 * outer(P1d2, d1)) + (1.0 / s) * outer(P1d2, d1)) * P2);
        o.L21 = ((-2.0 * kb) / (l1 * l2 * D * s)) *
                ((th * (P2 - cs2 * outer(P2d1, d2)) + (1.0 / s) * outer(P2d1, d2)) * P1);
        o.L22 = ((2.0 * kb) / (pow(l2, 2.0) * D * s)) *
                (th * (outer(P2d1, d2) + cs2 * A2 + c * P2 + outerP(d2, d1, P2)) - (1.0 / s) * A2);
        T c1 = (2.0 * kb * th) / (l1 * pow(D, 2.0) * s), c2 = (2.0 * kb * th) / (l2 * pow(D, 2.0) * s);
        o.L1E1 = c1 * P1d2;
        o.L2E1 = c2 * P2d1;
        o.E1L1 = c1 * vTm(d2, P1);
        o.E1L2 = c2 * vTm(d1, P2);
    }
    o.Fu2 = (-kb * th) / pow(D, 2.0);
    o.Fu1 = -o.Fu2;
    o.Fx0 = -(o.Fx1 + o.Fx2);
    o.E1E1 = (-2.0 * kb * pow(th, 2.0)) / pow(D, 3.0);
    return o;
}

// bend along yarn direction c centred on node (uc, vc): n2 = previous, n1 = next
// parts: bit 0 the bend of the crossing nodes, bit 1 its crimp twin (the adjoint passes
// that seed the normals need only the twin)
template <class T, class Sink>
YS_DEV void bend_element(const ClothDev& C, Sink& S, int uc, int vc, int c, int yarn, int owner, int parts = 3) {
    typedef v3<T> V;
    typedef mat3<T> M;
    int du = c == 0 ? 1 : 0, dv = c == 1 ? 1 : 0;
    int u2 = uc - du, v2 = vc - dv, u1 = uc + du, v1 = vc + dv;
    int k2 = u2 + v2 * C.W, k0 = uc + vc * C.W, k1 = u1 + v1 * C.W;
    bool i2 = interior(C, u2, v2), i0 = interior(C, uc, vc), i1 = interior(C, u1, v1);
    T kb = par<T>(C.kb[yarn], C.dkb[yarn]);
    // Bending_F / Bending_D on the crossing nodes
    if (parts & 1) {
        T D = uvc<T>(C, k1, c) - uvc<T>(C, k2, c);
        BendVals<T> o = bend_values<T>(pos<T>(C, k2), pos<T>(C, k0), pos<T>(C, k1), D, kb);
        S.FL(k2, o.Fx2); S.FL(k0, o.Fx0); S.FL(k1, o.Fx1);
        if (i2) S.FE(k2, c, o.Fu2, owner);
        if (i1) S.FE(k1, c, o.Fu1, owner);
        M L10 = -(o.L11 + o.L12), L20 = -(o.L21 + o.L22);
        M L01 = -(o.L11 + o.L21), L02 = -(o.L12 + o.L22);
        M L00 = -(L10 + L20);
        S.LL(KIND_KX, k2, k2, o.L22); S.LL(KIND_KX, k2, k0, L20); S.LL(KIND_KX, k2, k1, o.L21);
        S.LL(KIND_KX, k0, k2, L02); S.LL(KIND_KX, k0, k0, L00); S.LL(KIND_KX, k0, k1, L01);
        S.LL(KIND_KX, k1, k2, o.L12); S.LL(KIND_KX, k1, k0, L10); S.LL(KIND_KX, k1, k1, o.L11);
        V L2E2 = -o.L2E1, E2L2 = -o.E1L2;
        V L0E1 = -(o.L1E1 + o.L2E1), L0E2 = -L0E1;
        V E1L0 = -(o.E1L1 + o.E1L2), E2L0 = -E1L0;
        if (i2) { S.LE(KIND_KX, k2, k2, c, L2E2); S.EL(KIND_KX, k2, k2, c, E2L2); S.EE(KIND_KX, k2, k2, c, o.E1E1); }
        if (i1) { S.LE(KIND_KX, k1, k1, c, o.L1E1); S.EL(KIND_KX, k1, k1, c, o.E1L1); S.EE(KIND_KX, k1, k1, c, o.E1E1); }
        if (i2 && i0) { S.LE(KIND_KX, k0, k2, c, L0E2); S.EL(KIND_KX, k2, k0, c, E2L0); }   // (Q8 placement)
        if (i0 && i1) { S.LE(KIND_KX, k0, k1, c, L0E1); S.EL(KIND_KX, k0, k1, c, E1L0); }
    }
    // Crimp_Bending_F / _D on the +-R n copies: compression inputs only (src/PhysicsCloth.cpp:824-987);
    // the copies keep their construction-time Eulerian coordinates (u L, v L)
    if (parts & 2) {
        T D = c == 0 ? (double)u1 * C.L - (double)u2 * C.L : (double)v1 * C.L - (double)v2 * C.L;
        BendVals<T> o = bend_values<T>(crimp_pos<T>(C, u2, v2, c), crimp_pos<T>(C, uc, vc, c), crimp_pos<T>(C, u1, v1, c), D, kb);
        T sg = c == 0 ? 1.0 : -1.0;
        S.SBCL(k2, sg * o.Fx2); S.SBCL(k0, sg * o.Fx0); S.SBCL(k1, sg * o.Fx1);
        if (i2) S.SBCLE(k2, c, sg * (-o.L2E1));
        if (i1) S.SBCLE(k1, c, sg * o.L1E1);
    }
}

// ---------------------------------------------------------------------------
// Element-major bends (value run): one thread per bend computes the bend and its crimp
// twin once and stores, for each of its three nodes (role 0: n2, 1: n0, 2: n1), the
// values that node's rows receive, in the order bend_element emits them.  A record of
// BEND_REC doubles per role, planes rec[(r * BEND_REC + f) * nb + b]; b is the bend's
// index in the reference's list (warp bends v * (W-2) + uc - 1, then weft bends).
// ---------------------------------------------------------------------------
constexpr int BEND_REC = 48;

YS_DEV void bend_decode(const ClothDev& C, int b, int& uc, int& vc, int& c, int& yarn) {
    int nb_warp = C.Lg * (C.W - 2);
    if (b < nb_warp) { c = 0; vc = b / (C.W - 2); uc = 1 + b % (C.W - 2); yarn = vc & 1; }
    else { int q = b - nb_warp; c = 1; uc = q / (C.Lg - 2); vc = 1 + q % (C.Lg - 2); yarn = uc & 1; }
}

struct RecW {   // strided writer of one role's record
    double* p; size_t st;
    YS_DEV void put(int f, double x) const { p[(size_t)f * st] = x; }
    YS_DEV void put3(int f, d3 v) const { put(f, v.x); put(f + 1, v.y); put(f + 2, v.z); }
    YS_DEV void putm(int f, const m33& m) const { for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) put(f + 3 * a + b, m.m[a][b]); }
};
struct RecR {   // strided reader
    const double* p; size_t st;
    YS_DEV double get(int f) const { return p[(size_t)f * st]; }
    YS_DEV d3 get3(int f) const { return D3(get(f), get(f + 1), get(f + 2)); }
    YS_DEV m33 getm(int f) const { m33 m; for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) m.m[a][b] = get(f + 3 * a + b); return m; }
};

__global__ void __launch_bounds__(256) k_bend_elem(ClothDev C, double* rec, int nb) {
    int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    int uc, vc, c, yarn;
    bend_decode(C, b, uc, vc, c, yarn);
    int du = c == 0 ? 1 : 0, dv = c == 1 ? 1 : 0;
    int u2 = uc - du, v2 = vc - dv, u1 = uc + du, v1 = vc + dv;
    int k2 = u2 + v2 * C.W, k0 = uc + vc * C.W, k1 = u1 + v1 * C.W;
    double kb = C.kb[yarn];
    RecW w0 = {rec + b, (size_t)nb}, w1 = {rec + (size_t)BEND_REC * nb + b, (size_t)nb},
         w2 = {rec + (size_t)2 * BEND_REC * nb + b, (size_t)nb};
    {
        double D = uvc(C, k1, c) - uvc(C, k2, c);
        BendVals<double> o = bend_values<double>(pos(C, k2), pos(C, k0), pos(C, k1), D, kb);
        m33 L10 = -(o.L11 + o.L12), L20 = -(o.L21 + o.L22);
        m33 L01 = -(o.L11 + o.L21), L02 = -(o.L12 + o.L22);
        m33 L00 = -(L10 + L20);
        d3 L2E2 = -o.L2E1, E2L2 = -o.E1L2;
        d3 L0E1 = -(o.L1E1 + o.L2E1), L0E2 = -L0E1;
        d3 E1L0 = -(o.E1L1 + o.E1L2), E2L0 = -E1L0;
        w0.put3(0, o.Fx2); w0.put(3, o.Fu2); w0.putm(4, o.L22); w0.putm(13, L20); w0.putm(22, o.L21);
        w0.put3(31, L2E2); w0.put3(34, E2L2); w0.put(37, o.E1E1);
        w1.put3(0, o.Fx0); w1.putm(3, L02); w1.putm(12, L00); w1.putm(21, L01);
        w1.put3(30, L0E2); w1.put3(33, E2L0); w1.put3(36, L0E1);
        w2.put3(0, o.Fx1); w2.put(3, o.Fu1); w2.putm(4, o.L12); w2.putm(13, L10); w2.putm(22, o.L11);
        w2.put3(31, o.L1E1); w2.put3(34, o.E1L1); w2.put(37, o.E1E1); w2.put3(38, E1L0);
    }
    {
        double D = c == 0 ? (double)u1 * C.L - (double)u2 * C.L : (double)v1 * C.L - (double)v2 * C.L;
        BendVals<double> o = bend_values<double>(crimp_pos<double>(C, u2, v2, c), crimp_pos<double>(C, uc, vc, c),
                                                 crimp_pos<double>(C, u1, v1, c), D, kb);
        double sg = c == 0 ? 1.0 : -1.0;
        w0.put3(38, sg * o.Fx2); w0.put3(41, sg * (-o.L2E1));
        w1.put3(39, sg * o.Fx0);
        w2.put3(41, sg * o.Fx1); w2.put3(44, sg * o.L1E1);
    }
}

// Replays bend b's record for node role r, in bend_element's order restricted to the
// calls that reach that node's rows.
template <class Sink>
YS_DEV void emit_bend_role(const ClothDev& C, Sink& S, const double* rec, int nb, int b, int r, int owner) {
    int uc, vc, c, yarn;
    bend_decode(C, b, uc, vc, c, yarn);
    int du = c == 0 ? 1 : 0, dv = c == 1 ? 1 : 0;
    int u2 = uc - du, v2 = vc - dv, u1 = uc + du, v1 = vc + dv;
    int k2 = u2 + v2 * C.W, k0 = uc + vc * C.W, k1 = u1 + v1 * C.W;
    bool i2 = interior(C, u2, v2), i0 = interior(C, uc, vc), i1 = interior(C, u1, v1);
    RecR R = {rec + (size_t)r * BEND_REC * nb + b, (size_t)nb};
    if (r == 0) {
        S.FL(k2, R.get3(0));
        if (i2) S.FE(k2, c, R.get(3), owner);
        S.LL(KIND_KX, k2, k2, R.getm(4)); S.LL(KIND_KX, k2, k0, R.getm(13)); S.LL(KIND_KX, k2, k1, R.getm(22));
        if (i2) { S.LE(KIND_KX, k2, k2, c, R.get3(31)); S.EL(KIND_KX, k2, k2, c, R.get3(34)); S.EE(KIND_KX, k2, k2, c, R.get(37)); }
        S.SBCL(k2, R.get3(38));
        if (i2) S.SBCLE(k2, c, R.get3(41));
    } else if (r == 1) {
        S.FL(k0, R.get3(0));
        S.LL(KIND_KX, k0, k2, R.getm(3)); S.LL(KIND_KX, k0, k0, R.getm(12)); S.LL(KIND_KX, k0, k1, R.getm(21));
        if (i2 && i0) { S.LE(KIND_KX, k0, k2, c, R.get3(30)); S.EL(KIND_KX, k2, k0, c, R.get3(33)); }
        if (i0 && i1) S.LE(KIND_KX, k0, k1, c, R.get3(36));
        S.SBCL(k0, R.get3(39));
    } else {
        S.FL(k1, R.get3(0));
        if (i1) S.FE(k1, c, R.get(3), owner);
        S.LL(KIND_KX, k1, k2, R.getm(4)); S.LL(KIND_KX, k1, k0, R.getm(13)); S.LL(KIND_KX, k1, k1, R.getm(22));
        if (i1) { S.LE(KIND_KX, k1, k1, c, R.get3(31)); S.EL(KIND_KX, k1, k1, c, R.get3(34)); S.EE(KIND_KX, k1, k1, c, R.get(37)); }
        if (i0 && i1) S.EL(KIND_KX, k0, k1, c, R.get3(38));
        S.SBCL(k1, R.get3(41));
        if (i1) S.SBCLE(k1, c, R.get3(44));
    }
}

// ---------------------------------------------------------------------------
// Constraint_F / Constraint_D of a handle node k (src/PhysicsCloth.cpp:232-248)
// ---------------------------------------------------------------------------
template <class T, class Sink>
YS_DEV void handle_element(const ClothDev& C, Sink& S, int k) {
    v3<T> xf = plane3<T>(C.xfix, C.dxfix, C.N, k);
    T hk = par<T>(C.handle_k, C.dhandle_k);
    S.FL(k, (-hk) * (pos<T>(C, k) - xf));
    S.LL(KIND_KX, k, k, I33T<T>(-hk));
}

// ---------------------------------------------------------------------------
// Wind_F on one face (src/PhysicsCloth.cpp:3091-3106; src/ClothUnit.h:268-273)
// ---------------------------------------------------------------------------
template <class T, class Sink>
YS_DEV void face_element(const ClothDev& C, Sink& S, int ka, int kb2, int kc) {
    typedef v3<T> V;
    V xa = pos<T>(C, ka), xb = pos<T>(C, kb2), xc = pos<T>(C, kc);
    V cr = cross(xb - xa, xc - xa);
    T cn = nrm(cr);
    T area = 0.5 * cn;
    V n = cn == 0.0 ? cr : V3<T>(cr.x / cn, cr.y / cn, cr.z / cn);
    V fv = (vel<T>(C, ka) + vel<T>(C, kb2) + vel<T>(C, kc)) / 3.0;
    V rel = V3<T>(C.wind_vel[0], C.wind_vel[1], C.wind_vel[2]) - fv;
    T vn = dot(n, rel);
    V vt = rel - vn * n;
    V wf = (C.wind_density * area * fabs(vn) * vn) * n + (C.wind_drag * area) * vt;
    V w3 = wf / 3.0;
    S.FL(ka, w3); S.FL(kb2, w3); S.FL(kc, w3);
}

// ---------------------------------------------------------------------------
// ConSlideFriction at the thread's own crossing (src/PhysicsCloth.cpp:2579-2641),
// run after every edge / bend of the node has been accumulated.  Returns comp.
// ---------------------------------------------------------------------------
// Components c_lo .. c_hi - 1 (both by default; the first assembly kernel's value run
// gives the warp and the weft component to two threads).
template <class T, class Sink>
YS_DEV T friction(const ClothDev& C, Sink& S, T* is_slide, int c_lo = 0, int c_hi = 2) {
    typedef v3<T> V;
    int k = S.me;
    V n = plane3<T>(C.nrm, C.dnrm, C.N, k);
    T comp = 0.5 * dot(n, S.SBC_L);
    comp = comp > 0.0 ? comp : 0.0;
    T lim = comp * par<T>(C.mu, C.dmu);
    for (int c = c_lo; c < c_hi; ++c) {
        T dl = uvc<T>(C, k, c) - uvbarc<T>(C, k, c);
        T cEE = 0.5 * dot(n, S.SBC_LE[c]);
        T FNS = S.F[3 + c];          // F_E before friction (F_E_No_Slide)
        T Fpt = S.FEpart[c];         // thread-partial F_Es[ID] (Q2)
        T th = tanh(10.0 * dl);
        is_slide[c] = 1.0 / (1.0 + exp(-(1e4 * (fabs(FNS) - lim))));
        T fric = (-0.5 * (par<T>(C.kf, C.dkf) * dl - th * comp * par<T>(C.mu, C.dmu)) * tanh(10.0 * (lim - Fpt)) +
                       0.5 * (par<T>(C.kf, C.dkf) * dl + th * comp * par<T>(C.mu, C.dmu))) - par<T>(C.df, C.ddf) * uvdc<T>(C, k, c);
        T tl = tanh(10.0 * (lim - FNS));
        T fpos = (-0.5 * (par<T>(C.kf, C.dkf) - (1.0 - pow(th, 2.0)) * par<T>(C.mu, C.dmu) * comp - th * par<T>(C.mu, C.dmu) * cEE)) * tl -
                      (0.5 * (par<T>(C.kf, C.dkf) * dl - th * par<T>(C.mu, C.dmu) * comp)) * (1.0 - pow(tl, 2.0)) * (S.KxEE[c] - par<T>(C.mu, C.dmu) * cEE) -
                      (0.5 * (par<T>(C.kf, C.dkf) + (1.0 - pow(th, 2.0)) * par<T>(C.mu, C.dmu) * comp + th * par<T>(C.mu, C.dmu) * cEE));
        T tq = tanh(10.0 * lim - FNS);   // (10*fric_lim - F)[c] (Q3)
        T fvel = (-0.5 * (par<T>(C.kf, C.dkf) * dl - th * par<T>(C.mu, C.dmu) * comp)) * (1.0 - pow(tq, 2.0)) * S.KvEE[c] - par<T>(C.df, C.ddf);
        S.addF(3 + c, fric);
        S.EE(KIND_KX, k, k, c, fpos);
        S.EE(KIND_KV, k, k, c, fvel);
    }
    return comp;
}

}  // namespace ys

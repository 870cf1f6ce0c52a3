// adjoint.hpp -- reverse-mode (adjoint) kernels of the implicit step, gfx950.  Included
// by yarnstep.hip after the assembly, shear, update and normals kernels.
//
// The reference differentiates a rollout with LibTorch autograd: SimStep builds
// b = h (F + f_ext - K_v^T v) + M v from torch ops and solves A x = b with EigenSolver,
// whose backward is a second CG solve and -lambda x^T for A (src/TrainHelper.cpp:15-54,
// src/Solvers.h:35-52, src/Solvers.cpp:33-58).  Here one step of the backward pass is:
//
//   1. the normals' adjoint: the derivative of the smallest eigenvector of each
//      crossing's scatter matrix, transposed (k_adj_nrm), at the state after the step;
//   2. the update's adjoint (Cloth::update, src/Cloth.cpp:111-148): the adjoint of the
//      solution g and of is_slide, and what the state carries to the step before
//      (k_adj_update);
//   3. the backward solve: A lambda = g (the reference's solve with A,
//      src/Solvers.cpp:51-53) or A^T lambda = g (the exact adjoint);
//   4. the vector-Jacobian product of the assembly: with
//      Phi = lambda . b  (- lambda . A xs in the exact mode, xs the step's solution),
//      dPhi/dz for every state input z and material parameter.
//
// Step 4 reuses the element code of the assembly in dual numbers: a node's thread
// evaluates every element that touches the node with the tangent seeded on one of the
// node's own inputs (tan_at: the seed sits in the thread's LDS slot), through AdjSink,
// which adds lambda-weighted tangents of every contribution the element emits to any
// node's rows.  The friction's inputs (compression sums, E force before friction, its
// thread partial, the E-E self entries) are sums over a node's elements; their adjoints
// (ybar) come first from the friction's own dual passes (k_adj_fric) and weigh those
// contributions.  The shear reads the crossing's compression, so its adjoint (comp-bar,
// k_adj_shear) comes before the friction's.  Parameter derivatives take only the rows
// of the thread's own node, so that every element counts once.

namespace {
using namespace ys;

// per-node state adjoint planes (the tangent planes' order): x 3, v 3, xfix 3, uv 2,
// uvd 2, uvbar 2, nrm 3
enum AdjPlane { G_X = 0, G_V = 3, G_XFIX = 6, G_UV = 9, G_UVD = 11, G_UVBAR = 13, G_NRM = 15, G_PLANES = 18 };

struct AdjW {
    const double* lam;    // padded 5-plane vector: the adjoint of b
    const double* xs;     // padded 5-plane vector: the step's solution (exact mode)
    const double* ybar;   // FRIC_IN planes of N: adjoints of the friction's inputs
    int Wp, P;
    double h;
};

// AdjSink: the assembly's sink interface (FL, FE, LL, LE, EL, EE, SBCL, SBCLE, addF and
// the friction's accumulators), adding to acc the tangent of
//   sum over rows r of lambda_r b_r  (- lambda_r (A xs)_r, exact mode)
//   + sum over friction inputs y of ybar_y y,
// over every row the element writes (own = false) or only node me's rows (own = true).
// REF: the reference's gradient (A, M and K_v held constant: M v and K_v^T v take their
// tangent through v only, and A contributes nothing).
template <bool REF>
struct AdjSink {
    const ClothDev* C;
    AdjW Wt;
    int me;
    bool own, fric;
    double coef[3];
    double acc;
    // the friction's accumulated inputs (k_adj_fric)
    dd F[5];
    v3<dd> SBC_L, SBC_LE[2];
    dd FEpart[2], KxEE[2], KvEE[2];

    YS_DEV void init(const ClothDev* c, const AdjW& w, int k, bool own_rows) {
        C = c; Wt = w; me = k; own = own_rows; fric = false;
        coef[0] = 1.0; coef[1] = -(w.h * w.h); coef[2] = -w.h;
        acc = 0.0;
    }
    YS_DEV bool row(int k) const { return !own || k == me; }
    YS_DEV size_t pk(int k) const { return (size_t)padidx(Wt.Wp, k % C->W, k / C->W); }
    YS_DEV double lam(int k, int c) const { return Wt.lam[(size_t)c * Wt.P + pk(k)]; }
    YS_DEV double xs(int k, int c) const { return Wt.xs[(size_t)c * Wt.P + pk(k)]; }
    YS_DEV double yb(int k, int i) const { return Wt.ybar[(size_t)i * C->N + k]; }
    YS_DEV dd velL(int k, int c) const { return par<dd>(C->v[c * C->N + k], tan_at(C->dv, C->N, k, c)); }
    YS_DEV dd velE(int k, int c) const { return uvdc<dd>(*C, k, c); }
    // tangent of a block entry times a velocity in b (the block held constant in REF)
    YS_DEV static double bprod(dd B, dd x) { return REF ? B.v * x.d : B.d * x.v + B.v * x.d; }
    YS_DEV int fric_owner(int k) const { return omp_owner(C->Nc, crossing(*C, k % C->W, k / C->W)); }

    YS_DEV void FL(int k, v3<dd> f) {
        if (!row(k)) return;
        acc += Wt.h * (lam(k, 0) * f.x.d + lam(k, 1) * f.y.d + lam(k, 2) * f.z.d);
    }
    YS_DEV void FE(int k, int c, dd f, int elem_owner) {
        if (!row(k)) return;
        double w = Wt.h * lam(k, 3 + c) + yb(k, FI_FNS + c);
        if (elem_owner == fric_owner(k)) w += yb(k, FI_FEP + c);
        acc += w * f.d;
    }
    YS_DEV void addF(int i, dd f) { acc += Wt.h * lam(me, i) * f.d; }
    YS_DEV void SBCL(int k, v3<dd> f) {
        if (!row(k)) return;
        acc += yb(k, FI_SBCL) * f.x.d + yb(k, FI_SBCL + 1) * f.y.d + yb(k, FI_SBCL + 2) * f.z.d;
    }
    YS_DEV void SBCLE(int k, int c, v3<dd> f) {
        if (!row(k)) return;
        const int o = FI_SBCLE + 3 * c;
        acc += yb(k, o) * f.x.d + yb(k, o + 1) * f.y.d + yb(k, o + 2) * f.z.d;
    }
    // LL block B of (force node F, dof node D): A[3F+v, 3D+u] += coef B[u][v]
    // (src/MatrixMp.cpp:65-76); M: b[3F+v] += B[u][v] v_D[u]; K_v: b[3D+u] -= h B[u][v] v_F[v]
    YS_DEV void LL(int kind, int kF, int kD, const mat3<dd>& B) {
        if (row(kF))
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) {
                    if (!REF) acc -= coef[kind] * lam(kF, v) * B.m[u][v].d * xs(kD, u);
                    if (kind == KIND_M) acc += lam(kF, v) * bprod(B.m[u][v], velL(kD, u));
                }
        if (kind == KIND_KV && row(kD))
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) acc -= Wt.h * lam(kD, u) * bprod(B.m[u][v], velL(kF, v));
    }
    // LE: A[3kL+v, E(kE, c)] += coef B[v]
    YS_DEV void LE(int kind, int kL, int kE, int c, v3<dd> B) {
        const dd b[3] = {B.x, B.y, B.z};
        if (row(kL))
            for (int v = 0; v < 3; ++v) {
                if (!REF) acc -= coef[kind] * lam(kL, v) * b[v].d * xs(kE, 3 + c);
                if (kind == KIND_M) acc += lam(kL, v) * bprod(b[v], velE(kE, c));
            }
        if (kind == KIND_KV && row(kE))
            for (int v = 0; v < 3; ++v) acc -= Wt.h * lam(kE, 3 + c) * bprod(b[v], velL(kL, v));
    }
    // EL: A[E(kE, c), 3kL+u] += coef B[u]
    YS_DEV void EL(int kind, int kL, int kE, int c, v3<dd> B) {
        const dd b[3] = {B.x, B.y, B.z};
        if (row(kE))
            for (int u = 0; u < 3; ++u) {
                if (!REF) acc -= coef[kind] * lam(kE, 3 + c) * b[u].d * xs(kL, u);
                if (kind == KIND_M) acc += lam(kE, 3 + c) * bprod(b[u], velL(kL, u));
            }
        if (kind == KIND_KV && row(kL))
            for (int u = 0; u < 3; ++u) acc -= Wt.h * lam(kL, u) * bprod(b[u], velE(kE, c));
    }
    // EE: A[E(kF, c), E(kD, c)] += coef v; the self entries of K_x and K_v also feed the
    // node's friction (not the friction's own entries, which come after it)
    YS_DEV void EE(int kind, int kF, int kD, int c, dd v) {
        if (row(kF)) {
            if (!REF) acc -= coef[kind] * lam(kF, 3 + c) * v.d * xs(kD, 3 + c);
            if (kind == KIND_M) acc += lam(kF, 3 + c) * bprod(v, velE(kD, c));
            // (REF: the friction reads these entries only into A's E-E entries and, for K_v,
            // into K_v^T v as a constant block, so their adjoints are zero)
            if (!REF && kD == kF && !fric) {
                if (kind == KIND_KX) acc += yb(kF, FI_KXEE + c) * v.d;
                if (kind == KIND_KV) acc += yb(kF, FI_KVEE + c) * v.d;
            }
        }
        if (kind == KIND_KV && row(kD)) acc -= Wt.h * lam(kD, 3 + c) * bprod(v, velE(kF, c));
    }
};

// Output of the adjoint kernels: the step's state adjoint (G_PLANES planes of N, added
// to), the per-node parameter partials (YC_NPARAM planes of N, added to), and the
// intermediate adjoints of the compression (N) and of the friction inputs (ybar).
struct AdjOut {
    double* g;
    double* gpar;
    double* gcomp;
    double* ybar;
};

__device__ __forceinline__ void add_g(const AdjOut& O, int N, int plane, int k, double x) {
    O.g[(size_t)plane * N + k] += x;
}

// The node-gather bodies below take the tile of their block: one 64-thread block per 8 x 8
// tile (tile_node), as the single-cloth kernels at the end of this file launch them, or one
// block per tile of every cloth of a batch (batch_adjoint.hpp).

// ---------------------------------------------------------------------------
// Elements other than the shear (handle, edges, bends with their crimp twins, faces):
// the state passes.  Pass p seeds one input of the thread's node: x (3), v (3), uv and
// u-dot v-dot (both components at once: an element along yarn c reads component c only,
// so each element's tangent belongs to its own component), normals (3, the crimp twins
// only), xfix (3, the handle only).
// ---------------------------------------------------------------------------
constexpr int ADJ_STATE_PASSES = 14;

template <bool REF>
__device__ __forceinline__ void adj_elems_tile(const ClothDev& C, const AdjW& Wt, const AdjOut& O, int tile) {
    const int k = tile_node(C, tile, threadIdx.x);
    if (k < 0) return;
    const int u = k % C.W, v = k / C.W;
    const bool wind = C.wind_density != 0.0 || C.wind_drag != 0.0;
    const int ne_warp = C.Lg * (C.W - 1), nb_warp = C.Lg * (C.W - 2);
    AdjSink<REF> S;
    S.init(&C, Wt, k, false);
#pragma unroll 1
    for (int pass = 0; pass < ADJ_STATE_PASSES; ++pass) {
        int field, comp, plane;
        if (pass < 3) { field = SEED_X; comp = pass; plane = G_X + pass; }
        else if (pass < 6) { field = SEED_V; comp = pass - 3; plane = G_V + pass - 3; }
        else if (pass == 6) { field = SEED_UV; comp = -1; plane = G_UV; }
        else if (pass == 7) { field = SEED_UVD; comp = -1; plane = G_UVD; }
        else if (pass < 11) { field = SEED_NRM; comp = pass - 8; plane = G_NRM + pass - 8; }
        else { field = SEED_XFIX; comp = pass - 11; plane = G_XFIX + pass - 11; }
        if (field == SEED_NRM && !interior(C, u, v)) continue;   // boundary normals are 0 and unread
        if (field == SEED_XFIX && !(C.flags[k] & 2)) continue;
        set_seed(k, field, comp < 0 ? 3 : 1 << comp);
        double out[2] = {0.0, 0.0};   // per yarn direction for the uv and uvd passes
        const bool by_dir = comp < 0;
        if ((field == SEED_X || field == SEED_XFIX) && (C.flags[k] & 2)) {
            S.acc = 0.0;
            handle_element<dd>(C, S, k);
            out[0] += S.acc;
        }
        if (field == SEED_X || field == SEED_V || field == SEED_UV || field == SEED_UVD) {
#pragma unroll 1
            for (int ei = 0; ei < 4; ++ei) {
                int eu, ev, ec, idx;
                if (ei == 0) { if (u == 0) continue; eu = u - 1; ev = v; ec = 0; idx = v * (C.W - 1) + u - 1; }
                else if (ei == 1) { if (u == C.W - 1) continue; eu = u; ev = v; ec = 0; idx = v * (C.W - 1) + u; }
                else if (ei == 2) { if (v == 0) continue; eu = u; ev = v - 1; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v - 1; }
                else { if (v == C.Lg - 1) continue; eu = u; ev = v; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v; }
                S.acc = 0.0;
                edge_element<dd>(C, S, eu, ev, ec, ec == 0 ? (v & 1) : (u & 1), omp_owner(C.n_edges, idx));
                out[by_dir ? ec : 0] += S.acc;
            }
        }
        if (field == SEED_X || field == SEED_UV || field == SEED_NRM) {
            const int parts = field == SEED_NRM ? 2 : (field == SEED_UV ? 1 : 3);
#pragma unroll 1
            for (int bi6 = 0; bi6 < 6; ++bi6) {
                int d = bi6 % 3 - 1, dir = bi6 / 3;
                int uc = dir == 0 ? u + d : u, vc = dir == 0 ? v : v + d;
                if (dir == 0 && (uc < 1 || uc > C.W - 2)) continue;
                if (dir == 1 && (vc < 1 || vc > C.Lg - 2)) continue;
                if (field == SEED_UV && d == 0) continue;   // the bend's rest length reads its ends only
                int bi = dir == 0 ? v * (C.W - 2) + uc - 1 : nb_warp + u * (C.Lg - 2) + vc - 1;
                S.acc = 0.0;
                bend_element<dd>(C, S, uc, vc, dir, dir == 0 ? (v & 1) : (u & 1), omp_owner(C.n_bends, bi), parts);
                out[by_dir ? dir : 0] += S.acc;
            }
        }
        if ((field == SEED_X || field == SEED_V) && wind) {
#pragma unroll 1
            for (int fi = 0; fi < 8; ++fi) {
                int cu = u - 1 + ((fi >> 2) & 1), cv = v - 1 + ((fi >> 1) & 1), q = fi & 1;
                if (cu < 0 || cv < 0 || cu > C.W - 2 || cv > C.Lg - 2) continue;
                int tl = cu + cv * C.W, tr = tl + 1, bl = tl + C.W, br = bl + 1;
                int fa, fb, fc;
                if (((cu + cv) & 1) == 0) { if (q == 0) { fa = tl; fb = tr; fc = br; } else { fa = tl; fb = br; fc = bl; } }
                else { if (q == 0) { fa = tl; fb = tr; fc = bl; } else { fa = tr; fb = br; fc = bl; } }
                if (fa != k && fb != k && fc != k) continue;
                S.acc = 0.0;
                face_element<dd>(C, S, fa, fb, fc);
                out[0] += S.acc;
            }
        }
        if (by_dir) {   // boundary nodes too: an edge's rest length reads its ends' uv and u-dot v-dot
            add_g(O, C.N, plane, k, out[0]);
            add_g(O, C.N, plane + 1, k, out[1]);
        } else {
            add_g(O, C.N, plane, k, out[0]);
        }
    }
}

// The parameter passes of the same elements: C carries unit tangents of the parameters
// of one pass (rho, kb and the handle stiffness; or ks; or kc), and each element's
// tangent goes to its own parameter, yarn by yarn.  Own rows only.
template <bool REF>
__device__ __forceinline__ void adj_elems_par_tile(const ClothDev& C, const AdjW& Wt, const AdjOut& O, int pass, int tile) {
    const int k = tile_node(C, tile, threadIdx.x);
    if (k < 0) return;
    const int u = k % C.W, v = k / C.W;
    const int ne_warp = C.Lg * (C.W - 1), nb_warp = C.Lg * (C.W - 2);
    set_seed(-1, 0, 0);
    AdjSink<REF> S;
    S.init(&C, Wt, k, true);
    double gp[YC_NPARAM];
    for (int i = 0; i < YC_NPARAM; ++i) gp[i] = 0.0;
    // pass 0: rho (edges), kb (bends), handle_k; pass 1: ks (edges); pass 2: kc (edges)
    if (pass == 0 && (C.flags[k] & 2)) {
        S.acc = 0.0;
        handle_element<dd>(C, S, k);
        gp[YC_P_HANDLE_K] += S.acc;
    }
#pragma unroll 1
    for (int ei = 0; ei < 4; ++ei) {
        int eu, ev, ec, idx;
        if (ei == 0) { if (u == 0) continue; eu = u - 1; ev = v; ec = 0; idx = v * (C.W - 1) + u - 1; }
        else if (ei == 1) { if (u == C.W - 1) continue; eu = u; ev = v; ec = 0; idx = v * (C.W - 1) + u; }
        else if (ei == 2) { if (v == 0) continue; eu = u; ev = v - 1; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v - 1; }
        else { if (v == C.Lg - 1) continue; eu = u; ev = v; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v; }
        const int yarn = ec == 0 ? (v & 1) : (u & 1);
        S.acc = 0.0;
        edge_element<dd>(C, S, eu, ev, ec, yarn, omp_owner(C.n_edges, idx));
        if (pass == 0) gp[yarn == 0 ? YC_P_RHO0 : YC_P_RHO1] += S.acc;
        else if (pass == 1) gp[yarn == 0 ? YC_P_Y0 : YC_P_Y1] += S.acc;
        else gp[YC_P_KC] += S.acc;
    }
    if (pass == 0) {
#pragma unroll 1
        for (int bi6 = 0; bi6 < 6; ++bi6) {
            int d = bi6 % 3 - 1, dir = bi6 / 3;
            int uc = dir == 0 ? u + d : u, vc = dir == 0 ? v : v + d;
            if (dir == 0 && (uc < 1 || uc > C.W - 2)) continue;
            if (dir == 1 && (vc < 1 || vc > C.Lg - 2)) continue;
            int bi = dir == 0 ? v * (C.W - 2) + uc - 1 : nb_warp + u * (C.Lg - 2) + vc - 1;
            const int yarn = dir == 0 ? (v & 1) : (u & 1);
            S.acc = 0.0;
            bend_element<dd>(C, S, uc, vc, dir, yarn, omp_owner(C.n_bends, bi));
            gp[yarn == 0 ? YC_P_B0 : YC_P_B1] += S.acc;
        }
    }
    for (int i = 0; i < YC_NPARAM; ++i)
        if (gp[i] != 0.0) O.gpar[(size_t)i * C.N + k] += gp[i];
}

// ---------------------------------------------------------------------------
// Shear segments touching the node (assemble2_node's enumeration).  State passes: x of
// the node (3, every segment that holds it) and the compression of its crossing (the
// four segments centred on it, into gcomp); parameter pass (C.dS = 1, own rows): S.
// ---------------------------------------------------------------------------
template <class Fn>
__device__ __forceinline__ void for_shear_segments(const ClothDev& C, int k, Fn fn) {
    const int u = k % C.W, v = k / C.W;
    auto K = [&](int uu, int vv) { return uu + vv * C.W; };
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
            fn(kc, k1, k3);
        }
    }
}

template <bool REF, class Sink>
__device__ __forceinline__ void shear_adj_segment(const ClothDev& C, Sink& S, const double* comp, int kc, int k1, int k3) {
    ShearOut<dd> o;
    shear_compute<dd>(C, kc, k1, k3, comp, seed_ptr(SEED_COMP), o);
    const int kk[3] = {kc, k1, k3};
    for (int r = 0; r < 3; ++r) emit_shear_role<dd>(S, kk, r, o.F[r], o.B[r]);
}

template <bool REF>
__device__ __forceinline__ void adj_shear_tile(const ClothDev& C, const AdjW& Wt, const AdjOut& O, const double* comp, int tile) {
    const int k = tile_node(C, tile, threadIdx.x);
    if (k < 0) return;
    AdjSink<REF> S;
    S.init(&C, Wt, k, false);
    const bool inter = interior(C, k % C.W, k / C.W);
#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
        if (pass == 3 && !inter) break;
        set_seed(k, pass < 3 ? SEED_X : SEED_COMP, pass < 3 ? 1 << pass : 1);
        S.acc = 0.0;
        for_shear_segments(C, k, [&](int kc, int k1, int k3) {
            if (pass == 3 && kc != k) return;
            shear_adj_segment<REF>(C, S, comp, kc, k1, k3);
        });
        if (pass < 3) add_g(O, C.N, G_X + pass, k, S.acc);
        else O.gcomp[k] = S.acc;
    }
}

// The shear's S derivative without dual numbers: every force and block of a segment is S
// times a function of the geometry and the compression (S enters only through kx and its
// jamming-branch derivatives, src/PhysicsCloth.cpp:2654, 2694-2703, 2747, 2974-2996, each
// with one factor S), so d/dS of the own rows' Phi is Phi evaluated with S = 1, which the
// host puts into C.  (A dual-number form of this pass, compiled at -O2 or -O3, returned values
// that differ from its host build and from -O1 on the same inputs, while the other
// exact-mode passes matched; this form needs a fraction of the registers.)
template <bool REF>
__device__ __forceinline__ void adj_shear_par_tile(const ClothDev& C, const AdjW& Wt, const AdjOut& O, const double* comp, int tile) {
    const int k = tile_node(C, tile, threadIdx.x);
    if (k < 0) return;
    const int W = C.W;
    auto lam = [&](int q, int c) { return Wt.lam[(size_t)c * Wt.P + padidx(Wt.Wp, q % W, q / W)]; };
    auto xs = [&](int q, int c) { return Wt.xs[(size_t)c * Wt.P + padidx(Wt.Wp, q % W, q / W)]; };
    const double cx = -(Wt.h * Wt.h);   // K_x's coefficient in A
    const double l0 = lam(k, 0), l1 = lam(k, 1), l2 = lam(k, 2);
    double acc = 0.0;
    for_shear_segments(C, k, [&](int kc, int k1, int k3) {
        ShearOut<double> o;
        shear_compute<double>(C, kc, k1, k3, comp, nullptr, o);
        const int kk[3] = {kc, k1, k3};
        const int r = kc == k ? 0 : (k1 == k ? 1 : 2);
        const d3 F = r == 0 ? o.F[0] : (r == 1 ? o.F[1] : o.F[2]);
        acc += Wt.h * (l0 * F.x + l1 * F.y + l2 * F.z);
        if (!REF)
            for (int j = 0; j < 3; ++j) {
                // row node k, column node kk[j]: A[3k+v, 3kk[j]+u] += cx B[u][v]
                const m33& B = r == 0 ? o.B[0][j] : (r == 1 ? o.B[1][j] : o.B[2][j]);
                const double x0 = xs(kk[j], 0), x1 = xs(kk[j], 1), x2 = xs(kk[j], 2);
                const double xu[3] = {x0, x1, x2};
                for (int u = 0; u < 3; ++u) acc -= cx * (l0 * B.m[u][0] + l1 * B.m[u][1] + l2 * B.m[u][2]) * xu[u];
            }
    });
    O.gpar[(size_t)YC_P_S * C.N + k] += acc;
}

// ---------------------------------------------------------------------------
// Friction at the node's crossing (src/PhysicsCloth.cpp:2579-2641) as a function of its
// accumulated inputs (read back from the assembly's fric_in), its normal, uv, uvbar,
// u-dot v-dot and mu, kf, df.  Its outputs: the friction force (into b), its E-E entries
// (into A, exact mode; the K_v one also into b through K_v^T v), the compression (read by
// the shear: weight gcomp) and is_slide (read by the update: weight gslide).
// State passes: the FRIC_IN inputs into ybar, then n (3), uv (2), uvbar (2), uvd (2).
// Parameter launch (fpar 0, 1, 2: C carries the unit tangent of mu, kf or df).
// ---------------------------------------------------------------------------
template <bool REF>
__device__ __forceinline__ double fric_adj_pass(const ClothDev& C, AdjSink<REF>& S, const double* fi, int seed_in, int k,
                                                double gcomp, const double* gsl) {
    const int N = C.N;
    auto in = [&](int i) { return dd(fi[(size_t)i * N + k], i == seed_in ? 1.0 : 0.0); };
    S.SBC_L = V3<dd>(in(FI_SBCL), in(FI_SBCL + 1), in(FI_SBCL + 2));
    for (int c = 0; c < 2; ++c) {
        S.SBC_LE[c] = V3<dd>(in(FI_SBCLE + 3 * c), in(FI_SBCLE + 3 * c + 1), in(FI_SBCLE + 3 * c + 2));
        S.F[3 + c] = in(FI_FNS + c);
        S.FEpart[c] = in(FI_FEP + c);
        S.KxEE[c] = in(FI_KXEE + c);
        S.KvEE[c] = in(FI_KVEE + c);
    }
    S.acc = 0.0;
    dd sl[2];
    dd comp = friction<dd>(C, S, sl);
    return S.acc + gcomp * comp.d + gsl[0] * sl[0].d + gsl[1] * sl[1].d;
}

template <bool REF>
__device__ __forceinline__ void adj_fric_tile(const ClothDev& C, const AdjW& Wt, const AdjOut& O, const double* fric_in,
                                              const double* gslide, int fpar, int tile) {
    const int k = tile_node(C, tile, threadIdx.x);
    if (k < 0) return;
    const int N = C.N;
    if (!interior(C, k % C.W, k / C.W)) {
        if (fpar < 0) for (int i = 0; i < FRIC_IN; ++i) O.ybar[(size_t)i * N + k] = 0.0;
        return;
    }
    AdjSink<REF> S;
    S.init(&C, Wt, k, true);
    S.fric = true;
    const double gc = O.gcomp[k];
    const double gsl[2] = {gslide[k], gslide[N + k]};
    if (fpar >= 0) {   // mu, kf or df: the tangent is in C
        set_seed(-1, 0, 0);
        const int p = fpar == 0 ? YC_P_MU : (fpar == 1 ? YC_P_KF : YC_P_DF);
        O.gpar[(size_t)p * N + k] += fric_adj_pass<REF>(C, S, fric_in, -1, k, gc, gsl);
        return;
    }
#pragma unroll 1
    for (int i = 0; i < FRIC_IN; ++i) {
        set_seed(-1, 0, 0);
        O.ybar[(size_t)i * N + k] = fric_adj_pass<REF>(C, S, fric_in, i, k, gc, gsl);
    }
#pragma unroll 1
    for (int pass = 0; pass < 9; ++pass) {
        int field, comp, plane;
        if (pass < 3) { field = SEED_NRM; comp = pass; plane = G_NRM + pass; }
        else if (pass < 5) { field = SEED_UV; comp = pass - 3; plane = G_UV + pass - 3; }
        else if (pass < 7) { field = SEED_UVBAR; comp = pass - 5; plane = G_UVBAR + pass - 5; }
        else { field = SEED_UVD; comp = pass - 7; plane = G_UVD + pass - 7; }
        set_seed(k, field, 1 << comp);
        add_g(O, N, plane, k, fric_adj_pass<REF>(C, S, fric_in, -1, k, gc, gsl));
    }
}

// ---------------------------------------------------------------------------
// The normals' adjoint at state x (the state the normals were computed from): for each
// interior node k, with n = sg e_min the signed unit eigenvector of the smallest
// eigenvalue of C = sum_i d_i d_i^T over its 5-point cross (d_i = p_i - mean),
//   dn = sg sum_{j != min} e_j (e_j^T dC e_min) / (lam_min - lam_j)   (k_normals_tan),
// so n-bar . dn = <Cbar, dC> with Cbar = sg sum_j (n-bar . e_j) / (lam_min - lam_j) e_j e_min^T,
// d-bar_i = (Cbar + Cbar^T) d_i, and p-bar_i = d-bar_i - mean(d-bar).  Written per cross
// point: pbar[(i * 3 + c) N + k], i over (k, k-1, k+1, k-W, k+W).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adj_nrm_node(const StateDev& s, const double* gnrm, double* pbar, int k) {
    if (k >= s.N) return;
    const int N = s.N, u = k % s.W, v = k / s.W;
    if (!(u > 0 && u < s.W - 1 && v > 0 && v < s.Lg - 1)) {
        for (int i = 0; i < 15; ++i) pbar[(size_t)i * N + k] = 0.0;
        return;
    }
    NodeFrame f;
    node_frame(s, k, f);
    double E[3][3];
    for (int j = 0; j < 3; ++j) {
        double nn = sqrt(f.V[0][j] * f.V[0][j] + f.V[1][j] * f.V[1][j] + f.V[2][j] * f.V[2][j]);
        for (int r = 0; r < 3; ++r) E[r][j] = f.V[r][j] / nn;
    }
    const int im = f.im;
    const double em[3] = {frame_col(E, 0, im), frame_col(E, 1, im), frame_col(E, 2, im)};
    const double lam_min = im == 0 ? f.lam[0] : (im == 1 ? f.lam[1] : f.lam[2]);
    const double gn[3] = {gnrm[k], gnrm[N + k], gnrm[2 * N + k]};
    double Cb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int j = 0; j < 3; ++j) {
        if (j == im) continue;
        const double w = f.sg * (gn[0] * E[0][j] + gn[1] * E[1][j] + gn[2] * E[2][j]) / (lam_min - f.lam[j]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) Cb[a][b] += w * E[a][j] * em[b];
    }
    double db[5][3], mb[3] = {0, 0, 0};
    for (int i = 0; i < 5; ++i)
        for (int a = 0; a < 3; ++a) {
            double x = 0.0;
            for (int b = 0; b < 3; ++b) x += (Cb[a][b] + Cb[b][a]) * f.d[i][b];
            db[i][a] = x;
            mb[a] += x;
        }
    for (int i = 0; i < 5; ++i)
        for (int a = 0; a < 3; ++a) pbar[(size_t)(i * 3 + a) * N + k] = db[i][a] - mb[a] * (1.0 / 5.0);
}

// ---------------------------------------------------------------------------
// The update's adjoint (Cloth::update, src/Cloth.cpp:111-148, as k_update): with gin the
// adjoint of the state after the step and pbar its normals' adjoint per cross point,
//   x-bar'  = gin.x + sum of pbar over the crosses that hold the node + gin.xfix (off the handles)
//   g_L     = gin.v + h x-bar'                       (the adjoint of the solution's L rows)
//   g_E     = gin.uvd + h gin.uv + h is_slide gin.uvbar   (interior crossings)
//   gslide  = h sol_E gin.uvbar
// and the part the state carries to the step before (x, the handles' xfix, uv, uvbar;
// u-dot v-dot on the boundary, where the update leaves them) goes to gout, which the
// assembly's adjoint then adds to.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void adj_update_node(const unsigned char* flags, int W, int Lg, const double* gin, const double* pbar,
                                                const double* is_slide, const double* sol, int Wp, int P, double h, double* rhs,
                                                double* gslide, double* gout, int k) {
    const int N = W * Lg;
    if (k >= N) return;
    const int u = k % W, v = k / W;
    const size_t pk = (size_t)padidx(Wp, u, v);
    const bool fixed = (flags[k] & 2) != 0;
    const bool inter = u > 0 && u < W - 1 && v > 0 && v < Lg - 1;
    for (int c = 0; c < 3; ++c) {
        double xb = gin[(size_t)(G_X + c) * N + k];
        // node k is point 0 of its own cross, point 1 (k-1) of k+1's, 2 (k+1) of k-1's,
        // 3 (k-W) of k+W's, 4 (k+W) of k-W's; pbar is zero for crosses off the interior
        xb += pbar[(size_t)(0 * 3 + c) * N + k];
        if (u + 1 < W) xb += pbar[(size_t)(1 * 3 + c) * N + k + 1];
        if (u > 0) xb += pbar[(size_t)(2 * 3 + c) * N + k - 1];
        if (v + 1 < Lg) xb += pbar[(size_t)(3 * 3 + c) * N + k + W];
        if (v > 0) xb += pbar[(size_t)(4 * 3 + c) * N + k - W];
        const double gf = gin[(size_t)(G_XFIX + c) * N + k];
        if (!fixed) xb += gf;
        rhs[(size_t)c * P + pk] = gin[(size_t)(G_V + c) * N + k] + h * xb;
        gout[(size_t)(G_X + c) * N + k] = xb;
        gout[(size_t)(G_V + c) * N + k] = 0.0;
        gout[(size_t)(G_XFIX + c) * N + k] = fixed ? gf : 0.0;
        gout[(size_t)(G_NRM + c) * N + k] = 0.0;
    }
    for (int c = 0; c < 2; ++c) {
        const double guv = gin[(size_t)(G_UV + c) * N + k], gud = gin[(size_t)(G_UVD + c) * N + k],
                     gub = gin[(size_t)(G_UVBAR + c) * N + k];
        gout[(size_t)(G_UV + c) * N + k] = guv;
        gout[(size_t)(G_UVBAR + c) * N + k] = gub;
        if (inter) {
            const double sl = is_slide[(size_t)c * N + k];
            rhs[(size_t)(3 + c) * P + pk] = gud + h * guv + h * sl * gub;
            gslide[(size_t)c * N + k] = h * sol[(size_t)(3 + c) * P + pk] * gub;
            gout[(size_t)(G_UVD + c) * N + k] = 0.0;
        } else {
            rhs[(size_t)(3 + c) * P + pk] = 0.0;
            gslide[(size_t)c * N + k] = 0.0;
            gout[(size_t)(G_UVD + c) * N + k] = gud;
        }
    }
}

// x-bar += pbar gathered over the crosses (the last normals' adjoint, for the initial
// state, whose normals yc_set_state computes from x)
__device__ __forceinline__ void adj_fold_nrm_node(int W, int Lg, const double* pbar, double* g, int k) {
    const int N = W * Lg;
    if (k >= N) return;
    const int u = k % W, v = k / W;
    for (int c = 0; c < 3; ++c) {
        double xb = pbar[(size_t)(0 * 3 + c) * N + k];
        if (u + 1 < W) xb += pbar[(size_t)(1 * 3 + c) * N + k + 1];
        if (u > 0) xb += pbar[(size_t)(2 * 3 + c) * N + k - 1];
        if (v + 1 < Lg) xb += pbar[(size_t)(3 * 3 + c) * N + k + W];
        if (v > 0) xb += pbar[(size_t)(4 * 3 + c) * N + k - W];
        g[(size_t)(G_X + c) * N + k] += xb;
    }
}

// The loss gradient of a state: g += gst (node-major (N,3) / (N,2) arrays in the
// yc_get_state order: lpos, lposfix, lvel, epos, eposbar, evel; 15 N doubles).
__global__ void __launch_bounds__(NBLK) k_adj_add_gstate(int N, const double* gst, double* g) {
    const int k = blockIdx.x * NBLK + threadIdx.x;
    if (k >= N) return;
    for (int c = 0; c < 3; ++c) {
        g[(size_t)(G_X + c) * N + k] += gst[3 * (size_t)k + c];
        g[(size_t)(G_XFIX + c) * N + k] += gst[3 * (size_t)N + 3 * (size_t)k + c];
        g[(size_t)(G_V + c) * N + k] += gst[6 * (size_t)N + 3 * (size_t)k + c];
    }
    for (int c = 0; c < 2; ++c) {
        g[(size_t)(G_UV + c) * N + k] += gst[9 * (size_t)N + 2 * (size_t)k + c];
        g[(size_t)(G_UVBAR + c) * N + k] += gst[11 * (size_t)N + 2 * (size_t)k + c];
        g[(size_t)(G_UVD + c) * N + k] += gst[13 * (size_t)N + 2 * (size_t)k + c];
    }
}

// The position loss of the reference's training (src/TrainHeter.h:211-226) on one state,
// against device targets (node-major (N,3) positions, (N,2) Eulerian coordinates):
// w (sum |x - X|^2 + sum |uv - U|^2), its gradient added to g, the node's terms to
// lpart (summed afterwards in a fixed order).
__device__ __forceinline__ void adj_target_loss_node(int N, const double* x, const double* uv, const double* X, const double* U,
                                                     double w, double* g, double* lpart, int k) {
    if (k >= N) return;
    double l = 0.0;
    for (int c = 0; c < 3; ++c) {
        const double d = x[(size_t)c * N + k] - X[3 * (size_t)k + c];
        l += d * d;
        g[(size_t)(G_X + c) * N + k] += 2.0 * w * d;
    }
    for (int c = 0; c < 2; ++c) {
        const double d = uv[(size_t)c * N + k] - U[2 * (size_t)k + c];
        l += d * d;
        g[(size_t)(G_UV + c) * N + k] += 2.0 * w * d;
    }
    lpart[k] += w * l;
}

// sum of each of `planes` planes of n doubles into out[plane], in a fixed order (one
// block per plane)
__global__ void __launch_bounds__(256) k_plane_sums(const double* a, int n, double* out) {
    __shared__ double s[256];
    const double* p = a + (size_t)blockIdx.x * n;
    double x = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) x += p[i];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// The 18 state planes of a context (x, v, xfix, uv, uvd, uvbar, nrm) to and from one
// contiguous checkpoint block.
__device__ __forceinline__ void ckpt_node(const StateDev& s, double* blk, int save, int k) {
    if (k >= s.N) return;
    const size_t N = s.N;
    double* planes[7] = {s.x, s.v, s.xfix, s.uv, s.uvd, s.uvbar, s.nrm};
    const int cnt[7] = {3, 3, 3, 2, 2, 2, 3};
    int q = 0;
    for (int f = 0; f < 7; ++f)
        for (int c = 0; c < cnt[f]; ++c, ++q) {
            if (save) blk[(size_t)q * N + k] = planes[f][(size_t)c * N + k];
            else planes[f][(size_t)c * N + k] = blk[(size_t)q * N + k];
        }
}

// ---------------------------------------------------------------------------
// Single-cloth kernels
// ---------------------------------------------------------------------------
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_adj_elems(ClothDev C, AdjW Wt, AdjOut O) { adj_elems_tile<REF>(C, Wt, O, blockIdx.x); }
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_adj_elems_par(ClothDev C, AdjW Wt, AdjOut O, int pass) {
    adj_elems_par_tile<REF>(C, Wt, O, pass, blockIdx.x);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_adj_shear(ClothDev C, AdjW Wt, AdjOut O, const double* comp) {
    adj_shear_tile<REF>(C, Wt, O, comp, blockIdx.x);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_adj_shear_par(ClothDev C, AdjW Wt, AdjOut O, const double* comp) {
    adj_shear_par_tile<REF>(C, Wt, O, comp, blockIdx.x);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_adj_fric(ClothDev C, AdjW Wt, AdjOut O, const double* fric_in, const double* gslide,
                                                     int fpar) {
    adj_fric_tile<REF>(C, Wt, O, fric_in, gslide, fpar, blockIdx.x);
}
__global__ void __launch_bounds__(NBLK) k_adj_nrm(StateDev s, const double* gnrm, double* pbar) {
    adj_nrm_node(s, gnrm, pbar, blockIdx.x * NBLK + threadIdx.x);
}
__global__ void __launch_bounds__(NBLK) k_adj_update(const unsigned char* flags, int W, int Lg, const double* gin,
                                                     const double* pbar, const double* is_slide, const double* sol, int Wp, int P,
                                                     double h, double* rhs, double* gslide, double* gout) {
    adj_update_node(flags, W, Lg, gin, pbar, is_slide, sol, Wp, P, h, rhs, gslide, gout, blockIdx.x * NBLK + threadIdx.x);
}
__global__ void __launch_bounds__(NBLK) k_adj_fold_nrm(int W, int Lg, const double* pbar, double* g) {
    adj_fold_nrm_node(W, Lg, pbar, g, blockIdx.x * NBLK + threadIdx.x);
}
__global__ void __launch_bounds__(NBLK) k_adj_target_loss(int N, const double* x, const double* uv, const double* X,
                                                          const double* U, double w, double* g, double* lpart) {
    adj_target_loss_node(N, x, uv, X, U, w, g, lpart, blockIdx.x * NBLK + threadIdx.x);
}
__global__ void __launch_bounds__(NBLK) k_ckpt(StateDev s, double* blk, int save) { ckpt_node(s, blk, save, blockIdx.x * NBLK + threadIdx.x); }

}  // namespace

// assemble1.hpp -- the first assembly kernel's work for one tile, shared by k_assemble1,
// its tangent form k_assemble1_dirs, the batched k_b_assemble1 and the host build of
// tests/asm_host.cpp.  Included after elements.hpp and assembly.hpp.
//
// T = double: the step's rows of A, F, b (src/PhysicsCloth.cpp:46-156,
// src/TrainHelper.cpp:19-20), two threads per node (blocks of 2 x 64 threads): half 0
// evaluates the warp-direction elements (handle, warp edges and bends), half 1 the
// weft-direction ones and the faces.  The two halves write disjoint slots of the node's
// rows except the L-L self block, which half 1 keeps apart; after a barrier half 0 adds
// it and half 1's accumulators in, then does the friction and b.  T = dd: their
// directional derivatives along the tangents held in C, into the d* outputs of O, one
// thread per node (blocks of 64 threads).  ASM1_REF: the reference gradient's tangent run.
#pragma once

namespace {
using namespace ys;

template <class T, bool ASM1_REF>
__device__ __forceinline__ void assemble1_block(const ClothDev& C, const AsmOut& O, double h, int tile) {
    constexpr bool tan = !std::is_same<T, double>::value;
    const int nl = threadIdx.x & (TILE_N - 1);              // node lane in the tile
    const int half = tan ? -1 : (int)(threadIdx.x >> 6);    // -1: every element (tangent run)
    const int k = tile_node(C, tile, nl);
    const bool active = k >= 0;
    RowSinkT<T, ASM1_REF> S;
    S.init(&C, active ? k : 0, tan ? nullptr : O.A, h, tan ? O.dA : nullptr, tan && O.detach);
    if constexpr (!tan) S.select_kind(O.only_kind);
    const int u = S.mu_, v = S.mv_;
    // value run: the node's 159 slots accumulate in LDS (the tile's slot planes, same
    // layout as in HBM) and go to HBM once at the end, instead of a global
    // read-modify-write per contribution; behind them half 1's self block (9 planes) and
    // accumulators (ASM1_ACC planes)
    extern __shared__ double lds_rows[];
    double* gA = S.A;
    double* self_w = lds_rows + (size_t)NSLOT * TILE_N;
    double* acc_w = self_w + (size_t)9 * TILE_N;
    if constexpr (!tan) {
        S.A = lds_rows + nl;
        for (int s = half; s < NSLOT; s += 2) S.A[(size_t)s * TILE_N] = 0.0;
        if (half == 1) {
            S.Aself = self_w + nl;
            for (int i = 0; i < 9; ++i) S.Aself[(size_t)i * TILE_N] = 0.0;
        }
        __syncthreads();
    } else {
        if (active && S.dA) for (int s = 0; s < NSLOT; ++s) S.dA[(size_t)s * TILE_N] = 0.0;
    }
    const bool warp_side = half <= 0, weft_side = half != 0;
    if (active) {
        // handle (src/PhysicsCloth.cpp:232-248)
        if (warp_side && (C.flags[k] & 2)) handle_element<T>(C, S, k);
        // edges (list order: warp rows, then weft columns; src/Cloth.cpp:190-228).  One
        // call site in a loop that is not unrolled: the inlined element code is large, and
        // four or six copies of it would not fit the instruction cache.
        int ne_warp = C.Lg * (C.W - 1);
    #pragma unroll 1
        for (int ei = warp_side ? 0 : 2; ei < (weft_side ? 4 : 2); ++ei) {
            int eu, ev, ec, idx;
            if (ei == 0) { if (u == 0) continue; eu = u - 1; ev = v; ec = 0; idx = v * (C.W - 1) + u - 1; }
            else if (ei == 1) { if (u == C.W - 1) continue; eu = u; ev = v; ec = 0; idx = v * (C.W - 1) + u; }
            else if (ei == 2) { if (v == 0) continue; eu = u; ev = v - 1; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v - 1; }
            else { if (v == C.Lg - 1) continue; eu = u; ev = v; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v; }
            // value run: the edge's entries in the node's rows summed in registers, then into
            // the slots (the tangent run has no registers to spare for the dual sums)
            const int uo = eu == u && ev == v ? eu + (ec == 0) : eu, vo = eu == u && ev == v ? ev + (ec == 1) : ev;
            if constexpr (!tan) S.edge_begin();
            edge_element<T>(C, S, eu, ev, ec, ec == 0 ? (v & 1) : (u & 1), omp_owner(C.n_edges, idx));
            if constexpr (!tan) S.edge_end(uo + vo * C.W, S.inter, S.inter && interior(C, uo, vo), ec);
        }
        // bends centred on u-1, u, u+1 (warp) and v-1, v, v+1 (weft) (src/Cloth.cpp:230-270):
        // a bend centred d after the node holds it as n2 (role 0), centred on it as n0
        // (role 1), centred d before it as n1 (role 2)
        int nb_warp = C.Lg * (C.W - 2);
    #pragma unroll 1
        for (int bi6 = warp_side ? 0 : 3; bi6 < (weft_side ? 6 : 3); ++bi6) {
            int d = bi6 % 3 - 1, dir = bi6 / 3;
            int uc = dir == 0 ? u + d : u, vc = dir == 0 ? v : v + d;
            if (dir == 0 && (uc < 1 || uc > C.W - 2)) continue;
            if (dir == 1 && (vc < 1 || vc > C.Lg - 2)) continue;
            int bi = dir == 0 ? v * (C.W - 2) + uc - 1 : nb_warp + u * (C.Lg - 2) + vc - 1;
            if constexpr (USE_BEND_REC && !tan) emit_bend_role(C, S, O.bend_rec, C.n_bends, bi, 1 - d, omp_owner(C.n_bends, bi));
            else bend_element<T>(C, S, uc, vc, dir, dir == 0 ? (v & 1) : (u & 1), omp_owner(C.n_bends, bi));
        }
        // faces of the up to four cells around the node (src/Cloth.cpp:272-286)
        if (weft_side && (C.wind_density != 0.0 || C.wind_drag != 0.0)) {
    #pragma unroll 1
            for (int fi = 0; fi < 8; ++fi) {
                int cu = u - 1 + ((fi >> 2) & 1), cv = v - 1 + ((fi >> 1) & 1), q = fi & 1;
                if (cu < 0 || cv < 0 || cu > C.W - 2 || cv > C.Lg - 2) continue;
                int tl = cu + cv * C.W, tr = tl + 1, bl = tl + C.W, br = bl + 1;
                int fa, fb, fc;
                if (((cu + cv) & 1) == 0) { if (q == 0) { fa = tl; fb = tr; fc = br; } else { fa = tl; fb = br; fc = bl; } }
                else { if (q == 0) { fa = tl; fb = tr; fc = bl; } else { fa = tr; fb = br; fc = bl; } }
                if (fa == k || fb == k || fc == k) face_element<T>(C, S, fa, fb, fc);
            }
        }
    }
    if constexpr (!tan) {
        // half 1 hands its accumulators over; half 0 adds them and the self block
        if (half == 1 && active) S.store_acc(acc_w + nl, TILE_N);
        __syncthreads();
        if (half == 0 && active) {
            S.add_acc(acc_w + nl, TILE_N);
            for (int rc = 0; rc < 3; ++rc)
                for (int cc = 0; cc < 3; ++cc)
                    S.A[(size_t)(rc * 45 + cc) * TILE_N] += self_w[(size_t)(rc * 3 + cc) * TILE_N + nl];
        }
    }
    if (active && half <= 0) {
        // friction at the node's own crossing (src/PhysicsCloth.cpp:114-123)
        T comp = T(0.0), sl[2] = {T(0.0), T(0.0)};
        if constexpr (!tan) if (O.fric_in) {
            const double fi[FRIC_IN] = {val(S.SBC_L.x), val(S.SBC_L.y), val(S.SBC_L.z),
                                        val(S.SBC_LE[0].x), val(S.SBC_LE[0].y), val(S.SBC_LE[0].z),
                                        val(S.SBC_LE[1].x), val(S.SBC_LE[1].y), val(S.SBC_LE[1].z),
                                        val(S.F[3]), val(S.F[4]), val(S.FEpart[0]), val(S.FEpart[1]),
                                        val(S.KxEE[0]), val(S.KxEE[1]), val(S.KvEE[0]), val(S.KvEE[1])};
            for (int i = 0; i < FRIC_IN; ++i) O.fric_in[(size_t)i * C.N + k] = S.inter ? fi[i] : 0.0;
        }
        if (S.inter) comp = friction<T>(C, S, sl);
        // b = h (F + f_ext - K_v^T v) + M v (src/TrainHelper.cpp:20, 63)
        int pk = padidx(O.Wp, u, v);
        int e = S.inter ? crossing(C, u, v) : 0;
        if constexpr (!tan) {
            for (int c = 0; c < 5; ++c) O.F[c * C.N + k] = val(S.F[c]);
            O.comp[k] = val(comp);
            O.is_slide[k] = val(sl[0]); O.is_slide[C.N + k] = val(sl[1]);
            O.fepart[k] = val(S.FEpart[0]); O.fepart[C.N + k] = val(S.FEpart[1]);
            for (int c = 0; c < 5; ++c) {
                double bv = 0.0;
                if (c < 3 || S.inter) {
                    double ext = 0.0;
                    if (O.ext) ext = c < 3 ? O.ext[3 * k + c] : O.ext[3 * C.N + 2 * e + (c - 3)];
                    bv = h * (val(S.F[c]) + ext - val(S.KvTv[c])) + val(S.Mv[c]);
                }
                O.b[c * O.P + pk] = bv;
            }
        } else {
            for (int c = 0; c < 5; ++c) O.dF[c * C.N + k] = tan_of(S.F[c]);
            O.dcomp[k] = tan_of(comp);
            O.dis_slide[k] = tan_of(sl[0]); O.dis_slide[C.N + k] = tan_of(sl[1]);
            for (int c = 0; c < 5; ++c) {
                double bv = 0.0;   // f_ext enters through its tangent, when one is set
                if (c < 3 || S.inter) {
                    double dext = 0.0;
                    if (O.dext) dext = c < 3 ? O.dext[3 * k + c] : O.dext[3 * C.N + 2 * e + (c - 3)];
                    bv = h * (tan_of(S.F[c]) + dext - tan_of(S.KvTv[c])) + tan_of(S.Mv[c]);
                }
                O.db[c * O.P + pk] = bv;
            }
        }
    }
    if constexpr (!tan) {
        // the rows to HBM, both halves, once the friction's entries are in, with the NaN/Inf
        // screen of the slots the shear replay does not touch (it screens its own)
        __syncthreads();
        if (active) {
            bool bad = false;
            for (int s = half; s < NSLOT; s += 2) {
                const double a = S.A[(size_t)s * TILE_N];
                gA[(size_t)s * TILE_N] = a;
                if (!isfinite(a)) bad = true;
            }
            if (bad) atomicOr(&O.flag[1], 1);
        }
    }
    if (active && S.bad) atomicOr(&O.flag[0], 1);
}

}  // namespace

// assemble1.hpp -- the first assembly kernel's work for one tile, shared by k_assemble1,
// its tangent form k_assemble1_dirs, the batched k_b_assemble1 and the host build of
// tests/asm_host.cpp.  Included after elements.hpp and assembly.hpp.
//
// T = double: the step's rows of A, F, b (src/PhysicsCloth.cpp:46-156,
// src/TrainHelper.cpp:19-20), four threads per node (blocks of 4 x 64 threads): group 0
// evaluates the warp edges, group 1 the warp bends, group 2 the weft edges and the faces,
// group 3 the weft bends.  Warp and weft elements write disjoint slots of the node's rows
// except the L-L self block, which the weft groups keep apart.  The bend groups write the
// slots as they go; the edge groups sum each edge's entries in registers and add them to
// private planes (the node and its two neighbours along the direction), which they fold
// into the slots after a barrier, so that an edge group and the bend group of its
// direction never update a slot at the same time; group 0 adds the handle.  After another
// barrier group 0 adds the other groups' accumulators and the weft self block in and does
// the warp component of the friction and b of the L and E_u rows; group 2 adds the weft
// groups' E_v sums, forms the compression sum in group 0's order and does the weft
// component and b of the E_v row.  The order of every sum is fixed, so the result does not
// depend on timing.  T = dd: the directional derivatives along the tangents held in C, into the d*
// outputs of O, one thread per node (blocks of 64 threads).  ASM1_REF: the reference
// gradient's tangent run.
#pragma once

namespace {
using namespace ys;

// Adds an edge group's private planes P (3 node columns x 16 entries, TILE_N apart) to the
// node's slots A, for the yarn direction CD (0: warp, along u; 1: weft, along v): column 0
// the node itself, 1 / 2 its neighbour before / after it along CD; has[col]: the column's
// edge exists, ee[col]: its E entries exist.  The slots are slot_of's, as constants, so the
// updates need no index arithmetic and do not wait on each other.  The weft group's L-L
// self block goes to its own planes Aself.
template <int CD>
__device__ __forceinline__ void fold_edge_planes(double* A, double* Aself, const double* P, const bool (&has)[3],
                                                 const bool (&ee)[3]) {
    constexpr int OLL[3] = {0, CD == 0 ? 1 : 3, CD == 0 ? 2 : 4};   // L-L stencil offsets
    constexpr int DLT[3] = {0, -1, 1};
    constexpr int ER = CD == 0 ? SLOT_EU : SLOT_EV;                 // the direction's E row
#pragma unroll
    for (int col = 0; col < 3; ++col) {
        if (!has[col]) continue;
        const double* Pc = P + (size_t)col * 16 * TILE_N;
#pragma unroll
        for (int rc = 0; rc < 3; ++rc)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const double x = Pc[(size_t)(rc * 3 + cc) * TILE_N];
                if (CD == 1 && col == 0) Aself[(size_t)(rc * 3 + cc) * TILE_N] += x;
                else A[(size_t)(rc * 45 + OLL[col] * 3 + cc) * TILE_N] += x;
            }
        if (!ee[col]) continue;
#pragma unroll
        for (int rc = 0; rc < 3; ++rc) A[(size_t)(rc * 45 + SLOT_LE + 3 * CD + DLT[col] + 1) * TILE_N] += Pc[(size_t)(9 + rc) * TILE_N];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) A[(size_t)(ER + (DLT[col] + 1) * 3 + cc) * TILE_N] += Pc[(size_t)(12 + cc) * TILE_N];
        A[(size_t)(ER + 10 + DLT[col]) * TILE_N] += Pc[(size_t)15 * TILE_N];
    }
}

template <class T, bool ASM1_REF>
__device__ __forceinline__ void assemble1_block(const ClothDev& C, const AsmOut& O, double h, int tile) {
    constexpr bool tan = !std::is_same<T, double>::value;
    const int nl = threadIdx.x & (TILE_N - 1);              // node lane in the tile
    const int grp = tan ? -1 : (int)(threadIdx.x >> 6);     // -1: every element (tangent run)
    const int k = tile_node(C, tile, nl);
    const bool active = k >= 0;
    RowSinkT<T, ASM1_REF> S;
    S.init(&C, active ? k : 0, tan ? nullptr : O.A, h, tan ? O.dA : nullptr, tan && O.detach);
    if constexpr (!tan) S.select_kind(O.only_kind);
    const int u = S.mu_, v = S.mv_;
    // value run: the node's 159 slots accumulate in LDS (the tile's slot planes, same
    // layout as in HBM) and go to HBM once at the end, instead of a global
    // read-modify-write per contribution; behind them the weft groups' self block (9
    // planes) and the edge groups' private planes (2 x ASM1_PRV), which later carry the
    // accumulators of groups 1-3 (3 x NACC planes)
    extern __shared__ double lds_rows[];
    double* gA = S.A;
    double* self_w = lds_rows + (size_t)NSLOT * TILE_N;
    double* prv = self_w + (size_t)9 * TILE_N;
    double* prv_g = prv + (size_t)(grp == 2 ? ASM1_PRV : 0) * TILE_N + nl;
    if constexpr (!tan) {
        S.A = lds_rows + nl;
        for (int s = grp; s < NSLOT; s += 4) S.A[(size_t)s * TILE_N] = 0.0;
        if (grp >= 2) S.Aself = self_w + nl;
        if (grp == 3)
            for (int i = 0; i < 9; ++i) S.Aself[(size_t)i * TILE_N] = 0.0;
        if (grp == 0 || grp == 2)
            for (int i = 0; i < ASM1_PRV; ++i) prv_g[(size_t)i * TILE_N] = 0.0;
        __syncthreads();
    } else {
        if (active && S.dA) for (int s = 0; s < NSLOT; ++s) S.dA[(size_t)s * TILE_N] = 0.0;
    }
    if (active) {
        // handle (src/PhysicsCloth.cpp:232-248); the value run's group 0 adds it after the
        // barrier below
        if (tan && (C.flags[k] & 2)) handle_element<T>(C, S, k);
        // edges (list order: warp rows, then weft columns; src/Cloth.cpp:190-228).  One
        // call site in a loop that is not unrolled: the inlined element code is large, and
        // four or six copies of it would not fit the instruction cache.
        int ne_warp = C.Lg * (C.W - 1);
        const int e0 = grp <= 0 ? 0 : 2, e1 = (grp < 0 || grp == 2) ? 4 : (grp == 0 ? 2 : 0);
    #pragma unroll 1
        for (int ei = e0; ei < e1; ++ei) {
            int eu, ev, ec, idx;
            if (ei == 0) { if (u == 0) continue; eu = u - 1; ev = v; ec = 0; idx = v * (C.W - 1) + u - 1; }
            else if (ei == 1) { if (u == C.W - 1) continue; eu = u; ev = v; ec = 0; idx = v * (C.W - 1) + u; }
            else if (ei == 2) { if (v == 0) continue; eu = u; ev = v - 1; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v - 1; }
            else { if (v == C.Lg - 1) continue; eu = u; ev = v; ec = 1; idx = ne_warp + u * (C.Lg - 1) + v; }
            // value run: the edge's entries in the node's rows summed in registers, then
            // added to the private planes (the other end precedes the node for ei = 0, 2)
            // (the tangent run has no registers to spare for the dual sums)
            if constexpr (!tan) S.edge_begin();
            edge_element<T>(C, S, eu, ev, ec, ec == 0 ? (v & 1) : (u & 1), omp_owner(C.n_edges, idx));
            if constexpr (!tan) S.edge_end_planes(prv_g, TILE_N, (ei & 1) ? 2 : 1);
        }
        // bends centred on u-1, u, u+1 (warp) and v-1, v, v+1 (weft) (src/Cloth.cpp:230-270):
        // a bend centred d after the node holds it as n2 (role 0), centred on it as n0
        // (role 1), centred d before it as n1 (role 2)
        int nb_warp = C.Lg * (C.W - 2);
        const int b0 = (grp < 0 || grp == 1) ? 0 : 3, b1 = (grp < 0 || grp == 3) ? 6 : (grp == 1 ? 3 : 0);
    #pragma unroll 1
        for (int bi6 = b0; bi6 < b1; ++bi6) {
            int d = bi6 % 3 - 1, dir = bi6 / 3;
            int uc = dir == 0 ? u + d : u, vc = dir == 0 ? v : v + d;
            if (dir == 0 && (uc < 1 || uc > C.W - 2)) continue;
            if (dir == 1 && (vc < 1 || vc > C.Lg - 2)) continue;
            int bi = dir == 0 ? v * (C.W - 2) + uc - 1 : nb_warp + u * (C.Lg - 2) + vc - 1;
            // value run: the bend centred on the node leaves its crimp twin (compression sums
            // only) to the edge group of its direction, which has less to do
            if constexpr (USE_BEND_REC && !tan) emit_bend_role(C, S, O.bend_rec, C.n_bends, bi, 1 - d, omp_owner(C.n_bends, bi));
            else bend_element<T>(C, S, uc, vc, dir, dir == 0 ? (v & 1) : (u & 1), omp_owner(C.n_bends, bi), (!tan && d == 0) ? 1 : 3);
        }
        if constexpr (!tan) {
            if (grp == 0 || grp == 2) {
                const int dir = grp == 0 ? 0 : 1;
                if (dir == 0 ? (u >= 1 && u <= C.W - 2) : (v >= 1 && v <= C.Lg - 2)) {
                    const int bi = dir == 0 ? v * (C.W - 2) + u - 1 : nb_warp + u * (C.Lg - 2) + v - 1;
                    bend_element<T>(C, S, u, v, dir, dir == 0 ? (v & 1) : (u & 1), omp_owner(C.n_bends, bi), 2);
                }
            }
        }
        // faces of the up to four cells around the node (src/Cloth.cpp:272-286)
        if ((grp < 0 || grp == 2) && (C.wind_density != 0.0 || C.wind_drag != 0.0)) {
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
        // the edge groups fold their private planes into the slots: the node's own column
        // (E entries when it is interior) and its neighbours' along the direction (E
        // entries when both are interior), the entries the edges' terms emit; group 2's
        // L-L self block goes to the weft self planes
        __syncthreads();
        if (active && (grp == 0 || grp == 2)) {
            const int c = grp == 0 ? 0 : 1;
            const int up = c == 0 ? u - 1 : u, vp = c == 0 ? v : v - 1, un = c == 0 ? u + 1 : u, vn = c == 0 ? v : v + 1;
            const bool has_p = up >= 0 && vp >= 0, has_n = un < C.W && vn < C.Lg;
            const bool has[3] = {has_p || has_n, has_p, has_n};
            const bool ee[3] = {S.inter, S.inter && interior(C, up, vp), S.inter && interior(C, un, vn)};
            if constexpr (!tan) {
                if (c == 0) fold_edge_planes<0>(S.A, nullptr, prv_g, has, ee);
                else fold_edge_planes<1>(S.A, S.Aself, prv_g, has, ee);
            }
            if (grp == 0 && (C.flags[k] & 2)) handle_element<T>(C, S, k);
        }
        // groups 1-3 hand their accumulators over, group 0 its compression sum; group 0 adds
        // them and the self block, group 2 the weft component's sums and the compression
        // sum in group 0's order ((g0 + g1) + g2) + g3, so that both hold the same bits
        __syncthreads();
        constexpr int NA = RowSinkT<T>::NACC;
        double* sbcl0 = prv + (size_t)3 * NA * TILE_N + nl;   // group 0's SBC_L, 3 planes
        if (grp > 0 && active) S.store_acc(prv + (size_t)(grp - 1) * NA * TILE_N + nl, TILE_N);
        if (grp == 0 && active) {
            sbcl0[0] = val(S.SBC_L.x); sbcl0[TILE_N] = val(S.SBC_L.y); sbcl0[2 * TILE_N] = val(S.SBC_L.z);
        }
        __syncthreads();
        if (grp == 0 && active) {
            for (int g = 0; g < 3; ++g) S.add_acc(prv + (size_t)g * NA * TILE_N + nl, TILE_N);
            for (int rc = 0; rc < 3; ++rc)
                for (int cc = 0; cc < 3; ++cc)
                    S.A[(size_t)(rc * 45 + cc) * TILE_N] += self_w[(size_t)(rc * 3 + cc) * TILE_N + nl];
        }
        if (grp == 2 && active) {
            auto sbcl = [&](const double* p) {
                return V3<T>(T(p[0]), T(p[(size_t)TILE_N]), T(p[(size_t)2 * TILE_N]));
            };
            const double* p1 = prv + (size_t)RowSinkT<T>::ACC_SBCL * TILE_N + nl;                   // group 1's
            const double* p3 = prv + (size_t)(2 * NA + RowSinkT<T>::ACC_SBCL) * TILE_N + nl;        // group 3's
            S.SBC_L = ((sbcl(sbcl0) + sbcl(p1)) + S.SBC_L) + sbcl(p3);
            S.add_acc_weft(prv + (size_t)2 * NA * TILE_N + nl, TILE_N);
        }
        // friction at the node's own crossing (src/PhysicsCloth.cpp:114-123): the warp
        // component by group 0, the weft one by group 2; each writes its component's outputs.
        // The component is a compile-time constant (a run-time index into the sink's
        // arrays would put them in scratch memory).
        auto fric_part = [&](auto CF) {
            constexpr int cf = decltype(CF)::value;
            T comp = T(0.0), sl[2] = {T(0.0), T(0.0)};
            if (O.fric_in) {
                const size_t N = C.N;
                auto put = [&](int i, T x) { O.fric_in[(size_t)i * N + k] = S.inter ? val(x) : 0.0; };
                if (cf == 0) {
                    put(FI_SBCL, S.SBC_L.x); put(FI_SBCL + 1, S.SBC_L.y); put(FI_SBCL + 2, S.SBC_L.z);
                }
                put(FI_SBCLE + 3 * cf, S.SBC_LE[cf].x); put(FI_SBCLE + 3 * cf + 1, S.SBC_LE[cf].y);
                put(FI_SBCLE + 3 * cf + 2, S.SBC_LE[cf].z);
                put(FI_FNS + cf, S.F[3 + cf]); put(FI_FEP + cf, S.FEpart[cf]);
                put(FI_KXEE + cf, S.KxEE[cf]); put(FI_KVEE + cf, S.KvEE[cf]);
            }
            if (S.inter) comp = friction<T>(C, S, sl, cf, cf + 1);
            // b = h (F + f_ext - K_v^T v) + M v (src/TrainHelper.cpp:20, 63)
            const int pk = padidx(O.Wp, u, v);
            const int e = S.inter ? crossing(C, u, v) : 0;
            if (cf == 0) {
                O.comp[k] = val(comp);
                for (int c = 0; c < 4; ++c) O.F[c * C.N + k] = val(S.F[c]);
            } else {
                O.F[4 * C.N + k] = val(S.F[4]);
            }
            O.is_slide[cf * C.N + k] = val(sl[cf]);
            O.fepart[cf * C.N + k] = val(S.FEpart[cf]);
#pragma unroll
            for (int c = cf == 0 ? 0 : 4; c < (cf == 0 ? 4 : 5); ++c) {
                double bv = 0.0;
                if (c < 3 || S.inter) {
                    double ext = 0.0;
                    if (O.ext) ext = c < 3 ? O.ext[3 * k + c] : O.ext[3 * C.N + 2 * e + (c - 3)];
                    bv = h * (val(S.F[c]) + ext - val(S.KvTv[c])) + val(S.Mv[c]);
                }
                O.b[c * O.P + pk] = bv;
            }
        };
        if (active && grp == 0) fric_part(std::integral_constant<int, 0>{});
        if (active && grp == 2) fric_part(std::integral_constant<int, 1>{});
    } else if (active) {
        // friction at the node's own crossing (src/PhysicsCloth.cpp:114-123)
        T comp = T(0.0), sl[2] = {T(0.0), T(0.0)};
        if (S.inter) comp = friction<T>(C, S, sl);
        // b = h (F + f_ext - K_v^T v) + M v (src/TrainHelper.cpp:20, 63)
        int pk = padidx(O.Wp, u, v);
        int e = S.inter ? crossing(C, u, v) : 0;
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
    if constexpr (!tan) {
        // the rows to HBM, all four groups, once the friction's entries are in, with the NaN/Inf
        // screen of the slots the shear replay does not touch (it screens its own)
        __syncthreads();
        if (active) {
            bool bad = false;
            for (int s = grp; s < NSLOT; s += 4) {
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

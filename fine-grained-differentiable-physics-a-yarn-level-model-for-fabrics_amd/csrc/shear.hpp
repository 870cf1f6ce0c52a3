// shear.hpp -- crossing shear element (src/PhysicsCloth.cpp:2643-3089), gfx950.
// Included by yarnstep.hip after the assembly kernels are declared.
#pragma once

namespace {

// One shear segment (n0 centre, n1 warp neighbour, n3 weft neighbour; geometry
// src/ClothUnit.h:240-248), stiffness scaled by the centre's compression.  The
// jamming branch keeps the reference's expressions as written, including the
// operator-precedence forms (1 - c^2 * H) and the kx second derivatives without PI.
// Forces and LL blocks of one segment: F[i] on node i and block B[i][j] = dF_i/dx_j, for
// i, j over (n0, n1, n3).
template <class T>
struct ShearOut {
    v3<T> F[3];
    mat3<T> B[3][3];
};

// Inlined into each kernel, so that a kernel that discards the blocks does not compute
// them and the outputs need no stack copy.
template <class T>
__device__ __forceinline__ void shear_compute(const ClothDev& C, int k0, int k1, int k3, const double* comp_arr, const double* dcomp_arr,
                              ShearOut<T>& out) {
    typedef v3<T> V;
    typedef mat3<T> M;
    const double c = 3.0;        // src/PhysicsCloth.cpp:2649
    const double sig = 0.6;      // src/PhysicsCloth.cpp:2650
    const double L = C.L, R = C.R;
    V x0 = pos<T>(C, k0);
    V a = pos<T>(C, k1) - x0, b3 = pos<T>(C, k3) - x0;
    T l1 = nrm(a), l3 = nrm(b3);
    V d1 = V3<T>(a.x / l1, a.y / l1, a.z / l1), d3v = V3<T>(b3.x / l3, b3.y / l3, b3.z / l3);
    M P1 = projector(d1), P3 = projector(d3v);
    T phi = acos(dot(d1, d3v));
    T cp = par<T>(comp_arr[k0], tan_at(dcomp_arr, 0, k0, 0));
    T kx = (cp + 1.0) * par<T>(C.S, C.dS) * R * R * PI_REF;
    T sn = sin(phi), cs = cos(phi);
    V P1d3 = P1 * d3v, P3d1 = P3 * d1;
    T cs2 = cs / pow(sn, 2.0);
    // recurring matrices
    M A13 = outerP(P1d3, d3v, P1);   // P1 V d3^T P1
    M A31 = outerP(P3d1, d1, P3);    // P3 d1 d1^T P3
    M B13 = outerP(P1d3, d1, P3);    // P1 V d1^T P3
    M B31 = outerP(P3d1, d3v, P1);   // P3 d1 d3^T P1
    bool jam = flip(C, k0) != flip(C, k1) && flip(C, k0) != flip(C, k3);   // src/ClothUnit.h:226-230
    M F11, F13, F31, F33;
    V Fx1, Fx3;
    if (jam) {
        const double phib = PI_REF / 2.0, phil = C.jam_thrd;
        V px1 = (-1.0 / (l1 * sn)) * P1d3;
        V px3 = (-1.0 / (l3 * sn)) * P3d1;
        M p11 = (1.0 / (pow(l1, 2.0) * sn)) * (outer(P1d3, d1) + outerP(d1, d3v, P1) - cs2 * A13 + cs * P1);
        M p33 = (1.0 / (pow(l3, 2.0) * sn)) * (outer(P3d1, d3v) + outerP(d3v, d1, P3) - cs2 * A31 + cs * P3);
        M p13 = (-1.0 / (l1 * l3 * sn)) * (cs2 * B13 + P1 * P3);
        M p31 = (-1.0 / (l3 * l1 * sn)) * (cs2 * outerP(P3d1, d3v, P3) + P3 * P1);
        const double LR = L / R;
        T ch = cos(0.5 * phi), sh = sin(0.5 * phi);
        T gamma = (L * sqrt(2.0) - 2.0 * L * sin(0.5 * phi)) / R;
        V gx1 = (-LR * ch) * px1, gx3 = (-LR * ch) * px3;
        M g11 = ((0.5 * LR * sh) / pow(l1 * sn, 2.0)) * A13 - (LR * ch) * p11;
        M g33 = ((0.5 * LR * sh) / pow(l3 * sn, 2.0)) * A31 - (LR * ch) * p33;
        T chh = cos(phi / 2.0), shh = sin(phi / 2.0);
        M g13 = ((0.5 * LR * shh) / (l1 * l3 * pow(sn, 2.0))) * B13 - (LR * chh) * p13;
        M g31 = ((0.5 * LR * shh) / (l3 * l1 * pow(sn, 2.0))) * B31 - (LR * chh) * p31;
        T pl = phi - phil, pb = phi - phib;
        T ps = phi * pl * pb;
        T dps = pl * pb + phi * pb + phi * pl;    // d(phi_s)/d(phi)
        V ps1 = dps * px1, ps3 = dps * px3;
        // second derivatives of phi_s, term for term as written (src/PhysicsCloth.cpp:2810-2848)
        T q2 = 2.0 * pb + 2.0 * pl + 2.0 * phi;   // weight of the phi_x phi_x^T terms
        M o11 = outer(px1, px1), o33 = outer(px3, px3), o13 = outer(px1, px3), o31 = outer(px3, px1);
        M ps11 = dps * p11 + q2 * o11;
        M ps33 = dps * p33 + q2 * o33;
        M ps13 = dps * p13 + (pb + pl + phi) * o13 + (pb + pl + phi) * o31;
        M ps31 = dps * p31 + (2.0 * pb + pl + 2.0 * phi) * o31 + pl * o31;
        const double pb5 = pow(phib, 5.0), pb4 = pow(phib, 4.0);
        T num = pb5 * pl;
        T den = pow(ps, 2.0) + pb4 * sig * sig;
        T cfx = num / den;
        T cf = tanh(cfx);
        V num1 = pb5 * px1, num3 = pb5 * px3;
        V den1 = (2.0 * ps) * ps1, den3 = (2.0 * ps) * ps3;
        V n1 = den * num1 - num * den1, n3 = den * num3 - num * den3;   // con_func_x_x*_num
        T dd2 = pow(den, 2.0);
        V cfx1 = (1.0 / dd2) * n1, cfx3 = (1.0 / dd2) * n3;
        T cf2 = pow(cf, 2.0);
        V cf1 = (1.0 - cf2) * cfx1, cf3 = (1.0 - cf2) * cfx3;
        M n1x1 = den * (pb5 * p11) - (2.0 * num) * outer(ps1, ps1) - (2.0 * num * ps) * ps11;
        M n3x3 = den * (pb5 * p33) - (2.0 * num) * outer(ps3, ps3) - (2.0 * num * ps) * ps33;
        M n1x3 = den * (pb5 * p13) + (2.0 * ps) * outer(num1, ps3) - (2.0 * ps) * outer(num3, ps1) -
                   (2.0 * num) * outer(ps1, ps3) - (2.0 * num * ps) * ps13;
        M n3x1 = den * (pb5 * p31) + (2.0 * ps) * outer(num3, ps1) - (2.0 * ps) * outer(num1, ps3) -
                   (2.0 * num) * outer(ps3, ps1) - (2.0 * num * ps) * ps31;
        T dd4 = pow(dd2, 2.0);
        M cfx11 = (1.0 / dd4) * (dd2 * n1x1 - outer(n1, (2.0 * den) * den1));
        M cfx33 = (1.0 / dd4) * (dd2 * n3x3 - outer(n3, (2.0 * den) * den3));
        M cfx13 = (1.0 / dd4) * (dd2 * n1x3 - outer(n1, (2.0 * den) * den3));
        M cfx31 = (1.0 / dd4) * (dd2 * n3x1 - outer(n3, (2.0 * den) * den1));
        T m2 = -2.0 * cf * (1.0 - cf2);
        M cf11 = m2 * outer(cfx1, cfx1) + addS(1.0, (-cf2) * cfx11);
        M cf33 = m2 * outer(cfx3, cfx3) + addS(1.0, (-cf2) * cfx33);
        M cf13 = m2 * outer(cfx3, cfx1) + (1.0 - cf2) * cfx13;
        M cf31 = m2 * outer(cfx1, cfx3) + (1.0 - cf2) * cfx31;
        T gc = pow(gamma, c), gc1 = pow(gamma, c - 1.0), gc2 = pow(gamma, c - 2.0);
        const double c1 = c * (c - 1.0);
        // kx second derivative terms (term1 - term2 + term3, src/PhysicsCloth.cpp:2928-2995)
        M k11 = (c1 * gc2 * outer(gx1, gx1) + (c * gc1) * g11) -
                  ((c1 * gc2 * cf) * outer(gx1, gx1) + (c * gc1 * cf) * g11 + (c * gc1) * outer(gx1, cf1)) +
                  ((-c * gc1) * outer(gx1, cf1) + (1.0 - gc) * cf11);
        M k33 = (c1 * gc2 * outer(gx3, gx3) + (c * gc1) * g33) -
                  ((c1 * gc2 * cf) * outer(gx3, gx3) + (c * gc1 * cf) * g33 + (c * gc1) * outer(gx3, cf3)) +
                  ((-c * gc1) * outer(gx3, cf3) + (1.0 - gc) * cf33);
        M k13 = (c1 * gc2 * outer(gx1, gx3) + (c * gc1) * g13) -
                  ((c1 * gc2 * cf) * outer(gx1, gx3) + (c * gc1 * cf) * g13 + (c * gc1) * outer(gx1, cf3)) +
                  ((-c * gc1) * outer(cf1, gx3) + (1.0 - gc) * cf13);
        M k31 = (c1 * gc2 * outer(gx3, gx1) + (c * gc1) * g31) -
                  ((c1 * gc2 * cf) * outer(gx3, gx1) + (c * gc1 * cf) * g31 + (c * gc1) * outer(gx3, cf1)) +
                  ((-c * gc1) * outer(cf3, gx1) + (1.0 - gc) * cf31);
        T base = 0.5 * (cp + 1.0) * par<T>(C.S, C.dS) * PI_REF * pow(R, 2.0);
        kx = base * ((1.0 + gc) + (1.0 - gc) * cf);
        V one = V3<T>(1.0, 1.0, 1.0);
        V kx1 = base * ((c * gc1) * gx1 - (c * gc1 * cf) * gx1 + (1.0 - gc) * (one - cf2 * cfx1));
        V kx3 = base * ((c * gc1) * gx3 - (c * gc1 * cf) * gx3 + (1.0 - gc) * (one - cf2 * cfx3));
        T base2 = 0.5 * (cp + 1.0) * par<T>(C.S, C.dS) * pow(R, 2.0);
        M kx11 = base2 * k11, kx33 = base2 * k33, kx13 = base2 * k13, kx31 = base2 * k31;
        T pb2 = pow(pb, 2.0);
        Fx1 = (-0.5 * L * pb2) * kx1 + ((kx * L * pb) / (l1 * sn)) * P1d3;
        Fx3 = (-0.5 * L * pb2) * kx3 + ((kx * L * pb) / (l3 * sn)) * P3d1;
        M F11o = ((kx * L) / (pow(l1, 2.0) * sn)) * (pb * (-outer(P1d3, d1) + cs2 * A13 - cs * P1 - outerP(d1, d3v, P1)) - (1.0 / sn) * A13);
        M F33o = ((kx * L) / (pow(l3, 2.0) * sn)) * (pb * (-outer(P3d1, d3v) + cs2 * A31 - cs * P3 - outerP(d3v, d1, P3)) - (1.0 / sn) * A31);
        M F13o = ((kx * L) / (l1 * l3 * sn)) * (pb * (cs2 * B13 + P1 * P3) - (1.0 / sn) * B13);
        M F31o = ((kx * L) / (l3 * l1 * sn)) * (pb * (cs2 * B31 + P3 * P1) - (1.0 / sn) * B31);
        F11 = (-0.5 * L * pb2) * kx11 - (2.0 * L * pb) * outer(kx1, px1) + F11o;
        F33 = (-0.5 * L * pb2) * kx33 - (2.0 * L * pb) * outer(kx3, px3) + F33o;
        F13 = (-0.5 * L * pb2) * kx13 - (L * pb) * outer(kx1, px3) - (L * pb) * outer(px1, kx3) + F13o;
        F31 = (-0.5 * L * pb2) * kx31 - (L * pb) * outer(kx3, px1) - (L * pb) * outer(px3, kx1) + F31o;
    } else {
        // the reference takes (phi - pi/2) through .item() here, so its gradient does not
        // see it (src/PhysicsCloth.cpp:2720-2721); ref_mode reproduces that
        T pb = C.ref_mode ? T(val(phi) - PI_REF / 2.0) : phi - PI_REF / 2.0;
        Fx1 = ((kx * L * pb) / (l1 * sn)) * P1d3;
        Fx3 = ((kx * L * pb) / (l3 * sn)) * P3d1;
        F11 = ((kx * L) / (pow(l1, 2.0) * sn)) * (pb * (-outer(P1d3, d1) + cs2 * A13 - cs * P1 - outerP(d1, d3v, P1)) - (1.0 / sn) * A13);
        F13 = ((kx * L) / (l3 * l1 * sn)) * ((pb * (cs2 * outer(P1d3, d1) - (1.0 / sn) * outer(P1d3, d1))) * P3);
        F31 = ((kx * L) / (l1 * l3 * sn)) * ((pb * (cs2 * outer(P3d1, d3v) - (1.0 / sn) * outer(P3d1, d3v))) * P1);
        F33 = ((kx * L) / (pow(l3, 2.0) * sn)) * (pb * (-outer(P3d1, d3v) + cs2 * A31 - cs * P3 - outerP(d3v, d1, P3)) - (1.0 / sn) * A31);
    }
    out.F[0] = -(Fx1 + Fx3); out.F[1] = Fx1; out.F[2] = Fx3;
    out.B[1][0] = -(F11 + F13); out.B[2][0] = -(F31 + F33);
    out.B[0][1] = -(F11 + F31); out.B[0][2] = -(F13 + F33);
    out.B[0][0] = -(out.B[1][0] + out.B[2][0]);
    out.B[1][1] = F11; out.B[1][2] = F13; out.B[2][1] = F31; out.B[2][2] = F33;
}

// The segment's contributions to the rows of node role r (0: n0, 1: n1, 2: n3), in the
// order the whole segment emits them.
template <class T, class Sink>
__device__ void emit_shear_role(Sink& S, const int (&k)[3], int r, const v3<T>& F, const mat3<T> (&B)[3]) {
    S.FL(k[r], F);
    for (int j = 0; j < 3; ++j) S.LL(KIND_KX, k[r], k[j], B[j]);
}

// Segment s = q Nc + e: segment q of interior crossing e, q in the reference's shear_segs
// order (left,bot), (left,top), (right,bot), (right,top) (src/Cloth.cpp:170-181), so that
// neighbouring crossings' records of one kind are adjacent (the node kernels' wavefronts
// read them for neighbouring nodes): its nodes.
__device__ __forceinline__ int shear_seg(const ClothDev& C, int e, int q) { return q * C.Nc + e; }
__device__ __forceinline__ void shear_nodes(const ClothDev& C, int s, int (&k)[3]) {
    int q = s / C.Nc, e = s - q * C.Nc;
    int cu = 1 + e % (C.W - 2), cv = 1 + e / (C.W - 2);
    int kc = cu + cv * C.W;
    k[0] = kc;
    k[1] = (q >> 1) == 0 ? kc - 1 : kc + 1;
    k[2] = (q & 1) == 0 ? kc + C.W : kc - C.W;
}

// Element-major shear: one thread per segment computes it once and stores, for each of
// its three nodes, the 30 doubles that node's rows receive (force 3 + three LL blocks
// 27), as planes rec[(r * 30 + f) * nseg + s].
__device__ __forceinline__ void shear_elem_seg(const ClothDev& C, const double* comp, double* rec, int nseg, int s) {
    if (s >= nseg) return;
    int k[3];
    shear_nodes(C, s, k);
    ShearOut<double> o;
    shear_compute<double>(C, k[0], k[1], k[2], comp, nullptr, o);
    for (int r = 0; r < 3; ++r) {
        double* base = rec + (size_t)r * 30 * nseg + s;
        base[0] = o.F[r].x; base[(size_t)nseg] = o.F[r].y; base[(size_t)2 * nseg] = o.F[r].z;
        for (int j = 0; j < 3; ++j)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) base[(size_t)(3 + j * 9 + a * 3 + b) * nseg] = o.B[r][j].m[a][b];
    }
}

__global__ void __launch_bounds__(256) k_shear_elem(ClothDev C, const double* comp, double* rec, int nseg) {
    shear_elem_seg(C, comp, rec, nseg, blockIdx.x * 256 + threadIdx.x);
}

// Tangent run: the tangents of the same 30 doubles per role, one thread per segment, along
// the direction of each block's descriptors (blockIdx.y: the direction; its state and
// parameter tangents in C, its compression tangents O.dcomp, its records O.shear_rec), in the
// value records' layout.
template <bool REF>
__global__ void __launch_bounds__(256) k_shear_elem_tan(const ClothDev* __restrict__ Cs, const AsmOut* __restrict__ Os,
                                                        const double* comp) {
    const ClothDev& C = Cs[blockIdx.y];
    const AsmOut& O = Os[blockIdx.y];
    const int nseg = O.nseg;
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nseg) return;
    int k[3];
    shear_nodes(C, s, k);
    ShearOut<dd> o;
    shear_compute<dd>(C, k[0], k[1], k[2], comp, O.dcomp, o);
    double* rec = const_cast<double*>(O.shear_rec);   // the direction's own buffer (the readers take it const)
    for (int r = 0; r < 3; ++r) {
        double* base = rec + (size_t)r * 30 * nseg + s;
        base[0] = o.F[r].x.d; base[(size_t)nseg] = o.F[r].y.d; base[(size_t)2 * nseg] = o.F[r].z.d;
        if constexpr (REF) continue;   // the reference gradient needs no block tangents
        for (int j = 0; j < 3; ++j)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) base[(size_t)(3 + j * 9 + a * 3 + b) * nseg] = o.B[r][j].m[a][b].d;
    }
}

// Reads the role-r record of segment s.
__device__ __forceinline__ void shear_load(const double* rec, int nseg, int s, int r, d3& F, m33 (&B)[3]) {
    const double* base = rec + (size_t)r * 30 * nseg + s;
    F = D3(base[0], base[(size_t)nseg], base[(size_t)2 * nseg]);
    for (int j = 0; j < 3; ++j)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) B[j].m[a][b] = base[(size_t)(3 + j * 9 + a * 3 + b) * nseg];
}

}  // namespace

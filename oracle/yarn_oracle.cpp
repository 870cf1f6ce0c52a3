// yarn_oracle.cpp -- CPU oracle (test infrastructure) for the implicit yarn-level step.
//
// TEST INFRASTRUCTURE ONLY: this is the checker, never the product.  It restates the
// reference's FP64 step function by function, quirks included, so that the HIP
// kernels can be compared against it on the same inputs.  Each function cites the
// reference lines it follows (paths relative to the reference tree).
//
// Storage differs from the reference on purpose: the reference keeps dense N^2 block
// arrays per OpenMP thread (src/PhysicsCloth.cpp:7-37) and a dense n x n matrix
// (src/MatrixMp.cpp:57); here every contribution is scattered straight into a sparse
// row accumulator at the (row, col) the reference's To_Sparse_Mat would place it, so
// the values are the same and W up to 256 fits in memory.  The one place where the
// reference's per-thread partial sums change the arithmetic -- the friction input
// F_Es[ID] (src/PhysicsCloth.cpp:119-122, 2624) -- is emulated exactly with the
// libgomp static partition of the edge, bend and crossing loops over 8 threads.

#include "yarn_oracle.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

#include <omp.h>

namespace {

// src/Constants.h:24-26
const double PI = 3.1415926535897;
const double EPS = std::numeric_limits<double>::min();
const int NTHRD = 8;

typedef std::array<double, 3> V3;
typedef std::array<std::array<double, 3>, 3> M3;
typedef std::array<std::array<double, 4>, 4> M4;
typedef std::array<std::array<double, 8>, 8> M8;
typedef std::array<double, 8> V8;

V3 v3(double a, double b, double c) { V3 r = {a, b, c}; return r; }
V3 add(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
V3 sub(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
V3 scl(double s, const V3& a) { return v3(s * a[0], s * a[1], s * a[2]); }
V3 neg(const V3& a) { return scl(-1.0, a); }
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
V3 cross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
M3 zero3() { M3 m; for (auto& r : m) r.fill(0.0); return m; }
M3 eye3() { M3 m = zero3(); m[0][0] = m[1][1] = m[2][2] = 1.0; return m; }
M3 madd(const M3& a, const M3& b) { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = a[i][j] + b[i][j]; return m; }
M3 msub(const M3& a, const M3& b) { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = a[i][j] - b[i][j]; return m; }
M3 mscl(double s, const M3& a) { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = s * a[i][j]; return m; }
M3 mneg(const M3& a) { return mscl(-1.0, a); }
M3 mm(const M3& a, const M3& b) {
    M3 m = zero3();
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) for (int k = 0; k < 3; ++k) m[i][j] += a[i][k] * b[k][j];
    return m;
}
V3 mv(const M3& a, const V3& x) {
    return v3(a[0][0] * x[0] + a[0][1] * x[1] + a[0][2] * x[2],
              a[1][0] * x[0] + a[1][1] * x[1] + a[1][2] * x[2],
              a[2][0] * x[0] + a[2][1] * x[1] + a[2][2] * x[2]);
}
V3 vm(const V3& x, const M3& a) {  // x^T a
    return v3(x[0] * a[0][0] + x[1] * a[1][0] + x[2] * a[2][0],
              x[0] * a[0][1] + x[1] * a[1][1] + x[2] * a[2][1],
              x[0] * a[0][2] + x[1] * a[1][2] + x[2] * a[2][2]);
}
M3 outer(const V3& a, const V3& b) { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = a[i] * b[j]; return m; }
M3 proj(const V3& d) { return msub(eye3(), outer(d, d)); }  // I - d d^T
V3 unit(int a) { V3 r = {0, 0, 0}; r[a] = 1.0; return r; }
V3 vadd1(double s, const V3& a) { return v3(s + a[0], s + a[1], s + a[2]); }  // scalar + vector (broadcast)
M3 madd1(double s, const M3& a) { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = s + a[i][j]; return m; }

// 4x4 block [[A, c], [r^T, s]] over (x, u) of one edge end.
M4 blk(const M3& A, const V3& c, const V3& r, double s) {
    M4 b;
    for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) b[i][j] = A[i][j]; b[i][3] = c[i]; b[3][i] = r[i]; }
    b[3][3] = s;
    return b;
}
// scale * cat({cat({2B, B},1), cat({B, 2B},1)},0): the two-node edge mass pattern.
M8 kron2(const M4& b, double scale) {
    M8 m;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) {
        m[i][j] = scale * (2.0 * b[i][j]);
        m[i][j + 4] = scale * b[i][j];
        m[i + 4][j] = scale * b[i][j];
        m[i + 4][j + 4] = scale * (2.0 * b[i][j]);
    }
    return m;
}
M8 m8add(const M8& a, const M8& b) { M8 m; for (int i = 0; i < 8; ++i) for (int j = 0; j < 8; ++j) m[i][j] = a[i][j] + b[i][j]; return m; }
M8 m8scl(double s, const M8& a) { M8 m; for (int i = 0; i < 8; ++i) for (int j = 0; j < 8; ++j) m[i][j] = s * a[i][j]; return m; }
V8 m8v(const M8& a, const V8& q) { V8 r; for (int i = 0; i < 8; ++i) { double s = 0; for (int j = 0; j < 8; ++j) s += a[i][j] * q[j]; r[i] = s; } return r; }
double quad(const V8& q, const M8& a) { V8 t = m8v(a, q); double s = 0; for (int i = 0; i < 8; ++i) s += q[i] * t[i]; return s; }
double relu(double x) { return x > 0 ? x : 0.0; }
double sigm(double x) { return 1.0 / (1.0 + std::exp(-x)); }

// libgomp static schedule: n iterations over T threads, q = n/T, r = n%T,
// threads t < r own q+1 iterations starting at t*(q+1), the rest q starting at t*q+r.
int owner(int n, int i) {
    int q = n / NTHRD, r = n % NTHRD;
    int cut = r * (q + 1);
    if (i < cut) return i / (q + 1);
    return r + (i - cut) / q;
}

// Sparse row accumulator: entries are placed at the global (row, col) of the
// reference's dense A (src/MatrixMp.cpp:51-110).
struct SpAcc {
    std::vector<std::vector<std::pair<int, double>>> rows;
    void init(int n) { rows.assign(n, {}); }
    void zero() { for (auto& r : rows) for (auto& e : r) e.second = 0.0; }
    void add(int r, int c, double v) {
        auto& row = rows[r];
        for (auto& e : row) if (e.first == c) { e.second += v; return; }
        row.push_back(std::make_pair(c, v));
    }
    double get(int r, int c) const {
        for (auto& e : rows[r]) if (e.first == c) return e.second;
        return 0.0;
    }
};

// Parallel assembly (checker option, yo_set_threads): thread t owns the grid rows
// v in [tl_vlo, tl_vhi) and writes only the rows of A, F and the per-node accumulators that
// belong to its nodes; every thread walks the element lists in the reference's order, so
// each row receives its contributions in the serial order and the result is bit for bit
// the serial one.
thread_local int tl_vlo = 0, tl_vhi = INT_MAX;

enum { EDGE_WARP = 0, EDGE_WEFT = 1 };

struct Edge { int n0, n1, type, yarn; };            // src/ClothUnit.h:85-160
struct Bend { int n2, n0, n1, type, yarn; };        // src/ClothUnit.h:162-220
struct Shear { int n0, n1, n3; bool jam; };         // src/ClothUnit.h:222-256
struct Face { int n1, n2, n3; };                    // src/ClothUnit.h:258-283

}  // namespace

struct yo_cloth {
    yo_params p;
    int W, Lg, N, Nc, n;
    double ks[2], kb[2];
    double jamming_thrd;
    // node state (src/ClothUnit.h:7-27)
    std::vector<V3> x, v, xfix, nrm, xwarp, xweft;
    std::vector<std::array<double, 2>> uv, uvd, uvbar, uv_crimp;
    std::vector<char> is_edge, is_fixed, flip;
    // element lists in the reference's construction order (src/Cloth.cpp:167-305)
    std::vector<Edge> edges;
    std::vector<Bend> bends;
    std::vector<Shear> shears;
    std::vector<Face> faces;
    // assembly outputs
    std::vector<double> FL, FE, FEpart, FE_noslide;
    std::vector<V3> StrhBendC_L;                    // per node
    std::vector<std::array<V3, 2>> StrhBendC_LEself; // per node: LE(k, e(k), {u,v})
    std::vector<double> is_slide, comp;
    SpAcc M, Kx, Kv;
    // CSR of A, M, Kx, Kv after assembly
    std::vector<int> rowptr, col;
    std::vector<double> Aval, Mval, Kxval, Kvval;
    std::vector<double> b, vel, F;
    double h_last;

    int kidx(int u, int v) const { return u + v * W; }
    int eidx(int u, int v) const { return (u - 1) + (v - 1) * (W - 2); }
    int eofk(int k) const { return eidx(k % W, k / W); }

    // ----- scatter helpers: (buffer, find_index) -> global (row, col) -----
    // the rows of node k belong to this thread (always true outside a parallel assembly)
    bool own(int k) const { int v = k / W; return v >= tl_vlo && v < tl_vhi; }
    // LL[F,D][u][v] -> A[3F+v, 3D+u] (transposed block, src/MatrixMp.cpp:65-76)
    void addLL(SpAcc& S, int F, int D, const M3& B) {
        if (!own(F)) return;
        for (int u = 0; u < 3; ++u) for (int v2 = 0; v2 < 3; ++v2) S.add(3 * F + v2, 3 * D + u, B[u][v2]);
    }
    // LE at find_index(L node kL, E node kE, "EL") + comp -> A[3kL+v, 3N+2e(kE)+comp] (src/MatrixMp.cpp:78-87)
    void addLE(SpAcc& S, int kL, int kE, int c, const V3& B) {
        if (!own(kL)) return;
        int col = 3 * N + 2 * eofk(kE) + c;
        for (int v2 = 0; v2 < 3; ++v2) S.add(3 * kL + v2, col, B[v2]);
    }
    // EL at find_index(L node kL, E node kE, "EL") + comp -> A[3N+2e(kE)+comp, 3kL+u] (src/MatrixMp.cpp:89-98)
    void addEL(SpAcc& S, int kL, int kE, int c, const V3& B) {
        if (!own(kE)) return;
        int row = 3 * N + 2 * eofk(kE) + c;
        for (int u = 0; u < 3; ++u) S.add(row, 3 * kL + u, B[u]);
    }
    // EE at find_index(F, D, "EE", uu|vv) -> A[3N+2e(F)+c, 3N+2e(D)+c] (src/PhysicsCloth.cpp:205-209)
    void addEE(SpAcc& S, int kF, int kD, int c, double val) {
        if (!own(kF)) return;
        S.add(3 * N + 2 * eofk(kF) + c, 3 * N + 2 * eofk(kD) + c, val);
    }
    void addFL(int k, const V3& f) { if (!own(k)) return; for (int a = 0; a < 3; ++a) FL[3 * k + a] += f[a]; }
    // the compression accumulators StrhBend_C_L and the self entries of StrhBend_C_LE
    void addSBC(int k, const V3& f) { if (own(k)) StrhBendC_L[k] = add(StrhBendC_L[k], f); }
    void addSBCLE(int k, int c, const V3& f) { if (own(k)) StrhBendC_LEself[k][c] = add(StrhBendC_LEself[k][c], f); }
    // E force on crossing of node k, component c; from_elem marks the Q2 side channel
    void addFE(int k, int c, double f, int owner_t) {
        if (!own(k)) return;
        int e = eofk(k);
        FE[2 * e + c] += f;
        if (owner_t >= 0 && owner_t == owner(Nc, e)) FEpart[2 * e + c] += f;
    }

    // ----- construction (src/Cloth.cpp:3-109, 167-305) -----
    void build();
    void normals_and_crimp();     // UpdateClothUnits (src/Cloth.cpp:307-348), NodeNorm (440-463)
    V3 node_norm(int kc, int kl, int kr, int kt, int kb2) const;

    // ----- element functions (src/PhysicsCloth.cpp) -----
    struct EdgeGeo { double duv; V3 w, d, dx; double l; M3 P; V8 qd; double rho, ks; int c; };
    EdgeGeo edge_geo(const Edge& e) const;
    void Mass(const Edge& e, const EdgeGeo& g);
    void Gravity_F(const Edge& e, const EdgeGeo& g, int t);
    void Gravity_D(const Edge& e, const EdgeGeo& g);
    void Stretch_F(const Edge& e, const EdgeGeo& g, int t);
    void Stretch_D(const Edge& e, const EdgeGeo& g);
    void Inertia_F(const Edge& e, const EdgeGeo& g, int t);
    void Inertia_D_Pos(const Edge& e, const EdgeGeo& g);
    void Inertia_D_Vel(const Edge& e, const EdgeGeo& g);
    void MDotVDot_F(const Edge& e, const EdgeGeo& g, int t);
    void MDotVDot_D_Pos(const Edge& e, const EdgeGeo& g);
    void MDotVDot_D_Vel(const Edge& e, const EdgeGeo& g);
    void ParaColli_F(const Edge& e, const EdgeGeo& g, int t);
    void ParaColli_D(const Edge& e, const EdgeGeo& g);
    void Bending(const Bend& b, int t);
    void Crimp_Bending(const Bend& b);
    void ConSlideFriction(int k);
    void Shear_FD(const Shear& s);
    // Checker option: hold the non-jamming shear's (phi - pi/2) at stored values, the
    // quantity the reference takes through .item() (src/PhysicsCloth.cpp:2720-2721), so
    // that differences of F reproduce the reference's detached derivative.
    std::vector<double> frozen_pb;
    double shear_phi(const Shear& s) const;
    void Wind_F(const Face& f);
    void Constraint(int k);
    void edge_blocks(SpAcc& S, const Edge& e, const M3& L0L0, const M3& L0L1, const M3& L1L0, const M3& L1L1,
                     const V3& L0E0, const V3& L0E1, const V3& L1E0, const V3& L1E1,
                     const V3& E0L0, const V3& E0L1, const V3& E1L0, const V3& E1L1,
                     double E0E0, double E0E1, double E1E0, double E1E1);

    int assemble(double h, const double* ext);
    template <class Mid> void element_loops(Mid mid);
    int threads = 1;                  // checker option: OpenMP threads of assembly and CG
    // checker option: which element terms an assembly evaluates (YO_TERM_* bits)
    unsigned terms = YO_TERM_ALL;
    bool on(unsigned t) const { return (terms & t) != 0; }
    // checker option: evaluate one element of one list only (YO_ELEM_* kind, list index)
    int only_kind = 0, only_index = -1;
    bool pick(int kind, int i) const { return only_kind == 0 || (only_kind == kind && only_index == i); }
};

// ---------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------

void yo_cloth::build() {
    W = p.width; Lg = p.length; N = W * Lg; Nc = (W - 2) * (Lg - 2); n = 3 * N + 2 * Nc;
    for (int y = 0; y < p.n_yarns; ++y) {  // src/Yarn.h:14-15
        ks[y] = p.Y[y] * PI * std::pow(p.R, 2);
        kb[y] = p.B[y] * PI * std::pow(p.R, 2);
    }
    jamming_thrd = 2.0 * std::asin(p.R / p.L);  // src/Cloth.cpp:29
    x.assign(N, V3{0, 0, 0}); v = x; xfix = x; nrm = x; xwarp = x; xweft = x;
    uv.assign(N, {0.0, 0.0}); uvd = uv; uvbar = uv; uv_crimp = uv;
    is_edge.assign(N, 0); is_fixed.assign(N, 0); flip.assign(N, 0);
    // src/Cloth.cpp:52-99
    for (int j = 0; j < Lg; ++j) for (int i = 0; i < W; ++i) {
        int k = kidx(i, j);
        V3 pos = v3(i * p.L, j * p.L, p.hight);
        if (p.pose == 0) pos = v3(i * p.L, 0.0, p.hight - j * p.L);
        x[k] = pos; xfix[k] = pos;
        uv[k] = {i * p.L, j * p.L}; uvbar[k] = uv[k]; uv_crimp[k] = uv[k];
        bool f = false;
        if (i == 0 || i == W - 1 || j == 0 || j == Lg - 1) { is_edge[k] = 1; }
        else {
            if (p.pattern == 0) { if ((i + j) % 2 != 0) f = true; }
            else if (p.pattern == 1) { int rv = j % 4; int ru = (i + rv) % 4; if (ru < 2) f = true; }
            else { int rv = (j - 1) % 4; int q = rv / 2; int r = rv % 2; int iu = q * 3 + r * 2; int ru = (i - 1 + iu) % 4; if (ru == 0) f = true; }
        }
        flip[k] = f;
    }
    // shear segments (src/Cloth.cpp:170-181): i over width outer, j over length inner
    for (int i = 1; i < W - 1; ++i) for (int j = 1; j < Lg - 1; ++j) {
        int k = kidx(i, j), kl = kidx(i - 1, j), kr = kidx(i + 1, j), kt = kidx(i, j - 1), kbot = kidx(i, j + 1);
        int nb[4][2] = {{kl, kbot}, {kl, kt}, {kr, kbot}, {kr, kt}};
        for (int s = 0; s < 4; ++s) {
            Shear sh; sh.n0 = k; sh.n1 = nb[s][0]; sh.n3 = nb[s][1];
            sh.jam = (flip[k] != flip[sh.n1]) && (flip[k] != flip[sh.n3]);  // src/ClothUnit.h:226-230
            shears.push_back(sh);
        }
    }
    int ny = p.n_yarns;
    // warp edges (src/Cloth.cpp:190-208), weft edges (src/Cloth.cpp:210-228)
    for (int i = 0; i < Lg; ++i) for (int j = 0; j < W - 1; ++j)
        edges.push_back(Edge{kidx(j, i), kidx(j + 1, i), EDGE_WARP, ny == 1 ? 0 : (i % 2)});
    for (int i = 0; i < W; ++i) for (int j = 0; j < Lg - 1; ++j)
        edges.push_back(Edge{kidx(i, j), kidx(i, j + 1), EDGE_WEFT, ny == 1 ? 0 : (i % 2)});
    // bends (src/Cloth.cpp:230-270); the crimp twin of bend b is bend b on the offset copies
    for (int i = 0; i < Lg; ++i) for (int j = 0; j < W - 2; ++j)
        bends.push_back(Bend{kidx(j, i), kidx(j + 1, i), kidx(j + 2, i), EDGE_WARP, ny == 1 ? 0 : (i % 2)});
    for (int i = 0; i < W; ++i) for (int j = 0; j < Lg - 2; ++j)
        bends.push_back(Bend{kidx(i, j), kidx(i, j + 1), kidx(i, j + 2), EDGE_WEFT, ny == 1 ? 0 : (i % 2)});
    // faces (src/Cloth.cpp:272-286)
    for (int i = 0; i < W - 1; ++i) for (int j = 0; j < Lg - 1; ++j) {
        int tl = kidx(i, j), tr = kidx(i + 1, j), bl = kidx(i, j + 1), br = kidx(i + 1, j + 1);
        if ((i + j) % 2 == 0) { faces.push_back(Face{tl, tr, br}); faces.push_back(Face{tl, br, bl}); }
        else { faces.push_back(Face{tl, tr, bl}); faces.push_back(Face{tr, br, bl}); }
    }
    for (int hnd = 0; hnd < p.n_handles; ++hnd) is_fixed[kidx(p.handles[2 * hnd], p.handles[2 * hnd + 1])] = 1;  // src/Cloth.cpp:386-394
    normals_and_crimp();
    FL.assign(3 * N, 0); FE.assign(2 * Nc, 0); FEpart = FE; FE_noslide = FE;
    StrhBendC_L.assign(N, V3{0, 0, 0});
    StrhBendC_LEself.assign(N, {V3{0, 0, 0}, V3{0, 0, 0}});
    is_slide.assign(2 * Nc, 0); comp.assign(Nc, 0);
    M.init(n); Kx.init(n); Kv.init(n);
}

// NodeNorm (src/Cloth.cpp:440-463): the right singular vector of the smallest
// singular value of the centred 5x3 point matrix, i.e. the eigenvector of the
// smallest eigenvalue of its 3x3 scatter matrix (cyclic Jacobi here), sign-aligned
// with (x_right - x_left) x (x_bot - x_top) and flipped when flip_norm is set.
V3 yo_cloth::node_norm(int kc, int kl, int kr, int kt, int kb2) const {
    const int ids[5] = {kc, kl, kr, kt, kb2};
    V3 mean = {0, 0, 0};
    for (int i = 0; i < 5; ++i) mean = add(mean, x[ids[i]]);
    mean = scl(1.0 / 5.0, mean);
    double C[3][3] = {{0}};
    for (int i = 0; i < 5; ++i) {
        V3 d = sub(x[ids[i]], mean);
        for (int a = 0; a < 3; ++a) for (int b2 = 0; b2 < 3; ++b2) C[a][b2] += d[a] * d[b2];
    }
    double Vm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = C[0][1] * C[0][1] + C[0][2] * C[0][2] + C[1][2] * C[1][2];
        double tr = C[0][0] * C[0][0] + C[1][1] * C[1][1] + C[2][2] * C[2][2];
        if (off <= 1e-34 * tr || off == 0.0) break;
        for (int pp = 0; pp < 2; ++pp) for (int q = pp + 1; q < 3; ++q) {
            if (C[pp][q] == 0.0) continue;
            double theta = (C[q][q] - C[pp][pp]) / (2.0 * C[pp][q]);
            double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            for (int r = 0; r < 3; ++r) {  // C <- C J
                double a = C[r][pp], b2 = C[r][q];
                C[r][pp] = c * a - s * b2; C[r][q] = s * a + c * b2;
            }
            for (int r = 0; r < 3; ++r) {  // C <- J^T C
                double a = C[pp][r], b2 = C[q][r];
                C[pp][r] = c * a - s * b2; C[q][r] = s * a + c * b2;
            }
            for (int r = 0; r < 3; ++r) {
                double a = Vm[r][pp], b2 = Vm[r][q];
                Vm[r][pp] = c * a - s * b2; Vm[r][q] = s * a + c * b2;
            }
        }
    }
    int imin = 0;
    for (int i = 1; i < 3; ++i) if (C[i][i] < C[imin][imin]) imin = i;
    V3 nr = v3(Vm[0][imin], Vm[1][imin], Vm[2][imin]);
    nr = scl(1.0 / norm(nr), nr);
    V3 vec_warp = sub(x[kr], x[kl]);
    V3 vec_weft = sub(x[kb2], x[kt]);
    V3 vc = cross(vec_warp, vec_weft);
    double cosv = dot(vc, nr) / (norm(vc) * norm(nr));
    if (cosv < 0) nr = neg(nr);
    if (flip[kc]) nr = neg(nr);
    return nr;
}

void yo_cloth::normals_and_crimp() {
    // src/Cloth.cpp:313-325: crossing normals and the +-R*n crimp copies
    for (int k = 0; k < N; ++k) { xwarp[k] = x[k]; xweft[k] = x[k]; }
#pragma omp parallel for num_threads(threads) if (threads > 1) schedule(static)
    for (int i = 1; i < W - 1; ++i) for (int j = 1; j < Lg - 1; ++j) {
        int k = kidx(i, j);
        V3 nr = node_norm(k, kidx(i - 1, j), kidx(i + 1, j), kidx(i, j - 1), kidx(i, j + 1));
        nrm[k] = nr;
        xwarp[k] = sub(x[k], scl(p.R, nr));
        xweft[k] = add(x[k], scl(p.R, nr));
    }
}

// ---------------------------------------------------------------------------
// Edge elements
// ---------------------------------------------------------------------------

// Edge::update (src/ClothUnit.h:134-154)
yo_cloth::EdgeGeo yo_cloth::edge_geo(const Edge& e) const {
    EdgeGeo g;
    g.c = e.type == EDGE_WARP ? 0 : 1;
    g.duv = uv[e.n1][g.c] - uv[e.n0][g.c];
    g.dx = sub(x[e.n1], x[e.n0]);
    g.w = scl(1.0 / g.duv, g.dx);
    g.l = norm(g.dx);
    g.d = scl(1.0 / g.l, g.dx);
    g.P = proj(g.d);
    g.qd = {v[e.n0][0], v[e.n0][1], v[e.n0][2], uvd[e.n0][g.c], v[e.n1][0], v[e.n1][1], v[e.n1][2], uvd[e.n1][g.c]};
    g.rho = p.rho[e.yarn];
    g.ks = ks[e.yarn];
    return g;
}

// The placement pattern shared by Mass, Stretch_D, Inertia_D_*, MDotVDot_D_* for one edge:
// LL four blocks, then LE/EL/EE self blocks of each interior end, then the cross blocks
// when both ends are interior (e.g. src/PhysicsCloth.cpp:1529-1599).
void yo_cloth::edge_blocks(SpAcc& S, const Edge& e, const M3& L0L0, const M3& L0L1, const M3& L1L0, const M3& L1L1,
                           const V3& L0E0, const V3& L0E1, const V3& L1E0, const V3& L1E1,
                           const V3& E0L0, const V3& E0L1, const V3& E1L0, const V3& E1L1,
                           double E0E0, double E0E1, double E1E0, double E1E1) {
    int c = e.type == EDGE_WARP ? 0 : 1;
    addLL(S, e.n0, e.n0, L0L0);
    addLL(S, e.n0, e.n1, L0L1);
    addLL(S, e.n1, e.n0, L1L0);
    addLL(S, e.n1, e.n1, L1L1);
    if (!is_edge[e.n0]) {
        addEL(S, e.n0, e.n0, c, E0L0);
        addLE(S, e.n0, e.n0, c, L0E0);
        addEE(S, e.n0, e.n0, c, E0E0);
    }
    if (!is_edge[e.n1]) {
        addEL(S, e.n1, e.n1, c, E1L1);
        addLE(S, e.n1, e.n1, c, L1E1);
        addEE(S, e.n1, e.n1, c, E1E1);
    }
    if (!is_edge[e.n0] && !is_edge[e.n1]) {
        addLE(S, e.n1, e.n0, c, L1E0);   // find_index(n1, n0, EL)
        addLE(S, e.n0, e.n1, c, L0E1);   // find_index(n0, n1, EL)
        addEL(S, e.n1, e.n0, c, E0L1);   // row E(n0), col L(n1)
        addEL(S, e.n0, e.n1, c, E1L0);   // row E(n1), col L(n0)
        addEE(S, e.n0, e.n1, c, E0E1);
        addEE(S, e.n1, e.n0, c, E1E0);
    }
}

// Mass (src/PhysicsCloth.cpp:250-338)
void yo_cloth::Mass(const Edge& e, const EdgeGeo& g) {
    double coeff = (1.0 / 6.0) * g.rho * g.duv;
    M3 L0L0 = mscl(2.0 * coeff, eye3()), L0L1 = mscl(coeff, eye3());
    V3 L0E0 = scl(-2.0 * coeff, g.w), E0L0 = scl(-2.0 * coeff, g.w);
    double E0E0 = 2.0 * coeff * dot(g.w, g.w);
    V3 L0E1 = scl(-1.0 * coeff, g.w), E1L0 = scl(-1.0 * coeff, g.w);
    double E0E1 = 0.5 * E0E0;
    edge_blocks(M, e, L0L0, L0L1, L0L1, L0L0, L0E0, L0E1, L0E1, L0E0, E0L0, E1L0, E1L0, E0L0,
                E0E0, E0E1, E0E1, E0E0);
}

// Gravity_F (src/PhysicsCloth.cpp:340-362)
void yo_cloth::Gravity_F(const Edge& e, const EdgeGeo& g, int t) {
    V3 G = v3(p.G[0], p.G[1], p.G[2]);
    V3 fl = scl(-0.5 * g.rho * g.duv, G);
    addFL(e.n0, fl); addFL(e.n1, fl);
    double gx = dot(G, add(x[e.n0], x[e.n1]));
    if (!is_edge[e.n0]) addFE(e.n0, g.c, -(-0.5) * g.rho * gx, t);
    if (!is_edge[e.n1]) addFE(e.n1, g.c, -0.5 * g.rho * gx, t);
}

// Gravity_D (src/PhysicsCloth.cpp:364-425)
void yo_cloth::Gravity_D(const Edge& e, const EdgeGeo& g) {
    V3 G = v3(p.G[0], p.G[1], p.G[2]);
    V3 L0E0 = scl(0.5 * g.rho, G), L1E0 = L0E0, L0E1 = neg(L0E0), L1E1 = neg(L0E0);
    V3 E0L0 = L0E0, E0L1 = E0L0, E1L0 = L0E1, E1L1 = E1L0;
    int c = g.c;
    if (!is_edge[e.n0]) { addEL(Kx, e.n0, e.n0, c, E0L0); addLE(Kx, e.n0, e.n0, c, L0E0); }
    if (!is_edge[e.n1]) { addEL(Kx, e.n1, e.n1, c, E1L1); addLE(Kx, e.n1, e.n1, c, L1E1); }
    if (!is_edge[e.n0] && !is_edge[e.n1]) {
        addLE(Kx, e.n1, e.n0, c, L1E0);
        addLE(Kx, e.n0, e.n1, c, L0E1);
        addEL(Kx, e.n1, e.n0, c, E0L1);
        addEL(Kx, e.n0, e.n1, c, E1L0);
    }
}

// Stretch_F (src/PhysicsCloth.cpp:427-464); the StrhBend_NC_* accumulators are never
// read (src/PhysicsCloth.cpp:2606-2617 is commented out) and are not kept.
void yo_cloth::Stretch_F(const Edge& e, const EdgeGeo& g, int t) {
    double nw = norm(g.w);
    V3 f = scl(g.ks * (nw - 1), g.d);
    addFL(e.n0, f); addFL(e.n1, neg(f));
    if (e.type == EDGE_WARP) { addSBC(e.n0, f); addSBC(e.n1, neg(f)); }
    else { addSBC(e.n0, neg(f)); addSBC(e.n1, f); }
    if (!is_edge[e.n0]) addFE(e.n0, g.c, -0.5 * g.ks * (std::pow(nw, 2) - 1), t);
    if (!is_edge[e.n1]) addFE(e.n1, g.c, 0.5 * g.ks * (std::pow(nw, 2) - 1), t);
}

// Stretch_D (src/PhysicsCloth.cpp:466-577)
void yo_cloth::Stretch_D(const Edge& e, const EdgeGeo& g) {
    double nw = norm(g.w);
    M3 L1L1 = msub(mscl(g.ks / g.l, g.P), mscl(g.ks / g.duv, eye3()));
    M3 L1L0 = mneg(L1L1);
    double E1E1 = -g.ks * (std::pow(nw, 2) / g.duv);
    double E1E0 = -E1E1;
    V3 L1E1 = scl(g.ks * (nw / g.duv), g.d), L1E0 = neg(L1E1);
    V3 E1L1 = scl(g.ks / g.duv, g.w), E1L0 = neg(E1L1);
    edge_blocks(Kx, e, L1L1, L1L0, L1L0, L1L1, L1E1, L1E0, L1E0, L1E1, E1L1, E1L0, E1L0, E1L1,
                E1E1, E1E0, E1E0, E1E1);
    // StrhBend_C_LE: only the self entries LE(k, e(k)) are read (src/PhysicsCloth.cpp:2588-2603)
    double sg = e.type == EDGE_WARP ? 1.0 : -1.0;
    if (!is_edge[e.n0]) addSBCLE(e.n0, g.c, scl(sg, L1E1));
    if (!is_edge[e.n1]) addSBCLE(e.n1, g.c, scl(sg, L1E1));
}

// ParaColli_F / ParaColli_D (src/PhysicsCloth.cpp:989-1069)
void yo_cloth::ParaColli_F(const Edge& e, const EdgeGeo& g, int t) {
    double d = (flip[e.n0] != flip[e.n1]) ? 4.0 * p.R : 2.0 * p.R;
    double Fu0 = p.kc * p.L * relu(d - g.duv);
    if (!is_edge[e.n0]) addFE(e.n0, g.c, Fu0, t);
    if (!is_edge[e.n1]) addFE(e.n1, g.c, -Fu0, t);
}
void yo_cloth::ParaColli_D(const Edge& e, const EdgeGeo& g) {
    double d = (flip[e.n0] != flip[e.n1]) ? 4.0 * p.R : 2.0 * p.R;
    double val = -p.kc * p.L * relu(d - g.duv);   // the value, not its derivative (Q6)
    if (!is_edge[e.n0]) addEE(Kx, e.n0, e.n0, g.c, val);
    if (!is_edge[e.n1]) addEE(Kx, e.n1, e.n1, g.c, val);
    if (!is_edge[e.n0] && !is_edge[e.n1]) { addEE(Kx, e.n0, e.n1, g.c, val); addEE(Kx, e.n1, e.n0, g.c, val); }
}

namespace {
// M_block(w) and its derivative blocks used by Inertia_* and MDotVDot_*
M4 mass_block(const V3& w) { return blk(eye3(), neg(w), neg(w), dot(w, w)); }
}  // namespace

// Inertia_F (src/PhysicsCloth.cpp:1071-1186)
void yo_cloth::Inertia_F(const Edge& e, const EdgeGeo& g, int t) {
    double D = g.duv, r = g.rho;
    double wTw = dot(g.w, g.w);
    M8 M_matrix = kron2(mass_block(g.w), 1.0);
    V3 T_x0, T_x1;
    for (int a = 0; a < 3; ++a) {
        V3 w_x0_a = scl(-1.0 / D, unit(a));
        double wTw_x0_a = (-2.0 / (D * D)) * g.dx[a];
        M8 M_x0_a = kron2(blk(zero3(), neg(w_x0_a), neg(w_x0_a), wTw_x0_a), 1.0 / 6.0 * r * D);
        M8 M_x1_a = m8scl(-1.0, M_x0_a);
        T_x0[a] = 0.5 * quad(g.qd, M_x0_a);
        T_x1[a] = 0.5 * quad(g.qd, M_x1_a);
    }
    M4 M_uv0_block = blk(zero3(), scl(-1.0 / D, g.w), scl(-1.0 / D, g.w), (2.0 / D) * wTw);
    M8 M_uv0 = m8add(m8scl(-(1.0 / 6.0) * r, M_matrix), kron2(M_uv0_block, (1.0 / 6.0) * r * D));
    M8 M_uv1 = m8scl(-1.0, M_uv0);
    double T_uv0 = 0.5 * quad(g.qd, M_uv0), T_uv1 = 0.5 * quad(g.qd, M_uv1);
    addFL(e.n0, T_x0); addFL(e.n1, T_x1);
    if (!is_edge[e.n0]) addFE(e.n0, g.c, T_uv0, t);
    if (!is_edge[e.n1]) addFE(e.n1, g.c, T_uv1, t);
}

// Inertia_D_Pos (src/PhysicsCloth.cpp:1188-1600)
void yo_cloth::Inertia_D_Pos(const Edge& e, const EdgeGeo& g) {
    double D = g.duv, r = g.rho;
    double wTw = dot(g.w, g.w);
    const V8& q = g.qd;
    M8 M_matrix_x0[3];
    for (int a = 0; a < 3; ++a) {
        V3 w_x0_a = scl(-1.0 / D, unit(a));
        M_matrix_x0[a] = kron2(blk(zero3(), neg(w_x0_a), neg(w_x0_a), (-2.0 / (D * D)) * g.dx[a]), 1.0);
    }
    // d2/dx0_a dx0_a: only the uv-uv entry 2/D^2, scaled by rho/6 (src/PhysicsCloth.cpp:1246-1252)
    M4 b11 = blk(zero3(), v3(0, 0, 0), v3(0, 0, 0), 2.0 / (D * D));
    M8 M_x0_1_x0_1 = kron2(b11, (1.0 / 6.0) * r);
    double T = 0.5 * quad(q, M_x0_1_x0_1);
    M3 L0L0 = mscl(T, eye3()), L0L1 = mscl(-T, eye3()), L1L0 = mscl(-T, eye3()), L1L1 = mscl(T, eye3());
    // mixed x-uv terms (src/PhysicsCloth.cpp:1307-1421); the lower-left sign of the
    // derivative block is taken as written at src/PhysicsCloth.cpp:1362
    V3 L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1;
    M8 M_matrix_x0_uv_0[3];
    for (int a = 0; a < 3; ++a) {
        V3 w_x0_a_uv_0 = scl(1.0 / (D * D), unit(a));
        double wTw_x0_a_uv_0 = (-4.0 / (D * D)) * g.w[a];
        M_matrix_x0_uv_0[a] = kron2(blk(zero3(), neg(w_x0_a_uv_0), w_x0_a_uv_0, wTw_x0_a_uv_0), 1.0);
        M8 M_x0_a_uv_0 = m8add(m8scl((-1.0 / 6.0) * r, M_matrix_x0[a]), m8scl((1.0 / 6.0) * r * D, M_matrix_x0_uv_0[a]));
        M8 M_x0_a_uv_1 = m8scl(-1.0, M_x0_a_uv_0);
        M8 M_x1_a_uv_0 = m8scl(-1.0, M_x0_a_uv_0);
        M8 M_x1_a_uv_1 = m8scl(-1.0, M_x1_a_uv_0);
        L0E0[a] = 0.5 * quad(q, M_x0_a_uv_0);
        L0E1[a] = 0.5 * quad(q, M_x0_a_uv_1);
        L1E0[a] = 0.5 * quad(q, M_x1_a_uv_0);
        L1E1[a] = 0.5 * quad(q, M_x1_a_uv_1);
        M8 M_uv_0_x0_a = m8add(m8scl((-1.0 / 6.0) * r, M_matrix_x0[a]), m8scl((1.0 / 6.0) * r * D, m8scl(-1.0, M_matrix_x0_uv_0[a])));
        M8 M_uv_0_x1_a = m8scl(-1.0, M_uv_0_x0_a);
        M8 M_uv_1_x0_a = m8scl(-1.0, M_uv_0_x0_a);
        M8 M_uv_1_x1_a = m8scl(-1.0, M_uv_1_x0_a);
        E0L0[a] = 0.5 * quad(q, M_uv_0_x0_a);
        E0L1[a] = 0.5 * quad(q, M_uv_0_x1_a);
        E1L0[a] = 0.5 * quad(q, M_uv_1_x0_a);
        E1L1[a] = 0.5 * quad(q, M_uv_1_x1_a);
    }
    M8 M_matrix_uv_0 = kron2(blk(zero3(), scl(-1.0 / D, g.w), scl(-1.0 / D, g.w), (2.0 / D) * wTw), 1.0);
    V3 w_uv_0_uv_0 = scl(2.0 / (D * D), g.w);
    double wTw_uv_0_uv_0 = (6.0 / (D * D)) * wTw;
    M8 M_matrix_uv_0_uv_0 = kron2(blk(zero3(), neg(w_uv_0_uv_0), neg(w_uv_0_uv_0), wTw_uv_0_uv_0), 1.0);
    M8 M_uv_0_uv_0 = m8add(m8add(m8scl((-1.0 / 6.0) * r, M_matrix_uv_0), m8scl(-(1.0 / 6.0) * r, M_matrix_uv_0)),
                           m8scl((1.0 / 6.0) * r * D, M_matrix_uv_0_uv_0));
    double E0E0 = 0.5 * quad(q, M_uv_0_uv_0);
    double E0E1 = -E0E0, E1E0 = -E0E0, E1E1 = E0E0;
    edge_blocks(Kx, e, L0L0, L0L1, L1L0, L1L1, L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1,
                E0E0, E0E1, E1E0, E1E1);
}

// Inertia_D_Vel (src/PhysicsCloth.cpp:1602-1901): block entry (i, j) is (M_i qdot)[j]
void yo_cloth::Inertia_D_Vel(const Edge& e, const EdgeGeo& g) {
    double D = g.duv, r = g.rho;
    double wTw = dot(g.w, g.w);
    const V8& q = g.qd;
    M8 M_matrix = kron2(mass_block(g.w), 1.0);
    V8 Mx0q[3], Mx1q[3];
    for (int a = 0; a < 3; ++a) {
        V3 w_x0_a = scl(-1.0 / D, unit(a));
        M8 M_x0_a = kron2(blk(zero3(), neg(w_x0_a), neg(w_x0_a), (-2.0 / (D * D)) * g.dx[a]), (1.0 / 6.0) * r * D);
        Mx0q[a] = m8v(M_x0_a, q);
        Mx1q[a] = m8v(m8scl(-1.0, M_x0_a), q);
    }
    M4 M_uv0_block = blk(zero3(), scl(-1.0 / D, g.w), scl(-1.0 / D, g.w), (2.0 / D) * wTw);
    M8 M_uv0 = m8add(m8scl(-(1.0 / 6.0) * r, M_matrix), kron2(M_uv0_block, (1.0 / 6.0) * r * D));
    V8 Muv0q = m8v(M_uv0, q), Muv1q = m8v(m8scl(-1.0, M_uv0), q);
    M3 L0L0, L0L1, L1L0, L1L1;
    V3 L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1;
    for (int a = 0; a < 3; ++a) {
        for (int b2 = 0; b2 < 3; ++b2) {
            L0L0[a][b2] = Mx0q[a][b2]; L0L1[a][b2] = Mx0q[a][4 + b2];
            L1L0[a][b2] = Mx1q[a][b2]; L1L1[a][b2] = Mx1q[a][4 + b2];
        }
        L0E0[a] = Mx0q[a][3]; L0E1[a] = Mx0q[a][7]; L1E0[a] = Mx1q[a][3]; L1E1[a] = Mx1q[a][7];
        E0L0[a] = Muv0q[a]; E0L1[a] = Muv0q[4 + a]; E1L0[a] = Muv1q[a]; E1L1[a] = Muv1q[4 + a];
    }
    edge_blocks(Kv, e, L0L0, L0L1, L1L0, L1L1, L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1,
                Muv0q[3], Muv0q[7], Muv1q[3], Muv1q[7]);
}

namespace {
struct MDot { double Dd; V3 wd; double wTwd; M8 M_matrix, M_dot; };
MDot mdot_terms(const yo_cloth::EdgeGeo& g, const V3& dv, double Dd) {
    MDot m;
    double D = g.duv, r = g.rho;
    m.Dd = Dd;
    m.wd = scl(1.0 / (D * D), sub(scl(D, dv), scl(Dd, g.dx)));   // src/PhysicsCloth.cpp:1918
    m.wTwd = dot(g.w, m.wd) + dot(m.wd, g.w);
    m.M_matrix = kron2(mass_block(g.w), 1.0);
    M4 M_dot_block = blk(zero3(), neg(m.wd), neg(m.wd), m.wTwd);
    m.M_dot = m8add(m8scl((1.0 / 6.0) * r * Dd, m.M_matrix), kron2(M_dot_block, (1.0 / 6.0) * r * D));
    return m;
}
}  // namespace

// MDotVDot_F (src/PhysicsCloth.cpp:1903-1961)
void yo_cloth::MDotVDot_F(const Edge& e, const EdgeGeo& g, int t) {
    double Dd = uvd[e.n1][g.c] - uvd[e.n0][g.c];
    MDot m = mdot_terms(g, sub(v[e.n1], v[e.n0]), Dd);
    V8 MdV = m8v(m.M_dot, g.qd);
    addFL(e.n0, v3(-MdV[0], -MdV[1], -MdV[2]));
    addFL(e.n1, v3(-MdV[4], -MdV[5], -MdV[6]));
    if (!is_edge[e.n0]) addFE(e.n0, g.c, -MdV[3], t);
    if (!is_edge[e.n1]) addFE(e.n1, g.c, -MdV[7], t);
}

namespace {
// Slice the 8x8 column-stacked derivative matrix into the edge blocks
// (src/PhysicsCloth.cpp:2212-2230): Q[i][j] = column j (derivative variable), row i.
void slice_blocks(const V8 cols[8], M3& L0L0, M3& L0L1, M3& L1L0, M3& L1L1,
                  V3& L0E0, V3& L0E1, V3& L1E0, V3& L1E1, V3& E0L0, V3& E0L1, V3& E1L0, V3& E1L1,
                  double& E0E0, double& E0E1, double& E1E0, double& E1E1) {
    auto Q = [&](int i, int j) { return cols[j][i]; };
    for (int a = 0; a < 3; ++a) {
        for (int b2 = 0; b2 < 3; ++b2) {
            L0L0[a][b2] = Q(a, b2); L0L1[a][b2] = Q(a, 4 + b2);
            L1L0[a][b2] = Q(4 + a, b2); L1L1[a][b2] = Q(4 + a, 4 + b2);
        }
        E0L0[a] = Q(3, a); E0L1[a] = Q(3, 4 + a); E1L0[a] = Q(7, a); E1L1[a] = Q(7, 4 + a);
        L0E0[a] = Q(a, 3); L0E1[a] = Q(a, 7); L1E0[a] = Q(4 + a, 3); L1E1[a] = Q(4 + a, 7);
    }
    E0E0 = Q(3, 3); E0E1 = Q(3, 7); E1E0 = Q(7, 3); E1E1 = Q(7, 7);
}
}  // namespace

// MDotVDot_D_Pos (src/PhysicsCloth.cpp:1963-2308); added with + into F_*_Pos as written
void yo_cloth::MDotVDot_D_Pos(const Edge& e, const EdgeGeo& g) {
    double D = g.duv, r = g.rho;
    double Dd = uvd[e.n1][g.c] - uvd[e.n0][g.c];
    V3 dv = sub(v[e.n1], v[e.n0]);
    MDot m = mdot_terms(g, dv, Dd);
    const V8& q = g.qd;
    V8 cols[8];
    for (int end = 0; end < 2; ++end) {
        double sgn = end == 0 ? -1.0 : 1.0;
        for (int a = 0; a < 3; ++a) {
            V3 w_x_a = scl(sgn / D, unit(a));
            double wTw_x_a = (sgn * 2.0 / (D * D)) * g.dx[a];
            V3 w_dot_x_a = scl(-sgn * (Dd / (D * D)), unit(a));
            double wTw_dot_x_a = 2 * dot(w_x_a, m.wd) + 2 * dot(g.w, w_dot_x_a);
            M8 M_matrix_x = kron2(blk(zero3(), neg(w_x_a), neg(w_x_a), wTw_x_a), 1.0);
            M8 M_dot_matrix_x = kron2(blk(zero3(), neg(w_dot_x_a), neg(w_dot_x_a), wTw_dot_x_a), 1.0);
            cols[4 * end + a] = m8v(m8add(m8scl((1.0 / 6.0) * r * Dd, M_matrix_x), m8scl((1.0 / 6.0) * r * D, M_dot_matrix_x)), q);
        }
        V3 w_uv = scl(sgn, scl(1.0 / D, g.w));                       // w_uv0 = -w/D, w_uv1 = w/D
        double wTw_uv = sgn * (2.0 / D) * dot(g.w, g.w);
        V3 w_dot_uv0 = sub(scl(1.0 / (D * D), dv), scl(4.0 * Dd / (D * D * D), g.dx));
        V3 w_dot_uv = end == 0 ? w_dot_uv0 : neg(w_dot_uv0);
        double wTw_dot_uv = 2 * dot(w_uv, m.wd) + 2 * dot(w_dot_uv, g.w);
        M8 M_matrix_uv = kron2(blk(zero3(), w_uv, w_uv, wTw_uv), 1.0);
        M8 M_dot_matrix_uv = kron2(blk(zero3(), neg(w_dot_uv), neg(w_dot_uv), wTw_dot_uv), 1.0);
        M8 Muv = m8add(m8add(m8scl((1.0 / 6.0) * r * Dd, M_matrix_uv), m8scl(-(1.0 / 6.0) * r, m.M_dot)),
                       m8scl((1.0 / 6.0) * r * D, M_dot_matrix_uv));
        cols[4 * end + 3] = m8v(Muv, q);
    }
    M3 L0L0, L0L1, L1L0, L1L1; V3 L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1;
    double E0E0, E0E1, E1E0, E1E1;
    slice_blocks(cols, L0L0, L0L1, L1L0, L1L1, L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1, E0E0, E0E1, E1E0, E1E1);
    edge_blocks(Kx, e, L0L0, L0L1, L1L0, L1L1, L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1, E0E0, E0E1, E1E0, E1E1);
}

// MDotVDot_D_Vel (src/PhysicsCloth.cpp:2310-2577); subtracted from F_*_Vel as written
void yo_cloth::MDotVDot_D_Vel(const Edge& e, const EdgeGeo& g) {
    double D = g.duv, r = g.rho;
    double Dd = uvd[e.n1][g.c] - uvd[e.n0][g.c];
    MDot m = mdot_terms(g, sub(v[e.n1], v[e.n0]), Dd);
    const V8& q = g.qd;
    V8 cols[8];
    for (int end = 0; end < 2; ++end) {
        double sgn = end == 0 ? -1.0 : 1.0;
        for (int a = 0; a < 3; ++a) {
            V3 w_dot_xd_a = scl(sgn / D, unit(a));
            double wTw_dot_xd_a = 2 * dot(g.w, w_dot_xd_a);
            M8 Mdm = kron2(blk(zero3(), neg(w_dot_xd_a), neg(w_dot_xd_a), wTw_dot_xd_a), 1.0);
            V8 t1 = m8v(Mdm, q);
            V8 col;
            for (int i = 0; i < 8; ++i) col[i] = (1.0 / 6.0) * r * D * t1[i] + m.M_dot[i][4 * end + a];
            cols[4 * end + a] = col;
        }
        V3 w_dot_uvd = scl(-sgn / D, g.w);        // uv0: (1/D) w, uv1: (-1/D) w
        double wTw_dot_uvd = 2 * dot(g.w, w_dot_uvd);
        M8 Mdm = kron2(blk(zero3(), neg(w_dot_uvd), neg(w_dot_uvd), wTw_dot_uvd), 1.0);
        M8 Mtot = m8add(m8scl(sgn * (1.0 / 6.0) * r, m.M_matrix), m8scl((1.0 / 6.0) * r * D, Mdm));
        V8 t1 = m8v(Mtot, q);
        V8 col;
        for (int i = 0; i < 8; ++i) col[i] = t1[i] + m.M_dot[i][4 * end + 3];
        cols[4 * end + 3] = col;
    }
    M3 L0L0, L0L1, L1L0, L1L1; V3 L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1;
    double E0E0, E0E1, E1E0, E1E1;
    slice_blocks(cols, L0L0, L0L1, L1L0, L1L1, L0E0, L0E1, L1E0, L1E1, E0L0, E0L1, E1L0, E1L1, E0E0, E0E1, E1E0, E1E1);
    edge_blocks(Kv, e, mneg(L0L0), mneg(L0L1), mneg(L1L0), mneg(L1L1), neg(L0E0), neg(L0E1), neg(L1E0), neg(L1E1),
                neg(E0L0), neg(E0L1), neg(E1L0), neg(E1L1), -E0E0, -E0E1, -E1E0, -E1E1);
}

// ---------------------------------------------------------------------------
// Bending (src/PhysicsCloth.cpp:579-987)
// ---------------------------------------------------------------------------

namespace {
struct BendGeo { double duv, l1, l2, theta; V3 d1, d2; M3 P1, P2; };
BendGeo bend_geo(const V3& x2, const V3& x0, const V3& x1, double duv) {  // Bend_Seg::update (src/ClothUnit.h:197-210)
    BendGeo g;
    g.duv = duv;
    V3 a = sub(x1, x0), b = sub(x2, x0);
    g.l1 = norm(a); g.l2 = norm(b);
    g.d1 = scl(1.0 / g.l1, a); g.d2 = scl(1.0 / g.l2, b);
    g.P1 = proj(g.d1); g.P2 = proj(g.d2);
    double cth = -dot(g.d1, g.d2);
    cth = std::min(std::max(cth, -1.0 + EPS), 1.0 - EPS);
    g.theta = std::acos(cth);
    return g;
}
struct BendOut { V3 F_x1, F_x2, F_x0; double F_u2, F_u1; M3 LL[3][3]; V3 L1E1, L2E1, E1L1, E1L2; double E1E1; };
// index order inside LL: 0 = n2, 1 = n0, 2 = n1
BendOut bend_eval(const BendGeo& b, double kb) {
    BendOut o;
    V3 z = {0, 0, 0};
    o.F_x1 = z; o.F_x2 = z;
    double th = b.theta, s = std::sin(th), c = std::cos(th), D = b.duv;
    M3 L1L1 = zero3(), L1L2 = zero3(), L2L1 = zero3(), L2L2 = zero3();
    o.L1E1 = z; o.L2E1 = z; o.E1L1 = z; o.E1L2 = z;
    if (s != 0) {
        o.F_x1 = scl((-2 * kb * th) / (b.l1 * D * s), mv(b.P1, b.d2));
        o.F_x2 = scl((-2 * kb * th) / (b.l2 * D * s), mv(b.P2, b.d1));
        V3 P1d2 = mv(b.P1, b.d2), P2d1 = mv(b.P2, b.d1);
        L1L1 = mscl((2 * kb) / (std::pow(b.l1, 2) * D * s),
                    msub(mscl(th, madd(madd(madd(outer(P1d2, b.d1), mscl(c / std::pow(s, 2), mm(outer(P1d2, b.d2), b.P1))),
                                            mscl(c, b.P1)), mm(outer(b.d1, b.d2), b.P1))),
                         mscl(1 / s, mm(outer(P1d2, b.d2), b.P1))));
        L1L2 = mscl((-2 * kb) / (b.l1 * b.l2 * D * s),
                    mm(madd(mscl(th, msub(b.P1, mscl(c / std::pow(s, 2), outer(P1d2, b.d1)))), mscl(1 / s, outer(P1d2, b.d1))), b.P2));
        L2L1 = mscl((-2 * kb) / (b.l1 * b.l2 * D * s),
                    mm(madd(mscl(th, msub(b.P2, mscl(c / std::pow(s, 2), outer(P2d1, b.d2)))), mscl(1 / s, outer(P2d1, b.d2))), b.P1));
        L2L2 = mscl((2 * kb) / (std::pow(b.l2, 2) * D * s),
                    msub(mscl(th, madd(madd(madd(outer(P2d1, b.d2), mscl(c / std::pow(s, 2), mm(outer(P2d1, b.d1), b.P2))),
                                            mscl(c, b.P2)), mm(outer(b.d2, b.d1), b.P2))),
                         mscl(1 / s, mm(outer(P2d1, b.d1), b.P2))));
        o.L1E1 = scl((2 * kb * th) / (b.l1 * std::pow(D, 2) * s), P1d2);
        o.L2E1 = scl((2 * kb * th) / (b.l2 * std::pow(D, 2) * s), P2d1);
        o.E1L1 = scl((2 * kb * th) / (b.l1 * std::pow(D, 2) * s), vm(b.d2, b.P1));
        o.E1L2 = scl((2 * kb * th) / (b.l2 * std::pow(D, 2) * s), vm(b.d1, b.P2));
    }
    o.F_u2 = (-kb * th) / std::pow(D, 2);
    o.F_u1 = -o.F_u2;
    o.F_x0 = neg(add(o.F_x1, o.F_x2));
    M3 L1L0 = mneg(madd(L1L1, L1L2)), L2L0 = mneg(madd(L2L1, L2L2));
    M3 L0L1 = mneg(madd(L1L1, L2L1)), L0L2 = mneg(madd(L1L2, L2L2));
    M3 L0L0 = mneg(madd(L1L0, L2L0));
    o.LL[0][0] = L2L2; o.LL[0][1] = L2L0; o.LL[0][2] = L2L1;
    o.LL[1][0] = L0L2; o.LL[1][1] = L0L0; o.LL[1][2] = L0L1;
    o.LL[2][0] = L1L2; o.LL[2][1] = L1L0; o.LL[2][2] = L1L1;
    o.E1E1 = (-2 * kb * std::pow(th, 2)) / std::pow(D, 3);
    return o;
}
}  // namespace

void yo_cloth::Bending(const Bend& bd, int t) {
    int c = bd.type == EDGE_WARP ? 0 : 1;
    double duv = uv[bd.n1][c] - uv[bd.n2][c];
    BendGeo g = bend_geo(x[bd.n2], x[bd.n0], x[bd.n1], duv);
    BendOut o = bend_eval(g, kb[bd.yarn]);
    // Bending_F (src/PhysicsCloth.cpp:579-641)
    addFL(bd.n2, o.F_x2); addFL(bd.n0, o.F_x0); addFL(bd.n1, o.F_x1);
    if (!is_edge[bd.n2]) addFE(bd.n2, c, o.F_u2, t);
    if (!is_edge[bd.n0]) addFE(bd.n0, c, 0.0, t);
    if (!is_edge[bd.n1]) addFE(bd.n1, c, o.F_u1, t);
    // Bending_D (src/PhysicsCloth.cpp:643-822)
    int ids[3] = {bd.n2, bd.n0, bd.n1};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) addLL(Kx, ids[i], ids[j], o.LL[i][j]);
    V3 L1E2 = neg(o.L1E1); (void)L1E2;
    V3 L2E2 = neg(o.L2E1);
    V3 L0E1 = neg(add(o.L1E1, o.L2E1));
    V3 L0E2 = neg(L0E1);
    V3 E2L2 = neg(o.E1L2);
    V3 E1L0 = neg(add(o.E1L1, o.E1L2));
    V3 E2L0 = neg(E1L0);
    double E2E2 = o.E1E1;
    if (!is_edge[bd.n2]) { addLE(Kx, bd.n2, bd.n2, c, L2E2); addEL(Kx, bd.n2, bd.n2, c, E2L2); addEE(Kx, bd.n2, bd.n2, c, E2E2); }
    if (!is_edge[bd.n1]) { addLE(Kx, bd.n1, bd.n1, c, o.L1E1); addEL(Kx, bd.n1, bd.n1, c, o.E1L1); addEE(Kx, bd.n1, bd.n1, c, o.E1E1); }
    if (!is_edge[bd.n2] && !is_edge[bd.n0]) {
        addLE(Kx, bd.n0, bd.n2, c, L0E2);   // find_index(n0, n2, EL)
        addEL(Kx, bd.n2, bd.n0, c, E2L0);   // find_index(n2, n0, EL): row E(n0), col L(n2)
    }
    if (!is_edge[bd.n0] && !is_edge[bd.n1]) {
        addLE(Kx, bd.n0, bd.n1, c, L0E1);
        addEL(Kx, bd.n0, bd.n1, c, E1L0);   // row E(n1), col L(n0)
    }
}

// Crimp_Bending_F / _D (src/PhysicsCloth.cpp:824-987): same element on the +-R*n copies,
// feeding only the compression accumulators.  The copies keep their construction-time
// Eulerian coordinates (src/Cloth.cpp:191-192 copies the node map once).
void yo_cloth::Crimp_Bending(const Bend& bd) {
    int c = bd.type == EDGE_WARP ? 0 : 1;
    const std::vector<V3>& xc = bd.type == EDGE_WARP ? xwarp : xweft;
    double duv = uv_crimp[bd.n1][c] - uv_crimp[bd.n2][c];
    BendGeo g = bend_geo(xc[bd.n2], xc[bd.n0], xc[bd.n1], duv);
    BendOut o = bend_eval(g, kb[bd.yarn]);
    double sg = bd.type == EDGE_WARP ? 1.0 : -1.0;
    addSBC(bd.n2, scl(sg, o.F_x2));
    addSBC(bd.n0, scl(sg, o.F_x0));
    addSBC(bd.n1, scl(sg, o.F_x1));
    V3 L2E2 = neg(o.L2E1);
    if (!is_edge[bd.n2]) addSBCLE(bd.n2, c, scl(sg, L2E2));
    if (!is_edge[bd.n1]) addSBCLE(bd.n1, c, scl(sg, o.L1E1));
    // the remaining crimp LE writes (src/PhysicsCloth.cpp:944-986) land on non-self
    // entries that ConSlideFriction never reads
}

// ---------------------------------------------------------------------------
// Friction, shear, wind, handles
// ---------------------------------------------------------------------------

// One component (u or v) of ConSlideFriction's arithmetic (src/PhysicsCloth.cpp:2619-2633):
// inputs uv - uvbar, the compression and its EE input, the no-slide E force, the
// thread-partial E force (Q2), the no-slide EE Jacobian diagonals, the Eulerian velocity
// and the cloth's mu, kf, df.
struct FricIn { double dlt, comp, compEE, FNS, Fpt, Fpos, Fvel, uvd, mu, kf, df; };
struct FricOut { double is_slide, fric, fpos, fvel; };
FricOut friction_eval(const FricIn& in) {
    FricOut o;
    const double comp = in.comp, mu = in.mu, kf = in.kf, df = in.df;
    const double lim = comp * mu;
    const double th = std::tanh(10 * in.dlt);
    o.is_slide = sigm(1e4 * (std::fabs(in.FNS) - lim));
    o.fric = (-0.5 * (kf * in.dlt - th * comp * mu) * std::tanh(10 * (lim - in.Fpt)) + 0.5 * (kf * in.dlt + th * comp * mu)) -
             df * in.uvd;
    const double tl = std::tanh(10 * (lim - in.FNS));
    o.fpos = (-0.5 * (kf - (1.0 - std::pow(th, 2)) * mu * comp - th * mu * in.compEE)) * tl -
             (0.5 * (kf * in.dlt - th * mu * comp)) * (1.0 - std::pow(tl, 2)) * (in.Fpos - mu * in.compEE) -
             (0.5 * (kf + (1.0 - std::pow(th, 2)) * mu * comp + th * mu * in.compEE));
    const double tq = std::tanh(10 * lim - in.FNS);   // (10*fric_lim - F)[k] as written (Q3)
    o.fvel = (-0.5 * (kf * in.dlt - th * mu * comp)) * (1.0 - std::pow(tq, 2)) * in.Fvel - df;
    return o;
}

// ConSlideFriction (src/PhysicsCloth.cpp:2579-2641) for the interior crossing of node k
void yo_cloth::ConSlideFriction(int k) {
    int e = eofk(k);
    double Fpos[2], Fvel[2], dlt[2], compEE[2];
    for (int c = 0; c < 2; ++c) {
        int r = 3 * N + 2 * e + c;
        Fpos[c] = Kx.get(r, r);
        Fvel[c] = Kv.get(r, r);
        dlt[c] = uv[k][c] - uvbar[k][c];
        compEE[c] = 0.5 * dot(nrm[k], StrhBendC_LEself[k][c]);
    }
    double compression = relu(0.5 * dot(nrm[k], StrhBendC_L[k]));
    comp[e] = compression;
    for (int c = 0; c < 2; ++c) {
        FricIn in{dlt[c], compression, compEE[c], FE_noslide[2 * e + c], FEpart[2 * e + c], Fpos[c], Fvel[c], uvd[k][c],
                  p.mu, p.kf, p.df};
        FricOut o = friction_eval(in);
        is_slide[2 * e + c] = o.is_slide;
        FE[2 * e + c] += o.fric;
        int r = 3 * N + 2 * e + c;
        Kx.add(r, r, o.fpos);
        Kv.add(r, r, o.fvel);
    }
}

// Shear_F / Shear_D (src/PhysicsCloth.cpp:2643-3089)
void yo_cloth::Shear_FD(const Shear& s) {
    const int c = 3;
    const double sig = 0.6;
    const double L = p.L, R = p.R, S = p.S;
    V3 a = sub(x[s.n1], x[s.n0]), b3 = sub(x[s.n3], x[s.n0]);
    double l1 = norm(a), l3 = norm(b3);
    V3 d1 = scl(1.0 / l1, a), d3 = scl(1.0 / l3, b3);
    M3 P1 = proj(d1), P3 = proj(d3);
    double phi = std::acos(dot(d1, d3));   // unclamped (src/ClothUnit.h:247)
    double comp_shear = comp[eofk(s.n0)];
    double kx = (comp_shear + 1.0) * S * R * R * PI;
    double sn = std::sin(phi), cs = std::cos(phi);
    V3 P1d3 = mv(P1, d3), P3d1 = mv(P3, d1);
    if (s.jam) {
        double phi_bar = PI / 2, phi_l = jamming_thrd;
        V3 phi_x1 = scl(-1.0 / (l1 * sn), P1d3);
        V3 phi_x3 = scl(-1.0 / (l3 * sn), P3d1);
        M3 phi_x1_x1 = mscl(1.0 / (std::pow(l1, 2) * sn),
                            madd(msub(madd(outer(P1d3, d1), mm(outer(d1, d3), P1)), mscl(cs / std::pow(sn, 2), mm(outer(P1d3, d3), P1))), mscl(cs, P1)));
        M3 phi_x3_x3 = mscl(1.0 / (std::pow(l3, 2) * sn),
                            madd(msub(madd(outer(P3d1, d3), mm(outer(d3, d1), P3)), mscl(cs / std::pow(sn, 2), mm(outer(P3d1, d1), P3))), mscl(cs, P3)));
        M3 phi_x1_x3 = mscl(-1.0 / (l1 * l3 * sn), madd(mscl(cs / std::pow(sn, 2), mm(outer(P1d3, d1), P3)), mm(P1, P3)));
        M3 phi_x3_x1 = mscl(-1.0 / (l3 * l1 * sn), madd(mscl(cs / std::pow(sn, 2), mm(outer(P3d1, d3), P3)), mm(P3, P1)));
        double gamma = (L * std::sqrt(2.0) - 2 * L * std::sin(0.5 * phi)) / R;
        V3 gamma_x1 = scl(-(L / R) * std::cos(0.5 * phi), phi_x1);
        V3 gamma_x3 = scl(-(L / R) * std::cos(0.5 * phi), phi_x3);
        M3 gamma_x1_x1 = msub(mscl((0.5 * (L / R) * std::sin(0.5 * phi)) / std::pow(l1 * sn, 2), mm(outer(P1d3, d3), P1)),
                              mscl((L / R) * std::cos(0.5 * phi), phi_x1_x1));
        M3 gamma_x3_x3 = msub(mscl((0.5 * (L / R) * std::sin(0.5 * phi)) / std::pow(l3 * sn, 2), mm(outer(P3d1, d1), P3)),
                              mscl((L / R) * std::cos(0.5 * phi), phi_x3_x3));
        M3 gamma_x1_x3 = msub(mscl((0.5 * (L / R) * std::sin(phi / 2)) / (l1 * l3 * std::pow(sn, 2)), mm(outer(P1d3, d1), P3)),
                              mscl((L / R) * std::cos(phi / 2), phi_x1_x3));
        M3 gamma_x3_x1 = msub(mscl((0.5 * (L / R) * std::sin(phi / 2)) / (l3 * l1 * std::pow(sn, 2)), mm(outer(P3d1, d3), P1)),
                              mscl((L / R) * std::cos(phi / 2), phi_x3_x1));
        double pl = phi - phi_l, pb = phi - phi_bar;
        double phi_s = phi * pl * pb;
        V3 phi_s_x1 = add(add(scl(pl * pb, phi_x1), scl(phi * pb, phi_x1)), scl(phi * pl, phi_x1));
        V3 phi_s_x3 = add(add(scl(pl * pb, phi_x3), scl(phi * pb, phi_x3)), scl(phi * pl, phi_x3));
        M3 o11 = outer(phi_x1, phi_x1), o33 = outer(phi_x3, phi_x3), o13 = outer(phi_x1, phi_x3), o31 = outer(phi_x3, phi_x1);
        // src/PhysicsCloth.cpp:2810-2848, term by term as written
        M3 phi_s_x1_x1 = mscl(pl * pb, phi_x1_x1);
        phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(pb, o11)); phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(pl, o11));
        phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(pb, o11)); phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(phi * pb, phi_x1_x1));
        phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(phi, o11)); phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(pl, o11));
        phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(phi, o11)); phi_s_x1_x1 = madd(phi_s_x1_x1, mscl(phi * pl, phi_x1_x1));
        M3 phi_s_x3_x3 = mscl(pl * pb, phi_x3_x3);
        phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(pb, o33)); phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(pl, o33));
        phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(pb, o33)); phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(phi * pb, phi_x3_x3));
        phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(phi, o33)); phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(pl, o33));
        phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(phi, o33)); phi_s_x3_x3 = madd(phi_s_x3_x3, mscl(phi * pl, phi_x3_x3));
        M3 phi_s_x1_x3 = mscl(pl * pb, phi_x1_x3);
        phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(pb, o13)); phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(pl, o13));
        phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(pb, o31)); phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(phi * pb, phi_x1_x3));
        phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(phi, o13)); phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(pl, o31));
        phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(phi, o31)); phi_s_x1_x3 = madd(phi_s_x1_x3, mscl(phi * pl, phi_x1_x3));
        M3 phi_s_x3_x1 = mscl(pl * pb, phi_x3_x1);
        phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(pb, o31)); phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(pl, o31));
        phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(pb, o31)); phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(phi * pb, phi_x3_x1));
        phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(phi, o31)); phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(pl, o31));
        phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(phi, o31)); phi_s_x3_x1 = madd(phi_s_x3_x1, mscl(phi * pl, phi_x3_x1));

        double pb5 = std::pow(phi_bar, 5), pb4 = std::pow(phi_bar, 4);
        double cfx_num = pb5 * pl;
        double cfx_den = std::pow(phi_s, 2) + pb4 * sig * sig;
        double cfx = cfx_num / cfx_den;
        double cf = std::tanh(cfx);
        V3 cfx_num_x1 = scl(pb5, phi_x1), cfx_num_x3 = scl(pb5, phi_x3);
        V3 cfx_den_x1 = scl(2 * phi_s, phi_s_x1), cfx_den_x3 = scl(2 * phi_s, phi_s_x3);
        V3 cfx_x1_num = sub(scl(cfx_den, cfx_num_x1), scl(cfx_num, cfx_den_x1));
        double cfx_x1_den = std::pow(cfx_den, 2);
        V3 cfx_x1 = scl(1.0 / cfx_x1_den, cfx_x1_num);
        V3 cf_x1 = scl(1.0 - std::pow(cf, 2), cfx_x1);
        V3 cfx_x3_num = sub(scl(cfx_den, cfx_num_x3), scl(cfx_num, cfx_den_x3));
        double cfx_x3_den = std::pow(cfx_den, 2);
        V3 cfx_x3 = scl(1.0 / cfx_x3_den, cfx_x3_num);
        V3 cf_x3 = scl(1.0 - std::pow(cf, 2), cfx_x3);
        M3 cfx_num_x1_x1 = mscl(pb5, phi_x1_x1), cfx_num_x3_x3 = mscl(pb5, phi_x3_x3);
        M3 cfx_num_x1_x3 = mscl(pb5, phi_x1_x3), cfx_num_x3_x1 = mscl(pb5, phi_x3_x1);
        M3 cfx_x1_num_x1 = msub(msub(mscl(cfx_den, cfx_num_x1_x1), mscl(2 * cfx_num, outer(phi_s_x1, phi_s_x1))), mscl(2 * cfx_num * phi_s, phi_s_x1_x1));
        V3 cfx_x1_den_x1 = scl(2 * cfx_den, cfx_den_x1);
        M3 cfx_x1_x1 = mscl(1.0 / std::pow(cfx_x1_den, 2), msub(mscl(cfx_x1_den, cfx_x1_num_x1), outer(cfx_x1_num, cfx_x1_den_x1)));
        M3 cfx_x3_num_x3 = msub(msub(mscl(cfx_den, cfx_num_x3_x3), mscl(2 * cfx_num, outer(phi_s_x3, phi_s_x3))), mscl(2 * cfx_num * phi_s, phi_s_x3_x3));
        V3 cfx_x3_den_x3 = scl(2 * cfx_den, cfx_den_x3);
        M3 cfx_x3_x3 = mscl(1.0 / std::pow(cfx_x3_den, 2), msub(mscl(cfx_x3_den, cfx_x3_num_x3), outer(cfx_x3_num, cfx_x3_den_x3)));
        M3 cfx_x1_num_x3 = msub(msub(madd(mscl(cfx_den, cfx_num_x1_x3), mscl(2 * phi_s, outer(cfx_num_x1, phi_s_x3))),
                                     mscl(2 * phi_s, outer(cfx_num_x3, phi_s_x1))),
                                madd(mscl(2 * cfx_num, outer(phi_s_x1, phi_s_x3)), mscl(2 * cfx_num * phi_s, phi_s_x1_x3)));
        V3 cfx_x1_den_x3 = scl(2 * cfx_den, cfx_den_x3);
        M3 cfx_x1_x3 = mscl(1.0 / std::pow(cfx_x1_den, 2), msub(mscl(cfx_x1_den, cfx_x1_num_x3), outer(cfx_x1_num, cfx_x1_den_x3)));
        M3 cfx_x3_num_x1 = msub(msub(madd(mscl(cfx_den, cfx_num_x3_x1), mscl(2 * phi_s, outer(cfx_num_x3, phi_s_x1))),
                                     mscl(2 * phi_s, outer(cfx_num_x1, phi_s_x3))),
                                madd(mscl(2 * cfx_num, outer(phi_s_x3, phi_s_x1)), mscl(2 * cfx_num * phi_s, phi_s_x3_x1)));
        V3 cfx_x3_den_x1 = scl(2 * cfx_den, cfx_den_x1);
        M3 cfx_x3_x1 = mscl(1.0 / std::pow(cfx_x3_den, 2), msub(mscl(cfx_x3_den, cfx_x3_num_x1), outer(cfx_x3_num, cfx_x3_den_x1)));
        double cf2 = std::pow(cf, 2);
        // (1 - cf^2 * H) parses as 1 - (cf^2 H) at src/PhysicsCloth.cpp:2914,2918 (Q4)
        M3 cf_x1_x1 = madd(mscl(-2 * cf * (1 - cf2), outer(cfx_x1, cfx_x1)), madd1(1.0, mscl(-cf2, cfx_x1_x1)));
        M3 cf_x3_x3 = madd(mscl(-2 * cf * (1 - cf2), outer(cfx_x3, cfx_x3)), madd1(1.0, mscl(-cf2, cfx_x3_x3)));
        M3 cf_x1_x3 = madd(mscl(-2 * cf * (1 - cf2), outer(cfx_x3, cfx_x1)), mscl(1 - cf2, cfx_x1_x3));
        M3 cf_x3_x1 = madd(mscl(-2 * cf * (1 - cf2), outer(cfx_x1, cfx_x3)), mscl(1 - cf2, cfx_x3_x1));
        double gc = std::pow(gamma, c), gc1 = std::pow(gamma, c - 1), gc2 = std::pow(gamma, c - 2);
        auto kterm = [&](const V3& ga, const V3& gb, const M3& gab, const V3& cfa_t2, const V3& cfb_t2,
                         const V3& cfa_t3, const V3& gb_t3, bool t3_cf_first, const M3& cfab) {
            // term1 - term2 + term3 of src/PhysicsCloth.cpp:2928-2972
            M3 t1 = madd(mscl(c * (c - 1) * gc2, outer(ga, gb)), mscl(c * gc1, gab));
            M3 t2 = madd(madd(mscl(c * (c - 1) * gc2 * cf, outer(ga, gb)), mscl(c * gc1 * cf, gab)), mscl(c * gc1, outer(cfa_t2, cfb_t2)));
            M3 t3 = madd(mscl(-c * gc1, t3_cf_first ? outer(cfa_t3, gb_t3) : outer(gb_t3, cfa_t3)), mscl(1 - gc, cfab));
            return madd(msub(t1, t2), t3);
        };
        // term2's outer is (gamma_a, con_func_b); term3's is (gamma_a, con_func_b) for the
        // x1x1/x3x3/x3x1... forms exactly as written below
        M3 k11 = kterm(gamma_x1, gamma_x1, gamma_x1_x1, gamma_x1, cf_x1, gamma_x1, cf_x1, true, cf_x1_x1);
        M3 k33 = kterm(gamma_x3, gamma_x3, gamma_x3_x3, gamma_x3, cf_x3, gamma_x3, cf_x3, true, cf_x3_x3);
        M3 k13 = kterm(gamma_x1, gamma_x3, gamma_x1_x3, gamma_x1, cf_x3, cf_x1, gamma_x3, true, cf_x1_x3);
        M3 k31 = kterm(gamma_x3, gamma_x1, gamma_x3_x1, gamma_x3, cf_x1, cf_x3, gamma_x1, true, cf_x3_x1);
        double base = 0.5 * (comp_shear + 1.0) * S * PI * std::pow(R, 2);
        kx = base * ((1 + gc) + (1 - gc) * cf);
        V3 kx_x1 = scl(base, add(sub(scl(c * gc1, gamma_x1), scl(c * gc1 * cf, gamma_x1)), scl(1 - gc, vadd1(1.0, scl(-cf2, cfx_x1)))));
        V3 kx_x3 = scl(base, add(sub(scl(c * gc1, gamma_x3), scl(c * gc1 * cf, gamma_x3)), scl(1 - gc, vadd1(1.0, scl(-cf2, cfx_x3)))));
        double base2 = 0.5 * (comp_shear + 1.0) * S * std::pow(R, 2);   // no PI (Q4)
        M3 kx_x1_x1 = mscl(base2, k11), kx_x3_x3 = mscl(base2, k33), kx_x1_x3 = mscl(base2, k13), kx_x3_x1 = mscl(base2, k31);
        // forces (Shear_F, src/PhysicsCloth.cpp:2705-2716)
        V3 F_x1 = add(scl(-0.5 * L * std::pow(pb, 2), kx_x1), scl((kx * L * pb) / (l1 * sn), P1d3));
        V3 F_x3 = add(scl(-0.5 * L * std::pow(pb, 2), kx_x3), scl((kx * L * pb) / (l3 * sn), P3d1));
        V3 F_x0 = neg(add(F_x1, F_x3));
        addFL(s.n0, F_x0); addFL(s.n1, F_x1); addFL(s.n3, F_x3);
        // Jacobian (src/PhysicsCloth.cpp:2997-3051)
        M3 F11o = mscl((kx * L) / (std::pow(l1, 2) * sn),
                       msub(mscl(pb, msub(msub(madd(mneg(outer(P1d3, d1)), mscl(cs / std::pow(sn, 2), mm(outer(P1d3, d3), P1))), mscl(cs, P1)), mm(outer(d1, d3), P1))),
                            mscl(1.0 / sn, mm(outer(P1d3, d3), P1))));
        M3 F33o = mscl((kx * L) / (std::pow(l3, 2) * sn),
                       msub(mscl(pb, msub(msub(madd(mneg(outer(P3d1, d3)), mscl(cs / std::pow(sn, 2), mm(outer(P3d1, d1), P3))), mscl(cs, P3)), mm(outer(d3, d1), P3))),
                            mscl(1.0 / sn, mm(outer(P3d1, d1), P3))));
        M3 F13o = mscl((kx * L) / (l1 * l3 * sn),
                       msub(mscl(pb, madd(mscl(cs / std::pow(sn, 2), mm(outer(P1d3, d1), P3)), mm(P1, P3))),
                            mscl(1.0 / sn, mm(outer(P1d3, d1), P3))));
        M3 F31o = mscl((kx * L) / (l3 * l1 * sn),
                       msub(mscl(pb, madd(mscl(cs / std::pow(sn, 2), mm(outer(P3d1, d3), P1)), mm(P3, P1))),
                            mscl(1.0 / sn, mm(outer(P3d1, d3), P1))));
        M3 F11 = madd(msub(mscl(-0.5 * L * std::pow(pb, 2), kx_x1_x1), mscl(2 * L * pb, outer(kx_x1, phi_x1))), F11o);
        M3 F33 = madd(msub(mscl(-0.5 * L * std::pow(pb, 2), kx_x3_x3), mscl(2 * L * pb, outer(kx_x3, phi_x3))), F33o);
        M3 F13 = madd(msub(msub(mscl(-0.5 * L * std::pow(pb, 2), kx_x1_x3), mscl(L * pb, outer(kx_x1, phi_x3))), mscl(L * pb, outer(phi_x1, kx_x3))), F13o);
        M3 F31 = madd(msub(msub(mscl(-0.5 * L * std::pow(pb, 2), kx_x3_x1), mscl(L * pb, outer(kx_x3, phi_x1))), mscl(L * pb, outer(phi_x3, kx_x1))), F31o);
        M3 F10 = mneg(madd(F11, F13)), F30 = mneg(madd(F31, F33));
        M3 F01 = mneg(madd(F11, F31)), F03 = mneg(madd(F13, F33));
        M3 F00 = mneg(madd(F10, F30));
        addLL(Kx, s.n0, s.n0, F00); addLL(Kx, s.n0, s.n1, F01); addLL(Kx, s.n0, s.n3, F03);
        addLL(Kx, s.n1, s.n0, F10); addLL(Kx, s.n1, s.n1, F11); addLL(Kx, s.n1, s.n3, F13);
        addLL(Kx, s.n3, s.n0, F30); addLL(Kx, s.n3, s.n1, F31); addLL(Kx, s.n3, s.n3, F33);
        return;
    }
    // non-jamming branch (src/PhysicsCloth.cpp:2720-2726, 3056-3088)
    double pb = phi - PI / 2;
    if (!frozen_pb.empty()) pb = frozen_pb[&s - shears.data()];
    V3 F_x1 = scl((kx * L * pb) / (l1 * sn), P1d3);
    V3 F_x3 = scl((kx * L * pb) / (l3 * sn), P3d1);
    V3 F_x0 = scl(-1, add(F_x1, F_x3));
    addFL(s.n0, F_x0); addFL(s.n1, F_x1); addFL(s.n3, F_x3);
    M3 L1L1 = mscl((kx * L) / (std::pow(l1, 2) * sn),
                   msub(mscl(pb, msub(msub(madd(mscl(-1, outer(P1d3, d1)), mscl(cs / std::pow(sn, 2), mm(outer(P1d3, d3), P1))), mscl(cs, P1)), mm(outer(d1, d3), P1))),
                        mscl(1 / sn, mm(outer(P1d3, d3), P1))));
    M3 L1L3 = mscl((kx * L) / (l3 * l1 * sn),
                   mm(mscl(pb, msub(mscl(cs / std::pow(sn, 2), outer(P1d3, d1)), mscl(1 / sn, outer(P1d3, d1)))), P3));
    M3 L3L1 = mscl((kx * L) / (l1 * l3 * sn),
                   mm(mscl(pb, msub(mscl(cs / std::pow(sn, 2), outer(P3d1, d3)), mscl(1 / sn, outer(P3d1, d3)))), P1));
    M3 L3L3 = mscl((kx * L) / (std::pow(l3, 2) * sn),
                   msub(mscl(pb, msub(msub(madd(mscl(-1, outer(P3d1, d3)), mscl(cs / std::pow(sn, 2), mm(outer(P3d1, d1), P3))), mscl(cs, P3)), mm(outer(d3, d1), P3))),
                        mscl(1 / sn, mm(outer(P3d1, d1), P3))));
    M3 L1L0 = mneg(madd(L1L1, L1L3)), L3L0 = mneg(madd(L3L1, L3L3));
    M3 L0L1 = mneg(madd(L1L1, L3L1)), L0L3 = mneg(madd(L1L3, L3L3));
    M3 L0L0 = mneg(madd(L1L0, L3L0));
    addLL(Kx, s.n0, s.n0, L0L0); addLL(Kx, s.n0, s.n1, L0L1); addLL(Kx, s.n0, s.n3, L0L3);
    addLL(Kx, s.n1, s.n0, L1L0); addLL(Kx, s.n1, s.n1, L1L1); addLL(Kx, s.n1, s.n3, L1L3);
    addLL(Kx, s.n3, s.n0, L3L0); addLL(Kx, s.n3, s.n1, L3L1); addLL(Kx, s.n3, s.n3, L3L3);
}

double yo_cloth::shear_phi(const Shear& s) const {
    V3 a = sub(x[s.n1], x[s.n0]), b3 = sub(x[s.n3], x[s.n0]);
    V3 d1 = scl(1.0 / norm(a), a), d3 = scl(1.0 / norm(b3), b3);
    return std::acos(dot(d1, d3));
}

// Wind_F (src/PhysicsCloth.cpp:3091-3106) with Face::update (src/ClothUnit.h:268-273)
void yo_cloth::Wind_F(const Face& f) {
    V3 cr = cross(sub(x[f.n2], x[f.n1]), sub(x[f.n3], x[f.n1]));
    double cn = norm(cr);
    double area = 0.5 * cn;
    V3 nf = cn == 0.0 ? cr : scl(1.0 / cn, cr);
    V3 fv = scl(1.0 / 3.0, add(add(v[f.n1], v[f.n2]), v[f.n3]));
    V3 rel = sub(v3(p.wind_vel[0], p.wind_vel[1], p.wind_vel[2]), fv);
    double vn = dot(nf, rel);
    V3 vt = sub(rel, scl(vn, nf));
    V3 wf = add(scl(p.wind_density * area * std::fabs(vn) * vn, nf), scl(p.wind_drag * area, vt));
    V3 w3 = scl(1.0 / 3.0, wf);
    addFL(f.n1, w3); addFL(f.n2, w3); addFL(f.n3, w3);
}

// Constraint_F / Constraint_D (src/PhysicsCloth.cpp:232-248)
void yo_cloth::Constraint(int k) {
    if (!is_fixed[k]) return;
    addFL(k, scl(-p.handle_k, sub(x[k], xfix[k])));
    addLL(Kx, k, k, mscl(-p.handle_k, eye3()));
}

// ---------------------------------------------------------------------------
// FillMV + system formation
// ---------------------------------------------------------------------------

// The element loops of FillMV (src/PhysicsCloth.cpp:46-131) over the rows this thread
// owns; `mid` runs between the friction and the shear loop (a barrier in a parallel
// assembly: the shear reads the compression of crossings other threads own).
template <class Mid>
void yo_cloth::element_loops(Mid mid) {
    const bool all = tl_vlo == 0 && tl_vhi == INT_MAX;
    auto touches = [&](int a, int b, int c) { return all || own(a) || own(b) || own(c); };
    // src/PhysicsCloth.cpp:56-62
    if (on(YO_TERM_HANDLE) && only_kind == 0) for (int k = 0; k < N; ++k) if (own(k)) Constraint(k);
    // src/PhysicsCloth.cpp:64-80 (edge loop; t = owner thread of edge i)
    int ne = (int)edges.size();
    for (int i = 0; i < ne; ++i) {
        if (!pick(YO_ELEM_EDGE, i)) continue;
        const Edge& e = edges[i];
        if (!touches(e.n0, e.n1, e.n1)) continue;
        int t = owner(ne, i);
        EdgeGeo g = edge_geo(e);
        if (on(YO_TERM_MASS)) Mass(e, g);
        if (on(YO_TERM_GRAVITY)) Gravity_F(e, g, t);
        if (on(YO_TERM_STRETCH)) Stretch_F(e, g, t);
        if (on(YO_TERM_INERTIA)) Inertia_F(e, g, t);
        if (on(YO_TERM_MDOT)) MDotVDot_F(e, g, t);
        if (on(YO_TERM_PARACOLLI)) ParaColli_F(e, g, t);
        if (on(YO_TERM_GRAVITY)) Gravity_D(e, g);
        if (on(YO_TERM_STRETCH)) Stretch_D(e, g);
        if (on(YO_TERM_INERTIA)) { Inertia_D_Pos(e, g); Inertia_D_Vel(e, g); }
        if (on(YO_TERM_MDOT)) { MDotVDot_D_Pos(e, g); MDotVDot_D_Vel(e, g); }
        if (on(YO_TERM_PARACOLLI)) ParaColli_D(e, g);
    }
    // src/PhysicsCloth.cpp:82-89
    int nb = (int)bends.size();
    for (int i = 0; i < nb; ++i) {
        if (!pick(YO_ELEM_BEND, i)) continue;
        if (!touches(bends[i].n2, bends[i].n0, bends[i].n1)) continue;
        if (on(YO_TERM_BEND)) Bending(bends[i], owner(nb, i));
        if (on(YO_TERM_CRIMP)) Crimp_Bending(bends[i]);
    }
    // src/PhysicsCloth.cpp:91-95
    if (on(YO_TERM_WIND))
        for (size_t i = 0; i < faces.size(); ++i)
            if (pick(YO_ELEM_FACE, (int)i) && touches(faces[i].n1, faces[i].n2, faces[i].n3)) Wind_F(faces[i]);
    // src/PhysicsCloth.cpp:97-112: no-slide snapshots (the EE diagonals are read from Kx/Kv directly)
    for (int j = 1; j < Lg - 1; ++j)
        if (j >= tl_vlo && j < tl_vhi)
            for (int i = 1; i < W - 1; ++i) for (int c = 0; c < 2; ++c) FE_noslide[2 * eidx(i, j) + c] = FE[2 * eidx(i, j) + c];
    // src/PhysicsCloth.cpp:114-123 (with the friction term off, the compression is still kept)
    for (int j = 1; j < Lg - 1; ++j) {
        if (j < tl_vlo || j >= tl_vhi) continue;
        for (int i = 1; i < W - 1; ++i) {
            if (on(YO_TERM_FRICTION) && pick(YO_ELEM_CROSSING, eidx(i, j))) ConSlideFriction(kidx(i, j));
            else comp[eidx(i, j)] = relu(0.5 * dot(nrm[kidx(i, j)], StrhBendC_L[kidx(i, j)]));
        }
    }
    mid();
    // src/PhysicsCloth.cpp:125-131
    if (on(YO_TERM_SHEAR))
        for (size_t i = 0; i < shears.size(); ++i)
            if (pick(YO_ELEM_SHEAR, (int)i) && touches(shears[i].n0, shears[i].n1, shears[i].n3)) Shear_FD(shears[i]);
}

int yo_cloth::assemble(double h, const double* ext) {
    std::fill(FL.begin(), FL.end(), 0.0);
    std::fill(FE.begin(), FE.end(), 0.0);
    std::fill(FEpart.begin(), FEpart.end(), 0.0);
    for (auto& s : StrhBendC_L) s = V3{0, 0, 0};
    for (auto& s : StrhBendC_LEself) s = {V3{0, 0, 0}, V3{0, 0, 0}};
    M.zero(); Kx.zero(); Kv.zero();
    if (threads <= 1) {
        element_loops([] {});
    } else {
#pragma omp parallel num_threads(threads)
        {
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();
            tl_vlo = (int)((long long)Lg * t / nt);
            tl_vhi = (int)((long long)Lg * (t + 1) / nt);
            element_loops([] {
#pragma omp barrier
            });
            tl_vlo = 0;
            tl_vhi = INT_MAX;
        }
    }

    // CSR pattern: union of the M, Kx, Kv rows, sorted.  Each accumulator row is sorted by
    // column first (its values are already summed per entry), so the union and the values
    // come from one merge per row.
    rowptr.assign(n + 1, 0);
    std::vector<std::array<std::vector<std::pair<int, double>>, 3>> srt(n);
    std::vector<std::vector<int>> pat(n);
#pragma omp parallel for num_threads(threads) if (threads > 1) schedule(static)
    for (int r = 0; r < n; ++r) {
        const SpAcc* src[3] = {&M, &Kx, &Kv};
        for (int m = 0; m < 3; ++m) {
            srt[r][m] = src[m]->rows[r];
            std::sort(srt[r][m].begin(), srt[r][m].end(),
                      [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
            for (auto& e : srt[r][m]) pat[r].push_back(e.first);
        }
        std::sort(pat[r].begin(), pat[r].end());
        pat[r].erase(std::unique(pat[r].begin(), pat[r].end()), pat[r].end());
    }
    for (int r = 0; r < n; ++r) rowptr[r + 1] = rowptr[r] + (int)pat[r].size();
    int nnz = rowptr[n];
    col.assign(nnz, 0);
    Aval.assign(nnz, 0); Mval.assign(nnz, 0); Kxval.assign(nnz, 0); Kvval.assign(nnz, 0);
    const double h2 = std::pow(h, 2);
#pragma omp parallel for num_threads(threads) if (threads > 1) schedule(static)
    for (int r = 0; r < n; ++r) {
        size_t it[3] = {0, 0, 0};
        std::vector<double>* dst[3] = {&Mval, &Kxval, &Kvval};
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            const int c = pat[r][j - rowptr[r]];
            col[j] = c;
            for (int m = 0; m < 3; ++m) {
                const auto& row = srt[r][m];
                (*dst[m])[j] = (it[m] < row.size() && row[it[m]].first == c) ? row[it[m]++].second : 0.0;
            }
            // A = M - h^2 K_x - h K_v (src/TrainHelper.cpp:19)
            Aval[j] = Mval[j] - Kxval[j] * h2 - Kvval[j] * h;
        }
    }
    // F and v in the global layout (get_SpaF / get_SpaVel, src/MatrixMp.cpp:120-122)
    F.assign(n, 0); vel.assign(n, 0);
    for (int k = 0; k < N; ++k) for (int a = 0; a < 3; ++a) { F[3 * k + a] = FL[3 * k + a]; vel[3 * k + a] = v[k][a]; }
    for (int j = 1; j < Lg - 1; ++j) for (int i = 1; i < W - 1; ++i) {
        int k = kidx(i, j), e = eidx(i, j);
        for (int c = 0; c < 2; ++c) { F[3 * N + 2 * e + c] = FE[2 * e + c]; vel[3 * N + 2 * e + c] = uvd[k][c]; }
    }
    // b = h (F [+ f_ext] - K_v^T v) + M v (src/TrainHelper.cpp:20, 63)
    std::vector<double> KvTv(n, 0.0), Mv(n, 0.0);
    for (int r = 0; r < n; ++r) for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
        KvTv[col[j]] += Kvval[j] * vel[r];
        Mv[r] += Mval[j] * vel[col[j]];
    }
    b.assign(n, 0);
    for (int i = 0; i < n; ++i) b[i] = h * (F[i] + (ext ? ext[i] : 0.0) - KvTv[i]) + Mv[i];
    h_last = h;
    for (int j = 0; j < nnz; ++j) if (!std::isfinite(Aval[j])) return 1;
    for (int i = 0; i < n; ++i) if (!std::isfinite(b[i])) return 1;
    return 0;
}

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------

extern "C" {

yo_cloth* yo_create(const yo_params* p) {
    if (!p || p->width < 3 || p->length < 3 || p->n_yarns < 1 || p->n_yarns > 2) return nullptr;
    yo_cloth* c = new yo_cloth();
    c->p = *p;
    c->build();
    return c;
}

void yo_destroy(yo_cloth* c) { delete c; }

int yo_ndof(const yo_cloth* c) { return c->n; }

int yo_owner(int n_iter, int i) { return owner(n_iter, i); }

// get_ClothState (src/Cloth.cpp:350-371).  The reference's state index is u + v*length
// (Q9), which only addresses a square cloth's nodes one to one (it collides or runs past
// the arrays when width != length); this index, u + v*width, equals it for square cloths
// and is the trajectory loader's order (src/Loader.cpp:160) for the others.
void yo_get_state(const yo_cloth* c, double* lpos, double* lposfix, double* lvel,
                  double* epos, double* eposbar, double* evel) {
    for (int v2 = 0; v2 < c->Lg; ++v2) for (int u = 0; u < c->W; ++u) {
        int k = c->kidx(u, v2), s = u + v2 * c->W;
        for (int a = 0; a < 3; ++a) {
            if (lpos) lpos[3 * s + a] = c->x[k][a];
            if (lposfix) lposfix[3 * s + a] = c->xfix[k][a];
            if (lvel) lvel[3 * s + a] = c->v[k][a];
        }
        for (int a = 0; a < 2; ++a) {
            if (epos) epos[2 * s + a] = c->uv[k][a];
            if (eposbar) eposbar[2 * s + a] = c->uvbar[k][a];
            if (evel) evel[2 * s + a] = c->uvd[k][a];
        }
    }
}

// set_ClothState (src/Cloth.cpp:421-438) followed by UpdateClothUnits
void yo_set_state(yo_cloth* c, const double* lpos, const double* lposfix, const double* lvel,
                  const double* epos, const double* eposbar, const double* evel) {
    for (int v2 = 0; v2 < c->Lg; ++v2) for (int u = 0; u < c->W; ++u) {
        int k = c->kidx(u, v2), s = u + v2 * c->W;
        for (int a = 0; a < 3; ++a) {
            c->x[k][a] = lpos[3 * s + a];
            c->xfix[k][a] = lposfix[3 * s + a];
            c->v[k][a] = lvel[3 * s + a];
        }
        for (int a = 0; a < 2; ++a) {
            c->uv[k][a] = epos[2 * s + a];
            c->uvbar[k][a] = eposbar[2 * s + a];
            c->uvd[k][a] = evel[2 * s + a];
        }
    }
    c->normals_and_crimp();
}

int yo_assemble(yo_cloth* c, double h, const double* ext_force) { return c->assemble(h, ext_force); }

void yo_set_terms(yo_cloth* c, unsigned terms) { c->terms = terms; }

void yo_set_element(yo_cloth* c, int kind, int index) { c->only_kind = kind; c->only_index = index; }

void yo_set_threads(yo_cloth* c, int threads) { c->threads = threads < 1 ? 1 : threads; }

void yo_friction_eval(const double* in, double* out) {
    FricIn i{in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], in[10]};
    FricOut o = friction_eval(i);
    out[0] = o.is_slide; out[1] = o.fric; out[2] = o.fpos; out[3] = o.fvel;
}

void yo_get_compression_parts(const yo_cloth* c, double* sbc_l, double* sbc_le) {
    for (int k = 0; k < c->N; ++k)
        for (int a = 0; a < 3; ++a) {
            if (sbc_l) sbc_l[3 * k + a] = c->StrhBendC_L[k][a];
            if (sbc_le) for (int q = 0; q < 2; ++q) sbc_le[6 * k + 3 * q + a] = c->StrhBendC_LEself[k][q][a];
        }
}

void yo_freeze_shear_angle(yo_cloth* c, int on) {
    c->frozen_pb.clear();
    if (!on) return;
    for (const Shear& s : c->shears) c->frozen_pb.push_back(c->shear_phi(s) - PI / 2);
}

void yo_set_frozen_shear_angle(yo_cloth* c, const double* pb, int n) {
    c->frozen_pb.assign(pb, pb + n);
}

int yo_get_frozen_shear_angle(const yo_cloth* c, double* pb) {
    for (size_t i = 0; i < c->frozen_pb.size(); ++i) pb[i] = c->frozen_pb[i];
    return (int)c->frozen_pb.size();
}

int yo_nnz(const yo_cloth* c) { return (int)c->col.size(); }

void yo_get_csr(const yo_cloth* c, int which, int* rowptr, int* col, double* val) {
    const std::vector<double>& src = which == 0 ? c->Aval : which == 1 ? c->Mval : which == 2 ? c->Kxval : c->Kvval;
    if (rowptr) std::memcpy(rowptr, c->rowptr.data(), sizeof(int) * c->rowptr.size());
    if (col) std::memcpy(col, c->col.data(), sizeof(int) * c->col.size());
    if (val) std::memcpy(val, src.data(), sizeof(double) * src.size());
}

void yo_get_dense(const yo_cloth* c, int which, double* dense) {
    const std::vector<double>& src = which == 0 ? c->Aval : which == 1 ? c->Mval : which == 2 ? c->Kxval : c->Kvval;
    int n = c->n;
    std::memset(dense, 0, sizeof(double) * (size_t)n * n);
    for (int r = 0; r < n; ++r) for (int j = c->rowptr[r]; j < c->rowptr[r + 1]; ++j) dense[(size_t)r * n + c->col[j]] = src[j];
}

void yo_get_vec(const yo_cloth* c, int which, double* out) {
    const std::vector<double>& src = which == 0 ? c->F : which == 1 ? c->b : c->vel;
    std::memcpy(out, src.data(), sizeof(double) * src.size());
}

void yo_get_crossing(const yo_cloth* c, double* is_slide, double* comp, double* fe_partial) {
    if (is_slide) std::memcpy(is_slide, c->is_slide.data(), sizeof(double) * c->is_slide.size());
    if (comp) std::memcpy(comp, c->comp.data(), sizeof(double) * c->comp.size());
    if (fe_partial) std::memcpy(fe_partial, c->FEpart.data(), sizeof(double) * c->FEpart.size());
}

void yo_get_normals(const yo_cloth* c, double* n) {
    for (int k = 0; k < c->N; ++k) for (int a = 0; a < 3; ++a) n[3 * k + a] = c->nrm[k][a];
}

// Jacobi-preconditioned CG (src/Solvers.cpp:22-24; the published Eigen 3.3
// ConjugateGradient + DiagonalPreconditioner algorithm, full matrix used as stored)
int yo_solve(yo_cloth* c, double tol, int maxit, const double* x0, double* xout, int* iters, double* relres) {
    int n = c->n;
    const std::vector<int>& rp = c->rowptr;
    const std::vector<int>& cl = c->col;
    const std::vector<double>& A = c->Aval;
    const std::vector<double>& b = c->b;
    const int T = c->threads;
    auto spmv = [&](const std::vector<double>& in, std::vector<double>& out) {
#pragma omp parallel for num_threads(T) if (T > 1) schedule(static)
        for (int r = 0; r < n; ++r) { double s = 0; for (int j = rp[r]; j < rp[r + 1]; ++j) s += A[j] * in[cl[j]]; out[r] = s; }
    };
    // inner products: in index order with one thread; with several, as the ordered sum of
    // 256 fixed chunks (the same value for any thread count)
    constexpr int NCH = 256;
    std::vector<double> part(NCH);
    auto dotv = [&](const std::vector<double>& x, const std::vector<double>& y) {
        if (T <= 1) { double s = 0; for (int i = 0; i < n; ++i) s += x[i] * y[i]; return s; }
#pragma omp parallel for num_threads(T) schedule(static)
        for (int q = 0; q < NCH; ++q) {
            const int i0 = (int)((long long)n * q / NCH), i1 = (int)((long long)n * (q + 1) / NCH);
            double s = 0;
            for (int i = i0; i < i1; ++i) s += x[i] * y[i];
            part[q] = s;
        }
        double s = 0;
        for (int q = 0; q < NCH; ++q) s += part[q];
        return s;
    };
    std::vector<double> invdiag(n, 1.0);
    for (int r = 0; r < n; ++r) for (int j = rp[r]; j < rp[r + 1]; ++j) if (cl[j] == r && A[j] != 0.0) invdiag[r] = 1.0 / A[j];
    std::vector<double> x(n, 0.0), r(n), p(n), z(n), tmp(n);
    if (x0) for (int i = 0; i < n; ++i) x[i] = x0[i];
    spmv(x, tmp);
    for (int i = 0; i < n; ++i) r[i] = b[i] - tmp[i];
    double rhsNorm2 = dotv(b, b);
    if (rhsNorm2 == 0) {
        for (int i = 0; i < n; ++i) xout[i] = 0;
        if (iters) *iters = 0;
        if (relres) *relres = 0;
        return 0;
    }
    double threshold = std::max(tol * tol * rhsNorm2, std::numeric_limits<double>::min());
    double res2 = dotv(r, r);
    int it = 0;
    if (res2 >= threshold) {
        for (int i = 0; i < n; ++i) p[i] = invdiag[i] * r[i];
        double absNew = dotv(r, p);
        while (it < maxit) {
            spmv(p, tmp);
            double pAp = dotv(p, tmp);
            double alpha = absNew / pAp;
#pragma omp parallel for num_threads(T) if (T > 1) schedule(static)
            for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
            res2 = dotv(r, r);
            if (res2 < threshold) break;
#pragma omp parallel for num_threads(T) if (T > 1) schedule(static)
            for (int i = 0; i < n; ++i) z[i] = invdiag[i] * r[i];
            double absOld = absNew;
            absNew = dotv(r, z);
            double beta = absNew / absOld;
#pragma omp parallel for num_threads(T) if (T > 1) schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            it++;
        }
    }
    for (int i = 0; i < n; ++i) xout[i] = x[i];
    if (iters) *iters = it;
    if (relres) *relres = std::sqrt(res2 / rhsNorm2);
    for (int i = 0; i < n; ++i) if (!std::isfinite(x[i])) return 2;
    return 0;
}

// Cloth::update (src/Cloth.cpp:111-148)
void yo_update(yo_cloth* c, const double* nv, double h) {
    int N = c->N;
    for (int i = 0; i < c->W; ++i) for (int j = 0; j < c->Lg; ++j) {
        int k = c->kidx(i, j);
        for (int a = 0; a < 3; ++a) {
            c->v[k][a] = nv[3 * k + a];
            c->x[k][a] = c->x[k][a] + c->v[k][a] * h;
        }
        if (!c->is_fixed[k]) c->xfix[k] = c->x[k];
        if (!c->is_edge[k]) {
            int e = c->eidx(i, j);
            for (int a = 0; a < 2; ++a) {
                c->uvd[k][a] = nv[3 * N + 2 * e + a];
                c->uv[k][a] = c->uv[k][a] + c->uvd[k][a] * h;
                c->uvbar[k][a] = c->uvbar[k][a] + c->is_slide[2 * e + a] * c->uvd[k][a] * h;
            }
        }
    }
    c->normals_and_crimp();
}

int yo_step(yo_cloth* c, double h, const double* ext_force, double tol, int maxit, int* iters, double* relres) {
    if (c->assemble(h, ext_force) != 0) return 1;
    std::vector<double> nv(c->n);
    int st = yo_solve(c, tol, maxit, nullptr, nv.data(), iters, relres);
    if (st != 0) return 2;
    yo_update(c, nv.data(), h);
    return 0;
}

// SimStep with the reference's dense storage (a CPU cost model of its layout, W <= ~25):
// per step the 8 per-thread block arrays of ResetMV (src/PhysicsCloth.cpp:158-197) are
// allocated zeroed -- LL N^2 x 9, LE/EL N x 2Nc x 3, EE (2Nc)^2 for M, F_*_Pos and F_*_Vel
// -- the assembled blocks are placed in thread 0's arrays at find_index's offsets and the
// FillMV reduction adds the other threads' arrays in (:134-156); To_Sparse_Mat builds the
// five dense n x n matrices SimStep asks for (get_SpaM twice, get_SpaFPos, get_SpaFVel
// twice, src/TrainHelper.cpp:19-20, src/MatrixMp.cpp:40-111), A and b are formed densely,
// A and b are scanned for NaN/Inf (:22-34), and EigenSolverForward's n^2 scan turns A into
// the sparse solver matrix (src/Solvers.cpp:6-12).  The element arithmetic is the oracle's,
// not LibTorch's per-element tensors, so this is a lower bound of the reference's cost.
// The matrix the scan produces holds the sparse path's values in its order, so the step
// equals yo_step bit for bit.
int yo_step_dense(yo_cloth* c, double h, const double* ext_force, double tol, int maxit, int* iters, double* relres) {
    if (c->assemble(h, ext_force) != 0) return 1;
    const int N = c->N, Nc = c->Nc, n = c->n, T = 8, nt = std::max(1, c->threads);
    const size_t nLL = (size_t)N * N * 9, nEL = (size_t)N * 2 * Nc * 3, nEE = (size_t)2 * Nc * 2 * Nc;
    // per kind (M, Pos, Vel) and thread: LL, EL, LE, EE
    std::vector<std::vector<double>> arr((size_t)3 * T * 4);
    auto A_ = [&](int kind, int t, int part) -> std::vector<double>& { return arr[((size_t)kind * T + t) * 4 + part]; };
    for (int kind = 0; kind < 3; ++kind)
        for (int t = 0; t < T; ++t) {
            const size_t sz[4] = {nLL, nEL, nEL, nEE};
            for (int part = 0; part < 4; ++part) {
                std::vector<double>& v = A_(kind, t, part);
                v.resize(sz[part]);
                double* d = v.data();
#pragma omp parallel for num_threads(nt) schedule(static)
                for (long long i = 0; i < (long long)sz[part]; ++i) d[i] = 0.0;
            }
        }
    // the assembled blocks at find_index's offsets (src/PhysicsCloth.cpp:199-216):
    // LL[F N + D][u][v] = A[3F+v, 3D+u]; EL/LE[F 2Nc + 2 e + c]; EE[(2 eF + c) 2Nc + 2 eD + c]
    const std::vector<double>* vals[3] = {&c->Mval, &c->Kxval, &c->Kvval};
    for (int kind = 0; kind < 3; ++kind)
        for (int r = 0; r < n; ++r)
            for (int j = c->rowptr[r]; j < c->rowptr[r + 1]; ++j) {
                const int col = c->col[j];
                const double val = (*vals[kind])[j];
                if (r < 3 * N && col < 3 * N)
                    A_(kind, 0, 0)[((size_t)(r / 3) * N + col / 3) * 9 + (col % 3) * 3 + r % 3] = val;
                else if (r >= 3 * N && col < 3 * N)
                    A_(kind, 0, 1)[((size_t)(col / 3) * 2 * Nc + (r - 3 * N)) * 3 + col % 3] = val;
                else if (r < 3 * N)
                    A_(kind, 0, 2)[((size_t)(r / 3) * 2 * Nc + (col - 3 * N)) * 3 + r % 3] = val;
                else
                    A_(kind, 0, 3)[(size_t)(r - 3 * N) * 2 * Nc + (col - 3 * N)] = val;
            }
    // the FillMV reduction of the per-thread arrays into thread 0's
    for (int kind = 0; kind < 3; ++kind)
        for (int t = 1; t < T; ++t)
            for (int part = 0; part < 4; ++part) {
                double* d0 = A_(kind, 0, part).data();
                const double* dt = A_(kind, t, part).data();
                const long long sz = (long long)A_(kind, 0, part).size();
#pragma omp parallel for num_threads(nt) schedule(static)
                for (long long i = 0; i < sz; ++i) d0[i] = d0[i] + dt[i];
            }
    // To_Sparse_Mat (src/MatrixMp.cpp:40-111) for M, F_*_Pos, F_*_Vel, M, F_*_Vel
    auto to_dense = [&](int kind, std::vector<double>& D) {
        D.assign((size_t)n * n, 0.0);
        const double* LL = A_(kind, 0, 0).data();
        const double* EL = A_(kind, 0, 1).data();
        const double* LE = A_(kind, 0, 2).data();
        const double* EE = A_(kind, 0, 3).data();
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int F = 0; F < N; ++F) {
            for (int Dn = 0; Dn < N; ++Dn)
                for (int u = 0; u < 3; ++u)
                    for (int v = 0; v < 3; ++v) D[(size_t)(3 * F + v) * n + 3 * Dn + u] = LL[((size_t)F * N + Dn) * 9 + u * 3 + v];
            for (int e = 0; e < 2 * Nc; ++e)
                for (int v = 0; v < 3; ++v) {
                    D[(size_t)(3 * F + v) * n + 3 * N + e] = LE[((size_t)F * 2 * Nc + e) * 3 + v];
                    D[(size_t)(3 * N + e) * n + 3 * F + v] = EL[((size_t)F * 2 * Nc + e) * 3 + v];
                }
        }
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int e = 0; e < 2 * Nc; ++e)
            for (int f = 0; f < 2 * Nc; ++f) D[(size_t)(3 * N + e) * n + 3 * N + f] = EE[(size_t)e * 2 * Nc + f];
    };
    std::vector<double> M1, Kx, Kv1, M2, Kv2;
    to_dense(0, M1); to_dense(1, Kx); to_dense(2, Kv1); to_dense(0, M2); to_dense(2, Kv2);
    // A = M - h^2 K_x - h K_v; b = h (F [+ ext] - K_v^T v) + M v, densely (src/TrainHelper.cpp:19-20)
    std::vector<double> Ad((size_t)n * n);
    const double h2 = std::pow(h, 2);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long i = 0; i < (long long)n * n; ++i) Ad[i] = M1[i] - Kx[i] * h2 - Kv1[i] * h;
    std::vector<double> bd(n);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) {
        double kv = 0.0, mv = 0.0;
        for (int r = 0; r < n; ++r) kv += Kv2[(size_t)r * n + i] * c->vel[r];
        for (int j = 0; j < n; ++j) mv += M2[(size_t)i * n + j] * c->vel[j];
        bd[i] = h * (c->F[i] + (ext_force ? ext_force[i] : 0.0) - kv) + mv;
    }
    // DetectNaN / DetectInf of A and b (src/TrainHelper.cpp:22-34)
    long long bad = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : bad) schedule(static)
    for (long long i = 0; i < (long long)n * n; ++i) bad += !std::isfinite(Ad[i]);
    for (int i = 0; i < n; ++i) bad += !std::isfinite(bd[i]);
    if (bad) return 1;
    // EigenSolverForward's scan of the dense A into the solver's matrix (src/Solvers.cpp:6-12);
    // the structural pattern keeps the sparse path's columns, so the solve is the same
    std::vector<int> rp(n + 1, 0), cl;
    std::vector<double> av;
    cl.reserve(c->col.size());
    av.reserve(c->col.size());
    for (int r = 0; r < n; ++r) {
        for (int j = c->rowptr[r]; j < c->rowptr[r + 1]; ++j) {
            cl.push_back(c->col[j]);
            av.push_back(Ad[(size_t)r * n + c->col[j]]);
        }
        rp[r + 1] = (int)cl.size();
    }
    double scanned = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : scanned) schedule(static)
    for (long long i = 0; i < (long long)n * n; ++i) scanned += Ad[i] != 0.0 ? 1.0 : 0.0;
    if ((long long)scanned > (long long)av.size()) return 1;   // an entry outside the stencil
    c->Aval = av;
    c->b = bd;
    std::vector<double> nv(n);
    int st = yo_solve(c, tol, maxit, nullptr, nv.data(), iters, relres);
    if (st != 0) return 2;
    yo_update(c, nv.data(), h);
    return 0;
}

}  // extern "C"

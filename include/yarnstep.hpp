/*
 * yarnstep.hpp -- C++17 host mirror of the reference's step API over the C ABI of
 * yarnstep.h (header only; links libyarnstep.so, nothing else).
 *
 * The reference's classes hold FP64 CPU tensors (src/Cloth.h:31-82,
 * src/PhysicsCloth.h:6-78, src/TrainHelper.h:10-12, src/Solvers.h:35-52).  Here the same
 * names hold std::vector<double> in the same layouts, and the reference's dense n x n
 * system matrices are CSR matrices with the yarn stencil's pattern, so that a driver's
 * step reads as the reference's:
 *
 *   phy.FillMV(h);
 *   Csr A = phy.get_SpaM() - phy.get_SpaFPos() * (h * h) - phy.get_SpaFVel() * h;
 *   Vec b = h * (phy.get_SpaF() - phy.get_SpaFVel().tmul(phy.get_SpaVel())) + phy.get_SpaM() * phy.get_SpaVel();
 *   Vec new_vel = EigenSolver::apply(cloth, A, b);
 *   cloth.update(new_vel, phy.F_E_is_Slides[0], h);
 *
 * (src/Simulate.h:86-96), or in one call SimStep(cloth, phy, h) (src/TrainHelper.cpp:15-97).
 * Errors of the C ABI other than the reference's NaN/Inf result are thrown as
 * yarnstep::Error.
 */
#ifndef YARNSTEP_HPP
#define YARNSTEP_HPP

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "yarnstep.h"

namespace yarnstep {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error("yarnstep error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void check(int rc) {
    if (rc != YC_OK) throw Error(rc, yc_last_error());
}

typedef std::vector<double> Vec;

inline Vec operator+(const Vec& a, const Vec& b) {
    if (a.size() != b.size()) throw Error(YC_ERR_ARG, "vector sizes differ");
    Vec r(a.size());
    for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] + b[i];
    return r;
}
inline Vec operator-(const Vec& a, const Vec& b) {
    if (a.size() != b.size()) throw Error(YC_ERR_ARG, "vector sizes differ");
    Vec r(a.size());
    for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] - b[i];
    return r;
}
inline Vec operator*(double s, const Vec& a) {
    Vec r(a.size());
    for (size_t i = 0; i < a.size(); ++i) r[i] = s * a[i];
    return r;
}

// Square CSR matrix in the reference's global layout.  Matrices exported from one
// context share one pattern, so sums and differences combine values entry by entry.
struct Csr {
    int n = 0;
    std::vector<int> rowptr, col;
    std::vector<double> val;

    Vec operator*(const Vec& x) const {   // A x
        if ((int)x.size() != n) throw Error(YC_ERR_ARG, "vector size differs from the matrix");
        Vec y(n, 0.0);
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) s += val[j] * x[col[j]];
            y[r] = s;
        }
        return y;
    }
    Vec tmul(const Vec& x) const {         // A^T x (the transpose(0,1) mm of src/Simulate.h:90)
        if ((int)x.size() != n) throw Error(YC_ERR_ARG, "vector size differs from the matrix");
        Vec y(n, 0.0);
        for (int r = 0; r < n; ++r)
            for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) y[col[j]] += val[j] * x[r];
        return y;
    }
    Csr operator*(double s) const {
        Csr m = *this;
        for (double& v : m.val) v *= s;
        return m;
    }
    Csr combine(const Csr& o, double so) const {
        if (n != o.n || rowptr != o.rowptr || col != o.col) throw Error(YC_ERR_ARG, "matrix patterns differ");
        Csr m = *this;
        for (size_t j = 0; j < val.size(); ++j) m.val[j] += so * o.val[j];
        return m;
    }
    Csr operator+(const Csr& o) const { return combine(o, 1.0); }
    Csr operator-(const Csr& o) const { return combine(o, -1.0); }
};

// ClothState (src/Cloth.h:22-29): N x 3 (LPos, LPosFix, LVel) and N x 2 (EPos, EPosBar,
// EVel) doubles, node index u + v * width.
struct ClothState {
    Vec LPos, LPosFix, LVel, EPos, EPosBar, EVel;
};

// Cloth (src/Cloth.h:31-82) together with the device-side PhysicsCloth state.
class Cloth {
public:
    explicit Cloth(const yc_desc& d) {
        check(yc_create(&d, &ctx_));
        check(yc_sizes(ctx_, &N_, &Nc_, &n_, nullptr));
    }
    ~Cloth() { yc_destroy(ctx_); }
    Cloth(const Cloth&) = delete;
    Cloth& operator=(const Cloth&) = delete;

    int num_nodes() const { return N_; }
    int num_cros_nodes() const { return Nc_; }
    int ndof() const { return n_; }
    yc_ctx* context() const { return ctx_; }

    void set_ClothState(const ClothState& s) {   // src/Cloth.cpp:421-438
        const size_t N = N_;
        if (s.LPos.size() != 3 * N || s.LPosFix.size() != 3 * N || s.LVel.size() != 3 * N || s.EPos.size() != 2 * N ||
            s.EPosBar.size() != 2 * N || s.EVel.size() != 2 * N)
            throw Error(YC_ERR_ARG, "ClothState arrays must hold N x 3 and N x 2 doubles");
        check(yc_set_state(ctx_, s.LPos.data(), s.LPosFix.data(), s.LVel.data(), s.EPos.data(), s.EPosBar.data(),
                           s.EVel.data()));
    }
    ClothState get_ClothState() const {          // src/Cloth.cpp:350-371
        ClothState s{Vec(3 * (size_t)N_), Vec(3 * (size_t)N_), Vec(3 * (size_t)N_),
                     Vec(2 * (size_t)N_), Vec(2 * (size_t)N_), Vec(2 * (size_t)N_)};
        check(yc_get_state(ctx_, s.LPos.data(), s.LPosFix.data(), s.LVel.data(), s.EPos.data(), s.EPosBar.data(),
                           s.EVel.data()));
        return s;
    }
    // Cloth::update(new_vel, F_E_is_Slide, h) (src/Cloth.cpp:111-148)
    void update(const Vec& new_vel, const Vec& is_slide, double h) {
        if ((int)new_vel.size() != n_ || (int)is_slide.size() != 2 * Nc_) throw Error(YC_ERR_ARG, "update: sizes");
        check(yc_update_vel(ctx_, new_vel.data(), is_slide.data(), h));
    }

private:
    yc_ctx* ctx_ = nullptr;
    int N_ = 0, Nc_ = 0, n_ = 0;
};

// PhysicsCloth (src/PhysicsCloth.h:6-78): FillMV and the get_Spa* exports.
class PhysicsCloth {
public:
    explicit PhysicsCloth(Cloth& c) : cloth_(&c) {}
    // FillMV (src/PhysicsCloth.cpp:46-156); h is the step A and b are formed with
    void FillMV(double h) {
        check(yc_assemble(cloth_->context(), h));
        F_E_is_Slides.assign(1, Vec(2 * (size_t)cloth_->num_cros_nodes()));
        check(yc_export_crossing(cloth_->context(), F_E_is_Slides[0].data(), nullptr, nullptr));
    }
    void ResetMV() {}   // the next FillMV overwrites the system

    Csr get_SpaM() const { return matrix(YC_MAT_M); }
    Csr get_SpaFPos() const { return matrix(YC_MAT_KX); }
    Csr get_SpaFVel() const { return matrix(YC_MAT_KV); }
    Vec get_SpaF() const { return vec(0); }
    Vec get_SpaPos() const { return vec(4); }
    Vec get_SpaVel() const { return vec(2); }

    std::vector<Vec> F_E_is_Slides;   // [0]: is_slide of every crossing (Nc x 2) after FillMV

private:
    Csr matrix(int which) const {
        Csr m;
        m.n = cloth_->ndof();
        int nnz = 0;
        check(yc_export_matrix(cloth_->context(), which, nullptr, nullptr, nullptr, &nnz));
        m.rowptr.resize(m.n + 1);
        m.col.resize(nnz);
        m.val.resize(nnz);
        check(yc_export_matrix(cloth_->context(), which, m.rowptr.data(), m.col.data(), m.val.data(), &nnz));
        return m;
    }
    Vec vec(int which) const {
        Vec v(cloth_->ndof());
        check(yc_export_vec(cloth_->context(), which, v.data()));
        return v;
    }
    Cloth* cloth_;
};

// EigenSolver::apply(A, b) (src/Solvers.h:35-52, forward src/Solvers.cpp:3-31): the
// reference's Jacobi-PCG on the device, on the cloth's context.
struct EigenSolver {
    static Vec apply(Cloth& cloth, const Csr& A, const Vec& b, double tol = 1e-8, int maxit = -1, int* iters = nullptr) {
        if (A.n != cloth.ndof() || (int)b.size() != A.n) throw Error(YC_ERR_ARG, "EigenSolver::apply: sizes");
        check(yc_load_system(cloth.context(), A.rowptr.data(), A.col.data(), A.val.data(), b.data()));
        int it = 0;
        double rel = 0.0;
        int rc = yc_solve(cloth.context(), nullptr, tol, maxit < 0 ? 2 * A.n : maxit, &it, &rel);
        if (rc != YC_OK && rc != YC_ERR_NOT_CONVERGED) check(rc);   // the reference never reports non-convergence
        if (iters) *iters = it;
        Vec x(A.n);
        check(yc_export_vec(cloth.context(), 3, x.data()));
        return x;
    }
};

// SimStep (src/TrainHelper.cpp:15-97): false on NaN/Inf in A, b or the new velocity.
inline bool SimStep(Cloth& cloth, PhysicsCloth& phy, double h, const Vec* ext_force = nullptr, double tol = 1e-8) {
    (void)phy;
    if (ext_force) check(yc_set_ext_force(cloth.context(), ext_force->data()));
    yc_step_info info;
    int rc = yc_step(cloth.context(), h, nullptr, tol, 2 * cloth.ndof(), &info);
    if (ext_force) check(yc_set_ext_force(cloth.context(), nullptr));
    if (rc == YC_ERR_NONFINITE) return false;
    if (rc != YC_OK && rc != YC_ERR_NOT_CONVERGED) check(rc);
    return true;
}

}  // namespace yarnstep

#endif

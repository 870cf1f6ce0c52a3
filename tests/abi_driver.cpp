// abi_driver.cpp -- a reference-style driver written against include/yarnstep.hpp (the
// C++ mirror of the reference's step API over the C ABI).  It runs `steps` steps of the
// reference's inline step (src/Simulate.h:86-96: FillMV, A and b formed from get_SpaM /
// get_SpaFPos / get_SpaFVel, EigenSolver::apply, Cloth::update) on one cloth and `steps`
// SimSteps (src/TrainHelper.cpp:15-97) on a second cloth from the same state, and writes
// both final states.  tests/test_gpu_boundary.py compares them with the oracle.
//
// usage: abi_driver <input> <output>
//   input:  yc_desc (raw struct), double h, double tol, int steps, then the ClothState
//           arrays LPos, LPosFix, LVel (N x 3 doubles each), EPos, EPosBar, EVel (N x 2)
//   output: the inline-step cloth's ClothState arrays, then the SimStep cloth's, then
//           the inline step's CG iterations per step (ints)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "yarnstep.hpp"

using namespace yarnstep;

static bool read_all(FILE* f, void* p, size_t n) { return std::fread(p, 1, n, f) == n; }
static void write_state(FILE* f, const ClothState& s) {
    for (const Vec* v : {&s.LPos, &s.LPosFix, &s.LVel, &s.EPos, &s.EPosBar, &s.EVel})
        std::fwrite(v->data(), sizeof(double), v->size(), f);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) { std::perror("input"); return 2; }
    yc_desc d;
    double h = 0.0, tol = 0.0;
    int steps = 0;
    bool ok = read_all(in, &d, sizeof d) && read_all(in, &h, sizeof h) && read_all(in, &tol, sizeof tol) &&
              read_all(in, &steps, sizeof steps);
    const size_t N = (size_t)d.width * d.length;
    ClothState s0{Vec(3 * N), Vec(3 * N), Vec(3 * N), Vec(2 * N), Vec(2 * N), Vec(2 * N)};
    for (Vec* v : {&s0.LPos, &s0.LPosFix, &s0.LVel, &s0.EPos, &s0.EPosBar, &s0.EVel})
        ok = ok && read_all(in, v->data(), v->size() * sizeof(double));
    std::fclose(in);
    if (!ok) { std::fprintf(stderr, "short input\n"); return 2; }
    try {
        Cloth cloth(d), cloth2(d);
        cloth.set_ClothState(s0);
        cloth2.set_ClothState(s0);
        PhysicsCloth phy(cloth), phy2(cloth2);
        std::vector<int> iters;
        for (int i = 0; i < steps; ++i) {
            // src/Simulate.h:86-94
            phy.FillMV(h);
            const Csr A = phy.get_SpaM() - phy.get_SpaFPos() * (h * h) - phy.get_SpaFVel() * h;
            const Vec b = h * (phy.get_SpaF() - phy.get_SpaFVel().tmul(phy.get_SpaVel())) + phy.get_SpaM() * phy.get_SpaVel();
            int it = 0;
            Vec new_vel = EigenSolver::apply(cloth, A, b, tol, -1, &it);
            cloth.update(new_vel, phy.F_E_is_Slides[0], h);
            phy.ResetMV();
            iters.push_back(it);
            if (!SimStep(cloth2, phy2, h, nullptr, tol)) { std::fprintf(stderr, "SimStep: NaN/Inf\n"); return 1; }
        }
        FILE* out = std::fopen(argv[2], "wb");
        if (!out) { std::perror("output"); return 2; }
        write_state(out, cloth.get_ClothState());
        write_state(out, cloth2.get_ClothState());
        std::fwrite(iters.data(), sizeof(int), iters.size(), out);
        std::fclose(out);
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::printf("abi_driver: %d steps of W=%d done\n", steps, d.width);
    return 0;
}

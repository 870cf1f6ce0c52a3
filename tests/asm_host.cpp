// Host build of the first assembly kernel's work for a tile (assemble1_block,
// fine-grained-..._amd/csrc/assemble1.hpp), for the undefined-behaviour check of that
// code: compiled with g++ under AddressSanitizer and UndefinedBehaviorSanitizer, every GPU
// thread of a block run as a std::thread, the block's LDS an exact-size global array (the
// sanitizer's red zones catch any index past it), blocks one after another.  The same
// function the GPU kernels call, through the same declarations (elements.hpp,
// assembly.hpp).
//
// Input file: yc_desc (raw), h, then the state in the kernels' planes: x (3N), v (3N),
// xfix (3N), uv (2N), uvd (2N), uvbar (2N), nrm (3N).  Output file: the count of A's
// stored slots, then their rows, columns (int32, the solver's global numbering: 3 L DOFs
// per node, then 2 E DOFs per interior crossing) and values, then F (5N), b (5 planes of
// P), comp (N), is_slide (2N), fepart (2N) and the two status flags as doubles.
#include <barrier>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#include "yarnstep.h"
#include "elements.hpp"
#include "assembly.hpp"
#include "assemble1.hpp"

using namespace ys;

thread_local HostDim3 threadIdx, blockIdx, blockDim;
static std::barrier<>* g_block_barrier = nullptr;
void host_syncthreads() { g_block_barrier->arrive_and_wait(); }

// the block's LDS, exactly the dynamic size the GPU kernel is launched with (in the
// unnamed namespace the kernel code's extern declaration refers to)
namespace {
double lds_rows[ASM1_LDS / sizeof(double)];
}


static bool read_all(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "usage: asm_host IN OUT\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    yc_desc d;
    double h;
    if (!read_all(f, &d, sizeof d) || !read_all(f, &h, sizeof h)) return 2;
    ClothDev C = cloth_params(d);
    const int N = C.N, W = C.W, Lg = C.Lg;
    std::vector<double> x(3 * N), v(3 * N), xfix(3 * N), uv(2 * N), uvd(2 * N), uvbar(2 * N), nrm(3 * N);
    for (auto* a : {&x, &v, &xfix, &uv, &uvd, &uvbar, &nrm})
        if (!read_all(f, a->data(), a->size() * sizeof(double))) return 2;
    fclose(f);
    std::vector<unsigned char> flags(N, 0);
    for (int j = 0; j < Lg; ++j)
        for (int i = 0; i < W; ++i)
            if (i > 0 && i < W - 1 && j > 0 && j < Lg - 1 && pattern_flip(d.pattern, i, j)) flags[i + j * W] |= 1;
    for (int q = 0; q < d.n_handles; ++q) flags[d.handles[2 * q] + d.handles[2 * q + 1] * W] |= 2;
    C.x = x.data(); C.v = v.data(); C.xfix = xfix.data(); C.uv = uv.data(); C.uvd = uvd.data(); C.uvbar = uvbar.data();
    C.nrm = nrm.data(); C.flags = flags.data();
    const int nty = (Lg + TILE - 1) / TILE, nT = C.ntx * nty;
    const int Wp = W + 4, Hp = Lg + 4, P = ((Wp * Hp + 63) / 64) * 64;
    std::vector<double> A((size_t)nT * NSLOT * TILE_N, 0.0), F(5 * N, 0.0), b(5 * P, 0.0), comp(N, 0.0), is_slide(2 * N, 0.0),
        fepart(2 * N, 0.0), invd(5 * P, 0.0);
    int flag[2] = {0, 0};
    AsmOut O;
    std::memset((void*)&O, 0, sizeof O);
    O.A = A.data(); O.F = F.data(); O.b = b.data(); O.is_slide = is_slide.data(); O.comp = comp.data(); O.fepart = fepart.data();
    O.invdiag = invd.data(); O.ext = nullptr; O.dext = nullptr; O.flag = flag; O.Wp = Wp; O.P = P;
    O.dA = O.dF = O.db = O.dis_slide = O.dcomp = nullptr; O.detach = 0;
    O.shear_rec = nullptr; O.nseg = 4 * C.Nc; O.bend_rec = nullptr; O.only_kind = -1;
    for (int tile = 0; tile < nT; ++tile) {
        std::barrier<> bar(ASM1_THREADS);
        g_block_barrier = &bar;
        std::vector<std::thread> th;
        for (int t = 0; t < ASM1_THREADS; ++t)
            th.emplace_back([&, t] {
                threadIdx = {(unsigned)t, 0u, 0u};
                blockIdx = {(unsigned)tile, 0u, 0u};
                blockDim = {(unsigned)ASM1_THREADS, 1u, 1u};
                assemble1_block<double, false>(C, O, h, tile);
            });
        for (auto& t : th) t.join();
    }
    // the stored slots in global numbering (slot_decode gives each slot's row component,
    // column node offset and column component)
    auto dof = [&](int u, int vv, int c) -> int {
        if (c < 3) return 3 * (u + vv * W) + c;
        if (u <= 0 || u >= W - 1 || vv <= 0 || vv >= Lg - 1) return -1;
        return 3 * N + 2 * ((u - 1) + (vv - 1) * (W - 2)) + (c - 3);
    };
    std::vector<int> ri, ci;
    std::vector<double> va;
    for (int vv = 0; vv < Lg; ++vv)
        for (int u = 0; u < W; ++u) {
            const size_t base = a_base(C.ntx, u, vv);
            for (int sl = 0; sl < NSLOT; ++sl) {
                const double a = A[base + (size_t)sl * TILE_N];
                if (a == 0.0) continue;
                int rc, du, dv, cc;
                slot_decode(sl, rc, du, dv, cc);
                const int r = dof(u, vv, rc), c = dof(u + du, vv + dv, cc);
                if (r < 0 || c < 0 || u + du < 0 || u + du >= W || vv + dv < 0 || vv + dv >= Lg) { flag[0] |= 4; continue; }
                ri.push_back(r); ci.push_back(c); va.push_back(a);
            }
        }
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    const long long nnz = (long long)va.size();
    fwrite(&nnz, sizeof nnz, 1, g);
    fwrite(ri.data(), sizeof(int), ri.size(), g);
    fwrite(ci.data(), sizeof(int), ci.size(), g);
    fwrite(va.data(), sizeof(double), va.size(), g);
    for (auto* a : {&F, &b, &comp, &is_slide, &fepart}) fwrite(a->data(), sizeof(double), a->size(), g);
    double fl[2] = {(double)flag[0], (double)flag[1]};
    fwrite(fl, sizeof(double), 2, g);
    fclose(g);
    return 0;
}

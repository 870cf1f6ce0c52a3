// batch_adjoint.hpp -- the reverse mode of many cloths at once (included at the end of
// yarnstep.hip, after batch.hpp).
//
// A batch records its members' rollouts in lockstep (each member's own adjoint recording,
// checkpoint interval 1) and runs their backward passes in lockstep: every phase of a
// backward step (normals' adjoint, state reload, the system's recomputation, the update's
// adjoint, the backward solve, the assembly's vector-Jacobian product, the loss gradient)
// is one launch over all cloths, through the batch's block -> cloth maps, and the solves
// are the batch's one solve launch (k_b_cgp).  Each cloth's arithmetic is its single-cloth
// backward pass's (the same device bodies on its own buffers).  The reference trains one
// cloth per forward/backward pass in its drivers (src/TrainHeter.h:161-231).

namespace {

// One member's adjoint buffers and descriptors, in the batch's layout order.
struct AdjItem {
    ClothDev C;        // seed codes (state passes)
    ClothDev Cp[3];    // element parameter passes: rho, kb and handle_k; ks; kc
    ClothDev Cf[3];    // friction parameter passes: mu, kf, df
    ClothDev Cs;       // the shear at S = 1 (its S derivative)
    AdjW w;
    double* g[2];      // state adjoints by parity
    double *gpar, *gcomp, *ybar, *pbar, *gslide, *fric_in, *lam, *xs, *cx, *rhs;
    const double *comp, *is_slide;
    StateDev st;       // the context's state planes
    double* ckpt;      // s_0, s_1, ... (one block per recorded step)
    double* fin;       // the state after the last recorded step
    const double *tx, *tu;
    double tw;
    double* lpart;
    SpmvArgs sp;       // the backward solve's matrix times cx, into cy
    const double* A;
    double* At;
    int t0, n0, adj_max;
};

__device__ __forceinline__ int ba_node(const AdjItem& a, int nb) { return ((int)blockIdx.x - a.n0) * nb + (int)threadIdx.x; }

__device__ __forceinline__ StateDev ba_block_state(const AdjItem& a, double* blk) {
    StateDev s = a.st;
    const size_t N = a.st.N;
    s.x = blk + G_X * N; s.v = blk + G_V * N; s.xfix = blk + G_XFIX * N; s.uv = blk + G_UV * N;
    s.uvd = blk + G_UVD * N; s.uvbar = blk + G_UVBAR * N; s.nrm = blk + G_NRM * N;
    return s;
}

__device__ __forceinline__ double* ba_state_blk(const AdjItem& a, int n, int T) {
    return n == T ? a.fin : a.ckpt + (size_t)n * G_PLANES * a.st.N;
}

// node kernels (blocks of nb threads, the batch's node map)
__global__ void __launch_bounds__(256) k_ba_zero(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb) {
    const AdjItem& a = it[map[blockIdx.x]];
    const int k = ba_node(a, nb), N = a.st.N;
    if (k >= N) return;
    for (int q = 0; q < G_PLANES; ++q) { a.g[0][(size_t)q * N + k] = 0.0; a.g[1][(size_t)q * N + k] = 0.0; }
    for (int q = 0; q < YC_NPARAM; ++q) a.gpar[(size_t)q * N + k] = 0.0;
    a.lpart[k] = 0.0;
    const size_t pk = (size_t)padidx(a.w.Wp, k % a.st.W, k / a.st.W);
    for (int c = 0; c < 5; ++c) a.lam[(size_t)c * a.w.P + pk] = 0.0;
}

__global__ void __launch_bounds__(256) k_ba_ckpt(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb, int n, int T,
                                                 int save) {
    const AdjItem& a = it[map[blockIdx.x]];
    const int k = ba_node(a, nb);
    if (save && n < T && n >= a.adj_max) return;   // a recording's last state lives in fin
    ckpt_node(a.st, ba_state_blk(a, n, T), save, k);
}

__global__ void __launch_bounds__(256) k_ba_nrm(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb, int n, int T,
                                                int p) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_nrm_node(ba_block_state(a, ba_state_blk(a, n, T)), a.g[p] + G_NRM * (size_t)a.st.N, a.pbar, ba_node(a, nb));
}

__global__ void __launch_bounds__(256) k_ba_fold(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb, int p) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_fold_nrm_node(a.st.W, a.st.Lg, a.pbar, a.g[p], ba_node(a, nb));
}

// the step's solution from the state after it (xs), the update's adjoint, and lambda of the
// step after it as the backward solve's initial guess (cx)
__global__ void __launch_bounds__(256) k_ba_update(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb, int n, int T,
                                                   int p, double h) {
    const AdjItem& a = it[map[blockIdx.x]];
    const int k = ba_node(a, nb), N = a.st.N, W = a.st.W, Lg = a.st.Lg;
    if (k >= N) return;
    const int u = k % W, v = k / W;
    const size_t pk = (size_t)padidx(a.w.Wp, u, v), P = a.w.P;
    const double* s1 = ba_state_blk(a, n + 1, T);
    const bool inter = u > 0 && u < W - 1 && v > 0 && v < Lg - 1;
    for (int c = 0; c < 3; ++c) a.xs[c * P + pk] = s1[(size_t)(G_V + c) * N + k];
    for (int c = 0; c < 2; ++c) a.xs[(3 + c) * P + pk] = inter ? s1[(size_t)(G_UVD + c) * N + k] : 0.0;
    adj_update_node(a.st.flags, W, Lg, a.g[p], a.pbar, a.is_slide, a.xs, a.w.Wp, a.w.P, h, a.rhs, a.gslide, a.g[1 - p], k);
    for (int c = 0; c < 5; ++c) a.cx[c * P + pk] = a.lam[c * P + pk];
}

__global__ void __launch_bounds__(256) k_ba_keep_lam(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb) {
    const AdjItem& a = it[map[blockIdx.x]];
    const int k = ba_node(a, nb);
    if (k >= a.st.N) return;
    const size_t pk = (size_t)padidx(a.w.Wp, k % a.st.W, k / a.st.W), P = a.w.P;
    for (int c = 0; c < 5; ++c) a.lam[c * P + pk] = a.cx[c * P + pk];
}

// the targets' loss of the state after step n (>= 1) into g[q]
__global__ void __launch_bounds__(256) k_ba_loss(const AdjItem* __restrict__ it, const int* __restrict__ map, int nb, int n, int T,
                                                 int q) {
    const AdjItem& a = it[map[blockIdx.x]];
    const size_t N = a.st.N;
    const double* s = ba_state_blk(a, n, T);
    adj_target_loss_node(a.st.N, s + G_X * N, s + G_UV * N, a.tx + (size_t)(n - 1) * 3 * N, a.tu + (size_t)(n - 1) * 2 * N, a.tw,
                         a.g[q], a.lpart, ba_node(a, nb));
}

// tile kernels
__global__ void __launch_bounds__(NBLK) k_ba_transpose(const AdjItem* __restrict__ it, const int* __restrict__ map) {
    const AdjItem& a = it[map[blockIdx.x]];
    transpose_tile(a.A, a.At, a.st.W, a.st.Lg, a.sp.ntx, blockIdx.x - a.t0);
}

__global__ void __launch_bounds__(BLK) k_ba_spmv(const AdjItem* __restrict__ it, const int* __restrict__ map) {
    const AdjItem& a = it[map[blockIdx.x]];
    spmv_tile(a.sp, nullptr, blockIdx.x - a.t0);
}

__global__ void __launch_bounds__(BLK) k_ba_cg_init(const BatchItem* __restrict__ it, const int* __restrict__ map) {
    const BatchItem& b = it[map[blockIdx.x]];
    cg_init_block(b.va, b.part, 1, blockIdx.x - b.v0);
}

__device__ __forceinline__ AdjOut ba_out(const AdjItem& a, int p) {
    AdjOut o;
    o.g = a.g[1 - p]; o.gpar = a.gpar; o.gcomp = a.gcomp; o.ybar = a.ybar;
    return o;
}

template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_ba_shear(const AdjItem* __restrict__ it, const int* __restrict__ map, int p) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_shear_tile<REF>(a.C, a.w, ba_out(a, p), a.comp, blockIdx.x - a.t0);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_ba_shear_par(const AdjItem* __restrict__ it, const int* __restrict__ map, int p) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_shear_par_tile<REF>(a.Cs, a.w, ba_out(a, p), a.comp, blockIdx.x - a.t0);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_ba_fric(const AdjItem* __restrict__ it, const int* __restrict__ map, int p, int fpar) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_fric_tile<REF>(fpar < 0 ? a.C : a.Cf[fpar], a.w, ba_out(a, p), a.fric_in, a.gslide, fpar, blockIdx.x - a.t0);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_ba_elems(const AdjItem* __restrict__ it, const int* __restrict__ map, int p) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_elems_tile<REF>(a.C, a.w, ba_out(a, p), blockIdx.x - a.t0);
}
template <bool REF>
__global__ void __launch_bounds__(TILE_N) k_ba_elems_par(const AdjItem* __restrict__ it, const int* __restrict__ map, int p, int pass) {
    const AdjItem& a = it[map[blockIdx.x]];
    adj_elems_par_tile<REF>(a.Cp[pass], a.w, ba_out(a, p), pass, blockIdx.x - a.t0);
}

// per member: the YC_NPARAM parameter sums and the loss (grid: members x (YC_NPARAM + 1))
__global__ void __launch_bounds__(256) k_ba_sums(const AdjItem* __restrict__ it, double* out) {
    const AdjItem& a = it[blockIdx.x];
    const int q = blockIdx.y, N = a.st.N;
    const double* src = q < YC_NPARAM ? a.gpar + (size_t)q * N : a.lpart;
    __shared__ double s[256];
    double x = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) x += src[i];
    s[threadIdx.x] = x;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * (YC_NPARAM + 1) + q] = s[0];
}

// The members' recordings must all be active, with checkpoint interval 1 and the same
// number of recorded steps.
int ba_check(yc_batch* b, int need_steps) {
    const int T = b->m[0]->adj_steps;
    for (yc_ctx* c : b->m) {
        if (c->adj_max <= 0) return fail(YC_ERR_ARG, "batch adjoint: a member has no recording (yc_adjoint_begin)");
        if (c->adj_K != 1) return fail(YC_ERR_ARG, "batch adjoint: members record with checkpoint interval 1");
        if (c->adj_steps != T) return fail(YC_ERR_ARG, "batch adjoint: members have recorded different numbers of steps");
        if (need_steps > 0 && c->adj_steps + need_steps > c->adj_max) return fail(YC_ERR_ARG, "more steps than yc_adjoint_begin allowed");
    }
    return YC_OK;
}

int ba_upload(yc_batch* b, int mode, double h, AdjItem** d_adj) {
    const int n = (int)b->m.size();
    std::vector<AdjItem> ai(n);
    for (int li = 0; li < n; ++li) {
        yc_ctx* c = b->m[b->order[li]];
        AdjItem& a = ai[li];
        a.C = make_adj_dev(c, mode);
        for (int p = 0; p < 3; ++p) {
            a.Cp[p] = a.C;
            a.Cf[p] = a.C;
        }
        a.Cp[0].drho[0] = a.Cp[0].drho[1] = 1.0; a.Cp[0].dkb[0] = a.Cp[0].dkb[1] = 1.0; a.Cp[0].dhandle_k = 1.0;
        a.Cp[1].dks[0] = a.Cp[1].dks[1] = 1.0;
        a.Cp[2].dkc = 1.0;
        a.Cf[0].dmu = 1.0; a.Cf[1].dkf = 1.0; a.Cf[2].ddf = 1.0;
        a.Cs = a.C;
        a.Cs.S = 1.0;
        a.w.lam = c->adj_lam; a.w.xs = c->adj_xs; a.w.ybar = c->adj_ybar; a.w.Wp = c->Wp; a.w.P = c->P; a.w.h = h;
        a.g[0] = c->adj_g0; a.g[1] = c->adj_g1;
        a.gpar = c->adj_gpar; a.gcomp = c->adj_gcomp; a.ybar = c->adj_ybar; a.pbar = c->adj_pbar; a.gslide = c->adj_gslide;
        a.fric_in = c->adj_fric_in; a.lam = c->adj_lam; a.xs = c->adj_xs; a.cx = c->cx; a.rhs = c->b;
        a.comp = c->comp; a.is_slide = c->is_slide;
        a.st = make_state(c);
        a.ckpt = c->adj_ckpt; a.fin = c->adj_final;
        a.tx = c->adj_tx; a.tu = c->adj_tu; a.tw = c->adj_tw; a.lpart = c->adj_lpart;
        a.A = c->A; a.At = c->At;
        a.sp = make_spmv(c, c->cx, c->cy, mode == YC_GRAD_EXACT ? c->At : c->A);
        a.t0 = b->items[li].t0; a.n0 = b->items[li].n0; a.adj_max = c->adj_max;
    }
    HIPCHK(hipMalloc((void**)d_adj, sizeof(AdjItem) * n));
    HIPCHK(hipMemcpyAsync(*d_adj, ai.data(), sizeof(AdjItem) * n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));   // ai is the source of the copy
    return YC_OK;
}

// the per-member status of the queued work (k_b_status), as yc_batch_run reports it
int ba_status(yc_batch* b, const BatchItem* items, long long* iters, int* status) {
    const int n = (int)b->m.size();
    hipLaunchKernelGGL(k_b_status, dim3((n + 63) / 64), dim3(64), 0, b->stream, items, n, b->d_status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->h_status, b->d_status, sizeof(long long) * BSTAT * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    int worst = YC_OK;
    auto rank = [](int code) { return code == YC_ERR_HIP ? 3 : code == YC_ERR_NONFINITE ? 2 : code == YC_ERR_NOT_CONVERGED ? 1 : 0; };
    for (int li = 0; li < n; ++li) {
        const long long* st = b->h_status + BSTAT * li;
        const int i = b->order[li];
        int code = YC_OK;
        if (st[5] || st[3]) code = YC_ERR_HIP;
        else if (st[4] || st[2]) code = YC_ERR_NONFINITE;
        else if (st[1]) code = YC_ERR_NOT_CONVERGED;
        if (st[5]) (void)hipMemsetAsync(b->m[i]->cg_err, 0, sizeof(int), b->stream);
        if (iters) iters[i] = st[0];
        if (status) status[i] = code;
        if (rank(code) > rank(worst)) worst = code;
    }
    (void)hipStreamSynchronize(b->stream);
    return worst;
}

}  // namespace

extern "C" {

int yc_batch_adjoint_run(yc_batch* b, double h, int steps, double tol, int maxit, long long* iters, int* status) {
    g_err.clear();
    if (!b) return fail(YC_ERR_ARG, "null batch");
    if (steps < 0 || !(tol >= 0.0)) return fail(YC_ERR_ARG, "batch: bad steps or tol");
    int rc = ba_check(b, steps);
    if (rc) return rc;
    const int n = (int)b->m.size();
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(b->device));
    struct Restore { int d; ~Restore() { (void)hipSetDevice(d); } } restore{cur};
    for (yc_ctx* c : b->m) HIPCHK(hipStreamSynchronize(c->stream));
    if (!b->laid_out && (rc = batch_layout(b))) return rc;
    if ((rc = batch_upload_items(b, tol, maxit))) return rc;
    AdjItem* d_adj = nullptr;
    if ((rc = ba_upload(b, YC_GRAD_REF, h, &d_adj))) return rc;
    const BatchItem* it = b->d_items;
    const int nb = b->opts.node_block;
    hipLaunchKernelGGL(k_b_reset, dim3((n + 63) / 64), dim3(64), 0, b->stream, it, n);
    const int t0 = b->m[0]->adj_steps;
    for (int s = 0; s < steps && !rc; ++s) {
        rc = batch_step(b, h);
        if (!rc) hipLaunchKernelGGL(k_ba_ckpt, dim3(b->n_node), dim3(nb), 0, b->stream, (const AdjItem*)d_adj, (const int*)b->map_node, nb,
                                    t0 + s + 1, 1 << 30, 1);
    }
    if (!rc) HIPCHK(hipGetLastError());
    int worst = rc ? rc : ba_status(b, it, iters, status);
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(d_adj);
    for (yc_ctx* c : b->m) {
        c->assembled = c->solved = false;
        c->adj_steps = worst ? 0 : t0 + steps;   // a failed step voids every member's recording
    }
    if (worst) return worst == rc ? rc : fail(worst, "batch adjoint recording: a member's step failed; the recordings are void");
    return YC_OK;
}

int yc_batch_adjoint_backward(yc_batch* b, int mode, double h, double tol, int maxit, double* dparam, double* loss, long long* iters,
                              int* status) {
    g_err.clear();
    if (!b) return fail(YC_ERR_ARG, "null batch");
    if (mode != YC_GRAD_REF && mode != YC_GRAD_EXACT) return fail(YC_ERR_ARG, "unknown gradient mode");
    int rc = ba_check(b, 0);
    if (rc) return rc;
    const int T = b->m[0]->adj_steps;
    if (T <= 0) return fail(YC_ERR_ARG, "batch adjoint: no recorded step");
    for (yc_ctx* c : b->m)
        if (!c->adj_tx) return fail(YC_ERR_ARG, "batch adjoint: a member has no targets (yc_adjoint_set_targets)");
    const int n = (int)b->m.size();
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(b->device));
    struct Restore { int d; ~Restore() { (void)hipSetDevice(d); } } restore{cur};
    for (yc_ctx* c : b->m) HIPCHK(hipStreamSynchronize(c->stream));
    const bool exact = mode == YC_GRAD_EXACT;
    for (yc_ctx* c : b->m) {
        if (exact && !c->At && (rc = dalloc(&c->At, c->a_size))) return rc;
        c->fric_out = c->adj_fric_in;   // the recomputed assemblies keep the friction inputs
        c->bj_of = nullptr;
    }
    if (!b->laid_out && (rc = batch_layout(b))) return rc;
    rc = batch_upload_items(b, tol, maxit);
    std::vector<BatchItem> items2 = b->items;   // the solves' matrix: A (reference gradient) or A^T
    BatchItem* d_items2 = nullptr;
    AdjItem* d_adj = nullptr;
    if (!rc) {
        for (int li = 0; li < n; ++li) if (exact) items2[li].cg.A = b->m[b->order[li]]->At;
        if (hipMalloc((void**)&d_items2, sizeof(BatchItem) * n) != hipSuccess) rc = fail(YC_ERR_HIP, "batch adjoint: descriptors");
        else if (hipMemcpy(d_items2, items2.data(), sizeof(BatchItem) * n, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(YC_ERR_HIP, "batch adjoint: descriptor upload");
    }
    if (!rc) rc = ba_upload(b, mode, h, &d_adj);
    const int nb = b->opts.node_block;
    const int* mt = b->map_tile;
    const int* mn = b->map_node;
    const AdjItem* ai = d_adj;
    const BatchItem* it2 = d_items2;
    auto nodes = [&](auto kern, auto... args) {
        hipLaunchKernelGGL(kern, dim3(b->n_node), dim3(nb), 0, b->stream, ai, mn, nb, args...);
    };
    auto tiles = [&](auto kern, int threads, auto... args) {
        hipLaunchKernelGGL(kern, dim3(b->n_tile), dim3(threads), 0, b->stream, ai, mt, args...);
    };
    if (!rc) {
        hipLaunchKernelGGL(k_b_reset, dim3((n + 63) / 64), dim3(64), 0, b->stream, it2, n);
        nodes(k_ba_ckpt, T, T, 1);    // the state after the last step, restored at the end
        nodes(k_ba_zero);
        nodes(k_ba_loss, T, T, 0);
        int p = 0;
        for (int s = T - 1; s >= 0 && !rc; --s) {
            nodes(k_ba_nrm, s + 1, T, p);
            nodes(k_ba_ckpt, s, T, 0);
            hipLaunchKernelGGL(k_b_assemble1, dim3(b->n_tile), dim3(ASM1_THREADS), ASM1_LDS, b->stream, it2, mt, h);
            if ((rc = batch_node_phase(b, h, 0))) break;   // the shear records of every member
            hipLaunchKernelGGL(k_b_assemble2, dim3(b->n_tile), dim3(ASM2_THREADS), 0, b->stream, it2, mt, h);
            if (exact) tiles(k_ba_transpose, NBLK);
            if (b->bj)
                for (yc_ctx* c : b->m) {
                    const int ngx = (c->W + 1) / 2, ngy = (c->Lg + 1) / 2, ng = ngx * ngy;
                    hipLaunchKernelGGL(k_bj_factor, dim3((ng + BJF_GROUPS - 1) / BJF_GROUPS), dim3(64), 0, b->stream,
                                       (const double*)(exact ? c->At : c->A), c->bj, c->W, c->Lg, c->Wp, c->P, c->ntx, ngx, ng);
                }
            nodes(k_ba_update, s, T, p, h);
            tiles(k_ba_spmv, BLK);
            hipLaunchKernelGGL(k_ba_cg_init, dim3(b->n_vec), dim3(BLK), 0, b->stream, it2, (const int*)b->map_vec);
            hipLaunchKernelGGL(k_b_epoch, dim3((n + 63) / 64), dim3(64), 0, b->stream, it2, n);
            {
                const int* gp = b->d_groups;
                int ng = (int)b->groups.size();
                void* args[] = {(void*)&it2, (void*)&mt, (void*)&gp, (void*)&ng};
                rc = persist_launch(b->device, b->stream, [&]() -> int {
                    if (hipLaunchKernel(b_cgp_fn(b->kp, b->bj), dim3(b->max_group), dim3(BLK), args, 0, b->stream) != hipSuccess)
                        return fail(YC_ERR_HIP, "batch adjoint: solve launch");
                    return YC_OK;
                });
                if (rc) break;
            }
            nodes(k_ba_keep_lam);
            if (exact) {
                tiles(k_ba_shear<false>, TILE_N, p);
                tiles(k_ba_shear_par<false>, TILE_N, p);
                tiles(k_ba_fric<false>, TILE_N, p, -1);
                for (int f = 0; f < 3; ++f) tiles(k_ba_fric<false>, TILE_N, p, f);
                tiles(k_ba_elems<false>, TILE_N, p);
                for (int q = 0; q < 3; ++q) tiles(k_ba_elems_par<false>, TILE_N, p, q);
            } else {
                tiles(k_ba_shear<true>, TILE_N, p);
                tiles(k_ba_shear_par<true>, TILE_N, p);
                tiles(k_ba_fric<true>, TILE_N, p, -1);
                for (int f = 0; f < 3; ++f) tiles(k_ba_fric<true>, TILE_N, p, f);
                tiles(k_ba_elems<true>, TILE_N, p);
                for (int q = 0; q < 3; ++q) tiles(k_ba_elems_par<true>, TILE_N, p, q);
            }
            if (s >= 1) nodes(k_ba_loss, s, T, 1 - p);
            if (hipGetLastError() != hipSuccess) { rc = fail(YC_ERR_HIP, "batch adjoint: launch"); break; }
            p ^= 1;
        }
        if (!rc) {   // the initial states' normals come from their x
            nodes(k_ba_nrm, 0, T, p);
            nodes(k_ba_fold, p);
        }
    }
    std::vector<double> sums((size_t)n * (YC_NPARAM + 1));
    double* d_sums = nullptr;
    if (!rc && hipMalloc((void**)&d_sums, sizeof(double) * sums.size()) != hipSuccess) rc = fail(YC_ERR_HIP, "batch adjoint: sums");
    if (!rc) {
        hipLaunchKernelGGL(k_ba_sums, dim3(n, YC_NPARAM + 1), dim3(256), 0, b->stream, ai, d_sums);
        if (hipMemcpyAsync(sums.data(), d_sums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
            rc = fail(YC_ERR_HIP, "batch adjoint: download");
    }
    int worst = rc;
    if (!rc) worst = ba_status(b, it2, iters, status);
    if (d_adj) nodes(k_ba_ckpt, T, T, 0);   // every member back to its last recorded state
    (void)hipStreamSynchronize(b->stream);
    if (d_sums) (void)hipFree(d_sums);
    if (d_adj) (void)hipFree(d_adj);
    if (d_items2) (void)hipFree(d_items2);
    for (yc_ctx* c : b->m) {
        c->fric_out = nullptr;
        c->bj_of = nullptr;
        c->assembled = c->solved = false;
    }
    if (worst) return worst == rc ? rc : fail(worst, "batch adjoint: a member's backward solve or system failed");
    for (int li = 0; li < n; ++li) {
        const int i = b->order[li];
        yc_ctx* c = b->m[i];
        double* sm = sums.data() + (size_t)li * (YC_NPARAM + 1);
        const double area = PI_REF * std::pow(c->d.R, 2);   // ks = Y pi R^2, kb = B pi R^2 (src/Yarn.h:14-15)
        sm[YC_P_Y0] *= area; sm[YC_P_Y1] *= area; sm[YC_P_B0] *= area; sm[YC_P_B1] *= area;
        if (c->d.n_yarns == 1) {
            sm[YC_P_RHO0] += sm[YC_P_RHO1]; sm[YC_P_Y0] += sm[YC_P_Y1]; sm[YC_P_B0] += sm[YC_P_B1];
            sm[YC_P_RHO1] = sm[YC_P_Y1] = sm[YC_P_B1] = 0.0;
        }
        if (dparam) for (int q = 0; q < YC_NPARAM; ++q) dparam[(size_t)i * YC_NPARAM + q] = sm[q];
        if (loss) loss[i] = sm[YC_NPARAM];
    }
    return YC_OK;
}

}  // extern "C"

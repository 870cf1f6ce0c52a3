// batch.hpp -- many cloths of uneven size stepped together (included at the end of
// yarnstep.hip, whose kernels' per-tile / per-node bodies it reuses).
//
// A batch is a list of contexts on one device.  Each phase of a step is one launch over
// all of them: a block looks up its cloth in a block -> cloth map and runs the same body
// as the single-cloth kernel on that cloth's buffers, so every cloth's arithmetic is the
// single-cloth path's, bit for bit.  A step is
//   k_b_epoch, k_b_assemble1, k_b_shear, k_b_assemble2, k_b_cg_init,
//   k_b_cgp (one launch whose blocks are all resident; they walk groups of cloths that
//   fit on the device at once, block b taking tile b of each group in turn),
//   k_b_update, k_b_normals
// with no host synchronisation; a run replays it from a hipGraph where the runtime
// captures the step, and launches it directly otherwise.  The reference
// steps one Cloth per PhysicsCloth (src/TrainHelper.cpp:15-97) and its training drivers
// loop over cloths (src/TrainHeter.h:81-146); batching has no counterpart there.

namespace {

struct BatchItem {
    ClothDev C;
    AsmOut O;
    VecArgs va;
    CGPart part;
    CgpArgs cg;
    StateDev st;
    unsigned long long* epoch;   // the context's device epoch counter
    double* shear_rec;
    int t0, s0, v0, n0;          // first block of the cloth in the tile, segment, vector and node launches
    int nseg;
};

// per-cloth outcome of a run: iterations, solves at maxit, non-finite solves, slot error,
// non-finite A/b, persistent-CG wait error
constexpr int BSTAT = 6;

__global__ void __launch_bounds__(64) k_b_reset(const BatchItem* __restrict__ it, int n) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const BatchItem& b = it[i];
    b.O.flag[0] = 0;
    b.O.flag[1] = 0;
    b.cg.sc->total_iters = 0;
    b.cg.sc->nonconv = 0;
    b.cg.sc->nonfinite = 0;
    *b.cg.err = 0;
}

__global__ void __launch_bounds__(64) k_b_epoch(const BatchItem* __restrict__ it, int n) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) *it[i].epoch += 1;
}

__global__ void __launch_bounds__(64) k_b_status(const BatchItem* __restrict__ it, int n, long long* out) {
    int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const BatchItem& b = it[i];
    out[BSTAT * i + 0] = b.cg.sc->total_iters;
    out[BSTAT * i + 1] = b.cg.sc->nonconv;
    out[BSTAT * i + 2] = b.cg.sc->nonfinite;
    out[BSTAT * i + 3] = b.O.flag[0];
    out[BSTAT * i + 4] = b.O.flag[1];
    out[BSTAT * i + 5] = *b.cg.err;
}

// the first assembly kernel's work (assemble1_block) on the block's cloth
__global__ void __launch_bounds__(ASM1_THREADS, 1) k_b_assemble1(const BatchItem* __restrict__ it, const int* __restrict__ map, double h) {
    const BatchItem& bi = it[map[blockIdx.x]];
    assemble1_block<double, false>(bi.C, bi.O, h, blockIdx.x - bi.t0);
}

__global__ void __launch_bounds__(ASM2_THREADS) k_b_assemble2(const BatchItem* __restrict__ it, const int* __restrict__ map, double h) {
    const BatchItem& b = it[map[blockIdx.x]];
    assemble2_value(b.C, b.O, h, blockIdx.x - b.t0);
}

template <int NB>
__global__ void __launch_bounds__(NB) k_b_shear(const BatchItem* __restrict__ it, const int* __restrict__ map) {
    const BatchItem& b = it[map[blockIdx.x]];
    shear_elem_seg(b.C, b.O.comp, b.shear_rec, b.nseg, (blockIdx.x - b.s0) * NB + threadIdx.x);
}

__global__ void __launch_bounds__(BLK) k_b_cg_init(const BatchItem* __restrict__ it, const int* __restrict__ map) {
    const BatchItem& b = it[map[blockIdx.x]];
    cg_init_block(b.va, b.part, 0, blockIdx.x - b.v0);
}

// The solves of all cloths in one launch, an ordinary one whose grid (the largest group)
// fits the device at the kernel's occupancy, so all of its blocks are resident (as for
// k_cgp, a cooperative launch would add a cross-queue wait on each side).  Group g holds tiles
// [grp[2g], grp[2g] + grp[2g+1]) of the tile numbering; block b solves tile b of group 0,
// then tile b of group 1, and so on.  A cloth's tiles are all in one group, so its
// blocks are distinct and resident; a block that reaches a cloth early waits (bounded)
// for peers still finishing an earlier group's cloth, which needs no block of a later
// group, so every wait ends.  Blocks freed by a cloth that converged early go on to the
// next group at once.
// With the block-Jacobi rows in registers a tile needs a whole SIMD's registers (one block
// per CU); the Jacobi form fits two blocks per CU.
template <int KP, bool BJ>
__global__ void __launch_bounds__(BLK, BJ ? 1 : 2) k_b_cgp(const BatchItem* __restrict__ it, const int* __restrict__ map,
                                               const int* __restrict__ grp, int ngroups) {
    for (int g = 0; g < ngroups; ++g) {
        if ((int)blockIdx.x < grp[2 * g + 1]) {
            const int blk = grp[2 * g] + blockIdx.x;
            const BatchItem& b = it[map[blk]];
            cgp_tile<KP, BJ>(b.cg, blk - b.t0, *b.epoch);
        }
        __syncthreads();   // the tile's LDS is reused by the next group's tile
    }
}

template <int NB>
__global__ void __launch_bounds__(NB) k_b_update(const BatchItem* __restrict__ it, const int* __restrict__ map, double h) {
    const BatchItem& b = it[map[blockIdx.x]];
    update_node(b.st, b.cg.x, b.O.is_slide, b.O.Wp, b.O.P, h, (blockIdx.x - b.n0) * NB + threadIdx.x);
}

template <int NB>
__global__ void __launch_bounds__(NB) k_b_normals(const BatchItem* __restrict__ it, const int* __restrict__ map) {
    const BatchItem& b = it[map[blockIdx.x]];
    normals_node(b.st, (blockIdx.x - b.n0) * NB + threadIdx.x);
}

// bj: the 2x2-block-Jacobi preconditioner (every cloth of the batch uses it), else Jacobi
const void* b_cgp_fn(int kp, bool bj) {
    if (bj) {
        switch (kp) {
            case 1: return (const void*)k_b_cgp<1, true>;
            case 2: return (const void*)k_b_cgp<2, true>;
            case 3: return (const void*)k_b_cgp<3, true>;
            default: return (const void*)k_b_cgp<4, true>;
        }
    }
    switch (kp) {
        case 1: return (const void*)k_b_cgp<1, false>;
        case 2: return (const void*)k_b_cgp<2, false>;
        case 3: return (const void*)k_b_cgp<3, false>;
        default: return (const void*)k_b_cgp<4, false>;
    }
}

}  // namespace

struct yc_batch {
    std::vector<yc_ctx*> m;        // members in the caller's order
    std::vector<int> order;        // layout order (indices into m)
    int device = 0;
    hipStream_t stream = nullptr;
    yc_batch_opts opts;
    int kp = 1, capacity = 0;
    bool bj = false;                           // every member uses the 2x2-block-Jacobi preconditioner
    int n_tile = 0, n_seg = 0, n_vec = 0, n_node = 0;
    std::vector<std::pair<int, int>> groups;   // (first tile, tiles) of each group of the solve launch
    int max_group = 0;
    std::vector<BatchItem> items;              // host copy, layout order
    BatchItem* d_items = nullptr;
    int* d_maps = nullptr;                     // tile, segment, vector, node maps and the groups, back to back
    int *map_tile = nullptr, *map_seg = nullptr, *map_vec = nullptr, *map_node = nullptr, *d_groups = nullptr;
    long long* d_status = nullptr;
    long long* h_status = nullptr;             // pinned
    hipGraphExec_t graph = nullptr;
    double graph_h = 0.0;
    std::vector<char> graph_ext;               // external-force flags the graph's per-cloth launches were captured with
    int graph_failed = 0, graph_used = 0;
    bool laid_out = false;
};

namespace {

int batch_free_layout(yc_batch* b) {
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    if (b->d_items) { (void)hipFree(b->d_items); b->d_items = nullptr; }
    if (b->d_maps) { (void)hipFree(b->d_maps); b->d_maps = nullptr; }
    b->laid_out = false;
    return YC_OK;
}

// Block offsets, block -> cloth maps and resident groups for the current options.
int batch_layout(yc_batch* b) {
    batch_free_layout(b);
    const int n = (int)b->m.size();
    const int nb = b->opts.node_block;
    b->order.resize(n);
    for (int i = 0; i < n; ++i) b->order[i] = i;
    if (b->opts.group_order == 1)
        std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return b->m[x]->nT > b->m[y]->nT; });
    b->kp = 1;
    for (yc_ctx* c : b->m) b->kp = std::max(b->kp, c->cg1_kp);
    int per_cu = 0, ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b_cgp_fn(b->kp, b->bj), BLK, 0));
    b->capacity = per_cu * ncu;
    std::vector<int> mt, ms, mv, mn;
    b->items.assign(n, BatchItem());
    b->groups.clear();
    int gstart = 0, gsize = 0;
    for (int li = 0; li < n; ++li) {
        yc_ctx* c = b->m[b->order[li]];
        if (c->nT > b->capacity)
            return fail(YC_ERR_ARG, "batch: a " + std::to_string(c->W) + "x" + std::to_string(c->Lg) +
                                        " cloth's solve does not fit on the device at once");
        if (gsize + c->nT > b->capacity) {
            b->groups.push_back({gstart, gsize});
            gstart += gsize;
            gsize = 0;
        }
        gsize += c->nT;
        BatchItem& it = b->items[li];
        it.t0 = (int)mt.size();
        it.s0 = (int)ms.size();
        it.v0 = (int)mv.size();
        it.n0 = (int)mn.size();
        it.nseg = 4 * c->Nc;
        mt.insert(mt.end(), c->nT, li);
        ms.insert(ms.end(), (it.nseg + nb - 1) / nb, li);
        mv.insert(mv.end(), c->part.nV, li);
        mn.insert(mn.end(), (c->N + nb - 1) / nb, li);
    }
    if (gsize) b->groups.push_back({gstart, gsize});
    b->max_group = 0;
    for (const auto& g : b->groups) b->max_group = std::max(b->max_group, g.second);
    b->n_tile = (int)mt.size(); b->n_seg = (int)ms.size(); b->n_vec = (int)mv.size(); b->n_node = (int)mn.size();
    std::vector<int> maps;
    maps.reserve(mt.size() + ms.size() + mv.size() + mn.size());
    maps.insert(maps.end(), mt.begin(), mt.end());
    maps.insert(maps.end(), ms.begin(), ms.end());
    maps.insert(maps.end(), mv.begin(), mv.end());
    maps.insert(maps.end(), mn.begin(), mn.end());
    for (const auto& g : b->groups) { maps.push_back(g.first); maps.push_back(g.second); }
    HIPCHK(hipMalloc((void**)&b->d_items, sizeof(BatchItem) * n));
    HIPCHK(hipMalloc((void**)&b->d_maps, sizeof(int) * maps.size()));
    HIPCHK(hipMemcpy(b->d_maps, maps.data(), sizeof(int) * maps.size(), hipMemcpyHostToDevice));
    b->map_tile = b->d_maps;
    b->map_seg = b->map_tile + mt.size();
    b->map_vec = b->map_seg + ms.size();
    b->map_node = b->map_vec + mv.size();
    b->d_groups = b->map_node + mn.size();
    b->laid_out = true;
    return YC_OK;
}

// Per-cloth descriptors for this run (current external forces, tolerance, maxit).
int batch_upload_items(yc_batch* b, double tol, int maxit) {
    for (size_t li = 0; li < b->items.size(); ++li) {
        yc_ctx* c = b->m[b->order[li]];
        BatchItem& it = b->items[li];
        c->cd = make_dev(c);
        it.C = c->cd;
        it.O = make_out(c);
        it.va = make_vec(c);
        it.part = c->part;
        CgpArgs& a = it.cg;
        a.A = c->A; a.invd = c->invd; a.bbpart = c->part.bb; a.rrpart = c->part.rr;
        a.brow = b->bj ? c->bj : nullptr;   // the members' common preconditioner
        a.x = c->cx;
        a.r0 = c->cr; a.w[0] = c->cy; a.w[1] = c->cw1;
        a.slot[0] = c->cgslot; a.slot[1] = c->cgslot + c->nT;
        a.sc = c->sc; a.err = c->cg_err;
        a.epoch = 0;   // taken from the device counter
        a.tol = tol; a.maxit = maxit < 0 ? 2 * c->n : maxit; a.nbb = c->part.nV;
        a.spin_limit = 1 << 22;   // a block may also wait for peers still solving an earlier group's cloth
        a.W = c->W; a.Lg = c->Lg; a.Wp = c->Wp; a.P = c->P; a.ntx = c->ntx; a.nT = c->nT;
        it.st = make_state(c);
        it.epoch = c->bepoch;
        it.shear_rec = c->shear_rec;
    }
    HIPCHK(hipMemcpy(b->d_items, b->items.data(), sizeof(BatchItem) * b->items.size(), hipMemcpyHostToDevice));
    return YC_OK;
}

template <int NB>
int batch_node_launches(yc_batch* b, double h, int phase) {
    if (phase == 0)
        hipLaunchKernelGGL(k_b_shear<NB>, dim3(b->n_seg), dim3(NB), 0, b->stream, (const BatchItem*)b->d_items,
                           (const int*)b->map_seg);
    else {
        hipLaunchKernelGGL(k_b_update<NB>, dim3(b->n_node), dim3(NB), 0, b->stream, (const BatchItem*)b->d_items,
                           (const int*)b->map_node, h);
        hipLaunchKernelGGL(k_b_normals<NB>, dim3(b->n_node), dim3(NB), 0, b->stream, (const BatchItem*)b->d_items,
                           (const int*)b->map_node);
    }
    HIPCHK(hipGetLastError());
    return YC_OK;
}

int batch_node_phase(yc_batch* b, double h, int phase) {
    switch (b->opts.node_block) {
        case 64: return batch_node_launches<64>(b, h, phase);
        case 128: return batch_node_launches<128>(b, h, phase);
        default: return batch_node_launches<256>(b, h, phase);
    }
}

// One step of every cloth, queued on the batch stream.
int batch_step(yc_batch* b, double h) {
    const int n = (int)b->m.size();
    for (yc_ctx* c : b->m) c->bj_of = nullptr;   // the batch rewrites every A
    const BatchItem* it = b->d_items;
    hipLaunchKernelGGL(k_b_epoch, dim3((n + 63) / 64), dim3(64), 0, b->stream, it, n);
    hipLaunchKernelGGL(k_b_assemble1, dim3(b->n_tile), dim3(ASM1_THREADS), ASM1_LDS, b->stream, it, (const int*)b->map_tile, h);
    HIPCHK(hipGetLastError());
    int rc = batch_node_phase(b, h, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_b_assemble2, dim3(b->n_tile), dim3(ASM2_THREADS), 0, b->stream, it, (const int*)b->map_tile, h);
    hipLaunchKernelGGL(k_b_cg_init, dim3(b->n_vec), dim3(BLK), 0, b->stream, it, (const int*)b->map_vec);
    if (b->bj)   // each cloth's block-Jacobi rows from its new A, on the batch stream
        for (yc_ctx* c : b->m) {
            const int ngx = (c->W + 1) / 2, ngy = (c->Lg + 1) / 2, ng = ngx * ngy;
            hipLaunchKernelGGL(k_bj_factor, dim3((ng + BJF_GROUPS - 1) / BJF_GROUPS), dim3(64), 0, b->stream,
                               (const double*)c->A, c->bj, c->W, c->Lg, c->Wp, c->P, c->ntx, ngx, ng);
        }
    HIPCHK(hipGetLastError());
    {
        const int* mp = b->map_tile;
        const int* gp = b->d_groups;
        int ng = (int)b->groups.size();
        void* args[] = {(void*)&it, (void*)&mp, (void*)&gp, (void*)&ng};
        auto launch = [&]() -> int {
            HIPCHK(hipLaunchKernel(b_cgp_fn(b->kp, b->bj), dim3(b->max_group), dim3(BLK), args, 0, b->stream));
            return YC_OK;
        };
        // while a step is captured into a graph the order is kept when the graph is launched
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(b->stream, &cap));
        int rc2 = cap == hipStreamCaptureStatusNone ? persist_launch(b->device, b->stream, launch) : launch();
        if (rc2) return rc2;
    }
    return batch_node_phase(b, h, 1);
}

// Captures one step into a graph; false (and no error) when the runtime cannot capture it.
bool batch_capture(yc_batch* b, double h) {
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    std::string saved = g_err;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    int rc = batch_step(b, h);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    bool ok = rc == YC_OK && e == hipSuccess && g != nullptr;
    if (ok && hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0) != hipSuccess) {
        b->graph = nullptr;
        ok = false;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    g_err = saved;
    if (ok) {
        b->graph_h = h;
        b->graph_ext.clear();
        for (yc_ctx* c : b->m) b->graph_ext.push_back(c->has_ext ? 1 : 0);
    }
    return ok;
}

}  // namespace

extern "C" {

int yc_batch_create(yc_ctx* const* ctxs, int n, yc_batch** out) {
    g_err.clear();
    if (!ctxs || !out || n <= 0) return fail(YC_ERR_ARG, "batch: null argument or empty batch");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return fail(YC_ERR_ARG, "batch: null context");
        if (ctxs[i]->device != ctxs[0]->device) return fail(YC_ERR_ARG, "batch: contexts on different devices");
        if (ctxs[i]->precond != ctxs[0]->precond)
            return fail(YC_ERR_ARG, "batch: the cloths of a batch use one preconditioner");
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(YC_ERR_ARG, "batch: a context appears twice");
    }
    yc_batch* b = new yc_batch();
    b->m.assign(ctxs, ctxs + n);
    b->device = ctxs[0]->device;
    b->bj = ctxs[0]->precond == YC_PRECOND_BLOCK2;
    b->opts.graph = 1;
    b->opts.group_order = 0;
    b->opts.node_block = 64;
    int cur = 0;
    auto bail = [&](int rc) { yc_batch_destroy(b); (void)hipSetDevice(cur); return rc; };
    if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(b->device) != hipSuccess) return bail(fail(YC_ERR_HIP, "batch: device"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(YC_ERR_HIP, "batch: stream"));
    if (hipFuncSetAttribute((const void*)k_b_assemble1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ASM1_LDS) != hipSuccess)
        return bail(fail(YC_ERR_HIP, "batch: assembly LDS size not supported"));
    if (hipMalloc((void**)&b->d_status, sizeof(long long) * BSTAT * n) != hipSuccess ||
        hipHostMalloc((void**)&b->h_status, sizeof(long long) * BSTAT * n) != hipSuccess)
        return bail(fail(YC_ERR_HIP, "batch: status buffers"));
    int rc = batch_layout(b);
    (void)hipSetDevice(cur);
    if (rc) return bail(rc);
    *out = b;
    return YC_OK;
}

void yc_batch_destroy(yc_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    persist_forget(b->device, b->stream);
    batch_free_layout(b);
    if (b->d_status) (void)hipFree(b->d_status);
    if (b->h_status) (void)hipHostFree(b->h_status);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int yc_batch_set_opts(yc_batch* b, const yc_batch_opts* o) {
    if (!b || !o) return fail(YC_ERR_ARG, "null argument");
    if (o->node_block != 64 && o->node_block != 128 && o->node_block != 256)
        return fail(YC_ERR_ARG, "batch: node_block must be 64, 128 or 256");
    if (o->group_order != 0 && o->group_order != 1) return fail(YC_ERR_ARG, "batch: group_order must be 0 or 1");
    if (o->graph != 0 && o->graph != 1) return fail(YC_ERR_ARG, "batch: graph must be 0 or 1");
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(b->device));
    (void)hipStreamSynchronize(b->stream);
    b->opts = *o;
    int rc = batch_layout(b);
    (void)hipSetDevice(cur);
    return rc;
}

int yc_batch_get_opts(const yc_batch* b, yc_batch_opts* o) {
    if (!b || !o) return fail(YC_ERR_ARG, "null argument");
    *o = b->opts;
    return YC_OK;
}

int yc_batch_info(const yc_batch* b, int* n_groups, int* graph_used, int* total_tiles, int* capacity) {
    if (!b) return fail(YC_ERR_ARG, "null batch");
    if (n_groups) *n_groups = (int)b->groups.size();
    if (graph_used) *graph_used = b->graph_used;
    if (total_tiles) *total_tiles = b->n_tile;
    if (capacity) *capacity = b->capacity;
    return YC_OK;
}

int yc_batch_run(yc_batch* b, double h, int steps, double tol, int maxit, long long* iters, int* status) {
    g_err.clear();
    if (!b) return fail(YC_ERR_ARG, "null batch");
    if (steps < 0 || !(tol >= 0.0)) return fail(YC_ERR_ARG, "batch: bad steps or tol");
    const int n = (int)b->m.size();
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(b->device));
    struct Restore { int d; ~Restore() { (void)hipSetDevice(d); } } restore{cur};
    for (yc_ctx* c : b->m) HIPCHK(hipStreamSynchronize(c->stream));
    if (!b->laid_out) {
        int rc = batch_layout(b);
        if (rc) return rc;
    }
    int rc = batch_upload_items(b, tol, maxit);
    if (rc) return rc;
    hipLaunchKernelGGL(k_b_reset, dim3((n + 63) / 64), dim3(64), 0, b->stream, (const BatchItem*)b->d_items, n);
    HIPCHK(hipGetLastError());
    b->graph_used = 0;
    if (steps > 0 && b->opts.graph && !b->graph_failed) {
        std::vector<char> ext;
        for (yc_ctx* c : b->m) ext.push_back(c->has_ext ? 1 : 0);
        if (!b->graph || b->graph_h != h || b->graph_ext != ext) {
            HIPCHK(hipStreamSynchronize(b->stream));
            if (!batch_capture(b, h)) b->graph_failed = 1;
        }
        b->graph_used = b->graph != nullptr;
    }
    for (int s = 0; s < steps; ++s) {
        if (b->graph_used) {
            rc = persist_launch(b->device, b->stream, [&]() -> int {
                HIPCHK(hipGraphLaunch(b->graph, b->stream));
                return YC_OK;
            });
            if (rc) return rc;
        }
        else if ((rc = batch_step(b, h))) return rc;
    }
    hipLaunchKernelGGL(k_b_status, dim3((n + 63) / 64), dim3(64), 0, b->stream, (const BatchItem*)b->d_items, n, b->d_status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->h_status, b->d_status, sizeof(long long) * BSTAT * n, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    int worst = YC_OK, worst_i = -1;
    auto rank = [](int code) { return code == YC_ERR_HIP ? 3 : code == YC_ERR_NONFINITE ? 2 : code == YC_ERR_NOT_CONVERGED ? 1 : 0; };
    for (int li = 0; li < n; ++li) {
        const long long* st = b->h_status + BSTAT * li;
        int i = b->order[li];
        yc_ctx* c = b->m[i];
        c->assembled = false;
        c->solved = false;
        int code = YC_OK;
        if (st[5] || st[3]) code = YC_ERR_HIP;
        else if (st[4] || st[2]) code = YC_ERR_NONFINITE;
        else if (st[1]) code = YC_ERR_NOT_CONVERGED;
        if (iters) iters[i] = st[0];
        if (status) status[i] = code;
        if (rank(code) > rank(worst)) { worst = code; worst_i = i; }
    }
    if (worst != YC_OK) {
        const long long* st = nullptr;
        for (int li = 0; li < n; ++li) if (b->order[li] == worst_i) st = b->h_status + BSTAT * li;
        std::string what = st && st[5] ? "persistent CG: a block waited past its spin limit"
                         : st && st[3] ? "assembly: stencil slot out of range"
                         : worst == YC_ERR_NONFINITE ? "NaN/Inf in A, b or a solution"
                         : "a CG solve stopped at maxit above tolerance";
        if (st && st[5]) {   // on the batch's stream, ahead of its next launches
            (void)hipMemsetAsync(b->m[worst_i]->cg_err, 0, sizeof(int), b->stream);
            (void)hipStreamSynchronize(b->stream);
        }
        return fail(worst, "batch cloth " + std::to_string(worst_i) + ": " + what);
    }
    return YC_OK;
}

}  // extern "C"

/*
 * yarnstep.h -- C ABI of the MI355X (gfx950) implicit yarn-level time step.
 *
 * One opaque context holds one cloth: its structure-of-arrays node state in HBM,
 * the 159-slot stencil system matrix, the right-hand side and the CG workspaces.
 * All entry points return an int status:
 *   YC_OK (0), YC_ERR_NONFINITE (1: NaN/Inf in A, b or the solution -- the cases
 *   where the reference's SimStep returns false, src/TrainHelper.cpp:22-48),
 *   YC_ERR_NOT_CONVERGED (2: CG stopped at maxit above tol), YC_ERR_HIP (3),
 *   YC_ERR_ARG (4: bad argument).  No C++ exception crosses this boundary.
 * Host arrays passed in are borrowed for the duration of the call.  A context is
 * bound to the device that was current at yc_create and uses one stream; it is not
 * thread-safe, separate contexts are independent.
 *
 * Index conventions are the reference's:
 *   node k = u + v*width, crossing e = (u-1) + (v-1)*(width-2),
 *   global vector [L: 3k+{x,y,z}] ++ [E: 3N + 2e+{u,v}]      (src/MatrixMp.cpp:120-122)
 *   state arrays use index u + v*width: get/set_ClothState's u + v*length
 *   (src/Cloth.cpp:362,426) for the square cloths it addresses one to one, and the
 *   trajectory loader's order (src/Loader.cpp:160) for non-square ones.
 */
#ifndef YARNSTEP_H
#define YARNSTEP_H

#ifdef __cplusplus
extern "C" {
#endif

#define YC_OK 0
#define YC_ERR_NONFINITE 1
#define YC_ERR_NOT_CONVERGED 2
#define YC_ERR_HIP 3
#define YC_ERR_ARG 4

#define YC_MAX_HANDLES 16

/* Cloth description: the arguments of the reference's Cloth constructor
 * (src/Cloth.h:55-60, src/Cloth.cpp:3-50), its Yarn list (src/Yarn.h:6-17), the
 * Environment's gravity and wind (src/Environment.h:6-22, src/Wind.h:6-35) and
 * the handle list of Cloth::set_handles (src/Cloth.cpp:386-394). */
typedef struct {
    int width, length;
    double hight, L, R;
    int n_yarns;                 /* 1 or 2 (warp row / weft column i uses yarn i%2) */
    double rho[2], Y[2], B[2];
    double kc, kf, df, mu, S, handle_k;
    int pattern;                 /* 0 plain, 1 twill, 2 satin */
    int pose;                    /* 0 upright, 1 flat */
    double G[3];
    double wind_vel[3], wind_density, wind_drag;
    int n_handles;
    int handles[2 * YC_MAX_HANDLES];
} yc_desc;

typedef struct yc_ctx yc_ctx;

typedef struct {
    int cg_iters;        /* iterations of the last solve (reference semantics: completed iterations) */
    double cg_relres;    /* |b - A x| / |b| estimate from the recursive residual */
    int status;
} yc_step_info;

/* Replaces: Cloth::Cloth + PhysicsCloth::PhysicsCloth + InitMV
 * (src/Cloth.cpp:3-50, src/PhysicsCloth.h:22-23, src/PhysicsCloth.cpp:3-44). */
int yc_create(const yc_desc* desc, yc_ctx** out);
/* Replaces: the destructors of Cloth / PhysicsCloth. */
void yc_destroy(yc_ctx* ctx);
/* Last error message of this thread (empty string when none). */
const char* yc_last_error(void);

/* Sizes: N nodes, Nc crossings, n = 3N + 2Nc DOFs, nnz slots of the stencil matrix. */
int yc_sizes(const yc_ctx* ctx, int* N, int* Nc, int* n, int* nslots);

/* Replaces: Cloth::set_ClothState (src/Cloth.cpp:421-438), followed by the
 * geometry refresh of UpdateClothUnits (src/Cloth.cpp:307-348).  Host arrays:
 * lpos/lposfix/lvel N*3 doubles, epos/eposbar/evel N*2 doubles. */
int yc_set_state(yc_ctx* ctx, const double* lpos, const double* lposfix, const double* lvel,
                 const double* epos, const double* eposbar, const double* evel);
/* Replaces: Cloth::get_ClothState (src/Cloth.cpp:350-371).  Any pointer may be NULL. */
int yc_get_state(yc_ctx* ctx, double* lpos, double* lposfix, double* lvel,
                 double* epos, double* eposbar, double* evel);

/* yc_set_state / yc_get_state with device arrays (on the context's device, same layouts):
 * the state moves between the caller's HBM buffers and the context without a host copy.
 * `stream` is the caller's hipStream_t (NULL: the null stream): the transfer starts after
 * the work queued on it so far, and work queued on it afterwards starts after the
 * transfer, so the caller may reuse or read its buffers in stream order; neither call
 * waits on the host.  No counterpart in the reference (its state is CPU tensors); for
 * callers that keep the state in device memory, such as a LibTorch or PyTorch training
 * loop on the GPU.  Every pointer of yc_set_state_device is required; any of
 * yc_get_state_device may be NULL. */
int yc_set_state_device(yc_ctx* ctx, const double* lpos, const double* lposfix, const double* lvel,
                        const double* epos, const double* eposbar, const double* evel, void* stream);
int yc_get_state_device(yc_ctx* ctx, double* lpos, double* lposfix, double* lvel,
                        double* epos, double* eposbar, double* evel, void* stream);

/* External generalized force added to b (the ext_force of
 * SimStep(cloth, phys, h, ext_force), src/TrainHelper.cpp:56-63).  n doubles in the
 * global layout, copied to the device; NULL clears it. */
int yc_set_ext_force(yc_ctx* ctx, const double* ext_force);

/* Replaces: PhysicsCloth::FillMV (src/PhysicsCloth.cpp:46-156) and the system
 * formation A = M - h^2 K_x - h K_v, b = h (F + f_ext - K_v^T v) + M v
 * (src/TrainHelper.cpp:19-20, 61-63).  Returns YC_ERR_NONFINITE on NaN/Inf in A or b. */
int yc_assemble(yc_ctx* ctx, double h);

/* Replaces: get_SpaM/get_SpaFPos/get_SpaFVel-derived A (src/PhysicsCloth.h:72-77,
 * src/MatrixMp.cpp:51-110) for inspection: the structural pattern of A as CSR of the
 * global n x n matrix.  rowptr n+1 ints, col/val nnz entries (nnz from yc_export_csr
 * with col == val == NULL). */
int yc_export_csr(yc_ctx* ctx, int* rowptr, int* col, double* val, int* nnz);
/* One matrix of the assembled system in the reference's global layout (dense in the
 * reference, CSR here, with A's structural pattern):
 *   YC_MAT_A   A = M - h^2 K_x - h K_v, as yc_export_csr;
 *   YC_MAT_M   get_SpaM (src/PhysicsCloth.h:72);
 *   YC_MAT_KX  get_SpaFPos, the position Jacobian of the forces (src/PhysicsCloth.h:74);
 *   YC_MAT_KV  get_SpaFVel, the velocity Jacobian (src/PhysicsCloth.h:75).
 * L-L blocks are placed transposed as To_Sparse_Mat places them (src/MatrixMp.cpp:65-76).
 * M, K_x and K_v are re-evaluated from the state and step of the last yc_assemble (a
 * second assembly pass; the context's system is left as assembled).  Pointers as
 * yc_export_csr: with rowptr, col and val NULL only nnz is returned. */
#define YC_MAT_A 0
#define YC_MAT_M 1
#define YC_MAT_KX 2
#define YC_MAT_KV 3
int yc_export_matrix(yc_ctx* ctx, int which, int* rowptr, int* col, double* val, int* nnz);
/* which: 0 F (get_SpaF), 1 b, 2 v (get_SpaVel), 3 last solution x, 4 positions
 * (get_SpaPos, src/PhysicsCloth.h:76: LPos, then EPos of the crossings).  out: n doubles. */
int yc_export_vec(yc_ctx* ctx, int which, double* out);
/* Per crossing: is_slide (Nc*2, the F_E_is_Slides[0] of src/PhysicsCloth.cpp:2633),
 * compression (Nc), friction input F_E partial (Nc*2).  Any pointer may be NULL. */
int yc_export_crossing(yc_ctx* ctx, double* is_slide, double* comp, double* fe_partial);
/* Crossing normals N*3 (zero on the boundary), as Cloth::NodeNorm (src/Cloth.cpp:440-463). */
int yc_export_normals(yc_ctx* ctx, double* nrm);

/* Replaces: EigenSolver::apply forward (src/Solvers.cpp:3-31): Jacobi-preconditioned
 * CG on the full A from x0 (n doubles, NULL = zero guess) until |r|^2 < tol^2 |b|^2 or
 * maxit iterations.  Returns YC_ERR_NOT_CONVERGED when stopped at maxit. */
int yc_solve(yc_ctx* ctx, const double* x0, double tol, int maxit, int* iters, double* relres);

/* Replaces: Cloth::update(new_vel, F_E_is_Slides[0], h) (src/Cloth.cpp:111-148) with
 * the last solution as new_vel, including the geometry refresh. */
int yc_update(yc_ctx* ctx, double h);

/* Replaces: Cloth::update(new_vel, F_E_is_Slide, h) (src/Cloth.cpp:111-148) with a
 * caller's velocity: new_vel n doubles in the global layout, is_slide Nc x 2 doubles
 * (NULL: the last assembly's F_E_is_Slides[0]).  For callers that form the step
 * themselves, as the reference's inline step does (src/Simulate.h:86-96). */
int yc_update_vel(yc_ctx* ctx, const double* new_vel, const double* is_slide, double h);

/* Replaces: the input side of EigenSolver::apply(A, b) (src/Solvers.h:35-52) for a system
 * the caller formed, e.g. from the yc_export_matrix pieces as src/Simulate.h:86-90 forms
 * it: A as CSR of the global n x n matrix (rowptr n+1, col/val by row; entries may repeat
 * and are summed) and b (n doubles) become the context's system, and yc_solve then solves
 * it with the reference's Jacobi preconditioner (its diagonal, 1/a_ii or 1 where zero).
 * A non-zero entry outside the yarn stencil (A's structural pattern) is YC_ERR_ARG;
 * NaN/Inf in A or b is YC_ERR_NONFINITE. */
int yc_load_system(yc_ctx* ctx, const int* rowptr, const int* col, const double* val, const double* b);

/* Replaces: SimStep (src/TrainHelper.cpp:15-97): assemble, solve, update.  x0 as in
 * yc_solve (NULL = zero guess).  info may be NULL. */
int yc_step(yc_ctx* ctx, double h, const double* x0, double tol, int maxit, yc_step_info* info);

/* CG variant of yc_solve / yc_step / yc_run and the warm-started forms.
 * YC_CG_REFERENCE (default): the reference solver's Jacobi-PCG recurrences in its order
 *   (src/Solvers.cpp:22-24), three kernels and two reductions per iteration.
 * YC_CG_FUSED: a single-reduction form of the same Jacobi-PCG (A u and A p carried as
 *   vectors, so p.Ap is known from the previous iteration's five inner products; no
 *   symmetry of A assumed) -- the same Krylov iterates in exact arithmetic, the same
 *   stopping rule on |r|^2 and the same iteration count convention -- with one kernel
 *   per iteration.  Solutions
 *   agree with the reference form to the solve tolerance; counts may differ by rounding. */
#define YC_CG_REFERENCE 0
#define YC_CG_FUSED 1
/* YC_CG_PERSISTENT: the fused recurrences in one launch per solve; each 8 x 8 tile's
 * block keeps its rows of A in registers for the whole solve.  Needs every tile resident
 * at once (YC_ERR_ARG otherwise, e.g. above about 128 x 128 crossings).  The blocks wait
 * on one another, so the persistent solves of all contexts and batches of a process run
 * one at a time per device (a solve queued on another stream waits for the work queued
 * there so far).  Kernels of other libraries or processes can only delay a solve; one
 * delayed past the spin limit (about a quarter second) fails with YC_ERR_HIP. */
#define YC_CG_PERSISTENT 2
int yc_set_cg_variant(yc_ctx* ctx, int variant);

/* Preconditioner of the CG solves.
 * YC_PRECOND_JACOBI (default): the reference solver's diagonal preconditioner
 *   (src/Solvers.cpp:3-31, Eigen's DiagonalPreconditioner).
 * YC_PRECOND_BLOCK2: block Jacobi over 2 x 2 groups of nodes (the exact inverse of A on
 *   each group's up to 20 DOFs, recomputed from A before every solve).  No counterpart
 *   in the reference: the same stopping rule on the unpreconditioned |r|^2, so the
 *   solution meets the same tolerance, in about half the iterations.  Persistent CG
 *   only: a solve with another variant, or on a cloth whose persistent CG does not fit,
 *   fails with YC_ERR_ARG. */
#define YC_PRECOND_JACOBI 0
#define YC_PRECOND_BLOCK2 1
int yc_set_preconditioner(yc_ctx* ctx, int kind);
/* Inspection hook: the block-Jacobi rows of the assembled A, n x 20 doubles, DOF i (global
 * layout) at rows[20 i]: the coefficients of its row of M^-1 over its 2 x 2 group, node
 * q = du + 2 dv from the group's origin (even u, v) and component c at 5 q + c, zero where
 * the group has no such DOF.  YC_PRECOND_BLOCK2 contexts only, after yc_assemble. */
int yc_export_precond(yc_ctx* ctx, double* rows);

/* Runs `steps` SimSteps back to back on the device with zero-guess CG; returns the
 * total CG iterations and the worst status.  With YC_CG_PERSISTENT the steps are queued
 * without any host synchronisation (assembly flags and solve outcomes accumulate on the
 * device), so a NaN/Inf or a solve stopped at maxit is reported after the last step;
 * the other variants synchronise once per CG chunk and stop at the first NaN/Inf. */
int yc_run(yc_ctx* ctx, double h, int steps, double tol, int maxit, long long* total_iters);

/* yc_run (persistent CG only) with a cheap initial guess in place of the zero one:
 * YC_GUESS_PREV the previous velocity v_n, YC_GUESS_EXTRAP its linear extrapolation
 * 2 v_n - v_{n-1}, YC_GUESS_QUADRATIC the quadratic one 3 v_n - 3 v_{n-1} + v_{n-2}
 * (earlier velocities from the previous steps of this context; a new state counts them
 * equal to v_n).  CG refines any of them to the same tolerance.  No counterpart in the
 * reference (its solver starts from zero, src/Solvers.cpp:22-24): the baselines a
 * learned warm start (yc_run_warm) is measured against. */
#define YC_GUESS_ZERO 0
#define YC_GUESS_PREV 1
#define YC_GUESS_EXTRAP 2
#define YC_GUESS_QUADRATIC 3
int yc_run_guess(yc_ctx* ctx, int guess, double h, int steps, double tol, int maxit, long long* total_iters);

/* Times `reps` launches of the CG loop's stencil product y = A p on the context's
 * stream with HIP events; avg_ms receives the mean duration of one launch.  Call after
 * yc_assemble; it overwrites the CG workspace. */
int yc_time_spmv(yc_ctx* ctx, int reps, double* avg_ms);

/* Algorithmic HBM bytes of one product launch: 8 bytes per structural nonzero of A
 * plus 16 bytes per DOF (p read, A p written). */
double yc_spmv_bytes(const yc_ctx* ctx);
/* Times `reps` persistent-CG solves of the assembled system (YC_CG_PERSISTENT only) with
 * an event pair around each solve's launch; avg_ms the mean launch duration, iters
 * the iteration count of the solve.  Overwrites the CG workspace and the solution. */
int yc_time_solve(yc_ctx* ctx, double tol, int maxit, int reps, double* avg_ms, int* iters);
/* Algorithmic bytes of one Jacobi-PCG iteration streamed from memory: 8 per structural
 * nonzero of A plus 120 per DOF (15 vector passes); with YC_PRECOND_BLOCK2 the 20
 * coefficients of each DOF's preconditioner row replace its diagonal entry (272 per DOF). */
double yc_cg_iteration_bytes(const yc_ctx* ctx);
/* Times `reps` assemblies of A and b (the kernels of yc_assemble, without its status
 * read-back) on the context's stream with one HIP event pair; avg_ms the mean per
 * assembly.  Measurement hook with no counterpart in the reference. */
int yc_time_assemble(yc_ctx* ctx, double h, int reps, double* avg_ms);

/* ---------------------------------------------------------------------------
 * Learned CG warm start.  No counterpart in the reference (whose only network,
 * src/ForceNet.h, predicts control forces): a 32-128-128-128-5 MLP with BF16 weights
 * and FP32 accumulation on MFMA predicts the velocity change of the implicit step
 * per node; CG then starts from v + dv and refines to the same tolerance as from a
 * zero guess (yc_step stays the zero-guess fallback and yardstick).
 * Parameters: one flat FP32 array of yc_mlp_param_count() values in the order
 * W1[128x32] b1[128] W2[128x128] b2[128] W3[128x128] b3[128] W4[5x128] b4[5]
 * (row-major [out][in]).  Inputs are standardised with (x - in_mean) / in_std; the
 * outputs are in units of out_std.
 * ------------------------------------------------------------------------- */
typedef struct yc_mlp yc_mlp;

int yc_mlp_param_count(void);
int yc_mlp_create(yc_mlp** out);
void yc_mlp_destroy(yc_mlp* m);
int yc_mlp_set_params(yc_mlp* m, const float* params);   /* also resets the Adam state */
int yc_mlp_get_params(yc_mlp* m, float* params);
int yc_mlp_set_norm(yc_mlp* m, const float* in_mean /*32*/, const float* in_std /*32*/, const float* out_std /*5*/);
/* Training set on the device: X rows x 32 raw features, T rows x 5 targets in units
 * of out_std, M rows x 5 masks (0 for DOFs a node does not have). */
int yc_mlp_set_data(yc_mlp* m, const float* X, const float* T, const float* M, int rows);
/* Inference on host features: Y rows x 5 (units of out_std). */
int yc_mlp_forward(yc_mlp* m, const float* X, int rows, float* Y);
/* yc_mlp_forward without any device allocation: the features and outputs are staged in
 * a file-scope device array of the library (one per device, held until the process
 * ends) instead of buffers allocated per call, for callers that cannot allocate (tight
 * serving loops, code that must not call hipMalloc).  No counterpart in the reference.
 * rows at most 65,536 (YC_ERR_ARG beyond); results equal yc_mlp_forward's bit for bit.
 * Calls on one device share the staging array: calls that overlap on one device (from
 * several threads or networks at once) race and must be serialised by the caller. */
int yc_mlp_forward_noalloc(yc_mlp* m, const float* X, int rows, float* Y);
/* Loss (mean of the masked squared error) and its gradient on dataset rows idx[0..rows). */
int yc_mlp_grad(yc_mlp* m, const int* idx, int rows, float* grad, float* loss);
/* Adam over idx[0..rows) in batches of `batch` rows (remainder dropped); mean batch loss out. */
int yc_mlp_train(yc_mlp* m, const int* idx, int rows, int batch, float lr, float* mean_loss);

/* Times `reps` launches of the network's forward kernel on `rows` rows with HIP events
 * on the network's stream; avg_ms receives the mean duration of one launch. */
int yc_mlp_time_forward(yc_mlp* m, int rows, int reps, double* avg_ms);

/* Network features of every node (N x 32 floats) after yc_assemble, and after yc_solve
 * the solution's difference from the quadratic extrapolation x_q = 3 v_n - 3 v_{n-1} +
 * v_{n-2} (N x 5):
 * the (input, target) pairs the simulator produces for training.  The features: the
 * Jacobi correction D^-1 (b - A x_q) at the node (5) and its x, y, z part at the four
 * neighbours (12), the last two steps' velocity changes (10), the sliding offset
 * (uv - uvbar) / L (2), and the interior, handle and flipped flags (3). */
int yc_features(yc_ctx* ctx, float* out);
int yc_targets(yc_ctx* ctx, float* out);
/* The network's initial guess x0 = x_q + out_std * y in the solution vector (read it with
 * yc_export_vec), after yc_assemble and without solving: for callers that solve with a
 * solver of their own. */
int yc_warm_guess(yc_ctx* ctx, yc_mlp* m);
/* Times `reps` warm guesses (x_q, A x_q, and the one-launch features + network +
 * correction) and `reps` launches of the last kernel alone, with HIP events on the
 * context's stream, after yc_assemble; the solution vector is left unspecified. */
int yc_time_warm_guess(yc_ctx* ctx, yc_mlp* m, int reps, double* avg_ms, double* avg_mlp_ms);
/* SimStep with the network's warm start (x0 = x_q + out_std * y on the device). */
int yc_step_warm(yc_ctx* ctx, yc_mlp* m, double h, double tol, int maxit, yc_step_info* info);
int yc_run_warm(yc_ctx* ctx, yc_mlp* m, double h, int steps, double tol, int maxit, long long* total_iters);

/* ---------------------------------------------------------------------------
 * Differentiation: forward-mode tangents (directional derivatives) of SimStep.
 * Replaces: the gradient the reference obtains by LibTorch autograd through SimStep
 * with EigenSolver::apply's backward (src/Solvers.h:35-52, src/Solvers.cpp:33-58),
 * used by the parameter-learning drivers (src/TrainHeter.h, src/YarnSimNetHeter.h).
 * The tangent of a scalar loss along a parameter direction is the gradient's component
 * along it, so YC_NPARAM directions give the full parameter gradient.
 *   YC_GRAD_EXACT: the derivative of the step as computed (A's dependence included).
 *   YC_GRAD_REF:   the reference's gradient: A, M, K_v held constant (they are built
 *     outside the autograd graph, src/MatrixMp.cpp:46-98), the non-jamming shear angle
 *     held constant (.item(), src/PhysicsCloth.cpp:2720-2721), and the solve's backward
 *     taken with A instead of A^T (src/Solvers.cpp:51-53); as a tangent map that is a
 *     forward solve with A^T.
 * Parameter order of a direction: */
#define YC_P_RHO0 0
#define YC_P_RHO1 1
#define YC_P_Y0 2
#define YC_P_Y1 3
#define YC_P_B0 4
#define YC_P_B1 5
#define YC_P_S 6
#define YC_P_MU 7
#define YC_P_KC 8
#define YC_P_KF 9
#define YC_P_DF 10
#define YC_P_HANDLE_K 11
#define YC_NPARAM 12
#define YC_GRAD_REF 0
#define YC_GRAD_EXACT 1

/* Starts ndir tangent directions: dparam ndir x YC_NPARAM, dstate ndir x 15N doubles in
 * the yc_set_state order (lpos, lposfix, lvel N*3 each; epos, eposbar, evel N*2 each) or
 * NULL for zero state tangents.  Replaces any directions already held. */
int yc_tangent_begin(yc_ctx* ctx, int ndir, const double* dparam, const double* dstate);
/* SimStep (as yc_step, zero CG guess) and the tangents of all directions through it;
 * the tangent solves use tol and maxit too. */
int yc_step_tangent(yc_ctx* ctx, double h, int mode, double tol, int maxit, yc_step_info* info);
/* Tangents of the external force for the following yc_step_tangent calls: dext holds
 * ndir x n doubles in the global layout (the yc_set_ext_force layout), NULL clears them.
 * The force enters b = h (F + f_ext - K_v^T v) + M v, so its tangent adds h dext to the
 * tangent of b in both gradient modes: the path of the reference's control parameters
 * (src/YarnSimNet.h:87-246, SimStep with ext_force, src/TrainHelper.cpp:56-97). */
int yc_tangent_set_ext(yc_ctx* ctx, const double* dext);
/* State tangent of direction dir, in the yc_get_state layout; any pointer may be NULL. */
int yc_tangent_get(yc_ctx* ctx, int dir, double* dlpos, double* dlposfix, double* dlvel,
                   double* depos, double* deposbar, double* devel);
/* Frees the tangent buffers. */
int yc_tangent_end(yc_ctx* ctx);

/* ---------------------------------------------------------------------------------
 * Differentiation: reverse mode (adjoint) of a recorded rollout.
 * Replaces: loss.backward() through a rollout of SimSteps (src/TrainHeter.h:194-231),
 * i.e. LibTorch autograd through b = h (F + f_ext - K_v^T v) + M v (src/TrainHelper.cpp:
 * 15-97) and EigenSolver::apply's backward (src/Solvers.h:35-52, src/Solvers.cpp:33-58).
 * One backward solve per step whatever the number of parameters: YC_GRAD_REF solves with
 * A as the reference's backward does, with A, M, K_v and the non-jamming shear angle held
 * constant; YC_GRAD_EXACT solves with A^T and adds A's dependence (-lambda^T dA x).
 * The recording stores the state every checkpoint_every steps; the backward pass
 * recomputes the states in between (bit for bit: the recorded steps use the zero CG
 * guess) and every step's system from its state.
 * --------------------------------------------------------------------------------- */
/* Starts a recording at the current state, for up to max_steps steps. */
int yc_adjoint_begin(yc_ctx* ctx, int max_steps, int checkpoint_every);
/* One recorded SimStep (as yc_step with x0 = NULL); the external force in effect is used.
 * A step whose solve stops above tolerance (YC_ERR_NOT_CONVERGED) or fails is neither
 * taken nor recorded, as in yc_step_tangent. */
int yc_adjoint_step(yc_ctx* ctx, double h, double tol, int maxit, yc_step_info* info);
/* steps recorded SimSteps queued without a host round trip (persistent CG; other forms
 * step one by one).  When any of them fails the recording is void (it holds no step
 * afterwards; start again with yc_adjoint_begin) and the status says why. */
int yc_adjoint_run(yc_ctx* ctx, double h, int steps, double tol, int maxit, long long* total_iters);
/* Device targets of the reference's position loss (src/TrainHeter.h:211-226): lpos
 * max_steps x N x 3, epos max_steps x N x 2 (node-major, the state after step t in row
 * t - 1), host arrays copied to the device; the loss of step t is
 * weight (sum_k |x_k - lpos_k|^2 + sum_k |uv_k - epos_k|^2).  NULL, NULL clears them. */
int yc_adjoint_set_targets(yc_ctx* ctx, const double* lpos, const double* epos, double weight);
/* The backward pass over the recorded steps, with the same h (and the CG tol, maxit for
 * the backward solves).  gstate: the loss gradient with respect to the state after each
 * step, steps x 15N doubles in the yc_get_state layout (host), or NULL for the targets'
 * loss (or none).  Outputs (any may be NULL): dparam[YC_NPARAM] the parameter gradient;
 * dstate0[15N] the gradient with respect to the recorded initial state (yc_set_state
 * layout); dext[steps x n] the gradient with respect to each step's external force;
 * loss the targets' loss; cg_iters the backward solves' iterations.  The context ends at
 * the state after the last recorded step. */
int yc_adjoint_backward(yc_ctx* ctx, int mode, double h, double tol, int maxit, const double* gstate, double* dparam,
                        double* dstate0, double* dext, double* loss, long long* cg_iters);
/* Frees the recording. */
int yc_adjoint_end(yc_ctx* ctx);

/* ---------------------------------------------------------------------------------
 * Batches: many cloths of uneven size stepped together, each phase of a step one launch
 * over all of them (the persistent CG one launch whose blocks walk groups of cloths that
 * fit on the device at once), replayed from a hipGraph where the runtime can capture
 * the step.  No counterpart in the reference, which steps one Cloth per
 * PhysicsCloth (src/TrainHelper.cpp:15-97) and loops over cloths in its training
 * drivers (src/TrainHeter.h:81-146).  Each cloth's arithmetic is yc_run's with the
 * persistent CG, bit for bit.
 * ------------------------------------------------------------------------------- */
typedef struct yc_batch yc_batch;
typedef struct {
    int graph;        /* 1: replay each step from a hipGraph (direct launches if capture is unsupported) */
    int group_order;  /* 0: cloths laid out in the given order; 1: by decreasing size */
    int node_block;   /* threads per block of the per-node and per-segment kernels: 64, 128 or 256 */
} yc_batch_opts;
/* ctxs: n distinct contexts on one device; they stay owned by the caller and must
 * outlive the batch.  Fails if one cloth's solve alone cannot be resident at once.  The
 * members share one preconditioner (yc_set_preconditioner before creating the batch: all
 * YC_PRECOND_JACOBI or all YC_PRECOND_BLOCK2); the batch keeps the one they had. */
int yc_batch_create(yc_ctx* const* ctxs, int n, yc_batch** out);
void yc_batch_destroy(yc_batch* b);
int yc_batch_set_opts(yc_batch* b, const yc_batch_opts* opts);
int yc_batch_get_opts(const yc_batch* b, yc_batch_opts* opts);
/* `steps` SimSteps of every cloth (zero CG guess, tol/maxit as yc_step, maxit < 0 for
 * each cloth's 2n, the reference solver's default; each cloth's current external force).  iters[i] (may be NULL): CG iterations of cloth i summed
 * over the steps; status[i] (may be NULL): its worst status.  Returns the worst status
 * over all cloths; yc_last_error names the cloth. */
int yc_batch_run(yc_batch* b, double h, int steps, double tol, int maxit, long long* iters, int* status);
/* Layout facts: resident groups, whether the last run replayed a graph, tiles of all
 * cloths, tiles resident at once. */
int yc_batch_info(const yc_batch* b, int* n_groups, int* graph_used, int* total_tiles, int* capacity);
/* Reverse mode of a batch: the members' adjoint recordings (yc_adjoint_begin on each, with
 * checkpoint interval 1 and the same number of steps recorded) advanced together and
 * their backward passes run together, each phase one launch over all cloths and the
 * backward solves one launch; each cloth's gradient is its single-cloth
 * backward pass's.  No counterpart in the reference, whose training drivers differentiate
 * one cloth per pass (src/TrainHeter.h:161-231).
 * yc_batch_adjoint_run: `steps` recorded SimSteps of every member (as yc_batch_run); a
 * failed step voids every member's recording.
 * yc_batch_adjoint_backward: the members' target losses (yc_adjoint_set_targets on each)
 * differentiated with respect to their parameters: dparam n x YC_NPARAM and loss[n] (may
 * be NULL), iters[i] the backward solves' iterations, status[i] as yc_batch_run.  The
 * members end at their last recorded states. */
int yc_batch_adjoint_run(yc_batch* b, double h, int steps, double tol, int maxit, long long* iters, int* status);
int yc_batch_adjoint_backward(yc_batch* b, int mode, double h, double tol, int maxit, double* dparam, double* loss,
                              long long* iters, int* status);

#ifdef __cplusplus
}
#endif
#endif

/*
 * yarn_oracle.h -- C API of the CPU oracle for the implicit yarn-level step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain FP64 CPU restatement of
 * the reference's step (src/PhysicsCloth.cpp, src/Cloth.cpp, src/MatrixMp.cpp,
 * src/TrainHelper.cpp, src/Solvers.cpp).  It is the checker the HIP product is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Layouts follow the reference:
 *   node index        k = u + v*width                      (src/Cloth.cpp:201)
 *   crossing index    e = (u-1) + (v-1)*(width-2)           (src/Cloth.cpp:221)
 *   global vector     [L: 3k+{x,y,z}] ++ [E: 3N + 2e+{u,v}] (src/MatrixMp.cpp:120)
 *   state arrays      index u + v*length, as get/set_ClothState do (src/Cloth.cpp:362,426)
 */
#ifndef YARN_ORACLE_H
#define YARN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define YO_MAX_HANDLES 16

typedef struct {
    int width, length;
    double hight, L, R;
    int n_yarns;                 /* 1 (homogeneous) or 2 (warp row / weft column i uses yarn i%2) */
    double rho[2], Y[2], B[2];   /* per yarn: density, Young's modulus, bending modulus */
    double kc, kf, df, mu, S, handle_k;
    int pattern;                 /* 0 plain, 1 twill, 2 satin (src/Cloth.cpp:70-88) */
    int pose;                    /* 0 upright, 1 flat (src/Cloth.cpp:58-59) */
    double G[3];
    double wind_vel[3], wind_density, wind_drag;
    int n_handles;
    int handles[2 * YO_MAX_HANDLES];   /* (u, v) pairs */
} yo_params;

typedef struct yo_cloth yo_cloth;

yo_cloth* yo_create(const yo_params* p);
void yo_destroy(yo_cloth* c);

/* n = 3N + 2Nc, N = width*length, Nc = (width-2)*(length-2) */
int yo_ndof(const yo_cloth* c);

/* get/set_ClothState (src/Cloth.cpp:350-438).  Arrays hold width*length entries
 * of 3 (LPos, LPosFix, LVel) or 2 (EPos, EPosBar, EVel) doubles. */
void yo_get_state(const yo_cloth* c, double* lpos, double* lposfix, double* lvel,
                  double* epos, double* eposbar, double* evel);
void yo_set_state(yo_cloth* c, const double* lpos, const double* lposfix, const double* lvel,
                  const double* epos, const double* eposbar, const double* evel);

/* FillMV + system formation (src/PhysicsCloth.cpp:46-156, src/TrainHelper.cpp:19-20,61-63).
 * ext_force may be NULL (n doubles otherwise).  Returns 0, or 1 on NaN/Inf in A or b. */
int yo_assemble(yo_cloth* c, double h, const double* ext_force);

/* Checker option: the element terms yo_assemble evaluates (default all), so that a test
 * can pin one term's Jacobian against differences of that term's force alone.  With
 * YO_TERM_FRICTION off the compression is still computed (the shear reads it). */
#define YO_TERM_MASS 1u
#define YO_TERM_GRAVITY 2u
#define YO_TERM_STRETCH 4u
#define YO_TERM_INERTIA 8u
#define YO_TERM_MDOT 16u
#define YO_TERM_PARACOLLI 32u
#define YO_TERM_BEND 64u
#define YO_TERM_CRIMP 128u
#define YO_TERM_FRICTION 256u
#define YO_TERM_SHEAR 512u
#define YO_TERM_WIND 1024u
#define YO_TERM_HANDLE 2048u
#define YO_TERM_ALL 4095u
void yo_set_terms(yo_cloth* c, unsigned terms);
/* Checker option: assemble one element only -- edge, bend (with its crimp twin), shear
 * segment or face `index` of the reference's element lists (src/Cloth.cpp:167-305), or
 * the friction of crossing `index` -- so that a test can take differences of that
 * element's force alone.  YO_ELEM_ALL restores the full assembly; the handles are
 * evaluated only then. */
#define YO_ELEM_ALL 0
#define YO_ELEM_EDGE 1
#define YO_ELEM_BEND 2
#define YO_ELEM_SHEAR 3
#define YO_ELEM_FACE 4
#define YO_ELEM_CROSSING 5
void yo_set_element(yo_cloth* c, int kind, int index);
/* OpenMP threads of this cloth's assembly, normals and CG (default 1).  The assembly is
 * bit for bit the single-thread one (each thread owns grid rows and walks the element
 * lists in the reference's order); the CG's inner products with several threads are
 * ordered sums of 256 fixed chunks, the same for any thread count above one. */
void yo_set_threads(yo_cloth* c, int threads);
/* ConSlideFriction's arithmetic for one crossing component, the function the assembly
 * calls (src/PhysicsCloth.cpp:2619-2633).  in: uv - uvbar, compression, its EE input
 * 0.5 n.StrhBend_C_LE, the no-slide E force, the thread-partial E force, the no-slide EE
 * position and velocity Jacobian diagonals, the Eulerian velocity, mu, kf, df.
 * out: is_slide, friction force, EE position entry, EE velocity entry. */
void yo_friction_eval(const double* in /*11*/, double* out /*4*/);
/* The compression accumulators of the last assembly: StrhBend_C_L per node (N x 3) and the
 * self entries of StrhBend_C_LE per node and yarn direction (N x 2 x 3). */
void yo_get_compression_parts(const yo_cloth* c, double* sbc_l, double* sbc_le);

/* Results of the last yo_assemble. */
int yo_nnz(const yo_cloth* c);
/* Checker option for the reference's detached gradient: hold the non-jamming shear's
 * (phi - pi/2) (src/PhysicsCloth.cpp:2720-2721) at the current state's values (on = 1),
 * at given values, or release it (on = 0). */
void yo_freeze_shear_angle(yo_cloth* c, int on);
void yo_set_frozen_shear_angle(yo_cloth* c, const double* pb, int n);
int yo_get_frozen_shear_angle(const yo_cloth* c, double* pb);

/* which: 0 = A, 1 = M, 2 = K_x (F_*_Pos), 3 = K_v (F_*_Vel); CSR of the global n x n matrix */
void yo_get_csr(const yo_cloth* c, int which, int* rowptr, int* col, double* val);
void yo_get_dense(const yo_cloth* c, int which, double* dense /* n*n row-major */);
/* which: 0 = F (get_SpaF), 1 = b, 2 = v (get_SpaVel) */
void yo_get_vec(const yo_cloth* c, int which, double* out);
/* per crossing: is_slide (2), comp (1), thread-partial friction input F_E (2) */
void yo_get_crossing(const yo_cloth* c, double* is_slide, double* comp, double* fe_partial);
/* per node normal (3 doubles, zero on the boundary) */
void yo_get_normals(const yo_cloth* c, double* n);

/* Jacobi-preconditioned CG with the reference solver's stopping rule
 * (squared residual below tol^2*|b|^2, at most maxit iterations, x0 or zero guess). */
int yo_solve(yo_cloth* c, double tol, int maxit, const double* x0, double* x,
             int* iters, double* relres);

/* Cloth::update (src/Cloth.cpp:111-148) with the is_slide of the last assembly. */
void yo_update(yo_cloth* c, const double* new_vel, double h);

/* SimStep (src/TrainHelper.cpp:15-97): assemble, solve, update.  Returns 0 ok,
 * 1 NaN/Inf in A/b, 2 NaN/Inf in the solution. */
int yo_step(yo_cloth* c, double h, const double* ext_force, double tol, int maxit,
            int* iters, double* relres);

/* yo_step with the reference's dense storage work (per-thread N^2 block arrays, their
 * reduction, dense To_Sparse_Mat matrices and products, the n^2 scan into the solver's
 * matrix): a cost model of the reference's layout for small cloths; same result. */
int yo_step_dense(yo_cloth* c, double h, const double* ext_force, double tol, int maxit,
                  int* iters, double* relres);

/* Static-schedule owner thread used for the friction input (8 threads). */
int yo_owner(int n_iter, int i);

#ifdef __cplusplus
}
#endif
#endif

/* bte.h -- C ABI of the MI355X (gfx950) steady-state phonon BTE source iteration.
 *
 * This is the drop-in boundary below the reference's C++ API.  The reference has no C ABI: its
 * public surface is the C++ class StaticBTESolver (includes/StaticBTESolver/StaticBTESolver.h:81-95)
 * whose solve() runs _preprocess (StaticBTESolver.cpp:372-688) and then the outer source
 * iteration _iteration (StaticBTESolver.cpp:690-808) with a per-(band, direction) inner solve
 * (_solve_matrix :270-308 on the CPU, class BICGSTAB src/bicgstab.hpp:33-321 on the GPU).  The
 * entry points below replace that iteration as a whole: the host keeps the geometric and angular
 * preprocessing, hands the arrays over once, and bte_solve() runs the whole iteration resident on
 * one GPU.  The C++ facade (include/StaticBTESolver/StaticBTESolver.h) calls them.
 *
 * Conventions: plain C, no exceptions cross the boundary, every call returns BTE_OK (0) or a
 * negative status and leaves a message for bte_last_error().  Arrays are caller-owned host
 * memory and are copied on upload.  One context owns one HIP stream on one device; a context is
 * not thread-safe, distinct contexts are independent.
 */
#ifndef STATICBTE_AMD_BTE_H
#define STATICBTE_AMD_BTE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bte_ctx bte_ctx;

enum bte_status {
    BTE_OK = 0,
    BTE_ERR_ARG = -1,      /* bad argument or size */
    BTE_ERR_HIP = -2,      /* HIP runtime error */
    BTE_ERR_STATE = -3,    /* call out of order (e.g. solve before geometry) */
    BTE_ERR_BC = -4,       /* unsupported boundary condition for this mesh / quadrature */
    BTE_ERR_MESH = -5,     /* mesh inconsistent (e.g. more incoming faces than the ELL holds) */
    BTE_ERR_NOCONV = -6,   /* an inner solve hit its iteration limit (reported, not fatal) */
    BTE_ERR_NODEVICE = -7  /* no HIP device */
};

enum bte_precision {
    BTE_PREC_FP64 = 0,  /* everything in FP64 (the plain form) */
    BTE_PREC_MIXED = 1  /* FP32 inner BiCGSTAB + FP64 residual refinement, quadrature and T update */
};

enum bte_precond {
    BTE_PC_SWEEP = 0,  /* downwind (level-scheduled) sweep; exact when the upwind graph is acyclic */
    BTE_PC_JACOBI = 1  /* the reference GPU solver's right-Jacobi BiCGSTAB semantics (bicgstab.hpp) */
};

typedef struct bte_opts {
    int precision;        /* enum bte_precision, default BTE_PREC_FP64 */
    int precond;          /* enum bte_precond, default BTE_PC_SWEEP */
    int max_iter;         /* outer iterations, StaticBTESolver::solve(max_iter); default 10000 */
    double stop_margin;   /* stop when margin <= stop_margin, StaticBTESolver.cpp:798; default 1e-5 */
    double inner_tol;     /* relative residual of each inner solve, bicgstab.hpp:699 tol; 1e-9 */
    int inner_maxit;      /* inner iteration limit, StaticBTESolver.cpp:699; default 1000 */
    int refine_steps;     /* BTE_PREC_MIXED: max FP64 refinement steps per system; default 4 */
    int block;            /* work-group size per system (64..1024), 0 = autotuned; default 0 */
    int use_graph;        /* capture each outer iteration as a hipGraph; default 1 */
    int verbose;          /* print the reference's per-iteration stdout lines; default 0 */
    int timing;           /* per-kernel hipEvent timing into bte_stats; default 0 */
    int learned;          /* learned correction by the network of bte_attach_gnn: proposed after
                             state steps 1, 1 + learned_every, ... and kept only if the residual of
                             the iteration from the corrected state does not grow; stops proposing
                             after two rejections in a row.  Works with either accel; default 0 */
    int learned_every;    /* default 8; 0 = no proposals (keeps the state bookkeeping that
                             bte_learn_features reads, e.g. to record the plain iteration) */
    int direct_sweep;     /* BTE_PC_SWEEP + FP64: when no upwind edge had to be lagged (acyclic
                             upwind graph), solve each system by one downwind sweep with the RHS
                             formed on the fly instead of BiCGSTAB (same solution); default 1 */
    int sweep_variant;    /* direct sweep kernel: 0 autotuned, 1 basic, 2 level-ordered operands,
                             3 level-ordered operands with the intensity vector held in LDS
                             (needs n_cell * 8 bytes of LDS), 4 pipelined (operands loaded three
                             levels ahead, recent intensities in an LDS ring), 5 as 4 with one
                             wave per system (no work-group barriers), 6 / 7 as 4 with 2 / 4 bands
                             of a direction per work-group sharing each operator load, 8 as 4 with
                             a ring of 16384 values (128 KB of LDS: one work-group per compute unit,
                             fewer upwind values read back from HBM); a form the mesh cannot use
                             falls back to 2; default 0 */
    int accel;            /* outer iteration: BTE_ACCEL_NONE (the reference's source iteration) or
                             BTE_ACCEL_ANDERSON (Anderson mixing of the temperature field over the
                             last accel_memory iterates); default BTE_ACCEL_NONE */
    int accel_memory;     /* Anderson history length; default 20 */
    int band_group;       /* FP64 BiCGSTAB with BTE_PC_JACOBI: bands of one direction solved together by
                             one work-group, each operator row loaded once for all of them (each
                             band keeps its own scalars and stopping test): 1 (one system per
                             work-group), 2 or 4; 0 autotuned.  Other forms use 1; default 0 */
    double inner_tol32;   /* BTE_PREC_MIXED: relative residual each FP32 inner solve aims for before
                             the FP64 refinement step (near FP32 rounding it can stall); 1e-5 */
    int intensity_layout; /* device layout of the intensities: 0 automatic (3D pipelined sweeps with
                             4 or more bands interleave them), 1 band-major ee[b][d][c] (the
                             reference's), 2 direction-major with each cell's bands interleaved,
                             ee[d][c][b], when the pipelined sweep (sweep_variant 4 or 8) runs, else
                             band-major.  Results do not depend on it except the margins' summation
                             order; bte_get_intensity returns band-major either way; default 0 */
} bte_opts;

enum bte_accel { BTE_ACCEL_NONE = 0, BTE_ACCEL_ANDERSON = 1 };

typedef struct bte_result {
    /* caller-allocated outputs, any may be NULL */
    double* T;             /* [n_cell] cell temperature (cell_temperature) */
    double* T_band;        /* [n_cell * n_band] cell-major (cell_band_temperature[c][b]) */
    double* bc_flux;       /* [n_bc] heat flux per BC (bc_heat_flux) */
    double* bc_band_flux;  /* [n_bc * n_band] (bc_band_heat_flux[k][b]) */
    double* margins;       /* [margins_capacity] margin of each outer iteration */
    int margins_capacity;
    /* filled by bte_solve */
    int iterations;           /* outer iterations run */
    int converged;            /* 1 if the last margin <= stop_margin */
    long long inner_iters;    /* total inner BiCGSTAB iterations over all systems */
    int inner_failures;       /* systems that hit inner_maxit (summed over outer iterations); in
                                 BTE_PREC_MIXED, systems whose FP64 re-solve (after an FP32 inner
                                 solve stalled) hit it */
    double solve_seconds;     /* wall time of the outer loop (like StaticBTESolver.cpp:865-876) */
    double setup_seconds;     /* assembly + sweep schedule + uploads */
    int learned_accepted;     /* learned corrections accepted / proposed */
    int learned_proposed;
} bte_result;

typedef struct bte_stats {
    int n_kernels;
    const char* names[16];
    double ms[16];        /* summed hipEvent time per kernel kind (opts.timing = 1) */
    long long calls[16];
    double spmv_bytes;    /* algorithmic bytes of all SpMV row sweeps in the last solve */
    long long spmv_rows;  /* rows processed by SpMVs in the last solve */
    int ell_width;        /* off-diagonal slots per row */
    int max_levels;       /* deepest sweep schedule over directions */
    int lagged_edges;     /* upwind edges dropped from the sweep to break cycles */
    int block;            /* work-group size used */
    int variant;          /* direct-sweep variant used (1..8), 0 for the Krylov paths */
    int band_group;       /* Krylov paths: bands per work-group used (opts.band_group); 0 direct */
    int intensity_layout; /* intensity layout of the last solve: 1 band-major, 2 bands interleaved */
} bte_stats;

void bte_default_opts(bte_opts* opts);

/* Contexts.  device < 0 uses the current HIP device. */
int bte_ctx_create(int device, bte_ctx** out);
void bte_ctx_destroy(bte_ctx* ctx);
const char* bte_last_error(const bte_ctx* ctx);

/* Geometry from _preprocess (StaticBTESolver.cpp:494-687): per cell volume; per cell face f
 * (f opposite vertex f) area, outward unit normal (x,y,z) and neighbour (>= 0 a cell, < 0 the
 * boundary entity k encoded as -k-1).  dim in {1,2,3}, n_face = dim+1. */
int bte_set_geometry(bte_ctx* ctx, int dim, int n_cell, int n_face, const double* volume,
                     const double* face_area, const double* face_normal, const int* neighbor);

/* Angular quadrature from _preprocess (StaticBTESolver.cpp:374-492): control angles omega[n_dir],
 * unit directions dir[n_dir*3], weighted directions S[n_dir*3]; num_phi is needed for the
 * specular mirror maps (:205-252); solid_angle and T_ref as set by setParam (:48-94). */
int bte_set_quadrature(bte_ctx* ctx, int DM, int n_dir, const double* omega, const double* dir,
                       const double* S, int num_theta, int num_phi, double solid_angle, double T_ref);

/* Bands (BTEBand.cpp:8-26): group velocity, relaxation time, Ctot, Lr. */
int bte_set_bands(bte_ctx* ctx, int n_band, const double* v, const double* tau, const double* C,
                  const double* Lr);

/* Boundary conditions (BTEBoundaryCondition.cpp:7-23): index (= -k-1), type in {1,2,31,32,33},
 * temperature. */
int bte_set_bcs(bte_ctx* ctx, int n_bc, const int* index, const int* type, const double* T);

/* Runs the source iteration (StaticBTESolver::_iteration + _get_heat_flux).  Returns BTE_OK, or
 * BTE_ERR_NOCONV when some inner solve hit inner_maxit (results are still filled in). */
int bte_solve(bte_ctx* ctx, const bte_opts* opts, bte_result* res);

/* bte_solve without device memory allocation.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The reductions (margin partial sums, squared norms, Anderson dot products, heat-flux staging) use
 * a 1 MiB array per device in the library's code object, with room for 56,992 (BC, band) pairs
 * (more are rejected with BTE_ERR_ARG); every other device buffer must already exist from
 * an earlier bte_solve of the same form with inputs at least as large, else BTE_ERR_STATE.  No
 * graphs are captured or replayed and the launch shape is not autotuned here.  All contexts on one
 * device share that device's array: calls that overlap on one device race.  Results equal
 * bte_solve's with use_graph = 0 and the same launch shape. */
int bte_solve_noalloc(bte_ctx* ctx, const bte_opts* opts, bte_result* res);

/* The same iteration in steps (bte_solve = bte_begin + bte_step(max_iter) + bte_finish), for
 * callers that inspect or correct the temperature field between outer iterations.
 * bte_begin: prepares and resets the state (e = 0, density = 0 as StaticBTESolver.cpp:702-705).
 * bte_step: runs up to n_iter outer iterations, stopping early once margin <= stop_margin.
 * bte_get_temperature: T[n_cell] and T_band[n_cell*n_band] (cell-major) of the last iteration.
 * bte_set_temperature: replaces T and the emission density C_b/Omega (T - T_ref) used by the next
 *   iteration (the boundary intensities of the last iteration are kept).
 * bte_finish: heat flux and outputs of the last iteration, as bte_solve. */
int bte_begin(bte_ctx* ctx, const bte_opts* opts);
int bte_step(bte_ctx* ctx, int n_iter, int* n_done, int* converged);
int bte_get_temperature(bte_ctx* ctx, double* T, double* T_band);
int bte_set_temperature(bte_ctx* ctx, const double* T);
int bte_finish(bte_ctx* ctx, bte_result* res);

/* Final intensities e[b][d][c] ([n_band*n_dir*n_cell], ee_curr of the last iteration). */
int bte_get_intensity(bte_ctx* ctx, double* ee);

int bte_get_stats(bte_ctx* ctx, bte_stats* stats);

/* ---- learned correction: mesh graph network (no counterpart in the reference) ----
 * Nodes are cells, edges interior faces.  16 input features per node (zero padded), hidden width
 * 64, `layers` message-passing layers with residual connections, linear read-out; BF16 MFMA
 * products with FP32 accumulation.  Parameters are an FP32 vector of bte_gnn_num_params() values
 * (layout restated in oracle/gnn_ref.py).  Training: bte_gnn_backward (forward + weighted squared
 * error + gradients) then bte_gnn_adam. */
typedef struct bte_gnn bte_gnn;
int bte_gnn_create(int device, int layers, unsigned long long seed, bte_gnn** out);
void bte_gnn_destroy(bte_gnn* g);
const char* bte_gnn_last_error(const bte_gnn* g);
int bte_gnn_num_params(const bte_gnn* g);
int bte_gnn_get_params(bte_gnn* g, float* params);
int bte_gnn_set_params(bte_gnn* g, const float* params);  /* also resets the Adam state */
int bte_gnn_get_grads(bte_gnn* g, float* grads);
/* nbr: [n_node * n_face] face table as in bte_set_geometry (negative = boundary) */
int bte_gnn_set_graph(bte_gnn* g, int n_node, int n_face, const int* nbr);
int bte_gnn_forward(bte_gnn* g, const float* x /*[n_node*16]*/, float* y /*[n_node], may be NULL*/);
int bte_gnn_backward(bte_gnn* g, const float* x, const float* target, const float* weight /*may be NULL*/,
                     double* loss);
int bte_gnn_adam(bte_gnn* g, float lr, float beta1, float beta2, float eps, float weight_decay);

/* Attaches the network used by bte_opts.learned (NULL detaches).  Not owned: it must outlive every
 * learned solve of the context and live on the same device.  bte_begin points its graph at the
 * context's mesh whenever that changed (which replaces any graph set by bte_gnn_set_graph). */
int bte_attach_gnn(bte_ctx* ctx, bte_gnn* g);

/* Network inputs of the current iteration state, for building training pairs: X [n_cell * 16]
 * (node-major), x_next [n_cell] the temperature part of the state the next iteration would start
 * from (may be NULL), span the wall-temperature span that scales the network output (correction =
 * span * y; may be NULL).  Valid after a bte_step of an iteration begun with accel or learned that
 * ran at least one iteration; the target for y is (T_converged - x_next) / span. */
int bte_learn_features(bte_ctx* ctx, float* X, double* x_next, double* span);

/* ---- uneven batch of problems (no counterpart in the reference: it runs one problem per process) ----
 * n contexts on one device, each fully loaded (geometry, quadrature, bands, BCs), solved together:
 * every outer iteration sweeps the systems of all still-iterating problems in ONE launch (work-groups
 * take the systems with the longest level chains first), then each problem runs its own outer-step
 * bookkeeping (stop test, Anderson mixing, learned corrections) and leaves the batch when it
 * converges.  Every problem must use the FP64 direct-sweep form.  The contexts must outlive the
 * batch and must not be used concurrently with it; results equal those of solving each alone. */
typedef struct bte_batch bte_batch;
int bte_batch_create(int n, bte_ctx* const* ctxs, bte_batch** out);
void bte_batch_destroy(bte_batch* b);
const char* bte_batch_last_error(const bte_batch* b);
/* results: [n] with caller-set output arrays as for bte_solve (may be NULL); each solve_seconds is the
 * wall time of the whole batch. */
int bte_batch_solve(bte_batch* b, const bte_opts* opts, bte_result* results);
/* Timing of the last bte_batch_solve run with opts.timing = 1 (else zeros): "rhs" (the diffuse and
 * right-hand-side launches of all problems) and "solve" (the batched sweep launch), summed hipEvent
 * milliseconds and launch counts, as bte_get_stats reports them for one context. */
int bte_batch_get_stats(const bte_batch* b, bte_stats* st);

/* ---- piecewise entry points (used by the parity tests; same kernels as bte_solve) ---- */

/* Assembly of system (band, dir), StaticBTESolver.cpp:122-160, in dense face-slot form:
 * diag[n_cell] and off[n_cell*n_face] (coefficient of neighbour[c*F+f], 0 where the face is
 * outgoing or on the boundary). */
int bte_debug_assemble(bte_ctx* ctx, int band, int dir, double* diag, double* off);

/* Right-hand sides of all systems, StaticBTESolver.cpp:162-269, for given ee_prev [B*D*N] and
 * density [N*B] (cell-major); out rhs [B*D*N]. */
int bte_debug_rhs(bte_ctx* ctx, const double* ee_prev, const double* density, double* rhs);

/* One batched inner solve of all systems for given ee_prev and density (as bte_debug_rhs);
 * out x [B*D*N] and per-system iteration counts iters [B*D] (may be NULL). */
int bte_debug_inner(bte_ctx* ctx, const bte_opts* opts, const double* ee_prev, const double* density,
                    double* x, int* iters);

#ifdef __cplusplus
}
#endif
#endif

/*
 * giftbte.h - C ABI of libgiftbte_hip.so, the MI355X solver for GiftBTE's
 * steady StaticBTESynthetic path (Order 2, 2limit or 1, IterativeScheme
 * Synthetic or Sequential, MatrixSolver LU or BICGSTAB) and its StaticFourier
 * warm start.
 *
 * The reference exposes this path as C++ classes with no FFI:
 *   StaticFourier(BTEMesh*, BTEBoundaryCondition*, BTEBand*, int, int) + _solve()
 *       includes/StaticFourier/StaticFourier.h:91-102
 *   StaticBTESynthetic(BTEMesh*, BTEBoundaryCondition*, BTEBand*, BTEAngle*, int, int)
 *       + solve(int Use_Backup, int max_iter, int USE_limiter, double tolT, double tolQ)
 *       includes/StaticBTESynthetic/StaticBTESynthetic.h:177,221
 *   driven by SolutionAll::_Fourier_Solver / _BTE_Solver
 *       src/Solution/Solution.cpp:236-414
 * The entry points below replace those calls one for one.  Every function
 * returns GBTE_OK or an error code; gbte_last_error() gives the message.
 * Nothing here calls exit(): the reference's exit() paths
 * (e.g. src/BTEMesh/BTEMesh.cpp:868-879, StaticBTESynthetic.cpp:2719-2723)
 * become status codes.  One handle per host thread; no MPI.
 */
#ifndef GIFTBTE_H
#define GIFTBTE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GBTE_OK = 0,
    GBTE_BAD_INPUT = 1,   /* deck / mesh / argument error                        */
    GBTE_DIVERGED = 2,    /* error grew 10 times in a row                       */
    GBTE_NAN = 3,         /* NaN / Inf in a residual                            */
    GBTE_HIP_ERROR = 4,   /* HIP runtime failure, or no GPU                     */
    GBTE_UNSUPPORTED = 5  /* a reference feature outside this build             */
};

/* The deck's Order, IterativeScheme and MatrixSolver (CONTROL), dispatched
 * by SolutionAll::_BTE_Solver (src/Solution/Solution.cpp:287-406) to
 * solve / solve_Iterative (StaticBTESyntheticSolver.cpp:12-486, Order 2 and
 * 2limit, Synthetic), solve_DOM / solve_DOM_Iterative (:968-1392, Sequential)
 * and the first-order variants (:487-966, :1394-1802).  gbte_solve runs
 * all of them (2limit: the limiter of StaticBTESynthetic.cpp:1372-1422; not
 * with GBTE_FP32_SWEEP, whose float rounding moves the limiter's switch).
 * BICGSTAB differs from LU only in the per-system linear solve (Eigen BiCGSTAB
 * with an incomplete-LU preconditioner at its default tolerance, machine
 * epsilon), so both run the exact solves here. */
enum { GBTE_ORDER_1 = 1, GBTE_ORDER_2 = 2, GBTE_ORDER_2LIMIT = 3 };
enum { GBTE_SCHEME_SYNTHETIC = 0, GBTE_SCHEME_SEQUENTIAL = 1 };
enum { GBTE_MATRIX_LU = 0, GBTE_MATRIX_BICGSTAB = 1 };

/* Storage of the per-(band, direction) fields: energy densities, boundary
 * values, LSQ gradients and sweep sources.  GBTE_FP32_SWEEP stores them in
 * float (half the HBM traffic of the transport kernels).  The LSQ gradient
 * product runs in float on the matrix cores; every other operation, the
 * temperature / flux recovery, the synthetic solve and the residual and
 * k_eff reductions stay in double.  The energy
 * densities are held as deviations from the equilibrium at the deck's Tref,
 * so float keeps about 7 significant digits of the deviation itself.
 * GBTE_FP32_SWEEP needs the tiled face kernel and the wave-specialised sweep
 * (gbte_create returns GBTE_BAD_INPUT with fallback_kernels or on a deck
 * those kernels do not fit).  The reference computes in double throughout. */
enum { GBTE_FP64 = 0, GBTE_FP32_SWEEP = 1 };

/* Linear solver of the synthetic (macro) and Fourier systems, which the
 * reference solves with Eigen::SparseLU (StaticBTESynthetic.cpp:2692-2734,
 * StaticFourier.cpp:934-976): an exact nested-dissection Cholesky factor
 * (factored once at gbte_create, triangular solves on the GPU), or
 * Jacobi-preconditioned CG to macro_rtol. */
enum { GBTE_LINSOLVE_DIRECT = 0, GBTE_LINSOLVE_PCG = 1 };

/* Read-only view of the flattened problem tables, laid out exactly as the
 * reference's StaticBTESynthetic constructor flattens BTEMesh/BTEBand/BTEAngle
 * (src/StaticBTESynthetic/StaticBTESynthetic.cpp:224-626): 6 face slots per
 * cell, 4 vertices per face slot, 8 vertices per cell, boundaries with two
 * (cell, face) sides, unique faces with two sides.  Owned by the problem. */
typedef struct gbte_tables {
    int dim;              /* geometry dimension (2 or 3)                       */
    int n_cells, n_nodes, n_bounds, n_faces, n_bands, n_dirs, n_matters;
    int material_dim;     /* MaterialDimension (2 or 3)                        */
    double Lx, Ly, Lz;
    /* cells */
    const int *face_size;          /* [n_cells]                   */
    const int *face_neighbor;      /* [n_cells*6]  -1 on boundary */
    const int *face_bound;         /* [n_cells*6]  -1 if interior */
    const int *face_index;         /* [n_cells*6]  unique face id */
    const double *face_area;       /* [n_cells*6]                 */
    const double *face_nx, *face_ny, *face_nz;   /* [n_cells*6]   */
    const double *face_cx, *face_cy, *face_cz;   /* [n_cells*6]   */
    const int *face_nvert;         /* [n_cells*6]                 */
    const int *face_vert;          /* [n_cells*6*4]               */
    const int *cell_nvert;         /* [n_cells]                   */
    const int *cell_vert;          /* [n_cells*8]                 */
    const double *volume, *cx, *cy, *cz, *heat_source;   /* [n_cells] */
    const int *matter;             /* [n_cells]                   */
    const int *nbr_ptr, *nbr_idx;  /* same-matter vertex neighbours, CSR */
    const double *node_x, *node_y, *node_z;              /* [n_nodes] */
    /* boundaries */
    const int *bound_cell, *bound_face;    /* [n_bounds*2], -1 if absent */
    const int *bound_type;                 /* [n_bounds]                */
    const double *bound_temp;              /* [n_bounds] deck value     */
    const int *bound_connect;              /* [n_bounds] periodic pair  */
    /* unique faces */
    const int *fface_cell, *fface_face;    /* [n_faces*2]               */
    const int *fface_bound;                /* [n_faces]                 */
    /* materials, bands, directions */
    const double *kappa_bulk, *capacity_bulk;           /* [n_matters]          */
    const double *vg, *tau, *cap, *lr, *hr;             /* [n_matters*n_bands] */
    const double *dir_x, *dir_y, *dir_z;                /* [n_dirs]             */
    const double *weight;          /* [n_dirs] angle weight / solid angle */
    const double *initial_temp;    /* [n_cells] Tref unless a file sets it */
    /* CONTROL values from the deck */
    int max_iter;
    double tol_temp, tol_flux, tref;
    int order, scheme, matrix_solver;      /* GBTE_ORDER_*, GBTE_SCHEME_*, GBTE_MATRIX_* */
} gbte_tables;

typedef struct gbte_problem gbte_problem;
typedef struct gbte_handle gbte_handle;

typedef struct gbte_options {
    int precision;        /* GBTE_FP64 or GBTE_FP32_SWEEP                       */
    int device;           /* HIP device ordinal                                 */
    int verbose;          /* print the reference's per-iteration log lines      */
    double macro_rtol;    /* relative residual of the PCG linear solver (0: default) */
    int linear_solver;    /* GBTE_LINSOLVE_DIRECT (default) or GBTE_LINSOLVE_PCG */
    int fallback_kernels; /* 1: use the kernels meant for decks the fast ones do not
                             fit (level-synchronous sweep, untiled face kernels);
                             same results, for testing those paths */
    int sweep_kernel;     /* tetrahedral meshes: GBTE_SWEEP_AUTO (default: the set-up
                             tuner times every sweep kernel that takes the mesh and
                             keeps the fastest) or one kernel family:
                             GBTE_SWEEP_WAVESPEC, GBTE_SWEEP_LEVEL (FP64 only),
                             GBTE_SWEEP_CHUNKED (a lane takes a position's band
                             chunk), GBTE_SWEEP_LANES (several lanes per position,
                             one 16-byte vector each); same results to rounding */
    int gradient_kernel;  /* tetrahedral meshes: GBTE_GRAD_AUTO (default: the set-up tuner
                             keeps the fastest) or GBTE_GRAD_GATHER (stencil gathers on the
                             vector units), GBTE_GRAD_MATRIX (tiles of four cells on the
                             matrix cores) or GBTE_GRAD_VERTEX (sums over the mesh
                             vertices, then the remaining stencil entries; tetrahedra of
                             one material); same results to rounding */
    int source_layout;    /* tetrahedral meshes with the chunked sweep: the layout of the
                             sweep source between the source kernel and the sweep,
                             GBTE_SLAYOUT_AUTO (default: the tuner times the pair in each)
                             or GBTE_SLAYOUT_CHUNK ([dir][band chunk][position][chunk]),
                             GBTE_SLAYOUT_POSITION ([dir][position][band]) or
                             GBTE_SLAYOUT_CELL ([cell][dir][band], e's layout); same
                             results */
} gbte_options;

enum { GBTE_SWEEP_AUTO = 0, GBTE_SWEEP_WAVESPEC = 1, GBTE_SWEEP_LEVEL = 2, GBTE_SWEEP_CHUNKED = 3, GBTE_SWEEP_LANES = 4 };
enum { GBTE_GRAD_AUTO = 0, GBTE_GRAD_GATHER = 1, GBTE_GRAD_MATRIX = 2, GBTE_GRAD_VERTEX = 3 };
enum { GBTE_SLAYOUT_AUTO = 0, GBTE_SLAYOUT_CHUNK = 1, GBTE_SLAYOUT_POSITION = 2, GBTE_SLAYOUT_CELL = 3 };

typedef struct gbte_stats {
    int iterations;       /* outer iterations run                               */
    int macro_passes;     /* synthetic (macro) solves, summed                   */
    int fourier_passes;   /* Fourier warm-start passes                          */
    int converged;        /* 1 if tolT and tolQ were met                        */
    double err_temp, err_flux;   /* last errors (StaticBTESynthetic.cpp:2565-2611) */
    double solve_ms;      /* wall time of the outer loop                        */
    double sweep_ms;      /* device time of the upwind sweep kernel (HIP events), summed */
    double transport_ms;  /* device time from the source kernel to the end of the face/moment kernel */
    double macro_ms;      /* device time of the synthetic (macro) loops, host waits included */
    int sweep_levels;     /* upwind levels of the deepest direction                  */
    int macro_cg_iters;   /* PCG iterations, summed over synthetic solves (PCG only) */
    double source_ms;     /* device time of the source kernel, summed                */
    double face_ms;       /* device time of the vertex/face/moment kernel(s), summed */
    double grad_ms;       /* of which the LSQ gradient kernel on tetrahedral meshes (0 in 2D,
                             where the face kernel forms the gradient) */
} gbte_stats;

/* Statistics of gbte_adjoint. */
typedef struct gbte_adjoint_stats {
    int iterations;   /* adjoint iterations run                                   */
    int converged;    /* 1 if the relative change of the T adjoint fell to tol    */
    double change;    /* last relative change                                     */
    double solve_ms;  /* wall time of the call                                    */
} gbte_adjoint_stats;

/* ---- problem construction (replaces BTEMain.cpp:87-362: decks, BTEBand,
 *      BTEBoundaryCondition, DistributeMesh, BTEAngle) ---------------------- */

/* Read input/CONTROL, input/GEOMETRY, input/PHONON and the files they name,
 * relative to deck_dir, and build the tables.  Replaces the deck scan in
 * src/BTEMain.cpp:87-315 and the object set-up in :327-362. */
int gbte_problem_load(const char *deck_dir, gbte_problem **out);
/* Pointer view of the problem's tables (valid until gbte_problem_free). */
int gbte_problem_tables(const gbte_problem *p, gbte_tables *out);
void gbte_problem_free(gbte_problem *p);

/* ---- solver (replaces StaticFourier / StaticBTESynthetic) --------------- */

/* Allocate device state for a problem.  Replaces the two solver
 * constructors (StaticFourier.cpp:82-268, StaticBTESynthetic.cpp:224-626). */
int gbte_create(const gbte_problem *p, const gbte_options *opts, gbte_handle **out);
/* Fourier warm start, StaticFourier::_solve (StaticFourier.cpp:934-976),
 * from T = initial temperature, as SolutionAll::_Fourier_Solver
 * (Solution.cpp:236-260).  Leaves T and q in the handle. */
int gbte_fourier_solve(gbte_handle *h, gbte_stats *stats);
/* Overwrite the starting T and q (Solution.cpp:268-273).  Any pointer may be
 * NULL to keep the current field. */
int gbte_set_initial(gbte_handle *h, const double *T, const double *qx, const double *qy, const double *qz);
/* StaticBTESynthetic::solve(0, max_iter, 0, tolT, tolQ)
 * (StaticBTESyntheticSolver.cpp:12-247), or solve_DOM for the Sequential
 * scheme (:968-1178: the same transport iteration without the synthetic
 * correction, T the recovered lattice temperature). */
int gbte_solve(gbte_handle *h, int max_iter, double tol_temp, double tol_flux, gbte_stats *stats);
/* Many independent problems at once: gbte_solve on each of the n distinct
 * handles, concurrently (one host thread per handle; each handle keeps its
 * own stream, graphs and buffers, so the device runs the solves side by
 * side).  Per-handle arguments, stats and status (any may be NULL except the
 * arguments); returns the first failing status or GBTE_OK.  No reference
 * counterpart: the reference solves one deck per process. */
int gbte_solve_batch(gbte_handle *const *hs, int n, const int *max_iter, const double *tol_temp,
                     const double *tol_flux, gbte_stats *stats, int *status);
/* Output fields (caller-owned, n_cells doubles each; T_band n_bands*n_cells;
 * NULL skips a field): macro (lattice) temperature, energy-equivalent
 * temperature (what Solution.cpp:408-413 copies back), heat flux, band
 * temperatures.  StaticBTESynthetic.cpp:2613-2653. */
int gbte_get_fields(gbte_handle *h, double *T_lattice, double *T_energy, double *qx, double *qy, double *qz,
                    double *T_band);
/* The reference's output files in directory dir (NULL: the working
 * directory): TempLattice.dat, Tempband.dat, Temperature.dat, HeatFlux.dat
 * and Result.vtk, as StaticBTESynthetic::_print_out
 * (StaticBTESynthetic.cpp:2613-2653) and SolutionAll::_print_out
 * (Solution.cpp:128-233) leave them after a steady run, from the handle's
 * fields.  gbte_write_fields writes the same files from caller fields in deck
 * order (n_cells each, T_band n_bands * n_cells) and needs no device. */
int gbte_write_outputs(gbte_handle *h, const gbte_problem *p, const char *dir);
int gbte_write_fields(const gbte_problem *p, const double *T_lattice, const double *T_energy, const double *qx,
                      const double *qy, const double *qz, const double *T_band, const char *dir);
/* Effective conductivity sum_i q_axis,i V_i / (A_perp dT), with A_perp the
 * domain extent normal to the axis (L_y for axis 0 in 2D).  No reference
 * function; the MATLAB check bin/examples/cross-plane/verify_kappa_crossplane.m:8
 * computes mean(q) L / dT, which equals this on full equal-area domains. */
int gbte_keff(gbte_handle *h, int axis, double delta_T, double *keff);
/* What the handle's set-up timed and chose among kernel variants (one line
 * per candidate with its milliseconds per launch, then the choice; empty
 * where nothing was tuned).  Valid until gbte_destroy.  No reference
 * counterpart: the reference has one implementation of each loop. */
const char *gbte_tuning_log(const gbte_handle *h);
/* k_eff from a flux field already on the device: *keff_dev =
 * sum_i q_dev[i] vol_dev[i] / (area delta_T) over n_cells cells (device
 * pointers; keff_dev one double), ordered on `stream` (a hipStream_t; NULL:
 * the null stream), in a fixed reduction order.  The partial sums live in a
 * buffer allocated and freed inside the call, which therefore waits for the
 * stream.  n_cells <= 16,777,216.  Errors: gbte_problem_error().  The same
 * quantity as gbte_keff (A_perp and delta_T supplied by the caller). */
int gbte_flux_keff(const double *q_dev, const double *vol_dev, int n_cells, double area, double delta_T,
                   double *keff_dev, void *stream);
/* gbte_flux_keff that allocates nothing: the partial sums go to a fixed
 * 32 KiB device array of the library (one per GPU, held until the process
 * ends), so the call only enqueues two kernels on `stream` and returns; it
 * can be captured into a HIP graph.  Same limit (n_cells <= 16,777,216,
 * larger is GBTE_BAD_INPUT) and bit-identical results.  Calls that overlap on
 * one GPU share the array and race: order them on one stream.  No reference
 * counterpart: for callers that cannot allocate device memory (graph
 * capture, serving loops). */
int gbte_flux_keff_noalloc(const double *q_dev, const double *vol_dev, int n_cells, double area, double delta_T,
                           double *keff_dev, void *stream);
/* Sensitivities of k_eff (the gbte_keff quantity along `axis` with the
 * nominal delta_T) at the converged state of the last gbte_solve, by the
 * discrete adjoint of the outer iteration (StaticBTESyntheticSolver.cpp:
 * 63-221 with its synthetic passes, or solve_DOM for Sequential): a
 * fixed-point iteration of the transposed map, its transport part a sweep of
 * the transposed upwind matrix in reversed upwind order on the GPU, to a
 * relative change of tol or max_iter iterations.
 *   dk_dtau_scale[n_bands]: d k_eff / d s_b for tau_b -> s_b tau_b, the lattice
 *     ratios following as lr = (C/tau) / sum (C/tau) and kappa_bulk = sum
 *     C v^2 tau / MaterialDimension, at s = 1;
 *   dk_dbound_temp[n_bounds] (may be NULL): d k_eff / d T of each isothermal
 *     (type 1) boundary face; 0 for the others.
 * Covers 2D meshes of one material in FP64 storage, Order 2 or 1, both
 * schemes, boundary types 1, 2, 3 and periodic; else GBTE_UNSUPPORTED.
 * No reference counterpart. */
int gbte_adjoint(gbte_handle *h, int axis, double delta_T, int max_iter, double tol, double *dk_dtau_scale,
                 double *dk_dbound_temp, gbte_adjoint_stats *stats);
const char *gbte_last_error(const gbte_handle *h);
/* The calling thread's last error message of a call without a handle. */
const char *gbte_problem_error(void);
void gbte_destroy(gbte_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* GIFTBTE_H */

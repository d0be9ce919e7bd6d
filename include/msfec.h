/*
 * msfec.h -- C-ABI of the MI355X multiscale basis library (libmsfec_hip.so).
 *
 * The library replaces, for a whole batch of coarse cells at once, the
 * per-cell object Q::QBasis of MsFEC and the parts of Q::QMultiscale that
 * consume it (msfec_q_*, msfec_coarse_solve), and the two-dimensional forms
 * of QNed::QNedBasis (msfec_ned_*) and RTDQ::RTDQBasis (msfec_rt_*).  Every
 * entry point below cites the reference member it stands in for (paths
 * relative to the MsFEC repository).  Plain C types only: no
 * C++ exceptions cross this boundary, pointers are host or device memory as
 * the accompanying *_loc argument says, streams are hipStream_t passed as
 * void*.  All GPU work of one handle runs on the stream given to its
 * compute call (0 = the default stream).  The set_* calls copy their input
 * on the stream of the handle's last compute call (the default stream
 * before the first) and return once the copy is done, so a compute or graph
 * replay still in flight there finishes with the old input; device-resident
 * inputs must be ready on that stream.  A handle is not thread-safe;
 * distinct handles may be used from distinct threads.
 *
 * Conventions shared with the reference (deal.II):
 *   - coarse vertex order: bit d of the vertex index selects the low/high side
 *     in direction d (GeometryInfo<dim> / FE_Q<dim>(1) local DoF order);
 *   - the coarse element matrix M is nb x nb row-major with row = test basis,
 *     nb = 2^dim, exactly like get_global_element_matrix();
 *   - fine nodes and fine cells of one coarse cell are numbered
 *     lexicographically, x fastest (deal.II numbers fine DoFs by cell traversal;
 *     values agree node by node at equal coordinates).
 */
#ifndef MSFEC_H
#define MSFEC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int msfec_status;
enum {
  MSFEC_OK = 0,
  MSFEC_ERR_INVALID = 1,     /* bad argument or unsupported combination    */
  MSFEC_ERR_HIP = 2,         /* a HIP runtime call failed                  */
  MSFEC_ERR_STATE = 3,       /* call out of order (e.g. get before compute) */
  MSFEC_ERR_UNSUPPORTED = 4, /* valid request this build does not handle   */
  MSFEC_ERR_CELL = 5         /* at least one cell failed; see cell status  */
};

/* where a pointer argument lives */
enum { MSFEC_HOST = 0, MSFEC_DEVICE = 1 };

/* diffusion coefficient layouts (per coarse cell, fine cells lexicographic):
 *   SCALAR_CELL: N^dim doubles, A = a I constant on each fine cell
 *   TENSOR_QP  : N^dim x 2^dim x ncomp doubles, symmetric tensor at each
 *                QGauss(2) point (points lexicographic, x fastest), ncomp = 3
 *                (xx,xy,yy) in 2-D and 6 (xx,xy,xz,yy,yz,zz) in 3-D.  This is
 *                what Diffusion_A::value_list produces at the quadrature points
 *                (source/equation_data/eqn_coeff_A.cc:163-191). */
enum { MSFEC_COEF_SCALAR_CELL = 0, MSFEC_COEF_TENSOR_QP = 1 };
/* load f: constant, one value per fine cell, or per fine cell per Gauss point */
enum { MSFEC_RHS_CONST = 0, MSFEC_RHS_CELL = 1, MSFEC_RHS_QP = 2 };
/* optional reaction R u (the reference's QBasis has none: 0 reproduces it) */
enum { MSFEC_REACTION_NONE = 0, MSFEC_REACTION_CONST = 1, MSFEC_REACTION_CELL = 2 };
/* local solver: banded Cholesky of the interior system, the analogue of
 * use_direct_solver_basis = true (source/Q/q_basis.cc:304-330):
 *   BAND_FP64      FP64, one pivot per step on the vector ALUs (the default
 *                  of the Python and C++ front ends);
 *   BAND_FP64_MFMA FP64, four pivots per step with FP64 matrix-core updates;
 *   BAND_MIXED     FP32 factor on the matrix cores and iterative refinement
 *                  with the residual in FP64: a cell stops refining once the
 *                  predicted error (|dx| / |x|)^2 is below 1e-11, after at
 *                  most max_refine steps (scalar coefficient per fine cell and
 *                  constant or per-cell load only);
 *   CG             matrix-free conjugate gradients in FP64 with the Jacobi
 *                  preconditioner, the analogue of use_direct_solver_basis =
 *                  false (solve_iterative, q_basis.cc:335-357, which uses SSOR):
 *                  one workgroup per (cell, basis), no bandwidth limit (3-D up
 *                  to n_refine_local = 4, 2-D up to 6), starts from the coarse
 *                  Q1 function of the basis (scalar coefficient per fine
 *                  cell: the solution when the coefficient is constant; zero
 *                  otherwise), stops at
 *                  ||r|| <= cg_tol ||b|| (default 1e-13) or after cg_max_iter
 *                  iterations (default 5000); the element matrix is second
 *                  order in the solve's error (the energy form of the Schur
 *                  term for a scalar coefficient per fine cell; Psi^T K Psi
 *                  with the assembled matrix for tensors at the Gauss points
 *                  or loads per Gauss point, where K p is applied from the
 *                  assembled stencil). */
enum { MSFEC_SOLVER_BAND_FP64 = 0, MSFEC_SOLVER_BAND_FP64_MFMA = 1, MSFEC_SOLVER_BAND_MIXED = 2,
       MSFEC_SOLVER_CG = 3 };
/* per-cell status codes.  NOT_CONVERGED: BAND_MIXED stopped after
 * max_refine steps with a predicted error above 1e-10 relative, or CG
 * reached cg_max_iter.  NOT_SPD also reports a CG breakdown (a diagonal
 * entry or p^T K p <= 0). */
enum { MSFEC_CELL_OK = 0, MSFEC_CELL_NOT_CONVERGED = 1, MSFEC_CELL_NOT_SPD = 2, MSFEC_CELL_NONFINITE = 3 };
/* flags */
enum { MSFEC_FLAG_TIMING = 1 /* record per-kernel HIP events in compute */ };

/* The fields of ParametersBasis the basis reads
 * (include/Q/q_parameters.h:78-101, source/Q/q_parameters.cc:252-262), plus
 * the problem data the reference reads from the parameter file
 * (source/equation_data/eqn_coeff_A.cc, eqn_rhs.cc). */
typedef struct {
  int dim;             /* 2 or 3                                           */
  int n_refine_local;  /* ParametersBasis::n_refine_local                  */
  int coef_mode;       /* MSFEC_COEF_*                                     */
  int rhs_mode;        /* MSFEC_RHS_*                                      */
  double rhs_const;    /* f for MSFEC_RHS_CONST                            */
  int reaction_mode;   /* MSFEC_REACTION_*                                 */
  double reaction_const;
  int solver;          /* MSFEC_SOLVER_* (replaces use_direct_solver)      */
  int flags;           /* MSFEC_FLAG_*                                     */
  int max_refine;      /* BAND_MIXED: refinement steps at most (0 = 3)     */
  double cg_tol;       /* CG: relative residual to reach (0 = 1e-13)       */
  int cg_max_iter;     /* CG: iterations at most (0 = 5000)                */
  int band_window;     /* BAND_FP64_MFMA, scalar coefficient: slots of the
                          band window, 16, 32 or 64 (it must hold the
                          half-bandwidth + 4); 0 = the smallest that does.
                          The other solvers keep 64.                       */
} msfec_q_params;

typedef struct {
  int n_cells, dim, n_refine_local;
  int n_basis;         /* nb = 2^dim                                       */
  int n_fine_nodes;    /* (2^l + 1)^dim                                    */
  int n_fine_cells;    /* 2^(l dim)                                        */
  int n_interior;      /* (2^l - 1)^dim                                    */
  int bandwidth;       /* half-bandwidth of the interior system            */
  int band_window;     /* slots of the matrix-core band window in use      */
  size_t coef_count;   /* doubles expected by msfec_q_set_coefficient      */
  size_t rhs_count;    /* doubles expected by msfec_q_set_rhs (0: none)    */
  size_t reaction_count;
  size_t device_bytes; /* device memory held by the handle                 */
} msfec_q_info;

typedef struct msfec_q_batch *msfec_q_handle;

/* Create a batch of n_cells coarse cells.  vertices: n_cells x 2^dim x dim
 * doubles; each cell must be an axis-aligned box (the coarse meshes of the
 * reference are refined hyper_cubes, q_global.cc:96-98).
 * Replaces QBasis::QBasis (source/Q/q_basis.cc:7-56, include/Q/q_basis.h:77-82)
 * for every cell of the map built in QMultiscale::initialize_and_compute_basis
 * (source/Q/q_global.cc:38-71). */
msfec_status msfec_q_create(const msfec_q_params *params, int n_cells,
                            const double *vertices, int vertices_loc,
                            msfec_q_handle *out);

/* Coefficient / load / reaction data (layouts above).  count must equal the
 * value reported by msfec_q_info.  Replaces the evaluation of Diffusion_A,
 * RightHandSideParsed and ReactionRate inside QBasis::assemble_system
 * (source/Q/q_basis.cc:196-220). */
msfec_status msfec_q_set_coefficient(msfec_q_handle h, const double *data,
                                     size_t count, int loc);
msfec_status msfec_q_set_rhs(msfec_q_handle h, const double *data, size_t count,
                             int loc);
msfec_status msfec_q_set_reaction(msfec_q_handle h, const double *data,
                                  size_t count, int loc);

/* Compute every basis of every cell: fine assembly, Dirichlet lifting of the
 * coarse Q1 shape functions, the 2^dim local solves and the coarse element
 * matrix/RHS.  Asynchronous on `stream`: kernel launches only (no memsets, no
 * host synchronisation, no allocation), so one call can be captured into a
 * HIP graph and replayed; replays read the handle's current coefficient,
 * load and reaction buffers.  The same holds for msfec_ned_compute and
 * msfec_rt_compute.
 * Replaces QBasis::run (source/Q/q_basis.cc:518-603), i.e. make_grid,
 * setup_system, assemble_system, the condense + solve_direct/solve_iterative
 * loop and assemble_global_element_matrix (:95-357). */
msfec_status msfec_q_compute(msfec_q_handle h, void *stream);

/* msfec_q_compute without a handle and without any allocation: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  It runs the kernels of msfec_q_compute
 * with MSFEC_SOLVER_BAND_FP64_MFMA (bit-identical results) and takes the
 * params of msfec_q_create restricted to that solver, a scalar coefficient
 * per fine cell, MSFEC_RHS_CONST and MSFEC_REACTION_NONE / _CONST.  Every
 * pointer is device memory owned by the caller: vertices (n_cells x 2^dim x
 * dim, axis-aligned boxes, not checked here), coef (n_cells x 2^(l dim)),
 * Psi (n_cells x nb x n_fine_nodes), M (n_cells x nb x nb), b (n_cells x
 * nb), status (n_cells ints).  Asynchronous on `stream`: kernel launches
 * only (no allocation, no memset, no host synchronisation).
 * Scratch (row table, panels, forward results) lives in one 128 MiB device
 * array per GPU, held from the library's first use on that GPU until the
 * process ends; a call processes its cells in chunks that fit it (3-D l = 3:
 * 498 cells per chunk, one after another on the stream).  Every scratch
 * word a chunk reads is written earlier in the same chunk, so no clearing is
 * needed; calls that overlap on one GPU (on different streams or from
 * different threads) share the array and race.  Rejected with
 * MSFEC_ERR_UNSUPPORTED: other solvers or modes, and levels beyond the band
 * window (3-D l > 3, 2-D l > 5). */
msfec_status msfec_q_compute_noalloc(const msfec_q_params *params, int n_cells,
                                     const double *vertices, const double *coef,
                                     double *Psi, double *M, double *b, int *status,
                                     void *stream);

/* Wait for the work of the last compute. */
msfec_status msfec_q_synchronize(msfec_q_handle h);

/* M: n_cells x nb x nb, b: n_cells x nb (either may be NULL).  Synchronizes.
 * Replaces QBasis::get_global_element_matrix / get_global_element_rhs
 * (source/Q/q_basis.cc:361-373). */
msfec_status msfec_q_get_element(msfec_q_handle h, double *M, double *b, int loc);

/* Psi: n_cells x nb x n_fine_nodes, basis-major.  Synchronizes.
 * Replaces the per-cell solution_vector array of QBasis
 * (include/Q/q_basis.h:253). */
msfec_status msfec_q_get_basis(msfec_q_handle h, double *Psi, int loc);

/* Per-cell status (MSFEC_CELL_*), n_cells ints on the host.  Synchronizes.
 * The reference aborts the run on a failed local solve (q_basis.cc:337); here
 * the failing cells are reported and the others stay valid. */
msfec_status msfec_q_get_cell_status(msfec_q_handle h, int *status);

/* u_fine[c][node] = sum_k w[c][k] Psi[c][k][node].  weights: n_cells x nb.
 * Replaces QBasis::set_global_weights (source/Q/q_basis.cc:394-416). */
msfec_status msfec_q_reconstruct(msfec_q_handle h, const double *weights,
                                 int weights_loc, double *u_fine, int out_loc);

/* Sizes of the batch. */
msfec_status msfec_q_get_info(msfec_q_handle h, msfec_q_info *info);

/* Per-kernel times (ms), averaged over every compute made with
 * MSFEC_FLAG_TIMING since the previous call (which this call then forgets):
 * [0] assembly (row table or stencil tables), [1] factor + solve (for
 * BAND_MIXED: the FP32 factor and first solution), [2] refinement residual
 * (BAND_MIXED), [3] refinement sweep and element matrix (BAND_MIXED),
 * [4] element matrix (tensor coefficients).  Entries that do not apply are
 * zero.  Synchronizes.  The number of entries written is returned in *n. */
msfec_status msfec_q_kernel_times(msfec_q_handle h, float *ms, int max, int *n);

void msfec_q_destroy(msfec_q_handle h);

/* Coarse multiscale problem on a structured grid of nc^dim coarse cells of
 * [0,1]^dim (cells and nodes lexicographic, x fastest) with homogeneous
 * Dirichlet data on the whole boundary.  Assembles the element matrices and
 * right-hand sides of a basis batch whose cells follow that order and solves
 * by Jacobi-preconditioned CG in FP64 on the device until
 * ||r|| <= max(abs_tol, rel_tol ||b||).  u (n_nodes doubles, loc) receives the
 * nodal solution, zero on the boundary; iterations and the final residual are
 * returned.  Replaces QMultiscale::assemble_system, solve_iterative and the
 * weight extraction of send_global_weights_to_cell
 * (source/Q/q_global.cc:198-297, 327-363, 368-392). */
msfec_status msfec_coarse_solve(msfec_q_handle h, int nc, double abs_tol,
                                double rel_tol, int max_iter, double *u,
                                int u_loc, double *weights, int weights_loc,
                                int *iterations, double *residual);

/* Adjoint gradient of the coarse misfit J = 1/2 sum over the free coarse
 * nodes of (u_H - u_target)^2 with respect to the log-coefficient g of every
 * fine cell (coefficient a = exp(g), MSFEC_COEF_SCALAR_CELL), for a batch laid
 * out as in msfec_coarse_solve and computed with BAND_FP64_MFMA or BAND_MIXED
 * (the stored panels give w = K_II^-1 F_I).  u_target: (nc+1)^dim nodal
 * values (boundary entries ignored); J on the host; grad: n_cells x N^dim in
 * the layout of the coefficient.  The coarse solves run to rel_tol.
 *   dJ/dg_e = -a_e vt_e^T Khat_e (ut_e + w_e)
 * with ut, vt the fine reconstructions of the coarse solution and of the
 * adjoint K_H v = u_H - u_target, and Khat_e the unit-coefficient element
 * stiffness.  No reference counterpart: the reference has no adjoint; this is
 * the differentiation target of the coarse-scale inverse problem. */
msfec_status msfec_q_adjoint_gradient(msfec_q_handle h, int nc, const double *u_target,
                                      int target_loc, double rel_tol, int max_iter,
                                      double *J, double *grad, int grad_loc);

/* ---------------------------------------------------------------------------
 * Uneven batches: coarse cells with their own refinement level.
 *
 * The reference refines every cell by the one n_refine_local of the
 * parameter file (source/Q/q_parameters.cc:229, q_basis.cc:101); here each
 * cell of one std::map<CellId, QBasis> (q_global.cc:38-87) may carry its own
 * level.  One handle holds one batch per level present, each with its own
 * local solver:
 *   - solver MSFEC_SOLVER_AUTO: where the interior band fits the band window
 *     (3-D l <= 3, 2-D l <= 5) the fastest of the FP64 solvers and tile
 *     shapes (matrix-core band with a 16-, 32- or 64-slot window, vector-ALU
 *     band, CG), timed on up to 256 of the level's own cells
 *     with the current coefficients (a candidate whose cells do not all come
 *     out MSFEC_CELL_OK is rejected); matrix-free CG beyond the window.  The
 *     choice is made by msfec_qu_tune, or by the first compute when no tuning
 *     was done (that compute synchronizes, so a graph capture comes after an
 *     eager compute); it is kept per handle until msfec_qu_tune is called
 *     again (for instance after a coefficient field of different contrast);
 *   - any other solver: that solver where it applies, CG beyond the window.
 * compute forks the levels onto streams of their own (default priority)
 * after the caller's stream and joins them back, so the levels overlap; the
 * fork and join are event waits, so a compute can be captured into a graph.
 * Layouts, cells in batch order:
 *   coefficient (and per-cell load / reaction): cell c holds its N_c^dim
 *     values (N_c = 2^levels[c]) at coef_offsets[c];
 *   Psi: cell c holds nb x (N_c + 1)^dim values at basis_offsets[c];
 *   M, b, status: n_cells x nb x nb, n_cells x nb, n_cells as for msfec_q.
 * Only MSFEC_COEF_SCALAR_CELL and MSFEC_RHS_CONST / _CELL are taken. */
enum { MSFEC_SOLVER_AUTO = 4 };

typedef struct msfec_qu_batch *msfec_qu_handle;

/* params as for msfec_q_create (n_refine_local is ignored), levels: n_cells
 * ints on the host.  Replaces QBasis::QBasis for every cell of the map
 * (source/Q/q_basis.cc:7-56, q_global.cc:38-71) with a level per cell. */
msfec_status msfec_qu_create(const msfec_q_params *params, int n_cells,
                             const double *vertices, int vertices_loc,
                             const int *levels, msfec_qu_handle *out);
/* n_cells + 1 prefix offsets (host) of the coefficient and Psi layouts;
 * either pointer may be NULL. */
msfec_status msfec_qu_offsets(msfec_qu_handle h, size_t *coef_offsets,
                              size_t *basis_offsets);
/* count = coef_offsets[n_cells]. */
msfec_status msfec_qu_set_coefficient(msfec_qu_handle h, const double *data,
                                      size_t count, int loc);
msfec_status msfec_qu_set_rhs(msfec_qu_handle h, const double *data, size_t count,
                              int loc);
msfec_status msfec_qu_set_reaction(msfec_qu_handle h, const double *data,
                                   size_t count, int loc);
/* Time the candidate solvers of every level (MSFEC_SOLVER_AUTO) on the
 * current coefficients and keep the fastest; synchronous. */
msfec_status msfec_qu_tune(msfec_qu_handle h);
/* The solver of a level present in the batch, its band window (slots; 0
 * for the solvers without one), its cell count and, when tuned, the timed
 * milliseconds of the 5 candidates (matrix-core band with a 16-, 32- and
 * 64-slot window, vector-ALU band, CG; -1 where not timed, not applicable or
 * rejected).  Any pointer may be NULL. */
msfec_status msfec_qu_level_info(msfec_qu_handle h, int level, int *solver, int *window,
                                 int *n_cells, float *ms5);
/* The batch of one level, for read-only queries through msfec_q_get_info,
 * msfec_q_kernel_times (params.flags MSFEC_FLAG_TIMING) and
 * msfec_q_get_cg_iterations; owned by h (do not destroy or set data). */
msfec_status msfec_qu_level_handle(msfec_qu_handle h, int level, msfec_q_handle *out);
/* QBasis::run (source/Q/q_basis.cc:518-603) for every cell of every level;
 * asynchronous on stream. */
msfec_status msfec_qu_compute(msfec_qu_handle h, void *stream);
msfec_status msfec_qu_synchronize(msfec_qu_handle h);
/* As msfec_q_get_element / _get_basis / _get_cell_status, cells in batch
 * order (Psi in the ragged layout above).  Synchronize. */
msfec_status msfec_qu_get_element(msfec_qu_handle h, double *M, double *b, int loc);
msfec_status msfec_qu_get_basis(msfec_qu_handle h, double *Psi, int loc);
msfec_status msfec_qu_get_cell_status(msfec_qu_handle h, int *status);
/* As msfec_coarse_solve for a batch whose cells are the nc^dim coarse
 * cells in lexicographic order, whatever their levels (the coarse problem
 * sees only M and b; q_global.cc:198-392). */
msfec_status msfec_qu_coarse_solve(msfec_qu_handle h, int nc, double abs_tol,
                                   double rel_tol, int max_iter, double *u, int u_loc,
                                   double *weights, int weights_loc, int *iterations,
                                   double *residual);
void msfec_qu_destroy(msfec_qu_handle h);

/* ---------------------------------------------------------------------------
 * Q_Ned: the mixed Q1 x lowest-order Nedelec multiscale basis, two-dimensional
 * (the reference's QNedBasis, source/Q_Ned/q_ned_basis.cc, is its 3-D form).
 * Per coarse box cell and fine cell (2^l x 2^l, lexicographic, x fastest):
 * a (curl-curl coefficient), binv (B^-1), r (reaction), f (2 components).
 * 8 bases: 0..3 vertex bases (deal.II vertex order: sigma = phi_n and
 * u x n = grad phi_n on the boundary, source (v, grad phi_n) when full_rhs),
 * 4..7 edge bases (deal.II line order x = x0, x = x1, y = y0, y = y1:
 * sigma = 0 and u x n = N_E on the boundary).  Fine degrees of freedom of a
 * basis (ns + nu per cell, m = 2^l): sigma at the (m+1)^2 nodes (lexicographic),
 * then u as line integrals of the tangential component (tangent +x / +y) on
 * the m (m+1) horizontal edges (edge (i,j) at j m + i) and the (m+1) m
 * vertical edges (edge (i,j) at j (m+1) + i).
 * ------------------------------------------------------------------------- */
typedef struct {
  int n_refine_local; /* l, 1..4 (the interior half-bandwidth must stay <= 60) */
  int full_rhs;       /* QNedBasis' full_rhs parameter (q_ned_basis.cc:432-444) */
  int flags;          /* MSFEC_FLAG_*                                          */
  int solver;         /* 0 or MSFEC_SOLVER_BAND_FP64_MFMA: FP64 LDL^T on the
                         matrix cores; MSFEC_SOLVER_BAND_MIXED: FP32 LDL^T on
                         the matrix cores and refinement with FP64 residuals
                         until the predicted error |dx| / |x| times the
                         contraction (|dx| / |x| over its previous value) is
                         below 1e-13 per cell, at most max_refine passes
                         (0 = 5), the cells still above 1e-12 flagged
                         MSFEC_CELL_NOT_CONVERGED                          */
  int max_refine;
} msfec_ned_params;

typedef struct msfec_ned_batch *msfec_ned_handle;

/* Create a batch.  vertices: n_cells x 4 x 2 doubles, axis-aligned boxes.
 * Replaces QNedBasis::QNedBasis + setup_grid/setup_system_matrix
 * (source/Q_Ned/q_ned_basis.cc:27-217, include/Q_Ned/q_ned_basis.h:98-112). */
msfec_status msfec_ned_create(const msfec_ned_params *params, int n_cells,
                              const double *vertices, int vertices_loc,
                              msfec_ned_handle *out);

/* Per-fine-cell data, n_cells x 4^l doubles each (f: x 2).  Stands for the
 * Diffusion_A / DiffusionInverse_B / ReactionRate / RightHandSide evaluations
 * of QNedBasis::assemble_system (q_ned_basis.cc:392-412, 445-452). */
msfec_status msfec_ned_set_coefficients(msfec_ned_handle h, const double *a,
                                        const double *binv, const double *r,
                                        const double *f, int loc);

/* All cells: assembly, lifting, the interior solves of the 8 bases and the
 * element matrix, enqueued on stream.  Replaces QNedBasis::run
 * (q_ned_basis.cc:1281-1429) with solve_direct (:570-625). */
msfec_status msfec_ned_compute(msfec_ned_handle h, void *stream);
msfec_status msfec_ned_synchronize(msfec_ned_handle h);

/* M: n_cells x 8 x 8 (row = test basis), b: n_cells x 8, as built by
 * QNedBasis::assemble_global_element_matrix (q_ned_basis.cc:841-939) and
 * returned by get_global_element_matrix/_rhs (include/Q_Ned/q_ned_basis.h:117-129). */
msfec_status msfec_ned_get_element(msfec_ned_handle h, double *M, double *b, int loc);

/* Psi: n_cells x 8 x (ns + nu), the fine coefficients of the bases
 * (basis_h1_v / basis_curl_v of QNedBasis, renumbered as above). */
msfec_status msfec_ned_get_basis(msfec_ned_handle h, double *Psi, int loc);

/* Per-cell status: MSFEC_CELL_OK, MSFEC_CELL_NOT_SPD for a zero pivot,
 * MSFEC_CELL_NONFINITE for non-finite data. */
msfec_status msfec_ned_get_cell_status(msfec_ned_handle h, int *status);

/* Mean per-kernel times in ms (rows, factor + solves, element) of the computes
 * made with MSFEC_FLAG_TIMING since the previous call. */
msfec_status msfec_ned_kernel_times(msfec_ned_handle h, float *ms, int max, int *n);
msfec_status msfec_ned_destroy(msfec_ned_handle h);

/* ---------------------------------------------------------------------------
 * RT_DQ: the mixed lowest-order Raviart-Thomas x DQ0 multiscale basis,
 * two-dimensional (the reference's RTDQBasis, source/RT_DQ/rt_dq_basis.cc, is
 * its 3-D form).  Per coarse box cell and fine cell (2^l x 2^l,
 * lexicographic, x fastest): a (scalar diffusion, or the diffusion tensor at
 * the Gauss points; the solve uses its inverse like DiffusionInverse_A) and
 * f (load).  4 bases, one per coarse face in deal.II
 * face order (x = x0, x = x1, y = y0, y = y1): the boundary flux is that of
 * the coarse Raviart-Thomas function of the face and the divergence -+1/|K|
 * (is_laplace, rt_dq_basis.cc:516-538).  Fine degrees of freedom of a basis
 * (ne = 2 m (m+1), m = 2^l): the flux through each fine edge in the +x / +y
 * direction, the m (m+1) vertical edges first (edge (P, q), x-node P and cell
 * row q, at q (m+1) + P), then the horizontal edges (edge (p, Q) at
 * m (m+1) + Q m + p).  The u part of every basis is the constant 1
 * (set_u_to_std, rt_dq_basis.cc:1039-1043) and is not stored.  The element
 * matrix is 5 x 5 with index 4 the DQ0 basis.
 * ------------------------------------------------------------------------- */
typedef struct {
  int n_refine_local; /* l, 1..5 (the reduced system's half-bandwidth 2^l <= 60) */
  int flags;          /* MSFEC_FLAG_*                                          */
  int coef_mode;      /* MSFEC_COEF_SCALAR_CELL: a per fine cell (A = a I);
                         MSFEC_COEF_TENSOR_QP: the symmetric A (xx, xy, yy) at
                         the 4 QGauss(2) points of each fine cell, points
                         x fastest (what Diffusion_A gives there); the solve
                         inverts it at each point like DiffusionInverse_A */
} msfec_rt_params;

typedef struct msfec_rt_batch *msfec_rt_handle;

/* Create a batch.  vertices: n_cells x 4 x 2 doubles, axis-aligned boxes.
 * Replaces RTDQBasis::RTDQBasis + setup_grid/setup_system_matrix
 * (source/RT_DQ/rt_dq_basis.cc:7-265, include/RT_DQ/rt_dq_basis.h:96-111). */
msfec_status msfec_rt_create(const msfec_rt_params *params, int n_cells,
                             const double *vertices, int vertices_loc,
                             msfec_rt_handle *out);

/* Per-fine-cell data: a, n_cells x 4^l doubles (MSFEC_COEF_SCALAR_CELL) or
 * n_cells x 4^l x 4 x 3 (MSFEC_COEF_TENSOR_QP); f, n_cells x 4^l, may be NULL
 * (f = 1).  Stands for the DiffusionInverse_A and RightHandSideParsed
 * evaluations of RTDQBasis::assemble_system (rt_dq_basis.cc:423-461). */
msfec_status msfec_rt_set_coefficients(msfec_rt_handle h, const double *a, const double *f,
                                       int loc);

/* All cells: the flux solves of the 4 bases and the element matrix, enqueued
 * on stream.  Replaces RTDQBasis::run (rt_dq_basis.cc:1114-1170) with
 * solve_direct (:582-612). */
msfec_status msfec_rt_compute(msfec_rt_handle h, void *stream);
msfec_status msfec_rt_synchronize(msfec_rt_handle h);

/* M: n_cells x 5 x 5 (row = test basis), b: n_cells x 5, as built by
 * RTDQBasis::assemble_global_element_matrix (rt_dq_basis.cc:822-914) and
 * returned by get_global_element_matrix/_rhs (rt_dq_basis.cc:1096-1106). */
msfec_status msfec_rt_get_element(msfec_rt_handle h, double *M, double *b, int loc);

/* sigma: n_cells x 4 x ne, the fine fluxes of the bases (the sigma blocks of
 * basis_div_v of RTDQBasis, renumbered as above). */
msfec_status msfec_rt_get_basis(msfec_rt_handle h, double *sigma, int loc);

/* The fine solution from the coarse weights w (n_cells x 5: 4 face weights,
 * then the cell weight): sigma = sum_i w_i sigma_i (n_cells x ne) and
 * u = w_4 on every fine cell (n_cells x 4^l); either output may be NULL.
 * Replaces RTDQBasis::set_global_weights (rt_dq_basis.cc:1011-1036).
 * Synchronizes. */
msfec_status msfec_rt_reconstruct(msfec_rt_handle h, const double *w, int w_loc, double *sigma,
                                  double *u, int out_loc);

/* Per-cell status: MSFEC_CELL_OK, MSFEC_CELL_NOT_SPD for a zero pivot,
 * MSFEC_CELL_NONFINITE for a non-finite element matrix. */
msfec_status msfec_rt_get_cell_status(msfec_rt_handle h, int *status);

/* Mean per-kernel times in ms (rows, factor + solves, element) of the computes
 * made with MSFEC_FLAG_TIMING since the previous call. */
msfec_status msfec_rt_kernel_times(msfec_rt_handle h, float *ms, int max, int *n);
msfec_status msfec_rt_destroy(msfec_rt_handle h);

/* ---------------------------------------------------------------------------
 * Ned_RT: the mixed lowest-order Nedelec x Raviart-Thomas multiscale basis in
 * three dimensions (the reference's NedRTBasis, source/Ned_RT/
 * ned_rt_basis.cc).  Per coarse box cell, on the 2^l x 2^l x 2^l fine grid:
 * sigma on the fine edges, u on the fine faces,
 *   (A^-1 sigma, tau) - (u, curl tau) = 0,
 *   (v, curl sigma) + (B div u, div v) + (R u, v) = (v, curl N_n),
 * 12 edge bases (deal.II line order; the tangential trace of the coarse
 * Nedelec function N_n on the boundary, the load only for them) and 6 face
 * bases (deal.II face order; the normal trace of the coarse Raviart-Thomas
 * function).  Fine degrees of freedom (m = 2^l, nx = m (m+1)^2, ne = 3 nx,
 * nf1 = m^2 (m+1), N = ne + 3 nf1):
 *   x-edge (i, j, k) [nodes (i,j,k)-(i+1,j,k)] at (k (m+1) + j) m + i,
 *   y-edge at nx + (k m + j) (m+1) + i, z-edge at 2 nx + (k (m+1) + j) (m+1) + i
 *   (the line integral of the tangential component along +x, +y, +z);
 *   x-face (i, j, k) [x = x0 + i hx] at ne + (k m + j) (m+1) + i,
 *   y-face at ne + nf1 + (k (m+1) + j) m + i, z-face at ne + 2 nf1 + (k m + j) m + i
 *   (the flux in the +x, +y, +z direction).
 * The element matrix is 18 x 18 (row = test basis; edge bases first).
 * ------------------------------------------------------------------------- */
typedef struct {
  int n_refine_local; /* l, 1..4                                                      */
  int flags;          /* MSFEC_FLAG_*                                                 */
  int solver;         /* 0: the dense direct solve for l <= 2 (the interior system of
                         up to 256 unknowns factored on the matrix cores) and MINRES
                         above; MSFEC_SOLVER_CG: matrix-free MINRES for every l      */
  int max_iter;       /* MINRES iteration limit, 0: 20000; a solve that reaches it
                         flags its cell MSFEC_CELL_NOT_CONVERGED                      */
  double rtol;        /* MINRES stop: preconditioned residual <= rtol x the initial
                         one, 0: 1e-14                                                */
} msfec_nrt_params;

typedef struct msfec_nrt_batch *msfec_nrt_handle;

/* Create a batch.  vertices: n_cells x 8 x 3 doubles, axis-aligned boxes in
 * deal.II vertex order.  Replaces NedRTBasis::NedRTBasis + setup_grid /
 * setup_system_matrix / setup_basis_dofs_curl / setup_basis_dofs_div
 * (source/Ned_RT/ned_rt_basis.cc:7-359, include/Ned_RT/ned_rt_basis.h).
 * MSFEC_ERR_UNSUPPORTED for l > 4.  The MINRES path holds 7 fine vectors per
 * basis and 177 doubles per fine cell in device memory (l = 4: 32 MiB a
 * cell). */
msfec_status msfec_nrt_create(const msfec_nrt_params *params, int n_cells,
                              const double *vertices, int vertices_loc,
                              msfec_nrt_handle *out);

/* Values at the 8 QGauss(2) points of every fine cell (fine cells x fastest,
 * then y, z; points x fastest): ainv, n_cells x 8^l x 8 x 6 (xx, xy, xz, yy,
 * yz, zz of DiffusionInverse_A); b, n_cells x 8^l x 8 (Diffusion_B); r, the
 * reaction rate, same shape, NULL for zero (the reference's ReactionRate,
 * eqn_coeff_R.cc:8-14); f, n_cells x 8^l x 8 x 3 (RightHandSideParsed), NULL
 * for zero.  Stands for the value_list calls of NedRTBasis::assemble_system
 * (ned_rt_basis.cc:453-461). */
msfec_status msfec_nrt_set_coefficients(msfec_nrt_handle h, const double *ainv, const double *b,
                                        const double *r, const double *f, int loc);

/* All cells: the 18 basis solves and the element matrix, enqueued on stream.
 * Replaces NedRTBasis::run (ned_rt_basis.cc:1280-1352) with solve_direct
 * (:579-634) and assemble_global_element_matrix (:850-948). */
msfec_status msfec_nrt_compute(msfec_nrt_handle h, void *stream);
msfec_status msfec_nrt_synchronize(msfec_nrt_handle h);

/* M: n_cells x 18 x 18, b: n_cells x 18 (zero for the edge bases), as built
 * by assemble_global_element_matrix and returned by
 * get_global_element_matrix/_rhs (ned_rt_basis.cc:1262-1272). */
msfec_status msfec_nrt_get_element(msfec_nrt_handle h, double *M, double *b, int loc);

/* X: n_cells x 18 x N, the full fine vectors of the bases (basis_curl_v and
 * basis_div_v of NedRTBasis, renumbered as above). */
msfec_status msfec_nrt_get_basis(msfec_nrt_handle h, double *X, int loc);

/* The fine solution from the coarse weights w (n_cells x 18): sigma =
 * sum_{i<12} w_i sigma_i (edge part of the edge bases, n_cells x ne) and u =
 * sum_{i>=12} w_i u_i (face part of the face bases, n_cells x (N - ne)); either
 * output may be NULL.  Replaces NedRTBasis::set_global_weights
 * (ned_rt_basis.cc:1156-1181).  Synchronizes. */
msfec_status msfec_nrt_reconstruct(msfec_nrt_handle h, const double *w, int w_loc, double *sigma,
                                   double *u, int out_loc);

/* Per-cell status: MSFEC_CELL_OK, MSFEC_CELL_NOT_SPD when a pivot of the
 * symmetric indefinite factorisation has the wrong sign (edge pivots must be
 * positive, face pivots negative), MSFEC_CELL_NOT_CONVERGED when a MINRES
 * solve reached max_iter, MSFEC_CELL_NONFINITE for non-finite data. */
msfec_status msfec_nrt_get_cell_status(msfec_nrt_handle h, int *status);

/* MINRES iterations of the last compute, n_cells x 18 ints on the host (the
 * reference's solver_control.last_step() of solve_iterative,
 * ned_rt_basis.cc:796-801); MSFEC_ERR_STATE for a handle on the dense path.
 * Synchronizes. */
msfec_status msfec_nrt_get_iterations(msfec_nrt_handle h, int *iterations);

/* Mean times in ms over the computes made with MSFEC_FLAG_TIMING since the
 * previous call: the whole compute (the dense path's one kernel) and, on the
 * MINRES path, its local-matrix + preconditioner, MINRES and element-matrix
 * parts (up to max values, *n of them written). */
msfec_status msfec_nrt_kernel_times(msfec_nrt_handle h, float *ms, int max, int *n);
msfec_status msfec_nrt_destroy(msfec_nrt_handle h);

/* ---------------------------------------------------------------------------
 * RT_DQ in three dimensions, as the reference's RTDQBasis
 * (source/RT_DQ/rt_dq_basis.cc): the 6 face bases of the mixed
 * Raviart-Thomas x DQ0 system per coarse box cell, on the 2^l x 2^l x 2^l
 * fine grid (fine faces numbered as the faces of the Ned_RT section without
 * the edge offset: nf = 3 m^2 (m+1)).  The flux of basis n is the coarse
 * Raviart-Thomas function of face n plus the curl of an interior Nedelec
 * field, solved by Jacobi-PCG on the curl-curl system; u is 1
 * (set_u_to_std).  No reaction term (the reference's ReactionRate is zero).
 * ------------------------------------------------------------------------- */
typedef struct {
  int n_refine_local; /* l, 1..4                                                  */
  int flags;          /* MSFEC_FLAG_*                                             */
  int max_iter;       /* CG iteration limit, 0: 20000 (MSFEC_CELL_NOT_CONVERGED) */
  double rtol;        /* CG stop: |r| <= rtol |b|, 0: 1e-13                       */
} msfec_rt3_params;

typedef struct msfec_rt3_batch *msfec_rt3_handle;

/* vertices: n_cells x 8 x 3, axis-aligned boxes in deal.II vertex order.
 * Replaces RTDQBasis::RTDQBasis + setup_grid / setup_system_matrix /
 * setup_basis_dofs (rt_dq_basis.cc:7-380). */
msfec_status msfec_rt3_create(const msfec_rt3_params *params, int n_cells,
                              const double *vertices, int vertices_loc,
                              msfec_rt3_handle *out);
/* ainv: n_cells x 8^l x 8 x 6 (xx, xy, xz, yy, yz, zz of DiffusionInverse_A at
 * the QGauss(2) points of every fine cell, fine cells x fastest); f:
 * n_cells x 8^l x 8 (the right-hand side at the same points), NULL for zero
 * (rt_dq_basis.cc:455-470). */
msfec_status msfec_rt3_set_coefficients(msfec_rt3_handle h, const double *ainv, const double *f,
                                        int loc);
/* All cells: the 6 basis solves and the element matrix, enqueued on stream.
 * Replaces RTDQBasis::run (rt_dq_basis.cc:1114-1170) and
 * assemble_global_element_matrix (:822-914). */
msfec_status msfec_rt3_compute(msfec_rt3_handle h, void *stream);
msfec_status msfec_rt3_synchronize(msfec_rt3_handle h);
/* M: n_cells x 7 x 7 (row = test basis: the 6 face bases in deal.II face
 * order, then the DQ0 basis), b: n_cells x 7 (zero but b_6 = (f, 1))
 * (get_global_element_matrix/_rhs, rt_dq_basis.cc:1096-1106). */
msfec_status msfec_rt3_get_element(msfec_rt3_handle h, double *M, double *b, int loc);
/* sigma: n_cells x 6 x nf, the fine fluxes of the face bases in the +x, +y,
 * +z directions (the sigma block of basis_div_v; its u block is 1). */
msfec_status msfec_rt3_get_basis(msfec_rt3_handle h, double *sigma, int loc);
/* MSFEC_CELL_OK, MSFEC_CELL_NOT_CONVERGED, MSFEC_CELL_NOT_SPD (a breakdown of
 * CG), MSFEC_CELL_NONFINITE.  Synchronizes. */
msfec_status msfec_rt3_get_cell_status(msfec_rt3_handle h, int *status);
/* CG iterations of the last compute, n_cells x 6 ints on the host.
 * Synchronizes. */
msfec_status msfec_rt3_get_iterations(msfec_rt3_handle h, int *iterations);
/* Mean times in ms over the computes made with MSFEC_FLAG_TIMING since the
 * previous call: whole, face masses + preconditioner, CG, element matrix. */
msfec_status msfec_rt3_kernel_times(msfec_rt3_handle h, float *ms, int max, int *n);
msfec_status msfec_rt3_destroy(msfec_rt3_handle h);

/* CG iterations of the last compute, n_cells x 2^dim ints on the host
 * (basis-minor), for a handle created with MSFEC_SOLVER_CG.  The reference
 * reports CG iterations per basis solve in verbose mode
 * (source/Q/q_basis.cc:349-353).  Without a reaction term the last basis is
 * not iterated: it is 1 minus the sum of the others (the bases form a
 * partition of unity), and its entry repeats the largest count of the
 * others. */
msfec_status msfec_q_get_cg_iterations(msfec_q_handle h, int *iterations);

/* Thread-local description of the last error of this thread. */
const char *msfec_last_error(void);
const char *msfec_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MSFEC_H */

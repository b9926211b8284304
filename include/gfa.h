/*
 * gfa.h -- C-ABI of the MI355X Green-function action (H - E S) G and its tfQMR solve.
 *
 * One plan object replaces the reference's action_plan_t + green_action::action_t + green_solver_t
 * trio.  Plain pointers and sizes only; device pointers are HIP device addresses.  Every entry point
 * returns a status (GFA_OK == 0) and never exits the process; gfa_last_error() explains the last
 * failure of the calling thread.
 *
 * Block layout of a Green function X (input and output of the action):
 *     real_t X[nnzb][2][64][64]      real_t = double (fp_bits 64) or float (fp_bits 32)
 *     [inzb][re/im plane][target grid point i = (z*4 + y)*4 + x][right-hand side j]
 * identical to the reference (include/green_action.hxx:141-142).
 */
#ifndef GFA_H
#define GFA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFA_OK            0
#define GFA_ERR_ARGUMENT  1   /* invalid argument or unsupported option          */
#define GFA_ERR_DEVICE    2   /* HIP runtime failure (no device, launch failure) */
#define GFA_ERR_STATE     3   /* call out of order (e.g. no energy set)           */
#define GFA_ERR_MEMORY    4   /* device or host allocation failed                  */

typedef struct gfa_plan gfa_plan;

/* Boundary conditions: 0 isolated, 1 periodic, 2 vacuum, 3 repeat (include/boundary_condition.hxx:17-40).
 * A periodic direction whose truncation sphere fits the cell is wrapped internally (code 5). */
typedef struct {
    uint32_t ng[3];             /* grid points per direction, multiples of 4                       */
    int8_t bc[3];               /* boundary condition per direction                                */
    int8_t reserved0;
    double hg[3];               /* grid spacing per direction (Bohr)                               */
    const double* xyzZinso;     /* [natoms*8]: x, y, z, Z, id, numax, sigma, unused (centred frame)*/
    int32_t natoms;
    int32_t echo;               /* log level, 0 = silent                                           */
    double r_trunc;             /* truncation radius (Bohr), green_function.truncation.radius      */
    double r_proj;              /* projection radius in units of sigma, green_function.projection.radius (6) */
    int32_t fd_range[3];        /* 0 default (8, or 4 if periodic), 4: 8th order, 8: 16th order   */
    int32_t nsources[3];        /* box of source blocks at the cell centre (green_function.sources), <0: all */
    const int32_t* sources;     /* optional explicit source block coordinates [nsources_explicit][3] */
    int32_t nsources_explicit;
    int32_t reserved1;
    double scale_T;             /* hamiltonian.scale.kinetic (1)                                   */
    const double* source_radii; /* optional [nsources_explicit] truncation radius of each source block
                                 * (batched (E, radius) instances); NULL: r_trunc for all. Sources may
                                 * repeat, each repetition is a column group of its own.            */
} gfa_plan_desc;

typedef struct {
    uint64_t nnzb;              /* non-zero 64x64 blocks of X                                      */
    uint32_t nRows, nCols;      /* target block rows, source block columns (64 RHS each)           */
    uint32_t nAtoms, nAtomImages;
    uint64_t natomcoeffs;       /* SHO coefficients over atom images                               */
    uint64_t ncoeffs;           /* SHO coefficients over atoms                                     */
    uint32_t nlists[3];         /* finite-difference lines per direction                           */
    int32_t fd_range[3];
    uint64_t nlist_entries[3];
    uint64_t sho_add_nnz, sho_prj_nnz, sho_sum_nnz;
    int8_t bc[3];               /* boundary conditions after wrapping (5 = wrap)                   */
    int8_t reserved;
    uint32_t nchunks;           /* projection work items                                           */
    uint64_t device_bytes;      /* device memory currently held by the plan's data and action workspace */
    uint64_t solver_bytes;      /* device memory currently held by the plan's solver state (gfa_solver_release frees it) */
} gfa_sizes;

/* arrays for gfa_plan_export (host copies, for verification against the reference) */
enum {
    GFA_EXPORT_ROW_START = 0,   /* uint32[nRows+1]    RowStart        src/green_function.cxx:801  */
    GFA_EXPORT_COLINDX,         /* uint16[nnzb]       colindx                                     */
    GFA_EXPORT_SUBSET,          /* uint32[nCols]      subset (diagonal blocks)                    */
    GFA_EXPORT_VEFF_INDEX,      /* int32[nnzb]        veff_index                                  */
    GFA_EXPORT_ROW_POS,         /* float[nRows][4]    rowCubePos                                  */
    GFA_EXPORT_COL_POS,         /* float[nCols][4]    colCubePos                                  */
    GFA_EXPORT_SOURCE_IDS,      /* int64[nCols]       global_source_indices                       */
    GFA_EXPORT_KIN_START_X,     /* uint32[nlists+1]   finite-difference list starts, x, y, z      */
    GFA_EXPORT_KIN_START_Y,
    GFA_EXPORT_KIN_START_Z,
    GFA_EXPORT_KIN_LIST_X,      /* int32[entries]     finite-difference lists, x, y, z            */
    GFA_EXPORT_KIN_LIST_Y,
    GFA_EXPORT_KIN_LIST_Z,
    GFA_EXPORT_IMAGE_POS,       /* double[nai][4]     AtomImagePos                                */
    GFA_EXPORT_IMAGE_LMAX,      /* int8[nai]                                                       */
    GFA_EXPORT_IMAGE_STARTS,    /* uint32[nai+1]                                                   */
    GFA_EXPORT_IMAGE_ATOM,      /* uint32[nai]        AtomImageIndex                              */
    GFA_EXPORT_IMAGE_SHIFT,     /* int8[nai][4]                                                    */
    GFA_EXPORT_ATOM_STARTS,     /* uint32[nAtoms+1]                                                */
    GFA_EXPORT_ATOM_LMAX,       /* int8[nAtoms]                                                    */
    GFA_EXPORT_ATOM_IDS,        /* int64[nAtoms]      global_atom_ids                             */
    GFA_EXPORT_ADD_START,       /* uint32[nnzb+1]     sparse_SHOadd                               */
    GFA_EXPORT_ADD_INDEX,       /* uint32[nnz]                                                     */
    GFA_EXPORT_PRJ_START,       /* uint32[nCols][nai+1] sparse_SHOprj, one CSR per column         */
    GFA_EXPORT_PRJ_INDEX,       /* uint32[nnz] concatenated over columns                          */
    GFA_EXPORT_SUM_START,       /* uint32[nAtoms+1]   sparse_SHOsum                               */
    GFA_EXPORT_SUM_INDEX,       /* uint32[nnz]                                                     */
    GFA_EXPORT_ROWINDX,         /* uint32[nnzb]       rowindx                                     */
    GFA_EXPORT_V_CELL,          /* double[ncell][64]  cell-indexed local potential               */
    GFA_EXPORT_ATOM_MATRICES,   /* double[sum 2 nc^2] (h - E s) dVol per contributing atom        */
    GFA_EXPORT_IMAGE_PHASE,     /* double[nai][2]     Bloch phase of each image                   */
    GFA_EXPORT_PHASE,           /* double[3][2][2]    kinetic halo phases                         */
    GFA_EXPORT_KIN_PREFACTOR,   /* double[3]                                                       */
    GFA_EXPORT_COUNT
};

/* ---- plan life cycle; replaces green_function::construct_Green_function (include/green_function.hxx:23-31)
 *      and the action_plan_t constructor/destructor (include/action_plan.hxx:84-95) ---- */
int gfa_plan_create(const gfa_plan_desc* desc, gfa_plan** plan);
int gfa_plan_destroy(gfa_plan* plan);
int gfa_plan_query(const gfa_plan* plan, gfa_sizes* sizes);
int gfa_plan_export(const gfa_plan* plan, int what, void* dst, uint64_t dst_bytes);

/* ---- per-system and per-(E,k) updates ----
 * gfa_set_potential replaces green_function::update_potential's potential part (include/green_function.hxx:33-40):
 *   V[ng_z][ng_y][ng_x] on the full grid, Hartree.
 * gfa_set_atom_matrices replaces its atom-matrix part plus the SHO prefactor rescaling of
 *   src/energy_contour.cxx:307-323: hs[2][nc][nc] = h then s of original atom ia in the L2-normalised
 *   SHO basis (zyx order), nc = nSHO(numax).
 * gfa_set_energy replaces green_function::update_energy_parameter (include/green_function.hxx:42-49).
 * gfa_set_kpoint replaces green_function::update_phases (include/green_function.hxx:51-56); k in units of 2 pi. */
int gfa_set_potential(gfa_plan* plan, const double* V);
int gfa_set_atom_matrices(gfa_plan* plan, int32_t ia, const double* hs, int32_t nc);
int gfa_set_energy(gfa_plan* plan, double E_re, double E_im, double dVol);
/* one energy per column group, E[nCols][2] (re, im): the batched form of update_energy_parameter for
 * instances of an energy contour (src/energy_contour.cxx:364-380 loops over the points one at a time) */
int gfa_set_energies(gfa_plan* plan, const double* E, double dVol);
int gfa_set_kpoint(gfa_plan* plan, const double k[3]);

/* ---- the action y = (H - E S) x on device pointers; replaces action_t::multiply
 *      (include/green_action.hxx:140-176).  stream is a hipStream_t (NULL: default stream).  The call is
 *      asynchronous on that stream.  *nflops (may be NULL) receives the reference's flop count.
 *      gfa_action_multiply has the reference's argument list: nnzb and nCols must match the plan; colIndex
 *      (device memory, as tfQMRgpu passes it) may be NULL, else it must hold the plan's colindx, which is
 *      checked with one blocking copy the first time a pointer is seen (GFA_ERR_ARGUMENT if it differs; a
 *      caller that rewrites the array behind the same pointer is not re-checked); l2nX is accepted and unused
 *      (it sizes tfQMRgpu's binary reductions); precondition must be 0 (has_preconditioner() is false). ---- */
int gfa_action(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, double* nflops);
int gfa_action_multiply(gfa_plan* plan, int fp_bits, void* y, const void* x, const uint16_t* colIndex,
                        uint32_t nnzb, uint32_t nCols, uint32_t l2nX, void* stream, int precondition,
                        double* nflops);

/* ---- the action without device allocation: no counterpart in the reference.  For callers that cannot
 *      allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 *      kernels and results as gfa_action, with the action's workspace (projection partial sums, atom
 *      coefficients, the zero block) in a fixed scratch array of 64 MiB that the library holds on every GPU it
 *      runs on until the process ends.  The plan must be on the device already (gfa_plan_prepare, or any earlier
 *      device call); a plan whose workspace exceeds 64 MiB is refused with GFA_ERR_ARGUMENT.  Calls that overlap
 *      on one GPU share the array and race: serialise them (one stream, or events). ---- */
int gfa_action_noalloc(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, double* nflops);

/* ---- the reference's action timing loop: action_t::solve with a negative iteration count and no tfQMRgpu
 *      (include/green_action.hxx:243-281, called by test_action at src/green_action.cxx:139-156).  Starting from
 *      x = 0 it applies y = A x `iterations` times, swapping x and y, with a device synchronisation after every
 *      call; *seconds receives the wall time of the loop and *nflops the reference's flop count of all calls.
 *      Allocates and frees its two vectors. ---- */
int gfa_action_benchmark(gfa_plan* plan, int fp_bits, int32_t iterations, double* seconds, double* nflops);

/* ---- one part of the action, for timing and tests: part 1 = SHO projection and atom-matrix multiply (into the
 *      plan's coefficient workspace), 2 = the local kernel (potential, kinetic and SHO add-back of the workspace
 *      coefficients, writes y), 3 = both (= gfa_action).  Part 2 alone uses the coefficients of the last part 1
 *      of the same precision.  Mirrors the reference's separate multiply of each term
 *      (green_potential / green_kinetic / green_dyadic::multiply, include/green_action.hxx:154-175). ---- */
int gfa_action_part(gfa_plan* plan, int fp_bits, int part, void* y, const void* x, void* stream);

/* ---- tile-shape autotuner: times every tile shape of the action on the caller's device vectors x, y (the
 *      layout of gfa_action; y is overwritten), reps applications each, and keeps the fastest for fp_bits.
 *      No counterpart in the reference, whose kernels have fixed launch shapes. ---- */
int gfa_autotune(gfa_plan* plan, int fp_bits, void* y, const void* x, void* stream, int reps, double* best_ms);
/* pins the tile shape instead: kind 0 (LDS accumulator): columns per work-group 16 or 32, kinetic lines per
 * thread in flight 1, 2 or 4; kind 1 (register accumulator per z plane): columns 32 or 64, lines ignored;
 * kind 2 (16-byte lanes, each thread owning one x-, y- and z-line): columns 16 or 32 (or 64 in FP32), lines = the
 * directions whose neighbour lines are loaded before the first stencil, 2 or 3 */
int gfa_set_tile_shape(gfa_plan* plan, int fp_bits, int kind, int columns, int lines);
/* the engine of the dyadic term's two contractions per precision, the SHO projection (c = P^T x over the blocks an
 * atom image touches) and the add-back (y += P d): 0 their sum-factorised forms on the vector ALUs, the add-back
 * inside the local kernel (the default); 1 their dense forms on the matrix cores (16x16x4 MFMA of the precision,
 * FP32 or FP64), the add-back in a kernel of its own after the local kernel.  Atoms with numax > 5 keep the vector
 * forms.  gfa_autotune times both with its fastest tile shape and keeps the faster.  No counterpart in the
 * reference. */
int gfa_set_dyadic_engine(gfa_plan* plan, int fp_bits, int engine);
int gfa_get_dyadic_engine(gfa_plan* plan, int fp_bits, int* engine);
/* the tile shape the action of fp_bits uses (the autotuner's choice, the pinned one or the default); *tuned = 1
 * when it was set by gfa_autotune or gfa_set_tile_shape */
int gfa_get_tile_shape(gfa_plan* plan, int fp_bits, int* kind, int* columns, int* lines, int* tuned);

/* ---- the solve (H - E S) X = B with B = unit columns on the diagonal blocks; replaces
 *      green_solver_t::solve / action_t::solve (include/green_solver.hxx:24-30, include/green_action.hxx:181-241)
 *      and tfqmrgpu::solve(action, buffer, 1e-9, maxiter, 0, true). ---- */
typedef struct {
    int32_t fp_bits;            /* 64: plain FP64 tfQMR; 32: FP32 tfQMR inside FP64 refinement        */
    int32_t max_iterations;     /* tfQMR iterations per (inner) solve                                 */
    double tolerance;           /* relative residual per column (1e-9)                                */
    int32_t max_refinements;    /* FP64 outer refinement steps for fp_bits 32 (<= 4 in the workloads) */
    int32_t check_every;        /* iterations between host convergence checks (0: 8)                  */
    uint64_t shadow_seed;       /* seed of the random shadow residual                                 */
    void* X_device;             /* optional: receives the FP64 solution [nnzb][2][64][64] (device)   */
    void* stream;               /* hipStream_t or NULL                                                */
    double inner_tolerance;     /* fp_bits 32: residual reduction asked of one FP32 solve (0: 1e-4)   */
    int32_t echo;               /* >0: one line per inner solve on stderr                             */
    int32_t eager;              /* 0: replay each iteration from a captured HIP graph; 1: plain launches */
    int32_t autotune;           /* 1: time the action's tile shapes on the solver's vectors first (once per precision) */
    int32_t fp32_max_iterations;/* fp_bits 32: half-steps allowed to one FP32 solve (0: max_iterations).  tfQMR:
                                 * after one FP32 solve an unconverged column group goes on in FP64 from its FP32
                                 * iterate, or from zero if the FP32 solve reduced its residual less than tenfold.
                                 * COCG: a group still unconverged when its FP32 solve has used this budget
                                 * continues in FP64 from the FP32 iterate                                        */
    double* residuals;          /* optional host output [64 nCols]: final relative residual of every column          */
    int32_t* halfsteps;         /* optional host output [64 nCols]: operator applications of every column (all solves) */
    int32_t method;             /* inner Krylov method: 0 tfQMR (as the reference), 1 COCG (conjugate orthogonal CG;
                                 * needs a complex-symmetric operator: real Bloch phases, e.g. Gamma or a wrapped
                                 * cell; one application per iteration, no squared residual polynomial, so its FP32
                                 * recurrences keep converging where tfQMR's stall)                                 */
    int32_t fp32_progress_check;/* fp_bits 32, COCG: after this many applications of the first FP32 solve (0: no
                                 * check), a column group whose recursive residual is still above tolerance^(1/3)
                                 * (less than a third of the way to the tolerance, in decades) leaves the FP32
                                 * solve and goes on in FP64 from its FP32 iterate, as at the FP32 budget           */
    const int32_t* group_max_iterations; /* optional [nCols]: a column group leaves the solve after this many
                                 * operator applications (0: no cap); with a tolerance far below reach every group
                                 * then stops at a fixed iterate (solves of nearby operators stop at the same
                                 * one: finite differences), while the others run on                         */
} gfa_solve_opts;

typedef struct {
    int32_t iterations;          /* tfQMR iterations done (sum over refinement steps)                */
    int32_t refinements;         /* FP64 refinement steps done                                       */
    int32_t converged_columns;   /* columns whose true FP64 residual met the tolerance               */
    int32_t total_columns;       /* 64 * nCols                                                       */
    double max_residual;         /* largest relative residual ||b - A x|| / ||b|| (FP64 action)      */
    double seconds;              /* wall time of the solve                                           */
    double flops;                /* action flops (reference counting)                                */
    int32_t actions;             /* number of action applications                                    */
    int32_t host_syncs;          /* host waits on the device (status polls, residual fetches, set-ups)     */
    int32_t graph_captures;      /* HIP graph captures of a solver step (one per change of the active set) */
    int32_t fp64_groups;         /* fp_bits 32: column groups that went on in FP64 (see fp32_max_iterations) */
} gfa_solve_stats;

/* rho[nCols][64][2]: Green function diagonal G_ii / pi (re, im) as include/green_action.hxx:226-238;
 * may be NULL.  The solver state (the FP64 solution, A x, the inner solver's work vectors) stays allocated with the
 * plan after the call, so that a following solve of the same plan allocates nothing; gfa_solver_release frees it. */
int gfa_solve(gfa_plan* plan, const gfa_solve_opts* opts, double* rho, gfa_solve_stats* stats);

/* ---- device memory of a solve.  gfa_solver_release frees the solver state of a plan (the reference frees the
 *      tfQMRgpu buffer with its action_t, include/green_action.hxx:118-123); the plan itself stays usable.
 *      gfa_solve_memory gives the device bytes of the solver state a solve with opts (fp_bits, method) holds at its
 *      largest -- for fp_bits 32 including the FP64 work vectors of a column group that falls back to FP64 -- and
 *      the bytes of the plan's device data (an estimate before the plan is first used on the device).
 *      gfa_plan_prepare uploads the plan and allocates the action's workspace of fp_bits, so that later actions
 *      of that precision allocate nothing (for timing, graph capture and gfa_action_noalloc). ---- */
int gfa_solver_release(gfa_plan* plan);
int gfa_solve_memory(gfa_plan* plan, const gfa_solve_opts* opts, uint64_t* solver_bytes, uint64_t* plan_bytes);
int gfa_plan_prepare(gfa_plan* plan, int fp_bits);

/* ---- the Green function of many source blocks in chunks: replaces the reference's distribution of the source
 *      columns over MPI ranks (src/green_function.cxx:255-354) by a loop over chunks of source blocks on one device.
 *      For the sources of desc (its box or explicit list, one truncation radius) at energy E and k-point k, every
 *      chunk is a plan of its own with V and the raw atom matrices hs (as gfa_contour_density takes them), solved
 *      with opts; the chunk size is the largest number of sources whose plan and solve state (gfa_solve_memory of
 *      the plan of all sources, per source) fit max_bytes (0: one chunk).  Column groups are independent, so the
 *      result equals the solve of one plan of all sources.  rho[nsrc][64][2] = G_ii / pi in source order. ---- */
typedef struct {
    int32_t sources;            /* source blocks of the description                                            */
    int32_t sources_per_chunk;
    int32_t chunks;
    int32_t converged_columns;
    int32_t total_columns;
    int32_t iterations;         /* the largest iteration count of a chunk                                      */
    double max_residual;
    double seconds;
    uint64_t largest_chunk_bytes; /* device memory of the largest chunk's plan and solve state                 */
} gfa_chunked_stats;

int gfa_solve_chunked(const gfa_plan_desc* desc, const double* V, const double* hs, double E_re, double E_im,
                      const double k[3], const gfa_solve_opts* opts, uint64_t max_bytes, double* rho,
                      gfa_chunked_stats* stats);

/* ---- adjoint gradient of L = Im sum_c w_c Tr[G_c], the trace over the diagonal block of column group c ----
 *  With dG = -A^-1 dA G:  dL = -Im sum_c w_c <lambda_c, dA G_c>,  lambda_c = (H - E_c S)^-H B_c.
 *  X: FP64 solution of (H - E_c S) X = B; Lambda: FP64 solution of the adjoint system (H - conj(E_c) S) Lambda = B
 *  (the same plan with gfa_set_energies(conj(E))), both device arrays in the plan layout.  E: the forward
 *  energies [nCols][2]; w: weights [nCols][2].  Outputs on the host (either may be NULL): grad_V[ng_z][ng_y][ng_x]
 *  = dL/dV(r); grad_hs = per original atom [2][nc][nc] concatenated in atom order, dL/dh and dL/ds of the matrices
 *  given to gfa_set_atom_matrices.  The V part is a fused conj(lambda) G reduction per block, folded per cell
 *  block in a fixed order; the matrix part contracts the projected coefficients of lambda and X.
 *  No counterpart in the reference (which has no derivatives). ---- */
int gfa_gradient(gfa_plan* plan, const void* X, const void* Lambda, const double* E, const double* w, double* grad_V,
                 double* grad_hs, void* stream);

/* ---- atomic density matrices <p_i|G|p_j>: replaces green_dyadic::get_projection_coefficients with SHOprj_right
 *      (include/green_dyadic.hxx:925-1091).  X: an FP64 Green function in the plan layout (device).  D receives, per
 *      contributing atom of the plan (gfa_plan_export ATOM_IDS order), the nc x nc matrix
 *      factor * Im(phase * sum_columns <p_i|G|s> p_j(s)), where the right projection runs over the points s of every
 *      source block of the plan; concatenated over atoms.  As in the reference, plans whose atoms have periodic
 *      images are rejected. ---- */
int gfa_atom_density_matrices(gfa_plan* plan, const void* X, double factor, double* D, void* stream);

/* ---- the energy contour around the solve (src/energy_contour.cxx) ----
 * gfa_gauss_legendre replaces Gauss_Legendre_quadrature (src/energy_contour.cxx:456): n points on [-1, 1] in
 *   ascending order, n rounded up as the reference's tables are (above 32 to a multiple of 4, above 64 to a
 *   multiple of 8); returns the number of points written (0 for n <= 0), or minus an error status beyond 112.
 * gfa_gauss_fermi_dirac replaces Gauss_Fermi_Dirac_quadrature (src/energy_contour.cxx:2156): the Gauss rule of
 *   the weight 1 / (1 + exp(30 x)) on [-1, infinity), n <= 16.
 * gfa_energy_mesh replaces get_energy_mesh (src/energy_contour.cxx:139-241): the contour points relative to the
 *   Fermi level and their complex weights, E[nE][2] and w[nE][2]; with E or w NULL only *nE is returned. */
int gfa_gauss_legendre(int32_t n, double* x, double* w);
int gfa_gauss_fermi_dirac(int32_t n, double* x, double* w);
int gfa_energy_mesh(double kBT, double eBot, int32_t nBot, int32_t nPar, int32_t nFer, int32_t nPol, double* E,
                    double* w, int32_t capacity, int32_t* nE);

typedef struct {
    const double* E;            /* [nE][2] energy points relative to the Fermi level (gfa_energy_mesh)          */
    const double* wE;           /* [nE][2] complex weights                                                     */
    int32_t nE;
    int32_t nk;                 /* k-points                                                                    */
    const double* k;            /* [nk][4]: k (units of 2 pi) and its weight (brillouin_zone::WEIGHT)          */
    double E_Fermi;             /* added to every energy point                                                 */
    double dVol;                /* grid volume element (0: hx hy hz)                                           */
    int32_t max_groups;         /* column groups per batched solve (0: 64); energies are batched per solve     */
    int32_t reserved;
    gfa_solve_opts solve;       /* options of every batched solve (fp_bits, tolerance, method, ...)            */
} gfa_contour_opts;

typedef struct {
    int32_t solves;             /* batched solves run (energy batches x k-points)                              */
    int32_t sources;            /* source blocks of the description                                            */
    int32_t converged_columns;  /* columns that met the tolerance, over all solves                             */
    int32_t total_columns;
    double max_residual;
    double seconds;
} gfa_contour_stats;

/* ---- contour integration: replaces energy_contour::Integrator::integrate (src/energy_contour.cxx:273-451) up to
 *      the 4x4x4 density.  For the sources of desc (one column group each, one truncation radius), the potential
 *      V[ng_z][ng_y][ng_x] and the raw atom matrices hs (per atom [2][nc][nc] in atom order, as
 *      gfa_set_atom_matrices takes them), it solves every (energy, k-point, source) instance, energies batched as
 *      column groups of one plan, and returns rho[nsources][64] = Re sum_E w_E sum_k w_k G_ii / pi and, if not NULL,
 *      the response density rho_response[nsources][64] of the reference (-sum over all but the last two energy
 *      points, divided by the same sum of their energies).  The 8x8x8 interpolation and the Fermi-level update
 *      stay with the caller, as in the reference. ---- */
int gfa_contour_density(const gfa_plan_desc* desc, const double* V, const double* hs, const gfa_contour_opts* opts,
                        double* rho, double* rho_response, gfa_contour_stats* stats);

/* text of the last error of the calling thread ("" if none) */
const char* gfa_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GFA_H */

/* tahoe_gpu.h -- C ABI of the MI355X (gfx950) continuum-solid Newton step.
 *
 * One handle (tg_ctx) holds one finite-element problem resident on one GPU:
 * mesh, equation numbering, the upper-triangle MSR stiffness, the solver's
 * block operator and all Newton/PCG vectors.  Every call on a handle is
 * ordered on the handle's own HIP stream; calls on one handle are not
 * thread-safe, different handles may be driven from different threads.
 *
 * Host arrays are plain C arrays; all sizes are element counts.
 * Node-major arrays are [nnd][3]; equation-space arrays are [n_eq] in the
 * reference's global equation order.  Connectivity is 0-based.
 *
 * Every function returns a status equal to the reference's ExceptionT::CodeT
 * (toolbox/src/common/ExceptionT.h:17-30): 0 ok, 1 general fail,
 * 3 out of memory, 4 out of range, 5 size mismatch, 6 bad input value,
 * 7 bad Jacobian determinant.  tg_last_error() gives the message.
 *
 * The reference interface each entry point replaces is cited beside it
 * (paths relative to the Tahoe source tree).
 */
#ifndef TAHOE_GPU_H
#define TAHOE_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TG_OK = 0,
    TG_GENERAL_FAIL = 1,
    TG_OUT_OF_MEMORY = 3,
    TG_OUT_OF_RANGE = 4,
    TG_SIZE_MISMATCH = 5,
    TG_BAD_INPUT = 6,
    TG_BAD_JACOBIAN = 7
};

/* element formulations: XML names total_lagrangian / updated_lagrangian_Q1P0
 * (tahoe/src/elements/ElementListT.cpp:744-745) */
enum { TG_TOTAL_LAGRANGIAN = 0, TG_Q1P0 = 1 };

/* arithmetic of the stiffness and the inner linear solve:
 * TG_FP64  everything in FP64;
 * TG_MIXED element stiffness assembled in FP32 and the PCG operator stored
 *          in FP32 (half the SpMV bytes); Krylov vectors, SpMV sums and dot
 *          products FP64; residual, Newton test and update FP64, so Newton
 *          acts as the iterative refinement of the FP32 operator. */
enum { TG_FP64 = 0, TG_MIXED = 1 };

typedef struct tg_ctx tg_ctx;

typedef struct {
    int iterations;      /* PCG iterations taken */
    double rel_residual; /* ||b - A x|| / ||b|| (recurrence) */
    int converged;
} tg_lin_stats;

/* NLSolver parameters (tahoe/src/solvers/NLSolver.cpp:675-757) plus the
 * Jacobi-PCG tolerance that replaces the direct factorisation. */
typedef struct {
    double rel_tolerance;        /* rel_tolerance */
    double abs_tolerance;        /* abs_tolerance */
    double divergence_tolerance; /* divergence_tolerance */
    int max_iterations;          /* max_iterations */
    double lin_rtol;             /* PCG ||r||/||b|| */
    int lin_maxit;
} tg_newton_opts;

typedef struct {
    int status;           /* 0 converged, 1 failed (diverged / iteration limit), 7 bad Jacobian */
    int iterations;       /* Newton corrections */
    double error0, error; /* ||R|| before the first and after the last correction */
    long lin_iterations;  /* PCG iterations summed over corrections */
} tg_newton_stats;

/* per-kernel-class device time, summed over launches since tg_timing_reset */
typedef struct {
    double residual_ms;  long residual_launches;
    double stiffness_ms; long stiffness_launches;
    double expand_ms;    long expand_launches;
    double spmv_ms;      long spmv_launches;
    double vector_ms;    long vector_launches;
    double precond_ms;   long precond_launches;   /* multigrid V-cycles (TG_PC_AMG) */
} tg_timing;

/* ---- handle ---------------------------------------------------------- */
int tg_create(tg_ctx** ctx, int device);
int tg_destroy(tg_ctx* ctx);
const char* tg_last_error(const tg_ctx* ctx);

/* ---- problem definition ---------------------------------------------- */
/* Mesh and element-block materials.  Replaces reading the geometry and
 * ContinuumElementT's element cards (ContinuumElementT.cpp:107). */
int tg_set_mesh(tg_ctx* ctx, int nnd, int nel, const double* X, const int* conn, const int* mat);
/* Simo_isotropic moduli for material id (IsotropicT.cpp:32-45 gives mu,kappa). */
int tg_set_material(tg_ctx* ctx, int id, double mu, double kappa);
/* TG_TOTAL_LAGRANGIAN or TG_Q1P0.  The reference's updated_lagrangian
 * element (UpdatedLagrangianT.cpp:83-92: shape functions of the current
 * configuration, weights det(dx/dxi)) integrates the same forms as
 * TotalLagrangianT (det(dx/dxi) = J det(dX/dxi)) and maps to
 * TG_TOTAL_LAGRANGIAN. */
int tg_set_formulation(tg_ctx* ctx, int formulation);
/* Material model of every material of the context, whose moduli come from
 * tg_set_material: TG_MAT_SIMO_ISOTROPIC (default, Simo_isotropic,
 * SimoIso3D.cpp:16-52) or TG_MAT_STVENANT (large_strain_StVenant: FDKStV
 * with FDHookeanMatT::s_ij/c_ijkl, FDHookeanMatT.cpp, and
 * IsotropicT::ComputeModuli, IsotropicT.cpp:153-169; total-Lagrangian element
 * only). */
enum { TG_MAT_SIMO_ISOTROPIC = 0, TG_MAT_STVENANT = 1 };
int tg_set_material_model(tg_ctx* ctx, int model);
/* TG_FP64 (default) or TG_MIXED; no counterpart in the reference, which is
 * FP64 only.  TG_MIXED: the total-Lagrangian stiffness is assembled in
 * FP32; the Q1P0 forward-difference stiffness is formed in FP64 and
 * narrowed to the FP32 inner operator. */
int tg_set_precision(tg_ctx* ctx, int precision);
/* Preconditioner of the inner PCG: TG_PC_JACOBI (default) is the
 * reference's DiagonalMatrixT (1/diag, DiagonalMatrixT.cpp:267-324);
 * TG_PC_BLOCK_JACOBI inverts each node's 3x3 diagonal block; TG_PC_AMG is
 * one V-cycle of aggregation multigrid (rigid-body modes, Chebyshev
 * smoothing, Galerkin coarse operators re-formed with every tangent) for a
 * single problem (no segments).  The last two have no counterpart in the
 * reference.  Any may be set before or after tg_initialize; the multigrid
 * structure is built on the host when first needed. */
enum { TG_PC_JACOBI = 0, TG_PC_BLOCK_JACOBI = 1, TG_PC_AMG = 2 };
int tg_set_preconditioner(tg_ctx* ctx, int pc);
/* Linear solver inside tg_newton, tg_newton_correction and tg_solve:
 * TG_LIN_PCG (default) is the segmented preconditioned CG; TG_LIN_GMRES is
 * the newton_krylov_solver's restarted GMRES(restart) with left
 * preconditioning (NLSolver_NK::GMRES, NLSolver_NK.cpp:121-238): a restart
 * cycle stops when the preconditioned residual falls below lin_rtol times
 * its value at the start of that cycle, and lin_maxit <= 0 means 3 n_eq
 * Arnoldi steps (max_linear_iter = 0).  GMRES needs a single problem (no
 * segments). */
enum { TG_LIN_PCG = 0, TG_LIN_GMRES = 1 };
int tg_set_linear_solver(tg_ctx* ctx, int solver, int restart);
/* Kinematic BC cards: prescribed value = values[k] * t (KBC_CardT::Value,
 * tahoe/src/primitives/KBC_CardT.cpp:53-64, linear schedule). */
int tg_set_kbc(tg_ctx* ctx, int nkbc, const int* nodes, const int* dofs, const double* values);
/* Nodal force cards (force_BC, FBC_CardT): the residual gains
 * values[k] * scale at (nodes[k], dofs[k]) for active dofs (FieldT::FormRHS,
 * FieldT.cpp:390-401, with FBC_CardT::CurrentValue, FBC_CardT.cpp:57-63:
 * scale = the card's schedule value, set by tg_apply_forces; 0 until then). */
int tg_set_forces(tg_ctx* ctx, int nforce, const int* nodes, const int* dofs, const double* values);
int tg_apply_forces(tg_ctx* ctx, double scale);

/* Batched problems: the mesh of tg_set_mesh is the disjoint union of nseg
 * independent sub-meshes; sub-mesh g owns nodes [node_begin[g],
 * node_begin[g+1]) and the elements on them.  node_begin[0] = 0,
 * node_begin[nseg] = nnd, and every other begin is a multiple of 64 (pad
 * with nodes no element uses, all three dofs carrying kinematic cards).
 * Each sub-problem gets its own Newton and PCG control; its arithmetic, and
 * so its result, is the same alone and in any batch.  No counterpart in the
 * reference (one problem per run); for callers that solve many small
 * problems (FE^2 micro-scale RVEs). */
int tg_set_segments(tg_ctx* ctx, int nseg, const int* node_begin);

/* Equation numbering, MSR pattern, element colouring and device storage.
 * Replaces FEManagerT::SetEquationSystem (FEManagerT.cpp:2248-2313) ->
 * NodeManagerT::SetEquationNumbers (NodeManagerT.cpp:712-760) and
 * GlobalMatrixT::Initialize -> MSRMatrixT::Initialize (MSRMatrixT.cpp:35-44).
 * The equation numbers and the colouring are formed on the host; the node
 * graph, the MSR pattern and the solver operator's structure on the device
 * (so tg_get_msr_bindx copies from the device).  A node shared by more than
 * 64 elements is refused (TG_OUT_OF_RANGE).  With the environment variable
 * TG_SETUP_TRACE set, the time of each setup phase is printed to stderr. */
int tg_initialize(tg_ctx* ctx, int* n_eq, long* msr_length);
/* FieldT::Equations(): [nnd][3], 1-based, -1 prescribed (print_eqnos). */
int tg_get_eqnos(const tg_ctx* ctx, int* eqnos);
/* Aztec bindx of the upper MSR (MSRBuilderT::GenerateMSR, MSRBuilderT.cpp:134-181). */
int tg_get_msr_bindx(const tg_ctx* ctx, int* bindx);

/* ---- state ------------------------------------------------------------ */
int tg_set_displacement(tg_ctx* ctx, const double* u);
int tg_get_displacement(tg_ctx* ctx, double* u);
/* Prescribed displacements at time t (nStaticIntegrator::ConsistentKBC,
 * tahoe/src/integrators/static/nStaticIntegrator.cpp:18-46). */
int tg_apply_kbc(tg_ctx* ctx, double t);
/* Q1P0 converged-state history (SimoQ1P0::CloseStep/ResetStep, SimoQ1P0.cpp:42-61). */
int tg_close_step(tg_ctx* ctx);
int tg_reset_step(tg_ctx* ctx);

/* ---- element group ---------------------------------------------------- */
/* Residual R = -f_int of the current displacement (ElementBaseT::FormRHS,
 * ElementBaseT.cpp:147 -> SolidElementT::ElementRHSDriver SolidElementT.cpp:1166,
 * assembled as SolverT::AssembleRHS SolverT.cpp:446-477).  R_out [n_eq] and
 * norm_out (SolverT::Residual, SolverT.cpp:848) may be NULL. */
int tg_form_rhs(tg_ctx* ctx, double* R_out, double* norm_out);
/* Stiffness into the device MSR (ElementBaseT::FormLHS ElementBaseT.cpp:114 ->
 * SolidElementT::ElementLHSDriver SolidElementT.cpp:1100 -> MSRMatrixT::Assemble
 * MSRMatrixT.cpp:66-165, after GlobalMatrixT::Clear). */
int tg_form_lhs(tg_ctx* ctx);
/* The tangent half of a Newton correction: K(u) into the inner solver's
 * operator (through the MSR or straight into the block operator, as
 * TG_TUNE_NEWTON_ASSEMBLY selects) and the preconditioner's data for it --
 * NLSolver::Solve's fFEManager.FormLHS (NLSolver.cpp:155) followed by the
 * factorisation half of fLHS->Solve (NLSolver.cpp:762 -> GlobalMatrixT::
 * Factorize, DiagonalMatrixT::Factorize DiagonalMatrixT.cpp:267-324 for
 * Jacobi).  A following tg_newton_correction forms its own tangent again. */
int tg_form_tangent(tg_ctx* ctx);

/* ---- global matrix ---------------------------------------------------- */
/* MSR values val[msr_length] (the data GenerateRCV/PrintLHS dump, MSRMatrixT.cpp:449-524). */
int tg_get_msr_values(tg_ctx* ctx, double* val);
/* b = K x in equation space (GlobalMatrixT::Multx, MSRMatrixT.cpp:385-420). */
int tg_multx(tg_ctx* ctx, const double* x, double* b);
/* MSRMatrixT::CopyDiagonal (MSRMatrixT.cpp:351-378). */
int tg_copy_diagonal(tg_ctx* ctx, double* diag);
/* Solve K x = b by Jacobi-PCG; rhs holds b on entry and x on exit
 * (GlobalMatrixT::Solve, GlobalMatrixT.cpp:77-114). */
int tg_solve(tg_ctx* ctx, double* rhs, double rtol, int maxit, tg_lin_stats* stats);

/* tg_solve that allocates nothing -- no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  b and x are DEVICE
 * arrays [n_eq] in equation space (x is written, b only read); the matrix is
 * the context's MSR from the last tg_form_lhs, as for tg_solve.  The PCG's
 * vectors, reduction partials, scalar slots and done count live in one
 * static device array of TG_NOALLOC_SCRATCH_BYTES per GPU, kept until the
 * process ends; a problem needing more (about 144 bytes per node) is
 * rejected with TG_SIZE_MISMATCH.  Supports TG_LIN_PCG with the Jacobi and
 * block-Jacobi preconditioners (TG_PC_AMG keeps its own buffers: rejected,
 * TG_BAD_INPUT); every launch is direct (no graph), and no event timing is
 * recorded during the call.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (two threads, two contexts) race and must
 * be serialised by the caller.  Returns after the result is in x. */
#define TG_NOALLOC_SCRATCH_BYTES (64L << 20)
int tg_solve_noalloc(tg_ctx* ctx, const double* b_dev, double* x_dev, double rtol, int maxit, tg_lin_stats* stats);

/* ---- solver ----------------------------------------------------------- */
/* One load step's Newton loop, device resident: R, ||R||, K, PCG, u += du
 * until ExitIteration's tests hold (NLSolver::Solve, NLSolver.cpp:57-264). */
int tg_newton(tg_ctx* ctx, const tg_newton_opts* opts, tg_newton_stats* stats);
/* tg_newton for every segment at once: per-segment ||R||, ExitIteration
 * tests and PCG; converged or failed segments stop updating while the rest
 * continue.  stats[nseg] receives each segment's outcome: status 0
 * converged, 1 diverged or out of iterations, 7 an element of that problem
 * inverted in a residual pass (the others carry on; the call itself then
 * returns 0).  A failed segment's displacement goes back to its last
 * converged state (tg_close_step) so that the element passes over the rest
 * see no inverted element; an inverted element in a stiffness pass still
 * fails the whole call with 7. */
int tg_newton_batch(tg_ctx* ctx, const tg_newton_opts* opts, tg_newton_stats* stats);
/* One Newton correction without the convergence loop: R, K, PCG, u += du.
 * stats->error0 gets ||R|| before the update (used by benchmarks). */
int tg_newton_correction(tg_ctx* ctx, const tg_newton_opts* opts, tg_newton_stats* stats);

/* ---- instrumentation -------------------------------------------------- */
int tg_synchronize(tg_ctx* ctx);
int tg_timing_enable(tg_ctx* ctx, int enable);
int tg_timing_reset(tg_ctx* ctx);
int tg_timing_get(tg_ctx* ctx, tg_timing* out);
/* plain device-to-device copy kernel (k_copy_calib) over `bytes` bytes, run
 * `reps` times; *ms gets the mean kernel time.  Calibrates the HBM counters
 * of a profiler against a kernel of known traffic (bytes read + written). */
int tg_calib_copy(int device, long bytes, int reps, double* ms);

/* ---- tuning ------------------------------------------------------------ */
/* Launch-shape knobs (no counterpart in the reference; set by the
 * autotuner in tahoe_amd/autotune.py or by hand):
 * TG_TUNE_STIFF_CHUNK  elements per range of the colour-ordered stiffness
 *                      launches (0: one launch per colour, -1: derived from
 *                      the matrix footprint);
 * TG_TUNE_STIFF_SHAPE  block shape of the stiffness kernel: 0..3 for 16,
 *                      12, 8 or 4 elements per workgroup with element-major
 *                      lanes, 4..5 for 16 or 8 with block-major lanes,
 *                      6..8 for 16, 8 or 4 in the D*B form (D_q B shared
 *                      by an element's lanes through LDS), 9..11 for 12, 8
 *                      or 4 with the MSR read-modify-write loads issued
 *                      before the block products, 12..14 for 12, 8 or 4
 *                      and 15 for 16 block-major with the additions issued
 *                      as memory-side atomics (order per entry unchanged:
 *                      colours); 16..17 for 8 or 4 elements per wave on the
 *                      matrix cores (16x16x4 MFMA tiles, FP32 in TG_MIXED,
 *                      FP64 otherwise; MSR assembly -- the straight-into-
 *                      operator Newton path runs shape 9 instead), 18 as 16
 *                      with registers bounded for 3 waves per SIMD; default
 *                      17 for the FP32 (TG_MIXED) stiffness, 9 for FP64;
 * TG_TUNE_NEWTON_ASSEMBLY  where tg_newton / tg_newton_batch /
 *                      tg_newton_correction put the tangent: 0 (default)
 *                      into the MSR and then expanded into the inner solve's
 *                      block operator, 1 straight into the block operator.
 *                      The operator is the same either way (same sums in
 *                      the same order); with 1 the MSR queries
 *                      (tg_get_msr_values, tg_multx, tg_copy_diagonal) see
 *                      the matrix of the last tg_form_lhs. */
/* TG_TUNE_PCG_GRAPH    1 (default): the PCG's 16-iteration chunk is captured
 *                      once as a HIP graph and replayed (one launch instead
 *                      of 80, for small, launch-bound problems); 0: direct
 *                      launches.  Per-kernel timing (tg_timing_enable)
 *                      always uses direct launches. */
/* TG_TUNE_SPMV_XCD     PCG SpMV variant, bit 0: each XCD (own L2) takes one
 *                      contiguous eighth of the rows instead of the hardware
 *                      order; bit 1: branch-free inner loop unrolled by 3
 *                      (padding slots multiply zeros); bit 2: each lower
 *                      block is read as the transpose of its upper partner
 *                      (the same values: the operator is symmetric), so HBM
 *                      streams about half the operator and the partner rows'
 *                      recent reads serve the rest from cache.  Every
 *                      variant gives bitwise the same product.  0 .. 7, or
 *                      -1 (default): 7 for a single problem of 2^17 nodes
 *                      or more, 2 for a batch of 2^17 nodes or more, else
 *                      0. */
/* TG_TUNE_AMG_DEGREE   multigrid: Chebyshev degree of each pre- and
 *                      post-smoothing pass, 2 (default) .. 4. */
/* TG_TUNE_AMG_WEIGHT   multigrid: weight of the coarse-grid correction in
 *                      percent, 100 .. 190 (default 150: unsmoothed
 *                      aggregation under-corrects; any weight in (0, 2)
 *                      keeps the V-cycle symmetric positive definite). */
/* TG_TUNE_PCG_FUSED    1: the direction update is formed inside the SpMV
 *                      (four kernels per PCG iteration instead of five, no
 *                      separate direction pass) for the Jacobi and
 *                      multigrid preconditioners, in any SpMV variant;
 *                      0: the five-kernel iteration; -1 (default): 0 for a
 *                      problem or batch of 2^17 nodes or more, else 1. */
/* TG_TUNE_STIFF_GATHER 0 (default): the total-Lagrangian stiffness is
 *                      added into the MSR element by element (colour-
 *                      ordered read-modify-write through the slot map);
 *                      1: each element's 300 upper entries are written to
 *                      an element buffer in launch order, and one gather
 *                      pass sums every MSR value's contributions in that
 *                      order and writes it once.  Same sums in the same
 *                      order: the values are bitwise equal.  Costs a
 *                      buffer of 300 values per element and a list of
 *                      contributions (4 bytes each, 300 per element) built
 *                      on first use. */
/* TG_TUNE_PCG_PERSIST  segments of at most this many slices (64 nodes
 *                      each) run their whole PCG inside one persistent
 *                      workgroup: no launch per iteration, the segment's
 *                      vectors and operator kept in the CU's caches; larger
 *                      segments use the kernel-per-step iteration.  The
 *                      iterates are bitwise the same either way.  0 off,
 *                      at most 16 (1024 nodes); -1 (default): 12. */
/* TG_TUNE_AMG_REUSE    multigrid: one numeric set-up of the coarse levels
 *                      (Galerkin products, smoother intervals, coarsest
 *                      inverse) serves up to this many further tangents of
 *                      the same Newton solve; the fine level always smooths
 *                      with the current operator.  tg_apply_kbc,
 *                      tg_set_displacement, tg_close_step and tg_reset_step
 *                      start a new solve (the next tangent sets the coarse
 *                      levels up again).  0: a new set-up for every
 *                      tangent; default 6. */
enum { TG_TUNE_STIFF_CHUNK = 1, TG_TUNE_STIFF_SHAPE = 2, TG_TUNE_NEWTON_ASSEMBLY = 3, TG_TUNE_PCG_GRAPH = 4,
       TG_TUNE_SPMV_XCD = 5, TG_TUNE_AMG_DEGREE = 6, TG_TUNE_AMG_WEIGHT = 7, TG_TUNE_PCG_FUSED = 8,
       TG_TUNE_STIFF_GATHER = 9, TG_TUNE_PCG_PERSIST = 10, TG_TUNE_AMG_REUSE = 11 };
int tg_set_tuning(tg_ctx* ctx, int knob, long value);
int tg_get_tuning(tg_ctx* ctx, int knob, long* value);

#ifdef __cplusplus
}
#endif
#endif /* TAHOE_GPU_H */

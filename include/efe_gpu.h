/*
 * efe_gpu.h -- C ABI of the MI355X (gfx950) implementation of EdelweissFE's
 * nonlinear implicit static (NIST) Newton step.
 *
 * Every pointer marked "device" is a device address on the context's GPU
 * (for instance a torch tensor's data_ptr()); every pointer marked "host" is
 * ordinary host memory.  All work is enqueued on the context's stream; calls
 * that return numbers to the host (norms, status, iteration counts)
 * synchronise that stream.  A context is not thread safe, like the
 * single-threaded NIST solver it serves.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   element loop     edelweissfe/solvers/nonlinearimplicitstatic.py:794-849 computeElements,
 *                    edelweissfe/solvers/nonlinearimplicitstaticparallelmk2.pyx:112-198,
 *                    native element ABI edelweissfe/elements/marmotelement/element.pxd:97-103
 *   material update  edelweissfe/materials/base/basehypoelasticmaterial.py:133-154 computeStress
 *   DOF/VIJ/CSR      edelweissfe/numerics/dofmanager.py:445-557, edelweissfe/numerics/csrgenerator.pyx:47-115
 *   Dirichlet on K   edelweissfe/solvers/nonlinearimplicitstatic.py:559-593 applyDirichletK
 *   linear solve     edelweissfe/solvers/nonlinearimplicitstatic.py:727-751 linearSolve,
 *                    edelweissfe/config/linsolve.py:46-86 getLinSolverByName
 *   convergence      edelweissfe/solvers/nonlinearimplicitstatic.py:625-725,771-792
 *   state accept     edelweissfe/models/femodel.py:245-258, elements/displacementelement/element.py:373-379
 */
#ifndef EFE_GPU_H
#define EFE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define EFE_OK 0
#define EFE_CUTBACK 1          /* a material asked for a smaller increment (CutbackRequest, vonmises.py:230-231) */
#define EFE_NAN 2              /* NaN in a result (DivergingSolution, nonlinearimplicitstatic.py:748-749) */
#define EFE_NOT_CONVERGED 3    /* Krylov solver hit its iteration limit */
#define EFE_ERR_ARG -1
#define EFE_ERR_HIP -2
#define EFE_ERR_STATE -3

/* element types (elements/library.py:212-291) */
#define EFE_C3D8 0    /* hexa8, 2x2x2 Gauss */
#define EFE_C3D20 1   /* hexa20, 3x3x3 Gauss */
#define EFE_C3D20R 2  /* hexa20, 2x2x2 Gauss */

/* materials */
#define EFE_LINEAR_ELASTIC 0 /* props [E, nu]                          linearelastic.py:59-61 */
#define EFE_VONMISES 1       /* props [E, nu, fy0, HLin, dfy, delta]   vonmises.py:105-113 */
#define EFE_DRUCKER_PRAGER 2 /* props [E, nu, phi_deg, psi_deg, c0, H] (no reference counterpart) */

/* precision of the stiffness contractions, the stored K and the Krylov solve */
#define EFE_FP64 0
#define EFE_FP32 1
/* FP32 matrix-free operator: the element pass stores per-QP J^-1 and the compact tangent
 * (FP32) and only the diagonal blocks; K p is applied element by element.  CG only. */
#define EFE_FP32_MATFREE 2

#define EFE_SOLVER_CG 0
#define EFE_SOLVER_GMRES 1
#define EFE_SOLVER_BICGSTAB 2  /* scalar CSR solves (efe_csr_linsolve*) only */
#define EFE_PRECOND_BLOCK_JACOBI 0
#define EFE_PRECOND_ILU0 1

typedef struct efe_ctx efe_ctx;

typedef struct {
    int32_t n_nodes;
    int32_t n_elements;
    int32_t element_type;       /* EFE_C3D8 | EFE_C3D20 | EFE_C3D20R */
    int32_t material;           /* EFE_LINEAR_ELASTIC | EFE_VONMISES | EFE_DRUCKER_PRAGER */
    const double* coords;       /* host [n_nodes][3] */
    const int32_t* conn;        /* host [n_elements][nodes_per_element], node positions, boxGen node order */
    int32_t n_props;            /* 2 (linear elastic) or 6 */
    int32_t props_per_element;  /* 0: props is [n_props]; 1: props is [n_elements][n_props] */
    const double* props;        /* host */
    const int32_t* specimen_of_node; /* host [n_nodes] or NULL: independent specimens (block-diagonal system) */
    int32_t n_specimens;        /* 1 when specimen_of_node is NULL */
    void* stream;               /* hipStream_t or NULL for the default stream */
    /* Several sections (sections/base/sectionbase.py:86-128 assigns a material to each element
     * set): host [n_elements] material of every element, or NULL (every element `material`).
     * When given, `material` is ignored, props must be per element with n_props = 6 (linear
     * elastic reads the first 2), and the element pass runs once per material present. */
    const int32_t* material_of_element;
} efe_mesh_desc;

typedef struct {
    int32_t solver;            /* EFE_SOLVER_CG | EFE_SOLVER_GMRES */
    int32_t precond;           /* EFE_PRECOND_BLOCK_JACOBI | EFE_PRECOND_ILU0 */
    int32_t precision;         /* EFE_FP64 | EFE_FP32 (must match the precision K was assembled in) */
    int32_t max_iter;
    int32_t restart;           /* GMRES restart length */
    int32_t check_every;       /* Krylov iterations between convergence readbacks */
    double rtol;               /* ||r|| <= rtol ||b|| */
    double atol;
} efe_solve_opts;

typedef struct {
    int32_t iterations;
    int32_t status;            /* EFE_OK | EFE_NOT_CONVERGED | EFE_NAN */
    double residual;           /* final ||r||_2 of the solver's own recursion */
    double rhs_norm;
} efe_solve_info;

/* last error text of the calling thread */
const char* efe_last_error(void);

/* Builds the DOF map (3 p + c), the bit-exact CSR pattern of csrgenerator.pyx:68-77,
 * the element->CSR gather map, and allocates the QP state (dofmanager.py:153-218,
 * csrgenerator.pyx:47-98, element.py:207-240). */
int efe_ctx_create(const efe_mesh_desc* desc, efe_ctx** out);
void efe_ctx_destroy(efe_ctx* ctx);

/* sizes: [n_dof, csr_nnz, n_qp_per_element, n_state_per_qp] (host out) */
int efe_get_sizes(efe_ctx* ctx, int64_t out[4]);

/* CSR pattern, identical to SciPy's canonical csr_matrix((0,(I,J))) (csrgenerator.pyx:68-77):
 * indptr host [n_dof+1], indices host [csr_nnz]. */
int efe_get_csr_pattern(efe_ctx* ctx, int32_t* indptr, int32_t* indices);

/* Marks the constrained DOFs of the current step (nonlinearimplicitstatic.py:1077-1084).
 * dofs host [n]; replaces the previous set.  Subsequent assemblies eliminate these
 * rows and columns symmetrically with a unit diagonal. */
int efe_set_dirichlet(efe_ctx* ctx, const int32_t* dofs, int32_t n);

/* One computeElements pass (nonlinearimplicitstatic.py:794-849) fused with
 * updateCSR (csrgenerator.pyx:100-115):
 *   for every element: dEps = B dU per QP, material update of the reference state
 *   into the temporary state, Ke = sum B^T C B detJ w, Pe = -sum B^T sigma detJ w;
 *   P = sum Pe, F = sum |Pe| (device [n_dof], overwritten);
 *   K (context-owned) = deterministic sum of Ke in element order, in `precision`,
 *   with the Dirichlet DOFs eliminated.
 * dir_values (device [n_dof] or NULL): prescribed corrections on the Dirichlet DOFs;
 *   when given, dir_correction (device [n_dof]) receives K_raw[:,dir] * dir_values on
 *   free rows and 0 on Dirichlet rows, so that b = R - dir_correction is the
 *   right-hand side of the eliminated system.
 * time host [2] (step time, total time) and dT mirror computeYourself's arguments.
 * Returns EFE_OK or EFE_CUTBACK (then *pNewDT = 0.5, vonmises.py:231). */
int efe_compute_elements(efe_ctx* ctx, const double* U, const double* dU, const double* time, double dT,
                         double* P, double* F, const double* dir_values, double* dir_correction,
                         int32_t precision, double* pNewDT);

/* The residual half of efe_compute_elements (nonlinearimplicitstatic.py:794-849 without the
 * tangent): the same material update of the reference state into the temporary state and
 * the same P and F, bit for bit, but no element tangents, operator or block-Jacobi inverse.
 * The operator stays that of the last efe_compute_elements.  For a convergence check that
 * needs no solve when it passes: the Newton iteration 0 of an increment that starts from an
 * extrapolated (or zero) increment, where the reference computes the tangent and drops it.
 * precision picks the element pass's launch shape (that of the following efe_compute_elements);
 * the residual is FP64 and the same in every precision.
 * Returns EFE_OK or EFE_CUTBACK (then *pNewDT = 0.5). */
int efe_compute_residual(efe_ctx* ctx, const double* U, const double* dU, const double time[2], double dT,
                         double* P, double* F, int32_t precision, double* pNewDT);

/* External body-force vector of a force per volume force3 (host [3]) on the elements
 * elements[i] (host [n]; NULL: all elements), added into PExt (device [n_dof]):
 * NIST.computeBodyForces (nonlinearimplicitstatic.py:510-557) over bForce.elementSet with the
 * element's computeBodyForce (element.py:348-371).  An element listed twice counts once. */
int efe_body_force(efe_ctx* ctx, const int32_t* elements, int32_t n, const double* force3, double* PExt);

/* Uniform surface traction traction3 (host [3], force per area) on the faces faces[i]
 * (1..6, Abaqus numbering by local node number, as the box generator's surfaces,
 * boxgen.py:438-446) of the elements elements[i] (host [n]): the consistent nodal forces
 * integral N_a dA * traction3 are added into PExt (device [n_dof]).  This is the element side of
 * NIST.computeDistributedLoads (nonlinearimplicitstatic.py:460-508) for load type "surface
 * traction" (stepactions/distributedload.py); the reference's own DisplacementElement has
 * no computeDistributedLoad (element.py:288). */
int efe_surface_traction(efe_ctx* ctx, const int32_t* elements, const int32_t* faces, int32_t n,
                         const double* traction3, double* PExt);

/* Geostatic initial stress (stepactions/geostatic.py:94-101 -> the element's
 * setInitialCondition("geostatic stress", [p1, h1, p2, h2, xLateral, zLateral])):
 * at every QP of the listed elements (host [n], or NULL for all elements) the reference
 * state's sigma_yy = p1 + (p2 - p1) (y - h1) / (h2 - h1) at the QP's y coordinate,
 * sigma_xx = xLateral sigma_yy, sigma_zz = zLateral sigma_yy; shear stresses and the
 * remaining state are kept.  The NIST solver applies it at every iteration start while
 * the geostatic step action is active (nonlinearimplicitstatic.py:413-414). */
int efe_set_geostatic(efe_ctx* ctx, const double def[6], const int32_t* elements, int32_t n);

/* The assembled K (before Dirichlet elimination only if no Dirichlet set is active)
 * in CSR value order (device [csr_nnz], FP64). */
int efe_export_csr_values(efe_ctx* ctx, double* values);

/* y = K x with the context's K (device [n_dof], FP64 vectors): the FP64 assembled operator,
 * or the FP32 matrix-free one after an EFE_FP32_MATFREE pass. */
int efe_spmv(efe_ctx* ctx, const double* x, double* y);

/* Solves K x = b with the context's (eliminated) K (device [n_dof]); nonlinearimplicitstatic.py:727-751.
 * For EFE_FP32 the Krylov iteration runs entirely in FP32; x on Dirichlet DOFs is b exactly. */
int efe_linsolve(efe_ctx* ctx, const double* b, double* x, const efe_solve_opts* opts, efe_solve_info* info);

/* Convergence quantities of nonlinearimplicitstatic.py:671-687 and :786-790 (host out[5]):
 * [max|R|, argmax|R| (first), ||ddU||_inf (0 if ddU is NULL), ||F||_1, any NaN in R or ddU]. */
int efe_convergence_norms(efe_ctx* ctx, const double* R, const double* ddU, const double* F, double out[5]);

/* Per-specimen convergence quantities for a batched context (host out[n_specimens][4]):
 * [max|R|, ||ddU||_inf, ||F||_1, NaN flag]. */
int efe_convergence_norms_batched(efe_ctx* ctx, const double* R, const double* ddU, const double* F, double* out);

/* Batched contexts (specimen_of_node given): one block-Jacobi PCG per specimen, run
 * together.  Each specimen has its own alpha/beta, its own stop rule
 * ||r_s||_2 <= max(spec_atol[s], rtol_s ||b_s||_2), rtol_s = spec_rtol[s] (host [n_specimens])
 * or opts->rtol when spec_rtol is NULL, and freezes once converged;
 * spec_active (host [n_specimens] or NULL = all) excludes specimens (x = 0 there);
 * spec_iterations (host [n_specimens] or NULL) receives the iteration counts.
 * Replaces one linearSolve per specimen model (nonlinearimplicitstatic.py:727-751). */
int efe_linsolve_batched(efe_ctx* ctx, const double* b, double* x, const efe_solve_opts* opts,
                         const double* spec_atol, const double* spec_rtol, const uint8_t* spec_active,
                         int32_t* spec_iterations, efe_solve_info* info);

/* Accepts the temporary state of the specimens with accept[s] != 0 (host [n_specimens]);
 * femodel.py:245-258 per specimen model. */
int efe_accept_state_batched(efe_ctx* ctx, const uint8_t* accept);

/* Options of a device-resident Newton round (efe_newton_round_batched). */
typedef struct efe_newton_opts {
    int32_t precision;       /* EFE_FP64, EFE_FP32 or EFE_FP32_MATFREE: element pass, operator and PCG */
    int32_t krylov_max_iter; /* PCG iteration limit; a specimen that reaches it fails its increment */
    int32_t check_every;     /* PCG iterations per replayed batch */
    int32_t mixed_first;     /* 1: a first iteration (prescribed increment applied) solves to rtol_first */
    int32_t alt_after;       /* iterations after which flux_tol_alt replaces flux_tol (nonlinearimplicitstatic.py:666-669: 15) */
    int32_t reserved;
    double rtol;             /* PCG relative tolerance */
    double rtol_first;       /* with mixed_first */
    double atol_factor;      /* PCG absolute tolerance = atol_factor x the specimen's flux threshold */
    double flux_tol;         /* fluxResidualTolerance (1e-8) */
    double flux_tol_alt;     /* fluxResidualToleranceAlternative (5e-3) */
    double correction_tol;   /* fieldCorrectionTolerance (1e-8) */
    double flux_floor;       /* lower bound of the flux threshold (1e-7) */
    double avg_floor;        /* lower bound of the average flux (1e-10) */
} efe_newton_opts;

#define EFE_NEWTON_ITERATING 0
#define EFE_NEWTON_CONVERGED 1
#define EFE_NEWTON_FAILED 2
#define EFE_NEWTON_IDLE 3
#define EFE_FAIL_NONE 0
#define EFE_FAIL_CUTBACK 1    /* a material asked for a smaller increment */
#define EFE_FAIL_NAN 2        /* NaN in the residual or the correction */
#define EFE_FAIL_ITERATIONS 3 /* maxIter or maxGrowingIter reached */
#define EFE_FAIL_KRYLOV 4     /* the PCG reached krylov_max_iter or returned a non-finite correction */

/* Batched contexts: the Newton iterations of one lockstep round, every specimen's convergence
 * test, divergence count and cutback decided on the device (the loop of
 * nonlinearimplicitstatic.py:405-458 and its tests :625-725, per specimen).  One pass is one
 * replayed hipGraph: the previous solve's correction applied, element pass and assembly over the
 * iterating specimens, residual, per-specimen norms and decisions, PCG set-up and the first
 * check_every PCG iterations; further PCG batches follow until every solving specimen has
 * converged.  The host reads two counts per PCG batch and nothing else.
 * dU, P, F, PExt (or NULL), deltaT: device [n_dof]; dU holds the increment's start (the
 * extrapolated increment, deltaT on prescribed DOFs of extrapolated specimens) and receives the
 * round's increments; P and F (or NULL) receive the last pass's.  deltaT: the prescribed
 * increments (taken on prescribed DOFs of a specimen's first iteration).  Host [n_specimens]:
 * live (0: the specimen idles), extrapolated, has_dirichlet, max_iter, max_grow (maxIter,
 * maxGrowingIter of its step), n_acc (element DOF count: the average-flux divisor).  Out, host
 * [n_specimens]: state (EFE_NEWTON_*), iterations, reason (EFE_FAIL_*), work [n_specimens][3]
 * (element passes, linear solves, PCG iterations of the round).  No counterpart in the
 * reference, which iterates one model on the host. */
int efe_newton_round_batched(efe_ctx* ctx, double* dU, double* P, double* F, const double* PExt, const double* deltaT,
                             const uint8_t* live, const uint8_t* extrapolated, const uint8_t* has_dirichlet,
                             const int32_t* max_iter, const int32_t* max_grow, const double* n_acc,
                             const efe_newton_opts* opts, int32_t* state, int32_t* iterations, int32_t* reason,
                             int32_t* work);

/* After an efe_compute_elements that returned EFE_CUTBACK on a batched context: out[s] = 1
 * for the specimens whose material asked for a smaller increment (host [n_specimens]);
 * the CutbackRequest of vonmises.py:230-231 raised per specimen model. */
int efe_get_specimen_status(efe_ctx* ctx, uint8_t* out);

/* Batched contexts: the specimens the following element passes cover (host [n_specimens],
 * 1 = computed; NULL = all).  A skipped specimen keeps the temporary QP state, element
 * residuals, P, F, operator and block-Jacobi rows of its last computed pass, which a pass with
 * an unchanged increment would reproduce bit for bit: the lockstep driver skips the specimens
 * that have converged, failed or idle in a round.  No counterpart in the reference, which
 * solves each model on its own. */
int efe_set_active_specimens(efe_ctx* ctx, const uint8_t* active);

/* Number of specimens of the context (1 for an ordinary context). */
int efe_get_specimen_count(efe_ctx* ctx);

/* The linear-solver plugin contract f(A: csr_matrix, b) -> x (config/linsolve.py:46-86, called
 * by NIST.linearSolve, nonlinearimplicitstatic.py:727-751) on the GPU, for a scalar CSR matrix
 * as NIST hands it over: K after applyDirichletK (nonlinearimplicitstatic.py:559-593), whose
 * constrained rows are unit rows while their columns stay.  A row whose only stored entry is a
 * unit diagonal is prescribed: x = b there, and its column moves to the right-hand side, so the
 * Krylov method solves the free block (symmetric when K is) K_ff x_f = b_f - K_fd b_d.
 * opts->solver: EFE_SOLVER_CG (symmetric K), EFE_SOLVER_GMRES (restarted GMRES(opts->restart),
 * non-symmetric K) or EFE_SOLVER_BICGSTAB; Jacobi preconditioner (opts->precond and precision are
 * ignored: FP64); stop ||r_f||_2 <= max(rtol ||b_f'||, atol).
 * indptr host [n+1] (0 first, non-decreasing), indices host [indptr[n]] in [0, n), data host
 * [indptr[n]], b and x host [n].  Allocates device memory for the call and frees it. */
int efe_csr_linsolve(int32_t n, const int32_t* indptr, const int32_t* indices, const double* data, const double* b,
                     double* x, const efe_solve_opts* opts, efe_solve_info* info);

/* efe_csr_linsolve on device arrays (indptr, indices, data, b, x device; x may not alias b),
 * enqueued on `stream` (hipStream_t, NULL = default); the pattern is validated on the device
 * first (EFE_ERR_ARG when indptr or an index is out of range).  Allocates its workspace for the
 * call. */
int efe_csr_linsolve_device(int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                            const double* b, double* x, void* stream, const efe_solve_opts* opts,
                            efe_solve_info* info);

/* efe_csr_linsolve_device without allocation: the same solve (the same kernels, bit for bit)
 * with its workspace in the library's static device array instead of a hipMalloc, for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid it).  It synchronises
 * `stream` (pattern check, convergence), so it is not for graph capture.  No counterpart in the
 * reference.  Limits: EFE_SOLVER_CG up to 1048576 unknowns,
 * EFE_SOLVER_BICGSTAB up to ~700000 (EFE_ERR_ARG beyond, and for EFE_SOLVER_GMRES), checked
 * before any device work.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (concurrent streams or threads) race. */
int efe_csr_linsolve_device_noalloc(int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                                    const double* b, double* x, void* stream, const efe_solve_opts* opts,
                                    efe_solve_info* info);

/* Accepts the temporary QP state as the new reference state (femodel.py:245-258). */
int efe_accept_state(efe_ctx* ctx);

/* Copies the reference (which=0) or temporary (which=1) QP state to the host,
 * layout [n_elements][n_qp][n_state] with [stress 6, strain 6, material state]
 * (element.py:222-236). */
int efe_get_qp_state(efe_ctx* ctx, int32_t which, double* out);

/* Element stiffness and residual of the last element pass for elements[i] (host [n]), in
 * the layout of the native element ABI (MarmotElement::computeYourself,
 * marmotelement/element.pxd:97-103): Ke device [n][nDof][nDof] column-major with DOFs
 * node-major (3a + i), Pe device [n][nDof] (may be NULL); nDof = 3 x nodes per element.
 * For a host-side assembly in the manner of NISTParallel (nonlinearimplicitstaticparallel.pyx),
 * or inspection.  Needs an assembled (FP64 or FP32) efe_compute_elements as the last element pass
 * (EFE_ERR_STATE after a matrix-free or residual-only pass), and for Pe no load computation
 * (efe_body_force, efe_surface_traction) since. */
int efe_get_element_matrices(efe_ctx* ctx, const int32_t* elements, int32_t n, double* Ke, double* Pe);

/* The native element ABI, MarmotElement::computeYourself(QTotal, dQ, Pe, Ke, time, dT, pNewDT)
 * (marmotelement/element.pxd:97-103; called per element by NISTParallelForMarmotElements,
 * solvers/nonlinearimplicitstaticparallel.pyx:65-167), for the n elements elements[i] (host [n])
 * of the context in one launch, with caller-owned device arrays in the ABI's layouts:
 *   dQ        [n][nDof]          DOF increments, node-major (3 a + c); QTotal [n][nDof] is not read
 *                                (the small-strain element uses the increment only, element.py:290-346)
 *   stateRef  [n][nQP][nState]   the element's state at the increment start (Marmot's state buffer
 *                                after _initializeStateVarsTemp; element.py:222-236 layout)
 *   stateTemp [n][nQP][nState]   receives the updated state (may equal stateRef)
 *   Pe        [n][nDof]          receives -sum B^T sigma detJ w (overwritten)
 *   Ke        [n][nDof][nDof]    receives sum B^T C B detJ w column-major (overwritten), or NULL
 * nState = n_state_per_qp of efe_get_sizes.  The element formulation, material and properties
 * are the context's.  A material asking for a smaller increment lowers *pNewDT (in/out, as the
 * caller's pNewDT initialised to 1e36) to at most 0.5 and returns EFE_CUTBACK. */
int efe_element_compute(efe_ctx* ctx, const int32_t* elements, int32_t n, const double* QTotal, const double* dQ,
                        const double* stateRef, double* stateTemp, double* Pe, double* Ke, const double time[2],
                        double dT, double* pNewDT);

/* Device gather of QP results (utils/elementresultcollector.pyx:37-132, which collects
 * element.getResultArray(result, qp) for a list of elements and quadrature points): out
 * (device [n][nq][count]) = components first .. first+count-1 of the reference (which=0) or
 * temporary (which=1) state, for elements[i] (host [n]; NULL: all elements in order) at the
 * quadrature points qps[j] (host [nq]).  Results: stress = components 0..5, strain = 6..11,
 * material state (J2 / Drucker-Prager kappa) = 12 (element.py:222-236). */
int efe_collect_results(efe_ctx* ctx, int32_t which, const int32_t* elements, int32_t n, const int32_t* qps,
                        int32_t nq, int32_t first, int32_t count, double* out);

/* Overwrites the reference QP state from the host (same layout); in == NULL sets it to
 * zero (the initial state of a fresh model, element.py:207-240). */
int efe_set_qp_state(efe_ctx* ctx, const double* in);

/* Elements per workgroup of the element kernel and threads per workgroup of the
 * Krylov kernels; 0 keeps the current choice. */
int efe_set_launch(efe_ctx* ctx, int32_t elements_per_workgroup, int32_t krylov_block);

/* Device times (ms) measured with HIP events on the context's streams (host out[16]):
 *   out[0..3]   the last call: element kernel, assembly, one Krylov iteration (average of the
 *               last solve), norms;
 *   out[4..14]  totals since the context was created: element kernel of tangent passes (ms),
 *               tangent passes (count), assembly (ms), residual-only passes (ms), residual-only
 *               passes (count), Krylov iterations (ms, without preconditioner set-up), Krylov
 *               iterations (count), solves (count), ILU(0) factorisations (ms), norms (ms),
 *               norm calls (count); out[15] = 0.
 * The totals divided by their counts are the history's per-call averages (the computation-time
 * categories of config/timing.py, measured on the device). */
int efe_get_phase_times(efe_ctx* ctx, double out[16]);

#ifdef __cplusplus
}
#endif

#endif

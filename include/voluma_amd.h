/* voluma_amd.h — C ABI of the MI355X (gfx950) LDU linear-solve library.
 *
 * Drop-in boundary for Voluma's linear-solve path. The reference interface this
 * replaces (paths relative to the Voluma source tree):
 *
 *   SolverPerformance solveLdu(const LduSystem&, std::vector<scalar>& x,
 *                              const SolverControls&);           src/fvm/linalg/LinearSolver.h:42
 *   void amul(const LduSystem&, std::vector<scalar>& x,
 *             std::vector<scalar>& y);                            src/fvm/linalg/LinearSolver.h:45
 *   struct SolverControls / SolverPerformance / LduSystem         src/fvm/linalg/LinearSolver.h:15-40
 *   GamgHierarchy buildHierarchy(const Mesh&) (cached per mesh)   src/fvm/linalg/GAMG.h:57-58
 *
 * The reference's LduSystem is a non-owning view {mesh, diag (boundary-folded), upper,
 * lower (== upper when symmetric), b}. Here the mesh addressing and the GAMG hierarchy
 * live in a vol_mesh handle (built once per mesh, like the reference's Mesh::cached
 * hierarchy), the coefficients in a vol_matrix handle (resident in HBM, so repeated
 * solves with one matrix — PISO correctors — reuse them), and b/x are passed per solve.
 *
 * Conventions: plain pointers and sizes, int32 labels, FP64 values; every function
 * returns VOL_OK (0) or a negative status and never throws; vol_last_error() gives
 * the message of the calling thread's last failure. Non-convergence is not an error
 * (perf.converged = 0), as in the reference (LinearSolver.cpp:283).
 * Host pointers are copied; with VOL_DEVICE the pointers are device memory of the
 * context's GPU and stay where they are. Handles keep their parents alive (matrix →
 * mesh → context, device allocations → context): destroy them in any order.
 * Calls on one context are serialised on its stream; do not call into one context
 * from two threads at once.
 */
#ifndef VOLUMA_AMD_H
#define VOLUMA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference exception sites in parentheses) */
#define VOL_OK 0
#define VOL_ERR_INVALID_ARGUMENT (-1)
#define VOL_ERR_UNKNOWN_SOLVER (-2)         /* LinearSolver.cpp:280 */
#define VOL_ERR_UNKNOWN_PRECONDITIONER (-3) /* LinearSolver.cpp:76 */
#define VOL_ERR_PCG_NONSYMMETRIC (-4)       /* LinearSolver.cpp:177 */
#define VOL_ERR_SINGULAR_COARSEST (-5)      /* GAMG.cpp:336 */
#define VOL_ERR_UNSUPPORTED (-6)            /* valid in the reference, not provided on the GPU */
#define VOL_ERR_HIP (-7)                    /* HIP runtime failure */
#define VOL_ERR_NO_GAMG_WEIGHTS (-8)        /* GAMG requested on a mesh created without weights */
#define VOL_ERR_TIMEOUT (-9)                /* vol_server_wait: not finished in time (no reference counterpart) */

/* ---- memory location of caller buffers */
#define VOL_HOST 0
#define VOL_DEVICE 1

/* ---- summation modes (vol_solver_controls.sums) */
#define VOL_SUMS_CONTEXT (-1)     /* follow the context: CFD_REPRODUCIBLE or vol_set_reproducible */
#define VOL_SUMS_FAST 0           /* deterministic tree sums (the reference's default mode analogue) */
#define VOL_SUMS_REPRODUCIBLE 1   /* reference reproducible mode: exact sums + preconditioner swap */
#define VOL_SUMS_EXACT 2          /* exact sums, no swap (deterministic GAMG; extension) */

/* ---- arithmetic of the solve (vol_solver_controls.precision) */
#define VOL_FP64 0
/* FP64 defect correction around an FP32 inner Krylov solve: each outer step forms
 * r = b − A·x in FP64, solves A·e = r with PCG/PBiCGStab whose Amul, GAMG smoothing,
 * restriction/prolongation and vectors are FP32 (dot products, norms and the coarsest
 * LU solve in FP64), until the inner residual falls by innerRelTol, then x += e.
 * PCG keeps its search direction across the outer steps (a reliable update: only the
 * recursive residual is replaced by the FP64 one); PBiCGStab restarts. Convergence is
 * judged on the FP64 residual with the usual normalisation. Fast sums only; no
 * counterpart in the reference. */
#define VOL_MIXED 1

typedef struct vol_context_s* vol_context;
typedef struct vol_mesh_s* vol_mesh;
typedef struct vol_matrix_s* vol_matrix;
typedef struct vol_batch_s* vol_batch;
typedef struct vol_server_s* vol_server;

/* SolverControls (LinearSolver.h:15-23) plus GPU extensions. Strings as in the
 * reference: solver PCG | PBiCGStab | GaussSeidel | symGaussSeidel | Jacobi;
 * preconditioner diagonal | DIC | DILU | GAMG | none. Call vol_controls_default()
 * to get the reference defaults. */
typedef struct {
    const char* solver;
    const char* preconditioner;
    double tolerance;   /* 1e-8 */
    double relTol;      /* 0 */
    int maxIter;        /* 1000 */
    int minIter;        /* 0 */
    int nSweeps;        /* 1 */
    /* extensions */
    int sums;           /* VOL_SUMS_*, default VOL_SUMS_CONTEXT */
    int precision;      /* VOL_FP64 | VOL_MIXED */
    int gamgPreSweeps;  /* GAMG smoother sweeps before the coarse correction (default 2, ≥ 1) */
    int gamgPostSweeps; /* ... and after it (default 2, ≥ 1) */
    double gamgOmega;   /* damped-Jacobi smoother: damping (default 0.8) */
    double innerRelTol; /* VOL_MIXED: residual reduction per inner solve (default 1e-2, in (0, 1)) */
    int gamgSmoother;   /* VOL_GAMG_JACOBI (default) | VOL_GAMG_GAUSS_SEIDEL */
    int reserved;
} vol_solver_controls;

/* ---- GAMG smoother (vol_solver_controls.gamgSmoother) */
/* damped Jacobi x ← x + ω(D⁻¹(r − (A − D)x) − x), parallel on every level */
#define VOL_GAMG_JACOBI 0
/* the reference's smoother (GAMG.cpp:363-384,423-442): x = 0, gamgPreSweeps forward
 * Gauss–Seidel sweeps, coarse correction, gamgPostSweeps backward sweeps, each level
 * run as level-scheduled sweeps; with 2/2 sweeps and exact sums the V-cycle is the
 * reference's bit for bit */
#define VOL_GAMG_GAUSS_SEIDEL 1

/* SolverPerformance (LinearSolver.h:25-31) */
typedef struct {
    char solver[32];
    double initialResidual;
    double finalResidual;
    double residualSum;
    double normFactor;
    int iterations;       /* Krylov iterations (VOL_MIXED: summed over the inner solves) */
    int converged;
    int outerIterations;  /* VOL_MIXED: defect-correction steps; 0 for VOL_FP64 */
    int reserved;
} vol_solver_performance;

void vol_controls_default(vol_solver_controls* c);

/* ---- context: one GPU, one HIP stream. Reads CFD_REPRODUCIBLE like
 * par::Environment does (src/fvm/parallel/Comm.cpp:48). */
int vol_context_create(int device, vol_context* out);
int vol_context_destroy(vol_context ctx);
int vol_set_reproducible(vol_context ctx, int on); /* par::setReproducible, Comm.h:38 */
int vol_synchronize(vol_context ctx);
const char* vol_last_error(void);

/* Device memory of the context's GPU for VOL_DEVICE buffers. */
int vol_device_alloc(vol_context ctx, uint64_t bytes, void** ptr);
int vol_device_free(vol_context ctx, void* ptr);
int vol_memcpy_to_device(vol_context ctx, void* dst, const void* src, uint64_t bytes);
int vol_memcpy_to_host(vol_context ctx, void* dst, const void* src, uint64_t bytes);
/* dst[i] = src[i] for n doubles on the GPU with the library's streaming kernel,
 * `repeats` times; *seconds = average time of one copy (HIP events). Measurement aid
 * (achievable HBM bandwidth, counter calibration); no reference counterpart. */
int vol_device_copy(vol_context ctx, double* dst, const double* src, int64_t n, int repeats, double* seconds);

/* ---- mesh: LDU addressing of nInternalFaces faces between nCells cells
 * (Mesh::owner/neighbour, internal part). gamgWeights (nullable) is the per-face
 * agglomeration weight |S_f|·nonOrthDeltaCoeff_f that buildHierarchy uses
 * (GAMG.cpp:210-218); with it the GAMG hierarchy is built here, once.
 * mergeLevels: pairings per level (CFD_GAMG_MERGE, GAMG.cpp:96-99; 0 → 2). */
int vol_mesh_create(vol_context ctx, int nCells, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                    const double* gamgWeights, int mergeLevels, vol_mesh* out);
int vol_mesh_destroy(vol_mesh mesh);
/* GAMG hierarchy inspection (GamgHierarchy::levels, GAMG.h:18-31): number of levels
 * and per-level cell counts (sizes has room for 40). */
int vol_mesh_gamg_levels(vol_mesh mesh, int* nLevels, int32_t* sizes);
/* level k's agg (cell → next-level cell), faceMap and faceFlip (−1 / 0 on the coarsest) */
int vol_mesh_gamg_level(vol_mesh mesh, int level, int32_t* nFaces, int32_t* agg, int32_t* faceMap,
                        int32_t* faceFlip);
/* How the hierarchy was built (no reference counterpart; measurement aid): seconds taken,
 * pairing passes and decision rounds run on the GPU (0 / 0 when built on the host:
 * meshes below 32768 cells, or VOLUMA_AGGLOMERATION=host). Any output may be NULL. */
int vol_mesh_gamg_setup_info(vol_mesh mesh, double* seconds, int* passes, int* rounds);

/* Boundary faces of the mesh (after the internal ones, in patch order), by owner cell:
 * Mesh::owner()[nInternalFaces + bf]. Needed by the assembly entry points. */
int vol_mesh_set_boundary(vol_mesh mesh, int nBoundaryFaces, const int32_t* boundaryOwner);

/* ---- matrix: coefficients over a mesh. diag must already include the boundary
 * contributions (LinearSolver.h:35, folded by FvMatrix::solve, FvMatrix.h:154-163);
 * lower == NULL or lower == upper means symmetric (LinearSolver.h:37). */
int vol_matrix_create(vol_mesh mesh, vol_matrix* out);
int vol_matrix_set(vol_matrix A, const double* diag, const double* upper, const double* lower, int location);
int vol_matrix_destroy(vol_matrix A);
/* read back diag (folded), upper and lower (== upper when symmetric); FvMatrix::diag()/
 * upper()/lowerPtr() (FvMatrix.h:32-44). Any output may be NULL. */
int vol_matrix_get(vol_matrix A, double* diag, double* upper, double* lower, int location);

/* y = A x (LinearSolver.cpp:121-133), x and y of length nCells. */
int vol_amul(vol_matrix A, const double* x, double* y, int location);

/* solveLdu (LinearSolver.cpp:135-285): x in/out, b of length nCells. */
int vol_solve(vol_matrix A, const double* b, double* x, const vol_solver_controls* controls,
              vol_solver_performance* perf, int location);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same solve as vol_solve with b and x in device memory (VOL_DEVICE), whose work
 * vectors, reduction partials and loop scalars live in a fixed array of
 * VOL_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of buffers the
 * library allocates. It calls no allocation routine. Limits: at most
 * VOL_NOALLOC_MAX_CELLS cells (VOL_ERR_INVALID_ARGUMENT beyond); solver PCG or
 * PBiCGStab, preconditioner none or diagonal, precision VOL_FP64, any sums mode
 * (VOL_ERR_UNSUPPORTED otherwise; reproducible mode swaps every preconditioner to
 * diagonal as vol_solve does). Results are bit-identical to vol_solve's. Every call on
 * one GPU uses that GPU's one array: calls that overlap on a GPU (two threads or two
 * contexts) race. */
#define VOL_NOALLOC_MAX_CELLS 262144
#define VOL_NOALLOC_SCRATCH_BYTES (24u * 1024u * 1024u)
int vol_solve_noalloc(vol_matrix A, const double* b, double* x, const vol_solver_controls* controls,
                      vol_solver_performance* perf);

/* ---- coefficient assembly on the GPU, followed by the boundary fold of FvMatrix::solve
 * (FvMatrix.h:154-163), for a scalar field whose patches are fixedValue (bcFixed[bf] = 1,
 * value bcValue[bf], nullable → 0) or zeroGradient (0) (BoundaryCondition.h:116-137).
 * Face arrays (faceGamma, magSf, ndc, flux) hold nInternalFaces + nBoundaryFaces entries.
 * A receives the folded diagonal and upper (and lower for div); bOut (nullable) gets
 * bIn (nullable → 0) plus the folded boundary source; internalCoeffs/boundarySource
 * (nullable, nBoundaryFaces) get the unfolded boundary terms.
 *   fvm::laplacian(γ, ψ) without the explicit non-orthogonal correction: Fvm.h:141-154,171-180
 *   (faceGamma nullable → uniform gamma)
 *   fvm::div(F, ψ, upwind): Fvm.h:74-106,126-135 */
int vol_assemble_laplacian(vol_matrix A, double gamma, const double* faceGamma, const double* magSf,
                           const double* ndc, const int32_t* bcFixed, const double* bcValue, const double* bIn,
                           double* bOut, double* internalCoeffs, double* boundarySource, int location);
int vol_assemble_div_upwind(vol_matrix A, const double* flux, const int32_t* bcFixed, const double* bcValue,
                            const double* bIn, double* bOut, double* internalCoeffs, double* boundarySource,
                            int location);

/* ---- FvMatrix algebra on the GPU (FvMatrix.h:21-108,154-163,232-244; Fvm.h:18-186).
 * A vol_matrix also holds the FvMatrix parts of one field component: the unfolded
 * diagonal, source (nCells), internalCoeffs and boundarySource (nBoundaryFaces) and the
 * face-flux correction (nInternalFaces + nBoundaryFaces); upper/lower are the matrix's.
 * vol_fvm_reset zeroes them, as FvMatrix(psi) does. Each term is built alone, as the
 * reference builds it in a fresh FvMatrix, and added with sign +1 (operator+=) or −1
 * (operator-=); the lower part appears when a term with one (div) is added. So
 *     UEqn = ddt(U) + div(phi, U) − laplacian(nu, U)   (IncompressibleFlow.cpp:183-185)
 * is vol_fvm_reset, vol_fvm_ddt_euler(+1), vol_fvm_div_upwind(+1),
 * vol_fvm_laplacian(−1), and the parts are bit-identical to the reference's.
 * vol_fvm_finish folds internalCoeffs into the diagonal and boundarySource into
 * b = source (FvMatrix.h:154-163) and makes A ready for vol_solve. Boundary faces are
 * fixedValue (bcFixed 1, value bcValue, nullable → 0) or zeroGradient (0). Face arrays
 * hold nInternalFaces + nBoundaryFaces entries (Sf, corrVecs: 3 per face). Needs
 * vol_mesh_set_boundary. */
typedef struct {
    const double* psi;      /* current field (nCells): its Gauss-linear gradient (Fvc.h:208-224) */
    const double* w;        /* face interpolation weights */
    const double* Sf;       /* face area vectors */
    const double* corrVecs; /* non-orthogonal correction vectors */
    const double* V;        /* cell volumes */
} vol_nonorth;
int vol_fvm_reset(vol_matrix A);
/* fvm::ddt, Euler (Fvm.h:18-31; Schemes.h:71-86): diag += V/dt, source += V·psi0/dt */
int vol_fvm_ddt_euler(vol_matrix A, int sign, double dt, const double* V, const double* psi0, int location);
/* fvm::ddt with ddtCoeffs (Fvm.h:18-31): diag += c·V, source += V·(c0·psi0 − c00·psi00);
 * Euler: c = c0 = 1/dt, c00 = 0; backward (variable-step BDF2, Schemes.h:75-81) with the
 * old-old field psi00 (nullable → psi0, as VolField::oldOld() is without one) */
int vol_fvm_ddt(vol_matrix A, int sign, double c, double c0, double c00, const double* V, const double* psi0,
                const double* psi00, int location);
/* fvm::div(F, ψ, upwind) (Fvm.h:74-137) */
int vol_fvm_div_upwind(vol_matrix A, int sign, const double* flux, const int32_t* bcFixed, const double* bcValue,
                       int location);
/* fvm::laplacian(γ, ψ, nonOrthCorr) (Fvm.h:141-186); faceGamma nullable → uniform gamma;
 * nonOrth nullable → no explicit correction; otherwise it is applied when some internal
 * face has |corrVecs|² > 1e-20 (Fvm.h:157), with fc = γ|S|·(corr·(w g_P + (1−w) g_N)). */
int vol_fvm_laplacian(vol_matrix A, int sign, double gamma, const double* faceGamma, const double* magSf,
                      const double* ndc, const int32_t* bcFixed, const double* bcValue, const vol_nonorth* nonOrth,
                      int location);
/* FvMatrix −= su (FvMatrix.h:93-97): source += su·V per cell */
int vol_fvm_source(vol_matrix A, const double* su, const double* V, int location);
/* copy out the parts (each nullable) */
int vol_fvm_get(vol_matrix A, double* diag, double* source, double* internalCoeffs, double* boundarySource,
                double* faceFluxCorrection, int location);
/* fold and finish (FvMatrix.h:154-163): A's solve diagonal = diag + Σ internalCoeffs,
 * b (nullable, nCells) = source + Σ boundarySource per owner cell */
int vol_fvm_finish(vol_matrix A, double* b, int location);
/* fold and finish of a copy of the matrix with −= su (FvMatrix.h:93-97), as the momentum
 * predictor solves FvMatrix M(UEqn); M −= ngp (IncompressibleFlow.cpp:214-226): b =
 * (source + su·V) + Σ boundarySource; the matrix's own source is left unchanged, so
 * vol_fvm_H afterwards sees UEqn without su. */
int vol_fvm_finish_source(vol_matrix A, const double* su, const double* V, double* b, int location);

/* ---- pressure-velocity coupling on the FvMatrix parts (FvMatrix.h:111-141,179-222).
 * Results are bit-identical to the reference's. A vector equation (UEqn) is handled one
 * component at a time, its parts assembled per component; nComponents = 3 then selects
 * the vector form of cmptAv, (v + v + v)/3, of the boundary coefficients
 * (FvMatrix.h:227-228), exact for patches whose coefficient is the same in every
 * component (fixedValue, zeroGradient). psi: the field's current cell values (nCells). */
/* FvMatrix::relax(alpha) (FvMatrix.h:111-131): D = max(|diag + Σ ic|, Σ|offdiag|),
 * diag += D/α − (diag + Σ ic), source += (that increment)·psi; alpha >= 1: no change.
 * Call vol_fvm_finish (again) before solving. */
int vol_fvm_relax(vol_matrix A, double alpha, const double* psi, int location);
/* FvMatrix::setReference(cell, value) (FvMatrix.h:134-141): source[cell] += diag·value,
 * diag[cell] doubled; global cell numbers are the mesh's own (single domain), and a cell
 * outside [0, nCells) changes nothing, as on a rank that does not own it. */
int vol_fvm_set_reference(vol_matrix A, int64_t cell, double value);
/* FvMatrix::A() (FvMatrix.h:179-187): out[c] = (diag + Σ cmptAv(ic)) / V (nCells) */
int vol_fvm_A(vol_matrix A, const double* V, int nComponents, double* out, int location);
/* FvMatrix::H() (FvMatrix.h:189-208), one component: out[c] = (source − Σ offdiag·psi_nb
 * + Σ bs − Σ (ic − cmptAv ic)·psi_c) / V (nCells) */
int vol_fvm_H(vol_matrix A, const double* psi, const double* V, int nComponents, double* out, int location);
/* FvMatrix::flux() (FvMatrix.h:210-222), one component: out (nInternalFaces +
 * nBoundaryFaces) = upper·psi_N − lower·psi_P on internal faces, ic·psi_P − bs on
 * boundary faces, + faceFluxCorrection when a laplacian term switched it on. */
int vol_fvm_flux(vol_matrix A, const double* psi, double* out, int location);

/* ---- mesh geometry on the GPU (Mesh.cpp:21-48 and the cell pyramid decomposition;
 * Mesh.cpp:600-630), bit-identical to the reference's Mesh: from points (interleaved
 * xyz) and faces (faceOffsets, nFaces + 1 entries, into facePoints; every face ≥ 3
 * points) with owner (nFaces) and neighbour (nInternalFaces), as RawMesh holds them;
 * outputs V, C (cells), Sf, Cf, d, magSf, w, ndc, corrVecs, skewVecs (faces; skew
 * nullable) and *maxSkewness (host pointer, nullable; the mesh is skewed above 1e-8,
 * Mesh.cpp:596).
 * Labels are int32. Inputs and outputs on the host or all on the device (location). */
int vol_geometry(vol_context ctx, int nPoints, const double* points, int nFaces, const int32_t* faceOffsets,
                 const int32_t* facePoints, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                 int nCells, double* V, double* C, double* Sf, double* Cf, double* d, double* magSf, double* w,
                 double* ndc, double* corrVecs, double* skewVecs, double* maxSkewness, int location);

/* ---- explicit operators of the pressure corrector (Fvc.h), on the GPU.
 * vol_mesh_set_geometry stores the mesh geometry the operators need (after
 * vol_mesh_set_boundary): V (nCells); per face (nInternalFaces + nBoundaryFaces): Sf and
 * d (interleaved xyz), magSf, w; skew (Mesh::skewVecs, interleaved xyz, nullable → 0)
 * and skewed (Mesh::skewed(): whether interpolate/flux apply the skew correction; the
 * Gauss-skew gradient always uses skew, as the reference does). Multi-component values
 * are interleaved per cell or face, as std::vector<Vec3>/<Tensor> hold them: nComponents
 * 1 (scalar field; its gradient a Vec3 per cell) or 3 (vector field; its gradient a
 * row-major Tensor per cell). Results are bit-identical to the reference's. A 2-D mesh
 * passes the faces the reference's Mesh keeps (empty patches dropped) and its empty
 * direction (vol_mesh_set_empty_dir). */
int vol_mesh_set_geometry(vol_mesh mesh, const double* V, const double* Sf, const double* magSf, const double* w,
                          const double* d, const double* skew, int skewed, int location);
/* A 2-D mesh (Mesh::twoD, emptyDir; Mesh.cpp:300-330): the empty direction (0, 1, 2; −1
 * for 3-D, the default). Every gradient then adds 1 to the empty diagonal of its
 * least-squares matrix and zeroes the empty component (Fvc.h:152,169-181): vol_fvc_grad
 * and the non-orthogonal gradient of vol_fvm_laplacian. */
int vol_mesh_set_empty_dir(vol_mesh mesh, int emptyDir);
/* fvc::interpolate (Fvc.h:13-31,72-81): w c_P + (1−w) c_N on internal faces, plus
 * gradDot(skew, w g_P + (1−w) g_N) when cellGrad is given and the mesh is skewed;
 * boundary faces take boundaryValues (nullable → the owner cell's value). */
int vol_fvc_interpolate(vol_mesh mesh, int nComponents, const double* cellValues, const double* cellGrad,
                        const double* boundaryValues, double* faceValues, int location);
/* fvc::grad (Fvc.h:185-291) with fixedValue faces (bFixesValue 1, face values bValue) and
 * zeroGradient faces otherwise: VOL_GRAD_LEAST_SQUARES (= fvc::gradLS; nComponents 1 or
 * 3), VOL_GRAD_GAUSS_LINEAR and VOL_GRAD_GAUSS_SKEW (nComponents 1). */
#define VOL_GRAD_GAUSS_LINEAR 0
#define VOL_GRAD_LEAST_SQUARES 1
#define VOL_GRAD_GAUSS_SKEW 2
int vol_fvc_grad(vol_mesh mesh, int scheme, int nComponents, const double* cellValues, const int32_t* bFixesValue,
                 const double* bValue, double* cellGrad, int location);
/* fvc::div(m, F) (Fvc.h:294-305): per cell Σ ±F / V */
int vol_fvc_div(vol_mesh mesh, const double* faceFlux, double* out, int location);
/* fvc::surfaceIntegrate(m, F) (Fvc.h:307-318): per cell Σ ±F · (1/V) */
int vol_fvc_surface_integrate(vol_mesh mesh, int nComponents, const double* faceValues, double* out, int location);
/* face flux of a cell vector field as phiHbyA is formed (IncompressibleFlow.cpp:276-291):
 * dot(U_f, S_f) with U_f = interpolate(U, gradU) on internal faces; on boundary faces
 * dot(Ub, S_f) where bFixesFlux, else dot(U_P, S_f). gradU nullable. */
int vol_fvc_flux(vol_mesh mesh, const double* U, const double* gradU, const int32_t* bFixesFlux, const double* Ub,
                 double* phi, int location);

/* ---- pressure-corrector steps (IncompressibleFlow.cpp:242-362), on a mesh with geometry.
 * Cell vectors are interleaved Vec3 (nCells × 3), face arrays cover nInternalFaces +
 * nBoundaryFaces, boundary arrays nBoundaryFaces; all bit-identical to the reference. */
/* rAU = 1/A, HbyA = rAU·H (IncompressibleFlow.cpp:245-249) */
int vol_piso_rau_hbya(vol_mesh mesh, const double* A, const double* H, double* rAU, double* HbyA, int location);
/* boundary values of HbyA for its least-squares gradient (IncompressibleFlow.cpp:261-270):
 * U_b + rAU_P ∇p_P on fixed-flux faces (bFixesFlux), zero elsewhere */
int vol_piso_hbya_boundary(vol_mesh mesh, const int32_t* bFixesFlux, const double* Ub, const double* rAU,
                           const double* gradp, double* bValue, int location);
/* face-wise combinations over all faces: out = a + b·c (phiHbyA += rAUf·φ_F,
 * IncompressibleFlow.cpp:294-295), a − b (φ_F − ∇p_f·S_f, :219-220; φ = phiHbyA − F,
 * :338-339), a − b/c (φ_F − pF/rAUf, :351-352). c unused for VOL_FACE_A_MINUS_B. */
#define VOL_FACE_A_PLUS_B_TIMES_C 0
#define VOL_FACE_A_MINUS_B 1
#define VOL_FACE_A_MINUS_B_OVER_C 2
int vol_piso_face_op(vol_mesh mesh, int op, const double* a, const double* b, const double* c, double* out,
                     int location);
/* ddtPhiCorr (IncompressibleFlow.cpp:296-316), internal faces, phiHbyA updated in place:
 * ddtPhiCoeff < 0 selects the automatic coupling; c0, c00 are ddtCoeffs (Schemes.h:71-86);
 * U00 and phiOldOld are read only when c00 != 0. */
int vol_piso_ddt_phi_corr(vol_mesh mesh, double ddtPhiCoeff, double c0, double c00, const double* U0,
                          const double* U00, const double* phiOld, const double* phiOldOld, const double* rAUf,
                          double* phiHbyA, int location);
/* constrainPressure (IncompressibleFlow.cpp:17-29): the fixedFluxPressure gradient of the
 * boundary faces [patchStart, patchStart + patchSize) (boundary numbering):
 * gradient[i] = (phiHbyA_f − U_b·S_f) / (|S_f| rAUf_f) */
int vol_piso_constrain_pressure(vol_mesh mesh, int patchStart, int patchSize, const double* phiHbyA,
                                const double* Ub, const double* rAUf, double* gradient, int location);
/* fvc::reconstruct of a face flux (Fvc.h:33-67); emptyDir: the 2-D empty direction or −1 */
int vol_fvc_reconstruct(vol_mesh mesh, const double* faceFlux, int emptyDir, double* out, int location);
/* velocity correction (IncompressibleFlow.cpp:347-360): U = HbyA − rAU·v (sign < 0, v = ∇p)
 * or HbyA + rAU·v (sign > 0, v = reconstruct(...)); component emptyDir (2-D) set to 0 */
int vol_piso_correct_U(vol_mesh mesh, const double* HbyA, const double* rAU, const double* v, int sign, int emptyDir,
                       double* U, int location);

/* boundary values of a field with fixedValue (bFixed 1: bValue) or zeroGradient (the
 * owner cell's value) faces, as VolField::correctBoundaryConditions evaluates them
 * (VolField.h:89-96, BoundaryCondition.h:98-138); nComponents 1 or 3 (interleaved) */
int vol_boundary_values(vol_mesh mesh, int nComponents, const int32_t* bFixed, const double* bValue,
                        const double* cellValues, double* out, int location);
/* device-side strided copy dst[i·dstStride] = scale·src[i·srcStride] (scale ±1 copies or
 * negates exactly), e.g. a component of an interleaved vector field (device pointers;
 * no reference counterpart) */
int vol_copy_strided(vol_context ctx, int64_t n, const double* src, int srcStride, double* dst, int dstStride,
                     double scale);

/* ---- incompressible flow time steps resident on the GPU (IncompressibleFlow.{h,cpp}).
 * vol_flow holds U (interleaved Vec3), p and the face flux φ of a mesh on the device with
 * their old time levels, the momentum and pressure matrices and scratch, and advances
 * them by IncompressibleFlow::step (IncompressibleFlow.cpp:144-370) for laminar flow with
 * upwind convection, fixedValue / zeroGradient boundaries and no body force: momentum
 * assembly and predictor, nCorr pressure correctors (A, H, rAU/HbyA, rAUf, phiHbyA,
 * ddtPhiCorr, pressure equation with nNonOrthCorr correctors and a reference cell when no
 * patch fixes p, φ update, velocity correction). Nothing leaves the device during a step
 * but the solvers' scalars. The mesh needs vol_mesh_set_boundary, vol_mesh_set_geometry
 * (and vol_mesh_set_empty_dir in 2-D), GAMG weights when a GAMG solve is asked for; the
 * boundary faces are the kept faces of the reference's Mesh (empty patches dropped). */
typedef struct vol_flow_s* vol_flow;
#define VOL_DDT_EULER 1
#define VOL_DDT_BACKWARD 2
typedef struct {
    double nu;                  /* kinematic viscosity (uniform ν_eff) */
    int nCorr;                  /* PimpleControls::nCorr (2) */
    int nNonOrthCorr;           /* (0) */
    int momentumPredictor;      /* (1) */
    int ddtScheme;              /* VOL_DDT_EULER / VOL_DDT_BACKWARD (IncompressibleFlow.h:75) */
    int gradPScheme;            /* VOL_GRAD_* for ∇p (Gauss linear) */
    int laplacianNonOrthCorr;   /* (1) */
    int64_t pRefCell;           /* (0) */
    double pRefValue;           /* (0) */
    double ddtPhiCoeff;         /* (0: off; < 0 automatic) */
    vol_solver_controls UControls, pControls, pFinalControls;
} vol_flow_options;
/* PimpleControls' defaults (IncompressibleFlow.h:18-36) with the GPU's controls: U
 * PBiCGStab + DILU 1e-8; p PCG + GAMG 1e-8 relTol 0.01 and final 1e-8, GAMG smoothed by
 * Gauss-Seidel (the reference's smoother) */
void vol_flow_options_default(vol_flow_options* o);
/* faces: nInternalFaces + nBoundaryFaces of ndc and corrVecs (3 per face, interleaved);
 * per boundary face: uFixed / uValue (3 per face) and pFixed / pValue */
int vol_flow_create(vol_mesh mesh, const vol_flow_options* options, const double* ndc, const double* corrVecs,
                    const int32_t* uFixed, const double* uValue, const int32_t* pFixed, const double* pValue,
                    int location, vol_flow* out);
/* U (nCells × 3), p (nCells), phi (faces); each nullable (left as it is) */
int vol_flow_set_fields(vol_flow flow, const double* U, const double* p, const double* phi, int location);
int vol_flow_get_fields(vol_flow flow, double* U, double* p, double* phi, int location);
/* one time step; perf (nullable) receives up to maxPerf solver performances in solve
 * order, *nPerf (nullable) their number */
int vol_flow_step(vol_flow flow, double dt, vol_solver_performance* perf, int maxPerf, int* nPerf);
int vol_flow_destroy(vol_flow flow);

/* ---- batches: many independent systems solved together (no reference counterpart).
 * Systems are concatenated: owner/neighbour/gamgWeights hold each system's internal
 * faces in turn with that system's own (0-based) cell numbers; diag/upper/lower, b and
 * x are the concatenations of the systems' arrays. Each system keeps the GAMG hierarchy
 * it has alone (built as vol_mesh_create builds it; every system needs more than 64
 * cells), its own Krylov scalars and its own convergence test, and one kernel launch
 * serves all systems in each phase. A system's result does not depend on the other
 * systems of the batch; it agrees with vol_solve of that system within the solve's
 * tolerance (the reduction trees differ). PCG with GAMG, diagonal or no
 * preconditioner; fast sums; VOL_FP64 or VOL_MIXED. perf: one entry per system. */
int vol_batch_create(vol_context ctx, int nSystems, const int32_t* nCells, const int32_t* nInternalFaces,
                     const int32_t* owner, const int32_t* neighbour, const double* gamgWeights, vol_batch* out);
int vol_batch_destroy(vol_batch batch);
/* total cells and faces; levels[s] = GAMG levels of system s (nullable outputs) */
int vol_batch_info(vol_batch batch, int64_t* nCellsTotal, int64_t* nFacesTotal, int32_t* levels);
/* Batched iterations after the first replay one PCG iteration captured as a HIP graph
 * (VOLUMA_BATCH_GRAPH=0 launches every kernel instead; the results are the same bits).
 * Counts since the last reset: graph replays and eagerly launched iterations. No
 * reference counterpart. */
int vol_batch_graph_stats(vol_batch batch, int* graphLaunches, int* eagerIterations, int reset);
int vol_batch_set(vol_batch batch, const double* diag, const double* upper, const double* lower, int location);
int vol_batch_solve(vol_batch batch, const double* b, double* x, const vol_solver_controls* controls,
                    vol_solver_performance* perf, int location);

/* ---- solve server (no reference counterpart): a request loop that keeps meshes, their
 * GAMG hierarchies and matrices resident and runs requests concurrently. `slots`
 * worker threads each own a context (HIP stream + scratch); a request is admitted, in
 * arrival order, when a worker is free and its device memory (a per-worker copy of the
 * mesh, a matrix and, for a mesh with a hierarchy, a batch of 4 copies; cached, and
 * evicted least recently used first) fits hbmBudgetBytes. An entry is counted at an
 * estimate until its first request ends, then at what its allocations measured.
 * A request names a mesh registered with vol_server_add_mesh and carries one matrix and
 * nRhs right-hand sides (e.g. PISO correctors) on that matrix, solved with `controls`.
 * Pressure requests (PCG + GAMG, damped-Jacobi V-cycle, fast sums) on meshes of at most
 * 262144 cells run as systems of the worker's batch of copies, together with the requests
 * queued behind them for the same mesh and controls (at most 4 right-hand sides); other
 * requests solve their right-hand sides one after another. All request buffers are host memory owned by the
 * caller and must stay valid until vol_server_wait returns for the ticket. A request's
 * result does not depend on what else runs. */
typedef struct {
    int mesh;                        /* id from vol_server_add_mesh */
    const double* diag;              /* boundary-folded diagonal (nCells) */
    const double* upper;             /* nInternalFaces */
    const double* lower;             /* nullable: symmetric */
    int nRhs;                        /* 1..8 */
    const double* const* b;          /* nRhs right-hand sides */
    double* const* x;                /* nRhs initial guesses, overwritten by the solutions */
    vol_solver_controls controls;
    vol_solver_performance* perf;    /* nullable, nRhs entries */
} vol_request;
int vol_server_create(int device, int slots, uint64_t hbmBudgetBytes, vol_server* out);
/* registers a mesh (builds its GAMG hierarchy once); boundaryOwner nullable */
int vol_server_add_mesh(vol_server s, int nCells, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                        const double* gamgWeights, const int32_t* boundaryOwner, int nBoundaryFaces, int* meshId);
int vol_server_submit(vol_server s, const vol_request* request, int64_t* ticket);   /* returns at once */
/* makes a registered mesh resident on every worker (copy, matrix storage, batch of
 * copies), queued behind the requests submitted before it; blocks until done
 * (VOL_ERR_TIMEOUT after timeoutSeconds). Lets a caller pay the first-use set-up
 * before its latency-sensitive stream. */
int vol_server_prefetch(vol_server s, int meshId, double timeoutSeconds);
/* blocks until the request is done (VOL_ERR_TIMEOUT after timeoutSeconds); returns the
 * request's own status, also in *requestStatus (nullable); times (nullable, 3 doubles):
 * when it was queued, started and finished, in seconds on the server's steady clock */
int vol_server_wait(vol_server s, int64_t ticket, double timeoutSeconds, int* requestStatus, double* times);
int vol_server_stats(vol_server s, uint64_t* bytesInUse, int* queued, int* cachedEntries);
int vol_server_destroy(vol_server s); /* stops the workers after their current request */

/* ---- measurement hooks (no reference counterpart): time every level-0 A·p
 * product issued by vol_solve or vol_amul with HIP events on the context stream
 * (recorded by the launch itself, so they bracket the kernel only). */
int vol_timing_enable(vol_context ctx, int on);
int vol_timing_read(vol_context ctx, double* amulTotalMs, int64_t* amulCount, int reset);

#ifdef __cplusplus
}
#endif
#endif

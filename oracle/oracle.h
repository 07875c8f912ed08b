/* oracle.h — CPU restatement (plain C) of the reference's LDU linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker / CPU baseline;
 * the product (voluma_amd) never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * Voluma source tree). Pinning is done in tests/: against fixtures written by the
 * real reference (oracle/_ref, tests/golden/), and against closed forms, dense
 * brute force and invariants.
 */
#ifndef VOLUMA_ORACLE_H
#define VOLUMA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- exact summation (src/fvm/parallel/Comm.h:42-93, Comm.cpp:107-164) */
double oracle_exact_sum(const double* v, long n);
/* sequential plain sum (the reference's non-reproducible SumAcc, Comm.h:96-116) */
double oracle_plain_sum(const double* v, long n);

/* ---------------- LDU addressing ("mesh") handle
 * n cells, nf internal faces, owner/neighbour over internal faces, optional GAMG
 * agglomeration weights |S_f|·ndc_f (NULL: no GAMG). mergeLevels = pairings per
 * GAMG level (reference default 2, GAMG.cpp:96-99). */
void* oracle_mesh_new(int n, int nf, const int* own, const int* nei, const double* wgt, int mergeLevels);
void oracle_mesh_free(void* h);

/* GAMG hierarchy queries (GAMG.cpp:204-261) */
int oracle_gamg_nlevels(void* h);
void oracle_gamg_level_sizes(void* h, int k, int* n, int* nf);
void oracle_gamg_level(void* h, int k, int* agg, int* faceMap, int* faceFlip, int* l, int* u, double* weight);

/* y = A x (LinearSolver.cpp:121-133) */
void oracle_amul(void* h, const double* diag, const double* upper, const double* lower, const double* x, double* y);

/* Galerkin coarse coefficients of level k (GAMG.cpp:263-301); arrays sized by level */
void oracle_gamg_coeffs(void* h, const double* diag, const double* upper, const double* lower, int k, double* cdiag,
                        double* cupper, double* clower);

/* Preconditioner application. smoother: 0 = reference Gauss–Seidel V-cycle
 * (GAMG.cpp:363-449), 1 = the project's damped-Jacobi V-cycle (nPre/nPost sweeps,
 * damping omega). Returns 0 or a negative error code. */
int oracle_gamg_apply(void* h, const double* diag, const double* upper, const double* lower, const double* r, double* w,
                      int smoother, double omega, int nPre, int nPost);

/* solveLdu (LinearSolver.cpp:135-285).
 *   mode 0: plain sequential sums, names used as given (reference default)
 *   mode 1: reference reproducible mode: exact sums + preconditioner/smoother swap (:138-148)
 *   mode 2: exact sums, no swap (project extension: deterministic GAMG)
 *   gamgSmoother/omega/nPre/nPost: V-cycle variant when preconditioner == "GAMG".
 *   hist (nullable, maxIter+1 entries): finalResidual after each iteration; hist[0] = initial.
 *   perf[4] = initialResidual, finalResidual, residualSum, normFactor.
 * Returns 0, or <0 on error with a message in err (reference exceptions). */
int oracle_solve(void* h, const double* diag, const double* upper, const double* lower, int symmetric,
                 const double* b, double* x, const char* solver, const char* precond, double tol, double relTol,
                 int maxIter, int minIter, int nSweeps, int mode, int gamgSmoother, double omega, int nPre, int nPost,
                 double* perf, int* iters, int* converged, double* hist, char* err, int errLen);

/* ---------------- assembly (src/fvm/discretization/Fvm.h)
 * nI internal faces, nB boundary faces; own has nI+nB entries (boundary faces after internal).
 * Boundary condition per boundary face: bcFixed[bf] != 0 → fixedValue(bcValue[bf]), else zeroGradient
 * (src/fvm/field/BoundaryCondition.h:116-137). */
/* fvm::laplacian(gamma_f, ψ) coefficients (Fvm.h:141-154,171-180); gamma NULL → 1 */
void oracle_laplacian(int n, int nI, int nB, const int* own, const int* nei, const double* magSf, const double* ndc,
                      const double* gamma, const int* bcFixed, const double* bcValue, double* diag, double* upper,
                      double* ic, double* bs);
/* explicit non-orthogonal correction given a cell gradient g (n×3): Fvm.h:155-170 */
void oracle_laplacian_nonorth(int n, int nI, const int* own, const int* nei, const double* magSf, const double* w,
                              const double* corr, const double* gamma, const double* g, double* source, double* ffc);
/* fvc::grad Gauss linear of a scalar with boundary face values bval (Fvc.h:215-225) */
void oracle_grad_gauss(int n, int nI, int nB, const int* own, const int* nei, const double* w, const double* Sf,
                       const double* V, const double* psi, const double* bval, double* g);
/* fvm::div(F, ψ, upwind) (Fvm.h:74-106,126-135) */
void oracle_div_upwind(int n, int nI, int nB, const int* own, const int* nei, const double* F, const int* bcFixed,
                       const double* bcValue, double* diag, double* upper, double* lower, double* ic, double* bs);
/* fvm::ddt Euler (Fvm.h:18-31, Schemes.h:71-86) */
void oracle_ddt_euler(int n, const double* V, double dt, const double* psi0, double* diag, double* source);
/* boundary fold done by FvMatrix::solve before solveLdu (FvMatrix.h:154-163) */
void oracle_fold(int n, int nI, int nB, const int* own, const double* diag, const double* source, const double* ic,
                 const double* bs, double* d, double* b);
/* FvMatrix::relax (FvMatrix.h:111-131), in place on diag and source */
void oracle_relax(int n, int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
                  const double* ic, const double* psi, double alpha, double* diag, double* source);
/* FvMatrix::setReference (FvMatrix.h:134-141) */
void oracle_set_reference(int n, long cell, double value, double* diag, double* source);
/* FvMatrix::A (FvMatrix.h:179-187); nCmpt 3: vector equation */
void oracle_A(int n, int nI, int nB, const int* own, const double* diag, const double* ic, const double* V, int nCmpt,
              double* a);
/* FvMatrix::H (FvMatrix.h:189-208), one component */
void oracle_H(int n, int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
              const double* source, const double* ic, const double* bs, const double* psi, const double* V, int nCmpt,
              double* h);
/* FvMatrix::flux (FvMatrix.h:210-222), one component; ffc NULL when the matrix has none */
void oracle_flux(int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
                 const double* ic, const double* bs, const double* ffc, const double* psi, double* F);
/* explicit corrector operators (Fvc.h); multi-component values interleaved, nc 1 or 3 */
void oracle_fvc_interpolate(int n, int nI, int nB, const int* own, const int* nei, const double* w,
                            const double* skew, int nc, const double* c, const double* grad, double* F);
void oracle_fvc_grad_ls(int n, int nI, int nB, const int* own, const int* nei, const double* d, const double* Sf,
                        const double* magSf, int nc, const double* c, const int* bFix, const double* bVal, double* g);
void oracle_fvc_grad_gauss_skew(int n, int nI, int nB, const int* own, const int* nei, const double* w,
                                const double* skew, const double* d, const double* Sf, const double* magSf,
                                const double* V, const double* psi, const int* bFix, const double* bVal, double* g);
void oracle_fvc_div(int n, int nI, int nB, const int* own, const int* nei, const double* V, const double* F,
                    double* r);
void oracle_fvc_surface_integrate(int n, int nI, int nB, const int* own, const int* nei, const double* V, int nc,
                                  const double* F, double* r);
void oracle_fvc_flux(int n, int nI, int nB, const int* own, const int* nei, const double* w, const double* skew,
                     const double* Sf, const double* U, const double* grad, const int* bFix, const double* Ub,
                     double* phi);

/* Pressure-corrector steps (IncompressibleFlow.cpp:17-29,242-362) and fvc::reconstruct
 * (Fvc.h:33-67). Face arrays cover nI + nB faces, boundary arrays nB; own covers all
 * faces. */
void oracle_piso_rau_hbya(int n, const double* A, const double* H, double* rAU, double* HbyA);
void oracle_piso_hbya_boundary(int nI, int nB, const int* own, const int* bFix, const double* Ub, const double* rAU,
                               const double* gp, double* bVal);
void oracle_piso_face_op(int nF, int op, const double* a, const double* b, const double* c, double* out);
void oracle_piso_ddt_phi_corr(int nI, const int* own, const int* nei, const double* w, const double* Sf, double coeff,
                              double c0, double c00, const double* U0, const double* U00, const double* phiOld,
                              const double* phiOldOld, const double* rAUf, double* phiHbyA);
void oracle_piso_constrain_pressure(int nI, int b0, int cnt, const double* Sf, const double* magSf,
                                    const double* phiHbyA, const double* Ub, const double* rAUf, double* g);
void oracle_fvc_reconstruct(int n, int nI, int nB, const int* own, const int* nei, const double* Sf,
                            const double* magSf, const double* ssf, int emptyDir, double* v);
void oracle_piso_correct_U(int n, const double* HbyA, const double* rAU, const double* v, int sign, int emptyDir,
                           double* U);

#ifdef __cplusplus
}
#endif
#endif

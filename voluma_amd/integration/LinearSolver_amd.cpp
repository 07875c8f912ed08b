// Drop-in replacement for Voluma's src/fvm/linalg/LinearSolver.cpp and GAMG.cpp.
//
// A Voluma build that compiles this file instead of those two gets cfd::solveLdu and
// cfd::amul (src/fvm/linalg/LinearSolver.h:42-45) on the MI355X through the C ABI of
// include/voluma_amd.h. Link with libvoluma_amd.so.
//
//  * the vol_mesh (LDU addressing + GAMG hierarchy built from |S_f|·nonOrthDeltaCoeff_f,
//    GAMG.cpp:210-218) is cached on the Mesh with Mesh::cached, as the reference caches
//    its hierarchy (GAMG.cpp:204);
//  * each solve uploads diag/upper/lower (the LduSystem is a non-owning view) and copies
//    b/x; reproducible mode (par::reproducible()) selects VOL_SUMS_REPRODUCIBLE, which
//    applies the reference's swap rules (LinearSolver.cpp:138-148) and exact sums;
//  * failures become std::runtime_error with the library's message, the reference's
//    error convention (LinearSolver.cpp:76,177,280; GAMG.cpp:336);
//  * GAMG smooths as the reference does (Gauss–Seidel, 2 forward + 2 backward sweeps,
//    GAMG.cpp:423-442) and merges CFD_GAMG_MERGE pairings per level (default 2,
//    GAMG.cpp:96-99). VOLUMA_GAMG_SMOOTHER=jacobi opts in to the library's damped-Jacobi
//    V-cycle (2/2 sweeps, ω 0.8), which converges in a different number of iterations.
// Serial meshes only (nGhost == 0): the multi-GPU halo path is not part of this build.

#include "fvm/linalg/LinearSolver.h"

#include "voluma_amd.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfd {

namespace {

void check(int rc) {
    if (rc != VOL_OK) throw std::runtime_error(std::string("voluma_amd: ") + vol_last_error());
}

// Pairing passes merged per GAMG level, as the reference reads it (GAMG.cpp:96-99).
int mergeLevels() {
    const char* e = std::getenv("CFD_GAMG_MERGE");
    return e ? std::max(1, std::atoi(e)) : 2;
}

// GAMG smoother: the reference's Gauss–Seidel unless the caller opts in to Jacobi.
int gamgSmoother() {
    const char* e = std::getenv("VOLUMA_GAMG_SMOOTHER");
    if (!e || !*e || std::strcmp(e, "gaussseidel") == 0) return VOL_GAMG_GAUSS_SEIDEL;
    if (std::strcmp(e, "jacobi") == 0) return VOL_GAMG_JACOBI;
    throw std::runtime_error(std::string("voluma_amd: VOLUMA_GAMG_SMOOTHER must be gaussseidel or jacobi, not ") + e);
}

vol_context context() {
    static vol_context ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] { check(vol_context_create(0, &ctx)); });
    return ctx;
}

struct AmdHandles {
    vol_mesh mesh = nullptr;
    vol_matrix A = nullptr;
    AmdHandles() = default;
    AmdHandles(const AmdHandles&) = delete;
    AmdHandles& operator=(const AmdHandles&) = delete;
    AmdHandles(AmdHandles&& o) noexcept : mesh(o.mesh), A(o.A) { o.mesh = nullptr; o.A = nullptr; }
    ~AmdHandles() {
        if (A) vol_matrix_destroy(A);
        if (mesh) vol_mesh_destroy(mesh);
    }
};

AmdHandles buildHandles(const Mesh& m) {
    if (m.nGhost() != 0) throw std::runtime_error("voluma_amd: distributed meshes are not supported");
    const label nI = m.nInternalFaces();
    std::vector<double> w(nI);
    for (label f = 0; f < nI; ++f) w[f] = m.magSf()[f] * m.nonOrthDeltaCoeffs()[f];
    AmdHandles h;
    check(vol_mesh_create(context(), m.nCells(), nI, m.owner().data(), m.neighbour().data(), w.data(), mergeLevels(),
                          &h.mesh));
    check(vol_matrix_create(h.mesh, &h.A));
    return h;
}

AmdHandles& handles(const Mesh& m) {
    // Mesh::cached returns const T&; the handles are only used through their C handles.
    return const_cast<AmdHandles&>(m.cached<AmdHandles>("voluma_amd", [&] { return buildHandles(m); }));
}

void setMatrix(const LduSystem& s, AmdHandles& h) {
    check(vol_matrix_set(h.A, s.diag.data(), s.upper, s.symmetric ? nullptr : s.lower, VOL_HOST));
}

} // namespace

void amul(const LduSystem& s, std::vector<scalar>& x, std::vector<scalar>& y) {
    AmdHandles& h = handles(s.mesh);
    setMatrix(s, h);
    check(vol_amul(h.A, x.data(), y.data(), VOL_HOST));
}

SolverPerformance solveLdu(const LduSystem& s, std::vector<scalar>& x, const SolverControls& ctrl) {
    AmdHandles& h = handles(s.mesh);
    setMatrix(s, h);
    vol_solver_controls c;
    vol_controls_default(&c);
    c.solver = ctrl.solver.c_str();
    c.preconditioner = ctrl.preconditioner.c_str();
    c.tolerance = ctrl.tolerance;
    c.relTol = ctrl.relTol;
    c.maxIter = ctrl.maxIter;
    c.minIter = ctrl.minIter;
    c.nSweeps = ctrl.nSweeps;
    c.sums = par::reproducible() ? VOL_SUMS_REPRODUCIBLE : VOL_SUMS_FAST;
    c.gamgSmoother = gamgSmoother();
    vol_solver_performance p;
    check(vol_solve(h.A, s.b.data(), x.data(), &c, &p, VOL_HOST));
    SolverPerformance perf;
    perf.solver = p.solver;
    perf.initialResidual = p.initialResidual;
    perf.finalResidual = p.finalResidual;
    perf.residualSum = p.residualSum;
    perf.normFactor = p.normFactor;
    perf.iterations = p.iterations;
    perf.converged = p.converged != 0;
    return perf;
}

} // namespace cfd

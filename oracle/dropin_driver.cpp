// Drop-in check driver: a small Voluma program (reference mesh, fields, fvm::laplacian,
// FvMatrix::solve) that solves a Poisson problem through cfd::solveLdu. Built twice by
// oracle/Makefile: once with the reference's own LinearSolver.cpp/GAMG.cpp
// (_ref/dropin_cpu) and once with voluma_amd/integration/LinearSolver_amd.cpp
// (_ref/dropin_amd, runs on the GPU). Prints one JSON line per solve; the test compares
// the two. Contains no reference code: it only calls the reference's public API.

#include "fvm/discretization/Fvm.h"
#include "fvm/linalg/FvMatrix.h"
#include "fvm/mesh/Mesh.h"
#include "fvm/mesh/RawMesh.h"
#include "fvm/parallel/Comm.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace cfd;

static uint64_t fnv1a(const std::vector<double>& v, label n) {
    uint64_t h = 1469598103934665603ull;
    for (label i = 0; i < n; ++i) {
        unsigned char b[8];
        std::memcpy(b, &v[i], 8);
        for (int k = 0; k < 8; ++k) h = (h ^ b[k]) * 1099511628211ull;
    }
    return h;
}

static int run(int argc, char** argv);

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {  // a failed solve is reported, not an abort
        std::fprintf(stderr, "dropin: %s\n", e.what());
        return 1;
    }
}

static int run(int argc, char** argv) {
    par::Environment env(argc, argv);
    const int nx = argc > 1 ? std::atoi(argv[1]) : 24;
    BoxSpec s;
    s.n[0] = nx;
    s.n[1] = nx - 4;
    s.n[2] = nx - 8;
    RawMesh raw = generateBox(s);
    MeshPtr mesh = Mesh::build(&raw);
    struct Case {
        const char* solver;
        const char* pre;
        bool repro;
    } cases[] = {{"PCG", "diagonal", true}, {"PBiCGStab", "diagonal", true}, {"PCG", "GAMG", false},
                 {"PCG", "diagonal", false}};
    for (const auto& c : cases) {
        par::setReproducible(c.repro);
        ScalarField p(mesh, "p", 0.0);
        for (const auto& pt : mesh->patches()) {
            if (pt.name == "xMax")
                p.fixedValue(pt.name, 0.0);
            else
                p.zeroGradient(pt.name);
        }
        p.correctBoundaryConditions();
        auto M = fvm::laplacian(1.0, p, false);
        std::vector<scalar> su(mesh->nCells());
        uint64_t z = 12345;
        for (label i = 0; i < mesh->nCells(); ++i) {  // splitmix64 stream in [-1, 1)
            uint64_t q = (z += 0x9E3779B97F4A7C15ull);
            q = (q ^ (q >> 30)) * 0xBF58476D1CE4E5B9ull;
            q = (q ^ (q >> 27)) * 0x94D049BB133111EBull;
            q ^= q >> 31;
            su[i] = 2.0 * (double(q >> 11) * 0x1.0p-53) - 1.0;
        }
        M -= su;
        SolverControls ctl;
        ctl.solver = c.solver;
        ctl.preconditioner = c.pre;
        ctl.tolerance = 1e-10;
        ctl.maxIter = 2000;
        auto perf = M.solve(ctl).at(0);
        double l2 = 0;
        for (label i = 0; i < mesh->nCells(); ++i) l2 += p[i] * p[i];
        std::printf("{\"solver\": \"%s\", \"preconditioner\": \"%s\", \"reproducible\": %d, \"cells\": %d, "
                    "\"iterations\": %d, \"initial\": %.17g, \"final\": %.17g, \"l2sq\": %.17g, \"fnv\": \"%016llx\", "
                    "\"p0\": %.17g, \"pmid\": %.17g}\n",
                    c.solver, c.pre, int(c.repro), mesh->nCells(), perf.iterations, perf.initialResidual,
                    perf.finalResidual, l2, (unsigned long long)fnv1a(p.internal(), mesh->nCells()), p[0],
                    p[mesh->nCells() / 2]);
    }
    return 0;
}

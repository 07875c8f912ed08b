// Test-side driver around the reference implementation (Voluma, built serial from
// its own sources by oracle/Makefile into oracle/_ref/libvolref.so).
//
// This file contains no reference code: it only calls the reference's public API
// (generateBox, Mesh::build, fvm::laplacian/div/ddt, FvMatrix::solve, solveLdu,
// amul, buildHierarchy, GamgPreconditioner, par::ExactSum) and copies results into
// caller-provided flat arrays, so that tests/golden/make_golden.py can write
// fixtures and tests can cross-check the C oracle against the real reference.
// It is test infrastructure: nothing in the product links or loads it.

#include "fvm/discretization/Fvm.h"
#include "fvm/linalg/FvMatrix.h"
#include "fvm/linalg/GAMG.h"
#include "fvm/linalg/LinearSolver.h"
#include "fvm/mesh/Mesh.h"
#include "fvm/mesh/MeshReaders.h"
#include "fvm/mesh/RawMesh.h"
#include "fvm/parallel/Comm.h"
#include "fvm/solvers/IncompressibleFlow.h"

#include <cstring>
#include <exception>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using namespace cfd;

namespace {

struct RefMesh {
    MeshPtr m;
};

thread_local std::string gErr;

int fail(const std::exception& e) {
    gErr = e.what();
    return -1;
}

// Boundary conditions for a scalar field: fixedPatch[p] != 0 → fixedValue(value[p]),
// otherwise zeroGradient.
void setScalarBCs(ScalarField& f, const int* fixedPatch, const double* fixedValue) {
    const Mesh& m = f.mesh();
    for (std::size_t p = 0; p < m.patches().size(); ++p) {
        const auto& name = m.patches()[p].name;
        if (fixedPatch && fixedPatch[p])
            f.fixedValue(name, fixedValue ? fixedValue[p] : 0.0);
        else
            f.zeroGradient(name);
    }
    f.correctBoundaryConditions();
}

SolverControls makeControls(const char* solver, const char* precond, double tol, double relTol, int maxIter,
                            int minIter, int nSweeps) {
    SolverControls c;
    c.solver = solver;
    c.preconditioner = precond;
    c.tolerance = tol;
    c.relTol = relTol;
    c.maxIter = maxIter;
    c.minIter = minIter;
    c.nSweeps = nSweeps;
    return c;
}

void writePerf(const SolverPerformance& p, double* perf, int* iters, int* conv) {
    perf[0] = p.initialResidual;
    perf[1] = p.finalResidual;
    perf[2] = p.residualSum;
    perf[3] = p.normFactor;
    *iters = p.iterations;
    *conv = p.converged ? 1 : 0;
}

} // namespace

template <class T> static double& cmptRef(T& v, int k);
template <> double& cmptRef<scalar>(scalar& v, int) { return v; }
template <> double& cmptRef<Vec3>(Vec3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

template <class T>
static void transportPost(MeshPtr mp, int nc, int pressure, double dt, double nu, const double* psi0, const double* F,
                          const int* fixedPatch, const double* faceValues, int nonOrth, double alpha, long refCell,
                          double refValue, double* diagRelaxed, double* sourceRelaxed, double* Aout, double* Hout,
                          double* flux, double* diagRef, double* sourceRef) {
    const Mesh& m = *mp;
    const label n = m.nCells(), nI = m.nInternalFaces(), nB = m.nBoundaryFaces(), nF = m.nFaces();
    VolField<T> p(mp, "U", Traits<T>::zero());
    for (label c = 0; c < n; ++c)
        for (int k = 0; k < nc; ++k) cmptRef<T>(p[c], k) = psi0[k * n + c];
    for (std::size_t q = 0; q < m.patches().size(); ++q) {
        const auto& name = m.patches()[q].name;
        if (fixedPatch[q])
            p.fixedValue(name, Traits<T>::zero());
        else
            p.zeroGradient(name);
    }
    for (std::size_t q = 0; q < m.patches().size(); ++q) {
        auto* fv = dynamic_cast<FixedValueBC<T>*>(&p.bc(label(q)));
        if (!fv) continue;
        const Patch& pt = m.patches()[q];
        for (label i = 0; i < pt.size; ++i)
            for (int k = 0; k < nc; ++k) cmptRef<T>(fv->values()[i], k) = faceValues[k * nB + pt.start + i - nI];
    }
    p.correctBoundaryConditions();
    p.storeOld();
    TimeState ts;
    ts.dt = dt;
    std::vector<scalar> Fv(F, F + nF);
    ConvectionScheme cs;
    cs.type = ConvectionScheme::Type::Upwind;
    FvMatrix<T> M = pressure ? fvm::laplacian(nu, p, nonOrth != 0) : fvm::ddt(p, DdtScheme::Euler, ts);
    if (!pressure) {
        M += fvm::div(Fv, p, cs);
        M -= fvm::laplacian(nu, p, nonOrth != 0);
    }
    M.relax(alpha);
    auto put = [&](const std::vector<T>& v, label len, double* out) {
        for (label i = 0; i < len; ++i)
            for (int k = 0; k < nc; ++k) out[k * len + i] = cmptRef<T>(const_cast<T&>(v[i]), k);
    };
    std::copy(M.diag().begin(), M.diag().begin() + n, diagRelaxed);
    put(M.source(), n, sourceRelaxed);
    { const auto a = M.A(); std::copy(a.begin(), a.begin() + n, Aout); }
    put(M.H(), n, Hout);
    put(M.flux(), nF, flux);
    M.setReference(refCell, refValue);
    std::copy(M.diag().begin(), M.diag().begin() + n, diagRef);
    put(M.source(), n, sourceRef);
}

// A RawMesh from read(options), or null with the message kept (templates cannot sit in
// the extern "C" block).
template <class F> void* readWith(F read, double scale, double depth2D, const char* patchTypes) {
    try {
        MeshImportOptions o;
        o.scale = scale;
        o.depth2D = depth2D;
        std::istringstream ss(patchTypes ? patchTypes : "");
        std::string line;
        while (std::getline(ss, line)) {
            const auto eq = line.rfind('=');
            if (line.empty() || eq == std::string::npos) continue;
            o.patchTypes[line.substr(0, eq)] = patchTypeFromString(line.substr(eq + 1));
        }
        return new RawMesh(read(o));
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}

extern "C" {

const char* vr_last_error() { return gErr.c_str(); }

void vr_set_reproducible(int on) { par::setReproducible(on != 0); }

void* vr_box(int nx, int ny, int nz) {
    try {
        BoxSpec s;
        s.n[0] = nx;
        s.n[1] = ny;
        s.n[2] = nz;
        RawMesh raw = generateBox(s);
        auto* r = new RefMesh;
        r->m = Mesh::build(&raw);
        return r;
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}

// Raw polyMesh arrays → reference Mesh. Patch names are '\n'-separated.
void* vr_raw(long nPoints, const double* pts, long nFaces, const long* faceOff, const long* facePts, const long* owner,
             long nInternal, const long* neighbour, int nPatches, const long* pStart, const long* pSize,
             const char* names) {
    try {
        RawMesh raw;
        raw.points.resize(nPoints);
        for (long i = 0; i < nPoints; ++i) raw.points[i] = Vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        raw.faceOffsets.assign(faceOff, faceOff + nFaces + 1);
        raw.facePoints.assign(facePts, facePts + faceOff[nFaces]);
        raw.owner.assign(owner, owner + nFaces);
        raw.neighbour.assign(neighbour, neighbour + nInternal);
        std::string all(names);
        std::size_t pos = 0;
        for (int p = 0; p < nPatches; ++p) {
            RawPatch rp;
            const std::size_t e = all.find('\n', pos);
            rp.name = all.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
            pos = e == std::string::npos ? all.size() : e + 1;
            rp.type = PatchType::Patch;
            rp.start = pStart[p];
            rp.size = pSize[p];
            raw.patches.push_back(rp);
        }
        glabel nc = 0;
        for (auto o : raw.owner) nc = std::max(nc, o + 1);
        for (auto o : raw.neighbour) nc = std::max(nc, o + 1);
        raw.nCells = nc;
        auto* r = new RefMesh;
        r->m = Mesh::build(&raw);
        return r;
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}

void vr_free(void* h) { delete static_cast<RefMesh*>(h); }

void vr_sizes(void* h, long* out) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    out[0] = m.nCells();
    out[1] = m.nInternalFaces();
    out[2] = m.nFaces();
    out[3] = long(m.patches().size());
}

void vr_patch(void* h, int p, long* start, long* size) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    *start = m.patches()[p].start;
    *size = m.patches()[p].size;
}

// fvc::reconstruct (Fvc.h:33-67) of a face flux on the reference Mesh: out nCells × 3
int vr_reconstruct(void* h, const double* ssf, double* out) {
    try {
        const Mesh& m = *static_cast<RefMesh*>(h)->m;
        const std::vector<scalar> f(ssf, ssf + m.nFaces());
        const auto v = fvc::reconstruct(m, f);
        for (label c = 0; c < m.nCells(); ++c) {
            out[3 * c] = v[c].x;
            out[3 * c + 1] = v[c].y;
            out[3 * c + 2] = v[c].z;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}
// ddtCoeffs (Schemes.h:71-86): scheme 0 steady, 1 Euler, 2 backward; out c, c0, c00
void vr_ddt_coeffs(int scheme, double dt, double dt0, int haveOldOld, double* out) {
    TimeState ts;
    ts.dt = dt;
    ts.dt0 = dt0;
    const DdtScheme s = scheme == 0 ? DdtScheme::Steady : scheme == 1 ? DdtScheme::Euler : DdtScheme::Backward;
    const DdtCoeffs k = ddtCoeffs(s, ts, haveOldOld != 0);
    out[0] = k.c;
    out[1] = k.c0;
    out[2] = k.c00;
}

// nSteps time steps of the reference's IncompressibleFlow (PISO, IncompressibleFlow.cpp:
// 144-370) from U0, p0 and the face flux phi0 (kept as a restart keeps it), upwind
// convection. Per patch: U type 1 fixedValue (uValue, 3 per patch) or 0 zeroGradient; p
// type 1 fixedValue (pValue) or 0 zeroGradient. iopt: nCorr, nNonOrthCorr,
// momentumPredictor, ddt scheme (1 Euler, 2 backward), gradPScheme (0 Gauss linear,
// 1 least squares, 2 Gauss skew), laplacianNonOrthCorr, pRefCell. dopt: ddtPhiCoeff,
// pRefValue, U tolerance, p tolerance, p relTol, pFinal tolerance. Out: U, p, phi.
int vr_piso_run(void* h, int nSteps, double nu, double dt, const int* iopt, const double* dopt, const int* uType,
                const double* uValue, const int* pType, const double* pValue, const double* U0, const double* p0,
                const double* phi0, double* U, double* p, double* phi) {
    try {
        MeshPtr mesh = static_cast<RefMesh*>(h)->m;
        const Mesh& m = *mesh;
        IncompressibleFlow flow(mesh, nu);
        for (label q = 0; q < label(m.patches().size()); ++q) {
            const std::string& name = m.patches()[q].name;
            if (uType[q]) flow.U().fixedValue(name, Vec3{uValue[3 * q], uValue[3 * q + 1], uValue[3 * q + 2]});
            else flow.U().zeroGradient(name);
            if (pType[q]) flow.p().fixedValue(name, pValue[q]);
            else flow.p().zeroGradient(name);
        }
        for (label c = 0; c < m.nCells(); ++c) {
            flow.U()[c] = Vec3{U0[3 * c], U0[3 * c + 1], U0[3 * c + 2]};
            flow.p()[c] = p0[c];
        }
        flow.phi().assign(phi0, phi0 + m.nFaces());
        PimpleControls& ct = flow.controls;
        ct.nCorr = iopt[0];
        ct.nNonOrthCorr = iopt[1];
        ct.momentumPredictor = iopt[2] != 0;
        flow.ddtScheme = iopt[3] == 2 ? DdtScheme::Backward : DdtScheme::Euler;
        flow.gradPScheme = iopt[4] == 1 ? GradScheme::LeastSquares : iopt[4] == 2 ? GradScheme::GaussSkew
                                                                                   : GradScheme::GaussLinear;
        flow.laplacianNonOrthCorr = iopt[5] != 0;
        ct.pRefCell = iopt[6];
        ct.ddtPhiCoeff = dopt[0];
        ct.pRefValue = dopt[1];
        ct.UControls = SolverControls{"PBiCGStab", "DILU", dopt[2], 0.0, 1000, 0, 1};
        ct.pControls = SolverControls{"PCG", "GAMG", dopt[3], dopt[4], 2000, 0, 1};
        ct.pFinalControls = SolverControls{"PCG", "GAMG", dopt[5], 0.0, 2000, 0, 1};
        flow.divScheme = ConvectionScheme::parse("upwind");
        flow.markRestarted();
        for (int k = 0; k < nSteps; ++k) flow.step(dt);
        for (label c = 0; c < m.nCells(); ++c) {
            U[3 * c] = flow.U()[c].x;
            U[3 * c + 1] = flow.U()[c].y;
            U[3 * c + 2] = flow.U()[c].z;
            p[c] = flow.p()[c];
        }
        std::copy(flow.phi().begin(), flow.phi().end(), phi);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// twoD, emptyDir, number of solved components, the solved components (Mesh.h:102-105)
void vr_layout(void* h, long* out) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    out[0] = m.twoD() ? 1 : 0;
    out[1] = m.emptyDir();
    out[2] = long(m.solvedComponents().size());
    for (std::size_t k = 0; k < m.solvedComponents().size() && k < 3; ++k) out[3 + k] = m.solvedComponents()[k];
}
void vr_patch_name(void* h, int p, char* out, int len) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    std::strncpy(out, m.patches()[p].name.c_str(), std::size_t(len) - 1);
    out[len - 1] = 0;
}

void vr_addressing(void* h, int* owner, int* neighbour) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    std::memcpy(owner, m.owner().data(), sizeof(int) * m.nFaces());
    std::memcpy(neighbour, m.neighbour().data(), sizeof(int) * m.nInternalFaces());
}

static void copyVec(const std::vector<Vec3>& v, double* out) {
    if (!out) return;
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}
static void copySc(const std::vector<double>& v, double* out) {
    if (out) std::memcpy(out, v.data(), sizeof(double) * v.size());
}

void vr_geometry(void* h, double* V, double* C, double* Sf, double* magSf, double* Cf, double* d, double* w,
                 double* ndc, double* corr) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    copySc(m.V(), V);
    copyVec(m.C(), C);
    copyVec(m.Sf(), Sf);
    copySc(m.magSf(), magSf);
    copyVec(m.Cf(), Cf);
    copyVec(m.d(), d);
    copySc(m.w(), w);
    copySc(m.nonOrthDeltaCoeffs(), ndc);
    copyVec(m.corrVecs(), corr);
}

// polyMesh directory read by the reference (readPolyMesh, RawMesh.cpp:223-336)
void* vr_read_polymesh(const char* dir) {
    try {
        return new RawMesh(readPolyMesh(dir));
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}
// Mesh files read by the reference's importers (MeshReaders.cpp): readGmsh
// (:236-482) and readMeshFile (:805-816), with MeshImportOptions
// {scale, depth2D, patchTypes from "name=type" lines (patchTypeFromString)}.
void* vr_read_gmsh(const char* file, double scale, double depth2D, const char* patchTypes) {
    return readWith([&](const MeshImportOptions& o) { return readGmsh(file, o); }, scale, depth2D, patchTypes);
}
void* vr_read_mesh_file(const char* path, double scale, double depth2D, const char* patchTypes) {
    return readWith([&](const MeshImportOptions& o) { return readMeshFile(path, o); }, scale, depth2D, patchTypes);
}
// generateBox (RawMesh.cpp:391-502) on the unit cube: periodic bits 1 x, 2 y, 4 z; twoD
void* vr_generate_box(int nx, int ny, int nz, int periodic, int twoD) {
    try {
        BoxSpec b;
        b.n[0] = nx;
        b.n[1] = ny;
        b.n[2] = nz;
        for (int d = 0; d < 3; ++d) b.periodic[d] = (periodic >> d & 1) != 0;
        b.twoD = twoD != 0;
        return new RawMesh(generateBox(b));
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}
// replace the point coordinates of a RawMesh (same count) before it is built
int vr_rawmesh_set_points(void* h, const double* pts, long nPoints) {
    RawMesh& r = *static_cast<RawMesh*>(h);
    if (nPoints != long(r.points.size())) return -1;
    for (long i = 0; i < nPoints; ++i) r.points[i] = Vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    return 0;
}
// nPoints, nFaces, nInternalFaces, nFacePoints, nCells, nPatches
void vr_rawmesh_sizes(void* h, long* out) {
    const RawMesh& r = *static_cast<RawMesh*>(h);
    out[0] = long(r.points.size());
    out[1] = long(r.owner.size());
    out[2] = long(r.neighbour.size());
    out[3] = long(r.facePoints.size());
    out[4] = long(r.nCells);
    out[5] = long(r.patches.size());
}
// names / types: 64 / 32 bytes per patch
void vr_rawmesh_arrays(void* h, double* pts, long* faceOff, long* facePts, long* owner, long* nei, long* pStart,
                       long* pSize, char* names, char* types) {
    const RawMesh& r = *static_cast<RawMesh*>(h);
    for (std::size_t i = 0; i < r.points.size(); ++i) {
        pts[3 * i] = r.points[i].x;
        pts[3 * i + 1] = r.points[i].y;
        pts[3 * i + 2] = r.points[i].z;
    }
    std::copy(r.faceOffsets.begin(), r.faceOffsets.end(), faceOff);
    std::copy(r.facePoints.begin(), r.facePoints.end(), facePts);
    std::copy(r.owner.begin(), r.owner.end(), owner);
    std::copy(r.neighbour.begin(), r.neighbour.end(), nei);
    for (std::size_t p = 0; p < r.patches.size(); ++p) {
        pStart[p] = long(r.patches[p].start);
        pSize[p] = long(r.patches[p].size);
        std::strncpy(names + 64 * p, r.patches[p].name.c_str(), 63);
        names[64 * p + 63] = 0;
        std::strncpy(types + 32 * p, toString(r.patches[p].type).c_str(), 31);
        types[32 * p + 31] = 0;
    }
}
// the reference Mesh built from it (Mesh::build), for vr_geometry / vr_skew
void* vr_rawmesh_build(void* h) {
    try {
        RawMesh raw = *static_cast<RawMesh*>(h);
        auto* r = new RefMesh;
        r->m = Mesh::build(&raw);
        return r;
    } catch (const std::exception& e) {
        fail(e);
        return nullptr;
    }
}
void vr_rawmesh_free(void* h) { delete static_cast<RawMesh*>(h); }

void vr_skew(void* h, double* skew, int* skewed) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    copyVec(m.skewVecs(), skew);
    *skewed = m.skewed() ? 1 : 0;
}

// The explicit operators of the pressure corrector (Fvc.h) on a scalar p and a vector U
// (interleaved xyz) with fixedValue patches (face values pB / UB per boundary face,
// interleaved for U) and zeroGradient elsewhere, and a face flux F:
//   gLSp, gLSU   fvc::gradLS (value faces: fixed patches)            Fvc.h:185-206
//   ip           fvc::interpolate(m, p)                              Fvc.h:24-31
//   ipSkew, iU   fvc::interpolate(m, c, gradLS(c)) (skew-corrected)  Fvc.h:72-81
//   gGL, gLS, gGS  fvc::grad(pField, GaussLinear / LeastSquares / GaussSkew)  Fvc.h:208-291
//   divF         fvc::div(m, F)                                      Fvc.h:294-305
//   siU          fvc::surfaceIntegrate(m, iU)                        Fvc.h:307-318
//   phi          dot(iU_f, Sf) on internal faces; on boundary faces dot(UB, Sf) on the
//                fixed patches and dot(U_P, Sf) elsewhere (phiHbyA, IncompressibleFlow.cpp:276-291)
int vr_fvc(void* h, const double* p, const double* U, const int* fixedPatch, const double* pB, const double* UB,
           const double* F, double* gLSp, double* gLSU, double* ip, double* ipSkew, double* iU, double* gGL,
           double* gLS, double* gGS, double* divF, double* siU, double* phi) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        const Mesh& m = *mp;
        const label n = m.nCells(), nI = m.nInternalFaces(), nB = m.nBoundaryFaces(), nF = m.nFaces();
        std::vector<scalar> pc(p, p + n);
        std::vector<Vec3> Uc(n);
        for (label c = 0; c < n; ++c) Uc[c] = Vec3{U[3 * c], U[3 * c + 1], U[3 * c + 2]};
        std::vector<char> bFix(nB, 0);
        std::vector<scalar> pBv(pB, pB + nB);
        std::vector<Vec3> UBv(nB);
        for (std::size_t q = 0; q < m.patches().size(); ++q) {
            const Patch& pt = m.patches()[q];
            for (label i = 0; i < pt.size; ++i) bFix[pt.start + i - nI] = fixedPatch[q] ? 1 : 0;
        }
        for (label b = 0; b < nB; ++b) UBv[b] = Vec3{UB[3 * b], UB[3 * b + 1], UB[3 * b + 2]};
        const auto gp = fvc::gradLS(m, pc, bFix, pBv);
        const auto gU = fvc::gradLS(m, Uc, bFix, UBv);
        for (label c = 0; c < n; ++c) {
            gLSp[3 * c] = gp[c].x;
            gLSp[3 * c + 1] = gp[c].y;
            gLSp[3 * c + 2] = gp[c].z;
            for (int k = 0; k < 9; ++k) gLSU[9 * c + k] = (&gU[c].xx)[k];
        }
        copySc(fvc::interpolate(m, pc), ip);
        copySc(fvc::interpolate(m, pc, gp), ipSkew);
        const auto iUv = fvc::interpolate(m, Uc, gU);
        copyVec(iUv, iU);
        ScalarField pf(mp, "p", 0.0);
        for (label c = 0; c < n; ++c) pf[c] = p[c];
        for (std::size_t q = 0; q < m.patches().size(); ++q) {
            const auto& name = m.patches()[q].name;
            if (fixedPatch[q]) {
                pf.fixedValue(name, 0.0);
                auto* fv = dynamic_cast<FixedValueBC<scalar>*>(&pf.bc(label(q)));
                const Patch& pt = m.patches()[q];
                for (label i = 0; i < pt.size; ++i) fv->values()[i] = pB[pt.start + i - nI];
            } else {
                pf.zeroGradient(name);
            }
        }
        pf.correctBoundaryConditions();
        auto put3 = [&](const std::vector<Vec3>& g, double* out) {
            for (label c = 0; c < n; ++c) {
                out[3 * c] = g[c].x;
                out[3 * c + 1] = g[c].y;
                out[3 * c + 2] = g[c].z;
            }
        };
        put3(fvc::grad(pf, GradScheme::GaussLinear), gGL);
        put3(fvc::grad(pf, GradScheme::LeastSquares), gLS);
        put3(fvc::grad(pf, GradScheme::GaussSkew), gGS);
        const std::vector<scalar> Fv(F, F + nF);
        const auto dv = fvc::div(m, Fv);
        std::copy(dv.begin(), dv.begin() + n, divF);
        put3(fvc::surfaceIntegrate(m, iUv), siU);
        for (label f = 0; f < nI; ++f) phi[f] = dot(iUv[f], m.Sf()[f]);
        for (label f = nI; f < nF; ++f)
            phi[f] = bFix[f - nI] ? dot(UBv[f - nI], m.Sf()[f]) : dot(Uc[m.owner()[f]], m.Sf()[f]);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void vr_cellfaces(void* h, int* off, int* list) {
    const Mesh& m = *static_cast<RefMesh*>(h)->m;
    std::memcpy(off, m.cellFaceOffsets().data(), sizeof(int) * m.cellFaceOffsets().size());
    std::memcpy(list, m.cellFaces().data(), sizeof(int) * m.cellFaces().size());
}

// fvm::laplacian(gamma (uniform), p) for a scalar p with the given field values (used
// by the non-orthogonal correction's gradient) and fixedValue/zeroGradient patches.
int vr_laplacian(void* h, double gamma, const int* fixedPatch, const double* fixedValue, const double* psi,
                 int nonOrth, int gradScheme, double* diag, double* upper, double* source, double* ic, double* bs,
                 double* ffc) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        mp->setGradScheme(GradScheme(gradScheme));
        ScalarField p(mp, "p", 0.0);
        if (psi)
            for (label c = 0; c < mp->nCells(); ++c) p[c] = psi[c];
        setScalarBCs(p, fixedPatch, fixedValue);
        auto M = fvm::laplacian(gamma, p, nonOrth != 0);
        copySc(M.diag(), diag);
        copySc(M.upper(), upper);
        copySc(M.source(), source);
        copySc(M.internalCoeffs(), ic);
        copySc(M.boundarySource(), bs);
        if (ffc) {
            if (M.hasFaceFluxCorrection())
                copySc(M.faceFluxCorrection(), ffc);
            else
                std::memset(ffc, 0, sizeof(double) * mp->nFaces());
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// fvm::div(F, psi, upwind) for a scalar psi with fixedValue/zeroGradient patches.
int vr_div_upwind(void* h, const double* F, const int* fixedPatch, const double* fixedValue, double* diag,
                  double* upper, double* lower, double* source, double* ic, double* bs) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        ScalarField p(mp, "U", 0.0);
        setScalarBCs(p, fixedPatch, fixedValue);
        std::vector<scalar> Fv(F, F + mp->nFaces());
        ConvectionScheme s;
        s.type = ConvectionScheme::Type::Upwind;
        auto M = fvm::div(Fv, p, s);
        copySc(M.diag(), diag);
        copySc(M.upper(), upper);
        std::memcpy(lower, M.lowerPtr(), sizeof(double) * mp->nInternalFaces());
        copySc(M.source(), source);
        copySc(M.internalCoeffs(), ic);
        copySc(M.boundarySource(), bs);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// fvm::ddt(psi) with the Euler scheme and step dt; psi0 is the old-time field.
int vr_ddt_euler(void* h, double dt, const double* psi0, double* diag, double* source) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        ScalarField p(mp, "U", 0.0);
        for (label c = 0; c < mp->nCells(); ++c) p[c] = psi0[c];
        setScalarBCs(p, nullptr, nullptr);
        p.storeOld();
        TimeState ts;
        ts.dt = dt;
        auto M = fvm::ddt(p, DdtScheme::Euler, ts);
        copySc(M.diag(), diag);
        copySc(M.source(), source);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// One component of the momentum matrix as IncompressibleFlow::step composes it
// (IncompressibleFlow.cpp:183-185): fvm::ddt (Euler) + fvm::div(F, upwind) −
// fvm::laplacian(nu, nonOrth), for a scalar field with old and current value psi0.
// The FvMatrix parts are returned as they stand, then the boundary fold that
// FvMatrix::solve applies (FvMatrix.h:154-163) into dF/bF.
// faceValues (nullable, nBoundaryFaces): per-face values of the fixedValue patches.
int vr_transport(void* h, double dt, double nu, const double* psi0, const double* F, const int* fixedPatch,
                 const double* fixedValue, const double* faceValues, int nonOrth, double* diag, double* upper,
                 double* lower, double* source, double* ic, double* bs, double* ffc, double* dF, double* bF) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        const Mesh& m = *mp;
        ScalarField p(mp, "U", 0.0);
        for (label c = 0; c < m.nCells(); ++c) p[c] = psi0[c];
        setScalarBCs(p, fixedPatch, fixedValue);
        if (faceValues) {
            for (std::size_t q = 0; q < m.patches().size(); ++q) {
                auto* fv = dynamic_cast<FixedValueBC<scalar>*>(&p.bc(label(q)));
                if (!fv) continue;
                const Patch& pt = m.patches()[q];
                for (label i = 0; i < pt.size; ++i) fv->values()[i] = faceValues[pt.start + i - m.nInternalFaces()];
            }
            p.correctBoundaryConditions();
        }
        p.storeOld();
        TimeState ts;
        ts.dt = dt;
        std::vector<scalar> Fv(F, F + m.nFaces());
        ConvectionScheme cs;
        cs.type = ConvectionScheme::Type::Upwind;
        auto M = fvm::ddt(p, DdtScheme::Euler, ts);
        M += fvm::div(Fv, p, cs);
        M -= fvm::laplacian(nu, p, nonOrth != 0);
        copySc(M.diag(), diag);
        copySc(M.upper(), upper);
        std::memcpy(lower, M.lowerPtr(), sizeof(double) * m.nInternalFaces());
        copySc(M.source(), source);
        copySc(M.internalCoeffs(), ic);
        copySc(M.boundarySource(), bs);
        if (M.hasFaceFluxCorrection())
            copySc(M.faceFluxCorrection(), ffc);
        else
            std::memset(ffc, 0, sizeof(double) * m.nFaces());
        for (label c = 0; c < m.nCells(); ++c) {
            dF[c] = M.diag()[c];
            bF[c] = M.source()[c];
        }
        const label nI = m.nInternalFaces();
        for (label bf = 0; bf < m.nBoundaryFaces(); ++bf) {
            const label c = m.owner()[nI + bf];
            dF[c] += M.internalCoeffs()[bf];
            bF[c] += M.boundarySource()[bf];
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// The momentum matrix of vr_transport followed by FvMatrix::relax(alpha) (FvMatrix.h:111-131)
// and the pressure-velocity coupling quantities A(), H(), flux() (FvMatrix.h:179-222) of
// the relaxed matrix, then setReference(refCell, refValue) (FvMatrix.h:134-141) on it.
// pressure != 0: the matrix is fvm::laplacian(nu) alone (a pressure equation, symmetric).
// nCmpt 1: a scalar field; 3: a vector field (component arrays one after another: psi0
// and sources n per component, faceValues nBoundary per component, flux nFaces).
int vr_transport_post(void* h, int nCmpt, int pressure, double dt, double nu, const double* psi0, const double* F,
                      const int* fixedPatch, const double* faceValues, int nonOrth, double alpha, long refCell,
                      double refValue, double* diagRelaxed, double* sourceRelaxed, double* Aout, double* Hout,
                      double* flux, double* diagRef, double* sourceRef) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        if (nCmpt == 1)
            transportPost<scalar>(mp, 1, pressure, dt, nu, psi0, F, fixedPatch, faceValues, nonOrth, alpha, refCell, refValue,
                                  diagRelaxed, sourceRelaxed, Aout, Hout, flux, diagRef, sourceRef);
        else if (nCmpt == 3)
            transportPost<Vec3>(mp, 3, pressure, dt, nu, psi0, F, fixedPatch, faceValues, nonOrth, alpha, refCell, refValue,
                                diagRelaxed, sourceRelaxed, Aout, Hout, flux, diagRef, sourceRef);
        else
            throw std::runtime_error("nCmpt must be 1 or 3");
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// solveLdu on a raw LDU system (diag already boundary-folded). lower may equal upper.
int vr_solve(void* h, const double* diag, const double* upper, const double* lower, int symmetric, const double* b,
             double* x, const char* solver, const char* precond, double tol, double relTol, int maxIter, int minIter,
             int nSweeps, double* perf, int* iters, int* conv) {
    try {
        const Mesh& m = *static_cast<RefMesh*>(h)->m;
        const label n = m.nTotalCells();
        std::vector<scalar> d(diag, diag + n), bb(b, b + n), xx(x, x + n);
        LduSystem sys{m, d, upper, lower, bb, symmetric != 0};
        auto p = solveLdu(sys, xx, makeControls(solver, precond, tol, relTol, maxIter, minIter, nSweeps));
        std::memcpy(x, xx.data(), sizeof(double) * n);
        writePerf(p, perf, iters, conv);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// FvMatrix<scalar>::solve for laplacian(gamma, p) − su (source += su·V) with the given
// patches: exercises the boundary fold done inside FvMatrix::solve.
int vr_fvm_poisson_solve(void* h, double gamma, const int* fixedPatch, const double* fixedValue, const double* su,
                         double* psi, const char* solver, const char* precond, double tol, double relTol, int maxIter,
                         double* perf, int* iters, int* conv) {
    try {
        MeshPtr mp = static_cast<RefMesh*>(h)->m;
        ScalarField p(mp, "p", 0.0);
        for (label c = 0; c < mp->nCells(); ++c) p[c] = psi[c];
        setScalarBCs(p, fixedPatch, fixedValue);
        auto M = fvm::laplacian(gamma, p, false);
        std::vector<scalar> s(su, su + mp->nCells());
        M -= s;
        auto out = M.solve(makeControls(solver, precond, tol, relTol, maxIter, 0, 1));
        for (label c = 0; c < mp->nCells(); ++c) psi[c] = p[c];
        writePerf(out.at(0), perf, iters, conv);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int vr_amul(void* h, const double* diag, const double* upper, const double* lower, const double* x, double* y) {
    try {
        const Mesh& m = *static_cast<RefMesh*>(h)->m;
        const label n = m.nTotalCells();
        std::vector<scalar> d(diag, diag + n), bb(n, 0.0), xx(x, x + n), yy(n);
        LduSystem sys{m, d, upper, lower, bb, upper == lower};
        amul(sys, xx, yy);
        std::memcpy(y, yy.data(), sizeof(double) * n);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// GAMG hierarchy. Call with out arrays == nullptr first to get nLevels and per-level
// sizes (levelN[k], levelNf[k]); then call again with arrays of total length Σ n
// (agg) and Σ nf (faceMap, faceFlip, l, u) and Σ nf doubles (weight).
int vr_gamg_hierarchy(void* h, int* nLevels, long* levelN, long* levelNf, int* agg, int* faceMap, int* faceFlip,
                      int* lArr, int* uArr, double* weight) {
    try {
        const Mesh& m = *static_cast<RefMesh*>(h)->m;
        const GamgHierarchy& H = gamgHierarchy(m);
        *nLevels = int(H.levels.size());
        long oa = 0, of = 0;
        for (std::size_t k = 0; k < H.levels.size(); ++k) {
            const auto& L = H.levels[k];
            if (levelN) levelN[k] = L.n;
            if (levelNf) levelNf[k] = long(L.l.size());
            if (agg) {
                for (std::size_t i = 0; i < L.agg.size(); ++i) agg[oa + i] = L.agg[i];
                for (std::size_t i = L.agg.size(); i < std::size_t(L.n); ++i) agg[oa + i] = -1;
            }
            if (faceMap)
                for (std::size_t f = 0; f < L.l.size(); ++f) {
                    faceMap[of + f] = f < L.faceMap.size() ? L.faceMap[f] : -2;
                    faceFlip[of + f] = f < L.faceFlip.size() ? L.faceFlip[f] : 0;
                    lArr[of + f] = L.l[f];
                    uArr[of + f] = L.u[f];
                    weight[of + f] = L.weight[f];
                }
            oa += L.n;
            of += long(L.l.size());
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// w = M⁻¹ r for the reference GAMG preconditioner (one V-cycle, Gauss–Seidel smoothing).
int vr_gamg_apply(void* h, const double* diag, const double* upper, const double* lower, const double* r, double* w) {
    try {
        const Mesh& m = *static_cast<RefMesh*>(h)->m;
        const label n = m.nTotalCells();
        std::vector<scalar> d(diag, diag + n), bb(n, 0.0), rr(r, r + n), ww(n, 0.0);
        LduSystem sys{m, d, upper, lower, bb, upper == lower};
        GamgPreconditioner P(sys);
        P.apply(rr, ww);
        std::memcpy(w, ww.data(), sizeof(double) * n);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

double vr_exact_sum(const double* v, long n) {
    par::ExactSum s;
    for (long i = 0; i < n; ++i) s.add(v[i]);
    return s.value();
}

} // extern "C"

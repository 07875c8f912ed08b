// Incompressible flow time steps resident on the GPU: the reference's
// IncompressibleFlow::step (src/fvm/solvers/IncompressibleFlow.cpp:144-370) for laminar
// flow with upwind convection, fixedValue / zeroGradient boundaries and no body force,
// composed of the library's own entry points on device pointers. The steps are taken in
// the reference's order with the reference's operations, so a step reproduces the
// reference's fields (bit for bit with reproducible sums, where both sides use the same
// partition-independent solvers).
//
// Fields: U interleaved (nCells × 3, as std::vector<Vec3> holds it) plus one contiguous
// copy per component for the per-component matrices and solves; p; φ over the kept faces
// (internal, then boundary). Old levels U⁰, U⁰⁰, φ⁰, φ⁰⁰ as VolField::storeOld and the
// solver keep them (VolField.h:108-117, IncompressibleFlow.cpp:153-160).
#include "host_common.hpp"

#include <cmath>

namespace {

void chk(int rc) {
    if (rc != VOL_OK) throw VolError(rc, vol_last_error());
}

// ddtCoeffs (Schemes.h:71-86)
void ddtCoeffs(int scheme, double dt, double dt0, bool haveOldOld, double& c, double& c0, double& c00) {
    const double rDt = 1.0 / dt;
    if (scheme == VOL_DDT_BACKWARD && haveOldOld && dt0 > 0) {
        const double ct = 1.0 + dt / (dt + dt0);
        const double ct00 = dt * dt / (dt0 * (dt + dt0));
        c = ct * rDt;
        c0 = (ct + ct00) * rDt;
        c00 = ct00 * rDt;
    } else {
        c = c0 = rDt;
        c00 = 0.0;
    }
}

}  // namespace

struct vol_flow_s {
    vol_mesh mesh = nullptr;
    vol_context ctx = nullptr;
    hipStream_t st = nullptr;
    int n = 0, ni = 0, nb = 0, nF = 0;
    vol_flow_options o{};
    bool needRef = true;
    vol_matrix UEqn[3] = {nullptr, nullptr, nullptr}, P = nullptr;
    // the mesh's geometry (owned by the mesh, which outlives the flow) and what the flow
    // needs beyond it (kept faces)
    const double *V = nullptr, *Sf = nullptr, *magSf = nullptr, *w = nullptr;
    DevBuf<double> ndc, corr;
    int emptyDir = -1;
    std::vector<int> solved;
    // boundary conditions
    DevBuf<int32_t> uFix, pFix, zeroFix;
    DevBuf<double> uVal, uValK[3], pVal, zeroVal;
    // fields and time levels
    DevBuf<double> U, Uc[3], p, phi, U0, U00, phiOld, phiOldOld;
    int nOld = 0, timeIndex = 0;
    double dt = -1.0;
    bool havePhiOldOld = false;
    // scratch
    DevBuf<double> Ub, pb, gp, su, b, Acell, H, Hk, rAU, HbyA, gradRAU, rAUf, bValH, gH, phiHbyA, divPhi, F, nuF;
    DevBuf<double> u0k, u00k;  // a component of U⁰ / U⁰⁰

    ~vol_flow_s() {
        for (auto& A : UEqn)
            if (A) vol_matrix_destroy(A);
        if (P) vol_matrix_destroy(P);
    }
    void copy(double* dst, const double* src, size_t cnt) {
        if (cnt) HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToDevice, st));
    }
    // component k of the interleaved field into its contiguous copy, and back
    void splitU() {
        for (int k = 0; k < 3; ++k) chk(vol_copy_strided(ctx, n, U.p + k, 3, Uc[k].p, 1, 1.0));
    }
    void joinU() {
        for (int k = 0; k < 3; ++k) chk(vol_copy_strided(ctx, n, Uc[k].p, 1, U.p + k, 3, 1.0));
    }
    void boundaryU() { chk(vol_boundary_values(mesh, 3, uFix.p, uVal.p, U.p, Ub.p, VOL_DEVICE)); }
    void boundaryP() { chk(vol_boundary_values(mesh, 1, pFix.p, pVal.p, p.p, pb.p, VOL_DEVICE)); }
    void gradP() {  // fvc::grad(p_, gradPScheme) with p's boundary values
        boundaryP();
        chk(vol_fvc_grad(mesh, o.gradPScheme, 1, p.p, pFix.p, pb.p, gp.p, VOL_DEVICE));
    }
    vol_nonorth nonorth(const double* psi) const {
        vol_nonorth no;
        no.psi = psi;
        no.w = w;
        no.Sf = Sf;
        no.corrVecs = corr.p;
        no.V = V;
        return no;
    }
    void solve(vol_matrix A, const double* rhs, double* x, const vol_solver_controls& c,
               std::vector<vol_solver_performance>& perf) {
        vol_solver_performance pf;
        chk(vol_solve(A, rhs, x, &c, &pf, VOL_DEVICE));
        perf.push_back(pf);
    }

    void step(double dtNew, std::vector<vol_solver_performance>& perf) {
        // new time level (:153-160)
        const double dt0 = timeIndex > 0 ? dt : -1.0;
        dt = dtNew;
        ++timeIndex;
        if (nOld >= 1) copy(U00.p, U0.p, 3 * size_t(n));
        copy(U0.p, U.p, 3 * size_t(n));
        nOld = std::min(nOld + 1, 2);
        havePhiOldOld = timeIndex > 1;  // φ⁰⁰ = the previous step's φ⁰
        if (havePhiOldOld) copy(phiOldOld.p, phiOld.p, nF);
        copy(phiOld.p, phi.p, nF);
        const double* U00p = nOld >= 2 ? U00.p : U0.p;  // VolField::oldOld()

        // momentum matrix per component (:183-185); relax(1) leaves it unchanged
        double c, c0, c00;
        ddtCoeffs(o.ddtScheme, dt, dt0, nOld >= 2, c, c0, c00);
        for (int k = 0; k < 3; ++k) {
            vol_matrix A = UEqn[k];
            chk(vol_copy_strided(ctx, n, U0.p + k, 3, u0k.p, 1, 1.0));
            chk(vol_copy_strided(ctx, n, U00p + k, 3, u00k.p, 1, 1.0));
            chk(vol_fvm_reset(A));
            chk(vol_fvm_ddt(A, +1, c, c0, c00, V, u0k.p, u00k.p, VOL_DEVICE));
            chk(vol_fvm_div_upwind(A, +1, phi.p, uFix.p, uValK[k].p, VOL_DEVICE));
            const vol_nonorth no = nonorth(Uc[k].p);
            chk(vol_fvm_laplacian(A, -1, 1.0, nuF.p, magSf, ndc.p, uFix.p, uValK[k].p,
                                  o.laplacianNonOrthCorr ? &no : nullptr, VOL_DEVICE));
        }
        // momentum predictor: M = UEqn; M −= −∇p; solve the solved components (:212-227)
        if (o.momentumPredictor) {
            gradP();
            for (int k : solved) {
                chk(vol_copy_strided(ctx, n, gp.p + k, 3, su.p, 1, -1.0));
                chk(vol_fvm_finish_source(UEqn[k], su.p, V, b.p, VOL_DEVICE));
                solve(UEqn[k], b.p, Uc[k].p, o.UControls, perf);
            }
            joinU();
        }

        // pressure correctors (:242-362)
        for (int corr = 0; corr < o.nCorr; ++corr) {
            const bool finalCorr = corr == o.nCorr - 1;
            chk(vol_fvm_A(UEqn[0], V, 3, Acell.p, VOL_DEVICE));
            for (int k = 0; k < 3; ++k) {
                chk(vol_fvm_H(UEqn[k], Uc[k].p, V, 3, Hk.p, VOL_DEVICE));
                chk(vol_copy_strided(ctx, n, Hk.p, 1, H.p + k, 3, 1.0));
            }
            chk(vol_piso_rau_hbya(mesh, Acell.p, H.p, rAU.p, HbyA.p, VOL_DEVICE));
            const bool skewed = meshSkewed;
            if (skewed) {
                chk(vol_fvc_grad(mesh, VOL_GRAD_LEAST_SQUARES, 1, rAU.p, zeroFix.p, zeroVal.p, gradRAU.p, VOL_DEVICE));
                chk(vol_fvc_interpolate(mesh, 1, rAU.p, gradRAU.p, nullptr, rAUf.p, VOL_DEVICE));
            } else {
                chk(vol_fvc_interpolate(mesh, 1, rAU.p, nullptr, nullptr, rAUf.p, VOL_DEVICE));
            }
            boundaryU();
            const double* gHp = nullptr;
            if (skewed) {
                gradP();
                chk(vol_piso_hbya_boundary(mesh, uFix.p, Ub.p, rAU.p, gp.p, bValH.p, VOL_DEVICE));
                chk(vol_fvc_grad(mesh, VOL_GRAD_LEAST_SQUARES, 3, HbyA.p, uFix.p, bValH.p, gH.p, VOL_DEVICE));
                gHp = gH.p;
            }
            chk(vol_fvc_flux(mesh, HbyA.p, gHp, uFix.p, Ub.p, phiHbyA.p, VOL_DEVICE));
            if (o.ddtPhiCoeff != 0.0) {
                double k, k0, k00;
                ddtCoeffs(o.ddtScheme, dt, dt0, nOld >= 2 && havePhiOldOld, k, k0, k00);
                chk(vol_piso_ddt_phi_corr(mesh, o.ddtPhiCoeff, k0, k00, U0.p, U00p, phiOld.p,
                                          havePhiOldOld ? phiOldOld.p : nullptr, rAUf.p, phiHbyA.p, VOL_DEVICE));
            }
            chk(vol_fvc_div(mesh, phiHbyA.p, divPhi.p, VOL_DEVICE));
            for (int no = 0; no <= o.nNonOrthCorr; ++no) {
                const bool finalNO = no == o.nNonOrthCorr;
                const vol_nonorth nn = nonorth(p.p);
                chk(vol_fvm_reset(P));
                chk(vol_fvm_laplacian(P, +1, 1.0, rAUf.p, magSf, ndc.p, pFix.p, pVal.p,
                                      o.laplacianNonOrthCorr ? &nn : nullptr, VOL_DEVICE));
                chk(vol_fvm_source(P, divPhi.p, V, VOL_DEVICE));
                if (needRef) chk(vol_fvm_set_reference(P, o.pRefCell, o.pRefValue));
                chk(vol_fvm_finish(P, b.p, VOL_DEVICE));
                solve(P, b.p, p.p, (finalCorr && finalNO) ? o.pFinalControls : o.pControls, perf);
                if (finalNO) {
                    chk(vol_fvm_flux(P, p.p, F.p, VOL_DEVICE));
                    chk(vol_piso_face_op(mesh, VOL_FACE_A_MINUS_B, phiHbyA.p, F.p, nullptr, phi.p, VOL_DEVICE));
                }
            }
            gradP();
            chk(vol_piso_correct_U(mesh, HbyA.p, rAU.p, gp.p, -1, emptyDir, U.p, VOL_DEVICE));
            splitU();
        }
        HIPCHECK(hipStreamSynchronize(st));
    }
    bool meshSkewed = false;
};

extern "C" {

void vol_flow_options_default(vol_flow_options* o) {
    if (!o) return;
    *o = vol_flow_options{};
    o->nu = 0.0;
    o->nCorr = 2;
    o->nNonOrthCorr = 0;
    o->momentumPredictor = 1;
    o->ddtScheme = VOL_DDT_BACKWARD;
    o->gradPScheme = VOL_GRAD_GAUSS_LINEAR;
    o->laplacianNonOrthCorr = 1;
    o->pRefCell = 0;
    o->pRefValue = 0.0;
    o->ddtPhiCoeff = 0.0;
    vol_controls_default(&o->UControls);
    o->UControls.solver = "PBiCGStab";
    o->UControls.preconditioner = "DILU";
    o->UControls.tolerance = 1e-8;
    vol_controls_default(&o->pControls);
    o->pControls.solver = "PCG";
    o->pControls.preconditioner = "GAMG";
    o->pControls.tolerance = 1e-8;
    o->pControls.relTol = 0.01;
    o->pControls.maxIter = 2000;
    o->pControls.gamgSmoother = VOL_GAMG_GAUSS_SEIDEL;
    o->pFinalControls = o->pControls;
    o->pFinalControls.relTol = 0.0;
}

int vol_flow_create(vol_mesh mesh, const vol_flow_options* options, const double* ndc, const double* corrVecs,
                    const int32_t* uFixed, const double* uValue, const int32_t* pFixed, const double* pValue,
                    int location, vol_flow* out) {
    API_BEGIN
    if (!mesh || !options || !ndc || !corrVecs || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const vol_flow_options& o = *options;
    if (!(o.nu >= 0.0) || o.nCorr < 1 || o.nNonOrthCorr < 0 ||
        (o.ddtScheme != VOL_DDT_EULER && o.ddtScheme != VOL_DDT_BACKWARD))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad flow options");
    auto f = std::make_unique<vol_flow_s>();
    f->mesh = mesh;
    f->ctx = meshContext(mesh);
    f->st = f->ctx->stream;
    meshSizes(mesh, f->n, f->ni, f->nb);
    if (f->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    f->nF = f->ni + f->nb;
    f->o = o;
    HIPCHECK(hipSetDevice(f->ctx->device));
    const size_t n = f->n, nF = f->nF, nb = f->nb;
    // mesh geometry as stored by vol_mesh_set_geometry, copied so the flow owns it
    int skewed = 0, emptyDir = -1;
    meshGeometry(mesh, f->V, f->Sf, f->magSf, f->w, skewed, emptyDir);
    f->meshSkewed = skewed != 0;
    f->emptyDir = emptyDir;
    for (int k = 0; k < 3; ++k)
        if (k != emptyDir) f->solved.push_back(k);
    auto up = [&](DevBuf<double>& d, const double* s, size_t cnt) {
        d.alloc(std::max<size_t>(cnt, 1));
        if (cnt)
            HIPCHECK(hipMemcpyAsync(d.p, s, sizeof(double) * cnt,
                                    location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->st));
    };
    auto upI = [&](DevBuf<int32_t>& d, const int32_t* s, size_t cnt) {
        d.alloc(std::max<size_t>(cnt, 1));
        if (cnt)
            HIPCHECK(hipMemcpyAsync(d.p, s, sizeof(int32_t) * cnt,
                                    location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->st));
    };
    if (nb > 0 && (!uFixed || !uValue || !pFixed || !pValue))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "boundary conditions required");
    up(f->ndc, ndc, nF);
    up(f->corr, corrVecs, 3 * nF);
    upI(f->uFix, uFixed, nb);
    up(f->uVal, uValue, 3 * nb);
    upI(f->pFix, pFixed, nb);
    up(f->pVal, pValue, nb);
    std::vector<int32_t> hp(nb, 0);
    if (nb) {
        if (location == VOL_DEVICE)
            HIPCHECK(hipMemcpyAsync(hp.data(), pFixed, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, f->st));
        else
            std::copy(pFixed, pFixed + nb, hp.begin());
        HIPCHECK(hipStreamSynchronize(f->st));
    }
    f->needRef = std::none_of(hp.begin(), hp.end(), [](int32_t v) { return v != 0; });  // :109-113
    for (int k = 0; k < 3; ++k) {
        f->uValK[k].alloc(std::max<size_t>(nb, 1));
        chk(vol_copy_strided(f->ctx, int64_t(nb), f->uVal.p + k, 3, f->uValK[k].p, 1, 1.0));
    }
    f->zeroFix.alloc(std::max<size_t>(nb, 1));
    HIPCHECK(hipMemsetAsync(f->zeroFix.p, 0, sizeof(int32_t) * std::max<size_t>(nb, 1), f->st));
    f->zeroVal.alloc(std::max<size_t>(nb, 1));
    HIPCHECK(hipMemsetAsync(f->zeroVal.p, 0, sizeof(double) * std::max<size_t>(nb, 1), f->st));
    std::vector<double> nu(nF, o.nu);  // nuEffFaces() without a turbulence model (:83-85)
    f->nuF.alloc(std::max<size_t>(nF, 1));
    if (nF) HIPCHECK(hipMemcpyAsync(f->nuF.p, nu.data(), sizeof(double) * nF, hipMemcpyHostToDevice, f->st));
    for (auto* d : {&f->U, &f->U0, &f->U00, &f->H, &f->HbyA, &f->gp, &f->gradRAU, &f->Ub}) d->alloc(3 * n + 3 * nb);
    for (auto* d : {&f->p, &f->su, &f->b, &f->Acell, &f->Hk, &f->rAU, &f->divPhi, &f->u0k, &f->u00k}) d->alloc(n);
    for (auto& c : f->Uc) c.alloc(n);
    for (auto* d : {&f->phi, &f->phiOld, &f->phiOldOld, &f->rAUf, &f->phiHbyA, &f->F}) d->alloc(nF);
    f->pb.alloc(std::max<size_t>(nb, 1));
    f->bValH.alloc(std::max<size_t>(3 * nb, 1));
    f->gH.alloc(9 * n);
    for (auto* d : {&f->U, &f->p, &f->phi}) HIPCHECK(hipMemsetAsync(d->p, 0, sizeof(double) * d->n, f->st));
    for (auto& c : f->Uc) HIPCHECK(hipMemsetAsync(c.p, 0, sizeof(double) * n, f->st));
    for (auto& A : f->UEqn) chk(vol_matrix_create(mesh, &A));
    chk(vol_matrix_create(mesh, &f->P));
    HIPCHECK(hipStreamSynchronize(f->st));
    *out = f.release();
    API_END
}

int vol_flow_set_fields(vol_flow f, const double* U, const double* p, const double* phi, int location) {
    API_BEGIN
    if (!f) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null flow");
    HIPCHECK(hipSetDevice(f->ctx->device));
    const hipMemcpyKind k = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (U && f->n) HIPCHECK(hipMemcpyAsync(f->U.p, U, sizeof(double) * 3 * f->n, k, f->st));
    if (p && f->n) HIPCHECK(hipMemcpyAsync(f->p.p, p, sizeof(double) * f->n, k, f->st));
    if (phi && f->nF) HIPCHECK(hipMemcpyAsync(f->phi.p, phi, sizeof(double) * f->nF, k, f->st));
    if (U) f->splitU();
    HIPCHECK(hipStreamSynchronize(f->st));
    API_END
}

int vol_flow_get_fields(vol_flow f, double* U, double* p, double* phi, int location) {
    API_BEGIN
    if (!f) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null flow");
    HIPCHECK(hipSetDevice(f->ctx->device));
    const hipMemcpyKind k = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (U && f->n) HIPCHECK(hipMemcpyAsync(U, f->U.p, sizeof(double) * 3 * f->n, k, f->st));
    if (p && f->n) HIPCHECK(hipMemcpyAsync(p, f->p.p, sizeof(double) * f->n, k, f->st));
    if (phi && f->nF) HIPCHECK(hipMemcpyAsync(phi, f->phi.p, sizeof(double) * f->nF, k, f->st));
    HIPCHECK(hipStreamSynchronize(f->st));
    API_END
}

int vol_flow_step(vol_flow f, double dt, vol_solver_performance* perf, int maxPerf, int* nPerf) {
    API_BEGIN
    if (!f) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null flow");
    if (!(dt > 0.0)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "dt must be > 0");
    HIPCHECK(hipSetDevice(f->ctx->device));
    std::vector<vol_solver_performance> pf;
    f->step(dt, pf);
    if (perf)
        for (int i = 0; i < int(pf.size()) && i < maxPerf; ++i) perf[i] = pf[i];
    if (nPerf) *nPerf = int(pf.size());
    API_END
}

int vol_flow_destroy(vol_flow f) {
    delete f;
    return VOL_OK;
}

}  // extern "C"

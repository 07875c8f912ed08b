// Coefficient assembly on the GPU: fvm::laplacian and fvm::div (upwind) for a scalar field with fixedValue / zeroGradient patches, followed by the boundary
// fold that FvMatrix::solve performs before solveLdu.
//
// The reference scatters face coefficients into the diagonal in face order
// (src/fvm/discretization/Fvm.h:88-106,149-154). Here each cell gathers its faces in
// ascending order through the SELL slot list of the mesh (the same per-cell order), so
// the result is bit-identical and needs no atomics.
#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

// c_f = γ_f |S_f| ndc_f (Fvm.h:150); γ uniform when gamma == nullptr.
__device__ __forceinline__ double lap_coef(const AsmArgs& a, int f) {
    const double g = a.gamma ? a.gamma[f] : a.gammaU;
    return g * a.magSf[f] * a.ndc[f];
}

__global__ void k_lap_faces(AsmArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    a.upper[f] = lap_coef(a, f);
}

// Boundary terms of fvm::laplacian (Fvm.h:171-180) with fixedValue gic = −ndc_b,
// gbc = ndc_b·v and zeroGradient gic = gbc = 0 (BoundaryCondition.h:116-117,128-137).
__device__ __forceinline__ void lap_boundary(const AsmArgs& a, int bf, double& ic, double& bs) {
    const int f = a.nf + bf;
    const double g = a.gamma ? a.gamma[f] : a.gammaU;
    const double gs = g * a.magSf[f];
    double gic = 0.0, gbc = 0.0;
    if (a.bcFixed[bf]) {
        gic = -a.ndc[f];
        gbc = a.ndc[f] * (a.bcValue ? a.bcValue[bf] : 0.0);
    }
    ic = 0.0 + gs * gic;
    bs = 0.0 - gs * gbc;
}

// Upwind convection (Fvm.h:88-106): lower = −w_up F, upper = lower + F.
__device__ __forceinline__ void div_face(const AsmArgs& a, int f, double& lo, double& up) {
    const double F = a.flux[f];
    const double wUp = F >= 0 ? 1.0 : 0.0;
    lo = -wUp * F;
    up = lo + F;
}
__device__ __forceinline__ void div_boundary(const AsmArgs& a, int bf, double& ic, double& bs) {
    const double F = a.flux[a.nf + bf];
    const bool fixed = a.bcFixed[bf] != 0;
    const double vic = fixed ? 0.0 : 1.0;
    const double vbc = fixed ? (a.bcValue ? a.bcValue[bf] : 0.0) : 0.0;
    ic = 0.0 + F * vic;  // Fvm.h:132
    bs = 0.0 - F * vbc;  // Fvm.h:133
}

__global__ void k_div_faces(AsmArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    double lo, up;
    div_face(a, f, lo, up);
    a.upper[f] = up;
    a.lower[f] = lo;
}

// Per cell: diagonal from its faces in ascending order, then the boundary fold in
// ascending boundary-face order (FvMatrix.h:159-163); b = bIn + folded boundary source.
template <int KIND>
__global__ void k_asm_cells(AsmArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const int slice = c / WAVE, lane = c % WAVE;
    const int s0 = a.S.sliceOff[slice], width = (a.S.sliceOff[slice + 1] - s0) / WAVE;
    double d = 0.0;
    for (int k = 0; k < width; ++k) {
        const int e = a.S.src[s0 + k * WAVE + lane];
        if (e < 0) continue;
        const int f = e >> 1;
        if constexpr (KIND == ASM_LAPLACIAN) {
            d -= lap_coef(a, f);  // Fvm.h:152-153
        } else {
            double lo, up;
            div_face(a, f, lo, up);
            d -= (e & 1) ? up : lo;  // owner row: −lower, neighbour row: −upper (Fvm.h:105-106)
        }
    }
    double bsum = a.bIn ? a.bIn[c] : 0.0;
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k];
        double ic, bs;
        if constexpr (KIND == ASM_LAPLACIAN)
            lap_boundary(a, bf, ic, bs);
        else
            div_boundary(a, bf, ic, bs);
        d += ic;
        bsum += bs;
        if (a.icOut) a.icOut[bf] = ic;
        if (a.bsOut) a.bsOut[bf] = bs;
    }
    a.diag[c] = d;
    if (a.bOut) a.bOut[c] = bsum;
}

void launch_assemble(hipStream_t st, int kind, const AsmArgs& a) {
    if (a.nf > 0) {
        if (kind == ASM_LAPLACIAN)
            hipLaunchKernelGGL(k_lap_faces, dim3((a.nf + 255) / 256), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(k_div_faces, dim3((a.nf + 255) / 256), dim3(256), 0, st, a);
    }
    if (a.n > 0) {
        if (kind == ASM_LAPLACIAN)
            hipLaunchKernelGGL(k_asm_cells<ASM_LAPLACIAN>, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(k_asm_cells<ASM_DIV>, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
    }
}

// ============================================================================ FvMatrix terms
// Laplacian coefficient of the term (γ_f |S_f| ndc_f).
__device__ __forceinline__ double fvm_lap_coef(const FvmArgs& a, int f) {
    const double g = a.gamma ? a.gamma[f] : a.gammaU;
    return g * a.magSf[f] * a.ndc[f];
}

// Fvm.h:157: any internal face with |corr|² > 1e-20 switches the correction on.
__global__ void k_any_corr(FvmArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    const double cx = a.corr[3 * f], cy = a.corr[3 * f + 1], cz = a.corr[3 * f + 2];
    if (cx * cx + cy * cy + cz * cz > 1e-20) atomicOr(a.anyCorr, 1);
}

// Gauss-linear gradient (Fvc.h:216-224): per cell its internal faces in ascending
// order (owner adds, neighbour subtracts Sf·φ_f with φ_f = w φ_P + (1−w) φ_N), then its
// boundary faces (Sf·φ_b), then the product with 1/V.
__global__ void k_grad_gauss(FvmArgs a) {
    if (!*a.anyCorr) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const int slice = c / WAVE, lane = c % WAVE;
    const int s0 = a.S.sliceOff[slice], width = (a.S.sliceOff[slice + 1] - s0) / WAVE;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int k = 0; k < width; ++k) {
        const int slot = s0 + k * WAVE + lane;
        const int e = a.S.src[slot];
        if (e < 0) continue;
        const int f = e >> 1, nb = a.S.col[slot];
        const bool isOwner = (e & 1) == 0;
        const double wf = a.w[f];
        const double phif = isOwner ? wf * a.psi[c] + (1.0 - wf) * a.psi[nb] : wf * a.psi[nb] + (1.0 - wf) * a.psi[c];
        const double vx = a.Sf[3 * f] * phif, vy = a.Sf[3 * f + 1] * phif, vz = a.Sf[3 * f + 2] * phif;
        if (isOwner) {
            gx += vx;
            gy += vy;
            gz += vz;
        } else {
            gx -= vx;
            gy -= vy;
            gz -= vz;
        }
    }
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k], f = a.nf + bf;
        const double pb = a.bcFixed[bf] ? (a.bcValue ? a.bcValue[bf] : 0.0) : a.psi[c];
        gx += a.Sf[3 * f] * pb;
        gy += a.Sf[3 * f + 1] * pb;
        gz += a.Sf[3 * f + 2] * pb;
    }
    const double r = 1.0 / a.V[c];
    a.grad[3 * c] = a.emptyDir == 0 ? 0.0 : gx * r;  // 2-D: empty component zeroed (Fvc.h:169-181)
    a.grad[3 * c + 1] = a.emptyDir == 1 ? 0.0 : gy * r;
    a.grad[3 * c + 2] = a.emptyDir == 2 ? 0.0 : gz * r;
}

// Off-diagonal part of a term and its addition (FvMatrix.h:232-239): the lower part
// is first copied from upper when the matrix has none yet, then lower += s·lower_term
// (the term's upper when it has no lower part) and upper += s·upper_term. The
// laplacian's correction flux fc (Fvm.h:162-167) is formed here and added to ffc.
template <int KIND>
__global__ void k_fvm_faces(FvmArgs a) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    double tu = 0.0, tl = 0.0;
    constexpr bool termLower = KIND == FVM_DIV;
    if constexpr (KIND == FVM_DIV) {
        const double F = a.flux[f];
        const double wUp = F >= 0 ? 1.0 : 0.0;
        tl = -wUp * F;
        tu = tl + F;
    } else if constexpr (KIND == FVM_LAP) {
        tu = fvm_lap_coef(a, f);
    }
    const double up = a.upper[f];
    if (a.lower) a.lower[f] = (a.hadLower ? a.lower[f] : up) + a.sign * (termLower ? tl : tu);
    a.upper[f] = up + a.sign * tu;
}

// Correction flux per internal face (Fvm.h:162-165): gf = w g_P + (1−w) g_N,
// fc = (γ|S|)·(corr·gf); ffc += s·fc (FvMatrix.h:240-244).
__global__ void k_fvm_fc(FvmArgs a, const int* own, const int* nei) {
    if (!*a.anyCorr) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    const int o = own[f], nb = nei[f];
    const double wf = a.w[f], wn = 1.0 - wf;
    const double gx = wf * a.grad[3 * o] + wn * a.grad[3 * nb];
    const double gy = wf * a.grad[3 * o + 1] + wn * a.grad[3 * nb + 1];
    const double gz = wf * a.grad[3 * o + 2] + wn * a.grad[3 * nb + 2];
    const double g = a.gamma ? a.gamma[f] : a.gammaU;
    const double fc = (g * a.magSf[f]) * (a.corr[3 * f] * gx + a.corr[3 * f + 1] * gy + a.corr[3 * f + 2] * gz);
    a.fc[f] = fc;
    a.ffc[f] = a.ffc[f] + a.sign * fc;
}

// Per cell: the term's diagonal and source from its faces in ascending order and its
// boundary faces' internalCoeffs/boundarySource, then the signed addition.
template <int KIND>
__global__ void k_fvm_cells(FvmArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    double td = 0.0, ts = 0.0;
    if constexpr (KIND == FVM_DDT) {
        const double Vc = a.V[c];
        td = 0.0 + a.ddtC * Vc;                                              // Fvm.h:27
        ts = 0.0 + Vc * (a.ddtC0 * a.psi0[c] - a.ddtC00 * a.psi00[c]);       // Fvm.h:28
    } else {
        const int slice = c / WAVE, lane = c % WAVE;
        const int s0 = a.S.sliceOff[slice], width = (a.S.sliceOff[slice + 1] - s0) / WAVE;
        const bool corr = KIND == FVM_LAP && a.psi && *a.anyCorr;
        for (int k = 0; k < width; ++k) {
            const int e = a.S.src[s0 + k * WAVE + lane];
            if (e < 0) continue;
            const int f = e >> 1;
            const bool isOwner = (e & 1) == 0;
            if constexpr (KIND == FVM_DIV) {
                const double F = a.flux[f];
                const double wUp = F >= 0 ? 1.0 : 0.0;
                const double lo = -wUp * F;
                td -= isOwner ? lo : lo + F;                       // Fvm.h:105-106
            } else {
                td -= fvm_lap_coef(a, f);                          // Fvm.h:152-153
                if (corr) {
                    if (isOwner)
                        ts -= a.fc[f];                             // Fvm.h:166
                    else
                        ts += a.fc[f];                             // Fvm.h:167
                }
            }
        }
    }
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k], f = a.nf + bf;
        double tic = 0.0, tbs = 0.0;
        if constexpr (KIND == FVM_DIV) {                           // Fvm.h:126-135
            const double F = a.flux[f];
            const bool fixed = a.bcFixed[bf] != 0;
            tic = 0.0 + F * (fixed ? 0.0 : 1.0);
            tbs = 0.0 - F * (fixed ? (a.bcValue ? a.bcValue[bf] : 0.0) : 0.0);
        } else if constexpr (KIND == FVM_LAP) {                    // Fvm.h:171-180
            const double g = a.gamma ? a.gamma[f] : a.gammaU;
            const double gs = g * a.magSf[f];
            double gic = 0.0, gbc = 0.0;
            if (a.bcFixed[bf]) {
                gic = -a.ndc[f];
                gbc = a.ndc[f] * (a.bcValue ? a.bcValue[bf] : 0.0);
            }
            tic = 0.0 + gs * gic;
            tbs = 0.0 - gs * gbc;
        }
        a.ic[bf] = a.ic[bf] + a.sign * tic;
        a.bs[bf] = a.bs[bf] + a.sign * tbs;
    }
    a.diag[c] = a.diag[c] + a.sign * td;
    a.source[c] = a.source[c] + a.sign * ts;
}

__global__ void k_fvm_source(int n, const double* su, const double* V, double* source) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) source[c] += su[c] * V[c];
}

__global__ void k_fvm_fold(FvmArgs a, double* dOut, double* bOut) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    double d = a.diag[c], b = a.source[c];
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k];
        d += a.ic[bf];
        b += a.bs[bf];
    }
    dOut[c] = d;
    if (bOut) bOut[c] = b;
}

// ============================================================================ pressure-velocity coupling
// cmptAv of a boundary coefficient (FvMatrix.h:227-228): the value for a scalar equation,
// (v + v + v)/3 for a vector equation with the same coefficient in every component.
__device__ __forceinline__ double cmpt_av(double v, int nCmpt) { return nCmpt == 3 ? (v + v + v) / 3.0 : v; }

// FvMatrix::relax (FvMatrix.h:111-131). Per cell: Σ|offdiag| over its faces in ascending
// order (|upper| on owner rows, |lower| on neighbour rows, as the reference's face loop
// adds them), Σ ic over its boundary faces in order, then D = max(|diag + Σ ic|, Σ|offdiag|),
// diag += D/α − (diag + Σ ic), source += that increment · ψ.
__global__ void k_fvm_relax(FvmArgs a, double alpha, const double* psi, const double* lower) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const int slice = c / WAVE, lane = c % WAVE;
    const int s0 = a.S.sliceOff[slice], width = (a.S.sliceOff[slice + 1] - s0) / WAVE;
    double sumOff = 0.0;
    for (int k = 0; k < width; ++k) {
        const int e = a.S.src[s0 + k * WAVE + lane];
        if (e < 0) continue;
        const int f = e >> 1;
        sumOff += fabs((e & 1) ? lower[f] : a.upper[f]);
    }
    double bD = 0.0;
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) bD += a.ic[a.bList[k]];
    const double d = a.diag[c];
    const double full = d + bD;
    const double mag = fabs(full);
    const double D = mag < sumOff ? sumOff : mag;  // std::max(|full|, sumOff)
    const double inc = D / alpha - full;
    a.diag[c] = d + inc;
    a.source[c] = a.source[c] + inc * psi[c];
}

// FvMatrix::setReference (FvMatrix.h:134-141) at one cell (checked on the host).
__global__ void k_fvm_set_reference(double* diag, double* source, int c, double value) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double d = diag[c];
    source[c] = source[c] + d * value;
    diag[c] = d + d;
}

// FvMatrix::A (FvMatrix.h:179-187): (diag + Σ_b cmptAv(ic)) / V.
__global__ void k_fvm_A(FvmArgs a, const double* V, int nCmpt, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    double d = a.diag[c];
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) d += cmpt_av(a.ic[a.bList[k]], nCmpt);
    out[c] = d / V[c];
}

// FvMatrix::H (FvMatrix.h:189-208): per cell source − Σ offdiag·ψ_nb over its faces in
// ascending order (owner rows upper·ψ_N, neighbour rows lower·ψ_P), then per boundary
// face + bs − (ic − cmptAv ic)·ψ, then / V.
__global__ void k_fvm_H(FvmArgs a, const double* lower, const double* psi, const double* V, int nCmpt, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const int slice = c / WAVE, lane = c % WAVE;
    const int s0 = a.S.sliceOff[slice], width = (a.S.sliceOff[slice + 1] - s0) / WAVE;
    double h = a.source[c];
    for (int k = 0; k < width; ++k) {
        const int slot = s0 + k * WAVE + lane;
        const int e = a.S.src[slot];
        if (e < 0) continue;
        const int f = e >> 1;
        h -= ((e & 1) ? lower[f] : a.upper[f]) * psi[a.S.col[slot]];
    }
    const double pc = psi[c];
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k];
        const double ic = a.ic[bf];
        h += a.bs[bf];
        h -= (ic - cmpt_av(ic, nCmpt)) * pc;
    }
    out[c] = h / V[c];
}

// FvMatrix::flux (FvMatrix.h:210-222) on the internal faces: upper·ψ_N − lower·ψ_P (+ ffc).
__global__ void k_fvm_flux_faces(FvmArgs a, const double* lower, const int* own, const int* nei, const double* psi,
                                 int hasFfc, double* out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    double F = a.upper[f] * psi[nei[f]] - lower[f] * psi[own[f]];
    if (hasFfc) F = F + a.ffc[f];
    out[f] = F;
}

// ... and on the boundary faces, per owner cell: ic·ψ_P − bs (+ ffc).
__global__ void k_fvm_flux_boundary(FvmArgs a, const double* psi, int hasFfc, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const double pc = psi[c];
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k];
        double F = a.ic[bf] * pc - a.bs[bf];
        if (hasFfc) F = F + a.ffc[a.nf + bf];
        out[a.nf + bf] = F;
    }
}

// fold of a copy of the matrix with an extra source term −= su (FvMatrix.h:93-97):
// b = (source + su·V) + Σ bs per cell; the matrix's own source is left as it is.
__global__ void k_fvm_fold_extra(FvmArgs a, const double* su, const double* V, double* dOut, double* bOut) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    double d = a.diag[c], b = a.source[c] + su[c] * V[c];
    for (int k = a.bOff[c]; k < a.bOff[c + 1]; ++k) {
        const int bf = a.bList[k];
        d += a.ic[bf];
        b += a.bs[bf];
    }
    dOut[c] = d;
    bOut[c] = b;
}

void launch_fvm_term(hipStream_t st, int kind, const FvmArgs& a, const int* own, const int* nei) {
    const dim3 fb((a.nf + 255) / 256), cb((a.n + 255) / 256), t(256);
    if (kind == FVM_LAP && a.psi) {
        (void)hipMemsetAsync(a.anyCorr, 0, sizeof(int), st);
        if (a.nf > 0) hipLaunchKernelGGL(k_any_corr, fb, t, 0, st, a);
        if (a.n > 0) hipLaunchKernelGGL(k_grad_gauss, cb, t, 0, st, a);
        if (a.nf > 0) hipLaunchKernelGGL(k_fvm_fc, fb, t, 0, st, a, own, nei);
    }
    if (a.nf > 0) {
        if (kind == FVM_DDT) hipLaunchKernelGGL(k_fvm_faces<FVM_DDT>, fb, t, 0, st, a);
        if (kind == FVM_DIV) hipLaunchKernelGGL(k_fvm_faces<FVM_DIV>, fb, t, 0, st, a);
        if (kind == FVM_LAP) hipLaunchKernelGGL(k_fvm_faces<FVM_LAP>, fb, t, 0, st, a);
    }
    if (a.n > 0) {
        if (kind == FVM_DDT) hipLaunchKernelGGL(k_fvm_cells<FVM_DDT>, cb, t, 0, st, a);
        if (kind == FVM_DIV) hipLaunchKernelGGL(k_fvm_cells<FVM_DIV>, cb, t, 0, st, a);
        if (kind == FVM_LAP) hipLaunchKernelGGL(k_fvm_cells<FVM_LAP>, cb, t, 0, st, a);
    }
}

void launch_fvm_source(hipStream_t st, int n, const double* su, const double* V, double* source) {
    if (n > 0) hipLaunchKernelGGL(k_fvm_source, dim3((n + 255) / 256), dim3(256), 0, st, n, su, V, source);
}

void launch_fvm_fold(hipStream_t st, const FvmArgs& a, double* dOut, double* bOut, const double* su, const double* V) {
    if (a.n <= 0) return;
    if (su)
        hipLaunchKernelGGL(k_fvm_fold_extra, dim3((a.n + 255) / 256), dim3(256), 0, st, a, su, V, dOut, bOut);
    else
        hipLaunchKernelGGL(k_fvm_fold, dim3((a.n + 255) / 256), dim3(256), 0, st, a, dOut, bOut);
}

void launch_fvm_relax(hipStream_t st, const FvmArgs& a, double alpha, const double* psi, const double* lower) {
    if (a.n > 0) hipLaunchKernelGGL(k_fvm_relax, dim3((a.n + 255) / 256), dim3(256), 0, st, a, alpha, psi, lower);
}

void launch_fvm_set_reference(hipStream_t st, double* diag, double* source, int cell, double value) {
    hipLaunchKernelGGL(k_fvm_set_reference, dim3(1), dim3(64), 0, st, diag, source, cell, value);
}

void launch_fvm_A(hipStream_t st, const FvmArgs& a, const double* V, int nCmpt, double* out) {
    if (a.n > 0) hipLaunchKernelGGL(k_fvm_A, dim3((a.n + 255) / 256), dim3(256), 0, st, a, V, nCmpt, out);
}

void launch_fvm_H(hipStream_t st, const FvmArgs& a, const double* lower, const double* psi, const double* V, int nCmpt,
                  double* out) {
    if (a.n > 0) hipLaunchKernelGGL(k_fvm_H, dim3((a.n + 255) / 256), dim3(256), 0, st, a, lower, psi, V, nCmpt, out);
}

void launch_fvm_flux(hipStream_t st, const FvmArgs& a, const double* lower, const int* own, const int* nei,
                     const double* psi, int hasFfc, double* out) {
    if (a.nf > 0)
        hipLaunchKernelGGL(k_fvm_flux_faces, dim3((a.nf + 255) / 256), dim3(256), 0, st, a, lower, own, nei, psi,
                           hasFfc, out);
    if (a.n > 0 && a.nb > 0)
        hipLaunchKernelGGL(k_fvm_flux_boundary, dim3((a.n + 255) / 256), dim3(256), 0, st, a, psi, hasFfc, out);
}

}  // namespace vol

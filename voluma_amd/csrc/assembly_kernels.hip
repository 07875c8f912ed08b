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

}  // namespace vol

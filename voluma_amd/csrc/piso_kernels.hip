// The element-wise and face-wise steps of the pressure corrector that sit between the
// FvMatrix and fvc operators (IncompressibleFlow.cpp:242-362), and fvc::reconstruct
// (Fvc.h:33-67):
//   rAU = 1/A, HbyA = rAU·H (:245-249);
//   the fixed-flux boundary values of HbyA for its least-squares gradient,
//   U_b + rAU_P ∇p_P (:261-270);
//   phiHbyA += rAUf·φ_F, φ_F − ∇p_f·S_f style differences and φ = phiHbyA − F (:294-295,
//   :219-220, :338-339, :351-352);
//   the ddtPhiCorr coupling (:296-316);
//   constrainPressure's fixed-flux gradient (:17-29);
//   U = HbyA ∓ rAU·v with the empty direction zeroed in 2-D (:353-360).
// Every value is formed with the reference's operations in its order; a cell's
// reconstruct sums run over its faces in ascending order (internal, then boundary), as
// the reference's face loop reaches them, so all results are bit-identical.
#include "device_common.hpp"
#include "vol_internal.hpp"
#include "voluma_amd.h"

namespace vol {

namespace {
constexpr int PISO_THREADS = 256;
constexpr double PISO_SMALL = 1.0e-15;  // SMALL (Types.h:18)
inline dim3 piso_grid(long long cnt) { return dim3(unsigned((cnt + PISO_THREADS - 1) / PISO_THREADS)); }
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
}  // namespace

__global__ void k_piso_rau_hbya(int n, const double* A, const double* H, double* rAU, double* HbyA) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double r = 1.0 / A[c];
    rAU[c] = r;
#pragma unroll
    for (int k = 0; k < 3; ++k) HbyA[3 * c + k] = r * H[3 * c + k];
}

// per cell over its boundary faces: fixed-flux faces U_b + rAU_P ∇p_P, others zero
__global__ void k_piso_hbya_boundary(FvcGeom G, const int* bFix, const double* Ub, const double* rAU,
                                     const double* gp, double* bVal) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.n) return;
    for (int k = G.bOff[c]; k < G.bOff[c + 1]; ++k) {
        const int bf = G.bList[k];
#pragma unroll
        for (int q = 0; q < 3; ++q) bVal[3 * bf + q] = bFix[bf] ? Ub[3 * bf + q] + rAU[c] * gp[3 * c + q] : 0.0;
    }
}

template <int OP>
__global__ void k_piso_face_op(int nF, const double* a, const double* b, const double* c, double* out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    if constexpr (OP == VOL_FACE_A_PLUS_B_TIMES_C) out[f] = a[f] + b[f] * c[f];
    else if constexpr (OP == VOL_FACE_A_MINUS_B) out[f] = a[f] - b[f];
    else out[f] = a[f] - b[f] / c[f];
}

// ddtPhiCorr on internal faces: corr0 = φ⁰ − (w U⁰_P + (1−w) U⁰_N)·S; coupling = coeff,
// or 1 − min(|corr0| / (|φ⁰| + SMALL), 1) for coeff < 0; dc = c0·corr0 − c00·(φ⁰⁰ − U⁰⁰_f·S)
// (the second term only when c00 ≠ 0); phiHbyA += rAUf·coupling·dc.
__global__ void k_piso_ddt_phi_corr(FvcGeom G, double coeff, double c0, double c00, const double* U0,
                                    const double* U00, const double* phiOld, const double* phiOldOld,
                                    const double* rAUf, double* phiHbyA) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= G.nf) return;
    const double wf = G.w[f];
    const int o = G.own[f], n = G.nei[f];
    const double* S = G.Sf + 3 * f;
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = wf * U0[3 * o + k] + (1.0 - wf) * U0[3 * n + k];
    const double corr0 = phiOld[f] - dot3(v, S);
    double coupling = coeff;
    if (coeff < 0) {
        const double q = fabs(corr0) / (fabs(phiOld[f]) + PISO_SMALL);
        coupling = 1.0 - (1.0 < q ? 1.0 : q);  // 1 − std::min(q, 1)
    }
    double dc = c0 * corr0;
    if (c00 != 0.0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = wf * U00[3 * o + k] + (1.0 - wf) * U00[3 * n + k];
        dc -= c00 * (phiOldOld[f] - dot3(v, S));
    }
    phiHbyA[f] += rAUf[f] * coupling * dc;
}

// fixedFluxPressure gradient of boundary faces [b0, b0 + cnt): (phiHbyA − U_b·S) / (|S| rAUf)
__global__ void k_piso_constrain_pressure(FvcGeom G, int b0, int cnt, const double* phiHbyA, const double* Ub,
                                          const double* rAUf, double* g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const int bf = b0 + i, f = G.nf + bf;
    g[i] = (phiHbyA[f] - dot3(Ub + 3 * bf, G.Sf + 3 * f)) / (G.magSf[f] * rAUf[f]);
}

// reconstruct: A = Σ S Sᵀ/|S|, b = Σ (φ/|S|) S over the cell's faces; in 2-D 1 added to the
// empty direction's diagonal; v = adj(A) b / det (zero when |det| < 1e-300).
namespace {
__device__ __forceinline__ void rec_add(const double* S, double ms, double phi, double* A, double* b) {
    const double r = 1.0 / ms;  // a Tensor divided by a scalar is multiplied by its reciprocal (Types.h:86)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) A[3 * i + j] += S[i] * S[j] * r;
    const double s = phi / ms;
#pragma unroll
    for (int i = 0; i < 3; ++i) b[i] += s * S[i];
}
}  // namespace

__global__ void k_fvc_reconstruct(FvcGeom G, const double* ssf, int emptyDir, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.n) return;
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, b[3] = {0, 0, 0};
    int s0, width, stride;
    {
        const SellView& S = G.S;
        if (S.csrG > 0) {
            s0 = S.sliceOff[c];
            width = S.sliceOff[c + 1] - s0;
            stride = 1;
        } else if (S.ellWidth > 0) {
            width = S.ellWidth;
            s0 = (c / WAVE) * WAVE * width + c % WAVE;
            stride = WAVE;
        } else {
            const int b0 = S.sliceOff[c / WAVE];
            width = (S.sliceOff[c / WAVE + 1] - b0) / WAVE;
            s0 = b0 + c % WAVE;
            stride = WAVE;
        }
    }
    for (int k = 0; k < width; ++k) {
        const int e = G.S.src[s0 + k * stride];
        if (e < 0) continue;
        const int f = e >> 1;
        rec_add(G.Sf + 3 * f, G.magSf[f], ssf[f], A, b);
    }
    for (int k = G.bOff[c]; k < G.bOff[c + 1]; ++k) {
        const int f = G.nf + G.bList[k];
        rec_add(G.Sf + 3 * f, G.magSf[f], ssf[f], A, b);
    }
    if (emptyDir >= 0) A[4 * emptyDir] += 1.0;
    const double xx = A[0], xy = A[1], xz = A[2], yx = A[3], yy = A[4], yz = A[5], zx = A[6], zy = A[7], zz = A[8];
    const double c00 = yy * zz - yz * zy, c01 = yz * zx - yx * zz, c02 = yx * zy - yy * zx;
    const double det = xx * c00 + xy * c01 + xz * c02;
    double v[3] = {0.0, 0.0, 0.0};
    if (!(fabs(det) < 1e-300)) {
        const double c10 = xz * zy - xy * zz, c11 = xx * zz - xz * zx, c12 = xy * zx - xx * zy;
        const double c20 = xy * yz - xz * yy, c21 = xz * yx - xx * yz, c22 = xx * yy - xy * yx;
        v[0] = (c00 * b[0] + c10 * b[1] + c20 * b[2]) / det;
        v[1] = (c01 * b[0] + c11 * b[1] + c21 * b[2]) / det;
        v[2] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) / det;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * c + k] = v[k];
}

__global__ void k_piso_correct_U(int n, const double* HbyA, const double* rAU, const double* v, int sign,
                                 int emptyDir, double* U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double t = rAU[c] * v[3 * c + k];
        const double u = sign < 0 ? HbyA[3 * c + k] - t : HbyA[3 * c + k] + t;
        U[3 * c + k] = k == emptyDir ? 0.0 : u;
    }
}

// fixedValue / zeroGradient boundary values (correctBoundaryConditions): per cell over its
// boundary faces, out = bValue on fixed faces, the cell's value elsewhere
template <int NC>
__global__ void k_boundary_values(FvcGeom G, const int* bFix, const double* bValue, const double* c, double* out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) out[NC * bf + q] = bFix[bf] ? bValue[NC * bf + q] : c[NC * cell + q];
    }
}

// dst[i·ds] = scale·src[i·ss] (scale ±1: a copy or an exact negation)
__global__ void k_copy_strided(long long n, const double* src, int ss, double* dst, int ds, double scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * ds] = scale * src[i * ss];
}

// ---------------------------------------------------------------- launchers
void launch_boundary_values(hipStream_t st, const FvcGeom& G, int nc, const int* bFix, const double* bValue,
                            const double* c, double* out) {
    if (G.n <= 0 || G.nb <= 0) return;
    if (nc == 1)
        hipLaunchKernelGGL(k_boundary_values<1>, piso_grid(G.n), dim3(PISO_THREADS), 0, st, G, bFix, bValue, c, out);
    else
        hipLaunchKernelGGL(k_boundary_values<3>, piso_grid(G.n), dim3(PISO_THREADS), 0, st, G, bFix, bValue, c, out);
}
void launch_copy_strided(hipStream_t st, long long n, const double* src, int ss, double* dst, int ds, double scale) {
    if (n > 0) hipLaunchKernelGGL(k_copy_strided, piso_grid(n), dim3(PISO_THREADS), 0, st, n, src, ss, dst, ds, scale);
}
void launch_piso_rau_hbya(hipStream_t st, int n, const double* A, const double* H, double* rAU, double* HbyA) {
    if (n > 0) hipLaunchKernelGGL(k_piso_rau_hbya, piso_grid(n), dim3(PISO_THREADS), 0, st, n, A, H, rAU, HbyA);
}
void launch_piso_hbya_boundary(hipStream_t st, const FvcGeom& G, const int* bFix, const double* Ub, const double* rAU,
                               const double* gp, double* bVal) {
    if (G.n > 0 && G.nb > 0)
        hipLaunchKernelGGL(k_piso_hbya_boundary, piso_grid(G.n), dim3(PISO_THREADS), 0, st, G, bFix, Ub, rAU, gp, bVal);
}
void launch_piso_face_op(hipStream_t st, int op, int nF, const double* a, const double* b, const double* c,
                         double* out) {
    if (nF <= 0) return;
    if (op == VOL_FACE_A_PLUS_B_TIMES_C)
        hipLaunchKernelGGL(k_piso_face_op<VOL_FACE_A_PLUS_B_TIMES_C>, piso_grid(nF), dim3(PISO_THREADS), 0, st, nF, a, b,
                           c, out);
    else if (op == VOL_FACE_A_MINUS_B)
        hipLaunchKernelGGL(k_piso_face_op<VOL_FACE_A_MINUS_B>, piso_grid(nF), dim3(PISO_THREADS), 0, st, nF, a, b, c,
                           out);
    else
        hipLaunchKernelGGL(k_piso_face_op<VOL_FACE_A_MINUS_B_OVER_C>, piso_grid(nF), dim3(PISO_THREADS), 0, st, nF, a,
                           b, c, out);
}
void launch_piso_ddt_phi_corr(hipStream_t st, const FvcGeom& G, double coeff, double c0, double c00, const double* U0,
                              const double* U00, const double* phiOld, const double* phiOldOld, const double* rAUf,
                              double* phiHbyA) {
    if (G.nf > 0)
        hipLaunchKernelGGL(k_piso_ddt_phi_corr, piso_grid(G.nf), dim3(PISO_THREADS), 0, st, G, coeff, c0, c00, U0, U00,
                           phiOld, phiOldOld, rAUf, phiHbyA);
}
void launch_piso_constrain_pressure(hipStream_t st, const FvcGeom& G, int b0, int cnt, const double* phiHbyA,
                                    const double* Ub, const double* rAUf, double* g) {
    if (cnt > 0)
        hipLaunchKernelGGL(k_piso_constrain_pressure, piso_grid(cnt), dim3(PISO_THREADS), 0, st, G, b0, cnt, phiHbyA,
                           Ub, rAUf, g);
}
void launch_fvc_reconstruct(hipStream_t st, const FvcGeom& G, const double* ssf, int emptyDir, double* out) {
    if (G.n > 0) hipLaunchKernelGGL(k_fvc_reconstruct, piso_grid(G.n), dim3(PISO_THREADS), 0, st, G, ssf, emptyDir, out);
}
void launch_piso_correct_U(hipStream_t st, int n, const double* HbyA, const double* rAU, const double* v, int sign,
                           int emptyDir, double* U) {
    if (n > 0)
        hipLaunchKernelGGL(k_piso_correct_U, piso_grid(n), dim3(PISO_THREADS), 0, st, n, HbyA, rAU, v, sign, emptyDir, U);
}

}  // namespace vol

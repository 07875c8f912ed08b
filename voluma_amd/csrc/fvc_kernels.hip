// Explicit operators of the pressure corrector (Fvc.h): face interpolation (plain and
// skew-corrected), the least-squares, Gauss-linear and skew-corrected Gauss gradients,
// the divergence of a face flux, the surface integral of face values, and the face flux
// of a cell vector field as phiHbyA is formed (IncompressibleFlow.cpp:276-291).
//
// The reference scatters face contributions into cells in face order. Here a cell
// gathers its faces through the level-0 slot list (internal faces, ascending) and then
// its boundary faces (ascending), which is the same per-cell order, so every sum is
// bit-identical; face-wise results are computed per face. Multi-component values are
// interleaved (a Vec3 per entry for vector fields and scalar gradients, a row-major
// Tensor for vector gradients), as std::vector<Vec3>/<Tensor> lay them out.
#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

namespace {
constexpr int FVC_THREADS = 256;
inline dim3 fvc_grid(int cnt) { return dim3((cnt + FVC_THREADS - 1) / FVC_THREADS); }

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Slots of row c of the level-0 layout: slot k at s0 + k·stride.
__device__ __forceinline__ void row_span(const SellView& S, int c, int& s0, int& width, int& stride) {
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

// gradDot(skew, w g_P + (1−w) g_N) (Types.h:156-157): sk·g for a Vec3 gradient, the
// Vec3 sk·T for a Tensor one.
template <int NC>
__device__ __forceinline__ void skew_term(const double* sk, double w, const double* gP, const double* gN, double* t) {
    double g[3 * NC];
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) g[k] = w * gP[k] + (1.0 - w) * gN[k];
    if constexpr (NC == 1) {
        t[0] = sk[0] * g[0] + sk[1] * g[1] + sk[2] * g[2];
    } else {
#pragma unroll
        for (int j = 0; j < 3; ++j) t[j] = sk[0] * g[j] + sk[1] * g[3 + j] + sk[2] * g[6 + j];
    }
}

// Inverse of a 3×3 tensor as Fvc.h:112-126 forms it.
__device__ __forceinline__ void inv3(const double* a, double* r) {
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    r[0] = (a[4] * a[8] - a[5] * a[7]) / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[3] = (a[5] * a[6] - a[3] * a[8]) / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[6] = (a[3] * a[7] - a[4] * a[6]) / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

// g += lsApply(G, v, dphi) (Fvc.h:129-130).
template <int NC>
__device__ __forceinline__ void ls_add(const double* G, const double* v, const double* dphi, double* g) {
    const double a[3] = {G[0] * v[0] + G[1] * v[1] + G[2] * v[2], G[3] * v[0] + G[4] * v[1] + G[5] * v[2],
                         G[6] * v[0] + G[7] * v[1] + G[8] * v[2]};
    if constexpr (NC == 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) g[k] += a[k] * dphi[0];
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) g[3 * i + j] += a[i] * dphi[j];
    }
}

// Least-squares constraint vector of boundary face i: d on value faces, else
// n·max(n·d, 0.05|d|) (Fvc.h:191-200).
__device__ __forceinline__ void ls_boundary_de(const FvcGeom& G, int i, bool fixes, double* de) {
    const double* d = G.d + 3 * i;
    if (fixes) {
        de[0] = d[0];
        de[1] = d[1];
        de[2] = d[2];
        return;
    }
    const double s = 1.0 / G.magSf[i];
    const double nv[3] = {G.Sf[3 * i] * s, G.Sf[3 * i + 1] * s, G.Sf[3 * i + 2] * s};
    const double nd0 = dot3(nv, d), lim = 0.05 * sqrt(dot3(d, d));
    const double nd = nd0 < lim ? lim : nd0;  // std::max(n·d, 0.05|d|)
#pragma unroll
    for (int k = 0; k < 3; ++k) de[k] = nv[k] * nd;
}
}  // namespace

// ---------------------------------------------------------------- interpolate (Fvc.h:13-31,72-81)
template <int NC>
__global__ void k_fvc_interp_faces(FvcGeom G, const double* c, const double* grad, double* out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= G.nf) return;
    const int o = G.own[f], nb = G.nei[f];
    const double w = G.w[f];
    double v[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] = w * c[NC * o + k] + (1.0 - w) * c[NC * nb + k];
    if (grad && G.skewed) {
        double t[NC];
        skew_term<NC>(G.skew + 3 * f, w, grad + 3 * NC * o, grad + 3 * NC * nb, t);
#pragma unroll
        for (int k = 0; k < NC; ++k) v[k] = v[k] + t[k];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) out[NC * f + k] = v[k];
}

// Boundary faces: the given face values, else the owner cell's value.
template <int NC>
__global__ void k_fvc_interp_boundary(FvcGeom G, const double* c, const double* bval, double* out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) out[NC * (G.nf + bf) + q] = bval ? bval[NC * bf + q] : c[NC * cell + q];
    }
}

// ---------------------------------------------------------------- gradients
// Weighted least squares (Fvc.h:134-206) per cell: Σ d⊗d/|d|² over its faces, the
// inverse, then Σ (G·d/|d|²)·Δφ with Δφ = φ_N − φ_P of each internal face and the
// boundary constraints (value faces: d, φ_b − φ_P; others: the zero normal gradient).
template <int NC>
__global__ void k_fvc_grad_ls(FvcGeom G, const double* c, const int* bFix, const double* bVal, double* out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    int s0, width, stride;
    row_span(G.S, cell, s0, width, stride);
    double dd[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) dd[k] = 0.0;
    for (int k = 0; k < width; ++k) {
        const int e = G.S.src[s0 + k * stride];
        if (e < 0) continue;
        const double* v = G.d + 3 * (e >> 1);
        const double s = 1.0 / dot3(v, v);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) dd[3 * a + b] += (v[a] * v[b]) * s;
    }
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k];
        double de[3];
        ls_boundary_de(G, G.nf + bf, bFix[bf] != 0, de);
        const double s = 1.0 / dot3(de, de);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) dd[3 * a + b] += (de[a] * de[b]) * s;
    }
    if (G.emptyDir >= 0) dd[4 * G.emptyDir] += 1.0;  // 2-D (Fvc.h:152)
    double Gi[9];
    inv3(dd, Gi);
    double g[3 * NC];
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) g[k] = 0.0;
    for (int k = 0; k < width; ++k) {
        const int slot = s0 + k * stride;
        const int e = G.S.src[slot];
        if (e < 0) continue;
        const int f = e >> 1, other = G.S.col[slot];
        const double* v = G.d + 3 * f;
        const double wgt = 1.0 / dot3(v, v);
        const double vw[3] = {v[0] * wgt, v[1] * wgt, v[2] * wgt};
        const int P = (e & 1) ? other : cell, N = (e & 1) ? cell : other;
        double dp[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) dp[q] = c[NC * N + q] - c[NC * P + q];
        ls_add<NC>(Gi, vw, dp, g);
    }
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k];
        const bool fixes = bFix[bf] != 0;
        double de[3];
        ls_boundary_de(G, G.nf + bf, fixes, de);
        const double s = 1.0 / dot3(de, de);
        const double vs[3] = {de[0] * s, de[1] * s, de[2] * s};
        double dphi[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) dphi[q] = fixes ? bVal[NC * bf + q] - c[NC * cell + q] : 0.0;
        ls_add<NC>(Gi, vs, dphi, g);
    }
    if (G.emptyDir >= 0)  // 2-D: derivatives along the empty direction zeroed (Fvc.h:169-181)
#pragma unroll
        for (int q = 0; q < NC; ++q) g[NC * G.emptyDir + q] = 0.0;
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) out[3 * NC * cell + k] = g[k];
}

// Gauss gradient of a scalar (Fvc.h:213-222, GaussLinear; 223-259, GaussSkew with the
// least-squares gradient gls): Σ S_f ψ_f over the cell's faces (owner +, neighbour −,
// boundary +) · (1/V).
template <bool SKEW>
__global__ void k_fvc_grad_gauss(FvcGeom G, const double* psi, const double* gls, const int* bFix,
                                 const double* bVal, double* out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    int s0, width, stride;
    row_span(G.S, cell, s0, width, stride);
    double g[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < width; ++k) {
        const int slot = s0 + k * stride;
        const int e = G.S.src[slot];
        if (e < 0) continue;
        const int f = e >> 1, other = G.S.col[slot];
        const int P = (e & 1) ? other : cell, N = (e & 1) ? cell : other;
        const double w = G.w[f];
        double phif = w * psi[P] + (1.0 - w) * psi[N];
        if constexpr (SKEW) {
            double t;
            skew_term<1>(G.skew + 3 * f, w, gls + 3 * P, gls + 3 * N, &t);
            phif = phif + t;
        }
        const double* S = G.Sf + 3 * f;
        if (e & 1) {
#pragma unroll
            for (int q = 0; q < 3; ++q) g[q] -= S[q] * phif;
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) g[q] += S[q] * phif;
        }
    }
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k], i = G.nf + bf;
        double phif;
        if (bFix[bf]) {
            phif = bVal[bf];
        } else if constexpr (SKEW) {
            const double* dv = G.d + 3 * i;
            const double s = 1.0 / G.magSf[i];
            const double nv[3] = {G.Sf[3 * i] * s, G.Sf[3 * i + 1] * s, G.Sf[3 * i + 2] * s};
            const double nd = dot3(nv, dv);
            const double dt[3] = {dv[0] - nv[0] * nd, dv[1] - nv[1] * nd, dv[2] - nv[2] * nd};
            const double gb = 0.0 * psi[cell] + 0.0;  // zeroGradient: gic·ψ_P + gbc
            phif = psi[cell] + dot3(dt, gls + 3 * cell) + gb * nd;
        } else {
            phif = psi[cell];
        }
        const double* S = G.Sf + 3 * i;
#pragma unroll
        for (int q = 0; q < 3; ++q) g[q] += S[q] * phif;
    }
    const double r = 1.0 / G.V[cell];
#pragma unroll
    for (int q = 0; q < 3; ++q) out[3 * cell + q] = q == G.emptyDir ? 0.0 : g[q] * r;
}

// ---------------------------------------------------------------- div, surfaceIntegrate
// Σ F over the cell's faces (owner +, neighbour −, boundary +); DIV: / V (Fvc.h:294-305),
// else · (1/V) (surfaceIntegrate, Fvc.h:307-318).
template <int NC, bool DIV>
__global__ void k_fvc_integrate(FvcGeom G, const double* F, double* out) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    int s0, width, stride;
    row_span(G.S, cell, s0, width, stride);
    double r[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) r[q] = 0.0;
    for (int k = 0; k < width; ++k) {
        const int e = G.S.src[s0 + k * stride];
        if (e < 0) continue;
        const int f = e >> 1;
        if (e & 1) {
#pragma unroll
            for (int q = 0; q < NC; ++q) r[q] -= F[NC * f + q];
        } else {
#pragma unroll
            for (int q = 0; q < NC; ++q) r[q] += F[NC * f + q];
        }
    }
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int i = G.nf + G.bList[k];
#pragma unroll
        for (int q = 0; q < NC; ++q) r[q] += F[NC * i + q];
    }
    if constexpr (DIV) {
#pragma unroll
        for (int q = 0; q < NC; ++q) out[NC * cell + q] = r[q] / G.V[cell];
    } else {
        const double s = 1.0 / G.V[cell];
#pragma unroll
        for (int q = 0; q < NC; ++q) out[NC * cell + q] = r[q] * s;
    }
}

// ---------------------------------------------------------------- phiHbyA-style face flux
__global__ void k_fvc_flux_faces(FvcGeom G, const double* U, const double* grad, double* phi) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= G.nf) return;
    const int o = G.own[f], nb = G.nei[f];
    const double w = G.w[f];
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = w * U[3 * o + k] + (1.0 - w) * U[3 * nb + k];
    if (grad && G.skewed) {
        double t[3];
        skew_term<3>(G.skew + 3 * f, w, grad + 9 * o, grad + 9 * nb, t);
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = v[k] + t[k];
    }
    phi[f] = dot3(v, G.Sf + 3 * f);
}

__global__ void k_fvc_flux_boundary(FvcGeom G, const double* U, const int* bFix, const double* Ub, double* phi) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= G.n) return;
    for (int k = G.bOff[cell]; k < G.bOff[cell + 1]; ++k) {
        const int bf = G.bList[k], i = G.nf + bf;
        phi[i] = dot3(bFix[bf] ? Ub + 3 * bf : U + 3 * cell, G.Sf + 3 * i);
    }
}

// ---------------------------------------------------------------- launchers
void launch_fvc_interpolate(hipStream_t st, const FvcGeom& G, int nc, const double* c, const double* grad,
                            const double* bval, double* out) {
    if (nc == 1) {
        if (G.nf > 0) hipLaunchKernelGGL(k_fvc_interp_faces<1>, fvc_grid(G.nf), dim3(FVC_THREADS), 0, st, G, c, grad, out);
        if (G.n > 0 && G.nb > 0)
            hipLaunchKernelGGL(k_fvc_interp_boundary<1>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, c, bval, out);
    } else {
        if (G.nf > 0) hipLaunchKernelGGL(k_fvc_interp_faces<3>, fvc_grid(G.nf), dim3(FVC_THREADS), 0, st, G, c, grad, out);
        if (G.n > 0 && G.nb > 0)
            hipLaunchKernelGGL(k_fvc_interp_boundary<3>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, c, bval, out);
    }
}

void launch_fvc_grad_ls(hipStream_t st, const FvcGeom& G, int nc, const double* c, const int* bFix,
                        const double* bVal, double* out) {
    if (G.n <= 0) return;
    if (nc == 1)
        hipLaunchKernelGGL(k_fvc_grad_ls<1>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, c, bFix, bVal, out);
    else
        hipLaunchKernelGGL(k_fvc_grad_ls<3>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, c, bFix, bVal, out);
}

void launch_fvc_grad_gauss(hipStream_t st, const FvcGeom& G, bool skew, const double* psi, const double* gls,
                           const int* bFix, const double* bVal, double* out) {
    if (G.n <= 0) return;
    if (skew)
        hipLaunchKernelGGL(k_fvc_grad_gauss<true>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, psi, gls, bFix, bVal,
                           out);
    else
        hipLaunchKernelGGL(k_fvc_grad_gauss<false>, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, psi, gls, bFix, bVal,
                           out);
}

void launch_fvc_integrate(hipStream_t st, const FvcGeom& G, int nc, bool div, const double* F, double* out) {
    if (G.n <= 0) return;
    if (div)
        hipLaunchKernelGGL((k_fvc_integrate<1, true>), fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, F, out);
    else if (nc == 1)
        hipLaunchKernelGGL((k_fvc_integrate<1, false>), fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, F, out);
    else
        hipLaunchKernelGGL((k_fvc_integrate<3, false>), fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, F, out);
}

void launch_fvc_flux(hipStream_t st, const FvcGeom& G, const double* U, const double* grad, const int* bFix,
                     const double* Ub, double* phi) {
    if (G.nf > 0) hipLaunchKernelGGL(k_fvc_flux_faces, fvc_grid(G.nf), dim3(FVC_THREADS), 0, st, G, U, grad, phi);
    if (G.n > 0 && G.nb > 0)
        hipLaunchKernelGGL(k_fvc_flux_boundary, fvc_grid(G.n), dim3(FVC_THREADS), 0, st, G, U, bFix, Ub, phi);
}

}  // namespace vol

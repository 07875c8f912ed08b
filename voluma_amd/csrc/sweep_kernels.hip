// Level-scheduled triangular sweeps: the DIC/DILU preconditioner (factor, forward and
// backward substitution, LinearSolver.cpp:56-105) and the Gauss–Seidel solver sweeps
// (LinearSolver.cpp:238-264).
//
// The reference runs these loops one cell (or one face) at a time. A cell's update reads
// only neighbours with a smaller index (forward) or a larger one (backward), so cells
// are grouped into levels: the forward level of a cell is one more than the largest
// forward level among its lower-numbered neighbours, and no two cells of one level are
// neighbours. One launch per level updates all its cells at once, each cell with the
// same operands, in the same order, as the sequential loop, so the results are
// bit-identical to it. Level lists and per-cell neighbour lists belong to the mesh;
// the factor and the coefficient products of the substitutions belong to the matrix.
#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

// DIC/DILU factor (LinearSolver.cpp:67-74) for the cells of one forward level:
// d_u = diag_u − Σ_{l<u, ascending (l, face)} (A_lu·A_ul) / d_l, rD_u = 1/d_u. The same
// thread then forms the coefficient products the substitutions use, rD_u·A_ul for its
// lower neighbours and rD_u·A_ul' for its upper ones (row u's own coefficients).
__global__ void k_dilu_factor(SweepView s, int lo, int hi, const double* diag, const double* upper,
                              const double* lower, double* dF, double* rD, double* pf, double* pb) {
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const int u = s.permF[i];
    const int k0 = s.lowOff[u], k1 = s.lowOff[u + 1];
    double d = diag[u];
    for (int k = k0; k < k1; ++k) {
        const int e = s.lowSrc[k], f = e >> 1;
        const double aul = (e & 1) ? lower[f] : upper[f];  // row u, column l
        const double alu = (e & 1) ? upper[f] : lower[f];  // row l, column u
        d = d - (alu * aul) / dF[s.lowCol[k]];
    }
    dF[u] = d;
    const double r = 1.0 / d;
    rD[u] = r;
    for (int k = k0; k < k1; ++k) {
        const int e = s.lowSrc[k], f = e >> 1;
        pf[k] = r * ((e & 1) ? lower[f] : upper[f]);
    }
    for (int k = s.upOff[u]; k < s.upOff[u + 1]; ++k) {
        const int e = s.upSrc[k], f = e >> 1;
        pb[k] = r * ((e & 1) ? lower[f] : upper[f]);
    }
}

// Forward substitution (LinearSolver.cpp:93-98) for one forward level:
// w_u = rD_u·r_u − Σ_{l<u, ascending (l, face)} (rD_u·A_ul)·w_l.
template <class T>
__global__ void k_dilu_forward(SweepView s, int lo, int hi, const T* rD, const T* pf, const T* r, T* w,
                               const int* done) {
    if (done && *done) return;
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const int u = s.permF[i];
    T acc = rD[u] * r[u];
    for (int k = s.lowOff[u]; k < s.lowOff[u + 1]; ++k) acc = acc - pf[k] * w[s.lowCol[k]];
    w[u] = acc;
}

// Backward substitution (LinearSolver.cpp:99-103) for one backward level:
// w_l −= (rD_l·A_lu)·w_u over the upper neighbours u > l in descending (u, face) order.
template <class T>
__global__ void k_dilu_backward(SweepView s, int lo, int hi, const T* pb, T* w, const int* done) {
    if (done && *done) return;
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const int l = s.permB[i];
    T acc = w[l];
    for (int k = s.upOff[l]; k < s.upOff[l + 1]; ++k) acc = acc - pb[k] * w[s.upCol[k]];
    w[l] = acc;
}

// Gauss–Seidel update of the cells of one level (LinearSolver.cpp:247-262; the GAMG
// smoother GAMG.cpp:363-384 has the same row), in place: x_c = (b_c − Σ coef·x_nb over
// its slots, which hold its faces in ascending order) / diag_c. Lower-numbered
// neighbours already hold this sweep's values (forward) or the previous pass's
// (backward), exactly as in the sequential loop. The slot walk covers the three
// layouts of SellView (CSR rows, ELL, SELL slices).
template <class T>
__global__ void k_gs_level(SellView S, const int* perm, int lo, int hi, const T* coef, const T* diag, const T* b,
                           T* x, const int* done) {
    if (done && *done) return;
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const int c = perm[i];
    int s0, width, stride;
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
    T r = b[c];
    for (int k = 0; k < width; ++k) {
        const int slot = s0 + k * stride;
        int j;
        if (S.col16) {
            const int d = int(S.col16[slot]);
            j = d == COL16_PAD ? -1 : c + d;
        } else {
            j = S.col[slot];
        }
        if (j >= 0) r = r - coef[slot] * x[j];
    }
    x[c] = r / diag[c];
}

// x = 0 and x[i] += xc[agg[i]] (GAMG.cpp:423,438) for the Gauss–Seidel V-cycle.
template <class T>
__global__ void k_fill_zero(int n, T* x, const int* done) {
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = T(0);
}
template <class T>
__global__ void k_prolong(int n, const int* agg, const T* xc, T* x, const int* done) {
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + xc[agg[i]];
}

// Coarsest-level solve in the reference's order (GAMG.cpp:406-421): forward substitution
// (rows ascending, each subtracting in ascending k: done column by column by all lanes,
// which keeps every row's order), then backward substitution row by row from the last,
// each row subtracting in ascending k before its division. A row's sum starts with the
// value the previous row just produced, so the backward pass is one dependent chain,
// run by lane 0 with the factors in LDS when they fit (n ≤ 64).
template <class T>
__global__ __launch_bounds__(64) void k_coarse_solve_rows(int n, const double* lu, const T* r, T* x,
                                                          const int* done) {
    if (done && *done) return;
    extern __shared__ double sh[];  // g (n), then the factors (n·n) when n ≤ 64
    const bool small = n <= 64;
    double* gv = sh;
    double* f = sh + n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) gv[i] = double(r[i]);
    if (small)
        for (int i = threadIdx.x; i < n * n; i += blockDim.x) f[i] = lu[i];
    __syncthreads();
    const double* a = small ? f : lu;
    for (int k = 0; k < n; ++k) {
        const double gk = gv[k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) gv[i] -= a[size_t(i) * n + k] * gk;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int i = n - 1; i >= 0; --i) {
            double acc = gv[i];
            const double* row = a + size_t(i) * n;
            for (int k = i + 1; k < n; ++k) acc -= row[k] * gv[k];
            gv[i] = acc / row[i];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = T(gv[i]);
}

namespace {
constexpr int SWEEP_THREADS = 256;
inline dim3 sweep_grid(int cnt) { return dim3((cnt + SWEEP_THREADS - 1) / SWEEP_THREADS); }
}  // namespace

void launch_dilu_factor(hipStream_t st, const SweepView& s, const int* levOffF, int nLevF, const double* diag,
                        const double* upper, const double* lower, double* dF, double* rD, double* pf, double* pb) {
    for (int L = 0; L < nLevF; ++L) {
        const int lo = levOffF[L], hi = levOffF[L + 1];
        if (hi > lo)
            hipLaunchKernelGGL(k_dilu_factor, sweep_grid(hi - lo), dim3(SWEEP_THREADS), 0, st, s, lo, hi, diag, upper,
                               lower, dF, rD, pf, pb);
    }
}

template <class T>
void launch_dilu_apply(hipStream_t st, const SweepView& s, const int* levOffF, int nLevF, const int* levOffB,
                       int nLevB, const T* rD, const T* pf, const T* pb, const T* r, T* w, const int* done) {
    for (int L = 0; L < nLevF; ++L) {
        const int lo = levOffF[L], hi = levOffF[L + 1];
        if (hi > lo)
            hipLaunchKernelGGL(k_dilu_forward<T>, sweep_grid(hi - lo), dim3(SWEEP_THREADS), 0, st, s, lo, hi, rD, pf, r,
                               w, done);
    }
    // backward level 0 holds the cells without upper neighbours: nothing to subtract
    for (int L = 1; L < nLevB; ++L) {
        const int lo = levOffB[L], hi = levOffB[L + 1];
        if (hi > lo)
            hipLaunchKernelGGL(k_dilu_backward<T>, sweep_grid(hi - lo), dim3(SWEEP_THREADS), 0, st, s, lo, hi, pb, w,
                               done);
    }
}
template void launch_dilu_apply<double>(hipStream_t, const SweepView&, const int*, int, const int*, int,
                                        const double*, const double*, const double*, const double*, double*,
                                        const int*);
template void launch_dilu_apply<float>(hipStream_t, const SweepView&, const int*, int, const int*, int, const float*,
                                       const float*, const float*, const float*, float*, const int*);

template <class T>
void launch_gs_sweep(hipStream_t st, const SellView& S, const int* perm, const int* levOff, int nLev, const T* coef,
                     const T* diag, const T* b, T* x, const int* done) {
    for (int L = 0; L < nLev; ++L) {
        const int lo = levOff[L], hi = levOff[L + 1];
        if (hi > lo)
            hipLaunchKernelGGL(k_gs_level<T>, sweep_grid(hi - lo), dim3(SWEEP_THREADS), 0, st, S, perm, lo, hi, coef,
                               diag, b, x, done);
    }
}
template void launch_gs_sweep<double>(hipStream_t, const SellView&, const int*, const int*, int, const double*,
                                      const double*, const double*, double*, const int*);
template void launch_gs_sweep<float>(hipStream_t, const SellView&, const int*, const int*, int, const float*,
                                     const float*, const float*, float*, const int*);

template <class T>
void launch_coarse_solve_rows(hipStream_t st, int n, const double* lu, const T* r, T* x, const int* done) {
    const size_t lds = sizeof(double) * (size_t(n) + (n <= 64 ? size_t(n) * n : 0));
    hipLaunchKernelGGL(k_coarse_solve_rows<T>, dim3(1), dim3(64), lds, st, n, lu, r, x, done);
}
template void launch_coarse_solve_rows<double>(hipStream_t, int, const double*, const double*, double*, const int*);
template void launch_coarse_solve_rows<float>(hipStream_t, int, const double*, const float*, float*, const int*);

template <class T>
void launch_fill_zero(hipStream_t st, int n, T* x, const int* done) {
    if (n > 0) hipLaunchKernelGGL(k_fill_zero<T>, sweep_grid(n), dim3(SWEEP_THREADS), 0, st, n, x, done);
}
template void launch_fill_zero<double>(hipStream_t, int, double*, const int*);
template void launch_fill_zero<float>(hipStream_t, int, float*, const int*);

template <class T>
void launch_prolong(hipStream_t st, int n, const int* agg, const T* xc, T* x, const int* done) {
    if (n > 0) hipLaunchKernelGGL(k_prolong<T>, sweep_grid(n), dim3(SWEEP_THREADS), 0, st, n, agg, xc, x, done);
}
template void launch_prolong<double>(hipStream_t, int, const int*, const double*, double*, const int*);
template void launch_prolong<float>(hipStream_t, int, const int*, const float*, float*, const int*);

}  // namespace vol

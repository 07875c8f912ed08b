// HIP kernels of the LDU linear-solve path for CDNA4 (gfx950, wave64).
//
// All kernels are memory-bound; none uses MFMA. Arithmetic is compiled with
// -ffp-contract=off so that every a·b + c is a rounded product followed by a
// rounded add, as in the reference (CMakeLists.txt:11).
#include <hip/hip_ext.h>

#include <type_traits>

#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

// Value of v in lane `lane` (the same for the whole wavefront) as a scalar read:
// two v_readlane, no LDS round trip as a general shuffle would take.
__device__ __forceinline__ double lane_bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b & 0xffffffffll), lane));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), lane));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// ============================================================================ SELL gather
// One wavefront per 64-row slice; lanes own rows. Block b is remapped so that each of
// the 8 XCDs (blocks are dispatched round-robin over XCDs) walks a contiguous range of
// slices: neighbouring rows share x gathers in that XCD's L2.
constexpr int SELL_WAVES = 4;  // 256 threads per block
constexpr int NXCD = 8;

// Grid rounded up to a multiple of the XCD count, and the logical block of this block
// under the XCD-contiguous remap: with blocks dispatched round-robin over the XCDs, XCD
// x gets logical blocks [x·nb/8, (x+1)·nb/8), so rows that share cache lines (of the
// vectors they write, read or gather) meet in one XCD's L2 instead of up to eight.
inline int xcd_grid(long blocks) { return int(((blocks + NXCD - 1) / NXCD) * NXCD); }
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x / NXCD;
    return (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
}

// ============================================================================ partial sums
constexpr int FIN_THREADS = 256;
constexpr int FIN_LOADS = 32;  // partials in flight per thread

// Sum of nparts block partials in a fixed order by a block of FIN_THREADS threads:
// FIN_LOADS partials in flight per thread, then the block sum (the value on every thread).
__device__ __forceinline__ double fin_sum(const double* pq, int nparts, double* sh) {
    double s = 0.0;
    for (int k0 = threadIdx.x; k0 < nparts; k0 += FIN_LOADS * blockDim.x) {
        double v[FIN_LOADS];
#pragma unroll
        for (int u = 0; u < FIN_LOADS; ++u) {
            const int k = k0 + u * blockDim.x;
            v[u] = k < nparts ? pq[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < FIN_LOADS; ++u) s += v[u];
    }
    return block_sum(s, sh);
}

__device__ void fin_step(int op, DevScalars& S, const DevScalars& I, double v0, double v1);

// Last-block finalize of a fused dot (SellArgsT::arrive): after its partial is stored,
// each block's thread 0 makes it visible device-wide and takes a ticket; the block that
// takes the last one reads every partial after an acquire, sums them as the finalize
// kernel does (the same block size and order, so the same bits), runs the scalar step
// and resets the counter for the next launch. A block of FIN_THREADS threads.
__device__ __forceinline__ void last_block_fin(const double* part, int nparts, int* arrive, DevScalars* S, int op) {
    __shared__ int isLast;
    __shared__ double sh[FIN_THREADS / WAVE];
    if (threadIdx.x == 0) {
        __threadfence();
        isLast = atomicAdd(arrive, 1) == nparts - 1;
    }
    __syncthreads();
    if (!isLast) return;
    __threadfence();
    const double t = fin_sum(part, nparts, sh);
    if (threadIdx.x == 0) {
        *arrive = 0;
        fin_step(op, *S, S[1], t, 0.0);
    }
}

int sell_grid_blocks(int nslices) { return xcd_grid((nslices + SELL_WAVES - 1) / SELL_WAVES); }

// Value of x at row j for the gather ops; T is the storage/arithmetic type of the level
// (double, or float in the mixed-precision inner solve).
template <int XL, class T>
__device__ __forceinline__ T xload(const SellArgsT<T>& a, int j, T xu) {
    const T om = T(a.omega);
    if constexpr (XL == XL_PLAIN) return a.x[j];
    if constexpr (XL == XL_UNIFORM) return xu;
    if constexpr (XL == XL_FROMZERO) return om * (a.r[j] * a.rD[j]);
    if constexpr (XL == XL_PROLONG) return a.x[j] + a.xc[a.agg[j]];
    if constexpr (XL == XL_ZEROPROLONG) return om * (a.r[j] * a.rD[j]) + a.xc[a.agg[j]];
    return T(0);
}


// Pushed prolongation: row I's new value v added to the finer cells of aggregate I,
// members g, g + G, … of its list (lane g of a group of G lanes that all hold v). Each
// finer cell belongs to one aggregate, so no two lanes write the same cell.
// Members in groups of 8: every index load of a group, then every z load, are in flight
// together.
template <class T>
__device__ __forceinline__ void push_prolongation(const SellArgsT<T>& a, int I, T v, int g, int G) {
    constexpr int U = 8;
    const int e = a.pushOff[I + 1];
    for (int k0 = a.pushOff[I] + g; k0 < e; k0 += U * G) {
        int m[U];
        T z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) m[u] = k0 + u * G < e ? a.pushList[k0 + u * G] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = m[u] >= 0 ? a.pushZ[m[u]] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (m[u] >= 0) a.pushZ[m[u]] = z[u] + v;
    }
}

// One row of a SELL gather op (the arithmetic shared by the grid kernel and the
// one-workgroup coarse-level kernel). Slots are taken in groups of U: all column and
// coefficient loads of a group, then all x gathers, are in flight together, and the
// accumulation keeps the slot order. pos: the row's position in the slices (the row
// itself unless the rows are sorted, SellView::perm).
template <int OP, int XL, int U, class T>
__device__ __forceinline__ T sell_row_at(const SellArgsT<T>& a, int pos, int row, T xu) {
    // SELL: slot k of the row at sliceOff[slice] + 64k + lane; CSR: at sliceOff[row] + k
    int s0, width, stride;
    if (a.S.csrG > 0) {
        s0 = a.S.sliceOff[row];
        width = a.S.sliceOff[row + 1] - s0;
        stride = 1;
    } else if (a.S.ellWidth > 0) {
        width = a.S.ellWidth;
        s0 = (pos / WAVE) * WAVE * width + pos % WAVE;
        stride = WAVE;
    } else {
        const int slice = pos / WAVE;
        const int b = a.S.sliceOff[slice];
        width = (a.S.sliceOff[slice + 1] - b) / WAVE;
        s0 = b + pos % WAVE;
        stride = WAVE;
    }
    T acc;
    if constexpr (OP == OP_AMUL) acc = a.diag[row] * xload<XL>(a, row, xu);
    else if constexpr (OP == OP_RESID) acc = a.r[row] - a.diag[row] * xload<XL>(a, row, xu);
    else acc = a.r[row];
    for (int k = 0; k < width; k += U) {
        int j[U];
        T av[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = s0 + (k + u) * stride;
            const bool in = k + u < width;
            if (a.S.col16) {
                const int d = in ? int(a.S.col16[idx]) : COL16_PAD;
                j[u] = d == COL16_PAD ? -1 : row + d;
            } else {
                j[u] = in ? a.S.col[idx] : -1;
            }
            av[u] = in ? a.coef[idx] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = j[u] >= 0 ? xload<XL>(a, j[u], xu) : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (j[u] >= 0) {
                const T t = av[u] * xv[u];
                if constexpr (OP == OP_AMUL) acc = acc + t;
                else acc = acc - t;
            }
        }
    }
    T outv;
    if constexpr (OP == OP_JACOBI) {
        const T J = acc / a.diag[row];
        const T xo = xload<XL>(a, row, xu);
        outv = xo + T(a.omega) * (J - xo);
    } else if constexpr (OP == OP_JACREF) {
        outv = acc / a.diag[row];
    } else {
        outv = acc;
    }
    if (a.out) a.out[row] = outv;  // (the fused restriction keeps the value only)
    return outv;
}

template <int OP, int XL, int U, class T>
__device__ __forceinline__ T sell_row(const SellArgsT<T>& a, int row, T xu) {
    return sell_row_at<OP, XL, U>(a, a.S.iperm ? a.S.iperm[row] : row, row, xu);
}

// One row of a gather op over the packed rows (SellView::colP): its column offsets in
// one vector load, its coefficients in ⌈W/2⌉ pair loads, then the x gathers; the
// accumulation keeps the slot order, so the result has the bits of sell_row. PW: offsets
// per row (4 or 8, ≥ ellWidth).
template <class T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

template <int OP, int XL, int PW, class T>
__device__ __forceinline__ T sell_row_packed(const SellArgsT<T>& a, int row, T xu) {
    using T2 = typename Pair<T>::type;
    const int W = a.S.ellWidth, P = (W + 1) / 2;
    int j[PW];
    {
        int w[PW / 2];
        if constexpr (PW == 4) {
            const int2 v = reinterpret_cast<const int2*>(a.S.colP)[row];
            w[0] = v.x;
            w[1] = v.y;
        } else {
            const int4 v = reinterpret_cast<const int4*>(a.S.colP)[row];
            w[0] = v.x;
            w[1] = v.y;
            w[2] = v.z;
            w[3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < PW; ++k) {
            const int d = (k & 1) ? (w[k / 2] >> 16) : int(short(w[k / 2] & 0xffff));
            j[k] = (k < W && d != COL16_PAD) ? row + d : -1;
        }
    }
    T av[PW];
    {
        const T2* cp = reinterpret_cast<const T2*>(a.coef + a.S.nslots) + size_t(row / WAVE) * WAVE * P + row % WAVE;
#pragma unroll
        for (int p = 0; p < PW / 2; ++p) {
            T2 v;
            v.x = T(0);
            v.y = T(0);
            if (p < P) v = cp[p * WAVE];
            av[2 * p] = v.x;
            av[2 * p + 1] = v.y;
        }
    }
    T acc;
    if constexpr (OP == OP_AMUL) acc = a.diag[row] * xload<XL>(a, row, xu);
    else if constexpr (OP == OP_RESID) acc = a.r[row] - a.diag[row] * xload<XL>(a, row, xu);
    else acc = a.r[row];
    T xv[PW];
#pragma unroll
    for (int k = 0; k < PW; ++k) xv[k] = j[k] >= 0 ? xload<XL>(a, j[k], xu) : T(0);
#pragma unroll
    for (int k = 0; k < PW; ++k) {
        if (j[k] >= 0) {
            const T t = av[k] * xv[k];
            if constexpr (OP == OP_AMUL) acc = acc + t;
            else acc = acc - t;
        }
    }
    T outv;
    if constexpr (OP == OP_JACOBI) {
        const T J = acc / a.diag[row];
        const T xo = xload<XL>(a, row, xu);
        outv = xo + T(a.omega) * (J - xo);
    } else if constexpr (OP == OP_JACREF) {
        outv = acc / a.diag[row];
    } else {
        outv = acc;
    }
    if (a.out) a.out[row] = outv;  // (the fused restriction keeps the value only)
    return outv;
}

// One row of a gather op over the packed SELL rows of a coarse level (SellView::colS):
// slots in groups of 16, each group's offsets in four 8-byte loads and its coefficients
// in eight pair loads, then its x gathers; the accumulation keeps the slot order, so the
// result has the bits of sell_row.
template <int OP, int XL, class T>
__device__ __forceinline__ T sell_row_pk(const SellArgsT<T>& a, int slice, int lane, int row, T xu) {
    using T2 = typename Pair<T>::type;
    const int4 d = a.S.sliceP[slice];
    const int W = d.z;
    const int2* cw = a.S.colS + d.x + lane;
    const T2* cp = reinterpret_cast<const T2*>(a.coef + a.S.nslots) + d.y + lane;
    T acc;
    if constexpr (OP == OP_AMUL) acc = a.diag[row] * xload<XL>(a, row, xu);
    else if constexpr (OP == OP_RESID) acc = a.r[row] - a.diag[row] * xload<XL>(a, row, xu);
    else acc = a.r[row];
    for (int k0 = 0; k0 < W; k0 += 16) {
        int j[16];
        T av[16], xv[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + 4 * q;
            int2 v = make_int2(0, 0);
            if (k < W) v = cw[WAVE * (k / 4)];
            const int o[4] = {int(short(v.x & 0xffff)), v.x >> 16, int(short(v.y & 0xffff)), v.y >> 16};
#pragma unroll
            for (int e = 0; e < 4; ++e) j[4 * q + e] = (k + e < W && o[e] != COL16_PAD) ? row + o[e] : -1;
        }
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int k = k0 + 2 * p;
            T2 v;
            v.x = T(0);
            v.y = T(0);
            if (k < W) v = cp[WAVE * (k / 2)];
            av[2 * p] = v.x;
            av[2 * p + 1] = v.y;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) xv[u] = j[u] >= 0 ? xload<XL>(a, j[u], xu) : T(0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (j[u] >= 0) {
                const T t = av[u] * xv[u];
                if constexpr (OP == OP_AMUL) acc = acc + t;
                else acc = acc - t;
            }
        }
    }
    T outv;
    if constexpr (OP == OP_JACOBI) {
        const T J = acc / a.diag[row];
        const T xo = xload<XL>(a, row, xu);
        outv = xo + T(a.omega) * (J - xo);
    } else if constexpr (OP == OP_JACREF) {
        outv = acc / a.diag[row];
    } else {
        outv = acc;
    }
    if (a.out) a.out[row] = outv;
    return outv;
}

// A gather op over packed SELL rows (no reduction), with the pushed prolongation after
// a post-sweep and the skip of finished batch systems like k_sell.
template <int OP, int XL, class T>
__global__ __launch_bounds__(256) void k_sell_pk(SellArgsT<T> a) {
    if (a.done && *a.done) return;
    const int slice = xcd_block() * SELL_WAVES + threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int row = slice * WAVE + lane;
    if (slice >= a.S.nslices) return;
    if (a.sliceSys) {  // batches: a slice whose systems are all done is skipped (wave-uniform)
        const int2 r = a.sliceSys[slice];
        bool all = true;
        for (int q = r.x; q <= r.y; ++q) all = all && a.sysS[q].done;
        if (all) return;
    }
    if (row >= a.S.n) return;
    T xu = T(0);
    if constexpr (XL == XL_UNIFORM) xu = *a.xuPtr;
    const T outv = sell_row_pk<OP, XL>(a, slice, lane, row, xu);
    if constexpr (OP == OP_JACOBI) {
        if (a.pushZ) push_prolongation(a, row, outv, 0, 1);
    }
}

// Body of k_sell_packed and k_sell_packed_fin. LASTFIN: the last-block finalize of the
// fused dot (SellArgsT::arrive), in a kernel of its own so that its registers do not
// lower the occupancy of the others.
template <int OP, int XL, bool DOT, int PW, class T, bool LASTFIN>
__device__ __forceinline__ void sell_packed_body(const SellArgsT<T>& a) {
    if (a.done && *a.done) return;
    const int lb = xcd_block();
    const int slice = lb * SELL_WAVES + threadIdx.x / WAVE;
    const int row = slice * WAVE + (threadIdx.x & (WAVE - 1));
    const bool valid = slice < a.S.nslices && row < a.S.n;
    T xu = T(0);
    if constexpr (XL == XL_UNIFORM) xu = *a.xuPtr;
    T outv = T(0);
    if (valid) outv = sell_row_packed<OP, XL, PW>(a, row, xu);
    if constexpr (DOT) {
        __shared__ double sh[256 / WAVE];
        const double v = block_sum(valid ? double(outv) * double(a.dotv[row]) : 0.0, sh);
        if (threadIdx.x == 0) a.part[lb] = v;
        if constexpr (LASTFIN) last_block_fin(a.part, int(gridDim.x), a.arrive, a.finS, a.finOp);
    }
}

template <int OP, int XL, bool DOT, int PW, class T>
__global__ __launch_bounds__(256) void k_sell_packed(SellArgsT<T> a) {
    sell_packed_body<OP, XL, DOT, PW, T, false>(a);
}
template <int OP, int XL, int PW, class T>
__global__ __launch_bounds__(256) void k_sell_packed_fin(SellArgsT<T> a) {
    sell_packed_body<OP, XL, true, PW, T, true>(a);
}

// Body of k_sell and k_sell_fin (LASTFIN as for the packed rows).
template <int OP, int XL, bool DOT, int U, class T, bool LASTFIN>
__device__ __forceinline__ void sell_body(const SellArgsT<T>& a) {
    if (a.done && *a.done) return;
    const int lb = xcd_block();
    const int slice = lb * SELL_WAVES + threadIdx.x / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    const int pos = slice * WAVE + lane;
    const bool valid = slice < a.S.nslices && pos < a.S.n;
    const int row = valid && a.S.perm ? a.S.perm[pos] : pos;
    if constexpr (!DOT) {  // batches: a slice whose systems are all done is skipped (wave-uniform)
        if (a.sliceSys && slice < a.S.nslices) {
            const int2 r = a.sliceSys[slice];
            bool all = true;
            for (int q = r.x; q <= r.y; ++q) all = all && a.sysS[q].done;
            if (all) return;
        }
    }
    T xu = T(0);
    if constexpr (XL == XL_UNIFORM) xu = *a.xuPtr;
    T outv = T(0);
    if (valid) outv = sell_row_at<OP, XL, U>(a, pos, row, xu);
    if constexpr (OP == OP_JACOBI) {
        if (a.pushZ && valid) push_prolongation(a, row, outv, 0, 1);
    }
    if constexpr (DOT) {  // Σ out·dotv, accumulated in double
        __shared__ double sh[256 / WAVE];
        const double v = block_sum(valid ? double(outv) * double(a.dotv[row]) : 0.0, sh);
        if (threadIdx.x == 0) a.part[lb] = v;
        if constexpr (LASTFIN) last_block_fin(a.part, int(gridDim.x), a.arrive, a.finS, a.finOp);
    }
}

template <int OP, int XL, bool DOT, int U, class T>
__global__ __launch_bounds__(256) void k_sell(SellArgsT<T> a) {
    sell_body<OP, XL, DOT, U, T, false>(a);
}
template <int OP, int XL, int U, class T>
__global__ __launch_bounds__(256) void k_sell_fin(SellArgsT<T> a) {
    sell_body<OP, XL, true, U, T, true>(a);
}

// One CSR row processed by G lanes (lane g takes slots g, g+G, …): all loads of the row
// are in flight at once, then every lane of the group accumulates the products in slot
// order by reading them from their lanes with cross-lane shuffles, so the arithmetic is
// identical to sell_row. Lane g == 0 writes the result. All lanes of the wave must call.
template <int OP, int XL, int G, int IT, class T>
__device__ __forceinline__ T csr_row(const SellArgsT<T>& a, int row, int g, bool valid, T xu) {
    int s0 = 0, len = 0;
    if (valid) {
        s0 = a.S.sliceOff[row];
        len = a.S.sliceOff[row + 1] - s0;
    }
    T p[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int k = g + G * it;
        p[it] = T(0);
        if (k < len) {
            const int j = a.S.col[s0 + k];
            p[it] = a.coef[s0 + k] * xload<XL>(a, j, xu);
        }
    }
    T acc = T(0);
    if (valid) {
        if constexpr (OP == OP_AMUL) acc = a.diag[row] * xload<XL>(a, row, xu);
        else if constexpr (OP == OP_RESID) acc = a.r[row] - a.diag[row] * xload<XL>(a, row, xu);
        else acc = a.r[row];
    }
    const int base = (threadIdx.x & (WAVE - 1)) - g;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T t = __shfl(p[it], base + q, WAVE);
            if (q + G * it < len) {
                if constexpr (OP == OP_AMUL) acc = acc + t;
                else acc = acc - t;
            }
        }
    }
    T outv = T(0);
    if (valid) {
        if constexpr (OP == OP_JACOBI) {
            const T J = acc / a.diag[row];
            const T xo = xload<XL>(a, row, xu);
            outv = xo + T(a.omega) * (J - xo);
        } else if constexpr (OP == OP_JACREF) {
            outv = acc / a.diag[row];
        } else {
            outv = acc;
        }
        if (g == 0 && a.out) a.out[row] = outv;
    }
    return outv;
}

template <int OP, int XL, int G, int IT, class T>
__global__ __launch_bounds__(256) void k_csrv(SellArgsT<T> a) {
    if (a.done && *a.done) return;
    const int gid = xcd_block() * blockDim.x + threadIdx.x;
    const int row = gid / G, g = gid % G;
    T xu = T(0);
    if constexpr (XL == XL_UNIFORM) xu = *a.xuPtr;
    bool valid = row < a.S.n;
    if (valid && a.rowSys) valid = !a.sysS[a.rowSys[row]].done;  // batches: skip finished systems
    const T v = csr_row<OP, XL, G, IT>(a, row, g, valid, xu);
    if constexpr (OP == OP_JACOBI) {
        if (a.pushZ && valid) push_prolongation(a, row, v, g, G);
    }
}

template <class T, class K> static void sell_launch(K kernel, int nb, hipStream_t st, const SellArgsT<T>& a) {
    if (a.evStart)
        hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, st, a.evStart, a.evStop, 0, a);
    else
        hipLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, st, a);
}

template <int OP, int XL, class T>
static void csr_dispatch(hipStream_t st, const SellArgsT<T>& a) {
    const long threads = long(a.S.n) * a.S.csrG;
    const int nb = xcd_grid((threads + 255) / 256);
    if (nb == 0) return;
    const int G = a.S.csrG, IT = a.S.csrIt;
    if (G == 4 && IT <= 4) sell_launch(k_csrv<OP, XL, 4, 4, T>, nb, st, a);
    else if (G == 8 && IT <= 4) sell_launch(k_csrv<OP, XL, 8, 4, T>, nb, st, a);
    else if (G == 16 && IT <= 4) sell_launch(k_csrv<OP, XL, 16, 4, T>, nb, st, a);
    else sell_launch(k_csrv<OP, XL, 16, 8, T>, nb, st, a);
}


template <int OP, int XL, class T>
static void sell_dispatch2(hipStream_t st, const SellArgsT<T>& a) {
    if (a.S.csrG > 0) {
        csr_dispatch<OP, XL>(st, a);
        return;
    }
    const int nb = sell_grid_blocks(a.S.nslices);
    if (nb == 0) return;
    if (a.S.colP && !a.sliceSys) {  // packed rows (level 0 of a single system)
        const bool fused = a.part != nullptr;
        if constexpr (OP == OP_AMUL && XL == XL_PLAIN) {
            if (fused && a.arrive) {
                if (a.S.colPW == 2) sell_launch(k_sell_packed_fin<OP, XL, 4, T>, nb, st, a);
                else sell_launch(k_sell_packed_fin<OP, XL, 8, T>, nb, st, a);
                return;
            }
        }
        if constexpr (OP == OP_AMUL || (OP == OP_JACOBI && XL != XL_UNIFORM && XL != XL_FROMZERO)) {
            if (fused) {
                if (a.S.colPW == 2) sell_launch(k_sell_packed<OP, XL, true, 4, T>, nb, st, a);
                else sell_launch(k_sell_packed<OP, XL, true, 8, T>, nb, st, a);
                return;
            }
        }
        if (!fused) {
            if (a.S.colPW == 2) sell_launch(k_sell_packed<OP, XL, false, 4, T>, nb, st, a);
            else sell_launch(k_sell_packed<OP, XL, false, 8, T>, nb, st, a);
        }
        return;
    }
    if (a.S.colS && !a.part) {  // packed SELL rows (coarse levels, batches)
        sell_launch(k_sell_pk<OP, XL, T>, nb, st, a);
        return;
    }
    const bool wide = a.S.maxWidth > 8;
    if (a.part) {  // fused Σ out·dotv: A·p, and the last sweep of the V-cycle (Σ w·r)
        if constexpr (OP == OP_AMUL && XL == XL_PLAIN) {
            if (a.arrive) {
                if (wide)
                    sell_launch(k_sell_fin<OP, XL, 16, T>, nb, st, a);
                else
                    sell_launch(k_sell_fin<OP, XL, 4, T>, nb, st, a);
                return;
            }
        }
        if constexpr (OP == OP_AMUL || (OP == OP_JACOBI && XL != XL_UNIFORM && XL != XL_FROMZERO)) {
            if (wide)
                sell_launch(k_sell<OP, XL, true, 16, T>, nb, st, a);
            else
                sell_launch(k_sell<OP, XL, true, 4, T>, nb, st, a);
        }
    } else {
        if (wide)
            sell_launch(k_sell<OP, XL, false, 16, T>, nb, st, a);
        else
            sell_launch(k_sell<OP, XL, false, 4, T>, nb, st, a);
    }
}

template <int OP, class T>
static void sell_dispatch1(hipStream_t st, int xl, const SellArgsT<T>& a) {
    switch (xl) {
    case XL_PLAIN: sell_dispatch2<OP, XL_PLAIN>(st, a); break;
    case XL_UNIFORM:
        if constexpr (OP == OP_AMUL) sell_dispatch2<OP, XL_UNIFORM>(st, a);
        break;
    case XL_FROMZERO: sell_dispatch2<OP, XL_FROMZERO>(st, a); break;
    case XL_PROLONG: sell_dispatch2<OP, XL_PROLONG>(st, a); break;
    case XL_ZEROPROLONG: sell_dispatch2<OP, XL_ZEROPROLONG>(st, a); break;
    }
}

template <class T>
static void launch_sell_t(hipStream_t st, int op, int xl, const SellArgsT<T>& a) {
    switch (op) {
    case OP_AMUL: sell_dispatch1<OP_AMUL>(st, xl, a); break;
    case OP_RESID: sell_dispatch1<OP_RESID>(st, xl, a); break;
    case OP_JACOBI: sell_dispatch1<OP_JACOBI>(st, xl, a); break;
    case OP_JACREF:
        if constexpr (sizeof(T) == 8) sell_dispatch1<OP_JACREF>(st, xl, a);
        break;
    }
}

void launch_sell(hipStream_t st, int op, int xl, const SellArgs& a) { launch_sell_t(st, op, xl, a); }
void launch_sell32(hipStream_t st, int op, int xl, const SellArgs32& a) { launch_sell_t(st, op, xl, a); }

// Slot coefficients from face arrays: coef[slot] = upper or lower of the slot's face.
__global__ void k_sell_coef(SellView S, const double* upper, const double* lower, double* coef, float* coef32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.nslots) return;
    const int s = S.src[i];
    const double v = s < 0 ? 0.0 : ((s & 1) ? lower[s >> 1] : upper[s >> 1]);
    if (coef) coef[i] = v;
    if (coef32) coef32[i] = float(v);
    if (S.colP) {  // the pair copy: ELL slot k of row 64·slice + lane → pair k/2, element k%2
        const int W = S.ellWidth, P = (W + 1) / 2;
        const int slice = i / (WAVE * W), rem = i - slice * WAVE * W, k = rem / WAVE, lane = rem % WAVE;
        const size_t j = size_t(S.nslots) + (size_t(slice) * WAVE * P + size_t(k / 2) * WAVE + lane) * 2 + (k & 1);
        if (coef) coef[j] = v;
        if (coef32) coef32[j] = float(v);
    }
    if (S.colS) {  // the packed SELL pair copy: the slice by bisection of the slice offsets
        int lo = 0, hi = S.nslices;  // sliceOff[lo] ≤ i < sliceOff[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (S.sliceOff[mid] <= i) lo = mid;
            else hi = mid;
        }
        const int4 d = S.sliceP[lo];
        const int rem = i - S.sliceOff[lo], k = rem / WAVE, lane = rem % WAVE;
        const size_t j = size_t(S.nslots) + (size_t(d.y) + size_t(k / 2) * WAVE + lane) * 2 + (k & 1);
        if (coef) coef[j] = v;
        if (coef32) coef32[j] = float(v);
        if ((d.z & 1) && k == d.z - 1) {  // the unused half of an odd-width slice's last pair
            if (coef) coef[j + 1] = 0.0;
            if (coef32) coef32[j + 1] = 0.0f;
        }
    }
}

void launch_sell_coef(hipStream_t st, const SellView& S, const double* upper, const double* lower, double* coef,
                      float* coef32) {
    if (S.nslots == 0) return;
    hipLaunchKernelGGL(k_sell_coef, dim3((S.nslots + 255) / 256), dim3(256), 0, st, S, upper, lower, coef, coef32);
}

// ============================================================================ vector kernels
// Grid of a fast vector op, which is also the partial count of its reduction. An element
// step stores before the next element's loads may issue (the vectors may alias), so a
// thread's elements are one memory round trip each: PCG's x/r update, the widest of
// them, runs up to 4·RED_BLOCKS blocks (within the partial scratch, maxBlk) like the ops
// without a reduction; other reductions use RED_BLOCKS blocks.
int vec_fast_blocks(int op, int n, int maxBlk) {
    const int rows = std::max((n + RED_THREADS - 1) / RED_THREADS, 1);
    if (op == V_PCG_XR) return std::max(std::min({rows, 4 * RED_BLOCKS, maxBlk}), 1);
    if (vec_has_reduction(op) > 0) return RED_BLOCKS;
    return std::min(rows, 4 * RED_BLOCKS);
}

int vec_has_reduction(int op) {
    switch (op) {
    case V_SUM_X: case V_PRECOND_DOT: case V_DOT: case V_PCG_XR: case V_BI_S: case V_BMINUSAX_NORM: return 1;
    case V_INIT: case V_BI_DOT2: return 2;
    case V_BI_XR: return 1;
    default: return 0;
    }
}

template <class T> __device__ __forceinline__ double ld(const T* p, int i) { return double(p[i]); }

// Element step of each vector op; returns the values to reduce in r0/r1. Values are
// widened to double, combined in double and rounded once when stored; for T = double
// this is exactly the reference's expression order.
template <int OP, class T>
__device__ __forceinline__ void vec_elem(const VecArgsT<T>& A, int i, double& r0, double& r1) {
    const DevScalars* S = A.S;
    if constexpr (OP == V_SUM_X) {
        r0 = ld(A.a, i);
    } else if constexpr (OP == V_INIT) {  // a = wA, b = b, c = pA; w = rA (y: ω·(rA·rD), nullable)
        const double rA = ld(A.b, i) - ld(A.a, i);
        A.w[i] = T(rA);
        if (A.y) A.y[i] = T(A.omega) * (T(rA) * A.rD[i]);
        r0 = fabs(ld(A.a, i) - ld(A.c, i)) + fabs(ld(A.b, i) - ld(A.c, i));
        r1 = fabs(rA);
    } else if constexpr (OP == V_PRECOND_DOT) {  // a = r → w
        const T w = T(A.rD ? ld(A.rD, i) * ld(A.a, i) : ld(A.a, i));
        A.w[i] = w;
        r0 = double(w) * ld(A.a, i);
    } else if constexpr (OP == V_DOT) {
        r0 = ld(A.a, i) * ld(A.b, i);
    } else if constexpr (OP == V_PCG_P) {  // a = w; x = p
        A.x[i] = T(S->iterations == 0 ? ld(A.a, i) : ld(A.a, i) + S->beta * ld(A.x, i));
    } else if constexpr (OP == V_PCG_XR) {  // a = p, b = q; x = x, y = r (w: ω·(r·rD), nullable)
        A.x[i] = T(ld(A.x, i) + S->alpha * ld(A.a, i));
        const T r = T(ld(A.y, i) - S->alpha * ld(A.b, i));
        A.y[i] = r;
        if (A.w) A.w[i] = T(A.omega) * (r * A.rD[i]);
        r0 = fabs(double(r));
    } else if constexpr (OP == V_BI_P) {  // a = r, b = Ay; x = p
        A.x[i] = T(S->iterations == 0 ? ld(A.a, i) : ld(A.a, i) + S->beta * (ld(A.x, i) - S->omegaB * ld(A.b, i)));
    } else if constexpr (OP == V_SCALE) {
        A.w[i] = T(A.rD ? ld(A.rD, i) * ld(A.a, i) : ld(A.a, i));
    } else if constexpr (OP == V_BI_S) {  // a = r, b = Ay; w = s
        const T s = T(ld(A.a, i) - S->alphaB * ld(A.b, i));
        A.w[i] = s;
        r0 = fabs(double(s));
    } else if constexpr (OP == V_BI_XEXIT) {  // a = y; x = x
        A.x[i] = T(ld(A.x, i) + S->alphaB * ld(A.a, i));
    } else if constexpr (OP == V_BI_DOT2) {  // a = t, b = s
        r0 = ld(A.a, i) * ld(A.a, i);
        r1 = ld(A.a, i) * ld(A.b, i);
    } else if constexpr (OP == V_BI_XR) {  // a = y, b = z, c = s, d = t; x = x, y = r
        A.x[i] = T(ld(A.x, i) + (S->alphaB * ld(A.a, i) + S->omegaB * ld(A.b, i)));
        const T r = T(ld(A.c, i) - S->omegaB * ld(A.d, i));
        A.y[i] = r;
        r0 = fabs(double(r));
    } else if constexpr (OP == V_BMINUSAX_NORM) {  // a = wA, b = b; w = r
        const double r = ld(A.b, i) - ld(A.a, i);
        A.w[i] = T(r);
        r0 = fabs(r);
    } else if constexpr (OP == V_COPY) {
        A.w[i] = A.a[i];
    } else if constexpr (OP == V_RECIP) {
        A.w[i] = T(1.0 / ld(A.a, i));
    } else if constexpr (OP == V_FROMZERO) {  // a = r, b = d
        A.w[i] = T(A.omega * (ld(A.a, i) / ld(A.b, i)));
    } else if constexpr (OP == V_FILL_XREF) {
        A.w[i] = T(S->xRef);
    } else if constexpr (OP == V_ZSCALE) {  // as xload's first sweep from zero computes it
        const T r = A.a[i];
        A.w[i] = T(A.omega) * (r * A.rD[i]);
    }
}

template <int OP, class T>
__device__ __forceinline__ bool vec_should_run(const VecArgsT<T>& A) {
    if constexpr (OP == V_BI_XEXIT) return A.S->exitPending != 0;
    return !(A.done && *A.done);
}

template <int OP, int NRED, class T>
__global__ __launch_bounds__(RED_THREADS) void k_vec_fast(VecArgsT<T> A, double* part, int maxBlk) {
    if (!vec_should_run<OP>(A)) return;
    double s0 = 0.0, s1 = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {
        double r0 = 0.0, r1 = 0.0;
        vec_elem<OP>(A, i, r0, r1);
        if constexpr (NRED > 0) s0 += r0;
        if constexpr (NRED > 1) s1 += r1;
    }
    if constexpr (NRED > 0) {
        __shared__ double sh[RED_THREADS / WAVE];
        const double t0 = block_sum(s0, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t0;
        if constexpr (NRED > 1) {
            const double t1 = block_sum(s1, sh);
            if (threadIdx.x == 0) part[maxBlk + blockIdx.x] = t1;
        }
    }
}

template <int OP, int NRED>
__global__ __launch_bounds__(RED_THREADS) void k_vec_exact(VecArgs A, long long* xpart, int* xnf, int maxBlk) {
    if (!vec_should_run<OP>(A)) return;
    __shared__ long long bins[NRED][NBINS];
    __shared__ int nf[NRED];
    for (int k = threadIdx.x; k < NRED * NBINS; k += blockDim.x) bins[k / NBINS][k % NBINS] = 0;
    if (threadIdx.x < NRED) nf[threadIdx.x] = 0;
    __syncthreads();
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < A.n; base += stride) {
        const int i = base + threadIdx.x;
        double r0 = 0.0, r1 = 0.0;
        const bool act = i < A.n;
        if (act) vec_elem<OP>(A, i, r0, r1);
        exact_add_wave(bins[0], &nf[0], r0, act);
        if constexpr (NRED > 1) exact_add_wave(bins[1], &nf[1], r1, act);
    }
    __syncthreads();
    if (threadIdx.x < NRED) exact_normalize(bins[threadIdx.x]);
    __syncthreads();
    for (int k = threadIdx.x; k < NRED * NBINS; k += blockDim.x) {
        const int q = k / NBINS, bn = k % NBINS;
        xpart[(size_t(q) * maxBlk + blockIdx.x) * NBINS + bn] = bins[q][bn];
    }
    if (threadIdx.x < NRED) xnf[threadIdx.x * maxBlk + blockIdx.x] = nf[threadIdx.x];
}

template <int OP, class T>
static void vec_dispatch(hipStream_t st, const VecArgsT<T>& a, bool exact, const RedScratch& R) {
    constexpr int NRED = (OP == V_INIT || OP == V_BI_DOT2) ? 2
                         : (OP == V_SUM_X || OP == V_PRECOND_DOT || OP == V_DOT || OP == V_PCG_XR || OP == V_BI_S ||
                            OP == V_BMINUSAX_NORM || OP == V_BI_XR)
                             ? 1
                             : 0;
    if constexpr (NRED > 0 && std::is_same_v<T, double>) {
        if (exact) {
            hipLaunchKernelGGL((k_vec_exact<OP, NRED>), dim3(RED_BLOCKS_EXACT), dim3(RED_THREADS), 0, st, a, R.xpart,
                               R.xnf, R.maxBlkExact);
            return;
        }
    }
    hipLaunchKernelGGL((k_vec_fast<OP, NRED, T>), dim3(vec_fast_blocks(OP, a.n, R.maxBlk)), dim3(RED_THREADS), 0, st,
                       a, R.part, R.maxBlk);
}

template <class T>
static void launch_vec_t(hipStream_t st, int op, const VecArgsT<T>& a, bool exact, const RedScratch& R) {
    switch (op) {
    case V_SUM_X: vec_dispatch<V_SUM_X>(st, a, exact, R); break;
    case V_INIT: vec_dispatch<V_INIT>(st, a, exact, R); break;
    case V_PRECOND_DOT: vec_dispatch<V_PRECOND_DOT>(st, a, exact, R); break;
    case V_DOT: vec_dispatch<V_DOT>(st, a, exact, R); break;
    case V_PCG_P: vec_dispatch<V_PCG_P>(st, a, exact, R); break;
    case V_PCG_XR: vec_dispatch<V_PCG_XR>(st, a, exact, R); break;
    case V_BI_P: vec_dispatch<V_BI_P>(st, a, exact, R); break;
    case V_SCALE: vec_dispatch<V_SCALE>(st, a, exact, R); break;
    case V_BI_S: vec_dispatch<V_BI_S>(st, a, exact, R); break;
    case V_BI_XEXIT: vec_dispatch<V_BI_XEXIT>(st, a, exact, R); break;
    case V_BI_DOT2: vec_dispatch<V_BI_DOT2>(st, a, exact, R); break;
    case V_BI_XR: vec_dispatch<V_BI_XR>(st, a, exact, R); break;
    case V_BMINUSAX_NORM: vec_dispatch<V_BMINUSAX_NORM>(st, a, exact, R); break;
    case V_COPY: vec_dispatch<V_COPY>(st, a, exact, R); break;
    case V_RECIP: vec_dispatch<V_RECIP>(st, a, exact, R); break;
    case V_FROMZERO: vec_dispatch<V_FROMZERO>(st, a, exact, R); break;
    case V_ZSCALE: vec_dispatch<V_ZSCALE>(st, a, exact, R); break;
    }
}

void launch_vec(hipStream_t st, int op, const VecArgs& a, bool exact, const RedScratch& R) {
    launch_vec_t(st, op, a, exact, R);
}
void launch_vec32(hipStream_t st, int op, const VecArgs32& a, const RedScratch& R) {
    launch_vec_t(st, op, a, false, R);
}

__global__ void k_mix(int op, int n, const double* a, float* w, float* w2, float* z, double* x, const float* e,
                      const int* done) {
    if (done && *done) return;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (op == MIX_TO32) {
            const float v = float(a[i]);
            w[i] = v;
            if (w2) w2[i] = v;
            if (z) z[i] = 0.0f;
        } else {
            x[i] += double(e[i]);
        }
    }
}

void launch_mix(hipStream_t st, int op, int n, const double* a, float* w, float* w2, float* z, double* x,
                const float* e, const int* done) {
    if (n <= 0) return;
    int nb = (n + RED_THREADS - 1) / RED_THREADS;
    if (nb > 4 * RED_BLOCKS) nb = 4 * RED_BLOCKS;
    hipLaunchKernelGGL(k_mix, dim3(nb), dim3(RED_THREADS), 0, st, op, n, a, w, w2, z, x, e, done);
}

// ============================================================================ finalize
// One block: reduce the per-block partials in a fixed order (fast) or sum the exact
// bins (exact), then run the solver's scalar step on thread 0.
__device__ __forceinline__ bool dev_converged(const DevScalars& S) {
    // LinearSolver.cpp:114-117
    if (S.iterations < S.minIter) return false;
    return S.finalResidual < S.tolerance || (S.relTol > 0 && S.finalResidual < S.relTol * S.initialResidual);
}

// The solver's scalar step after a reduction (v0, v1 = reduced values). I is the inner
// solve's scalars of mixed precision (FIN_OUTER).
__device__ void fin_step(int op, DevScalars& S, const DevScalars& I, double v0, double v1) {
    switch (op) {
    case FIN_XREF: S.xRef = v0 / double(S.n); break;
    case FIN_INIT:
        S.normFactor = v0 + 1e-20;
        S.residualSum = v1;
        S.initialResidual = S.residualSum / S.normFactor;
        S.finalResidual = S.initialResidual;
        S.iterations = 0;
        if (S.initialResidual < S.tolerance && S.minIter <= 0) S.done = 1;
        break;
    case FIN_PCG_WRA:
        S.wArAold = S.wArA;
        S.wArA = v0;
        if (S.iterations != 0) S.beta = S.wArA / S.wArAold;
        break;
    case FIN_PCG_WAPA:
        S.wApA = v0;
        if (fabs(S.wApA) < VSMALL) {
            S.done = 1;
        } else {
            S.alpha = S.wArA / S.wApA;
        }
        break;
    case FIN_PCG_NORM:
        S.finalResidual = v0 / S.normFactor;
        ++S.iterations;
        if (!(S.iterations < S.maxIter && !dev_converged(S))) S.done = 1;
        break;
    case FIN_BI_RHO:
        S.rhoOld = S.rho;
        S.rho = v0;
        if (S.iterations != 0) {
            if (fabs(S.rhoOld) < VSMALL || fabs(S.omegaB) < VSMALL) {
                S.done = 1;
            } else {
                S.beta = (S.rho / S.rhoOld) * (S.alphaB / S.omegaB);
            }
        }
        break;
    case FIN_BI_ALPHA:
        if (fabs(v0) < VSMALL) {
            S.done = 1;
        } else {
            S.alphaB = S.rho / v0;
        }
        break;
    case FIN_BI_S:
        ++S.iterations;
        S.finalResidual = v0 / S.normFactor;
        if (dev_converged(S)) {
            S.exitPending = 1;
            S.done = 1;
        }
        break;
    case FIN_BI_OMEGA:
        S.tAtA = v0;
        S.omegaB = S.tAtA > VSMALL ? v1 / S.tAtA : 0.0;
        break;
    case FIN_BI_NORM:
        S.finalResidual = v0 / S.normFactor;
        if (!(S.iterations < S.maxIter && !dev_converged(S))) S.done = 1;
        break;
    case FIN_JAC_NORM:
        S.iterations += S.nSweeps;
        S.finalResidual = v0 / S.normFactor;
        if (!(S.iterations < S.maxIter && !dev_converged(S))) S.done = 1;
        break;
    case FIN_OUTER: {  // I holds the inner solve segment that just ended
        const int seg = I.iterations - I.itStart;
        S.iterations += seg;
        ++S.outer;
        S.finalResidual = v0 / S.normFactor;
        // stop on convergence, the iteration budget, or an inner segment that made no progress
        if (!(S.iterations < S.maxIter && !dev_converged(S)) || seg == 0) S.done = 1;
        break;
    }
    default:
        break;
    }
}

// Mixed precision: I ← a fresh inner solve for the current outer residual of S, or done.
__device__ void inner_start(const DevScalars& S, DevScalars& I) {
    const double target = fmax(S.tolerance, S.relTol > 0 ? S.relTol * S.initialResidual : 0.0);
    if (S.keepKrylov && S.outer > 0 && !S.done && !I.exitPending) {
        // reliable update: the recurrence goes on from the replaced (true) residual with
        // the same search direction and ρ; only the segment's stopping point is reset
        I.tolerance = fmax(S.innerRel * S.finalResidual, 0.5 * target);
        I.itStart = I.iterations;
        I.maxIter = I.iterations + (S.maxIter - S.iterations);
        I.done = 0;
        return;
    }
    DevScalars z = {};
    z.normFactor = S.normFactor;
    z.tolerance = fmax(S.innerRel * S.finalResidual, 0.5 * target);
    z.relTol = 0.0;
    z.maxIter = S.maxIter - S.iterations;
    z.minIter = 0;
    z.nSweeps = 1;
    z.n = S.n;
    z.wArA = GREAT;
    z.done = S.done;
    I = z;
}

// Copy of the scalars into host-visible memory (the solve loop watches it).
__device__ __forceinline__ void publish(const DevScalars& S, DevScalars* hostOut) {
    if (hostOut) *hostOut = S;
}

__global__ __launch_bounds__(FIN_THREADS) void k_finalize(int op, int nred, int exact, int nparts, const double* part,
                                                  int maxBlk, const long long* xpart, const int* xnf, int maxBlkExact,
                                                  DevScalars* Sp, double* result, DevScalars* hostOut) {
    DevScalars& S = *Sp;
    if (op == FIN_INNER_START) {  // no reduction: S[1] ← fresh inner solve, or done with S[0]
        if (threadIdx.x == 0) inner_start(S, Sp[1]);
        return;
    }
    if (op != FIN_STORE && S.done) {
        if (threadIdx.x == 0) publish(S, hostOut);
        return;
    }
    __shared__ double vals[MAX_RED];
    __shared__ double sh[FIN_THREADS / WAVE];
    __shared__ long long bins[NBINS];
    __shared__ int nfl;
    for (int q = 0; q < nred; ++q) {
        if (!exact) {
            const double t = fin_sum(part + size_t(q) * maxBlk, nparts, sh);
            if (threadIdx.x == 0) vals[q] = t;
        } else {
            if (threadIdx.x == 0) nfl = 0;
            __syncthreads();
            if (threadIdx.x < NBINS) {
                long long s = 0;
                for (int k = 0; k < nparts; ++k) s += xpart[(size_t(q) * maxBlkExact + k) * NBINS + threadIdx.x];
                bins[threadIdx.x] = s;
            }
            for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
                const int f = xnf[q * maxBlkExact + k];
                if (f) atomicOr(&nfl, f);
            }
            __syncthreads();
            if (threadIdx.x == 0) vals[q] = exact_to_double(bins, nfl);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double v0 = vals[0], v1 = nred > 1 ? vals[1] : 0.0;
    if (op == FIN_STORE) {
        if (result) {
            result[0] = v0;
            if (nred > 1) result[1] = v1;
        }
        return;
    }
    fin_step(op, S, Sp[1], v0, v1);
    publish(S, hostOut);
}

void launch_finalize(hipStream_t st, int op, int nred, bool exact, int nparts, const RedScratch& R, DevScalars* S,
                     DevScalars* hostOut) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, st, op, nred, exact ? 1 : 0, nparts, R.part, R.maxBlk,
                       R.xpart, R.xnf, R.maxBlkExact, S, R.result, hostOut);
}

__global__ void k_init_scalars(DevScalars* S, double tol, double relTol, int maxIter, int minIter, int nSweeps, int n,
                               double innerRel, int keepKrylov) {
    DevScalars z = {};
    z.innerRel = innerRel;
    z.keepKrylov = keepKrylov;
    z.tolerance = tol;
    z.relTol = relTol;
    z.maxIter = maxIter;
    z.minIter = minIter;
    z.nSweeps = nSweeps;
    z.n = n;
    z.wArA = GREAT;  // LinearSolver.cpp:180
    *S = z;
}

__global__ void k_binit_scalars(DevScalars* S, int nSys, const int* n, double tol, double relTol, int maxIter,
                                int minIter, int nSweeps, double innerRel, int keepKrylov) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nSys) return;
    DevScalars z = {};
    z.innerRel = innerRel;
    z.keepKrylov = keepKrylov;
    z.tolerance = tol;
    z.relTol = relTol;
    z.maxIter = maxIter;
    z.minIter = minIter;
    z.nSweeps = nSweeps;
    z.n = n[s];
    z.wArA = GREAT;  // LinearSolver.cpp:180
    S[s] = z;
}

void launch_binit_scalars(hipStream_t st, DevScalars* S, int nSys, const int* n, double tol, double relTol,
                          int maxIter, int minIter, int nSweeps, double innerRel, int keepKrylov) {
    hipLaunchKernelGGL(k_binit_scalars, dim3((nSys + 63) / 64), dim3(64), 0, st, S, nSys, n, tol, relTol, maxIter,
                       minIter, nSweeps, innerRel, keepKrylov);
}

__global__ void k_clear_exit(DevScalars* S) { S->exitPending = 0; }

void launch_clear_exit(hipStream_t st, DevScalars* S) { hipLaunchKernelGGL(k_clear_exit, dim3(1), dim3(1), 0, st, S); }

void launch_init_scalars(hipStream_t st, DevScalars* S, double tol, double relTol, int maxIter, int minIter,
                         int nSweeps, int n, double innerRel, int keepKrylov) {
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, st, S, tol, relTol, maxIter, minIter, nSweeps, n,
                       innerRel, keepKrylov);
}

// ============================================================================ GAMG
// Restriction: rc[I] = Σ res over the fine cells of aggregate I in ascending order,
// which is the order of the reference's `rc[agg[i]] += res[i]` loop (GAMG.cpp:434-435).
template <class T>
__global__ void k_restrict(int nc, const int* mOff, const int* mList, const T* res, T* rc, const int* done, T* zc,
                           const T* rDc, double omega) {
    if (done && *done) return;
    const int I = xcd_block() * blockDim.x + threadIdx.x;
    if (I >= nc) return;
    // members in groups of 8: every index load of a group, then every value load, are in
    // flight together; the sum still adds them one by one in ascending order
    constexpr int U = 8;
    const int b = mOff[I], e = mOff[I + 1];
    T s = T(0);
    for (int k0 = b; k0 < e; k0 += U) {
        int j[U];
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = k0 + u < e ? mList[k0 + u] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = j[u] >= 0 ? res[j[u]] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j[u] >= 0) s += v[u];
    }
    rc[I] = s;
    if (zc) zc[I] = T(omega) * (s * rDc[I]);  // as the first sweep from zero computes it (xload)
}

void launch_restrict(hipStream_t st, int nc, const int* mOff, const int* mList, const double* res, double* rc,
                     const int* done, double* zc, const double* rDc, double omega) {
    if (nc == 0) return;
    hipLaunchKernelGGL(k_restrict<double>, dim3(xcd_grid((nc + 255) / 256)), dim3(256), 0, st, nc, mOff, mList, res,
                       rc, done, zc, rDc, omega);
}
void launch_restrict32(hipStream_t st, int nc, const int* mOff, const int* mList, const float* res, float* rc,
                       const int* done, float* zc, const float* rDc, double omega) {
    if (nc == 0) return;
    hipLaunchKernelGGL(k_restrict<float>, dim3(xcd_grid((nc + 255) / 256)), dim3(256), 0, st, nc, mOff, mList, res,
                       rc, done, zc, rDc, omega);
}

// Galerkin coarsening (GAMG.cpp:279-301) as gathers over fine contributions listed in
// the reference's accumulation order: per coarse cell the fine diagonals (ascending
// cell), then its intra-aggregate faces (ascending face); per coarse face its fine faces
// (ascending face), swapped when the fine face is flipped. cfList entries are
// (fineFace << 1) | flip.
__global__ void k_galerkin_cells(int nc, const int* mOff, const int* mList, const int* iOff, const int* iList,
                                 const double* Fd, const double* Fu, const double* Fl, double* Cd) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nc) return;
    double s = 0.0;
    for (int k = mOff[I]; k < mOff[I + 1]; ++k) s += Fd[mList[k]];
    for (int k = iOff[I]; k < iOff[I + 1]; ++k) {
        const int f = iList[k];
        s += Fu[f] + Fl[f];
    }
    Cd[I] = s;
}

__global__ void k_galerkin_faces(int ncf, const int* cfOff, const int* cfList, const double* Fu, const double* Fl,
                                 double* Cu, double* Cl) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncf) return;
    double u = 0.0, l = 0.0;
    for (int k = cfOff[c]; k < cfOff[c + 1]; ++k) {
        const int e = cfList[k], f = e >> 1;
        if (!(e & 1)) {
            u += Fu[f];
            l += Fl[f];
        } else {
            u += Fl[f];
            l += Fu[f];
        }
    }
    Cu[c] = u;
    Cl[c] = l;
}

void launch_galerkin(hipStream_t st, int nc, int ncf, const int* mOff, const int* mList, const int* iOff,
                     const int* iList, const int* cfOff, const int* cfList, const double* Fd, const double* Fu,
                     const double* Fl, double* Cd, double* Cu, double* Cl) {
    if (nc > 0)
        hipLaunchKernelGGL(k_galerkin_cells, dim3((nc + 255) / 256), dim3(256), 0, st, nc, mOff, mList, iOff, iList,
                           Fd, Fu, Fl, Cd);
    if (ncf > 0)
        hipLaunchKernelGGL(k_galerkin_faces, dim3((ncf + 255) / 256), dim3(256), 0, st, ncf, cfOff, cfList, Fu, Fl, Cu,
                           Cl);
}

// Level-0 coefficients oriented to (l, u) = (min, max) cell (GAMG.cpp:266-278).
__global__ void k_orient0(int nf, const char* ownerIsL, const double* upper, const double* lower, double* Fu,
                          double* Fl) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const bool o = ownerIsL[f] != 0;
    Fu[f] = o ? upper[f] : lower[f];
    Fl[f] = o ? lower[f] : upper[f];
}

void launch_orient0(hipStream_t st, int nf, const char* ownerIsL, const double* upper, const double* lower,
                    double* Fu, double* Fl) {
    if (nf == 0) return;
    hipLaunchKernelGGL(k_orient0, dim3((nf + 255) / 256), dim3(256), 0, st, nf, ownerIsL, upper, lower, Fu, Fl);
}

// Dense coarsest matrix (GAMG.cpp:321-326): lu[i][i] = d[i]; lu[l][u] += upper; lu[u][l] += lower.
// Dense coarsest matrix: diagonal, then the faces in order (rows/cols relative to base).
__device__ void coarse_assemble_body(int n, int nf, const int* l, const int* u, int base, const double* d,
                                     const double* cu, const double* cl, double* lu) {
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) lu[k] = 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) lu[size_t(i) * n + i] = d[i];
    __syncthreads();
    if (threadIdx.x == 0)  // faces in order: a pair (l,u) may repeat only across levels, keep the reference order
        for (int f = 0; f < nf; ++f) {
            const int a = l[f] - base, b = u[f] - base;
            lu[size_t(a) * n + b] += cu[f];
            lu[size_t(b) * n + a] += cl[f];
        }
    __syncthreads();
}

__global__ void k_coarse_assemble(int n, int nf, const int* l, const int* u, const double* d, const double* cu,
                                  const double* cl, double* lu) {
    coarse_assemble_body(n, nf, l, u, 0, d, cu, cl, lu);
}

void launch_coarse_assemble(hipStream_t st, int n, int nf, const int* l, const int* u, const double* d,
                            const double* cu, const double* cl, double* lu) {
    hipLaunchKernelGGL(k_coarse_assemble, dim3(1), dim3(256), 0, st, n, nf, l, u, d, cu, cl, lu);
}

// No-pivot Doolittle LU (GAMG.cpp:334-343). Within step k every (i, j) update is
// independent, so the parallel form performs the same rounded operations.
__device__ void coarse_lu_body(int n, double* lu, double* luT, int* status) {
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double piv = lu[size_t(k) * n + k];
        if (fabs(piv) < VSMALL) {
            if (threadIdx.x == 0) bad = 1;
            break;
        }
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) lu[size_t(i) * n + k] = lu[size_t(i) * n + k] / piv;
        __syncthreads();
        const int m = n - k - 1;
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            const int i = k + 1 + t / m, j = k + 1 + t % m;
            const double fct = lu[size_t(i) * n + k];
            if (fct == 0.0) continue;
            lu[size_t(i) * n + j] -= fct * lu[size_t(k) * n + j];
        }
        __syncthreads();
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) luT[t] = lu[size_t(t % n) * n + t / n];
    if (threadIdx.x == 0 && status) status[0] = bad ? -1 : 0;
}

__global__ void k_coarse_lu(int n, double* lu, double* luT, int* status) { coarse_lu_body(n, lu, luT, status); }

void launch_coarse_lu(hipStream_t st, int n, double* lu, double* luT, int* status) {
    hipLaunchKernelGGL(k_coarse_lu, dim3(1), dim3(256), 0, st, n, lu, luT, status);
}

// Coarsest solve with the LU factors: forward substitution column by column (for every
// row the subtractions run in ascending k, as in GAMG.cpp:413-414), then backward
// substitution column by column from the last one (descending k per row, then the
// division by the pivot).
template <class T>
__device__ void coarse_solve_body(int n, const double* lu, const T* r, T* x, double* g) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) g[i] = double(r[i]);
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double gk = g[k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) g[i] -= lu[size_t(i) * n + k] * gk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        if (threadIdx.x == 0) g[k] /= lu[size_t(k) * n + k];
        __syncthreads();
        const double gk = g[k];
        for (int i = threadIdx.x; i < k; i += blockDim.x) g[i] -= lu[size_t(i) * n + k] * gk;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = T(g[i]);
}

template <class T>
__global__ void k_coarse_solve(int n, const double* lu, const T* r, T* x, const int* done) {
    if (done && *done) return;
    extern __shared__ double g[];
    coarse_solve_body(n, lu, r, x, g);
}

// n ≤ 64: one wavefront, lane i holds g_i and reads row i of the factors (row(k): the
// factor in column k); the column steps exchange g_k by a cross-lane broadcast. Same
// operations and order as k_coarse_solve.
// The steps are branch-free: every lane computes, and a select keeps the result only
// where the step applies (lane i > k forward; lane k's division, lanes i < k backward),
// so a lane's value sees exactly the operations of the branching form. Steps k ≥ n
// change nothing: no lane is both > k and < n, and lanes ≥ n hold g = 0.
template <class T, class Row>
__device__ __forceinline__ void coarse_subst64(int n, Row row, double g, T* x) {
    const int i = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 64; ++k) {
        const double gk = lane_bcast(g, k);
        const double t = g - row(k) * gk;
        g = (i > k && i < n) ? t : g;
    }
#pragma unroll
    for (int k = 63; k >= 0; --k) {
        const double q = g / row(k);
        g = (i == k && k < n) ? q : g;
        const double gk = lane_bcast(g, k);
        const double t = g - row(k) * gk;
        g = (i < k && k < n) ? t : g;
    }
    if (i < n) x[i] = T(g);
}

// ... with the factors of row i in registers
template <class T>
__device__ void coarse_solve64_g(int n, const double* luT, double g, T* x) {
    const int i = threadIdx.x;
    double row[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) row[k] = (k < n && i < n) ? luT[size_t(k) * n + i] : 0.0;
    coarse_subst64(n, [&](int k) { return row[k]; }, g, x);
}

template <class T>
__device__ void coarse_solve64_body(int n, const double* luT, const T* r, T* x) {
    const int i = threadIdx.x;
    coarse_solve64_g(n, luT, i < n ? double(r[i]) : 0.0, x);
}

template <class T>
__global__ __launch_bounds__(64) void k_coarse_solve64(int n, const double* luT, const T* r, T* x, const int* done) {
    if (done && *done) return;
    coarse_solve64_body(n, luT, r, x);
}

// The restriction into a coarsest level of ≤ 64 cells and its substitution in one
// wavefront: lane I sums the residuals of aggregate I (ascending, as k_restrict) into
// its right-hand side and the substitution starts from it (as k_coarse_solve64).
template <class T>
__global__ __launch_bounds__(64) void k_restrict_coarse64(int nc, const int* mOff, const int* mList, const T* res,
                                                          const double* luT, T* x, const int* done) {
    if (done && *done) return;
    const int I = threadIdx.x;
    T s = T(0);
    if (I < nc) {
        for (int k = mOff[I]; k < mOff[I + 1]; ++k) s += res[mList[k]];
    }
    coarse_solve64_g(nc, luT, I < nc ? double(s) : 0.0, x);
}

template <class T>
void launch_restrict_coarse64(hipStream_t st, int nc, const int* mOff, const int* mList, const T* res,
                              const double* luT, T* x, const int* done) {
    hipLaunchKernelGGL(k_restrict_coarse64<T>, dim3(1), dim3(64), 0, st, nc, mOff, mList, res, luT, x, done);
}
template void launch_restrict_coarse64<double>(hipStream_t, int, const int*, const int*, const double*, const double*,
                                               double*, const int*);
template void launch_restrict_coarse64<float>(hipStream_t, int, const int*, const int*, const float*, const double*,
                                              float*, const int*);

// ============================================================================ residual + restriction
// Residual (r − A·x) of CSR row `row` (slots s0 .. s0 + width) by G lanes of a
// wavefront: lane g loads slots g, g+G, …, then every lane of the group subtracts the
// products in slot order, read from their lanes, so the result has the bits of the
// residual gather (sell_row / csr_row). All lanes of the wavefront call.
template <int XL, int G, int IT, class T>
__device__ __forceinline__ T resid_row(const SellArgsT<T>& a, int row, int s0, int width, bool valid) {
    const int g = threadIdx.x % G;
    T p[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int k = g + G * it;
        p[it] = T(0);
        if (k < width) p[it] = a.coef[s0 + k] * xload<XL>(a, a.S.col[s0 + k], T(0));
    }
    T acc = T(0);
    if (valid) acc = a.r[row] - a.diag[row] * xload<XL>(a, row, T(0));
    const int base = (threadIdx.x & (WAVE - 1)) - g;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T t = __shfl(p[it], base + q, WAVE);
            if (q + G * it < width) acc = acc - t;
        }
    }
    return acc;
}

// Residuals of the member rows at positions [p0, p1) of the restriction lists into LDS
// (res[p − p0]); G lanes per row, the whole block of NT threads in step.
template <int XL, int G, int IT, int NT, class T>
__device__ __forceinline__ void resid_members(const SellArgsT<T>& a, const int4* rows, int p0, int p1, T* res) {
    const int grp = threadIdx.x / G;
    for (int q = p0; q < p1; q += NT / G) {
        const int p = q + grp;
        const bool valid = p < p1;
        const int4 rw = valid ? rows[p] : make_int4(0, 0, 0, 0);
        const T v = resid_row<XL, G, IT>(a, rw.x, rw.y, rw.z, valid);
        if (valid && threadIdx.x % G == 0) res[p - p0] = v;
    }
}

// rc[I] for the aggregates of block b's range: member residuals into LDS, then one lane
// per aggregate adds its members in ascending order (as k_restrict).
template <int XL, int G, int IT, class T>
__global__ __launch_bounds__(256) void k_resid_restrict(SellArgsT<T> a, RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    const int b = xcd_block();
    if (b >= R.nBlocks) return;
    __shared__ T res[RR_POS];
    const int2 lo = R.blk[b], hi = R.blk[b + 1];
    const int I = lo.x + int(threadIdx.x);
    int mb = 0, me = 0;
    if (I < hi.x) {  // the aggregate's member range, loaded ahead of the residuals
        mb = R.mOff[I] - lo.y;
        me = R.mOff[I + 1] - lo.y;
    }
    resid_members<XL, G, IT, 256>(a, R.rows, lo.y, hi.y, res);
    __syncthreads();
    if (I < hi.x) {
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        R.rc[I] = s;
        if (R.zc) R.zc[I] = T(R.omega) * (s * R.rDc[I]);
    }
}

// Every member residual of the level into LDS (512 threads, the LU factors copied into
// LDS alongside), then wavefront 0 restricts into the coarsest level (lane I: aggregate
// I) and runs its substitution from registers (as k_restrict_coarse64).
template <int XL, int G, int IT, class T>
__global__ __launch_bounds__(RR_COARSE_THREADS) void k_resid_restrict_coarse64(SellArgsT<T> a, RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    __shared__ T res[RR_COARSE_ROWS];
    __shared__ double lu[64 * 64];
    const int i = threadIdx.x, n = R.nc;
    // the factors (≤ 64·64) in flight together with the residuals' loads, stored after
    constexpr int LU_PER = 64 * 64 / RR_COARSE_THREADS;
    double luv[LU_PER];
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = i + u * RR_COARSE_THREADS;
        luv[u] = k < n * n ? R.luT[k] : 0.0;
    }
    int mb = 0, me = 0;
    if (i < n) {
        mb = R.mOff[i];
        me = R.mOff[i + 1];
    }
    resid_members<XL, G, IT, RR_COARSE_THREADS>(a, R.rows, 0, a.S.n, res);
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = i + u * RR_COARSE_THREADS;
        if (k < n * n) lu[k] = luv[u];
    }
    __syncthreads();
    if (i < WAVE) {
        double row[64];  // LU row i, all read from LDS before the substitution starts
#pragma unroll
        for (int k = 0; k < 64; ++k) row[k] = (k < n && i < n) ? lu[k * n + i] : 0.0;
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        coarse_subst64(n, [&](int k) { return row[k]; }, i < n ? double(s) : 0.0, R.rc);
    }
}

// The same from x = z (XL_PLAIN), followed by the level's damped-Jacobi post-sweep from
// x = z + xc[agg]: one row per lane group, which keeps its row's coefficients and the
// gathered z and agg of its neighbours from the residual for the sweep, and reads xc from
// LDS. Each row's arithmetic is that of the residual and prolongation-sweep gathers
// (sell_row / csr_row with XL_PLAIN, then OP_JACOBI with XL_PROLONG).
template <int G, int IT, class T>
__global__ __launch_bounds__(RR_COARSE_THREADS) void k_resid_restrict_coarse64_post(SellArgsT<T> a,
                                                                                   RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    __shared__ T res[RR_COARSE_THREADS / G];
    __shared__ double lu[64 * 64];
    __shared__ T xcL[64];
    const int i = threadIdx.x, n = R.nc;
    constexpr int LU_PER = 64 * 64 / RR_COARSE_THREADS;
    double luv[LU_PER];
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = i + u * RR_COARSE_THREADS;
        luv[u] = k < n * n ? R.luT[k] : 0.0;
    }
    int mb = 0, me = 0;
    if (i < n) {
        mb = R.mOff[i];
        me = R.mOff[i + 1];
    }
    // this lane group's member row (position p of the restriction lists)
    const int p = i / G, g = i % G;
    const bool valid = p < a.S.n;
    const int4 rw = valid ? R.rows[p] : make_int4(0, 0, 0, 0);
    T cf[IT], zv[IT];
    int ag[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int k = g + G * it;
        cf[it] = T(0);
        zv[it] = T(0);
        ag[it] = 0;
        if (k < rw.z) {
            const int j = a.S.col[rw.y + k];
            cf[it] = a.coef[rw.y + k];
            zv[it] = a.x[j];
            ag[it] = a.agg[j];
        }
    }
    T rr = T(0), dd = T(1), zr = T(0), acc = T(0);
    int agr = 0;
    if (valid) {
        rr = a.r[rw.x];
        dd = a.diag[rw.x];
        zr = a.x[rw.x];
        agr = a.agg[rw.x];
        acc = rr - dd * zr;
    }
    const int base = (i & (WAVE - 1)) - g;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const T pr = cf[it] * zv[it];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T t = __shfl(pr, base + q, WAVE);
            if (q + G * it < rw.z) acc = acc - t;
        }
    }
    if (valid && g == 0) res[p] = acc;
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = i + u * RR_COARSE_THREADS;
        if (k < n * n) lu[k] = luv[u];
    }
    __syncthreads();
    if (i < WAVE) {
        double row[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) row[k] = (k < n && i < n) ? lu[k * n + i] : 0.0;
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        coarse_subst64(n, [&](int k) { return row[k]; }, i < n ? double(s) : 0.0, xcL);
        if (i < n) R.rc[i] = xcL[i];
    }
    __syncthreads();
    T acc2 = rr;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const T pr = cf[it] * (zv[it] + xcL[ag[it]]);
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T t = __shfl(pr, base + q, WAVE);
            if (q + G * it < rw.z) acc2 = acc2 - t;
        }
    }
    if (valid) {
        const T J = acc2 / dd;
        const T xo = zr + xcL[agr];
        const T v = xo + T(a.omega) * (J - xo);
        if (g == 0 && a.out) a.out[rw.x] = v;
        if (a.pushZ) push_prolongation(a, rw.x, v, g, G);
    }
}

// The same on the packed level-0 rows (SellView::colP): one lane per member row, reached
// through the restriction list (R.mList), with sell_row_packed's arithmetic.
template <int XL, int PW, class T>
__global__ __launch_bounds__(256) void k_resid_restrict_packed(SellArgsT<T> a, RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    const int b = xcd_block();
    if (b >= R.nBlocks) return;
    __shared__ T res[RR_POS];
    const int2 lo = R.blk[b], hi = R.blk[b + 1];
    const int I = lo.x + int(threadIdx.x);
    int mb = 0, me = 0;
    if (I < hi.x) {
        mb = R.mOff[I] - lo.y;
        me = R.mOff[I + 1] - lo.y;
    }
    for (int p = lo.y + int(threadIdx.x); p < hi.y; p += 256) res[p - lo.y] = sell_row_packed<OP_RESID, XL, PW>(a, R.mList[p], T(0));
    __syncthreads();
    if (I < hi.x) {
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        R.rc[I] = s;
        if (R.zc) R.zc[I] = T(R.omega) * (s * R.rDc[I]);
    }
}

// The same on a SELL level's own rows (one lane per member row, sell_row's arithmetic).
template <int XL, int U, class T>
__global__ __launch_bounds__(256) void k_resid_restrict_sell(SellArgsT<T> a, RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    const int b = xcd_block();
    if (b >= R.nBlocks) return;
    __shared__ T res[RR_POS];
    const int2 lo = R.blk[b], hi = R.blk[b + 1];
    const int I = lo.x + int(threadIdx.x);
    int mb = 0, me = 0;
    if (I < hi.x) {
        mb = R.mOff[I] - lo.y;
        me = R.mOff[I + 1] - lo.y;
    }
    for (int p = lo.y + int(threadIdx.x); p < hi.y; p += 256) res[p - lo.y] = sell_row<OP_RESID, XL, U>(a, R.mList[p], T(0));
    __syncthreads();
    if (I < hi.x) {
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        R.rc[I] = s;
        if (R.zc) R.zc[I] = T(R.omega) * (s * R.rDc[I]);
    }
}

// The same on a SELL level's packed rows (SellView::colS), sell_row_pk's arithmetic.
template <int XL, class T>
__global__ __launch_bounds__(256) void k_resid_restrict_pk(SellArgsT<T> a, RestrictArgsT<T> R) {
    if (a.done && *a.done) return;
    const int b = xcd_block();
    if (b >= R.nBlocks) return;
    __shared__ T res[RR_POS];
    const int2 lo = R.blk[b], hi = R.blk[b + 1];
    const int I = lo.x + int(threadIdx.x);
    int mb = 0, me = 0;
    if (I < hi.x) {
        mb = R.mOff[I] - lo.y;
        me = R.mOff[I + 1] - lo.y;
    }
    for (int p = lo.y + int(threadIdx.x); p < hi.y; p += 256) {
        const int r = R.mList[p];
        res[p - lo.y] = sell_row_pk<OP_RESID, XL>(a, r / WAVE, r % WAVE, r, T(0));
    }
    __syncthreads();
    if (I < hi.x) {
        T s = T(0);
        for (int k = mb; k < me; ++k) s += res[k];
        R.rc[I] = s;
        if (R.zc) R.zc[I] = T(R.omega) * (s * R.rDc[I]);
    }
}

template <int XL, class T>
static void rr_sell(hipStream_t st, const SellArgsT<T>& a, const RestrictArgsT<T>& R) {
    if (a.S.colS) {
        hipLaunchKernelGGL((k_resid_restrict_pk<XL, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
        return;
    }
    if (a.S.maxWidth > 8)
        hipLaunchKernelGGL((k_resid_restrict_sell<XL, 16, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
    else
        hipLaunchKernelGGL((k_resid_restrict_sell<XL, 4, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
}

// ============================================================================ last two levels
// coarse_subst64's steps (same operations, order and selects) with the factors read
// from LDS (lu[k·n + i]) and the loops unrolled four steps at a time, for a workgroup
// whose register budget cannot hold a factor row.
template <class T>
__device__ __forceinline__ void coarse_subst64_lds(int n, const double* lu, double g, T* x) {
    const int i = threadIdx.x, ic = i < n ? i : 0;
#pragma unroll 4
    for (int k = 0; k < 64; ++k) {
        const double gk = lane_bcast(g, k);
        const double rk = k < n ? lu[k * n + ic] : 0.0;
        const double t = g - rk * gk;
        g = (i > k && i < n) ? t : g;
    }
#pragma unroll 4
    for (int k = 63; k >= 0; --k) {
        const double rk = k < n ? lu[k * n + ic] : 0.0;
        const double q = g / rk;
        g = (i == k && k < n) ? q : g;
        const double gk = lane_bcast(g, k);
        const double t = g - rk * gk;
        g = (i < k && k < n) ? t : g;
    }
    if (i < n) x[i] = T(g);
}

// acc0 − Σ coef·x(j) over CSR row `row`, by the 2 lanes g of a group (lane g takes slots
// g, g+2, … of each batch of 2·LAST2_IT; every lane subtracts the products in slot
// order, read from their lanes): csr_row's arithmetic. All lanes of the wavefront call,
// with the same batch count.
template <class T, class XF>
__device__ __forceinline__ T last2_row(const int* off, const int* col, const T* coef, int row, bool valid, int g,
                                       T acc0, int batches, XF xf) {
    int s0 = 0, len = 0;
    if (valid) {
        s0 = off[row];
        len = off[row + 1] - s0;
    }
    T acc = acc0;
    const int base = (threadIdx.x & (WAVE - 1)) - g;
    for (int b = 0; b < batches; ++b) {
        const int k0 = b * 2 * LAST2_IT;
        T p[LAST2_IT];
#pragma unroll
        for (int it = 0; it < LAST2_IT; ++it) {
            const int k = k0 + g + 2 * it;
            p[it] = T(0);
            if (k < len) p[it] = coef[s0 + k] * xf(col[s0 + k]);
        }
#pragma unroll
        for (int it = 0; it < LAST2_IT; ++it) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const T t = __shfl(p[it], base + q, WAVE);
                if (k0 + q + 2 * it < len) acc = acc - t;
            }
        }
    }
    return acc;
}

template <class T>
__global__ __launch_bounds__(LAST2_THREADS) void k_vcycle_last2(Last2ArgsT<T> a) {
    if (a.done && *a.done) return;
    __shared__ T sZA[LAST2_ROWS], sRes[LAST2_ROWS], sRB[LAST2_ROWS], sZB[LAST2_ROWS], sXB[LAST2_ROWS];
    __shared__ int sAggA[LAST2_ROWS], sAggB[LAST2_ROWS];
    __shared__ double sLU[64 * 64];
    __shared__ T sXC[64];
    const int t = threadIdx.x, grp = t / 2, g = t % 2;
    const int nA = a.nA, nB = a.nB, nC = a.nC;
    const T om = T(a.omega);
    // staging: z_A, the aggregate maps, the LU factors, and each row's r and diagonals
    constexpr int LU_PER = 64 * 64 / LAST2_THREADS;
    double luv[LU_PER];
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = t + u * LAST2_THREADS;
        luv[u] = k < nC * nC ? a.luT[k] : 0.0;
    }
    if (t < nA) {
        sZA[t] = a.zA[t];
        sAggA[t] = a.aggA[t];
    }
    if (t < nB) sAggB[t] = a.aggB[t];
    const bool vA = grp < nA, vB = grp < nB;
    T rA = T(0), dA = T(1), dB = T(1), rDB = T(0);
    if (vA) {
        rA = a.rA[grp];
        dA = a.diagA[grp];
    }
    if (vB) dB = a.diagB[grp];
    if (t < nB) rDB = a.rDB[t];
#pragma unroll
    for (int u = 0; u < LU_PER; ++u) {
        const int k = t + u * LAST2_THREADS;
        if (k < nC * nC) sLU[k] = luv[u];
    }
    __syncthreads();
    {  // A's residual
        const T res = last2_row(a.offA, a.colA, a.coefA, grp, vA, g, vA ? rA - dA * sZA[grp] : T(0), a.batchesA,
                                [&](int j) { return sZA[j]; });
        if (vA && g == 0) sRes[grp] = res;
    }
    __syncthreads();
    if (t < nB) {  // restriction A → B, members ascending, and z_B
        T s = T(0);
        for (int k = a.mOffA[t]; k < a.mOffA[t + 1]; ++k) s += sRes[a.mListA[k]];
        sRB[t] = s;
        sZB[t] = om * (s * rDB);
    }
    __syncthreads();
    {  // B's residual
        const T res = last2_row(a.offB, a.colB, a.coefB, grp, vB, g, vB ? sRB[grp] - dB * sZB[grp] : T(0), a.batchesB,
                                [&](int j) { return sZB[j]; });
        __syncthreads();  // every lane has read sRes (A's residuals) before it is overwritten
        if (vB && g == 0) sRes[grp] = res;
    }
    __syncthreads();
    if (t < WAVE) {  // restriction B → C and the substitution (factors read from LDS)
        T s = T(0);
        if (t < nC)
            for (int k = a.mOffB[t]; k < a.mOffB[t + 1]; ++k) s += sRes[a.mListB[k]];
        coarse_subst64_lds(nC, sLU, t < nC ? double(s) : 0.0, sXC);
    }
    __syncthreads();
    {  // B's post-sweep from z_B + x_C[agg]
        auto xf = [&](int j) { return sZB[j] + sXC[sAggB[j]]; };
        const T acc = last2_row(a.offB, a.colB, a.coefB, grp, vB, g, vB ? sRB[grp] : T(0), a.batchesB, xf);
        if (vB && g == 0) {
            const T J = acc / dB;
            const T xo = xf(grp);
            sXB[grp] = xo + om * (J - xo);
        }
    }
    __syncthreads();
    {  // A's post-sweep from z_A + x_B[agg]
        auto xf = [&](int j) { return sZA[j] + sXB[sAggA[j]]; };
        const T acc = last2_row(a.offA, a.colA, a.coefA, grp, vA, g, vA ? rA : T(0), a.batchesA, xf);
        if (vA) {
            const T J = acc / dA;
            const T xo = xf(grp);
            const T v = xo + om * (J - xo);
            if (g == 0 && a.xA) a.xA[grp] = v;
            if (a.pushZ) {  // members g, g + 2, … of the finer cells of aggregate grp
                for (int k = a.pushOff[grp] + g, e = a.pushOff[grp + 1]; k < e; k += 2) {
                    const int m = a.pushList[k];
                    a.pushZ[m] = a.pushZ[m] + v;
                }
            }
        }
    }
}

template <class T> void launch_vcycle_last2(hipStream_t st, const Last2ArgsT<T>& a) {
    hipLaunchKernelGGL(k_vcycle_last2<T>, dim3(1), dim3(LAST2_THREADS), 0, st, a);
}
template void launch_vcycle_last2<double>(hipStream_t, const Last2ArgsT<double>&);
template void launch_vcycle_last2<float>(hipStream_t, const Last2ArgsT<float>&);

template <int XL, class T>
static void rr_packed(hipStream_t st, const SellArgsT<T>& a, const RestrictArgsT<T>& R) {
    if (a.S.colPW == 2)
        hipLaunchKernelGGL((k_resid_restrict_packed<XL, 4, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
    else
        hipLaunchKernelGGL((k_resid_restrict_packed<XL, 8, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
}

template <int XL, int G, int IT, class T>
static void rr_launch(hipStream_t st, bool coarse, const SellArgsT<T>& a, const RestrictArgsT<T>& R) {
    if (coarse && XL == XL_PLAIN && R.post)
        hipLaunchKernelGGL((k_resid_restrict_coarse64_post<G, IT, T>), dim3(1), dim3(RR_COARSE_THREADS), 0, st, a, R);
    else if (coarse)
        hipLaunchKernelGGL((k_resid_restrict_coarse64<XL, G, IT, T>), dim3(1), dim3(RR_COARSE_THREADS), 0, st, a, R);
    else
        hipLaunchKernelGGL((k_resid_restrict<XL, G, IT, T>), dim3(xcd_grid(R.nBlocks)), dim3(256), 0, st, a, R);
}
// the grid form takes the level's lanes per row (the block ranges are cut for them); the
// coarsest form as few lanes per row as its longest row allows, so that its one block
// has every row in flight at once
template <int XL, class T>
static void rr_dispatch(hipStream_t st, bool coarse, int G, int IT, const SellArgsT<T>& a, const RestrictArgsT<T>& R) {
    if (coarse) {
        const int w = a.S.maxWidth;
        if (w <= 16) rr_launch<XL, 4, 4>(st, true, a, R);
        else if (w <= 32) rr_launch<XL, 4, 8>(st, true, a, R);
        else rr_launch<XL, 8, 8>(st, true, a, R);  // rows ≤ 64 slots (the host checks)
        static_assert(RR_COARSE_THREADS / 4 >= 64, "lane groups of the coarsest form");
    } else if (G == 4 && IT <= 4) rr_launch<XL, 4, 4>(st, false, a, R);
    else if (G == 8 && IT <= 4) rr_launch<XL, 8, 4>(st, false, a, R);
    else if (G == 16 && IT <= 4) rr_launch<XL, 16, 4>(st, false, a, R);
    else rr_launch<XL, 16, 8>(st, false, a, R);
}
template <class T>
static void rr_any(hipStream_t st, bool coarse, int xl, int G, int IT, const SellArgsT<T>& a,
                   const RestrictArgsT<T>& R) {
    if (xl == XL_FROMZERO) rr_dispatch<XL_FROMZERO>(st, coarse, G, IT, a, R);
    else rr_dispatch<XL_PLAIN>(st, coarse, G, IT, a, R);
}
template <class T>
void launch_resid_restrict(hipStream_t st, int xl, int G, int IT, const SellArgsT<T>& a, const RestrictArgsT<T>& R) {
    if (R.nBlocks == 0) return;
    if (a.S.colP) {  // packed level-0 rows
        if (xl == XL_FROMZERO) rr_packed<XL_FROMZERO>(st, a, R);
        else rr_packed<XL_PLAIN>(st, a, R);
        return;
    }
    if (a.S.csrG == 0) {  // a SELL level's own rows
        if (xl == XL_FROMZERO) rr_sell<XL_FROMZERO>(st, a, R);
        else rr_sell<XL_PLAIN>(st, a, R);
        return;
    }
    rr_any(st, false, xl, G, IT, a, R);
}
template <class T>
void launch_resid_restrict_coarse64(hipStream_t st, int xl, int G, int IT, const SellArgsT<T>& a,
                                    const RestrictArgsT<T>& R) {
    rr_any(st, true, xl, G, IT, a, R);
}
template void launch_resid_restrict<double>(hipStream_t, int, int, int, const SellArgs&, const RestrictArgsT<double>&);
template void launch_resid_restrict<float>(hipStream_t, int, int, int, const SellArgs32&, const RestrictArgsT<float>&);
template void launch_resid_restrict_coarse64<double>(hipStream_t, int, int, int, const SellArgs&,
                                                     const RestrictArgsT<double>&);
template void launch_resid_restrict_coarse64<float>(hipStream_t, int, int, int, const SellArgs32&,
                                                    const RestrictArgsT<float>&);

template <class T>
static void coarse_solve_t(hipStream_t st, int n, const double* lu, const double* luT, const T* r, T* x,
                           const int* done) {
    if (n <= 64)
        hipLaunchKernelGGL(k_coarse_solve64<T>, dim3(1), dim3(64), 0, st, n, luT, r, x, done);
    else
        hipLaunchKernelGGL(k_coarse_solve<T>, dim3(1), dim3(256), sizeof(double) * n, st, n, lu, r, x, done);
}
void launch_coarse_solve(hipStream_t st, int n, const double* lu, const double* luT, const double* r, double* x,
                         const int* done) {
    coarse_solve_t(st, n, lu, luT, r, x, done);
}
void launch_coarse_solve32(hipStream_t st, int n, const double* lu, const double* luT, const float* r, float* x,
                           const int* done) {
    coarse_solve_t(st, n, lu, luT, r, x, done);
}

// ============================================================================ batches
template <int OP, int NRED, class T>
__global__ __launch_bounds__(RED_THREADS) void k_bvec(VecArgsT<T> A, const int4* chunks, double* part, int maxChunks) {
    const int4 ch = chunks[blockIdx.x];
    VecArgsT<T> a = A;
    a.S = A.S + ch.x;
    if constexpr (OP == V_BI_XEXIT) {
        if (!a.S->exitPending) return;
    } else {
        if (a.S->done) return;
    }
    double s0 = 0.0, s1 = 0.0;
    for (int i = ch.y + threadIdx.x; i < ch.z; i += blockDim.x) {
        double r0 = 0.0, r1 = 0.0;
        vec_elem<OP>(a, i, r0, r1);
        if constexpr (NRED > 0) s0 += r0;
        if constexpr (NRED > 1) s1 += r1;
    }
    if constexpr (NRED > 0) {
        __shared__ double sh[RED_THREADS / WAVE];
        const double t0 = block_sum(s0, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = t0;
        if constexpr (NRED > 1) {
            const double t1 = block_sum(s1, sh);
            if (threadIdx.x == 0) part[maxChunks + blockIdx.x] = t1;
        }
    }
}

template <int OP, class T>
static void bvec_dispatch(hipStream_t st, const VecArgsT<T>& a, const BatchRed& R) {
    constexpr int NRED = (OP == V_INIT || OP == V_BI_DOT2) ? 2
                         : (OP == V_SUM_X || OP == V_PRECOND_DOT || OP == V_DOT || OP == V_PCG_XR || OP == V_BI_S ||
                            OP == V_BMINUSAX_NORM || OP == V_BI_XR)
                             ? 1
                             : 0;
    if (R.nChunks == 0) return;
    hipLaunchKernelGGL((k_bvec<OP, NRED, T>), dim3(R.nChunks), dim3(RED_THREADS), 0, st, a, R.chunks, R.part,
                       R.maxChunks);
}

template <class T>
static void launch_bvec_t(hipStream_t st, int op, const VecArgsT<T>& a, const BatchRed& R) {
    switch (op) {
    case V_SUM_X: bvec_dispatch<V_SUM_X>(st, a, R); break;
    case V_INIT: bvec_dispatch<V_INIT>(st, a, R); break;
    case V_PRECOND_DOT: bvec_dispatch<V_PRECOND_DOT>(st, a, R); break;
    case V_DOT: bvec_dispatch<V_DOT>(st, a, R); break;
    case V_PCG_P: bvec_dispatch<V_PCG_P>(st, a, R); break;
    case V_PCG_XR: bvec_dispatch<V_PCG_XR>(st, a, R); break;
    case V_BI_P: bvec_dispatch<V_BI_P>(st, a, R); break;
    case V_SCALE: bvec_dispatch<V_SCALE>(st, a, R); break;
    case V_BI_S: bvec_dispatch<V_BI_S>(st, a, R); break;
    case V_BI_XEXIT: bvec_dispatch<V_BI_XEXIT>(st, a, R); break;
    case V_BI_DOT2: bvec_dispatch<V_BI_DOT2>(st, a, R); break;
    case V_BI_XR: bvec_dispatch<V_BI_XR>(st, a, R); break;
    case V_BMINUSAX_NORM: bvec_dispatch<V_BMINUSAX_NORM>(st, a, R); break;
    case V_COPY: bvec_dispatch<V_COPY>(st, a, R); break;
    case V_FILL_XREF: bvec_dispatch<V_FILL_XREF>(st, a, R); break;
    }
}
void launch_bvec(hipStream_t st, int op, const VecArgs& a, const BatchRed& R) { launch_bvec_t(st, op, a, R); }
void launch_bvec32(hipStream_t st, int op, const VecArgs32& a, const BatchRed& R) { launch_bvec_t(st, op, a, R); }

__global__ void k_bfinalize(int op, int nred, const double* part, int maxChunks, const int* sysChunkOff,
                            DevScalars* S, DevScalars* I, int nSys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nSys) return;
    if (op == FIN_INNER_START) {
        inner_start(S[s], I[s]);
        return;
    }
    const int done = S[s].done;
    // the system's chunk partials in order, 16 loads in flight at a time
    constexpr int U = 16;
    double v[MAX_RED] = {0.0, 0.0};
    const int b = sysChunkOff[s], e = sysChunkOff[s + 1];
    for (int q = 0; q < nred; ++q)
        for (int k0 = b; k0 < e; k0 += U) {
            double t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) t[u] = k0 + u < e ? part[size_t(q) * maxChunks + k0 + u] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k0 + u < e) v[q] += t[u];
        }
    if (done) return;
    fin_step(op, S[s], I ? I[s] : S[s], v[0], v[1]);
}

void launch_bfinalize(hipStream_t st, int op, int nred, const BatchRed& R, DevScalars* S, DevScalars* I) {
    if (R.nSys == 0) return;
    hipLaunchKernelGGL(k_bfinalize, dim3((R.nSys + 63) / 64), dim3(64), 0, st, op, nred, R.part, R.maxChunks,
                       R.sysChunkOff, S, I, R.nSys);
}

__global__ void k_bcount(const DevScalars* S, int nSys, int* nDone) {
    int c = 0;
    for (int s = threadIdx.x; s < nSys; s += blockDim.x) c += S[s].done ? 1 : 0;
    __shared__ double sh[256 / WAVE];
    const double t = block_sum(double(c), sh);
    if (threadIdx.x == 0) nDone[0] = int(t);
}

void launch_bcount(hipStream_t st, const DevScalars* S, int nSys, int* nDone) {
    hipLaunchKernelGGL(k_bcount, dim3(1), dim3(256), 0, st, S, nSys, nDone);
}

__global__ void k_bmix(int op, const int4* chunks, const DevScalars* S, const double* a, float* w, float* w2,
                       float* z, double* x, const float* e) {
    const int4 ch = chunks[blockIdx.x];
    if (S[ch.x].done) return;
    for (int i = ch.y + threadIdx.x; i < ch.z; i += blockDim.x) {
        if (op == MIX_TO32) {
            const float v = float(a[i]);
            w[i] = v;
            if (w2) w2[i] = v;
            if (z) z[i] = 0.0f;
        } else {
            x[i] += double(e[i]);
        }
    }
}

void launch_bmix(hipStream_t st, int op, const BatchRed& R, const DevScalars* S, const double* a, float* w,
                 float* w2, float* z, double* x, const float* e) {
    if (R.nChunks == 0) return;
    hipLaunchKernelGGL(k_bmix, dim3(R.nChunks), dim3(RED_THREADS), 0, st, op, R.chunks, S, a, w, w2, z, x, e);
}

__global__ __launch_bounds__(256) void k_bcoarse_factor(const CoarseDesc* D, double* lu, double* luT, int* status) {
    const CoarseDesc d = D[blockIdx.x];
    coarse_assemble_body(d.n, d.nf, d.l, d.u, d.rowBase, d.d + d.rowBase, d.cu, d.cl, lu + d.luOff);
    coarse_lu_body(d.n, lu + d.luOff, luT + d.luOff, status + blockIdx.x);
}

void launch_bcoarse_factor(hipStream_t st, const CoarseDesc* D, int nSys, double* lu, double* luT, int* status) {
    if (nSys > 0) hipLaunchKernelGGL(k_bcoarse_factor, dim3(nSys), dim3(256), 0, st, D, lu, luT, status);
}

// One workgroup per system; n ≤ 64 uses the one-wave register form, larger n the LDS
// form (dynamic LDS of maxN doubles).
template <class T>
__global__ __launch_bounds__(256) void k_bcoarse_solve(const CoarseDesc* D, const double* lu, const double* luT,
                                                       const T* const* rLev, T* const* xLev, const DevScalars* S) {
    if (S[blockIdx.x].done) return;
    const CoarseDesc d = D[blockIdx.x];
    const T* r = rLev[d.level] + d.rowBase;
    T* x = xLev[d.level] + d.rowBase;
    if (d.n <= 64) {
        if (threadIdx.x < WAVE) coarse_solve64_body(d.n, luT + d.luOff, r, x);
    } else {
        extern __shared__ double g[];
        coarse_solve_body(d.n, lu + d.luOff, r, x, g);
    }
}

template <class T>
static void bcoarse_solve_t(hipStream_t st, const CoarseDesc* D, int nSys, int maxN, const double* lu,
                            const double* luT, const T* const* rLev, T* const* xLev, const DevScalars* S) {
    if (nSys == 0) return;
    const size_t shm = maxN > 64 ? sizeof(double) * size_t(maxN) : 0;
    hipLaunchKernelGGL(k_bcoarse_solve<T>, dim3(nSys), dim3(256), shm, st, D, lu, luT, rLev, xLev, S);
}
void launch_bcoarse_solve(hipStream_t st, const CoarseDesc* D, int nSys, int maxN, const double* lu,
                          const double* luT, const double* const* rLev, double* const* xLev, const DevScalars* S) {
    bcoarse_solve_t(st, D, nSys, maxN, lu, luT, rLev, xLev, S);
}
void launch_bcoarse_solve32(hipStream_t st, const CoarseDesc* D, int nSys, int maxN, const double* lu,
                            const double* luT, const float* const* rLev, float* const* xLev, const DevScalars* S) {
    bcoarse_solve_t(st, D, nSys, maxN, lu, luT, rLev, xLev, S);
}

// ============================================================================ V-cycle tail
// The small levels of the V-cycle in one workgroup: every phase (sweep, residual,
// restriction, coarsest solve, prolongation sweep) loops over the level's rows with
// the block's threads and ends with a barrier. Per-row arithmetic is sell_row, so the
// results equal the per-level kernels bit for bit.
template <int OP, int XL, class T>
__device__ __forceinline__ void tail_phase(const SellArgsT<T>& a) {
    for (int row = threadIdx.x; row < a.S.n; row += blockDim.x) sell_row<OP, XL, 16>(a, row, T(0));
    __syncthreads();
}

// LDS holds [coarsest LU (n ≤ 64) or substitution vector, doubles] then the tail levels'
// vectors (r, two work buffers, x) in the level precision T. The coarsest solve runs in
// FP64 whatever T is.
// A tail level with its matrix pointers redirected into LDS when it was preloaded.
template <class T>
__device__ __forceinline__ TailLevelT<T> tail_level(const TailLevelT<T>& G, char* base) {
    TailLevelT<T> L = G;
    if (L.ldsMat) {
        L.S.sliceOff = reinterpret_cast<const int*>(base + L.ldsRowOff);
        L.S.col = reinterpret_cast<const int*>(base + L.ldsCol);
        L.S.col16 = nullptr;
        L.coef = reinterpret_cast<const T*>(base + L.ldsCoef);
        L.diag = reinterpret_cast<const T*>(base + L.ldsDiag);
        L.rD = reinterpret_cast<const T*>(base + L.ldsRD);
        if (L.agg) L.agg = reinterpret_cast<const int*>(base + L.ldsAgg);
        if (L.mOff) L.mOff = reinterpret_cast<const int*>(base + L.ldsMOff);
        if (L.mList) L.mList = reinterpret_cast<const int*>(base + L.ldsMList);
    }
    return L;
}

template <class T>
__global__ __launch_bounds__(TAIL_THREADS) void k_vcycle_tail(TailArgsT<T> t) {
    if (t.done && *t.done) return;
    extern __shared__ double lds[];
    T* ldsT = reinterpret_cast<T*>(lds);
    char* base = reinterpret_cast<char*>(lds);
    const int nl = t.nLevels;
    const int cur = pre_smoothed_buffer(t.nPre), freeB = 1 - cur;
    {
        // matrices (and the coarsest factors, and rTop when its address is fixed) into
        // LDS: every later phase then reads LDS only. One chunk per wave at a time, all
        // its loads in flight before its stores.
        constexpr int PER = TAIL_CHUNK_WORDS / WAVE;
        unsigned* lw = reinterpret_cast<unsigned*>(lds);
        const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
        for (int c = wave; c < t.nChunks; c += nw) {
            const TailChunk ch = t.chunks[c];
            unsigned v[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = lane + u * WAVE;
                v[u] = i < ch.n ? ch.src[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = lane + u * WAVE;
                if (i < ch.n) lw[ch.dst + i] = v[u];
            }
        }
        if (!t.rInChunks) {
            const TailLevelT<T>& L0 = t.L[0];
            for (int i = threadIdx.x; i < L0.S.n; i += blockDim.x) ldsT[L0.ldsR + i] = t.rTop[i];
        }
        __syncthreads();
    }
    for (int l = 0; l + 1 < nl; ++l) {
        const TailLevelT<T> L = tail_level(t.L[l], base);
        SellArgsT<T> a;
        a.S = L.S;
        a.coef = L.coef;
        a.diag = L.diag;
        a.rD = L.rD;
        a.r = ldsT + L.ldsR;
        a.omega = t.omega;
        T* bufs[2] = {ldsT + L.ldsA, ldsT + L.ldsB};
        int c = 0;
        for (int s = 1; s < t.nPre; ++s) {
            const int dst = (s == 1) ? 0 : 1 - c;
            a.x = bufs[c];
            a.out = bufs[dst];
            if (s == 1)
                tail_phase<OP_JACOBI, XL_FROMZERO>(a);
            else
                tail_phase<OP_JACOBI, XL_PLAIN>(a);
            c = dst;
        }
        a.x = bufs[cur];
        a.out = bufs[freeB];
        if (t.nPre == 1)
            tail_phase<OP_RESID, XL_FROMZERO>(a);
        else
            tail_phase<OP_RESID, XL_PLAIN>(a);
        const T* res = bufs[freeB];
        T* rc = ldsT + t.L[l + 1].ldsR;
        for (int I = threadIdx.x; I < L.nNext; I += blockDim.x) {
            T sum = T(0);
            for (int k = L.mOff[I]; k < L.mOff[I + 1]; ++k) sum += res[L.mList[k]];
            rc[I] = sum;
        }
        __syncthreads();
    }
    {  // coarsest level: LU substitution in FP64 (same operations as k_coarse_solve)
        const TailLevelT<T>& LC = t.L[nl - 1];
        const T* r = ldsT + LC.ldsR;
        T* x = (nl == 1) ? t.xTop : ldsT + LC.ldsX;
        const int n = t.nCoarse;
        double* cl = lds;
        if (n <= 64) {  // the factors were preloaded into cl
            if (threadIdx.x < WAVE) {
                const int i = threadIdx.x;
                double g = i < n ? double(r[i]) : 0.0;
                for (int k = 0; k < n; ++k) {
                    const double gk = lane_bcast(g, k);
                    if (i > k && i < n) g -= cl[k * n + i] * gk;
                }
                for (int k = n - 1; k >= 0; --k) {
                    if (i == k) g /= cl[k * n + k];
                    const double gk = lane_bcast(g, k);
                    if (i < k) g -= cl[k * n + i] * gk;
                }
                if (i < n) x[i] = T(g);
            }
        } else {
            double* g = cl;
            for (int i = threadIdx.x; i < n; i += blockDim.x) g[i] = double(r[i]);
            __syncthreads();
            for (int k = 0; k < n; ++k) {
                const double gk = g[k];
                for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) g[i] -= t.lu[size_t(i) * n + k] * gk;
                __syncthreads();
            }
            for (int k = n - 1; k >= 0; --k) {
                if (threadIdx.x == 0) g[k] /= t.lu[size_t(k) * n + k];
                __syncthreads();
                const double gk = g[k];
                for (int i = threadIdx.x; i < k; i += blockDim.x) g[i] -= t.lu[size_t(i) * n + k] * gk;
                __syncthreads();
            }
            for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = T(g[i]);
        }
        __syncthreads();
    }
    for (int l = nl - 2; l >= 0; --l) {
        const TailLevelT<T> L = tail_level(t.L[l], base);
        SellArgsT<T> a;
        a.S = L.S;
        a.coef = L.coef;
        a.diag = L.diag;
        a.rD = L.rD;
        a.r = ldsT + L.ldsR;
        a.omega = t.omega;
        a.xc = ldsT + t.L[l + 1].ldsX;
        a.agg = L.agg;
        T* bufs[2] = {ldsT + L.ldsA, ldsT + L.ldsB};
        T* xout = (l == 0) ? t.xTop : ldsT + L.ldsX;
        int src = cur, other = freeB;
        for (int s = 0; s < t.nPost; ++s) {
            a.out = (s == t.nPost - 1) ? xout : bufs[other];
            a.x = bufs[src];
            if (s == 0) {
                if (t.nPre == 1)
                    tail_phase<OP_JACOBI, XL_ZEROPROLONG>(a);
                else
                    tail_phase<OP_JACOBI, XL_PROLONG>(a);
            } else {
                tail_phase<OP_JACOBI, XL_PLAIN>(a);
            }
            if (s != t.nPost - 1) {
                const int tmp = src;
                src = other;
                other = tmp;
            }
        }
    }
}

template <class T>
static void launch_tail_t(hipStream_t st, const TailArgsT<T>& t) {
    static const bool attr = [] {  // once per process (thread-safe initialisation)
        (void)hipFuncSetAttribute((const void*)k_vcycle_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  TAIL_LDS_BYTES);
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL(k_vcycle_tail<T>, dim3(1), dim3(TAIL_THREADS), t.ldsBytes, st, t);
}

void launch_vcycle_tail(hipStream_t st, const TailArgs& t) { launch_tail_t(st, t); }
void launch_vcycle_tail32(hipStream_t st, const TailArgs32& t) { launch_tail_t(st, t); }


}  // namespace vol

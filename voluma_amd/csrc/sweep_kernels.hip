// Level-scheduled triangular sweeps: the DIC/DILU preconditioner (factor, forward and
// backward substitution, LinearSolver.cpp:56-105), the Gauss–Seidel solver sweeps
// (LinearSolver.cpp:238-264) and the Gauss–Seidel GAMG smoother (GAMG.cpp:363-384).
//
// The reference runs these loops one cell (or one face) at a time. A cell's update reads
// only neighbours with a smaller index (forward) or a larger one (backward), so cells
// are grouped into levels: the forward level of a cell is one more than the largest
// forward level among its lower-numbered neighbours, and no two cells of one level are
// neighbours. All cells of a level are updated at once, each with the same operands, in
// the same order, as in the sequential loop, so the results are bit-identical to it.
//
// Levels are short (10²–10⁴ cells on the meshes here) and there are hundreds of them,
// so the cost of a level is the latency of its dependent loads, plus the launch when it
// has one. The substitutions and Gauss–Seidel sweeps therefore read a level-ordered ELL
// copy of their rows (LevelEll): the cells of a level are contiguous, and entry k of
// the level's row t sits at base + k·size + t, so a row needs its cell index, its
// entries and the gathers of its neighbours' values, and no per-row offsets. Runs of
// consecutive levels of at most 1024 cells are swept by one workgroup of 1024 threads,
// a row per thread, in a single launch, with the run's part of x in LDS (every cell its
// rows write or read: loaded when the run starts, updated in place, stored to x as it is
// computed), a barrier on the LDS writes only between levels (a full barrier would also
// wait for the next levels' operand loads in flight), and the next level's operands
// loaded while the current level's LDS gathers are read. Larger levels get a grid launch
// each: a level of ~10⁴ rows moves ~1 MB, more than one CU streams in the time of a
// launch (measured: 16 rows per thread in one workgroup was 3× slower than grid
// launches on 2M-cell meshes). The factor, run once per matrix, walks the neighbour
// lists level by level (LevelSched).
#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

namespace {
constexpr int SWEEP_THREADS = 256;
constexpr int RUN_THREADS = 1024;
inline dim3 sweep_grid(int cnt) { return dim3((cnt + SWEEP_THREADS - 1) / SWEEP_THREADS); }
}  // namespace

// One level as a grid: thread i updates cell perm[i].
template <class Row>
__global__ __launch_bounds__(SWEEP_THREADS) void k_level(Row op, const int* perm, int lo, int hi, const int* done) {
    if (done && *done) return;
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < hi) op(perm[i]);
}

// Levels [L0, L1) by one workgroup, a barrier after each level.
template <class Row>
__global__ __launch_bounds__(RUN_THREADS) void k_level_run(Row op, const int* perm, const int* off, int L0, int L1,
                                                           const int* done) {
    if (done && *done) return;
    for (int L = L0; L < L1; ++L) {
        const int lo = off[L], hi = off[L + 1];
        for (int i = lo + int(threadIdx.x); i < hi; i += RUN_THREADS) op(perm[i]);
        __syncthreads();
    }
}

template <class Row>
static void launch_levels(hipStream_t st, const LevelSched& ls, int firstLevel, const Row& op, const int* done) {
    for (int g = 0; g < ls.nGrp; ++g) {
        const int L0 = std::max(ls.grpH[g], firstLevel), L1 = ls.grpH[g + 1];
        if (L0 >= L1) continue;
        if (ls.grpRunH[g]) {
            hipLaunchKernelGGL(k_level_run<Row>, dim3(1), dim3(RUN_THREADS), 0, st, op, ls.perm, ls.offD, L0, L1,
                               done);
        } else {
            for (int L = L0; L < L1; ++L) {
                const int lo = ls.offH[L], hi = ls.offH[L + 1];
                if (hi > lo)
                    hipLaunchKernelGGL(k_level<Row>, sweep_grid(hi - lo), dim3(SWEEP_THREADS), 0, st, op, ls.perm, lo,
                                       hi, done);
            }
        }
    }
}

// DIC/DILU factor (LinearSolver.cpp:67-74) of cell u: d_u = diag_u − Σ_{l<u, ascending
// (l, face)} (A_lu·A_ul) / d_l, rD_u = 1/d_u. The same thread then forms the coefficient
// products the substitutions use: rD_u·A_ul for its lower-list entries and rD_u times
// its own row's coefficient for its upper-list entries.
struct DiluFactorRow {
    SweepView s;
    const double *diag, *upper, *lower;
    double *dF, *rD, *pf, *pb;
    __device__ void operator()(int u) const {
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
};

void launch_dilu_factor(hipStream_t st, const SweepView& s, const LevelSched& fwd, const double* diag,
                        const double* upper, const double* lower, double* dF, double* rD, double* pf, double* pb) {
    launch_levels(st, fwd, 0, DiluFactorRow{s, diag, upper, lower, dF, rD, pf, pb}, nullptr);
}

// ---------------------------------------------------------------- level-ordered ELL sweeps
// Row operations (T: double, or float in the mixed-precision inner solve):
//   SW_GS   Gauss–Seidel (LinearSolver.cpp:247-262, GAMG.cpp:368-379):
//           x_c = (b_c − Σ coef·x_nb) / diag_c, entries in ascending face order
//   SW_FWD  DILU forward substitution (LinearSolver.cpp:93-98):
//           w_u = rD_u·r_u − Σ (rD_u·A_ul)·w_l, entries in ascending (l, face) order
//   SW_BWD  DILU backward substitution (LinearSolver.cpp:99-103):
//           w_l = w_l − Σ (rD_l·A_lu)·w_u, entries in descending (u, face) order
// What a row needs that the sweep does not change: its cell, its first PF entries
// (column and coefficient) and the operands of its start value (GS: b and diag; forward:
// rD·r; backward: its own w, which only its own level writes).
//
// Every load is issued whatever the row's width: a thread without a row, or an entry
// past its row, reads a valid stand-in (the level's first cell, entry 0, x[0]) whose
// value is then discarded. With no branches around them the loads of a phase go out
// together and the compiler waits for exactly those a result needs (vector loads return
// in order), not for everything in flight.
template <int PF, class T> struct RowPre {
    int c;     // the row's cell (a stand-in when !has)
    int s;     // runs: the LDS slot of the row's cell
    bool has;  // the thread has a row in this level
    T i0, i1;
    int col[PF];  // −1: no entry
    T cf[PF];
};

__device__ __forceinline__ int row_cell(const LevelEll& E, int pos0, int size, int t) {
    return E.cell[pos0 + (t < size ? t : 0)];
}
template <int PF, class T>
__device__ __forceinline__ void row_load_entries(const LevelEll& E, const SweepArgs<T>& a, int size, int width, int base,
                                                 int t, RowPre<PF, T>& p) {
    p.has = t < size;
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const bool in = p.has && k < width;
        const int e = in ? base + k * size + t : 0;
        const int j = E.col[e];
        const T v = a.coef[e];
        p.col[k] = in ? j : -1;
        p.cf[k] = v;
    }
}
template <int OP, int PF, class T> __device__ __forceinline__ void row_load_start(const SweepArgs<T>& a, RowPre<PF, T>& p) {
    if constexpr (OP == SW_GS) {
        p.i0 = a.b[p.c];
        p.i1 = a.diag[p.c];
    } else if constexpr (OP == SW_FWD) {
        p.i0 = a.rD[p.c] * a.r[p.c];
        p.i1 = T(1);
    } else {
        p.i0 = a.x[p.c];
        p.i1 = T(1);
    }
}
template <int PF, class T> __device__ __forceinline__ void row_gather(const SweepArgs<T>& a, const RowPre<PF, T>& p, T* xv) {
#pragma unroll
    for (int k = 0; k < PF; ++k) xv[k] = a.x[p.col[k] >= 0 ? p.col[k] : 0];
}
// acc = start − Σ coef·x over the row's entries in order (padding entries, −1, only
// trail a row); entries k ≥ PF in steps of SWEEP_TAIL with their loads issued together.
// Writes the row's new value.
constexpr int SWEEP_TAIL = 4;
template <int OP, int PF, class T>
__device__ __forceinline__ void row_finish(const LevelEll& E, const SweepArgs<T>& a, const RowPre<PF, T>& p, const T* xv,
                                           int size, int width, int base, int t) {
    T acc = p.i0;
#pragma unroll
    for (int k = 0; k < PF; ++k)
        if (p.col[k] >= 0) acc = acc - p.cf[k] * xv[k];
    for (int k0 = PF; k0 < width; k0 += SWEEP_TAIL) {
        int j[SWEEP_TAIL];
        T cf[SWEEP_TAIL], xt[SWEEP_TAIL];
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q) {
            const bool in = k0 + q < width;
            const int e = in ? base + (k0 + q) * size + t : 0;
            const int jj = E.col[e];
            cf[q] = a.coef[e];
            j[q] = in ? jj : -1;
        }
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q) xt[q] = a.x[j[q] >= 0 ? j[q] : 0];
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q)
            if (j[q] >= 0) acc = acc - cf[q] * xt[q];
    }
    if constexpr (OP == SW_GS) a.x[p.c] = acc / p.i1;
    else a.x[p.c] = acc;
}

// Runs: entries hold LDS slots of x (xs), so the gathers and the tail's reads come from
// LDS; the row writes its value to its slot and to x (read by later launches only).
template <int PF, class T> __device__ __forceinline__ void row_gather_lds(const T* xs, const RowPre<PF, T>& p, T* xv) {
#pragma unroll
    for (int k = 0; k < PF; ++k) xv[k] = xs[p.col[k] >= 0 ? p.col[k] : 0];
}
template <int OP, int PF, class T>
__device__ __forceinline__ void row_finish_lds(const LevelEll& E, const SweepArgs<T>& a, T* xs, const RowPre<PF, T>& p,
                                               const T* xv, int size, int width, int base, int t) {
    T acc = p.i0;
#pragma unroll
    for (int k = 0; k < PF; ++k)
        if (p.col[k] >= 0) acc = acc - p.cf[k] * xv[k];
    for (int k0 = PF; k0 < width; k0 += SWEEP_TAIL) {
        int j[SWEEP_TAIL];
        T cf[SWEEP_TAIL], xt[SWEEP_TAIL];
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q) {
            const bool in = k0 + q < width;
            const int e = in ? base + (k0 + q) * size + t : 0;
            const int jj = E.col[e];
            cf[q] = a.coef[e];
            j[q] = in ? jj : -1;
        }
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q) xt[q] = xs[j[q] >= 0 ? j[q] : 0];
#pragma unroll
        for (int q = 0; q < SWEEP_TAIL; ++q)
            if (j[q] >= 0) acc = acc - cf[q] * xt[q];
    }
    const T v = OP == SW_GS ? acc / p.i1 : acc;
    xs[p.s] = v;
    a.x[p.c] = v;
}
// Barrier for values exchanged through LDS only: waits for the LDS writes, not for the
// global loads in flight (the next levels' operands) or the stores to x.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One level as a grid: thread t updates the level's row t.
constexpr int GRID_PF = 8;
template <int OP, class T>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ell_level(LevelEll E, SweepArgs<T> a, int pos0, int size, int width,
                                                             int base, const int* done) {
    if (done && *done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= size) return;
    RowPre<GRID_PF, T> p;
    p.c = row_cell(E, pos0, size, t);
    row_load_entries(E, a, size, width, base, t, p);
    row_load_start<OP>(a, p);
    T xv[GRID_PF];
    row_gather(a, p, xv);
    row_finish<OP>(E, a, p, xv, size, width, base, t);
}

// Levels [L0, L1) by one workgroup of THREADS threads (every level has at most THREADS
// cells), a row per thread, x of the run's cells (cells[0 .. nCells), entries holding
// their slots) in LDS, a barrier on the LDS writes after each level. The levels' offsets,
// widths and bases are staged in LDS first. In level i a thread first reads its row's
// neighbour values from LDS, then loads level i+1's entries and start operands (from the
// cell it loaded in level i−1) and level i+2's cell and slot: none of them is written
// before its own level, and they stay in flight across the barrier. Sums keep the entry order: the result is the plain loop's.
template <int OP, int THREADS, int PF, class T>
__global__ __launch_bounds__(THREADS) void k_ell_run(LevelEll E, SweepArgs<T> a, int L0, int L1, const int* cells,
                                                     int nCells, const int* done) {
    if (done && *done) return;
    __shared__ int sOff[ELL_RUN_LEVELS + 1], sWidth[ELL_RUN_LEVELS], sBase[ELL_RUN_LEVELS];
    __shared__ T xs[RUN_LDS_CELLS];
    const int nl = L1 - L0, t = int(threadIdx.x);
    // staging in batches: every load of a batch in flight together, then the stores (a
    // plain loop waits for each element's load, or pair of dependent loads, in turn)
    constexpr int SB = 16;
    for (int i0 = t; i0 <= nl; i0 += SB * THREADS) {
        int o[SB], w[SB], b[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int i = i0 + u * THREADS;
            o[u] = i <= nl ? E.offD[L0 + i] : 0;
            w[u] = i < nl ? E.widthD[L0 + i] : 0;
            b[u] = i < nl ? E.baseD[L0 + i] : 0;
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int i = i0 + u * THREADS;
            if (i <= nl) sOff[i] = o[u];
            if (i < nl) {
                sWidth[i] = w[u];
                sBase[i] = b[u];
            }
        }
    }
    for (int i0 = t; i0 < nCells; i0 += SB * THREADS) {
        int c[SB];
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) c[u] = i0 + u * THREADS < nCells ? cells[i0 + u * THREADS] : -1;
#pragma unroll
        for (int u = 0; u < SB; ++u) v[u] = c[u] >= 0 ? a.x[c[u]] : T(0);
#pragma unroll
        for (int u = 0; u < SB; ++u)
            if (c[u] >= 0) xs[i0 + u * THREADS] = v[u];
    }
    __syncthreads();
    // past the last level, a thread loads the last level's data again (unused)
    auto lev = [nl](int i) { return i < nl ? i : nl - 1; };
    RowPre<PF, T> cur, nxt;
    int size = sOff[1] - sOff[0], width = sWidth[0], base = sBase[0];
    cur.c = row_cell(E, sOff[0], size, t);
    cur.s = E.slot[sOff[0] + (t < size ? t : 0)];
    row_load_entries(E, a, size, width, base, t, cur);
    row_load_start<OP>(a, cur);
    const int p1 = sOff[lev(1)], z1 = sOff[lev(1) + 1] - p1;
    int cNext = row_cell(E, p1, z1, t), sNext = E.slot[p1 + (t < z1 ? t : 0)];
    for (int i = 0; i < nl; ++i) {
        T xv[PF];
        row_gather_lds(xs, cur, xv);
        asm volatile("" ::: "memory");  // the gathers go out before the loads below
        const int n1 = lev(i + 1), n2 = lev(i + 2);
        const int nSize = sOff[n1 + 1] - sOff[n1], nWidth = sWidth[n1], nBase = sBase[n1];
        nxt.c = cNext;
        nxt.s = sNext;
        row_load_start<OP>(a, nxt);
        row_load_entries(E, a, nSize, nWidth, nBase, t, nxt);
        const int p2 = sOff[n2], z2 = sOff[n2 + 1] - p2;
        cNext = row_cell(E, p2, z2, t);
        sNext = E.slot[p2 + (t < z2 ? t : 0)];
        if (cur.has) row_finish_lds<OP>(E, a, xs, cur, xv, size, width, base, t);
        lds_barrier();
        cur = nxt;
        size = nSize;
        width = nWidth;
        base = nBase;
    }
}

template <int OP, class T>
static void launch_ell(hipStream_t st, const LevelEll& E, int firstLevel, const SweepArgs<T>& a, const int* done) {
    for (int g = 0; g < E.nGrp; ++g) {
        const int L0 = std::max(E.grpH[g], firstLevel), L1 = E.grpH[g + 1];
        if (L0 >= L1) continue;
        switch (E.grpKindH[g]) {
        case ELL_RUN256:
            hipLaunchKernelGGL((k_ell_run<OP, 256, 24, T>), dim3(1), dim3(256), 0, st, E, a, L0, L1,
                               E.runCells + E.runOffH[g], E.runCntH[g], done);
            break;
        case ELL_RUN512:
            hipLaunchKernelGGL((k_ell_run<OP, 512, 16, T>), dim3(1), dim3(512), 0, st, E, a, L0, L1,
                               E.runCells + E.runOffH[g], E.runCntH[g], done);
            break;
        case ELL_RUN1024:
            hipLaunchKernelGGL((k_ell_run<OP, 1024, 12, T>), dim3(1), dim3(1024), 0, st, E, a, L0, L1,
                               E.runCells + E.runOffH[g], E.runCntH[g], done);
            break;
        default:
            for (int L = L0; L < L1; ++L) {
                const int pos0 = E.offH[L], size = E.offH[L + 1] - pos0;
                if (size > 0)
                    hipLaunchKernelGGL((k_ell_level<OP, T>), sweep_grid(size), dim3(SWEEP_THREADS), 0, st, E, a, pos0,
                                       size, E.widthH[L], E.baseH[L], done);
            }
        }
    }
}

template <class T>
void launch_dilu_apply(hipStream_t st, const LevelEll& F, const LevelEll& B, const T* rD, const T* ef, const T* eb,
                       const T* r, T* w, const int* done) {
    SweepArgs<T> a;
    a.rD = rD;
    a.r = r;
    a.x = w;
    a.coef = ef;
    launch_ell<SW_FWD, T>(st, F, 0, a, done);
    a.coef = eb;
    // backward level 0 holds the cells without upper neighbours: nothing to subtract
    launch_ell<SW_BWD, T>(st, B, 1, a, done);
}
template void launch_dilu_apply<double>(hipStream_t, const LevelEll&, const LevelEll&, const double*, const double*,
                                        const double*, const double*, double*, const int*);
template void launch_dilu_apply<float>(hipStream_t, const LevelEll&, const LevelEll&, const float*, const float*,
                                       const float*, const float*, float*, const int*);

template <class T>
void launch_gs_sweep(hipStream_t st, const LevelEll& E, const T* coef, const T* diag, const T* b, T* x,
                     const int* done) {
    SweepArgs<T> a;
    a.coef = coef;
    a.diag = diag;
    a.b = b;
    a.x = x;
    launch_ell<SW_GS, T>(st, E, 0, a, done);
}
template void launch_gs_sweep<double>(hipStream_t, const LevelEll&, const double*, const double*, const double*,
                                      double*, const int*);
template void launch_gs_sweep<float>(hipStream_t, const LevelEll&, const float*, const float*, const float*, float*,
                                     const int*);

// ---------------------------------------------------------------- sync-free sweeps
// One launch per sweep. The work items are the wavefronts of E.waves (64 rows of one
// level each, in level order), and a wavefront takes the next item from a counter when
// it starts, not from its block index: the order in which workgroups start is not the
// order of their indices across the XCDs. Every value a row waits for then belongs to an
// item taken earlier, by a wavefront that has started and is resident or finished, and
// never to one of its own lanes, so the earliest unfinished item can always proceed. A row reads the new value of a neighbour on the sweep's side (GS forward and
// the DILU forward substitution: lower-numbered; backward: higher-numbered) from xNew,
// which starts filled with a sentinel NaN that arithmetic does not produce, and waits
// until it is no longer the sentinel; other neighbours' values come from xOld. Values
// are published by an atomic exchange (a read-modify-write is performed where all XCDs
// see it; a plain relaxed store may stay in the writer's L2) and read with device-scope
// atomic loads: the value is its own flag, so no fence is needed. A wait that runs past
// its limit sets *fault, and once it is set no wait spins any more, so every launch ends
// soon. Row arithmetic and order are row_finish's.
constexpr int FREE_GROUP = 8;                    // entries whose loads go out together
constexpr unsigned FREE_SPIN_LIMIT = 1u << 18;   // ≈ 10–20 ms of s_sleep(1) per value

template <class T> struct Sentinel;
template <> struct Sentinel<double> {
    static constexpr unsigned long long bits = 0x7FF4A5A5DEAD0001ull;  // signalling NaN
    __device__ static bool is(double v) { return (unsigned long long)__double_as_longlong(v) == bits; }
    __device__ static double value() { return __longlong_as_double((long long)bits); }
};
template <> struct Sentinel<float> {
    static constexpr unsigned bits = 0x7FA5DEADu;  // signalling NaN
    __device__ static bool is(float v) { return (unsigned)__float_as_uint(v) == bits; }
    __device__ static float value() { return __uint_as_float(bits); }
};

// xNew ← sentinel, and the sweep's work counter ← 0 (stream-ordered before the sweep).
template <class T>
__global__ void k_fill_sentinel(int n, T* x, int* counter, const int* done) {
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = Sentinel<T>::value();
    if (i == 0) *counter = 0;
}

template <class T>
__device__ __forceinline__ T free_wait(T* xNew, int j, int* fault) {
    T v = __hip_atomic_load(&xNew[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (Sentinel<T>::is(v)) {
        if ((spins & 1023) == 0 && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(&xNew[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (++spins > FREE_SPIN_LIMIT) {
            __hip_atomic_fetch_or(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
    return v;
}

template <int OP, bool BACK, class T>
__global__ __launch_bounds__(SWEEP_THREADS) void k_ell_free(LevelEll E, SweepArgs<T> a, const T* xOld, T* xNew,
                                                            int* counter, int* fault, const int* done) {
    if (done && *done) return;
    int w = 0;
    if ((threadIdx.x % WAVE) == 0) w = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w = __builtin_amdgcn_readfirstlane(w);  // lane 0's item for the whole wavefront
    if (w >= E.nWaves) return;              // whole wavefronts
    const int2 wl = E.waves[w];
    const int L = wl.x;
    const int pos0 = E.offD[L], size = E.offD[L + 1] - pos0, width = E.widthD[L], base = E.baseD[L];
    const int rb = E.runBaseD[L];  // ≥ 0: entries hold LDS slots of a run (cell = runCells[rb + slot])
    const int t = wl.y + int(threadIdx.x % WAVE);
    const bool has = t < size;
    const int c = E.cell[pos0 + (has ? t : 0)];
    T acc, dg = T(1);
    if constexpr (OP == SW_GS) {
        acc = a.b[c];
        dg = a.diag[c];
    } else if constexpr (OP == SW_FWD) {
        acc = a.rD[c] * a.r[c];
    } else {
        acc = xOld[c];
    }
    for (int k0 = 0; k0 < width; k0 += FREE_GROUP) {
        int j[FREE_GROUP];
        T cf[FREE_GROUP], v[FREE_GROUP];
#pragma unroll
        for (int u = 0; u < FREE_GROUP; ++u) {
            const bool in = has && k0 + u < width;
            const int e = in ? base + (k0 + u) * size + t : 0;
            const int jj = E.col[e];
            cf[u] = a.coef[e];
            j[u] = in ? jj : -1;
        }
        if (rb >= 0) {  // LDS slots → cells
#pragma unroll
            for (int u = 0; u < FREE_GROUP; ++u)
                if (j[u] >= 0) j[u] = E.runCells[rb + j[u]];
        }
        // first reads of the group together; then wait on the ones not published yet
#pragma unroll
        for (int u = 0; u < FREE_GROUP; ++u) {
            const bool fresh = j[u] >= 0 && (BACK ? j[u] > c : j[u] < c);
            v[u] = j[u] < 0 ? T(0)
                   : fresh ? __hip_atomic_load(&xNew[j[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : xOld[j[u]];
        }
#pragma unroll
        for (int u = 0; u < FREE_GROUP; ++u) {
            const bool fresh = j[u] >= 0 && (BACK ? j[u] > c : j[u] < c);
            if (fresh && Sentinel<T>::is(v[u])) v[u] = free_wait(xNew, j[u], fault);
        }
#pragma unroll
        for (int u = 0; u < FREE_GROUP; ++u)
            if (j[u] >= 0) acc = acc - cf[u] * v[u];
    }
    if (has) {
        const T res = OP == SW_GS ? acc / dg : acc;
        (void)__hip_atomic_exchange(&xNew[c], res, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// fault points at two ints: [0] the fault flag, [1] the work counter.
template <int OP, bool BACK, class T>
static void launch_free(hipStream_t st, const LevelEll& E, int n, const SweepArgs<T>& a, const T* xOld, T* xNew,
                        int* fault, const int* done) {
    if (n <= 0 || E.nWaves == 0) return;
    int* counter = fault + 1;
    hipLaunchKernelGGL(k_fill_sentinel<T>, sweep_grid(n), dim3(SWEEP_THREADS), 0, st, n, xNew, counter, done);
    const long threads = long(E.nWaves) * WAVE;
    hipLaunchKernelGGL((k_ell_free<OP, BACK, T>), dim3(unsigned((threads + SWEEP_THREADS - 1) / SWEEP_THREADS)),
                       dim3(SWEEP_THREADS), 0, st, E, a, xOld, xNew, counter, fault, done);
}

template <class T>
void launch_gs_sweep_free(hipStream_t st, const LevelEll& E, bool backward, const T* coef, const T* diag, const T* b,
                          const T* xOld, T* xNew, int* fault, const int* done) {
    SweepArgs<T> a;
    a.coef = coef;
    a.diag = diag;
    a.b = b;
    const int n = E.offH[E.nLev];
    if (backward)
        launch_free<SW_GS, true, T>(st, E, n, a, xOld, xNew, fault, done);
    else
        launch_free<SW_GS, false, T>(st, E, n, a, xOld, xNew, fault, done);
}
template void launch_gs_sweep_free<double>(hipStream_t, const LevelEll&, bool, const double*, const double*,
                                           const double*, const double*, double*, int*, const int*);
template void launch_gs_sweep_free<float>(hipStream_t, const LevelEll&, bool, const float*, const float*,
                                          const float*, const float*, float*, int*, const int*);

template <class T>
void launch_dilu_apply_free(hipStream_t st, const LevelEll& F, const LevelEll& B, const T* rD, const T* ef,
                            const T* eb, const T* r, T* tmp, T* w, int* fault, const int* done) {
    SweepArgs<T> a;
    a.rD = rD;
    a.r = r;
    a.coef = ef;
    const int n = F.offH[F.nLev];
    launch_free<SW_FWD, false, T>(st, F, n, a, nullptr, tmp, fault, done);  // tmp: forward result
    a.coef = eb;
    launch_free<SW_BWD, true, T>(st, B, n, a, tmp, w, fault, done);  // every cell, level 0 included
}
template void launch_dilu_apply_free<double>(hipStream_t, const LevelEll&, const LevelEll&, const double*,
                                             const double*, const double*, const double*, double*, double*, int*,
                                             const int*);
template void launch_dilu_apply_free<float>(hipStream_t, const LevelEll&, const LevelEll&, const float*, const float*,
                                            const float*, const float*, float*, float*, int*, const int*);

// Per-entry coefficients of a LevelEll from their source array: out[e] = from[src[e]]
// (0 for padding, src < 0).
template <class T>
__global__ void k_gather_entries(size_t n, const int* src, const T* from, T* out) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) {
        const int s = src[e];
        out[e] = s >= 0 ? from[s] : T(0);
    }
}
template <class T> void launch_gather_entries(hipStream_t st, size_t n, const int* src, const T* from, T* out) {
    if (n > 0)
        hipLaunchKernelGGL(k_gather_entries<T>, dim3(unsigned((n + SWEEP_THREADS - 1) / SWEEP_THREADS)),
                           dim3(SWEEP_THREADS), 0, st, n, src, from, out);
}
template void launch_gather_entries<double>(hipStream_t, size_t, const int*, const double*, double*);
template void launch_gather_entries<float>(hipStream_t, size_t, const int*, const float*, float*);

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
    if (small) {  // 16 loads in flight per lane, then their stores
        constexpr int SB = 16;
        for (int i0 = threadIdx.x; i0 < n * n; i0 += SB * 64) {
            double v[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) v[u] = i0 + u * 64 < n * n ? lu[i0 + u * 64] : 0.0;
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (i0 + u * 64 < n * n) f[i0 + u * 64] = v[u];
        }
    }
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

template <class T>
void launch_coarse_solve_rows(hipStream_t st, int n, const double* lu, const T* r, T* x, const int* done) {
    const size_t lds = sizeof(double) * (size_t(n) + (n <= 64 ? size_t(n) * n : 0));
    hipLaunchKernelGGL(k_coarse_solve_rows<T>, dim3(1), dim3(64), lds, st, n, lu, r, x, done);
}
template void launch_coarse_solve_rows<double>(hipStream_t, int, const double*, const double*, double*, const int*);
template void launch_coarse_solve_rows<float>(hipStream_t, int, const double*, const float*, float*, const int*);

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

// Internal interfaces between the host-side solver (api.cpp) and the HIP kernels
// (solver_kernels.hip). Nothing here is part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vol {

constexpr int WAVE = 64;   // CDNA wavefront
constexpr int NBINS = 72;  // exact-sum bins (par::ExactSum::NBINS, Comm.h:44)

// ------------------------------------------------------------------ SELL-64 layout
// Off-diagonal rows of an LDU matrix in sliced-ELLPACK form with slice height 64
// (one wavefront): rows [64s, 64s+64) occupy slots sliceOff[s] .. sliceOff[s+1),
// slot k of a row in slice s at sliceOff[s] + 64k + lane. Slots of a row keep
// ascending face order, which is the per-row accumulation order of the reference's
// face loop (src/fvm/linalg/LinearSolver.cpp:121-133). Padding slots have col = −1.
//
// Coarse GAMG levels, whose rows are long, use a CSR layout instead (csrG > 0): row r
// occupies slots sliceOff[r] .. sliceOff[r+1) contiguously and is processed by csrG
// lanes together (csrIt slots per lane).
struct SellView {
    int n = 0;          // rows
    int nslices = 0;
    int nslots = 0;     // padded slot count
    int maxWidth = 0;   // longest row (slots)
    int csrG = 0;       // > 0: CSR layout, lanes per row
    int csrIt = 0;      // slots per lane (csrG·csrIt ≥ maxWidth)
    const int* sliceOff = nullptr;
    const int* col = nullptr;  // neighbour row, −1 = padding
    const int* src = nullptr;  // (face << 1) | side: side 0 → upper[face], 1 → lower[face]; −1 padding
    // SELL levels whose column offsets (neighbour − row) all fit int16 also carry them
    // as int16 (COL16_PAD = padding); the gather kernels then read 2 bytes per slot
    // instead of 4. Null otherwise.
    const int16_t* col16 = nullptr;
    // ELL form (ellWidth > 0): every slice has ellWidth slots, so slot k of row r sits at
    // (r/64)·64·ellWidth + 64k + r%64 and no slice offset has to be loaded first.
    int ellWidth = 0;
    // Packed rows (level 0 in ELL form, ellWidth ≤ 8, every neighbour within ±32767 rows;
    // colP != null): row r's column offsets (neighbour − r, COL16_PAD = padding) as int16
    // in colPW words at colP + r·colPW (2 offsets per word, colPW = 2 or 4), and the
    // coefficient array holds, after its nslots slot-major values, nslotsP more: the same
    // values in pairs, pair p of row r at (r/64)·64·P + 64p + r%64 counted in pairs
    // (P = ⌈ellWidth/2⌉). A gather then reads one vector load of offsets and P pair loads
    // per row instead of 2·ellWidth scalar loads, and 2 bytes per column instead of 4.
    const int* colP = nullptr;
    int colPW = 0;
    int nslotsP = 0;
    // Sorted SELL rows (perm != null): position p of the slices (slice p/64, lane p%64)
    // holds row perm[p], and row r sits at position iperm[r]. Rows are sorted by length
    // within windows of consecutive rows so that the rows sharing a slice have nearly
    // equal lengths (less padding); each row keeps its slot order, so every row's result
    // is unchanged. The vectors stay in row order.
    const int* perm = nullptr;
    const int* iperm = nullptr;
    // Packed SELL rows (a coarse level in SELL form, or a batch level, whose column offsets
    // all fit int16; colS != null): slice s has sliceP[s] = (word base, pair base, width w, 0); the column
    // offsets (neighbour − row, COL16_PAD = padding) of lane l, four per int2 word, word q
    // at colS[word base + 64q + l]; and the coefficient array holds, after its nslots
    // slot-major values, nslotsP more: the same values in pairs, pair p of lane l at pair
    // index pair base + 64p + l (⌈w/2⌉ pairs per row). The slot-major arrays stay.
    const int2* colS = nullptr;
    const int4* sliceP = nullptr;
};
constexpr int COL16_PAD = -32768;

// ------------------------------------------------------------------ device scalars
// Krylov scalars and loop control of one solve, kept on the device so that the
// iteration needs no host round trip. Kernels after a finalize read them.
struct DevScalars {
    double normFactor, initialResidual, finalResidual, residualSum, xRef;
    double wArA, wArAold, wApA, alpha, beta;           // PCG
    double rho, rhoOld, alphaB, omegaB, tAtA, tAsA;    // PBiCGStab
    double tolerance, relTol;
    int maxIter, minIter, nSweeps, n;
    int iterations;
    double innerRel;  // mixed precision: residual reduction asked of each inner solve
    int done;         // loop finished (converged, maxIter or breakdown): later kernels skip
    int exitPending;  // PBiCGStab exit on s: x += αy still to apply
    int status;
    int outer;        // mixed precision: defect-correction steps taken
    int itStart;      // mixed, inner scalars: iteration count when the current segment began
    int keepKrylov;   // mixed, outer scalars: inner PCG keeps its search direction across steps
};
// Mixed precision keeps two DevScalars side by side: [0] the FP64 outer loop (the
// reported performance), [1] the current FP32 inner Krylov solve.


enum FinOp : int {
    FIN_XREF = 0,      // xRef = Σx / n
    FIN_INIT = 1,      // normFactor, residualSum, initial residual, early exit
    FIN_PCG_WRA = 2,
    FIN_PCG_WAPA = 3,
    FIN_PCG_NORM = 4,
    FIN_BI_RHO = 5,
    FIN_BI_ALPHA = 6,
    FIN_BI_S = 7,
    FIN_BI_OMEGA = 8,
    FIN_BI_NORM = 9,
    FIN_JAC_NORM = 10,
    FIN_STORE = 11,    // only store the reduced values (tests/diagnostics)
    FIN_OUTER = 12,    // mixed: outer residual Σ|b − Ax| after x += e; convergence of the whole solve
    FIN_INNER_START = 13,  // mixed: set up the next inner solve in S[1] from S[0]
};

// Reduction scratch: fast mode uses `nblk` double partials per reduced quantity;
// exact mode uses `nblk` × 72 int64 bins + a non-finite flag per block.
struct RedScratch {
    double* part = nullptr;     // [maxRed][maxBlk]
    long long* xpart = nullptr; // [maxRed][maxBlkExact][72]
    int* xnf = nullptr;         // [maxRed][maxBlkExact]
    double* result = nullptr;   // [maxRed] (FIN_STORE)
    int* arrive = nullptr;      // ticket counter of the last-block finalize (nullable; 0 between launches)
    int maxBlk = 0, maxBlkExact = 0;
};

constexpr int RED_BLOCKS = 1024;       // grid of fast reduction kernels
constexpr int RED_BLOCKS_EXACT = 512;  // grid of exact reduction kernels
constexpr int RED_THREADS = 256;
constexpr int MAX_RED = 2;

// ------------------------------------------------------------------ SELL gather kernels
enum SellOp : int {
    OP_AMUL = 0,    // out = diag∘x + Σ a·x_nb                       (LinearSolver.cpp:121-133)
    OP_RESID = 1,   // out = r − diag∘x − Σ a·x_nb, in that order     (GAMG.cpp:387-395)
    OP_JACOBI = 2,  // out = x + ω((r − Σ a·x_nb)/d − x)             (project smoother)
    OP_JACREF = 3,  // out = (r − Σ a·x_nb)/d                        (LinearSolver.cpp:247-267)
};
enum XLoad : int {
    XL_PLAIN = 0,        // x[j]
    XL_UNIFORM = 1,      // a scalar
    XL_FROMZERO = 2,     // ω·(r[j]·rD[j]), rD = 1/d: the first Jacobi sweep from x = 0
    XL_PROLONG = 3,      // x[j] + xc[agg[j]]
    XL_ZEROPROLONG = 4,  // ω·(r[j]·rD[j]) + xc[agg[j]]
};

template <class T> struct SellArgsT {
    SellView S;
    const T* coef = nullptr;
    const T* diag = nullptr;
    const T* rD = nullptr;          // 1/diag (XL_FROMZERO / XL_ZEROPROLONG)
    const T* x = nullptr;
    const T* r = nullptr;
    const T* xc = nullptr;
    const int* agg = nullptr;
    const T* xuPtr = nullptr;       // XL_UNIFORM value (device scalar)
    double omega = 1.0;
    T* out = nullptr;
    const T* dotv = nullptr;        // fused dot: Σ out·dotv (in double) into part (fast mode only)
    double* part = nullptr;
    const int* done = nullptr;
    // batches: rows of systems whose scalars say done are skipped (SELL: per slice the
    // first and last system of its rows; CSR: the system of each row)
    const int2* sliceSys = nullptr;
    const int* rowSys = nullptr;
    const DevScalars* sysS = nullptr;
    // prolongation pushed down (OP_JACOBI, nullable): row I's result is also added to
    // pushZ[m] for the finer cells m of aggregate I (pushList[pushOff[I] .. pushOff[I+1])),
    // which is the value x[m] + xc[agg[m]] the finer level's sweep would gather
    T* pushZ = nullptr;
    const int* pushOff = nullptr;
    const int* pushList = nullptr;
    // host side: when set, the launch records these events at the kernel's start and end
    // (hipExtLaunchKernelGGL), which times the kernel itself
    hipEvent_t evStart = nullptr, evStop = nullptr;
    // fused dot only (nullable): every block takes a ticket from arrive after storing its
    // partial; the last one sums the partials in the finalize kernel's order and runs the
    // scalar step finOp on finS, so no separate finalize launch follows
    int* arrive = nullptr;
    DevScalars* finS = nullptr;
    int finOp = 0;
};
using SellArgs = SellArgsT<double>;
using SellArgs32 = SellArgsT<float>;

void launch_sell(hipStream_t st, int op, int xl, const SellArgs& a);
void launch_sell32(hipStream_t st, int op, int xl, const SellArgs32& a);
int sell_grid_blocks(int nslices);  // blocks used by launch_sell (== partial count when fused)

void launch_sell_coef(hipStream_t st, const SellView& S, const double* upper, const double* lower, double* coef,
                      float* coef32);

// ------------------------------------------------------------------ vector kernels
// Reductions: results reduced per block into scratch, then a 1-block finalize.
// `exact` selects the ExactSum accumulator.
enum VecOp : int {
    V_SUM_X = 0,          // red0 = Σ a
    V_INIT = 1,           // rA = b − wA; red0 = Σ(|wA − pA| + |b − pA|), red1 = Σ|rA|; y (nullable) = ω·(rA·rD)
    V_PRECOND_DOT = 2,    // w = rD∘r (rD == null: w = r); red0 = Σ w·r
    V_DOT = 3,            // red0 = Σ a·b
    V_PCG_P = 4,          // p = (it == 0) ? w : w + β p
    V_PCG_XR = 5,         // x += α p; r −= α q; red0 = Σ|r|; w (nullable) = ω·(r·rD)
    V_BI_P = 6,           // p = (it == 0) ? r : r + β(p − ω Ay)
    V_SCALE = 7,          // w = rD∘r (or copy)
    V_BI_S = 8,           // s = r − α Ay; red0 = Σ|s|
    V_BI_XEXIT = 9,       // if exitPending: x += α y
    V_BI_DOT2 = 10,       // red0 = Σ t·t, red1 = Σ t·s
    V_BI_XR = 11,         // x += α y + ω z; r = s − ω t; red0 = Σ|r|
    V_BMINUSAX_NORM = 12, // r = b − wA; red0 = Σ|r|
    V_COPY = 13,          // a → w
    V_RECIP = 14,         // w = 1/a
    V_FROMZERO = 15,      // w = ω·(r/d)
    V_FILL_XREF = 16,     // w = xRef of the row's system (batches)
    V_ZSCALE = 17,        // w = ω·(a·rD) in the vector type (the first Jacobi sweep from zero)
};

// Vector element type T (double, or float for the inner solve of mixed precision);
// element arithmetic and all reductions run in double either way.
template <class T> struct VecArgsT {
    int n = 0;
    const T *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *e = nullptr;
    T *w = nullptr, *x = nullptr, *y = nullptr;
    const T* rD = nullptr;
    double omega = 1.0;
    DevScalars* S = nullptr;
    const int* done = nullptr;
};
using VecArgs = VecArgsT<double>;
using VecArgs32 = VecArgsT<float>;

void launch_vec(hipStream_t st, int op, const VecArgs& a, bool exact, const RedScratch& R);
void launch_vec32(hipStream_t st, int op, const VecArgs32& a, const RedScratch& R);  // fast sums only

// Precision changes of mixed precision (predicated on *done when done != nullptr).
enum MixOp : int {
    MIX_TO32 = 0,  // w = float(a); w2 = w (nullable); z = 0 (nullable)
    MIX_ADD = 1,   // x += double(e)
};
void launch_mix(hipStream_t st, int op, int n, const double* a, float* w, float* w2, float* z, double* x,
                const float* e, const int* done);
// hostOut (nullable, host-visible memory): receives a copy of *S after the step
void launch_finalize(hipStream_t st, int op, int nred, bool exact, int nparts, const RedScratch& R, DevScalars* S,
                     DevScalars* hostOut = nullptr);
int vec_has_reduction(int op);
// blocks of a fast-sum vector op (= partials of its reduction; maxBlk: the scratch's)
int vec_fast_blocks(int op, int n, int maxBlk);

void launch_clear_exit(hipStream_t st, DevScalars* S);
void launch_init_scalars(hipStream_t st, DevScalars* S, double tol, double relTol, int maxIter, int minIter,
                         int nSweeps, int n, double innerRel, int keepKrylov = 0);

// ------------------------------------------------------------------ GAMG kernels
// rc[I] = Σ res over the fine cells of I (ascending); zc (nullable): ω·(rc·rDc) as well
void launch_restrict(hipStream_t st, int nc, const int* mOff, const int* mList, const double* res, double* rc,
                     const int* done, double* zc = nullptr, const double* rDc = nullptr, double omega = 1.0);
void launch_restrict32(hipStream_t st, int nc, const int* mOff, const int* mList, const float* res, float* rc,
                       const int* done, float* zc = nullptr, const float* rDc = nullptr, double omega = 1.0);
void launch_galerkin(hipStream_t st, int nc, int ncf, const int* mOff, const int* mList, const int* iOff,
                     const int* iList, const int* cfOff, const int* cfList, const double* Fd, const double* Fu,
                     const double* Fl, double* Cd, double* Cu, double* Cl);
void launch_orient0(hipStream_t st, int nf, const char* ownerIsL, const double* upper, const double* lower,
                    double* Fu, double* Fl);
// Dense LU of the coarsest level, one workgroup. lu: n×n row-major (in/out). status[0] = −1 if singular.
void launch_coarse_assemble(hipStream_t st, int n, int nf, const int* l, const int* u, const double* d,
                            const double* cu, const double* cl, double* lu);
void launch_coarse_lu(hipStream_t st, int n, double* lu, double* luT, int* status);
void launch_coarse_solve(hipStream_t st, int n, const double* lu, const double* luT, const double* r, double* x,
                         const int* done);
// Restriction into a coarsest level of nc ≤ 64 cells fused with its substitution.
template <class T>
void launch_restrict_coarse64(hipStream_t st, int nc, const int* mOff, const int* mList, const T* res,
                              const double* luT, T* x, const int* done);
// Residual and restriction of a level in one launch: rc[I] = Σ over the fine cells i of
// aggregate I, ascending, of (r − A·x)_i, each row's residual computed in slot order as
// the residual gather does (x: `a.x` under XL_PLAIN, ω·(r·rD) under XL_FROMZERO); zc
// (nullable) as launch_restrict. No residual vector is written. CSR levels only.
// blk: per block (first aggregate, first member position) of its range (nBlocks + 1
// entries; a range holds ≤ RR_POS member rows unless it is one aggregate, and ≤ RR_POS
// aggregates); rows: per member position (row, its first slot, its slot count), in the
// order of the restriction lists. G lanes per row, G·IT ≥ the longest row.
constexpr int RR_POS = 256;
constexpr int RR_COARSE_ROWS = 1024;  // rows of a level whose restriction goes into the coarsest
template <class T> struct RestrictArgsT {
    int nc = 0;
    const int2* blk = nullptr;
    int nBlocks = 0;
    const int4* rows = nullptr;
    const int* mOff = nullptr;
    const int* mList = nullptr;  // packed level-0 rows: the member rows by position
    T* rc = nullptr;
    T* zc = nullptr;
    const T* rDc = nullptr;
    double omega = 1.0;
    const double* luT = nullptr;  // coarsest form: LU factors (transposed), rc is then x
    // coarsest form, XL_PLAIN: also the level's one damped-Jacobi post-sweep into a.out from
    // x = a.x + xc[a.agg] (xc: the coarsest solution); needs level rows ≤ RR_COARSE_THREADS
    // / rrCoarseLanes(longest row)
    int post = 0;
};
constexpr int RR_COARSE_THREADS = 512;
// lanes per row of the coarsest form
inline int rrCoarseLanes(int maxWidth) { return maxWidth <= 32 ? 4 : 8; }
// (a.S.colP set: the packed level-0 rows, one lane per member row through R.mList; a.out
// must be null)
template <class T> void launch_resid_restrict(hipStream_t st, int xl, int G, int IT, const SellArgsT<T>& a,
                                              const RestrictArgsT<T>& R);
// The same into a coarsest level of nc ≤ 64 cells, followed in the same launch by its
// substitution (as launch_restrict_coarse64); one workgroup, level rows ≤ RR_COARSE_ROWS,
// each of ≤ 64 slots.
template <class T> void launch_resid_restrict_coarse64(hipStream_t st, int xl, int G, int IT, const SellArgsT<T>& a,
                                                       const RestrictArgsT<T>& R);
// The last two levels above the coarsest (A = coarsest − 2, B = coarsest − 1, both CSR,
// ≤ LAST2_ROWS rows of ≤ LAST2_BATCHES·2·LAST2_IT slots) and the coarsest C (≤ 64 cells) of a 1/1
// damped-Jacobi V-cycle with z kept, in one workgroup: A's residual, restriction into B
// (with z_B), B's residual, restriction into C, C's substitution, B's post-sweep, A's
// post-sweep, every row's arithmetic and every sum's order as in the separate launches.
// A's result goes to xA (nullable) and, with pushZ, is added to the finer level's z.
constexpr int LAST2_THREADS = 1024, LAST2_ROWS = 512, LAST2_IT = 8, LAST2_BATCHES = 2;
template <class T> struct Last2ArgsT {
    int nA = 0, nB = 0, nC = 0;
    int batchesA = 1, batchesB = 1;  // ⌈longest row / (2·LAST2_IT)⌉ of A and B
    const int *offA = nullptr, *colA = nullptr, *aggA = nullptr, *mOffA = nullptr, *mListA = nullptr;
    const T *coefA = nullptr, *diagA = nullptr, *rA = nullptr, *zA = nullptr;
    const int *offB = nullptr, *colB = nullptr, *aggB = nullptr, *mOffB = nullptr, *mListB = nullptr;
    const T *coefB = nullptr, *diagB = nullptr, *rDB = nullptr;
    const double* luT = nullptr;
    double omega = 1.0;
    T* xA = nullptr;
    T* pushZ = nullptr;
    const int *pushOff = nullptr, *pushList = nullptr;
    const int* done = nullptr;
};
template <class T> void launch_vcycle_last2(hipStream_t st, const Last2ArgsT<T>& a);
// Same substitution in FP64 on FP32 vectors (r widened, x rounded).
void launch_coarse_solve32(hipStream_t st, int n, const double* lu, const double* luT, const float* r, float* x,
                           const int* done);

// Coarse levels of the V-cycle run inside one workgroup (k_vcycle_tail): the levels
// from `first` down to the coarsest, with block barriers between the phases.
constexpr int TAIL_MAX_CELLS = 4096;  // largest first level the tail kernel takes (VOLUMA_TAIL_MAX_CELLS)
constexpr int TAIL_THREADS = 1024;
// The tail's vectors (r, two work buffers, x per level) live in LDS at the given
// offsets (in doubles); only matrix data is read from global memory.
constexpr int TAIL_LDS_BYTES = 160 * 1024;  // gfx950 LDS per workgroup
template <class T> struct TailLevelT {
    SellView S;
    const T* coef = nullptr;
    const T* diag = nullptr;
    const T* rD = nullptr;
    const int* agg = nullptr; // → next level
    const int* mOff = nullptr;
    const int* mList = nullptr;
    int nNext = 0;
    int ldsR = 0, ldsA = 0, ldsB = 0, ldsX = 0;  // offsets in elements of T
    // matrix data copied into LDS at kernel start when it fits (ldsMat = 1): byte
    // offsets of row offsets, columns, coefficients, diagonal, 1/diagonal, agg and the
    // restriction lists
    int ldsMat = 0;
    int nRowOff = 0;
    int ldsRowOff = 0, ldsCol = 0, ldsCoef = 0, ldsDiag = 0, ldsRD = 0, ldsAgg = 0, ldsMOff = 0, ldsMList = 0;
};
// A run of 4-byte words that the tail kernel copies from global memory into its LDS at
// start (dst: word offset in LDS). Every wave loads a whole chunk with all its loads in
// flight, and the waves take different chunks, so the preload costs a few memory round
// trips instead of one per array.
constexpr int TAIL_CHUNK_WORDS = 1024;
struct TailChunk {
    const unsigned* src = nullptr;
    int dst = 0;
    int n = 0;  // ≤ TAIL_CHUNK_WORDS
};
template <class T> struct TailArgsT {
    int nLevels = 0;                  // levels in the tail, the last is the coarsest
    const TailLevelT<T>* L = nullptr; // device array
    const T* rTop = nullptr;          // right-hand side of the first tail level (global)
    T* xTop = nullptr;                // its V-cycle result (global)
    int nCoarse = 0;
    const double* lu = nullptr;       // coarsest LU factors, always FP64
    const double* luT = nullptr;
    int ldsBytes = 0;                 // LDS: [coarsest LU / substitution vector (doubles)] [level vectors (T)] [matrices]
    int nPre = 2, nPost = 2;
    double omega = 0.8;
    const int* done = nullptr;
    const TailChunk* chunks = nullptr;  // device array: matrices, coarsest factors (n ≤ 64), rTop
    int nChunks = 0;
    int rInChunks = 0;                  // rTop is among the chunks (its address is fixed)
};
using TailLevel = TailLevelT<double>;
using TailArgs = TailArgsT<double>;
using TailLevel32 = TailLevelT<float>;
using TailArgs32 = TailArgsT<float>;
void launch_vcycle_tail(hipStream_t st, const TailArgs& t);
void launch_vcycle_tail32(hipStream_t st, const TailArgs32& t);
// buffer (0 = bufA, 1 = bufB) holding the pre-smoothed x after nPre sweeps
__host__ __device__ inline int pre_smoothed_buffer(int nPre) { return nPre >= 2 ? ((nPre - 2) & 1) : 0; }

// ------------------------------------------------------------------ assembly
enum AsmKind : int { ASM_LAPLACIAN = 0, ASM_DIV = 1 };
struct AsmArgs {
    int n = 0, nf = 0;                  // cells, internal faces
    SellView S;                         // level-0 slots of the mesh (src = face << 1 | side)
    const int* bOff = nullptr;          // cell → boundary faces (local index bf), ascending
    const int* bList = nullptr;
    const double* gamma = nullptr;      // per face (internal then boundary), or uniform gammaU
    double gammaU = 1.0;
    const double* magSf = nullptr;      // per face
    const double* ndc = nullptr;        // per face
    const double* flux = nullptr;       // per face (div)
    const int* bcFixed = nullptr;       // per boundary face: fixedValue (1) / zeroGradient (0)
    const double* bcValue = nullptr;    // per boundary face (nullable → 0)
    double* diag = nullptr;             // out: folded diagonal
    double* upper = nullptr;            // out
    double* lower = nullptr;            // out (div)
    double* icOut = nullptr;            // out (nullable): internalCoeffs
    double* bsOut = nullptr;            // out (nullable): boundarySource
    const double* bIn = nullptr;        // in (nullable): source before the fold
    double* bOut = nullptr;             // out (nullable): folded source
};
void launch_assemble(hipStream_t st, int kind, const AsmArgs& a);

// ------------------------------------------------------------------ FvMatrix algebra
// One term (fvm::ddt Euler, fvm::div upwind, fvm::laplacian) is built per element in
// registers exactly as the reference builds it in a fresh FvMatrix, then added to the
// matrix parts with sign ±1 (FvMatrix::operator+= / -=, FvMatrix.h:57-80,232-239).
enum FvmKind : int { FVM_DDT = 0, FVM_DIV = 1, FVM_LAP = 2 };
struct FvmArgs {
    int n = 0, nf = 0, nb = 0;          // cells, internal faces, boundary faces
    SellView S;                         // level-0 slots (src = face << 1 | side)
    const int* bOff = nullptr;          // cell → boundary faces, ascending
    const int* bList = nullptr;
    double sign = 1.0;
    // term inputs (face arrays: internal then boundary faces)
    double ddtC = 0.0, ddtC0 = 0.0, ddtC00 = 0.0;  // ddt: ddtCoeffs c, c0, c00 (Schemes.h:68-86)
    const double* V = nullptr;          // ddt, non-orthogonal gradient: cell volumes
    const double* psi0 = nullptr;       // ddt: old-time field
    const double* psi00 = nullptr;      // ddt: old-old-time field (psi0 when there is none)
    const double* flux = nullptr;       // div
    double gammaU = 1.0;                // laplacian: uniform γ unless gamma != nullptr
    const double* gamma = nullptr;
    const double* magSf = nullptr;
    const double* ndc = nullptr;
    const int* bcFixed = nullptr;       // per boundary face: fixedValue 1 / zeroGradient 0
    const double* bcValue = nullptr;    // per boundary face (nullable → 0)
    // laplacian non-orthogonal correction (psi != nullptr)
    const double* psi = nullptr;
    const double* w = nullptr;          // per face
    const double* Sf = nullptr;         // 3 per face
    const double* corr = nullptr;       // 3 per face
    int* anyCorr = nullptr;             // device flag: some internal |corr|² > 1e-20
    double* grad = nullptr;             // 3 per cell (scratch)
    int emptyDir = -1;                  // 2-D: the gradient's empty component is zeroed (Fvc.h:169-181)
    double* fc = nullptr;               // per internal face (scratch)
    // FvMatrix parts (in/out)
    double* diag = nullptr;
    double* upper = nullptr;
    double* lower = nullptr;            // non-null once the matrix has a lower part
    int hadLower = 0;                   // lower held values before this term
    double* source = nullptr;
    double* ic = nullptr;               // internalCoeffs
    double* bs = nullptr;               // boundarySource
    double* ffc = nullptr;              // face-flux correction (per face; zeros = none yet)
};
// own/nei: device owner/neighbour of the internal faces (non-orthogonal correction)
void launch_fvm_term(hipStream_t st, int kind, const FvmArgs& a, const int* own, const int* nei);
// FvMatrix −= su (FvMatrix.h:93-97): source[c] += su[c]·V[c]
void launch_fvm_source(hipStream_t st, int n, const double* su, const double* V, double* source);
// ------------------------------------------------------------------ mesh geometry
// Inputs: points (interleaved xyz), faces as offsets into facePts (int32), owner of every
// face, neighbour of the internal faces, and the cell → face lists (faces ascending).
// Outputs per cell: V, C; per face: Sf, Cf, d, magSf, w, ndc, corr, skew (nullable);
// maxSkewBits (nullable): atomic max of the bits of |skew|/|d| over internal faces.
struct GeoArgs {
    int nCells = 0, nFaces = 0, nInternal = 0;
    const double* points = nullptr;
    const int* faceOff = nullptr;
    const int* facePts = nullptr;
    const int* owner = nullptr;
    const int* neighbour = nullptr;
    const int* cfOff = nullptr;
    const int* cfList = nullptr;
    double *V = nullptr, *C = nullptr, *Sf = nullptr, *Cf = nullptr, *d = nullptr, *magSf = nullptr, *w = nullptr,
           *ndc = nullptr, *corr = nullptr, *skew = nullptr;
    unsigned long long* maxSkewBits = nullptr;
};
void launch_geometry(hipStream_t st, const GeoArgs& g);

// ------------------------------------------------------------------ explicit corrector operators
// Mesh view of the fvc kernels: level-0 slots and boundary lists for the per-cell
// gathers, owner/neighbour for the per-face ones, and the geometry (face arrays hold
// nf + nb entries; Sf, d, skew interleaved xyz).
struct FvcGeom {
    int n = 0, nf = 0, nb = 0;
    SellView S;
    const int* bOff = nullptr;
    const int* bList = nullptr;
    const int* own = nullptr;
    const int* nei = nullptr;
    const double *V = nullptr, *Sf = nullptr, *magSf = nullptr, *w = nullptr, *d = nullptr, *skew = nullptr;
    int skewed = 0;
    int emptyDir = -1;  // 2-D meshes: the empty direction (gradients: 1 on its LS diagonal, component zeroed)
};
// fvc::interpolate (Fvc.h:13-31,72-81); grad nullable (skew correction when G.skewed);
// bval nullable (boundary faces take the owner's value)
void launch_fvc_interpolate(hipStream_t st, const FvcGeom& G, int nc, const double* c, const double* grad,
                            const double* bval, double* out);
// fvc::gradLS / grad(LeastSquares) (Fvc.h:134-206)
void launch_fvc_grad_ls(hipStream_t st, const FvcGeom& G, int nc, const double* c, const int* bFix,
                        const double* bVal, double* out);
// fvc::grad GaussLinear / GaussSkew of a scalar (Fvc.h:213-259); gls: least-squares gradient (GaussSkew)
void launch_fvc_grad_gauss(hipStream_t st, const FvcGeom& G, bool skew, const double* psi, const double* gls,
                           const int* bFix, const double* bVal, double* out);
// fvc::div (div = true, nc = 1) or fvc::surfaceIntegrate (Fvc.h:294-318)
void launch_fvc_integrate(hipStream_t st, const FvcGeom& G, int nc, bool div, const double* F, double* out);
// face flux of a cell vector field, phiHbyA form (IncompressibleFlow.cpp:276-291)
void launch_fvc_flux(hipStream_t st, const FvcGeom& G, const double* U, const double* grad, const int* bFix,
                     const double* Ub, double* phi);
// pressure-corrector steps (piso_kernels.hip)
void launch_piso_rau_hbya(hipStream_t st, int n, const double* A, const double* H, double* rAU, double* HbyA);
void launch_piso_hbya_boundary(hipStream_t st, const FvcGeom& G, const int* bFix, const double* Ub, const double* rAU,
                               const double* gp, double* bVal);
void launch_piso_face_op(hipStream_t st, int op, int nF, const double* a, const double* b, const double* c,
                         double* out);
void launch_piso_ddt_phi_corr(hipStream_t st, const FvcGeom& G, double coeff, double c0, double c00, const double* U0,
                              const double* U00, const double* phiOld, const double* phiOldOld, const double* rAUf,
                              double* phiHbyA);
void launch_piso_constrain_pressure(hipStream_t st, const FvcGeom& G, int b0, int cnt, const double* phiHbyA,
                                    const double* Ub, const double* rAUf, double* g);
void launch_fvc_reconstruct(hipStream_t st, const FvcGeom& G, const double* ssf, int emptyDir, double* out);
void launch_piso_correct_U(hipStream_t st, int n, const double* HbyA, const double* rAU, const double* v, int sign,
                           int emptyDir, double* U);
void launch_boundary_values(hipStream_t st, const FvcGeom& G, int nc, const int* bFix, const double* bValue,
                            const double* c, double* out);
void launch_copy_strided(hipStream_t st, long long n, const double* src, int ss, double* dst, int ds, double scale);

// ------------------------------------------------------------------ level-scheduled sweeps
// Lower / upper neighbour lists of the cells (device arrays). Lower list of cell u: neighbours l < u in ascending (l, face) order;
// upper list of cell l: neighbours u > l in descending (u, face) order; src = face << 1 |
// side of the list's own row (0: upper[face], 1: lower[face]).
struct SweepView {
    const int* lowOff = nullptr;
    const int* lowCol = nullptr;
    const int* lowSrc = nullptr;
    const int* upOff = nullptr;
    const int* upCol = nullptr;
    const int* upSrc = nullptr;
};
// A level order prepared for launching (DIC/DILU factor): cells of level L are
// perm[off[L] .. off[L+1]); groups [grp[g], grp[g+1]) of consecutive levels are either
// a run of levels of at most SWEEP_RUN_MAX cells each, swept by one workgroup in one
// launch (grpRun[g] = 1), or a single larger level with a grid launch.
constexpr int SWEEP_RUN_MAX = 1024;
struct LevelSched {
    const int* perm = nullptr;   // device
    const int* offD = nullptr;   // device level offsets (nLev + 1)
    const int* offH = nullptr;   // host copy of the offsets
    int nLev = 0;
    const int* grpH = nullptr;   // host: nGrp + 1 group boundaries (level indices)
    const char* grpRunH = nullptr;
    int nGrp = 0;
};
// DIC/DILU factor over the forward levels: dF, rD = 1/dF per cell, pf per lower-list
// entry, pb per upper-list entry
void launch_dilu_factor(hipStream_t st, const SweepView& s, const LevelSched& fwd, const double* diag,
                        const double* upper, const double* lower, double* dF, double* rD, double* pf, double* pb);
// Level-ordered ELL rows of a triangular sweep: positions in level order, the cell at
// each position, and entry k of row t of level L at base[L] + k·size_L + t (neighbour
// col, −1 padding; the coefficients come per matrix, in the same layout). Groups
// [grp[g], grp[g+1]) of levels are one level of more than 1024 cells swept by a grid
// (ELL_GRID), or a run of at most ELL_RUN_LEVELS consecutive levels swept by one
// workgroup, a row per thread, whose levels hold at most 256, 512 or 1024 cells
// (ELL_RUN256 < ELL_RUN512 < ELL_RUN1024: the larger class covers the smaller).
// A run keeps x in LDS: every cell its rows write or read has a slot there (at most
// RUN_LDS_CELLS, runCells[runOff[g] ..] in slot order), filled from x when the run
// starts and updated in place; entries of run rows hold the neighbour's slot instead of
// its cell, and slot[pos] is the slot of the row's own cell.
constexpr int ELL_GRID = 0, ELL_RUN256 = 1, ELL_RUN512 = 2, ELL_RUN1024 = 3;
constexpr int ELL_RUN_LEVELS = 2048;
constexpr int RUN_LDS_CELLS = 16384;
struct LevelEll {
    const int* cell = nullptr;    // device
    const int* slot = nullptr;    // device: per position, the LDS slot of its cell (runs)
    const int* runCells = nullptr;  // device: cells of each run's LDS slots
    const int* runOffH = nullptr;   // host: per group, first entry in runCells
    const int* runCntH = nullptr;   // host: per group, number of slots (0: grid)
    const int* col = nullptr;     // device
    const int* offD = nullptr;    // device: nLev + 1 level offsets
    const int* widthD = nullptr;  // device: entries per row of each level
    const int* baseD = nullptr;   // device: first entry of each level
    const int* offH = nullptr;    // host copies
    const int* widthH = nullptr;
    const int* baseH = nullptr;
    int nLev = 0;
    const int* grpH = nullptr;
    const char* grpKindH = nullptr;
    int nGrp = 0;
    // sync-free sweeps: per wavefront (64 rows of one level) its level and first row
    const int2* waves = nullptr;  // device
    int nWaves = 0;
    // per level, the first runCells entry of the LDS run it belongs to (−1: none): a run
    // level's entries hold LDS slots, cell = runCells[runBase + slot]
    const int* runBaseD = nullptr;  // device
};
enum SweepOp { SW_GS = 0, SW_FWD = 1, SW_BWD = 2 };
template <class T> struct SweepArgs {
    const T* coef = nullptr;  // per entry (LevelEll layout)
    const T* b = nullptr;     // SW_GS
    const T* diag = nullptr;  // SW_GS
    const T* rD = nullptr;    // SW_FWD
    const T* r = nullptr;     // SW_FWD
    T* x = nullptr;           // updated in place (w for DILU)
};
// Sync-free forms of the sweeps (one launch per sweep, no level barriers): a row waits
// for each neighbour whose new value it reads until that value is published in xNew,
// which starts filled with a sentinel; values it reads old come from xOld (GS: the
// neighbours on the other side; backward DILU: its own forward value). The arithmetic of
// every row is the level-scheduled one, so the bits are too. fault (two device ints):
// [0] set when a wait ran past its spin limit (the sweep's results are then void), [1]
// the launch's work counter.
template <class T>
void launch_gs_sweep_free(hipStream_t st, const LevelEll& E, bool backward, const T* coef, const T* diag, const T* b,
                          const T* xOld, T* xNew, int* fault, const int* done);
template <class T>
void launch_dilu_apply_free(hipStream_t st, const LevelEll& F, const LevelEll& B, const T* rD, const T* ef,
                            const T* eb, const T* r, T* tmp, T* w, int* fault, const int* done);
// w = M⁻¹ r of DIC/DILU: forward then backward substitution (ef, eb: rD·A products per
// entry of F and B), every launch a no-op once *done
template <class T>
void launch_dilu_apply(hipStream_t st, const LevelEll& F, const LevelEll& B, const T* rD, const T* ef, const T* eb,
                       const T* r, T* w, const int* done);
// one Gauss–Seidel pass over the levels of E (forward or backward order)
template <class T>
void launch_gs_sweep(hipStream_t st, const LevelEll& E, const T* coef, const T* diag, const T* b, T* x,
                     const int* done);
// out[e] = from[src[e]] (0 where src[e] < 0)
template <class T> void launch_gather_entries(hipStream_t st, size_t n, const int* src, const T* from, T* out);
// coarsest solve with the LU factors in the reference's row order (Gauss–Seidel V-cycle)
template <class T>
void launch_coarse_solve_rows(hipStream_t st, int n, const double* lu, const T* r, T* x, const int* done);
// x = 0; x[i] += xc[agg[i]] (Gauss–Seidel V-cycle), no-ops once *done
template <class T> void launch_fill_zero(hipStream_t st, int n, T* x, const int* done);
template <class T> void launch_prolong(hipStream_t st, int n, const int* agg, const T* xc, T* x, const int* done);

// FvMatrix::solve fold (FvMatrix.h:154-163): d = diag + Σ ic, b = source + Σ bs per cell;
// su != nullptr: the fold of a copy of the matrix with −= su (b = source + su·V + Σ bs)
void launch_fvm_fold(hipStream_t st, const FvmArgs& a, double* dOut, double* bOut, const double* su = nullptr,
                     const double* V = nullptr);
// FvMatrix::relax (FvMatrix.h:111-131) in place on a.diag / a.source; lower = upper when
// the matrix has no lower part
void launch_fvm_relax(hipStream_t st, const FvmArgs& a, double alpha, const double* psi, const double* lower);
// FvMatrix::setReference (FvMatrix.h:134-141) at cell (0 <= cell < n)
void launch_fvm_set_reference(hipStream_t st, double* diag, double* source, int cell, double value);
// FvMatrix::A (FvMatrix.h:179-187); nCmpt 3: vector equation
void launch_fvm_A(hipStream_t st, const FvmArgs& a, const double* V, int nCmpt, double* out);
// FvMatrix::H (FvMatrix.h:189-208), one component
void launch_fvm_H(hipStream_t st, const FvmArgs& a, const double* lower, const double* psi, const double* V, int nCmpt,
                  double* out);
// FvMatrix::flux (FvMatrix.h:210-222): out has nf + nb entries
void launch_fvm_flux(hipStream_t st, const FvmArgs& a, const double* lower, const int* own, const int* nei,
                     const double* psi, int hasFfc, double* out);

// ------------------------------------------------------------------ batches of systems
// Independent systems concatenated (cells, faces and every GAMG level). Vector work is
// cut into chunks that never straddle two systems: chunk c = (system, first row, end
// row); a segmented reduction leaves one partial per chunk and the per-system finalize
// sums its chunks in order, so each system's arithmetic is independent of the others.
constexpr int BATCH_CHUNK = 2048;  // rows per chunk
struct BatchRed {
    const int4* chunks = nullptr;  // x = system, y = first row, z = end row
    int nChunks = 0;
    const int* sysChunkOff = nullptr;  // system → its chunks [off[s], off[s+1])
    int nSys = 0;
    double* part = nullptr;            // [MAX_RED][maxChunks]
    int maxChunks = 0;
};
// Coarsest level of one system: its rows and faces inside the concatenated level arrays.
struct CoarseDesc {
    int n = 0, nf = 0, rowBase = 0, level = 0;
    size_t luOff = 0;
    const int* l = nullptr;  // global rows of the level
    const int* u = nullptr;
    const double* d = nullptr;  // diagonal (global rows)
    const double* cu = nullptr; // face coefficients (this system's faces)
    const double* cl = nullptr;
};
// Vector op over all chunks (S = the systems' scalars; a system that is done is skipped).
void launch_bvec(hipStream_t st, int op, const VecArgs& a, const BatchRed& R);
void launch_bvec32(hipStream_t st, int op, const VecArgs32& a, const BatchRed& R);
void launch_binit_scalars(hipStream_t st, DevScalars* S, int nSys, const int* n, double tol, double relTol,
                          int maxIter, int minIter, int nSweeps, double innerRel, int keepKrylov = 0);
// Per-system scalar step after a segmented reduction (I: inner scalars of mixed precision).
void launch_bfinalize(hipStream_t st, int op, int nred, const BatchRed& R, DevScalars* S, DevScalars* I);
// nDone[0] = number of systems whose scalars say done.
void launch_bcount(hipStream_t st, const DevScalars* S, int nSys, int* nDone);
// Precision changes per chunk (see MixOp), skipped for systems that are done.
void launch_bmix(hipStream_t st, int op, const BatchRed& R, const DevScalars* S, const double* a, float* w,
                 float* w2, float* z, double* x, const float* e);
// Dense assembly + no-pivot LU of every system's coarsest level (one workgroup each).
void launch_bcoarse_factor(hipStream_t st, const CoarseDesc* D, int nSys, double* lu, double* luT, int* status);
// Coarsest substitutions: r and x of system s live at rLev[D.level] + D.rowBase.
void launch_bcoarse_solve(hipStream_t st, const CoarseDesc* D, int nSys, int maxN, const double* lu,
                          const double* luT, const double* const* rLev, double* const* xLev, const DevScalars* S);
void launch_bcoarse_solve32(hipStream_t st, const CoarseDesc* D, int nSys, int maxN, const double* lu,
                            const double* luT, const float* const* rLev, float* const* xLev, const DevScalars* S);

}  // namespace vol

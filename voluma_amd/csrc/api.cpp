// Host side (C++20) of the LDU linear-solve library: handles, setup and the Krylov
// drivers, exported through the C ABI of include/voluma_amd.h.
//
// The solve loop keeps every scalar on the device (DevScalars): each reduction is a
// kernel writing per-block partials plus a one-block finalize that performs the
// solver's scalar step (α, β, convergence test) and raises a `done` flag; all later
// kernels of the iteration test that flag and exit. The host therefore enqueues
// iteration k+1 before waiting for iteration k, and never stalls the GPU.
#include "../../include/voluma_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <tuple>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "host_common.hpp"

using namespace vol;
using namespace vol::host;

namespace {

thread_local std::string gLastError;

// ---------------------------------------------------------------- no-allocation scratch
// Work memory of vol_solve_noalloc: one fixed array per GPU, held until the process
// ends. Layout (256-byte aligned pieces): 11 vectors of VOL_NOALLOC_MAX_CELLS doubles,
// the fast-sum partials (MAX_RED × NOALLOC_PART_BLOCKS doubles), the exact-sum bins
// (MAX_RED × RED_BLOCKS_EXACT × 72 int64) and non-finite flags, the stored results
// and the solve's two DevScalars. Every piece is written by the solve before it is read
// (the partials and bins by each reducing block, DevScalars by launch_init_scalars), so
// nothing has to start at zero; calls that overlap on one GPU share it and race.
constexpr int NOALLOC_VECTORS = 11;
constexpr int NOALLOC_PART_BLOCKS = 1024;  // ≥ max(RED_BLOCKS, SELL grid blocks at the cell limit)
constexpr size_t noallocRound(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_VEC_BYTES = noallocRound(sizeof(double) * size_t(VOL_NOALLOC_MAX_CELLS));
constexpr size_t NOALLOC_PART_OFF = NOALLOC_VECTORS * NOALLOC_VEC_BYTES;
constexpr size_t NOALLOC_XPART_OFF = NOALLOC_PART_OFF + noallocRound(sizeof(double) * MAX_RED * NOALLOC_PART_BLOCKS);
constexpr size_t NOALLOC_XNF_OFF =
    NOALLOC_XPART_OFF + noallocRound(sizeof(long long) * MAX_RED * RED_BLOCKS_EXACT * NBINS);
constexpr size_t NOALLOC_RESULT_OFF = NOALLOC_XNF_OFF + noallocRound(sizeof(int) * MAX_RED * RED_BLOCKS_EXACT);
constexpr size_t NOALLOC_SCAL_OFF = NOALLOC_RESULT_OFF + noallocRound(sizeof(double) * MAX_RED);
constexpr size_t NOALLOC_BYTES = NOALLOC_SCAL_OFF + noallocRound(2 * sizeof(DevScalars));
static_assert(NOALLOC_BYTES <= VOL_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the published size");

static __device__ __align__(256) unsigned char scratch[VOL_NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on each GPU (each holds its own copy), looked up once
// per device with that device current. The host-side name `scratch` is not a device
// address and is never passed to a kernel or a copy.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex gScratchMu;
unsigned char* gScratchPtr[NOALLOC_MAX_DEVICES] = {};

// Called by vol_context_create with `device` current.
void registerScratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;  // such devices get no no-allocation solve
    std::lock_guard<std::mutex> lk(gScratchMu);
    if (gScratchPtr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    gScratchPtr[device] = static_cast<unsigned char*>(p);
}

unsigned char* scratchOf(int device) {
    std::lock_guard<std::mutex> lk(gScratchMu);
    return device >= 0 && device < NOALLOC_MAX_DEVICES ? gScratchPtr[device] : nullptr;
}

}  // namespace

namespace vol::host {
void setLastError(const std::string& msg) { gLastError = msg; }
}  // namespace vol::host

// ============================================================================ handles

struct vol_mesh_s {
    std::atomic<int> refs{1};  // caller handle + matrices
    vol_context ctx = nullptr;
    int n = 0, nf = 0;
    std::vector<int> own, nei;
    SellDev sell;  // level 0 / Amul
    bool hasGamg = false;
    std::vector<HLevel> H;  // host hierarchy (kept for inspection)
    AgglomerationStats aggStats;  // pairing passes / decision rounds of the device build
    double aggSeconds = 0.0;      // time to build H
    std::vector<std::unique_ptr<LevelDev>> lv;
    DevBuf<char> ownerIsL;  // level-0 orientation
    // boundary faces (for assembly): cell → boundary-face CSR
    int nb = -1;
    DevBuf<int> bOff, bList;
    DevBuf<int> dOwn, dNei;  // device owner/neighbour, uploaded on first use
    // level schedules and neighbour lists of the triangular sweeps (DIC/DILU,
    // Gauss–Seidel), built on first use
    std::mutex sweepMu;
    bool sweepReady = false;
    // geometry of the explicit corrector operators (vol_mesh_set_geometry) and their
    // staging buffers for host pointers
    bool geomSet = false;
    int skewed = 0;
    int emptyDir = -1;  // 2-D meshes (vol_mesh_set_empty_dir)
    DevBuf<double> gV, gSf, gMagSf, gW, gD, gSkew;
    DevBuf<double> fst[8], fout, fgls;
    DevBuf<int> fsti;
    LevelOrderDev ordF;  // forward level order of level 0 (DIC/DILU factor)
    DevBuf<int> lowOff, lowCol, lowSrc, upOff, upCol, upSrc;
    LevelEllDev diluF, diluB;  // substitution rows (entry source: list index)
    LevelEllDev gsF, gsB;      // Gauss–Seidel rows of level 0 (entry source: SELL slot)
    // Gauss–Seidel GAMG smoother: rows of the coarse GAMG levels 1 .. nL−2 (index =
    // GAMG level; level 0 uses gsF/gsB)
    struct LevelOrder {
        LevelEllDev F, B;
    };
    bool gsLevelsReady = false;
    std::vector<std::unique_ptr<LevelOrder>> gsLev;
    // k consecutive Gauss–Seidel sweeps of one direction as one level schedule, per
    // (GAMG level, backward, k); built on first use
    std::map<std::tuple<int, int, int>, std::unique_ptr<LevelEllDev>> gsMulti;
    SweepView sweepView() const {
        SweepView v;
        v.lowOff = lowOff.p;
        v.lowCol = lowCol.p;
        v.lowSrc = lowSrc.p;
        v.upOff = upOff.p;
        v.upCol = upCol.p;
        v.upSrc = upSrc.p;
        return v;
    }
};

struct MatLevel {
    DevBuf<double> diag, rD, Fu, Fl, coef, bufA, bufB, r, x;
    DevBuf<double> z;  // ω·(r·rD) of the restricted residual: the first smoothing from zero
    DevBuf<float> diag32, rD32, coef32, bufA32, bufB32, r32, x32, z32;  // mixed precision
    DevBuf<double> gsCF, gsCB;  // Gauss–Seidel smoother coefficients (LevelEll layouts)
    DevBuf<float> gsCF32, gsCB32;
    DevBuf<double> coefR;  // the level's rows in restriction order (LevelDev::rrCsr)
    DevBuf<float> coefR32;
};

struct vol_matrix_s {
    vol_mesh mesh = nullptr;
    bool symmetric = true;
    bool set = false;
    DevBuf<double> diag, upper, lower, coef, rD;
    DevBuf<double> z0;  // level-0 first Jacobi sweep from zero, ω·(r·rD), written with r
    // GAMG coefficients per level, built lazily after each vol_matrix_set
    bool gamgReady = false;
    std::vector<std::unique_ptr<MatLevel>> L;
    DevBuf<double> lu, luT;
    int nCoarse = 0;
    int tailStart = -1;  // first GAMG level handled by the tail kernel (−1: none)
    int tailLdsBytes = 0, tailLdsBytes32 = 0;
    DevBuf<TailLevel> tail;
    DevBuf<TailChunk> tailChunks, tailChunks32;  // LDS preload of the tail kernel
    int tailNChunks = 0, tailNChunks32 = 0;
    DevBuf<int> status;
    // FP32 copies for mixed precision, built lazily after each vol_matrix_set
    bool mixedReady = false, mixedGamg = false;
    // DIC/DILU factor (per cell) and substitution coefficients (per list entry), built
    // lazily after each vol_matrix_set; FP32 copies for mixed precision
    bool diluReady = false, dilu32Ready = false;
    DevBuf<double> dluD, dluR, dluPf, dluPb, dluEf, dluEb;  // Ef/Eb: pf/pb in the LevelEll layouts
    DevBuf<float> dluR32, dluEf32, dluEb32;
    // Gauss–Seidel coefficients in the LevelEll layouts: level 0 (solver and smoother)
    // and the coarse levels (smoother), FP32 copies for mixed precision
    bool gsReady = false, gsLevReady = false, gs32Ready = false;
    DevBuf<double> gsCF, gsCB;
    std::map<std::tuple<int, int, int>, DevBuf<double>> gsMultiCoef;  // per gsMulti schedule
    DevBuf<float> gsCF32, gsCB32;
    DevBuf<float> coef32, diag32, rD32;
    DevBuf<TailLevel32> tail32;
    DevBuf<float> ir, iw, ip, iq, ie, ir0, iy, iAy, is, iz, it;  // inner-solve vectors
    DevBuf<float> iz0;  // inner PCG: ω·(r·rD) kept with the inner residual
    DevBuf<double> diluTmp;  // sync-free DILU: the forward substitution's result
    DevBuf<float> diluTmp32;
    // solve work vectors
    DevBuf<double> vx, vb, vr, vw, vp, vq, wA, pA, r0, y, Ay, s, z, t;
    DevBuf<double> hx, hy;  // staging for host-pointer calls
    DevBuf<double> st0, st1, st2, st3, st4, st5;  // assembly staging
    DevBuf<int> sti;
    // FvMatrix parts of device assembly (vol_fvm_*): unfolded diag, source,
    // internalCoeffs, boundarySource, face-flux correction; upper/lower are shared
    bool fvmActive = false, fvmHasLower = false;
    bool fvmHasFfc = false;  // some laplacian term switched its non-orthogonal correction on
    DevBuf<double> fdiag, fsrc, fic, fbs, fffc, fgrad, ffcT;
    DevBuf<double> fout;  // staging for host-pointer results
    DevBuf<int> anyCorr;
    DevBuf<double> stg[9];  // staging for host inputs
};

namespace {



void setDevice(vol_context ctx) { HIPCHECK(hipSetDevice(ctx->device)); }

// ------------------------------------------------------------------ mesh setup
void uploadLevels(vol_mesh m, const std::vector<int>* layout = nullptr);

void setupMesh(vol_mesh m, const double* w, int mergeLevels) {
    hipStream_t st = m->ctx->stream;
    const int n = m->n, nf = m->nf;
    // level-0 rows: faces ascending; owner row → upper[f] (side 0), neighbour row → lower[f] (side 1)
    std::vector<int> off(n + 1, 0);
    for (int f = 0; f < nf; ++f) {
        ++off[m->own[f] + 1];
        ++off[m->nei[f] + 1];
    }
    for (int c = 0; c < n; ++c) off[c + 1] += off[c];
    std::vector<int> col(off[n]), src(off[n]);
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int f = 0; f < nf; ++f) {
            int k = fill[m->own[f]]++;
            col[k] = m->nei[f];
            src[k] = (f << 1) | 0;
            k = fill[m->nei[f]]++;
            col[k] = m->own[f];
            src[k] = (f << 1) | 1;
        }
    }
    buildSell(n, off, col, src, m->sell, st, true, true);
    if (!w) return;
    m->hasGamg = true;
    // agglomeration on the GPU for meshes of ≥ 32768 cells (below, the launches of the
    // decision rounds cost more than the host loop); VOLUMA_AGGLOMERATION=host / device
    // fixes the choice. Both give the same hierarchy.
    const char* where = std::getenv("VOLUMA_AGGLOMERATION");
    const std::string wh = where ? where : "";
    const bool onHost = wh == "host" || (wh != "device" && n < 32768);
    const auto t0 = std::chrono::steady_clock::now();
    if (onHost) {
        m->H = buildHierarchy(n, nf, m->own.data(), m->nei.data(), w, mergeLevels);
        m->aggStats = AgglomerationStats{};
    } else {
        m->H = buildHierarchyDevice(st, n, nf, m->own.data(), m->nei.data(), w, mergeLevels, &m->aggStats);
    }
    m->aggSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    uploadLevels(m);
}

// Device structures of the GAMG levels of m->H (SELL/CSR rows, transfer lists).
// layout (nullable): the coarse layout of every level, as chosen for the mesh this one
// copies (otherwise measured level by level).
void uploadLevels(vol_mesh m, const std::vector<int>* layout) {
    hipStream_t st = m->ctx->stream;
    const int nf = m->nf;
    const int nL = int(m->H.size());
    std::vector<char> oil(nf);
    for (int f = 0; f < nf; ++f) oil[f] = char(m->own[f] < m->nei[f]);
    m->ownerIsL.upload(oil, st);
    for (int k = 0; k < nL; ++k) {
        const HLevel& F = m->H[k];
        auto D = std::make_unique<LevelDev>();
        D->n = F.n;
        D->nf = int(F.l.size());
        D->l.upload(F.l, st);
        D->u.upload(F.u, st);
        if (k > 0) {  // SELL rows in adjacency (face) order: A(i, j) = upper if i is the l side
            std::vector<int> srcL(F.adjFace.size());
            for (size_t q = 0; q < srcL.size(); ++q) srcL[q] = (F.adjFace[q] << 1) | (F.adjIsL[q] ? 0 : 1);
            D->layout = buildCoarseLayout(F.n, F.off, F.adjCell, srcL, D->sell, st, layout ? (*layout)[k] : -1,
                                          k + 1 < nL ? &F.agg : nullptr, m->ctx->sellSort, m->ctx->sellPairs);
        }
        if (k + 1 < nL) {
            const HLevel& C = m->H[k + 1];
            D->agg.upload(F.agg, st);
            const GatherLists G = buildGatherLists(F.n, C.n, F.agg, F.l, F.faceMap, F.faceFlip, int(C.l.size()));
            G.upload(*D, st);
            // fused residual + restriction: CSR levels read their own rows; SELL levels up
            // to the size limit a copy of their rows in the order of the restriction lists
            int lanes = 0;
            std::vector<int4> rows(G.mList.size());
            if (k == 0 && m->sell.colP.p && m->ctx->fusedLevel0) {
                lanes = 1;  // packed level-0 rows: one lane per member row, through mList
                rows.clear();
            } else if (k > 0 && D->layout == 1 && m->ctx->fusedSellDirect && F.n > m->ctx->fusedSellMax) {
                lanes = 1;  // a SELL level's own rows, likewise
                rows.clear();
            } else if (k > 0 && D->layout == 0) {
                lanes = D->sell.csrG;
                for (size_t p = 0; p < rows.size(); ++p) {
                    const int i = G.mList[p];
                    rows[p] = make_int4(i, F.off[i], F.off[i + 1] - F.off[i], 0);
                }
            } else if (k > 0 && D->layout == 1 && F.n <= m->ctx->fusedSellMax) {
                std::vector<int> offP(1, 0), colP, srcP;
                offP.reserve(F.n + 1);
                colP.reserve(F.adjCell.size());
                srcP.reserve(F.adjCell.size());
                for (size_t p = 0; p < rows.size(); ++p) {
                    const int i = G.mList[p];
                    rows[p] = make_int4(i, offP.back(), F.off[i + 1] - F.off[i], 0);
                    for (int q = F.off[i]; q < F.off[i + 1]; ++q) {
                        colP.push_back(F.adjCell[q]);
                        srcP.push_back((F.adjFace[q] << 1) | (F.adjIsL[q] ? 0 : 1));
                    }
                    offP.push_back(int(colP.size()));
                }
                if (buildCsr(int(rows.size()), offP, colP, srcP, D->rrCsr, st)) lanes = D->rrCsr.csrG;
            }
            const std::vector<int> blk = lanes ? restrictBlocks(G.mOff, lanes) : std::vector<int>();
            if (!blk.empty()) {
                std::vector<int2> b2(blk.size());
                for (size_t q = 0; q < blk.size(); ++q) b2[q] = make_int2(blk[q], G.mOff[blk[q]]);
                D->rrBlk.upload(b2, st);
                if (!rows.empty()) D->rrRows.upload(rows, st);
                D->rrBlocks = int(blk.size()) - 1;
            }
        }
        m->lv.push_back(std::move(D));
    }
    if (m->H.back().n > 4096)
        throw VolError(VOL_ERR_UNSUPPORTED, "GAMG: coarsest level larger than 4096 cells (agglomeration stalled)");
}

// ------------------------------------------------------------------ matrix / GAMG coefficients
// LDS layout of the tail kernel: [coarsest LU or substitution vector (doubles)]
// [r, two work buffers, x per level (T)] [per level, when it fits: row offsets, columns,
// coefficients, diagonal, 1/diagonal, agg, restriction lists]. Returns the bytes used.
template <class T> int tailLayout(std::vector<TailLevelT<T>>& V, int coarseN) {
    const size_t coarseBytes = size_t(coarseN <= 64 ? coarseN * coarseN : coarseN) * sizeof(double);
    size_t off = coarseBytes / sizeof(T);  // in elements of T
    for (auto& t : V) {
        t.ldsR = int(off);
        t.ldsA = int(off + t.S.n);
        t.ldsB = int(off + 2 * size_t(t.S.n));
        t.ldsX = int(off + 3 * size_t(t.S.n));
        off += 4 * size_t(t.S.n);
    }
    size_t bytes = off * sizeof(T);
    auto al = [](size_t b) { return (b + 15) & ~size_t(15); };
    for (size_t k = 0; k + 1 < V.size(); ++k) {
        TailLevelT<T>& t = V[k];
        const size_t n = t.S.n, ns = t.S.nslots, nr = t.S.csrG > 0 ? n + 1 : size_t(t.S.nslices) + 1;
        size_t o = al(bytes);
        const size_t oRow = o;
        o = al(o + 4 * nr);
        const size_t oCol = o;
        o = al(o + 4 * ns);
        const size_t oCoef = o;
        o = al(o + sizeof(T) * ns);
        const size_t oDiag = o;
        o = al(o + sizeof(T) * n);
        const size_t oRD = o;
        o = al(o + sizeof(T) * n);
        const size_t oAgg = o;
        o = al(o + (t.agg ? 4 * n : 0));
        const size_t oMOff = o;
        o = al(o + (t.mOff ? 4 * (size_t(t.nNext) + 1) : 0));
        const size_t oMList = o;
        o = al(o + (t.mList ? 4 * n : 0));
        if (o > size_t(TAIL_LDS_BYTES)) continue;  // this level's matrix stays in global memory
        t.ldsMat = 1;
        t.nRowOff = int(nr);
        t.ldsRowOff = int(oRow);
        t.ldsCol = int(oCol);
        t.ldsCoef = int(oCoef);
        t.ldsDiag = int(oDiag);
        t.ldsRD = int(oRD);
        t.ldsAgg = int(oAgg);
        t.ldsMOff = int(oMOff);
        t.ldsMList = int(oMList);
        bytes = o;
    }
    return int(bytes);
}

// The chunks the tail kernel preloads into LDS: the matrices of the levels that fit
// (layout of tailLayout), the coarsest LU factors when the coarsest has ≤ 64 cells, and
// rTop (the first tail level's r) when given (its address is fixed below level 0).
template <class T>
std::vector<TailChunk> tailChunks(const std::vector<TailLevelT<T>>& V, int coarseN, const double* luT, const T* rTop) {
    std::vector<TailChunk> C;
    auto add = [&](const void* src, size_t byteOff, size_t bytes) {
        const unsigned* s = static_cast<const unsigned*>(src);
        const int dst = int(byteOff / 4);
        const int words = int(bytes / 4);
        for (int w = 0; w < words; w += TAIL_CHUNK_WORDS) {
            TailChunk c;
            c.src = s + w;
            c.dst = dst + w;
            c.n = std::min(TAIL_CHUNK_WORDS, words - w);
            C.push_back(c);
        }
    };
    if (coarseN <= 64) add(luT, 0, sizeof(double) * size_t(coarseN) * coarseN);
    if (rTop) add(rTop, sizeof(T) * size_t(V[0].ldsR), sizeof(T) * size_t(V[0].S.n));
    for (size_t k = 0; k + 1 < V.size(); ++k) {
        const TailLevelT<T>& t = V[k];
        if (!t.ldsMat) continue;
        const size_t n = t.S.n, ns = t.S.nslots;
        add(t.S.sliceOff, t.ldsRowOff, 4 * size_t(t.nRowOff));
        add(t.S.col, t.ldsCol, 4 * ns);
        add(t.coef, t.ldsCoef, sizeof(T) * ns);
        add(t.diag, t.ldsDiag, sizeof(T) * n);
        add(t.rD, t.ldsRD, sizeof(T) * n);
        if (t.agg) add(t.agg, t.ldsAgg, 4 * n);
        if (t.mOff) add(t.mOff, t.ldsMOff, 4 * (size_t(t.nNext) + 1));
        if (t.mList) add(t.mList, t.ldsMList, 4 * n);
    }
    return C;
}

void buildGamgCoeffs(vol_matrix A) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    const int nL = int(m->lv.size());
    if (A->L.size() != size_t(nL)) {
        A->L.clear();
        for (int k = 0; k < nL; ++k) {
            auto M = std::make_unique<MatLevel>();
            const LevelDev& D = *m->lv[k];
            if (k > 0) {
                M->diag.alloc(D.n);
                M->rD.alloc(D.n);
                M->coef.alloc(size_t(D.sell.nslots) + D.sell.nslotsP);  // (+ the pair copy)
                M->r.alloc(D.n);
                M->x.alloc(D.n);
                M->z.alloc(D.n);
            }
            M->Fu.alloc(D.nf);
            M->Fl.alloc(D.nf);
            M->bufA.alloc(D.n);
            M->bufB.alloc(D.n);
            A->L.push_back(std::move(M));
        }
    }
    launch_orient0(st, m->nf, m->ownerIsL.p, A->upper.p, A->symmetric ? A->upper.p : A->lower.p, A->L[0]->Fu.p,
                   A->L[0]->Fl.p);
    for (int k = 0; k + 1 < nL; ++k) {
        const LevelDev& D = *m->lv[k];
        const LevelDev& DC = *m->lv[k + 1];
        const double* Fd = k == 0 ? A->diag.p : A->L[k]->diag.p;
        launch_galerkin(st, DC.n, DC.nf, D.mOff.p, D.mList.p, D.iOff.p, D.iList.p, D.cfOff.p, D.cfList.p, Fd,
                        A->L[k]->Fu.p, A->L[k]->Fl.p, A->L[k + 1]->diag.p, A->L[k + 1]->Fu.p, A->L[k + 1]->Fl.p);
        launch_sell_coef(st, DC.sell.view(), A->L[k + 1]->Fu.p, A->L[k + 1]->Fl.p, A->L[k + 1]->coef.p, nullptr);
        if (DC.rrCsr.nslots > 0) {
            A->L[k + 1]->coefR.alloc(DC.rrCsr.nslots);
            launch_sell_coef(st, DC.rrCsr.view(), A->L[k + 1]->Fu.p, A->L[k + 1]->Fl.p, A->L[k + 1]->coefR.p, nullptr);
        }
        VecArgs v;  // 1/diag of the coarse level
        v.n = DC.n;
        v.a = A->L[k + 1]->diag.p;
        v.w = A->L[k + 1]->rD.p;
        launch_vec(st, V_RECIP, v, false, m->ctx->red);
    }
    const LevelDev& DL = *m->lv[nL - 1];
    A->nCoarse = DL.n;
    A->lu.alloc(size_t(DL.n) * DL.n);
    A->luT.alloc(size_t(DL.n) * DL.n);
    A->status.alloc(1);
    const double* cd = nL == 1 ? A->diag.p : A->L[nL - 1]->diag.p;
    launch_coarse_assemble(st, DL.n, DL.nf, DL.l.p, DL.u.p, cd, A->L[nL - 1]->Fu.p, A->L[nL - 1]->Fl.p, A->lu.p);
    launch_coarse_lu(st, DL.n, A->lu.p, A->luT.p, A->status.p);
    // levels at or below the tail threshold (and their coarsest) go to one workgroup,
    // provided their vectors fit in LDS
    A->tailStart = -1;
    auto ldsNeed = [&](int k0) {
        size_t d = 0;
        for (int k = k0; k < nL; ++k) d += 4 * size_t(m->lv[k]->n);
        const size_t nc = size_t(DL.n);
        return (nc <= 64 ? nc * nc : nc) * sizeof(double) + d * sizeof(double);
    };
    for (int k = 0; k < nL; ++k)
        if (m->lv[k]->n <= m->ctx->tailMax && ldsNeed(k) <= size_t(TAIL_LDS_BYTES)) {
            A->tailStart = k;
            break;
        }
    if (A->tailStart >= 0) {
        std::vector<TailLevel> T;
        for (int k = A->tailStart; k < nL; ++k) {
            const LevelDev& D = *m->lv[k];
            MatLevel& M = *A->L[k];
            TailLevel t;
            t.S = k == 0 ? m->sell.view() : D.sell.view();
            t.coef = k == 0 ? A->coef.p : M.coef.p;
            t.diag = k == 0 ? A->diag.p : M.diag.p;
            t.rD = k == 0 ? A->rD.p : M.rD.p;
            if (k + 1 < nL) {
                t.agg = D.agg.p;
                t.mOff = D.mOff.p;
                t.mList = D.mList.p;
                t.nNext = m->lv[k + 1]->n;
            }
            T.push_back(t);
        }
        A->tailLdsBytes = tailLayout(T, DL.n);
        A->tail.upload(T, st);
        const double* rTop = A->tailStart > 0 ? A->L[A->tailStart]->r.p : nullptr;
        const std::vector<TailChunk> ch = tailChunks<double>(T, DL.n, A->luT.p, rTop);
        A->tailChunks.upload(ch, st);
        A->tailNChunks = int(ch.size());
    }
    int status = 0;
    HIPCHECK(hipMemcpyAsync(&status, A->status.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (status != 0) throw VolError(VOL_ERR_SINGULAR_COARSEST, "GAMG: singular coarsest matrix");
    A->gamgReady = true;
    A->mixedReady = false;
}

void toFloat(hipStream_t st, const double* a, float* w, size_t n) {
    launch_mix(st, MIX_TO32, int(n), a, w, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// FP32 copies of the level-0 matrix and, with GAMG, of every coarse level and the tail
// descriptors (the coarsest LU stays FP64). Values are rounded from the FP64 ones.
// Level schedules of the triangular sweeps (sweep_kernels.hip). Lower list of cell u:
// its neighbours l < u in ascending (l, face) order, the order of the reference's
// (u, l)-sorted face list (LinearSolver.cpp:29-51); upper list of cell l: its neighbours
// u > l in descending (u, face) order, the order of the backward face loop.
void buildSweep(vol_mesh m) {
    std::lock_guard<std::mutex> guard(m->sweepMu);
    if (m->sweepReady) return;
    const int n = m->n, nf = m->nf;
    std::vector<int> lowOff(n + 1, 0), upOff(n + 1, 0);
    for (int f = 0; f < nf; ++f) {
        const int o = m->own[f], e = m->nei[f];
        ++lowOff[std::max(o, e) + 1];
        ++upOff[std::min(o, e) + 1];
    }
    for (int c = 0; c < n; ++c) {
        lowOff[c + 1] += lowOff[c];
        upOff[c + 1] += upOff[c];
    }
    // entries (neighbour, face); side of the list's own row: 0 when it owns the face
    std::vector<std::pair<int, int>> low(nf), up(nf);
    {
        std::vector<int> fl(lowOff.begin(), lowOff.end() - 1), fu(upOff.begin(), upOff.end() - 1);
        for (int f = 0; f < nf; ++f) {
            const int o = m->own[f], e = m->nei[f];
            const int l = std::min(o, e), u = std::max(o, e);
            low[fl[u]++] = {l, f};
            up[fu[l]++] = {u, f};
        }
    }
    std::vector<int> lowCol(nf), lowSrc(nf), upCol(nf), upSrc(nf), levF(n, 0), levB(n, 0);
    for (int c = 0; c < n; ++c) {
        std::sort(low.begin() + lowOff[c], low.begin() + lowOff[c + 1]);
        std::sort(up.begin() + upOff[c], up.begin() + upOff[c + 1], std::greater<>());
        for (int k = lowOff[c]; k < lowOff[c + 1]; ++k) {
            const int f = low[k].second;
            lowCol[k] = low[k].first;
            lowSrc[k] = (f << 1) | (m->own[f] == c ? 0 : 1);
            levF[c] = std::max(levF[c], levF[low[k].first] + 1);
        }
        for (int k = upOff[c]; k < upOff[c + 1]; ++k) {
            const int f = up[k].second;
            upCol[k] = up[k].first;
            upSrc[k] = (f << 1) | (m->own[f] == c ? 0 : 1);
        }
    }
    for (int c = n - 1; c >= 0; --c)
        for (int k = upOff[c]; k < upOff[c + 1]; ++k) levB[c] = std::max(levB[c], levB[upCol[k]] + 1);
    auto order = [n](const std::vector<int>& lev, std::vector<int>& off, std::vector<int>& perm) {
        const int nl = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
        off.assign(nl + 1, 0);
        for (int c = 0; c < n; ++c) ++off[lev[c] + 1];
        for (int L = 0; L < nl; ++L) off[L + 1] += off[L];
        std::vector<int> fill(off.begin(), off.end() - 1);
        perm.assign(n, 0);
        for (int c = 0; c < n; ++c) perm[fill[lev[c]]++] = c;
    };
    std::vector<int> permF, permB, offF, offB;
    order(levF, offF, permF);
    order(levB, offB, permB);
    hipStream_t st = m->ctx->stream;
    m->ordF.build(offF, permF, st);
    std::vector<int> idx(nf);
    for (int k = 0; k < nf; ++k) idx[k] = k;
    m->diluF.build(offF, permF, lowOff, lowCol, idx, st);
    m->diluB.build(offB, permB, upOff, upCol, idx, st);
    {
        std::vector<int> ro, rc, rs;
        sellRows(m->sell, ro, rc, rs);
        m->gsF.build(offF, permF, ro, rc, rs, st);
        m->gsB.build(offB, permB, ro, rc, rs, st);
    }
    m->lowOff.upload(lowOff, st);
    m->upOff.upload(upOff, st);
    if (nf) {
        m->lowCol.upload(lowCol, st);
        m->lowSrc.upload(lowSrc, st);
        m->upCol.upload(upCol, st);
        m->upSrc.upload(upSrc, st);
    }
    HIPCHECK(hipStreamSynchronize(st));
    m->sweepReady = true;
}

// Forward / backward level order of a graph given by its edges (l[k] < u[k]): cells by
// level, then by index; offF/offB hold the level offsets.
void levelOrder(int n, const std::vector<int>& l, const std::vector<int>& u, std::vector<int>& offF,
                std::vector<int>& permF, std::vector<int>& offB, std::vector<int>& permB) {
    const int nf = int(l.size());
    std::vector<int> lo(n + 1, 0), hi(n + 1, 0), lc(nf), hc(nf);
    for (int k = 0; k < nf; ++k) {
        ++lo[u[k] + 1];
        ++hi[l[k] + 1];
    }
    for (int c = 0; c < n; ++c) {
        lo[c + 1] += lo[c];
        hi[c + 1] += hi[c];
    }
    {
        std::vector<int> fl(lo.begin(), lo.end() - 1), fh(hi.begin(), hi.end() - 1);
        for (int k = 0; k < nf; ++k) {
            lc[fl[u[k]]++] = l[k];
            hc[fh[l[k]]++] = u[k];
        }
    }
    std::vector<int> levF(n, 0), levB(n, 0);
    for (int c = 0; c < n; ++c)
        for (int k = lo[c]; k < lo[c + 1]; ++k) levF[c] = std::max(levF[c], levF[lc[k]] + 1);
    for (int c = n - 1; c >= 0; --c)
        for (int k = hi[c]; k < hi[c + 1]; ++k) levB[c] = std::max(levB[c], levB[hc[k]] + 1);
    auto order = [n](const std::vector<int>& lev, std::vector<int>& off, std::vector<int>& perm) {
        const int nl = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
        off.assign(nl + 1, 0);
        for (int c = 0; c < n; ++c) ++off[lev[c] + 1];
        for (int L = 0; L < nl; ++L) off[L + 1] += off[L];
        std::vector<int> fill(off.begin(), off.end() - 1);
        perm.assign(n, 0);
        for (int c = 0; c < n; ++c) perm[fill[lev[c]]++] = c;
    };
    order(levF, offF, permF);
    order(levB, offB, permB);
}

// Level orders of the coarse GAMG levels for the Gauss–Seidel smoother.
void buildGsLevels(vol_mesh m) {
    buildSweep(m);
    std::lock_guard<std::mutex> guard(m->sweepMu);
    if (m->gsLevelsReady) return;
    const int nL = int(m->lv.size());
    m->gsLev.clear();
    m->gsLev.resize(std::max(nL, 1));
    for (int k = 1; k + 1 < nL; ++k) {
        auto o = std::make_unique<vol_mesh_s::LevelOrder>();
        std::vector<int> pf, pb, of, ob, ro, rc, rs;
        levelOrder(m->H[k].n, m->H[k].l, m->H[k].u, of, pf, ob, pb);
        sellRows(m->lv[k]->sell, ro, rc, rs);
        o->F.build(of, pf, ro, rc, rs, m->ctx->stream);
        o->B.build(ob, pb, ro, rc, rs, m->ctx->stream);
        m->gsLev[k] = std::move(o);
    }
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    m->gsLevelsReady = true;
}

// DIC/DILU factor of the matrix (LinearSolver.cpp:66-78) on the forward levels.
void buildDilu(vol_matrix A) {
    vol_mesh m = A->mesh;
    buildSweep(m);
    hipStream_t st = m->ctx->stream;
    const size_t n = m->n, nf = m->nf;
    A->dluD.alloc(n);
    A->dluR.alloc(n);
    A->dluPf.alloc(std::max<size_t>(nf, 1));
    A->dluPb.alloc(std::max<size_t>(nf, 1));
    launch_dilu_factor(st, m->sweepView(), m->ordF.view(), A->diag.p, A->upper.p,
                       A->symmetric ? A->upper.p : A->lower.p, A->dluD.p, A->dluR.p, A->dluPf.p, A->dluPb.p);
    A->dluEf.alloc(std::max<size_t>(m->diluF.nEntries, 1));
    A->dluEb.alloc(std::max<size_t>(m->diluB.nEntries, 1));
    launch_gather_entries<double>(st, m->diluF.nEntries, m->diluF.src.p, A->dluPf.p, A->dluEf.p);
    launch_gather_entries<double>(st, m->diluB.nEntries, m->diluB.src.p, A->dluPb.p, A->dluEb.p);
    A->diluReady = true;
    A->dilu32Ready = false;
}

// Gauss–Seidel coefficients of the matrix in the LevelEll layouts: level 0, and the
// coarse GAMG levels when the GAMG smoother needs them.
void buildGs(vol_matrix A, bool gamgLevels) {
    vol_mesh m = A->mesh;
    buildSweep(m);
    hipStream_t st = m->ctx->stream;
    if (!A->gsReady) {
        A->gsCF.alloc(std::max<size_t>(m->gsF.nEntries, 1));
        A->gsCB.alloc(std::max<size_t>(m->gsB.nEntries, 1));
        launch_gather_entries<double>(st, m->gsF.nEntries, m->gsF.src.p, A->coef.p, A->gsCF.p);
        launch_gather_entries<double>(st, m->gsB.nEntries, m->gsB.src.p, A->coef.p, A->gsCB.p);
        A->gsReady = true;
        A->gs32Ready = false;
    }
    if (gamgLevels && !A->gsLevReady) {
        const int nL = int(m->lv.size());
        for (int k = 1; k + 1 < nL; ++k) {
            const vol_mesh_s::LevelOrder& o = *m->gsLev[k];
            MatLevel& M = *A->L[k];
            M.gsCF.alloc(std::max<size_t>(o.F.nEntries, 1));
            M.gsCB.alloc(std::max<size_t>(o.B.nEntries, 1));
            launch_gather_entries<double>(st, o.F.nEntries, o.F.src.p, M.coef.p, M.gsCF.p);
            launch_gather_entries<double>(st, o.B.nEntries, o.B.src.p, M.coef.p, M.gsCB.p);
        }
        A->gsLevReady = true;
        A->gs32Ready = false;
    }
}

// k consecutive Gauss–Seidel sweeps over a graph (edges l[e] < u[e]) as one level
// schedule. A forward sweep updates cells in ascending order, each reading its lower
// neighbours' values of this sweep and its upper neighbours' of the previous one; so
// (c, s) follows (j, s) for lower neighbours j, (j, s − 1) for upper ones and (c, s − 1):
//   L(c, s) = 1 + max(L(j, s) over lower j, L(j, s − 1) over upper j, L(c, s − 1)).
// Every read then sees the value the sequential sweeps would see and every write comes
// after the reads of the value it replaces (a lower neighbour's (j, s) precedes (c, s),
// an upper neighbour's (j, s − 1) precedes (c, s), and (c, s) precedes the upper
// neighbours' (j, s)), so the sweeps give the sequential loop's bits with far fewer
// levels than k separate sweeps (the second sweep adds only a few levels to the first).
// Backward sweeps mirror it (descending order, upper neighbours updated first). Levels
// list their cells (one entry per (cell, sweep)) in ascending cell order.
void multiSweepOrder(int n, const std::vector<int>& l, const std::vector<int>& u, bool backward, int k,
                     std::vector<int>& off, std::vector<int>& perm) {
    const int ne = int(l.size());
    std::vector<int> bo(n + 1, 0), ao(n + 1, 0), bc(ne), ac(ne);  // before / after neighbours
    auto before = [&](int e, int& c, int& j) {
        c = backward ? l[e] : u[e];
        j = backward ? u[e] : l[e];
    };
    for (int e = 0; e < ne; ++e) {
        int c, j;
        before(e, c, j);
        ++bo[c + 1];
        ++ao[j + 1];
    }
    for (int c = 0; c < n; ++c) {
        bo[c + 1] += bo[c];
        ao[c + 1] += ao[c];
    }
    {
        std::vector<int> fb(bo.begin(), bo.end() - 1), fa(ao.begin(), ao.end() - 1);
        for (int e = 0; e < ne; ++e) {
            int c, j;
            before(e, c, j);
            bc[fb[c]++] = j;
            ac[fa[j]++] = c;
        }
    }
    std::vector<int> prev(n, -1), cur(n, 0), lev;  // lev[s·n + c]
    lev.resize(size_t(k) * n);
    for (int s = 0; s < k; ++s) {
        for (int t = 0; t < n; ++t) {
            const int c = backward ? n - 1 - t : t;
            int L = prev[c] + 1;
            for (int q = bo[c]; q < bo[c + 1]; ++q) L = std::max(L, cur[bc[q]] + 1);
            for (int q = ao[c]; q < ao[c + 1]; ++q) L = std::max(L, prev[ac[q]] + 1);
            cur[c] = L;
        }
        std::copy(cur.begin(), cur.end(), lev.begin() + size_t(s) * n);
        prev = cur;
    }
    const int nl = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    off.assign(nl + 1, 0);
    for (int v : lev) ++off[v + 1];
    for (int L = 0; L < nl; ++L) off[L + 1] += off[L];
    std::vector<int> fill(off.begin(), off.end() - 1);
    perm.assign(lev.size(), 0);
    for (int c = 0; c < n; ++c)  // a cell's sweeps fall in distinct levels
        for (int s = 0; s < k; ++s) perm[fill[lev[size_t(s) * n + c]]++] = c;
}

// The k-sweep schedule of GAMG level lvl (0: the mesh) and direction, built on first use.
const LevelEllDev& gsMultiEll(vol_mesh m, int lvl, bool backward, int k) {
    std::lock_guard<std::mutex> guard(m->sweepMu);
    auto& slot = m->gsMulti[{lvl, backward ? 1 : 0, k}];
    if (slot) return *slot;
    std::vector<int> l, u;
    int n;
    if (lvl == 0) {
        n = m->n;
        l.resize(m->nf);
        u.resize(m->nf);
        for (int f = 0; f < m->nf; ++f) {
            l[f] = std::min(m->own[f], m->nei[f]);
            u[f] = std::max(m->own[f], m->nei[f]);
        }
    } else {
        n = m->H[lvl].n;
        l = m->H[lvl].l;
        u = m->H[lvl].u;
    }
    std::vector<int> off, perm, ro, rc, rs;
    multiSweepOrder(n, l, u, backward, k, off, perm);
    sellRows(lvl == 0 ? m->sell : m->lv[lvl]->sell, ro, rc, rs);
    auto e = std::make_unique<LevelEllDev>();
    e->build(off, perm, ro, rc, rs, m->ctx->stream);
    slot = std::move(e);
    return *slot;
}

// The matrix's coefficients in that schedule's layout (level coefficients gathered by the
// schedule's entry sources), built after each change of the matrix.
const double* gsMultiCoef(vol_matrix A, int lvl, bool backward, int k) {
    const LevelEllDev& E = gsMultiEll(A->mesh, lvl, backward, k);
    auto& buf = A->gsMultiCoef[{lvl, backward ? 1 : 0, k}];
    if (!buf.p) {
        buf.alloc(std::max<size_t>(E.nEntries, 1));
        launch_gather_entries<double>(A->mesh->ctx->stream, E.nEntries, E.src.p,
                                      lvl == 0 ? A->coef.p : A->L[lvl]->coef.p, buf.p);
    }
    return buf.p;
}

void buildGs32(vol_matrix A) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    A->gsCF32.alloc(std::max<size_t>(m->gsF.nEntries, 1));
    A->gsCB32.alloc(std::max<size_t>(m->gsB.nEntries, 1));
    launch_gather_entries<float>(st, m->gsF.nEntries, m->gsF.src.p, A->coef32.p, A->gsCF32.p);
    launch_gather_entries<float>(st, m->gsB.nEntries, m->gsB.src.p, A->coef32.p, A->gsCB32.p);
    const int nL = int(m->lv.size());
    for (int k = 1; k + 1 < nL; ++k) {
        const vol_mesh_s::LevelOrder& o = *m->gsLev[k];
        MatLevel& M = *A->L[k];
        M.gsCF32.alloc(std::max<size_t>(o.F.nEntries, 1));
        M.gsCB32.alloc(std::max<size_t>(o.B.nEntries, 1));
        launch_gather_entries<float>(st, o.F.nEntries, o.F.src.p, M.coef32.p, M.gsCF32.p);
        launch_gather_entries<float>(st, o.B.nEntries, o.B.src.p, M.coef32.p, M.gsCB32.p);
    }
    A->gs32Ready = true;
}

void buildDilu32(vol_matrix A) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    const size_t n = m->n, ef = std::max<size_t>(m->diluF.nEntries, 1), eb = std::max<size_t>(m->diluB.nEntries, 1);
    A->dluR32.alloc(n);
    A->dluEf32.alloc(ef);
    A->dluEb32.alloc(eb);
    toFloat(st, A->dluR.p, A->dluR32.p, n);
    toFloat(st, A->dluEf.p, A->dluEf32.p, ef);
    toFloat(st, A->dluEb.p, A->dluEb32.p, eb);
    A->dilu32Ready = true;
}

void buildMixed(vol_matrix A, bool gamg) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    A->coef32.alloc(size_t(m->sell.nslots) + m->sell.nslotsP);
    A->diag32.alloc(m->n);
    A->rD32.alloc(m->n);
    toFloat(st, A->coef.p, A->coef32.p, size_t(m->sell.nslots) + m->sell.nslotsP);
    toFloat(st, A->diag.p, A->diag32.p, m->n);
    toFloat(st, A->rD.p, A->rD32.p, m->n);
    if (gamg) {
        const int nL = int(m->lv.size());
        for (int k = 0; k < nL; ++k) {
            const LevelDev& D = *m->lv[k];
            MatLevel& M = *A->L[k];
            if (k > 0) {
                M.diag32.alloc(D.n);
                M.rD32.alloc(D.n);
                M.coef32.alloc(size_t(D.sell.nslots) + D.sell.nslotsP);
                M.r32.alloc(D.n);
                M.x32.alloc(D.n);
                M.z32.alloc(D.n);
                toFloat(st, M.diag.p, M.diag32.p, D.n);
                toFloat(st, M.rD.p, M.rD32.p, D.n);
                toFloat(st, M.coef.p, M.coef32.p, size_t(D.sell.nslots) + D.sell.nslotsP);
                if (D.rrCsr.nslots > 0) {
                    M.coefR32.alloc(D.rrCsr.nslots);
                    toFloat(st, M.coefR.p, M.coefR32.p, D.rrCsr.nslots);
                }
            }
            M.bufA32.alloc(D.n);
            M.bufB32.alloc(D.n);
        }
        if (A->tailStart >= 0) {
            const LevelDev& DL = *m->lv[nL - 1];
            std::vector<TailLevel32> T;
            for (int k = A->tailStart; k < nL; ++k) {
                const LevelDev& D = *m->lv[k];
                MatLevel& M = *A->L[k];
                TailLevel32 t;
                t.S = k == 0 ? m->sell.view() : D.sell.view();
                t.coef = k == 0 ? A->coef32.p : M.coef32.p;
                t.diag = k == 0 ? A->diag32.p : M.diag32.p;
                t.rD = k == 0 ? A->rD32.p : M.rD32.p;
                if (k + 1 < nL) {
                    t.agg = D.agg.p;
                    t.mOff = D.mOff.p;
                    t.mList = D.mList.p;
                    t.nNext = m->lv[k + 1]->n;
                }
                T.push_back(t);
            }
            A->tailLdsBytes32 = tailLayout(T, DL.n);
            A->tail32.upload(T, st);
            const float* rTop = A->tailStart > 0 ? A->L[A->tailStart]->r32.p : nullptr;
            const std::vector<TailChunk> ch = tailChunks<float>(T, DL.n, A->luT.p, rTop);
            A->tailChunks32.upload(ch, st);
            A->tailNChunks32 = int(ch.size());
        }
    }
    HIPCHECK(hipStreamSynchronize(st));
    A->mixedReady = true;
    A->mixedGamg = gamg;
}

// SELL slot coefficients and 1/diag after diag/upper/lower changed.
void matrixFinish(vol_matrix A) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    A->coef.alloc(size_t(m->sell.nslots) + m->sell.nslotsP);  // slot-major values (+ the pair copy)
    launch_sell_coef(st, m->sell.view(), A->upper.p, A->symmetric ? A->upper.p : A->lower.p, A->coef.p, nullptr);
    A->rD.alloc(m->n);
    VecArgs v;
    v.n = m->n;
    v.a = A->diag.p;
    v.w = A->rD.p;
    launch_vec(st, V_RECIP, v, false, m->ctx->red);
    A->gamgReady = false;
    A->mixedReady = false;
    A->diluReady = false;
    A->dilu32Ready = false;
    A->gsReady = false;
    A->gsLevReady = false;
    A->gs32Ready = false;
    A->gsMultiCoef.clear();
    A->set = true;
}

// Device copy of a caller array (host → staging, device → used in place).
template <class T> const T* stage(DevBuf<T>& buf, const T* p, size_t n, int location, hipStream_t st) {
    if (!p || location == VOL_DEVICE) return p;
    buf.alloc(n);
    if (n) HIPCHECK(hipMemcpyAsync(buf.p, p, sizeof(T) * n, hipMemcpyHostToDevice, st));
    return buf.p;
}

// fvm::laplacian (kind ASM_LAPLACIAN) or fvm::div upwind (ASM_DIV) into A, with the fold.
void assemble(vol_matrix A, int kind, double gammaU, const double* faceGamma, const double* magSf, const double* ndc,
              const double* flux, const int32_t* bcFixed, const double* bcValue, const double* bIn, double* bOut,
              double* icOut, double* bsOut, int location) {
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    if (m->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    if (!bcFixed && m->nb > 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bcFixed required");
    setDevice(c);
    hipStream_t st = c->stream;
    const int n = m->n, nf = m->nf, nb = m->nb;
    const size_t nF = size_t(nf) + nb;
    AsmArgs a;
    a.n = n;
    a.nf = nf;
    a.S = m->sell.view();
    a.bOff = m->bOff.p;
    a.bList = m->bList.p;
    a.gammaU = gammaU;
    a.gamma = stage(A->st0, faceGamma, nF, location, st);
    if (kind == ASM_LAPLACIAN) {
        if (!magSf || !ndc) throw VolError(VOL_ERR_INVALID_ARGUMENT, "magSf and ndc required");
        a.magSf = stage(A->st1, magSf, nF, location, st);
        a.ndc = stage(A->st2, ndc, nF, location, st);
    } else {
        if (!flux) throw VolError(VOL_ERR_INVALID_ARGUMENT, "flux required");
        a.flux = stage(A->st1, flux, nF, location, st);
    }
    a.bcFixed = stage(A->sti, bcFixed, size_t(nb), location, st);
    a.bcValue = stage(A->st3, bcValue, size_t(nb), location, st);
    a.bIn = stage(A->st4, bIn, size_t(n), location, st);
    A->symmetric = kind == ASM_LAPLACIAN;
    A->diag.alloc(n);
    A->upper.alloc(nf);
    a.diag = A->diag.p;
    a.upper = A->upper.p;
    if (!A->symmetric) {
        A->lower.alloc(nf);
        a.lower = A->lower.p;
    }
    const bool hostOut = location != VOL_DEVICE;
    DevBuf<double> tb, tic, tbs;
    if (bOut) {
        if (hostOut) tb.alloc(n);
        a.bOut = hostOut ? tb.p : bOut;
    }
    if (icOut) {
        if (hostOut) tic.alloc(nb);
        a.icOut = hostOut ? tic.p : icOut;
    }
    if (bsOut) {
        if (hostOut) tbs.alloc(nb);
        a.bsOut = hostOut ? tbs.p : bsOut;
    }
    launch_assemble(st, kind, a);
    matrixFinish(A);
    if (hostOut) {
        if (bOut) HIPCHECK(hipMemcpyAsync(bOut, tb.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        if (icOut && nb) HIPCHECK(hipMemcpyAsync(icOut, tic.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
        if (bsOut && nb) HIPCHECK(hipMemcpyAsync(bsOut, tbs.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------ FvMatrix assembly
const double* stageK(vol_matrix A, size_t k, const double* p, size_t n, int location) {
    return stage(A->stg[k], p, n, location, A->mesh->ctx->stream);
}

void fvmCheck(vol_matrix A, int sign) {
    if (!A->fvmActive) throw VolError(VOL_ERR_INVALID_ARGUMENT, "call vol_fvm_reset first");
    if (sign != 1 && sign != -1) throw VolError(VOL_ERR_INVALID_ARGUMENT, "sign must be +1 or -1");
}

FvmArgs fvmArgs(vol_matrix A, int sign) {
    vol_mesh m = A->mesh;
    FvmArgs a;
    a.n = m->n;
    a.nf = m->nf;
    a.nb = m->nb;
    a.S = m->sell.view();
    a.bOff = m->bOff.p;
    a.bList = m->bList.p;
    a.sign = double(sign);
    a.emptyDir = m->emptyDir;
    a.diag = A->fdiag.p;
    a.upper = A->upper.p;
    a.source = A->fsrc.p;
    a.ic = A->fic.p;
    a.bs = A->fbs.p;
    a.ffc = A->fffc.p;
    a.hadLower = A->fvmHasLower ? 1 : 0;
    return a;
}

void fvmReset(vol_matrix A) {
    vol_mesh m = A->mesh;
    if (m->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    hipStream_t st = m->ctx->stream;
    const size_t n = m->n, nf = m->nf, nb = m->nb;
    A->fdiag.alloc(n);
    A->upper.alloc(nf);
    A->lower.alloc(nf);
    A->fsrc.alloc(n);
    A->fic.alloc(nb);
    A->fbs.alloc(nb);
    A->fffc.alloc(nf + nb);
    const std::pair<DevBuf<double>*, size_t> parts[] = {{&A->fdiag, n}, {&A->upper, nf}, {&A->fsrc, n},
                                                       {&A->fic, nb},  {&A->fbs, nb},  {&A->fffc, nf + nb}};
    for (const auto& [b, cnt] : parts)
        if (cnt) HIPCHECK(hipMemsetAsync(b->p, 0, sizeof(double) * cnt, st));
    A->fvmActive = true;
    A->fvmHasLower = false;
    A->fvmHasFfc = false;
    A->set = false;
    A->gamgReady = false;
    A->mixedReady = false;
}

// device owner/neighbour of the internal faces, uploaded on first need
void ensureOwnNei(vol_mesh m) {
    if (!m->dOwn.p && m->nf) {
        m->dOwn.upload(m->own, m->ctx->stream);
        m->dNei.upload(m->nei, m->ctx->stream);
    }
}

void fvmTerm(vol_matrix A, int kind, FvmArgs& a, bool termLower) {
    vol_mesh m = A->mesh;
    hipStream_t st = m->ctx->stream;
    const bool newLower = A->fvmHasLower || termLower;
    a.lower = newLower ? A->lower.p : nullptr;
    const int* own = nullptr;
    const int* nei = nullptr;
    if (kind == FVM_LAP && a.psi) {
        ensureOwnNei(m);
        own = m->dOwn.p;
        nei = m->dNei.p;
        A->fgrad.alloc(3 * size_t(m->n));
        A->ffcT.alloc(m->nf);
        A->anyCorr.alloc(1);
        a.grad = A->fgrad.p;
        a.fc = A->ffcT.p;
        a.anyCorr = A->anyCorr.p;
    }
    launch_fvm_term(st, kind, a, own, nei);
    A->fvmHasLower = newLower;
    if (kind == FVM_LAP && a.psi) {
        // the matrix holds a face-flux correction from here on (FvMatrix.h:50-53,240-244)
        int any = 0;
        HIPCHECK(hipMemcpyAsync(&any, A->anyCorr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (any) A->fvmHasFfc = true;
    }
    HIPCHECK(hipStreamSynchronize(st));
}

// lower part of the FvMatrix parts: upper when the matrix has none (FvMatrix.h:44)
const double* fvmLower(vol_matrix A) { return A->fvmHasLower ? A->lower.p : A->upper.p; }

// Result buffer: the caller's on the device, else the staging buffer copied back by fvmOut.
double* fvmOutBuf(vol_matrix A, double* out, size_t cnt, int location) {
    if (location == VOL_DEVICE) return out;
    A->fout.alloc(cnt);
    return A->fout.p;
}
void fvmOut(vol_matrix A, double* out, size_t cnt, int location) {
    hipStream_t st = A->mesh->ctx->stream;
    if (location != VOL_DEVICE && cnt)
        HIPCHECK(hipMemcpyAsync(out, A->fout.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

void fvmFinish(vol_matrix A, double* b, const double* su, const double* V, int location) {
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    hipStream_t st = m->ctx->stream;
    const int n = m->n;
    A->diag.alloc(n);
    FvmArgs a = fvmArgs(A, 1);
    DevBuf<double> tb;
    double* bd = b;
    if (b && location != VOL_DEVICE) {
        tb.alloc(n);
        bd = tb.p;
    }
    launch_fvm_fold(st, a, A->diag.p, bd, su ? stageK(A, 0, su, n, location) : nullptr,
                    su ? stageK(A, 1, V, n, location) : nullptr);
    A->symmetric = !A->fvmHasLower;
    matrixFinish(A);
    if (b && location != VOL_DEVICE) HIPCHECK(hipMemcpyAsync(b, tb.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------ solve
struct Plan {
    std::string solver, precond;
    bool exact = false;
    bool gamgGS = false;  // Gauss–Seidel GAMG smoother
    int nPre = 2, nPost = 2;
    double omega = 0.8;
};

struct Ctx {
    vol_context c;
    hipStream_t st;
    const int* done;
    DevScalars* S;
    const RedScratch* red = nullptr;  // reduction scratch: the context's unless set
    bool noTiming = false;            // no event recording (no event objects created)
    // set by runLoop: the host-visible slot the iteration's last finalize copies the
    // scalars into (so that no separate copy is needed to watch the loop)
    DevScalars* hostOut = nullptr;
};

const RedScratch& redOf(const Ctx& C) { return C.red ? *C.red : C.c->red; }

void launchSell(hipStream_t st, int op, int xl, const SellArgs& a) { launch_sell(st, op, xl, a); }
void launchSell(hipStream_t st, int op, int xl, const SellArgs32& a) { launch_sell32(st, op, xl, a); }

// Where a level's final x goes besides its output: added to the finer level's vector z
// over the aggregates' member lists (the pushed prolongation).
template <class T> struct PushTarget {
    T* z = nullptr;
    const int* mOff = nullptr;
    const int* mList = nullptr;
};

template <class T>
void sell(const Ctx& C, int op, int xl, const SellView& V, const T* coef, const T* diag, const T* x, const T* r,
          T* out, double omega = 1.0, const T* xc = nullptr, const int* agg = nullptr, const T* dotv = nullptr,
          double* part = nullptr, const T* xu = nullptr, const T* rD = nullptr,
          const PushTarget<T>* push = nullptr) {
    SellArgsT<T> a;
    if (push) {
        a.pushZ = push->z;
        a.pushOff = push->mOff;
        a.pushList = push->mList;
    }
    a.S = V;
    a.coef = coef;
    a.diag = diag;
    a.rD = rD;
    a.x = x;
    a.r = r;
    a.xc = xc;
    a.agg = agg;
    a.omega = omega;
    a.out = out;
    a.dotv = dotv;
    a.part = part;
    a.done = C.done;
    a.xuPtr = xu;
    launchSell(C.st, op, xl, a);
}

template <class T>
void vec(const Ctx& C, int op, int n, bool exact, const T* a, const T* b = nullptr, const T* c = nullptr,
         const T* d = nullptr, T* w = nullptr, T* x = nullptr, T* y = nullptr, const T* rD = nullptr,
         double omega = 1.0) {
    VecArgsT<T> v;
    v.omega = omega;
    v.n = n;
    v.a = a;
    v.b = b;
    v.c = c;
    v.d = d;
    v.w = w;
    v.x = x;
    v.y = y;
    v.rD = rD;
    v.S = C.S;
    v.done = C.done;
    if constexpr (std::is_same_v<T, double>)
        launch_vec(C.st, op, v, exact, redOf(C));
    else
        launch_vec32(C.st, op, v, redOf(C));
}

void fin(const Ctx& C, int op, int nred, bool exact, int nparts, DevScalars* hostOut = nullptr) {
    launch_finalize(C.st, op, nred, exact, nparts, redOf(C), C.S, hostOut);
}

// Level-0 matrix in the solve's working precision.
template <class T> struct Level0 {
    const T* coef;
    const T* diag;
    const T* rD;
};
Level0<double> level0(vol_matrix A, double*) { return {A->coef.p, A->diag.p, A->rD.p}; }
Level0<float> level0(vol_matrix A, float*) { return {A->coef32.p, A->diag32.p, A->rD32.p}; }

// Level-0 A·p with optional fused dot (fast mode), timed when timing is on. With
// finOp and the context's lastBlockFin, the launch's last block also runs that
// finalize step; returns whether it does (then the caller launches no finalize).
template <class T>
bool amulTimed(const Ctx& C, vol_matrix A, const T* p, T* q, const T* dotv, bool fused, int finOp = 0) {
    vol_context c = C.c;
    const SellView V = A->mesh->sell.view();
    const Level0<T> L0 = level0(A, (T*)nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing && !C.noTiming) {
        if (c->evUsed + 2 > c->evPool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                HIPCHECK(hipEventCreate(&e));
                c->evPool.push_back(e);
            }
        }
        e0 = c->evPool[c->evUsed++];
        e1 = c->evPool[c->evUsed++];
    }
    SellArgsT<T> a;  // the events bracket the kernel itself (recorded by the launch)
    a.S = V;
    a.coef = L0.coef;
    a.diag = L0.diag;
    a.x = p;
    a.out = q;
    a.dotv = fused ? dotv : nullptr;
    a.part = fused ? redOf(C).part : nullptr;
    a.done = C.done;
    a.evStart = e0;
    a.evStop = e1;
    const bool last = fused && finOp != 0 && c->lastBlockFin && redOf(C).arrive != nullptr;
    if (last) {
        a.arrive = redOf(C).arrive;
        a.finS = C.S;
        a.finOp = finOp;
    }
    launchSell(C.st, OP_AMUL, XL_PLAIN, a);
    return last;
}

void collectTiming(vol_context c) {
    if (!c->timing) return;
    for (size_t k = 0; k + 1 < c->evUsed; k += 2) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, c->evPool[k], c->evPool[k + 1]));
        c->amulMs += ms;
        ++c->amulCount;
    }
    c->evUsed = 0;
}

// Per-level data of the V-cycle in precision T.
template <class T> struct LevelPtrs {
    const T *coef, *diag, *rD;
    T *bufA, *bufB, *r, *x, *z;
};
LevelPtrs<double> levelPtrs(vol_matrix A, int l, double*) {
    MatLevel& M = *A->L[l];
    if (l == 0) return {A->coef.p, A->diag.p, A->rD.p, M.bufA.p, M.bufB.p, nullptr, nullptr, nullptr};
    return {M.coef.p, M.diag.p, M.rD.p, M.bufA.p, M.bufB.p, M.r.p, M.x.p, M.z.p};
}
LevelPtrs<float> levelPtrs(vol_matrix A, int l, float*) {
    MatLevel& M = *A->L[l];
    if (l == 0) return {A->coef32.p, A->diag32.p, A->rD32.p, M.bufA32.p, M.bufB32.p, nullptr, nullptr, nullptr};
    return {M.coef32.p, M.diag32.p, M.rD32.p, M.bufA32.p, M.bufB32.p, M.r32.p, M.x32.p, M.z32.p};
}
const double* restrictCoef(const MatLevel* M, double*) { return M->coefR.p; }
const float* restrictCoef(const MatLevel* M, float*) { return M->coefR32.p; }
void tailLaunch(const Ctx& C, vol_matrix A, const Plan& P, int l, const double* r, double* xout) {
    TailArgs t;
    t.nLevels = int(A->mesh->lv.size()) - l;
    t.L = A->tail.p;
    t.rTop = r;
    t.xTop = xout;
    t.nCoarse = A->nCoarse;
    t.lu = A->lu.p;
    t.luT = A->luT.p;
    t.ldsBytes = A->tailLdsBytes;
    t.nPre = P.nPre;
    t.nPost = P.nPost;
    t.omega = P.omega;
    t.done = C.done;
    t.chunks = A->tailChunks.p;
    t.nChunks = A->tailNChunks;
    t.rInChunks = l > 0 && r == A->L[l]->r.p ? 1 : 0;
    launch_vcycle_tail(C.st, t);
}
void tailLaunch(const Ctx& C, vol_matrix A, const Plan& P, int l, const float* r, float* xout) {
    TailArgs32 t;
    const int nL = int(A->mesh->lv.size());
    t.nLevels = nL - l;
    t.L = A->tail32.p;
    t.rTop = r;
    t.xTop = xout;
    t.nCoarse = A->nCoarse;
    t.lu = A->lu.p;
    t.luT = A->luT.p;
    t.ldsBytes = A->tailLdsBytes32;
    t.nPre = P.nPre;
    t.nPost = P.nPost;
    t.omega = P.omega;
    t.done = C.done;
    t.chunks = A->tailChunks32.p;
    t.nChunks = A->tailNChunks32;
    t.rInChunks = l > 0 && r == A->L[l]->r32.p ? 1 : 0;
    launch_vcycle_tail32(C.st, t);
}
void coarseSolve(const Ctx& C, vol_matrix A, const double* r, double* x) {
    launch_coarse_solve(C.st, A->nCoarse, A->lu.p, A->luT.p, r, x, C.done);
}
void coarseSolve(const Ctx& C, vol_matrix A, const float* r, float* x) {
    launch_coarse_solve32(C.st, A->nCoarse, A->lu.p, A->luT.p, r, x, C.done);
}
// rc = restriction of res; with zc: also zc = ω·(rc·rDc), the first Jacobi sweep from
// zero of the coarse level, so that its sweeps gather one vector instead of two
void restrictLevel(const Ctx& C, int nc, const LevelDev& D, const double* res, double* rc, double* zc = nullptr,
                   const double* rDc = nullptr, double omega = 1.0) {
    launch_restrict(C.st, nc, D.mOff.p, D.mList.p, res, rc, C.done, zc, rDc, omega);
}
void restrictLevel(const Ctx& C, int nc, const LevelDev& D, const float* res, float* rc, float* zc = nullptr,
                   const float* rDc = nullptr, double omega = 1.0) {
    launch_restrict32(C.st, nc, D.mOff.p, D.mList.p, res, rc, C.done, zc, rDc, omega);
}

// Does level l of the Jacobi V-cycle get its z = ω·(r·rD) from the restriction? (a
// level smoothed by grid launches, below level 0, with one pre-sweep)
bool levelUsesZ(vol_matrix A, const Plan& P, int l) {
    const int nL = int(A->mesh->lv.size());
    return P.nPre == 1 && l >= 1 && l < nL - 1 && l != A->tailStart && A->mesh->ctx->coarseZ;
}

// One V-cycle at level l: r → x (GAMG.cpp:403-443 structure, damped-Jacobi smoothing).
// With part != nullptr (level 0 outside the tail), the last post-sweep also leaves the
// per-block partials of Σ x·dotv in part. With push, the launch that produces the final
// x also adds it to the finer level's vector over the aggregates (returns true; false
// when it was not pushed and the finer level has to gather x itself).
template <class T>
bool vcycle(const Ctx& C, vol_matrix A, const Plan& P, int l, const T* r, T* xout, const T* dotv = nullptr,
            double* part = nullptr, T* z0 = nullptr, const PushTarget<T>* push = nullptr) {
    vol_mesh m = A->mesh;
    const int nL = int(m->lv.size());
    if (l == A->tailStart) {
        tailLaunch(C, A, P, l, r, xout);
        return false;
    }
    if (l == nL - 1) {
        coarseSolve(C, A, r, xout);
        return false;
    }
    const LevelDev& D = *m->lv[l];
    const SellView V = l == 0 ? m->sell.view() : D.sell.view();
    const LevelPtrs<T> L = levelPtrs(A, l, (T*)nullptr);
    const bool useZ = levelUsesZ(A, P, l);  // then L.z holds the first sweep from zero
    if (useZ && l + 3 == nL && m->ctx->last2 && P.nPost == 1 && A->tailStart < 0 && A->nCoarse <= 64) {
        // this level, the next and the coarsest in one workgroup
        const LevelDev& DB = *m->lv[l + 1];
        constexpr int W = LAST2_BATCHES * 2 * LAST2_IT;
        if (D.layout == 0 && DB.layout == 0 && D.n <= LAST2_ROWS && DB.n <= LAST2_ROWS && D.sell.maxWidth <= W &&
            DB.sell.maxWidth <= W) {
            const LevelPtrs<T> LB = levelPtrs(A, l + 1, (T*)nullptr);
            Last2ArgsT<T> a;
            a.nA = D.n;
            a.nB = DB.n;
            a.nC = A->nCoarse;
            a.batchesA = std::max(1, (D.sell.maxWidth + 2 * LAST2_IT - 1) / (2 * LAST2_IT));
            a.batchesB = std::max(1, (DB.sell.maxWidth + 2 * LAST2_IT - 1) / (2 * LAST2_IT));
            a.offA = D.sell.sliceOff.p;
            a.colA = D.sell.col.p;
            a.aggA = D.agg.p;
            a.mOffA = D.mOff.p;
            a.mListA = D.mList.p;
            a.coefA = L.coef;
            a.diagA = L.diag;
            a.rA = r;
            a.zA = L.z;
            a.offB = DB.sell.sliceOff.p;
            a.colB = DB.sell.col.p;
            a.aggB = DB.agg.p;
            a.mOffB = DB.mOff.p;
            a.mListB = DB.mList.p;
            a.coefB = LB.coef;
            a.diagB = LB.diag;
            a.rDB = LB.rD;
            a.luT = A->luT.p;
            a.omega = P.omega;
            a.xA = push ? nullptr : xout;  // pushed: nothing reads this level's x
            if (push) {
                a.pushZ = push->z;
                a.pushOff = push->mOff;
                a.pushList = push->mList;
            }
            a.done = C.done;
            launch_vcycle_last2<T>(C.st, a);
            return push != nullptr;
        }
    }
    T* const z = useZ ? L.z : (l == 0 && P.nPre == 1 ? z0 : nullptr);  // given: ω·(r·rD)
    T* bufs[2] = {L.bufA, L.bufB};
    int cur = 0;  // index of the buffer holding the pre-smoothed x (nPre ≥ 2)
    for (int s = 1; s < P.nPre; ++s) {
        const int dst = (s == 1) ? 0 : 1 - cur;
        sell<T>(C, OP_JACOBI, s == 1 ? XL_FROMZERO : XL_PLAIN, V, L.coef, L.diag, s == 1 ? nullptr : bufs[cur], r,
                bufs[dst], P.omega, nullptr, nullptr, nullptr, nullptr, nullptr, L.rD);
        cur = dst;
    }
    const int freeB = 1 - cur;
    const LevelPtrs<T> LC = levelPtrs(A, l + 1, (T*)nullptr);
    const bool toCoarsest = l + 1 == nL - 1 && A->nCoarse <= 64 && A->tailStart != l + 1;
    bool postDone = false;  // the post-sweep ran in the coarsest launch
    // the vector the first post-sweep gathers with the prolongation (x + xc[agg]): the
    // coarser level may push xc into it instead
    T* const prolongX = z ? z : (P.nPre == 1 ? nullptr : bufs[cur]);
    PushTarget<T> down;
    if (m->ctx->pushProlong && prolongX && !(l == 0 && m->ctx->prolongSplit)) {
        down.z = prolongX;
        down.mOff = D.mOff.p;
        down.mList = D.mList.p;
    }
    bool pushed = false;  // xc already added to prolongX
    if (m->ctx->fusedRestrict && D.rrBlocks > 0 &&
        (!toCoarsest || (l > 0 && D.rrRows.p && D.n <= RR_COARSE_ROWS && V.maxWidth <= 64))) {
        // residual and restriction in one launch (no residual vector); into the coarsest
        // level also its substitution
        SellArgsT<T> a;
        // its own rows (CSR, packed level 0, SELL read directly) or the restriction-order copy
        const bool own = D.rrCsr.nslots == 0;
        a.S = own ? V : D.rrCsr.view();
        a.coef = own ? L.coef : restrictCoef(A->L[l].get(), (T*)nullptr);
        a.diag = L.diag;
        a.rD = L.rD;
        a.x = z ? z : (P.nPre == 1 ? nullptr : bufs[cur]);
        a.r = r;
        a.omega = P.omega;
        a.done = C.done;
        const int xl = (z || P.nPre != 1) ? XL_PLAIN : XL_FROMZERO;
        RestrictArgsT<T> R;
        R.nc = m->lv[l + 1]->n;
        R.blk = D.rrBlk.p;
        R.nBlocks = D.rrBlocks;
        R.rows = D.rrRows.p;
        R.mOff = D.mOff.p;
        R.mList = D.mList.p;
        R.omega = P.omega;
        if (toCoarsest) {
            R.rc = LC.x;
            R.luT = A->luT.p;
            // one post-sweep from z above level 0 (no fused dot): in the same launch
            postDone = z && l > 0 && P.nPost == 1 && V.n <= RR_COARSE_THREADS / rrCoarseLanes(a.S.maxWidth);
            if (postDone) {
                R.post = 1;
                a.agg = D.agg.p;
                a.out = push ? nullptr : xout;  // pushed: nothing reads this level's x
                if (push) {
                    a.pushZ = push->z;
                    a.pushOff = push->mOff;
                    a.pushList = push->mList;
                }
            }
            launch_resid_restrict_coarse64<T>(C.st, xl, a.S.csrG, a.S.csrIt, a, R);
        } else {
            R.rc = LC.r;
            if (levelUsesZ(A, P, l + 1)) {
                R.zc = LC.z;
                R.rDc = LC.rD;
            }
            launch_resid_restrict<T>(C.st, xl, a.S.csrG, a.S.csrIt, a, R);
            pushed = vcycle<T>(C, A, P, l + 1, LC.r, LC.x, nullptr, nullptr, nullptr, down.z ? &down : nullptr);
        }
    } else {
        // residual of the pre-smoothed x, then restriction
        if (z)
            sell<T>(C, OP_RESID, XL_PLAIN, V, L.coef, L.diag, z, r, bufs[freeB], P.omega);
        else
            sell<T>(C, OP_RESID, P.nPre == 1 ? XL_FROMZERO : XL_PLAIN, V, L.coef, L.diag,
                    P.nPre == 1 ? nullptr : bufs[cur], r, bufs[freeB], P.omega, nullptr, nullptr, nullptr, nullptr,
                    nullptr, L.rD);
        if (toCoarsest) {
            // the coarsest level: restriction and substitution in one launch
            launch_restrict_coarse64<T>(C.st, A->nCoarse, D.mOff.p, D.mList.p, bufs[freeB], A->luT.p, LC.x, C.done);
        } else {
            if (levelUsesZ(A, P, l + 1))
                restrictLevel(C, m->lv[l + 1]->n, D, bufs[freeB], LC.r, LC.z, LC.rD, P.omega);
            else
                restrictLevel(C, m->lv[l + 1]->n, D, bufs[freeB], LC.r);
            pushed = vcycle<T>(C, A, P, l + 1, LC.r, LC.x, nullptr, nullptr, nullptr, down.z ? &down : nullptr);
        }
    }
    if (postDone) return push != nullptr;
    // prolongation fused into the first post-sweep (or already pushed into prolongX)
    int src = cur, other = freeB;
    for (int s = 0; s < P.nPost; ++s) {
        const bool last = s == P.nPost - 1;
        const PushTarget<T>* pu = last ? push : nullptr;
        T* dst = last ? (pu ? nullptr : xout) : bufs[other];  // pushed: nothing reads this level's x
        const T* dv = last ? dotv : nullptr;
        double* pt = last ? part : nullptr;
        if (s == 0 && pushed) {
            sell<T>(C, OP_JACOBI, XL_PLAIN, V, L.coef, L.diag, prolongX, r, dst, P.omega, nullptr, nullptr, dv, pt,
                    nullptr, nullptr, pu);
        } else if (s == 0 && z && l == 0 && m->ctx->prolongSplit) {
            // z += xc[agg] streamed first, so that the sweep gathers one vector
            launch_prolong<T>(C.st, m->n, D.agg.p, LC.x, z, C.done);
            sell<T>(C, OP_JACOBI, XL_PLAIN, V, L.coef, L.diag, z, r, dst, P.omega, nullptr, nullptr, dv, pt);
        } else if (s == 0 && z)
            sell<T>(C, OP_JACOBI, XL_PROLONG, V, L.coef, L.diag, z, r, dst, P.omega, LC.x, D.agg.p, dv, pt, nullptr,
                    nullptr, pu);
        else if (s == 0)
            sell<T>(C, OP_JACOBI, P.nPre == 1 ? XL_ZEROPROLONG : XL_PROLONG, V, L.coef, L.diag,
                    P.nPre == 1 ? nullptr : bufs[src], r, dst, P.omega, LC.x, D.agg.p, dv, pt, nullptr, L.rD, pu);
        else
            sell<T>(C, OP_JACOBI, XL_PLAIN, V, L.coef, L.diag, bufs[src], r, dst, P.omega, nullptr, nullptr, dv, pt,
                    nullptr, nullptr, pu);
        if (s != P.nPost - 1) {
            const int t = src;
            src = other;
            other = t;
        }
    }
    return push != nullptr;
}

// After a solve with sync-free sweeps (stream synchronised): a wait that ran past its
// limit voids the results.
void checkSweepFault(vol_context c, hipStream_t st) {
    int f = 0;
    HIPCHECK(hipMemcpyAsync(&f, c->sweepFault.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (f) {
        HIPCHECK(hipMemsetAsync(c->sweepFault.p, 0, sizeof(int), st));
        HIPCHECK(hipStreamSynchronize(st));
        throw VolError(VOL_ERR_HIP, "sync-free sweep: a wait for a neighbour's value timed out");
    }
}

// `sweeps` sync-free Gauss–Seidel sweeps of x (in place from the caller's view):
// alternately x → tmp and tmp → x, and a copy back after an odd count.
template <class T>
void gsSweepsFree(const Ctx& C, const LevelEll& E, bool backward, const T* coef, const T* diag, const T* b, T* x, T* tmp,
                  int n, int sweeps) {
    T* src = x;
    T* dst = tmp;
    for (int s = 0; s < sweeps; ++s) {
        launch_gs_sweep_free<T>(C.st, E, backward, coef, diag, b, src, dst, C.c->sweepFault.p, C.done);
        std::swap(src, dst);
    }
    if (src != x) vec<T>(C, V_COPY, n, false, src, nullptr, nullptr, nullptr, x);
}

// One V-cycle at level l with the reference's Gauss–Seidel smoother (GAMG.cpp:403-443):
// x = 0, nPre forward sweeps, residual r − A·x, restriction, recursion, x += xc[agg],
// nPost backward sweeps. x is xout itself.
template <class T> void vcycleGS(const Ctx& C, vol_matrix A, const Plan& P, int l, const T* r, T* xout) {
    vol_mesh m = A->mesh;
    const int nL = int(m->lv.size());
    if (l == nL - 1) {
        launch_coarse_solve_rows<T>(C.st, A->nCoarse, A->lu.p, r, xout, C.done);
        return;
    }
    const LevelDev& D = *m->lv[l];
    const SellView V = l == 0 ? m->sell.view() : D.sell.view();
    const LevelPtrs<T> L = levelPtrs(A, l, (T*)nullptr);
    const int n = l == 0 ? m->n : D.n;
    const LevelEll sF = l == 0 ? m->gsF.view() : m->gsLev[l]->F.view();
    const LevelEll sB = l == 0 ? m->gsB.view() : m->gsLev[l]->B.view();
    const T *cF, *cB;
    if constexpr (std::is_same_v<T, float>) {
        cF = l == 0 ? A->gsCF32.p : A->L[l]->gsCF32.p;
        cB = l == 0 ? A->gsCB32.p : A->L[l]->gsCB32.p;
    } else {
        cF = l == 0 ? A->gsCF.p : A->L[l]->gsCF.p;
        cB = l == 0 ? A->gsCB.p : A->L[l]->gsCB.p;
    }
    launch_fill_zero<T>(C.st, n, xout, C.done);
    if (C.c->freeSweeps) {  // sync-free sweeps alternating between xout and bufB
        gsSweepsFree<T>(C, sF, false, cF, L.diag, r, xout, L.bufB, n, P.nPre);
    } else if constexpr (std::is_same_v<T, double>) {  // the sweeps as one pipelined schedule
        if (P.nPre > 1) {
            launch_gs_sweep<T>(C.st, gsMultiEll(m, l, false, P.nPre).view(), gsMultiCoef(A, l, false, P.nPre),
                               L.diag, r, xout, C.done);
        } else {
            for (int s = 0; s < P.nPre; ++s) launch_gs_sweep<T>(C.st, sF, cF, L.diag, r, xout, C.done);
        }
    } else {
        for (int s = 0; s < P.nPre; ++s) launch_gs_sweep<T>(C.st, sF, cF, L.diag, r, xout, C.done);
    }
    sell<T>(C, OP_RESID, XL_PLAIN, V, L.coef, L.diag, xout, r, L.bufA, P.omega, nullptr, nullptr, nullptr, nullptr,
            nullptr, L.rD);
    const LevelPtrs<T> LC = levelPtrs(A, l + 1, (T*)nullptr);
    restrictLevel(C, m->lv[l + 1]->n, D, L.bufA, LC.r);
    vcycleGS<T>(C, A, P, l + 1, LC.r, LC.x);
    launch_prolong<T>(C.st, n, D.agg.p, LC.x, xout, C.done);
    if (C.c->freeSweeps) {
        gsSweepsFree<T>(C, sB, true, cB, L.diag, r, xout, L.bufB, n, P.nPost);
    } else if constexpr (std::is_same_v<T, double>) {
        if (P.nPost > 1) {
            launch_gs_sweep<T>(C.st, gsMultiEll(m, l, true, P.nPost).view(), gsMultiCoef(A, l, true, P.nPost),
                               L.diag, r, xout, C.done);
        } else {
            for (int s = 0; s < P.nPost; ++s) launch_gs_sweep<T>(C.st, sB, cB, L.diag, r, xout, C.done);
        }
    } else {
        for (int s = 0; s < P.nPost; ++s) launch_gs_sweep<T>(C.st, sB, cB, L.diag, r, xout, C.done);
    }
}

template <class T>
void precondApply(const Ctx& C, vol_matrix A, const Plan& P, const T* r, T* w, int n) {
    if (P.precond == "GAMG" && P.gamgGS) {
        vcycleGS<T>(C, A, P, 0, r, w);
    } else if (P.precond == "GAMG") {
        vcycle<T>(C, A, P, 0, r, w);
    } else if (P.precond == "DIC" || P.precond == "DILU") {
        vol_mesh m = A->mesh;
        const T *rD, *ef, *eb;
        if constexpr (std::is_same_v<T, float>) {
            rD = A->dluR32.p;
            ef = A->dluEf32.p;
            eb = A->dluEb32.p;
        } else {
            rD = A->dluR.p;
            ef = A->dluEf.p;
            eb = A->dluEb.p;
        }
        if (C.c->freeSweeps) {
            T* tmp;
            if constexpr (std::is_same_v<T, float>) {
                A->diluTmp32.alloc(n);
                tmp = A->diluTmp32.p;
            } else {
                A->diluTmp.alloc(n);
                tmp = A->diluTmp.p;
            }
            launch_dilu_apply_free<T>(C.st, m->diluF.view(), m->diluB.view(), rD, ef, eb, r, tmp, w,
                                      C.c->sweepFault.p, C.done);
        } else {
            launch_dilu_apply<T>(C.st, m->diluF.view(), m->diluB.view(), rD, ef, eb, r, w, C.done);
        }
    } else
        vec<T>(C, V_SCALE, n, false, r, nullptr, nullptr, nullptr, w, nullptr, nullptr,
               P.precond == "diagonal" ? level0(A, (T*)nullptr).rD : nullptr);
}

void readScalars(vol_context c, const DevScalars* S, DevScalars& out) {
    HIPCHECK(hipMemcpyAsync(c->hostScal, S, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    out = *c->hostScal;
}

bool hostConverged(const vol_solver_controls& ct, const DevScalars& S) {
    if (S.iterations < ct.minIter) return false;
    return S.finalResidual < ct.tolerance || (ct.relTol > 0 && S.finalResidual < ct.relTol * S.initialResidual);
}

// Vectors of one Krylov solve in precision T. x is the unknown being updated (the
// solution in FP64, the correction e in the inner solve of mixed precision); r is its
// residual, updated recursively.
template <class T> struct Krylov {
    T *x, *r, *w, *p, *q;            // PCG (q = A·p)
    T *r0, *y, *Ay, *s, *z, *t;      // PBiCGStab
    T* z0 = nullptr;                 // ω·(r·rD) kept with r (PCG + Jacobi GAMG, one pre-sweep)
};

// One PCG or PBiCGStab iteration, every kernel predicated on C.done.
template <class T>
void krylovIteration(const Ctx& C, vol_matrix A, const Plan& P, const Krylov<T>& k, bool isPCG, bool ex) {
    const int n = A->mesh->n;
    const bool fuse = !ex;
    const int nparts = ex ? RED_BLOCKS_EXACT : RED_BLOCKS;
    const int sellParts = sell_grid_blocks(A->mesh->sell.nslices);
    const Level0<T> L0 = level0(A, (T*)nullptr);
    if (isPCG) {  // LinearSolver.cpp:176-201
        if (P.precond == "GAMG" && !P.gamgGS && fuse && A->tailStart != 0) {  // Σ w·r fused into the last sweep
            vcycle<T>(C, A, P, 0, k.r, k.w, k.r, redOf(C).part, k.z0);
            fin(C, FIN_PCG_WRA, 1, false, sellParts);
        } else {
            if (P.precond == "GAMG" || P.precond == "DIC" || P.precond == "DILU") {
                precondApply<T>(C, A, P, k.r, k.w, n);
                vec<T>(C, V_DOT, n, ex, k.w, k.r);
            } else {
                vec<T>(C, V_PRECOND_DOT, n, ex, k.r, nullptr, nullptr, nullptr, k.w, nullptr, nullptr,
                       P.precond == "diagonal" ? L0.rD : nullptr);
            }
            fin(C, FIN_PCG_WRA, 1, ex, nparts);
        }
        vec<T>(C, V_PCG_P, n, ex, k.w, nullptr, nullptr, nullptr, nullptr, k.p);
        const bool finDone = amulTimed<T>(C, A, k.p, k.q, k.p, fuse, FIN_PCG_WAPA);
        if (fuse) {
            if (!finDone) fin(C, FIN_PCG_WAPA, 1, false, sellParts);
        } else {
            vec<T>(C, V_DOT, n, ex, k.q, k.p);
            fin(C, FIN_PCG_WAPA, 1, ex, nparts);
        }
        vec<T>(C, V_PCG_XR, n, ex, k.p, k.q, nullptr, nullptr, k.z0, k.x, k.r, k.z0 ? L0.rD : nullptr, P.omega);
        fin(C, FIN_PCG_NORM, 1, ex, ex ? nparts : vec_fast_blocks(V_PCG_XR, n, redOf(C).maxBlk), C.hostOut);
    } else {  // PBiCGStab, LinearSolver.cpp:202-237
        vec<T>(C, V_DOT, n, ex, k.r0, k.r);
        fin(C, FIN_BI_RHO, 1, ex, nparts);
        vec<T>(C, V_BI_P, n, ex, k.r, k.Ay, nullptr, nullptr, nullptr, k.p);
        precondApply<T>(C, A, P, k.p, k.y, n);
        const bool finDone = amulTimed<T>(C, A, k.y, k.Ay, k.r0, fuse, FIN_BI_ALPHA);
        if (fuse) {
            if (!finDone) fin(C, FIN_BI_ALPHA, 1, false, sellParts);
        } else {
            vec<T>(C, V_DOT, n, ex, k.r0, k.Ay);
            fin(C, FIN_BI_ALPHA, 1, ex, nparts);
        }
        vec<T>(C, V_BI_S, n, ex, k.r, k.Ay, nullptr, nullptr, k.s);
        fin(C, FIN_BI_S, 1, ex, nparts);
        vec<T>(C, V_BI_XEXIT, n, ex, k.y, nullptr, nullptr, nullptr, nullptr, k.x);
        launch_clear_exit(C.st, C.S);
        precondApply<T>(C, A, P, k.s, k.z, n);
        sell<T>(C, OP_AMUL, XL_PLAIN, A->mesh->sell.view(), L0.coef, L0.diag, k.z, nullptr, k.t);
        vec<T>(C, V_BI_DOT2, n, ex, k.t, k.s);
        fin(C, FIN_BI_OMEGA, 2, ex, nparts);
        vec<T>(C, V_BI_XR, n, ex, k.y, k.z, k.s, k.t, nullptr, k.x, k.r);
        fin(C, FIN_BI_NORM, 1, ex, nparts, C.hostOut);
    }
}

// Runs `iterate` until C.S->done, enqueuing one iteration ahead of the host: iteration
// k+1 is queued before the host reads the scalars after iteration k, and runs as
// no-ops if k finished. The iteration's last finalize writes the scalars into the
// pinned slot C.hostOut (two slots, alternating), so watching the loop costs an event
// per iteration and no copy.
template <class F> void runLoop(vol_context c, Ctx& C, int maxLoops, F&& iterate) {
    hipStream_t st = C.st;
    hipEvent_t* evs = c->loopEv;
    DevScalars* hsl = c->hostScal;
    C.hostOut = &hsl[0];
    iterate();
    HIPCHECK(hipEventRecord(evs[0], st));
    for (int k = 1; k <= maxLoops; ++k) {
        C.hostOut = &hsl[k & 1];
        iterate();
        HIPCHECK(hipEventRecord(evs[k & 1], st));
        HIPCHECK(hipEventSynchronize(evs[(k - 1) & 1]));
        if (hsl[(k - 1) & 1].done) break;
    }
    HIPCHECK(hipStreamSynchronize(st));
    C.hostOut = nullptr;
}

// Work buffers of one FP64 solve given by the caller instead of the matrix's and the
// context's own (the no-allocation entry point): vectors of n doubles, reduction scratch
// and the solve's two DevScalars. Solvers PCG/PBiCGStab with preconditioner none or
// diagonal only (the others keep per-matrix state).
struct Workspace {
    double *wA, *pA, *vr, *vw, *vp, *r0, *y, *Ay, *s, *z, *t;
    RedScratch red;
    DevScalars* scal;
};

void solveDevice(vol_matrix A, const double* b, double* x, const vol_solver_controls& ct, vol_solver_performance& perf,
                 const Workspace* ws = nullptr) {
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    const int n = m->n;
    Plan P;
    P.solver = ct.solver ? ct.solver : "PCG";
    P.precond = ct.preconditioner ? ct.preconditioner : "diagonal";
    int sums = ct.sums == VOL_SUMS_CONTEXT ? (c->reproducible ? VOL_SUMS_REPRODUCIBLE : VOL_SUMS_FAST) : ct.sums;
    if (sums < VOL_SUMS_FAST || sums > VOL_SUMS_EXACT) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad sums mode");
    if (sums == VOL_SUMS_REPRODUCIBLE) {  // LinearSolver.cpp:138-148
        if (P.precond != "none") P.precond = "diagonal";
        if (P.solver == "GaussSeidel" || P.solver == "symGaussSeidel") P.solver = "Jacobi";
    }
    P.exact = sums != VOL_SUMS_FAST;
    P.nPre = ct.gamgPreSweeps;
    P.nPost = ct.gamgPostSweeps;
    P.omega = ct.gamgOmega;
    if (ct.gamgSmoother != VOL_GAMG_JACOBI && ct.gamgSmoother != VOL_GAMG_GAUSS_SEIDEL)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "gamgSmoother must be VOL_GAMG_JACOBI or VOL_GAMG_GAUSS_SEIDEL");
    P.gamgGS = ct.gamgSmoother == VOL_GAMG_GAUSS_SEIDEL;
    if (ct.precision != VOL_FP64 && ct.precision != VOL_MIXED)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "precision must be VOL_FP64 or VOL_MIXED");
    const bool mixed = ct.precision == VOL_MIXED;
    std::memset(&perf, 0, sizeof(perf));
    std::strncpy(perf.solver, P.solver.c_str(), sizeof(perf.solver) - 1);

    const bool isPCG = P.solver == "PCG", isBi = P.solver == "PBiCGStab", isJac = P.solver == "Jacobi";
    const bool isGS = P.solver == "GaussSeidel" || P.solver == "symGaussSeidel";
    if (!isPCG && !isBi && !isJac && !isGS) throw VolError(VOL_ERR_UNKNOWN_SOLVER, "unknown linear solver " + P.solver);
    if (isPCG && !A->symmetric) throw VolError(VOL_ERR_PCG_NONSYMMETRIC, "PCG requires a symmetric matrix; use PBiCGStab");
    std::string pc = P.precond;
    if (isBi && pc == "DIC") pc = "DILU";  // LinearSolver.cpp:203
    if (isPCG || isBi) {
        if (pc != "none" && pc != "diagonal" && pc != "GAMG" && pc != "DIC" && pc != "DILU")
            throw VolError(VOL_ERR_UNKNOWN_PRECONDITIONER, "unknown preconditioner " + pc);
        if (pc == "GAMG" && !m->hasGamg)
            throw VolError(VOL_ERR_NO_GAMG_WEIGHTS, "GAMG needs a mesh created with agglomeration weights");
        if (pc == "GAMG" && (P.nPre < 1 || P.nPost < 1 || !(P.omega > 0.0)))
            throw VolError(VOL_ERR_INVALID_ARGUMENT, "GAMG smoother needs ≥ 1 pre/post sweep and ω > 0");
    }
    P.precond = pc;
    if (ct.maxIter < 0 || ct.nSweeps < 1) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad maxIter/nSweeps");
    if (mixed) {
        if (P.exact) throw VolError(VOL_ERR_UNSUPPORTED, "VOL_MIXED runs with fast sums only");
        if (isJac || isGS) throw VolError(VOL_ERR_UNSUPPORTED, "VOL_MIXED needs PCG or PBiCGStab");
        if (!(ct.innerRelTol > 0.0 && ct.innerRelTol < 1.0))
            throw VolError(VOL_ERR_INVALID_ARGUMENT, "innerRelTol must lie in (0, 1)");
    }

    if (ws) {
        const bool simple = P.precond == "none" || P.precond == "diagonal";
        if (mixed || !(isPCG || isBi) || !simple)
            throw VolError(VOL_ERR_UNSUPPORTED, "the no-allocation solve runs PCG or PBiCGStab with preconditioner "
                                                "none or diagonal in FP64");
    }
    setDevice(c);
    hipStream_t st = c->stream;
    DevScalars* scal = ws ? ws->scal : c->scal.p;
    Ctx C{c, st, &scal->done, scal};
    if (ws) {
        C.red = &ws->red;
        C.noTiming = true;
    } else {
        c->ensureScratch(sell_grid_blocks(m->sell.nslices));
        A->wA.alloc(n);
        A->pA.alloc(n);
        A->vr.alloc(n);
    }
    double* wA = ws ? ws->wA : A->wA.p;
    double* pA = ws ? ws->pA : A->pA.p;
    double* vr = ws ? ws->vr : A->vr.p;
    const bool ex = P.exact;
    const SellView V = m->sell.view();
    launch_init_scalars(st, C.S, ct.tolerance, ct.relTol, ct.maxIter, ct.minIter, ct.nSweeps, n,
                        mixed ? ct.innerRelTol : 0.0, mixed && isPCG ? 1 : 0);
    // initial residual and normalisation (LinearSolver.cpp:150-169)
    sell<double>(C, OP_AMUL, XL_PLAIN, V, A->coef.p, A->diag.p, x, nullptr, wA);
    const int nparts = ex ? RED_BLOCKS_EXACT : RED_BLOCKS;
    vec<double>(C, V_SUM_X, n, ex, x);
    fin(C, FIN_XREF, 1, ex, nparts);
    sell<double>(C, OP_AMUL, XL_UNIFORM, V, A->coef.p, A->diag.p, nullptr, nullptr, pA, 1.0, nullptr, nullptr,
                 nullptr, nullptr, &C.S->xRef);
    // FP64 PCG with the fused Jacobi V-cycle keeps z0 = ω·(r·rD) with r (prologue and
    // every x/r update), the level-0 first sweep from zero
    const bool useZ0 = isPCG && !mixed && P.precond == "GAMG" && !P.gamgGS && !ex && P.nPre == 1 && c->level0Z && !ws;
    if (useZ0) A->z0.alloc(n);
    vec<double>(C, V_INIT, n, ex, wA, b, pA, nullptr, vr, nullptr, useZ0 ? A->z0.p : nullptr,
                useZ0 ? A->rD.p : nullptr, P.omega);
    fin(C, FIN_INIT, 2, ex, nparts);
    DevScalars hs;
    readScalars(c, C.S, hs);
    if (!hs.done) {
        if (P.precond == "GAMG" && !A->gamgReady) buildGamgCoeffs(A);
        if (P.precond == "GAMG" && P.gamgGS) buildGsLevels(m);
        if ((P.precond == "GAMG" && P.gamgGS) || isGS) buildGs(A, P.precond == "GAMG" && P.gamgGS);
        const bool dilu = (isPCG || isBi) && (P.precond == "DIC" || P.precond == "DILU");
        if (dilu && !A->diluReady) buildDilu(A);
        if (dilu && mixed && !A->dilu32Ready) buildDilu32(A);
        if (isGS) buildSweep(m);
        const int maxLoops = std::max(ct.maxIter, 1) + 2;
        if (mixed) {
            if (!A->mixedReady || (P.precond == "GAMG" && !A->mixedGamg)) buildMixed(A, P.precond == "GAMG");
            if (P.precond == "GAMG" && P.gamgGS && !A->gs32Ready) buildGs32(A);
            for (DevBuf<float>* v : {&A->ir, &A->iw, &A->ip, &A->iq, &A->ie}) v->alloc(n);
            if (isBi)
                for (DevBuf<float>* v : {&A->ir0, &A->iy, &A->iAy, &A->is, &A->iz, &A->it}) v->alloc(n);
            Krylov<float> k{A->ie.p, A->ir.p, A->iw.p, A->ip.p, A->iq.p, A->ir0.p, A->iy.p, A->iAy.p, A->is.p, A->iz.p,
                            A->it.p};
            DevScalars* SI = c->scal.p + 1;
            Ctx Ci{c, st, &SI->done, SI};
            // the inner PCG with the fused Jacobi V-cycle keeps z0 = ω·(r·rD) with r, as
            // the FP64 one does (set from each inner right-hand side, then by x/r updates)
            const bool useZ0f = isPCG && P.precond == "GAMG" && !P.gamgGS && P.nPre == 1 && c->level0Z;
            if (useZ0f) {
                A->iz0.alloc(n);
                k.z0 = A->iz0.p;
            }
            auto innerRhs = [&] {  // inner right-hand side: the FP64 residual rounded; e = 0
                fin(C, FIN_INNER_START, 0, false, 0);
                launch_mix(st, MIX_TO32, n, A->vr.p, A->ir.p, isBi ? A->ir0.p : nullptr, A->ie.p, nullptr, nullptr,
                           C.done);
                if (useZ0f)
                    vec<float>(C, V_ZSCALE, n, false, A->ir.p, nullptr, nullptr, nullptr, A->iz0.p, nullptr, nullptr,
                               A->rD32.p, P.omega);
            };
            innerRhs();
            for (int outer = 0; outer < maxLoops; ++outer) {
                runLoop(c, Ci, maxLoops, [&] { krylovIteration<float>(Ci, A, P, k, isPCG, false); });
                // x += e; r = b − A·x in FP64; convergence; next inner solve
                launch_mix(st, MIX_ADD, n, nullptr, nullptr, nullptr, nullptr, x, A->ie.p, C.done);
                sell<double>(C, OP_AMUL, XL_PLAIN, V, A->coef.p, A->diag.p, x, nullptr, A->wA.p);
                vec<double>(C, V_BMINUSAX_NORM, n, false, A->wA.p, b, nullptr, nullptr, A->vr.p);
                fin(C, FIN_OUTER, 1, false, RED_BLOCKS);
                innerRhs();
                readScalars(c, C.S, hs);
                if (hs.done) break;
            }
        } else if (isPCG || isBi) {
            if (!ws) {
                A->vw.alloc(n);
                A->vp.alloc(n);
            }
            Krylov<double> k{x, vr, ws ? ws->vw : A->vw.p, ws ? ws->vp : A->vp.p, wA, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr};
            if (useZ0 && isPCG) k.z0 = A->z0.p;
            if (isBi) {
                if (ws) {
                    k.r0 = ws->r0;
                    k.y = ws->y;
                    k.Ay = ws->Ay;
                    k.s = ws->s;
                    k.z = ws->z;
                    k.t = ws->t;
                } else {
                    for (DevBuf<double>* v : {&A->r0, &A->y, &A->Ay, &A->s, &A->z, &A->t}) v->alloc(n);
                    k.r0 = A->r0.p;
                    k.y = A->y.p;
                    k.Ay = A->Ay.p;
                    k.s = A->s.p;
                    k.z = A->z.p;
                    k.t = A->t.p;
                }
                // r0 = r, p = 0, Ay = 0 by kernels (no memset/copy nodes)
                vec<double>(C, V_COPY, n, false, vr, nullptr, nullptr, nullptr, k.r0);
                launch_fill_zero<double>(st, n, k.p, nullptr);
                launch_fill_zero<double>(st, n, k.Ay, nullptr);
            } else if (P.precond != "GAMG") {
                launch_fill_zero<double>(st, n, k.p, nullptr);
            }
            runLoop(c, C, maxLoops, [&] { krylovIteration<double>(C, A, P, k, isPCG, ex); });
        } else {  // Jacobi and Gauss–Seidel solvers, LinearSolver.cpp:238-278
            A->vw.alloc(n);
            const bool sym = P.solver == "symGaussSeidel";
            runLoop(c, C, maxLoops, [&] {
                for (int sw = 0; sw < ct.nSweeps; ++sw) {
                    if (isGS && c->freeSweeps) {  // x → vw (forward), then vw → x (backward) or a copy
                        int* f = c->sweepFault.p;
                        launch_gs_sweep_free<double>(st, m->gsF.view(), false, A->gsCF.p, A->diag.p, b, x, A->vw.p, f,
                                                     C.done);
                        if (sym)
                            launch_gs_sweep_free<double>(st, m->gsB.view(), true, A->gsCB.p, A->diag.p, b, A->vw.p, x,
                                                         f, C.done);
                        else
                            vec<double>(C, V_COPY, n, false, A->vw.p, nullptr, nullptr, nullptr, x);
                        continue;
                    }
                    if (isGS) {
                        launch_gs_sweep<double>(st, m->gsF.view(), A->gsCF.p, A->diag.p, b, x, C.done);
                        if (sym) launch_gs_sweep<double>(st, m->gsB.view(), A->gsCB.p, A->diag.p, b, x, C.done);
                        continue;
                    }
                    sell<double>(C, OP_JACREF, XL_PLAIN, V, A->coef.p, A->diag.p, x, b, A->vw.p);
                    vec<double>(C, V_COPY, n, ex, A->vw.p, nullptr, nullptr, nullptr, x);
                }
                sell<double>(C, OP_AMUL, XL_PLAIN, V, A->coef.p, A->diag.p, x, nullptr, A->wA.p);
                vec<double>(C, V_BMINUSAX_NORM, n, ex, A->wA.p, b, nullptr, nullptr, A->vr.p);
                fin(C, FIN_JAC_NORM, 1, ex, nparts, C.hostOut);
            });
        }
        readScalars(c, C.S, hs);
        if (!hs.done) throw VolError(VOL_ERR_HIP, "solver loop did not terminate");
    }
    if (c->freeSweeps && (isGS || P.precond == "DIC" || P.precond == "DILU" || P.gamgGS)) checkSweepFault(c, st);
    collectTiming(c);
    perf.initialResidual = hs.initialResidual;
    perf.finalResidual = hs.finalResidual;
    perf.residualSum = hs.residualSum;
    perf.normFactor = hs.normFactor;
    perf.iterations = hs.iterations;
    perf.converged = hostConverged(ct, hs) ? 1 : 0;
    perf.outerIterations = hs.outer;
}

}  // namespace

// Handles keep their parents alive (matrix → mesh → context), so callers may destroy
// them in any order; the last release frees.
// Copy of mesh src for context dst: same addressing, hierarchy and boundary lists,
// device structures built for dst (the host hierarchy is reused, not recomputed).
vol_mesh meshCloneTo(vol_mesh src, vol_context dst) {
    setDevice(dst);
    auto m = std::make_unique<vol_mesh_s>();
    m->ctx = dst;
    m->n = src->n;
    m->nf = src->nf;
    m->own = src->own;
    m->nei = src->nei;
    setupMesh(m.get(), nullptr, 2);
    if (src->hasGamg) {
        m->hasGamg = true;
        m->H = src->H;
        const std::vector<int> layout = meshView(src).layout;
        uploadLevels(m.get(), &layout);
    }
    if (src->nb >= 0) {
        std::vector<int> off(src->n + 1), list(src->nb);
        HIPCHECK(hipMemcpy(off.data(), src->bOff.p, sizeof(int) * off.size(), hipMemcpyDeviceToHost));
        if (src->nb) HIPCHECK(hipMemcpy(list.data(), src->bList.p, sizeof(int) * list.size(), hipMemcpyDeviceToHost));
        m->bOff.upload(off, dst->stream);
        m->bList.upload(list, dst->stream);
        m->nb = src->nb;
    }
    HIPCHECK(hipStreamSynchronize(dst->stream));
    dst->refs.fetch_add(1);
    return m.release();
}

void ctxRelease(vol_context c) {
    if (c->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto e : c->evPool) (void)hipEventDestroy(e);
    for (auto e : c->loopEv)
        if (e) (void)hipEventDestroy(e);
    if (c->hostScal) (void)hipHostFree(c->hostScal);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

MeshView meshView(vol_mesh m) {
    MeshView v;
    v.n = m->n;
    v.nf = m->nf;
    v.own = m->own.data();
    v.nei = m->nei.data();
    v.H = m->hasGamg ? &m->H : nullptr;
    for (const auto& D : m->lv) v.layout.push_back(D->layout);
    return v;
}

void meshRelease(vol_mesh m) {
    if (m->refs.fetch_sub(1) != 1) return;
    vol_context c = m->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete m;
    ctxRelease(c);
}

// ============================================================================ C ABI
// ------------------------------------------------------------------ explicit corrector operators
namespace {
template <class T> void devCopy(DevBuf<T>& dst, const T* src, size_t cnt, int location, hipStream_t st) {
    dst.alloc(std::max<size_t>(cnt, 1));
    if (cnt)
        HIPCHECK(hipMemcpyAsync(dst.p, src, sizeof(T) * cnt,
                                location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
}

FvcGeom fvcGeom(vol_mesh m) {
    if (!m->geomSet) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh geometry not set (vol_mesh_set_geometry)");
    ensureOwnNei(m);
    FvcGeom G;
    G.n = m->n;
    G.nf = m->nf;
    G.nb = m->nb;
    G.S = m->sell.view();
    G.bOff = m->bOff.p;
    G.bList = m->bList.p;
    G.own = m->dOwn.p;
    G.nei = m->dNei.p;
    G.V = m->gV.p;
    G.Sf = m->gSf.p;
    G.magSf = m->gMagSf.p;
    G.w = m->gW.p;
    G.d = m->gD.p;
    G.skew = m->gSkew.p;
    G.skewed = m->skewed;
    G.emptyDir = m->emptyDir;
    return G;
}

const double* fvcIn(vol_mesh m, int k, const double* p, size_t cnt, int location) {
    if (!p || location == VOL_DEVICE) return p;
    devCopy(m->fst[k], p, cnt, location, m->ctx->stream);
    return m->fst[k].p;
}
const int* fvcInI(vol_mesh m, const int32_t* p, size_t cnt, int location) {
    if (!p || location == VOL_DEVICE) return p;
    devCopy(m->fsti, p, cnt, location, m->ctx->stream);
    return m->fsti.p;
}
double* fvcOutBuf(vol_mesh m, double* out, size_t cnt, int location) {
    if (location == VOL_DEVICE) return out;
    m->fout.alloc(std::max<size_t>(cnt, 1));
    return m->fout.p;
}
void fvcOut(vol_mesh m, double* out, size_t cnt, int location) {
    hipStream_t st = m->ctx->stream;
    if (location != VOL_DEVICE && cnt)
        HIPCHECK(hipMemcpyAsync(out, m->fout.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}
void fvcComponents(int nc) {
    if (nc != 1 && nc != 3) throw VolError(VOL_ERR_INVALID_ARGUMENT, "nComponents must be 1 or 3");
}
}  // namespace

namespace vol::host {
vol_context meshContext(vol_mesh m) { return m->ctx; }
void meshSizes(vol_mesh m, int& n, int& nInternal, int& nBoundary) {
    n = m->n;
    nInternal = m->nf;
    nBoundary = m->nb;
}
bool meshHasGamg(vol_mesh m) { return m->hasGamg; }
void meshGeometry(vol_mesh m, const double*& V, const double*& Sf, const double*& magSf, const double*& w,
                  int& skewed, int& emptyDir) {
    if (!m->geomSet) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh geometry not set (vol_mesh_set_geometry)");
    V = m->gV.p;
    Sf = m->gSf.p;
    magSf = m->gMagSf.p;
    w = m->gW.p;
    skewed = m->skewed;
    emptyDir = m->emptyDir;
}
}  // namespace vol::host

extern "C" {

void vol_controls_default(vol_solver_controls* c) {
    if (!c) return;
    c->solver = "PCG";
    c->preconditioner = "diagonal";
    c->tolerance = 1e-8;
    c->relTol = 0.0;
    c->maxIter = 1000;
    c->minIter = 0;
    c->nSweeps = 1;
    c->sums = VOL_SUMS_CONTEXT;
    c->precision = VOL_FP64;
    c->gamgPreSweeps = 2;
    c->gamgPostSweeps = 2;
    c->gamgOmega = 0.8;
    c->innerRelTol = 1e-2;
    c->gamgSmoother = VOL_GAMG_JACOBI;
    c->reserved = 0;
}

const char* vol_last_error(void) { return gLastError.c_str(); }

int vol_context_create(int device, vol_context* out) {
    API_BEGIN
    if (!out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null output");
    int count = 0;
    HIPCHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw VolError(VOL_ERR_INVALID_ARGUMENT, "no such GPU");
    auto c = std::make_unique<vol_context_s>();
    c->device = device;
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if (const char* e = std::getenv("CFD_REPRODUCIBLE")) c->reproducible = std::string(e) != "0" && std::string(e) != "";
    if (const char* e = std::getenv("VOLUMA_TAIL_MAX_CELLS")) c->tailMax = std::min(std::atoi(e), TAIL_MAX_CELLS);
    auto envOff = [](const char* name) {
        const char* e = std::getenv(name);
        return e && std::string(e) == "0";
    };
    c->coarseZ = !envOff("VOLUMA_COARSE_Z");
    c->level0Z = !envOff("VOLUMA_LEVEL0_Z");
    if (const char* e = std::getenv("VOLUMA_PROLONG_SPLIT")) c->prolongSplit = std::string(e) == "1";
    c->fusedRestrict = !envOff("VOLUMA_FUSED_RESTRICT");
    c->pushProlong = !envOff("VOLUMA_PUSH_PROLONG");
    c->fusedLevel0 = !envOff("VOLUMA_FUSED_LEVEL0");
    if (const char* e = std::getenv("VOLUMA_LAST2")) c->last2 = std::string(e) == "1";
    if (const char* e = std::getenv("VOLUMA_SELL_SORT")) c->sellSort = std::max(0, std::atoi(e));
    c->sellPairs = !envOff("VOLUMA_SELL_PAIRS");
    if (const char* e = std::getenv("VOLUMA_LAST_BLOCK_FIN")) c->lastBlockFin = std::string(e) == "1";
    if (const char* e = std::getenv("VOLUMA_FUSED_SELL_DIRECT")) c->fusedSellDirect = std::string(e) == "1";
    if (const char* e = std::getenv("VOLUMA_FUSED_SELL_MAX")) c->fusedSellMax = std::atoi(e);
    if (const char* e = std::getenv("VOLUMA_SWEEP")) c->freeSweeps = std::string(e) == "free";
    c->sweepFault.alloc(2);  // fault flag, work counter
    HIPCHECK(hipMemsetAsync(c->sweepFault.p, 0, 2 * sizeof(int), c->stream));
    c->scal.alloc(2);
    HIPCHECK(hipHostMalloc((void**)&c->hostScal, 2 * sizeof(DevScalars), hipHostMallocDefault));
    c->ensureScratch(RED_BLOCKS);
    for (hipEvent_t& e : c->loopEv) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    registerScratch(device);
    *out = c.release();
    API_END
}

int vol_context_destroy(vol_context ctx) {
    API_BEGIN
    if (!ctx) return VOL_OK;
    ctxRelease(ctx);
    API_END
}

int vol_set_reproducible(vol_context ctx, int on) {
    API_BEGIN
    if (!ctx) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null context");
    ctx->reproducible = on != 0;
    API_END
}

int vol_synchronize(vol_context ctx) {
    API_BEGIN
    if (!ctx) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null context");
    setDevice(ctx);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    API_END
}

int vol_device_alloc(vol_context ctx, uint64_t bytes, void** ptr) {
    API_BEGIN
    if (!ctx || !ptr) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(ctx);
    HIPCHECK(hipMalloc(ptr, std::max<uint64_t>(bytes, 1)));
    ctx->refs.fetch_add(1);
    API_END
}

int vol_device_free(vol_context ctx, void* ptr) {
    API_BEGIN
    if (!ctx) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null context");
    if (!ptr) return VOL_OK;
    setDevice(ctx);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipFree(ptr));
    ctxRelease(ctx);
    API_END
}

int vol_device_copy(vol_context ctx, double* dst, const double* src, int64_t n, int repeats, double* seconds) {
    API_BEGIN
    if (!ctx || !dst || !src || n < 0 || n > (int64_t(1) << 31) - 1 || repeats < 1)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad copy arguments");
    setDevice(ctx);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    VecArgs v;
    v.n = int(n);
    v.a = src;
    v.w = dst;
    HIPCHECK(hipEventRecord(e0, ctx->stream));
    for (int k = 0; k < repeats; ++k) launch_vec(ctx->stream, V_COPY, v, false, ctx->red);
    HIPCHECK(hipEventRecord(e1, ctx->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = double(ms) * 1e-3 / repeats;
    API_END
}

int vol_memcpy_to_device(vol_context ctx, void* dst, const void* src, uint64_t bytes) {
    API_BEGIN
    if (!ctx || ((!dst || !src) && bytes)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(ctx);
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    API_END
}

int vol_memcpy_to_host(vol_context ctx, void* dst, const void* src, uint64_t bytes) {
    API_BEGIN
    if (!ctx || ((!dst || !src) && bytes)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(ctx);
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    API_END
}

int vol_mesh_create(vol_context ctx, int nCells, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                    const double* gamgWeights, int mergeLevels, vol_mesh* out) {
    API_BEGIN
    if (!ctx || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (nCells < 1 || nInternalFaces < 0 || (nInternalFaces > 0 && (!owner || !neighbour)))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad mesh sizes");
    if (int64_t(nInternalFaces) * 2 >= (int64_t(1) << 30)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh too large");
    for (int f = 0; f < nInternalFaces; ++f) {
        if (owner[f] < 0 || owner[f] >= nCells || neighbour[f] < 0 || neighbour[f] >= nCells || owner[f] == neighbour[f])
            throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad owner/neighbour at face " + std::to_string(f));
    }
    setDevice(ctx);
    auto m = std::make_unique<vol_mesh_s>();
    m->ctx = ctx;
    m->n = nCells;
    m->nf = nInternalFaces;
    m->own.assign(owner, owner + nInternalFaces);
    m->nei.assign(neighbour, neighbour + nInternalFaces);
    setupMesh(m.get(), gamgWeights, mergeLevels > 0 ? mergeLevels : 2);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->refs.fetch_add(1);
    *out = m.release();
    API_END
}

int vol_mesh_destroy(vol_mesh mesh) {
    API_BEGIN
    if (!mesh) return VOL_OK;
    meshRelease(mesh);
    API_END
}

int vol_mesh_gamg_levels(vol_mesh mesh, int* nLevels, int32_t* sizes) {
    API_BEGIN
    if (!mesh || !nLevels) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    *nLevels = int(mesh->H.size());
    if (sizes)
        for (size_t k = 0; k < mesh->H.size(); ++k) sizes[k] = mesh->H[k].n;
    API_END
}

int vol_mesh_gamg_setup_info(vol_mesh mesh, double* seconds, int* passes, int* rounds) {
    API_BEGIN
    if (!mesh) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null mesh");
    if (seconds) *seconds = mesh->aggSeconds;
    if (passes) *passes = mesh->aggStats.passes;
    if (rounds) *rounds = mesh->aggStats.rounds;
    API_END
}

int vol_mesh_gamg_level(vol_mesh mesh, int level, int32_t* nFaces, int32_t* agg, int32_t* faceMap,
                        int32_t* faceFlip) {
    API_BEGIN
    if (!mesh || level < 0 || level >= int(mesh->H.size())) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad level");
    const HLevel& L = mesh->H[level];
    if (nFaces) *nFaces = int(L.l.size());
    const bool coarsest = L.agg.empty();
    if (agg)
        for (int i = 0; i < L.n; ++i) agg[i] = coarsest ? -1 : L.agg[i];
    for (size_t f = 0; f < L.l.size(); ++f) {
        if (faceMap) faceMap[f] = coarsest ? -1 : L.faceMap[f];
        if (faceFlip) faceFlip[f] = coarsest ? 0 : L.faceFlip[f];
    }
    API_END
}

int vol_mesh_set_boundary(vol_mesh mesh, int nBoundaryFaces, const int32_t* boundaryOwner) {
    API_BEGIN
    if (!mesh || nBoundaryFaces < 0 || (nBoundaryFaces > 0 && !boundaryOwner))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad boundary");
    const int n = mesh->n;
    std::vector<int> off(n + 1, 0), list(nBoundaryFaces);
    for (int bf = 0; bf < nBoundaryFaces; ++bf) {
        if (boundaryOwner[bf] < 0 || boundaryOwner[bf] >= n) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad boundary owner");
        ++off[boundaryOwner[bf] + 1];
    }
    for (int c = 0; c < n; ++c) off[c + 1] += off[c];
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int bf = 0; bf < nBoundaryFaces; ++bf) list[fill[boundaryOwner[bf]]++] = bf;
    }
    setDevice(mesh->ctx);
    mesh->bOff.upload(off, mesh->ctx->stream);
    mesh->bList.upload(list, mesh->ctx->stream);
    mesh->nb = nBoundaryFaces;
    HIPCHECK(hipStreamSynchronize(mesh->ctx->stream));
    API_END
}


int vol_geometry(vol_context ctx, int nPoints, const double* points, int nFaces, const int32_t* faceOffsets,
                 const int32_t* facePoints, int nInternalFaces, const int32_t* owner, const int32_t* neighbour,
                 int nCells, double* V, double* C, double* Sf, double* Cf, double* d, double* magSf, double* w,
                 double* ndc, double* corrVecs, double* skewVecs, double* maxSkewness, int location) {
    API_BEGIN
    if (!ctx || !points || !faceOffsets || !facePoints || !owner || (nInternalFaces > 0 && !neighbour) || !V || !C ||
        !Sf || !Cf || !d || !magSf || !w || !ndc || !corrVecs)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (nPoints < 0 || nFaces < 0 || nCells <= 0 || nInternalFaces < 0 || nInternalFaces > nFaces)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad mesh sizes");
    setDevice(ctx);
    hipStream_t st = ctx->stream;
    const bool dev = location == VOL_DEVICE;
    // addressing on the host: checked here (every index the kernels follow) and used for
    // the cell → face lists
    auto host = [&](const int32_t* p, size_t cnt) {
        std::vector<int> v(cnt);
        if (cnt)
            HIPCHECK(hipMemcpy(v.data(), p, sizeof(int) * cnt, dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
        return v;
    };
    const std::vector<int> fo = host(faceOffsets, size_t(nFaces) + 1), own = host(owner, nFaces),
                           nei = host(neighbour, nInternalFaces);
    if (fo[0] != 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "faceOffsets must start at 0");
    for (int f = 0; f < nFaces; ++f)
        if (fo[f + 1] - fo[f] < 3) throw VolError(VOL_ERR_INVALID_ARGUMENT, "face with fewer than 3 points");
    const std::vector<int> fp = host(facePoints, size_t(fo[nFaces]));
    for (int k : fp)
        if (k < 0 || k >= nPoints) throw VolError(VOL_ERR_INVALID_ARGUMENT, "face point out of range");
    std::vector<int> cfOff(size_t(nCells) + 1, 0);
    for (int f = 0; f < nFaces; ++f) {
        if (own[f] < 0 || own[f] >= nCells || (f < nInternalFaces && (nei[f] < 0 || nei[f] >= nCells)))
            throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad owner/neighbour at face " + std::to_string(f));
        ++cfOff[own[f] + 1];
        if (f < nInternalFaces) ++cfOff[nei[f] + 1];
    }
    for (int c = 0; c < nCells; ++c) {
        if (cfOff[c + 1] == 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "cell without faces");
        cfOff[c + 1] += cfOff[c];
    }
    std::vector<int> cfList(cfOff[nCells]);
    {
        std::vector<int> fill(cfOff.begin(), cfOff.end() - 1);
        for (int f = 0; f < nFaces; ++f) {
            cfList[fill[own[f]]++] = f;
            if (f < nInternalFaces) cfList[fill[nei[f]]++] = f;
        }
    }
    DevBuf<double> dPts, o[10];
    DevBuf<int> dFo, dFp, dOwn, dNei, dCfo, dCfl;
    DevBuf<unsigned long long> dMax;
    dFo.upload(fo, st);
    dFp.upload(fp, st);
    dOwn.upload(own, st);
    dNei.upload(nei, st);
    dCfo.upload(cfOff, st);
    dCfl.upload(cfList, st);
    GeoArgs g;
    g.nCells = nCells;
    g.nFaces = nFaces;
    g.nInternal = nInternalFaces;
    if (dev) {
        g.points = points;
    } else {
        dPts.alloc(std::max<size_t>(3 * size_t(nPoints), 1));
        if (nPoints)
            HIPCHECK(hipMemcpyAsync(dPts.p, points, sizeof(double) * 3 * nPoints, hipMemcpyHostToDevice, st));
        g.points = dPts.p;
    }
    g.faceOff = dFo.p;
    g.facePts = dFp.p;
    g.owner = dOwn.p;
    g.neighbour = dNei.p;
    g.cfOff = dCfo.p;
    g.cfList = dCfl.p;
    double* outs[10] = {V, C, Sf, Cf, d, magSf, w, ndc, corrVecs, skewVecs};
    const size_t cnt[10] = {size_t(nCells),     3 * size_t(nCells), 3 * size_t(nFaces), 3 * size_t(nFaces),
                            3 * size_t(nFaces), size_t(nFaces),     size_t(nFaces),     size_t(nFaces),
                            3 * size_t(nFaces), 3 * size_t(nFaces)};
    double* dst[10];
    for (int k = 0; k < 10; ++k) {
        if (!outs[k]) {
            dst[k] = nullptr;
        } else if (dev) {
            dst[k] = outs[k];
        } else {
            o[k].alloc(std::max<size_t>(cnt[k], 1));
            dst[k] = o[k].p;
        }
    }
    g.V = dst[0];
    g.C = dst[1];
    g.Sf = dst[2];
    g.Cf = dst[3];
    g.d = dst[4];
    g.magSf = dst[5];
    g.w = dst[6];
    g.ndc = dst[7];
    g.corr = dst[8];
    g.skew = dst[9];
    if (maxSkewness) {
        dMax.alloc(1);
        HIPCHECK(hipMemsetAsync(dMax.p, 0, sizeof(unsigned long long), st));
        g.maxSkewBits = dMax.p;
    }
    launch_geometry(st, g);
    if (!dev)
        for (int k = 0; k < 10; ++k)
            if (outs[k] && cnt[k])
                HIPCHECK(hipMemcpyAsync(outs[k], dst[k], sizeof(double) * cnt[k], hipMemcpyDeviceToHost, st));
    if (maxSkewness) {
        unsigned long long bits = 0;
        HIPCHECK(hipMemcpyAsync(&bits, dMax.p, sizeof(bits), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::memcpy(maxSkewness, &bits, sizeof(double));
    }
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_mesh_set_geometry(vol_mesh mesh, const double* V, const double* Sf, const double* magSf, const double* w,
                          const double* d, const double* skew, int skewed, int location) {
    API_BEGIN
    if (!mesh || !V || !Sf || !magSf || !w || !d) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (mesh->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    setDevice(mesh->ctx);
    hipStream_t st = mesh->ctx->stream;
    const size_t n = mesh->n, nF = size_t(mesh->nf) + mesh->nb;
    devCopy(mesh->gV, V, n, location, st);
    devCopy(mesh->gSf, Sf, 3 * nF, location, st);
    devCopy(mesh->gMagSf, magSf, nF, location, st);
    devCopy(mesh->gW, w, nF, location, st);
    devCopy(mesh->gD, d, 3 * nF, location, st);
    if (skew) {
        devCopy(mesh->gSkew, skew, 3 * nF, location, st);
    } else {
        mesh->gSkew.alloc(std::max<size_t>(3 * nF, 1));
        HIPCHECK(hipMemsetAsync(mesh->gSkew.p, 0, sizeof(double) * std::max<size_t>(3 * nF, 1), st));
    }
    mesh->skewed = skewed ? 1 : 0;
    mesh->geomSet = true;
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_boundary_values(vol_mesh mesh, int nComponents, const int32_t* bFixed, const double* bValue,
                        const double* cellValues, double* out, int location) {
    API_BEGIN
    if (!mesh || !cellValues || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvcComponents(nComponents);
    if (mesh->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    if (mesh->nb > 0 && (!bFixed || !bValue)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bFixed and bValue required");
    setDevice(mesh->ctx);
    FvcGeom G;
    G.n = mesh->n;
    G.nf = mesh->nf;
    G.nb = mesh->nb;
    G.bOff = mesh->bOff.p;
    G.bList = mesh->bList.p;
    const size_t nc = nComponents, nb = mesh->nb;
    launch_boundary_values(mesh->ctx->stream, G, nComponents, fvcInI(mesh, bFixed, nb, location),
                           fvcIn(mesh, 0, bValue, nc * nb, location), fvcIn(mesh, 1, cellValues, nc * G.n, location),
                           fvcOutBuf(mesh, out, nc * nb, location));
    fvcOut(mesh, out, nc * nb, location);
    API_END
}

int vol_copy_strided(vol_context ctx, int64_t n, const double* src, int srcStride, double* dst, int dstStride,
                     double scale) {
    API_BEGIN
    if (!ctx || (n > 0 && (!src || !dst))) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 0 || srcStride < 1 || dstStride < 1) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad count or stride");
    setDevice(ctx);
    launch_copy_strided(ctx->stream, n, src, srcStride, dst, dstStride, scale);
    HIPCHECK(hipGetLastError());
    API_END
}

int vol_mesh_set_empty_dir(vol_mesh mesh, int emptyDir) {
    API_BEGIN
    if (!mesh) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null mesh");
    if (emptyDir < -1 || emptyDir > 2) throw VolError(VOL_ERR_INVALID_ARGUMENT, "emptyDir must be -1, 0, 1 or 2");
    mesh->emptyDir = emptyDir;
    API_END
}

int vol_fvc_interpolate(vol_mesh mesh, int nComponents, const double* cellValues, const double* cellGrad,
                        const double* boundaryValues, double* faceValues, int location) {
    API_BEGIN
    if (!mesh || !cellValues || !faceValues) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvcComponents(nComponents);
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    const size_t nc = nComponents, n = G.n, nF = size_t(G.nf) + G.nb;
    launch_fvc_interpolate(mesh->ctx->stream, G, nComponents, fvcIn(mesh, 0, cellValues, nc * n, location),
                           fvcIn(mesh, 1, cellGrad, 3 * nc * n, location),
                           fvcIn(mesh, 2, boundaryValues, nc * G.nb, location),
                           fvcOutBuf(mesh, faceValues, nc * nF, location));
    fvcOut(mesh, faceValues, nc * nF, location);
    API_END
}

int vol_fvc_grad(vol_mesh mesh, int scheme, int nComponents, const double* cellValues, const int32_t* bFixesValue,
                 const double* bValue, double* cellGrad, int location) {
    API_BEGIN
    if (!mesh || !cellValues || !cellGrad) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvcComponents(nComponents);
    if (scheme != VOL_GRAD_GAUSS_LINEAR && scheme != VOL_GRAD_LEAST_SQUARES && scheme != VOL_GRAD_GAUSS_SKEW)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "unknown gradient scheme");
    if (scheme != VOL_GRAD_LEAST_SQUARES && nComponents != 1)
        throw VolError(VOL_ERR_UNSUPPORTED, "Gauss gradients are provided for scalar fields");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    if (G.nb > 0 && (!bFixesValue || !bValue))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bFixesValue and bValue required");
    hipStream_t st = mesh->ctx->stream;
    const size_t nc = nComponents, n = G.n;
    const double* c = fvcIn(mesh, 0, cellValues, nc * n, location);
    const int* bf = fvcInI(mesh, bFixesValue, G.nb, location);
    const double* bv = fvcIn(mesh, 1, bValue, nc * G.nb, location);
    double* out = fvcOutBuf(mesh, cellGrad, 3 * nc * n, location);
    if (scheme == VOL_GRAD_LEAST_SQUARES) {
        launch_fvc_grad_ls(st, G, nComponents, c, bf, bv, out);
    } else if (scheme == VOL_GRAD_GAUSS_LINEAR) {
        launch_fvc_grad_gauss(st, G, false, c, nullptr, bf, bv, out);
    } else {
        mesh->fgls.alloc(std::max<size_t>(3 * n, 1));
        launch_fvc_grad_ls(st, G, 1, c, bf, bv, mesh->fgls.p);
        launch_fvc_grad_gauss(st, G, true, c, mesh->fgls.p, bf, bv, out);
    }
    fvcOut(mesh, cellGrad, 3 * nc * n, location);
    API_END
}

int vol_fvc_div(vol_mesh mesh, const double* faceFlux, double* out, int location) {
    API_BEGIN
    if (!mesh || !faceFlux || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    launch_fvc_integrate(mesh->ctx->stream, G, 1, true, fvcIn(mesh, 0, faceFlux, size_t(G.nf) + G.nb, location),
                         fvcOutBuf(mesh, out, G.n, location));
    fvcOut(mesh, out, G.n, location);
    API_END
}

int vol_fvc_surface_integrate(vol_mesh mesh, int nComponents, const double* faceValues, double* out, int location) {
    API_BEGIN
    if (!mesh || !faceValues || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvcComponents(nComponents);
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    const size_t nc = nComponents;
    launch_fvc_integrate(mesh->ctx->stream, G, nComponents, false,
                         fvcIn(mesh, 0, faceValues, nc * (size_t(G.nf) + G.nb), location),
                         fvcOutBuf(mesh, out, nc * G.n, location));
    fvcOut(mesh, out, nc * G.n, location);
    API_END
}

int vol_fvc_flux(vol_mesh mesh, const double* U, const double* gradU, const int32_t* bFixesFlux, const double* Ub,
                 double* phi, int location) {
    API_BEGIN
    if (!mesh || !U || !phi) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    if (G.nb > 0 && (!bFixesFlux || !Ub)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bFixesFlux and Ub required");
    const size_t n = G.n, nF = size_t(G.nf) + G.nb;
    launch_fvc_flux(mesh->ctx->stream, G, fvcIn(mesh, 0, U, 3 * n, location), fvcIn(mesh, 1, gradU, 9 * n, location),
                    fvcInI(mesh, bFixesFlux, G.nb, location), fvcIn(mesh, 2, Ub, 3 * size_t(G.nb), location),
                    fvcOutBuf(mesh, phi, nF, location));
    fvcOut(mesh, phi, nF, location);
    API_END
}

int vol_piso_rau_hbya(vol_mesh mesh, const double* A, const double* H, double* rAU, double* HbyA, int location) {
    API_BEGIN
    if (!mesh || !A || !H || !rAU || !HbyA) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(mesh->ctx);
    hipStream_t st = mesh->ctx->stream;
    const size_t n = mesh->n;
    const double* a = fvcIn(mesh, 0, A, n, location);
    const double* h = fvcIn(mesh, 1, H, 3 * n, location);
    double* r = rAU;
    double* hb = HbyA;
    if (location != VOL_DEVICE) {
        mesh->fst[2].alloc(std::max<size_t>(n, 1));
        mesh->fst[3].alloc(std::max<size_t>(3 * n, 1));
        r = mesh->fst[2].p;
        hb = mesh->fst[3].p;
    }
    launch_piso_rau_hbya(st, int(n), a, h, r, hb);
    if (location != VOL_DEVICE && n) {
        HIPCHECK(hipMemcpyAsync(rAU, r, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(HbyA, hb, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_piso_hbya_boundary(vol_mesh mesh, const int32_t* bFixesFlux, const double* Ub, const double* rAU,
                           const double* gradp, double* bValue, int location) {
    API_BEGIN
    if (!mesh || !rAU || !gradp || !bValue) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    if (G.nb > 0 && (!bFixesFlux || !Ub)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bFixesFlux and Ub required");
    const size_t n = G.n, nb = G.nb;
    launch_piso_hbya_boundary(mesh->ctx->stream, G, fvcInI(mesh, bFixesFlux, nb, location),
                              fvcIn(mesh, 0, Ub, 3 * nb, location), fvcIn(mesh, 1, rAU, n, location),
                              fvcIn(mesh, 2, gradp, 3 * n, location), fvcOutBuf(mesh, bValue, 3 * nb, location));
    fvcOut(mesh, bValue, 3 * nb, location);
    API_END
}

int vol_piso_face_op(vol_mesh mesh, int op, const double* a, const double* b, const double* c, double* out,
                     int location) {
    API_BEGIN
    if (!mesh || !a || !b || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (op != VOL_FACE_A_PLUS_B_TIMES_C && op != VOL_FACE_A_MINUS_B && op != VOL_FACE_A_MINUS_B_OVER_C)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "unknown face operation");
    if (op != VOL_FACE_A_MINUS_B && !c) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (mesh->nb < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "mesh boundary not set (vol_mesh_set_boundary)");
    setDevice(mesh->ctx);
    const size_t nF = size_t(mesh->nf) + mesh->nb;
    launch_piso_face_op(mesh->ctx->stream, op, int(nF), fvcIn(mesh, 0, a, nF, location), fvcIn(mesh, 1, b, nF, location),
                        op == VOL_FACE_A_MINUS_B ? nullptr : fvcIn(mesh, 2, c, nF, location),
                        fvcOutBuf(mesh, out, nF, location));
    fvcOut(mesh, out, nF, location);
    API_END
}

int vol_piso_ddt_phi_corr(vol_mesh mesh, double ddtPhiCoeff, double c0, double c00, const double* U0,
                          const double* U00, const double* phiOld, const double* phiOldOld, const double* rAUf,
                          double* phiHbyA, int location) {
    API_BEGIN
    if (!mesh || !U0 || !phiOld || !rAUf || !phiHbyA) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (c00 != 0.0 && (!U00 || !phiOldOld)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "U00 and phiOldOld required");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    hipStream_t st = mesh->ctx->stream;
    const size_t n = G.n, nF = size_t(G.nf) + G.nb;
    double* io = fvcOutBuf(mesh, phiHbyA, nF, location);
    if (location != VOL_DEVICE && nF)
        HIPCHECK(hipMemcpyAsync(io, phiHbyA, sizeof(double) * nF, hipMemcpyHostToDevice, st));
    const bool two = c00 != 0.0;
    launch_piso_ddt_phi_corr(st, G, ddtPhiCoeff, c0, c00, fvcIn(mesh, 0, U0, 3 * n, location),
                             two ? fvcIn(mesh, 1, U00, 3 * n, location) : nullptr, fvcIn(mesh, 2, phiOld, nF, location),
                             two ? fvcIn(mesh, 3, phiOldOld, nF, location) : nullptr,
                             fvcIn(mesh, 4, rAUf, nF, location), io);
    fvcOut(mesh, phiHbyA, nF, location);
    API_END
}

int vol_piso_constrain_pressure(vol_mesh mesh, int patchStart, int patchSize, const double* phiHbyA,
                                const double* Ub, const double* rAUf, double* gradient, int location) {
    API_BEGIN
    if (!mesh || !phiHbyA || !Ub || !rAUf || (patchSize > 0 && !gradient))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    if (patchStart < 0 || patchSize < 0 || int64_t(patchStart) + patchSize > G.nb)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "patch range outside the boundary faces");
    const size_t nF = size_t(G.nf) + G.nb;
    launch_piso_constrain_pressure(mesh->ctx->stream, G, patchStart, patchSize, fvcIn(mesh, 0, phiHbyA, nF, location),
                                   fvcIn(mesh, 1, Ub, 3 * size_t(G.nb), location), fvcIn(mesh, 2, rAUf, nF, location),
                                   fvcOutBuf(mesh, gradient, size_t(patchSize), location));
    fvcOut(mesh, gradient, size_t(patchSize), location);
    API_END
}

int vol_fvc_reconstruct(vol_mesh mesh, const double* faceFlux, int emptyDir, double* out, int location) {
    API_BEGIN
    if (!mesh || !faceFlux || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (emptyDir < -1 || emptyDir > 2) throw VolError(VOL_ERR_INVALID_ARGUMENT, "emptyDir must be -1, 0, 1 or 2");
    setDevice(mesh->ctx);
    const FvcGeom G = fvcGeom(mesh);
    const size_t n = G.n, nF = size_t(G.nf) + G.nb;
    launch_fvc_reconstruct(mesh->ctx->stream, G, fvcIn(mesh, 0, faceFlux, nF, location), emptyDir,
                           fvcOutBuf(mesh, out, 3 * n, location));
    fvcOut(mesh, out, 3 * n, location);
    API_END
}

int vol_piso_correct_U(vol_mesh mesh, const double* HbyA, const double* rAU, const double* v, int sign, int emptyDir,
                       double* U, int location) {
    API_BEGIN
    if (!mesh || !HbyA || !rAU || !v || !U) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    if (emptyDir < -1 || emptyDir > 2) throw VolError(VOL_ERR_INVALID_ARGUMENT, "emptyDir must be -1, 0, 1 or 2");
    setDevice(mesh->ctx);
    const size_t n = mesh->n;
    launch_piso_correct_U(mesh->ctx->stream, int(n), fvcIn(mesh, 0, HbyA, 3 * n, location),
                          fvcIn(mesh, 1, rAU, n, location), fvcIn(mesh, 2, v, 3 * n, location), sign, emptyDir,
                          fvcOutBuf(mesh, U, 3 * n, location));
    fvcOut(mesh, U, 3 * n, location);
    API_END
}

int vol_assemble_laplacian(vol_matrix A, double gamma, const double* faceGamma, const double* magSf,
                           const double* ndc, const int32_t* bcFixed, const double* bcValue, const double* bIn,
                           double* bOut, double* internalCoeffs, double* boundarySource, int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    assemble(A, ASM_LAPLACIAN, gamma, faceGamma, magSf, ndc, nullptr, bcFixed, bcValue, bIn, bOut, internalCoeffs,
             boundarySource, location);
    API_END
}

int vol_assemble_div_upwind(vol_matrix A, const double* flux, const int32_t* bcFixed, const double* bcValue,
                            const double* bIn, double* bOut, double* internalCoeffs, double* boundarySource,
                            int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    assemble(A, ASM_DIV, 1.0, nullptr, nullptr, nullptr, flux, bcFixed, bcValue, bIn, bOut, internalCoeffs,
             boundarySource, location);
    API_END
}

int vol_fvm_reset(vol_matrix A) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    setDevice(A->mesh->ctx);
    fvmReset(A);
    HIPCHECK(hipStreamSynchronize(A->mesh->ctx->stream));
    API_END
}

int vol_fvm_ddt_euler(vol_matrix A, int sign, double dt, const double* V, const double* psi0, int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, sign);
    if (!V || !psi0 || !(dt > 0.0)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "ddt needs V, psi0 and dt > 0");
    setDevice(A->mesh->ctx);
    const size_t n = A->mesh->n;
    FvmArgs a = fvmArgs(A, sign);
    a.ddtC = a.ddtC0 = 1.0 / dt;  // Schemes.h:74,83
    a.ddtC00 = 0.0;
    a.V = stageK(A, 0, V, n, location);
    a.psi0 = a.psi00 = stageK(A, 1, psi0, n, location);
    fvmTerm(A, FVM_DDT, a, false);
    API_END
}

int vol_fvm_ddt(vol_matrix A, int sign, double c, double c0, double c00, const double* V, const double* psi0,
                const double* psi00, int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, sign);
    if (!V || !psi0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "ddt needs V and psi0");
    setDevice(A->mesh->ctx);
    const size_t n = A->mesh->n;
    FvmArgs a = fvmArgs(A, sign);
    a.ddtC = c;
    a.ddtC0 = c0;
    a.ddtC00 = c00;
    a.V = stageK(A, 0, V, n, location);
    a.psi0 = stageK(A, 1, psi0, n, location);
    a.psi00 = psi00 ? stageK(A, 2, psi00, n, location) : a.psi0;
    fvmTerm(A, FVM_DDT, a, false);
    API_END
}

int vol_fvm_div_upwind(vol_matrix A, int sign, const double* flux, const int32_t* bcFixed, const double* bcValue,
                       int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, sign);
    vol_mesh m = A->mesh;
    if (!flux || (!bcFixed && m->nb > 0)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "div needs flux and bcFixed");
    setDevice(m->ctx);
    const size_t nF = size_t(m->nf) + m->nb;
    FvmArgs a = fvmArgs(A, sign);
    a.flux = stageK(A, 0, flux, nF, location);
    a.bcFixed = stage(A->sti, bcFixed, size_t(m->nb), location, m->ctx->stream);
    a.bcValue = stageK(A, 1, bcValue, size_t(m->nb), location);
    fvmTerm(A, FVM_DIV, a, true);
    API_END
}

int vol_fvm_laplacian(vol_matrix A, int sign, double gamma, const double* faceGamma, const double* magSf,
                      const double* ndc, const int32_t* bcFixed, const double* bcValue, const vol_nonorth* nonOrth,
                      int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, sign);
    vol_mesh m = A->mesh;
    if (!magSf || !ndc || (!bcFixed && m->nb > 0))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "laplacian needs magSf, ndc and bcFixed");
    if (nonOrth && (!nonOrth->psi || !nonOrth->w || !nonOrth->Sf || !nonOrth->corrVecs || !nonOrth->V))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "non-orthogonal correction needs psi, w, Sf, corrVecs and V");
    setDevice(m->ctx);
    const size_t n = m->n, nF = size_t(m->nf) + m->nb;
    FvmArgs a = fvmArgs(A, sign);
    a.gammaU = gamma;
    a.gamma = stageK(A, 0, faceGamma, nF, location);
    a.magSf = stageK(A, 1, magSf, nF, location);
    a.ndc = stageK(A, 2, ndc, nF, location);
    a.bcFixed = stage(A->sti, bcFixed, size_t(m->nb), location, m->ctx->stream);
    a.bcValue = stageK(A, 3, bcValue, size_t(m->nb), location);
    if (nonOrth) {
        a.psi = stageK(A, 4, nonOrth->psi, n, location);
        a.w = stageK(A, 5, nonOrth->w, nF, location);
        a.Sf = stageK(A, 6, nonOrth->Sf, 3 * nF, location);
        a.corr = stageK(A, 7, nonOrth->corrVecs, 3 * nF, location);
        a.V = stageK(A, 8, nonOrth->V, n, location);
    }
    fvmTerm(A, FVM_LAP, a, false);
    API_END
}

int vol_fvm_source(vol_matrix A, const double* su, const double* V, int location) {
    API_BEGIN
    if (!A || !su || !V) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    const size_t n = m->n;
    launch_fvm_source(m->ctx->stream, m->n, stageK(A, 0, su, n, location), stageK(A, 1, V, n, location), A->fsrc.p);
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    API_END
}

int vol_fvm_get(vol_matrix A, double* diag, double* source, double* ic, double* bs, double* ffc, int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, 1);
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    hipStream_t st = m->ctx->stream;
    const hipMemcpyKind kind = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    auto cp = [&](double* dst, const DevBuf<double>& b, size_t cnt) {
        if (dst && cnt) HIPCHECK(hipMemcpyAsync(dst, b.p, sizeof(double) * cnt, kind, st));
    };
    const size_t n = m->n, nf = m->nf, nb = m->nb;
    cp(diag, A->fdiag, n);
    cp(source, A->fsrc, n);
    cp(ic, A->fic, nb);
    cp(bs, A->fbs, nb);
    cp(ffc, A->fffc, nf + nb);
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_fvm_finish(vol_matrix A, double* b, int location) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, 1);
    fvmFinish(A, b, nullptr, nullptr, location);
    API_END
}

int vol_fvm_finish_source(vol_matrix A, const double* su, const double* V, double* b, int location) {
    API_BEGIN
    if (!A || !su || !V || !b) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    fvmFinish(A, b, su, V, location);
    API_END
}

int vol_fvm_relax(vol_matrix A, double alpha, const double* psi, int location) {
    API_BEGIN
    if (!A || !psi) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    if (alpha >= 1.0) return VOL_OK;  // FvMatrix.h:112
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    FvmArgs a = fvmArgs(A, 1);
    launch_fvm_relax(m->ctx->stream, a, alpha, stageK(A, 0, psi, size_t(m->n), location), fvmLower(A));
    A->set = false;  // the solver matrix comes from the next vol_fvm_finish
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    API_END
}

int vol_fvm_set_reference(vol_matrix A, int64_t cell, double value) {
    API_BEGIN
    if (!A) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null matrix");
    fvmCheck(A, 1);
    vol_mesh m = A->mesh;
    if (cell < 0 || cell >= m->n) return VOL_OK;  // no cell of this mesh carries that number
    setDevice(m->ctx);
    launch_fvm_set_reference(m->ctx->stream, A->fdiag.p, A->fsrc.p, int(cell), value);
    A->set = false;
    HIPCHECK(hipStreamSynchronize(m->ctx->stream));
    API_END
}

int vol_fvm_A(vol_matrix A, const double* V, int nComponents, double* out, int location) {
    API_BEGIN
    if (!A || !V || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    if (nComponents != 1 && nComponents != 3) throw VolError(VOL_ERR_INVALID_ARGUMENT, "nComponents must be 1 or 3");
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    const size_t n = m->n;
    FvmArgs a = fvmArgs(A, 1);
    launch_fvm_A(m->ctx->stream, a, stageK(A, 1, V, n, location), nComponents, fvmOutBuf(A, out, n, location));
    fvmOut(A, out, n, location);
    API_END
}

int vol_fvm_H(vol_matrix A, const double* psi, const double* V, int nComponents, double* out, int location) {
    API_BEGIN
    if (!A || !psi || !V || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    if (nComponents != 1 && nComponents != 3) throw VolError(VOL_ERR_INVALID_ARGUMENT, "nComponents must be 1 or 3");
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    const size_t n = m->n;
    FvmArgs a = fvmArgs(A, 1);
    launch_fvm_H(m->ctx->stream, a, fvmLower(A), stageK(A, 0, psi, n, location), stageK(A, 1, V, n, location),
                 nComponents, fvmOutBuf(A, out, n, location));
    fvmOut(A, out, n, location);
    API_END
}

int vol_fvm_flux(vol_matrix A, const double* psi, double* out, int location) {
    API_BEGIN
    if (!A || !psi || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    fvmCheck(A, 1);
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    ensureOwnNei(m);
    const size_t nF = size_t(m->nf) + m->nb;
    FvmArgs a = fvmArgs(A, 1);
    launch_fvm_flux(m->ctx->stream, a, fvmLower(A), m->dOwn.p, m->dNei.p, stageK(A, 0, psi, size_t(m->n), location),
                    A->fvmHasFfc ? 1 : 0, fvmOutBuf(A, out, nF, location));
    fvmOut(A, out, nF, location);
    API_END
}

int vol_matrix_create(vol_mesh mesh, vol_matrix* out) {
    API_BEGIN
    if (!mesh || !out) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    auto A = std::make_unique<vol_matrix_s>();
    A->mesh = mesh;
    mesh->refs.fetch_add(1);
    *out = A.release();
    API_END
}

int vol_matrix_destroy(vol_matrix A) {
    API_BEGIN
    if (!A) return VOL_OK;
    vol_mesh m = A->mesh;
    setDevice(m->ctx);
    (void)hipStreamSynchronize(m->ctx->stream);
    delete A;
    meshRelease(m);
    API_END
}

int vol_matrix_set(vol_matrix A, const double* diag, const double* upper, const double* lower, int location) {
    API_BEGIN
    if (!A || !diag || (!upper && A->mesh->nf > 0)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    setDevice(c);
    hipStream_t st = c->stream;
    const int n = m->n, nf = m->nf;
    A->symmetric = (lower == nullptr || lower == upper);
    A->diag.alloc(n);
    A->upper.alloc(nf);
    const hipMemcpyKind kind = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHECK(hipMemcpyAsync(A->diag.p, diag, sizeof(double) * n, kind, st));
    if (nf) HIPCHECK(hipMemcpyAsync(A->upper.p, upper, sizeof(double) * nf, kind, st));
    if (!A->symmetric) {
        A->lower.alloc(nf);
        if (nf) HIPCHECK(hipMemcpyAsync(A->lower.p, lower, sizeof(double) * nf, kind, st));
    }
    matrixFinish(A);
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_matrix_get(vol_matrix A, double* diag, double* upper, double* lower, int location) {
    API_BEGIN
    if (!A || !A->set) throw VolError(VOL_ERR_INVALID_ARGUMENT, "matrix not set");
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    setDevice(c);
    hipStream_t st = c->stream;
    const hipMemcpyKind kind = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (diag) HIPCHECK(hipMemcpyAsync(diag, A->diag.p, sizeof(double) * m->n, kind, st));
    if (upper && m->nf) HIPCHECK(hipMemcpyAsync(upper, A->upper.p, sizeof(double) * m->nf, kind, st));
    if (lower && m->nf)
        HIPCHECK(hipMemcpyAsync(lower, A->symmetric ? A->upper.p : A->lower.p, sizeof(double) * m->nf, kind, st));
    HIPCHECK(hipStreamSynchronize(st));
    API_END
}

int vol_amul(vol_matrix A, const double* x, double* y, int location) {
    API_BEGIN
    if (!A || !A->set || !x || !y) throw VolError(VOL_ERR_INVALID_ARGUMENT, "matrix not set or null vector");
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    setDevice(c);
    hipStream_t st = c->stream;
    const int n = m->n;
    const double* dx = x;
    double* dy = y;
    if (location != VOL_DEVICE) {
        A->hx.alloc(n);
        A->hy.alloc(n);
        HIPCHECK(hipMemcpyAsync(A->hx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
        dx = A->hx.p;
        dy = A->hy.p;
    }
    Ctx C{c, st, nullptr, c->scal.p};
    amulTimed<double>(C, A, dx, dy, nullptr, false);  // timed when timing is on
    if (location != VOL_DEVICE) HIPCHECK(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    collectTiming(c);
    API_END
}

int vol_solve(vol_matrix A, const double* b, double* x, const vol_solver_controls* controls,
              vol_solver_performance* perf, int location) {
    API_BEGIN
    if (!A || !A->set || !b || !x) throw VolError(VOL_ERR_INVALID_ARGUMENT, "matrix not set or null vector");
    vol_solver_controls ct;
    if (controls)
        ct = *controls;
    else
        vol_controls_default(&ct);
    vol_solver_performance pf;
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    setDevice(c);
    const int n = m->n;
    if (location == VOL_DEVICE) {
        solveDevice(A, b, x, ct, pf);
    } else {
        hipStream_t st = c->stream;
        A->vb.alloc(n);
        A->vx.alloc(n);
        HIPCHECK(hipMemcpyAsync(A->vb.p, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(A->vx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
        solveDevice(A, A->vb.p, A->vx.p, ct, pf);
        HIPCHECK(hipMemcpyAsync(x, A->vx.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    if (perf) *perf = pf;
    API_END
}

int vol_solve_noalloc(vol_matrix A, const double* b, double* x, const vol_solver_controls* controls,
                      vol_solver_performance* perf) {
    API_BEGIN
    if (!A || !A->set || !b || !x) throw VolError(VOL_ERR_INVALID_ARGUMENT, "matrix not set or null vector");
    vol_mesh m = A->mesh;
    vol_context c = m->ctx;
    const int n = m->n;
    if (n > VOL_NOALLOC_MAX_CELLS)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "vol_solve_noalloc: " + std::to_string(n) +
                                                     " cells exceed its scratch limit of " +
                                                     std::to_string(VOL_NOALLOC_MAX_CELLS));
    if (sell_grid_blocks(m->sell.nslices) > NOALLOC_PART_BLOCKS)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "vol_solve_noalloc: matrix layout exceeds the scratch partials");
    unsigned char* base = scratchOf(c->device);
    if (!base) throw VolError(VOL_ERR_UNSUPPORTED, "vol_solve_noalloc: no scratch on this device");
    vol_solver_controls ct;
    if (controls)
        ct = *controls;
    else
        vol_controls_default(&ct);
    Workspace ws;
    double* v[NOALLOC_VECTORS];
    for (int k = 0; k < NOALLOC_VECTORS; ++k) v[k] = reinterpret_cast<double*>(base + k * NOALLOC_VEC_BYTES);
    ws.wA = v[0];
    ws.pA = v[1];
    ws.vr = v[2];
    ws.vw = v[3];
    ws.vp = v[4];
    ws.r0 = v[5];
    ws.y = v[6];
    ws.Ay = v[7];
    ws.s = v[8];
    ws.z = v[9];
    ws.t = v[10];
    ws.red.part = reinterpret_cast<double*>(base + NOALLOC_PART_OFF);
    ws.red.maxBlk = NOALLOC_PART_BLOCKS;
    ws.red.xpart = reinterpret_cast<long long*>(base + NOALLOC_XPART_OFF);
    ws.red.xnf = reinterpret_cast<int*>(base + NOALLOC_XNF_OFF);
    ws.red.result = reinterpret_cast<double*>(base + NOALLOC_RESULT_OFF);
    ws.red.maxBlkExact = RED_BLOCKS_EXACT;
    ws.scal = reinterpret_cast<DevScalars*>(base + NOALLOC_SCAL_OFF);
    vol_solver_performance pf;
    solveDevice(A, b, x, ct, pf, &ws);
    if (perf) *perf = pf;
    API_END
}

int vol_timing_enable(vol_context ctx, int on) {
    API_BEGIN
    if (!ctx) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null context");
    ctx->timing = on != 0;
    API_END
}

int vol_timing_read(vol_context ctx, double* amulTotalMs, int64_t* amulCount, int reset) {
    API_BEGIN
    if (!ctx) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null context");
    if (amulTotalMs) *amulTotalMs = ctx->amulMs;
    if (amulCount) *amulCount = ctx->amulCount;
    if (reset) {
        ctx->amulMs = 0;
        ctx->amulCount = 0;
    }
    API_END
}

}  // extern "C"

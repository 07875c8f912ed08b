// Batches of independent LDU systems solved together (vol_batch_*).
//
// The systems are concatenated: level 0 is the block-diagonal matrix of all systems
// (cells and faces offset per system), and every GAMG level l is the concatenation of
// the systems' own level-l matrices, each system keeping the hierarchy it has alone
// (the reference's stopping rule applies per system, GAMG.cpp:223-232). Within level l
// the systems that still smooth there come first, those whose coarsest level is l after
// them, so the smoothing kernels cover one prefix of the rows. One launch per phase then
// serves every system; the coarsest solves of all systems are one launch with one
// workgroup per system. Reductions are segmented by system (chunks that never straddle
// two systems, summed per system in order) and each system has its own Krylov scalars
// and convergence flag, so a system's arithmetic does not depend on its batch mates.
#include "host_common.hpp"

using namespace vol;
using namespace vol::host;

struct BLevel {
    int nRows = 0, nSmooth = 0, nFaces = 0, nFacesSmooth = 0;
    SellDev sell;                         // rows [0, nSmooth)
    DevBuf<int2> sliceSys;                // SELL: first/last system of each slice
    DevBuf<int> rowSys;                   // CSR: system of each row
    DevBuf<int> agg;                      // smoothing rows → next level (global)
    DevBuf<int> mOff, mList, iOff, iList, cfOff, cfList;
    DevBuf<int> l, u;                     // faces (global rows)
    DevBuf<double> diag, rD, Fu, Fl, coef, bufA, bufB, r, x;
    DevBuf<float> diag32, rD32, coef32, bufA32, bufB32, r32, x32;
};

struct vol_batch_s {
    vol_context ctx = nullptr;
    int nSys = 0;
    std::vector<int> n, nf, nLev;
    std::vector<int> cellOff, faceOff;  // level 0
    int nTotal = 0, nfTotal = 0, nLevMax = 0;
    std::vector<std::unique_ptr<BLevel>> L;
    DevBuf<char> ownerIsL;
    // segmented reductions over level 0
    DevBuf<int4> chunks;
    DevBuf<int> sysChunkOff, sysN;
    DevBuf<double> part;
    BatchRed red;
    // coarsest levels
    std::vector<CoarseDesc> cdesc;
    DevBuf<CoarseDesc> dCdesc;
    DevBuf<double> lu, luT;
    DevBuf<int> status;
    int maxCoarse = 0;
    DevBuf<const double*> rLev;
    DevBuf<double*> xLev;
    DevBuf<const float*> rLev32;
    DevBuf<float*> xLev32;
    // scalars: outer / inner (mixed precision) per system
    DevBuf<DevScalars> S, SI;
    DevBuf<int> nDone;
    int* hostDone = nullptr;  // pinned, 2 slots
    // matrix
    bool set = false, symmetric = true, mixedReady = false;
    DevBuf<double> diag, upper, lower;
    DevBuf<double> vr, vw, vp, vq, wA, pA, hb, hx;
    DevBuf<float> ir, iw, ip, iq, ie;
    std::vector<int> layoutHint;  // coarse layout per level (0 CSR, 1 SELL) when known
    // one PCG iteration captured as a HIP graph per key (plan, precision, buffers),
    // replayed every iteration; at most MAX_GRAPHS kept, the oldest dropped first
    static constexpr size_t MAX_GRAPHS = 4;
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    hipGraphExec_t graph = nullptr;  // the one of the running solve
    int graphLaunches = 0, eagerIterations = 0;  // statistics of the last solve
    ~vol_batch_s() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    }
};

namespace {

// ------------------------------------------------------------------ setup
// pre (nullable): per system a hierarchy built already (the same as buildHierarchy of
// its addressing and weights), used instead of building one
void buildBatch(vol_batch B, const int32_t* own, const int32_t* nei, const double* w,
                const std::vector<const std::vector<HLevel>*>* pre = nullptr) {
    hipStream_t st = B->ctx->stream;
    const int nS = B->nSys;
    std::vector<std::vector<HLevel>> built(pre ? 0 : nS);
    std::vector<const std::vector<HLevel>*> Hp(nS);
    auto H = [&](int s) -> const std::vector<HLevel>& { return *Hp[s]; };
    B->cellOff.assign(nS + 1, 0);
    B->faceOff.assign(nS + 1, 0);
    for (int s = 0; s < nS; ++s) {
        B->cellOff[s + 1] = B->cellOff[s] + B->n[s];
        B->faceOff[s + 1] = B->faceOff[s] + B->nf[s];
    }
    B->nTotal = B->cellOff[nS];
    B->nfTotal = B->faceOff[nS];
    for (int s = 0; s < nS; ++s) {
        const int fo = B->faceOff[s];
        for (int f = 0; f < B->nf[s]; ++f) {
            const int o = own[fo + f], q = nei[fo + f];
            if (o < 0 || o >= B->n[s] || q < 0 || q >= B->n[s] || o == q)
                throw VolError(VOL_ERR_INVALID_ARGUMENT, "system " + std::to_string(s) + ": bad owner/neighbour");
        }
        if (pre) {
            Hp[s] = (*pre)[s];
        } else {
            built[s] = buildHierarchy(B->n[s], B->nf[s], own + fo, nei + fo, w + fo, 2);
            Hp[s] = &built[s];
        }
        B->nLev[s] = int(H(s).size());
        if (B->nLev[s] < 2)
            throw VolError(VOL_ERR_INVALID_ARGUMENT,
                           "system " + std::to_string(s) + ": a batched system needs more than 64 cells");
        if (H(s).back().n > 4096)
            throw VolError(VOL_ERR_UNSUPPORTED, "system " + std::to_string(s) + ": coarsest level above 4096 cells");
        B->nLevMax = std::max(B->nLevMax, B->nLev[s]);
    }
    const int LM = B->nLevMax;
    // per-level order: systems that smooth at l, then systems whose coarsest level is l
    std::vector<std::vector<int>> order(LM), rowOff(LM, std::vector<int>(nS, -1)), faceOffL(LM, std::vector<int>(nS, -1));
    B->L.clear();
    for (int l = 0; l < LM; ++l) {
        auto Lv = std::make_unique<BLevel>();
        for (int pass = 0; pass < 2; ++pass)
            for (int s = 0; s < nS; ++s) {
                if (B->nLev[s] <= l) continue;
                const bool smooth = B->nLev[s] - 1 > l;
                if (smooth != (pass == 0)) continue;
                order[l].push_back(s);
                rowOff[l][s] = Lv->nRows;
                faceOffL[l][s] = Lv->nFaces;
                Lv->nRows += H(s)[l].n;
                Lv->nFaces += int(H(s)[l].l.size());
                if (smooth) {
                    Lv->nSmooth = Lv->nRows;
                    Lv->nFacesSmooth = Lv->nFaces;
                }
            }
        B->L.push_back(std::move(Lv));
    }
    // level 0 keeps the callers' order (every system smooths there)
    for (int s = 0; s < nS; ++s)
        if (rowOff[0][s] != B->cellOff[s]) throw VolError(VOL_ERR_INVALID_ARGUMENT, "internal: level-0 order");
    for (int l = 0; l < LM; ++l) {
        BLevel& Lv = *B->L[l];
        std::vector<int> gl(Lv.nFaces), gu(Lv.nFaces);
        for (int s : order[l]) {
            const HLevel& F = H(s)[l];
            for (size_t f = 0; f < F.l.size(); ++f) {
                gl[faceOffL[l][s] + f] = F.l[f] + rowOff[l][s];
                gu[faceOffL[l][s] + f] = F.u[f] + rowOff[l][s];
            }
        }
        Lv.l.upload(gl, st);
        Lv.u.upload(gu, st);
        // rows of the smoothing prefix
        std::vector<int> off(Lv.nSmooth + 1, 0), col, src;
        if (l == 0) {
            for (int s = 0; s < nS; ++s)
                for (int f = 0; f < B->nf[s]; ++f) {
                    ++off[own[B->faceOff[s] + f] + B->cellOff[s] + 1];
                    ++off[nei[B->faceOff[s] + f] + B->cellOff[s] + 1];
                }
            for (int c = 0; c < Lv.nSmooth; ++c) off[c + 1] += off[c];
            col.resize(off[Lv.nSmooth]);
            src.resize(off[Lv.nSmooth]);
            std::vector<int> fill(off.begin(), off.end() - 1);
            for (int s = 0; s < nS; ++s)
                for (int f = 0; f < B->nf[s]; ++f) {
                    const int gf = B->faceOff[s] + f;
                    const int o = own[gf] + B->cellOff[s], q = nei[gf] + B->cellOff[s];
                    int k = fill[o]++;
                    col[k] = q;
                    src[k] = (gf << 1) | 0;
                    k = fill[q]++;
                    col[k] = o;
                    src[k] = (gf << 1) | 1;
                }
            buildSell(Lv.nSmooth, off, col, src, Lv.sell, st, true, false, 0, B->ctx->sellPairs);
            std::vector<char> oil(B->nfTotal);
            for (int s = 0; s < nS; ++s)
                for (int f = 0; f < B->nf[s]; ++f)
                    oil[B->faceOff[s] + f] = char(own[B->faceOff[s] + f] < nei[B->faceOff[s] + f]);
            B->ownerIsL.upload(oil, st);
        } else if (Lv.nSmooth > 0) {
            for (int s : order[l]) {
                if (B->nLev[s] - 1 <= l) continue;
                const HLevel& F = H(s)[l];
                for (int i = 0; i < F.n; ++i) off[rowOff[l][s] + i + 1] = F.off[i + 1] - F.off[i];
            }
            for (int c = 0; c < Lv.nSmooth; ++c) off[c + 1] += off[c];
            col.resize(off[Lv.nSmooth]);
            src.resize(off[Lv.nSmooth]);
            for (int s : order[l]) {
                if (B->nLev[s] - 1 <= l) continue;
                const HLevel& F = H(s)[l];
                for (int i = 0; i < F.n; ++i) {
                    int k = off[rowOff[l][s] + i];
                    for (int q = F.off[i]; q < F.off[i + 1]; ++q, ++k) {
                        col[k] = F.adjCell[q] + rowOff[l][s];
                        src[k] = ((F.adjFace[q] + faceOffL[l][s]) << 1) | (F.adjIsL[q] ? 0 : 1);
                    }
                }
            }
            buildCoarseLayout(Lv.nSmooth, off, col, src, Lv.sell, st,
                              l < int(B->layoutHint.size()) ? B->layoutHint[l] : -1, nullptr, 0, B->ctx->sellPairs);
        }
        // system of every smoothing row, for skipping finished systems
        if (Lv.nSmooth > 0) {
            std::vector<int> sysOf(Lv.nSmooth);
            for (int s : order[l]) {
                if (B->nLev[s] - 1 <= l) continue;
                for (int i = 0; i < H(s)[l].n; ++i) sysOf[rowOff[l][s] + i] = s;
            }
            if (Lv.sell.csrG > 0) {
                Lv.rowSys.upload(sysOf, st);
            } else {
                std::vector<int2> ss(Lv.sell.nslices);
                for (int q = 0; q < Lv.sell.nslices; ++q) {
                    const int r0 = q * WAVE, r1 = std::min(Lv.nSmooth, r0 + WAVE) - 1;
                    ss[q] = make_int2(std::min(sysOf[r0], sysOf[r1]), std::max(sysOf[r0], sysOf[r1]));
                }
                Lv.sliceSys.upload(ss, st);
            }
        }
        // transfer to level l+1
        if (l + 1 < LM && Lv.nSmooth > 0) {
            const BLevel& C = *B->L[l + 1];
            std::vector<int> agg(Lv.nRows, -1), fmap(Lv.nFaces, -1);
            std::vector<char> flip(Lv.nFaces, 0);
            for (int s : order[l]) {
                if (B->nLev[s] - 1 <= l) continue;
                const HLevel& F = H(s)[l];
                for (int i = 0; i < F.n; ++i) agg[rowOff[l][s] + i] = F.agg[i] + rowOff[l + 1][s];
                for (size_t f = 0; f < F.faceMap.size(); ++f) {
                    const int cf = F.faceMap[f];
                    fmap[faceOffL[l][s] + f] = cf < 0 ? -1 : cf + faceOffL[l + 1][s];
                    flip[faceOffL[l][s] + f] = F.faceFlip[f];
                }
            }
            const GatherLists G = buildGatherLists(Lv.nRows, C.nRows, agg, gl, fmap, flip, C.nFaces);
            G.upload(Lv, st);
            agg.resize(Lv.nSmooth);
            Lv.agg.upload(agg, st);

        }
        // coefficient and work storage
        if (l > 0) {
            Lv.diag.alloc(Lv.nRows);
            Lv.rD.alloc(Lv.nRows);
            Lv.r.alloc(Lv.nRows);
            Lv.x.alloc(Lv.nRows);
            if (Lv.nSmooth > 0) Lv.coef.alloc(size_t(Lv.sell.nslots) + Lv.sell.nslotsP);  // (+ the pair copy)
        }
        Lv.Fu.alloc(Lv.nFaces);
        Lv.Fl.alloc(Lv.nFaces);
        Lv.bufA.alloc(Lv.nSmooth);
        Lv.bufB.alloc(Lv.nSmooth);
    }
    // coarsest levels
    B->cdesc.assign(nS, CoarseDesc{});
    size_t luSize = 0;
    for (int s = 0; s < nS; ++s) {
        const int c = B->nLev[s] - 1;
        BLevel& Lv = *B->L[c];
        CoarseDesc& d = B->cdesc[s];
        d.n = H(s)[c].n;
        d.nf = int(H(s)[c].l.size());
        d.rowBase = rowOff[c][s];
        d.level = c;
        d.luOff = luSize;
        d.l = Lv.l.p + faceOffL[c][s];
        d.u = Lv.u.p + faceOffL[c][s];
        d.d = Lv.diag.p;
        d.cu = Lv.Fu.p + faceOffL[c][s];
        d.cl = Lv.Fl.p + faceOffL[c][s];
        luSize += size_t(d.n) * d.n;
        B->maxCoarse = std::max(B->maxCoarse, d.n);
    }
    B->dCdesc.upload(B->cdesc, st);
    B->lu.alloc(luSize);
    B->luT.alloc(luSize);
    B->status.alloc(nS);
    std::vector<const double*> rl(LM, nullptr);
    std::vector<double*> xl(LM, nullptr);
    for (int l = 1; l < LM; ++l) {
        rl[l] = B->L[l]->r.p;
        xl[l] = B->L[l]->x.p;
    }
    B->rLev.upload(rl, st);
    B->xLev.upload(xl, st);
    // level-0 chunks
    std::vector<int4> ch;
    std::vector<int> sco(nS + 1, 0);
    for (int s = 0; s < nS; ++s) {
        for (int r0 = 0; r0 < B->n[s]; r0 += BATCH_CHUNK)
            ch.push_back(make_int4(s, B->cellOff[s] + r0, B->cellOff[s] + std::min(B->n[s], r0 + BATCH_CHUNK), 0));
        sco[s + 1] = int(ch.size());
    }
    B->chunks.upload(ch, st);
    B->sysChunkOff.upload(sco, st);
    B->sysN.upload(B->n, st);
    B->part.alloc(size_t(MAX_RED) * ch.size());
    B->red.chunks = B->chunks.p;
    B->red.nChunks = int(ch.size());
    B->red.sysChunkOff = B->sysChunkOff.p;
    B->red.nSys = nS;
    B->red.part = B->part.p;
    B->red.maxChunks = int(ch.size());
    B->S.alloc(nS);
    B->SI.alloc(nS);
    B->nDone.alloc(1);
    HIPCHECK(hipHostMalloc((void**)&B->hostDone, 2 * sizeof(int), hipHostMallocDefault));
    HIPCHECK(hipStreamSynchronize(st));
}

void toFloatB(hipStream_t st, const double* a, float* w, size_t n) {
    launch_mix(st, MIX_TO32, int(n), a, w, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// Coarse-level coefficients (Galerkin), SELL slot values, 1/diag and the coarsest LUs.
void buildCoeffs(vol_batch B) {
    hipStream_t st = B->ctx->stream;
    const int LM = B->nLevMax;
    BLevel& L0 = *B->L[0];
    L0.coef.alloc(size_t(L0.sell.nslots) + L0.sell.nslotsP);
    launch_sell_coef(st, L0.sell.view(), B->upper.p, B->symmetric ? B->upper.p : B->lower.p, L0.coef.p, nullptr);
    L0.rD.alloc(B->nTotal);
    VecArgs v;
    v.n = B->nTotal;
    v.a = B->diag.p;
    v.w = L0.rD.p;
    launch_vec(st, V_RECIP, v, false, B->ctx->red);
    launch_orient0(st, B->nfTotal, B->ownerIsL.p, B->upper.p, B->symmetric ? B->upper.p : B->lower.p, L0.Fu.p, L0.Fl.p);
    for (int l = 0; l + 1 < LM; ++l) {
        BLevel& F = *B->L[l];
        BLevel& C = *B->L[l + 1];
        if (F.nSmooth == 0) break;
        const double* Fd = l == 0 ? B->diag.p : F.diag.p;
        launch_galerkin(st, C.nRows, C.nFaces, F.mOff.p, F.mList.p, F.iOff.p, F.iList.p, F.cfOff.p, F.cfList.p, Fd,
                        F.Fu.p, F.Fl.p, C.diag.p, C.Fu.p, C.Fl.p);
        if (C.nSmooth > 0) launch_sell_coef(st, C.sell.view(), C.Fu.p, C.Fl.p, C.coef.p, nullptr);
        VecArgs w;
        w.n = C.nRows;
        w.a = C.diag.p;
        w.w = C.rD.p;
        launch_vec(st, V_RECIP, w, false, B->ctx->red);
    }
    launch_bcoarse_factor(st, B->dCdesc.p, B->nSys, B->lu.p, B->luT.p, B->status.p);
    std::vector<int> status(B->nSys);
    HIPCHECK(hipMemcpyAsync(status.data(), B->status.p, sizeof(int) * B->nSys, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int s = 0; s < B->nSys; ++s)
        if (status[s] != 0)
            throw VolError(VOL_ERR_SINGULAR_COARSEST, "GAMG: singular coarsest matrix in system " + std::to_string(s));
    B->mixedReady = false;
}

void buildMixedB(vol_batch B) {
    hipStream_t st = B->ctx->stream;
    for (int l = 0; l < B->nLevMax; ++l) {
        BLevel& Lv = *B->L[l];
        const double* d = l == 0 ? B->diag.p : Lv.diag.p;
        const size_t nr = l == 0 ? size_t(B->nTotal) : size_t(Lv.nRows);
        Lv.diag32.alloc(nr);
        Lv.rD32.alloc(nr);
        toFloatB(st, d, Lv.diag32.p, nr);
        toFloatB(st, Lv.rD.p, Lv.rD32.p, nr);
        if (Lv.nSmooth > 0) {
            Lv.coef32.alloc(size_t(Lv.sell.nslots) + Lv.sell.nslotsP);
            toFloatB(st, Lv.coef.p, Lv.coef32.p, size_t(Lv.sell.nslots) + Lv.sell.nslotsP);
        }
        Lv.bufA32.alloc(Lv.nSmooth);
        Lv.bufB32.alloc(Lv.nSmooth);
        if (l > 0) {
            Lv.r32.alloc(Lv.nRows);
            Lv.x32.alloc(Lv.nRows);
        }
    }
    std::vector<const float*> rl(B->nLevMax, nullptr);
    std::vector<float*> xl(B->nLevMax, nullptr);
    for (int l = 1; l < B->nLevMax; ++l) {
        rl[l] = B->L[l]->r32.p;
        xl[l] = B->L[l]->x32.p;
    }
    B->rLev32.upload(rl, st);
    B->xLev32.upload(xl, st);
    HIPCHECK(hipStreamSynchronize(st));
    B->mixedReady = true;
}

// ------------------------------------------------------------------ solve
struct BPlan {
    std::string precond;
    int nPre = 2, nPost = 2;
    double omega = 0.8;
};

template <class T> struct BLevelPtrs {
    const T *coef, *diag, *rD;
    T *bufA, *bufB, *r, *x;
};
BLevelPtrs<double> blevel(vol_batch B, int l, double*) {
    BLevel& L = *B->L[l];
    return {L.coef.p, l == 0 ? B->diag.p : L.diag.p, L.rD.p, L.bufA.p, L.bufB.p, L.r.p, L.x.p};
}
BLevelPtrs<float> blevel(vol_batch B, int l, float*) {
    BLevel& L = *B->L[l];
    return {L.coef32.p, L.diag32.p, L.rD32.p, L.bufA32.p, L.bufB32.p, L.r32.p, L.x32.p};
}

void bsellLaunch(hipStream_t st, int op, int xl, const SellArgs& a) { launch_sell(st, op, xl, a); }
void bsellLaunch(hipStream_t st, int op, int xl, const SellArgs32& a) { launch_sell32(st, op, xl, a); }

template <class T>
void bsell(hipStream_t st, int op, int xl, const SellView& V, const T* coef, const T* diag, const T* rD, const T* x,
           const T* r, T* out, double omega = 1.0, const T* xc = nullptr, const int* agg = nullptr,
           const BLevel* skip = nullptr, const DevScalars* S = nullptr) {
    SellArgsT<T> a;
    if (skip && S) {
        a.sliceSys = skip->sliceSys.p;
        a.rowSys = skip->rowSys.p;
        a.sysS = S;
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
    bsellLaunch(st, op, xl, a);
}

void brestrict(hipStream_t st, int nc, const BLevel& F, const double* res, double* rc) {
    launch_restrict(st, nc, F.mOff.p, F.mList.p, res, rc, nullptr);
}
void brestrict(hipStream_t st, int nc, const BLevel& F, const float* res, float* rc) {
    launch_restrict32(st, nc, F.mOff.p, F.mList.p, res, rc, nullptr);
}
void bcoarse(vol_batch B, const double*, const DevScalars* S) {
    launch_bcoarse_solve(B->ctx->stream, B->dCdesc.p, B->nSys, B->maxCoarse, B->lu.p, B->luT.p, B->rLev.p, B->xLev.p,
                         S);
}
void bcoarse(vol_batch B, const float*, const DevScalars* S) {
    launch_bcoarse_solve32(B->ctx->stream, B->dCdesc.p, B->nSys, B->maxCoarse, B->lu.p, B->luT.p, B->rLev32.p,
                           B->xLev32.p, S);
}

// V-cycle of every system at once: r (level 0) → x. Same per-system operations as the
// single-system V-cycle without the one-workgroup tail.
template <class T>
void bvcycle(vol_batch B, const BPlan& P, const T* r0, T* x0, const DevScalars* S) {
    hipStream_t st = B->ctx->stream;
    const int LM = B->nLevMax;
    std::vector<int> curL(LM, 0);
    int depth = 0;  // levels with smoothing rows
    for (int l = 0; l + 1 < LM && B->L[l]->nSmooth > 0; ++l) {
        BLevel& Lv = *B->L[l];
        const SellView V = Lv.sell.view();
        const BLevelPtrs<T> p = blevel(B, l, (T*)nullptr);
        const T* r = l == 0 ? r0 : p.r;
        T* bufs[2] = {p.bufA, p.bufB};
        int cur = 0;
        for (int s = 1; s < P.nPre; ++s) {
            const int dst = (s == 1) ? 0 : 1 - cur;
            bsell<T>(st, OP_JACOBI, s == 1 ? XL_FROMZERO : XL_PLAIN, V, p.coef, p.diag, p.rD,
                     s == 1 ? nullptr : bufs[cur], r, bufs[dst], P.omega, nullptr, nullptr, &Lv, S);
            cur = dst;
        }
        const int freeB = 1 - cur;
        bsell<T>(st, OP_RESID, P.nPre == 1 ? XL_FROMZERO : XL_PLAIN, V, p.coef, p.diag, p.rD,
                 P.nPre == 1 ? nullptr : bufs[cur], r, bufs[freeB], P.omega, nullptr, nullptr, &Lv, S);
        const BLevelPtrs<T> pc = blevel(B, l + 1, (T*)nullptr);
        brestrict(st, B->L[l + 1]->nRows, Lv, bufs[freeB], pc.r);
        curL[l] = cur;
        depth = l + 1;
    }
    bcoarse(B, r0, S);
    for (int l = depth - 1; l >= 0; --l) {
        BLevel& Lv = *B->L[l];
        const SellView V = Lv.sell.view();
        const BLevelPtrs<T> p = blevel(B, l, (T*)nullptr);
        const BLevelPtrs<T> pc = blevel(B, l + 1, (T*)nullptr);
        const T* r = l == 0 ? r0 : p.r;
        T* xout = l == 0 ? x0 : p.x;
        T* bufs[2] = {p.bufA, p.bufB};
        int src = curL[l], other = 1 - curL[l];
        for (int s = 0; s < P.nPost; ++s) {
            T* dst = (s == P.nPost - 1) ? xout : bufs[other];
            if (s == 0)
                bsell<T>(st, OP_JACOBI, P.nPre == 1 ? XL_ZEROPROLONG : XL_PROLONG, V, p.coef, p.diag, p.rD,
                         P.nPre == 1 ? nullptr : bufs[src], r, dst, P.omega, pc.x, Lv.agg.p, &Lv, S);
            else
                bsell<T>(st, OP_JACOBI, XL_PLAIN, V, p.coef, p.diag, p.rD, bufs[src], r, dst, P.omega, nullptr,
                         nullptr, &Lv, S);
            if (s != P.nPost - 1) std::swap(src, other);
        }
    }
}

template <class T>
void bvec(vol_batch B, int op, DevScalars* S, const T* a, const T* b = nullptr, const T* c = nullptr,
          T* w = nullptr, T* x = nullptr, T* y = nullptr, const T* rD = nullptr) {
    VecArgsT<T> v;
    v.n = B->nTotal;
    v.a = a;
    v.b = b;
    v.c = c;
    v.w = w;
    v.x = x;
    v.y = y;
    v.rD = rD;
    v.S = S;
    if constexpr (std::is_same_v<T, double>)
        launch_bvec(B->ctx->stream, op, v, B->red);
    else
        launch_bvec32(B->ctx->stream, op, v, B->red);
}

template <class T> struct BKrylov {
    T *x, *r, *w, *p, *q;
};

// One PCG iteration of every system that is not done (LinearSolver.cpp:176-201).
template <class T>
void bpcgIteration(vol_batch B, const BPlan& P, const BKrylov<T>& k, DevScalars* S) {
    hipStream_t st = B->ctx->stream;
    const BLevelPtrs<T> p0 = blevel(B, 0, (T*)nullptr);
    if (P.precond == "GAMG") {
        bvcycle<T>(B, P, k.r, k.w, S);
        bvec<T>(B, V_DOT, S, k.w, k.r);
    } else {
        bvec<T>(B, V_PRECOND_DOT, S, k.r, nullptr, nullptr, k.w, nullptr, nullptr,
                P.precond == "diagonal" ? p0.rD : nullptr);
    }
    launch_bfinalize(st, FIN_PCG_WRA, 1, B->red, S, nullptr);
    bvec<T>(B, V_PCG_P, S, k.w, nullptr, nullptr, nullptr, k.p);
    bsell<T>(st, OP_AMUL, XL_PLAIN, B->L[0]->sell.view(), p0.coef, p0.diag, nullptr, k.p, nullptr, k.q, 1.0, nullptr,
             nullptr, B->L[0].get(), S);
    bvec<T>(B, V_DOT, S, k.q, k.p);
    launch_bfinalize(st, FIN_PCG_WAPA, 1, B->red, S, nullptr);
    bvec<T>(B, V_PCG_XR, S, k.p, k.q, nullptr, nullptr, k.x, k.r);
    launch_bfinalize(st, FIN_PCG_NORM, 1, B->red, S, nullptr);
    launch_bcount(st, S, B->nSys, B->nDone.p);
}

// Whether batched iterations replay a captured graph (VOLUMA_BATCH_GRAPH=0: launch
// every kernel of every iteration instead; both give the same bits).
bool useGraph() {
    const char* e = std::getenv("VOLUMA_BATCH_GRAPH");
    return !(e && std::string(e) == "0");
}

// The iteration `iterate` captured into B->graph (once per key). Every launch in it is
// a kernel predicated on the systems' device flags, so the same graph serves every
// iteration; nothing in it is a memset or copy node.
template <class F> void captureIteration(vol_batch B, const std::string& key, F&& iterate) {
    for (auto& g : B->graphs)
        if (g.first == key) {
            B->graph = g.second;
            return;
        }
    if (B->graphs.size() >= vol_batch_s::MAX_GRAPHS) {
        (void)hipGraphExecDestroy(B->graphs.front().second);
        B->graphs.erase(B->graphs.begin());
    }
    B->graph = nullptr;
    hipStream_t st = B->ctx->stream;
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        iterate();
    } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCHECK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
    B->graphs.push_back({key, ex});
    B->graph = ex;
}

// Runs `iterate` until every system is done, one iteration ahead of the host. With a
// graph key, each iteration is a replay of the captured iteration.
template <class F> void bloop(vol_batch B, int maxLoops, F&& iterate, const std::string& graphKey = "") {
    hipStream_t st = B->ctx->stream;
    hipEvent_t evs[2];
    HIPCHECK(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
    int* hd = B->hostDone;
    bool graph = false;  // the first iteration runs eagerly (it makes any lazy allocations)
    auto step = [&] {
        if (graph) {
            HIPCHECK(hipGraphLaunch(B->graph, st));
            ++B->graphLaunches;
        } else {
            iterate();
            ++B->eagerIterations;
        }
    };
    step();
    HIPCHECK(hipMemcpyAsync(&hd[0], B->nDone.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipEventRecord(evs[0], st));
    if (!graphKey.empty()) {
        captureIteration(B, graphKey, iterate);
        graph = true;
    }
    for (int k = 1; k <= maxLoops; ++k) {
        step();
        HIPCHECK(hipMemcpyAsync(&hd[k & 1], B->nDone.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(evs[k & 1], st));
        HIPCHECK(hipEventSynchronize(evs[(k - 1) & 1]));
        if (hd[(k - 1) & 1] == B->nSys) break;
    }
    HIPCHECK(hipStreamSynchronize(st));
    (void)hipEventDestroy(evs[0]);
    (void)hipEventDestroy(evs[1]);
}

// Key of a captured iteration: everything its launches depend on (the plan and the
// buffers they name; x is the caller's solution vector in FP64).
std::string graphKey(vol_batch B, const BPlan& P, const char* prec, const double* x) {
    if (!useGraph()) return "";
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s %s %d %d %.17g %p %p %p %p %p %p %p %p %p", prec, P.precond.c_str(), P.nPre,
                  P.nPost, P.omega, (const void*)x, (void*)B->vr.p, (void*)B->vw.p, (void*)B->vp.p, (void*)B->vq.p,
                  (void*)B->ir.p, (void*)B->ie.p, (void*)B->ip.p, (void*)B->iq.p);
    return buf;
}

int readDone(vol_batch B, const DevScalars* S) {
    hipStream_t st = B->ctx->stream;
    launch_bcount(st, S, B->nSys, B->nDone.p);
    HIPCHECK(hipMemcpyAsync(B->hostDone, B->nDone.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return B->hostDone[0];
}

void bsolve(vol_batch B, const double* b, double* x, const vol_solver_controls& ct, vol_solver_performance* perf) {
    hipStream_t st = B->ctx->stream;
    const int n = B->nTotal;
    BPlan P;
    const std::string solver = ct.solver ? ct.solver : "PCG";
    P.precond = ct.preconditioner ? ct.preconditioner : "diagonal";
    if (solver != "PCG") throw VolError(VOL_ERR_UNSUPPORTED, "batched solves provide PCG only");
    if (!B->symmetric) throw VolError(VOL_ERR_PCG_NONSYMMETRIC, "PCG requires symmetric matrices");
    if (P.precond != "GAMG" && P.precond != "diagonal" && P.precond != "none") {
        if (P.precond == "DIC" || P.precond == "DILU")
            throw VolError(VOL_ERR_UNSUPPORTED, "batched solves provide the GAMG, diagonal and none preconditioners");
        throw VolError(VOL_ERR_UNKNOWN_PRECONDITIONER, "unknown preconditioner " + P.precond);
    }
    const int sums = ct.sums == VOL_SUMS_CONTEXT ? (B->ctx->reproducible ? VOL_SUMS_REPRODUCIBLE : VOL_SUMS_FAST)
                                                  : ct.sums;
    if (sums != VOL_SUMS_FAST) throw VolError(VOL_ERR_UNSUPPORTED, "batched solves use fast sums only");
    if (ct.precision != VOL_FP64 && ct.precision != VOL_MIXED)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "precision must be VOL_FP64 or VOL_MIXED");
    const bool mixed = ct.precision == VOL_MIXED;
    if (mixed && !(ct.innerRelTol > 0.0 && ct.innerRelTol < 1.0))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "innerRelTol must lie in (0, 1)");
    P.nPre = ct.gamgPreSweeps;
    P.nPost = ct.gamgPostSweeps;
    P.omega = ct.gamgOmega;
    if (P.precond == "GAMG" && (P.nPre < 1 || P.nPost < 1 || !(P.omega > 0.0)))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "GAMG smoother needs ≥ 1 pre/post sweep and ω > 0");
    if (P.precond == "GAMG" && ct.gamgSmoother != VOL_GAMG_JACOBI)
        throw VolError(VOL_ERR_UNSUPPORTED, "batched solves smooth GAMG levels with damped Jacobi only");
    if (ct.maxIter < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad maxIter");
    if (mixed && !B->mixedReady) buildMixedB(B);
    DevScalars* S = B->S.p;
    launch_binit_scalars(st, S, B->nSys, B->sysN.p, ct.tolerance, ct.relTol, ct.maxIter, ct.minIter, ct.nSweeps,
                         mixed ? ct.innerRelTol : 0.0, mixed ? 1 : 0);
    for (DevBuf<double>* v : {&B->vr, &B->vw, &B->vp, &B->vq, &B->wA, &B->pA}) v->alloc(n);
    const BLevel& L0 = *B->L[0];
    const SellView V = L0.sell.view();
    // prologue (LinearSolver.cpp:150-169), per system
    bsell<double>(st, OP_AMUL, XL_PLAIN, V, L0.coef.p, B->diag.p, nullptr, x, nullptr, B->wA.p);
    bvec<double>(B, V_SUM_X, S, x);
    launch_bfinalize(st, FIN_XREF, 1, B->red, S, nullptr);
    bvec<double>(B, V_FILL_XREF, S, nullptr, nullptr, nullptr, B->vq.p);
    bsell<double>(st, OP_AMUL, XL_PLAIN, V, L0.coef.p, B->diag.p, nullptr, B->vq.p, nullptr, B->pA.p);
    bvec<double>(B, V_INIT, S, B->wA.p, b, B->pA.p, B->vr.p);
    launch_bfinalize(st, FIN_INIT, 2, B->red, S, nullptr);
    const int maxLoops = std::max(ct.maxIter, 1) + 2;
    if (readDone(B, S) < B->nSys) {
        if (!mixed) {
            BKrylov<double> k{x, B->vr.p, B->vw.p, B->vp.p, B->vq.p};
            bloop(B, maxLoops, [&] { bpcgIteration<double>(B, P, k, S); }, graphKey(B, P, "fp64", x));
        } else {
            for (DevBuf<float>* v : {&B->ir, &B->iw, &B->ip, &B->iq, &B->ie}) v->alloc(n);
            DevScalars* SI = B->SI.p;
            BKrylov<float> k{B->ie.p, B->ir.p, B->iw.p, B->ip.p, B->iq.p};
            launch_bfinalize(st, FIN_INNER_START, 0, B->red, S, SI);
            launch_bmix(st, MIX_TO32, B->red, S, B->vr.p, B->ir.p, nullptr, B->ie.p, nullptr, nullptr);
            for (int outer = 0; outer < maxLoops; ++outer) {
                bloop(B, maxLoops, [&] { bpcgIteration<float>(B, P, k, SI); }, graphKey(B, P, "mixed", nullptr));
                launch_bmix(st, MIX_ADD, B->red, S, nullptr, nullptr, nullptr, nullptr, x, B->ie.p);
                bsell<double>(st, OP_AMUL, XL_PLAIN, V, L0.coef.p, B->diag.p, nullptr, x, nullptr, B->wA.p, 1.0,
                              nullptr, nullptr, &L0, S);
                bvec<double>(B, V_BMINUSAX_NORM, S, B->wA.p, b, nullptr, B->vr.p);
                launch_bfinalize(st, FIN_OUTER, 1, B->red, S, SI);
                launch_bfinalize(st, FIN_INNER_START, 0, B->red, S, SI);
                launch_bmix(st, MIX_TO32, B->red, S, B->vr.p, B->ir.p, nullptr, B->ie.p, nullptr, nullptr);
                if (readDone(B, S) == B->nSys) break;
            }
        }
        if (readDone(B, S) < B->nSys) throw VolError(VOL_ERR_HIP, "batched solver loop did not terminate");
    }
    std::vector<DevScalars> hs(B->nSys);
    HIPCHECK(hipMemcpyAsync(hs.data(), S, sizeof(DevScalars) * B->nSys, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (perf)
        for (int s = 0; s < B->nSys; ++s) {
            vol_solver_performance& p = perf[s];
            std::memset(&p, 0, sizeof(p));
            std::strncpy(p.solver, "PCG", sizeof(p.solver) - 1);
            const DevScalars& h = hs[s];
            p.initialResidual = h.initialResidual;
            p.finalResidual = h.finalResidual;
            p.residualSum = h.residualSum;
            p.normFactor = h.normFactor;
            p.iterations = h.iterations;
            const bool conv = h.iterations >= ct.minIter &&
                              (h.finalResidual < ct.tolerance ||
                               (ct.relTol > 0 && h.finalResidual < ct.relTol * h.initialResidual));
            p.converged = conv ? 1 : 0;
            p.outerIterations = h.outer;
        }
}

}  // namespace

// (declared with ctxRelease in host_common.hpp, outside the namespaces)
vol_batch batchOfCopies(vol_context ctx, vol_mesh m, int K) {
    const MeshView v = meshView(m);
    if (!v.H) throw VolError(VOL_ERR_NO_GAMG_WEIGHTS, "batched copies need a mesh with a GAMG hierarchy");
    if (K < 1 || int64_t(v.n) * K > (int64_t(1) << 30) || int64_t(v.nf) * K > (int64_t(1) << 29))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad number of copies");
    auto B = std::make_unique<vol_batch_s>();
    B->ctx = ctx;
    B->nSys = K;
    B->n.assign(K, v.n);
    B->nf.assign(K, v.nf);
    B->nLev.assign(K, 0);
    std::vector<int32_t> own(size_t(v.nf) * K), nei(size_t(v.nf) * K);
    for (int s = 0; s < K; ++s) {
        std::copy(v.own, v.own + v.nf, own.begin() + size_t(s) * v.nf);
        std::copy(v.nei, v.nei + v.nf, nei.begin() + size_t(s) * v.nf);
    }
    const std::vector<const std::vector<HLevel>*> pre(K, v.H);
    B->layoutHint = v.layout;  // the mesh's coarse layouts suit its copies
    HIPCHECK(hipSetDevice(ctx->device));
    buildBatch(B.get(), own.data(), nei.data(), nullptr, &pre);
    ctx->refs.fetch_add(1);
    return B.release();
}

void batchSetSystems(vol_batch B, const std::vector<const double*>& diag, const std::vector<const double*>& upper) {
    if (int(diag.size()) != B->nSys || int(upper.size()) != B->nSys)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "one matrix per system");
    HIPCHECK(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    B->symmetric = true;
    B->diag.alloc(B->nTotal);
    B->upper.alloc(B->nfTotal);
    for (int s = 0; s < B->nSys; ++s) {
        double* d = B->diag.p + B->cellOff[s];
        double* u = B->upper.p + B->faceOff[s];
        const size_t n = size_t(B->n[s]), nf = size_t(B->nf[s]);
        if (s > 0 && diag[s] == diag[s - 1] && upper[s] == upper[s - 1] && B->n[s] == B->n[s - 1]) {
            // the same matrix as the previous system: copy on the device
            HIPCHECK(hipMemcpyAsync(d, B->diag.p + B->cellOff[s - 1], sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            if (nf)
                HIPCHECK(hipMemcpyAsync(u, B->upper.p + B->faceOff[s - 1], sizeof(double) * nf,
                                        hipMemcpyDeviceToDevice, st));
        } else {
            HIPCHECK(hipMemcpyAsync(d, diag[s], sizeof(double) * n, hipMemcpyHostToDevice, st));
            if (nf) HIPCHECK(hipMemcpyAsync(u, upper[s], sizeof(double) * nf, hipMemcpyHostToDevice, st));
        }
    }
    B->set = false;
    buildCoeffs(B);
    B->set = true;
}

extern "C" {

int vol_batch_create(vol_context ctx, int nSystems, const int32_t* nCells, const int32_t* nInternalFaces,
                     const int32_t* owner, const int32_t* neighbour, const double* gamgWeights, vol_batch* out) {
    API_BEGIN
    if (!ctx || !out || nSystems < 1 || !nCells || !nInternalFaces || !owner || !neighbour || !gamgWeights)
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument or no systems");
    auto B = std::make_unique<vol_batch_s>();
    B->ctx = ctx;
    B->nSys = nSystems;
    B->n.assign(nCells, nCells + nSystems);
    B->nf.assign(nInternalFaces, nInternalFaces + nSystems);
    B->nLev.assign(nSystems, 0);
    int64_t tot = 0, totf = 0;
    for (int s = 0; s < nSystems; ++s) {
        if (B->n[s] < 1 || B->nf[s] < 0) throw VolError(VOL_ERR_INVALID_ARGUMENT, "bad system size");
        tot += B->n[s];
        totf += B->nf[s];
    }
    if (tot > (int64_t(1) << 30) || totf > (int64_t(1) << 29))
        throw VolError(VOL_ERR_INVALID_ARGUMENT, "batch too large for int32 rows/faces");
    HIPCHECK(hipSetDevice(ctx->device));
    buildBatch(B.get(), owner, neighbour, gamgWeights);
    ctx->refs.fetch_add(1);
    *out = B.release();
    API_END
}

int vol_batch_destroy(vol_batch B) {
    API_BEGIN
    if (!B) return VOL_OK;
    vol_context c = B->ctx;
    (void)hipSetDevice(c->device);
    if (B->hostDone) (void)hipHostFree(B->hostDone);
    delete B;
    ctxRelease(c);
    API_END
}

int vol_batch_graph_stats(vol_batch B, int* graphLaunches, int* eagerIterations, int reset) {
    API_BEGIN
    if (!B) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null batch");
    if (graphLaunches) *graphLaunches = B->graphLaunches;
    if (eagerIterations) *eagerIterations = B->eagerIterations;
    if (reset) B->graphLaunches = B->eagerIterations = 0;
    API_END
}

int vol_batch_info(vol_batch B, int64_t* nCellsTotal, int64_t* nFacesTotal, int32_t* levels) {
    API_BEGIN
    if (!B) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null batch");
    if (nCellsTotal) *nCellsTotal = B->nTotal;
    if (nFacesTotal) *nFacesTotal = B->nfTotal;
    if (levels)
        for (int s = 0; s < B->nSys; ++s) levels[s] = B->nLev[s];
    API_END
}

int vol_batch_set(vol_batch B, const double* diag, const double* upper, const double* lower, int location) {
    API_BEGIN
    if (!B || !diag || !upper) throw VolError(VOL_ERR_INVALID_ARGUMENT, "null argument");
    HIPCHECK(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    B->symmetric = (lower == nullptr || lower == upper);
    B->diag.alloc(B->nTotal);
    B->upper.alloc(B->nfTotal);
    const hipMemcpyKind kind = location == VOL_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIPCHECK(hipMemcpyAsync(B->diag.p, diag, sizeof(double) * B->nTotal, kind, st));
    if (B->nfTotal) HIPCHECK(hipMemcpyAsync(B->upper.p, upper, sizeof(double) * B->nfTotal, kind, st));
    if (!B->symmetric) {
        B->lower.alloc(B->nfTotal);
        if (B->nfTotal) HIPCHECK(hipMemcpyAsync(B->lower.p, lower, sizeof(double) * B->nfTotal, kind, st));
    }
    B->set = false;
    buildCoeffs(B);
    B->set = true;
    API_END
}

int vol_batch_solve(vol_batch B, const double* b, double* x, const vol_solver_controls* controls,
                    vol_solver_performance* perf, int location) {
    API_BEGIN
    if (!B || !B->set || !b || !x) throw VolError(VOL_ERR_INVALID_ARGUMENT, "batch not set or null vector");
    vol_solver_controls ct;
    if (controls)
        ct = *controls;
    else
        vol_controls_default(&ct);
    HIPCHECK(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    const size_t n = B->nTotal;
    if (location == VOL_DEVICE) {
        bsolve(B, b, x, ct, perf);
    } else {
        B->hb.alloc(n);
        B->hx.alloc(n);
        HIPCHECK(hipMemcpyAsync(B->hb.p, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(B->hx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
        bsolve(B, B->hb.p, B->hx.p, ct, perf);
        HIPCHECK(hipMemcpyAsync(x, B->hx.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    API_END
}

}  // extern "C"

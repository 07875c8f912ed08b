// Host-side building blocks shared by the library's translation units (api.cpp,
// batch.cpp): errors, device buffers, the SELL/CSR row layouts, the GAMG hierarchy
// setup and the context handle. Internal; not part of the C ABI.
#pragma once

#include "../../include/voluma_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "vol_internal.hpp"

namespace vol::host {

struct VolError : std::runtime_error {
    int code;
    VolError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw VolError(VOL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) hipCheck((x), #x)

// Device bytes held by DevBufs, net of frees, counted per host thread (the solve
// server's workers read it around their work to account what each cached entry holds).
inline thread_local int64_t tlDeviceBytes = 0;

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            tlDeviceBytes -= int64_t(sizeof(T) * std::max<size_t>(n, 1));
        }
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        HIPCHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
        n = count;
        tlDeviceBytes += int64_t(sizeof(T) * std::max<size_t>(count, 1));
    }
    void upload(const std::vector<T>& v, hipStream_t st) {
        alloc(v.size());
        if (!v.empty()) HIPCHECK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    }
};

// One level order of the triangular sweeps in host and device form (LevelSched view):
// offsets and permutation, and the launch groups — runs of consecutive levels of at most
// SWEEP_RUN_MAX cells (one workgroup) and single larger levels (a grid each).
struct LevelOrderDev {
    std::vector<int> off, grp;
    std::vector<char> run;
    DevBuf<int> perm, offD;
    void build(const std::vector<int>& offH, const std::vector<int>& permH, hipStream_t st) {
        off = offH;
        const int nl = int(off.size()) - 1;
        grp.clear();
        run.clear();
        for (int L = 0; L < nl;) {
            grp.push_back(L);
            if (off[L + 1] - off[L] > SWEEP_RUN_MAX) {
                run.push_back(0);
                ++L;
            } else {
                run.push_back(1);
                while (L < nl && off[L + 1] - off[L] <= SWEEP_RUN_MAX) ++L;
            }
        }
        grp.push_back(std::max(nl, 0));
        perm.upload(permH, st);
        offD.upload(off, st);
    }
    LevelSched view() const {
        LevelSched v;
        v.perm = perm.p;
        v.offD = offD.p;
        v.offH = off.data();
        v.nLev = int(off.size()) - 1;
        v.grpH = grp.data();
        v.grpRunH = run.data();
        v.nGrp = int(grp.size()) - 1;
        return v;
    }
};

// Level-ordered ELL rows of a triangular sweep (LevelEll view). Rows are given as a
// CSR by cell (rowOff, rowCol, rowSrc: neighbour and coefficient source index of each
// entry, in the sweep's order); src keeps the source index of every entry (−1 padding)
// for the per-matrix coefficient gather.
struct LevelEllDev {
    std::vector<int> off, width, base, grp, runOff, runCnt;
    std::vector<char> kind;
    DevBuf<int> cell, slot, runCells, col, src, offD, widthD, baseD;
    DevBuf<int2> waves;  // sync-free sweeps: (level, first row) per wavefront
    DevBuf<int> runBaseD;  // per level: first runCells entry of its run (−1: not in a run)
    int nWaves = 0;
    size_t nEntries = 0;
    void build(const std::vector<int>& levOff, const std::vector<int>& perm, const std::vector<int>& rowOff,
               const std::vector<int>& rowCol, const std::vector<int>& rowSrc, hipStream_t st) {
        off = levOff;
        const int nl = int(off.size()) - 1;
        width.assign(std::max(nl, 0), 0);
        base.assign(std::max(nl, 0), 0);
        size_t total = 0;
        for (int L = 0; L < nl; ++L) {
            if (off[L + 1] <= off[L]) throw VolError(VOL_ERR_INVALID_ARGUMENT, "empty sweep level");
            int w = 0;
            for (int i = off[L]; i < off[L + 1]; ++i) w = std::max(w, rowOff[perm[i] + 1] - rowOff[perm[i]]);
            width[L] = w;
            base[L] = int(total);
            total += size_t(w) * size_t(off[L + 1] - off[L]);
        }
        if (total >= (size_t(1) << 31)) throw VolError(VOL_ERR_INVALID_ARGUMENT, "sweep rows exceed 2^31 entries");
        nEntries = total;
        {  // wavefronts of one level each, in level order (the sync-free sweeps' work order)
            std::vector<int2> wv;
            for (int L = 0; L < nl; ++L)
                for (int t0 = 0; t0 < off[L + 1] - off[L]; t0 += WAVE) wv.push_back(make_int2(L, t0));
            nWaves = int(wv.size());
            waves.upload(wv, st);
        }
        // at least one entry, so a sweep's stand-in reads of entry 0 stay in bounds
        std::vector<int> ec(std::max<size_t>(total, 1), -1), es(total, -1);
        for (int L = 0; L < nl; ++L) {
            const int size = off[L + 1] - off[L];
            for (int t = 0; t < size; ++t) {
                const int c = perm[off[L] + t];
                for (int k = 0; k < rowOff[c + 1] - rowOff[c]; ++k) {
                    const size_t e = size_t(base[L]) + size_t(k) * size + t;
                    ec[e] = rowCol[rowOff[c] + k];
                    es[e] = rowSrc[rowOff[c] + k];
                }
            }
        }
        grp.clear();
        kind.clear();
        auto kindOf = [&](int L) {
            const int size = off[L + 1] - off[L];
            return size <= 256 ? ELL_RUN256 : size <= 512 ? ELL_RUN512 : size <= 1024 ? ELL_RUN1024 : ELL_GRID;
        };
        for (int L = 0; L < nl;) {
            if (kindOf(L) == ELL_GRID) {
                grp.push_back(L);
                kind.push_back(char(ELL_GRID));
                ++L;
                continue;
            }
            // a stretch of run levels [L, R): segments of one class, a segment of fewer
            // than 16 levels merged into a neighbour (the merged run takes the larger
            // class), each segment cut into runs of at most ELL_RUN_LEVELS levels
            int R = L;
            while (R < nl && kindOf(R) != ELL_GRID) ++R;
            std::vector<int> segS, segK;
            for (int i = L; i < R; ++i)
                if (segK.empty() || segK.back() != kindOf(i)) {
                    segS.push_back(i);
                    segK.push_back(kindOf(i));
                }
            segS.push_back(R);
            for (bool merged = true; merged && segK.size() > 1;) {
                merged = false;
                for (size_t g = 0; g < segK.size(); ++g) {
                    if (segS[g + 1] - segS[g] >= 16) continue;
                    const size_t a = g > 0 ? g - 1 : 0;  // merge segments a and a + 1
                    segK[a] = std::max(segK[a], segK[a + 1]);
                    segK.erase(segK.begin() + long(a) + 1);
                    segS.erase(segS.begin() + long(a) + 1);
                    merged = true;
                    break;
                }
            }
            for (size_t g = 0; g < segK.size(); ++g)
                for (int s0 = segS[g]; s0 < segS[g + 1]; s0 += ELL_RUN_LEVELS) {
                    grp.push_back(s0);
                    kind.push_back(char(segK[g]));
                }
            L = R;
        }
        grp.push_back(std::max(nl, 0));
        // LDS-resident runs: a run takes consecutive levels while the cells its rows
        // write or read fit RUN_LDS_CELLS slots (a level that alone does not fit is swept
        // as a grid level); slots in first-appearance order (a row's cell, then its
        // neighbours), run rows' entries and own cells re-coded as slots
        std::vector<int> slotv(std::max<size_t>(perm.size(), 1), 0), cells;
        std::vector<int> runBase(std::max(nl, 1), -1);  // entries of run levels hold slots
        {
            std::vector<int> g2, ro, rn;
            std::vector<char> k2;
            std::unordered_map<int, int> map, seen;
            std::vector<int> fresh;
            auto close = [&](int L, char k, int start, int cnt) {
                g2.push_back(L);
                k2.push_back(k);
                ro.push_back(start);
                rn.push_back(cnt);
            };
            for (size_t g = 0; g + 1 < grp.size(); ++g) {
                const int A = grp[g], B = grp[g + 1];
                if (kind[g] == char(ELL_GRID)) {
                    close(A, kind[g], int(cells.size()), 0);
                    continue;
                }
                for (int L = A; L < B;) {
                    map.clear();
                    const int start = int(cells.size());
                    int R = L;
                    for (; R < B; ++R) {
                        fresh.clear();
                        seen.clear();
                        for (int i = off[R]; i < off[R + 1]; ++i) {
                            const int c = perm[i];
                            auto take = [&](int q) {
                                if (!map.count(q) && seen.emplace(q, 0).second) fresh.push_back(q);
                            };
                            take(c);
                            for (int k = rowOff[c]; k < rowOff[c + 1]; ++k) take(rowCol[k]);
                        }
                        if (map.size() + fresh.size() > size_t(RUN_LDS_CELLS)) break;
                        for (int q : fresh) {
                            map.emplace(q, int(cells.size()) - start);
                            cells.push_back(q);
                        }
                    }
                    if (R == L) {  // one level alone exceeds the slots
                        close(L, char(ELL_GRID), start, 0);
                        ++L;
                        continue;
                    }
                    close(L, kind[g], start, int(cells.size()) - start);
                    for (int Lq = L; Lq < R; ++Lq) {
                        const int size = off[Lq + 1] - off[Lq];
                        runBase[Lq] = start;
                        for (int t = 0; t < size; ++t) {
                            const int c = perm[off[Lq] + t];
                            slotv[off[Lq] + t] = map.at(c);
                            for (int k = 0; k < rowOff[c + 1] - rowOff[c]; ++k) {
                                const size_t e = size_t(base[Lq]) + size_t(k) * size + t;
                                ec[e] = map.at(ec[e]);
                            }
                        }
                    }
                    L = R;
                }
            }
            g2.push_back(std::max(nl, 0));
            grp = std::move(g2);
            kind = std::move(k2);
            runOff = std::move(ro);
            runCnt = std::move(rn);
        }
        cell.upload(perm, st);
        slot.upload(slotv, st);
        if (cells.empty()) cells.push_back(0);
        runCells.upload(cells, st);
        runBaseD.upload(runBase, st);
        col.upload(ec, st);
        src.upload(es, st);
        offD.upload(off, st);
        widthD.upload(width, st);
        baseD.upload(base, st);
        HIPCHECK(hipStreamSynchronize(st));
    }
    LevelEll view() const {
        LevelEll v;
        v.cell = cell.p;
        v.slot = slot.p;
        v.runCells = runCells.p;
        v.runOffH = runOff.data();
        v.runCntH = runCnt.data();
        v.col = col.p;
        v.offD = offD.p;
        v.widthD = widthD.p;
        v.baseD = baseD.p;
        v.offH = off.data();
        v.widthH = width.data();
        v.baseH = base.data();
        v.nLev = int(off.size()) - 1;
        v.grpH = grp.data();
        v.grpKindH = kind.data();
        v.nGrp = int(grp.size()) - 1;
        v.waves = waves.p;
        v.nWaves = nWaves;
        v.runBaseD = runBaseD.p;
        return v;
    }
};

struct SellDev {
    int n = 0, nslices = 0, nslots = 0, maxWidth = 0, csrG = 0, csrIt = 0, ellWidth = 0;
    DevBuf<int> sliceOff, col, src;
    DevBuf<int16_t> col16;
    DevBuf<int> colP;  // packed rows (SellView::colP)
    int colPW = 0, nslotsP = 0;
    DevBuf<int> perm, iperm;  // sorted rows (SellView::perm)
    DevBuf<int2> colS;        // packed SELL rows (SellView::colS)
    DevBuf<int4> sliceP;
    SellView view() const {
        SellView v;
        v.n = n;
        v.nslices = nslices;
        v.nslots = nslots;
        v.maxWidth = maxWidth;
        v.csrG = csrG;
        v.csrIt = csrIt;
        v.sliceOff = sliceOff.p;
        v.col = col.p;
        v.src = src.p;
        v.col16 = col16.p;
        v.ellWidth = ellWidth;
        v.colP = colP.p;
        v.colPW = colPW;
        v.nslotsP = nslotsP;
        v.perm = perm.p;
        v.iperm = iperm.p;
        v.colS = colS.p;
        v.sliceP = sliceP.p;
        return v;
    }
};

// Row lists (row → (col, src) in the given order) → SELL-64 on the device. With try16,
// int16 slot-major column offsets under VOLUMA_COL16=1 when every neighbour lies within
// ±32767 rows; with packed, the packed rows (SellView::colP) when the form allows them
// (VOLUMA_PACKED=0: never); the owner of the coefficients then allocates nslots + nslotsP.
// sortWindow > 0 (levels of more than sortWindow rows): rows sorted by length, longest
// first and ties in row order, within each window of sortWindow consecutive rows
// (SellView::perm); the int16 and packed forms are then not built. pairs: the packed
// SELL rows as well (SellView::colS) when every column offset fits int16; the owner of
// the coefficients then allocates nslots + nslotsP.
inline void buildSell(int n, const std::vector<int>& off, const std::vector<int>& colL, const std::vector<int>& srcL,
               SellDev& S, hipStream_t st, bool try16 = false, bool packed = false, int sortWindow = 0,
               bool pairs = false) {
    S.n = n;
    S.nslices = (n + WAVE - 1) / WAVE;
    std::vector<int> posRow(n);  // position → row
    for (int r = 0; r < n; ++r) posRow[r] = r;
    const bool sorted = sortWindow > 0 && n > sortWindow;
    if (sorted) {
        for (int w0 = 0; w0 < n; w0 += sortWindow)
            std::stable_sort(posRow.begin() + w0, posRow.begin() + std::min(n, w0 + sortWindow),
                             [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
        try16 = packed = pairs = false;
    }
    std::vector<int> sliceOff(S.nslices + 1, 0);
    for (int s = 0; s < S.nslices; ++s) {
        int w = 0;
        for (int p = s * WAVE; p < std::min(n, (s + 1) * WAVE); ++p)
            w = std::max(w, off[posRow[p] + 1] - off[posRow[p]]);
        sliceOff[s + 1] = sliceOff[s] + w * WAVE;
        S.maxWidth = std::max(S.maxWidth, w);
    }
    // ELL (every slice at the longest row's width) when it costs at most 10 % more slots
    if (size_t(S.nslices) * WAVE * S.maxWidth <= size_t(sliceOff[S.nslices]) * 11 / 10) {
        S.ellWidth = S.maxWidth;
        for (int s = 0; s <= S.nslices; ++s) sliceOff[s] = s * WAVE * S.maxWidth;
    }
    S.nslots = sliceOff[S.nslices];
    std::vector<int> col(S.nslots, -1), src(S.nslots, -1);
    for (int p = 0; p < n; ++p) {
        const int s = p / WAVE, lane = p % WAVE, r = posRow[p];
        for (int k = off[r]; k < off[r + 1]; ++k) {
            const int idx = sliceOff[s] + (k - off[r]) * WAVE + lane;
            col[idx] = colL[k];
            src[idx] = srcL[k];
        }
    }
    S.sliceOff.upload(sliceOff, st);
    S.col.upload(col, st);
    S.src.upload(src, st);
    if (sorted) {
        std::vector<int> rowPos(n);
        for (int p = 0; p < n; ++p) rowPos[posRow[p]] = p;
        S.perm.upload(posRow, st);
        S.iperm.upload(rowPos, st);
    }
    if (pairs) {
        bool fits = true;
        for (int r = 0; r < n && fits; ++r)
            for (int k = off[r]; k < off[r + 1] && fits; ++k) {
                const int d = colL[k] - r;
                fits = d > COL16_PAD && d <= 32767;
            }
        if (fits) {
            const int padWord = int((uint32_t(uint16_t(COL16_PAD)) << 16) | uint16_t(COL16_PAD));
            std::vector<int4> sp(S.nslices);
            std::vector<int2> words;
            int ob = 0, pb = 0;
            for (int s = 0; s < S.nslices; ++s) {
                const int w = (sliceOff[s + 1] - sliceOff[s]) / WAVE, Q = (w + 3) / 4, P = (w + 1) / 2;
                sp[s] = make_int4(ob, pb, w, 0);
                words.resize(size_t(ob) + size_t(WAVE) * Q, make_int2(padWord, padWord));
                for (int lane = 0; lane < WAVE; ++lane) {
                    const int r = s * WAVE + lane;
                    if (r >= n) break;
                    for (int k = off[r]; k < off[r + 1]; ++k) {
                        const int q = (k - off[r]) / 4, e = (k - off[r]) % 4;
                        int& word = e < 2 ? words[size_t(ob) + size_t(WAVE) * q + lane].x
                                          : words[size_t(ob) + size_t(WAVE) * q + lane].y;
                        const uint32_t h = uint16_t(int16_t(colL[k] - r));
                        word = (e & 1) ? int((uint32_t(word) & 0xffffu) | (h << 16))
                                       : int((uint32_t(word) & 0xffff0000u) | h);
                    }
                }
                ob += WAVE * Q;
                pb += WAVE * P;
            }
            S.colS.upload(words, st);
            S.sliceP.upload(sp, st);
            S.nslotsP = 2 * pb;
        }
    }
    if (try16 && std::getenv("VOLUMA_COL16") && std::string(std::getenv("VOLUMA_COL16")) == "1") {
        std::vector<int16_t> c16(S.nslots, int16_t(COL16_PAD));
        bool fits = true;
        for (int r = 0; r < n && fits; ++r) {
            const int s = r / WAVE, lane = r % WAVE;
            for (int k = off[r]; k < off[r + 1]; ++k) {
                const int d = colL[k] - r;
                if (d <= COL16_PAD || d > 32767) {
                    fits = false;
                    break;
                }
                c16[sliceOff[s] + (k - off[r]) * WAVE + lane] = int16_t(d);
            }
        }
        if (fits) S.col16.upload(c16, st);
    }
    const char* packEnv = std::getenv("VOLUMA_PACKED");
    if (packed && S.ellWidth > 0 && S.ellWidth <= 8 && !(packEnv && std::string(packEnv) == "0")) {
        const int W = S.ellWidth, PW = W <= 4 ? 4 : 8;  // offsets per row, padded
        std::vector<int16_t> p16(size_t(S.nslices) * WAVE * PW, int16_t(COL16_PAD));
        bool fits = true;
        for (int r = 0; r < n && fits; ++r)
            for (int k = off[r]; k < off[r + 1]; ++k) {
                const int d = colL[k] - r;
                if (d <= COL16_PAD || d > 32767) {
                    fits = false;
                    break;
                }
                p16[size_t(r) * PW + (k - off[r])] = int16_t(d);
            }
        if (fits) {
            std::vector<int> words(p16.size() / 2);
            for (size_t q = 0; q < words.size(); ++q)
                words[q] = int((uint32_t(uint16_t(p16[2 * q + 1])) << 16) | uint16_t(p16[2 * q]));
            S.colP.upload(words, st);
            S.colPW = PW / 2;
            S.nslotsP = S.nslices * WAVE * 2 * ((W + 1) / 2);
        }
    }
}

// Row lists → CSR with G lanes per row (coarse levels): G·IT ≥ longest row, IT ≤ 8.
// Returns false when rows are too long for it (then SELL is used).
// Entries of every row of S in slot order (the order of its gathers): neighbour and slot
// index, as a CSR by row (read back from the device).
inline void sellRows(const SellDev& S, std::vector<int>& off, std::vector<int>& col, std::vector<int>& slot) {
    const int n = S.n;
    const int nOff = S.csrG > 0 ? n + 1 : S.nslices + 1;
    std::vector<int> so(nOff), sc(S.nslots), ip;
    HIPCHECK(hipMemcpy(so.data(), S.sliceOff.p, sizeof(int) * so.size(), hipMemcpyDeviceToHost));
    if (S.nslots) HIPCHECK(hipMemcpy(sc.data(), S.col.p, sizeof(int) * sc.size(), hipMemcpyDeviceToHost));
    if (S.iperm.p) {
        ip.resize(n);
        HIPCHECK(hipMemcpy(ip.data(), S.iperm.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    off.assign(n + 1, 0);
    col.clear();
    slot.clear();
    for (int r = 0; r < n; ++r) {
        int s0, width, stride;
        const int q = ip.empty() ? r : ip[r];  // the row's position in the slices
        if (S.csrG > 0) {
            s0 = so[r];
            width = so[r + 1] - s0;
            stride = 1;
        } else if (S.ellWidth > 0) {
            width = S.ellWidth;
            s0 = (q / WAVE) * WAVE * width + q % WAVE;
            stride = WAVE;
        } else {
            width = (so[q / WAVE + 1] - so[q / WAVE]) / WAVE;
            s0 = so[q / WAVE] + q % WAVE;
            stride = WAVE;
        }
        for (int k = 0; k < width; ++k) {
            const int sl = s0 + k * stride;
            if (sc[sl] >= 0) {
                col.push_back(sc[sl]);
                slot.push_back(sl);
            }
        }
        off[r + 1] = int(col.size());
    }
}

inline bool buildCsr(int n, const std::vector<int>& off, const std::vector<int>& colL, const std::vector<int>& srcL,
              SellDev& S, hipStream_t st) {
    int maxW = 0;
    for (int r = 0; r < n; ++r) maxW = std::max(maxW, off[r + 1] - off[r]);
    int G = 4;
    while (G < 16 && G * 4 < maxW) G *= 2;
    const int IT = (maxW + G - 1) / G;
    if (IT > 8) return false;
    S.n = n;
    S.nslices = n;
    S.nslots = off[n];
    S.maxWidth = maxW;
    S.csrG = G;
    S.csrIt = std::max(IT, 1);
    S.sliceOff.upload(off, st);
    S.col.upload(colL, st);
    S.src.upload(srcL, st);
    return true;
}

// Time the two sweeps a 1/1 V-cycle runs on a level in the forms the solve uses (the
// residual gathering z; the post-sweep gathering z + xc[agg]) over layout S on stream st,
// with the caches flushed before each launch (in a V-cycle the finer levels have evicted
// most of the level's data): median of five, seconds. Values are zero; aggDev is the
// level's map to the next level (nullable: four consecutive rows per aggregate).
inline double timeLayout(const SellDev& S, hipStream_t st, void* flush, size_t flushBytes, const int* aggDev) {
    DevBuf<double> coef, v;
    DevBuf<int> agg;
    const size_t n = S.n, nc = n;
    coef.alloc(S.nslots);
    v.alloc(5 * n + nc);
    HIPCHECK(hipMemsetAsync(coef.p, 0, sizeof(double) * std::max<size_t>(S.nslots, 1), st));
    HIPCHECK(hipMemsetAsync(v.p, 0, sizeof(double) * (5 * n + nc), st));
    if (!aggDev) {
        std::vector<int> ag(n);
        for (size_t i = 0; i < n; ++i) ag[i] = int(i / 4);
        agg.upload(ag, st);
        aggDev = agg.p;
    }
    SellArgs a;
    a.S = S.view();
    a.coef = coef.p;
    a.diag = v.p;
    a.rD = v.p + n;
    a.r = v.p + 2 * n;
    a.x = v.p + 3 * n;  // z
    a.out = v.p + 4 * n;
    a.xc = v.p + 5 * n;
    a.agg = aggDev;
    a.omega = 0.9;
    hipEvent_t e[4];
    for (auto& x : e) HIPCHECK(hipEventCreate(&x));
    constexpr int REPS = 5;
    float t[REPS];
    for (int k = 0; k < REPS; ++k) {
        float t1 = 0, t2 = 0;
        HIPCHECK(hipMemsetAsync(flush, k + 1, flushBytes, st));
        a.evStart = e[0];
        a.evStop = e[1];
        launch_sell(st, OP_RESID, XL_PLAIN, a);
        HIPCHECK(hipMemsetAsync(flush, k + 2, flushBytes, st));
        a.evStart = e[2];
        a.evStop = e[3];
        launch_sell(st, OP_JACOBI, XL_PROLONG, a);
        HIPCHECK(hipEventSynchronize(e[3]));
        HIPCHECK(hipEventElapsedTime(&t1, e[0], e[1]));
        HIPCHECK(hipEventElapsedTime(&t2, e[2], e[3]));
        t[k] = t1 + t2;
    }
    for (auto& x : e) (void)hipEventDestroy(x);
    std::sort(t, t + REPS);
    return t[REPS / 2] / 1e3;
}

// Layout of a coarse GAMG level: the CSR-vector form (G lanes per row) or SELL-64 (both
// give the same results, row by row in slot order). The two sweeps of a 1/1 V-cycle are
// timed in each from cold caches on this GPU, and CSR is taken when it is at least 5 %
// faster: SELL reads one wavefront per 64 rows and measured faster inside the solve on
// the big coarse levels when the cold timings were close. Levels above 4M rows use SELL.
// `choice` (0 CSR, 1 SELL) skips the measurement: copies of a mesh take the choice made
// for the original. VOLUMA_COARSE_LAYOUT=csr / sell fixes the choice. agg (nullable):
// the level's map to the next level, for the timing. Returns the layout built (0 CSR,
// 1 SELL).
inline int buildCoarseLayout(int n, const std::vector<int>& off, const std::vector<int>& col,
                             const std::vector<int>& src, SellDev& S, hipStream_t st, int choice = -1,
                             const std::vector<int>* agg = nullptr, int sortWindow = 0, bool pairs = false) {
    constexpr int TUNE_MAX_ROWS = 1 << 22;
    const char* force = std::getenv("VOLUMA_COARSE_LAYOUT");
    const std::string f = force ? force : "";
    if (f == "csr") choice = 0;
    if (f == "sell" || n > TUNE_MAX_ROWS) choice = 1;
    if (choice < 0) {
        choice = 1;
        SellDev a, b;
        if (buildCsr(n, off, col, src, a, st)) {
            buildSell(n, off, col, src, b, st, false, false, sortWindow);
            const size_t fb = size_t(512) << 20;  // larger than L2 + the 256 MB Infinity Cache
            DevBuf<char> flush;
            flush.alloc(fb);
            DevBuf<int> aggD;
            if (agg && int(agg->size()) == n) aggD.upload(*agg, st);
            const double tc = timeLayout(a, st, flush.p, fb, aggD.p);
            const double ts = timeLayout(b, st, flush.p, fb, aggD.p);
            if (tc < 0.95 * ts) choice = 0;
        }
    }
    if (choice == 0 && buildCsr(n, off, col, src, S, st)) return 0;
    buildSell(n, off, col, src, S, st, false, false, sortWindow, pairs);
    return 1;
}

// ------------------------------------------------------------------ GAMG hierarchy (host setup)
// Pairwise agglomeration on geometric weights, same decisions as the reference's
// agglomerate/buildHierarchy (src/fvm/linalg/GAMG.cpp:40-94, 206-252): cells visited in
// index order, pair with the unaggregated neighbour of strictly largest weight (first
// in adjacency order on ties), otherwise join the strongest neighbour's aggregate;
// coarse faces numbered by first appearance in fine-face order, weights summed in that
// order; mergeLevels pairings per level; stop at ≤ 64 cells or when no progress.
struct HLevel {
    int n = 0;
    std::vector<int> l, u;
    std::vector<double> weight;
    std::vector<int> off, adjCell, adjFace;
    std::vector<char> adjIsL;
    std::vector<int> agg, faceMap;
    std::vector<char> faceFlip;

    void adjacency() {
        const int nf = int(l.size());
        off.assign(n + 1, 0);
        for (int f = 0; f < nf; ++f) {
            ++off[l[f] + 1];
            ++off[u[f] + 1];
        }
        for (int i = 0; i < n; ++i) off[i + 1] += off[i];
        adjCell.resize(off[n]);
        adjFace.resize(off[n]);
        adjIsL.resize(off[n]);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int f = 0; f < nf; ++f) {
            int k = fill[l[f]]++;
            adjCell[k] = u[f];
            adjFace[k] = f;
            adjIsL[k] = 1;
            k = fill[u[f]]++;
            adjCell[k] = l[f];
            adjFace[k] = f;
            adjIsL[k] = 0;
        }
    }

    HLevel pair() {
        agg.assign(n, -1);
        int nc = 0;
        for (int i = 0; i < n; ++i) {
            if (agg[i] >= 0) continue;
            int best = -1;
            double bw = -1;
            for (int k = off[i]; k < off[i + 1]; ++k) {
                const int j = adjCell[k];
                const double wk = weight[adjFace[k]];
                if (agg[j] < 0 && j != i && wk > bw) {
                    bw = wk;
                    best = j;
                }
            }
            if (best >= 0) {
                agg[i] = agg[best] = nc++;
                continue;
            }
            for (int k = off[i]; k < off[i + 1]; ++k) {
                const double wk = weight[adjFace[k]];
                if (wk > bw) {
                    bw = wk;
                    best = adjCell[k];
                }
            }
            agg[i] = best >= 0 ? agg[best] : nc++;
        }
        HLevel C;
        C.n = nc;
        const int nf = int(l.size());
        faceMap.assign(nf, -1);
        faceFlip.assign(nf, 0);
        std::unordered_map<uint64_t, int> ids;
        ids.reserve(size_t(nf) / 2 + 16);
        for (int f = 0; f < nf; ++f) {
            const int a = agg[l[f]], b = agg[u[f]];
            if (a == b) continue;
            const int lo = std::min(a, b), hi = std::max(a, b);
            const uint64_t key = (uint64_t(uint32_t(lo)) << 32) | uint32_t(hi);
            auto [it, fresh] = ids.try_emplace(key, int(C.l.size()));
            if (fresh) {
                C.l.push_back(lo);
                C.u.push_back(hi);
                C.weight.push_back(0.0);
            }
            faceMap[f] = it->second;
            faceFlip[f] = char(a != lo);
            C.weight[it->second] += weight[f];
        }
        return C;
    }
};

inline std::vector<HLevel> buildHierarchy(int n, int nf, const int* own, const int* nei, const double* w, int mergeLevels) {
    std::vector<HLevel> H;
    HLevel L0;
    L0.n = n;
    L0.l.resize(nf);
    L0.u.resize(nf);
    L0.weight.assign(w, w + nf);
    for (int f = 0; f < nf; ++f) {
        L0.l[f] = std::min(own[f], nei[f]);
        L0.u[f] = std::max(own[f], nei[f]);
    }
    L0.adjacency();
    H.push_back(std::move(L0));
    const int nCoarsest = 64;
    while (H.size() < 40) {
        HLevel& F = H.back();
        const int nF = F.n;
        if (nF <= nCoarsest) break;
        HLevel C = F.pair();
        for (int k = 1; k < mergeLevels && C.n > nCoarsest; ++k) {
            C.adjacency();
            HLevel C2 = C.pair();
            for (auto& a : F.agg) a = C.agg[a];
            for (size_t f = 0; f < F.faceMap.size(); ++f) {
                const int cf = F.faceMap[f];
                if (cf < 0) continue;
                F.faceMap[f] = C.faceMap[cf];
                F.faceFlip[f] = char(F.faceFlip[f] != C.faceFlip[cf]);
            }
            C = std::move(C2);
        }
        if (C.n >= nF) {
            F.agg.clear();
            break;
        }
        C.adjacency();
        H.push_back(std::move(C));
    }
    return H;
}

}  // namespace vol::host

namespace vol {
struct AgglomerationStats {
    int passes = 0;  // pairing passes
    int rounds = 0;  // decision rounds over all passes
};
// buildHierarchy computed on the GPU of stream st (agglomeration.hip): the same levels,
// maps, coarse faces, weights and adjacency.
std::vector<host::HLevel> buildHierarchyDevice(hipStream_t st, int n, int nf, const int* own, const int* nei,
                                               const double* w, int mergeLevels, AgglomerationStats* stats);
}  // namespace vol

namespace vol::host {

struct LevelDev {
    int n = 0, nf = 0;
    SellDev sell;                       // coarse levels (level 0 uses the mesh SELL)
    int layout = -1;                    // of sell: 0 CSR, 1 SELL (-1 at level 0)
    DevBuf<int> agg;                    // → next level
    DevBuf<int> mOff, mList;            // next-level cell → its fine cells (ascending)
    DevBuf<int> iOff, iList;            // next-level cell → intra-aggregate fine faces (ascending)
    DevBuf<int> cfOff, cfList;          // next-level face → (fine face << 1 | flip), ascending
    DevBuf<int> l, u;                   // faces (coarsest: dense assembly)
    // fused residual + restriction: per block (first aggregate, first member position) of
    // its range; per member position (row, first slot, slot count) in the level's CSR
    // layout, or, for a SELL level, in rrCsr: the level's rows again, in the order of the
    // restriction lists, as CSR (its coefficients are filled beside the SELL ones)
    DevBuf<int2> rrBlk;
    DevBuf<int4> rrRows;
    int rrBlocks = 0;
    SellDev rrCsr;
};

// Block ranges of the fused residual + restriction over the aggregates of a level with G
// lanes per row: consecutive aggregates while their member rows number at most one per
// lane group (256/G), and at most RR_POS aggregates; an aggregate alone takes a block
// whatever its size. Empty when an aggregate has more than RR_POS member rows.
inline std::vector<int> restrictBlocks(const std::vector<int>& mOff, int G) {
    const int nc = int(mOff.size()) - 1, cap = 256 / G;
    std::vector<int> blk{0};
    int I = 0;
    while (I < nc) {
        if (mOff[I + 1] - mOff[I] > RR_POS) return {};
        const int start = I++;
        while (I < nc && I - start < RR_POS && mOff[I + 1] - mOff[start] <= cap) ++I;
        blk.push_back(I);
    }
    return blk;
}


// Per-thread message of the last failed C ABI call (vol_last_error).
void setLastError(const std::string& msg);

// Restriction and Galerkin gather lists from fine level F to coarse level C:
//   mOff/mList:   coarse cell → its fine cells, ascending
//   iOff/iList:   coarse cell → fine faces inside the aggregate (faceMap < 0), ascending
//   cfOff/cfList: coarse face → (fine face << 1 | flip), ascending
// Fine cells with agg < 0 and faces of such cells are skipped (batches: systems whose
// coarsest level is F).
struct GatherLists {
    std::vector<int> mOff, mList, iOff, iList, cfOff, cfList;
    template <class LD> void upload(LD& D, hipStream_t st) const {
        D.mOff.upload(mOff, st);
        D.mList.upload(mList, st);
        D.iOff.upload(iOff, st);
        D.iList.upload(iList, st);
        D.cfOff.upload(cfOff, st);
        D.cfList.upload(cfList, st);
    }
};
inline GatherLists buildGatherLists(int nFine, int nCoarse, const std::vector<int>& agg, const std::vector<int>& fl,
                                    const std::vector<int>& faceMap, const std::vector<char>& flip, int nCoarseFaces) {
    GatherLists G;
    const int fnf = int(faceMap.size());
    G.mOff.assign(nCoarse + 1, 0);
    for (int i = 0; i < nFine; ++i)
        if (agg[i] >= 0) ++G.mOff[agg[i] + 1];
    for (int I = 0; I < nCoarse; ++I) G.mOff[I + 1] += G.mOff[I];
    G.mList.resize(G.mOff[nCoarse]);
    {
        std::vector<int> fill(G.mOff.begin(), G.mOff.end() - 1);
        for (int i = 0; i < nFine; ++i)
            if (agg[i] >= 0) G.mList[fill[agg[i]]++] = i;
    }
    G.iOff.assign(nCoarse + 1, 0);
    for (int f = 0; f < fnf; ++f)
        if (faceMap[f] < 0 && agg[fl[f]] >= 0) ++G.iOff[agg[fl[f]] + 1];
    for (int I = 0; I < nCoarse; ++I) G.iOff[I + 1] += G.iOff[I];
    G.iList.resize(G.iOff[nCoarse]);
    {
        std::vector<int> fill(G.iOff.begin(), G.iOff.end() - 1);
        for (int f = 0; f < fnf; ++f)
            if (faceMap[f] < 0 && agg[fl[f]] >= 0) G.iList[fill[agg[fl[f]]]++] = f;
    }
    G.cfOff.assign(nCoarseFaces + 1, 0);
    for (int f = 0; f < fnf; ++f)
        if (faceMap[f] >= 0) ++G.cfOff[faceMap[f] + 1];
    for (int c = 0; c < nCoarseFaces; ++c) G.cfOff[c + 1] += G.cfOff[c];
    G.cfList.resize(G.cfOff[nCoarseFaces]);
    {
        std::vector<int> fill(G.cfOff.begin(), G.cfOff.end() - 1);
        for (int f = 0; f < fnf; ++f)
            if (faceMap[f] >= 0) G.cfList[fill[faceMap[f]]++] = (f << 1) | (flip[f] ? 1 : 0);
    }
    return G;
}

// mesh accessors for translation units that do not see vol_mesh_s (api.cpp)
vol_context meshContext(vol_mesh m);
void meshSizes(vol_mesh m, int& n, int& nInternal, int& nBoundary);
bool meshHasGamg(vol_mesh m);
// the geometry vol_mesh_set_geometry stored (device pointers owned by the mesh)
void meshGeometry(vol_mesh m, const double*& V, const double*& Sf, const double*& magSf, const double*& w,
                  int& skewed, int& emptyDir);

inline int fail(const VolError& e) {
    setLastError(e.what());
    return e.code;
}
inline int failStd(const std::exception& e) {
    setLastError(e.what());
    return VOL_ERR_INVALID_ARGUMENT;
}

}  // namespace vol::host

// Body of every C ABI function: exceptions become status codes.
#define API_BEGIN try {
#define API_END                                           \
    }                                                     \
    catch (const vol::host::VolError& e) {                \
        return vol::host::fail(e);                        \
    }                                                     \
    catch (const std::bad_alloc&) {                       \
        vol::host::setLastError("host out of memory");    \
        return VOL_ERR_INVALID_ARGUMENT;                  \
    }                                                     \
    catch (const std::exception& e) {                     \
        return vol::host::failStd(e);                     \
    }                                                     \
    return VOL_OK;

using namespace vol;  // internal header: the library's own namespaces
using namespace vol::host;

struct vol_context_s {
    std::atomic<int> refs{1};  // caller handle + meshes + device allocations
    int device = 0;
    hipStream_t stream = nullptr;
    bool reproducible = false;
    // GAMG levels this small run in the one-workgroup tail kernel (VOLUMA_TAIL_MAX_CELLS;
    // 0 = off, the default: on MI355X separate launches for the small levels measured faster)
    int tailMax = 0;
    // Jacobi V-cycle forms that give the same bits (measurement switches, read from the
    // environment at context creation): coarse levels take z = ω·(r·rD) from the
    // restriction (VOLUMA_COARSE_Z=0: gather r and 1/diag instead); FP64 PCG keeps the
    // level-0 z with the residual (VOLUMA_LEVEL0_Z=0); VOLUMA_PROLONG_SPLIT=1: the
    // level-0 post-sweep reads z + xc[agg] from a streaming prolongation pass instead of
    // gathering it per neighbour (measured slower on the 2.06M-cell tets: 24.75 against
    // 24.60 ms per solve)
    bool coarseZ = true, level0Z = true, prolongSplit = false;
    // levels in the CSR layout compute their residual and its restriction in one launch
    // (VOLUMA_FUSED_RESTRICT=0: a residual gather, then the restriction); so can SELL
    // levels of at most fusedSellMax rows (VOLUMA_FUSED_SELL_MAX, default 0), from a copy
    // of their rows in restriction order: measured slower on the 2.06M-cell tets (level 1
    // 31.0 µs against 23.5 for the residual and restriction launches, and the following
    // post-sweeps slower as the SELL rows are no longer in cache; 22.8 against 21.4 ms)
    bool fusedRestrict = true;
    // the last post-sweep of a coarse level adds its result to the finer level's vector
    // over the aggregates, so that the finer post-sweep gathers one vector instead of
    // x + xc[agg] per neighbour (VOLUMA_PUSH_PROLONG=0: gather)
    bool pushProlong = true;
    // level 0 on packed rows computes its residual and restriction in one launch too,
    // reading the rows in restriction order (VOLUMA_FUSED_LEVEL0=0: separate launches;
    // read when a mesh is set up)
    bool fusedLevel0 = true;
    // the SELL coarse levels likewise, reading their own SELL rows in restriction order
    // (VOLUMA_FUSED_SELL_DIRECT=1; default off: measured slower on the 2.06M-cell tets,
    // level 1 35.8 µs against 21 µs for the two launches, as a wave's rows then share few
    // slot cache lines; 20.9 against 19.7 ms per solve)
    bool fusedSellDirect = false;
    // the last two levels above the coarsest and the coarsest in one workgroup when they
    // are small CSR levels of a 1/1 V-cycle with z kept (VOLUMA_LAST2=1; default off:
    // measured slower on the 2.06M-cell tets, 41.1 µs against 29.9 for the three launches
    // it replaces, 19.7 against 19.1 ms per solve: one workgroup pays every phase's
    // global-load round trips in turn)
    bool last2 = false;
    int fusedSellMax = 0;
    // SELL coarse levels: rows sorted by length within windows of this many rows
    // (SellView::perm; VOLUMA_SELL_SORT=<rows>, default 0: natural row order). Measured
    // slower on the 2.06M-cell tets (19.8 against 19.0 ms per solve): the sort cuts the
    // padded slots of levels 1–2 from 1.31–1.33× the entries to 1.03×, but each row's own
    // vectors are then read and written out of order, and the timed layout choice takes
    // CSR for level 1
    int sellSort = 0;
    // SELL levels also in the packed SELL form (SellView::colS): the coarse levels of a
    // mesh, and every level of a batch (VOLUMA_SELL_PAIRS=0: slot-major only)
    bool sellPairs = true;
    // triangular sweeps (Gauss–Seidel, DIC/DILU substitutions) in their sync-free form,
    // one launch per sweep (VOLUMA_SWEEP=free; the same bits; measured slower than the
    // default level-scheduled launches); sweepFault: [0] set by a sync-free wait that ran
    // past its limit, [1] the sync-free sweeps' work counter
    bool freeSweeps = false;
    DevBuf<int> sweepFault;
    RedScratch red;
    DevBuf<double> part, result;
    DevBuf<long long> xpart;
    DevBuf<int> xnf, arrive;
    // the fast-mode A·p finishes its own reduction in its last block (SellArgsT::arrive;
    // VOLUMA_LAST_BLOCK_FIN=1; same bits as the finalize launch)
    bool lastBlockFin = false;
    DevBuf<DevScalars> scal;
    DevScalars* hostScal = nullptr;  // pinned, 2 slots
    hipEvent_t loopEv[2] = {nullptr, nullptr};  // the solve loop's look-ahead events
    // timing
    bool timing = false;
    std::vector<hipEvent_t> evPool;
    size_t evUsed = 0;
    double amulMs = 0;
    int64_t amulCount = 0;

    void ensureScratch(int blocks) {
        const int need = std::max(blocks, RED_BLOCKS);
        if (red.maxBlk < need) {
            part.alloc(size_t(MAX_RED) * need);
            red.part = part.p;
            red.maxBlk = need;
        }
        if (!red.xpart) {
            xpart.alloc(size_t(MAX_RED) * RED_BLOCKS_EXACT * NBINS);
            xnf.alloc(size_t(MAX_RED) * RED_BLOCKS_EXACT);
            result.alloc(MAX_RED);
            red.xpart = xpart.p;
            red.xnf = xnf.p;
            red.result = result.p;
            red.maxBlkExact = RED_BLOCKS_EXACT;
        }
        if (!red.arrive) {
            arrive.alloc(1);
            HIPCHECK(hipMemset(arrive.p, 0, sizeof(int)));
            red.arrive = arrive.p;
        }
    }
};

// Drops one reference of a context and frees it with the last one (api.cpp).
void ctxRelease(vol_context c);
// Addressing and GAMG hierarchy of a mesh handle (api.cpp).
struct MeshView {
    int n = 0, nf = 0;
    const int* own = nullptr;
    const int* nei = nullptr;
    const std::vector<HLevel>* H = nullptr;  // null without agglomeration weights
    std::vector<int> layout;                 // per level: 0 CSR, 1 SELL (-1 at level 0)
};
MeshView meshView(vol_mesh m);
// A batch of K copies of mesh m's system on context ctx, its hierarchy reused (batch.cpp).
vol_batch batchOfCopies(vol_context ctx, vol_mesh m, int K);
// Symmetric matrices of the batch's systems from host arrays (diag folded, upper), then
// the GAMG coefficients of every system (batch.cpp).
void batchSetSystems(vol_batch B, const std::vector<const double*>& diag, const std::vector<const double*>& upper);
// Copy of a mesh (addressing, GAMG hierarchy, boundary lists) for another context (api.cpp).
vol_mesh meshCloneTo(vol_mesh src, vol_context dst);

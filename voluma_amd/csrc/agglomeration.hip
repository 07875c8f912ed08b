// GAMG pairwise agglomeration on the GPU, with the maps of the sequential pairing
// (src/fvm/linalg/GAMG.cpp:40-94: cells in index order, each unaggregated cell pairs
// with its unaggregated neighbour of strictly largest weight, first in adjacency order
// on ties, or else joins its strongest neighbour's aggregate; coarse faces numbered by
// first appearance in fine-face order, weights summed in that order; GAMG.cpp:206-252:
// mergeLevels pairings per level, stop at ≤ 64 cells or without progress).
//
// The sequential loop is a greedy matching in index order. Its decisions are computed
// here in rounds that are exact, not speculative: a cell's state goes once from
// undecided to final, and a cell decides only from final states:
//   * it is taken when a lower-numbered neighbour has already picked it;
//   * it leads when every lower-numbered neighbour is final and none picked it; it
//     then picks the first-maximum-weight neighbour j > i among those not known to be
//     taken before i (taken: a neighbour m < i of j picked j) as soon as that
//     neighbour's own status at time i is known (all its neighbours m < i final), or
//     finds no candidate and joins/stays single.
// Each round is a launch over all cells; a round only adds decisions, and reading a
// state decided in the same round is as valid as reading it a round later, so the
// launches need no ordering beyond the stream's. On tet70 the first pass finishes in
// ~480 rounds (the wavefront of decisions sweeps the index order), box64 in ~320.
// Aggregate numbers (creation order = leader index order) come from a prefix sum,
// joining cells follow their strongest neighbour's aggregate by pointer chasing, coarse
// faces from a stable radix sort of (coarse pair, fine face) keys, the next level's
// adjacency from a stable sort of its face sides by row.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "host_common.hpp"

namespace vol {

namespace {

constexpr int ST_UNDEC = -3;   // not decided yet
constexpr int ST_TAKEN = -2;   // picked by a lower-numbered neighbour
constexpr int ST_NOPICK = -1;  // leads, no untaken neighbour: joins or stays single
                               // ≥ 0: leads and picked that neighbour
constexpr int AGG_THREADS = 256;

inline int blocksFor(long n) { return int(std::max<long>(1, (n + AGG_THREADS - 1) / AGG_THREADS)); }

__global__ void k_fill_i32(int n, int* a, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// One round of decisions. done: set by the host's check when every cell is final.
__global__ __launch_bounds__(AGG_THREADS) void k_pair_round(int n, const int* __restrict__ off,
                                                           const int* __restrict__ adj,
                                                           const int* __restrict__ adjFace,
                                                           const double* __restrict__ w, int* st, int* decided) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int dec = 0;
    if (i < n && st[i] == ST_UNDEC) {
        const int k0 = off[i], k1 = off[i + 1];
        bool unknownLow = false, taken = false;
        for (int k = k0; k < k1; ++k) {
            const int j = adj[k];
            if (j < i) {
                const int s = st[j];
                if (s == ST_UNDEC) unknownLow = true;
                else if (s == i) taken = true;
            }
        }
        int res = ST_UNDEC;
        if (taken) {
            res = ST_TAKEN;
        } else if (!unknownLow) {
            double bw = -1.0;
            int best = -1;
            bool bestUnknown = false;
            for (int k = k0; k < k1; ++k) {
                const int j = adj[k];
                if (j <= i) continue;
                bool tk = false, unk = false;  // j's status when cell i is visited
                for (int b = off[j], e = off[j + 1]; b < e; ++b) {
                    const int m = adj[b];
                    if (m < i) {
                        const int s = st[m];
                        if (s == j) tk = true;
                        else if (s == ST_UNDEC) unk = true;
                    }
                }
                if (tk) continue;
                const double wk = w[adjFace[k]];
                if (wk > bw) {
                    bw = wk;
                    best = j;
                    bestUnknown = unk;
                }
            }
            if (best < 0) res = ST_NOPICK;
            else if (!bestUnknown) res = best;
        }
        if (res != ST_UNDEC) {
            st[i] = res;
            dec = 1;
        }
    }
    // one atomic per wavefront
    const unsigned long long ballot = __ballot(dec);
    if ((threadIdx.x & (WAVE - 1)) == 0 && ballot) atomicAdd(decided, __popcll(ballot));
}

// Aggregate creators: leaders that picked a neighbour, and isolated cells.
__global__ void k_pair_creators(int n, const int* off, const int* st, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (st[i] >= 0 || (st[i] == ST_NOPICK && off[i + 1] == off[i])) ? 1 : 0;
}

// agg of creators and the cells they picked; joiners get −1 and their strongest
// neighbour (first maximum over all neighbours, GAMG.cpp:79-85) in target.
__global__ void k_pair_assign(int n, const int* off, const int* adj, const int* adjFace, const double* w,
                              const int* st, const int* flag, const int* id, int* agg, int* target) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = st[i];
    if (flag[i]) {
        agg[i] = id[i];
        if (s >= 0) agg[s] = id[i];
        target[i] = -1;
    } else if (s == ST_NOPICK) {
        double bw = -1.0;
        int best = -1;
        for (int k = off[i]; k < off[i + 1]; ++k) {
            const double wk = w[adjFace[k]];
            if (wk > bw) {
                bw = wk;
                best = adj[k];
            }
        }
        agg[i] = -1;
        target[i] = best;
    } else {
        target[i] = -1;  // taken: written by the cell that picked it
    }
}

// Joiners copy their target's aggregate once it is known; pending counts the rest.
__global__ void k_pair_join(int n, const int* target, int* agg, int* pending) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int p = 0;
    if (i < n && target[i] >= 0 && agg[i] < 0) {
        const int a = agg[target[i]];
        if (a >= 0) agg[i] = a;
        else p = 1;
    }
    const unsigned long long ballot = __ballot(p);
    if ((threadIdx.x & (WAVE - 1)) == 0 && ballot) atomicAdd(pending, __popcll(ballot));
}

constexpr unsigned long long NO_FACE = ~0ull;

// Coarse pair (lo << 32 | hi) of every fine face, NO_FACE inside an aggregate.
__global__ void k_cface_keys(int nf, const int* l, const int* u, const int* agg, unsigned long long* keys,
                             int* vals) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int a = agg[l[f]], b = agg[u[f]];
    keys[f] = a == b ? NO_FACE
                     : ((unsigned long long)(unsigned)min(a, b) << 32) | (unsigned long long)(unsigned)max(a, b);
    vals[f] = f;
}

// First fine face of each coarse pair (sorted keys, fine faces ascending within a key).
__global__ void k_cface_first(int nf, const unsigned long long* keys, const int* vals, int* first) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    if (keys[p] != NO_FACE && (p == 0 || keys[p - 1] != keys[p])) first[vals[p]] = 1;
}

// Per coarse pair (thread of its first sorted position): its number, l/u, the weight
// sum in fine-face order, and faceMap/faceFlip of its fine faces; faces inside an
// aggregate get −1 / 0.
__global__ void k_cface_groups(int nf, const unsigned long long* keys, const int* vals, const int* cfId,
                               const double* w, const int* l, const int* agg, int* faceMap, char* flip, int* cl,
                               int* cu, double* cw) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nf) return;
    const unsigned long long key = keys[p];
    if (key == NO_FACE) {
        faceMap[vals[p]] = -1;
        flip[vals[p]] = 0;
        return;
    }
    if (p > 0 && keys[p - 1] == key) return;
    const int cf = cfId[vals[p]];
    const int lo = int(key >> 32), hi = int(key & 0xffffffffull);
    double sum = 0.0;
    for (int q = p; q < nf && keys[q] == key; ++q) {
        const int f = vals[q];
        faceMap[f] = cf;
        flip[f] = char(agg[l[f]] != lo);
        sum += w[f];
    }
    cl[cf] = lo;
    cu[cf] = hi;
    cw[cf] = sum;
}

// Face sides keyed by row: entry 2f (row l[f]) and 2f + 1 (row u[f]).
__global__ void k_adj_keys(int nf, const int* l, const int* u, int* rows, int* ent) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    rows[2 * f] = l[f];
    ent[2 * f] = 2 * f;
    rows[2 * f + 1] = u[f];
    ent[2 * f + 1] = 2 * f + 1;
}

__global__ void k_adj_count(int m, const int* rows, int* cnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) atomicAdd(&cnt[rows[p] + 1], 1);
}

__global__ void k_adj_fill(int m, const int* ent, const int* l, const int* u, int* adj, int* adjFace) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int e = ent[p], f = e >> 1;
    adj[p] = (e & 1) ? l[f] : u[f];
    adjFace[p] = f;
}

// Merged pairing: fine maps composed with the next pass's.
__global__ void k_compose_agg(int n, int* agg, const int* cagg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) agg[i] = cagg[agg[i]];
}
__global__ void k_compose_faces(int nf, int* faceMap, char* flip, const int* cMap, const char* cFlip) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int cf = faceMap[f];
    if (cf < 0) return;
    faceMap[f] = cMap[cf];
    flip[f] = char(flip[f] != cFlip[cf]);
}

using host::DevBuf;

// A level on the device: faces (l < u) with weights, adjacency in face order, and the
// maps to the next level once paired.
struct DLevel {
    int n = 0, nf = 0;
    DevBuf<int> l, u, off, adj, adjFace, agg, faceMap;
    DevBuf<double> w;
    DevBuf<char> flip;
};

struct Scratch {
    DevBuf<char> temp;
    DevBuf<int> st, flag, id, target, counter, vals, first, cfId, rows, ent, rowsOut, entOut, cnt;
    DevBuf<unsigned long long> keys, keysOut;
    int* hostCounter = nullptr;
    Scratch() { HIPCHECK(hipHostMalloc((void**)&hostCounter, sizeof(int), hipHostMallocDefault)); }
    ~Scratch() { (void)hipHostFree(hostCounter); }
    int readCounter(hipStream_t st) {
        HIPCHECK(hipMemcpyAsync(hostCounter, counter.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return *hostCounter;
    }
    void zeroCounter(hipStream_t st) { HIPCHECK(hipMemsetAsync(counter.p, 0, sizeof(int), st)); }
};

void exclusiveScan(Scratch& S, hipStream_t st, const int* in, int* out, int n) {
    size_t bytes = 0;
    HIPCHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, size_t(n), rocprim::plus<int>(), st));
    S.temp.alloc(bytes);
    HIPCHECK(rocprim::exclusive_scan(S.temp.p, bytes, in, out, 0, size_t(n), rocprim::plus<int>(), st));
}

void inclusiveScan(Scratch& S, hipStream_t st, const int* in, int* out, int n) {
    size_t bytes = 0;
    HIPCHECK(rocprim::inclusive_scan(nullptr, bytes, in, out, size_t(n), rocprim::plus<int>(), st));
    S.temp.alloc(bytes);
    HIPCHECK(rocprim::inclusive_scan(S.temp.p, bytes, in, out, size_t(n), rocprim::plus<int>(), st));
}

int readInt(Scratch& S, hipStream_t st, const int* p) {
    HIPCHECK(hipMemcpyAsync(S.hostCounter, p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return *S.hostCounter;
}

// CSR adjacency of L from its faces, the entries of a row in face order (GAMG.cpp:15-38):
// the 2·nf face sides sorted stably by row.
void adjacency(DLevel& L, Scratch& S, hipStream_t st) {
    const int m = 2 * L.nf;
    L.off.alloc(L.n + 1);
    L.adj.alloc(m);
    L.adjFace.alloc(m);
    S.cnt.alloc(L.n + 1);
    HIPCHECK(hipMemsetAsync(S.cnt.p, 0, sizeof(int) * (L.n + 1), st));
    if (m > 0) {
        S.rows.alloc(m);
        S.ent.alloc(m);
        S.rowsOut.alloc(m);
        S.entOut.alloc(m);
        hipLaunchKernelGGL(k_adj_keys, dim3(blocksFor(L.nf)), dim3(AGG_THREADS), 0, st, L.nf, L.l.p, L.u.p,
                           S.rows.p, S.ent.p);
        size_t bytes = 0;
        HIPCHECK(rocprim::radix_sort_pairs(nullptr, bytes, S.rows.p, S.rowsOut.p, S.ent.p, S.entOut.p, size_t(m),
                                           0, 32, st));
        S.temp.alloc(bytes);
        HIPCHECK(rocprim::radix_sort_pairs(S.temp.p, bytes, S.rows.p, S.rowsOut.p, S.ent.p, S.entOut.p, size_t(m),
                                           0, 32, st));
        hipLaunchKernelGGL(k_adj_count, dim3(blocksFor(m)), dim3(AGG_THREADS), 0, st, m, S.rowsOut.p, S.cnt.p);
        hipLaunchKernelGGL(k_adj_fill, dim3(blocksFor(m)), dim3(AGG_THREADS), 0, st, m, S.entOut.p, L.l.p, L.u.p,
                           L.adj.p, L.adjFace.p);
    }
    inclusiveScan(S, st, S.cnt.p, L.off.p, L.n + 1);  // cnt[0] = 0, cnt[r + 1] = degree of r
}

// One pairing pass over L: L.agg, L.faceMap, L.flip and the next level's faces (C.n,
// C.nf, C.l, C.u, C.w; no adjacency yet). Returns the number of decision rounds.
int pairPass(DLevel& L, DLevel& C, Scratch& S, hipStream_t st) {
    const int n = L.n, nf = L.nf;
    S.st.alloc(n);
    S.counter.alloc(1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocksFor(n)), dim3(AGG_THREADS), 0, st, n, S.st.p, ST_UNDEC);
    S.zeroCounter(st);
    int rounds = 0, decided = 0, batch = 16;
    while (decided < n) {
        for (int r = 0; r < batch; ++r)
            hipLaunchKernelGGL(k_pair_round, dim3(blocksFor(n)), dim3(AGG_THREADS), 0, st, n, L.off.p, L.adj.p,
                               L.adjFace.p, L.w.p, S.st.p, S.counter.p);
        rounds += batch;
        decided = S.readCounter(st);
        // the lowest undecided cell decides in every round, so n rounds always suffice
        if (rounds > n + 2 * batch) throw host::VolError(VOL_ERR_HIP, "agglomeration rounds did not finish");
        batch = std::min(batch * 2, 64);
    }
    // aggregate numbers in leader order
    S.flag.alloc(n);
    S.id.alloc(n);
    S.target.alloc(n);
    L.agg.alloc(n);
    hipLaunchKernelGGL(k_pair_creators, dim3(blocksFor(n)), dim3(AGG_THREADS), 0, st, n, L.off.p, S.st.p, S.flag.p);
    exclusiveScan(S, st, S.flag.p, S.id.p, n);
    const int nc = n > 0 ? readInt(S, st, S.id.p + (n - 1)) + readInt(S, st, S.flag.p + (n - 1)) : 0;
    hipLaunchKernelGGL(k_pair_assign, dim3(blocksFor(n)), dim3(AGG_THREADS), 0, st, n, L.off.p, L.adj.p,
                       L.adjFace.p, L.w.p, S.st.p, S.flag.p, S.id.p, L.agg.p, S.target.p);
    for (int it = 0;; ++it) {
        S.zeroCounter(st);
        hipLaunchKernelGGL(k_pair_join, dim3(blocksFor(n)), dim3(AGG_THREADS), 0, st, n, S.target.p, L.agg.p,
                           S.counter.p);
        if (S.readCounter(st) == 0) break;
        if (it > 1000) throw host::VolError(VOL_ERR_HIP, "agglomeration: joining cells did not resolve");
    }
    // coarse faces by first appearance
    L.faceMap.alloc(nf);
    L.flip.alloc(nf);
    int ncf = 0;
    if (nf > 0) {
        S.keys.alloc(nf);
        S.keysOut.alloc(nf);
        S.vals.alloc(nf);
        S.entOut.alloc(nf);
        S.first.alloc(nf);
        S.cfId.alloc(nf);
        hipLaunchKernelGGL(k_cface_keys, dim3(blocksFor(nf)), dim3(AGG_THREADS), 0, st, nf, L.l.p, L.u.p, L.agg.p,
                           S.keys.p, S.vals.p);
        size_t bytes = 0;
        HIPCHECK(rocprim::radix_sort_pairs(nullptr, bytes, S.keys.p, S.keysOut.p, S.vals.p, S.entOut.p, size_t(nf),
                                           0, 64, st));
        S.temp.alloc(bytes);
        HIPCHECK(rocprim::radix_sort_pairs(S.temp.p, bytes, S.keys.p, S.keysOut.p, S.vals.p, S.entOut.p, size_t(nf),
                                           0, 64, st));
        HIPCHECK(hipMemsetAsync(S.first.p, 0, sizeof(int) * nf, st));
        hipLaunchKernelGGL(k_cface_first, dim3(blocksFor(nf)), dim3(AGG_THREADS), 0, st, nf, S.keysOut.p,
                           S.entOut.p, S.first.p);
        exclusiveScan(S, st, S.first.p, S.cfId.p, nf);
        ncf = readInt(S, st, S.cfId.p + (nf - 1)) + readInt(S, st, S.first.p + (nf - 1));
    }
    C.n = nc;
    C.nf = ncf;
    C.l.alloc(ncf);
    C.u.alloc(ncf);
    C.w.alloc(ncf);
    if (nf > 0)
        hipLaunchKernelGGL(k_cface_groups, dim3(blocksFor(nf)), dim3(AGG_THREADS), 0, st, nf, S.keysOut.p, S.entOut.p,
                           S.cfId.p, L.w.p, L.l.p, L.agg.p, L.faceMap.p, L.flip.p, C.l.p, C.u.p, C.w.p);
    return rounds;
}

template <class T> std::vector<T> download(const DevBuf<T>& b, size_t n, hipStream_t st) {
    std::vector<T> v(n);
    if (n) HIPCHECK(hipMemcpyAsync(v.data(), b.p, sizeof(T) * n, hipMemcpyDeviceToHost, st));
    return v;
}

}  // namespace

// The GAMG hierarchy of buildHierarchy (host_common.hpp), computed on the GPU of stream
// st: the same levels, maps, coarse faces, weights and adjacency.
std::vector<host::HLevel> buildHierarchyDevice(hipStream_t st, int n, int nf, const int* own, const int* nei,
                                               const double* w, int mergeLevels, AgglomerationStats* stats) {
    constexpr int nCoarsest = 64;
    Scratch S;
    std::vector<std::unique_ptr<DLevel>> D;
    {
        auto L0 = std::make_unique<DLevel>();
        L0->n = n;
        L0->nf = nf;
        std::vector<int> l(nf), u(nf);
        for (int f = 0; f < nf; ++f) {
            l[f] = std::min(own[f], nei[f]);
            u[f] = std::max(own[f], nei[f]);
        }
        L0->l.upload(l, st);
        L0->u.upload(u, st);
        L0->w.alloc(nf);
        if (nf) HIPCHECK(hipMemcpyAsync(L0->w.p, w, sizeof(double) * nf, hipMemcpyHostToDevice, st));
        adjacency(*L0, S, st);
        D.push_back(std::move(L0));
    }
    std::vector<char> paired;  // per level: its maps are valid
    int passes = 0, rounds = 0;
    while (D.size() < 40) {
        DLevel& F = *D.back();
        const int nF = F.n;
        if (nF <= nCoarsest) break;
        auto C = std::make_unique<DLevel>();
        rounds += pairPass(F, *C, S, st);
        ++passes;
        for (int k = 1; k < mergeLevels && C->n > nCoarsest; ++k) {
            adjacency(*C, S, st);
            auto C2 = std::make_unique<DLevel>();
            rounds += pairPass(*C, *C2, S, st);
            ++passes;
            hipLaunchKernelGGL(k_compose_agg, dim3(blocksFor(F.n)), dim3(AGG_THREADS), 0, st, F.n, F.agg.p, C->agg.p);
            hipLaunchKernelGGL(k_compose_faces, dim3(blocksFor(F.nf)), dim3(AGG_THREADS), 0, st, F.nf, F.faceMap.p,
                               F.flip.p, C->faceMap.p, C->flip.p);
            C = std::move(C2);
        }
        if (C->n >= nF) break;  // no progress: F is the coarsest (its maps are dropped)
        paired.resize(D.size(), 0);
        paired.back() = 1;
        adjacency(*C, S, st);
        D.push_back(std::move(C));
    }
    paired.resize(D.size(), 0);
    std::vector<host::HLevel> H(D.size());
    for (size_t k = 0; k < D.size(); ++k) {
        DLevel& L = *D[k];
        host::HLevel& h = H[k];
        h.n = L.n;
        h.l = download(L.l, L.nf, st);
        h.u = download(L.u, L.nf, st);
        h.weight = download(L.w, L.nf, st);
        h.off = download(L.off, size_t(L.n) + 1, st);
        h.adjCell = download(L.adj, size_t(2) * L.nf, st);
        h.adjFace = download(L.adjFace, size_t(2) * L.nf, st);
        if (paired[k]) {
            h.agg = download(L.agg, L.n, st);
            h.faceMap = download(L.faceMap, L.nf, st);
            h.faceFlip = download(L.flip, L.nf, st);
        }
    }
    HIPCHECK(hipStreamSynchronize(st));
    for (auto& h : H) {  // side flags of the adjacency entries: entry of row r is the l side if l[face] == r
        h.adjIsL.resize(h.adjCell.size());
        for (int r = 0; r < h.n; ++r)
            for (int k = h.off[r]; k < h.off[r + 1]; ++k) h.adjIsL[k] = char(h.l[h.adjFace[k]] == r);
    }
    if (stats) {
        stats->passes = passes;
        stats->rounds = rounds;
    }
    return H;
}

}  // namespace vol

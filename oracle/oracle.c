/* oracle.c — CPU restatement of the reference's LDU linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Written from the reference's behaviour;
 * each function names the reference file:line it follows. Compiled with
 * -ffp-contract=off like the reference (CMakeLists.txt:11), so a*b+c is never fused.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define VSMALL 1.0e-300 /* core/Types.h:19 */
#define GREAT 1.0e15    /* core/Types.h:20 */
#define NBINS 72        /* parallel/Comm.h:44 */

/* ======================================================================== ExactSum
 * parallel/Comm.h:65-93 (add), Comm.cpp:112-164 (clear/normalize/binsToDouble/value). */
typedef struct {
    int64_t bins[NBINS];
    int64_t pending;
    int nonFinite;
    double nonFiniteValue;
} ExactSum;

static void es_clear(ExactSum* s) {
    memset(s->bins, 0, sizeof(s->bins));
    s->pending = 0;
    s->nonFinite = 0;
    s->nonFiniteValue = 0.0;
}

static void es_normalize(ExactSum* s) {
    for (int k = 0; k < NBINS - 1; ++k) {
        const int64_t carry = s->bins[k] >> 32;
        s->bins[k] -= carry * ((int64_t)1 << 32);
        s->bins[k + 1] += carry;
    }
    s->pending = 0;
}

static void es_add(ExactSum* s, double v) {
    uint64_t bits;
    memcpy(&bits, &v, sizeof(bits));
    const int ex = (int)((bits >> 52) & 0x7ff);
    if (ex == 0x7ff) {
        s->nonFinite = 1;
        s->nonFiniteValue += v;
        return;
    }
    int64_t M = (int64_t)(bits & (((uint64_t)1 << 52) - 1));
    if (ex != 0) M |= (int64_t)1 << 52;
    if (M == 0) return;
    const int p = (ex == 0 ? 1 : ex) - 1; /* v = ±M·2^(p−1074) */
    const int i = p >> 5, sh = p & 31;
    const int64_t sgn = -(int64_t)(bits >> 63);
    int64_t lo = (M & (((int64_t)1 << (32 - sh)) - 1)) << sh;
    const int64_t r = M >> (32 - sh);
    int64_t mid = r & 0xFFFFFFFF;
    int64_t hi = r >> 32;
    lo = (lo ^ sgn) - sgn;
    mid = (mid ^ sgn) - sgn;
    hi = (hi ^ sgn) - sgn;
    s->bins[i] += lo;
    s->bins[i + 1] += mid;
    s->bins[i + 2] += hi;
    if (++s->pending >= ((int64_t)1 << 29)) es_normalize(s);
}

/* Comm.cpp:135-156: carry-normalise, fold the sign, convert top bin first with ldexp. */
static double es_bins_to_double(int64_t* b, int n) {
    for (int k = 0; k < n - 1; ++k) {
        const int64_t c = b[k] >> 32;
        b[k] -= c * ((int64_t)1 << 32);
        b[k + 1] += c;
    }
    double sign = 1.0;
    if (b[n - 1] < 0) {
        sign = -1.0;
        for (int k = 0; k < n; ++k) b[k] = -b[k];
        for (int k = 0; k < n - 1; ++k) {
            const int64_t c = b[k] >> 32;
            b[k] -= c * ((int64_t)1 << 32);
            b[k + 1] += c;
        }
    }
    double r = 0.0;
    for (int k = n - 1; k >= 0; --k)
        if (b[k] != 0) r += ldexp((double)b[k], 32 * k - 1074);
    return sign * r;
}

static double es_value(const ExactSum* s) {
    if (s->nonFinite) return s->nonFiniteValue;
    int64_t b[NBINS];
    memcpy(b, s->bins, sizeof(b));
    return es_bins_to_double(b, NBINS);
}

double oracle_exact_sum(const double* v, long n) {
    ExactSum s;
    es_clear(&s);
    for (long i = 0; i < n; ++i) es_add(&s, v[i]);
    return es_value(&s);
}

double oracle_plain_sum(const double* v, long n) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += v[i];
    return s;
}

/* SumAcc (Comm.h:96-116): exact in reproducible mode, else a sequential double sum. */
typedef struct {
    int exact;
    ExactSum ex;
    double plain;
} SumAcc;
static void sa_init(SumAcc* s, int exact) {
    s->exact = exact;
    s->plain = 0.0;
    if (exact) es_clear(&s->ex);
}
static void sa_add(SumAcc* s, double v) {
    if (s->exact)
        es_add(&s->ex, v);
    else
        s->plain += v;
}
static double sa_value(SumAcc* s) { return s->exact ? es_value(&s->ex) : s->plain; }

/* ======================================================================== mesh handle */
typedef struct {
    int n, nf;
    int *l, *u;
    double* weight;
    int *off, *adjCell, *adjFace;
    char* adjIsL;
    int* agg; /* NULL on the coarsest level */
    int* faceMap;
    char* faceFlip;
} Level;

typedef struct {
    int n, nf;
    int *own, *nei;
    double* wgt;
    int *cfOff, *cfList; /* cell → internal faces, ascending (Mesh.cpp:630-641, internal part) */
    /* DIC/DILU face order (LinearSolver.cpp:29-51) */
    int npf;
    int *pfFace, *pfL, *pfU;
    char* pfOwnerIsL;
    /* GAMG */
    int mergeLevels;
    int nLevels;
    Level* levels;
    int* meshFace;
    char* ownerIsL;
} OMesh;

static void* xmalloc(size_t b) {
    void* p = malloc(b ? b : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}
static void* xcalloc(size_t n, size_t s) {
    void* p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* GAMG.cpp:15-37 */
static void build_adjacency(Level* L) {
    L->off = (int*)xcalloc((size_t)L->n + 1, sizeof(int));
    for (int f = 0; f < L->nf; ++f) {
        ++L->off[L->l[f] + 1];
        ++L->off[L->u[f] + 1];
    }
    for (int i = 0; i < L->n; ++i) L->off[i + 1] += L->off[i];
    const int nnz = L->off[L->n];
    L->adjCell = (int*)xmalloc(sizeof(int) * (size_t)nnz);
    L->adjFace = (int*)xmalloc(sizeof(int) * (size_t)nnz);
    L->adjIsL = (char*)xmalloc((size_t)nnz);
    int* fill = (int*)xmalloc(sizeof(int) * ((size_t)L->n + 1));
    memcpy(fill, L->off, sizeof(int) * (size_t)L->n);
    for (int f = 0; f < L->nf; ++f) {
        int k = fill[L->l[f]]++;
        L->adjCell[k] = L->u[f];
        L->adjFace[k] = f;
        L->adjIsL[k] = 1;
        k = fill[L->u[f]]++;
        L->adjCell[k] = L->l[f];
        L->adjFace[k] = f;
        L->adjIsL[k] = 0;
    }
    free(fill);
}

static void free_level(Level* L) {
    free(L->l);
    free(L->u);
    free(L->weight);
    free(L->off);
    free(L->adjCell);
    free(L->adjFace);
    free(L->adjIsL);
    free(L->agg);
    free(L->faceMap);
    free(L->faceFlip);
    memset(L, 0, sizeof(*L));
}

typedef struct {
    uint64_t key;
    int f;
} KeyFace;
static int cmp_keyface(const void* a, const void* b) {
    const KeyFace* x = (const KeyFace*)a;
    const KeyFace* y = (const KeyFace*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->f > y->f) - (x->f < y->f);
}
typedef struct {
    int first;
    int group;
} FirstGroup;
static int cmp_first(const void* a, const void* b) {
    const FirstGroup* x = (const FirstGroup*)a;
    const FirstGroup* y = (const FirstGroup*)b;
    return (x->first > y->first) - (x->first < y->first);
}

/* GAMG.cpp:40-94: greedy pairwise agglomeration in cell order; coarse faces are
 * numbered by first appearance in fine-face order, weights summed in face order. */
static Level agglomerate(Level* L) {
    L->agg = (int*)xmalloc(sizeof(int) * (size_t)(L->n ? L->n : 1));
    for (int i = 0; i < L->n; ++i) L->agg[i] = -1;
    int nc = 0;
    for (int i = 0; i < L->n; ++i) {
        if (L->agg[i] >= 0) continue;
        int best = -1;
        double bw = -1;
        for (int k = L->off[i]; k < L->off[i + 1]; ++k) {
            const int j = L->adjCell[k];
            if (L->agg[j] < 0 && j != i && L->weight[L->adjFace[k]] > bw) {
                bw = L->weight[L->adjFace[k]];
                best = j;
            }
        }
        if (best >= 0) {
            L->agg[i] = L->agg[best] = nc++;
            continue;
        }
        for (int k = L->off[i]; k < L->off[i + 1]; ++k) {
            if (L->weight[L->adjFace[k]] > bw) {
                bw = L->weight[L->adjFace[k]];
                best = L->adjCell[k];
            }
        }
        L->agg[i] = best >= 0 ? L->agg[best] : nc++;
    }
    Level C;
    memset(&C, 0, sizeof(C));
    C.n = nc;
    L->faceMap = (int*)xmalloc(sizeof(int) * (size_t)(L->nf ? L->nf : 1));
    L->faceFlip = (char*)xcalloc((size_t)(L->nf ? L->nf : 1), 1);
    KeyFace* kf = (KeyFace*)xmalloc(sizeof(KeyFace) * (size_t)(L->nf ? L->nf : 1));
    int m = 0;
    for (int f = 0; f < L->nf; ++f) {
        L->faceMap[f] = -1;
        const int a = L->agg[L->l[f]], b = L->agg[L->u[f]];
        if (a == b) continue;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        kf[m].key = ((uint64_t)(uint32_t)lo << 32) | (uint64_t)(uint32_t)hi;
        kf[m].f = f;
        ++m;
        L->faceFlip[f] = (char)(a != lo);
    }
    qsort(kf, (size_t)m, sizeof(KeyFace), cmp_keyface);
    /* groups of equal key; order groups by their first (smallest) fine face */
    int ng = 0;
    for (int a = 0; a < m; ++a)
        if (a == 0 || kf[a].key != kf[a - 1].key) ++ng;
    FirstGroup* fg = (FirstGroup*)xmalloc(sizeof(FirstGroup) * (size_t)(ng ? ng : 1));
    int* gStart = (int*)xmalloc(sizeof(int) * ((size_t)ng + 1));
    ng = 0;
    for (int a = 0; a < m; ++a)
        if (a == 0 || kf[a].key != kf[a - 1].key) {
            gStart[ng] = a;
            fg[ng].first = kf[a].f;
            fg[ng].group = ng;
            ++ng;
        }
    gStart[ng] = m;
    qsort(fg, (size_t)ng, sizeof(FirstGroup), cmp_first);
    C.nf = ng;
    C.l = (int*)xmalloc(sizeof(int) * (size_t)(ng ? ng : 1));
    C.u = (int*)xmalloc(sizeof(int) * (size_t)(ng ? ng : 1));
    C.weight = (double*)xmalloc(sizeof(double) * (size_t)(ng ? ng : 1));
    for (int cf = 0; cf < ng; ++cf) {
        const int g = fg[cf].group;
        C.l[cf] = (int)(kf[gStart[g]].key >> 32);
        C.u[cf] = (int)(kf[gStart[g]].key & 0xFFFFFFFFu);
        double w = 0.0;
        for (int a = gStart[g]; a < gStart[g + 1]; ++a) {
            L->faceMap[kf[a].f] = cf;
            w += L->weight[kf[a].f];
        }
        C.weight[cf] = w;
    }
    free(kf);
    free(fg);
    free(gStart);
    return C;
}

/* GAMG.cpp:206-252 (serial): level 0 from owned internal faces, then mergeLevels
 * pairings per level until ≤ 64 cells or no progress; at most 40 levels. */
static void build_hierarchy(OMesh* m) {
    const int nCoarsest = 64;
    Level* lv = (Level*)xcalloc(40, sizeof(Level));
    int nL = 0;
    Level* L0 = &lv[0];
    L0->n = m->n;
    L0->l = (int*)xmalloc(sizeof(int) * (size_t)(m->nf ? m->nf : 1));
    L0->u = (int*)xmalloc(sizeof(int) * (size_t)(m->nf ? m->nf : 1));
    L0->weight = (double*)xmalloc(sizeof(double) * (size_t)(m->nf ? m->nf : 1));
    m->meshFace = (int*)xmalloc(sizeof(int) * (size_t)(m->nf ? m->nf : 1));
    m->ownerIsL = (char*)xmalloc((size_t)(m->nf ? m->nf : 1));
    int k0 = 0;
    for (int f = 0; f < m->nf; ++f) {
        const int o = m->own[f], nb = m->nei[f];
        if (o >= m->n || nb >= m->n || o == nb) continue;
        L0->l[k0] = o < nb ? o : nb;
        L0->u[k0] = o < nb ? nb : o;
        L0->weight[k0] = m->wgt[f];
        m->meshFace[k0] = f;
        m->ownerIsL[k0] = (char)(o < nb);
        ++k0;
    }
    L0->nf = k0;
    build_adjacency(L0);
    nL = 1;
    while (nL < 40) {
        Level* F = &lv[nL - 1];
        const double nF = (double)F->n;
        if (nF <= 64.0) break;
        Level C = agglomerate(F);
        for (int k = 1; k < m->mergeLevels && C.n > nCoarsest; ++k) {
            build_adjacency(&C);
            Level C2 = agglomerate(&C);
            for (int i = 0; i < F->n; ++i) F->agg[i] = C.agg[F->agg[i]];
            for (int f = 0; f < F->nf; ++f) {
                const int cf = F->faceMap[f];
                if (cf < 0) continue;
                F->faceMap[f] = C.faceMap[cf];
                F->faceFlip[f] = (char)(F->faceFlip[f] != C.faceFlip[cf]);
            }
            free_level(&C);
            C = C2;
        }
        if ((double)C.n >= nF) {
            free(F->agg);
            F->agg = NULL;
            free_level(&C);
            break;
        }
        build_adjacency(&C);
        lv[nL++] = C;
    }
    m->levels = lv;
    m->nLevels = nL;
}

typedef struct {
    int u, l, f;
} PF;
static int cmp_pf(const void* a, const void* b) {
    const PF* x = (const PF*)a;
    const PF* y = (const PF*)b;
    if (x->u != y->u) return x->u < y->u ? -1 : 1;
    if (x->l != y->l) return x->l < y->l ? -1 : 1;
    return (x->f > y->f) - (x->f < y->f);
}

void* oracle_mesh_new(int n, int nf, const int* own, const int* nei, const double* wgt, int mergeLevels) {
    OMesh* m = (OMesh*)xcalloc(1, sizeof(OMesh));
    m->n = n;
    m->nf = nf;
    m->own = (int*)xmalloc(sizeof(int) * (size_t)(nf ? nf : 1));
    m->nei = (int*)xmalloc(sizeof(int) * (size_t)(nf ? nf : 1));
    memcpy(m->own, own, sizeof(int) * (size_t)nf);
    memcpy(m->nei, nei, sizeof(int) * (size_t)nf);
    if (wgt) {
        m->wgt = (double*)xmalloc(sizeof(double) * (size_t)(nf ? nf : 1));
        memcpy(m->wgt, wgt, sizeof(double) * (size_t)nf);
    }
    m->mergeLevels = mergeLevels < 1 ? 1 : mergeLevels;
    m->cfOff = (int*)xcalloc((size_t)n + 1, sizeof(int));
    for (int f = 0; f < nf; ++f) {
        ++m->cfOff[own[f] + 1];
        ++m->cfOff[nei[f] + 1];
    }
    for (int c = 0; c < n; ++c) m->cfOff[c + 1] += m->cfOff[c];
    m->cfList = (int*)xmalloc(sizeof(int) * (size_t)(m->cfOff[n] ? m->cfOff[n] : 1));
    int* fill = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fill, m->cfOff, sizeof(int) * (size_t)n);
    for (int f = 0; f < nf; ++f) {
        m->cfList[fill[own[f]]++] = f;
        m->cfList[fill[nei[f]]++] = f;
    }
    free(fill);
    /* LinearSolver.cpp:34-51: faces with both cells owned, o != n, sorted by (u, l) */
    PF* tmp = (PF*)xmalloc(sizeof(PF) * (size_t)(nf ? nf : 1));
    int np = 0;
    for (int f = 0; f < nf; ++f) {
        const int o = own[f], nb = nei[f];
        if (o >= n || nb >= n || o == nb) continue;
        tmp[np].u = o > nb ? o : nb;
        tmp[np].l = o < nb ? o : nb;
        tmp[np].f = f;
        ++np;
    }
    qsort(tmp, (size_t)np, sizeof(PF), cmp_pf);
    m->npf = np;
    m->pfFace = (int*)xmalloc(sizeof(int) * (size_t)(np ? np : 1));
    m->pfL = (int*)xmalloc(sizeof(int) * (size_t)(np ? np : 1));
    m->pfU = (int*)xmalloc(sizeof(int) * (size_t)(np ? np : 1));
    m->pfOwnerIsL = (char*)xmalloc((size_t)(np ? np : 1));
    for (int k = 0; k < np; ++k) {
        m->pfFace[k] = tmp[k].f;
        m->pfU[k] = tmp[k].u;
        m->pfL[k] = tmp[k].l;
        m->pfOwnerIsL[k] = (char)(own[tmp[k].f] == tmp[k].l);
    }
    free(tmp);
    if (m->wgt) build_hierarchy(m);
    return m;
}

void oracle_mesh_free(void* h) {
    OMesh* m = (OMesh*)h;
    if (!m) return;
    free(m->own);
    free(m->nei);
    free(m->wgt);
    free(m->cfOff);
    free(m->cfList);
    free(m->pfFace);
    free(m->pfL);
    free(m->pfU);
    free(m->pfOwnerIsL);
    for (int k = 0; k < m->nLevels; ++k) free_level(&m->levels[k]);
    free(m->levels);
    free(m->meshFace);
    free(m->ownerIsL);
    free(m);
}

int oracle_gamg_nlevels(void* h) { return ((OMesh*)h)->nLevels; }

void oracle_gamg_level_sizes(void* h, int k, int* n, int* nf) {
    OMesh* m = (OMesh*)h;
    *n = m->levels[k].n;
    *nf = m->levels[k].nf;
}

void oracle_gamg_level(void* h, int k, int* agg, int* faceMap, int* faceFlip, int* l, int* u, double* weight) {
    const Level* L = &((OMesh*)h)->levels[k];
    for (int i = 0; i < L->n; ++i) agg[i] = L->agg ? L->agg[i] : -1;
    for (int f = 0; f < L->nf; ++f) {
        faceMap[f] = L->faceMap && L->agg ? L->faceMap[f] : -2;
        faceFlip[f] = L->faceFlip && L->agg ? L->faceFlip[f] : 0;
        l[f] = L->l[f];
        u[f] = L->u[f];
        weight[f] = L->weight[f];
    }
}

/* ======================================================================== Amul
 * LinearSolver.cpp:121-133: y = diag∘x, then faces ascending. */
void oracle_amul(void* h, const double* diag, const double* upper, const double* lower, const double* x, double* y) {
    const OMesh* m = (const OMesh*)h;
    for (int c = 0; c < m->n; ++c) y[c] = diag[c] * x[c];
    for (int f = 0; f < m->nf; ++f) {
        y[m->own[f]] += upper[f] * x[m->nei[f]];
        y[m->nei[f]] += lower[f] * x[m->own[f]];
    }
}

/* ======================================================================== GAMG preconditioner */
typedef struct {
    const OMesh* m;
    double** cd; /* per-level diag */
    double** cu; /* per-level upper (l,u) */
    double** cl; /* per-level lower */
    double* lu;  /* coarsest dense LU, row-major */
    int nCoarse;
    double **rbuf, **xbuf, **tmp, **xold;
    double* g;
    int smoother; /* 0 GS, 1 Jacobi */
    double omega;
    int nPre, nPost;
} Gamg;

static void gamg_free(Gamg* G) {
    if (!G) return;
    const int nL = G->m->nLevels;
    for (int k = 0; k < nL; ++k) {
        free(G->cd[k]);
        free(G->cu[k]);
        free(G->cl[k]);
        free(G->rbuf[k]);
        free(G->xbuf[k]);
        free(G->tmp[k]);
        free(G->xold[k]);
    }
    free(G->cd);
    free(G->cu);
    free(G->cl);
    free(G->rbuf);
    free(G->xbuf);
    free(G->tmp);
    free(G->xold);
    free(G->lu);
    free(G->g);
    free(G);
}

/* GAMG.cpp:263-353 (serial): reorient level-0 coefficients to (l,u), Galerkin
 * coarsening in fine order, dense no-pivot Doolittle LU of the coarsest level. */
static Gamg* gamg_new(const OMesh* m, const double* diag, const double* upper, const double* lower, int* err) {
    *err = 0;
    const int nL = m->nLevels;
    Gamg* G = (Gamg*)xcalloc(1, sizeof(Gamg));
    G->m = m;
    G->cd = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->cu = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->cl = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->rbuf = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->xbuf = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->tmp = (double**)xcalloc((size_t)nL, sizeof(double*));
    G->xold = (double**)xcalloc((size_t)nL, sizeof(double*));
    {
        const Level* L = &m->levels[0];
        G->cd[0] = (double*)xmalloc(sizeof(double) * (size_t)L->n);
        memcpy(G->cd[0], diag, sizeof(double) * (size_t)L->n);
        G->cu[0] = (double*)xmalloc(sizeof(double) * (size_t)(L->nf ? L->nf : 1));
        G->cl[0] = (double*)xmalloc(sizeof(double) * (size_t)(L->nf ? L->nf : 1));
        for (int k = 0; k < L->nf; ++k) {
            const int f = m->meshFace[k];
            G->cu[0][k] = m->ownerIsL[k] ? upper[f] : lower[f];
            G->cl[0][k] = m->ownerIsL[k] ? lower[f] : upper[f];
        }
    }
    for (int lev = 0; lev + 1 < nL; ++lev) {
        const Level* L = &m->levels[lev];
        const Level* LC = &m->levels[lev + 1];
        double* Cd = (double*)xcalloc((size_t)LC->n, sizeof(double));
        double* Cu = (double*)xcalloc((size_t)(LC->nf ? LC->nf : 1), sizeof(double));
        double* Cl = (double*)xcalloc((size_t)(LC->nf ? LC->nf : 1), sizeof(double));
        for (int i = 0; i < L->n; ++i) Cd[L->agg[i]] += G->cd[lev][i];
        for (int f = 0; f < L->nf; ++f) {
            const int cf = L->faceMap[f];
            if (cf < 0) {
                Cd[L->agg[L->l[f]]] += G->cu[lev][f] + G->cl[lev][f];
            } else if (!L->faceFlip[f]) {
                Cu[cf] += G->cu[lev][f];
                Cl[cf] += G->cl[lev][f];
            } else {
                Cu[cf] += G->cl[lev][f];
                Cl[cf] += G->cu[lev][f];
            }
        }
        G->cd[lev + 1] = Cd;
        G->cu[lev + 1] = Cu;
        G->cl[lev + 1] = Cl;
    }
    const Level* LC = &m->levels[nL - 1];
    const int n = LC->n;
    G->nCoarse = n;
    G->lu = (double*)xcalloc((size_t)n * (size_t)n, sizeof(double));
    double* lu = G->lu;
    for (int i = 0; i < n; ++i) lu[(size_t)i * n + i] = G->cd[nL - 1][i];
    for (int f = 0; f < LC->nf; ++f) {
        lu[(size_t)LC->l[f] * n + LC->u[f]] += G->cu[nL - 1][f];
        lu[(size_t)LC->u[f] * n + LC->l[f]] += G->cl[nL - 1][f];
    }
    for (int k = 0; k < n; ++k) {
        const double piv = lu[(size_t)k * n + k];
        if (fabs(piv) < VSMALL) {
            *err = -4; /* GAMG: singular coarsest matrix (GAMG.cpp:336) */
            gamg_free(G);
            return NULL;
        }
        for (int i = k + 1; i < n; ++i) {
            const double fct = lu[(size_t)i * n + k] / piv;
            lu[(size_t)i * n + k] = fct;
            if (fct == 0.0) continue;
            for (int j = k + 1; j < n; ++j) lu[(size_t)i * n + j] -= fct * lu[(size_t)k * n + j];
        }
    }
    for (int lev = 0; lev < nL; ++lev) {
        const int nl = m->levels[lev].n;
        G->rbuf[lev] = (double*)xcalloc((size_t)nl, sizeof(double));
        G->xbuf[lev] = (double*)xcalloc((size_t)nl, sizeof(double));
        G->tmp[lev] = (double*)xcalloc((size_t)nl, sizeof(double));
        G->xold[lev] = (double*)xcalloc((size_t)nl, sizeof(double));
    }
    G->g = (double*)xcalloc((size_t)(n ? n : 1), sizeof(double));
    G->smoother = 0;
    G->omega = 1.0;
    G->nPre = G->nPost = 2;
    return G;
}

/* one Gauss–Seidel row (GAMG.cpp:368-380, serial part) */
static void gs_row(const Level* L, const double* Ad, const double* Au, const double* Al, const double* r, double* x,
                   int i) {
    double s = r[i];
    for (int k = L->off[i]; k < L->off[i + 1]; ++k) {
        const int f = L->adjFace[k];
        s -= (L->adjIsL[k] ? Au[f] : Al[f]) * x[L->adjCell[k]];
    }
    x[i] = s / Ad[i];
}

/* Project smoother: damped Jacobi x_i ← x_i + ω·(J_i − x_i), J_i = (r_i − Σ_k a_ik x_k)/d_i with the
 * off-diagonal sum in the same row order as the reference's smoother rows (GAMG.cpp:368-379).
 * fromZero: the sweep that starts from x = 0 is defined as x_i = ω·(r_i·rD_i) with rD_i = 1/d_i
 * (the reciprocal the diagonal preconditioner uses, LinearSolver.cpp:78). */
static void jacobi_sweep(const Level* L, const double* Ad, const double* Au, const double* Al, const double* r,
                         double* x, double* xold, double omega, int fromZero) {
    if (fromZero) {
        for (int i = 0; i < L->n; ++i) {
            const double rD = 1.0 / Ad[i];
            x[i] = omega * (r[i] * rD);
        }
        return;
    }
    memcpy(xold, x, sizeof(double) * (size_t)L->n);
    for (int i = 0; i < L->n; ++i) {
        double s = r[i];
        for (int k = L->off[i]; k < L->off[i + 1]; ++k) {
            const int f = L->adjFace[k];
            s -= (L->adjIsL[k] ? Au[f] : Al[f]) * xold[L->adjCell[k]];
        }
        const double J = s / Ad[i];
        x[i] = xold[i] + omega * (J - xold[i]);
    }
}

/* GAMG.cpp:387-401 (serial) */
static void level_residual(const Level* L, const double* Ad, const double* Au, const double* Al, const double* r,
                           const double* x, double* res) {
    for (int i = 0; i < L->n; ++i) res[i] = r[i] - Ad[i] * x[i];
    for (int f = 0; f < L->nf; ++f) {
        res[L->l[f]] -= Au[f] * x[L->u[f]];
        res[L->u[f]] -= Al[f] * x[L->l[f]];
    }
}

/* GAMG.cpp:403-443 */
static void vcycle(Gamg* G, int lev, const double* r, double* x) {
    const OMesh* m = G->m;
    const int nL = m->nLevels;
    const Level* L = &m->levels[lev];
    if (lev == nL - 1) {
        const int n = G->nCoarse;
        double* g = G->g;
        for (int i = 0; i < n; ++i) g[i] = 0.0;
        for (int i = 0; i < L->n; ++i) g[i] = r[i];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) g[i] -= G->lu[(size_t)i * n + k] * g[k];
        if (G->smoother == 0) {
            for (int i = n - 1; i >= 0; --i) {
                for (int k = i + 1; k < n; ++k) g[i] -= G->lu[(size_t)i * n + k] * g[k];
                g[i] /= G->lu[(size_t)i * n + i];
            }
        } else {
            /* project variant: backward substitution by columns (each row subtracts in
             * descending k, then divides), the order of the one-workgroup GPU solve */
            for (int k = n - 1; k >= 0; --k) {
                g[k] /= G->lu[(size_t)k * n + k];
                for (int i = 0; i < k; ++i) g[i] -= G->lu[(size_t)i * n + k] * g[k];
            }
        }
        for (int i = 0; i < L->n; ++i) x[i] = g[i];
        return;
    }
    const double *Ad = G->cd[lev], *Au = G->cu[lev], *Al = G->cl[lev];
    if (G->smoother == 0) {
        for (int i = 0; i < L->n; ++i) x[i] = 0.0;
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < L->n; ++i) gs_row(L, Ad, Au, Al, r, x, i);
    } else {
        for (int s = 0; s < G->nPre; ++s) jacobi_sweep(L, Ad, Au, Al, r, x, G->xold[lev], G->omega, s == 0);
        if (G->nPre == 0)
            for (int i = 0; i < L->n; ++i) x[i] = 0.0;
    }
    double* res = G->tmp[lev];
    level_residual(L, Ad, Au, Al, r, x, res);
    double* rc = G->rbuf[lev + 1];
    double* xc = G->xbuf[lev + 1];
    const int nc = m->levels[lev + 1].n;
    for (int i = 0; i < nc; ++i) rc[i] = 0.0;
    for (int i = 0; i < L->n; ++i) rc[L->agg[i]] += res[i];
    vcycle(G, lev + 1, rc, xc);
    for (int i = 0; i < L->n; ++i) x[i] += xc[L->agg[i]];
    if (G->smoother == 0) {
        for (int s = 0; s < 2; ++s)
            for (int i = L->n - 1; i >= 0; --i) gs_row(L, Ad, Au, Al, r, x, i);
    } else {
        for (int s = 0; s < G->nPost; ++s) jacobi_sweep(L, Ad, Au, Al, r, x, G->xold[lev], G->omega, 0);
    }
}

/* GAMG.cpp:445-449 */
static void gamg_apply(Gamg* G, const double* r, double* w) {
    double* x0 = G->xbuf[0];
    vcycle(G, 0, r, x0);
    memcpy(w, x0, sizeof(double) * (size_t)G->m->levels[0].n);
}

void oracle_gamg_coeffs(void* h, const double* diag, const double* upper, const double* lower, int k, double* cdiag,
                        double* cupper, double* clower) {
    const OMesh* m = (const OMesh*)h;
    int err = 0;
    Gamg* G = gamg_new(m, diag, upper, lower, &err);
    if (!G) return;
    memcpy(cdiag, G->cd[k], sizeof(double) * (size_t)m->levels[k].n);
    memcpy(cupper, G->cu[k], sizeof(double) * (size_t)m->levels[k].nf);
    memcpy(clower, G->cl[k], sizeof(double) * (size_t)m->levels[k].nf);
    gamg_free(G);
}

int oracle_gamg_apply(void* h, const double* diag, const double* upper, const double* lower, const double* r, double* w,
                      int smoother, double omega, int nPre, int nPost) {
    const OMesh* m = (const OMesh*)h;
    if (!m->nLevels) return -5;
    int err = 0;
    Gamg* G = gamg_new(m, diag, upper, lower, &err);
    if (!G) return err;
    G->smoother = smoother;
    G->omega = omega;
    G->nPre = nPre;
    G->nPost = nPost;
    gamg_apply(G, r, w);
    gamg_free(G);
    return 0;
}

/* ======================================================================== preconditioner
 * LinearSolver.cpp:56-112 */
typedef struct {
    int type; /* 0 none, 1 diagonal, 2 DIC, 3 DILU, 4 GAMG */
    const OMesh* m;
    const double *upper, *lower;
    double* rD;
    Gamg* gamg;
} Precond;

static int precond_init(Precond* P, const OMesh* m, const double* diag, const double* upper, const double* lower,
                        const char* type, int smoother, double omega, int nPre, int nPost) {
    memset(P, 0, sizeof(*P));
    P->m = m;
    P->upper = upper;
    P->lower = lower;
    const int n = m->n;
    P->rD = (double*)xcalloc((size_t)n, sizeof(double));
    if (!strcmp(type, "none")) {
        P->type = 0;
        return 0;
    }
    if (!strcmp(type, "GAMG")) {
        if (!m->nLevels) return -5; /* no agglomeration weights given */
        int err = 0;
        P->type = 4;
        P->gamg = gamg_new(m, diag, upper, lower, &err);
        if (!P->gamg) return err;
        P->gamg->smoother = smoother;
        P->gamg->omega = omega;
        P->gamg->nPre = nPre;
        P->gamg->nPost = nPost;
        return 0;
    }
    for (int c = 0; c < n; ++c) P->rD[c] = diag[c];
    if (!strcmp(type, "DIC") || !strcmp(type, "DILU")) {
        P->type = !strcmp(type, "DIC") ? 2 : 3;
        for (int k = 0; k < m->npf; ++k) {
            const int f = m->pfFace[k];
            const double Alu = m->pfOwnerIsL[k] ? upper[f] : lower[f];
            const double Aul = m->pfOwnerIsL[k] ? lower[f] : upper[f];
            P->rD[m->pfU[k]] -= Alu * Aul / P->rD[m->pfL[k]];
        }
    } else if (!strcmp(type, "diagonal")) {
        P->type = 1;
    } else {
        return -2; /* unknown preconditioner (LinearSolver.cpp:76) */
    }
    for (int c = 0; c < n; ++c) P->rD[c] = 1.0 / P->rD[c];
    return 0;
}

static void precond_free(Precond* P) {
    free(P->rD);
    gamg_free(P->gamg);
    P->rD = NULL;
    P->gamg = NULL;
}

static void precond_apply(const Precond* P, const double* r, double* w) {
    const OMesh* m = P->m;
    const int n = m->n;
    if (P->type == 0) {
        for (int c = 0; c < n; ++c) w[c] = r[c];
        return;
    }
    if (P->type == 4) {
        gamg_apply(P->gamg, r, w);
        return;
    }
    for (int c = 0; c < n; ++c) w[c] = P->rD[c] * r[c];
    if (P->type == 1) return;
    for (int k = 0; k < m->npf; ++k) {
        const int f = m->pfFace[k];
        const double Aul = m->pfOwnerIsL[k] ? P->lower[f] : P->upper[f];
        w[m->pfU[k]] -= P->rD[m->pfU[k]] * Aul * w[m->pfL[k]];
    }
    for (int k = m->npf - 1; k >= 0; --k) {
        const int f = m->pfFace[k];
        const double Alu = m->pfOwnerIsL[k] ? P->upper[f] : P->lower[f];
        w[m->pfL[k]] -= P->rD[m->pfL[k]] * Alu * w[m->pfU[k]];
    }
}

/* ======================================================================== solveLdu */
typedef struct {
    double tolerance, relTol;
    int maxIter, minIter;
} Ctrl;
typedef struct {
    double initialResidual, finalResidual;
    int iterations;
} Perf;

/* LinearSolver.cpp:114-117 */
static int converged(const Ctrl* c, const Perf* p) {
    if (p->iterations < c->minIter) return 0;
    return p->finalResidual < c->tolerance || (c->relTol > 0 && p->finalResidual < c->relTol * p->initialResidual);
}

static double sum_prod(const double* a, const double* b, int n, int exact) {
    SumAcc s;
    sa_init(&s, exact);
    for (int c = 0; c < n; ++c) sa_add(&s, a[c] * b[c]);
    return sa_value(&s);
}
static double sum_mag(const double* a, int n, int exact) {
    SumAcc s;
    sa_init(&s, exact);
    for (int c = 0; c < n; ++c) sa_add(&s, fabs(a[c]));
    return sa_value(&s);
}

static void set_err(char* err, int len, const char* msg) {
    if (err && len > 0) {
        strncpy(err, msg, (size_t)len - 1);
        err[len - 1] = 0;
    }
}

int oracle_solve(void* h, const double* diag, const double* upper, const double* lower, int symmetric,
                 const double* b, double* x, const char* solverIn, const char* precondIn, double tol, double relTol,
                 int maxIter, int minIter, int nSweeps, int mode, int gamgSmoother, double omega, int nPre, int nPost,
                 double* perfOut, int* itersOut, int* convOut, double* hist, char* err, int errLen) {
    const OMesh* m = (const OMesh*)h;
    const int n = m->n;
    const int exact = mode != 0;
    char solver[64], precond[64];
    strncpy(solver, solverIn, sizeof(solver) - 1);
    solver[sizeof(solver) - 1] = 0;
    strncpy(precond, precondIn, sizeof(precond) - 1);
    precond[sizeof(precond) - 1] = 0;
    if (mode == 1) { /* LinearSolver.cpp:138-148 */
        if (strcmp(precond, "none")) strcpy(precond, "diagonal");
        if (!strcmp(solver, "GaussSeidel") || !strcmp(solver, "symGaussSeidel")) strcpy(solver, "Jacobi");
    }
    Ctrl ctrl = {tol, relTol, maxIter, minIter};
    Perf perf = {0, 0, 0};
    int rc = 0;

    double* wA = (double*)xcalloc((size_t)n, sizeof(double));
    double* rA = (double*)xcalloc((size_t)n, sizeof(double));
    double* pA = (double*)xcalloc((size_t)n, sizeof(double));
    double* xr = (double*)xcalloc((size_t)n, sizeof(double));
    oracle_amul(h, diag, upper, lower, x, wA);
    for (int c = 0; c < n; ++c) rA[c] = b[c] - wA[c];
    /* normalisation factor, LinearSolver.cpp:157-166 */
    SumAcc xs;
    sa_init(&xs, exact);
    for (int c = 0; c < n; ++c) sa_add(&xs, x[c]);
    const double xRef = sa_value(&xs) / (double)n;
    for (int c = 0; c < n; ++c) xr[c] = xRef;
    oracle_amul(h, diag, upper, lower, xr, pA);
    SumAcc nfs;
    sa_init(&nfs, exact);
    for (int c = 0; c < n; ++c) sa_add(&nfs, fabs(wA[c] - pA[c]) + fabs(b[c] - pA[c]));
    const double normFactor = sa_value(&nfs) + 1e-20;
    const double residualSum = sum_mag(rA, n, exact);
    perf.initialResidual = residualSum / normFactor;
    perf.finalResidual = perf.initialResidual;
    if (hist) hist[0] = perf.initialResidual;
    int earlyDone = perf.initialResidual < ctrl.tolerance && ctrl.minIter <= 0;

    if (!earlyDone) {
        if (!strcmp(solver, "PCG")) { /* LinearSolver.cpp:176-201 */
            if (!symmetric) {
                set_err(err, errLen, "PCG requires a symmetric matrix; use PBiCGStab");
                rc = -3;
                goto done;
            }
            Precond P;
            rc = precond_init(&P, m, diag, upper, lower, precond, gamgSmoother, omega, nPre, nPost);
            if (rc) {
                set_err(err, errLen, rc == -2 ? "unknown preconditioner" : "GAMG setup failed");
                precond_free(&P);
                goto done;
            }
            double* w = (double*)xcalloc((size_t)n, sizeof(double));
            double* p = (double*)xcalloc((size_t)n, sizeof(double));
            double wArA = GREAT, wArAold;
            do {
                wArAold = wArA;
                precond_apply(&P, rA, w);
                wArA = sum_prod(w, rA, n, exact);
                if (perf.iterations == 0) {
                    for (int c = 0; c < n; ++c) p[c] = w[c];
                } else {
                    const double beta = wArA / wArAold;
                    for (int c = 0; c < n; ++c) p[c] = w[c] + beta * p[c];
                }
                oracle_amul(h, diag, upper, lower, p, wA);
                const double wApA = sum_prod(wA, p, n, exact);
                if (fabs(wApA) < VSMALL) break;
                const double alpha = wArA / wApA;
                for (int c = 0; c < n; ++c) {
                    x[c] += alpha * p[c];
                    rA[c] -= alpha * wA[c];
                }
                perf.finalResidual = sum_mag(rA, n, exact) / normFactor;
                ++perf.iterations;
                if (hist) hist[perf.iterations] = perf.finalResidual;
            } while (perf.iterations < ctrl.maxIter && !converged(&ctrl, &perf));
            free(w);
            free(p);
            precond_free(&P);
        } else if (!strcmp(solver, "PBiCGStab")) { /* LinearSolver.cpp:202-237 */
            Precond P;
            rc = precond_init(&P, m, diag, upper, lower, !strcmp(precond, "DIC") ? "DILU" : precond, gamgSmoother,
                              omega, nPre, nPost);
            if (rc) {
                set_err(err, errLen, rc == -2 ? "unknown preconditioner" : "GAMG setup failed");
                precond_free(&P);
                goto done;
            }
            double* rA0 = (double*)xmalloc(sizeof(double) * (size_t)n);
            memcpy(rA0, rA, sizeof(double) * (size_t)n);
            double* pA2 = (double*)xcalloc((size_t)n, sizeof(double));
            double* yA = (double*)xcalloc((size_t)n, sizeof(double));
            double* AyA = (double*)xcalloc((size_t)n, sizeof(double));
            double* sA = (double*)xcalloc((size_t)n, sizeof(double));
            double* zA = (double*)xcalloc((size_t)n, sizeof(double));
            double* tA = (double*)xcalloc((size_t)n, sizeof(double));
            double rA0rA = 0, alpha = 0, omegaB = 0;
            do {
                const double rA0rAold = rA0rA;
                rA0rA = sum_prod(rA0, rA, n, exact);
                if (perf.iterations == 0) {
                    for (int c = 0; c < n; ++c) pA2[c] = rA[c];
                } else {
                    if (fabs(rA0rAold) < VSMALL || fabs(omegaB) < VSMALL) break;
                    const double beta = (rA0rA / rA0rAold) * (alpha / omegaB);
                    for (int c = 0; c < n; ++c) pA2[c] = rA[c] + beta * (pA2[c] - omegaB * AyA[c]);
                }
                precond_apply(&P, pA2, yA);
                oracle_amul(h, diag, upper, lower, yA, AyA);
                const double rA0AyA = sum_prod(rA0, AyA, n, exact);
                if (fabs(rA0AyA) < VSMALL) break;
                alpha = rA0rA / rA0AyA;
                for (int c = 0; c < n; ++c) sA[c] = rA[c] - alpha * AyA[c];
                ++perf.iterations;
                perf.finalResidual = sum_mag(sA, n, exact) / normFactor;
                if (converged(&ctrl, &perf)) {
                    for (int c = 0; c < n; ++c) x[c] += alpha * yA[c];
                    if (hist) hist[perf.iterations] = perf.finalResidual;
                    break;
                }
                precond_apply(&P, sA, zA);
                oracle_amul(h, diag, upper, lower, zA, tA);
                const double tAtA = sum_prod(tA, tA, n, exact);
                omegaB = tAtA > VSMALL ? sum_prod(tA, sA, n, exact) / tAtA : 0.0;
                for (int c = 0; c < n; ++c) {
                    x[c] += alpha * yA[c] + omegaB * zA[c];
                    rA[c] = sA[c] - omegaB * tA[c];
                }
                perf.finalResidual = sum_mag(rA, n, exact) / normFactor;
                if (hist) hist[perf.iterations] = perf.finalResidual;
            } while (perf.iterations < ctrl.maxIter && !converged(&ctrl, &perf));
            free(rA0);
            free(pA2);
            free(yA);
            free(AyA);
            free(sA);
            free(zA);
            free(tA);
            precond_free(&P);
        } else if (!strcmp(solver, "GaussSeidel") || !strcmp(solver, "symGaussSeidel") || !strcmp(solver, "Jacobi")) {
            /* LinearSolver.cpp:238-278 */
            const int jac = !strcmp(solver, "Jacobi");
            const int sym = !strcmp(solver, "symGaussSeidel");
            double* xOld = (double*)xcalloc((size_t)n, sizeof(double));
            do {
                for (int sw = 0; sw < nSweeps; ++sw) {
                    const double* xs = x;
                    if (jac) {
                        memcpy(xOld, x, sizeof(double) * (size_t)n);
                        xs = xOld;
                    }
                    for (int pass = 0; pass < (sym ? 2 : 1); ++pass) {
                        for (int t = 0; t < n; ++t) {
                            const int c = pass == 0 ? t : n - 1 - t;
                            double r = b[c];
                            for (int k = m->cfOff[c]; k < m->cfOff[c + 1]; ++k) {
                                const int f = m->cfList[k];
                                if (m->own[f] == c && m->nei[f] == c)
                                    r -= (upper[f] + lower[f]) * xs[c];
                                else if (m->own[f] == c)
                                    r -= upper[f] * xs[m->nei[f]];
                                else
                                    r -= lower[f] * xs[m->own[f]];
                            }
                            x[c] = r / diag[c];
                        }
                    }
                }
                perf.iterations += nSweeps;
                oracle_amul(h, diag, upper, lower, x, wA);
                for (int c = 0; c < n; ++c) rA[c] = b[c] - wA[c];
                perf.finalResidual = sum_mag(rA, n, exact) / normFactor;
                if (hist && perf.iterations <= maxIter) hist[perf.iterations] = perf.finalResidual;
            } while (perf.iterations < ctrl.maxIter && !converged(&ctrl, &perf));
            free(xOld);
        } else {
            set_err(err, errLen, "unknown linear solver");
            rc = -1;
            goto done;
        }
    }
done:
    if (perfOut) {
        perfOut[0] = perf.initialResidual;
        perfOut[1] = perf.finalResidual;
        perfOut[2] = residualSum;
        perfOut[3] = normFactor;
    }
    if (itersOut) *itersOut = perf.iterations;
    if (convOut) *convOut = earlyDone ? 1 : converged(&ctrl, &perf);
    free(wA);
    free(rA);
    free(pA);
    free(xr);
    return rc;
}

/* ======================================================================== assembly */
/* Fvm.h:141-154 and :171-180 with fixedValue gic = −ndc_b, gbc = ndc_b·v and zeroGradient 0, 0
 * (BoundaryCondition.h:116-117,128-137). */
void oracle_laplacian(int n, int nI, int nB, const int* own, const int* nei, const double* magSf, const double* ndc,
                      const double* gamma, const int* bcFixed, const double* bcValue, double* diag, double* upper,
                      double* ic, double* bs) {
    for (int c = 0; c < n; ++c) diag[c] = 0.0;
    for (int f = 0; f < nI; ++f) {
        const double g = gamma ? gamma[f] : 1.0;
        const double c = g * magSf[f] * ndc[f];
        upper[f] = c;
        diag[own[f]] -= c;
        diag[nei[f]] -= c;
    }
    for (int bf = 0; bf < nB; ++bf) {
        const int f = nI + bf;
        const double g = gamma ? gamma[f] : 1.0;
        const double gs = g * magSf[f];
        double gic = 0.0, gbc = 0.0;
        if (bcFixed[bf]) {
            gic = -ndc[f];
            gbc = ndc[f] * bcValue[bf];
        }
        ic[bf] = 0.0 + gs * gic;
        bs[bf] = 0.0 - gs * gbc;
    }
}

/* Fvm.h:160-169 for a scalar field: gf = w g_P + (1−w) g_N; fc = (γ|S|)·(corr·gf). */
void oracle_laplacian_nonorth(int n, int nI, const int* own, const int* nei, const double* magSf, const double* w,
                              const double* corr, const double* gamma, const double* g, double* source, double* ffc) {
    (void)n;
    for (int f = 0; f < nI; ++f) {
        const int o = own[f], nb = nei[f];
        const double wf = w[f], wn = 1.0 - w[f];
        const double gx = wf * g[3 * o] + wn * g[3 * nb];
        const double gy = wf * g[3 * o + 1] + wn * g[3 * nb + 1];
        const double gz = wf * g[3 * o + 2] + wn * g[3 * nb + 2];
        const double gm = gamma ? gamma[f] : 1.0;
        const double fc = (gm * magSf[f]) * (corr[3 * f] * gx + corr[3 * f + 1] * gy + corr[3 * f + 2] * gz);
        ffc[f] = fc;
        source[o] -= fc;
        source[nb] += fc;
    }
}

/* Fvc.h:215-225: Gauss linear gradient of a scalar; bval = boundary face values. */
void oracle_grad_gauss(int n, int nI, int nB, const int* own, const int* nei, const double* w, const double* Sf,
                       const double* V, const double* psi, const double* bval, double* g) {
    for (int c = 0; c < 3 * n; ++c) g[c] = 0.0;
    for (int i = 0; i < nI; ++i) {
        const double phif = w[i] * psi[own[i]] + (1.0 - w[i]) * psi[nei[i]];
        for (int d = 0; d < 3; ++d) {
            const double c = Sf[3 * i + d] * phif;
            g[3 * own[i] + d] += c;
            g[3 * nei[i] + d] -= c;
        }
    }
    for (int i = nI; i < nI + nB; ++i)
        for (int d = 0; d < 3; ++d) g[3 * own[i] + d] += Sf[3 * i + d] * bval[i - nI];
    for (int c = 0; c < n; ++c) {
        const double r = 1.0 / V[c];
        for (int d = 0; d < 3; ++d) g[3 * c + d] = g[3 * c + d] * r;
    }
}

/* Fvm.h:88-106,126-135, upwind weights; fixedValue vic=0, vbc=v; zeroGradient vic=1, vbc=0. */
void oracle_div_upwind(int n, int nI, int nB, const int* own, const int* nei, const double* F, const int* bcFixed,
                       const double* bcValue, double* diag, double* upper, double* lower, double* ic, double* bs) {
    for (int c = 0; c < n; ++c) diag[c] = 0.0;
    for (int f = 0; f < nI; ++f) {
        const double Ff = F[f];
        const double wUp = Ff >= 0 ? 1.0 : 0.0;
        lower[f] = -wUp * Ff;
        upper[f] = lower[f] + Ff;
        diag[own[f]] -= lower[f];
        diag[nei[f]] -= upper[f];
    }
    for (int bf = 0; bf < nB; ++bf) {
        const int f = nI + bf;
        const double vic = bcFixed[bf] ? 0.0 : 1.0;
        const double vbc = bcFixed[bf] ? bcValue[bf] : 0.0;
        ic[bf] = 0.0 + F[f] * vic;
        bs[bf] = 0.0 - F[f] * vbc;
    }
}

/* Fvm.h:24-29 with Euler c = c0 = 1/dt, c00 = 0 (Schemes.h:72-85). */
void oracle_ddt_euler(int n, const double* V, double dt, const double* psi0, double* diag, double* source) {
    const double rDt = 1.0 / dt;
    for (int c = 0; c < n; ++c) {
        const double Vc = V[c];
        diag[c] = 0.0 + rDt * Vc;
        source[c] = 0.0 + Vc * (rDt * psi0[c] - 0.0 * psi0[c]);
    }
}

/* FvMatrix.h:154-163 */
void oracle_fold(int n, int nI, int nB, const int* own, const double* diag, const double* source, const double* ic,
                 const double* bs, double* d, double* b) {
    for (int c = 0; c < n; ++c) {
        d[c] = diag[c];
        b[c] = source[c];
    }
    for (int bf = 0; bf < nB; ++bf) {
        const int c = own[nI + bf];
        d[c] += ic[bf];
        b[c] += bs[bf];
    }
}

/* ---------------------------------------------------------------- pressure-velocity coupling
 * FvMatrix parts: diag and source per cell, upper/lower per internal face (lower == upper for
 * a matrix without a lower part), internalCoeffs/boundarySource per boundary face, and for
 * vector equations one call per component. */

/* cmptAv of the internalCoeff (FvMatrix.h:227-228): the value itself for a scalar equation,
 * (v + v + v)/3 for a vector equation whose boundary coefficient is the same in every
 * component (fixedValue and zeroGradient patches). */
static double cmpt_av(double v, int nCmpt) { return nCmpt == 3 ? (v + v + v) / 3.0 : v; }

/* FvMatrix::relax (FvMatrix.h:111-131): D = max(|diag + Σ_b ic|, Σ|offdiag|) per cell,
 * diag += D/α − (diag + Σ_b ic) and source += that increment · ψ. The off-diagonal sums
 * run in face order, owner rows over |upper|, neighbour rows over |lower|. */
void oracle_relax(int n, int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
                  const double* ic, const double* psi, double alpha, double* diag, double* source) {
    if (alpha >= 1.0) return;
    double* sumOff = (double*)calloc((size_t)n + 1, sizeof(double));
    double* bD = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int f = 0; f < nI; ++f) {
        sumOff[own[f]] += fabs(upper[f]);
        sumOff[nei[f]] += fabs(lower[f]);
    }
    for (int bf = 0; bf < nB; ++bf) bD[own[nI + bf]] += ic[bf];
    for (int c = 0; c < n; ++c) {
        const double full = diag[c] + bD[c];
        const double mag = fabs(full);
        const double D = mag < sumOff[c] ? sumOff[c] : mag;
        const double inc = D / alpha - full;
        diag[c] += inc;
        source[c] += inc * psi[c];
    }
    free(sumOff);
    free(bD);
}

/* FvMatrix::setReference (FvMatrix.h:134-141) with global cell numbers equal to the local
 * ones: source += diag·value, diag doubled; a cell outside [0, n) changes nothing. */
void oracle_set_reference(int n, long cell, double value, double* diag, double* source) {
    if (cell < 0 || cell >= n) return;
    source[cell] += diag[cell] * value;
    diag[cell] += diag[cell];
}

/* FvMatrix::A (FvMatrix.h:179-187): (diag + Σ_b cmptAv(ic)) / V */
void oracle_A(int n, int nI, int nB, const int* own, const double* diag, const double* ic, const double* V, int nCmpt,
              double* a) {
    for (int c = 0; c < n; ++c) a[c] = diag[c];
    for (int bf = 0; bf < nB; ++bf) a[own[nI + bf]] += cmpt_av(ic[bf], nCmpt);
    for (int c = 0; c < n; ++c) a[c] /= V[c];
}

/* FvMatrix::H (FvMatrix.h:189-208): (source − Σ offdiag·ψ_nb + Σ_b bs − Σ_b (ic − cmptAv ic)·ψ) / V,
 * the face terms in face order, then the boundary faces in order. */
void oracle_H(int n, int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
              const double* source, const double* ic, const double* bs, const double* psi, const double* V, int nCmpt,
              double* h) {
    for (int c = 0; c < n; ++c) h[c] = source[c];
    for (int f = 0; f < nI; ++f) {
        h[own[f]] -= upper[f] * psi[nei[f]];
        h[nei[f]] -= lower[f] * psi[own[f]];
    }
    for (int bf = 0; bf < nB; ++bf) {
        const int c = own[nI + bf];
        h[c] += bs[bf];
        h[c] -= (ic[bf] - cmpt_av(ic[bf], nCmpt)) * psi[c];
    }
    for (int c = 0; c < n; ++c) h[c] = h[c] / V[c];
}

/* FvMatrix::flux (FvMatrix.h:210-222): upper·ψ_N − lower·ψ_P on internal faces,
 * ic·ψ_P − bs on boundary faces, then + ffc on every face when the matrix has one
 * (ffc == NULL: none). */
void oracle_flux(int nI, int nB, const int* own, const int* nei, const double* upper, const double* lower,
                 const double* ic, const double* bs, const double* ffc, const double* psi, double* F) {
    for (int f = 0; f < nI; ++f) F[f] = upper[f] * psi[nei[f]] - lower[f] * psi[own[f]];
    for (int bf = 0; bf < nB; ++bf) F[nI + bf] = ic[bf] * psi[own[nI + bf]] - bs[bf];
    if (ffc)
        for (int f = 0; f < nI + nB; ++f) F[f] += ffc[f];
}

/* ---------------------------------------------------------------- explicit corrector operators
 * Fvc.h. Multi-component values are interleaved per cell / face, as std::vector<Vec3>
 * and std::vector<Tensor> lay them out: a Vec3 per entry for vector fields and for the
 * gradient of a scalar, a row-major Tensor (xx xy xz yx yy yz zx zy zz) for the gradient
 * of a vector. nc is 1 (scalar field) or 3 (vector field); face arrays have nI + nB
 * entries. Meshes are 3-D (no empty direction). */

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* gradDot(sk, w g_P + (1−w) g_N) (Types.h:156-157): the scalar sk·g for a Vec3 gradient,
 * the Vec3 sk·T (column j: sk_x T_xj + sk_y T_yj + sk_z T_zj) for a Tensor gradient. */
static void skew_term(int nc, const double* sk, double w, const double* gP, const double* gN, double* out) {
    double g[9];
    for (int k = 0; k < 3 * nc; ++k) g[k] = w * gP[k] + (1.0 - w) * gN[k];
    if (nc == 1) {
        out[0] = sk[0] * g[0] + sk[1] * g[1] + sk[2] * g[2];
    } else {
        for (int j = 0; j < 3; ++j) out[j] = sk[0] * g[j] + sk[1] * g[3 + j] + sk[2] * g[6 + j];
    }
}

/* fvc::interpolate(m, c) (Fvc.h:24-31): w c_P + (1−w) c_N on internal faces, the owner's
 * value on boundary faces; with grad and skew (a skewed mesh) the skew-corrected form
 * (Fvc.h:72-81) adds gradDot(skew, w g_P + (1−w) g_N) on internal faces. */
void oracle_fvc_interpolate(int n, int nI, int nB, const int* own, const int* nei, const double* w,
                            const double* skew, int nc, const double* c, const double* grad, double* F) {
    (void)n;
    for (int i = 0; i < nI; ++i)
        for (int k = 0; k < nc; ++k) F[nc * i + k] = w[i] * c[nc * own[i] + k] + (1.0 - w[i]) * c[nc * nei[i] + k];
    for (int i = nI; i < nI + nB; ++i)
        for (int k = 0; k < nc; ++k) F[nc * i + k] = c[nc * own[i] + k];
    if (grad && skew) {
        for (int i = 0; i < nI; ++i) {
            double t[3];
            skew_term(nc, skew + 3 * i, w[i], grad + 3 * nc * own[i], grad + 3 * nc * nei[i], t);
            for (int k = 0; k < nc; ++k) F[nc * i + k] += t[k];
        }
    }
}

/* Tensor inverse of Fvc.h:112-126 (cofactors over the determinant, entry by entry). */
static void inv3(const double* a, double* r) {
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

/* g += lsApply(G, v, dphi) (Fvc.h:129-130): (G·v)·dphi for a scalar, outer(G·v, dphi)
 * for a vector. */
static void ls_add(int nc, const double* G, const double* v, const double* dphi, double* g) {
    const double a[3] = {G[0] * v[0] + G[1] * v[1] + G[2] * v[2], G[3] * v[0] + G[4] * v[1] + G[5] * v[2],
                         G[6] * v[0] + G[7] * v[1] + G[8] * v[2]};
    if (nc == 1) {
        for (int k = 0; k < 3; ++k) g[k] += a[k] * dphi[0];
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) g[3 * i + j] += a[i] * dphi[j];
    }
}

/* Weighted least-squares gradient (Fvc.h:134-167) with the boundary constraints de/dphi
 * per boundary face. */
static void ls_grad(int n, int nI, int nB, const int* own, const int* nei, const double* d, int nc, const double* c,
                    const double* de, const double* dphi, double* g) {
    const int ng = 3 * nc;
    double* dd = (double*)xcalloc((size_t)n * 9 + 1, sizeof(double));
    double* Gi = (double*)xcalloc((size_t)n * 9 + 1, sizeof(double));
    for (int k = 0; k < n * ng; ++k) g[k] = 0.0;
    for (int i = 0; i < nI + nB; ++i) {
        const double* v = i < nI ? d + 3 * i : de + 3 * (i - nI);
        const double s = 1.0 / dot3(v, v);
        double t[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) t[3 * a + b] = (v[a] * v[b]) * s;
        for (int k = 0; k < 9; ++k) dd[9 * own[i] + k] += t[k];
        if (i < nI)
            for (int k = 0; k < 9; ++k) dd[9 * nei[i] + k] += t[k];
    }
    for (int cc = 0; cc < n; ++cc) inv3(dd + 9 * cc, Gi + 9 * cc);
    for (int i = 0; i < nI; ++i) {
        const double* v = d + 3 * i;
        const double wgt = 1.0 / dot3(v, v);
        const double vw[3] = {v[0] * wgt, v[1] * wgt, v[2] * wgt};
        double dp[3];
        for (int k = 0; k < nc; ++k) dp[k] = c[nc * nei[i] + k] - c[nc * own[i] + k];
        ls_add(nc, Gi + 9 * own[i], vw, dp, g + ng * own[i]);
        ls_add(nc, Gi + 9 * nei[i], vw, dp, g + ng * nei[i]);
    }
    for (int i = nI; i < nI + nB; ++i) {
        const double* v = de + 3 * (i - nI);
        const double s = 1.0 / dot3(v, v);
        const double vs[3] = {v[0] * s, v[1] * s, v[2] * s};
        ls_add(nc, Gi + 9 * own[i], vs, dphi + nc * (i - nI), g + ng * own[i]);
    }
    free(dd);
    free(Gi);
}

/* Boundary constraint of a gradient face (Fvc.h:196-200, 276-281): n·max(n·d, 0.05|d|). */
static void grad_face_de(const double* Sf, double magSf, const double* d, double* de) {
    const double s = 1.0 / magSf;
    const double nv[3] = {Sf[0] * s, Sf[1] * s, Sf[2] * s};
    const double nd0 = dot3(nv, d), lim = 0.05 * sqrt(dot3(d, d));
    const double nd = nd0 < lim ? lim : nd0;
    for (int k = 0; k < 3; ++k) de[k] = nv[k] * nd;
}

/* fvc::gradLS(m, c, bFixesValue, bValue) (Fvc.h:185-206): value faces constrain with d and
 * bValue − c_P, the others with the zero normal gradient. */
void oracle_fvc_grad_ls(int n, int nI, int nB, const int* own, const int* nei, const double* d, const double* Sf,
                        const double* magSf, int nc, const double* c, const int* bFix, const double* bVal, double* g) {
    double* de = (double*)xcalloc((size_t)nB * 3 + 1, sizeof(double));
    double* dphi = (double*)xcalloc((size_t)nB * 3 + 1, sizeof(double));
    for (int i = nI; i < nI + nB; ++i) {
        const int bf = i - nI;
        if (bFix[bf]) {
            for (int k = 0; k < 3; ++k) de[3 * bf + k] = d[3 * i + k];
            for (int k = 0; k < nc; ++k) dphi[nc * bf + k] = bVal[nc * bf + k] - c[nc * own[i] + k];
        } else {
            grad_face_de(Sf + 3 * i, magSf[i], d + 3 * i, de + 3 * bf);
        }
    }
    ls_grad(n, nI, nB, own, nei, d, nc, c, de, dphi, g);
    free(de);
    free(dphi);
}

/* fvc::grad(ψ, GaussSkew) of a scalar with fixedValue (bFix, face values bVal) and
 * zeroGradient faces (Fvc.h:223-259): face values w ψ_P + (1−w) ψ_N + skew·(w g_P +
 * (1−w) g_N) with g the least-squares gradient; on zeroGradient faces ψ_P + d_t·g_P +
 * gb (n·d) with gb = 0·ψ_P + 0; Σ S_f ψ_f (owner +, neighbour −) · (1/V). */
void oracle_fvc_grad_gauss_skew(int n, int nI, int nB, const int* own, const int* nei, const double* w,
                                const double* skew, const double* d, const double* Sf, const double* magSf,
                                const double* V, const double* psi, const int* bFix, const double* bVal, double* g) {
    double* gls = (double*)xcalloc((size_t)n * 3 + 1, sizeof(double));
    oracle_fvc_grad_ls(n, nI, nB, own, nei, d, Sf, magSf, 1, psi, bFix, bVal, gls);
    for (int k = 0; k < 3 * n; ++k) g[k] = 0.0;
    for (int i = 0; i < nI; ++i) {
        double t;
        skew_term(1, skew + 3 * i, w[i], gls + 3 * own[i], gls + 3 * nei[i], &t);
        const double phif = w[i] * psi[own[i]] + (1.0 - w[i]) * psi[nei[i]] + t;
        for (int k = 0; k < 3; ++k) {
            const double cc = Sf[3 * i + k] * phif;
            g[3 * own[i] + k] += cc;
            g[3 * nei[i] + k] -= cc;
        }
    }
    for (int i = nI; i < nI + nB; ++i) {
        const int bf = i - nI, P = own[i];
        double phif;
        if (bFix[bf]) {
            phif = bVal[bf];
        } else {
            const double* dv = d + 3 * i;
            const double s = 1.0 / magSf[i];
            const double nv[3] = {Sf[3 * i] * s, Sf[3 * i + 1] * s, Sf[3 * i + 2] * s};
            const double nd = dot3(nv, dv);
            const double dt[3] = {dv[0] - nv[0] * nd, dv[1] - nv[1] * nd, dv[2] - nv[2] * nd};
            const double gb = 0.0 * psi[P] + 0.0;
            phif = psi[P] + dot3(dt, gls + 3 * P) + gb * nd;
        }
        for (int k = 0; k < 3; ++k) g[3 * P + k] += Sf[3 * i + k] * phif;
    }
    for (int c = 0; c < n; ++c) {
        const double r = 1.0 / V[c];
        for (int k = 0; k < 3; ++k) g[3 * c + k] = g[3 * c + k] * r;
    }
    free(gls);
}

/* fvc::div(m, F) (Fvc.h:294-305): owner +F, neighbour −F, boundary +F, then / V. */
void oracle_fvc_div(int n, int nI, int nB, const int* own, const int* nei, const double* V, const double* F,
                    double* r) {
    for (int c = 0; c < n; ++c) r[c] = 0.0;
    for (int i = 0; i < nI; ++i) {
        r[own[i]] += F[i];
        r[nei[i]] -= F[i];
    }
    for (int i = nI; i < nI + nB; ++i) r[own[i]] += F[i];
    for (int c = 0; c < n; ++c) r[c] /= V[c];
}

/* fvc::surfaceIntegrate(m, F) (Fvc.h:307-318), nc components per face: the same sums,
 * then · (1/V). */
void oracle_fvc_surface_integrate(int n, int nI, int nB, const int* own, const int* nei, const double* V, int nc,
                                  const double* F, double* r) {
    for (int k = 0; k < n * nc; ++k) r[k] = 0.0;
    for (int i = 0; i < nI; ++i)
        for (int k = 0; k < nc; ++k) {
            r[nc * own[i] + k] += F[nc * i + k];
            r[nc * nei[i] + k] -= F[nc * i + k];
        }
    for (int i = nI; i < nI + nB; ++i)
        for (int k = 0; k < nc; ++k) r[nc * own[i] + k] += F[nc * i + k];
    for (int c = 0; c < n; ++c) {
        const double s = 1.0 / V[c];
        for (int k = 0; k < nc; ++k) r[nc * c + k] = r[nc * c + k] * s;
    }
}

/* Face flux of a cell vector field as phiHbyA is formed (IncompressibleFlow.cpp:276-291):
 * internal faces dot(U_f, S_f) with U_f the (skew-corrected when grad) interpolate,
 * boundary faces dot(Ub, S_f) where bFix, else dot(U_P, S_f). */
void oracle_fvc_flux(int n, int nI, int nB, const int* own, const int* nei, const double* w, const double* skew,
                     const double* Sf, const double* U, const double* grad, const int* bFix, const double* Ub,
                     double* phi) {
    double* Uf = (double*)xcalloc((size_t)(nI + nB) * 3 + 1, sizeof(double));
    oracle_fvc_interpolate(n, nI, nB, own, nei, w, skew, 3, U, grad, Uf);
    for (int i = 0; i < nI; ++i) phi[i] = dot3(Uf + 3 * i, Sf + 3 * i);
    for (int i = nI; i < nI + nB; ++i)
        phi[i] = bFix[i - nI] ? dot3(Ub + 3 * (i - nI), Sf + 3 * i) : dot3(U + 3 * own[i], Sf + 3 * i);
    free(Uf);
}

/* ---------------------------------------------------------------- pressure-corrector steps */

/* rAU = 1.0 / A; HbyA = rAU * H (IncompressibleFlow.cpp:245-249) */
void oracle_piso_rau_hbya(int n, const double* A, const double* H, double* rAU, double* HbyA) {
    for (int c = 0; c < n; ++c) {
        rAU[c] = 1.0 / A[c];
        for (int k = 0; k < 3; ++k) HbyA[3 * c + k] = rAU[c] * H[3 * c + k];
    }
}

/* bVal = U_b + rAU_o * gradp_o on fixed-flux faces, else 0 (IncompressibleFlow.cpp:261-270) */
void oracle_piso_hbya_boundary(int nI, int nB, const int* own, const int* bFix, const double* Ub, const double* rAU,
                               const double* gp, double* bVal) {
    for (int i = 0; i < nB; ++i) {
        const int o = own[nI + i];
        for (int k = 0; k < 3; ++k) bVal[3 * i + k] = bFix[i] ? Ub[3 * i + k] + rAU[o] * gp[3 * o + k] : 0.0;
    }
}

/* op 0: a + b*c (:294-295); 1: a - b (:219-220, :338-339); 2: a - b/c (:351-352) */
void oracle_piso_face_op(int nF, int op, const double* a, const double* b, const double* c, double* out) {
    for (int f = 0; f < nF; ++f) out[f] = op == 0 ? a[f] + b[f] * c[f] : op == 1 ? a[f] - b[f] : a[f] - b[f] / c[f];
}

/* ddtPhiCorr (IncompressibleFlow.cpp:296-316) */
void oracle_piso_ddt_phi_corr(int nI, const int* own, const int* nei, const double* w, const double* Sf, double coeff,
                              double c0, double c00, const double* U0, const double* U00, const double* phiOld,
                              const double* phiOldOld, const double* rAUf, double* phiHbyA) {
    for (int f = 0; f < nI; ++f) {
        const double wf = w[f];
        const int o = own[f], n = nei[f];
        double v[3];
        for (int k = 0; k < 3; ++k) v[k] = wf * U0[3 * o + k] + (1.0 - wf) * U0[3 * n + k];
        const double Uf0 = dot3(v, Sf + 3 * f);
        const double corr0 = phiOld[f] - Uf0;
        double coupling = coeff;
        if (coeff < 0) {
            const double q = fabs(corr0) / (fabs(phiOld[f]) + 1.0e-15);
            coupling = 1.0 - (1.0 < q ? 1.0 : q);
        }
        double dc = c0 * corr0;
        if (c00 != 0.0) {
            for (int k = 0; k < 3; ++k) v[k] = wf * U00[3 * o + k] + (1.0 - wf) * U00[3 * n + k];
            dc -= c00 * (phiOldOld[f] - dot3(v, Sf + 3 * f));
        }
        phiHbyA[f] += rAUf[f] * coupling * dc;
    }
}

/* constrainPressure (IncompressibleFlow.cpp:17-29) on boundary faces [b0, b0 + cnt) */
void oracle_piso_constrain_pressure(int nI, int b0, int cnt, const double* Sf, const double* magSf,
                                    const double* phiHbyA, const double* Ub, const double* rAUf, double* g) {
    for (int i = 0; i < cnt; ++i) {
        const int f = nI + b0 + i;
        g[i] = (phiHbyA[f] - dot3(Ub + 3 * (b0 + i), Sf + 3 * f)) / (magSf[f] * rAUf[f]);
    }
}

/* fvc::reconstruct (Fvc.h:33-67): face loop scattering S S^T/|S| and (phi/|S|) S into
 * owner and neighbour, 2-D diagonal fix, adjugate inverse */
static void rec_acc(double* A, double* b, const double* S, double ms, double phi) {
    const double r = 1.0 / ms; /* Tensor / scalar multiplies by the reciprocal (Types.h:86) */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[3 * i + j] += S[i] * S[j] * r;
    const double s = phi / ms;
    for (int i = 0; i < 3; ++i) b[i] += s * S[i];
}
void oracle_fvc_reconstruct(int n, int nI, int nB, const int* own, const int* nei, const double* Sf,
                            const double* magSf, const double* ssf, int emptyDir, double* v) {
    double* A = (double*)xcalloc((size_t)n * 9 + 1, sizeof(double));
    double* b = (double*)xcalloc((size_t)n * 3 + 1, sizeof(double));
    for (int f = 0; f < nI; ++f) {
        rec_acc(A + 9 * own[f], b + 3 * own[f], Sf + 3 * f, magSf[f], ssf[f]);
        rec_acc(A + 9 * nei[f], b + 3 * nei[f], Sf + 3 * f, magSf[f], ssf[f]);
    }
    for (int f = nI; f < nI + nB; ++f) rec_acc(A + 9 * own[f], b + 3 * own[f], Sf + 3 * f, magSf[f], ssf[f]);
    for (int c = 0; c < n; ++c) {
        double* t = A + 9 * c;
        const double* r = b + 3 * c;
        v[3 * c] = v[3 * c + 1] = v[3 * c + 2] = 0.0;
        if (emptyDir >= 0) t[4 * emptyDir] += 1.0;
        const double c00 = t[4] * t[8] - t[5] * t[7], c01 = t[5] * t[6] - t[3] * t[8], c02 = t[3] * t[7] - t[4] * t[6];
        const double det = t[0] * c00 + t[1] * c01 + t[2] * c02;
        if (fabs(det) < 1e-300) continue;
        const double c10 = t[2] * t[7] - t[1] * t[8], c11 = t[0] * t[8] - t[2] * t[6], c12 = t[1] * t[6] - t[0] * t[7];
        const double c20 = t[1] * t[5] - t[2] * t[4], c21 = t[2] * t[3] - t[0] * t[5], c22 = t[0] * t[4] - t[1] * t[3];
        v[3 * c] = (c00 * r[0] + c10 * r[1] + c20 * r[2]) / det;
        v[3 * c + 1] = (c01 * r[0] + c11 * r[1] + c21 * r[2]) / det;
        v[3 * c + 2] = (c02 * r[0] + c12 * r[1] + c22 * r[2]) / det;
    }
    free(A);
    free(b);
}

/* U = HbyA - rAU*v (sign < 0) or HbyA + rAU*v; the empty direction zeroed (:347-360) */
void oracle_piso_correct_U(int n, const double* HbyA, const double* rAU, const double* v, int sign, int emptyDir,
                           double* U) {
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < 3; ++k) {
            const double t = rAU[c] * v[3 * c + k];
            U[3 * c + k] = k == emptyDir ? 0.0 : sign < 0 ? HbyA[3 * c + k] - t : HbyA[3 * c + k] + t;
        }
}

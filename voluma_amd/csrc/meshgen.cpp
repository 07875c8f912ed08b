// Mesh synthesis and finite-volume geometry for the solver's inputs (host C++20).
//
// The linear-solve path consumes LDU addressing (owner/neighbour) and, for GAMG,
// the per-face agglomeration weight |S_f|·nonOrthDeltaCoeff_f. This module makes
// those inputs for the benchmark meshes:
//   * the box mesh with generateBox's point/face order (reference
//     src/fvm/mesh/RawMesh.cpp:391-502): internal faces owner-major +x,+y,+z, then
//     patches xMin,xMax,yMin,yMax,zMin,zMax;
//   * a jittered tetrahedral mesh: an nx×ny×nz hex lattice, each hex split into the 6
//     Kuhn tetrahedra around its main diagonal, interior vertices displaced uniformly
//     by ±jitter·spacing per axis from a splitmix64 stream; internal faces in
//     upper-triangular (owner, neighbour) order, boundary triangles grouped in the
//     same six patches.
// Geometry follows the reference's face/cell decomposition with the same operation
// order (src/fvm/mesh/Mesh.cpp:21-48 faces, :135-154 cells, :600-628 face weights),
// compiled without FP contraction, so weights are bit-identical to the reference's.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <array>
#include <filesystem>
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
};
inline V3 operator+(V3 a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, const V3& a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double mag(const V3& a) { return std::sqrt(dot(a, a)); }

constexpr double VSMALL = 1.0e-300;

struct Raw {
    std::vector<V3> points;
    std::vector<int64_t> faceOff{0}, facePts, owner, neighbour;
    std::vector<int64_t> pStart, pSize;
    std::vector<std::string> pName, pType;  // patch names / types
    std::vector<std::string> pNeighbour;     // cyclic patches: the partner patch's name
    int64_t nCells = 0;
    // geometry (filled by computeGeometry)
    bool haveGeom = false;
    std::vector<double> V, magSf, w, ndc;
    std::vector<V3> C, Sf, Cf, d, corr;
    std::vector<V3> skew;  // skew vectors of the internal faces (zero on boundary faces)
    double skewMax = 0.0;
    void addFace(std::initializer_list<int64_t> p) {
        facePts.insert(facePts.end(), p.begin(), p.end());
        faceOff.push_back(int64_t(facePts.size()));
    }
    int64_t nFaces() const { return int64_t(owner.size()); }
    int64_t nInternal() const { return int64_t(neighbour.size()); }
};

// generateBox (RawMesh.cpp:391-502) on the unit cube: periodic bit d (1 x, 2 y, 4 z)
// makes the two patches of that direction a cyclic pair; twoD (nz = 1) makes zMin and
// zMax one empty patch "frontAndBack".
Raw makeBox(int nx, int ny, int nz, int periodic = 0, bool twoD = false) {
    for (int d = 0; d < 3; ++d) {
        const int nd = d == 0 ? nx : d == 1 ? ny : nz;
        if ((periodic >> d & 1) && nd < 2 && !(d == 2 && twoD))
            throw std::runtime_error("box: a periodic direction needs at least 2 cells");
    }
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("box: bad resolution");
    if (twoD && nz != 1) throw std::invalid_argument("box: twoD requires nz == 1");
    Raw m;
    auto P = [&](int i, int j, int k) -> int64_t { return i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k); };
    auto C = [&](int i, int j, int k) -> int64_t { return i + int64_t(nx) * (j + int64_t(ny) * k); };
    // unit box: coordinate = 0 + (1 − 0)·(i/n), the reference's lo + (hi − lo)·t
    auto coord = [](int i, int n) { return 0.0 + (1.0 - 0.0) * (double(i) / n); };
    m.points.resize(std::size_t(nx + 1) * (ny + 1) * (nz + 1));
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) m.points[P(i, j, k)] = V3{coord(i, nx), coord(j, ny), coord(k, nz)};
    m.nCells = int64_t(nx) * ny * nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t c = C(i, j, k);
                if (i + 1 < nx) {
                    m.addFace({P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1)});
                    m.owner.push_back(c);
                    m.neighbour.push_back(C(i + 1, j, k));
                }
                if (j + 1 < ny) {
                    m.addFace({P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k)});
                    m.owner.push_back(c);
                    m.neighbour.push_back(C(i, j + 1, k));
                }
                if (k + 1 < nz) {
                    m.addFace({P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1)});
                    m.owner.push_back(c);
                    m.neighbour.push_back(C(i, j, k + 1));
                }
            }
    static const char* names[6] = {"xMin", "xMax", "yMin", "yMax", "zMin", "zMax"};
    auto beginP = [&](const std::string& name, const std::string& type, const std::string& nb) {
        m.pName.push_back(name);
        m.pType.push_back(type);
        m.pNeighbour.push_back(nb);
        m.pStart.push_back(m.nFaces());
    };
    auto patchOf = [&](int q, int d) {  // patch q of direction d: cyclic with its partner, or a plain patch
        const bool cyc = periodic >> d & 1;
        beginP(names[q], cyc ? "cyclic" : "patch", cyc ? names[q ^ 1] : "");
    };
    int q0 = 0;
    auto begin = [&]() { patchOf(q0, q0 / 2); };
    auto end = [&]() {
        m.pSize.push_back(m.nFaces() - m.pStart.back());
        ++q0;
    };
    for (int side = 0; side < 2; ++side) {  // xMin, xMax
        begin();
        const int i = side == 0 ? 0 : nx;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                if (side == 0)
                    m.addFace({P(i, j, k), P(i, j, k + 1), P(i, j + 1, k + 1), P(i, j + 1, k)});
                else
                    m.addFace({P(i, j, k), P(i, j + 1, k), P(i, j + 1, k + 1), P(i, j, k + 1)});
                m.owner.push_back(C(side == 0 ? 0 : nx - 1, j, k));
            }
        end();
    }
    for (int side = 0; side < 2; ++side) {  // yMin, yMax
        begin();
        const int j = side == 0 ? 0 : ny;
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                if (side == 0)
                    m.addFace({P(i, j, k), P(i + 1, j, k), P(i + 1, j, k + 1), P(i, j, k + 1)});
                else
                    m.addFace({P(i, j, k), P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j, k)});
                m.owner.push_back(C(i, side == 0 ? 0 : ny - 1, k));
            }
        end();
    }
    if (twoD) beginP("frontAndBack", "empty", "");
    for (int side = 0; side < 2; ++side) {  // zMin, zMax (one empty patch in 2-D)
        if (!twoD) begin();
        const int k = side == 0 ? 0 : nz;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (side == 0)
                    m.addFace({P(i, j, k), P(i, j + 1, k), P(i + 1, j + 1, k), P(i + 1, j, k)});
                else
                    m.addFace({P(i, j, k), P(i + 1, j, k), P(i + 1, j + 1, k), P(i, j + 1, k)});
                m.owner.push_back(C(i, j, side == 0 ? 0 : nz - 1));
            }
        if (!twoD) end();
    }
    if (twoD) m.pSize.push_back(m.nFaces() - m.pStart.back());
    return m;
}

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

// Kuhn split of an nx×ny×nz lattice on the unit box into 6 tets per hex.
Raw makeTet(int nx, int ny, int nz, double jitter, uint64_t seed) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("tet: bad resolution");
    if (!(jitter >= 0.0 && jitter < 0.25)) throw std::invalid_argument("tet: jitter must be in [0, 0.25)");
    Raw m;
    const int64_t NX = nx + 1, NY = ny + 1;
    auto P = [&](int64_t i, int64_t j, int64_t k) -> int64_t { return i + NX * (j + NY * k); };
    m.points.resize(std::size_t(NX) * NY * (nz + 1));
    SplitMix64 rng{seed};
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                V3 p{double(i) / nx, double(j) / ny, double(k) / nz};
                if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz && jitter > 0) {
                    p.x += (2.0 * rng.uniform() - 1.0) * jitter * hx;
                    p.y += (2.0 * rng.uniform() - 1.0) * jitter * hy;
                    p.z += (2.0 * rng.uniform() - 1.0) * jitter * hz;
                }
                m.points[P(i, j, k)] = p;
            }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int64_t nHex = int64_t(nx) * ny * nz;
    m.nCells = 6 * nHex;
    // tet vertices as lattice indices (exact integers for orientation tests)
    struct TetV {
        int64_t v[4];
        int32_t L[4][3];
    };
    auto tetOf = [&](int64_t cell) {
        const int64_t hex = cell / 6, t = cell % 6;
        const int i = int(hex % nx), j = int((hex / nx) % ny), k = int(hex / (int64_t(nx) * ny));
        TetV T;
        int c[3] = {i, j, k};
        for (int s = 0; s < 4; ++s) {
            if (s > 0) ++c[perms[t][s - 1]];
            T.L[s][0] = c[0];
            T.L[s][1] = c[1];
            T.L[s][2] = c[2];
            T.v[s] = P(c[0], c[1], c[2]);
        }
        return T;
    };
    // every tet face: key = sorted vertex triple packed in 63 bits
    const int64_t nPts = int64_t(m.points.size());
    int bits = 1;
    while ((int64_t(1) << bits) < nPts) ++bits;
    if (3 * bits > 63) throw std::invalid_argument("tet: mesh too large for face keys");
    struct FI {
        uint64_t key;
        int64_t cell;
        int32_t opp;  // local index of the vertex opposite the face
    };
    std::vector<FI> fi;
    fi.reserve(std::size_t(m.nCells) * 4);
    for (int64_t c = 0; c < m.nCells; ++c) {
        const TetV T = tetOf(c);
        for (int o = 0; o < 4; ++o) {
            int64_t a[3], q = 0;
            for (int s = 0; s < 4; ++s)
                if (s != o) a[q++] = T.v[s];
            std::sort(a, a + 3);
            fi.push_back({(uint64_t(a[0]) << (2 * bits)) | (uint64_t(a[1]) << bits) | uint64_t(a[2]), c, o});
        }
    }
    std::sort(fi.begin(), fi.end(), [](const FI& x, const FI& y) {
        return x.key != y.key ? x.key < y.key : x.cell < y.cell;
    });
    // oriented triangle for the face of `cell` opposite local vertex `opp`: normal points away from opp
    auto orient = [&](int64_t cell, int opp, int64_t out[3]) {
        const TetV T = tetOf(cell);
        int idx[3], q = 0;
        for (int s = 0; s < 4; ++s)
            if (s != opp) idx[q++] = s;
        auto vec = [&](int a, int b, int64_t r[3]) {
            for (int d = 0; d < 3; ++d) r[d] = int64_t(T.L[b][d]) - T.L[a][d];
        };
        int64_t e1[3], e2[3], e3[3];
        vec(idx[0], idx[1], e1);
        vec(idx[0], idx[2], e2);
        vec(opp, idx[0], e3);
        const int64_t nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                e1[0] * e2[1] - e1[1] * e2[0]};
        const int64_t s = nrm[0] * e3[0] + nrm[1] * e3[1] + nrm[2] * e3[2];
        if (s == 0) throw std::logic_error("tet: degenerate face");
        out[0] = T.v[idx[0]];
        out[1] = s > 0 ? T.v[idx[1]] : T.v[idx[2]];
        out[2] = s > 0 ? T.v[idx[2]] : T.v[idx[1]];
    };
    struct IF {
        int64_t own, nei;
        int32_t opp;
    };
    std::vector<IF> internal;
    struct BF {
        int patch;
        int64_t own;
        uint64_t key;
        int32_t opp;
    };
    std::vector<BF> boundary;
    internal.reserve(fi.size() / 2);
    for (std::size_t a = 0; a < fi.size();) {
        if (a + 1 < fi.size() && fi[a + 1].key == fi[a].key) {
            internal.push_back({fi[a].cell, fi[a + 1].cell, fi[a].opp});  // cells ascending: own < nei
            a += 2;
        } else {
            // boundary plane from lattice coordinates of its three vertices
            const TetV T = tetOf(fi[a].cell);
            int patch = -1;
            const int lim[3] = {nx, ny, nz};
            for (int d = 0; d < 3 && patch < 0; ++d)
                for (int side = 0; side < 2 && patch < 0; ++side) {
                    const int val = side == 0 ? 0 : lim[d];
                    bool all = true;
                    for (int s = 0; s < 4; ++s)
                        if (s != fi[a].opp && T.L[s][d] != val) all = false;
                    if (all) patch = 2 * d + side;
                }
            if (patch < 0) throw std::logic_error("tet: unmatched interior face");
            boundary.push_back({patch, fi[a].cell, fi[a].key, fi[a].opp});
            a += 1;
        }
    }
    std::sort(internal.begin(), internal.end(), [](const IF& x, const IF& y) {
        return x.own != y.own ? x.own < y.own : x.nei < y.nei;
    });
    std::sort(boundary.begin(), boundary.end(), [](const BF& x, const BF& y) {
        if (x.patch != y.patch) return x.patch < y.patch;
        if (x.own != y.own) return x.own < y.own;
        return x.key < y.key;
    });
    for (const auto& f : internal) {
        int64_t t[3];
        orient(f.own, f.opp, t);
        m.addFace({t[0], t[1], t[2]});
        m.owner.push_back(f.own);
        m.neighbour.push_back(f.nei);
    }
    for (int p = 0; p < 6; ++p) {
        m.pStart.push_back(m.nFaces());
        for (const auto& f : boundary) {
            if (f.patch != p) continue;
            int64_t t[3];
            orient(f.own, f.opp, t);
            m.addFace({t[0], t[1], t[2]});
            m.owner.push_back(f.own);
        }
        m.pSize.push_back(m.nFaces() - m.pStart.back());
    }
    return m;
}

// Face and cell geometry in the reference's operation order.
void computeGeometry(Raw& m) {
    const int64_t nF = m.nFaces(), nI = m.nInternal(), nC = m.nCells;
    m.Sf.assign(nF, V3{});
    m.Cf.assign(nF, V3{});
    for (int64_t f = 0; f < nF; ++f) {
        const int64_t b = m.faceOff[f], e = m.faceOff[f + 1], n = e - b;
        const auto& P = m.points;
        const int64_t* fp = m.facePts.data() + b;
        if (n == 3) {
            m.Cf[f] = (P[fp[0]] + P[fp[1]] + P[fp[2]]) / 3.0;
            m.Sf[f] = 0.5 * cross(P[fp[1]] - P[fp[0]], P[fp[2]] - P[fp[0]]);
            continue;
        }
        V3 pAvg;
        for (int64_t k = 0; k < n; ++k) pAvg += P[fp[k]];
        pAvg = pAvg / double(n);
        V3 sumN, sumAc;
        double sumA = 0;
        for (int64_t k = 0; k < n; ++k) {
            const V3& p0 = P[fp[k]];
            const V3& p1 = P[fp[(k + 1) % n]];
            const V3 c = p0 + p1 + pAvg;
            const V3 nk = cross(p1 - p0, pAvg - p0);
            const double a = mag(nk);
            sumN += nk;
            sumA += a;
            sumAc += a * c;
        }
        m.Cf[f] = sumA < VSMALL ? pAvg : sumAc / (3.0 * sumA);
        m.Sf[f] = 0.5 * sumN;
    }
    // cell → face CSR over all faces in ascending face order
    std::vector<int64_t> off(nC + 1, 0), list;
    for (int64_t f = 0; f < nF; ++f) {
        ++off[m.owner[f] + 1];
        if (f < nI) ++off[m.neighbour[f] + 1];
    }
    for (int64_t c = 0; c < nC; ++c) off[c + 1] += off[c];
    list.resize(off[nC]);
    {
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int64_t f = 0; f < nF; ++f) {
            list[fill[m.owner[f]]++] = f;
            if (f < nI) list[fill[m.neighbour[f]]++] = f;
        }
    }
    m.C.assign(nC, V3{});
    m.V.assign(nC, 0.0);
    for (int64_t c = 0; c < nC; ++c) {
        V3 cEst;
        const int64_t nfc = off[c + 1] - off[c];
        for (int64_t k = off[c]; k < off[c + 1]; ++k) cEst += m.Cf[list[k]];
        cEst = cEst / double(nfc);
        V3 cSum;
        double vSum = 0;
        for (int64_t k = off[c]; k < off[c + 1]; ++k) {
            const int64_t f = list[k];
            const V3 S = (m.owner[f] == c) ? m.Sf[f] : V3{-m.Sf[f].x, -m.Sf[f].y, -m.Sf[f].z};
            const double pv = std::max(dot(S, m.Cf[f] - cEst), VSMALL);
            cSum += pv * (0.75 * m.Cf[f] + 0.25 * cEst);
            vSum += pv;
        }
        m.C[c] = cSum / vSum;
        m.V[c] = vSum / 3.0;
    }
    m.d.assign(nF, V3{});
    for (int64_t f = 0; f < nF; ++f)
        m.d[f] = f < nI ? m.C[m.neighbour[f]] - m.C[m.owner[f]] : m.Cf[f] - m.C[m.owner[f]];
    m.magSf.assign(nF, 0.0);
    m.w.assign(nF, 0.0);
    m.ndc.assign(nF, 0.0);
    m.corr.assign(nF, V3{});
    m.skew.assign(nF, V3{});
    m.skewMax = 0.0;
    for (int64_t f = 0; f < nF; ++f) {
        m.magSf[f] = mag(m.Sf[f]);
        const V3 n = m.Sf[f] / m.magSf[f];
        const V3& dd = m.d[f];
        if (f < nI) {
            const V3& CP = m.C[m.owner[f]];
            const double sOwn = std::abs(dot(m.Sf[f], m.Cf[f] - CP));
            const double sNei = std::abs(dot(m.Sf[f], CP + dd - m.Cf[f]));
            m.w[f] = sNei / (sOwn + sNei);
            // face centre minus the linear-interpolation point (Mesh.cpp:618-621)
            m.skew[f] = m.Cf[f] - (CP + dd * (1.0 - m.w[f]));
            m.skewMax = std::max(m.skewMax, mag(m.skew[f]) / mag(dd));
        } else {
            m.w[f] = 1.0;
        }
        m.ndc[f] = 1.0 / std::max(dot(n, dd), 0.05 * mag(dd));
        m.corr[f] = n - dd * m.ndc[f];
    }
    m.haveGeom = true;
}

// ------------------------------------------------------------------ OpenFOAM polyMesh reader
// One polyMesh file (RawMesh.cpp:223-336 reads the same files): the FoamFile header
// (format ascii|binary, class, arch "LSB;label=NN;scalar=NN"), then lists "N ( ... )"
// of labels, vectors "(x y z)", faces "k(p0 … pk-1)", or in binary form N raw little-
// endian elements between the parentheses. C and C++ comments are skipped in ASCII.
struct FoamFile {
    std::string path, text, cls;
    std::size_t pos = 0;
    bool binary = false;
    int labelBytes = 4, scalarBytes = 8;

    explicit FoamFile(const std::string& p) : path(p) {
        std::ifstream in(p, std::ios::binary);
        if (!in) throw std::runtime_error(p + ": cannot open");
        std::ostringstream ss;
        ss << in.rdbuf();
        text = ss.str();
        const std::size_t h = text.find("FoamFile");
        if (h == std::string::npos) throw std::runtime_error(p + ": no FoamFile header");
        const std::size_t open = text.find('{', h), close = text.find('}', h);
        if (open == std::string::npos || close == std::string::npos) throw std::runtime_error(p + ": bad header");
        const std::string hdr = text.substr(open + 1, close - open - 1);
        binary = value(hdr, "format") == "binary";
        cls = value(hdr, "class");
        const std::string arch = value(hdr, "arch");
        if (arch.find("MSB") != std::string::npos) throw std::runtime_error(p + ": big-endian binary not supported");
        const std::size_t l = arch.find("label="), sc = arch.find("scalar=");
        if (l != std::string::npos) labelBytes = std::atoi(arch.c_str() + l + 6) / 8;
        if (sc != std::string::npos) scalarBytes = std::atoi(arch.c_str() + sc + 7) / 8;
        if ((labelBytes != 4 && labelBytes != 8) || (scalarBytes != 4 && scalarBytes != 8))
            throw std::runtime_error(p + ": unsupported label/scalar size");
        pos = close + 1;
    }
    static std::string value(const std::string& hdr, const std::string& key) {
        std::size_t k = 0;
        while ((k = hdr.find(key, k)) != std::string::npos) {
            const bool startOk = k == 0 || std::isspace(static_cast<unsigned char>(hdr[k - 1]));
            const std::size_t e = k + key.size();
            if (startOk && e < hdr.size() && std::isspace(static_cast<unsigned char>(hdr[e]))) {
                const std::size_t semi = hdr.find(';', e);
                std::string v = hdr.substr(e, semi - e);
                const auto b = v.find_first_not_of(" \t\r\n\""), en = v.find_last_not_of(" \t\r\n\"");
                return b == std::string::npos ? std::string() : v.substr(b, en - b + 1);
            }
            k = e;
        }
        return {};
    }
    void skip() {
        while (pos < text.size()) {
            const char c = text[pos];
            if (std::isspace(static_cast<unsigned char>(c))) {
                ++pos;
            } else if (text.compare(pos, 2, "//") == 0) {
                pos = text.find('\n', pos);
                if (pos == std::string::npos) pos = text.size();
            } else if (text.compare(pos, 2, "/*") == 0) {
                pos = text.find("*/", pos);
                pos = pos == std::string::npos ? text.size() : pos + 2;
            } else {
                break;
            }
        }
    }
    [[noreturn]] void bad(const std::string& what) const {
        throw std::runtime_error(path + ": " + what + " at byte " + std::to_string(pos));
    }
    void expect(char c) {
        skip();
        if (pos >= text.size() || text[pos] != c) bad(std::string("expected '") + c + "'");
        ++pos;
    }
    int64_t integer() {
        skip();
        char* e = nullptr;
        const long long v = std::strtoll(text.c_str() + pos, &e, 10);
        if (e == text.c_str() + pos) bad("expected an integer");
        pos = std::size_t(e - text.c_str());
        return v;
    }
    double real() {
        skip();
        char* e = nullptr;
        const double v = std::strtod(text.c_str() + pos, &e);
        if (e == text.c_str() + pos) bad("expected a number");
        pos = std::size_t(e - text.c_str());
        return v;
    }
    // list header "N (" (binary: the raw elements follow directly)
    int64_t listStart() {
        const int64_t n = integer();
        if (n < 0) bad("negative list size");
        expect('(');
        return n;
    }
    const char* rawBlock(int64_t n, int elemBytes) {
        const std::size_t bytes = std::size_t(n) * std::size_t(elemBytes);
        if (pos + bytes >= text.size() || text[pos + bytes] != ')') bad("binary list size mismatch");
        const char* d = text.data() + pos;
        pos += bytes + 1;
        return d;
    }
    std::vector<int64_t> labels() {
        const int64_t n = listStart();
        std::vector<int64_t> v(static_cast<std::size_t>(n));
        if (binary) {
            const char* d = rawBlock(n, labelBytes);
            for (int64_t i = 0; i < n; ++i) {
                if (labelBytes == 4) {
                    int32_t x;
                    std::memcpy(&x, d + 4 * i, 4);
                    v[i] = x;
                } else {
                    int64_t x;
                    std::memcpy(&x, d + 8 * i, 8);
                    v[i] = x;
                }
            }
            return v;
        }
        for (auto& x : v) x = integer();
        expect(')');
        return v;
    }
    std::vector<V3> vectors() {
        const int64_t n = listStart();
        std::vector<V3> v(static_cast<std::size_t>(n));
        if (binary) {
            const char* d = rawBlock(n, 3 * scalarBytes);
            for (int64_t i = 0; i < n; ++i) {
                double c[3];
                for (int k = 0; k < 3; ++k) {
                    if (scalarBytes == 8) {
                        std::memcpy(&c[k], d + 24 * i + 8 * k, 8);
                    } else {
                        float f;
                        std::memcpy(&f, d + 12 * i + 4 * k, 4);
                        c[k] = f;
                    }
                }
                v[i] = V3{c[0], c[1], c[2]};
            }
            return v;
        }
        for (auto& p : v) {
            expect('(');
            p.x = real();
            p.y = real();
            p.z = real();
            expect(')');
        }
        expect(')');
        return v;
    }
    // next word of the boundary dictionary: a name, a value, or one of { } ;
    std::string word() {
        skip();
        if (pos >= text.size()) bad("unexpected end of file");
        const char c = text[pos];
        if (c == '{' || c == '}' || c == ';' || c == '(' || c == ')') {
            ++pos;
            return std::string(1, c);
        }
        const std::size_t b = pos;
        while (pos < text.size() && !std::isspace(static_cast<unsigned char>(text[pos])) &&
               std::string("{};()").find(text[pos]) == std::string::npos)
            ++pos;
        std::string w = text.substr(b, pos - b);
        if (w.size() >= 2 && w.front() == '"' && w.back() == '"') w = w.substr(1, w.size() - 2);
        return w;
    }
};

Raw readPolyMesh(const std::string& dir) {
    {
        std::ifstream probe(dir + "/points");
        std::ifstream gz(dir + "/points.gz");
        if (!probe && gz) throw std::runtime_error(dir + ": compressed polyMesh (*.gz) is not supported");
    }
    Raw m;
    m.faceOff.clear();
    {
        FoamFile f(dir + "/points");
        m.points = f.vectors();
    }
    {
        FoamFile f(dir + "/faces");
        if (f.cls == "faceCompactList") {
            m.faceOff = f.labels();
            m.facePts = f.labels();
        } else {
            if (f.binary) throw std::runtime_error(dir + "/faces: binary faces must be faceCompactList");
            const int64_t n = f.listStart();
            m.faceOff.reserve(std::size_t(n) + 1);
            m.faceOff.push_back(0);
            for (int64_t i = 0; i < n; ++i) {
                const int64_t np = f.integer();
                f.expect('(');
                for (int64_t k = 0; k < np; ++k) m.facePts.push_back(f.integer());
                f.expect(')');
                m.faceOff.push_back(int64_t(m.facePts.size()));
            }
            f.expect(')');
        }
    }
    {
        FoamFile f(dir + "/owner");
        m.owner = f.labels();
    }
    {
        FoamFile f(dir + "/neighbour");
        m.neighbour = f.labels();
    }
    {
        FoamFile f(dir + "/boundary");
        const int64_t n = f.listStart();
        for (int64_t p = 0; p < n; ++p) {
            std::string name = f.word(), type, nbPatch;
            int64_t start = 0, size = 0;
            if (f.word() != "{") f.bad("expected '{' after patch name");
            int depth = 1;
            while (depth > 0) {
                const std::string key = f.word();
                if (key == "}") {
                    --depth;
                    continue;
                }
                if (key == "{") {
                    ++depth;
                    continue;
                }
                if (key == ";") continue;
                std::vector<std::string> vals;
                for (std::string v = f.word(); v != ";"; v = f.word()) {
                    if (v == "{") {  // a nested dictionary as the value: skipped
                        int d = 1;
                        while (d > 0) {
                            const std::string w = f.word();
                            if (w == "{") ++d;
                            if (w == "}") --d;
                        }
                        break;
                    }
                    vals.push_back(v);
                }
                if (vals.empty()) continue;
                if (key == "type") type = vals[0];
                else if (key == "nFaces") size = std::strtoll(vals[0].c_str(), nullptr, 10);
                else if (key == "startFace") start = std::strtoll(vals[0].c_str(), nullptr, 10);
                else if (key == "neighbourPatch") nbPatch = vals[0];
            }
            // the type in the reference's spelling (patchTypeFromString, then toString)
            type = type == "wall" ? "wall"
                   : (type == "symmetry" || type == "symmetryPlane") ? "symmetryPlane"
                   : type == "empty" ? "empty"
                   : type == "cyclic" ? "cyclic"
                                      : "patch";
            m.pName.push_back(name);
            m.pType.push_back(type);
            m.pNeighbour.push_back(nbPatch);
            m.pStart.push_back(start);
            m.pSize.push_back(size);
        }
    }
    int64_t nc = 0;
    for (int64_t o : m.owner) nc = std::max(nc, o + 1);
    for (int64_t o : m.neighbour) nc = std::max(nc, o + 1);
    m.nCells = nc;
    if (m.faceOff.size() != m.owner.size() + 1) throw std::runtime_error(dir + ": faces and owner sizes differ");
    if (m.faceOff.empty() || m.faceOff.front() != 0 || m.faceOff.back() != int64_t(m.facePts.size()))
        throw std::runtime_error(dir + ": inconsistent face offsets");
    for (int64_t k : m.facePts)
        if (k < 0 || k >= int64_t(m.points.size())) throw std::runtime_error(dir + ": face point out of range");
    return m;
}

// ------------------------------------------------------------------ Gmsh reader (3-D)
// ASCII Gmsh 2.2 and 4.x files as the reference's readGmsh reads them
// (MeshReaders.cpp:192-482): nodes, elements of the highest dimension become cells,
// their faces are matched by sorted corner nodes (first element = c0), boundary faces
// take the physical group of the 2-D element with the same nodes (else patch
// "defaultFaces"), unused points are dropped, and the face list is assembled by
// assembleFaces below. 2-D meshes (extruded by the reference) are not supported here.
// ---------------------------------------------------------------- mesh import
// Shared by the Gmsh and Fluent readers (MeshReaders.cpp): face keys, face-list assembly,
// 2-D meshes extruded to one layer, import options.
namespace meshio {

// Element type → (dimension, corner nodes); higher-order types keep their corners
// (the element type numbers are Gmsh's).
bool elementType(int t, int& dim, int& corners) {
    static const std::map<int, std::pair<int, int>> table = {
        {15, {0, 1}}, {1, {1, 2}},   {8, {1, 2}},   {26, {1, 2}},  {27, {1, 2}},  {28, {1, 2}},  {2, {2, 3}},
        {9, {2, 3}},  {20, {2, 3}},  {21, {2, 3}},  {22, {2, 3}},  {23, {2, 3}},  {24, {2, 3}},  {25, {2, 3}},
        {3, {2, 4}},  {10, {2, 4}},  {16, {2, 4}},  {36, {2, 4}},  {37, {2, 4}},  {38, {2, 4}},  {4, {3, 4}},
        {11, {3, 4}}, {29, {3, 4}},  {30, {3, 4}},  {31, {3, 4}},  {5, {3, 8}},   {12, {3, 8}},  {17, {3, 8}},
        {92, {3, 8}}, {93, {3, 8}},  {6, {3, 6}},   {13, {3, 6}},  {18, {3, 6}},  {90, {3, 6}},  {91, {3, 6}},
        {7, {3, 5}},  {14, {3, 5}},  {19, {3, 5}},  {118, {3, 5}}, {119, {3, 5}}};
    const auto it = table.find(t);
    if (it == table.end()) return false;
    dim = it->second.first;
    corners = it->second.second;
    return true;
}

// Faces of the 3-D element kinds by corner index in Gmsh's node order (their orientation
// is corrected from the geometry afterwards).
const std::vector<std::vector<int>>& elementFaces(int corners) {
    static const std::vector<std::vector<int>> tet{{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
    static const std::vector<std::vector<int>> pyramid{{0, 3, 2, 1}, {0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4}};
    static const std::vector<std::vector<int>> prism{
        {0, 2, 1}, {3, 4, 5}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}};
    static const std::vector<std::vector<int>> hex{{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                                   {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    switch (corners) {
    case 4: return tet;
    case 5: return pyramid;
    case 6: return prism;
    default: return hex;
    }
}

struct Key {
    std::array<int64_t, 4> v{-1, -1, -1, -1};
    bool operator==(const Key& o) const { return v == o.v; }
};
struct KeyHash {
    std::size_t operator()(const Key& k) const {
        std::size_t h = 0x9e3779b97f4a7c15ull;
        for (int64_t x : k.v) h = (h ^ std::size_t(x + 1)) * 0x100000001b3ull;
        return h;
    }
};
Key keyOf(const std::vector<int64_t>& n) {
    if (n.size() > 4) throw std::runtime_error("mesh import: element face with more than 4 nodes");
    Key k;
    std::copy(n.begin(), n.end(), k.v.begin());
    std::sort(k.v.begin(), k.v.begin() + n.size());
    return k;
}

struct Element {
    int dim, phys;
    std::vector<int64_t> nodes;
};
struct Face {
    std::vector<int64_t> nodes;
    int64_t c0, c1;
    int patch;
};
struct Patch {
    std::string name, type;
};

// Point average and half the summed fan cross products about it (the polygon's
// centre estimate and area vector used to orient faces).
V3 polygonNormal(const std::vector<V3>& P, const std::vector<int64_t>& f, V3& centre) {
    V3 c;
    for (int64_t i : f) c += P[i];
    c = c / double(f.size());
    V3 n;
    for (std::size_t k = 0; k < f.size(); ++k) n += cross(P[f[k]] - c, P[f[(k + 1) % f.size()]] - c);
    centre = c;
    return 0.5 * n;
}

// Face list → Raw (MeshReaders.cpp:131-190): cell centres estimated as the mean of their
// face centres; internal faces get c0 < c1 and a normal from c0 to c1, boundary faces an
// outward normal (nodes reversed otherwise); internal faces sorted by (c0, c1), then
// boundary faces patch by patch in input order; patches without faces are dropped.
Raw assembleFaces(std::vector<V3> points, int64_t nCells, std::vector<Face> faces, const std::vector<Patch>& patches) {
    std::vector<V3> cc(static_cast<std::size_t>(nCells), V3{});
    std::vector<int> nf(std::size_t(nCells), 0);
    std::vector<V3> fc(faces.size()), fn(faces.size());
    for (std::size_t f = 0; f < faces.size(); ++f) {
        Face& F = faces[f];
        if (F.c0 < 0 && F.c1 >= 0) std::swap(F.c0, F.c1);
        if (F.c0 < 0 || F.c0 >= nCells || F.c1 >= nCells) throw std::runtime_error("mesh import: face with invalid cells");
        fn[f] = polygonNormal(points, F.nodes, fc[f]);
        cc[F.c0] += fc[f];
        ++nf[F.c0];
        if (F.c1 >= 0) {
            cc[F.c1] += fc[f];
            ++nf[F.c1];
        }
    }
    for (int64_t c = 0; c < nCells; ++c) {
        if (nf[c] < 4) throw std::runtime_error("mesh import: cell " + std::to_string(c) + " has fewer than 4 faces");
        cc[c] = cc[c] / double(nf[c]);
    }
    std::vector<std::size_t> internal;
    std::vector<std::vector<std::size_t>> bnd(patches.size());
    for (std::size_t f = 0; f < faces.size(); ++f) {
        Face& F = faces[f];
        if (F.c1 >= 0) {
            if (F.c1 < F.c0) std::swap(F.c0, F.c1);
            if (dot(fn[f], cc[F.c1] - cc[F.c0]) < 0) std::reverse(F.nodes.begin(), F.nodes.end());
            internal.push_back(f);
        } else {
            if (F.patch < 0 || F.patch >= int(patches.size()))
                throw std::runtime_error("mesh import: boundary face without patch");
            if (dot(fn[f], fc[f] - cc[F.c0]) < 0) std::reverse(F.nodes.begin(), F.nodes.end());
            bnd[F.patch].push_back(f);
        }
    }
    std::sort(internal.begin(), internal.end(), [&](std::size_t a, std::size_t b) {
        return faces[a].c0 != faces[b].c0 ? faces[a].c0 < faces[b].c0 : faces[a].c1 < faces[b].c1;
    });
    Raw r;
    r.points = std::move(points);
    r.nCells = nCells;
    auto add = [&](const Face& F) {
        r.facePts.insert(r.facePts.end(), F.nodes.begin(), F.nodes.end());
        r.faceOff.push_back(int64_t(r.facePts.size()));
        r.owner.push_back(F.c0);
    };
    for (std::size_t f : internal) {
        add(faces[f]);
        r.neighbour.push_back(faces[f].c1);
    }
    for (std::size_t p = 0; p < patches.size(); ++p) {
        if (bnd[p].empty()) continue;
        const int64_t start = r.nFaces();
        for (std::size_t f : bnd[p]) add(faces[f]);
        r.pName.push_back(patches[p].name);
        r.pType.push_back(patches[p].type);
        r.pStart.push_back(start);
        r.pSize.push_back(r.nFaces() - start);
    }
    return r;
}

std::string lower(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// MeshImportOptions (MeshReaders.h:20-24): coordinate scale, 2-D extrusion depth and
// patch types by name that replace the reader's choice.
struct Options {
    double scale = 1.0;
    double depth2D = 0.1;
    std::map<std::string, std::string> patchTypes;
};
// A patch type string in the reference's spelling (patchTypeFromString then toString,
// RawMesh.cpp:14-30): wall, symmetryPlane (also from "symmetry"), empty, cyclic, patch.
std::string typeName(const std::string& s) {
    if (s == "wall") return "wall";
    if (s == "symmetry" || s == "symmetryPlane") return "symmetryPlane";
    if (s == "empty") return "empty";
    if (s == "cyclic") return "cyclic";
    return "patch";
}
// The type a patch gets: the override for its name if there is one (MeshReaders.cpp:57-61).
std::string patchType(const std::string& name, const std::string& t, const Options& o) {
    const auto it = o.patchTypes.find(name);
    return it != o.patchTypes.end() ? it->second : t;
}

// A planar polygon mesh: points (x, y used) and each cell's points in order.
struct Mesh2D {
    std::vector<V3> points;
    std::vector<std::vector<int64_t>> cells;
};
// Twice the signed area of a cell, summed as the reference sums it (x_k·y_k+1 − x_k+1·y_k).
double twiceArea(const Mesh2D& m, const std::vector<int64_t>& c) {
    double a = 0;
    for (std::size_t k = 0; k < c.size(); ++k) {
        const V3& p = m.points[c[k]];
        const V3& q = m.points[c[(k + 1) % c.size()]];
        a += p.x * q.y - q.x * p.y;
    }
    return a;
}

// One layer of cells from a counter-clockwise 2-D mesh (Extrude2D.cpp:59-166): points
// z = 0 then z = depth; internal faces from the edges shared by two cells (owner the
// lower cell, its edge direction), sorted by (owner, neighbour); then each patch's
// boundary edges sorted by cell; then "frontAndBack" (empty): per cell its reversed
// polygon at z = 0 and its polygon at z = depth. classify(a, b) gives the patch of a
// boundary edge. The edges are visited in (min, max) point order and sorted with
// std::sort, as the reference does, so equal keys keep the reference's order.
template <class Classify>
Raw extrude2D(const Mesh2D& m, const std::vector<Patch>& patches, Classify classify, double depth) {
    const int64_t nP = int64_t(m.points.size()), nC = int64_t(m.cells.size());
    for (int64_t c = 0; c < nC; ++c) {
        if (m.cells[c].size() < 3) throw std::runtime_error("extrude2D: degenerate cell");
        if (twiceArea(m, m.cells[c]) <= 0)
            throw std::runtime_error("extrude2D: cell " + std::to_string(c) + " is not counter-clockwise");
    }
    struct Use {
        int64_t cell, a, b;
    };
    struct Internal {
        int64_t own, nei, a, b;
    };
    std::map<std::pair<int64_t, int64_t>, std::vector<Use>> edges;
    for (int64_t c = 0; c < nC; ++c) {
        const auto& pl = m.cells[c];
        for (std::size_t k = 0; k < pl.size(); ++k) {
            const int64_t a = pl[k], b = pl[(k + 1) % pl.size()];
            edges[{std::min(a, b), std::max(a, b)}].push_back({c, a, b});
        }
    }
    std::vector<Internal> internal;
    std::vector<std::vector<Use>> bnd(patches.size());
    for (const auto& kv : edges) {
        const auto& u = kv.second;
        if (u.size() == 2) {
            const Use& u0 = u[0].cell < u[1].cell ? u[0] : u[1];
            const Use& u1 = u[0].cell < u[1].cell ? u[1] : u[0];
            internal.push_back({u0.cell, u1.cell, u0.a, u0.b});
        } else if (u.size() == 1) {
            const int p = classify(u[0].a, u[0].b);
            if (p < 0 || p >= int(patches.size())) throw std::runtime_error("extrude2D: unclassified boundary edge");
            bnd[p].push_back(u[0]);
        } else {
            throw std::runtime_error("extrude2D: non-manifold edge");
        }
    }
    std::sort(internal.begin(), internal.end(), [](const Internal& x, const Internal& y) {
        return x.own != y.own ? x.own < y.own : x.nei < y.nei;
    });
    for (auto& list : bnd) std::sort(list.begin(), list.end(), [](const Use& x, const Use& y) { return x.cell < y.cell; });
    Raw r;
    r.points.resize(std::size_t(2 * nP));
    for (int64_t i = 0; i < nP; ++i) {
        r.points[i] = {m.points[i].x, m.points[i].y, 0.0};
        r.points[nP + i] = {m.points[i].x, m.points[i].y, depth};
    }
    r.nCells = nC;
    for (const Internal& f : internal) {
        r.addFace({f.a, f.b, nP + f.b, nP + f.a});
        r.owner.push_back(f.own);
        r.neighbour.push_back(f.nei);
    }
    auto patch = [&](const std::string& name, const std::string& type) {
        r.pName.push_back(name);
        r.pType.push_back(type);
        r.pStart.push_back(r.nFaces());
    };
    for (std::size_t p = 0; p < patches.size(); ++p) {
        patch(patches[p].name, patches[p].type);
        for (const Use& e : bnd[p]) {
            r.addFace({e.a, e.b, nP + e.b, nP + e.a});
            r.owner.push_back(e.cell);
        }
        r.pSize.push_back(r.nFaces() - r.pStart.back());
    }
    patch("frontAndBack", "empty");
    for (int64_t c = 0; c < nC; ++c) {
        const auto& pl = m.cells[c];
        for (auto it = pl.rbegin(); it != pl.rend(); ++it) r.facePts.push_back(*it);
        r.faceOff.push_back(int64_t(r.facePts.size()));
        r.owner.push_back(c);
        for (int64_t q : pl) r.facePts.push_back(nP + q);
        r.faceOff.push_back(int64_t(r.facePts.size()));
        r.owner.push_back(c);
    }
    r.pSize.push_back(r.nFaces() - r.pStart.back());
    return r;
}

// An imported 2-D mesh (MeshReaders.cpp:64-96): clockwise cells reversed, the patches
// plus "defaultFaces" (wall) for boundary edges in no group, extruded by depth2D;
// patches without faces dropped.
Raw extrudeImported(Mesh2D m, const std::vector<Patch>& patches,
                    const std::map<std::pair<int64_t, int64_t>, int>& edgePatch, const Options& o) {
    for (auto& c : m.cells)
        if (twiceArea(m, c) < 0) std::reverse(c.begin(), c.end());
    std::vector<Patch> specs;
    for (const Patch& p : patches) specs.push_back({p.name, patchType(p.name, p.type, o)});
    const int defaultIdx = int(specs.size());
    specs.push_back({"defaultFaces", patchType("defaultFaces", "wall", o)});
    Raw r = extrude2D(
        m, specs,
        [&](int64_t a, int64_t b) {
            const auto it = edgePatch.find({std::min(a, b), std::max(a, b)});
            return it != edgePatch.end() ? it->second : defaultIdx;
        },
        o.depth2D);
    Raw k = r;
    k.pName.clear();
    k.pType.clear();
    k.pStart.clear();
    k.pSize.clear();
    for (std::size_t p = 0; p < r.pSize.size(); ++p)
        if (r.pSize[p] > 0) {
            k.pName.push_back(r.pName[p]);
            k.pType.push_back(r.pType[p]);
            k.pStart.push_back(r.pStart[p]);
            k.pSize.push_back(r.pSize[p]);
        }
    return k;
}

}  // namespace meshio

// ---------------------------------------------------------------- Gmsh
namespace gmsh {
using namespace meshio;

Raw read(const std::string& file, const Options& opt) {
    const double scale = opt.scale;
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    double version = 0;
    std::map<std::pair<int, int>, std::string> physNames;
    std::map<std::pair<int, int>, std::vector<int>> entityPhys;
    std::unordered_map<int64_t, int64_t> nodeIndex;
    std::vector<V3> points;
    std::vector<Element> elems;
    std::string line;
    auto endSection = [&](const std::string& name) {
        while (std::getline(in, line))
            if (line.rfind("$End" + name, 0) == 0) return;
        throw std::runtime_error(file + ": missing $End" + name);
    };
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == "$MeshFormat") {
            int fileType = 0, dataSize = 0;
            in >> version >> fileType >> dataSize;
            if (fileType != 0) throw std::runtime_error(file + ": binary Gmsh files are not supported");
            if (version < 2 || version >= 5) throw std::runtime_error(file + ": unsupported Gmsh version");
            endSection("MeshFormat");
        } else if (line == "$PhysicalNames") {
            int n;
            in >> n;
            for (int i = 0; i < n; ++i) {
                int d, t;
                in >> d >> t;
                std::getline(in, line);
                const auto a = line.find('"'), b = line.rfind('"');
                physNames[{d, t}] = (a != std::string::npos && b > a) ? line.substr(a + 1, b - a - 1) : std::to_string(t);
            }
            endSection("PhysicalNames");
        } else if (line == "$Entities" && version >= 4) {
            std::size_t n[4];
            in >> n[0] >> n[1] >> n[2] >> n[3];
            for (int d = 0; d < 4; ++d)
                for (std::size_t i = 0; i < n[d]; ++i) {
                    int tag;
                    double x;
                    in >> tag;
                    for (int k = 0; k < (d == 0 ? 3 : 6); ++k) in >> x;
                    std::size_t np;
                    in >> np;
                    std::vector<int> ph(np);
                    for (int& q : ph) in >> q;
                    for (int& q : ph) q = std::abs(q);
                    entityPhys[{d, tag}] = ph;
                    if (d > 0) {
                        std::size_t nb;
                        in >> nb;
                        for (std::size_t k = 0; k < nb; ++k) in >> tag;
                    }
                }
            endSection("Entities");
        } else if (line == "$Nodes") {
            if (version >= 4) {
                std::size_t nBlocks, nNodes, minTag, maxTag;
                in >> nBlocks >> nNodes >> minTag >> maxTag;
                points.reserve(nNodes);
                for (std::size_t b = 0; b < nBlocks; ++b) {
                    int d, tag, param;
                    std::size_t n;
                    in >> d >> tag >> param >> n;
                    std::vector<int64_t> tags(n);
                    for (int64_t& t : tags) in >> t;
                    for (std::size_t i = 0; i < n; ++i) {
                        V3 p;
                        in >> p.x >> p.y >> p.z;
                        double u;
                        for (int k = 0; k < (param ? d : 0); ++k) in >> u;
                        nodeIndex[tags[i]] = int64_t(points.size());
                        points.push_back(p * scale);
                    }
                }
            } else {
                std::size_t n;
                in >> n;
                points.reserve(n);
                for (std::size_t i = 0; i < n; ++i) {
                    int64_t t;
                    V3 p;
                    in >> t >> p.x >> p.y >> p.z;
                    nodeIndex[t] = int64_t(points.size());
                    points.push_back(p * scale);
                }
            }
            endSection("Nodes");
        } else if (line == "$Elements") {
            if (version >= 4) {
                std::size_t nBlocks, nEl, minTag, maxTag;
                in >> nBlocks >> nEl >> minTag >> maxTag;
                elems.reserve(nEl);
                std::getline(in, line);
                for (std::size_t b = 0; b < nBlocks; ++b) {
                    int d, tag, type;
                    std::size_t n;
                    in >> d >> tag >> type >> n;
                    std::getline(in, line);
                    int dim, corners;
                    if (!elementType(type, dim, corners))
                        throw std::runtime_error(file + ": unsupported Gmsh element type " + std::to_string(type));
                    const auto it = entityPhys.find({d, tag});
                    const int phys = (it != entityPhys.end() && !it->second.empty()) ? it->second[0] : -1;
                    for (std::size_t i = 0; i < n; ++i) {
                        std::getline(in, line);
                        std::istringstream ls(line);
                        int64_t id, v;
                        ls >> id;
                        Element e{dim, phys, {}};
                        for (int k = 0; k < corners && ls >> v; ++k) e.nodes.push_back(v);
                        elems.push_back(std::move(e));
                    }
                }
            } else {
                std::size_t n;
                in >> n;
                std::getline(in, line);
                elems.reserve(n);
                for (std::size_t i = 0; i < n; ++i) {
                    std::getline(in, line);
                    std::istringstream ls(line);
                    int64_t id;
                    int type, nt;
                    ls >> id >> type >> nt;
                    std::vector<int> tags(std::max(nt, 0));
                    for (int& t : tags) ls >> t;
                    int dim, corners;
                    if (!elementType(type, dim, corners))
                        throw std::runtime_error(file + ": unsupported Gmsh element type " + std::to_string(type));
                    Element e{dim, nt > 0 ? tags[0] : -1, {}};
                    if (e.phys == 0) e.phys = -1;
                    int64_t v;
                    for (int k = 0; k < corners && ls >> v; ++k) e.nodes.push_back(v);
                    elems.push_back(std::move(e));
                }
            }
            endSection("Elements");
        }
    }
    if (points.empty() || elems.empty()) throw std::runtime_error(file + ": no nodes or elements");
    auto node = [&](int64_t t) {
        const auto it = nodeIndex.find(t);
        if (it == nodeIndex.end()) throw std::runtime_error(file + ": element refers to unknown node " + std::to_string(t));
        return it->second;
    };
    int meshDim = 0;
    for (const Element& e : elems) meshDim = std::max(meshDim, e.dim);
    if (meshDim < 2) throw std::runtime_error(file + ": no 2D or 3D elements");
    // patches: the physical groups of dimension meshDim − 1
    std::map<int, int> physToPatch;
    std::vector<Patch> patches;
    for (const Element& e : elems)
        if (e.dim == meshDim - 1 && e.phys >= 0 && !physToPatch.count(e.phys)) {
            Patch p;
            const auto it = physNames.find({meshDim - 1, e.phys});
            p.name = it != physNames.end() ? it->second : "patch" + std::to_string(e.phys);
            const std::string l = lower(p.name);
            p.type = l.find("symmetry") != std::string::npos ? "symmetryPlane"
                     : l.find("wall") != std::string::npos   ? "wall"
                                                             : "patch";
            physToPatch[e.phys] = int(patches.size());
            patches.push_back(p);
        }
    for (std::size_t a = 0; a < patches.size(); ++a)
        for (std::size_t b = a + 1; b < patches.size(); ++b)
            if (patches[a].name == patches[b].name)
                throw std::runtime_error(file + ": duplicate physical name " + patches[a].name);
    if (meshDim == 2) {  // MeshReaders.cpp:406-426: every node kept, cells from 2-D elements
        Mesh2D m2;
        m2.points = points;
        const double z0 = points[0].z;
        for (const V3& p : points)
            if (std::abs(p.z - z0) > 1e-9 * (1 + std::abs(z0)))
                throw std::runtime_error(file + ": 2D Gmsh mesh must lie in a plane z = const");
        std::map<std::pair<int64_t, int64_t>, int> edgePatch;
        for (const Element& e : elems) {
            if (e.dim == 2) {
                std::vector<int64_t> c;
                for (int64_t t : e.nodes) c.push_back(node(t));
                m2.cells.push_back(c);
            } else if (e.dim == 1 && e.phys >= 0) {
                const int64_t a = node(e.nodes[0]), b = node(e.nodes[1]);
                edgePatch[{std::min(a, b), std::max(a, b)}] = physToPatch[e.phys];
            }
        }
        for (Patch& p : patches) p.type = patchType(p.name, p.type, opt);
        return extrudeImported(std::move(m2), patches, edgePatch, opt);
    }
    std::unordered_map<Key, int, KeyHash> bFacePatch;
    for (const Element& e : elems)
        if (e.dim == 2 && e.phys >= 0) {
            std::vector<int64_t> f;
            for (int64_t t : e.nodes) f.push_back(node(t));
            bFacePatch[keyOf(f)] = physToPatch[e.phys];
        }
    const int defaultIdx = int(patches.size());
    patches.push_back({"defaultFaces", "wall"});
    for (Patch& p : patches) p.type = patchType(p.name, p.type, opt);
    std::vector<Face> faces;
    std::unordered_map<Key, std::size_t, KeyHash> faceOf;
    int64_t nCells = 0;
    for (const Element& e : elems) {
        if (e.dim != 3) continue;
        const int64_t c = nCells++;
        std::vector<int64_t> cn;
        for (int64_t t : e.nodes) cn.push_back(node(t));
        for (const auto& fd : elementFaces(int(cn.size()))) {
            std::vector<int64_t> f;
            for (int k : fd) f.push_back(cn[k]);
            const Key k = keyOf(f);
            const auto it = faceOf.find(k);
            if (it == faceOf.end()) {
                faceOf.emplace(k, faces.size());
                faces.push_back({f, c, -1, -1});
            } else {
                Face& F = faces[it->second];
                if (F.c1 >= 0) throw std::runtime_error(file + ": face shared by more than two cells");
                F.c1 = c;
            }
        }
    }
    for (Face& F : faces)
        if (F.c1 < 0) {
            const auto it = bFacePatch.find(keyOf(F.nodes));
            F.patch = it != bFacePatch.end() ? it->second : defaultIdx;
        }
    std::vector<int64_t> used(points.size(), -1);
    std::vector<V3> pts;
    for (Face& F : faces)
        for (int64_t& v : F.nodes) {
            if (used[v] < 0) {
                used[v] = int64_t(pts.size());
                pts.push_back(points[v]);
            }
            v = used[v];
        }
    return assembleFaces(std::move(pts), nCells, std::move(faces), patches);
}

}  // namespace gmsh

// ---------------------------------------------------------------- Fluent
// Fluent mesh files (MeshReaders.cpp:484-803): parenthesised sections "(index ...)" with
// hexadecimal headers; index 10 nodes, 12 cells, 13 faces, 39/45 zone names and types,
// 2 dimension; index + 2000 (single precision, 32-bit integers) or + 3000 (double,
// 64-bit) marks a binary body ending in "End of Binary Section". Other sections are
// skipped; periodic pairs (18) import as ordinary patches; non-conformal sections are
// rejected.
namespace fluent {
using namespace meshio;

struct Zone {
    int id, bc;
    std::vector<Face> faces;  // nodes and cells 0-based, −1 for no cell
};

class Reader {
public:
    explicit Reader(const std::string& file) : file_(file) {
        std::ifstream in(file, std::ios::binary);
        if (!in) throw std::runtime_error("cannot open " + file);
        std::ostringstream ss;
        ss << in.rdbuf();
        s_ = ss.str();
    }
    int dim = 3;
    std::vector<V3> points;
    int64_t nCells = 0;
    std::vector<Zone> zones;
    std::map<int, std::pair<std::string, std::string>> zoneInfo;  // id → (type, name)

    void parse() {
        for (;;) {
            ws();
            if (p_ >= s_.size()) return;
            if (s_[p_++] != '(') continue;
            section(int(integer(10)));
        }
    }

private:
    std::string file_, s_;
    std::size_t p_ = 0;

    [[noreturn]] void fail(const std::string& m) const {
        throw std::runtime_error(file_ + ": " + m + " (at byte " + std::to_string(p_) + ")");
    }
    void ws() {
        while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
    }
    bool at(char c) {
        ws();
        return p_ < s_.size() && s_[p_] == c;
    }
    void expect(char c) {
        if (!at(c)) fail(std::string("expected '") + c + "'");
        ++p_;
    }
    long long integer(int base) {
        ws();
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        const long long v = std::strtoll(b, &e, base);
        if (e == b) fail("expected integer");
        p_ += std::size_t(e - b);
        return v;
    }
    double real() {
        ws();
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) fail("expected number");
        p_ += std::size_t(e - b);
        return v;
    }
    std::string word() {
        ws();
        const std::size_t b = p_;
        if (p_ < s_.size() && s_[p_] == '"') {
            const std::size_t e = s_.find('"', p_ + 1);
            p_ = e + 1;
            return s_.substr(b + 1, e - b - 1);
        }
        while (p_ < s_.size() && !std::isspace(static_cast<unsigned char>(s_[p_])) && s_[p_] != '(' && s_[p_] != ')')
            ++p_;
        return s_.substr(b, p_ - b);
    }
    // past the ')' that closes the current level (quoted strings skipped whole)
    void closeLevel() {
        for (int depth = 1; p_ < s_.size() && depth > 0;) {
            const char c = s_[p_++];
            if (c == '"') {
                const std::size_t e = s_.find('"', p_);
                p_ = e == std::string::npos ? s_.size() : e + 1;
            } else if (c == '(') {
                ++depth;
            } else if (c == ')') {
                --depth;
            }
        }
    }
    void closeBinary() {
        const std::size_t e = s_.find("End of Binary Section", p_);
        if (e == std::string::npos) fail("missing 'End of Binary Section'");
        p_ = s_.find(')', e);
        if (p_ == std::string::npos) fail("unterminated binary section");
        ++p_;
    }
    template <class T> T raw() {
        if (p_ + sizeof(T) > s_.size()) fail("unexpected end of binary data");
        T v;
        std::memcpy(&v, s_.data() + p_, sizeof(T));
        p_ += sizeof(T);
        return v;
    }
    std::vector<long long> header() {
        expect('(');
        std::vector<long long> h;
        while (!at(')')) h.push_back(integer(16));
        ++p_;
        return h;
    }
    void closeBody(int bin) {
        if (bin) {
            closeBinary();
        } else {
            expect(')');
            expect(')');
        }
    }

    void section(int idx) {
        const int kind = idx % 1000, bin = idx / 1000;
        switch (kind) {
        case 0:
        case 1: closeLevel(); return;
        case 2:
            dim = int(integer(10));
            expect(')');
            return;
        case 10: nodes(bin); return;
        case 12: cells(bin); return;
        case 13: faces(bin); return;
        case 39:
        case 45: zone(); return;
        case 58:
        case 59:
        case 61:
        case 62:
        case 63:
            fail("non-conformal / hanging-node Fluent meshes (section " + std::to_string(idx) + ") are not supported");
        default: break;  // 18 (periodic pairs) and anything else: skipped
        }
        if (bin) closeBinary();
        else closeLevel();
    }
    void nodes(int bin) {
        const auto h = header();
        if (h.size() < 4) fail("bad node header");
        if (h[0] == 0) {
            points.resize(std::size_t(h[2]));
            expect(')');
            return;
        }
        const int64_t first = h[1] - 1, last = h[2] - 1;
        const int nd = h.size() >= 5 ? int(h[4]) : dim;
        if (int64_t(points.size()) <= last) points.resize(std::size_t(last) + 1);
        expect('(');
        for (int64_t i = first; i <= last; ++i) {
            double x[3] = {0, 0, 0};
            for (int k = 0; k < nd; ++k) x[k] = bin == 0 ? real() : bin == 2 ? double(raw<float>()) : raw<double>();
            points[i] = {x[0], x[1], x[2]};
        }
        closeBody(bin);
    }
    void cells(int bin) {
        const auto h = header();
        if (h.size() < 3) fail("bad cell header");
        nCells = std::max<int64_t>(nCells, h[2]);
        if (h[0] != 0 && at('(')) {  // per-cell element types: not needed
            ++p_;
            if (bin) {
                closeBinary();
                return;
            }
            closeLevel();
        }
        expect(')');
    }
    void faces(int bin) {
        const auto h = header();
        if (h.size() < 5) {
            if (h.size() >= 3 && h[0] == 0) {
                expect(')');
                return;
            }
            fail("bad face header");
        }
        if (h[0] == 0) {
            expect(')');
            return;
        }
        Zone z{int(h[0]), int(h[3]), {}};
        const long long n = h[2] - h[1] + 1;
        const int faceType = int(h[4]);
        z.faces.reserve(std::size_t(std::max(n, 0LL)));
        expect('(');
        auto next = [&]() -> long long {
            if (bin == 0) return integer(16);
            return bin == 3 ? raw<int64_t>() : raw<int32_t>();
        };
        for (long long i = 0; i < n; ++i) {
            long long nn = faceType;
            if (faceType == 0 || faceType == 5) {  // mixed: a type per face; polygon: a count
                nn = next();
                if (faceType == 0 && nn == 5) nn = next();
            }
            Face F{{}, -1, -1, -1};
            for (long long k = 0; k < nn; ++k) F.nodes.push_back(next() - 1);
            F.c0 = next() - 1;
            F.c1 = next() - 1;
            z.faces.push_back(std::move(F));
        }
        closeBody(bin);
        zones.push_back(std::move(z));
    }
    void zone() {
        expect('(');
        const int id = int(integer(10));
        const std::string type = word();
        const std::string name = word();
        closeLevel();
        zoneInfo[id] = {type, name};
        if (at('(')) {
            ++p_;
            closeLevel();
        }
        expect(')');
    }
};

// Patch type of a boundary zone from its type word or, without one, its bc number.
std::string zoneType(const std::string& type, int bc) {
    const std::string t = lower(type);
    if (t == "wall" || (t.empty() && bc == 3)) return "wall";
    if (t == "symmetry" || t == "axis" || (t.empty() && (bc == 7 || bc == 37))) return "symmetryPlane";
    return "patch";
}

// 2-D cell polygons from the edges of each cell, chained from its first edge in file
// order (MeshReaders.cpp:98-126).
std::vector<std::vector<int64_t>> polygons(int64_t nCells, const std::vector<std::array<int64_t, 3>>& cellEdges) {
    std::vector<std::vector<std::pair<int64_t, int64_t>>> e(static_cast<std::size_t>(nCells));
    for (const auto& ce : cellEdges) e[ce[0]].push_back({ce[1], ce[2]});
    std::vector<std::vector<int64_t>> cells(static_cast<std::size_t>(nCells));
    for (int64_t c = 0; c < nCells; ++c) {
        const auto& E = e[c];
        const std::string cs = "mesh import: 2D cell " + std::to_string(c);
        if (E.size() < 3) throw std::runtime_error(cs + " has < 3 edges");
        std::vector<char> used(E.size(), 0);
        std::vector<int64_t> poly{E[0].first, E[0].second};
        used[0] = 1;
        for (std::size_t k = 1; k < E.size(); ++k) {
            const int64_t last = poly.back();
            std::size_t j = 0;
            for (; j < E.size(); ++j)
                if (!used[j] && (E[j].first == last || E[j].second == last)) break;
            if (j == E.size()) throw std::runtime_error(cs + " edges do not form a loop");
            used[j] = 1;
            poly.push_back(E[j].first == last ? E[j].second : E[j].first);
        }
        if (poly.back() != poly.front()) throw std::runtime_error(cs + " is not closed");
        poly.pop_back();
        cells[c] = std::move(poly);
    }
    return cells;
}

Raw read(const std::string& file, const Options& opt) {
    Reader fr(file);
    fr.parse();
    if (fr.points.empty() || fr.zones.empty()) throw std::runtime_error(file + ": no nodes or faces found");
    for (V3& p : fr.points) p = p * opt.scale;
    int64_t nc = fr.nCells;
    for (const Zone& z : fr.zones)
        for (const Face& F : z.faces) nc = std::max({nc, F.c0 + 1, F.c1 + 1});
    // every zone with a one-sided face becomes a patch, in zone order
    std::vector<Patch> patches;
    std::map<int, int> zoneToPatch;
    for (const Zone& z : fr.zones) {
        bool boundary = false;
        for (const Face& F : z.faces) boundary = boundary || F.c0 < 0 || F.c1 < 0;
        if (!boundary) continue;
        const auto it = fr.zoneInfo.find(z.id);
        Patch p;
        p.name = it != fr.zoneInfo.end() ? it->second.second : "zone" + std::to_string(z.id);
        p.type = patchType(p.name, zoneType(it != fr.zoneInfo.end() ? it->second.first : "", z.bc), opt);
        zoneToPatch[z.id] = int(patches.size());
        patches.push_back(p);
    }
    if (fr.dim == 2) {
        Mesh2D m2;
        m2.points = fr.points;
        std::vector<std::array<int64_t, 3>> cellEdges;
        std::map<std::pair<int64_t, int64_t>, int> edgePatch;
        for (const Zone& z : fr.zones)
            for (const Face& F : z.faces) {
                if (F.nodes.size() != 2)
                    throw std::runtime_error(file + ": 2D face with " + std::to_string(F.nodes.size()) + " nodes");
                const int64_t a = F.nodes[0], b = F.nodes[1];
                if (F.c0 >= 0) cellEdges.push_back({F.c0, a, b});
                if (F.c1 >= 0) cellEdges.push_back({F.c1, a, b});
                if (F.c0 < 0 || F.c1 < 0) edgePatch[{std::min(a, b), std::max(a, b)}] = zoneToPatch.at(z.id);
            }
        m2.cells = polygons(nc, cellEdges);
        return extrudeImported(std::move(m2), patches, edgePatch, opt);
    }
    std::vector<Face> faces;
    for (Zone& z : fr.zones)
        for (Face& F : z.faces) {
            if (F.c0 < 0 || F.c1 < 0) F.patch = zoneToPatch.at(z.id);
            faces.push_back(std::move(F));
        }
    return assembleFaces(std::move(fr.points), nc, std::move(faces), patches);
}

}  // namespace fluent

// A mesh path as readMeshFile chooses (MeshReaders.cpp:805-816): a directory is a
// polyMesh, a file starting "$MeshFormat" Gmsh, a .msh / .cas file or one whose first
// line holds '(' Fluent.
Raw readMeshFile(const std::string& path, const meshio::Options& opt) {
    namespace fs = std::filesystem;
    if (fs::is_directory(path)) return readPolyMesh(path);
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string first;
    std::getline(in, first);
    if (first.rfind("$MeshFormat", 0) == 0) return gmsh::read(path, opt);
    const std::string ext = meshio::lower(fs::path(path).extension().string());
    if (ext == ".msh" || ext == ".cas" || first.find('(') != std::string::npos) return fluent::read(path, opt);
    throw std::runtime_error(path + ": unknown mesh format");
}

// Import options from the C arguments; patchTypes holds "name=type" lines.
meshio::Options importOptions(double scale, double depth2D, const char* patchTypes) {
    meshio::Options o;
    o.scale = scale;
    o.depth2D = depth2D;
    std::istringstream ss(patchTypes ? patchTypes : "");
    std::string line;
    while (std::getline(ss, line)) {
        const auto eq = line.rfind('=');
        if (line.empty()) continue;
        if (eq == std::string::npos) throw std::runtime_error("patch type override '" + line + "' is not name=type");
        o.patchTypes[line.substr(0, eq)] = meshio::typeName(line.substr(eq + 1));
    }
    return o;
}

thread_local std::string gErr;

void copyV3(const std::vector<V3>& v, double* out) {
    if (!out) return;
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}
void copyD(const std::vector<double>& v, double* out) {
    if (out && !v.empty()) std::memcpy(out, v.data(), sizeof(double) * v.size());
}

} // namespace

// A new Raw from f(), or null with the message kept for vm_last_error.
template <class F> void* guarded(F f) {
    try {
        return new Raw(f());
    } catch (const std::exception& e) {
        gErr = e.what();
        return nullptr;
    }
}

extern "C" {

const char* vm_last_error() { return gErr.c_str(); }

void* vm_box(int nx, int ny, int nz) {
    try {
        return new Raw(makeBox(nx, ny, nz));
    } catch (const std::exception& e) {
        gErr = e.what();
        return nullptr;
    }
}

// generateBox with periodic directions (bit 1 x, 2 y, 4 z: cyclic patch pairs) and 2-D
// (nz = 1, frontAndBack empty)
void* vm_box_spec(int nx, int ny, int nz, int periodic, int twoD) {
    return guarded([&] { return makeBox(nx, ny, nz, periodic, twoD != 0); });
}

// the partner patch of a cyclic patch p (empty otherwise); -1 for a bad index
int vm_patch_neighbour(void* h, int p, char* out, int len) {
    const Raw& m = *static_cast<Raw*>(h);
    if (p < 0 || p >= int(m.pStart.size()) || !out || len <= 0) return -1;
    const std::string s = p < int(m.pNeighbour.size()) ? m.pNeighbour[p] : std::string();
    std::strncpy(out, s.c_str(), std::size_t(len) - 1);
    out[len - 1] = 0;
    return 0;
}

// replace the point coordinates of a mesh (same count); its geometry is recomputed
int vm_set_points(void* h, const double* pts, int64_t nPoints) {
    Raw& m = *static_cast<Raw*>(h);
    if (!pts || nPoints != int64_t(m.points.size())) {
        gErr = "vm_set_points: point count differs from the mesh's";
        return -1;
    }
    for (int64_t i = 0; i < nPoints; ++i) m.points[i] = V3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    m.haveGeom = false;
    return 0;
}

void* vm_tet(int nx, int ny, int nz, double jitter, uint64_t seed) {
    try {
        return new Raw(makeTet(nx, ny, nz, jitter, seed));
    } catch (const std::exception& e) {
        gErr = e.what();
        return nullptr;
    }
}

// polyMesh directory (ASCII faceList / faceCompactList, or binary faceCompactList)
void* vm_read_polymesh(const char* dir) {
    try {
        return new Raw(readPolyMesh(dir));
    } catch (const std::exception& e) {
        gErr = e.what();
        return nullptr;
    }
}

// Mesh files as the reference's importers read them (MeshImportOptions: coordinates
// multiplied by scale, 2-D meshes extruded by depth2D, patch types overridden by the
// "name=type" lines of patchTypes, which may be null): an ASCII Gmsh 2.2 / 4.x file,
// a Fluent file (ASCII or binary, 2-D or 3-D), or any of these or a polyMesh directory
// chosen by readMeshFile's rules.
void* vm_read_gmsh(const char* file, double scale, double depth2D, const char* patchTypes) {
    return guarded([&] { return gmsh::read(file, importOptions(scale, depth2D, patchTypes)); });
}
void* vm_read_fluent(const char* file, double scale, double depth2D, const char* patchTypes) {
    return guarded([&] { return fluent::read(file, importOptions(scale, depth2D, patchTypes)); });
}
void* vm_read_mesh_file(const char* path, double scale, double depth2D, const char* patchTypes) {
    return guarded([&] { return readMeshFile(path, importOptions(scale, depth2D, patchTypes)); });
}

// Name and type of patch p (empty for generated meshes); returns -1 for a bad index.
int vm_patch_info(void* h, int p, char* name, int nameLen, char* type, int typeLen) {
    const Raw& m = *static_cast<Raw*>(h);
    if (p < 0 || p >= int(m.pStart.size())) return -1;
    auto put = [](const std::string& s, char* out, int len) {
        if (!out || len <= 0) return;
        std::strncpy(out, s.c_str(), std::size_t(len) - 1);
        out[len - 1] = 0;
    };
    put(p < int(m.pName.size()) ? m.pName[p] : std::string(), name, nameLen);
    put(p < int(m.pType.size()) ? m.pType[p] : std::string(), type, typeLen);
    return 0;
}

void vm_free(void* h) { delete static_cast<Raw*>(h); }

// out: nPoints, nFaces, nInternal, nFacePoints, nCells, nPatches
void vm_sizes(void* h, int64_t* out) {
    const Raw& m = *static_cast<Raw*>(h);
    out[0] = int64_t(m.points.size());
    out[1] = m.nFaces();
    out[2] = m.nInternal();
    out[3] = int64_t(m.facePts.size());
    out[4] = m.nCells;
    out[5] = int64_t(m.pStart.size());
}

void vm_raw(void* h, double* pts, int64_t* faceOff, int64_t* facePts, int64_t* owner, int64_t* neighbour,
            int64_t* pStart, int64_t* pSize) {
    const Raw& m = *static_cast<Raw*>(h);
    copyV3(m.points, pts);
    std::memcpy(faceOff, m.faceOff.data(), sizeof(int64_t) * m.faceOff.size());
    std::memcpy(facePts, m.facePts.data(), sizeof(int64_t) * m.facePts.size());
    std::memcpy(owner, m.owner.data(), sizeof(int64_t) * m.owner.size());
    if (!m.neighbour.empty()) std::memcpy(neighbour, m.neighbour.data(), sizeof(int64_t) * m.neighbour.size());
    std::memcpy(pStart, m.pStart.data(), sizeof(int64_t) * m.pStart.size());
    std::memcpy(pSize, m.pSize.data(), sizeof(int64_t) * m.pSize.size());
}

// Any output pointer may be null. Vectors are interleaved xyz.
int vm_geometry(void* h, double* V, double* C, double* Sf, double* magSf, double* Cf, double* d, double* w,
                double* ndc, double* corr) {
    try {
        Raw& m = *static_cast<Raw*>(h);
        if (!m.haveGeom) computeGeometry(m);
        copyD(m.V, V);
        copyV3(m.C, C);
        copyV3(m.Sf, Sf);
        copyD(m.magSf, magSf);
        copyV3(m.Cf, Cf);
        copyV3(m.d, d);
        copyD(m.w, w);
        copyD(m.ndc, ndc);
        copyV3(m.corr, corr);
        return 0;
    } catch (const std::exception& e) {
        gErr = e.what();
        return -1;
    }
}

// Skew vectors (interleaved xyz, nFaces; zero on boundary faces) and the largest
// |skew|/|d| over internal faces (the reference calls a mesh skewed above 1e-8).
int vm_skew(void* h, double* skew, double* skewMax) {
    try {
        Raw& m = *static_cast<Raw*>(h);
        if (!m.haveGeom) computeGeometry(m);
        copyV3(m.skew, skew);
        if (skewMax) *skewMax = m.skewMax;
        return 0;
    } catch (const std::exception& e) {
        gErr = e.what();
        return -1;
    }
}

} // extern "C"

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
#include <charconv>
#include <iterator>
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
// Gmsh MSH files in ASCII, format versions 2.2 and 4.x, as the file-format chapter of
// the Gmsh reference manual describes them: $MeshFormat, $PhysicalNames, $Entities
// (4.x: the physical tags of every geometric entity), $Nodes, $Elements; any other
// section is skipped. Volume elements become the cells; a boundary face takes the
// physical group of the surface element over the same corners.
//
// Face order rules. These fix owner/neighbour, the face order and the point order, and
// through them every later bit (geometry, agglomeration, sums), so they are the ones of
// the reference's importer (readGmsh and assembleFaceMesh, src/fvm/mesh/
// MeshReaders.cpp:131-190,236-482):
//  (1) cells are the volume elements in file order; each lists its faces in the order
//      of its kind's local face table (faceTable below, corner indices in Gmsh's node
//      order), each face's nodes as the table lists them;
//  (2) a face is known by its sorted corner set: its first appearance fixes its node
//      order and first cell, a second appearance its other cell; faces are numbered by
//      first appearance;
//  (3) points are renumbered by first use along that face sequence, unused ones dropped;
//  (4) an internal face has owner < neighbour and its nodes are reversed when the polygon
//      normal (fan about the point average) points against neighbour − owner centre; a
//      boundary face's nodes are reversed when the normal points into its cell (cell
//      centre = mean of the centres of its faces, in face order);
//  (5) internal faces are ordered by (owner, neighbour), then the boundary faces patch
//      by patch in patch order, each patch's in face order; patches with no face are
//      dropped;
//  (6) the patches are the physical groups of the surface elements in order of first
//      use, named by $PhysicalNames (else "patch<tag>") and typed symmetryPlane if the
//      lower-case name contains "symmetry", else wall if it contains "wall", else patch;
//      then "defaultFaces" (wall) for boundary faces in no group. When two surface
//      elements cover the same corners the later one's group counts.
namespace gmsh {

// Element type number → (dimension, corner count). Higher-order types are read by
// their corners, which Gmsh lists first (manual, "MSH node ordering").
bool kindOf(int type, int& dim, int& corners) {
    struct Row { int type, dim, corners; };
    static const Row rows[] = {
        {15, 0, 1},  {1, 1, 2},   {8, 1, 2},   {26, 1, 2},  {27, 1, 2},  {28, 1, 2},  {2, 2, 3},   {9, 2, 3},
        {20, 2, 3},  {21, 2, 3},  {22, 2, 3},  {23, 2, 3},  {24, 2, 3},  {25, 2, 3},  {3, 2, 4},   {10, 2, 4},
        {16, 2, 4},  {36, 2, 4},  {37, 2, 4},  {38, 2, 4},  {4, 3, 4},   {11, 3, 4},  {29, 3, 4},  {30, 3, 4},
        {31, 3, 4},  {5, 3, 8},   {12, 3, 8},  {17, 3, 8},  {92, 3, 8},  {93, 3, 8},  {6, 3, 6},   {13, 3, 6},
        {18, 3, 6},  {90, 3, 6},  {91, 3, 6},  {7, 3, 5},   {14, 3, 5},  {19, 3, 5},  {118, 3, 5}, {119, 3, 5}};
    for (const Row& r : rows)
        if (r.type == type) {
            dim = r.dim;
            corners = r.corners;
            return true;
        }
    return false;
}

// Local faces of a volume element with `corners` corners (4 tet, 5 pyramid, 6 prism,
// 8 hexahedron): per face its corner count and corner indices.
struct LocalFace { int n; int v[4]; };
const std::vector<LocalFace>& faceTable(int corners) {
    static const std::vector<LocalFace> tet = {{3, {0, 2, 1}}, {3, {0, 1, 3}}, {3, {0, 3, 2}}, {3, {1, 2, 3}}};
    static const std::vector<LocalFace> pyramid = {
        {4, {0, 3, 2, 1}}, {3, {0, 1, 4}}, {3, {1, 2, 4}}, {3, {2, 3, 4}}, {3, {3, 0, 4}}};
    static const std::vector<LocalFace> prism = {
        {3, {0, 2, 1}}, {3, {3, 4, 5}}, {4, {0, 1, 4, 3}}, {4, {1, 2, 5, 4}}, {4, {2, 0, 3, 5}}};
    static const std::vector<LocalFace> hex = {{4, {0, 3, 2, 1}}, {4, {4, 5, 6, 7}}, {4, {0, 1, 5, 4}},
                                               {4, {1, 2, 6, 5}}, {4, {2, 3, 7, 6}}, {4, {3, 0, 4, 7}}};
    return corners == 4 ? tet : corners == 5 ? pyramid : corners == 6 ? prism : hex;
}

// The file as lines of whitespace-separated tokens, read in one piece.
class Tokens {
public:
    explicit Tokens(const std::string& path) : path_(path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error(path + ": cannot be opened");
        text_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    const std::string& path() const { return path_; }
    bool atEnd() { skipSpace(); return pos_ >= text_.size(); }
    // the rest of the current line (without CR/LF), advancing past it
    std::string restOfLine() {
        const size_t e = text_.find('\n', pos_);
        const size_t stop = e == std::string::npos ? text_.size() : e;
        std::string s = text_.substr(pos_, stop - pos_);
        if (!s.empty() && s.back() == '\r') s.pop_back();
        pos_ = e == std::string::npos ? text_.size() : e + 1;
        return s;
    }
    std::string word() {
        skipSpace();
        const size_t b = pos_;
        while (pos_ < text_.size() && !std::isspace(static_cast<unsigned char>(text_[pos_]))) ++pos_;
        if (b == pos_) fail("unexpected end of file");
        return text_.substr(b, pos_ - b);
    }
    int64_t integer() {
        const std::string w = word();
        int64_t v = 0;
        const auto r = std::from_chars(w.data(), w.data() + w.size(), v);
        if (r.ec != std::errc() || r.ptr != w.data() + w.size()) fail("expected an integer, found '" + w + "'");
        return v;
    }
    double real() {
        const std::string w = word();
        char* end = nullptr;
        const double v = std::strtod(w.c_str(), &end);
        if (end != w.c_str() + w.size()) fail("expected a number, found '" + w + "'");
        return v;
    }
    // integers of the current line (up to its end)
    std::vector<int64_t> lineIntegers() {
        std::vector<int64_t> out;
        const std::string s = restOfLine();
        const char* p = s.data();
        const char* e = p + s.size();
        while (true) {
            while (p < e && std::isspace(static_cast<unsigned char>(*p))) ++p;
            if (p == e) break;
            int64_t v = 0;
            const auto r = std::from_chars(p, e, v);
            if (r.ec != std::errc()) fail("bad integer in line '" + s + "'");
            out.push_back(v);
            p = r.ptr;
        }
        return out;
    }
    // skips to the line after "$End<name>"
    void closeSection(const std::string& name) {
        const std::string tag = "$End" + name;
        while (pos_ < text_.size()) {
            const std::string s = restOfLine();
            if (s.compare(0, tag.size(), tag) == 0) return;
        }
        fail("section $" + name + " has no " + tag);
    }
    [[noreturn]] void fail(const std::string& what) const { throw std::runtime_error(path_ + ": " + what); }

private:
    void skipSpace() {
        while (pos_ < text_.size() && std::isspace(static_cast<unsigned char>(text_[pos_]))) ++pos_;
    }
    std::string path_, text_;
    size_t pos_ = 0;
};

// What the sections hold: points by node tag, elements (dimension, physical group or
// −1, corner node tags) in file order, physical names by (dimension, tag).
struct Content {
    double version = 0;
    std::map<std::pair<int, int64_t>, std::string> physName;
    std::map<std::pair<int, int64_t>, int64_t> entityGroup;  // 4.x: first physical tag of an entity, or −1
    std::unordered_map<int64_t, int64_t> pointOfTag;
    std::vector<V3> points;
    std::vector<int8_t> dim;
    std::vector<int64_t> group;
    std::vector<int64_t> cornerOff{0}, corners;
};

void sectionFormat(Tokens& t, Content& c) {
    c.version = t.real();
    const int64_t fileType = t.integer();
    t.integer();  // data size
    if (fileType != 0) t.fail("binary MSH files are not read; write the mesh as ASCII");
    if (!(c.version >= 2.0 && c.version < 5.0)) t.fail("MSH version " + std::to_string(c.version) + " is not 2.x or 4.x");
    t.closeSection("MeshFormat");
}

void sectionPhysicalNames(Tokens& t, Content& c) {
    for (int64_t k = t.integer(); k > 0; --k) {
        const int d = int(t.integer());
        const int64_t tag = t.integer();
        const std::string rest = t.restOfLine();
        const size_t a = rest.find('"'), b = rest.rfind('"');
        c.physName[{d, tag}] = a != std::string::npos && b > a ? rest.substr(a + 1, b - a - 1) : std::to_string(tag);
    }
    t.closeSection("PhysicalNames");
}

void sectionEntities(Tokens& t, Content& c) {
    int64_t count[4];
    for (int64_t& n : count) n = t.integer();
    for (int d = 0; d < 4; ++d)
        for (int64_t k = 0; k < count[d]; ++k) {
            const int64_t tag = t.integer();
            for (int j = 0; j < (d == 0 ? 3 : 6); ++j) t.real();  // point coordinates / bounding box
            int64_t first = -1;
            for (int64_t j = 0, n = t.integer(); j < n; ++j) {
                const int64_t g = std::abs(t.integer());
                if (j == 0) first = g;
            }
            c.entityGroup[{d, tag}] = first;
            if (d > 0)
                for (int64_t j = 0, n = t.integer(); j < n; ++j) t.integer();  // bounding entities
        }
    t.closeSection("Entities");
}

void addPoint(Content& c, int64_t tag, V3 p) {
    c.pointOfTag[tag] = int64_t(c.points.size());
    c.points.push_back(p);
}

void sectionNodes(Tokens& t, Content& c, double scale) {
    if (c.version >= 4.0) {
        const int64_t blocks = t.integer();
        t.integer();  // node count
        t.integer();  // smallest tag
        t.integer();  // largest tag
        for (int64_t b = 0; b < blocks; ++b) {
            const int64_t d = t.integer();
            t.integer();  // entity tag
            const int64_t parametric = t.integer();
            const int64_t n = t.integer();
            std::vector<int64_t> tags(n);
            for (int64_t& g : tags) g = t.integer();
            for (int64_t i = 0; i < n; ++i) {
                V3 p;
                p.x = t.real();
                p.y = t.real();
                p.z = t.real();
                for (int64_t j = 0; j < (parametric ? d : 0); ++j) t.real();
                addPoint(c, tags[i], p * scale);
            }
        }
    } else {
        for (int64_t k = t.integer(); k > 0; --k) {
            const int64_t tag = t.integer();
            V3 p;
            p.x = t.real();
            p.y = t.real();
            p.z = t.real();
            addPoint(c, tag, p * scale);
        }
    }
    t.closeSection("Nodes");
}

void addElement(Tokens& t, Content& c, int type, int64_t group, const int64_t* nodes, size_t available) {
    int d = 0, nc = 0;
    if (!kindOf(type, d, nc)) t.fail("element type " + std::to_string(type) + " is not supported");
    c.dim.push_back(int8_t(d));
    c.group.push_back(group);
    for (size_t k = 0; k < size_t(nc) && k < available; ++k) c.corners.push_back(nodes[k]);
    c.cornerOff.push_back(int64_t(c.corners.size()));
}

void sectionElements(Tokens& t, Content& c) {
    if (c.version >= 4.0) {
        const int64_t blocks = t.integer();
        t.integer();  // element count
        t.integer();  // smallest tag
        t.restOfLine();
        for (int64_t b = 0; b < blocks; ++b) {
            const std::vector<int64_t> head = t.lineIntegers();  // entity dim, entity tag, type, count
            if (head.size() < 4) t.fail("bad element block header");
            const auto it = c.entityGroup.find({int(head[0]), head[1]});
            const int64_t group = it != c.entityGroup.end() ? it->second : -1;
            for (int64_t i = 0; i < head[3]; ++i) {
                const std::vector<int64_t> v = t.lineIntegers();  // tag, nodes
                if (v.empty()) t.fail("empty element line");
                addElement(t, c, int(head[2]), group, v.data() + 1, v.size() - 1);
            }
        }
    } else {
        const int64_t n = t.integer();
        t.restOfLine();
        for (int64_t i = 0; i < n; ++i) {
            const std::vector<int64_t> v = t.lineIntegers();  // tag, type, #tags, tags..., nodes
            if (v.size() < 3 || v.size() < size_t(3 + v[2])) t.fail("bad element line");
            const int64_t nt = v[2];
            const int64_t group = nt > 0 && v[3] != 0 ? v[3] : -1;
            addElement(t, c, int(v[1]), group, v.data() + 3 + nt, v.size() - 3 - nt);
        }
    }
    t.closeSection("Elements");
}

Content parse(const std::string& path, double scale) {
    Tokens t(path);
    Content c;
    while (!t.atEnd()) {
        const std::string line = t.restOfLine();
        if (line == "$MeshFormat")
            sectionFormat(t, c);
        else if (line == "$PhysicalNames")
            sectionPhysicalNames(t, c);
        else if (line == "$Entities" && c.version >= 4.0)
            sectionEntities(t, c);
        else if (line == "$Nodes")
            sectionNodes(t, c, scale);
        else if (line == "$Elements")
            sectionElements(t, c);
    }
    if (c.points.empty() || c.dim.empty()) t.fail("no nodes or no elements");
    return c;
}

std::string lowerCase(std::string s) {
    for (char& ch : s) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    return s;
}

// Reads a 3-D mesh; `typeOf(name, type)` gives a patch's final type (overrides).
template <class TypeOf> Raw read(const std::string& path, double scale, TypeOf typeOf) {
    const Content c = parse(path, scale);
    const int64_t nEl = int64_t(c.dim.size());
    int maxDim = 0;
    for (int8_t d : c.dim) maxDim = std::max(maxDim, int(d));
    if (maxDim < 3) throw std::runtime_error(path + ": no volume elements (2-D meshes are not read here)");
    auto pointIndex = [&](int64_t tag) {
        const auto it = c.pointOfTag.find(tag);
        if (it == c.pointOfTag.end()) throw std::runtime_error(path + ": an element names node " + std::to_string(tag) +
                                                               ", which $Nodes does not define");
        return it->second;
    };
    // patches: groups of surface elements in order of first use
    std::vector<std::string> pName, pType;
    std::map<int64_t, int> patchOfGroup;
    for (int64_t e = 0; e < nEl; ++e) {
        if (c.dim[e] != 2 || c.group[e] < 0 || patchOfGroup.count(c.group[e])) continue;
        const auto it = c.physName.find({2, c.group[e]});
        const std::string name = it != c.physName.end() ? it->second : "patch" + std::to_string(c.group[e]);
        const std::string low = lowerCase(name);
        const std::string type = low.find("symmetry") != std::string::npos ? "symmetryPlane"
                                 : low.find("wall") != std::string::npos   ? "wall"
                                                                           : "patch";
        for (const std::string& other : pName)
            if (other == name) throw std::runtime_error(path + ": two physical groups are named " + name);
        patchOfGroup[c.group[e]] = int(pName.size());
        pName.push_back(name);
        pType.push_back(type);
    }
    const int defaultPatch = int(pName.size());
    pName.push_back("defaultFaces");
    pType.push_back("wall");
    for (size_t p = 0; p < pName.size(); ++p) pType[p] = typeOf(pName[p], pType[p]);

    // every (cell, local face) in generation order, keyed by its sorted corners
    using Key = std::array<int64_t, 4>;
    auto keyOf = [](const int64_t* v, int n) {
        Key k{-1, -1, -1, -1};
        std::copy(v, v + n, k.begin());
        std::sort(k.begin(), k.begin() + n);
        return k;
    };
    struct Gen { Key key; int64_t cell; int n; int64_t v[4]; };
    std::vector<Gen> gen;
    int64_t nCells = 0;
    for (int64_t e = 0; e < nEl; ++e) {
        if (c.dim[e] != 3) continue;
        const int64_t* cn = c.corners.data() + c.cornerOff[e];
        const int nc = int(c.cornerOff[e + 1] - c.cornerOff[e]);
        for (const LocalFace& lf : faceTable(nc)) {
            Gen g;
            g.cell = nCells;
            g.n = lf.n;
            for (int k = 0; k < lf.n; ++k) {
                if (lf.v[k] >= nc) throw std::runtime_error(path + ": a volume element has too few nodes");
                g.v[k] = pointIndex(cn[lf.v[k]]);
            }
            g.key = keyOf(g.v, g.n);
            gen.push_back(g);
        }
        ++nCells;
    }
    // match by sorting (key, generation index): the first of a run is the face
    std::vector<int64_t> order(gen.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int64_t(i);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return gen[a].key != gen[b].key ? gen[a].key < gen[b].key : a < b;
    });
    std::vector<int64_t> other(gen.size(), -2);  // per first appearance: the second cell (−1 none); −2 not first
    for (size_t i = 0; i < order.size();) {
        size_t j = i + 1;
        while (j < order.size() && gen[order[j]].key == gen[order[i]].key) ++j;
        if (j - i > 2) throw std::runtime_error(path + ": a face is shared by more than two cells");
        other[order[i]] = j - i == 2 ? gen[order[i + 1]].cell : -1;
        i = j;
    }
    // boundary groups of the surface elements, the later element winning
    std::map<Key, int> patchOfKey;
    for (int64_t e = 0; e < nEl; ++e) {
        if (c.dim[e] != 2 || c.group[e] < 0) continue;
        const int n = int(c.cornerOff[e + 1] - c.cornerOff[e]);
        int64_t v[4];
        for (int k = 0; k < n; ++k) v[k] = pointIndex(c.corners[c.cornerOff[e] + k]);
        patchOfKey[keyOf(v, n)] = patchOfGroup.at(c.group[e]);
    }
    // faces by first appearance; points renumbered by first use
    struct Face { int64_t c0, c1; int patch; std::vector<int64_t> v; };
    std::vector<Face> faces;
    std::vector<int64_t> newIndex(c.points.size(), -1);
    std::vector<V3> pts;
    int64_t unmatched = 0;
    for (size_t i = 0; i < gen.size(); ++i) {
        if (other[i] == -2) continue;
        Face f;
        f.c0 = gen[i].cell;
        f.c1 = other[i];
        f.patch = -1;
        if (f.c1 < 0) {
            const auto it = patchOfKey.find(gen[i].key);
            f.patch = it != patchOfKey.end() ? it->second : (++unmatched, defaultPatch);
        }
        for (int k = 0; k < gen[i].n; ++k) {
            int64_t& ni = newIndex[gen[i].v[k]];
            if (ni < 0) {
                ni = int64_t(pts.size());
                pts.push_back(c.points[gen[i].v[k]]);
            }
            f.v.push_back(ni);
        }
        faces.push_back(std::move(f));
    }
    // orientation from estimated centres
    auto centreAndNormal = [&](const std::vector<int64_t>& v, V3& centre) {
        V3 s;
        for (int64_t p : v) s += pts[p];
        centre = s / double(v.size());
        V3 n;
        for (size_t k = 0; k < v.size(); ++k) n += cross(pts[v[k]] - centre, pts[v[(k + 1) % v.size()]] - centre);
        return 0.5 * n;
    };
    std::vector<V3> cellC(nCells), faceC(faces.size()), faceN(faces.size());
    std::vector<int> faceCount(nCells, 0);
    for (size_t f = 0; f < faces.size(); ++f) {
        faceN[f] = centreAndNormal(faces[f].v, faceC[f]);
        cellC[faces[f].c0] += faceC[f];
        ++faceCount[faces[f].c0];
        if (faces[f].c1 >= 0) {
            cellC[faces[f].c1] += faceC[f];
            ++faceCount[faces[f].c1];
        }
    }
    for (int64_t k = 0; k < nCells; ++k) {
        if (faceCount[k] < 4) throw std::runtime_error(path + ": cell " + std::to_string(k) + " has only " +
                                                       std::to_string(faceCount[k]) + " faces");
        cellC[k] = cellC[k] / double(faceCount[k]);
    }
    std::vector<int64_t> internal;
    std::vector<std::vector<int64_t>> byPatch(pName.size());
    for (size_t f = 0; f < faces.size(); ++f) {
        Face& F = faces[f];
        if (F.c1 >= 0) {
            if (F.c1 < F.c0) std::swap(F.c0, F.c1);
            if (dot(faceN[f], cellC[F.c1] - cellC[F.c0]) < 0) std::reverse(F.v.begin(), F.v.end());
            internal.push_back(int64_t(f));
        } else {
            if (dot(faceN[f], faceC[f] - cellC[F.c0]) < 0) std::reverse(F.v.begin(), F.v.end());
            byPatch[F.patch].push_back(int64_t(f));
        }
    }
    std::stable_sort(internal.begin(), internal.end(), [&](int64_t a, int64_t b) {
        return faces[a].c0 != faces[b].c0 ? faces[a].c0 < faces[b].c0 : faces[a].c1 < faces[b].c1;
    });
    Raw r;
    r.points = std::move(pts);
    r.nCells = nCells;
    auto emit = [&](const Face& F) {
        r.facePts.insert(r.facePts.end(), F.v.begin(), F.v.end());
        r.faceOff.push_back(int64_t(r.facePts.size()));
        r.owner.push_back(F.c0);
    };
    for (int64_t f : internal) {
        emit(faces[f]);
        r.neighbour.push_back(faces[f].c1);
    }
    for (size_t p = 0; p < pName.size(); ++p) {
        if (byPatch[p].empty()) continue;
        const int64_t start = r.nFaces();
        for (int64_t f : byPatch[p]) emit(faces[f]);
        r.pName.push_back(pName[p]);
        r.pType.push_back(pType[p]);
        r.pNeighbour.push_back("");
        r.pStart.push_back(start);
        r.pSize.push_back(r.nFaces() - start);
    }
    (void)unmatched;
    return r;
}

}  // namespace gmsh

// Patch type spellings accepted in overrides, as the reference's patch types print:
// wall, symmetryPlane (also "symmetry"), empty, cyclic, patch (anything else).
std::string canonicalType(const std::string& s) {
    if (s == "wall" || s == "empty" || s == "cyclic") return s;
    if (s == "symmetry" || s == "symmetryPlane") return "symmetryPlane";
    return "patch";
}

// "name=type" lines → overrides of patch types by name.
std::map<std::string, std::string> typeOverrides(const char* lines) {
    std::map<std::string, std::string> out;
    std::istringstream ss(lines ? lines : "");
    std::string line;
    while (std::getline(ss, line)) {
        if (line.empty()) continue;
        const auto eq = line.rfind('=');
        if (eq == std::string::npos) throw std::runtime_error("patch type override '" + line + "' has no '='");
        out[line.substr(0, eq)] = canonicalType(line.substr(eq + 1));
    }
    return out;
}

Raw readGmshFile(const std::string& path, double scale, const char* overrides) {
    const auto over = typeOverrides(overrides);
    return gmsh::read(path, scale, [&](const std::string& name, const std::string& type) {
        const auto it = over.find(name);
        return it != over.end() ? it->second : type;
    });
}

// A directory is read as a polyMesh, a file whose first line is "$MeshFormat" as Gmsh.
Raw readMeshPath(const std::string& path, double scale, const char* overrides) {
    namespace fs = std::filesystem;
    if (fs::is_directory(path)) return readPolyMesh(path);
    std::ifstream in(path);
    if (!in) throw std::runtime_error(path + ": cannot be opened");
    std::string first;
    std::getline(in, first);
    if (first.rfind("$MeshFormat", 0) == 0) return readGmshFile(path, scale, overrides);
    throw std::runtime_error(path + ": neither a polyMesh directory nor a Gmsh MSH file");
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

// A 3-D Gmsh MSH file (ASCII, 2.2 or 4.x): coordinates multiplied by scale, patch types
// replaced by the "name=type" lines of patchTypes (nullable).
void* vm_read_gmsh(const char* file, double scale, const char* patchTypes) {
    return guarded([&] { return readGmshFile(file, scale, patchTypes); });
}
// A polyMesh directory or a Gmsh file, told apart by what the path is.
void* vm_read_mesh_file(const char* path, double scale, const char* patchTypes) {
    return guarded([&] { return readMeshPath(path, scale, patchTypes); });
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

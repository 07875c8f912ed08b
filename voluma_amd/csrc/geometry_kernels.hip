// Mesh geometry on the GPU from points and faces, with the reference's arithmetic
// (Mesh.cpp:21-48 faces, the pyramid decomposition of cells, Mesh.cpp:600-630 face
// quantities): face area vectors and centres (a triangle directly, other polygons by the
// fan about the point average), cell centres and volumes (pyramids on the estimated
// centre, faces in ascending order), then per face d, |S|, w, ndc, corr and the skew
// vector. Every sum runs in the reference's order, so the results are bit-identical.
#include "device_common.hpp"
#include "vol_internal.hpp"

namespace vol {

namespace {
constexpr int GEO_THREADS = 256;
constexpr double GEO_VSMALL = 1.0e-300;
inline dim3 geo_grid(long long cnt) { return dim3(unsigned((cnt + GEO_THREADS - 1) / GEO_THREADS)); }

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 ld3(const double* p, long long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double* p, long long i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 dvd(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double mag(V3 a) { return sqrt(dot(a, a)); }
}  // namespace

// Face area vector and centre (Mesh.cpp:21-48).
__global__ void k_geo_faces(GeoArgs g) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.nFaces) return;
    const int b = g.faceOff[f], e = g.faceOff[f + 1], n = e - b;
    const int* fp = g.facePts + b;
    if (n == 3) {
        const V3 p0 = ld3(g.points, fp[0]), p1 = ld3(g.points, fp[1]), p2 = ld3(g.points, fp[2]);
        st3(g.Cf, f, dvd(add(add(p0, p1), p2), 3.0));
        st3(g.Sf, f, mul(cross(sub(p1, p0), sub(p2, p0)), 0.5));
        return;
    }
    V3 pAvg = {0.0, 0.0, 0.0};
    for (int k = 0; k < n; ++k) pAvg = add(pAvg, ld3(g.points, fp[k]));
    pAvg = dvd(pAvg, double(n));
    V3 sumN = {0.0, 0.0, 0.0}, sumAc = {0.0, 0.0, 0.0};
    double sumA = 0.0;
    for (int k = 0; k < n; ++k) {
        const V3 p0 = ld3(g.points, fp[k]);
        const V3 p1 = ld3(g.points, fp[k + 1 < n ? k + 1 : 0]);
        const V3 c = add(add(p0, p1), pAvg);
        const V3 nk = cross(sub(p1, p0), sub(pAvg, p0));
        const double a = mag(nk);
        sumN = add(sumN, nk);
        sumA += a;
        sumAc = add(sumAc, mul(c, a));
    }
    st3(g.Cf, f, sumA < GEO_VSMALL ? pAvg : dvd(sumAc, 3.0 * sumA));
    st3(g.Sf, f, mul(sumN, 0.5));
}

// Cell centre and volume: the mean of the face centres as the apex estimate, then the
// pyramid on each face (faces ascending; the area vector flipped on neighbour faces):
// pv = max(S·(Cf − cEst), VSMALL), C = Σ pv (¾Cf + ¼cEst) / Σ pv, V = Σ pv / 3.
__global__ void k_geo_cells(GeoArgs g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.nCells) return;
    const int k0 = g.cfOff[c], k1 = g.cfOff[c + 1];
    V3 cEst = {0.0, 0.0, 0.0};
    for (int k = k0; k < k1; ++k) cEst = add(cEst, ld3(g.Cf, g.cfList[k]));
    cEst = dvd(cEst, double(k1 - k0));
    V3 cSum = {0.0, 0.0, 0.0};
    double vSum = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int f = g.cfList[k];
        V3 S = ld3(g.Sf, f);
        if (g.owner[f] != c) S = {-S.x, -S.y, -S.z};
        const V3 Cf = ld3(g.Cf, f);
        const double d = dot(S, sub(Cf, cEst));
        const double pv = d < GEO_VSMALL ? GEO_VSMALL : d;  // std::max(d, VSMALL)
        cSum = add(cSum, mul(add(mul(Cf, 0.75), mul(cEst, 0.25)), pv));
        vSum += pv;
    }
    st3(g.C, c, dvd(cSum, vSum));
    g.V[c] = vSum / 3.0;
}

// Face quantities (Mesh.cpp:600-630): d, |S|, w (1 on boundary faces), ndc, corr and the
// skew vector of internal faces; the largest |skew|/|d| by an atomic max on the bits of
// the (non-negative) ratio.
__global__ void k_geo_face_derived(GeoArgs g) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.nFaces) return;
    const V3 Sf = ld3(g.Sf, f), Cf = ld3(g.Cf, f);
    const V3 CP = ld3(g.C, g.owner[f]);
    const V3 dd = f < g.nInternal ? sub(ld3(g.C, g.neighbour[f]), CP) : sub(Cf, CP);
    st3(g.d, f, dd);
    const double ms = mag(Sf);
    g.magSf[f] = ms;
    const V3 nv = dvd(Sf, ms);
    double w = 1.0;
    V3 sk = {0.0, 0.0, 0.0};
    if (f < g.nInternal) {
        const double sOwn = fabs(dot(Sf, sub(Cf, CP)));
        const double sNei = fabs(dot(Sf, sub(add(CP, dd), Cf)));
        w = sNei / (sOwn + sNei);
        sk = sub(Cf, add(CP, mul(dd, 1.0 - w)));
        const double ratio = mag(sk) / mag(dd);
        if (g.maxSkewBits) atomicMax(g.maxSkewBits, (unsigned long long)__double_as_longlong(ratio));
    }
    g.w[f] = w;
    if (g.skew) st3(g.skew, f, sk);
    const double nd = dot(nv, dd), lim = 0.05 * mag(dd);
    const double ndc = 1.0 / (nd < lim ? lim : nd);  // 1 / max(n·d, 0.05|d|)
    g.ndc[f] = ndc;
    st3(g.corr, f, sub(nv, mul(dd, ndc)));
}

void launch_geometry(hipStream_t st, const GeoArgs& g) {
    if (g.nFaces > 0) hipLaunchKernelGGL(k_geo_faces, geo_grid(g.nFaces), dim3(GEO_THREADS), 0, st, g);
    if (g.nCells > 0) hipLaunchKernelGGL(k_geo_cells, geo_grid(g.nCells), dim3(GEO_THREADS), 0, st, g);
    if (g.nFaces > 0) hipLaunchKernelGGL(k_geo_face_derived, geo_grid(g.nFaces), dim3(GEO_THREADS), 0, st, g);
}

}  // namespace vol

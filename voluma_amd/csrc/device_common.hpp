// Device-side building blocks shared by the solver kernels (gfx950, wave64).
//
//  * wave/block sum reductions for double;
//  * the exact (order-independent) FP64 accumulator: the same fixed-point layout as
//    the reference's par::ExactSum (src/fvm/parallel/Comm.h:42-93) — 72 int64 bins of
//    32 bits, each double split into three 32-bit pieces at bin p>>5 — so that the
//    reduced integer bins, converted top-down with ldexp (Comm.cpp:135-156), give the
//    reference's reproducible-mode sum bit for bit, whatever the reduction order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vol_internal.hpp"

namespace vol {

constexpr double VSMALL = 1.0e-300;
constexpr double GREAT = 1.0e15;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Deterministic block sum (fixed tree: lanes by xor-shuffle, then waves in order).
// `sh` needs blockDim.x/64 doubles. Result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < nw; ++k) t += sh[k];
    return t;
}

// ------------------------------------------------------------------ exact sums
// Piece decomposition of one double, as ExactSum::add does it.
struct ExactPieces {
    int bin;  // −1: contributes nothing (zero); −2: non-finite
    long long lo, mid, hi;
};

__device__ __forceinline__ ExactPieces exact_split(double v) {
    ExactPieces P;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const int ex = int((bits >> 52) & 0x7ff);
    if (ex == 0x7ff) {
        P.bin = -2;
        P.lo = P.mid = P.hi = 0;
        return P;
    }
    long long M = (long long)(bits & ((1ull << 52) - 1));
    if (ex != 0) M |= 1ll << 52;
    if (M == 0) {
        P.bin = -1;
        P.lo = P.mid = P.hi = 0;
        return P;
    }
    const int p = (ex == 0 ? 1 : ex) - 1;
    const int s = p & 31;
    const long long sgn = -(long long)(bits >> 63);
    long long lo = (M & ((1ll << (32 - s)) - 1)) << s;
    const long long r = M >> (32 - s);
    long long mid = r & 0xFFFFFFFFll;
    long long hi = r >> 32;
    P.bin = p >> 5;
    P.lo = (lo ^ sgn) - sgn;
    P.mid = (mid ^ sgn) - sgn;
    P.hi = (hi ^ sgn) - sgn;
    return P;
}

// Non-finite tracking: the reference sums the non-finite values in order; the result
// depends only on which of NaN, +Inf, −Inf occurred (any NaN or both infinities → NaN).
enum : int { NF_NAN = 1, NF_PINF = 2, NF_NINF = 4 };
__device__ __forceinline__ int nonfinite_flag(double v) {
    if (v != v) return NF_NAN;
    return v > 0 ? NF_PINF : NF_NINF;
}

// Block-level exact accumulator in LDS: bins[72] int64 + a non-finite flag word.
// Every thread of the block calls exact_add_wave with one value (active=false to skip).
// Within a wave, when all active values fall in a 5-bin window the pieces are summed
// across lanes first (≤ 64·2^32 < 2^39, no overflow) and added once per bin.
__device__ __forceinline__ void exact_add_wave(long long* bins, int* nfFlag, double v, bool active) {
    ExactPieces P = active ? exact_split(v) : ExactPieces{-1, 0, 0, 0};
    if (P.bin == -2) {
        atomicOr(nfFlag, nonfinite_flag(v));
        P.bin = -1;
    }
    const bool has = P.bin >= 0;
    const int bmin = wave_min_i(has ? P.bin : 1 << 20);
    if (bmin == (1 << 20)) return;  // nothing in this wave (uniform branch)
    const int bmax = wave_max_i(has ? P.bin : -1);
    const int lane = threadIdx.x & (WAVE - 1);
    if (bmax - bmin <= 2) {
        long long w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0;
        if (has) {
            const int o = P.bin - bmin;
            w0 = o == 0 ? P.lo : 0;
            w1 = o == 0 ? P.mid : (o == 1 ? P.lo : 0);
            w2 = o == 0 ? P.hi : (o == 1 ? P.mid : P.lo);
            w3 = o == 1 ? P.hi : (o == 2 ? P.mid : 0);
            w4 = o == 2 ? P.hi : 0;
        }
        w0 = wave_sum_ll(w0);
        w1 = wave_sum_ll(w1);
        w2 = wave_sum_ll(w2);
        w3 = wave_sum_ll(w3);
        w4 = wave_sum_ll(w4);
        if (lane == 0) {
            atomicAdd((unsigned long long*)&bins[bmin], (unsigned long long)w0);
            atomicAdd((unsigned long long*)&bins[bmin + 1], (unsigned long long)w1);
            atomicAdd((unsigned long long*)&bins[bmin + 2], (unsigned long long)w2);
            atomicAdd((unsigned long long*)&bins[bmin + 3], (unsigned long long)w3);
            atomicAdd((unsigned long long*)&bins[bmin + 4], (unsigned long long)w4);
        }
    } else if (has) {
        atomicAdd((unsigned long long*)&bins[P.bin], (unsigned long long)P.lo);
        atomicAdd((unsigned long long*)&bins[P.bin + 1], (unsigned long long)P.mid);
        atomicAdd((unsigned long long*)&bins[P.bin + 2], (unsigned long long)P.hi);
    }
}

// Carry-normalise bins (Comm.cpp:114-121): afterwards bins[0..70] ∈ [0, 2^32).
__device__ __forceinline__ void exact_normalize(long long* b) {
    for (int k = 0; k < NBINS - 1; ++k) {
        const long long c = b[k] >> 32;
        b[k] -= c * (1ll << 32);
        b[k + 1] += c;
    }
}

// binsToDouble (Comm.cpp:135-156) on a private copy.
__device__ __forceinline__ double exact_to_double(long long* b, int nfFlag) {
    if (nfFlag) {
        if ((nfFlag & NF_NAN) || ((nfFlag & NF_PINF) && (nfFlag & NF_NINF))) return __longlong_as_double(0x7ff8000000000000ll);
        return (nfFlag & NF_PINF) ? __longlong_as_double(0x7ff0000000000000ll) : __longlong_as_double((long long)0xfff0000000000000ull);
    }
    exact_normalize(b);
    double sign = 1.0;
    if (b[NBINS - 1] < 0) {
        sign = -1.0;
        for (int k = 0; k < NBINS; ++k) b[k] = -b[k];
        exact_normalize(b);
    }
    double r = 0.0;
    for (int k = NBINS - 1; k >= 0; --k)
        if (b[k] != 0) r += ldexp((double)b[k], 32 * k - 1074);
    return sign * r;
}

}  // namespace vol

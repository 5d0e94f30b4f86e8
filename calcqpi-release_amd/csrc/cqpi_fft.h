// In-LDS FFT helpers shared by the forward transform and its adjoint.
#pragma once

#include "cqpi_common.h"

namespace cqpi {

__device__ inline int imod(int a, int m)
{
    int r = a % m;
    return r < 0 ? r + m : r;
}

__device__ inline int brev(int j, int lgK)
{
    return lgK ? (int)(__brev((unsigned)j) >> (32 - lgK)) : 0;
}

// In-place decimation-in-time transform of np lines of length K (power of
// two) stored [q*K + idx] in bit-reversed order: out[m] = sum_j in[j] w^(jm)
// with w = exp(sign * 2 pi i / K).  tw[m] = w^m for m < K/2 (the table must
// match the sign).  Radix-4 stages (two radix-2 stages fused), plus one
// radix-2 stage when log2 K is odd.  Ends with a barrier.
__device__ inline void fft_lines(double2 *S, int np, int K, int lgK, const double2 *__restrict__ tw, int sign)
{
    int lg = 1;
    if (lgK & 1) {  // one radix-2 stage (span 2) first
        const int halfK = K >> 1;
        for (int t = threadIdx.x; t < np * halfK; t += blockDim.x) {
            int q = t / halfK, u = t - q * halfK;
            double2 *L = S + (size_t)q * K + 2 * u;
            double2 a = L[0], b = L[1];
            L[0] = cadd(a, b);
            L[1] = csub(a, b);
        }
        __syncthreads();
        lg = 2;
    }
    // fused pairs of radix-2 stages: spans L = 2^lg and 2L
    const int quarterK = K >> 2;
    for (; lg < lgK; lg += 2) {
        const int L = 1 << lg, hL = L >> 1;
        const int ts1 = K >> lg, ts2 = K >> (lg + 1);
        for (int t = threadIdx.x; t < np * quarterK; t += blockDim.x) {
            int q = t / quarterK, u = t - q * quarterK;
            int blk = u >> (lg - 1), r = u & (hL - 1);
            double2 *X = S + (size_t)q * K + blk * 2 * L + r;
            double2 w1 = tw[r * ts1];  // w_L^r
            double2 w2 = tw[r * ts2];  // w_2L^r
            // w_2L^(r + L/2) = w_2L^r * (+-i)
            double2 w3 = sign > 0 ? make_double2(-w2.y, w2.x) : make_double2(w2.y, -w2.x);
            double2 a0 = X[0], a1 = cmul(X[hL], w1), a2 = X[L], a3 = cmul(X[L + hL], w1);
            double2 b0 = cadd(a0, a1), b1 = csub(a0, a1), b2 = cadd(a2, a3), b3 = csub(a2, a3);
            double2 c2 = cmul(b2, w2), c3 = cmul(b3, w3);
            X[0] = cadd(b0, c2);
            X[L] = csub(b0, c2);
            X[hL] = cadd(b1, c3);
            X[L + hL] = csub(b1, c3);
        }
        __syncthreads();
    }
}

// one output point of a line (natural order) when K is not a power of two;
// tw[m] = w^m for all m < K
__device__ inline double2 dft_point(const double2 *L, int K, int Rm, const double2 *__restrict__ tw)
{
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < K; j++) acc = cadd(acc, cmul(L[j], tw[(int)(((long)j * Rm) % K)]));
    return acc;
}

}  // namespace cqpi

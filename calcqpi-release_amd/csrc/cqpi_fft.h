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


// ---------------------------------------------------------------------------
// Register-staged Stockham transform of one line per lane group.
//
// A line of N = 2^lg points (natural order, in LDS) is transformed in place
// by a group of LN lanes of one wave: out[m] = sum_j in[j] w^(jm), w =
// exp(sign 2 pi i / N).  Each pass of radix R (8 while possible, then 4 or 2)
// loads R points per butterfly into registers (stride N/R: consecutive lanes
// read consecutive addresses), applies the twiddles w_{LR}^{j r} from the
// table tw[m] = w^m (m < N), does the radix-R DFT in registers and stores
// the outputs in Stockham order, so the result is in natural order and no
// bit reversal is needed.  Only the lanes of one wave touch a line, so the
// passes need no workgroup barrier: a wavefront-scope fence orders the LDS
// accesses between passes.
// ---------------------------------------------------------------------------

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// multiply by sign * i
template <typename C>
__device__ inline C mul_si(C a, int sign)
{
    C r;
    r.x = sign > 0 ? -a.y : a.y;
    r.y = sign > 0 ? a.x : -a.x;
    return r;
}

// in-register DFT of R points (C = double2 or float2)
template <int R, typename C>
__device__ inline void dft_small(C *x, int sign)
{
    if constexpr (R == 2) {
        C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 4) {
        C a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
        C a2 = cadd(x[1], x[3]), a3 = mul_si(csub(x[1], x[3]), sign);
        x[0] = cadd(a0, a2);
        x[2] = csub(a0, a2);
        x[1] = cadd(a1, a3);
        x[3] = csub(a1, a3);
    } else {
        C e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
        dft_small<4>(e, sign);
        dft_small<4>(o, sign);
        using Re = decltype(x[0].x);
        constexpr Re r2 = (Re)0.70710678118654752440;
        // w8^q o_q, w8 = (1 + sign i)/sqrt2
        C o1 = cscale(r2, cadd(o[1], mul_si(o[1], sign)));
        C o2 = mul_si(o[2], sign);
        C o3 = cscale(r2, csub(mul_si(o[3], sign), o[3]));
        x[0] = cadd(e[0], o[0]);
        x[4] = csub(e[0], o[0]);
        x[1] = cadd(e[1], o1);
        x[5] = csub(e[1], o1);
        x[2] = cadd(e[2], o2);
        x[6] = csub(e[2], o2);
        x[3] = cadd(e[3], o3);
        x[7] = csub(e[3], o3);
    }
}

// radix of the first pass for 2^lg remaining: 8 while possible, 4+4 for 16
constexpr int stockham_radix(int lg) { return (lg >= 3 && lg != 4) ? 8 : (lg >= 2 ? 4 : 2); }
constexpr int ilog2c(int r) { return r == 8 ? 3 : (r == 4 ? 2 : 1); }
constexpr int ilog2n(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}
// largest butterfly count of any pass = lanes that one line keeps busy (<= 64)
constexpr int stockham_lanes(int N)
{
    int lg = ilog2n(N);
    int mx = 1;
    while (lg > 0) {
        int r = stockham_radix(lg);
        int nb = N / r;
        if (nb > mx) mx = nb;
        lg -= ilog2c(r);
    }
    return mx < 64 ? mx : 64;
}

// Lines are stored padded so that the strided Stockham stores spread over the
// LDS banks.  FP64 lines (16-byte points): point i at i + i/8 + i/64.  FP32
// lines (8-byte points): point i at i + i/8, which takes the stores of the
// first pass (stride 8 points: 16 lanes on one bank) and of the second
// (blocks of 8 at stride 64) to two lanes a bank, while the loads at stride
// N/8 stay contiguous; the line stride is odd, so that the lanes reading
// one slot of several lines fall on different banks.  A padded line of N
// points takes padded_len(N) slots.
#ifndef CQPI_FFT32_PAD
#define CQPI_FFT32_PAD 1
#endif
// LP = false keeps FP32 lines unpadded (with the natural twiddle table), as
// the first pass does for K = 1024: there its workgroups take one energy each,
// and the padded layout measured slower.
template <typename C, bool LP = true>
__host__ __device__ constexpr int pidx(int i)
{
    return sizeof(C) == 16 ? i + (i >> 3) + (i >> 6) : (LP && CQPI_FFT32_PAD ? i + (i >> 3) : i);
}
template <typename C, bool LP = true>
__host__ __device__ constexpr int padded_len(int N)
{
    return sizeof(C) == 16 ? N + (N >> 3) + (N >> 6) : (LP && CQPI_FFT32_PAD ? N + (N >> 3) + 1 : N);
}

// FP32 lines take their twiddles from a per-pass table in LDS: the pass of
// span L and radix R (L > 1) holds w^(j r N/(L R)) for j < L, r = 1..R-1 at
// [tw_pass_off(N, L) + j (R-1) + r - 1], so that the R-1 twiddles of a
// butterfly are adjacent and the lanes of a wave (consecutive j) read
// addresses R-1 points apart, not j r apart.  At most N entries in all.
#ifndef CQPI_FFT32_TWP
#define CQPI_FFT32_TWP 1
#endif
template <typename C, bool LP = true>
__host__ __device__ constexpr bool pass_twiddles()
{
    return sizeof(C) == 8 && LP && CQPI_FFT32_TWP;
}
__host__ __device__ constexpr int tw_pass_off(int N, int Lstop)
{
    int lg = 0;
    while ((1 << lg) < N) lg++;
    int L = 1, off = 0;
    while (lg > 0 && L < Lstop) {
        const int R = (lg >= 3 && lg != 4) ? 8 : (lg >= 2 ? 4 : 2);
        if (L > 1) off += L * (R - 1);
        L *= R;
        lg -= R == 8 ? 3 : (R == 4 ? 2 : 1);
    }
    return off;
}
// fills the per-pass table dst (tw_pass_off(N, N) entries) from the natural
// table tw[m] = w^m, m < N, with the threads of the workgroup
template <int N, typename C>
__device__ inline void build_pass_twiddles(C *dst, const C *__restrict__ tw)
{
    constexpr int total = tw_pass_off(N, N);
    for (int m = threadIdx.x; m < total; m += blockDim.x) {
        int lg = 0;
        while ((1 << lg) < N) lg++;
        int L = 1, off = 0, idx = 0;
        while (lg > 0) {
            const int R = (lg >= 3 && lg != 4) ? 8 : (lg >= 2 ? 4 : 2);
            if (L > 1) {
                const int sz = L * (R - 1);
                if (m < off + sz) {
                    const int u = m - off, j = u / (R - 1), r = u - j * (R - 1) + 1;
                    idx = j * r * (N / (L * R));
                    break;
                }
                off += sz;
            }
            L *= R;
            lg -= R == 8 ? 3 : (R == 4 ? 2 : 1);
        }
        dst[m] = tw[idx];
    }
}

template <int N, int L, int R, int LN, bool LP, typename C>
__device__ inline void stockham_pass(C *X, const C *__restrict__ tw, int ll, int sign)
{
    constexpr int NB = N / R;
    constexpr int BPL = (NB + LN - 1) / LN;
    C v[BPL][R];
#pragma unroll
    for (int b = 0; b < BPL; b++) {
        const int k = ll + LN * b;
        if (k < NB)
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = X[pidx<C, LP>(k + r * NB)];
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < BPL; b++) {
        const int k = ll + LN * b;
        if (k < NB) {
            const int j = k & (L - 1);
            if constexpr (L > 1) {
                if constexpr (pass_twiddles<C, LP>()) {
                    constexpr int toff = tw_pass_off(N, L);
                    const C *tp = tw + toff + j * (R - 1);
#pragma unroll
                    for (int r = 1; r < R; r++) v[b][r] = cmul(v[b][r], tp[r - 1]);
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++) v[b][r] = cmul(v[b][r], tw[j * r * (N / (L * R))]);
                }
            }
            dft_small<R>(v[b], sign);
            const int base = (k - j) * R + j;
#pragma unroll
            for (int q = 0; q < R; q++) X[pidx<C, LP>(base + q * L)] = v[b][q];
        }
    }
    wave_sync();
}

template <int N, int L, int LGREM, int LN, bool LP, typename C>
__device__ inline void stockham_rest(C *X, const C *__restrict__ tw, int ll, int sign)
{
    if constexpr (LGREM > 0) {
        constexpr int R = stockham_radix(LGREM);
        stockham_pass<N, L, R, LN, LP>(X, tw, ll, sign);
        stockham_rest<N, L * R, LGREM - ilog2c(R), LN, LP>(X, tw, ll, sign);
    }
}

// transform the padded line X (N points, C = double2 or float2, twiddle table
// of the same type: the natural table, or for padded FP32 lines the per-pass
// table of build_pass_twiddles) with lanes ll in [0, stockham_lanes(N))
template <int N, bool LP = true, typename C>
__device__ inline void wave_fft(C *X, const C *__restrict__ tw, int ll, int sign)
{
    constexpr int LN = stockham_lanes(N);
    constexpr int lg = ilog2n(N);
    stockham_rest<N, 1, lg, LN, LP>(X, tw, ll, sign);
}

}  // namespace cqpi

// Resolvent and k->R transform (FP64).
//
// G0(k,w) = sum_n v_n v_n^H / (w - e_n + i eta)   (src/tightbinding.cc:282-289)
// G0(R)   = (1/K^2) sum_k G0(k) exp(+2 pi i k.R)   (src/qpi.cc:557-578, FFTW_BACKWARD)
//
// The reference forms G0(k) for the whole grid on the host, transforms all
// K x K points of every orbital pair and then copies only the window of R the
// contraction reads (src/qpihipkernel.cc:383-393).  Here the resolvent is
// formed inside the first transform pass, in LDS, and each pass keeps only
// the window R in [-h, h] (side nw = 2h+1):
//   pass 1 (one workgroup per (kx row i, energy, pair chunk)):
//       S[pair][ky] = G0(k)  ->  FFT along ky  ->  F1[e][i][Ry slot][pair]
//   pass 2 (one workgroup per (Ry slot, energy, pair chunk)):
//       S[pair][kx] = F1     ->  FFT along kx  ->  G0w[e][Rx slot][Ry slot][pair] / K^2
// Power-of-two K use an in-LDS radix-2 transform (bit-reversed load, log2 K
// butterfly stages); other K fall back to a direct DFT of the window outputs.
// Twiddles tw[m] = exp(+2 pi i m/K) come from a table computed on the host.
#include "cqpi_common.h"

namespace {

constexpr int kThreads = 256;

__device__ inline int imod(int a, int m)
{
    int r = a % m;
    return r < 0 ? r + m : r;
}

// in-place radix-2 DIT over np lines of length K stored [q*K + idx] (bit-reversed input)
__device__ void fft_lines(double2 *S, int np, int K, int lgK, const double2 *__restrict__ tw)
{
    const int halfK = K >> 1;
    for (int lg = 1; lg <= lgK; lg++) {
        const int half = 1 << (lg - 1);
        const int tstride = K >> lg;
        for (int t = threadIdx.x; t < np * halfK; t += blockDim.x) {
            int q = t / halfK;
            int u = t - q * halfK;
            int grp = u >> (lg - 1);
            int qq = u & (half - 1);
            int i0 = (grp << lg) + qq;
            double2 *L = S + (size_t)q * K;
            double2 a = L[i0];
            double2 b = cmul(L[i0 + half], tw[qq * tstride]);
            L[i0] = cadd(a, b);
            L[i0 + half] = csub(a, b);
        }
        __syncthreads();
    }
}

template <int B>
__global__ void __launch_bounds__(kThreads) k_resolvent_ky(Geom g, const double *__restrict__ eval,
                                                         const double2 *__restrict__ evec,
                                                         const double *__restrict__ omegas,
                                                         double eta, const double2 *__restrict__ tw,
                                                         int np, double2 *__restrict__ F1)
{
    extern __shared__ double2 S[];
    const int K = g.K, i = blockIdx.x, e = blockIdx.y, p0 = blockIdx.z * np;
    const int NP = B * B;
    const int npc = min(np, NP - p0);
    const long KK = (long)K * K;
    const bool pow2 = (K & (K - 1)) == 0;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    const double omega = omegas[e];
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        long k = (long)i * K + j;
        double2 f[B];
#pragma unroll
        for (int nn = 0; nn < B; nn++)
            f[nn] = cinv(make_double2(omega - eval[nn * KK + k], eta));
        int dst = pow2 ? (lgK ? (int)(__brev((unsigned)j) >> (32 - lgK)) : 0) : j;
        if constexpr (B <= 6) {
            double2 V[B * B];
#pragma unroll
            for (int r = 0; r < B * B; r++) V[r] = evec[r * KK + k];
            for (int q = 0; q < npc; q++) {
                int pr = p0 + q, r = pr / B, c = pr - r * B;
                double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                for (int nn = 0; nn < B; nn++) {
                    double2 vr = V[0], vc = V[0];
#pragma unroll
                    for (int x = 0; x < B; x++) {
                        if (x == r) vr = V[x * B + nn];
                        if (x == c) vc = V[x * B + nn];
                    }
                    acc = cadd(acc, cmul(f[nn], cmulc(vr, vc)));
                }
                S[(size_t)q * K + dst] = acc;
            }
        } else {
            for (int q = 0; q < npc; q++) {
                int pr = p0 + q, r = pr / B, c = pr - r * B;
                double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                for (int nn = 0; nn < B; nn++) {
                    double2 vr = evec[(r * B + nn) * KK + k], vc = evec[(c * B + nn) * KK + k];
                    acc = cadd(acc, cmul(f[nn], cmulc(vr, vc)));
                }
                S[(size_t)q * K + dst] = acc;
            }
        }
    }
    __syncthreads();
    const int nw = g.nw, h = g.h;
    double2 *out = F1 + ((size_t)e * K + i) * (size_t)nw * NP;
    if (pow2) {
        fft_lines(S, npc, K, lgK, tw);
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int slot = t / npc, q = t - slot * npc;
            out[(size_t)slot * NP + p0 + q] = S[(size_t)q * K + imod(slot - h, K)];
        }
    } else {
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int slot = t / npc, q = t - slot * npc;
            int Rm = imod(slot - h, K);
            double2 acc = make_double2(0.0, 0.0);
            const double2 *L = S + (size_t)q * K;
            for (int j = 0; j < K; j++) acc = cadd(acc, cmul(L[j], tw[(int)(((long)j * Rm) % K)]));
            out[(size_t)slot * NP + p0 + q] = acc;
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_transform_kx(Geom g, const double2 *__restrict__ F1,
                                                         const double2 *__restrict__ tw, int np,
                                                         double2 *__restrict__ G0w)
{
    extern __shared__ double2 S[];
    const int K = g.K, sy = blockIdx.x, e = blockIdx.y, NP = g.B * g.B, p0 = blockIdx.z * np;
    const int npc = min(np, NP - p0);
    const int nw = g.nw, h = g.h;
    const bool pow2 = (K & (K - 1)) == 0;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    const double2 *in = F1 + (size_t)e * K * nw * NP;
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        int i = t / npc, q = t - i * npc;
        int dst = pow2 ? (lgK ? (int)(__brev((unsigned)i) >> (32 - lgK)) : 0) : i;
        S[(size_t)q * K + dst] = in[((size_t)i * nw + sy) * NP + p0 + q];
    }
    __syncthreads();
    const double inv = 1.0 / ((double)K * (double)K);
    double2 *out = G0w + (size_t)e * nw * nw * NP;
    if (pow2) {
        fft_lines(S, npc, K, lgK, tw);
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            out[((size_t)sx * nw + sy) * NP + p0 + q] = cscale(inv, S[(size_t)q * K + imod(sx - h, K)]);
        }
    } else {
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            int Rm = imod(sx - h, K);
            double2 acc = make_double2(0.0, 0.0);
            const double2 *L = S + (size_t)q * K;
            for (int i = 0; i < K; i++) acc = cadd(acc, cmul(L[i], tw[(int)(((long)i * Rm) % K)]));
            out[((size_t)sx * nw + sy) * NP + p0 + q] = cscale(inv, acc);
        }
    }
}

constexpr size_t kLdsBudget = 80 * 1024;

int pairs_per_chunk(int K, int NP)
{
    int np = (int)(kLdsBudget / (sizeof(double2) * (size_t)K));
    if (np < 1) np = 1;
    return np < NP ? np : NP;
}

template <int B>
int launch_res_t(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                 double eta, const double2 *tw, double2 *F1, hipStream_t s)
{
    int NP = B * B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    size_t lds = sizeof(double2) * (size_t)np * g.K;
    hipLaunchKernelGGL(k_resolvent_ky<B>, dim3(g.K, E, nch), dim3(kThreads), lds, s, g, eval, evec,
                       omegas, eta, tw, np, F1);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace

namespace cqpi {

int launch_resolvent_ky(const Geom &g, const double *eval, const double2 *evec, const double *omegas,
                        int E, double eta, const double2 *tw, double2 *F1, hipStream_t s)
{
    if ((size_t)g.K * sizeof(double2) > kLdsBudget) return CQPI_E_ARG;
    switch (g.B) {
#define CASE(b) \
    case b: return launch_res_t<b>(g, eval, evec, omegas, E, eta, tw, F1, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

int launch_transform_kx(const Geom &g, const double2 *F1, int E, const double2 *tw, double2 *G0w,
                        hipStream_t s)
{
    if ((size_t)g.K * sizeof(double2) > kLdsBudget) return CQPI_E_ARG;
    int NP = g.B * g.B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    size_t lds = sizeof(double2) * (size_t)np * g.K;
    hipLaunchKernelGGL(k_transform_kx, dim3(g.nw, E, nch), dim3(kThreads), lds, s, g, F1, tw, np, G0w);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

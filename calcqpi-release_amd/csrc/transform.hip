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
//   pass 1 (one workgroup per kx row i and batch of energies; the row's eigen
//          data is read once for the whole batch):
//       S[pair][ky] = G0(k)  ->  FFT along ky  ->  F1[e][i][Ry slot][pair]
//   pass 2 (one workgroup per (Ry slot, energy, pair chunk)):
//       S[pair][kx] = F1     ->  FFT along kx  ->  G0w[e][Rx slot][Ry slot][pair] / K^2
//       and, for the mixed-precision contraction, the fp16 hi/lo planes.
// Power-of-two K use an in-LDS decimation-in-time transform (bit-reversed
// load, radix-4 stages plus one radix-2 stage when log2 K is odd); other K
// fall back to a direct DFT of the window outputs.  Twiddles
// tw[m] = exp(+2 pi i m/K) come from a table computed on the host.
#include "cqpi_common.h"
#include "cqpi_fft.h"
#include "mixed.h"

namespace {

using cqpi::brev;
using cqpi::dft_point;
using cqpi::imod;

constexpr int kThreads = 256;

__device__ inline void fft_lines(double2 *S, int np, int K, int lgK, const double2 *__restrict__ tw)
{
    cqpi::fft_lines(S, np, K, lgK, tw, +1);
}

template <int B, int JPT>
__global__ void __launch_bounds__(kThreads) k_resolvent_ky(Geom g, const double *__restrict__ eval,
                                                         const double2 *__restrict__ evec,
                                                         const double *__restrict__ omegas, int E,
                                                         int EB, double eta, const double2 *__restrict__ tw,
                                                         int np, double2 *__restrict__ F1)
{
    extern __shared__ double2 S[];
    constexpr int NP = B * B;
    // eigen data of the thread's JPT k-points kept in registers across the energy loop
    constexpr bool kRegs = B <= 4 && JPT > 0;
    constexpr int kMaxJ = JPT > 0 ? JPT : 1;
    const int K = g.K, i = blockIdx.x, p0 = blockIdx.z * np;
    const int e0 = blockIdx.y * EB, e1 = min(E, e0 + EB);
    const int npc = min(np, NP - p0);
    const long KK = (long)K * K;
    const bool pow2 = (K & (K - 1)) == 0;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    const int nw = g.nw, h = g.h;
    double2 *Tw = S + (size_t)np * K;  // twiddles w^m, m < K/2, staged in LDS
    if (pow2)
        for (int m = threadIdx.x; m < (K >> 1); m += kThreads) Tw[m] = tw[m];

    double ev_r[kRegs ? kMaxJ : 1][B];
    double2 vv_r[kRegs ? kMaxJ : 1][kRegs ? NP : 1];
    const bool regs = kRegs && K <= kMaxJ * kThreads;
    if (regs) {
#pragma unroll
        for (int jj = 0; jj < (kRegs ? kMaxJ : 1); jj++) {
            int j = threadIdx.x + jj * kThreads;
            if (j < K) {
                long k = (long)i * K + j;
#pragma unroll
                for (int nn = 0; nn < B; nn++) ev_r[jj][nn] = eval[nn * KK + k];
                if constexpr (kRegs)
#pragma unroll
                    for (int r = 0; r < NP; r++) vv_r[jj][r] = evec[r * KK + k];
            }
        }
    }
    for (int e = e0; e < e1; e++) {
        const double omega = omegas[e];
        if (regs) {
            if constexpr (kRegs) {
#pragma unroll
                for (int jj = 0; jj < kMaxJ; jj++) {
                    int j = threadIdx.x + jj * kThreads;
                    if (j >= K) break;
                    double2 f[B];
#pragma unroll
                    for (int nn = 0; nn < B; nn++) f[nn] = cinv(make_double2(omega - ev_r[jj][nn], eta));
                    int dst = pow2 ? brev(j, lgK) : j;
                    for (int q = 0; q < npc; q++) {
                        int pr = p0 + q, r = pr / B, c = pr - r * B;
                        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                        for (int nn = 0; nn < B; nn++) {
                            double2 vr = vv_r[jj][0], vc = vv_r[jj][0];
#pragma unroll
                            for (int x = 0; x < B; x++) {
                                if (x == r) vr = vv_r[jj][x * B + nn];
                                if (x == c) vc = vv_r[jj][x * B + nn];
                            }
                            acc = cadd(acc, cmul(f[nn], cmulc(vr, vc)));
                        }
                        S[(size_t)q * K + dst] = acc;
                    }
                }
            }
        } else {
            for (int j = threadIdx.x; j < K; j += kThreads) {
                long k = (long)i * K + j;
                double2 f[B];
#pragma unroll
                for (int nn = 0; nn < B; nn++) f[nn] = cinv(make_double2(omega - eval[nn * KK + k], eta));
                int dst = pow2 ? brev(j, lgK) : j;
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
        double2 *out = F1 + ((size_t)e * K + i) * (size_t)nw * NP;
        if (pow2) {
            fft_lines(S, npc, K, lgK, Tw);
            for (int t = threadIdx.x; t < nw * npc; t += kThreads) {
                int slot = t / npc, q = t - slot * npc;
                out[(size_t)slot * NP + p0 + q] = S[(size_t)q * K + imod(slot - h, K)];
            }
        } else {
            for (int t = threadIdx.x; t < nw * npc; t += kThreads) {
                int slot = t / npc, q = t - slot * npc;
                out[(size_t)slot * NP + p0 + q] = dft_point(S + (size_t)q * K, K, imod(slot - h, K), tw);
            }
        }
        __syncthreads();  // S is rewritten for the next energy
    }
}

__global__ void __launch_bounds__(kThreads) k_transform_kx(Geom g, MixGeom mg, const double2 *__restrict__ F1,
                                                         const double2 *__restrict__ tw, int np, double sG,
                                                         double2 *__restrict__ G0w, _Float16 *__restrict__ P)
{
    extern __shared__ double2 S[];
    const int K = g.K, sy = blockIdx.x, e = blockIdx.y, B = g.B, NP = B * B, p0 = blockIdx.z * np;
    const int npc = min(np, NP - p0);
    const int nw = g.nw, h = g.h;
    const bool pow2 = (K & (K - 1)) == 0;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    const double2 *in = F1 + (size_t)e * K * nw * NP;
    double2 *Tw = S + (size_t)np * K;  // twiddles w^m, m < K/2, staged in LDS
    if (pow2)
        for (int m = threadIdx.x; m < (K >> 1); m += blockDim.x) Tw[m] = tw[m];
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        int i = t / npc, q = t - i * npc;
        int dst = pow2 ? brev(i, lgK) : i;
        S[(size_t)q * K + dst] = in[((size_t)i * nw + sy) * NP + p0 + q];
    }
    __syncthreads();
    const double inv = 1.0 / ((double)K * (double)K);
    double2 *out = G0w + (size_t)e * nw * nw * NP;
    if (pow2) {
        fft_lines(S, npc, K, lgK, Tw);
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            out[((size_t)sx * nw + sy) * NP + p0 + q] = cscale(inv, S[(size_t)q * K + imod(sx - h, K)]);
        }
    } else {
        // direct DFT: transform the needed window points in place of the line
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            out[((size_t)sx * nw + sy) * NP + p0 + q] =
                cscale(inv, dft_point(S + (size_t)q * K, K, imod(sx - h, K), tw));
        }
        if (!P) return;
        // the planes below read S at window points: make S hold them
        __syncthreads();
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            S[(size_t)q * K + imod(sx - h, K)] = make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            int sx = t / npc, q = t - sx * npc;
            S[(size_t)q * K + imod(sx - h, K)] = out[((size_t)sx * nw + sy) * NP + p0 + q];
        }
        __syncthreads();
    }
    if (!P) return;
    // fp16 hi/lo planes for the mixed contraction (rows along x):
    //   a planes: row sy,      column sx  <- G0(sx, sy)[o][p]
    //   b planes: row 2h - sy, column sx  <- G0(2h - sx, sy)[p][o]   (= G0(-R)^T)
    const size_t plane = (size_t)mg.Ly * nw;
    const double s = sG * (pow2 ? inv : 1.0);
    _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    for (int t = threadIdx.x; t < 2 * npc * 2 * mg.Ly; t += blockDim.x) {
        int idx = t % mg.Ly;
        int rest = t / mg.Ly;
        int ri = rest & 1;
        rest >>= 1;
        int q = rest % npc, ab = rest / npc;
        int pr = p0 + q, r = pr / B, c = pr - r * B;
        int sx = idx - mg.OFF;
        double v = 0.0;
        int row, pl;
        if (ab == 0) {
            row = sy;
            pl = (r * B + c) * 2 + ri;  // o = r, p = c
            if (sx >= 0 && sx < nw) {
                double2 z = S[(size_t)q * K + imod(sx - h, K)];
                v = (ri ? z.y : z.x) * s;
            }
        } else {
            row = 2 * h - sy;
            pl = (c * B + r) * 2 + ri;  // G0[r][c] = G0(-R)^T[c][r]: o = c, p = r
            if (sx >= 0 && sx < nw) {
                double2 z = S[(size_t)q * K + imod(h - sx, K)];
                v = (ri ? z.y : z.x) * s;
            }
        }
        _Float16 hi = (_Float16)v;
        _Float16 lo = (_Float16)((v - (double)hi) * 4096.0);
        size_t o0 = ((size_t)(ab * 2 + 0) * mg.NPL + pl) * plane + (size_t)row * mg.Ly + idx;
        size_t o1 = ((size_t)(ab * 2 + 1) * mg.NPL + pl) * plane + (size_t)row * mg.Ly + idx;
        Pe[o0] = hi;
        Pe[o1] = lo;
    }
}

constexpr size_t kLdsBudget = 80 * 1024;   // largest line buffer a workgroup may take
constexpr size_t kLdsTarget = 32 * 1024;   // preferred: several workgroups per CU

// orbital pairs per workgroup: as many lines as fit the preferred LDS size,
// split evenly over the chunks
int pairs_per_chunk(int K, int NP)
{
    int npmax = (int)(kLdsTarget / (sizeof(double2) * (size_t)K));
    if (npmax < 1) npmax = 1;
    int nch = (NP + npmax - 1) / npmax;
    return (NP + nch - 1) / nch;
}

template <int B>
int launch_res_t(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                 double eta, const double2 *tw, double2 *F1, hipStream_t s)
{
    int NP = B * B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    // energies per workgroup: enough workgroups to fill the chip (>= 4 per CU)
    int rows = g.K * nch;
    int EB = std::max(1, std::min(E, (rows * E) / 1024));
    int nb = (E + EB - 1) / EB;
    size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    dim3 grid(g.K, nb, nch), blk(kThreads);
    // k-points per thread whose eigen data stays in registers (small B only)
    int jpt = (g.K + kThreads - 1) / kThreads;
    if (B <= 4 && jpt == 1)
        hipLaunchKernelGGL((k_resolvent_ky<B, 1>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1);
    else if (B <= 4 && jpt == 2)
        hipLaunchKernelGGL((k_resolvent_ky<B, 2>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1);
    else if (B <= 4 && jpt <= 4)
        hipLaunchKernelGGL((k_resolvent_ky<B, 4>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1);
    else
        hipLaunchKernelGGL((k_resolvent_ky<B, 0>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1);
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

int launch_transform_kx(const Geom &g, const MixGeom *mg, const double2 *F1, int E, const double2 *tw,
                        double sG, double2 *G0w, _Float16 *P, hipStream_t s)
{
    if ((size_t)g.K * sizeof(double2) > kLdsBudget) return CQPI_E_ARG;
    int NP = g.B * g.B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    MixGeom m{};
    if (mg) m = *mg;
    hipLaunchKernelGGL(k_transform_kx, dim3(g.nw, E, nch), dim3(kThreads), lds, s, g, m, F1, tw, np, sG,
                       G0w, mg ? P : nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

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
//   pass 1 (one workgroup per kx row i, batch of energies and chunk of
//          orbital pairs; the row's eigen data is read once for the batch):
//       S[pair][ky] = G0(k)  ->  FFT along ky  ->  F1
//   pass 2 (one workgroup per (Ry slot, energy, pair chunk)):
//       S[pair][kx] = F1     ->  FFT along kx  ->  G0w[e][Rx slot][Ry slot][pair] / K^2
//       and, for the mixed-precision contraction, the fp16 hi/lo planes.
// F1 layout: the direct-DFT and radix-4 paths use [e][Ry slot][i][pair], which
// pass 2 reads as one contiguous block.  The Stockham path uses one block of
// nw * pairs values per (e, kx row i), holding pass 1's chunk c (pairs
// [p0, p0 + npc)) at p0 * nw as [Ry slot][pair - p0], so that each pass-1
// workgroup writes one contiguous run per energy (the slot-major layout
// scattered its writes in 40-byte pieces, 3.7x slower in isolation); pass 2
// gathers a slot's pairs from npc-point pieces that neighbouring slots share.
// Power-of-two K from 64 to 1024 use the register-staged Stockham transform
// (one line per lane group of one wave, radix 8/4/2 passes, no workgroup
// barrier between passes); smaller powers of two use the in-LDS radix-4
// transform; other K a direct DFT of the window outputs.  The twiddle table
// tw[m] = exp(+2 pi i m/K), m < K, is computed on the host.
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
                                                         int np, double2 *__restrict__ F1,
                                                         const double2 *__restrict__ Gk)
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
        if (Gk) {
            // G(k) of energy e from the workspace Gk[e][pair][kx][ky]
            const double2 *gk = Gk + ((size_t)e * NP + p0) * KK + (size_t)i * K;
            for (int t = threadIdx.x; t < npc * K; t += kThreads) {
                const int q = t / K, j = t - q * K;
                S[(size_t)q * K + (pow2 ? brev(j, lgK) : j)] = gk[(size_t)q * KK + j];
            }
        } else if (regs) {
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
        double2 *out = F1 + (size_t)e * nw * K * NP + (size_t)i * NP;
        if (pow2) {
            fft_lines(S, npc, K, lgK, Tw);
            for (int t = threadIdx.x; t < nw * npc; t += kThreads) {
                int slot = t / npc, q = t - slot * npc;
                out[(size_t)slot * K * NP + p0 + q] = S[(size_t)q * K + imod(slot - h, K)];
            }
        } else {
            for (int t = threadIdx.x; t < nw * npc; t += kThreads) {
                int slot = t / npc, q = t - slot * npc;
                out[(size_t)slot * K * NP + p0 + q] = dft_point(S + (size_t)q * K, K, imod(slot - h, K), tw);
            }
        }
        __syncthreads();  // S is rewritten for the next energy
    }
}

typedef _Float16 planes_h8 __attribute__((ext_vector_type(8)));

// fp16 hi/lo planes for the mixed contraction (rows along x), from the
// transformed lines S[q][m] (window point R at m = R mod K), scaled by s:
//   a planes: row sy,      column sx  <- G0(sx, sy)[o][p]
//   b planes: row 2h - sy, column sx  <- G0(2h - sx, sy)[p][o]   (= G0(-R)^T)
// Each item fills 8 consecutive columns, one column group of plane_off (Ly
// is a multiple of 8), and stores them as one 16-byte hi and one 16-byte lo
// vector.
// multiply-high divisors (udiv_magic) of the item decomposition, computed on
// the host: column groups of a plane, pairs of the chunk, bands
struct PlaneDiv {
    unsigned long long ng, npc, b;
};

template <typename C, bool PAD>
__device__ inline void write_planes(const C *S, int K, int sy, int e, int p0, int npc, int B, int nw, int h,
                                    const MixGeom &mg, double s, _Float16 *__restrict__ P, const PlaneDiv &dv)
{
    const size_t plane = (size_t)mg.Ly * nw;
    _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int ng = mg.Ly >> 3;
    const bool pow2 = (K & (K - 1)) == 0;
    for (int t = threadIdx.x; t < 2 * npc * 2 * ng; t += blockDim.x) {
        int rest = udiv(t, dv.ng);
        const int gi = t - rest * ng;
        const int ri = rest & 1;
        rest >>= 1;
        const int ab = udiv(rest, dv.npc), q = rest - ab * npc;
        const int pr = p0 + q, r = udiv(pr, dv.b), c = pr - r * B;
        // a: o = r, p = c;  b: G0[r][c] = G0(-R)^T[c][r], o = c, p = r
        const int row = ab ? 2 * h - sy : sy;
        const int pl = (ab ? (c * B + r) : (r * B + c)) * 2 + ri;
        const C *line = S + (size_t)q * (PAD ? cqpi::padded_len<C>(K) : K);
        // split in the precision of the lines: for FP32 lines (s a power of
        // two) v - hi is exact in FP32, so hi + lo/2^12 keeps 22 bits of v
        using Re = decltype(line[0].x);
        const Re sr = (Re)s;
        planes_h8 hv, lv;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int sx = gi * 8 + u - mg.OFF;
            Re v = 0;
            if (sx >= 0 && sx < nw) {
                const int rr = ab ? h - sx : sx - h;
                const int m = pow2 ? (rr & (K - 1)) : imod(rr, K);
                const C z = line[PAD ? cqpi::pidx<C>(m) : m];
                v = (ri ? z.y : z.x) * sr;
            }
            const _Float16 hi = (_Float16)v;
            hv[u] = hi;
            lv[u] = (_Float16)((v - (Re)hi) * (Re)4096);
        }
        const size_t o = plane_off(row, gi * 8, nw);
        *reinterpret_cast<planes_h8 *>(Pe + ((size_t)(ab * 2 + 0) * mg.NPL + pl) * plane + o) = hv;
        *reinterpret_cast<planes_h8 *>(Pe + ((size_t)(ab * 2 + 1) * mg.NPL + pl) * plane + o) = lv;
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
    const double2 *in = F1 + ((size_t)e * nw + sy) * K * NP;
    double2 *Tw = S + (size_t)np * K;  // twiddles w^m, m < K/2, staged in LDS
    if (pow2)
        for (int m = threadIdx.x; m < (K >> 1); m += blockDim.x) Tw[m] = tw[m];
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        int i = t / npc, q = t - i * npc;
        int dst = pow2 ? brev(i, lgK) : i;
        S[(size_t)q * K + dst] = in[(size_t)i * NP + p0 + q];
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
    const PlaneDiv dv{udiv_magic((unsigned)(mg.Ly >> 3)), udiv_magic((unsigned)npc), udiv_magic((unsigned)B)};
    write_planes<double2, false>(S, K, sy, e, p0, npc, B, nw, h, mg, sG * (pow2 ? inv : 1.0), P, dv);
}

// ---------------------------------------------------------------------------
// Stockham path (power-of-two K = N, 64 <= N <= 1024).  C is the complex type
// of the lines, F1 and the twiddles: double2 (FP64 mode) or float2 (the
// reduced-precision modes: the resolvent factor 1/(w - e + i eta) is formed in
// FP64 from the FP64 eigenvalues and then rounded).
// ---------------------------------------------------------------------------
constexpr int kMinWaveK = 64, kMaxWaveK = 1024;
constexpr int kMaxRegB = 6;   // band counts whose per-k eigen data stays in registers
constexpr int kMaxLines = 16; // orbital pairs (lines) per workgroup
constexpr int kMaxSymPairs = 64;  // b^2 of the paired-row pass 2 (b <= 8)
constexpr int kWaveMaxThreads = 512;
// FP32 lines take up to 16 waves a workgroup, so that every orbital pair of
// up to 4 bands is one line of one chunk (3 bands: 9 waves for 9 lines; with
// 8 waves the pairs were split 5 + 4 over two chunks, each re-reading the
// eigen data, and half the waves idled in the line transforms); FP64 lines
// keep 8 waves (9 FP64 lines do not leave room for two workgroups a CU)
#ifndef CQPI_WAVE32_MAX
#define CQPI_WAVE32_MAX 1024
#endif
constexpr int kWaveMaxThreads32 = CQPI_WAVE32_MAX;
__host__ __device__ constexpr int wave_max_threads(size_t csize)
{
    return csize == 8 ? kWaveMaxThreads32 : kWaveMaxThreads;
}
// the first pass pads its FP32 lines (and takes per-pass twiddles) for
// K <= 512; at K = 1024 (one energy per workgroup, lines read from the G(k)
// workspace) the unpadded layout measured faster: 0.78 against 1.0 ms per
// energy for 10 bands
__host__ __device__ constexpr bool res_line_padded(int N) { return N <= 512; }
// LDS slots of one line of the first pass
inline int res_line_len(int N, size_t csize)
{
    return csize == 16 ? cqpi::padded_len<double2>(N)
                       : (res_line_padded(N) ? cqpi::padded_len<float2, true>(N) : cqpi::padded_len<float2, false>(N));
}
// FP32 pass-1 workgroups are held to 80 VGPRs a lane, so that 6 waves share
// a SIMD (three workgroups of 512 per CU): measured 20% faster than the
// compiler's 94 VGPRs (5 waves, two workgroups); FP64 ones keep theirs
constexpr int kResWaves32 = 6;

// resolvent factor 1/(d + i eta), d = w - e formed in FP64 by the caller
// (it cancels); FP64 result for double2 lines, FP32 (one reciprocal) for float2
template <typename C>
__device__ inline C res_factor(double d, double eta)
{
    if constexpr (sizeof(C) == 16) {
        const double2 f = cinv(make_double2(d, eta));
        C r;
        r.x = f.x;
        r.y = f.y;
        return r;
    } else {
        const float x = (float)d, y = (float)eta;
        const float inv = __builtin_amdgcn_rcpf(x * x + y * y);
        C r;
        r.x = x * inv;
        r.y = -y * inv;
        return r;
    }
}

template <typename C>
__device__ inline C to_c(double2 z)
{
    C r;
    r.x = z.x;
    r.y = z.y;
    return r;
}

// pass 1: block = lane groups of stockham_lanes(N) lanes, one line (orbital
// pair of the chunk) per group.  0 < B <= kMaxRegB: the eigen data of the
// thread's k-point (N <= blockDim) stays in registers for the whole energy
// batch; otherwise (B == 0: band count g.B) it is read per energy.
template <int B, int N>
constexpr bool res_regs() { return B > 0 && B <= kMaxRegB && N <= kWaveMaxThreads; }

// G0(k) of every orbital pair of one k-point from its eigen data held in
// registers: P[n] = v_n v_n^H in its upper triangle (diagonal real), so that
// G_rc = sum_n f_n P_rc and G_cr = sum_n f_n conj(P_rc) share their products
template <typename C, int B>
struct EigenPairs {
    using R = decltype(C{}.x);
    R d[B][B];                                // d[n][r] = |v_rn|^2
    C o[B][B * (B - 1) / 2 > 0 ? B * (B - 1) / 2 : 1];  // o[n][upper (r, c)] = v_rn conj(v_cn)
};

template <typename C, int B>
__device__ inline void pairs_from_vectors(EigenPairs<C, B> &P, const double2 *v)  // v[r*B + n]
{
    using R = typename EigenPairs<C, B>::R;
#pragma unroll
    for (int n = 0; n < B; n++) {
        int u = 0;
#pragma unroll
        for (int r = 0; r < B; r++) {
            const double2 a = v[r * B + n];
            P.d[n][r] = (R)(a.x * a.x + a.y * a.y);
#pragma unroll
            for (int c = r + 1; c < B; c++, u++) {
                const double2 z = cmulc(a, v[c * B + n]);
                P.o[n][u].x = (R)z.x;
                P.o[n][u].y = (R)z.y;
            }
        }
    }
}

// S[(r*B + c) * NL + pidx(j)] = sum_n f_n P[n]_rc for all B*B pairs, or
// (UP) S[upper_index(r, c) * NL + pidx(j)] for the pairs r <= c only
template <typename C, int B, int NL, bool LP, bool UP = false>
__device__ inline void resolvent_all_pairs(C *S, int j, const EigenPairs<C, B> &P, const C *f)
{
    using R = typename EigenPairs<C, B>::R;
#pragma unroll
    for (int r = 0; r < B; r++) {
        R gx = 0, gy = 0;
#pragma unroll
        for (int n = 0; n < B; n++) {
            gx += f[n].x * P.d[n][r];
            gy += f[n].y * P.d[n][r];
        }
        C gz;
        gz.x = gx;
        gz.y = gy;
        S[(size_t)(UP ? upper_index(r, r, B) : r * B + r) * NL + cqpi::pidx<C, LP>(j)] = gz;
    }
    int u = 0;
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = r + 1; c < B; c++, u++) {
            // f P = (a - b) + i (c' + d), f conj(P) = (a + b) + i (d - c')
            R a = 0, b = 0, cc = 0, d = 0;
#pragma unroll
            for (int n = 0; n < B; n++) {
                a += f[n].x * P.o[n][u].x;
                b += f[n].y * P.o[n][u].y;
                cc += f[n].x * P.o[n][u].y;
                d += f[n].y * P.o[n][u].x;
            }
            C up;
            up.x = a - b;
            up.y = cc + d;
            if constexpr (UP) {
                S[(size_t)upper_index(r, c, B) * NL + cqpi::pidx<C, LP>(j)] = up;
            } else {
                C lo;
                lo.x = a + b;
                lo.y = d - cc;
                S[(size_t)(r * B + c) * NL + cqpi::pidx<C, LP>(j)] = up;
                S[(size_t)(c * B + r) * NL + cqpi::pidx<C, LP>(j)] = lo;
            }
        }
}

// waves per SIMD pass 1 is compiled for: FP32 lines 6 (80 VGPRs), but 5 for
// the 9-wave workgroups of one chunk of every pair (two of them fill a CU
// either way, and 96 VGPRs hold the pair products without spilling) and 4
// for 4 and 5 bands (their eigenvectors alone take 32 and 50 VGPRs)
// (UPT: the 6 pairs r <= c of 3 bands under the pair symmetry, 8-wave
// workgroups at 6 waves per SIMD: three a CU)
template <typename C, bool ALLP, int B, bool UPT = false>
constexpr int res_min_waves() { return sizeof(C) == 8 ? (UPT ? 6 : (ALLP ? 5 : (B >= 4 ? 4 : kResWaves32))) : 1; }

// UPT: one chunk of the pairs r <= c only (g.ts), resolved at compile time
template <typename C, int B, int N, bool ALLP = false, bool UPT = false>
__global__ void __launch_bounds__(wave_max_threads(sizeof(C)), (res_min_waves<C, ALLP, B, UPT>())) k_res_ky_s(Geom g, const double *__restrict__ eval,
                                                              const double2 *__restrict__ evec,
                                                              const double *__restrict__ omegas, int E, int EB,
                                                              double eta, const C *__restrict__ tw, int np,
                                                              C *__restrict__ F1, const C *__restrict__ Gk)
{
    extern __shared__ __align__(16) unsigned char smem_s[];
    C *S = reinterpret_cast<C *>(smem_s);
    constexpr int LN = cqpi::stockham_lanes(N);
    constexpr bool LP = res_line_padded(N);            // padded FP32 lines (K <= 512)
    constexpr int NL = cqpi::padded_len<C, LP>(N);  // LDS slots per line
    // FP32 lines read the twiddles from an LDS copy; FP64 ones (which measured
    // slower that way) from the cached global table
    C *Twl = S + (size_t)np * NL;
    if constexpr (cqpi::pass_twiddles<C, LP>())
        cqpi::build_pass_twiddles<N>(Twl, tw);
    else if constexpr (sizeof(C) == 8)
        for (int m = threadIdx.x; m < N; m += blockDim.x) Twl[m] = tw[m];
    const C *Tw = sizeof(C) == 8 ? Twl : tw;
    constexpr bool kRegs = res_regs<B, N>();  // blockDim >= N (launcher)
    const int nb = B > 0 ? B : g.B, NP = nb * nb, NT = transform_pairs(g);
    const int i = blockIdx.x, p0 = blockIdx.z * np, npc = min(np, NT - p0);
    const int e0 = blockIdx.y * EB, e1 = min(E, e0 + EB);
    const long KK = (long)N * N;
    const int nw = g.nw, h = g.h;
    const int grp = threadIdx.x / LN, ll = threadIdx.x - grp * LN, ngrp = blockDim.x / LN;
    const unsigned long long dm = udiv_magic(npc);

    // ALLP (one chunk of every pair): G0(k) from the Hermitian pair
    // products; a partial chunk from the vectors
    constexpr bool allp = kRegs && ALLP;
    double ev_r[kRegs ? B : 1];
    C vv_r[kRegs && !allp ? B * B : 1];
    EigenPairs<C, allp ? B : 1> pp;
    if constexpr (kRegs) {
        const int j = threadIdx.x;
        if (j < N) {
            const long k = (long)i * N + j;
#pragma unroll
            for (int n = 0; n < B; n++) ev_r[n] = eval[n * KK + k];
            if constexpr (allp) {
                double2 v[B * B];
#pragma unroll
                for (int r = 0; r < B * B; r++) v[r] = evec[r * KK + k];
                pairs_from_vectors<C, B>(pp, v);
            } else {
#pragma unroll
                for (int r = 0; r < B * B; r++) vv_r[r] = to_c<C>(evec[r * KK + k]);
            }
        }
    }
    for (int e = e0; e < e1; e++) {
        const double omega = omegas[e];
        if constexpr (kRegs) {
            const int j = threadIdx.x;
            if (j < N) {
                C f[B];
#pragma unroll
                for (int n = 0; n < B; n++) f[n] = res_factor<C>(omega - ev_r[n], eta);
                if constexpr (allp && UPT) {
                    resolvent_all_pairs<C, B, NL, LP, true>(S, j, pp, f);
                } else if constexpr (allp) {
                    if (g.ts)
                        resolvent_all_pairs<C, B, NL, LP, true>(S, j, pp, f);
                    else
                        resolvent_all_pairs<C, B, NL, LP>(S, j, pp, f);
                } else {
                    // the chunk's pairs p0 .. p0 + npc - 1: every (r, c) unrolled,
                    // so that the eigenvector components are static registers,
                    // with a test on the chunk (uniform over the workgroup);
                    // G_rc = sum_n (f_n v_rn) conj(v_cn).  With g.ts the pairs
                    // are those with r <= c, numbered by upper_index.
                    const bool ts = g.ts;
#pragma unroll
                    for (int r = 0; r < B; r++) {
                        const int rlo = ts ? upper_index(r, r, B) : r * B, rhi = ts ? rlo + B - r : r * B + B;
                        if (rhi <= p0 || rlo >= p0 + npc) continue;
                        C fv[B];
#pragma unroll
                        for (int n = 0; n < B; n++) fv[n] = cmul(f[n], vv_r[r * B + n]);
#pragma unroll
                        for (int c = 0; c < B; c++) {
                            const int pr = ts ? (c >= r ? upper_index(r, c, B) : -1) : r * B + c;
                            if (pr < p0 || pr >= p0 + npc) continue;
                            C acc = cmulc(fv[0], vv_r[c * B]);
#pragma unroll
                            for (int n = 1; n < B; n++) acc = cadd(acc, cmulc(fv[n], vv_r[c * B + n]));
                            S[(size_t)(pr - p0) * NL + cqpi::pidx<C, LP>(j)] = acc;
                        }
                    }
                }
            }
        }
        if constexpr (!kRegs) {
            // G(k) of this energy and the chunk's pairs from the workspace
            // (k_gk), one line per pair, read along ky
            const C *gk = Gk + (size_t)e * NP * KK + (size_t)i * N;
            for (int t = threadIdx.x; t < npc * N; t += blockDim.x) {
                const int q = t / N, j = t - q * N;
                int r, c;
                transform_pair_rc(p0 + q, nb, g.ts, r, c);
                S[(size_t)q * NL + cqpi::pidx<C, LP>(j)] = gk[(size_t)(r * nb + c) * KK + j];
            }
        }
        lds_barrier();
        for (int q = grp; q < npc; q += ngrp) cqpi::wave_fft<N, LP>(S + (size_t)q * NL, Tw, ll, +1);
        lds_barrier();
        // the chunk's part of the (e, kx row) block: one contiguous run
        C *out = F1 + ((size_t)e * N + i) * nw * NT + (size_t)p0 * nw;
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            const int slot = udiv(t, dm), q = t - slot * npc;
            out[t] = S[(size_t)q * NL + cqpi::pidx<C, LP>((slot - h) & (N - 1))];
        }
        lds_barrier();  // S is rewritten for the next energy
    }
}

// G(k) of every orbital pair for band counts and grids that the register
// path of pass 1 does not cover: one workgroup per TK consecutive k-points of
// one kx row, their eigen data staged in LDS once for a block of energies,
// thread (kk, r) forms row r, G_rc = sum_n f_n v_rn conj(v_cn) (f_n the
// resolvent factor, formed in FP64 and rounded as in pass 1).  Writes
// Gk[e][pair][k], contiguous along ky for pass 1's line loads.
template <typename C, int B>
constexpr int gk_tile()
{
    // k-points per workgroup: the staged vectors within 48 KB, 8 to 64
    int t = 64;
    while (t > 8 && (size_t)t * B * B * sizeof(C) + (size_t)t * B * sizeof(double) > 48 * 1024) t >>= 1;
    return t;
}

template <typename C, int B>
__global__ void __launch_bounds__((gk_tile<C, B>() * B)) k_gk(int N, const double *__restrict__ eval,
                                                            const double2 *__restrict__ evec,
                                                            const double *__restrict__ omegas, int E, double eta,
                                                            C *__restrict__ Gk)
{
    constexpr int TK = gk_tile<C, B>();
    __shared__ C Vs[B * B][TK];
    __shared__ double Es[B][TK];
    const long KK = (long)N * N;
    const long k0 = (long)blockIdx.y * N + (long)blockIdx.x * TK;
    for (int t = threadIdx.x; t < B * B * TK; t += blockDim.x) {
        const int kk = t % TK, rn = t / TK;
        Vs[rn][kk] = to_c<C>(evec[rn * KK + k0 + kk]);
    }
    for (int t = threadIdx.x; t < B * TK; t += blockDim.x) {
        const int kk = t % TK, n = t / TK;
        Es[n][kk] = eval[n * KK + k0 + kk];
    }
    __syncthreads();
    const int kk = threadIdx.x % TK, r = threadIdx.x / TK;
    for (int e = 0; e < E; e++) {
        const double omega = omegas[e];
        C acc[B];
#pragma unroll
        for (int c = 0; c < B; c++) acc[c] = to_c<C>(make_double2(0.0, 0.0));
#pragma unroll 1
        for (int n = 0; n < B; n++) {
            const C a = cmul(res_factor<C>(omega - Es[n][kk], eta), Vs[r * B + n][kk]);
#pragma unroll
            for (int c = 0; c < B; c++) acc[c] = cadd(acc[c], cmulc(a, Vs[c * B + n][kk]));
        }
        C *out = Gk + ((size_t)e * B * B + (size_t)r * B) * KK + k0 + kk;
#pragma unroll
        for (int c = 0; c < B; c++) out[(size_t)c * KK] = acc[c];
    }
}

template <typename C, int B>
int launch_gk_t(int N, const double *eval, const double2 *evec, const double *omegas, int E, double eta, C *Gk,
                hipStream_t s)
{
    constexpr int TK = gk_tile<C, B>();
    if (N % TK) return CQPI_E_ARG;
    hipLaunchKernelGGL((k_gk<C, B>), dim3(N / TK, N), dim3(TK * B), 0, s, N, eval, evec, omegas, E, eta, Gk);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <typename C>
int launch_gk(int B, int N, const double *eval, const double2 *evec, const double *omegas, int E, double eta, C *Gk,
              hipStream_t s)
{
    switch (B) {
#define CASE(b) \
    case b: return launch_gk_t<C, b>(N, eval, evec, omegas, E, eta, Gk, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

// window slots of transformed pair q: pf = r B + c at R, and (ts, r < c)
// pm = c B + r at -R; pm = -1 where there is no mirrored copy
__device__ inline void kx_pair_slots(int q, int B, int ts, int &pf, int &pm)
{
    pm = -1;
    if (!ts) {
        pf = q;
        return;
    }
    int r, c;
    transform_pair_rc(q, B, ts, r, c);
    pf = r * B + c;
    if (r != c) pm = c * B + r;
}

// pass 2 on the Stockham path: one workgroup per (Ry slot, energy, pair chunk);
// the window is written in FP64 whatever C is.  FP32 lines are compiled for 7
// waves per SIMD (72 VGPRs: three 9-wave workgroups a CU; 6 % faster than the
// compiler's 78 VGPRs and two workgroups; 8 waves spilled)
template <typename C, int N>
__global__ void __launch_bounds__(N >= 1024 ? 512 : 1024, (sizeof(C) == 8 ? 7 : 1)) k_kx_s(Geom g, MixGeom mg, const C *__restrict__ F1,
                                                                 const C *__restrict__ tw, int np, int np1, int rwin,
                                                                 double sG, double2 *__restrict__ G0w,
                                                                 _Float16 *__restrict__ P, PlaneDiv dv,
                                                                 float2 *__restrict__ W32)
{
    extern __shared__ __align__(16) unsigned char smem_s[];
    C *S = reinterpret_cast<C *>(smem_s);
    constexpr int LN = cqpi::stockham_lanes(N);
    constexpr int NL = cqpi::padded_len<C>(N);  // LDS slots per line
    // FP32 lines read the twiddles from an LDS copy; FP64 ones (which measured
    // slower that way) from the cached global table
    C *Twl = S + (size_t)np * NL;
    if constexpr (cqpi::pass_twiddles<C>())
        cqpi::build_pass_twiddles<N>(Twl, tw);
    else if constexpr (sizeof(C) == 8)
        for (int m = threadIdx.x; m < N; m += blockDim.x) Twl[m] = tw[m];
    const C *Tw = sizeof(C) == 8 ? Twl : tw;
    // neighbouring Ry slots read the same cache lines of F1: keep them on one XCD
    int sy, e, z;
    xcd_block(sy, e, z);
    // NT transformed pairs per (energy, row) of F1 (NP, or with g.ts the
    // pairs r <= c); the window holds all NP
    const int B = g.B, NP = B * B, NT = transform_pairs(g), p0 = z * np;
    const int npc = min(np, NT - p0);
    const int nw = g.nw, h = g.h;
    const int grp = threadIdx.x / LN, ll = threadIdx.x - grp * LN, ngrp = blockDim.x / LN;
    // pair p sits in pass-1 chunk c = p / np1, at c np1 nw + slot npc1(c) + p - c np1
    // of its (e, kx row) block
    const C *in = F1 + (size_t)e * N * nw * NT;
    if (np1 == NT && npc == NT) {
        // one pass-1 chunk: the slot's pairs are one run of NT values per kx
        // row; a batch of loads is in flight before their LDS stores
        constexpr int U = 8;
        const int tot = N * NT;
        const C *src = in + (size_t)sy * NT;
        for (int t0 = threadIdx.x; t0 < tot; t0 += U * blockDim.x) {
            C v[U];
            // every v[u] loaded (the index clamped), so that the batch stays
            // in registers (a conditionally written array went to scratch in FP64)
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = min(t0 + u * (int)blockDim.x, tot - 1);
                const int i = udiv(t, dv.npc);
                v[u] = src[(size_t)i * nw * NT + (t - i * NT)];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * blockDim.x;
                if (t < tot) {
                    const int i = udiv(t, dv.npc);
                    S[(size_t)(t - i * NT) * NL + cqpi::pidx<C>(i)] = v[u];
                }
            }
        }
    } else {
        const unsigned long long dm = udiv_magic(npc);
        for (int t = threadIdx.x; t < N * npc; t += blockDim.x) {
            const int i = udiv(t, dm), q = t - i * npc, p = p0 + q;
            int c0 = 0;
            while (c0 + np1 <= p) c0 += np1;
            const int npc1 = min(np1, NT - c0);
            S[(size_t)q * NL + cqpi::pidx<C>(i)] = in[(size_t)i * nw * NT + (size_t)c0 * nw + sy * npc1 + (p - c0)];
        }
    }
    lds_barrier();
    for (int q = grp; q < npc; q += ngrp) cqpi::wave_fft<N>(S + (size_t)q * NL, Tw, ll, +1);
    lds_barrier();
    const double inv = 1.0 / ((double)N * (double)N);
    double2 *out = G0w + (size_t)e * nw * nw * NP;
    // the window in FP64: whole, or (rwin < h) only |R| <= rwin.  With g.ts
    // a pair (r, c), r < c, is also G0(-R)_cr: slot (2h - sx, 2h - sy)
    if (abs(sy - h) <= rwin) {
        const int sx0 = h - rwin, nx = 2 * rwin + 1;
        for (int t = threadIdx.x; t < nx * npc; t += blockDim.x) {
            const int sx = sx0 + t / npc, q = t - (sx - sx0) * npc;
            const C z = S[(size_t)q * NL + cqpi::pidx<C>((sx - h) & (N - 1))];
            const double2 v = make_double2(inv * (double)z.x, inv * (double)z.y);
            int pf, pm;
            kx_pair_slots(p0 + q, B, g.ts, pf, pm);
            out[((size_t)sx * nw + sy) * NP + pf] = v;
            if (pm >= 0) out[((size_t)(2 * h - sx) * nw + (2 * h - sy)) * NP + pm] = v;
        }
    }
    if (P) {
        PlaneDiv d2 = dv;
        if (npc != np) d2.npc = udiv_magic((unsigned)npc);  // the last, shorter chunk
        write_planes<C, true>(S, N, sy, e, p0, npc, B, nw, h, mg, sG * inv, P, d2);
    }
    if (W32) {
        // the whole window row sy in FP32, W32[e][sy][sx][pair] (one contiguous
        // run); with g.ts the pairs r < c also go to row 2h - sy as (c, r)
        const float invf = (float)inv;
        const unsigned long long mq = udiv_magic((unsigned)npc);
        float2 *ow = W32 + ((size_t)e * nw + sy) * nw * NP;
        float2 *om = W32 + ((size_t)e * nw + (2 * h - sy)) * nw * NP;
        for (int t = threadIdx.x; t < nw * npc; t += blockDim.x) {
            const int sx = udiv(t, mq), q = t - sx * npc;
            const C z = S[(size_t)q * NL + cqpi::pidx<C>((sx - h) & (N - 1))];
            const float2 v = make_float2((float)z.x * invf, (float)z.y * invf);
            int pf, pm;
            kx_pair_slots(p0 + q, B, g.ts, pf, pm);
            ow[(size_t)sx * NP + pf] = v;
            if (pm >= 0) om[(size_t)(2 * h - sx) * NP + pm] = v;
        }
    }
}

// pass 2 with the pair symmetry (g.ts), one workgroup per mirror pair of Ry
// slots {sy, 2h - sy} (sy <= h) and energy: the NT lines r <= c of both rows
// are transformed together (NT for the centre row sy = h), so that every
// window record of the two rows, all b^2 pairs of a slot, is written whole
// by one workgroup, G0(R)_cr taken as G0(-R)_rc from the other row at -Rx.
// (Writing the pairs c > r from the workgroup of the mirror row left every
// record split between two workgroups: 4x the window's bytes written in
// FP64.)  F1 holds the NT pairs in one pass-1 chunk; no fp16 planes.
template <typename C, int N>
__global__ void __launch_bounds__(N >= 1024 ? 512 : 1024, (sizeof(C) == 8 ? 6 : 1)) k_kx_sym(Geom g, const C *__restrict__ F1,
                                                                   const C *__restrict__ tw, int rwin,
                                                                   double2 *__restrict__ G0w,
                                                                   float2 *__restrict__ W32)
{
    extern __shared__ __align__(16) unsigned char smem_s[];
    __shared__ int ptab[kMaxSymPairs];
    C *S = reinterpret_cast<C *>(smem_s);
    constexpr int LN = cqpi::stockham_lanes(N);
    constexpr int NL = cqpi::padded_len<C>(N);  // LDS slots per line
    const int B = g.B, NP = B * B, NT = transform_pairs(g);
    C *Twl = S + (size_t)2 * NT * NL;
    if constexpr (cqpi::pass_twiddles<C>())
        cqpi::build_pass_twiddles<N>(Twl, tw);
    else if constexpr (sizeof(C) == 8)
        for (int m = threadIdx.x; m < N; m += blockDim.x) Twl[m] = tw[m];
    const C *Tw = sizeof(C) == 8 ? Twl : tw;
    // pair pf = r B + c of a record: line upper_index(min, max), from the
    // mirror row at -Rx where r > c (bit 16)
    for (int pf = threadIdx.x; pf < NP; pf += blockDim.x) {
        const int r = pf / B, c = pf - r * B;
        ptab[pf] = r <= c ? upper_index(r, c, B) : (upper_index(c, r, B) | (1 << 16));
    }
    int sy, e, z;
    xcd_block(sy, e, z);
    const int nw = g.nw, h = g.h, sm = 2 * h - sy;
    const int nrow = sm != sy ? 2 : 1;
    const int grp = threadIdx.x / LN, ll = threadIdx.x - grp * LN, ngrp = blockDim.x / LN;
    const C *in = F1 + (size_t)e * N * nw * NT;
    // lines t NT + q: pair q of row t (t = 0: sy, t = 1: sm); a batch of
    // loads in flight before their LDS stores
    {
        constexpr int U = 8;
        const unsigned long long mq = udiv_magic((unsigned)NT);
        const int tot = N * NT;
        for (int t = 0; t < nrow; t++) {
            const C *src = in + (size_t)(t ? sm : sy) * NT;
            C *dst = S + (size_t)t * NT * NL;
            for (int t0 = threadIdx.x; t0 < tot; t0 += U * blockDim.x) {
                C v[U];
                // every v[u] loaded (the index clamped), so that the batch
                // stays in registers (a conditionally written array went to
                // scratch in FP64)
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = min(t0 + u * (int)blockDim.x, tot - 1);
                    const int i = udiv(k, mq);
                    v[u] = src[(size_t)i * nw * NT + (k - i * NT)];
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = t0 + u * blockDim.x;
                    if (k < tot) {
                        const int i = udiv(k, mq);
                        dst[(size_t)(k - i * NT) * NL + cqpi::pidx<C>(i)] = v[u];
                    }
                }
            }
        }
    }
    lds_barrier();
    for (int q = grp; q < nrow * NT; q += ngrp) cqpi::wave_fft<N>(S + (size_t)q * NL, Tw, ll, +1);
    lds_barrier();
    const double inv = 1.0 / ((double)N * (double)N);
    const unsigned long long mp = udiv_magic((unsigned)NP);
    for (int t = 0; t < nrow; t++) {
        const int y = t ? sm : sy;
        const C *own = S + (size_t)t * NT * NL, *mir = S + (size_t)(nrow - 1 - t) * NT * NL;
        // FP64 window: whole, or (rwin < h) only |R| <= rwin
        if (abs(y - h) <= rwin) {
            const int sx0 = h - rwin, nx = 2 * rwin + 1;
            double2 *out = G0w + ((size_t)e * nw * nw + y) * NP;
            for (int k = threadIdx.x; k < nx * NP; k += blockDim.x) {
                const int ix = udiv(k, mp), pf = k - ix * NP, sx = sx0 + ix;
                const int pt = ptab[pf], q = pt & 0xffff;
                const C zv = (pt >> 16) ? mir[(size_t)q * NL + cqpi::pidx<C>((h - sx) & (N - 1))]
                                        : own[(size_t)q * NL + cqpi::pidx<C>((sx - h) & (N - 1))];
                out[(size_t)sx * nw * NP + pf] = make_double2(inv * (double)zv.x, inv * (double)zv.y);
            }
        }
        if (W32) {
            // the whole window row y in FP32, W32[e][y][sx][pair] (one contiguous run)
            const float invf = (float)inv;
            float2 *ow = W32 + ((size_t)e * nw + y) * nw * NP;
            for (int k = threadIdx.x; k < nw * NP; k += blockDim.x) {
                const int sx = udiv(k, mp), pf = k - sx * NP;
                const int pt = ptab[pf], q = pt & 0xffff;
                const C zv = (pt >> 16) ? mir[(size_t)q * NL + cqpi::pidx<C>((h - sx) & (N - 1))]
                                        : own[(size_t)q * NL + cqpi::pidx<C>((sx - h) & (N - 1))];
                ow[k] = make_float2((float)zv.x * invf, (float)zv.y * invf);
            }
        }
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

// Stockham path: pairs per chunk from the LDS budget (one line of N complex
// values per pair), workgroup = enough lane groups for the chunk's lines and,
// on pass 1, at least one thread per k-point (up to kWaveMaxThreads)
constexpr size_t kWaveLds = 96 * 1024;
// FP32 lines: the lines and the twiddle table of a workgroup within 72 KB, so
// that two workgroups share a CU also at K = 1024 (padded lines of 1153 points)
#ifndef CQPI_LDS32_KB
#define CQPI_LDS32_KB 72
#endif
constexpr size_t kWaveLds32 = CQPI_LDS32_KB * 1024;
// FP64 lines of the paired-row pass 2: up to 12 (3 bands at K = 512) in one
// workgroup, one workgroup a CU
constexpr size_t kSymLds64 = 128 * 1024;
size_t line_lds_budget(int N, size_t csize)
{
    return csize == 16 ? kWaveLds : kWaveLds32 - (size_t)N * csize;
}

int wave_pairs_max(int N, int NP, size_t csize, int maxthreads)
{
    const int groups = maxthreads / cqpi::stockham_lanes(N);
    const int lines = std::max(1, std::min(std::min(groups, kMaxLines), (int)(line_lds_budget(N, csize) / (csize * (size_t)res_line_len(N, csize)))));
    const int nch = (NP + lines - 1) / lines;
    return (NP + nch - 1) / nch;
}

// FP32 lines: the larger workgroups only where they leave fewer idle lane
// groups (chunks x lines per chunk), e.g. 9 pairs as one chunk of 9, not 5 + 4
int wave_pairs(int N, int NP, size_t csize)
{
    const int np = wave_pairs_max(N, NP, csize, kWaveMaxThreads);
    if (csize == 16) return np;
    const int npw = wave_pairs_max(N, NP, csize, kWaveMaxThreads32);
    const int slots = (NP + np - 1) / np * np, slotsw = (NP + npw - 1) / npw * npw;
    return slotsw < slots ? npw : np;
}

int wave_threads(int N, int np, bool pass1, size_t csize)
{
    int t = np * cqpi::stockham_lanes(N);
    if (pass1) t = std::max(t, std::min(N, kWaveMaxThreads));
    t = (t + 63) / 64 * 64;
    return std::min(t, wave_max_threads(csize));
}

// energies per G(k) block of the workspace path
constexpr size_t kGkBudget = (size_t)4 << 30;
int gk_block(const Geom &g, int E, size_t csize)
{
    const size_t per_e = (size_t)g.B * g.B * g.K * g.K * csize;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)E, kGkBudget / per_e));
}

template <typename C, int B, int N>
int launch_res_s(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E, double eta,
                 const C *tw, C *F1, void *work, size_t work_bytes, hipStream_t s)
{
    const int NP = g.B * g.B, NT = transform_pairs(g);
    const int np = wave_pairs(N, NT, sizeof(C));
    const int nch = (NT + np - 1) / np;
    const int thr = wave_threads(N, np, true, sizeof(C));
    if (thr / cqpi::stockham_lanes(N) < 1) return CQPI_E_ARG;
    const size_t lds = sizeof(C) * ((size_t)np * res_line_len(N, sizeof(C)) + (sizeof(C) == 8 ? N : 0));
    if constexpr (!res_regs<B, N>()) {
        // G(k) through the workspace, a block of energies at a time
        const size_t per_e = (size_t)NP * N * N * sizeof(C);
        if (!work || work_bytes < per_e) return CQPI_E_ARG;
        const int EBa = (int)std::min<size_t>((size_t)E, work_bytes / per_e);
        C *Gk = static_cast<C *>(work);
        for (int ea = 0; ea < E; ea += EBa) {
            const int eb = std::min(EBa, E - ea);
            const int r = launch_gk<C>(g.B, N, eval, evec, omegas + ea, eb, eta, Gk, s);
            if (r) return r;
            hipLaunchKernelGGL((k_res_ky_s<C, B, N>), dim3(N, eb, nch), dim3(thr), lds, s, g, eval, evec,
                               omegas + ea, eb, 1, eta, tw, np, F1 + (size_t)ea * N * g.nw * NT, (const C *)Gk);
            if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
        }
        return 0;
    }
    // energies per workgroup: FP32 lines, with several workgroups resident
    // per CU, want ~16 workgroups per CU (a short tail); FP64 ones ~2; the
    // 9-wave workgroups of one chunk of every pair ~4 (two blocks of energies
    // for 512 rows: 0.265 ms against 0.29 with four, 3 orbitals, K = 512).
    // The energy blocks are made even (51 energies: 4 blocks of 13, not 12,12,12,12,3)
    const int rows = N * nch;
    const int wg_target = sizeof(C) == 8 ? (nch == 1 && NT > 4 ? 1024 : 4096) : 512;
    const int nb_want = std::max(1, std::min(E, (wg_target + rows / 2) / rows));
    const int EB = g.eb > 0 ? std::min(E, g.eb) : (E + nb_want - 1) / nb_want;
    const int nb = (E + EB - 1) / EB;
    // one chunk of every pair: G0(k) from the Hermitian pair products, except
    // for the 10 and 15 pairs r <= c of 4 and 5 bands, whose products do not
    // fit the registers (the pair products of 5 bands spilled: 0.74 -> 1.40 ms
    // for 40 energies, K = 512)
    if constexpr (B == 3 && sizeof(C) == 8) {
        if (nch == 1 && g.ts) {
            hipLaunchKernelGGL((k_res_ky_s<C, B, N, true, true>), dim3(N, nb, nch), dim3(thr), lds, s, g, eval, evec,
                               omegas, E, EB, eta, tw, np, F1, (const C *)nullptr);
            return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
        }
    }
    if (nch == 1 && (B <= 3 || !g.ts))
        hipLaunchKernelGGL((k_res_ky_s<C, B, N, true>), dim3(N, nb, nch), dim3(thr), lds, s, g, eval, evec, omegas, E,
                           EB, eta, tw, np, F1, (const C *)nullptr);
    else
        hipLaunchKernelGGL((k_res_ky_s<C, B, N>), dim3(N, nb, nch), dim3(thr), lds, s, g, eval, evec, omegas, E, EB,
                           eta, tw, np, F1, (const C *)nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

// pass 1 alone, from G(k) of every energy already in the workspace layout
// Gk[e][pair][kx][ky] (the surface Green's function writes it)
template <typename C, int N>
int launch_res_gk(const Geom &g, const C *Gk, const double *omegas, int E, const C *tw, C *F1, hipStream_t s)
{
    const int NT = transform_pairs(g);
    const int np = wave_pairs(N, NT, sizeof(C));
    const int nch = (NT + np - 1) / np;
    const int thr = wave_threads(N, np, true, sizeof(C));
    if (thr / cqpi::stockham_lanes(N) < 1) return CQPI_E_ARG;
    const size_t lds = sizeof(C) * ((size_t)np * res_line_len(N, sizeof(C)) + (sizeof(C) == 8 ? N : 0));
    hipLaunchKernelGGL((k_res_ky_s<C, 0, N>), dim3(N, E, nch), dim3(thr), lds, s, g, (const double *)nullptr,
                       (const double2 *)nullptr, omegas, E, 1, 0.0, tw, np, F1, Gk);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <typename C, int N>
int launch_res_sb(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E, double eta,
                  const C *tw, C *F1, void *work, size_t work_bytes, hipStream_t s)
{
    if (N > kWaveMaxThreads) return launch_res_s<C, 0, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    switch (g.B) {
    case 1: return launch_res_s<C, 1, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 2: return launch_res_s<C, 2, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 3: return launch_res_s<C, 3, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 4: return launch_res_s<C, 4, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 5: return launch_res_s<C, 5, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 6: return launch_res_s<C, 6, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    default: return launch_res_s<C, 0, N>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    }
}

// pass 2 reads whole [kx][pair] rows of F1: one chunk of all pairs whenever
// they fit the LDS budget (a split would fetch each row's cache lines twice)
int kx_pairs(int N, int NP, size_t csize)
{
    const int lines = std::max(1, std::min(1024 / cqpi::stockham_lanes(N), (int)(line_lds_budget(N, csize) / (csize * (size_t)(csize == 16 ? cqpi::padded_len<double2>(N) : cqpi::padded_len<float2>(N))))));
    const int nch = (NP + lines - 1) / lines;
    return (NP + nch - 1) / nch;
}

template <typename C, int N>
int launch_kx_s(const Geom &g, const MixGeom &m, const C *F1, int E, const C *tw, double sG, double2 *G0w,
                _Float16 *P, int rwin, hipStream_t s, float2 *W32 = nullptr)
{
    // pairs with planes are only written without the pair symmetry
    if (g.ts && P) return CQPI_E_ARG;
    const int NT = transform_pairs(g), np = kx_pairs(N, NT, sizeof(C)), nch = (NT + np - 1) / np;
    const int np1 = wave_pairs(N, NT, sizeof(C));  // pass 1's chunks fix the F1 layout
    rwin = std::min(std::max(rwin, 0), g.h);
    if (g.ts && np1 == NT && g.B * g.B <= kMaxSymPairs) {
        // both rows of a mirror pair in one workgroup where their lines fit
        const int tsym = (2 * NT * cqpi::stockham_lanes(N) + 63) / 64 * 64;
        const size_t lsym = sizeof(C) * ((size_t)2 * NT * cqpi::padded_len<C>(N) + (sizeof(C) == 8 ? N : 0));
        if (tsym <= (N >= 1024 ? 512 : 1024) && lsym <= (sizeof(C) == 8 ? kWaveLds32 : kSymLds64)) {
            hipLaunchKernelGGL((k_kx_sym<C, N>), dim3(g.h + 1, E, 1), dim3(tsym), lsym, s, g, F1, tw, rwin, G0w, W32);
            return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
        }
    }
    const int thr = std::min(N >= 1024 ? 512 : 1024, (np * cqpi::stockham_lanes(N) + 63) / 64 * 64);
    const size_t lds = sizeof(C) * ((size_t)np * cqpi::padded_len<C>(N) + (sizeof(C) == 8 ? N : 0));
    const PlaneDiv dv{udiv_magic((unsigned)std::max(m.Ly >> 3, 1)), udiv_magic((unsigned)np), udiv_magic((unsigned)g.B)};
    hipLaunchKernelGGL((k_kx_s<C, N>), dim3(g.nw, E, nch), dim3(thr), lds, s, g, m, F1, tw, np, np1, rwin, sG, G0w,
                       P, dv, W32);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

bool wave_path(int K) { return (K & (K - 1)) == 0 && K >= kMinWaveK && K <= kMaxWaveK; }

template <int B>
int launch_res_t(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                 double eta, const double2 *tw, double2 *F1, hipStream_t s, const double2 *Gk = nullptr)
{
    int NP = B * B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    // energies per workgroup: enough workgroups to fill the chip (>= 4 per CU)
    int rows = g.K * nch;
    int EB = g.eb > 0 ? std::min(E, g.eb) : std::max(1, std::min(E, (rows * E) / 1024));
    int nb = (E + EB - 1) / EB;
    size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    dim3 grid(g.K, nb, nch), blk(kThreads);
    // k-points per thread whose eigen data stays in registers (small B only)
    int jpt = (g.K + kThreads - 1) / kThreads;
    if (Gk)
        hipLaunchKernelGGL((k_resolvent_ky<B, 0>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1, Gk);
    else if (B <= 4 && jpt == 1)
        hipLaunchKernelGGL((k_resolvent_ky<B, 1>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1, Gk);
    else if (B <= 4 && jpt == 2)
        hipLaunchKernelGGL((k_resolvent_ky<B, 2>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1, Gk);
    else if (B <= 4 && jpt <= 4)
        hipLaunchKernelGGL((k_resolvent_ky<B, 4>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1, Gk);
    else
        hipLaunchKernelGGL((k_resolvent_ky<B, 0>), grid, blk, lds, s, g, eval, evec, omegas, E, EB, eta, tw, np, F1, Gk);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace

namespace cqpi {

int launch_resolvent_ky(const Geom &g, const double *eval, const double2 *evec, const double *omegas,
                        int E, double eta, const double2 *tw, double2 *F1, void *work, size_t work_bytes,
                        hipStream_t s)
{
    if (wave_path(g.K)) {
        switch (g.K) {
        case 64: return launch_res_sb<double2, 64>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
        case 128: return launch_res_sb<double2, 128>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
        case 256: return launch_res_sb<double2, 256>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
        case 512: return launch_res_sb<double2, 512>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
        case 1024: return launch_res_sb<double2, 1024>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
        }
    }
    // the LDS-line passes transform every pair
    if ((size_t)g.K * sizeof(double2) > kLdsBudget || g.ts) return CQPI_E_ARG;
    switch (g.B) {
#define CASE(b) \
    case b: return launch_res_t<b>(g, eval, evec, omegas, E, eta, tw, F1, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

int launch_resolvent_ky_gk(const Geom &g, const void *Gk, const double *omegas, int E, const void *tw, void *F1,
                           bool fp32, hipStream_t s)
{
    if (!wave_path(g.K)) {
        // other grids: the LDS-line pass (FP64 only)
        if (fp32 || (size_t)g.K * sizeof(double2) > kLdsBudget || g.ts) return CQPI_E_ARG;
        const double2 *gk = static_cast<const double2 *>(Gk);
        const double2 *t = static_cast<const double2 *>(tw);
        double2 *f = static_cast<double2 *>(F1);
        switch (g.B) {
#define CASE(b) \
    case b: return launch_res_t<b>(g, nullptr, nullptr, omegas, E, 0.0, t, f, s, gk);
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
            CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
        default: return CQPI_E_ARG;
        }
    }
#define GK_CASE(n)                                                                                             \
    case n:                                                                                                    \
        return fp32 ? launch_res_gk<float2, n>(g, static_cast<const float2 *>(Gk), omegas, E,                 \
                                               static_cast<const float2 *>(tw), static_cast<float2 *>(F1), s) \
                    : launch_res_gk<double2, n>(g, static_cast<const double2 *>(Gk), omegas, E,               \
                                                static_cast<const double2 *>(tw), static_cast<double2 *>(F1), s);
    switch (g.K) {
        GK_CASE(64) GK_CASE(128) GK_CASE(256) GK_CASE(512) GK_CASE(1024)
    }
#undef GK_CASE
    return CQPI_E_ARG;
}

bool transform32_supported(const Geom &g) { return wave_path(g.K); }

int launch_resolvent_ky32(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                          double eta, const float2 *tw, float2 *F1, void *work, size_t work_bytes, hipStream_t s)
{
    switch (g.K) {
    case 64: return launch_res_sb<float2, 64>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 128: return launch_res_sb<float2, 128>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 256: return launch_res_sb<float2, 256>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 512: return launch_res_sb<float2, 512>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    case 1024: return launch_res_sb<float2, 1024>(g, eval, evec, omegas, E, eta, tw, F1, work, work_bytes, s);
    default: return CQPI_E_ARG;
    }
}

size_t resolvent_work_bytes(const Geom &g, int E, bool fp32)
{
    if (!wave_path(g.K) || (g.B <= kMaxRegB && g.K <= kWaveMaxThreads)) return 0;
    const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
    return (size_t)gk_block(g, E, cs) * g.B * g.B * g.K * g.K * cs;
}

int launch_transform_kx32(const Geom &g, const MixGeom *mg, const float2 *F1, int E, const float2 *tw, double sG,
                          double2 *G0w, _Float16 *P, int rwin, hipStream_t s, float2 *W32)
{
    MixGeom m{};
    if (mg) m = *mg;
    _Float16 *Pm = mg && !W32 ? P : nullptr;
    switch (g.K) {
    case 64: return launch_kx_s<float2, 64>(g, m, F1, E, tw, sG, G0w, Pm, rwin, s, W32);
    case 128: return launch_kx_s<float2, 128>(g, m, F1, E, tw, sG, G0w, Pm, rwin, s, W32);
    case 256: return launch_kx_s<float2, 256>(g, m, F1, E, tw, sG, G0w, Pm, rwin, s, W32);
    case 512: return launch_kx_s<float2, 512>(g, m, F1, E, tw, sG, G0w, Pm, rwin, s, W32);
    case 1024: return launch_kx_s<float2, 1024>(g, m, F1, E, tw, sG, G0w, Pm, rwin, s, W32);
    default: return CQPI_E_ARG;
    }
}

int launch_transform_kx(const Geom &g, const MixGeom *mg, const double2 *F1, int E, const double2 *tw,
                        double sG, double2 *G0w, _Float16 *P, hipStream_t s)
{
    MixGeom m{};
    if (mg) m = *mg;
    _Float16 *Pm = mg ? P : nullptr;
    if (wave_path(g.K)) {
        switch (g.K) {
        case 64: return launch_kx_s<double2, 64>(g, m, F1, E, tw, sG, G0w, Pm, g.h, s);
        case 128: return launch_kx_s<double2, 128>(g, m, F1, E, tw, sG, G0w, Pm, g.h, s);
        case 256: return launch_kx_s<double2, 256>(g, m, F1, E, tw, sG, G0w, Pm, g.h, s);
        case 512: return launch_kx_s<double2, 512>(g, m, F1, E, tw, sG, G0w, Pm, g.h, s);
        case 1024: return launch_kx_s<double2, 1024>(g, m, F1, E, tw, sG, G0w, Pm, g.h, s);
        }
    }
    if ((size_t)g.K * sizeof(double2) > kLdsBudget || g.ts) return CQPI_E_ARG;
    int NP = g.B * g.B;
    int np = pairs_per_chunk(g.K, NP);
    int nch = (NP + np - 1) / np;
    size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    hipLaunchKernelGGL(k_transform_kx, dim3(g.nw, E, nch), dim3(kThreads), lds, s, g, m, F1, tw, np, sG,
                       G0w, mg ? P : nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

// Surface / bulk Green's function of a layered model by decimation
// (tmatrix::calcsurfacegreensfunction, src/qpi.cc:488-554), FP64.
//
// Per k (fractional k of the grid, kz = 0) the layers form a semi-infinite
// chain with in-plane blocks eps_s (surface layer), eps_b (bulk layer) and
// the coupling alpha to the next layer, beta = alpha^H
// (tightbind::setmatrix with kzdepinplane / kzdepoutofplane,
// src/tightbinding.cc:184-208: in-plane = hops with rz == 0 plus the -/+ efermi
// diagonal, coupling = hops with rz < 0).  Each step, with
// g = (w + i eta - eps_b)^-1:
//   eps_s += alpha g beta;  eps_b += alpha g beta + beta g alpha;
//   alpha <- alpha g alpha; beta <- beta g beta;
// until both Frobenius norms are <= epserr (src/qpi.cc:541); then
// G(k) = (w + i eta - eps_s)^-1 (surface) or with eps_b (bulk).  One thread per
// (k, energy) with the b x b blocks in registers for b <= 4, one 16-lane
// group per (k, energy) with a row per lane for 5 <= b <= 16 (no scratch up
// to 14 bands); the result goes to the G(k)
// workspace layout Gk[e][pair][kx][ky] that the first transform pass reads.
// A run stops a k-point after maxit steps and reports it in stat.
#include "cqpi_common.h"

namespace {

struct SurfSets {
    const DevRGroup *g[3];  // 0: bulk in-plane, 1: surface in-plane, 2: coupling (rz < 0)
    const double2 *t[3];
    int n[3];
};

template <int B>
__device__ inline void build_h(const DevRGroup *__restrict__ groups, int ngroups, const double2 *__restrict__ tmat,
                               const double2 *__restrict__ tw, int i, int j, int K, bool diag, double efermi,
                               int scmax, double2 *H)
{
#pragma unroll(B <= 5 ? 64 : 1)
    for (int r = 0; r < B * B; r++) H[r] = make_double2(0.0, 0.0);
    if (diag)
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < B; r++) H[r * B + r].x = r < scmax ? -efermi : efermi;
    const double kx = (i > (K >> 1)) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
    const double ky = (j > (K >> 1)) ? ((double)j - (double)K) / (double)K : (double)j / (double)K;
    const unsigned kmask = (K & (K - 1)) == 0 ? (unsigned)K - 1u : 0u;
    for (int gi = 0; gi < ngroups; gi++) {
        const DevRGroup G = groups[gi];
        double2 e;
        if (G.integral) {
            const unsigned m = (unsigned)G.irx * (unsigned)i + (unsigned)G.iry * (unsigned)j;
            e = tw[kmask ? (m & kmask) : (m % (unsigned)K)];
        } else {
            double sn, cs;
            sincos(2.0 * M_PI * (G.rx * kx + G.ry * ky), &sn, &cs);
            e = make_double2(cs, sn);
        }
        const double2 *T = tmat + (size_t)gi * B * B;
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < B * B; r++) H[r] = cadd(H[r], cmul(T[r], e));
    }
}

// G = (w + i eta - H)^-1, Gauss-Jordan with partial pivoting; every index
// static (the row exchange by compares)
template <int B>
__device__ inline void resinv(const double2 *H, double om, double eta, double2 *G)
{
    double2 M[B * B];
#pragma unroll(B <= 5 ? 64 : 1)
    for (int r = 0; r < B; r++)
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) {
            M[r * B + c] = make_double2((r == c ? om : 0.0) - H[r * B + c].x, (r == c ? eta : 0.0) - H[r * B + c].y);
            G[r * B + c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
        }
#pragma unroll(B <= 5 ? 64 : 1)
    for (int j = 0; j < B; j++) {
        int piv = j;
        double best = cabs2(M[j * B + j]);
#pragma unroll(B <= 5 ? 64 : 1)
        for (int i = j + 1; i < B; i++)
            if (cabs2(M[i * B + j]) > best) {
                best = cabs2(M[i * B + j]);
                piv = i;
            }
#pragma unroll(B <= 5 ? 64 : 1)
        for (int i = j + 1; i < B; i++)
            if (i == piv)
#pragma unroll(B <= 5 ? 64 : 1)
                for (int c = 0; c < B; c++) {
                    double2 t = M[j * B + c];
                    M[j * B + c] = M[i * B + c];
                    M[i * B + c] = t;
                    t = G[j * B + c];
                    G[j * B + c] = G[i * B + c];
                    G[i * B + c] = t;
                }
        const double2 ip = cinv(M[j * B + j]);
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) {
            M[j * B + c] = cmul(M[j * B + c], ip);
            G[j * B + c] = cmul(G[j * B + c], ip);
        }
#pragma unroll(B <= 5 ? 64 : 1)
        for (int i = 0; i < B; i++) {
            if (i == j) continue;
            const double2 l = M[i * B + j];
#pragma unroll(B <= 5 ? 64 : 1)
            for (int c = 0; c < B; c++) {
                M[i * B + c] = csub(M[i * B + c], cmul(l, M[j * B + c]));
                G[i * B + c] = csub(G[i * B + c], cmul(l, G[j * B + c]));
            }
        }
    }
}

// P = X Y (B x B)
template <int B>
__device__ inline void mm(const double2 *X, const double2 *Y, double2 *P)
{
#pragma unroll(B <= 5 ? 64 : 1)
    for (int r = 0; r < B; r++)
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) {
            double2 a = make_double2(0.0, 0.0);
#pragma unroll(B <= 5 ? 64 : 1)
            for (int k = 0; k < B; k++) a = cadd(a, cmul(X[r * B + k], Y[k * B + c]));
            P[r * B + c] = a;
        }
}

// X <- X Y, row by row (row r of X Y needs only row r of X)
template <int B>
__device__ inline void mm_left_inplace(double2 *X, const double2 *Y)
{
#pragma unroll(B <= 5 ? 64 : 1)
    for (int r = 0; r < B; r++) {
        double2 row[B];
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) {
            double2 a = make_double2(0.0, 0.0);
#pragma unroll(B <= 5 ? 64 : 1)
            for (int k = 0; k < B; k++) a = cadd(a, cmul(X[r * B + k], Y[k * B + c]));
            row[c] = a;
        }
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) X[r * B + c] = row[c];
    }
}

__device__ inline void store_c(double2 *p, double2 v) { *p = v; }
__device__ inline void store_c(float2 *p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

template <typename C, int B>
__global__ void __launch_bounds__(64) k_surface_gk(int K, SurfSets sets, const double2 *__restrict__ tw, double efermi,
                                                   int scmax, const double *__restrict__ omegas, double eta,
                                                   double epserr, int maxit, int surface, C *__restrict__ Gk,
                                                   int *__restrict__ stat)
{
    constexpr int NP = B * B;
    const long KK = (long)K * K;
    const long k0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    // lanes past the grid repeat the last point (their results are not stored),
    // so that the whole wave takes part in the step-count reduction below
    const bool active = k0 < KK;
    const long k = active ? k0 : KK - 1;
    const int i = (int)(k / K), j = (int)(k - (long)i * K);
    double2 es[NP], eb[NP], al[NP], be[NP], g[NP], t1[NP];
    build_h<B>(sets.g[1], sets.n[1], sets.t[1], tw, i, j, K, true, efermi, scmax, es);
    build_h<B>(sets.g[0], sets.n[0], sets.t[0], tw, i, j, K, true, efermi, scmax, eb);
    build_h<B>(sets.g[2], sets.n[2], sets.t[2], tw, i, j, K, false, efermi, scmax, al);
#pragma unroll(B <= 5 ? 64 : 1)
    for (int r = 0; r < B; r++)
#pragma unroll(B <= 5 ? 64 : 1)
        for (int c = 0; c < B; c++) be[r * B + c] = cconj(al[c * B + r]);
    const double om = omegas[e];
    const double eps2 = epserr * epserr;
    int it = 0;
    for (;;) {
        resinv<B>(eb, om, eta, g);
        mm<B>(g, be, t1);  // t1 = g beta
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < B; r++)
#pragma unroll(B <= 5 ? 64 : 1)
            for (int c = 0; c < B; c++) {
                double2 a = make_double2(0.0, 0.0);  // (alpha g beta)_rc
#pragma unroll(B <= 5 ? 64 : 1)
                for (int q = 0; q < B; q++) a = cadd(a, cmul(al[r * B + q], t1[q * B + c]));
                es[r * B + c] = cadd(es[r * B + c], a);
                eb[r * B + c] = cadd(eb[r * B + c], a);
            }
        mm_left_inplace<B>(g, al);  // g <- g alpha
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < B; r++)
#pragma unroll(B <= 5 ? 64 : 1)
            for (int c = 0; c < B; c++) {
                double2 a = make_double2(0.0, 0.0);  // (beta g alpha)_rc
#pragma unroll(B <= 5 ? 64 : 1)
                for (int q = 0; q < B; q++) a = cadd(a, cmul(be[r * B + q], g[q * B + c]));
                eb[r * B + c] = cadd(eb[r * B + c], a);
            }
        mm_left_inplace<B>(be, t1);  // beta <- beta g beta
        mm_left_inplace<B>(al, g);   // alpha <- alpha g alpha
        it++;
        double na = 0.0, nb = 0.0;
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < NP; r++) {
            na += cabs2(al[r]);
            nb += cabs2(be[r]);
        }
        if (!(na > eps2 || nb > eps2)) break;  // both Frobenius norms <= epserr
        if (it >= maxit) {
            it = maxit + 1;  // not converged
            break;
        }
    }
    resinv<B>(surface ? es : eb, om, eta, g);
    if (active)
#pragma unroll(B <= 5 ? 64 : 1)
        for (int r = 0; r < NP; r++) store_c(Gk + ((size_t)e * NP + r) * KK + k, g[r]);
    // the wave's largest step count, one atomic per wave (blocks are one wave)
    int wmax = it;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    if (threadIdx.x == 0) atomicMax(stat + e, wmax);
}

// ---------------------------------------------------------------------------
// 5 <= B <= 16 bands: one 16-lane group per (k, energy), lane r holding row r
// of eps_s, eps_b, alpha, beta and g in registers (no scratch).  A product
// X Y stages Y's rows in the group's LDS and each lane forms its row of X Y
// from them; the inverse is Gauss-Jordan with partial pivoting, the pivot
// found by a shuffle reduction over the group, rows exchanged through LDS
// and the pivot row broadcast through LDS.  Lanes r >= B shadow row B-1
// and store nothing.  Four items a wave; a wave loops until all four have
// converged (a converged group skips the work).
// ---------------------------------------------------------------------------
constexpr int kGroup = 16;

__device__ inline void group_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row rr of the in-plane block (diag: with the -/+ efermi diagonal) or coupling
template <int B>
__device__ inline void build_h_row(const DevRGroup *__restrict__ groups, int ngroups, const double2 *__restrict__ tmat,
                                   const double2 *__restrict__ tw, int i, int j, int K, bool diag, double efermi,
                                   int scmax, int rr, double2 *H)
{
#pragma unroll
    for (int c = 0; c < B; c++) H[c] = make_double2(0.0, 0.0);
    if (diag)
#pragma unroll
        for (int c = 0; c < B; c++)
            if (c == rr) H[c].x = rr < scmax ? -efermi : efermi;
    const double kx = (i > (K >> 1)) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
    const double ky = (j > (K >> 1)) ? ((double)j - (double)K) / (double)K : (double)j / (double)K;
    const unsigned kmask = (K & (K - 1)) == 0 ? (unsigned)K - 1u : 0u;
    for (int gi = 0; gi < ngroups; gi++) {
        const DevRGroup G = groups[gi];
        double2 e;
        if (G.integral) {
            const unsigned m = (unsigned)G.irx * (unsigned)i + (unsigned)G.iry * (unsigned)j;
            e = tw[kmask ? (m & kmask) : (m % (unsigned)K)];
        } else {
            double sn, cs;
            sincos(2.0 * M_PI * (G.rx * kx + G.ry * ky), &sn, &cs);
            e = make_double2(cs, sn);
        }
        const double2 *T = tmat + ((size_t)gi * B + rr) * B;
#pragma unroll
        for (int c = 0; c < B; c++) H[c] = cadd(H[c], cmul(T[c], e));
    }
}

// P = X Y, X the lane's row, Y staged row-major in S (B x B).  The k loop
// stays rolled (X[k] picked by compares): unrolled, the loads of all B staged
// rows were hoisted ahead and held B^2 values in registers (spills for b > 5)
template <int B>
__device__ inline void row_times(const double2 *X, const double2 *S, double2 *P)
{
#pragma unroll
    for (int c = 0; c < B; c++) P[c] = make_double2(0.0, 0.0);
#pragma unroll 1
    for (int k = 0; k < B; k++) {
        double2 x = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q == k) x = X[q];
#pragma unroll
        for (int c = 0; c < B; c++) P[c] = cadd(P[c], cmul(x, S[k * B + c]));
    }
}

template <int B>
__device__ inline void stage_row(double2 *S, int r, const double2 *row)
{
    if (r < B)
#pragma unroll
        for (int c = 0; c < B; c++) S[r * B + c] = row[c];
}

// G = (w + i eta - H)^-1 by rows; S1, S2 (B x B) and PV (2B) are the group's LDS
template <int B>
__device__ inline void row_resinv(const double2 *H, double om, double eta, int r, int rr, double2 *G, double2 *S1,
                                  double2 *S2, double2 *PV)
{
    double2 M[B];
#pragma unroll
    for (int c = 0; c < B; c++) {
        M[c] = make_double2((c == rr ? om : 0.0) - H[c].x, (c == rr ? eta : 0.0) - H[c].y);
        G[c] = make_double2(c == rr ? 1.0 : 0.0, 0.0);
    }
#pragma unroll
    for (int j = 0; j < B; j++) {
        // pivot: the largest |M_ij| over the rows i >= j
        double v = (r < B && rr >= j) ? cabs2(M[j]) : -1.0;
        int piv = rr;
#pragma unroll
        for (int off = kGroup / 2; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off, kGroup);
            const int op = __shfl_xor(piv, off, kGroup);
            if (ov > v || (ov == v && op < piv)) {
                v = ov;
                piv = op;
            }
        }
        if (piv != j) {
            // exchange rows j and piv through LDS
            stage_row<B>(S1, r, M);
            stage_row<B>(S2, r, G);
            group_sync();
            const int src = rr == j ? piv : (rr == piv ? j : rr);
            if (rr == j || rr == piv)
#pragma unroll
                for (int c = 0; c < B; c++) {
                    M[c] = S1[src * B + c];
                    G[c] = S2[src * B + c];
                }
            group_sync();
        }
        if (rr == j && r < B) {
            const double2 ip = cinv(M[j]);
#pragma unroll
            for (int c = 0; c < B; c++) {
                M[c] = cmul(M[c], ip);
                G[c] = cmul(G[c], ip);
                PV[c] = M[c];
                PV[B + c] = G[c];
            }
        }
        group_sync();
        if (rr != j) {
            const double2 l = M[j];
#pragma unroll
            for (int c = 0; c < B; c++) {
                M[c] = csub(M[c], cmul(l, PV[c]));
                G[c] = csub(G[c], cmul(l, PV[B + c]));
            }
        }
        group_sync();
    }
}

template <typename C, int B>
__global__ void __launch_bounds__(64) k_surface_gk_g(int K, SurfSets sets, const double2 *__restrict__ tw,
                                                     double efermi, int scmax, const double *__restrict__ omegas,
                                                     double eta, double epserr, int maxit, int surface,
                                                     C *__restrict__ Gk, int *__restrict__ stat)
{
    constexpr int NP = B * B;
    __shared__ double2 sm[64 / kGroup][2 * NP + 2 * B];
    const int lane = threadIdx.x & 63, grp = lane / kGroup, r = lane % kGroup;
    const int rr = r < B ? r : B - 1;
    const long KK = (long)K * K;
    const long k0 = (long)blockIdx.x * (64 / kGroup) + grp;
    const int e = blockIdx.y;
    const bool active = k0 < KK;
    const long k = active ? k0 : KK - 1;
    const int i = (int)(k / K), j = (int)(k - (long)i * K);
    double2 *S1 = sm[grp], *S2 = S1 + NP, *PV = S2 + NP;
    double2 es[B], eb[B], al[B], be[B], g[B], t1[B];
    build_h_row<B>(sets.g[1], sets.n[1], sets.t[1], tw, i, j, K, true, efermi, scmax, rr, es);
    build_h_row<B>(sets.g[0], sets.n[0], sets.t[0], tw, i, j, K, true, efermi, scmax, rr, eb);
    build_h_row<B>(sets.g[2], sets.n[2], sets.t[2], tw, i, j, K, false, efermi, scmax, rr, al);
    // beta = alpha^H: row rr of beta is the conjugated column rr of alpha
    stage_row<B>(S1, r, al);
    group_sync();
#pragma unroll
    for (int c = 0; c < B; c++) be[c] = cconj(S1[c * B + rr]);
    group_sync();
    const double om = omegas[e];
    const double eps2 = epserr * epserr;
    int it = 0;
    bool done = false;
    for (;;) {
        if (!done) {
            row_resinv<B>(eb, om, eta, r, rr, g, S1, S2, PV);  // g = (w + i eta - eps_b)^-1
            stage_row<B>(S1, r, be);
            group_sync();
            row_times<B>(g, S1, t1);  // t1 = g beta
            group_sync();
            stage_row<B>(S2, r, t1);
            group_sync();
            double2 a[B];
            row_times<B>(al, S2, a);  // alpha g beta
#pragma unroll
            for (int c = 0; c < B; c++) {
                es[c] = cadd(es[c], a[c]);
                eb[c] = cadd(eb[c], a[c]);
            }
            stage_row<B>(S1, r, al);
            group_sync();
            row_times<B>(g, S1, a);  // g alpha
            group_sync();
#pragma unroll
            for (int c = 0; c < B; c++) g[c] = a[c];
            stage_row<B>(S1, r, g);
            group_sync();
            row_times<B>(be, S1, a);  // beta g alpha
#pragma unroll
            for (int c = 0; c < B; c++) eb[c] = cadd(eb[c], a[c]);
            row_times<B>(be, S2, a);  // beta <- beta g beta
#pragma unroll
            for (int c = 0; c < B; c++) be[c] = a[c];
            row_times<B>(al, S1, a);  // alpha <- alpha g alpha
#pragma unroll
            for (int c = 0; c < B; c++) al[c] = a[c];
            group_sync();
            it++;
            double na = 0.0, nb = 0.0;
            if (r < B)
#pragma unroll
                for (int c = 0; c < B; c++) {
                    na += cabs2(al[c]);
                    nb += cabs2(be[c]);
                }
#pragma unroll
            for (int off = kGroup / 2; off > 0; off >>= 1) {
                na += __shfl_xor(na, off, kGroup);
                nb += __shfl_xor(nb, off, kGroup);
            }
            if (!(na > eps2 || nb > eps2)) done = true;  // both Frobenius norms <= epserr
            else if (it >= maxit) {
                it = maxit + 1;  // not converged
                done = true;
            }
        }
        if (__all(done)) break;
    }
    row_resinv<B>(surface ? es : eb, om, eta, r, rr, g, S1, S2, PV);
    if (active && r < B)
#pragma unroll
        for (int c = 0; c < B; c++) store_c(Gk + ((size_t)e * NP + r * B + c) * KK + k, g[c]);
    int wmax = it;
    for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    if (threadIdx.x == 0) atomicMax(stat + e, wmax);
}

__global__ void k_zero_int(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

namespace cqpi {

int launch_surface_gk(int B, int K, const DevRGroup *const g[3], const double2 *const t[3], const int n[3],
                      const double2 *tw, double efermi, int scmax, const double *omegas, int E, double eta,
                      double epserr, int maxit, int surface, void *Gk, bool fp32, int *stat, hipStream_t s)
{
    SurfSets sets;
    for (int q = 0; q < 3; q++) {
        sets.g[q] = g[q];
        sets.t[q] = t[q];
        sets.n[q] = n[q];
    }
    if (E < 1 || K < 1) return CQPI_E_ARG;
    hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((E + 63) / 64)), dim3(64), 0, s, stat, E);
    const long KK = (long)K * K;
    const dim3 grid((unsigned)((KK + 63) / 64), (unsigned)E);
    const dim3 gridg((unsigned)((KK + 64 / kGroup - 1) / (64 / kGroup)), (unsigned)E);  // four (k, e) a wave
    switch (B) {
#define CASE(b)                                                                                                    \
    case b:                                                                                                        \
        if (fp32)                                                                                                  \
            hipLaunchKernelGGL((k_surface_gk<float2, b>), grid, dim3(64), 0, s, K, sets, tw, efermi, scmax, omegas, \
                               eta, epserr, maxit, surface, static_cast<float2 *>(Gk), stat);                      \
        else                                                                                                       \
            hipLaunchKernelGGL((k_surface_gk<double2, b>), grid, dim3(64), 0, s, K, sets, tw, efermi, scmax,        \
                               omegas, eta, epserr, maxit, surface, static_cast<double2 *>(Gk), stat);             \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
#define CASEG(b)                                                                                                   \
    case b:                                                                                                        \
        if (fp32)                                                                                                  \
            hipLaunchKernelGGL((k_surface_gk_g<float2, b>), gridg, dim3(64), 0, s, K, sets, tw, efermi, scmax,      \
                               omegas, eta, epserr, maxit, surface, static_cast<float2 *>(Gk), stat);              \
        else                                                                                                       \
            hipLaunchKernelGGL((k_surface_gk_g<double2, b>), gridg, dim3(64), 0, s, K, sets, tw, efermi, scmax,     \
                               omegas, eta, epserr, maxit, surface, static_cast<double2 *>(Gk), stat);             \
        break;
        CASEG(5) CASEG(6) CASEG(7) CASEG(8) CASEG(9) CASEG(10) CASEG(11) CASEG(12) CASEG(14) CASEG(16)
#undef CASEG
    default: return CQPI_E_ARG;
    }
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

// Batched per-k Hermitian eigensolve of H(k) on the K x K grid (FP64).
//
// One thread per k-point.  H(k) = sum_R t(R) exp(+-2 pi i R.k) - efermi
// (src/tightbinding.cc:210-219, src/tightbinding.h:128-135), accumulated as
// one dense B x B block per lattice vector, on the grid
// k_i = i/K or (i-K)/K for i > K/2 (src/qpi.cc:465-466), diagonalised by
// cyclic complex Jacobi rotations.  Eigen data is written structure-of-arrays
// so that consecutive threads (consecutive k) store consecutive addresses:
//   eval[n*KK + k], evec[(r*B + n)*KK + k]  (column n = eigenvector n).
// This is computed once per model and cached across energies, as the
// reference caches evect/eval (src/qpi.cc:479-480).
#include "cqpi_common.h"

namespace {

__device__ inline double kcoord(int i, int K)
{
    return (i > (K >> 1)) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
}

// One cyclic Jacobi rotation on (p, q) of the b x b Hermitian A, accumulating
// the eigenvectors in V: A <- U^H A U, V <- V U with
// U = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] on rows/columns p, q, where
// A_pq = r e^{i phi} and t = s/c solves t^2 + 2 theta t - 1 = 0 with
// theta = (A_qq - A_pp) / (2r) (smaller root).
__device__ inline void rotate(double2 *A, double2 *V, int b, int p, int q)
{
    double2 apq = A[p * b + q];
    double r = sqrt(cabs2(apq));
    if (r == 0.0) return;
    double app = A[p * b + p].x, aqq = A[q * b + q].x;
    // one reciprocal of r for the phase and theta, one of tau for t and c:
    // t = sign(theta) / tau, tau = |theta| + sqrt(theta^2 + 1) >= 1, and
    // c = 1 / sqrt(t^2 + 1) = tau / sqrt(tau^2 + 1)
    const double ir = 1.0 / r;
    double2 ph = make_double2(apq.x * ir, apq.y * ir);  // e^{i phi}
    double theta = (aqq - app) * (0.5 * ir);
    double t, c;
    if (fabs(theta) > 1e150) {
        t = 0.5 / theta;
        c = 1.0 / sqrt(t * t + 1.0);
    } else {
        const double tau = fabs(theta) + sqrt(theta * theta + 1.0);
        const double it = 1.0 / tau;
        t = theta >= 0.0 ? it : -it;
        c = 1.0 / sqrt(it * it + 1.0);
    }
    double s = t * c;
    double2 sph = cscale(s, ph);          // s e^{i phi}
    double2 sphc = cconj(sph);            // s e^{-i phi}
    double2 cphc = cscale(c, cconj(ph));  // c e^{-i phi}
    double2 cph = cscale(c, ph);
    for (int k = 0; k < b; k++) {
        double2 akp = A[k * b + p], akq = A[k * b + q];
        A[k * b + p] = csub(cscale(c, akp), cmul(sphc, akq));
        A[k * b + q] = cadd(cscale(s, akp), cmul(cphc, akq));
    }
    for (int k = 0; k < b; k++) {
        double2 apk = A[p * b + k], aqk = A[q * b + k];
        A[p * b + k] = csub(cscale(c, apk), cmul(sph, aqk));
        A[q * b + k] = cadd(cscale(s, apk), cmul(cph, aqk));
    }
    A[p * b + q] = make_double2(0.0, 0.0);
    A[q * b + p] = make_double2(0.0, 0.0);
    A[p * b + p] = make_double2(app - t * r, 0.0);
    A[q * b + q] = make_double2(aqq + t * r, 0.0);
    for (int k = 0; k < b; k++) {
        double2 vkp = V[k * b + p], vkq = V[k * b + q];
        V[k * b + p] = csub(cscale(c, vkp), cmul(sphc, vkq));
        V[k * b + q] = cadd(cscale(s, vkp), cmul(cphc, vkq));
    }
}

__device__ inline bool converged(const double2 *A, int b)
{
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < b; p++) {
        diag += A[p * b + p].x * A[p * b + p].x;
        for (int q = p + 1; q < b; q++) off += cabs2(A[p * b + q]);
    }
    return off == 0.0 || off < 1e-34 * diag;
}

// Cyclic-by-row Jacobi, one thread per k-point.  Small B: everything unrolled
// so A and V stay in registers.  B = 14, 16: rolled loops (A and V live in
// private memory).  6 <= B <= 12 use the lane-group kernel k_eigen_g below.
constexpr int kUnrollB = 5;

template <int B>
__device__ void jacobi(double2 (&A)[B * B], double2 (&V)[B * B])
{
    for (int i = 0; i < B * B; i++) V[i] = make_double2((i % (B + 1)) == 0 ? 1.0 : 0.0, 0.0);
    for (int i = 0; i < B; i++) A[i * B + i].y = 0.0;
    if constexpr (B <= kUnrollB) {
        for (int sweep = 0; sweep < 60; sweep++) {
            bool done;
            {
                double off = 0.0, diag = 0.0;
#pragma unroll
                for (int p = 0; p < B; p++) {
                    diag += A[p * B + p].x * A[p * B + p].x;
#pragma unroll
                    for (int q = p + 1; q < B; q++) off += cabs2(A[p * B + q]);
                }
                done = off == 0.0 || off < 1e-34 * diag;
            }
            if (done) break;
#pragma unroll
            for (int p = 0; p < B - 1; p++)
#pragma unroll
                for (int q = p + 1; q < B; q++) rotate(A, V, B, p, q);
        }
    } else {
        for (int sweep = 0; sweep < 60; sweep++) {
            if (converged(A, B)) break;
#pragma unroll 1
            for (int p = 0; p < B - 1; p++)
#pragma unroll 1
                for (int q = p + 1; q < B; q++) rotate(A, V, B, p, q);
        }
    }
}

template <int B>
__global__ void __launch_bounds__(128) k_eigen(const DevRGroup *__restrict__ groups, int ngroups,
                                               const double2 *__restrict__ tmat, const double2 *__restrict__ tw,
                                               double efermi, int scmax, int K, double *__restrict__ eval,
                                               double2 *__restrict__ evec)
{
    const long KK = (long)K * K;
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= KK) return;
    int i = (int)(idx / K), j = (int)(idx % K);
    double kx = kcoord(i, K), ky = kcoord(j, K);
    double2 A[B * B];
    double2 V[B * B];
#pragma unroll
    for (int r = 0; r < B * B; r++) A[r] = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < B; r++) A[r * B + r].x = (r < scmax) ? -efermi : efermi;
    const unsigned kmask = (K & (K - 1)) == 0 ? (unsigned)K - 1u : 0u;
    for (int g = 0; g < ngroups; g++) {
        const DevRGroup G = groups[g];
        double2 e;
        if (G.integral) {
            // irx, iry < K <= 32768: the sum stays below 2^31
            const unsigned m = (unsigned)G.irx * (unsigned)i + (unsigned)G.iry * (unsigned)j;
            e = tw[kmask ? (m & kmask) : (m % (unsigned)K)];
        } else {
            double sn, cs;
            sincos(2.0 * M_PI * (G.rx * kx + G.ry * ky), &sn, &cs);
            e = make_double2(cs, sn);
        }
        const double2 *T = tmat + (size_t)g * B * B;
        if constexpr (B <= kUnrollB) {
#pragma unroll
            for (int r = 0; r < B * B; r++) A[r] = cadd(A[r], cmul(T[r], e));
        } else {
            for (int r = 0; r < B * B; r++) A[r] = cadd(A[r], cmul(T[r], e));
        }
    }
    jacobi<B>(A, V);
#pragma unroll
    for (int nn = 0; nn < B; nn++) eval[nn * KK + idx] = A[nn * B + nn].x;
#pragma unroll
    for (int r = 0; r < B * B; r++) evec[r * KK + idx] = V[r];
}

// ---------------------------------------------------------------------------
// 6 <= b <= 12: one lane group of 16 lanes per k-point, a parallel cyclic
// Jacobi.  Lane r holds row r of A = V^H H V and row r of V (b rounded up to
// an even M with a decoupled zero row).  Each round rotates the M/2 disjoint
// pairs (r, M-1-r) at once; between rounds the positions 1..M-1 move round a
// circle (position 0 fixed), so that M-1 rounds meet every pair once and
// return every index to its place (a sweep).  Rows move between lanes by
// shuffles, columns by a fixed permutation of each lane's registers, so every
// register index is static.  The rotation of a pair is the one of the
// per-thread kernel (explicit zero for A_pq, diagonal from t).
// ---------------------------------------------------------------------------
constexpr int kGroup = 16;

__device__ inline double2 shfl2(double2 v, int src)
{
    return make_double2(__shfl(v.x, src, kGroup), __shfl(v.y, src, kGroup));
}

// source position of new position i after one turn of the circle
__host__ __device__ constexpr int circle_src(int i, int M) { return i == 0 ? 0 : (i == 1 ? M - 1 : i - 1); }

template <int B>
__global__ void __launch_bounds__(256) k_eigen_g(const DevRGroup *__restrict__ groups, int ngroups,
                                                 const double2 *__restrict__ tmat, const double2 *__restrict__ tw,
                                                 double efermi, int scmax, int K, double *__restrict__ eval,
                                                 double2 *__restrict__ evec)
{
    constexpr int M = B + (B & 1);
    static_assert(M <= kGroup, "lane group too small");
    const long KK = (long)K * K;
    const int r = threadIdx.x & (kGroup - 1);
    const long gid = ((long)blockIdx.x * blockDim.x + threadIdx.x) / kGroup;  // k-point of the group
    const bool kok = gid < KK;   // uniform over the group
    const long k = kok ? gid : 0;
    const int i = (int)(k / K), j = (int)(k - (long)i * K);
    const bool act = r < M;
    const int s = act ? M - 1 - r : r;          // partner position
    const bool isp = r < M / 2;
    const int pl = isp ? r : s, ql = isp ? s : r;  // p and q lanes of my pair

    double2 Ar[M], Vr[M];
#pragma unroll
    for (int c = 0; c < M; c++) {
        Ar[c] = make_double2(0.0, 0.0);
        Vr[c] = make_double2(c == r ? 1.0 : 0.0, 0.0);
    }
    if (r < B) {
        const double dia = r < scmax ? -efermi : efermi;
#pragma unroll
        for (int c = 0; c < M; c++)
            if (c == r) Ar[c].x = dia;
        const double kx = (i > (K >> 1)) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
        const double ky = (j > (K >> 1)) ? ((double)j - (double)K) / (double)K : (double)j / (double)K;
        const unsigned kmask = (K & (K - 1)) == 0 ? (unsigned)K - 1u : 0u;
        for (int g = 0; g < ngroups; g++) {
            const DevRGroup G = groups[g];
            double2 e;
            if (G.integral) {
                const unsigned m = (unsigned)G.irx * (unsigned)i + (unsigned)G.iry * (unsigned)j;
                e = tw[kmask ? (m & kmask) : (m % (unsigned)K)];
            } else {
                double sn, cs;
                sincos(2.0 * M_PI * (G.rx * kx + G.ry * ky), &sn, &cs);
                e = make_double2(cs, sn);
            }
            const double2 *T = tmat + ((size_t)g * B + r) * B;
#pragma unroll
            for (int c = 0; c < B; c++) Ar[c] = cadd(Ar[c], cmul(T[c], e));
        }
#pragma unroll
        for (int c = 0; c < M; c++)
            if (c == r) Ar[c].y = 0.0;
    }

    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int c = 0; c < M; c++) {
            if (c == r) dg += Ar[c].x * Ar[c].x;
            else off += cabs2(Ar[c]);
        }
        for (int o = 1; o < kGroup; o <<= 1) {
            off += __shfl_xor(off, o, kGroup);
            dg += __shfl_xor(dg, o, kGroup);
        }
        if (off == 0.0 || off < 1e-34 * dg) break;  // uniform over the group
        for (int rnd = 0; rnd < M - 1; rnd++) {
            // my diagonal and my pair element
            double2 d = make_double2(0.0, 0.0), x = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < M; c++) {
                if (c == r) d = Ar[c];
                if (c == s) x = Ar[c];
            }
            // the pair's rotation, from the p lane's A_pp, A_pq and the q lane's A_qq
            const double app = __shfl(d.x, pl, kGroup), aqq = __shfl(d.x, ql, kGroup);
            const double2 apq = shfl2(x, pl);
            double cr = 1.0, sr = 0.0, tr = 0.0;
            double2 ph = make_double2(1.0, 0.0);
            const double rr = sqrt(cabs2(apq));
            if (act && rr != 0.0) {
                ph = make_double2(apq.x / rr, apq.y / rr);
                const double theta = (aqq - app) / (2.0 * rr);
                double t;
                if (fabs(theta) > 1e150)
                    t = 0.5 / theta;
                else
                    t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                cr = 1.0 / sqrt(t * t + 1.0);
                sr = t * cr;
                tr = t * rr;
            }
            // columns p = jp, q = M-1-jp of every row (A and V), with pair jp's rotation
#pragma unroll
            for (int jp = 0; jp < M / 2; jp++) {
                const double cj = __shfl(cr, jp, kGroup), sj = __shfl(sr, jp, kGroup);
                const double2 phj = shfl2(ph, jp);
                const double2 sphc = cscale(sj, cconj(phj)), cphc = cscale(cj, cconj(phj));
                const int p = jp, q = M - 1 - jp;
                double2 a = Ar[p], b = Ar[q];
                Ar[p] = csub(cscale(cj, a), cmul(sphc, b));
                Ar[q] = cadd(cscale(sj, a), cmul(cphc, b));
                a = Vr[p];
                b = Vr[q];
                Vr[p] = csub(cscale(cj, a), cmul(sphc, b));
                Vr[q] = cadd(cscale(sj, a), cmul(cphc, b));
            }
            // rows p and q: A_p <- c A_p - s ph A_q,  A_q <- s A_p + c ph A_q
            {
                const double2 sph = cscale(sr, ph), cph = cscale(cr, ph);
#pragma unroll
                for (int c = 0; c < M; c++) {
                    const double2 o = shfl2(Ar[c], s);
                    Ar[c] = isp ? csub(cscale(cr, Ar[c]), cmul(sph, o)) : cadd(cscale(sr, o), cmul(cph, Ar[c]));
                }
            }
            // exact zero for the rotated pair, diagonal from t
#pragma unroll
            for (int c = 0; c < M; c++) {
                if (c == s && act) Ar[c] = make_double2(0.0, 0.0);
                if (c == r && act) Ar[c] = make_double2(isp ? app - tr : aqq + tr, 0.0);
            }
            // one turn of the circle: rows by shuffle, columns by a fixed permutation
            {
                const int src = act ? circle_src(r, M) : r;
                double2 t[M];
#pragma unroll
                for (int c = 0; c < M; c++) t[c] = shfl2(Ar[c], src);
#pragma unroll
                for (int c = 0; c < M; c++) Ar[c] = t[circle_src(c, M)];
#pragma unroll
                for (int c = 0; c < M; c++) t[c] = Vr[c];
#pragma unroll
                for (int c = 0; c < M; c++) Vr[c] = t[circle_src(c, M)];
            }
        }
    }
    if (kok && r < B) {
        double dv = 0.0;
#pragma unroll
        for (int c = 0; c < M; c++)
            if (c == r) dv = Ar[c].x;
        eval[(long)r * KK + k] = dv;
#pragma unroll
        for (int n = 0; n < B; n++) evec[((long)r * B + n) * KK + k] = Vr[n];
    }
}

// Band-structure histogram (tightbind::calcbst3dhisthr, src/tightbinding.cc:
// 1158-1193, ux = uy = 1, all bands): the eigenvalues of H(k) at the
// (pix os)^2 points k = (io/(pix os - 1) - 1/2, jo/(pix os - 1) - 1/2), each
// added to pixel (io/os, jo/os) of the layers nearest to it as a Lorentzian
// of full width 2 (ue - le)/layers.  One thread per k; hist[(x + y pix) +
// layer pix^2] accumulates with atomics (the order of the sums is free).
template <int B>
__global__ void __launch_bounds__(128) k_bands_hist(const DevRGroup *__restrict__ groups, int ngroups,
                                                    const double2 *__restrict__ tmat, double efermi, int scmax,
                                                    int pix, int os, double le, double ue, int layers,
                                                    double *__restrict__ hist)
{
    const long ospix = (long)pix * os;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ospix * ospix) return;
    const int io = (int)(idx / ospix), jo = (int)(idx - (long)io * ospix);
    const double kx = (double)io / (double)(ospix - 1) - 0.5, ky = (double)jo / (double)(ospix - 1) - 0.5;
    double2 A[B * B];
    double2 V[B * B];
#pragma unroll
    for (int r = 0; r < B * B; r++) A[r] = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < B; r++) A[r * B + r].x = (r < scmax) ? -efermi : efermi;
    for (int g = 0; g < ngroups; g++) {
        const DevRGroup G = groups[g];
        double sn, cs;
        sincos(2.0 * M_PI * (G.rx * kx + G.ry * ky), &sn, &cs);
        const double2 e = make_double2(cs, sn);
        const double2 *T = tmat + (size_t)g * B * B;
        for (int r = 0; r < B * B; r++) A[r] = cadd(A[r], cmul(T[r], e));
    }
    jacobi<B>(A, V);
    const double bin = (ue - le) / (double)(layers - 1), w = 2.0 * (ue - le) / (double)layers;
    double *px = hist + (long)(io / os) + (long)(jo / os) * pix;
    const long plane = (long)pix * pix;
    for (int n = 0; n < B; n++) {
        const double ev = A[n * B + n].x;
        if (!(ev > le - 2.0 * bin && ev < ue + 2.0 * bin)) continue;
        const int lay = (int)round((ev - le) / (ue - le) * (double)(layers - 1));
        for (int nl = lay - 1; nl <= lay + 1; nl++) {
            if (nl < 0 || nl >= layers) continue;
            const double x = (ue - le) * (double)nl / (double)(layers - 1) + le - ev;
            atomicAdd(px + nl * plane, w / 2.0 / (x * x + w * w / 4.0) / M_PI);
        }
    }
}

template <int B>
int launch_eigen_t(const DevRGroup *groups, int ngroups, const double2 *tmat, const double2 *tw, double efermi,
                   int scmax, int K, double *eval, double2 *evec, hipStream_t s)
{
    long KK = (long)K * K;
    if constexpr (B >= 6 && B <= 12) {
        const long blocks = (KK * kGroup + 255) / 256;
        hipLaunchKernelGGL(k_eigen_g<B>, dim3((unsigned)blocks), dim3(256), 0, s, groups, ngroups, tmat, tw, efermi,
                           scmax, K, eval, evec);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    int threads = 128;
    long blocks = (KK + threads - 1) / threads;
    hipLaunchKernelGGL(k_eigen<B>, dim3((unsigned)blocks), dim3(threads), 0, s, groups, ngroups,
                       tmat, tw, efermi, scmax, K, eval, evec);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace

namespace cqpi {

int launch_eigen(int B, const DevRGroup *groups, int ngroups, const double2 *tmat, const double2 *tw,
                 double efermi, int scmax, int K, double *eval, double2 *evec, hipStream_t s)
{
    switch (B) {
#define CASE(b) \
    case b: return launch_eigen_t<b>(groups, ngroups, tmat, tw, efermi, scmax, K, eval, evec, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

int launch_bands_hist(int B, const DevRGroup *groups, int ngroups, const double2 *tmat, double efermi, int scmax,
                      int pix, int os, double le, double ue, int layers, double *hist, hipStream_t s)
{
    if (pix < 1 || os < 1 || (long)pix * os < 2 || layers < 2 || !(ue > le)) return CQPI_E_ARG;
    const long n = (long)pix * os * pix * os;
    const dim3 grid((unsigned)((n + 127) / 128));
    switch (B) {
#define CASE(b)                                                                                                 \
    case b:                                                                                                     \
        hipLaunchKernelGGL(k_bands_hist<b>, grid, dim3(128), 0, s, groups, ngroups, tmat, efermi, scmax, pix, os, \
                           le, ue, layers, hist);                                                               \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

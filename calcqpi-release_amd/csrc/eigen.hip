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
    double2 ph = make_double2(apq.x / r, apq.y / r);  // e^{i phi}
    double theta = (aqq - app) / (2.0 * r);
    double t;
    if (fabs(theta) > 1e150)
        t = 0.5 / theta;
    else
        t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
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

// Cyclic-by-row Jacobi.  Small B: everything unrolled so A and V stay in
// registers.  Larger B: rolled loops (A and V live in private memory).
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

template <int B>
int launch_eigen_t(const DevRGroup *groups, int ngroups, const double2 *tmat, const double2 *tw, double efermi,
                   int scmax, int K, double *eval, double2 *evec, hipStream_t s)
{
    long KK = (long)K * K;
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

}  // namespace cqpi

// Spectral-function maps (SPF, unfolded SPF) and the density of states: the
// reference's reductions of G0(k) and G0(R=0) (FP64).
//
//   SPF  (gpucalcspf, src/qpihipkernel.cc:317-329; CPU src/qpi.cc:657-681):
//        pixel ((i - n/2) mod n, (j - n/2) mod n) = -sum_{o<maxband} Im G0(k_ij)_oo,
//        i, j < n on the k grid (the driver sets n = K).
//   USPF (gpucalcuspf[spin], src/qpihipkernel.cc:331-370; CPU src/qpi.cc:731-775):
//        pixel (i, j) = -sum_{k,l} z_k z_l Im[ G0(k_g)_{l,k} e^{2 pi i q.(x_k - x_l)} ]
//        with q = ((i - n/2)/K, (j - n/2)/K) and grid index
//        k_g = ((i + n/2) mod n) mod K (per spin half with spin).
//   DOS  (tmatrix::calcdos, src/qpi.cc:856-862): -sum_{o<maxband} Im G0(R=0)_oo,
//        with G0(R=0) = K^-2 sum_k G0(k).
// G0(k) comes from the host (the drop-in calls hand over the k grid), from
// the cached eigen data, G0(k)_{l,m} = sum_n v_ln conj(v_mn) / (w + i eta - e_n),
// or from the G(k) workspace Gk[e][pair][k] that the surface Green's function fills.
#include <algorithm>

#include "cqpi_common.h"

namespace {

struct SpecArgs {
    int K, n, B, mb, spin, mode;  // mode 1: SPF, 2: USPF
};

// G0(k)_{l,m} at flat grid index kk from the eigen data (eval[n][KK], evec[(r B + n)][KK])
__device__ inline double2 g_eig(const double *__restrict__ eval, const double2 *__restrict__ evec, long KK, long kk,
                                int B, int l, int m, double om, double eta)
{
    double2 acc = make_double2(0.0, 0.0);
    for (int nn = 0; nn < B; nn++) {
        const double2 f = cinv(make_double2(om - eval[(long)nn * KK + kk], eta));
        const double2 vl = evec[((long)l * B + nn) * KK + kk], vm = evec[((long)m * B + nn) * KK + kk];
        acc = cadd(acc, cmul(cmulc(vl, vm), f));
    }
    return acc;
}

enum { kSrcHost = 0, kSrcEigen = 1, kSrcWork = 2 };

// SRC kSrcHost: G0(k) blocks g0k[kk][l*B + m] (drop-in, one energy);
// kSrcEigen: from the eigen data for energy omegas[blockIdx.y];
// kSrcWork: g0k is the workspace Gk[e][l*B + m][kk]
template <int SRC>
__global__ void __launch_bounds__(256) k_spectral(SpecArgs a, const double2 *__restrict__ g0k,
                                                  const double *__restrict__ eval, const double2 *__restrict__ evec,
                                                  const double *__restrict__ omegas, double eta,
                                                  const double *__restrict__ pos, double *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    const int n = a.n, K = a.K, B = a.B, n2 = n >> 1;
    if (p >= n * n) return;
    const long KK = (long)K * K;
    const double om = SRC == kSrcEigen ? omegas[e] : 0.0;
    auto G = [&](long kk, int l, int m) -> double2 {
        if (SRC == kSrcEigen) return g_eig(eval, evec, KK, kk, B, l, m, om, eta);
        if (SRC == kSrcWork) return g0k[((long)e * B * B + l * B + m) * KK + kk];
        return g0k[kk * B * B + l * B + m];
    };
    double tr = 0.0;
    if (a.mode == 1) {
        const int X = p % n, Y = p / n;
        const int i = (X + n2) % n, j = (Y + n2) % n;  // pixel X = (i - n/2) mod n
        const long kk = (long)i * K + j;
        for (int o = 0; o < a.mb; o++) tr -= G(kk, o, o).y;
    } else {
        const int i = p % n, j = p / n;
        const double qx = ((double)i - n2) / (double)K, qy = ((double)j - n2) / (double)K;
        const long kk = (long)(((i + n2) % n) % K) * K + ((j + n2) % n) % K;
        const int sb = a.mb >> 1, nsets = a.spin ? 2 : 1, np = a.spin ? sb : a.mb;
        for (int sg = 0; sg < nsets; sg++) {
            const int off = sg ? sb : 0;
            for (int k = off; k < off + np; k++)
                for (int l = off; l < off + np; l++) {
                    double sn, cs;
                    sincos(2.0 * M_PI * (qx * (pos[3 * k] - pos[3 * l]) + qy * (pos[3 * k + 1] - pos[3 * l + 1])), &sn,
                           &cs);
                    const double2 g = G(kk, l, k);
                    tr -= pos[3 * k + 2] * pos[3 * l + 2] * (g.x * sn + g.y * cs);  // Im(g e^{i ph})
                }
        }
    }
    out[(size_t)e * n * n + p] = tr;
}

constexpr int kDosThreads = 256;

// partial[e][c] = -sum over the grid points of chunk c of sum_{o<mb} Im G0(k)_oo
__global__ void __launch_bounds__(kDosThreads) k_dos_partial(int K, int B, int mb, const double *__restrict__ eval,
                                                             const double2 *__restrict__ evec,
                                                             const double *__restrict__ omegas, double eta,
                                                             double *__restrict__ partial)
{
    __shared__ double red[kDosThreads];
    const long KK = (long)K * K;
    const int e = blockIdx.y, c = blockIdx.x, nc = gridDim.x;
    const long k0 = KK * c / nc, k1 = KK * (c + 1) / nc;
    const double om = omegas[e];
    double acc = 0.0;
    for (long kk = k0 + threadIdx.x; kk < k1; kk += blockDim.x)
        for (int nn = 0; nn < B; nn++) {
            const double d = om - eval[(long)nn * KK + kk];
            const double lor = eta / (d * d + eta * eta);  // -Im 1/(w + i eta - e)
            double w2 = 0.0;
            for (int o = 0; o < mb; o++) w2 += cabs2(evec[((long)o * B + nn) * KK + kk]);
            acc += w2 * lor;
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kDosThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)e * nc + c] = red[0];
}

// the same partial sums from the G(k) workspace Gk[e][pair][k]
__global__ void __launch_bounds__(kDosThreads) k_dos_partial_gk(int K, int B, int mb, const double2 *__restrict__ Gk,
                                                                double *__restrict__ partial)
{
    __shared__ double red[kDosThreads];
    const long KK = (long)K * K;
    const int e = blockIdx.y, c = blockIdx.x, nc = gridDim.x;
    const long k0 = KK * c / nc, k1 = KK * (c + 1) / nc;
    double acc = 0.0;
    for (int o = 0; o < mb; o++) {
        const double2 *row = Gk + ((size_t)e * B * B + (size_t)o * B + o) * KK;
        for (long kk = k0 + threadIdx.x; kk < k1; kk += blockDim.x) acc -= row[kk].y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kDosThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)e * nc + c] = red[0];
}

__global__ void k_dos_final(int E, int nc, double scale, const double *__restrict__ partial, double *__restrict__ dos)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    for (int c = 0; c < nc; c++) s += partial[(size_t)e * nc + c];
    dos[e] = s * scale;
}

}  // namespace

namespace cqpi {

int launch_spectral(int K, int n, int B, int mb, int spin, int mode, const double2 *g0k, const double *eval,
                    const double2 *evec, const double *omegas, int E, double eta, const double *pos, double *out,
                    hipStream_t s)
{
    if ((mode != 1 && mode != 2) || n > K || mb > B || E < 1) return CQPI_E_ARG;
    const SpecArgs a{K, n, B, mb, spin, mode};
    const dim3 grid((unsigned)((n * n + 255) / 256), (unsigned)E);
    if (g0k)
        hipLaunchKernelGGL(k_spectral<kSrcHost>, grid, dim3(256), 0, s, a, g0k, eval, evec, omegas, eta, pos, out);
    else
        hipLaunchKernelGGL(k_spectral<kSrcEigen>, grid, dim3(256), 0, s, a, g0k, eval, evec, omegas, eta, pos, out);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_spectral_gk(int K, int n, int B, int mb, int spin, int mode, const double2 *Gk, int E, const double *pos,
                       double *out, hipStream_t s)
{
    if ((mode != 1 && mode != 2) || n > K || mb > B || E < 1) return CQPI_E_ARG;
    const SpecArgs a{K, n, B, mb, spin, mode};
    const dim3 grid((unsigned)((n * n + 255) / 256), (unsigned)E);
    hipLaunchKernelGGL(k_spectral<kSrcWork>, grid, dim3(256), 0, s, a, Gk, nullptr, nullptr, nullptr, 0.0, pos, out);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_dos_gk(int K, int B, int mb, const double2 *Gk, int E, double *partial, double *dos, hipStream_t s)
{
    if (mb > B || E < 1) return CQPI_E_ARG;
    const int nc = dos_chunks(K);
    hipLaunchKernelGGL(k_dos_partial_gk, dim3((unsigned)nc, (unsigned)E), dim3(kDosThreads), 0, s, K, B, mb, Gk,
                       partial);
    if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    hipLaunchKernelGGL(k_dos_final, dim3((unsigned)((E + 63) / 64)), dim3(64), 0, s, E, nc,
                       1.0 / ((double)K * (double)K), partial, dos);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int dos_chunks(int K) { return std::max(1, std::min(256, (K * K) / 4096)); }

int launch_dos(int K, int B, int mb, const double *eval, const double2 *evec, const double *omegas, int E,
               double eta, double *partial, double *dos, hipStream_t s)
{
    if (mb > B || E < 1) return CQPI_E_ARG;
    const int nc = dos_chunks(K);
    hipLaunchKernelGGL(k_dos_partial, dim3((unsigned)nc, (unsigned)E), dim3(kDosThreads), 0, s, K, B, mb, eval, evec,
                       omegas, eta, partial);
    if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    hipLaunchKernelGGL(k_dos_final, dim3((unsigned)((E + 63) / 64)), dim3(64), 0, s, E, nc,
                       1.0 / ((double)K * (double)K), partial, dos);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

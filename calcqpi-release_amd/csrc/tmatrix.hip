// T-matrix and the periodic (bulk) term of the LDOS (FP64).
//
// T = (1 - V G0(R=0))^{-1} V   (tmatrix::setscatteringmatrix, src/qpi.cc:833-854)
// solved per energy by Gaussian elimination with partial pivoting on
// [1 - V G0(0) | V].
//
// The bulk term of the reference's contraction,
//   -Im sum_{R1,R2,o1,o2} w_o1(x-R1) w_o2(x-R2) G0(R2-R1)_{o2,o1}
// (src/qpi.cc:921, src/qpihipkernel.cc:94), depends on the sub-pixel (ii,jj)
// only, not on the cell: it is computed once per energy as an os*os table.
#include "cqpi_common.h"

namespace {

constexpr int kMaxB = 16;

// one thread per energy.  MB: compile-time matrix side (the arrays stay in
// registers for small MB: the row exchange of the pivoting is done by compares
// so that every index is static); MB = kMaxB serves any B <= kMaxB
template <int MB>
__device__ inline void tmatrix_one(const Geom &g, const double2 *__restrict__ G0w, const double2 *__restrict__ V,
                                   int e, double2 *__restrict__ Tout, int *__restrict__ status)
{
    const int B = MB < kMaxB ? MB : g.B, NP = B * B;
    const double2 *G = G0w + ((size_t)e * g.nw * g.nw + (size_t)g.h * g.nw + g.h) * NP;
    double2 M[MB * MB], X[MB * MB];
#pragma unroll
    for (int r = 0; r < MB; r++)
#pragma unroll
        for (int c = 0; c < MB; c++) {
            if (r < B && c < B) {
                double2 acc = make_double2(r == c ? 1.0 : 0.0, 0.0);
                for (int k = 0; k < B; k++) acc = csub(acc, cmul(V[r * B + k], G[k * B + c]));
                M[r * MB + c] = acc;
                X[r * MB + c] = V[r * B + c];
            }
        }
    int singular = 0;
#pragma unroll
    for (int j = 0; j < MB; j++) {
        if (j >= B || singular) break;
        int piv = j;
        double best = cabs2(M[j * MB + j]);
#pragma unroll
        for (int i = j + 1; i < MB; i++) {
            if (i < B) {
                double v = cabs2(M[i * MB + j]);
                if (v > best) { best = v; piv = i; }
            }
        }
        if (best == 0.0) { singular = 1; break; }
#pragma unroll
        for (int i = j + 1; i < MB; i++) {
            if (i < B && i == piv) {
#pragma unroll
                for (int c = 0; c < MB; c++) {
                    if (c < B) {
                        double2 t = M[j * MB + c]; M[j * MB + c] = M[i * MB + c]; M[i * MB + c] = t;
                        t = X[j * MB + c]; X[j * MB + c] = X[i * MB + c]; X[i * MB + c] = t;
                    }
                }
            }
        }
        double2 invp = cinv(M[j * MB + j]);
#pragma unroll
        for (int i = 0; i < MB; i++) {
            if (i == j || i >= B) continue;
            double2 l = cmul(M[i * MB + j], invp);
            if (l.x == 0.0 && l.y == 0.0) continue;
#pragma unroll
            for (int c = 0; c < MB; c++)
                if (c >= j && c < B) M[i * MB + c] = csub(M[i * MB + c], cmul(l, M[j * MB + c]));
#pragma unroll
            for (int c = 0; c < MB; c++)
                if (c < B) X[i * MB + c] = csub(X[i * MB + c], cmul(l, X[j * MB + c]));
        }
#pragma unroll
        for (int c = 0; c < MB; c++)
            if (c < B) X[j * MB + c] = cmul(X[j * MB + c], invp);
#pragma unroll
        for (int c = 0; c < MB; c++)
            if (c >= j && c < B) M[j * MB + c] = cmul(M[j * MB + c], invp);
    }
    double2 *T = Tout + (size_t)e * NP;
#pragma unroll
    for (int r = 0; r < MB; r++)
#pragma unroll
        for (int c = 0; c < MB; c++)
            if (r < B && c < B) T[r * B + c] = X[r * MB + c];
    status[e] = singular;
}

template <int MB>
__global__ void __launch_bounds__(64) k_tmatrix(Geom g, const double2 *__restrict__ G0w, const double2 *__restrict__ V,
                          int E, double2 *__restrict__ Tout, int *__restrict__ status)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    tmatrix_one<MB>(g, G0w, V, e, Tout, status);
}

constexpr int kBulkThreads = 256;

// J = false: bulk[e][pix] = sum w_o1 w_o2 Im G0(R2-R1)_{o2,o1} (double).
// J = true: the Josephson map's complex table, column o1 + maxband (double2).
// One workgroup per (sub-pixel, energy); the tile values w_o(x - R) of its
// sub-pixel are staged in LDS first ([o][R], maxb (2w+1)^2 doubles).
// MB > 0: the workgroups with blockIdx.x == os^2 compute the T-matrix of
// their energy instead (thread 0), so that one launch serves both.
template <bool J, int MB>
__global__ void __launch_bounds__(kBulkThreads) k_bulk(Geom g, const double2 *__restrict__ G0w,
                                                      const double *__restrict__ tiles, void *__restrict__ out,
                                                      const double2 *__restrict__ V, double2 *__restrict__ Tout,
                                                      int *__restrict__ status)
{
    __shared__ double red[kBulkThreads], redr[J ? kBulkThreads : 1];
    extern __shared__ double wv[];
    const int pix = blockIdx.x, e = blockIdx.y;
    const int os = g.os, w = g.w, T = g.T, B = g.B, NP = B * B, nw = g.nw, h = g.h;
    if constexpr (MB > 0) {
        if (pix == os * os) {
            if (threadIdx.x == 0) tmatrix_one<MB>(g, G0w, V, e, Tout, status);
            return;
        }
    }
    const int ii = pix / os, jj = pix - ii * os;
    const int x = ii + w * os, y = jj + w * os;
    const int side = 2 * w + 1, NW = side * side;
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    const int cofs = J ? g.maxb : 0;
    for (int t = threadIdx.x; t < g.maxb * NW; t += blockDim.x) {
        const int o = t / NW, a = t - o * NW;
        const int rx = a / side - w, ry = a - (a / side) * side - w;
        wv[t] = tiles[((size_t)o * T + (x - rx * os)) * T + (y - ry * os)];
    }
    __syncthreads();
    const double2 *G0 = G0w + (size_t)e * nw * nw * NP;
    double acc = 0.0, accr = 0.0;
    for (int t = threadIdx.x; t < NW * NW; t += blockDim.x) {
        int a1 = t / NW, a2 = t - a1 * NW;
        int r1x = a1 / side - w, r1y = a1 % side - w;
        int r2x = a2 / side - w, r2y = a2 % side - w;
        const double2 *G = G0 + ((size_t)(r2x - r1x + h) * nw + (r2y - r1y + h)) * NP;
        for (int sg = 0; sg < nsets; sg++) {
            int off = sg * nper;
            for (int o1 = off; o1 < off + nper; o1++) {
                const double f1 = wv[o1 * NW + a1];
                if (f1 == 0.0) continue;
                for (int o2 = off; o2 < off + nper; o2++) {
                    const double f2 = wv[o2 * NW + a2];
                    const double2 gv = G[o2 * B + o1 + cofs];
                    acc += f1 * f2 * gv.y;
                    if (J) accr += f1 * f2 * gv.x;
                }
            }
        }
    }
    red[threadIdx.x] = acc;
    if (J) redr[threadIdx.x] = accr;
    __syncthreads();
    for (int s = kBulkThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] += red[threadIdx.x + s];
            if (J) redr[threadIdx.x] += redr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (J)
            static_cast<double2 *>(out)[(size_t)e * os * os + pix] = make_double2(redr[0], red[0]);
        else
            static_cast<double *>(out)[(size_t)e * os * os + pix] = red[0];
    }
}

size_t bulk_lds(const Geom &g) { return (size_t)g.maxb * (2 * g.w + 1) * (2 * g.w + 1) * sizeof(double); }

}  // namespace

namespace cqpi {

int launch_tmatrix(const Geom &g, const double2 *G0w, const double2 *V, int E, double2 *T, int *status,
                   hipStream_t s)
{
    if (g.B > kMaxB) return CQPI_E_ARG;
    int threads = 64;
    const dim3 grid((E + threads - 1) / threads);
    switch (g.B) {
    case 1: hipLaunchKernelGGL(k_tmatrix<1>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    case 2: hipLaunchKernelGGL(k_tmatrix<2>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    case 3: hipLaunchKernelGGL(k_tmatrix<3>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    case 4: hipLaunchKernelGGL(k_tmatrix<4>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    case 5: hipLaunchKernelGGL(k_tmatrix<5>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    default: hipLaunchKernelGGL(k_tmatrix<kMaxB>, grid, dim3(threads), 0, s, g, G0w, V, E, T, status); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_bulk(const Geom &g, const double2 *G0w, const double *tiles, int E, double *bulk,
                hipStream_t s)
{
    hipLaunchKernelGGL((k_bulk<false, 0>), dim3(g.os * g.os, E), dim3(kBulkThreads), bulk_lds(g), s, g, G0w, tiles,
                       bulk, nullptr, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_bulk_josephson(const Geom &g, const double2 *G0w, const double *tiles, int E, double2 *bulk,
                          hipStream_t s)
{
    if (2 * g.maxb > g.B) return CQPI_E_ARG;
    hipLaunchKernelGGL((k_bulk<true, 0>), dim3(g.os * g.os, E), dim3(kBulkThreads), bulk_lds(g), s, g, G0w, tiles,
                       bulk, nullptr, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

// the T-matrix and the periodic (bulk) term of E energies in one launch
int launch_tmatrix_bulk(const Geom &g, const double2 *G0w, const double2 *V, int E, double2 *T, int *status,
                        const double *tiles, void *bulk, bool josephson, hipStream_t s)
{
    if (g.B > kMaxB || (josephson && 2 * g.maxb > g.B)) return CQPI_E_ARG;
    const dim3 grid(g.os * g.os + 1, E), blk(kBulkThreads);
    const size_t lds = bulk_lds(g);
#define CQPI_TB(mb)                                                                                               \
    if (josephson)                                                                                                \
        hipLaunchKernelGGL((k_bulk<true, mb>), grid, blk, lds, s, g, G0w, tiles, bulk, V, T, status);             \
    else                                                                                                          \
        hipLaunchKernelGGL((k_bulk<false, mb>), grid, blk, lds, s, g, G0w, tiles, bulk, V, T, status);
    switch (g.B) {
    case 1: CQPI_TB(1) break;
    case 2: CQPI_TB(2) break;
    case 3: CQPI_TB(3) break;
    case 4: CQPI_TB(4) break;
    case 5: CQPI_TB(5) break;
    default: CQPI_TB(kMaxB) break;
    }
#undef CQPI_TB
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}


}  // namespace cqpi

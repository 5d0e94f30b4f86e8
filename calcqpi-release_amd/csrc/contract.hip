// Continuum contraction (FP64): rho(r, w) for every sub-pixel of the map.
//
// The reference evaluates, per cell pos and sub-pixel x (src/qpi.cc:889-945,
// src/qpihipkernel.cc:60-100),
//   rho = - sum_{R1,R2 in [-w,w]^2} sum_{o1,o2} w_o1(x-R1) w_o2(x-R2)
//           Im[ (G0(pos+R2) T G0(-pos-R1))_{o2,o1} + G0(R2-R1)_{o2,o1} ]
// with (2w+1)^2 + (2w+1)^4 b x b complex products per cell.  The sum factorises
// exactly: with
//   a_p(x) = sum_{R2,o2} w_o2(x-R2) G0(pos+R2)_{o2,p}
//   b_q(x) = sum_{R1,o1} G0(-pos-R1)_{q,o1} w_o1(x-R1)
// the impurity part is -Im sum_{p,q} a_p T_pq b_q and the periodic part is the
// per-sub-pixel bulk table.  With spin, o1 and o2 run over one spin block at a
// time (src/qpi.cc:924-940) and the two blocks are summed.
//
// Mapping: a workgroup covers an 8 x 8 block of cells (one cell per lane) and
// 4 waves; each wave takes groups of P sub-pixels, so the tile value
// w_o(x-R) is uniform across a wave while G0 rows vary per lane.  Each lane
// accumulates a and b for its P sub-pixels in registers.
#include "cqpi_common.h"

namespace {

constexpr int kCells = 8;     // cells per block side
constexpr int kWaves = 4;

template <int B, int P>
__global__ void __launch_bounds__(64 * kWaves) k_contract64(Geom g, const double2 *__restrict__ G0w,
                                                          const double2 *__restrict__ Tm,
                                                          const double *__restrict__ bulk,
                                                          const double *__restrict__ tiles,
                                                          double *__restrict__ ldos)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx = blockIdx.x * kCells + (lane & (kCells - 1));
    const int cy = blockIdx.y * kCells + (lane >> 3);
    const int e = blockIdx.z;
    if (cx >= g.n || cy >= g.n) return;
    const int os = g.os, w = g.w, T = g.T, nw = g.nw, h = g.h, W = g.W;
    const int NP = B * B;
    const int n2 = g.n >> 1;
    const int px = cx - n2, py = cy - n2;
    const int npix = os * os;
    const int npg = (npix + P - 1) / P;
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    const double2 *G0 = G0w + (size_t)e * nw * nw * NP;
    const double2 *Te = Tm + (size_t)e * NP;
    double *out = ldos + (size_t)e * W * W;

    for (int pg = wave; pg < npg; pg += kWaves) {
        int ii[P], jj[P];
        bool ok[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            int pix = pg * P + p;
            ok[p] = pix < npix;
            pix = ok[p] ? pix : 0;
            ii[p] = pix / os;
            jj[p] = pix - ii[p] * os;
        }
        double val[P];
#pragma unroll
        for (int p = 0; p < P; p++) val[p] = 0.0;
        for (int sg = 0; sg < nsets; sg++) {
            const int olo = sg * nper, ohi = olo + nper;
            double2 acca[P][B], accb[P][B];
#pragma unroll
            for (int p = 0; p < P; p++)
#pragma unroll
                for (int q = 0; q < B; q++) {
                    acca[p][q] = make_double2(0.0, 0.0);
                    accb[p][q] = make_double2(0.0, 0.0);
                }
            for (int rx = -w; rx <= w; rx++) {
                for (int ry = -w; ry <= w; ry++) {
                    const double2 *GA = G0 + ((size_t)(px + rx + h) * nw + (py + ry + h)) * NP;
                    const double2 *GB = G0 + ((size_t)(-px - rx + h) * nw + (-py - ry + h)) * NP;
                    const int tx = (w - rx) * os, ty = (w - ry) * os;
                    for (int o = olo; o < ohi; o++) {
                        double wt[P];
#pragma unroll
                        for (int p = 0; p < P; p++)
                            wt[p] = tiles[((size_t)o * T + ii[p] + tx) * T + jj[p] + ty];
#pragma unroll
                        for (int q = 0; q < B; q++) {
                            double2 ga = GA[o * B + q];
                            double2 gb = GB[q * B + o];
#pragma unroll
                            for (int p = 0; p < P; p++) {
                                acca[p][q].x = fma(wt[p], ga.x, acca[p][q].x);
                                acca[p][q].y = fma(wt[p], ga.y, acca[p][q].y);
                                accb[p][q].x = fma(wt[p], gb.x, accb[p][q].x);
                                accb[p][q].y = fma(wt[p], gb.y, accb[p][q].y);
                            }
                        }
                    }
                }
            }
            // Im sum_pq a_p T_pq b_q
#pragma unroll
            for (int q1 = 0; q1 < B; q1++) {
                double2 trow[B];
#pragma unroll
                for (int q2 = 0; q2 < B; q2++) trow[q2] = Te[q1 * B + q2];
#pragma unroll
                for (int p = 0; p < P; p++) {
                    double2 tb = make_double2(0.0, 0.0);
#pragma unroll
                    for (int q2 = 0; q2 < B; q2++) tb = cadd(tb, cmul(trow[q2], accb[p][q2]));
                    val[p] += acca[p][q1].x * tb.y + acca[p][q1].y * tb.x;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (!ok[p]) continue;
            int X = cx * os + ii[p], Y = cy * os + jj[p];
            out[(size_t)X + (size_t)Y * W] = -(val[p] + bulk[(size_t)e * npix + ii[p] * os + jj[p]]);
        }
    }
}

template <int B>
int launch_c64(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
               const double *tiles, int E, double *ldos, hipStream_t s)
{
    constexpr int P = B <= 1 ? 8 : (B <= 3 ? 4 : (B <= 6 ? 2 : 1));
    int nb = (g.n + kCells - 1) / kCells;
    hipLaunchKernelGGL((k_contract64<B, P>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0w, T, bulk,
                       tiles, ldos);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace

namespace cqpi {

int launch_contract_fp64(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
                         const double *tiles, int E, double *ldos, hipStream_t s)
{
    switch (g.B) {
#define CASE(b) \
    case b: return launch_c64<b>(g, G0w, T, bulk, tiles, E, ldos, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

}  // namespace cqpi

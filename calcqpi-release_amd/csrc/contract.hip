// Continuum contraction (FP64, and the same loop nest in FP32): rho(r, w) for
// every sub-pixel of the map.
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
// accumulates a and b for its P sub-pixels in registers.  k_contract_st stages
// the workgroup's window rows in LDS and is the one launched; k_contract,
// which gathers them from memory per lane, serves windows whose staged rows
// exceed the LDS.
#include "cqpi_common.h"

namespace {

constexpr int kCells = 8;     // cells per block side
constexpr int kWaves = 4;

template <typename Real>
struct Cplx;
template <>
struct Cplx<double> {
    using type = double2;
};
template <>
struct Cplx<float> {
    using type = float2;
};

// Real = double: the FP64 contraction.  Real = float: the same loop nest in
// FP32 (G0 window, T and tiles narrowed to FP32, FP32 accumulation; the
// periodic table stays FP64).
//
// J: the Josephson map (the b gather from the particle-hole column o +
// maxband, complex periodic table and tip factor; as k_contract_st below), for
// windows whose staged rows do not fit the LDS.
template <typename Real, int B, int P, bool J>
__global__ void __launch_bounds__(64 * kWaves) k_contract(Geom g, const typename Cplx<Real>::type *__restrict__ G0w,
                                                        const typename Cplx<Real>::type *__restrict__ Tm,
                                                        const double *__restrict__ bulk,
                                                        const Real *__restrict__ tiles, double *__restrict__ ldos,
                                                        const double2 *__restrict__ tips)
{
    using C = typename Cplx<Real>::type;
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
    const C *G0 = G0w + (size_t)e * nw * nw * NP;
    const C *Te = Tm + (size_t)e * NP;
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
        Real val[P], valr[P];  // Im and (J) Re of sum_s a^T T b
#pragma unroll
        for (int p = 0; p < P; p++) val[p] = valr[p] = 0;
        const int jo = J ? g.maxb : 0;  // column offset of the b gather
        for (int sg = 0; sg < nsets; sg++) {
            const int olo = sg * nper, ohi = olo + nper;
            C acca[P][B], accb[P][B];
#pragma unroll
            for (int p = 0; p < P; p++)
#pragma unroll
                for (int q = 0; q < B; q++) {
                    acca[p][q].x = acca[p][q].y = 0;
                    accb[p][q].x = accb[p][q].y = 0;
                }
            for (int rx = -w; rx <= w; rx++) {
                for (int ry = -w; ry <= w; ry++) {
                    const C *GA = G0 + ((size_t)(px + rx + h) * nw + (py + ry + h)) * NP;
                    const C *GB = G0 + ((size_t)(-px - rx + h) * nw + (-py - ry + h)) * NP;
                    const int tx = (w - rx) * os, ty = (w - ry) * os;
                    for (int o = olo; o < ohi; o++) {
                        Real wt[P];
#pragma unroll
                        for (int p = 0; p < P; p++)
                            wt[p] = tiles[((size_t)o * T + ii[p] + tx) * T + jj[p] + ty];
#pragma unroll
                        for (int q = 0; q < B; q++) {
                            const C ga = GA[o * B + q];
                            const C gb = GB[q * B + o + jo];
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
                C trow[B];
#pragma unroll
                for (int q2 = 0; q2 < B; q2++) trow[q2] = Te[q1 * B + q2];
#pragma unroll
                for (int p = 0; p < P; p++) {
                    Real tbx = 0, tby = 0;
#pragma unroll
                    for (int q2 = 0; q2 < B; q2++) {
                        tbx += trow[q2].x * accb[p][q2].x - trow[q2].y * accb[p][q2].y;
                        tby += trow[q2].x * accb[p][q2].y + trow[q2].y * accb[p][q2].x;
                    }
                    val[p] += acca[p][q1].x * tby + acca[p][q1].y * tbx;
                    if (J) valr[p] += acca[p][q1].x * tbx - acca[p][q1].y * tby;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (!ok[p]) continue;
            int X = cx * os + ii[p], Y = cy * os + jj[p];
            if (J) {
                const double2 bz = reinterpret_cast<const double2 *>(bulk)[(size_t)e * npix + ii[p] * os + jj[p]];
                const double2 tp = tips[e];
                const double zr = (double)valr[p] + bz.x, zi = (double)val[p] + bz.y;
                out[(size_t)X + (size_t)Y * W] = -(tp.x * zi + tp.y * zr);
            } else {
                out[(size_t)X + (size_t)Y * W] = -((double)val[p] + bulk[(size_t)e * npix + ii[p] * os + jj[p]]);
            }
        }
    }
}

// The same loop nest with the window rows staged in LDS.  Per spin block and
// orbital o, the workgroup copies G0(pos+R)_{o,q} and G0(-pos-R)_{q,o} for its
// 8 x 8 cells and every |R| <= w (two (8+2w)^2 x b blocks) into LDS; the lanes
// then read them from there instead of gathering a b^2 slot block per lane
// from memory, where neighbouring cells re-read each other's slots.  The
// copies are redone for each group of sub-pixels the waves take.
//
// J: the Josephson map of a superconducting model (src/qpihipkernel.cc:171-223,
// spin :225-271; CPU src/qpi.cc:953-1033).  The b gather reads the
// particle-hole column o + maxband, and per energy
//   rho = - sum_s Im[ tip (a^T T b' + bulk') ]
// with the complex periodic table bulk' (bulk holds E x os^2 double2) and the
// tip factor tips[e].
template <typename Real, int B, int P, bool J>
__global__ void __launch_bounds__(64 * kWaves) k_contract_st(Geom g, const typename Cplx<Real>::type *__restrict__ G0w,
                                                           const typename Cplx<Real>::type *__restrict__ Tm,
                                                           const double *__restrict__ bulk,
                                                           const Real *__restrict__ tiles, double *__restrict__ ldos,
                                                           const double2 *__restrict__ tips)
{
    using C = typename Cplx<Real>::type;
    extern __shared__ __align__(16) unsigned char smem_c[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx0 = blockIdx.x * kCells, cy0 = blockIdx.y * kCells;
    const int cx = cx0 + (lane & (kCells - 1));
    const int cy = cy0 + (lane >> 3);
    const int e = blockIdx.z;
    const bool cell_ok = cx < g.n && cy < g.n;
    const int os = g.os, w = g.w, T = g.T, nw = g.nw, h = g.h, W = g.W;
    constexpr int NP = B * B;
    const int n2 = g.n >> 1;
    const int SA = kCells + 2 * w;
    C *GAs = reinterpret_cast<C *>(smem_c);   // [lx][ly][q]
    C *GBs = GAs + (size_t)SA * SA * B;       // [mx][my][q]
    const int npix = os * os;
    const int npg = (npix + P - 1) / P;
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    const C *G0 = G0w + (size_t)e * nw * nw * NP;
    const C *Te = Tm + (size_t)e * NP;
    double *out = ldos + (size_t)e * W * W;
    // slot origins of the staged blocks: GA slot (ax0 + lx, ay0 + ly), GB slot (bx0 + mx, by0 + my)
    const int ax0 = cx0 - n2 - w + h, ay0 = cy0 - n2 - w + h;
    const int bx0 = h - (cx0 - n2) - (kCells - 1) - w, by0 = h - (cy0 - n2) - (kCells - 1) - w;
    const int lxo = cx - cx0, lyo = cy - cy0;              // this cell in the block

    for (int it = 0; it * kWaves < npg; it++) {
        const int pg = it * kWaves + wave;
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
        Real val[P], valr[P];  // Im and (J) Re of sum_s a^T T b
#pragma unroll
        for (int p = 0; p < P; p++) val[p] = valr[p] = 0;
        for (int sg = 0; sg < nsets; sg++) {
            const int olo = sg * nper, ohi = olo + nper;
            C acca[P][B], accb[P][B];
#pragma unroll
            for (int p = 0; p < P; p++)
#pragma unroll
                for (int q = 0; q < B; q++) {
                    acca[p][q] = C{0, 0};
                    accb[p][q] = C{0, 0};
                }
            for (int o = olo; o < ohi; o++) {
                __syncthreads();  // the previous orbital's rows are read
                stage_window_rows<C, B>(G0, nw, o, J ? o + g.maxb : o, SA, ax0, ay0, bx0, by0, GAs, GBs);
                __syncthreads();
                if (pg < npg && cell_ok) {
                    for (int rx = -w; rx <= w; rx++) {
                        for (int ry = -w; ry <= w; ry++) {
                            // GA slot (pos + R), GB slot (-pos - R), in block coordinates
                            const C *ga = GAs + ((size_t)(lxo + rx + w) * SA + (lyo + ry + w)) * B;
                            const C *gb = GBs + ((size_t)(kCells - 1 - lxo - rx + w) * SA +
                                                       (kCells - 1 - lyo - ry + w)) * B;
                            const int tx = (w - rx) * os, ty = (w - ry) * os;
                            Real wt[P];
#pragma unroll
                            for (int p = 0; p < P; p++)
                                wt[p] = tiles[((size_t)o * T + ii[p] + tx) * T + jj[p] + ty];
#pragma unroll
                            for (int q = 0; q < B; q++) {
                                const C a = ga[q], b = gb[q];
#pragma unroll
                                for (int p = 0; p < P; p++) {
                                    acca[p][q].x = fma(wt[p], a.x, acca[p][q].x);
                                    acca[p][q].y = fma(wt[p], a.y, acca[p][q].y);
                                    accb[p][q].x = fma(wt[p], b.x, accb[p][q].x);
                                    accb[p][q].y = fma(wt[p], b.y, accb[p][q].y);
                                }
                            }
                        }
                    }
                }
            }
            // Im sum_pq a_p T_pq b_q
#pragma unroll
            for (int q1 = 0; q1 < B; q1++) {
#pragma unroll
                for (int p = 0; p < P; p++) {
                    Real tbx = 0, tby = 0;
#pragma unroll
                    for (int q2 = 0; q2 < B; q2++) {
                        const C tq = Te[q1 * B + q2];
                        tbx += tq.x * accb[p][q2].x - tq.y * accb[p][q2].y;
                        tby += tq.x * accb[p][q2].y + tq.y * accb[p][q2].x;
                    }
                    val[p] += acca[p][q1].x * tby + acca[p][q1].y * tbx;
                    if (J) valr[p] += acca[p][q1].x * tbx - acca[p][q1].y * tby;
                }
            }
        }
        if (pg < npg && cell_ok) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (!ok[p]) continue;
                const int X = cx * os + ii[p], Y = cy * os + jj[p];
                if (J) {
                    // -Im[tip z], z = sum_s a^T T b' + bulk'
                    const double2 bz = reinterpret_cast<const double2 *>(bulk)[(size_t)e * npix + ii[p] * os + jj[p]];
                    const double2 tp = tips[e];
                    const double zr = (double)valr[p] + bz.x, zi = (double)val[p] + bz.y;
                    out[(size_t)X + (size_t)Y * W] = -(tp.x * zi + tp.y * zr);
                } else {
                    out[(size_t)X + (size_t)Y * W] = -((double)val[p] + bulk[(size_t)e * npix + ii[p] * os + jj[p]]);
                }
            }
        }
    }
}

size_t contract_st_lds(const Geom &g, size_t csize)
{
    return (size_t)2 * (kCells + 2 * g.w) * (kCells + 2 * g.w) * g.B * csize;
}

// FP64 -> FP32 narrowing of the G0 windows and T of a chunk
__global__ void k_narrow(const double2 *__restrict__ a, size_t na, const double2 *__restrict__ t, size_t nt,
                         float2 *__restrict__ af, float2 *__restrict__ tf)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) af[i] = make_float2((float)a[i].x, (float)a[i].y);
    if (i < nt) tf[i] = make_float2((float)t[i].x, (float)t[i].y);
}

template <int B>
int launch_c64(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
               const double *tiles, int E, double *ldos, const double2 *tips, hipStream_t s)
{
    constexpr int P = B <= 1 ? 8 : (B <= 3 ? 4 : (B <= 12 ? 2 : 1));
    int nb = (g.n + kCells - 1) / kCells;
    const size_t lds = contract_st_lds(g, sizeof(double2));
    if (tips && lds > 150 * 1024) {
        hipLaunchKernelGGL((k_contract<double, B, P, true>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0w, T,
                           bulk, tiles, ldos, tips);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (tips) {
        hipLaunchKernelGGL((k_contract_st<double, B, P, true>), dim3(nb, nb, E), dim3(64 * kWaves), lds, s, g, G0w,
                           T, bulk, tiles, ldos, tips);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (lds <= 150 * 1024) {
        hipLaunchKernelGGL((k_contract_st<double, B, P, false>), dim3(nb, nb, E), dim3(64 * kWaves), lds, s, g, G0w,
                           T, bulk, tiles, ldos, nullptr);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    hipLaunchKernelGGL((k_contract<double, B, P, false>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0w, T, bulk,
                       tiles, ldos, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <int B>
int launch_c32(const Geom &g, const float2 *G0f, const float2 *Tf, const double *bulk, const float *tiles, int E,
               double *ldos, const double2 *tips, hipStream_t s)
{
    // four sub-pixel groups of 4 for os = 4 (one per wave, the rows staged once)
    constexpr int P = B <= 6 ? 4 : 2;
    int nb = (g.n + kCells - 1) / kCells;
    const size_t lds = contract_st_lds(g, sizeof(float2));
    if (tips && lds > 150 * 1024) {
        hipLaunchKernelGGL((k_contract<float, B, P, true>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0f, Tf,
                           bulk, tiles, ldos, tips);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (tips) {
        hipLaunchKernelGGL((k_contract_st<float, B, P, true>), dim3(nb, nb, E), dim3(64 * kWaves), lds, s, g, G0f, Tf,
                           bulk, tiles, ldos, tips);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (lds <= 150 * 1024) {
        hipLaunchKernelGGL((k_contract_st<float, B, P, false>), dim3(nb, nb, E), dim3(64 * kWaves), lds, s, g, G0f,
                           Tf, bulk, tiles, ldos, nullptr);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    hipLaunchKernelGGL((k_contract<float, B, P, false>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0f, Tf, bulk,
                       tiles, ldos, nullptr);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

// ---------------------------------------------------------------------------
// Threshold-list contraction (src/qpihipkernel.cc:273-315, CPU src/qpi.cc:1088-1135):
//   rho(pos, x) = - sum_(R1,R2) sum_(o1,o2) F[R1,R2,o2,o1,x]
//                   Im[ (G0(pos+R2) T G0(-pos-R1))_{o2,o1} + G0(R2-R1)_{o2,o1} ]
// with F the list's factors summed per active window pair (pair k = (r1, r2)
// = lpair[k], pairs of one r1 contiguous from r1off[r1]) and per orbital pair
// present: combos coff[k B + o1] .. coff[k B + o1 + 1] - 1 hold o2 = co2[cb]
// and the factors of every sub-pixel, ccoef[cb][pix] (the periodic part
// reads the dense table coef[k][o2][o1][pix]).  The list selects arbitrary (R1, o1, R2, o2)
// pairs, so the sum does not factorise as the dense map does: per pair the
// lane forms C = G0(pos+R2) Y with Y = T G0(-pos-R1) kept from its R1.
// Lanes are the 8 x 8 cells, the waves split the R1 values, the sub-pixels
// are accumulated in registers (at most kListPx per pass) and summed over the
// waves in LDS.  The periodic part (position independent) comes from
// k_bulk_list.
// ---------------------------------------------------------------------------
constexpr int kListPx = 16;

// FP64 for B <= 5 and os <= 4, the loads regrouped: per active
// pair the lane loads its b x b block G0(pos+R2) at once (independent loads)
// and forms, per o1 present, Im (G0(pos+R2) Y)_{o2,o1} for every o2; a combo
// then selects its o2, so no load waits on the combo's orbital.  The combo
// factors are uniform (scalar) loads of a compile-time count NPIX = os^2, so
// that they issue together (a run-time bound puts a branch, and a wait,
// before each).
template <int B, int NPIX>
__global__ void __launch_bounds__(64 * kWaves) k_contract_list_s(Geom g, const double2 *__restrict__ G0w,
                                                               const double2 *__restrict__ Tm,
                                                               const double *__restrict__ bulk,
                                                               const int *__restrict__ coff,
                                                               const int *__restrict__ co2,
                                                               const double *__restrict__ ccoef,
                                                               const int *__restrict__ r1off,
                                                               const int *__restrict__ lpair,
                                                               double *__restrict__ ldos)
{
    __shared__ double red[kWaves - 1][NPIX][64];
    constexpr int NP = B * B;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx = blockIdx.x * kCells + (lane & (kCells - 1));
    const int cy = blockIdx.y * kCells + (lane >> 3);
    const int e = blockIdx.z;
    const bool cell_ok = cx < g.n && cy < g.n;
    const int os = g.os, w = g.w, nw = g.nw, h = g.h, W = g.W;
    const int side = 2 * w + 1, NR = side * side;
    const int n2 = g.n >> 1;
    const int px = (cell_ok ? cx : 0) - n2, py = (cell_ok ? cy : 0) - n2;
    const double2 *G0 = G0w + (size_t)e * nw * nw * NP;
    const double2 *Te = Tm + (size_t)e * NP;
    double *out = ldos + (size_t)e * W * W;

    double val[NPIX];
#pragma unroll
    for (int q = 0; q < NPIX; q++) val[q] = 0.0;
    for (int r1 = wave; r1 < NR; r1 += kWaves) {
        const int k0 = r1off[r1], k1 = r1off[r1 + 1];
        if (k0 == k1) continue;
        const int rx1 = r1 / side - w, ry1 = r1 % side - w;
        const double2 *GB = G0 + ((size_t)(-px - rx1 + h) * nw + (-py - ry1 + h)) * NP;
        double2 Y[NP];  // T G0(-pos-R1)
#pragma unroll
        for (int q = 0; q < B; q++)
#pragma unroll
            for (int o1 = 0; o1 < B; o1++) {
                double2 acc = make_double2(0.0, 0.0);
#pragma unroll
                for (int p = 0; p < B; p++) acc = cadd(acc, cmul(Te[q * B + p], GB[p * B + o1]));
                Y[q * B + o1] = acc;
            }
        for (int k = k0; k < k1; k++) {
            const int r2 = lpair[k] - r1 * NR;
            const int rx2 = r2 / side - w, ry2 = r2 % side - w;
            const double2 *GA = G0 + ((size_t)(px + rx2 + h) * nw + (py + ry2 + h)) * NP;
            double2 ga[NP];
#pragma unroll
            for (int t = 0; t < NP; t++) ga[t] = GA[t];
#pragma unroll
            for (int o1 = 0; o1 < B; o1++) {
                const int c0 = coff[k * B + o1], c1 = coff[k * B + o1 + 1];
                if (c0 == c1) continue;
                double cimv[B];  // Im (G0(pos+R2) Y)_{o2,o1} for every o2
#pragma unroll
                for (int o2 = 0; o2 < B; o2++) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < B; q++)
                        acc += ga[o2 * B + q].x * Y[q * B + o1].y + ga[o2 * B + q].y * Y[q * B + o1].x;
                    cimv[o2] = acc;
                }
                for (int cb = c0; cb < c1; cb++) {
                    const int o2 = co2[cb];
                    double cim = cimv[0];
#pragma unroll
                    for (int o = 1; o < B; o++)
                        if (o2 == o) cim = cimv[o];
                    const double *cp = ccoef + (size_t)cb * NPIX;
#pragma unroll
                    for (int q = 0; q < NPIX; q++) val[q] += cp[q] * cim;
                }
            }
        }
    }
    __syncthreads();
    if (wave > 0)
#pragma unroll
        for (int q = 0; q < NPIX; q++) red[wave - 1][q][lane] = val[q];
    __syncthreads();
    if (wave == 0 && cell_ok) {
#pragma unroll
        for (int q = 0; q < NPIX; q++) {
            double v = val[q];
#pragma unroll
            for (int v2 = 0; v2 < kWaves - 1; v2++) v += red[v2][q][lane];
            const int ii = q / os, jj = q - ii * os;
            out[(size_t)(cx * os + ii) + (size_t)(cy * os + jj) * W] = -(v + bulk[(size_t)e * NPIX + q]);
        }
    }
}

// The list map for any band count up to kListMaxB, in FP64 or (C = float2,
// windows and T narrowed) FP32: the same sum with one column o1 of
// Y = T G0(-pos-R1) in registers at a time, so that its size does not grow
// with b^2; the row G0(pos+R2)_{o2,.} is then read once per o1 (from cache).
constexpr int kListMaxB = 16;

// NPX > 0: os^2 = NPX <= 16 sub-pixels, the factor loads of a combo
// unconditional (they issue together); 0: any os, in passes of kListPx
template <typename C, int NPX>
__global__ void __launch_bounds__(64 * kWaves) k_contract_list_g(Geom g, const C *__restrict__ G0w,
                                                               const C *__restrict__ Tm,
                                                               const double *__restrict__ bulk,
                                                               const int *__restrict__ coff,
                                                               const int *__restrict__ co2,
                                                               const double *__restrict__ ccoef,
                                                               const int *__restrict__ r1off,
                                                               const int *__restrict__ lpair, double *__restrict__ ldos)
{
    using Re = decltype(C{}.x);
    constexpr int PX = NPX > 0 ? NPX : kListPx;  // sub-pixels per pass
    __shared__ double red[kWaves - 1][PX][64];
    const int B = g.B, NP = B * B;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx = blockIdx.x * kCells + (lane & (kCells - 1));
    const int cy = blockIdx.y * kCells + (lane >> 3);
    const int e = blockIdx.z;
    const bool cell_ok = cx < g.n && cy < g.n;
    const int os = g.os, w = g.w, nw = g.nw, h = g.h, W = g.W;
    const int side = 2 * w + 1, NR = side * side, npix = os * os;
    const int n2 = g.n >> 1;
    // cells past the map edge compute with cell 0 (in the window) and store nothing
    const int px = (cell_ok ? cx : 0) - n2, py = (cell_ok ? cy : 0) - n2;
    const C *G0 = G0w + (size_t)e * nw * nw * NP;
    const C *Te = Tm + (size_t)e * NP;
    double *out = ldos + (size_t)e * W * W;

    for (int pc = 0; pc < npix; pc += PX) {
        double val[PX];
#pragma unroll
        for (int q = 0; q < PX; q++) val[q] = 0.0;
        for (int r1 = wave; r1 < NR; r1 += kWaves) {
            const int rx1 = r1 / side - w, ry1 = r1 % side - w;
            const C *GB = G0 + ((size_t)(-px - rx1 + h) * nw + (-py - ry1 + h)) * NP;
            const int k0 = r1off[r1], k1 = r1off[r1 + 1];
            if (k0 == k1) continue;
            for (int o1 = 0; o1 < B; o1++) {
                C Y[kListMaxB];  // column o1 of T G0(-pos-R1)
#pragma unroll
                for (int q = 0; q < kListMaxB; q++) {
                    Re yx = 0, yy = 0;
                    if (q < B)
                        for (int p = 0; p < B; p++) {
                            const C t = Te[q * B + p], gb = GB[p * B + o1];
                            yx += t.x * gb.x - t.y * gb.y;
                            yy += t.x * gb.y + t.y * gb.x;
                        }
                    Y[q].x = yx;
                    Y[q].y = yy;
                }
                for (int k = k0; k < k1; k++) {
                    const int c1 = coff[k * B + o1 + 1];
                    if (coff[k * B + o1] == c1) continue;
                    const int r2 = lpair[k] - r1 * NR;
                    const int rx2 = r2 / side - w, ry2 = r2 % side - w;
                    const C *GA = G0 + ((size_t)(px + rx2 + h) * nw + (py + ry2 + h)) * NP;
                    // only the o2 present with this o1 in the pair's entries
                    for (int cb = coff[k * B + o1]; cb < c1; cb++) {
                        const int o2 = co2[cb];
                        Re cim = 0;  // Im (G0(pos+R2) Y)_{o2,o1}
#pragma unroll
                        for (int q = 0; q < kListMaxB; q++)
                            if (q < B) {
                                const C a = GA[o2 * B + q];
                                cim += a.x * Y[q].y + a.y * Y[q].x;
                            }
                        const double *cp = ccoef + (size_t)cb * (NPX > 0 ? NPX : npix) + pc;
#pragma unroll
                        for (int q = 0; q < PX; q++)
                            if (NPX > 0 || pc + q < npix) val[q] += cp[q] * (double)cim;
                    }
                }
            }
        }
        __syncthreads();  // the previous pass's partial sums are read
        if (wave > 0)
#pragma unroll
            for (int q = 0; q < PX; q++) red[wave - 1][q][lane] = val[q];
        __syncthreads();
        if (wave == 0 && cell_ok) {
#pragma unroll
            for (int q = 0; q < PX; q++) {
                const int pix = pc + q;
                if (pix >= npix) break;
                double v = val[q];
#pragma unroll
                for (int v2 = 0; v2 < kWaves - 1; v2++) v += red[v2][q][lane];
                const int ii = pix / os, jj = pix - ii * os;
                out[(size_t)(cx * os + ii) + (size_t)(cy * os + jj) * W] = -(v + bulk[(size_t)e * npix + pix]);
            }
        }
    }
}

// periodic part of the list map: bulk[e][pix] = sum_(pairs, o1, o2) F Im G0(R2-R1)_{o2,o1}
constexpr int kListBulkThreads = 256;

__global__ void __launch_bounds__(kListBulkThreads) k_bulk_list(Geom g, const double2 *__restrict__ G0w,
                                                              const double *__restrict__ coef,
                                                              const int *__restrict__ lpair, int nact,
                                                              double *__restrict__ bulk)
{
    __shared__ double red[kListBulkThreads];
    const int pix = blockIdx.x, e = blockIdx.y;
    const int B = g.B, NP = B * B, nw = g.nw, h = g.h, w = g.w;
    const int side = 2 * w + 1, NR = side * side, npix = g.os * g.os;
    const double2 *G0 = G0w + (size_t)e * nw * nw * NP;
    double acc = 0.0;
    for (int t = threadIdx.x; t < nact * NP; t += blockDim.x) {
        const int k = t / NP, oo = t - k * NP;
        const int r1 = lpair[k] / NR, r2 = lpair[k] - r1 * NR;
        const int dx = (r2 / side) - (r1 / side), dy = (r2 % side) - (r1 % side);  // R2 - R1
        acc += coef[((size_t)k * NP + oo) * npix + pix] * G0[((size_t)(dx + h) * nw + (dy + h)) * NP + oo].y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kListBulkThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) bulk[(size_t)e * npix + pix] = red[0];
}

template <typename C>
void list_g(const Geom &g, int nb, int E, const C *G0w, const C *T, const double *bulk, const int *coff,
            const int *co2, const double *ccoef, const int *r1off, const int *lpair, double *ldos, hipStream_t s)
{
    const dim3 grid(nb, nb, E), blk(64 * kWaves);
    switch (g.os) {
    case 1: hipLaunchKernelGGL((k_contract_list_g<C, 1>), grid, blk, 0, s, g, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos); break;
    case 2: hipLaunchKernelGGL((k_contract_list_g<C, 4>), grid, blk, 0, s, g, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos); break;
    case 3: hipLaunchKernelGGL((k_contract_list_g<C, 9>), grid, blk, 0, s, g, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos); break;
    case 4: hipLaunchKernelGGL((k_contract_list_g<C, 16>), grid, blk, 0, s, g, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos); break;
    default: hipLaunchKernelGGL((k_contract_list_g<C, 0>), grid, blk, 0, s, g, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos); break;
    }
}

}  // namespace

namespace cqpi {

int launch_contract_fp64(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
                         const double *tiles, int E, double *ldos, hipStream_t s, const double2 *tips)
{
    if (tips && 2 * g.maxb > g.B) return CQPI_E_ARG;
    switch (g.B) {
#define CASE(b) \
    case b: return launch_c64<b>(g, G0w, T, bulk, tiles, E, ldos, tips, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

int launch_contract_list(const Geom &g, const double2 *G0w, const double2 *T, const double *coef,
                         const int *coff, const int *co2, const double *ccoef,
                         const int *r1off, const int *lpair, int nact, int E, double *bulk, double *ldos,
                         hipStream_t s, float2 *G0f, float2 *Tf)
{
    if (g.B > kListMaxB) return CQPI_E_ARG;
    hipLaunchKernelGGL(k_bulk_list, dim3(g.os * g.os, E), dim3(kListBulkThreads), 0, s, g, G0w, coef, lpair, nact,
                       bulk);
    if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    const int nb = (g.n + kCells - 1) / kCells;
    if (G0f) {
        // FP32: windows and T narrowed, the periodic part stays FP64
        const size_t na = (size_t)E * g.nw * g.nw * g.B * g.B, nt = (size_t)E * g.B * g.B;
        hipLaunchKernelGGL(k_narrow, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, G0w, na, T, nt, G0f, Tf);
        if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
        list_g<float2>(g, nb, E, (const float2 *)G0f, (const float2 *)Tf, bulk, coff, co2, ccoef, r1off, lpair,
                       ldos, s);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (g.B > 5) {
        list_g<double2>(g, nb, E, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos, s);
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    if (g.os <= 4) {
        switch (g.B * 8 + g.os) {
#define CASE(b, o)                                                                                                 \
    case b * 8 + o:                                                                                                \
        hipLaunchKernelGGL((k_contract_list_s<b, o * o>), dim3(nb, nb, E), dim3(64 * kWaves), 0, s, g, G0w, T,     \
                           bulk, coff, co2, ccoef, r1off, lpair, ldos);                                            \
        break;
#define CASES(b) CASE(b, 1) CASE(b, 2) CASE(b, 3) CASE(b, 4)
            CASES(1) CASES(2) CASES(3) CASES(4) CASES(5)
#undef CASES
#undef CASE
        default: return CQPI_E_ARG;
        }
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    }
    // more sub-pixels than one pass holds: the generic kernel in passes
    list_g<double2>(g, nb, E, G0w, T, bulk, coff, co2, ccoef, r1off, lpair, ldos, s);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_contract_fp32(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
                         const float *tilesf, int E, double *ldos, float2 *G0f, float2 *Tf, hipStream_t s,
                         const double2 *tips)
{
    if (tips && 2 * g.maxb > g.B) return CQPI_E_ARG;
    const size_t na = (size_t)E * g.nw * g.nw * g.B * g.B, nt = (size_t)E * g.B * g.B;
    hipLaunchKernelGGL(k_narrow, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, G0w, na, T, nt, G0f, Tf);
    if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    switch (g.B) {
#define CASE(b) \
    case b: return launch_c32<b>(g, G0f, Tf, bulk, tilesf, E, ldos, tips, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(14) CASE(16)
#undef CASE
    default: return CQPI_E_ARG;
    }
}

}  // namespace cqpi

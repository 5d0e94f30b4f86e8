// Hand-written adjoint of the continuum-LDOS chain (FP64).
//
// Forward (per energy):  H(k) = sum_h t_h e^{+-2 pi i R_h.k} -/+ mu
//   G(k) = (w + i eta - H(k))^{-1};  G(R) = K^-2 sum_k G(k) e^{+2 pi i k.R}
//   T = (1 - V G(0))^{-1} V;  rho(x) = -Im sum_s a^T T b - bulk(x)
// with a_p = sum w_o G(pos+R)_{op}, b_q = sum G(-pos-R)_{qo} w_o,
// bulk = sum w_o1 w_o2 Im G(R2-R1)_{o2 o1}.
// Given rhobar = dL/drho, this returns dL/dt_h (real and imaginary parts),
// dL/dmu and dL/dV.  Complex adjoints use dL = Re sum conj(Zbar) dZ:
//   sbar = -i rhobar;  abar_p = sbar conj((T b)_p);  bbar_q = sbar conj((a^T T)_q)
//   Tbar += sbar conj(a b^T)
//   Gbar(pos+R)_{op} += sum_x w_o abar_p;  Gbar(-pos-R)_{qo} += sum_x w_o bbar_q
//   Gbar(R2-R1)_{o2 o1} += -i rhobar w_o1 w_o2            (periodic term)
//   Gbar(0) += T^H Tbar T^H;  Vbar = M^-H Tbar (1 + G(0) T)^H, M = 1 - V G(0)
//   Gbar(k) = K^-2 sum_R Gbar(R) e^{-2 pi i k.R}           (adjoint transform)
//   Hbar(k) = G(k)^H Gbar(k) G(k)^H
//   dL/dt_h = sum_k conj(Hbar_{o1 o2}(k)) phi_h(k);  dL/dmu = -/+ Re tr Hbar.
// G(k) is never stored: the last pass rebuilds it per energy from the cached
// eigen data.  The R-window of G and T come from the forward run.
#include "cqpi_common.h"
#include "cqpi_fft.h"

namespace {

using cqpi::brev;
using cqpi::imod;

constexpr int kCells = 8;
constexpr int kWaves = 4;

__device__ inline double wsum_reduce(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// abar, bbar per pixel and per spin block; partial Tbar per workgroup.  The
// gathers a, b are the forward contraction's, with each orbital's window rows
// for the workgroup's 8 x 8 cells staged in LDS (stage_window_rows; at most
// 2 (8+2w)^2 b complex values, 33 KB for the largest window and band count
// the adjoint accepts).
template <int B, int P>
__global__ void __launch_bounds__(64 * kWaves) k_adj_pix(Geom g, const double2 *__restrict__ G0w,
                                                       const double2 *__restrict__ Tm,
                                                       const double *__restrict__ tiles,
                                                       const double *__restrict__ rhobar,
                                                       double2 *__restrict__ abar, double2 *__restrict__ bbar,
                                                       double2 *__restrict__ Tpart)
{
    __shared__ double2 red[kWaves][B * B];
    extern __shared__ __align__(16) unsigned char smem_p[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cx0 = blockIdx.x * kCells, cy0 = blockIdx.y * kCells;
    const int cx = cx0 + (lane & (kCells - 1));
    const int cy = cy0 + (lane >> 3);
    const int e = blockIdx.z;
    const bool active = cx < g.n && cy < g.n;
    const int os = g.os, w = g.w, T = g.T, nw = g.nw, h = g.h, W = g.W;
    const int NP = B * B;
    const int n2 = g.n >> 1;
    const int npix = os * os;
    const int npg = (npix + P - 1) / P;
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    const double2 *G0 = G0w + (size_t)e * nw * nw * NP;
    const double2 *Te = Tm + (size_t)e * NP;
    // staged blocks: slot origins as in the forward contraction
    const int SA = kCells + 2 * w;
    double2 *GAs = reinterpret_cast<double2 *>(smem_p);
    double2 *GBs = GAs + (size_t)SA * SA * B;
    const int ax0 = cx0 - n2 - w + h, ay0 = cy0 - n2 - w + h;
    const int bx0 = h - (cx0 - n2) - (kCells - 1) - w, by0 = h - (cy0 - n2) - (kCells - 1) - w;
    const int lxo = cx - cx0, lyo = cy - cy0;
    double2 tacc[B * B];
#pragma unroll
    for (int i = 0; i < NP; i++) tacc[i] = make_double2(0.0, 0.0);

    // every thread runs the same number of iterations (the staging barriers)
    for (int it = 0; it * kWaves < npg; it++) {
        const int pg = it * kWaves + wave;
        const bool run = pg < npg;
        int ii[P], jj[P];
        bool ok[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            int pix = pg * P + p;
            ok[p] = active && pix < npix;
            pix = pix < npix ? pix : 0;
            ii[p] = pix / os;
            jj[p] = pix - ii[p] * os;
        }
        for (int sg = 0; sg < nsets; sg++) {
            const int olo = sg * nper, ohi = olo + nper;
            double2 a[P][B], b[P][B];
#pragma unroll
            for (int p = 0; p < P; p++)
#pragma unroll
                for (int q = 0; q < B; q++) {
                    a[p][q] = make_double2(0.0, 0.0);
                    b[p][q] = make_double2(0.0, 0.0);
                }
            for (int o = olo; o < ohi; o++) {
                __syncthreads();  // the previous orbital's rows are read
                stage_window_rows<double2, B>(G0, nw, o, o, SA, ax0, ay0, bx0, by0, GAs, GBs);
                __syncthreads();
                if (!run) continue;
                for (int rx = -w; rx <= w; rx++)
                    for (int ry = -w; ry <= w; ry++) {
                        const double2 *ga = GAs + ((size_t)(lxo + rx + w) * SA + (lyo + ry + w)) * B;
                        const double2 *gb =
                            GBs + ((size_t)(kCells - 1 - lxo - rx + w) * SA + (kCells - 1 - lyo - ry + w)) * B;
                        const int tx = (w - rx) * os, ty = (w - ry) * os;
                        double wt[P];
#pragma unroll
                        for (int p = 0; p < P; p++)
                            wt[p] = tiles[((size_t)o * T + ii[p] + tx) * T + jj[p] + ty];
#pragma unroll
                        for (int q = 0; q < B; q++) {
                            const double2 va = ga[q], vb = gb[q];
#pragma unroll
                            for (int p = 0; p < P; p++) {
                                a[p][q] = cadd(a[p][q], cscale(wt[p], va));
                                b[p][q] = cadd(b[p][q], cscale(wt[p], vb));
                            }
                        }
                    }
            }
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (!ok[p]) continue;
                const int X = cx * os + ii[p], Y = cy * os + jj[p];
                const size_t pixid = (size_t)X + (size_t)Y * W;
                const double rb = rhobar[(size_t)e * W * W + pixid];
                const double2 sbar = make_double2(0.0, -rb);
                // planar [e][sg][p][pixel]
                double2 *ab = abar + ((size_t)e * nsets + sg) * B * W * W + pixid;
                double2 *bb = bbar + ((size_t)e * nsets + sg) * B * W * W + pixid;
#pragma unroll
                for (int q1 = 0; q1 < B; q1++) {
                    double2 tb = make_double2(0.0, 0.0), at = make_double2(0.0, 0.0);
#pragma unroll
                    for (int q2 = 0; q2 < B; q2++) {
                        tb = cadd(tb, cmul(Te[q1 * B + q2], b[p][q2]));
                        at = cadd(at, cmul(a[p][q2], Te[q2 * B + q1]));
                    }
                    ab[(size_t)q1 * W * W] = cmul(sbar, cconj(tb));
                    bb[(size_t)q1 * W * W] = cmul(sbar, cconj(at));
                }
#pragma unroll
                for (int q1 = 0; q1 < B; q1++)
#pragma unroll
                    for (int q2 = 0; q2 < B; q2++)
                        tacc[q1 * B + q2] = cadd(tacc[q1 * B + q2], cmul(sbar, cconj(cmul(a[p][q1], b[p][q2]))));
            }
        }
    }
    // Tbar partial: wave reduction, then across waves
#pragma unroll
    for (int i = 0; i < NP; i++) {
        double re = wsum_reduce(tacc[i].x), im = wsum_reduce(tacc[i].y);
        if (lane == 0) red[wave][i] = make_double2(re, im);
    }
    __syncthreads();
    if (threadIdx.x < NP) {
        double2 s = make_double2(0.0, 0.0);
        for (int v = 0; v < kWaves; v++) s = cadd(s, red[v][threadIdx.x]);
        size_t blk = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        Tpart[((size_t)e * gridDim.x * gridDim.y + blk) * NP + threadIdx.x] = s;
    }
}

// Tbar[e] = sum of the per-workgroup partials; rhosum[e][pix] = sum over cells of rhobar
__global__ void k_adj_reduce(Geom g, const double2 *__restrict__ Tpart, int nblk, const double *__restrict__ rhobar,
                             double2 *__restrict__ Tbar, double *__restrict__ rhosum)
{
    __shared__ double r1[256], r2[256];
    const int e = blockIdx.y, item = blockIdx.x;
    const int NP = g.B * g.B, os = g.os, W = g.W;
    double s1 = 0.0, s2 = 0.0;
    if (item < NP) {
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
            double2 v = Tpart[((size_t)e * nblk + b) * NP + item];
            s1 += v.x;
            s2 += v.y;
        }
    } else {
        const int pix = item - NP, ii = pix / os, jj = pix % os;
        for (int c = threadIdx.x; c < g.n * g.n; c += blockDim.x) {
            int cx = c % g.n, cy = c / g.n;
            s1 += rhobar[(size_t)e * W * W + (size_t)(cx * os + ii) + (size_t)(cy * os + jj) * W];
        }
    }
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = s2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            r1[threadIdx.x] += r1[threadIdx.x + s];
            r2[threadIdx.x] += r2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (item < NP)
            Tbar[(size_t)e * NP + item] = make_double2(r1[0], r2[0]);
        else
            rhosum[(size_t)e * os * os + item - NP] = r1[0];
    }
}

// Weight overlaps of the periodic term, independent of energy:
//   Q[r*b+c][S][x] = sum_{R1} w_c(x;R1) w_r(x;S+R1)   for |S| <= 2w (both R1, S+R1 in the window),
// zero unless r and c carry Wannier functions in one spin block.
// Layout [pair][Sx+2w][Sy+2w][pixel], pixel = ii*os+jj.
__global__ void k_adj_q(Geom g, const double *__restrict__ tiles, double *__restrict__ Q)
{
    const int B = g.B, os = g.os, w = g.w, T = g.T, D = 4 * w + 1, npix = os * os;
    const long total = (long)B * B * D * D * npix;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int pix = (int)(idx % npix);
    const int sy = (int)((idx / npix) % D), sx = (int)((idx / ((long)npix * D)) % D);
    const int pr = (int)(idx / ((long)npix * D * D));
    const int r = pr / B, c = pr % B, ii = pix / os, jj = pix % os;
    const int Sx = sx - 2 * w, Sy = sy - 2 * w;
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    double v = 0.0;
    if (r < nsets * nper && c < nsets * nper && r / nper == c / nper) {
        const size_t TT = (size_t)T * T;
        for (int r1x = max(-w, -w - Sx); r1x <= min(w, w - Sx); r1x++)
            for (int r1y = max(-w, -w - Sy); r1y <= min(w, w - Sy); r1y++)
                v += tiles[c * TT + (size_t)(ii + (w - r1x) * os) * T + jj + (w - r1y) * os] *
                     tiles[r * TT + (size_t)(ii + (w - Sx - r1x) * os) * T + jj + (w - Sy - r1y) * os];
    }
    Q[idx] = v;
}

constexpr int kWinTX = 16;  // slots per workgroup along x
constexpr int kWinTY = 8;   // slots per workgroup along y
constexpr int kWinPix = 4;  // sub-pixel rows staged per phase
constexpr int kStageCols = kWinTX * kWinTY / kWinPix;  // cell columns staged at once

// Gbar on the slot window, one thread per slot, split over workgroups by
// gather and band p (a workgroup per energy, gather and p; 2 waves each, so
// that the launch has about 6b times the waves of one holding every b x b
// entry of a slot, which ran at under two waves per SIMD):
//   gather a: Gbar_a(S)_{o,p} = sum_R sum_x w_o(x;R) abar^{s(o)}_p(cell S-R, x)
//                               - i sum_x rhosum(x) Q[o,p][S][x]   (periodic term)
//   gather b: Gbar_b(S)_{p,o} = sum_R sum_x w_o(x;R) bbar^{s(o)}_p(cell -S-R, x)
// Gbar = Gbar_a + Gbar_b (k_adj_kx reads the sum).  Cells are relative to the
// map centre (cell index = position + n/2), s(o) is the spin block of orbital
// o.  A workgroup owns kWinTX x kWinTY slots; the adjoints of the cells they
// reach (the tile plus a w halo; for b its mirror image) are staged in LDS
// one strip of kWinPix x 1 sub-pixels at a time, so each adjoint value is
// read from memory about once per tile.  Every entry of both halves is
// written by exactly one workgroup (plain stores, no accumulation).
template <int B>
__global__ void __launch_bounds__(kWinTX *kWinTY) k_adj_win(Geom g, const double *__restrict__ tiles,
                                                            const double2 *__restrict__ abar,
                                                            const double2 *__restrict__ bbar,
                                                            const double *__restrict__ rhosum,
                                                            const double *__restrict__ Q,
                                                            double2 *__restrict__ Gbar_a,
                                                            double2 *__restrict__ Gbar_b)
{
    extern __shared__ double2 stage[];  // [kWinPix][RY][RX] (planes padded to RC), weights [R][kWinPix][B]
    const int os = g.os, w = g.w, T = g.T, W = g.W, n = g.n, nw = g.nw, h = g.h, n2 = n >> 1;
    // RC: stride of the staged sub-pixel planes, padded by two values so the
    // planes of one cell start on different LDS banks
    const int RX = kWinTX + 2 * w, RY = kWinTY + 2 * w, RC = RX * RY + 2;
    const int side = 2 * w + 1;
    double *wl = reinterpret_cast<double *>(stage + kWinPix * RC);
    const int tx = threadIdx.x % kWinTX, ty = threadIdx.x / kWinTX;
    const int S0x = (int)blockIdx.x * kWinTX - h, S0y = (int)blockIdx.y * kWinTY - h;
    const int Sx = S0x + tx, Sy = S0y + ty;
    const int z = blockIdx.z, p = z % B, ab = (z / B) & 1, e = z / (2 * B);
    const int nsets = g.spin ? 2 : 1;
    const int nper = g.spin ? (g.maxb >> 1) : g.maxb;
    // cell origin of the staged region: a reads cells S-R, b reads cells -S-R;
    // this thread's offset inside it for R = 0
    const int x0 = ab ? -S0x - (kWinTX - 1) - w + n2 : S0x - w + n2;
    const int y0 = ab ? -S0y - (kWinTY - 1) - w + n2 : S0y - w + n2;
    const int off = ab ? (kWinTY - 1 - ty + w) * RX + (kWinTX - 1 - tx) + w : (ty + w) * RX + tx + w;
    const size_t TT = (size_t)T * T;
    // staging lanes: sub-pixel row of the strip, first cell column; for the
    // weights (row wi, orbital wo) and the first R row wr0
    const int si = threadIdx.x % kWinPix, scx = threadIdx.x / kWinPix;
    constexpr int kWgtStep = kWinTX * kWinTY / (kWinPix * B);
    const bool wact = threadIdx.x < kWgtStep * kWinPix * B;
    const int wio = threadIdx.x % (kWinPix * B), wr0 = threadIdx.x / (kWinPix * B);
    const int wsi = wio / B, wso = wio % B;
    double2 acc[B];
#pragma unroll
    for (int i = 0; i < B; i++) acc[i] = make_double2(0.0, 0.0);

    for (int sg = 0; sg < nsets; sg++) {
        const int olo = sg * nper, ohi = olo + nper;
        const double2 *src = (ab ? bbar : abar) + (((size_t)e * nsets + sg) * B + p) * W * W;
        for (int jj = 0; jj < os; jj++)
            for (int ii0 = 0; ii0 < os; ii0 += kWinPix) {
                const int npx = min(kWinPix, os - ii0);
                __syncthreads();
                // lanes: the strip's sub-pixel row si fastest, then the cell
                // column, so a wave reads consecutive pixels of a map row
                double2 *dst = stage + (size_t)si * RC;
                for (int ry = 0; ry < RY; ry++) {
                    const int cy = y0 + ry;
                    const bool rowok = si < npx && cy >= 0 && cy < n;
                    for (int rx = scx; rx < RX; rx += kStageCols) {
                        const int cx = x0 + rx;
                        double2 v = make_double2(0.0, 0.0);
                        if (rowok && cx >= 0 && cx < n)
                            v = src[(size_t)(cx * os + ii0 + si) + (size_t)(cy * os + jj) * W];
                        dst[ry * RX + rx] = v;
                    }
                }
                // weights w_o(x;R) of this strip as wl[R][row][o], zero outside
                // the spin block and past the strip's last sub-pixel; lanes
                // (row, o) fastest, then the R column
                if (wact)
                    for (int rxw = 0; rxw < side; rxw++)
                        for (int ryw = wr0; ryw < side; ryw += kWgtStep) {
                            double v = 0.0;
                            if (wsi < npx && wso >= olo && wso < ohi)
                                v = tiles[(size_t)wso * TT + (size_t)(ii0 + wsi + (2 * w - rxw) * os) * T + jj +
                                          (2 * w - ryw) * os];
                            wl[(rxw * side + ryw) * (kWinPix * B) + wio] = v;
                        }
                __syncthreads();
                // all kWinPix rows of the strip per R, so that each wait on
                // LDS covers 2 kWinPix B FMAs; rows past the strip's last
                // sub-pixel are staged as zeros with zero weights
                for (int rxi = -w; rxi <= w; rxi++) {
                    for (int ryi = -w; ryi <= w; ryi++) {
                        const int d = -ryi * RX - rxi;
                        double2 va[kWinPix];
#pragma unroll
                        for (int i = 0; i < kWinPix; i++) va[i] = stage[i * RC + off + d];
                        // the kWinPix x B weights of this R, contiguous
                        const double *wr = wl + ((rxi + w) * side + (ryi + w)) * (kWinPix * B);
#pragma unroll
                        for (int i = 0; i < kWinPix; i++) {
#pragma unroll
                            for (int o = 0; o < B; o++) acc[o] = cadd(acc[o], cscale(wr[i * B + o], va[i]));
                        }
                    }
                }
            }
    }
    if (Sx > h || Sy > h) return;
    if (!ab && abs(Sx) <= 2 * w && abs(Sy) <= 2 * w) {
        const int D = 4 * w + 1, npix = os * os;
        const double *rs = rhosum + (size_t)e * npix;
        const double *q = Q + ((size_t)(Sx + 2 * w) * D + (Sy + 2 * w)) * npix;
#pragma unroll
        for (int o = 0; o < B; o++) {
            const double *qp = q + (size_t)(o * B + p) * D * D * npix;
            double re = 0.0;
            for (int pix = 0; pix < npix; pix++) re += rs[pix] * qp[pix];
            acc[o].y -= re;
        }
    }
    double2 *out = (ab ? Gbar_b : Gbar_a) + ((size_t)e * nw * nw + (size_t)(Sx + h) * nw + (Sy + h)) * (B * B);
#pragma unroll
    for (int o = 0; o < B; o++) out[ab ? p * B + o : o * B + p] = acc[o];
}

// Gbar(0) += T^H Tbar T^H;  Vbar[e] = M^-H Tbar (1 + G(0) T)^H.  One thread
// per energy; B is a compile-time constant and every loop unrolled, so the
// matrices stay in registers (the pivot row exchange goes by compares).
template <int B>
__global__ void __launch_bounds__(64) k_adj_tmatrix(Geom g, const double2 *__restrict__ G0w,
                                                    const double2 *__restrict__ V, const double2 *__restrict__ Tm,
                                                    const double2 *__restrict__ Tbar, int E,
                                                    double2 *__restrict__ Gbar, double2 *__restrict__ Vbar)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    constexpr int NP = B * B;
    const size_t off0 = ((size_t)e * g.nw * g.nw + (size_t)g.h * g.nw + g.h) * NP;
    const double2 *G = G0w + off0;
    const double2 *T = Tm + (size_t)e * NP;
    const double2 *Tb = Tbar + (size_t)e * NP;
    double2 M[NP], Mi[NP], X[NP];
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
            double2 acc = make_double2(r == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int k = 0; k < B; k++) acc = csub(acc, cmul(V[r * B + k], G[k * B + c]));
            M[r * B + c] = acc;
            Mi[r * B + c] = make_double2(r == c ? 1.0 : 0.0, 0.0);
        }
    bool sing = false;
#pragma unroll
    for (int j = 0; j < B; j++) {  // Gauss-Jordan, partial pivoting
        int piv = j;
        double best = cabs2(M[j * B + j]);
#pragma unroll
        for (int i = j + 1; i < B; i++)
            if (cabs2(M[i * B + j]) > best) {
                best = cabs2(M[i * B + j]);
                piv = i;
            }
        // a zero pivot (an energy the forward flagged singular) ends the elimination
        sing = sing || best == 0.0;
        if (sing) continue;
#pragma unroll
        for (int i = j + 1; i < B; i++)
            if (i == piv)
#pragma unroll
                for (int c = 0; c < B; c++) {
                    double2 t = M[j * B + c];
                    M[j * B + c] = M[i * B + c];
                    M[i * B + c] = t;
                    t = Mi[j * B + c];
                    Mi[j * B + c] = Mi[i * B + c];
                    Mi[i * B + c] = t;
                }
        const double2 ip = cinv(M[j * B + j]);
#pragma unroll
        for (int i = 0; i < B; i++) {
            if (i == j) continue;
            const double2 l = cmul(M[i * B + j], ip);
#pragma unroll
            for (int c = 0; c < B; c++) {
                M[i * B + c] = csub(M[i * B + c], cmul(l, M[j * B + c]));
                Mi[i * B + c] = csub(Mi[i * B + c], cmul(l, Mi[j * B + c]));
            }
        }
#pragma unroll
        for (int c = 0; c < B; c++) {
            M[j * B + c] = cmul(M[j * B + c], ip);
            Mi[j * B + c] = cmul(Mi[j * B + c], ip);
        }
    }
    // X = 1 + G T
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
            double2 acc = make_double2(r == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int k = 0; k < B; k++) acc = cadd(acc, cmul(G[r * B + k], T[k * B + c]));
            X[r * B + c] = acc;
        }
    // Vbar = Mi^H Tb X^H
    double2 *Vb = Vbar + (size_t)e * NP;
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < B; k++) {
                double2 tx = make_double2(0.0, 0.0);  // (Tb X^H)_{kc}
#pragma unroll
                for (int l = 0; l < B; l++) tx = cadd(tx, cmul(Tb[k * B + l], cconj(X[c * B + l])));
                acc = cadd(acc, cmul(cconj(Mi[k * B + r]), tx));
            }
            Vb[r * B + c] = acc;
        }
    // Gbar(0) += T^H Tb T^H
    double2 *G0b = Gbar + off0;
#pragma unroll
    for (int r = 0; r < B; r++)
#pragma unroll
        for (int c = 0; c < B; c++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < B; k++) {
                double2 tt = make_double2(0.0, 0.0);  // (Tb T^H)_{kc}
#pragma unroll
                for (int l = 0; l < B; l++) tt = cadd(tt, cmul(Tb[k * B + l], cconj(T[c * B + l])));
                acc = cadd(acc, cmul(cconj(T[k * B + r]), tt));
            }
            G0b[r * B + c] = cadd(G0b[r * B + c], acc);
        }
}

constexpr int kThreads = 256;

// Resolvent adjoint in the eigenbasis.  With G(k) = V F V^H (V the
// eigenvectors, F = diag f_n, f_n = 1/(w + i eta - e_n)),
//   Hbar = G^H Gbar G^H = V F* (V^H Gbar V) F* V^H,
// and V does not depend on the energy, so a k-point accumulates
//   Z += F* (V^H Gbar V) F*          (2 b^3 complex products per energy)
// and forms Hbar = V Z V^H once, after its last energy.  vv[r*B+n] = V_rn,
// gb(p) returns Gbar_rl at p = r*B + l.
template <int B, typename GbAt>
__device__ inline void adj_res_accumulate(const double2 *vv, const double2 *cf, GbAt gb, double2 *zacc)
{
#pragma unroll
    for (int n = 0; n < B; n++) {
        double2 wn[B];  // conj(f_n) (Gbar V)[:, n]
#pragma unroll
        for (int r = 0; r < B; r++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int l = 0; l < B; l++) acc = cadd(acc, cmul(gb(r * B + l), vv[l * B + n]));
            wn[r] = cmul(acc, cf[n]);
        }
#pragma unroll
        for (int m = 0; m < B; m++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int r = 0; r < B; r++) acc = cadd(acc, cmul(cconj(vv[r * B + m]), wn[r]));
            zacc[m * B + n] = cadd(zacc[m * B + n], cmul(cf[m], acc));
        }
    }
}

// hb = V Z V^H
template <int B>
__device__ inline void adj_res_finish(const double2 *vv, const double2 *zacc, double2 *hb)
{
#pragma unroll
    for (int c = 0; c < B; c++) {
        double2 t[B];  // (Z V^H)[:, c]
#pragma unroll
        for (int m = 0; m < B; m++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int n = 0; n < B; n++) acc = cadd(acc, cmulc(zacc[m * B + n], vv[c * B + n]));
            t[m] = acc;
        }
#pragma unroll
        for (int r = 0; r < B; r++) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int m = 0; m < B; m++) acc = cadd(acc, cmul(vv[r * B + m], t[m]));
            hb[r * B + c] = acc;
        }
    }
}

// conj(f_n) of the energy omega
template <int B>
__device__ inline void adj_res_cf(const double *ev, double omega, double eta, double2 *cf)
{
#pragma unroll
    for (int nn = 0; nn < B; nn++) cf[nn] = cconj(cinv(make_double2(omega - ev[nn], eta)));
}

// adjoint of the kx pass: F1bar[e][i][sy][pair] = K^-2 sum_sx Gbar[e][sx][sy][pair] e^{-2 pi i i (sx-h)/K}
__global__ void __launch_bounds__(kThreads) k_adj_kx(Geom g, const double2 *__restrict__ Gbar,
                                                   const double2 *__restrict__ Gbar_b,
                                                   const double2 *__restrict__ twn, int np,
                                                   double2 *__restrict__ F1bar)
{
    extern __shared__ double2 S[];
    const int K = g.K, sy = blockIdx.x, e = blockIdx.y, NP = g.B * g.B, p0 = blockIdx.z * np;
    const int npc = min(np, NP - p0);
    const int nw = g.nw, h = g.h;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    double2 *Tw = S + (size_t)np * K;
    for (int m = threadIdx.x; m < (K >> 1); m += blockDim.x) Tw[m] = twn[m];
    const double2 *in = Gbar + (size_t)e * nw * nw * NP, *inb = Gbar_b + (size_t)e * nw * nw * NP;
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        int m = t / npc, q = t - m * npc;
        double2 acc = make_double2(0.0, 0.0);
        for (int sx = imod(m + h, K); sx < nw; sx += K) {
            const size_t idx = ((size_t)sx * nw + sy) * NP + p0 + q;
            acc = cadd(acc, cadd(in[idx], inb[idx]));
        }
        S[(size_t)q * K + brev(m, lgK)] = acc;
    }
    __syncthreads();
    cqpi::fft_lines(S, npc, K, lgK, Tw, -1);
    const double inv = 1.0 / ((double)K * (double)K);
    double2 *out = F1bar + (size_t)e * K * nw * NP;
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        int i = t / npc, q = t - i * npc;
        out[((size_t)i * nw + sy) * NP + p0 + q] = cscale(inv, S[(size_t)q * K + i]);
    }
}

// adjoint of the ky pass fused with the resolvent adjoint, summed over energies:
// Gbar(k) for one kx row, then Hacc(k) += G(k)^H Gbar(k) G(k)^H, accumulated
// in the eigenbasis (adj_res_accumulate).  One workgroup per row, all orbital
// pairs in LDS.
template <int B, int JPT>
__global__ void __launch_bounds__(kThreads) k_adj_ky(Geom g, const double *__restrict__ eval,
                                                   const double2 *__restrict__ evec,
                                                   const double *__restrict__ omegas, int E, int EG, double eta,
                                                   const double2 *__restrict__ F1bar,
                                                   const double2 *__restrict__ twn, double2 *__restrict__ Hacc)
{
    extern __shared__ double2 S[];
    constexpr int NP = B * B;
    const int K = g.K, i = blockIdx.x, nw = g.nw, h = g.h;
    const int e0 = blockIdx.y * EG, e1 = min(E, e0 + EG);
    const long KK = (long)K * K;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    double2 *Tw = S + (size_t)NP * K;
    for (int m = threadIdx.x; m < (K >> 1); m += blockDim.x) Tw[m] = twn[m];
    // eigen data of the thread's k-points, kept for the energy loop
    double ev[JPT][B];
    double2 vv[JPT][NP], zacc[JPT][NP];
#pragma unroll
    for (int t = 0; t < JPT; t++) {
        const int j = threadIdx.x + t * kThreads;
        const long k = (long)i * K + (j < K ? j : 0);
#pragma unroll
        for (int nn = 0; nn < B; nn++) ev[t][nn] = eval[nn * KK + k];
#pragma unroll
        for (int r = 0; r < NP; r++) {
            vv[t][r] = evec[r * KK + k];
            zacc[t][r] = make_double2(0.0, 0.0);
        }
    }
    for (int e = e0; e < e1; e++) {
        __syncthreads();
        const double2 *in = F1bar + ((size_t)e * K + i) * (size_t)nw * NP;
        for (int t = threadIdx.x; t < K * NP; t += blockDim.x) {
            int m = t / NP, q = t - m * NP;
            double2 acc = make_double2(0.0, 0.0);
            for (int sy = imod(m + h, K); sy < nw; sy += K) acc = cadd(acc, in[(size_t)sy * NP + q]);
            S[(size_t)q * K + brev(m, lgK)] = acc;
        }
        __syncthreads();
        cqpi::fft_lines(S, NP, K, lgK, Tw, -1);
        const double omega = omegas[e];
#pragma unroll
        for (int t = 0; t < JPT; t++) {
            const int j = threadIdx.x + t * kThreads;
            if (j >= K) break;
            double2 cf[B];
            adj_res_cf<B>(ev[t], omega, eta, cf);
            adj_res_accumulate<B>(vv[t], cf, [&](int p) { return S[(size_t)p * K + j]; }, zacc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < JPT; t++) {
        const int j = threadIdx.x + t * kThreads;
        if (j >= K) break;
        const long k = (long)i * K + j;
        double2 hb[NP];
        adj_res_finish<B>(vv[t], zacc[t], hb);
#pragma unroll
        for (int p = 0; p < NP; p++) Hacc[((size_t)blockIdx.y * NP + p) * KK + k] = hb[p];
    }
}

// The same adjoint on the Stockham path (power-of-two K = N <= 512, b <= 3):
// one thread per k-point of the row (eigen data and the Hbar accumulators in
// registers), one lane group per orbital-pair line, padded FP64 lines.
constexpr int kAdjWaveMaxThreads = 512;

template <int B, int N>
__global__ void __launch_bounds__(kAdjWaveMaxThreads) k_adj_ky_s(Geom g, const double *__restrict__ eval,
                                                                 const double2 *__restrict__ evec,
                                                                 const double *__restrict__ omegas, int E, int EG,
                                                                 double eta, const double2 *__restrict__ F1bar,
                                                                 const double2 *__restrict__ twn,
                                                                 double2 *__restrict__ Hacc)
{
    extern __shared__ __align__(16) unsigned char smem_a[];
    double2 *S = reinterpret_cast<double2 *>(smem_a);
    constexpr int NP = B * B;
    constexpr int LN = cqpi::stockham_lanes(N);
    constexpr int NL = cqpi::padded_len<double2>(N);
    const int i = blockIdx.x, nw = g.nw, h = g.h;
    const int e0 = blockIdx.y * EG, e1 = min(E, e0 + EG);
    const long KK = (long)N * N;
    const int grp = threadIdx.x / LN, ll = threadIdx.x - grp * LN, ngrp = blockDim.x / LN;
    const int j = threadIdx.x;  // this thread's k-point (blockDim >= N)
    const bool has_k = j < N;
    const long k = (long)i * N + (has_k ? j : 0);
    // the eigenvectors and the eigenbasis accumulators in registers
    double ev[B];
    double2 vv[NP], zacc[NP];
#pragma unroll
    for (int nn = 0; nn < B; nn++) ev[nn] = eval[nn * KK + k];
#pragma unroll
    for (int r = 0; r < NP; r++) {
        vv[r] = evec[r * KK + k];
        zacc[r] = make_double2(0.0, 0.0);
    }
    for (int e = e0; e < e1; e++) {
        const double2 *in = F1bar + ((size_t)e * N + i) * (size_t)nw * NP;
        for (int t = threadIdx.x; t < N * NP; t += blockDim.x) {
            const int m = t / NP, q = t - m * NP;
            double2 acc = make_double2(0.0, 0.0);
            for (int sy = (m + h) & (N - 1); sy < nw; sy += N) acc = cadd(acc, in[(size_t)sy * NP + q]);
            S[(size_t)q * NL + cqpi::pidx<double2>(m)] = acc;
        }
        lds_barrier();
        for (int q = grp; q < NP; q += ngrp) cqpi::wave_fft<N>(S + (size_t)q * NL, twn, ll, -1);
        lds_barrier();
        if (has_k) {
            double2 cf[B];
            adj_res_cf<B>(ev, omegas[e], eta, cf);
            const double2 *gb = S + cqpi::pidx<double2>(j);
            adj_res_accumulate<B>(vv, cf, [&](int p) { return gb[(size_t)p * NL]; }, zacc);
        }
        lds_barrier();  // S is rewritten for the next energy
    }
    if (has_k) {
        double2 hb[NP];
        adj_res_finish<B>(vv, zacc, hb);
#pragma unroll
        for (int p = 0; p < NP; p++) Hacc[((size_t)blockIdx.y * NP + p) * KK + k] = hb[p];
    }
}

// The ky adjoint for rows whose b^2 lines do not fit one workgroup (b >= 4
// with K >= 256, or K = 1024), in two kernels over a block of energies:
//   k_adj_ky_lines: Gbar(k) lines, a chunk of orbital pairs per workgroup,
//                   into the workspace Gk[el][pair][kx][ky];
//   k_adj_res_k:    Hbar(k) = G(k)^H Gbar(k) G(k)^H per k-point (one thread
//                   each), summed over the block's energies of its group and
//                   added to Hacc[group][pair][k].
__global__ void __launch_bounds__(kThreads) k_adj_ky_lines(Geom g, const double2 *__restrict__ F1bar,
                                                         const double2 *__restrict__ twn, int np,
                                                         double2 *__restrict__ Gk)
{
    extern __shared__ double2 S[];
    const int K = g.K, i = blockIdx.x, el = blockIdx.y, NP = g.B * g.B, p0 = blockIdx.z * np;
    const int npc = min(np, NP - p0);
    const int nw = g.nw, h = g.h;
    const long KK = (long)K * K;
    int lgK = 0;
    while ((1 << lgK) < K) lgK++;
    double2 *Tw = S + (size_t)np * K;
    for (int m = threadIdx.x; m < (K >> 1); m += blockDim.x) Tw[m] = twn[m];
    const double2 *in = F1bar + ((size_t)el * K + i) * (size_t)nw * NP + p0;
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        const int m = t / npc, q = t - m * npc;
        double2 acc = make_double2(0.0, 0.0);
        for (int sy = (m + h) & (K - 1); sy < nw; sy += K) acc = cadd(acc, in[(size_t)sy * NP + q]);
        S[(size_t)q * K + brev(m, lgK)] = acc;
    }
    __syncthreads();
    cqpi::fft_lines(S, npc, K, lgK, Tw, -1);
    double2 *out = Gk + ((size_t)el * NP + p0) * KK + (size_t)i * K;
    for (int t = threadIdx.x; t < K * npc; t += blockDim.x) {
        const int q = t / K, j = t - q * K;
        out[(size_t)q * KK + j] = S[(size_t)q * K + j];
    }
}

template <int B>
__global__ void __launch_bounds__(256) k_adj_res_k(Geom g, const double *__restrict__ eval,
                                                   const double2 *__restrict__ evec,
                                                   const double *__restrict__ omegas, int eb, double eta,
                                                   const double2 *__restrict__ Gk, int add,
                                                   double2 *__restrict__ Hacc)
{
    constexpr int NP = B * B;
    const long KK = (long)g.K * g.K;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int grp = blockIdx.y, NG = gridDim.y;
    if (k >= KK) return;
    double ev[B];
    double2 vv[NP], zacc[NP];
#pragma unroll
    for (int nn = 0; nn < B; nn++) ev[nn] = eval[nn * KK + k];
#pragma unroll
    for (int r = 0; r < NP; r++) {
        vv[r] = evec[r * KK + k];
        zacc[r] = make_double2(0.0, 0.0);
    }
    for (int el = grp; el < eb; el += NG) {
        double2 cf[B];
        adj_res_cf<B>(ev, omegas[el], eta, cf);
        const double2 *gb = Gk + (size_t)el * NP * KK + k;
        adj_res_accumulate<B>(vv, cf, [&](int p) { return gb[(size_t)p * KK]; }, zacc);
    }
    double2 hb[NP];
    adj_res_finish<B>(vv, zacc, hb);
    double2 *out = Hacc + (size_t)grp * NP * KK + k;
#pragma unroll
    for (int p = 0; p < NP; p++) out[(size_t)p * KK] = add ? cadd(out[(size_t)p * KK], hb[p]) : hb[p];
}

__device__ inline double kcoord(int i, int K)
{
    return (i > (K >> 1)) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
}

constexpr int kMaxRy = 16;

__device__ inline double2 block_sum2(double2 v, double *r1, double *r2)
{
    r1[threadIdx.x] = v.x;
    r2[threadIdx.x] = v.y;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            r1[threadIdx.x] += r1[threadIdx.x + s];
            r2[threadIdx.x] += r2[threadIdx.x + s];
        }
        __syncthreads();
    }
    double2 out = make_double2(r1[0], r2[0]);
    __syncthreads();
    return out;
}

// e^{-2 pi i x kf} for a hopping coordinate x and the grid coordinate kf of
// index k: from the table twn[m] = e^{-2 pi i m/K} when x is an integer (then
// kf = k/K and (k-K)/K give the same phase), else by sincos
__device__ inline double2 hop_phase(double x, int k, int K, const double2 *__restrict__ twn)
{
    const double xr = rint(x);
    if (xr == x && fabs(xr) < 1e6) return twn[imod((int)(((long)xr * k) % K), K)];
    double sn, cs;
    sincos(-2.0 * M_PI * x * kcoord(k, K), &sn, &cs);
    return make_double2(cs, sn);
}

// Hopping sums S(R) = sum_k Hbar(k) e^{-2 pi i R.k} factorise over kx and ky.
// The y phases first, once per call: Pt[r][ky] = e^{-2 pi i ry_r ky}.
__global__ void k_adj_phase(Geom g, const double *__restrict__ ry, int nry, const double2 *__restrict__ twn,
                            double2 *__restrict__ Pt)
{
    const int K = g.K;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nry * K) return;
    const int r = idx / K, ky = idx - r * K;
    Pt[idx] = hop_phase(ry[r], ky, K, twn);
}

// Stage 1, one wave per (kx row, orbital pair): summing the energy groups,
//   U[kx][r][pair] = sum_ky Hbar(kx, ky) Pt[r][ky]   for the distinct ry,
//   U[kx][nry][pair] = sum_ky Hbar(kx, ky)            (for the trace)
// Hacc is [group][pair][k], so a wave reads a contiguous row.
__global__ void __launch_bounds__(64) k_adj_hrow(Geom g, const double2 *__restrict__ Hacc, int NG,
                                                 const double2 *__restrict__ Pt, int nry,
                                                 double2 *__restrict__ U)
{
    const int kx = blockIdx.x, pr = blockIdx.y;
    const int NP = g.B * g.B, K = g.K;
    const long KK = (long)K * K;
    double2 acc[kMaxRy + 1];
#pragma unroll
    for (int r = 0; r <= kMaxRy; r++) acc[r] = make_double2(0.0, 0.0);
    for (int ky = threadIdx.x; ky < K; ky += 64) {
        double2 hv = make_double2(0.0, 0.0);
        for (int gi = 0; gi < NG; gi++) hv = cadd(hv, Hacc[((size_t)gi * NP + pr) * KK + (long)kx * K + ky]);
#pragma unroll
        for (int r = 0; r < kMaxRy; r++)
            if (r < nry) acc[r] = cadd(acc[r], cmul(hv, Pt[r * K + ky]));
        acc[kMaxRy] = cadd(acc[kMaxRy], hv);
    }
#pragma unroll
    for (int r = 0; r <= kMaxRy; r++) {
        if (r < nry || r == kMaxRy) {
            const double re = wsum_reduce(acc[r].x), im = wsum_reduce(acc[r].y);
            const int slot = r == kMaxRy ? nry : r;
            if (threadIdx.x == 0) U[((size_t)kx * (nry + 1) + slot) * NP + pr] = make_double2(re, im);
        }
    }
}

// Stage 2, one workgroup per (R, pair): SR[R][pair] = sum_kx U[kx][iry(R)][pair] e^{-2 pi i rx kx};
// Rl[R] = (rx, iry); the last entry (0, nry) gives the plain sums used for the trace (mu).
__global__ void __launch_bounds__(256) k_adj_hcol(Geom g, const double2 *__restrict__ U, int nry,
                                                  const double2 *__restrict__ Rl, const double2 *__restrict__ twn,
                                                  double2 *__restrict__ SR)
{
    __shared__ double r1[256], r2[256];
    const int ridx = blockIdx.x, pr = blockIdx.y;
    const int NP = g.B * g.B, K = g.K;
    const double rx = Rl[ridx].x;
    const int iry = (int)Rl[ridx].y;
    double2 acc = make_double2(0.0, 0.0);
    for (int kx = threadIdx.x; kx < K; kx += blockDim.x)
        acc = cadd(acc, cmul(U[((size_t)kx * (nry + 1) + iry) * NP + pr], hop_phase(rx, kx, K, twn)));
    double2 s = block_sum2(acc, r1, r2);
    if (threadIdx.x == 0) SR[(size_t)ridx * NP + pr] = s;
}

// energy groups of the ky adjoint (row workgroups x groups): enough
// workgroups to fill the chip.  The Stockham kernel (b <= 3, 64 <= K <= 512)
// holds a CU with one workgroup, so two workgroups per CU (512) suffice; fewer
// groups also shrink the Hacc slices written there and read by k_adj_hrow.
int cqpi_energy_groups(const Geom &g, int E)
{
    const bool stockham = g.B <= 3 && g.K >= 64 && g.K <= 512;
    const int ng = (stockham ? 512 : 2048) / std::max(g.K, 1);
    return std::max(1, std::min(E, ng));
}

template <int B>
int launch_pix_t(const Geom &g, const double2 *G0w, const double2 *T, const double *tiles, const double *rhobar,
                 int E, double2 *abar, double2 *bbar, double2 *Tpart, int &nblk, hipStream_t s)
{
    constexpr int P = B <= 1 ? 8 : (B <= 3 ? 4 : (B <= 6 ? 2 : 1));
    int nb = (g.n + kCells - 1) / kCells;
    nblk = nb * nb;
    const size_t lds = (size_t)2 * (kCells + 2 * g.w) * (kCells + 2 * g.w) * B * sizeof(double2);
    if (lds > 64 * 1024) return CQPI_E_ARG;  // beyond the windows adjoint_supported accepts
    hipLaunchKernelGGL((k_adj_pix<B, P>), dim3(nb, nb, E), dim3(64 * kWaves), lds, s, g, G0w, T, tiles, rhobar,
                       abar, bbar, Tpart);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

size_t win_lds(const Geom &g)
{
    return sizeof(double2) * kWinPix * ((size_t)(kWinTX + 2 * g.w) * (kWinTY + 2 * g.w) + 2) +
           sizeof(double) * kWinPix * (size_t)(2 * g.w + 1) * (2 * g.w + 1) * g.B;
}

template <int B>
int launch_win_t(const Geom &g, const double *tiles, const double2 *abar, const double2 *bbar, const double *rhosum,
                 double *Q, int E, double2 *Gbar, hipStream_t s)
{
    const long nq = (long)cqpi::adjoint_q_size(g);
    hipLaunchKernelGGL(k_adj_q, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, g, tiles, Q);
    // z: (energy, gather, band p); Gbar holds the a half, then the b half
    dim3 grid((g.nw + kWinTX - 1) / kWinTX, (g.nw + kWinTY - 1) / kWinTY, E * 2 * B);
    hipLaunchKernelGGL(k_adj_win<B>, grid, dim3(kWinTX * kWinTY), win_lds(g), s, g, tiles, abar, bbar, rhosum, Q,
                       Gbar, Gbar + (size_t)E * g.nw * g.nw * B * B);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <int B, int N>
int launch_ky_s(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E, double eta,
                const double2 *F1bar, const double2 *twn, double2 *Hacc, hipStream_t s)
{
    constexpr int NP = B * B;
    const int NG = cqpi_energy_groups(g, E), EG = (E + NG - 1) / NG;
    int thr = std::max(NP * cqpi::stockham_lanes(N), N);
    thr = std::min((thr + 63) / 64 * 64, kAdjWaveMaxThreads);  // >= N: one thread per k-point
    const size_t lds = sizeof(double2) * (size_t)NP * cqpi::padded_len<double2>(N);
    hipLaunchKernelGGL((k_adj_ky_s<B, N>), dim3(N, NG), dim3(thr), lds, s, g, eval, evec, omegas, E, EG, eta, F1bar,
                       twn, Hacc);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <int B>
int launch_ky_t(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E, double eta,
                const double2 *F1bar, const double2 *twn, double2 *Hacc, hipStream_t s)
{
    if constexpr (B <= 3) {
        switch (g.K) {
        case 64: return launch_ky_s<B, 64>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s);
        case 128: return launch_ky_s<B, 128>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s);
        case 256: return launch_ky_s<B, 256>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s);
        case 512: return launch_ky_s<B, 512>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s);
        default: break;
        }
    }
    size_t lds = sizeof(double2) * ((size_t)B * B * g.K + (size_t)(g.K >> 1));
    const int NG = cqpi_energy_groups(g, E), EG = (E + NG - 1) / NG;
    dim3 grid(g.K, NG);
    const int jpt = (g.K + kThreads - 1) / kThreads;
    if (jpt <= 1)
        hipLaunchKernelGGL((k_adj_ky<B, 1>), grid, dim3(kThreads), lds, s, g, eval, evec, omegas, E, EG, eta, F1bar,
                           twn, Hacc);
    else if (jpt == 2)
        hipLaunchKernelGGL((k_adj_ky<B, 2>), grid, dim3(kThreads), lds, s, g, eval, evec, omegas, E, EG, eta, F1bar,
                           twn, Hacc);
    else
        hipLaunchKernelGGL((k_adj_ky<B, 4>), grid, dim3(kThreads), lds, s, g, eval, evec, omegas, E, EG, eta, F1bar,
                           twn, Hacc);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

// the ky adjoint in one kernel per row for b <= 3: the Stockham kernel
// (64 <= K <= 512) or the row's b^2 lines and twiddles within 96 KB of LDS;
// otherwise (and for every b >= 4, whose per-thread accumulators would spill
// beside the row's lines) the two-kernel path through the Gbar(k) workspace
bool ky_fused(const Geom &g)
{
    const size_t lds = sizeof(double2) * ((size_t)g.B * g.B * g.K + (size_t)(g.K >> 1));
    return g.B <= 3 && ((g.K >= 64 && g.K <= 512) || (lds <= 96 * 1024 && g.K <= 4 * kThreads));
}

constexpr int kAdjMaxB = 5;
constexpr size_t kAdjGkBudget = (size_t)1 << 30;  // Gbar(k) workspace of the two-kernel ky path

template <int B>
int launch_ky_chunked(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                      double eta, const double2 *F1bar, const double2 *twn, double2 *Hacc, double2 *Gk,
                      size_t gk_bytes, hipStream_t s)
{
    constexpr int NP = B * B;
    const long KK = (long)g.K * g.K;
    const size_t per_e = (size_t)NP * KK * sizeof(double2);
    if (!Gk || gk_bytes < per_e) return CQPI_E_ARG;
    const int EBw = (int)std::min<size_t>((size_t)E, gk_bytes / per_e);
    int npmax = (int)((32 * 1024) / (sizeof(double2) * (size_t)g.K));
    if (npmax < 1) npmax = 1;
    const int nch = (NP + npmax - 1) / npmax, np = (NP + nch - 1) / nch;
    const size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    const int NG = cqpi_energy_groups(g, E);
    const unsigned nkb = (unsigned)((KK + 255) / 256);
    for (int ea = 0; ea < E; ea += EBw) {
        const int eb = std::min(EBw, E - ea);
        hipLaunchKernelGGL(k_adj_ky_lines, dim3(g.K, eb, nch), dim3(kThreads), lds, s, g,
                           F1bar + (size_t)ea * g.K * g.nw * NP, twn, np, Gk);
        hipLaunchKernelGGL(k_adj_res_k<B>, dim3(nkb, NG), dim3(256), 0, s, g, eval, evec, omegas + ea, eb, eta,
                           (const double2 *)Gk, ea > 0 ? 1 : 0, Hacc);
        if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    }
    return 0;
}

}  // namespace

namespace cqpi {

size_t adjoint_q_size(const Geom &g)
{
    return (size_t)g.B * g.B * (4 * g.w + 1) * (4 * g.w + 1) * g.os * g.os;
}

size_t adjoint_pix_blocks(const Geom &g)
{
    size_t nb = (g.n + kCells - 1) / kCells;
    return nb * nb;
}

bool adjoint_supported(const Geom &g)
{
    const bool pow2 = (g.K & (g.K - 1)) == 0;
    const size_t pix = (size_t)2 * (kCells + 2 * g.w) * (kCells + 2 * g.w) * g.B * sizeof(double2);
    return pow2 && g.B <= kAdjMaxB && g.w <= 4 && g.K <= 4 * kThreads && pix <= 64 * 1024 &&
           win_lds(g) <= 64 * 1024;
}

size_t adjoint_workspace_bytes(const Geom &g, int E)
{
    if (ky_fused(g)) return 0;
    const size_t per_e = (size_t)g.B * g.B * g.K * g.K * sizeof(double2);
    const size_t eb = std::max<size_t>(1, std::min<size_t>((size_t)E, kAdjGkBudget / per_e));
    return eb * per_e;
}

int adjoint_energy_groups(const Geom &g, int E) { return cqpi_energy_groups(g, E); }

int launch_adjoint(const Geom &g, const double2 *G0w, const double2 *T, const double2 *V, const double *tiles,
                   const double *rhobar, const double *eval, const double2 *evec, const double *omegas, int E,
                   double eta, const double2 *twn, const double *ryvals, int nry, const double2 *Rl, int nRl,
                   double2 *abar, double2 *bbar, double2 *Tpart, double2 *Tbar, double *rhosum, double2 *Gbar,
                   double2 *Vbar, double2 *F1bar, double2 *Hacc, double2 *U, double2 *SR, double *Q,
                   double2 *Gk, size_t gk_bytes, hipStream_t s)
{
    if (!adjoint_supported(g) || nry > kMaxRy) return CQPI_E_ARG;
    int nblk = 0, r = CQPI_E_ARG;
    switch (g.B) {
    case 1: r = launch_pix_t<1>(g, G0w, T, tiles, rhobar, E, abar, bbar, Tpart, nblk, s); break;
    case 2: r = launch_pix_t<2>(g, G0w, T, tiles, rhobar, E, abar, bbar, Tpart, nblk, s); break;
    case 3: r = launch_pix_t<3>(g, G0w, T, tiles, rhobar, E, abar, bbar, Tpart, nblk, s); break;
    case 4: r = launch_pix_t<4>(g, G0w, T, tiles, rhobar, E, abar, bbar, Tpart, nblk, s); break;
    case 5: r = launch_pix_t<5>(g, G0w, T, tiles, rhobar, E, abar, bbar, Tpart, nblk, s); break;
    }
    if (r) return r;
    const int NP = g.B * g.B;
    hipLaunchKernelGGL(k_adj_reduce, dim3(NP + g.os * g.os, E), dim3(256), 0, s, g, Tpart, nblk, rhobar, Tbar, rhosum);
    switch (g.B) {
    case 1: r = launch_win_t<1>(g, tiles, abar, bbar, rhosum, Q, E, Gbar, s); break;
    case 2: r = launch_win_t<2>(g, tiles, abar, bbar, rhosum, Q, E, Gbar, s); break;
    case 3: r = launch_win_t<3>(g, tiles, abar, bbar, rhosum, Q, E, Gbar, s); break;
    case 4: r = launch_win_t<4>(g, tiles, abar, bbar, rhosum, Q, E, Gbar, s); break;
    case 5: r = launch_win_t<5>(g, tiles, abar, bbar, rhosum, Q, E, Gbar, s); break;
    }
    if (r) return r;
    switch (g.B) {
#define CASE(b)                                                                                                   \
    case b:                                                                                                       \
        hipLaunchKernelGGL(k_adj_tmatrix<b>, dim3((E + 63) / 64), dim3(64), 0, s, g, G0w, V, T, Tbar, E, Gbar, Vbar); \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
    default: return CQPI_E_ARG;
    }
    // adjoint transform along kx (pairs chunked to the LDS budget), then ky + resolvent
    int npmax = (int)((32 * 1024) / (sizeof(double2) * (size_t)g.K));
    if (npmax < 1) npmax = 1;
    int nch = (NP + npmax - 1) / npmax, np = (NP + nch - 1) / nch;
    size_t lds = sizeof(double2) * ((size_t)np * g.K + (size_t)(g.K >> 1));
    hipLaunchKernelGGL(k_adj_kx, dim3(g.nw, E, nch), dim3(kThreads), lds, s, g, Gbar,
                       Gbar + (size_t)E * g.nw * g.nw * NP, twn, np, F1bar);
    if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    if (ky_fused(g)) {
        switch (g.B) {
        case 1: r = launch_ky_t<1>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s); break;
        case 2: r = launch_ky_t<2>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s); break;
        case 3: r = launch_ky_t<3>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, s); break;
        default: r = CQPI_E_ARG;
        }
    } else {
#define CASE(b) \
    case b: r = launch_ky_chunked<b>(g, eval, evec, omegas, E, eta, F1bar, twn, Hacc, Gk, gk_bytes, s); break;
        switch (g.B) {
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
        default: r = CQPI_E_ARG;
        }
#undef CASE
    }
    if (r) return r;
    // the y phase table follows the K (nry + 1) b^2 sums in U
    double2 *Pt = U + (size_t)g.K * (nry + 1) * NP;
    if (nry > 0)
        hipLaunchKernelGGL(k_adj_phase, dim3((nry * g.K + 255) / 256), dim3(256), 0, s, g, ryvals, nry, twn, Pt);
    hipLaunchKernelGGL(k_adj_hrow, dim3(g.K, NP), dim3(64), 0, s, g, Hacc, cqpi_energy_groups(g, E), Pt, nry, U);
    hipLaunchKernelGGL(k_adj_hcol, dim3(nRl, NP), dim3(256), 0, s, g, U, nry, Rl, twn, SR);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace cqpi

// Mixed-precision continuum contraction on MFMA (gfx950).
//
// Same factorisation as the FP64 kernel, rho = -Im(a^T T b) - bulk, but the
// gathers a and b run as fp16 matrix products with FP32 accumulation.
// Operands are split into hi + lo fp16 parts (x = hi + lo/2^12, each part
// scaled by a power of two so that max|x| lies in [2^12, 2^13)): the products
// hi*hi + (hi*lo' + lo'*hi)/2^12 recover ~22 significant bits per product.
// The T sandwich and the bulk table stay FP64.
//
// GEMM shape (per wave): one "tile" is CB consecutive cells along x at one
// cell row cy and one sub-pixel row jj.  Its M = CB*os rows are the pixels
// (c, ii); K runs over (dy, o, dx): the lattice rows cy+dy-w, the orbitals o
// of the current spin block and the CB+2w lattice columns the tile touches,
// in groups of 8 consecutive columns; N runs over (p, re/im) of
// G0(pos+R)[o][p] (the a gather) and of G0(-pos-R)[p][o] (the b gather).
// The tile matrix A[(c,ii)][(dy,o,dx)] = w_o(ii+(2w-dx+c)os, jj+(2w-dy)os)
// (zero outside the band 0 <= dx-c <= 2w) is the same for every tile and is
// precomputed on the host in MFMA fragment order.  G0 is stored per energy as
// "planes" (one per (o, p, re/im)) with rows along x, so a lane's 8 K values
// of a B fragment are 8 consecutive halves in memory.
//
// v_mfma_f32_16x16x32_f16 lane layout: A: lane l holds row l%16, k 8*(l/16)..+7;
// B: lane l holds column l%16, k 8*(l/16)..+7; C: lane l holds rows
// 4*(l/16)..+3 of column l%16.
#include <cmath>
#include <vector>

#include "cqpi_common.h"
#include "mixed.h"

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kLoScale = 4096;  // lo parts are stored times 2^12

// per-energy power-of-two scale: max|G0 window| * s in [2^12, 2^13).  The
// maximum is taken by several workgroups per energy (grid x), each folding its
// slice and raising scale[e], read as a 64-bit integer, with one atomic max:
// non-negative doubles order like their bit patterns.  k_wscale_clear zeroes
// scale[] first and k_wscale_fin turns each maximum into its scale in place.
// (One 1024-lane workgroup per energy took 78.6 us for the 133-row window of
// 3 bands and one energy; one workgroup per 2048 values, 6.2 us plus 4.7 us
// for the clear and the finish.)
__global__ void k_wscale_clear(double *__restrict__ scale, int E)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) scale[e] = 0.0;
}

__global__ void __launch_bounds__(256) k_wscale(const double2 *__restrict__ G0w, size_t per_e,
                                               double *__restrict__ scale)
{
    __shared__ double red[256];
    const double2 *G = G0w + (size_t)blockIdx.y * per_e;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_e; i += stride) {
        const double2 v = G[i];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0.0)
        atomicMax(reinterpret_cast<unsigned long long *>(scale) + blockIdx.y,
                  (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_wscale_fin(double *__restrict__ scale, int E)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const double mx = scale[e];
    double sc = 1.0;
    if (mx > 0.0) {
        int ex;
        frexp(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
        sc = ldexp(1.0, 13 - ex);
    }
    scale[e] = sc;
}

// planes P[e][ab][hl][(o*B+p)*2+ri] of nw rows sy and Ly columns idx (laid
// out by plane_off), rows along x: ab=0 holds G0(R)[o][p], ab=1 holds
// G0(-R)[p][o], R = (idx-OFF-h, sy-h); entries outside the window are zero.
// (The whole-chain path writes these planes from the transform kernel; this
// kernel serves the GPUQPI drop-in path.)  One workgroup per (row sy, energy,
// chunk of kPackPairs source pairs): the a row sy and the b row 2h - sy of the
// chunk's pairs are read as contiguous runs of their window records into LDS,
// then each item builds 8 consecutive columns of one plane and stores them as
// one 16-byte hi and one 16-byte lo vector.  A gather of each value from its
// own record, one plane row per workgroup, took 2.1 ms more for the 6-band
// Josephson map of 51 energies.
// kp pairs per chunk: 8, fewer where the two staged rows would pass 64 KB of LDS
constexpr int kPackPairs = 8;
constexpr size_t kPackLds = 64 * 1024;

__global__ void __launch_bounds__(256) k_pack_rows(Geom g, MixGeom mg, const double2 *__restrict__ G0w,
                                                   const double *__restrict__ scale, _Float16 *__restrict__ P,
                                                   int kp)
{
    extern __shared__ double2 Sp[];  // [ab][sx][kp]
    const int sy = blockIdx.x, e = blockIdx.z;
    const int B = g.B, NP = B * B, nw = g.nw, h = g.h;
    const int q0 = blockIdx.y * kp, nq = min(kp, NP - q0);
    const double s = scale[e];
    const double2 *G = G0w + (size_t)e * nw * nw * NP;
    const size_t plane = (size_t)mg.Ly * nw;
    // ab = 0: G(sx, sy)[q]; ab = 1: G(2h - sx, 2h - sy)[q], q = q0 + j
    for (int t = threadIdx.x; t < 2 * nw * nq; t += blockDim.x) {
        const int ab = t / (nw * nq), rest = t - ab * nw * nq, sx = rest / nq, j = rest - sx * nq;
        const size_t rec = ab ? ((size_t)(2 * h - sx) * nw + (2 * h - sy)) : ((size_t)sx * nw + sy);
        Sp[((size_t)ab * nw + sx) * kp + j] = G[rec * NP + q0 + j];
    }
    __syncthreads();
    const int ng = mg.Ly >> 3;
    // items (ab, j, re|im, column group); source pair q = q0 + j = r B + c
    // goes to plane (r B + c) for a (o = r, p = c) and (c B + r) for b
    // (the b plane (o, p) holds G(-R)[p][o])
    for (int t = threadIdx.x; t < 2 * nq * 2 * ng; t += blockDim.x) {
        const int gi = t % ng, rest = t / ng, ri = rest & 1, abj = rest >> 1, ab = abj / nq, j = abj - ab * nq;
        const int q = q0 + j, r = q / B, c = q - r * B;
        const int pl = (ab ? (c * B + r) : (r * B + c)) * 2 + ri;
        half8 hv, lv;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int sx = gi * 8 + u - mg.OFF;
            double v = 0.0;
            if (sx >= 0 && sx < nw) {
                const double2 z = Sp[((size_t)ab * nw + sx) * kp + j];
                v = (ri ? z.y : z.x) * s;
            }
            const _Float16 hh = (_Float16)v;
            hv[u] = hh;
            lv[u] = (_Float16)((v - (double)hh) * kLoScale);
        }
        const size_t o = plane_off(sy, gi * 8, nw);
        *reinterpret_cast<half8 *>(P + ((((size_t)e * 2 + ab) * 2 + 0) * mg.NPL + pl) * plane + o) = hv;
        *reinterpret_cast<half8 *>(P + ((((size_t)e * 2 + ab) * 2 + 1) * mg.NPL + pl) * plane + o) = lv;
    }
}

// Folds T into the b planes for the LDOS map: the planes (o, p) of ab = 1
// become c = T b, (T G0(-R))[p][o] = sum_q T_pq G0(-R)[q][o], so that the
// contraction's epilogue needs no per-pixel T sandwich.  One lane per
// (row, column, orbital o) of one energy; the b values of its column are
// rebuilt from hi + lo 2^-12 in FP32, the sums taken in FP32 and split
// again.  cscale[e] receives the new scale of the c values: scale[e]
// 2^-ex with max_p sum_q |T_pq| < 2^ex, so that they stay below 2^13.
__global__ void k_fold_c(Geom g, MixGeom mg, const double *__restrict__ scale, const double2 *__restrict__ Tm,
                         _Float16 *__restrict__ P, double *__restrict__ cscale)
{
    constexpr int kMaxFoldB = 16;
    const int e = blockIdx.y, B = g.B, nw = g.nw, mb = g.maxb;
    const double2 *Te = Tm + (size_t)e * B * B;
    double ts = 0.0;
    for (int p = 0; p < B; p++) {
        double r = 0.0;
        for (int q = 0; q < B; q++) r += fabs(Te[p * B + q].x) + fabs(Te[p * B + q].y);
        ts = fmax(ts, r);
    }
    int ex = 0;
    if (ts > 0.0) frexp(ts, &ex);
    const float f = (float)ldexp(1.0, -ex);  // c scale / b scale
    if (blockIdx.x == 0 && threadIdx.x == 0) cscale[e] = scale[e] * (double)f;
    // items in the planes' own order ([column group][row][8 columns]), so
    // that the lanes of a wave touch contiguous memory
    const size_t item = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t cells = (size_t)nw * mg.Ly;
    if (item >= cells * mb) return;
    const int o = (int)(item / cells);
    const size_t off = item - (size_t)o * cells;
    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    _Float16 *Pw = P + (size_t)e * 4 * mg.NPL * plane;
    const _Float16 *bh = Pe + (size_t)(2 * mg.NPL) * plane, *bl = Pe + (size_t)(3 * mg.NPL) * plane;
    float br[kMaxFoldB], bi[kMaxFoldB];
    for (int q = 0; q < B; q++) {
        const size_t pr = (size_t)((o * B + q) * 2) * plane + off, pi = pr + plane;
        br[q] = (float)bh[pr] + (float)bl[pr] * (1.0f / kLoScale);
        bi[q] = (float)bh[pi] + (float)bl[pi] * (1.0f / kLoScale);
    }
    for (int p = 0; p < B; p++) {
        float cr = 0.f, ci = 0.f;
        for (int q = 0; q < B; q++) {
            const float tr = (float)Te[p * B + q].x, ti = (float)Te[p * B + q].y;
            cr += tr * br[q] - ti * bi[q];
            ci += tr * bi[q] + ti * br[q];
        }
        cr *= f;
        ci *= f;
        const size_t pr = (size_t)((o * B + p) * 2) * plane + off, pi = pr + plane;
        const _Float16 hr = (_Float16)cr, hi = (_Float16)ci;
        Pw[(size_t)(2 * mg.NPL) * plane + pr] = hr;
        Pw[(size_t)(3 * mg.NPL) * plane + pr] = (_Float16)((cr - (float)hr) * (float)kLoScale);
        Pw[(size_t)(2 * mg.NPL) * plane + pi] = hi;
        Pw[(size_t)(3 * mg.NPL) * plane + pi] = (_Float16)((ci - (float)hi) * (float)kLoScale);
    }
}

// the 8 halves at columns col .. col+7 of row `row` of plane pl (col a
// multiple of LDW; the run may span two column groups of plane_off)
template <int LDW>
__device__ inline half8 load8(const _Float16 *pl, int row, int col, int nw)
{
    half8 r;
    if constexpr (LDW == 8) {
        uint4 u = *(const uint4 *)(pl + plane_off(row, col, nw));
        __builtin_memcpy(&r, &u, 16);
    } else if constexpr (LDW == 4) {
        uint2 u0 = *(const uint2 *)(pl + plane_off(row, col, nw)), u1 = *(const uint2 *)(pl + plane_off(row, col + 4, nw));
        uint4 u = make_uint4(u0.x, u0.y, u1.x, u1.y);
        __builtin_memcpy(&r, &u, 16);
    } else if constexpr (LDW == 2) {
        uint4 u = make_uint4(*(const unsigned *)(pl + plane_off(row, col, nw)),
                             *(const unsigned *)(pl + plane_off(row, col + 2, nw)),
                             *(const unsigned *)(pl + plane_off(row, col + 4, nw)),
                             *(const unsigned *)(pl + plane_off(row, col + 6, nw)));
        __builtin_memcpy(&r, &u, 16);
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = pl[plane_off(row, col + j, nw)];
    }
    return r;
}

constexpr int kWaves = 4;
// waves per SIMD the register-resident contraction is compiled for: three for
// up to 3 orbitals (the LDS of three workgroups fits a CU; the compiler then
// keeps the accumulators in VGPRs instead of AGPRs), one for more
constexpr int areg_waves(int bt) { return bt <= 3 ? 3 : 1; }
constexpr int kJB = 4;  // sub-pixel rows (jj) per accumulator block

// one wave = one tile: CB cells along x starting at cx0, one cell row cy,
// all sub-pixels.  Rows of the GEMM are the pixels (c, ii), c < CB, ii < os,
// of one sub-pixel row jj; the jj loop runs in blocks of kJB accumulators.
//
// J: the Josephson map of a superconducting model (src/qpihipkernel.cc:171-271,
// CPU src/qpi.cc:953-1033): the b gather reads the particle-hole column
// o + maxband (the planes of G0(-R)[p][o + maxband]), both the real and the
// imaginary part of sum_s a^T T b' are kept, and per pixel
//   rho = -Im[ tip (sum_s a^T T b' + bulk') ]
// with the complex periodic table bulk' (E x os^2 double2) and tips[e].
//
// cscale != NULL (LDOS map): the b planes hold c = T b (k_fold_c, scale
// cscale[e]); the epilogue then sums Im(a_p c_p) with no T sandwich.
template <int NT, int LDW, int JB, bool J>
__global__ void __launch_bounds__(64 * kWaves) k_cmix(Geom g, MixGeom mg, const _Float16 *__restrict__ P,
                                                    const half8 *__restrict__ Afrag,
                                                    const double *__restrict__ scale, double sA,
                                                    const double2 *__restrict__ Tm,
                                                    const double *__restrict__ bulk,
                                                    double *__restrict__ ldos,
                                                    const double2 *__restrict__ tips,
                                                    const double *__restrict__ cscale)
{
    __shared__ double2 Ts[16 * 16];
    constexpr int NB = 16 / JB;  // jj blocks (os <= 16)
    __shared__ float Cs[kWaves][JB][16][16 * NT + 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int bx, by, e;
    xcd_block(bx, by, e);
    const int B = g.B, os = g.os, w = g.w, n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cy_raw = by * kWaves + wave;
    const bool cy_ok = cy_raw < n;
    const int cy = cy_ok ? cy_raw : n - 1;
    const int cx0 = bx * mg.CB;
    const double2 *Te = Tm + (size_t)e * B * B;
    for (int i = threadIdx.x; i < B * B; i += blockDim.x) Ts[i] = Te[i];
    __syncthreads();

    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int col = lane & 15, kq = lane >> 4;
    const int col0 = cx0 - n2 - w + h + mg.OFF;
    const int rows = mg.CB * os;
    double val[NB], valr[NB];  // Im and (J) Re of sum_s a^T T b
#pragma unroll
    for (int jb = 0; jb < NB; jb++) val[jb] = valr[jb] = 0.0;
    const int jo = J ? g.maxb : 0;  // orbital offset of the b planes

    for (int sg = 0; sg < mg.nsets; sg++) {
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            const int jj0 = jb * JB;
            if (jj0 >= os) break;
            float4v hh[JB][NT], cr[JB][NT];
#pragma unroll
            for (int i = 0; i < JB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    hh[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                    cr[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                }
            for (int ks = 0; ks < mg.KS; ks++) {
                const int kg = ks * 4 + kq;
                half8 bh[NT], bl[NT];
                const bool kok = kg < mg.KG;
                int dy = 0, ol = 0, g8 = 0;
                if (kok) {
                    g8 = kg % mg.NUg;
                    int r = kg / mg.NUg;
                    ol = r % mg.nper;
                    dy = r / mg.nper;
                }
                const int o = sg * mg.nper + ol;
                const int row = cy - n2 + dy - w + h;
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int nn = t * 16 + col;
                    int ab = -1, pl = 0;
                    if (nn < 2 * B) {
                        ab = 0;
                        pl = (o * B + (nn >> 1)) * 2 + (nn & 1);
                    } else if (nn < 4 * B) {
                        ab = 1;
                        int m2 = nn - 2 * B;
                        pl = ((o + jo) * B + (m2 >> 1)) * 2 + (m2 & 1);
                    }
                    if (kok && ab >= 0) {
                        bh[t] = load8<LDW>(Pe + ((size_t)(ab * 2 + 0) * mg.NPL + pl) * plane, row, col0 + g8 * 8, nw);
                        bl[t] = load8<LDW>(Pe + ((size_t)(ab * 2 + 1) * mg.NPL + pl) * plane, row, col0 + g8 * 8, nw);
                    } else {
                        bh[t] = half8{0, 0, 0, 0, 0, 0, 0, 0};
                        bl[t] = bh[t];
                    }
                }
#pragma unroll
                for (int i = 0; i < JB; i++) {
                    const int jj = jj0 + i;
                    if (jj >= os) break;
                    const size_t fa = ((((size_t)sg * os + jj) * mg.KS + ks) * 2) * 64;
                    const half8 ah = Afrag[fa + lane];
                    const half8 al = Afrag[fa + 64 + lane];
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        hh[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[t], hh[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[t], cr[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[t], cr[i][t], 0, 0, 0);
                    }
                }
            }
            // epilogue (wave-local): C -> LDS, then one pixel per lane:
            // Im sum_pq a_p T_pq b_q in FP64
#pragma unroll
            for (int i = 0; i < JB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        Cs[wave][i][kq * 4 + r][t * 16 + col] = hh[i][t][r] + cr[i][t][r] * (1.0f / kLoScale);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                const int i = lane >> 4, m = lane & 15;
                if (i < JB && jj0 + i < os && m < rows) {
                    const float *Cr = Cs[wave][i][m];
                    double acc = 0.0, accr = 0.0;
                    if (!J && cscale) {
                        for (int p = 0; p < B; p++)
                            acc += (double)Cr[2 * p] * (double)Cr[2 * B + 2 * p + 1] +
                                   (double)Cr[2 * p + 1] * (double)Cr[2 * B + 2 * p];
                    } else
                    for (int p = 0; p < B; p++) {
                        double2 tb = make_double2(0.0, 0.0);
                        for (int q = 0; q < B; q++)
                            tb = cadd(tb, cmul(Ts[p * B + q], make_double2(Cr[2 * B + 2 * q], Cr[2 * B + 2 * q + 1])));
                        acc += (double)Cr[2 * p] * tb.y + (double)Cr[2 * p + 1] * tb.x;
                        if (J) accr += (double)Cr[2 * p] * tb.x - (double)Cr[2 * p + 1] * tb.y;
                    }
                    val[jb] += acc;
                    if (J) valr[jb] += accr;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    const double unscale = 1.0 / (sA * sA * scale[e] * (!J && cscale ? cscale[e] : scale[e]));
    const int W = g.W;
    double *out = ldos + (size_t)e * W * W;
#pragma unroll
    for (int jb = 0; jb < NB; jb++) {
        const int i = lane >> 4, jj = jb * JB + i, m = lane & 15;
        if (i >= JB || jj >= os || m >= rows || !cy_ok) continue;
        const int c = m / os, ii = m - c * os, cx = cx0 + c;
        if (cx >= n) continue;
        const int X = cx * os + ii, Y = cy * os + jj;
        if (J) {
            const double2 bz = reinterpret_cast<const double2 *>(bulk)[(size_t)e * os * os + ii * os + jj];
            const double2 tp = tips[e];
            const double zr = valr[jb] * unscale + bz.x, zi = val[jb] * unscale + bz.y;
            out[(size_t)X + (size_t)Y * W] = -(tp.x * zi + tp.y * zr);
        } else {
            out[(size_t)X + (size_t)Y * W] = -(val[jb] * unscale + bulk[(size_t)e * os * os + ii * os + jj]);
        }
    }
}

// Specialised kernel for small sub-pixel counts and K depths (compile-time OS,
// KS): the A fragments of every (jj, ks) stay in registers and each wave walks
// TPW tiles along y, so A is read from memory once per wave instead of once
// per tile.  Non-spin models only (one orbital block).
template <int NT, int LDW, int OS, int KS, int TPW, int BT, int JS>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(areg_waves(BT), areg_waves(BT)))) k_cmix_areg(Geom g, MixGeom mg, const _Float16 *__restrict__ P,
                                                         const half8 *__restrict__ Afrag,
                                                         const double *__restrict__ scale, double sA,
                                                         const double2 *__restrict__ Tm,
                                                         const double *__restrict__ bulk,
                                                         double *__restrict__ ldos)
{
    // JS waves share a cell row, wave w taking the sub-pixel rows
    // jj in [j0, j0 + OSW): fewer tile fragments per wave, more waves per SIMD
    constexpr int OSW = OS / JS, CPS = kWaves / JS;  // cell rows per step
    __shared__ float2 Tf[BT * BT];  // T in FP32 for the epilogue
    __shared__ float Cs[kWaves][OSW][16][16 * NT + 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j0 = (wave % JS) * OSW, wrow = wave / JS;
    int bx, by, e;
    xcd_block(bx, by, e);
    const int B = g.B, w = g.w, n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cx0 = bx * mg.CB;
    const double2 *Te = Tm + (size_t)e * B * B;
    for (int i = threadIdx.x; i < BT * BT; i += blockDim.x) Tf[i] = make_float2((float)Te[i].x, (float)Te[i].y);

    half8 ah[OSW][KS], al[OSW][KS];
#pragma unroll
    for (int jj = 0; jj < OSW; jj++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const size_t fa = (((size_t)(j0 + jj) * KS + ks) * 2) * 64;
            ah[jj][ks] = Afrag[fa + lane];
            al[jj][ks] = Afrag[fa + 64 + lane];
        }
    __syncthreads();

    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int col = lane & 15, kq = lane >> 4;
    const int col0 = cx0 - n2 - w + h + mg.OFF;
    const int rows = mg.CB * OS;
    const double sc = sA * scale[e];
    const double unscale = 1.0 / (sc * sc);
    const int W = g.W;
    double *out = ldos + (size_t)e * W * W;

    // Stage the plane rows this workgroup's tiles read: rows of the cells
    // cy_lo .. cy_lo + TPW*kWaves - 1 plus the w halo, the NUg groups of 8
    // columns from col0, all 4*NPL planes.  Layout [plane][row (RP)][NUg*8],
    // RP odd so that the 16 planes one MFMA column group reads fall on
    // different banks.
    extern __shared__ __align__(16) unsigned char smem_m[];
    _Float16 *Pl = reinterpret_cast<_Float16 *>(smem_m);
    const int RW = TPW * CPS + 2 * w, RP = RW | 1, seg = mg.NUg;
    const int cy_lo = by * TPW * CPS;
    const int row_lo = cy_lo - n2 - w + h;
    // (consecutive threads take consecutive rows: contiguous in the planes;
    // multiply-high divisions instead of runtime ones)
    {
        const unsigned long long mrw = udiv_magic((unsigned)RW), mseg = udiv_magic((unsigned)seg);
        for (int l = threadIdx.x; l < 4 * mg.NPL * RW * seg; l += blockDim.x) {
            const int rest = udiv(l, mrw), r = l - rest * RW;
            const int pid = udiv(rest, mseg), g8 = rest - pid * seg;
            const int row = row_lo + r;
            half8 v = half8{0, 0, 0, 0, 0, 0, 0, 0};
            if (row < nw) v = load8<LDW>(Pe + (size_t)pid * plane, row, col0 + g8 * 8, nw);
            *reinterpret_cast<half8 *>(Pl + ((size_t)pid * RP + r) * seg * 8 + g8 * 8) = v;
        }
    }
    if (threadIdx.x == 0) *reinterpret_cast<half8 *>(Pl + 4 * mg.NPL * RP * seg * 8) = half8{0, 0, 0, 0, 0, 0, 0, 0};
    __syncthreads();

    // per lane and K step: the staged offset (halves) of its B fragment in the
    // hi plane at tile row 0, or -1 where the lane's K group or N column is
    // padding (then it reads the zero slot after the staged planes)
    const int LOFF = mg.NPL * RP * seg * 8;   // hi -> lo plane of the same (o, p)
    const int ZOFF = 4 * mg.NPL * RP * seg * 8;
    int qoff[KS][NT];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int kg = ks * 4 + kq;
        const bool kok = kg < mg.KG;
        const int g8 = kg % mg.NUg, r = kg / mg.NUg;
        const int dy = r / mg.nper, o = r % mg.nper;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int nn = t * 16 + col;
            int ab = -1, pl = 0;
            if (nn < 2 * B) {
                ab = 0;
                pl = (o * B + (nn >> 1)) * 2 + (nn & 1);
            } else if (nn < 4 * B) {
                ab = 1;
                int m2 = nn - 2 * B;
                pl = (o * B + (m2 >> 1)) * 2 + (m2 & 1);
            }
            const int pid = (ab * 2) * mg.NPL + pl;
            qoff[ks][t] = (kok && ab >= 0) ? (pid * RP + dy) * seg * 8 + g8 * 8 : -1;
        }
    }
    // the periodic (bulk) term of each pixel this lane finishes
    constexpr int NJB = (OSW + 3) / 4;
    double bk[NJB];
#pragma unroll
    for (int jb = 0; jb < NJB; jb++) {
        const int jl = jb * 4 + (lane >> 4), ii = (lane & 15) % OS;
        bk[jb] = (jl < OSW) ? bulk[(size_t)e * OS * OS + ii * OS + j0 + jl] : 0.0;
    }

    for (int tt = 0; tt < TPW; tt++) {
        const int cy_raw = (by * TPW + tt) * CPS + wrow;
        if (cy_raw >= n) break;  // uniform per wave; no block-wide barrier follows
        const int cy = cy_raw;
        const int trow = (cy - cy_lo) * seg * 8;
        // all B fragments of the tile first (unconditional LDS reads), then the MFMAs
        half8 bh[KS][NT], bl[KS][NT];
#pragma unroll
        for (int ks = 0; ks < KS; ks++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int q = qoff[ks][t];
                const int a0 = q >= 0 ? q + trow : ZOFF, a1 = q >= 0 ? q + trow + LOFF : ZOFF;
                bh[ks][t] = *reinterpret_cast<const half8 *>(Pl + a0);
                bl[ks][t] = *reinterpret_cast<const half8 *>(Pl + a1);
            }
        float4v hh[OSW][NT], cr[OSW][NT];
#pragma unroll
        for (int i = 0; i < OSW; i++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                hh[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                cr[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int i = 0; i < OSW; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    hh[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i][ks], bh[ks][t], hh[i][t], 0, 0, 0);
                    cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i][ks], bl[ks][t], cr[i][t], 0, 0, 0);
                    cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i][ks], bh[ks][t], cr[i][t], 0, 0, 0);
                }
        }
        // epilogue (wave-local)
#pragma unroll
        for (int i = 0; i < OSW; i++)
#pragma unroll
            for (int t = 0; t < NT; t++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    Cs[wave][i][kq * 4 + r][t * 16 + col] = hh[i][t][r] + cr[i][t][r] * (1.0f / kLoScale);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int jb = 0; jb < (OSW + 3) / 4; jb++) {
            const int jl = jb * 4 + (lane >> 4), jj = j0 + jl, m = lane & 15;
            if (jl < OSW && m < rows) {
                // Im sum_pq a_p T_pq b_q: (T b)_p in FP32, the sum over p in FP64
                const float *Cr = Cs[wave][jl][m];
                float cv[4 * BT];
#pragma unroll
                for (int k = 0; k < 4 * BT; k++) cv[k] = Cr[k];
                double acc = 0.0;
#pragma unroll
                for (int p = 0; p < BT; p++) {
                    float tbx = 0.f, tby = 0.f;
#pragma unroll
                    for (int q = 0; q < BT; q++) {
                        const float2 t = Tf[p * BT + q];
                        const float bx = cv[2 * BT + 2 * q], by = cv[2 * BT + 2 * q + 1];
                        tbx += t.x * bx - t.y * by;
                        tby += t.x * by + t.y * bx;
                    }
                    acc += (double)cv[2 * p] * (double)tby + (double)cv[2 * p + 1] * (double)tbx;
                }
                const int c = m / OS, ii = m - c * OS, cx = cx0 + c;
                if (cx < n) {
                    const int X = cx * OS + ii, Y = cy * OS + jj;
                    out[(size_t)X + (size_t)Y * W] = -(acc * unscale + bk[jb]);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

constexpr int kTPW = 8;  // default cell rows per wave

template <int NT, int LDW, int OS, int KS, int TPW, int BT, int JS>
void launch_areg(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale,
                 double sA, const double2 *T, const double *bulk, int E, double *ldos, hipStream_t s)
{
    constexpr int CPS = kWaves / JS;
    dim3 grid((g.n + mg.CB - 1) / mg.CB, (g.n + CPS * TPW - 1) / (CPS * TPW), E);
    const int RP = (TPW * CPS + 2 * g.w) | 1;
    const size_t lds = ((size_t)4 * mg.NPL * RP * mg.NUg * 8 + 8) * sizeof(_Float16);  // + the zero slot
    hipLaunchKernelGGL((k_cmix_areg<NT, LDW, OS, KS, TPW, BT, JS>), grid, dim3(64 * kWaves), lds, s, g, mg, P, A, scale,
                       sA, T, bulk, ldos);
}

template <int NT, int LDW, int OS, int KS, int BT, int JS>
bool try_areg(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale,
              double sA, const double2 *T, const double *bulk, int E, double *ldos, hipStream_t s)
{
    if (g.os != OS || g.B != BT || mg.KS != KS || mg.NT != NT || mg.nsets != 1 || mg.LDW != LDW) return false;
    switch (g.tpw) {
    case 2: launch_areg<NT, LDW, OS, KS, 2, BT, JS>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s); break;
    case 4: launch_areg<NT, LDW, OS, KS, 4, BT, JS>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s); break;
    default: launch_areg<NT, LDW, OS, KS, kTPW * JS, BT, JS>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s); break;
    }
    return true;
}

// T-folded contraction for one orbital block (LDOS maps, BT <= 5 bands, every
// orbital with a Wannier function).  With c = T b formed before the products,
//   c_p = sum_{R1,o1} (T G0(-pos-R1))[p][o1] w_o1,   rho = -Im sum_p a_p c_p - bulk,
// so no per-pixel T sandwich remains: the workgroup stages, from the FP32
// window W32[e][sy][sx][pair] (pair = r*BT + c holds G0(R)[r][c],
// R = (sx - h, sy - h)), the a rows G0(pos+R)[o][p] and the folded c rows
// (T G0(-pos-R))[p][o] = sum_q T_pq G0(-pos-R)[q][o], split into fp16 hi/lo.
// The MFMA then forms C^T[ch][pix]: its A operand is a window fragment whose
// row is the channel ch = 4p + j (j: a_re, a_im, c_re, c_im), its B operand
// the tile fragment whose column is the pixel, so that lane (kq, col) holds
// the four channels of p = 4t + kq of pixel col and the pixel's sum over p is
// a sum over the four rows of lanes.  The lo parts are stored unscaled (an
// fp16 residual is exact to 2^-24 absolute, far below the hi part's scale), so
// that the three products hi*hi, lo*hi and hi*lo share one accumulator.
// Geometry is compile-time: OS sub-pixels, window WW, BT bands, CB cells per
// tile (CB*OS = 16 rows), K = (dy, o, dx) in groups of 8 columns.
template <int OS, int WW>
struct TfGeom {
    static constexpr int CB = 16 / OS;
    static constexpr int NU = CB + 2 * WW;
    static constexpr int NUG = (NU + 7) / 8;
    static constexpr int W2 = 2 * WW + 1;
};

// waves per workgroup and per SIMD of k_cmix_tf (eight-wave workgroups at
// four waves a SIMD measured 6 % slower for 3 bands)
constexpr int tf_wg_waves(int) { return kWaves; }
constexpr int tf_simd_waves(int bt) { return areg_waves(bt); }

template <int OS, int WW, int BT, int TPW, int JS>
__global__ void __launch_bounds__(64 * tf_wg_waves(BT)) __attribute__((amdgpu_waves_per_eu(tf_simd_waves(BT), tf_simd_waves(BT)))) k_cmix_tf(Geom g, const float2 *__restrict__ W32,
                                                         const half8 *__restrict__ Afrag,
                                                         const double *__restrict__ scale, double sA,
                                                         const double2 *__restrict__ Tm,
                                                         const double *__restrict__ bulk,
                                                         double *__restrict__ ldos)
{
    using TG = TfGeom<OS, WW>;
    constexpr int CB = TG::CB, NUG = TG::NUG, SEG8 = NUG * 8;
    constexpr int KG = TG::W2 * BT * NUG, KS = (KG + 3) / 4, NT = (4 * BT + 15) / 16;
    constexpr int OSW = OS / JS, CPS = tf_wg_waves(BT) / JS;
    constexpr int NP = BT * BT, NPL = NP * 2;
    constexpr int RW = TPW * CPS + 2 * WW, RP = RW | 1;
    constexpr int PSTR = RP * SEG8;             // halves per staged plane
    constexpr int LOFF = NPL * PSTR, ZOFF = 4 * NPL * PSTR;
    __shared__ float2 Tf[NP];
    __shared__ float sTsh;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j0 = (wave % JS) * OSW, wrow = wave / JS;
    int bx, by, e;
    xcd_block(bx, by, e);
    const int n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cx0 = bx * CB;
    const double sG = scale[e];
    const double2 *Te = Tm + (size_t)e * NP;
    if (threadIdx.x < NP) Tf[threadIdx.x] = make_float2((float)Te[threadIdx.x].x, (float)Te[threadIdx.x].y);
    if (threadIdx.x == 0) {
        // power-of-two scale of the folded rows: |(T G0)_po| <= max_p sum_q |T_pq| * max|G0|,
        // and max|G0| sG < 2^13
        double ts = 0.0;
        for (int p = 0; p < BT; p++) {
            double r = 0.0;
            for (int q = 0; q < BT; q++) r += fabs(Te[p * BT + q].x) + fabs(Te[p * BT + q].y);
            ts = fmax(ts, r);
        }
        double sT = 1.0;
        if (ts > 0.0) {
            int ex;
            frexp(ts * 8192.0 / sG, &ex);
            sT = ldexp(1.0, 13 - ex);
        }
        sTsh = (float)sT;
    }

    __syncthreads();
    const float sGf = (float)sG, sTf = sTsh;
    const double unscale = 1.0 / (sA * sA * sG * (double)sTf);
    const int W = g.W;
    double *out = ldos + (size_t)e * W * W;

    // Stage the rows of the workgroup's tiles: plane rows row_lo .. + RW - 1,
    // window columns sx0 .. + 8 NUG - 1, for every (o, p): the a rows and the
    // folded c rows, hi and lo, re and im.  LDS layout [plane][row (RP)][8 NUG]
    // with plane = ((a|c) * 2 + hi|lo) * NPL + (o BT + p) * 2 + re|im.
    extern __shared__ __align__(16) unsigned char smem_m[];
    _Float16 *Pl = reinterpret_cast<_Float16 *>(smem_m);
    const int cy_lo = by * TPW * CPS;
    const int row_lo = cy_lo - n2 - WW + h;
    const int sx0 = cx0 - n2 - WW + h;
    const float2 *We = W32 + (size_t)e * nw * nw * NP;
    // one item per (row, column group, orbital o): the b row G0(-pos-R)[q][o]
    // is loaded once for the BT folds (T G0)[p][o], p < BT
    for (int it = threadIdx.x; it < RW * NUG * BT; it += blockDim.x) {
        const int o = it % BT, rest = it / BT;
        const int g8 = rest % NUG, r = rest / NUG;
        const int row = row_lo + r;
        const int rowc = min(max(row, 0), nw - 1);
        const bool rok = row == rowc;
        const float2 *arow = We + (size_t)rowc * nw * NP + o * BT;
        const float2 *brow = We + (size_t)(2 * h - rowc) * nw * NP + o;
        // every load first (indices clamped into the window, values outside it
        // dropped below), then the arithmetic
        float2 av[8][BT], bv[8][BT];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int sxc = min(max(sx0 + g8 * 8 + u, 0), nw - 1);
#pragma unroll
            for (int q = 0; q < BT; q++) {
                av[u][q] = arow[(size_t)sxc * NP + q];
                bv[u][q] = brow[(size_t)(2 * h - sxc) * NP + q * BT];
            }
        }
        bool okv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int sx = sx0 + g8 * 8 + u;
            okv[u] = rok && sx >= 0 && sx < nw;
        }
#pragma unroll
        for (int p = 0; p < BT; p++) {
            float2 c0[BT];
#pragma unroll
            for (int q = 0; q < BT; q++) c0[q] = Tf[p * BT + q];
            half8 ahr, alr, ahi, ali, chr, clr, chi, cli;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const bool ok = okv[u];
                float yr = 0.f, yi = 0.f;
#pragma unroll
                for (int q = 0; q < BT; q++) {
                    yr += c0[q].x * bv[u][q].x - c0[q].y * bv[u][q].y;
                    yi += c0[q].x * bv[u][q].y + c0[q].y * bv[u][q].x;
                }
                const float xr = ok ? av[u][p].x * sGf : 0.f, xi = ok ? av[u][p].y * sGf : 0.f;
                yr = ok ? yr * sTf : 0.f;
                yi = ok ? yi * sTf : 0.f;
                // v - hi is exact in FP32 (power-of-two scales)
                _Float16 t;
                t = (_Float16)xr; ahr[u] = t; alr[u] = (_Float16)(xr - (float)t);
                t = (_Float16)xi; ahi[u] = t; ali[u] = (_Float16)(xi - (float)t);
                t = (_Float16)yr; chr[u] = t; clr[u] = (_Float16)(yr - (float)t);
                t = (_Float16)yi; chi[u] = t; cli[u] = (_Float16)(yi - (float)t);
            }
            _Float16 *dst = Pl + (size_t)((o * BT + p) * 2) * PSTR + r * SEG8 + g8 * 8;
            *reinterpret_cast<half8 *>(dst) = ahr;
            *reinterpret_cast<half8 *>(dst + PSTR) = ahi;
            *reinterpret_cast<half8 *>(dst + LOFF) = alr;
            *reinterpret_cast<half8 *>(dst + LOFF + PSTR) = ali;
            *reinterpret_cast<half8 *>(dst + 2 * LOFF) = chr;
            *reinterpret_cast<half8 *>(dst + 2 * LOFF + PSTR) = chi;
            *reinterpret_cast<half8 *>(dst + 3 * LOFF) = clr;
            *reinterpret_cast<half8 *>(dst + 3 * LOFF + PSTR) = cli;
        }
    }
    if (threadIdx.x == 0) *reinterpret_cast<half8 *>(Pl + ZOFF) = half8{0, 0, 0, 0, 0, 0, 0, 0};
    __syncthreads();

    half8 ah[OSW][KS], al[OSW][KS];
#pragma unroll
    for (int jj = 0; jj < OSW; jj++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const size_t fa = (((size_t)(j0 + jj) * KS + ks) * 2) * 64;
            ah[jj][ks] = Afrag[fa + lane];
            al[jj][ks] = Afrag[fa + 64 + lane];
        }
    // per lane and K step: the staged offset of its window fragment (channel
    // ch = t*16 + col) in the hi plane at tile row 0, or the zero slot where
    // the K group or the channel is padding
    const int col = lane & 15, kq = lane >> 4;
    int qoff[KS][NT];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int kg = ks * 4 + kq;
        const int g8 = kg % NUG, r = kg / NUG;
        const int dy = r / BT, o = r % BT;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int ch = t * 16 + col, p = ch >> 2, j = ch & 3;
            const int pid = (j >> 1) * 2 * NPL + (o * BT + p) * 2 + (j & 1);
            qoff[ks][t] = (kg < KG && p < BT) ? (pid * RP + dy) * SEG8 + g8 * 8 : -1;
        }
    }
    // the pixel of this lane's column and the periodic term of its sub-pixel rows
    const int c = col / OS, ii = col - c * OS, cx = cx0 + c;
    double bk[OSW];
#pragma unroll
    for (int i = 0; i < OSW; i++) bk[i] = bulk[(size_t)e * OS * OS + ii * OS + j0 + i];

    for (int tt = 0; tt < TPW; tt++) {
        const int cy = (by * TPW + tt) * CPS + wrow;
        if (cy >= n) break;  // uniform per wave; no block-wide barrier follows
        const int trow = (cy - cy_lo) * SEG8;
        half8 bh[KS][NT], bl[KS][NT];
#pragma unroll
        for (int ks = 0; ks < KS; ks++)
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int q = qoff[ks][t];
                const int a0 = q >= 0 ? q + trow : ZOFF, a1 = q >= 0 ? q + trow + LOFF : ZOFF;
                bh[ks][t] = *reinterpret_cast<const half8 *>(Pl + a0);
                bl[ks][t] = *reinterpret_cast<const half8 *>(Pl + a1);
            }
        float4v acc[OSW][NT];
#pragma unroll
        for (int i = 0; i < OSW; i++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int i = 0; i < OSW; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh[ks][t], ah[i][ks], acc[i][t], 0, 0, 0);
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bl[ks][t], ah[i][ks], acc[i][t], 0, 0, 0);
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh[ks][t], al[i][ks], acc[i][t], 0, 0, 0);
                }
        }
        // epilogue: a_re c_im + a_im c_re of this lane's p, summed over the
        // four lane rows (p = 4t + kq).  FP32 adds rounding of the order of the
        // accumulation's own; the row sums use the lane swaps of gfx950
        // (a swap's two results are, per lane, x[l] and x[l ^ 16] (x[l ^ 32]))
        float val[OSW];
#pragma unroll
        for (int i = 0; i < OSW; i++) {
            float part = acc[i][0][0] * acc[i][0][3] + acc[i][0][1] * acc[i][0][2];
#pragma unroll
            for (int t = 1; t < NT; t++) part += acc[i][t][0] * acc[i][t][3] + acc[i][t][1] * acc[i][t][2];
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            val[i] = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
        }
        // lane row kq stores the sub-pixel rows j0 + 4 ib + kq
#pragma unroll
        for (int ib = 0; ib < (OSW + 3) / 4; ib++) {
            float mine = 0.f;
            double b = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++)
                if (kq == k2 && ib * 4 + k2 < OSW) {
                    mine = val[ib * 4 + k2];
                    b = bk[ib * 4 + k2];
                }
            const int i = ib * 4 + kq;
            if (i < OSW && cx < n) {
                const int X = cx * OS + ii, Y = cy * OS + j0 + i;
                out[(size_t)X + (size_t)Y * W] = -((double)mine * unscale + b);
            }
        }
    }
}

template <int OS, int WW, int BT, int TPW, int JS>
void launch_tf(const Geom &g, const float2 *W32, const half8 *A, const double *scale, double sA, const double2 *T,
               const double *bulk, int E, double *ldos, hipStream_t s)
{
    using TG = TfGeom<OS, WW>;
    constexpr int CPS = tf_wg_waves(BT) / JS, RW = TPW * CPS + 2 * WW, RP = RW | 1;
    dim3 grid((g.n + TG::CB - 1) / TG::CB, (g.n + CPS * TPW - 1) / (CPS * TPW), E);
    const size_t lds = ((size_t)4 * 2 * BT * BT * RP * TG::NUG * 8 + 8) * sizeof(_Float16);  // + the zero slot
    hipLaunchKernelGGL((k_cmix_tf<OS, WW, BT, TPW, JS>), grid, dim3(64 * tf_wg_waves(BT)), lds, s, g, W32, A, scale, sA, T,
                       bulk, ldos);
}

template <int OS, int WW, int BT, int JS>
bool tf_shape(const Geom &g, const MixGeom &mg)
{
    using TG = TfGeom<OS, WW>;
    // (the compile-time geometry must be the context's)
    return g.os == OS && g.w == WW && g.B == BT && g.maxb == BT && !g.spin && mg.nsets == 1 && mg.CB == TG::CB &&
           mg.NUg == TG::NUG && mg.nper == BT;
}

template <int OS, int WW, int BT, int JS>
void run_tf(const Geom &g, const float2 *W32, const half8 *A, const double *scale, double sA, const double2 *T,
            const double *bulk, int E, double *ldos, hipStream_t s)
{
    switch (g.tpw) {
    case 2: launch_tf<OS, WW, BT, 2, JS>(g, W32, A, scale, sA, T, bulk, E, ldos, s); break;
    case 4: launch_tf<OS, WW, BT, 4, JS>(g, W32, A, scale, sA, T, bulk, E, ldos, s); break;
    default: launch_tf<OS, WW, BT, kTPW * JS * 4 / tf_wg_waves(BT), JS>(g, W32, A, scale, sA, T, bulk, E, ldos, s); break;
    }
}

// JS: the shape's default waves per cell row; g.js (autotuner) may pick 1 or 2
template <int OS, int WW, int BT, int JS>
bool try_tf(const Geom &g, const MixGeom &mg, const float2 *W32, const half8 *A, const double *scale, double sA,
            const double2 *T, const double *bulk, int E, double *ldos, hipStream_t s)
{
    if (!tf_shape<OS, WW, BT, JS>(g, mg)) return false;
    if ((g.js ? g.js : JS) == 1)
        run_tf<OS, WW, BT, 1>(g, W32, A, scale, sA, T, bulk, E, ldos, s);
    else
        run_tf<OS, WW, BT, 2>(g, W32, A, scale, sA, T, bulk, E, ldos, s);
    return true;
}

// the FP64 window [e][sx][sy][pair] of the drop-in path as the FP32 window
// W32[e][sy][sx][pair] of the T-folded contraction
__global__ void k_pack32(Geom g, const double2 *__restrict__ G0w, float2 *__restrict__ W32)
{
    const int sy = blockIdx.x, e = blockIdx.y, nw = g.nw, NP = g.B * g.B;
    const double2 *G = G0w + (size_t)e * nw * nw * NP;
    float2 *o = W32 + ((size_t)e * nw + sy) * nw * NP;
    for (int t = threadIdx.x; t < nw * NP; t += blockDim.x) {
        const int sx = t / NP, q = t - sx * NP;
        const double2 z = G[((size_t)sx * nw + sy) * NP + q];
        o[t] = make_float2((float)z.x, (float)z.y);
    }
}

// Generic shapes with the B operand staged in LDS (spin blocks, many bands,
// any os): one workgroup of kStWaves waves covers kStWaves consecutive cell
// rows (one per wave) and one tile of CB cells; for each spin block it copies
// the plane rows those tiles reach (kStWaves + 2w rows, the spin block's
// orbitals only) into LDS, then every wave reads its B fragments from there,
// padding lanes from a zero slot.  A fragments come from global memory (L2),
// as in k_cmix.
constexpr int kStWaves = 8;

size_t cmix_st_lds(const MixGeom &mg, const Geom &g, int NT, int JB)
{
    const int RP = (kStWaves + 2 * g.w) | 1;
    const size_t planes = (size_t)4 * mg.nper * g.B * 2;  // (a|b, hi|lo) x the block's (o, p, re/im)
    return (planes * RP * mg.NUg * 8 + 8) * sizeof(_Float16) + (size_t)kStWaves * JB * 16 * (16 * NT + 1) * sizeof(float);
}

// J: the Josephson map, as in k_cmix (the b planes of the spin block staged
// from the particle-hole columns o + maxband).
// cscale: as k_cmix.
template <int NT, int LDW, int JB, bool J>
__global__ void __launch_bounds__(64 * kStWaves) k_cmix_st(Geom g, MixGeom mg, const _Float16 *__restrict__ P,
                                                          const half8 *__restrict__ Afrag,
                                                          const double *__restrict__ scale, double sA,
                                                          const double2 *__restrict__ Tm,
                                                          const double *__restrict__ bulk,
                                                          double *__restrict__ ldos,
                                                          const double2 *__restrict__ tips,
                                                          const double *__restrict__ cscale)
{
    constexpr int NB = 16 / JB;  // jj blocks (os <= 16)
    __shared__ double2 Ts[16 * 16];
    extern __shared__ __align__(16) unsigned char smem_st[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int bx, by, e;
    xcd_block(bx, by, e);
    const int B = g.B, os = g.os, w = g.w, n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cy_lo = by * kStWaves;
    const int cy_raw = cy_lo + wave;
    const bool cy_ok = cy_raw < n;
    const int cy = cy_ok ? cy_raw : n - 1;
    const int cx0 = bx * mg.CB;
    const double2 *Te = Tm + (size_t)e * B * B;
    for (int i = threadIdx.x; i < B * B; i += blockDim.x) Ts[i] = Te[i];

    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int col = lane & 15, kq = lane >> 4;
    const int col0 = cx0 - n2 - w + h + mg.OFF;
    const int rows = mg.CB * os;
    const int RW = kStWaves + 2 * w, RP = RW | 1, seg = mg.NUg;
    const int row_lo = cy_lo - n2 - w + h;
    const int NPS = mg.nper * B * 2;                    // planes of one spin block per (a|b, hi|lo)
    _Float16 *Pl = reinterpret_cast<_Float16 *>(smem_st);
    const int LOFF = NPS * RP * seg * 8;                // hi -> lo plane of the same (o, p)
    const int ZOFF = 4 * NPS * RP * seg * 8;            // 16-byte zero slot
    float *Csw = reinterpret_cast<float *>(Pl + ZOFF + 8) + (size_t)wave * JB * 16 * (16 * NT + 1);
    const int trow = (cy - cy_lo) * seg * 8;
    const unsigned long long mrw = udiv_magic((unsigned)RW), mseg = udiv_magic((unsigned)seg),
                             mnps = udiv_magic((unsigned)NPS);
    double val[NB], valr[NB];  // Im and (J) Re of sum_s a^T T b
#pragma unroll
    for (int jb = 0; jb < NB; jb++) val[jb] = valr[jb] = 0.0;
    const int jo = J ? g.maxb : 0;  // orbital offset of the b planes

    for (int sg = 0; sg < mg.nsets; sg++) {
        __syncthreads();  // the previous spin block's fragments are read
        // (consecutive threads take consecutive rows: contiguous in the planes)
        for (int l = threadIdx.x; l < 4 * NPS * RW * seg; l += blockDim.x) {
            const int rest = udiv(l, mrw), r = l - rest * RW;
            const int pid = udiv(rest, mseg), g8 = rest - pid * seg;
            const int abhl = udiv(pid, mnps);
            const int pl = sg * NPS + (pid - abhl * NPS) + (abhl >= 2 ? jo * B * 2 : 0);
            const int row = row_lo + r;
            half8 v = half8{0, 0, 0, 0, 0, 0, 0, 0};
            if (row < nw) v = load8<LDW>(Pe + ((size_t)abhl * mg.NPL + pl) * plane, row, col0 + g8 * 8, nw);
            *reinterpret_cast<half8 *>(Pl + ((size_t)pid * RP + r) * seg * 8 + g8 * 8) = v;
        }
        if (threadIdx.x == 0) *reinterpret_cast<half8 *>(Pl + ZOFF) = half8{0, 0, 0, 0, 0, 0, 0, 0};
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            const int jj0 = jb * JB;
            if (jj0 >= os) break;
            float4v hh[JB][NT], cr[JB][NT];
#pragma unroll
            for (int i = 0; i < JB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    hh[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                    cr[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                }
            // the tile fragments of the next K step are loaded while this
            // step's MFMAs run (they come from L2, one pair per sub-pixel row)
            half8 ahn[JB], aln[JB];
#pragma unroll
            for (int i = 0; i < JB; i++) {
                const int jj = min(jj0 + i, os - 1);
                const size_t fa = ((((size_t)sg * os + jj) * mg.KS + 0) * 2) * 64;
                ahn[i] = Afrag[fa + lane];
                aln[i] = Afrag[fa + 64 + lane];
            }
            for (int ks = 0; ks < mg.KS; ks++) {
                half8 ahc[JB], alc[JB];
#pragma unroll
                for (int i = 0; i < JB; i++) {
                    ahc[i] = ahn[i];
                    alc[i] = aln[i];
                }
                if (ks + 1 < mg.KS) {
#pragma unroll
                    for (int i = 0; i < JB; i++) {
                        const int jj = min(jj0 + i, os - 1);
                        const size_t fa = ((((size_t)sg * os + jj) * mg.KS + ks + 1) * 2) * 64;
                        ahn[i] = Afrag[fa + lane];
                        aln[i] = Afrag[fa + 64 + lane];
                    }
                }
                const int kg = ks * 4 + kq;
                const bool kok = kg < mg.KG;
                const int g8 = kg % mg.NUg, r = kg / mg.NUg;
                const int ol = r % mg.nper, dy = r / mg.nper;
                half8 bh[NT], bl[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int nn = t * 16 + col;
                    int ab = -1, pl = 0;
                    if (nn < 2 * B) {
                        ab = 0;
                        pl = (ol * B + (nn >> 1)) * 2 + (nn & 1);
                    } else if (nn < 4 * B) {
                        ab = 1;
                        int m2 = nn - 2 * B;
                        pl = (ol * B + (m2 >> 1)) * 2 + (m2 & 1);
                    }
                    const bool ok = kok && ab >= 0;
                    const int q = (((ab * 2) * NPS + pl) * RP + dy) * seg * 8 + g8 * 8 + trow;
                    bh[t] = *reinterpret_cast<const half8 *>(Pl + (ok ? q : ZOFF));
                    bl[t] = *reinterpret_cast<const half8 *>(Pl + (ok ? q + LOFF : ZOFF));
                }
#pragma unroll
                for (int i = 0; i < JB; i++) {
                    const int jj = jj0 + i;
                    if (jj >= os) break;
                    const half8 ah = ahc[i], al = alc[i];
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        hh[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[t], hh[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[t], cr[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[t], cr[i][t], 0, 0, 0);
                    }
                }
            }
            // epilogue (wave-local): C -> LDS, then one pixel per lane:
            // Im sum_pq a_p T_pq b_q in FP64
#pragma unroll
            for (int i = 0; i < JB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        Csw[((size_t)i * 16 + kq * 4 + r) * (16 * NT + 1) + t * 16 + col] =
                            hh[i][t][r] + cr[i][t][r] * (1.0f / kLoScale);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                const int i = lane >> 4, m = lane & 15;
                if (i < JB && jj0 + i < os && m < rows) {
                    const float *Cr = Csw + ((size_t)i * 16 + m) * (16 * NT + 1);
                    double acc = 0.0, accr = 0.0;
                    if (!J && cscale) {
                        for (int p = 0; p < B; p++)
                            acc += (double)Cr[2 * p] * (double)Cr[2 * B + 2 * p + 1] +
                                   (double)Cr[2 * p + 1] * (double)Cr[2 * B + 2 * p];
                    } else
                    for (int p = 0; p < B; p++) {
                        double2 tb = make_double2(0.0, 0.0);
                        for (int q = 0; q < B; q++)
                            tb = cadd(tb, cmul(Ts[p * B + q], make_double2(Cr[2 * B + 2 * q], Cr[2 * B + 2 * q + 1])));
                        acc += (double)Cr[2 * p] * tb.y + (double)Cr[2 * p + 1] * tb.x;
                        if (J) accr += (double)Cr[2 * p] * tb.x - (double)Cr[2 * p + 1] * tb.y;
                    }
                    val[jb] += acc;
                    if (J) valr[jb] += accr;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    const double unscale = 1.0 / (sA * sA * scale[e] * (!J && cscale ? cscale[e] : scale[e]));
    const int W = g.W;
    double *out = ldos + (size_t)e * W * W;
#pragma unroll
    for (int jb = 0; jb < NB; jb++) {
        const int i = lane >> 4, jj = jb * JB + i, m = lane & 15;
        if (i >= JB || jj >= os || m >= rows || !cy_ok) continue;
        const int c = m / os, ii = m - c * os, cx = cx0 + c;
        if (cx >= n) continue;
        const int X = cx * os + ii, Y = cy * os + jj;
        if (J) {
            const double2 bz = reinterpret_cast<const double2 *>(bulk)[(size_t)e * os * os + ii * os + jj];
            const double2 tp = tips[e];
            const double zr = valr[jb] * unscale + bz.x, zi = val[jb] * unscale + bz.y;
            out[(size_t)X + (size_t)Y * W] = -(tp.x * zi + tp.y * zr);
        } else {
            out[(size_t)X + (size_t)Y * W] = -(val[jb] * unscale + bulk[(size_t)e * os * os + ii * os + jj]);
        }
    }
}

// LDOS maps with T folded into the b planes (k_fold_c), staged as k_cmix_st,
// with the channels of band p on four consecutive columns (a_re, a_im, c_re,
// c_im): Im(a_p c_p) = a_re c_im + a_im c_re is formed inside each group of
// four lanes by lane shuffles, summed over the four bands of a column tile by
// two more and over the tiles in registers.  No LDS transpose: the LDS holds
// only the staged planes, so JB sub-pixel rows (up to the whole cell) share
// each B fragment.  Lane (kq, col < 4) owns pixel 4 kq + col of the tile.
template <int NT, int LDW, int JB>
__global__ void __launch_bounds__(64 * kStWaves) k_cmix_stq(Geom g, MixGeom mg, const _Float16 *__restrict__ P,
                                                           const half8 *__restrict__ Afrag,
                                                           const double *__restrict__ scale, double sA,
                                                           const double *__restrict__ bulk,
                                                           double *__restrict__ ldos,
                                                           const double *__restrict__ cscale)
{
    constexpr int NB = (16 + JB - 1) / JB;  // jj blocks (os <= 16)
    extern __shared__ __align__(16) unsigned char smem_st[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int bx, by, e;
    xcd_block(bx, by, e);
    const int B = g.B, os = g.os, w = g.w, n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cy_lo = by * kStWaves;
    const int cy_raw = cy_lo + wave;
    const bool cy_ok = cy_raw < n;
    const int cy = cy_ok ? cy_raw : n - 1;
    const int cx0 = bx * mg.CB;
    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int col = lane & 15, kq = lane >> 4;
    const int col0 = cx0 - n2 - w + h + mg.OFF;
    const int rows = mg.CB * os;
    const int RW = kStWaves + 2 * w, RP = RW | 1, seg = mg.NUg;
    const int row_lo = cy_lo - n2 - w + h;
    const int NPS = mg.nper * B * 2;                    // planes of one spin block per (a|b, hi|lo)
    _Float16 *Pl = reinterpret_cast<_Float16 *>(smem_st);
    const int LOFF = NPS * RP * seg * 8;                // hi -> lo plane of the same (o, p)
    const int ZOFF = 4 * NPS * RP * seg * 8;            // 16-byte zero slot
    const int trow = (cy - cy_lo) * seg * 8;
    const unsigned long long mrw = udiv_magic((unsigned)RW), mseg = udiv_magic((unsigned)seg),
                             mnps = udiv_magic((unsigned)NPS);
    // this lane's pixel; its sum over the spin blocks is kept in the map
    // itself (one lane owns each pixel): raw after the first block, scaled
    // with the periodic term after the last
    const double unscale = 1.0 / (sA * sA * scale[e] * cscale[e]);
    const int W = g.W;
    double *out = ldos + (size_t)e * W * W;
    const int m = 4 * kq + col;
    const int mc = m / os, ii = m - mc * os, cx = cx0 + mc;
    const bool owner = col < 4 && m < rows && cy_ok && cx < n;

    for (int sg = 0; sg < mg.nsets; sg++) {
        __syncthreads();  // the previous spin block's fragments are read
        for (int l = threadIdx.x; l < 4 * NPS * RW * seg; l += blockDim.x) {
            const int rest = udiv(l, mrw), r = l - rest * RW;
            const int pid = udiv(rest, mseg), g8 = rest - pid * seg;
            const int abhl = udiv(pid, mnps);
            const int pl = sg * NPS + (pid - abhl * NPS);
            const int row = row_lo + r;
            half8 v = half8{0, 0, 0, 0, 0, 0, 0, 0};
            if (row < nw) v = load8<LDW>(Pe + ((size_t)abhl * mg.NPL + pl) * plane, row, col0 + g8 * 8, nw);
            *reinterpret_cast<half8 *>(Pl + ((size_t)pid * RP + r) * seg * 8 + g8 * 8) = v;
        }
        if (threadIdx.x == 0) *reinterpret_cast<half8 *>(Pl + ZOFF) = half8{0, 0, 0, 0, 0, 0, 0, 0};
        __syncthreads();
#pragma unroll 1
        for (int jb = 0; jb < NB; jb++) {
            const int jj0 = jb * JB;
            if (jj0 >= os) break;
            float4v hh[JB][NT], cr[JB][NT];
#pragma unroll
            for (int i = 0; i < JB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    hh[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                    cr[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                }
            // the tile fragments of the next K step are loaded while this
            // step's MFMAs run
            half8 ahn[JB], aln[JB];
#pragma unroll
            for (int i = 0; i < JB; i++) {
                const int jj = min(jj0 + i, os - 1);
                const size_t fa = ((((size_t)sg * os + jj) * mg.KS + 0) * 2) * 64;
                ahn[i] = Afrag[fa + lane];
                aln[i] = Afrag[fa + 64 + lane];
            }
            for (int ks = 0; ks < mg.KS; ks++) {
                half8 ahc[JB], alc[JB];
#pragma unroll
                for (int i = 0; i < JB; i++) {
                    ahc[i] = ahn[i];
                    alc[i] = aln[i];
                }
                if (ks + 1 < mg.KS) {
#pragma unroll
                    for (int i = 0; i < JB; i++) {
                        const int jj = min(jj0 + i, os - 1);
                        const size_t fa = ((((size_t)sg * os + jj) * mg.KS + ks + 1) * 2) * 64;
                        ahn[i] = Afrag[fa + lane];
                        aln[i] = Afrag[fa + 64 + lane];
                    }
                }
                const int kg = ks * 4 + kq;
                const bool kok = kg < mg.KG;
                const int g8 = kg % mg.NUg, r = kg / mg.NUg;
                const int ol = r % mg.nper, dy = r / mg.nper;
                half8 bh[NT], bl[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    // column t*16 + col = 4 p + j: j = a_re, a_im, c_re, c_im
                    const int nn = t * 16 + col, p = nn >> 2, j = nn & 3;
                    const bool ok = kok && p < B;
                    const int pl = (ol * B + p) * 2 + (j & 1);
                    const int q = ((((j >> 1) * 2) * NPS + pl) * RP + dy) * seg * 8 + g8 * 8 + trow;
                    bh[t] = *reinterpret_cast<const half8 *>(Pl + (ok ? q : ZOFF));
                    bl[t] = *reinterpret_cast<const half8 *>(Pl + (ok ? q + LOFF : ZOFF));
                }
#pragma unroll
                for (int i = 0; i < JB; i++) {
                    const half8 ah = ahc[i], al = alc[i];
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        hh[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[t], hh[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[t], cr[i][t], 0, 0, 0);
                        cr[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[t], cr[i][t], 0, 0, 0);
                    }
                }
            }
            // epilogue: lane (kq, col) holds pixel rows 4 kq + r of column
            // t*16 + col; every lane of a group of four gets the band's
            // a_re c_im + a_im c_re, the groups' sums then add over xor 4, 8
#pragma unroll
            for (int i = 0; i < JB; i++) {
                float sr[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float v = hh[i][t][r] + cr[i][t][r] * (1.0f / kLoScale);
                        const float x1 = __shfl_xor(v, 1), x2 = __shfl_xor(v, 2), x3 = __shfl_xor(v, 3);
                        sr[r] += v * x3 + x1 * x2;
                    }
                float mine = 0.f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    float x = sr[r];
                    x += __shfl_xor(x, 4);
                    x += __shfl_xor(x, 8);
                    if ((col & 3) == r) mine = x;
                }
                const int jj = jj0 + i;
                if (owner && jj < os) {
                    double *o = out + (size_t)(cx * os + ii) + (size_t)(cy * os + jj) * W;
                    double v = (double)mine;
                    if (sg > 0) v += *o;
                    *o = sg + 1 < mg.nsets ? v : -(v * unscale + bulk[(size_t)e * os * os + ii * os + jj]);
                }
            }
        }
    }
}

size_t cmix_stq_lds(const MixGeom &mg, const Geom &g)
{
    const int RP = (kStWaves + 2 * g.w) | 1;
    const size_t planes = (size_t)4 * mg.nper * g.B * 2;
    return (planes * RP * mg.NUg * 8 + 8) * sizeof(_Float16);
}

template <int NT, int JB>
void launch_stq(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale, double sA,
                const double *bulk, int E, double *ldos, hipStream_t s, const double *cscale)
{
    dim3 grid((g.n + mg.CB - 1) / mg.CB, (g.n + kStWaves - 1) / kStWaves, E);
    dim3 blk(64 * kStWaves);
    const size_t lds = cmix_stq_lds(mg, g);
#define CQPI_STQ(ldw) \
    hipLaunchKernelGGL((k_cmix_stq<NT, ldw, JB>), grid, blk, lds, s, g, mg, P, A, scale, sA, bulk, ldos, cscale)
    switch (mg.LDW) {
    case 8: CQPI_STQ(8); break;
    case 4: CQPI_STQ(4); break;
    case 2: CQPI_STQ(2); break;
    default: CQPI_STQ(1); break;
    }
#undef CQPI_STQ
}

template <int NT, int JB, bool J>
void launch_st_j(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale, double sA,
                 const double2 *T, const double *bulk, int E, double *ldos, const double2 *tips, hipStream_t s,
                 const double *cscale)
{
    dim3 grid((g.n + mg.CB - 1) / mg.CB, (g.n + kStWaves - 1) / kStWaves, E);
    dim3 blk(64 * kStWaves);
    const size_t lds = cmix_st_lds(mg, g, NT, JB);
#define CQPI_ST(ldw) \
    hipLaunchKernelGGL((k_cmix_st<NT, ldw, JB, J>), grid, blk, lds, s, g, mg, P, A, scale, sA, T, bulk, ldos, tips, \
                       cscale)
    switch (mg.LDW) {
    case 8: CQPI_ST(8); break;
    case 4: CQPI_ST(4); break;
    case 2: CQPI_ST(2); break;
    default: CQPI_ST(1); break;
    }
#undef CQPI_ST
}

// folded LDOS maps: sub-pixel rows per B fragment of the shuffle epilogue
#ifndef CQPI_STQ_JB
#define CQPI_STQ_JB 3
#endif

template <int NT, int JB>
int launch_st(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale, double sA,
              const double2 *T, const double *bulk, int E, double *ldos, const double2 *tips, hipStream_t s,
              const double *cscale)
{
    if (tips)
        launch_st_j<NT, JB, true>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, nullptr);
    else if (cscale)
        launch_stq<NT, CQPI_STQ_JB>(g, mg, P, A, scale, sA, bulk, E, ldos, s, cscale);
    else
        launch_st_j<NT, JB, false>(g, mg, P, A, scale, sA, T, bulk, E, ldos, nullptr, s, cscale);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

template <int NT, bool J>
void launch_nt_j(const Geom &g, const MixGeom &mg, const _Float16 *P, const half8 *A, const double *scale, double sA,
                 const double2 *T, const double *bulk, int E, double *ldos, const double2 *tips, hipStream_t s,
                 const double *cscale)
{
    dim3 grid((g.n + mg.CB - 1) / mg.CB, (g.n + kWaves - 1) / kWaves, E);
    dim3 blk(64 * kWaves);
    // accumulator blocks of JB sub-pixel rows: fewer for wide N, so that the
    // accumulators leave room for more than one wave per SIMD
    constexpr int JB = NT >= 3 ? 2 : 4;
#define CQPI_NT(ldw) \
    hipLaunchKernelGGL((k_cmix<NT, ldw, JB, J>), grid, blk, 0, s, g, mg, P, A, scale, sA, T, bulk, ldos, tips, cscale)
    switch (mg.LDW) {
    case 8: CQPI_NT(8); break;
    case 4: CQPI_NT(4); break;
    case 2: CQPI_NT(2); break;
    default: CQPI_NT(1); break;
    }
#undef CQPI_NT
}

template <int NT>
int launch_nt(const Geom &g, const MixGeom &mg, const _Float16 *P, const void *Afrag, const double *scale,
              double sA, const double2 *T, const double *bulk, int E, double *ldos, const double2 *tips, hipStream_t s,
              const double *cscale)
{
    const half8 *A = (const half8 *)Afrag;
    if (tips)
        launch_nt_j<NT, true>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, nullptr);
    else
        launch_nt_j<NT, false>(g, mg, P, A, scale, sA, T, bulk, E, ldos, nullptr, s, cscale);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

}  // namespace

namespace cqpi {

MixGeom mix_geometry(const Geom &g)
{
    MixGeom mg{};
    mg.CB = std::max(1, 16 / g.os);
    if (mg.CB > 1 && (mg.CB & 1)) mg.CB -= 1;  // even CB keeps plane rows 4-byte aligned
    mg.nsets = g.spin ? 2 : 1;
    mg.nper = g.spin ? (g.maxb >> 1) : g.maxb;
    mg.NU = mg.CB + 2 * g.w;
    mg.NUg = (mg.NU + 7) / 8;
    mg.KG = (2 * g.w + 1) * mg.nper * mg.NUg;
    mg.KS = (mg.KG + 3) / 4;
    mg.NT = (4 * g.B + 15) / 16;
    mg.NPL = g.B * g.B * 2;
    mg.LDW = (mg.CB % 8 == 0) ? 8 : (mg.CB % 4 == 0) ? 4 : (mg.CB % 2 == 0) ? 2 : 1;
    int base = g.h - (g.n >> 1) - g.w;  // plane index of tile row 0 before the offset
    mg.OFF = ((-base) % mg.LDW + mg.LDW) % mg.LDW;
    int ntile_y = (g.n + mg.CB - 1) / mg.CB;
    int last = (ntile_y - 1) * mg.CB + base + mg.OFF + mg.NUg * 8 - 1;
    int need = std::max(last + 1, mg.OFF + g.nw);
    mg.Ly = (need + 7) / 8 * 8;
    return mg;
}

size_t mix_plane_halves(const Geom &g, const MixGeom &mg)
{
    return (size_t)4 * mg.NPL * g.nw * mg.Ly;
}

// A fragments [sg][jj][ks][hl][lane][8] (fp16) and the power-of-two scale sA
std::vector<_Float16> mix_tile_fragments(const Geom &g, const MixGeom &mg, const double *tiles, double &sA)
{
    double mx = 0.0;
    size_t nt = (size_t)g.maxb * g.T * g.T;
    for (size_t i = 0; i < nt; i++) mx = std::max(mx, std::fabs(tiles[i]));
    sA = 1.0;
    if (mx > 0.0) {
        int ex;
        std::frexp(mx, &ex);
        sA = std::ldexp(1.0, 13 - ex);
    }
    // the T-folded contraction keeps lo parts unscaled (one accumulator for the three products)
    const double loscale = mix_tfold(g, mg) ? 1.0 : (double)kLoScale;
    std::vector<_Float16> F((size_t)mg.nsets * g.os * mg.KS * 2 * 64 * 8);
    const int rows = mg.CB * g.os, T = g.T, w = g.w, os = g.os;
    for (int sg = 0; sg < mg.nsets; sg++)
        for (int jj = 0; jj < os; jj++)
            for (int ks = 0; ks < mg.KS; ks++)
                for (int lane = 0; lane < 64; lane++)
                    for (int j = 0; j < 8; j++) {
                        int m = lane & 15, kg = ks * 4 + (lane >> 4);
                        double v = 0.0;
                        if (m < rows && kg < mg.KG) {
                            int g8 = kg % mg.NUg, r = kg / mg.NUg, ol = r % mg.nper, dy = r / mg.nper;
                            int dx = g8 * 8 + j, c = m / os, ii = m % os;
                            if (dx < mg.NU && dx - c >= 0 && dx - c <= 2 * w) {
                                int o = sg * mg.nper + ol;
                                int x = ii + (2 * w - dx + c) * os, y = jj + (2 * w - dy) * os;
                                v = tiles[((size_t)o * T + x) * T + y] * sA;
                            }
                        }
                        _Float16 hi = (_Float16)v;
                        _Float16 lo = (_Float16)((v - (double)hi) * loscale);
                        size_t base = ((((size_t)sg * os + jj) * mg.KS + ks) * 2) * 64 * 8;
                        F[base + (size_t)lane * 8 + j] = hi;
                        F[base + 64 * 8 + (size_t)lane * 8 + j] = lo;
                    }
    return F;
}

// shapes of the T-folded contraction (k_cmix_tf): one orbital block, os 4
// (w=2: K steps 2/4/7 for 1/3/5 orbitals) or os 8 (w=1, 1 orbital)
// (os, w, bands, waves per cell row)
#define CQPI_TF_SHAPES(X) X(4, 2, 3, 2) X(4, 2, 1, 1) X(4, 2, 5, 1) X(8, 1, 1, 1)

bool mix_tfold(const Geom &g, const MixGeom &mg)
{
#define CQPI_TF_IS(os, ww, bt, js) if (tf_shape<os, ww, bt, js>(g, mg)) return true;
    CQPI_TF_SHAPES(CQPI_TF_IS)
#undef CQPI_TF_IS
    return false;
}

int launch_mix_prepare(const Geom &g, const MixGeom &mg, const double2 *G0w, int E, double *scale,
                       _Float16 *P, hipStream_t s, bool tfold)
{
    size_t per_e = (size_t)g.nw * g.nw * g.B * g.B;
    // workgroups per energy: one per 2048 values, at most 256
    const int nblk = (int)std::max<size_t>(1, std::min<size_t>(256, per_e / 2048));
    hipLaunchKernelGGL(k_wscale_clear, dim3((E + 255) / 256), dim3(256), 0, s, scale, E);
    hipLaunchKernelGGL(k_wscale, dim3(nblk, E), dim3(256), 0, s, G0w, per_e, scale);
    hipLaunchKernelGGL(k_wscale_fin, dim3((E + 255) / 256), dim3(256), 0, s, scale, E);
    if (tfold)
        hipLaunchKernelGGL(k_pack32, dim3(g.nw, E), dim3(256), 0, s, g, G0w, reinterpret_cast<float2 *>(P));
    else {
        const size_t row = sizeof(double2) * 2 * g.nw;  // LDS of one staged pair
        if (row > 160 * 1024) return CQPI_E_ARG;
        const int kp = (int)std::max<size_t>(1, std::min<size_t>(kPackPairs, kPackLds / row));
        hipLaunchKernelGGL(k_pack_rows, dim3(g.nw, (g.B * g.B + kp - 1) / kp, E), dim3(256), row * kp, s, g, mg,
                           G0w, scale, P, kp);
    }
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_contract_mixed(const Geom &g, const MixGeom &mg, const _Float16 *P, const void *Afrag,
                          const double *scale, double sA, const double2 *T, const double *bulk, int E,
                          double *ldos, hipStream_t s, const double2 *tips, bool tfold, double *cscale)
{
    const half8 *A = (const half8 *)Afrag;
    if (tips && 2 * g.maxb > g.B) return CQPI_E_ARG;
    if (tfold) {
        // P holds the FP32 window W32 (see k_cmix_tf)
        if (tips) return CQPI_E_ARG;
        const float2 *W32 = reinterpret_cast<const float2 *>(P);
#define CQPI_TF_RUN(os, ww, bt, js) \
        if (try_tf<os, ww, bt, js>(g, mg, W32, A, scale, sA, T, bulk, E, ldos, s)) \
            return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
        CQPI_TF_SHAPES(CQPI_TF_RUN)
#undef CQPI_TF_RUN
        return CQPI_E_ARG;
    }
    // register-resident tile fragments for the common small shapes:
    // os=4 (w=2: K steps 2/4/7 for 1/3/5 orbitals) and os=8 (w=1, 1 orbital).
    // 3 orbitals: two waves share a cell row (half the fragments each, three
    // waves a SIMD instead of two: 4% faster)
    // (LDOS maps only: the Josephson map takes the staged or the generic kernel)
    if (!tips && (try_areg<1, 4, 4, 4, 3, 2>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s) ||
        try_areg<1, 4, 4, 2, 1, 1>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s) ||
        try_areg<2, 4, 4, 7, 5, 1>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s) ||
        try_areg<1, 2, 8, 1, 1, 1>(g, mg, P, A, scale, sA, T, bulk, E, ldos, s)))
        return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
    // LDOS maps: T folded into the b planes first (no per-pixel T sandwich)
    if (!tips && cscale) {
        const size_t items = (size_t)g.nw * mg.Ly * g.maxb;
        hipLaunchKernelGGL(k_fold_c, dim3((unsigned)((items + 255) / 256), E), dim3(256), 0, s, g, mg, scale, T,
                           const_cast<_Float16 *>(P), cscale);
        if (hipGetLastError() != hipSuccess) return CQPI_E_HIP;
    }
    // staged B operand when the spin block's plane rows fit the LDS
    constexpr size_t kStLdsMax = 150 * 1024;
    if (g.os <= 16 && g.B <= 16) {
        switch (mg.NT) {
        case 1: if (cmix_st_lds(mg, g, 1, 4) <= kStLdsMax) return launch_st<1, 4>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, cscale); break;
        case 2: if (cmix_st_lds(mg, g, 2, 4) <= kStLdsMax) return launch_st<2, 4>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, cscale); break;
        case 3: if (cmix_st_lds(mg, g, 3, 2) <= kStLdsMax) return launch_st<3, 2>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, cscale); break;
        case 4: if (cmix_st_lds(mg, g, 4, 2) <= kStLdsMax) return launch_st<4, 2>(g, mg, P, A, scale, sA, T, bulk, E, ldos, tips, s, cscale); break;
        default: break;
        }
    }
    switch (mg.NT) {
    case 1: return launch_nt<1>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, tips, s, cscale);
    case 2: return launch_nt<2>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, tips, s, cscale);
    case 3: return launch_nt<3>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, tips, s, cscale);
    case 4: return launch_nt<4>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, tips, s, cscale);
    default: return CQPI_E_ARG;
    }
}

}  // namespace cqpi

// Mixed-precision continuum contraction on MFMA (gfx950).
//
// Same factorisation as the FP64 kernel, rho = -Im(a^T T b) - bulk, but the
// gathers a and b run as fp16 matrix products with FP32 accumulation.
// Operands are split into hi + lo fp16 parts (x = hi + lo/2^12, each part
// scaled by a power of two so that max|x| lies in [2^12, 2^13)): the products
// hi*hi + (hi*lo' + lo'*hi)/2^12 recover ~22 significant bits per product.
// The T sandwich and the bulk table stay FP64.
//
// GEMM shape (per wave): one "tile" is CB consecutive cells along y at one
// cell column cx and one sub-pixel column ii.  Its M = CB*os rows are the
// pixels (c, jj); K runs over (dx, o, dy): the lattice columns cx+dx-w, the
// orbitals o of the current spin block and the CB+2w lattice rows the tile
// touches, in groups of 8 consecutive rows; N runs over (p, re/im) of
// G0(pos+R)[o][p] (the a gather) and of G0(-pos-R)[p][o] (the b gather).
// The tile matrix A[(c,jj)][(dx,o,dy)] = w_o(ii+(2w-dx)os, jj+(2w-dy+c)os)
// (zero outside the band 0 <= dy-c <= 2w) is the same for every tile and is
// precomputed on the host in MFMA fragment order.  G0 is repacked per energy
// into "planes" (one per (o, p, re/im)) along y, so a lane's 8 K values of a
// B fragment are 8 consecutive halves in memory.
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

// per-energy power-of-two scale: max|G0 window| * s in [2^12, 2^13)
__global__ void __launch_bounds__(1024) k_wscale(const double2 *__restrict__ G0w, size_t per_e,
                                                double *__restrict__ scale)
{
    __shared__ double red[1024];
    const double2 *G = G0w + (size_t)blockIdx.x * per_e;
    double m = 0.0;
    for (size_t i = threadIdx.x; i < per_e; i += blockDim.x) {
        double2 v = G[i];
        m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double mx = red[0];
        double sc = 1.0;
        if (mx > 0.0) {
            int ex;
            frexp(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
            sc = ldexp(1.0, 13 - ex);
        }
        scale[blockIdx.x] = sc;
    }
}

// planes P[e][ab][hl][(o*B+p)*2+ri][sx][Ly]: ab=0 holds G0(R)[o][p], ab=1 holds
// G0(-R)[p][o], R = (sx-h, idx-OFF-h); entries outside the window are zero
__global__ void k_pack(Geom g, MixGeom mg, const double2 *__restrict__ G0w,
                       const double *__restrict__ scale, _Float16 *__restrict__ P)
{
    const int sx = blockIdx.x, e = blockIdx.z;
    const int B = g.B, NP = B * B, nw = g.nw;
    const int ab = blockIdx.y / mg.NPL, pl = blockIdx.y % mg.NPL;
    const int ri = pl & 1, op = pl >> 1, o = op / B, p = op % B;
    const double s = scale[e];
    const double2 *G = G0w + (size_t)e * nw * nw * NP;
    const size_t plane = (size_t)mg.Ly * nw;
    _Float16 *hi = P + ((((size_t)e * 2 + ab) * 2 + 0) * mg.NPL + pl) * plane + (size_t)sx * mg.Ly;
    _Float16 *lo = P + ((((size_t)e * 2 + ab) * 2 + 1) * mg.NPL + pl) * plane + (size_t)sx * mg.Ly;
    for (int idx = threadIdx.x; idx < mg.Ly; idx += blockDim.x) {
        int sy = idx - mg.OFF;
        double v = 0.0;
        if (sy >= 0 && sy < nw) {
            double2 z = ab == 0 ? G[((size_t)sx * nw + sy) * NP + o * B + p]
                                : G[((size_t)(2 * g.h - sx) * nw + (2 * g.h - sy)) * NP + p * B + o];
            v = (ri ? z.y : z.x) * s;
        }
        _Float16 h = (_Float16)v;
        hi[idx] = h;
        lo[idx] = (_Float16)((v - (double)h) * kLoScale);
    }
}

template <int LDW>
__device__ inline half8 load8(const _Float16 *p)
{
    half8 r;
    if constexpr (LDW == 8) {
        uint4 u = *(const uint4 *)p;
        __builtin_memcpy(&r, &u, 16);
    } else if constexpr (LDW == 4) {
        uint2 u0 = *(const uint2 *)p, u1 = *(const uint2 *)(p + 4);
        uint4 u = make_uint4(u0.x, u0.y, u1.x, u1.y);
        __builtin_memcpy(&r, &u, 16);
    } else if constexpr (LDW == 2) {
        const unsigned *q = (const unsigned *)p;
        uint4 u = make_uint4(q[0], q[1], q[2], q[3]);
        __builtin_memcpy(&r, &u, 16);
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = p[j];
    }
    return r;
}

constexpr int kWaves = 4;
constexpr int kIIB = 4;  // sub-pixel columns per accumulator block

template <int NT, int LDW>
__global__ void __launch_bounds__(64 * kWaves) k_cmix(Geom g, MixGeom mg, const _Float16 *__restrict__ P,
                                                    const half8 *__restrict__ Afrag,
                                                    const double *__restrict__ scale, double sA,
                                                    const double2 *__restrict__ Tm,
                                                    const double *__restrict__ bulk,
                                                    double *__restrict__ ldos)
{
    __shared__ double2 Ts[16 * 16];
    __shared__ float Cs[kWaves][16][16 * NT + 1];
    __shared__ double vals[kWaves][16][16];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int e = blockIdx.z;
    const int B = g.B, os = g.os, w = g.w, n = g.n, h = g.h, nw = g.nw;
    const int n2 = n >> 1;
    const int cx_raw = blockIdx.x * kWaves + wave;
    const bool cx_ok = cx_raw < n;
    const int cx = cx_ok ? cx_raw : n - 1;
    const int cy0 = blockIdx.y * mg.CB;
    const double2 *Te = Tm + (size_t)e * B * B;
    for (int i = threadIdx.x; i < B * B; i += blockDim.x) Ts[i] = Te[i];
    for (int i = lane; i < 16 * 16; i += 64) vals[wave][i / 16][i % 16] = 0.0;
    __syncthreads();

    const size_t plane = (size_t)mg.Ly * nw;
    const _Float16 *Pe = P + (size_t)e * 4 * mg.NPL * plane;
    const int col = lane & 15, kq = lane >> 4;
    const int sy0 = cy0 - n2 - w + h + mg.OFF;

    for (int sg = 0; sg < mg.nsets; sg++) {
        for (int ii0 = 0; ii0 < os; ii0 += kIIB) {
            float4v hh[kIIB][NT], cr[kIIB][NT];
#pragma unroll
            for (int i = 0; i < kIIB; i++)
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    hh[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                    cr[i][t] = float4v{0.f, 0.f, 0.f, 0.f};
                }
            for (int ks = 0; ks < mg.KS; ks++) {
                const int kg = ks * 4 + kq;
                half8 bh[NT], bl[NT];
                const bool kok = kg < mg.KG;
                int dx = 0, ol = 0, g8 = 0;
                if (kok) {
                    g8 = kg % mg.NUg;
                    int r = kg / mg.NUg;
                    ol = r % mg.nper;
                    dx = r / mg.nper;
                }
                const int o = sg * mg.nper + ol;
                const int sx = cx - n2 + dx - w + h;
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
                    if (kok && ab >= 0) {
                        const size_t off = (size_t)sx * mg.Ly + sy0 + g8 * 8;
                        bh[t] = load8<LDW>(Pe + ((size_t)(ab * 2 + 0) * mg.NPL + pl) * plane + off);
                        bl[t] = load8<LDW>(Pe + ((size_t)(ab * 2 + 1) * mg.NPL + pl) * plane + off);
                    } else {
                        bh[t] = half8{0, 0, 0, 0, 0, 0, 0, 0};
                        bl[t] = bh[t];
                    }
                }
#pragma unroll
                for (int i = 0; i < kIIB; i++) {
                    const int ii = ii0 + i;
                    if (ii >= os) break;
                    const size_t fa = ((((size_t)sg * os + ii) * mg.KS + ks) * 2) * 64;
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
            // epilogue: Im sum_pq a_p T_pq b_q per pixel row, FP64
#pragma unroll
            for (int i = 0; i < kIIB; i++) {
                const int ii = ii0 + i;
                if (ii >= os) break;
#pragma unroll
                for (int t = 0; t < NT; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        Cs[wave][kq * 4 + r][t * 16 + col] = hh[i][t][r] + cr[i][t][r] * (1.0f / kLoScale);
                __syncthreads();
                const int m = lane >> 2, part = lane & 3;
                double acc = 0.0;
                for (int p = part; p < B; p += 4) {
                    double2 tb = make_double2(0.0, 0.0);
                    for (int q = 0; q < B; q++) {
                        double2 bq = make_double2(Cs[wave][m][2 * B + 2 * q], Cs[wave][m][2 * B + 2 * q + 1]);
                        tb = cadd(tb, cmul(Ts[p * B + q], bq));
                    }
                    double ar = Cs[wave][m][2 * p], ai = Cs[wave][m][2 * p + 1];
                    acc += ar * tb.y + ai * tb.x;
                }
                acc += __shfl_xor(acc, 1);
                acc += __shfl_xor(acc, 2);
                if (part == 0) vals[wave][ii][m] += acc;
                __syncthreads();
            }
        }
    }
    const double sc = sA * scale[e];
    const double unscale = 1.0 / (sc * sc);
    const int W = g.W, rows = mg.CB * os;
    double *out = ldos + (size_t)e * W * W;
    for (int idx = lane; idx < os * 16; idx += 64) {
        const int ii = idx >> 4, m = idx & 15;
        if (m >= rows) continue;
        const int c = m / os, jj = m - c * os, cy = cy0 + c;
        if (!cx_ok || cy >= n) continue;
        const int X = cx * os + ii, Y = cy * os + jj;
        out[(size_t)X + (size_t)Y * W] =
            -(vals[wave][ii][m] * unscale + bulk[(size_t)e * os * os + ii * os + jj]);
    }
}

template <int NT>
int launch_nt(const Geom &g, const MixGeom &mg, const _Float16 *P, const void *Afrag, const double *scale,
              double sA, const double2 *T, const double *bulk, int E, double *ldos, hipStream_t s)
{
    dim3 grid((g.n + kWaves - 1) / kWaves, (g.n + mg.CB - 1) / mg.CB, E);
    dim3 blk(64 * kWaves);
    const half8 *A = (const half8 *)Afrag;
    switch (mg.LDW) {
    case 8: hipLaunchKernelGGL((k_cmix<NT, 8>), grid, blk, 0, s, g, mg, P, A, scale, sA, T, bulk, ldos); break;
    case 4: hipLaunchKernelGGL((k_cmix<NT, 4>), grid, blk, 0, s, g, mg, P, A, scale, sA, T, bulk, ldos); break;
    case 2: hipLaunchKernelGGL((k_cmix<NT, 2>), grid, blk, 0, s, g, mg, P, A, scale, sA, T, bulk, ldos); break;
    default: hipLaunchKernelGGL((k_cmix<NT, 1>), grid, blk, 0, s, g, mg, P, A, scale, sA, T, bulk, ldos); break;
    }
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

// A fragments [sg][ii][ks][hl][lane][8] (fp16) and the power-of-two scale sA
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
    std::vector<_Float16> F((size_t)mg.nsets * g.os * mg.KS * 2 * 64 * 8);
    const int rows = mg.CB * g.os, T = g.T, w = g.w, os = g.os;
    for (int sg = 0; sg < mg.nsets; sg++)
        for (int ii = 0; ii < os; ii++)
            for (int ks = 0; ks < mg.KS; ks++)
                for (int lane = 0; lane < 64; lane++)
                    for (int j = 0; j < 8; j++) {
                        int m = lane & 15, kg = ks * 4 + (lane >> 4);
                        double v = 0.0;
                        if (m < rows && kg < mg.KG) {
                            int g8 = kg % mg.NUg, r = kg / mg.NUg, ol = r % mg.nper, dx = r / mg.nper;
                            int dy = g8 * 8 + j, c = m / os, jj = m % os;
                            if (dy < mg.NU && dy - c >= 0 && dy - c <= 2 * w) {
                                int o = sg * mg.nper + ol;
                                int x = ii + (2 * w - dx) * os, y = jj + (2 * w - dy + c) * os;
                                v = tiles[((size_t)o * T + x) * T + y] * sA;
                            }
                        }
                        _Float16 hi = (_Float16)v;
                        _Float16 lo = (_Float16)((v - (double)hi) * kLoScale);
                        size_t base = ((((size_t)sg * os + ii) * mg.KS + ks) * 2) * 64 * 8;
                        F[base + (size_t)lane * 8 + j] = hi;
                        F[base + 64 * 8 + (size_t)lane * 8 + j] = lo;
                    }
    return F;
}

int launch_mix_prepare(const Geom &g, const MixGeom &mg, const double2 *G0w, int E, double *scale,
                       _Float16 *P, hipStream_t s)
{
    size_t per_e = (size_t)g.nw * g.nw * g.B * g.B;
    hipLaunchKernelGGL(k_wscale, dim3(E), dim3(1024), 0, s, G0w, per_e, scale);
    hipLaunchKernelGGL(k_pack, dim3(g.nw, 2 * mg.NPL, E), dim3(128), 0, s, g, mg, G0w, scale, P);
    return hipGetLastError() == hipSuccess ? 0 : CQPI_E_HIP;
}

int launch_contract_mixed(const Geom &g, const MixGeom &mg, const _Float16 *P, const void *Afrag,
                          const double *scale, double sA, const double2 *T, const double *bulk, int E,
                          double *ldos, hipStream_t s)
{
    switch (mg.NT) {
    case 1: return launch_nt<1>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, s);
    case 2: return launch_nt<2>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, s);
    case 3: return launch_nt<3>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, s);
    case 4: return launch_nt<4>(g, mg, P, Afrag, scale, sA, T, bulk, E, ldos, s);
    default: return CQPI_E_ARG;
    }
}

}  // namespace cqpi

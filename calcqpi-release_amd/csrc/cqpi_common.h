// Shared device/host helpers of the MI355X continuum-LDOS library.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/cqpi_amd.h"

// complex FP64 as double2 (x = re, y = im)
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ inline double2 cmulc(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
// 1 / z
__host__ __device__ inline double2 cinv(double2 z)
{
    double d = 1.0 / (z.x * z.x + z.y * z.y);
    return make_double2(z.x * d, -z.y * d);
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) { return cmul(a, cinv(b)); }

// complex FP32 as float2 (the reduced-precision transform)
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline float2 cscale(float s, float2 a) { return make_float2(s * a.x, s * a.y); }

// Hopping terms grouped by lattice vector R, 'anti' terms (phase
// exp(-2 pi i R.k)) folded onto -R, each group's amplitudes summed into one
// dense B x B block.  For integral R (kept mod K in irx, iry) the phase
// exp(+2 pi i R.k) at grid point (i, j) is the twiddle exp(+2 pi i m/K),
// m = (irx i + iry j) mod K.
struct DevRGroup {
    double rx, ry;
    int irx, iry;
    int integral;
    int pad;
};

// Geometry shared by the transform, T-matrix and contraction kernels.
struct Geom {
    int K;        // k-grid side
    int n;        // cells per side
    int w;        // neighbour window
    int os;       // oversampling
    int B;        // bands
    int maxb;     // bands carrying a Wannier function
    int spin;     // same-spin contraction blocks
    int h;        // window half width: G0(R) stored for R in [-h, h]^2
    int nw;       // 2h+1
    int T;        // tile side (2w+1)*os
    int W;        // map side os*n
    // launch tuning (0: built-in heuristic), chosen by the autotuner
    int tpw;      // mixed contraction: cell rows per wave (2, 4, 8)
    int eb;       // first transform pass: energies per workgroup
    int js;       // T-folded mixed contraction: waves per cell row (1, 2; 0: per shape)
    int streams;  // eager runs: streams an energy chunk is split over (1, 2; 0: 2 from 8 energies)
    // transform passes of the Stockham path: 1 where H(k)^T = H(-k) (real
    // hoppings), so that G0(k)^T = G0(-k) and G0(R)_cr = G0(-R)_rc; the passes
    // then transform only the pairs r <= c and pass 2 writes each of them
    // also as (c, r) at -R.  0: every pair is transformed.
    int ts;
};

// pairs (lines) that the transform passes carry per (energy, row)
__host__ __device__ inline int transform_pairs(const Geom &g)
{
    return g.ts ? g.B * (g.B + 1) / 2 : g.B * g.B;
}

// index of the pair (r, c), r <= c, in the row-major list of such pairs
__host__ __device__ constexpr int upper_index(int r, int c, int B)
{
    return r * B - r * (r - 1) / 2 + (c - r);
}

// orbitals (r, c) of transformed pair q: q = r B + c, or with ts the q-th
// pair r <= c in row-major order
__device__ inline void transform_pair_rc(int q, int B, int ts, int &r, int &c)
{
    if (!ts) {
        r = q / B;
        c = q - r * B;
        return;
    }
    int row = 0, len = B;
    while (q >= len) {
        q -= len;
        row++;
        len--;
    }
    r = row;
    c = row + q;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads also waits
// for the workgroup's outstanding global stores; where a barrier only guards
// the reuse of an LDS buffer this one lets those stores drain in the
// background.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// t / d by a multiply-high with m = udiv_magic(d) = ceil(2^32 / d).  With
// m d = 2^32 + r, 0 <= r < d: t m / 2^32 = t/d + t r / (d 2^32), and the
// second term stays below 1/d (so the floor is exact) for 0 <= t < 2^32 / d.
__host__ __device__ inline unsigned long long udiv_magic(unsigned d)
{
    return ((1ull << 32) + d - 1) / d;
}
__device__ inline int udiv(int t, unsigned long long m)
{
    return (int)(((unsigned long long)(unsigned)t * m) >> 32);
}

// Workgroup coordinates of a 3-D grid renumbered for the 8 XCDs of the chip.
// The dispatcher hands consecutive workgroup ids to the XCDs round-robin, so
// id L runs on XCD L % 8.  The renumbering v = (ids before XCD L % 8) + L / 8
// is a bijection of [0, total) that gives each XCD a contiguous range of v;
// decoding v with z slowest keeps the workgroups of one z slice (one energy)
// on few XCDs, so that slice's data stays in those XCDs' L2.
__device__ inline void xcd_block(int &bx, int &by, int &bz)
{
    const unsigned nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
    const unsigned total = nx * ny * nz;
    const unsigned L = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const unsigned x = L % 8u, q = total / 8u, r = total % 8u;
    const unsigned v = x * q + (x < r ? x : r) + L / 8u;
    const unsigned plane = nx * ny;
    bz = (int)(v / plane);
    const unsigned rem = v - (unsigned)bz * plane;
    by = (int)(rem / nx);
    bx = (int)(rem - (unsigned)by * nx);
}

// Window rows of one orbital o for a square block of cells, staged in LDS by
// the whole workgroup.  Slot coordinates are window slots (R + h).  For a
// block whose cells need G0(pos+R) at slots (ax0 + lx, ay0 + ly) and
// G0(-pos-R) at slots (bx0 + lx, by0 + ly), 0 <= lx, ly < SA:
//   GAs[(lx SA + ly) B + q] = G0(slot a)_{o,q},  GBs[...] = G0(slot b)_{q,ob},
// zero where a slot falls outside the nw x nw window (ob = o for the LDOS,
// o + maxband for the particle-hole column of the Josephson map).  G0 points
// at one energy's window, [slot x][slot y][pair].
template <typename C, int B>
__device__ inline void stage_window_rows(const C *__restrict__ G0, int nw, int o, int ob, int SA, int ax0, int ay0,
                                         int bx0, int by0, C *GAs, C *GBs)
{
    for (int t = threadIdx.x; t < SA * SA * B; t += blockDim.x) {
        const int q = t % B, sl = t / B, ly = sl % SA, lx = sl / SA;
        const int axs = ax0 + lx, ays = ay0 + ly, bxs = bx0 + lx, bys = by0 + ly;
        C va = C{0, 0}, vb = C{0, 0};
        if (axs >= 0 && axs < nw && ays >= 0 && ays < nw) va = G0[((size_t)axs * nw + ays) * (B * B) + o * B + q];
        if (bxs >= 0 && bxs < nw && bys >= 0 && bys < nw) vb = G0[((size_t)bxs * nw + bys) * (B * B) + q * B + ob];
        GAs[t] = va;
        GBs[t] = vb;
    }
}

struct MixGeom;

// launchers (defined next to their kernels)
namespace cqpi {

int launch_eigen(int B, const DevRGroup *groups, int ngroups, const double2 *tmat, const double2 *tw,
                 double efermi, int scmax, int K, double *eval, double2 *evec, hipStream_t s);
// band-structure histogram hist[(x + y pix) + layer pix^2] (zeroed by the caller)
int launch_bands_hist(int B, const DevRGroup *groups, int ngroups, const double2 *tmat, double efermi, int scmax,
                      int pix, int os, double le, double ue, int layers, double *hist, hipStream_t s);

// resolvent G0(k,w) from eigen data fused with the transform along ky; writes
// F1 (the Ry window of every kx row).  omegas: device array of the chunk.
// work: workspace of resolvent_work_bytes(g, E, ...) bytes (G(k) of a block
// of energies, for band counts or grids beyond the register path; may be
// NULL when that is 0).
int launch_resolvent_ky(const Geom &g, const double *eval, const double2 *evec,
                        const double *omegas, int E, double eta, const double2 *tw,
                        double2 *F1, void *work, size_t work_bytes, hipStream_t s);
size_t resolvent_work_bytes(const Geom &g, int E, bool fp32);
// the ky pass from G(k) of E energies already in the workspace layout
// Gk[e][pair][kx][ky] (double2, or float2 with fp32; power-of-two K in [64, 1024])
int launch_resolvent_ky_gk(const Geom &g, const void *Gk, const double *omegas, int E, const void *tw, void *F1,
                           bool fp32, hipStream_t s);
// surface (surface != 0) or bulk Green's function of a layered model by
// decimation into that layout; groups g/t/n: bulk in-plane, surface
// in-plane, coupling (rz < 0).  stat[e]: the most decimation steps any k
// needed (maxit + 1: not converged).  B <= 8.
int launch_surface_gk(int B, int K, const DevRGroup *const g[3], const double2 *const t[3], const int n[3],
                      const double2 *tw, double efermi, int scmax, const double *omegas, int E, double eta,
                      double epserr, int maxit, int surface, void *Gk, bool fp32, int *stat, hipStream_t s);
// transform along kx, window extraction and 1/K^2: G0w[e][sx][sy][pair]; with
// mg != NULL also the fp16 hi/lo planes of the mixed contraction (scale sG)
int launch_transform_kx(const Geom &g, const MixGeom *mg, const double2 *F1, int E, const double2 *tw,
                        double sG, double2 *G0w, _Float16 *P, hipStream_t s);
// the same two passes with FP32 lines, F1 and twiddles (power-of-two K in
// [64, 1024]); the resolvent factor is formed in FP64, the window written in FP64
bool transform32_supported(const Geom &g);
int launch_resolvent_ky32(const Geom &g, const double *eval, const double2 *evec, const double *omegas, int E,
                          double eta, const float2 *tw, float2 *F1, void *work, size_t work_bytes, hipStream_t s);
// rwin: half-width of the part of the window written to G0w (g.h: all of it;
// the mixed contraction needs only |R| <= 2w, for the T-matrix and the bulk term)
// W32: write the whole window in FP32 as W32[e][sy][sx][pair] instead of the
// fp16 planes (the T-folded contraction's input)
int launch_transform_kx32(const Geom &g, const MixGeom *mg, const float2 *F1, int E, const float2 *tw, double sG,
                          double2 *G0w, _Float16 *P, int rwin, hipStream_t s, float2 *W32 = nullptr);
// T[e] = (1 - V G0w[e](0))^{-1} V; status[e] = 1 if singular
int launch_tmatrix(const Geom &g, const double2 *G0w, const double2 *V, int E, double2 *T,
                   int *status, hipStream_t s);
// periodic (bulk) term per sub-pixel: bulk[e][ii*os+jj]
int launch_bulk(const Geom &g, const double2 *G0w, const double *tiles, int E, double *bulk,
                hipStream_t s);
// the Josephson map's complex periodic table: bulk[e][ii*os+jj] =
// sum w_o1 w_o2 G0(R2-R1)_{o2, o1+maxband}
int launch_bulk_josephson(const Geom &g, const double2 *G0w, const double *tiles, int E, double2 *bulk,
                          hipStream_t s);
// launch_tmatrix and launch_bulk (or launch_bulk_josephson) in one launch
int launch_tmatrix_bulk(const Geom &g, const double2 *G0w, const double2 *V, int E, double2 *T, int *status,
                        const double *tiles, void *bulk, bool josephson, hipStream_t s);
// adjoint of the whole chain (see adjoint.hip)
size_t adjoint_pix_blocks(const Geom &g);
// doubles of the energy-independent weight-overlap table of the periodic term
size_t adjoint_q_size(const Geom &g);
bool adjoint_supported(const Geom &g);
int adjoint_energy_groups(const Geom &g, int E);
// bytes of the Gbar(k) workspace the ky adjoint needs for E energies (0 when
// each row's lines fit one workgroup); launch_adjoint takes it as Gk
size_t adjoint_workspace_bytes(const Geom &g, int E);
// Rl[r] = (rx, index into ryvals) of each distinct hopping vector, plus a last
// entry (0, nry) for the plain k sums; SR[r][pair] receives the sums; U holds
// K (nry + 1) b^2 + nry K values (the kx-row sums, then the y phase table)
int launch_adjoint(const Geom &g, const double2 *G0w, const double2 *T, const double2 *V, const double *tiles,
                   const double *rhobar, const double *eval, const double2 *evec, const double *omegas, int E,
                   double eta, const double2 *twn, const double *ryvals, int nry, const double2 *Rl, int nRl,
                   double2 *abar, double2 *bbar, double2 *Tpart, double2 *Tbar, double *rhosum, double2 *Gbar,
                   double2 *Vbar, double2 *F1bar, double2 *Hacc, double2 *U, double2 *SR, double *Q,
                   double2 *Gk, size_t gk_bytes, hipStream_t s);
// contraction: ldos[e][x + y*W].  With tips != NULL (device, one per energy)
// the Josephson map instead: particle-hole column o + maxband, bulk the
// complex periodic table of launch_bulk_josephson (needs 2 maxband <= b)
int launch_contract_fp64(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
                         const double *tiles, int E, double *ldos, hipStream_t s, const double2 *tips = nullptr);
// threshold-list contraction: coef[k][o2][o1][pix] per active window
// pair k = lpair[k] = r1*(2w+1)^2 + r2, pairs of one r1 contiguous from
// r1off[r1]; the same factors per orbital pair present, the combos
// coff[k*b + o1] .. coff[k*b + o1 + 1] - 1 with o2 = co2[cb] and factors
// ccoef[cb][pix]; writes the periodic part to bulk (E x os^2) and the maps to ldos
// (up to 16 bands; G0f, Tf != NULL: FP32, with E*nw*nw*b*b and E*b*b
// elements of scratch for the narrowed windows and T)
int launch_contract_list(const Geom &g, const double2 *G0w, const double2 *T, const double *coef,
                         const int *coff, const int *co2, const double *ccoef,
                         const int *r1off, const int *lpair, int nact, int E, double *bulk, double *ldos,
                         hipStream_t s, float2 *G0f = nullptr, float2 *Tf = nullptr);
// spectral-function maps n x n per energy (mode 1: SPF, 2: USPF; pos: mb x 3 =
// x, y, prefactor): from G0(k) blocks g0k[K*K][b*b] (one energy) or, with
// g0k == NULL, from the eigen data for the E energies omegas (device)
int launch_spectral(int K, int n, int B, int mb, int spin, int mode, const double2 *g0k, const double *eval,
                    const double2 *evec, const double *omegas, int E, double eta, const double *pos, double *out,
                    hipStream_t s);
// density of states -sum_{o<mb} Im G0(R=0)_oo per energy; partial: E * dos_chunks(K) doubles
int dos_chunks(int K);
// the SPF / USPF maps and the DOS from the G(k) workspace Gk[e][pair][k] of E
// energies (the surface Green's function); partial: E * dos_chunks(K) doubles
int launch_spectral_gk(int K, int n, int B, int mb, int spin, int mode, const double2 *Gk, int E, const double *pos,
                       double *out, hipStream_t s);
int launch_dos_gk(int K, int B, int mb, const double2 *Gk, int E, double *partial, double *dos, hipStream_t s);
int launch_dos(int K, int B, int mb, const double *eval, const double2 *evec, const double *omegas, int E,
               double eta, double *partial, double *dos, hipStream_t s);
// FP32 contraction: narrows the E windows and T into G0f (E*nw*nw*b*b) and Tf
// (E*b*b) first; tilesf are the tiles in FP32
int launch_contract_fp32(const Geom &g, const double2 *G0w, const double2 *T, const double *bulk,
                         const float *tilesf, int E, double *ldos, float2 *G0f, float2 *Tf, hipStream_t s,
                         const double2 *tips = nullptr);
}  // namespace cqpi

/*
 * oracle.c -- FP64 CPU restatement of calcQPI's continuum-LDOS (cLDOS) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, and by nothing else.  The product never
 * links, loads or calls it.
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows.  The reference itself cannot be compiled in
 * this image (it needs GSL and FFTW3, neither of which is installed), so the
 * linear algebra and the FFT are written here from the mathematics:
 *   - Hermitian eigensolver: cyclic complex Jacobi (replaces gsl_eigen_hermv)
 *   - 2-D DFT with e^{+2 pi i k.R}, normalised by 1/K^2 (replaces the FFTW
 *     FFTW_BACKWARD plan), radix-2 for power-of-two K, direct DFT otherwise
 *   - LU with partial pivoting (replaces gsl_linalg_complex_LU_decomp/invert)
 *   - small complex matrix products (replace gsl_blas_zgemm / zgerc)
 *
 * Storage conventions (all complex numbers are interleaved re,im doubles):
 *   H, evec, g0 blocks:  b x b row-major, element (r,c) at [(r*b+c)*2]
 *   evec column n is eigenvector n (gsl convention)
 *   k-grid / R-grid arrays: [i][j] with i the x index, block at ((i*K+j)*b*b)
 *   wf tiles: [o][x][y], side T=(2w+1)*os
 *   ldos: x + y*W   (src/qpi.cc:821-831)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct {
    double rx, ry, rz;
    int o1, o2;       /* 0-based */
    double tre, tim;  /* already divided by the degeneracy */
    int anti;
} or_hop;

/* ------------------------------------------------------------------------- */
/* _hr.dat reader: src/tightbinding.cc:70-149                                 */
/* ------------------------------------------------------------------------- */

static char *or_getline(FILE *f, char **buf, size_t *cap)
{
    ssize_t n = getline(buf, cap, f);
    if (n < 0) return NULL;
    /* strip trailing newline characters only */
    while (n > 0 && ((*buf)[n - 1] == '\n' || (*buf)[n - 1] == '\r')) (*buf)[--n] = 0;
    return *buf;
}

/*
 * Reads a Wannier90 _hr.dat file.  Returns 0 on success, nonzero on error.
 * hops is malloc'ed (caller frees with or_free).  name receives line 1.
 * Semantics follow src/tightbinding.cc:70-149:
 *   line 1 name, line 2 num_wann, line 3 nrpts, then nrpts degeneracies spread
 *   over as many lines as needed (:84-107), then "x y z u v Re Im" lines.
 *   A new R is counted whenever (x,y,z) differs from the previous line
 *   (:121-127) and each t is divided by the degeneracy of that R (:134-135).
 *   Orbital indices are 1-based (:136-141); anti = both orbitals >= scmaxbands
 *   (:137).
 */
int or_read_hr(const char *path, int scmodel, or_hop **hops_out, size_t *nh_out,
               size_t *bands_out, char *name, size_t namelen)
{
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    char *buf = NULL;
    size_t cap = 0;
    int err = 0;
    size_t n = 0, nrpts = 0;
    if (!or_getline(f, &buf, &cap)) { err = 2; goto done; }
    if (name && namelen) {
        strncpy(name, buf, namelen - 1);
        name[namelen - 1] = 0;
    }
    if (!or_getline(f, &buf, &cap) || sscanf(buf, "%zu", &n) != 1) { err = 3; goto done; }
    if (!or_getline(f, &buf, &cap) || sscanf(buf, "%zu", &nrpts) != 1) { err = 4; goto done; }
    size_t scmax = scmodel ? (n >> 1) : n;
    long *deg = (long *)calloc(nrpts ? nrpts : 1, sizeof(long));
    size_t got = 0;
    while (got < nrpts) {
        if (!or_getline(f, &buf, &cap)) { err = 5; break; }
        char *p = buf;
        for (;;) {
            while (*p && !isdigit((unsigned char)*p)) p++;
            if (!*p || got >= nrpts) break;
            char *end;
            deg[got++] = strtol(p, &end, 10);
            p = end;
        }
    }
    if (err) { free(deg); goto done; }
    size_t capn = 64, nh = 0;
    or_hop *hops = (or_hop *)malloc(capn * sizeof(or_hop));
    size_t k = 0;
    int ox = 0, oy = 0, oz = 0;
    while (or_getline(f, &buf, &cap)) {
        if (!buf[0]) continue;
        int x, y, z;
        long u, v;
        double hr, hc;
        int cnt = sscanf(buf, "%d %d %d %ld %ld %lf %lf", &x, &y, &z, &u, &v, &hr, &hc);
        if (cnt != 7) { err = 6; break; }
        if (k) {
            if (x != ox || y != oy || z != oz) { k++; ox = x; oy = y; oz = z; }
        } else { ox = x; oy = y; oz = z; k++; }
        if (k > nrpts) { err = 7; break; }
        if (u <= 0 || v <= 0 || (size_t)u > n || (size_t)v > n) { err = 8; break; }
        hr /= (double)deg[k - 1];
        hc /= (double)deg[k - 1];
        if (nh == capn) { capn *= 2; hops = (or_hop *)realloc(hops, capn * sizeof(or_hop)); }
        or_hop h;
        h.rx = x; h.ry = y; h.rz = z;
        h.o1 = (int)(u - 1); h.o2 = (int)(v - 1);
        h.tre = hr; h.tim = hc;
        h.anti = ((size_t)(u - 1) >= scmax) && ((size_t)(v - 1) >= scmax);
        hops[nh++] = h;
    }
    free(deg);
    if (err) { free(hops); goto done; }
    *hops_out = hops;
    *nh_out = nh;
    *bands_out = n;
done:
    free(buf);
    fclose(f);
    return err;
}

void or_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* H(k): src/tightbinding.cc:210-219 with hoppingterm::addhop                 */
/*       src/tightbinding.h:128-135                                           */
/* ------------------------------------------------------------------------- */
void or_setmatrix(const or_hop *hops, size_t nh, size_t b, double efermi, size_t scmax,
                  double kx, double ky, double *Hout)
{
    cplx *H = (cplx *)Hout;
    for (size_t i = 0; i < b * b; i++) H[i] = 0;
    for (size_t i = 0; i < b; i++) H[i * b + i] = (i < scmax) ? -efermi : efermi;
    const double kz = 0.0;
    for (size_t m = 0; m < nh; m++) {
        const or_hop *h = &hops[m];
        double arg = 2.0 * M_PI * (h->rx * kx + h->ry * ky + h->rz * kz);
        if (h->anti) arg = -arg;
        cplx t = h->tre + I * h->tim;
        H[h->o1 * b + h->o2] += t * (cos(arg) + I * sin(arg));
    }
}

/* ------------------------------------------------------------------------- */
/* Hermitian eigensolver (cyclic complex Jacobi).  Replaces gsl_eigen_hermv   */
/* (src/qpi.cc:470).  On return eval[n] and column n of evec (row-major b*b). */
/* ------------------------------------------------------------------------- */
void or_eigh(size_t b, const double *Hin, double *eval, double *evec_out)
{
    cplx *A = (cplx *)malloc(b * b * sizeof(cplx));
    cplx *V = (cplx *)evec_out;
    memcpy(A, Hin, b * b * sizeof(cplx));
    for (size_t i = 0; i < b; i++)
        for (size_t j = 0; j < b; j++) V[i * b + j] = (i == j) ? 1.0 : 0.0;
    for (size_t i = 0; i < b; i++) A[i * b + i] = creal(A[i * b + i]);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, diag = 0.0;
        for (size_t p = 0; p < b; p++) {
            diag += creal(A[p * b + p]) * creal(A[p * b + p]);
            for (size_t q = p + 1; q < b; q++) {
                double m = cabs(A[p * b + q]);
                off += m * m;
            }
        }
        if (off == 0.0 || off < 1e-34 * diag) break;
        for (size_t p = 0; p + 1 < b; p++) {
            for (size_t q = p + 1; q < b; q++) {
                cplx apq = A[p * b + q];
                double r = cabs(apq);
                if (r == 0.0) continue;
                double app = creal(A[p * b + p]), aqq = creal(A[q * b + q]);
                cplx ph = apq / r; /* e^{i phi} */
                double theta = (aqq - app) / (2.0 * r);
                double t;
                if (fabs(theta) > 1e150) t = 0.5 / theta;
                else t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                cplx phc = conj(ph);
                /* columns: A <- A U with U = [[c, s],[-s e^{-i phi}, c e^{-i phi}]] */
                for (size_t k = 0; k < b; k++) {
                    cplx akp = A[k * b + p], akq = A[k * b + q];
                    A[k * b + p] = c * akp - s * phc * akq;
                    A[k * b + q] = s * akp + c * phc * akq;
                }
                /* rows: A <- U^H A */
                for (size_t k = 0; k < b; k++) {
                    cplx apk = A[p * b + k], aqk = A[q * b + k];
                    A[p * b + k] = c * apk - s * ph * aqk;
                    A[q * b + k] = s * apk + c * ph * aqk;
                }
                A[p * b + q] = 0;
                A[q * b + p] = 0;
                A[p * b + p] = app - t * r;
                A[q * b + q] = aqq + t * r;
                for (size_t k = 0; k < b; k++) {
                    cplx vkp = V[k * b + p], vkq = V[k * b + q];
                    V[k * b + p] = c * vkp - s * phc * vkq;
                    V[k * b + q] = s * vkp + c * phc * vkq;
                }
            }
        }
    }
    for (size_t i = 0; i < b; i++) eval[i] = creal(A[i * b + i]);
    free(A);
}

/* k-grid: src/qpi.cc:456-471 */
static double or_kcoord(size_t i, size_t K)
{
    size_t cut = K >> 1;
    return (i > cut) ? ((double)i - (double)K) / (double)K : (double)i / (double)K;
}

/* tmatrix::calchamiltonian, src/qpi.cc:454-475 (cached eigen data) */
void or_calchamiltonian(const or_hop *hops, size_t nh, size_t b, double efermi, size_t scmax,
                        size_t K, double *eval, double *evec)
{
#pragma omp parallel
    {
        double *H = (double *)malloc(b * b * 2 * sizeof(double));
#pragma omp for schedule(static)
        for (size_t i = 0; i < K; i++)
            for (size_t j = 0; j < K; j++) {
                or_setmatrix(hops, nh, b, efermi, scmax, or_kcoord(i, K), or_kcoord(j, K), H);
                size_t idx = i * K + j;
                or_eigh(b, H, eval + idx * b, evec + idx * b * b * 2);
            }
        free(H);
    }
}

/* G0(k,w) = sum_n v_n v_n^H / (w - e_n + i eta): src/tightbinding.cc:282-289,
 * looped over the k-grid as in tmatrix::setgreensfunction, src/qpi.cc:478-486 */
void or_setgreensfunction(size_t K, size_t b, const double *eval, const double *evec_in,
                          double omega, double eta, double *g0_out)
{
    const cplx *evec = (const cplx *)evec_in;
    cplx *g0 = (cplx *)g0_out;
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < K * K; idx++) {
        const cplx *V = evec + idx * b * b;
        const double *e = eval + idx * b;
        cplx *G = g0 + idx * b * b;
        for (size_t r = 0; r < b * b; r++) G[r] = 0;
        for (size_t nn = 0; nn < b; nn++) {
            cplx f = 1.0 / ((omega - e[nn]) + I * eta);
            for (size_t r = 0; r < b; r++)
                for (size_t c = 0; c < b; c++)
                    G[r * b + c] += f * V[r * b + nn] * conj(V[c * b + nn]);
        }
    }
}

/* 1-D transform of length K with stride, sign +1: out[m] = sum_j in[j] e^{+2 pi i j m/K} */
static void or_dft_line(cplx *x, size_t K, size_t stride, cplx *scratch, const cplx *tw)
{
    int pow2 = (K & (K - 1)) == 0;
    for (size_t j = 0; j < K; j++) scratch[j] = x[j * stride];
    if (pow2 && K > 1) {
        /* bit reversal */
        size_t lg = 0;
        while (((size_t)1 << lg) < K) lg++;
        for (size_t j = 0; j < K; j++) {
            size_t r = 0;
            for (size_t bit = 0; bit < lg; bit++)
                if (j & ((size_t)1 << bit)) r |= (size_t)1 << (lg - 1 - bit);
            if (r > j) { cplx t = scratch[j]; scratch[j] = scratch[r]; scratch[r] = t; }
        }
        for (size_t len = 2; len <= K; len <<= 1) {
            size_t half = len >> 1, step = K / len;
            for (size_t s0 = 0; s0 < K; s0 += len)
                for (size_t q = 0; q < half; q++) {
                    cplx w = tw[q * step];
                    cplx u = scratch[s0 + q], v = scratch[s0 + q + half] * w;
                    scratch[s0 + q] = u + v;
                    scratch[s0 + q + half] = u - v;
                }
        }
        for (size_t j = 0; j < K; j++) x[j * stride] = scratch[j];
    } else {
        for (size_t m = 0; m < K; m++) {
            cplx acc = 0;
            for (size_t j = 0; j < K; j++) acc += scratch[j] * tw[(j * m) % K];
            x[m * stride] = acc;
        }
    }
}

/* G0(R) = (1/K^2) sum_k G0(k) e^{+2 pi i k.R}, in place, result at [Rx mod K][Ry mod K]:
 * tmatrix::calcrealspacegreensfunction src/qpi.cc:557-578 with the in-place
 * FFTW_BACKWARD plan of src/qpi.h:105 */
void or_realspace(size_t K, size_t b, double *g0_io)
{
    cplx *g0 = (cplx *)g0_io;
    cplx *tw = (cplx *)malloc(K * sizeof(cplx));
    for (size_t m = 0; m < K; m++) {
        double a = 2.0 * M_PI * (double)m / (double)K;
        tw[m] = cos(a) + I * sin(a);
    }
    size_t bb = b * b;
    double norm = (double)K * (double)K;
    for (size_t pr = 0; pr < bb; pr++) {
        /* scale first as the reference does (src/qpi.cc:561-565) */
#pragma omp parallel for schedule(static)
        for (size_t idx = 0; idx < K * K; idx++) g0[idx * bb + pr] /= norm;
        /* transform along j (second index) for every i */
#pragma omp parallel
        {
            cplx *scr = (cplx *)malloc(K * sizeof(cplx));
#pragma omp for schedule(static)
            for (size_t i = 0; i < K; i++) or_dft_line(g0 + (i * K) * bb + pr, K, bb, scr, tw);
#pragma omp for schedule(static)
            for (size_t j = 0; j < K; j++) or_dft_line(g0 + j * bb + pr, K, K * bb, scr, tw);
            free(scr);
        }
    }
    free(tw);
}

/* ------------------------------------------------------------------------- */
/* T = (1 - V G0(R=0))^{-1} V : tmatrix::setscatteringmatrix src/qpi.cc:833-854 */
/* V is a full b x b complex matrix here; the reference uses V = diag(vscat)  */
/* (src/qpi.cc:840-841), which is the special case of a diagonal V.           */
/* Returns nonzero if the matrix is singular.                                 */
/* ------------------------------------------------------------------------- */
int or_tmatrix(size_t b, const double *Vin, const double *g00_in, double *Tout)
{
    const cplx *V = (const cplx *)Vin, *G = (const cplx *)g00_in;
    cplx *M = (cplx *)malloc(b * b * sizeof(cplx));
    cplx *Inv = (cplx *)malloc(b * b * sizeof(cplx));
    size_t *perm = (size_t *)malloc(b * sizeof(size_t));
    int err = 0;
    for (size_t r = 0; r < b; r++)
        for (size_t c = 0; c < b; c++) {
            cplx acc = (r == c) ? 1.0 : 0.0;
            for (size_t k = 0; k < b; k++) acc -= V[r * b + k] * G[k * b + c];
            M[r * b + c] = acc;
        }
    /* LU decomposition with partial pivoting (in place), P M = L U */
    for (size_t i = 0; i < b; i++) perm[i] = i;
    for (size_t j = 0; j < b; j++) {
        size_t piv = j;
        double best = cabs(M[j * b + j]);
        for (size_t i = j + 1; i < b; i++)
            if (cabs(M[i * b + j]) > best) { best = cabs(M[i * b + j]); piv = i; }
        if (best == 0.0) { err = 1; break; }
        if (piv != j) {
            for (size_t c = 0; c < b; c++) { cplx t = M[j * b + c]; M[j * b + c] = M[piv * b + c]; M[piv * b + c] = t; }
            size_t t = perm[j]; perm[j] = perm[piv]; perm[piv] = t;
        }
        for (size_t i = j + 1; i < b; i++) {
            cplx l = M[i * b + j] / M[j * b + j];
            M[i * b + j] = l;
            for (size_t c = j + 1; c < b; c++) M[i * b + c] -= l * M[j * b + c];
        }
    }
    if (!err) {
        /* inverse column by column: solve L U x = P e_c */
        for (size_t c = 0; c < b; c++) {
            cplx *x = (cplx *)malloc(b * sizeof(cplx));
            for (size_t i = 0; i < b; i++) x[i] = (perm[i] == c) ? 1.0 : 0.0;
            for (size_t i = 0; i < b; i++)
                for (size_t k = 0; k < i; k++) x[i] -= M[i * b + k] * x[k];
            for (size_t ii = b; ii-- > 0;) {
                for (size_t k = ii + 1; k < b; k++) x[ii] -= M[ii * b + k] * x[k];
                x[ii] /= M[ii * b + ii];
            }
            for (size_t i = 0; i < b; i++) Inv[i * b + c] = x[i];
            free(x);
        }
        cplx *T = (cplx *)Tout;
        for (size_t r = 0; r < b; r++)
            for (size_t c = 0; c < b; c++) {
                cplx acc = 0;
                for (size_t k = 0; k < b; k++) acc += Inv[r * b + k] * V[k * b + c];
                T[r * b + c] = acc;
            }
    }
    free(M); free(Inv); free(perm);
    return err;
}

/* ------------------------------------------------------------------------- */
/* Continuum contraction, reference loop nest: tmatrix::calcwannierldos        */
/* src/qpi.cc:876-950 (CPU branch, including the spin branch :924-940 and the */
/* zero-orbital skip :916-918).                                               */
/* g0 holds G0(R) on the full K x K grid.  wfzero[o] != 0 marks a zero tile.  */
/* ------------------------------------------------------------------------- */
static inline void zgemm_small(size_t b, const cplx *A, const cplx *B, cplx *C)
{
    for (size_t r = 0; r < b; r++)
        for (size_t c = 0; c < b; c++) {
            cplx acc = 0;
            for (size_t k = 0; k < b; k++) acc += A[r * b + k] * B[k * b + c];
            C[r * b + c] = acc;
        }
}

/* The same loop nest restricted to the cells i in [i0, i1), j in [j0, j1) of
 * the n x n map (the other pixels stay 0): a sample of a map too large for
 * the whole loop nest on the host. */
void or_wannierldos_cells(size_t K, size_t n, size_t window, size_t oversamp, size_t b, size_t maxband,
                          int spin, const double *Tin, const double *g0_in, const double *wf,
                          const int *wfzero, size_t i0, size_t i1, size_t j0, size_t j1, double *ldos);

void or_wannierldos(size_t K, size_t n, size_t window, size_t oversamp, size_t b, size_t maxband,
                    int spin, const double *Tin, const double *g0_in, const double *wf,
                    const int *wfzero, double *ldos)
{
    or_wannierldos_cells(K, n, window, oversamp, b, maxband, spin, Tin, g0_in, wf, wfzero, 0, n, 0, n, ldos);
}

void or_wannierldos_cells(size_t K, size_t n, size_t window, size_t oversamp, size_t b, size_t maxband,
                          int spin, const double *Tin, const double *g0_in, const double *wf,
                          const int *wfzero, size_t i0, size_t i1, size_t j0, size_t j1, double *ldos)
{
    const cplx *T = (const cplx *)Tin, *g0 = (const cplx *)g0_in;
    size_t W = oversamp * n, TS = (2 * window + 1) * oversamp, bb = b * b;
    if (!maxband) maxband = b;
    size_t n2 = n >> 1, spinbands = maxband >> 1;
    int w = (int)window;
    for (size_t i = 0; i < W * W; i++) ldos[i] = 0.0;
#define G0AT(ix, iy) (g0 + ((size_t)(ix) * K + (size_t)(iy)) * bb)
#define WF(o, x, y) wf[((size_t)(o) * TS + (size_t)(x)) * TS + (size_t)(y)]
#pragma omp parallel
    {
        cplx *buffer = (cplx *)malloc(bb * sizeof(cplx)), *cl = (cplx *)malloc(bb * sizeof(cplx));
#pragma omp for schedule(static) collapse(2)
        for (size_t i = i0; i < i1; i++)
            for (size_t j = j0; j < j1; j++) {
                long ipos = (long)i - (long)n2, jpos = (long)j - (long)n2;
                for (int nni1 = -w; nni1 <= w; nni1++)
                    for (int nnj1 = -w; nnj1 <= w; nnj1++) {
                        /* T G0(0,R) */
                        size_t ax = (size_t)((2 * (long)K - (ipos + nni1)) % (long)K);
                        size_t ay = (size_t)((2 * (long)K - (jpos + nnj1)) % (long)K);
                        zgemm_small(b, T, G0AT(ax, ay), buffer);
                        for (int nni2 = -w; nni2 <= w; nni2++)
                            for (int nnj2 = -w; nnj2 <= w; nnj2++) {
                                size_t bx = (size_t)(((long)K + ipos + nni2) % (long)K);
                                size_t by = (size_t)(((long)K + jpos + nnj2) % (long)K);
                                zgemm_small(b, G0AT(bx, by), buffer, cl);
                                size_t cx = (size_t)(((long)K - nni1 + nni2) % (long)K);
                                size_t cy = (size_t)(((long)K - nnj1 + nnj2) % (long)K);
                                const cplx *gb = G0AT(cx, cy);
                                for (size_t ii = 0; ii < oversamp; ii++)
                                    for (size_t jj = 0; jj < oversamp; jj++) {
                                        size_t x = ii + window * oversamp, y = jj + window * oversamp;
                                        double *out = &ldos[((i * oversamp + ii) % W) + ((j * oversamp + jj) % W) * W];
                                        size_t x1 = x - nni1 * oversamp, y1 = y - nnj1 * oversamp;
                                        size_t x2 = x - nni2 * oversamp, y2 = y - nnj2 * oversamp;
                                        int nsets = spin ? 2 : 1;
                                        size_t nper = spin ? spinbands : maxband;
                                        for (int sg = 0; sg < nsets; sg++) {
                                            size_t off = sg ? spinbands : 0;
                                            for (size_t o1 = off; o1 < off + nper; o1++) {
                                                if (wfzero && wfzero[o1]) continue;
                                                for (size_t o2 = off; o2 < off + nper; o2++) {
                                                    if (wfzero && wfzero[o2]) continue;
                                                    double factor = WF(o1, x1, y1) * WF(o2, x2, y2);
                                                    *out -= cimag(cl[o2 * b + o1]) * factor;
                                                    *out -= cimag(gb[o2 * b + o1]) * factor;
                                                }
                                            }
                                        }
                                    }
                            }
                    }
            }
        free(buffer);
        free(cl);
    }
#undef G0AT
#undef WF
}

/* ------------------------------------------------------------------------- */
/* Josephson map, reference loop nest: tmatrix::calcwannierjosephson          */
/* src/qpi.cc:953-1033 (CPU branch; GPU kernels src/qpihipkernel.cc:171-271).  */
/* maxband orbitals carry tiles (the reference uses bands/2); the particle-   */
/* hole column o1 + maxband is read and the sum is multiplied by tip before   */
/* the imaginary part is taken.  maxband = 0 means bands/2.                   */
/* ------------------------------------------------------------------------- */
void or_wannierjosephson(size_t K, size_t n, size_t window, size_t oversamp, size_t b, size_t maxband,
                         int spin, const double *Tin, const double *g0_in, const double *wf,
                         const int *wfzero, double tip_re, double tip_im, double *ldos)
{
    const cplx *T = (const cplx *)Tin, *g0 = (const cplx *)g0_in;
    const cplx tip = tip_re + I * tip_im;
    size_t W = oversamp * n, TS = (2 * window + 1) * oversamp, bb = b * b;
    if (!maxband) maxband = b >> 1;
    size_t n2 = n >> 1, spinbands = maxband >> 1;
    int w = (int)window;
    for (size_t i = 0; i < W * W; i++) ldos[i] = 0.0;
#define G0AT(ix, iy) (g0 + ((size_t)(ix) * K + (size_t)(iy)) * bb)
#define WF(o, x, y) wf[((size_t)(o) * TS + (size_t)(x)) * TS + (size_t)(y)]
#pragma omp parallel
    {
        cplx *buffer = (cplx *)malloc(bb * sizeof(cplx)), *cl = (cplx *)malloc(bb * sizeof(cplx));
#pragma omp for schedule(static)
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                long ipos = (long)i - (long)n2, jpos = (long)j - (long)n2;
                for (int nni1 = -w; nni1 <= w; nni1++)
                    for (int nnj1 = -w; nnj1 <= w; nnj1++) {
                        size_t ax = (size_t)((2 * (long)K - (ipos + nni1)) % (long)K);
                        size_t ay = (size_t)((2 * (long)K - (jpos + nnj1)) % (long)K);
                        zgemm_small(b, T, G0AT(ax, ay), buffer);
                        for (int nni2 = -w; nni2 <= w; nni2++)
                            for (int nnj2 = -w; nnj2 <= w; nnj2++) {
                                size_t bx = (size_t)(((long)K + ipos + nni2) % (long)K);
                                size_t by = (size_t)(((long)K + jpos + nnj2) % (long)K);
                                zgemm_small(b, G0AT(bx, by), buffer, cl);
                                size_t cx = (size_t)(((long)K - nni1 + nni2) % (long)K);
                                size_t cy = (size_t)(((long)K - nnj1 + nnj2) % (long)K);
                                const cplx *gb = G0AT(cx, cy);
                                for (size_t ii = 0; ii < oversamp; ii++)
                                    for (size_t jj = 0; jj < oversamp; jj++) {
                                        size_t x = ii + window * oversamp, y = jj + window * oversamp;
                                        double *out = &ldos[((i * oversamp + ii) % W) + ((j * oversamp + jj) % W) * W];
                                        size_t x1 = x - nni1 * oversamp, y1 = y - nnj1 * oversamp;
                                        size_t x2 = x - nni2 * oversamp, y2 = y - nnj2 * oversamp;
                                        int nsets = spin ? 2 : 1;
                                        size_t nper = spin ? spinbands : maxband;
                                        for (int sg = 0; sg < nsets; sg++) {
                                            size_t off = sg ? spinbands : 0;
                                            for (size_t o1 = off; o1 < off + nper; o1++) {
                                                if (wfzero && wfzero[o1]) continue;
                                                for (size_t o2 = off; o2 < off + nper; o2++) {
                                                    if (wfzero && wfzero[o2]) continue;
                                                    double factor = WF(o1, x1, y1) * WF(o2, x2, y2);
                                                    size_t col = o1 + maxband;
                                                    cplx val = (cl[o2 * b + col] + gb[o2 * b + col]) * tip;
                                                    *out -= cimag(val) * factor;
                                                }
                                            }
                                        }
                                    }
                            }
                    }
            }
        free(buffer);
        free(cl);
    }
#undef G0AT
#undef WF
}

/* ------------------------------------------------------------------------- */
/* Threshold list: mkthresholdlist src/qpi.cc:139-211.  Entries {i, j, o1, o2, */
/* factor} for every window pair (xn1, yn1, xn2, yn2) in [0, 2w]^4 whose tile */
/* product exceeds threshold * max(w^2) in magnitude, in the reference's     */
/* order (spin: the (o1, o2) entry, then the (sb+o1, sb+o2) entry).  counts   */
/* receives the (2w+1)^4 list lengths in IDX4CC order; out (may be NULL)      */
/* receives up to cap entries.  Returns the total count.                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t i, j, o1, o2;
    double factor;
} or_flag;

size_t or_mkthresholdlist(const double *wf, size_t norb, size_t window, size_t oversamp, double threshold,
                          size_t maxband, int spin, or_flag *out, size_t cap, size_t *counts)
{
    size_t TS = (2 * window + 1) * oversamp, wins = 2 * window + 1, count = 0;
    double maxval = 0.0;
    if (!maxband) maxband = norb;
    size_t spinbands = maxband >> 1;
#define WF(o, x, y) wf[((size_t)(o) * TS + (size_t)(x)) * TS + (size_t)(y)]
    for (size_t l = 0; l < norb; l++)
        for (size_t i = 0; i < TS; i++)
            for (size_t j = 0; j < TS; j++) {
                double v = WF(l, i, j) * WF(l, i, j);
                if (v > maxval) maxval = v;
            }
    size_t np = spin ? spinbands : maxband;
    for (size_t xn1 = 0; xn1 < wins; xn1++)
        for (size_t yn1 = 0; yn1 < wins; yn1++)
            for (size_t xn2 = 0; xn2 < wins; xn2++)
                for (size_t yn2 = 0; yn2 < wins; yn2++) {
                    size_t before = count;
                    for (size_t o1 = 0; o1 < np; o1++)
                        for (size_t o2 = 0; o2 < np; o2++)
                            for (size_t i = 0; i < oversamp; i++)
                                for (size_t j = 0; j < oversamp; j++) {
                                    size_t x = i + (wins - 1) * oversamp, y = j + (wins - 1) * oversamp;
                                    for (int part = 0; part < (spin ? 2 : 1); part++) {
                                        size_t a = part ? spinbands + o1 : o1, c = part ? spinbands + o2 : o2;
                                        double f = WF(a, x - xn1 * oversamp, y - yn1 * oversamp) *
                                                   WF(c, x - xn2 * oversamp, y - yn2 * oversamp);
                                        if (fabs(f) > threshold * maxval) {
                                            if (out && count < cap) {
                                                or_flag e = {(uint32_t)i, (uint32_t)j, (uint32_t)a, (uint32_t)c, f};
                                                out[count] = e;
                                            }
                                            count++;
                                        }
                                    }
                                }
                    if (counts) counts[((xn1 * wins + yn1) * wins + xn2) * wins + yn2] = count - before;
                }
#undef WF
    return count;
}

/* ------------------------------------------------------------------------- */
/* Threshold list with spin labels: mkthresholdlist(..., spinarr)            */
/* src/qpi.cc:213-253.  spinarr is resized to norb (src/qpi.cc:217: missing  */
/* labels 0); entries for every o1 with a nonzero label and every o2 with the */
/* same label.  Returns the total count.                                      */
/* ------------------------------------------------------------------------- */
size_t or_mkthresholdlist_spinarr(const double *wf, size_t norb, size_t window, size_t oversamp, double threshold,
                                  const int *spinarr, size_t nspinarr, or_flag *out, size_t cap, size_t *counts)
{
    size_t TS = (2 * window + 1) * oversamp, wins = 2 * window + 1, count = 0;
    double maxval = 0.0;
    int *sp = (int *)calloc(norb ? norb : 1, sizeof(int));
    for (size_t o = 0; o < norb && o < nspinarr; o++) sp[o] = spinarr[o];
#define WF(o, x, y) wf[((size_t)(o) * TS + (size_t)(x)) * TS + (size_t)(y)]
    for (size_t l = 0; l < norb; l++)
        for (size_t i = 0; i < TS; i++)
            for (size_t j = 0; j < TS; j++) {
                double v = WF(l, i, j) * WF(l, i, j);
                if (v > maxval) maxval = v;
            }
    for (size_t xn1 = 0; xn1 < wins; xn1++)
        for (size_t yn1 = 0; yn1 < wins; yn1++)
            for (size_t xn2 = 0; xn2 < wins; xn2++)
                for (size_t yn2 = 0; yn2 < wins; yn2++) {
                    size_t before = count;
                    for (size_t o1 = 0; o1 < norb; o1++)
                        if (sp[o1])
                            for (size_t o2 = 0; o2 < norb; o2++)
                                if (sp[o1] == sp[o2])
                                    for (size_t i = 0; i < oversamp; i++)
                                        for (size_t j = 0; j < oversamp; j++) {
                                            size_t x = i + (wins - 1) * oversamp, y = j + (wins - 1) * oversamp;
                                            double f = WF(o1, x - xn1 * oversamp, y - yn1 * oversamp) *
                                                       WF(o2, x - xn2 * oversamp, y - yn2 * oversamp);
                                            if (fabs(f) > threshold * maxval) {
                                                if (out && count < cap) {
                                                    or_flag e = {(uint32_t)i, (uint32_t)j, (uint32_t)o1, (uint32_t)o2, f};
                                                    out[count] = e;
                                                }
                                                count++;
                                            }
                                        }
                    if (counts) counts[((xn1 * wins + yn1) * wins + xn2) * wins + yn2] = count - before;
                }
#undef WF
    free(sp);
    return count;
}

/* ------------------------------------------------------------------------- */
/* LDOS from a threshold list: tmatrix::calcwannierldos(oversamp, window,      */
/* flaglist) src/qpi.cc:1088-1135 (GPU kernel src/qpihipkernel.cc:273-315).   */
/* flags/counts as produced by or_mkthresholdlist (any list of that layout).  */
/* ------------------------------------------------------------------------- */
void or_wannierldoslist(size_t K, size_t n, size_t window, size_t oversamp, size_t b, const double *Tin,
                        const double *g0_in, const or_flag *flags, const size_t *counts, double *ldos)
{
    const cplx *T = (const cplx *)Tin, *g0 = (const cplx *)g0_in;
    size_t W = oversamp * n, bb = b * b, wins = 2 * window + 1, npair = wins * wins * wins * wins;
    size_t n2 = n >> 1;
    int w = (int)window;
    size_t *offs = (size_t *)malloc((npair + 1) * sizeof(size_t));
    offs[0] = 0;
    for (size_t p = 0; p < npair; p++) offs[p + 1] = offs[p] + counts[p];
    for (size_t i = 0; i < W * W; i++) ldos[i] = 0.0;
#define G0AT(ix, iy) (g0 + ((size_t)(ix) * K + (size_t)(iy)) * bb)
#pragma omp parallel
    {
        cplx *buffer = (cplx *)malloc(bb * sizeof(cplx)), *cl = (cplx *)malloc(bb * sizeof(cplx));
#pragma omp for schedule(static)
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                long ipos = (long)i - (long)n2, jpos = (long)j - (long)n2;
                for (int nni1 = -w; nni1 <= w; nni1++)
                    for (int nnj1 = -w; nnj1 <= w; nnj1++) {
                        size_t ax = (size_t)((2 * (long)K - (ipos + nni1)) % (long)K);
                        size_t ay = (size_t)((2 * (long)K - (jpos + nnj1)) % (long)K);
                        zgemm_small(b, T, G0AT(ax, ay), buffer);
                        for (int nni2 = -w; nni2 <= w; nni2++)
                            for (int nnj2 = -w; nnj2 <= w; nnj2++) {
                                size_t bx = (size_t)(((long)K + ipos + nni2) % (long)K);
                                size_t by = (size_t)(((long)K + jpos + nnj2) % (long)K);
                                zgemm_small(b, G0AT(bx, by), buffer, cl);
                                size_t cx = (size_t)(((long)K - nni1 + nni2) % (long)K);
                                size_t cy = (size_t)(((long)K - nnj1 + nnj2) % (long)K);
                                const cplx *gb = G0AT(cx, cy);
                                size_t p = (((size_t)(nni1 + w) * wins + (size_t)(nnj1 + w)) * wins +
                                            (size_t)(nni2 + w)) * wins + (size_t)(nnj2 + w);
                                for (size_t ind = offs[p]; ind < offs[p + 1]; ind++) {
                                    const or_flag *f = &flags[ind];
                                    double *out = &ldos[((i * oversamp + f->i) % W) + ((j * oversamp + f->j) % W) * W];
                                    *out -= cimag(cl[f->o2 * b + f->o1]) * f->factor;
                                    *out -= cimag(gb[f->o2 * b + f->o1]) * f->factor;
                                }
                            }
                    }
            }
        free(buffer);
        free(cl);
    }
#undef G0AT
    free(offs);
}

/* ------------------------------------------------------------------------- */
/* Spectral function maps from G0(k) on the K x K grid (block (i, j) at        */
/* ((i*K + j)*b*b), row-major b x b).  Map n x n, x + y*n.                     */
/* SPF: gpucalcspf src/qpihipkernel.cc:317-329 (CPU calcspf src/qpi.cc:657-681 */
/* with shift, n = K): k index (i, j), i, j < n, at pixel ((i-n/2) mod n, ...) */
/* of -sum_{o<maxband} Im G0(k)_oo.                                            */
/* USPF: calcuspf src/qpi.cc:731-775 (GPU :331-370): pixel (i, j), k = ((i -   */
/* n/2)/K, (j - n/2)/K) in the phase, grid index ((i + n/2) mod n) mod K,      */
/* -sum_{k,l} z_k z_l Im[ G0_{l,k} e^{2 pi i k.(x_k - x_l)} ] over maxband      */
/* orbitals (spin: each half).  pos: maxband x 2 (x, y), z: maxband.           */
/* ------------------------------------------------------------------------- */
void or_spf(size_t K, size_t n, size_t b, size_t maxband, const double *g0k_in, double *out)
{
    const cplx *g0 = (const cplx *)g0k_in;
    size_t n2 = n >> 1;
    if (!maxband) maxband = b;
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
            const cplx *G = g0 + (i * K + j) * b * b;
            double tr = 0.0;
            for (size_t o = 0; o < maxband; o++) tr -= cimag(G[o * b + o]);
            out[((n + i - n2) % n) + ((n + j - n2) % n) * n] = tr;
        }
}

void or_uspf(size_t K, size_t n, size_t b, size_t maxband, int spin, const double *pos, const double *z,
             const double *g0k_in, double *out)
{
    const cplx *g0 = (const cplx *)g0k_in;
    size_t n2 = n >> 1;
    if (!maxband) maxband = b;
    size_t sb = maxband >> 1;
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
            double kx = ((double)i - (double)n2) / (double)K, ky = ((double)j - (double)n2) / (double)K;
            size_t ki = ((i + n2) % n) % K, kj = ((j + n2) % n) % K;
            const cplx *G = g0 + (ki * K + kj) * b * b;
            double tr = 0.0;
            int nsets = spin ? 2 : 1;
            size_t np = spin ? sb : maxband;
            for (int sg = 0; sg < nsets; sg++) {
                size_t off = sg ? sb : 0;
                for (size_t k = off; k < off + np; k++)
                    for (size_t l = off; l < off + np; l++) {
                        double ph = 2.0 * M_PI * (kx * (pos[2 * k] - pos[2 * l]) + ky * (pos[2 * k + 1] - pos[2 * l + 1]));
                        tr -= z[k] * z[l] * cimag(G[l * b + k] * cexp(I * ph));
                    }
            }
            out[i + j * n] = tr;
        }
}

/* ------------------------------------------------------------------------- */
/* Surface / bulk Green's function by decimation:                             */
/* tmatrix::calcsurfacegreensfunction src/qpi.cc:488-554 with                 */
/* tightbind::setmatrix(k, H, kzdep) src/tightbinding.cc:184-208 (in-plane:   */
/* hops with rz == 0 plus the -/+ efermi diagonal; out-of-plane: hops with    */
/* rz < 0, no diagonal), matnorm src/qpi.cc:55-61 (Frobenius norm) and        */
/* calcgreensfunctioninv src/tightbinding.h:75-84 ((w + i eta - H)^-1).       */
/* shops (may be NULL): the surface layer's model for eps_s.  Writes G0(k) of */
/* the K x K grid; returns the largest iteration count used (or -1 if maxit   */
/* was reached somewhere).                                                    */
/* ------------------------------------------------------------------------- */
static void or_setmatrix_kz(const or_hop *hops, size_t nh, size_t b, double efermi, size_t scmax, double kx,
                            double ky, int outofplane, cplx *H)
{
    for (size_t i = 0; i < b * b; i++) H[i] = 0;
    if (!outofplane)
        for (size_t i = 0; i < b; i++) H[i * b + i] = (i < scmax) ? -efermi : efermi;
    for (size_t m = 0; m < nh; m++) {
        const or_hop *h = &hops[m];
        if (outofplane ? !(h->rz < 0.0) : !(h->rz == 0.0)) continue;
        double arg = 2.0 * M_PI * (h->rx * kx + h->ry * ky);
        if (h->anti) arg = -arg;
        H[h->o1 * b + h->o2] += (h->tre + I * h->tim) * (cos(arg) + I * sin(arg));
    }
}

/* (w + i eta - H)^-1 by Gauss-Jordan with partial pivoting */
static void or_resinv(size_t b, const cplx *H, double omega, double eta, cplx *G)
{
    cplx *M = (cplx *)malloc(b * b * sizeof(cplx));
    for (size_t r = 0; r < b; r++)
        for (size_t c = 0; c < b; c++) {
            M[r * b + c] = (r == c ? omega + I * eta : 0) - H[r * b + c];
            G[r * b + c] = r == c ? 1 : 0;
        }
    for (size_t j = 0; j < b; j++) {
        size_t piv = j;
        for (size_t i = j + 1; i < b; i++)
            if (cabs(M[i * b + j]) > cabs(M[piv * b + j])) piv = i;
        if (piv != j)
            for (size_t c = 0; c < b; c++) {
                cplx t = M[j * b + c]; M[j * b + c] = M[piv * b + c]; M[piv * b + c] = t;
                t = G[j * b + c]; G[j * b + c] = G[piv * b + c]; G[piv * b + c] = t;
            }
        cplx ip = 1.0 / M[j * b + j];
        for (size_t c = 0; c < b; c++) {
            M[j * b + c] *= ip;
            G[j * b + c] *= ip;
        }
        for (size_t i = 0; i < b; i++) {
            if (i == j) continue;
            cplx l = M[i * b + j];
            for (size_t c = 0; c < b; c++) {
                M[i * b + c] -= l * M[j * b + c];
                G[i * b + c] -= l * G[j * b + c];
            }
        }
    }
    free(M);
}

int or_surfacegreen(const or_hop *hops, size_t nh, const or_hop *shops, size_t nsh, size_t b, double efermi,
                    size_t scmax, size_t K, double omega, double eta, double epserr, int surface, int maxit,
                    double *g0k_out)
{
    cplx *g0 = (cplx *)g0k_out;
    size_t bb = b * b;
    int worst = 0;
#pragma omp parallel
    {
        cplx *es = (cplx *)malloc(bb * sizeof(cplx)), *eb = (cplx *)malloc(bb * sizeof(cplx)),
             *al = (cplx *)malloc(bb * sizeof(cplx)), *be = (cplx *)malloc(bb * sizeof(cplx)),
             *g = (cplx *)malloc(bb * sizeof(cplx)), *t1 = (cplx *)malloc(bb * sizeof(cplx)),
             *t2 = (cplx *)malloc(bb * sizeof(cplx));
        int local = 0;
#pragma omp for schedule(static)
        for (size_t i = 0; i < K; i++)
            for (size_t j = 0; j < K; j++) {
                double kx = or_kcoord(i, K), ky = or_kcoord(j, K);
                if (shops) or_setmatrix_kz(shops, nsh, b, efermi, scmax, kx, ky, 0, es);
                else or_setmatrix_kz(hops, nh, b, efermi, scmax, kx, ky, 0, es);
                or_setmatrix_kz(hops, nh, b, efermi, scmax, kx, ky, 0, eb);
                or_setmatrix_kz(hops, nh, b, efermi, scmax, kx, ky, 1, al);
                for (size_t r = 0; r < b; r++)
                    for (size_t c = 0; c < b; c++) be[r * b + c] = conj(al[c * b + r]);
                int it = 0;
                for (;;) {
                    or_resinv(b, eb, omega, eta, g);
                    zgemm_small(b, g, be, t1);           /* g beta */
                    zgemm_small(b, al, t1, t2);          /* alpha g beta */
                    for (size_t m = 0; m < bb; m++) {
                        es[m] += t2[m];
                        eb[m] += t2[m];
                    }
                    zgemm_small(b, g, al, t1);           /* g alpha */
                    zgemm_small(b, be, t1, t2);          /* beta g alpha */
                    for (size_t m = 0; m < bb; m++) eb[m] += t2[m];
                    zgemm_small(b, al, t1, t2);          /* alpha g alpha */
                    memcpy(al, t2, bb * sizeof(cplx));
                    zgemm_small(b, g, be, t1);
                    zgemm_small(b, be, t1, t2);          /* beta g beta */
                    memcpy(be, t2, bb * sizeof(cplx));
                    it++;
                    double na = 0, nb2 = 0;
                    for (size_t m = 0; m < bb; m++) {
                        na += creal(al[m]) * creal(al[m]) + cimag(al[m]) * cimag(al[m]);
                        nb2 += creal(be[m]) * creal(be[m]) + cimag(be[m]) * cimag(be[m]);
                    }
                    if (!(sqrt(na) > epserr || sqrt(nb2) > epserr)) break;
                    if (it >= maxit) {
                        it = -1;
                        break;
                    }
                }
                or_resinv(b, surface ? es : eb, omega, eta, g0 + (i * K + j) * bb);
                if (it < 0) local = -1;
                else if (local >= 0 && it > local) local = it;
            }
#pragma omp critical
        {
            if (local < 0 || worst < 0) worst = -1;
            else if (local > worst) worst = local;
        }
        free(es); free(eb); free(al); free(be); free(g); free(t1); free(t2);
    }
    return worst;
}

/* ------------------------------------------------------------------------- */
/* Whole path for a list of energies (the energy loop of src/calcqpi.cc:671-760 */
/* with green=normal and output=wannier).  ldos_out: E x W x W, x + y*W.       */
/* Returns nonzero if a T-matrix inversion failed.                            */
/* ------------------------------------------------------------------------- */
int or_cldos_run(const or_hop *hops, size_t nh, size_t b, double efermi, int scmodel,
                 size_t K, size_t n, size_t window, size_t oversamp, size_t maxband, int spin,
                 double eta, const double *omegas, size_t E, const double *V,
                 const double *wf, const int *wfzero, double *ldos_out)
{
    size_t scmax = scmodel ? (b >> 1) : b;
    size_t W = oversamp * n;
    double *eval = (double *)malloc(K * K * b * sizeof(double));
    double *evec = (double *)malloc(K * K * b * b * 2 * sizeof(double));
    double *g0 = (double *)malloc(K * K * b * b * 2 * sizeof(double));
    double *T = (double *)malloc(b * b * 2 * sizeof(double));
    int err = 0;
    or_calchamiltonian(hops, nh, b, efermi, scmax, K, eval, evec);
    for (size_t e = 0; e < E && !err; e++) {
        or_setgreensfunction(K, b, eval, evec, omegas[e], eta, g0);
        or_realspace(K, b, g0);
        err = or_tmatrix(b, V, g0, T);
        if (!err)
            or_wannierldos(K, n, window, oversamp, b, maxband, spin, T, g0, wf, wfzero,
                           ldos_out + e * W * W);
    }
    free(eval); free(evec); free(g0); free(T);
    return err;
}

/* Energies of the layers: idl::getbias, src/idl.cc:969-973 */
double or_getbias(double lb, double hb, size_t layers, size_t i)
{
    if (layers == 1) return lb;
    return (hb - lb) * (double)i / ((double)layers - 1.0) + lb;
}

/* ------------------------------------------------------------------------- */
/* idl map file: idl::write src/idl.cc:78-95, idl::read src/idl.cc:50-74      */
/* data[x + y*sx + z*sx*sy] as doubles, stored as float32, x fastest.         */
/* ------------------------------------------------------------------------- */
int or_idl_write(const char *path, const char *name, const char *date, size_t sx, size_t sy,
                 size_t sz, const double hdr[7], const double *data)
{
    FILE *f = fopen(path, "wb");
    if (!f) return 1;
    /* header: name, date, sizes, scan range (2), offset (2), bias (2), setpoint,
     * each on its own line; doubles printed as a default C++ ostream would (%g) */
    fprintf(f, "%s\n%s\n%zu\n%zu\n%zu\n%g\n%g\n%g\n%g\n%g\n%g\n%g\n", name, date, sx, sy, sz,
            hdr[0], hdr[1], hdr[2], hdr[3], hdr[4], hdr[5], hdr[6]);
    size_t tot = sx * sy * sz;
    for (size_t i = 0; i < tot; i++) {
        float v = (float)data[i];
        fwrite(&v, sizeof(float), 1, f);
    }
    int bad = ferror(f);
    fclose(f);
    return bad ? 2 : 0;
}

int or_idl_read_header(const char *path, size_t dims[3], double hdr[7])
{
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    char *buf = NULL;
    size_t cap = 0;
    int err = 0;
    if (!or_getline(f, &buf, &cap) || !or_getline(f, &buf, &cap)) err = 2;
    if (!err && fscanf(f, "%zu %zu %zu %lf %lf %lf %lf %lf %lf %lf", &dims[0], &dims[1], &dims[2],
                       &hdr[0], &hdr[1], &hdr[2], &hdr[3], &hdr[4], &hdr[5], &hdr[6]) != 10)
        err = 3;
    free(buf);
    fclose(f);
    return err;
}

int or_idl_read(const char *path, double *data, size_t maxcount)
{
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    char *buf = NULL;
    size_t cap = 0;
    size_t dims[3];
    double hdr[7];
    int err = 0;
    if (!or_getline(f, &buf, &cap) || !or_getline(f, &buf, &cap)) err = 2;
    if (!err && fscanf(f, "%zu %zu %zu %lf %lf %lf %lf %lf %lf %lf", &dims[0], &dims[1], &dims[2],
                       &hdr[0], &hdr[1], &hdr[2], &hdr[3], &hdr[4], &hdr[5], &hdr[6]) != 10)
        err = 3;
    if (!err) {
        fgetc(f); /* is.ignore(1,'\n') */
        size_t tot = dims[0] * dims[1] * dims[2];
        if (tot > maxcount) err = 4;
        for (size_t i = 0; i < tot && !err; i++) {
            float v;
            if (fread(&v, sizeof(float), 1, f) != 1) { err = 5; break; }
            data[i] = v;
        }
    }
    free(buf);
    fclose(f);
    return err;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

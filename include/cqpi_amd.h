/*
 * cqpi_amd.h -- C ABI of the MI355X continuum-LDOS library (libcqpi_amd.so).
 *
 * Plain C types only: sizes, pointers to double (complex numbers are
 * interleaved re,im pairs), status codes.  Every entry point returns 0 on
 * success and a nonzero CQPI_E_* code on failure; cqpi_last_error() then
 * holds a message (per thread).
 *
 * The first group replaces the reference's GPU plug-in, class GPUQPI
 * (include/gpuqpi.h:18-32) as implemented by HipQPI (src/qpihipkernel.h:64-88,
 * src/qpihipkernel.cc:395-704).  The second group extends it: the whole
 * chain (eigensolve, resolvent, k->R transform, T-matrix, contraction) runs on
 * the device, which the reference does on the host in tmatrix
 * (src/qpi.cc:454-486, :557-578, :833-854).
 *
 * Layouts:
 *   complex b x b matrices: row-major, element (r,c) at [2*(r*ld+c)] with a
 *     leading dimension ld (= b unless stated)
 *   tiles wf: maxband x T x T doubles, T = (2*window+1)*oversamp, element
 *     (o,x,y) at [(o*T+x)*T+y]  (HipQPI packs IDX3C(o,x,y,T), src/qpihipkernel.cc:414-418)
 *   LDOS map: W x W doubles, W = oversamp*n, element (x,y) at [x + y*W]
 *     (src/qpihipkernel.cc:96, src/qpi.cc:821-831); a batch of E maps is
 *     E consecutive maps
 *
 * Threading: a context is used by one host thread at a time (the reference's
 * HipQPI is not thread-safe either).  All device work of a context runs on
 * the context's own HIP stream.
 */
#ifndef CQPI_AMD_H
#define CQPI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CQPI_OK = 0,
    CQPI_E_ARG = 1,      /* invalid argument / shape */
    CQPI_E_HIP = 2,      /* HIP runtime error (allocation, launch, copy) */
    CQPI_E_STATE = 3,    /* call out of order (e.g. run before set_model) */
    CQPI_E_SINGULAR = 4, /* 1 - V G0(0) singular at some energy */
    CQPI_E_IO = 5,       /* file could not be read or parsed */
    CQPI_E_NODEV = 6,    /* no HIP device */
    CQPI_E_CONVERGE = 7  /* surface Green's function not converged */
};

/* Precision of a context's runs.  In every mode the eigensolve, the
 * T-matrix solve and the periodic (bulk) term are FP64, and the G0(R) window
 * that the T-matrix reads is stored in FP64.  In the FP32 and mixed modes the
 * resolvent and both k->R transform passes run on FP32 lines (for
 * power-of-two k-grids from 64 to 1024; other grids keep FP64 lines), and the
 * surface Green's function G(k) is stored in FP32; the contraction then runs
 * as below, with T applied in FP32 and the final sum in FP64. */
enum {
    CQPI_PREC_FP64 = 0,  /* FP64 everywhere */
    CQPI_PREC_FP32 = 1,  /* contraction in FP32 on the vector ALUs */
    CQPI_PREC_MIXED = 2  /* contraction on MFMA: split-FP16 hi+lo operands, FP32 accumulation */
};

/* One hopping term, orbitals 0-based, amplitude already divided by the
 * Wannier90 degeneracy (hoppingterm, src/tightbinding.h:111-164). 56 bytes. */
typedef struct {
    double rx, ry, rz;
    int32_t o1, o2;
    double tre, tim;
    int32_t anti;     /* 1: conjugate phase, both orbitals in the hole block */
    int32_t reserved;
} cqpi_hopping;

typedef struct cqpi_ctx cqpi_ctx;

/* One entry of a threshold list: sub-pixel (i, j) of the cell, orbitals
 * o1, o2 and the tile product factor; the layout of the reference's
 * qpigpuflaglist (src/qpihipkernel.h:47-50) with 32-bit indices. */
typedef struct {
    uint32_t i, j, o1, o2;
    double factor;
} cqpi_flagentry;

/* ---------------------------------------------------------------------- */
/* Status                                                                  */
/* ---------------------------------------------------------------------- */
/* Message of the last failed call on this thread ("" if none). */
const char *cqpi_last_error(void);
/* Library version string. */
const char *cqpi_version(void);
/* Number of visible HIP devices (0 if none); never fails. */
int cqpi_device_count(void);

/* ---------------------------------------------------------------------- */
/* GPUQPI drop-in (one energy per call, host-computed G0 and T)            */
/* ---------------------------------------------------------------------- */

/* Replaces HipQPI::HipQPI(wanniern, kpoints, n, window, oversamp, bands,
 * maxband, spin, wf) (src/qpihipkernel.h:75, src/qpihipkernel.cc:395-429).
 * wanniern must equal oversamp*n.  wf: maxband tiles (layout above); zero
 * tiles stand for orbitals without a Wannier function.  Allocates the device
 * buffers and uploads the tiles once. */
int cqpi_create(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp,
                size_t bands, size_t maxband, int spin, const double *wf, cqpi_ctx **out);

/* Replaces HipQPI::wannierldos(scat, g0) (include/gpuqpi.h:23,
 * src/qpihipkernel.cc:577-598).  scat: T, b x b complex row-major (ld=b).
 * g0: table of kpoints*kpoints pointers, entry [ix*kpoints+iy] pointing to the
 * b x b complex matrix G0(R = (ix, iy) mod kpoints) with leading dimension
 * tda (complex elements), i.e. exactly the reference's
 * vector<vector<gsl_matrix_complex*>> g0[ix][iy] with ->data and ->tda.
 * Only the window of R actually used is read.  The data is copied before the
 * call returns (the caller overwrites g0 for the next energy, src/calcqpi.cc:721-750);
 * the contraction runs asynchronously. */
int cqpi_wannierldos(cqpi_ctx *ctx, const double *scat, const double *const *g0, size_t tda);

/* Replaces HipQPI::wannierjosephson(scat, g0, tip) (include/gpuqpi.h:24,
 * src/qpihipkernel.cc:600-622; kernels :171-271, CPU src/qpi.cc:953-1033):
 * the Josephson map of a superconducting model,
 *   rho(x) = - sum Im[ tip (G0(pos+R2) T G0(-pos-R1) + G0(R2-R1))_{o2, o1+maxband} ]
 *            w_o1(x-R1) w_o2(x-R2)
 * over o1, o2 < maxband (one spin block at a time with spin), i.e. the
 * particle-hole block.  The context must have been created with
 * 2*maxband <= bands (the reference passes maxband = bands/2).  Arguments as
 * cqpi_wannierldos; tip = tip_re + i tip_im (the driver's energy-dependent
 * tip factor, src/calcqpi.cc:705-709).  Any precision mode; retrieve the map
 * with cqpi_retrieve. */
int cqpi_wannierjosephson(cqpi_ctx *ctx, const double *scat, const double *const *g0, size_t tda, double tip_re,
                          double tip_im);

/* Size in bytes of the fixed per-device scratch of cqpi_wannierldos_device. */
#define CQPI_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cqpi_wannierldos for E energies whose inputs are
 * already in device memory, with no allocation of any kind.
 *   d_T      E matrices T, b x b complex row-major each
 *   d_g0win  E windows of G0(R), the layout of cqpi_retrieve_g0_window
 *            (side nw = 2h+1, h from cqpi_window_half), one after the other
 *   d_ldos   E maps of W x W doubles (output)
 * The context's precision applies (FP64, FP32 or mixed; set it beforehand:
 * cqpi_set_precision allocates what the mode keeps per context).  The
 * periodic table, and the fp16 planes and scales (mixed) or FP32 copies
 * (FP32), live in one fixed scratch array of CQPI_NOALLOC_SCRATCH_BYTES per
 * device that the library holds until the process ends; a call needing more
 * returns CQPI_E_ARG (pass fewer energies per call).  Every byte of scratch a
 * call reads is written by that call first.  The kernels run on the
 * context's stream and the call returns without waiting (cqpi_retrieve does
 * not apply: the maps are in d_ldos).  Calls that overlap in time on one GPU,
 * from any context or thread, share that scratch and race. */
int cqpi_wannierldos_device(cqpi_ctx *ctx, const double *d_T, const double *d_g0win, size_t E, double *d_ldos);

/* Replaces HipQPI::HipQPI(wanniern, kpoints, n, window, oversamp, bands,
 * flaglist) (src/qpihipkernel.h:76, src/qpihipkernel.cc:431-493): a context
 * for the threshold-list contraction.  counts holds the (2w+1)^4 list
 * lengths of the window pairs (xn1, yn1, xn2, yn2), R = xn - w, in the order
 * ((xn1*(2w+1) + yn1)*(2w+1) + xn2)*(2w+1) + yn2 (IDX4CC); flags holds the
 * entries of all lists one after the other, as the reference flattens them
 * (src/qpihipkernel.cc:452-480).  Up to 16 bands; FP64 or FP32.  The context
 * serves cqpi_wannierldoslist and cqpi_run (maps from the list). */
int cqpi_create_list(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp, size_t bands,
                     const cqpi_flagentry *flags, const size_t *counts, cqpi_ctx **out);

/* Replaces HipQPI::wannierldoslist(scat, g0) (include/gpuqpi.h:25,
 * src/qpihipkernel.cc:624-636; kernel :273-315, CPU src/qpi.cc:1088-1135):
 *   rho(x) = - sum over the list entries of (R1, R2) with sub-pixel x of
 *            factor Im[ (G0(pos+R2) T G0(-pos-R1) + G0(R2-R1))_{o2,o1} ].
 * Arguments as cqpi_wannierldos; the context must come from
 * cqpi_create_list.  Retrieve the map with cqpi_retrieve. */
int cqpi_wannierldoslist(cqpi_ctx *ctx, const double *scat, const double *const *g0, size_t tda);

/* The driver's list builder mkthresholdlist (src/qpi.cc:139-211; called at
 * src/calcqpi.cc:620-634): entries whose tile product exceeds
 * threshold * max(w^2) in magnitude, over all window pairs, orbitals
 * o1, o2 < maxband (0: norb; with spin both in one half) and sub-pixels, in
 * the reference's order.  wf: norb tiles, layout as cqpi_create.  counts
 * receives the (2w+1)^4 list lengths, *total the entry count; out (may be
 * NULL) receives up to cap entries (CQPI_E_ARG if the list is longer; call
 * with out = NULL first to size it). */
int cqpi_make_threshold_list(const double *wf, size_t norb, size_t window, size_t oversamp, double threshold,
                             size_t maxband, int spin, cqpi_flagentry *out, size_t cap, size_t *counts,
                             size_t *total);

/* The driver's list builder with a spin label per orbital, mkthresholdlist(
 * ..., spinarr) (src/qpi.cc:213-253; called at src/calcqpi.cc:626-628 when
 * the input file has spinarr=): entries over every orbital pair o1, o2 < norb
 * whose labels are equal and nonzero, in the reference's order.  spinarr has
 * nspinarr labels; it is resized to norb (missing labels are 0, and an
 * orbital labelled 0 has no entries).  Other arguments and results as
 * cqpi_make_threshold_list. */
int cqpi_make_threshold_list_spinarr(const double *wf, size_t norb, size_t window, size_t oversamp,
                                     double threshold, const int *spinarr, size_t nspinarr, cqpi_flagentry *out,
                                     size_t cap, size_t *counts, size_t *total);

/* Replaces HipQPI::HipQPI(kpoints, n, bands, maxbands, spin, pos, prearr)
 * (src/qpihipkernel.h:77, src/qpihipkernel.cc:498-533): a context for the
 * spectral-function maps, n x n (n <= kpoints; the driver sets n = kpoints
 * for SPF, src/calcqpi.cc:333-334).  pos: npos x 2 orbital positions x, y
 * (npos = 0: all at the origin, else npos >= maxbands); prearr: nprearr
 * prefactors, orbital i uses prearr[i % nprearr] (nprearr = 0: all 1).
 * maxbands = 0 means bands.  The maps are retrieved with cqpi_retrieve. */
int cqpi_create_spectral(size_t kpoints, size_t n, size_t bands, size_t maxbands, int spin, const double *pos,
                         size_t npos, const double *prearr, size_t nprearr, cqpi_ctx **out);

/* Replaces HipQPI::spf(g0) (include/gpuqpi.h:26, src/qpihipkernel.cc:638-643;
 * kernel :317-329, CPU src/qpi.cc:657-681): the spectral function
 *   A(k) = -sum_{o<maxbands} Im G0(k)_oo
 * of grid point (i, j), i, j < n, at pixel ((i - n/2) mod n, (j - n/2) mod n).
 * g0: the kpoints*kpoints pointer table of cqpi_wannierldos, here holding
 * G0(k) (the reference does not transform it for this mode); copied before
 * the call returns. */
int cqpi_spf(cqpi_ctx *ctx, const double *const *g0, size_t tda);

/* Replaces HipQPI::uspf(g0) (include/gpuqpi.h:27, src/qpihipkernel.cc:645-653;
 * kernels :331-370, CPU src/qpi.cc:731-775): the unfolded spectral function
 * at pixel (i, j), q = ((i - n/2)/kpoints, (j - n/2)/kpoints),
 *   -sum_{k,l} z_k z_l Im[ G0(k_g)_{l,k} exp(2 pi i q.(x_k - x_l)) ]
 * over maxbands orbitals (with spin: within each half), k_g the grid point
 * ((i + n/2) mod n, (j + n/2) mod n) mod kpoints.  Arguments as cqpi_spf. */
int cqpi_uspf(cqpi_ctx *ctx, const double *const *g0, size_t tda);

/* Replaces HipQPI::retrieveResult(double *ldos) (src/qpihipkernel.cc:690-704):
 * waits for the device and copies the W x W map of the last wannierldos call
 * (or layer 0 of the last cqpi_run) to ldos. */
int cqpi_retrieve(cqpi_ctx *ctx, double *ldos);

/* Replaces HipQPI::printinfo (src/qpihipkernel.cc:535-560): device name and
 * memory use as text; writes at most len bytes including the terminator. */
int cqpi_printinfo(cqpi_ctx *ctx, char *buf, size_t len);

/* Replaces HipQPI::~HipQPI (src/qpihipkernel.cc:562-575).  NULL is a no-op. */
void cqpi_destroy(cqpi_ctx *ctx);

/* Sets the contraction precision (CQPI_PREC_*) for later calls. */
int cqpi_set_precision(cqpi_ctx *ctx, int precision);

/* ---------------------------------------------------------------------- */
/* Whole chain on the device (extension of the plug-in)                    */
/* ---------------------------------------------------------------------- */

/* Tight-binding model: H(k) = sum t exp(+-2 pi i R.k) - efermi (hole block
 * +efermi when scmodel), src/tightbinding.cc:210-219.  Invalidates cached
 * eigen data.  nhops may be 0 (then H(k) is the constant diagonal). */
int cqpi_set_model(cqpi_ctx *ctx, const cqpi_hopping *hops, size_t nhops, double efermi,
                   int scmodel);

/* Impurity potential V (b x b complex row-major).  The reference's V is
 * diagonal, V = diag(vscat) (src/qpi.cc:840-841, src/qpi.h:107-130); a full
 * matrix is accepted here. */
int cqpi_set_impurity(cqpi_ctx *ctx, const double *V);

/* Per-k eigensolve of H(k) on the k-grid, cached until the model changes
 * (tmatrix::calchamiltonian, src/qpi.cc:454-475).  cqpi_run calls it when
 * needed; calling it explicitly lets a caller time it separately. */
int cqpi_compute_eigen(cqpi_ctx *ctx);

/* Full cLDOS for E energies with broadening eta: resolvent, k->R transform,
 * T-matrix and contraction, energies processed in device-sized chunks.  The
 * maps stay on the device (retrievable with cqpi_retrieve_layers); if
 * ldos_host is not NULL the E maps are also copied there (E*W*W doubles).
 * Returns CQPI_E_SINGULAR if 1 - V G0(0) is singular at some energy. */
int cqpi_run(cqpi_ctx *ctx, const double *omegas, size_t E, double eta, double *ldos_host);

/* Copies maps [first, first+count) of the last cqpi_run to ldos_host. */
int cqpi_retrieve_layers(cqpi_ctx *ctx, size_t first, size_t count, double *ldos_host);

/* Device pointer of the maps of the last cqpi_run (E*W*W doubles, owned by
 * the context, valid until the next run or destroy), for zero-copy hand-off. */
int cqpi_device_ldos(cqpi_ctx *ctx, const double **dptr, size_t *count);

/* Copies G0(R) of energy index e of the last cqpi_run for the window of R
 * used by the contraction: side nw = 2h+1 with R = slot - h, h as reported
 * by cqpi_window_half; out is nw*nw*b*b complex, [(sx*nw+sy)*b*b + r*b + c].
 * A mixed-precision run keeps only |R| <= 2w of the window in FP64 (all the
 * T-matrix and the periodic term read); after such a run this returns
 * CQPI_E_STATE unless cqpi_set_keep_windows(ctx, 1) was set for the run. */
int cqpi_window_half(cqpi_ctx *ctx, size_t *h);
int cqpi_retrieve_g0_window(cqpi_ctx *ctx, size_t e, double *out);
/* Copies T (b x b complex) of energy index e of the last cqpi_run. */
int cqpi_retrieve_tmatrix(cqpi_ctx *ctx, size_t e, double *out);

/* Per-stage device time of the last cqpi_run / cqpi_wannierldos, measured
 * with HIP events on the context stream.  Stages: 0 eigensolve, 1 resolvent +
 * transform along y, 2 transform along x, 3 T-matrix + periodic term,
 * 4 contraction.  ms[s] = total milliseconds, launches[s] = launch count. */
#define CQPI_NSTAGES 5
int cqpi_stage_times(cqpi_ctx *ctx, double *ms, int64_t *launches);
/* Enables/disables the per-stage event timing (default off). */
int cqpi_set_timing(cqpi_ctx *ctx, int on);

/* Selects the map cqpi_run computes: on = 0 the LDOS map (default), on = 1
 * the Josephson map of a superconducting model (output=josephson in the
 * reference driver, src/calcqpi.cc:705-709 and src/qpi.cc:953-1033), with
 * the tip factor of energy w
 *   tip(w) = i (-sign(w) pi deltat) / sqrt((w + i etat)^2 - deltat^2)
 * (principal root; etat NaN: the run's eta, as the driver's default).  Needs
 * 2*maxband <= bands (any precision mode); a run with it on cannot feed
 * cqpi_gradient. */
int cqpi_set_josephson(cqpi_ctx *ctx, int on, double deltat, double etat);

/* Band-structure histogram of the context's model (bsfile= in the reference
 * driver: tightbind::calcbst3dhisthr, src/tightbinding.cc:1158-1193, called
 * from src/calcqpi.cc:308-326 with ux = uy = 1 and all bands).  The
 * eigenvalues of H(k) at the (pix oversamp)^2 points k = (io/(pix oversamp - 1)
 * - 1/2, jo/(...) - 1/2) are added to pixel (io/oversamp, jo/oversamp) of the
 * three layers around round((e - le)/(ue - le) (layers - 1)) as the Lorentzian
 * w/2 / ((bias - e)^2 + w^2/4) / pi, w = 2 (ue - le)/layers (bias of layer l:
 * le + (ue - le) l/(layers - 1)), for eigenvalues within two layer spacings of
 * [le, ue].  hist: pix x pix x layers doubles, x fastest.  The sums run in
 * any order (atomics): equal to the reference's up to rounding. */
int cqpi_bands_histogram(cqpi_ctx *ctx, size_t pix, size_t oversamp, double le, double ue, size_t layers,
                         double *hist);

/* Green's function the LDOS runs use (green= in the reference driver,
 * src/calcqpi.cc:129, :207-208, :367-375). */
enum {
    CQPI_GREEN_NORMAL = 0,  /* G0(k) of the bulk model from its eigen data (default) */
    CQPI_GREEN_SURFACE = 1, /* surface layer of a semi-infinite stack of layers */
    CQPI_GREEN_BULK = 2     /* a layer deep inside that stack */
};
/* Replaces G0(k) of cqpi_run by the surface or bulk Green's function of the
 * layered model (tmatrix::calcsurfacegreensfunction, src/qpi.cc:488-554,
 * at kz = 0): per k the in-plane block (hoppings with rz == 0 plus the
 * -/+ efermi diagonal) and the coupling to the next layer (rz < 0) are
 * decimated until the Frobenius norms of both couplings are <= epserr; shops
 * (NULL/0: the model's own hoppings) gives the surface layer's in-plane block
 * (stbfile=).  Call after cqpi_set_model; bands 1-12, 14, 16.  It applies to every
 * output of the context: LDOS and Josephson maps, threshold-list maps, the
 * SPF / USPF maps and cqpi_dos.  cqpi_run and cqpi_dos return
 * CQPI_E_CONVERGE if some k needs more steps than the limit
 * (cqpi_set_surface_steps, default 500).  A run with it cannot
 * feed cqpi_gradient.  mode CQPI_GREEN_NORMAL switches it off. */
int cqpi_set_surface(cqpi_ctx *ctx, int mode, double epserr, const cqpi_hopping *shops, size_t nshops);
/* Step limit of the decimation (default 500; the reference loops until it
 * converges, src/qpi.cc:515-541, which a device kernel cannot).  A run in
 * which some k-point needs more steps returns CQPI_E_CONVERGE. */
int cqpi_set_surface_steps(cqpi_ctx *ctx, int maxit);

/* Pair symmetry of the k->R transform (no counterpart in the reference, which
 * transforms every orbital pair, src/qpi.cc:557-578).  Where every hopping
 * block at an integral R (the hole block's terms at their conjugate-phase R)
 * has its exact transpose at -R (real hoppings of a Hermitian model, also as
 * the normal blocks of a particle-hole model with real pairing), H(k)^T = H(-k), so
 * G0(k)^T = G0(-k) and G0(R)_cr = G0(-R)_rc: the transform passes of
 * cqpi_run then carry only the b(b+1)/2 pairs r <= c and write each pair
 * also as (c, r) at -R.  cqpi_set_model detects the property; enable (default
 * 1) lets runs use it (power-of-two k-grids from 64 to 1024, G0 from the
 * eigen data, and in mixed mode the T-folded contraction).  The maps agree
 * with those of enable = 0 to rounding.  cqpi_get_pair_symmetry: model = 1
 * where the model has the property, used = 1 where the last run used it
 * (either pointer may be NULL). */
int cqpi_set_pair_symmetry(cqpi_ctx *ctx, int enable);
int cqpi_get_pair_symmetry(const cqpi_ctx *ctx, int *model, int *used);

/* For a context of cqpi_create_spectral: mode 1 (SPF) or 2 (USPF) makes
 * cqpi_run compute those maps (n x n per energy) from the device's eigen
 * data instead of host G0(k) (output=spf|uspf of the reference driver,
 * src/calcqpi.cc:712-717); 0 selects none. */
int cqpi_set_spectral(cqpi_ctx *ctx, int mode);

/* Density of states -sum_{o<maxband} Im G0(R=0)_oo (tmatrix::calcdos,
 * src/qpi.cc:856-862; the driver's dosfile, src/calcqpi.cc:336-391) for E
 * energies, G0(R=0) = kpoints^-2 sum_k G0(k) from the device's eigen data;
 * dos: E doubles on the host.  Needs cqpi_set_model. */
int cqpi_dos(cqpi_ctx *ctx, const double *omegas, size_t E, double eta, double *dos);

/* ---------------------------------------------------------------------- */
/* Gradient (extension; the reference has no differentiation)              */
/* ---------------------------------------------------------------------- */

/* When on, the next cqpi_run keeps the G0(R) window and T of every energy
 * (all energies in one chunk) so that cqpi_gradient can run; G0(k) itself is
 * never stored.  Default off. */
int cqpi_set_keep_windows(cqpi_ctx *ctx, int on);

/* Hand-written adjoint of the whole chain for the maps of the last cqpi_run
 * (which must have kept its windows), FP64.  rhobar = dL/drho for those maps
 * (E*W*W doubles, layout of the maps).  Outputs:
 *   g_hops[2*h], g_hops[2*h+1] = dL/dRe t_h, dL/dIm t_h for the hoppings in
 *     the order given to cqpi_set_model (amplitudes as passed, i.e. after the
 *     degeneracy division);
 *   *g_mu = dL/defermi;
 *   g_V (b x b complex) with dL = Re sum conj(g_V) dV, i.e. dL/dRe V_ij =
 *     Re g_V_ij and dL/dIm V_ij = Im g_V_ij.
 * Any output pointer may be NULL.  Needs a power-of-two k-grid (<= 1024),
 * bands <= 5, window <= 4.  Where a k row's b^2 lines exceed one workgroup's
 * LDS (b >= 4 with K >= 256, or K = 1024) the call grows the G(k) workspace
 * by a block of energies of Gbar(k), at most 1 GiB. */
int cqpi_gradient(cqpi_ctx *ctx, const double *rhobar, double *g_hops, double *g_mu, double *g_V);

/* As cqpi_gradient with rhobar in device memory (e.g. computed by the caller
 * on the GPU from cqpi_device_ldos); the caller's writes to it must be
 * complete (synchronised) before the call. */
int cqpi_gradient_device(cqpi_ctx *ctx, const double *d_rhobar, double *g_hops, double *g_mu, double *g_V);

/* ---------------------------------------------------------------------- */
/* Launch tuning (extension; the reference has fixed launch shapes)         */
/* ---------------------------------------------------------------------- */

/* Sets the launch tuning of later runs: tpw = cell rows per wave of the
 * mixed-precision contraction (0 = default 8, or 2, 4, 8); eb = energies per
 * workgroup of the first transform pass (0 = built-in heuristic).  Results
 * do not depend on either. */
int cqpi_set_tuning(cqpi_ctx *ctx, int tpw, int eb);
int cqpi_get_tuning(cqpi_ctx *ctx, int *tpw, int *eb);
/* The same with the other two tuned knobs: js = waves per cell row of the
 * T-folded mixed contraction (0 = the shape's default, or 1, 2), streams =
 * streams an eager run splits an energy chunk over (0 = two from 8 energies,
 * 1 = one, 2 = two).  Results do not depend on any of them. */
int cqpi_set_tuning_ext(cqpi_ctx *ctx, int tpw, int eb, int js, int streams);
int cqpi_get_tuning_ext(cqpi_ctx *ctx, int *tpw, int *eb, int *js, int *streams);

/* Autotuner: times candidate tunings on the device for this context's shape
 * and E energies (reps timed runs each, eigen data cached), keeps the fastest
 * and remembers it for the process (contexts of the same shape then reuse it
 * without timing).  Knobs, each timed with the others fixed: eb, then (mixed)
 * tpw and js, then streams.  The maps of the last timed run are left in the
 * context.  *tpw and *eb (may be NULL) receive the choice; cqpi_get_tuning_ext
 * reads all four. */
int cqpi_autotune(cqpi_ctx *ctx, const double *omegas, size_t E, double eta, int reps, int *tpw, int *eb);

/* ---------------------------------------------------------------------- */
/* Batches of independent instances in one graph (extension)                */
/* ---------------------------------------------------------------------- */

/* One instance of a batch: geometry as in cqpi_create (wanniern is
 * oversamp*n), model as in cqpi_set_model, V as in cqpi_set_impurity, and its
 * own energies.  The pointers are read during cqpi_batch_create only. */
typedef struct {
    size_t kpoints, n, window, oversamp, bands, maxband;
    int spin, scmodel;
    const cqpi_hopping *hops;
    size_t nhops;
    double efermi;
    const double *V;      /* b x b complex */
    const double *wf;     /* maxband x T x T tiles */
    const double *omegas; /* E energies */
    size_t E;
    double eta;
} cqpi_instance;

typedef struct cqpi_batch cqpi_batch;

/* Creates a context per instance (all with the given precision). */
int cqpi_batch_create(const cqpi_instance *inst, size_t count, int precision, cqpi_batch **out);
/* Runs cqpi_autotune for every instance (shapes seen before reuse their
 * tuning).  Only before the first cqpi_batch_run. */
int cqpi_batch_autotune(cqpi_batch *b, int reps);
/* Whole batch on the device.  The first call runs every instance once
 * (allocation, validation) and captures the launches of all instances -- per
 * instance the eigensolve, resolvent, both transforms, T-matrix and
 * contraction -- into one HIP graph with parallel branches; every call then
 * launches that graph once and waits for it.  CQPI_E_SINGULAR names the first
 * instance and energy with a singular 1 - V G0(0). */
int cqpi_batch_run(cqpi_batch *b);
int cqpi_batch_size(cqpi_batch *b, size_t *count);
/* Context of instance i (owned by the batch): its maps are read with
 * cqpi_retrieve_layers / cqpi_device_ldos after cqpi_batch_run. */
int cqpi_batch_context(cqpi_batch *b, size_t i, cqpi_ctx **ctx);
/* Device time of the last graph launch, graph node count, device memory. */
int cqpi_batch_info(cqpi_batch *b, double *last_ms, size_t *graph_nodes, size_t *gpumem);
void cqpi_batch_destroy(cqpi_batch *b);

/* ---------------------------------------------------------------------- */
/* Input formats (host)                                                     */
/* ---------------------------------------------------------------------- */

/* Reads a Wannier90 _hr.dat file (tightbind::readfile, src/tightbinding.cc:70-149).
 * On success *hops is allocated with malloc (free with cqpi_free) and holds
 * *nhops terms, *bands is num_wann, name receives line 1 (namelen bytes). */
int cqpi_read_hr(const char *path, int scmodel, cqpi_hopping **hops, size_t *nhops,
                 size_t *bands, char *name, size_t namelen);
void cqpi_free(void *p);

/* One analytic orbital of the driver's orbitals= mode (src/calcqpi.cc:519-575)
 * sampled on the Wannier tile grid: tile[i*T + j], T = (2 window + 1) oversamp,
 * at x = i/oversamp - window - 1/2 (y likewise with j), as
 * wannierfunctions::precalculate (src/wannierfunctions.cc:436-442) with the
 * orbital wannierfunctions(name, radius, height, xpos, ypos, theta, 0, prefact)
 * (src/wannierfunctions.h:56; waves :21-124, shift and rotation :34-40).
 * name: s, px, py, pz, dx2, dxy, dxz, dyz, dr2, fy3x2, fxyz, fyz2, fz3, fxz2,
 * fzx2, fxx2, n (4-fold lobe), z (zero).  theta in radians.  Host only;
 * CQPI_E_ARG for an unknown name or radius <= 0. */
int cqpi_orbital_tile(const char *name, double radius, double height, double xpos, double ypos, double theta,
                      double prefact, size_t window, size_t oversamp, double *tile);
/* Wannier tile of a Wannier90/XCrySDen .xsf orbital file (orbitalfiles= in
 * the reference driver, src/calcqpi.cc:576-601; wannierfunctions(filename,
 * height, window, oversamp, prefact), src/wannierfunctions.cc:444-461, with
 * xcfloader, src/loadxcf.cc:33-349): tile[i*T + j] = prefact * the value of
 * the data-grid node nearest (nearest in each spanning-vector coordinate, 0
 * outside) to xu a1 + yu a2 at height zh, xu = i/oversamp - window, yu
 * likewise, a1, a2 the CONVVEC rows; zh = highest atom z + height (height > 0)
 * or lowest atom z + height (height < 0).  Host only; CQPI_E_IO if the file
 * cannot be read or has no 3D_field data grid. */
int cqpi_xsf_tile(const char *path, double height, size_t window, size_t oversamp, double prefact, double *tile);
/* radius sqrt(ln(1/p)): where exp(-r^2/radius^2) falls to p
 * (wannierfunctions::getrange, src/wannierfunctions.cc:382-384); the driver
 * sizes window = 0 (auto) from it with p = 0.01. */
double cqpi_orbital_range(double radius, double p);

#ifdef __cplusplus
}
#endif

#endif /* CQPI_AMD_H */

// gpuqpi_amd_shim.h -- GPUQPI plug-in for calcQPI's tmatrix that forwards to
// libcqpi_amd.so.  Compiled inside the calcQPI source tree (it uses the
// reference's GSL matrices, wannierfunctions and idl classes), selected in
// src/qpi.cc the way HipQPI is selected (src/qpi.cc:34-43, :290-321).
//
// Behaviour follows HipQPI (src/qpihipkernel.cc:395-704): one context per
// tmatrix, created at the first energy and reused; wannierldos copies G0
// before returning and runs the contraction asynchronously; retrieveResult
// synchronises.  Errors print the library's message and exit with its code,
// as gpuAssert does (src/qpihipkernel.cc:372-380).
#ifndef GPUQPI_AMD_SHIM_H
#define GPUQPI_AMD_SHIM_H

#include <cstdlib>
#include <iostream>
#include <vector>

#include "cqpi_amd.h"
#include "gpuqpi.h"
#include "idl.h"
#include "qpi.h"  // flaglist (src/qpi.h:33-39), as qpihipkernel.h includes it
#include "wannierfunctions.h"

class AmdQPI : public GPUQPI
{
    cqpi_ctx *ctx = nullptr;
    size_t W, K, bands;
    std::vector<const double *> table;  // g0[ix][iy]->data, built once (matrices are reused)
    std::vector<double> scatbuf, ldosbuf;

    static void check(int rc)
    {
        if (rc) {
            std::cerr << "CPU HIP error: " << cqpi_last_error() << std::endl;
            exit(rc);
        }
    }

public:
    AmdQPI(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp, size_t bands_,
           size_t maxband, bool spin, std::vector<wannierfunctions> &wf)
        : W(wanniern), K(kpoints), bands(bands_)
    {
        const size_t T = (2 * window + 1) * oversamp;
        std::vector<double> tiles(maxband * T * T, 0.0);  // zero tiles for padded "z" orbitals
        for (size_t o = 0; o < maxband && o < wf.size(); o++)
            if (!wf[o].iszero())
                for (size_t i = 0; i < T; i++)
                    for (size_t j = 0; j < T; j++) tiles[(o * T + i) * T + j] = wf[o].getwave_cached(i, j);
        check(cqpi_create(wanniern, kpoints, n, window, oversamp, bands, maxband, spin, tiles.data(), &ctx));
        scatbuf.resize(2 * bands * bands);
    }
    // the threshold-list constructor (src/qpihipkernel.cc:431-493): the lists
    // flattened in IDX4CC order, as HipQPI flattens them
    AmdQPI(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp, size_t bands_, flaglist &flist)
        : W(wanniern), K(kpoints), bands(bands_)
    {
        const size_t winn = 2 * window + 1;
        std::vector<cqpi_flagentry> flat;
        std::vector<size_t> counts;
        for (size_t i = 0; i < winn; i++)
            for (size_t j = 0; j < winn; j++)
                for (size_t k = 0; k < winn; k++)
                    for (size_t l = 0; l < winn; l++) {
                        counts.push_back(flist[i][j][k][l].size());
                        for (const flaglistentry &f : flist[i][j][k][l])
                            flat.push_back(cqpi_flagentry{(uint32_t)f.i, (uint32_t)f.j, (uint32_t)f.o1, (uint32_t)f.o2,
                                                          f.factor});
                    }
        check(cqpi_create_list(wanniern, kpoints, n, window, oversamp, bands, flat.data(), counts.data(), &ctx));
        scatbuf.resize(2 * bands * bands);
    }
    // the spectral-function constructor (src/qpihipkernel.cc:498-533): maps n x n
    AmdQPI(size_t kpoints, size_t n, size_t bands_, size_t maxbands, bool spin, std::vector<std::vector<double> > &pos,
           std::vector<double> &prearr)
        : W(n), K(kpoints), bands(bands_)
    {
        std::vector<double> xy;
        for (const std::vector<double> &p : pos) {
            xy.push_back(p.size() > 0 ? p[0] : 0.0);
            xy.push_back(p.size() > 1 ? p[1] : 0.0);
        }
        check(cqpi_create_spectral(kpoints, n, bands, maxbands, spin, xy.empty() ? nullptr : xy.data(), pos.size(),
                                   prearr.empty() ? nullptr : prearr.data(), prearr.size(), &ctx));
        scatbuf.resize(2 * bands * bands);
    }
    void printinfo(std::ostream &os)
    {
        char buf[2048];
        check(cqpi_printinfo(ctx, buf, sizeof buf));
        os << buf;
    }
    // the pointer table of g0 (built once: the matrices are reused)
    void table_of(std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        if (table.empty()) {
            table.resize(K * K);
            for (size_t ix = 0; ix < K; ix++)
                for (size_t iy = 0; iy < K; iy++) table[ix * K + iy] = g0[ix][iy]->data;
        }
    }
    // T packed with ld = bands, and the pointer table
    void pack(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        for (size_t i = 0; i < bands; i++)
            for (size_t j = 0; j < bands; j++) {
                scatbuf[2 * (i * bands + j)] = scat->data[2 * (i * scat->tda + j)];
                scatbuf[2 * (i * bands + j) + 1] = scat->data[2 * (i * scat->tda + j) + 1];
            }
        table_of(g0);
    }
    void wannierldos(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        pack(scat, g0);
        check(cqpi_wannierldos(ctx, scatbuf.data(), table.data(), g0[0][0]->tda));
    }
    void wannierjosephson(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0,
                          gsl_complex tip)
    {
        pack(scat, g0);
        check(cqpi_wannierjosephson(ctx, scatbuf.data(), table.data(), g0[0][0]->tda, GSL_REAL(tip), GSL_IMAG(tip)));
    }
    void wannierldoslist(gsl_matrix_complex *scat, std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        pack(scat, g0);
        check(cqpi_wannierldoslist(ctx, scatbuf.data(), table.data(), g0[0][0]->tda));
    }
    // g0 holds G0(k) in these two modes (the reference skips the transform)
    void spf(std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        table_of(g0);
        check(cqpi_spf(ctx, table.data(), g0[0][0]->tda));
    }
    void uspf(std::vector<std::vector<gsl_matrix_complex *> > &g0)
    {
        table_of(g0);
        check(cqpi_uspf(ctx, table.data(), g0[0][0]->tda));
    }
    void retrieveResult(double *ldos) { check(cqpi_retrieve(ctx, ldos)); }
    void retrieveResult(idl &map, size_t layer)
    {
        ldosbuf.resize(W * W);
        check(cqpi_retrieve(ctx, ldosbuf.data()));
        for (size_t i = 0; i < W; i++)
            for (size_t j = 0; j < W; j++) map.set(i, j, layer, ldosbuf[i + j * W]);
    }
    void retrieveResult(std::vector<std::vector<double> > &ldos)
    {
        ldosbuf.resize(W * W);
        check(cqpi_retrieve(ctx, ldosbuf.data()));
        ldos.assign(W, std::vector<double>(W));
        for (size_t i = 0; i < W; i++)
            for (size_t j = 0; j < W; j++) ldos[i][j] = ldosbuf[i + j * W];
    }
    ~AmdQPI() { cqpi_destroy(ctx); }
};

#endif

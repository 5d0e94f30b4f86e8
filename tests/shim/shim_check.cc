// Test driver for the GPUQPI plug-in shim (calcqpi-release_amd/host/
// gpuqpi_amd_shim.h): reads a case written by tests/test_gpu_shim.py, runs the
// shim's constructors and virtuals as the reference's tmatrix calls them
// (src/qpi.cc:290-321, :876-883, :1088-1091, :660-667; retrieveResult
// :605-606, :803-804, :822-823), and writes the maps back.
//
//   shim_check DIR
// DIR holds params.txt (W K n w os b maxband spin threshold), tiles.bin,
// g0.bin (G0(R), K x K x b x b complex), scat.bin (T), optional flags.bin +
// counts.bin (a threshold list) and g0k.bin (G0(k) for the SPF).  Output:
// out_ptr.bin, out_idl.bin, out_vec.bin (the three retrieveResult overloads,
// x + y W each), out_list.bin, out_spf.bin.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "gpuqpi_amd_shim.h"

static std::vector<double> load(const std::string &path, size_t n)
{
    std::vector<double> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f || !f.read(reinterpret_cast<char *>(v.data()), n * sizeof(double))) {
        std::cerr << "cannot read " << path << std::endl;
        exit(3);
    }
    return v;
}

static bool exists(const std::string &path) { return std::ifstream(path).good(); }

static void save(const std::string &path, const std::vector<double> &v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
}

// a K x K table of b x b matrices with leading dimension ld > b (as GSL
// allocates them inside larger blocks), filled from [ix][iy][r][c] complex
struct Grid {
    size_t K, b, ld;
    std::vector<double> store;
    std::vector<gsl_matrix_complex> mats;
    std::vector<std::vector<gsl_matrix_complex *> > g;
    Grid(size_t K_, size_t b_, const std::vector<double> &src) : K(K_), b(b_), ld(b_ + 1)
    {
        store.assign(K * K * b * ld * 2, 0.0);
        mats.resize(K * K);
        g.assign(K, std::vector<gsl_matrix_complex *>(K));
        for (size_t ix = 0; ix < K; ix++)
            for (size_t iy = 0; iy < K; iy++) {
                gsl_matrix_complex &m = mats[ix * K + iy];
                m.size1 = m.size2 = b;
                m.tda = ld;
                m.data = store.data() + (ix * K + iy) * b * ld * 2;
                g[ix][iy] = &m;
            }
        fill(src);
    }
    void fill(const std::vector<double> &src)
    {
        for (size_t k = 0; k < K * K; k++)
            for (size_t r = 0; r < b; r++)
                for (size_t c = 0; c < b; c++) {
                    mats[k].data[2 * (r * ld + c)] = src[2 * ((k * b + r) * b + c)];
                    mats[k].data[2 * (r * ld + c) + 1] = src[2 * ((k * b + r) * b + c) + 1];
                }
    }
    void poison() { std::fill(store.begin(), store.end(), std::numeric_limits<double>::quiet_NaN()); }
};

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: shim_check DIR" << std::endl;
        return 2;
    }
    const std::string d = std::string(argv[1]) + "/";
    size_t W, K, n, w, os, b, maxb;
    int spin;
    double threshold;
    {
        std::ifstream p(d + "params.txt");
        if (!(p >> W >> K >> n >> w >> os >> b >> maxb >> spin >> threshold)) {
            std::cerr << "bad params.txt" << std::endl;
            return 3;
        }
    }
    const size_t T = (2 * w + 1) * os;
    const std::vector<double> tiles = load(d + "tiles.bin", maxb * T * T);
    const std::vector<double> g0src = load(d + "g0.bin", K * K * b * b * 2);
    const std::vector<double> tsrc = load(d + "scat.bin", b * b * 2);

    // T as a GSL matrix with a leading dimension beyond b
    std::vector<double> tstore((b + 2) * b * 2, 0.0);
    gsl_matrix_complex scat{b, b, b + 2, tstore.data()};
    for (size_t r = 0; r < b; r++)
        for (size_t c = 0; c < b; c++) {
            tstore[2 * (r * (b + 2) + c)] = tsrc[2 * (r * b + c)];
            tstore[2 * (r * (b + 2) + c) + 1] = tsrc[2 * (r * b + c) + 1];
        }
    std::vector<wannierfunctions> wf;
    for (size_t o = 0; o < maxb; o++) {
        bool zero = true;
        for (size_t i = 0; i < T * T; i++) zero = zero && tiles[o * T * T + i] == 0.0;
        wf.emplace_back(T, tiles.data() + o * T * T, zero);
    }
    Grid g0(K, b, g0src);
    {
        GPUQPI *q = new AmdQPI(W, K, n, w, os, b, maxb, spin != 0, wf);
        q->printinfo(std::cout);
        std::cout << std::endl;
        q->wannierldos(&scat, g0.g);
        g0.poison();  // the caller overwrites g0 for the next energy once wannierldos returns
        std::vector<double> a(W * W);
        q->retrieveResult(a.data());
        idl map(W, W, 3);
        q->retrieveResult(map, 2);
        std::vector<std::vector<double> > v;
        q->retrieveResult(v);
        std::vector<double> bi(W * W), bv(W * W);
        for (size_t x = 0; x < W; x++)
            for (size_t y = 0; y < W; y++) {
                bi[x + y * W] = map.v[(2 * W + y) * W + x];
                bv[x + y * W] = v[x][y];
            }
        save(d + "out_ptr.bin", a);
        save(d + "out_idl.bin", bi);
        save(d + "out_vec.bin", bv);
        delete q;
    }
    g0.fill(g0src);
    if (exists(d + "counts.bin")) {
        const size_t wins = 2 * w + 1, npairs = wins * wins * wins * wins;
        const std::vector<double> cnt = load(d + "counts.bin", npairs);
        size_t total = 0;
        for (double c : cnt) total += (size_t)c;
        const std::vector<double> fl = load(d + "flags.bin", total * 5);  // i, j, o1, o2, factor
        flaglist flist(wins, std::vector<std::vector<std::vector<std::vector<flaglistentry> > > >(
                                 wins, std::vector<std::vector<std::vector<flaglistentry> > >(
                                           wins, std::vector<std::vector<flaglistentry> >(wins))));
        size_t e = 0, k = 0;
        for (size_t i = 0; i < wins; i++)
            for (size_t j = 0; j < wins; j++)
                for (size_t kk = 0; kk < wins; kk++)
                    for (size_t l = 0; l < wins; l++, k++)
                        for (size_t u = 0; u < (size_t)cnt[k]; u++, e++)
                            flist[i][j][kk][l].push_back(flaglistentry{(size_t)fl[5 * e], (size_t)fl[5 * e + 1],
                                                                       (size_t)fl[5 * e + 2], (size_t)fl[5 * e + 3],
                                                                       fl[5 * e + 4]});
        GPUQPI *q = new AmdQPI(W, K, n, w, os, b, flist);
        q->wannierldoslist(&scat, g0.g);
        std::vector<double> a(W * W);
        q->retrieveResult(a.data());
        save(d + "out_list.bin", a);
        delete q;
    }
    if (exists(d + "g0k.bin")) {
        Grid gk(K, b, load(d + "g0k.bin", K * K * b * b * 2));
        std::vector<std::vector<double> > pos;
        std::vector<double> prearr;
        GPUQPI *q = new AmdQPI(K, K, b, maxb, spin != 0, pos, prearr);
        q->spf(gk.g);
        std::vector<double> a(K * K);
        q->retrieveResult(a.data());
        save(d + "out_spf.bin", a);
        delete q;
    }
    std::cout << "shim_check done" << std::endl;
    return 0;
}

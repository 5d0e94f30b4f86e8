// Wannier tiles from Wannier90 / XCrySDen .xsf orbital files: the calcqpi
// driver's orbitalfiles= mode (src/calcqpi.cc:576-601), host code.
//
// The file's CRYSTAL block gives CONVVEC (lattice vectors, rows) and the
// atoms of PRIMCOORD; BEGIN_BLOCK_DATAGRID_3D / 3D_field /
// BEGIN_DATAGRID_3D_UNKNOWN holds nx ny nz, the origin, three spanning
// vectors and nx*ny*nz values, x fastest (xcfloader::readfile,
// src/loadxcf.cc:33-185).  A tile point (xu, yu) in lattice units is the
// Cartesian point xu a1 + yu a2 (a1, a2 the first two CONVVEC rows, x and y
// components) at height zh; it is written in the spanning vectors,
// p - origin = f1 s1 + f2 s2 + f3 s3, and takes the value of node
// (round(f1 nx), round(f2 ny), round(f3 nz)), or 0 outside the grid
// (xcfloader::getvalue and operator(), :220-236, :330-332).  zh is the
// highest atom's z plus height for height > 0, the lowest atom's z plus
// height for height < 0 (getzheight, :334-349).  The tile samples
// (i/oversamp - window, j/oversamp - window) (wannierfunctions(filename, ...),
// src/wannierfunctions.cc:444-461).
#include <cfloat>
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cqpi_amd.h"

namespace {

struct XsfGrid {
    size_t nx = 0, ny = 0, nz = 0;
    double origin[3] = {0, 0, 0};
    double span[3][3] = {{0}};  // span[a] = spanning vector a
    double conv[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::vector<double> atom_z;
    std::vector<double> v;  // v[i + nx (j + ny k)]
};

bool read3(std::istream &is, double *d)
{
    std::string line;
    if (!std::getline(is, line)) return false;
    std::istringstream ss(line);
    return (bool)(ss >> d[0] >> d[1] >> d[2]);
}

// returns an empty string on success, else what went wrong
std::string read_xsf(const char *path, XsfGrid &g)
{
    std::ifstream is(path);
    if (!is) return "could not open file";
    std::string line;
    bool crystal = false, grid = false;
    while (std::getline(is, line)) {
        if (line.find('#') != std::string::npos) continue;
        if (line.find("CRYSTAL") != std::string::npos) {
            crystal = true;
        } else if (crystal && line.find("CONVVEC") != std::string::npos) {
            for (int r = 0; r < 3; r++)
                if (!read3(is, g.conv[r])) return "bad CONVVEC";
        } else if (crystal && line.find("PRIMCOORD") != std::string::npos) {
            if (!std::getline(is, line)) return "bad PRIMCOORD";
            std::istringstream ss(line);
            size_t n = 0;
            ss >> n;
            for (size_t a = 0; a < n; a++) {
                if (!std::getline(is, line)) return "PRIMCOORD ends early";
                std::istringstream as(line);
                std::string name;
                double x, y, z;
                if (!(as >> name >> x >> y >> z)) return "bad atom line";
                g.atom_z.push_back(z);
            }
        } else if (line.find("BEGIN_BLOCK_DATAGRID_3D") != std::string::npos) {
            std::string kw;
            if (!std::getline(is, line) || !(std::istringstream(line) >> kw) || kw != "3D_field")
                return "unknown data type";
            if (!std::getline(is, line) || !(std::istringstream(line) >> kw) || kw != "BEGIN_DATAGRID_3D_UNKNOWN")
                return "unknown data grid identifier";
            if (!std::getline(is, line)) return "missing grid size";
            std::istringstream ss(line);
            if (!(ss >> g.nx >> g.ny >> g.nz) || !g.nx || !g.ny || !g.nz) return "bad grid size";
            if (!read3(is, g.origin)) return "bad origin";
            for (int a = 0; a < 3; a++)
                if (!read3(is, g.span[a])) return "bad spanning vector";
            const size_t n = g.nx * g.ny * g.nz;
            g.v.resize(n);
            for (size_t q = 0; q < n; q++)
                if (!(is >> g.v[q])) return "data grid ends early";
            std::string end;
            if (!(is >> end) || end != "END_DATAGRID_3D") return "missing END_DATAGRID_3D";
            if (!(is >> end) || end != "END_BLOCK_DATAGRID_3D") return "missing END_BLOCK_DATAGRID_3D";
            grid = true;
        }
    }
    return grid ? "" : "no 3D data grid";
}

// f with sum_a f[a] span[a] = d (Cramer's rule)
bool solve3(const double s[3][3], const double d[3], double f[3])
{
    auto det = [](const double *a, const double *b, const double *c) {
        return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
               a[2] * (b[0] * c[1] - b[1] * c[0]);
    };
    const double D = det(s[0], s[1], s[2]);
    if (D == 0.0) return false;
    f[0] = det(d, s[1], s[2]) / D;
    f[1] = det(s[0], d, s[2]) / D;
    f[2] = det(s[0], s[1], d) / D;
    return true;
}

double node_value(const XsfGrid &g, double x, double y, double z)
{
    const double d[3] = {x - g.origin[0], y - g.origin[1], z - g.origin[2]};
    double f[3];
    if (!solve3(g.span, d, f)) return 0.0;
    const double fi = std::round(f[0] * (double)g.nx), fj = std::round(f[1] * (double)g.ny),
                 fk = std::round(f[2] * (double)g.nz);
    if (!(fi >= 0 && fj >= 0 && fk >= 0 && fi < (double)g.nx && fj < (double)g.ny && fk < (double)g.nz)) return 0.0;
    return g.v[(size_t)fi + g.nx * ((size_t)fj + g.ny * (size_t)fk)];
}

}  // namespace

extern "C" int cqpi_xsf_tile(const char *path, double height, size_t window, size_t oversamp, double prefact,
                             double *tile)
{
    if (!path || !tile || !oversamp) return CQPI_E_ARG;
    XsfGrid g;
    if (!read_xsf(path, g).empty()) return CQPI_E_IO;
    double zh = NAN;
    if (height > 0.0) {
        double m = -DBL_MAX;
        for (double z : g.atom_z) m = std::fmax(m, z);
        zh = m + height;
    } else if (height < 0.0) {
        double m = DBL_MAX;
        for (double z : g.atom_z) m = std::fmin(m, z);
        zh = m + height;
    }
    const size_t T = (2 * window + 1) * oversamp;
    for (size_t i = 0; i < T; i++)
        for (size_t j = 0; j < T; j++) {
            const double xu = (double)i / (double)oversamp - (double)window;
            const double yu = (double)j / (double)oversamp - (double)window;
            const double x = xu * g.conv[0][0] + yu * g.conv[1][0], y = xu * g.conv[0][1] + yu * g.conv[1][1];
            tile[i * T + j] = prefact * node_value(g, x, y, zh);
        }
    return 0;
}

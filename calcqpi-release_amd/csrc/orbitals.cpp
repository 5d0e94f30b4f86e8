// Analytic atomic-like orbitals sampled on the Wannier tile grid: the
// calcqpi driver's orbitals= mode (src/calcqpi.cc:519-575), host code.
//
// An orbital is a real spherical harmonic Y(x/r, y/r, z/r) times the
// Slater-type radial part prefact * exp(-r / radius) (wannierfunctions::swave ...
// fxx2wave, src/wannierfunctions.cc:21-124).  It sits at (xpos, ypos) at
// height `height` below the tip plane: a tile point (x, y) is evaluated at
// (x - xpos, y - ypos, -height) (shift, src/wannierfunctions.h:40), turned by
// theta about z when theta != 0 (rotate with phi = 0, :34-38).  The tile
// samples x = i/oversamp - window - 1/2, i < (2 window + 1) oversamp
// (wannierfunctions::precalculate, src/wannierfunctions.cc:436-442).
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/cqpi_amd.h"

namespace {

enum Kind { kS, kPx, kPy, kPz, kDx2, kDxy, kDxz, kDyz, kDr2, kFy3x2, kFxyz, kFyz2, kFz3, kFxz2, kFzx2, kFxx2, kN, kZero };

struct Named {
    const char *name;
    Kind kind;
};

// the names settype accepts (src/wannierfunctions.cc:395-415)
constexpr Named kNames[] = {{"s", kS},         {"px", kPx},     {"py", kPy},     {"pz", kPz},     {"dx2", kDx2},
                            {"dxy", kDxy},     {"dxz", kDxz},   {"dyz", kDyz},   {"dr2", kDr2},   {"fy3x2", kFy3x2},
                            {"fxyz", kFxyz},   {"fyz2", kFyz2}, {"fz3", kFz3},   {"fxz2", kFxz2}, {"fzx2", kFzx2},
                            {"fxx2", kFxx2},   {"n", kN},       {"z", kZero}};

// angular factor of each real harmonic in the unit vector (u, v, w), with
// the normalisation sum over the 2l+1 members of Y^2 = (2l+1)/(4 pi)
double angular(Kind k, double u, double v, double w)
{
    const double inv4pi = 0.25 / M_PI;
    switch (k) {
    case kS: return std::sqrt(inv4pi);
    case kPx: return std::sqrt(3.0 * inv4pi) * u;
    case kPy: return std::sqrt(3.0 * inv4pi) * v;
    case kPz: return std::sqrt(3.0 * inv4pi) * w;
    case kDx2: return 0.5 * std::sqrt(15.0 * inv4pi) * (u * u - v * v);
    case kDxy: return std::sqrt(15.0 * inv4pi) * u * v;
    case kDxz: return std::sqrt(15.0 * inv4pi) * u * w;
    case kDyz: return std::sqrt(15.0 * inv4pi) * v * w;
    case kDr2: return 0.25 * std::sqrt(5.0 / M_PI) * (3.0 * w * w - 1.0);
    case kFy3x2: return 0.25 * std::sqrt(17.5 / M_PI) * v * (3.0 * u * u - v * v);
    case kFxyz: return std::sqrt(105.0 * inv4pi) * u * v * w;
    case kFyz2: return 0.25 * std::sqrt(10.5 / M_PI) * v * (5.0 * w * w - 1.0);
    case kFz3: return 0.25 * std::sqrt(7.0 / M_PI) * (5.0 * w * w * w - 3.0 * w);
    case kFxz2: return 0.25 * std::sqrt(10.5 / M_PI) * u * (5.0 * w * w - 1.0);
    case kFzx2: return 0.25 * std::sqrt(105.0 / M_PI) * w * (u * u - v * v);
    case kFxx2: return 0.25 * std::sqrt(17.5 / M_PI) * u * (u * u - 3.0 * v * v);
    default: return 0.0;
    }
}

double orbital_value(Kind k, double radius, double prefact, double symmetry, double x, double y, double z)
{
    if (k == kZero) return 0.0;
    const double r = std::sqrt(x * x + y * y + z * z);
    const double radial = prefact * std::exp(-r / radius);
    if (k == kN) {
        // n-fold symmetric lobe (nsymmwave, src/wannierfunctions.cc:116-120)
        const double c = std::cos(0.5 * symmetry * std::atan2(x, y));
        return radial * c * c * r / radius;
    }
    if (k == kS) return radial * angular(kS, 0.0, 0.0, 0.0);
    if (r == 0.0) return 0.0;
    return radial * angular(k, x / r, y / r, z / r);
}

}  // namespace

extern "C" int cqpi_orbital_tile(const char *name, double radius, double height, double xpos, double ypos,
                                 double theta, double prefact, size_t window, size_t oversamp, double *tile)
{
    if (!name || !tile || !oversamp || !(radius > 0.0)) return CQPI_E_ARG;
    int kind = -1;
    for (const Named &n : kNames)
        if (std::strcmp(name, n.name) == 0) kind = n.kind;
    if (kind < 0) return CQPI_E_ARG;
    const Kind k = (Kind)kind;
    const size_t T = (2 * window + 1) * oversamp;
    const double cs = std::cos(theta), sn = std::sin(theta);
    for (size_t i = 0; i < T; i++)
        for (size_t j = 0; j < T; j++) {
            double x = (double)i / (double)oversamp - (double)window - 0.5 - xpos;
            double y = (double)j / (double)oversamp - (double)window - 0.5 - ypos;
            const double z = -height;
            if (theta != 0.0 && k != kS) {
                const double xr = x * cs + y * sn;
                y = -x * sn + y * cs;
                x = xr;
            }
            tile[i * T + j] = orbital_value(k, radius, prefact, 4.0, x, y, z);
        }
    return 0;
}

extern "C" double cqpi_orbital_range(double radius, double p) { return radius * std::sqrt(std::log(1.0 / p)); }

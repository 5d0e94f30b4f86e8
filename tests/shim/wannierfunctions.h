// Test-only stand-in for calcQPI's wannierfunctions: a tile of T x T samples
// with the two queries the plug-in makes (iszero, getwave_cached).
#pragma once
#include <cstddef>
#include <vector>

class wannierfunctions
{
    size_t T;
    std::vector<double> tile;
    bool zero;

public:
    wannierfunctions(size_t t, const double *data, bool z) : T(t), tile(data, data + t * t), zero(z) {}
    bool iszero() const { return zero; }
    double getwave_cached(size_t i, size_t j) const { return tile[i * T + j]; }
};

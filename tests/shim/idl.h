// Test-only stand-in for calcQPI's idl map class: only set(x, y, layer, v),
// the call the plug-in's retrieveResult(idl&, layer) makes, stored as
// [layer][y][x].
#pragma once
#include <cstddef>
#include <vector>

class idl
{
public:
    size_t nx, ny, nl;
    std::vector<double> v;
    idl(size_t x, size_t y, size_t l) : nx(x), ny(y), nl(l), v(x * y * l, 0.0) {}
    void set(size_t x, size_t y, size_t layer, double val) { v[(layer * ny + y) * nx + x] = val; }
};

// Wannier90 _hr.dat reader of the library (host code).
//
// Follows tightbind::readfile (src/tightbinding.cc:70-149):
//   line 1: model name; line 2: num_wann; line 3: nrpts;
//   then nrpts degeneracies, as many per line as the file holds (:84-107);
//   then one "x y z u v Re Im" line per matrix element.  A new lattice vector
//   starts whenever (x,y,z) differs from the previous line (:121-127), so the
//   file must list each R's elements together; the amplitude is divided by the
//   degeneracy of its R (:134-135); orbitals are 1-based (:136-141); a term is
//   "anti" (conjugated phase) when both orbitals lie in the hole block of a
//   superconducting model (:137).
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cqpi_amd.h"

namespace cqpi {
int set_error(int code, const std::string &msg);
}

extern "C" int cqpi_read_hr(const char *path, int scmodel, cqpi_hopping **hops, size_t *nhops,
                            size_t *bands, char *name, size_t namelen)
{
    using cqpi::set_error;
    if (!path || !hops || !nhops || !bands) return set_error(CQPI_E_ARG, "NULL argument");
    std::ifstream is(path);
    if (!is) return set_error(CQPI_E_IO, std::string("cannot open ") + path);
    std::string line;
    size_t lineno = 0;
    auto next = [&](std::string &s) -> bool {
        if (!std::getline(is, s)) return false;
        lineno++;
        while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
        return true;
    };
    if (!next(line)) return set_error(CQPI_E_IO, "empty model file");
    if (name && namelen) {
        std::strncpy(name, line.c_str(), namelen - 1);
        name[namelen - 1] = 0;
    }
    long nb = 0, nrpts = 0;
    if (!next(line) || !(std::istringstream(line) >> nb) || nb <= 0)
        return set_error(CQPI_E_IO, "bad num_wann line");
    if (!next(line) || !(std::istringstream(line) >> nrpts) || nrpts < 0)
        return set_error(CQPI_E_IO, "bad nrpts line");
    std::vector<long> deg;
    deg.reserve(nrpts);
    while ((long)deg.size() < nrpts) {
        if (!next(line)) return set_error(CQPI_E_IO, "file ends inside the degeneracy list");
        const char *p = line.c_str();
        while (*p && (long)deg.size() < nrpts) {
            while (*p && !std::isdigit((unsigned char)*p)) p++;
            if (!*p) break;
            char *end;
            deg.push_back(std::strtol(p, &end, 10));
            p = end;
        }
    }
    for (long d : deg)
        if (d <= 0) return set_error(CQPI_E_IO, "non-positive degeneracy");
    size_t scmax = scmodel ? ((size_t)nb >> 1) : (size_t)nb;
    std::vector<cqpi_hopping> out;
    long k = 0;
    int ox = 0, oy = 0, oz = 0;
    while (next(line)) {
        if (line.empty()) continue;
        std::istringstream iss(line);
        int x, y, z;
        long u, v;
        double hr, hc;
        if (!(iss >> x >> y >> z >> u >> v >> hr >> hc))
            return set_error(CQPI_E_IO, "malformed hopping line " + std::to_string(lineno) + ": " + line);
        if (k == 0 || x != ox || y != oy || z != oz) {
            k++;
            ox = x;
            oy = y;
            oz = z;
        }
        if (k > nrpts)
            return set_error(CQPI_E_IO, "more lattice vectors than nrpts at line " + std::to_string(lineno));
        if (u <= 0 || v <= 0 || u > nb || v > nb)
            return set_error(CQPI_E_IO, "invalid orbital index at line " + std::to_string(lineno));
        cqpi_hopping h;
        h.rx = x;
        h.ry = y;
        h.rz = z;
        h.o1 = (int32_t)(u - 1);
        h.o2 = (int32_t)(v - 1);
        h.tre = hr / (double)deg[k - 1];
        h.tim = hc / (double)deg[k - 1];
        h.anti = ((size_t)(u - 1) >= scmax && (size_t)(v - 1) >= scmax) ? 1 : 0;
        h.reserved = 0;
        out.push_back(h);
    }
    cqpi_hopping *buf = (cqpi_hopping *)std::malloc(std::max<size_t>(out.size(), 1) * sizeof(cqpi_hopping));
    if (!buf) return set_error(CQPI_E_ARG, "out of host memory");
    if (!out.empty()) std::memcpy(buf, out.data(), out.size() * sizeof(cqpi_hopping));
    *hops = buf;
    *nhops = out.size();
    *bands = (size_t)nb;
    return 0;
}

// Geometry and launchers of the mixed-precision (split-fp16 MFMA) contraction.
#pragma once

#include <vector>

#include "cqpi_common.h"

struct MixGeom {
    int CB;     // cells per tile along y (CB*os <= 16 pixel rows)
    int nsets;  // spin blocks
    int nper;   // orbitals per spin block
    int NU;     // lattice rows touched by a tile: CB + 2w
    int NUg;    // groups of 8 rows
    int KG;     // K groups of 8 per spin block: (2w+1) * nper * NUg
    int KS;     // K steps of 32
    int NT;     // N tiles of 16 columns (4b columns: a and b, re/im)
    int NPL;    // planes per (a|b, hi|lo): b*b*2
    int LDW;    // alignment of a fragment's 8 halves (in halves)
    int OFF;    // plane row offset that realises that alignment
    int Ly;     // plane row length (halves)
};

// Element (row, col) of one fp16 plane of nw rows and Ly columns.  Columns
// come in groups of 8 and each group keeps its rows contiguous,
// [col / 8][row][col % 8], so that the contraction's staging of a strip of
// columns over a run of rows reads contiguous memory.
__host__ __device__ inline size_t plane_off(int row, int col, int nw)
{
    return ((size_t)(col >> 3) * nw + row) * 8 + (col & 7);
}

namespace cqpi {
MixGeom mix_geometry(const Geom &g);
size_t mix_plane_halves(const Geom &g, const MixGeom &mg);
std::vector<_Float16> mix_tile_fragments(const Geom &g, const MixGeom &mg, const double *tiles, double &sA);
// shapes served by the T-folded contraction, whose operand buffer P holds the
// FP32 window W32[e][sy][sx][pair] instead of the fp16 planes
bool mix_tfold(const Geom &g, const MixGeom &mg);
int launch_mix_prepare(const Geom &g, const MixGeom &mg, const double2 *G0w, int E, double *scale,
                       _Float16 *P, hipStream_t s, bool tfold = false);
// tips: device tip factors of the E energies for the Josephson map (bulk then
// holds the complex periodic table, E x os^2 double2); NULL: the LDOS map.
// cscale (LDOS maps, E doubles of device scratch, may be NULL): the kernels
// other than the T-folded one first fold T into the b planes (in place) and
// keep their scale there.
int launch_contract_mixed(const Geom &g, const MixGeom &mg, const _Float16 *P, const void *Afrag,
                          const double *scale, double sA, const double2 *T, const double *bulk, int E,
                          double *ldos, hipStream_t s, const double2 *tips = nullptr, bool tfold = false,
                          double *cscale = nullptr);
}  // namespace cqpi

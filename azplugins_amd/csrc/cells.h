// Cell-list helpers shared by the single-system and the batched neighbor
// list builds.  Both builds must bin and test pairs identically, so the
// arithmetic here is evaluated without contraction whatever the translation
// unit's flags are.
#pragma once

#include <algorithm>
#include <cmath>

#include "device_common.h"

namespace azp
{

struct CellGrid
    {
    int n[3];
    double w[3];
    };

// Cells of width >= r_list (with a 1e-9 relative margin) per axis, at most
// 1024 per axis and about 4 per particle (dilute systems).
inline CellGrid cell_grid_for(const double L[3], double rmax, uint32_t N)
    {
    CellGrid g;
    size_t ncell = 1;
    for (int k = 0; k < 3; ++k)
        {
        int n = (rmax > 0.0) ? (int)std::floor(L[k] / (rmax * (1.0 + 1e-9))) : 1;
        n = std::max(1, std::min(n, 1024));
        g.n[k] = n;
        g.w[k] = L[k] / n;
        ncell *= (size_t)n;
        }
    while (ncell > 4 * (size_t)std::max<uint32_t>(N, 64) + 64)
        {
        int kk = 0;
        for (int k = 1; k < 3; ++k)
            if (g.n[k] > g.n[kk])
                kk = k;
        g.n[kk] = std::max(1, g.n[kk] / 2);
        g.w[kk] = L[kk] / g.n[kk];
        ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
        }
    return g;
    }

// Cell index along one axis for x in [-L/2, L/2) on an n-cell grid.
__device__ __forceinline__ int cell_coord(double x, double L, double Linv, int n)
    {
#pragma clang fp contract(off)
    const double f = (x + 0.5 * L) * Linv;
    int c = (int)floor(f * (double)n);
    c %= n;
    if (c < 0)
        c += n;
    return c;
    }

// Type index of a particle, or ntypes when pos.w is not a valid type index.
__device__ __forceinline__ uint32_t checked_type(double w, uint32_t ntypes)
    {
    return (w >= 0.0 && w < (double)ntypes) ? (uint32_t)w : ntypes;
    }

// The neighbor cells of cell c along one axis (c-1, c, c+1 with periodic
// wrap), with duplicates removed when the grid is narrower than 3 cells.
// Returns how many distinct cells were written to out.
__device__ __forceinline__ int stencil_axis(int c, int n, int out[3])
    {
    int m = 0;
    for (int d = -1; d <= 1; ++d)
        {
        const int x = (c + d + n) % n;
        bool dup = false;
        for (int k = 0; k < m; ++k)
            dup |= (out[k] == x);
        if (!dup)
            out[m++] = x;
        }
    return m;
    }

} // namespace azp

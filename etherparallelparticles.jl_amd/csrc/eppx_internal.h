// Internal declarations shared by the host API (api.hip) and the kernels
// (kernels.hip) of libeppx.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eppx.h"

namespace eppx {

constexpr int kBlock = 256;
constexpr int kScanTile = 2048;  // elements per block in the cell-count scan
constexpr int kMaxTag = 32;

// Sticky device flags (OR-ed by kernels, read at the next waiting call).
enum : uint32_t {
    kFlagCellOverflow = 1u,       // a cell holds more than max_neighbour_number particles
    kFlagNeighbourOverflow = 2u,  // a particle has more than max_neighbour_number neighbours
    kFlagRankFallback = 4u        // a cell was too crowded for the in-cell ordering pass
};

// The domain's cell grid as the kernels need it (src/Class/Domain/Domain2D.jl).
struct Grid {
    float first_x, first_y, last_x, last_y;
    float gap_x_inv, gap_y_inv, gap_square;
    int32_t n_x, n_y, n_cells;
};

// One particle system's structure-of-arrays state; slot s holds reference
// row id[s].  After a search, slots [0, n_alive) are alive and sorted by
// (cell, id); dead particles follow.
struct Fields {
    float2 *pos, *dvel;
    float4 *dyn;  // (u, v, density, pressure): what a neighbour's pair term reads, one 16-byte record
    float *mass, *vol, *drho, *h, *wv, *wvrho;
    int32_t *tag, *mov, *id, *cell, *alive;
};

// Neighbour-walk inputs shared by the pair kernels.
struct Walk {
    Grid g;
    const int32_t *start;     // (n_cells + 2) exclusive scan of cell counts; slot ranges
    const float4 *geo;        // (x, y, h, m) at the last search, sorted order
    const int32_t *tag;       // tags, sorted order
    const int32_t *cell;
    const int32_t *n_alive;   // device scalar
    uint32_t adj[kMaxTag + 1];  // adj[t]: bit (u-1) set when tag t is adjacent to tag u
    int32_t criterion;
    float rmin2, rmax2;       // criterion accepts every dr2 < rmin2, rejects every dr2 > rmax2
    uint32_t tags_present;    // tag bits of the particles (a superset)
    int32_t check_all;        // the band is not empty, or some tag lies outside 1..kMaxTag
    uint32_t *fallback;       // blocks that walked global memory instead of LDS (counter)
};

struct Physics {
    int32_t fluid_tag, wall_tag, kernel, eos, visc;
    float dt, rho0, p0, c0, tait_b, gx, gy, mu, alpha, beta, av_c, coefficient;
};

inline Physics to_physics(const eppx_physics &p) {
    return Physics{p.fluid_tag, p.wall_tag, p.kernel, p.eos, p.visc, p.dt, p.rho0, p.p0, p.c0, p.tait_b,
                   p.gx, p.gy, p.mu, p.alpha, p.beta, p.av_c, p.coefficient};
}

// Column map of the reference tables (0-based, -1 absent) used by the
// upload / download conversion kernels.
struct Columns {
    int32_t n_capacity, n_int, n_float;
    int32_t Tag, IsMovable, PositionVec, VelocityVec, dVelocityVec, Mass, Volume, Density, dDensity, Pressure, H;
    int32_t SumWeight, SumWeightedDensity;
};

// Neighbour-column export targets (reference layout, column-major) for the
// download path; a null pointer skips that column group.
struct NeighbourExport {
    int32_t n_capacity, M;
    int32_t *nCount, *nIndex;          // int table columns (already offset)
    float *nRVec, *nR, *nW, *nDW, *nHInv;  // float table columns (already offset)
    int32_t kernel;                    // kernel of the last search, for nW / nDW / nHInv
    int32_t with_kernel;               // 1 when the last search ran EPPX_SEARCH_KERNEL_CONTINUITY
    int32_t fluid_tag, wall_tag;       // pair classes that carry nW / nDW / nHInv
};

// ------------------------------------------------------------ launchers
// All launchers enqueue on `stream` and return the hipError_t of the launch.
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);
hipError_t launch_reset_search(int32_t *counts, int32_t n_counts, unsigned long long *pairs_last, hipStream_t s);
hipError_t launch_key_count(Fields f, int32_t n, Grid g, int32_t *counts, int32_t *key, int32_t *tmpslot,
                            uint32_t *flags, int32_t max_per_cell, hipStream_t s);
hipError_t launch_scan(const int32_t *counts, int32_t n, int32_t *start, int32_t *block_sums, hipStream_t s);
hipError_t launch_fill_members(const Fields f, int32_t n, int32_t n_cells, const int32_t *key,
                               const int32_t *tmpslot, const int32_t *start, int32_t *members, hipStream_t s);
hipError_t launch_permute(Fields cur, Fields nxt, float4 *geo, int32_t n, int32_t cap, int32_t n_cells,
                          const int32_t *key, const int32_t *tmpslot, const int32_t *start,
                          const int32_t *members, int32_t *n_alive, uint32_t *flags, hipStream_t s);
hipError_t launch_search(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph, int32_t fast,
                         int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                         unsigned long long *pairs_total, uint32_t *flags, hipStream_t s);
hipError_t launch_snapshot_mass(float4 *geo, Fields f, int32_t n, hipStream_t s);
hipError_t launch_self(Fields f, int32_t n, int32_t action, const Physics &ph, const eppx_field_op &op,
                       hipStream_t s);
hipError_t launch_interaction(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph,
                              const eppx_field_op &op, int32_t fast, hipStream_t s);
hipError_t launch_export(const Walk &w, Fields f, int32_t n, const int64_t *row_start, int32_t *js,
                         const int32_t *ncount, NeighbourExport ex, hipStream_t s);
hipError_t launch_from_ref(Fields f, int32_t cap, Columns c, const int32_t *ip, const float *fp,
                           const int32_t *is_alive, const int32_t *cell_index, hipStream_t s);
hipError_t launch_to_ref(Fields f, int32_t cap, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                         int32_t *cell_index, hipStream_t s);
hipError_t launch_stencil(Grid g, int32_t *count, int32_t *list, hipStream_t s);
hipError_t launch_finite(Fields f, int32_t n, uint32_t *bad, hipStream_t s);
hipError_t launch_h_range(Fields f, int32_t n, float *hmin_hmax, hipStream_t s);

}  // namespace eppx

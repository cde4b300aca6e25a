// Internal declarations of the 3D path of libeppx (api3d.hip, kernels3d.hip).
//
// The reference leaves every 3D layer a stub (AbstractDomain.jl:119,
// NeighbourSystemBase.jl:50-57, NeighbourSearch.jl:49 and :229,
// Geometry3D.jl:89-94), while its SPH library is written for any dimension
// (the loops over 0:(N - 1) of Library.jl:21-34, Pressure.jl, Viscosity.jl,
// Motion.jl, and the 3D sGravity! of State.jl:49-62).  The 3D path extends
// the 2D one axis by axis: the cell grid of Domain2D with a z axis (cell
// c = i + n_x (j - 1) + n_x n_y (k - 1), 1-based), the 27-cell stencil in
// the 2D visit order extended (x offset outer, then y, then z), and the same
// criterion, insert order and pair arithmetic.
#pragma once

#include "eppx_internal.h"

namespace eppx {

// The 3D cell grid as the kernels need it.
struct Grid3 {
    float first_x, first_y, first_z, last_x, last_y, last_z;
    float gap_x_inv, gap_y_inv, gap_z_inv, gap_square;
    int32_t n_x, n_y, n_z, n_cells;
};

// Structure-of-arrays state of a 3D particle system; slot s holds reference
// row id[s].  After a search, slots [0, n_alive) are alive and sorted by
// (cell, id); dead particles follow.  Every float4 array holds 64 slots more
// than the capacity (the candidate tests read up to three slots past a cell).
struct Fields3 {
    float4 *pos;   // (x, y, z, Tag as int bits)
    float4 *vel;   // (u, v, w, Density)
    float4 *hmvp;  // (H, Mass, Volume, Pressure)
    float4 *dvel;  // (du, dv, dw, 0)
    float *drho;
    int32_t *mov, *id, *cell, *alive;
};

struct Physics3 {
    Physics p;
    float gz;
};

// Neighbour-walk inputs of the 3D pair kernels.
struct Walk3 {
    Grid3 g;
    const int32_t *start;        // exclusive scan of the cell counts (n_cells + 2 entries + total)
    const uint32_t *cell_tags;   // tag bits of each cell's alive particles
    const float4 *geo;           // (x, y, z, Tag bits) at the last search, sorted order
    const float4 *vel, *hmvp;    // current records, sorted order
    const int32_t *cell;         // 1-based cell of each sorted slot
    const int32_t *n_alive;      // device scalars: [0] alive slots, [1] kept slots
    uint32_t adj[kMaxTag + 1];   // adj[t]: bit (u-1) set when tag t is adjacent to tag u
    uint32_t tags_present;       // tag bits of the uploaded particles
    int32_t odd_tags;            // some particle has a tag outside 1..kMaxTag (adjacent to nothing)
    int32_t criterion;
    float rmin2, rmax2;          // the criterion accepts every dr2 < rmin2, rejects every dr2 > rmax2
    float hinv_u;                // > 0: every alive particle has one H, and this is its pair hinv
    float m_u;                   // > 0: every alive particle has this one Mass
    int32_t stripe;              // groups of 4 cells per stripe of the pair-kernel grid
    int32_t stripes_per_xcd;     // stripes each of the 8 XCDs works through
    // Accept masks of the last search, one 64-bit mask per (stencil cell s, slot i) at
    // masks[s * mask_stride + i], for neighbour cells of at most 64 particles.  mask_mode:
    // 0 the pair kernel tests every candidate; 1 it tests them and stores the masks (the
    // search); 2 it reads the stored masks instead (momentum after that search).
    unsigned long long *masks;
    int64_t mask_stride;
    int32_t mask_mode;
};

// Neighbour-column export of the 3D download (nRVec holds 3 floats per pair).
using NeighbourExport3 = NeighbourExport;

hipError_t launch3_from_ref(Fields3 f, Columns c, const int32_t *ip, const float *fp, const int32_t *is_alive,
                            const int32_t *cell_index, hipStream_t s);
hipError_t launch3_to_ref(Fields3 f, int32_t n, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                          int32_t *cell_index, hipStream_t s);
// Cell keys (0-based cell of alive particles, n_cells for dead ones, which
// die outside the domain) and per-cell counts with each slot's provisional
// place in its cell.
hipError_t launch3_key_count(Fields3 f, int32_t n, Grid3 g, int32_t *counts, int32_t *key, int32_t *tmpslot,
                             uint32_t *flags, int32_t max_per_cell, hipStream_t s);
hipError_t launch3_fill_members(Fields3 f, int32_t n, int32_t n_cells, const int32_t *key, const int32_t *tmpslot,
                                const int32_t *start, int32_t *members, hipStream_t s);
// Moves every slot to its sorted place in nxt (cell start + members of the
// cell with a smaller particle index) and writes the snapshot geo.
hipError_t launch3_permute(Fields3 cur, Fields3 nxt, float4 *geo, int32_t n, int32_t n_cells, const int32_t *key,
                           const int32_t *tmpslot, const int32_t *start, const int32_t *members, int32_t *n_alive,
                           uint32_t *flags, hipStream_t s);
hipError_t launch3_cell_tags(const float4 *geo, int32_t n_cells, const int32_t *start, uint32_t *cell_tags,
                             hipStream_t s);
// Neighbour search pass (counts per slot into ncount, pair totals into
// the slots at kPairLast (this search) and kPairTotal (running)) with the fused action.
hipError_t launch3_search(const Walk3 &w, Fields3 f, int32_t action, const Physics &ph, int32_t fast, int32_t M,
                          int32_t *ncount, unsigned long long *pairs, uint32_t *flags, int32_t n_slots,
                          hipStream_t s);
hipError_t launch3_momentum(const Walk3 &w, Fields3 f, const Physics &ph, int32_t fast, int32_t n_slots,
                            hipStream_t s);
hipError_t launch3_self(Fields3 f, int32_t n, int32_t action, const Physics3 &ph, hipStream_t s);
hipError_t launch3_export(const Walk3 &w, Fields3 f, int32_t n_slots, const int64_t *row_start, int32_t *js,
                          const int32_t *ncount, NeighbourExport3 ex, hipStream_t s);

}  // namespace eppx

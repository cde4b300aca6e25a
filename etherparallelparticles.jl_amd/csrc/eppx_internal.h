// Internal declarations shared by the host API (api.hip) and the kernels
// (kernels.hip) of libeppx.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eppx.h"

namespace eppx {

constexpr int kBlock = 256;
constexpr int kScanTile = 1024;  // elements per block in the cell-count scan (735 blocks for 752k cells)
constexpr int kMaxTag = 32;

// Pair counters.  A search adds the pairs of each block (each wave in 3D) to
// one of kPairSlots slots chosen by the block (wave) index, each slot on a
// 128-byte line of its own: adds of thousands of blocks to one address queue
// one after another at its memory channel (on the 2,030,018-particle search
// that queue cost 9 us at 512 threads per block and 90 us at 256).  The
// count of the last search is the sum of the slots at kPairLast, the running
// total the sum of those at kPairTotal; kPairWords 64-bit words in all.
constexpr int kPairSlots = 256;
constexpr int kPairStride = 16;  // 64-bit words per slot
constexpr int kPairLast = 0;
constexpr int kPairTotal = kPairSlots * kPairStride;
constexpr int kPairWords = 2 * kPairSlots * kPairStride;

// Pair-kernel block headers: kStageInfo ints of range bounds, then a table of
// 16-bit cell starts, 6 rows of kStartRow256 / kStartRow512 entries (as ints).
// Int kHeaderTable of the bounds: 1 when the search staged the block (the table is
// then written wherever it fits).
constexpr int kStageInfo = 24;
constexpr int kHeaderTable = 21;
constexpr int kStartRow256 = 40, kStartRow512 = 112;
constexpr int kHeaderInts256 = kStageInfo + 3 * kStartRow256;
constexpr int kHeaderInts512 = kStageInfo + 3 * kStartRow512;

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

// One case (an independent particle system) of a set of device buffers.  A
// particle system holds one case; a batch holds many, side by side in one
// segmented cell list: case c owns the global cells
// [cell_base, cell_base + g.n_cells), and its particles sort before those of
// every later case.
struct CaseDesc {
    Grid g;
    int32_t cell_base;
    float rmin2, rmax2;     // criterion accepts every dr2 < rmin2, rejects every dr2 > rmax2
    uint32_t tags_present;  // tag bits of the case's particles (a superset)
    int32_t check_all;      // the band is not empty, or some tag lies outside 1..kMaxTag
    float hinv_u;           // > 0: every alive particle has one H, and this is (H + H) / (2 H H); 0: H varies
    float m_u;              // > 0: every alive particle has this one Mass (with hinv_u > 0: the FAST terms
                            // hoist it out of the pair sums); 0: Mass varies
    float dt;               // batch: the case's time step (a particle system uses the call's physics)
    int32_t steps_left;     // > 0: the case advances; a batch step decrements it
    int32_t fresh;          // batch: the case waits for its first search
    int32_t removed;        // batch: the case's slots are released at the next cell build
};

// Structure-of-arrays particle state; slot s holds row id[s] of case cs[s].
// After a search, slots [0, n_alive) are alive and sorted by (global cell,
// id); dead particles follow, then (batch) released ones.
struct Fields {
    float2 *pos, *dvel;
    float4 *dyn;  // (u, v, density, pressure): what a neighbour's pair term reads, one 16-byte record
    float *mass, *vol, *drho, *h, *wv, *wvrho, *wvp;  // wvp: SumWeightedPressure
    int32_t *tag, *mov, *id, *cell, *alive, *cs;
};

// Neighbour-walk inputs shared by the pair kernels.  A particle system
// passes its one case by value (one, cases == nullptr); a batch passes its
// device case table and each slot's case.
struct Walk {
    CaseDesc one;
    const CaseDesc *cases;
    const int32_t *start;     // exclusive scan of the global cell counts; slot ranges
    const float4 *geo;        // (x, y, h, m) at the last search, sorted order
    const int32_t *tag;       // tags, sorted order
    const int32_t *cell;      // cell within the particle's case (1-based), sorted order
    const int32_t *cs;        // case of each slot, sorted order
    const int32_t *n_alive;   // device scalars: [0] alive slots, [1] kept slots
    int32_t cap;              // slots allocated in every per-slot array
    uint32_t adj[kMaxTag + 1];  // adj[t]: bit (u-1) set when tag t is adjacent to tag u
    int32_t criterion;
    uint32_t *fallback;       // counters: [0] blocks that read global memory instead of LDS,
                              // [1] particle walks whose neighbour rows exceed 64 slots
    // The search's accept masks of the three neighbour rows of each slot
    // (masks[r * mask_stride + i]): mask_mode 1 stores them, 2 reads them
    // instead of testing the candidates again (same pair set), 0 neither.
    unsigned long long *masks;
    int32_t mask_stride, mask_mode;
    // Block headers of the pair kernels (a block's range bounds and the cell
    // starts of its rows, as its tile stager found them), one buffer per
    // block size (256 and 512 threads): header_mode 1 writes them (a search),
    // 2 reads them instead of finding them again (a momentum pass on the same
    // sorted layout, with the search's block size), 0 neither.
    int32_t *headers256, *headers512;
    int32_t header_mode;
    // 1: dDensity is zero on every alive slot (a step's searches after its
    // reset): the continuity sum starts at zero and the fused EOS search
    // leaves the zero in place
    int32_t drho_zero;
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
    int32_t SumWeight, SumWeightedDensity, SumWeightedPressure;
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
// Cell keys: cell_base + cell - 1 of the particle's case (`one`, or
// cases[cs[s]] when cases is given) for alive particles, n_cells (all cases)
// for dead ones, n_cells + 1 for released ones.
hipError_t launch_key_count(Fields f, int32_t n, const CaseDesc &one, const CaseDesc *cases, int32_t n_cells,
                            int32_t *counts, int32_t *key, int32_t *tmpslot, uint32_t *flags, int32_t max_per_cell,
                            hipStream_t s);
hipError_t launch_scan(const int32_t *counts, int32_t n, int32_t *start, int32_t *block_sums, hipStream_t s);
hipError_t launch_fill_members(const Fields f, int32_t n, int32_t n_cells, const int32_t *key,
                               const int32_t *tmpslot, const int32_t *start, int32_t *members, hipStream_t s);
// opts: kPermuteSums moves SumWeight, SumWeightedDensity and
// SumWeightedPressure (without it they are zero in both buffers and stay);
// kPermuteNoDrho / kPermuteNoDvel leave dDensity / dVelocity of the alive
// slots alone (zero there in both buffers); kPermuteOneMassH: every slot of both buffers
// holds Mass mass_one and H h_one (the same bits), so neither is read or moved and the
// snapshot takes them.  dyn_src / vol_src (null: cur's) hold the slots'
// (u, v, rho, p) records and volumes when a fused kernel wrote them aside.
constexpr int32_t kPermuteSums = 1, kPermuteNoDrho = 2, kPermuteNoDvel = 4, kPermuteOneMassH = 8;
hipError_t launch_permute(Fields cur, Fields nxt, float4 *geo, int32_t n, int32_t batch, int32_t n_cells,
                          const int32_t *key, const int32_t *tmpslot, const int32_t *start,
                          const int32_t *members, int32_t *n_alive, uint32_t *flags, int32_t opts,
                          const float4 *dyn_src, const float *vol_src, float mass_one, float h_one,
                          hipStream_t s);
// which = 0: the action runs for particles of cases with steps left; 1: of fresh cases.
// block: threads per pair block (256 or 512); stage: 1 when pass 2 reads
// its per-neighbour data from the block's LDS tile.  EXACT math uses 256, 0.
hipError_t launch_search(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph, int32_t fast,
                         int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                         unsigned long long *pairs_total, uint32_t *flags, int32_t which, int32_t block,
                         int32_t stage, hipStream_t s);
hipError_t launch_snapshot_mass(float4 *geo, Fields f, int32_t n, hipStream_t s);
// cases == nullptr: every alive particle, dt from ph; otherwise particles of
// cases with steps left, dt from their case.
hipError_t launch_self(Fields f, int32_t n, int32_t action, const Physics &ph, const eppx_field_op &op,
                       const CaseDesc *cases, hipStream_t s);
hipError_t launch_tick(CaseDesc *cases, int32_t n_cases, hipStream_t s);
// Fused kernels of the time step: continuity + EOS with the cell counts of the
// coming search zeroed; gravity + move with that search's cell key and count.
hipError_t launch_eos_reset(Fields f, int32_t n, const Physics &ph, const CaseDesc *cases, int32_t *counts,
                            int32_t n_counts, unsigned long long *pairs_last, hipStream_t s);
hipError_t launch_move_key(Fields f, int32_t n, const Physics &ph, const CaseDesc &one, const CaseDesc *cases,
                           int32_t n_cells, int32_t *counts, int32_t *key, int32_t *tmpslot, uint32_t *flags,
                           int32_t max_per_cell, hipStream_t s);
hipError_t launch_set_case(CaseDesc *d, const CaseDesc &v, hipStream_t s);

// Fused time step of a particle system (256 / 512 threads, gathered shapes).
// Momentum with gravity + accelerate-move and the coming search's cell keys
// and counts in its epilogue: the moved (u, v, rho, p) records go to dyn_new
// (the neighbours still read the old ones), the positions, dvel, cell and
// alive in place; pairs_last is zeroed.
struct MoveKey {
    int32_t n_slots, n_cells, max_per_cell;
    int32_t *counts, *key, *tmpslot;
    uint32_t *flags;
    float4 *dyn_new;
    unsigned long long *pairs_last;
    // 1: dVelocity is zero on every alive slot (a step after the first of a
    // call): the sum starts at zero and the move leaves the zero in place
    int32_t dvel_zero;
};
hipError_t launch_momentum_move(const Walk &w, Fields f, int32_t n, const Physics &ph, const MoveKey &mk,
                                int32_t fast, int32_t block, hipStream_t s);
// The search with the next step's sContinuity! in its epilogue: rho, p and
// dDensity in place, the volumes (of every slot) to vol_new (the neighbours
// still read the old ones), and the cell counts zeroed for the next step.
struct EosOut {
    int32_t n_slots, n_counts;
    float *vol_new;
    int32_t *counts;
};
hipError_t launch_search_eos(const Walk &w, Fields f, int32_t n, const Physics &ph, const EosOut &eo, int32_t fast,
                             int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                             unsigned long long *pairs_total, uint32_t *flags, int32_t block, hipStream_t s);
// The pair-kernel instantiations, one translation unit per math mode and
// kernel function (search.hip, momentum.hip compiled with EPPX_TU_FAST and
// EPPX_TU_KERNEL).  eos non-null: the fused next-step continuity (particle
// systems); mk non-null: the fused move and cell keys (particle systems).
struct PairLaunch {
    int32_t n, M;
    int32_t *ncount;
    unsigned long long *pairs_last, *pairs_total;
    uint32_t *flags;
    int32_t which, block, stage;
    const EosOut *eos;
};
template <bool FAST, int K>
hipError_t launch_search_pair(const Walk &w, Fields f, const Physics &ph, const PairLaunch &pl, hipStream_t s);
template <bool FAST, int K>
hipError_t launch_momentum_pair(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t block, int32_t stage,
                                const MoveKey *mk, hipStream_t s);
hipError_t launch_interaction(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph,
                              const eppx_field_op &op, int32_t fast, int32_t block, int32_t stage, hipStream_t s);
hipError_t launch_export(const Walk &w, Fields f, int32_t n, const int64_t *row_start, int32_t *js,
                         const int32_t *ncount, NeighbourExport ex, hipStream_t s);
// Rows [0, c.n_capacity) of one case's reference tables into slots
// [offset, offset + c.n_capacity), as case `cs`.
hipError_t launch_from_ref(Fields f, int32_t offset, Columns c, const int32_t *ip, const float *fp,
                           const int32_t *is_alive, const int32_t *cell_index, int32_t cs, hipStream_t s);
// Slots [0, n) of case `cs` (every slot when cs < 0) back into reference rows id[s].
hipError_t launch_to_ref(Fields f, int32_t n, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                         int32_t *cell_index, int32_t cs, hipStream_t s);
hipError_t launch_stencil(Grid g, int32_t *count, int32_t *list, hipStream_t s);
hipError_t launch_finite(Fields f, int32_t n, uint32_t *bad, hipStream_t s);
hipError_t launch_valu_calibration(float *out, int32_t blocks, int32_t iters, hipStream_t s);
hipError_t launch_h_range(Fields f, int32_t n, float *hmin_hmax, hipStream_t s);

}  // namespace eppx

// Host side of libeppx: the C ABI declared in include/eppx.h.
//
// A context owns one HIP stream on one device.  A particle system owns its
// structure-of-arrays state twice (the cell-list build moves every particle
// from one copy to the other in cell order), the positions snapshot of the
// last search and the cell arrays of that search.  Everything is enqueued on
// the context's stream; waiting calls synchronise that stream only.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"

using namespace eppx;
using namespace eppx::host;


struct eppx_ns {
    eppx_ctx *ctx = nullptr;
    eppx_domain2d domain{};
    Grid grid{};
    uint32_t adj[kMaxTag + 1] = {};
    int32_t max_tag = 0;
    int32_t M = 0;
    uint64_t serial = 0;  // unique per created system (graphs recorded against it check it)
};

// One time step recorded as a hipGraph (eppx_step).  The step moves the
// state from buffer `cur` to the other one, so each starting parity has its
// own graph; kernel arguments are recorded by value, so a graph is replayed
// only for the same neighbour system, physics and math mode, and dropped when
// the particle data are uploaded again.
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    uint64_t ns_serial = 0;
    Physics ph{};
    int32_t math = -1;
};

struct eppx_ps {
    eppx_ctx *ctx = nullptr;
    int32_t cap = 0, n = 0;
    eppx_layout layout{};
    Fields buf[2]{};
    int cur = 0;
    float4 *geo = nullptr;  // (x, y, h, m) snapshot of the last search, sorted order
    // cell arrays of the last search
    int32_t n_cells = -1;
    int32_t *counts = nullptr, *start = nullptr, *tile_sums = nullptr;
    int32_t *key = nullptr, *tmpslot = nullptr, *members = nullptr, *ncount = nullptr, *n_alive = nullptr;
    unsigned long long *pairs = nullptr;  // pair counters: kPairWords words (kPairLast, kPairTotal slots)
    uint32_t *fallback = nullptr;         // pair blocks that walked global memory
    float *h_range = nullptr;
    float rmin2 = -1.0f, rmax2 = INFINITY;  // symmetry-criterion bounds from the H range
    float hinv_u = 0.0f;                    // one H for every alive particle: its pair hinv; else 0
    float m_u = 0.0f;                       // one Mass for every alive particle: that mass; else 0
    // every slot of both buffers holds Mass mass_one and H h_one (the same bits): the permute moves neither
    bool mh_one = false;
    float mass_one = 0.0f, h_one = 0.0f;
    uint32_t tags_present = 0xffffffffu;    // tag bits of the uploaded particles
    bool odd_tags = true;                   // some tag outside 1..kMaxTag
    // search state
    eppx_ns *last_ns = nullptr;
    int32_t last_crit = 0, last_action = 0, last_kernel = 0, last_fluid = 0, last_wall = 0;
    bool searched = false;
    // recorded steps, per buffer parity and per kind of step (StepKind)
    StepGraph graphs[2][8];
    // the fused step's records and volumes written aside (allocated on first use)
    float4 *dyn_new = nullptr;
    float *vol_new = nullptr;
    // pair-kernel shapes (eppx_autotune picks them): threads per block, and
    // whether pass 2 reads its per-neighbour data from the LDS tile
    int32_t search_block = 256, momentum_block = 256, search_stage = 0, momentum_stage = 0;
    unsigned long long *masks = nullptr;  // the last search's accept masks (3 per slot)
    bool masks_valid = false;             // they describe every alive slot's pairs
    // the pair kernels' block headers (256- and 512-thread blocks) and which
    // of them the searches on the current layout wrote (bit 0: 256, bit 1: 512)
    int32_t *headers256 = nullptr, *headers512 = nullptr;
    int32_t header_blocks = 0;
    // SumWeight / SumWeightedDensity / SumWeightedPressure may be non-zero in
    // some buffer (an upload brought non-zero values, or a filter or
    // extrapolation interaction ran); while false the permute skips them
    bool accum_live = false;
    // dVelocity of the alive particles: -2 zero on all of them, t >= 0 zero
    // except on particles of tag t (the fluid of the momentum passes since),
    // -1 unknown
    int32_t dvel_nonzero_tag = -1;
    // columns the kernels do not own, kept on the host between upload and download
    std::vector<int32_t> pass_int_cols, pass_float_cols;
    std::vector<int32_t> pass_int;
    std::vector<float> pass_float;
    Fields &f() { return buf[cur]; }
};

// Fixed device scratch of eppx_ps_download_noalloc: the staging of the
// reference tables, for callers that cannot allocate device memory.  Every
// device the library runs on holds one copy for the life of the process.
// In host code `scratch` names a host-side stand-in, not a device address:
// the address of each device's copy is taken once per device, with that
// device current (eppx_ctx_create), and only that address reaches kernels
// and copies.  Calls on one device share it, so overlapping calls race.
constexpr size_t kScratchBytes = (size_t)128 << 20;  // 128 MiB
constexpr size_t kScratchWords = kScratchBytes / sizeof(uint32_t);
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_tried[kMaxDevices] = {};

// Takes the device address of `scratch` on `device` (current) once; returns
// the hipError_t of hipGetSymbolAddress on the first call for the device.
hipError_t init_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_tried[device]) return hipSuccess;
    g_scratch_tried[device] = true;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = static_cast<uint32_t *>(p);
    return e;
}

uint32_t *scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace

namespace {



hipError_t alloc_fields(Fields &f, int32_t cap) {
    hipError_t e = hipSuccess;
    auto a = [&](auto **p) {
        if (e == hipSuccess) e = dalloc(p, (size_t)cap);
    };
    a(&f.pos); a(&f.dvel); a(&f.dyn);
    a(&f.mass); a(&f.vol); a(&f.drho); a(&f.h); a(&f.wv); a(&f.wvrho); a(&f.wvp);
    a(&f.tag); a(&f.mov); a(&f.id); a(&f.cell); a(&f.alive); a(&f.cs);
    return e;
}

void free_fields(Fields &f) {
    dfree(f.pos); dfree(f.dvel); dfree(f.dyn);
    dfree(f.mass); dfree(f.vol); dfree(f.drho); dfree(f.h); dfree(f.wv); dfree(f.wvrho); dfree(f.wvp);
    dfree(f.tag); dfree(f.mov); dfree(f.id); dfree(f.cell); dfree(f.alive); dfree(f.cs);
}

bool owned_float(const eppx_layout &L, int c) {
    auto in = [&](int start, int width) { return start >= 0 && c >= start && c < start + width; };
    return in(L.PositionVec, 2) || in(L.VelocityVec, 2) || in(L.dVelocityVec, 2) || in(L.Mass, 1) ||
           in(L.Volume, 1) || in(L.Density, 1) || in(L.dDensity, 1) || in(L.Pressure, 1) || in(L.H, 1) ||
           in(L.SumWeight, 1) || in(L.SumWeightedDensity, 1) || in(L.SumWeightedPressure, 1) ||
           in(L.nRVec, 2 * L.max_neighbour_number) ||
           in(L.nR, L.max_neighbour_number) || in(L.nW, L.max_neighbour_number) ||
           in(L.nDW, L.max_neighbour_number) || in(L.nHInv, L.max_neighbour_number);
}

bool owned_int(const eppx_layout &L, int c) {
    auto in = [&](int start, int width) { return start >= 0 && c >= start && c < start + width; };
    return in(L.Tag, 1) || in(L.IsMovable, 1) || in(L.nCount, 1) || in(L.nIndex, L.max_neighbour_number);
}

Columns columns_of(const eppx_ps *ps) {
    const eppx_layout &L = ps->layout;
    return Columns{ps->cap, L.n_int, L.n_float, L.Tag, L.IsMovable, L.PositionVec, L.VelocityVec, L.dVelocityVec,
                   L.Mass, L.Volume, L.Density, L.dDensity, L.Pressure, L.H, L.SumWeight, L.SumWeightedDensity,
                   L.SumWeightedPressure};
}


int32_t ensure_cells(eppx_ps *ps, int32_t n_cells) {
    if (ps->n_cells == n_cells) return EPPX_OK;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ps->counts);
    dfree(ps->start);
    dfree(ps->tile_sums);
    const int64_t nb = (int64_t)n_cells + 2;  // cells + dead bucket + released bucket
    HIP_OK(ctx, dalloc(&ps->counts, (size_t)nb));
    HIP_OK(ctx, dalloc(&ps->start, (size_t)nb + 1));
    HIP_OK(ctx, dalloc(&ps->tile_sums, (size_t)((nb + kScanTile - 1) / kScanTile)));
    // empty cells until the next cell build: a walk before it finds no neighbours
    // instead of reading uninitialised ranges
    HIP_OK(ctx, hipMemsetAsync(ps->start, 0, ((size_t)nb + 1) * sizeof(int32_t), ctx->stream));
    ps->n_cells = n_cells;
    ps->masks_valid = false;
    return EPPX_OK;
}


// Criterion bounds: the criterion accepts every dr2 < rmin2 and rejects
// every dr2 > rmax2.  defaultSearchCriterion: dr2 <= gap_square, i.e.
// dr2 < the next float up, with an empty band; symmetry: the bounds the H
// range gives.
void criterion_bounds(int32_t criterion, const Grid &g, float sym_min2, float sym_max2, float *rmin2, float *rmax2) {
    if (criterion == EPPX_CRIT_DEFAULT) {
        const float gs = g.gap_square;
        *rmin2 = std::isfinite(gs) ? nextafterf(gs, INFINITY) : -1.0f;
        *rmax2 = std::isfinite(gs) ? -INFINITY : INFINITY;
    } else {
        *rmin2 = sym_min2;
        *rmax2 = sym_max2;
    }
}

// The particle system's one case: its grid, criterion bounds and tags; it
// always advances.
CaseDesc case_of(const eppx_ps *ps, const eppx_ns *ns, int32_t criterion) {
    CaseDesc c{};
    c.g = ns->grid;
    criterion_bounds(criterion, ns->grid, ps->rmin2, ps->rmax2, &c.rmin2, &c.rmax2);
    c.tags_present = ps->tags_present;
    c.check_all = (c.rmin2 <= c.rmax2) || ps->odd_tags;
    c.hinv_u = ps->hinv_u;
    c.m_u = ps->m_u;
    c.steps_left = 1;
    return c;
}

Walk make_walk(const eppx_ps *ps, const eppx_ns *ns, int32_t criterion) {
    Walk w{};
    const Fields &f = ps->buf[ps->cur];
    w.one = case_of(ps, ns, criterion);
    w.cases = nullptr;
    w.start = ps->start;
    w.geo = ps->geo;
    w.cell = f.cell;
    w.tag = f.tag;
    w.cs = f.cs;
    w.fallback = ps->fallback;
    w.n_alive = ps->n_alive;
    w.cap = ps->cap;
    memcpy(w.adj, ns->adj, sizeof(w.adj));
    w.criterion = criterion;
    w.masks = ps->masks;
    w.headers256 = ps->headers256;
    w.headers512 = ps->headers512;
    w.mask_stride = ps->cap;
    w.mask_mode = 0;
    return w;
}

bool has(int32_t col) { return col >= 0; }

// Columns an action reads or writes must exist in the layout (the reference
// fails to compile an action that names a missing PM field).
int32_t check_columns(eppx_ps *ps, std::initializer_list<std::pair<int32_t, const char *>> cols, const char *what) {
    for (auto &c : cols)
        if (!has(c.first))
            return set_error(ps->ctx, EPPX_E_ARG, std::string(what) + " needs column " + c.second);
    return EPPX_OK;
}

int32_t physics_arg(eppx_ctx *ctx, const void *params, size_t size, Physics *out) {
    if (!params || size != sizeof(eppx_physics))
        return set_error(ctx, EPPX_E_ARG, "parameter block must be an eppx_physics");
    const eppx_physics *p = static_cast<const eppx_physics *>(params);
    if (p->kernel != EPPX_KERNEL_CUBIC && p->kernel != EPPX_KERNEL_WENDLAND_C2)
        return set_error(ctx, EPPX_E_ARG, "unknown kernel");
    if (p->eos != EPPX_EOS_LINEAR && p->eos != EPPX_EOS_TAIT) return set_error(ctx, EPPX_E_ARG, "unknown eos");
    if (p->visc != EPPX_VISC_CLASSIC && p->visc != EPPX_VISC_ARTIFICIAL)
        return set_error(ctx, EPPX_E_ARG, "unknown viscosity");
    if (p->fluid_tag < 1 || p->fluid_tag > kMaxTag || p->wall_tag < 1 || p->wall_tag > kMaxTag)
        return set_error(ctx, EPPX_E_ARG, "fluid_tag and wall_tag must lie in 1..32");
    *out = to_physics(*p);
    return EPPX_OK;
}

int32_t field_arg(eppx_ctx *ctx, const eppx_ps *ps, const void *params, size_t size, eppx_field_op *out) {
    if (!params || size != sizeof(eppx_field_op))
        return set_error(ctx, EPPX_E_ARG, "parameter block must be an eppx_field_op");
    *out = *static_cast<const eppx_field_op *>(params);
    const eppx_layout &L = ps->layout;
    int32_t col;
    switch (out->field) {
        case EPPX_FIELD_POSITION: col = L.PositionVec; break;
        case EPPX_FIELD_VELOCITY: col = L.VelocityVec; break;
        case EPPX_FIELD_DVELOCITY: col = L.dVelocityVec; break;
        case EPPX_FIELD_MASS: col = L.Mass; break;
        case EPPX_FIELD_VOLUME: col = L.Volume; break;
        case EPPX_FIELD_DENSITY: col = L.Density; break;
        case EPPX_FIELD_DDENSITY: col = L.dDensity; break;
        case EPPX_FIELD_PRESSURE: col = L.Pressure; break;
        default: return set_error(ctx, EPPX_E_ARG, "unknown field");
    }
    const int32_t width = out->field <= EPPX_FIELD_DVELOCITY ? 2 : 1;
    if (col < 0 || out->component < 0 || out->component >= width)
        return set_error(ctx, EPPX_E_ARG, "field not in the layout or component out of range");
    return EPPX_OK;
}

// Bit of a pair-kernel block size in eppx_ps::header_blocks.
inline int32_t block_bit(int32_t block) { return block >= 512 ? 2 : 1; }

// clean! + insert + cell-order move + find + fused action.
// keys_done: the cell counts, keys and provisional slots of this search were
// already taken (k_move_key in a step).
// drho_zero (a step's searches): dDensity is zero on every alive slot of the
// current buffer and, with permute_no_drho, of the other one as well;
// permute_no_dvel: dVelocity is zero on every alive slot of both.
int32_t do_search(eppx_ps *ps, eppx_ns *ns, int32_t criterion, int32_t action, const Physics &ph,
                  bool keys_done = false, const float4 *dyn_src = nullptr, const float *vol_src = nullptr,
                  const EosOut *eos = nullptr, bool drho_zero = false, bool permute_no_drho = false,
                  bool permute_no_dvel = false) {
    eppx_ctx *ctx = ps->ctx;
    int32_t rc = ensure_cells(ps, ns->grid.n_cells);
    if (rc) return rc;
    const int32_t nc = ns->grid.n_cells;
    hipStream_t s = ctx->stream;
    Fields cur = ps->buf[ps->cur], nxt = ps->buf[ps->cur ^ 1];
    if (!keys_done) {
        HIP_OK(ctx, timed(ctx, "cell_reset", [&] { return launch_reset_search(ps->counts, nc + 2, ps->pairs, s); }));
        HIP_OK(ctx, timed(ctx, "cell_key", [&] {
                   return launch_key_count(cur, ps->n, case_of(ps, ns, criterion), nullptr, nc, ps->counts, ps->key,
                                           ps->tmpslot, ctx->d_flags, ns->M, s);
               }));
    }
    HIP_OK(ctx, timed(ctx, "cell_scan", [&] { return launch_scan(ps->counts, nc + 2, ps->start, ps->tile_sums, s); }));
    HIP_OK(ctx, timed(ctx, "cell_members", [&] {
               return launch_fill_members(cur, ps->n, nc, ps->key, ps->tmpslot, ps->start, ps->members, s);
           }));
    HIP_OK(ctx, timed(ctx, "cell_permute", [&] {
               return launch_permute(cur, nxt, ps->geo, ps->n, 0, nc, ps->key, ps->tmpslot, ps->start,
                                     ps->members, ps->n_alive, ctx->d_flags,
                                     (ps->accum_live ? kPermuteSums : 0) | (permute_no_drho ? kPermuteNoDrho : 0) |
                                         (permute_no_dvel ? kPermuteNoDvel : 0) | (ps->mh_one ? kPermuteOneMassH : 0),
                                     dyn_src, vol_src, ps->mass_one, ps->h_one, s);
           }));
    ps->cur ^= 1;
    Walk w = make_walk(ps, ns, criterion);
    w.drho_zero = drho_zero && action == EPPX_SEARCH_KERNEL_CONTINUITY;
    w.mask_mode = action == EPPX_SEARCH_KERNEL_CONTINUITY ? 1 : 0;  // keep the masks for the interactions
    w.header_mode = w.mask_mode;  // and the block headers
    const int fast = ctx->math_mode == EPPX_MATH_FAST;
    HIP_OK(ctx, timed(ctx, "search", [&] {
               if (eos)
                   return launch_search_eos(w, ps->f(), ps->n, ph, *eos, fast, ns->M, ps->ncount,
                                            ps->pairs + kPairLast, ps->pairs + kPairTotal, ctx->d_flags,
                                            ps->search_block, s);
               return launch_search(w, ps->f(), ps->n, action, ph, fast, ns->M, ps->ncount,
                                    ps->pairs + kPairLast, ps->pairs + kPairTotal, ctx->d_flags, 0, ps->search_block,
                                    ps->search_stage, s);
           }));
    ps->last_ns = ns;
    ps->last_crit = criterion;
    ps->last_action = action;
    ps->last_kernel = ph.kernel;
    ps->last_fluid = ph.fluid_tag;
    ps->last_wall = ph.wall_tag;
    ps->searched = true;
    ps->masks_valid = action == EPPX_SEARCH_KERNEL_CONTINUITY;
    ps->header_blocks = ps->masks_valid ? block_bit(fast ? ps->search_block : 256) : 0;
    return EPPX_OK;
}

void drop_graphs(eppx_ps *ps) {
    bool any = false;
    for (auto &row : ps->graphs)
        for (auto &g : row) any = any || g.exec;
    if (any) (void)hipStreamSynchronize(ps->ctx->stream);
    for (auto &row : ps->graphs)
        for (auto &g : row) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = StepGraph{};
        }
}

// The pair kernels read m_j from the search snapshot; an operation that
// changes Mass after the search refreshes it there too.  The masses may no
// longer be one value, so the FAST terms stop hoisting it (recorded steps
// carry the old case table and are dropped).
int32_t refresh_snapshot_mass(eppx_ps *ps, int32_t field) {
    if (field != EPPX_FIELD_MASS) return EPPX_OK;
    if (ps->m_u != 0.0f || ps->mh_one) {  // recorded permutes may skip the Mass column
        ps->m_u = 0.0f;
        ps->mh_one = false;
        drop_graphs(ps);
    }
    if (!ps->searched) return EPPX_OK;
    HIP_OK(ps->ctx, launch_snapshot_mass(ps->geo, ps->f(), ps->n, ps->ctx->stream));
    return EPPX_OK;
}

// dVelocity after a momentum pass (it writes the fluid's), a move (it zeroes
// the fluid's) or a field op on it (anything).
void note_dvel_momentum(eppx_ps *ps, int32_t fluid_tag) {
    const int32_t t = ps->dvel_nonzero_tag;
    ps->dvel_nonzero_tag = (t == -2 || t == fluid_tag) ? fluid_tag : -1;
}
void note_dvel_moved(eppx_ps *ps, int32_t fluid_tag) {
    if (ps->dvel_nonzero_tag == fluid_tag) ps->dvel_nonzero_tag = -2;
}

int32_t do_self(eppx_ps *ps, int32_t action, const Physics &ph, const eppx_field_op &op) {
    eppx_ctx *ctx = ps->ctx;
    if (action == EPPX_SELF_GRAVITY_MOVE) note_dvel_moved(ps, ph.fluid_tag);
    if (action == EPPX_SELF_SCALE_FIELD && op.field == EPPX_FIELD_DVELOCITY) ps->dvel_nonzero_tag = -1;
    const char *name = action == EPPX_SELF_CONTINUITY_EOS ? "continuity_eos"
                       : action == EPPX_SELF_GRAVITY_MOVE ? "gravity_move"
                       : action == EPPX_SELF_FILTER       ? "filter_self"
                       : action == EPPX_SELF_EXTRAPOLATE_PRESSURE ? "extrapolate_self"
                                                          : "scale_field";
    HIP_OK(ctx, timed(ctx, name, [&] { return launch_self(ps->f(), ps->n, action, ph, op, nullptr, ctx->stream); }));
    return refresh_snapshot_mass(ps, action == EPPX_SELF_SCALE_FIELD ? op.field : -1);
}

int32_t do_interaction(eppx_ps *ps, eppx_ns *ns, int32_t action, Physics ph, const eppx_field_op &op) {
    eppx_ctx *ctx = ps->ctx;
    if (action == EPPX_INTER_MOMENTUM) note_dvel_momentum(ps, ph.fluid_tag);
    if (action == EPPX_INTER_ADD_PER_NEIGHBOUR && op.field == EPPX_FIELD_DVELOCITY) ps->dvel_nonzero_tag = -1;
    if ((action == EPPX_INTER_FILTER || action == EPPX_INTER_EXTRAPOLATE_PRESSURE) && !ps->accum_live) {
        ps->accum_live = true;  // from now on the permute moves the sums (a recorded step moved none)
        drop_graphs(ps);
    }
    ph.kernel = ps->last_kernel;  // the stored nW / nDW of the reference come from the search
    Walk w = make_walk(ps, ns, ps->last_crit);
    // the search's masks hold the momentum pass's pairs when its pair classes are the search's
    if (action == EPPX_INTER_MOMENTUM && ps->masks_valid && ph.fluid_tag == ps->last_fluid &&
        ph.wall_tag == ps->last_wall)
        w.mask_mode = 2;
    const int fast = ctx->math_mode == EPPX_MATH_FAST;
    if (w.mask_mode == 2 && (ps->header_blocks & block_bit(fast ? ps->momentum_block : 256))) w.header_mode = 2;
    const char *name = action == EPPX_INTER_MOMENTUM                ? "momentum"
                       : action == EPPX_INTER_FILTER                ? "filter_inter"
                       : action == EPPX_INTER_EXTRAPOLATE_PRESSURE ? "extrapolate_inter"
                                                                    : "add_per_neighbour";
    HIP_OK(ctx, timed(ctx, name, [&] {
               return launch_interaction(w, ps->f(), ps->n, action, ph, op, fast, ps->momentum_block,
                                         ps->momentum_stage, ctx->stream);
           }));
    return refresh_snapshot_mass(ps, action == EPPX_INTER_ADD_PER_NEIGHBOUR ? op.field : -1);
}

// iMomentum! with sMomentum! and the coming search's cell keys fused (a
// particle system's time step; the search's masks give the pair set).
int32_t do_momentum_move(eppx_ps *ps, eppx_ns *ns, Physics ph, const MoveKey &mk) {
    eppx_ctx *ctx = ps->ctx;
    ph.kernel = ps->last_kernel;
    Walk w = make_walk(ps, ns, ps->last_crit);
    if (ps->masks_valid && ph.fluid_tag == ps->last_fluid && ph.wall_tag == ps->last_wall) w.mask_mode = 2;
    const int fast = ctx->math_mode == EPPX_MATH_FAST;
    if (w.mask_mode == 2 && (ps->header_blocks & block_bit(fast ? ps->momentum_block : 256))) w.header_mode = 2;
    HIP_OK(ctx, timed(ctx, "momentum", [&] {
               return launch_momentum_move(w, ps->f(), ps->n, ph, mk, fast, ps->momentum_block, ctx->stream);
           }));
    return EPPX_OK;
}

// Autotuner of the pair-kernel shape: times the FAST search action and the
// momentum pass on the current state for each candidate shape (256 or 512
// threads per block; per-neighbour data gathered from HBM or staged in the
// block's LDS tile), the first trial of each as a warm-up, keeps the fastest
// of each kernel, and restores what the trials wrote (dDensity, dVelocity,
// the pair counters), so the state does not change.  us[k] (k = 0..3 search,
// 4..7 momentum; shapes (256, gathered), (512, gathered), (256, staged),
// (512, staged)) receives the average microseconds.  EXACT math keeps
// (256, gathered).
constexpr int32_t kShapes = 4;
constexpr int32_t kShapeBlock[kShapes] = {256, 512, 256, 512};
constexpr int32_t kShapeStage[kShapes] = {0, 0, 1, 1};

int32_t tune_pair_blocks(eppx_ps *ps, const Walk &w, const Physics &ph, int32_t M, int32_t which, int32_t trials,
                         bool masks_valid, double *us) {
    eppx_ctx *ctx = ps->ctx;
    for (int k = 0; k < 2 * kShapes; ++k) us[k] = 0.0;
    if (ctx->math_mode != EPPX_MATH_FAST || ps->n == 0) {
        ps->search_block = ps->momentum_block = 256;
        ps->search_stage = ps->momentum_stage = 0;
        return EPPX_OK;
    }
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)ps->n;
    const Fields f = ps->f();
    float *drho = nullptr;
    float2 *dvel = nullptr;
    unsigned long long *pairs = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = dalloc(&drho, n);
    if (e == hipSuccess) e = dalloc(&dvel, n);
    if (e == hipSuccess) e = dalloc(&pairs, kPairWords);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpyAsync(drho, f.drho, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dvel, f.dvel, n * sizeof(float2), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pairs, ps->pairs, kPairWords * sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
    const eppx_field_op op{};
    Walk ws = w;  // the search trials write the block headers of every shape on this layout
    ws.header_mode = 1;
    Walk wm = w;  // the momentum pass reads the search's masks and headers as it will in a step
    wm.mask_mode = masks_valid ? 2 : 0;
    wm.header_mode = masks_valid ? 2 : 0;
    for (int kind = 0; kind < 2 && e == hipSuccess; ++kind) {
        for (int v = 0; v < kShapes && e == hipSuccess; ++v) {
            double total = 0.0;
            for (int t = 0; t <= trials && e == hipSuccess; ++t) {
                if (kind == 0) e = hipMemcpyAsync(f.drho, drho, n * sizeof(float), hipMemcpyDeviceToDevice, s);
                else e = hipMemcpyAsync(f.dvel, dvel, n * sizeof(float2), hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                if (e == hipSuccess)
                    e = kind == 0 ? launch_search(ws, f, ps->n, EPPX_SEARCH_KERNEL_CONTINUITY, ph, 1, M, ps->ncount,
                                                  ps->pairs + kPairLast, ps->pairs + kPairTotal, ctx->d_flags, which,
                                                  kShapeBlock[v], kShapeStage[v], s)
                                  : launch_interaction(wm, f, ps->n, EPPX_INTER_MOMENTUM, ph, op, 1, kShapeBlock[v],
                                                       kShapeStage[v], s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (t > 0) total += 1000.0 * ms;  // trial 0 warms up
            }
            us[kShapes * kind + v] = trials > 0 ? total / trials : 0.0;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(f.drho, drho, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(f.dvel, dvel, n * sizeof(float2), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ps->pairs, pairs, kPairWords * sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(drho);
    dfree(dvel);
    dfree(pairs);
    if (e != hipSuccess) return hip_fail(ctx, e, "autotune");
    if (masks_valid) ps->header_blocks = block_bit(256) | block_bit(512);
    if (trials > 0) {
        int best_s = 0, best_m = 0;
        for (int v = 1; v < kShapes; ++v) {
            if (us[v] < us[best_s]) best_s = v;
            if (us[kShapes + v] < us[kShapes + best_m]) best_m = v;
        }
        // In a step the momentum pass reads the search's block headers only when both kernels have one block
        // size (the trials gave every size its headers): the fastest pair of one size, when its total is
        // within 2 % of the fastest pair of any sizes.
        int pair_s = -1, pair_m = -1;
        for (int v = 0; v < kShapes; ++v)
            for (int m = 0; m < kShapes; ++m)
                if (kShapeBlock[v] == kShapeBlock[m] &&
                    (pair_s < 0 || us[v] + us[kShapes + m] < us[pair_s] + us[kShapes + pair_m])) {
                    pair_s = v;
                    pair_m = m;
                }
        if (pair_s >= 0 && us[pair_s] + us[kShapes + pair_m] <= 1.02 * (us[best_s] + us[kShapes + best_m])) {
            best_s = pair_s;
            best_m = pair_m;
        }
        ps->search_block = kShapeBlock[best_s];
        ps->search_stage = kShapeStage[best_s];
        ps->momentum_block = kShapeBlock[best_m];
        ps->momentum_stage = kShapeStage[best_m];
    }
    return EPPX_OK;
}

void shape_out(const eppx_ps *ps, int32_t *choice) {
    choice[0] = ps->search_block;
    choice[1] = ps->search_stage;
    choice[2] = ps->momentum_block;
    choice[3] = ps->momentum_stage;
}

}  // namespace

// ===================================================================== ABI
extern "C" {

int32_t eppx_abi_version(void) { return EPPX_ABI_VERSION; }

int32_t eppx_ctx_create(int32_t device, eppx_ctx **out) {
    if (!out) return EPPX_E_ARG;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || device < 0 || device >= count) return EPPX_E_HIP;
    eppx_ctx *ctx = new eppx_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        dalloc(&ctx->d_flags, 1) != hipSuccess || hipMemset(ctx->d_flags, 0, sizeof(uint32_t)) != hipSuccess) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        dfree(ctx->d_flags);
        delete ctx;
        return EPPX_E_HIP;
    }
    // the no-alloc download's scratch address on this device (once per device; a
    // failure is reported by that entry point, which then has no address)
    (void)init_scratch(device);
    *out = ctx;
    return EPPX_OK;
}

// A context with live particle / neighbour systems is freed when the last
// of them is destroyed, so handles may be released in any order.
int32_t eppx_ctx_destroy(eppx_ctx *ctx) {
    if (!ctx || ctx->closing) return EPPX_E_ARG;
    ctx->closing = true;
    release_ctx(ctx);
    return EPPX_OK;
}

const char *eppx_last_error(const eppx_ctx *ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int32_t eppx_sync(eppx_ctx *ctx) {
    if (!ctx) return EPPX_E_ARG;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return read_flags(ctx);
}

int32_t eppx_set_math_mode(eppx_ctx *ctx, int32_t mode) {
    if (!ctx || (mode != EPPX_MATH_EXACT && mode != EPPX_MATH_FAST)) return EPPX_E_ARG;
    ctx->math_mode = mode;
    return EPPX_OK;
}

int32_t eppx_set_timing(eppx_ctx *ctx, int32_t enable) {
    if (!ctx) return EPPX_E_ARG;
    ctx->timing = enable != 0;
    return EPPX_OK;
}

int32_t eppx_kernel_time(eppx_ctx *ctx, const char *kernel, double *avg_us, int64_t *launches, int32_t reset) {
    if (!ctx || !kernel) return EPPX_E_ARG;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    auto it = ctx->timed.find(kernel);
    const double sum = it == ctx->timed.end() ? 0.0 : it->second.first;
    const int64_t cnt = it == ctx->timed.end() ? 0 : it->second.second;
    if (avg_us) *avg_us = cnt ? sum / (double)cnt : 0.0;
    if (launches) *launches = cnt;
    if (reset) ctx->timed.clear();
    return EPPX_OK;
}

// Domain2D{Int32, Float32}(gap, first_x, first_y, last_x, last_y), Domain2D.jl:28-67.
int32_t eppx_domain2d_init(double gap, double first_x, double first_y, double last_x, double last_y,
                           eppx_domain2d *d) {
    if (!d) return EPPX_E_ARG;
    const float g = (float)gap;
    d->gap = g;
    d->gap_square = g * g;
    d->first_x = (float)first_x;
    d->first_y = (float)first_y;
    d->last_x = (float)last_x;
    d->last_y = (float)last_y;
    d->span_x = d->last_x - d->first_x;
    d->span_y = d->last_y - d->first_y;
    const float nx = floorf(d->span_x / g), ny = floorf(d->span_y / g);
    if (!(nx >= 1.0f) || !(ny >= 1.0f) || nx > 2.0e9f || ny > 2.0e9f || (double)nx * (double)ny > 2.0e9) {
        return EPPX_E_ARG;
    }
    d->n_x = (int32_t)nx;
    d->n_y = (int32_t)ny;
    d->n = d->n_x * d->n_y;
    d->gap_x = d->span_x / (float)d->n_x;
    d->gap_y = d->span_y / (float)d->n_y;
    d->gap_x_inv = 1.0f / d->gap_x;
    d->gap_y_inv = 1.0f / d->gap_y;
    return EPPX_OK;
}

// inside / indexCartesianFromPosition / indexLinearFromPosition, Domain2D.jl:110-137.
int32_t eppx_domain2d_locate(const eppx_domain2d *d, float x, float y, int32_t *inside, int32_t *pi, int32_t *pj,
                             int32_t *linear) {
    if (!d) return EPPX_E_ARG;
    const int in = d->first_x <= x && x <= d->last_x && d->first_y <= y && y <= d->last_y;
    int32_t i = (int32_t)floorf((x - d->first_x) * d->gap_x_inv) + 1;
    i = std::max(1, std::min(d->n_x, i));
    int32_t j = (int32_t)floorf((y - d->first_y) * d->gap_y_inv) + 1;
    j = std::max(1, std::min(d->n_y, j));
    if (inside) *inside = in;
    if (pi) *pi = i;
    if (pj) *pj = j;
    if (linear) *linear = i + d->n_x * (j - 1);
    return EPPX_OK;
}

int32_t eppx_ps_create(eppx_ctx *ctx, int32_t n_capacity, const eppx_layout *layout, eppx_ps **out) {
    if (!ctx || !layout || !out || n_capacity < 0) return set_error(ctx, EPPX_E_ARG, "eppx_ps_create: bad argument");
    // the pair kernels address slot records with 32-bit byte offsets
    if (n_capacity >= (1 << 28))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps_create: n_capacity must be below 2^28 (268,435,456)");
    *out = nullptr;
    const eppx_layout &L = *layout;
    if (L.n_int < 0 || L.n_float < 0 || L.max_neighbour_number < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_ps_create: negative table width");
    auto fits = [](int32_t col, int32_t width, int32_t n) { return col < 0 || (col + width <= n); };
    const int32_t M = L.max_neighbour_number;
    if (!fits(L.Tag, 1, L.n_int) || !fits(L.IsMovable, 1, L.n_int) || !fits(L.nCount, 1, L.n_int) ||
        !fits(L.nIndex, M, L.n_int) || !fits(L.PositionVec, 2, L.n_float) || !fits(L.VelocityVec, 2, L.n_float) ||
        !fits(L.dVelocityVec, 2, L.n_float) || !fits(L.Mass, 1, L.n_float) || !fits(L.Volume, 1, L.n_float) ||
        !fits(L.Density, 1, L.n_float) || !fits(L.dDensity, 1, L.n_float) || !fits(L.Pressure, 1, L.n_float) ||
        !fits(L.H, 1, L.n_float) || !fits(L.Gap, 1, L.n_float) || !fits(L.SumWeight, 1, L.n_float) ||
        !fits(L.SumWeightedDensity, 1, L.n_float) || !fits(L.SumWeightedPressure, 1, L.n_float) ||
        !fits(L.nRVec, 2 * M, L.n_float) ||
        !fits(L.nR, M, L.n_float) || !fits(L.nW, M, L.n_float) || !fits(L.nDW, M, L.n_float) ||
        !fits(L.nHInv, M, L.n_float))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps_create: a column lies outside its table");
    if (!has(L.PositionVec) || !has(L.Tag))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps_create: PositionVec and Tag are required");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if (ctx->closing) return set_error(ctx, EPPX_E_STATE, "eppx_ps_create: context is closed");
    eppx_ps *ps = new eppx_ps();
    ps->ctx = ctx;
    ctx->children += 1;
    ps->cap = n_capacity;
    ps->layout = L;
    hipError_t e = alloc_fields(ps->buf[0], n_capacity);
    if (e == hipSuccess) e = alloc_fields(ps->buf[1], n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->geo, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->fallback, 2);
    if (e == hipSuccess) e = hipMemsetAsync(ps->fallback, 0, 2 * sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess) e = dalloc(&ps->key, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->tmpslot, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->members, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->ncount, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->n_alive, 2);
    if (e == hipSuccess) e = dalloc(&ps->masks, 3 * (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->headers256, ((size_t)n_capacity / 256 + 1) * kHeaderInts256);
    if (e == hipSuccess) e = dalloc(&ps->headers512, ((size_t)n_capacity / 512 + 1) * kHeaderInts512);
    if (e == hipSuccess) e = dalloc(&ps->pairs, kPairWords);
    if (e == hipSuccess) e = dalloc(&ps->h_range, 8);
    if (e == hipSuccess) e = hipMemsetAsync(ps->pairs, 0, kPairWords * sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ps->n_alive, 0, 2 * sizeof(int32_t), ctx->stream);
    // the filter and extrapolation sums start at zero in both buffers (the
    // permute does not move them while they stay zero)
    for (int b = 0; b < 2 && e == hipSuccess; ++b)
        for (float *p : {ps->buf[b].wv, ps->buf[b].wvrho, ps->buf[b].wvp})
            if (e == hipSuccess && n_capacity)
                e = hipMemsetAsync(p, 0, (size_t)n_capacity * sizeof(float), ctx->stream);
    if (e != hipSuccess) {
        eppx_ps_destroy(ps);
        return hip_fail(ctx, e, "eppx_ps_create");
    }
    for (int c = 0; c < L.n_int; ++c)
        if (!owned_int(L, c)) ps->pass_int_cols.push_back(c);
    for (int c = 0; c < L.n_float; ++c)
        if (!owned_float(L, c)) ps->pass_float_cols.push_back(c);
    *out = ps;
    return EPPX_OK;
}

int32_t eppx_ps_destroy(eppx_ps *ps) {
    if (!ps) return EPPX_E_ARG;
    (void)hipSetDevice(ps->ctx->device);
    (void)hipStreamSynchronize(ps->ctx->stream);
    drop_graphs(ps);
    free_fields(ps->buf[0]);
    free_fields(ps->buf[1]);
    dfree(ps->geo);
    dfree(ps->fallback);
    dfree(ps->counts); dfree(ps->start); dfree(ps->tile_sums);
    dfree(ps->key); dfree(ps->tmpslot); dfree(ps->members); dfree(ps->ncount); dfree(ps->n_alive); dfree(ps->masks);
    dfree(ps->headers256); dfree(ps->headers512);
    dfree(ps->pairs); dfree(ps->h_range);
    dfree(ps->dyn_new); dfree(ps->vol_new);
    eppx_ctx *ctx = ps->ctx;
    delete ps;
    ctx->children -= 1;
    release_ctx(ctx);
    return EPPX_OK;
}

// Any non-zero value (NaN included) in rows [0, rows) of the SumWeight,
// SumWeightedDensity and SumWeightedPressure columns of a float table.
bool columns_zero(int32_t c0, int32_t width, const float *fp, size_t rows) {
    if (c0 < 0) return true;
    for (int32_t c = c0; c < c0 + width; ++c) {
        const float *col = fp + rows * (size_t)c;
        for (size_t r = 0; r < rows; ++r)
            if (!(col[r] == 0.0f)) return false;
    }
    return true;
}

// Whether column c holds one bit pattern on every row (value: that pattern; an absent column is zero).
bool column_one_value(int32_t c, const float *fp, size_t rows, float *value) {
    *value = 0.0f;
    if (c < 0 || rows == 0) return true;
    const float *col = fp + rows * (size_t)c;
    uint32_t first;
    memcpy(&first, col, sizeof(first));
    for (size_t r = 1; r < rows; ++r) {
        uint32_t v;
        memcpy(&v, col + r, sizeof(v));
        if (v != first) return false;
    }
    memcpy(value, &first, sizeof(first));
    return true;
}

bool accumulators_nonzero(const eppx_layout &L, const float *fp, size_t rows) {
    for (int32_t c : {L.SumWeight, L.SumWeightedDensity, L.SumWeightedPressure}) {
        if (c < 0) continue;
        const float *col = fp + rows * (size_t)c;
        for (size_t r = 0; r < rows; ++r)
            if (!(col[r] == 0.0f)) return true;
    }
    return false;
}

int32_t eppx_ps_upload(eppx_ps *ps, int32_t n_particles, const int32_t *is_alive, const int32_t *cell_index,
                       const int32_t *int_props, const float *float_props) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    if (n_particles < 0 || n_particles > ps->cap || !is_alive || !cell_index || (L.n_int && !int_props) ||
        (L.n_float && !float_props))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps_upload: bad argument (n_particles must be <= capacity)");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t cap = (size_t)ps->cap;
    int32_t *d_ip = nullptr, *d_alive = nullptr, *d_cell = nullptr;
    float *d_fp = nullptr;
    HIP_OK(ctx, dalloc(&d_ip, cap * (size_t)L.n_int));
    HIP_OK(ctx, dalloc(&d_fp, cap * (size_t)L.n_float));
    HIP_OK(ctx, dalloc(&d_alive, cap));
    HIP_OK(ctx, dalloc(&d_cell, cap));
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (cap) {
        if (L.n_int) e = hipMemcpyAsync(d_ip, int_props, cap * L.n_int * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess && L.n_float)
            e = hipMemcpyAsync(d_fp, float_props, cap * L.n_float * sizeof(float), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_alive, is_alive, cap * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_cell, cell_index, cap * sizeof(int32_t), hipMemcpyHostToDevice, s);
        // both buffers: the cell-order move rewrites only the slots in use
        for (int b = 0; b < 2 && e == hipSuccess; ++b)
            e = launch_from_ref(ps->buf[b], 0, columns_of(ps), d_ip, d_fp, d_alive, d_cell, 0, s);
    }
    ps->cur = 0;
    ps->n = n_particles;
    // H range over alive particles for the criterion's exact-safe bounds, and
    // the tags present.
    const float init[8] = {INFINITY, 0.0f, 0.0f, 0.0f, 0.0f, INFINITY, 0.0f, 0.0f};
    float range[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (e == hipSuccess) e = hipMemcpyAsync(ps->h_range, init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n_particles) e = launch_h_range(ps->buf[0], n_particles, ps->h_range, s);
    if (e == hipSuccess) e = hipMemcpyAsync(range, ps->h_range, sizeof(range), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_ip);
    dfree(d_fp);
    dfree(d_alive);
    dfree(d_cell);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_ps_upload");
    uint32_t bad, odd;
    memcpy(&bad, &range[2], sizeof(bad));
    memcpy(&ps->tags_present, &range[3], sizeof(uint32_t));
    memcpy(&odd, &range[4], sizeof(odd));
    ps->odd_tags = odd != 0;
    symmetry_bounds(!bad && has(L.H), range[0], range[1], &ps->rmin2, &ps->rmax2);
    ps->hinv_u = uniform_hinv(!bad && has(L.H), range[0], range[1]);
    uint32_t mbad;
    memcpy(&mbad, &range[7], sizeof(mbad));
    ps->m_u = uniform_mass(!mbad && has(L.Mass), range[5], range[6]);
    ps->accum_live = accumulators_nonzero(L, float_props, cap);  // both buffers now hold these rows
    ps->mh_one = column_one_value(L.Mass, float_props, cap, &ps->mass_one) &&
                 column_one_value(L.H, float_props, cap, &ps->h_one);
    ps->dvel_nonzero_tag = columns_zero(L.dVelocityVec, 2, float_props, cap) ? -2 : -1;
    // keep the columns the kernels do not own
    ps->pass_int.assign(ps->pass_int_cols.size() * cap, 0);
    ps->pass_float.assign(ps->pass_float_cols.size() * cap, 0.0f);
    for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
        memcpy(&ps->pass_int[k * cap], int_props + (size_t)ps->pass_int_cols[k] * cap, cap * sizeof(int32_t));
    for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
        memcpy(&ps->pass_float[k * cap], float_props + (size_t)ps->pass_float_cols[k] * cap, cap * sizeof(float));
    ps->searched = false;
    ps->masks_valid = false;
    ps->last_ns = nullptr;
    drop_graphs(ps);  // recorded kernel arguments (particle count, H bounds) may have changed
    return EPPX_OK;
}

}  // extern "C"

namespace {

// Device staging of a download: the reference tables in their host layout.
struct Staging {
    int32_t *ip, *alive, *cell;
    float *fp;
};

// Download of ps into the reference tables through device staging `st`
// (cap * (n_int + n_float + 2) words); fill: clear the staging with a fill
// kernel instead of hipMemsetAsync.
int32_t download_via(eppx_ps *ps, eppx_ns *ns, int32_t *is_alive, int32_t *cell_index, int32_t *int_props,
                     float *float_props, const Staging &st, bool fill) {
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    const size_t cap = (size_t)ps->cap;
    int32_t *d_ip = st.ip, *d_alive = st.alive, *d_cell = st.cell;
    float *d_fp = st.fp;
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (fill) {
        // int 0 and float +0.0 share their bits
        if (cap * L.n_int) e = launch_fill_i32(d_ip, 0, (int64_t)(cap * L.n_int), s);
        if (e == hipSuccess && cap * L.n_float)
            e = launch_fill_i32(reinterpret_cast<int32_t *>(d_fp), 0, (int64_t)(cap * L.n_float), s);
    } else {
        e = hipMemsetAsync(d_ip, 0, std::max<size_t>(1, cap * L.n_int) * sizeof(int32_t), s);
        if (e == hipSuccess) e = hipMemsetAsync(d_fp, 0, std::max<size_t>(1, cap * L.n_float) * sizeof(float), s);
    }
    if (e == hipSuccess) e = launch_to_ref(ps->f(), ps->cap, columns_of(ps), d_ip, d_fp, d_alive, d_cell, -1, s);
    const int32_t M = L.max_neighbour_number;
    if (e == hipSuccess && ns && ps->searched && ps->last_ns == ns &&
        (has(L.nCount) || has(L.nIndex) || has(L.nRVec) || has(L.nR) || has(L.nW) || has(L.nDW) || has(L.nHInv))) {
        NeighbourExport ex{};
        ex.n_capacity = ps->cap;
        ex.M = M;
        ex.nCount = has(L.nCount) ? d_ip + cap * L.nCount : nullptr;
        ex.nIndex = has(L.nIndex) && M ? d_ip + cap * L.nIndex : nullptr;
        ex.nRVec = has(L.nRVec) && M ? d_fp + cap * L.nRVec : nullptr;
        ex.nR = has(L.nR) && M ? d_fp + cap * L.nR : nullptr;
        ex.nW = has(L.nW) && M ? d_fp + cap * L.nW : nullptr;
        ex.nDW = has(L.nDW) && M ? d_fp + cap * L.nDW : nullptr;
        ex.nHInv = has(L.nHInv) && M ? d_fp + cap * L.nHInv : nullptr;
        ex.kernel = ps->last_kernel;
        ex.with_kernel = ps->last_action == EPPX_SEARCH_KERNEL_CONTINUITY;
        ex.fluid_tag = ps->last_fluid;
        ex.wall_tag = ps->last_wall;
        e = launch_export(make_walk(ps, ns, ps->last_crit), ps->f(), ps->n, nullptr, nullptr, ps->ncount, ex, s);
    }
    if (e == hipSuccess && L.n_int)
        e = hipMemcpyAsync(int_props, d_ip, cap * L.n_int * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && L.n_float)
        e = hipMemcpyAsync(float_props, d_fp, cap * L.n_float * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(is_alive, d_alive, cap * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(cell_index, d_cell, cap * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_ps_download");
    if (ps->pass_int.size() == ps->pass_int_cols.size() * cap)
        for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
            memcpy(int_props + (size_t)ps->pass_int_cols[k] * cap, &ps->pass_int[k * cap], cap * sizeof(int32_t));
    if (ps->pass_float.size() == ps->pass_float_cols.size() * cap)
        for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
            memcpy(float_props + (size_t)ps->pass_float_cols[k] * cap, &ps->pass_float[k * cap], cap * sizeof(float));
    return EPPX_OK;
}

int32_t download_args(eppx_ps *ps, int32_t *is_alive, int32_t *cell_index, int32_t *int_props, float *float_props) {
    const eppx_layout &L = ps->layout;
    if (!is_alive || !cell_index || (L.n_int && !int_props) || (L.n_float && !float_props))
        return set_error(ps->ctx, EPPX_E_ARG, "eppx_ps_download: bad argument");
    return EPPX_OK;
}

}  // namespace

extern "C" {

int32_t eppx_ps_download(eppx_ps *ps, eppx_ns *ns, int32_t *n_particles, int32_t *is_alive, int32_t *cell_index,
                         int32_t *int_props, float *float_props) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    int32_t rc = download_args(ps, is_alive, cell_index, int_props, float_props);
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t cap = (size_t)ps->cap;
    if (n_particles) *n_particles = ps->n;
    if (!cap) return read_flags(ctx);
    Staging st{};
    hipError_t e = dalloc(&st.ip, cap * (size_t)L.n_int);
    if (e == hipSuccess) e = dalloc(&st.fp, cap * (size_t)L.n_float);
    if (e == hipSuccess) e = dalloc(&st.alive, cap);
    if (e == hipSuccess) e = dalloc(&st.cell, cap);
    rc = e == hipSuccess ? download_via(ps, ns, is_alive, cell_index, int_props, float_props, st, false)
                         : hip_fail(ctx, e, "eppx_ps_download");
    dfree(st.ip);
    dfree(st.fp);
    dfree(st.alive);
    dfree(st.cell);
    return rc ? rc : read_flags(ctx);
}

int32_t eppx_ps_download_noalloc(eppx_ps *ps, eppx_ns *ns, int32_t *n_particles, int32_t *is_alive,
                                 int32_t *cell_index, int32_t *int_props, float *float_props) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    int32_t rc = download_args(ps, is_alive, cell_index, int_props, float_props);
    if (rc) return rc;
    const size_t cap = (size_t)ps->cap;
    // staging words: the int and float tables, is_alive and cell_index, each table 64-word aligned
    auto up = [](size_t w) { return (w + 63) & ~(size_t)63; };
    const size_t w_ip = up(cap * (size_t)L.n_int), w_fp = up(cap * (size_t)L.n_float), w_v = up(cap);
    const size_t words = w_ip + w_fp + 2 * w_v;
    if (words > kScratchWords) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "eppx_ps_download_noalloc: the tables need %zu staging words (capacity %zu x %d columns + 2); "
                 "the fixed device scratch holds %zu (use eppx_ps_download)",
                 words, cap, L.n_int + L.n_float, (size_t)kScratchWords);
        return set_error(ctx, EPPX_E_ARG, buf);
    }
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if (n_particles) *n_particles = ps->n;
    if (!cap) return read_flags(ctx);
    uint32_t *base = scratch_ptr(ctx->device);
    if (!base) return set_error(ctx, EPPX_E_HIP, "eppx_ps_download_noalloc: no device address for the scratch buffer");
    Staging st{};
    st.ip = reinterpret_cast<int32_t *>(base);
    st.fp = reinterpret_cast<float *>(base + w_ip);
    st.alive = reinterpret_cast<int32_t *>(base + w_ip + w_fp);
    st.cell = reinterpret_cast<int32_t *>(base + w_ip + w_fp + w_v);
    rc = download_via(ps, ns, is_alive, cell_index, int_props, float_props, st, true);
    return rc ? rc : read_flags(ctx);
}

int32_t eppx_ps_alive_count(eppx_ps *ps, int32_t *out) {
    if (!ps || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> alive((size_t)ps->n);
    if (ps->n)
        HIP_OK(ctx, hipMemcpyAsync(alive.data(), ps->f().alive, alive.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    int32_t c = 0;
    for (int32_t a : alive) c += a == 1;
    *out = c;
    return EPPX_OK;
}

int32_t eppx_ns_create(eppx_ctx *ctx, const eppx_domain2d *domain, const int32_t *adjacency, int32_t max_tag,
                       int32_t max_neighbour_number, eppx_ns **out) {
    if (!ctx || !domain || !adjacency || !out || max_tag < 1 || max_tag > kMaxTag || max_neighbour_number < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_ns_create: bad argument (1 <= max_tag <= 32)");
    if (domain->n_x < 1 || domain->n_y < 1 || (int64_t)domain->n_x * domain->n_y != domain->n)
        return set_error(ctx, EPPX_E_ARG, "eppx_ns_create: inconsistent domain");
    if (ctx->closing) return set_error(ctx, EPPX_E_STATE, "eppx_ns_create: context is closed");
    eppx_ns *ns = new eppx_ns();
    ns->ctx = ctx;
    ctx->children += 1;
    ns->domain = *domain;
    const eppx_domain2d &d = *domain;
    ns->grid = Grid{d.first_x, d.first_y, d.last_x, d.last_y, d.gap_x_inv, d.gap_y_inv, d.gap_square,
                    d.n_x, d.n_y, d.n};
    ns->max_tag = max_tag;
    ns->M = max_neighbour_number;
    static std::atomic<uint64_t> next_serial{1};
    ns->serial = next_serial++;
    for (int t = 1; t <= max_tag; ++t)
        for (int u = 1; u <= max_tag; ++u)
            if (adjacency[(t - 1) + (size_t)max_tag * (u - 1)] == 1) ns->adj[t] |= 1u << (u - 1);
    *out = ns;
    return EPPX_OK;
}

int32_t eppx_ns_destroy(eppx_ns *ns) {
    if (!ns) return EPPX_E_ARG;
    eppx_ctx *ctx = ns->ctx;
    delete ns;
    ctx->children -= 1;
    release_ctx(ctx);
    return EPPX_OK;
}

int32_t eppx_ns_cell_stencil(eppx_ns *ns, int32_t *count, int32_t *list) {
    if (!ns || !count || !list) return EPPX_E_ARG;
    eppx_ctx *ctx = ns->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t nc = (size_t)ns->grid.n_cells;
    int32_t *d_count = nullptr, *d_list = nullptr;
    HIP_OK(ctx, dalloc(&d_count, nc));
    HIP_OK(ctx, dalloc(&d_list, nc * 9));
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemsetAsync(d_list, 0, nc * 9 * sizeof(int32_t), s);
    if (e == hipSuccess) e = launch_stencil(ns->grid, d_count, d_list, s);
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_count, nc * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(list, d_list, nc * 9 * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_count);
    dfree(d_list);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_ns_cell_stencil");
    return EPPX_OK;
}

int32_t eppx_search(eppx_ps *ps, eppx_ns *ns, int32_t criterion, int32_t action, const void *params,
                    size_t params_size) {
    if (!ps || !ns) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (ns->ctx != ctx) return set_error(ctx, EPPX_E_ARG, "eppx_search: systems belong to different contexts");
    if (criterion != EPPX_CRIT_DEFAULT && criterion != EPPX_CRIT_SYMMETRY)
        return set_error(ctx, EPPX_E_ARG, "eppx_search: unknown criterion");
    const eppx_layout &L = ps->layout;
    if (criterion == EPPX_CRIT_SYMMETRY && !has(L.H))
        return set_error(ctx, EPPX_E_ARG, "eppx_search: the symmetry criterion needs column H");
    Physics ph{};
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        int32_t rc = physics_arg(ctx, params, params_size, &ph);
        if (rc) return rc;
        rc = check_columns(ps, {{L.H, "H"}, {L.VelocityVec, "VelocityVec"}, {L.Mass, "Mass"}, {L.Volume, "Volume"},
                                {L.Density, "Density"}, {L.dDensity, "dDensity"}},
                           "sComputeKernelAndiContinuity");
        if (rc) return rc;
    } else if (action != EPPX_SEARCH_NONE) {
        return set_error(ctx, EPPX_E_ARG, "eppx_search: unknown action");
    }
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_search(ps, ns, criterion, action, ph);
}

int32_t eppx_selfaction(eppx_ps *ps, int32_t action, const void *params, size_t params_size) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    Physics ph{};
    eppx_field_op op{};
    int32_t rc;
    switch (action) {
        case EPPX_SELF_CONTINUITY_EOS:
            rc = physics_arg(ctx, params, params_size, &ph);
            if (!rc)
                rc = check_columns(ps, {{L.Density, "Density"}, {L.dDensity, "dDensity"}, {L.Volume, "Volume"},
                                        {L.Mass, "Mass"}, {L.Pressure, "Pressure"}},
                                   "sContinuity");
            break;
        case EPPX_SELF_GRAVITY_MOVE:
            rc = physics_arg(ctx, params, params_size, &ph);
            if (!rc)
                rc = check_columns(ps, {{L.VelocityVec, "VelocityVec"}, {L.dVelocityVec, "dVelocityVec"}},
                                   "sMomentum");
            break;
        case EPPX_SELF_FILTER:
            rc = physics_arg(ctx, params, params_size, &ph);
            if (!rc)
                rc = check_columns(ps, {{L.SumWeight, "SumWeight"}, {L.SumWeightedDensity, "SumWeightedDensity"},
                                        {L.H, "H"}, {L.Volume, "Volume"}, {L.Mass, "Mass"}, {L.Density, "Density"},
                                        {L.Pressure, "Pressure"}},
                                   "sFilter");
            break;
        case EPPX_SELF_EXTRAPOLATE_PRESSURE:
            rc = physics_arg(ctx, params, params_size, &ph);
            if (!rc)
                rc = check_columns(ps, {{L.SumWeight, "SumWeight"}, {L.SumWeightedPressure, "SumWeightedPressure"},
                                        {L.Pressure, "Pressure"}},
                                   "sExtrapolatePressure");
            break;
        case EPPX_SELF_SCALE_FIELD:
            rc = field_arg(ctx, ps, params, params_size, &op);
            break;
        default:
            rc = set_error(ctx, EPPX_E_ARG, "eppx_selfaction: unknown action");
    }
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_self(ps, action, ph, op);
}

int32_t eppx_interaction(eppx_ps *ps, eppx_ns *ns, int32_t action, const void *params, size_t params_size) {
    if (!ps || !ns) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns)
        return set_error(ctx, EPPX_E_STATE, "eppx_interaction: no search on this neighbour system yet");
    const eppx_layout &L = ps->layout;
    Physics ph{};
    eppx_field_op op{};
    int32_t rc;
    switch (action) {
        case EPPX_INTER_MOMENTUM:
        case EPPX_INTER_FILTER:
        case EPPX_INTER_EXTRAPOLATE_PRESSURE:
            if (ps->last_action != EPPX_SEARCH_KERNEL_CONTINUITY)
                return set_error(ctx, EPPX_E_STATE,
                                 "eppx_interaction: kernel values come from a search with "
                                 "EPPX_SEARCH_KERNEL_CONTINUITY");
            rc = physics_arg(ctx, params, params_size, &ph);
            if (!rc && action == EPPX_INTER_MOMENTUM)
                rc = check_columns(ps, {{L.Density, "Density"}, {L.Pressure, "Pressure"}, {L.Mass, "Mass"},
                                        {L.VelocityVec, "VelocityVec"}, {L.dVelocityVec, "dVelocityVec"}},
                                   "iMomentum");
            if (!rc && action == EPPX_INTER_FILTER)
                rc = check_columns(ps, {{L.SumWeight, "SumWeight"}, {L.SumWeightedDensity, "SumWeightedDensity"},
                                        {L.Volume, "Volume"}, {L.Mass, "Mass"}},
                                   "iFilter");
            if (!rc && action == EPPX_INTER_EXTRAPOLATE_PRESSURE)
                rc = check_columns(ps, {{L.SumWeight, "SumWeight"}, {L.SumWeightedPressure, "SumWeightedPressure"},
                                        {L.Volume, "Volume"}, {L.Density, "Density"}, {L.Pressure, "Pressure"}},
                                   "iExtrapolatePressure");
            break;
        case EPPX_INTER_ADD_PER_NEIGHBOUR:
            rc = field_arg(ctx, ps, params, params_size, &op);
            break;
        default:
            rc = set_error(ctx, EPPX_E_ARG, "eppx_interaction: unknown action");
    }
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_interaction(ps, ns, action, ph, op);
}

int32_t eppx_step(eppx_ps *ps, eppx_ns *ns, const eppx_physics *phys, int32_t n_steps, int32_t filter_interval,
                  int64_t step_offset) {
    if (!ps || !ns || !phys || n_steps < 0) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns || ps->last_action != EPPX_SEARCH_KERNEL_CONTINUITY ||
        ps->last_crit != EPPX_CRIT_SYMMETRY)
        return set_error(ctx, EPPX_E_STATE,
                         "eppx_step: needs a prior eppx_search(symmetry, kernel+continuity) on this system");
    Physics ph{};
    int32_t rc = physics_arg(ctx, phys, sizeof(*phys), &ph);
    if (rc) return rc;
    const eppx_layout &L = ps->layout;
    rc = check_columns(ps, {{L.Density, "Density"}, {L.dDensity, "dDensity"}, {L.Volume, "Volume"}, {L.Mass, "Mass"},
                            {L.Pressure, "Pressure"}, {L.VelocityVec, "VelocityVec"},
                            {L.dVelocityVec, "dVelocityVec"}, {L.H, "H"}},
                       "eppx_step");
    if (rc) return rc;
    if (filter_interval > 0)
        rc = check_columns(ps, {{L.SumWeight, "SumWeight"}, {L.SumWeightedDensity, "SumWeightedDensity"}},
                           "eppx_step filter");
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_cells(ps, ns->grid.n_cells))) return rc;  // no allocation while a step is recorded
    const eppx_field_op op{};
    const int32_t nc = ns->grid.n_cells;
    hipStream_t s = ctx->stream;
    // The fused step (gathered pair shapes): momentum runs sMomentum! and the
    // cell keys in its epilogue, and the search may run the next step's
    // sContinuity! in its own; their outputs for the permute go to dyn_new /
    // vol_new.
    const bool fuse_move = ps->momentum_stage == 0, fuse_eos_ok = ps->search_stage == 0;
    if ((fuse_move || fuse_eos_ok) && !ps->dyn_new) {  // no allocation while a step is recorded
        HIP_OK(ctx, dalloc(&ps->dyn_new, (size_t)ps->cap));
        HIP_OK(ctx, dalloc(&ps->vol_new, (size_t)ps->cap));
    }
    // With every particle fluid or wall, the step's sContinuity! leaves
    // dDensity zero on every alive slot, and so does the fused search that
    // applies the next one: from the call's first permute on, both buffers
    // hold zeros there (slots only die), so the searches start their sums at
    // zero and the later permutes leave dDensity alone.
    auto tag_bit_h = [](int32_t t) { return (uint32_t)t - 1u < (uint32_t)kMaxTag ? 1u << (t - 1) : 0u; };
    const bool drho_zero = !ps->odd_tags &&
                           (ps->tags_present & ~(tag_bit_h(ph.fluid_tag) | tag_bit_h(ph.wall_tag))) == 0u;
    // With dVelocity zero on every particle but the fluid, the first step's
    // move zeroes the fluid's and the permute copies the zeros: after it both
    // buffers hold zeros on every alive slot, so the later steps start the
    // momentum sums at zero and neither the move nor the permute writes them.
    const bool dvel_skip = fuse_move && (ps->dvel_nonzero_tag == -2 || ps->dvel_nonzero_tag == ph.fluid_tag);
    // eos_first: the step starts with sContinuity! (first step of the call,
    // or after a step whose search did not apply it); fuse_eos: its search
    // applies the next step's; dvel_zero: dVelocity is zero on every alive slot.
    auto enqueue_step = [&](bool eos_first, bool fuse_eos, bool filter_now, bool dvel_zero) -> int32_t {
        int32_t r;
        if (eos_first)
            HIP_OK(ctx, timed(ctx, "continuity_eos", [&] {
                       return launch_eos_reset(ps->f(), ps->n, ph, nullptr, ps->counts, nc + 2, ps->pairs + kPairLast,
                                               s);
                   }));
        if (fuse_move) {
            const MoveKey mk{ps->n, nc, ns->M, ps->counts, ps->key, ps->tmpslot, ctx->d_flags, ps->dyn_new,
                             ps->pairs + kPairLast, dvel_zero ? 1 : 0};
            if ((r = do_momentum_move(ps, ns, ph, mk))) return r;
        } else {
            if ((r = do_interaction(ps, ns, EPPX_INTER_MOMENTUM, ph, op))) return r;
            HIP_OK(ctx, timed(ctx, "gravity_move", [&] {
                       return launch_move_key(ps->f(), ps->n, ph, case_of(ps, ns, EPPX_CRIT_SYMMETRY), nullptr, nc,
                                              ps->counts, ps->key, ps->tmpslot, ctx->d_flags, ns->M, s);
                   }));
        }
        const EosOut eo{ps->n, nc + 2, ps->vol_new, ps->counts};
        if ((r = do_search(ps, ns, EPPX_CRIT_SYMMETRY, EPPX_SEARCH_KERNEL_CONTINUITY, ph, true,
                           fuse_move ? ps->dyn_new : nullptr, eos_first ? nullptr : ps->vol_new,
                           fuse_eos ? &eo : nullptr, drho_zero, drho_zero && !eos_first, dvel_zero)))
            return r;
        if (filter_now) {
            if ((r = do_interaction(ps, ns, EPPX_INTER_FILTER, ph, op))) return r;
            if ((r = do_self(ps, EPPX_SELF_FILTER, ph, op))) return r;
        }
        return EPPX_OK;
    };
    bool eos_pending = true;
    for (int32_t k = 1; k <= n_steps; ++k) {
        const bool filter_now = filter_interval > 0 && (step_offset + k) % filter_interval == 0;
        // the last step of the call leaves the state as the reference's loop
        // body does (its dDensity not yet applied); a filter step too
        const bool fuse = fuse_eos_ok && k < n_steps && !filter_now;
        const bool eos_first = eos_pending;
        eos_pending = !fuse;
        const bool dvel_zero = dvel_skip && k > 1;
        if (filter_now || ctx->timing || !ps->masks_valid) {  // filter steps and per-kernel timing run eagerly
            if ((rc = enqueue_step(eos_first, fuse, filter_now, dvel_zero))) return rc;
            continue;
        }
        StepGraph &g = ps->graphs[ps->cur][(eos_first ? 0 : 2) + (fuse ? 1 : 0) + (dvel_zero ? 4 : 0)];
        if (!g.exec || g.ns_serial != ns->serial || g.math != ctx->math_mode || memcmp(&g.ph, &ph, sizeof(ph)) != 0) {
            if (g.exec) {
                (void)hipStreamSynchronize(ctx->stream);  // an earlier launch of it may still run
                (void)hipGraphExecDestroy(g.exec);
            }
            g = StepGraph{};
            const int cur0 = ps->cur;
            HIP_OK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            rc = enqueue_step(eos_first, fuse, false, dvel_zero);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            ps->cur = cur0;  // recorded, not run yet
            if (rc == EPPX_OK && e != hipSuccess) rc = hip_fail(ctx, e, "eppx_step: hipStreamEndCapture");
            if (rc == EPPX_OK) {
                const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                if (ei != hipSuccess) rc = hip_fail(ctx, ei, "eppx_step: hipGraphInstantiate");
            }
            if (graph) (void)hipGraphDestroy(graph);
            if (rc) {
                g = StepGraph{};
                return rc;
            }
            g.ns_serial = ns->serial;
            g.ph = ph;
            g.math = ctx->math_mode;
        }
        HIP_OK(ctx, hipGraphLaunch(g.exec, ctx->stream));
        ps->cur ^= 1;
    }
    if (n_steps > 0) note_dvel_moved(ps, ph.fluid_tag);  // every step ends with the fluid's dVelocity zeroed
    return EPPX_OK;
}

int32_t eppx_autotune(eppx_ps *ps, eppx_ns *ns, const eppx_physics *phys, int32_t trials, int32_t *choice,
                      double *us) {
    if (!ps || !ns || !phys || trials < 0 || !choice) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns)
        return set_error(ctx, EPPX_E_STATE, "eppx_autotune: needs a prior eppx_search on this neighbour system");
    Physics ph{};
    int32_t rc = physics_arg(ctx, phys, sizeof(*phys), &ph);
    if (rc) return rc;
    ph.kernel = ps->last_kernel;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    double t[2 * kShapes];
    rc = tune_pair_blocks(ps, make_walk(ps, ns, ps->last_crit), ph, ns->M, 0, trials,
                          ps->masks_valid && ph.fluid_tag == ps->last_fluid && ph.wall_tag == ps->last_wall, t);
    if (rc) return rc;
    drop_graphs(ps);  // recorded steps carry the old shapes
    shape_out(ps, choice);
    if (us) memcpy(us, t, sizeof(t));
    return read_flags(ctx);
}

int32_t eppx_set_pair_shape(eppx_ps *ps, const int32_t *choice) {
    auto ok = [](int32_t b, int32_t st) { return (b == 256 || b == 512) && (st == 0 || st == 1); };
    if (!ps || !choice || !ok(choice[0], choice[1]) || !ok(choice[2], choice[3])) return EPPX_E_ARG;
    ps->search_block = choice[0];
    ps->search_stage = choice[1];
    ps->momentum_block = choice[2];
    ps->momentum_stage = choice[3];
    drop_graphs(ps);
    return EPPX_OK;
}

int32_t eppx_pair_count(eppx_ps *ps, eppx_ns *ns, int64_t *out) {
    if (!ps || !ns || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns) return set_error(ctx, EPPX_E_STATE, "eppx_pair_count: no search yet");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    unsigned long long v = 0;
    HIP_OK(ctx, read_pair_slots(ctx, ps->pairs, kPairLast, &v));
    *out = (int64_t)v;
    return read_flags(ctx);
}

int32_t eppx_pair_total(eppx_ps *ps, int64_t *out, int32_t reset) {
    if (!ps || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    unsigned long long v = 0;
    HIP_OK(ctx, read_pair_slots(ctx, ps->pairs, kPairTotal, &v));
    if (reset) {
        HIP_OK(ctx, hipMemsetAsync(ps->pairs + kPairTotal, 0, kPairSlots * kPairStride * sizeof(unsigned long long),
                                   ctx->stream));
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    }
    *out = (int64_t)v;
    return EPPX_OK;
}

int32_t eppx_export_pairs(eppx_ps *ps, eppx_ns *ns, int64_t *offsets, int32_t *js, int64_t capacity) {
    if (!ps || !ns || !offsets || (capacity > 0 && !js)) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns) return set_error(ctx, EPPX_E_STATE, "eppx_export_pairs: no search yet");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    int32_t na = 0;
    HIP_OK(ctx, hipMemcpyAsync(&na, ps->n_alive, sizeof(na), hipMemcpyDeviceToHost, s));
    HIP_OK(ctx, hipStreamSynchronize(s));
    std::vector<int32_t> cnt((size_t)na), ids((size_t)na);
    if (na) {
        HIP_OK(ctx, hipMemcpyAsync(cnt.data(), ps->ncount, na * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(ctx, hipMemcpyAsync(ids.data(), ps->f().id, na * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_OK(ctx, hipStreamSynchronize(s));
    }
    const int32_t n = ps->n;
    std::vector<int64_t> by_row((size_t)n + 1, 0);
    for (int32_t k = 0; k < na; ++k) by_row[(size_t)ids[k] + 1] = cnt[k];
    for (int32_t r = 0; r < n; ++r) by_row[(size_t)r + 1] += by_row[(size_t)r];
    memcpy(offsets, by_row.data(), by_row.size() * sizeof(int64_t));
    const int64_t total = by_row[(size_t)n];
    if (total > capacity) {
        char buf[160];
        snprintf(buf, sizeof(buf), "eppx_export_pairs: %lld pairs exceed capacity %lld", (long long)total,
                 (long long)capacity);
        return set_error(ctx, EPPX_E_CAPACITY, buf);
    }
    if (!total) return EPPX_OK;
    std::vector<int64_t> row_start((size_t)na);
    for (int32_t k = 0; k < na; ++k) row_start[k] = by_row[(size_t)ids[k]];
    int64_t *d_rs = nullptr;
    int32_t *d_js = nullptr;
    HIP_OK(ctx, dalloc(&d_rs, (size_t)na));
    HIP_OK(ctx, dalloc(&d_js, (size_t)total));
    NeighbourExport ex{};
    ex.n_capacity = ps->cap;
    hipError_t e = hipMemcpyAsync(d_rs, row_start.data(), na * sizeof(int64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_export(make_walk(ps, ns, ps->last_crit), ps->f(), ps->n, d_rs, d_js, ps->ncount, ex, s);
    if (e == hipSuccess) e = hipMemcpyAsync(js, d_js, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_rs);
    dfree(d_js);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_export_pairs");
    return read_flags(ctx);
}

int32_t eppx_pair_fallback_blocks(eppx_ps *ps, int64_t *out, int32_t reset) {
    if (!ps || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    uint32_t v = 0;
    HIP_OK(ctx, hipMemcpyAsync(&v, ps->fallback, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    if (reset) HIP_OK(ctx, hipMemsetAsync(ps->fallback, 0, sizeof(uint32_t), ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    *out = v;
    return EPPX_OK;
}

int32_t eppx_pair_row_fallback(eppx_ps *ps, int64_t *out, int32_t reset) {
    if (!ps || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    uint32_t v = 0;
    HIP_OK(ctx, hipMemcpyAsync(&v, ps->fallback + 1, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    if (reset) HIP_OK(ctx, hipMemsetAsync(ps->fallback + 1, 0, sizeof(uint32_t), ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    *out = v;
    return EPPX_OK;
}

int32_t eppx_valu_calibration(eppx_ctx *ctx, int32_t blocks, int32_t iters, double *us) {
    if (!ctx || blocks < 1 || iters < 1) return set_error(ctx, EPPX_E_ARG, "eppx_valu_calibration: bad argument");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    float *d_out = nullptr;
    HIP_OK(ctx, dalloc(&d_out, (size_t)blocks));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_valu_calibration(d_out, blocks, iters, ctx->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    if (e == hipSuccess) e = launch_valu_calibration(d_out, blocks, iters, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(d_out);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_valu_calibration");
    if (us) *us = 1000.0 * ms;
    return EPPX_OK;
}

int32_t eppx_check_finite(eppx_ps *ps) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    uint32_t *d_bad = nullptr, bad = 0;
    HIP_OK(ctx, dalloc(&d_bad, 1));
    hipError_t e = hipMemsetAsync(d_bad, 0, sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess && ps->n) e = launch_finite(ps->f(), ps->n, d_bad, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(d_bad);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_check_finite");
    if (bad) return set_error(ctx, EPPX_E_NONFINITE, "non-finite density, pressure, position or velocity");
    return EPPX_OK;
}

}  // extern "C"

// ===================================================================== batch
// Many independent cases in one set of device buffers (an eppx_ps that owns
// the slots) and one segmented cell list: case c owns the global cells
// [cell_base, cell_base + n_cells) and its particles sort after those of
// every case with a smaller cell_base.  One launch per kernel per step
// advances every case that has steps left; each case stops after exactly its
// own step budget.  The physics is shared except for dt.

namespace {

struct BatchCase {
    bool used = false;
    int32_t n = 0;                   // rows of the case's reference tables
    std::vector<int32_t> pass_int;   // columns the kernels do not own, n rows each
    std::vector<float> pass_float;
};

}  // namespace

struct eppx_batch {
    eppx_ctx *ctx = nullptr;
    eppx_ps *ps = nullptr;   // the slots of every case
    eppx_ns ns{};            // adjacency and max_neighbour_number (the grids are per case)
    int32_t max_cases = 0;
    Physics ph{};
    bool ph_set = false;
    CaseDesc *d_cases = nullptr;
    std::vector<CaseDesc> cases;  // what d_cases holds
    std::vector<BatchCase> meta;
    int32_t total_cells = 0;
    uint64_t generation = 1;  // membership changes; recorded steps are only replayed for the same one
    StepGraph graphs[2];
};

namespace {

bool same_physics_class(const Physics &a, const Physics &b) {
    Physics x = a, y = b;
    x.dt = y.dt = 0.0f;
    return memcmp(&x, &y, sizeof(x)) == 0;
}

void drop_batch_graphs(eppx_batch *b) {
    if (b->graphs[0].exec || b->graphs[1].exec) (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &g : b->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = StepGraph{};
    }
}

int32_t upload_cases(eppx_batch *b) {
    eppx_ctx *ctx = b->ctx;
    HIP_OK(ctx, hipMemcpyAsync(b->d_cases, b->cases.data(), b->cases.size() * sizeof(CaseDesc),
                               hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));  // the host vector may change right after
    return EPPX_OK;
}

// The device decrements steps_left; the host copy takes those counts before
// it changes the table and writes it back.
int32_t pull_steps(eppx_batch *b) {
    eppx_ctx *ctx = b->ctx;
    std::vector<CaseDesc> now(b->max_cases);
    HIP_OK(ctx, hipMemcpyAsync(now.data(), b->d_cases, now.size() * sizeof(CaseDesc), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < b->max_cases; ++c) b->cases[c].steps_left = now[c].steps_left;
    return EPPX_OK;
}

// Cell bases in slot order of the cases in use, the cell arrays sized for
// them, and the table on the device.
int32_t relayout(eppx_batch *b) {
    int64_t base = 0;
    for (int c = 0; c < b->max_cases; ++c) {
        if (!b->meta[c].used) continue;
        b->cases[c].cell_base = (int32_t)base;
        base += b->cases[c].g.n_cells;
    }
    if (base > INT32_MAX - 3) return set_error(b->ctx, EPPX_E_CAPACITY, "batch: more than 2^31 cells");
    b->total_cells = (int32_t)base;
    int32_t rc = ensure_cells(b->ps, b->total_cells);
    if (rc) return rc;
    b->generation += 1;
    drop_batch_graphs(b);
    return upload_cases(b);
}

Walk batch_walk(const eppx_batch *b) {
    const eppx_ps *ps = b->ps;
    Walk w{};
    w.cases = b->d_cases;
    w.start = ps->start;
    w.geo = ps->geo;
    const Fields &f = ps->buf[ps->cur];
    w.cell = f.cell;
    w.tag = f.tag;
    w.cs = f.cs;
    w.fallback = ps->fallback;
    w.n_alive = ps->n_alive;
    w.cap = ps->cap;
    memcpy(w.adj, b->ns.adj, sizeof(w.adj));
    w.criterion = EPPX_CRIT_SYMMETRY;
    w.masks = ps->masks;
    w.headers256 = ps->headers256;
    w.headers512 = ps->headers512;
    w.mask_stride = ps->cap;
    w.mask_mode = 0;
    return w;
}

// The segmented cell build and (search) the fused kernel + continuity action
// for the cases `which` selects (0: with steps left, 1: waiting for a first
// search).
int32_t batch_cell_build(eppx_batch *b, bool search, int32_t which, bool keys_done = false) {
    eppx_ps *ps = b->ps;
    eppx_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    const int32_t nc = b->total_cells;
    Fields cur = ps->buf[ps->cur], nxt = ps->buf[ps->cur ^ 1];
    if (!keys_done) {
        HIP_OK(ctx, timed(ctx, "cell_reset", [&] { return launch_reset_search(ps->counts, nc + 2, ps->pairs, s); }));
        HIP_OK(ctx, timed(ctx, "cell_key", [&] {
                   return launch_key_count(cur, ps->n, CaseDesc{}, b->d_cases, nc, ps->counts, ps->key, ps->tmpslot,
                                           ctx->d_flags, b->ns.M, s);
               }));
    }
    HIP_OK(ctx, timed(ctx, "cell_scan", [&] { return launch_scan(ps->counts, nc + 2, ps->start, ps->tile_sums, s); }));
    HIP_OK(ctx, timed(ctx, "cell_members", [&] {
               return launch_fill_members(cur, ps->n, nc, ps->key, ps->tmpslot, ps->start, ps->members, s);
           }));
    HIP_OK(ctx, timed(ctx, "cell_permute", [&] {
               return launch_permute(cur, nxt, ps->geo, ps->n, 1, nc, ps->key, ps->tmpslot, ps->start, ps->members,
                                     ps->n_alive, ctx->d_flags, ps->accum_live ? kPermuteSums : 0, nullptr, nullptr,
                                     0.0f, 0.0f, s);
           }));
    ps->cur ^= 1;
    if (!search) {
        ps->masks_valid = false;  // slots moved, masks did not
        return EPPX_OK;
    }
    Walk w = batch_walk(b);
    w.mask_mode = 1;
    const int fast = ctx->math_mode == EPPX_MATH_FAST;
    HIP_OK(ctx, timed(ctx, "search", [&] {
               return launch_search(w, ps->f(), ps->n, EPPX_SEARCH_KERNEL_CONTINUITY, b->ph, fast, b->ns.M,
                                    ps->ncount, ps->pairs + kPairLast, ps->pairs + kPairTotal, ctx->d_flags, which,
                                    ps->search_block, ps->search_stage, s);
           }));
    ps->masks_valid = which == 0;  // a first search (1) leaves the other cases' moved slots without masks
    return EPPX_OK;
}

// One step of every case with steps left (main.jl:166-183 per case).
int32_t batch_enqueue_step(eppx_batch *b) {
    eppx_ps *ps = b->ps;
    eppx_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    const eppx_field_op op{};
    const int fast = ctx->math_mode == EPPX_MATH_FAST;
    HIP_OK(ctx, timed(ctx, "continuity_eos", [&] {
               return launch_eos_reset(ps->f(), ps->n, b->ph, b->d_cases, ps->counts, b->total_cells + 2,
                                       ps->pairs + kPairLast, s);
           }));
    Walk w = batch_walk(b);
    w.mask_mode = ps->masks_valid ? 2 : 0;
    HIP_OK(ctx, timed(ctx, "momentum", [&] {
               return launch_interaction(w, ps->f(), ps->n, EPPX_INTER_MOMENTUM, b->ph, op, fast, ps->momentum_block,
                                         ps->momentum_stage, s);
           }));
    HIP_OK(ctx, timed(ctx, "gravity_move", [&] {
               return launch_move_key(ps->f(), ps->n, b->ph, CaseDesc{}, b->d_cases, b->total_cells, ps->counts,
                                      ps->key, ps->tmpslot, ctx->d_flags, b->ns.M, s);
           }));
    int32_t rc = batch_cell_build(b, true, 0, true);
    if (rc) return rc;
    HIP_OK(ctx, launch_tick(b->d_cases, b->max_cases, s));
    return EPPX_OK;
}

}  // namespace

extern "C" {

int32_t eppx_batch_bytes_per_particle(int32_t *out) {
    if (!out) return EPPX_E_ARG;
    // two buffers of 80 B (the 76 B of particle state and the case id), the
    // 16 B search snapshot, key, provisional slot, cell member and count, and
    // the search's three 64-bit accept masks
    *out = 2 * 80 + 16 + 4 * 4 + 24;
    return EPPX_OK;
}

int32_t eppx_batch_create(eppx_ctx *ctx, const eppx_layout *layout, int32_t particle_capacity, int32_t case_capacity,
                          const int32_t *adjacency, int32_t max_tag, int32_t max_neighbour_number, eppx_batch **out) {
    if (!ctx || !layout || !out || particle_capacity < 0 || case_capacity < 1 || !adjacency || max_tag < 1 ||
        max_tag > kMaxTag || max_neighbour_number < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_batch_create: bad argument");
    *out = nullptr;
    eppx_ps *ps = nullptr;
    int32_t rc = eppx_ps_create(ctx, particle_capacity, layout, &ps);
    if (rc) return rc;
    eppx_batch *b = new eppx_batch();
    b->ctx = ctx;
    b->ps = ps;
    ps->n = 0;
    b->max_cases = case_capacity;
    b->ns.ctx = ctx;
    b->ns.max_tag = max_tag;
    b->ns.M = max_neighbour_number;
    for (int t = 1; t <= max_tag; ++t)
        for (int u = 1; u <= max_tag; ++u)
            if (adjacency[(t - 1) + (size_t)max_tag * (u - 1)] == 1) b->ns.adj[t] |= 1u << (u - 1);
    b->cases.assign(case_capacity, CaseDesc{});
    b->meta.assign(case_capacity, BatchCase{});
    hipError_t e = dalloc(&b->d_cases, (size_t)case_capacity);
    if (e != hipSuccess) {
        eppx_batch_destroy(b);
        return hip_fail(ctx, e, "eppx_batch_create");
    }
    if ((rc = relayout(b))) {
        eppx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return EPPX_OK;
}

int32_t eppx_batch_destroy(eppx_batch *b) {
    if (!b) return EPPX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    drop_batch_graphs(b);
    dfree(b->d_cases);
    eppx_ps *ps = b->ps;
    delete b;
    return ps ? eppx_ps_destroy(ps) : EPPX_OK;
}

int32_t eppx_batch_add(eppx_batch *b, const eppx_domain2d *domain, const eppx_physics *phys, int32_t n_particles,
                       const int32_t *is_alive, const int32_t *cell_index, const int32_t *int_props,
                       const float *float_props, int32_t steps, int32_t *case_id) {
    if (!b || !domain || !phys || !case_id || n_particles < 0 || steps < 0 || !is_alive || !cell_index)
        return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    eppx_ps *ps = b->ps;
    const eppx_layout &L = ps->layout;
    if ((L.n_int && !int_props) || (L.n_float && !float_props))
        return set_error(ctx, EPPX_E_ARG, "eppx_batch_add: missing property table");
    Physics ph{};
    int32_t rc = physics_arg(ctx, phys, sizeof(*phys), &ph);
    if (rc) return rc;
    if (b->ph_set && !same_physics_class(ph, b->ph))
        return set_error(ctx, EPPX_E_ARG, "eppx_batch_add: physics differs from the batch's (only dt may differ)");
    rc = check_columns(ps, {{L.Density, "Density"}, {L.dDensity, "dDensity"}, {L.Volume, "Volume"},
                            {L.Mass, "Mass"}, {L.Pressure, "Pressure"}, {L.VelocityVec, "VelocityVec"},
                            {L.dVelocityVec, "dVelocityVec"}, {L.H, "H"}},
                       "eppx_batch_add");
    if (rc) return rc;
    if (domain->n_x < 1 || domain->n_y < 1 || domain->n != domain->n_x * domain->n_y)
        return set_error(ctx, EPPX_E_ARG, "eppx_batch_add: bad domain");
    int c = 0;
    while (c < b->max_cases && b->meta[c].used) ++c;
    if (c == b->max_cases) return set_error(ctx, EPPX_E_CAPACITY, "eppx_batch_add: every case slot is in use");
    if ((int64_t)ps->n + n_particles > ps->cap) {
        char buf[160];
        snprintf(buf, sizeof(buf), "eppx_batch_add: %d particles do not fit (%d of %d slots in use)", n_particles,
                 ps->n, ps->cap);
        return set_error(ctx, EPPX_E_CAPACITY, buf);
    }
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if ((rc = pull_steps(b))) return rc;
    // H range over alive particles and the tags present, on the host
    const size_t rows = (size_t)n_particles;
    float hmin = INFINITY, hmax = 0.0f, mmin = INFINITY, mmax = 0.0f;
    bool bad = false, odd = false, mbad = !has(L.Mass);
    uint32_t tags = 0;
    for (size_t r = 0; r < rows; ++r) {
        const int32_t t = int_props[r + rows * (size_t)L.Tag];
        const uint32_t t1 = (uint32_t)t - 1u;
        if (t1 < (uint32_t)kMaxTag) tags |= 1u << t1;
        else odd = true;
        if (is_alive[r] != 1) continue;
        const float h = float_props[r + rows * (size_t)L.H];
        if (!(h > 0.0f) || !std::isfinite(h)) bad = true;
        else {
            hmin = std::min(hmin, h);
            hmax = std::max(hmax, h);
        }
        const float m = float_props[r + rows * (size_t)L.Mass];
        if (!(m > 0.0f) || !std::isfinite(m)) mbad = true;
        else {
            mmin = std::min(mmin, m);
            mmax = std::max(mmax, m);
        }
    }
    CaseDesc cd{};
    const eppx_domain2d &d = *domain;
    cd.g = Grid{d.first_x, d.first_y, d.last_x, d.last_y, d.gap_x_inv, d.gap_y_inv, d.gap_square, d.n_x, d.n_y, d.n};
    symmetry_bounds(!bad, hmin, hmax, &cd.rmin2, &cd.rmax2);
    cd.tags_present = tags;
    cd.check_all = (cd.rmin2 <= cd.rmax2) || odd;
    cd.hinv_u = uniform_hinv(!bad, hmin, hmax);
    cd.m_u = uniform_mass(!mbad, mmin, mmax);
    cd.dt = ph.dt;
    cd.steps_left = steps;
    cd.fresh = 1;
    // the case's tables into slots [n, n + n_particles) of the current buffer
    if (n_particles) {
        int32_t *d_ip = nullptr, *d_alive = nullptr, *d_cell = nullptr;
        float *d_fp = nullptr;
        HIP_OK(ctx, dalloc(&d_ip, rows * (size_t)L.n_int));
        HIP_OK(ctx, dalloc(&d_fp, rows * (size_t)L.n_float));
        HIP_OK(ctx, dalloc(&d_alive, rows));
        HIP_OK(ctx, dalloc(&d_cell, rows));
        hipStream_t s = ctx->stream;
        hipError_t e = hipSuccess;
        if (L.n_int) e = hipMemcpyAsync(d_ip, int_props, rows * L.n_int * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess && L.n_float)
            e = hipMemcpyAsync(d_fp, float_props, rows * L.n_float * sizeof(float), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_alive, is_alive, rows * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_cell, cell_index, rows * sizeof(int32_t), hipMemcpyHostToDevice, s);
        Columns cols = columns_of(ps);
        cols.n_capacity = n_particles;
        if (e == hipSuccess) e = launch_from_ref(ps->f(), ps->n, cols, d_ip, d_fp, d_alive, d_cell, c, s);
        if (e == hipSuccess && !ps->accum_live && accumulators_nonzero(L, float_props, rows)) {
            ps->accum_live = true;  // the permute moves the sums from now on
            drop_batch_graphs(b);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        dfree(d_ip);
        dfree(d_fp);
        dfree(d_alive);
        dfree(d_cell);
        if (e != hipSuccess) return hip_fail(ctx, e, "eppx_batch_add");
    }
    BatchCase &m = b->meta[c];
    m.used = true;
    m.n = n_particles;
    m.pass_int.assign(ps->pass_int_cols.size() * rows, 0);
    m.pass_float.assign(ps->pass_float_cols.size() * rows, 0.0f);
    for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
        memcpy(&m.pass_int[k * rows], int_props + (size_t)ps->pass_int_cols[k] * rows, rows * sizeof(int32_t));
    for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
        memcpy(&m.pass_float[k * rows], float_props + (size_t)ps->pass_float_cols[k] * rows, rows * sizeof(float));
    ps->n += n_particles;
    b->cases[c] = cd;
    if (!b->ph_set) {
        b->ph = ph;
        b->ph_set = true;
    }
    if ((rc = relayout(b))) return rc;
    // the case's first search (main.jl:160), before it takes part in steps
    if ((rc = batch_cell_build(b, true, 1))) return rc;
    b->cases[c].fresh = 0;
    if ((rc = upload_cases(b))) return rc;
    *case_id = c;
    return read_flags(ctx);
}

int32_t eppx_batch_step(eppx_batch *b, int32_t n_steps) {
    if (!b || n_steps < 0) return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    eppx_ps *ps = b->ps;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if (!b->ph_set || ps->n == 0) return EPPX_OK;
    for (int32_t k = 0; k < n_steps; ++k) {
        int32_t rc;
        if (ctx->timing || !ps->masks_valid) {  // the step after a membership change finds its own pairs
            if ((rc = batch_enqueue_step(b))) return rc;
            continue;
        }
        StepGraph &g = b->graphs[ps->cur];
        if (!g.exec || g.ns_serial != b->generation || g.math != ctx->math_mode) {
            if (g.exec) {
                (void)hipStreamSynchronize(ctx->stream);  // an earlier launch of it may still run
                (void)hipGraphExecDestroy(g.exec);
            }
            g = StepGraph{};
            const int cur0 = ps->cur;
            HIP_OK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
            rc = batch_enqueue_step(b);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            ps->cur = cur0;  // recorded, not run yet
            if (rc == EPPX_OK && e != hipSuccess) rc = hip_fail(ctx, e, "eppx_batch_step: hipStreamEndCapture");
            if (rc == EPPX_OK) {
                const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                if (ei != hipSuccess) rc = hip_fail(ctx, ei, "eppx_batch_step: hipGraphInstantiate");
            }
            if (graph) (void)hipGraphDestroy(graph);
            if (rc) {
                g = StepGraph{};
                return rc;
            }
            g.ns_serial = b->generation;
            g.math = ctx->math_mode;
        }
        HIP_OK(ctx, hipGraphLaunch(g.exec, ctx->stream));
        ps->cur ^= 1;
    }
    return EPPX_OK;
}

int32_t eppx_batch_steps_left(eppx_batch *b, int32_t *steps_left) {
    if (!b || !steps_left) return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    std::vector<CaseDesc> now(b->max_cases);
    HIP_OK(ctx, hipMemcpyAsync(now.data(), b->d_cases, now.size() * sizeof(CaseDesc), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < b->max_cases; ++c) {
        steps_left[c] = b->meta[c].used ? now[c].steps_left : -1;
        b->cases[c].steps_left = now[c].steps_left;
    }
    return read_flags(ctx);
}

int32_t eppx_batch_download(eppx_batch *b, int32_t case_id, int32_t *n_particles, int32_t *is_alive,
                            int32_t *cell_index, int32_t *int_props, float *float_props) {
    if (!b || case_id < 0 || case_id >= b->max_cases || !n_particles || !is_alive || !cell_index) return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    eppx_ps *ps = b->ps;
    const BatchCase &m = b->meta[case_id];
    if (!m.used) return set_error(ctx, EPPX_E_STATE, "eppx_batch_download: no such case");
    const eppx_layout &L = ps->layout;
    if ((L.n_int && !int_props) || (L.n_float && !float_props)) return EPPX_E_ARG;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t rows = (size_t)m.n;
    *n_particles = m.n;
    if (!rows) return EPPX_OK;
    int32_t *d_ip = nullptr, *d_alive = nullptr, *d_cell = nullptr;
    float *d_fp = nullptr;
    HIP_OK(ctx, dalloc(&d_ip, std::max<size_t>(1, rows * (size_t)L.n_int)));
    HIP_OK(ctx, dalloc(&d_fp, std::max<size_t>(1, rows * (size_t)L.n_float)));
    HIP_OK(ctx, dalloc(&d_alive, rows));
    HIP_OK(ctx, dalloc(&d_cell, rows));
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemsetAsync(d_ip, 0, std::max<size_t>(1, rows * L.n_int) * sizeof(int32_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(d_fp, 0, std::max<size_t>(1, rows * L.n_float) * sizeof(float), s);
    Columns cols = columns_of(ps);
    cols.n_capacity = m.n;
    if (e == hipSuccess) e = launch_to_ref(ps->f(), ps->n, cols, d_ip, d_fp, d_alive, d_cell, case_id, s);
    if (e == hipSuccess && L.n_int)
        e = hipMemcpyAsync(int_props, d_ip, rows * L.n_int * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && L.n_float)
        e = hipMemcpyAsync(float_props, d_fp, rows * L.n_float * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(is_alive, d_alive, rows * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(cell_index, d_cell, rows * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_ip);
    dfree(d_fp);
    dfree(d_alive);
    dfree(d_cell);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_batch_download");
    for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
        memcpy(int_props + (size_t)ps->pass_int_cols[k] * rows, &m.pass_int[k * rows], rows * sizeof(int32_t));
    for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
        memcpy(float_props + (size_t)ps->pass_float_cols[k] * rows, &m.pass_float[k * rows], rows * sizeof(float));
    return read_flags(ctx);
}

int32_t eppx_batch_remove(eppx_batch *b, int32_t case_id) {
    if (!b || case_id < 0 || case_id >= b->max_cases) return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    eppx_ps *ps = b->ps;
    if (!b->meta[case_id].used) return set_error(ctx, EPPX_E_STATE, "eppx_batch_remove: no such case");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    int32_t rc = pull_steps(b);
    if (rc) return rc;
    // the case's slots sort last at a cell build, then the slot count drops to the kept ones
    b->cases[case_id].removed = 1;
    b->cases[case_id].steps_left = 0;
    if ((rc = upload_cases(b))) return rc;
    if (ps->n) {
        if ((rc = batch_cell_build(b, false, 0))) return rc;
        int32_t kept = 0;
        HIP_OK(ctx, hipMemcpyAsync(&kept, ps->n_alive + 1, sizeof(kept), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
        ps->n = kept;
    }
    b->meta[case_id] = BatchCase{};
    b->cases[case_id] = CaseDesc{};
    if ((rc = relayout(b))) return rc;
    // the cases after the removed one moved to lower cells: the cell list of
    // the new layout, before a step walks it
    if (ps->n && (rc = batch_cell_build(b, false, 0))) return rc;
    return read_flags(ctx);
}

int32_t eppx_batch_autotune(eppx_batch *b, int32_t trials, int32_t *choice, double *us) {
    if (!b || trials < 0 || !choice) return EPPX_E_ARG;
    eppx_ctx *ctx = b->ctx;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    double t[2 * kShapes] = {};
    if (b->ph_set && b->ps->n) {
        const int32_t rc = tune_pair_blocks(b->ps, batch_walk(b), b->ph, b->ns.M, 0, trials, b->ps->masks_valid, t);
        if (rc) return rc;
    }
    drop_batch_graphs(b);
    shape_out(b->ps, choice);
    if (us) memcpy(us, t, sizeof(t));
    return read_flags(ctx);
}

int32_t eppx_batch_info(eppx_batch *b, int32_t *slots_in_use, int32_t *cases_in_use) {
    if (!b) return EPPX_E_ARG;
    if (slots_in_use) *slots_in_use = b->ps->n;
    if (cases_in_use) {
        int32_t k = 0;
        for (const auto &m : b->meta) k += m.used;
        *cases_in_use = k;
    }
    return EPPX_OK;
}

int32_t eppx_batch_pair_total(eppx_batch *b, int64_t *out, int32_t reset) {
    if (!b) return EPPX_E_ARG;
    return eppx_pair_total(b->ps, out, reset);
}

}  // extern "C"

// Host side of the 3D path of libeppx: the eppx_*3d entry points of
// include/eppx.h.
//
// The reference's 3D layers are stubs (AbstractDomain.jl:119,
// NeighbourSystemBase.jl:50-57, NeighbourSearch.jl:49 and :229), so this
// path extends the 2D one axis by axis and is checked against the 3D
// restatement of the oracle, not against reference output.  A 3D particle
// system keeps the same handles' life cycle as the 2D one: a context, a
// neighbour system over a Domain3D grid, a particle system whose state lives
// twice in HBM (the cell-list build moves every particle from one copy to
// the other in cell order), the snapshot of the last search and its cells.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "d3_internal.h"
#include "host_common.h"

using namespace eppx;
using namespace eppx::host;

struct eppx_ns3d {
    eppx_ctx *ctx = nullptr;
    eppx_domain3d domain{};
    Grid3 grid{};
    uint32_t adj[kMaxTag + 1] = {};
    int32_t max_tag = 0;
    int32_t M = 0;
};

struct eppx_ps3d {
    eppx_ctx *ctx = nullptr;
    int32_t cap = 0, n = 0;
    eppx_layout layout{};
    Fields3 buf[2]{};
    int cur = 0;
    float4 *geo = nullptr;  // (x, y, z, tag) snapshot of the last search, sorted order
    int32_t n_cells = -1;
    int32_t *counts = nullptr, *start = nullptr, *tile_sums = nullptr;
    uint32_t *cell_tags = nullptr;
    int32_t *key = nullptr, *tmpslot = nullptr, *members = nullptr, *ncount = nullptr, *n_alive = nullptr;
    unsigned long long *pairs = nullptr;  // pair counters: kPairWords words (kPairLast, kPairTotal slots)
    float rmin2 = -1.0f, rmax2 = INFINITY;
    float hinv_u = 0.0f;
    float m_u = 0.0f;  // one Mass for every alive particle: that mass; else 0 (the 3D actions never change Mass)
    uint32_t tags_present = 0xffffffffu;
    bool odd_tags = true;
    eppx_ns3d *last_ns = nullptr;
    int32_t last_crit = 0, last_action = 0, last_kernel = 0, last_fluid = 0, last_wall = 0;
    bool searched = false;
    // the last search's accept masks, 27 per slot (allocated by the first search)
    unsigned long long *masks = nullptr;
    bool masks_valid = false;
    std::vector<int32_t> pass_int_cols, pass_float_cols;
    std::vector<int32_t> pass_int;
    std::vector<float> pass_float;
    Fields3 &f() { return buf[cur]; }
};

namespace {

constexpr size_t kPad = 64;  // slots past the capacity in every float4 array (the candidate tests read ahead)

bool has(int32_t col) { return col >= 0; }

hipError_t alloc_fields3(Fields3 &f, int32_t cap) {
    hipError_t e = hipSuccess;
    auto a4 = [&](float4 **p) {
        if (e == hipSuccess) e = dalloc(p, (size_t)cap + kPad);
        if (e == hipSuccess) e = hipMemset(*p, 0, ((size_t)cap + kPad) * sizeof(float4));
    };
    auto a = [&](auto **p) {
        if (e == hipSuccess) e = dalloc(p, (size_t)cap);
    };
    a4(&f.pos); a4(&f.vel); a4(&f.hmvp); a4(&f.dvel);
    a(&f.drho); a(&f.mov); a(&f.id); a(&f.cell); a(&f.alive);
    return e;
}

void free_fields3(Fields3 &f) {
    dfree(f.pos); dfree(f.vel); dfree(f.hmvp); dfree(f.dvel);
    dfree(f.drho); dfree(f.mov); dfree(f.id); dfree(f.cell); dfree(f.alive);
}

// Columns the 3D kernels own (vectors are 3 wide, nRVec 3 per neighbour).
bool owned_float3(const eppx_layout &L, int c) {
    auto in = [&](int start, int width) { return start >= 0 && c >= start && c < start + width; };
    const int M = L.max_neighbour_number;
    return in(L.PositionVec, 3) || in(L.VelocityVec, 3) || in(L.dVelocityVec, 3) || in(L.Mass, 1) ||
           in(L.Volume, 1) || in(L.Density, 1) || in(L.dDensity, 1) || in(L.Pressure, 1) || in(L.H, 1) ||
           in(L.nRVec, 3 * M) || in(L.nR, M) || in(L.nW, M) || in(L.nDW, M) || in(L.nHInv, M);
}

bool owned_int3(const eppx_layout &L, int c) {
    auto in = [&](int start, int width) { return start >= 0 && c >= start && c < start + width; };
    return in(L.Tag, 1) || in(L.IsMovable, 1) || in(L.nCount, 1) || in(L.nIndex, L.max_neighbour_number);
}

Columns columns3_of(const eppx_ps3d *ps) {
    const eppx_layout &L = ps->layout;
    return Columns{ps->cap, L.n_int, L.n_float, L.Tag, L.IsMovable, L.PositionVec, L.VelocityVec, L.dVelocityVec,
                   L.Mass, L.Volume, L.Density, L.dDensity, L.Pressure, L.H, -1, -1, -1};
}

int32_t physics3_arg(eppx_ctx *ctx, const void *params, size_t size, Physics3 *out) {
    if (!params || size != sizeof(eppx_physics3d))
        return set_error(ctx, EPPX_E_ARG, "parameter block must be an eppx_physics3d");
    const eppx_physics3d *p = static_cast<const eppx_physics3d *>(params);
    const eppx_physics &b = p->base;
    if (b.kernel != EPPX_KERNEL_CUBIC && b.kernel != EPPX_KERNEL_WENDLAND_C2)
        return set_error(ctx, EPPX_E_ARG, "unknown kernel");
    if (b.eos != EPPX_EOS_LINEAR && b.eos != EPPX_EOS_TAIT) return set_error(ctx, EPPX_E_ARG, "unknown eos");
    if (b.visc != EPPX_VISC_CLASSIC && b.visc != EPPX_VISC_ARTIFICIAL)
        return set_error(ctx, EPPX_E_ARG, "unknown viscosity");
    if (b.fluid_tag < 1 || b.fluid_tag > kMaxTag || b.wall_tag < 1 || b.wall_tag > kMaxTag)
        return set_error(ctx, EPPX_E_ARG, "fluid_tag and wall_tag must lie in 1..32");
    out->p = to_physics(b);
    out->gz = p->gz;
    return EPPX_OK;
}

int32_t check_columns3(eppx_ps3d *ps, std::initializer_list<std::pair<int32_t, const char *>> cols, const char *what) {
    for (auto &c : cols)
        if (c.first < 0) {
            char buf[200];
            snprintf(buf, sizeof(buf), "%s needs the %s column", what, c.second);
            return set_error(ps->ctx, EPPX_E_ARG, buf);
        }
    return EPPX_OK;
}

int32_t ensure_cells3(eppx_ps3d *ps, int32_t n_cells) {
    if (ps->n_cells == n_cells) return EPPX_OK;
    eppx_ctx *ctx = ps->ctx;
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ps->counts);
    dfree(ps->start);
    dfree(ps->tile_sums);
    dfree(ps->cell_tags);
    const int64_t nb = (int64_t)n_cells + 2;  // cells + dead bucket + an empty bucket
    HIP_OK(ctx, dalloc(&ps->counts, (size_t)nb));
    HIP_OK(ctx, dalloc(&ps->start, (size_t)nb + 1));
    HIP_OK(ctx, dalloc(&ps->tile_sums, (size_t)((nb + kScanTile - 1) / kScanTile)));
    HIP_OK(ctx, dalloc(&ps->cell_tags, (size_t)n_cells));
    HIP_OK(ctx, hipMemsetAsync(ps->start, 0, ((size_t)nb + 1) * sizeof(int32_t), ctx->stream));
    HIP_OK(ctx, hipMemsetAsync(ps->cell_tags, 0, (size_t)std::max(n_cells, 1) * sizeof(uint32_t), ctx->stream));
    ps->n_cells = n_cells;
    return EPPX_OK;
}

Walk3 make_walk3(const eppx_ps3d *ps, const eppx_ns3d *ns, int32_t criterion) {
    Walk3 w{};
    const Fields3 &f = ps->buf[ps->cur];
    w.g = ns->grid;
    w.start = ps->start;
    w.cell_tags = ps->cell_tags;
    w.geo = ps->geo;
    w.vel = f.vel;
    w.hmvp = f.hmvp;
    w.cell = f.cell;
    w.n_alive = ps->n_alive;
    memcpy(w.adj, ns->adj, sizeof(w.adj));
    w.tags_present = ps->tags_present;
    w.odd_tags = ps->odd_tags ? 1 : 0;
    w.criterion = criterion;
    if (criterion == EPPX_CRIT_DEFAULT) {
        // defaultSearchCriterion: dr2 <= gap_square, i.e. dr2 < the next float up
        const float gs = ns->grid.gap_square;
        w.rmin2 = std::isfinite(gs) ? nextafterf(gs, INFINITY) : -1.0f;
        w.rmax2 = std::isfinite(gs) ? -INFINITY : INFINITY;
    } else {
        w.rmin2 = ps->rmin2;
        w.rmax2 = ps->rmax2;
    }
    w.hinv_u = ps->hinv_u;
    w.m_u = ps->m_u;
    w.masks = ps->masks;
    w.mask_stride = ps->cap;
    w.mask_mode = 0;
    const int32_t groups = (ns->grid.n_cells + 3) / 4;  // 4 cells per pair block
    w.stripe = 8;                                       // 32 consecutive cells per stripe
    w.stripes_per_xcd = std::max(1, (groups + 8 * w.stripe - 1) / (8 * w.stripe));
    return w;
}

// Algorithm.search! in 3D: clean + insert (cell keys, counts, scan, the
// serial-insert order) + find with the fused action.
int32_t do_search3(eppx_ps3d *ps, eppx_ns3d *ns, int32_t criterion, int32_t action, const Physics &ph) {
    eppx_ctx *ctx = ps->ctx;
    hipStream_t s = ctx->stream;
    const int32_t nc = ns->grid.n_cells;
    int32_t rc = ensure_cells3(ps, nc);
    if (rc) return rc;
    const int32_t nb = nc + 2;
    Fields3 cur = ps->buf[ps->cur], nxt = ps->buf[ps->cur ^ 1];
    HIP_OK(ctx, timed(ctx, "cell_key", [&] {
               hipError_t e = launch_reset_search(ps->counts, nb, ps->pairs + kPairLast, s);
               if (e == hipSuccess)
                   e = launch3_key_count(cur, ps->n, ns->grid, ps->counts, ps->key, ps->tmpslot, ctx->d_flags,
                                         ns->M, s);
               return e;
           }));
    HIP_OK(ctx, timed(ctx, "cell_scan", [&] { return launch_scan(ps->counts, nb, ps->start, ps->tile_sums, s); }));
    HIP_OK(ctx, timed(ctx, "cell_permute", [&] {
               hipError_t e = launch3_fill_members(cur, ps->n, nc, ps->key, ps->tmpslot, ps->start, ps->members, s);
               if (e == hipSuccess)
                   e = launch3_permute(cur, nxt, ps->geo, ps->n, nc, ps->key, ps->tmpslot, ps->start, ps->members,
                                       ps->n_alive, ctx->d_flags, s);
               if (e == hipSuccess) e = launch3_cell_tags(ps->geo, nc, ps->start, ps->cell_tags, s);
               return e;
           }));
    ps->cur ^= 1;
    if (!ps->masks) HIP_OK(ctx, dalloc(&ps->masks, (size_t)27 * (size_t)ps->cap));
    Walk3 w = make_walk3(ps, ns, criterion);
    w.mask_mode = 1;
    ps->masks_valid = false;
    HIP_OK(ctx, timed(ctx, "search", [&] {
               return launch3_search(w, ps->f(), action, ph, ctx->math_mode == EPPX_MATH_FAST, ns->M, ps->ncount,
                                     ps->pairs, ctx->d_flags, ps->n, s);
           }));
    ps->last_ns = ns;
    ps->last_crit = criterion;
    ps->last_action = action;
    ps->last_kernel = ph.kernel;
    ps->last_fluid = ph.fluid_tag;
    ps->last_wall = ph.wall_tag;
    ps->searched = true;
    ps->masks_valid = true;
    return EPPX_OK;
}

int32_t do_momentum3(eppx_ps3d *ps, eppx_ns3d *ns, const Physics &ph) {
    eppx_ctx *ctx = ps->ctx;
    Walk3 w = make_walk3(ps, ns, ps->last_crit);
    if (ps->masks_valid) w.mask_mode = 2;
    HIP_OK(ctx, timed(ctx, "momentum", [&] {
               return launch3_momentum(w, ps->f(), ph, ctx->math_mode == EPPX_MATH_FAST, ps->n, ctx->stream);
           }));
    return EPPX_OK;
}

int32_t do_self3(eppx_ps3d *ps, int32_t action, const Physics3 &ph) {
    eppx_ctx *ctx = ps->ctx;
    const char *name = action == EPPX_SELF_CONTINUITY_EOS ? "continuity_eos" : "gravity_move";
    HIP_OK(ctx, timed(ctx, name, [&] { return launch3_self(ps->f(), ps->n, action, ph, ctx->stream); }));
    return EPPX_OK;
}

struct Staging3 {
    int32_t *ip = nullptr, *alive = nullptr, *cell = nullptr;
    float *fp = nullptr;
    void release() {
        dfree(ip);
        dfree(fp);
        dfree(alive);
        dfree(cell);
    }
};

}  // namespace

extern "C" {

int32_t eppx_domain3d_init(double gap, double first_x, double first_y, double first_z, double last_x, double last_y,
                           double last_z, eppx_domain3d *d) {
    if (!d) return EPPX_E_ARG;
    const float g = (float)gap;
    d->gap = g;
    d->gap_square = g * g;
    d->first_x = (float)first_x;
    d->first_y = (float)first_y;
    d->first_z = (float)first_z;
    d->last_x = (float)last_x;
    d->last_y = (float)last_y;
    d->last_z = (float)last_z;
    d->span_x = d->last_x - d->first_x;
    d->span_y = d->last_y - d->first_y;
    d->span_z = d->last_z - d->first_z;
    const float nx = floorf(d->span_x / g), ny = floorf(d->span_y / g), nz = floorf(d->span_z / g);
    if (!(nx >= 1.0f) || !(ny >= 1.0f) || !(nz >= 1.0f) || (double)nx * (double)ny * (double)nz > 2.0e9)
        return EPPX_E_ARG;
    d->n_x = (int32_t)nx;
    d->n_y = (int32_t)ny;
    d->n_z = (int32_t)nz;
    d->n = d->n_x * d->n_y * d->n_z;
    d->gap_x = d->span_x / (float)d->n_x;
    d->gap_y = d->span_y / (float)d->n_y;
    d->gap_z = d->span_z / (float)d->n_z;
    d->gap_x_inv = 1.0f / d->gap_x;
    d->gap_y_inv = 1.0f / d->gap_y;
    d->gap_z_inv = 1.0f / d->gap_z;
    return EPPX_OK;
}

int32_t eppx_domain3d_locate(const eppx_domain3d *d, float x, float y, float z, int32_t *inside, int32_t *pi,
                             int32_t *pj, int32_t *pk, int32_t *linear) {
    if (!d) return EPPX_E_ARG;
    const int in = d->first_x <= x && x <= d->last_x && d->first_y <= y && y <= d->last_y && d->first_z <= z &&
                   z <= d->last_z;
    int32_t i = (int32_t)floorf((x - d->first_x) * d->gap_x_inv) + 1;
    i = std::max(1, std::min(d->n_x, i));
    int32_t j = (int32_t)floorf((y - d->first_y) * d->gap_y_inv) + 1;
    j = std::max(1, std::min(d->n_y, j));
    int32_t k = (int32_t)floorf((z - d->first_z) * d->gap_z_inv) + 1;
    k = std::max(1, std::min(d->n_z, k));
    if (inside) *inside = in;
    if (pi) *pi = i;
    if (pj) *pj = j;
    if (pk) *pk = k;
    if (linear) *linear = i + d->n_x * ((j - 1) + d->n_y * (k - 1));
    return EPPX_OK;
}

int32_t eppx_ns3d_create(eppx_ctx *ctx, const eppx_domain3d *domain, const int32_t *adjacency, int32_t max_tag,
                         int32_t max_neighbour_number, eppx_ns3d **out) {
    if (!ctx || !domain || !adjacency || !out || max_tag < 1 || max_tag > kMaxTag || max_neighbour_number < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_ns3d_create: bad argument (1 <= max_tag <= 32)");
    const eppx_domain3d &d = *domain;
    if (d.n_x < 1 || d.n_y < 1 || d.n_z < 1 || (int64_t)d.n_x * d.n_y * d.n_z != d.n)
        return set_error(ctx, EPPX_E_ARG, "eppx_ns3d_create: inconsistent domain");
    if (ctx->closing) return set_error(ctx, EPPX_E_STATE, "eppx_ns3d_create: context is closed");
    eppx_ns3d *ns = new eppx_ns3d();
    ns->ctx = ctx;
    ctx->children += 1;
    ns->domain = d;
    ns->grid = Grid3{d.first_x, d.first_y, d.first_z, d.last_x, d.last_y, d.last_z, d.gap_x_inv, d.gap_y_inv,
                     d.gap_z_inv, d.gap_square, d.n_x, d.n_y, d.n_z, d.n};
    ns->max_tag = max_tag;
    ns->M = max_neighbour_number;
    for (int t = 1; t <= max_tag; ++t)
        for (int u = 1; u <= max_tag; ++u)
            if (adjacency[(t - 1) + (size_t)max_tag * (u - 1)] == 1) ns->adj[t] |= 1u << (u - 1);
    *out = ns;
    return EPPX_OK;
}

int32_t eppx_ns3d_destroy(eppx_ns3d *ns) {
    if (!ns) return EPPX_E_ARG;
    eppx_ctx *ctx = ns->ctx;
    delete ns;
    ctx->children -= 1;
    release_ctx(ctx);
    return EPPX_OK;
}

int32_t eppx_ns3d_cell_stencil(eppx_ns3d *ns, int32_t *count, int32_t *list) {
    if (!ns || !count || !list) return EPPX_E_ARG;
    const Grid3 &g = ns->grid;
    const size_t nc = (size_t)g.n_cells;
    for (int32_t I = 1; I <= g.n_cells; ++I) {
        // indexLinearToCartesian extended: i = mod1(I, n_x), j = mod1(cld(I, n_x), n_y), k = cld(I, n_x n_y)
        const int32_t c = I - 1;
        const int32_t i = c % g.n_x + 1, j = (c / g.n_x) % g.n_y + 1, k = c / (g.n_x * g.n_y) + 1;
        int32_t m = 0;
        for (int32_t di = -1; di <= 1; ++di)
            for (int32_t dj = -1; dj <= 1; ++dj)
                for (int32_t dk = -1; dk <= 1; ++dk) {
                    const int32_t ii = i + di, jj = j + dj, kk = k + dk;
                    if (ii >= 1 && ii <= g.n_x && jj >= 1 && jj <= g.n_y && kk >= 1 && kk <= g.n_z)
                        list[(size_t)c + nc * (size_t)(m++)] = ii + g.n_x * ((jj - 1) + g.n_y * (kk - 1));
                }
        for (int32_t r = m; r < 27; ++r) list[(size_t)c + nc * (size_t)r] = 0;
        count[c] = m;
    }
    return EPPX_OK;
}

int32_t eppx_ps3d_destroy(eppx_ps3d *ps);

int32_t eppx_ps3d_create(eppx_ctx *ctx, int32_t n_capacity, const eppx_layout *layout, eppx_ps3d **out) {
    if (!ctx || !layout || !out || n_capacity < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_create: bad argument");
    if (n_capacity >= (1 << 27))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_create: n_capacity must be below 2^27 (134,217,728)");
    *out = nullptr;
    const eppx_layout &L = *layout;
    if (L.n_int < 0 || L.n_float < 0 || L.max_neighbour_number < 0)
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_create: negative table width");
    auto fits = [](int32_t col, int64_t width, int32_t n) { return col < 0 || (col + width <= n); };
    const int64_t M = L.max_neighbour_number;
    if (!fits(L.Tag, 1, L.n_int) || !fits(L.IsMovable, 1, L.n_int) || !fits(L.nCount, 1, L.n_int) ||
        !fits(L.nIndex, M, L.n_int) || !fits(L.PositionVec, 3, L.n_float) || !fits(L.VelocityVec, 3, L.n_float) ||
        !fits(L.dVelocityVec, 3, L.n_float) || !fits(L.Mass, 1, L.n_float) || !fits(L.Volume, 1, L.n_float) ||
        !fits(L.Density, 1, L.n_float) || !fits(L.dDensity, 1, L.n_float) || !fits(L.Pressure, 1, L.n_float) ||
        !fits(L.H, 1, L.n_float) || !fits(L.nRVec, 3 * M, L.n_float) || !fits(L.nR, M, L.n_float) ||
        !fits(L.nW, M, L.n_float) || !fits(L.nDW, M, L.n_float) || !fits(L.nHInv, M, L.n_float))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_create: a column lies outside its table");
    if (!has(L.PositionVec) || !has(L.Tag))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_create: PositionVec and Tag are required");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    if (ctx->closing) return set_error(ctx, EPPX_E_STATE, "eppx_ps3d_create: context is closed");
    eppx_ps3d *ps = new eppx_ps3d();
    ps->ctx = ctx;
    ctx->children += 1;
    ps->cap = n_capacity;
    ps->layout = L;
    hipError_t e = alloc_fields3(ps->buf[0], n_capacity);
    if (e == hipSuccess) e = alloc_fields3(ps->buf[1], n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->geo, (size_t)n_capacity + kPad);
    if (e == hipSuccess) e = hipMemset(ps->geo, 0, ((size_t)n_capacity + kPad) * sizeof(float4));
    if (e == hipSuccess) e = dalloc(&ps->key, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->tmpslot, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->members, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->ncount, (size_t)n_capacity);
    if (e == hipSuccess) e = dalloc(&ps->n_alive, 2);
    if (e == hipSuccess) e = dalloc(&ps->pairs, kPairWords);
    if (e == hipSuccess) e = hipMemsetAsync(ps->pairs, 0, kPairWords * sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ps->n_alive, 0, 2 * sizeof(int32_t), ctx->stream);
    if (e != hipSuccess) {
        eppx_ps3d_destroy(ps);
        return hip_fail(ctx, e, "eppx_ps3d_create");
    }
    for (int c = 0; c < L.n_int; ++c)
        if (!owned_int3(L, c)) ps->pass_int_cols.push_back(c);
    for (int c = 0; c < L.n_float; ++c)
        if (!owned_float3(L, c)) ps->pass_float_cols.push_back(c);
    *out = ps;
    return EPPX_OK;
}

int32_t eppx_ps3d_destroy(eppx_ps3d *ps) {
    if (!ps) return EPPX_E_ARG;
    (void)hipSetDevice(ps->ctx->device);
    (void)hipStreamSynchronize(ps->ctx->stream);
    free_fields3(ps->buf[0]);
    free_fields3(ps->buf[1]);
    dfree(ps->geo);
    dfree(ps->counts); dfree(ps->start); dfree(ps->tile_sums); dfree(ps->cell_tags);
    dfree(ps->key); dfree(ps->tmpslot); dfree(ps->members); dfree(ps->ncount); dfree(ps->n_alive);
    dfree(ps->pairs);
    dfree(ps->masks);
    eppx_ctx *ctx = ps->ctx;
    delete ps;
    ctx->children -= 1;
    release_ctx(ctx);
    return EPPX_OK;
}

int32_t eppx_ps3d_upload(eppx_ps3d *ps, int32_t n_particles, const int32_t *is_alive, const int32_t *cell_index,
                         const int32_t *int_props, const float *float_props) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    if (n_particles < 0 || n_particles > ps->cap || !is_alive || !cell_index || (L.n_int && !int_props) ||
        (L.n_float && !float_props))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_upload: bad argument");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t cap = (size_t)ps->cap;
    hipStream_t s = ctx->stream;
    HIP_OK(ctx, hipStreamSynchronize(s));
    Staging3 st;
    hipError_t e = hipSuccess;
    if (cap) {
        e = dalloc(&st.ip, cap * (size_t)L.n_int);
        if (e == hipSuccess) e = dalloc(&st.fp, cap * (size_t)L.n_float);
        if (e == hipSuccess) e = dalloc(&st.alive, cap);
        if (e == hipSuccess) e = dalloc(&st.cell, cap);
        if (e == hipSuccess && L.n_int)
            e = hipMemcpyAsync(st.ip, int_props, cap * L.n_int * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess && L.n_float)
            e = hipMemcpyAsync(st.fp, float_props, cap * L.n_float * sizeof(float), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(st.alive, is_alive, cap * sizeof(int32_t), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(st.cell, cell_index, cap * sizeof(int32_t), hipMemcpyHostToDevice, s);
        for (int b = 0; b < 2 && e == hipSuccess; ++b)
            e = launch3_from_ref(ps->buf[b], columns3_of(ps), st.ip, st.fp, st.alive, st.cell, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    st.release();
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_ps3d_upload");
    ps->cur = 0;
    ps->n = n_particles;
    // H range and tags of the alive particles: the criterion's exact-safe
    // bounds, the one-H pair hinv and the tag tests the walk needs
    float hmin = INFINITY, hmax = 0.0f, mmin = INFINITY, mmax = 0.0f;
    bool bad = !has(L.H), mbad = !has(L.Mass);
    uint32_t tags = 0u;
    bool odd = false;
    for (int32_t r = 0; r < n_particles; ++r) {
        if (is_alive[r] != 1) continue;
        if (has(L.H)) {
            const float h = float_props[(size_t)r + cap * L.H];
            if (!(h > 0.0f) || !std::isfinite(h)) bad = true;
            hmin = std::min(hmin, h);
            hmax = std::max(hmax, h);
        }
        if (has(L.Mass)) {
            const float m = float_props[(size_t)r + cap * L.Mass];
            if (!(m > 0.0f) || !std::isfinite(m)) mbad = true;
            mmin = std::min(mmin, m);
            mmax = std::max(mmax, m);
        }
        const int32_t t = int_props[(size_t)r + cap * L.Tag];
        if (t >= 1 && t <= kMaxTag) tags |= 1u << (t - 1);
        else odd = true;
    }
    ps->tags_present = tags;
    ps->odd_tags = odd;
    symmetry_bounds(!bad, hmin, hmax, &ps->rmin2, &ps->rmax2);
    ps->hinv_u = uniform_hinv(!bad, hmin, hmax);
    ps->m_u = uniform_mass(!mbad, mmin, mmax);
    ps->pass_int.assign(ps->pass_int_cols.size() * cap, 0);
    ps->pass_float.assign(ps->pass_float_cols.size() * cap, 0.0f);
    for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
        memcpy(&ps->pass_int[k * cap], int_props + (size_t)ps->pass_int_cols[k] * cap, cap * sizeof(int32_t));
    for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
        memcpy(&ps->pass_float[k * cap], float_props + (size_t)ps->pass_float_cols[k] * cap, cap * sizeof(float));
    ps->searched = false;
    ps->masks_valid = false;
    ps->last_ns = nullptr;
    return EPPX_OK;
}

int32_t eppx_ps3d_download(eppx_ps3d *ps, eppx_ns3d *ns, int32_t *n_particles, int32_t *is_alive,
                           int32_t *cell_index, int32_t *int_props, float *float_props) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    if (!is_alive || !cell_index || (L.n_int && !int_props) || (L.n_float && !float_props))
        return set_error(ctx, EPPX_E_ARG, "eppx_ps3d_download: bad argument");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    const size_t cap = (size_t)ps->cap;
    if (n_particles) *n_particles = ps->n;
    if (!cap) return read_flags(ctx);
    hipStream_t s = ctx->stream;
    Staging3 st;
    hipError_t e = dalloc(&st.ip, cap * (size_t)L.n_int);
    if (e == hipSuccess) e = dalloc(&st.fp, cap * (size_t)L.n_float);
    if (e == hipSuccess) e = dalloc(&st.alive, cap);
    if (e == hipSuccess) e = dalloc(&st.cell, cap);
    if (e == hipSuccess) e = hipMemsetAsync(st.ip, 0, std::max<size_t>(1, cap * L.n_int) * sizeof(int32_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(st.fp, 0, std::max<size_t>(1, cap * L.n_float) * sizeof(float), s);
    if (e == hipSuccess) e = launch3_to_ref(ps->f(), ps->cap, columns3_of(ps), st.ip, st.fp, st.alive, st.cell, s);
    const int32_t M = L.max_neighbour_number;
    if (e == hipSuccess && ns && ps->searched && ps->last_ns == ns &&
        (has(L.nCount) || has(L.nIndex) || has(L.nRVec) || has(L.nR) || has(L.nW) || has(L.nDW) || has(L.nHInv))) {
        NeighbourExport3 ex{};
        ex.n_capacity = ps->cap;
        ex.M = M;
        ex.nCount = has(L.nCount) ? st.ip + cap * L.nCount : nullptr;
        ex.nIndex = has(L.nIndex) && M ? st.ip + cap * L.nIndex : nullptr;
        ex.nRVec = has(L.nRVec) && M ? st.fp + cap * L.nRVec : nullptr;
        ex.nR = has(L.nR) && M ? st.fp + cap * L.nR : nullptr;
        ex.nW = has(L.nW) && M ? st.fp + cap * L.nW : nullptr;
        ex.nDW = has(L.nDW) && M ? st.fp + cap * L.nDW : nullptr;
        ex.nHInv = has(L.nHInv) && M ? st.fp + cap * L.nHInv : nullptr;
        ex.kernel = ps->last_kernel;
        ex.with_kernel = ps->last_action == EPPX_SEARCH_KERNEL_CONTINUITY;
        ex.fluid_tag = ps->last_fluid;
        ex.wall_tag = ps->last_wall;
        e = launch3_export(make_walk3(ps, ns, ps->last_crit), ps->f(), ps->n, nullptr, nullptr, ps->ncount, ex, s);
    }
    if (e == hipSuccess && L.n_int)
        e = hipMemcpyAsync(int_props, st.ip, cap * L.n_int * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && L.n_float)
        e = hipMemcpyAsync(float_props, st.fp, cap * L.n_float * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(is_alive, st.alive, cap * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(cell_index, st.cell, cap * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    st.release();
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_ps3d_download");
    if (ps->pass_int.size() == ps->pass_int_cols.size() * cap)
        for (size_t k = 0; k < ps->pass_int_cols.size(); ++k)
            memcpy(int_props + (size_t)ps->pass_int_cols[k] * cap, &ps->pass_int[k * cap], cap * sizeof(int32_t));
    if (ps->pass_float.size() == ps->pass_float_cols.size() * cap)
        for (size_t k = 0; k < ps->pass_float_cols.size(); ++k)
            memcpy(float_props + (size_t)ps->pass_float_cols[k] * cap, &ps->pass_float[k * cap], cap * sizeof(float));
    return read_flags(ctx);
}

int32_t eppx_ps3d_alive_count(eppx_ps3d *ps, int32_t *out) {
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

int32_t eppx_search3d(eppx_ps3d *ps, eppx_ns3d *ns, int32_t criterion, int32_t action, const void *params,
                      size_t params_size) {
    if (!ps || !ns || ps->ctx != ns->ctx) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (criterion != EPPX_CRIT_DEFAULT && criterion != EPPX_CRIT_SYMMETRY)
        return set_error(ctx, EPPX_E_ARG, "eppx_search3d: unknown criterion");
    const eppx_layout &L = ps->layout;
    Physics3 ph{};
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        int32_t rc = physics3_arg(ctx, params, params_size, &ph);
        if (rc) return rc;
        rc = check_columns3(ps, {{L.H, "H"}, {L.VelocityVec, "VelocityVec"}, {L.Mass, "Mass"}, {L.Volume, "Volume"},
                                 {L.Density, "Density"}, {L.dDensity, "dDensity"}},
                            "sComputeKernelAndiContinuity");
        if (rc) return rc;
    } else if (action != EPPX_SEARCH_NONE) {
        return set_error(ctx, EPPX_E_ARG, "eppx_search3d: unknown action");
    }
    if (criterion == EPPX_CRIT_SYMMETRY && !has(L.H))
        return set_error(ctx, EPPX_E_ARG, "symmetrySearchCriterion needs the H column");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_search3(ps, ns, criterion, action, ph.p);
}

int32_t eppx_selfaction3d(eppx_ps3d *ps, int32_t action, const void *params, size_t params_size) {
    if (!ps) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    const eppx_layout &L = ps->layout;
    Physics3 ph{};
    int32_t rc = EPPX_OK;
    if (action == EPPX_SELF_CONTINUITY_EOS) {
        rc = physics3_arg(ctx, params, params_size, &ph);
        if (!rc)
            rc = check_columns3(ps, {{L.Density, "Density"}, {L.dDensity, "dDensity"}, {L.Volume, "Volume"},
                                     {L.Mass, "Mass"}, {L.Pressure, "Pressure"}},
                                "sContinuity");
    } else if (action == EPPX_SELF_GRAVITY_MOVE) {
        rc = physics3_arg(ctx, params, params_size, &ph);
        if (!rc)
            rc = check_columns3(ps, {{L.VelocityVec, "VelocityVec"}, {L.dVelocityVec, "dVelocityVec"}}, "sMomentum");
    } else {
        return set_error(ctx, EPPX_E_ARG, "eppx_selfaction3d: the 3D path runs sContinuity! and sMomentum! only");
    }
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_self3(ps, action, ph);
}

int32_t eppx_interaction3d(eppx_ps3d *ps, eppx_ns3d *ns, int32_t action, const void *params, size_t params_size) {
    if (!ps || !ns) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (action != EPPX_INTER_MOMENTUM)
        return set_error(ctx, EPPX_E_ARG, "eppx_interaction3d: the 3D path runs iMomentum! only");
    if (!ps->searched || ps->last_ns != ns)
        return set_error(ctx, EPPX_E_STATE, "eppx_interaction3d: needs a prior eppx_search3d on this neighbour system");
    Physics3 ph{};
    int32_t rc = physics3_arg(ctx, params, params_size, &ph);
    if (rc) return rc;
    const eppx_layout &L = ps->layout;
    rc = check_columns3(ps, {{L.H, "H"}, {L.VelocityVec, "VelocityVec"}, {L.dVelocityVec, "dVelocityVec"},
                             {L.Mass, "Mass"}, {L.Density, "Density"}, {L.Pressure, "Pressure"}},
                        "iMomentum");
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    return do_momentum3(ps, ns, ph.p);
}

int32_t eppx_step3d(eppx_ps3d *ps, eppx_ns3d *ns, const eppx_physics3d *phys, int32_t n_steps) {
    if (!ps || !ns || !phys || n_steps < 0) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns)
        return set_error(ctx, EPPX_E_STATE, "eppx_step3d: needs a prior eppx_search3d on this neighbour system");
    Physics3 ph{};
    int32_t rc = physics3_arg(ctx, phys, sizeof(*phys), &ph);
    if (rc) return rc;
    const eppx_layout &L = ps->layout;
    rc = check_columns3(ps, {{L.H, "H"}, {L.VelocityVec, "VelocityVec"}, {L.dVelocityVec, "dVelocityVec"},
                             {L.Mass, "Mass"}, {L.Volume, "Volume"}, {L.Density, "Density"},
                             {L.dDensity, "dDensity"}, {L.Pressure, "Pressure"}},
                        "eppx_step3d");
    if (rc) return rc;
    HIP_OK(ctx, hipSetDevice(ctx->device));
    // the loop body of main.jl:166-183: sContinuity!, iMomentum!, sMomentum!, search!
    for (int32_t k = 0; k < n_steps; ++k) {
        if ((rc = do_self3(ps, EPPX_SELF_CONTINUITY_EOS, ph))) return rc;
        if ((rc = do_momentum3(ps, ns, ph.p))) return rc;
        if ((rc = do_self3(ps, EPPX_SELF_GRAVITY_MOVE, ph))) return rc;
        if ((rc = do_search3(ps, ns, EPPX_CRIT_SYMMETRY, EPPX_SEARCH_KERNEL_CONTINUITY, ph.p))) return rc;
    }
    return EPPX_OK;
}

int32_t eppx_pair_count3d(eppx_ps3d *ps, eppx_ns3d *ns, int64_t *out) {
    if (!ps || !ns || !out) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns) return set_error(ctx, EPPX_E_STATE, "eppx_pair_count3d: no search yet");
    HIP_OK(ctx, hipSetDevice(ctx->device));
    unsigned long long v = 0;
    HIP_OK(ctx, read_pair_slots(ctx, ps->pairs, kPairLast, &v));
    *out = (int64_t)v;
    return read_flags(ctx);
}

int32_t eppx_pair_total3d(eppx_ps3d *ps, int64_t *out, int32_t reset) {
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

int32_t eppx_export_pairs3d(eppx_ps3d *ps, eppx_ns3d *ns, int64_t *offsets, int32_t *js, int64_t capacity) {
    if (!ps || !ns || !offsets || (capacity > 0 && !js)) return EPPX_E_ARG;
    eppx_ctx *ctx = ps->ctx;
    if (!ps->searched || ps->last_ns != ns)
        return set_error(ctx, EPPX_E_STATE, "eppx_export_pairs3d: no search yet");
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
        snprintf(buf, sizeof(buf), "eppx_export_pairs3d: %lld pairs exceed capacity %lld", (long long)total,
                 (long long)capacity);
        return set_error(ctx, EPPX_E_CAPACITY, buf);
    }
    if (!total) return read_flags(ctx);
    std::vector<int64_t> row_start((size_t)na);
    for (int32_t k = 0; k < na; ++k) row_start[k] = by_row[(size_t)ids[k]];
    int64_t *d_rs = nullptr;
    int32_t *d_js = nullptr;
    HIP_OK(ctx, dalloc(&d_rs, (size_t)na));
    hipError_t e = dalloc(&d_js, (size_t)total);
    NeighbourExport3 ex{};
    ex.n_capacity = ps->cap;
    if (e == hipSuccess) e = hipMemcpyAsync(d_rs, row_start.data(), na * sizeof(int64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch3_export(make_walk3(ps, ns, ps->last_crit), ps->f(), ps->n, d_rs, d_js, ps->ncount, ex, s);
    if (e == hipSuccess) e = hipMemcpyAsync(js, d_js, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_rs);
    dfree(d_js);
    if (e != hipSuccess) return hip_fail(ctx, e, "eppx_export_pairs3d");
    return read_flags(ctx);
}

}  // extern "C"

// Device code of libeppx: cell-list build, pair kernels, self kernels and
// the reference-layout conversion kernels.  gfx950, wave64.
//
// Floating point: the library is compiled with -ffp-contract=off and
// correctly rounded f32 divide / sqrt, so the EXACT variants below evaluate
// each expression in the reference's written order (left-to-right products,
// no FMA).  The FAST variants contract explicitly with fmaf and use the
// hardware reciprocal / reciprocal-sqrt; they never touch the criterion.
#include <math.h>

#include "eppx_internal.h"

namespace eppx {
namespace {

constexpr int kRankLimit = 1024;  // largest cell ordered by particle index

// sigma of the 2D kernels: FT(10/7/pi) (CubicSpline.jl:14), FT(7/4/pi) (WendlandC2.jl:13)
constexpr float kSigmaCubic = (float)(10.0 / 7.0 / 3.14159265358979323846);
constexpr float kSigmaWendland = (float)(7.0 / 4.0 / 3.14159265358979323846);

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Mean.harmonic / invharmonic (src/Math/Mean.jl:28-34)
__device__ __forceinline__ float harmonic(float a, float b) { return 2.0f * a * b / (a + b); }
__device__ __forceinline__ float invharmonic(float a, float b) { return (a + b) / (2.0f * a * b); }

// _gradient of CubicSpline.jl:41-55 / WendlandC2.jl:37-47 (2D), written order.
__device__ __forceinline__ float grad_exact(int kernel, float r, float hinv) {
    const float q = r * hinv;
    const float h3 = hinv * hinv * hinv;
    if (kernel == EPPX_KERNEL_CUBIC) {
        if (q < 1.0f) return kSigmaCubic * h3 * q * (3.0f * q - 4.0f) * 0.75f;
        if (q < 2.0f) {
            const float t = 2.0f - q;
            return -kSigmaCubic * h3 * t * t * 0.75f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return -kSigmaWendland * h3 * 0.625f * q * (t * t * t);
    }
    return 0.0f;
}

// _value of CubicSpline.jl:17-31 / WendlandC2.jl:16-26 (2D), written order.
__device__ __forceinline__ float value_exact(int kernel, float r, float hinv) {
    const float q = r * hinv;
    if (kernel == EPPX_KERNEL_CUBIC) {
        if (q < 1.0f) return kSigmaCubic * (hinv * hinv) * (3.0f * q * q * (q - 2.0f) + 4.0f) * 0.25f;
        if (q < 2.0f) {
            const float t = 2.0f - q;
            return kSigmaCubic * (hinv * hinv) * (t * t * t) * 0.25f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        const float t = 2.0f - q;
        const float t2 = t * t;
        return kSigmaWendland * (t2 * t2) * (hinv * hinv) * (1.0f + 2.0f * q) * 0.0625f;
    }
    return 0.0f;
}

// Same functions, contracted: dw = sigma h^3 * poly(q).
__device__ __forceinline__ float grad_fast(int kernel, float r, float hinv) {
    const float q = r * hinv;
    const float h3 = hinv * hinv * hinv;
    if (kernel == EPPX_KERNEL_CUBIC) {
        if (q < 1.0f) return (kSigmaCubic * 0.75f) * h3 * q * fmaf(3.0f, q, -4.0f);
        if (q < 2.0f) {
            const float t = 2.0f - q;
            return (-kSigmaCubic * 0.75f) * h3 * (t * t);
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return (-kSigmaWendland * 0.625f) * h3 * q * (t * t * t);
    }
    return 0.0f;
}

// Equation of state of the driver's sContinuity!: linear c0^2 (rho - rho0) + p0
// (main.jl:64) or Tait B ((rho/rho0)^7 - 1) + p0 with power(x, Val(7)) =
// (x*x*x)^2 * x (src/Math/OptimizedFunction.jl:32-34).
__device__ __forceinline__ float eos(const Physics &ph, float rho) {
    if (ph.eos == EPPX_EOS_LINEAR) {
        const float c02 = ph.c0 * ph.c0;
        return c02 * (rho - ph.rho0) + ph.p0;
    }
    const float x = rho / ph.rho0;
    const float t = x * x * x;
    return ph.tait_b * (t * t * x - 1.0f) + ph.p0;
}

// ------------------------------------------------------------------ walk
// Visits the accepted neighbours of sorted slot i in the reference's list
// order: neighbour cells in device_initializeNeighbourSystem! order (x offset
// outer, y offset inner; NeighbourSystemBase.jl:37-47), then ascending
// particle index inside a cell (the order of a serial insert).  The test is
// device_findNeighbourParticlesFromCells! (NeighbourSearch.jl:189-222):
// I != J, adjacency, dx = x_I - x_J, dr2 = dx*dx + dy*dy, criterion.
template <class Visit>
__device__ __forceinline__ void walk(const Walk &w, int i, float2 xi, int ti, float hi, Visit &&visit) {
    const int c = w.cell[i] - 1;
    const int cx = c % w.g.n_x, cy = c / w.g.n_x;
    const uint32_t adj = (ti >= 1 && ti <= kMaxTag) ? w.adj[ti] : 0u;
    for (int di = -1; di <= 1; ++di) {
        const int ii = cx + di;
        if (ii < 0 || ii >= w.g.n_x) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = cy + dj;
            if (jj < 0 || jj >= w.g.n_y) continue;
            const int k = ii + w.g.n_x * jj;
            const int b = w.start[k], e = w.start[k + 1];
            for (int j = b; j < e; ++j) {
                if (j == i) continue;
                const int tj = w.tag[j];
                if (tj < 1 || tj > kMaxTag || !((adj >> (tj - 1)) & 1u)) continue;
                const float2 xj = w.pos_nb[j];
                const float dx = xi.x - xj.x;
                const float dy = xi.y - xj.y;
                const float dr2 = dx * dx + dy * dy;
                bool ok;
                if (w.criterion == EPPX_CRIT_DEFAULT) {
                    ok = dr2 <= w.g.gap_square;  // defaultSearchCriterion
                } else if (dr2 > w.rmax2) {
                    ok = false;  // beyond (2 max h)^2 with margin: rejected by the exact test too
                } else if (dr2 < w.rmin2) {
                    ok = true;   // inside (2 min h)^2 with margin: accepted by the exact test too
                } else {
                    const float h2 = harmonic(hi, w.h[j]) * 2.0f;  // symmetrySearchCriterion
                    ok = dr2 <= h2 * h2;
                }
                if (ok) visit(j, tj, dx, dy, dr2);
            }
        }
    }
}

__device__ __forceinline__ int block_sum_i32(int v, int *lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) lds[wave] = v;
    __syncthreads();
    int total = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) total += lds[k];
    return total;  // valid in thread 0
}

// ------------------------------------------------------------- fills
__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_reset_search(int32_t *counts, int32_t n, unsigned long long *pairs_last) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) counts[i] = 0;
    if (i == 0) *pairs_last = 0ull;
}

// ------------------------------------------------------------- cell list
// device_insertParticlesIntoCells! (NeighbourSearch.jl:12-47): cell of each
// alive particle (cached for immovable ones), kill outside the domain.  The
// key is the 0-based cell, n_cells for dead particles.  Counts are taken with
// one atomic per distinct key in the wave (the slots of a sorted wave share
// few cells), which also yields each particle's provisional slot.
__global__ __launch_bounds__(kBlock) void k_key_count(Fields f, int32_t n, Grid g, int32_t *counts, int32_t *key,
                                                       int32_t *tmpslot, uint32_t *flags, int32_t max_per_cell) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = s < n;
    int32_t k = -1;
    if (valid) {
        k = g.n_cells;
        if (f.alive[s] == 1) {
            int32_t c = f.cell[s];
            // a cached cell outside this grid (caller data, or another domain) is recomputed
            if (c < 0 || c > g.n_cells) c = -1;
            if (!(f.mov[s] == 0 && c > 0)) {
                const float2 x = f.pos[s];
                if (g.first_x <= x.x && x.x <= g.last_x && g.first_y <= x.y && x.y <= g.last_y) {
                    int32_t i = (int32_t)floorf((x.x - g.first_x) * g.gap_x_inv) + 1;
                    i = max(1, min(g.n_x, i));
                    int32_t j = (int32_t)floorf((x.y - g.first_y) * g.gap_y_inv) + 1;
                    j = max(1, min(g.n_y, j));
                    c = i + g.n_x * (j - 1);
                    f.cell[s] = c;
                } else {
                    f.alive[s] = 0;
                    f.cell[s] = 0;
                    c = 0;
                }
            }
            if (c > 0) k = c - 1;
        }
        key[s] = k;
    }
    unsigned long long active = __ballot(valid);
    int32_t slot = 0;
    while (active) {
        const int leader = __ffsll(active) - 1;
        const int32_t lk = __shfl(k, leader);
        const unsigned long long same = __ballot(k == lk) & active;
        int32_t base = 0;
        if (lane_id() == leader) {
            const int32_t cnt = __popcll(same);
            base = atomicAdd(&counts[lk], cnt);
            if (lk < g.n_cells && base + cnt > max_per_cell) atomicOr(flags, kFlagCellOverflow);
        }
        base = __shfl(base, leader);
        if (valid && k == lk) slot = base + __popcll(same & ((1ull << lane_id()) - 1ull));
        active &= ~same;
    }
    if (valid) tmpslot[s] = slot;
}

// Exclusive scan of cell counts: per-tile sums, one-block scan of the tile
// sums, per-tile rescan.  No inter-block waiting.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int32_t *in, int32_t n, int32_t *tile_sums) {
    __shared__ int lds[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int v = 0;
    for (int k = 0; k < kScanTile / kBlock; ++k) {
        const int64_t idx = base + (int64_t)k * kBlock + threadIdx.x;
        if (idx < n) v += in[idx];
    }
    const int t = block_sum_i32(v, lds);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = t;
}

__device__ __forceinline__ int block_inclusive_scan(int v, int *lds) {
    const int lane = lane_id(), wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    if (wave == 0) {
        int x = lane < n_waves ? lds[lane] : 0;
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(x, off);
            if (lane >= off) x += t;
        }
        if (lane < n_waves) lds[lane] = x;
    }
    __syncthreads();
    if (wave > 0) v += lds[wave - 1];
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t *tile_sums, int32_t n_tiles, int32_t *start, int32_t n) {
    __shared__ int lds[16];
    __shared__ int chunk_total;
    int carry = 0;
    for (int base = 0; base < n_tiles; base += 1024) {
        const int idx = base + threadIdx.x;
        const int v = idx < n_tiles ? tile_sums[idx] : 0;
        const int inc = block_inclusive_scan(v, lds);
        if (idx < n_tiles) tile_sums[idx] = carry + inc - v;
        if (threadIdx.x == 1023) chunk_total = inc;
        __syncthreads();
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[n] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const int32_t *in, int32_t n, const int32_t *tile_offsets,
                                                      int32_t *start) {
    __shared__ int lds[kBlock / 64];
    constexpr int kPer = kScanTile / kBlock;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kPer;
    int v[kPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    const int inc = block_inclusive_scan(sum, lds);
    int run = tile_offsets[blockIdx.x] + inc - sum;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (base + k < n) start[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_members(Fields f, int32_t n, int32_t n_cells, const int32_t *key,
                                                          const int32_t *tmpslot, const int32_t *start,
                                                          int32_t *members) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int32_t k = key[s];
    if (k < n_cells) members[start[k] + tmpslot[s]] = f.id[s];
}

// Final position of each slot: cell start + number of cell members with a
// smaller particle index (the order a serial insert gives); dead particles
// keep their provisional slot after the alive ones.  Moves every field.
__global__ __launch_bounds__(kBlock) void k_permute(Fields cur, Fields nxt, float2 *pos_nb, int32_t n, int32_t cap,
                                                     int32_t n_cells, const int32_t *key, const int32_t *tmpslot,
                                                     const int32_t *start, const int32_t *members,
                                                     int32_t *n_alive, uint32_t *flags) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= cap) return;
    if (s == 0) *n_alive = start[n_cells];
    int32_t dest = s;
    int32_t alive = cur.alive[s];
    if (s < n) {
        const int32_t k = key[s];
        const int32_t b = start[k];
        if (k < n_cells) {
            const int32_t e = start[k + 1];
            if (e - b <= kRankLimit) {
                const int32_t me = cur.id[s];
                int32_t r = 0;
                for (int32_t q = b; q < e; ++q) r += members[q] < me;
                dest = b + r;
            } else {
                dest = b + tmpslot[s];
                atomicOr(flags, kFlagRankFallback);
            }
            alive = 1;
        } else {
            dest = b + tmpslot[s];
            alive = cur.alive[s] == 1 ? 0 : cur.alive[s];
        }
    }
    const float2 x = cur.pos[s];
    nxt.pos[dest] = x;
    pos_nb[dest] = x;
    nxt.vel[dest] = cur.vel[s];
    nxt.dvel[dest] = cur.dvel[s];
    nxt.mass[dest] = cur.mass[s];
    nxt.vol[dest] = cur.vol[s];
    nxt.rho[dest] = cur.rho[s];
    nxt.drho[dest] = cur.drho[s];
    nxt.p[dest] = cur.p[s];
    nxt.h[dest] = cur.h[s];
    nxt.wv[dest] = cur.wv[s];
    nxt.wvrho[dest] = cur.wvrho[s];
    nxt.tag[dest] = cur.tag[s];
    nxt.mov[dest] = cur.mov[s];
    nxt.id[dest] = cur.id[s];
    nxt.cell[dest] = cur.cell[s];
    nxt.alive[dest] = alive;
}

// ------------------------------------------------------------ pair kernels
// Search pass: find the neighbours of every alive particle and run the fused
// action (EPPX_SEARCH_KERNEL_CONTINUITY: sComputeKernelAndiContinuity!,
// main.jl:74-90 -> iValueGradient! Library/Kernel.jl:114-128,
// iClassicContinuity! / iBalancedContinuity! Continuity.jl:23-47).
template <bool FAST, int ACTION>
__global__ __launch_bounds__(kBlock) void k_search(Walk w, Fields f, Physics ph, int32_t M, int32_t *ncount,
                                                    unsigned long long *pairs_last,
                                                    unsigned long long *pairs_total, uint32_t *flags) {
    __shared__ int lds[kBlock / 64];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int na = *w.n_alive;
    int cnt = 0;
    if (i < na) {
        const float2 xi = w.pos_nb[i];
        const int ti = w.tag[i];
        const float hi = w.h[i];
        if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) {
            const float2 ui = f.vel[i];
            const float rho_i = f.rho[i];
            float acc = f.drho[i];
            const bool i_fluid = ti == ph.fluid_tag, i_wall = ti == ph.wall_tag;
            walk(w, i, xi, ti, hi, [&](int j, int tj, float dx, float dy, float dr2) {
                ++cnt;
                const bool ff = i_fluid && tj == ph.fluid_tag;
                const bool mixed = (i_fluid && tj == ph.wall_tag) || (i_wall && tj == ph.fluid_tag);
                if (!(ff || mixed)) return;
                const float hj = w.h[j];
                const float2 uj = f.vel[j];
                if (!FAST) {
                    const float r = sqrtf(dr2);
                    const float hinv = invharmonic(hi, hj);
                    const float dw = grad_exact(ph.kernel, r, hinv);
                    float vdotx = 0.0f;
                    vdotx += dx * (ui.x - uj.x);
                    vdotx += dy * (ui.y - uj.y);
                    const float term = ff ? f.mass[j] * vdotx * dw / r : rho_i * f.vol[j] * vdotx * dw / r;
                    acc += term;
                } else {
                    const float rinv = frsq(dr2);
                    const float r = dr2 * rinv;
                    const float hinv = (hi + hj) * frcp(2.0f * hi * hj);
                    const float dw = grad_fast(ph.kernel, r, hinv);
                    const float vdotx = fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x));
                    const float mj = ff ? f.mass[j] : rho_i * f.vol[j];
                    acc = fmaf(mj * vdotx, dw * rinv, acc);
                }
            });
            f.drho[i] = acc;
        } else {
            walk(w, i, xi, ti, hi, [&](int, int, float, float, float) { ++cnt; });
        }
        ncount[i] = cnt;
        if (cnt > M) atomicOr(flags, kFlagNeighbourOverflow);
    }
    const int total = block_sum_i32(cnt, lds);
    if (threadIdx.x == 0 && total) {
        atomicAdd(pairs_last, (unsigned long long)total);
        atomicAdd(pairs_total, (unsigned long long)total);
    }
}

// Interaction pass over the last search's neighbour lists
// (device_interaction!, Action.jl:121-138).
//  EPPX_INTER_MOMENTUM: iMomentum! main.jl:107-118 -> iClassicPressure! /
//    iBalancedPressure! (Pressure.jl:24-60), iClassicViscosity! /
//    sArtificialViscosity! (Viscosity.jl:10-52).
//  EPPX_INTER_FILTER: iFilter main.jl:128-134 -> iKernelFilter! (Filter.jl:10-22).
//  EPPX_INTER_ADD_PER_NEIGHBOUR: field[I] += value per neighbour.
template <bool FAST, int ACTION>
__global__ __launch_bounds__(kBlock) void k_interaction(Walk w, Fields f, Physics ph, eppx_field_op op) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *w.n_alive) return;
    const float2 xi = w.pos_nb[i];
    const int ti = w.tag[i];
    const float hi = w.h[i];
    if (ACTION == EPPX_INTER_MOMENTUM) {
        if (ti != ph.fluid_tag) return;
        const float2 ui = f.vel[i];
        const float rho_i = f.rho[i], p_i = f.p[i];
        float2 acc = f.dvel[i];
        const float pr_i = FAST ? p_i * frcp(rho_i * rho_i) : 0.0f;
        walk(w, i, xi, ti, hi, [&](int j, int tj, float dx, float dy, float dr2) {
            const bool classic = tj == ph.fluid_tag;
            if (!classic && tj != ph.wall_tag) return;
            const float hj = w.h[j];
            const float rho_j = f.rho[j], p_j = f.p[j], m_j = f.mass[j];
            const float2 uj = f.vel[j];
            const float dux = ui.x - uj.x, duy = ui.y - uj.y;
            if (!FAST) {
                const float r = sqrtf(dr2);
                const float hinv = invharmonic(hi, hj);
                const float dw = grad_exact(ph.kernel, r, hinv);
                float prho2 = classic ? p_i / (rho_i * rho_i) + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
                prho2 += fabsf(prho2) * ph.coefficient;
                prho2 *= -m_j * dw / r;
                acc.x += prho2 * dx;
                acc.y += prho2 * dy;
                if (ph.visc == EPPX_VISC_CLASSIC) {
                    const float h = 1.0f / hinv;
                    const float az = r * r + 0.01f * h * h;
                    const float force = 2.0f * ph.mu * m_j * r * dw / (rho_i * rho_j * az);
                    acc.x += force * dux;
                    acc.y += force * duy;
                } else {
                    float vdotx = 0.0f;
                    vdotx += dx * dux;
                    vdotx += dy * duy;
                    if (vdotx > 0.0f) return;
                    const float rho = (rho_i + rho_j) * 0.5f;
                    const float h = 1.0f / hinv;
                    const float az = r * r + 0.01f * h * h;
                    const float phi = vdotx / (hinv * az);
                    float force = (-ph.alpha * ph.av_c + ph.beta * phi) * phi / rho;
                    force *= -m_j / (rho_i * rho_j * r) * dw;
                    acc.x += force * dx;
                    acc.y += force * dy;
                }
            } else {
                const float rinv = frsq(dr2);
                const float r = dr2 * rinv;
                const float hinv = (hi + hj) * frcp(2.0f * hi * hj);
                const float dw = grad_fast(ph.kernel, r, hinv);
                const float rirj = rho_i * rho_j;
                float prho2 = classic ? fmaf(p_j, frcp(rho_j * rho_j), pr_i) : (p_i + p_j) * frcp(rirj);
                prho2 = fmaf(fabsf(prho2), ph.coefficient, prho2);
                const float mdw = -m_j * dw;
                prho2 *= mdw * rinv;
                acc.x = fmaf(prho2, dx, acc.x);
                acc.y = fmaf(prho2, dy, acc.y);
                if (ph.visc == EPPX_VISC_CLASSIC) {
                    const float h = frcp(hinv);
                    const float az = fmaf(r, r, 0.01f * h * h);
                    const float force = (2.0f * ph.mu) * m_j * r * dw * frcp(rirj * az);
                    acc.x = fmaf(force, dux, acc.x);
                    acc.y = fmaf(force, duy, acc.y);
                } else {
                    const float vdotx = fmaf(dy, duy, dx * dux);
                    if (vdotx > 0.0f) return;
                    const float h = frcp(hinv);
                    const float az = fmaf(r, r, 0.01f * h * h);
                    const float phi = vdotx * frcp(hinv * az);
                    const float force = fmaf(ph.beta, phi, -ph.alpha * ph.av_c) * phi * frcp(0.5f * (rho_i + rho_j)) *
                                        (mdw * frcp(rirj * r));
                    acc.x = fmaf(force, dx, acc.x);
                    acc.y = fmaf(force, dy, acc.y);
                }
            }
        });
        f.dvel[i] = acc;
    } else if (ACTION == EPPX_INTER_FILTER) {
        if (ti != ph.fluid_tag) return;
        float wv = f.wv[i], wvrho = f.wvrho[i];
        walk(w, i, xi, ti, hi, [&](int j, int tj, float, float, float dr2) {
            if (tj != ph.fluid_tag) return;
            const float r = sqrtf(dr2);
            const float wgt = value_exact(ph.kernel, r, invharmonic(hi, w.h[j]));
            wv += wgt * f.vol[j];
            wvrho += wgt * f.mass[j];
        });
        f.wv[i] = wv;
        f.wvrho[i] = wvrho;
    } else {
        float *col = nullptr;
        switch (op.field) {
            case EPPX_FIELD_POSITION: col = reinterpret_cast<float *>(f.pos) + op.component; break;
            case EPPX_FIELD_VELOCITY: col = reinterpret_cast<float *>(f.vel) + op.component; break;
            case EPPX_FIELD_DVELOCITY: col = reinterpret_cast<float *>(f.dvel) + op.component; break;
            case EPPX_FIELD_MASS: col = f.mass; break;
            case EPPX_FIELD_VOLUME: col = f.vol; break;
            case EPPX_FIELD_DENSITY: col = f.rho; break;
            case EPPX_FIELD_DDENSITY: col = f.drho; break;
            default: col = f.p; break;
        }
        const int stride = op.field <= EPPX_FIELD_DVELOCITY ? 2 : 1;
        float v = col[(int64_t)i * stride];
        walk(w, i, xi, ti, hi, [&](int, int, float, float, float) { v += op.value; });
        col[(int64_t)i * stride] = v;
    }
}

// ------------------------------------------------------------ self kernels
// device_selfaction! (Action.jl:39-51) with
//  EPPX_SELF_CONTINUITY_EOS: sContinuity! main.jl:92-105 -> Continuity.jl:10-21,
//    sVolume! State.jl:10-19, eos;
//  EPPX_SELF_GRAVITY_MOVE: sMomentum! main.jl:120-126 -> sGravity! State.jl:35-47,
//    sAccelerateMove! Motion.jl:61-75;
//  EPPX_SELF_FILTER: sFilter main.jl:136-144 -> sKernelFilter! Filter.jl:24-38;
//  EPPX_SELF_SCALE_FIELD: field[I] *= value.
template <int ACTION>
__global__ __launch_bounds__(kBlock) void k_self(Fields f, int32_t n, Physics ph, eppx_field_op op) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const int tag = f.tag[s];
    if (ACTION == EPPX_SELF_CONTINUITY_EOS) {
        if (tag != ph.fluid_tag && tag != ph.wall_tag) return;
        const float rho = f.rho[s] + f.drho[s] * ph.dt;
        f.rho[s] = rho;
        f.drho[s] = 0.0f;
        f.vol[s] = f.mass[s] / rho;
        f.p[s] = eos(ph, rho);
    } else if (ACTION == EPPX_SELF_GRAVITY_MOVE) {
        if (tag != ph.fluid_tag) return;
        float2 du = f.dvel[s];
        du.x += ph.gx;
        du.y += ph.gy;
        float2 x = f.pos[s], u = f.vel[s];
        x.x += (u.x + du.x * ph.dt * 0.5f) * ph.dt;
        u.x += du.x * ph.dt;
        x.y += (u.y + du.y * ph.dt * 0.5f) * ph.dt;
        u.y += du.y * ph.dt;
        f.pos[s] = x;
        f.vel[s] = u;
        f.dvel[s] = make_float2(0.0f, 0.0f);
    } else if (ACTION == EPPX_SELF_FILTER) {
        if (tag != ph.fluid_tag) return;
        const float hinv = 1.0f / f.h[s];
        const float w0 = (ph.kernel == EPPX_KERNEL_CUBIC ? kSigmaCubic : kSigmaWendland) * (hinv * hinv);
        const float wv = f.wv[s] + w0 * f.vol[s];
        const float wvrho = f.wvrho[s] + w0 * f.mass[s];
        const float rho = wvrho / wv;
        f.rho[s] = rho;
        f.wv[s] = 0.0f;
        f.wvrho[s] = 0.0f;
        f.vol[s] = f.mass[s] / rho;
        f.p[s] = eos(ph, rho);
    } else {
        float *col;
        int stride = 1;
        switch (op.field) {
            case EPPX_FIELD_POSITION: col = reinterpret_cast<float *>(f.pos) + op.component; stride = 2; break;
            case EPPX_FIELD_VELOCITY: col = reinterpret_cast<float *>(f.vel) + op.component; stride = 2; break;
            case EPPX_FIELD_DVELOCITY: col = reinterpret_cast<float *>(f.dvel) + op.component; stride = 2; break;
            case EPPX_FIELD_MASS: col = f.mass; break;
            case EPPX_FIELD_VOLUME: col = f.vol; break;
            case EPPX_FIELD_DENSITY: col = f.rho; break;
            case EPPX_FIELD_DDENSITY: col = f.drho; break;
            default: col = f.p; break;
        }
        col[(int64_t)s * stride] *= op.value;
    }
}

// ------------------------------------------------------- layout conversion
// Neighbour lists of the last search, written as a CSR by particle index
// and / or into the reference's per-particle neighbour columns
// (nCount, nIndex, nRVec, nR, nW, nDW, nHInv; NeighbourSearch.jl:215-220,
// Library/Kernel.jl:123-126).
__global__ __launch_bounds__(kBlock) void k_export(Walk w, Fields f, const int64_t *row_start, int32_t *js,
                                                    const int32_t *ncount, NeighbourExport ex) {
    const int32_t fluid_tag = ex.fluid_tag, wall_tag = ex.wall_tag;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *w.n_alive) return;
    const int row = f.id[i];
    const int ti = w.tag[i];
    const float hi = w.h[i];
    const size_t cap = (size_t)ex.n_capacity;
    const int64_t o = row_start ? row_start[i] : 0;
    const int limit = js ? ncount[i] : 0;  // the row length the search counted
    int k = 0;
    walk(w, i, w.pos_nb[i], ti, hi, [&](int j, int tj, float dx, float dy, float dr2) {
        if (js && k < limit) js[o + k] = f.id[j];
        if (k < ex.M) {
            if (ex.nIndex) ex.nIndex[row + cap * k] = f.id[j] + 1;
            if (ex.nRVec) {
                ex.nRVec[row + cap * (2 * k)] = dx;
                ex.nRVec[row + cap * (2 * k + 1)] = dy;
            }
            const float r = sqrtf(dr2);
            if (ex.nR) ex.nR[row + cap * k] = r;
            const bool kernel_pair = (ti == fluid_tag && (tj == fluid_tag || tj == wall_tag)) ||
                                     (ti == wall_tag && tj == fluid_tag);
            if (ex.with_kernel && kernel_pair) {
                const float hinv = invharmonic(hi, w.h[j]);
                if (ex.nHInv) ex.nHInv[row + cap * k] = hinv;
                if (ex.nW) ex.nW[row + cap * k] = value_exact(ex.kernel, r, hinv);
                if (ex.nDW) ex.nDW[row + cap * k] = grad_exact(ex.kernel, r, hinv);
            }
        }
        ++k;
    });
    if (ex.nCount) ex.nCount[row] = k;
}

__global__ __launch_bounds__(kBlock) void k_from_ref(Fields f, int32_t cap, Columns c, const int32_t *ip,
                                                      const float *fp, const int32_t *is_alive,
                                                      const int32_t *cell_index) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= cap) return;
    const size_t C = (size_t)cap;
    auto fcol = [&](int col) { return col >= 0 ? fp[s + C * col] : 0.0f; };
    auto icol = [&](int col) { return col >= 0 ? ip[s + C * col] : 0; };
    f.pos[s] = make_float2(fcol(c.PositionVec), c.PositionVec >= 0 ? fcol(c.PositionVec + 1) : 0.0f);
    f.vel[s] = make_float2(fcol(c.VelocityVec), c.VelocityVec >= 0 ? fcol(c.VelocityVec + 1) : 0.0f);
    f.dvel[s] = make_float2(fcol(c.dVelocityVec), c.dVelocityVec >= 0 ? fcol(c.dVelocityVec + 1) : 0.0f);
    f.mass[s] = fcol(c.Mass);
    f.vol[s] = fcol(c.Volume);
    f.rho[s] = fcol(c.Density);
    f.drho[s] = fcol(c.dDensity);
    f.p[s] = fcol(c.Pressure);
    f.h[s] = fcol(c.H);
    f.wv[s] = fcol(c.SumWeight);
    f.wvrho[s] = fcol(c.SumWeightedDensity);
    f.tag[s] = icol(c.Tag);
    f.mov[s] = icol(c.IsMovable);
    f.id[s] = s;
    f.alive[s] = is_alive[s];
    f.cell[s] = cell_index[s];
}

__global__ __launch_bounds__(kBlock) void k_to_ref(Fields f, int32_t cap, Columns c, int32_t *ip, float *fp,
                                                    int32_t *is_alive, int32_t *cell_index) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= cap) return;
    const size_t C = (size_t)cap;
    const int row = f.id[s];
    auto put = [&](int col, float v) {
        if (col >= 0) fp[row + C * col] = v;
    };
    auto put2 = [&](int col, float2 v) {
        if (col >= 0) {
            fp[row + C * col] = v.x;
            fp[row + C * (col + 1)] = v.y;
        }
    };
    put2(c.PositionVec, f.pos[s]);
    put2(c.VelocityVec, f.vel[s]);
    put2(c.dVelocityVec, f.dvel[s]);
    put(c.Mass, f.mass[s]);
    put(c.Volume, f.vol[s]);
    put(c.Density, f.rho[s]);
    put(c.dDensity, f.drho[s]);
    put(c.Pressure, f.p[s]);
    put(c.H, f.h[s]);
    put(c.SumWeight, f.wv[s]);
    put(c.SumWeightedDensity, f.wvrho[s]);
    if (c.Tag >= 0) ip[row + C * c.Tag] = f.tag[s];
    if (c.IsMovable >= 0) ip[row + C * c.IsMovable] = f.mov[s];
    is_alive[row] = f.alive[s];
    cell_index[row] = f.cell[s];
}

// device_initializeNeighbourSystem! (NeighbourSystemBase.jl:28-48), 1-based.
__global__ __launch_bounds__(kBlock) void k_stencil(Grid g, int32_t *count, int32_t *list) {
    const int I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= g.n_cells) return;
    const int i = I % g.n_x + 1, j = I / g.n_x + 1;
    int k = 0;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj) {
            const int ii = i + di, jj = j + dj;
            if (ii >= 1 && ii <= g.n_x && jj >= 1 && jj <= g.n_y) {
                list[I + (size_t)g.n_cells * k] = ii + g.n_x * (jj - 1);
                ++k;
            }
        }
    count[I] = k;
}

__global__ __launch_bounds__(kBlock) void k_finite(Fields f, int32_t n, uint32_t *bad) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const float2 x = f.pos[s], u = f.vel[s];
    if (!isfinite(f.rho[s]) || !isfinite(f.p[s]) || !isfinite(x.x) || !isfinite(x.y) || !isfinite(u.x) ||
        !isfinite(u.y))
        atomicOr(bad, 1u);
}

// min / max of H over alive particles; h is positive so the float bit
// patterns order like unsigned integers (non-positive or NaN h marks the
// range invalid via the sentinel written by the host).
__global__ __launch_bounds__(kBlock) void k_h_range(Fields f, int32_t n, float *range) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const float h = f.h[s];
    unsigned int *r = reinterpret_cast<unsigned int *>(range);
    if (!(h > 0.0f) || !isfinite(h)) {
        atomicMax(&r[2], 1u);
        return;
    }
    atomicMin(&r[0], __float_as_uint(h));
    atomicMax(&r[1], __float_as_uint(h));
}

inline dim3 grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return dim3((unsigned)(g > 0 ? g : 1));
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, grid_for(n), dim3(kBlock), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_reset_search(int32_t *counts, int32_t n, unsigned long long *pairs_last, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_search, grid_for(n), dim3(kBlock), 0, s, counts, n, pairs_last);
    return hipGetLastError();
}

hipError_t launch_key_count(Fields f, int32_t n, Grid g, int32_t *counts, int32_t *key, int32_t *tmpslot,
                            uint32_t *flags, int32_t max_per_cell, hipStream_t s) {
    hipLaunchKernelGGL(k_key_count, grid_for(n), dim3(kBlock), 0, s, f, n, g, counts, key, tmpslot, flags,
                       max_per_cell);
    return hipGetLastError();
}

hipError_t launch_scan(const int32_t *counts, int32_t n, int32_t *start, int32_t *tile_sums, hipStream_t s) {
    const int32_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kBlock), 0, s, counts, n, tile_sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tile_sums, tiles, start, n);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kBlock), 0, s, counts, n, tile_sums, start);
    return hipGetLastError();
}

hipError_t launch_fill_members(const Fields f, int32_t n, int32_t n_cells, const int32_t *key,
                               const int32_t *tmpslot, const int32_t *start, int32_t *members, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_members, grid_for(n), dim3(kBlock), 0, s, f, n, n_cells, key, tmpslot, start,
                       members);
    return hipGetLastError();
}

hipError_t launch_permute(Fields cur, Fields nxt, float2 *pos_nb, int32_t n, int32_t cap, int32_t n_cells,
                          const int32_t *key, const int32_t *tmpslot, const int32_t *start,
                          const int32_t *members, int32_t *n_alive, uint32_t *flags, hipStream_t s) {
    hipLaunchKernelGGL(k_permute, grid_for(cap), dim3(kBlock), 0, s, cur, nxt, pos_nb, n, cap, n_cells, key,
                       tmpslot, start, members, n_alive, flags);
    return hipGetLastError();
}

hipError_t launch_search(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph, int32_t fast,
                         int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                         unsigned long long *pairs_total, uint32_t *flags, hipStream_t s) {
    const dim3 g = grid_for(n);
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        if (fast)
            hipLaunchKernelGGL((k_search<true, EPPX_SEARCH_KERNEL_CONTINUITY>), g, dim3(kBlock), 0, s, w, f, ph, M,
                               ncount, pairs_last, pairs_total, flags);
        else
            hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_KERNEL_CONTINUITY>), g, dim3(kBlock), 0, s, w, f, ph, M,
                               ncount, pairs_last, pairs_total, flags);
    } else {
        hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_NONE>), g, dim3(kBlock), 0, s, w, f, ph, M, ncount,
                           pairs_last, pairs_total, flags);
    }
    return hipGetLastError();
}

hipError_t launch_self(Fields f, int32_t n, int32_t action, const Physics &ph, const eppx_field_op &op,
                       hipStream_t s) {
    const dim3 g = grid_for(n);
    switch (action) {
        case EPPX_SELF_CONTINUITY_EOS:
            hipLaunchKernelGGL(k_self<EPPX_SELF_CONTINUITY_EOS>, g, dim3(kBlock), 0, s, f, n, ph, op);
            break;
        case EPPX_SELF_GRAVITY_MOVE:
            hipLaunchKernelGGL(k_self<EPPX_SELF_GRAVITY_MOVE>, g, dim3(kBlock), 0, s, f, n, ph, op);
            break;
        case EPPX_SELF_FILTER:
            hipLaunchKernelGGL(k_self<EPPX_SELF_FILTER>, g, dim3(kBlock), 0, s, f, n, ph, op);
            break;
        default:
            hipLaunchKernelGGL(k_self<EPPX_SELF_SCALE_FIELD>, g, dim3(kBlock), 0, s, f, n, ph, op);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_interaction(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph,
                              const eppx_field_op &op, int32_t fast, hipStream_t s) {
    const dim3 g = grid_for(n);
    switch (action) {
        case EPPX_INTER_MOMENTUM:
            if (fast)
                hipLaunchKernelGGL((k_interaction<true, EPPX_INTER_MOMENTUM>), g, dim3(kBlock), 0, s, w, f, ph, op);
            else
                hipLaunchKernelGGL((k_interaction<false, EPPX_INTER_MOMENTUM>), g, dim3(kBlock), 0, s, w, f, ph, op);
            break;
        case EPPX_INTER_FILTER:
            hipLaunchKernelGGL((k_interaction<false, EPPX_INTER_FILTER>), g, dim3(kBlock), 0, s, w, f, ph, op);
            break;
        default:
            hipLaunchKernelGGL((k_interaction<false, EPPX_INTER_ADD_PER_NEIGHBOUR>), g, dim3(kBlock), 0, s, w, f,
                               ph, op);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_export(const Walk &w, Fields f, int32_t n, const int64_t *row_start, int32_t *js,
                         const int32_t *ncount, NeighbourExport ex, hipStream_t s) {
    hipLaunchKernelGGL(k_export, grid_for(n), dim3(kBlock), 0, s, w, f, row_start, js, ncount, ex);
    return hipGetLastError();
}

hipError_t launch_from_ref(Fields f, int32_t cap, Columns c, const int32_t *ip, const float *fp,
                           const int32_t *is_alive, const int32_t *cell_index, hipStream_t s) {
    hipLaunchKernelGGL(k_from_ref, grid_for(cap), dim3(kBlock), 0, s, f, cap, c, ip, fp, is_alive, cell_index);
    return hipGetLastError();
}

hipError_t launch_to_ref(Fields f, int32_t cap, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                         int32_t *cell_index, hipStream_t s) {
    hipLaunchKernelGGL(k_to_ref, grid_for(cap), dim3(kBlock), 0, s, f, cap, c, ip, fp, is_alive, cell_index);
    return hipGetLastError();
}

hipError_t launch_stencil(Grid g, int32_t *count, int32_t *list, hipStream_t s) {
    hipLaunchKernelGGL(k_stencil, grid_for(g.n_cells), dim3(kBlock), 0, s, g, count, list);
    return hipGetLastError();
}

hipError_t launch_finite(Fields f, int32_t n, uint32_t *bad, hipStream_t s) {
    hipLaunchKernelGGL(k_finite, grid_for(n), dim3(kBlock), 0, s, f, n, bad);
    return hipGetLastError();
}

hipError_t launch_h_range(Fields f, int32_t n, float *range, hipStream_t s) {
    hipLaunchKernelGGL(k_h_range, grid_for(n), dim3(kBlock), 0, s, f, n, range);
    return hipGetLastError();
}

}  // namespace eppx

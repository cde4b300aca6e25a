// Device code of libeppx outside the pair kernels: cell-list build,
// interactions over global memory, self kernels and the reference-layout
// conversion kernels, with their launchers and the pair-kernel dispatch.
// gfx950, wave64.  The pair-kernel templates live in pair_device.h and are
// instantiated in search.hip and momentum.hip.
#include "pair_device.h"

namespace eppx {
namespace {

// ------------------------------------------------------------- fills
__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// (a launch has at least kBlock >= kPairSlots threads)
__global__ void k_reset_search(int32_t *counts, int32_t n, unsigned long long *pairs_last) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) counts[i] = 0;
    if (i < kPairSlots) pairs_last[i * kPairStride] = 0ull;
}


__global__ __launch_bounds__(kBlock) void k_key_count(Fields f, int32_t n, CaseDesc one, const CaseDesc *cases,
                                                       int32_t n_cells, int32_t *counts, int32_t *key,
                                                       int32_t *tmpslot, uint32_t *flags, int32_t max_per_cell) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = s < n;
    int32_t k = -1;
    if (valid) {
        k = cell_key(f, s, cases ? cases[f.cs[s]] : one, n_cells);
        key[s] = k;
    }
    const int32_t slot = count_key(valid, k, n_cells, counts, flags, max_per_cell);
    if (valid) tmpslot[s] = slot;
}

// Exclusive scan of cell counts: per-tile sums, then a per-tile rescan whose
// block first adds up the sums of the tiles before it.  No inter-block
// waiting.
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

__global__ __launch_bounds__(kBlock) void k_scan_down(const int32_t *in, int32_t n, const int32_t *tile_sums,
                                                      int32_t *start) {
    __shared__ int lds[kBlock / 64];
    __shared__ int tile_offset;
    // the sum of the earlier tiles (their sums are at most a few thousand values)
    int pre = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kBlock) pre += tile_sums[t];
    pre = block_sum_i32(pre, lds);
    if (threadIdx.x == 0) {
        tile_offset = pre;
        if (blockIdx.x == gridDim.x - 1) start[n] = pre + tile_sums[blockIdx.x];  // the total
    }
    __syncthreads();
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
    int run = tile_offset + inc - sum;
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

// Final position of each of the n slots in use: cell start + number of cell
// members with a smaller particle index (the order a serial insert gives);
// dead, then released particles keep their provisional slot after the alive
// ones.  Moves every field and writes the search snapshot geo = (x, y, h, m).
__global__ __launch_bounds__(kBlock) void k_permute(Fields cur, Fields nxt, float4 *geo, int32_t n, int32_t batch,
                                                     int32_t n_cells, const int32_t *key, const int32_t *tmpslot,
                                                     const int32_t *start, const int32_t *members,
                                                     int32_t *n_alive, uint32_t *flags, int32_t opts,
                                                     const float4 *dyn_src, const float *vol_src, float mass_one,
                                                     float h_one) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s == 0) {
        n_alive[0] = start[n_cells];      // alive slots
        n_alive[1] = start[n_cells + 1];  // alive and dead slots (released ones follow)
    }
    if (s >= n) return;
    // Every load of the slot is issued before the first store (the two
    // buffers could alias as far as the compiler knows, which would otherwise
    // serialise each field's load and store), so the whole record is in
    // flight while the rank is taken.
    const int32_t k = key[s];
    const int32_t me = cur.id[s];
    const int32_t tmp = tmpslot[s];
    const float2 x = cur.pos[s];
    const float4 dyn = (dyn_src ? dyn_src : cur.dyn)[s];
    const bool move_accum = opts & kPermuteSums;
    const bool move_drho = !(opts & kPermuteNoDrho) || k >= n_cells;  // dead slots keep theirs
    const bool move_dvel = !(opts & kPermuteNoDvel) || k >= n_cells;
    const float2 dvel = move_dvel ? cur.dvel[s] : make_float2(0.0f, 0.0f);
    const bool one_mh = opts & kPermuteOneMassH;
    const float mass = one_mh ? mass_one : cur.mass[s], vol = (vol_src ? vol_src : cur.vol)[s];
    const float h = one_mh ? h_one : cur.h[s];
    const float drho = move_drho ? cur.drho[s] : 0.0f;
    // the filter and extrapolation sums, when they may be non-zero
    const float wv = move_accum ? cur.wv[s] : 0.0f, wvrho = move_accum ? cur.wvrho[s] : 0.0f;
    const float wvp = move_accum ? cur.wvp[s] : 0.0f;
    const int32_t tag = cur.tag[s], mov = cur.mov[s];
    // an alive particle of a particle system sits in cell key + 1 (cell_key_of);
    // the cell of a batch slot or a dead one, and the state of a dead one, are read
    const bool plain = !batch && k < n_cells;
    const int32_t cell = plain ? k + 1 : cur.cell[s];
    const int32_t cs = batch ? cur.cs[s] : 0;
    const int32_t b = start[k];
    int32_t dest, alive;
    if (k < n_cells) {
        const int32_t e = start[k + 1];
        if (e - b <= kRankLimit) {
            // four independent member loads in flight per iteration
            int32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, q = b;
            for (; q + 4 <= e; q += 4) {
                const int32_t a0 = members[q], a1 = members[q + 1], a2 = members[q + 2], a3 = members[q + 3];
                r0 += a0 < me;
                r1 += a1 < me;
                r2 += a2 < me;
                r3 += a3 < me;
            }
            for (; q < e; ++q) r0 += members[q] < me;
            dest = b + r0 + r1 + r2 + r3;
        } else {
            dest = b + tmp;
            atomicOr(flags, kFlagRankFallback);
        }
        alive = 1;
    } else {
        dest = b + tmp;
        const int32_t alive_in = cur.alive[s];
        alive = alive_in == 1 ? 0 : alive_in;
    }
    nxt.pos[dest] = x;
    geo[dest] = make_float4(x.x, x.y, h, mass);
    nxt.dyn[dest] = dyn;
    if (move_dvel) nxt.dvel[dest] = dvel;
    if (!one_mh) nxt.mass[dest] = mass;
    nxt.vol[dest] = vol;
    if (move_drho) nxt.drho[dest] = drho;
    if (!one_mh) nxt.h[dest] = h;
    if (move_accum) {
        nxt.wv[dest] = wv;
        nxt.wvrho[dest] = wvrho;
        nxt.wvp[dest] = wvp;
    }
    nxt.tag[dest] = tag;
    nxt.mov[dest] = mov;
    nxt.id[dest] = me;
    nxt.cell[dest] = cell;
    nxt.alive[dest] = alive;
    if (batch) nxt.cs[dest] = cs;  // a particle system's slots all hold case 0
}

// Other interactions over the last search's lists (device_interaction!,
// Action.jl:121-138), walking global memory:
//  EPPX_INTER_FILTER: iFilter main.jl:128-134 -> iKernelFilter! (Filter.jl:10-22).
//  EPPX_INTER_ADD_PER_NEIGHBOUR: field[I] += value per neighbour.
// Component op.component of a named field as a strided float column.
__device__ __forceinline__ float *field_column(const Fields &f, const eppx_field_op &op, int &stride) {
    switch (op.field) {
        case EPPX_FIELD_POSITION: stride = 2; return reinterpret_cast<float *>(f.pos) + op.component;
        case EPPX_FIELD_VELOCITY: stride = 4; return reinterpret_cast<float *>(f.dyn) + op.component;
        case EPPX_FIELD_DVELOCITY: stride = 2; return reinterpret_cast<float *>(f.dvel) + op.component;
        case EPPX_FIELD_MASS: stride = 1; return f.mass;
        case EPPX_FIELD_VOLUME: stride = 1; return f.vol;
        case EPPX_FIELD_DENSITY: stride = 4; return reinterpret_cast<float *>(f.dyn) + 2;
        case EPPX_FIELD_DDENSITY: stride = 1; return f.drho;
        default: stride = 4; return reinterpret_cast<float *>(f.dyn) + 3;  // EPPX_FIELD_PRESSURE
    }
}

// Julia's max(x, y) of two floats: NaN if either is NaN, and +0 above -0.
__device__ __forceinline__ float jl_max(float x, float y) {
    const bool take_y = (y > x) || (signbit(y) < signbit(x));
    return take_y ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

template <int ACTION>
__global__ __launch_bounds__(kBlock) void k_interaction(Walk w, Fields f, Physics ph, eppx_field_op op) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *w.n_alive) return;
    const float4 gi = w.geo[i];
    const int ti = f.tag[i];
    const CaseDesc cd = w.cases ? w.cases[w.cs[i]] : w.one;
    if (ACTION == EPPX_INTER_FILTER) {
        if (ti != ph.fluid_tag) return;
        float wv = f.wv[i], wvrho = f.wvrho[i];
        const uint32_t fluid_bit = tag_bit(ph.fluid_tag);
        walk(w, cd, i, gi, [&](int j, uint32_t tbj, float, float, float dr2) {
            if (tbj != fluid_bit) return;
            const float r = sqrtf(dr2);
            const float wgt = value_exact(ph.kernel, r, invharmonic(gi.z, w.geo[j].z));
            wv += wgt * f.vol[j];
            wvrho += wgt * f.mass[j];
        });
        f.wv[i] = wv;
        f.wvrho[i] = wvrho;
    } else if (ACTION == EPPX_INTER_EXTRAPOLATE_PRESSURE) {
        // iExtrapolatePressure! (Pressure.jl:81-101) for wall i and fluid j, with
        // w the pair's kernel value (the search's nW), p0 and g from ph
        if (ti != ph.wall_tag) return;
        float wv = f.wv[i], wvp = f.wvp[i];
        const uint32_t fluid_bit = tag_bit(ph.fluid_tag);
        walk(w, cd, i, gi, [&](int j, uint32_t tbj, float dx, float dy, float dr2) {
            if (tbj != fluid_bit) return;
            const float r = sqrtf(dr2);
            const float wgt = value_exact(ph.kernel, r, invharmonic(gi.z, w.geo[j].z));
            float g_dot_x = 0.0f;
            g_dot_x += dx * ph.gx;
            g_dot_x += dy * ph.gy;
            const float4 dj = f.dyn[j];
            const float wvj = wgt * f.vol[j];
            wv += wvj;
            wvp += wvj * (jl_max(dj.w, ph.p0) + jl_max(dj.z * g_dot_x, ph.p0));
        });
        f.wv[i] = wv;
        f.wvp[i] = wvp;
    } else {
        int stride;
        float *col = field_column(f, op, stride);
        float v = col[(int64_t)i * stride];
        walk(w, cd, i, gi, [&](int, uint32_t, float, float, float) { v += op.value; });
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
//  EPPX_SELF_EXTRAPOLATE_PRESSURE: sExtrapolatePressure! (Pressure.jl:125-143) on
//    wall particles: p = wv_p / wv and both sums cleared when wv > 0, else p = p0;
//  EPPX_SELF_SCALE_FIELD: field[I] *= value.
// sContinuity! (main.jl:92-105): rho += drho dt, drho = 0, V = m / rho, p = eos(rho).
__device__ __forceinline__ void continuity_eos_one(Fields &f, int s, int tag, const Physics &ph) {
    if (tag != ph.fluid_tag && tag != ph.wall_tag) return;
    float4 d = f.dyn[s];
    const float vol = continuity_eos_regs(d, f.drho[s], f.mass[s], ph);
    f.drho[s] = 0.0f;
    f.vol[s] = vol;
    f.dyn[s] = d;
}

// sMomentum! (main.jl:120-126): du += g, then x += (u + du dt / 2) dt and
// u += du dt per component, du = 0.
__device__ __forceinline__ void gravity_move_one(Fields &f, int s, int tag, const Physics &ph) {
    if (tag != ph.fluid_tag) return;
    float2 x = f.pos[s];
    float4 d = f.dyn[s];
    gravity_move_regs(x, d, f.dvel[s], ph);
    f.pos[s] = x;
    f.dyn[s] = d;
    f.dvel[s] = make_float2(0.0f, 0.0f);
}

template <int ACTION>
__global__ __launch_bounds__(kBlock) void k_self(Fields f, int32_t n, Physics ph, eppx_field_op op,
                                                 const CaseDesc *cases) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    if (cases) {  // a batch: cases that have finished stay as they are; dt is the case's
        const CaseDesc &cd = cases[f.cs[s]];
        if (cd.steps_left <= 0) return;
        ph.dt = cd.dt;
    }
    const int tag = f.tag[s];
    if (ACTION == EPPX_SELF_CONTINUITY_EOS) {
        continuity_eos_one(f, s, tag, ph);
    } else if (ACTION == EPPX_SELF_GRAVITY_MOVE) {
        gravity_move_one(f, s, tag, ph);
    } else if (ACTION == EPPX_SELF_FILTER) {
        if (tag != ph.fluid_tag) return;
        const float hinv = 1.0f / f.h[s];
        const float w0 = (ph.kernel == EPPX_KERNEL_CUBIC ? kSigmaCubic : kSigmaWendland) * (hinv * hinv);
        const float wv = f.wv[s] + w0 * f.vol[s];
        const float wvrho = f.wvrho[s] + w0 * f.mass[s];
        const float rho = wvrho / wv;
        f.wv[s] = 0.0f;
        f.wvrho[s] = 0.0f;
        f.vol[s] = f.mass[s] / rho;
        float4 d = f.dyn[s];
        d.z = rho;
        d.w = eos(ph, rho);
        f.dyn[s] = d;
    } else if (ACTION == EPPX_SELF_EXTRAPOLATE_PRESSURE) {
        if (tag != ph.wall_tag) return;
        const float wv = f.wv[s];
        float4 d = f.dyn[s];
        if (wv > 0.0f) {
            d.w = f.wvp[s] / wv;
            f.wv[s] = 0.0f;
            f.wvp[s] = 0.0f;
        } else {
            d.w = ph.p0;
        }
        f.dyn[s] = d;
    } else {
        int stride;
        float *col = field_column(f, op, stride);
        col[(int64_t)s * stride] *= op.value;
    }
}

// Mass column of the search snapshot after Mass changed (slots in sorted order).
__global__ __launch_bounds__(kBlock) void k_snapshot_mass(float4 *geo, Fields f, int32_t n) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n) geo[s].w = f.mass[s];
}

// The step's first pass: sContinuity! for every slot, and the cell counts
// and last-search pair counter of the coming cell build zeroed (slots
// s < n_counts).  Launched over max(n, n_counts) threads.
__global__ __launch_bounds__(kBlock) void k_eos_reset(Fields f, int32_t n, Physics ph, const CaseDesc *cases,
                                                       int32_t *counts, int32_t n_counts,
                                                       unsigned long long *pairs_last) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n_counts) counts[s] = 0;
    if (s < kPairSlots) pairs_last[s * kPairStride] = 0ull;
    if (s >= n) return;
    // every load of the slot up front, so that they are in flight together
    const int32_t alive = f.alive[s], tag = f.tag[s];
    const int32_t cs = cases ? f.cs[s] : 0;
    float4 d = f.dyn[s];
    const float drho = f.drho[s], mass = f.mass[s];
    if (alive != 1 || (tag != ph.fluid_tag && tag != ph.wall_tag)) return;
    if (cases) {
        const CaseDesc &cd = cases[cs];
        if (cd.steps_left <= 0) return;
        ph.dt = cd.dt;
    }
    const float vol = continuity_eos_regs(d, drho, mass, ph);
    f.drho[s] = 0.0f;
    f.vol[s] = vol;
    f.dyn[s] = d;
}

// sMomentum! fused with the cell key and count of the search that follows:
// the key is taken from the position just written.
__global__ __launch_bounds__(kBlock) void k_move_key(Fields f, int32_t n, Physics ph, CaseDesc one,
                                                      const CaseDesc *cases, int32_t n_cells, int32_t *counts,
                                                      int32_t *key, int32_t *tmpslot, uint32_t *flags,
                                                      int32_t max_per_cell) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = s < n;
    int32_t k = -1;
    if (valid) {
        // every load of the slot up front, so that they are in flight together
        const int32_t alive = f.alive[s], tag = f.tag[s], mov = f.mov[s], cell = f.cell[s];
        const int32_t cs = cases ? f.cs[s] : 0;
        float2 x = f.pos[s];
        float4 d = f.dyn[s];
        const float2 du = f.dvel[s];
        const CaseDesc cd = cases ? cases[cs] : one;
        if (alive == 1 && cd.steps_left > 0 && tag == ph.fluid_tag) {
            Physics p = ph;
            if (cases) p.dt = cd.dt;
            gravity_move_regs(x, d, du, p);
            f.pos[s] = x;
            f.dyn[s] = d;
            f.dvel[s] = make_float2(0.0f, 0.0f);
        }
        k = cell_key_of(f, s, cd, n_cells, alive, mov, cell, x);
        key[s] = k;
    }
    const int32_t slot = count_key(valid, k, n_cells, counts, flags, max_per_cell);
    if (valid) tmpslot[s] = slot;
}

// Writes one case descriptor (a kernel, so that a recorded step may carry it).
__global__ void k_set_case(CaseDesc *d, CaseDesc v) { *d = v; }

// One batch step done: every case with steps left has one fewer.
__global__ __launch_bounds__(kBlock) void k_tick(CaseDesc *cases, int32_t n_cases) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n_cases && cases[c].steps_left > 0) cases[c].steps_left -= 1;
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
    const float4 gi = w.geo[i];
    const int ti = f.tag[i];
    const float hi = gi.z;
    const uint32_t fluid_bit = tag_bit(fluid_tag), wall_bit = tag_bit(wall_tag);
    const size_t cap = (size_t)ex.n_capacity;
    const int64_t o = row_start ? row_start[i] : 0;
    const int limit = js ? ncount[i] : 0;  // the row length the search counted
    int k = 0;
    walk(w, w.cases ? w.cases[w.cs[i]] : w.one, i, gi, [&](int j, uint32_t tbj, float dx, float dy, float dr2) {
        if (js && k < limit) js[o + k] = f.id[j];
        if (k < ex.M) {
            if (ex.nIndex) ex.nIndex[row + cap * k] = f.id[j] + 1;
            if (ex.nRVec) {
                ex.nRVec[row + cap * (2 * k)] = dx;
                ex.nRVec[row + cap * (2 * k + 1)] = dy;
            }
            const float r = sqrtf(dr2);
            if (ex.nR) ex.nR[row + cap * k] = r;
            const bool kernel_pair = (ti == fluid_tag && (tbj == fluid_bit || tbj == wall_bit)) ||
                                     (ti == wall_tag && tbj == fluid_bit);
            if (ex.with_kernel && kernel_pair) {
                const float hinv = invharmonic(hi, w.geo[j].z);
                if (ex.nHInv) ex.nHInv[row + cap * k] = hinv;
                if (ex.nW) ex.nW[row + cap * k] = value_exact(ex.kernel, r, hinv);
                if (ex.nDW) ex.nDW[row + cap * k] = grad_exact_rt(ex.kernel, r, hinv);
            }
        }
        ++k;
    });
    if (ex.nCount) ex.nCount[row] = k;
}

__global__ __launch_bounds__(kBlock) void k_from_ref(Fields f, int32_t offset, Columns c, const int32_t *ip,
                                                      const float *fp, const int32_t *is_alive,
                                                      const int32_t *cell_index, int32_t cs) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= c.n_capacity) return;
    const size_t C = (size_t)c.n_capacity;
    auto fcol = [&](int col) { return col >= 0 ? fp[row + C * col] : 0.0f; };
    auto icol = [&](int col) { return col >= 0 ? ip[row + C * col] : 0; };
    const int s = offset + row;
    f.pos[s] = make_float2(fcol(c.PositionVec), c.PositionVec >= 0 ? fcol(c.PositionVec + 1) : 0.0f);
    f.dyn[s] = make_float4(fcol(c.VelocityVec), c.VelocityVec >= 0 ? fcol(c.VelocityVec + 1) : 0.0f,
                           fcol(c.Density), fcol(c.Pressure));
    f.dvel[s] = make_float2(fcol(c.dVelocityVec), c.dVelocityVec >= 0 ? fcol(c.dVelocityVec + 1) : 0.0f);
    f.mass[s] = fcol(c.Mass);
    f.vol[s] = fcol(c.Volume);
    f.drho[s] = fcol(c.dDensity);
    f.h[s] = fcol(c.H);
    f.wv[s] = fcol(c.SumWeight);
    f.wvrho[s] = fcol(c.SumWeightedDensity);
    f.wvp[s] = fcol(c.SumWeightedPressure);
    f.tag[s] = icol(c.Tag);
    f.mov[s] = icol(c.IsMovable);
    f.id[s] = row;
    f.alive[s] = is_alive[row];
    f.cell[s] = cell_index[row];
    f.cs[s] = cs;
}

__global__ __launch_bounds__(kBlock) void k_to_ref(Fields f, int32_t n, Columns c, int32_t *ip, float *fp,
                                                    int32_t *is_alive, int32_t *cell_index, int32_t cs) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || (cs >= 0 && f.cs[s] != cs)) return;
    const size_t C = (size_t)c.n_capacity;
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
    const float4 d = f.dyn[s];
    put2(c.VelocityVec, make_float2(d.x, d.y));
    put2(c.dVelocityVec, f.dvel[s]);
    put(c.Density, d.z);
    put(c.Pressure, d.w);
    put(c.Mass, f.mass[s]);
    put(c.Volume, f.vol[s]);
    put(c.dDensity, f.drho[s]);
    put(c.H, f.h[s]);
    put(c.SumWeight, f.wv[s]);
    put(c.SumWeightedDensity, f.wvrho[s]);
    put(c.SumWeightedPressure, f.wvp[s]);
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

// Profiler calibration: `iters` x 8 independent v_fma_f32 per lane (eight
// chains, so every wave can issue one VALU instruction per issue slot), the
// known VALU instruction count against which rocprofv3's VALU counters and
// the kernel duration are scaled.
__global__ __launch_bounds__(kBlock) void k_valu_calibration(float *out, int32_t iters) {
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
    const float b = 1.0000001f, c = 1e-7f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], b, c);
    }
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += a[k];
    if (t == 12345.0f) out[blockIdx.x] = t;  // keeps the chains live; never true for these inputs
}

__global__ __launch_bounds__(kBlock) void k_finite(Fields f, int32_t n, uint32_t *bad) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const float2 x = f.pos[s];
    const float4 d = f.dyn[s];
    if (!isfinite(d.z) || !isfinite(d.w) || !isfinite(x.x) || !isfinite(x.y) || !isfinite(d.x) || !isfinite(d.y))
        atomicOr(bad, 1u);
}

// min / max of H and of Mass over alive particles; both are positive, so the
// float bit patterns order like unsigned integers.  range[2] (range[7]) is
// set when some alive h (mass) is not a positive finite number.  range[3]
// collects the tag bits of particles 0..n-1 (alive or not) and range[4] is set
// when one of them has a tag outside 1..kMaxTag; range[5], range[6]: the mass
// min and max.  Grid-stride over the particles, one set of atomics per block.
__global__ __launch_bounds__(kBlock) void k_h_range(Fields f, int32_t n, float *range) {
    __shared__ unsigned int part[8][kBlock / 64];
    unsigned int lo = 0xffffffffu, hi = 0u, bad = 0u, tags = 0u, odd = 0u, mlo = 0xffffffffu, mhi = 0u, mbad = 0u;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
        const uint32_t tb = tag_bit(f.tag[s]);
        tags |= tb;
        odd |= tb == 0u;
        if (f.alive[s] == 1) {
            const float h = f.h[s];
            if (!(h > 0.0f) || !isfinite(h)) {
                bad = 1u;
            } else {
                lo = min(lo, __float_as_uint(h));
                hi = max(hi, __float_as_uint(h));
            }
            const float m = f.mass[s];
            if (!(m > 0.0f) || !isfinite(m)) {
                mbad = 1u;
            } else {
                mlo = min(mlo, __float_as_uint(m));
                mhi = max(mhi, __float_as_uint(m));
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned int)__shfl_down((int)lo, off));
        hi = max(hi, (unsigned int)__shfl_down((int)hi, off));
        bad |= (unsigned int)__shfl_down((int)bad, off);
        tags |= (unsigned int)__shfl_down((int)tags, off);
        odd |= (unsigned int)__shfl_down((int)odd, off);
        mlo = min(mlo, (unsigned int)__shfl_down((int)mlo, off));
        mhi = max(mhi, (unsigned int)__shfl_down((int)mhi, off));
        mbad |= (unsigned int)__shfl_down((int)mbad, off);
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        part[0][wave] = lo;
        part[1][wave] = hi;
        part[2][wave] = bad;
        part[3][wave] = tags;
        part[4][wave] = odd;
        part[5][wave] = mlo;
        part[6][wave] = mhi;
        part[7][wave] = mbad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            lo = min(lo, part[0][k]);
            hi = max(hi, part[1][k]);
            bad |= part[2][k];
            tags |= part[3][k];
            odd |= part[4][k];
            mlo = min(mlo, part[5][k]);
            mhi = max(mhi, part[6][k]);
            mbad |= part[7][k];
        }
        unsigned int *r = reinterpret_cast<unsigned int *>(range);
        if (lo != 0xffffffffu) atomicMin(&r[0], lo);
        if (hi) atomicMax(&r[1], hi);
        if (bad) atomicMax(&r[2], 1u);
        if (tags) atomicOr(&r[3], tags);
        if (odd) atomicMax(&r[4], 1u);
        if (mlo != 0xffffffffu) atomicMin(&r[5], mlo);
        if (mhi) atomicMax(&r[6], mhi);
        if (mbad) atomicMax(&r[7], 1u);
    }
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

hipError_t launch_key_count(Fields f, int32_t n, const CaseDesc &one, const CaseDesc *cases, int32_t n_cells,
                            int32_t *counts, int32_t *key, int32_t *tmpslot, uint32_t *flags, int32_t max_per_cell,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_key_count, grid_for(n), dim3(kBlock), 0, s, f, n, one, cases, n_cells, counts, key,
                       tmpslot, flags, max_per_cell);
    return hipGetLastError();
}

hipError_t launch_scan(const int32_t *counts, int32_t n, int32_t *start, int32_t *tile_sums, hipStream_t s) {
    const int32_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kBlock), 0, s, counts, n, tile_sums);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kBlock), 0, s, counts, n, tile_sums, start);
    return hipGetLastError();
}

hipError_t launch_fill_members(const Fields f, int32_t n, int32_t n_cells, const int32_t *key,
                               const int32_t *tmpslot, const int32_t *start, int32_t *members, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_members, grid_for(n), dim3(kBlock), 0, s, f, n, n_cells, key, tmpslot, start,
                       members);
    return hipGetLastError();
}

hipError_t launch_permute(Fields cur, Fields nxt, float4 *geo, int32_t n, int32_t batch, int32_t n_cells,
                          const int32_t *key, const int32_t *tmpslot, const int32_t *start,
                          const int32_t *members, int32_t *n_alive, uint32_t *flags, int32_t opts,
                          const float4 *dyn_src, const float *vol_src, float mass_one, float h_one,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_permute, grid_for(n), dim3(kBlock), 0, s, cur, nxt, geo, n, batch, n_cells, key, tmpslot,
                       start, members, n_alive, flags, opts, dyn_src, vol_src, mass_one, h_one);
    return hipGetLastError();
}

hipError_t launch_search(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph, int32_t fast,
                         int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                         unsigned long long *pairs_total, uint32_t *flags, int32_t which, int32_t block,
                         int32_t stage, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        const PairLaunch pl{n, M, ncount, pairs_last, pairs_total, flags, which, block, stage, nullptr};
        const bool cubic = ph.kernel == KC;
        if (fast && cubic) return launch_search_pair<true, KC>(w, f, ph, pl, s);
        if (fast) return launch_search_pair<true, KW>(w, f, ph, pl, s);
        if (cubic) return launch_search_pair<false, KC>(w, f, ph, pl, s);
        return launch_search_pair<false, KW>(w, f, ph, pl, s);
    }
    const dim3 g = grid_for(n, 256);
    const EosOut none{};
    if (w.cases)
        hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_NONE, KW, true, 256, 0, false>), g, dim3(256), 0, s, w, f, ph,
                           M, ncount, pairs_last, pairs_total, flags, which, none);
    else
        hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_NONE, KW, false, 256, 0, false>), g, dim3(256), 0, s, w, f,
                           ph, M, ncount, pairs_last, pairs_total, flags, which, none);
    return hipGetLastError();
}

hipError_t launch_search_eos(const Walk &w, Fields f, int32_t n, const Physics &ph, const EosOut &eo, int32_t fast,
                             int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                             unsigned long long *pairs_total, uint32_t *flags, int32_t block, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    const PairLaunch pl{n, M, ncount, pairs_last, pairs_total, flags, 0, block, 0, &eo};
    const bool cubic = ph.kernel == KC;
    if (fast && cubic) return launch_search_pair<true, KC>(w, f, ph, pl, s);
    if (fast) return launch_search_pair<true, KW>(w, f, ph, pl, s);
    if (cubic) return launch_search_pair<false, KC>(w, f, ph, pl, s);
    return launch_search_pair<false, KW>(w, f, ph, pl, s);
}

hipError_t launch_self(Fields f, int32_t n, int32_t action, const Physics &ph, const eppx_field_op &op,
                       const CaseDesc *cases, hipStream_t s) {
    const dim3 g = grid_for(n);
    switch (action) {
        case EPPX_SELF_CONTINUITY_EOS:
            hipLaunchKernelGGL(k_self<EPPX_SELF_CONTINUITY_EOS>, g, dim3(kBlock), 0, s, f, n, ph, op, cases);
            break;
        case EPPX_SELF_GRAVITY_MOVE:
            hipLaunchKernelGGL(k_self<EPPX_SELF_GRAVITY_MOVE>, g, dim3(kBlock), 0, s, f, n, ph, op, cases);
            break;
        case EPPX_SELF_FILTER:
            hipLaunchKernelGGL(k_self<EPPX_SELF_FILTER>, g, dim3(kBlock), 0, s, f, n, ph, op, cases);
            break;
        case EPPX_SELF_EXTRAPOLATE_PRESSURE:
            hipLaunchKernelGGL(k_self<EPPX_SELF_EXTRAPOLATE_PRESSURE>, g, dim3(kBlock), 0, s, f, n, ph, op, cases);
            break;
        default:
            hipLaunchKernelGGL(k_self<EPPX_SELF_SCALE_FIELD>, g, dim3(kBlock), 0, s, f, n, ph, op, cases);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_momentum_move(const Walk &w, Fields f, int32_t n, const Physics &ph, const MoveKey &mk,
                                int32_t fast, int32_t block, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    const bool cubic = ph.kernel == KC;
    if (fast && cubic) return launch_momentum_pair<true, KC>(w, f, ph, n, block, 0, &mk, s);
    if (fast) return launch_momentum_pair<true, KW>(w, f, ph, n, block, 0, &mk, s);
    if (cubic) return launch_momentum_pair<false, KC>(w, f, ph, n, block, 0, &mk, s);
    return launch_momentum_pair<false, KW>(w, f, ph, n, block, 0, &mk, s);
}

hipError_t launch_interaction(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph,
                              const eppx_field_op &op, int32_t fast, int32_t block, int32_t stage, hipStream_t s) {
    switch (action) {
        case EPPX_INTER_MOMENTUM: {
            constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
            const bool cubic = ph.kernel == KC;
            if (fast && cubic) return launch_momentum_pair<true, KC>(w, f, ph, n, block, stage, nullptr, s);
            if (fast) return launch_momentum_pair<true, KW>(w, f, ph, n, block, stage, nullptr, s);
            if (cubic) return launch_momentum_pair<false, KC>(w, f, ph, n, block, stage, nullptr, s);
            return launch_momentum_pair<false, KW>(w, f, ph, n, block, stage, nullptr, s);
        }
        case EPPX_INTER_FILTER:
            hipLaunchKernelGGL(k_interaction<EPPX_INTER_FILTER>, grid_for(n), dim3(kBlock), 0, s, w, f, ph, op);
            break;
        case EPPX_INTER_EXTRAPOLATE_PRESSURE:
            hipLaunchKernelGGL(k_interaction<EPPX_INTER_EXTRAPOLATE_PRESSURE>, grid_for(n), dim3(kBlock), 0, s, w, f,
                               ph, op);
            break;
        default:
            hipLaunchKernelGGL(k_interaction<EPPX_INTER_ADD_PER_NEIGHBOUR>, grid_for(n), dim3(kBlock), 0, s, w, f, ph,
                               op);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_export(const Walk &w, Fields f, int32_t n, const int64_t *row_start, int32_t *js,
                         const int32_t *ncount, NeighbourExport ex, hipStream_t s) {
    hipLaunchKernelGGL(k_export, grid_for(n), dim3(kBlock), 0, s, w, f, row_start, js, ncount, ex);
    return hipGetLastError();
}

hipError_t launch_from_ref(Fields f, int32_t offset, Columns c, const int32_t *ip, const float *fp,
                           const int32_t *is_alive, const int32_t *cell_index, int32_t cs, hipStream_t s) {
    hipLaunchKernelGGL(k_from_ref, grid_for(c.n_capacity), dim3(kBlock), 0, s, f, offset, c, ip, fp, is_alive,
                       cell_index, cs);
    return hipGetLastError();
}

hipError_t launch_to_ref(Fields f, int32_t n, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                         int32_t *cell_index, int32_t cs, hipStream_t s) {
    hipLaunchKernelGGL(k_to_ref, grid_for(n), dim3(kBlock), 0, s, f, n, c, ip, fp, is_alive, cell_index, cs);
    return hipGetLastError();
}

hipError_t launch_eos_reset(Fields f, int32_t n, const Physics &ph, const CaseDesc *cases, int32_t *counts,
                            int32_t n_counts, unsigned long long *pairs_last, hipStream_t s) {
    hipLaunchKernelGGL(k_eos_reset, grid_for(std::max(n, n_counts)), dim3(kBlock), 0, s, f, n, ph, cases, counts,
                       n_counts, pairs_last);
    return hipGetLastError();
}

hipError_t launch_move_key(Fields f, int32_t n, const Physics &ph, const CaseDesc &one, const CaseDesc *cases,
                           int32_t n_cells, int32_t *counts, int32_t *key, int32_t *tmpslot, uint32_t *flags,
                           int32_t max_per_cell, hipStream_t s) {
    hipLaunchKernelGGL(k_move_key, grid_for(n), dim3(kBlock), 0, s, f, n, ph, one, cases, n_cells, counts, key,
                       tmpslot, flags, max_per_cell);
    return hipGetLastError();
}

hipError_t launch_set_case(CaseDesc *d, const CaseDesc &v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_case, dim3(1), dim3(1), 0, s, d, v);
    return hipGetLastError();
}

hipError_t launch_tick(CaseDesc *cases, int32_t n_cases, hipStream_t s) {
    hipLaunchKernelGGL(k_tick, grid_for(n_cases), dim3(kBlock), 0, s, cases, n_cases);
    return hipGetLastError();
}

hipError_t launch_stencil(Grid g, int32_t *count, int32_t *list, hipStream_t s) {
    hipLaunchKernelGGL(k_stencil, grid_for(g.n_cells), dim3(kBlock), 0, s, g, count, list);
    return hipGetLastError();
}

hipError_t launch_valu_calibration(float *out, int32_t blocks, int32_t iters, hipStream_t s) {
    hipLaunchKernelGGL(k_valu_calibration, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, iters);
    return hipGetLastError();
}

hipError_t launch_finite(Fields f, int32_t n, uint32_t *bad, hipStream_t s) {
    hipLaunchKernelGGL(k_finite, grid_for(n), dim3(kBlock), 0, s, f, n, bad);
    return hipGetLastError();
}

hipError_t launch_snapshot_mass(float4 *geo, Fields f, int32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_snapshot_mass, grid_for(n), dim3(kBlock), 0, s, geo, f, n);
    return hipGetLastError();
}

hipError_t launch_h_range(Fields f, int32_t n, float *range, hipStream_t s) {
    const dim3 g((unsigned)std::min<int64_t>(grid_for(n).x, 1024));  // grid-stride: one atomic set per block
    hipLaunchKernelGGL(k_h_range, g, dim3(kBlock), 0, s, f, n, range);
    return hipGetLastError();
}

}  // namespace eppx

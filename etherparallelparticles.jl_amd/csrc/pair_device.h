// Device code shared by the kernel translation units of libeppx: helpers,
// the smoothing kernels, the neighbour walks, the LDS tile stager, the
// cell-key helpers and the pair-kernel templates (k_search, k_momentum).
// gfx950, wave64.
//
// Floating point: the library is compiled with -ffp-contract=off and
// correctly rounded f32 divide / sqrt, so the EXACT variants below evaluate
// each expression in the reference's written order (left-to-right products,
// no FMA).  The FAST variants contract explicitly with fmaf and use the
// hardware reciprocal / reciprocal-sqrt; they never touch the criterion.
//
// Pair kernels (search + continuity, momentum) keep one thread per particle
// and split each neighbour cell into two passes: first every candidate of
// the cell is tested (one 16-byte load of x, y, h, tag each) into a 64-bit
// accept mask, then the pair physics runs over the set bits in ascending
// order.  The physics therefore executes once per accepted pair instead of
// once per candidate under a divergent branch, and the summation order stays
// the reference's list order.
#pragma once

#include <math.h>

#include <algorithm>
#include <type_traits>

#include "eppx_internal.h"

namespace eppx {
namespace {

constexpr int kRankLimit = 1024;  // largest cell ordered by particle index
// Pair-kernel shapes the autotuner picks from: BLOCK threads per block (one
// particle each) and the neighbour slots per row the block stages in LDS.
template <int BLOCK>
struct PairShape {
    static constexpr int tile_row = BLOCK + BLOCK / 2;
};
// The tile arrays pass 1 reads keep their last kTilePad entries free (the
// wave-uniform walk reads up to 63 positions past a lane's row range), so the
// ranges a tile stages hold at most 3 tile_row - kTilePad slots.
constexpr int kTilePad = 64;

// sigma of the 2D kernels: FT(10/7/pi) (CubicSpline.jl:14), FT(7/4/pi) (WendlandC2.jl:13)
constexpr float kSigmaCubic = (float)(10.0 / 7.0 / 3.14159265358979323846);
constexpr float kSigmaWendland = (float)(7.0 / 4.0 / 3.14159265358979323846);

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// The dispatcher hands workgroup b to XCD b % 8 (8 XCDs, each with its own
// L2).  Pair blocks are renumbered so that each XCD works through one
// contiguous run of cell-sorted slots: the neighbour rows a block stages are
// then mostly those its XCD-mates staged, and stay in that XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int g) {
    const int x = b % kXcds, k = b / kXcds, per = g / kXcds, rem = g % kXcds;
    return x * per + min(x, rem) + k;
}
// Record k of a slot array, addressed by a 32-bit byte offset from the array
// base (a system holds fewer than 2^28 slots, so k * 16 fits): the gather
// then takes the scalar-base + vector-offset form, with no 64-bit address
// arithmetic per pair.
template <class T>
__device__ __forceinline__ T ld_slot(const T *base, int k) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint32_t)k * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st_slot(T *base, int k, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)k * (uint32_t)sizeof(T)) = v;
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Row cy = c / n and column cx = c % n of cell c on a grid n cells wide.  For
// 0 <= c < 2^22 the float estimate c * rcp(n) is within one of the quotient
// (rcp within an ulp, the product rounded once), so one remainder correction
// makes it exact; larger cells take the integer division.
__device__ __forceinline__ void cell_xy(int c, int n, int &cx, int &cy) {
    if (c >= 0 && c < (1 << 22)) {
        int q = (int)((float)c * frcp((float)n));
        int r = c - q * n;
        if (r < 0) {
            --q;
            r += n;
        } else if (r >= n) {
            ++q;
            r -= n;
        }
        cy = q;
        cx = r;
    } else {
        cy = c / n;
        cx = c - cy * n;
    }
}

// Mean.harmonic / invharmonic (src/Math/Mean.jl:28-34)
__device__ __forceinline__ float harmonic(float a, float b) { return 2.0f * a * b / (a + b); }
__device__ __forceinline__ float invharmonic(float a, float b) { return (a + b) / (2.0f * a * b); }

// _gradient of CubicSpline.jl:41-55 / WendlandC2.jl:37-47 (2D), written order.
template <int K>
__device__ __forceinline__ float grad_exact(float r, float hinv) {
    const float q = r * hinv;
    const float h3 = hinv * hinv * hinv;
    if (K == EPPX_KERNEL_CUBIC) {
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

// Same function contracted: dw = sigma h^3 poly(q).
template <int K>
__device__ __forceinline__ float grad_fast(float r, float hinv) {
    const float q = r * hinv;
    const float h3 = hinv * hinv * hinv;
    if (K == EPPX_KERNEL_CUBIC) {
        const float t = 2.0f - q;
        const float inner = q * fmaf(3.0f, q, -4.0f);
        const float outer = -(t * t);
        return (kSigmaCubic * 0.75f) * h3 * (q < 1.0f ? inner : (q < 2.0f ? outer : 0.0f));
    }
    const float t = fmaxf(2.0f - q, 0.0f);
    return (-kSigmaWendland * 0.625f) * h3 * q * (t * t * t);
}

// dw / r of the same kernel, for the FAST pair terms (rinv = 1 / r,
// h3 = hinv^3).  Wendland C2's dw carries a factor q = r hinv, so dw / r
// needs no reciprocal: (-sigma 5/8) h3 hinv (2 - q)^3.
template <int K>
__device__ __forceinline__ float grad_over_r_fast(float r, float rinv, float hinv, float h3) {
    const float q = r * hinv;
    if (K == EPPX_KERNEL_CUBIC) return grad_fast<K>(r, hinv) * rinv;
    const float t = fmaxf(2.0f - q, 0.0f);
    return ((-kSigmaWendland * 0.625f) * h3 * hinv) * (t * t * t);
}

// The pair's hinv = (h_i + h_j) / (2 h_i h_j) and hinv^3 for the FAST terms
// (a case with one H takes UniformTerms instead).
struct PairH {
    float hinv, h3;
    __device__ __forceinline__ PairH(float hi, float hj) {
        hinv = (hi + hj) * frcp(2.0f * hi * hj);
        h3 = hinv * hinv * hinv;
    }
};

// One H in the case (FAST): every pair term is the kernel's constant factor
// kfac times the rest, so the pair sums carry the rest and the caller adds
// kfac x sum once.  shape(r, rinv) = dw / r / kfac:
//   cubic spline: kfac = sigma 0.75 hinv^3, shape = q (3q - 4) or -(2 - q)^2, over r;
//   Wendland C2:  kfac = -sigma 0.625 hinv^4, shape = (2 - q)^3.
// The artificial viscosity's hinv (v = min(u.x, 0) hinv) and the 1/2 of the
// mean density fold into its two constants: av_b = 2 beta hinv^2,
// av_a = 2 alpha c hinv.
template <int K>
struct UniformTerms {
    float hinv, hinv2, kfac, av_b, av_a;
    __device__ __forceinline__ UniformTerms(float hinv_u, const Physics &ph) {
        hinv = hinv_u;
        hinv2 = hinv * hinv;
        const float h3 = hinv * hinv * hinv;
        kfac = K == EPPX_KERNEL_CUBIC ? (kSigmaCubic * 0.75f) * h3 : ((-kSigmaWendland * 0.625f) * h3) * hinv;
        av_b = ph.beta * (2.0f * hinv * hinv);
        av_a = (ph.alpha * ph.av_c) * (2.0f * hinv);
    }
    __device__ __forceinline__ float shape(float r, float rinv) const {
        if (K == EPPX_KERNEL_CUBIC) {
            const float q = r * hinv;
            const float t = 2.0f - q;
            return (q < 1.0f ? q * fmaf(3.0f, q, -4.0f) : (q < 2.0f ? -(t * t) : 0.0f)) * rinv;
        }
        const float t = fmaxf(fmaf(r, -hinv, 2.0f), 0.0f);  // 2 - q
        return t * (t * t);
    }
    // shape at dr2: Wendland C2 takes r from one hardware square root (one
    // instruction fewer than r = dr2 rsq(dr2), and finite at dr2 = 0)
    __device__ __forceinline__ float shape2(float dr2) const {
        if (K == EPPX_KERNEL_CUBIC) {
            const float rinv = frsq(dr2);
            return shape(dr2 * rinv, rinv);
        }
        return shape(__builtin_amdgcn_sqrtf(dr2), 0.0f);
    }
};

// p_j / rho_j^2 of the classic pressure term (FAST): one expression wherever
// it is taken (per pair, or once per slot of a staged tile), so the pair sums
// carry the same bits for every pair shape.
__device__ __forceinline__ float pressure_ratio(float p, float rho) { return p * frcp(rho * rho); }

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

__device__ __forceinline__ float grad_exact_rt(int kernel, float r, float hinv) {
    return kernel == EPPX_KERNEL_CUBIC ? grad_exact<EPPX_KERNEL_CUBIC>(r, hinv)
                                       : grad_exact<EPPX_KERNEL_WENDLAND_C2>(r, hinv);
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

// sContinuity! (main.jl:92-105) on values in registers: rho += drho dt,
// p = eos(rho); d = (u, v, rho, p) updated; returns the volume m / rho.
__device__ __forceinline__ float continuity_eos_regs(float4 &d, float drho, float mass, const Physics &ph) {
    const float rho = d.z + drho * ph.dt;
    d.z = rho;
    d.w = eos(ph, rho);
    return mass / rho;
}
// sMomentum! (main.jl:120-126) on values in registers: du += g, then per
// component x += (u + du dt / 2) dt and u += du dt.
__device__ __forceinline__ void gravity_move_regs(float2 &x, float4 &d, float2 du, const Physics &ph) {
    du.x += ph.gx;
    du.y += ph.gy;
    x.x += (d.x + du.x * ph.dt * 0.5f) * ph.dt;
    d.x += du.x * ph.dt;
    x.y += (d.y + du.y * ph.dt * 0.5f) * ph.dt;
    d.y += du.y * ph.dt;
}

// Tags enter the neighbour walk as the bit 1 << (t - 1); a tag outside
// 1..kMaxTag gets 0, which no tag is adjacent to.
__device__ __forceinline__ uint32_t tag_bit(int t) {
    const uint32_t t1 = (uint32_t)t - 1u;
    return t1 < (uint32_t)kMaxTag ? 1u << t1 : 0u;
}
// ActivePair row (adjacency bits) of the tag whose bit is tb.
__device__ __forceinline__ uint32_t adj_of(const Walk &w, uint32_t tb) {
    return tb ? w.adj[__builtin_ctz(tb) + 1] : 0u;
}

// The accept / reject decision of device_findNeighbourParticlesFromCells!
// (NeighbourSearch.jl:199-214) for a candidate at dr2 = dx*dx + dy*dy
// (dx = x_I - x_J; two roundings, no FMA).  Exact in every mode: the host
// sets rmin2 / rmax2 so that the criterion accepts every dr2 < rmin2 and
// rejects every dr2 > rmax2 (rmax2 < rmin2 when that decides all of them);
// only dr2 in the band between them evaluates the criterion itself.
__device__ __forceinline__ bool criterion_exact(const Walk &w, const CaseDesc &cd, float hi, float hj, float dr2) {
    if (w.criterion == EPPX_CRIT_DEFAULT) return dr2 <= cd.g.gap_square;  // defaultSearchCriterion
    const float h2 = harmonic(hi, hj) * 2.0f;                            // symmetrySearchCriterion
    return dr2 <= h2 * h2;
}
__device__ __forceinline__ bool accept_dr2(const Walk &w, const CaseDesc &cd, float hi, float hj, float dr2) {
    if (dr2 < cd.rmin2) return true;
    if (!(dr2 <= cd.rmax2)) return false;
    return criterion_exact(w, cd, hi, hj, dr2);
}

// ------------------------------------------------------------------ walk
// Visits the accepted neighbours of sorted slot i, reading global memory, in
// the reference's list order: neighbour cells in device_initializeNeighbourSystem!
// order (x offset outer, y offset inner; NeighbourSystemBase.jl:37-47), then
// ascending particle index inside a cell (the order of a serial insert).
// visit(j, tag bit of j, dx, dy, dr2).
template <class Visit>
__device__ __forceinline__ void walk(const Walk &w, const CaseDesc &cd, int i, float4 gi, Visit &&visit) {
    const int c = w.cell[i] - 1;
    const int cx = c % cd.g.n_x, cy = c / cd.g.n_x;
    const uint32_t adj = adj_of(w, tag_bit(w.tag[i]));
    const float2 xi = make_float2(gi.x, gi.y);
    for (int di = -1; di <= 1; ++di) {
        const int ii = cx + di;
        if (ii < 0 || ii >= cd.g.n_x) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = cy + dj;
            if (jj < 0 || jj >= cd.g.n_y) continue;
            const int k = cd.cell_base + ii + cd.g.n_x * jj;
            const int b = w.start[k], e = w.start[k + 1];
            for (int j = b; j < e; ++j) {
                if (j == i) continue;
                const uint32_t tbj = tag_bit(w.tag[j]);
                if (!(adj & tbj)) continue;
                const float4 gj = w.geo[j];
                const float dx = xi.x - gj.x, dy = xi.y - gj.y;
                const float dr2 = dx * dx + dy * dy;
                if (accept_dr2(w, cd, gi.z, gj.z, dr2)) visit(j, tbj, dx, dy, dr2);
            }
        }
    }
}

// ------------------------------------------------------------ row walk
// Two passes over the 3x3 cell neighbourhood of slot i.
//  Pass 1: the three cells of each neighbour row are contiguous in the
//    cell-sorted order, so each row is one slot range [b_r, e_r) whose
//    candidates are tested, four at a time, into a 64-bit accept mask.
//  Pass 2: for every accepted neighbour j whose tag is in keep_tags (the
//    tag filter is folded into the pass-1 masks), p = fetch(j, k, tag bit,
//    src) gathers what the pair term needs (slot k of src when src stages
//    it, else global memory) and
//    apply(j, tag bit, h_j, m_j, dx, dy, dr2, p) adds the term.  ORDERED
//    visits them in the reference's list order (x offset outer, y offset
//    inner, ascending slot inside a cell), so sums are bit-identical to the
//    reference; otherwise row by row, in one loop that keeps every lane busy.
// A lane whose row range exceeds 64 slots walks global memory one candidate
// at a time instead.  With COUNT, returns the number of accepted neighbours
// of every tag (nCount); otherwise 0.
__device__ __forceinline__ unsigned long long sel3(int r, unsigned long long a, unsigned long long b,
                                                   unsigned long long c) {
    return r == 0 ? a : (r == 1 ? b : c);
}
__device__ __forceinline__ int sel3(int r, int a, int b, int c) { return r == 0 ? a : (r == 1 ? b : c); }

// Accept bits of N candidates g[0..N) (bit k for candidate k).  Without
// CHECK only dr2 < rmin2 is tested: the caller has established that the
// band is empty and every candidate tag is adjacent.  With CHECK, candidates
// in the band take the exact criterion (one branch per group, rarely taken)
// and non-adjacent tags are cleared.
// g: the candidates' (x, y); their h comes from src (slots k0 + k), read only
// in the band.
template <bool CHECK, int N, class Src>
__device__ __forceinline__ uint32_t accept_bits(const Walk &w, const CaseDesc &cd, float2 xi, float hi, uint32_t adj,
                                                const float2 (&g)[N], const uint32_t (&tb)[N], const Src &src,
                                                int k0) {
    float d[N];
    uint32_t ok = 0u, band = 0u;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float dx = xi.x - g[k].x, dy = xi.y - g[k].y;
        d[k] = dx * dx + dy * dy;
        ok |= (uint32_t)(d[k] < cd.rmin2) << k;
        if (CHECK) band |= (uint32_t)(d[k] >= cd.rmin2 && d[k] <= cd.rmax2) << k;
    }
    if (CHECK) {
        if (band) {
#pragma unroll
            for (int k = 0; k < N; ++k)
                if ((band >> k) & 1u) ok |= (uint32_t)criterion_exact(w, cd, hi, src.h_of(k0 + k), d[k]) << k;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) ok &= ~((uint32_t)((adj & tb[k]) == 0u) << k);
    }
    return ok;
}

// Candidate geometry of the row walk: slot j of neighbour row r, read from
// global memory or from the block's LDS copy of its three row ranges.
// get(k): the snapshot (x, y, h, m) of slot k; xy_of(k), h_of(k): its parts
// (pass 1 reads (x, y) only); tbit(k): its tag bit.
// kStage: what else a tile holds for pass 2 (0 nothing, 1 the (u, v, rho, p)
// record, 2 its (u, v) half, 3 the (u, v) half in place of (h, m), which the
// terms of a case with one H and one Mass do not read).
struct GlobalRows {
    static constexpr int kStage = 0;
    static constexpr bool kTile = false;
    const float4 *geo;
    const int32_t *tag;
    __device__ __forceinline__ int off(int) const { return 0; }
    __device__ __forceinline__ float4 get(int k) const { return geo[k]; }
    __device__ __forceinline__ float2 xy_of(int k) const { return ld_slot(reinterpret_cast<const float2 *>(geo), 2 * k); }
    __device__ __forceinline__ float h_of(int k) const { return geo[k].z; }
    __device__ __forceinline__ uint32_t tbit(int k) const { return tag_bit(tag[k]); }
    __device__ __forceinline__ uint32_t tbit_adj(int k) const { return tag_bit(tag[k]); }
    __device__ __forceinline__ float4 at(int r, int j) const { return get(j + off(r)); }
};
template <int STAGE>
struct TileRows {
    static constexpr int kStage = STAGE;
    static constexpr bool kTile = true;
    const float2 *xy;     // LDS: (x, y) of row r, slot j at xy[j + off(r)]
    const float2 *hm;     // LDS: (h, m) of the same slot
    const uint8_t *tix;   // LDS: tag - 1 of the same slot (63 outside 1..kMaxTag)
    const float4 *dyn;    // LDS (STAGE 1): (u, v, rho, p) of the same slot
    const float2 *uv;     // LDS (STAGE 2): (u, v) of the same slot
    int off0, off1, off2;
    __device__ __forceinline__ int off(int r) const { return sel3(r, off0, off1, off2); }
    __device__ __forceinline__ float4 get(int k) const {
        const float2 a = xy[k], b = hm[k];
        return make_float4(a.x, a.y, b.x, b.y);
    }
    __device__ __forceinline__ float2 xy_of(int k) const { return xy[k]; }
    __device__ __forceinline__ float h_of(int k) const { return hm[k].x; }
    // a 64-bit shift: slot index 63 lands outside the low 32 bits, giving 0
    __device__ __forceinline__ uint32_t tbit(int k) const { return (uint32_t)(1ull << tix[k]); }
    // tag bit of an accepted neighbour: its tag is adjacent, so in 1..kMaxTag
    __device__ __forceinline__ uint32_t tbit_adj(int k) const { return 1u << tix[k]; }
    __device__ __forceinline__ float4 at(int r, int j) const { return get(j + off(r)); }
    // the same tile read with (u, v) in the second array (staged in place of (h, m))
    __device__ __forceinline__ TileRows<3> uv_in_hm() const {
        TileRows<3> t;
        t.xy = xy;
        t.hm = hm;
        t.tix = tix;
        t.dyn = dyn;
        t.uv = hm;
        t.off0 = off0;
        t.off1 = off1;
        t.off2 = off2;
        return t;
    }
};

template <bool CHECK, class Src>
__device__ __forceinline__ unsigned long long accept_row(const Walk &w, const CaseDesc &cd, const Src &src, int r,
                                                         int b, int e, float2 xi, float hi, uint32_t adj) {
    unsigned long long m = 0ull;
    int j = b;
    const int o = src.off(r);
    for (; j + 4 <= e; j += 4) {  // four independent loads in flight per iteration
        const float2 g[4] = {src.xy_of(j + o), src.xy_of(j + 1 + o), src.xy_of(j + 2 + o), src.xy_of(j + 3 + o)};
        uint32_t tb[4] = {0u, 0u, 0u, 0u};
        if (CHECK) {
#pragma unroll
            for (int k = 0; k < 4; ++k) tb[k] = src.tbit(j + k + o);
        }
        m |= (unsigned long long)accept_bits<CHECK, 4>(w, cd, xi, hi, adj, g, tb, src, j + o) << (j - b);
    }
    for (; j < e; ++j) {
        const float2 g[1] = {src.xy_of(j + o)};
        const uint32_t tb[1] = {CHECK ? src.tbit(j + o) : 0u};
        m |= (unsigned long long)accept_bits<CHECK, 1>(w, cd, xi, hi, adj, g, tb, src, j + o) << (j - b);
    }
    return m;
}

// Largest v over the wave's active lanes, for 0 <= v < 128: a binary search
// on ballots, so the result is wave-uniform (scalar registers).
__device__ __forceinline__ int wave_max_u7(int v) {
    int L = 0;
#pragma unroll
    for (int bit = 6; bit >= 0; --bit) {
        const int c = L | (1 << bit);
        if (__ballot(v >= c)) L = c;
    }
    return L;
}

// Pass 1 of one row for a lane whose rows are staged in LDS, when the
// criterion is one threshold (no band) and every tag lies in 1..kMaxTag.
// The wave steps through positions 0..L-1 of its lanes' row ranges together
// (L: the longest range among its active lanes), four candidates a step, the
// loop control in scalar registers.  Per candidate a lane shifts the sign bit
// of dr2 - rmin2 into a 32-bit word (v_alignbit): the difference of two
// finite floats is negative exactly when dr2 < rmin2 (it is zero only when
// they are equal), and dr2 is the reference's two-rounding sum, so the
// decision is the threshold test's bit for bit.  With TAGS the candidate's
// adjacency enters a second word the same way (adj_rev has bit 31 - t set
// when tag t + 1 is adjacent).  The first candidate shifted lies highest, so
// a word is bit-reversed at the end; positions past the lane's n slots are
// cleared.  s: tile index of the lane's first slot of the row; the tile
// keeps kTilePad entries past the staged ranges, so reads up to s + 63 stay
// inside it.
template <bool TAGS, class Src>
__device__ __forceinline__ unsigned long long accept_row_tiled(const Src &src, int s, int n, int L, float2 xi,
                                                               float rmin2, uint32_t adj_rev) {
    uint32_t word[2] = {0u, 0u};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k0 = 32 * h;
        if (L > k0) {  // wave-uniform
            const int cnt = min(L - k0, 32);
            const int sk = s + k0;
            uint32_t acc = 0u, tacc = 0u;
            int k = 0;
            do {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float2 g = src.xy_of(sk + k + q);
                    const float dx = xi.x - g.x, dy = xi.y - g.y;
                    const float d = dx * dx + dy * dy;
                    acc = __builtin_amdgcn_alignbit(acc, __float_as_uint(d - rmin2), 31);
                    if (TAGS) tacc = __builtin_amdgcn_alignbit(tacc, adj_rev << (src.tix[sk + k + q] & 31u), 31);
                }
                k += 4;
            } while (k < cnt);
            // k candidates shifted (a multiple of 4, at most 32): candidate q sits at bit k - 1 - q
            uint32_t wd = __builtin_bitreverse32(acc) >> (32 - k);
            if (TAGS) wd &= __builtin_bitreverse32(tacc) >> (32 - k);
            word[h] = wd;
        }
    }
    const unsigned long long m = ((unsigned long long)word[1] << 32) | word[0];
    return n > 0 ? m & (~0ull >> (64 - n)) : 0ull;
}

// A particle's cell (cx, cy) within its case and, per neighbour row r (cell
// row cy - 1 + r), the slot range [b[r], e[r]) of cells cx - 1 .. cx + 1
// (clamped to the grid; empty outside it); tbit: its tag bit.
struct PairRows {
    int cx, cy;
    int b[3], e[3];
    uint32_t tbit;
};
// The same from the cell table in global memory (c: 0-based cell).
__device__ __forceinline__ PairRows pair_rows_global(const Walk &w, const CaseDesc &cd, int c, int tag) {
    PairRows pr;
    const int nx = cd.g.n_x, base = cd.cell_base;
    cell_xy(c, nx, pr.cx, pr.cy);
    const int x0 = max(pr.cx - 1, 0), x1 = min(pr.cx + 1, nx - 1);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int jj = pr.cy + r - 1;
        pr.b[r] = pr.e[r] = 0;
        if (jj >= 0 && jj < cd.g.n_y) {
            pr.b[r] = w.start[base + x0 + nx * jj];
            pr.e[r] = w.start[base + x1 + 1 + nx * jj];
        }
    }
    pr.tbit = tag_bit(tag);
    return pr;
}

template <bool ORDERED, bool COUNT, class Src, class Fetch, class Apply>
__device__ __forceinline__ int walk_rows(const Walk &w, const CaseDesc &cd, const Src &src, int i, float4 gi,
                                         const PairRows &pr, uint32_t keep_tags, Fetch &&fetch, Apply &&apply) {
    const int nx = cd.g.n_x, base = cd.cell_base;
    const int cx = pr.cx, cy = pr.cy;
    const int b[3] = {pr.b[0], pr.b[1], pr.b[2]}, e[3] = {pr.e[0], pr.e[1], pr.e[2]};
    if (e[0] - b[0] > 64 || e[1] - b[1] > 64 || e[2] - b[2] > 64) {
        // counted once per wave: the lanes that take this path
        const unsigned long long lanes = __ballot(1);
        if (lane_id() == __ffsll((long long)lanes) - 1) atomicAdd(&w.fallback[1], (uint32_t)__popcll(lanes));
        int cnt = 0;
        walk(w, cd, i, gi, [&](int j, uint32_t tbj, float dx, float dy, float dr2) {
            ++cnt;
            if (tbj & keep_tags) {
                const float4 gj = w.geo[j];
                apply(j, tbj, gj.z, gj.w, dx, dy, dr2, fetch(j, j, tbj, GlobalRows{w.geo, w.tag}));
            }
        });
        return COUNT ? cnt : 0;
    }
    const uint32_t adj = adj_of(w, pr.tbit);
    const uint32_t kadj = adj & keep_tags;
    const float2 xi = make_float2(gi.x, gi.y);
    const unsigned long long self = 1ull << (i - b[1]);
    unsigned long long m0 = 0ull, m1 = 0ull, m2 = 0ull;
    const size_t ms = (size_t)w.mask_stride;
    if (!COUNT && w.mask_mode == 2) {  // the last search's masks: the same pair set, the same keep classes
        m0 = w.masks[i];
        m1 = w.masks[i + ms];
        m2 = w.masks[i + 2 * ms];
    } else {
        // The band and the tag test are skipped for the whole wave when no
        // lane needs them: the band is empty and every tag present is in kadj.
        const bool check = __ballot(cd.check_all || (cd.tags_present & ~kadj) != 0u) != 0ull;
        bool done = false;
        if constexpr (Src::kTile) {
            if (__ballot(cd.check_all) == 0ull) {  // one threshold, tags in 1..kMaxTag: the wave-uniform pass 1
                const int n0 = e[0] - b[0], n1 = e[1] - b[1], n2 = e[2] - b[2];
                const int L = wave_max_u7(max(n0, max(n1, n2)));
                const int s0 = n0 > 0 ? b[0] + src.off(0) : 0;
                const int s1 = n1 > 0 ? b[1] + src.off(1) : 0;
                const int s2 = n2 > 0 ? b[2] + src.off(2) : 0;
                if (check) {
                    const uint32_t adj_rev = __builtin_bitreverse32(kadj);
                    m0 = accept_row_tiled<true>(src, s0, n0, L, xi, cd.rmin2, adj_rev);
                    m1 = accept_row_tiled<true>(src, s1, n1, L, xi, cd.rmin2, adj_rev);
                    m2 = accept_row_tiled<true>(src, s2, n2, L, xi, cd.rmin2, adj_rev);
                } else {
                    m0 = accept_row_tiled<false>(src, s0, n0, L, xi, cd.rmin2, 0u);
                    m1 = accept_row_tiled<false>(src, s1, n1, L, xi, cd.rmin2, 0u);
                    m2 = accept_row_tiled<false>(src, s2, n2, L, xi, cd.rmin2, 0u);
                }
                done = true;
            }
        }
        if (done) {
        } else if (check) {
            m0 = accept_row<true>(w, cd, src, 0, b[0], e[0], xi, gi.z, kadj);
            m1 = accept_row<true>(w, cd, src, 1, b[1], e[1], xi, gi.z, kadj);
            m2 = accept_row<true>(w, cd, src, 2, b[2], e[2], xi, gi.z, kadj);
        } else {
            m0 = accept_row<false>(w, cd, src, 0, b[0], e[0], xi, gi.z, kadj);
            m1 = accept_row<false>(w, cd, src, 1, b[1], e[1], xi, gi.z, kadj);
            m2 = accept_row<false>(w, cd, src, 2, b[2], e[2], xi, gi.z, kadj);
        }
        m1 &= ~self;  // I != J
        if (w.mask_mode == 1) {
            w.masks[i] = m0;
            w.masks[i + ms] = m1;
            w.masks[i + 2 * ms] = m2;
        }
    }
    int cnt = 0;
    if (COUNT) {
        if (adj & ~keep_tags) {  // some adjacent tags are counted but not visited
            cnt = __popcll(accept_row<true>(w, cd, src, 0, b[0], e[0], xi, gi.z, adj)) +
                  __popcll(accept_row<true>(w, cd, src, 1, b[1], e[1], xi, gi.z, adj) & ~self) +
                  __popcll(accept_row<true>(w, cd, src, 2, b[2], e[2], xi, gi.z, adj));
        } else {
            cnt = __popcll(m0) + __popcll(m1) + __popcll(m2);
        }
    }
    // dr2 of an accepted pair: the reference's two roundings in the ordered
    // (EXACT) walk; one FMA in the FAST one, where it only feeds the pair terms
    // (the accept decision above stays the reference's)
    auto visit = [&](int j, int k, uint32_t tbj, const auto &p) {  // slot j, tile index k
        const float4 gj = src.get(k);
        const float dx = xi.x - gj.x, dy = xi.y - gj.y;
        apply(j, tbj, gj.z, gj.w, dx, dy, ORDERED ? dx * dx + dy * dy : fmaf(dx, dx, dy * dy), p);
    };
    if (ORDERED) {
        for (int s9 = 0; s9 < 9; ++s9) {
            const int ii = cx + s9 / 3 - 1, r = s9 % 3, jj = cy + r - 1;
            if (ii < 0 || ii >= nx || jj < 0 || jj >= cd.g.n_y) continue;
            const int k = base + ii + nx * jj;
            const int cb = w.start[k], len = w.start[k + 1] - cb;
            if (len == 0) continue;
            const int lo = cb - sel3(r, b[0], b[1], b[2]);  // < 64: the cell lies inside its row range
            unsigned long long m = sel3(r, m0, m1, m2) >> lo;
            if (len < 64) m &= (1ull << len) - 1ull;
            const int roff = src.off(r);
            while (m) {
                const int j = cb + __builtin_ctzll(m);
                m &= m - 1ull;
                const uint32_t tbj = src.tbit_adj(j + roff);
                visit(j, j + roff, tbj, fetch(j, j + roff, tbj, src));
            }
        }
    } else {
        // Rows 0 and 2 (the rows below and above) share one loop: a particle
        // low in its cell has many neighbours below and few above, so their
        // sum varies little across a wave.  The middle row follows.
        unsigned long long na = m0, nb = m2;
        int ba = b[0], bb = b[2];
        int oa = src.off(0), ob = src.off(2);
        if (!na) { na = nb; ba = bb; oa = ob; nb = 0ull; }
        while (na) {
            const int j = ba + __builtin_ctzll(na), jo = oa;
            na &= na - 1ull;
            const uint32_t tbj = src.tbit_adj(j + jo);
            visit(j, j + jo, tbj, fetch(j, j + jo, tbj, src));
            // the switch after the term: nothing of this pair is live across it
            if (!na) { na = nb; ba = bb; oa = ob; nb = 0ull; }
        }
        unsigned long long nm = m1;
        const int bm = b[1], om = src.off(1);
        while (nm) {
            const int j = bm + __builtin_ctzll(nm);
            nm &= nm - 1ull;
            const uint32_t tbj = src.tbit_adj(j + om);
            visit(j, j + om, tbj, fetch(j, j + om, tbj, src));
        }
    }
    return cnt;
}

// Stages the geometry of a pair block's neighbourhood in LDS.  The block's
// sorted slots [i0, i0 + BLOCK) lie in cells c0 <= c <= c1 of one case, in
// one cell row or in two consecutive ones.  The slots of each cell row (a
// group) have their three neighbour rows inside the rows' slot ranges over
// the group's cells widened by one on either side; those ranges (three per
// group, at most six) are copied to the tile one after another, and each
// thread's row offsets point into its own group's ranges.  Returns false,
// with nothing staged, when the block spans two cases or more than two cell
// rows or the ranges do not fit the tile; the block then reads global
// memory.  only_tag: set when every staged slot (the block's own particles
// among them) carries the tag `tag`.  info: kStageInfo ints of LDS.  Every
// thread of the block calls it (two barriers).  stage_pr (momentum, one H and
// one Mass): the second tile array holds (x_j, rho_j) instead of (h_j, m_j),
// with x_j = pressure_ratio(p_j, rho_j) for slots of tag `tag` (fluid) and p_j
// for the others -- the pair terms then read them instead of h and m.
//
// The block waits on three round trips to memory: the caller's loads of the
// alive count, of the cells of slots i0 and iB = max(min(i0 + BLOCK, cap) - 1, i0)
// (cell_a, cell_b, 1-based; the cell of the last alive slot is loaded again
// in the one block that holds it) and of the thread's own records; the range
// bounds; then the tile together with the cell starts of the groups' rows.
// Those starts (tstart, relative to their range's first slot) give each
// thread its three row ranges (pr) without another load; a group wider than
// the table (a sparse block) takes them from global memory instead.
// cell_i, tag_i: the thread's own cell (1-based) and tag (slot min(i, n - 1)).
// hdr_out: this block's header, written (a search); hdr_in: read instead of
// finding the ranges and the start table (a pass on the layout of the search
// that wrote it, with the same block size): one round trip fewer.
// UV and stage_uv (search, one H and one Mass): the second tile array holds
// (u_j, v_j) instead of (h_j, m_j), so pass 2 gathers nothing from global
// memory for fluid-fluid pairs.
template <int BLOCK>
struct StartTable {  // 16-bit cell starts per group row: what the LDS left by the tile holds
    static constexpr int row = BLOCK >= 512 ? kStartRow512 : kStartRow256;
    static constexpr int header_ints = BLOCK >= 512 ? kHeaderInts512 : kHeaderInts256;  // info, then the table
};
template <int BLOCK>
__device__ __forceinline__ int32_t *block_header(const Walk &w, int i0) {
    int32_t *base = BLOCK >= 512 ? w.headers512 : w.headers256;
    return base + (size_t)(i0 / BLOCK) * StartTable<BLOCK>::header_ints;
}
template <int BLOCK, int STAGE, bool UV = false, int TROW = PairShape<BLOCK>::tile_row>
__device__ __forceinline__ bool stage_rows(const Walk &w, const float4 *dyn, int i0, int n, int cell_a, int cell_b,
                                           int cell_i, int tag_i, float2 *txy, float2 *thm, uint8_t *tix,
                                           float4 *tdyn, float2 *tuv, int *info, uint16_t *tstart,
                                           TileRows<STAGE> &rows, PairRows &pr, int tag, bool &only_tag,
                                           bool stage_pr = false, const int32_t *hdr_in = nullptr,
                                           int32_t *hdr_out = nullptr, bool stage_uv = false) {
    constexpr int kCap = 3 * TROW - kTilePad;
    constexpr int kTab = StartTable<BLOCK>::row;
    constexpr int kHdr = StartTable<BLOCK>::header_ints;
    const int t = threadIdx.x;
    if (hdr_in) {  // the block header a search wrote on this layout: info and the start table
        for (int q = t; q < kHdr; q += BLOCK) {
            const int v = hdr_in[q];
            if (q < kStageInfo) info[q] = v;
            else reinterpret_cast<int32_t *>(tstart)[q - kStageInfo] = v;
        }
    } else if (t < 6) {
        const int g = t / 3, r = t % 3;  // group (cell row) and neighbour row
        const int last = min(i0 + BLOCK, n) - 1;
        const CaseDesc cd = w.cases ? w.cases[w.cs[i0]] : w.one;
        const int nx = cd.g.n_x;
        const int c0 = cell_a - 1;
        const int c1 = (last == max(min(i0 + BLOCK, w.cap) - 1, i0) ? cell_b : w.cell[last]) - 1;
        const int cy0 = c0 / nx, cy1 = c1 / nx;
        const int ok = (cy1 == cy0 || cy1 == cy0 + 1) && (!w.cases || w.cs[last] == w.cs[i0]);
        int b = 0, len = 0, xlo = 0, cnt = 0;
        if (ok && (g == 0 || cy1 != cy0)) {
            const int xa = g == 0 ? c0 % nx : 0;                          // the group's first x cell
            const int xb = (g == 1 || cy1 == cy0) ? c1 % nx : nx - 1;     // and its last
            xlo = max(xa - 1, 0);
            cnt = min(xb + 1, nx - 1) + 2 - xlo;  // cell starts xlo .. min(xb + 1, nx - 1) + 1
            const int jj = cy0 + g + r - 1;
            if (jj >= 0 && jj < cd.g.n_y) {
                b = w.start[cd.cell_base + xlo + nx * jj];
                len = w.start[cd.cell_base + xlo + cnt - 1 + nx * jj] - b;
            }
        }
        info[t] = b;
        info[6 + t] = len;
        if (r == 0) {
            info[15 + g] = xlo;
            info[17 + g] = cnt;
        }
        if (t == 0) {
            info[12] = ok;
            info[13] = cy0;
            info[14] = nx;
            info[19] = cd.g.n_y;
            info[20] = cd.cell_base;
            info[kHeaderTable] = 0;
        }
    }
    __syncthreads();
    if (hdr_out && t < kStageInfo && t != kHeaderTable) hdr_out[t] = info[t];
    // block-uniform: range starts and lengths, and where each range goes in the tile
    auto u = [&](int k) { return __builtin_amdgcn_readfirstlane(info[k]); };
    const int d1 = u(6), d2 = d1 + u(7), d3 = d2 + u(8), d4 = d3 + u(9), d5 = d4 + u(10), total = d5 + u(11);
    // a header holds the start table (where one fits) only for a block the search that wrote it
    // staged (its tile may be smaller than this one)
    const bool ok = u(12) && total <= kCap &&
                    (u(6) | u(7) | u(8) | u(9) | u(10) | u(11)) >= 0 &&  // no range runs backwards
                    (!hdr_in || u(kHeaderTable));
    const int nx = u(14), cy0 = u(13), ny = u(19), cbase = u(20);
    const bool table = u(17) <= kTab && u(18) <= kTab;
    int other = 0;  // this thread staged a slot of another tag
    if (ok) {
        // One flat pass over the six ranges laid end to end: each thread loads
        // all of its tile slots (at most kIter) and table entries before it
        // writes any, so the block waits for one round trip to memory rather
        // than one per range.  Tile index q of range k holds slot q + e_k.
        constexpr int kIter = (kCap + BLOCK - 1) / BLOCK;
        constexpr int kTabIter = (6 * kTab + BLOCK - 1) / BLOCK;
        const int e0 = u(0), e1 = u(1) - d1, e2 = u(2) - d2, e3 = u(3) - d3, e4 = u(4) - d4, e5 = u(5) - d5;
        float4 g[kIter], dd[kIter];
        float2 rp[kIter], uv[kIter];
        int tg[kIter], ts[kTabIter];
#pragma unroll
        for (int it = 0; it < kIter; ++it) {
            const int q = min(t + it * BLOCK, total - 1);  // total > 0: the block's own slots are staged
            const int sl = q + (q >= d5 ? e5 : q >= d4 ? e4 : q >= d3 ? e3 : q >= d2 ? e2 : q >= d1 ? e1 : e0);
            g[it] = w.geo[sl];
            tg[it] = w.tag[sl];
            if constexpr (STAGE == 2) {  // the whole record in one load: (u, v), and (rho, p) for stage_pr
                const float4 d = dyn[sl];
                uv[it] = make_float2(d.x, d.y);
                rp[it] = make_float2(d.z, d.w);
            } else if (stage_pr) {
                rp[it] = reinterpret_cast<const float2 *>(dyn)[2 * (size_t)sl + 1];  // (rho, p)
            }
            if constexpr (UV) {  // rp is read under the condition it is written on
                if (stage_uv) rp[it] = reinterpret_cast<const float2 *>(dyn)[2 * (size_t)sl];  // (u, v)
            }
            if (STAGE == 1) dd[it] = dyn[sl];
        }
        // table entry x of group row k = (g, r): the start of cell xlo_g + x of cell row cy0 + g + r - 1
#pragma unroll
        for (int it = 0; it < kTabIter; ++it) {
            const int q = t + it * BLOCK, k = q / kTab, x = q - k * kTab, gk = k >= 3, jj = cy0 + k - 1 - 2 * gk;
            ts[it] = 0;
            if (!hdr_in && table && q < 6 * kTab && x < info[17 + gk] && jj >= 0 && jj < ny)  // (k varies by thread)
                ts[it] = w.start[cbase + info[15 + gk] + x + nx * jj] - info[k];
        }
#pragma unroll
        for (int it = 0; it < kIter; ++it) {
            const int q = t + it * BLOCK;
            if (q < total) {
                txy[q] = make_float2(g[it].x, g[it].y);
                if (UV && stage_uv) thm[q] = rp[it];
                else
                    thm[q] = stage_pr ? make_float2(tg[it] == tag ? pressure_ratio(rp[it].y, rp[it].x) : rp[it].y,
                                                    rp[it].x)
                                      : make_float2(g[it].z, g[it].w);
                other |= tg[it] != tag;
                const uint32_t t1 = (uint32_t)tg[it] - 1u;
                tix[q] = (uint8_t)(t1 < (uint32_t)kMaxTag ? t1 : 63u);
                if (STAGE == 1) tdyn[q] = dd[it];
                if (STAGE == 2) tuv[q] = uv[it];
            }
        }
        if (!hdr_in) {
#pragma unroll
            for (int it = 0; it < kTabIter; ++it) {
                const int q = t + it * BLOCK;
                if (q < 6 * kTab) tstart[q] = (uint16_t)ts[it];
            }
        }
    } else if (t == 0) {
        atomicAdd(w.fallback, 1u);
    }
    only_tag = ok && !__syncthreads_or(other);
    if (hdr_out && ok && table)
        for (int q = t; q < 3 * kTab; q += BLOCK) hdr_out[kStageInfo + q] = reinterpret_cast<const int32_t *>(tstart)[q];
    if (hdr_out && t == 0) hdr_out[kHeaderTable] = ok;  // the table is there when one fits
    if (ok) {
        int cxi, cyi;
        cell_xy(cell_i - 1, nx, cxi, cyi);
        const int gi = cyi != cy0;  // this thread's slot is in the second row
        rows.xy = txy;
        rows.hm = thm;
        rows.tix = tix;
        rows.dyn = tdyn;
        rows.uv = tuv;
        rows.off0 = gi ? d3 - u(3) : -u(0);
        rows.off1 = gi ? d4 - u(4) : d1 - u(1);
        rows.off2 = gi ? d5 - u(5) : d2 - u(2);
        if (table) {
            pr.cx = cxi;
            pr.cy = cyi;
            const int xlo = gi ? u(16) : u(15);
            const int xl = max(cxi - 1, 0) - xlo, xr = min(cxi + 1, nx - 1) + 1 - xlo;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int k = 3 * gi + r, jj = cyi + r - 1;
                const int rb = gi ? u(3 + r) : u(r);  // the range's first slot
                pr.b[r] = pr.e[r] = 0;
                if (jj >= 0 && jj < ny) {
                    pr.b[r] = rb + tstart[k * kTab + xl];
                    pr.e[r] = rb + tstart[k * kTab + xr];
                }
            }
            pr.tbit = tag_bit(tag_i);
        } else {
            const CaseDesc cd = w.cases ? w.cases[w.cs[i0]] : w.one;
            pr = pair_rows_global(w, cd, cell_i - 1, tag_i);
        }
    }
    return ok;
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

// ------------------------------------------------------------- cell list
// device_insertParticlesIntoCells! (NeighbourSearch.jl:12-47): cell of each
// alive particle (cached for immovable ones), kill outside the domain.  The
// key is the 0-based cell, n_cells for dead particles.
// Cell key of slot s (updates its cell, and kills it outside its domain).
// alive, mov, cell and x are the slot's current values, already loaded.
__device__ __forceinline__ int32_t cell_key_of(Fields &f, int32_t s, const CaseDesc &cd, int32_t n_cells,
                                               int32_t alive, int32_t mov, int32_t cell, float2 x) {
    const Grid &g = cd.g;
    int32_t k = cd.removed ? n_cells + 1 : n_cells;
    if (!cd.removed && alive == 1) {
        int32_t c = cell;
        // a cached cell outside this grid (caller data, or another domain) is recomputed
        if (c < 0 || c > g.n_cells) c = -1;
        if (!(mov == 0 && c > 0)) {
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
        if (c > 0) k = cd.cell_base + c - 1;
    }
    return k;
}
__device__ __forceinline__ int32_t cell_key(Fields &f, int32_t s, const CaseDesc &cd, int32_t n_cells) {
    return cell_key_of(f, s, cd, n_cells, f.alive[s], f.mov[s], f.cell[s], f.pos[s]);
}

// Per-key counts: the valid lanes of a wave split into runs of consecutive
// lanes with equal keys (a sorted wave mostly spans a few cells); the first
// lane of each run adds the run's length to its key's count, all runs in one
// atomic instruction, and a lane's provisional slot is its run's base plus
// its place in the run.  A key that appears in two runs of a wave gets two
// bases.  Every lane of the wave calls it.
__device__ __forceinline__ int32_t count_key(bool valid, int32_t k, int32_t n_cells, int32_t *counts,
                                             uint32_t *flags, int32_t max_per_cell) {
    const int lane = lane_id();
    const int32_t prev = __shfl_up(k, 1);
    const unsigned long long vmask = __ballot(valid);
    const bool is_head = valid && (lane == 0 || prev != k || !((vmask >> (lane - 1)) & 1ull));
    const unsigned long long heads = __ballot(is_head);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
    // a valid lane that is no head follows a valid lane of its key, so a head lies at or below it
    const int head = valid ? 63 - __clzll(heads & upto) : lane;
    int32_t base = 0;
    if (is_head) {
        const unsigned long long stops = (heads | ~vmask) & ~upto;  // next head or invalid lane
        const int32_t cnt = (stops ? __ffsll(stops) - 1 : 64) - lane;
        base = atomicAdd(&counts[k], cnt);
        if (k < n_cells && base + cnt > max_per_cell) atomicOr(flags, kFlagCellOverflow);
    }
    base = __shfl(base, head);
    return valid ? base + (lane - head) : 0;
}

// ------------------------------------------------------------ pair physics
// iValueGradient! + iClassicContinuity! / iBalancedContinuity!
// (Library/Kernel.jl:114-128, Continuity.jl:23-47): one term of drho_i;
// mj_or_rhoVj is m_j (fluid-fluid) or rho_i * V_j (fluid-wall, wall-fluid).
template <int K>
__device__ __forceinline__ float continuity_fast(const PairH &ph, float2 ui, float2 uj, float mj_or_rhoVj,
                                                 float dx, float dy, float dr2) {
    const float rinv = frsq(dr2);
    const float r = dr2 * rinv;
    const float dwr = grad_over_r_fast<K>(r, rinv, ph.hinv, ph.h3);
    const float vdotx = fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x));
    return (mj_or_rhoVj * vdotx) * dwr;
}

// The same term over kfac (one H in the case).
template <int K>
__device__ __forceinline__ float continuity_fast_u(const UniformTerms<K> &ut, float2 ui, float2 uj, float mj_or_rhoVj,
                                                   float dx, float dy, float dr2) {
    const float vdotx = fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x));
    return (mj_or_rhoVj * vdotx) * ut.shape2(dr2);
}

// The fluid-fluid term over kfac m_u (one H and one Mass in the case).
template <int K>
__device__ __forceinline__ float vdotx_shape_u(const UniformTerms<K> &ut, float2 ui, float2 uj, float dx, float dy,
                                               float dr2) {
    const float vdotx = fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x));
    return vdotx * ut.shape2(dr2);
}

// The reference's written order, given the pair's hinv = invharmonic(h_i, h_j).
template <int K>
__device__ __forceinline__ float continuity_exact(float hinv, float2 ui, float2 uj, float mj_or_rhoVj, float dx,
                                                  float dy, float dr2) {
    const float r = sqrtf(dr2);
    const float dw = grad_exact<K>(r, hinv);
    float vdotx = 0.0f;
    vdotx += dx * (ui.x - uj.x);
    vdotx += dy * (ui.y - uj.y);
    return mj_or_rhoVj * vdotx * dw / r;
}

template <bool FAST, int K>
__device__ __forceinline__ float continuity_term(float hi, float hj, float2 ui, float2 uj, float mj_or_rhoVj,
                                                 float dx, float dy, float dr2) {
    if (!FAST) return continuity_exact<K>(invharmonic(hi, hj), ui, uj, mj_or_rhoVj, dx, dy, dr2);
    return continuity_fast<K>(PairH(hi, hj), ui, uj, mj_or_rhoVj, dx, dy, dr2);
}

struct ContinuityGather {
    float2 uv;
    float vol;
};

// Search pass: find the neighbours of every alive particle and run the fused
// action (EPPX_SEARCH_KERNEL_CONTINUITY: sComputeKernelAndiContinuity!,
// main.jl:74-90).
// EOS (particle systems, in a time step, with the continuity action): the
// epilogue also applies the next step's sContinuity! (main.jl:92-105) to the
// particle's new dDensity -- rho, p and dDensity in place (the pair terms of
// other threads read only (u, v) of a record), the volume to eo.vol_new
// (they read V_j), for every slot -- and zeroes the cell counts of the next
// step (the scan has consumed them).
template <bool FAST, int ACTION, int K, bool BATCH, int BLOCK, int STAGE, bool EOS>
// Eight waves per SIMD for one-case launches (they spill only on rare paths);
// the batch kernels need more registers and keep the compiler's choice.
__global__ __launch_bounds__(BLOCK, BATCH ? 1 : 8) void k_search(Walk w, Fields f, Physics ph, int32_t M,
                                                                 int32_t *ncount, unsigned long long *pairs_last,
                                                                 unsigned long long *pairs_total, uint32_t *flags,
                                                                 int32_t which, EosOut eo) {
    // 256 threads, gathered shapes: a tile of 3 x 352 slots, four staging loads per thread instead of
    // five (search 136.4 -> 130.5 us on the 2M dam break; 0.01 % more of its blocks read global memory)
    constexpr int kTileRow = (STAGE == 0 && BLOCK == 256 && !BATCH) ? 352 : PairShape<BLOCK>::tile_row;
    __shared__ int lds[BLOCK / 64];
    __shared__ float2 txy[3 * kTileRow];
    __shared__ float2 thm[3 * kTileRow];
    __shared__ uint8_t tix[3 * kTileRow];
    __shared__ float2 tuv[STAGE == 2 ? 3 * kTileRow : 1];
    __shared__ int info[kStageInfo];
    __shared__ __align__(16) uint16_t tstart[6 * StartTable<BLOCK>::row];
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * BLOCK;
    // the cells of the block's first and last slots, loaded with the alive count (i0 < cap: the grid
    // covers the slots)
    const int cell_a = w.cell[i0], cell_b = w.cell[max(min(i0 + BLOCK, w.cap) - 1, i0)];
    const int n = *w.n_alive;
    if constexpr (EOS) {
        for (int c = blockIdx.x * BLOCK + threadIdx.x; c < eo.n_counts; c += gridDim.x * BLOCK) eo.counts[c] = 0;
        if (i0 >= n) {  // slots of dead particles only: their volumes as they are
            const int s = i0 + threadIdx.x;
            if (s < eo.n_slots) eo.vol_new[s] = f.vol[s];
            return;
        }
    }
    if (i0 >= n) return;  // the whole block
    const int i = i0 + threadIdx.x;
    if (EOS && i >= n && i < eo.n_slots) eo.vol_new[i] = f.vol[i];
    // the particle's own records, requested before the tile is staged (their loads overlap its loads)
    const int ic = min(i, n - 1);
    const float4 gi = w.geo[ic];
    const int ti = f.tag[ic];
    const float4 di = ACTION == EPPX_SEARCH_KERNEL_CONTINUITY ? f.dyn[ic] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float drho_i = w.drho_zero ? 0.0f : ld_slot(f.drho, ic);
    const int cell_i = w.cell[ic];
    TileRows<STAGE> rows;
    PairRows pr;
    bool all_fluid = false;  // every particle of the block's neighbourhood is fluid
    // the tile holds (u_j, v_j) in place of (h_j, m_j): FAST, gathered shapes, one H and one Mass
    constexpr bool kUV = FAST && STAGE == 0 && !BATCH && ACTION == EPPX_SEARCH_KERNEL_CONTINUITY;
    const bool stage_uv = kUV && w.one.hinv_u > 0.0f && w.one.m_u > 0.0f;
    const bool tiled = ACTION == EPPX_SEARCH_KERNEL_CONTINUITY &&
                       stage_rows<BLOCK, STAGE, kUV, kTileRow>(w, f.dyn, i0, n, cell_a, cell_b, cell_i, ti, txy, thm, tix,
                                                     nullptr, tuv, info, tstart, rows, pr, ph.fluid_tag, all_fluid,
                                                     false, nullptr,
                                                     !BATCH && w.header_mode == 1 ? block_header<BLOCK>(w, i0)
                                                                                  : nullptr,
                                                     stage_uv);
    int cnt = 0;
    const CaseDesc cd = BATCH ? w.cases[w.cs[ic]] : w.one;
    if (!tiled) pr = pair_rows_global(w, cd, cell_i - 1, ti);
    // a batch runs the action for the cases that advance (which = 0) or wait for a first search (1)
    if (i < n && (which == 0 ? cd.steps_left > 0 : cd.fresh != 0)) {
        const bool i_fluid = ti == ph.fluid_tag, i_wall = ti == ph.wall_tag;
        if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) {
            const float2 ui = make_float2(di.x, di.y);
            const float rho_i = di.z;
            float acc = drho_i;
            // pair classes of the continuity action: fluid-fluid (classic), fluid-wall / wall-fluid (balanced)
            const uint32_t keep = i_fluid ? tag_bit(ph.fluid_tag) | tag_bit(ph.wall_tag)
                                          : (i_wall ? tag_bit(ph.fluid_tag) : 0u);
            const float2 *uv = reinterpret_cast<const float2 *>(f.dyn);  // (u, v) of slot j at uv[2 j]
            const uint32_t fluid_bit = tag_bit(ph.fluid_tag);
            // u_j, and V_j for mixed pairs only (fluid-fluid terms use m_j from the snapshot)
            auto fetch = [&](int j, int k, uint32_t tbj, const auto &src) {
                ContinuityGather p;
                if constexpr (std::decay_t<decltype(src)>::kStage >= 2) p.uv = src.uv[k];
                else p.uv = ld_slot(uv, 2 * j);
                p.vol = (i_fluid && tbj == fluid_bit) ? 0.0f : ld_slot(f.vol, j);
                return p;
            };
            auto term = [&](int, uint32_t tbj, float hj, float mj, float dx, float dy, float dr2,
                            const ContinuityGather &p) {
                const float m = (i_fluid && tbj == fluid_bit) ? mj : rho_i * p.vol;
                acc += continuity_term<FAST, K>(gi.z, hj, ui, p.uv, m, dx, dy, dr2);
            };
            bool walked = false;
            if constexpr (FAST) {
                // One H in the case: the pair's hinv is loop invariant.  Untiled
                // blocks take the same form (through global memory), so every
                // pair shape and a case alone or in a batch give the same bits.
                if (cd.hinv_u > 0.0f && cd.m_u > 0.0f) {
                    // One Mass as well: the fluid-fluid terms (classic continuity) leave out m_j too,
                    // the mixed ones (balanced) carry their rho_i V_j; both sums are taken in every
                    // block shape, so a fluid-fluid pair adds the same bits everywhere.
                    const UniformTerms<K> ut(cd.hinv_u, ph);
                    float s_ff = 0.0f, s_mx = 0.0f;
                    if (all_fluid) {
                        auto fetch_f = [&](int j, int k, uint32_t, const auto &src) {
                            ContinuityGather p;
                            if constexpr (std::decay_t<decltype(src)>::kStage >= 2) p.uv = src.uv[k];
                            else p.uv = ld_slot(uv, 2 * j);
                            p.vol = 0.0f;
                            return p;
                        };
                        auto term_f = [&](int, uint32_t, float, float, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            s_ff += vdotx_shape_u<K>(ut, ui, p.uv, dx, dy, dr2);
                        };
                        if (stage_uv)
                            cnt = walk_rows<false, true>(w, cd, rows.uv_in_hm(), i, gi, pr, keep, fetch_f, term_f);
                        else cnt = walk_rows<false, true>(w, cd, rows, i, gi, pr, keep, fetch_f, term_f);
                    } else {
                        auto term_m = [&](int, uint32_t tbj, float, float, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            const bool ff = i_fluid && tbj == fluid_bit;
                            const float shape = ut.shape2(dr2);
                            const float vdotx = fmaf(dy, ui.y - p.uv.y, dx * (ui.x - p.uv.x));
                            s_ff += ff ? vdotx * shape : 0.0f;
                            s_mx += ff ? 0.0f : ((rho_i * p.vol) * vdotx) * shape;
                        };
                        if (tiled && stage_uv)
                            cnt = walk_rows<false, true>(w, cd, rows.uv_in_hm(), i, gi, pr, keep, fetch, term_m);
                        else if (tiled) cnt = walk_rows<false, true>(w, cd, rows, i, gi, pr, keep, fetch, term_m);
                        else
                            cnt = walk_rows<false, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term_m);
                    }
                    acc = fmaf(ut.kfac, s_mx, fmaf(ut.kfac * cd.m_u, s_ff, acc));
                    walked = true;
                } else if (cd.hinv_u > 0.0f) {
                    const UniformTerms<K> ut(cd.hinv_u, ph);
                    float sum = 0.0f;  // the pair terms over kfac
                    if (all_fluid) {  // fluid-fluid pairs only (classic continuity): no tag per pair (tiled)
                        auto fetch_f = [&](int j, int k, uint32_t, const auto &src) {
                            ContinuityGather p;
                            if constexpr (std::decay_t<decltype(src)>::kStage >= 2) p.uv = src.uv[k];
                            else p.uv = ld_slot(uv, 2 * j);
                            p.vol = 0.0f;
                            return p;
                        };
                        auto term_f = [&](int, uint32_t, float, float mj, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            sum += continuity_fast_u<K>(ut, ui, p.uv, mj, dx, dy, dr2);
                        };
                        cnt = walk_rows<false, true>(w, cd, rows, i, gi, pr, keep, fetch_f, term_f);
                    } else {
                        auto term_u = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            const float m = (i_fluid && tbj == fluid_bit) ? mj : rho_i * p.vol;
                            sum += continuity_fast_u<K>(ut, ui, p.uv, m, dx, dy, dr2);
                        };
                        if (tiled) cnt = walk_rows<false, true>(w, cd, rows, i, gi, pr, keep, fetch, term_u);
                        else cnt = walk_rows<false, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term_u);
                    }
                    acc = fmaf(ut.kfac, sum, acc);
                    walked = true;
                }
            } else {
                if (cd.hinv_u > 0.0f) {
                    // one H in the case: the pair's hinv is the particle's own (the same bits, once)
                    const float hinv_e = invharmonic(gi.z, gi.z);
                    auto term_e = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2,
                                      const ContinuityGather &p) {
                        const float m = (i_fluid && tbj == fluid_bit) ? mj : rho_i * p.vol;
                        acc += continuity_exact<K>(hinv_e, ui, p.uv, m, dx, dy, dr2);
                    };
                    if (tiled) cnt = walk_rows<true, true>(w, cd, rows, i, gi, pr, keep, fetch, term_e);
                    else cnt = walk_rows<true, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term_e);
                    walked = true;
                }
            }
            if (walked) {
            } else if (tiled) {
                cnt = walk_rows<!FAST, true>(w, cd, rows, i, gi, pr, keep, fetch, term);
            } else {
                cnt = walk_rows<!FAST, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term);
            }
            if (EOS && (i_fluid || i_wall)) {
                float4 d = di;
                const float vol = continuity_eos_regs(d, acc, gi.w, ph);  // gi.w: Mass (the snapshot follows it)
                // 32-bit slot offsets from the scalar bases: nothing 64-bit per lane outlives the walk
                if (!w.drho_zero) st_slot(f.drho, i, 0.0f);
                st_slot(reinterpret_cast<float2 *>(f.dyn), 2 * i + 1, make_float2(d.z, d.w));  // (rho, p) half
                st_slot(eo.vol_new, i, vol);
            } else {
                st_slot(f.drho, i, acc);
                if (EOS) st_slot(eo.vol_new, i, ld_slot(f.vol, i));
            }
        } else {
            walk(w, cd, i, gi, [&](int, uint32_t, float, float, float) { ++cnt; });
        }
        ncount[i] = cnt;
        if (cnt > M) atomicOr(flags, kFlagNeighbourOverflow);
    }
    const int total = block_sum_i32(cnt, lds);
    if (threadIdx.x == 0 && total) {
        const int slot = (blockIdx.x % kPairSlots) * kPairStride;
        atomicAdd(pairs_last + slot, (unsigned long long)total);
        atomicAdd(pairs_total + slot, (unsigned long long)total);
    }
}

// iMomentum! (main.jl:107-118): iClassicPressure! (fluid-fluid) or
// iBalancedPressure! (fluid-wall) (Pressure.jl:24-60), then
// iClassicViscosity! or sArtificialViscosity! (Viscosity.jl:10-52), added to
// the acceleration of fluid particle i.
// FAST form of the same terms, written with dwr = dw / r:
//  pressure: prho2 (-m_j dwr) (dx, dy);
//  classic viscosity: 2 mu m_j r dw / (rho_i rho_j (r^2 + 0.01 h^2)) (du)
//    = -2 mu q^2 (-m_j dwr) / (rho_i rho_j qa) (du), with qa = q^2 + 0.01
//    (r^2 + 0.01 h^2 = qa / hinv^2);
//  artificial viscosity: with v = min(vdotx, 0) hinv, phi = v / qa, and
//    (beta phi - alpha c) phi (-m_j dwr) / (rho_bar rho_i rho_j) (dx, dy)
//    = (beta v - alpha c qa) v (-m_j dwr) / (qa^2 rho_bar rho_i rho_j) (dx, dy),
//    one reciprocal; a separating pair (vdotx > 0) gives v = 0.
template <int K, int VISC>
__device__ __forceinline__ void momentum_fast(const Physics &ph, float2 &acc, bool classic, const PairH &h,
                                              float2 ui, float rho_i, float p_i, float pr_i, float2 uj, float rho_j,
                                              float p_j, float m_j, float dx, float dy, float dr2) {
    const float rinv = frsq(dr2);
    const float r = dr2 * rinv;
    const float mdwr = -m_j * grad_over_r_fast<K>(r, rinv, h.hinv, h.h3);
    const float rirj = rho_i * rho_j;
    // classic: p_i / rho_i^2 + p_j / rho_j^2 (pr_i = p_i / rho_i^2); balanced: (p_i + p_j) / (rho_i rho_j)
    float prho2 = classic ? fmaf(p_j, frcp(rho_j * rho_j), pr_i) : (p_i + p_j) * frcp(rirj);
    prho2 = fmaf(fabsf(prho2), ph.coefficient, prho2);
    float radial = prho2 * mdwr;
    const float q = r * h.hinv;
    const float qa = fmaf(q, q, 0.01f);
    if (VISC == EPPX_VISC_CLASSIC) {
        const float force = (-2.0f * ph.mu) * (q * q) * mdwr * frcp(rirj * qa);
        acc.x = fmaf(force, ui.x - uj.x, acc.x);
        acc.y = fmaf(force, ui.y - uj.y, acc.y);
    } else {
        const float v = fminf(fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x)), 0.0f) * h.hinv;
        const float den = (qa * qa) * ((0.5f * (rho_i + rho_j)) * rirj);
        radial = fmaf(fmaf(ph.beta, v, -(ph.alpha * ph.av_c) * qa) * v, mdwr * frcp(den), radial);
    }
    acc.x = fmaf(radial, dx, acc.x);
    acc.y = fmaf(radial, dy, acc.y);
}

// The same terms over kfac (one H in the case), added to sum:
//   artificial viscosity (beta v - alpha c qa) v (-m_j dwr) / (qa^2 rho_bar rho_i rho_j)
//     = (av_b m - av_a qa) m (-m_j dwr) / (qa^2 (rho_i + rho_j) rho_i rho_j), m = min(u.x, 0).
template <int K, int VISC>
__device__ __forceinline__ void momentum_fast_u(const Physics &ph, float2 &sum, bool classic,
                                                const UniformTerms<K> &ut, float2 ui, float rho_i, float p_i,
                                                float pr_i, float2 uj, float rho_j, float p_j, float m_j, float dx,
                                                float dy, float dr2) {
    const float rinv = frsq(dr2);
    const float r = dr2 * rinv;
    const float mdwr = -m_j * ut.shape(r, rinv);
    const float rirj = rho_i * rho_j;
    float prho2 = classic ? fmaf(p_j, frcp(rho_j * rho_j), pr_i) : (p_i + p_j) * frcp(rirj);
    prho2 = fmaf(fabsf(prho2), ph.coefficient, prho2);
    float radial = prho2 * mdwr;
    const float qa = fmaf(dr2, ut.hinv2, 0.01f);  // q^2 + 0.01
    if (VISC == EPPX_VISC_CLASSIC) {
        const float force = (-2.0f * ph.mu) * (dr2 * ut.hinv2) * mdwr * frcp(rirj * qa);
        sum.x = fmaf(force, ui.x - uj.x, sum.x);
        sum.y = fmaf(force, ui.y - uj.y, sum.y);
    } else {
        const float m = fminf(fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x)), 0.0f);
        const float den = (qa * qa) * ((rho_i + rho_j) * rirj);
        radial = fmaf(fmaf(ut.av_b, m, -ut.av_a * qa) * m, mdwr * frcp(den), radial);
    }
    sum.x = fmaf(radial, dx, sum.x);
    sum.y = fmaf(radial, dy, sum.y);
}

// One H and one Mass in the case, and no pressure correction (coefficient 0,
// as in the driver) (FAST): the sums leave out -m_u as well (every term
// carries shape where the terms above carry mdwr = -m_j shape), and carry
// rho_i times each term, so no term forms rho_i rho_j; the caller adds
// -m_u kfac / rho_i x sum once.  x_j: p_j / rho_j^2 (pressure_ratio) for a
// classic (fluid) neighbour, p_j for a balanced (wall) one; pir_i = p_i / rho_i.
//   classic pressure  rho_i (p_i / rho_i^2 + x_j) = pir_i + rho_i x_j
//   balanced pressure rho_i (p_i + x_j) / (rho_i rho_j) = (p_i + x_j) / rho_j
//   artificial viscosity (av_b m - av_a qa) m / (qa^2 (rho_i + rho_j) rho_j)
// and radial = shape (pressure + viscosity): one product by shape.
template <int K, int VISC>
__device__ __forceinline__ void momentum_fast_um(const Physics &ph, float2 &sum, bool classic,
                                                 const UniformTerms<K> &ut, float2 ui, float rho_i, float p_i,
                                                 float pir_i, float2 uj, float rho_j, float x_j, float dx, float dy,
                                                 float dr2) {
    const float shape = ut.shape2(dr2);
    const float prho = classic ? fmaf(rho_i, x_j, pir_i) : (p_i + x_j) * frcp(rho_j);  // (coefficient 0)
    const float qa = fmaf(dr2, ut.hinv2, 0.01f);  // q^2 + 0.01
    float radial;
    if (VISC == EPPX_VISC_CLASSIC) {
        radial = prho * shape;
        const float force = (-2.0f * ph.mu) * (dr2 * ut.hinv2) * shape * frcp(rho_j * qa);
        sum.x = fmaf(force, ui.x - uj.x, sum.x);
        sum.y = fmaf(force, ui.y - uj.y, sum.y);
    } else {
        const float m = fminf(fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x)), 0.0f);
        const float den = (qa * qa) * ((rho_i + rho_j) * rho_j);
        radial = shape * fmaf(fmaf(ut.av_b, m, -ut.av_a * qa) * m, frcp(den), prho);
    }
    sum.x = fmaf(radial, dx, sum.x);
    sum.y = fmaf(radial, dy, sum.y);
}

// The reference's written order (Pressure.jl:24-60, Viscosity.jl:10-52),
// given the pair's hinv = invharmonic(h_i, h_j), the particle's
// pr_i = p_i / (rho_i rho_i) and, when the case has one H, h = 1 / hinv
// (h_or_0 = 0: taken here, for the pairs that need it).
template <int K, int VISC>
__device__ __forceinline__ void momentum_exact(const Physics &ph, float2 &acc, bool classic, float hinv, float h_or_0,
                                               float2 ui, float rho_i, float p_i, float pr_i, float2 uj, float rho_j,
                                               float p_j, float m_j, float dx, float dy, float dr2) {
    const float dux = ui.x - uj.x, duy = ui.y - uj.y;
    const float r = sqrtf(dr2);
    const float dw = grad_exact<K>(r, hinv);
    float prho2 = classic ? pr_i + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
    prho2 += fabsf(prho2) * ph.coefficient;
    prho2 *= -m_j * dw / r;
    acc.x += prho2 * dx;
    acc.y += prho2 * dy;
    if (VISC == EPPX_VISC_CLASSIC) {
        const float h = h_or_0 > 0.0f ? h_or_0 : 1.0f / hinv;
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
        const float h = h_or_0 > 0.0f ? h_or_0 : 1.0f / hinv;
        const float az = r * r + 0.01f * h * h;
        const float phi = vdotx / (hinv * az);
        float force = (-ph.alpha * ph.av_c + ph.beta * phi) * phi / rho;
        force *= -m_j / (rho_i * rho_j * r) * dw;
        acc.x += force * dx;
        acc.y += force * dy;
    }
}

template <bool FAST, int K, int VISC>
__device__ __forceinline__ void momentum_pair(const Physics &ph, float2 &acc, bool classic, float hi, float hj,
                                              float2 ui, float rho_i, float p_i, float pr_i, float2 uj, float rho_j,
                                              float p_j, float m_j, float dx, float dy, float dr2) {
    if (!FAST) {
        momentum_exact<K, VISC>(ph, acc, classic, invharmonic(hi, hj), 0.0f, ui, rho_i, p_i, pr_i, uj, rho_j, p_j, m_j,
                                dx, dy, dr2);
        return;
    }
    momentum_fast<K, VISC>(ph, acc, classic, PairH(hi, hj), ui, rho_i, p_i, pr_i, uj, rho_j, p_j,
                           m_j, dx, dy, dr2);
}

// MOVE (particle systems, in a time step): the epilogue also runs
// sMomentum! (gravity + accelerate-move) on the result and takes the cell key
// and count of every slot for the coming search (MoveKey).
template <bool FAST, int K, int VISC, bool BATCH, int BLOCK, int STAGE, bool MOVE>
// Eight waves per SIMD; the EXACT batch kernels would spill at that bound and keep the compiler's choice.
__global__ __launch_bounds__(BLOCK, (BATCH && !FAST) ? 1 : 8) void k_momentum(Walk w, Fields f, Physics ph,
                                                                               MoveKey mk) {
    // 256 threads with (u, v) staged: a tile of 3 x 336 slots (25 B each) keeps the block at 26 KB of
    // LDS, six blocks per CU instead of five (and four staging loads per thread instead of five); the
    // blocks whose ranges exceed its 944 slots (0.6 % of them on the 2M dam break) read global memory
    constexpr int kTileRow = (STAGE == 2 && BLOCK == 256) ? 336 : PairShape<BLOCK>::tile_row;
    __shared__ float2 txy[3 * kTileRow];
    __shared__ float2 thm[3 * kTileRow];
    __shared__ uint8_t tix[3 * kTileRow];
    __shared__ float4 tdyn[STAGE == 1 ? 3 * kTileRow : 1];
    __shared__ float2 tuv[STAGE == 2 ? 3 * kTileRow : 1];
    __shared__ int info[kStageInfo];
    __shared__ __align__(16) uint16_t tstart[6 * StartTable<BLOCK>::row];
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * BLOCK;
    // the cells of the block's first and last slots, loaded with the alive count (i0 < cap: the grid
    // covers the slots)
    const int cell_a = w.cell[i0], cell_b = w.cell[max(min(i0 + BLOCK, w.cap) - 1, i0)];
    const int n = *w.n_alive;
    if (!MOVE && i0 >= n) return;  // the whole block
    if (MOVE && blockIdx.x == 0 && threadIdx.x < kPairSlots) mk.pairs_last[threadIdx.x * kPairStride] = 0ull;
    const int i = i0 + threadIdx.x;
    // the particle's own records, requested before the tile is staged (their loads overlap its loads)
    const int ic = min(i, max(n - 1, 0));
    const int tag_i = f.tag[ic];
    const float4 gi_own = w.geo[ic], di_own = f.dyn[ic];
    const float2 dvel_own = (MOVE && mk.dvel_zero) ? make_float2(0.0f, 0.0f) : f.dvel[ic];
    const int cell_i = w.cell[ic];
    TileRows<STAGE> rows;
    PairRows pr;
    bool all_fluid = false;  // every particle of the block's neighbourhood is fluid
    bool tiled = false;
    bool stage_pr = false;   // the tile holds (x_j, rho_j) (the one-H, one-Mass FAST terms)
    if (i0 < n) {  // the same for the whole block
        if (FAST && (STAGE == 0 || STAGE == 2)) {
            // a tiled block lies in one case: the case of its first slot
            const CaseDesc c0 = BATCH ? w.cases[w.cs[i0]] : w.one;
            stage_pr = c0.hinv_u > 0.0f && c0.m_u > 0.0f && ph.coefficient == 0.0f;
        }
        tiled = stage_rows<BLOCK, STAGE, false, kTileRow>(w, f.dyn, i0, n, cell_a, cell_b, cell_i, tag_i, txy, thm, tix, tdyn,
                                         tuv, info, tstart, rows, pr, ph.fluid_tag, all_fluid, stage_pr,
                                         !BATCH && w.header_mode == 2 ? block_header<BLOCK>(w, i0) : nullptr);
    }
    if (!MOVE && i >= n) return;
    const CaseDesc cd = BATCH ? w.cases[w.cs[ic]] : w.one;
    // iMomentum! acts on fluid particles only (of cases that advance)
    const bool act = i < n && tag_i == ph.fluid_tag && cd.steps_left > 0;
    if (!MOVE && !act) return;
    float4 di = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 acc = make_float2(0.0f, 0.0f);

    if (act) {
    if (!tiled) pr = pair_rows_global(w, cd, cell_i - 1, tag_i);
    const float4 gi = gi_own;
    di = di_own;
    const float2 ui = make_float2(di.x, di.y);
    acc = dvel_own;
    // p_i / rho_i^2 of the classic pressure, once (EXACT: the reference's expression, the same bits)
    const float pr_i = FAST ? pressure_ratio(di.w, di.z) : di.w / (di.z * di.z);
    const uint32_t fluid_bit = tag_bit(ph.fluid_tag);
    auto fetch = [&](int j, int k, uint32_t, const auto &src) -> float4 {
        if constexpr (std::decay_t<decltype(src)>::kStage == 1) return src.dyn[k];
        else return ld_slot(f.dyn, j);
    };
    auto term = [&](int, uint32_t tbj, float hj, float mj, float dx, float dy, float dr2, const float4 &dj) {
        momentum_pair<FAST, K, VISC>(ph, acc, tbj == fluid_bit, gi.z, hj, ui, di.z, di.w, pr_i,
                                     make_float2(dj.x, dj.y), dj.z, dj.w, mj, dx, dy, dr2);
    };
    const uint32_t keep = tag_bit(ph.fluid_tag) | tag_bit(ph.wall_tag);
    bool walked = false;
    if constexpr (FAST) {
        // One H in the case: the pair's hinv is loop invariant.  Untiled blocks
        // take the same form (through global memory), so every pair shape and
        // a case alone or in a batch give the same bits.
        if (cd.hinv_u > 0.0f && cd.m_u > 0.0f && ph.coefficient == 0.0f) {
            // One Mass as well (and no pressure correction): the sums leave out -m_u, and every
            // neighbour's x_j (p_j / rho_j^2 or p_j) comes from the tile or is taken per pair with the
            // same expression.
            const UniformTerms<K> ut(cd.hinv_u, ph);
            float2 sum = make_float2(0.0f, 0.0f);  // the pair terms over -m_u kfac / rho_i
            const float pir_i = di.w / di.z;       // p_i / rho_i
            auto fetch_um = [&](int j, int k, uint32_t tbj, const auto &src) -> float4 {
                using Src = std::decay_t<decltype(src)>;
                if constexpr (Src::kTile && (Src::kStage == 0 || Src::kStage == 2)) {
                    // (x_j, rho_j) from the tile; (u, v) from the tile (STAGE 2) or gathered
                    float2 uvj;
                    if constexpr (Src::kStage == 2) uvj = src.uv[k];
                    else uvj = ld_slot(reinterpret_cast<const float2 *>(f.dyn), 2 * j);
                    const float2 xr = src.hm[k];
                    return make_float4(uvj.x, uvj.y, xr.y, xr.x);
                } else {
                    float4 d;
                    if constexpr (Src::kStage == 1) d = src.dyn[k];
                    else d = ld_slot(f.dyn, j);
                    d.w = tbj == fluid_bit ? pressure_ratio(d.w, d.z) : d.w;
                    return d;
                }
            };
            if (all_fluid) {  // fluid-fluid pairs only (classic pressure): no tag per pair (tiled blocks)
                auto term_f = [&](int, uint32_t, float, float, float dx, float dy, float dr2, const float4 &dj) {
                    momentum_fast_um<K, VISC>(ph, sum, true, ut, ui, di.z, di.w, pir_i, make_float2(dj.x, dj.y), dj.z,
                                              dj.w, dx, dy, dr2);
                };
                walk_rows<false, false>(w, cd, rows, i, gi, pr, keep, fetch_um, term_f);
            } else {
                auto term_m = [&](int, uint32_t tbj, float, float, float dx, float dy, float dr2, const float4 &dj) {
                    momentum_fast_um<K, VISC>(ph, sum, tbj == fluid_bit, ut, ui, di.z, di.w, pir_i,
                                              make_float2(dj.x, dj.y), dj.z, dj.w, dx, dy, dr2);
                };
                if (tiled) walk_rows<false, false>(w, cd, rows, i, gi, pr, keep, fetch_um, term_m);
                else walk_rows<false, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch_um, term_m);
            }
            const float km = -(cd.m_u * ut.kfac) / di.z;  // the terms carry rho_i
            acc.x = fmaf(km, sum.x, acc.x);
            acc.y = fmaf(km, sum.y, acc.y);
            walked = true;
        } else if (cd.hinv_u > 0.0f) {
            const UniformTerms<K> ut(cd.hinv_u, ph);
            float2 sum = make_float2(0.0f, 0.0f);  // the pair terms over kfac
            if (all_fluid) {  // fluid-fluid pairs only (classic pressure): no tag per pair (tiled blocks)
                auto term_f = [&](int, uint32_t, float, float mj, float dx, float dy, float dr2, const float4 &dj) {
                    momentum_fast_u<K, VISC>(ph, sum, true, ut, ui, di.z, di.w, pr_i, make_float2(dj.x, dj.y), dj.z,
                                             dj.w, mj, dx, dy, dr2);
                };
                walk_rows<false, false>(w, cd, rows, i, gi, pr, keep, fetch, term_f);
            } else {
                auto term_u = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2,
                                  const float4 &dj) {
                    momentum_fast_u<K, VISC>(ph, sum, tbj == fluid_bit, ut, ui, di.z, di.w, pr_i,
                                             make_float2(dj.x, dj.y), dj.z, dj.w, mj, dx, dy, dr2);
                };
                if (tiled) walk_rows<false, false>(w, cd, rows, i, gi, pr, keep, fetch, term_u);
                else walk_rows<false, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term_u);
            }
            acc.x = fmaf(ut.kfac, sum.x, acc.x);
            acc.y = fmaf(ut.kfac, sum.y, acc.y);
            walked = true;
        }
    } else {
        if (cd.hinv_u > 0.0f) {
            // one H in the case: the pair's hinv, and its h, are the particle's own
            // (the same bits as per pair, once)
            const float hinv_e = invharmonic(gi.z, gi.z), h_e = 1.0f / hinv_e;
            auto term_e = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2, const float4 &dj) {
                momentum_exact<K, VISC>(ph, acc, tbj == fluid_bit, hinv_e, h_e, ui, di.z, di.w, pr_i,
                                        make_float2(dj.x, dj.y), dj.z, dj.w, mj, dx, dy, dr2);
            };
            if (tiled) walk_rows<true, false>(w, cd, rows, i, gi, pr, keep, fetch, term_e);
            else walk_rows<true, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term_e);
            walked = true;
        }
    }
    if (walked) {
    } else if (tiled) {
        walk_rows<!FAST, false>(w, cd, rows, i, gi, pr, keep, fetch, term);
    } else {
        walk_rows<!FAST, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, pr, keep, fetch, term);
    }
    if (!MOVE) f.dvel[i] = acc;
    }
    if constexpr (MOVE) {
        // The neighbours of other threads still read this slot's record, so
        // the moved one goes to dyn_new; positions, dvel, cell and alive are
        // not read by the pair terms.
        const bool valid = i < mk.n_slots;
        int32_t k = -1;
        if (valid) {
            const int32_t alive = f.alive[i], mov = f.mov[i], cell = f.cell[i];
            float2 x = f.pos[i];
            float4 d = f.dyn[i];
            if (act) {
                gravity_move_regs(x, d, acc, ph);  // du = the momentum result
                f.pos[i] = x;
                if (!mk.dvel_zero) f.dvel[i] = make_float2(0.0f, 0.0f);
            }
            mk.dyn_new[i] = d;
            k = cell_key_of(f, i, cd, mk.n_cells, alive, mov, cell, x);
            mk.key[i] = k;
        }
        const int32_t slot = count_key(valid, k, mk.n_cells, mk.counts, mk.flags, mk.max_per_cell);
        if (valid) mk.tmpslot[i] = slot;
    }
}


inline dim3 grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return dim3((unsigned)(g > 0 ? g : 1));
}

}  // namespace
}  // namespace eppx

// Device code of libeppx: cell-list build, pair kernels, self kernels and
// the reference-layout conversion kernels.  gfx950, wave64.
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
    float hinv, kfac, av_b, av_a;
    __device__ __forceinline__ UniformTerms(float hinv_u, const Physics &ph) {
        hinv = hinv_u;
        const float h3 = hinv * hinv * hinv;
        kfac = K == EPPX_KERNEL_CUBIC ? (kSigmaCubic * 0.75f) * h3 : ((-kSigmaWendland * 0.625f) * h3) * hinv;
        av_b = ph.beta * (2.0f * hinv * hinv);
        av_a = (ph.alpha * ph.av_c) * (2.0f * hinv);
    }
    __device__ __forceinline__ float shape(float r, float rinv) const {
        const float q = r * hinv;
        if (K == EPPX_KERNEL_CUBIC) {
            const float t = 2.0f - q;
            return (q < 1.0f ? q * fmaf(3.0f, q, -4.0f) : (q < 2.0f ? -(t * t) : 0.0f)) * rinv;
        }
        const float t = fmaxf(2.0f - q, 0.0f);
        return t * (t * t);
    }
};

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
// record, 2 its (u, v) half).
struct GlobalRows {
    static constexpr int kStage = 0;
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

template <bool ORDERED, bool COUNT, class Src, class Fetch, class Apply>
__device__ __forceinline__ int walk_rows(const Walk &w, const CaseDesc &cd, const Src &src, int i, float4 gi,
                                         uint32_t keep_tags, Fetch &&fetch, Apply &&apply) {
    const int c = w.cell[i] - 1;
    const int nx = cd.g.n_x, base = cd.cell_base;
    int cx, cy;
    cell_xy(c, nx, cx, cy);
    const int x0 = max(cx - 1, 0), x1 = min(cx + 1, nx - 1);
    int b[3], e[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int jj = cy + r - 1;
        b[r] = e[r] = 0;
        if (jj >= 0 && jj < cd.g.n_y) {
            b[r] = w.start[base + x0 + nx * jj];
            e[r] = w.start[base + x1 + 1 + nx * jj];
        }
    }
    if (e[0] - b[0] > 64 || e[1] - b[1] > 64 || e[2] - b[2] > 64) {
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
    const uint32_t adj = adj_of(w, tag_bit(w.tag[i]));
    const uint32_t kadj = adj & keep_tags;
    const float2 xi = make_float2(gi.x, gi.y);
    const unsigned long long self = 1ull << (i - b[1]);
    unsigned long long m0, m1, m2;
    const size_t ms = (size_t)w.mask_stride;
    if (!COUNT && w.mask_mode == 2) {  // the last search's masks: the same pair set, the same keep classes
        m0 = w.masks[i];
        m1 = w.masks[i + ms];
        m2 = w.masks[i + 2 * ms];
    } else {
        // The band and the tag test are skipped for the whole wave when no
        // lane needs them: the band is empty and every tag present is in kadj.
        const bool check = __ballot(cd.check_all || (cd.tags_present & ~kadj) != 0u) != 0ull;
        if (check) {
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
    auto visit = [&](int j, int roff, uint32_t tbj, const auto &p) {
        const float4 gj = src.get(j + roff);
        const float dx = xi.x - gj.x, dy = xi.y - gj.y;
        apply(j, tbj, gj.z, gj.w, dx, dy, dx * dx + dy * dy, p);
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
                visit(j, roff, tbj, fetch(j, j + roff, tbj, src));
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
            visit(j, jo, tbj, fetch(j, j + jo, tbj, src));
            // the switch after the term: nothing of this pair is live across it
            if (!na) { na = nb; ba = bb; oa = ob; nb = 0ull; }
        }
        unsigned long long nm = m1;
        const int bm = b[1], om = src.off(1);
        while (nm) {
            const int j = bm + __builtin_ctzll(nm);
            nm &= nm - 1ull;
            const uint32_t tbj = src.tbit_adj(j + om);
            visit(j, om, tbj, fetch(j, j + om, tbj, src));
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
// thread of the block calls it (two barriers).
constexpr int kStageInfo = 16;
template <int BLOCK, int STAGE>
__device__ __forceinline__ bool stage_rows(const Walk &w, const float4 *dyn, int i0, int n, float2 *txy,
                                           float2 *thm, uint8_t *tix, float4 *tdyn, float2 *tuv, int *info,
                                           TileRows<STAGE> &rows, int tag, bool &only_tag) {
    constexpr int kCap = 3 * PairShape<BLOCK>::tile_row;
    const int t = threadIdx.x;
    if (t < 6) {
        const int g = t / 3, r = t % 3;  // group (cell row) and neighbour row
        const int last = min(i0 + BLOCK, n) - 1;
        const CaseDesc cd = w.cases ? w.cases[w.cs[i0]] : w.one;
        const int nx = cd.g.n_x;
        const int c0 = w.cell[i0] - 1, c1 = w.cell[last] - 1;
        const int cy0 = c0 / nx, cy1 = c1 / nx;
        const int ok = (cy1 == cy0 || cy1 == cy0 + 1) && (!w.cases || w.cs[last] == w.cs[i0]);
        int b = 0, len = 0;
        if (ok && (g == 0 || cy1 != cy0)) {
            const int xa = g == 0 ? c0 % nx : 0;                          // the group's first x cell
            const int xb = (g == 1 || cy1 == cy0) ? c1 % nx : nx - 1;     // and its last
            const int jj = cy0 + g + r - 1;
            if (jj >= 0 && jj < cd.g.n_y) {
                b = w.start[cd.cell_base + max(xa - 1, 0) + nx * jj];
                len = w.start[cd.cell_base + min(xb + 1, nx - 1) + 1 + nx * jj] - b;
            }
        }
        info[t] = b;
        info[6 + t] = len;
        if (t == 0) {
            info[12] = ok;
            info[13] = cy0;
            info[14] = nx;
        }
    }
    __syncthreads();
    // block-uniform: range starts and lengths, and where each range goes in the tile
    auto u = [&](int k) { return __builtin_amdgcn_readfirstlane(info[k]); };
    const int d1 = u(6), d2 = d1 + u(7), d3 = d2 + u(8), d4 = d3 + u(9), d5 = d4 + u(10), total = d5 + u(11);
    const bool ok = u(12) && total <= kCap &&
                    (u(6) | u(7) | u(8) | u(9) | u(10) | u(11)) >= 0;  // no range runs backwards
    int other = 0;  // this thread staged a slot of another tag
    if (ok) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int b = u(k), len = u(6 + k);
            const int d = k == 0 ? 0 : k == 1 ? d1 : k == 2 ? d2 : k == 3 ? d3 : k == 4 ? d4 : d5;
            for (int m = t; m < len; m += BLOCK) {
                const float4 g = w.geo[b + m];
                txy[d + m] = make_float2(g.x, g.y);
                thm[d + m] = make_float2(g.z, g.w);
                const int tg = w.tag[b + m];
                other |= tg != tag;
                const uint32_t t1 = (uint32_t)tg - 1u;
                tix[d + m] = (uint8_t)(t1 < (uint32_t)kMaxTag ? t1 : 63u);
                if (STAGE == 1) tdyn[d + m] = dyn[b + m];
                if (STAGE == 2) tuv[d + m] = reinterpret_cast<const float2 *>(dyn)[2 * (size_t)(b + m)];
            }
        }
        const int i = min(i0 + t, n - 1);
        int cxi, cyi;
        cell_xy(w.cell[i] - 1, u(14), cxi, cyi);
        const bool second = cyi != u(13);  // this thread's slot is in the second row
        rows.xy = txy;
        rows.hm = thm;
        rows.tix = tix;
        rows.dyn = tdyn;
        rows.uv = tuv;
        rows.off0 = second ? d3 - u(3) : -u(0);
        rows.off1 = second ? d4 - u(4) : d1 - u(1);
        rows.off2 = second ? d5 - u(5) : d2 - u(2);
    } else if (t == 0) {
        atomicAdd(w.fallback, 1u);
    }
    only_tag = ok && !__syncthreads_or(other);
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
                                                     const float4 *dyn_src, const float *vol_src) {
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
    const float mass = cur.mass[s], vol = (vol_src ? vol_src : cur.vol)[s], h = cur.h[s];
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
    nxt.mass[dest] = mass;
    nxt.vol[dest] = vol;
    if (move_drho) nxt.drho[dest] = drho;
    nxt.h[dest] = h;
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
    const float rinv = frsq(dr2);
    const float vdotx = fmaf(dy, ui.y - uj.y, dx * (ui.x - uj.x));
    return (mj_or_rhoVj * vdotx) * ut.shape(dr2 * rinv, rinv);
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
    constexpr int kTileRow = PairShape<BLOCK>::tile_row;
    __shared__ int lds[BLOCK / 64];
    __shared__ float2 txy[3 * kTileRow];
    __shared__ float2 thm[3 * kTileRow];
    __shared__ uint8_t tix[3 * kTileRow];
    __shared__ float2 tuv[STAGE == 2 ? 3 * kTileRow : 1];
    __shared__ int info[kStageInfo];
    const int n = *w.n_alive;
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * BLOCK;
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
    TileRows<STAGE> rows;
    bool all_fluid = false;  // every particle of the block's neighbourhood is fluid
    const bool tiled = ACTION == EPPX_SEARCH_KERNEL_CONTINUITY &&
                       stage_rows<BLOCK, STAGE>(w, f.dyn, i0, n, txy, thm, tix, nullptr, tuv, info, rows, ph.fluid_tag,
                                                all_fluid);
    int cnt = 0;
    const CaseDesc cd = BATCH ? w.cases[w.cs[min(i, n - 1)]] : w.one;
    // a batch runs the action for the cases that advance (which = 0) or wait for a first search (1)
    if (i < n && (which == 0 ? cd.steps_left > 0 : cd.fresh != 0)) {
        const float4 gi = w.geo[i];
        const int ti = f.tag[i];
        const bool i_fluid = ti == ph.fluid_tag, i_wall = ti == ph.wall_tag;
        if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) {
            const float4 di = f.dyn[i];
            const float2 ui = make_float2(di.x, di.y);
            const float rho_i = di.z;
            float acc = w.drho_zero ? 0.0f : ld_slot(f.drho, i);
            // pair classes of the continuity action: fluid-fluid (classic), fluid-wall / wall-fluid (balanced)
            const uint32_t keep = i_fluid ? tag_bit(ph.fluid_tag) | tag_bit(ph.wall_tag)
                                          : (i_wall ? tag_bit(ph.fluid_tag) : 0u);
            const float2 *uv = reinterpret_cast<const float2 *>(f.dyn);  // (u, v) of slot j at uv[2 j]
            const uint32_t fluid_bit = tag_bit(ph.fluid_tag);
            // u_j, and V_j for mixed pairs only (fluid-fluid terms use m_j from the snapshot)
            auto fetch = [&](int j, int k, uint32_t tbj, const auto &src) {
                ContinuityGather p;
                if constexpr (std::decay_t<decltype(src)>::kStage == 2) p.uv = src.uv[k];
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
                if (cd.hinv_u > 0.0f) {
                    const UniformTerms<K> ut(cd.hinv_u, ph);
                    float sum = 0.0f;  // the pair terms over kfac
                    if (all_fluid) {  // fluid-fluid pairs only (classic continuity): no tag per pair (tiled)
                        auto fetch_f = [&](int j, int k, uint32_t, const auto &src) {
                            ContinuityGather p;
                            if constexpr (std::decay_t<decltype(src)>::kStage == 2) p.uv = src.uv[k];
                            else p.uv = ld_slot(uv, 2 * j);
                            p.vol = 0.0f;
                            return p;
                        };
                        auto term_f = [&](int, uint32_t, float, float mj, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            sum += continuity_fast_u<K>(ut, ui, p.uv, mj, dx, dy, dr2);
                        };
                        cnt = walk_rows<false, true>(w, cd, rows, i, gi, keep, fetch_f, term_f);
                    } else {
                        auto term_u = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2,
                                          const ContinuityGather &p) {
                            const float m = (i_fluid && tbj == fluid_bit) ? mj : rho_i * p.vol;
                            sum += continuity_fast_u<K>(ut, ui, p.uv, m, dx, dy, dr2);
                        };
                        if (tiled) cnt = walk_rows<false, true>(w, cd, rows, i, gi, keep, fetch, term_u);
                        else cnt = walk_rows<false, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term_u);
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
                    if (tiled) cnt = walk_rows<true, true>(w, cd, rows, i, gi, keep, fetch, term_e);
                    else cnt = walk_rows<true, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term_e);
                    walked = true;
                }
            }
            if (walked) {
            } else if (tiled) {
                cnt = walk_rows<!FAST, true>(w, cd, rows, i, gi, keep, fetch, term);
            } else {
                cnt = walk_rows<!FAST, true>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term);
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
        atomicAdd(pairs_last, (unsigned long long)total);
        atomicAdd(pairs_total, (unsigned long long)total);
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
    const float q = r * ut.hinv;
    const float qa = fmaf(q, q, 0.01f);
    if (VISC == EPPX_VISC_CLASSIC) {
        const float force = (-2.0f * ph.mu) * (q * q) * mdwr * frcp(rirj * qa);
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
    constexpr int kTileRow = PairShape<BLOCK>::tile_row;
    __shared__ float2 txy[3 * kTileRow];
    __shared__ float2 thm[3 * kTileRow];
    __shared__ uint8_t tix[3 * kTileRow];
    __shared__ float4 tdyn[STAGE == 1 ? 3 * kTileRow : 1];
    __shared__ int info[kStageInfo];
    const int n = *w.n_alive;
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * BLOCK;
    if (!MOVE && i0 >= n) return;  // the whole block
    if (MOVE && blockIdx.x == 0 && threadIdx.x == 0) *mk.pairs_last = 0ull;
    TileRows<STAGE> rows;
    bool all_fluid = false;  // every particle of the block's neighbourhood is fluid
    bool tiled = false;
    if (i0 < n)  // the same for the whole block
        tiled = stage_rows<BLOCK, STAGE>(w, f.dyn, i0, n, txy, thm, tix, tdyn, nullptr, info, rows, ph.fluid_tag,
                                         all_fluid);
    const int i = i0 + threadIdx.x;
    if (!MOVE && i >= n) return;
    const CaseDesc cd = BATCH ? w.cases[w.cs[min(i, n - 1)]] : w.one;
    // iMomentum! acts on fluid particles only (of cases that advance)
    const bool act = i < n && f.tag[i] == ph.fluid_tag && cd.steps_left > 0;
    if (!MOVE && !act) return;
    float4 di = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 acc = make_float2(0.0f, 0.0f);

    if (act) {
    const float4 gi = w.geo[i];
    di = f.dyn[i];
    const float2 ui = make_float2(di.x, di.y);
    acc = (MOVE && mk.dvel_zero) ? make_float2(0.0f, 0.0f) : f.dvel[i];
    // p_i / rho_i^2 of the classic pressure, once (EXACT: the reference's expression, the same bits)
    const float pr_i = FAST ? di.w * frcp(di.z * di.z) : di.w / (di.z * di.z);
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
        if (cd.hinv_u > 0.0f) {
            const UniformTerms<K> ut(cd.hinv_u, ph);
            float2 sum = make_float2(0.0f, 0.0f);  // the pair terms over kfac
            if (all_fluid) {  // fluid-fluid pairs only (classic pressure): no tag per pair (tiled blocks)
                auto term_f = [&](int, uint32_t, float, float mj, float dx, float dy, float dr2, const float4 &dj) {
                    momentum_fast_u<K, VISC>(ph, sum, true, ut, ui, di.z, di.w, pr_i, make_float2(dj.x, dj.y), dj.z,
                                             dj.w, mj, dx, dy, dr2);
                };
                walk_rows<false, false>(w, cd, rows, i, gi, keep, fetch, term_f);
            } else {
                auto term_u = [&](int, uint32_t tbj, float, float mj, float dx, float dy, float dr2,
                                  const float4 &dj) {
                    momentum_fast_u<K, VISC>(ph, sum, tbj == fluid_bit, ut, ui, di.z, di.w, pr_i,
                                             make_float2(dj.x, dj.y), dj.z, dj.w, mj, dx, dy, dr2);
                };
                if (tiled) walk_rows<false, false>(w, cd, rows, i, gi, keep, fetch, term_u);
                else walk_rows<false, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term_u);
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
            if (tiled) walk_rows<true, false>(w, cd, rows, i, gi, keep, fetch, term_e);
            else walk_rows<true, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term_e);
            walked = true;
        }
    }
    if (walked) {
    } else if (tiled) {
        walk_rows<!FAST, false>(w, cd, rows, i, gi, keep, fetch, term);
    } else {
        walk_rows<!FAST, false>(w, cd, GlobalRows{w.geo, w.tag}, i, gi, keep, fetch, term);
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
    if (s == 0) *pairs_last = 0ull;
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

__global__ __launch_bounds__(kBlock) void k_finite(Fields f, int32_t n, uint32_t *bad) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const float2 x = f.pos[s];
    const float4 d = f.dyn[s];
    if (!isfinite(d.z) || !isfinite(d.w) || !isfinite(x.x) || !isfinite(x.y) || !isfinite(d.x) || !isfinite(d.y))
        atomicOr(bad, 1u);
}

// min / max of H over alive particles; h is positive so the float bit
// patterns order like unsigned integers.  range[2] is set when some alive
// h is not a positive finite number.  Grid-stride over the particles, one
// set of atomics per block.  range[3] collects the tag bits of particles 0..n-1 (alive or not) and
// range[4] is set when one of them has a tag outside 1..kMaxTag.
__global__ __launch_bounds__(kBlock) void k_h_range(Fields f, int32_t n, float *range) {
    __shared__ unsigned int part[5][kBlock / 64];
    unsigned int lo = 0xffffffffu, hi = 0u, bad = 0u, tags = 0u, odd = 0u;
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
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned int)__shfl_down((int)lo, off));
        hi = max(hi, (unsigned int)__shfl_down((int)hi, off));
        bad |= (unsigned int)__shfl_down((int)bad, off);
        tags |= (unsigned int)__shfl_down((int)tags, off);
        odd |= (unsigned int)__shfl_down((int)odd, off);
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        part[0][wave] = lo;
        part[1][wave] = hi;
        part[2][wave] = bad;
        part[3][wave] = tags;
        part[4][wave] = odd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            lo = min(lo, part[0][k]);
            hi = max(hi, part[1][k]);
            bad |= part[2][k];
            tags |= part[3][k];
            odd |= part[4][k];
        }
        unsigned int *r = reinterpret_cast<unsigned int *>(range);
        if (lo != 0xffffffffu) atomicMin(&r[0], lo);
        if (hi) atomicMax(&r[1], hi);
        if (bad) atomicMax(&r[2], 1u);
        if (tags) atomicOr(&r[3], tags);
        if (odd) atomicMax(&r[4], 1u);
    }
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
                          const float4 *dyn_src, const float *vol_src, hipStream_t s) {
    hipLaunchKernelGGL(k_permute, grid_for(n), dim3(kBlock), 0, s, cur, nxt, geo, n, batch, n_cells, key, tmpslot,
                       start, members, n_alive, flags, opts, dyn_src, vol_src);
    return hipGetLastError();
}

template <bool FAST, int K, int BLOCK, int STAGE>
static void launch_search_kb(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t M, int32_t *ncount,
                             unsigned long long *pairs_last, unsigned long long *pairs_total, uint32_t *flags,
                             int32_t which, hipStream_t s) {
    constexpr int KCONT = EPPX_SEARCH_KERNEL_CONTINUITY;
    const dim3 g = grid_for(n, BLOCK);
    const EosOut none{};
    if (w.cases)
        hipLaunchKernelGGL((k_search<FAST, KCONT, K, true, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s, w, f, ph, M,
                           ncount, pairs_last, pairs_total, flags, which, none);
    else
        hipLaunchKernelGGL((k_search<FAST, KCONT, K, false, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s, w, f, ph, M,
                           ncount, pairs_last, pairs_total, flags, which, none);
}

// FAST kernels come in every tuned shape (256 or 512 threads; (u, v) staged
// in LDS or gathered); EXACT ones (the parity mode) in 256 threads, gathered.
template <bool FAST, int K>
static void launch_search_k(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t M, int32_t *ncount,
                            unsigned long long *pairs_last, unsigned long long *pairs_total, uint32_t *flags,
                            int32_t which, int32_t block, int32_t stage, hipStream_t s) {
    if (FAST && block == 512 && stage)
        launch_search_kb<FAST, K, 512, 2>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, s);
    else if (FAST && block == 512)
        launch_search_kb<FAST, K, 512, 0>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, s);
    else if (FAST && stage)
        launch_search_kb<FAST, K, 256, 2>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, s);
    else
        launch_search_kb<FAST, K, 256, 0>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, s);
}

hipError_t launch_search(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph, int32_t fast,
                         int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                         unsigned long long *pairs_total, uint32_t *flags, int32_t which, int32_t block,
                         int32_t stage, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        const bool cubic = ph.kernel == KC;
        if (fast && cubic)
            launch_search_k<true, KC>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, block, stage, s);
        else if (fast)
            launch_search_k<true, KW>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, block, stage, s);
        else if (cubic)
            launch_search_k<false, KC>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, block, stage, s);
        else
            launch_search_k<false, KW>(w, f, ph, n, M, ncount, pairs_last, pairs_total, flags, which, block, stage, s);
    } else {
        const dim3 g = grid_for(n, 256);
        const EosOut none{};
        if (w.cases)
            hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_NONE, KW, true, 256, 0, false>), g, dim3(256), 0, s, w, f,
                               ph, M, ncount, pairs_last, pairs_total, flags, which, none);
        else
            hipLaunchKernelGGL((k_search<false, EPPX_SEARCH_NONE, KW, false, 256, 0, false>), g, dim3(256), 0, s, w, f,
                               ph, M, ncount, pairs_last, pairs_total, flags, which, none);
    }
    return hipGetLastError();
}

template <bool FAST, int K, int BLOCK>
static void launch_search_eos_kb(const Walk &w, Fields f, const Physics &ph, const EosOut &eo, int32_t n, int32_t M,
                                 int32_t *ncount, unsigned long long *pairs_last, unsigned long long *pairs_total,
                                 uint32_t *flags, hipStream_t s) {
    hipLaunchKernelGGL((k_search<FAST, EPPX_SEARCH_KERNEL_CONTINUITY, K, false, BLOCK, 0, true>), grid_for(n, BLOCK),
                       dim3(BLOCK), 0, s, w, f, ph, M, ncount, pairs_last, pairs_total, flags, 0, eo);
}

hipError_t launch_search_eos(const Walk &w, Fields f, int32_t n, const Physics &ph, const EosOut &eo, int32_t fast,
                             int32_t M, int32_t *ncount, unsigned long long *pairs_last,
                             unsigned long long *pairs_total, uint32_t *flags, int32_t block, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    const bool cubic = ph.kernel == KC;
    if (fast && block == 512 && cubic)
        launch_search_eos_kb<true, KC, 512>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    else if (fast && block == 512)
        launch_search_eos_kb<true, KW, 512>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    else if (fast && cubic)
        launch_search_eos_kb<true, KC, 256>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    else if (fast)
        launch_search_eos_kb<true, KW, 256>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    else if (cubic)
        launch_search_eos_kb<false, KC, 256>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    else
        launch_search_eos_kb<false, KW, 256>(w, f, ph, eo, n, M, ncount, pairs_last, pairs_total, flags, s);
    return hipGetLastError();
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

template <bool FAST, int K, bool BATCH, int BLOCK, int STAGE>
static void launch_momentum_b(const Walk &w, Fields f, const Physics &ph, int32_t n, hipStream_t s) {
    const dim3 g = grid_for(n, BLOCK);
    const MoveKey none{};
    if (ph.visc == EPPX_VISC_CLASSIC)
        hipLaunchKernelGGL((k_momentum<FAST, K, EPPX_VISC_CLASSIC, BATCH, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s,
                           w, f, ph, none);
    else
        hipLaunchKernelGGL((k_momentum<FAST, K, EPPX_VISC_ARTIFICIAL, BATCH, BLOCK, STAGE, false>), g, dim3(BLOCK), 0,
                           s, w, f, ph, none);
}

template <bool FAST, int K, bool BATCH>
static void launch_momentum_s(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t block, int32_t stage,
                              hipStream_t s) {
    if (FAST && block == 512 && stage) launch_momentum_b<FAST, K, BATCH, 512, 1>(w, f, ph, n, s);
    else if (FAST && block == 512) launch_momentum_b<FAST, K, BATCH, 512, 0>(w, f, ph, n, s);
    else if (FAST && stage) launch_momentum_b<FAST, K, BATCH, 256, 1>(w, f, ph, n, s);
    else launch_momentum_b<FAST, K, BATCH, 256, 0>(w, f, ph, n, s);
}

template <bool FAST, int K>
static void launch_momentum_v(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t block, int32_t stage,
                              hipStream_t s) {
    if (w.cases) launch_momentum_s<FAST, K, true>(w, f, ph, n, block, stage, s);
    else launch_momentum_s<FAST, K, false>(w, f, ph, n, block, stage, s);
}

template <bool FAST, int K, int BLOCK>
static void launch_momentum_move_kb(const Walk &w, Fields f, const Physics &ph, int32_t n, const MoveKey &mk,
                                    hipStream_t s) {
    const dim3 g = grid_for(n, BLOCK);
    if (ph.visc == EPPX_VISC_CLASSIC)
        hipLaunchKernelGGL((k_momentum<FAST, K, EPPX_VISC_CLASSIC, false, BLOCK, 0, true>), g, dim3(BLOCK), 0, s, w, f,
                           ph, mk);
    else
        hipLaunchKernelGGL((k_momentum<FAST, K, EPPX_VISC_ARTIFICIAL, false, BLOCK, 0, true>), g, dim3(BLOCK), 0, s, w,
                           f, ph, mk);
}

hipError_t launch_momentum_move(const Walk &w, Fields f, int32_t n, const Physics &ph, const MoveKey &mk,
                                int32_t fast, int32_t block, hipStream_t s) {
    constexpr int KC = EPPX_KERNEL_CUBIC, KW = EPPX_KERNEL_WENDLAND_C2;
    const bool cubic = ph.kernel == KC;
    if (fast && block == 512 && cubic) launch_momentum_move_kb<true, KC, 512>(w, f, ph, n, mk, s);
    else if (fast && block == 512) launch_momentum_move_kb<true, KW, 512>(w, f, ph, n, mk, s);
    else if (fast && cubic) launch_momentum_move_kb<true, KC, 256>(w, f, ph, n, mk, s);
    else if (fast) launch_momentum_move_kb<true, KW, 256>(w, f, ph, n, mk, s);
    else if (cubic) launch_momentum_move_kb<false, KC, 256>(w, f, ph, n, mk, s);
    else launch_momentum_move_kb<false, KW, 256>(w, f, ph, n, mk, s);
    return hipGetLastError();
}

hipError_t launch_interaction(const Walk &w, Fields f, int32_t n, int32_t action, const Physics &ph,
                              const eppx_field_op &op, int32_t fast, int32_t block, int32_t stage, hipStream_t s) {
    switch (action) {
        case EPPX_INTER_MOMENTUM: {
            const bool cubic = ph.kernel == EPPX_KERNEL_CUBIC;
            if (fast && cubic) launch_momentum_v<true, EPPX_KERNEL_CUBIC>(w, f, ph, n, block, stage, s);
            else if (fast) launch_momentum_v<true, EPPX_KERNEL_WENDLAND_C2>(w, f, ph, n, block, stage, s);
            else if (cubic) launch_momentum_v<false, EPPX_KERNEL_CUBIC>(w, f, ph, n, block, stage, s);
            else launch_momentum_v<false, EPPX_KERNEL_WENDLAND_C2>(w, f, ph, n, block, stage, s);
            break;
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

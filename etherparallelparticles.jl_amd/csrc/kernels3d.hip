// Device code of the 3D path of libeppx: the cell-list build, the pair
// kernels (neighbour search + kernel + continuity, momentum), the self
// kernels and the reference-layout conversion.  gfx950, wave64.
//
// Pair kernels keep one thread per particle.  A particle walks its 27
// neighbour cells in the stencil order (x offset outer, y, z inner) and each
// cell's slots in ascending particle index -- the reference's list order, so
// EXACT sums are the restatement's bit for bit.  A cell range is taken in
// chunks of 64 candidates: pass 1 tests every candidate of the chunk (one
// 16-byte load of x, y, z, tag each) into a 64-bit accept mask, pass 2 runs
// the pair physics over the set bits.  A cell holding no particle of a tag
// adjacent to the particle's is skipped as a whole (wall particles away from
// the fluid have no pairs and cost only the cell-tag reads).
#include "d3_internal.h"
#include "pair_device.h"

namespace eppx {
namespace {

constexpr int kBlock3 = 256;

// sigma of the 3D kernels: FT(1/pi) (CubicSpline.jl:15), FT(21/16/pi) (WendlandC2.jl:15)
constexpr float kSigmaCubic3 = (float)(1.0 / 3.14159265358979323846);
constexpr float kSigmaWendland3 = (float)(21.0 / 16.0 / 3.14159265358979323846);

// _value with N = 3 (CubicSpline.jl:17-31, WendlandC2.jl:16-26) in the
// written order; power(hinv, Val(3)) = hinv * hinv * hinv,
// power(2 - q, Val(4)) = (t * t) * (t * t) (OptimizedFunction.jl:10-31).
__device__ __forceinline__ float value3_exact(int kernel, float r, float hinv) {
    const float q = r * hinv;
    const float h3 = hinv * hinv * hinv;
    if (kernel == EPPX_KERNEL_CUBIC) {
        if (q < 1.0f) return kSigmaCubic3 * h3 * (3.0f * q * q * (q - 2.0f) + 4.0f) * 0.25f;
        if (q < 2.0f) {
            const float t = 2.0f - q;
            return kSigmaCubic3 * h3 * (t * t * t) * 0.25f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        const float t = 2.0f - q;
        const float t2 = t * t;
        return kSigmaWendland3 * (t2 * t2) * h3 * (1.0f + 2.0f * q) * 0.0625f;
    }
    return 0.0f;
}

// _gradient with N = 3 (CubicSpline.jl:41-55, WendlandC2.jl:37-47), written
// order; power(hinv, Val(4)) = (hinv * hinv) * (hinv * hinv).
template <int K>
__device__ __forceinline__ float grad3_exact(float r, float hinv) {
    const float q = r * hinv;
    const float hh = hinv * hinv;
    const float h4 = hh * hh;
    if (K == EPPX_KERNEL_CUBIC) {
        if (q < 1.0f) return kSigmaCubic3 * h4 * q * (3.0f * q - 4.0f) * 0.75f;
        if (q < 2.0f) {
            const float t = 2.0f - q;
            return -kSigmaCubic3 * h4 * t * t * 0.75f;
        }
        return 0.0f;
    }
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return -kSigmaWendland3 * h4 * 0.625f * q * (t * t * t);
    }
    return 0.0f;
}

// dw / r of the same kernels, contracted (FAST): Wendland C2's dw carries q =
// r hinv, so dw / r = (-sigma 5/8) hinv^5 (2 - q)^3 needs no reciprocal.
template <int K>
__device__ __forceinline__ float grad3_over_r_fast(float r, float rinv, float hinv) {
    const float hh = hinv * hinv;
    const float h4 = hh * hh;
    const float q = r * hinv;
    if (K == EPPX_KERNEL_CUBIC) {
        const float t = 2.0f - q;
        const float inner = (kSigmaCubic3 * 0.75f) * (h4 * hinv) * fmaf(3.0f, q, -4.0f);
        const float outer = (-kSigmaCubic3 * 0.75f) * h4 * (t * t) * rinv;
        return q < 1.0f ? inner : (q < 2.0f ? outer : 0.0f);
    }
    const float t = fmaxf(2.0f - q, 0.0f);
    return (-kSigmaWendland3 * 0.625f) * (h4 * hinv) * (t * (t * t));
}

// The criterion of device_findNeighbourParticlesFromCells! for a pair in the
// band the host bounds leave open (NeighbourSearch.jl:141-164).
__device__ __forceinline__ bool criterion3(const Walk3 &w, float hi, float hj, float dr2) {
    if (w.criterion == EPPX_CRIT_DEFAULT) return dr2 <= w.g.gap_square;
    const float h2 = harmonic(hi, hj) * 2.0f;
    return dr2 <= h2 * h2;
}

// dr2 = dx*dx + dy*dy + dz*dz, left to right, no FMA (the 2D sum with the z term added).
__device__ __forceinline__ float dist2(float dx, float dy, float dz) { return dx * dx + dy * dy + dz * dz; }

// Accept mask of candidates [base, base + len), len <= 64, of slot i: bit k
// when base + k != i, (CHECK) the tags are adjacent and the criterion
// accepts.  Candidates are taken four at a time; the slots read past the
// range stay inside the padded arrays and are masked off.
template <bool CHECK, bool BAND>
__device__ __forceinline__ unsigned long long accept3(const Walk3 &w, int i, float4 gi, float hi, uint32_t adj_i,
                                                      int base, int len) {
    unsigned long long mask = 0ull;
    for (int k = 0; k < len; k += 4) {
        uint32_t nib = 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 gj = ld_slot(w.geo, base + k + u);
            const float dr2 = dist2(gi.x - gj.x, gi.y - gj.y, gi.z - gj.z);
            bool ok = dr2 < w.rmin2;
            if (BAND && !ok && dr2 <= w.rmax2) ok = criterion3(w, hi, ld_slot(w.hmvp, base + k + u).x, dr2);
            if (CHECK) ok = ok && (adj_i & tag_bit(__float_as_int(gj.w))) != 0u;
            nib |= ok ? (1u << u) : 0u;
        }
        mask |= (unsigned long long)nib << k;
    }
    if (len < 64) mask &= (1ull << len) - 1ull;
    const unsigned d = (unsigned)(i - base);
    if (d < 64u) mask &= ~(1ull << d);
    return mask;
}

// Visits the accepted neighbours of sorted slot i in the reference's list
// order: visit(j, geo_j, dx, dy, dz, dr2).
template <bool CHECK, bool BAND, class Visit>
__device__ __forceinline__ void walk3(const Walk3 &w, int i, float4 gi, float hi, uint32_t adj_i, Visit &&visit) {
    const Grid3 &g = w.g;
    const int c = w.cell[i] - 1;
    const int cx = c % g.n_x, t = c / g.n_x;
    const int cy = t % g.n_y, cz = t / g.n_y;
    for (int di = -1; di <= 1; ++di) {
        const int x = cx + di;
        if (x < 0 || x >= g.n_x) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int y = cy + dj;
            if (y < 0 || y >= g.n_y) continue;
            for (int dk = -1; dk <= 1; ++dk) {
                const int z = cz + dk;
                if (z < 0 || z >= g.n_z) continue;
                const int nb = x + g.n_x * (y + g.n_y * z);
                if (!(w.cell_tags[nb] & adj_i)) continue;
                const int b = w.start[nb], e = w.start[nb + 1];
                for (int base = b; base < e; base += 64) {
                    unsigned long long m = accept3<CHECK, BAND>(w, i, gi, hi, adj_i, base, min(64, e - base));
                    while (m) {
                        const int j = base + (__ffsll((long long)m) - 1);
                        m &= m - 1ull;
                        const float4 gj = ld_slot(w.geo, j);
                        const float dx = gi.x - gj.x, dy = gi.y - gj.y, dz = gi.z - gj.z;
                        visit(j, gj, dx, dy, dz, dist2(dx, dy, dz));
                    }
                }
            }
        }
    }
}

// walk3 with the candidate tag test only in waves holding a particle that
// is not adjacent to every present tag, and the exact criterion band only
// when the host bounds leave one.
template <class Visit>
__device__ __forceinline__ void walk3_any(const Walk3 &w, int i, float4 gi, float hi, uint32_t adj_i, bool check,
                                          Visit &&visit) {
    const bool band = w.rmin2 <= w.rmax2;
    if (__ballot(check)) {
        if (band) walk3<true, true>(w, i, gi, hi, adj_i, visit);
        else walk3<true, false>(w, i, gi, hi, adj_i, visit);
    } else {
        if (band) walk3<false, true>(w, i, gi, hi, adj_i, visit);
        else walk3<false, false>(w, i, gi, hi, adj_i, visit);
    }
}

__device__ __forceinline__ uint32_t adj3(const Walk3 &w, uint32_t tb) { return tb ? w.adj[__builtin_ctz(tb) + 1] : 0u; }

// iValueGradient! + iClassicContinuity! / iBalancedContinuity!
// (Library/Kernel.jl:114-128, Continuity.jl:23-47, vdotx Library.jl:21-34
// with N = 3); coef is m_j (fluid-fluid) or rho_i V_j (fluid-wall, wall-fluid).
template <bool FAST, int K>
__device__ __forceinline__ float continuity3(float hinv, float4 vi, float4 vj, float coef, float dx, float dy,
                                             float dz, float dr2) {
    if (!FAST) {
        const float r = sqrtf(dr2);
        const float dw = grad3_exact<K>(r, hinv);
        float vdotx = 0.0f;
        vdotx += dx * (vi.x - vj.x);
        vdotx += dy * (vi.y - vj.y);
        vdotx += dz * (vi.z - vj.z);
        return coef * vdotx * dw / r;
    }
    const float rinv = frsq(dr2);
    const float dwr = grad3_over_r_fast<K>(dr2 * rinv, rinv, hinv);
    const float vdotx = fmaf(dz, vi.z - vj.z, fmaf(dy, vi.y - vj.y, dx * (vi.x - vj.x)));
    return (coef * vdotx) * dwr;
}

// iClassicPressure! / iBalancedPressure! (Pressure.jl:24-60) and
// iClassicViscosity! / sArtificialViscosity! (Viscosity.jl:10-52) with N = 3,
// in the order iMomentum! (main.jl:107-118) adds them to du.  pr_i = p_i / rho_i^2.
template <bool FAST, int K, int VISC>
__device__ __forceinline__ void momentum3(const Physics &ph, float4 &acc, bool classic, float hinv, float4 vi,
                                          float p_i, float pr_i, float4 vj, float4 qj, float dx, float dy, float dz,
                                          float dr2) {
    const float rho_i = vi.w, rho_j = vj.w, p_j = qj.w, m_j = qj.y;
    const float dux = vi.x - vj.x, duy = vi.y - vj.y, duz = vi.z - vj.z;
    if (!FAST) {
        const float r = sqrtf(dr2);
        const float dw = grad3_exact<K>(r, hinv);
        float prho2 = classic ? pr_i + p_j / (rho_j * rho_j) : (p_i + p_j) / (rho_i * rho_j);
        prho2 += fabsf(prho2) * ph.coefficient;
        prho2 *= -m_j * dw / r;
        acc.x += prho2 * dx;
        acc.y += prho2 * dy;
        acc.z += prho2 * dz;
        const float h = 1.0f / hinv;
        const float az = r * r + 0.01f * h * h;  // avoidzero, Library.jl:17-19
        if (VISC == EPPX_VISC_CLASSIC) {
            const float force = 2.0f * ph.mu * m_j * r * dw / (rho_i * rho_j * az);
            acc.x += force * dux;
            acc.y += force * duy;
            acc.z += force * duz;
        } else {
            float vdotx = 0.0f;
            vdotx += dx * dux;
            vdotx += dy * duy;
            vdotx += dz * duz;
            if (vdotx > 0.0f) return;
            const float rho = (rho_i + rho_j) * 0.5f;
            const float phi = vdotx / (hinv * az);
            float force = (-ph.alpha * ph.av_c + ph.beta * phi) * phi / rho;
            force *= -m_j / (rho_i * rho_j * r) * dw;
            acc.x += force * dx;
            acc.y += force * dy;
            acc.z += force * dz;
        }
        return;
    }
    // FAST: dw / r from the contracted kernel; r^2 + 0.01 h^2 = h^2 (q^2 + 0.01)
    const float rinv = frsq(dr2);
    const float mdwr = -m_j * grad3_over_r_fast<K>(dr2 * rinv, rinv, hinv);
    float prho2 = classic ? fmaf(p_j, frcp(rho_j * rho_j), pr_i) : (p_i + p_j) * frcp(rho_i * rho_j);
    prho2 = fmaf(fabsf(prho2), ph.coefficient, prho2);
    const float radial = prho2 * mdwr;
    acc.x = fmaf(radial, dx, acc.x);
    acc.y = fmaf(radial, dy, acc.y);
    acc.z = fmaf(radial, dz, acc.z);
    const float q2 = dr2 * (hinv * hinv);
    const float qa = q2 + 0.01f;
    if (VISC == EPPX_VISC_CLASSIC) {
        const float force = (-2.0f * ph.mu) * mdwr * q2 * frcp(rho_i * rho_j * qa);
        acc.x = fmaf(force, dux, acc.x);
        acc.y = fmaf(force, duy, acc.y);
        acc.z = fmaf(force, duz, acc.z);
    } else {
        // a separating pair (vdotx > 0) gets phi = 0 and with it no term
        const float vdotx = fminf(fmaf(dz, duz, fmaf(dy, duy, dx * dux)), 0.0f);
        const float phi = vdotx * hinv * frcp(qa);
        const float force = fmaf(ph.beta, phi, -ph.alpha * ph.av_c) * phi * mdwr *
                            frcp((rho_i + rho_j) * 0.5f * (rho_i * rho_j));
        acc.x = fmaf(force, dx, acc.x);
        acc.y = fmaf(force, dy, acc.y);
        acc.z = fmaf(force, dz, acc.z);
    }
}

// One H and one Mass in the case (FAST): every pair term is the 3D kernel's
// constant factor kfac (Wendland C2: -sigma 0.625 hinv^5, cubic spline:
// sigma 0.75 hinv^4) times the shape of UniformTerms, so the pair sums carry
// the rest and the caller adds kfac x sum (and m_u) once.
template <int K>
struct Uniform3 : UniformTerms<K> {
    __device__ __forceinline__ Uniform3(float hinv_u, const Physics &ph) : UniformTerms<K>(hinv_u, ph) {
        const float h4 = (hinv_u * hinv_u) * (hinv_u * hinv_u);
        this->kfac = K == EPPX_KERNEL_CUBIC ? (kSigmaCubic3 * 0.75f) * h4 : (-kSigmaWendland3 * 0.625f) * (h4 * hinv_u);
    }
};

// momentum3 (FAST) over -m_u kfac, one H and one Mass, coefficient 0, with the
// staged record qj of StageX: (Mass, p_j / rho_j^2) for a classic (fluid)
// neighbour, (1 / rho_j, p_j / rho_j) for a balanced (wall) one, whose
// (p_i + p_j) / (rho_i rho_j) is then (p_i / rho_j + p_j / rho_j) / rho_i.
// take = false: the pair adds nothing (its term is computed all the same and
// dropped by a select, so not even a NaN of it gets through).
template <int K, int VISC>
__device__ __forceinline__ void momentum_fast_um3(const Physics &ph, float4 &sum, bool classic, bool take,
                                                  const UniformTerms<K> &ut, float4 vi, float p_i, float pr_i,
                                                  float inv_rho_i, float4 vj, float4 qj, float dx, float dy, float dz,
                                                  float dr2) {
    const float rho_i = vi.w, rho_j = vj.w;
    const float rinv = frsq(dr2);
    const float shape = ut.shape(dr2 * rinv, rinv);
    const float rirj = rho_i * rho_j;
    const float prho2 = classic ? pr_i + qj.w : fmaf(p_i, qj.y, qj.w) * inv_rho_i;
    float radial = prho2 * shape;
    const float qa = fmaf(dr2, ut.hinv2, 0.01f);  // q^2 + 0.01
    const float dux = vi.x - vj.x, duy = vi.y - vj.y, duz = vi.z - vj.z;
    if (VISC == EPPX_VISC_CLASSIC) {
        float force = (-2.0f * ph.mu) * (dr2 * ut.hinv2) * shape * frcp(rirj * qa);
        force = take ? force : 0.0f;
        sum.x = fmaf(force, dux, sum.x);
        sum.y = fmaf(force, duy, sum.y);
        sum.z = fmaf(force, duz, sum.z);
    } else {
        // a separating pair (vdotx > 0) gets no term
        const float m = fminf(fmaf(dz, duz, fmaf(dy, duy, dx * dux)), 0.0f);
        const float den = (qa * qa) * ((rho_i + rho_j) * rirj);
        radial = fmaf(fmaf(ut.av_b, m, -ut.av_a * qa) * m, shape * frcp(den), radial);
    }
    radial = take ? radial : 0.0f;
    sum.x = fmaf(radial, dx, sum.x);
    sum.y = fmaf(radial, dy, sum.y);
    sum.z = fmaf(radial, dz, sum.z);
}

// ------------------------------------------------------------ pair kernels
// One wave per cell (four cells per 256-thread block, each wave on its own):
// lane l holds the cell's l-th particle (cells of more than 64 particles take
// their particles 64 at a time).  The particles of a cell share their 27
// neighbour cells, so the wave stages each neighbour cell's records (geo,
// vel, hmvp: 48 bytes a slot, read once and coalesced) in its own LDS area
// and every lane tests and gathers from there: pass 1 reads one candidate for
// all lanes at once (a broadcast), pass 2 each lane's own accepted
// candidates.  Cells are walked in the stencil order and each cell's slots
// in ascending particle index, the reference's list order.
constexpr int kCellBlock = 64;     // lanes of a wave: particles of one cell at a time
constexpr int kCellWaves = 4;      // waves (cells) per block

// Block -> group of kCellWaves consecutive cells.  The groups are cut into
// stripes of `stripe` consecutive groups; the dispatcher hands block b to XCD
// b % 8, and XCD x works through stripes x, x + 8, x + 16, ... in order.  The
// eight XCDs thus work on neighbouring stripes at any time, which keeps them
// equally loaded where the particles fill only part of the grid (one
// contiguous range per XCD would leave XCDs idle), and each stripe's cells
// share most of their neighbour cells in that XCD's L2.  The grid is
// 8 x stripes_per_xcd x stripe blocks; groups past the last cell have no work.
__device__ __forceinline__ int stripe_group(int b, int stripe) {
    const int x = b % kXcds, l = b / kXcds;
    return (x + kXcds * (l / stripe)) * stripe + l % stripe;
}

// Orders a wave's LDS stores before its later LDS loads (the waves of a block
// never share a staging area, so no block barrier is needed).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave's staging area: the slots of up to two neighbour cells (64 each),
// the second cell's right after the first's.
struct Stage3 {
    float4 g[2 * kCellBlock];  // (x, y, z, tag bit) of the staged slots
    float4 v[2 * kCellBlock];  // (u, v, w, rho)
    float4 q[2 * kCellBlock];  // (H, Mass, Volume, Pressure); see StageX
};

// Adjacency of any tag present in cell c (a cell holding no particle of a tag
// in it is skipped by every lane of the block).
__device__ __forceinline__ uint32_t adj_of_cell(const Walk3 &w, uint32_t tags) {
    uint32_t a = 0u;
    while (tags) {
        const int t = __builtin_ctz(tags);
        tags &= tags - 1u;
        a |= w.adj[t + 1];
    }
    return a;
}

// The reach filter of a wave: the bounding box of its particles' positions
// (the lanes holding one) and the largest dr2 the criterion can accept,
// widened by 1e-4 relative -- far more than the float32 rounding of dr2 or of
// the box distance.  A candidate farther than that from the box is no lane's
// neighbour, so it is not staged.  Of the 27 cells around a full cell of a
// lattice at h = 2dx (1,728 candidates) about 970 are within reach.
struct Reach3 {
    float lx, ly, lz, hx, hy, hz, lim2;
};

__device__ __forceinline__ Reach3 wave_reach(const Walk3 &w, float4 gi, bool valid) {
    Reach3 r;
    r.lx = valid ? gi.x : INFINITY;
    r.ly = valid ? gi.y : INFINITY;
    r.lz = valid ? gi.z : INFINITY;
    r.hx = valid ? gi.x : -INFINITY;
    r.hy = valid ? gi.y : -INFINITY;
    r.hz = valid ? gi.z : -INFINITY;
    for (int off = 32; off > 0; off >>= 1) {
        r.lx = fminf(r.lx, __shfl_xor(r.lx, off));
        r.ly = fminf(r.ly, __shfl_xor(r.ly, off));
        r.lz = fminf(r.lz, __shfl_xor(r.lz, off));
        r.hx = fmaxf(r.hx, __shfl_xor(r.hx, off));
        r.hy = fmaxf(r.hy, __shfl_xor(r.hy, off));
        r.hz = fmaxf(r.hz, __shfl_xor(r.hz, off));
    }
    r.lim2 = fmaxf(w.rmin2, w.rmax2) * 1.0001f;
    return r;
}

// What the one-H, one-Mass momentum stages in place of a slot's (Mass,
// Pressure): for the tag bit `fluid`, (Mass, p / rho^2) (pressure_ratio); for
// `wall`, (1 / rho, p / rho).  Zero bits: the records as they are.
struct StageX {
    uint32_t fluid = 0u, wall = 0u;
};

// Stages the slots of [base, base + len) (len <= 64) within reach, in slot
// order, from position off: slot base + L goes to off + popcount(bits of the
// ballot below L).  Returns the ballot.  A staged g.w holds the slot's tag
// bit; sx: see StageX.  Every lane of the wave calls it.
__device__ __forceinline__ unsigned long long stage_reach(const Walk3 &w, Stage3 &st, const Reach3 &r, int base,
                                                          int len, int off, StageX sx) {
    const int lane = lane_id();
    float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool in = false;
    if (lane < len) {
        g = w.geo[base + lane];
        const float dx = fmaxf(0.0f, fmaxf(r.lx - g.x, g.x - r.hx));
        const float dy = fmaxf(0.0f, fmaxf(r.ly - g.y, g.y - r.hy));
        const float dz = fmaxf(0.0f, fmaxf(r.lz - g.z, g.z - r.hz));
        in = !(dx * dx + dy * dy + dz * dz > r.lim2);
    }
    const unsigned long long bal = __ballot(in);
    if (in) {
        const int pos = off + __popcll(bal & ((1ull << lane) - 1ull));
        const uint32_t tb = tag_bit(__float_as_int(g.w));
        g.w = __uint_as_float(tb);
        const float4 v = w.vel[base + lane];
        float4 q = w.hmvp[base + lane];
        if (sx.fluid && tb == sx.fluid) {
            q.w = pressure_ratio(q.w, v.w);
        } else if (sx.wall && tb == sx.wall) {
            const float ir = frcp(v.w);
            q.y = ir;
            q.w *= ir;
        }
        st.g[pos] = g;
        st.v[pos] = v;
        st.q[pos] = q;
    }
    return bal;
}

// Pass 1 over the staged slots [off, off + len) against this lane's particle
// (bit k: slot off + k).
template <bool CHECK, bool BAND>
__device__ __forceinline__ unsigned long long accept_staged(const Walk3 &w, const Stage3 &st, int off, float4 gi,
                                                            float hi, uint32_t adj_i, int len) {
    unsigned long long mask = 0ull;
    for (int k = 0; k < len; k += 4) {  // reads up to 3 slots past the range, inside the area (off <= 64)
        uint32_t nib = 0u;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 gj = st.g[off + k + u];
            const float dr2 = dist2(gi.x - gj.x, gi.y - gj.y, gi.z - gj.z);
            bool ok = dr2 < w.rmin2;
            if (BAND && !ok && dr2 <= w.rmax2) ok = criterion3(w, hi, st.q[off + k + u].x, dr2);
            if (CHECK) ok = ok && (adj_i & __float_as_uint(gj.w)) != 0u;
            nib |= ok ? (1u << u) : 0u;
        }
        mask |= (unsigned long long)nib << k;
    }
    return len < 64 ? mask & ((1ull << len) - 1ull) : mask;
}

// This lane's own slot among the staged ones (bal: the staging ballot of the
// range starting at base): its bit, or 0 when it is not in the range.
__device__ __forceinline__ unsigned long long self_bit(unsigned long long bal, int i, int base) {
    const unsigned d = (unsigned)(i - base);
    if (d >= 64u || !((bal >> d) & 1ull)) return 0ull;
    return 1ull << __popcll(bal & ((1ull << d) - 1ull));
}

// The walk of one 64-particle chunk of cell c: visit(geo_j, vel_j, hmvp_j,
// dx, dy, dz, dr2) for each accepted neighbour j (geo_j.w: its tag bit).
// `active`: this lane holds a particle that takes pairs; every lane of the
// wave calls it, with the wave's reach filter.  Stored masks (mask_mode 1 and
// 2) index the staged order of a cell of at most 64 particles.
template <bool CHECK, bool BAND, class Visit>
__device__ __forceinline__ void cell_walk3(const Walk3 &w, Stage3 &st, int c, uint32_t adj_block, int i, bool active,
                                           const Reach3 &reach, StageX sx, float4 gi, float hi, uint32_t adj_i,
                                           Visit &&visit) {
    const Grid3 &g = w.g;
    const int cx = c % g.n_x, t = c / g.n_x;
    const int cy = t % g.n_y, cz = t / g.n_y;
    for (int di = -1; di <= 1; ++di) {
        const int x = cx + di;
        if (x < 0 || x >= g.n_x) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int y = cy + dj;
            if (y < 0 || y >= g.n_y) continue;
            for (int dk = -1; dk <= 1; ++dk) {
                const int z = cz + dk;
                if (z < 0 || z >= g.n_z) continue;
                const int nb = x + g.n_x * (y + g.n_y * z);
                if (!(w.cell_tags[nb] & adj_block)) continue;
                const int b = w.start[nb], e = w.start[nb + 1];
                const size_t mi = (size_t)((di + 1) * 9 + (dj + 1) * 3 + dk + 1) * (size_t)w.mask_stride + i;
                const bool small = e - b <= kCellBlock;  // one chunk: its mask is stored
                const bool load = w.mask_mode == 2 && small;
                unsigned long long stored = 0ull;
                if (load) {
                    if (active) stored = w.masks[mi];
                    if (!__ballot(stored != 0ull)) continue;  // no lane has a pair in this cell
                }
                for (int base = b; base < e; base += kCellBlock) {
                    const int raw = min(kCellBlock, e - base);
                    const unsigned long long bal = stage_reach(w, st, reach, base, raw, 0, sx);
                    const int len = __popcll(bal);
                    wave_sync();
                    if (active) {
                        unsigned long long m;
                        if (load) {
                            m = stored;
                        } else {
                            m = len ? accept_staged<CHECK, BAND>(w, st, 0, gi, hi, adj_i, len) & ~self_bit(bal, i, base)
                                    : 0ull;
                            if (w.mask_mode == 1 && small) w.masks[mi] = m;
                        }
                        while (m) {
                            const int k = __ffsll((long long)m) - 1;
                            m &= m - 1ull;
                            // the three records at once: one LDS round trip per pair
                            const float4 gj = st.g[k], vj = st.v[k], qj = st.q[k];
                            const float dx = gi.x - gj.x, dy = gi.y - gj.y, dz = gi.z - gj.z;
                            visit(gj, vj, qj, dx, dy, dz, dist2(dx, dy, dz));
                        }
                    }
                    wave_sync();
                }
            }
        }
    }
}

// FAST walk of one 64-particle chunk of cell c: the neighbour cells in 14
// groups -- the cell itself, then each offset together with its mirror (the
// stencil cells s and 26 - s) -- with both cells of a group staged at once,
// one after the other, and their accepted pairs walked as one 128-bit mask.
// A particle near one face of its cell has many pairs in the cell beyond that
// face and few in the cell beyond the opposite one, so a group's pair count
// varies much less across the wave than a cell's: the wave runs about half
// the pass-2 iterations of the cell-by-cell walk.  Sums then run in another
// order than the reference's list order (FAST math only).
template <bool CHECK, bool BAND, class Visit>
__device__ __forceinline__ void cell_walk3_paired(const Walk3 &w, Stage3 &st, int c, uint32_t adj_block, int i,
                                                  bool active, const Reach3 &reach, StageX sx, float4 gi,
                                                  float hi, uint32_t adj_i, Visit &&visit) {
    const Grid3 &g = w.g;
    const int cx = c % g.n_x, t = c / g.n_x;
    const int cy = t % g.n_y, cz = t / g.n_y;
    // slot range of stencil cell s (x offset s / 9 - 1, y (s / 3) % 3 - 1, z s % 3 - 1), empty when the cell
    // lies outside the grid or holds no particle of a tag adjacent to the wave's
    auto range = [&](int sidx, int &b, int &e) {
        const int x = cx + sidx / 9 - 1, y = cy + (sidx / 3) % 3 - 1, z = cz + sidx % 3 - 1;
        b = e = 0;
        if (x < 0 || x >= g.n_x || y < 0 || y >= g.n_y || z < 0 || z >= g.n_z) return;
        const int nb = x + g.n_x * (y + g.n_y * z);
        if (!(w.cell_tags[nb] & adj_block)) return;
        b = w.start[nb];
        e = w.start[nb + 1];
    };
    const size_t ms = (size_t)w.mask_stride;
    for (int grp = 0; grp < 14; ++grp) {
        int ba, ea, bb, eb;
        range(13 + grp, ba, ea);
        if (grp) range(13 - grp, bb, eb);
        else bb = eb = 0;
        const int na = ea - ba, nbb = eb - bb;
        // a cell of at most 64 particles is one chunk, whose mask the search stores
        const bool small_a = na > 0 && na <= kCellBlock, small_b = nbb > 0 && nbb <= kCellBlock;
        const size_t mia = (size_t)(13 + grp) * ms + i, mib = (size_t)(13 - grp) * ms + i;
        const bool load_a = w.mask_mode == 2 && small_a, load_b = w.mask_mode == 2 && small_b;
        unsigned long long sa = 0ull, sb = 0ull;
        if (active) {
            if (load_a) sa = w.masks[mia];
            if (load_b) sb = w.masks[mib];
        }
        // a stored mask that is empty in every lane: the cell needs no staging
        const bool need_a = na > 0 && (!load_a || __ballot(sa != 0ull) != 0ull);
        const bool need_b = nbb > 0 && (!load_b || __ballot(sb != 0ull) != 0ull);
        if (!need_a && !need_b) continue;
        const int chunks = (max(need_a ? na : 0, need_b ? nbb : 0) + kCellBlock - 1) / kCellBlock;
        for (int ch = 0; ch < chunks; ++ch) {
            const int base_a = ba + ch * kCellBlock, base_b = bb + ch * kCellBlock;
            const int raw_a = need_a ? min(kCellBlock, max(0, ea - base_a)) : 0;
            const int raw_b = need_b ? min(kCellBlock, max(0, eb - base_b)) : 0;
            const unsigned long long bal_a = stage_reach(w, st, reach, base_a, raw_a, 0, sx);
            const int len_a = __popcll(bal_a);
            const unsigned long long bal_b = stage_reach(w, st, reach, base_b, raw_b, len_a, sx);
            const int len_b = __popcll(bal_b);
            wave_sync();
            if (active) {
                unsigned long long ma = 0ull, mb = 0ull;
                if (load_a) {
                    ma = raw_a ? sa : 0ull;
                } else if (raw_a) {
                    ma = len_a ? accept_staged<CHECK, BAND>(w, st, 0, gi, hi, adj_i, len_a) : 0ull;
                    if (grp == 0) ma &= ~self_bit(bal_a, i, base_a);
                    if (w.mask_mode == 1 && small_a) w.masks[mia] = ma;
                }
                if (load_b) {
                    mb = raw_b ? sb : 0ull;
                } else if (raw_b) {
                    mb = len_b ? accept_staged<CHECK, BAND>(w, st, len_a, gi, hi, adj_i, len_b) : 0ull;
                    if (w.mask_mode == 1 && small_b) w.masks[mib] = mb;
                }
                // one 128-bit mask over the staged slots [0, len_a + len_b): bits of the
                // current 64-slot word in cur, the rest in nxt
                unsigned long long cur = ma, nxt = 0ull;
                if (len_a == 0) {
                    cur = mb;
                } else if (len_a < 64) {
                    cur |= mb << len_a;
                    nxt = mb >> (64 - len_a);
                } else {
                    nxt = mb;
                }
                int off = 0;
                if (!cur) {
                    cur = nxt;
                    nxt = 0ull;
                    off = 64;
                }
                while (cur) {
                    const int k = off + __builtin_ctzll(cur);
                    cur &= cur - 1ull;
                    if (!cur) {  // the second word, if any, after the last bit of the first
                        cur = nxt;
                        nxt = 0ull;
                        off = 64;
                    }
                    const float4 gj = st.g[k], vj = st.v[k], qj = st.q[k];
                    const float dx = gi.x - gj.x, dy = gi.y - gj.y, dz = gi.z - gj.z;
                    visit(gj, vj, qj, dx, dy, dz, dist2(dx, dy, dz));
                }
            }
            wave_sync();
        }
    }
}

template <class Visit>
__device__ __forceinline__ void cell_walk3_any(const Walk3 &w, Stage3 &st, int c, uint32_t adj_block, int i,
                                               bool active, const Reach3 &r, StageX sx, float4 gi, float hi,
                                               uint32_t adj_i, bool check, Visit &&visit) {
    const bool band = w.rmin2 <= w.rmax2;
    if (__ballot(check)) {
        if (band) cell_walk3<true, true>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
        else cell_walk3<true, false>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
    } else {
        if (band) cell_walk3<false, true>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
        else cell_walk3<false, false>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
    }
}

// valid: this lane holds a particle of the chunk (its position joins the
// wave's reach filter, which the search and momentum over the same cell list
// then compute alike, so stored masks index the same staged order).
template <bool FAST, class Visit>
__device__ __forceinline__ void cell_walk3_sel(const Walk3 &w, Stage3 &st, int c, uint32_t adj_block, int i,
                                               bool valid, bool active, StageX sx, float4 gi, float hi,
                                               uint32_t adj_i, bool check, Visit &&visit) {
    const Reach3 r = wave_reach(w, gi, valid);
    if (!FAST) {
        cell_walk3_any(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, check, visit);
        return;
    }
    const bool band = w.rmin2 <= w.rmax2;
    if (__ballot(check)) {
        if (band) cell_walk3_paired<true, true>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
        else cell_walk3_paired<true, false>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
    } else {
        if (band) cell_walk3_paired<false, true>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
        else cell_walk3_paired<false, false>(w, st, c, adj_block, i, active, r, sx, gi, hi, adj_i, visit);
    }
}


// Search pass: the neighbours of every alive particle, counted into ncount,
// with the fused action (EPPX_SEARCH_KERNEL_CONTINUITY:
// sComputeKernelAndiContinuity!, main.jl:74-90, onto the existing dDensity).
// UM (FAST, continuity action): every alive particle has one H and one Mass
// (w.hinv_u, w.m_u), and the terms leave them out (a kernel of its own, so
// the general path's registers do not limit it, nor its the general path).
// Six waves per SIMD (at most 80 VGPRs; a few spill on rare paths) run the
// search faster than five; momentum spills in its pair loop at six.
template <bool FAST, int ACTION, int K, bool UM>
__global__ __launch_bounds__(kCellBlock * kCellWaves, 6) void k3_search(Walk3 w, Fields3 f, Physics ph, int32_t M,
                                                                     int32_t *ncount, unsigned long long *pairs,
                                                                     uint32_t *flags) {
    __shared__ Stage3 stage[kCellWaves];
    const int wave = threadIdx.x / kCellBlock;
    Stage3 &st = stage[wave];
    const int c = stripe_group(blockIdx.x, w.stripe) * kCellWaves + wave;
    if (c >= w.g.n_cells) return;
    const int c0 = w.start[c], c1 = w.start[c + 1];
    if (c0 >= c1) return;
    const uint32_t cell_tags = w.cell_tags[c];
    const uint32_t adj_block = adj_of_cell(w, cell_tags);
    const uint32_t fluid_bit = tag_bit(ph.fluid_tag), wall_bit = tag_bit(ph.wall_tag);
    int total = 0;
    for (int i0 = c0; i0 < c1; i0 += kCellBlock) {
        const int i = i0 + lane_id();
        const bool active = i < c1;
        int cnt = 0;
        float4 gi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = gi;
        float hi = 0.0f, drho = 0.0f;
        uint32_t tbi = 0u;
        if (active) {
            gi = w.geo[i];
            vi = w.vel[i];
            hi = w.hmvp[i].x;
            tbi = tag_bit(__float_as_int(gi.w));
            if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) drho = f.drho[i];
        }
        const uint32_t adj_i = adj3(w, tbi);
        const bool fluid_i = tbi && tbi == fluid_bit, wall_i = tbi && tbi == wall_bit;
        const bool check = active && (w.odd_tags || (adj_i & w.tags_present) != w.tags_present);
        if constexpr (UM) {
            // The fluid-fluid terms (classic continuity) leave out kfac and m_j, the mixed ones
            // (balanced) kfac only
            const Uniform3<K> ut(w.hinv_u, ph);
            float s_ff = 0.0f, s_mx = 0.0f;
            cell_walk3_sel<FAST>(w, st, c, adj_block, i, active, active && adj_i != 0u, StageX{}, gi, hi, adj_i, check,
                           [&](float4 gj, float4 vj, float4 qj, float dx, float dy, float dz, float dr2) {
                               ++cnt;
                               const uint32_t tbj = __float_as_uint(gj.w);
                               const bool classic = fluid_i && tbj == fluid_bit;
                               const bool balanced = (fluid_i && tbj == wall_bit) || (wall_i && tbj == fluid_bit);
                               // (selects, not a branch: the pair's records are read at once)
                               const float rinv = frsq(dr2);
                               const float shape = ut.shape(dr2 * rinv, rinv);
                               const float vdotx = fmaf(dz, vi.z - vj.z, fmaf(dy, vi.y - vj.y, dx * (vi.x - vj.x)));
                               s_ff += classic ? vdotx * shape : 0.0f;
                               s_mx += balanced ? ((vi.w * qj.z) * vdotx) * shape : 0.0f;
                           });
            drho = fmaf(ut.kfac, s_mx, fmaf(ut.kfac * w.m_u, s_ff, drho));
        } else {
            cell_walk3_sel<FAST>(w, st, c, adj_block, i, active, active && adj_i != 0u, StageX{}, gi, hi, adj_i, check,
                           [&](float4 gj, float4 vj, float4 qj, float dx, float dy, float dz, float dr2) {
                               ++cnt;
                               if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) {
                                   const uint32_t tbj = __float_as_uint(gj.w);
                                   const bool classic = fluid_i && tbj == fluid_bit;
                                   const bool balanced =
                                       (fluid_i && tbj == wall_bit) || (wall_i && tbj == fluid_bit);
                                   if (classic || balanced) {
                                       const float hinv = w.hinv_u > 0.0f ? w.hinv_u : invharmonic(hi, qj.x);
                                       const float coef = classic ? qj.y : vi.w * qj.z;
                                       drho += continuity3<FAST, K>(hinv, vi, vj, coef, dx, dy, dz, dr2);
                                   }
                               }
                           });
        }
        if (active) {
            ncount[i] = cnt;
            if (cnt > M) atomicOr(flags, kFlagNeighbourOverflow);
            if (ACTION == EPPX_SEARCH_KERNEL_CONTINUITY) f.drho[i] = drho;
        }
        total += cnt;
    }
    for (int off = 32; off > 0; off >>= 1) total += __shfl_down(total, off);
    if (lane_id() == 0 && total) {
        const int slot = ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kPairSlots) * kPairStride;
        atomicAdd(pairs + kPairLast + slot, (unsigned long long)total);
        atomicAdd(pairs + kPairTotal + slot, (unsigned long long)total);
    }
}

// Momentum pass (iMomentum!, main.jl:107-118): fluid particles add the
// pressure and viscosity terms of their fluid and wall neighbours to du;
// other particles take no terms (cells without fluid have no work).
// UM (FAST): one H and one Mass (w.hinv_u, w.m_u) and pressure coefficient 0;
// the terms leave them out.
template <bool FAST, int K, int VISC, bool UM>
__global__ __launch_bounds__(kCellBlock * kCellWaves) void k3_momentum(Walk3 w, Fields3 f, Physics ph) {
    __shared__ Stage3 stage[kCellWaves];
    const int wave = threadIdx.x / kCellBlock;
    Stage3 &st = stage[wave];
    const int c = stripe_group(blockIdx.x, w.stripe) * kCellWaves + wave;
    if (c >= w.g.n_cells) return;
    const int c0 = w.start[c], c1 = w.start[c + 1];
    const uint32_t fluid_bit = tag_bit(ph.fluid_tag), wall_bit = tag_bit(ph.wall_tag);
    if (c0 >= c1 || !(w.cell_tags[c] & fluid_bit)) return;
    const uint32_t adj_block = adj3(w, fluid_bit);
    if (!adj_block) return;  // a fluid adjacent to nothing has no pairs (and no stored masks)
    for (int i0 = c0; i0 < c1; i0 += kCellBlock) {
        const int i = i0 + lane_id();
        const bool valid = i < c1;
        bool active = valid;
        float4 gi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = gi, qi = gi, acc = gi;
        if (active) {
            gi = w.geo[i];
            active = tag_bit(__float_as_int(gi.w)) == fluid_bit;
        }
        if (active) {
            vi = w.vel[i];
            qi = w.hmvp[i];
            acc = f.dvel[i];
        }
        const float p_i = qi.w;
        const float pr_i = FAST ? p_i * frcp(vi.w * vi.w) : p_i / (vi.w * vi.w);
        const bool check = active && (w.odd_tags || (adj_block & w.tags_present) != w.tags_present);
        if constexpr (UM) {
            // The sums leave out -m_u kfac, and the staged pressure of a fluid neighbour is its
            // p / rho^2
            // every accepted neighbour is fluid or wall unless a third tag is adjacent to the fluid
            const bool fw_only = (adj_block & ~(fluid_bit | wall_bit)) == 0u;
            const Uniform3<K> ut(w.hinv_u, ph);
            const float inv_rho_i = frcp(vi.w);
            float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            cell_walk3_sel<FAST>(w, st, c, adj_block, i, valid, active, StageX{fluid_bit, wall_bit}, gi, qi.x,
                                 adj_block, check,
                           [&](float4 gj, float4 vj, float4 qj, float dx, float dy, float dz, float dr2) {
                               const uint32_t tbj = __float_as_uint(gj.w);
                               const bool classic = tbj == fluid_bit;
                               const bool take = fw_only || classic || tbj == wall_bit;
                               momentum_fast_um3<K, VISC>(ph, sum, classic, take, ut, vi, p_i, pr_i, inv_rho_i, vj,
                                                          qj, dx, dy, dz, dr2);
                           });
            const float km = -(w.m_u * ut.kfac);
            acc.x = fmaf(km, sum.x, acc.x);
            acc.y = fmaf(km, sum.y, acc.y);
            acc.z = fmaf(km, sum.z, acc.z);
        } else {
            cell_walk3_sel<FAST>(w, st, c, adj_block, i, valid, active, StageX{}, gi, qi.x, adj_block, check,
                           [&](float4 gj, float4 vj, float4 qj, float dx, float dy, float dz, float dr2) {
                               const uint32_t tbj = __float_as_uint(gj.w);
                               const bool classic = tbj == fluid_bit;
                               if (!classic && tbj != wall_bit) return;
                               const float hinv = w.hinv_u > 0.0f ? w.hinv_u : invharmonic(qi.x, qj.x);
                               momentum3<FAST, K, VISC>(ph, acc, classic, hinv, vi, p_i, pr_i, vj, qj, dx, dy, dz,
                                                        dr2);
                           });
        }
        if (active) f.dvel[i] = acc;
    }
}

// Neighbour lists of the last search as a CSR by slot (row_start) and / or
// in the reference's per-particle neighbour columns (nCount, nIndex, nRVec
// with 3 floats per pair, nR, nW, nDW, nHInv; NeighbourSearch.jl:215-220,
// Library/Kernel.jl:123-126).
__global__ __launch_bounds__(kBlock3) void k3_export(Walk3 w, Fields3 f, const int64_t *row_start, int32_t *js,
                                                      const int32_t *ncount, NeighbourExport ex) {
    const int i = blockIdx.x * kBlock3 + threadIdx.x;
    if (i >= w.n_alive[0]) return;
    const int row = f.id[i];
    const float4 gi = w.geo[i];
    const float hi = w.hmvp[i].x;
    const int ti = __float_as_int(gi.w);
    const uint32_t tbi = tag_bit(ti), adj_i = adj3(w, tbi);
    const uint32_t fluid_bit = tag_bit(ex.fluid_tag), wall_bit = tag_bit(ex.wall_tag);
    const size_t cap = (size_t)ex.n_capacity;
    const int64_t o = row_start ? row_start[i] : 0;
    const int limit = js ? ncount[i] : 0;
    int k = 0;
    walk3<true, true>(w, i, gi, hi, adj_i, [&](int j, float4 gj, float dx, float dy, float dz, float dr2) {
        if (js && k < limit) js[o + k] = f.id[j];
        if (k < ex.M) {
            if (ex.nIndex) ex.nIndex[row + cap * k] = f.id[j] + 1;
            if (ex.nRVec) {
                ex.nRVec[row + cap * (3 * k)] = dx;
                ex.nRVec[row + cap * (3 * k + 1)] = dy;
                ex.nRVec[row + cap * (3 * k + 2)] = dz;
            }
            const float r = sqrtf(dr2);
            if (ex.nR) ex.nR[row + cap * k] = r;
            const uint32_t tbj = tag_bit(__float_as_int(gj.w));
            const bool kernel_pair = (tbi == fluid_bit && (tbj == fluid_bit || tbj == wall_bit)) ||
                                     (tbi == wall_bit && tbj == fluid_bit);
            if (ex.with_kernel && kernel_pair && tbi) {
                const float hinv = invharmonic(hi, ld_slot(w.hmvp, j).x);
                if (ex.nHInv) ex.nHInv[row + cap * k] = hinv;
                if (ex.nW) ex.nW[row + cap * k] = value3_exact(ex.kernel, r, hinv);
                if (ex.nDW)
                    ex.nDW[row + cap * k] = ex.kernel == EPPX_KERNEL_CUBIC ? grad3_exact<EPPX_KERNEL_CUBIC>(r, hinv)
                                                                         : grad3_exact<EPPX_KERNEL_WENDLAND_C2>(r, hinv);
            }
        }
        ++k;
    });
    if (ex.nCount) ex.nCount[row] = k;
}

// ------------------------------------------------------------- cell list
// device_insertParticlesIntoCells! (NeighbourSearch.jl:12-47) in 3D: the cell
// of each alive particle (cached for immovable ones), death outside the
// domain (inclusive bounds); key = 0-based cell, n_cells for dead particles.
__global__ __launch_bounds__(kBlock3) void k3_key_count(Fields3 f, int32_t n, Grid3 g, int32_t *counts,
                                                         int32_t *key, int32_t *tmpslot, uint32_t *flags,
                                                         int32_t max_per_cell) {
    const int32_t s = blockIdx.x * kBlock3 + threadIdx.x;
    const bool valid = s < n;
    int32_t k = g.n_cells;
    if (valid && f.alive[s] == 1) {
        int32_t c = f.cell[s];
        if (c < 0 || c > g.n_cells) c = -1;  // a cached cell outside this grid is recomputed
        if (!(f.mov[s] == 0 && c > 0)) {
            const float4 x = f.pos[s];
            if (g.first_x <= x.x && x.x <= g.last_x && g.first_y <= x.y && x.y <= g.last_y && g.first_z <= x.z &&
                x.z <= g.last_z) {
                int32_t ci = (int32_t)floorf((x.x - g.first_x) * g.gap_x_inv) + 1;
                ci = max(1, min(g.n_x, ci));
                int32_t cj = (int32_t)floorf((x.y - g.first_y) * g.gap_y_inv) + 1;
                cj = max(1, min(g.n_y, cj));
                int32_t ck = (int32_t)floorf((x.z - g.first_z) * g.gap_z_inv) + 1;
                ck = max(1, min(g.n_z, ck));
                c = ci + g.n_x * ((cj - 1) + g.n_y * (ck - 1));
                f.cell[s] = c;
            } else {
                f.alive[s] = 0;
                f.cell[s] = 0;
                c = 0;
            }
        }
        if (c > 0) k = c - 1;
    }
    if (valid) key[s] = k;
    const int32_t slot = count_key(valid, k, g.n_cells, counts, flags, max_per_cell);
    if (valid) tmpslot[s] = slot;
}

__global__ __launch_bounds__(kBlock3) void k3_fill_members(Fields3 f, int32_t n, int32_t n_cells, const int32_t *key,
                                                            const int32_t *tmpslot, const int32_t *start,
                                                            int32_t *members) {
    const int32_t s = blockIdx.x * kBlock3 + threadIdx.x;
    if (s >= n) return;
    const int32_t k = key[s];
    if (k < n_cells) members[start[k] + tmpslot[s]] = f.id[s];
}

// Final place of each slot: cell start + number of cell members with a
// smaller particle index (the order of a serial insert); dead particles
// follow the alive ones.  Moves every field and writes geo = (x, y, z, tag).
__global__ __launch_bounds__(kBlock3) void k3_permute(Fields3 cur, Fields3 nxt, float4 *geo, int32_t n,
                                                       int32_t n_cells, const int32_t *key, const int32_t *tmpslot,
                                                       const int32_t *start, const int32_t *members,
                                                       int32_t *n_alive, uint32_t *flags) {
    const int32_t s = blockIdx.x * kBlock3 + threadIdx.x;
    if (s == 0) {
        n_alive[0] = start[n_cells];
        n_alive[1] = start[n_cells + 1];
    }
    if (s >= n) return;
    const int32_t k = key[s];
    const int32_t me = cur.id[s];
    const int32_t tmp = tmpslot[s];
    const float4 x = cur.pos[s], v = cur.vel[s], q = cur.hmvp[s], d = cur.dvel[s];
    const float drho = cur.drho[s];
    const int32_t mov = cur.mov[s], cell_in = cur.cell[s], alive_in = cur.alive[s];
    const int32_t b = start[k];
    int32_t dest, alive, cell;
    if (k < n_cells) {
        const int32_t e = start[k + 1];
        if (e - b <= kRankLimit) {
            int32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, p = b;
            for (; p + 4 <= e; p += 4) {
                const int32_t a0 = members[p], a1 = members[p + 1], a2 = members[p + 2], a3 = members[p + 3];
                r0 += a0 < me;
                r1 += a1 < me;
                r2 += a2 < me;
                r3 += a3 < me;
            }
            for (; p < e; ++p) r0 += members[p] < me;
            dest = b + r0 + r1 + r2 + r3;
        } else {
            dest = b + tmp;
            atomicOr(flags, kFlagRankFallback);
        }
        alive = 1;
        cell = k + 1;
    } else {
        dest = b + tmp;
        alive = alive_in == 1 ? 0 : alive_in;
        cell = cell_in;
    }
    nxt.pos[dest] = x;
    geo[dest] = x;
    nxt.vel[dest] = v;
    nxt.hmvp[dest] = q;
    nxt.dvel[dest] = d;
    nxt.drho[dest] = drho;
    nxt.mov[dest] = mov;
    nxt.id[dest] = me;
    nxt.cell[dest] = cell;
    nxt.alive[dest] = alive;
}

// Tag bits of each cell's (alive, sorted) particles.
__global__ __launch_bounds__(kBlock3) void k3_cell_tags(const float4 *geo, int32_t n_cells, const int32_t *start,
                                                         uint32_t *cell_tags) {
    const int32_t c = blockIdx.x * kBlock3 + threadIdx.x;
    if (c >= n_cells) return;
    uint32_t bits = 0u;
    const int32_t e = start[c + 1];
    for (int32_t s = start[c]; s < e; ++s) bits |= tag_bit(__float_as_int(geo[s].w));
    cell_tags[c] = bits;
}

// ---------------------------------------------------------- self kernels
// EPPX_SELF_CONTINUITY_EOS: sContinuity! (main.jl:92-105; Continuity.jl:10-21,
// sVolume! State.jl:10-19) on fluid and wall particles; EPPX_SELF_GRAVITY_MOVE:
// sMomentum! (main.jl:120-126; the 3D sGravity! State.jl:49-63,
// sAccelerateMove! Motion.jl:61-75) on fluid particles.
__global__ __launch_bounds__(kBlock3) void k3_self(Fields3 f, int32_t n, int32_t action, Physics3 P) {
    const int32_t s = blockIdx.x * kBlock3 + threadIdx.x;
    if (s >= n || f.alive[s] != 1) return;
    const Physics &ph = P.p;
    float4 x = f.pos[s];
    const int32_t tag = __float_as_int(x.w);
    if (action == EPPX_SELF_CONTINUITY_EOS) {
        if (tag != ph.fluid_tag && tag != ph.wall_tag) return;
        float4 v = f.vel[s], q = f.hmvp[s];
        v.w += f.drho[s] * ph.dt;
        f.drho[s] = 0.0f;
        q.z = q.y / v.w;
        q.w = eos(ph, v.w);
        f.vel[s] = v;
        f.hmvp[s] = q;
    } else if (action == EPPX_SELF_GRAVITY_MOVE) {
        if (tag != ph.fluid_tag) return;
        float4 v = f.vel[s], d = f.dvel[s];
        d.x += ph.gx;
        d.y += ph.gy;
        d.z += P.gz;
        x.x += (v.x + d.x * ph.dt * 0.5f) * ph.dt;
        v.x += d.x * ph.dt;
        x.y += (v.y + d.y * ph.dt * 0.5f) * ph.dt;
        v.y += d.y * ph.dt;
        x.z += (v.z + d.z * ph.dt * 0.5f) * ph.dt;
        v.z += d.z * ph.dt;
        f.pos[s] = x;
        f.vel[s] = v;
        f.dvel[s] = make_float4(0.0f, 0.0f, 0.0f, d.w);
    }
}

// ------------------------------------------------------- layout conversion
__global__ __launch_bounds__(kBlock3) void k3_from_ref(Fields3 f, Columns c, const int32_t *ip, const float *fp,
                                                        const int32_t *is_alive, const int32_t *cell_index) {
    const int row = blockIdx.x * kBlock3 + threadIdx.x;
    if (row >= c.n_capacity) return;
    const size_t C = (size_t)c.n_capacity;
    auto fcol = [&](int col, int k) { return col >= 0 ? fp[row + C * (col + k)] : 0.0f; };
    auto icol = [&](int col) { return col >= 0 ? ip[row + C * col] : 0; };
    f.pos[row] = make_float4(fcol(c.PositionVec, 0), fcol(c.PositionVec, 1), fcol(c.PositionVec, 2),
                             __int_as_float(icol(c.Tag)));
    f.vel[row] = make_float4(fcol(c.VelocityVec, 0), fcol(c.VelocityVec, 1), fcol(c.VelocityVec, 2),
                             fcol(c.Density, 0));
    f.hmvp[row] = make_float4(fcol(c.H, 0), fcol(c.Mass, 0), fcol(c.Volume, 0), fcol(c.Pressure, 0));
    f.dvel[row] = make_float4(fcol(c.dVelocityVec, 0), fcol(c.dVelocityVec, 1), fcol(c.dVelocityVec, 2), 0.0f);
    f.drho[row] = fcol(c.dDensity, 0);
    f.mov[row] = icol(c.IsMovable);
    f.id[row] = row;
    f.alive[row] = is_alive[row];
    f.cell[row] = cell_index[row];
}

__global__ __launch_bounds__(kBlock3) void k3_to_ref(Fields3 f, int32_t n, Columns c, int32_t *ip, float *fp,
                                                      int32_t *is_alive, int32_t *cell_index) {
    const int s = blockIdx.x * kBlock3 + threadIdx.x;
    if (s >= n) return;
    const size_t C = (size_t)c.n_capacity;
    const int row = f.id[s];
    auto put = [&](int col, int k, float v) {
        if (col >= 0) fp[row + C * (col + k)] = v;
    };
    const float4 x = f.pos[s], v = f.vel[s], q = f.hmvp[s], d = f.dvel[s];
    put(c.PositionVec, 0, x.x);
    put(c.PositionVec, 1, x.y);
    put(c.PositionVec, 2, x.z);
    put(c.VelocityVec, 0, v.x);
    put(c.VelocityVec, 1, v.y);
    put(c.VelocityVec, 2, v.z);
    put(c.dVelocityVec, 0, d.x);
    put(c.dVelocityVec, 1, d.y);
    put(c.dVelocityVec, 2, d.z);
    put(c.Density, 0, v.w);
    put(c.H, 0, q.x);
    put(c.Mass, 0, q.y);
    put(c.Volume, 0, q.z);
    put(c.Pressure, 0, q.w);
    put(c.dDensity, 0, f.drho[s]);
    if (c.Tag >= 0) ip[row + C * c.Tag] = __float_as_int(x.w);
    if (c.IsMovable >= 0) ip[row + C * c.IsMovable] = f.mov[s];
    is_alive[row] = f.alive[s];
    cell_index[row] = f.cell[s];
}

inline dim3 grid3_for(int64_t n) {
    const int64_t g = (n + kBlock3 - 1) / kBlock3;
    return dim3((unsigned)(g > 0 ? g : 1));
}

template <bool FAST, int K>
hipError_t search_as(const Walk3 &w, Fields3 f, int32_t action, const Physics &ph, int32_t M, int32_t *ncount,
                     unsigned long long *pairs, uint32_t *flags, int32_t n_slots, hipStream_t s) {
    const dim3 grid((unsigned)(kXcds * w.stripes_per_xcd * w.stripe));
    const dim3 block(kCellBlock * kCellWaves);
    if (action == EPPX_SEARCH_KERNEL_CONTINUITY) {
        if (FAST && w.hinv_u > 0.0f && w.m_u > 0.0f)
            hipLaunchKernelGGL((k3_search<FAST, EPPX_SEARCH_KERNEL_CONTINUITY, K, FAST>), grid, block, 0, s, w, f, ph,
                               M, ncount, pairs, flags);
        else
            hipLaunchKernelGGL((k3_search<FAST, EPPX_SEARCH_KERNEL_CONTINUITY, K, false>), grid, block, 0, s, w, f, ph,
                               M, ncount, pairs, flags);
    } else {
        hipLaunchKernelGGL((k3_search<false, EPPX_SEARCH_NONE, EPPX_KERNEL_CUBIC, false>), grid, block, 0, s, w, f, ph,
                           M, ncount, pairs, flags);
    }
    return hipGetLastError();
}

template <bool FAST, int K>
hipError_t momentum_as(const Walk3 &w, Fields3 f, const Physics &ph, int32_t n_slots, hipStream_t s) {
    const dim3 grid((unsigned)(kXcds * w.stripes_per_xcd * w.stripe));
    const dim3 block(kCellBlock * kCellWaves);
    const bool um = FAST && w.hinv_u > 0.0f && w.m_u > 0.0f && ph.coefficient == 0.0f;
    if (ph.visc == EPPX_VISC_CLASSIC) {
        if (um) hipLaunchKernelGGL((k3_momentum<FAST, K, EPPX_VISC_CLASSIC, FAST>), grid, block, 0, s, w, f, ph);
        else hipLaunchKernelGGL((k3_momentum<FAST, K, EPPX_VISC_CLASSIC, false>), grid, block, 0, s, w, f, ph);
    } else {
        if (um) hipLaunchKernelGGL((k3_momentum<FAST, K, EPPX_VISC_ARTIFICIAL, FAST>), grid, block, 0, s, w, f, ph);
        else hipLaunchKernelGGL((k3_momentum<FAST, K, EPPX_VISC_ARTIFICIAL, false>), grid, block, 0, s, w, f, ph);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch3_from_ref(Fields3 f, Columns c, const int32_t *ip, const float *fp, const int32_t *is_alive,
                            const int32_t *cell_index, hipStream_t s) {
    hipLaunchKernelGGL(k3_from_ref, grid3_for(c.n_capacity), dim3(kBlock3), 0, s, f, c, ip, fp, is_alive,
                       cell_index);
    return hipGetLastError();
}

hipError_t launch3_to_ref(Fields3 f, int32_t n, Columns c, int32_t *ip, float *fp, int32_t *is_alive,
                          int32_t *cell_index, hipStream_t s) {
    hipLaunchKernelGGL(k3_to_ref, grid3_for(n), dim3(kBlock3), 0, s, f, n, c, ip, fp, is_alive, cell_index);
    return hipGetLastError();
}

hipError_t launch3_key_count(Fields3 f, int32_t n, Grid3 g, int32_t *counts, int32_t *key, int32_t *tmpslot,
                             uint32_t *flags, int32_t max_per_cell, hipStream_t s) {
    hipLaunchKernelGGL(k3_key_count, grid3_for(n), dim3(kBlock3), 0, s, f, n, g, counts, key, tmpslot, flags,
                       max_per_cell);
    return hipGetLastError();
}

hipError_t launch3_fill_members(Fields3 f, int32_t n, int32_t n_cells, const int32_t *key, const int32_t *tmpslot,
                                const int32_t *start, int32_t *members, hipStream_t s) {
    hipLaunchKernelGGL(k3_fill_members, grid3_for(n), dim3(kBlock3), 0, s, f, n, n_cells, key, tmpslot, start,
                       members);
    return hipGetLastError();
}

hipError_t launch3_permute(Fields3 cur, Fields3 nxt, float4 *geo, int32_t n, int32_t n_cells, const int32_t *key,
                           const int32_t *tmpslot, const int32_t *start, const int32_t *members, int32_t *n_alive,
                           uint32_t *flags, hipStream_t s) {
    hipLaunchKernelGGL(k3_permute, grid3_for(n), dim3(kBlock3), 0, s, cur, nxt, geo, n, n_cells, key, tmpslot, start,
                       members, n_alive, flags);
    return hipGetLastError();
}

hipError_t launch3_cell_tags(const float4 *geo, int32_t n_cells, const int32_t *start, uint32_t *cell_tags,
                             hipStream_t s) {
    hipLaunchKernelGGL(k3_cell_tags, grid3_for(n_cells), dim3(kBlock3), 0, s, geo, n_cells, start, cell_tags);
    return hipGetLastError();
}

hipError_t launch3_search(const Walk3 &w, Fields3 f, int32_t action, const Physics &ph, int32_t fast, int32_t M,
                          int32_t *ncount, unsigned long long *pairs, uint32_t *flags, int32_t n_slots,
                          hipStream_t s) {
    const bool cubic = ph.kernel == EPPX_KERNEL_CUBIC;
    if (fast)
        return cubic ? search_as<true, EPPX_KERNEL_CUBIC>(w, f, action, ph, M, ncount, pairs, flags, n_slots, s)
                     : search_as<true, EPPX_KERNEL_WENDLAND_C2>(w, f, action, ph, M, ncount, pairs, flags, n_slots, s);
    return cubic ? search_as<false, EPPX_KERNEL_CUBIC>(w, f, action, ph, M, ncount, pairs, flags, n_slots, s)
                 : search_as<false, EPPX_KERNEL_WENDLAND_C2>(w, f, action, ph, M, ncount, pairs, flags, n_slots, s);
}

hipError_t launch3_momentum(const Walk3 &w, Fields3 f, const Physics &ph, int32_t fast, int32_t n_slots,
                            hipStream_t s) {
    const bool cubic = ph.kernel == EPPX_KERNEL_CUBIC;
    if (fast)
        return cubic ? momentum_as<true, EPPX_KERNEL_CUBIC>(w, f, ph, n_slots, s)
                     : momentum_as<true, EPPX_KERNEL_WENDLAND_C2>(w, f, ph, n_slots, s);
    return cubic ? momentum_as<false, EPPX_KERNEL_CUBIC>(w, f, ph, n_slots, s)
                 : momentum_as<false, EPPX_KERNEL_WENDLAND_C2>(w, f, ph, n_slots, s);
}

hipError_t launch3_self(Fields3 f, int32_t n, int32_t action, const Physics3 &ph, hipStream_t s) {
    hipLaunchKernelGGL(k3_self, grid3_for(n), dim3(kBlock3), 0, s, f, n, action, ph);
    return hipGetLastError();
}

hipError_t launch3_export(const Walk3 &w, Fields3 f, int32_t n_slots, const int64_t *row_start, int32_t *js,
                          const int32_t *ncount, NeighbourExport3 ex, hipStream_t s) {
    hipLaunchKernelGGL(k3_export, grid3_for(n_slots), dim3(kBlock3), 0, s, w, f, row_start, js, ncount, ex);
    return hipGetLastError();
}

}  // namespace eppx

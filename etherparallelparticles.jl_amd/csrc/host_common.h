// Host-side state and helpers shared by the C-ABI translation units of
// libeppx (api.hip: the 2D path and the batch; api3d.hip: the 3D path).
#pragma once

#include <math.h>
#include <string.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "eppx_internal.h"

struct PendingTiming {
    const char *name;
    hipEvent_t start, stop;
};

struct eppx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    int32_t math_mode = EPPX_MATH_EXACT;
    bool timing = false;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> timed;  // name -> (sum us, launches)
    uint32_t *d_flags = nullptr;
    int32_t children = 0;   // live particle / neighbour systems
    bool closing = false;   // destroy requested while children were alive
};

namespace eppx {
namespace host {

inline int32_t set_error(eppx_ctx *ctx, int32_t code, const std::string &msg) {
    if (ctx) ctx->error = msg;
    return code;
}


// Frees a context once it was asked to close and no system refers to it.
inline void release_ctx(eppx_ctx *ctx) {
    if (!ctx->closing || ctx->children > 0) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->d_flags) (void)hipFree(ctx->d_flags);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

inline int32_t hip_fail(eppx_ctx *ctx, hipError_t e, const char *what) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    return set_error(ctx, e == hipErrorOutOfMemory ? EPPX_E_NOMEM : EPPX_E_HIP, buf);
}

#define HIP_OK(ctx, expr)                                         \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr);  \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void **)p, count * sizeof(T));
}

template <class T>
inline void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Sum of the pair-counter slots of region kPairLast or kPairTotal of the
// counter buffer d (waits for the stream).
inline hipError_t read_pair_slots(eppx_ctx *ctx, const unsigned long long *d, int region, unsigned long long *out) {
    std::vector<unsigned long long> h((size_t)kPairSlots * kPairStride);
    hipError_t e = hipMemcpyAsync(h.data(), d + region, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    unsigned long long sum = 0;
    for (int k = 0; k < kPairSlots; ++k) sum += h[(size_t)k * kPairStride];
    *out = sum;
    return e;
}

// Event-bracketed launch when the context times kernels.
template <class F>
inline hipError_t timed(eppx_ctx *ctx, const char *name, F &&launch) {
    if (!ctx->timing) return launch();
    auto take = [&]() {
        hipEvent_t e = nullptr;
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else {
            (void)hipEventCreate(&e);
        }
        return e;
    };
    PendingTiming t{name, take(), take()};
    (void)hipEventRecord(t.start, ctx->stream);
    hipError_t e = launch();
    (void)hipEventRecord(t.stop, ctx->stream);
    ctx->pending.push_back(t);
    return e;
}

inline void resolve_timing(eppx_ctx *ctx) {
    for (auto &t : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto &slot = ctx->timed[t.name];
            slot.first += 1000.0 * ms;
            slot.second += 1;
        }
        ctx->event_pool.push_back(t.start);
        ctx->event_pool.push_back(t.stop);
    }
    ctx->pending.clear();
}

inline int32_t read_flags(eppx_ctx *ctx) {
    uint32_t flags = 0;
    HIP_OK(ctx, hipMemcpyAsync(&flags, ctx->d_flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->timing) resolve_timing(ctx);
    if (!flags) return EPPX_OK;
    HIP_OK(ctx, hipMemsetAsync(ctx->d_flags, 0, sizeof(uint32_t), ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    std::string msg = "capacity exceeded:";
    if (flags & kFlagCellOverflow) msg += " a cell holds more than max_neighbour_number particles;";
    if (flags & kFlagNeighbourOverflow) msg += " a particle has more than max_neighbour_number neighbours;";
    if (flags & kFlagRankFallback) msg += " a cell holds more than 1024 particles (in-cell order unspecified);";
    return set_error(ctx, EPPX_E_CAPACITY, msg);
}

// Bounds of symmetrySearchCriterion for alive particles whose H lies in
// [hmin, hmax] (valid: every such H is a positive finite number).
inline void symmetry_bounds(bool valid, float hmin_f, float hmax_f, float *rmin2, float *rmax2) {
    const double hmin = hmin_f, hmax = hmax_f;
    if (!valid || !(hmin >= 1e-30) || !(hmax <= 1e30) || !(hmin <= hmax)) {
        *rmin2 = -1.0f;
        *rmax2 = INFINITY;
    } else if (hmin == hmax) {
        // One H for all: the criterion compares dr2 with one threshold,
        // evaluated here with the device's operations (Mean.harmonic, then
        // *2 and squared; IEEE f32, no contraction).
        const float h = hmin_f;
        const float h2 = 2.0f * h * h / (h + h) * 2.0f;
        const float t = h2 * h2;
        *rmin2 = std::isfinite(t) ? nextafterf(t, INFINITY) : -1.0f;
        *rmax2 = std::isfinite(t) ? -INFINITY : INFINITY;
    } else {
        const double eps = 1.0 / 1048576.0;  // 2^-20, well above the criterion's rounding
        *rmin2 = nextafterf((float)(4.0 * hmin * hmin * (1.0 - eps)), 0.0f);
        *rmax2 = nextafterf((float)(4.0 * hmax * hmax * (1.0 + eps)), INFINITY);
    }
}

// The pair hinv (H + H) / (2 H H) when every alive particle has one valid H
// (the FAST pair terms then hoist it), else 0.
inline float uniform_hinv(bool valid, float hmin, float hmax) {
    if (!valid || !(hmin == hmax) || !(hmin >= 1e-30f) || !(hmin <= 1e30f)) return 0.0f;
    const float h = hmin;
    const float hinv = (h + h) / (2.0f * h * h);
    return std::isfinite(hinv) && hinv > 0.0f ? hinv : 0.0f;
}

// The one Mass of every alive particle (the FAST terms then hoist it), else 0.
inline float uniform_mass(bool valid, float mmin, float mmax) {
    if (!valid || !(mmin == mmax) || !(mmin > 0.0f) || !std::isfinite(mmin)) return 0.0f;
    return mmin;
}

}  // namespace host
}  // namespace eppx

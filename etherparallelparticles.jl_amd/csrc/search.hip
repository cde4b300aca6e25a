// Instantiations of the search pair kernel (k_search, neighbour search +
// kernel + continuity) for one math mode and one smoothing kernel: the build
// compiles this file once per (EPPX_TU_FAST, EPPX_TU_KERNEL).
#include "pair_device.h"

#if !defined(EPPX_TU_FAST) || !defined(EPPX_TU_KERNEL)
#error "compile with -DEPPX_TU_FAST=0|1 -DEPPX_TU_KERNEL=0|1"
#endif

namespace eppx {
namespace {

template <bool FAST, int K, int BLOCK, int STAGE>
void launch_kb(const Walk &w, Fields f, const Physics &ph, const PairLaunch &pl, hipStream_t s) {
    constexpr int KCONT = EPPX_SEARCH_KERNEL_CONTINUITY;
    const dim3 g = grid_for(pl.n, BLOCK);
    if (pl.eos)  // a particle system's step: the next step's continuity in the epilogue
        hipLaunchKernelGGL((k_search<FAST, KCONT, K, false, BLOCK, 0, true>), g, dim3(BLOCK), 0, s, w, f, ph, pl.M,
                           pl.ncount, pl.pairs_last, pl.pairs_total, pl.flags, 0, *pl.eos);
    else if (w.cases)
        hipLaunchKernelGGL((k_search<FAST, KCONT, K, true, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s, w, f, ph, pl.M,
                           pl.ncount, pl.pairs_last, pl.pairs_total, pl.flags, pl.which, EosOut{});
    else
        hipLaunchKernelGGL((k_search<FAST, KCONT, K, false, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s, w, f, ph,
                           pl.M, pl.ncount, pl.pairs_last, pl.pairs_total, pl.flags, pl.which, EosOut{});
}

}  // namespace

// FAST kernels come in every tuned shape (256 or 512 threads; (u, v) staged
// in LDS or gathered; the fused-continuity variant gathered only); EXACT ones
// (the parity mode) in 256 threads, gathered.
template <bool FAST, int K>
hipError_t launch_search_pair(const Walk &w, Fields f, const Physics &ph, const PairLaunch &pl, hipStream_t s) {
    if (!FAST) launch_kb<false, K, 256, 0>(w, f, ph, pl, s);
    else if (pl.block == 512 && pl.stage && !pl.eos) launch_kb<FAST, K, 512, 2>(w, f, ph, pl, s);
    else if (pl.block == 512) launch_kb<FAST, K, 512, 0>(w, f, ph, pl, s);
    else if (pl.stage && !pl.eos) launch_kb<FAST, K, 256, 2>(w, f, ph, pl, s);
    else launch_kb<FAST, K, 256, 0>(w, f, ph, pl, s);
    return hipGetLastError();
}

template hipError_t launch_search_pair<EPPX_TU_FAST != 0, EPPX_TU_KERNEL>(const Walk &, Fields, const Physics &,
                                                                         const PairLaunch &, hipStream_t);

}  // namespace eppx

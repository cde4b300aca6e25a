// Instantiations of the momentum pair kernel (k_momentum: pressure +
// viscosity, with the move and cell keys fused for a particle system's step)
// for one math mode and one smoothing kernel: the build compiles this file
// once per (EPPX_TU_FAST, EPPX_TU_KERNEL).
#include "pair_device.h"

#if !defined(EPPX_TU_FAST) || !defined(EPPX_TU_KERNEL)
#error "compile with -DEPPX_TU_FAST=0|1 -DEPPX_TU_KERNEL=0|1"
#endif

namespace eppx {
namespace {

template <bool FAST, int K, int VISC, int BLOCK, int STAGE>
void launch_kb(const Walk &w, Fields f, const Physics &ph, int32_t n, const MoveKey *mk, hipStream_t s) {
    const dim3 g = grid_for(n, BLOCK);
    if (mk)  // a particle system's step: gravity + move and the coming search's cell keys in the epilogue
        hipLaunchKernelGGL((k_momentum<FAST, K, VISC, false, BLOCK, STAGE, true>), g, dim3(BLOCK), 0, s, w, f, ph,
                           *mk);
    else if (w.cases)  // a batch gathers (u, v) (its kernels need more registers; staging them measured slower)
        hipLaunchKernelGGL((k_momentum<FAST, K, VISC, true, BLOCK, STAGE == 2 ? 0 : STAGE, false>), g, dim3(BLOCK), 0,
                           s, w, f, ph, MoveKey{});
    else
        hipLaunchKernelGGL((k_momentum<FAST, K, VISC, false, BLOCK, STAGE, false>), g, dim3(BLOCK), 0, s, w, f, ph,
                           MoveKey{});
}

template <bool FAST, int K, int VISC>
void launch_shape(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t block, int32_t stage,
                  const MoveKey *mk, hipStream_t s) {
    if (!FAST) launch_kb<false, K, VISC, 256, 0>(w, f, ph, n, mk, s);
    else if (block == 512 && stage && !mk) launch_kb<FAST, K, VISC, 512, 1>(w, f, ph, n, mk, s);
    else if (block == 512) launch_kb<FAST, K, VISC, 512, 2>(w, f, ph, n, mk, s);
    else if (stage && !mk) launch_kb<FAST, K, VISC, 256, 1>(w, f, ph, n, mk, s);
    else launch_kb<FAST, K, VISC, 256, 2>(w, f, ph, n, mk, s);
}

}  // namespace

// FAST kernels come in every tuned shape: 256 or 512 threads, and the
// neighbours' whole records staged in LDS (STAGE 1) or their (u, v) staged
// beside the geometry (STAGE 2: with one H and one Mass the tile then holds
// all a pair term reads, and pass 2 gathers nothing from global memory; the
// fused-move variant takes this one).  EXACT ones (the parity mode) in 256
// threads, gathered.
template <bool FAST, int K>
hipError_t launch_momentum_pair(const Walk &w, Fields f, const Physics &ph, int32_t n, int32_t block, int32_t stage,
                                const MoveKey *mk, hipStream_t s) {
    if (ph.visc == EPPX_VISC_CLASSIC) launch_shape<FAST, K, EPPX_VISC_CLASSIC>(w, f, ph, n, block, stage, mk, s);
    else launch_shape<FAST, K, EPPX_VISC_ARTIFICIAL>(w, f, ph, n, block, stage, mk, s);
    return hipGetLastError();
}

template hipError_t launch_momentum_pair<EPPX_TU_FAST != 0, EPPX_TU_KERNEL>(const Walk &, Fields, const Physics &,
                                                                           int32_t, int32_t, int32_t, const MoveKey *,
                                                                           hipStream_t);

}  // namespace eppx

// The interpreter tile-pass kernels: persistent workgroups, a tile's
// amplitudes in registers, register-layout changes through LDS, each op's
// kind read from the program at run time.  Included by qa_fused.hip
// (complex64 instances) and qa_fused_f64.hip (complex128 instances,
// compiled with their own scheduling options).
#pragma once

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_tile.hpp"

namespace qa {
namespace fuse {

// the prefetching form needs 16 more registers of amplitudes per thread;
// the lean form is held to a register budget of three waves per SIMD
template <typename T, int NT>
__global__ void __launch_bounds__(512) k_tilePass(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    tilePassLoop<T, NT, 1, Interpreter>(amps, a, ops);
}
template <typename T, int NT>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3)))
k_tilePassLean(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    tilePassLoop<T, NT, 0, Interpreter>(amps, a, ops);
}

// [precision FP32/FP64][non-temporal][prefetch]
template <typename T, int NT, int PF>
void launchTile(unsigned grid, unsigned block, size_t lds, void* amps, const TileArgs& a, const TOp* ops) {
    if constexpr (PF) k_tilePass<T, NT><<<grid, block, lds, stream()>>>((T*)amps, a, ops);
    else k_tilePassLean<T, NT><<<grid, block, lds, stream()>>>((T*)amps, a, ops);
}

// the kernel and launch for one precision: nt = non-temporal HBM
// accesses, pf = prefetching form
template <typename T>
const void* tileKernelT(bool nt, bool pf) {
    return nt ? (pf ? (const void*)k_tilePass<T, 1> : (const void*)k_tilePassLean<T, 1>)
              : (pf ? (const void*)k_tilePass<T, 0> : (const void*)k_tilePassLean<T, 0>);
}
template <typename T>
void launchTileT(bool nt, bool pf, unsigned grid, unsigned block, size_t lds, void* amps, const TileArgs& a,
                 const TOp* ops) {
    if (nt) pf ? launchTile<T, 1, 1>(grid, block, lds, amps, a, ops) : launchTile<T, 1, 0>(grid, block, lds, amps, a, ops);
    else pf ? launchTile<T, 0, 1>(grid, block, lds, amps, a, ops) : launchTile<T, 0, 0>(grid, block, lds, amps, a, ops);
}

// complex128 instances (qa_fused_f64.hip)
const void* tileKernelF64(bool nt, bool pf);
void launchTileF64(bool nt, bool pf, unsigned grid, unsigned block, size_t lds, void* amps, const TileArgs& a,
                   const TOp* ops);

}  // namespace fuse
}  // namespace qa

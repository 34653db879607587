// The tile-pass kernels: one workgroup per tile (persistent), the tile's
// amplitudes in registers, register-layout changes through LDS.  Included
// by qa_fused.hip (complex64 instances) and qa_fused_f64.hip (complex128
// instances, compiled with their own scheduling options).
#pragma once

#include "qa_common.hpp"
#include "qa_fused.hpp"
#include "qa_tile.hpp"

namespace qa {
namespace fuse {

// NT: non-temporal HBM loads/stores, used when the contiguous low run of a
// tile row is at least 2^STREAM_LOW_BIT amplitudes.  Persistent workgroups;
// PF = 1: each prefetches its next tile into registers before working on
// the current one, so HBM traffic overlaps the register and LDS work (16
// more registers of amplitudes per thread); PF = 0: load, work, store.
template <typename T, int NT, int PF>
__device__ __forceinline__ void tilePass(T* __restrict__ amps, const TileArgs& a, const TOp* __restrict__ ops) {
    extern __shared__ __align__(16) unsigned char smem[];
    ct<T>* tile = reinterpret_cast<ct<T>*>(smem);
    const uint32_t tid = threadIdx.x;
    RegLayout row, first;
    passLayouts(row, first, a, ops, tid);
    const RegLayout& loadL = a.firstDirect ? first : row;

    // tile t's base is t deposited over the non-local positions; the next
    // tile of this workgroup (t + gridDim.x) is one masked add away
    auto nextBase = [&](uint64_t b) { return ((b | ~a.nonLocalMask) + a.tileStep) & a.nonLocalMask; };
    uint64_t t = blockIdx.x;
    uint64_t base = deposit(t, a.nonLocalPos, a.numNonLocal);
    if constexpr (PF) {
        ct<T> rn[RA];
        if (t < a.numTiles) RegIO<T, NT, 0>::hbmLoad(rn, amps, base | loadL.tg, loadL, a.localPos);
        for (; t < a.numTiles; t += gridDim.x) {
            ct<T> r[RA];
#pragma unroll
            for (int v = 0; v < RA; v++) r[v] = rn[v];
            const uint64_t baseN = nextBase(base);
            if (t + gridDim.x < a.numTiles) RegIO<T, NT, 0>::hbmLoad(rn, amps, baseN | loadL.tg, loadL, a.localPos);
            tileBody<T, NT>(amps, a, ops, base, r, tile, row, first, tid);
            base = baseN;
            // the next tile's first LDS writes must not overtake this tile's reads
            if (a.usesLds) __syncthreads();
        }
    } else {
        for (; t < a.numTiles; t += gridDim.x) {
            ct<T> r[RA];
            RegIO<T, NT, 0>::hbmLoad(r, amps, base | loadL.tg, loadL, a.localPos);
            tileBody<T, NT>(amps, a, ops, base, r, tile, row, first, tid);
            base = nextBase(base);
            if (a.usesLds) __syncthreads();
        }
    }
}

// the prefetching form needs 16 more registers of amplitudes per thread;
// the lean form is held to a register budget of three waves per SIMD
template <typename T, int NT>
__global__ void __launch_bounds__(512) k_tilePass(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    tilePass<T, NT, 1>(amps, a, ops);
}
template <typename T, int NT>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3)))
k_tilePassLean(T* __restrict__ amps, TileArgs a, const TOp* __restrict__ ops) {
    tilePass<T, NT, 0>(amps, a, ops);
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

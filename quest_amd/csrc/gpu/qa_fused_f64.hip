// complex128 instances of the tile-pass kernels (qa_tilepass.hpp), in a
// translation unit of their own so that the build can schedule them for
// instruction-level parallelism (their occupancy is fixed by the LDS tile),
// which measured faster for complex128 and slower for complex64.
#include "qa_tilepass.hpp"

namespace qa {
namespace fuse {

const void* tileKernelF64(bool nt, bool pf) { return tileKernelT<double>(nt, pf); }

void launchTileF64(bool nt, bool pf, unsigned grid, unsigned block, size_t lds, void* amps, const TileArgs& a,
                   const TOp* ops) {
    launchTileT<double>(nt, pf, grid, block, lds, amps, a, ops);
}

}  // namespace fuse
}  // namespace qa

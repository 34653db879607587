// Shared device/host helpers of the gfx950 kernel layer: amplitude vector
// types, index bit algebra, launch shapes, error plumbing.
//
// Index algebra follows quest/src/core/bitwise.hpp:34-254 (insertBits over
// ascending qubit positions, setBits over an unsorted target list, parity of
// masked bits); it is restated here for 64-bit device indices.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "qa_dev.hpp"
#include "quest_amd_gpu.h"

namespace qa {

// ---- launch shapes -----------------------------------------------------
constexpr int BLOCK = 256;   // 4 wave64s per workgroup
constexpr int ITEMS = 4;     // work items per thread, strided by BLOCK

// Workgroups covering `work` items, `items` per thread.  AMD dispatch
// packets count work-items in 32 bits, so a grid of more than 2^30
// work-items folds into a second dimension (2^22 workgroups per row);
// kernels number their workgroups with blockId() and skip items past the
// end (states of 33-34 qubits need the fold).
inline dim3 gridFor(uint64_t work, int items = ITEMS) {
    const uint64_t per = (uint64_t)BLOCK * (uint64_t)items;
    uint64_t g = (work + per - 1) / per;
    if (!g) g = 1;
    const uint64_t maxX = 1ull << 22;
    if (g <= maxX) return dim3((unsigned)g, 1, 1);
    return dim3((unsigned)maxX, (unsigned)((g + maxX - 1) / maxX), 1);
}
__device__ __forceinline__ uint64_t blockId() { return (uint64_t)blockIdx.y * gridDim.x + blockIdx.x; }

// ---- errors ------------------------------------------------------------
void setError(const char* fmt, ...);
hipStream_t stream();
bool isInit();
// state epoch: bumped by every entry point that may change amplitudes
// (QA_REQUIRE_INIT), left alone by the read-only reductions
// (QA_REQUIRE_INIT_RO); equal epochs mean no state has changed in between
void bumpEpoch();
// reduction scratch: device partials and a pinned host landing area
double* devScratch(size_t numDoubles);
void trimDevScratch();  // frees the device partials if they grew past 32 MiB
double* hostScratch(size_t numDoubles);
// the device the library is bound to (qa_init), -1 before
int device();
// takes the device address of the no-allocation outcome-table scratch for
// `device`, which must be current (qa_reduce.hip; called by qa_init)
int initNoAllocScratch(int device);

}  // namespace qa

#define QA_REQUIRE_INIT_RO()                                               \
    do {                                                                   \
        if (!qa::isInit()) {                                               \
            qa::setError("%s: qa_init() has not been called", __func__);   \
            return QA_ERR_NOT_INIT;                                        \
        }                                                                  \
    } while (0)

#define QA_REQUIRE_INIT()     \
    do {                      \
        QA_REQUIRE_INIT_RO(); \
        qa::bumpEpoch();      \
    } while (0)

#define QA_HIP(cmd)                                                                   \
    do {                                                                              \
        hipError_t e_ = (cmd);                                                        \
        if (e_ != hipSuccess) {                                                       \
            qa::setError("%s: %s failed: %s", __func__, #cmd, hipGetErrorString(e_)); \
            return QA_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

#define QA_LAUNCH_CHECK() QA_HIP(hipGetLastError())

#define QA_ARG_CHECK(cond, ...)       \
    do {                              \
        if (!(cond)) {                \
            qa::setError(__VA_ARGS__); \
            return QA_ERR_ARG;        \
        }                             \
    } while (0)

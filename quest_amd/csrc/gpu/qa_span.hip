// Span-pass kernel instances (qa_span.hpp) and their launches.
#include "qa_common.hpp"
#include "qa_span.hpp"

namespace qa {
namespace fuse {

const void* spanKernel(int prec, bool nt) {
    if (prec == QA_FP64) return nt ? (const void*)k_spanPass<double, 1> : (const void*)k_spanPass<double, 0>;
    return nt ? (const void*)k_spanPass<float, 1> : (const void*)k_spanPass<float, 0>;
}

void launchSpanAny(int prec, bool nt, unsigned grid, unsigned block, size_t lds, void* amps, const SpanArgs& a,
                   const SpanOp* ops) {
    if (prec == QA_FP64) {
        if (nt) k_spanPass<double, 1><<<grid, block, lds, stream()>>>((double*)amps, a, ops);
        else k_spanPass<double, 0><<<grid, block, lds, stream()>>>((double*)amps, a, ops);
    } else {
        if (nt) k_spanPass<float, 1><<<grid, block, lds, stream()>>>((float*)amps, a, ops);
        else k_spanPass<float, 0><<<grid, block, lds, stream()>>>((float*)amps, a, ops);
    }
}

}  // namespace fuse
}  // namespace qa

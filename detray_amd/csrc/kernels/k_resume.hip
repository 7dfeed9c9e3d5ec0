// The per-lane and gathered kernels with a step budget per call
// (dtr_propagate_resumable) on the toy's geometry.
#include "kernels.hpp"

namespace dtr {
#define DTR_RESUME(P)                                                                                          \
    template int launch_resumable<P>(const dtr_detector*, const dtr_field*, const DevConfig&,                 \
                                     const dtr_track_batch*, const DevResults&, uint32_t, const uint32_t*, int, \
                                     uint32_t, const PauseView&, hipStream_t);
DTR_RESUME(PolicyFP64)
DTR_RESUME(PolicyMixed)
#undef DTR_RESUME
}  // namespace dtr

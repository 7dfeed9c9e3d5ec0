// The direct navigator's kernels.
#include "kernels.hpp"

namespace dtr {
#define DTR_DIRECT(P)                                                                                           \
    template int launch_direct<P>(const dtr_detector*, const dtr_field*, const DevConfig&, const DevConfig&,   \
                                  const dtr_track_batch*, const DevResults&, const SeqView&, uint32_t,         \
                                  const uint32_t*, hipStream_t);
DTR_DIRECT(PolicyFP64)
DTR_DIRECT(PolicyMixed)
DTR_DIRECT(WithExtendedGeometry<PolicyFP64>)
DTR_DIRECT(WithExtendedGeometry<PolicyMixed>)
#undef DTR_DIRECT
}  // namespace dtr

// Per-lane kernels with the field gradient in the transport Jacobian.
#include "kernels.hpp"

namespace dtr {
#define DTR_LANE(P)                                                                                               \
    template int launch_lane<P>(const dtr_detector*, const dtr_field*, const DevConfig&, const dtr_track_batch*, \
                                const DevResults&, uint32_t, const uint32_t*, int, hipStream_t);
DTR_LANE(WithFieldGradient<PolicyFP64>)
DTR_LANE(WithFieldGradient<PolicyMixed>)
DTR_LANE(WithFieldGradient<WithExtendedGeometry<PolicyFP64>>)
DTR_LANE(WithFieldGradient<WithExtendedGeometry<PolicyMixed>>)
#undef DTR_LANE
}  // namespace dtr

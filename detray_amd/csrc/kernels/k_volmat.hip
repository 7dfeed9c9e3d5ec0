// Kernels of detectors with homogeneous volume material: per-lane and
// gathered navigation, and the direct navigator, with the continuous energy
// loss in the stepper.
#include "kernels.hpp"

namespace dtr {
#define DTR_LANE(P)                                                                                               \
    template int launch_lane<P>(const dtr_detector*, const dtr_field*, const DevConfig&, const dtr_track_batch*, \
                                const DevResults&, uint32_t, const uint32_t*, int, hipStream_t);
DTR_LANE(WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>)
DTR_LANE(WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>)
#undef DTR_LANE
#define DTR_DIRECT(P)                                                                                           \
    template int launch_direct<P>(const dtr_detector*, const dtr_field*, const DevConfig&, const DevConfig&,   \
                                  const dtr_track_batch*, const DevResults&, const SeqView&, uint32_t,         \
                                  const uint32_t*, hipStream_t);
DTR_DIRECT(WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>)
DTR_DIRECT(WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>)
#undef DTR_DIRECT
}  // namespace dtr

// The adjoint's recording forward and backward kernels on the extended
// geometry (wires, annuli, placed cylinders, rods), mixed policy.
#include "kernels.hpp"

namespace dtr {
#define DTR_ADJ(W)                                                                                                 \
    template int launch_adjoint_record<WithExtendedGeometry<PolicyMixed>, W>(                                    \
        const dtr_detector*, const dtr_field*, const DevConfig&, const AdjointIO&, uint32_t, const uint32_t*,      \
        hipStream_t);                                                                                              \
    template int launch_adjoint_backward<WithExtendedGeometry<PolicyMixed>, W>(                                  \
        const dtr_detector*, const dtr_field*, const DevConfig&, const AdjointIO&, uint32_t, const uint32_t*, int, \
        double*, hipStream_t);
DTR_ADJ(1)
DTR_ADJ(2)
#undef DTR_ADJ
}  // namespace dtr

// The batched kernels, FP64 policy, one and two waves per SIMD.
#include "kernels.hpp"

namespace dtr {
template int launch_batched<PolicyFP64, 1>(const dtr_detector*, const dtr_field*, const DevConfig&,
                                        const dtr_track_batch*, const DevResults&, const dtr_launch_config&,
                                        hipStream_t);
template int launch_batched<PolicyFP64, 2>(const dtr_detector*, const dtr_field*, const DevConfig&,
                                        const dtr_track_batch*, const DevResults&, const dtr_launch_config&,
                                        hipStream_t);
}  // namespace dtr

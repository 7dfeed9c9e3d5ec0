// The plain FP64 kernel: toy / telescope geometry and the extended geometry.
#include "kernels.hpp"

namespace dtr {
template int launch_plain<false>(const dtr_detector*, const dtr_field*, const DevConfig&, const dtr_track_batch*,
                                 const DevResults&, hipStream_t);
template int launch_plain<true>(const dtr_detector*, const dtr_field*, const DevConfig&, const dtr_track_batch*,
                                const DevResults&, hipStream_t);
}  // namespace dtr

// Device-resident handles of the C ABI and the launchers of the propagation
// kernels.  The launchers are templates over the precision policy and the
// register budget; each group of instantiations is compiled in a translation
// unit of its own (k_*.hip), so the library builds in parallel, and capi.hip
// (argument checks, dispatch) only declares them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "detray_amd.h"
#include "device/adjoint.hpp"
#include "device/direct_nav.hpp"
#include "device/engine.hpp"

struct dtr_detector {
    dtr::DevDetector dev;
    dtr::FastDetector fast;
    void* blocks[13];  // the eight tables, surface records, bin records, brute-force records, grid tolerances,
                       // material maps
    uint32_t n_volumes;
    uint32_t n_surfaces;
    int device;
    uint32_t n_cu;
    // wire, annulus or placed-cylinder surfaces, material rods, rings with
    // material, grid bins of several surfaces or volume material: the kernels
    // are instantiated with the extended geometry
    bool wires;
    bool volume_material;  // a volume holds homogeneous material
    bool density_effect;   // a material carries density-effect data
};

struct dtr_field {
    dtr::DevField dev;
    float* values;
};

namespace dtr {

inline int to_status(hipError_t e) {
    if (e == hipSuccess) return DTR_SUCCESS;
    if (e == hipErrorOutOfMemory) return DTR_ERROR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return DTR_ERROR_NO_DEVICE;
    return DTR_ERROR_HIP;
}

inline TrackInputs track_inputs(const dtr_track_batch* in) {
    return TrackInputs{in->bound_params ? nullptr : in->free_params, in->initial_cov, in->bound_params,
                       in->start_surface, in->n};
}

// LDS bytes of a batched block: the candidate cache (16 B per slot and lane)
// and the exchange slots and FIFO of each wave (cooperative navigation)
size_t lds_bytes(uint32_t block);

// The plain FP64 kernel (dtr_propagate): one work-item per track, block 256.
template <bool kExt>
int launch_plain(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                 const DevResults& r, hipStream_t stream);

// The batched kernels (dtr_propagate_ex): grid schedule with per-lane or
// pooled navigation, or persistent waves over the work queue L.queue.
template <class P, int W>
int launch_batched(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                   const DevResults& r, const dtr_launch_config& L, hipStream_t stream);

// One work-item per track with per-lane (gather 0) or gathered (1)
// navigation, two waves per SIMD (use_field_gradient).
template <class P>
int launch_lane(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                const DevResults& r, uint32_t block, const uint32_t* order, int gather, hipStream_t stream);

// The per-lane (gather 0) or gathered (1) kernels with a step budget per
// call and the lanes' state in pv (dtr_propagate_resumable), two waves per
// SIMD.
template <class P>
int launch_resumable(const dtr_detector* det, const dtr_field* field, const DevConfig& dc,
                     const dtr_track_batch* in, const DevResults& r, uint32_t block, const uint32_t* order,
                     int gather, uint32_t xcd, const PauseView& pv, hipStream_t stream);

// The direct navigator (dtr_propagate_direct).
template <class P>
int launch_direct(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const DevConfig& dd,
                  const dtr_track_batch* in, const DevResults& r, const SeqView& sq, uint32_t block,
                  const uint32_t* order, hipStream_t stream);

// The adjoint's recording forward pass and backward pass.
template <class P, int W>
int launch_adjoint_record(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const AdjointIO& io,
                          uint32_t block, const uint32_t* order, hipStream_t stream);
template <class P, int W>
int launch_adjoint_backward(const dtr_detector* det, const dtr_field* field, const DevConfig& dc,
                            const AdjointIO& io, uint32_t block, const uint32_t* order, int exact, double* grad,
                            hipStream_t stream);

}  // namespace dtr

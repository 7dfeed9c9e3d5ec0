// The propagation kernels and the definitions of their launchers
// (launch.hpp).  Included only by the k_*.hip translation units, each of
// which instantiates one group of launchers explicitly.
//
// dtr_propagate runs one work-item per track through detray's propagate()
// loop (core/include/detray/propagator/propagator.hpp:262-400) with the actor
// chain pathlimit_aborter, parameter_updater<pointwise_material_interactor>,
// surface_sequencer.
#pragma once

#include <type_traits>

#include "launch.hpp"
#include "device/coop_nav.hpp"
#include "device/propagator.hpp"

namespace dtr {

// kExt: detectors with wires, annuli, placed cylinders, rods, multi-surface
// bins or rings with material.  Two waves per SIMD (256 registers per lane,
// VGPRs and AGPRs together): left to itself the compiler allocates past 256
// registers for this kernel and runs one wave per SIMD, which hides less of
// its memory latency.
template <bool kExt>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) propagate_fp64_kernel(
    DevDetector det, DevField field, DevConfig cfg, const double* __restrict__ fp,
    const double* __restrict__ init_cov, const double* __restrict__ bound0, const uint32_t* __restrict__ start_sf,
    DevResults out, uint32_t n) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    propagate_track<kExt>(det, field, cfg, fp, init_cov, bound0, start_sf, out, n, gid);
}

// LDS of a block: the candidate cache, slot-major ([8][blockDim] paths, then
// [8][blockDim] surface indices and [8][blockDim] link|status words), then
// the exchange slots and FIFO of each wave for the cooperative navigation.
// With kState the lanes' track states ([blockDim][kStateStride] floats,
// WithLdsState policies) follow the cache.
template <bool kState = false>
__device__ inline CacheView lds_cache(unsigned char* smem) {
    const int nb = static_cast<int>(blockDim.x);
    double* path = reinterpret_cast<double*>(smem);
    uint32_t* sf = reinterpret_cast<uint32_t*>(path + kCache * nb);
    uint32_t* ls = sf + kCache * nb;
    const int t = static_cast<int>(threadIdx.x);
    CacheView c{path + t, sf + t, ls + t, nb};
    if constexpr (kState) c.state = reinterpret_cast<float*>(ls + kCache * nb) + t * kStateStride;
    return c;
}

// LDS bytes of a one-work-item-per-track block of policy P: the cache and,
// for WithLdsState policies, the track-state slots
template <class P>
inline size_t lane_lds_bytes(uint32_t block) {
    return static_cast<size_t>(block) * (kCache * 16u + (P::kLdsState ? kStateStride * 4u : 0u));
}

// The policy the per-lane grid kernel runs for P: FP32 stage arithmetic
// keeps its Jacobian and covariance in LDS
template <class P>
using LanePolicy = std::conditional_t<(sizeof(typename P::S) == sizeof(float)) && !P::kLdsState, WithLdsState<P>, P>;

__device__ inline CoopShared* lds_shared(unsigned char* smem) {
    CoopShared* all = reinterpret_cast<CoopShared*>(smem + static_cast<size_t>(blockDim.x) * kCache * 16u);
    return all + threadIdx.x / 64u;
}

// The wave interface of the cooperative navigation on CDNA (64 lanes).
struct WaveHip {
    static __device__ uint32_t lane() { return __lane_id(); }
    template <class T>
    static __device__ T shfl(T v, int src) {
        return __shfl(v, src);
    }
    static __device__ uint64_t ballot(bool p) { return __ballot(p); }
    static __device__ bool any(bool p) { return __any(p) != 0; }
    static __device__ void sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    static __device__ uint32_t atomic_add(uint32_t* p, uint32_t v) { return atomicAdd(p, v); }
};

// The batched kernels are compiled for W resident waves per SIMD: the
// register budget per lane is 512 / W (W = 2 trades a few spills for twice
// the waves that hide memory latency).
#define DTR_BATCHED_KERNEL __global__ void __launch_bounds__(256) \
    __attribute__((amdgpu_waves_per_eu(W, W)))

// The block of the track order that hardware block `b` of `nb` runs.  The
// hardware deals blocks to the 8 XCDs in turn; with chunks of G > 1, XCD x
// runs blocks x G .. x G + G - 1 of each round of 8 G consecutive blocks, so
// the waves resident on one XCD (and sharing its L2) work on neighbouring
// tracks of the (sorted) order.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t G) {
    if (G <= 1u) return b;
    const uint32_t g = 8u;
    // chunks of G consecutive blocks per XCD: hardware block b = 8k + x;
    // the blocks past the last whole round of g chunks keep hardware order
    const uint32_t full = (nb / (g * G)) * (g * G);
    if (b >= full) return b;
    const uint32_t x = b % g, k = b / g;
    return (k / G) * (g * G) + x * G + (k % G);
}

// One track per lane with the volume searches (local navigations) of the
// wave gathered: a lane whose navigation asks for a search waits while the
// other lanes go on with their own propagate() loops, and once every lane
// still running waits, they search together.  Between two such points the
// lanes run their actor halves, navigation updates and steps side by side,
// and the searches, the most expensive and most divergent part of the loop,
// run with the whole wave active.  Each lane runs exactly the operations of
// propagate() in its own order; only the interleaving across lanes changes,
// so the results are those of the per-lane kernel.
//
// kPause: a step budget per call (propagate_resumable): a lane pauses where
// its actor half would begin once it has taken pv.max_steps steps in this
// call, and a resuming call starts the paused lanes there.
template <class P, bool kPause = false>
__device__ __attribute__((always_inline)) inline void propagate_gathered(
    const FastDetector& det, const DevField& field, const DevConfig& cfg, const TrackInputs& in,
    const DevResults& out, bool has_track, uint32_t track, CacheView c, const PauseView& pv = PauseView{}) {
    enum : uint8_t { kStart, kActor, kNav1, kNav2, kDone };
    Lane<P> lane;
    float os = 0.f;
    uint8_t phase = kDone;
    bool waiting = false;
    uint32_t s0 = 0u;
    if constexpr (kPause) {
        if (has_track && pv.resume) {
            if (pv.wv(kPwState, track) == kLanePaused) {
                lane.load(pv, c, track);
                s0 = lane.tk.n_steps;
                phase = kActor;
            }
            has_track = false;
        }
    }
    if (has_track) {
        lane.start_begin(det, field, cfg, in, c, track, os);
        phase = kStart;
        waiting = true;  // the navigator's initialisation is a volume search
    }
    // a finished track: its results, and its state as done
    auto done = [&]() {
        lane.finish(out, in.n, cfg);
        if constexpr (kPause) pv.wv(kPwState, lane.gid) = kLaneDone;
        phase = kDone;
    };
    for (;;) {
        const uint64_t running = __ballot(phase != kDone);
        if (running == 0ull) break;
        if (__ballot(waiting) == running) {
            // every running lane waits: the searches, then each lane's
            // navigation cascade goes on (it may ask for another search);
            // lanes that are done (or have no track) sit this out
            if (!waiting) continue;
            // the one place the searches happen: inline
            lane.apply_local_navigation(
                local_navigation<P>(det, cfg, lane.tk.pos, lane.tk.dir, lane.tk.ext_tol, lane.tk.volume, os, c));
            if (phase == kStart) {
                lane.start_end();
                waiting = false;
                phase = kActor;
                continue;
            }
            waiting = lane.nav_resume(det, cfg, c, os);
            if (waiting) continue;
            // the navigation update is complete
            const bool end = lane.nav_end();
            if (phase == kNav2) {
                phase = kActor;
                continue;
            }
            if (end) {
                done();
                continue;
            }
            lane.step_half(det, field, cfg, c);
            phase = kNav2;
            waiting = lane.nav_begin(det, cfg, c, os);
            if (!waiting) {
                lane.nav_end();
                phase = kActor;
            }
            continue;
        }
        if (phase == kActor && !waiting) {
            if constexpr (kPause) {
                if (lane.tk.n_steps - s0 >= pv.max_steps) {
                    lane.pause(pv, c, out);
                    phase = kDone;
                    continue;
                }
            }
            // actor half, navigation update, step half, navigation update
            // (Lane::advance), up to the first request for a search
            if (lane.actor_half(det, field, cfg, in, out, c)) {
                done();
                continue;
            }
            phase = kNav1;
            waiting = lane.nav_begin(det, cfg, c, os);
            if (waiting) continue;
            if (lane.nav_end()) {
                done();
                continue;
            }
            lane.step_half(det, field, cfg, c);
            phase = kNav2;
            waiting = lane.nav_begin(det, cfg, c, os);
            if (waiting) continue;
            lane.nav_end();
            phase = kActor;
        }
    }
}

template <class P, int W>
DTR_BATCHED_KERNEL propagate_grid_kernel(FastDetector det, DevField field, DevConfig cfg, TrackInputs in,
                                         DevResults out, const uint32_t* __restrict__ order, uint32_t xcd) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(xcd_block(blockIdx.x, gridDim.x, xcd)) * blockDim.x + threadIdx.x;
    if (k >= in.n) return;
    const uint32_t track = order ? order[k] : static_cast<uint32_t>(k);
    propagate_fast<P>(det, field, cfg, in, out, track, lds_cache<P::kLdsState>(smem));
}

// One track per lane, the wave's volume searches gathered.  Every lane of
// the wave takes part in the gathering, also the lanes past the end of the
// batch (without a track).
template <class P, int W>
DTR_BATCHED_KERNEL propagate_gather_kernel(FastDetector det, DevField field, DevConfig cfg, TrackInputs in,
                                           DevResults out, const uint32_t* __restrict__ order, uint32_t xcd) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(xcd_block(blockIdx.x, gridDim.x, xcd)) * blockDim.x + threadIdx.x;
    const bool has = k < in.n;
    const uint32_t track = has ? (order ? order[k] : static_cast<uint32_t>(k)) : 0u;
    propagate_gathered<P>(det, field, cfg, in, out, has, track, lds_cache<P::kLdsState>(smem));
}

// The per-lane (G = 0) and gathered (G = 1) kernels with a step budget per
// call (dtr_propagate_resumable).
template <class P, int W, int G>
DTR_BATCHED_KERNEL propagate_resumable_kernel(FastDetector det, DevField field, DevConfig cfg, TrackInputs in,
                                              DevResults out, const uint32_t* __restrict__ order, uint32_t xcd,
                                              PauseView pv) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(xcd_block(blockIdx.x, gridDim.x, xcd)) * blockDim.x + threadIdx.x;
    if constexpr (G == 0) {
        if (k >= in.n) return;
        const uint32_t track = order ? order[k] : static_cast<uint32_t>(k);
        propagate_resumable<P>(det, field, cfg, in, out, track, lds_cache<P::kLdsState>(smem), pv);
    } else {
        const bool has = k < in.n;
        const uint32_t track = has ? (order ? order[k] : static_cast<uint32_t>(k)) : 0u;
        propagate_gathered<P, true>(det, field, cfg, in, out, has, track, lds_cache<P::kLdsState>(smem), pv);
    }
}

// One track per lane, navigation searches pooled across the wave.  Every
// lane of the last wave takes part in the pooled searches, also the lanes
// past the end of the batch.
template <class P, int W>
DTR_BATCHED_KERNEL propagate_coop_kernel(FastDetector det, DevField field, DevConfig cfg, TrackInputs in,
                                         DevResults out, const uint32_t* __restrict__ order) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool has = k < in.n;
    const uint32_t track = has ? (order ? order[k] : static_cast<uint32_t>(k)) : 0u;
    coop_run_grid<P, WaveHip>(det, field, cfg, in, out, has, track, lds_cache(smem), lds_shared(smem));
}

// Persistent waves over a device work queue with pooled navigation searches:
// a lane whose track has ended takes the next queued track, so a wave stays
// full while its tracks have very different step counts.  A wave claims
// `claim` entries at a time with one atomic on *queue; it ends when the queue
// is drained and its lanes are done.
template <class P, int W>
DTR_BATCHED_KERNEL propagate_coop_persistent_kernel(FastDetector det, DevField field, DevConfig cfg,
                                                    TrackInputs in, DevResults out,
                                                    const uint32_t* __restrict__ order, uint32_t* __restrict__ queue,
                                                    uint32_t claim, uint32_t min_idle) {
    extern __shared__ __align__(16) unsigned char smem[];
    coop_run_persistent<P, WaveHip>(det, field, cfg, in, out, order, queue, claim, min_idle, lds_cache(smem),
                                    lds_shared(smem));
}

// resets the persistent work queue on the stream (a kernel, not a memset, so
// the launch can be graph-captured); one copy per translation unit
static __global__ void queue_reset_kernel(uint32_t* queue) { *queue = 0u; }

// Propagation with the direct navigator: one work-item per track.
template <class P, int W>
DTR_BATCHED_KERNEL propagate_direct_kernel(FastDetector det, DevField field, DevConfig cfg, DevConfig dcfg,
                                           TrackInputs in, DevResults out, SeqView sq,
                                           const uint32_t* __restrict__ order) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= in.n) return;
    const uint32_t track = order ? order[k] : static_cast<uint32_t>(k);
    propagate_direct<P>(det, field, cfg, dcfg, in, out, sq, track, lds_cache(smem));
}

// Recording forward pass of the adjoint: one work-item per track (work-item
// k runs track order[k] and owns tape column k), lane navigation with the
// candidate cache in LDS.
template <class P, int W>
DTR_BATCHED_KERNEL adjoint_record_kernel(FastDetector det, DevField field, DevConfig cfg, AdjointIO io,
                                         const uint32_t* __restrict__ order) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= io.n) return;
    const uint32_t track = order ? order[k] : static_cast<uint32_t>(k);
    adjoint_record<P>(det, field, cfg, io, track, static_cast<uint32_t>(k), lds_cache(smem));
}

// Backward pass: one work-item per track, in the forward pass's order.
template <class P, int W>
DTR_BATCHED_KERNEL adjoint_backward_kernel(FastDetector det, DevField field, DevConfig cfg, AdjointIO io,
                                           const uint32_t* __restrict__ order, int exact,
                                           double* __restrict__ grad) {
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= io.n) return;
    const uint32_t track = order ? order[k] : static_cast<uint32_t>(k);
    adjoint_backward<P>(det, field, cfg, io, exact != 0, track, static_cast<uint32_t>(k), grad);
}

inline uint32_t grid_of(uint64_t n, uint32_t block) {
    return static_cast<uint32_t>((n + block - 1) / block);
}

template <bool kExt>
int launch_plain(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                 const DevResults& r, hipStream_t stream) {
    const uint32_t block = 256;
    hipLaunchKernelGGL(propagate_fp64_kernel<kExt>, dim3(grid_of(in->n, block)), dim3(block), 0, stream, det->dev,
                       field->dev, dc, in->free_params, in->initial_cov, in->bound_params, in->start_surface, r,
                       in->n);
    return to_status(hipGetLastError());
}

template <class P, int W>
int launch_batched(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                   const DevResults& r, const dtr_launch_config& L, hipStream_t stream) {
    const TrackInputs ti = track_inputs(in);
    const uint32_t block = L.block_size;
    const size_t shmem = lds_bytes(block);
    if (L.schedule == DTR_SCHEDULE_GRID) {
        const uint32_t grid = grid_of(in->n, block);
        if (L.navigation == DTR_NAVIGATION_LANE)
            hipLaunchKernelGGL((propagate_grid_kernel<LanePolicy<P>, W>), dim3(grid), dim3(block),
                               lane_lds_bytes<LanePolicy<P>>(block), stream, det->fast, field->dev, dc, ti, r,
                               L.order, L.xcd_chunk);
        else if (L.navigation == DTR_NAVIGATION_GATHER)
            hipLaunchKernelGGL((propagate_gather_kernel<LanePolicy<P>, W>), dim3(grid), dim3(block),
                               lane_lds_bytes<LanePolicy<P>>(block), stream, det->fast, field->dev, dc, ti, r,
                               L.order, L.xcd_chunk);
        else
            hipLaunchKernelGGL((propagate_coop_kernel<P, W>), dim3(grid), dim3(block), shmem, stream, det->fast,
                               field->dev, dc, ti, r, L.order);
        return to_status(hipGetLastError());
    }
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, stream, L.queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return to_status(e);
    // resident waves: CUs x 4 SIMDs x waves_per_simd, never more than the
    // tracks need
    const uint64_t waves = static_cast<uint64_t>(det->n_cu) * 4u * L.waves_per_simd;
    const uint64_t need = (static_cast<uint64_t>(in->n) + 63u) / 64u;
    const uint64_t w = waves < need ? waves : need;
    const uint32_t wpb = block / 64u;
    const uint32_t grid = static_cast<uint32_t>((w + wpb - 1) / wpb);
    hipLaunchKernelGGL((propagate_coop_persistent_kernel<P, W>), dim3(grid), dim3(block), shmem, stream, det->fast,
                       field->dev, dc, ti, r, L.order, L.queue, L.tracks_per_claim, L.min_idle);
    return to_status(hipGetLastError());
}

template <class P>
int launch_lane(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const dtr_track_batch* in,
                const DevResults& r, uint32_t block, const uint32_t* order, int gather, hipStream_t stream) {
    if (gather)
        hipLaunchKernelGGL((propagate_gather_kernel<LanePolicy<P>, 2>), dim3(grid_of(in->n, block)), dim3(block),
                           lane_lds_bytes<LanePolicy<P>>(block), stream, det->fast, field->dev, dc, track_inputs(in),
                           r, order, 0u);
    else
        hipLaunchKernelGGL((propagate_grid_kernel<LanePolicy<P>, 2>), dim3(grid_of(in->n, block)), dim3(block),
                           lane_lds_bytes<LanePolicy<P>>(block), stream, det->fast, field->dev, dc, track_inputs(in),
                           r, order, 0u);
    return to_status(hipGetLastError());
}

template <class P>
int launch_resumable(const dtr_detector* det, const dtr_field* field, const DevConfig& dc,
                     const dtr_track_batch* in, const DevResults& r, uint32_t block, const uint32_t* order,
                     int gather, uint32_t xcd, const PauseView& pv, hipStream_t stream) {
    const size_t lds = lane_lds_bytes<LanePolicy<P>>(block);
    if (gather)
        hipLaunchKernelGGL((propagate_resumable_kernel<LanePolicy<P>, 2, 1>), dim3(grid_of(in->n, block)),
                           dim3(block), lds, stream, det->fast, field->dev, dc, track_inputs(in), r, order, xcd, pv);
    else
        hipLaunchKernelGGL((propagate_resumable_kernel<LanePolicy<P>, 2, 0>), dim3(grid_of(in->n, block)),
                           dim3(block), lds, stream, det->fast, field->dev, dc, track_inputs(in), r, order, xcd, pv);
    return to_status(hipGetLastError());
}

template <class P>
int launch_direct(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const DevConfig& dd,
                  const dtr_track_batch* in, const DevResults& r, const SeqView& sq, uint32_t block,
                  const uint32_t* order, hipStream_t stream) {
    const size_t shmem = static_cast<size_t>(block) * kCache * 16u;
    hipLaunchKernelGGL((propagate_direct_kernel<P, 2>), dim3(grid_of(in->n, block)), dim3(block), shmem, stream,
                       det->fast, field->dev, dc, dd, track_inputs(in), r, sq, order);
    return to_status(hipGetLastError());
}

template <class P, int W>
int launch_adjoint_record(const dtr_detector* det, const dtr_field* field, const DevConfig& dc, const AdjointIO& io,
                          uint32_t block, const uint32_t* order, hipStream_t stream) {
    const size_t shmem = static_cast<size_t>(block) * kCache * 16u;
    hipLaunchKernelGGL((adjoint_record_kernel<P, W>), dim3(grid_of(io.n, block)), dim3(block), shmem, stream,
                       det->fast, field->dev, dc, io, order);
    return to_status(hipGetLastError());
}

template <class P, int W>
int launch_adjoint_backward(const dtr_detector* det, const dtr_field* field, const DevConfig& dc,
                            const AdjointIO& io, uint32_t block, const uint32_t* order, int exact, double* grad,
                            hipStream_t stream) {
    hipLaunchKernelGGL((adjoint_backward_kernel<P, W>), dim3(grid_of(io.n, block)), dim3(block), 0, stream,
                       det->fast, field->dev, dc, io, order, exact, grad);
    return to_status(hipGetLastError());
}

}  // namespace dtr

/*
 * detray_amd — MI355X-native track propagation behind a C ABI.
 *
 * This header is the drop-in boundary for detray's propagation hot path:
 *   detray::propagator<rk_stepper, caching_navigator, actor_chain>::propagate
 *   (core/include/detray/propagator/propagator.hpp:260-400)
 * run over a batch of tracks on one GPU.  Every entry point takes plain
 * pointers and sizes, returns an int status code (never exits the process,
 * unlike DETRAY_HIP_ERROR_CHECK, core/include/detray/definitions/detail/
 * hip_definitions.hpp:20-31) and keeps detray's vocabulary: detector tables,
 * field, propagation config, actor chain, bound/free track parameters.
 *
 * Units follow detray (core/include/detray/definitions/units.hpp):
 * mm, GeV, e; magnetic field in GeV/(e mm) (1 T = 0.000299792458).
 *
 * Track parametrisation (core/include/detray/definitions/track_parametrization.hpp:18-64):
 *   free  vector [x, y, z, t, dx, dy, dz, q/p]
 *   bound vector [loc0, loc1, phi, theta, q/p, t]
 */
#ifndef DETRAY_AMD_H
#define DETRAY_AMD_H

#include <stddef.h>
#include <stdint.h>

/* Only the C ABI is exported from the shared library (it is built with
 * hidden visibility), so that its internal C++ symbols never interpose
 * with those of other libraries in the process. */
#if defined(__GNUC__) || defined(__clang__)
#define DTR_API __attribute__((visibility("default")))
#else
#define DTR_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
enum dtr_status_code {
    DTR_SUCCESS = 0,
    DTR_ERROR_INVALID_ARGUMENT = 1,
    DTR_ERROR_HIP = 2,
    DTR_ERROR_OUT_OF_MEMORY = 3,
    DTR_ERROR_NO_DEVICE = 4,
    DTR_ERROR_UNSUPPORTED = 5,
    DTR_ERROR_IO = 6           /* file cannot be opened, read or written, or is not JSON */
};

/* Per-track propagation outcome (navigation::status at the end of
 * propagate(), core/include/detray/definitions/navigation.hpp:24-31). */
enum dtr_track_status {
    DTR_TRACK_EXIT = 0,        /* left the world: propagate() returned true */
    DTR_TRACK_ABORT = 1,       /* navigation aborted (stall, no surfaces) */
    DTR_TRACK_PATH_LIMIT = 2,  /* pathlimit_aborter fired */
    DTR_TRACK_OVERFLOW = 3,    /* surface_sequencer capacity reached */
    DTR_TRACK_MOMENTUM_LIMIT = 4, /* momentum_aborter fired (p or pT at or below its limit) */
    DTR_TRACK_TARGET = 5,      /* target_aborter fired (the target surface was reached) */
    DTR_TRACK_PAUSED = 6,      /* paused (dtr_propagate_resumable): propagate() returned with the
                                  navigation alive, is_paused() (propagator.hpp:238-243) */
    DTR_TRACK_NOT_RUN = 255
};

/* ------------------------------------------------------------------------ */
/* Detector tables (flattened detray::detector<toy_metadata|telescope>)      */
/* ------------------------------------------------------------------------ */

/* Mask shapes (toy_metadata mask_id, detectors/include/detray/detectors/
 * toy_metadata.hpp:48-53; the telescope uses rectangles only). */
enum dtr_mask_shape {
    DTR_RECTANGLE2D = 0,          /* v = {half_x, half_y}                      */
    DTR_TRAPEZOID2D = 1,          /* v = {half_x_low, half_x_high, half_y, 1/(2 half_y)} */
    DTR_CONCENTRIC_CYLINDER2D = 2,/* v = {r, lower_z, upper_z}                  */
    DTR_RING2D = 3,               /* v = {inner_r, outer_r}                     */
    /* wires (shapes/line.hpp, line2D frame: loc0 = signed distance of closest
     * approach, loc1 = z along the wire; wire_chamber_metadata.hpp) */
    DTR_LINE_SQUARE = 4,          /* drift cell, v = {half cell size, half_z}  */
    DTR_LINE_CIRCULAR = 5,        /* straw tube, v = {radius, half_z}          */
    /* stereo annulus (shapes/annulus2D.hpp, polar2D frame like the ring):
     * v = {min_r, max_r, min_phi_rel, max_phi_rel, average_phi, shift_x, shift_y} */
    DTR_ANNULUS2D = 6,
    /* cylinder at any placement (shapes/cylinder2D.hpp, cylindrical2D frame:
     * loc0 = r phi, loc1 = z; two intersections per ray), v = {r, lower_z, upper_z} */
    DTR_CYLINDER2D = 7
};

/* Surface categories (core/include/detray/definitions/geometry.hpp:40-46). */
enum dtr_surface_id { DTR_PORTAL = 0, DTR_SENSITIVE = 1, DTR_PASSIVE = 2 };

/* Placement: global = R * local + t.  r is row-major, its columns are the
 * local x, y, z axes in global coordinates (generic_transform3.hpp:79-101). */
typedef struct dtr_transform {
    double r[9];
    double t[3];
} dtr_transform;

typedef struct dtr_mask {
    double v[7];          /* boundary values in the shape's order (64 B per mask) */
    uint16_t volume_link; /* 0xffff = leaves the world */
    uint8_t shape;        /* dtr_mask_shape */
    uint8_t pad[5];
} dtr_mask;

/* Homogeneous material slab (core/include/detray/material/material_slab.hpp,
 * material.hpp).  Derived quantities are filled by the builder.  The same
 * record describes a volume's homogeneous material (dtr_volume::material;
 * the thickness is then unused). */
enum dtr_material_kind {
    DTR_MATERIAL_SLAB = 0, /* material_slab.hpp: path thickness / cos(incidence) */
    DTR_MATERIAL_ROD = 1   /* material_rod.hpp (on wires): path
                              2 sqrt((r^2 - d^2) / (1 - cos^2)), d = loc0 */
};

typedef struct dtr_material_slab {
    double X0, L0, A, Z, mass_rho;
    double molar_rho;     /* mass_rho / (A * u) in mol/mm^3 */
    double mean_exc;      /* I = 16 eV * Z^0.9, or the density-effect table's I */
    double thickness;     /* mm (slab thickness, or rod radius) */
    double thickness_in_X0;
    uint32_t kind;        /* dtr_material_kind */
    uint32_t has_density; /* 1: density[] holds the Sternheimer density-effect
                             parameters (density_effect_data.hpp), which then
                             replace the approximate density correction of the
                             Bethe-Bloch formula (relativistic_quantities.hpp) */
    double density[5];    /* a, k (the exponent m), x0, x1, C-bar */
} dtr_material_slab;

/* Material map (material_map.hpp: grid<axes<shape>, bins::single<
 * material_slab>>): one slab per bin over the surface's bound local
 * coordinates (loc0, loc1) — (phi, z) on a concentric cylinder, (r, phi) on
 * a ring or annulus, (x, y) on a rectangle or trapezoid, (r phi, z) on a
 * placed cylinder — regular axes, a circular axis wrapping (phi over
 * [-pi, pi]), a closed one clamping to its edge bins (axis_bounds.hpp:
 * 107-206, axis_binning.hpp:80-115).  Bin (i0, i1) holds
 * materials[slabs + i0 * n1 + i1]; a slab of zero thickness is no material
 * (pointwise_material_interactor.hpp:84-88).  A surface uses a map with
 * material = DTR_MATERIAL_MAP | map index. */
#define DTR_MATERIAL_MAP 0x80000000u
typedef struct dtr_material_map {
    uint32_t n0, n1;      /* bins per axis */
    uint32_t slabs;       /* first of the n0 * n1 slabs in materials[] */
    uint8_t circular0, circular1;
    uint8_t pad[2];
    double min0, max0, min1, max1;
} dtr_material_map;

typedef struct dtr_surface {
    uint64_t barcode;   /* detray::geometry::identifier (identifier.hpp:153-157) */
    uint32_t transform; /* into transforms[] */
    uint32_t mask;      /* first mask in masks[] */
    uint32_t material;  /* into materials[], DTR_MATERIAL_MAP | index into
                           material_maps[], 0xffffffff = none */
    uint16_t volume;
    uint8_t n_masks;
    uint8_t sf_id;      /* dtr_surface_id */
    uint8_t shape;      /* dtr_mask_shape of all masks of this surface */
    uint8_t pad[7];
} dtr_surface;

/* Grid types on the volume's sensitive surfaces (toy_metadata.hpp:119-160) */
enum dtr_grid_type {
    DTR_GRID_CYLINDER = 0, /* axes (phi circular [-pi,pi), z closed) */
    DTR_GRID_RING = 1      /* axes (r closed, phi circular [-pi,pi)) */
};

typedef struct dtr_grid {
    double min0, max0, min1, max1; /* axis spans */
    double radius;                 /* reference cylinder radius (cylinder grids) */
    uint32_t n0, n1;               /* bins per axis */
    uint32_t bins;                 /* offset into grid_bins[]: n0*n1*capacity entries, entry k of
                                      bin (i0,i1) at (i0*n1 + i1)*capacity + k, unused entries
                                      0xffffffff */
    uint8_t type;                  /* dtr_grid_type */
    uint8_t capacity;              /* surfaces per bin, 0 means 1 (the toy's static_array<., 1>;
                                      the wire chamber's dynamic bins hold up to 3) */
    uint8_t pad[2];
} dtr_grid;

typedef struct dtr_volume {
    uint32_t transform;         /* placement */
    uint32_t bf_begin, bf_end;  /* brute-force surface list range in bf_surfaces[] */
    uint32_t sf_begin, sf_end;  /* all surfaces of the volume */
    int32_t grid;               /* -1: none */
    uint32_t material;          /* homogeneous volume material (tracking_volume.hpp:
                                   118, 285; homogeneous_volume_material_builder):
                                   0 = vacuum, k > 0: materials[k - 1] (kind
                                   DTR_MATERIAL_SLAB, thickness unused).  One-based
                                   so that zero-initialised records are vacuum. */
} dtr_volume;

typedef struct dtr_detector_desc {
    uint32_t n_volumes, n_surfaces, n_transforms, n_masks;
    uint32_t n_materials, n_grids, n_bf, n_bins;
    const dtr_volume* volumes;
    const dtr_surface* surfaces;
    const dtr_transform* transforms;
    const dtr_mask* masks;
    const dtr_material_slab* materials;
    const dtr_grid* grids;
    const uint32_t* bf_surfaces;
    const uint32_t* grid_bins; /* surface index or 0xffffffff */
    uint32_t n_material_maps;
    uint32_t pad;
    const dtr_material_map* material_maps;
} dtr_detector_desc;

/* ------------------------------------------------------------------------ */
/* Magnetic field                                                           */
/* ------------------------------------------------------------------------ */
enum dtr_field_kind { DTR_FIELD_CONSTANT = 0, DTR_FIELD_GRID = 1 };

/* Regular xyz grid of float3 samples, trilinear interpolation with clamping
 * at the border (the covfie affine -> linear -> clamp -> strided composition
 * used by tests/include/detray/test/common/bfield.hpp:44-47). */
typedef struct dtr_field_desc {
    int32_t kind;
    int32_t pad;
    double b[3];               /* constant field */
    uint32_t n[3];             /* grid points per axis */
    uint32_t pad2;
    double min[3];             /* position of sample (0,0,0) */
    double pitch[3];           /* sample spacing */
    const float* values;       /* [(ix*ny + iy)*nz + iz][3] */
} dtr_field_desc;

/* ------------------------------------------------------------------------ */
/* Optional actors (actor_chain.hpp:53-118, composite_actor.hpp:68-91)      */
/* ------------------------------------------------------------------------ */
/* The actor chain of every propagation is, in this order:
 *   pathlimit_aborter                              (aborters.hpp:23-75)
 *   parameter_updater<pointwise_material_interactor, step_tracer?>
 *                                                  (parameter_updater.hpp:
 *                                                   314-627, composite)
 *   surface_sequencer                              (surface_sequencer.hpp)
 *   momentum_aborter?                              (aborters.hpp:80-143)
 *   target_aborter?                                (aborters.hpp:146-175)
 * The actors marked ? are switched on by dtr_actor_states.actors; each
 * keeps its reference semantics (an aborter does nothing once the
 * navigation has finished; the step tracer observes the parameter
 * transporter, after the material interactor, before the setter). */
enum dtr_actor {
    DTR_ACTOR_MOMENTUM_ABORTER = 1u << 0,
    DTR_ACTOR_TARGET_ABORTER = 1u << 1,
    DTR_ACTOR_STEP_TRACER = 1u << 2
};

/* step_tracer output (tests/include/detray/test/utils/step_tracer.hpp:
 * step_data): one entry per parameter update (sensitive or material
 * surface), entry k of track i at k * n + i, DEVICE memory (HOST memory for
 * the oracle).  Any array but n may be NULL. */
typedef struct dtr_step_trace {
    uint32_t capacity;      /* entries per track; later ones are counted only */
    uint32_t pad;
    uint32_t* n;            /* [n] entries of each track (may exceed capacity) */
    uint32_t* surface;      /* [capacity][n] surface index */
    double* step_size;      /* [capacity][n] stepping.step_size() */
    double* path_length;    /* [capacity][n] stepping.path_length() */
    uint32_t* n_trials;     /* [capacity][n] stepping.n_total_trials() */
    double* free_params;    /* [8][capacity][n] stepping state on the surface */
    double* bound_params;   /* [6][capacity][n] after the material interactor */
    double* jacobian;       /* [64][capacity][n] stepping.transport_jacobian(),
                               free 8x8 row-major (already reset to the identity
                               after a transport, as in the reference) */
} dtr_step_trace;

/* States of the optional actors (the actor_states tuple of
 * propagator::propagate, propagator.hpp:260-264). */
typedef struct dtr_actor_states {
    uint32_t actors;         /* dtr_actor bits */
    float min_p;             /* momentum_aborter::state (10 MeV) */
    float min_pT;            /* momentum_aborter::state (10 MeV) */
    uint32_t pad;
    const uint32_t* target_surface; /* [n] target_aborter::state: surface index per track */
    dtr_step_trace trace;    /* step_tracer::state */
} dtr_actor_states;

/* Defaults of dtr_actor_states (no optional actor, the reference's limits). */
DTR_API void dtr_default_actors(dtr_actor_states* actors);

/* ------------------------------------------------------------------------ */
/* Propagation configuration (propagation::config, propagation_config.hpp;  */
/* stepping_config.hpp:27-48, navigation_config.hpp:23-40,                  */
/* intersection_config.hpp:20-36).  Same field types and defaults.          */
/* ------------------------------------------------------------------------ */
typedef struct dtr_prop_config {
    /* stepping::config */
    float min_stepsize;       /* 1e-4 mm */
    float rk_error_tol;       /* 1e-4 mm */
    float step_constraint;    /* FLT_MAX (accuracy constraint of constrained_step) */
    float path_limit;         /* 5 m (pathlimit_aborter) */
    uint32_t max_rk_updates;  /* 10000 */
    uint8_t use_mean_loss;    /* 1: in volume material the q/p of the RK
                                 stages 2-4 follows the energy loss of the
                                 previous stage (rk_stepper.hpp:534-565) */
    uint8_t use_eloss_gradient;    /* 0: 1 carries d(dqop/ds)/dqop of the
                                      volume material into the transport
                                      Jacobian (rk_stepper.hpp:195-231,
                                      767-800); no effect in vacuum */
    uint8_t use_field_gradient;    /* 0: 1 adds dB/dr (central differences,
                                      +-0.1 mm, at the end of each step) to
                                      the Jacobian's position columns,
                                      rk_stepper.hpp:878-998 */
    uint8_t do_covariance_transport; /* 1 */
    /* intersection::config */
    float min_mask_tolerance; /* 1e-5 mm */
    float max_mask_tolerance; /* 3 mm */
    float mask_tolerance_scalor; /* 5e-2 */
    float path_tolerance;     /* 1 um */
    float overstep_tolerance; /* -1000 um */
    /* navigation::config */
    uint32_t search_window[2];/* {0, 0} */
    float accumulated_error;  /* 0.001 */
    int32_t n_scattering_stddev; /* 3 */
    uint8_t estimate_scattering_noise; /* 1 */
    /* pointwise_material_interactor::state */
    uint8_t do_energy_loss;        /* 1 */
    uint8_t do_multiple_scattering;/* 1 */
    uint8_t use_material;          /* 1: run parameter_updater<pointwise_material_interactor> */
    /* surface_sequencer capacity (0 disables recording) */
    uint32_t max_surfaces;
    uint32_t pad;
    /* optional actors and their states, NULL for none (the adjoint passes
     * reject them with DTR_ERROR_UNSUPPORTED) */
    const dtr_actor_states* actors;
} dtr_prop_config;

/* ------------------------------------------------------------------------ */
/* Track batches (structure of arrays, parameter-major: x[n], y[n], ...)     */
/* ------------------------------------------------------------------------ */
typedef struct dtr_track_batch {
    uint32_t n;
    uint32_t pad;
    /* Free start (propagator::state(free_track_parameters, ...),
     * propagator.hpp:79-95): free_params [8][n], tracks start in volume 0.
     * initial_cov: NULL or [21][n] upper-triangle bound covariance used at
     * the first surface in place of the identity placeholder
     * (parameter_updater.hpp:392-397). */
    const double* free_params;
    const double* initial_cov;
    /* Bound start (propagator::state(bound_track_parameters, ...),
     * propagator.hpp:121-162): when bound_params is not NULL, track i starts
     * on surface start_surface[i] with bound parameters [6][n] and
     * covariance initial_cov (identity if NULL); free_params is ignored. */
    const double* bound_params;
    const uint32_t* start_surface;
} dtr_track_batch;

typedef struct dtr_result_batch {
    double* bound_params;   /* [6][n]  at the last sensitive/material surface */
    double* bound_cov;      /* [21][n] upper triangle, row-major (0,0),(0,1)..(5,5) */
    double* free_params;    /* [8][n]  at the end of propagation */
    double* path_length;    /* [n] */
    uint32_t* surfaces;     /* [max_surfaces][n] surface indices, in order */
    uint32_t* n_surfaces;   /* [n] */
    uint32_t* n_steps;      /* [n] RK steps taken */
    uint32_t* n_trials;     /* [n] RK trials (stepping.n_total_trials()) */
    uint8_t* status;        /* [n] dtr_track_status */
} dtr_result_batch;

/* ------------------------------------------------------------------------ */
/* Handles                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct dtr_detector dtr_detector; /* device-resident, immutable */
typedef struct dtr_field dtr_field;

/* Defaults of propagation::config (and the interactor state). */
DTR_API void dtr_default_config(dtr_prop_config* cfg);

/* Copy the detector tables to the current HIP device.
 * Replaces detray::get_buffer(detector) + detector(view)
 * (core/include/detray/core/detector.hpp:136-202). */
DTR_API int dtr_detector_create(const dtr_detector_desc* desc, dtr_detector** out);
DTR_API int dtr_detector_destroy(dtr_detector* det);

/* Field view on the device: constant or sampled grid.
 * Replaces covfie::field<...>::view_t (tests/include/detray/test/common/bfield.hpp:35-88). */
DTR_API int dtr_field_create(const dtr_field_desc* desc, dtr_field** out);
DTR_API int dtr_field_destroy(dtr_field* field);

/* Propagate a batch of tracks on the device.
 * in/out point to DEVICE memory (SoA, see above).  The actor chain is
 *   pathlimit_aborter, parameter_updater<pointwise_material_interactor>,
 *   surface_sequencer
 * as in tests/include/detray/test/device/propagator_test.hpp:87-99 and
 * core/include/detray/propagator/actors/surface_sequencer.hpp:25-106.
 * The stepper is rk_stepper<field, algebra, constrained_step> with the
 * caching_navigator (cache size 8).  Work is enqueued on `stream`
 * (a hipStream_t, NULL = default stream); the call does not synchronise. */
DTR_API int dtr_propagate(const dtr_detector* det, const dtr_field* field,
                  const dtr_prop_config* cfg, const dtr_track_batch* in,
                  dtr_result_batch* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched propagation: arithmetic and schedule of the launch               */
/* ------------------------------------------------------------------------ */
enum dtr_precision {
    DTR_PRECISION_FP64 = 0,  /* FP64 throughout */
    /* FP32 Runge-Kutta stages, field interpolation, transport Jacobian,
     * covariance transport and material noise terms; FP64 free track state,
     * path length, the step-size error estimate, the mean energy loss and
     * navigation geometry */
    DTR_PRECISION_MIXED = 1
};

enum dtr_schedule {
    DTR_SCHEDULE_GRID = 0,       /* one work-item per track, one launch */
    /* persistent waves over a device work queue: a lane whose track has
     * ended takes the next queued track */
    DTR_SCHEDULE_PERSISTENT = 1
};

enum dtr_navigation_mode {
    /* lanes that need a local navigation (volume search) pool their
     * candidate lists and the whole wave evaluates them */
    DTR_NAVIGATION_WAVE = 0,
    /* every lane searches on its own (grid schedule only) */
    DTR_NAVIGATION_LANE = 1,
    /* every lane searches on its own, and the lanes of a wave run their
     * searches together: a lane that needs one waits until every running
     * lane of its wave does (grid schedule only) */
    DTR_NAVIGATION_GATHER = 2
};

typedef struct dtr_launch_config {
    uint32_t precision;        /* dtr_precision (default MIXED) */
    uint32_t schedule;         /* dtr_schedule (default PERSISTENT) */
    uint32_t block_size;       /* 64, 128, 192 or 256 (default 256) */
    uint32_t waves_per_simd;   /* resident waves per SIMD the kernel is built for: 1
                                  (512 registers per lane) or 2 (256); the
                                  persistent schedule launches this many waves
                                  per SIMD; other values are rejected with
                                  DTR_ERROR_INVALID_ARGUMENT (default 2) */
    uint32_t tracks_per_claim; /* persistent: queue entries a wave claims per atomic,
                                  multiple of 64 (default 64) */
    uint32_t min_idle;         /* persistent: idle lanes a wave waits for before it
                                  takes new tracks, 1..64 (default 1) */
    uint32_t navigation;       /* dtr_navigation_mode (default WAVE) */
    /* grid schedule, per-lane or gathered navigation: consecutive blocks of
     * the track order that run on one accelerator complex die (XCD) at a
     * time.  The hardware deals blocks to the 8 XCDs of an MI355X in turn;
     * with chunks the waves that share an XCD's L2 work on neighbouring
     * tracks of the (sorted) batch.  0 or 1: hardware order; at most 65536
     * (default 64) */
    uint32_t xcd_chunk;
    /* optional DEVICE permutation [n]: the k-th queue entry / work-item
     * propagates track order[k]; results stay in track order */
    const uint32_t* order;
    /* persistent: DEVICE workspace of one uint32, reset on `stream` by the
     * call (a kernel, no memset), so the launch can be graph-captured */
    uint32_t* queue;
} dtr_launch_config;

/* Defaults of dtr_launch_config. */
DTR_API void dtr_default_launch(dtr_launch_config* launch);

/* dtr_propagate with a chosen precision and schedule.  Same actor chain,
 * navigation and results as dtr_propagate (the same
 * propagator::propagate per track, propagator.hpp:262-400); the surface
 * sequence of a track does not depend on the launch configuration.  Does
 * not allocate and does not synchronise; with the persistent schedule,
 * calls that share one `queue` must not overlap.  use_field_gradient runs
 * on the grid schedule with per-lane navigation only (DTR_ERROR_UNSUPPORTED
 * otherwise).  A detector with the extended geometry (wire surfaces,
 * annuli, placed cylinders, material rods, rings with material or grid bins
 * of several surfaces) runs on every schedule and navigation mode, compiled
 * for two waves per SIMD (use_field_gradient: per-lane navigation on the grid
 * schedule, as for any detector). */
DTR_API int dtr_propagate_ex(const dtr_detector* det, const dtr_field* field,
                     const dtr_prop_config* cfg, const dtr_track_batch* in,
                     dtr_result_batch* out, const dtr_launch_config* launch,
                     void* stream);

/* Track schedule: a permutation of the batch that groups tracks by
 * direction cell (pseudorapidity bins of width 1/eta_bins_per_unit over
 * |eta| < eta_max, phi_bins slices of phi), for dtr_launch_config.order.
 * A counting sort on the device: kernels on `stream` (clear, count, a
 * two-level scan, scatter), no allocation and no memset, so it can be
 * captured in a HIP graph together with the propagation.  At most 2^26
 * cells.  The order of tracks inside one cell is not specified (the results
 * of a track do not depend on it). */
typedef struct dtr_sort_config {
    uint32_t eta_bins_per_unit; /* default 20 */
    uint32_t phi_bins;          /* default 128 */
    float eta_max;              /* default 5; |eta| beyond goes to the edge bins */
    uint32_t pad;
} dtr_sort_config;

DTR_API void dtr_default_sort(dtr_sort_config* cfg);
/* bytes of device workspace dtr_sort_tracks needs for this configuration */
DTR_API size_t dtr_sort_workspace_size(const dtr_sort_config* cfg);
/* free_params: DEVICE [8][n]; order: DEVICE [n] output; workspace: DEVICE,
 * dtr_sort_workspace_size bytes. */
DTR_API int dtr_sort_tracks(const dtr_sort_config* cfg, const double* free_params, uint32_t n,
                            uint32_t* order, void* workspace, size_t workspace_bytes,
                            void* stream);

/* dtr_sort_tracks without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a library-owned
 * device array of at most 2^20 cells (4 MiB, held on every GPU the library
 * runs on until the process ends); configurations with more cells return
 * DTR_ERROR_INVALID_ARGUMENT.  All calls on one GPU share that array (its
 * counts are cleared by each call's first kernel): calls that overlap on
 * one GPU, on different streams, race.  Same results as dtr_sort_tracks. */
DTR_API int dtr_sort_tracks_static(const dtr_sort_config* cfg, const double* free_params, uint32_t n,
                                   uint32_t* order, void* stream);

/* ------------------------------------------------------------------------ */
/* Direct navigation: follow a given surface sequence                       */
/* ------------------------------------------------------------------------ */
/* Surface sequences of a batch, DEVICE memory: surfaces [max_surfaces][n]
 * (entry k of track i at k * n + i, the surface_sequencer layout of
 * dtr_result_batch), n_surfaces [n]. */
enum dtr_nav_direction {            /* navigation::direction (navigation_state.hpp:178) */
    DTR_DIRECTION_FORWARD = 0,      /* e_forward */
    DTR_DIRECTION_BACKWARD = 1      /* e_backward */
};

typedef struct dtr_surface_sequence {
    uint32_t max_surfaces;
    uint32_t direction;             /* dtr_nav_direction */
    const uint32_t* surfaces;
    const uint32_t* n_surfaces;
} dtr_surface_sequence;

/* propagator<rk_stepper, direct_navigator, actor chain>::propagate per track
 * (direct_navigator.hpp:29-298): instead of searching the detector, each
 * track steps to the next surface of its sequence, intersected with
 * infinite mask tolerance (the target is always hit); the navigation exits
 * after the last surface and aborts if a target cannot be reached or is not
 * a surface of the detector.  seq->direction DTR_DIRECTION_BACKWARD is the
 * reference's navigation().set_direction(e_backward) + reset()
 * (tests/integration_tests/cpu/propagator/propagation_direct_navigator.cpp:
 * 184-190): the targets are the sequence's entries from the last to the
 * first, intersected against the track direction, the stepper takes
 * negative steps and the material interactor gives the energy back
 * (pointwise_material_interactor.hpp:240-262); a track usually starts bound
 * on the sequence's last surface.  Same stepper, actor chain and results as
 * dtr_propagate_ex; the launch gives the precision, block size and track
 * order (grid schedule only, no use_field_gradient). */
DTR_API int dtr_propagate_direct(const dtr_detector* det, const dtr_field* field,
                                 const dtr_prop_config* cfg, const dtr_track_batch* in,
                                 const dtr_surface_sequence* seq, dtr_result_batch* out,
                                 const dtr_launch_config* launch, void* stream);

/* Pause and resume (propagator.hpp:238-250, 274-282: is_paused, resume, and
 * propagate() continuing a paused state without initialising the
 * navigation).  The reference pauses a track when an actor stops its
 * heartbeat while the navigation is alive; here a step budget per call
 * gives the pause: a track that has taken max_steps Runge-Kutta steps in the
 * call stops where the actor half of propagate()'s loop would begin, with
 * status DTR_TRACK_PAUSED and its step, trial and surface counts so far;
 * the surface sequence is written as it grows.  The next call with resume
 * = 1 continues every paused track from that point (tracks that finished
 * are left alone, their results stay as written).  A propagation done in
 * several calls gives the results of one dtr_propagate_ex call bit for bit.
 * Between calls the caller may copy, reorder or drop nothing in the state:
 * it is indexed by track.  data: DEVICE buffer of
 * dtr_propagation_state_size(n) bytes (about 0.9 KB per track). */
typedef struct dtr_propagation_state {
    void* data;
    uint32_t max_steps; /* steps a track takes in one call before it pauses, >= 1 */
    uint32_t resume;    /* 0: start every track from the batch's inputs; 1: continue the paused ones */
} dtr_propagation_state;

DTR_API size_t dtr_propagation_state_size(uint32_t n);

/* dtr_propagate_ex with pause and resume: grid schedule, per-lane or
 * gathered navigation (DTR_ERROR_UNSUPPORTED otherwise, and with
 * use_field_gradient), every call with the same detector, field, config,
 * track batch (its inputs are read by the first call; the initial
 * covariance of a free start also when a later call reaches the first
 * surface) and result batch.  No allocation, no synchronisation. */
DTR_API int dtr_propagate_resumable(const dtr_detector* det, const dtr_field* field,
                                    const dtr_prop_config* cfg, const dtr_track_batch* in,
                                    dtr_result_batch* out, const dtr_launch_config* launch,
                                    const dtr_propagation_state* state, void* stream);

/* ------------------------------------------------------------------------ */
/* Differentiation: d(chi2 of hit residuals)/d(initial bound parameters)    */
/* ------------------------------------------------------------------------ */
/* chi2 = sum over the hits matched on the track's sensitive surfaces of
 * (loc0 - m0)^2 / sigma_loc0^2 + (loc1 - m1)^2 / sigma_loc1^2, with loc the
 * bound parameters at the surface before its material (the quantities the
 * reference's parameter_updater.hpp:422-572 computes there).  A hit is
 * matched when the track crosses its surface, in the track's hit order.
 * The reference has no adjoint: it propagates Jacobians forward
 * (parameter_updater.hpp:422-572, rk_stepper.hpp:633-956); the gradient the
 * chain of those Jacobians gives is returned by dtr_adjoint_forward beside
 * chi2 and is what DTR_ADJOINT_TRANSPORT reproduces. */
enum dtr_adjoint_model {
    /* derivative of the propagation itself: field gradient of the
     * trilinear map, derivative of the mean energy loss, normalisation of
     * the direction */
    DTR_ADJOINT_EXACT = 0,
    /* the derivative the reference's transport Jacobians describe
     * (stepping::config use_field_gradient = use_eloss_gradient = false) */
    DTR_ADJOINT_TRANSPORT = 1
};

typedef struct dtr_adjoint_config {
    uint32_t max_steps;        /* tape capacity: Runge-Kutta steps per track (512) */
    uint32_t max_events;       /* tape capacity: sensitive / material surfaces per track (64) */
    uint32_t checkpoint_every; /* K: the free state is stored every K steps (32) */
    uint32_t precision;        /* dtr_precision of both passes (MIXED) */
    uint32_t model;            /* dtr_adjoint_model of the backward pass (EXACT) */
    uint32_t block_size;       /* work-items per block, 64..256, multiple of 64 (64) */
    uint32_t waves_per_simd;   /* register budget the kernels are built for: 1 or 2 (2);
                                  0 means 2, other values are rejected */
    uint32_t pad;
    double sigma_loc0;         /* hit resolution in loc0, mm (0.010) */
    double sigma_loc1;         /* hit resolution in loc1, mm (0.050) */
} dtr_adjoint_config;

/* Hits of a batch, DEVICE memory: surface [max_hits][n] (in crossing
 * order), loc [2][max_hits][n] (loc0 block, then loc1 block), n_hits [n]. */
typedef struct dtr_hits {
    uint32_t max_hits;
    uint32_t pad;
    uint32_t* surface;
    double* loc;
    uint32_t* n_hits;
} dtr_hits;

/* Tracks of a differentiated batch, DEVICE memory: bound start parameters
 * [6][n] on start_surface [n] (propagator::state(bound_track_parameters),
 * propagator.hpp:121-162), and the tape of dtr_adjoint_tape_size bytes that
 * the forward pass writes and the backward pass reads.  order: optional
 * permutation [n] (work-item k runs track order[k]); the backward pass
 * must get the order of the forward pass.  Results stay in track order. */
typedef struct dtr_adjoint_batch {
    uint32_t n;
    uint32_t pad;
    const double* bound_params;
    const uint32_t* start_surface;
    void* tape;
    size_t tape_bytes;
    const uint32_t* order;
} dtr_adjoint_batch;

/* Per-track tape status written by dtr_adjoint_forward. */
enum dtr_tape_status {
    DTR_TAPE_OK = 0,
    DTR_TAPE_STEP_OVERFLOW = 1,   /* more than max_steps steps */
    DTR_TAPE_EVENT_OVERFLOW = 2,  /* more than max_events surfaces */
    /* not differentiable here: the track stopped in material, or it
     * started free */
    DTR_TAPE_UNSUPPORTED = 3
};

/* Defaults of dtr_adjoint_config. */
DTR_API void dtr_default_adjoint(dtr_adjoint_config* cfg);
/* Bytes of tape for n tracks (about 26 KB per track at the defaults). */
DTR_API size_t dtr_adjoint_tape_size(const dtr_adjoint_config* cfg, uint32_t n);

/* Forward pass: propagator::propagate of every track from its bound start
 * (same stepper, navigator and actor chain as dtr_propagate, without the
 * surface sequencer), writing the tape.  Optional DEVICE outputs (NULL to
 * skip): chi2 [n] against `hits`; grad_transport [6][n], the gradient of
 * chi2 from the chain of the reference's bound-to-bound transport
 * Jacobians; status [n] (dtr_tape_status).  `record`, if not NULL, receives
 * the bound loc0/loc1 of every sensitive surface crossed (true
 * intersections, for making hits).  `hits` may be NULL (chi2 = 0).
 * cfg->use_field_gradient is rejected with DTR_ERROR_UNSUPPORTED: the
 * exact model already carries the full field derivative. */
DTR_API int dtr_adjoint_forward(const dtr_detector* det, const dtr_field* field,
                                const dtr_prop_config* cfg, const dtr_adjoint_config* acfg,
                                const dtr_adjoint_batch* batch, const dtr_hits* hits,
                                dtr_hits* record, double* chi2, double* grad_transport,
                                uint32_t* status, void* stream);

/* Backward pass: grad [6][n] (DEVICE) = d chi2 / d(bound start parameters)
 * of every track, from the tape of the last dtr_adjoint_forward on this
 * batch (same det, field, cfg, acfg precision and hits).  The segment of K
 * steps before each checkpoint is recomputed from it.  Tracks whose tape
 * status is not DTR_TAPE_OK get NaN. */
DTR_API int dtr_adjoint_backward(const dtr_detector* det, const dtr_field* field,
                                 const dtr_prop_config* cfg, const dtr_adjoint_config* acfg,
                                 const dtr_adjoint_batch* batch, const dtr_hits* hits,
                                 double* grad, void* stream);

/* Host-side builders (detector emitters and inputs).  The returned tables
 * are owned by the library; free them with dtr_host_detector_free. */
typedef struct dtr_toy_config {
    uint32_t n_brl_layers;    /* 4 (build_toy_detector.hpp:117) */
    uint32_t n_edc_layers;    /* 3 (build_toy_detector.hpp:119) */
    uint8_t use_material;     /* 0: none; 1: Si slabs on modules, Be beampipe;
                                 2: material maps on every portal and the
                                 beampipe, none on modules
                                 (toy_det_config::use_material_maps) */
    uint8_t pad[3];
} dtr_toy_config;

typedef struct dtr_telescope_config {
    uint32_t n_planes;
    uint32_t pad;
    const double* positions; /* along the pilot ray (0,0,0)->(0,0,1), mm */
    double half_x, half_y;   /* rectangle half lengths */
    double thickness;        /* silicon_tml slab thickness, 0 = no material */
    double envelope;         /* cuboid portal envelope (tel_det_config::envelope) */
    /* tel_det_config::volume_material (build_telescope_detector.hpp:82,
     * 262-267): homogeneous material of the telescope volume, NULL = vacuum */
    const dtr_material_slab* volume_material;
} dtr_telescope_config;

/* build_wire_chamber.hpp:120-331 (wire_chamber_config defaults in
 * parentheses): an empty inner volume, then n_layers cylinder layers of
 * 2*cell_size thickness, each with one ring of wires at its central radius,
 * alternating stereo angle, tungsten rods on the wires, a 100 x 1 phi/z grid
 * per layer. */
typedef struct dtr_wire_chamber_config {
    uint32_t n_layers;           /* 10 */
    uint32_t square_cells;       /* 1: drift cells (line_square), 0: straw tubes */
    double first_layer_inner_radius; /* 500 mm */
    double half_z;               /* 1000 mm */
    double cell_size;            /* 10 mm (wire_layer_generator_config) */
    double stereo_angle;         /* 50 mrad */
    double mat_radius;           /* 15 um rod radius */
} dtr_wire_chamber_config;

DTR_API int dtr_build_toy_detector(const dtr_toy_config* cfg, dtr_detector_desc* out);
DTR_API int dtr_build_wire_chamber(const dtr_wire_chamber_config* cfg, dtr_detector_desc* out);
DTR_API int dtr_build_telescope_detector(const dtr_telescope_config* cfg,
                                 dtr_detector_desc* out);
DTR_API void dtr_host_detector_free(dtr_detector_desc* desc);

/* The table checks dtr_detector_create applies, on the host: every index the
 * navigation follows in range (volumes, surfaces, masks, materials and
 * material maps, grid bins), known shapes and ids, maps whose slabs exist
 * and are slabs.  The part of detail::check_consistency
 * (utils/consistency_checker.hpp:314-380) that flat tables can break.
 * DTR_SUCCESS, or DTR_ERROR_INVALID_ARGUMENT with the first problem in
 * `message` (may be NULL). */
DTR_API int dtr_check_detector(const dtr_detector_desc* desc, char* message, size_t message_size);

/* A material of predefined_materials.hpp by its name there ("silicon",
 * "silicon_with_ded", "silicon_tml", "beryllium_tml", "tungsten",
 * "isobutane", "argon_liquid", "iron_with_ded", "copper_with_ded",
 * "cesium_iodide", "cesium_iodide_with_ded"), as a slab of the given
 * thickness (0 for a volume material).  DTR_ERROR_INVALID_ARGUMENT for an
 * unknown name. */
DTR_API int dtr_predefined_material(const char* name, double thickness, dtr_material_slab* out);

/* Detector tables from detray's JSON detector files: the geometry file and,
 * optionally (NULL to skip), the surface-grid, homogeneous-material and
 * material-map files.  Replaces io::read_detector with the json
 * geometry_reader, surface_grid_reader, homogeneous_material_reader and
 * material_map_reader (io/include/detray/io/frontend/detector_reader.hpp,
 * io/include/detray/io/backend/geometry_reader.hpp:51-170,
 * detail/grid_reader.hpp:55-420, homogeneous_material_reader.hpp:40-120,
 * material_map_reader.hpp:51-185): the tables are those the reference's
 * detector builder makes of the files (volume order, surface order by
 * index_in_coll or by portal/shape factories, cylinder z shifts folded into
 * the masks, one slab per material surface, one dtr_material_map per mapped
 * surface with its slabs after the volume's homogeneous ones, bins the file
 * leaves out holding no material).  Material maps need regular closed or
 * circular axes and the map type of the surface's frame (a surface with
 * both a slab and a map is an error).  Shapes, grids and materials the
 * tables cannot hold return
 * DTR_ERROR_UNSUPPORTED; malformed files DTR_ERROR_IO or
 * DTR_ERROR_INVALID_ARGUMENT, with a description in `message` (may be NULL).
 * Free the tables with dtr_host_detector_free. */
DTR_API int dtr_read_detector_json(const char* geometry_file, const char* grids_file,
                                   const char* material_file, const char* material_maps_file,
                                   dtr_detector_desc* out, char* message, size_t message_size);

/* In-memory detector view: the reference's IO payloads
 * (io/include/detray/io/frontend/payloads.hpp:86-330) as plain arrays, so a
 * detector held by the reference (detray::detector<metadata>) reaches the
 * tables without files: io::geometry_writer::convert, the surface-grid
 * writer's convert and homogeneous_material_writer::convert make the payloads
 * of a detector (geometry_writer.hpp, surface_grid_writer.hpp,
 * homogeneous_material_writer.hpp), the C++ adapter (detray_amd.hpp,
 * flatten()) copies them into these records and dtr_flatten_detector builds
 * the tables with the semantics of dtr_read_detector_json.  Enumerations keep
 * the reference's numbering (io::shape_id, io::accel_id, io::material_id,
 * axis::binning/bounds/label, surface_id). */
typedef struct dtr_view_mask {      /* mask_payload */
    uint32_t shape;                 /* io::shape_id */
    uint32_t n_boundaries;          /* entries of boundaries[] in use */
    uint64_t volume_link;           /* single_link_payload: volume index, 2^64-1 = leaves the world */
    double boundaries[8];
} dtr_view_mask;

typedef struct dtr_view_transform { /* transform_payload */
    double tr[3];
    double rot[9];                  /* column major */
} dtr_view_transform;

typedef struct dtr_view_surface {   /* surface_payload */
    uint32_t type;                  /* surface_id: 0 portal, 1 sensitive, 2 passive */
    uint32_t n_masks;
    uint64_t index_in_coll;         /* 2^64-1: not given */
    dtr_view_transform transform;
    const dtr_view_mask* masks;
} dtr_view_surface;

typedef struct dtr_view_volume {    /* volume_payload */
    dtr_view_transform transform;
    uint32_t n_surfaces;
    uint32_t pad;
    const dtr_view_surface* surfaces;
} dtr_view_volume;

typedef struct dtr_view_axis {      /* axis_payload (regular axes: the first and last edge) */
    uint32_t binning, bounds, label, bins;
    double edges[2];
} dtr_view_axis;

typedef struct dtr_view_grid {      /* grid_payload<std::size_t, io::accel_id> */
    uint64_t owner_link;            /* volume index */
    uint32_t type;                  /* grid_link.type: io::accel_id */
    uint32_t n_axes;
    dtr_view_axis axes[2];
    uint32_t n_bins;
    uint32_t pad;
    const uint32_t* loc_index;      /* [n_bins][2] grid_bin_payload::loc_index */
    const uint64_t* content_offset; /* [n_bins + 1]: bin k holds content[offset[k] .. offset[k+1]) */
    const uint64_t* content;        /* volume-local surface indices */
} dtr_view_grid;

typedef struct dtr_view_material {  /* surface_material_payload of material_volume_payload */
    uint64_t volume_link;           /* material_volume_payload::volume_link */
    uint64_t surface;               /* volume-local surface index (surface_material_payload::surface) */
    uint64_t index_in_coll;         /* 2^64-1: not given */
    uint32_t type;                  /* io::material_id: slab or rod */
    uint32_t pad;
    double thickness;               /* a rod's radius */
    double params[7];               /* material_param_payload: X0, L0, Ar, Z, mass density, molar density, state */
} dtr_view_material;

typedef struct dtr_detector_view {
    uint32_t n_volumes, n_grids, n_materials, pad;
    const dtr_view_volume* volumes;     /* detector_payload::volumes */
    const dtr_view_grid* grids;         /* detector_grids_payload::grids, any order */
    const dtr_view_material* materials; /* detector_homogeneous_material_payload, any order */
} dtr_detector_view;

/* Tables of an in-memory detector view (above), as dtr_read_detector_json
 * makes them of the same payloads read from files (same errors and
 * messages).  Free the tables with dtr_host_detector_free. */
DTR_API int dtr_flatten_detector(const dtr_detector_view* view, dtr_detector_desc* out, char* message,
                                 size_t message_size);

/* Write detector tables as detray JSON files (geometry; surface grids,
 * homogeneous material and material maps, NULL to skip), as
 * io::write_detector with the json geometry_writer, surface_grid_writer,
 * homogeneous_material_writer and material_map_writer
 * (io/include/detray/io/frontend/detector_writer.hpp,
 * io/include/detray/io/backend/geometry_writer.hpp:40-150,
 * surface_grid_writer.hpp:40-85, homogeneous_material_writer.hpp:40-190,
 * material_map_writer.hpp:45-110) write them.  volume_names may be NULL
 * (empty names). */
DTR_API int dtr_write_detector_json(const dtr_detector_desc* det, const char* detector_name,
                                    const char* const* volume_names, const char* geometry_file,
                                    const char* grids_file, const char* material_file,
                                    const char* material_maps_file, char* message, size_t message_size);

/* Sample the analytic finite solenoid (on-axis 2 T, radius 1.2 m, half
 * length 3 m) on a regular grid.  `values` must hold nx*ny*nz*3 floats. */
DTR_API int dtr_sample_solenoid(double b_center, double radius, double half_length,
                        const uint32_t n[3], const double min[3],
                        const double pitch[3], float* values);

/* Seeded track generators (counter-based SplitMix64; identical bits on
 * every host).  `free_params` is host memory [8][n]. */
typedef struct dtr_track_gen_config {
    uint64_t seed;
    uint32_t n;
    uint32_t kind;            /* 0: telescope beam, 1: toy muons */
    /* kind 0: |p| uniform [p_min,p_max], polar angle |N(0, sigma_theta)| */
    double p_min, p_max, sigma_theta;
    /* kind 1: eta uniform, phi uniform, pT log-uniform, Gaussian vertex */
    double eta_min, eta_max, pt_min, pt_max;
    double sigma_xy, sigma_z;
} dtr_track_gen_config;

DTR_API int dtr_generate_tracks(const dtr_track_gen_config* cfg, double* free_params);

/* Library version string. */
DTR_API const char* dtr_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DETRAY_AMD_H */

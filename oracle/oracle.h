/*
 * Oracle: scalar FP64 CPU restatement of detray's propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product (detray_amd) never
 * links or calls it.
 *
 * It consumes the same flat detector / field / config / track tables as the
 * device path (include/detray_amd.h) and writes results in the same SoA
 * layout, but in HOST memory.
 */
#ifndef DETRAY_AMD_ORACLE_H
#define DETRAY_AMD_ORACLE_H

#include "detray_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-track debugging trace of the navigation (optional). */
typedef struct oracle_trace {
    uint32_t capacity;       /* entries per track */
    uint32_t* volume;        /* [n][capacity] volume after each navigator update */
    uint32_t* surface;       /* [n][capacity] current surface (or 0xffffffff) */
    uint32_t* n;             /* [n] */
} oracle_trace;

/* Propagate tracks on the host with n_threads OpenMP threads (0 = all,
 * dynamic schedule, chunk 1 as the reference CPU benchmark,
 * tests/benchmarks/include/detray/benchmarks/cpu/propagation_benchmark.hpp:72-148).
 * All pointers of `in`/`out` are HOST pointers.  trace may be NULL. */
int oracle_propagate(const dtr_detector_desc* det, const dtr_field_desc* field,
                     const dtr_prop_config* cfg, const dtr_track_batch* in,
                     dtr_result_batch* out, int n_threads, oracle_trace* trace);

/* oracle_propagate that also counts, per track, the ray-surface
 * intersections the navigation computed (n_isect [n], HOST memory): the
 * N_isect of the algorithmic flop model. */
int oracle_propagate_counted(const dtr_detector_desc* det, const dtr_field_desc* field,
                             const dtr_prop_config* cfg, const dtr_track_batch* in,
                             dtr_result_batch* out, int n_threads, uint32_t* n_isect);

/* Bound loc0 / loc1 of the updater at every sensitive surface a track
 * crosses (before the material interaction), HOST memory: surface
 * [max_hits][n], loc [2][max_hits][n], n_hits [n]. */
typedef struct oracle_hits {
    uint32_t max_hits;
    uint32_t pad;
    uint32_t* surface;
    double* loc;
    uint32_t* n_hits;
} oracle_hits;

/* oracle_propagate that also records the sensitive-surface parameters
 * (hits may be NULL). */
int oracle_propagate_hits(const dtr_detector_desc* det, const dtr_field_desc* field,
                          const dtr_prop_config* cfg, const dtr_track_batch* in,
                          dtr_result_batch* out, oracle_hits* hits, int n_threads,
                          oracle_trace* trace);

/* Propagation with the direct navigator (direct_navigator.hpp): every track
 * follows its surface sequence [max_sequence][n] (n_sequence[n] entries) instead
 * of searching the detector; same stepper and actors.  direction
 * DTR_DIRECTION_BACKWARD visits the sequence from its last entry to its first,
 * stepping against the track direction. */
int oracle_propagate_direct(const dtr_detector_desc* det, const dtr_field_desc* field,
                            const dtr_prop_config* cfg, const dtr_track_batch* in,
                            const uint32_t* sequence, const uint32_t* n_sequence, uint32_t max_sequence,
                            uint32_t direction, dtr_result_batch* out, int n_threads);

/* Evaluate the field at n points (host), for field tests. */
int oracle_field_at(const dtr_field_desc* field, uint32_t n, const double* xyz,
                    double* bxyz);

/* Material physics probes (interaction.hpp), for KAT tests.
 * Returns the Bethe-Bloch stopping power -dE/dx [GeV/mm] of a muon with
 * q/p = qop in slab material `mat`, the Landau sigma(q/p) over path `path`
 * and the Highland theta0 for path/X0 = x_over_x0. */
int oracle_material_probe(const dtr_material_slab* mat, double qop,
                          double path, double x_over_x0, double out[3]);

/* Ray-wire intersection probe (ray_line_intersector.hpp with a fixed mask
 * tolerance `tol`, overstep tolerance 0), for KAT tests:
 * out = {path, loc0, loc1, inside}.  Returns 1 for a valid intersection. */
/* Energy loss of the muon hypothesis in a homogeneous material:
 * out = {-dE/dx (Bethe-Bloch), d(qop)/ds, d(d(qop)/ds)/dqop}
 * (interaction.hpp, rk_stepper.hpp:157-231). */
int oracle_eloss_probe(const dtr_material_slab* mat, double qop, double out[3]);

/* The oracle's stepper alone, as rk_stepper.cpp:220-263 drives it: from the
 * free parameters fp[8], n_steps Runge-Kutta steps of ds (step size reset
 * each time) in a homogeneous volume material (NULL: vacuum).  out[5 * (n_steps
 * + 1)]: per row (start, then after each step) qop, d(qop)/ds, d2(qop)/ds dqop,
 * path length and the transport Jacobian element J(qop, qop).  jac (NULL or
 * 72 values): the final 8x8 free transport Jacobian, then the final free
 * parameters. */
int oracle_rk_steps(const dtr_field_desc* field, const dtr_material_slab* vol_mat, const dtr_prop_config* cfg,
                    const double fp[8], double ds, uint32_t n_steps, double* out, double* jac);

int oracle_line_intersect(const double ro[3], const double rd[3], const dtr_transform* trf,
                          const dtr_mask* mask, double tol, double out[4]);

/* Candidate intersections of a ray with surface `sf` of a detector as the
 * navigator's candidate search makes them (every solution a mask accepts,
 * fixed mask tolerance `tol`, external tolerance `ext_tol`, overstep
 * tolerance `overstep`): out[k] = {path, x, y, z, status (0 inside, 1 edge)}
 * for k < return value (at most 2). */
int oracle_surface_candidates(const dtr_detector_desc* det, uint32_t sf, const double ro[3],
                              const double rd[3], double tol, double ext_tol, double overstep,
                              double out[2][5]);

#ifdef __cplusplus
}
#endif

#endif

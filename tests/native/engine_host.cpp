// Host build of the batched engine (detray_amd/csrc/device/engine.hpp, the
// code the GPU kernels run), for CPU tests.  TEST INFRASTRUCTURE: only
// tests/ load this library.
//
//  * dtr_test_prefilter: the FP32 grid prefilter against the exact
//    candidate evaluation for rays x every grid-bin record of a detector;
//    the prefilter may only reject surfaces the exact mask check rejects.
//  * dtr_test_engine_propagate: the engine's per-lane propagate loop
//    (propagate_fast) run track by track on the host, with the same tables
//    and results as dtr_propagate_ex, to compare with the oracle on the CPU.
//  * dtr_test_engine_propagate_direct: the direct navigator's loop
//    (propagate_direct) the same way, forward or backward.
//  * dtr_test_adjoint: the adjoint's recording pass and backward pass
//    (adjoint_record, adjoint_backward) track by track on the host.
//  * dtr_test_engine_propagate_resumable: the lane loop with a step budget
//    per call (propagate_resumable), pausing and resuming.
#include <cstdint>
#include <cstring>
#include <vector>

#include "device/adjoint.hpp"
#include "device/direct_nav.hpp"
#include "device/engine.hpp"

namespace {

using namespace dtr;

// The device-side tables of a detector, built on the host as
// dtr_detector_create builds them.
struct HostDetector {
    std::vector<NavSurface> sf;
    std::vector<NavRecord> bin_rec, bf_rec;
    std::vector<float> tolk;
    FastDetector fast{};

    explicit HostDetector(const dtr_detector_desc& d)
        : sf(d.n_surfaces), bin_rec(d.n_bins), bf_rec(d.n_bf), tolk(d.n_grids) {
        for (uint32_t k = 0; k < d.n_surfaces; ++k) sf[k] = make_nav_surface(d, k);
        make_nav_records(d, sf.data(), bin_rec.data(), bf_rec.data(), tolk.data());
        fast.sf = sf.data();
        fast.masks = d.masks;
        fast.materials = d.materials;
        fast.mat_maps = d.material_maps;
        fast.volumes = d.volumes;
        fast.grids = d.grids;
        fast.transforms = d.transforms;
        fast.bf = d.bf_surfaces;
        fast.bins = d.grid_bins;
        fast.bin_rec = bin_rec.data();
        fast.bf_rec = bf_rec.data();
        fast.grid_tolk = tolk.data();
    }
};

bool has_volume_material(const dtr_detector_desc& d) {
    for (uint32_t k = 0; k < d.n_volumes; ++k)
        if (d.volumes[k].material != 0u) return true;
    return false;
}

bool has_ext(const dtr_detector_desc& d) {
    if (has_volume_material(d)) return true;
    for (uint32_t k = 0; k < d.n_surfaces; ++k) {
        const dtr_surface& s = d.surfaces[k];
        if (s.shape >= DTR_LINE_SQUARE) return true;
        if (s.shape == DTR_RING2D && (s.sf_id == DTR_SENSITIVE || s.material != 0xffffffffu)) return true;
    }
    for (uint32_t k = 0; k < d.n_materials; ++k)
        if (d.materials[k].kind == DTR_MATERIAL_ROD) return true;
    for (uint32_t k = 0; k < d.n_grids; ++k)
        if (d.grids[k].capacity > 1u) return true;
    return false;
}

template <class P>
void prefilter_pairs(const dtr_detector_desc& d, const HostDetector& hd, const DevConfig& cfg, const double* rays,
                     uint32_t n_rays, double ext_tol, uint64_t out[3], uint32_t first_miss[2]) {
    const DevField no_field{};
    double path[kCache];
    uint32_t sfs[kCache], ls[kCache];
    const CacheView c{path, sfs, ls, 1};
    for (uint32_t r = 0; r < n_rays; ++r) {
        NavLite nl{};
        nl.ext_tol = ext_tol;
        nl.pos = {rays[6 * r + 0], rays[6 * r + 1], rays[6 * r + 2]};
        nl.dir = {rays[6 * r + 3], rays[6 * r + 4], rays[6 * r + 5]};
        FastEngine<P, NavLite> e{hd.fast, no_field, cfg, nl, c};
        const float ox = static_cast<float>(nl.pos.x), oy = static_cast<float>(nl.pos.y),
                    oz = static_cast<float>(nl.pos.z);
        const float dx = static_cast<float>(nl.dir.x), dy = static_cast<float>(nl.dir.y),
                    dz = static_cast<float>(nl.dir.z);
        const float ft = static_cast<float>(e.tol_max());
        for (uint32_t g = 0; g < d.n_grids; ++g) {
            const dtr_grid& gr = d.grids[g];
            const float margin = e.sphere_margin(static_cast<int>(g));
            const uint32_t nb = gr.n0 * gr.n1 * (gr.capacity ? gr.capacity : 1u);
            for (uint32_t k = 0; k < nb; ++k) {
                const NavRecord& rec = hd.bin_rec[gr.bins + k];
                if (rec.rec.index == kInvalid) continue;
                out[0]++;
                Cand cd{};
                bool exact;
                if (P::kExt && rec.rec.shape == DTR_CYLINDER2D) {
                    // both solutions, as the candidate search takes them
                    nl.next = 0;
                    nl.last = -1;
                    for (int i = 0; i < kCache; ++i) path[i] = kMax;
                    e.candidate_search(rec.rec, cfg.overstep_tol);
                    exact = nl.last >= 0;
                } else {
                    exact = e.evaluate(rec.rec, cfg.overstep_tol, cd);
                }
                if (!exact) continue;
                out[1]++;
                if (!FastEngine<P, NavLite>::prefilter(rec.box, ox, oy, oz, dx, dy, dz, margin, ft)) {
                    if (out[2] == 0) {
                        first_miss[0] = r;
                        first_miss[1] = rec.rec.index;
                    }
                    out[2]++;
                }
            }
        }
    }
}

template <class P>
void propagate_all(const HostDetector& hd, const DevField& field, const DevConfig& cfg, const TrackInputs& in,
                   const DevResults& out) {
    double path[kCache];
    uint32_t sfs[kCache], ls[kCache];
    const CacheView c{path, sfs, ls, 1};
    for (uint64_t t = 0; t < in.n; ++t)
        propagate_fast<P>(hd.fast, field, cfg, in, out, static_cast<uint32_t>(t), c);
}

template <class P>
void propagate_direct_all(const HostDetector& hd, const DevField& field, const DevConfig& cfg, const TrackInputs& in,
                          const DevResults& out, const SeqView& sq) {
    double path[kCache];
    uint32_t sfs[kCache], ls[kCache];
    const CacheView c{path, sfs, ls, 1};
    const DevConfig dd = direct_config(cfg);
    for (uint64_t t = 0; t < in.n; ++t)
        propagate_direct<P>(hd.fast, field, cfg, dd, in, out, sq, static_cast<uint32_t>(t), c);
}

template <class P>
void adjoint_all(const HostDetector& hd, const DevField& field, const DevConfig& cfg, const AdjointIO& io,
                 bool exact, double* grad) {
    double path[kCache];
    uint32_t sfs[kCache], ls[kCache];
    const CacheView c{path, sfs, ls, 1};
    for (uint64_t t = 0; t < io.n; ++t)
        adjoint_record<P>(hd.fast, field, cfg, io, static_cast<uint32_t>(t), static_cast<uint32_t>(t), c);
    if (grad)
        for (uint64_t t = 0; t < io.n; ++t)
            adjoint_backward<P>(hd.fast, field, cfg, io, exact, static_cast<uint32_t>(t), static_cast<uint32_t>(t),
                                grad);
}

template <class P>
void propagate_resumable_all(const HostDetector& hd, const DevField& field, const DevConfig& cfg,
                             const TrackInputs& in, const DevResults& out, const PauseView& pv) {
    double path[kCache];
    uint32_t sfs[kCache], ls[kCache];
    const CacheView c{path, sfs, ls, 1};
    for (uint64_t t = 0; t < in.n; ++t)
        propagate_resumable<P>(hd.fast, field, cfg, in, out, static_cast<uint32_t>(t), c, pv);
}

DevField host_field(const dtr_field_desc* f) {
    DevField field{};
    field.kind = f->kind;
    for (int i = 0; i < 3; ++i) {
        field.b[i] = f->b[i];
        field.n[i] = f->n[i];
        field.min[i] = f->min[i];
        field.pitch[i] = f->pitch[i];
    }
    field.values = f->values;
    return field;
}

}  // namespace

extern "C" {

// rays: [n_rays][6] origin and unit direction.  out = {evaluated pairs,
// exactly accepted pairs, accepted pairs the prefilter rejected};
// first_miss = {ray, surface} of the first miss.
__attribute__((visibility("default"))) int dtr_test_prefilter(const dtr_detector_desc* d,
                                                              const dtr_prop_config* cfg, const double* rays,
                                                              uint32_t n_rays, double ext_tol, uint64_t out[3],
                                                              uint32_t first_miss[2]) {
    if (!d || !cfg || !rays || !out || !first_miss) return DTR_ERROR_INVALID_ARGUMENT;
    HostDetector hd(*d);
    const DevConfig dc = make_dev_config(cfg);
    out[0] = out[1] = out[2] = 0u;
    if (has_ext(*d))
        prefilter_pairs<WithExtendedGeometry<PolicyFP64>>(*d, hd, dc, rays, n_rays, ext_tol, out, first_miss);
    else
        prefilter_pairs<PolicyFP64>(*d, hd, dc, rays, n_rays, ext_tol, out, first_miss);
    return DTR_SUCCESS;
}

// precision: 0 FP64 policy, 1 mixed.  Host memory throughout.
__attribute__((visibility("default"))) int dtr_test_engine_propagate(const dtr_detector_desc* d,
                                                                     const dtr_field_desc* f,
                                                                     const dtr_prop_config* cfg,
                                                                     const dtr_track_batch* in,
                                                                     dtr_result_batch* out, int precision) {
    if (!d || !f || !cfg || !in || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    HostDetector hd(*d);
    DevField field{};
    field.kind = f->kind;
    for (int i = 0; i < 3; ++i) {
        field.b[i] = f->b[i];
        field.n[i] = f->n[i];
        field.min[i] = f->min[i];
        field.pitch[i] = f->pitch[i];
    }
    field.values = f->values;
    const DevConfig dc = make_dev_config(cfg);
    const TrackInputs ti{in->bound_params ? nullptr : in->free_params, in->initial_cov, in->bound_params,
                         in->start_surface, in->n};
    const DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                       cfg->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                       out->n_trials, out->status};
    const bool ext = has_ext(*d), vol = has_volume_material(*d);
    if (precision == 0) {
        if (vol) propagate_all<WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>>(hd, field, dc, ti, r);
        else if (ext) propagate_all<WithExtendedGeometry<PolicyFP64>>(hd, field, dc, ti, r);
        else propagate_all<PolicyFP64>(hd, field, dc, ti, r);
    } else {
        if (vol) propagate_all<WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>>(hd, field, dc, ti, r);
        else if (ext) propagate_all<WithExtendedGeometry<PolicyMixed>>(hd, field, dc, ti, r);
        else propagate_all<PolicyMixed>(hd, field, dc, ti, r);
    }
    return DTR_SUCCESS;
}

// The plain kernel's per-track propagation (propagate_track, the FP64
// kernel of dtr_propagate), extended geometry when the detector needs it as
// dtr_detector_create decides.  Host memory throughout.
__attribute__((visibility("default"))) int dtr_test_plain_propagate(const dtr_detector_desc* d,
                                                                    const dtr_field_desc* f,
                                                                    const dtr_prop_config* cfg,
                                                                    const dtr_track_batch* in,
                                                                    dtr_result_batch* out) {
    if (!d || !f || !cfg || !in || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->use_field_gradient || cfg->actors) return DTR_ERROR_UNSUPPORTED;
    const DevDetector det{d->volumes, d->surfaces, d->transforms, d->masks, d->materials,
                          d->grids, d->bf_surfaces, d->grid_bins, d->material_maps};
    const DevField field = host_field(f);
    const DevConfig dc = make_dev_config(cfg);
    const DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                       cfg->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                       out->n_trials, out->status};
    const bool ext = has_ext(*d);
    for (uint32_t gid = 0; gid < in->n; ++gid) {
        if (ext)
            propagate_track<true>(det, field, dc, in->bound_params ? nullptr : in->free_params, in->initial_cov,
                                  in->bound_params, in->start_surface, r, in->n, gid);
        else
            propagate_track<false>(det, field, dc, in->bound_params ? nullptr : in->free_params, in->initial_cov,
                                   in->bound_params, in->start_surface, r, in->n, gid);
    }
    return DTR_SUCCESS;
}

// The direct navigator: sequence [max_sequence][n] surfaces, n_sequence [n],
// direction DTR_DIRECTION_FORWARD | BACKWARD.  Host memory throughout.
__attribute__((visibility("default"))) int dtr_test_engine_propagate_direct(
    const dtr_detector_desc* d, const dtr_field_desc* f, const dtr_prop_config* cfg, const dtr_track_batch* in,
    const dtr_surface_sequence* seq, dtr_result_batch* out, int precision) {
    if (!d || !f || !cfg || !in || !out || !seq) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->use_field_gradient || seq->direction > DTR_DIRECTION_BACKWARD) return DTR_ERROR_UNSUPPORTED;
    HostDetector hd(*d);
    const DevField field = host_field(f);
    const DevConfig dc = make_dev_config(cfg);
    const TrackInputs ti{in->bound_params ? nullptr : in->free_params, in->initial_cov, in->bound_params,
                         in->start_surface, in->n};
    const DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                       cfg->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                       out->n_trials, out->status};
    const SeqView sq{seq->surfaces, seq->n_surfaces, seq->max_surfaces, d->n_surfaces,
                     seq->direction == DTR_DIRECTION_BACKWARD ? 1u : 0u};
    const bool ext = has_ext(*d), vol = has_volume_material(*d);
    if (precision == 0) {
        if (vol) propagate_direct_all<WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>>(hd, field, dc, ti, r, sq);
        else if (ext) propagate_direct_all<WithExtendedGeometry<PolicyFP64>>(hd, field, dc, ti, r, sq);
        else propagate_direct_all<PolicyFP64>(hd, field, dc, ti, r, sq);
    } else {
        if (vol) propagate_direct_all<WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>>(hd, field, dc, ti, r, sq);
        else if (ext) propagate_direct_all<WithExtendedGeometry<PolicyMixed>>(hd, field, dc, ti, r, sq);
        else propagate_direct_all<PolicyMixed>(hd, field, dc, ti, r, sq);
    }
    return DTR_SUCCESS;
}

// Both passes of the adjoint over host arrays (the tape is allocated here):
// chi2 [n], grad_transport [6][n] (or null: no Jacobian chain), grad [6][n]
// (or null: forward pass only), status [n].  The detector may hold the
// extended geometry but no volume material.
__attribute__((visibility("default"))) int dtr_test_adjoint(const dtr_detector_desc* d, const dtr_field_desc* f,
                                                            const dtr_prop_config* cfg,
                                                            const dtr_adjoint_config* a, uint32_t n,
                                                            const double* bound, const uint32_t* start,
                                                            const dtr_hits* hits, double* chi2,
                                                            double* grad_transport, double* grad, uint32_t* status) {
    if (!d || !f || !cfg || !a || !bound || !start || !chi2 || !status) return DTR_ERROR_INVALID_ARGUMENT;
    if (has_volume_material(*d) || cfg->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    if (a->max_steps == 0u || a->max_events == 0u || a->checkpoint_every == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    HostDetector hd(*d);
    const DevField field = host_field(f);
    DevConfig dc = make_dev_config(cfg);
    dc.max_surfaces = 0u;
    if (!grad_transport && !cfg->estimate_scattering_noise) dc.do_cov = 0u;
    std::vector<unsigned char> tape(tape_bytes(n, a->max_steps, a->max_events, a->checkpoint_every));
    AdjointIO io{};
    io.bound0 = bound;
    io.start_sf = start;
    io.n = n;
    if (hits) io.hits = HitsView{hits->max_hits, hits->surface, hits->loc, hits->n_hits};
    io.iv0 = 1.0 / (a->sigma_loc0 * a->sigma_loc0);
    io.iv1 = 1.0 / (a->sigma_loc1 * a->sigma_loc1);
    io.tape = tape_carve(tape.data(), n, a->max_steps, a->max_events, a->checkpoint_every);
    io.chi2 = chi2;
    io.grad_transport = grad_transport;
    io.status = status;
    const bool exact = a->model == DTR_ADJOINT_EXACT;
    const bool ext = has_ext(*d);
    if (a->precision == DTR_PRECISION_FP64) {
        if (ext) adjoint_all<WithExtendedGeometry<PolicyFP64>>(hd, field, dc, io, exact, grad);
        else adjoint_all<PolicyFP64>(hd, field, dc, io, exact, grad);
    } else {
        if (ext) adjoint_all<WithExtendedGeometry<PolicyMixed>>(hd, field, dc, io, exact, grad);
        else adjoint_all<PolicyMixed>(hd, field, dc, io, exact, grad);
    }
    return DTR_SUCCESS;
}

// One call of the resumable propagation over host arrays: state is the
// buffer of dtr_propagation_state_size(n) bytes (doubles, then words).
__attribute__((visibility("default"))) int dtr_test_engine_propagate_resumable(
    const dtr_detector_desc* d, const dtr_field_desc* f, const dtr_prop_config* cfg, const dtr_track_batch* in,
    dtr_result_batch* out, void* state, uint32_t max_steps, uint32_t resume, int precision) {
    if (!d || !f || !cfg || !in || !out || !state || max_steps == 0u) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->use_field_gradient) return DTR_ERROR_UNSUPPORTED;
    HostDetector hd(*d);
    const DevField field = host_field(f);
    const DevConfig dc = make_dev_config(cfg);
    const TrackInputs ti{in->bound_params ? nullptr : in->free_params, in->initial_cov, in->bound_params,
                         in->start_surface, in->n};
    const DevResults r{out->bound_params, out->bound_cov, out->free_params, out->path_length,
                       cfg->max_surfaces ? out->surfaces : nullptr, out->n_surfaces, out->n_steps,
                       out->n_trials, out->status};
    PauseView pv;
    pv.n = in->n;
    pv.d = static_cast<double*>(state);
    pv.w = reinterpret_cast<uint32_t*>(pv.d + static_cast<uint64_t>(kPauseD) * in->n);
    pv.max_steps = max_steps;
    pv.resume = resume;
    const bool ext = has_ext(*d), vol = has_volume_material(*d);
    if (precision == 0) {
        if (vol) propagate_resumable_all<WithVolumeMaterial<WithExtendedGeometry<PolicyFP64>>>(hd, field, dc, ti, r, pv);
        else if (ext) propagate_resumable_all<WithExtendedGeometry<PolicyFP64>>(hd, field, dc, ti, r, pv);
        else propagate_resumable_all<PolicyFP64>(hd, field, dc, ti, r, pv);
    } else {
        if (vol) propagate_resumable_all<WithVolumeMaterial<WithExtendedGeometry<PolicyMixed>>>(hd, field, dc, ti, r, pv);
        else if (ext) propagate_resumable_all<WithExtendedGeometry<PolicyMixed>>(hd, field, dc, ti, r, pv);
        else propagate_resumable_all<PolicyMixed>(hd, field, dc, ti, r, pv);
    }
    return DTR_SUCCESS;
}

}  // extern "C"

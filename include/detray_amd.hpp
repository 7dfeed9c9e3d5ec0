/* detray_amd.hpp: C++20 adapter over the C ABI (detray_amd.h) in the
 * reference's propagation vocabulary.
 *
 *   propagator<rk_stepper, caching_navigator, actor_chain<...>>::propagate
 *       core/include/detray/propagator/propagator.hpp:32-60 (the type),
 *       260-400 (propagate); here over a batch of tracks in device memory
 *       instead of one propagation state.
 *   actor_chain<pathlimit_aborter, parameter_updater<pointwise_material_
 *   interactor>, surface_sequencer, momentum_aborter, target_aborter,
 *   step_tracer>
 *       actor_chain.hpp:53-118; the actors of actors/aborters.hpp:44-175,
 *       parameter_updater.hpp, pointwise_material_interactor.hpp,
 *       surface_sequencer.hpp and step_tracer.hpp, each with its state.
 *   flatten(detector_payload, detector_grids_payload*, detector_homogeneous_
 *   material_payload*)
 *       a detector held by the reference (core/include/detray/core/
 *       detector.hpp:136-202) to device tables without files: the
 *       reference's own writers convert the detector to these payloads
 *       (io/include/detray/io/frontend/payloads.hpp:86-330;
 *       geometry_writer::convert, the surface-grid writer, homogeneous_
 *       material_writer::convert) and flatten() hands them to
 *       dtr_flatten_detector.
 *
 * Header only; link libdetray_amd.so.  Errors are thrown as
 * detray_amd::error carrying the dtr_status_code. */
#pragma once

#include "detray_amd.h"

#include <cstdint>
#include <limits>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace detray_amd {

// --------------------------------------------------------------- errors --
class error : public std::runtime_error {
   public:
    error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept { return code_; }

   private:
    int code_;
};

inline const char* status_name(int rc) {
    switch (rc) {
        case DTR_SUCCESS: return "DTR_SUCCESS";
        case DTR_ERROR_INVALID_ARGUMENT: return "DTR_ERROR_INVALID_ARGUMENT";
        case DTR_ERROR_HIP: return "DTR_ERROR_HIP";
        case DTR_ERROR_OUT_OF_MEMORY: return "DTR_ERROR_OUT_OF_MEMORY";
        case DTR_ERROR_NO_DEVICE: return "DTR_ERROR_NO_DEVICE";
        case DTR_ERROR_UNSUPPORTED: return "DTR_ERROR_UNSUPPORTED";
        case DTR_ERROR_IO: return "DTR_ERROR_IO";
        default: return "unknown status";
    }
}

inline void check(int rc, const char* what, const char* detail = nullptr) {
    if (rc == DTR_SUCCESS) return;
    std::string m = std::string(what) + ": " + status_name(rc);
    if (detail && *detail) m += " (" + std::string(detail) + ")";
    throw error(rc, m);
}

// ------------------------------------------------------------- detector --
/* Detector tables in host memory, owned (dtr_host_detector_free). */
class host_detector {
   public:
    host_detector() = default;
    host_detector(const host_detector&) = delete;
    host_detector& operator=(const host_detector&) = delete;
    host_detector(host_detector&& o) noexcept : d_(o.d_) { o.d_ = dtr_detector_desc{}; }
    host_detector& operator=(host_detector&& o) noexcept {
        if (this != &o) {
            release();
            d_ = o.d_;
            o.d_ = dtr_detector_desc{};
        }
        return *this;
    }
    ~host_detector() { release(); }

    const dtr_detector_desc& desc() const noexcept { return d_; }

    /* build_toy_detector (tests/include/detray/test/common/build_toy_detector.hpp) */
    static host_detector toy(uint32_t n_brl_layers = 4u, uint32_t n_edc_layers = 3u, bool use_material = true) {
        dtr_toy_config c{};
        c.n_brl_layers = n_brl_layers;
        c.n_edc_layers = n_edc_layers;
        c.use_material = use_material ? 1u : 0u;
        host_detector h;
        check(dtr_build_toy_detector(&c, &h.d_), "dtr_build_toy_detector");
        return h;
    }
    /* build_telescope_detector with the pilot ray along z */
    static host_detector telescope(const std::vector<double>& positions, double half_x, double half_y,
                                   double thickness = 0.15, double envelope = 60.0,
                                   const dtr_material_slab* volume_material = nullptr) {
        dtr_telescope_config c{};
        c.n_planes = static_cast<uint32_t>(positions.size());
        c.positions = positions.data();
        c.half_x = half_x;
        c.half_y = half_y;
        c.thickness = thickness;
        c.envelope = envelope;
        c.volume_material = volume_material;
        host_detector h;
        check(dtr_build_telescope_detector(&c, &h.d_), "dtr_build_telescope_detector");
        return h;
    }
    /* build_wire_chamber (build_wire_chamber.hpp) */
    static host_detector wire_chamber(const dtr_wire_chamber_config& c) {
        host_detector h;
        check(dtr_build_wire_chamber(&c, &h.d_), "dtr_build_wire_chamber");
        return h;
    }
    /* io::read_detector with the JSON readers (grids, material and material
     * maps may be null) */
    static host_detector from_json(const char* geometry, const char* grids = nullptr, const char* material = nullptr,
                                   const char* material_maps = nullptr) {
        host_detector h;
        char msg[512] = {0};
        check(dtr_read_detector_json(geometry, grids, material, material_maps, &h.d_, msg, sizeof msg),
              "dtr_read_detector_json", msg);
        return h;
    }
    /* tables of an in-memory detector view (see flatten()) */
    static host_detector from_view(const dtr_detector_view& v) {
        host_detector h;
        char msg[512] = {0};
        check(dtr_flatten_detector(&v, &h.d_, msg, sizeof msg), "dtr_flatten_detector", msg);
        return h;
    }

   private:
    void release() noexcept {
        if (d_.n_surfaces || d_.n_volumes) dtr_host_detector_free(&d_);
        d_ = dtr_detector_desc{};
    }
    dtr_detector_desc d_{};
};

/* The detector's device copy (dtr_detector_create / destroy). */
class device_detector {
   public:
    explicit device_detector(const host_detector& h) { check(dtr_detector_create(&h.desc(), &h_), "dtr_detector_create"); }
    explicit device_detector(const dtr_detector_desc& d) { check(dtr_detector_create(&d, &h_), "dtr_detector_create"); }
    device_detector(const device_detector&) = delete;
    device_detector& operator=(const device_detector&) = delete;
    device_detector(device_detector&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    ~device_detector() {
        if (h_) dtr_detector_destroy(h_);
    }
    const dtr_detector* get() const noexcept { return h_; }

   private:
    dtr_detector* h_ = nullptr;
};

/* The magnetic field's device copy: constant (const_field_t) or a sampled
 * grid read with trilinear interpolation (the covfie field of bfield.hpp). */
class device_field {
   public:
    explicit device_field(const dtr_field_desc& d) { check(dtr_field_create(&d, &h_), "dtr_field_create"); }
    static device_field constant(double bx, double by, double bz) {
        dtr_field_desc d{};
        d.kind = DTR_FIELD_CONSTANT;
        d.b[0] = bx;
        d.b[1] = by;
        d.b[2] = bz;
        return device_field(d);
    }
    device_field(const device_field&) = delete;
    device_field& operator=(const device_field&) = delete;
    device_field(device_field&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    ~device_field() {
        if (h_) dtr_field_destroy(h_);
    }
    const dtr_field* get() const noexcept { return h_; }

   private:
    dtr_field* h_ = nullptr;
};

// -------------------------------------------------------------- actors --
/* The actor types of the reference's chain (actors/aborters.hpp,
 * parameter_updater.hpp, surface_sequencer.hpp, step_tracer.hpp), each with
 * the state a caller passes to propagate(). */
namespace actor {

struct pathlimit_aborter {  // aborters.hpp:44-75
    struct state {
        float path_limit = 5.f * 1000.f;  // 5 m
    };
};

struct pointwise_material_interactor {  // pointwise_material_interactor.hpp
    struct state {
        bool do_energy_loss = true;
        bool do_multiple_scattering = true;
    };
};

template <class interactor_t>
struct parameter_updater {  // parameter_updater.hpp (transporter, interactor, resetter)
    static_assert(std::is_same_v<interactor_t, pointwise_material_interactor>,
                  "the parameter updater runs with the pointwise material interactor");
    using state = typename interactor_t::state;
};

struct surface_sequencer {  // surface_sequencer.hpp:78-106
    struct state {
        uint32_t capacity = 64u;
    };
};

struct momentum_aborter {  // aborters.hpp:80-143
    struct state {
        float min_p = 10.f * 1e-3f;   // 10 MeV
        float min_pT = 10.f * 1e-3f;  // 10 MeV
    };
};

struct target_aborter {  // aborters.hpp:146-175
    struct state {
        const uint32_t* target_surface = nullptr;  // device [n]: the target surface of each track
    };
};

struct step_tracer {  // step_tracer.hpp, an observer of the parameter updater
    struct state {
        dtr_step_trace trace{};  // device buffers and capacity
    };
};

}  // namespace actor

namespace detail {
template <class T, class... Ts>
inline constexpr bool contains_v = (std::is_same_v<T, Ts> || ...);

template <class T>
inline constexpr int chain_rank = -1;
template <>
inline constexpr int chain_rank<actor::pathlimit_aborter> = 0;
template <>
inline constexpr int chain_rank<actor::parameter_updater<actor::pointwise_material_interactor>> = 1;
template <>
inline constexpr int chain_rank<actor::step_tracer> = 2;
template <>
inline constexpr int chain_rank<actor::surface_sequencer> = 3;
template <>
inline constexpr int chain_rank<actor::momentum_aborter> = 4;
template <>
inline constexpr int chain_rank<actor::target_aborter> = 5;

template <class... Ts>
constexpr bool ordered() {
    constexpr int r[] = {chain_rank<Ts>..., 99};
    for (std::size_t i = 0; i + 1 < sizeof...(Ts); ++i)
        if (r[i] >= r[i + 1]) return false;
    return true;
}
}  // namespace detail

/* actor_chain<actors...> (actor_chain.hpp:53-118): the chain the kernels run
 * is pathlimit_aborter, parameter_updater<pointwise_material_interactor>
 * (with the step tracer observing it), surface_sequencer, momentum_aborter,
 * target_aborter; a chain names a subsequence of it, in that order. */
template <class... actors_t>
struct actor_chain {
    static_assert(((detail::chain_rank<actors_t> >= 0) && ...), "an actor the device chain does not have");
    static_assert(detail::ordered<actors_t...>(), "actors out of the device chain's order, or repeated");
    static_assert(!detail::contains_v<actor::step_tracer, actors_t...> ||
                      detail::contains_v<actor::parameter_updater<actor::pointwise_material_interactor>, actors_t...>,
                  "the step tracer observes the parameter updater");

    template <class A>
    static constexpr bool has = detail::contains_v<A, actors_t...>;

    using state_tuple = std::tuple<typename actors_t::state&...>;
};

// --------------------------------------------------- stepper, navigator --
/* rk_stepper (rk_stepper.hpp): the field is a runtime argument;
 * use_field_gradient and the energy-loss options are in the config. */
struct rk_stepper {};
/* caching_navigator (caching_navigator.hpp) */
struct caching_navigator {};
/* direct_navigator (direct_navigator.hpp:29-298): replays surface sequences */
struct direct_navigator {};

/* propagation::config (propagation_config.hpp) with the reference defaults */
inline dtr_prop_config default_config() {
    dtr_prop_config c;
    dtr_default_config(&c);
    return c;
}
/* launch options of the batched kernels (no counterpart in the reference) */
inline dtr_launch_config default_launch() {
    dtr_launch_config l;
    dtr_default_launch(&l);
    return l;
}

/* propagator<stepper_t, navigator_t, actor_chain_t> (propagator.hpp:32-60)
 * over batches: propagate() runs propagator::propagate (propagator.hpp:
 * 260-400) for every track of `tracks` (device memory, structure of arrays)
 * into `results`.  Without a launch configuration the plain FP64 kernel
 * runs (dtr_propagate); with one the batched kernels (dtr_propagate_ex). */
template <class stepper_t, class navigator_t, class actor_chain_t>
class propagator {
    static_assert(std::is_same_v<stepper_t, rk_stepper>, "the device stepper is the Runge-Kutta-Nystrom stepper");
    static_assert(std::is_same_v<navigator_t, caching_navigator> || std::is_same_v<navigator_t, direct_navigator>,
                  "caching_navigator or direct_navigator");

   public:
    using stepper_type = stepper_t;
    using navigator_type = navigator_t;
    using actor_chain_type = actor_chain_t;

    explicit propagator(const dtr_prop_config& cfg, std::optional<dtr_launch_config> launch = std::nullopt)
        : cfg_(cfg), launch_(launch) {}

    const dtr_prop_config& config() const noexcept { return cfg_; }

    /* The config the chain and its states make of the propagator's config. */
    template <class... states_t>
    dtr_prop_config chain_config(dtr_actor_states& as, const std::tuple<states_t&...>& states) const {
        dtr_prop_config c = cfg_;
        using AC = actor_chain_t;
        dtr_default_actors(&as);
        c.path_limit = std::numeric_limits<float>::max();  // no pathlimit_aborter: no limit
        c.use_material = 0u;
        c.max_surfaces = 0u;
        c.actors = &as;
        std::apply([&](auto&... s) { (apply_state(c, as, s), ...); }, states);
        (void)AC{};
        return c;
    }

    /* caching navigator: propagate every track of the batch */
    template <class... states_t>
    void propagate(const device_detector& det, const device_field& field, const dtr_track_batch& tracks,
                   dtr_result_batch& results, std::tuple<states_t&...> states, void* stream = nullptr) const
        requires std::is_same_v<navigator_t, caching_navigator>
    {
        static_assert(std::is_same_v<std::tuple<states_t&...>, typename actor_chain_t::state_tuple>,
                      "one state per actor of the chain, in the chain's order");
        dtr_actor_states as;
        const dtr_prop_config c = chain_config(as, states);
        if (launch_) {
            check(dtr_propagate_ex(det.get(), field.get(), &c, &tracks, &results, &*launch_, stream),
                  "dtr_propagate_ex");
        } else {
            check(dtr_propagate(det.get(), field.get(), &c, &tracks, &results, stream), "dtr_propagate");
        }
    }

    /* caching navigator with pause and resume (propagator.hpp:238-250,
     * 274-282): every track takes at most pause.max_steps steps per call;
     * the first call (pause.resume == 0) starts the tracks, later ones
     * (pause.resume == 1) continue those left paused (status
     * DTR_TRACK_PAUSED, is_paused()).  pause.data: caller-owned device
     * buffer of dtr_propagation_state_size(n) bytes.  Needs a launch with
     * the grid schedule and lane or gathered navigation. */
    template <class... states_t>
    void propagate(const device_detector& det, const device_field& field, const dtr_track_batch& tracks,
                   dtr_result_batch& results, std::tuple<states_t&...> states, const dtr_propagation_state& pause,
                   void* stream = nullptr) const
        requires std::is_same_v<navigator_t, caching_navigator>
    {
        static_assert(std::is_same_v<std::tuple<states_t&...>, typename actor_chain_t::state_tuple>,
                      "one state per actor of the chain, in the chain's order");
        if (!launch_) throw error(DTR_ERROR_INVALID_ARGUMENT, "pause and resume need a launch configuration");
        dtr_actor_states as;
        const dtr_prop_config c = chain_config(as, states);
        check(dtr_propagate_resumable(det.get(), field.get(), &c, &tracks, &results, &*launch_, &pause, stream),
              "dtr_propagate_resumable");
    }

    /* propagator::is_paused (propagator.hpp:238-243) of a track's status */
    static constexpr bool is_paused(uint8_t status) noexcept { return status == DTR_TRACK_PAUSED; }

    /* direct navigator: replay a surface sequence per track */
    template <class... states_t>
    void propagate(const device_detector& det, const device_field& field, const dtr_track_batch& tracks,
                   const dtr_surface_sequence& sequence, dtr_result_batch& results,
                   std::tuple<states_t&...> states, void* stream = nullptr) const
        requires std::is_same_v<navigator_t, direct_navigator>
    {
        static_assert(std::is_same_v<std::tuple<states_t&...>, typename actor_chain_t::state_tuple>,
                      "one state per actor of the chain, in the chain's order");
        dtr_actor_states as;
        const dtr_prop_config c = chain_config(as, states);
        dtr_launch_config l = launch_ ? *launch_ : default_launch();
        if (!launch_) {
            l.precision = DTR_PRECISION_FP64;
            l.schedule = DTR_SCHEDULE_GRID;
            l.block_size = 64u;
        }
        check(dtr_propagate_direct(det.get(), field.get(), &c, &tracks, &sequence, &results, &l, stream),
              "dtr_propagate_direct");
    }

   private:
    static void apply_state(dtr_prop_config& c, dtr_actor_states&, const actor::pathlimit_aborter::state& s) {
        c.path_limit = s.path_limit;
    }
    static void apply_state(dtr_prop_config& c, dtr_actor_states&, const actor::pointwise_material_interactor::state& s) {
        c.use_material = 1u;
        c.do_energy_loss = s.do_energy_loss ? 1u : 0u;
        c.do_multiple_scattering = s.do_multiple_scattering ? 1u : 0u;
    }
    static void apply_state(dtr_prop_config& c, dtr_actor_states&, const actor::surface_sequencer::state& s) {
        c.max_surfaces = s.capacity;
    }
    static void apply_state(dtr_prop_config&, dtr_actor_states& as, const actor::momentum_aborter::state& s) {
        as.actors |= DTR_ACTOR_MOMENTUM_ABORTER;
        as.min_p = s.min_p;
        as.min_pT = s.min_pT;
    }
    static void apply_state(dtr_prop_config&, dtr_actor_states& as, const actor::target_aborter::state& s) {
        as.actors |= DTR_ACTOR_TARGET_ABORTER;
        as.target_surface = s.target_surface;
    }
    static void apply_state(dtr_prop_config&, dtr_actor_states& as, const actor::step_tracer::state& s) {
        as.actors |= DTR_ACTOR_STEP_TRACER;
        as.trace = s.trace;
    }

    dtr_prop_config cfg_;
    std::optional<dtr_launch_config> launch_;
};

// ----------------------------------------------- detector-view flattener --
namespace detail {
template <class T>
uint64_t opt_index(const T& v, uint64_t none = ~0ull) {
    if constexpr (requires { v.has_value(); }) {
        return v.has_value() ? static_cast<uint64_t>(*v) : none;
    } else {
        return static_cast<uint64_t>(v);
    }
}
template <class T>
void copy_transform(const T& t, dtr_view_transform& out) {
    for (int k = 0; k < 3; ++k) out.tr[k] = static_cast<double>(t.tr[k]);
    for (int k = 0; k < 9; ++k) out.rot[k] = static_cast<double>(t.rot[k]);
}
}  // namespace detail

/* Arrays of a dtr_detector_view, filled from the reference's payloads. */
struct view_storage {
    std::vector<dtr_view_volume> volumes;
    std::vector<std::vector<dtr_view_surface>> surfaces;
    std::vector<std::vector<std::vector<dtr_view_mask>>> masks;
    std::vector<dtr_view_grid> grids;
    std::vector<std::vector<uint32_t>> loc_index;
    std::vector<std::vector<uint64_t>> offsets, content;
    std::vector<dtr_view_material> materials;
    dtr_detector_view view{};
};

struct no_payload {};

/* The payloads of a reference detector (io/include/detray/io/frontend/
 * payloads.hpp: detector_payload, detector_grids_payload,
 * detector_homogeneous_material_payload, with the field names there) as a
 * dtr_detector_view.  The storage owns the arrays the view points into. */
template <class geometry_payload_t, class grids_payload_t = no_payload, class material_payload_t = no_payload>
void make_view(view_storage& st, const geometry_payload_t& geo, const grids_payload_t* grids = nullptr,
               const material_payload_t* material = nullptr) {
    const std::size_t nv = geo.volumes.size();
    st.volumes.assign(nv, dtr_view_volume{});
    st.surfaces.assign(nv, {});
    st.masks.assign(nv, {});
    for (std::size_t v = 0; v < nv; ++v) {
        const auto& vp = geo.volumes[v];
        detail::copy_transform(vp.transform, st.volumes[v].transform);
        auto& sfs = st.surfaces[v];
        auto& msk = st.masks[v];
        sfs.assign(vp.surfaces.size(), dtr_view_surface{});
        msk.assign(vp.surfaces.size(), {});
        for (std::size_t s = 0; s < vp.surfaces.size(); ++s) {
            const auto& sp = vp.surfaces[s];
            dtr_view_surface& out = sfs[s];
            out.type = static_cast<uint32_t>(sp.type);
            out.index_in_coll = detail::opt_index(sp.index_in_coll);
            detail::copy_transform(sp.transform, out.transform);
            for (const auto& mp : sp.masks) {
                dtr_view_mask m{};
                m.shape = static_cast<uint32_t>(mp.shape);
                m.volume_link = static_cast<uint64_t>(mp.volume_link.link);
                if (mp.boundaries.size() > 8u) throw error(DTR_ERROR_INVALID_ARGUMENT, "more than 8 mask boundaries");
                m.n_boundaries = static_cast<uint32_t>(mp.boundaries.size());
                for (std::size_t k = 0; k < mp.boundaries.size(); ++k)
                    m.boundaries[k] = static_cast<double>(mp.boundaries[k]);
                msk[s].push_back(m);
            }
        }
    }
    // the inner vectors are final: point into them
    for (std::size_t v = 0; v < nv; ++v) {
        for (std::size_t s = 0; s < st.surfaces[v].size(); ++s) {
            st.surfaces[v][s].masks = st.masks[v][s].data();
            st.surfaces[v][s].n_masks = static_cast<uint32_t>(st.masks[v][s].size());
        }
        st.volumes[v].surfaces = st.surfaces[v].data();
        st.volumes[v].n_surfaces = static_cast<uint32_t>(st.surfaces[v].size());
    }
    st.grids.clear();
    st.loc_index.clear();
    st.offsets.clear();
    st.content.clear();
    if constexpr (!std::is_same_v<grids_payload_t, no_payload>) {
        if (grids) {
            for (const auto& [vol, list] : grids->grids) {
                (void)vol;
                for (const auto& gp : list) {
                    dtr_view_grid g{};
                    g.owner_link = static_cast<uint64_t>(gp.owner_link.link);
                    g.type = static_cast<uint32_t>(gp.grid_link.type);
                    if (gp.axes.size() > 2u) throw error(DTR_ERROR_UNSUPPORTED, "grids of more than 2 axes");
                    g.n_axes = static_cast<uint32_t>(gp.axes.size());
                    for (std::size_t a = 0; a < gp.axes.size(); ++a) {
                        const auto& ap = gp.axes[a];
                        g.axes[a].binning = static_cast<uint32_t>(ap.binning);
                        g.axes[a].bounds = static_cast<uint32_t>(ap.bounds);
                        g.axes[a].label = static_cast<uint32_t>(ap.label);
                        g.axes[a].bins = static_cast<uint32_t>(ap.bins);
                        g.axes[a].edges[0] = ap.edges.empty() ? 0.0 : static_cast<double>(ap.edges.front());
                        g.axes[a].edges[1] = ap.edges.empty() ? 0.0 : static_cast<double>(ap.edges.back());
                    }
                    std::vector<uint32_t> li;
                    std::vector<uint64_t> off{0u}, con;
                    for (const auto& bp : gp.bins) {
                        if (bp.loc_index.size() != 2u) throw error(DTR_ERROR_INVALID_ARGUMENT, "a bin needs 2 local indices");
                        li.push_back(static_cast<uint32_t>(bp.loc_index[0]));
                        li.push_back(static_cast<uint32_t>(bp.loc_index[1]));
                        for (const auto& c : bp.content) con.push_back(static_cast<uint64_t>(c));
                        off.push_back(con.size());
                    }
                    g.n_bins = static_cast<uint32_t>(gp.bins.size());
                    st.grids.push_back(g);
                    st.loc_index.push_back(std::move(li));
                    st.offsets.push_back(std::move(off));
                    st.content.push_back(std::move(con));
                }
            }
            for (std::size_t k = 0; k < st.grids.size(); ++k) {
                st.grids[k].loc_index = st.loc_index[k].data();
                st.grids[k].content_offset = st.offsets[k].data();
                st.grids[k].content = st.content[k].data();
            }
        }
    }
    st.materials.clear();
    if constexpr (!std::is_same_v<material_payload_t, no_payload>) {
        if (material) {
            for (const auto& mv : material->volumes) {
                for (const auto& sm : mv.surface_mat) {
                    dtr_view_material m{};
                    m.volume_link = static_cast<uint64_t>(mv.volume_link.link);
                    m.surface = static_cast<uint64_t>(sm.surface.link);
                    m.index_in_coll = detail::opt_index(sm.index_in_coll);
                    m.type = static_cast<uint32_t>(sm.type);
                    m.thickness = static_cast<double>(sm.thickness);
                    for (int k = 0; k < 7; ++k) m.params[k] = static_cast<double>(sm.mat.params[k]);
                    st.materials.push_back(m);
                }
            }
        }
    }
    st.view = dtr_detector_view{};
    st.view.n_volumes = static_cast<uint32_t>(st.volumes.size());
    st.view.volumes = st.volumes.data();
    st.view.n_grids = static_cast<uint32_t>(st.grids.size());
    st.view.grids = st.grids.data();
    st.view.n_materials = static_cast<uint32_t>(st.materials.size());
    st.view.materials = st.materials.data();
}

/* Detector tables of the reference's payloads: e.g.
 *   auto geo = detray::io::geometry_writer::convert(det, names);
 *   auto tables = detray_amd::flatten(geo, &grids, &material);
 * with the grid and material payloads of the surface-grid and homogeneous-
 * material writers (or null). */
template <class geometry_payload_t, class grids_payload_t = no_payload, class material_payload_t = no_payload>
host_detector flatten(const geometry_payload_t& geo, const grids_payload_t* grids = nullptr,
                      const material_payload_t* material = nullptr) {
    view_storage st;
    make_view(st, geo, grids, material);
    return host_detector::from_view(st.view);
}

}  // namespace detray_amd

// Compiled test of the C++20 adapter (include/detray_amd.hpp).  TEST
// INFRASTRUCTURE: run by tests/test_adapter.py (CPU) and
// tests/test_gpu_adapter.py (GPU) as a subprocess.
//
//   adapter_test cpu <dir>   builders, error codes, the actor-chain config,
//                            and the detector-view flattener: payload
//                            structs with the reference's field names and
//                            enum types (io/include/detray/io/frontend/
//                            payloads.hpp), filled from the reference-format
//                            JSON files the library writes (as the
//                            reference's JSON readers fill them), flattened
//                            in memory, equal to the tables read from the
//                            same files and to the builder's tables
//   adapter_test gpu         propagator<rk_stepper, caching_navigator,
//                            actor_chain<...>>::propagate on the toy
//                            detector (plain FP64 kernel; batched mixed
//                            kernel; with the momentum aborter), the direct
//                            navigator, against the oracle
//
// Prints one line per check and exits non-zero on the first failure.
#include "detray_amd.hpp"
#include "host/json.hpp"
#include "oracle.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

namespace {

int g_failed = 0;

void expect(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    std::fflush(stdout);
    if (!ok) {
        g_failed = 1;
        std::exit(1);
    }
}

template <class F>
int error_code_of(F&& f) {
    try {
        f();
    } catch (const detray_amd::error& e) {
        return e.code();
    }
    return DTR_SUCCESS;
}

// ------------------------------------------------------------------------
// The reference's payload types: same names, member names and types
// (io/include/detray/io/frontend/payloads.hpp:86-330, definitions.hpp,
// definitions/geometry.hpp, definitions/grid_axis.hpp).
// ------------------------------------------------------------------------
namespace ref_io {
enum class shape_id : unsigned int {
    annulus2 = 0, cuboid3 = 1, cylinder2 = 2, cylinder3 = 3, portal_cylinder2 = 4, rectangle2 = 5, ring2 = 6,
    trapezoid2 = 7, cell_wire = 8, straw_tube = 9, single1 = 10, single2 = 11, single3 = 12, unknown = 13
};
enum class material_id : unsigned int { slab = 6, rod = 7, unknown = 17 };
enum class accel_id : unsigned int { brute_force = 0, polar2_grid = 3, concentric_cylinder2_grid = 4, unknown = 7 };
enum class surface_id : unsigned int { e_portal = 0, e_sensitive = 1, e_passive = 2 };
namespace axis {
enum class binning : unsigned int { e_regular = 0, e_irregular = 1 };
enum class bounds : unsigned int { e_open = 0, e_closed = 1, e_circular = 2 };
enum class label : unsigned int { e_x = 0, e_y = 1, e_z = 2, e_r = 0, e_phi = 1 };
}  // namespace axis
using scalar = double;

struct single_link_payload {
    std::size_t link{std::numeric_limits<std::size_t>::max()};
};
template <typename type_id_t>
struct typed_link_payload {
    type_id_t type{type_id_t::unknown};
    std::size_t index{std::numeric_limits<std::size_t>::max()};
};
struct transform_payload {
    std::array<scalar, 3u> tr{};
    std::array<scalar, 9u> rot{};
};
struct mask_payload {
    shape_id shape{shape_id::unknown};
    single_link_payload volume_link{};
    std::vector<scalar> boundaries{};
};
struct surface_payload {
    std::optional<std::size_t> index_in_coll;
    transform_payload transform{};
    std::vector<mask_payload> masks{};
    std::optional<typed_link_payload<material_id>> material;
    std::uint64_t source{};
    std::optional<std::uint64_t> identifier;
    surface_id type{surface_id::e_sensitive};
};
struct volume_payload {
    std::string name{};
    transform_payload transform{};
    std::vector<surface_payload> surfaces{};
    single_link_payload index{};
};
struct detector_payload {
    std::vector<volume_payload> volumes = {};
};
struct axis_payload {
    axis::binning binning{axis::binning::e_regular};
    axis::bounds bounds{axis::bounds::e_closed};
    axis::label label{axis::label::e_r};
    std::size_t bins{0u};
    std::vector<scalar> edges{};
};
struct grid_bin_payload {
    std::vector<unsigned int> loc_index{};
    std::vector<std::size_t> content{};
};
struct grid_payload {
    single_link_payload owner_link{};
    typed_link_payload<accel_id> grid_link{};
    std::vector<axis_payload> axes{};
    std::vector<grid_bin_payload> bins{};
};
struct detector_grids_payload {
    std::map<std::size_t, std::vector<grid_payload>> grids = {};
};
struct material_param_payload {
    std::array<scalar, 7u> params{};
};
struct surface_material_payload {
    material_id type{material_id::unknown};
    std::optional<std::size_t> index_in_coll;
    single_link_payload surface{};
    scalar thickness{std::numeric_limits<scalar>::max()};
    material_param_payload mat{};
};
struct material_volume_payload {
    single_link_payload volume_link{};
    std::vector<surface_material_payload> surface_mat{};
};
struct detector_homogeneous_material_payload {
    std::vector<material_volume_payload> volumes{};
};
}  // namespace ref_io

// ------------------------------------------------------------------------
// JSON file -> payloads, as the reference's json_{geometry,grids,material}_io
// from_json functions fill them.
// ------------------------------------------------------------------------
using dtr_json::Value;

Value load(const std::string& path) {
    std::ifstream f(path);
    std::stringstream ss;
    ss << f.rdbuf();
    return dtr_json::parse(ss.str());
}
const Value& at(const Value& v, const char* k) {
    const Value* p = v.find(k);
    if (!p) throw std::runtime_error(std::string("missing ") + k);
    return *p;
}
std::size_t idx(const Value& v) { return static_cast<std::size_t>(v.u); }

ref_io::transform_payload transform_of(const Value& v) {
    ref_io::transform_payload t;
    for (int k = 0; k < 3; ++k) t.tr[k] = at(v, "translation").a[k].d;
    for (int k = 0; k < 9; ++k) t.rot[k] = at(v, "rotation").a[k].d;
    return t;
}

ref_io::detector_payload geometry_of(const Value& root) {
    ref_io::detector_payload det;
    for (const Value& v : at(at(root, "data"), "volumes").a) {
        ref_io::volume_payload vp;
        vp.name = v.find("name") ? v.find("name")->s : "";
        vp.transform = transform_of(at(v, "transform"));
        vp.index.link = idx(at(v, "index"));
        for (const Value& s : at(v, "surfaces").a) {
            ref_io::surface_payload sp;
            if (const Value* ic = s.find("index_in_coll")) sp.index_in_coll = idx(*ic);
            sp.transform = transform_of(at(s, "transform"));
            sp.type = static_cast<ref_io::surface_id>(at(s, "type").u);
            std::vector<const Value*> ms;
            if (const Value* m = s.find("mask")) ms.push_back(m);
            if (const Value* m = s.find("masks"))
                for (const Value& x : m->a) ms.push_back(&x);
            for (const Value* m : ms) {
                ref_io::mask_payload mp;
                mp.shape = static_cast<ref_io::shape_id>(at(*m, "shape").u);
                mp.volume_link.link = idx(at(*m, "volume_link"));
                for (const Value& b : at(*m, "boundaries").a) mp.boundaries.push_back(b.d);
                sp.masks.push_back(mp);
            }
            vp.surfaces.push_back(sp);
        }
        det.volumes.push_back(vp);
    }
    return det;
}

ref_io::detector_grids_payload grids_of(const Value& root) {
    ref_io::detector_grids_payload out;
    for (const Value& entry : at(at(root, "data"), "grids").a) {
        for (const Value& g : at(entry, "grid_data").a) {
            ref_io::grid_payload gp;
            gp.owner_link.link = idx(at(g, "owner_link"));
            gp.grid_link.type = static_cast<ref_io::accel_id>(at(at(g, "grid_link"), "type").u);
            for (const Value& a : at(g, "axes").a) {
                ref_io::axis_payload ap;
                ap.binning = static_cast<ref_io::axis::binning>(at(a, "binning").u);
                ap.bounds = static_cast<ref_io::axis::bounds>(at(a, "bounds").u);
                ap.label = static_cast<ref_io::axis::label>(at(a, "label").u);
                ap.bins = idx(at(a, "bins"));
                for (const Value& e : at(a, "edges").a) ap.edges.push_back(e.d);
                gp.axes.push_back(ap);
            }
            for (const Value& b : at(g, "bins").a) {
                ref_io::grid_bin_payload bp;
                for (const Value& i : at(b, "loc_index").a) bp.loc_index.push_back(static_cast<unsigned>(i.u));
                for (const Value& c : at(b, "content").a) bp.content.push_back(idx(c));
                gp.bins.push_back(bp);
            }
            out.grids[gp.owner_link.link].push_back(gp);
        }
    }
    return out;
}

ref_io::detector_homogeneous_material_payload material_of(const Value& root) {
    ref_io::detector_homogeneous_material_payload out;
    for (const Value& mv : at(at(root, "data"), "volumes").a) {
        ref_io::material_volume_payload vp;
        vp.volume_link.link = idx(at(mv, "volume_link"));
        if (const Value* list = mv.find("surface_material")) {
            for (const Value& e : list->a) {
                ref_io::surface_material_payload sp;
                sp.type = static_cast<ref_io::material_id>(at(e, "type").u);
                if (const Value* ic = e.find("index_in_coll")) sp.index_in_coll = idx(*ic);
                sp.surface.link = idx(at(e, "surface_idx"));
                sp.thickness = at(e, "thickness").d;
                for (int k = 0; k < 7; ++k) sp.mat.params[k] = at(at(e, "material"), "params").a[k].d;
                vp.surface_mat.push_back(sp);
            }
        }
        out.volumes.push_back(vp);
    }
    return out;
}

// ------------------------------------------------------------------------
template <class T>
bool same_bytes(const T* a, const T* b, uint32_t n) {
    return n == 0 || std::memcmp(a, b, sizeof(T) * n) == 0;
}
bool same_tables(const dtr_detector_desc& a, const dtr_detector_desc& b) {
    return a.n_volumes == b.n_volumes && a.n_surfaces == b.n_surfaces && a.n_transforms == b.n_transforms &&
           a.n_masks == b.n_masks && a.n_materials == b.n_materials && a.n_grids == b.n_grids &&
           a.n_bf == b.n_bf && a.n_bins == b.n_bins && same_bytes(a.volumes, b.volumes, a.n_volumes) &&
           same_bytes(a.surfaces, b.surfaces, a.n_surfaces) && same_bytes(a.transforms, b.transforms, a.n_transforms) &&
           same_bytes(a.masks, b.masks, a.n_masks) && same_bytes(a.materials, b.materials, a.n_materials) &&
           same_bytes(a.grids, b.grids, a.n_grids) && same_bytes(a.bf_surfaces, b.bf_surfaces, a.n_bf) &&
           same_bytes(a.grid_bins, b.grid_bins, a.n_bins);
}

// Equal detectors up to the sharing of masks and materials between
// surfaces (the builders share them, the reference's JSON readers give every
// surface its own): per surface the same identifier, placement, masks and
// material; the same volumes, grids and lists.
bool same_detector(const dtr_detector_desc& a, const dtr_detector_desc& b) {
    if (a.n_volumes != b.n_volumes || a.n_surfaces != b.n_surfaces || a.n_transforms != b.n_transforms ||
        a.n_grids != b.n_grids || a.n_bf != b.n_bf || a.n_bins != b.n_bins)
        return false;
    if (!same_bytes(a.volumes, b.volumes, a.n_volumes) || !same_bytes(a.transforms, b.transforms, a.n_transforms) ||
        !same_bytes(a.grids, b.grids, a.n_grids) || !same_bytes(a.bf_surfaces, b.bf_surfaces, a.n_bf) ||
        !same_bytes(a.grid_bins, b.grid_bins, a.n_bins))
        return false;
    for (uint32_t k = 0; k < a.n_surfaces; ++k) {
        const dtr_surface &x = a.surfaces[k], &y = b.surfaces[k];
        if (x.barcode != y.barcode || x.transform != y.transform || x.volume != y.volume || x.n_masks != y.n_masks ||
            x.sf_id != y.sf_id || x.shape != y.shape)
            return false;
        for (uint32_t m = 0; m < x.n_masks; ++m)
            if (std::memcmp(&a.masks[x.mask + m], &b.masks[y.mask + m], sizeof(dtr_mask)) != 0) return false;
        if ((x.material == 0xffffffffu) != (y.material == 0xffffffffu)) return false;
        if (x.material != 0xffffffffu &&
            std::memcmp(&a.materials[x.material], &b.materials[y.material], sizeof(dtr_material_slab)) != 0)
            return false;
    }
    return true;
}

int run_cpu(const std::string& dir) {
    using namespace detray_amd;
    // builders
    {
        host_detector toy = host_detector::toy(4u, 3u);
        expect(toy.desc().n_volumes == 22u && toy.desc().n_surfaces == 3230u && toy.desc().n_transforms == 3252u,
               "toy detector (4 barrel, 3 endcap layers): 22 volumes, 3230 surfaces, 3252 transforms");
        host_detector tel = host_detector::telescope({100.0, 200.0, 300.0}, 50.0, 50.0);
        expect(tel.desc().n_volumes == 1u && tel.desc().n_surfaces == 3u + 6u, "telescope: 3 planes and 6 portals");
        dtr_material_slab csi{};
        check(dtr_predefined_material("cesium_iodide_with_ded", 0.0, &csi), "dtr_predefined_material");
        host_detector telv = host_detector::telescope({100.0, 200.0}, 50.0, 50.0, 0.0, 60.0, &csi);
        expect(telv.desc().volumes[0].material == 1u && telv.desc().materials[0].has_density == 1u,
               "telescope with CsI volume material (density-effect data)");
        dtr_wire_chamber_config wc{10u, 1u, 500.0, 1000.0, 10.0, 0.05, 0.015};
        host_detector ch = host_detector::wire_chamber(wc);
        expect(ch.desc().n_volumes == 11u, "wire chamber: 11 volumes");
        host_detector moved = std::move(toy);
        expect(moved.desc().n_surfaces == 3230u && toy.desc().n_surfaces == 0u, "host_detector moves its tables");
    }
    // error codes
    expect(error_code_of([] { host_detector::telescope({}, 50.0, 50.0); }) == DTR_ERROR_INVALID_ARGUMENT,
           "telescope without planes: DTR_ERROR_INVALID_ARGUMENT");
    expect(error_code_of([] { host_detector::from_json("/nonexistent/geometry.json"); }) == DTR_ERROR_IO,
           "missing JSON file: DTR_ERROR_IO");
    {
        dtr_material_slab m{};
        expect(dtr_predefined_material("unobtainium", 0.0, &m) == DTR_ERROR_INVALID_ARGUMENT,
               "unknown material name: DTR_ERROR_INVALID_ARGUMENT");
    }
    {
        // a mask link past the last volume
        host_detector tel = host_detector::telescope({100.0}, 50.0, 50.0);
        std::vector<dtr_mask> masks(tel.desc().masks, tel.desc().masks + tel.desc().n_masks);
        masks[0].volume_link = 7u;
        dtr_detector_desc bad = tel.desc();
        bad.masks = masks.data();
        const int rc = error_code_of([&] { device_detector d(bad); });
        expect(rc == DTR_ERROR_INVALID_ARGUMENT, "inconsistent tables: DTR_ERROR_INVALID_ARGUMENT before any device call");
    }
    {
        int n = 0;
        const bool gpu = hipGetDeviceCount(&n) == hipSuccess && n > 0;
        host_detector tel = host_detector::telescope({100.0}, 50.0, 50.0);
        const int rc = error_code_of([&] { device_detector d(tel); });
        expect(gpu ? rc == DTR_SUCCESS : rc == DTR_ERROR_NO_DEVICE,
               gpu ? "device detector created" : "no device: DTR_ERROR_NO_DEVICE");
    }
    // actor chain -> config
    {
        using chain = actor_chain<actor::pathlimit_aborter, actor::parameter_updater<actor::pointwise_material_interactor>,
                                  actor::surface_sequencer, actor::momentum_aborter>;
        propagator<rk_stepper, caching_navigator, chain> p(default_config());
        actor::pathlimit_aborter::state pl{1234.f};
        actor::pointwise_material_interactor::state mi{true, false};
        actor::surface_sequencer::state sq{16u};
        actor::momentum_aborter::state ma{0.5f, 0.25f};
        dtr_actor_states as;
        const dtr_prop_config c = p.chain_config(as, std::tie(pl, mi, sq, ma));
        expect(c.path_limit == 1234.f && c.use_material == 1u && c.do_energy_loss == 1u &&
                   c.do_multiple_scattering == 0u && c.max_surfaces == 16u && c.actors == &as &&
                   as.actors == DTR_ACTOR_MOMENTUM_ABORTER && as.min_p == 0.5f && as.min_pT == 0.25f,
               "actor chain states -> propagation config and actor states");
        using bare = actor_chain<actor::surface_sequencer>;
        propagator<rk_stepper, caching_navigator, bare> q(default_config());
        dtr_actor_states bs;
        actor::surface_sequencer::state s2{8u};
        const dtr_prop_config d = q.chain_config(bs, std::tie(s2));
        expect(d.use_material == 0u && d.max_surfaces == 8u && bs.actors == 0u &&
                   d.path_limit == std::numeric_limits<float>::max(),
               "a chain without aborter and updater: no path limit, no transport");
    }
    // detector-view flattener: the reference-format files of the toy
    // detector (as the library writes them) -> payloads -> tables
    {
        host_detector toy = host_detector::toy(4u, 7u);
        const std::string g = dir + "/toy_geometry.json", s = dir + "/toy_surface_grids.json",
                          m = dir + "/toy_homogeneous_material.json";
        char msg[256];
        check(dtr_write_detector_json(&toy.desc(), "toy", nullptr, g.c_str(), s.c_str(), m.c_str(), nullptr, msg,
                                      sizeof msg),
              "dtr_write_detector_json", msg);
        const auto geo = geometry_of(load(g));
        const auto grids = grids_of(load(s));
        const auto mat = material_of(load(m));
        host_detector flat = flatten(geo, &grids, &mat);
        host_detector read = host_detector::from_json(g.c_str(), s.c_str(), m.c_str());
        expect(same_tables(flat.desc(), read.desc()), "flatten(payloads) equals the JSON reader's tables");
        expect(same_detector(flat.desc(), toy.desc()),
               "flatten(payloads) equals the toy builder's detector (surface by surface)");
        host_detector geo_only = flatten(geo);
        expect(geo_only.desc().n_grids == 0u && geo_only.desc().n_materials == 0u &&
                   geo_only.desc().n_surfaces == toy.desc().n_surfaces,
               "flatten(geometry only): no grids, no material");
        auto broken = geo;
        broken.volumes[3].surfaces[0].masks[0].shape = ref_io::shape_id::cuboid3;
        const int rc = error_code_of([&] { (void)flatten(broken); });
        expect(rc == DTR_ERROR_UNSUPPORTED, "a shape the tables cannot hold: DTR_ERROR_UNSUPPORTED");
    }
    return g_failed;
}

// ------------------------------------------------------------------------
// GPU: the adapter's propagate against the oracle
// ------------------------------------------------------------------------
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc");
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct Results {
    std::vector<double> bound, cov, free_p, path;
    std::vector<uint32_t> sf, n_sf, steps, trials;
    std::vector<uint8_t> status;
    Results(uint32_t n, uint32_t ms)
        : bound(6u * n), cov(21u * n), free_p(8u * n), path(n), sf(static_cast<size_t>(ms) * n, 0xffffffffu),
          n_sf(n), steps(n), trials(n), status(n) {}
    dtr_result_batch host_batch() {
        return dtr_result_batch{bound.data(), cov.data(), free_p.data(), path.data(), sf.data(),
                                n_sf.data(), steps.data(), trials.data(), status.data()};
    }
};

struct DeviceResults {
    uint32_t n, ms;
    DeviceBuffer bound, cov, free_p, path, sf, n_sf, steps, trials, status;
    DeviceResults(uint32_t n_, uint32_t ms_)
        : n(n_), ms(ms_), bound(48u * n_), cov(168u * n_), free_p(64u * n_), path(8u * n_),
          sf(4u * static_cast<size_t>(ms_) * n_), n_sf(4u * n_), steps(4u * n_), trials(4u * n_), status(n_) {}
    dtr_result_batch batch() const {
        return dtr_result_batch{bound.as<double>(), cov.as<double>(), free_p.as<double>(), path.as<double>(),
                                sf.as<uint32_t>(), n_sf.as<uint32_t>(), steps.as<uint32_t>(), trials.as<uint32_t>(),
                                status.as<uint8_t>()};
    }
    Results fetch() const {
        Results r(n, ms);
        auto cp = [](void* dst, const DeviceBuffer& b, size_t bytes) {
            if (hipMemcpy(dst, b.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
        };
        cp(r.bound.data(), bound, 48u * n);
        cp(r.cov.data(), cov, 168u * n);
        cp(r.free_p.data(), free_p, 64u * n);
        cp(r.path.data(), path, 8u * n);
        cp(r.sf.data(), sf, 4u * static_cast<size_t>(ms) * n);
        cp(r.n_sf.data(), n_sf, 4u * n);
        cp(r.steps.data(), steps, 4u * n);
        cp(r.trials.data(), trials, 4u * n);
        cp(r.status.data(), status, n);
        return r;
    }
};

double max_rel(const std::vector<double>& a, const std::vector<double>& b) {
    double m = 0.0;
    for (size_t k = 0; k < a.size(); ++k) m = std::max(m, std::fabs(a[k] - b[k]) / std::max(std::fabs(b[k]), 1.0));
    return m;
}

bool same_sequences(const Results& a, const Results& b, uint32_t n) {
    if (a.n_sf != b.n_sf || a.status != b.status) return false;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < a.n_sf[i]; ++k)
            if (a.sf[static_cast<size_t>(k) * n + i] != b.sf[static_cast<size_t>(k) * n + i]) return false;
    return true;
}

int run_gpu() {
    using namespace detray_amd;
    const uint32_t n = 4096u, ms = 64u;
    host_detector toy = host_detector::toy(4u, 7u);
    device_detector det(toy);
    const double T = 0.000299792458;  // 1 T in GeV / (e mm)
    dtr_field_desc fd{};
    fd.kind = DTR_FIELD_CONSTANT;
    fd.b[2] = 2.0 * T;
    device_field field(fd);

    dtr_track_gen_config gc{};
    gc.seed = 43u;
    gc.n = n;
    gc.kind = 1u;
    gc.eta_min = -2.5;
    gc.eta_max = 2.5;
    gc.pt_min = 1.0;
    gc.pt_max = 100.0;
    gc.sigma_xy = 0.01;
    gc.sigma_z = 50.0;
    std::vector<double> fp(8u * n);
    check(dtr_generate_tracks(&gc, fp.data()), "dtr_generate_tracks");
    DeviceBuffer dfp(fp.size() * sizeof(double));
    if (hipMemcpy(dfp.p, fp.data(), fp.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 2;

    dtr_prop_config cfg = default_config();
    cfg.search_window[0] = cfg.search_window[1] = 3u;
    cfg.estimate_scattering_noise = 0u;

    // the oracle with the reference's default chain
    const dtr_track_batch host_in{n, 0u, fp.data(), nullptr, nullptr, nullptr};
    Results ref(n, ms);
    dtr_result_batch rb = ref.host_batch();
    check(oracle_propagate(&toy.desc(), &fd, &cfg, &host_in, &rb, 0, nullptr), "oracle_propagate");

    using chain = actor_chain<actor::pathlimit_aborter, actor::parameter_updater<actor::pointwise_material_interactor>,
                              actor::surface_sequencer>;
    actor::pathlimit_aborter::state pl{cfg.path_limit};
    actor::pointwise_material_interactor::state mi{};
    actor::surface_sequencer::state sq{ms};
    const dtr_track_batch in{n, 0u, dfp.as<double>(), nullptr, nullptr, nullptr};
    {
        propagator<rk_stepper, caching_navigator, chain> p(cfg);
        DeviceResults out(n, ms);
        dtr_result_batch ob = out.batch();
        p.propagate(det, field, in, ob, std::tie(pl, mi, sq));
        (void)hipDeviceSynchronize();
        const Results got = out.fetch();
        expect(same_sequences(got, ref, n), "propagator (plain FP64 kernel): surface sequences equal the oracle's");
        const double e = max_rel(got.bound, ref.bound);
        expect(e < 1e-9, "propagator (plain FP64 kernel): bound parameters within 1e-9 (" + std::to_string(e) + ")");
    }
    {
        dtr_launch_config l = default_launch();
        l.precision = DTR_PRECISION_MIXED;
        l.schedule = DTR_SCHEDULE_GRID;
        l.navigation = DTR_NAVIGATION_GATHER;
        l.block_size = 64u;
        l.waves_per_simd = 2u;
        propagator<rk_stepper, caching_navigator, chain> p(cfg, l);
        DeviceResults out(n, ms);
        dtr_result_batch ob = out.batch();
        p.propagate(det, field, in, ob, std::tie(pl, mi, sq));
        (void)hipDeviceSynchronize();
        const Results got = out.fetch();
        expect(same_sequences(got, ref, n), "propagator (batched mixed kernel): surface sequences equal the oracle's");
        const double e = max_rel(got.bound, ref.bound);
        expect(e < 1e-4, "propagator (batched mixed kernel): bound parameters within 1e-4 (" + std::to_string(e) + ")");
    }
    {
        // pause and resume: budgets of 4 steps until no track is paused
        // equal one call of the same launch bit for bit
        dtr_launch_config l = default_launch();
        l.precision = DTR_PRECISION_MIXED;
        l.schedule = DTR_SCHEDULE_GRID;
        l.navigation = DTR_NAVIGATION_GATHER;
        l.block_size = 64u;
        propagator<rk_stepper, caching_navigator, chain> p(cfg, l);
        DeviceResults one(n, ms), part(n, ms);
        dtr_result_batch ob = one.batch(), pb = part.batch();
        p.propagate(det, field, in, ob, std::tie(pl, mi, sq));
        DeviceBuffer state(dtr_propagation_state_size(n));
        dtr_propagation_state ps{state.p, 4u, 0u};
        uint32_t calls = 0, first_paused = 0;
        for (; calls < 1000u; ++calls) {
            p.propagate(det, field, in, pb, std::tie(pl, mi, sq), ps);
            ps.resume = 1u;
            (void)hipDeviceSynchronize();
            std::vector<uint8_t> st(n);
            (void)hipMemcpy(st.data(), part.status.p, n, hipMemcpyDeviceToHost);
            uint32_t paused = 0;
            for (uint8_t v : st) paused += decltype(p)::is_paused(v) ? 1u : 0u;
            if (calls == 0u) first_paused = paused;
            if (paused == 0u) break;
        }
        (void)hipDeviceSynchronize();
        const Results a = one.fetch(), b = part.fetch();
        expect(first_paused > n / 2u, "pause: most tracks take more than 4 steps (" + std::to_string(first_paused) + ")");
        expect(calls >= 3u && a.bound == b.bound && a.cov == b.cov && a.free_p == b.free_p && a.path == b.path &&
                   a.steps == b.steps && a.trials == b.trials && same_sequences(a, b, n),
               "pause and resume: " + std::to_string(calls + 1u) + " calls equal one call bit for bit");
    }
    {
        // momentum aborter at 3 GeV: the oracle with the same actor states
        using mchain = actor_chain<actor::pathlimit_aborter, actor::parameter_updater<actor::pointwise_material_interactor>,
                                   actor::surface_sequencer, actor::momentum_aborter>;
        actor::momentum_aborter::state ma{3.f, 0.f};
        propagator<rk_stepper, caching_navigator, mchain> p(cfg);
        dtr_actor_states as;
        const dtr_prop_config oc = p.chain_config(as, std::tie(pl, mi, sq, ma));
        Results mref(n, ms);
        dtr_result_batch mb = mref.host_batch();
        check(oracle_propagate(&toy.desc(), &fd, &oc, &host_in, &mb, 0, nullptr), "oracle_propagate");
        DeviceResults out(n, ms);
        dtr_result_batch ob = out.batch();
        p.propagate(det, field, in, ob, std::tie(pl, mi, sq, ma));
        (void)hipDeviceSynchronize();
        const Results got = out.fetch();
        uint32_t stopped = 0;
        for (uint32_t i = 0; i < n; ++i) stopped += mref.status[i] == DTR_TRACK_MOMENTUM_LIMIT ? 1u : 0u;
        expect(stopped > n / 10u, "momentum aborter stops tracks below 3 GeV (" + std::to_string(stopped) + ")");
        expect(same_sequences(got, mref, n), "propagator with the momentum aborter: statuses and sequences equal the oracle's");
    }
    {
        // direct navigator replaying the oracle's sequences
        using dchain = chain;
        propagator<rk_stepper, direct_navigator, dchain> p(cfg);
        DeviceBuffer dsf(ref.sf.size() * 4u), dn(ref.n_sf.size() * 4u);
        (void)hipMemcpy(dsf.p, ref.sf.data(), ref.sf.size() * 4u, hipMemcpyHostToDevice);
        (void)hipMemcpy(dn.p, ref.n_sf.data(), ref.n_sf.size() * 4u, hipMemcpyHostToDevice);
        const dtr_surface_sequence seq{ms, 0u, dsf.as<uint32_t>(), dn.as<uint32_t>()};
        DeviceResults out(n, ms);
        dtr_result_batch ob = out.batch();
        p.propagate(det, field, in, seq, ob, std::tie(pl, mi, sq));
        (void)hipDeviceSynchronize();
        const Results got = out.fetch();
        Results dref(n, ms);
        dtr_result_batch db = dref.host_batch();
        check(oracle_propagate_direct(&toy.desc(), &fd, &cfg, &host_in, ref.sf.data(), ref.n_sf.data(), ms,
                                      DTR_DIRECTION_FORWARD, &db, 0),
              "oracle_propagate_direct");
        expect(same_sequences(got, dref, n), "direct navigator: sequences equal the oracle's");
        const double e = max_rel(got.bound, dref.bound);
        expect(e < 1e-9, "direct navigator: bound parameters within 1e-9 (" + std::to_string(e) + ")");
    }
    return g_failed;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s cpu <dir> | gpu\n", argv[0]);
        return 2;
    }
    try {
        if (std::strcmp(argv[1], "cpu") == 0 && argc >= 3) return run_cpu(argv[2]);
        if (std::strcmp(argv[1], "gpu") == 0) return run_gpu();
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: %s cpu <dir> | gpu\n", argv[0]);
    return 2;
}

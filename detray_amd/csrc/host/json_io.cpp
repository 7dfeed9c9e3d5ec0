// Detector tables from and to detray's JSON detector files: geometry,
// surface grids and homogeneous material.
//
// Format: the payloads of io/include/detray/io/frontend/payloads.hpp as
// serialised by io/include/detray/io/json/detail/json_{common,algebra,
// geometry,grids,material}_io.hpp; each file is {"header": ..., "data": ...}
// (io/include/detray/io/json/json_converter.hpp).
//
// Reading follows what the reference's readers and builders make of a file
// (io/include/detray/io/backend/geometry_reader.hpp, surface_grid_reader.hpp,
// detail/grid_reader.hpp, homogeneous_material_reader.hpp with
// core/include/detray/builders/volume_builder.hpp, surface_factory.hpp,
// grid_builder.hpp, homogeneous_material_factory.hpp,
// homogeneous_material_builder.hpp):
//  * volumes are built in file order (the "index" field is not used);
//  * a surface with "index_in_coll" is placed at that position of its
//    volume; without it, surfaces are grouped into factories, portals first,
//    then sensitive and passive surfaces, each group ordered by shape id
//    (cylinder2 counts as portal_cylinder2), file order within a group;
//  * the placement transform of a volume is followed by the transforms of
//    its surfaces; the surface identifier is rebuilt from volume, type,
//    index and transform;
//  * a cylinder mask takes the z translation of its transform into its
//    z bounds and the transform becomes the identity;
//  * grid bins hold volume-local surface indices; a volume with a grid keeps
//    only its portals and passive surfaces in the brute-force list;
//  * grid bins keep all their entries in file order (the largest occupancy
//    becomes the grid's bin capacity);
//  * homogeneous material: entry j of a volume's list goes to the surface
//    at (smallest surface_idx + j), one slab or rod per surface, placed at
//    "index_in_coll" in the volume's slab or rod list; volumes append their
//    slabs, then their rods, in volume order.
// What the tables cannot hold (other shapes, irregular or open grid axes,
// several grids per volume) is
// rejected with DTR_ERROR_UNSUPPORTED; inconsistent files (gaps or
// duplicates in index_in_coll, out-of-range links) with
// DTR_ERROR_INVALID_ARGUMENT, where the reference would leave holes or
// overwrite entries.
//
// Writing produces the files the reference's writers produce for the same
// detector (geometry_writer.hpp, surface_grid_writer.hpp with
// detail/grid_writer.hpp, homogeneous_material_writer.hpp): every surface
// carries its index_in_coll, every material surface its own slab.

#include "detray_amd.h"
#include "host/json.hpp"
#include "host/tables.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

using dtr_json::Value;
using namespace dtr_host;

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void bad(const std::string& m) { throw IoError{DTR_ERROR_INVALID_ARGUMENT, m}; }
[[noreturn]] void unsupported(const std::string& m) { throw IoError{DTR_ERROR_UNSUPPORTED, m}; }

// io::shape_id (io/include/detray/io/frontend/definitions.hpp)
constexpr unsigned kIoAnnulus2 = 0u, kIoCylinder2 = 2u, kIoPortalCylinder2 = 4u, kIoRectangle2 = 5u, kIoRing2 = 6u,
                   kIoTrapezoid2 = 7u, kIoDriftCell = 8u, kIoStrawTube = 9u, kIoShapes = 13u;
const char* const kIoShapeNames[kIoShapes] = {
    "annulus2", "cuboid3", "cylinder2", "cylinder3", "portal_cylinder2", "rectangle2", "ring2",
    "trapezoid2", "drift_cell", "straw_tube", "single1", "single2", "single3"};
// io::material_id
constexpr unsigned kIoSlab = 6u, kIoRod = 7u;
// io::material_id of the 2D material maps (definitions.hpp:56-65)
constexpr unsigned kIoPolar2Map = 0u, kIoCart2Map = 1u, kIoConcCyl2Map = 3u, kIoCyl2Map = 4u;
// io::accel_id
constexpr unsigned kIoPolar2Grid = 3u, kIoConcCyl2Grid = 4u, kIoAccels = 7u;
// axis::bounds, axis::binning, axis::label (core/include/detray/definitions/grid_axis.hpp)
constexpr unsigned kClosed = 1u, kCircular = 2u, kRegular = 0u;
// volume_id (core/include/detray/definitions/geometry.hpp)
constexpr unsigned kVolCylinder = 0u, kVolCuboid = 4u;

// ---------------------------------------------------------------------------
// Field access
// ---------------------------------------------------------------------------
const Value& need(const Value& obj, const char* key, const std::string& where) {
    const Value* v = obj.find(key);
    if (!v) bad(where + ": missing \"" + key + "\"");
    return *v;
}

uint64_t as_index(const Value& v, const std::string& where) {
    if (v.kind != Value::UInt) bad(where + ": expected a non-negative integer");
    return v.u;
}

double as_number(const Value& v, const std::string& where) {
    if (!v.is_number()) bad(where + ": expected a number");
    return v.d;
}

const std::vector<Value>& as_array(const Value& v, const std::string& where) {
    if (v.kind != Value::Array) bad(where + ": expected an array");
    return v.a;
}

void numbers(const Value& v, size_t n, double* out, const std::string& where) {
    const auto& a = as_array(v, where);
    if (a.size() != n) bad(where + ": expected " + std::to_string(n) + " numbers");
    for (size_t k = 0; k < n; ++k) out[k] = as_number(a[k], where);
}

// transform_payload: translation, rotation as the x, y, z axes (column-major)
dtr_transform read_transform(const Value& t, const std::string& where) {
    double tr[3], rot[9];
    numbers(need(t, "translation", where), 3, tr, where + ".translation");
    numbers(need(t, "rotation", where), 9, rot, where + ".rotation");
    dtr_transform out{};
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) out.r[row * 3 + col] = rot[col * 3 + row];
    for (int k = 0; k < 3; ++k) out.t[k] = tr[k];
    return out;
}

Value write_transform(const dtr_transform& t) {
    Value v = Value::object();
    Value tr = Value::array();
    for (int k = 0; k < 3; ++k) tr.push(Value::number(t.t[k]));
    Value rot = Value::array();
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) rot.push(Value::number(t.r[row * 3 + col]));
    v.set("translation", std::move(tr));
    v.set("rotation", std::move(rot));
    return v;
}

dtr_transform identity() {
    dtr_transform t{};
    t.r[0] = t.r[4] = t.r[8] = 1.0;
    return t;
}

Value load(const char* path, const char* what) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) throw IoError{DTR_ERROR_IO, std::string("cannot open ") + what + " file " + path};
    std::stringstream ss;
    ss << f.rdbuf();
    if (f.bad()) throw IoError{DTR_ERROR_IO, std::string("cannot read ") + what + " file " + path};
    try {
        return dtr_json::parse(ss.str());
    } catch (const dtr_json::Error& e) {
        throw IoError{DTR_ERROR_IO, std::string(what) + " file " + path + ": " + e.what()};
    }
}

void save(const char* path, const Value& v, const char* what) {
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f) throw IoError{DTR_ERROR_IO, std::string("cannot create ") + what + " file " + path};
    const std::string text = dtr_json::dump(v);
    f.write(text.data(), static_cast<std::streamsize>(text.size()));
    if (!f) throw IoError{DTR_ERROR_IO, std::string("cannot write ") + what + " file " + path};
}

// The "data" section; null (what an empty payload serialises to) or absent
// means no entries.
const Value* data_of(const Value& root, const char* what) {
    if (root.kind != Value::Object) bad(std::string(what) + " file: not a JSON object");
    const Value* d = root.find("data");
    if (!d || d->kind == Value::Null) return nullptr;
    if (d->kind != Value::Object) bad(std::string(what) + " file: \"data\" is not an object");
    return d;
}

// ---------------------------------------------------------------------------
// Reading: files and in-memory views become payloads, payloads tables
// ---------------------------------------------------------------------------
// Payloads (io/include/detray/io/frontend/payloads.hpp:86-330), the fields
// the tables are built from.  Files are parsed into them (parse_*), in-memory
// views copied into them (from_view); ingest_* then does what the reference's
// readers and builders make of them.
struct MaskP {
    uint64_t shape = 0;
    uint64_t volume_link = 0;
    std::vector<double> boundaries;
};
struct SurfaceP {
    uint64_t type = 0;
    dtr_transform trf{};
    std::vector<MaskP> masks;
    bool has_index = false;
    uint64_t index = 0;
};
struct VolumeP {
    dtr_transform placement{};
    std::vector<SurfaceP> surfaces;
};
struct AxisP {
    uint64_t binning = 0, bounds = 0, bins = 0;
    std::vector<double> edges;
};
struct BinP {
    std::vector<uint64_t> loc_index;
    std::vector<uint64_t> content;
};
struct GridP {
    uint64_t owner = 0, type = 0;
    std::vector<AxisP> axes;
    std::vector<BinP> bins;
};
struct MaterialP {
    uint64_t type = 0, surface = 0;
    bool has_index = false;
    uint64_t index = 0;
    double thickness = 0.0;
    double params[7] = {};
};
struct MaterialVolumeP {
    uint64_t volume = 0;
    std::vector<MaterialP> entries;
};
// detector_grids_payload<surface_material_payload, material_id>
// (payloads.hpp, json_grids_io.hpp:120-157): one slab per bin
struct MapP {
    uint64_t volume = 0, owner = 0, type = 0;
    std::vector<AxisP> axes;
    std::vector<std::vector<uint64_t>> loc_index;
    std::vector<std::vector<MaterialP>> content;
};
struct DetectorP {
    std::vector<VolumeP> volumes;
    std::vector<GridP> grids;
    std::vector<MaterialVolumeP> materials;
    std::vector<MapP> maps;
};

// --- files -> payloads --------------------------------------------------------
SurfaceP parse_surface(const Value& s, const std::string& where) {
    SurfaceP out;
    out.type = as_index(need(s, "type", where), where + ".type");
    out.trf = read_transform(need(s, "transform", where), where + ".transform");
    std::vector<const Value*> masks;
    if (const Value* m = s.find("mask")) {
        masks.push_back(m);  // single-mask files
    } else if (const Value* ms = s.find("masks")) {
        for (const Value& m : as_array(*ms, where + ".masks")) masks.push_back(&m);
    }
    for (size_t k = 0; k < masks.size(); ++k) {
        const std::string mw = where + ".masks[" + std::to_string(k) + "]";
        MaskP m;
        m.shape = as_index(need(*masks[k], "shape", mw), mw + ".shape");
        for (const Value& b : as_array(need(*masks[k], "boundaries", mw), mw + ".boundaries"))
            m.boundaries.push_back(as_number(b, mw + ".boundaries"));
        m.volume_link = as_index(need(*masks[k], "volume_link", mw), mw + ".volume_link");
        out.masks.push_back(std::move(m));
    }
    if (const Value* ic = s.find("index_in_coll")) {
        out.has_index = true;
        out.index = as_index(*ic, where + ".index_in_coll");
    }
    return out;
}

void parse_geometry(const Value& root, DetectorP& P) {
    const Value* data = data_of(root, "geometry");
    const Value* jv = data ? data->find("volumes") : nullptr;
    if (!jv) return;
    for (size_t vi = 0; vi < as_array(*jv, "volumes").size(); ++vi) {
        const Value& v = jv->a[vi];
        const std::string where = "volume " + std::to_string(vi);
        VolumeP V;
        V.placement = read_transform(need(v, "transform", where), where + ".transform");
        const auto& js = as_array(need(v, "surfaces", where), where + ".surfaces");
        for (size_t k = 0; k < js.size(); ++k)
            V.surfaces.push_back(parse_surface(js[k], where + " surface " + std::to_string(k)));
        P.volumes.push_back(std::move(V));
    }
}

void parse_grids(const Value& root, DetectorP& P) {
    const Value* data = data_of(root, "grids");
    const Value* jg = data ? data->find("grids") : nullptr;
    if (!jg) return;
    for (const Value& entry : as_array(*jg, "grids")) {
        const auto& list = as_array(need(entry, "grid_data", "grids entry"), "grid_data");
        for (const Value& g : list) {
            GridP G;
            G.owner = as_index(need(g, "owner_link", "grid"), "grid.owner_link");
            const std::string where = "grid of volume " + std::to_string(G.owner);
            const Value& link = need(g, "grid_link", where);
            G.type = as_index(need(link, "type", where + ".grid_link"), where + ".grid_link.type");
            const auto& axes = as_array(need(g, "axes", where), where + ".axes");
            for (size_t k = 0; k < axes.size(); ++k) {
                const std::string aw = where + ".axes[" + std::to_string(k) + "]";
                AxisP A;
                A.binning = as_index(need(axes[k], "binning", aw), aw + ".binning");
                A.bounds = as_index(need(axes[k], "bounds", aw), aw + ".bounds");
                A.bins = as_index(need(axes[k], "bins", aw), aw + ".bins");
                for (const Value& e : as_array(need(axes[k], "edges", aw), aw + ".edges"))
                    A.edges.push_back(as_number(e, aw + ".edges"));
                G.axes.push_back(std::move(A));
            }
            for (const Value& bin : as_array(need(g, "bins", where), where + ".bins")) {
                BinP B;
                for (const Value& i : as_array(need(bin, "loc_index", where), where + ".loc_index"))
                    B.loc_index.push_back(as_index(i, where + ".loc_index"));
                for (const Value& c : as_array(need(bin, "content", where), where + ".content"))
                    B.content.push_back(as_index(c, where + ".content"));
                G.bins.push_back(std::move(B));
            }
            P.grids.push_back(std::move(G));
        }
    }
}

// material_slab_payload / surface_material_payload (json_material_io.hpp)
MaterialP parse_material_entry(const Value& e, const std::string& ew, bool with_surface) {
    MaterialP m;
    m.type = as_index(need(e, "type", ew), ew + ".type");
    numbers(need(need(e, "material", ew), "params", ew), 7, m.params, ew + ".material.params");
    m.thickness = as_number(need(e, "thickness", ew), ew + ".thickness");
    if (with_surface) m.surface = as_index(need(e, "surface_idx", ew), ew + ".surface_idx");
    if (const Value* ic = e.find("index_in_coll")) {
        m.has_index = true;
        m.index = as_index(*ic, ew + ".index_in_coll");
    }
    return m;
}

// material_map_reader.hpp:51-185 over json_grids_io.hpp:120-157
void parse_material_maps(const Value& root, DetectorP& P) {
    const Value* data = data_of(root, "material maps");
    const Value* jg = data ? data->find("grids") : nullptr;
    if (!jg) return;
    for (const Value& entry : as_array(*jg, "material maps")) {
        const uint64_t vol = as_index(need(entry, "volume_link", "material maps entry"), "material maps volume_link");
        const auto& list = as_array(need(entry, "grid_data", "material maps entry"), "grid_data");
        for (size_t q = 0; q < list.size(); ++q) {
            const Value& g = list[q];
            MapP M;
            M.volume = vol;
            const std::string where = "material map " + std::to_string(q) + " of volume " + std::to_string(vol);
            M.owner = as_index(need(g, "owner_link", where), where + ".owner_link");
            M.type = as_index(need(need(g, "grid_link", where), "type", where + ".grid_link"), where + ".grid_link.type");
            const auto& axes = as_array(need(g, "axes", where), where + ".axes");
            for (size_t k = 0; k < axes.size(); ++k) {
                const std::string aw = where + ".axes[" + std::to_string(k) + "]";
                AxisP A;
                A.binning = as_index(need(axes[k], "binning", aw), aw + ".binning");
                A.bounds = as_index(need(axes[k], "bounds", aw), aw + ".bounds");
                A.bins = as_index(need(axes[k], "bins", aw), aw + ".bins");
                for (const Value& e : as_array(need(axes[k], "edges", aw), aw + ".edges"))
                    A.edges.push_back(as_number(e, aw + ".edges"));
                M.axes.push_back(std::move(A));
            }
            const auto& bins = as_array(need(g, "bins", where), where + ".bins");
            for (size_t b = 0; b < bins.size(); ++b) {
                const std::string bw = where + ".bins[" + std::to_string(b) + "]";
                std::vector<uint64_t> li;
                for (const Value& i : as_array(need(bins[b], "loc_index", bw), bw + ".loc_index"))
                    li.push_back(as_index(i, bw + ".loc_index"));
                std::vector<MaterialP> content;
                const auto& cs = as_array(need(bins[b], "content", bw), bw + ".content");
                for (size_t k = 0; k < cs.size(); ++k)
                    content.push_back(parse_material_entry(cs[k], bw + ".content[" + std::to_string(k) + "]", false));
                M.loc_index.push_back(std::move(li));
                M.content.push_back(std::move(content));
            }
            P.maps.push_back(std::move(M));
        }
    }
}

void parse_material(const Value& root, DetectorP& P) {
    const Value* data = data_of(root, "material");
    const Value* jv = data ? data->find("volumes") : nullptr;
    if (!jv) return;
    for (const Value& mv : as_array(*jv, "material volumes")) {
        MaterialVolumeP M;
        M.volume = as_index(need(mv, "volume_link", "material volume"), "material volume_link");
        const std::string where = "material of volume " + std::to_string(M.volume);
        const Value* list = mv.find("surface_material");
        if (!list) list = mv.find("material_slabs");  // older files
        if (list) {
            const auto& entries = as_array(*list, where);
            for (size_t k = 0; k < entries.size(); ++k)
                M.entries.push_back(parse_material_entry(entries[k], where + " entry " + std::to_string(k), true));
        }
        P.materials.push_back(std::move(M));
    }
}

// --- in-memory view -> payloads -------------------------------------------------
constexpr uint64_t kNotGiven = ~0ull;

dtr_transform view_transform(const dtr_view_transform& t) {
    dtr_transform out{};
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) out.r[row * 3 + col] = t.rot[col * 3 + row];
    for (int k = 0; k < 3; ++k) out.t[k] = t.tr[k];
    return out;
}

void from_view(const dtr_detector_view& v, DetectorP& P) {
    if (v.n_volumes && !v.volumes) bad("view: volumes missing");
    if (v.n_grids && !v.grids) bad("view: grids missing");
    if (v.n_materials && !v.materials) bad("view: materials missing");
    for (uint32_t vi = 0; vi < v.n_volumes; ++vi) {
        const dtr_view_volume& vv = v.volumes[vi];
        const std::string where = "volume " + std::to_string(vi);
        if (vv.n_surfaces && !vv.surfaces) bad(where + ": surfaces missing");
        VolumeP V;
        V.placement = view_transform(vv.transform);
        for (uint32_t k = 0; k < vv.n_surfaces; ++k) {
            const dtr_view_surface& vs = vv.surfaces[k];
            const std::string sw = where + " surface " + std::to_string(k);
            if (vs.n_masks && !vs.masks) bad(sw + ": masks missing");
            SurfaceP S;
            S.type = vs.type;
            S.trf = view_transform(vs.transform);
            for (uint32_t m = 0; m < vs.n_masks; ++m) {
                const dtr_view_mask& vm = vs.masks[m];
                if (vm.n_boundaries > 8u) bad(sw + ": more than 8 mask boundaries");
                MaskP M;
                M.shape = vm.shape;
                M.volume_link = vm.volume_link;
                M.boundaries.assign(vm.boundaries, vm.boundaries + vm.n_boundaries);
                S.masks.push_back(std::move(M));
            }
            S.has_index = vs.index_in_coll != kNotGiven;
            S.index = vs.index_in_coll;
            V.surfaces.push_back(std::move(S));
        }
        P.volumes.push_back(std::move(V));
    }
    for (uint32_t gi = 0; gi < v.n_grids; ++gi) {
        const dtr_view_grid& vg = v.grids[gi];
        const std::string where = "grid of volume " + std::to_string(vg.owner_link);
        if (vg.n_axes > 2u) bad(where + ": more than 2 axes");
        if (vg.n_bins && (!vg.loc_index || !vg.content_offset)) bad(where + ": bins missing");
        GridP G;
        G.owner = vg.owner_link;
        G.type = vg.type;
        for (uint32_t k = 0; k < vg.n_axes; ++k) {
            AxisP A;
            A.binning = vg.axes[k].binning;
            A.bounds = vg.axes[k].bounds;
            A.bins = vg.axes[k].bins;
            A.edges = {vg.axes[k].edges[0], vg.axes[k].edges[1]};
            G.axes.push_back(std::move(A));
        }
        for (uint32_t b = 0; b < vg.n_bins; ++b) {
            BinP B;
            B.loc_index = {vg.loc_index[2u * b], vg.loc_index[2u * b + 1u]};
            const uint64_t lo = vg.content_offset[b], hi = vg.content_offset[b + 1u];
            if (hi < lo || (hi > lo && !vg.content)) bad(where + ": bad content offsets");
            B.content.assign(vg.content + lo, vg.content + hi);
            G.bins.push_back(std::move(B));
        }
        P.grids.push_back(std::move(G));
    }
    // material entries grouped by volume in first-seen order
    for (uint32_t k = 0; k < v.n_materials; ++k) {
        const dtr_view_material& vm = v.materials[k];
        MaterialVolumeP* M = nullptr;
        for (auto& mv : P.materials)
            if (mv.volume == vm.volume_link) M = &mv;
        if (!M) {
            P.materials.emplace_back();
            M = &P.materials.back();
            M->volume = vm.volume_link;
        }
        MaterialP m;
        m.type = vm.type;
        m.surface = vm.surface;
        m.has_index = vm.index_in_coll != kNotGiven;
        m.index = vm.index_in_coll;
        m.thickness = vm.thickness;
        for (int i = 0; i < 7; ++i) m.params[i] = vm.params[i];
        M->entries.push_back(m);
    }
}

// --- payloads -> volumes ----------------------------------------------------------
struct SurfaceIn {
    uint8_t sf_id = 0;
    uint8_t shape = 0;     // dtr_mask_shape
    unsigned io_key = 0;   // factory key: io shape id, cylinder2 -> portal_cylinder2
    dtr_transform trf{};
    std::vector<std::array<double, 7>> bounds;
    std::vector<uint64_t> links;  // as in the file: a volume index or the world exit 2^64-1
    bool has_index = false;
    uint64_t index = 0;
};

struct VolumeIn {
    dtr_transform placement{};
    std::vector<SurfaceIn> surfaces;  // final volume-local order
    bool has_grid = false;
    dtr_grid grid{};
    std::vector<std::vector<uint32_t>> bins;  // volume-local surface indices per bin
    // per surface: kind and position in that kind's list, kind -1 none
    std::vector<std::pair<int, size_t>> mat_of;
    std::vector<dtr_material_slab> mats[2];  // slabs, rods (separate index_in_coll spaces)
    std::vector<char> mat_set[2];
    // per surface: index into maps, -1 none
    std::vector<int> map_of;
    std::vector<std::pair<dtr_material_map, std::vector<dtr_material_slab>>> maps;
};

SurfaceIn ingest_surface(const SurfaceP& s, const std::string& where) {
    SurfaceIn out;
    if (s.type > 2u) bad(where + ": surface type " + std::to_string(s.type) + " is not portal/sensitive/passive");
    out.sf_id = static_cast<uint8_t>(s.type);
    out.trf = s.trf;
    if (s.masks.empty()) bad(where + ": surface without mask");
    if (s.masks.size() > 255u) unsupported(where + ": more than 255 masks on one surface");
    // the factory is chosen by the first mask's shape; all masks take it
    const uint64_t sid = s.masks[0].shape;
    if (sid >= kIoShapes) bad(where + ": unknown mask shape id " + std::to_string(sid));
    size_t n_bounds = 0;
    switch (sid) {
        case kIoAnnulus2:
            out.shape = DTR_ANNULUS2D;
            out.io_key = kIoAnnulus2;
            n_bounds = 7;
            break;
        case kIoCylinder2:
        case kIoPortalCylinder2:
            out.shape = DTR_CONCENTRIC_CYLINDER2D;
            out.io_key = kIoPortalCylinder2;
            n_bounds = 3;
            break;
        case kIoRectangle2:
            out.shape = DTR_RECTANGLE2D;
            out.io_key = kIoRectangle2;
            n_bounds = 2;
            break;
        case kIoRing2:
            out.shape = DTR_RING2D;
            out.io_key = kIoRing2;
            n_bounds = 2;
            break;
        case kIoTrapezoid2:
            out.shape = DTR_TRAPEZOID2D;
            out.io_key = kIoTrapezoid2;
            n_bounds = 4;
            break;
        case kIoDriftCell:
            out.shape = DTR_LINE_SQUARE;
            out.io_key = kIoDriftCell;
            n_bounds = 2;
            break;
        case kIoStrawTube:
            out.shape = DTR_LINE_CIRCULAR;
            out.io_key = kIoStrawTube;
            n_bounds = 2;
            break;
        default:
            unsupported(where + ": mask shape " + kIoShapeNames[sid] +
                        " (the tables hold annulus2, rectangle2, trapezoid2, ring2, cylinder2, drift_cell "
                        "and straw_tube masks)");
    }
    for (size_t k = 0; k < s.masks.size(); ++k) {
        const std::string mw = where + ".masks[" + std::to_string(k) + "]";
        const MaskP& m = s.masks[k];
        if (m.boundaries.size() != n_bounds)
            bad(mw + ".boundaries: expected " + std::to_string(n_bounds) + " numbers");
        std::array<double, 7> b{};
        for (size_t i = 0; i < n_bounds; ++i) b[i] = m.boundaries[i];
        out.bounds.push_back(b);
        // checked against the volume count once every volume is read
        out.links.push_back(m.volume_link);
    }
    if (out.shape == DTR_CONCENTRIC_CYLINDER2D) {
        const double dz = out.trf.t[2];
        for (auto& b : out.bounds) {
            b[1] += dz;
            b[2] += dz;
        }
        out.trf = identity();
    }
    out.has_index = s.has_index;
    out.index = s.index;
    return out;
}

void ingest_geometry(const DetectorP& P, std::vector<VolumeIn>& vols) {
    if (P.volumes.empty()) bad("geometry file: no volumes");
    for (size_t vi = 0; vi < P.volumes.size(); ++vi) {
        const VolumeP& v = P.volumes[vi];
        const std::string where = "volume " + std::to_string(vi);
        VolumeIn V;
        V.placement = v.placement;
        if (v.surfaces.empty()) bad(where + ": volume without surfaces");
        std::vector<SurfaceIn> in;
        in.reserve(v.surfaces.size());
        for (size_t k = 0; k < v.surfaces.size(); ++k)
            in.push_back(ingest_surface(v.surfaces[k], where + " surface " + std::to_string(k)));
        const size_t n = in.size();
        size_t n_indexed = 0;
        for (const auto& s : in) n_indexed += s.has_index ? 1u : 0u;
        if (n_indexed == n && n > 0) {
            V.surfaces.resize(n);
            std::vector<char> seen(n, 0);
            for (const auto& s : in) {
                if (s.index >= n || seen[s.index])
                    bad(where + ": index_in_coll values are not a permutation of 0.." + std::to_string(n - 1));
                seen[s.index] = 1;
                V.surfaces[s.index] = s;
            }
        } else if (n_indexed == 0) {
            // factories: portals first, then the others, each by shape key
            for (int portals = 1; portals >= 0; --portals) {
                for (unsigned key = 0; key < kIoShapes; ++key) {
                    for (const auto& s : in) {
                        if ((s.sf_id == DTR_PORTAL) == (portals == 1) && s.io_key == key) V.surfaces.push_back(s);
                    }
                }
            }
        } else {
            bad(where + ": index_in_coll given for some surfaces only");
        }
        V.mat_of.assign(V.surfaces.size(), {-1, 0});
        vols.push_back(std::move(V));
    }
}

void ingest_grids(const DetectorP& P, std::vector<VolumeIn>& vols) {
    for (const GridP& g : P.grids) {
        const uint64_t vi = g.owner;
        const std::string where = "grid of volume " + std::to_string(vi);
        if (vi >= vols.size()) bad(where + ": the volume is not in the geometry");
        VolumeIn& V = vols[vi];
        if (g.type >= kIoAccels) bad(where + ": unknown grid type id " + std::to_string(g.type));
        dtr_grid gr{};
        unsigned want0, want1;
        if (g.type == kIoConcCyl2Grid) {
            gr.type = DTR_GRID_CYLINDER;
            want0 = kCircular;
            want1 = kClosed;
        } else if (g.type == kIoPolar2Grid) {
            gr.type = DTR_GRID_RING;
            want0 = kClosed;
            want1 = kCircular;
        } else {
            unsupported(where + ": grid type " + std::to_string(g.type) +
                        " (the tables hold concentric_cylinder2 and polar2 grids)");
        }
        if (V.has_grid) unsupported(where + ": more than one surface grid in a volume");
        if (g.axes.size() != 2u) bad(where + ": a 2D grid needs 2 axes");
        uint32_t nb[2];
        double lo[2], hi[2];
        for (int k = 0; k < 2; ++k) {
            const std::string aw = where + ".axes[" + std::to_string(k) + "]";
            const AxisP& a = g.axes[k];
            if (a.binning != kRegular) unsupported(aw + ": irregular binning");
            if (a.bounds != (k == 0 ? want0 : want1))
                unsupported(aw + ": axis bounds " + std::to_string(a.bounds) + " (the tables hold " +
                            (gr.type == DTR_GRID_CYLINDER ? "circular phi x closed z" : "closed r x circular phi") + ")");
            if (a.bins == 0u || a.bins > (1u << 20)) bad(aw + ": bad number of bins");
            if (a.edges.size() < 2u) bad(aw + ": an axis span needs 2 edges");
            nb[k] = static_cast<uint32_t>(a.bins);
            lo[k] = a.edges.front();
            hi[k] = a.edges.back();
        }
        gr.n0 = nb[0];
        gr.n1 = nb[1];
        gr.min0 = lo[0];
        gr.max0 = hi[0];
        gr.min1 = lo[1];
        gr.max1 = hi[1];
        if (static_cast<uint64_t>(gr.n0) * gr.n1 > (1ull << 26)) unsupported(where + ": more than 2^26 bins");
        V.bins.assign(static_cast<size_t>(gr.n0) * gr.n1, {});
        size_t cap = 1;
        for (const BinP& bin : g.bins) {
            if (bin.loc_index.size() != 2u) bad(where + ": a bin needs 2 local indices");
            const uint64_t i0 = bin.loc_index[0], i1 = bin.loc_index[1];
            if (i0 >= gr.n0 || i1 >= gr.n1)
                bad(where + ": bin (" + std::to_string(i0) + ", " + std::to_string(i1) + ") out of bounds");
            std::vector<uint32_t>& entries = V.bins[static_cast<size_t>(i0) * gr.n1 + i1];
            for (const uint64_t sf : bin.content) {
                if (static_cast<uint32_t>(sf) == kInvalid) continue;  // invalid entries are skipped
                if (sf >= V.surfaces.size()) bad(where + ": surface " + std::to_string(sf) + " not in the volume");
                entries.push_back(static_cast<uint32_t>(sf));
            }
            cap = std::max(cap, entries.size());
        }
        // the bins are stored with the largest occupancy as capacity
        if (cap > 255u) unsupported(where + ": more than 255 surfaces in a bin");
        if (static_cast<uint64_t>(gr.n0) * gr.n1 * cap > (1ull << 26))
            unsupported(where + ": more than 2^26 bin entries");
        gr.capacity = static_cast<uint8_t>(cap);
        V.has_grid = true;
        V.grid = gr;
    }
}

void ingest_material(const DetectorP& P, std::vector<VolumeIn>& vols) {
    for (const MaterialVolumeP& mv : P.materials) {
        const uint64_t vi = mv.volume;
        const std::string where = "material of volume " + std::to_string(vi);
        if (vi >= vols.size()) bad(where + ": the volume is not in the geometry");
        VolumeIn& V = vols[vi];
        const size_t n = mv.entries.size();
        if (n == 0) continue;
        struct Entry {
            uint64_t sf;
            bool has_index;
            uint64_t index;
            dtr_material_slab slab;
        };
        std::vector<Entry> es;
        for (size_t k = 0; k < n; ++k) {
            const std::string ew = where + " entry " + std::to_string(k);
            const MaterialP& e = mv.entries[k];
            if (e.type != kIoSlab && e.type != kIoRod)
                bad(ew + ": material type " + std::to_string(e.type) + " is not homogeneous");
            const double* p = e.params;
            Entry en{e.surface, e.has_index, e.index, make_slab(p[0], p[1], p[2], p[3], p[4], e.thickness)};
            // a rod's "thickness" is its radius (material_rod.hpp:59-60)
            if (e.type == kIoRod) en.slab.kind = DTR_MATERIAL_ROD;
            es.push_back(en);
        }
        // pairing by offset from the smallest surface index
        std::vector<uint64_t> idx(n);
        for (size_t k = 0; k < n; ++k) idx[k] = es[k].sf;
        std::vector<uint64_t> sorted = idx;
        std::sort(sorted.begin(), sorted.end());
        const uint64_t first = sorted.front();
        for (size_t k = 0; k < n; ++k)
            if (sorted[k] != first + k) bad(where + ": surface_idx values are not a contiguous range");
        if (first + n > V.surfaces.size()) bad(where + ": surface_idx beyond the volume's surfaces");
        for (size_t k = 0; k < n; ++k) {
            const uint64_t sf = idx[k];
            if (V.mat_of[sf].first >= 0) continue;  // already has material: kept
            const Entry& en = es[sf - first];
            const int kind = en.slab.kind == DTR_MATERIAL_ROD ? 1 : 0;
            auto& list = V.mats[kind];
            auto& set = V.mat_set[kind];
            size_t pos;
            if (en.has_index) {
                pos = static_cast<size_t>(en.index);
                if (pos > (1u << 24)) bad(where + ": index_in_coll out of range");
                if (list.size() <= pos) {
                    list.resize(pos + 1);
                    set.resize(pos + 1, 0);
                }
                if (set[pos]) bad(where + ": two materials at index_in_coll " + std::to_string(pos));
            } else {
                pos = list.size();
                list.emplace_back();
                set.push_back(0);
            }
            list[pos] = en.slab;
            set[pos] = 1;
            V.mat_of[sf] = {kind, pos};
        }
    }
    for (size_t vi = 0; vi < vols.size(); ++vi)
        for (int kind = 0; kind < 2; ++kind)
            for (size_t k = 0; k < vols[vi].mat_set[kind].size(); ++k)
                if (!vols[vi].mat_set[kind][k])
                    bad("material of volume " + std::to_string(vi) + ": no " + (kind ? "rod" : "slab") +
                        " at index_in_coll " + std::to_string(k));
}

// material_map_reader.hpp:95-185, material_map_builder.hpp:280-330: a
// regular closed or circular axis per bound local coordinate of the
// surface's frame, one slab per bin (unlisted bins hold no material)
void ingest_material_maps(const DetectorP& P, std::vector<VolumeIn>& vols) {
    for (const MapP& M : P.maps) {
        const std::string where = "material map of volume " + std::to_string(M.volume) + " surface " +
                                  std::to_string(M.owner);
        if (M.volume >= vols.size()) bad(where + ": the volume is not in the geometry");
        VolumeIn& V = vols[M.volume];
        if (V.map_of.size() != V.surfaces.size()) V.map_of.assign(V.surfaces.size(), -1);
        if (M.owner >= V.surfaces.size()) bad(where + ": owner_link beyond the volume's surfaces");
        if (V.map_of[M.owner] >= 0) bad(where + ": two material maps on one surface");
        if (V.mat_of[M.owner].first >= 0) bad(where + ": the surface also has homogeneous material");
        const uint8_t shape = V.surfaces[M.owner].shape;
        unsigned want;
        switch (shape) {
            case DTR_RING2D:
            case DTR_ANNULUS2D: want = kIoPolar2Map; break;
            case DTR_RECTANGLE2D:
            case DTR_TRAPEZOID2D: want = kIoCart2Map; break;
            case DTR_CONCENTRIC_CYLINDER2D: want = kIoConcCyl2Map; break;
            case DTR_CYLINDER2D: want = kIoCyl2Map; break;
            default: unsupported(where + ": material maps on line surfaces");
        }
        if (M.type != want)
            bad(where + ": grid_link.type " + std::to_string(M.type) + " is not the map of the surface's frame (" +
                std::to_string(want) + ")");
        if (M.axes.size() != 2u) unsupported(where + ": only 2D material maps");
        dtr_material_map m{};
        for (int k = 0; k < 2; ++k) {
            const AxisP& a = M.axes[k];
            const std::string aw = where + " axis " + std::to_string(k);
            if (a.binning != kRegular) unsupported(aw + ": irregular binning");
            if (a.bounds != kClosed && a.bounds != kCircular) unsupported(aw + ": open bounds");
            if (a.bins == 0u || a.bins > (1u << 20)) bad(aw + ": bin count");
            if (a.edges.size() != 2u || !(a.edges[1] > a.edges[0])) bad(aw + ": edges are not [min, max]");
            (k ? m.n1 : m.n0) = static_cast<uint32_t>(a.bins);
            (k ? m.circular1 : m.circular0) = a.bounds == kCircular ? 1u : 0u;
            (k ? m.min1 : m.min0) = a.edges[0];
            (k ? m.max1 : m.max0) = a.edges[1];
        }
        const size_t nb = static_cast<size_t>(m.n0) * m.n1;
        if (nb > (1u << 24)) unsupported(where + ": more than 2^24 bins");
        std::vector<dtr_material_slab> slabs(nb, dtr_material_slab{});
        std::vector<char> set(nb, 0);
        for (size_t b = 0; b < M.loc_index.size(); ++b) {
            const std::string bw = where + " bin " + std::to_string(b);
            const auto& li = M.loc_index[b];
            if (li.size() != 2u || li[0] >= m.n0 || li[1] >= m.n1) bad(bw + ": loc_index out of range");
            const size_t at = static_cast<size_t>(li[0]) * m.n1 + li[1];
            if (set[at]) bad(bw + ": bin given twice");
            set[at] = 1;
            const auto& c = M.content[b];
            if (c.size() > 1u) bad(bw + ": more than one slab per bin");
            if (c.empty()) continue;
            if (c[0].type != kIoSlab) bad(bw + ": bin content is not a material slab");
            const double* p = c[0].params;
            slabs[at] = make_slab(p[0], p[1], p[2], p[3], p[4], c[0].thickness);
        }
        V.map_of[M.owner] = static_cast<int>(V.maps.size());
        V.maps.emplace_back(m, std::move(slabs));
    }
}

void assemble(const std::vector<VolumeIn>& vols, Tables& T) {
    if (vols.size() > 4096u) unsupported("more than 4096 volumes");
    for (uint32_t vi = 0; vi < vols.size(); ++vi) {
        const VolumeIn& V = vols[vi];
        dtr_volume dv{};
        dv.transform = static_cast<uint32_t>(T.transforms.size());
        T.transforms.push_back(V.placement);
        dv.sf_begin = static_cast<uint32_t>(T.surfaces.size());
        const uint32_t mat_offset = static_cast<uint32_t>(T.materials.size());
        for (uint32_t k = 0; k < V.surfaces.size(); ++k) {
            const SurfaceIn& S = V.surfaces[k];
            dtr_surface ds{};
            const uint32_t idx = static_cast<uint32_t>(T.surfaces.size());
            ds.transform = static_cast<uint32_t>(T.transforms.size());
            T.transforms.push_back(S.trf);
            ds.mask = static_cast<uint32_t>(T.masks.size());
            ds.n_masks = static_cast<uint8_t>(S.bounds.size());
            ds.shape = S.shape;
            for (size_t m = 0; m < S.bounds.size(); ++m) {
                dtr_mask dm{};
                for (int q = 0; q < 7; ++q) dm.v[q] = S.bounds[m][q];
                // the navigation link is a 16-bit value: the world exit (the
                // invalid value of a 16-bit, 32-bit or 64-bit link,
                // basic_converter.hpp:21-31) -> 0xffff; any other link names
                // a volume of the file
                const uint64_t link = S.links[m];
                if (link == ~0ull || link == 0xffffffffull || link == kEndOfWorld) {
                    dm.volume_link = kEndOfWorld;
                } else if (link < vols.size()) {
                    dm.volume_link = static_cast<uint16_t>(link);
                } else {
                    bad("volume " + std::to_string(vi) + " surface " + std::to_string(k) + ": mask volume_link " +
                        std::to_string(link) + " is neither a volume of the file (" + std::to_string(vols.size()) +
                        ") nor the world exit");
                }
                dm.shape = S.shape;
                T.masks.push_back(dm);
            }
            // the volume's slabs, then its rods
            const auto mo = V.mat_of[k];
            ds.material = mo.first < 0 ? kInvalid
                                       : mat_offset + static_cast<uint32_t>(mo.first ? V.mats[0].size() : 0u) +
                                             static_cast<uint32_t>(mo.second);
            if (!V.map_of.empty() && V.map_of[k] >= 0)
                ds.material = DTR_MATERIAL_MAP | static_cast<uint32_t>(T.maps.size() + V.map_of[k]);
            ds.volume = static_cast<uint16_t>(vi);
            ds.sf_id = S.sf_id;
            ds.barcode = make_barcode(vi, S.sf_id, idx, ds.transform);
            T.surfaces.push_back(ds);
        }
        if (T.surfaces.size() >= (1u << 21) || T.transforms.size() >= (1u << 21))
            unsupported("more than 2^21 surfaces or transforms");
        T.materials.insert(T.materials.end(), V.mats[0].begin(), V.mats[0].end());
        T.materials.insert(T.materials.end(), V.mats[1].begin(), V.mats[1].end());
        // then the slabs of its material maps
        for (const auto& mp : V.maps) {
            dtr_material_map m = mp.first;
            m.slabs = static_cast<uint32_t>(T.materials.size());
            T.materials.insert(T.materials.end(), mp.second.begin(), mp.second.end());
            T.maps.push_back(m);
        }
        dv.sf_end = static_cast<uint32_t>(T.surfaces.size());
        dv.bf_begin = static_cast<uint32_t>(T.bf.size());
        for (uint32_t s = dv.sf_begin; s < dv.sf_end; ++s)
            if (!V.has_grid || T.surfaces[s].sf_id != DTR_SENSITIVE) T.bf.push_back(s);
        dv.bf_end = static_cast<uint32_t>(T.bf.size());
        dv.grid = -1;
        if (V.has_grid) {
            dtr_grid g = V.grid;
            g.bins = static_cast<uint32_t>(T.bins.size());
            const size_t cap = g.capacity ? g.capacity : 1u;
            for (const auto& entries : V.bins)
                for (size_t k = 0; k < cap; ++k)
                    T.bins.push_back(k < entries.size() ? dv.sf_begin + entries[k] : kInvalid);
            // grid_builder.hpp:232-272: reference radius between the portal cylinders
            if (g.type == DTR_GRID_CYLINDER) {
                double rmin = 0.0, rmax = 0.0;
                bool any = false;
                for (uint32_t s = dv.sf_begin; s < dv.sf_end; ++s) {
                    const dtr_surface& ds = T.surfaces[s];
                    if (ds.sf_id != DTR_PORTAL || ds.shape != DTR_CONCENTRIC_CYLINDER2D) continue;
                    const double r = T.masks[ds.mask].v[0];
                    rmin = any ? std::min(rmin, r) : r;
                    rmax = any ? std::max(rmax, r) : r;
                    any = true;
                }
                g.radius = any ? 0.5 * (rmin + rmax) : 0.0;
            }
            dv.grid = static_cast<int32_t>(T.grids.size());
            T.grids.push_back(g);
        }
        T.volumes.push_back(dv);
    }
}

// ---------------------------------------------------------------------------
// Writing
// ---------------------------------------------------------------------------
void check_desc(const dtr_detector_desc& d) {
    if (d.n_volumes == 0u || !d.volumes || !d.surfaces || !d.transforms || !d.masks)
        bad("detector tables are empty");
    const std::string why = validate_desc(d);
    if (!why.empty()) bad(why);
    uint32_t next = 0;
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        if (v.sf_begin != next || v.sf_end < v.sf_begin || v.sf_end > d.n_surfaces)
            bad("volume " + std::to_string(vi) + ": surface range does not follow the previous volume");
        if (v.transform >= d.n_transforms) bad("volume " + std::to_string(vi) + ": transform out of range");
        for (uint32_t s = v.sf_begin; s < v.sf_end; ++s) {
            const dtr_surface& ds = d.surfaces[s];
            if (ds.volume != vi) bad("surface " + std::to_string(s) + ": volume link does not match its range");
            if (ds.transform >= d.n_transforms || ds.n_masks == 0u ||
                static_cast<uint64_t>(ds.mask) + ds.n_masks > d.n_masks)
                bad("surface " + std::to_string(s) + ": transform or mask out of range");
            if (ds.material != kInvalid && !(ds.material & DTR_MATERIAL_MAP) &&
                (ds.material >= d.n_materials || !d.materials))
                bad("surface " + std::to_string(s) + ": material out of range");
        }
        if (v.grid >= 0) {
            if (static_cast<uint32_t>(v.grid) >= d.n_grids || !d.grids)
                bad("volume " + std::to_string(vi) + ": grid out of range");
            const dtr_grid& g = d.grids[v.grid];
            if (static_cast<uint64_t>(g.bins) + static_cast<uint64_t>(g.n0) * g.n1 * (g.capacity ? g.capacity : 1u) >
                d.n_bins)
                bad("volume " + std::to_string(vi) + ": grid bins out of range");
        }
        next = v.sf_end;
    }
    if (next != d.n_surfaces) bad("surfaces not covered by the volume ranges");
}

unsigned io_shape(uint8_t shape) {
    switch (shape) {
        case DTR_RECTANGLE2D: return kIoRectangle2;
        case DTR_TRAPEZOID2D: return kIoTrapezoid2;
        case DTR_CONCENTRIC_CYLINDER2D: return kIoPortalCylinder2;
        case DTR_RING2D: return kIoRing2;
        case DTR_LINE_SQUARE: return kIoDriftCell;
        case DTR_LINE_CIRCULAR: return kIoStrawTube;
        case DTR_ANNULUS2D: return kIoAnnulus2;
        // the reference's reader takes cylinder2 as a portal cylinder
        // (geometry_reader.hpp:110-113): a placed cylinder has no file form
        case DTR_CYLINDER2D: unsupported("cylinder2 surfaces cannot be written: the reader makes portal cylinders of them");
        default: bad("unknown mask shape " + std::to_string(shape));
    }
}

size_t n_bounds(uint8_t shape) {
    switch (shape) {
        case DTR_ANNULUS2D: return 7;
        case DTR_TRAPEZOID2D: return 4;
        case DTR_CONCENTRIC_CYLINDER2D: return 3;
        default: return 2;
    }
}

std::string today() {
    const std::time_t now = std::time(nullptr);
    std::tm tm{};
    localtime_r(&now, &tm);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%d/%m/%Y %H:%M", &tm);
    return buf;
}

Value header_common(const char* name, const char* tag) {
    Value c = Value::object();
    c.set("version", Value::string("detray_amd"));
    c.set("detector", Value::string(name ? name : ""));
    c.set("date", Value::string(today()));
    c.set("tag", Value::string(tag));
    return c;
}

// Per-type grid ordinal (the index of a grid in the accelerator collection
// of its type)
std::vector<uint32_t> grid_ordinals(const dtr_detector_desc& d) {
    std::vector<uint32_t> ord(d.n_grids, 0);
    uint32_t count[2] = {0, 0};
    for (uint32_t g = 0; g < d.n_grids; ++g) ord[g] = count[d.grids[g].type == DTR_GRID_RING ? 1 : 0]++;
    return ord;
}

unsigned io_grid(uint8_t type) { return type == DTR_GRID_RING ? kIoPolar2Grid : kIoConcCyl2Grid; }

// io::material_id of the map on a surface of this shape
unsigned io_map(uint8_t shape) {
    switch (shape) {
        case DTR_RING2D:
        case DTR_ANNULUS2D: return kIoPolar2Map;
        case DTR_RECTANGLE2D:
        case DTR_TRAPEZOID2D: return kIoCart2Map;
        case DTR_CONCENTRIC_CYLINDER2D: return kIoConcCyl2Map;
        default: return kIoCyl2Map;
    }
}

// Per-type ordinal of each material map (its index in the map collection of
// its type), in surface order
std::vector<uint32_t> map_ordinals(const dtr_detector_desc& d) {
    std::vector<uint32_t> ord(d.n_material_maps, 0);
    uint32_t count[8] = {};
    for (uint32_t s = 0; s < d.n_surfaces; ++s) {
        const uint32_t m = d.surfaces[s].material;
        if (m == kInvalid || !(m & DTR_MATERIAL_MAP)) continue;
        ord[m & ~DTR_MATERIAL_MAP] = count[io_map(d.surfaces[s].shape)]++;
    }
    return ord;
}

Value write_geometry(const dtr_detector_desc& d, const char* name, const char* const* names) {
    const std::vector<uint32_t> map_ord = map_ordinals(d);
    Value root = Value::object();
    Value& h = root.set("header", Value::object());
    h.set("common", header_common(name, "geometry"));
    h.set("volume_count", Value::uint(d.n_volumes));
    h.set("surface_count", Value::uint(d.n_surfaces));
    Value& data = root.set("data", Value::object());
    Value& jv = data.set("volumes", Value::array());
    const std::vector<uint32_t> ord = grid_ordinals(d);
    uint64_t slab = 0, rods = 0;  // one slab or rod per material surface, in surface order
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        Value vol = Value::object();
        vol.set("name", Value::string((names && names[vi]) ? names[vi] : ""));
        vol.set("index", Value::uint(vi));
        unsigned vtype = kVolCuboid;
        for (uint32_t s = v.sf_begin; s < v.sf_end; ++s)
            if (d.surfaces[s].sf_id == DTR_PORTAL && (d.surfaces[s].shape == DTR_CONCENTRIC_CYLINDER2D ||
                                                      d.surfaces[s].shape == DTR_RING2D))
                vtype = kVolCylinder;
        vol.set("type", Value::uint(vtype));
        vol.set("transform", write_transform(d.transforms[v.transform]));
        Value& sfs = vol.set("surfaces", Value::array());
        for (uint32_t s = v.sf_begin; s < v.sf_end; ++s) {
            const dtr_surface& ds = d.surfaces[s];
            Value sf = Value::object();
            sf.set("identifier", Value::uint(ds.barcode));
            sf.set("type", Value::uint(ds.sf_id));
            sf.set("source", Value::uint(0));
            sf.set("transform", write_transform(d.transforms[ds.transform]));
            Value& masks = sf.set("masks", Value::array());
            for (uint32_t m = ds.mask; m < ds.mask + ds.n_masks; ++m) {
                const dtr_mask& dm = d.masks[m];
                Value jm = Value::object();
                jm.set("shape", Value::uint(io_shape(ds.shape)));
                jm.set("volume_link", Value::uint(dm.volume_link));
                Value& b = jm.set("boundaries", Value::array());
                for (size_t q = 0; q < n_bounds(ds.shape); ++q) b.push(Value::number(dm.v[q]));
                masks.push(std::move(jm));
            }
            if (ds.material != kInvalid && (ds.material & DTR_MATERIAL_MAP)) {
                const uint32_t mi = ds.material & ~DTR_MATERIAL_MAP;
                Value ml = Value::object();
                ml.set("type", Value::uint(io_map(ds.shape)));
                ml.set("index", Value::uint(map_ord[mi]));
                sf.set("material", std::move(ml));
            } else if (ds.material != kInvalid) {
                const bool rod = d.materials[ds.material].kind == DTR_MATERIAL_ROD;
                Value ml = Value::object();
                ml.set("type", Value::uint(rod ? kIoRod : kIoSlab));
                ml.set("index", Value::uint(rod ? rods++ : slab++));
                sf.set("material", std::move(ml));
            }
            sf.set("index_in_coll", Value::uint(s - v.sf_begin));
            sfs.push(std::move(sf));
        }
        if (v.grid >= 0) {
            Value links = Value::array();
            Value l = Value::object();
            l.set("type", Value::uint(io_grid(d.grids[v.grid].type)));
            l.set("index", Value::uint(ord[v.grid]));
            links.push(std::move(l));
            vol.set("acc_links", std::move(links));
        }
        jv.push(std::move(vol));
    }
    return root;
}

Value write_axis(unsigned label, unsigned bounds, uint32_t bins, double lo, double hi) {
    Value a = Value::object();
    a.set("label", Value::uint(label));
    a.set("bounds", Value::uint(bounds));
    a.set("binning", Value::uint(kRegular));
    a.set("bins", Value::uint(bins));
    Value e = Value::array();
    e.push(Value::number(lo));
    e.push(Value::number(hi));
    a.set("edges", std::move(e));
    return a;
}

Value write_grids(const dtr_detector_desc& d, const char* name) {
    Value root = Value::object();
    Value& h = root.set("header", Value::object());
    h.set("common", header_common(name, "surface_grids"));
    h.set("grid_count", Value::uint(d.n_grids));
    const std::vector<uint32_t> ord = grid_ordinals(d);
    Value grids = Value::array();
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        if (v.grid < 0) continue;
        const dtr_grid& g = d.grids[v.grid];
        Value gd = Value::object();
        gd.set("owner_link", Value::uint(vi));
        Value link = Value::object();
        link.set("type", Value::uint(io_grid(g.type)));
        link.set("index", Value::uint(ord[v.grid]));
        gd.set("grid_link", std::move(link));
        Value axes = Value::array();
        const bool cyl = g.type != DTR_GRID_RING;
        axes.push(write_axis(0u, cyl ? kCircular : kClosed, g.n0, g.min0, g.max0));
        axes.push(write_axis(1u, cyl ? kClosed : kCircular, g.n1, g.min1, g.max1));
        gd.set("axes", std::move(axes));
        Value bins = Value::array();
        // serial bin order of simple_serializer<2>: axis 0 runs fastest
        for (uint32_t i1 = 0; i1 < g.n1; ++i1) {
            for (uint32_t i0 = 0; i0 < g.n0; ++i0) {
                Value b = Value::object();
                Value li = Value::array();
                li.push(Value::uint(i0));
                li.push(Value::uint(i1));
                b.set("loc_index", std::move(li));
                Value content = Value::array();
                const uint32_t cap = g.capacity ? g.capacity : 1u;
                for (uint32_t k = 0; k < cap; ++k) {
                    const uint32_t sf = d.grid_bins[g.bins + (static_cast<size_t>(i0) * g.n1 + i1) * cap + k];
                    if (sf == kInvalid) continue;
                    if (sf < v.sf_begin || sf >= v.sf_end)
                        bad("grid of volume " + std::to_string(vi) + ": bin holds a surface of another volume");
                    content.push(Value::uint(sf - v.sf_begin));
                }
                b.set("content", std::move(content));
                bins.push(std::move(b));
            }
        }
        gd.set("bins", std::move(bins));
        Value entry = Value::object();
        entry.set("volume_link", Value::uint(vi));
        Value list = Value::array();
        list.push(std::move(gd));
        entry.set("grid_data", std::move(list));
        grids.push(std::move(entry));
    }
    if (grids.a.empty()) {
        root.set("data", Value::null());
    } else {
        Value& data = root.set("data", Value::object());
        data.set("grids", std::move(grids));
    }
    return root;
}

Value write_material(const dtr_detector_desc& d, const char* name) {
    uint64_t n_slabs = 0, n_rods = 0;
    for (uint32_t s = 0; s < d.n_surfaces; ++s) {
        if (d.surfaces[s].material == kInvalid || (d.surfaces[s].material & DTR_MATERIAL_MAP)) continue;
        if (d.materials[d.surfaces[s].material].kind == DTR_MATERIAL_ROD) ++n_rods;
        else ++n_slabs;
    }
    Value root = Value::object();
    Value& h = root.set("header", Value::object());
    h.set("common", header_common(name, "homogeneous_material"));
    h.set("slab_count", Value::uint(n_slabs));
    h.set("rod_count", Value::uint(n_rods));
    h.set("slab_surface_count", Value::uint(n_slabs));
    h.set("rod_surface_count", Value::uint(n_rods));
    Value& data = root.set("data", Value::object());
    Value& jv = data.set("volumes", Value::array());
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        Value mv = Value::object();
        mv.set("volume_link", Value::uint(vi));
        Value list = Value::array();
        uint64_t slab = 0, rods = 0;
        for (uint32_t s = v.sf_begin; s < v.sf_end; ++s) {
            const dtr_surface& ds = d.surfaces[s];
            if (ds.material == kInvalid || (ds.material & DTR_MATERIAL_MAP)) continue;
            const dtr_material_slab& m = d.materials[ds.material];
            const bool rod = m.kind == DTR_MATERIAL_ROD;
            Value e = Value::object();
            e.set("type", Value::uint(rod ? kIoRod : kIoSlab));
            e.set("surface_idx", Value::uint(s - v.sf_begin));
            e.set("thickness", Value::number(m.thickness));
            Value mat = Value::object();
            Value p = Value::array();
            for (double x : {m.X0, m.L0, m.A, m.Z, m.mass_rho, m.molar_rho, 0.0}) p.push(Value::number(x));
            mat.set("params", std::move(p));
            e.set("material", std::move(mat));
            e.set("index_in_coll", Value::uint(rod ? rods++ : slab++));
            list.push(std::move(e));
        }
        if (!list.a.empty()) mv.set("surface_material", std::move(list));
        jv.push(std::move(mv));
    }
    return root;
}

Value slab_entry(const dtr_material_slab& m, uint32_t surface_idx) {
    Value e = Value::object();
    e.set("type", Value::uint(kIoSlab));
    e.set("surface_idx", Value::uint(surface_idx));
    e.set("thickness", Value::number(m.thickness));
    Value mat = Value::object();
    Value p = Value::array();
    for (double x : {m.X0, m.L0, m.A, m.Z, m.mass_rho, m.molar_rho, 0.0}) p.push(Value::number(x));
    mat.set("params", std::move(p));
    e.set("material", std::move(mat));
    return e;
}

// material_map_writer.hpp:45-110 over grid_writer.hpp:56-160: per volume,
// one grid payload per mapped surface, owner_link its index in the volume,
// the bins in the serial order of the grid (axis 0 fastest), one slab each
Value write_material_maps(const dtr_detector_desc& d, const char* name) {
    const std::vector<uint32_t> ord = map_ordinals(d);
    Value root = Value::object();
    Value& h = root.set("header", Value::object());
    h.set("common", header_common(name, "material_maps"));
    h.set("grid_count", Value::uint(d.n_material_maps));
    Value grids = Value::array();
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        Value list = Value::array();
        for (uint32_t s = v.sf_begin; s < v.sf_end; ++s) {
            const dtr_surface& ds = d.surfaces[s];
            if (ds.material == kInvalid || !(ds.material & DTR_MATERIAL_MAP)) continue;
            const uint32_t mi = ds.material & ~DTR_MATERIAL_MAP;
            const dtr_material_map& m = d.material_maps[mi];
            Value gd = Value::object();
            gd.set("owner_link", Value::uint(s - v.sf_begin));
            Value link = Value::object();
            link.set("type", Value::uint(io_map(ds.shape)));
            link.set("index", Value::uint(ord[mi]));
            gd.set("grid_link", std::move(link));
            Value axes = Value::array();
            axes.push(write_axis(0u, m.circular0 ? kCircular : kClosed, m.n0, m.min0, m.max0));
            axes.push(write_axis(1u, m.circular1 ? kCircular : kClosed, m.n1, m.min1, m.max1));
            gd.set("axes", std::move(axes));
            Value bins = Value::array();
            for (uint32_t i1 = 0; i1 < m.n1; ++i1) {
                for (uint32_t i0 = 0; i0 < m.n0; ++i0) {
                    Value b = Value::object();
                    Value li = Value::array();
                    li.push(Value::uint(i0));
                    li.push(Value::uint(i1));
                    b.set("loc_index", std::move(li));
                    Value content = Value::array();
                    content.push(slab_entry(d.materials[m.slabs + static_cast<size_t>(i0) * m.n1 + i1], s - v.sf_begin));
                    b.set("content", std::move(content));
                    bins.push(std::move(b));
                }
            }
            gd.set("bins", std::move(bins));
            list.push(std::move(gd));
        }
        if (list.a.empty()) continue;
        Value entry = Value::object();
        entry.set("volume_link", Value::uint(vi));
        entry.set("grid_data", std::move(list));
        grids.push(std::move(entry));
    }
    if (grids.a.empty()) {
        root.set("data", Value::null());
    } else {
        Value& data = root.set("data", Value::object());
        data.set("grids", std::move(grids));
    }
    return root;
}

// payloads -> validated tables, exported into `out`
void build_tables(const DetectorP& P, dtr_detector_desc* out) {
    std::vector<VolumeIn> vols;
    ingest_geometry(P, vols);
    ingest_grids(P, vols);
    ingest_material(P, vols);
    ingest_material_maps(P, vols);
    Tables T;
    assemble(vols, T);
    dtr_detector_desc view{};
    view.n_volumes = static_cast<uint32_t>(T.volumes.size());
    view.n_surfaces = static_cast<uint32_t>(T.surfaces.size());
    view.n_transforms = static_cast<uint32_t>(T.transforms.size());
    view.n_masks = static_cast<uint32_t>(T.masks.size());
    view.n_materials = static_cast<uint32_t>(T.materials.size());
    view.n_grids = static_cast<uint32_t>(T.grids.size());
    view.n_bf = static_cast<uint32_t>(T.bf.size());
    view.n_bins = static_cast<uint32_t>(T.bins.size());
    view.volumes = T.volumes.data();
    view.surfaces = T.surfaces.data();
    view.transforms = T.transforms.data();
    view.masks = T.masks.data();
    view.materials = T.materials.data();
    view.grids = T.grids.data();
    view.bf_surfaces = T.bf.data();
    view.grid_bins = T.bins.data();
    view.n_material_maps = static_cast<uint32_t>(T.maps.size());
    view.material_maps = T.maps.data();
    const std::string why = validate_desc(view);
    if (!why.empty()) bad("detector tables read from the files: " + why);
    export_tables(T, out);
}

void message(char* out, size_t n, const std::string& m) {
    if (!out || n == 0) return;
    const size_t k = std::min(n - 1, m.size());
    std::memcpy(out, m.data(), k);
    out[k] = '\0';
}

}  // namespace

extern "C" {

int dtr_read_detector_json(const char* geometry_file, const char* grids_file, const char* material_file,
                           const char* material_maps_file, dtr_detector_desc* out, char* msg, size_t msg_size) {
    message(msg, msg_size, "");
    if (!geometry_file || !out) {
        message(msg, msg_size, "geometry file and output tables are required");
        return DTR_ERROR_INVALID_ARGUMENT;
    }
    try {
        DetectorP P;
        parse_geometry(load(geometry_file, "geometry"), P);
        if (grids_file) parse_grids(load(grids_file, "surface grids"), P);
        if (material_file) parse_material(load(material_file, "homogeneous material"), P);
        if (material_maps_file) parse_material_maps(load(material_maps_file, "material maps"), P);
        build_tables(P, out);
        return DTR_SUCCESS;
    } catch (const IoError& e) {
        message(msg, msg_size, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        message(msg, msg_size, "out of host memory");
        return DTR_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        message(msg, msg_size, e.what());
        return DTR_ERROR_INVALID_ARGUMENT;
    }
}

int dtr_flatten_detector(const dtr_detector_view* view, dtr_detector_desc* out, char* msg, size_t msg_size) {
    message(msg, msg_size, "");
    if (!view || !out) {
        message(msg, msg_size, "detector view and output tables are required");
        return DTR_ERROR_INVALID_ARGUMENT;
    }
    try {
        DetectorP P;
        from_view(*view, P);
        build_tables(P, out);
        return DTR_SUCCESS;
    } catch (const IoError& e) {
        message(msg, msg_size, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        message(msg, msg_size, "out of host memory");
        return DTR_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        message(msg, msg_size, e.what());
        return DTR_ERROR_INVALID_ARGUMENT;
    }
}

int dtr_write_detector_json(const dtr_detector_desc* det, const char* detector_name,
                            const char* const* volume_names, const char* geometry_file,
                            const char* grids_file, const char* material_file,
                            const char* material_maps_file, char* msg, size_t msg_size) {
    message(msg, msg_size, "");
    if (!det) {
        message(msg, msg_size, "detector tables are required");
        return DTR_ERROR_INVALID_ARGUMENT;
    }
    try {
        check_desc(*det);
        for (uint32_t vi = 0; vi < det->n_volumes; ++vi)
            if (det->volumes[vi].material != 0u) unsupported("volume material is not written");
        if (geometry_file) save(geometry_file, write_geometry(*det, detector_name, volume_names), "geometry");
        if (grids_file) save(grids_file, write_grids(*det, detector_name), "surface grids");
        if (material_file) save(material_file, write_material(*det, detector_name), "homogeneous material");
        if (material_maps_file) save(material_maps_file, write_material_maps(*det, detector_name), "material maps");
        return DTR_SUCCESS;
    } catch (const IoError& e) {
        message(msg, msg_size, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        message(msg, msg_size, "out of host memory");
        return DTR_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        message(msg, msg_size, e.what());
        return DTR_ERROR_INVALID_ARGUMENT;
    }
}

}  // extern "C"

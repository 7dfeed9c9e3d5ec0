// Host-side emitters of the flat detector tables consumed by the device
// propagation: detray's toy detector and telescope detector, a sampled
// finite-solenoid field and seeded track generators.
//
// The toy emitter reproduces the volume layout, surface numbering, mask and
// material assignment of tests/include/detray/test/common/
// build_toy_detector.hpp:50-1356 (barrel_generator.hpp, endcap_generator.hpp,
// cylinder_portal_generator.hpp, grid_builder.hpp) so that surface indices
// and barcodes match the reference detector; the telescope emitter follows
// build_telescope_detector.hpp:140-287, telescope_generator.hpp and
// cuboid_portal_generator.hpp.  Values the reference writes as float
// literals are converted the same way (float -> double).

#include "detray_amd.h"
#include "host/tables.hpp"

#include <algorithm>
#include <array>
#include <stdexcept>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

namespace {

using namespace dtr_host;

constexpr double kPi = M_PI;

inline double F(double x) { return static_cast<double>(static_cast<float>(x)); }

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x,
            a.x * b.y - b.x * a.y};
}
inline double perp(V3 v) { return std::sqrt(std::fma(v.x, v.x, v.y * v.y)); }
inline double phi_of(V3 v) { return std::atan2(v.y, v.x); }

// transform3(t, x, y, z): columns of the rotation are the local axes
dtr_transform make_trf(V3 t, V3 x, V3 y, V3 z) {
    dtr_transform tr{};
    tr.r[0] = x.x; tr.r[1] = y.x; tr.r[2] = z.x;
    tr.r[3] = x.y; tr.r[4] = y.y; tr.r[5] = z.y;
    tr.r[6] = x.z; tr.r[7] = y.z; tr.r[8] = z.z;
    tr.t[0] = t.x; tr.t[1] = t.y; tr.t[2] = t.z;
    return tr;
}
// transform3(t, z, x): y = z x x (generic_transform3.hpp:112-114)
dtr_transform make_trf_zx(V3 t, V3 z, V3 x) { return make_trf(t, x, cross(z, x), z); }
dtr_transform make_translation(V3 t) {
    return make_trf(t, {1, 0, 0}, {0, 1, 0}, {0, 0, 1});
}
V3 to_global(const dtr_transform& tr, V3 v) {
    // rotate() accumulates column by column (generic_transform3.hpp:228-260)
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    r0 += tr.r[0] * v.x; r1 += tr.r[3] * v.x; r2 += tr.r[6] * v.x;
    r0 += tr.r[1] * v.y; r1 += tr.r[4] * v.y; r2 += tr.r[7] * v.y;
    r0 += tr.r[2] * v.z; r1 += tr.r[5] * v.z; r2 += tr.r[8] * v.z;
    return {r0 + tr.t[0], r1 + tr.t[1], r2 + tr.t[2]};
}
V3 to_local(const dtr_transform& tr, V3 p) {
    const V3 d{p.x - tr.t[0], p.y - tr.t[1], p.z - tr.t[2]};
    return {tr.r[0] * d.x + tr.r[3] * d.y + tr.r[6] * d.z,
            tr.r[1] * d.x + tr.r[4] * d.y + tr.r[7] * d.z,
            tr.r[2] * d.x + tr.r[5] * d.y + tr.r[8] * d.z};
}
V3 axis(const dtr_transform& tr, int c) { return {tr.r[c], tr.r[3 + c], tr.r[6 + c]}; }

struct Material {
    double X0, L0, A, Z, rho;
};
// predefined_materials.hpp:178-189
Material silicon_tml() {
    return {F(95.7) * kMM, F(465.2) * kMM, F(28.03), F(14.0),
            2.32 * kGram / kCm3};
}
Material beryllium_tml() {
    return {F(352.8) * kMM, F(407.0) * kMM, F(9.012), F(4.0),
            1.848 * kGram / kCm3};
}

// predefined_materials.hpp:159-163
Material tungsten() {
    return {F(3.504) * kMM, F(99.46) * kMM, F(183.84), F(74.0), 19.3 * kGram / kCm3};
}

dtr_material_slab make_slab(const Material& m, double thickness) {
    return dtr_host::make_slab(m.X0, m.L0, m.A, m.Z, m.rho, thickness);
}

// The materials of predefined_materials.hpp that dtr_predefined_material
// names: X0, L0, A, Z, mass density, and for the *_with_ded entries the
// density-effect parameters (a, k, x0, x1, I [eV], C-bar); float literals
// there are converted float -> double as in the reference.
struct NamedMaterial {
    const char* name;
    double X0, L0, A, Z, rho_g_cm3;
    bool ded;
    double a, k, x0, x1, I_eV, cbar;
};
const NamedMaterial kMaterials[] = {
    {"silicon", F(93.7), F(465.2), F(28.0855), F(14.0), 2.329, false, 0, 0, 0, 0, 0, 0},
    {"silicon_with_ded", F(93.7), F(465.2), F(28.0855), F(14.0), 2.329, true,
     F(0.1492), F(3.2546), F(0.2015), F(2.8716), F(173.0), F(4.4355)},
    {"silicon_tml", F(95.7), F(465.2), F(28.03), F(14.0), 2.32, false, 0, 0, 0, 0, 0, 0},
    {"beryllium_tml", F(352.8), F(407.0), F(9.012), F(4.0), 1.848, false, 0, 0, 0, 0, 0, 0},
    {"tungsten", F(3.504), F(99.46), F(183.84), F(74.0), 19.3, false, 0, 0, 0, 0, 0, 0},
    {"isobutane", F(1693e2), F(288.3), F(58.124), F(34.0), 2.67, false, 0, 0, 0, 0, 0, 0},
    {"argon_liquid", F(14.0) * 10.0, F(85.77) * 10.0, F(39.948), F(18.0), 1.396, false, 0, 0, 0, 0, 0, 0},
    {"iron_with_ded", F(1.757) * 10.0, F(16.77) * 10.0, F(55.845), F(26.0), 7.874, true,
     F(0.14680), F(2.9632), F(-0.0012), F(3.1531), F(286.0), F(4.2911)},
    {"copper_with_ded", F(1.436) * 10.0, F(15.32) * 10.0, F(63.546), F(29.0), 8.960, true,
     F(0.14339), F(2.9044), F(-0.0254), F(3.2792), F(322.0), F(4.4190)},
    {"cesium_iodide", F(1.86) * 10.0, F(38.04) * 10.0, F(259.81), F(108.0), F(4.510), false, 0, 0, 0, 0, 0, 0},
    {"cesium_iodide_with_ded", F(1.86) * 10.0, F(38.04) * 10.0, F(259.81), F(108.0), F(4.510), true,
     F(0.25381), F(2.6657), F(0.0395), F(3.3353), F(553.1), F(6.2807)},
};

// material_rod.hpp:26-32: the radius takes the place of the thickness
dtr_material_slab make_rod(const Material& m, double radius) {
    dtr_material_slab r = dtr_host::make_slab(m.X0, m.L0, m.A, m.Z, m.rho, radius);
    r.kind = DTR_MATERIAL_ROD;
    return r;
}

// A mask as the builders create it: a shape and its boundary values
struct MaskDef {
    uint8_t shape;
    std::vector<double> v;
    uint16_t link;
};

struct SurfaceDef {
    dtr_transform trf;
    std::vector<MaskDef> masks;  // >1 for multi-mask portals
    uint8_t sf_id;
    int material = -1;           // index into the material list of the emitter
    int material_map = -1;       // index into the material maps of the emitter
    int mask_share = -1;         // share the masks of an earlier surface
};

struct GridDef {
    bool present = false;
    uint8_t type = 0;
    double min0, max0, min1, max1;
    uint32_t n0, n1;
    uint32_t capacity = 1;  // surfaces per bin
};

struct VolumeDef {
    dtr_transform placement = make_translation({0, 0, 0});
    std::vector<SurfaceDef> surfaces;
    GridDef grid;
    int material = -1;  // homogeneous volume material, index into the material table
    bool endcap_layer = false;  // thin cylinder portal maps (build_toy_detector.hpp:919-922)
};

// Axis-aligned bounding box, utils/bounding_volume.hpp
struct AABB {
    double mn[3], mx[3];
};

// Local minimal box of a planar mask (rectangle2D.hpp:179-189,
// trapezoid2D.hpp:209-224), transformed to global (bounding_volume.hpp:237-301)
AABB planar_box(const MaskDef& m, const dtr_transform& tr, double env) {
    double xb, yb;
    if (m.shape == DTR_RECTANGLE2D) {
        xb = m.v[0] + env;
        yb = m.v[1] + env;
    } else {
        xb = (m.v[0] > m.v[1] ? m.v[0] : m.v[1]) + env;
        yb = m.v[2] + env;
    }
    const V3 lmin{-xb, -yb, -env}, lmax{xb, yb, env};
    const double sx = lmax.x - lmin.x, sy = lmax.y - lmin.y, sz = lmax.z - lmin.z;
    const V3 bx = sx * axis(tr, 0), by = sy * axis(tr, 1), bz = sz * axis(tr, 2);
    V3 c[8];
    c[0] = to_global(tr, lmin);
    c[1] = to_global(tr, lmax);
    c[2] = c[0] + bx;
    c[3] = c[0] + by;
    c[4] = c[0] + bz;
    c[5] = c[1] - bx;
    c[6] = c[1] - by;
    c[7] = c[1] - bz;
    const double inv = std::numeric_limits<double>::max();
    AABB b{{inv, inv, inv}, {-inv, -inv, -inv}};
    for (const V3& p : c) {
        b.mn[0] = p.x < b.mn[0] ? p.x : b.mn[0];
        b.mx[0] = p.x > b.mx[0] ? p.x : b.mx[0];
        b.mn[1] = p.y < b.mn[1] ? p.y : b.mn[1];
        b.mx[1] = p.y > b.mx[1] ? p.y : b.mx[1];
        b.mn[2] = p.z < b.mn[2] ? p.z : b.mn[2];
        b.mx[2] = p.z > b.mx[2] ? p.z : b.mx[2];
    }
    return b;
}

AABB world_box(const std::vector<AABB>& boxes, double env) {
    const double inv = std::numeric_limits<double>::max();
    double mn[3] = {inv, inv, inv}, mx[3] = {-inv, -inv, -inv};
    for (const AABB& b : boxes) {
        for (int i = 0; i < 3; ++i) {
            mn[i] = b.mn[i] < mn[i] ? b.mn[i] : mn[i];
            mx[i] = b.mx[i] > mx[i] ? b.mx[i] : mx[i];
        }
    }
    AABB w;
    for (int i = 0; i < 3; ++i) {
        w.mn[i] = mn[i] - env;
        w.mx[i] = mx[i] + env;
    }
    return w;
}

std::vector<AABB> module_boxes(const std::vector<SurfaceDef>& sfs, double env) {
    std::vector<AABB> boxes;
    for (const auto& s : sfs) {
        for (const auto& m : s.masks) boxes.push_back(planar_box(m, s.trf, env));
    }
    return boxes;
}

SurfaceDef cyl_portal(double r, double z0, double z1, uint16_t link) {
    SurfaceDef s;
    s.trf = make_translation({0, 0, 0});
    s.masks.push_back({DTR_CONCENTRIC_CYLINDER2D,
                       {r, std::min(z0, z1), std::max(z0, z1)}, link});
    s.sf_id = DTR_PORTAL;
    return s;
}
SurfaceDef disc_portal(double rin, double rout, double z, uint16_t link) {
    SurfaceDef s;
    s.trf = make_translation({0, 0, z});
    s.masks.push_back({DTR_RING2D, {std::min(rin, rout), std::max(rin, rout)}, link});
    s.sf_id = DTR_PORTAL;
    return s;
}

struct Boundaries {
    double inner_r = 0, outer_r = 0, lower_z = 0, upper_z = 0;
};

// cylinder_portal_generator.hpp:150-285 with autofit around the modules
Boundaries add_cylinder_portal_box(VolumeDef& vol, double env, double fixed_hz,
                                   double fixed_rin, double fixed_rout,
                                   uint16_t north, uint16_t south, uint16_t east,
                                   uint16_t west) {
    const std::vector<AABB> boxes = module_boxes(vol.surfaces, env);
    const AABB w = world_box(boxes, env);
    double center[3];
    for (int i = 0; i < 3; ++i) {
        const double c = 0.5 * (w.mx[i] + w.mn[i]);
        center[i] = (c == std::numeric_limits<double>::max()) ? 0.0 : c;
    }
    const double hx = std::fabs(0.5 * (w.mx[0] - w.mn[0]));
    const double hy = std::fabs(0.5 * (w.mx[1] - w.mn[1]));
    const double hz = std::fabs(0.5 * (w.mx[2] - w.mn[2]));
    double mean = 0.0;
    for (const auto& s : vol.surfaces) mean += perp({s.trf.t[0], s.trf.t[1], s.trf.t[2]});
    const double mean_r = mean / static_cast<double>(vol.surfaces.size());
    Boundaries b;
    b.outer_r = std::max(hx, hy);
    b.inner_r = mean_r - (b.outer_r - mean_r);
    b.lower_z = std::min(center[2] - hz, center[2] + hz);
    b.upper_z = std::max(center[2] - hz, center[2] + hz);
    if (fixed_rin > 0.0) b.inner_r = fixed_rin;
    if (fixed_rout > 0.0) b.outer_r = fixed_rout;
    if (fixed_hz > 0.0) {
        b.lower_z = -fixed_hz;
        b.upper_z = fixed_hz;
    }
    vol.surfaces.push_back(cyl_portal(b.inner_r, b.lower_z, b.upper_z, south));
    vol.surfaces.push_back(cyl_portal(b.outer_r, b.lower_z, b.upper_z, north));
    vol.surfaces.push_back(disc_portal(b.inner_r, b.outer_r, b.lower_z, west));
    vol.surfaces.push_back(disc_portal(b.inner_r, b.outer_r, b.upper_z, east));
    return b;
}

// build_toy_detector.hpp:432-490 add_cylinder_portals
void add_gap_portals(VolumeDef& vol, double lower_z, double upper_z,
                     double inner_r, double outer_r, uint16_t north,
                     uint16_t south, uint16_t east, uint16_t west) {
    const double min_r = std::min(inner_r, outer_r), max_r = std::max(inner_r, outer_r);
    const double min_z = std::min(lower_z, upper_z), max_z = std::max(lower_z, upper_z);
    vol.surfaces.push_back(cyl_portal(min_r, min_z, max_z, south));
    vol.surfaces.push_back(cyl_portal(max_r, min_z, max_z, north));
    vol.surfaces.push_back(disc_portal(min_r, max_r, min_z, west));
    vol.surfaces.push_back(disc_portal(min_r, max_r, max_z, east));
}

struct Extent {
    uint32_t vol;
    double lower, upper;
};

// ---------------------------------------------------------------------------
// Table assembly: volume builders are built in index order; each volume
// pushes its placement transform, then the transforms of its surfaces
// (volume_builder.hpp build/add_to_detector).
// ---------------------------------------------------------------------------
void assemble(const std::vector<VolumeDef>& vols,
              const std::vector<dtr_material_slab>& mats, Tables& T,
              const std::vector<dtr_material_map>& maps = {}) {
    T.materials = mats;
    T.maps = maps;
    for (uint32_t vi = 0; vi < vols.size(); ++vi) {
        const VolumeDef& V = vols[vi];
        dtr_volume dv{};
        dv.material = V.material >= 0 ? static_cast<uint32_t>(V.material) + 1u : 0u;
        dv.transform = static_cast<uint32_t>(T.transforms.size());
        T.transforms.push_back(V.placement);
        dv.sf_begin = static_cast<uint32_t>(T.surfaces.size());
        std::vector<uint32_t> first_mask_of(V.surfaces.size(), 0);
        for (uint32_t k = 0; k < V.surfaces.size(); ++k) {
            const SurfaceDef& S = V.surfaces[k];
            dtr_surface ds{};
            const uint32_t idx = static_cast<uint32_t>(T.surfaces.size());
            ds.transform = static_cast<uint32_t>(T.transforms.size());
            T.transforms.push_back(S.trf);
            if (S.mask_share >= 0) {
                ds.mask = first_mask_of[S.mask_share];
                ds.n_masks = static_cast<uint8_t>(V.surfaces[S.mask_share].masks.size());
                ds.shape = V.surfaces[S.mask_share].masks[0].shape;
            } else {
                ds.mask = static_cast<uint32_t>(T.masks.size());
                ds.n_masks = static_cast<uint8_t>(S.masks.size());
                ds.shape = S.masks[0].shape;
                for (const MaskDef& m : S.masks) {
                    dtr_mask dm{};
                    for (size_t i = 0; i < m.v.size() && i < 7; ++i) dm.v[i] = m.v[i];
                    dm.volume_link = m.link;
                    dm.shape = m.shape;
                    T.masks.push_back(dm);
                }
            }
            first_mask_of[k] = ds.mask;
            ds.material = (S.material_map >= 0) ? (DTR_MATERIAL_MAP | static_cast<uint32_t>(S.material_map))
                          : (S.material >= 0)   ? static_cast<uint32_t>(S.material)
                                                : kInvalid;
            ds.volume = static_cast<uint16_t>(vi);
            ds.sf_id = S.sf_id;
            ds.barcode = make_barcode(vi, S.sf_id, idx, ds.transform);
            T.surfaces.push_back(ds);
        }
        dv.sf_end = static_cast<uint32_t>(T.surfaces.size());
        // Brute force: all surfaces, or only portals/passives if a grid exists
        dv.bf_begin = static_cast<uint32_t>(T.bf.size());
        for (uint32_t s = dv.sf_begin; s < dv.sf_end; ++s) {
            if (!V.grid.present || T.surfaces[s].sf_id != DTR_SENSITIVE) T.bf.push_back(s);
        }
        dv.bf_end = static_cast<uint32_t>(T.bf.size());
        dv.grid = -1;
        if (V.grid.present) {
            dtr_grid g{};
            g.type = V.grid.type;
            g.min0 = V.grid.min0;
            g.max0 = V.grid.max0;
            g.min1 = V.grid.min1;
            g.max1 = V.grid.max1;
            g.n0 = V.grid.n0;
            g.n1 = V.grid.n1;
            g.bins = static_cast<uint32_t>(T.bins.size());
            g.capacity = static_cast<uint8_t>(V.grid.capacity);
            const uint32_t cap = V.grid.capacity;
            T.bins.resize(T.bins.size() + static_cast<size_t>(g.n0) * g.n1 * cap, kInvalid);
            const dtr_transform& vt = V.placement;
            // grid_builder.hpp:232-272: radius between the portal cylinders
            if (g.type == DTR_GRID_CYLINDER) {
                double rmin = std::numeric_limits<double>::max(), rmax = -rmin;
                bool any = false;
                for (uint32_t s = dv.sf_begin; s < dv.sf_end; ++s) {
                    const dtr_surface& ds = T.surfaces[s];
                    if (ds.sf_id != DTR_PORTAL || ds.shape != DTR_CONCENTRIC_CYLINDER2D) continue;
                    const double r = T.masks[ds.mask].v[0];
                    rmin = std::min(rmin, r);
                    rmax = std::max(rmax, r);
                    any = true;
                }
                g.radius = any ? 0.5 * (rmin + rmax) : 0.0;
            }
            // bin_fillers.hpp:65-94 fill_by_pos
            const double w0 = (g.max0 - g.min0) / static_cast<double>(g.n0);
            const double w1 = (g.max1 - g.min1) / static_cast<double>(g.n1);
            for (uint32_t s = dv.sf_begin; s < dv.sf_end; ++s) {
                const dtr_surface& ds = T.surfaces[s];
                if (ds.sf_id != DTR_SENSITIVE) continue;
                const dtr_transform& st = T.transforms[ds.transform];
                const V3 t{st.t[0], st.t[1], st.t[2]};
                double l0, l1;
                if (g.type == DTR_GRID_CYLINDER) {
                    l0 = phi_of(t);
                    l1 = t.z;
                } else {
                    const V3 lp = to_local(vt, t);
                    l0 = perp(lp);
                    l1 = phi_of(lp);
                }
                int b0 = static_cast<int>((l0 - g.min0) / w0 + 1.0) - 1;
                int b1 = static_cast<int>((l1 - g.min1) / w1 + 1.0) - 1;
                // closed axes clamp, circular axes wrap
                auto clamp = [](int b, int n) { return b <= 0 ? 0 : (b >= n ? n - 1 : b); };
                auto wrap = [](int b, int n) { return (n + (b % n)) % n; };
                if (g.type == DTR_GRID_CYLINDER) {
                    b0 = wrap(b0, static_cast<int>(g.n0));
                    b1 = clamp(b1, static_cast<int>(g.n1));
                } else {
                    b0 = clamp(b0, static_cast<int>(g.n0));
                    b1 = wrap(b1, static_cast<int>(g.n1));
                }
                uint32_t* bin = &T.bins[g.bins + (static_cast<size_t>(b0) * g.n1 + b1) * cap];
                if (cap == 1u) {
                    if (bin[0] == kInvalid) bin[0] = s;  // static_array<sf, 1>: first fill wins
                } else {
                    // dynamic bins (populators.hpp attach): appended in surface order
                    uint32_t k = 0;
                    while (k < cap && bin[k] != kInvalid) ++k;
                    if (k == cap) throw std::length_error("grid bin capacity exceeded");
                    bin[k] = s;
                }
            }
            dv.grid = static_cast<int32_t>(T.grids.size());
            T.grids.push_back(g);
        }
        T.volumes.push_back(dv);
    }
}


// ---------------------------------------------------------------------------
// Toy detector (build_toy_detector.hpp)
// ---------------------------------------------------------------------------
struct ToyCfg {
    unsigned n_brl = 4, n_edc = 3;
    bool material = true;
    bool maps = false;  // toy_det_config::use_material_maps
    double outer_radius = F(180.0);
    double beampipe_vol_radius = F(25.0);
    double envelope = F(2.0);
    double beampipe_thickness = F(0.8);
    double module_thickness = F(1.5);
    std::vector<double> brl_radii{F(19.0), F(32.0), F(72.0), F(116.0), F(172.0)};
    std::vector<std::pair<unsigned, unsigned>> brl_binning{
        {0u, 0u}, {16u, 14u}, {32u, 14u}, {52u, 14u}, {78u, 14u}};
    std::vector<double> edc_positions{F(600.0), F(700.0), F(820.0), F(960.0),
                                      F(1100.0), F(1300.0), F(1500.0)};
    // barrel_generator_config (build_toy_detector.hpp:56-61)
    double brl_half_z = F(500.0);
    double brl_hx = F(8.4), brl_hy = F(36.0);
    double tilt_phi = F(0.14), radial_stagger = F(0.5), z_overlap = F(2.0);
    // endcap_generator_config (build_toy_detector.hpp:64-75)
    std::vector<std::vector<double>> edc_bounds{{F(3.0), F(9.5), F(39.0)},
                                                {F(6.0), F(10.0), F(39.0)}};
    double ring_stagger = F(2.0);
    std::vector<double> phi_stagger{F(4.0), F(4.0)};
    std::vector<double> phi_sub_stagger{F(0.5), F(0.5)};
    std::vector<unsigned> edc_binning{40u, 68u};
};

// barrel_generator.hpp:110-190
void barrel_modules(VolumeDef& vol, const ToyCfg& c, double radius,
                    unsigned n_phi, unsigned n_z, int mat) {
    const double phi_step = 2.0 * kPi / static_cast<double>(n_phi);
    const double min_phi = -kPi + 0.5 * phi_step;
    const double z_start = -0.5 * static_cast<double>(n_z - 1u) * (2.0 * c.brl_hy - c.z_overlap);
    const double z_step = (std::fabs(z_start) - z_start) / static_cast<double>(n_z - 1u);
    std::vector<V3> centers;
    for (unsigned zb = 0; zb < n_z; ++zb) {
        const double mod_z = z_start + static_cast<double>(zb) * z_step;
        const double mod_r = (zb % 2u) != 0u ? radius - 0.5 * c.radial_stagger
                                             : radius + 0.5 * c.radial_stagger;
        for (unsigned pb = 0; pb < n_phi; ++pb) {
            const double mod_phi = min_phi + static_cast<double>(pb) * phi_step;
            centers.push_back({mod_r * std::cos(mod_phi), mod_r * std::sin(mod_phi), mod_z});
        }
    }
    const int first = static_cast<int>(vol.surfaces.size());
    for (const V3& mc : centers) {
        SurfaceDef s;
        const double mod_phi = phi_of(mc);
        const V3 lz{std::cos(mod_phi + c.tilt_phi), std::sin(mod_phi + c.tilt_phi), 0.0};
        const V3 lx{-std::sin(mod_phi + c.tilt_phi), std::cos(mod_phi + c.tilt_phi), 0.0};
        s.trf = make_trf_zx(mc, lz, lx);
        s.sf_id = DTR_SENSITIVE;
        s.material = mat;
        if (static_cast<int>(vol.surfaces.size()) == first) {
            s.masks.push_back({DTR_RECTANGLE2D, {c.brl_hx, c.brl_hy}, 0});
        } else {
            s.mask_share = first;
            s.masks = vol.surfaces[first].masks;
        }
        vol.surfaces.push_back(s);
    }
}

// endcap_generator.hpp:140-300
void endcap_modules(VolumeDef& vol, const ToyCfg& c, double center_abs,
                    double inner_r, double outer_r, int side, int mat) {
    std::vector<double> radii;
    const double delta_r = outer_r - inner_r;
    double tot_length = 0.0;
    for (const auto& b : c.edc_bounds) tot_length += 2.0 * b[2] + 1.0 * kMM;
    const double r_overlap = (tot_length - delta_r) / static_cast<double>(c.edc_bounds.size());
    double prev_r = inner_r + r_overlap, prev_hl = 0.0, prev_ol = 0.0;
    for (const auto& b : c.edc_bounds) {
        const double hl = b[2];
        radii.push_back(prev_r + prev_hl - prev_ol + hl);
        prev_r = radii.back();
        prev_ol = 2.0 * r_overlap;
        prev_hl = hl;
    }
    for (unsigned ir = 0; ir < radii.size(); ++ir) {
        const double center = center_abs;
        const double stag = (ir % 2u) ? center + 0.5 * c.ring_stagger
                                      : center - 0.5 * c.ring_stagger;
        const double rz = radii.size() == 1u ? center : stag;
        const double sub = c.phi_sub_stagger.size() > 1u ? c.phi_sub_stagger[ir] : 0.0;
        const unsigned nphi = c.edc_binning[ir];
        const double phi_step = 2.0 * kPi / static_cast<double>(nphi);
        const double min_phi = -kPi + 0.5 * phi_step;
        const int first = static_cast<int>(vol.surfaces.size());
        for (unsigned ip = 0; ip < nphi; ++ip) {
            double rzs = 0.0;
            if (sub != 0.0 && !(nphi % 4u)) {
                if (!(ip % 4u)) rzs = sub;
                else if (!((ip + 1u) % 4u)) rzs = -sub;
            }
            const double phi = min_phi + static_cast<double>(ip) * phi_step;
            const double z = (ip % 2u) ? rz - 0.5 * c.phi_stagger[ir] : rz + 0.5 * c.phi_stagger[ir];
            V3 pos{radii[ir] * std::cos(phi), radii[ir] * std::sin(phi), z + rzs};
            V3 mc = pos;
            mc.z *= static_cast<double>(side);
            const double mphi = phi_of(pos);
            const V3 ly{std::cos(mphi), std::sin(mphi), 0.0};
            const V3 lz{0.0, 0.0, static_cast<double>(side)};
            const V3 lx = cross(ly, lz);
            SurfaceDef s;
            s.trf = make_trf_zx(mc, lz, lx);
            s.sf_id = DTR_SENSITIVE;
            s.material = mat;
            if (ip == 0) {
                const auto& b = c.edc_bounds[ir];
                s.masks.push_back({DTR_TRAPEZOID2D, {b[0], b[1], b[2], 1.0 / (2.0 * b[2])}, 0});
            } else {
                s.mask_share = first;
                s.masks = vol.surfaces[first].masks;
            }
            vol.surfaces.push_back(s);
        }
    }
}

std::vector<Extent> add_endcap(std::vector<VolumeDef>& vols, ToyCfg& c, int side,
                               uint32_t beampipe_idx, int si_mat) {
    std::vector<Extent> sizes;
    const uint32_t connector_link = static_cast<uint32_t>(vols.size()) + 1u;
    const double sign = static_cast<double>(side);
    const uint16_t link_north = kEndOfWorld;
    const uint16_t link_south = static_cast<uint16_t>(beampipe_idx);
    const double inner_radius = c.beampipe_vol_radius;
    const double outer_radius = c.outer_radius;
    double gap_east_z = sign * c.brl_half_z;
    Boundaries vb;
    bool is_gap = true;
    for (uint32_t i = 0; i < 2u * c.n_edc; ++i) {
        vols.emplace_back();
        const uint32_t vol_idx = static_cast<uint32_t>(vols.size()) - 1u;
        VolumeDef& vol = vols.back();
        if (i == 1u) {
            const double gap_west_z = sign * std::min(std::fabs(vb.upper_z), std::fabs(vb.lower_z));
            sizes.push_back({vol_idx, gap_east_z, gap_west_z});
            gap_east_z = sign * std::max(std::fabs(vb.upper_z), std::fabs(vb.lower_z));
            is_gap = !is_gap;
            continue;
        }
        is_gap = !is_gap;
        if (is_gap) {
            const uint16_t east = static_cast<uint16_t>(static_cast<int>(vol_idx) + side - 2);
            const uint16_t west = static_cast<uint16_t>(static_cast<int>(vol_idx) - side - 2);
            const double gap_west_z = sign * std::min(std::fabs(vb.upper_z), std::fabs(vb.lower_z));
            vol.placement = make_translation({0.0, 0.0, 0.5 * (gap_east_z + gap_west_z)});
            add_gap_portals(vol, gap_west_z, gap_east_z, inner_radius, outer_radius,
                            link_north, link_south, east, west);
            sizes.push_back({vol_idx, gap_east_z, gap_west_z});
            gap_east_z = sign * std::max(std::fabs(vb.upper_z), std::fabs(vb.lower_z));
        } else {
            const uint32_t j = i / 2u;
            uint32_t east = static_cast<uint32_t>(static_cast<int>(vol_idx) + side + 2);
            uint32_t west = static_cast<uint32_t>(static_cast<int>(vol_idx) - side + 2);
            if (sign < 0) {
                east = (i == 0u) ? connector_link : east;
                west = (j == c.n_edc - 1u) ? kEndOfWorld : west;
            } else {
                west = (i == 0u) ? connector_link : west;
                east = (j == c.n_edc - 1u) ? kEndOfWorld : east;
            }
            const double center_z = sign * c.edc_positions.at(j);
            vol.placement = make_translation({0.0, 0.0, center_z});
            vol.endcap_layer = true;
            endcap_modules(vol, c, std::fabs(center_z), inner_radius,
                           outer_radius - c.envelope, side, si_mat);
            vb = add_cylinder_portal_box(vol, c.envelope, 0.0, inner_radius, outer_radius,
                                         link_north, link_south,
                                         static_cast<uint16_t>(east),
                                         static_cast<uint16_t>(west));
            sizes.push_back({vol_idx, vb.lower_z, vb.upper_z});
            vol.grid.present = true;
            vol.grid.type = DTR_GRID_RING;
            vol.grid.min0 = inner_radius;
            vol.grid.max0 = outer_radius;
            vol.grid.min1 = -kPi;
            vol.grid.max1 = kPi;
            vol.grid.n0 = static_cast<uint32_t>(c.edc_binning.size());
            vol.grid.n1 = c.edc_binning.back();
        }
    }
    return sizes;
}

std::vector<Extent> add_barrel(std::vector<VolumeDef>& vols, ToyCfg& c,
                               uint32_t beampipe_idx, int si_mat) {
    std::vector<Extent> sizes;
    uint16_t east = kEndOfWorld, west = kEndOfWorld;
    const uint32_t nv = static_cast<uint32_t>(vols.size());
    if (c.n_edc > 0) {
        east = static_cast<uint16_t>(nv + 2u * c.n_brl + 2u);
        west = static_cast<uint16_t>(nv - 2u * c.n_edc + 1u);
    }
    const double h_z = c.brl_half_z;
    double gap_inner_r = c.beampipe_vol_radius;
    Boundaries vb;
    bool is_gap = true;
    for (uint32_t i = 0; i < 2u * c.n_brl; ++i) {
        vols.emplace_back();
        const uint32_t vol_idx = static_cast<uint32_t>(vols.size()) - 1u;
        VolumeDef& vol = vols.back();
        is_gap = !is_gap;
        if (is_gap) {
            uint32_t north = vol_idx - 1u;
            uint32_t south = vol_idx - 3u;
            south = (i == 1u) ? beampipe_idx : south;
            add_gap_portals(vol, -h_z, h_z, gap_inner_r, vb.inner_r,
                            static_cast<uint16_t>(north), static_cast<uint16_t>(south), east, west);
            sizes.push_back({vol_idx, gap_inner_r, vb.inner_r});
            gap_inner_r = vb.outer_r;
        } else {
            const uint32_t north = std::min(vol_idx + 3u, 2u * c.n_edc + 2u * c.n_brl + 1u);
            const uint32_t south = vol_idx + 1u;
            const unsigned j = (i + 2u) / 2u;
            const auto bin = c.brl_binning.at(j);
            barrel_modules(vol, c, c.brl_radii.at(j), bin.first, bin.second, si_mat);
            vb = add_cylinder_portal_box(vol, c.envelope, h_z, 0.0, 0.0,
                                         static_cast<uint16_t>(north),
                                         static_cast<uint16_t>(south), east, west);
            sizes.push_back({vol_idx, vb.inner_r, vb.outer_r});
            vol.grid.present = true;
            vol.grid.type = DTR_GRID_CYLINDER;
            vol.grid.min0 = -kPi;
            vol.grid.max0 = kPi;
            vol.grid.min1 = -h_z;
            vol.grid.max1 = h_z;
            vol.grid.n0 = bin.first;
            vol.grid.n1 = bin.second;
        }
    }
    vols.emplace_back();
    const uint32_t vol_idx = static_cast<uint32_t>(vols.size()) - 1u;
    add_gap_portals(vols.back(), -h_z, h_z, vb.outer_r, c.outer_radius, kEndOfWorld,
                    static_cast<uint16_t>(vol_idx - 2u), east, west);
    sizes.push_back({vol_idx, vb.outer_r, c.outer_radius});
    return sizes;
}

// build_toy_detector.hpp:1146-1210
void add_beampipe_edc_portals(VolumeDef& bp, const ToyCfg& c,
                              const std::vector<Extent>& edc) {
    const double side = std::copysign(1.0, edc.front().lower);
    const double outer_r = c.beampipe_vol_radius;
    double disc_pos_z = -side * std::numeric_limits<double>::max();
    SurfaceDef cyl;
    cyl.trf = make_translation({0, 0, 0});
    cyl.sf_id = DTR_PORTAL;
    for (const Extent& e : edc) {
        const double min_z = std::min(e.lower, e.upper), max_z = std::max(e.lower, e.upper);
        if (side < 0.0) disc_pos_z = (disc_pos_z > min_z) ? min_z : disc_pos_z;
        else disc_pos_z = (disc_pos_z < max_z) ? max_z : disc_pos_z;
        cyl.masks.push_back({DTR_CONCENTRIC_CYLINDER2D, {outer_r, min_z, max_z},
                             static_cast<uint16_t>(e.vol)});
    }
    bp.surfaces.push_back(disc_portal(0.0, outer_r, disc_pos_z, kEndOfWorld));
    bp.surfaces.push_back(cyl);
}

// build_toy_detector.hpp:999-1081
void add_connector(std::vector<VolumeDef>& vols, const ToyCfg& c, uint32_t beampipe_idx,
                   const std::vector<Extent>& edc, const std::vector<Extent>& brl) {
    const Extent& gap = edc[1];
    const double side = std::copysign(1.0, gap.lower);
    const double east_z = gap.lower, west_z = gap.upper;
    const double min_z = std::min(east_z, west_z), max_z = std::max(east_z, west_z);
    VolumeDef& vol = vols[gap.vol];
    vol.placement = make_translation({0.0, 0.0, 0.5 * (east_z + west_z)});
    SurfaceDef facing;
    facing.trf = make_translation({0.0, 0.0, (side < 0.0) ? max_z : min_z});
    facing.sf_id = DTR_PORTAL;
    for (const Extent& e : brl) {
        facing.masks.push_back({DTR_RING2D, {std::min(e.lower, e.upper), std::max(e.lower, e.upper)},
                                static_cast<uint16_t>(e.vol)});
    }
    vol.surfaces.push_back(facing);
    vol.surfaces.push_back(disc_portal(c.beampipe_vol_radius, c.outer_radius,
                                       (side < 0.0) ? min_z : max_z,
                                       static_cast<uint16_t>(gap.vol - 1u)));
    vol.surfaces.push_back(cyl_portal(c.beampipe_vol_radius, min_z, max_z,
                                      static_cast<uint16_t>(beampipe_idx)));
    vol.surfaces.push_back(cyl_portal(c.outer_radius, min_z, max_z, kEndOfWorld));
}

// predefined_materials.hpp:93-97
Material aluminium() {
    return {F(88.97) * kMM, F(397.0) * kMM, F(26.98), F(13.0), 2.699 * kGram / kCm3};
}

// mixture.hpp:33-80: fraction-weighted A, Z and mass density, X0 and L0
// from the fraction-weighted inverse lengths
Material mixture(const Material& a, double fa, const Material& b, double fb) {
    return {1.0 / (fa / a.X0 + fb / b.X0), 1.0 / (fa / a.L0 + fb / b.L0), a.A * fa + b.A * fb,
            a.Z * fa + b.Z * fb, a.rho * fa + b.rho * fb};
}

// Material maps of the toy detector (toy_det_config::use_material_maps,
// build_toy_detector.hpp:84-110, 371-409; material_map_generator.hpp:
// 34-83, 177-253): every portal and the beampipe carry a map instead of the
// module slabs.  Cylinder portals: 20 phi x 20 z bins of Si(90 %)/Al(10 %),
// thickness 5 mm (0.15 mm on the portals of endcap layer volumes) + 1e-8 z^2
// per z bin; disc portals: 5 r x 20 phi bins of the same mixture, 1 mm +
// 1e-4 r per r bin; the beampipe: 20 x 20 bins of 0.8 mm beryllium.  The
// bin positions z (r) run from the lowest to the highest edge over all masks
// of the surface in n - 1 accumulated steps.  The map axes are the merged
// mask extents: (phi in [-pi, pi] circular, z closed) on cylinders, (r
// closed, phi circular) on discs (grid_factory.hpp:243-280, 356-400).
void toy_material_maps(std::vector<VolumeDef>& vols, std::vector<dtr_material_slab>& mats,
                       std::vector<dtr_material_map>& maps) {
    const Material mix = mixture(silicon_tml(), 0.9, aluminium(), 0.1);
    for (VolumeDef& vol : vols) {
        for (SurfaceDef& S : vol.surfaces) {
            const std::vector<MaskDef>& ms = S.mask_share >= 0 ? vol.surfaces[S.mask_share].masks : S.masks;
            const uint8_t shape = ms[0].shape;
            const bool cyl = shape == DTR_CONCENTRIC_CYLINDER2D;
            if (!(cyl || shape == DTR_RING2D)) continue;
            if (S.sf_id != DTR_PORTAL && !(cyl && S.sf_id == DTR_PASSIVE)) continue;
            const bool beampipe = S.sf_id == DTR_PASSIVE;
            const Material mat = beampipe ? beryllium_tml() : mix;
            const double t = beampipe ? F(0.8) : cyl ? (vol.endcap_layer ? F(0.15) : F(5.0)) : F(1.0);
            const double scalor = beampipe ? 0.0 : cyl ? static_cast<double>(1e-8f) : static_cast<double>(1e-4f);
            dtr_material_map m{};
            double lo = std::numeric_limits<double>::max(), hi = cyl ? -lo : 0.0;
            for (const MaskDef& mk : ms) {
                lo = std::min(lo, cyl ? mk.v[1] : mk.v[0]);
                hi = std::max(hi, cyl ? mk.v[2] : mk.v[1]);
            }
            const uint32_t nvar = cyl ? 20u : 5u;
            std::vector<double> thick;
            double x = lo;
            const double dx = (cyl ? std::fabs(hi - lo) : (hi - lo)) / static_cast<double>(nvar - 1u);
            for (uint32_t k = 0; k < nvar; ++k) {
                thick.push_back(cyl ? scalor * x * x + t : scalor * x + t);
                x += dx;
            }
            if (cyl) {
                m.n0 = 20u;
                m.n1 = 20u;
                m.min0 = -kPi;
                m.max0 = kPi;
                m.circular0 = 1u;
                m.min1 = lo;
                m.max1 = hi;
            } else {
                m.n0 = 5u;
                m.n1 = 20u;
                m.min0 = lo;
                m.max0 = hi;
                m.min1 = -kPi;
                m.max1 = kPi;
                m.circular1 = 1u;
            }
            m.slabs = static_cast<uint32_t>(mats.size());
            for (uint32_t b0 = 0; b0 < m.n0; ++b0)
                for (uint32_t b1 = 0; b1 < m.n1; ++b1) mats.push_back(make_slab(mat, thick[cyl ? b1 : b0]));
            S.material = -1;
            S.material_map = static_cast<int>(maps.size());
            maps.push_back(m);
        }
    }
}

void build_toy(ToyCfg c, Tables& T) {
    std::vector<VolumeDef> vols;
    std::vector<dtr_material_slab> mats;
    int si_mat = -1, be_mat = -1;
    if (c.material) {
        // One slab per material surface, as homogeneous_material_generator
        // does; identical values, so a single entry per kind suffices.
        mats.push_back(make_slab(silicon_tml(), c.module_thickness));
        mats.push_back(make_slab(beryllium_tml(), c.beampipe_thickness));
        si_mat = 0;
        be_mat = 1;
    }
    vols.emplace_back();
    const uint32_t bp_idx = 0;
    {
        const double max_z = c.n_edc == 0u ? c.brl_half_z : c.edc_positions.at(c.n_edc - 1u);
        SurfaceDef s = cyl_portal(c.brl_radii.at(0), -max_z, max_z, static_cast<uint16_t>(bp_idx));
        s.sf_id = DTR_PASSIVE;
        s.material = be_mat;
        vols[0].surfaces.push_back(s);
    }
    std::vector<Extent> neg, brl, pos;
    if (c.n_edc > 0u) {
        neg = add_endcap(vols, c, -1, bp_idx, si_mat);
        add_beampipe_edc_portals(vols[0], c, neg);
    }
    if (c.n_brl > 0u) brl = add_barrel(vols, c, bp_idx, si_mat);
    {
        // build_toy_detector.hpp:1094-1142 (barrel section of the beampipe)
        const double h_z = c.brl_half_z;
        const double outer_r = c.beampipe_vol_radius;
        if (c.n_edc == 0u) {
            vols[0].surfaces.push_back(disc_portal(0.0, outer_r, -h_z, kEndOfWorld));
            vols[0].surfaces.push_back(disc_portal(0.0, outer_r, h_z, kEndOfWorld));
        }
        const uint16_t first_barrel = c.n_brl == 0u ? kEndOfWorld
                                                    : static_cast<uint16_t>(2u * c.n_edc + 2u);
        const double eps = std::numeric_limits<double>::epsilon();
        vols[0].surfaces.push_back(cyl_portal(outer_r, -h_z + eps, h_z - eps, first_barrel));
    }
    if (c.n_edc > 0u) {
        pos = add_endcap(vols, c, 1, bp_idx, si_mat);
        add_beampipe_edc_portals(vols[0], c, pos);
        add_connector(vols, c, bp_idx, neg, brl);
        add_connector(vols, c, bp_idx, pos, brl);
    }
    std::vector<dtr_material_map> maps;
    if (c.maps) toy_material_maps(vols, mats, maps);
    assemble(vols, mats, T, maps);
}

// ---------------------------------------------------------------------------
// Wire chamber (build_wire_chamber.hpp:150-331, wire_layer_generator.hpp)
// The wire layer generator works in the reference's default scalar (float):
// the number of wires per layer comes from a float phi accumulation, so the
// wire placement is computed in float and stored as double.
// ---------------------------------------------------------------------------
void wire_layer(VolumeDef& vol, const dtr_wire_chamber_config& cfg, float inner_layer_rad, float stereo,
                uint16_t vol_idx, int mat) {
    const float cell = static_cast<float>(cfg.cell_size);
    const float half_z = static_cast<float>(cfg.half_z);
    const float central = inner_layer_rad + cell;
    const float two_pi = 2.f * static_cast<float>(kPi);
    const float delta = 2.f * cell / central;
    std::vector<std::array<float, 3>> centers;
    float phi = 0.f;
    while (phi <= two_pi) {
        centers.push_back({central * std::cos(phi), central * std::sin(phi), 0.f});
        phi += delta;
    }
    const float cs = std::cos(stereo), sn = std::sin(stereo);
    for (const auto& c : centers) {
        // r axis, then the z axis rotated about it by the stereo angle
        // (axis_rotation.hpp:33-46 applied to (0, 0, 1))
        const float inv = 1.f / std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const float ux = c[0] * inv, uy = c[1] * inv, uz = c[2] * inv;
        const float zx = cs * 0.f + sn * uy + (1.f - cs) * (ux * uz);
        const float zy = cs * 0.f + sn * -ux + (1.f - cs) * (uy * uz);
        const float zz = cs * 1.f + sn * 0.f + (1.f - cs) * (uz * uz);
        // unit_vectors.hpp:28-48 make_curvilinear_unit_u
        float xx = -zy, xy = zx;
        const float scale = std::sqrt(xx * xx + xy * xy);
        if (scale < 1e-6f) {
            xx = 1.f;
            xy = 0.f;
        } else {
            xx = xx * (1.f / scale);
            xy = xy * (1.f / scale);
        }
        // transform3(t, z, x): y = z x x
        const float yx = zy * 0.f - zz * xy, yy = zz * xx - zx * 0.f, yz = zx * xy - zy * xx;
        SurfaceDef sf;
        sf.trf = make_trf({c[0], c[1], c[2]}, {xx, xy, 0.0}, {yx, yy, yz}, {zx, zy, zz});
        sf.sf_id = DTR_SENSITIVE;
        sf.material = mat;
        sf.masks.push_back({static_cast<uint8_t>(cfg.square_cells ? DTR_LINE_SQUARE : DTR_LINE_CIRCULAR),
                            {cell, half_z}, vol_idx});
        vol.surfaces.push_back(sf);
    }
}

void build_wire_chamber(const dtr_wire_chamber_config& cfg, Tables& T) {
    std::vector<VolumeDef> vols(1u + cfg.n_layers);
    std::vector<dtr_material_slab> mats;
    int rod = -1;
    if (cfg.mat_radius > 0.0) {
        mats.push_back(make_rod(tungsten(), cfg.mat_radius));
        rod = 0;
    }
    const double half_z = cfg.half_z;
    const float inner_rad = static_cast<float>(cfg.first_layer_inner_radius);
    const float cell = static_cast<float>(cfg.cell_size);
    // inner volume: no inner portal (build_inner(false))
    {
        VolumeDef& v = vols[0];
        v.surfaces.push_back(cyl_portal(inner_rad, -half_z, half_z, 1u));
        v.surfaces.push_back(disc_portal(0.0, inner_rad, -half_z, kEndOfWorld));
        v.surfaces.push_back(disc_portal(0.0, inner_rad, half_z, kEndOfWorld));
    }
    float stereo = static_cast<float>(cfg.stereo_angle);
    for (uint32_t i = 0; i < cfg.n_layers; ++i) {
        const uint16_t vi = static_cast<uint16_t>(i + 1u);
        VolumeDef& v = vols[vi];
        const float in_r = inner_rad + static_cast<float>(i) * 2.f * cell;
        const float out_r = inner_rad + static_cast<float>(i + 1u) * 2.f * cell;
        const uint16_t north = (i == cfg.n_layers - 1u) ? kEndOfWorld : static_cast<uint16_t>(vi + 1u);
        add_gap_portals(v, -half_z, half_z, in_r, out_r, north, static_cast<uint16_t>(vi - 1u), kEndOfWorld,
                        kEndOfWorld);
        // the sign alternates from layer to layer; the magnitude is kept
        stereo = ((i % 2u == 0u) ? 1.f : -1.f) * std::fabs(stereo);
        wire_layer(v, cfg, in_r, stereo, vi, rod);
        v.grid.present = true;
        v.grid.type = DTR_GRID_CYLINDER;
        v.grid.min0 = -F(kPi);
        v.grid.max0 = F(kPi);
        v.grid.min1 = -half_z;
        v.grid.max1 = half_z;
        v.grid.n0 = 100u;
        v.grid.n1 = 1u;
        v.grid.capacity = 3u;
    }
    assemble(vols, mats, T);
}

// ---------------------------------------------------------------------------
// Telescope (build_telescope_detector.hpp, cuboid_portal_generator.hpp)
// ---------------------------------------------------------------------------
void build_telescope(const dtr_telescope_config& cfg, Tables& T) {
    std::vector<VolumeDef> vols(1);
    std::vector<dtr_material_slab> mats;
    int mat = -1;
    if (cfg.thickness > 0.0) {
        mats.push_back(make_slab(silicon_tml(), cfg.thickness));
        mat = 0;
    }
    VolumeDef& vol = vols[0];
    // homogeneous_volume_material_builder decoration
    // (build_telescope_detector.hpp:262-267)
    if (cfg.volume_material) {
        dtr_material_slab vm = *cfg.volume_material;
        vm.kind = DTR_MATERIAL_SLAB;
        vm.thickness = 0.0;
        vm.thickness_in_X0 = 0.0;
        vol.material = static_cast<int>(mats.size());
        mats.push_back(vm);
    }
    for (uint32_t i = 0; i < cfg.n_planes; ++i) {
        const double s = cfg.positions[i];
        SurfaceDef sf;
        const V3 pos{0.0, 0.0, s};
        const V3 lz{0.0, 0.0, 1.0};
        V3 lx{-lz.y, lz.x, 0.0};
        const double scale = std::sqrt(lx.x * lx.x + lx.y * lx.y + lx.z * lx.z);
        if (scale < F(1e-6)) lx = {1.0, 0.0, 0.0};
        else lx = (1.0 / scale) * lx;
        sf.trf = make_trf_zx(pos, lz, lx);
        sf.sf_id = DTR_SENSITIVE;
        sf.material = mat;
        sf.masks.push_back({DTR_RECTANGLE2D, {cfg.half_x, cfg.half_y}, 0});
        vol.surfaces.push_back(sf);
    }
    const double env = cfg.envelope;
    const AABB w = world_box(module_boxes(vol.surfaces, env), env);
    V3 center;
    {
        double c3[3];
        for (int i = 0; i < 3; ++i) c3[i] = 0.5 * (w.mx[i] + w.mn[i]);
        center = {c3[0], c3[1], c3[2]};
    }
    const double hx = std::fabs(0.5 * (w.mx[0] - w.mn[0]));
    const double hy = std::fabs(0.5 * (w.mx[1] - w.mn[1]));
    const double hz = std::fabs(0.5 * (w.mx[2] - w.mn[2]));
    auto portal = [&](const dtr_transform& tr, double a, double b, int share) {
        SurfaceDef s;
        s.trf = tr;
        s.sf_id = DTR_PORTAL;
        s.masks.push_back({DTR_RECTANGLE2D, {a, b}, kEndOfWorld});
        s.mask_share = share;
        vol.surfaces.push_back(s);
    };
    int first = static_cast<int>(vol.surfaces.size());
    portal(make_translation(center + V3{0, 0, hz}), hx, hy, -1);
    portal(make_translation(center - V3{0, 0, hz}), hx, hy, first);
    first = static_cast<int>(vol.surfaces.size());
    portal(make_trf_zx(center + V3{0, hy, 0}, {0, -1, 0}, {1, 0, 0}), hx, hz, -1);
    portal(make_trf_zx(center - V3{0, hy, 0}, {0, -1, 0}, {1, 0, 0}), hx, hz, first);
    first = static_cast<int>(vol.surfaces.size());
    portal(make_trf_zx(center + V3{hx, 0, 0}, {1, 0, 0}, {0, 0, -1}), hz, hy, -1);
    portal(make_trf_zx(center - V3{hx, 0, 0}, {1, 0, 0}, {0, 0, -1}), hz, hy, first);
    assemble(vols, mats, T);
}

// ---------------------------------------------------------------------------
// Analytic finite solenoid: a current sheet of radius R between z = -L..L.
// Field from the published closed form with complete elliptic integrals
// (e.g. Callaghan & Maslen, NASA TN D-465 (1960); Derby & Olbert,
// Am. J. Phys. 78 (2010) 229), normalised to b_center on the axis at z = 0.
// ---------------------------------------------------------------------------
// Generalised complete elliptic integral by the Bulirsch cel algorithm:
// cel(kc, p, a, b) = int_0^{pi/2} (a cos^2 + b sin^2) / ((cos^2 + p sin^2)
//                    sqrt(cos^2 + kc^2 sin^2)) dphi
double cel(double kc, double p, double c, double s) {
    if (kc == 0.0) return std::numeric_limits<double>::infinity();
    const double errtol = 1e-12;
    double k = std::fabs(kc), pp = p, cc = c, ss = s, em = 1.0;
    if (p > 0.0) {
        pp = std::sqrt(p);
        ss = s / pp;
    } else {
        double f = kc * kc;
        double q = 1.0 - f;
        const double g = 1.0 - pp;
        f = f - pp;
        q = q * (ss - c * pp);
        pp = std::sqrt(f / g);
        cc = (c - ss) / g;
        ss = -q / (g * g * pp) + cc * pp;
    }
    double f = cc;
    cc = cc + ss / pp;
    double g = k / pp;
    ss = 2.0 * (ss + f * g);
    pp = g + pp;
    g = em;
    em = k + em;
    double kk = k;
    for (int it = 0; it < 200 && std::fabs(g - k) > g * errtol; ++it) {
        k = 2.0 * std::sqrt(kk);
        kk = k * em;
        f = cc;
        cc = cc + ss / pp;
        g = kk / pp;
        ss = 2.0 * (ss + f * g);
        pp = g + pp;
        g = em;
        em = k + em;
    }
    return 0.5 * kPi * (ss + cc * em) / (em * (em + pp));
}

// Field of a finite solenoid (unit surface current n*I*mu0 = 1) at (rho, z)
void solenoid_unit(double R, double L, double rho, double z, double& br, double& bz) {
    br = 0.0;
    bz = 0.0;
    for (int s = 0; s < 2; ++s) {
        const double zeta = (s == 0) ? z + L : z - L;
        const double sgn = (s == 0) ? 1.0 : -1.0;
        const double alpha = 1.0 / std::sqrt(zeta * zeta + (rho + R) * (rho + R));
        const double beta = zeta * alpha;
        const double gamma = (R - rho) / (R + rho);
        const double k2 = (zeta * zeta + (R - rho) * (R - rho)) /
                          (zeta * zeta + (R + rho) * (R + rho));
        const double kc = std::sqrt(std::max(k2, 0.0));
        // Derby & Olbert eqs. (3)-(4) (divided by mu0 n I / pi)
        br += sgn * alpha * cel(kc, 1.0, 1.0, -1.0) * R;
        bz += sgn * beta * cel(kc, gamma * gamma, 1.0, gamma) * R / (R + rho);
    }
    br /= kPi;
    bz /= kPi;
}

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
// Uniform in (0, 1) from track index and stream
inline double u01(uint64_t seed, uint64_t track, uint64_t stream) {
    const uint64_t h = splitmix64(splitmix64(seed ^ (track * 0x100000001b3ull)) + stream);
    return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
inline double gauss(uint64_t seed, uint64_t track, uint64_t stream) {
    const double u1 = u01(seed, track, stream), u2 = u01(seed, track, stream + 1000);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

}  // namespace

extern "C" {

int dtr_build_toy_detector(const dtr_toy_config* cfg, dtr_detector_desc* out) {
    if (!cfg || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->n_edc_layers > 7u || cfg->n_brl_layers > 4u) return DTR_ERROR_INVALID_ARGUMENT;
    if (cfg->n_edc_layers > 0u && cfg->n_brl_layers < 4u) return DTR_ERROR_INVALID_ARGUMENT;
    try {
        ToyCfg c;
        c.n_brl = cfg->n_brl_layers;
        c.n_edc = cfg->n_edc_layers;
        if (cfg->use_material > 2u) return DTR_ERROR_INVALID_ARGUMENT;
        c.material = cfg->use_material == 1u;
        c.maps = cfg->use_material == 2u;
        Tables T;
        build_toy(c, T);
        export_tables(T, out);
    } catch (const std::bad_alloc&) {
        return DTR_ERROR_OUT_OF_MEMORY;
    } catch (...) {
        return DTR_ERROR_INVALID_ARGUMENT;
    }
    return DTR_SUCCESS;
}

int dtr_build_wire_chamber(const dtr_wire_chamber_config* cfg, dtr_detector_desc* out) {
    if (!cfg || !out || cfg->n_layers == 0u || cfg->n_layers > 1000u || !(cfg->first_layer_inner_radius > 0.0) ||
        !(cfg->half_z > 0.0) || !(cfg->cell_size > 0.0) || cfg->mat_radius < 0.0)
        return DTR_ERROR_INVALID_ARGUMENT;
    try {
        Tables T;
        build_wire_chamber(*cfg, T);
        export_tables(T, out);
    } catch (const std::bad_alloc&) {
        return DTR_ERROR_OUT_OF_MEMORY;
    } catch (const std::length_error&) {
        return DTR_ERROR_UNSUPPORTED;
    }
    return DTR_SUCCESS;
}

int dtr_build_telescope_detector(const dtr_telescope_config* cfg, dtr_detector_desc* out) {
    if (!cfg || !out || cfg->n_planes == 0u || !cfg->positions || !(cfg->envelope > 0.0))
        return DTR_ERROR_INVALID_ARGUMENT;
    try {
        Tables T;
        build_telescope(*cfg, T);
        export_tables(T, out);
    } catch (const std::bad_alloc&) {
        return DTR_ERROR_OUT_OF_MEMORY;
    }
    return DTR_SUCCESS;
}

int dtr_predefined_material(const char* name, double thickness, dtr_material_slab* out) {
    if (!name || !out || !(thickness >= 0.0)) return DTR_ERROR_INVALID_ARGUMENT;
    for (const NamedMaterial& m : kMaterials) {
        if (std::strcmp(m.name, name) != 0) continue;
        *out = dtr_host::make_slab(m.X0, m.L0, m.A, m.Z, m.rho_g_cm3 * kGram / kCm3, thickness);
        if (m.ded) {
            out->has_density = 1u;
            out->density[0] = m.a;
            out->density[1] = m.k;
            out->density[2] = m.x0;
            out->density[3] = m.x1;
            out->density[4] = m.cbar;
            out->mean_exc = m.I_eV * kEV;
        }
        return DTR_SUCCESS;
    }
    return DTR_ERROR_INVALID_ARGUMENT;
}

int dtr_check_detector(const dtr_detector_desc* d, char* msg, size_t msg_size) {
    const std::string why = d ? validate_desc(*d) : std::string("no detector tables");
    if (msg && msg_size) {
        const size_t k = std::min(msg_size - 1, why.size());
        std::memcpy(msg, why.data(), k);
        msg[k] = '\0';
    }
    return why.empty() ? DTR_SUCCESS : DTR_ERROR_INVALID_ARGUMENT;
}

void dtr_host_detector_free(dtr_detector_desc* d) {
    if (!d) return;
    std::free(const_cast<dtr_volume*>(d->volumes));
    std::free(const_cast<dtr_surface*>(d->surfaces));
    std::free(const_cast<dtr_transform*>(d->transforms));
    std::free(const_cast<dtr_mask*>(d->masks));
    std::free(const_cast<dtr_material_slab*>(d->materials));
    std::free(const_cast<dtr_grid*>(d->grids));
    std::free(const_cast<uint32_t*>(d->bf_surfaces));
    std::free(const_cast<uint32_t*>(d->grid_bins));
    std::free(const_cast<dtr_material_map*>(d->material_maps));
    std::memset(d, 0, sizeof(*d));
}

int dtr_sample_solenoid(double b_center, double radius, double half_length,
                        const uint32_t n[3], const double mn[3], const double pitch[3],
                        float* values) {
    if (!n || !mn || !pitch || !values || radius <= 0.0 || half_length <= 0.0)
        return DTR_ERROR_INVALID_ARGUMENT;
    double br0, bz0;
    solenoid_unit(radius, half_length, 0.0, 0.0, br0, bz0);
    const double scale = b_center / bz0;
    for (uint32_t ix = 0; ix < n[0]; ++ix) {
        for (uint32_t iy = 0; iy < n[1]; ++iy) {
            const double x = mn[0] + ix * pitch[0], y = mn[1] + iy * pitch[1];
            const double rho = std::sqrt(x * x + y * y);
            for (uint32_t iz = 0; iz < n[2]; ++iz) {
                const double z = mn[2] + iz * pitch[2];
                double br, bz;
                if (std::fabs(rho - radius) < 1e-9 * radius) {
                    // on the current sheet: average of both sides
                    double a1, b1, a2, b2;
                    solenoid_unit(radius, half_length, rho * (1.0 - 1e-6), z, a1, b1);
                    solenoid_unit(radius, half_length, rho * (1.0 + 1e-6), z, a2, b2);
                    br = 0.5 * (a1 + a2);
                    bz = 0.5 * (b1 + b2);
                } else {
                    solenoid_unit(radius, half_length, rho, z, br, bz);
                }
                br *= scale;
                bz *= scale;
                const double bx = rho > 0.0 ? br * x / rho : 0.0;
                const double by = rho > 0.0 ? br * y / rho : 0.0;
                float* v = values + 3ull * ((static_cast<uint64_t>(ix) * n[1] + iy) * n[2] + iz);
                v[0] = static_cast<float>(bx);
                v[1] = static_cast<float>(by);
                v[2] = static_cast<float>(bz);
            }
        }
    }
    return DTR_SUCCESS;
}

int dtr_generate_tracks(const dtr_track_gen_config* cfg, double* fp) {
    if (!cfg || !fp) return DTR_ERROR_INVALID_ARGUMENT;
    const uint64_t n = cfg->n;
    for (uint64_t i = 0; i < n; ++i) {
        double pos[3] = {0, 0, 0}, dir[3], qop;
        const double q = (u01(cfg->seed, i, 1) < 0.5) ? -1.0 : 1.0;
        if (cfg->kind == 0u) {
            const double p = cfg->p_min + (cfg->p_max - cfg->p_min) * u01(cfg->seed, i, 2);
            const double theta = std::fabs(cfg->sigma_theta * gauss(cfg->seed, i, 3));
            const double phi = -kPi + 2.0 * kPi * u01(cfg->seed, i, 5);
            dir[0] = std::cos(phi) * std::sin(theta);
            dir[1] = std::sin(phi) * std::sin(theta);
            dir[2] = std::cos(theta);
            qop = q / p;
        } else {
            const double eta = cfg->eta_min + (cfg->eta_max - cfg->eta_min) * u01(cfg->seed, i, 2);
            const double phi = -kPi + 2.0 * kPi * u01(cfg->seed, i, 3);
            const double lpt = std::log(cfg->pt_min) +
                               (std::log(cfg->pt_max) - std::log(cfg->pt_min)) * u01(cfg->seed, i, 4);
            const double pt = std::exp(lpt);
            const double theta = 2.0 * std::atan(std::exp(-eta));
            dir[0] = std::cos(phi) * std::sin(theta);
            dir[1] = std::sin(phi) * std::sin(theta);
            dir[2] = std::cos(theta);
            const double p = pt / std::sin(theta);
            qop = q / p;
            pos[0] = cfg->sigma_xy * gauss(cfg->seed, i, 5);
            pos[1] = cfg->sigma_xy * gauss(cfg->seed, i, 7);
            pos[2] = cfg->sigma_z * gauss(cfg->seed, i, 9);
        }
        fp[0 * n + i] = pos[0];
        fp[1 * n + i] = pos[1];
        fp[2 * n + i] = pos[2];
        fp[3 * n + i] = 0.0;
        fp[4 * n + i] = dir[0];
        fp[5 * n + i] = dir[1];
        fp[6 * n + i] = dir[2];
        fp[7 * n + i] = qop;
    }
    return DTR_SUCCESS;
}

}  // extern "C"

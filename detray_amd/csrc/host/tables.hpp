// Flat detector tables on the host, shared by the detector emitters
// (builders.cpp) and the JSON reader/writer (json_io.cpp): the growing
// table set, surface identifiers, material slabs with their derived
// quantities, and the export into a dtr_detector_desc owned by the library.
#pragma once

#include "detray_amd.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace dtr_host {

constexpr uint16_t kEndOfWorld = 0xffffu;
constexpr uint32_t kInvalid = 0xffffffffu;

// detray unit conventions (core/include/detray/definitions/units.hpp)
constexpr double kMM = 1.0;
constexpr double kEV = 1e-9;
constexpr double kGram = 1.0 / 1.782662e-24;
constexpr double kCm3 = 1e3;
constexpr double kMol = 1.0;

struct Tables {
    std::vector<dtr_volume> volumes;
    std::vector<dtr_surface> surfaces;
    std::vector<dtr_transform> transforms;
    std::vector<dtr_mask> masks;
    std::vector<dtr_material_slab> materials;
    std::vector<dtr_grid> grids;
    std::vector<uint32_t> bf;
    std::vector<uint32_t> bins;
    std::vector<dtr_material_map> maps;
};

// geometry::identifier bit fields (core/include/detray/geometry/identifier.hpp:
// volume 12 bits, id 4, index 21, transform 21, extra 6)
inline uint64_t make_barcode(uint64_t vol, uint64_t id, uint64_t index, uint64_t trf) {
    return ((vol & 0xfffull) << 52) | ((id & 0xfull) << 48) |
           ((index & 0x1fffffull) << 27) | ((trf & 0x1fffffull) << 6);
}

// A homogeneous slab from the material parameters and the thickness
// (material.hpp: molar density from the mass density, mean excitation energy
// I = 16 eV * Z^0.9 without density-effect data; material_slab.hpp:
// thickness in X0)
inline dtr_material_slab make_slab(double X0, double L0, double A, double Z, double rho,
                                   double thickness) {
    dtr_material_slab s{};
    s.X0 = X0;
    s.L0 = L0;
    s.A = A;
    s.Z = Z;
    s.mass_rho = rho;
    // material.hpp:226-234 mass_to_molar_density
    const double molar_mass = A * kGram / kMol;
    s.molar_rho = (rho == 0.0) ? 0.0 : rho / molar_mass;
    s.mean_exc = 16.0 * kEV * std::pow(Z, 0.9);
    s.thickness = thickness;
    s.thickness_in_X0 = thickness / X0;
    return s;
}

template <typename T>
T* dup(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(sizeof(T) * v.size()));
    if (p) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

// Every index the navigation follows is in range: volume transforms,
// surface ranges, brute-force ranges and grids; surface transforms, masks,
// materials and volumes; mask volume links (a volume, or the world exit
// 0xffff); grid bin ranges and the surfaces in bins and brute-force lists.
// Returns an empty string for consistent tables, otherwise what is wrong.
inline std::string validate_desc(const dtr_detector_desc& d) {
    auto at = [](const char* what, uint64_t i) { return std::string(what) + " " + std::to_string(i) + ": "; };
    if (d.n_volumes == 0u || d.n_surfaces == 0u) return "no volumes or no surfaces";
    if (d.n_volumes > kEndOfWorld) return "more volumes than a 16-bit volume link can name";
    if (!d.volumes || !d.surfaces || (d.n_transforms && !d.transforms) || (d.n_masks && !d.masks) ||
        (d.n_materials && !d.materials) || (d.n_grids && !d.grids) || (d.n_bf && !d.bf_surfaces) ||
        (d.n_bins && !d.grid_bins) || (d.n_material_maps && !d.material_maps))
        return "a table with a non-zero count has no data";
    for (uint32_t vi = 0; vi < d.n_volumes; ++vi) {
        const dtr_volume& v = d.volumes[vi];
        if (v.transform >= d.n_transforms) return at("volume", vi) + "transform out of range";
        if (v.sf_begin > v.sf_end || v.sf_end > d.n_surfaces) return at("volume", vi) + "surface range out of range";
        if (v.bf_begin > v.bf_end || v.bf_end > d.n_bf) return at("volume", vi) + "brute-force range out of range";
        if (v.grid < -1 || (v.grid >= 0 && static_cast<uint32_t>(v.grid) >= d.n_grids))
            return at("volume", vi) + "grid out of range";
        if (v.material != 0u &&
            (v.material > d.n_materials || d.materials[v.material - 1u].kind != DTR_MATERIAL_SLAB))
            return at("volume", vi) + "material out of range or not homogeneous";
    }
    for (uint32_t s = 0; s < d.n_surfaces; ++s) {
        const dtr_surface& sf = d.surfaces[s];
        if (sf.transform >= d.n_transforms) return at("surface", s) + "transform out of range";
        if (sf.n_masks == 0u || static_cast<uint64_t>(sf.mask) + sf.n_masks > d.n_masks)
            return at("surface", s) + "mask range out of range";
        if (sf.shape > DTR_CYLINDER2D) return at("surface", s) + "unknown shape";
        if (sf.sf_id > DTR_PASSIVE) return at("surface", s) + "unknown surface id";
        if (sf.material != kInvalid) {
            if (sf.material & DTR_MATERIAL_MAP) {
                if ((sf.material & ~DTR_MATERIAL_MAP) >= d.n_material_maps)
                    return at("surface", s) + "material map out of range";
            } else if (sf.material >= d.n_materials) {
                return at("surface", s) + "material out of range";
            }
        }
        if (sf.volume >= d.n_volumes) return at("surface", s) + "volume out of range";
    }
    for (uint32_t m = 0; m < d.n_masks; ++m) {
        const uint16_t l = d.masks[m].volume_link;
        if (l != kEndOfWorld && l >= d.n_volumes) return at("mask", m) + "volume link out of range";
        if (d.masks[m].shape > DTR_CYLINDER2D) return at("mask", m) + "unknown shape";
    }
    for (uint32_t k = 0; k < d.n_materials; ++k) {
        if (d.materials[k].kind > DTR_MATERIAL_ROD) return at("material", k) + "unknown kind";
        if (d.materials[k].has_density > 1u) return at("material", k) + "has_density is not 0 or 1";
    }
    for (uint32_t k = 0; k < d.n_material_maps; ++k) {
        const dtr_material_map& m = d.material_maps[k];
        if (m.n0 == 0u || m.n1 == 0u || !(m.max0 > m.min0) || !(m.max1 > m.min1) || m.circular0 > 1u ||
            m.circular1 > 1u)
            return at("material map", k) + "empty or inverted axis";
        if (static_cast<uint64_t>(m.slabs) + static_cast<uint64_t>(m.n0) * m.n1 > d.n_materials)
            return at("material map", k) + "slabs out of range";
        for (uint64_t b = 0; b < static_cast<uint64_t>(m.n0) * m.n1; ++b)
            if (d.materials[m.slabs + b].kind != DTR_MATERIAL_SLAB) return at("material map", k) + "holds a rod";
    }
    for (uint32_t gi = 0; gi < d.n_grids; ++gi) {
        const dtr_grid& g = d.grids[gi];
        const uint64_t cap = g.capacity ? g.capacity : 1u;
        if (g.n0 == 0u || g.n1 == 0u || g.type > DTR_GRID_RING) return at("grid", gi) + "empty axis or unknown type";
        if (static_cast<uint64_t>(g.bins) + static_cast<uint64_t>(g.n0) * g.n1 * cap > d.n_bins)
            return at("grid", gi) + "bins out of range";
    }
    for (uint32_t k = 0; k < d.n_bins; ++k)
        if (d.grid_bins[k] != kInvalid && d.grid_bins[k] >= d.n_surfaces) return at("grid bin", k) + "surface out of range";
    for (uint32_t k = 0; k < d.n_bf; ++k)
        if (d.bf_surfaces[k] >= d.n_surfaces) return at("brute-force entry", k) + "surface out of range";
    return std::string();
}

inline void export_tables(const Tables& T, dtr_detector_desc* out) {
    std::memset(out, 0, sizeof(*out));
    out->n_volumes = static_cast<uint32_t>(T.volumes.size());
    out->n_surfaces = static_cast<uint32_t>(T.surfaces.size());
    out->n_transforms = static_cast<uint32_t>(T.transforms.size());
    out->n_masks = static_cast<uint32_t>(T.masks.size());
    out->n_materials = static_cast<uint32_t>(T.materials.size());
    out->n_grids = static_cast<uint32_t>(T.grids.size());
    out->n_bf = static_cast<uint32_t>(T.bf.size());
    out->n_bins = static_cast<uint32_t>(T.bins.size());
    out->volumes = dup(T.volumes);
    out->surfaces = dup(T.surfaces);
    out->transforms = dup(T.transforms);
    out->masks = dup(T.masks);
    out->materials = dup(T.materials);
    out->grids = dup(T.grids);
    out->bf_surfaces = dup(T.bf);
    out->grid_bins = dup(T.bins);
    out->n_material_maps = static_cast<uint32_t>(T.maps.size());
    out->material_maps = dup(T.maps);
}

}  // namespace dtr_host

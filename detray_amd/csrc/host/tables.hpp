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
}

}  // namespace dtr_host

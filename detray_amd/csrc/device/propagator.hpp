// Device implementation of detray's propagation hot path for one track per
// work-item: rk_stepper + caching_navigator + actor chain
//   pathlimit_aborter, parameter_updater<pointwise_material_interactor>,
//   surface_sequencer
// (core/include/detray/propagator/propagator.hpp:262-400).
//
// Layout choices for CDNA (64-wide waves):
//  * track inputs/outputs are structure-of-arrays so that a wave's loads and
//    stores of one parameter are contiguous;
//  * the transport Jacobian is kept in the reduced 25-value form of
//    codegen/transport_jacobian.hpp:18-325 (fixed identity/zero blocks are
//    implicit) and the covariance as its 21 independent values;
//  * the 8-slot candidate cache (caching_navigator.hpp:29) is one small array
//    per work-item.
// The detector tables are read-only and stay L2-resident.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "detray_amd.h"

namespace dtr {

#define DTR_HD __host__ __device__ inline

constexpr double kMax = 1.7976931348623157e308;
constexpr uint32_t kInvalid = 0xffffffffu;
constexpr uint16_t kEndOfWorld = 0xffffu;
constexpr int kCache = 8;

struct DevDetector {
    const dtr_volume* volumes;
    const dtr_surface* surfaces;
    const dtr_transform* transforms;
    const dtr_mask* masks;
    const dtr_material_slab* materials;
    const dtr_grid* grids;
    const uint32_t* bf;
    const uint32_t* bins;
    const dtr_material_map* mat_maps;
};

struct DevField {
    int32_t kind;
    uint32_t n[3];
    double b[3];
    double min[3];
    double pitch[3];
    const float* values;
};

// step_tracer outputs (dtr_step_trace), entry k of track i at k * n + i
struct TraceView {
    uint32_t cap;
    uint32_t* n;
    uint32_t* sf;
    double* step;
    double* path;
    uint32_t* trials;
    double* free;
    double* bound;
    double* jac;
};

// Propagation config as doubles where the reference converts float config
// values to the scalar type at the point of use.
struct DevConfig {
    double min_stepsize, rk_error_tol, step_constraint, path_limit;
    uint32_t max_rk_updates;
    float min_mask_tol, max_mask_tol, mask_scalor, path_tol, overstep_tol;
    uint32_t win0, win1;
    double accumulated_error;
    double n_stddev;
    uint32_t max_surfaces;
    uint8_t do_cov, estimate_noise, do_eloss, do_ms, use_material;
    uint8_t use_mean_loss, use_eloss_gradient;
    // optional actors (dtr_actor bits) and their states
    uint32_t actors;
    double min_p, min_pT;
    const uint32_t* target;
    TraceView trace;
};

// The device config of a propagation config (host).
inline DevConfig make_dev_config(const dtr_prop_config* c) {
    DevConfig dc{};
    dc.min_stepsize = c->min_stepsize;
    dc.rk_error_tol = c->rk_error_tol;
    dc.step_constraint = c->step_constraint;
    dc.path_limit = c->path_limit;
    dc.max_rk_updates = c->max_rk_updates;
    dc.min_mask_tol = c->min_mask_tolerance;
    dc.max_mask_tol = c->max_mask_tolerance;
    dc.mask_scalor = c->mask_tolerance_scalor;
    dc.path_tol = c->path_tolerance;
    dc.overstep_tol = c->overstep_tolerance;
    dc.win0 = c->search_window[0];
    dc.win1 = c->search_window[1];
    dc.accumulated_error = c->accumulated_error;
    dc.n_stddev = static_cast<double>(c->n_scattering_stddev);
    dc.max_surfaces = c->max_surfaces;
    dc.do_cov = c->do_covariance_transport;
    dc.estimate_noise = c->estimate_scattering_noise;
    dc.do_eloss = c->do_energy_loss;
    dc.do_ms = c->do_multiple_scattering;
    dc.use_material = c->use_material;
    dc.use_mean_loss = c->use_mean_loss;
    dc.use_eloss_gradient = c->use_eloss_gradient;
    if (const dtr_actor_states* a = c->actors) {
        dc.actors = a->actors;
        dc.min_p = static_cast<double>(a->min_p);
        dc.min_pT = static_cast<double>(a->min_pT);
        dc.target = a->target_surface;
        const dtr_step_trace& t = a->trace;
        dc.trace = TraceView{t.capacity, t.n, t.surface, t.step_size, t.path_length, t.n_trials,
                                  t.free_params, t.bound_params, t.jacobian};
    }
    return dc;
}

struct DevResults {
    double* bound_params;
    double* bound_cov;
    double* free_params;
    double* path_length;
    uint32_t* surfaces;
    uint32_t* n_surfaces;
    uint32_t* n_steps;
    uint32_t* n_trials;
    uint8_t* status;
};

// ---------------------------------------------------------------- algebra --
struct V3 {
    double x, y, z;
};
DTR_HD V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
DTR_HD V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DTR_HD V3 mul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
DTR_HD double dot(V3 a, V3 b) {
    double r = 0.0;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
DTR_HD V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
DTR_HD double norm(V3 a) { return sqrt(dot(a, a)); }
DTR_HD V3 normalize(V3 a) {
    const double n = norm(a);
    return {a.x / n, a.y / n, a.z / n};
}
DTR_HD double perp(V3 a) { return sqrt(fma(a.x, a.x, a.y * a.y)); }
DTR_HD double phi_of(V3 a) { return atan2(a.y, a.x); }
DTR_HD double theta_of(V3 a) { return atan2(perp(a), a.z); }
DTR_HD double dmin(double a, double b) { return a < b ? a : b; }
DTR_HD double dmax(double a, double b) { return a > b ? a : b; }

DTR_HD V3 trf_axis(const dtr_transform& t, int c) { return {t.r[c], t.r[3 + c], t.r[6 + c]}; }
DTR_HD V3 trf_translation(const dtr_transform& t) { return {t.t[0], t.t[1], t.t[2]}; }
DTR_HD V3 to_local(const dtr_transform& t, V3 p) {
    const V3 d = sub(p, trf_translation(t));
    return {t.r[0] * d.x + t.r[3] * d.y + t.r[6] * d.z,
            t.r[1] * d.x + t.r[4] * d.y + t.r[7] * d.z,
            t.r[2] * d.x + t.r[5] * d.y + t.r[8] * d.z};
}
DTR_HD V3 to_global(const dtr_transform& t, V3 l) {
    return {t.r[0] * l.x + t.r[1] * l.y + t.r[2] * l.z + t.t[0],
            t.r[3] * l.x + t.r[4] * l.y + t.r[5] * l.z + t.t[1],
            t.r[6] * l.x + t.r[7] * l.y + t.r[8] * l.z + t.t[2]};
}
DTR_HD V3 vec_to_global(const dtr_transform& t, V3 l) {
    return {t.r[0] * l.x + t.r[1] * l.y + t.r[2] * l.z,
            t.r[3] * l.x + t.r[4] * l.y + t.r[5] * l.z,
            t.r[6] * l.x + t.r[7] * l.y + t.r[8] * l.z};
}

// ------------------------------------------------------------------ field --
DTR_HD V3 field_at(const DevField& f, V3 p) {
    if (f.kind == DTR_FIELD_CONSTANT) return {f.b[0], f.b[1], f.b[2]};
    const double pc[3] = {p.x, p.y, p.z};
    int i0[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
        const int n = static_cast<int>(f.n[a]);
        double u = (pc[a] - f.min[a]) / f.pitch[a];
        // clamp (NaN positions clamp to the first sample)
        u = (u > 0.0) ? (u > n - 1 ? static_cast<double>(n - 1) : u) : 0.0;
        int i = static_cast<int>(floor(u));
        i = i > n - 2 ? n - 2 : i;
        i0[a] = i;
        t[a] = u - i;
    }
    const uint64_t ny = f.n[1], nz = f.n[2];
    const uint64_t base = (static_cast<uint64_t>(i0[0]) * ny + i0[1]) * nz + i0[2];
    const uint64_t sx = ny * nz, sy = nz;
    const float* v000 = f.values + 3 * base;
    double out[3];
    for (int c = 0; c < 3; ++c) {
        const double a000 = v000[c], a100 = v000[3 * sx + c];
        const double a001 = v000[3 + c], a101 = v000[3 * sx + 3 + c];
        const double a010 = v000[3 * sy + c], a110 = v000[3 * (sx + sy) + c];
        const double a011 = v000[3 * sy + 3 + c], a111 = v000[3 * (sx + sy) + 3 + c];
        const double c00 = a000 * (1.0 - t[0]) + a100 * t[0];
        const double c01 = a001 * (1.0 - t[0]) + a101 * t[0];
        const double c10 = a010 * (1.0 - t[0]) + a110 * t[0];
        const double c11 = a011 * (1.0 - t[0]) + a111 * t[0];
        const double c0 = c00 * (1.0 - t[1]) + c10 * t[1];
        const double c1 = c01 * (1.0 - t[1]) + c11 * t[1];
        out[c] = c0 * (1.0 - t[2]) + c1 * t[2];
    }
    return {out[0], out[1], out[2]};
}

// ------------------------------------------------------------ navigation --
enum : uint8_t { kInside = 0, kEdge = 1, kOutside = 2 };
enum : int8_t { kAbort = -3, kExit = -2, kUnknown = -1, kTowards = 0, kOnObject = 1, kOnPortal = 2 };
enum : uint8_t { kNoTrust = 0, kFair = 1, kHigh = 3, kFull = 4 };

struct Cand {
    double path;
    uint32_t sf;
    uint16_t link;
    uint8_t status;
    uint8_t pad;
};

struct NavTol {
    float min_mask_tol, max_mask_tol, mask_scalor, path_tol, overstep_tol;
};

struct Track {
    // navigation state (navigation_state.hpp:636-667)
    Cand c[kCache];
    double ext_tol;
    int8_t next, last, status;
    uint8_t trust;
    uint16_t volume;
    bool path_limit_hit;
    bool overflow;
    // stepper state (base_stepper.hpp:63-292, rk_stepper.hpp:86-263)
    V3 pos, dir;
    double time, qop;
    double J[25];  // reduced transport Jacobian
    double step_size, next_step_size, path_length, abs_path_length;
    double constraint_acc, constraint_abort;
    V3 dtds3;
    double dqopds3;  // d(qop)/ds of the last step's fourth stage
    double mass, charge;
    uint32_t n_trials, n_steps;
    // parameter_updater_state
    double bound[6];
    double cov[21];
    uint32_t up_sf;
    uint32_t n_seq;
    uint32_t n_trace;  // step_tracer entries
    bool heartbeat;
    bool stopped;  // energy loss brought the particle to rest
    bool momentum_hit, target_hit;  // the optional aborters fired
};

// Reduced Jacobian index helpers: rows 0-2 and 4-6 over columns 4..7, (7,7)
DTR_HD int jpos(int r, int c) { return r * 4 + (c - 4); }     // r in 0..2
DTR_HD int jdir(int r, int c) { return 12 + r * 4 + (c - 4); }  // r in 0..2 (rows 4..6)
constexpr int kJqq = 24;

// dtr_track_status of a finished track: exit (or sequencer overflow), else
// the first aborter that fired
DTR_HD uint8_t track_status(bool exit, bool overflow, bool path_limit, bool momentum, bool target) {
    if (exit) return overflow ? DTR_TRACK_OVERFLOW : DTR_TRACK_EXIT;
    if (path_limit) return DTR_TRACK_PATH_LIMIT;
    if (momentum) return DTR_TRACK_MOMENTUM_LIMIT;
    if (target) return DTR_TRACK_TARGET;
    return DTR_TRACK_ABORT;
}

// Element (r, c) of the free 8x8 transport Jacobian from its reduced form:
// identity position and time columns, rows 0-2 and 4-6 of the direction
// and q/p columns, (7, 7) (transport_jacobian.hpp:324)
template <typename T>
DTR_HD double free_jacobian(const T& J, int r, int c) {
    if (c < 4) return r == c ? 1.0 : 0.0;
    if (r < 3) return static_cast<double>(J[r * 4 + (c - 4)]);
    if (r >= 4 && r < 7) return static_cast<double>(J[12 + (r - 4) * 4 + (c - 4)]);
    if (r == 7) return c == 7 ? static_cast<double>(J[24]) : 0.0;
    return 0.0;  // time row
}

DTR_HD int covi(int r, int c) {
    // upper-triangle row-major index of (r, c), r <= c
    if (r > c) {
        const int t = r;
        r = c;
        c = t;
    }
    return r * 6 - (r * (r - 1)) / 2 + (c - r);
}

// Mask checks of the extended shapes on the local point `loc` (the frame of
// the surface's placement): inside with base_tol, edge with base_tol +
// ext_tol when ext_tol > 0.
DTR_HD void ext_mask_check(uint8_t shape, const double* v, V3 loc, double base_tol, double ext_tol,
                           bool& inside, bool& edge) {
    edge = false;
    if (shape == DTR_LINE_SQUARE || shape == DTR_LINE_CIRCULAR) {
        // line.hpp:86-121: square or circular cross section, half length
        const double c = fabs(loc.z);
        double a, b2;
        if (shape == DTR_LINE_SQUARE) {
            a = fabs(loc.x);
            b2 = fabs(loc.y);
            inside = (a <= (v[0] + base_tol) && b2 <= (v[0] + base_tol) && c <= (v[1] + base_tol));
        } else {
            a = fabs(perp(loc));
            b2 = 0.0;
            inside = (a <= (v[0] + base_tol) && c <= (v[1] + base_tol));
        }
        if (ext_tol > 0.0) {
            const double ft = base_tol + ext_tol;
            edge = (a <= (v[0] + ft) && b2 <= (v[0] + ft) && c <= (v[1] + ft));
        }
    } else if (shape == DTR_ANNULUS2D) {
        // annulus2D.hpp:140-190: beam-frame radius, strip-frame phi
        const double nx = loc.x + v[5], ny = loc.y + v[6];
        const double r_beam = sqrt(fma(nx, nx, ny * ny));
        double phi_focal = 0.0;
        bool have_phi = false;
        inside = ((r_beam + base_tol) >= v[0]) && (r_beam <= (v[1] + base_tol));
        if (inside) {
            phi_focal = phi_of(loc) - v[4];
            have_phi = true;
            const double pt = r_beam > 0.0 ? base_tol / r_beam : base_tol;
            inside = (phi_focal >= (v[2] - pt)) && (phi_focal <= (v[3] + pt));
        }
        if (ext_tol > 0.0) {
            const double ft = base_tol + ext_tol;
            edge = ((r_beam + ft) >= v[0]) && (r_beam <= (v[1] + ft));
            if (edge) {
                if (!have_phi) phi_focal = phi_of(loc) - v[4];
                const double pt = r_beam > 0.0 ? ft / r_beam : ft;
                edge = (phi_focal >= (v[2] - pt)) && (phi_focal <= (v[3] + pt));
            }
        }
    } else {
        // cylinder2D.hpp:63-90: local z within the bounds
        inside = ((v[1] - base_tol) <= loc.z && loc.z <= (v[2] + base_tol));
        if (ext_tol > 0.0) {
            const double ft = base_tol + ext_tol;
            edge = ((v[1] - ft) <= loc.z && loc.z <= (v[2] + ft));
        }
    }
}

// ------------------------------------------------------------ energy loss --
// relativistic_quantities.hpp:49-84 for a particle of mass m and charge q
struct RelQ {
    double qop, q2b2, beta2, bg, gamma, gamma2, E, Wmax;
};
DTR_HD RelQ rel_quantities(double mass, double qop, double q) {
    RelQ r;
    r.qop = qop;
    r.q2b2 = q * q + (mass * qop) * (mass * qop);
    const double mOverP = mass * ((q != 0.0) ? fabs(qop / q) : fabs(qop));
    const double pOverM = 1.0 / mOverP;
    r.beta2 = 1.0 / (1.0 + mOverP * mOverP);
    r.bg = pOverM;
    r.gamma = sqrt(1.0 + pOverM * pOverM);
    r.gamma2 = r.gamma * r.gamma;
    r.E = r.gamma * mass;
    const double m_e = 0.51099895069 * 1e-3;
    const double mfrac = m_e / mass;
    r.Wmax = (2.0 * m_e * r.bg * r.bg) / (1.0 + 2.0 * r.gamma * mfrac + mfrac * mfrac);
    return r;
}
// compute_delta_half (relativistic_quantities.hpp:137-200): the approximate
// correction above beta gamma = 10 without density-effect data (ned: molar
// electron density), Sternheimer's parameterisation with it
DTR_HD double delta_half(const dtr_material_slab& m, double bg, double ned) {
    if (!m.has_density) {
        if (!(bg >= 10.0)) return 0.0;
        const double plasma = (28.816 * 1e-9) * sqrt(1000.0 * ned);
        return log(bg * plasma / m.mean_exc) - 0.5;
    }
    const double x = log10(bg);
    double delta = 0.0;
    if (x >= m.density[2]) {
        delta = 2.0 * 2.302585092994045684 * x - m.density[4];
        if (x < m.density[3]) delta += m.density[0] * pow(m.density[3] - x, m.density[1]);
    }
    return 0.5 * delta;
}
// derive_delta_half (relativistic_quantities.hpp:202-240); -1/qop without
// density-effect data at any beta gamma, as there
DTR_HD double derive_delta_half(const dtr_material_slab& m, const RelQ& r) {
    if (!m.has_density) return -1.0 / r.qop;
    const double x = log10(r.bg);
    double delta = 0.0;
    if (x >= m.density[2]) {
        delta = -2.0 / r.qop;
        if (x < m.density[3])
            delta += m.density[0] * m.density[1] / (r.qop * 2.302585092994045684) *
                     pow(m.density[3] - x, m.density[1] - 1.0);
    }
    return 0.5 * delta;
}
// compute_bethe_bloch (interaction.hpp:44-58): the muon's stopping power
DTR_HD double stopping_power(const dtr_material_slab& m, const RelQ& r) {
    const double K = 0.307075 * 1e-3 * 1e2;
    const double ned = m.Z * m.molar_rho;
    const double eps_pl = 0.5 * K * ned * r.q2b2;
    if (!(eps_pl > 0.0)) return 0.0;
    const double dhalf = delta_half(m, r.bg, ned);
    const double m_e = 0.51099895069 * 1e-3;
    const double u = 2.0 * m_e * r.bg * r.bg;
    const double A = 0.5 * log(u * r.Wmax / (m.mean_exc * m.mean_exc));
    return 2.0 * eps_pl * (A - r.beta2 - dhalf);
}
// derive_bethe_bloch (interaction.hpp:110-154)
DTR_HD double derive_stopping_power(const dtr_material_slab& m, const RelQ& r) {
    const double K = 0.307075 * 1e-3 * 1e2;
    const double eps_pl = 0.5 * K * (m.Z * m.molar_rho) * r.q2b2;
    if (!(eps_pl > 0.0)) return 0.0;
    const double first = 2.0 / (r.qop * r.gamma2) * stopping_power(m, r);
    const double dA = -1.0 / (2.0 * r.qop) * (4.0 - r.Wmax / r.E);
    const double dB = -2.0 * r.beta2 / (r.qop * r.gamma2);
    const double dC = derive_delta_half(m, r);
    return first + 2.0 * eps_pl * (dA - dB - dC);
}
// rk_stepper.hpp:157-189: d(qop)/ds = qop^3 E (-dE/ds) / q^2, zero in vacuum
DTR_HD double eloss_dqopds(const dtr_material_slab* vm, double qop, double mass, double q) {
    if (!vm) return 0.0;
    const double p = q / qop;
    const double E = sqrt(p * p + mass * mass);
    const double sp = stopping_power(*vm, rel_quantities(mass, qop, q));
    return qop * qop * qop * E * sp / (q * q);
}
// rk_stepper.hpp:194-231: d(d(qop)/ds)/dqop (JINST 4 P04016, Eq. 3.12)
DTR_HD double eloss_d2qopdsdqop(const dtr_material_slab* vm, double qop, double mass, double q) {
    if (!vm) return 0.0;
    const double p = q / qop;
    const double p2 = p * p;
    const double E2 = p2 + mass * mass;
    const RelQ r = rel_quantities(mass, qop, q);
    const double g = -1.0 * stopping_power(*vm, r);
    const double dgdqop = -1.0 * derive_stopping_power(*vm, r);
    const double dq = eloss_dqopds(vm, qop, mass, q);
    return dq * (1.0 / qop * (3.0 - p2 / E2) + 1.0 / g * dgdqop);
}
// The slab a material surface applies at the bound local position (l0, l1)
// (material_accessor.hpp:30-48): its own slab, or the bin of its material
// map (material_map.hpp) — regular axes, bin int((x - min) / width + 1) - 1
// (axis_binning.hpp:80-82), a circular axis wrapping, a closed one clamping
// (axis_bounds.hpp:107-206)
DTR_HD int material_map_bin(double x, double mn, double mx, uint32_t n, uint8_t circular) {
    const int nb = static_cast<int>(n);
    const int b = static_cast<int>((x - mn) / ((mx - mn) / static_cast<double>(n)) + 1.0) - 1;
    if (circular) {
        if (b >= 0 && b < nb) return b;
        return (nb + (b % nb)) % nb;
    }
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}
DTR_HD const dtr_material_slab& surface_material(const dtr_material_slab* mats, const dtr_material_map* maps,
                                                 uint32_t link, double l0, double l1) {
    if (!(link & DTR_MATERIAL_MAP)) return mats[link];
    const dtr_material_map& m = maps[link & ~DTR_MATERIAL_MAP];
    const int i0 = material_map_bin(l0, m.min0, m.max0, m.n0, m.circular0);
    const int i1 = material_map_bin(l1, m.min1, m.max1, m.n1, m.circular1);
    return mats[m.slabs + static_cast<uint32_t>(i0) * m.n1 + static_cast<uint32_t>(i1)];
}

// The homogeneous material of volume v (0xffff: none), propagator.hpp:320-325
DTR_HD const dtr_material_slab* volume_material(const dtr_volume* vols, const dtr_material_slab* mats,
                                                uint16_t v) {
    if (v == kEndOfWorld) return nullptr;
    const uint32_t m = vols[v].material;
    return m ? &mats[m - 1u] : nullptr;
}

// kExt: the engine also handles wires, annuli, placed cylinders, rods,
// multi-surface bins and rings with material (the extended geometry); the
// instantiation without them is the plain FP64 kernel of detectors made of
// rectangles, trapezoids and portal cylinders/rings
template <bool kExt>
struct EngineT {
    const DevDetector& det;
    const DevField& field;
    const DevConfig& cfg;
    Track& tk;

    DTR_HD bool on_surface() const { return tk.status == kOnObject || tk.status == kOnPortal; }
    DTR_HD int n_cand() const { return tk.last - tk.next + 1; }
    DTR_HD bool exhausted() const { return n_cand() == 0; }
    DTR_HD int begin_idx() const { return (on_surface() && tk.next >= 1) ? tk.next - 1 : tk.next; }
    DTR_HD double nav_dist() const { return tk.c[tk.next].path; }
    DTR_HD void set_high_trust() { tk.trust = tk.trust < kHigh ? tk.trust : kHigh; }
    DTR_HD void set_fair_trust() { tk.trust = tk.trust < kFair ? tk.trust : kFair; }
    DTR_HD bool is_portal(uint32_t s) const { return det.surfaces[s].sf_id == DTR_PORTAL; }

    // resolve_mask (intersector_base.hpp:171-252)
    DTR_HD void resolve_mask(Cand& is, double path, V3 glob, double z2d, uint32_t sf_idx,
                             const dtr_surface& sf, const dtr_mask& m, const dtr_transform& t,
                             const NavTol& tol) const {
        is.status = kOutside;
        if (!(path >= static_cast<double>(tol.overstep_tol))) return;
        double base_tol = 0.0, ext_tol = 0.0;
        if (sf.sf_id != DTR_PORTAL || tol.min_mask_tol == 3.40282347e+38f) {
            ext_tol = tk.ext_tol;
            base_tol = dmax(static_cast<double>(tol.min_mask_tol),
                            dmin(static_cast<double>(tol.max_mask_tol),
                                 static_cast<double>(tol.mask_scalor) * fabs(path)));
        }
        bool inside, edge;
        if (m.shape == DTR_CONCENTRIC_CYLINDER2D) {
            inside = (m.v[1] - base_tol <= z2d && z2d <= m.v[2] + base_tol);
            edge = inside;
        } else {
            const V3 loc = to_local(t, glob);
            if (m.shape == DTR_RECTANGLE2D) {
                const double l0 = fabs(loc.x), l1 = fabs(loc.y);
                inside = (l0 <= (m.v[0] + base_tol) && l1 <= (m.v[1] + base_tol));
                edge = false;
                if (ext_tol > 0.0) {
                    const double ft = base_tol + ext_tol;
                    edge = (l0 <= (m.v[0] + ft) && l1 <= (m.v[1] + ft));
                }
            } else if (m.shape == DTR_TRAPEZOID2D) {
                const double rel_y = (m.v[2] + loc.y) * m.v[3];
                const double bound_x = m.v[0] + rel_y * (m.v[1] - m.v[0]);
                inside = (fabs(loc.x) <= (bound_x + base_tol) && fabs(loc.y) <= (m.v[2] + base_tol));
                edge = false;
                if (ext_tol > 0.0) {
                    const double ft = base_tol + ext_tol;
                    edge = (fabs(loc.x) <= (bound_x + ft) && fabs(loc.y) <= (m.v[2] + ft));
                }
            } else if (kExt && m.shape >= DTR_LINE_SQUARE) {
                ext_mask_check(m.shape, m.v, loc, base_tol, ext_tol, inside, edge);
            } else {
                const double r = perp(loc);
                inside = ((r + base_tol) >= m.v[0] && r <= (m.v[1] + base_tol));
                edge = inside;
            }
        }
        if (edge) is.status = kEdge;
        if (inside) is.status = kInside;
        is.path = path;
        is.sf = sf_idx;
        is.link = m.volume_link;
    }

    // ray_plane_intersector.hpp:60-87
    static DTR_HD bool plane_intersect(V3 ro, V3 rd, const dtr_transform& t, double& s, V3& glob) {
        const V3 sn = trf_axis(t, 2);
        const V3 st = trf_translation(t);
        const double denom = dot(rd, sn);
        if (denom == 0.0) {
            s = kMax;
            glob = {0.0, 0.0, 0.0};
            return false;
        }
        s = dot(sn, sub(st, ro)) / denom;
        glob = add(ro, mul(s, rd));
        return fabs(s) < 10.0 * 1000.0;
    }

    // ray_line_intersector.hpp:44-84
    static DTR_HD bool line_intersect(V3 ro, V3 rd, const dtr_transform& t, double& s, V3& glob) {
        const V3 z = trf_axis(t, 2);
        const double zd = dot(z, rd);
        const double denom = 1.0 - (zd * zd);
        if (denom < static_cast<double>(1e-5f)) {
            s = kMax;
            glob = {0.0, 0.0, 0.0};
            return false;
        }
        const V3 t2l = sub(trf_translation(t), ro);
        const double t2l_on_line = dot(t2l, z);
        const double t2l_on_track = dot(t2l, rd);
        s = 1.0 / denom * (t2l_on_track - t2l_on_line * zd);
        glob = add(ro, mul(s, rd));
        return fabs(s) < 10.0 * 1000.0;
    }

    // ray_cylinder_intersector.hpp:40-110, quadratic_equation.hpp:30-75:
    // ascending solutions, n = 0, 1 or 2
    static DTR_HD int cyl2_intersect(V3 ro, V3 rd, const dtr_transform& t, double r, double s[2]) {
        const V3 sz = trf_axis(t, 2);
        const V3 pc_cross_sz = cross(sub(ro, trf_translation(t)), sz);
        const V3 rd_cross_sz = cross(rd, sz);
        const double a = dot(rd_cross_sz, rd_cross_sz);
        const double b = 2.0 * dot(rd_cross_sz, pc_cross_sz);
        const double c = dot(pc_cross_sz, pc_cross_sz) - (r * r);
        const double eps = 2.220446049250313e-16;
        s[0] = s[1] = kMax;
        if (fabs(a) <= eps) {
            if (fabs(b) <= eps) return 0;
            s[0] = -c / b;
            return 1;
        }
        const double disc = b * b - 4.0 * a * c;
        if (disc > eps) {
            const double q = -0.5 * (b + copysign(sqrt(disc), b));
            const double s0 = q / a, s1 = c / q;
            s[0] = s0 > s1 ? s1 : s0;
            s[1] = s0 > s1 ? s0 : s1;
            return 2;
        }
        if (disc >= 0.0) {
            s[0] = -0.5 * b / a;
            return 1;
        }
        return 0;
    }

    // ray_cylinder_portal_intersector.hpp:68-132
    static DTR_HD bool cyl_intersect(V3 ro, V3 rd, double r, double overstep, double& path, double& z) {
        path = kMax;
        const double rd_perp_2 = rd.x * rd.x + rd.y * rd.y;
        if (rd_perp_2 < 2.220446049250313e-16) {
            z = 0.0;
            return false;
        }
        const double ro_perp_2 = ro.x * ro.x + ro.y * ro.y;
        const double rad_diff = r * r - ro_perp_2;
        if (fabs(rad_diff) > 1.0 * 1e-3) {
            const double inv2 = 1.0 / rd_perp_2;
            const double k = -inv2 * (ro.x * rd.x + ro.y * rd.y);
            const double discr = inv2 * rad_diff + k * k;
            if (discr < 0.0) {
                z = 0.0;
                return false;
            }
            const double sq = sqrt(discr);
            const double s1 = k + sq, s2 = k - sq;
            path = (fabs(s1) < fabs(s2)) ? s1 : s2;
            if (path < overstep && sq > 0.0) path = (fabs(s1) >= fabs(s2)) ? s1 : s2;
        } else {
            path = rad_diff;
        }
        if (path < overstep) path = kMax;
        z = ro.z + path * rd.z;
        return fabs(path) < 10.0 * 1000.0;
    }

    DTR_HD bool intersect(uint32_t s, V3 ro, V3 rd, double overstep, double& path, V3& glob,
                          double& z) const {
        const dtr_surface& sf = det.surfaces[s];
        if (sf.shape == DTR_CONCENTRIC_CYLINDER2D)
            return cyl_intersect(ro, rd, det.masks[sf.mask].v[0], overstep, path, z);
        if (kExt && (sf.shape == DTR_LINE_SQUARE || sf.shape == DTR_LINE_CIRCULAR))
            return line_intersect(ro, rd, det.transforms[sf.transform], path, glob);
        if (kExt && sf.shape == DTR_CYLINDER2D) {
            // intersection_update takes the smaller solution
            double sol[2];
            if (cyl2_intersect(ro, rd, det.transforms[sf.transform], det.masks[sf.mask].v[0], sol) == 0)
                return false;
            path = sol[0];
            glob = add(ro, mul(sol[0], rd));
            return fabs(sol[0]) < 10.0 * 1000.0;
        }
        return plane_intersect(ro, rd, det.transforms[sf.transform], path, glob);
    }

    // caching_navigator.hpp:149-229
    DTR_HD void insert(int pos, const Cand& nc) {
        if (pos == kCache) return;
        if (n_cand() == 0) {
            tk.c[0] = nc;
            tk.last = static_cast<int8_t>(tk.last + 1);
            return;
        }
        int idx = pos;
        const uint64_t nbar = det.surfaces[nc.sf].barcode;
        auto overlap = [&](int i) { return fabs(tk.c[i].path - nc.path) <= 1.0 * 1e-3; };
        auto clash = [&](int i) {
            const uint64_t a = tk.c[i].sf == kInvalid ? ~0ull : det.surfaces[tk.c[i].sf].barcode;
            return a == nbar && overlap(i);
        };
        if (clash(idx) || (idx > 0 && clash(idx - 1))) return;
        const int shift_begin = tk.last < kCache - 2 ? tk.last : kCache - 2;
        if (overlap(idx) && (nc.status == kEdge || is_portal(nc.sf))) {
            idx++;
            idx = idx > shift_begin ? shift_begin : idx;
        }
        for (int i = shift_begin; i >= idx; --i) tk.c[i + 1] = tk.c[i];
        tk.c[idx] = nc;
        tk.last = static_cast<int8_t>(tk.last + 1 < kCache - 1 ? tk.last + 1 : kCache - 1);
    }

    // navigation_functions.hpp:44-68 + intersection_kernel.hpp:24-168
    DTR_HD void candidate_search(uint32_t s, const NavTol& tol) {
        const V3 rd = tk.dir;  // forward navigation
        const dtr_surface& sf = det.surfaces[s];
        const dtr_transform& t = det.transforms[sf.transform];
        if (kExt && sf.shape == DTR_CYLINDER2D) {
            // intersection_initialize: every valid solution a mask accepts,
            // at most two
            double sol[2];
            const int n = cyl2_intersect(tk.pos, rd, t, det.masks[sf.mask].v[0], sol);
            for (uint32_t k = 0; k < sf.n_masks; ++k) {
                int found = 0;
                for (int i = 0; i < n; ++i) {
                    if (!(fabs(sol[i]) < 10.0 * 1000.0)) continue;
                    Cand is;
                    is.path = kMax;
                    is.sf = kInvalid;
                    is.link = 0;
                    is.pad = 0;
                    resolve_mask(is, sol[i], add(tk.pos, mul(sol[i], rd)), 0.0, s, sf, det.masks[sf.mask + k], t,
                                 tol);
                    if (is.status <= kEdge) {
                        int p = begin_idx();
                        const int e = tk.last + 1;
                        while (p < e && !(fabs(is.path) < fabs(tk.c[p].path))) ++p;
                        insert(p, is);
                        if (++found == 2) return;
                    }
                }
            }
            return;
        }
        double path, z;
        V3 glob;
        if (!intersect(s, tk.pos, rd, tol.overstep_tol, path, glob, z)) return;
        for (uint32_t k = 0; k < sf.n_masks; ++k) {
            Cand is;
            is.path = kMax;
            is.sf = kInvalid;
            is.link = 0;
            is.pad = 0;
            resolve_mask(is, path, glob, z, s, sf, det.masks[sf.mask + k], t, tol);
            if (is.status <= kEdge) {
                int p = begin_idx();
                const int e = tk.last + 1;
                while (p < e && !(fabs(is.path) < fabs(tk.c[p].path))) ++p;
                insert(p, is);
                return;
            }
        }
    }

    // navigation_functions.hpp:99-129 + intersection_kernel.hpp:172-272
    DTR_HD bool update_candidate(Cand& cand) const {
        if (cand.sf == kInvalid) return false;
        double path, z;
        V3 glob;
        const NavTol tol{cfg.min_mask_tol, cfg.max_mask_tol, cfg.mask_scalor, cfg.path_tol,
                         cfg.overstep_tol};
        if (!intersect(cand.sf, tk.pos, tk.dir, tol.overstep_tol, path, glob, z)) return false;
        const dtr_surface& sf = det.surfaces[cand.sf];
        const dtr_transform& t = det.transforms[sf.transform];
        for (uint32_t k = 0; k < sf.n_masks; ++k) {
            resolve_mask(cand, path, glob, z, cand.sf, sf, det.masks[sf.mask + k], t, tol);
            if (cand.status <= kEdge) return true;
        }
        return false;
    }

    // spatial_grid.hpp:100-202, axis_binning.hpp:80-115, axis_bounds.hpp
    DTR_HD void grid_search(const dtr_volume& vol, const NavTol& tol) {
        const dtr_grid& g = det.grids[vol.grid];
        double l0, l1;
        if (g.type == DTR_GRID_CYLINDER) {
            double path, z;
            cyl_intersect(tk.pos, tk.dir, g.radius, -kMax, path, z);
            l0 = phi_of(add(tk.pos, mul(path, tk.dir)));
            l1 = z;
        } else {
            const dtr_transform& vt = det.transforms[vol.transform];
            double s;
            V3 glob;
            plane_intersect(tk.pos, tk.dir, vt, s, glob);
            const V3 loc = to_local(vt, glob);
            l0 = perp(loc);
            l1 = phi_of(loc);
        }
        const double w0 = (g.max0 - g.min0) / static_cast<double>(g.n0);
        const double w1 = (g.max1 - g.min1) / static_cast<double>(g.n1);
        // non-finite lookup point: no defined bin, no grid candidates
        const double f0 = (l0 - g.min0) / w0 + 1.0, f1 = (l1 - g.min1) / w1 + 1.0;
        if (!(fabs(f0) < 1e9) || !(fabs(f1) < 1e9)) return;
        const int ib0 = static_cast<int>((l0 - g.min0) / w0 + 1.0) - 1;
        const int ib1 = static_cast<int>((l1 - g.min1) / w1 + 1.0) - 1;
        const int lb0 = ib0 - static_cast<int>(cfg.win0), ub0 = ib0 + static_cast<int>(cfg.win1);
        const int lb1 = ib1 - static_cast<int>(cfg.win0), ub1 = ib1 + static_cast<int>(cfg.win1);
        const bool cyl = (g.type == DTR_GRID_CYLINDER);
        const int n0 = static_cast<int>(g.n0), n1 = static_cast<int>(g.n1);
        int a0, e0, a1, e1;
        if (cyl) {
            a0 = lb0;
            e0 = ub0 + 1;
            int mn = lb1 > 0 ? lb1 : 0;
            mn = mn >= n1 ? n1 - 1 : mn;
            int mx = ub1 >= n1 ? n1 - 1 : ub1;
            mx = mx < 0 ? 0 : mx;
            a1 = mn;
            e1 = mx + 1;
        } else {
            int mn = lb0 > 0 ? lb0 : 0;
            mn = mn >= n0 ? n0 - 1 : mn;
            int mx = ub0 >= n0 ? n0 - 1 : ub0;
            mx = mx < 0 ? 0 : mx;
            a0 = mn;
            e0 = mx + 1;
            a1 = lb1;
            e1 = ub1 + 1;
        }
        for (int b0 = a0; b0 < e0; ++b0) {
            const int i0 = cyl ? (n0 + (b0 % n0)) % n0 : b0;
            for (int b1 = a1; b1 < e1; ++b1) {
                const int i1 = cyl ? b1 : (n1 + (b1 % n1)) % n1;
                // every entry of the bin in fill order (capacity 0 means 1)
                const uint32_t cap = kExt ? (g.capacity ? g.capacity : 1u) : 1u;
                const uint32_t first = g.bins + (static_cast<uint32_t>(i0) * g.n1 + i1) * cap;
                for (uint32_t k = 0; k < cap; ++k) {
                    const uint32_t s = det.bins[first + k];
                    if (s != kInvalid) candidate_search(s, tol);
                }
            }
        }
    }

    // navigation_functions.hpp:147-177
    DTR_HD void update_status() {
        if (!exhausted() && fabs(tk.c[tk.next].path) < static_cast<double>(cfg.path_tol)) {
            tk.status = is_portal(tk.c[tk.next].sf) ? kOnPortal : kOnObject;
            tk.next = static_cast<int8_t>(tk.next + 1);
        } else {
            tk.status = kTowards;
        }
        tk.trust = (exhausted() || tk.status == kOnPortal) ? kNoTrust : kFull;
    }

    // navigation_functions.hpp:193-246
    DTR_HD void local_navigation(float overstep_tol, bool resolve_overstepping) {
        const dtr_volume& vol = det.volumes[tk.volume];
        for (int i = 0; i < kCache; ++i) tk.c[i].path = kMax;
        tk.next = 0;
        tk.last = -1;
        const NavTol tol{cfg.min_mask_tol, cfg.max_mask_tol, cfg.mask_scalor, cfg.path_tol,
                         resolve_overstepping ? overstep_tol : -cfg.path_tol};
        for (uint32_t k = vol.bf_begin; k < vol.bf_end; ++k) candidate_search(det.bf[k], tol);
        if (vol.grid >= 0) grid_search(vol, tol);
        update_status();
        if (tk.trust != kFull && tk.status == kOnPortal)
            tk.trust = (tk.c[tk.next - 1].link == kEndOfWorld) ? kFull : kNoTrust;
    }

    DTR_HD void volume_switch() {
        const uint16_t link = tk.c[tk.next - 1].link;
        if (link == kEndOfWorld) {
            tk.status = kExit;
            return;
        }
        if (link != tk.volume) tk.trust = kNoTrust;
        tk.volume = link;
        local_navigation(cfg.overstep_tol, true);
        tk.trust = kFull;
    }

    DTR_HD void init_loose_cfg() {
        const float a = 100.f * cfg.overstep_tol, b = -10.f * cfg.max_mask_tol;
        local_navigation(a < b ? a : b, true);
        if (tk.trust != kFull || exhausted()) tk.status = kAbort;
    }

    // caching_navigator.hpp:250-388
    DTR_HD bool update_impl() {
        if (tk.trust == kHigh) {
            if (!update_candidate(tk.c[tk.next])) {
                tk.status = kUnknown;
                set_fair_trust();
            } else {
                update_status();
                if (tk.status == kTowards || tk.status == kOnPortal) return false;
                if (tk.trust == kFull && update_candidate(tk.c[tk.next])) return false;
                set_fair_trust();
            }
        }
        if (tk.trust == kFair && !exhausted()) {
            const int b = begin_idx(), e = tk.last + 1;
            for (int i = b; i < e; ++i) {
                if (!update_candidate(tk.c[i])) tk.c[i].path = kMax;
            }
            for (int i = b; i < e; ++i) {
                for (int j = i + 1; j < e; ++j) {
                    if (fabs(tk.c[j].path) < fabs(tk.c[i].path)) {
                        const Cand tmp = tk.c[i];
                        tk.c[i] = tk.c[j];
                        tk.c[j] = tmp;
                    }
                }
            }
            tk.next = static_cast<int8_t>(b);
            int inv = kCache;
            for (int i = kCache - 1; i >= 0; --i) {
                if (tk.c[i].path == kMax) inv = i;
            }
            tk.last = static_cast<int8_t>(inv - 1);
            update_status();
            if (exhausted()) tk.trust = kNoTrust;
        }
        if (tk.trust == kNoTrust) {
            local_navigation(cfg.overstep_tol, true);
            return true;
        }
        return false;
    }

    // navigator_base.hpp:97-154
    DTR_HD bool nav_update() {
        if (tk.trust == kFull) return false;
        bool is_init = update_impl();
        if (tk.status == kOnPortal) {
            volume_switch();
            is_init = true;
            if (!(tk.status > kUnknown)) return false;
        }
        if (tk.trust != kFull || exhausted()) {
            is_init = true;
            init_loose_cfg();
        }
        return is_init;
    }

    // -------------------------------------------------------------- stepper --
    DTR_HD const dtr_material_slab* vol_mat() const {
        return kExt ? volume_material(det.volumes, det.materials, tk.volume) : nullptr;
    }
    // rk_stepper.hpp:147-155: the last step's fourth-stage d(qop)/ds, or
    // its value at the current q/p before the first step
    DTR_HD double state_dqopds() const {
        if (tk.path_length == 0.0) return eloss_dqopds(vol_mat(), tk.qop, tk.mass, tk.charge);
        return tk.dqopds3;
    }

    // rk_stepper.hpp:273-530.  In volume material (extended geometry only)
    // the stage q/p follow the energy loss (evaluate_dqopds,
    // rk_stepper.hpp:534-565); in vacuum every stage sees the track's q/p.
    DTR_HD void step(double dist, bool do_reset) {
        if (fabs(dist) < static_cast<double>(1.f * 1e-3f)) {
            if (fabs(tk.next_step_size) < cfg.min_stepsize)
                tk.next_step_size = copysign(cfg.min_stepsize, dist);
            return;
        }
        if (do_reset) {
            tk.step_size = dist;
        } else if (tk.next_step_size > 0) {
            tk.step_size = dmin(tk.next_step_size, dist);
        } else {
            tk.step_size = dmax(tk.next_step_size, dist);
        }
        const double min_step = dmin(fabs(tk.step_size), cfg.min_stepsize);
        const V3 pos = tk.pos;
        const V3 dir = tk.dir;
        const double qop = tk.qop;
        const dtr_material_slab* vm = vol_mat();
        // stage q/p and d(qop)/ds
        double q[4] = {qop, qop, qop, qop}, dq[4] = {0.0, 0.0, 0.0, 0.0};
        if (kExt && vm) dq[0] = eloss_dqopds(vm, qop, tk.mass, tk.charge);
        const V3 b1 = field_at(field, pos);
        const V3 k1 = mul(qop, cross(dir, b1));
        V3 b2, b3, t2, t3, t4, k2, k3, k4;
        double error = 1e20;
        for (uint32_t i = 0; i < cfg.max_rk_updates; ++i) {
            tk.n_trials++;
            const double h = tk.step_size;
            const double h2 = h * h;
            const double half_h = h * 0.5;
            const V3 pos1 = add(add(pos, mul(half_h, dir)), mul(h2 * 0.125, k1));
            b2 = field_at(field, pos1);
            if (kExt && vm) {
                if (cfg.use_mean_loss) q[1] = qop + half_h * dq[0];
                dq[1] = eloss_dqopds(vm, q[1], tk.mass, tk.charge);
            }
            t2 = (half_h == 0.0) ? dir : add(dir, mul(half_h, k1));
            k2 = mul(q[1], cross(t2, b2));
            if (kExt && vm) {
                if (cfg.use_mean_loss) q[2] = qop + half_h * dq[1];
                dq[2] = eloss_dqopds(vm, q[2], tk.mass, tk.charge);
            }
            t3 = (half_h == 0.0) ? dir : add(dir, mul(half_h, k2));
            k3 = mul(q[2], cross(t3, b2));
            const V3 pos2 = add(add(pos, mul(h, dir)), mul(h2 * 0.5, k3));
            b3 = field_at(field, pos2);
            if (kExt && vm) {
                if (cfg.use_mean_loss) q[3] = qop + h * dq[2];
                dq[3] = eloss_dqopds(vm, q[3], tk.mass, tk.charge);
            }
            t4 = (h == 0.0) ? dir : add(dir, mul(h, k3));
            k4 = mul(q[3], cross(t4, b3));
            const V3 d = add(sub(sub(k1, k2), k3), k4);
            const V3 err = mul((1.0 / 6.0) * h2, d);
            error = dmax(norm(err), 1e-20);
            if (error <= 4.0 * cfg.rk_error_tol) break;
            tk.step_size = tk.step_size * dmin(dmax(sqrt(sqrt(cfg.rk_error_tol / error)), 0.25), 4.0);
        }
        tk.dtds3 = k4;
        tk.dqopds3 = dq[3];
        // constrained_step (accuracy, aborter)
        const double cmin = dmin(tk.constraint_acc, tk.constraint_abort);
        const double max_step = (tk.step_size >= 0.0 ? 1.0 : -1.0) * cmin;
        if (fabs(tk.step_size) > fabs(max_step)) tk.step_size = max_step;
        if (fabs(tk.step_size) < min_step) tk.step_size = copysign(min_step, tk.step_size);
        tk.next_step_size =
            tk.step_size * dmin(dmax(sqrt(sqrt(cfg.rk_error_tol / error)), 0.25), 4.0);
        if (fabs(tk.next_step_size) < cfg.min_stepsize)
            tk.next_step_size = copysign(cfg.min_stepsize, tk.next_step_size);
        // advance_track (rk_stepper.hpp:594-629)
        const double h = tk.step_size;
        const double h_6 = h * (1.0 / 6.0);
        double qn = tk.qop;
        if (kExt && vm) {
            qn = tk.qop + h_6 * (dq[0] + 2.0 * (dq[1] + dq[2]) + dq[3]);
            // brought to rest within the step: not taken, the track stops
            if (!(fabs(qn) <= kMax && qn * tk.qop > 0.0)) {
                tk.stopped = true;
                return;
            }
        }
        tk.pos = add(pos, mul(h, add(dir, mul(h_6, add(add(k1, k2), k3)))));
        const V3 nd = add(dir, mul(h_6, add(add(k1, mul(2.0, add(k2, k3))), k4)));
        const double nn = norm(nd);
        tk.dir = {nd.x / nn, nd.y / nn, nd.z / nn};
        tk.qop = qn;
        tk.path_length += h;
        tk.abs_path_length += fabs(h);
        if (cfg.do_cov) advance_jacobian(dir, t2, t3, t4, b1, b2, b3, q, vm);
        tk.n_steps++;
    }

    static DTR_HD void cwc(const double M[9], V3 b, double out[9]) {
        for (int j = 0; j < 3; ++j) {
            const V3 col{M[j], M[3 + j], M[6 + j]};
            const V3 r = cross(col, b);
            out[j] = r.x;
            out[3 + j] = r.y;
            out[6 + j] = r.z;
        }
    }

    // rk_stepper.hpp:633-956 (no field gradient) and
    // codegen/update_rk_transport_jacobian.hpp:18-199 on the reduced storage.
    // q: stage q/p; with use_eloss_gradient dqop_n/dqop1 and d(qop)/d(qop1)
    // follow d(d(qop)/ds)/dqop of the volume material (rk_stepper.hpp:
    // 767-800), 1 otherwise and in vacuum.
    DTR_HD void advance_jacobian(V3 t1, V3 t2, V3 t3, V3 t4, V3 b1, V3 b2, V3 b3, const double q[4],
                                 const dtr_material_slab* vm) {
        const double h = tk.step_size;
        const double half_h = h * 0.5;
        const double h_6 = h * static_cast<double>(1.f / 6.f);
        double dqn[4] = {1.0, 1.0, 1.0, 1.0};
        double dqopqop = 1.0;
        if (kExt && vm && cfg.use_eloss_gradient) {
            const double e1 = eloss_d2qopdsdqop(vm, q[0], tk.mass, tk.charge);
            dqn[1] = 1.0 + half_h * e1;
            const double e2 = eloss_d2qopdsdqop(vm, q[1], tk.mass, tk.charge) * dqn[1];
            dqn[2] = 1.0 + half_h * e2;
            const double e3 = eloss_d2qopdsdqop(vm, q[2], tk.mass, tk.charge) * dqn[2];
            dqn[3] = 1.0 + h * e3;
            const double e4 = eloss_d2qopdsdqop(vm, q[3], tk.mass, tk.charge) * dqn[3];
            dqopqop = 1.0 + h_6 * (e1 + 2.0 * (e2 + e3) + e4);
        }
        V3 dk0 = mul(dqn[0], cross(t1, b1));
        V3 dk1 = add(mul(dqn[1], cross(t2, b2)), mul(q[1] * half_h, cross(dk0, b2)));
        V3 dk2 = add(mul(dqn[2], cross(t3, b2)), mul(q[2] * half_h, cross(dk1, b2)));
        V3 dk3 = add(mul(dqn[3], cross(t4, b3)), mul(q[3] * h, cross(dk2, b3)));
        const V3 dFdqop = mul(h * h_6, add(add(dk0, dk1), dk2));
        const V3 dGdqop = mul(h_6, add(add(dk0, mul(2.0, add(dk1, dk2))), dk3));
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double d0[9], d1[9], d2[9], d3[9], tmp[9];
        cwc(I, b1, tmp);
        for (int i = 0; i < 9; ++i) d0[i] = q[0] * tmp[i];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + half_h * d0[i];
        cwc(tmp, b2, d1);
        for (int i = 0; i < 9; ++i) d1[i] *= q[1];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + half_h * d1[i];
        cwc(tmp, b2, d2);
        for (int i = 0; i < 9; ++i) d2[i] *= q[2];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + h * d2[i];
        cwc(tmp, b3, d3);
        for (int i = 0; i < 9; ++i) d3[i] *= q[3];
        double dFdt[9], dGdt[9];
        for (int i = 0; i < 9; ++i) {
            dFdt[i] = h * (I[i] + h_6 * (d0[i] + d1[i] + d2[i]));
            dGdt[i] = I[i] + h_6 * (d0[i] + 2.0 * (d1[i] + d2[i]) + d3[i]);
        }
        // new_J = D * J on the structural non-zeros (columns 4..7)
        double nj[25];
        for (int c = 4; c < 8; ++c) {
            const double jd0 = tk.J[jdir(0, c)], jd1 = tk.J[jdir(1, c)], jd2 = tk.J[jdir(2, c)];
            const double jq = (c == 7) ? tk.J[kJqq] : 0.0;
            for (int r = 0; r < 3; ++r) {
                double s = 0.0;
                s += tk.J[jpos(r, c)];
                s += dFdt[r * 3 + 0] * jd0;
                s += dFdt[r * 3 + 1] * jd1;
                s += dFdt[r * 3 + 2] * jd2;
                nj[jpos(r, c)] = s;
                double g = 0.0;
                g += dGdt[r * 3 + 0] * jd0;
                g += dGdt[r * 3 + 1] * jd1;
                g += dGdt[r * 3 + 2] * jd2;
                nj[jdir(r, c)] = g;
            }
            if (c == 7) {
                const double dfq[3] = {dFdqop.x, dFdqop.y, dFdqop.z};
                const double dgq[3] = {dGdqop.x, dGdqop.y, dGdqop.z};
                for (int r = 0; r < 3; ++r) {
                    nj[jpos(r, 7)] += dfq[r] * jq;
                    nj[jdir(r, 7)] += dgq[r] * jq;
                }
            }
        }
        nj[kJqq] = dqopqop * tk.J[kJqq];
        for (int i = 0; i < 25; ++i) tk.J[i] = nj[i];
    }

    // policies.hpp:99-139
    DTR_HD void rk_policy() {
        const double d = nav_dist();
        if (fabs(d) <= static_cast<double>(1.f * 1e-3f)) {
            set_high_trust();
            return;
        }
        const double rel = (tk.step_size - d) / d;
        if (rel > 0.1) tk.trust = kNoTrust;
        else if (rel > 0.05) set_fair_trust();
        else set_high_trust();
    }

    // ------------------------------------------------------ frames, actors --
    static DTR_HD bool is_line(uint8_t shape) { return shape == DTR_LINE_SQUARE || shape == DTR_LINE_CIRCULAR; }
    // line2D.hpp:29-47: signed distance of closest approach, z along the wire
    static DTR_HD void line_local(const dtr_transform& t, V3 pos, V3 dir, double& l0, double& l1) {
        const V3 local3 = to_local(t, pos);
        const V3 r = cross(trf_axis(t, 2), dir);
        const double sign = dot(r, sub(trf_translation(t), pos)) > 0.0 ? -1.0 : 1.0;
        l0 = sign * perp(local3);
        l1 = local3.z;
    }
    // jacobian_cylindrical.hpp:24-45: y = cylinder axis, z = normal at the
    // local point (phi from r phi / r), x = y x z
    static DTR_HD void cyl2_frame(const dtr_transform& t, V3 pos, V3& nx, V3& ny, V3& nz) {
        const V3 l = to_local(t, pos);
        const double r = perp(l);
        const double ph = (r * phi_of(l)) / r;
        ny = trf_axis(t, 2);
        nz = vec_to_global(t, {cos(ph), sin(ph), 0.0});
        nx = cross(ny, nz);
    }
    // jacobian_line.hpp:38-62: y along the wire, x = y x dir, z = x x y
    static DTR_HD void line_frame(const dtr_transform& t, V3 dir, V3& nx, V3& ny, V3& nz) {
        ny = trf_axis(t, 2);
        nx = normalize(cross(ny, dir));
        nz = cross(nx, ny);
    }

    DTR_HD void free_to_bound(uint32_t s, double out[6]) const {
        const dtr_surface& sf = det.surfaces[s];
        if (kExt && is_line(sf.shape)) {
            line_local(det.transforms[sf.transform], tk.pos, tk.dir, out[0], out[1]);
        } else if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            out[0] = phi_of(tk.pos);
            out[1] = tk.pos.z;
        } else if (kExt && sf.shape == DTR_CYLINDER2D) {
            const V3 l = to_local(det.transforms[sf.transform], tk.pos);
            out[0] = perp(l) * phi_of(l);
            out[1] = l.z;
        } else if (sf.shape == DTR_RING2D || (kExt && sf.shape == DTR_ANNULUS2D)) {
            const V3 l = to_local(det.transforms[sf.transform], tk.pos);
            out[0] = perp(l);
            out[1] = phi_of(l);
        } else {
            const V3 l = to_local(det.transforms[sf.transform], tk.pos);
            out[0] = l.x;
            out[1] = l.y;
        }
        out[2] = phi_of(tk.dir);
        out[3] = theta_of(tk.dir);
        out[4] = tk.qop;
        out[5] = tk.time;
    }
    static DTR_HD V3 bound_dir(const double b[6]) {
        const double s = sin(b[3]);
        return {cos(b[2]) * s, sin(b[2]) * s, cos(b[3])};
    }
    DTR_HD V3 bound_pos(uint32_t s, const double b[6]) const {
        const dtr_surface& sf = det.surfaces[s];
        if (kExt && is_line(sf.shape)) {
            // line2D.hpp:80-99 with the track direction
            const dtr_transform& t = det.transforms[sf.transform];
            const V3 r = cross(trf_axis(t, 2), bound_dir(b));
            return add(to_global(t, {0.0, 0.0, b[1]}), mul(b[0], normalize(r)));
        }
        if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double r = det.masks[sf.mask].v[0];
            return {r * cos(b[0]), r * sin(b[0]), b[1]};
        }
        if (kExt && sf.shape == DTR_CYLINDER2D) {
            const double r = det.masks[sf.mask].v[0];
            const double ph = b[0] / r;
            return to_global(det.transforms[sf.transform], {r * cos(ph), r * sin(ph), b[1]});
        }
        if (sf.shape == DTR_RING2D || (kExt && sf.shape == DTR_ANNULUS2D))
            return to_global(det.transforms[sf.transform], {b[0] * cos(b[1]), b[0] * sin(b[1]), 0.0});
        return to_global(det.transforms[sf.transform], {b[0], b[1], 0.0});
    }

    DTR_HD V3 dtds() const {
        if (tk.path_length == 0.0) return mul(tk.qop, cross(tk.dir, field_at(field, tk.pos)));
        return tk.dtds3;
    }

    // parameter_updater.hpp:422-572 get_full_jacobian, then
    // covariance_transport.hpp C' = J C J^T
    DTR_HD void transport_covariance(uint32_t dest) {
        const uint32_t dep = tk.up_sf;
        const dtr_surface& dsf = det.surfaces[dep];
        // bound-to-free at the departure surface: A (3x2 dpos/dloc), C (3x2 ddir/dangle)
        double A[6];
        // dpos/dangle of a line (jacobian_line.hpp:102-172), zero otherwise
        double Pa[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        if (kExt && is_line(dsf.shape)) {
            const dtr_transform& t = det.transforms[dsf.transform];
            const V3 ddir = bound_dir(tk.bound);
            const V3 dp = bound_pos(dep, tk.bound);
            V3 nx, ny, nz;
            line_frame(t, ddir, nx, ny, nz);
            A[0] = nx.x; A[1] = ny.x;
            A[2] = nx.y; A[3] = ny.y;
            A[4] = nx.z; A[5] = ny.z;
            double l0, l1;
            line_local(t, dp, ddir, l0, l1);
            const double ct = cos(tk.bound[3]), stt = sin(tk.bound[3]);
            const double cp = cos(tk.bound[2]), sp = sin(tk.bound[2]);
            const V3 yp = cross(ny, {-stt * sp, stt * cp, 0.0});
            const V3 yt = cross(ny, {ct * cp, ct * sp, -stt});
            const double C = 1.0 / dot(ddir, nz) * l0;
            const V3 dphi = mul(C, sub(yp, mul(dot(nx, yp), nx)));
            const V3 dtheta = mul(C, sub(yt, mul(dot(nx, yt), nx)));
            Pa[0][0] = dphi.x; Pa[1][0] = dphi.y; Pa[2][0] = dphi.z;
            Pa[0][1] = dtheta.x; Pa[1][1] = dtheta.y; Pa[2][1] = dtheta.z;
        } else if (dsf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const V3 dp = bound_pos(dep, tk.bound);
            const double r = perp(dp), ph = phi_of(dp);
            A[0] = -r * sin(ph); A[1] = 0.0;
            A[2] = r * cos(ph);  A[3] = 0.0;
            A[4] = 0.0;          A[5] = 1.0;
        } else if (kExt && dsf.shape == DTR_CYLINDER2D) {
            V3 nx, ny, nz;
            cyl2_frame(det.transforms[dsf.transform], bound_pos(dep, tk.bound), nx, ny, nz);
            A[0] = nx.x; A[1] = ny.x;
            A[2] = nx.y; A[3] = ny.y;
            A[4] = nx.z; A[5] = ny.z;
        } else if (kExt && (dsf.shape == DTR_RING2D || dsf.shape == DTR_ANNULUS2D)) {
            // jacobian_polar.hpp:44-72
            const dtr_transform& t = det.transforms[dsf.transform];
            const V3 l = to_local(t, bound_pos(dep, tk.bound));
            const double lr = perp(l), lp = phi_of(l), c = cos(lp), sn = sin(lp);
            const V3 x = trf_axis(t, 0), y = trf_axis(t, 1);
            const V3 col0 = add(mul(c, x), mul(sn, y));
            const V3 col1 = mul(lr, sub(mul(c, y), mul(sn, x)));
            A[0] = col0.x; A[1] = col1.x;
            A[2] = col0.y; A[3] = col1.y;
            A[4] = col0.z; A[5] = col1.z;
        } else {
            const dtr_transform& t = det.transforms[dsf.transform];
            A[0] = t.r[0]; A[1] = t.r[1];
            A[2] = t.r[3]; A[3] = t.r[4];
            A[4] = t.r[6]; A[5] = t.r[7];
        }
        double Cd[6];
        {
            const double th = tk.bound[3], ph = tk.bound[2];
            const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
            Cd[0] = -st * sp; Cd[1] = ct * cp;
            Cd[2] = st * cp;  Cd[3] = ct * sp;
            Cd[4] = 0.0;      Cd[5] = -st;
        }
        // M = J_transport * B2F  (8x6, rows 0-2 pos, 4-6 dir, 7 qop; row 3 = time)
        // columns: 0,1 loc; 2,3 angles; 4 qop; 5 time
        double Mp[3][6], Md[3][6];
        for (int r = 0; r < 3; ++r) {
            // pos rows: J_pos = [I | Jp(4..7)]: I*A for loc, Jp[4..6]*C for angles, Jp7 for qop
            Mp[r][0] = A[r * 2 + 0];
            Mp[r][1] = A[r * 2 + 1];
            Mp[r][2] = tk.J[jpos(r, 4)] * Cd[0] + tk.J[jpos(r, 5)] * Cd[2] + tk.J[jpos(r, 6)] * Cd[4];
            Mp[r][3] = tk.J[jpos(r, 4)] * Cd[1] + tk.J[jpos(r, 5)] * Cd[3] + tk.J[jpos(r, 6)] * Cd[5];
            if (kExt) {
                Mp[r][2] += Pa[r][0];
                Mp[r][3] += Pa[r][1];
            }
            Mp[r][4] = tk.J[jpos(r, 7)];
            Mp[r][5] = 0.0;
            Md[r][0] = 0.0;
            Md[r][1] = 0.0;
            Md[r][2] = tk.J[jdir(r, 4)] * Cd[0] + tk.J[jdir(r, 5)] * Cd[2] + tk.J[jdir(r, 6)] * Cd[4];
            Md[r][3] = tk.J[jdir(r, 4)] * Cd[1] + tk.J[jdir(r, 5)] * Cd[3] + tk.J[jdir(r, 6)] * Cd[5];
            Md[r][4] = tk.J[jdir(r, 7)];
            Md[r][5] = 0.0;
        }
        const double Mq4 = tk.J[kJqq];  // qop row: only column 4
        // d(qop)/ds of the path correction, with the volume at the surface
        // (parameter_updater.hpp:518-526)
        const double dqs = kExt ? state_dqopds() : 0.0;
        // path correction: P = (I + p2f f2p) applied on the left;
        // p2f = [dir, 0, dtds, 0], f2p = [n_pos, 0...]
        const V3 k = dtds();
        const dtr_surface& esf = det.surfaces[dest];
        V3 f2p, f2d{0.0, 0.0, 0.0};
        if (kExt && is_line(esf.shape)) {
            // jacobian_line.hpp:64-92: position and direction parts with dt/ds
            const dtr_transform& t = det.transforms[esf.transform];
            const V3 pc = sub(tk.pos, trf_translation(t));
            const V3 lz = trf_axis(t, 2);
            const double pz = dot(pc, lz);
            const double dz = dot(lz, tk.dir);
            const V3 pc_x = sub(pc, mul(pz, lz));
            const double nrm = -1.0 / (1.0 - dz * dz + dot(pc_x, k));
            f2p = mul(nrm, sub(tk.dir, mul(dz, lz)));
            f2d = mul(nrm, pc_x);
        } else {
            V3 n;
            if (esf.shape == DTR_CONCENTRIC_CYLINDER2D) {
                const double ph = phi_of(tk.pos);
                n = {cos(ph), sin(ph), 0.0};
            } else if (kExt && esf.shape == DTR_CYLINDER2D) {
                V3 nx, ny;
                cyl2_frame(det.transforms[esf.transform], tk.pos, nx, ny, n);
            } else {
                n = trf_axis(det.transforms[esf.transform], 2);
            }
            f2p = mul(-1.0 / dot(n, tk.dir), n);
        }
        double proj[6];  // f2p . M (per column)
        for (int c = 0; c < 6; ++c) {
            proj[c] = f2p.x * Mp[0][c] + f2p.y * Mp[1][c] + f2p.z * Mp[2][c];
            if (kExt) proj[c] += f2d.x * Md[0][c] + f2d.y * Md[1][c] + f2d.z * Md[2][c];
        }
        const double p2f_pos[3] = {tk.dir.x, tk.dir.y, tk.dir.z};
        const double p2f_dir[3] = {k.x, k.y, k.z};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 6; ++c) {
                Mp[r][c] += p2f_pos[r] * proj[c];
                Md[r][c] += p2f_dir[r] * proj[c];
            }
        }
        // free-to-bound at the destination
        double Bl[6];
        if (kExt && is_line(esf.shape)) {
            // jacobian_line.hpp:174-182: rows x and y of the line's frame
            V3 nx, ny, nz;
            line_frame(det.transforms[esf.transform], tk.dir, nx, ny, nz);
            Bl[0] = nx.x; Bl[1] = nx.y; Bl[2] = nx.z;
            Bl[3] = ny.x; Bl[4] = ny.y; Bl[5] = ny.z;
        } else if (esf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double rinv = 1.0 / perp(tk.pos), ph = phi_of(tk.pos);
            Bl[0] = -rinv * sin(ph); Bl[1] = rinv * cos(ph); Bl[2] = 0.0;
            Bl[3] = 0.0;             Bl[4] = 0.0;            Bl[5] = 1.0;
        } else if (kExt && esf.shape == DTR_CYLINDER2D) {
            V3 nx, ny, nz;
            cyl2_frame(det.transforms[esf.transform], tk.pos, nx, ny, nz);
            Bl[0] = nx.x; Bl[1] = nx.y; Bl[2] = nx.z;
            Bl[3] = ny.x; Bl[4] = ny.y; Bl[5] = ny.z;
        } else if (kExt && (esf.shape == DTR_RING2D || esf.shape == DTR_ANNULUS2D)) {
            // jacobian_polar.hpp:81-110
            const dtr_transform& t = det.transforms[esf.transform];
            const V3 l = to_local(t, tk.pos);
            const double lr = perp(l), lp = phi_of(l), c = cos(lp), sn = sin(lp);
            const V3 x = trf_axis(t, 0), y = trf_axis(t, 1);
            const V3 row0 = add(mul(c, x), mul(sn, y));
            const V3 row1 = mul(1.0 / lr, sub(mul(c, y), mul(sn, x)));
            Bl[0] = row0.x; Bl[1] = row0.y; Bl[2] = row0.z;
            Bl[3] = row1.x; Bl[4] = row1.y; Bl[5] = row1.z;
        } else {
            const dtr_transform& t = det.transforms[esf.transform];
            Bl[0] = t.r[0]; Bl[1] = t.r[3]; Bl[2] = t.r[6];
            Bl[3] = t.r[1]; Bl[4] = t.r[4]; Bl[5] = t.r[7];
        }
        double Ba[6];
        {
            const double th = theta_of(tk.dir), ph = phi_of(tk.dir);
            const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
            Ba[0] = -sp / st; Ba[1] = cp / st; Ba[2] = 0.0;
            Ba[3] = cp * ct;  Ba[4] = sp * ct; Ba[5] = -st;
        }
        double Jf[6][6];
        for (int c = 0; c < 6; ++c) {
            Jf[0][c] = Bl[0] * Mp[0][c] + Bl[1] * Mp[1][c] + Bl[2] * Mp[2][c];
            Jf[1][c] = Bl[3] * Mp[0][c] + Bl[4] * Mp[1][c] + Bl[5] * Mp[2][c];
            Jf[2][c] = Ba[0] * Md[0][c] + Ba[1] * Md[1][c] + Ba[2] * Md[2][c];
            Jf[3][c] = Ba[3] * Md[0][c] + Ba[4] * Md[1][c] + Ba[5] * Md[2][c];
            Jf[4][c] = (c == 4) ? Mq4 : 0.0;
            // the q/p row of the path correction (jacobian_engine.hpp:260-273)
            if (kExt) Jf[4][c] += dqs * proj[c];
            Jf[5][c] = (c == 5) ? 1.0 : 0.0;
        }
        // C' = Jf C Jf^T
        double Cf[6][6];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) Cf[r][c] = tk.cov[covi(r, c)];
        double T[6][6];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
                for (int m = 0; m < 6; ++m) s += Jf[r][m] * Cf[m][c];
                T[r][c] = s;
            }
        for (int r = 0; r < 6; ++r)
            for (int c = r; c < 6; ++c) {
                double s = 0.0;
                for (int m = 0; m < 6; ++m) s += T[r][m] * Jf[c][m];
                tk.cov[covi(r, c)] = s;
            }
    }

    DTR_HD bool on_sensitive() const {
        return tk.status == kOnObject && det.surfaces[tk.c[tk.next - 1].sf].sf_id == DTR_SENSITIVE;
    }
    DTR_HD bool on_material() const {
        return on_surface() && det.surfaces[tk.c[tk.next - 1].sf].material != kInvalid;
    }

    DTR_HD void reset_J() {
        for (int i = 0; i < 25; ++i) tk.J[i] = 0.0;
        tk.J[jdir(0, 4)] = 1.0;
        tk.J[jdir(1, 5)] = 1.0;
        tk.J[jdir(2, 6)] = 1.0;
        tk.J[kJqq] = 1.0;
    }

    // relativistic_quantities.hpp + interaction.hpp
    DTR_HD void material_interaction(uint32_t dest) {
        const dtr_material_slab& m =
            surface_material(det.materials, det.mat_maps, det.surfaces[dest].material, tk.bound[0], tk.bound[1]);
        V3 nrm;
        const dtr_surface& msf = det.surfaces[dest];
        if (msf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            nrm = {cos(tk.bound[0]), sin(tk.bound[0]), 0.0};
        } else if (kExt && msf.shape == DTR_CYLINDER2D) {
            // cylindrical2D.hpp:62-70: phi = loc0 / r of the mask
            const double ph = tk.bound[0] / det.masks[msf.mask].v[0];
            nrm = vec_to_global(det.transforms[msf.transform], {cos(ph), sin(ph), 0.0});
        } else {
            nrm = trf_axis(det.transforms[msf.transform], 2);
        }
        const double cos_inc = fabs(dot(bound_dir(tk.bound), nrm));
        if (!(m.thickness > 2.220446049250313e-16) || tk.bound[4] == 0.0) return;
        const double qop = tk.bound[4];
        double path_seg, path_x0;
        if (kExt && m.kind == DTR_MATERIAL_ROD) {
            // material_rod.hpp:62-78, approach = bound loc0
            const double a = tk.bound[0];
            path_seg = 0.0;
            if (!(fabs(a) > m.thickness))
                path_seg = 2.0 * sqrt((m.thickness * m.thickness - a * a) / (1.0 - cos_inc * cos_inc));
            path_x0 = path_seg / m.X0;
        } else {
            path_seg = m.thickness / cos_inc;
            path_x0 = m.thickness_in_X0 / cos_inc;
        }
        const double mass = tk.mass, q = tk.charge;
        const double q2b2 = q * q + (mass * qop) * (mass * qop);
        const double mOverP = mass * ((q != 0.0) ? fabs(qop / q) : fabs(qop));
        const double pOverM = 1.0 / mOverP;
        const double beta2 = 1.0 / (1.0 + mOverP * mOverP);
        const double bg = pOverM;
        const double gamma = sqrt(1.0 + pOverM * pOverM);
        const double m_e = 0.51099895069 * 1e-3;
        const double mfrac = m_e / mass;
        const double Wmax = (2.0 * m_e * bg * bg) / (1.0 + 2.0 * gamma * mfrac + mfrac * mfrac);
        const double K = 0.307075 * 1e-3 * 1e2;
        const double ned = m.Z * m.molar_rho;
        const double eps_pl = 0.5 * K * ned * q2b2;
        double e_loss = 0.0, sigma_qop = 0.0, theta0 = 0.0;
        if (cfg.do_eloss) {
            double bb = 0.0;
            if (eps_pl > 0.0) {
                const double dhalf = delta_half(m, bg, ned);
                const double u = 2.0 * m_e * bg * bg;
                const double A = 0.5 * log(u * Wmax / (m.mean_exc * m.mean_exc));
                bb = 2.0 * eps_pl * (A - beta2 - dhalf);
            }
            e_loss = path_seg * bb;
            if (cfg.do_cov) {
                const double eps = eps_pl * path_seg;
                const double fwhm = 4.0 * eps;
                const double sigmaE = 0.5 * 0.70710678118654752440 * fwhm / sqrt(0.69314718055994530942);
                const double pInv = qop / q;
                sigma_qop = sqrt(q2b2) * pInv * pInv * sigmaE;
            }
        }
        if (cfg.do_ms) {
            const double x0 = path_x0;
            const double pinv = fabs(qop / q);
            if (x0 > 0.0) {
                const double t = sqrt(x0 * q2b2);
                theta0 = static_cast<double>(13.6f) * 1e-3 * pinv * t *
                         (1.0 + static_cast<double>(0.038f) * 2.0 * log(t));
            }
        }
        if (cfg.do_eloss) {
            const double p = q / tk.bound[4];
            const double next_e = sqrt(mass * mass + p * p) - copysign(e_loss, 1.0);
            const double next_p = (mass < next_e) ? sqrt(next_e * next_e - mass * mass) : 0.0;
            const double next_qop = (q != 0.0) ? q / next_p : 1.0 / next_p;
            tk.bound[4] = (next_p == 0.0) ? kMax : next_qop;
            if (next_p == 0.0) tk.stopped = true;
            if (cfg.do_cov) tk.cov[covi(4, 4)] += sigma_qop * sigma_qop;
        }
        if (cfg.do_cov) {
            const V3 d = bound_dir(tk.bound);
            const double var = theta0 * theta0;
            tk.cov[covi(2, 2)] += (d.z == 1.0) ? kMax : var / (1.0 - d.z * d.z);
            tk.cov[covi(3, 3)] += var;
        }
        // parameter_setter (parameter_updater.hpp:585-627)
        tk.pos = bound_pos(dest, tk.bound);
        tk.dir = bound_dir(tk.bound);
        tk.time = tk.bound[5];
        tk.qop = tk.bound[4];
        if (cfg.estimate_noise) estimate_noise();
    }

    // noise_estimation.hpp:23-107
    DTR_HD void estimate_noise() {
        const double ns = cfg.n_stddev;
        const double v0 = tk.cov[covi(0, 0)], v1 = tk.cov[covi(1, 1)];
        const double dphi = ns * sqrt(tk.cov[covi(2, 2)]);
        const double dth = ns * sqrt(tk.cov[covi(3, 3)]);
        const double pe = tk.bound[2] + dphi, te = tk.bound[3] + dth;
        const double ste = sin(te);
        V3 displ{cos(pe) * ste, sin(pe) * ste, cos(te)};
        const double max_tol = 5.0;
        const double path = exhausted() ? max_tol : fabs(tk.c[tk.next].path);
        displ = mul(path, sub(displ, tk.dir));
        const double p = tk.charge / tk.qop;
        const double acc = 1.0 / p * cfg.accumulated_error * tk.path_length;
        const double ext = sqrt(ns * ns * (v0 + v1) + dot(displ, displ) + acc * acc);
        tk.ext_tol = ext > max_tol ? max_tol : ext;
    }

    // step_tracer observer (step_tracer.hpp: step_data): free state on the
    // surface, bound parameters after the material interactor, transport
    // Jacobian expanded to the free 8x8 matrix
    DTR_HD void trace_update(uint32_t dest, const double fp[8], uint32_t n, uint32_t gid) {
        const TraceView& t = cfg.trace;
        const uint32_t k = tk.n_trace++;
        if (k >= t.cap) return;
        const uint64_t o = static_cast<uint64_t>(k) * n + gid, M = static_cast<uint64_t>(t.cap) * n;
        if (t.sf) t.sf[o] = dest;
        if (t.step) t.step[o] = tk.step_size;
        if (t.path) t.path[o] = tk.path_length;
        if (t.trials) t.trials[o] = tk.n_trials;
        if (t.free)
            for (int i = 0; i < 8; ++i) t.free[i * M + o] = fp[i];
        if (t.bound)
            for (int i = 0; i < 6; ++i) t.bound[i * M + o] = tk.bound[i];
        if (t.jac)
            for (int r = 0; r < 8; ++r)
                for (int c = 0; c < 8; ++c) t.jac[(r * 8 + c) * M + o] = free_jacobian(tk.J, r, c);
    }

    // momentum_aborter (aborters.hpp:80-143): pT, then p, at or below the limits
    DTR_HD void momentum_aborter() {
        if (tk.status == kExit) return;
        const double q = tk.charge;
        const double pT = fabs(q / tk.qop * perp(tk.dir));
        if (pT <= cfg.min_pT || q / tk.qop <= cfg.min_p) {
            if (tk.status != kAbort) tk.momentum_hit = true;
            tk.status = kAbort;
            tk.heartbeat = false;
        }
    }

    // target_aborter (aborters.hpp:146-175)
    DTR_HD void target_aborter(uint32_t target) {
        if (on_surface() && tk.c[tk.next - 1].sf == target && tk.path_length > 0.0) {
            if (tk.status != kAbort) tk.target_hit = true;
            tk.status = kAbort;
            tk.heartbeat = false;
        }
    }

    DTR_HD void parameter_updater(const double* init_cov, uint32_t n, uint32_t gid) {
        if (!(on_sensitive() || on_material())) return;
        const uint32_t dest = tk.c[tk.next - 1].sf;
        const double fp[8] = {tk.pos.x, tk.pos.y, tk.pos.z, tk.time, tk.dir.x, tk.dir.y, tk.dir.z, tk.qop};
        if (tk.up_sf != kInvalid) {
            if (tk.up_sf != dest) {
                transport_covariance(dest);
                if (fabs(tk.path_length) > 0.0) reset_J();
                free_to_bound(dest, tk.bound);
                tk.up_sf = dest;
            }
        } else {
            if (init_cov) {
                for (int k = 0; k < 21; ++k) tk.cov[k] = init_cov[static_cast<uint64_t>(k) * n + gid];
            } else {
                for (int r = 0; r < 6; ++r)
                    for (int c = r; c < 6; ++c) tk.cov[covi(r, c)] = (r == c) ? 1.0 : 0.0;
            }
            free_to_bound(dest, tk.bound);
            tk.up_sf = dest;
        }
        if (on_material()) material_interaction(dest);
        if (cfg.actors & DTR_ACTOR_STEP_TRACER) trace_update(dest, fp, n, gid);
    }
};
using Engine = EngineT<false>;

// One track: propagator::propagate (propagator.hpp:262-400)
template <bool kExt>
DTR_HD void propagate_track(const DevDetector& det, const DevField& field,
                            const DevConfig& cfg, const double* __restrict__ fp,
                            const double* __restrict__ init_cov,
                            const double* __restrict__ bound0,
                            const uint32_t* __restrict__ start_sf, const DevResults& out,
                            uint32_t n, uint32_t gid) {
    Track tk;
    for (int i = 0; i < kCache; ++i) {
        tk.c[i].path = kMax;
        tk.c[i].sf = kInvalid;
        tk.c[i].link = 0;
        tk.c[i].status = kOutside;
        tk.c[i].pad = 0;
    }
    tk.ext_tol = 0.0;
    tk.next = 0;
    tk.last = -1;
    tk.status = kUnknown;
    tk.trust = kNoTrust;
    tk.volume = 0;
    tk.path_limit_hit = false;
    tk.overflow = false;
    tk.step_size = 0.0;
    tk.next_step_size = 0.0;
    tk.path_length = 0.0;
    tk.abs_path_length = 0.0;
    tk.constraint_acc = dmin(kMax, fabs(cfg.step_constraint));
    tk.constraint_abort = kMax;
    tk.dtds3 = {0.0, 0.0, 0.0};
    tk.dqopds3 = 0.0;
    // muon hypothesis, charge-conjugated to the track (pdg_particle.hpp:61-64)
    tk.mass = static_cast<double>(105.6583755f * 1e-3f);
    tk.n_trials = 0;
    tk.n_steps = 0;
    for (int i = 0; i < 6; ++i) tk.bound[i] = 0.0;
    for (int i = 0; i < 21; ++i) tk.cov[i] = 0.0;
    tk.up_sf = kInvalid;
    tk.n_seq = 0;
    tk.n_trace = 0;
    tk.momentum_hit = false;
    tk.target_hit = false;
    tk.stopped = false;

    EngineT<kExt> e{det, field, cfg, tk};
    const uint64_t N = n;
    if (bound0) {
        // bound start (propagator.hpp:121-162, base_stepper.hpp:84-108)
        const uint32_t s0 = start_sf[gid];
        for (int k = 0; k < 6; ++k) tk.bound[k] = bound0[k * N + gid];
        for (int r = 0; r < 6; ++r)
            for (int c = r; c < 6; ++c) {
                const int k = covi(r, c);
                tk.cov[k] = init_cov ? init_cov[k * N + gid] : (r == c ? 1.0 : 0.0);
            }
        tk.up_sf = s0;
        tk.pos = e.bound_pos(s0, tk.bound);
        tk.dir = Engine::bound_dir(tk.bound);
        tk.time = tk.bound[5];
        tk.qop = tk.bound[4];
        tk.volume = det.surfaces[s0].volume;
        init_cov = nullptr;
    } else {
        tk.pos = {fp[0 * N + gid], fp[1 * N + gid], fp[2 * N + gid]};
        tk.time = fp[3 * N + gid];
        tk.dir = {fp[4 * N + gid], fp[5 * N + gid], fp[6 * N + gid]};
        tk.qop = fp[7 * N + gid];
    }
    tk.charge = (-1.0 * tk.qop > 0.0) ? -1.0 : 1.0;
    e.reset_J();

    // navigator_base::init -> local_navigation without overstepping resolution
    e.local_navigation(cfg.overstep_tol, false);
    tk.heartbeat = tk.status > kUnknown;
    bool is_init = true;
    double path_length = 0.0;
    unsigned stall = 0;
    for (unsigned i = 0; i % 2 == 0 || tk.heartbeat; ++i) {
        if (i % 2 == 0) {
            // pathlimit_aborter (aborters.hpp:44-75)
            if (tk.status != kExit) {
                const double step_limit = cfg.path_limit - fabs(tk.abs_path_length);
                if (step_limit <= 0.0) {
                    tk.status = kAbort;
                    tk.path_limit_hit = true;
                    tk.heartbeat = false;
                }
                tk.constraint_abort = dmin(tk.constraint_abort, fabs(step_limit));
            }
            // parameter_updater<pointwise_material_interactor>
            if (cfg.use_material) e.parameter_updater(init_cov, n, gid);
            if (tk.stopped && tk.status > kUnknown) {
                tk.status = kAbort;
                tk.heartbeat = false;
            }
            // surface_sequencer (surface_sequencer.hpp:78-106)
            if (cfg.max_surfaces != 0u && (e.on_sensitive() || e.on_material())) {
                if (tk.n_seq == cfg.max_surfaces) {
                    tk.overflow = true;
                    tk.status = kExit;
                    tk.heartbeat = false;
                } else {
                    out.surfaces[static_cast<uint64_t>(tk.n_seq) * n + gid] = tk.c[tk.next - 1].sf;
                    tk.n_seq++;
                }
            }
            if (cfg.actors & DTR_ACTOR_MOMENTUM_ABORTER) e.momentum_aborter();
            if (cfg.actors & DTR_ACTOR_TARGET_ABORTER) e.target_aborter(cfg.target[gid]);
            if (!tk.heartbeat) continue;
            path_length = tk.path_length;
        } else {
            const bool reset = e.on_surface() || is_init;
            e.step(e.nav_dist(), reset);
            e.rk_policy();
            is_init = false;
            if (fabs(tk.path_length) <= fabs(path_length) + static_cast<double>(cfg.path_tol)) {
                if (stall >= 10u) {
                    tk.heartbeat = false;
                    tk.status = kAbort;
                }
                stall++;
            } else {
                stall = 0;
            }
        }
        is_init |= e.nav_update();
        tk.heartbeat = tk.heartbeat && (tk.status > kUnknown);
    }

    if (cfg.actors & DTR_ACTOR_STEP_TRACER) cfg.trace.n[gid] = tk.n_trace;
    if (out.status) out.status[gid] = track_status(tk.status == kExit, tk.overflow, tk.path_limit_hit,
                                                   tk.momentum_hit, tk.target_hit);
    if (out.bound_params)
        for (int k = 0; k < 6; ++k) out.bound_params[k * N + gid] = tk.bound[k];
    if (out.bound_cov)
        for (int k = 0; k < 21; ++k) out.bound_cov[k * N + gid] = tk.cov[k];
    if (out.free_params) {
        out.free_params[0 * N + gid] = tk.pos.x;
        out.free_params[1 * N + gid] = tk.pos.y;
        out.free_params[2 * N + gid] = tk.pos.z;
        out.free_params[3 * N + gid] = tk.time;
        out.free_params[4 * N + gid] = tk.dir.x;
        out.free_params[5 * N + gid] = tk.dir.y;
        out.free_params[6 * N + gid] = tk.dir.z;
        out.free_params[7 * N + gid] = tk.qop;
    }
    if (out.path_length) out.path_length[gid] = tk.path_length;
    if (out.n_surfaces) out.n_surfaces[gid] = tk.n_seq;
    if (out.n_steps) out.n_steps[gid] = tk.n_steps;
    if (out.n_trials) out.n_trials[gid] = tk.n_trials;
}


}  // namespace dtr

// Oracle: scalar FP64 CPU restatement of detray::propagator::propagate.
//
// TEST INFRASTRUCTURE.  This file restates the reference algorithm in plain
// C++ so that the device path (detray_amd) can be checked against it.  Each
// function cites the reference file:line it follows.  It is only loaded by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
//
// Scope (the propagation chain of tests/include/detray/test/device/
// propagator_test.hpp:87-99 plus surface_sequencer):
//   rk_stepper<field, algebra, constrained_step, stepper_rk_policy>
//   caching_navigator<detector, 8>
//   actor_chain<pathlimit_aborter,
//               parameter_updater<pointwise_material_interactor>,
//               surface_sequencer>
// No volume material (toy / telescope detectors have none), so d(qop)/ds = 0
// inside volumes (rk_stepper.hpp:148-188 returns 0 without material).

#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

constexpr double kMax = std::numeric_limits<double>::max();
constexpr uint32_t kInvalid = 0xffffffffu;
constexpr uint16_t kEndOfWorld = 0xffffu;
constexpr int kCache = 8;  // navigation::default_cache_size

// ----------------------------------------------------------------------------
// Small vector algebra (generic_vector.hpp, array_vector.hpp)
// ----------------------------------------------------------------------------
struct V3 {
    double v[3];
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
inline V3 add(const V3& a, const V3& b) { return {{a[0] + b[0], a[1] + b[1], a[2] + b[2]}}; }
inline V3 sub(const V3& a, const V3& b) { return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
inline V3 mul(double s, const V3& a) { return {{s * a[0], s * a[1], s * a[2]}}; }
// array_vector.hpp:104-116: tmp products summed left to right from 0
inline double dot(const V3& a, const V3& b) {
    double r = 0.0;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}
inline V3 cross(const V3& a, const V3& b) {
    return {{a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]}};
}
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
// generic_vector.hpp:35-55
inline double perp(const V3& a) { return std::sqrt(std::fma(a[0], a[0], a[1] * a[1])); }
inline double phi(const V3& a) { return std::atan2(a[1], a[0]); }
inline double theta(const V3& a) { return std::atan2(perp(a), a[2]); }
inline V3 normalize(const V3& a) {
    const double n = norm(a);
    return {{a[0] / n, a[1] / n, a[2] / n}};
}

// Transform helpers: global = R local + t
inline V3 trf_axis(const dtr_transform& t, int c) { return {{t.r[c], t.r[3 + c], t.r[6 + c]}}; }
inline V3 trf_translation(const dtr_transform& t) { return {{t.t[0], t.t[1], t.t[2]}}; }
inline V3 point_to_local(const dtr_transform& t, const V3& p) {
    const V3 d = sub(p, trf_translation(t));
    return {{t.r[0] * d[0] + t.r[3] * d[1] + t.r[6] * d[2],
             t.r[1] * d[0] + t.r[4] * d[1] + t.r[7] * d[2],
             t.r[2] * d[0] + t.r[5] * d[1] + t.r[8] * d[2]}};
}
inline V3 point_to_global(const dtr_transform& t, const V3& l) {
    return {{t.r[0] * l[0] + t.r[1] * l[1] + t.r[2] * l[2] + t.t[0],
             t.r[3] * l[0] + t.r[4] * l[1] + t.r[5] * l[2] + t.t[1],
             t.r[6] * l[0] + t.r[7] * l[1] + t.r[8] * l[2] + t.t[2]}};
}
inline V3 vector_to_global(const dtr_transform& t, const V3& l) {
    return {{t.r[0] * l[0] + t.r[1] * l[1] + t.r[2] * l[2],
             t.r[3] * l[0] + t.r[4] * l[1] + t.r[5] * l[2],
             t.r[6] * l[0] + t.r[7] * l[1] + t.r[8] * l[2]}};
}

// ----------------------------------------------------------------------------
// Field (tests/include/detray/test/common/bfield.hpp:35-47: constant or
// affine -> linear -> clamp -> strided xyz grid)
// ----------------------------------------------------------------------------
struct Field {
    const dtr_field_desc* d;
    V3 at(const V3& p) const {
        if (d->kind == DTR_FIELD_CONSTANT) return {{d->b[0], d->b[1], d->b[2]}};
        int i0[3];
        double t[3];
        for (int a = 0; a < 3; ++a) {
            const int n = static_cast<int>(d->n[a]);
            double u = (p[a] - d->min[a]) / d->pitch[a];
            // clamp (NaN positions clamp to the first sample)
            u = (u > 0.0) ? (u > n - 1 ? static_cast<double>(n - 1) : u) : 0.0;
            int i = static_cast<int>(std::floor(u));
            i = i > n - 2 ? n - 2 : i;
            i0[a] = i;
            t[a] = u - i;
        }
        const uint64_t ny = d->n[1], nz = d->n[2];
        V3 out{{0, 0, 0}};
        for (int c = 0; c < 3; ++c) {
            auto v = [&](int dx, int dy, int dz) {
                const uint64_t idx = ((static_cast<uint64_t>(i0[0] + dx) * ny + (i0[1] + dy)) * nz +
                                      (i0[2] + dz));
                return static_cast<double>(d->values[3 * idx + c]);
            };
            const double c00 = v(0, 0, 0) * (1.0 - t[0]) + v(1, 0, 0) * t[0];
            const double c01 = v(0, 0, 1) * (1.0 - t[0]) + v(1, 0, 1) * t[0];
            const double c10 = v(0, 1, 0) * (1.0 - t[0]) + v(1, 1, 0) * t[0];
            const double c11 = v(0, 1, 1) * (1.0 - t[0]) + v(1, 1, 1) * t[0];
            const double c0 = c00 * (1.0 - t[1]) + c10 * t[1];
            const double c1 = c01 * (1.0 - t[1]) + c11 * t[1];
            out[c] = c0 * (1.0 - t[2]) + c1 * t[2];
        }
        return out;
    }
};

// ----------------------------------------------------------------------------
// Detector view
// ----------------------------------------------------------------------------
struct Det {
    const dtr_detector_desc* d;
    const dtr_surface& sf(uint32_t i) const { return d->surfaces[i]; }
    const dtr_transform& trf(uint32_t s) const { return d->transforms[d->surfaces[s].transform]; }
    const dtr_mask& mask(uint32_t m) const { return d->masks[m]; }
};

// ----------------------------------------------------------------------------
// Particle (pdg_particle.hpp:61-91): muon / antimuon
// ----------------------------------------------------------------------------
struct Particle {
    double mass;
    double charge;
    int pdg;
};
inline Particle muon_for(double qop) {
    const double m = static_cast<double>(105.6583755f * 1e-3f);
    Particle mu{m, -1.0, 13};
    if (mu.charge * qop > 0.0) return mu;
    return {m, 1.0, -13};  // charge_conjugation
}

// ----------------------------------------------------------------------------
// Intersections (intersection.hpp:30-35, 190-430)
// ----------------------------------------------------------------------------
enum IStatus : uint8_t { kInside = 0, kEdge = 1, kOutside = 2 };

struct Cand {
    double path = kMax;
    uint32_t sf = kInvalid;  // surface index, invalid for never-filled slots
    uint16_t vol_link = 0;
    uint8_t status = kOutside;
    uint8_t dir = 1;
    bool probably_inside() const { return status <= kEdge; }
};

enum NavStatus : int8_t {
    kAbort = -3, kExit = -2, kUnknown = -1, kTowards = 0, kOnObject = 1, kOnPortal = 2
};
enum Trust : uint8_t { kNoTrust = 0, kFair = 1, kHigh = 3, kFull = 4 };

struct NavCfg {  // intersection::config + navigation::config (floats, as reference)
    float min_mask_tol, max_mask_tol, mask_scalor, path_tol, overstep_tol;
    uint32_t win[2];
};

struct Nav {
    Cand c[kCache];
    int8_t next = 0, last = -1;
    int8_t status = kUnknown;
    uint8_t trust = kNoTrust;
    int8_t direction = 1;
    uint16_t volume = 0;
    double ext_tol = 0.0;
    bool exited_path_limit = false;

    bool is_alive() const { return status > kUnknown; }
    bool on_surface() const { return status == kOnObject || status == kOnPortal; }
    bool on_portal() const { return status == kOnPortal; }
    int n_candidates() const { return last - next + 1; }
    bool cache_exhausted() const { return n_candidates() == 0; }
    Cand& current() { return c[next - 1]; }
    const Cand& current() const { return c[next - 1]; }
    Cand& target() { return c[next]; }
    const Cand& target() const { return c[next]; }
    // navigation_state.hpp begin()/end()
    int begin_idx() const { return (on_surface() && next >= 1) ? next - 1 : next; }
    int end_idx() const { return last + 1; }
    double operator()() const { return static_cast<double>(direction) * target().path; }
    // caching_navigator.hpp:126-146 trust setters only lower the level
    void set_high_trust() { trust = trust < kHigh ? trust : static_cast<uint8_t>(kHigh); }
    void set_fair_trust() { trust = trust < kFair ? trust : static_cast<uint8_t>(kFair); }
    void set_no_trust() { trust = kNoTrust; }
    void set_volume(uint16_t v) {
        if (v != volume) set_no_trust();
        volume = v;
    }
    void clear_cache() {
        for (int i = 0; i < kCache; ++i) c[i].path = kMax;
        next = 0;
        last = -1;
    }
    void exit_nav() { status = kExit; }
    void abort_nav() { status = kAbort; }
};

// ----------------------------------------------------------------------------
// Masks (shapes/*.hpp check_boundaries; mask.hpp:173-192)
// ----------------------------------------------------------------------------
// Returns {inside, inside_with_edge}
inline void check_planar(const dtr_mask& m, const dtr_transform& t, const V3& glob,
                         double tol, double edge_tol, bool& inside, bool& edge) {
    const V3 loc = point_to_local(t, glob);
    if (m.shape == DTR_RECTANGLE2D) {
        // rectangle2D.hpp:104-125
        const double l0 = std::fabs(loc[0]), l1 = std::fabs(loc[1]);
        inside = (l0 <= (m.v[0] + tol) && l1 <= (m.v[1] + tol));
        edge = false;
        if (edge_tol > 0.0) {
            const double ft = tol + edge_tol;
            edge = (l0 <= (m.v[0] + ft) && l1 <= (m.v[1] + ft));
        }
    } else if (m.shape == DTR_TRAPEZOID2D) {
        // trapezoid2D.hpp:121-150
        const double rel_y = (m.v[2] + loc[1]) * m.v[3];
        const double bound_x = m.v[0] + rel_y * (m.v[1] - m.v[0]);
        inside = (std::fabs(loc[0]) <= (bound_x + tol) && std::fabs(loc[1]) <= (m.v[2] + tol));
        edge = false;
        if (edge_tol > 0.0) {
            const double ft = tol + edge_tol;
            edge = (std::fabs(loc[0]) <= (bound_x + ft) && std::fabs(loc[1]) <= (m.v[2] + ft));
        }
    } else if (m.shape == DTR_ANNULUS2D) {
        // annulus2D.hpp:140-190 (global point): radius in the beam frame
        // (shifted), phi in the strip frame relative to the average phi,
        // phi tolerance = tol / r (shape_utils.hpp:26-28)
        const double nx = loc[0] + m.v[5], ny = loc[1] + m.v[6];
        const double r_beam = std::sqrt(std::fma(nx, nx, ny * ny));
        auto phi_tol = [r_beam](double t) { return r_beam > 0.0 ? t / r_beam : t; };
        double phi_focal = 0.0;
        bool have_phi = false;
        inside = ((r_beam + tol) >= m.v[0]) && (r_beam <= (m.v[1] + tol));
        if (inside) {
            phi_focal = phi(loc) - m.v[4];
            have_phi = true;
            const double pt = phi_tol(tol);
            inside = (phi_focal >= (m.v[2] - pt)) && (phi_focal <= (m.v[3] + pt));
        }
        edge = false;
        if (edge_tol > 0.0) {
            const double ft = tol + edge_tol;
            edge = ((r_beam + ft) >= m.v[0]) && (r_beam <= (m.v[1] + ft));
            if (edge) {
                if (!have_phi) phi_focal = phi(loc) - m.v[4];
                const double pt = phi_tol(ft);
                edge = (phi_focal >= (m.v[2] - pt)) && (phi_focal <= (m.v[3] + pt));
            }
        }
    } else if (m.shape == DTR_CYLINDER2D) {
        // cylinder2D.hpp:63-90: the local z within [lower_z, upper_z]
        inside = ((m.v[1] - tol) <= loc[2] && loc[2] <= (m.v[2] + tol));
        edge = false;
        if (edge_tol > 0.0) {
            const double ft = tol + edge_tol;
            edge = ((m.v[1] - ft) <= loc[2] && loc[2] <= (m.v[2] + ft));
        }
    } else {
        // ring2D.hpp:85-120: radius of the point in the local cartesian frame
        const double r = perp(loc);
        inside = ((r + tol) >= m.v[0] && r <= (m.v[1] + tol));
        edge = inside;
    }
}

// line.hpp:86-121 check_boundaries on the global point: the local cartesian
// position inside the square cross section (drift cell) or the circle
// (straw tube) and within the half length
inline void check_line(const dtr_mask& m, const dtr_transform& t, const V3& glob, double tol,
                       double edge_tol, bool& inside, bool& edge) {
    const V3 loc = point_to_local(t, glob);
    double a, b;
    const double c = std::fabs(loc[2]);
    if (m.shape == DTR_LINE_SQUARE) {
        a = std::fabs(loc[0]);
        b = std::fabs(loc[1]);
        inside = (a <= (m.v[0] + tol) && b <= (m.v[0] + tol) && c <= (m.v[1] + tol));
    } else {
        a = std::fabs(perp(loc));
        b = 0.0;
        inside = (a <= (m.v[0] + tol) && c <= (m.v[1] + tol));
    }
    edge = false;
    if (edge_tol > 0.0) {
        const double ft = tol + edge_tol;
        edge = (a <= (m.v[0] + ft) && b <= (m.v[0] + ft) && c <= (m.v[1] + ft));
    }
}

inline bool is_line(uint8_t shape) { return shape == DTR_LINE_SQUARE || shape == DTR_LINE_CIRCULAR; }

// intersector_base.hpp:171-252 resolve_mask (ip.point is global 3D for
// planar intersectors and (inv, z) for the concentric cylinder)
inline void resolve_mask(Cand& is, double path, const V3& glob, double z2d,
                         uint32_t sf_idx, const dtr_surface& sf, const dtr_mask& m,
                         const dtr_transform& t, const NavCfg& cfg, double ext_tol_in) {
    is.status = kOutside;
    if (!(path >= static_cast<double>(cfg.overstep_tol))) return;
    double base_tol = 0.0, ext_tol = 0.0;
    if (sf.sf_id != DTR_PORTAL || cfg.min_mask_tol == std::numeric_limits<float>::max()) {
        ext_tol = ext_tol_in;
        base_tol = std::max(static_cast<double>(cfg.min_mask_tol),
                            std::min(static_cast<double>(cfg.max_mask_tol),
                                     static_cast<double>(cfg.mask_scalor) * std::fabs(path)));
    }
    bool inside, edge;
    if (m.shape == DTR_CONCENTRIC_CYLINDER2D) {
        // concentric_cylinder2D.hpp:110-120 (z only)
        inside = (m.v[1] - base_tol <= z2d && z2d <= m.v[2] + base_tol);
        edge = inside;
    } else if (is_line(m.shape)) {
        check_line(m, t, glob, base_tol, ext_tol, inside, edge);
    } else {
        check_planar(m, t, glob, base_tol, ext_tol, inside, edge);
    }
    if (edge) is.status = kEdge;
    if (inside) is.status = kInside;
    is.path = path;
    is.sf = sf_idx;
    is.dir = !std::signbit(path);
    is.vol_link = m.volume_link;
}

// ray_plane_intersector.hpp:60-87.  Returns validity (|s| < 10 m).
inline bool plane_intersect(const V3& ro, const V3& rd, const dtr_transform& t,
                            double& s, V3& glob) {
    const V3 sn = trf_axis(t, 2);
    const V3 st = trf_translation(t);
    const double denom = dot(rd, sn);
    if (denom == 0.0) {
        s = kMax;
        return false;
    }
    s = dot(sn, sub(st, ro)) / denom;
    glob = add(ro, mul(s, rd));
    return std::fabs(s) < 10.0 * 1000.0;
}

// ray_line_intersector.hpp:44-84: closest approach of the ray to the wire
// axis; no solution for a ray (nearly) parallel to the wire
inline bool line_intersect(const V3& ro, const V3& rd, const dtr_transform& t, double& s, V3& glob) {
    const V3 z = trf_axis(t, 2);
    const V3 tt = trf_translation(t);
    const double zd = dot(z, rd);
    const double denom = 1.0 - (zd * zd);
    if (denom < static_cast<double>(1e-5f)) {
        s = kMax;
        return false;
    }
    const V3 t2l = sub(tt, ro);
    const double t2l_on_line = dot(t2l, z);
    const double t2l_on_track = dot(t2l, rd);
    s = 1.0 / denom * (t2l_on_track - t2l_on_line * zd);
    glob = add(ro, mul(s, rd));
    return std::fabs(s) < 10.0 * 1000.0;
}

// ray_cylinder_intersector.hpp:40-110 with quadratic_equation.hpp:30-75:
// the solutions of |(ro + s rd - c) x z|^2 = r^2, ascending; n = 0, 1 or 2
inline int cyl2_intersect(const V3& ro, const V3& rd, const dtr_transform& t, double r, double s[2]) {
    const V3 sz = trf_axis(t, 2);
    const V3 sc = trf_translation(t);
    const V3 pc_cross_sz = cross(sub(ro, sc), sz);
    const V3 rd_cross_sz = cross(rd, sz);
    const double a = dot(rd_cross_sz, rd_cross_sz);
    const double b = 2.0 * dot(rd_cross_sz, pc_cross_sz);
    const double c = dot(pc_cross_sz, pc_cross_sz) - (r * r);
    const double eps = std::numeric_limits<double>::epsilon();
    s[0] = s[1] = kMax;
    if (std::fabs(a) <= eps) {
        if (std::fabs(b) <= eps) return 0;
        s[0] = -c / b;
        return 1;
    }
    const double disc = b * b - 4.0 * a * c;
    if (disc > eps) {
        const double q = -0.5 * (b + std::copysign(std::sqrt(disc), b));
        s[0] = q / a;
        s[1] = c / q;
        if (s[0] > s[1]) std::swap(s[0], s[1]);
        return 2;
    }
    if (disc >= 0.0) {
        s[0] = -0.5 * b / a;
        return 1;
    }
    return 0;
}

// ray_cylinder_portal_intersector.hpp:68-132
inline bool cyl_portal_intersect(const V3& ro, const V3& rd, double r, double overstep,
                                 double& path, double& z) {
    const double inv = kMax;
    path = inv;
    const double rd_perp_2 = rd[0] * rd[0] + rd[1] * rd[1];
    if (rd_perp_2 < std::numeric_limits<double>::epsilon()) {
        z = 0.0;
        return false;
    }
    const double ro_perp_2 = ro[0] * ro[0] + ro[1] * ro[1];
    const double rad_diff = r * r - ro_perp_2;
    if (std::fabs(rad_diff) > 1.0 * 1e-3) {
        const double rd_perp_inv_2 = 1.0 / rd_perp_2;
        const double k = -rd_perp_inv_2 * (ro[0] * rd[0] + ro[1] * rd[1]);
        const double discr = rd_perp_inv_2 * rad_diff + k * k;
        if (discr < 0.0) {
            z = 0.0;
            return false;
        }
        const double sq = std::sqrt(discr);
        const double s1 = k + sq, s2 = k - sq;
        path = (std::fabs(s1) < std::fabs(s2)) ? s1 : s2;
        if (path < overstep && sq > 0.0) path = (std::fabs(s1) >= std::fabs(s2)) ? s1 : s2;
    } else {
        path = rad_diff;
    }
    if (path < overstep) path = inv;
    z = ro[2] + path * rd[2];
    return std::fabs(path) < 10.0 * 1000.0;
}

// Compute the intersection of a ray with a surface (the point of
// intersection used by all of its masks).
struct IPoint {
    double path;
    V3 glob;
    double z;
};
// Ray-surface intersections computed by the navigation of the current track
// (the N_isect of the algorithmic flop count), per thread.
thread_local uint32_t tl_isect = 0u;
inline bool intersect_surface(const Det& det, uint32_t sf_idx, const V3& ro, const V3& rd,
                              double overstep, IPoint& ip) {
    ++tl_isect;
    const dtr_surface& s = det.sf(sf_idx);
    const dtr_transform& t = det.trf(sf_idx);
    if (s.shape == DTR_CONCENTRIC_CYLINDER2D) {
        const double r = det.mask(s.mask).v[0];
        return cyl_portal_intersect(ro, rd, r, overstep, ip.path, ip.z);
    }
    if (is_line(s.shape)) return line_intersect(ro, rd, t, ip.path, ip.glob);
    if (s.shape == DTR_CYLINDER2D) {
        // intersection_update (intersection_kernel.hpp:150-200) takes the
        // smaller solution
        double sol[2];
        if (cyl2_intersect(ro, rd, t, det.mask(s.mask).v[0], sol) == 0) return false;
        ip.path = sol[0];
        ip.glob = add(ro, mul(sol[0], rd));
        return std::fabs(sol[0]) < 10.0 * 1000.0;
    }
    return plane_intersect(ro, rd, t, ip.path, ip.glob);
}

struct Engine;  // forward

// ----------------------------------------------------------------------------
// The propagation engine for one track
// ----------------------------------------------------------------------------
struct PropCfg {
    // stepping
    double min_stepsize, rk_error_tol, step_constraint, path_limit;
    uint32_t max_rk_updates;
    bool do_cov;
    bool use_field_gradient;
    bool use_mean_loss = true, use_eloss_gradient = false;
    // navigation
    NavCfg nav;
    double accumulated_error;
    int n_stddev;
    bool estimate_noise;
    // interactor
    bool do_eloss, do_ms, use_material;
    uint32_t max_surfaces;
    // optional actors (dtr_actor bits) and the momentum_aborter limits
    uint32_t actors = 0u;
    double min_p = 0.0, min_pT = 0.0;
};

struct Stepper {
    // free track parameters
    V3 pos, dir;
    double time, qop;
    // transport jacobian, dense 8x8 row-major
    double J[64];
    Particle ptc;
    uint64_t n_trials = 0;
    uint32_t n_steps = 0;
    double step_size = 0.0, next_step_size = 0.0;
    double path_length = 0.0, abs_path_length = 0.0;
    double constraints[4] = {kMax, kMax, kMax, kMax};  // accuracy, actor, aborter, user
    V3 dtds3{{0, 0, 0}};
    double dqopds3 = 0.0;

    void reset_J() {
        for (int i = 0; i < 64; ++i) J[i] = 0.0;
        for (int i = 0; i < 8; ++i) J[i * 8 + i] = 1.0;
    }
    // constrained_step.hpp:74-125
    void set_constraint(int type, double s) { constraints[type] = std::min(constraints[type], std::fabs(s)); }
    double constraint_size() const {
        double m = kMax;
        for (int i = 0; i < 4; ++i) m = std::min(m, constraints[i]);
        const double dir_sign = step_size >= 0.0 ? 1.0 : -1.0;
        return dir_sign * m;
    }
};

struct Updater {
    double bound[6];
    double cov[36];
    uint32_t sf = kInvalid;  // surface link (invalid: curvilinear/free start)
};

struct Engine {
    Det det;
    Field field;
    PropCfg cfg;
    Nav nav;
    Stepper st;
    Updater up;
    bool heartbeat = false;
    bool stopped = false;  // energy loss brought the particle to rest
    const double* init_cov = nullptr;  // 21 values or null
    // sequencer
    uint32_t* seq = nullptr;
    uint32_t seq_stride = 0;
    uint32_t n_seq = 0;
    bool overflow = false;
    // bound loc0/loc1 at the sensitive surfaces (oracle_propagate_hits)
    uint32_t* hit_sf = nullptr;
    double* hit_loc = nullptr;  // loc0 at [k * hit_stride], loc1 at [(hit_cap + k) * hit_stride]
    uint64_t hit_stride = 0;
    uint32_t hit_cap = 0, n_hit = 0;
    // trace
    uint32_t* tr_vol = nullptr;
    uint32_t* tr_sf = nullptr;
    uint32_t tr_cap = 0, tr_n = 0;
    // optional actors: target_aborter::state, step_tracer outputs of this
    // track (entry k at [k * st_stride], arrays of several values at
    // [(v * cap + k) * st_stride])
    uint32_t target = kInvalid;
    bool momentum_hit = false, target_hit = false;
    dtr_step_trace steps{};
    uint64_t st_stride = 0;
    uint32_t n_trace = 0;

    // ---------------------------------------------------------------- nav --
    bool is_portal(uint32_t s) const { return det.sf(s).sf_id == DTR_PORTAL; }

    void insert(int pos, const Cand& nc) {
        // caching_navigator.hpp:149-229
        if (pos == kCache) return;
        if (nav.n_candidates() == 0) {
            nav.c[0] = nc;
            nav.last = static_cast<int8_t>(nav.last + 1);
            return;
        }
        int idx = pos;
        auto overlap = [&](int i) { return std::fabs(nav.c[i].path - nc.path) <= 1.0 * 1e-3; };
        auto clash = [&](int i) {
            const uint64_t a = nav.c[i].sf == kInvalid ? ~0ull : det.sf(nav.c[i].sf).barcode;
            return a == det.sf(nc.sf).barcode && overlap(i);
        };
        if (clash(idx) || (idx > 0 && clash(idx - 1))) return;
        const int shift_max = kCache - 2;
        const int shift_begin = std::min(static_cast<int>(nav.last), shift_max);
        if (overlap(idx) && (nc.status == kEdge || is_portal(nc.sf))) {
            idx++;
            idx = idx > shift_begin ? shift_begin : idx;
        }
        for (int i = shift_begin; i >= idx; --i) nav.c[i + 1] = nav.c[i];
        nav.c[idx] = nc;
        nav.last = static_cast<int8_t>(std::min(nav.last + 1, kCache - 1));
    }

    // intersection_kernel.hpp:24-168 intersection_initialize + insert_sorted
    void candidate_search(uint32_t sf_idx, const NavCfg& icfg) {
        const V3 rd = mul(static_cast<double>(nav.direction), st.dir);
        const dtr_surface& s = det.sf(sf_idx);
        const dtr_transform& t = det.trf(sf_idx);
        if (s.shape == DTR_CYLINDER2D) {
            // intersection_initialize (intersection_kernel.hpp:24-110): every
            // valid solution that a mask accepts is inserted, at most two
            double sol[2];
            ++tl_isect;
            const int n = cyl2_intersect(st.pos, rd, t, det.mask(s.mask).v[0], sol);
            for (uint32_t k = 0; k < s.n_masks; ++k) {
                int found = 0;
                for (int i = 0; i < n; ++i) {
                    if (!(std::fabs(sol[i]) < 10.0 * 1000.0)) continue;
                    Cand is;
                    resolve_mask(is, sol[i], add(st.pos, mul(sol[i], rd)), 0.0, sf_idx, s,
                                 det.mask(s.mask + k), t, icfg, nav.ext_tol);
                    if (is.probably_inside()) {
                        int p = nav.begin_idx();
                        const int e = nav.end_idx();
                        while (p < e && !(std::fabs(is.path) < std::fabs(nav.c[p].path))) ++p;
                        insert(p, is);
                        if (++found == 2) return;
                    }
                }
            }
            return;
        }
        IPoint ip;
        if (!intersect_surface(det, sf_idx, st.pos, rd, icfg.overstep_tol, ip)) return;
        for (uint32_t k = 0; k < s.n_masks; ++k) {
            Cand is;
            resolve_mask(is, ip.path, ip.glob, ip.z, sf_idx, s, det.mask(s.mask + k), t, icfg,
                         nav.ext_tol);
            if (is.probably_inside()) {
                // upper_bound over [begin, end) with |path| ordering
                int p = nav.begin_idx();
                const int e = nav.end_idx();
                while (p < e && !(std::fabs(is.path) < std::fabs(nav.c[p].path))) ++p;
                insert(p, is);
                return;
            }
        }
    }

    // navigation_functions.hpp:99-129 update_candidate + intersection_update
    bool update_candidate(Cand& cand) {
        if (cand.sf == kInvalid) return false;
        const V3 rd = mul(static_cast<double>(nav.direction), st.dir);
        IPoint ip;
        if (!intersect_surface(det, cand.sf, st.pos, rd, cfg.nav.overstep_tol, ip)) return false;
        const dtr_surface& s = det.sf(cand.sf);
        const dtr_transform& t = det.trf(cand.sf);
        for (uint32_t k = 0; k < s.n_masks; ++k) {
            resolve_mask(cand, ip.path, ip.glob, ip.z, cand.sf, s, det.mask(s.mask + k), t,
                         cfg.nav, nav.ext_tol);
            if (cand.probably_inside()) return true;
        }
        return false;
    }

    // update_candidate with an explicit intersection config
    bool update_candidate(Cand& cand, const NavCfg& icfg) {
        if (cand.sf == kInvalid) return false;
        const V3 rd = mul(static_cast<double>(nav.direction), st.dir);
        IPoint ip;
        if (!intersect_surface(det, cand.sf, st.pos, rd, icfg.overstep_tol, ip)) return false;
        const dtr_surface& s = det.sf(cand.sf);
        const dtr_transform& t = det.trf(cand.sf);
        for (uint32_t k = 0; k < s.n_masks; ++k) {
            resolve_mask(cand, ip.path, ip.glob, ip.z, cand.sf, s, det.mask(s.mask + k), t, icfg, nav.ext_tol);
            if (cand.probably_inside()) return true;
        }
        return false;
    }

    // ------------------------------------------------------ direct navigator --
    // direct_navigator.hpp:29-298: the target is always cache entry 1, the
    // current surface entry 0; the targets are the given surface sequence.
    const uint32_t* dseq = nullptr;
    uint64_t dseq_stride = 0;
    uint32_t dseq_len = 0, dseq_next = 0;

    // the k-th target: entry k of the sequence going forward, entry
    // len - 1 - k going backward (direct_navigator.hpp:79-94 reset() starts
    // at the first or the last entry and advance() moves in the navigation
    // direction, :125-151)
    uint32_t dseq_at(uint32_t k) const {
        const uint32_t e = nav.direction > 0 ? k : dseq_len - 1u - k;
        return dseq[static_cast<uint64_t>(e) * dseq_stride];
    }

    // direct_navigator::update (direct_navigator.hpp:210-298)
    bool direct_update() {
        if (dseq_next >= dseq_len) {
            nav.exit_nav();
            return false;
        }
        if (nav.trust == kFull) return false;
        // infinite tolerances: the target is always hit
        NavCfg icfg = cfg.nav;
        icfg.min_mask_tol = std::numeric_limits<float>::max();
        icfg.max_mask_tol = std::numeric_limits<float>::max();
        icfg.mask_scalor = std::numeric_limits<float>::max();
        icfg.overstep_tol = -std::numeric_limits<float>::max();
        // a sequence entry that is not a surface of the detector ends the
        // navigation like an unreachable target
        if (nav.c[1].sf >= det.d->n_surfaces || !update_candidate(nav.c[1], icfg)) {
            nav.abort_nav();
            return false;
        }
        // navigation::update_status (navigation_functions.hpp:148-177) with
        // the direct state's advance()
        if (std::fabs(nav.c[1].path) < static_cast<double>(cfg.nav.path_tol)) {
            nav.status = is_portal(nav.c[1].sf) ? kOnPortal : kOnObject;
            nav.c[0] = nav.c[1];
            dseq_next++;
            if (dseq_next < dseq_len) nav.c[1].sf = dseq_at(dseq_next);
        } else {
            nav.status = kTowards;
        }
        nav.trust = kFull;
        if (nav.status == kTowards) return false;
        if (dseq_next < dseq_len && (nav.c[1].sf >= det.d->n_surfaces || !update_candidate(nav.c[1], icfg))) {
            nav.abort_nav();
            return false;
        }
        nav.set_volume(det.sf(nav.c[0].sf).volume);
        nav.trust = kFull;
        return true;
    }

    // spatial_grid.hpp:100-202 + axis.hpp:417-424 + axis_binning.hpp:80-115
    void grid_search(const dtr_volume& vol, const NavCfg& icfg) {
        const dtr_grid& g = det.d->grids[vol.grid];
        const dtr_transform& vt = det.d->transforms[vol.transform];
        double l0, l1;
        ++tl_isect;  // the grid lookup intersects the grid's reference surface
        if (g.type == DTR_GRID_CYLINDER) {
            double path, z;
            cyl_portal_intersect(st.pos, st.dir, g.radius, -kMax, path, z);
            const V3 p = add(st.pos, mul(path, st.dir));
            l0 = phi(p);
            l1 = z;
        } else {
            double s;
            V3 glob{{0, 0, 0}};
            plane_intersect(st.pos, st.dir, vt, s, glob);
            const V3 loc = point_to_local(vt, glob);
            l0 = perp(loc);
            l1 = phi(loc);
        }
        const double w0 = (g.max0 - g.min0) / static_cast<double>(g.n0);
        const double w1 = (g.max1 - g.min1) / static_cast<double>(g.n1);
        // a non-finite lookup point has no defined bin (the reference casts
        // it to int): no sensitive candidates from the grid then
        const double f0 = (l0 - g.min0) / w0 + 1.0, f1 = (l1 - g.min1) / w1 + 1.0;
        if (!(std::fabs(f0) < 1e9) || !(std::fabs(f1) < 1e9)) return;
        const int ib0 = static_cast<int>((l0 - g.min0) / w0 + 1.0) - 1;
        const int ib1 = static_cast<int>((l1 - g.min1) / w1 + 1.0) - 1;
        int r0[2] = {ib0 - static_cast<int>(icfg.win[0]), ib0 + static_cast<int>(icfg.win[1])};
        int r1[2] = {ib1 - static_cast<int>(icfg.win[0]), ib1 + static_cast<int>(icfg.win[1])};
        auto closed = [](int lb, int ub, int bins, int out[2]) {
            int mn = (lb > 0) ? lb : 0;
            mn = (mn >= bins) ? bins - 1 : mn;
            int mx = (ub >= bins) ? bins - 1 : ub;
            mx = (mx < 0) ? 0 : mx;
            out[0] = mn;
            out[1] = mx + 1;
        };
        auto circular = [](int lb, int ub, int out[2]) {
            out[0] = lb;
            out[1] = ub + 1;
        };
        int a0[2], a1[2];
        const bool circ0 = (g.type == DTR_GRID_CYLINDER);
        if (circ0) {
            circular(r0[0], r0[1], a0);
            closed(r1[0], r1[1], static_cast<int>(g.n1), a1);
        } else {
            closed(r0[0], r0[1], static_cast<int>(g.n0), a0);
            circular(r1[0], r1[1], a1);
        }
        auto wrap = [](int b, int n) { return (n + (b % n)) % n; };
        for (int b0 = a0[0]; b0 < a0[1]; ++b0) {
            for (int b1 = a1[0]; b1 < a1[1]; ++b1) {
                const int i0 = circ0 ? wrap(b0, static_cast<int>(g.n0)) : b0;
                const int i1 = circ0 ? b1 : wrap(b1, static_cast<int>(g.n1));
                // every entry of the bin, in the order the bin was filled
                const uint32_t cap = g.capacity ? g.capacity : 1u;
                const uint64_t first = g.bins + (static_cast<uint64_t>(i0) * g.n1 + i1) * cap;
                for (uint32_t k = 0; k < cap; ++k) {
                    const uint32_t s = det.d->grid_bins[first + k];
                    if (s != kInvalid) candidate_search(s, icfg);
                }
            }
        }
    }

    // navigation_functions.hpp:147-177
    void update_status() {
        if (!nav.cache_exhausted() &&
            std::fabs(nav.target().path) < static_cast<double>(cfg.nav.path_tol)) {
            nav.status = is_portal(nav.target().sf) ? kOnPortal : kOnObject;
            nav.next = static_cast<int8_t>(nav.next + 1);
        } else {
            nav.status = kTowards;
        }
        nav.trust = (nav.cache_exhausted() || nav.on_portal()) ? kNoTrust : kFull;
    }

    // navigation_functions.hpp:193-246
    void local_navigation(const NavCfg& ncfg, bool resolve_overstepping) {
        const dtr_volume& vol = det.d->volumes[nav.volume];
        nav.clear_cache();
        NavCfg icfg = ncfg;
        icfg.overstep_tol = resolve_overstepping ? ncfg.overstep_tol : -ncfg.path_tol;
        // tracking_volume.hpp:263-278: brute force (portals/passives) first,
        // then the sensitive-surface grid
        for (uint32_t k = vol.bf_begin; k < vol.bf_end; ++k) candidate_search(det.d->bf_surfaces[k], icfg);
        if (vol.grid >= 0) grid_search(vol, icfg);
        update_status();
        if (nav.trust != kFull) {
            if (nav.on_portal()) {
                nav.trust = (nav.current().vol_link == kEndOfWorld) ? kFull : kNoTrust;
            }
        }
    }

    // navigation_functions.hpp:264-288
    void volume_switch() {
        if (nav.current().vol_link == kEndOfWorld) {
            nav.exit_nav();
            return;
        }
        nav.set_volume(nav.current().vol_link);
        local_navigation(cfg.nav, true);
        nav.trust = kFull;
    }

    // navigation_functions.hpp:304-331
    void init_loose_cfg() {
        NavCfg loose = cfg.nav;
        loose.overstep_tol = std::min(100.f * loose.overstep_tol, -10.f * loose.max_mask_tol);
        local_navigation(loose, true);
        if (nav.trust != kFull || nav.cache_exhausted()) nav.abort_nav();
    }

    // caching_navigator.hpp:250-388
    bool update_impl() {
        const bool is_init = true;
        if (nav.trust == kHigh) {
            if (!update_candidate(nav.target())) {
                nav.status = kUnknown;
                nav.set_fair_trust();
            } else {
                update_status();
                if (nav.status == kTowards || nav.on_portal()) return !is_init;
                if (nav.trust == kFull && update_candidate(nav.target())) return !is_init;
                nav.set_fair_trust();
            }
        }
        if (nav.trust == kFair && !nav.cache_exhausted()) {
            const int b = nav.begin_idx(), e = nav.end_idx();
            for (int i = b; i < e; ++i) {
                if (!update_candidate(nav.c[i])) nav.c[i].path = kMax;
            }
            // sort.hpp:41-58 sequential (selection) sort over [b, e)
            for (int i = b; i < e; ++i) {
                for (int j = i + 1; j < e; ++j) {
                    if (std::fabs(nav.c[j].path) < std::fabs(nav.c[i].path)) {
                        const Cand tmp = nav.c[i];
                        nav.c[i] = nav.c[j];
                        nav.c[j] = tmp;
                    }
                }
            }
            nav.next = static_cast<int8_t>(b);
            int inv = kCache;
            for (int i = 0; i < kCache; ++i) {
                if (nav.c[i].path == kMax) {
                    inv = i;
                    break;
                }
            }
            nav.last = static_cast<int8_t>(inv - 1);
            update_status();
            if (nav.cache_exhausted()) nav.set_no_trust();
        }
        if (nav.trust == kNoTrust) {
            local_navigation(cfg.nav, true);
            return is_init;
        }
        return !is_init;
    }

    // navigator_base.hpp:97-154
    bool nav_update() {
        if (nav.trust == kFull) return false;
        bool is_init = update_impl();
        if (nav.on_portal()) {
            volume_switch();
            is_init = true;
            if (!nav.is_alive()) return false;
        }
        if (nav.trust != kFull || nav.cache_exhausted()) {
            is_init = true;
            init_loose_cfg();
        }
        return is_init;
    }

    // ------------------------------------------------------------ stepper --
    // rk_stepper.hpp:569-590
    static void dtds_of(const V3& dir, const V3& b, double h, const V3& prev, bool first,
                        double qop, V3& t, V3& k) {
        t = (first || h == 0.0) ? dir : add(dir, mul(h, prev));
        k = mul(qop, cross(t, b));
    }

    // The current volume's homogeneous material, or null (propagator.hpp:
    // 320-325: vol.has_material() ? vol.material_parameters(pos) : nullptr)
    const dtr_material_slab* volume_material() const {
        if (nav.volume >= det.d->n_volumes) return nullptr;
        const uint32_t m = det.d->volumes[nav.volume].material;
        return m ? &det.d->materials[m - 1u] : nullptr;
    }

    // rk_stepper.hpp:157-189: d(qop)/ds = qop^3 E (-dE/ds) / q^2, zero in vacuum
    double dqopds(double qop, const dtr_material_slab* vm) const {
        if (!vm) return 0.0;
        const double q = st.ptc.charge;
        const double p = q / qop;
        const double mass = st.ptc.mass;
        const double E = std::sqrt(p * p + mass * mass);
        const double stopping_power = bethe_bloch(*vm, make_rq(mass, qop, q));
        return qop * qop * qop * E * stopping_power / (q * q);
    }
    // rk_stepper.hpp:194-231: d(dqop/ds)/dqop (Eq. 3.12 of JINST 4 P04016)
    double d2qopdsdqop(double qop, const dtr_material_slab* vm) const {
        if (!vm) return 0.0;
        const double q = st.ptc.charge;
        const double p = q / qop;
        const double p2 = p * p;
        const double mass = st.ptc.mass;
        const double E2 = p2 + mass * mass;
        const RQ r = make_rq(mass, qop, q);
        const double g = -1.0 * bethe_bloch(*vm, r);
        const double dgdqop = -1.0 * derive_bethe_bloch(*vm, r);
        const double dq = dqopds(qop, vm);
        return dq * (1.0 / qop * (3.0 - p2 / E2) + 1.0 / g * dgdqop);
    }
    // rk_stepper.hpp:147-155: the last stage's value after a step
    double state_dqopds(const dtr_material_slab* vm) const {
        if (st.path_length == 0.0) return dqopds(st.qop, vm);
        return st.dqopds3;
    }
    // rk_stepper.hpp:534-565: (dqop/ds, qop) of stage i
    void evaluate_dqopds(int i, double h, double dq_prev, const dtr_material_slab* vm, double& dq,
                         double& q) const {
        if (!vm) {
            q = st.qop;
            dq = 0.0;
        } else if (cfg.use_mean_loss && i != 0) {
            q = st.qop + h * dq_prev;
            dq = dqopds(q, vm);
        } else {
            q = st.qop;
            dq = dqopds(q, vm);
        }
    }

    // rk_stepper.hpp:273-530
    bool step(double dist_to_next, bool do_reset) {
        const PropCfg& c = cfg;
        if (std::fabs(dist_to_next) < static_cast<double>(1.f * 1e-3f)) {
            if (std::fabs(st.next_step_size) < c.min_stepsize)
                st.next_step_size = std::copysign(c.min_stepsize, dist_to_next);
            return true;
        }
        if (do_reset) {
            st.step_size = dist_to_next;
        } else if (st.next_step_size > 0) {
            st.step_size = std::min(st.next_step_size, dist_to_next);
        } else {
            st.step_size = std::max(st.next_step_size, dist_to_next);
        }
        const double min_step = std::min(std::fabs(st.step_size), c.min_stepsize);
        const V3 pos = st.pos;
        const dtr_material_slab* vm = volume_material();
        V3 b_first = field.at(pos), b_mid{{0, 0, 0}}, b_last{{0, 0, 0}};
        double qop[4], dq[4];
        V3 t[4], k[4];
        evaluate_dqopds(0, 0.0, 0.0, vm, dq[0], qop[0]);
        dtds_of(st.dir, b_first, 0.0, V3{{0, 0, 0}}, true, qop[0], t[0], k[0]);
        auto estimate_error = [&](double h) {
            const double h2 = h * h;
            const double half_h = h * 0.5;
            const V3 pos1 = add(add(pos, mul(half_h, t[0])), mul(h2 * 0.125, k[0]));
            b_mid = field.at(pos1);
            evaluate_dqopds(1, half_h, dq[0], vm, dq[1], qop[1]);
            dtds_of(st.dir, b_mid, half_h, k[0], false, qop[1], t[1], k[1]);
            evaluate_dqopds(2, half_h, dq[1], vm, dq[2], qop[2]);
            dtds_of(st.dir, b_mid, half_h, k[1], false, qop[2], t[2], k[2]);
            const V3 pos2 = add(add(pos, mul(h, t[0])), mul(h2 * 0.5, k[2]));
            b_last = field.at(pos2);
            evaluate_dqopds(3, h, dq[2], vm, dq[3], qop[3]);
            dtds_of(st.dir, b_last, h, k[2], false, qop[3], t[3], k[3]);
            const double one_sixth = 1.0 / 6.0;
            const V3 d = add(sub(sub(k[0], k[1]), k[2]), k[3]);
            const V3 err = mul(one_sixth * h2, d);
            return norm(err);
        };
        auto scaling = [&](double err) {
            return std::min(std::max(std::sqrt(std::sqrt(c.rk_error_tol / err)), 0.25), 4.0);
        };
        double error = 1e20;
        for (uint32_t i = 0; i < c.max_rk_updates; ++i) {
            st.n_trials++;
            error = std::max(estimate_error(st.step_size), 1e-20);
            if (error <= 4.0 * c.rk_error_tol) break;
            st.step_size = st.step_size * scaling(error);
        }
        st.dtds3 = k[3];
        st.dqopds3 = dq[3];
        const double max_step = st.constraint_size();
        if (std::fabs(st.step_size) > std::fabs(max_step)) st.step_size = max_step;
        if (std::fabs(st.step_size) < min_step) st.step_size = std::copysign(min_step, st.step_size);
        st.next_step_size = st.step_size * scaling(error);
        if (std::fabs(st.next_step_size) < c.min_stepsize)
            st.next_step_size = std::copysign(c.min_stepsize, st.next_step_size);
        // the energy loss would bring the particle to rest within this step
        // (its q/p diverges or changes sign): the reference integrates on into
        // non-finite parameters; here the step is not taken and the track
        // stops as at a surface that absorbs it
        if (vm) {
            const double h_6 = st.step_size * (1.0 / 6.0);
            const double qn = st.qop + h_6 * (dq[0] + 2.0 * (dq[1] + dq[2]) + dq[3]);
            if (!(std::isfinite(qn) && qn * st.qop > 0.0)) {
                stopped = true;
                return true;
            }
        }
        advance_track(t, k, dq, vm);
        if (c.do_cov) advance_jacobian(t, k, qop, b_first, b_mid, b_last, vm);
        st.n_steps++;
        return true;
    }

    // rk_stepper.hpp:594-629
    void advance_track(const V3 t[4], const V3 k[4], const double dq[4], const dtr_material_slab* vm) {
        const double h = st.step_size;
        const double h_6 = h * (1.0 / 6.0);
        st.pos = add(st.pos, mul(h, add(t[0], mul(h_6, add(add(k[0], k[1]), k[2])))));
        V3 d = add(st.dir, mul(h_6, add(add(k[0], mul(2.0, add(k[1], k[2]))), k[3])));
        st.dir = normalize(d);
        if (vm) st.qop = st.qop + h_6 * (dq[0] + 2.0 * (dq[1] + dq[2]) + dq[3]);
        st.path_length += h;
        st.abs_path_length += std::fabs(h);
    }

    // rk_stepper.hpp:958-998 evaluate_field_gradient: central differences of
    // the field over +-0.1 mm along each axis; G[r * 3 + i] = dB_r / dx_i
    void field_gradient(const V3& p, double G[9]) const {
        const double delta = static_cast<double>(1e-1f);
        for (int i = 0; i < 3; ++i) {
            V3 p1 = p, p2 = p;
            p1[i] += delta;
            p2[i] -= delta;
            const V3 b1 = field.at(p1), b2 = field.at(p2);
            for (int r = 0; r < 3; ++r) G[r * 3 + i] = (b1[r] - b2[r]) * (1.0 / (2.0 * delta));
        }
    }

    static void mat3_mul(const double A[9], const double B[9], double out[9]) {
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
                double s = 0.0;
                for (int m = 0; m < 3; ++m) s += A[r * 3 + m] * B[m * 3 + cc];
                out[r * 3 + cc] = s;
            }
    }

    static void cwc(const double M[9], const V3& b, double out[9]) {
        // column-wise cross product: column j of out = (column j of M) x b
        for (int j = 0; j < 3; ++j) {
            const V3 col{{M[0 * 3 + j], M[1 * 3 + j], M[2 * 3 + j]}};
            const V3 r = cross(col, b);
            out[0 * 3 + j] = r[0];
            out[1 * 3 + j] = r[1];
            out[2 * 3 + j] = r[2];
        }
    }

    // rk_stepper.hpp:633-956; J <- D(h) J with the ATL-SOFT-PUB-2009-002
    // blocks.  dqop_n/dqop1 and d(qop)/d(qop1) (rk_stepper.hpp:767-800) are
    // 1 unless use_eloss_gradient is set; the stage q/p values qop[n] carry
    // the volume material's energy loss either way.
    void advance_jacobian(const V3 t[4], const V3 k[4], const double qop[4], const V3& b1,
                          const V3& b2, const V3& b3, const dtr_material_slab* vm) {
        (void)k;
        const double h = st.step_size;
        const double half_h = h * 0.5;
        // rk_stepper.hpp:664 uses the single-precision literal 1.f / 6.f here
        const double h_6 = h * static_cast<double>(1.f / 6.f);
        double dqn[4] = {1.0, 1.0, 1.0, 1.0};
        double dqopqop = 1.0;
        if (cfg.use_eloss_gradient) {
            const double d1 = d2qopdsdqop(qop[0], vm);
            dqn[1] = 1.0 + half_h * d1;
            const double d2 = d2qopdsdqop(qop[1], vm) * dqn[1];
            dqn[2] = 1.0 + half_h * d2;
            const double d3 = d2qopdsdqop(qop[2], vm) * dqn[2];
            dqn[3] = 1.0 + h * d3;
            const double d4 = d2qopdsdqop(qop[3], vm) * dqn[3];
            dqopqop = 1.0 + h_6 * (d1 + 2.0 * (d2 + d3) + d4);
        }
        // dk_n / dqop1
        V3 dk[4];
        dk[0] = mul(dqn[0], cross(t[0], b1));
        dk[1] = add(mul(dqn[1], cross(t[1], b2)), mul(qop[1] * half_h, cross(dk[0], b2)));
        dk[2] = add(mul(dqn[2], cross(t[2], b2)), mul(qop[2] * half_h, cross(dk[1], b2)));
        dk[3] = add(mul(dqn[3], cross(t[3], b3)), mul(qop[3] * h, cross(dk[2], b3)));
        const V3 dFdqop = mul(h * h_6, add(add(dk[0], dk[1]), dk[2]));
        const V3 dGdqop = mul(h_6, add(add(dk[0], mul(2.0, add(dk[1], dk[2]))), dk[3]));
        // dk_n / dt1
        double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double dkdt[4][9];
        double tmp[9];
        cwc(I, b1, tmp);
        for (int i = 0; i < 9; ++i) dkdt[0][i] = qop[0] * tmp[i];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + half_h * dkdt[0][i];
        cwc(tmp, b2, dkdt[1]);
        for (int i = 0; i < 9; ++i) dkdt[1][i] *= qop[1];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + half_h * dkdt[1][i];
        cwc(tmp, b2, dkdt[2]);
        for (int i = 0; i < 9; ++i) dkdt[2][i] *= qop[2];
        for (int i = 0; i < 9; ++i) tmp[i] = I[i] + h * dkdt[2][i];
        cwc(tmp, b3, dkdt[3]);
        for (int i = 0; i < 9; ++i) dkdt[3][i] *= qop[3];
        double dFdt[9], dGdt[9];
        for (int i = 0; i < 9; ++i) {
            dFdt[i] = h * (I[i] + h_6 * (dkdt[0][i] + dkdt[1][i] + dkdt[2][i]));
            dGdt[i] = I[i] + h_6 * (dkdt[0][i] + 2.0 * (dkdt[1][i] + dkdt[2][i]) + dkdt[3][i]);
        }
        // dk_n / dr1 with the field gradient (rk_stepper.hpp:878-937); the
        // gradient points follow the reference, which evaluates them after
        // advance_track, i.e. around the end point of the step
        double dFdr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, dGdr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (cfg.use_field_gradient) {
            const double h2 = h * h;
            const V3 r_ini = st.pos;
            const V3 r_mid = add(add(r_ini, mul(half_h, t[0])), mul(h2 * 0.125, k[0]));
            const V3 r_fin = add(add(r_ini, mul(h, t[0])), mul(h2 * 0.5, k[2]));
            double G_ini[9], G_mid[9], G_fin[9];
            field_gradient(r_ini, G_ini);
            field_gradient(r_mid, G_mid);
            field_gradient(r_fin, G_fin);
            double dkdr[4][9], m1[9], m2[9], tmp0[9], tmp3[9];
            cwc(G_ini, t[0], m1);
            for (int i = 0; i < 9; ++i) dkdr[0][i] = -qop[0] * m1[i];
            for (int i = 0; i < 9; ++i) tmp0[i] = I[i] + h2 * 0.125 * dkdr[0][i];
            mat3_mul(G_mid, tmp0, m2);  // dB_mid/dr1
            for (int n = 1; n <= 2; ++n) {
                double hd[9];
                for (int i = 0; i < 9; ++i) hd[i] = half_h * dkdr[n - 1][i];
                cwc(hd, b2, m1);
                double m3[9];
                cwc(m2, t[n], m3);
                for (int i = 0; i < 9; ++i) dkdr[n][i] = qop[n] * m1[i] - qop[n] * m3[i];
            }
            {
                double hd[9];
                for (int i = 0; i < 9; ++i) hd[i] = h * dkdr[2][i];
                cwc(hd, b3, m1);
                for (int i = 0; i < 9; ++i) tmp3[i] = I[i] + h2 * 0.5 * dkdr[2][i];
                double gb[9], m3[9];
                mat3_mul(G_fin, tmp3, gb);
                cwc(gb, t[3], m3);
                for (int i = 0; i < 9; ++i) dkdr[3][i] = qop[3] * m1[i] - qop[3] * m3[i];
            }
            for (int i = 0; i < 9; ++i) {
                dFdr[i] = I[i] + h * h_6 * (dkdr[0][i] + dkdr[1][i] + dkdr[2][i]);
                dGdr[i] = h_6 * (dkdr[0][i] + 2.0 * (dkdr[1][i] + dkdr[2][i]) + dkdr[3][i]);
            }
        }
        // D (8x8): rows/cols [pos 0-2, t 3, dir 4-6, qop 7]
        double D[64] = {0};
        for (int i = 0; i < 8; ++i) D[i * 8 + i] = 1.0;
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
                D[r * 8 + cc] = dFdr[r * 3 + cc];
                D[(4 + r) * 8 + cc] = dGdr[r * 3 + cc];
            }
        for (int r = 0; r < 3; ++r) {
            for (int cc = 0; cc < 3; ++cc) {
                D[r * 8 + 4 + cc] = dFdt[r * 3 + cc];
                D[(4 + r) * 8 + 4 + cc] = dGdt[r * 3 + cc];
            }
            D[r * 8 + 7] = dFdqop[r];
            D[(4 + r) * 8 + 7] = dGdqop[r];
        }
        D[7 * 8 + 7] = dqopqop;
        double NJ[64];
        for (int r = 0; r < 8; ++r) {
            for (int cc = 0; cc < 8; ++cc) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += D[r * 8 + m] * st.J[m * 8 + cc];
                NJ[r * 8 + cc] = s;
            }
        }
        std::memcpy(st.J, NJ, sizeof(NJ));
    }

    // stepper_rk_policy (policies.hpp:99-139)
    void rk_policy() {
        if (std::fabs(nav()) <= static_cast<double>(1.f * 1e-3f)) {
            nav.set_high_trust();
            return;
        }
        const double rel = (st.step_size - nav()) / nav();
        if (rel > 0.1) nav.set_no_trust();
        else if (rel > 0.05) nav.set_fair_trust();
        else nav.set_high_trust();
    }

    // ------------------------------------------------- frames & transforms --
    // line2D.hpp:29-47 global_to_local_3D: signed distance of closest
    // approach (negative on the side of cross(z, dir)) and z along the wire
    static void line_local(const dtr_transform& t, const V3& pos, const V3& dir, double& l0, double& l1) {
        const V3 local3 = point_to_local(t, pos);
        const V3 z = trf_axis(t, 2);
        const V3 r = cross(z, dir);
        const double sign = dot(r, sub(trf_translation(t), pos)) > 0.0 ? -1.0 : 1.0;
        l0 = sign * perp(local3);
        l1 = local3[2];
    }
    // jacobian_line.hpp:38-62 reference_frame: y along the wire, x = y x dir
    static void line_frame(const dtr_transform& t, const V3& dir, V3& nx, V3& ny, V3& nz) {
        ny = trf_axis(t, 2);
        nx = normalize(cross(ny, dir));
        nz = cross(nx, ny);
    }

    // free_to_bound_vector (transform_track_parameters.hpp:24-38)
    void free_to_bound(uint32_t s, const V3& pos, const V3& dir, double time, double qop,
                       double out[6]) const {
        const dtr_surface& sf = det.sf(s);
        if (is_line(sf.shape)) {
            line_local(det.trf(s), pos, dir, out[0], out[1]);
        } else if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            out[0] = phi(pos);
            out[1] = pos[2];
        } else if (sf.shape == DTR_CYLINDER2D) {
            // cylindrical2D.hpp:33-41: r phi and z in the local frame
            const V3 l = point_to_local(det.trf(s), pos);
            out[0] = perp(l) * phi(l);
            out[1] = l[2];
        } else if (sf.shape == DTR_RING2D || sf.shape == DTR_ANNULUS2D) {
            const V3 l = point_to_local(det.trf(s), pos);
            out[0] = perp(l);
            out[1] = phi(l);
        } else {
            const V3 l = point_to_local(det.trf(s), pos);
            out[0] = l[0];
            out[1] = l[1];
        }
        out[2] = phi(dir);
        out[3] = theta(dir);
        out[4] = qop;
        out[5] = time;
    }
    static V3 bound_dir(const double b[6]) {
        const double s = std::sin(b[3]);
        return {{std::cos(b[2]) * s, std::sin(b[2]) * s, std::cos(b[3])}};
    }
    // bound_to_free_vector (transform_track_parameters.hpp:40-73)
    void bound_to_free(uint32_t s, const double b[6], V3& pos, V3& dir, double& time,
                       double& qop) const {
        const dtr_surface& sf = det.sf(s);
        if (is_line(sf.shape)) {
            // line2D.hpp:80-99 local_to_global with the track direction
            const dtr_transform& t = det.trf(s);
            const V3 d = bound_dir(b);
            const V3 r = cross(trf_axis(t, 2), d);
            pos = add(point_to_global(t, V3{{0.0, 0.0, b[1]}}), mul(b[0], normalize(r)));
        } else if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double r = det.mask(sf.mask).v[0];
            pos = {{r * std::cos(b[0]), r * std::sin(b[0]), b[1]}};
        } else if (sf.shape == DTR_CYLINDER2D) {
            // cylindrical2D.hpp:43-60 at the mask radius
            const double r = det.mask(sf.mask).v[0];
            const double ph = b[0] / r;
            pos = point_to_global(det.trf(s), V3{{r * std::cos(ph), r * std::sin(ph), b[1]}});
        } else if (sf.shape == DTR_RING2D || sf.shape == DTR_ANNULUS2D) {
            pos = point_to_global(det.trf(s), V3{{b[0] * std::cos(b[1]), b[0] * std::sin(b[1]), 0.0}});
        } else {
            pos = point_to_global(det.trf(s), V3{{b[0], b[1], 0.0}});
        }
        dir = bound_dir(b);
        time = b[5];
        qop = b[4];
    }
    // normal used for the incidence angle: the wire direction for lines
    // (line2D.hpp:101-112)
    V3 surface_normal(uint32_t s, const double b[6]) const {
        const dtr_surface& sf = det.sf(s);
        if (sf.shape == DTR_CONCENTRIC_CYLINDER2D)
            return {{std::cos(b[0]), std::sin(b[0]), 0.0}};
        if (sf.shape == DTR_CYLINDER2D) {
            // cylindrical2D.hpp:62-70: phi = loc0 / r of the mask
            const double ph = b[0] / det.mask(sf.mask).v[0];
            return vector_to_global(det.trf(s), V3{{std::cos(ph), std::sin(ph), 0.0}});
        }
        return trf_axis(det.trf(s), 2);
    }
    // jacobian_cylindrical.hpp:24-45: y = cylinder axis, z = the normal at
    // the local point (phi from r phi / r), x = y x z
    void cyl2_frame(uint32_t s, const V3& pos, V3& nx, V3& ny, V3& nz) const {
        const dtr_transform& t = det.trf(s);
        const V3 l = point_to_local(t, pos);
        const double r = perp(l);
        const double ph = (r * phi(l)) / r;
        ny = trf_axis(t, 2);
        nz = vector_to_global(t, V3{{std::cos(ph), std::sin(ph), 0.0}});
        nx = cross(ny, nz);
    }
    // dpos/dloc (3x2) jacobian_cartesian.hpp:69-79 / concentric_cylindrical /
    // jacobian_line.hpp:94-100 (the x and y axes of the line's frame)
    void dpos_dloc(uint32_t s, const V3& pos, const V3& dir, double A[6]) const {
        const dtr_surface& sf = det.sf(s);
        if (is_line(sf.shape)) {
            V3 nx, ny, nz;
            line_frame(det.trf(s), dir, nx, ny, nz);
            for (int r = 0; r < 3; ++r) {
                A[r * 2 + 0] = nx[r];
                A[r * 2 + 1] = ny[r];
            }
        } else if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double r = perp(pos), ph = phi(pos);
            A[0] = -r * std::sin(ph); A[1] = 0.0;
            A[2] = r * std::cos(ph);  A[3] = 0.0;
            A[4] = 0.0;               A[5] = 1.0;
        } else if (sf.shape == DTR_CYLINDER2D) {
            V3 nx, ny, nz;
            cyl2_frame(s, pos, nx, ny, nz);
            for (int r = 0; r < 3; ++r) {
                A[r * 2 + 0] = nx[r];
                A[r * 2 + 1] = ny[r];
            }
        } else if (sf.shape == DTR_RING2D || sf.shape == DTR_ANNULUS2D) {
            // jacobian_polar.hpp:44-72
            const dtr_transform& t = det.trf(s);
            const V3 l = point_to_local(t, pos);
            const double lr = perp(l), lp = phi(l), c = std::cos(lp), sn = std::sin(lp);
            const V3 x = trf_axis(t, 0), y = trf_axis(t, 1);
            const V3 col0 = add(mul(c, x), mul(sn, y));
            const V3 col1 = mul(lr, sub(mul(c, y), mul(sn, x)));
            for (int r = 0; r < 3; ++r) {
                A[r * 2 + 0] = col0[r];
                A[r * 2 + 1] = col1[r];
            }
        } else {
            const dtr_transform& t = det.trf(s);
            A[0] = t.r[0]; A[1] = t.r[1];
            A[2] = t.r[3]; A[3] = t.r[4];
            A[4] = t.r[6]; A[5] = t.r[7];
        }
    }
    // dpos/dangle (3x2) of a line (jacobian_line.hpp:102-172); zero for the
    // other frames
    void line_dpos_dangle(uint32_t s, const V3& pos, const V3& dir, const double b[6], double A[6]) const {
        const dtr_transform& t = det.trf(s);
        double l0, l1;
        line_local(t, pos, dir, l0, l1);
        V3 nx, ny, nz;
        line_frame(t, dir, nx, ny, nz);
        const double ipdn = 1.0 / dot(dir, nz);
        const double ct = std::cos(b[3]), stt = std::sin(b[3]), cp = std::cos(b[2]), sp = std::sin(b[2]);
        const V3 dn_dphi{{-stt * sp, stt * cp, 0.0}};
        const V3 dn_dtheta{{ct * cp, ct * sp, -stt}};
        const V3 yp = cross(ny, dn_dphi);
        const V3 yt = cross(ny, dn_dtheta);
        const double C = ipdn * l0;
        const V3 dphi = mul(C, sub(yp, mul(dot(nx, yp), nx)));
        const V3 dtheta = mul(C, sub(yt, mul(dot(nx, yt), nx)));
        for (int r = 0; r < 3; ++r) {
            A[r * 2 + 0] = dphi[r];
            A[r * 2 + 1] = dtheta[r];
        }
    }
    // dloc/dpos (2x3)
    void dloc_dpos(uint32_t s, const V3& pos, const V3& dir, double A[6]) const {
        const dtr_surface& sf = det.sf(s);
        if (is_line(sf.shape)) {
            // jacobian_line.hpp:174-182: the first two rows of the frame's transpose
            V3 nx, ny, nz;
            line_frame(det.trf(s), dir, nx, ny, nz);
            for (int c = 0; c < 3; ++c) {
                A[0 * 3 + c] = nx[c];
                A[1 * 3 + c] = ny[c];
            }
        } else if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double rinv = 1.0 / perp(pos), ph = phi(pos);
            A[0] = -rinv * std::sin(ph); A[1] = rinv * std::cos(ph); A[2] = 0.0;
            A[3] = 0.0;                  A[4] = 0.0;                 A[5] = 1.0;
        } else if (sf.shape == DTR_CYLINDER2D) {
            V3 nx, ny, nz;
            cyl2_frame(s, pos, nx, ny, nz);
            for (int c = 0; c < 3; ++c) {
                A[0 * 3 + c] = nx[c];
                A[1 * 3 + c] = ny[c];
            }
        } else if (sf.shape == DTR_RING2D || sf.shape == DTR_ANNULUS2D) {
            // jacobian_polar.hpp:81-110
            const dtr_transform& t = det.trf(s);
            const V3 l = point_to_local(t, pos);
            const double lr = perp(l), lp = phi(l), c = std::cos(lp), sn = std::sin(lp);
            const V3 x = trf_axis(t, 0), y = trf_axis(t, 1);
            const V3 row0 = add(mul(c, x), mul(sn, y));
            const V3 row1 = mul(1.0 / lr, sub(mul(c, y), mul(sn, x)));
            for (int k = 0; k < 3; ++k) {
                A[0 * 3 + k] = row0[k];
                A[1 * 3 + k] = row1[k];
            }
        } else {
            const dtr_transform& t = det.trf(s);
            A[0] = t.r[0]; A[1] = t.r[3]; A[2] = t.r[6];
            A[3] = t.r[1]; A[4] = t.r[4]; A[5] = t.r[7];
        }
    }
    // free_to_path derivative (jacobian_cartesian.hpp:54-66,
    // jacobian_concentric_cylindrical.hpp:45-67: position part only;
    // jacobian_line.hpp:64-92: position and direction parts, with dt/ds)
    void path_derivative(uint32_t s, const V3& pos, const V3& dir, const V3& dtds_v, double f2p[8]) const {
        const dtr_surface& sf = det.sf(s);
        for (int i = 0; i < 8; ++i) f2p[i] = 0.0;
        if (is_line(sf.shape)) {
            const dtr_transform& t = det.trf(s);
            const V3 pc = sub(pos, trf_translation(t));
            const V3 lz = trf_axis(t, 2);
            const double pz = dot(pc, lz);
            const double dz = dot(lz, dir);
            const V3 pc_x = sub(pc, mul(pz, lz));
            const double nrm = -1.0 / (1.0 - dz * dz + dot(pc_x, dtds_v));
            const V3 pos_term = mul(nrm, sub(dir, mul(dz, lz)));
            const V3 dir_term = mul(nrm, pc_x);
            for (int i = 0; i < 3; ++i) {
                f2p[i] = pos_term[i];
                f2p[4 + i] = dir_term[i];
            }
            return;
        }
        V3 n;
        if (sf.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double ph = phi(pos);
            n = {{std::cos(ph), std::sin(ph), 0.0}};
        } else if (sf.shape == DTR_CYLINDER2D) {
            // jacobian_cylindrical.hpp:47-62
            V3 nx, ny;
            cyl2_frame(s, pos, nx, ny, n);
        } else {
            n = trf_axis(det.trf(s), 2);
        }
        const V3 d = mul(-1.0 / dot(n, dir), n);
        for (int i = 0; i < 3; ++i) f2p[i] = d[i];
    }

    // stepping.dtds() (rk_stepper.hpp:121-140)
    V3 dtds() const {
        if (st.path_length == 0.0) return mul(st.qop, cross(st.dir, field.at(st.pos)));
        return st.dtds3;
    }

    // parameter_transporter::get_full_jacobian (parameter_updater.hpp:422-572)
    void full_jacobian(uint32_t dest, double Jf[36]) const {
        const uint32_t dep = up.sf;
        V3 dpos, ddir;
        double dt, dq;
        bound_to_free(dep, up.bound, dpos, ddir, dt, dq);
        // B2F 8x6
        double B2F[48] = {0};
        double A[6];
        dpos_dloc(dep, dpos, ddir, A);
        for (int r = 0; r < 3; ++r) {
            B2F[r * 6 + 0] = A[r * 2 + 0];
            B2F[r * 6 + 1] = A[r * 2 + 1];
        }
        if (is_line(det.sf(dep).shape)) {
            line_dpos_dangle(dep, dpos, ddir, up.bound, A);
            for (int r = 0; r < 3; ++r) {
                B2F[r * 6 + 2] = A[r * 2 + 0];
                B2F[r * 6 + 3] = A[r * 2 + 1];
            }
        }
        {
            const double th = up.bound[3], ph = up.bound[2];
            const double ct = std::cos(th), stt = std::sin(th), cp = std::cos(ph), sp = std::sin(ph);
            B2F[4 * 6 + 2] = -stt * sp;
            B2F[4 * 6 + 3] = ct * cp;
            B2F[5 * 6 + 2] = stt * cp;
            B2F[5 * 6 + 3] = ct * sp;
            B2F[6 * 6 + 3] = -stt;
        }
        B2F[3 * 6 + 5] = 1.0;
        B2F[7 * 6 + 4] = 1.0;
        // path correction
        const V3 k = dtds();
        // path_to_free_derivative with the volume's d(qop)/ds
        // (parameter_updater.hpp:518-526, jacobian_engine.hpp:260-273)
        double p2f[8] = {st.dir[0], st.dir[1], st.dir[2], 0.0, k[0], k[1], k[2],
                         state_dqopds(volume_material())};
        double f2p[8];
        path_derivative(dest, st.pos, st.dir, k, f2p);
        double C[64];
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 8; ++c) C[r * 8 + c] = (r == c ? 1.0 : 0.0) + p2f[r] * f2p[c];
        // F2B 6x8
        double F2B[48] = {0};
        double Bl[6];
        dloc_dpos(dest, st.pos, st.dir, Bl);
        for (int c = 0; c < 3; ++c) {
            F2B[0 * 8 + c] = Bl[0 * 3 + c];
            F2B[1 * 8 + c] = Bl[1 * 3 + c];
        }
        {
            const double th = theta(st.dir), ph = phi(st.dir);
            const double ct = std::cos(th), stt = std::sin(th), cp = std::cos(ph), sp = std::sin(ph);
            F2B[2 * 8 + 4] = -sp / stt;
            F2B[2 * 8 + 5] = cp / stt;
            F2B[3 * 8 + 4] = cp * ct;
            F2B[3 * 8 + 5] = sp * ct;
            F2B[3 * 8 + 6] = -stt;
        }
        F2B[4 * 8 + 7] = 1.0;
        F2B[5 * 8 + 3] = 1.0;
        // J_full = F2B * C * J * B2F
        double T1[64], T2[48], T3[36];
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 8; ++c) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += C[r * 8 + m] * st.J[m * 8 + c];
                T1[r * 8 + c] = s;
            }
        for (int r = 0; r < 8; ++r)
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += T1[r * 8 + m] * B2F[m * 6 + c];
                T2[r * 6 + c] = s;
            }
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += F2B[r * 8 + m] * T2[m * 6 + c];
                T3[r * 6 + c] = s;
            }
        std::memcpy(Jf, T3, sizeof(T3));
    }

    // --------------------------------------------------------------- actors --
    bool on_sensitive() const {
        return nav.status == kOnObject && det.sf(nav.current().sf).sf_id == DTR_SENSITIVE;
    }
    bool encountered_material() const {
        return nav.on_surface() && det.sf(nav.current().sf).material != kInvalid;
    }

    // aborters.hpp:44-75
    void pathlimit_aborter() {
        if (nav.status == kExit) return;
        const double step_limit = cfg.path_limit - std::fabs(st.abs_path_length);
        if (step_limit <= 0.0) {
            nav.abort_nav();
            nav.exited_path_limit = true;
            heartbeat = false;
        }
        st.set_constraint(2, step_limit);
    }

    // relativistic_quantities.hpp + interaction.hpp
    struct RQ {
        double qOverP, q2OverBeta2, beta2, betaGamma, gamma, gamma2, E, Wmax;
    };
    static RQ make_rq(double mass, double qop, double q) {
        RQ r;
        r.qOverP = qop;
        r.q2OverBeta2 = q * q + (mass * qop) * (mass * qop);
        const double mOverP = mass * ((q != 0.0) ? std::fabs(qop / q) : std::fabs(qop));
        const double pOverM = 1.0 / mOverP;
        r.beta2 = 1.0 / (1.0 + mOverP * mOverP);
        r.betaGamma = pOverM;
        r.gamma = std::sqrt(1.0 + pOverM * pOverM);
        r.gamma2 = r.gamma * r.gamma;
        r.E = r.gamma * mass;
        const double m_e = 0.51099895069 * 1e-3;
        const double mfrac = m_e / mass;
        r.Wmax = (2.0 * m_e * r.betaGamma * r.betaGamma) / (1.0 + 2.0 * r.gamma * mfrac + mfrac * mfrac);
        return r;
    }
    // relativistic_quantities.hpp compute_delta_half: the approximate
    // correction above beta gamma = 10 without density-effect data, the
    // Sternheimer parameterisation with it
    static double delta_half(const dtr_material_slab& m, const RQ& r) {
        if (!m.has_density) {
            if (r.betaGamma < 10.0) return 0.0;
            const double plasma = (28.816 * 1e-9) * std::sqrt(1000.0 * (m.Z * m.molar_rho));
            return std::log(r.betaGamma * plasma / m.mean_exc) - 0.5;
        }
        const double a = m.density[0], k = m.density[1], x0 = m.density[2], x1 = m.density[3],
                     cbar = m.density[4];
        const double x = std::log10(r.betaGamma);
        double delta = 0.0;
        if (x >= x0) {
            delta = 2.0 * M_LN10 * x - cbar;
            if (x < x1) delta += a * std::pow(x1 - x, k);
        }
        return 0.5 * delta;
    }
    // relativistic_quantities.hpp derive_delta_half: d(delta/2)/dqop
    // (-1/qop without density-effect data at any beta gamma, as there)
    static double derive_delta_half(const dtr_material_slab& m, const RQ& r) {
        if (!m.has_density) return -1.0 / r.qOverP;
        const double a = m.density[0], k = m.density[1], x0 = m.density[2], x1 = m.density[3];
        const double x = std::log10(r.betaGamma);
        double delta = 0.0;
        if (x >= x0) {
            delta = -2.0 / r.qOverP;
            if (x < x1) delta += a * k / (r.qOverP * M_LN10) * std::pow(x1 - x, k - 1.0);
        }
        return 0.5 * delta;
    }
    // interaction.hpp:110-154 derive_bethe_bloch: d(-dE/dx)/dqop
    static double derive_bethe_bloch(const dtr_material_slab& m, const RQ& r) {
        const double K = 0.307075 * 1e-3 * 1e2;
        const double eps_per_length = 0.5 * K * (m.Z * m.molar_rho) * r.q2OverBeta2;
        if (eps_per_length <= 0.0) return 0.0;
        const double first = 2.0 / (r.qOverP * r.gamma2) * bethe_bloch(m, r);
        const double dA = -1.0 / (2.0 * r.qOverP) * (4.0 - r.Wmax / r.E);
        const double dB = -2.0 * r.beta2 / (r.qOverP * r.gamma2);
        const double dC = derive_delta_half(m, r);
        return first + 2.0 * eps_per_length * (dA - dB - dC);
    }
    static double bethe_bloch(const dtr_material_slab& m, const RQ& r) {
        const double K = 0.307075 * 1e-3 * 1e2;  // MeV cm^2 in GeV mm^2
        const double eps_per_length = 0.5 * K * (m.Z * m.molar_rho) * r.q2OverBeta2;
        if (eps_per_length <= 0.0) return 0.0;
        const double dhalf = delta_half(m, r);
        const double m_e = 0.51099895069 * 1e-3;
        const double u = 2.0 * m_e * r.betaGamma * r.betaGamma;
        const double A = 0.5 * std::log(u * r.Wmax / (m.mean_exc * m.mean_exc));
        const double running = A - r.beta2 - dhalf;
        return 2.0 * eps_per_length * running;
    }
    static double landau_sigma_qop(const dtr_material_slab& m, const RQ& r, double path,
                                   double charge) {
        const double K = 0.307075 * 1e-3 * 1e2;
        const double eps = 0.5 * K * (m.Z * m.molar_rho) * r.q2OverBeta2 * path;
        const double fwhm = 4.0 * eps;
        const double sigmaE = 0.5 * M_SQRT1_2 * fwhm / std::sqrt(M_LN2);
        const double pInv = r.qOverP / charge;
        return std::sqrt(r.q2OverBeta2) * pInv * pInv * sigmaE;
    }
    static double highland(double xOverX0, double qop, double charge, double q2OverBeta2) {
        const double momentumInv = std::fabs(qop / charge);
        if (xOverX0 <= 0.0) return 0.0;
        const double t = std::sqrt(xOverX0 * q2OverBeta2);
        return static_cast<double>(13.6f) * 1e-3 * momentumInv * t *
               (1.0 + static_cast<double>(0.038f) * 2.0 * std::log(t));
    }

    // parameter_updater (parameter_updater.hpp:314-627) + interactor
    // (pointwise_material_interactor.hpp:147-326)
    void parameter_updater() {
        if (!(on_sensitive() || encountered_material())) return;
        const uint32_t dest = nav.current().sf;
        const double fp[8] = {st.pos[0], st.pos[1], st.pos[2], st.time, st.dir[0], st.dir[1], st.dir[2], st.qop};
        if (up.sf != kInvalid) {
            if (up.sf == dest) {
                // on the initial surface: notify observers without transport
                record_hit(dest);
                material_and_setter(dest, fp);
                return;
            }
            double Jf[36];
            full_jacobian(dest, Jf);
            if (std::fabs(st.path_length) > 0.0) st.reset_J();
            double T[36], NC[36];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int m = 0; m < 6; ++m) s += Jf[r * 6 + m] * up.cov[m * 6 + c];
                    T[r * 6 + c] = s;
                }
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int m = 0; m < 6; ++m) s += T[r * 6 + m] * Jf[c * 6 + m];
                    NC[r * 6 + c] = s;
                }
            std::memcpy(up.cov, NC, sizeof(NC));
        } else {
            // placeholder covariance at the first surface
            if (init_cov) {
                int k = 0;
                for (int r = 0; r < 6; ++r)
                    for (int c = r; c < 6; ++c, ++k) {
                        up.cov[r * 6 + c] = init_cov[k];
                        up.cov[c * 6 + r] = init_cov[k];
                    }
            } else {
                for (int i = 0; i < 36; ++i) up.cov[i] = (i % 7 == 0) ? 1.0 : 0.0;
            }
        }
        free_to_bound(dest, st.pos, st.dir, st.time, st.qop, up.bound);
        up.sf = dest;
        record_hit(dest);
        material_and_setter(dest, fp);
    }

    // the bound parameters a measurement on this surface would be compared
    // with: those of the updater before the material interaction
    void record_hit(uint32_t dest) {
        if (!hit_sf || !on_sensitive() || n_hit >= hit_cap) return;
        hit_sf[n_hit * hit_stride] = dest;
        hit_loc[n_hit * hit_stride] = up.bound[0];
        hit_loc[(static_cast<uint64_t>(hit_cap) + n_hit) * hit_stride] = up.bound[1];
        n_hit++;
    }

    // material_accessor::get (material/detail/material_accessor.hpp:30-48):
    // a homogeneous slab, or the slab of the material map's bin at the bound
    // local position, grid::bin of a regular 2D grid whose axes clamp
    // (closed) or wrap (circular): axis_binning.hpp:80-82 (bin = int((v -
    // min) / width + 1) - 1), axis_bounds.hpp:107-206
    const dtr_material_slab& surface_slab(uint32_t link, double l0, double l1) const {
        if (!(link & DTR_MATERIAL_MAP)) return det.d->materials[link];
        const dtr_material_map& mp = det.d->material_maps[link & ~DTR_MATERIAL_MAP];
        auto bin = [](double v, double lo, double hi, uint32_t n, bool circular) {
            const double width = (hi - lo) / static_cast<double>(n);
            int b = static_cast<int>((v - lo) / width + 1.0) - 1;
            const int nb = static_cast<int>(n);
            if (circular) {
                b %= nb;
                return b < 0 ? b + nb : b;
            }
            return std::min(std::max(b, 0), nb - 1);
        };
        const int i0 = bin(l0, mp.min0, mp.max0, mp.n0, mp.circular0 != 0u);
        const int i1 = bin(l1, mp.min1, mp.max1, mp.n1, mp.circular1 != 0u);
        return det.d->materials[mp.slabs + static_cast<uint32_t>(i0) * mp.n1 + static_cast<uint32_t>(i1)];
    }

    void material_and_setter(uint32_t dest, const double fp[8]) {
        bool success = false;
        if (encountered_material()) {
            const dtr_material_slab& m = surface_slab(det.sf(dest).material, up.bound[0], up.bound[1]);
            const V3 n = surface_normal(dest, up.bound);
            const double cos_inc = std::fabs(dot(bound_dir(up.bound), n));
            double e_loss = 0.0, sigma_qop = 0.0, theta0 = 0.0;
            if (m.thickness > std::numeric_limits<double>::epsilon() && up.bound[4] != 0.0) {
                const double qop = up.bound[4];
                double path_seg, path_x0;
                if (m.kind == DTR_MATERIAL_ROD) {
                    // material_rod.hpp:62-78, approach = bound loc0
                    const double a = up.bound[0];
                    path_seg = 0.0;
                    if (!(std::fabs(a) > m.thickness)) {
                        const double sin2 = 1.0 - cos_inc * cos_inc;
                        path_seg = 2.0 * std::sqrt((m.thickness * m.thickness - a * a) / sin2);
                    }
                    path_x0 = path_seg / m.X0;
                } else {
                    path_seg = m.thickness / cos_inc;
                    path_x0 = m.thickness_in_X0 / cos_inc;
                }
                const RQ r = make_rq(st.ptc.mass, qop, st.ptc.charge);
                if (cfg.do_eloss) e_loss = path_seg * bethe_bloch(m, r);
                if (cfg.do_eloss && cfg.do_cov) sigma_qop = landau_sigma_qop(m, r, path_seg, st.ptc.charge);
                if (cfg.do_ms) theta0 = highland(path_x0, qop, st.ptc.charge, r.q2OverBeta2);
                success = true;
            }
            if (success) {
                if (cfg.do_eloss) {
                    // update_qop (pointwise_material_interactor.hpp:240-262)
                    const double mass = st.ptc.mass, q = st.ptc.charge;
                    const double p = q / up.bound[4];
                    const double next_e =
                        std::sqrt(mass * mass + p * p) - std::copysign(e_loss, static_cast<double>(nav.direction));
                    const double next_p = (mass < next_e) ? std::sqrt(next_e * next_e - mass * mass) : 0.0;
                    const double next_qop = (q != 0.0) ? q / next_p : 1.0 / next_p;
                    up.bound[4] = (next_p == 0.0) ? kMax : next_qop;
                    // particle at rest: nothing left to propagate
                    if (next_p == 0.0) stopped = true;
                    if (cfg.do_cov) up.cov[4 * 6 + 4] += sigma_qop * sigma_qop;
                }
                if (cfg.do_cov) {
                    const V3 d = bound_dir(up.bound);
                    const double var = theta0 * theta0;
                    up.cov[2 * 6 + 2] += (d[2] == 1.0) ? kMax : var / (1.0 - d[2] * d[2]);
                    up.cov[3 * 6 + 3] += var;
                }
            }
        }
        if (cfg.actors & DTR_ACTOR_STEP_TRACER) trace_update(dest, fp);
        if (!success) return;
        // parameter_setter (parameter_updater.hpp:585-627)
        bound_to_free(up.sf, up.bound, st.pos, st.dir, st.time, st.qop);
        if (cfg.estimate_noise) estimate_noise();
    }

    // noise_estimation.hpp:23-107
    void estimate_noise() {
        const double ns = static_cast<double>(cfg.n_stddev);
        const double v0 = up.cov[0], v1 = up.cov[7];
        const double dphi = ns * std::sqrt(up.cov[2 * 6 + 2]);
        const double dth = ns * std::sqrt(up.cov[3 * 6 + 3]);
        const double pe = up.bound[2] + dphi, te = up.bound[3] + dth;
        const double ste = std::sin(te);
        V3 displ{{std::cos(pe) * ste, std::sin(pe) * ste, std::cos(te)}};
        const double max_tol = static_cast<double>(5.f);
        const double path = nav.cache_exhausted() ? max_tol : std::fabs(nav.target().path);
        displ = mul(path, sub(displ, st.dir));
        const double q = st.ptc.charge;
        const double p = q / st.qop;
        const double acc = 1.0 / p * cfg.accumulated_error * st.path_length;
        const double ext = std::sqrt(ns * ns * (v0 + v1) + dot(displ, displ) + acc * acc);
        nav.ext_tol = ext > max_tol ? max_tol : ext;
    }

    // step_tracer (tests/include/detray/test/utils/step_tracer.hpp:
    // collect_data as observer of the parameter transporter, after the
    // material interactor and before the setter): fp is the stepping state
    // on the surface
    void trace_update(uint32_t dest, const double fp[8]) {
        const uint32_t k = n_trace++;
        if (k >= steps.capacity) return;
        const uint64_t o = k * st_stride, M = static_cast<uint64_t>(steps.capacity) * st_stride;
        if (steps.surface) steps.surface[o] = dest;
        if (steps.step_size) steps.step_size[o] = st.step_size;
        if (steps.path_length) steps.path_length[o] = st.path_length;
        if (steps.n_trials) steps.n_trials[o] = static_cast<uint32_t>(st.n_trials);
        if (steps.free_params)
            for (int v = 0; v < 8; ++v) steps.free_params[v * M + o] = fp[v];
        if (steps.bound_params)
            for (int v = 0; v < 6; ++v) steps.bound_params[v * M + o] = up.bound[v];
        if (steps.jacobian)
            for (int v = 0; v < 64; ++v) steps.jacobian[v * M + o] = st.J[v];
    }

    // momentum_aborter (aborters.hpp:80-143)
    void momentum_aborter() {
        if (nav.status == kExit) return;
        const double q = st.ptc.charge;
        const double pT = std::fabs(q / st.qop * perp(st.dir));
        if (pT <= cfg.min_pT || q / st.qop <= cfg.min_p) {
            if (nav.status != kAbort) momentum_hit = true;
            nav.abort_nav();
            heartbeat = false;
        }
    }

    // target_aborter (aborters.hpp:146-175)
    void target_aborter() {
        if (nav.on_surface() && nav.current().sf == target && st.path_length > 0.0) {
            if (nav.status != kAbort) target_hit = true;
            nav.abort_nav();
            heartbeat = false;
        }
    }

    // surface_sequencer.hpp:78-106
    void sequencer() {
        if (cfg.max_surfaces == 0u) return;
        if (!(on_sensitive() || encountered_material())) return;
        if (n_seq == cfg.max_surfaces) {
            overflow = true;
            nav.exit_nav();
            heartbeat = false;
            return;
        }
        seq[static_cast<uint64_t>(n_seq) * seq_stride] = nav.current().sf;
        n_seq++;
    }

    // the chain pathlimit_aborter, parameter_updater<pointwise_material_
    // interactor, step_tracer?>, surface_sequencer, momentum_aborter?,
    // target_aborter?
    void run_actors() {
        pathlimit_aborter();
        if (cfg.use_material) parameter_updater();
        if (stopped && nav.is_alive()) {
            nav.abort_nav();
            heartbeat = false;
        }
        sequencer();
        if (cfg.actors & DTR_ACTOR_MOMENTUM_ABORTER) momentum_aborter();
        if (cfg.actors & DTR_ACTOR_TARGET_ABORTER) target_aborter();
    }

    void trace() {
        if (!tr_vol || tr_n >= tr_cap) return;
        tr_vol[tr_n] = nav.volume;
        tr_sf[tr_n] = nav.on_surface() ? nav.current().sf : kInvalid;
        tr_n++;
    }

    // propagator.hpp:262-400; with a surface sequence the direct navigator
    // (direct_navigator.hpp:198-206 init: clear the cache, keep the target)
    bool propagate() {
        if (dseq) {
            nav.clear_cache();
            nav.next = 1;
            nav.last = 1;
            if (dseq_len > 0) nav.c[1].sf = dseq_at(0);
            direct_update();
        } else {
            local_navigation(cfg.nav, false);  // navigator_base::init
        }
        heartbeat = nav.is_alive();
        trace();
        bool is_init = true;
        double path_length = 0.0;
        unsigned stall = 0;
        for (unsigned i = 0; i % 2 == 0 || heartbeat; ++i) {
            if (i % 2 == 0) {
                run_actors();
                if (!heartbeat) continue;
                path_length = st.path_length;
            } else {
                const bool reset = nav.on_surface() || is_init;
#ifdef ORACLE_DEBUG
                std::fprintf(stderr, "i=%u vol=%u status=%d trust=%u next=%d last=%d dist=%.6g pos=(%.6g %.6g %.6g) r=%.6g dir=(%.4g %.4g %.4g) path=%.6g\n",
                             i, nav.volume, nav.status, nav.trust, nav.next, nav.last, nav(), st.pos[0], st.pos[1], st.pos[2],
                             perp(st.pos), st.dir[0], st.dir[1], st.dir[2], st.path_length);
#endif
                heartbeat = heartbeat && step(nav(), reset);
                rk_policy();
                if (i > 0) is_init = false;
                if (std::fabs(st.path_length) <= std::fabs(path_length) + static_cast<double>(cfg.nav.path_tol)) {
                    if (stall >= 10u) {
                        heartbeat = false;
                        nav.abort_nav();
                    }
                    stall++;
                } else {
                    stall = 0;
                }
            }
            is_init |= dseq ? direct_update() : nav_update();
            heartbeat = heartbeat && nav.is_alive();
            trace();
        }
        return nav.status == kExit;
    }
};

}  // namespace oracle

extern "C" {

static oracle::PropCfg make_prop_cfg(const dtr_prop_config* cfg) {
    oracle::PropCfg pc;
    pc.min_stepsize = cfg->min_stepsize;
    pc.rk_error_tol = cfg->rk_error_tol;
    pc.step_constraint = cfg->step_constraint;
    pc.path_limit = cfg->path_limit;
    pc.max_rk_updates = cfg->max_rk_updates;
    pc.do_cov = cfg->do_covariance_transport != 0;
    pc.use_field_gradient = cfg->use_field_gradient != 0;
    pc.use_mean_loss = cfg->use_mean_loss != 0;
    pc.use_eloss_gradient = cfg->use_eloss_gradient != 0;
    pc.nav.min_mask_tol = cfg->min_mask_tolerance;
    pc.nav.max_mask_tol = cfg->max_mask_tolerance;
    pc.nav.mask_scalor = cfg->mask_tolerance_scalor;
    pc.nav.path_tol = cfg->path_tolerance;
    pc.nav.overstep_tol = cfg->overstep_tolerance;
    pc.nav.win[0] = cfg->search_window[0];
    pc.nav.win[1] = cfg->search_window[1];
    pc.accumulated_error = cfg->accumulated_error;
    pc.n_stddev = cfg->n_scattering_stddev;
    pc.estimate_noise = cfg->estimate_scattering_noise != 0;
    pc.do_eloss = cfg->do_energy_loss != 0;
    pc.do_ms = cfg->do_multiple_scattering != 0;
    pc.use_material = cfg->use_material != 0;
    pc.max_surfaces = cfg->max_surfaces;
    return pc;
}

static int propagate_impl(const dtr_detector_desc* det, const dtr_field_desc* field,
                          const dtr_prop_config* cfg, const dtr_track_batch* in,
                          dtr_result_batch* out, oracle_hits* hits, int n_threads,
                          oracle_trace* trace, const uint32_t* seq, const uint32_t* seq_n,
                          uint32_t seq_max, uint32_t* n_isect = nullptr, int8_t direction = 1) {
    using namespace oracle;
    if (hits && (hits->max_hits == 0u || !hits->surface || !hits->loc || !hits->n_hits))
        return DTR_ERROR_INVALID_ARGUMENT;
    if (!det || !field || !cfg || !in || !out) return DTR_ERROR_INVALID_ARGUMENT;
    if (!in->bound_params && !in->free_params) return DTR_ERROR_INVALID_ARGUMENT;
    if (in->bound_params && !in->start_surface) return DTR_ERROR_INVALID_ARGUMENT;
    PropCfg pc = make_prop_cfg(cfg);
    pc.max_surfaces = out->surfaces ? cfg->max_surfaces : 0u;
    const dtr_actor_states* actors = cfg->actors;
    if (actors) {
        if ((actors->actors & DTR_ACTOR_TARGET_ABORTER) && !actors->target_surface) return DTR_ERROR_INVALID_ARGUMENT;
        if ((actors->actors & DTR_ACTOR_STEP_TRACER) && !actors->trace.n) return DTR_ERROR_INVALID_ARGUMENT;
        pc.actors = actors->actors;
        pc.min_p = static_cast<double>(actors->min_p);
        pc.min_pT = static_cast<double>(actors->min_pT);
    }
    const int64_t n = in->n;
#ifdef _OPENMP
    const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#else
    (void)n_threads;
#endif
    for (int64_t i = 0; i < n; ++i) {
        Engine e;
        e.det.d = det;
        e.field.d = field;
        e.cfg = pc;
        for (int k = 0; k < 6; ++k) e.up.bound[k] = 0.0;
        for (int k = 0; k < 36; ++k) e.up.cov[k] = 0.0;
        double ic[21];
        if (in->initial_cov) {
            for (int k = 0; k < 21; ++k) ic[k] = in->initial_cov[k * n + i];
            e.init_cov = ic;
        }
        if (in->bound_params) {
            // bound start: base_stepper.hpp:84-108 (free = bound_to_free on the
            // start surface), propagator.hpp:121-162 (volume of the surface),
            // parameter_updater_state(cfg, bound_params)
            const uint32_t s0 = in->start_surface[i];
            for (int k = 0; k < 6; ++k) e.up.bound[k] = in->bound_params[k * n + i];
            e.up.sf = s0;
            int k = 0;
            for (int r = 0; r < 6; ++r)
                for (int c = r; c < 6; ++c, ++k) {
                    const double v = in->initial_cov ? ic[k] : (r == c ? 1.0 : 0.0);
                    e.up.cov[r * 6 + c] = v;
                    e.up.cov[c * 6 + r] = v;
                }
            e.init_cov = nullptr;
            e.bound_to_free(s0, e.up.bound, e.st.pos, e.st.dir, e.st.time, e.st.qop);
            e.nav.set_volume(det->surfaces[s0].volume);
        } else {
            const double* fp = in->free_params;
            e.st.pos = {{fp[0 * n + i], fp[1 * n + i], fp[2 * n + i]}};
            e.st.time = fp[3 * n + i];
            e.st.dir = {{fp[4 * n + i], fp[5 * n + i], fp[6 * n + i]}};
            e.st.qop = fp[7 * n + i];
        }
        e.st.ptc = muon_for(e.st.qop);
        e.st.reset_J();
        e.st.set_constraint(0, static_cast<double>(cfg->step_constraint));
        if (pc.max_surfaces) {
            e.seq = out->surfaces + i;
            e.seq_stride = static_cast<uint32_t>(n);
        }
        if (seq) {
            e.dseq = seq + i;
            e.dseq_stride = static_cast<uint64_t>(n);
            e.dseq_len = seq_n[i] < seq_max ? seq_n[i] : seq_max;
            // navigation::direction (navigation_state.hpp:178): intersections
            // along direction * dir, signed step sizes, and the material
            // interactor's energy loss turned into a gain going backward
            e.nav.direction = direction;
        }
        if (hits) {
            e.hit_sf = hits->surface + i;
            e.hit_loc = hits->loc + i;
            e.hit_stride = static_cast<uint64_t>(n);
            e.hit_cap = hits->max_hits;
        }
        if (pc.actors & DTR_ACTOR_TARGET_ABORTER) e.target = actors->target_surface[i];
        if (pc.actors & DTR_ACTOR_STEP_TRACER) {
            // this track's column of the step_tracer arrays
            const dtr_step_trace& t = actors->trace;
            auto col = [&](auto* p) { return p ? p + i : p; };
            e.steps = dtr_step_trace{t.capacity, 0u, col(t.n), col(t.surface), col(t.step_size), col(t.path_length),
                                     col(t.n_trials), col(t.free_params), col(t.bound_params), col(t.jacobian)};
            e.st_stride = static_cast<uint64_t>(n);
        }
        if (trace) {
            e.tr_vol = trace->volume + static_cast<uint64_t>(i) * trace->capacity;
            e.tr_sf = trace->surface + static_cast<uint64_t>(i) * trace->capacity;
            e.tr_cap = trace->capacity;
        }
        tl_isect = 0u;
        e.propagate();
        if (n_isect) n_isect[i] = tl_isect;
        if (pc.actors & DTR_ACTOR_STEP_TRACER) actors->trace.n[i] = e.n_trace;
        uint8_t status;
        if (e.nav.status == kExit) status = e.overflow ? DTR_TRACK_OVERFLOW : DTR_TRACK_EXIT;
        else if (e.nav.exited_path_limit) status = DTR_TRACK_PATH_LIMIT;
        else if (e.momentum_hit) status = DTR_TRACK_MOMENTUM_LIMIT;
        else if (e.target_hit) status = DTR_TRACK_TARGET;
        else status = DTR_TRACK_ABORT;
        if (out->status) out->status[i] = status;
        if (out->bound_params)
            for (int k = 0; k < 6; ++k) out->bound_params[k * n + i] = e.up.bound[k];
        if (out->bound_cov) {
            int k = 0;
            for (int r = 0; r < 6; ++r)
                for (int c = r; c < 6; ++c, ++k) out->bound_cov[k * n + i] = e.up.cov[r * 6 + c];
        }
        if (out->free_params) {
            const double f[8] = {e.st.pos[0], e.st.pos[1], e.st.pos[2], e.st.time,
                                 e.st.dir[0], e.st.dir[1], e.st.dir[2], e.st.qop};
            for (int k = 0; k < 8; ++k) out->free_params[k * n + i] = f[k];
        }
        if (out->path_length) out->path_length[i] = e.st.path_length;
        if (out->n_surfaces) out->n_surfaces[i] = e.n_seq;
        if (out->n_steps) out->n_steps[i] = e.st.n_steps;
        if (out->n_trials) out->n_trials[i] = static_cast<uint32_t>(e.st.n_trials);
        if (trace) trace->n[i] = e.tr_n;
        if (hits) hits->n_hits[i] = e.n_hit;
    }
    return DTR_SUCCESS;
}

int oracle_propagate_hits(const dtr_detector_desc* det, const dtr_field_desc* field,
                          const dtr_prop_config* cfg, const dtr_track_batch* in,
                          dtr_result_batch* out, oracle_hits* hits, int n_threads,
                          oracle_trace* trace) {
    return propagate_impl(det, field, cfg, in, out, hits, n_threads, trace, nullptr, nullptr, 0u);
}

int oracle_propagate(const dtr_detector_desc* det, const dtr_field_desc* field,
                     const dtr_prop_config* cfg, const dtr_track_batch* in,
                     dtr_result_batch* out, int n_threads, oracle_trace* trace) {
    return propagate_impl(det, field, cfg, in, out, nullptr, n_threads, trace, nullptr, nullptr, 0u);
}

int oracle_propagate_counted(const dtr_detector_desc* det, const dtr_field_desc* field,
                             const dtr_prop_config* cfg, const dtr_track_batch* in,
                             dtr_result_batch* out, int n_threads, uint32_t* n_isect) {
    return propagate_impl(det, field, cfg, in, out, nullptr, n_threads, nullptr, nullptr, nullptr, 0u, n_isect);
}

int oracle_propagate_direct(const dtr_detector_desc* det, const dtr_field_desc* field,
                            const dtr_prop_config* cfg, const dtr_track_batch* in,
                            const uint32_t* sequence, const uint32_t* n_sequence, uint32_t max_sequence,
                            uint32_t direction, dtr_result_batch* out, int n_threads) {
    if (in && in->n && (!sequence || !n_sequence || max_sequence == 0u)) return DTR_ERROR_INVALID_ARGUMENT;
    if (direction > DTR_DIRECTION_BACKWARD) return DTR_ERROR_INVALID_ARGUMENT;
    return propagate_impl(det, field, cfg, in, out, nullptr, n_threads, nullptr, sequence, n_sequence,
                          max_sequence, nullptr, direction == DTR_DIRECTION_BACKWARD ? int8_t{-1} : int8_t{1});
}

int oracle_field_at(const dtr_field_desc* field, uint32_t n, const double* xyz, double* b) {
    if (!field || !xyz || !b) return DTR_ERROR_INVALID_ARGUMENT;
    oracle::Field f{field};
    for (uint32_t i = 0; i < n; ++i) {
        const oracle::V3 v = f.at(oracle::V3{{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}});
        b[3 * i] = v[0];
        b[3 * i + 1] = v[1];
        b[3 * i + 2] = v[2];
    }
    return DTR_SUCCESS;
}

int oracle_material_probe(const dtr_material_slab* mat, double qop, double path,
                          double x_over_x0, double out[3]) {
    using oracle::Engine;
    if (!mat || !out) return DTR_ERROR_INVALID_ARGUMENT;
    const oracle::Particle p = oracle::muon_for(qop);
    const Engine::RQ r = Engine::make_rq(p.mass, qop, p.charge);
    out[0] = Engine::bethe_bloch(*mat, r);
    out[1] = Engine::landau_sigma_qop(*mat, r, path, p.charge);
    out[2] = Engine::highland(x_over_x0, qop, p.charge, r.q2OverBeta2);
    return DTR_SUCCESS;
}

int oracle_eloss_probe(const dtr_material_slab* mat, double qop, double out[3]) {
    using oracle::Engine;
    if (!mat || !out || qop == 0.0) return DTR_ERROR_INVALID_ARGUMENT;
    Engine e;
    e.st.ptc = oracle::muon_for(qop);
    out[0] = Engine::bethe_bloch(*mat, Engine::make_rq(e.st.ptc.mass, qop, e.st.ptc.charge));
    out[1] = e.dqopds(qop, mat);
    out[2] = e.d2qopdsdqop(qop, mat);
    return DTR_SUCCESS;
}

int oracle_rk_steps(const dtr_field_desc* field, const dtr_material_slab* vol_mat, const dtr_prop_config* cfg,
                    const double fp[8], double ds, uint32_t n_steps, double* out, double* jac) {
    using namespace oracle;
    if (!field || !cfg || !fp || !out || ds == 0.0) return DTR_ERROR_INVALID_ARGUMENT;
    // a one-volume detector that holds the volume material
    dtr_volume vol{};
    vol.material = vol_mat ? 1u : 0u;
    dtr_detector_desc d{};
    d.n_volumes = 1;
    d.volumes = &vol;
    d.n_materials = vol_mat ? 1u : 0u;
    d.materials = vol_mat;
    Engine e;
    e.det.d = &d;
    e.field.d = field;
    e.cfg = make_prop_cfg(cfg);
    e.nav.set_volume(0);
    e.st.pos = {{fp[0], fp[1], fp[2]}};
    e.st.time = fp[3];
    e.st.dir = {{fp[4], fp[5], fp[6]}};
    e.st.qop = fp[7];
    e.st.ptc = muon_for(e.st.qop);
    e.st.reset_J();
    auto record = [&](uint32_t row) {
        double* o = out + 5u * row;
        o[0] = e.st.qop;
        o[1] = e.dqopds(e.st.qop, vol_mat);
        o[2] = e.d2qopdsdqop(e.st.qop, vol_mat);
        o[3] = e.st.path_length;
        o[4] = e.st.J[7 * 8 + 7];
    };
    record(0);
    for (uint32_t i = 0; i < n_steps; ++i) {
        e.step(ds, true);
        record(i + 1u);
    }
    if (jac) {
        for (int k = 0; k < 64; ++k) jac[k] = e.st.J[k];
        jac[64 + 0] = e.st.pos[0];
        jac[64 + 1] = e.st.pos[1];
        jac[64 + 2] = e.st.pos[2];
        jac[64 + 3] = e.st.time;
        jac[64 + 4] = e.st.dir[0];
        jac[64 + 5] = e.st.dir[1];
        jac[64 + 6] = e.st.dir[2];
        jac[64 + 7] = e.st.qop;
    }
    return DTR_SUCCESS;
}

int oracle_surface_candidates(const dtr_detector_desc* det, uint32_t sf, const double ro[3],
                              const double rd[3], double tol, double ext_tol, double overstep,
                              double out[2][5]) {
    using namespace oracle;
    if (!det || !ro || !rd || !out || sf >= det->n_surfaces) return 0;
    const Det D{det};
    const V3 o{{ro[0], ro[1], ro[2]}}, d{{rd[0], rd[1], rd[2]}};
    NavCfg cfg{static_cast<float>(tol), static_cast<float>(tol), 0.f, 1e-3f, static_cast<float>(overstep),
               {0u, 0u}};
    const dtr_surface& s = D.sf(sf);
    const dtr_transform& t = D.trf(sf);
    int found = 0;
    auto record = [&](const Cand& is, const V3& glob) {
        out[found][0] = is.path;
        for (int i = 0; i < 3; ++i) out[found][1 + i] = glob[i];
        out[found][4] = static_cast<double>(is.status);
        ++found;
    };
    if (s.shape == DTR_CYLINDER2D) {
        double sol[2];
        const int n = cyl2_intersect(o, d, t, D.mask(s.mask).v[0], sol);
        for (uint32_t k = 0; k < s.n_masks && found < 2; ++k) {
            for (int i = 0; i < n && found < 2; ++i) {
                if (!(std::fabs(sol[i]) < 10.0 * 1000.0)) continue;
                Cand is;
                const V3 glob = add(o, mul(sol[i], d));
                resolve_mask(is, sol[i], glob, 0.0, sf, s, D.mask(s.mask + k), t, cfg, ext_tol);
                if (is.probably_inside()) record(is, glob);
            }
        }
        return found;
    }
    IPoint ip;
    if (!intersect_surface(D, sf, o, d, cfg.overstep_tol, ip)) return 0;
    const V3 glob = add(o, mul(ip.path, d));
    for (uint32_t k = 0; k < s.n_masks; ++k) {
        Cand is;
        resolve_mask(is, ip.path, ip.glob, ip.z, sf, s, D.mask(s.mask + k), t, cfg, ext_tol);
        if (is.probably_inside()) {
            record(is, glob);
            break;
        }
    }
    return found;
}

int oracle_line_intersect(const double ro[3], const double rd[3], const dtr_transform* trf,
                          const dtr_mask* mask, double tol, double out[4]) {
    using namespace oracle;
    if (!ro || !rd || !trf || !mask || !out) return 0;
    const V3 o{{ro[0], ro[1], ro[2]}}, d{{rd[0], rd[1], rd[2]}};
    double s;
    V3 glob{{0, 0, 0}};
    if (!line_intersect(o, d, *trf, s, glob) || !(s >= 0.0)) return 0;
    bool inside, edge;
    check_line(*mask, *trf, glob, tol, 0.0, inside, edge);
    out[0] = s;
    Engine::line_local(*trf, glob, d, out[1], out[2]);
    out[3] = inside ? 1.0 : 0.0;
    return 1;
}

}  // extern "C"

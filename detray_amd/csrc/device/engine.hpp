// Batched propagation engine: the same algorithm as propagator.hpp (detray's
// propagate loop, rk_stepper, caching_navigator, actor chain), re-laid out for
// CDNA throughput:
//
//  * compact per-surface navigation records (NavSurface): translation,
//    rotation, first mask and links in one 160-B record, so a candidate costs
//    one dependent gather (bin -> record) instead of three
//    (bin -> surface -> transform/mask);
//  * the 8-slot candidate cache lives in LDS, slot-major ([slot][lane]), so a
//    wave's accesses to one slot hit 64 consecutive banks;
//  * grid searches gather one bin row (<= 8 bins) at a time with independent
//    loads and run an exact, conservative prefilter (distance of the surface
//    centre from the tangent line against the mask circumradius plus the
//    largest tolerance) before the full intersection + mask check;
//  * a precision policy P: FP64 everywhere, or mixed (FP32 Runge-Kutta
//    stages, field interpolation, Jacobian and covariance transport and
//    material noise terms; FP64 free track state, path length, step-size
//    error estimate, mean energy loss and navigation geometry).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>

#include "detray_amd.h"
#include "device/propagator.hpp"

// The batched engine lets the compiler contract a * b + c into fused
// multiply-adds (the library is otherwise built with -ffp-contract=off so
// that the plain FP64 kernel rounds in source order): its contract with the
// FP64 reference is a tolerance (1e-9 relative for the FP64 policy), and
// fused operations shorten the FP32 stage and transport arithmetic.
// Contraction is per source expression (`on`, not `fast`), so a track
// rounds the same way in every kernel that instantiates this code, whatever
// the surrounding code and register allocation.
#pragma clang fp contract(on)

namespace dtr {

// Track-state code is forced inline so that the per-track state stays in
// registers (a call that takes its address would move it to scratch).
#define DTR_AI __host__ __device__ inline __attribute__((always_inline))

// Compact navigation record of one surface (built from the detector tables
// when the detector is uploaded).
struct alignas(16) NavSurface {
    double t[3];       // translation
    double r[9];       // rotation, row-major (columns = local axes)
    double m[4];       // values of the first mask
    double radius;     // circumradius of the (planar) first mask, 0 if none
    uint64_t barcode;
    uint32_t mask;     // first mask in the mask table
    uint32_t material; // slab index or kInvalid
    uint16_t link;     // volume link of the first mask
    uint16_t volume;
    uint8_t shape, sf_id, n_masks, pad;
    // multiple of the largest mask tolerance by which a probably-inside
    // intersection can lie beyond `radius` (0 disables the prefilter)
    float tolk;
    uint32_t index;  // surface index in the detector
};
static_assert(sizeof(NavSurface) == 176, "NavSurface layout");

// FP32 prefilter of a surface, conservative: a surface it rejects cannot be
// probably-inside.  Sphere: centre and enlarged circumradius (r < 0 for an
// empty grid bin, r = +inf for a surface that is always evaluated).  Box:
// the plane (normal, local axes) and the half lengths of the mask's
// bounding rectangle (hx < 0: no box test); kx is the tolerance multiple of
// the local-x bound (1 for rectangles, 1 + slope for trapezoids).  Wires
// (kx < 0): n is the wire axis and hx bounds the distance of an accepted
// point from it, sqrt(2) times the cross section.
struct alignas(16) NavBox {
    float c[3];
    float r;
    float n[3];
    float hx;
    float ex[3];
    float hy;
    float ey[3];
    float kx;
};

// A surface as stored in the candidate lists of the navigation (one per
// grid bin, one per brute-force entry): the FP32 prefilter in the first
// 64 B, then the whole navigation record, so that a candidate costs one
// gather (list entry -> record) and the prefilter touches the same lines.
struct alignas(16) NavRecord {
    NavBox box;
    NavSurface rec;
};
static_assert(sizeof(NavRecord) == 240, "NavRecord layout");

struct FastDetector {
    const NavSurface* sf;
    const dtr_mask* masks;
    const dtr_material_slab* materials;
    const dtr_material_map* mat_maps;
    const dtr_volume* volumes;
    const dtr_grid* grids;
    const dtr_transform* transforms;
    const uint32_t* bf;
    const uint32_t* bins;
    const NavRecord* bin_rec;  // [n_bins] record of the surface in each grid bin
    const NavRecord* bf_rec;   // [n_bf] record of each brute-force entry
    const float* grid_tolk;    // [n_grids] largest tolerance multiple of the grid's surfaces
};

struct PolicyFP64 {
    using S = double;
    static constexpr bool kMixed = false;
    static constexpr bool kGrad = false;
    static constexpr bool kExt = false;
    static constexpr bool kLdsState = false;
    static constexpr bool kVolMat = false;
};
struct PolicyMixed {
    using S = float;
    static constexpr bool kMixed = true;
    static constexpr bool kGrad = false;
    static constexpr bool kExt = false;
    static constexpr bool kLdsState = false;
    static constexpr bool kVolMat = false;
};
// Detectors with wires, annuli, placed cylinders, material rods, rings with
// material and grid bins of several surfaces (the extended geometry): the
// engine resolves these shapes, their local frames and Jacobians, and every
// entry of a bin.
template <class Base>
struct WithExtendedGeometry : Base {
    static constexpr bool kExt = true;
};
// stepping::config::use_field_gradient: the transport Jacobian also carries
// the position columns (rk_stepper.hpp:878-937)
template <class Base>
struct WithFieldGradient : Base {
    static constexpr bool kGrad = true;
};
// Detectors whose volumes hold homogeneous material: the stepper follows the
// continuous energy loss (stage q/p, d(qop)/ds in the path correction and,
// with use_eloss_gradient, in the transport Jacobian).  Used with the
// extended geometry only.
template <class Base>
struct WithVolumeMaterial : Base {
    static constexpr bool kVolMat = true;
};
// The reduced transport Jacobian and the covariance (46 values of the stage
// precision) live in the lane's LDS slots instead of registers: they are
// read and written once per step (the Jacobian) or per surface (the
// covariance), while the registers they free keep the stepping and
// navigation temporaries out of scratch memory.
template <class Base>
struct WithLdsState : Base {
    static constexpr bool kLdsState = true;
};
constexpr int kStateSlots = 25 + 21;
// LDS words per lane of the track state: the lane's values are contiguous
// (constant offsets from one base address); an odd stride keeps the 64
// lanes of a wave on distinct banks when they read the same element.
constexpr int kStateStride = kStateSlots + 1;

// A per-lane array of the track state: in registers, or (kLds) in LDS,
// element i of this lane at p[i].
template <typename S, int N, bool kLds>
struct LaneArray {
    S v[N];
    DTR_AI S& operator[](int i) { return v[i]; }
    DTR_AI const S& operator[](int i) const { return v[i]; }
};
template <typename S, int N>
struct LaneArray<S, N, true> {
    S* p;
    DTR_AI S& operator[](int i) const { return p[i]; }
};

// Position columns of the transport Jacobian (rows pos 0-2 then dir 0-2),
// stored only when the policy transports them.
template <typename S, bool kOn>
struct PosColumns {
    S v[18];
};
template <typename S>
struct PosColumns<S, false> {};

// Host: the navigation record of surface s of a flattened detector.  The
// prefilter bound (radius, tolk) covers every point a mask check of the
// first mask can accept: rectangles |x| <= hx + t, |y| <= hy + t lie within
// hypot(hx, hy) + sqrt(2) t; trapezoids, whose x bound grows with y beyond
// the mask edge by t |hx_high - hx_low| / (2 hy), within hypot(max hx, hy) +
// sqrt((1 + slope)^2 + 1) t; rings within r_outer + t.
inline NavSurface make_nav_surface(const dtr_detector_desc& d, uint32_t s) {
    const dtr_surface& sf = d.surfaces[s];
    const dtr_transform& t = d.transforms[sf.transform];
    const dtr_mask& m = d.masks[sf.mask];
    NavSurface r{};
    for (int i = 0; i < 3; ++i) r.t[i] = t.t[i];
    for (int i = 0; i < 9; ++i) r.r[i] = t.r[i];
    for (int i = 0; i < 4; ++i) r.m[i] = m.v[i];
    r.barcode = sf.barcode;
    r.mask = sf.mask;
    r.material = sf.material;
    r.link = m.volume_link;
    r.volume = sf.volume;
    r.shape = sf.shape;
    r.sf_id = sf.sf_id;
    r.n_masks = sf.n_masks;
    r.radius = 0.0;
    r.tolk = 0.f;
    r.index = s;
    if (sf.n_masks == 1) {
        if (sf.shape == DTR_RECTANGLE2D) {
            r.radius = sqrt(m.v[0] * m.v[0] + m.v[1] * m.v[1]);
            r.tolk = 2.f;
        } else if (sf.shape == DTR_TRAPEZOID2D) {
            const double hx = m.v[0] > m.v[1] ? m.v[0] : m.v[1];
            const double slope = fabs(m.v[1] - m.v[0]) * fabs(m.v[3]);
            r.radius = sqrt(hx * hx + m.v[2] * m.v[2]);
            r.tolk = static_cast<float>(sqrt((1.0 + slope) * (1.0 + slope) + 1.0) * 1.01 + 0.01);
        } else if (sf.shape == DTR_RING2D) {
            r.radius = m.v[1];
            r.tolk = 2.f;
        } else if (sf.shape == DTR_LINE_SQUARE || sf.shape == DTR_LINE_CIRCULAR) {
            // a wire accepts points within sqrt(2) (r + t) of its axis and
            // hz + t along it: within hypot(hz, sqrt(2) r) + sqrt(3) t of
            // its centre
            r.radius = sqrt(m.v[1] * m.v[1] + 2.0 * m.v[0] * m.v[0]);
            r.tolk = 2.f;
        }
    }
    return r;
}

// Host: the candidate-list records (grid bins and brute-force entries) and
// the largest tolerance multiple per grid.  FP32 rounding of the sphere
// centre, of the track position and of the FP32 distance computation is
// covered by enlarging the radius by 1e-5 |c|_1 here and by 1e-5 |p|_1 +
// 1e-3 mm at the test; the box test adds its own error margin.
inline float round_up_float(double v) {
    float f = static_cast<float>(v);
    if (static_cast<double>(f) < v) f = nextafterf(f, std::numeric_limits<float>::infinity());
    return f;
}

inline NavBox make_box(const NavSurface& r) {
    NavBox b{};
    b.r = std::numeric_limits<float>::infinity();
    b.hx = -1.f;
    if (r.tolk == 0.f) return b;
    for (int i = 0; i < 3; ++i) b.c[i] = static_cast<float>(r.t[i]);
    b.r = round_up_float(r.radius + 1e-5 * (fabs(r.t[0]) + fabs(r.t[1]) + fabs(r.t[2])));
    if (r.shape == DTR_RECTANGLE2D || r.shape == DTR_TRAPEZOID2D || r.shape == DTR_LINE_SQUARE ||
        r.shape == DTR_LINE_CIRCULAR) {
        for (int i = 0; i < 3; ++i) {
            b.ex[i] = static_cast<float>(r.r[3 * i + 0]);
            b.ey[i] = static_cast<float>(r.r[3 * i + 1]);
            b.n[i] = static_cast<float>(r.r[3 * i + 2]);
        }
        if (r.shape == DTR_RECTANGLE2D) {
            b.hx = round_up_float(r.m[0]);
            b.hy = round_up_float(r.m[1]);
            b.kx = 1.f;
        } else if (r.shape == DTR_LINE_SQUARE || r.shape == DTR_LINE_CIRCULAR) {
            b.hx = round_up_float(1.4142135623730951 * r.m[0] * (1.0 + 1e-7));
            b.hy = 0.f;
            b.kx = -1.f;
        } else {
            const double hx = r.m[0] > r.m[1] ? r.m[0] : r.m[1];
            b.hx = round_up_float(hx);
            b.hy = round_up_float(r.m[2]);
            b.kx = round_up_float(1.0 + fabs(r.m[1] - r.m[0]) * fabs(r.m[3]));
        }
    }
    return b;
}

inline void make_nav_records(const dtr_detector_desc& d, const NavSurface* nav, NavRecord* bin_rec,
                             NavRecord* bf_rec, float* grid_tolk) {
    for (uint32_t k = 0; k < d.n_bins; ++k) {
        const uint32_t s = d.grid_bins[k];
        NavRecord e{};
        e.box.r = -1.f;
        e.box.hx = -1.f;
        e.rec.index = 0xffffffffu;
        if (s != 0xffffffffu) {
            e.rec = nav[s];
            e.box = make_box(nav[s]);
        }
        bin_rec[k] = e;
    }
    for (uint32_t k = 0; k < d.n_bf; ++k) {
        NavRecord e{};
        e.rec = nav[d.bf_surfaces[k]];
        e.box = make_box(e.rec);
        bf_rec[k] = e;
    }
    for (uint32_t gi = 0; gi < d.n_grids; ++gi) {
        const dtr_grid& g = d.grids[gi];
        float t = 0.f;
        for (uint32_t k = 0; k < g.n0 * g.n1 * (g.capacity ? g.capacity : 1u); ++k) {
            const uint32_t s = d.grid_bins[g.bins + k];
            if (s != 0xffffffffu && nav[s].tolk > t) t = nav[s].tolk;
        }
        grid_tolk[gi] = t;
    }
}

template <typename S>
struct T3 {
    S x, y, z;
};
template <typename S>
DTR_HD T3<S> tadd(T3<S> a, T3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename S>
DTR_HD T3<S> tsub(T3<S> a, T3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename S>
DTR_HD T3<S> tmul(S s, T3<S> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename S>
DTR_HD T3<S> tcross(T3<S> a, T3<S> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <typename S>
DTR_HD S tdot(T3<S> a, T3<S> b) {
    S r = S(0);
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
template <typename S>
DTR_HD T3<S> tcast(V3 v) { return {static_cast<S>(v.x), static_cast<S>(v.y), static_cast<S>(v.z)}; }
template <typename S>
DTR_HD V3 tdbl(T3<S> v) { return {static_cast<double>(v.x), static_cast<double>(v.y), static_cast<double>(v.z)}; }

// Field lookup in the stage precision
template <typename S>
DTR_HD T3<S> field_at_s(const DevField& f, V3 p) {
    if (f.kind == DTR_FIELD_CONSTANT)
        return {static_cast<S>(f.b[0]), static_cast<S>(f.b[1]), static_cast<S>(f.b[2])};
    const double pc[3] = {p.x, p.y, p.z};
    int i0[3];
    S t[3];
    for (int a = 0; a < 3; ++a) {
        const int n = static_cast<int>(f.n[a]);
        double u = (pc[a] - f.min[a]) / f.pitch[a];
        u = (u > 0.0) ? (u > n - 1 ? static_cast<double>(n - 1) : u) : 0.0;
        int i = static_cast<int>(floor(u));
        i = i > n - 2 ? n - 2 : i;
        i0[a] = i;
        t[a] = static_cast<S>(u - i);
    }
    const uint64_t ny = f.n[1], nz = f.n[2];
    const uint64_t base = (static_cast<uint64_t>(i0[0]) * ny + i0[1]) * nz + i0[2];
    const uint64_t sx = ny * nz, sy = nz;
    const float* v000 = f.values + 3 * base;
    S out[3];
    for (int c = 0; c < 3; ++c) {
        const S a000 = v000[c], a100 = v000[3 * sx + c];
        const S a001 = v000[3 + c], a101 = v000[3 * sx + 3 + c];
        const S a010 = v000[3 * sy + c], a110 = v000[3 * (sx + sy) + c];
        const S a011 = v000[3 * sy + 3 + c], a111 = v000[3 * (sx + sy) + 3 + c];
        const S one = S(1);
        const S c00 = a000 * (one - t[0]) + a100 * t[0];
        const S c01 = a001 * (one - t[0]) + a101 * t[0];
        const S c10 = a010 * (one - t[0]) + a110 * t[0];
        const S c11 = a011 * (one - t[0]) + a111 * t[0];
        const S c0 = c00 * (one - t[1]) + c10 * t[1];
        const S c1 = c01 * (one - t[1]) + c11 * t[1];
        out[c] = c0 * (one - t[2]) + c1 * t[2];
    }
    return {out[0], out[1], out[2]};
}

// Candidate cache view: slot i of this lane at base[i * stride].  `state`:
// the lane's kStateSlots LDS words of the track state (WithLdsState),
// element i at state[i], or null.
struct CacheView {
    double* path;
    uint32_t* sf;
    uint32_t* ls;  // volume link (bits 0-15) | status (16-23) | surface id (24-31)
    int stride;
    float* state = nullptr;
    DTR_HD double& p(int i) const { return path[i * stride]; }
    DTR_HD uint32_t& s(int i) const { return sf[i * stride]; }
    DTR_HD uint32_t& l(int i) const { return ls[i * stride]; }
    DTR_HD uint8_t status(int i) const { return static_cast<uint8_t>(l(i) >> 16); }
    DTR_HD uint16_t link(int i) const { return static_cast<uint16_t>(l(i) & 0xffffu); }
    DTR_HD uint8_t sf_id(int i) const { return static_cast<uint8_t>(l(i) >> 24); }
    DTR_HD void copy(int dst, int src) const {
        p(dst) = p(src);
        s(dst) = s(src);
        l(dst) = l(src);
    }
    DTR_HD void set(int i, const Cand& c) const {
        p(i) = c.path;
        s(i) = c.sf;
        l(i) = static_cast<uint32_t>(c.link) | (static_cast<uint32_t>(c.status) << 16) |
               (static_cast<uint32_t>(c.pad) << 24);
    }
    DTR_HD Cand get(int i) const {
        Cand c;
        c.path = p(i);
        c.sf = s(i);
        c.link = link(i);
        c.status = status(i);
        c.pad = sf_id(i);
        return c;
    }
};

// Muon hypothesis, charge-conjugated to the track (pdg_particle.hpp:61-64):
// the mass is a constant and the charge follows the sign of q/p, which no
// interaction changes (a track brought to rest ends its propagation).
constexpr double kMuonMass = static_cast<double>(105.6583755f * 1e-3f);
DTR_AI double track_charge(double qop) { return (-1.0 * qop > 0.0) ? -1.0 : 1.0; }

template <class P>
struct FastTrack {
    using S = typename P::S;
    // navigation
    double ext_tol;
    int8_t next, last, status;
    uint8_t trust;
    uint16_t volume;
    bool path_limit_hit, overflow, stopped, heartbeat;
    bool momentum_hit, target_hit;  // the optional aborters fired
    // navigation::direction as a sign (navigation_state.hpp:178): +1, or -1
    // for a backward direct navigation (intersections along -dir, negative
    // steps, energy gained in material); +1 in every other kernel, where it
    // folds away
    double nav_dir;
    // free track parameters (FP64 state)
    V3 pos, dir;
    double time, qop;
    // stepping
    double step_size, next_step_size, path_length, abs_path_length;
    double constraint_abort;
    T3<S> dtds3;
    double dqopds3;  // d(qop)/ds of the last step's fourth stage (volume material)
    double h_stage;  // step size the stages of the last step were evaluated at
    uint32_t n_trials, n_steps;
    // transport Jacobian (reduced) and covariance in the stage precision
    LaneArray<S, 25, P::kLdsState> J;
    PosColumns<S, P::kGrad> Jx;
    LaneArray<S, 21, P::kLdsState> cov;
    double bound[6];
    uint32_t up_sf;
    uint32_t n_seq;
    uint32_t n_trace;  // step_tracer entries
};

// The navigation fields of a track: what local navigation reads and writes.
struct NavLite {
    double ext_tol;
    int8_t next, last, status;
    uint8_t trust;
    uint16_t volume;
    V3 pos, dir;
};

template <class P>
DTR_HD uint32_t local_navigation_ool(const FastDetector& det, const DevConfig& cfg, V3 pos, V3 dir,
                                     double ext_tol, uint16_t volume, float overstep, CacheView c);

// TK is the track type: FastTrack<P> for the propagation loop, NavLite inside
// the out-of-line local navigation.  kJacOut: transport_covariance also
// writes the full 6x6 bound-to-bound Jacobian to jf_out (recording pass of
// the adjoint).
template <class P, class TK = FastTrack<P>, bool kJacOut = false>
struct FastEngine {
    using S = typename P::S;
    const FastDetector& det;
    const DevField& field;
    const DevConfig& cfg;
    TK& tk;
    CacheView c;
    double* jf_out = nullptr;

    DTR_AI bool on_surface() const { return tk.status == kOnObject || tk.status == kOnPortal; }
    DTR_AI int n_cand() const { return tk.last - tk.next + 1; }
    DTR_AI bool exhausted() const { return n_cand() == 0; }
    DTR_AI int begin_idx() const { return (on_surface() && tk.next >= 1) ? tk.next - 1 : tk.next; }
    // the track's navigation direction sign; NavLite (the searches of the
    // caching navigator) always looks forward
    DTR_AI double dir_sign() const {
        if constexpr (requires { tk.nav_dir; }) return tk.nav_dir;
        else return 1.0;
    }
    // navigation_state::operator() (navigation_state.hpp:355-359): the signed
    // distance to the target
    DTR_AI double nav_dist() const { return dir_sign() * c.p(tk.next); }
    DTR_AI void set_high_trust() { tk.trust = tk.trust < kHigh ? tk.trust : kHigh; }
    DTR_AI void set_fair_trust() { tk.trust = tk.trust < kFair ? tk.trust : kFair; }
    DTR_AI bool is_portal(uint32_t s) const { return det.sf[s].sf_id == DTR_PORTAL; }
    DTR_AI uint32_t cur_sf() const { return c.s(tk.next - 1); }

    // -------------------------------------------------------- intersections --
    static DTR_AI V3 rec_axis(const NavSurface& r, int k) { return {r.r[k], r.r[3 + k], r.r[6 + k]}; }
    static DTR_AI V3 rec_to_local(const NavSurface& r, V3 p) {
        const V3 d{p.x - r.t[0], p.y - r.t[1], p.z - r.t[2]};
        return {r.r[0] * d.x + r.r[3] * d.y + r.r[6] * d.z, r.r[1] * d.x + r.r[4] * d.y + r.r[7] * d.z,
                r.r[2] * d.x + r.r[5] * d.y + r.r[8] * d.z};
    }
    static DTR_AI V3 rec_to_global(const NavSurface& r, V3 l) {
        return {r.r[0] * l.x + r.r[1] * l.y + r.r[2] * l.z + r.t[0],
                r.r[3] * l.x + r.r[4] * l.y + r.r[5] * l.z + r.t[1],
                r.r[6] * l.x + r.r[7] * l.y + r.r[8] * l.z + r.t[2]};
    }

    static DTR_AI dtr_transform rec_trf(const NavSurface& r) {
        dtr_transform t;
        for (int i = 0; i < 9; ++i) t.r[i] = r.r[i];
        for (int i = 0; i < 3; ++i) t.t[i] = r.t[i];
        return t;
    }
    static DTR_AI bool is_line(uint8_t shape) { return shape == DTR_LINE_SQUARE || shape == DTR_LINE_CIRCULAR; }
    // mask values of the first mask (annuli have seven, the record holds four)
    DTR_AI const double* first_mask(const NavSurface& rec) const {
        return (P::kExt && rec.shape == DTR_ANNULUS2D) ? det.masks[rec.mask].v : rec.m;
    }
    // intersection_update of a placed cylinder: the smaller solution
    DTR_AI bool cyl2_update(const NavSurface& rec, double& path, V3& glob) const {
        double sol[2];
        const V3 rd = mul(dir_sign(), tk.dir);
        if (Engine::cyl2_intersect(tk.pos, rd, rec_trf(rec), rec.m[0], sol) == 0) return false;
        path = sol[0];
        glob = add(tk.pos, mul(sol[0], rd));
        return fabs(sol[0]) < 10.0 * 1000.0;
    }

    // ray_plane_intersector.hpp:60-87 on a record
    static DTR_AI bool plane_rec(V3 ro, V3 rd, const NavSurface& r, double& s, V3& glob) {
        const V3 sn = rec_axis(r, 2);
        const V3 st{r.t[0], r.t[1], r.t[2]};
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

    // intersector_base.hpp:171-252 resolve_mask
    DTR_AI void resolve(Cand& is, double path, V3 glob, double z2d, uint32_t sf_idx, uint8_t sf_id,
                        const double* mv, uint8_t shape, uint16_t link, const NavSurface& rec,
                        float overstep) const {
        is.status = kOutside;
        if (!(path >= static_cast<double>(overstep))) return;
        double base_tol = 0.0, ext_tol = 0.0;
        if (sf_id != DTR_PORTAL || cfg.min_mask_tol == 3.40282347e+38f) {
            ext_tol = tk.ext_tol;
            base_tol = dmax(static_cast<double>(cfg.min_mask_tol),
                            dmin(static_cast<double>(cfg.max_mask_tol),
                                 static_cast<double>(cfg.mask_scalor) * fabs(path)));
        }
        bool inside, edge;
        if (shape == DTR_CONCENTRIC_CYLINDER2D) {
            inside = (mv[1] - base_tol <= z2d && z2d <= mv[2] + base_tol);
            edge = inside;
        } else {
            const V3 loc = rec_to_local(rec, glob);
            if (shape == DTR_RECTANGLE2D) {
                const double l0 = fabs(loc.x), l1 = fabs(loc.y);
                inside = (l0 <= (mv[0] + base_tol) && l1 <= (mv[1] + base_tol));
                edge = false;
                if (ext_tol > 0.0) {
                    const double ft = base_tol + ext_tol;
                    edge = (l0 <= (mv[0] + ft) && l1 <= (mv[1] + ft));
                }
            } else if (shape == DTR_TRAPEZOID2D) {
                const double rel_y = (mv[2] + loc.y) * mv[3];
                const double bound_x = mv[0] + rel_y * (mv[1] - mv[0]);
                inside = (fabs(loc.x) <= (bound_x + base_tol) && fabs(loc.y) <= (mv[2] + base_tol));
                edge = false;
                if (ext_tol > 0.0) {
                    const double ft = base_tol + ext_tol;
                    edge = (fabs(loc.x) <= (bound_x + ft) && fabs(loc.y) <= (mv[2] + ft));
                }
            } else if (P::kExt && shape >= DTR_LINE_SQUARE) {
                ext_mask_check(shape, mv, loc, base_tol, ext_tol, inside, edge);
            } else {
                const double r = perp(loc);
                inside = ((r + base_tol) >= mv[0] && r <= (mv[1] + base_tol));
                edge = inside;
            }
        }
        if (edge) is.status = kEdge;
        if (inside) is.status = kInside;
        is.path = path;
        is.sf = sf_idx;
        is.link = link;
        is.pad = sf_id;  // the surface id travels with the candidate
    }

    // Intersect surface s and resolve its masks until one is probably inside
    // (intersection_kernel.hpp:24-168 for init, 172-272 for update).
    // Returns true if a probably-inside result was produced in `out`.
    DTR_AI bool evaluate(const NavSurface& rec, float overstep, Cand& out) const {
        const uint32_t s = rec.index;
        double path, z = 0.0;
        V3 glob{0.0, 0.0, 0.0};
        bool valid;
        // the ray along the navigation direction (navigation_functions.hpp:
        // 119-122)
        const V3 rd = mul(dir_sign(), tk.dir);
        if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            valid = Engine::cyl_intersect(tk.pos, rd, rec.m[0], overstep, path, z);
        } else if (P::kExt && is_line(rec.shape)) {
            valid = Engine::line_intersect(tk.pos, rd, rec_trf(rec), path, glob);
        } else if (P::kExt && rec.shape == DTR_CYLINDER2D) {
            valid = cyl2_update(rec, path, glob);
        } else {
            valid = plane_rec(tk.pos, rd, rec, path, glob);
        }
        if (!valid) return false;
        if (rec.n_masks == 1) {
            resolve(out, path, glob, z, s, rec.sf_id, first_mask(rec), rec.shape, rec.link, rec, overstep);
            return out.status <= kEdge;
        }
        for (uint32_t k = 0; k < rec.n_masks; ++k) {
            const dtr_mask& m = det.masks[rec.mask + k];
            resolve(out, path, glob, z, s, rec.sf_id, m.v, rec.shape, m.volume_link, rec, overstep);
            if (out.status <= kEdge) return true;
        }
        return false;
    }

    // caching_navigator.hpp:149-229 on the LDS cache
    DTR_AI void insert(int pos, const Cand& nc) {
        if (pos == kCache) return;
        if (n_cand() == 0) {
            c.set(0, nc);
            tk.last = static_cast<int8_t>(tk.last + 1);
            return;
        }
        int idx = pos;
        // barcodes are unique per surface (the surface index is part of
        // them), so equal barcodes are equal surface indices
        auto overlap = [&](int i) { return fabs(c.p(i) - nc.path) <= 1.0 * 1e-3; };
        auto clash = [&](int i) { return c.s(i) == nc.sf && overlap(i); };
        if (clash(idx) || (idx > 0 && clash(idx - 1))) return;
        const int shift_begin = tk.last < kCache - 2 ? tk.last : kCache - 2;
        if (overlap(idx) && (nc.status == kEdge || nc.pad == DTR_PORTAL)) {
            idx++;
            idx = idx > shift_begin ? shift_begin : idx;
        }
        for (int i = shift_begin; i >= idx; --i) c.copy(i + 1, i);
        c.set(idx, nc);
        tk.last = static_cast<int8_t>(tk.last + 1 < kCache - 1 ? tk.last + 1 : kCache - 1);
    }

    DTR_AI void insert_sorted(const Cand& is) {
        int p = begin_idx();
        const int e = tk.last + 1;
        while (p < e && !(fabs(is.path) < fabs(c.p(p)))) ++p;
        insert(p, is);
    }

    DTR_AI void candidate_search(const NavSurface& rec, float overstep) {
        if constexpr (P::kExt) {
            if (rec.shape == DTR_CYLINDER2D) {
                cyl2_search(rec, overstep);
                return;
            }
        }
        Cand is;
        is.path = kMax;
        is.sf = kInvalid;
        is.link = 0;
        is.status = kOutside;
        is.pad = 0;
        if (evaluate(rec, overstep, is)) insert_sorted(is);
    }

    // intersection_initialize of a placed cylinder: every valid solution a
    // mask accepts, at most two (intersection_kernel.hpp:24-168)
    DTR_AI void cyl2_search(const NavSurface& rec, float overstep) {
        double sol[2];
        const int n = Engine::cyl2_intersect(tk.pos, tk.dir, rec_trf(rec), rec.m[0], sol);
        for (uint32_t k = 0; k < rec.n_masks; ++k) {
            const dtr_mask& m = det.masks[rec.mask + k];
            int found = 0;
            for (int i = 0; i < n; ++i) {
                if (!(fabs(sol[i]) < 10.0 * 1000.0)) continue;
                Cand is;
                is.path = kMax;
                is.sf = kInvalid;
                is.link = 0;
                is.status = kOutside;
                is.pad = 0;
                resolve(is, sol[i], add(tk.pos, mul(sol[i], tk.dir)), 0.0, rec.index, rec.sf_id, m.v, rec.shape,
                        m.volume_link, rec, overstep);
                if (is.status <= kEdge) {
                    insert_sorted(is);
                    if (++found == 2) return;
                }
            }
        }
    }

    DTR_AI bool update_candidate(int slot) {
        const uint32_t s = c.s(slot);
        if (s == kInvalid) return false;
        Cand cd = c.get(slot);
        const NavSurface& rec = det.sf[s];
        double path, z = 0.0;
        V3 glob{0.0, 0.0, 0.0};
        bool valid;
        if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            valid = Engine::cyl_intersect(tk.pos, tk.dir, rec.m[0], cfg.overstep_tol, path, z);
        } else if (P::kExt && is_line(rec.shape)) {
            valid = Engine::line_intersect(tk.pos, tk.dir, rec_trf(rec), path, glob);
        } else if (P::kExt && rec.shape == DTR_CYLINDER2D) {
            valid = cyl2_update(rec, path, glob);
        } else {
            valid = plane_rec(tk.pos, tk.dir, rec, path, glob);
        }
        if (!valid) return false;
        bool ok = false;
        if (rec.n_masks == 1) {
            resolve(cd, path, glob, z, s, rec.sf_id, first_mask(rec), rec.shape, rec.link, rec, cfg.overstep_tol);
            ok = cd.status <= kEdge;
        } else {
            for (uint32_t k = 0; k < rec.n_masks; ++k) {
                const dtr_mask& m = det.masks[rec.mask + k];
                resolve(cd, path, glob, z, s, rec.sf_id, m.v, rec.shape, m.volume_link, rec,
                        cfg.overstep_tol);
                if (cd.status <= kEdge) {
                    ok = true;
                    break;
                }
            }
        }
        c.set(slot, cd);
        return ok;
    }

    // spatial_grid.hpp:100-202 with a row-batched gather and an exact
    // conservative prefilter: a surface whose centre is farther from the
    // tangent line than its mask circumradius plus twice the largest
    // tolerance cannot be probably-inside, so it is not evaluated.
    // The bin window of the grid lookup (spatial_grid.hpp:100-202,
    // axis_binning.hpp:80-115, axis_bounds.hpp): bins [a0, e0) x [a1, e1)
    // before wrapping of the circular axis.  False for a non-finite lookup
    // point (no grid candidates).
    DTR_AI bool grid_window(const dtr_volume& vol, const dtr_grid& g, int& a0, int& e0, int& a1,
                            int& e1) const {
        double l0, l1;
        if (g.type == DTR_GRID_CYLINDER) {
            double path, z;
            Engine::cyl_intersect(tk.pos, tk.dir, g.radius, -kMax, path, z);
            l0 = phi_of(add(tk.pos, mul(path, tk.dir)));
            l1 = z;
        } else {
            const dtr_transform& vt = det.transforms[vol.transform];
            double s;
            V3 glob;
            Engine::plane_intersect(tk.pos, tk.dir, vt, s, glob);
            const V3 loc = to_local(vt, glob);
            l0 = perp(loc);
            l1 = phi_of(loc);
        }
        const double w0 = (g.max0 - g.min0) / static_cast<double>(g.n0);
        const double w1 = (g.max1 - g.min1) / static_cast<double>(g.n1);
        const double f0 = (l0 - g.min0) / w0 + 1.0, f1 = (l1 - g.min1) / w1 + 1.0;
        if (!(fabs(f0) < 1e9) || !(fabs(f1) < 1e9)) return false;
        const int ib0 = static_cast<int>((l0 - g.min0) / w0 + 1.0) - 1;
        const int ib1 = static_cast<int>((l1 - g.min1) / w1 + 1.0) - 1;
        const int lb0 = ib0 - static_cast<int>(cfg.win0), ub0 = ib0 + static_cast<int>(cfg.win1);
        const int lb1 = ib1 - static_cast<int>(cfg.win0), ub1 = ib1 + static_cast<int>(cfg.win1);
        const int n0 = static_cast<int>(g.n0), n1 = static_cast<int>(g.n1);
        if (g.type == DTR_GRID_CYLINDER) {
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
        return true;
    }

    // Prefilter margin of a grid's surfaces for this track: the largest
    // mask tolerance times the grid's tolerance multiple plus the FP32
    // rounding allowance of the track position.
    DTR_AI float sphere_margin(int grid) const {
        return static_cast<float>(static_cast<double>(det.grid_tolk[grid]) * tol_max() +
                                  1e-5 * (fabs(tk.pos.x) + fabs(tk.pos.y) + fabs(tk.pos.z)) + 1e-3);
    }

    // largest tolerance a mask check can use
    DTR_AI double tol_max() const {
        return dmax(static_cast<double>(cfg.min_mask_tol), static_cast<double>(cfg.max_mask_tol)) + tk.ext_tol;
    }

    // FP32 prefilter of one record for the ray (o, d): false only if the
    // surface cannot be probably-inside.  The sphere test bounds the distance
    // of the centre from the line; the box test intersects the plane (when
    // the incidence is not grazing, |n.d| >= 0.05) and bounds the local
    // coordinates by the mask's rectangle plus tolerance and an FP32 error
    // margin of 2e-4 |c - o|_1 + 1e-5 |c|_1: with unit roundoff u = 6e-8 and
    // |n.d| >= 0.05 the local coordinates are off by at most about
    // 1.4e3 u |c - o| + 25 u (|c| + |o|) (s = n.(c-o)/n.d, |s| <= 20 |c-o|).
    static DTR_AI bool prefilter(const NavBox& b, float ox, float oy, float oz, float dx, float dy, float dz,
                                 float margin, float ft) {
        const float cx = b.c[0] - ox, cy = b.c[1] - oy, cz = b.c[2] - oz;
        const float xx = cy * dz - dy * cz, xy = cz * dx - dz * cx, xz = cx * dy - dx * cy;
        const float dist2 = xx * xx + xy * xy + xz * xz;
        const float rb = b.r + margin;
        if (!(b.r >= 0.f) || dist2 > rb * rb) return false;
        if (b.hx < 0.f) return true;
        if (P::kExt && b.kx < 0.f) {
            // wire: distance of its axis from the ray, |(c - o).(d x a)| /
            // |d x a|, against sqrt(2) (r + t); near-parallel rays (|d x a|
            // < 0.1) pass.  FP32 error margin 1e-5 (|c - o|_1 + |c|_1) / |d x a|
            // plus the caller's (at least 1e-3 mm), over 100 times the
            // rounding error of the FP32 evaluation.
            const float wx = dy * b.n[2] - dz * b.n[1], wy = dz * b.n[0] - dx * b.n[2],
                        wz = dx * b.n[1] - dy * b.n[0];
            const float w2 = wx * wx + wy * wy + wz * wz;
            if (!(w2 >= 0.01f)) return true;
            const float wn = sqrtf(w2);
            const float dist = fabsf(cx * wx + cy * wy + cz * wz) / wn;
            const float m = 1e-5f * (fabsf(cx) + fabsf(cy) + fabsf(cz) + fabsf(b.c[0]) + fabsf(b.c[1]) +
                                     fabsf(b.c[2])) / wn + margin;
            return !(dist > b.hx + 1.5f * ft + m);
        }
        const float denom = b.n[0] * dx + b.n[1] * dy + b.n[2] * dz;
        if (!(fabsf(denom) >= 0.05f)) return true;
        const float s = (b.n[0] * cx + b.n[1] * cy + b.n[2] * cz) / denom;
        const float px = s * dx - cx, py = s * dy - cy, pz = s * dz - cz;
        const float lx = px * b.ex[0] + py * b.ex[1] + pz * b.ex[2];
        const float ly = px * b.ey[0] + py * b.ey[1] + pz * b.ey[2];
        const float m = 2e-4f * (fabsf(cx) + fabsf(cy) + fabsf(cz)) +
                        1e-5f * (fabsf(b.c[0]) + fabsf(b.c[1]) + fabsf(b.c[2])) + margin;
        return !(fabsf(lx) > b.hx + b.kx * ft + m) && !(fabsf(ly) > b.hy + ft + m);
    }

    DTR_AI bool sphere_pass(uint32_t bin, float margin) const {
        const float ft = static_cast<float>(tol_max());
        return prefilter(det.bin_rec[bin].box, static_cast<float>(tk.pos.x), static_cast<float>(tk.pos.y),
                         static_cast<float>(tk.pos.z), static_cast<float>(tk.dir.x),
                         static_cast<float>(tk.dir.y), static_cast<float>(tk.dir.z), margin, ft);
    }

    // The volume's brute-force surfaces, evaluated two at a time (independent
    // record loads and intersections) and inserted in list order.
    DTR_AI void bf_search(const dtr_volume& vol, float overstep) {
        if constexpr (P::kExt) {
            for (uint32_t k = vol.bf_begin; k < vol.bf_end; ++k) candidate_search(det.bf_rec[k].rec, overstep);
            return;
        }
        uint32_t k = vol.bf_begin;
        for (; k + 2 <= vol.bf_end; k += 2) {
            Cand c0, c1;
            c0.pad = c1.pad = 0;
            const bool a0 = evaluate(det.bf_rec[k].rec, overstep, c0);
            const bool a1 = evaluate(det.bf_rec[k + 1].rec, overstep, c1);
            if (a0) insert_sorted(c0);
            if (a1) insert_sorted(c1);
        }
        if (k < vol.bf_end) candidate_search(det.bf_rec[k].rec, overstep);
    }

    // The window's bins in row-major order, prefiltered kChunk at a time
    // (independent loads, so a chunk costs one memory latency), the passing
    // ones then evaluated and inserted in order.
    DTR_AI void grid_search(const dtr_volume& vol, float overstep) {
        const dtr_grid& g = det.grids[vol.grid];
        int a0, e0, a1, e1;
        if (!grid_window(vol, g, a0, e0, a1, e1)) return;
        const bool cyl = (g.type == DTR_GRID_CYLINDER);
        const int n0 = static_cast<int>(g.n0), n1 = static_cast<int>(g.n1);
        const int w = e1 - a1;
        const int total = (e0 - a0) * w;
        const float margin = sphere_margin(vol.grid);
        const float ft = static_cast<float>(tol_max());
        const float ox = static_cast<float>(tk.pos.x), oy = static_cast<float>(tk.pos.y),
                    oz = static_cast<float>(tk.pos.z);
        const float dx = static_cast<float>(tk.dir.x), dy = static_cast<float>(tk.dir.y),
                    dz = static_cast<float>(tk.dir.z);
        if constexpr (P::kExt) {
            // every entry of each bin in fill order (capacity 0 means 1);
            // empty entries fail the prefilter
            const uint32_t cap = g.capacity ? g.capacity : 1u;
            for (int b0 = a0; b0 < e0; ++b0) {
                const int i0 = cyl ? wrap_bin(b0, n0) : b0;
                for (int b1 = a1; b1 < e1; ++b1) {
                    const int i1 = cyl ? b1 : wrap_bin(b1, n1);
                    const uint32_t first =
                        g.bins + (static_cast<uint32_t>(i0) * g.n1 + static_cast<uint32_t>(i1)) * cap;
                    for (uint32_t k = 0; k < cap; ++k) {
                        const NavRecord& e = det.bin_rec[first + k];
                        if (prefilter(e.box, ox, oy, oz, dx, dy, dz, margin, ft)) candidate_search(e.rec, overstep);
                    }
                }
            }
            return;
        }
        auto bin_of = [&](int j) {
            const int b0 = a0 + j / w, b1 = a1 + j % w;
            const int i0 = cyl ? wrap_bin(b0, n0) : b0;
            const int i1 = cyl ? b1 : wrap_bin(b1, n1);
            return g.bins + static_cast<uint32_t>(i0) * g.n1 + static_cast<uint32_t>(i1);
        };
        constexpr int kChunk = 8;
        for (int j0 = 0; j0 < total; j0 += kChunk) {
            uint32_t pass = 0u;
            int b0 = a0 + j0 / w, b1 = a1 + j0 % w;
#pragma unroll
            for (int k = 0; k < kChunk; ++k) {
                if (j0 + k < total) {
                    const int i0 = cyl ? wrap_bin(b0, n0) : b0;
                    const int i1 = cyl ? b1 : wrap_bin(b1, n1);
                    const uint32_t bin = g.bins + static_cast<uint32_t>(i0) * g.n1 + static_cast<uint32_t>(i1);
                    if (prefilter(det.bin_rec[bin].box, ox, oy, oz, dx, dy, dz, margin, ft)) pass |= (1u << k);
                }
                if (++b1 == e1) {
                    b1 = a1;
                    ++b0;
                }
            }
            while (pass) {
                const int k = __builtin_ctz(pass);
                pass &= pass - 1u;
                candidate_search(det.bin_rec[bin_of(j0 + k)].rec, overstep);
            }
        }
    }

    // circular bin index: b mod n in [0, n)
    static DTR_AI int wrap_bin(int b, int n) {
        if (b >= 0 && b < n) return b;
        if (b < 0 && b >= -n) return b + n;
        if (b >= n && b < 2 * n) return b - n;
        return (n + (b % n)) % n;
    }

    // navigation_functions.hpp:147-177
    DTR_AI void update_status() {
        if (!exhausted() && fabs(c.p(tk.next)) < static_cast<double>(cfg.path_tol)) {
            tk.status = (c.sf_id(tk.next) == DTR_PORTAL) ? kOnPortal : kOnObject;
            tk.next = static_cast<int8_t>(tk.next + 1);
        } else {
            tk.status = kTowards;
        }
        tk.trust = (exhausted() || tk.status == kOnPortal) ? kNoTrust : kFull;
    }

    // navigation_functions.hpp:193-246: the volume's portals and passives
    // (brute force) and the grid neighbourhood of its sensitives
    DTR_AI void local_navigation_body(float os) {
        const dtr_volume& vol = det.volumes[tk.volume];
        for (int i = 0; i < kCache; ++i) c.p(i) = kMax;
        tk.next = 0;
        tk.last = -1;
        bf_search(vol, os);
        if (vol.grid >= 0) grid_search(vol, os);
        update_status();
        if (tk.trust != kFull && tk.status == kOnPortal)
            tk.trust = (c.link(tk.next - 1) == kEndOfWorld) ? kFull : kNoTrust;
    }

    DTR_AI void apply_local_navigation(uint32_t r) {
        tk.next = static_cast<int8_t>(r & 0xffu);
        tk.last = static_cast<int8_t>((r >> 8) & 0xffu);
        tk.status = static_cast<int8_t>((r >> 16) & 0xffu);
        tk.trust = static_cast<uint8_t>(r >> 24);
    }

    // caching_navigator.hpp:250-388 up to its local navigation: returns true
    // when the cache is not trusted at all and the volume must be searched.
    DTR_AI bool update_impl_head() {
        if (tk.trust == kHigh) {
            if (!update_candidate(tk.next)) {
                tk.status = kUnknown;
                set_fair_trust();
            } else {
                update_status();
                if (tk.status == kTowards || tk.status == kOnPortal) return false;
                if (tk.trust == kFull && update_candidate(tk.next)) return false;
                set_fair_trust();
            }
        }
        if (tk.trust == kFair && !exhausted()) {
            const int b = begin_idx(), e = tk.last + 1;
            for (int i = b; i < e; ++i) {
                if (!update_candidate(i)) c.p(i) = kMax;
            }
            for (int i = b; i < e; ++i) {
                for (int j = i + 1; j < e; ++j) {
                    if (fabs(c.p(j)) < fabs(c.p(i))) {
                        const Cand tmp = c.get(i);
                        c.copy(i, j);
                        c.set(j, tmp);
                    }
                }
            }
            tk.next = static_cast<int8_t>(b);
            int inv = kCache;
            for (int i = kCache - 1; i >= 0; --i) {
                if (c.p(i) == kMax) inv = i;
            }
            tk.last = static_cast<int8_t>(inv - 1);
            update_status();
            if (exhausted()) tk.trust = kNoTrust;
        }
        return tk.trust == kNoTrust;
    }

    // navigator_base.hpp:97-154 (update) as a resumable cascade.  Each call
    // returns true when the lane needs a local navigation of tk.volume with
    // overstep tolerance `os`; once it has been applied
    // (apply_local_navigation) the cascade continues with nav_resume().
    // `ret` is the value update() returns (whether the navigation was
    // re-initialised).  Stages: 1 after update_impl's search, 2 after the
    // volume switch's search, 3 after init_loose_cfg's search.
    DTR_AI bool nav_begin(uint8_t& stage, bool& ret, float& os) {
        ret = false;
        stage = 0;
        if (tk.trust == kFull) return false;
        if (update_impl_head()) {
            ret = true;
            stage = 1;
            os = cfg.overstep_tol;
            return true;
        }
        return nav_portal(stage, ret, os);
    }

    // volume_switch (navigation_functions.hpp:265-290)
    DTR_AI bool nav_portal(uint8_t& stage, bool& ret, float& os) {
        if (tk.status == kOnPortal) {
            const uint16_t link = c.link(tk.next - 1);
            if (link == kEndOfWorld) {
                tk.status = kExit;
                ret = false;
                stage = 0;
                return false;
            }
            if (link != tk.volume) tk.trust = kNoTrust;
            tk.volume = link;
            stage = 2;
            os = cfg.overstep_tol;
            return true;
        }
        return nav_loose(stage, ret, os);
    }

    // init_loose_cfg (navigation_functions.hpp:305-330)
    DTR_AI bool nav_loose(uint8_t& stage, bool& ret, float& os) {
        if (tk.trust != kFull || exhausted()) {
            ret = true;
            stage = 3;
            const float a = 100.f * cfg.overstep_tol, b = -10.f * cfg.max_mask_tol;
            os = a < b ? a : b;
            return true;
        }
        stage = 0;
        return false;
    }

    DTR_AI bool nav_resume(uint8_t& stage, bool& ret, float& os) {
        if (stage == 1) return nav_portal(stage, ret, os);
        if (stage == 2) {
            tk.trust = kFull;
            ret = true;
            if (!(tk.status > kUnknown)) {
                ret = false;
                stage = 0;
                return false;
            }
            return nav_loose(stage, ret, os);
        }
        if (tk.trust != kFull || exhausted()) tk.status = kAbort;
        stage = 0;
        return false;
    }

    // ----------------------------------------------------------- stepping --
    // The current volume's homogeneous material (WithVolumeMaterial only)
    DTR_AI const dtr_material_slab* vol_mat() const {
        if constexpr (P::kVolMat) return volume_material(det.volumes, det.materials, tk.volume);
        return nullptr;
    }

    // rk_stepper.hpp:273-530: stages and field in S, free state, path length
    // and the error estimate in FP64.  In volume material (WithVolumeMaterial)
    // the stage q/p follow the energy loss, evaluated in FP64
    // (evaluate_dqopds, rk_stepper.hpp:534-565).
    DTR_AI void step(double dist, bool do_reset) {
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
        const T3<S> dir = tcast<S>(tk.dir);
        const S qop = static_cast<S>(tk.qop);
        const T3<S> b1 = field_at_s<S>(field, pos);
        const T3<S> k1 = tmul(qop, tcross(dir, b1));
        const V3 k1d = tdbl(k1);
        T3<S> b2, b3, t2, t3, t4, k2, k3, k4;
        // stage q/p (FP64 and S) and d(qop)/ds in volume material
        [[maybe_unused]] const dtr_material_slab* vm = vol_mat();
        [[maybe_unused]] double qd[4] = {tk.qop, tk.qop, tk.qop, tk.qop}, dq[4] = {0.0, 0.0, 0.0, 0.0};
        S qs[4] = {qop, qop, qop, qop};
        if constexpr (P::kVolMat) {
            if (vm) dq[0] = eloss_dqopds(vm, tk.qop, kMuonMass, track_charge(tk.qop));
        }
        double error = 1e20;
        for (uint32_t i = 0; i < cfg.max_rk_updates; ++i) {
            tk.n_trials++;
            const double h = tk.step_size;
            tk.h_stage = h;
            const double h2 = h * h;
            const double half_h = h * 0.5;
            const V3 pos1 = add(add(pos, mul(half_h, tk.dir)), mul(h2 * 0.125, k1d));
            b2 = field_at_s<S>(field, pos1);
            const S hh = static_cast<S>(half_h), hs = static_cast<S>(h);
            if constexpr (P::kVolMat) {
                if (vm) {
                    const double q = track_charge(tk.qop);
                    if (cfg.use_mean_loss) qd[1] = tk.qop + half_h * dq[0];
                    dq[1] = eloss_dqopds(vm, qd[1], kMuonMass, q);
                    if (cfg.use_mean_loss) qd[2] = tk.qop + half_h * dq[1];
                    dq[2] = eloss_dqopds(vm, qd[2], kMuonMass, q);
                    if (cfg.use_mean_loss) qd[3] = tk.qop + h * dq[2];
                    dq[3] = eloss_dqopds(vm, qd[3], kMuonMass, q);
                    for (int n = 1; n < 4; ++n) qs[n] = static_cast<S>(qd[n]);
                }
            }
            t2 = (half_h == 0.0) ? dir : tadd(dir, tmul(hh, k1));
            k2 = tmul(qs[1], tcross(t2, b2));
            t3 = (half_h == 0.0) ? dir : tadd(dir, tmul(hh, k2));
            k3 = tmul(qs[2], tcross(t3, b2));
            const V3 pos2 = add(add(pos, mul(h, tk.dir)), mul(h2 * 0.5, tdbl(k3)));
            b3 = field_at_s<S>(field, pos2);
            t4 = (h == 0.0) ? dir : tadd(dir, tmul(hs, k3));
            k4 = tmul(qs[3], tcross(t4, b3));
            const V3 d = add(sub(sub(k1d, tdbl(k2)), tdbl(k3)), tdbl(k4));
            const V3 err = mul((1.0 / 6.0) * h2, d);
            error = dmax(norm(err), 1e-20);
            if (error <= 4.0 * cfg.rk_error_tol) break;
            tk.step_size = tk.step_size * dmin(dmax(sqrt(sqrt(cfg.rk_error_tol / error)), 0.25), 4.0);
        }
        tk.dtds3 = k4;
        if constexpr (P::kVolMat) tk.dqopds3 = dq[3];
        const double cmin = dmin(dmin(kMax, fabs(cfg.step_constraint)), tk.constraint_abort);
        const double max_step = (tk.step_size >= 0.0 ? 1.0 : -1.0) * cmin;
        if (fabs(tk.step_size) > fabs(max_step)) tk.step_size = max_step;
        if (fabs(tk.step_size) < min_step) tk.step_size = copysign(min_step, tk.step_size);
        tk.next_step_size = tk.step_size * dmin(dmax(sqrt(sqrt(cfg.rk_error_tol / error)), 0.25), 4.0);
        if (fabs(tk.next_step_size) < cfg.min_stepsize)
            tk.next_step_size = copysign(cfg.min_stepsize, tk.next_step_size);
        // advance_track (rk_stepper.hpp:594-629)
        const double h = tk.step_size;
        const double h_6 = h * (1.0 / 6.0);
        [[maybe_unused]] double qn = tk.qop;
        if constexpr (P::kVolMat) {
            if (vm) {
                qn = tk.qop + h_6 * (dq[0] + 2.0 * (dq[1] + dq[2]) + dq[3]);
                // brought to rest within the step: not taken, the track stops
                if (!(fabs(qn) <= kMax && qn * tk.qop > 0.0)) {
                    tk.stopped = true;
                    return;
                }
            }
        }
        const V3 ksum = add(add(k1d, tdbl(k2)), tdbl(k3));
        tk.pos = add(pos, mul(h, add(tk.dir, mul(h_6, ksum))));
        const V3 kd = add(add(k1d, mul(2.0, add(tdbl(k2), tdbl(k3)))), tdbl(k4));
        const V3 nd = add(tk.dir, mul(h_6, kd));
        const double nn = norm(nd);
        tk.dir = {nd.x / nn, nd.y / nn, nd.z / nn};
        if constexpr (P::kVolMat) tk.qop = qn;
        tk.path_length += h;
        tk.abs_path_length += fabs(h);
        if (cfg.do_cov) {
            if constexpr (P::kGrad) advance_jacobian_grad(dir, k1, k3, t2, t3, t4, b1, b2, b3, qop);
            else advance_jacobian(dir, t2, t3, t4, b1, b2, b3, qs, qd, vm);
        }
        tk.n_steps++;
    }

    // rk_stepper.hpp:633-956 (no field gradient), applied column by column
    // on the reduced Jacobian.  q: stage q/p in S (qd in FP64); with
    // use_eloss_gradient in volume material, dqop_n/dqop1 and d(qop)/d(qop1)
    // follow d(d(qop)/ds)/dqop (rk_stepper.hpp:767-800), 1 otherwise.
    DTR_AI void advance_jacobian(T3<S> t1, T3<S> t2, T3<S> t3, T3<S> t4, T3<S> b1, T3<S> b2,
                                 T3<S> b3, const S q[4], [[maybe_unused]] const double qd[4],
                                 [[maybe_unused]] const dtr_material_slab* vm) {
        const S h = static_cast<S>(tk.step_size);
        const S half_h = h * S(0.5);
        const S h_6 = h * static_cast<S>(1.f / 6.f);
        T3<S> c0 = tcross(t1, b1), c1 = tcross(t2, b2), c2 = tcross(t3, b2), c3 = tcross(t4, b3);
        [[maybe_unused]] double dqopqop = 1.0;
        if constexpr (P::kVolMat) {
            if (vm && cfg.use_eloss_gradient) {
                const double hd = tk.step_size, hh = hd * 0.5, h6 = hd * static_cast<double>(1.f / 6.f);
                const double q = track_charge(tk.qop);
                const double e1 = eloss_d2qopdsdqop(vm, qd[0], kMuonMass, q);
                const double n1 = 1.0 + hh * e1;
                const double e2 = eloss_d2qopdsdqop(vm, qd[1], kMuonMass, q) * n1;
                const double n2 = 1.0 + hh * e2;
                const double e3 = eloss_d2qopdsdqop(vm, qd[2], kMuonMass, q) * n2;
                const double n3 = 1.0 + hd * e3;
                const double e4 = eloss_d2qopdsdqop(vm, qd[3], kMuonMass, q) * n3;
                dqopqop = 1.0 + h6 * (e1 + 2.0 * (e2 + e3) + e4);
                c1 = tmul(static_cast<S>(n1), c1);
                c2 = tmul(static_cast<S>(n2), c2);
                c3 = tmul(static_cast<S>(n3), c3);
            }
        }
        const T3<S> dk0 = c0;
        const T3<S> dk1 = tadd(c1, tmul(q[1] * half_h, tcross(dk0, b2)));
        const T3<S> dk2 = tadd(c2, tmul(q[2] * half_h, tcross(dk1, b2)));
        const T3<S> dk3 = tadd(c3, tmul(q[3] * h, tcross(dk2, b3)));
        const T3<S> dFdqop = tmul(h * h_6, tadd(tadd(dk0, dk1), dk2));
        const T3<S> dGdqop = tmul(h_6, tadd(tadd(dk0, tmul(S(2), tadd(dk1, dk2))), dk3));
        // dk_n/dt1 applied to the columns 4..6 of a unit vector e_j:
        // column j of dFdt / dGdt
        S dFdt[9], dGdt[9];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T3<S> e{j == 0 ? S(1) : S(0), j == 1 ? S(1) : S(0), j == 2 ? S(1) : S(0)};
            const T3<S> d0 = tmul(q[0], tcross(e, b1));
            const T3<S> d1 = tmul(q[1], tcross(tadd(e, tmul(half_h, d0)), b2));
            const T3<S> d2 = tmul(q[2], tcross(tadd(e, tmul(half_h, d1)), b2));
            const T3<S> d3 = tmul(q[3], tcross(tadd(e, tmul(h, d2)), b3));
            const T3<S> f = tmul(h, tadd(e, tmul(h_6, tadd(tadd(d0, d1), d2))));
            const T3<S> gg = tadd(e, tmul(h_6, tadd(tadd(d0, tmul(S(2), tadd(d1, d2))), d3)));
            dFdt[0 * 3 + j] = f.x;
            dFdt[1 * 3 + j] = f.y;
            dFdt[2 * 3 + j] = f.z;
            dGdt[0 * 3 + j] = gg.x;
            dGdt[1 * 3 + j] = gg.y;
            dGdt[2 * 3 + j] = gg.z;
        }
        const S dfq[3] = {dFdqop.x, dFdqop.y, dFdqop.z};
        const S dgq[3] = {dGdqop.x, dGdqop.y, dGdqop.z};
        const S jq = tk.J[kJqq];
#pragma unroll
        for (int cc = 4; cc < 8; ++cc) {
            const S jd0 = tk.J[jdir(0, cc)], jd1 = tk.J[jdir(1, cc)], jd2 = tk.J[jdir(2, cc)];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                S s = S(0);
                s += tk.J[jpos(r, cc)];
                s += dFdt[r * 3 + 0] * jd0;
                s += dFdt[r * 3 + 1] * jd1;
                s += dFdt[r * 3 + 2] * jd2;
                if (cc == 7) s += dfq[r] * jq;
                tk.J[jpos(r, cc)] = s;
            }
            S g[3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                S v = S(0);
                v += dGdt[r * 3 + 0] * jd0;
                v += dGdt[r * 3 + 1] * jd1;
                v += dGdt[r * 3 + 2] * jd2;
                if (cc == 7) v += dgq[r] * jq;
                g[r] = v;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) tk.J[jdir(r, cc)] = g[r];
        }
        if constexpr (P::kVolMat) tk.J[kJqq] = static_cast<S>(dqopqop) * jq;
    }

    // Field gradient by central differences over +-0.1 mm (rk_stepper.hpp:
    // 958-998), G[r][i] = dB_r / dx_i, in the stage precision.
    DTR_AI void field_gradient_fd(V3 p, S G[3][3]) const {
        const double delta = static_cast<double>(1e-1f);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            V3 p1 = p, p2 = p;
            if (i == 0) { p1.x += delta; p2.x -= delta; }
            if (i == 1) { p1.y += delta; p2.y -= delta; }
            if (i == 2) { p1.z += delta; p2.z -= delta; }
            const T3<S> b1 = field_at_s<S>(field, p1), b2 = field_at_s<S>(field, p2);
            const S inv = static_cast<S>(1.0 / (2.0 * delta));
            G[0][i] = (b1.x - b2.x) * inv;
            G[1][i] = (b1.y - b2.y) * inv;
            G[2][i] = (b1.z - b2.z) * inv;
        }
    }

    // advance_jacobian with the field gradient (rk_stepper.hpp:633-956,
    // use_field_gradient): the full update J <- [[dF/dr, dF/dt, dF/dqop],
    // [dG/dr, dG/dt, dG/dqop], [0, 0, 1]] J over the position, direction and
    // q/p columns.  As in the reference, the gradient enters dk/dr only, and
    // its points are taken around the end point of the step (the reference
    // evaluates them after advance_track).
    DTR_AI void advance_jacobian_grad(T3<S> t1, T3<S> k1, T3<S> k3, T3<S> t2, T3<S> t3, T3<S> t4, T3<S> b1,
                                      T3<S> b2, T3<S> b3, S qop) {
        const double hd = tk.step_size;
        const S h = static_cast<S>(hd);
        const S half_h = h * S(0.5);
        const S h_6 = h * static_cast<S>(1.f / 6.f);
        const T3<S> dk0 = tcross(t1, b1);
        const T3<S> dk1 = tadd(tcross(t2, b2), tmul(qop * half_h, tcross(dk0, b2)));
        const T3<S> dk2 = tadd(tcross(t3, b2), tmul(qop * half_h, tcross(dk1, b2)));
        const T3<S> dk3 = tadd(tcross(t4, b3), tmul(qop * h, tcross(dk2, b3)));
        const T3<S> dFdqop = tmul(h * h_6, tadd(tadd(dk0, dk1), dk2));
        const T3<S> dGdqop = tmul(h_6, tadd(tadd(dk0, tmul(S(2), tadd(dk1, dk2))), dk3));
        // the 6x7 block [dF/dr dF/dt dF/dqop; dG/dr dG/dt dG/dqop] by columns
        S D[6][7];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T3<S> e{j == 0 ? S(1) : S(0), j == 1 ? S(1) : S(0), j == 2 ? S(1) : S(0)};
            const T3<S> d0 = tmul(qop, tcross(e, b1));
            const T3<S> d1 = tmul(qop, tcross(tadd(e, tmul(half_h, d0)), b2));
            const T3<S> d2 = tmul(qop, tcross(tadd(e, tmul(half_h, d1)), b2));
            const T3<S> d3 = tmul(qop, tcross(tadd(e, tmul(h, d2)), b3));
            const T3<S> f = tmul(h, tadd(e, tmul(h_6, tadd(tadd(d0, d1), d2))));
            const T3<S> g = tadd(e, tmul(h_6, tadd(tadd(d0, tmul(S(2), tadd(d1, d2))), d3)));
            D[0][3 + j] = f.x; D[1][3 + j] = f.y; D[2][3 + j] = f.z;
            D[3][3 + j] = g.x; D[4][3 + j] = g.y; D[5][3 + j] = g.z;
        }
        D[0][6] = dFdqop.x; D[1][6] = dFdqop.y; D[2][6] = dFdqop.z;
        D[3][6] = dGdqop.x; D[4][6] = dGdqop.y; D[5][6] = dGdqop.z;
        {
            const V3 r_ini = tk.pos;
            const V3 t0d = tdbl(t1), k1d = tdbl(k1), k3d = tdbl(k3);
            const double h2 = hd * hd;
            const V3 r_mid = add(add(r_ini, mul(hd * 0.5, t0d)), mul(h2 * 0.125, k1d));
            const V3 r_fin = add(add(r_ini, mul(hd, t0d)), mul(h2 * 0.5, k3d));
            S Gi[3][3], Gm[3][3], Gf[3][3];
            field_gradient_fd(r_ini, Gi);
            field_gradient_fd(r_mid, Gm);
            field_gradient_fd(r_fin, Gf);
            const S hh2 = static_cast<S>(h2);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // column j of dk_n/dr1
                const T3<S> gi{Gi[0][j], Gi[1][j], Gi[2][j]};
                const T3<S> r0 = tmul(qop, tcross(t1, gi));  // -q (dB/dr)_j x t = q t x (dB/dr)_j
                // dr_mid/dr1 column j = e_j + h^2/8 r0; dB_mid/dr1 = Gm (e_j + h^2/8 r0)
                const T3<S> cm{(j == 0 ? S(1) : S(0)) + hh2 * S(0.125) * r0.x,
                               (j == 1 ? S(1) : S(0)) + hh2 * S(0.125) * r0.y,
                               (j == 2 ? S(1) : S(0)) + hh2 * S(0.125) * r0.z};
                const T3<S> gm{Gm[0][0] * cm.x + Gm[0][1] * cm.y + Gm[0][2] * cm.z,
                               Gm[1][0] * cm.x + Gm[1][1] * cm.y + Gm[1][2] * cm.z,
                               Gm[2][0] * cm.x + Gm[2][1] * cm.y + Gm[2][2] * cm.z};
                const T3<S> r1 = tadd(tmul(qop, tcross(tmul(half_h, r0), b2)), tmul(qop, tcross(t2, gm)));
                const T3<S> r2 = tadd(tmul(qop, tcross(tmul(half_h, r1), b2)), tmul(qop, tcross(t3, gm)));
                const T3<S> cf{(j == 0 ? S(1) : S(0)) + hh2 * S(0.5) * r2.x,
                               (j == 1 ? S(1) : S(0)) + hh2 * S(0.5) * r2.y,
                               (j == 2 ? S(1) : S(0)) + hh2 * S(0.5) * r2.z};
                const T3<S> gf{Gf[0][0] * cf.x + Gf[0][1] * cf.y + Gf[0][2] * cf.z,
                               Gf[1][0] * cf.x + Gf[1][1] * cf.y + Gf[1][2] * cf.z,
                               Gf[2][0] * cf.x + Gf[2][1] * cf.y + Gf[2][2] * cf.z};
                const T3<S> r3 = tadd(tmul(qop, tcross(tmul(h, r2), b3)), tmul(qop, tcross(t4, gf)));
                const T3<S> e{j == 0 ? S(1) : S(0), j == 1 ? S(1) : S(0), j == 2 ? S(1) : S(0)};
                const T3<S> f = tadd(e, tmul(h * h_6, tadd(tadd(r0, r1), r2)));
                const T3<S> g = tmul(h_6, tadd(tadd(r0, tmul(S(2), tadd(r1, r2))), r3));
                D[0][j] = f.x; D[1][j] = f.y; D[2][j] = f.z;
                D[3][j] = g.x; D[4][j] = g.y; D[5][j] = g.z;
            }
        }
        // new columns: 0-2 (position, Jx) and 4-7 (direction, q/p, J)
        const S jq = tk.J[kJqq];
        S nx[18];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            S old[6];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                old[r] = tk.Jx.v[r * 3 + cc];
                old[3 + r] = tk.Jx.v[9 + r * 3 + cc];
            }
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                S s2 = S(0);
#pragma unroll
                for (int m = 0; m < 6; ++m) s2 += D[r][m] * old[m];
                nx[(r < 3 ? 0 : 9) + (r % 3) * 3 + cc] = s2;
            }
        }
#pragma unroll
        for (int cc = 4; cc < 8; ++cc) {
            S old[6];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                old[r] = tk.J[jpos(r, cc)];
                old[3 + r] = tk.J[jdir(r, cc)];
            }
            const S q = (cc == 7) ? jq : S(0);
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                S s2 = S(0);
#pragma unroll
                for (int m = 0; m < 6; ++m) s2 += D[r][m] * old[m];
                s2 += D[r][6] * q;
                if (r < 3) tk.J[jpos(r, cc)] = s2;
                else tk.J[jdir(r - 3, cc)] = s2;
            }
        }
#pragma unroll
        for (int i = 0; i < 18; ++i) tk.Jx.v[i] = nx[i];
    }

    // policies.hpp:99-139
    DTR_AI void rk_policy() {
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

    // --------------------------------------------------------------- actors --
    DTR_AI bool on_sensitive() const {
        return tk.status == kOnObject && det.sf[cur_sf()].sf_id == DTR_SENSITIVE;
    }
    DTR_AI bool on_material() const {
        return on_surface() && det.sf[cur_sf()].material != kInvalid;
    }
    DTR_AI void reset_J() {
#pragma unroll
        for (int i = 0; i < 25; ++i) tk.J[i] = S(0);
        tk.J[jdir(0, 4)] = S(1);
        tk.J[jdir(1, 5)] = S(1);
        tk.J[jdir(2, 6)] = S(1);
        tk.J[kJqq] = S(1);
        if constexpr (P::kGrad) {
#pragma unroll
            for (int i = 0; i < 18; ++i) tk.Jx.v[i] = (i < 9 && i % 4 == 0) ? S(1) : S(0);
        }
    }
    DTR_AI void free_to_bound(uint32_t s, double out[6]) const {
        const NavSurface& rec = det.sf[s];
        if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            out[0] = phi_of(tk.pos);
            out[1] = tk.pos.z;
        } else if (P::kExt && is_line(rec.shape)) {
            Engine::line_local(rec_trf(rec), tk.pos, tk.dir, out[0], out[1]);
        } else if (P::kExt && rec.shape == DTR_CYLINDER2D) {
            const V3 l = rec_to_local(rec, tk.pos);
            out[0] = perp(l) * phi_of(l);
            out[1] = l.z;
        } else if (rec.shape == DTR_RING2D || (P::kExt && rec.shape == DTR_ANNULUS2D)) {
            const V3 l = rec_to_local(rec, tk.pos);
            out[0] = perp(l);
            out[1] = phi_of(l);
        } else {
            const V3 l = rec_to_local(rec, tk.pos);
            out[0] = l.x;
            out[1] = l.y;
        }
        out[2] = phi_of(tk.dir);
        out[3] = theta_of(tk.dir);
        out[4] = tk.qop;
        out[5] = tk.time;
    }
    DTR_AI V3 bound_pos(uint32_t s, const double b[6]) const {
        const NavSurface& rec = det.sf[s];
        if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const double r = rec.m[0];
            return {r * cos(b[0]), r * sin(b[0]), b[1]};
        }
        if constexpr (P::kExt) {
            if (is_line(rec.shape)) {
                // line2D.hpp:80-99 with the track direction
                const V3 r = cross(rec_axis(rec, 2), Engine::bound_dir(b));
                return add(rec_to_global(rec, {0.0, 0.0, b[1]}), mul(b[0], normalize(r)));
            }
            if (rec.shape == DTR_CYLINDER2D) {
                const double r = rec.m[0];
                const double ph = b[0] / r;
                return rec_to_global(rec, {r * cos(ph), r * sin(ph), b[1]});
            }
        }
        if (rec.shape == DTR_RING2D || (P::kExt && rec.shape == DTR_ANNULUS2D))
            return rec_to_global(rec, {b[0] * cos(b[1]), b[0] * sin(b[1]), 0.0});
        return rec_to_global(rec, {b[0], b[1], 0.0});
    }
    DTR_AI T3<S> dtds() const {
        if (tk.path_length == 0.0) {
            const T3<S> b = field_at_s<S>(field, tk.pos);
            return tmul(static_cast<S>(tk.qop), tcross(tcast<S>(tk.dir), b));
        }
        return tk.dtds3;
    }

    // normal of a placed cylinder at the bound position (cylindrical2D.hpp:62-70:
    // phi = loc0 / r of the mask)
    DTR_AI V3 cyl2_normal(const NavSurface& rec) const {
        const double ph = tk.bound[0] / rec.m[0];
        return vec_to_global(rec_trf(rec), {cos(ph), sin(ph), 0.0});
    }

    // Bound-to-free of the extended frames at the departure surface:
    // dpos/dloc (A, 3x2 row-major) and, for lines, dpos/dangle (Pa):
    // jacobian_line.hpp:102-172, jacobian_cylindrical.hpp, jacobian_polar.hpp:44-72
    DTR_AI void ext_bound_to_free(uint32_t dep, S A[6], S Pa[3][2]) const {
        const NavSurface& rec = det.sf[dep];
        const dtr_transform t = rec_trf(rec);
        V3 c0, c1;
        if (is_line(rec.shape)) {
            const V3 ddir = Engine::bound_dir(tk.bound);
            const V3 dp = bound_pos(dep, tk.bound);
            V3 nx, ny, nz;
            Engine::line_frame(t, ddir, nx, ny, nz);
            c0 = nx;
            c1 = ny;
            double l0, l1;
            Engine::line_local(t, dp, ddir, l0, l1);
            const double ct = cos(tk.bound[3]), stt = sin(tk.bound[3]);
            const double cp = cos(tk.bound[2]), sp = sin(tk.bound[2]);
            const V3 yp = cross(ny, {-stt * sp, stt * cp, 0.0});
            const V3 yt = cross(ny, {ct * cp, ct * sp, -stt});
            const double C = 1.0 / dot(ddir, nz) * l0;
            const V3 dphi = mul(C, sub(yp, mul(dot(nx, yp), nx)));
            const V3 dtheta = mul(C, sub(yt, mul(dot(nx, yt), nx)));
            Pa[0][0] = static_cast<S>(dphi.x); Pa[1][0] = static_cast<S>(dphi.y); Pa[2][0] = static_cast<S>(dphi.z);
            Pa[0][1] = static_cast<S>(dtheta.x); Pa[1][1] = static_cast<S>(dtheta.y); Pa[2][1] = static_cast<S>(dtheta.z);
        } else if (rec.shape == DTR_CYLINDER2D) {
            V3 nz;
            Engine::cyl2_frame(t, bound_pos(dep, tk.bound), c0, c1, nz);
        } else {
            const V3 l = rec_to_local(rec, bound_pos(dep, tk.bound));
            const double lr = perp(l), lp = phi_of(l), c = cos(lp), sn = sin(lp);
            const V3 x = rec_axis(rec, 0), y = rec_axis(rec, 1);
            c0 = add(mul(c, x), mul(sn, y));
            c1 = mul(lr, sub(mul(c, y), mul(sn, x)));
        }
        A[0] = static_cast<S>(c0.x); A[1] = static_cast<S>(c1.x);
        A[2] = static_cast<S>(c0.y); A[3] = static_cast<S>(c1.y);
        A[4] = static_cast<S>(c0.z); A[5] = static_cast<S>(c1.z);
    }

    // Free-to-bound of the extended frames at the destination: rows 0-1 of
    // the position block (Bl), and the path correction.  Returns f2p itself
    // for lines (with its direction part in f2d, jacobian_line.hpp:64-92 and
    // 174-182), the surface normal otherwise (cylindrical and polar frames,
    // jacobian_cylindrical.hpp, jacobian_polar.hpp:81-110).
    DTR_AI V3 ext_free_to_bound(uint32_t dest, T3<S> k, S Bl[6], T3<S>& f2d) const {
        const NavSurface& rec = det.sf[dest];
        const dtr_transform t = rec_trf(rec);
        V3 r0, r1, out;
        if (is_line(rec.shape)) {
            const V3 kd = tdbl(k);
            const V3 pc = sub(tk.pos, {rec.t[0], rec.t[1], rec.t[2]});
            const V3 lz = rec_axis(rec, 2);
            const double pz = dot(pc, lz);
            const double dz = dot(lz, tk.dir);
            const V3 pc_x = sub(pc, mul(pz, lz));
            const double nrm = -1.0 / (1.0 - dz * dz + dot(pc_x, kd));
            out = mul(nrm, sub(tk.dir, mul(dz, lz)));
            f2d = tcast<S>(mul(nrm, pc_x));
            V3 nz;
            Engine::line_frame(t, tk.dir, r0, r1, nz);
        } else if (rec.shape == DTR_CYLINDER2D) {
            Engine::cyl2_frame(t, tk.pos, r0, r1, out);
        } else {
            const V3 l = rec_to_local(rec, tk.pos);
            const double lr = perp(l), lp = phi_of(l), c = cos(lp), sn = sin(lp);
            const V3 x = rec_axis(rec, 0), y = rec_axis(rec, 1);
            r0 = add(mul(c, x), mul(sn, y));
            r1 = mul(1.0 / lr, sub(mul(c, y), mul(sn, x)));
            out = rec_axis(rec, 2);
        }
        Bl[0] = static_cast<S>(r0.x); Bl[1] = static_cast<S>(r0.y); Bl[2] = static_cast<S>(r0.z);
        Bl[3] = static_cast<S>(r1.x); Bl[4] = static_cast<S>(r1.y); Bl[5] = static_cast<S>(r1.z);
        return out;
    }

    // parameter_updater.hpp:422-572 + covariance_transport.hpp in S
    DTR_AI void transport_covariance(uint32_t dest) {
        const uint32_t dep = tk.up_sf;
        const NavSurface& drec = det.sf[dep];
        S A[6];
        S Pa[3][2] = {{S(0), S(0)}, {S(0), S(0)}, {S(0), S(0)}};  // dpos/dangle of a line
        if (drec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            const V3 dp = bound_pos(dep, tk.bound);
            const double r = perp(dp), ph = phi_of(dp);
            A[0] = static_cast<S>(-r * sin(ph)); A[1] = S(0);
            A[2] = static_cast<S>(r * cos(ph));  A[3] = S(0);
            A[4] = S(0);                         A[5] = S(1);
        } else if (P::kExt && (drec.shape >= DTR_LINE_SQUARE || drec.shape == DTR_RING2D)) {
            ext_bound_to_free(dep, A, Pa);
        } else {
            A[0] = static_cast<S>(drec.r[0]); A[1] = static_cast<S>(drec.r[1]);
            A[2] = static_cast<S>(drec.r[3]); A[3] = static_cast<S>(drec.r[4]);
            A[4] = static_cast<S>(drec.r[6]); A[5] = static_cast<S>(drec.r[7]);
        }
        S Cd[6];
        {
            const double th = tk.bound[3], ph = tk.bound[2];
            const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
            Cd[0] = static_cast<S>(-st * sp); Cd[1] = static_cast<S>(ct * cp);
            Cd[2] = static_cast<S>(st * cp);  Cd[3] = static_cast<S>(ct * sp);
            Cd[4] = S(0);                     Cd[5] = static_cast<S>(-st);
        }
        S Mp[3][5], Md[3][5];  // column 5 (time) is zero for both blocks
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if constexpr (P::kGrad) {
                Mp[r][0] = tk.Jx.v[r * 3 + 0] * A[0] + tk.Jx.v[r * 3 + 1] * A[2] + tk.Jx.v[r * 3 + 2] * A[4];
                Mp[r][1] = tk.Jx.v[r * 3 + 0] * A[1] + tk.Jx.v[r * 3 + 1] * A[3] + tk.Jx.v[r * 3 + 2] * A[5];
            } else {
                Mp[r][0] = A[r * 2 + 0];
                Mp[r][1] = A[r * 2 + 1];
            }
            Mp[r][2] = tk.J[jpos(r, 4)] * Cd[0] + tk.J[jpos(r, 5)] * Cd[2] + tk.J[jpos(r, 6)] * Cd[4];
            Mp[r][3] = tk.J[jpos(r, 4)] * Cd[1] + tk.J[jpos(r, 5)] * Cd[3] + tk.J[jpos(r, 6)] * Cd[5];
            if constexpr (P::kExt && P::kGrad) {
                // the line's dpos/dangle through the position columns
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    Mp[r][2] += tk.Jx.v[r * 3 + m] * Pa[m][0];
                    Mp[r][3] += tk.Jx.v[r * 3 + m] * Pa[m][1];
                }
            } else if constexpr (P::kExt) {
                Mp[r][2] += Pa[r][0];
                Mp[r][3] += Pa[r][1];
            }
            Mp[r][4] = tk.J[jpos(r, 7)];
            if constexpr (P::kGrad) {
                Md[r][0] = tk.Jx.v[9 + r * 3 + 0] * A[0] + tk.Jx.v[9 + r * 3 + 1] * A[2] + tk.Jx.v[9 + r * 3 + 2] * A[4];
                Md[r][1] = tk.Jx.v[9 + r * 3 + 0] * A[1] + tk.Jx.v[9 + r * 3 + 1] * A[3] + tk.Jx.v[9 + r * 3 + 2] * A[5];
            } else {
                Md[r][0] = S(0);
                Md[r][1] = S(0);
            }
            Md[r][2] = tk.J[jdir(r, 4)] * Cd[0] + tk.J[jdir(r, 5)] * Cd[2] + tk.J[jdir(r, 6)] * Cd[4];
            Md[r][3] = tk.J[jdir(r, 4)] * Cd[1] + tk.J[jdir(r, 5)] * Cd[3] + tk.J[jdir(r, 6)] * Cd[5];
            if constexpr (P::kExt && P::kGrad) {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    Md[r][2] += tk.Jx.v[9 + r * 3 + m] * Pa[m][0];
                    Md[r][3] += tk.Jx.v[9 + r * 3 + m] * Pa[m][1];
                }
            }
            Md[r][4] = tk.J[jdir(r, 7)];
        }
        const S Mq4 = tk.J[kJqq];
        // the q/p row of J B2F and of the path correction with the volume's
        // d(qop)/ds (parameter_updater.hpp:518-526, rk_stepper.hpp:147-155)
        [[maybe_unused]] S Jq[5] = {S(0), S(0), S(0), S(0), Mq4};
        [[maybe_unused]] S dqs = S(0);
        if constexpr (P::kVolMat)
            dqs = static_cast<S>(tk.path_length == 0.0
                                     ? eloss_dqopds(vol_mat(), tk.qop, kMuonMass, track_charge(tk.qop))
                                     : tk.dqopds3);
        const T3<S> k = dtds();
        const NavSurface& erec = det.sf[dest];
        const bool ext_dest = P::kExt && (erec.shape >= DTR_LINE_SQUARE || erec.shape == DTR_RING2D);
        S Bl[6];
        T3<S> f2d{S(0), S(0), S(0)};
        V3 n;
        if (ext_dest) {
            n = ext_free_to_bound(dest, k, Bl, f2d);
        } else if (erec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            if constexpr (P::kMixed) {
                // (cos, sin) of phi of the position, without trigonometry
                const double rinv = 1.0 / perp(tk.pos);
                n = {tk.pos.x * rinv, tk.pos.y * rinv, 0.0};
            } else {
                const double ph = phi_of(tk.pos);
                n = {cos(ph), sin(ph), 0.0};
            }
        } else {
            n = rec_axis(erec, 2);
        }
        const T3<S> f2p = (P::kExt && is_line(erec.shape)) ? tcast<S>(n) : tcast<S>(mul(-1.0 / dot(n, tk.dir), n));
        const S p2f_pos[3] = {static_cast<S>(tk.dir.x), static_cast<S>(tk.dir.y), static_cast<S>(tk.dir.z)};
        const S p2f_dir[3] = {k.x, k.y, k.z};
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) {
            S proj = f2p.x * Mp[0][cc] + f2p.y * Mp[1][cc] + f2p.z * Mp[2][cc];
            if constexpr (P::kExt) proj += f2d.x * Md[0][cc] + f2d.y * Md[1][cc] + f2d.z * Md[2][cc];
            if constexpr (P::kVolMat) Jq[cc] += dqs * proj;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                Mp[r][cc] += p2f_pos[r] * proj;
                Md[r][cc] += p2f_dir[r] * proj;
            }
        }
        if (ext_dest) {
            // rows set by ext_free_to_bound
        } else if (erec.shape == DTR_CONCENTRIC_CYLINDER2D) {
            if constexpr (P::kMixed) {
                const double rinv2 = 1.0 / (tk.pos.x * tk.pos.x + tk.pos.y * tk.pos.y);
                Bl[0] = static_cast<S>(-tk.pos.y * rinv2); Bl[1] = static_cast<S>(tk.pos.x * rinv2); Bl[2] = S(0);
                Bl[3] = S(0);                              Bl[4] = S(0);                             Bl[5] = S(1);
            } else {
                const double rinv = 1.0 / perp(tk.pos), ph = phi_of(tk.pos);
                Bl[0] = static_cast<S>(-rinv * sin(ph)); Bl[1] = static_cast<S>(rinv * cos(ph)); Bl[2] = S(0);
                Bl[3] = S(0);                            Bl[4] = S(0);                           Bl[5] = S(1);
            }
        } else {
            Bl[0] = static_cast<S>(erec.r[0]); Bl[1] = static_cast<S>(erec.r[3]); Bl[2] = static_cast<S>(erec.r[6]);
            Bl[3] = static_cast<S>(erec.r[1]); Bl[4] = static_cast<S>(erec.r[4]); Bl[5] = static_cast<S>(erec.r[7]);
        }
        S Ba[6];
        if constexpr (P::kMixed) {
            // sin(theta) = |d_T|, cos(theta) = d_z, (cos, sin)(phi) = d_T / |d_T|
            const double st = perp(tk.dir);
            const double inv = 1.0 / st;
            Ba[0] = static_cast<S>(-tk.dir.y * inv * inv); Ba[1] = static_cast<S>(tk.dir.x * inv * inv);
            Ba[2] = S(0);
            Ba[3] = static_cast<S>(tk.dir.x * inv * tk.dir.z); Ba[4] = static_cast<S>(tk.dir.y * inv * tk.dir.z);
            Ba[5] = static_cast<S>(-st);
        } else {
            const double th = theta_of(tk.dir), ph = phi_of(tk.dir);
            const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
            Ba[0] = static_cast<S>(-sp / st); Ba[1] = static_cast<S>(cp / st); Ba[2] = S(0);
            Ba[3] = static_cast<S>(cp * ct);  Ba[4] = static_cast<S>(sp * ct); Ba[5] = static_cast<S>(-st);
        }
        // Jf rows 0..3 over columns 0..4; row 4 = Mq4 e4; row 5 = e5
        S Jf[4][5];
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) {
            Jf[0][cc] = Bl[0] * Mp[0][cc] + Bl[1] * Mp[1][cc] + Bl[2] * Mp[2][cc];
            Jf[1][cc] = Bl[3] * Mp[0][cc] + Bl[4] * Mp[1][cc] + Bl[5] * Mp[2][cc];
            Jf[2][cc] = Ba[0] * Md[0][cc] + Ba[1] * Md[1][cc] + Ba[2] * Md[2][cc];
            Jf[3][cc] = Ba[3] * Md[0][cc] + Ba[4] * Md[1][cc] + Ba[5] * Md[2][cc];
        }
        if constexpr (kJacOut) {
            if (jf_out)
            for (int r = 0; r < 6; ++r)
                for (int cc = 0; cc < 6; ++cc) {
                    double v = 0.0;
                    if (r < 4 && cc < 5) v = static_cast<double>(Jf[r][cc]);
                    else if (r == 4 && cc < 5) v = static_cast<double>(P::kVolMat ? Jq[cc] : (cc == 4 ? Mq4 : S(0)));
                    else if (r == 5 && cc == 5) v = 1.0;
                    jf_out[r * 6 + cc] = v;
                }
        }
        // C' = Jf C Jf^T with the block structure of Jf
        auto C = [&](int r, int cc) { return tk.cov[covi(r, cc)]; };
        S T[6][6];  // T = Jf C
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int cc = 0; cc < 6; ++cc) {
                S s = S(0);
#pragma unroll
                for (int m = 0; m < 5; ++m) s += Jf[r][m] * C(m, cc);
                T[r][cc] = s;
            }
        }
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) {
            if constexpr (P::kVolMat) {
                S s = S(0);
#pragma unroll
                for (int m = 0; m < 5; ++m) s += Jq[m] * C(m, cc);
                T[4][cc] = s;
            } else {
                T[4][cc] = Mq4 * C(4, cc);
            }
            T[5][cc] = C(5, cc);
        }
        S NC[21];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
#pragma unroll
            for (int cc = r; cc < 6; ++cc) {
                S s = S(0);
                if (cc < 4) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) s += T[r][m] * Jf[cc][m];
                } else if (cc == 4) {
                    if constexpr (P::kVolMat) {
#pragma unroll
                        for (int m = 0; m < 5; ++m) s += T[r][m] * Jq[m];
                    } else {
                        s = T[r][4] * Mq4;
                    }
                } else {
                    s = T[r][5];
                }
                NC[covi(r, cc)] = s;
            }
        }
#pragma unroll
        for (int i = 0; i < 21; ++i) tk.cov[i] = NC[i];
    }

    // pointwise_material_interactor.hpp:147-326 + setter.  The mean energy
    // loss changes the track state and is evaluated in FP64; the noise terms
    // (Landau sigma of q/p, Highland theta0) only enter the covariance and are
    // evaluated in S.
    DTR_AI void material_interaction(uint32_t dest) {
        const NavSurface& rec = det.sf[dest];
        const dtr_material_slab& m =
            surface_material(det.materials, det.mat_maps, rec.material, tk.bound[0], tk.bound[1]);
        double cos_inc;
        if constexpr (P::kMixed) {
            // incidence from the free direction and position (the bound
            // angles and loc0 were just computed from them)
            V3 nrm;
            if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
                const double rinv = 1.0 / perp(tk.pos);
                nrm = {tk.pos.x * rinv, tk.pos.y * rinv, 0.0};
            } else if (P::kExt && rec.shape == DTR_CYLINDER2D) {
                nrm = cyl2_normal(rec);
            } else {
                nrm = rec_axis(rec, 2);
            }
            cos_inc = fabs(dot(tk.dir, nrm));
        } else {
            V3 nrm;
            if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) nrm = {cos(tk.bound[0]), sin(tk.bound[0]), 0.0};
            else if (P::kExt && rec.shape == DTR_CYLINDER2D) nrm = cyl2_normal(rec);
            else nrm = rec_axis(rec, 2);
            const V3 bd = Engine::bound_dir(tk.bound);
            cos_inc = fabs(dot(bd, nrm));
        }
        if (!(m.thickness > 2.220446049250313e-16) || tk.bound[4] == 0.0) return;
        const double qop = tk.bound[4];
        double path_seg = m.thickness / cos_inc;
        double path_x0 = m.thickness_in_X0 / cos_inc;
        if constexpr (P::kExt) {
            if (m.kind == DTR_MATERIAL_ROD) {
                // material_rod.hpp:62-78, approach = bound loc0
                const double a = tk.bound[0];
                path_seg = 0.0;
                if (!(fabs(a) > m.thickness))
                    path_seg = 2.0 * sqrt((m.thickness * m.thickness - a * a) / (1.0 - cos_inc * cos_inc));
                path_x0 = path_seg / m.X0;
            }
        }
        const double mass = kMuonMass, q = track_charge(tk.qop);
        const double q2b2 = q * q + (mass * qop) * (mass * qop);
        const double mOverP = mass * fabs(qop / q);
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
        double e_loss = 0.0;
        S sigma_qop = S(0), theta0 = S(0);
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
                const S eps = static_cast<S>(eps_pl * path_seg);
                const S fwhm = S(4) * eps;
                const S sigmaE = S(0.5) * static_cast<S>(0.70710678118654752440) * fwhm /
                                 sqrt(static_cast<S>(0.69314718055994530942));
                const S pInv = static_cast<S>(qop / q);
                sigma_qop = sqrt(static_cast<S>(q2b2)) * pInv * pInv * sigmaE;
            }
        }
        if (cfg.do_ms) {
            const S x0 = static_cast<S>(path_x0);
            const S pinv = static_cast<S>(fabs(qop / q));
            if (x0 > S(0)) {
                const S t = sqrt(x0 * static_cast<S>(q2b2));
                theta0 = static_cast<S>(13.6f) * static_cast<S>(1e-3) * pinv * t *
                         (S(1) + static_cast<S>(0.038f) * S(2) * log(t));
            }
        }
        if (cfg.do_eloss) {
            const double massd = kMuonMass, qd = track_charge(tk.qop);
            const double p = qd / tk.bound[4];
            // update_qop (pointwise_material_interactor.hpp:240-262): an
            // energy gain going backward
            const double next_e = sqrt(massd * massd + p * p) - copysign(e_loss, tk.nav_dir);
            const double next_p = (massd < next_e) ? sqrt(next_e * next_e - massd * massd) : 0.0;
            const double next_qop = (qd != 0.0) ? qd / next_p : 1.0 / next_p;
            tk.bound[4] = (next_p == 0.0) ? kMax : next_qop;
            if (next_p == 0.0) tk.stopped = true;
            if (cfg.do_cov) tk.cov[covi(4, 4)] += sigma_qop * sigma_qop;
        }
        if (cfg.do_cov) {
            const double dz = P::kMixed ? tk.dir.z : Engine::bound_dir(tk.bound).z;
            const S var = theta0 * theta0;
            tk.cov[covi(2, 2)] += (dz == 1.0) ? std::numeric_limits<S>::max()
                                              : static_cast<S>(static_cast<double>(var) / (1.0 - dz * dz));
            tk.cov[covi(3, 3)] += var;
        }
        // parameter_setter (parameter_updater.hpp:585-627): the free state
        // is re-derived from the bound parameters, which puts the position
        // onto the surface.  The mixed policy does the same projection
        // without trigonometry: onto the plane (planar masks), onto the
        // radius (cylinders), and renormalises the direction.
        if (P::kExt && (is_line(det.sf[dest].shape) || det.sf[dest].shape == DTR_CYLINDER2D)) {
            tk.pos = bound_pos(dest, tk.bound);
            tk.dir = Engine::bound_dir(tk.bound);
        } else if constexpr (P::kMixed) {
            const NavSurface& r = det.sf[dest];
            if (r.shape == DTR_CONCENTRIC_CYLINDER2D) {
                const double k = r.m[0] / perp(tk.pos);
                tk.pos = {tk.pos.x * k, tk.pos.y * k, tk.pos.z};
            } else {
                const V3 l = rec_to_local(r, tk.pos);
                tk.pos = rec_to_global(r, {l.x, l.y, 0.0});
            }
            const double dn = 1.0 / norm(tk.dir);
            tk.dir = {tk.dir.x * dn, tk.dir.y * dn, tk.dir.z * dn};
        } else {
            tk.pos = bound_pos(dest, tk.bound);
            tk.dir = Engine::bound_dir(tk.bound);
        }
        tk.time = tk.bound[5];
        tk.qop = tk.bound[4];
        if (cfg.estimate_noise) estimate_noise();
    }

    DTR_AI void estimate_noise() {
        const double ns = cfg.n_stddev;
        const double v0 = tk.cov[covi(0, 0)], v1 = tk.cov[covi(1, 1)];
        const double dphi = ns * sqrt(static_cast<double>(tk.cov[covi(2, 2)]));
        const double dth = ns * sqrt(static_cast<double>(tk.cov[covi(3, 3)]));
        const double pe = tk.bound[2] + dphi, te = tk.bound[3] + dth;
        const double ste = sin(te);
        V3 displ{cos(pe) * ste, sin(pe) * ste, cos(te)};
        const double max_tol = 5.0;
        const double path = exhausted() ? max_tol : fabs(c.p(tk.next));
        displ = mul(path, sub(displ, tk.dir));
        const double p = track_charge(tk.qop) / tk.qop;
        const double acc = 1.0 / p * cfg.accumulated_error * tk.path_length;
        const double ext = sqrt(ns * ns * (v0 + v1) + dot(displ, displ) + acc * acc);
        tk.ext_tol = ext > max_tol ? max_tol : ext;
    }

    // step_tracer observer (step_tracer.hpp: step_data): free state on the
    // surface, bound parameters after the material interactor, transport
    // Jacobian expanded to the free 8x8 matrix
    DTR_AI void trace_update(uint32_t dest, const double fp[8], uint64_t n, uint32_t gid) {
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
                for (int c = 0; c < 8; ++c) {
                    double v = free_jacobian(tk.J, r, c);
                    if constexpr (P::kGrad) {
                        // position columns carried by the gradient policy
                        if (c < 3 && r < 3) v = static_cast<double>(tk.Jx.v[r * 3 + c]);
                        if (c < 3 && r >= 4 && r < 7) v = static_cast<double>(tk.Jx.v[9 + (r - 4) * 3 + c]);
                    }
                    t.jac[(r * 8 + c) * M + o] = v;
                }
    }

    // momentum_aborter (aborters.hpp:80-143): pT, then p, at or below the limits
    DTR_AI void momentum_aborter() {
        if (tk.status == kExit) return;
        const double q = track_charge(tk.qop);
        const double pT = fabs(q / tk.qop * perp(tk.dir));
        if (pT <= cfg.min_pT || q / tk.qop <= cfg.min_p) {
            if (tk.status != kAbort) tk.momentum_hit = true;
            tk.status = kAbort;
            tk.heartbeat = false;
        }
    }

    // target_aborter (aborters.hpp:146-175)
    DTR_AI void target_aborter(uint32_t target) {
        if (on_surface() && cur_sf() == target && tk.path_length > 0.0) {
            if (tk.status != kAbort) tk.target_hit = true;
            tk.status = kAbort;
            tk.heartbeat = false;
        }
    }

    DTR_AI void parameter_updater(const double* init_cov, uint64_t n, uint32_t gid) {
        if (!(on_sensitive() || on_material())) return;
        const uint32_t dest = cur_sf();
        double fp[8];
        if (cfg.actors & DTR_ACTOR_STEP_TRACER) {
            fp[0] = tk.pos.x; fp[1] = tk.pos.y; fp[2] = tk.pos.z; fp[3] = tk.time;
            fp[4] = tk.dir.x; fp[5] = tk.dir.y; fp[6] = tk.dir.z; fp[7] = tk.qop;
        }
        if (tk.up_sf != kInvalid) {
            if (tk.up_sf != dest) {
                transport_covariance(dest);
                if (fabs(tk.path_length) > 0.0) reset_J();
                free_to_bound(dest, tk.bound);
                tk.up_sf = dest;
            }
        } else {
            if (init_cov) {
                for (int k = 0; k < 21; ++k) tk.cov[k] = static_cast<S>(init_cov[k * n + gid]);
            } else {
                for (int r = 0; r < 6; ++r)
                    for (int cc = r; cc < 6; ++cc) tk.cov[covi(r, cc)] = (r == cc) ? S(1) : S(0);
            }
            free_to_bound(dest, tk.bound);
            tk.up_sf = dest;
        }
        if (on_material()) material_interaction(dest);
        if (cfg.actors & DTR_ACTOR_STEP_TRACER) trace_update(dest, fp, n, gid);
    }
};

// Local navigation of a volume (navigation_functions.hpp:193-246) on the
// navigation fields only: it returns next | last << 8 | status << 16 |
// trust << 24, so the caller's track state is never addressed.
template <class P>
DTR_AI uint32_t local_navigation(const FastDetector& det, const DevConfig& cfg, V3 pos, V3 dir, double ext_tol,
                                 uint16_t volume, float overstep, CacheView c) {
    NavLite nl;
    nl.ext_tol = ext_tol;
    nl.next = 0;
    nl.last = -1;
    nl.status = kUnknown;
    nl.trust = kNoTrust;
    nl.volume = volume;
    nl.pos = pos;
    nl.dir = dir;
    const DevField no_field{};
    FastEngine<P, NavLite> e{det, no_field, cfg, nl, c};
    e.local_navigation_body(overstep);
    return static_cast<uint32_t>(static_cast<uint8_t>(nl.next)) |
           (static_cast<uint32_t>(static_cast<uint8_t>(nl.last)) << 8) |
           (static_cast<uint32_t>(static_cast<uint8_t>(nl.status)) << 16) |
           (static_cast<uint32_t>(nl.trust) << 24);
}

// The same out of line, for the per-lane loop, which needs it from four
// places of the navigation cascade (it is the largest function).
template <class P>
__host__ __device__ __attribute__((noinline)) uint32_t local_navigation_ool(
    const FastDetector& det, const DevConfig& cfg, V3 pos, V3 dir, double ext_tol, uint16_t volume,
    float overstep, CacheView c) {
    return local_navigation<P>(det, cfg, pos, dir, ext_tol, volume, overstep, c);
}

// Track inputs of one batch (structure of arrays, parameter-major).
struct TrackInputs {
    const double* fp;        // [8][n] free start, or null
    const double* init_cov;  // [21][n] or null
    const double* bound0;    // [6][n] bound start, or null
    const uint32_t* start_sf;
    uint64_t n;
};

// A propagation that pauses and resumes (propagator.hpp:238-250, 274-282):
// the reference pauses a track when an actor stops its heartbeat while the
// navigation is alive, and a later propagate() call continues it without
// initialising the navigation.  Here the pause comes from a budget of
// Runge-Kutta steps per call, checked where the actor half of the loop
// begins (where such an actor, first in the chain, would run), and the
// lane's whole state goes to a device buffer, structure of arrays over the
// tracks of the batch: kPauseD doubles and kPauseW words per track.
enum PauseSlotD : int {
    kPdPos = 0, kPdDir = 3, kPdTime = 6, kPdQop = 7, kPdStep = 8, kPdNextStep = 9, kPdPath = 10,
    kPdAbsPath = 11, kPdConstraint = 12, kPdDtds = 13, kPdDqopds = 16, kPdHStage = 17, kPdExtTol = 18,
    kPdNavDir = 19, kPdBound = 20, kPdJ = 26, kPdCov = 51, kPdJx = 72, kPdCache = 90, kPauseD = 98
};
enum PauseSlotW : int {
    kPwTrials = 0, kPwSteps = 1, kPwUpSf = 2, kPwSeq = 3, kPwTrace = 4, kPwNav = 5, kPwVolFlags = 6,
    kPwStall = 7, kPwState = 8, kPwCacheSf = 9, kPwCacheLs = 17, kPauseW = 25
};
// kPwState: the first call writes kPaused or kDone for every track; a
// resuming call runs the kPaused ones
enum : uint32_t { kLanePaused = 1u, kLaneDone = 2u };
struct PauseView {
    double* d;    // [kPauseD][n]
    uint32_t* w;  // [kPauseW][n]
    uint64_t n;
    uint32_t max_steps;  // steps a track takes in one call before it pauses
    uint32_t resume;     // 0: start the tracks from the inputs; 1: continue the paused ones
    DTR_AI double& dv(int slot, uint32_t t) const { return d[static_cast<uint64_t>(slot) * n + t]; }
    DTR_AI uint32_t& wv(int slot, uint32_t t) const { return w[static_cast<uint64_t>(slot) * n + t]; }
};

// propagator::propagate (propagator.hpp:262-400) for one track, split into
// phases so that the lanes of a wave can meet for the local navigations
// (each lane's navigation cascade stops where it needs one) and a persistent
// wave can swap a finished lane's track for a new one.  `c` is the lane's
// view of the LDS candidate cache.
//
// The loop of propagate() alternates an actor half (even i) and a step half
// (odd i), with `i % 2 == 0 || heartbeat` as its condition, so an actor half
// always follows a step half:
//   actor_half()          -> ends the loop if the heartbeat is gone
//   navigation update     -> ends the loop if the heartbeat is gone
//   step_half()
//   navigation update
template <class P>
struct Lane {
    using S = typename P::S;
    FastTrack<P> tk;
    bool is_init;
    bool use_init_cov;
    bool nav_ret;
    uint8_t nav_stage;
    double last_path;
    uint32_t stall;
    uint32_t gid;

    // Track set-up (propagator.hpp:79-162); the navigator's initialisation
    // (navigator_base.hpp:57-66, a local navigation without overstepping
    // resolution, overstep tolerance `os`) is then requested from the caller.
    DTR_AI void start_begin(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                            const TrackInputs& in, CacheView c, uint32_t track, float& os) {
        gid = track;
        if constexpr (P::kLdsState) {
            static_assert(sizeof(S) == sizeof(float), "LDS track state holds FP32 values");
            tk.J.p = c.state;
            tk.cov.p = c.state + 25;
        }
        for (int i = 0; i < kCache; ++i) {
            c.p(i) = kMax;
            c.s(i) = kInvalid;
            c.l(i) = static_cast<uint32_t>(kOutside) << 16;
        }
        tk.ext_tol = 0.0;
        tk.next = 0;
        tk.last = -1;
        tk.status = kUnknown;
        tk.trust = kNoTrust;
        tk.volume = 0;
        tk.path_limit_hit = false;
        tk.overflow = false;
        tk.stopped = false;
        tk.momentum_hit = false;
        tk.target_hit = false;
        tk.nav_dir = 1.0;
        tk.n_trace = 0;
        tk.step_size = 0.0;
        tk.h_stage = 0.0;
        tk.next_step_size = 0.0;
        tk.path_length = 0.0;
        tk.abs_path_length = 0.0;
        tk.constraint_abort = kMax;
        tk.dtds3 = {S(0), S(0), S(0)};
        tk.dqopds3 = 0.0;
        tk.n_trials = 0;
        tk.n_steps = 0;
        for (int i = 0; i < 6; ++i) tk.bound[i] = 0.0;
        for (int i = 0; i < 21; ++i) tk.cov[i] = S(0);
        tk.up_sf = kInvalid;
        tk.n_seq = 0;
        FastEngine<P> e{det, field, cfg, tk, c};
        const uint64_t N = in.n;
        use_init_cov = in.init_cov != nullptr;
        if (in.bound0) {
            // bound start (propagator.hpp:121-162, base_stepper.hpp:84-108)
            const uint32_t s0 = in.start_sf[gid];
            for (int k = 0; k < 6; ++k) tk.bound[k] = in.bound0[k * N + gid];
            for (int r = 0; r < 6; ++r)
                for (int cc = r; cc < 6; ++cc) {
                    const int k = covi(r, cc);
                    tk.cov[k] = use_init_cov ? static_cast<S>(in.init_cov[k * N + gid])
                                             : (r == cc ? S(1) : S(0));
                }
            tk.up_sf = s0;
            tk.pos = e.bound_pos(s0, tk.bound);
            tk.dir = Engine::bound_dir(tk.bound);
            tk.time = tk.bound[5];
            tk.qop = tk.bound[4];
            tk.volume = det.sf[s0].volume;
            use_init_cov = false;
        } else {
            tk.pos = {in.fp[0 * N + gid], in.fp[1 * N + gid], in.fp[2 * N + gid]};
            tk.time = in.fp[3 * N + gid];
            tk.dir = {in.fp[4 * N + gid], in.fp[5 * N + gid], in.fp[6 * N + gid]};
            tk.qop = in.fp[7 * N + gid];
        }
        e.reset_J();
        os = -cfg.path_tol;
    }

    // after the initial local navigation has been applied
    DTR_AI void start_end() {
        tk.heartbeat = tk.status > kUnknown;
        is_init = true;
        last_path = 0.0;
        stall = 0;
    }

    // pathlimit_aborter (aborters.hpp:44-75), parameter_updater<pointwise_
    // material_interactor>, surface_sequencer (surface_sequencer.hpp:78-106).
    // Returns true when propagate()'s loop ends here.
    DTR_AI bool actor_half(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                           const TrackInputs& in, const DevResults& out, CacheView c) {
        FastEngine<P> e{det, field, cfg, tk, c};
        if (tk.status != kExit) {
            const double step_limit = cfg.path_limit - fabs(tk.abs_path_length);
            if (step_limit <= 0.0) {
                tk.status = kAbort;
                tk.path_limit_hit = true;
                tk.heartbeat = false;
            }
            tk.constraint_abort = dmin(tk.constraint_abort, fabs(step_limit));
        }
        if (cfg.use_material) e.parameter_updater(use_init_cov ? in.init_cov : nullptr, in.n, gid);
        if (tk.stopped && tk.status > kUnknown) {
            tk.status = kAbort;
            tk.heartbeat = false;
        }
        if (cfg.max_surfaces != 0u && (e.on_sensitive() || e.on_material())) {
            if (tk.n_seq == cfg.max_surfaces) {
                tk.overflow = true;
                tk.status = kExit;
                tk.heartbeat = false;
            } else {
                out.surfaces[static_cast<uint64_t>(tk.n_seq) * in.n + gid] = e.cur_sf();
                tk.n_seq++;
            }
        }
        if (cfg.actors & DTR_ACTOR_MOMENTUM_ABORTER) e.momentum_aborter();
        if (cfg.actors & DTR_ACTOR_TARGET_ABORTER) e.target_aborter(cfg.target[gid]);
        if (!tk.heartbeat) return true;
        last_path = tk.path_length;
        return false;
    }

    // rk_stepper::step + the step-size policy + the stall guard
    DTR_AI void step_half(const FastDetector& det, const DevField& field, const DevConfig& cfg, CacheView c) {
        FastEngine<P> e{det, field, cfg, tk, c};
        const bool reset = e.on_surface() || is_init;
        e.step(e.nav_dist(), reset);
        e.rk_policy();
        is_init = false;
        if (fabs(tk.path_length) <= fabs(last_path) + static_cast<double>(cfg.path_tol)) {
            if (stall >= 10u) {
                tk.heartbeat = false;
                tk.status = kAbort;
            }
            stall++;
        } else {
            stall = 0;
        }
    }

    // navigation update: nav_begin, then (apply_local_navigation, nav_resume)
    // while a local navigation is requested, then nav_end.
    DTR_AI bool nav_begin(const FastDetector& det, const DevConfig& cfg, CacheView c, float& os) {
        const DevField no_field{};
        FastEngine<P> e{det, no_field, cfg, tk, c};
        return e.nav_begin(nav_stage, nav_ret, os);
    }
    DTR_AI bool nav_resume(const FastDetector& det, const DevConfig& cfg, CacheView c, float& os) {
        const DevField no_field{};
        FastEngine<P> e{det, no_field, cfg, tk, c};
        return e.nav_resume(nav_stage, nav_ret, os);
    }
    DTR_AI void apply_local_navigation(uint32_t r) {
        tk.next = static_cast<int8_t>(r & 0xffu);
        tk.last = static_cast<int8_t>((r >> 8) & 0xffu);
        tk.status = static_cast<int8_t>((r >> 16) & 0xffu);
        tk.trust = static_cast<uint8_t>(r >> 24);
    }
    // returns true when propagate()'s loop ends here (heartbeat gone)
    DTR_AI bool nav_end() {
        is_init |= nav_ret;
        tk.heartbeat = tk.heartbeat && (tk.status > kUnknown);
        return !tk.heartbeat;
    }

    DTR_AI void finish(const DevResults& out, uint64_t N, const DevConfig& cfg) const {
        if (cfg.actors & DTR_ACTOR_STEP_TRACER) cfg.trace.n[gid] = tk.n_trace;
        if (out.status)
            out.status[gid] = track_status(tk.status == kExit, tk.overflow, tk.path_limit_hit, tk.momentum_hit,
                                           tk.target_hit);
        if (out.bound_params)
            for (int k = 0; k < 6; ++k) out.bound_params[k * N + gid] = tk.bound[k];
        if (out.bound_cov)
            for (int k = 0; k < 21; ++k) out.bound_cov[k * N + gid] = static_cast<double>(tk.cov[k]);
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

    // The lane's state at the start of an actor half, to the pause buffer
    // and back; tk.bound, the FP32 Jacobian and covariance of the stage
    // precision and the candidate cache included.
    DTR_AI void save(const PauseView& s, CacheView c) const {
        const uint32_t t = gid;
        s.dv(kPdPos + 0, t) = tk.pos.x;
        s.dv(kPdPos + 1, t) = tk.pos.y;
        s.dv(kPdPos + 2, t) = tk.pos.z;
        s.dv(kPdDir + 0, t) = tk.dir.x;
        s.dv(kPdDir + 1, t) = tk.dir.y;
        s.dv(kPdDir + 2, t) = tk.dir.z;
        s.dv(kPdTime, t) = tk.time;
        s.dv(kPdQop, t) = tk.qop;
        s.dv(kPdStep, t) = tk.step_size;
        s.dv(kPdNextStep, t) = tk.next_step_size;
        s.dv(kPdPath, t) = tk.path_length;
        s.dv(kPdAbsPath, t) = tk.abs_path_length;
        s.dv(kPdConstraint, t) = tk.constraint_abort;
        s.dv(kPdDtds + 0, t) = static_cast<double>(tk.dtds3.x);
        s.dv(kPdDtds + 1, t) = static_cast<double>(tk.dtds3.y);
        s.dv(kPdDtds + 2, t) = static_cast<double>(tk.dtds3.z);
        s.dv(kPdDqopds, t) = tk.dqopds3;
        s.dv(kPdHStage, t) = tk.h_stage;
        s.dv(kPdExtTol, t) = tk.ext_tol;
        s.dv(kPdNavDir, t) = tk.nav_dir;
        for (int i = 0; i < 6; ++i) s.dv(kPdBound + i, t) = tk.bound[i];
        for (int i = 0; i < 25; ++i) s.dv(kPdJ + i, t) = static_cast<double>(tk.J[i]);
        for (int i = 0; i < 21; ++i) s.dv(kPdCov + i, t) = static_cast<double>(tk.cov[i]);
        if constexpr (P::kGrad)
            for (int i = 0; i < 18; ++i) s.dv(kPdJx + i, t) = static_cast<double>(tk.Jx.v[i]);
        for (int i = 0; i < kCache; ++i) {
            s.dv(kPdCache + i, t) = c.p(i);
            s.wv(kPwCacheSf + i, t) = c.s(i);
            s.wv(kPwCacheLs + i, t) = c.l(i);
        }
        s.wv(kPwTrials, t) = tk.n_trials;
        s.wv(kPwSteps, t) = tk.n_steps;
        s.wv(kPwUpSf, t) = tk.up_sf;
        s.wv(kPwSeq, t) = tk.n_seq;
        s.wv(kPwTrace, t) = tk.n_trace;
        s.wv(kPwNav, t) = static_cast<uint32_t>(static_cast<uint8_t>(tk.next)) |
                          (static_cast<uint32_t>(static_cast<uint8_t>(tk.last)) << 8) |
                          (static_cast<uint32_t>(static_cast<uint8_t>(tk.status)) << 16) |
                          (static_cast<uint32_t>(tk.trust) << 24);
        s.wv(kPwVolFlags, t) = static_cast<uint32_t>(tk.volume) | (tk.path_limit_hit ? 1u << 16 : 0u) |
                               (tk.overflow ? 1u << 17 : 0u) | (tk.stopped ? 1u << 18 : 0u) |
                               (tk.heartbeat ? 1u << 19 : 0u) | (tk.momentum_hit ? 1u << 20 : 0u) |
                               (tk.target_hit ? 1u << 21 : 0u) | (is_init ? 1u << 22 : 0u) |
                               (use_init_cov ? 1u << 23 : 0u);
        s.wv(kPwStall, t) = stall;
    }
    DTR_AI void load(const PauseView& s, CacheView c, uint32_t track) {
        gid = track;
        const uint32_t t = track;
        if constexpr (P::kLdsState) {
            tk.J.p = c.state;
            tk.cov.p = c.state + 25;
        }
        tk.pos = {s.dv(kPdPos + 0, t), s.dv(kPdPos + 1, t), s.dv(kPdPos + 2, t)};
        tk.dir = {s.dv(kPdDir + 0, t), s.dv(kPdDir + 1, t), s.dv(kPdDir + 2, t)};
        tk.time = s.dv(kPdTime, t);
        tk.qop = s.dv(kPdQop, t);
        tk.step_size = s.dv(kPdStep, t);
        tk.next_step_size = s.dv(kPdNextStep, t);
        tk.path_length = s.dv(kPdPath, t);
        tk.abs_path_length = s.dv(kPdAbsPath, t);
        tk.constraint_abort = s.dv(kPdConstraint, t);
        tk.dtds3 = {static_cast<S>(s.dv(kPdDtds + 0, t)), static_cast<S>(s.dv(kPdDtds + 1, t)),
                    static_cast<S>(s.dv(kPdDtds + 2, t))};
        tk.dqopds3 = s.dv(kPdDqopds, t);
        tk.h_stage = s.dv(kPdHStage, t);
        tk.ext_tol = s.dv(kPdExtTol, t);
        tk.nav_dir = s.dv(kPdNavDir, t);
        for (int i = 0; i < 6; ++i) tk.bound[i] = s.dv(kPdBound + i, t);
        for (int i = 0; i < 25; ++i) tk.J[i] = static_cast<S>(s.dv(kPdJ + i, t));
        for (int i = 0; i < 21; ++i) tk.cov[i] = static_cast<S>(s.dv(kPdCov + i, t));
        if constexpr (P::kGrad)
            for (int i = 0; i < 18; ++i) tk.Jx.v[i] = static_cast<S>(s.dv(kPdJx + i, t));
        for (int i = 0; i < kCache; ++i) {
            c.p(i) = s.dv(kPdCache + i, t);
            c.s(i) = s.wv(kPwCacheSf + i, t);
            c.l(i) = s.wv(kPwCacheLs + i, t);
        }
        tk.n_trials = s.wv(kPwTrials, t);
        tk.n_steps = s.wv(kPwSteps, t);
        tk.up_sf = s.wv(kPwUpSf, t);
        tk.n_seq = s.wv(kPwSeq, t);
        tk.n_trace = s.wv(kPwTrace, t);
        const uint32_t nav = s.wv(kPwNav, t);
        tk.next = static_cast<int8_t>(nav & 0xffu);
        tk.last = static_cast<int8_t>((nav >> 8) & 0xffu);
        tk.status = static_cast<int8_t>((nav >> 16) & 0xffu);
        tk.trust = static_cast<uint8_t>(nav >> 24);
        const uint32_t vf = s.wv(kPwVolFlags, t);
        tk.volume = static_cast<uint16_t>(vf & 0xffffu);
        tk.path_limit_hit = (vf >> 16) & 1u;
        tk.overflow = (vf >> 17) & 1u;
        tk.stopped = (vf >> 18) & 1u;
        tk.heartbeat = (vf >> 19) & 1u;
        tk.momentum_hit = (vf >> 20) & 1u;
        tk.target_hit = (vf >> 21) & 1u;
        is_init = (vf >> 22) & 1u;
        use_init_cov = (vf >> 23) & 1u;
        stall = s.wv(kPwStall, t);
        last_path = tk.path_length;
        nav_ret = false;
        nav_stage = 0;
    }
    // pause here: the state to the buffer, status DTR_TRACK_PAUSED and the
    // counts so far to the results (the surface sequence is already there)
    DTR_AI void pause(const PauseView& s, CacheView c, const DevResults& out) const {
        save(s, c);
        s.wv(kPwState, gid) = kLanePaused;
        if (out.status) out.status[gid] = DTR_TRACK_PAUSED;
        if (out.n_surfaces) out.n_surfaces[gid] = tk.n_seq;
        if (out.n_steps) out.n_steps[gid] = tk.n_steps;
        if (out.n_trials) out.n_trials[gid] = tk.n_trials;
    }

    // one lane on its own, with out-of-line local navigations
    DTR_AI void navigate_alone(const FastDetector& det, const DevConfig& cfg, CacheView c) {
        float os;
        bool req = nav_begin(det, cfg, c, os);
        while (req) {
            apply_local_navigation(local_navigation_ool<P>(det, cfg, tk.pos, tk.dir, tk.ext_tol, tk.volume, os, c));
            req = nav_resume(det, cfg, c, os);
        }
    }

    DTR_AI void start(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                      const TrackInputs& in, CacheView c, uint32_t track) {
        float os;
        start_begin(det, field, cfg, in, c, track, os);
        apply_local_navigation(local_navigation_ool<P>(det, cfg, tk.pos, tk.dir, tk.ext_tol, tk.volume, os, c));
        start_end();
    }

    // one actor half and one step half; true once propagate()'s loop ended
    DTR_AI bool advance(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                        const TrackInputs& in, const DevResults& out, CacheView c) {
        if (actor_half(det, field, cfg, in, out, c)) return true;
        navigate_alone(det, cfg, c);
        if (nav_end()) return true;
        step_half(det, field, cfg, c);
        navigate_alone(det, cfg, c);
        nav_end();
        return false;
    }
};

// One track from start to finish, lane on its own.  Forced inline so that
// every kernel instantiation allocates registers for its own occupancy.
template <class P>
DTR_AI void propagate_fast(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                           const TrackInputs& in, const DevResults& out, uint32_t track, CacheView c) {
    Lane<P> lane;
    lane.start(det, field, cfg, in, c, track);
    while (!lane.advance(det, field, cfg, in, out, c)) {
    }
    lane.finish(out, in.n, cfg);
}

// propagate_fast with a step budget per call: the track starts from the
// inputs (or, resuming, from its paused state), pauses once it has taken
// max_steps steps in this call, and finishes as propagate_fast would.
template <class P>
DTR_AI void propagate_resumable(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                                const TrackInputs& in, const DevResults& out, uint32_t track, CacheView c,
                                const PauseView& pv) {
    Lane<P> lane;
    if (pv.resume) {
        if (pv.wv(kPwState, track) != kLanePaused) return;
        lane.load(pv, c, track);
    } else {
        lane.start(det, field, cfg, in, c, track);
    }
    const uint32_t s0 = lane.tk.n_steps;
    for (;;) {
        if (lane.tk.n_steps - s0 >= pv.max_steps) {
            lane.pause(pv, c, out);
            return;
        }
        if (lane.advance(det, field, cfg, in, out, c)) break;
    }
    lane.finish(out, in.n, cfg);
    pv.wv(kPwState, track) = kLaneDone;
}

}  // namespace dtr

#pragma clang fp contract(off)


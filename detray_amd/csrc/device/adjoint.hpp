// Discrete adjoint of the propagation: d(chi2 of hit residuals)/d(initial
// bound parameters).
//
// Forward (recording) pass: the propagation of engine.hpp for one track
// from a bound start, which also writes a tape: the stage and advance step
// sizes of every Runge-Kutta step, the free state every K steps
// (checkpoints), and every surface event of the parameter updater (free
// state before, bound parameters, free state after the setter).  At the
// surfaces that carry the track's hits it accumulates chi2 = sum r^T V^-1 r
// of the local residuals and the gradient the reference's transport
// Jacobians give: the chain of full bound-to-bound Jacobians of
// parameter_updater.hpp:422-572 from the start to each hit.
//
// Backward pass, from the last event to the start, one segment of K steps
// at a time: the segment is replayed from its checkpoint (events inside it
// restore their recorded states), then the adjoint of the free state runs
// back through each step (vector-Jacobian product of the RK4 step, stages
// included) and each surface event (bound-to-free map, the mean energy loss
// as a function of the bound parameters, path-corrected free-to-bound map,
// residual).
//
// Two models:
//  * exact: the derivative of the propagation itself (field gradient of the
//    trilinear map, energy-loss derivative, normalisation of the direction,
//    the free state continuing unchanged through a surface without
//    material, and the length of every step that the navigator sized to
//    reach a surface, event or not, as a function of the state its distance
//    was computed from);
//  * transport: the derivative the reference's Jacobians describe
//    (rk_stepper.hpp:633-956 without field gradient, stepping_config.hpp:
//    27-48 use_field_gradient / use_eloss_gradient off; no normalisation;
//    stage derivatives at the advance step size; every event restarts from
//    its bound parameters), so its gradient equals the Jacobian-chain one.
#pragma once

#include "device/engine.hpp"

namespace dtr {

// ------------------------------------------------------------------- tape --
// Caller-allocated, structure of arrays over the n tracks of a batch, one
// column per work-item: with a track order, column k holds track order[k],
// so that the lanes of a wave, which run neighbouring tracks of the order,
// read and write neighbouring columns.
struct Tape {
    uint64_t n;
    uint32_t max_steps, max_events, every, n_ckpt;
    double* h;          // [max_steps][2][n] stage and advance step size
    double* ckpt;       // [n_ckpt][7][n] free state (pos, dir, q/p) before step j*every+1
    double* ev_pre;     // [max_events][10][n] pos, dir, q/p, dt/ds before the event
    double* ev_bound;   // [max_events][6][n] bound parameters before material
    double* ev_post;    // [max_events][7][n] free state after the event
    double* seg;        // [every][7][n] replay of one segment
    double* jcum;       // [36][n] chain of bound-to-bound Jacobians
    uint32_t* ev_step;  // [max_events][n] steps taken before the event
    uint32_t* ev_sf;    // [max_events][n] surface
    int32_t* ev_hit;    // [max_events][n] matched hit or -1
    uint32_t* ev_flags; // [max_events][n] kEvFromFree | kEvMaterial | kEvReused
    uint32_t* step_land;  // [max_steps][n] surface step j+1 was sized to reach (its length is the
                          // navigator's distance to it), or kInvalid
    uint32_t* n_steps;  // [n]
    uint32_t* n_events; // [n]
    uint32_t* status;   // [n] kTape*
};

enum : uint32_t { kTapeOk = 0, kTapeSteps = 1, kTapeEvents = 2, kTapeUnsupported = 3 };
// kEvReused: the track reached the surface its bound parameters are on
// (a placed cylinder crossed twice, a return to the start surface): the
// parameter transporter keeps them (parameter_updater.hpp:314-419), so the
// event works on the bound parameters of the last bound event
enum : uint32_t { kEvFromFree = 1u, kEvMaterial = 2u, kEvReused = 4u };

DTR_HD uint64_t tape_doubles(uint32_t max_steps, uint32_t max_events, uint32_t every) {
    const uint64_t n_ckpt = max_steps / every + 1;
    return 2ull * max_steps + 7ull * n_ckpt + 23ull * max_events + 7ull * every + 36ull;
}

// bytes of a tape for n tracks
DTR_HD size_t tape_bytes(uint64_t n, uint32_t max_steps, uint32_t max_events, uint32_t every) {
    const uint64_t words = 4ull * max_events + max_steps + 3ull;
    return static_cast<size_t>(n * (8ull * tape_doubles(max_steps, max_events, every) + 4ull * words));
}

DTR_HD Tape tape_carve(void* base, uint64_t n, uint32_t max_steps, uint32_t max_events, uint32_t every) {
    Tape t;
    t.n = n;
    t.max_steps = max_steps;
    t.max_events = max_events;
    t.every = every;
    t.n_ckpt = max_steps / every + 1;
    double* d = static_cast<double*>(base);
    t.h = d;
    d += 2ull * max_steps * n;
    t.ckpt = d;
    d += 7ull * t.n_ckpt * n;
    t.ev_pre = d;
    d += 10ull * max_events * n;
    t.ev_bound = d;
    d += 6ull * max_events * n;
    t.ev_post = d;
    d += 7ull * max_events * n;
    t.seg = d;
    d += 7ull * every * n;
    t.jcum = d;
    d += 36ull * n;
    uint32_t* w = reinterpret_cast<uint32_t*>(d);
    const uint64_t me = static_cast<uint64_t>(max_events) * n;
    t.ev_step = w;
    w += me;
    t.ev_sf = w;
    w += me;
    t.ev_hit = reinterpret_cast<int32_t*>(w);
    w += me;
    t.ev_flags = w;
    w += me;
    t.step_land = w;
    w += static_cast<uint64_t>(max_steps) * n;
    t.n_steps = w;
    w += n;
    t.n_events = w;
    w += n;
    t.status = w;
    return t;
}

// Hits of the tracks: surface [max_hits][n] in the order the track crosses
// them, loc [2][max_hits][n], count [n].
struct HitsView {
    uint32_t max_hits;
    const uint32_t* surface;
    const double* loc;
    const uint32_t* n_hits;
};
struct HitsOut {
    uint32_t max_hits;
    uint32_t* surface;
    double* loc;
    uint32_t* n_hits;
};

struct AdjointIO {
    const double* bound0;      // [6][n]
    const uint32_t* start_sf;  // [n]
    uint64_t n;
    HitsView hits;             // surface null: no hits
    HitsOut record;            // surface null: off
    double iv0, iv1;           // inverse variances of loc0, loc1
    double* chi2;              // [n] or null
    double* grad_transport;    // [6][n] or null
    uint32_t* status;          // [n] or null: copy of the tape status
    Tape tape;
};

// ------------------------------------------------------------ dual number --
// value and partial derivatives with respect to (loc0, phi, theta, q/p)
struct D4 {
    double v, d[4];
};
DTR_HD D4 dconst(double v) { return {v, {0.0, 0.0, 0.0, 0.0}}; }
DTR_HD D4 dvar(double v, int i) {
    D4 r = dconst(v);
    r.d[i] = 1.0;
    return r;
}
DTR_HD D4 dchain(double v, double dv, D4 a) {  // f(a) with f(a.v) = v, f'(a.v) = dv
    return {v, {a.d[0] * dv, a.d[1] * dv, a.d[2] * dv, a.d[3] * dv}};
}
DTR_HD D4 operator+(D4 a, D4 b) {
    return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2], a.d[3] + b.d[3]}};
}
DTR_HD D4 operator-(D4 a, D4 b) {
    return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2], a.d[3] - b.d[3]}};
}
DTR_HD D4 operator*(D4 a, D4 b) {
    D4 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
DTR_HD D4 operator/(D4 a, D4 b) {
    D4 r;
    r.v = a.v / b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
    return r;
}
DTR_HD D4 dscale(D4 a, double s) { return dchain(a.v * s, s, a); }
DTR_HD D4 dsqrt(D4 a) {
    const double s = sqrt(a.v);
    return dchain(s, 0.5 / s, a);
}
DTR_HD D4 dlog(D4 a) { return dchain(log(a.v), 1.0 / a.v, a); }
DTR_HD D4 dfabs(D4 a) { return a.v < 0.0 ? dscale(a, -1.0) : a; }
DTR_HD D4 dcos(D4 a) { return dchain(cos(a.v), -sin(a.v), a); }
DTR_HD D4 dsin(D4 a) { return dchain(sin(a.v), cos(a.v), a); }

DTR_HD D4 dpow(D4 a, double k) { return dchain(pow(a.v, k), k * pow(a.v, k - 1.0), a); }

// d/dx of half the density correction (relativistic_quantities.hpp:
// 176-200): the approximate high-energy form without density-effect data,
// Sternheimer's parametrisation (density_effect_data.hpp) with it.
DTR_HD D4 delta_half_d(const dtr_material_slab& m, D4 bg, double ned) {
    if (!m.has_density) {
        if (!(bg.v >= 10.0)) return dconst(0.0);
        const double plasma = (28.816 * 1e-9) * sqrt(1000.0 * ned);
        return dlog(dscale(bg, plasma / m.mean_exc)) - dconst(0.5);
    }
    const double ln10 = 2.302585092994045684;
    const D4 x = dscale(dlog(bg), 1.0 / ln10);
    if (!(x.v >= m.density[2])) return dconst(0.0);
    D4 delta = dscale(x, 2.0 * ln10) - dconst(m.density[4]);
    if (x.v < m.density[3]) delta = delta + dscale(dpow(dconst(m.density[3]) - x, m.density[1]), m.density[0]);
    return dscale(delta, 0.5);
}

// q/p after the mean energy loss (pointwise_material_interactor.hpp:147-326,
// the expressions of FastEngine::material_interaction) as a function of the
// bound parameters before it: loc0 (the normal of a concentric or placed
// cylinder, the closest approach to a rod), phi, theta and q/p.
DTR_HD D4 eloss_qop(const NavSurface& rec, const dtr_material_slab& m, const double b[6], double mass, double q) {
    const D4 l0 = dvar(b[0], 0), ph = dvar(b[2], 1), th = dvar(b[3], 2), qop = dvar(b[4], 3);
    const D4 st = dsin(th);
    const D4 dx = dcos(ph) * st, dy = dsin(ph) * st, dz = dcos(th);
    D4 cos_inc;
    if (rec.shape == DTR_CONCENTRIC_CYLINDER2D) {
        cos_inc = dfabs(dx * dcos(l0) + dy * dsin(l0));
    } else if (rec.shape == DTR_CYLINDER2D) {
        // normal at phi = loc0 / r (cylindrical2D.hpp:62-70)
        const D4 a = dscale(l0, 1.0 / rec.m[0]);
        const D4 ca = dcos(a), sa = dsin(a);
        const D4 nx = dscale(ca, rec.r[0]) + dscale(sa, rec.r[1]);
        const D4 ny = dscale(ca, rec.r[3]) + dscale(sa, rec.r[4]);
        const D4 nz = dscale(ca, rec.r[6]) + dscale(sa, rec.r[7]);
        cos_inc = dfabs(dx * nx + dy * ny + dz * nz);
    } else {
        // plane normal, or the wire axis of a line
        cos_inc = dfabs(dscale(dx, rec.r[2]) + dscale(dy, rec.r[5]) + dscale(dz, rec.r[8]));
    }
    D4 path_seg;
    if (m.kind == DTR_MATERIAL_ROD) {
        // material_rod.hpp:62-78 with the closest approach loc0
        path_seg = dconst(0.0);
        if (!(fabs(b[0]) > m.thickness))
            path_seg = dscale(dsqrt((dconst(m.thickness * m.thickness) - l0 * l0) / (dconst(1.0) - cos_inc * cos_inc)),
                              2.0);
    } else {
        path_seg = dconst(m.thickness) / cos_inc;
    }
    const D4 mq = dscale(qop, mass);
    const D4 q2b2 = dconst(q * q) + mq * mq;
    const D4 mOverP = dscale(dfabs(qop), mass / fabs(q));
    const D4 pOverM = dconst(1.0) / mOverP;
    const D4 beta2 = dconst(1.0) / (dconst(1.0) + mOverP * mOverP);
    const D4 bg = pOverM;
    const D4 gamma = dsqrt(dconst(1.0) + pOverM * pOverM);
    const double m_e = 0.51099895069 * 1e-3;
    const double mfrac = m_e / mass;
    const D4 Wmax = dscale(bg * bg, 2.0 * m_e) / (dconst(1.0 + mfrac * mfrac) + dscale(gamma, 2.0 * mfrac));
    const double K = 0.307075 * 1e-3 * 1e2;
    const double ned = m.Z * m.molar_rho;
    const D4 eps_pl = dscale(q2b2, 0.5 * K * ned);
    D4 bb = dconst(0.0);
    if (eps_pl.v > 0.0) {
        const D4 dhalf = delta_half_d(m, bg, ned);
        const D4 u = dscale(bg * bg, 2.0 * m_e);
        const D4 A = dscale(dlog(dscale(u * Wmax, 1.0 / (m.mean_exc * m.mean_exc))), 0.5);
        bb = dscale(eps_pl, 2.0) * (A - beta2 - dhalf);
    }
    const D4 e_loss = dfabs(path_seg * bb);
    const D4 p = dconst(q) / qop;
    const D4 next_e = dsqrt(dconst(mass * mass) + p * p) - e_loss;
    const D4 next_p = dsqrt(next_e * next_e - dconst(mass * mass));
    return dconst(q) / next_p;
}

// ---------------------------------------------------- field and its gradient
// G[i][a] = dB_i/dp_a of the trilinear map at p (zero along a clamped axis)
template <typename S>
DTR_HD void field_grad_s(const DevField& f, V3 p, S G[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) G[i][a] = S(0);
    if (f.kind == DTR_FIELD_CONSTANT) return;
    const double pc[3] = {p.x, p.y, p.z};
    int i0[3];
    S t[3], inside[3];
    for (int a = 0; a < 3; ++a) {
        const int n = static_cast<int>(f.n[a]);
        double u = (pc[a] - f.min[a]) / f.pitch[a];
        inside[a] = (u > 0.0 && u < n - 1) ? static_cast<S>(1.0 / f.pitch[a]) : S(0);
        u = (u > 0.0) ? (u > n - 1 ? static_cast<double>(n - 1) : u) : 0.0;
        int i = static_cast<int>(floor(u));
        i = i > n - 2 ? n - 2 : i;
        i0[a] = i;
        t[a] = static_cast<S>(u - i);
    }
    const uint64_t ny = f.n[1], nz = f.n[2];
    const uint64_t base = (static_cast<uint64_t>(i0[0]) * ny + i0[1]) * nz + i0[2];
    const uint64_t sx = ny * nz, sy = nz;
    for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
                const float* v = f.values + 3 * (base + cx * sx + cy * sy + cz);
                const S wx = cx ? t[0] : S(1) - t[0], wy = cy ? t[1] : S(1) - t[1];
                const S wz = cz ? t[2] : S(1) - t[2];
                const S gx = (cx ? S(1) : S(-1)) * inside[0], gy = (cy ? S(1) : S(-1)) * inside[1];
                const S gz = (cz ? S(1) : S(-1)) * inside[2];
                for (int i = 0; i < 3; ++i) {
                    const S b = static_cast<S>(v[i]);
                    G[i][0] += gx * wy * wz * b;
                    G[i][1] += wx * gy * wz * b;
                    G[i][2] += wx * wy * gz * b;
                }
            }
}

template <typename S>
DTR_HD T3<S> tmatT(const S G[3][3], T3<S> v) {  // G^T v
    return {G[0][0] * v.x + G[1][0] * v.y + G[2][0] * v.z, G[0][1] * v.x + G[1][1] * v.y + G[2][1] * v.z,
            G[0][2] * v.x + G[1][2] * v.y + G[2][2] * v.z};
}

// ----------------------------------------------------------- RK4 step -----
// Stages of the accepted trial of rk_stepper::step at hs (the expressions of
// FastEngine::step, so that a replay reproduces the recorded trajectory).
template <typename S>
struct Stages {
    T3<S> d, B1, B2, B3, k1, k2, k3, k4, t2, t3, t4;
    V3 p1, p2;
};

template <typename S>
DTR_HD Stages<S> rk_stages(const DevField& field, V3 pos, V3 dirv, double qopd, double hs) {
    Stages<S> st;
    st.d = tcast<S>(dirv);
    const S qop = static_cast<S>(qopd);
    st.B1 = field_at_s<S>(field, pos);
    st.k1 = tmul(qop, tcross(st.d, st.B1));
    const V3 k1d = tdbl(st.k1);
    const double h = hs, h2 = h * h, half_h = h * 0.5;
    st.p1 = add(add(pos, mul(half_h, dirv)), mul(h2 * 0.125, k1d));
    st.B2 = field_at_s<S>(field, st.p1);
    const S hh = static_cast<S>(half_h), hS = static_cast<S>(h);
    st.t2 = (half_h == 0.0) ? st.d : tadd(st.d, tmul(hh, st.k1));
    st.k2 = tmul(qop, tcross(st.t2, st.B2));
    st.t3 = (half_h == 0.0) ? st.d : tadd(st.d, tmul(hh, st.k2));
    st.k3 = tmul(qop, tcross(st.t3, st.B2));
    st.p2 = add(add(pos, mul(h, dirv)), mul(h2 * 0.5, tdbl(st.k3)));
    st.B3 = field_at_s<S>(field, st.p2);
    st.t4 = (h == 0.0) ? st.d : tadd(st.d, tmul(hS, st.k3));
    st.k4 = tmul(qop, tcross(st.t4, st.B3));
    return st;
}

// advance_track (rk_stepper.hpp:594-629) at ha; returns |nd|
template <typename S>
DTR_HD double rk_advance(const Stages<S>& st, V3 pos, V3 dirv, double ha, V3& pos_out, V3& dir_out) {
    const V3 k1d = tdbl(st.k1);
    const double h_6 = ha * (1.0 / 6.0);
    const V3 ksum = add(add(k1d, tdbl(st.k2)), tdbl(st.k3));
    pos_out = add(pos, mul(ha, add(dirv, mul(h_6, ksum))));
    const V3 kd = add(add(k1d, mul(2.0, add(tdbl(st.k2), tdbl(st.k3)))), tdbl(st.k4));
    const V3 nd = add(dirv, mul(h_6, kd));
    const double nn = norm(nd);
    dir_out = {nd.x / nn, nd.y / nn, nd.z / nn};
    return nn;
}

// Vector-Jacobian product of one step.  On entry (pb, db) are the adjoints
// of the state after the step, on exit those of the state before it; qb
// accumulates the q/p adjoint.  Stage derivatives in S, accumulation in
// FP64.  exact: field gradient and normalisation included, stage
// derivatives at hs; otherwise the reference's Jacobian (no field
// gradient, no normalisation, stage derivatives at ha).
template <typename S>
DTR_HD void rk_vjp(const DevField& field, V3 pos, V3 dirv, double qopd, double hs, double ha, bool exact, V3& pb,
                   V3& db, double& qb) {
    const Stages<S> st = rk_stages<S>(field, pos, dirv, qopd, hs);
    const S q = static_cast<S>(qopd);
    V3 ndb = db;
    if (exact) {
        V3 po, dn;
        const double nn = rk_advance(st, pos, dirv, ha, po, dn);
        ndb = mul(1.0 / nn, sub(db, mul(dot(dn, db), dn)));
    }
    const double hd = exact ? hs : ha;
    const double h_6 = ha * (1.0 / 6.0);
    // nd = d + ha/6 (k1 + 2 k2 + 2 k3 + k4); p' = p + ha d + ha^2/6 (k1 + k2 + k3)
    const V3 dacc = add(ndb, mul(ha, pb));
    const V3 pacc = pb;
    const T3<S> ndbS = tcast<S>(ndb), pbS = tcast<S>(pb);
    const S a6 = static_cast<S>(h_6), a3 = static_cast<S>(2.0 * h_6), a66 = static_cast<S>(ha * h_6);
    T3<S> k1b = tadd(tmul(a6, ndbS), tmul(a66, pbS));
    T3<S> k2b = tadd(tmul(a3, ndbS), tmul(a66, pbS));
    T3<S> k3b = tadd(tmul(a3, ndbS), tmul(a66, pbS));
    const T3<S> k4b = tmul(a6, ndbS);
    const S hh = static_cast<S>(hd * 0.5), hS = static_cast<S>(hd);
    S G[3][3];
    T3<S> dbs{S(0), S(0), S(0)}, pbs{S(0), S(0), S(0)};
    S qbs = S(0);
    // k4 = q (t4 x B3), t4 = d + h k3, B3 = B(p2), p2 = p + h d + h^2/2 k3
    qbs += tdot(k4b, tcross(st.t4, st.B3));
    const T3<S> t4b = tmul(q, tcross(st.B3, k4b));
    dbs = tadd(dbs, t4b);
    k3b = tadd(k3b, tmul(hS, t4b));
    if (exact) {
        field_grad_s<S>(field, st.p2, G);
        const T3<S> p2b = tmatT(G, tmul(q, tcross(k4b, st.t4)));
        pbs = tadd(pbs, p2b);
        dbs = tadd(dbs, tmul(hS, p2b));
        k3b = tadd(k3b, tmul(static_cast<S>(hd * hd * 0.5), p2b));
    }
    // k3 = q (t3 x B2), t3 = d + h/2 k2
    qbs += tdot(k3b, tcross(st.t3, st.B2));
    const T3<S> t3b = tmul(q, tcross(st.B2, k3b));
    T3<S> B2b = tmul(q, tcross(k3b, st.t3));
    dbs = tadd(dbs, t3b);
    k2b = tadd(k2b, tmul(hh, t3b));
    // k2 = q (t2 x B2), t2 = d + h/2 k1
    qbs += tdot(k2b, tcross(st.t2, st.B2));
    const T3<S> t2b = tmul(q, tcross(st.B2, k2b));
    B2b = tadd(B2b, tmul(q, tcross(k2b, st.t2)));
    dbs = tadd(dbs, t2b);
    k1b = tadd(k1b, tmul(hh, t2b));
    // B2 = B(p1), p1 = p + h/2 d + h^2/8 k1
    if (exact) {
        field_grad_s<S>(field, st.p1, G);
        const T3<S> p1b = tmatT(G, B2b);
        pbs = tadd(pbs, p1b);
        dbs = tadd(dbs, tmul(hh, p1b));
        k1b = tadd(k1b, tmul(static_cast<S>(hd * hd * 0.125), p1b));
    }
    // k1 = q (d x B1), B1 = B(p)
    qbs += tdot(k1b, tcross(st.d, st.B1));
    dbs = tadd(dbs, tmul(q, tcross(st.B1, k1b)));
    if (exact) {
        field_grad_s<S>(field, pos, G);
        pbs = tadd(pbs, tmatT(G, tmul(q, tcross(k1b, st.d))));
    }
    pb = add(pacc, tdbl(pbs));
    db = add(dacc, tdbl(dbs));
    qb += static_cast<double>(qbs);
}

// ------------------------------------------------- surface-event adjoints --
DTR_HD V3 rec_col(const NavSurface& r, int k) { return {r.r[k], r.r[3 + k], r.r[6 + k]}; }

DTR_HD bool is_line_shape(uint8_t shape) { return shape == DTR_LINE_SQUARE || shape == DTR_LINE_CIRCULAR; }

// b_bar = J_B2F(b)^T (pos_bar, dir_bar, qop_bar): the adjoint of the setter
// (position and direction of the bound parameters b on surface r; the
// frames of bound_pos: cartesian, polar for rings and annuli, concentric
// and placed cylindrical, and line2D, whose position also depends on the
// direction, jacobian_line.hpp:102-172)
DTR_HD void b2f_transpose(const NavSurface& r, const double b[6], V3 pb, V3 db, double qb, double bb[6]) {
    const double cp = cos(b[2]), sp = sin(b[2]), ct = cos(b[3]), st = sin(b[3]);
    const V3 dphi{-sp * st, cp * st, 0.0}, dtheta{cp * ct, sp * ct, -st};
    V3 c0, c1;
    double ang[2] = {0.0, 0.0};
    if (r.shape == DTR_CONCENTRIC_CYLINDER2D) {
        const double R = r.m[0];
        c0 = {-R * sin(b[0]), R * cos(b[0]), 0.0};
        c1 = {0.0, 0.0, 1.0};
    } else if (r.shape == DTR_RING2D || r.shape == DTR_ANNULUS2D) {
        const V3 e0 = rec_col(r, 0), e1 = rec_col(r, 1);
        const double c = cos(b[1]), s = sin(b[1]);
        c0 = add(mul(c, e0), mul(s, e1));
        c1 = mul(b[0], sub(mul(c, e1), mul(s, e0)));
    } else if (r.shape == DTR_CYLINDER2D) {
        // pos = T (R cos(loc0/R), R sin(loc0/R), loc1)
        const double ph = b[0] / r.m[0];
        c0 = add(mul(-sin(ph), rec_col(r, 0)), mul(cos(ph), rec_col(r, 1)));
        c1 = rec_col(r, 2);
    } else if (is_line_shape(r.shape)) {
        // pos = t + loc1 a + loc0 x, x = (a x dir) / |a x dir|; dx/dangle =
        // (I - x x^T)(a x ddir/dangle) / |a x dir|
        const V3 a = rec_col(r, 2);
        const V3 w = cross(a, V3{cp * st, sp * st, ct});
        const double rho = sqrt(dot(w, w));
        const V3 x = mul(1.0 / rho, w);
        c0 = x;
        c1 = a;
        const V3 dd[2] = {dphi, dtheta};
        for (int i = 0; i < 2; ++i) {
            const V3 u = cross(a, dd[i]);
            ang[i] = (b[0] / rho) * dot(sub(u, mul(dot(x, u), x)), pb);
        }
    } else {
        c0 = rec_col(r, 0);
        c1 = rec_col(r, 1);
    }
    bb[0] = dot(c0, pb);
    bb[1] = dot(c1, pb);
    bb[2] = dot(dphi, db) + ang[0];
    bb[3] = dot(dtheta, db) + ang[1];
    bb[4] = qb;
    bb[5] = 0.0;
}

// (pos_bar, dir_bar, qop_bar) = J^T b_bar for the free-to-bound map F onto
// surface r at the free state (p, d).  exact: the derivative of the frame's
// global-to-local map as the engine evaluates it (a line's loc0 is the
// signed distance from the wire axis, a placed cylinder's loc0 is r phi of
// the point); otherwise the reference's free-to-bound rows (line and
// cylindrical frames of jacobian_line.hpp:38-62, jacobian_cylindrical.hpp:
// 24-45).  land: the surface the last step was sized to reach, or null.
// With a landing surface the position is first moved along the trajectory
// onto it (a change of the state before the step changes the step length):
// J = F (I + p2f f2p), p2f = (d, k, 0) with dt/ds = k, f2p = (-n / (n.d), 0,
// 0) with the surface normal n, for a line the path derivative of
// jacobian_line.hpp:64-92 (position and direction parts)
// (parameter_updater.hpp:463-514 with land = r).
DTR_HD void f2b_transpose(const NavSurface& r, V3 p, V3 d, V3 k, const NavSurface* land, bool exact,
                          const double bb[6], V3& pb, V3& db, double& qb) {
    V3 vp;
    const V3 tr{r.t[0], r.t[1], r.t[2]};
    if (r.shape == DTR_CONCENTRIC_CYLINDER2D) {
        const double rho2 = p.x * p.x + p.y * p.y;
        vp = {-p.y * bb[0] / rho2, p.x * bb[0] / rho2, bb[1]};
    } else if (r.shape == DTR_RING2D || r.shape == DTR_ANNULUS2D) {
        const V3 e0 = rec_col(r, 0), e1 = rec_col(r, 1);
        const V3 dl = sub(p, tr);
        const double lx = dot(e0, dl), ly = dot(e1, dl);
        const double rl2 = lx * lx + ly * ly, rl = sqrt(rl2);
        vp = add(mul(bb[0] / rl, add(mul(lx, e0), mul(ly, e1))), mul(bb[1] / rl2, sub(mul(lx, e1), mul(ly, e0))));
    } else if (r.shape == DTR_CYLINDER2D) {
        const V3 e0 = rec_col(r, 0), e1 = rec_col(r, 1), e2 = rec_col(r, 2);
        const V3 dl = sub(p, tr);
        const double lx = dot(e0, dl), ly = dot(e1, dl);
        const double rl = sqrt(lx * lx + ly * ly), ph = atan2(ly, lx);
        const V3 rad = add(mul(lx / rl, e0), mul(ly / rl, e1));  // normal at the point
        const V3 tan = add(mul(-ly / rl, e0), mul(lx / rl, e1));
        // exact: d(rl phi) = phi drl + rl dphi; the reference: the tangent
        vp = add(mul(bb[0], exact ? add(mul(ph, rad), tan) : tan), mul(bb[1], e2));
    } else if (is_line_shape(r.shape)) {
        const V3 a = rec_col(r, 2);
        V3 x;
        if (exact) {
            // loc0 = sign |(lx, ly)| of the local point, the sign of line2D's
            // global_to_local; the local axes are taken as stored (a rotation
            // built in single precision is orthonormal to 1e-7 only, which
            // the drift distance of a point 100 mm along the wire feels)
            const V3 e0 = rec_col(r, 0), e1 = rec_col(r, 1);
            const V3 dl = sub(p, tr);
            const double lx = dot(e0, dl), ly = dot(e1, dl);
            const double sign = dot(cross(a, d), sub(tr, p)) > 0.0 ? -1.0 : 1.0;
            x = mul(sign / sqrt(lx * lx + ly * ly), add(mul(lx, e0), mul(ly, e1)));
        } else {
            x = normalize(cross(a, d));
        }
        vp = add(mul(bb[0], x), mul(bb[1], a));
    } else {
        vp = add(mul(bb[0], rec_col(r, 0)), mul(bb[1], rec_col(r, 1)));
    }
    const double rt2 = d.x * d.x + d.y * d.y, rt = sqrt(rt2), dd = rt2 + d.z * d.z;
    const V3 vd = add(mul(bb[2], V3{-d.y / rt2, d.x / rt2, 0.0}),
                      mul(bb[3], V3{d.z * d.x / (rt * dd), d.z * d.y / (rt * dd), -rt / dd}));
    pb = vp;
    db = vd;
    if (land) {
        // (I + p2f f2p)^T v: position and direction adjoints gain
        // f2p (p2f . v)
        V3 fp, fd{0.0, 0.0, 0.0};
        const V3 lt{land->t[0], land->t[1], land->t[2]};
        if (is_line_shape(land->shape)) {
            const V3 a = rec_col(*land, 2);
            const V3 dl = sub(p, lt);
            const V3 pcx = sub(dl, mul(dot(dl, a), a));
            const double dz = dot(a, d);
            const double nrm = -1.0 / (1.0 - dz * dz + dot(pcx, k));
            fp = mul(nrm, sub(d, mul(dz, a)));
            fd = mul(nrm, pcx);
        } else {
            V3 n;
            if (land->shape == DTR_CONCENTRIC_CYLINDER2D) {
                const double rho = sqrt(p.x * p.x + p.y * p.y);
                n = {p.x / rho, p.y / rho, 0.0};
            } else if (land->shape == DTR_CYLINDER2D) {
                const V3 e0 = rec_col(*land, 0), e1 = rec_col(*land, 1);
                const V3 dl = sub(p, lt);
                const double lx = dot(e0, dl), ly = dot(e1, dl);
                const double ph = atan2(ly, lx);
                n = add(mul(cos(ph), e0), mul(sin(ph), e1));
            } else {
                n = rec_col(*land, 2);
            }
            fp = mul(-1.0 / dot(n, d), n);
        }
        const double s = dot(d, vp) + dot(k, vd);
        pb = add(vp, mul(s, fp));
        db = add(vd, mul(s, fd));
    }
    qb = bb[4];
}

// ------------------------------------------------------------ recording --
DTR_AI void tape_put_state(double* base, uint64_t slot, uint64_t n, uint32_t gid, V3 p, V3 d, double qop) {
    double* q = base + slot * n + gid;
    q[0 * n] = p.x;
    q[1 * n] = p.y;
    q[2 * n] = p.z;
    q[3 * n] = d.x;
    q[4 * n] = d.y;
    q[5 * n] = d.z;
    q[6 * n] = qop;
}

DTR_AI void tape_get_state(const double* base, uint64_t slot, uint64_t n, uint32_t gid, V3& p, V3& d, double& qop) {
    const double* q = base + slot * n + gid;
    p = {q[0 * n], q[1 * n], q[2 * n]};
    d = {q[3 * n], q[4 * n], q[5 * n]};
    qop = q[6 * n];
}

// parameter_updater<pointwise_material_interactor> (parameter_updater.hpp:
// 585-627, FastEngine::parameter_updater) with the event written to the
// tape, the Jacobian chain, chi2 and its Jacobian-chain gradient.
template <class P>
DTR_AI void record_update(const FastDetector& det, const DevField& field, const DevConfig& cfg, const AdjointIO& io,
                          FastTrack<P>& tk, CacheView c, uint32_t gid, uint32_t col, uint32_t& n_ev,
                          uint32_t& next_hit,
                          uint32_t& n_rec, double& chi2, uint32_t& status) {
    using S = typename P::S;
    FastEngine<P, FastTrack<P>, true> e{det, field, cfg, tk, c};
    const bool sens = e.on_sensitive(), mat = e.on_material();
    if (!(sens || mat)) return;
    const Tape& tp = io.tape;
    const uint64_t N = io.n;
    const uint32_t dest = e.cur_sf();
    const bool keep = n_ev < tp.max_events;
    if (!keep) status = kTapeEvents;
    const uint64_t ev = keep ? n_ev : 0u;
    if (keep) {
        const T3<S> k = e.dtds();
        double* q = tp.ev_pre + ev * 10u * N + col;
        q[0 * N] = tk.pos.x;
        q[1 * N] = tk.pos.y;
        q[2 * N] = tk.pos.z;
        q[3 * N] = tk.dir.x;
        q[4 * N] = tk.dir.y;
        q[5 * N] = tk.dir.z;
        q[6 * N] = tk.qop;
        q[7 * N] = static_cast<double>(k.x);
        q[8 * N] = static_cast<double>(k.y);
        q[9 * N] = static_cast<double>(k.z);
    }
    double jf[36];
    e.jf_out = io.grad_transport ? jf : nullptr;
    bool from_free = false, reused = false;
    if (tk.up_sf != kInvalid) {
        if (tk.up_sf != dest) {
            // the covariance is not an output of the adjoint; it is
            // transported when the navigation reads it (noise estimation)
            // or the Jacobian chain is asked for
            if (cfg.do_cov || io.grad_transport) e.transport_covariance(dest);
            if (fabs(tk.path_length) > 0.0) e.reset_J();
            e.free_to_bound(dest, tk.bound);
            tk.up_sf = dest;
            from_free = true;
        } else if (n_ev > 0u) {
            reused = true;
        }
    } else {
        status = kTapeUnsupported;  // free start: no bound parameters to differentiate
    }
    if (from_free && io.grad_transport) {
        // Jcum <- Jf Jcum
        double* J = tp.jcum + col;
        for (int j = 0; j < 6; ++j) {
            double jc[6], nc[6];
            for (int m = 0; m < 6; ++m) jc[m] = J[(m * 6 + j) * N];
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int m = 0; m < 6; ++m) s += jf[r * 6 + m] * jc[m];
                nc[r] = s;
            }
            for (int r = 0; r < 6; ++r) J[(r * 6 + j) * N] = nc[r];
        }
    }
    if (keep)
        for (int k = 0; k < 6; ++k) tp.ev_bound[(ev * 6u + k) * N + col] = tk.bound[k];
    int32_t hit = -1;
    if (sens && io.hits.surface) {
        const uint32_t nh = io.hits.n_hits[gid] < io.hits.max_hits ? io.hits.n_hits[gid] : io.hits.max_hits;
        for (uint32_t k = next_hit; k < nh; ++k) {
            if (io.hits.surface[static_cast<uint64_t>(k) * N + gid] == dest) {
                hit = static_cast<int32_t>(k);
                next_hit = k + 1u;
                break;
            }
        }
        if (hit >= 0) {
            const uint64_t mh = io.hits.max_hits;
            const double r0 = tk.bound[0] - io.hits.loc[(0u * mh + hit) * N + gid];
            const double r1 = tk.bound[1] - io.hits.loc[(1u * mh + hit) * N + gid];
            chi2 += io.iv0 * r0 * r0 + io.iv1 * r1 * r1;
            if (io.grad_transport) {
                const double w0 = 2.0 * io.iv0 * r0, w1 = 2.0 * io.iv1 * r1;
                for (int k = 0; k < 6; ++k)
                    io.grad_transport[k * N + gid] +=
                        tp.jcum[(0 * 6 + k) * N + col] * w0 + tp.jcum[(1 * 6 + k) * N + col] * w1;
            }
        }
    }
    if (sens && io.record.surface && n_rec < io.record.max_hits) {
        const uint64_t mh = io.record.max_hits;
        io.record.surface[static_cast<uint64_t>(n_rec) * N + gid] = dest;
        io.record.loc[(0u * mh + n_rec) * N + gid] = tk.bound[0];
        io.record.loc[(1u * mh + n_rec) * N + gid] = tk.bound[1];
        n_rec++;
    }
    if (mat) e.material_interaction(dest);
    if (tk.stopped) status = kTapeUnsupported;
    if (keep) {
        tape_put_state(tp.ev_post, ev * 7u, N, col, tk.pos, tk.dir, tk.qop);
        tp.ev_step[ev * N + col] = tk.n_steps;
        tp.ev_sf[ev * N + col] = dest;
        tp.ev_hit[ev * N + col] = hit;
        tp.ev_flags[ev * N + col] =
            (from_free ? kEvFromFree : 0u) | (mat ? kEvMaterial : 0u) | (reused ? kEvReused : 0u);
        n_ev++;
    }
}

// propagator::propagate (propagator.hpp:262-400) for one track from its
// bound start with the tape written; the surface sequencer is replaced by
// the event list.
template <class P>
DTR_AI void adjoint_record(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                           const AdjointIO& io, uint32_t gid, uint32_t col, CacheView c) {
    const Tape& tp = io.tape;
    const uint64_t N = io.n;
    const TrackInputs in{nullptr, nullptr, io.bound0, io.start_sf, N};
    if (io.grad_transport) {
        for (int k = 0; k < 6; ++k) io.grad_transport[k * N + gid] = 0.0;
        for (int k = 0; k < 36; ++k) tp.jcum[k * N + col] = (k % 7 == 0) ? 1.0 : 0.0;
    }
    Lane<P> lane;
    lane.start(det, field, cfg, in, c, gid);
    FastTrack<P>& tk = lane.tk;
    uint32_t n_ev = 0u, next_hit = 0u, n_rec = 0u, status = kTapeOk;
    double chi2 = 0.0;
    while (true) {
        // actor half (Lane::actor_half)
        if (tk.status != kExit) {
            const double step_limit = cfg.path_limit - fabs(tk.abs_path_length);
            if (step_limit <= 0.0) {
                tk.status = kAbort;
                tk.path_limit_hit = true;
                tk.heartbeat = false;
            }
            tk.constraint_abort = dmin(tk.constraint_abort, fabs(step_limit));
        }
        if (cfg.use_material)
            record_update<P>(det, field, cfg, io, tk, c, gid, col, n_ev, next_hit, n_rec, chi2, status);
        if (tk.stopped && tk.status > kUnknown) {
            tk.status = kAbort;
            tk.heartbeat = false;
        }
        if (!tk.heartbeat) break;
        lane.last_path = tk.path_length;
        lane.navigate_alone(det, cfg, c);
        if (lane.nav_end()) break;
        // step half with the step sizes and checkpoints on the tape
        const uint32_t s0 = tk.n_steps;
        if (s0 % tp.every == 0u && s0 / tp.every < tp.n_ckpt)
            tape_put_state(tp.ckpt, static_cast<uint64_t>(s0 / tp.every) * 7u, N, col, tk.pos, tk.dir, tk.qop);
        // a step whose length is the navigator's distance lands on the
        // navigator's target surface
        const double dist = c.p(tk.next);
        const uint32_t target = c.s(tk.next);
        lane.step_half(det, field, cfg, c);
        if (tk.n_steps != s0) {
            if (s0 < tp.max_steps) {
                tp.h[(2ull * s0 + 0u) * N + col] = tk.h_stage;
                tp.h[(2ull * s0 + 1u) * N + col] = tk.step_size;
                tp.step_land[static_cast<uint64_t>(s0) * N + col] = (tk.step_size == dist) ? target : kInvalid;
            } else {
                status = kTapeSteps;
            }
        }
        lane.navigate_alone(det, cfg, c);
        lane.nav_end();
    }
    tp.n_steps[col] = tk.n_steps;
    tp.n_events[col] = n_ev;
    tp.status[col] = status;
    if (io.status) io.status[gid] = status;
    if (io.chi2) io.chi2[gid] = chi2;
    if (io.record.surface) io.record.n_hits[gid] = n_rec;
}

// -------------------------------------------------------------- backward --
// Straight-line distance s from (p, d) to the target surface r with the
// arithmetic of the navigator's candidate path (FastEngine::plane_rec,
// line_intersect, cyl2_intersect, cyl_intersect), and its gradient (gp, gd)
// with respect to p and d.  Cylinders have two solutions: the one nearest to
// `near`.
// Planes: s = n.(t - p) / (n.d).  Lines (closest approach to the wire axis
// a): s = ((t - p).d - ((t - p).a)(a.d)) / (1 - (a.d)^2).  Cylinders: the
// root of |P (p - t + s d)|^2 = R^2, P the projection across the axis, with
// the gradient of the implicit function, -P u / (P u . d) (1, s),
// u = p - t + s d.
DTR_HD bool target_path(const NavSurface& r, V3 p, V3 d, double near, double& s, V3& gp, V3& gd) {
    const V3 tr{r.t[0], r.t[1], r.t[2]};
    if (is_line_shape(r.shape)) {
        const V3 a = rec_col(r, 2);
        const double c = dot(a, d);
        const double den = 1.0 - c * c;
        if (den < static_cast<double>(1e-5f)) return false;
        const V3 t2l = sub(tr, p);
        const double on_line = dot(t2l, a);
        s = 1.0 / den * (dot(t2l, d) - on_line * c);
        gp = mul(-1.0 / den, sub(d, mul(c, a)));
        gd = mul(1.0 / den, add(sub(t2l, mul(on_line, a)), mul(2.0 * c * s, a)));
        return true;
    }
    if (r.shape == DTR_CONCENTRIC_CYLINDER2D) {
        // the portal intersector (ray_cylinder_portal_intersector.hpp:68-132):
        // within 1e-3 of R^2 its path is R^2 - rho^2 itself
        const double R = r.m[0];
        const double rd2 = d.x * d.x + d.y * d.y;
        if (rd2 < 2.220446049250313e-16) return false;
        const double rad_diff = R * R - (p.x * p.x + p.y * p.y);
        if (!(fabs(rad_diff) > 1.0 * 1e-3)) {
            s = rad_diff;
            gp = {-2.0 * p.x, -2.0 * p.y, 0.0};
            gd = {0.0, 0.0, 0.0};
            return true;
        }
        const double inv2 = 1.0 / rd2;
        const double kk = -inv2 * (p.x * d.x + p.y * d.y);
        const double discr = inv2 * rad_diff + kk * kk;
        if (discr < 0.0) return false;
        const double sq = sqrt(discr);
        s = fabs(kk + sq - near) <= fabs(kk - sq - near) ? kk + sq : kk - sq;
        const V3 pu{p.x + s * d.x, p.y + s * d.y, 0.0};
        const double den = dot(pu, d);
        gp = mul(-1.0 / den, pu);
        gd = mul(-s / den, pu);
        return true;
    }
    if (r.shape == DTR_CYLINDER2D) {
        const V3 a = rec_col(r, 2);
        double sol[2];
        dtr_transform t;
        for (int i = 0; i < 9; ++i) t.r[i] = r.r[i];
        for (int i = 0; i < 3; ++i) t.t[i] = r.t[i];
        const int ns = Engine::cyl2_intersect(p, d, t, r.m[0], sol);
        if (ns == 0) return false;
        s = (ns == 2 && fabs(sol[1] - near) < fabs(sol[0] - near)) ? sol[1] : sol[0];
        const V3 u = add(sub(p, tr), mul(s, d));
        const V3 pu = sub(u, mul(dot(u, a), a));
        const double den = dot(pu, d);
        gp = mul(-1.0 / den, pu);
        gd = mul(-s / den, pu);
        return true;
    }
    const V3 n = rec_col(r, 2);
    const double nd = dot(n, d);
    if (nd == 0.0) return false;
    s = dot(n, sub(tr, p)) / nd;
    gp = mul(-1.0 / nd, n);
    gd = mul(-s / nd, n);
    return true;
}

// A step sized by the navigator has the distance to its target surface as
// its length, h = s(p_c, d_c) (target_path), computed at an earlier state
// (p_c, d_c): usually the step's start, but the distance is not refreshed
// after a setter of a surface reached without moving.  hbar, the adjoint of
// h, is dx_end/dh . lambda_end with lambda_end the adjoint of the state the
// step ended in; (p_c, d_c) receive hbar (ds/dp_c, ds/dd_c).  Every such
// step counts, also one that ends on a portal or short of its target: the
// next step starts where it ended, and a wire's closest approach from a
// point of the trajectory moves with the direction there (by loc0 times
// the change of angle), so the dependence does not cancel.
//
// dx_end/dh: the tangent of the trajectory at the end of the step, (d,
// q/p d x B), when the stages were taken at h (the Runge-Kutta map follows
// the trajectory to the integration tolerance); the derivative of the
// advance at fixed stages when the step was clipped after its stages were
// taken at hs != h (rk_stepper.hpp:474-511).
template <typename S>
DTR_HD void step_tangent(const DevField& field, V3 pos, V3 dirv, double qopd, double hs, double ha, V3& td, V3& tk) {
    const Stages<S> st = rk_stages<S>(field, pos, dirv, qopd, hs);
    V3 pe, de;
    const double nn = rk_advance(st, pos, dirv, ha, pe, de);
    if (hs == ha) {
        const T3<double> B = field_at_s<double>(field, pe);
        td = de;
        tk = mul(qopd, cross(de, V3{B.x, B.y, B.z}));
        return;
    }
    const V3 k1 = tdbl(st.k1), k2 = tdbl(st.k2), k3 = tdbl(st.k3), k4 = tdbl(st.k4);
    td = add(dirv, mul(ha * (1.0 / 3.0), add(add(k1, k2), k3)));
    const V3 dnd = mul(1.0 / 6.0, add(add(k1, mul(2.0, add(k2, k3))), k4));
    tk = mul(1.0 / nn, sub(dnd, mul(dot(de, dnd), de)));
}

struct Pending {
    int32_t ev;  // event whose state is (p_c, d_c), -1: none
    bool post;   // its state after (true) or before (false) the event
    V3 pc, dc;   // adjoint contribution to add there
};
// Adjoint of event ev: (pb, db, qb) enter as the adjoint of the state after
// the event and leave as that of the state before it.  Returns false for
// the start event (its bound parameters are the initial ones), whose bound
// adjoint is written to grad.  pbb: the adjoint of the bound parameters
// (after material) of the last bound event, collected from later events
// that reused them (kEvReused) and handed to that event.
DTR_AI bool event_adjoint(const FastDetector& det, const DevField& field, const DevConfig& cfg, const AdjointIO& io,
                          bool exact, uint32_t gid, uint32_t col, uint64_t ev, double mass, double charge, V3& pb,
                          V3& db, double& qb, double grad[6], double pbb[6]) {
    const Tape& tp = io.tape;
    const uint64_t N = io.n;
    const uint32_t sfi = tp.ev_sf[ev * N + col];
    const uint32_t flags = tp.ev_flags[ev * N + col];
    const int32_t hit = tp.ev_hit[ev * N + col];
    const NavSurface& rec = det.sf[sfi];
    double b[6], bp[6], bb[6];
    for (int k = 0; k < 6; ++k) b[k] = tp.ev_bound[(ev * 6u + k) * N + col];
    for (int k = 0; k < 6; ++k) bp[k] = b[k];
    bp[4] = tp.ev_post[(ev * 7u + 6u) * N + col];
    const bool mat = (flags & kEvMaterial) != 0u;
    const bool reused = (flags & kEvReused) != 0u;
    // the state after the event is B2F(bound): after the setter of a
    // material event, and in the transport model after every event that
    // binds anew (a reused event continues the reference's Jacobian chain)
    const bool restart = reused ? (mat && exact) : (mat || !exact);
    V3 cpb{0.0, 0.0, 0.0}, cdb{0.0, 0.0, 0.0};
    double cqb = 0.0;
    for (int k = 0; k < 6; ++k) bb[k] = 0.0;
    if (restart) {
        b2f_transpose(rec, bp, pb, db, qb, bb);
    } else {
        cpb = pb;
        cdb = db;
        cqb = qb;
    }
    if (!reused) {
        // later events that reused these bound parameters (after material)
        for (int k = 0; k < 6; ++k) {
            bb[k] += pbb[k];
            pbb[k] = 0.0;
        }
    }
    if (mat && exact && cfg.do_eloss) {
        // the slab of the bin the bound position is in (a map's bins are
        // piecewise constant in loc0, loc1)
        const dtr_material_slab& m = surface_material(det.materials, det.mat_maps, rec.material, b[0], b[1]);
        if (m.thickness > 2.220446049250313e-16 && b[4] != 0.0) {
            const D4 q2 = eloss_qop(rec, m, b, mass, charge);
            const double g4 = bb[4];
            bb[0] += g4 * q2.d[0];
            bb[2] += g4 * q2.d[1];
            bb[3] += g4 * q2.d[2];
            bb[4] = g4 * q2.d[3];
        }
    }
    if (hit >= 0 && io.hits.surface) {
        const uint64_t mh = io.hits.max_hits;
        bb[0] += 2.0 * io.iv0 * (b[0] - io.hits.loc[(0u * mh + hit) * N + gid]);
        bb[1] += 2.0 * io.iv1 * (b[1] - io.hits.loc[(1u * mh + hit) * N + gid]);
    }
    if (reused) {
        // the free state before the event is discarded by a setter (or
        // passes through); the bound adjoint belongs to the last bound event
        for (int k = 0; k < 6; ++k) pbb[k] += bb[k];
        pb = cpb;
        db = cdb;
        qb = cqb;
        return true;
    }
    if (!(flags & kEvFromFree)) {
        // start event: b is the initial bound vector
        double cb[6];
        b2f_transpose(rec, b, cpb, cdb, cqb, cb);
        for (int k = 0; k < 6; ++k) grad[k] = bb[k] + cb[k];
        return false;
    }
    const double* q = tp.ev_pre + ev * 10u * N + col;
    const V3 p{q[0 * N], q[1 * N], q[2 * N]}, d{q[3 * N], q[4 * N], q[5 * N]};
    const V3 k{q[7 * N], q[8 * N], q[9 * N]};
    // exact: a plain projection (the length of the step that reached the
    // surface carries the dependence on the earlier state); transport: the
    // reference's path correction onto this surface
    V3 fpb, fdb;
    double fqb;
    f2b_transpose(rec, p, d, k, exact ? nullptr : &rec, exact, bb, fpb, fdb, fqb);
    pb = add(cpb, fpb);
    db = add(cdb, fdb);
    qb = cqb + fqb;
    return true;
}

// event_adjoint with a pending step-length contribution added at the state
// it belongs to (after the event: before its adjoint; before the event:
// after it, through the bound-to-free map for the start event).
DTR_AI bool event_step_adjoint(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                               const AdjointIO& io, bool exact, uint32_t gid, uint32_t col, int32_t e, double mass,
                               double charge,
                               V3& pb, V3& db, double& qb, double grad[6], double pbb[6], Pending& pend,
                               const double b0[6]) {
    if (pend.ev == e && pend.post) {
        pb = add(pb, pend.pc);
        db = add(db, pend.dc);
        pend.ev = -1;
    }
    const bool more = event_adjoint(det, field, cfg, io, exact, gid, col, static_cast<uint64_t>(e), mass, charge, pb,
                                    db, qb, grad, pbb);
    if (pend.ev == e && !pend.post) {
        if (more) {
            pb = add(pb, pend.pc);
            db = add(db, pend.dc);
        } else {
            double cb[6];
            b2f_transpose(det.sf[io.start_sf[gid]], b0, pend.pc, pend.dc, 0.0, cb);
            for (int k = 0; k < 6; ++k) grad[k] += cb[k];
        }
        pend.ev = -1;
    }
    return more;
}

template <class P>
DTR_AI void adjoint_backward(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                             const AdjointIO& io, bool exact, uint32_t gid, uint32_t col, double* grad_out) {
    using S = typename P::S;
    const Tape& tp = io.tape;
    const uint64_t N = io.n;
    double grad[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (tp.status[col] != kTapeOk) {
        const double nan = __builtin_nan("");
        for (int k = 0; k < 6; ++k) grad_out[k * N + gid] = nan;
        return;
    }
    const uint32_t n_steps = tp.n_steps[col];
    const int32_t n_ev = static_cast<int32_t>(tp.n_events[col]);
    double b0[6];
    for (int k = 0; k < 6; ++k) b0[k] = io.bound0[k * N + gid];
    // muon hypothesis, charge-conjugated to the track (as Lane::start_begin)
    const double charge = (-1.0 * b0[4] > 0.0) ? -1.0 : 1.0;
    const double mass = static_cast<double>(105.6583755f * 1e-3f);
    V3 pb{0.0, 0.0, 0.0}, db{0.0, 0.0, 0.0};
    double qb = 0.0;
    int32_t e = n_ev - 1;
    bool open = true;  // the start event has not been reached
    Pending pend{-1, false, V3{0.0, 0.0, 0.0}, V3{0.0, 0.0, 0.0}};
    double pbb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t K = tp.every;
    const uint32_t n_seg = (n_steps + K - 1u) / K;
    for (int32_t sgi = static_cast<int32_t>(n_seg) - 1; sgi >= 0; --sgi) {
        const uint32_t j0 = static_cast<uint32_t>(sgi) * K;
        const uint32_t j1 = (j0 + K < n_steps) ? j0 + K : n_steps;
        // replay the segment from its checkpoint: the state before each step
        V3 p, d;
        double qop;
        tape_get_state(tp.ckpt, static_cast<uint64_t>(sgi) * 7u, N, col, p, d, qop);
        int32_t ef = e;
        while (ef >= 0 && tp.ev_step[static_cast<uint64_t>(ef) * N + col] > j0) ef--;
        ef++;
        for (uint32_t j = j0 + 1u; j <= j1; ++j) {
            tape_put_state(tp.seg, static_cast<uint64_t>(j - j0 - 1u) * 7u, N, col, p, d, qop);
            if (j == j1) break;
            const double hs = tp.h[(2ull * (j - 1u) + 0u) * N + col];
            const double ha = tp.h[(2ull * (j - 1u) + 1u) * N + col];
            const Stages<S> st = rk_stages<S>(field, p, d, qop, hs);
            V3 pn, dn;
            rk_advance(st, p, d, ha, pn, dn);
            p = pn;
            d = dn;
            while (ef < n_ev && tp.ev_step[static_cast<uint64_t>(ef) * N + col] == j) {
                tape_get_state(tp.ev_post, static_cast<uint64_t>(ef) * 7u, N, col, p, d, qop);
                ef++;
            }
        }
        // back through the segment: the events after step j, then step j
        for (uint32_t j = j1; j > j0; --j) {
            while (e >= 0 && tp.ev_step[static_cast<uint64_t>(e) * N + col] == j) {
                event_step_adjoint(det, field, cfg, io, exact, gid, col, e, mass, charge, pb, db, qb, grad, pbb,
                                   pend, b0);
                e--;
            }
            V3 sp, sd;
            double sq;
            tape_get_state(tp.seg, static_cast<uint64_t>(j - j0 - 1u) * 7u, N, col, sp, sd, sq);
            const double hs = tp.h[(2ull * (j - 1u) + 0u) * N + col];
            const double ha = tp.h[(2ull * (j - 1u) + 1u) * N + col];
            // a step sized by the navigator: hbar from the adjoint of the
            // state it ended in (complete once the events after it are done)
            const uint32_t tgt = exact ? tp.step_land[static_cast<uint64_t>(j - 1u) * N + col] : kInvalid;
            double hbar = 0.0;
            if (tgt != kInvalid) {
                V3 td, tk;
                step_tangent<S>(field, sp, sd, sq, hs, ha, td, tk);
                hbar = dot(td, pb) + dot(tk, db);
            }
            rk_vjp<S>(field, sp, sd, sq, hs, ha, exact, pb, db, qb);
            if (tgt != kInvalid) {
                // the state the step length was computed from: the step's
                // start, or a state of the events after the previous step
                const NavSurface& tr = det.sf[tgt];
                V3 gp, gd;
                // whether (p, d) gives the step length; its gradient in gp, gd
                auto same = [&](V3 p, V3 d) {
                    double s;
                    return target_path(tr, p, d, ha, s, gp, gd) && fabs(s - ha) <= 1e-9 * fabs(ha) + 1e-15;
                };
                bool found = same(sp, sd);
                if (found) {
                    pb = add(pb, mul(hbar, gp));
                    db = add(db, mul(hbar, gd));
                }
                for (int32_t e2 = e; !found && e2 >= 0 && tp.ev_step[static_cast<uint64_t>(e2) * N + col] == j - 1u;
                     --e2) {
                    for (int w = 0; w < 2 && !found; ++w) {
                        V3 p, d;
                        double qq;
                        if (w == 0) {
                            tape_get_state(tp.ev_post, static_cast<uint64_t>(e2) * 7u, N, col, p, d, qq);
                        } else {
                            const double* q = tp.ev_pre + static_cast<uint64_t>(e2) * 10u * N + col;
                            p = {q[0 * N], q[1 * N], q[2 * N]};
                            d = {q[3 * N], q[4 * N], q[5 * N]};
                        }
                        if (same(p, d)) {
                            found = true;
                            pend.ev = e2;
                            pend.post = (w == 0);
                            pend.pc = mul(hbar, gp);
                            pend.dc = mul(hbar, gd);
                        }
                    }
                }
                if (!found) {
                    // not identified: taken from the step's start
                    double s;
                    if (target_path(tr, sp, sd, ha, s, gp, gd)) {
                        pb = add(pb, mul(hbar, gp));
                        db = add(db, mul(hbar, gd));
                    }
                }
            }
        }
    }
    while (e >= 0 && open) {
        open = event_step_adjoint(det, field, cfg, io, exact, gid, col, e, mass, charge, pb, db, qb, grad, pbb,
                                  pend, b0);
        e--;
    }
    if (open) {
        // no event on the start surface: the initial state is B2F(b0), and
        // events that reused the start parameters add their adjoint
        b2f_transpose(det.sf[io.start_sf[gid]], b0, pb, db, qb, grad);
        for (int k = 0; k < 6; ++k) grad[k] += pbb[k];
    }
    for (int k = 0; k < 6; ++k) grad_out[k * N + gid] = grad[k];
}

}  // namespace dtr

// Direct navigator (direct_navigator.hpp:29-298) on the batched engine: the
// track follows a given sequence of surfaces instead of searching the
// detector.  Candidate cache entry 0 holds the current surface, entry 1 the
// target; targets are intersected with infinite mask tolerances and no
// overstepping limit, so the target is always hit.
#pragma once

#include "device/engine.hpp"

namespace dtr {

// Surface sequences of a batch: surfaces [max][n] (index k of track i at
// k * n + i), counts [n].  backward: the targets are taken from the last
// entry to the first (direct_navigator.hpp:79-94, 125-151 with
// navigation::direction::e_backward).
struct SeqView {
    const uint32_t* sf;
    const uint32_t* count;
    uint32_t max;
    uint32_t n_surfaces;  // surfaces in the detector: larger indices abort
    uint32_t backward;
};

// The k-th target of track i's sequence of len entries.
DTR_AI uint32_t seq_target(const SeqView& sq, uint64_t N, uint32_t track, uint32_t k, uint32_t len) {
    const uint32_t e = sq.backward ? len - 1u - k : k;
    return sq.sf[static_cast<uint64_t>(e) * N + track];
}

// The intersection config of direct_navigator::update (direct_navigator.hpp:
// 233-236): {inf, inf, inf, path_tolerance, -inf}.
inline DevConfig direct_config(const DevConfig& cfg) {
    DevConfig d = cfg;
    d.min_mask_tol = 3.40282347e+38f;
    d.max_mask_tol = 3.40282347e+38f;
    d.mask_scalor = 3.40282347e+38f;
    d.overstep_tol = -3.40282347e+38f;
    return d;
}

// direct_navigator::update (direct_navigator.hpp:210-298); returns true when
// the stepper has to reset its step size (a surface was reached).
template <class P>
DTR_AI bool direct_update(const FastDetector& det, const DevConfig& cfg, const DevConfig& dcfg, FastTrack<P>& tk,
                          CacheView c, const SeqView& sq, uint64_t N, uint32_t track, uint32_t& next_ext,
                          uint32_t len) {
    if (next_ext >= len) {
        tk.status = kExit;
        return false;
    }
    if (tk.trust == kFull) return false;
    const DevField no_field{};
    FastEngine<P> e{det, no_field, dcfg, tk, c};
    Cand cd;
    // update_candidate on the target; an unreachable target (or an index
    // outside the detector) ends the navigation
    if (c.s(1) >= sq.n_surfaces || !e.evaluate(det.sf[c.s(1)], dcfg.overstep_tol, cd)) {
        tk.status = kAbort;
        return false;
    }
    c.set(1, cd);
    // navigation::update_status (navigation_functions.hpp:148-177) with the
    // direct state's advance (direct_navigator.hpp:129-151)
    if (fabs(c.p(1)) < static_cast<double>(cfg.path_tol)) {
        tk.status = (c.sf_id(1) == DTR_PORTAL) ? kOnPortal : kOnObject;
        c.copy(0, 1);
        next_ext++;
        if (next_ext < len) c.s(1) = seq_target(sq, N, track, next_ext, len);
    } else {
        tk.status = kTowards;
    }
    tk.trust = kFull;
    if (tk.status == kTowards) return false;
    if (next_ext < len) {
        if (c.s(1) >= sq.n_surfaces || !e.evaluate(det.sf[c.s(1)], dcfg.overstep_tol, cd)) {
            tk.status = kAbort;
            return false;
        }
        c.set(1, cd);
    }
    tk.volume = det.sf[c.s(0)].volume;
    tk.trust = kFull;
    return true;
}

// propagator::propagate (propagator.hpp:262-400) with the direct navigator
// for one track, one lane on its own.
template <class P>
DTR_AI void propagate_direct(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                             const DevConfig& dcfg, const TrackInputs& in, const DevResults& out, const SeqView& sq,
                             uint32_t track, CacheView c) {
    Lane<P> lane;
    float os;
    lane.start_begin(det, field, cfg, in, c, track, os);
    FastTrack<P>& tk = lane.tk;
    const uint64_t N = in.n;
    const uint32_t cnt = sq.count[track];
    const uint32_t len = cnt < sq.max ? cnt : sq.max;
    uint32_t next_ext = 0u;
    if (sq.backward) tk.nav_dir = -1.0;
    // direct_navigator::init (direct_navigator.hpp:196-206): target = first
    // surface of the sequence in the navigation direction, next = last = 1
    tk.next = 1;
    tk.last = 1;
    if (len > 0u) c.s(1) = seq_target(sq, N, track, 0u, len);
    direct_update<P>(det, cfg, dcfg, tk, c, sq, N, track, next_ext, len);
    lane.start_end();
    for (;;) {
        if (lane.actor_half(det, field, cfg, in, out, c)) break;
        lane.nav_ret = direct_update<P>(det, cfg, dcfg, tk, c, sq, N, track, next_ext, len);
        if (lane.nav_end()) break;
        lane.step_half(det, field, cfg, c);
        lane.nav_ret = direct_update<P>(det, cfg, dcfg, tk, c, sq, N, track, next_ext, len);
        lane.nav_end();
    }
    lane.finish(out, N, cfg);
}

}  // namespace dtr

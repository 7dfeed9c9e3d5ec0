// Wave-cooperative navigation and the kernel bodies built on it.
//
// A local navigation (navigation_functions.hpp:193-246) evaluates every
// portal/passive of the volume and every surface in the grid window around
// the track: ~46 candidates in the toy detector, ~42 of them grid bins of
// which the sphere prefilter passes ~12.  Lanes of a wave need it at
// different iterations of the propagation loop, so a lane-by-lane search
// runs with few lanes active and pays for the lane with the most prefilter
// passes in every grid row.  Here the grid bins of all requesting lanes are
// pooled: the wave runs the sphere test on 64 (lane, bin) items per round
// and the exact intersection + mask check on 64 passing items per round,
// every lane working for whichever lane owns the item, and each owner
// inserts its accepted candidates in the original order, which is what
// makes the result identical to the sequential search.
//
// W is the wave interface: lane(), shfl(), ballot(), any(), sync() (a
// wave-level memory fence for LDS exchange) and atomic_add().
#pragma once

#include "device/engine.hpp"

namespace dtr {

// Per-wave LDS of the cooperative search: 64 result slots and a FIFO of
// (owner lane, grid bin) entries that passed the sphere test.
constexpr uint32_t kFifo = 128;
struct alignas(16) CoopSlot {
    double path;
    uint32_t sf;
    uint32_t ls;  // volume link | status << 16 | surface id << 18 | owner lane << 24
};
struct CoopShared {
    CoopSlot slot[64];
    uint32_t fifo[kFifo];  // bin | owner << 26
};

// Cooperative local navigation.  Every lane of the wave calls it (convergent
// control flow); lanes with req == true get their cache rebuilt for
// (volume, pos, dir, ext_tol) with overstep tolerance `os` and receive
// next | last << 8 | status << 16 | trust << 24.
//
// The brute-force list (portals, passives) is searched by each requesting
// lane on its own: lanes of one volume walk the same list in step.  The grid
// window is pooled: sphere-test rounds over all (lane, bin) items, the
// passing ones appended in item order to a FIFO, and evaluation rounds of 64
// FIFO entries at a time, whose accepted candidates each owner inserts in
// FIFO order.  Insertion order is therefore the sequential one.
template <class P, class W>
__host__ __device__ __attribute__((noinline)) uint32_t coop_local_navigation(
    const FastDetector& det, const DevConfig& cfg, bool req, V3 pos, V3 dir, double ext_tol,
    uint16_t volume, float os, CacheView c, CoopShared* sh) {
    const uint32_t lane = W::lane();
    const uint64_t lt_mask = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    const DevField no_field{};
    NavLite nl;
    nl.ext_tol = ext_tol;
    nl.next = 0;
    nl.last = -1;
    nl.status = kUnknown;
    nl.trust = kNoTrust;
    nl.volume = volume;
    nl.pos = pos;
    nl.dir = dir;
    FastEngine<P, NavLite> me{det, no_field, cfg, nl, c};

    // brute-force surfaces, each lane on its own; then the grid window
    uint32_t ng = 0u, cap = 1u;
    int a0 = 0, a1 = 0, width = 1, grid = -1;
    float margin = 0.f, ft = 0.f;
    if (req) {
        for (int i = 0; i < kCache; ++i) c.p(i) = kMax;
        const dtr_volume& vol = det.volumes[volume];
        me.bf_search(vol, os);
        if (vol.grid >= 0) {
            int e0, e1;
            if (me.grid_window(vol, det.grids[vol.grid], a0, e0, a1, e1)) {
                width = e1 - a1;
                // items are the entries of the window's bins (capacity 0 means 1)
                if constexpr (P::kExt) cap = det.grids[vol.grid].capacity ? det.grids[vol.grid].capacity : 1u;
                ng = static_cast<uint32_t>((e0 - a0) * width) * cap;
                grid = vol.grid;
                margin = me.sphere_margin(vol.grid);
                ft = static_cast<float>(me.tol_max());
            }
        }
    }
    uint32_t incl = ng;
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        const uint32_t v = W::shfl(incl, lane >= d ? static_cast<int>(lane - d) : static_cast<int>(lane));
        if (lane >= d) incl += v;
    }
    const uint32_t excl = incl - ng;
    const uint32_t total = W::shfl(incl, 63);
    const float fpx = static_cast<float>(pos.x), fpy = static_cast<float>(pos.y), fpz = static_cast<float>(pos.z);
    const float fdx = static_cast<float>(dir.x), fdy = static_cast<float>(dir.y), fdz = static_cast<float>(dir.z);

    uint32_t head = 0u, tail = 0u;  // wave-uniform FIFO positions
    for (uint32_t base = 0u; base < total || tail > head; base += 64u) {
        if (base < total) {
            // sphere round: item -> owner lane and bin
            const uint32_t item = base + lane;
            int lo = 0, hi = 63;
            for (int it = 0; it < 6; ++it) {
                const int mid = (lo + hi) >> 1;
                const uint32_t v = W::shfl(incl, mid);
                if (lo < hi) {
                    if (v > item) hi = mid;
                    else lo = mid + 1;
                }
            }
            const int o = lo;
            const uint32_t oexcl = W::shfl(excl, o);
            const int oa0 = W::shfl(a0, o), oa1 = W::shfl(a1, o), owidth = W::shfl(width, o);
            const int ogrid = W::shfl(grid, o);
            const uint32_t ocap = P::kExt ? W::shfl(cap, o) : 1u;
            const float omargin = W::shfl(margin, o), oft = W::shfl(ft, o);
            const float ox = W::shfl(fpx, o), oy = W::shfl(fpy, o), oz = W::shfl(fpz, o);
            const float dx = W::shfl(fdx, o), dy = W::shfl(fdy, o), dz = W::shfl(fdz, o);
            bool pass = false;
            uint32_t bin = 0u;
            if (item < total) {
                const dtr_grid& g = det.grids[ogrid];
                const int jj = static_cast<int>((item - oexcl) / ocap);
                const uint32_t entry = (item - oexcl) % ocap;
                const int b0 = oa0 + jj / owidth, b1 = oa1 + jj % owidth;
                const bool cyl = g.type == DTR_GRID_CYLINDER;
                const int n0 = static_cast<int>(g.n0), n1 = static_cast<int>(g.n1);
                const int i0 = cyl ? FastEngine<P, NavLite>::wrap_bin(b0, n0) : b0;
                const int i1 = cyl ? b1 : FastEngine<P, NavLite>::wrap_bin(b1, n1);
                bin = g.bins + (static_cast<uint32_t>(i0) * g.n1 + static_cast<uint32_t>(i1)) * ocap + entry;
                pass = FastEngine<P, NavLite>::prefilter(det.bin_rec[bin].box, ox, oy, oz, dx, dy, dz, omargin,
                                                         oft);
            }
            const uint64_t pm = W::ballot(pass);
            if (pass) {
                const uint32_t at = tail + static_cast<uint32_t>(__builtin_popcountll(pm & lt_mask));
                sh->fifo[at % kFifo] = bin | (static_cast<uint32_t>(o) << 26);
            }
            tail += static_cast<uint32_t>(__builtin_popcountll(pm));
            W::sync();
        }
        // evaluation rounds: full rounds while sphere rounds remain, then
        // whatever is left
        while (tail - head >= 64u || (base + 64u >= total && tail > head)) {
            const uint32_t k = head + lane;
            const bool valid = k < tail;
            const uint32_t e = valid ? sh->fifo[k % kFifo] : 0u;
            const int o = static_cast<int>(e >> 26);
            NavLite wl;
            wl.ext_tol = W::shfl(ext_tol, o);
            wl.next = 0;
            wl.last = -1;
            wl.status = kUnknown;
            wl.trust = kNoTrust;
            wl.volume = 0;
            wl.pos = {W::shfl(pos.x, o), W::shfl(pos.y, o), W::shfl(pos.z, o)};
            wl.dir = {W::shfl(dir.x, o), W::shfl(dir.y, o), W::shfl(dir.z, o)};
            const float oos = W::shfl(os, o);
            bool acc = false;
            Cand cd;
            cd.path = kMax;
            cd.sf = kInvalid;
            cd.link = 0;
            cd.status = kOutside;
            cd.pad = 0;
            // a placed cylinder can insert two candidates: its owner searches
            // it at its place in the order (flag bit 20, the entry in sf)
            const bool own = P::kExt && valid && det.bin_rec[e & 0x3ffffffu].rec.shape == DTR_CYLINDER2D;
            if (valid && !own) {
                FastEngine<P, NavLite> we{det, no_field, cfg, wl, c};
                acc = we.evaluate(det.bin_rec[e & 0x3ffffffu].rec, oos, cd);
            }
            if (acc) {
                sh->slot[lane].path = cd.path;
                sh->slot[lane].sf = cd.sf;
                sh->slot[lane].ls = static_cast<uint32_t>(cd.link) | (static_cast<uint32_t>(cd.status) << 16) |
                                    (static_cast<uint32_t>(cd.pad) << 18) | (static_cast<uint32_t>(o) << 24);
            } else if (own) {
                acc = true;
                sh->slot[lane].path = kMax;
                sh->slot[lane].sf = e & 0x3ffffffu;
                sh->slot[lane].ls = (1u << 20) | (static_cast<uint32_t>(o) << 24);
            }
            W::sync();
            uint64_t m = W::ballot(acc);
            if (req) {
                while (m) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1ull;
                    const uint32_t ls = sh->slot[j].ls;
                    if ((ls >> 24) != lane) continue;
                    if (P::kExt && ((ls >> 20) & 1u)) {
                        me.candidate_search(det.bin_rec[sh->slot[j].sf].rec, os);
                        continue;
                    }
                    Cand x;
                    x.path = sh->slot[j].path;
                    x.sf = sh->slot[j].sf;
                    x.link = static_cast<uint16_t>(ls & 0xffffu);
                    x.status = static_cast<uint8_t>((ls >> 16) & 0x3u);
                    x.pad = static_cast<uint8_t>((ls >> 18) & 0x3u);
                    me.insert_sorted(x);
                }
            }
            head = (tail - head >= 64u) ? head + 64u : tail;
            W::sync();
        }
    }
    if (!req) return 0u;
    me.update_status();
    if (nl.trust != kFull && nl.status == kOnPortal)
        nl.trust = (c.link(nl.next - 1) == kEndOfWorld) ? kFull : kNoTrust;
    return static_cast<uint32_t>(static_cast<uint8_t>(nl.next)) |
           (static_cast<uint32_t>(static_cast<uint8_t>(nl.last)) << 8) |
           (static_cast<uint32_t>(static_cast<uint8_t>(nl.status)) << 16) |
           (static_cast<uint32_t>(nl.trust) << 24);
}

// The navigation update of every active lane, local navigations pooled.
template <class P, class W>
DTR_AI void coop_navigate(const FastDetector& det, const DevConfig& cfg, Lane<P>& lane, bool active,
                          CacheView c, CoopShared* slots) {
    float os = 0.f;
    bool req = active && lane.nav_begin(det, cfg, c, os);
    while (W::any(req)) {
        const uint32_t r = coop_local_navigation<P, W>(det, cfg, req, lane.tk.pos, lane.tk.dir, lane.tk.ext_tol,
                                                       lane.tk.volume, os, c, slots);
        if (req) {
            lane.apply_local_navigation(r);
            req = lane.nav_resume(det, cfg, c, os);
        }
    }
}

// One iteration of propagate()'s loop for every active lane; lanes whose
// loop ends are finished and deactivated.
template <class P, class W>
DTR_AI void coop_iteration(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                           const TrackInputs& in, const DevResults& out, Lane<P>& lane, bool& active,
                           CacheView c, CoopShared* slots) {
    if (active && lane.actor_half(det, field, cfg, in, out, c)) {
        lane.finish(out, in.n, cfg);
        active = false;
    }
    coop_navigate<P, W>(det, cfg, lane, active, c, slots);
    if (active && lane.nav_end()) {
        lane.finish(out, in.n, cfg);
        active = false;
    }
    if (active) lane.step_half(det, field, cfg, c);
    coop_navigate<P, W>(det, cfg, lane, active, c, slots);
    if (active) lane.nav_end();
}

template <class P>
DTR_AI void clear_lane(Lane<P>& lane) {
    lane.tk.pos = {0.0, 0.0, 0.0};
    lane.tk.dir = {0.0, 0.0, 1.0};
    lane.tk.ext_tol = 0.0;
    lane.tk.volume = 0;
}

// One track per lane (has_track false for lanes past the end of the batch).
template <class P, class W>
DTR_AI void coop_run_grid(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                          const TrackInputs& in, const DevResults& out, bool has_track, uint32_t track,
                          CacheView c, CoopShared* slots) {
    Lane<P> lane;
    clear_lane(lane);
    bool active = has_track;
    float os = 0.f;
    if (active) lane.start_begin(det, field, cfg, in, c, track, os);
    const uint32_t r = coop_local_navigation<P, W>(det, cfg, active, lane.tk.pos, lane.tk.dir, lane.tk.ext_tol,
                                                   lane.tk.volume, os, c, slots);
    if (active) {
        lane.apply_local_navigation(r);
        lane.start_end();
    }
    while (W::any(active)) coop_iteration<P, W>(det, field, cfg, in, out, lane, active, c, slots);
}

// Persistent lanes over the work queue *queue: a wave takes `claim` entries
// at a time and hands them to its idle lanes once at least min_idle (or all)
// of them are idle; it ends when the queue is drained and its lanes are done.
template <class P, class W>
DTR_AI void coop_run_persistent(const FastDetector& det, const DevField& field, const DevConfig& cfg,
                                const TrackInputs& in, const DevResults& out, const uint32_t* order,
                                uint32_t* queue, uint32_t claim, uint32_t min_idle, CacheView c,
                                CoopShared* slots) {
    const uint64_t n = in.n;
    const uint32_t lane_id = W::lane();
    const uint64_t lt_mask = (lane_id == 0u) ? 0ull : (~0ull >> (64u - lane_id));
    Lane<P> lane;
    clear_lane(lane);
    bool active = false;
    bool drained = false;
    uint64_t pool_next = 0, pool_end = 0;  // wave-uniform
    for (;;) {
        const bool want = !active;
        uint64_t need = W::ballot(want);
        const uint32_t idle = static_cast<uint32_t>(__builtin_popcountll(need));
        if (idle < min_idle && idle < 64u) need = 0ull;
        uint32_t track = kInvalid;
        while (need != 0ull && !drained) {
            if (pool_next >= pool_end) {
                const int leader = __builtin_ctzll(need);
                uint32_t base = 0u;
                if (static_cast<int>(lane_id) == leader) base = W::atomic_add(queue, claim);
                base = W::shfl(base, leader);
                if (base >= n) {
                    drained = true;
                    break;
                }
                pool_next = base;
                pool_end = (base + static_cast<uint64_t>(claim) < n) ? base + claim : n;
            }
            const uint64_t take = pool_end - pool_next;
            const uint64_t rank = static_cast<uint64_t>(__builtin_popcountll(need & lt_mask));
            if (((need >> lane_id) & 1ull) && rank < take) {
                const uint64_t q = pool_next + rank;
                track = order ? order[q] : static_cast<uint32_t>(q);
            }
            const uint64_t cnt = static_cast<uint64_t>(__builtin_popcountll(need));
            pool_next += cnt < take ? cnt : take;
            need = W::ballot(want && track == kInvalid);
        }
        const bool starting = track != kInvalid;
        if (W::any(starting)) {
            float os = 0.f;
            if (starting) lane.start_begin(det, field, cfg, in, c, track, os);
            const uint32_t r = coop_local_navigation<P, W>(det, cfg, starting, lane.tk.pos, lane.tk.dir,
                                                           lane.tk.ext_tol, lane.tk.volume, os, c, slots);
            if (starting) {
                lane.apply_local_navigation(r);
                lane.start_end();
                active = true;
            }
        }
        if (!W::any(active)) break;
        coop_iteration<P, W>(det, field, cfg, in, out, lane, active, c, slots);
    }
}

}  // namespace dtr

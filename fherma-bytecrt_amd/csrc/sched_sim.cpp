// sched_sim: the server's WaveScheduler against a simulated backend (no GPU).
//
//   sched_sim seed requests plans cap0 cap1 cap2 max_wave max_groups max_inflight rate
//
// Requests of random plans (each plan a fixed footprint in the three rings)
// arrive as a Poisson stream of `rate` per tick; formed waves "run" for a time
// that grows with their size and retire oldest first, as the server's retire
// thread does.  After every step the checker verifies: in-flight regions lie
// inside each ring and never overlap; waves respect max_wave, max_groups and
// max_inflight; waves take requests strictly in arrival order; a wave's ring
// regions equal the sum of its requests' footprints; and every request is
// served within a bounded number of ticks.  Prints one JSON line; exit status
// 1 on any violation.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <string>
#include <vector>

#include "bytecrt_sched.hpp"

namespace {

struct Live {
    size_t start[bcrt::kRings], end[bcrt::kRings];
    uint64_t done_tick;
};

[[noreturn]] void fail(const std::string& why) {
    std::printf("{\"ok\": false, \"error\": \"%s\"}\n", why.c_str());
    std::exit(1);
}

// Ring space of the live waves: each region inside [0, cap), no two overlap.
void check_rings(const std::deque<Live>& live, const size_t* cap) {
    for (int k = 0; k < bcrt::kRings; ++k) {
        for (size_t i = 0; i < live.size(); ++i) {
            const Live& a = live[i];
            if (a.end[k] > cap[k] || a.start[k] > a.end[k]) fail("region outside ring " + std::to_string(k));
            for (size_t j = i + 1; j < live.size(); ++j) {
                const Live& b = live[j];
                const bool disjoint = a.end[k] <= b.start[k] || b.end[k] <= a.start[k] ||
                                      a.start[k] == a.end[k] || b.start[k] == b.end[k];
                if (!disjoint) fail("overlapping regions in ring " + std::to_string(k));
            }
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 11) {
        std::fprintf(stderr,
                     "usage: sched_sim seed requests plans cap0 cap1 cap2 max_wave max_groups max_inflight rate\n");
        return 2;
    }
    const uint64_t seed = std::strtoull(argv[1], nullptr, 10);
    const size_t requests = std::strtoull(argv[2], nullptr, 10);
    const size_t nplans = std::strtoull(argv[3], nullptr, 10);
    const size_t cap[bcrt::kRings] = {std::strtoull(argv[4], nullptr, 10), std::strtoull(argv[5], nullptr, 10),
                                      std::strtoull(argv[6], nullptr, 10)};
    const uint32_t max_wave = uint32_t(std::atoi(argv[7])), max_groups = uint32_t(std::atoi(argv[8]));
    const uint32_t max_inflight = uint32_t(std::atoi(argv[9]));
    const double rate = std::atof(argv[10]);

    std::mt19937_64 rng(seed);
    // Plan footprints: plan p needs up to cap / (1 + p % 8) of the arena (from
    // nearly the whole ring down to small requests), staging proportionally.
    std::vector<std::array<size_t, bcrt::kRings>> foot(nplans);
    for (size_t p = 0; p < nplans; ++p) {
        const size_t dev = 1 + rng() % (cap[0] / (1 + p % 8));
        foot[p] = {dev, std::max<size_t>(1, dev * cap[1] / cap[0]), std::max<size_t>(1, dev * cap[2] / cap[0])};
    }
    std::vector<int> plan_tag(nplans);  // addresses stand for plans

    bcrt::WaveScheduler sched(cap, max_wave, max_groups, max_inflight);
    std::poisson_distribution<int> arrivals(rate);
    std::deque<Live> live;
    size_t submitted = 0, served = 0, next_expected = 0, waves = 0, biggest = 0, peak_inflight = 0;
    size_t peak_use[bcrt::kRings] = {0, 0, 0};
    uint64_t tick = 0;
    const uint64_t tick_limit = 1000000;
    while (served < requests) {
        if (++tick > tick_limit) fail("no progress");
        for (int a = arrivals(rng); a > 0 && submitted < requests; --a) {
            bcrt::SchedItem it;
            const size_t p = rng() % nplans;
            it.plan = &plan_tag[p];
            for (int k = 0; k < bcrt::kRings; ++k) it.need[k] = foot[p][k];
            it.user = reinterpret_cast<void*>(submitted);
            if (!sched.fits_alone(it)) fail("footprint beyond capacity");
            sched.push(it);
            ++submitted;
        }
        bcrt::SchedWave w;
        while (sched.form(&w)) {
            ++waves;
            if (w.count == 0 || w.count > max_wave) fail("wave size");
            if (w.plans.size() > max_groups) fail("too many plans in a wave");
            std::vector<size_t> ids;
            size_t need[bcrt::kRings] = {0, 0, 0};
            for (size_t i = 0; i < w.plans.size(); ++i)
                for (const auto& it : w.items[i]) {
                    if (it.plan != w.plans[i]) fail("request grouped under the wrong plan");
                    ids.push_back(reinterpret_cast<size_t>(it.user));
                    for (int k = 0; k < bcrt::kRings; ++k) need[k] += it.need[k];
                }
            std::sort(ids.begin(), ids.end());
            for (size_t id : ids)
                if (id != next_expected++) fail("wave does not take requests in arrival order");
            Live l;
            for (int k = 0; k < bcrt::kRings; ++k) {
                if (w.end[k] - w.start[k] != need[k]) fail("region size differs from footprints");
                l.start[k] = w.start[k];
                l.end[k] = w.end[k];
            }
            l.done_tick = tick + 1 + w.count + rng() % 3;
            live.push_back(l);
            if (live.size() > max_inflight || sched.inflight() != live.size()) fail("in-flight bound");
            check_rings(live, cap);
            biggest = std::max(biggest, w.count);
            peak_inflight = std::max(peak_inflight, live.size());
            for (int k = 0; k < bcrt::kRings; ++k) {
                size_t use = 0;
                for (const Live& x : live) use += x.end[k] - x.start[k];
                peak_use[k] = std::max(peak_use[k], use);
            }
            served += w.count;
        }
        if (live.empty() && sched.pending() > 0) fail("idle rings but a pending request cannot start");
        // Waves finish in any order but release their space oldest first.
        while (!live.empty() && live.front().done_tick <= tick) {
            live.pop_front();
            sched.retire_oldest();
        }
    }
    std::printf("{\"ok\": true, \"requests\": %zu, \"waves\": %zu, \"max_wave\": %zu, \"peak_inflight\": %zu, "
                "\"peak_use\": [%zu, %zu, %zu], \"ticks\": %llu}\n",
                requests, waves, biggest, peak_inflight, peak_use[0], peak_use[1], peak_use[2],
                static_cast<unsigned long long>(tick));
    return 0;
}

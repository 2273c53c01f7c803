// Wave scheduling policy of the product server, free of any device code so it
// can be exercised on the CPU against a simulated backend.
//
// Requests wait in arrival order.  A wave is the longest prefix of the
// pending queue that (a) has at most max_wave requests, (b) spans at most
// max_groups distinct plans, and (c) fits, as one contiguous region, into each
// of three FIFO rings (device arena, pinned input staging, pinned output
// staging) next to the waves already in flight; at most max_inflight waves are
// in flight.  Waves release their ring space oldest first.
#pragma once

#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <vector>

namespace bcrt {

constexpr int kRings = 3;

struct SchedItem {
    const void* plan = nullptr;  // requests of one plan run as one group
    size_t need[kRings] = {0, 0, 0};
    void* user = nullptr;
};

struct SchedWave {
    std::vector<const void*> plans;              // in order of first appearance
    std::vector<std::vector<SchedItem>> items;   // per plan, in arrival order
    size_t start[kRings] = {0, 0, 0}, end[kRings] = {0, 0, 0};
    size_t count = 0;
};

class WaveScheduler {
   public:
    WaveScheduler(const size_t (&cap)[kRings], uint32_t max_wave, uint32_t max_groups, uint32_t max_inflight)
        : max_wave_(max_wave), max_groups_(max_groups), max_inflight_(max_inflight) {
        for (int k = 0; k < kRings; ++k) cap_[k] = cap[k];
    }

    // Whether the request could ever be admitted (it fits each empty ring).
    bool fits_alone(const SchedItem& it) const {
        for (int k = 0; k < kRings; ++k)
            if (it.need[k] > cap_[k]) return false;
        return true;
    }
    void push(const SchedItem& it) { pending_.push_back(it); }
    size_t pending() const { return pending_.size(); }
    size_t inflight() const { return inflight_.size(); }
    const std::deque<SchedItem>& pending_items() const { return pending_; }
    size_t capacity(int k) const { return cap_[k]; }

    // Forms the next wave (recorded as in flight) if a slot is free and the
    // first pending request fits.
    bool form(SchedWave* out) {
        if (pending_.empty() || inflight_.size() >= max_inflight_) return false;
        std::map<const void*, size_t> index;
        SchedWave w;
        size_t need[kRings] = {0, 0, 0};
        size_t taken = 0;
        for (const SchedItem& it : pending_) {
            if (taken == max_wave_) break;
            const bool new_plan = !index.count(it.plan);
            if (new_plan && w.plans.size() == max_groups_) break;
            bool fits = true;
            for (int k = 0; k < kRings && fits; ++k) {
                size_t probe;
                fits = ring_alloc(k, need[k] + it.need[k], &probe);
            }
            if (!fits) break;
            for (int k = 0; k < kRings; ++k) need[k] += it.need[k];
            if (new_plan) {
                index[it.plan] = w.plans.size();
                w.plans.push_back(it.plan);
                w.items.emplace_back();
            }
            w.items[index[it.plan]].push_back(it);
            ++taken;
        }
        if (taken == 0) return false;
        for (int k = 0; k < kRings; ++k) {
            ring_alloc(k, need[k], &w.start[k]);
            w.end[k] = w.start[k] + need[k];
        }
        w.count = taken;
        pending_.erase(pending_.begin(), pending_.begin() + taken);
        Region r;
        for (int k = 0; k < kRings; ++k) r.start[k] = w.start[k], r.end[k] = w.end[k];
        inflight_.push_back(r);
        *out = std::move(w);
        return true;
    }

    // Releases the ring space of the oldest in-flight wave.
    void retire_oldest() {
        if (!inflight_.empty()) inflight_.pop_front();
    }

   private:
    struct Region {
        size_t start[kRings], end[kRings];
    };

    // Start of a free region of `size` bytes in ring k after the in-flight waves.
    bool ring_alloc(int k, size_t size, size_t* start) const {
        if (size > cap_[k]) return false;
        if (inflight_.empty()) {
            *start = 0;
            return true;
        }
        const size_t first = inflight_.front().start[k], last = inflight_.back().end[k];
        if (last > first) {  // not wrapped: room after `last`, else before `first`
            if (cap_[k] - last >= size) {
                *start = last;
                return true;
            }
            if (first >= size) {
                *start = 0;
                return true;
            }
            return false;
        }
        if (first - last >= size) {  // wrapped (last <= first): room between them
            *start = last;
            return true;
        }
        return false;
    }

    size_t cap_[kRings];
    uint32_t max_wave_, max_groups_, max_inflight_;
    std::deque<SchedItem> pending_;
    std::deque<Region> inflight_;
};

}  // namespace bcrt

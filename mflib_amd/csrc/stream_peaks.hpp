// PeakFinder over a signal that arrives in consecutive segments (a
// streaming handle's stacks, one segment per push).  The peaks it returns,
// concatenated over the pushes, are exactly PeakFinder's list for the whole
// signal (src/peakFinder.cpp:468-586): threshold and local maximum (strict
// on the right, >= on the left, the first and last samples compared with
// their one neighbour, a lone sample kept), then greedy pruning in order of
// decreasing value, ties to the smaller index, keeping a peak only when
// every kept one is more than minDist samples away.
//
// Two things are decided late:
//   - the last sample of a segment is a candidate or not once its right
//     neighbour arrives (or, at the end of the signal, by the end rule);
//   - a candidate's greedy fate depends on the candidates of higher
//     priority within minDist, which may still arrive.  A candidate is
//     settled when it lies more than minDist before the first sample a
//     later segment can contribute and every higher-priority candidate
//     within minDist of it is settled; its fate is then final.  Settled
//     peaks are returned in increasing index, never ahead of an unsettled
//     candidate, so the returned indices increase across pushes.
// Settled candidates are kept while an unsettled or future candidate can
// lie within minDist of them, and dropped afterwards.
// Pure host code (no HIP): the engine feeds it the device scan's
// candidates of each segment.
#pragma once

#include <algorithm>
#include <cstdint>
#include <limits>
#include <set>
#include <utility>
#include <vector>

namespace mfstream {

template <typename T>
class StreamPeaks {
public:
    StreamPeaks(T threshold, int minDist) : thr_(threshold), md_(minDist) {}

    T threshold() const { return thr_; }

    // One segment of n >= 1 samples at global indices [g0, g0 + n),
    // following the previous segment with no gap (the first segment starts
    // the signal).  cand: the candidates a scan of the segment alone found
    // (local index, value; any order).  Those at the segment's first and
    // last samples are decided here instead.  x0, x1: the segment's first
    // two samples; xm, xl: its last two (x1 and xm unused when n == 1).
    // Appends the peaks that became final to out, in increasing index.
    void push(long long g0, int n, std::vector<std::pair<T, int>> cand, T x0, T x1, T xm, T xl,
              std::vector<long long> &out)
    {
        if (n < 1) { return; }
        // the previous segment's last sample, now that its right neighbour is known
        if (havePrev_) {
            const bool left = havePrev2_ ? prev1_ >= prev2_ : true;  // the signal's first sample has no left
            if (prev1_ > thr_ && left && x0 < prev1_) { add(g0 - 1, prev1_); }
        }
        // the segment's first sample (n >= 2; a lone sample waits as the last)
        if (n >= 2) {
            const bool left = havePrev_ ? x0 >= prev1_ : true;
            if (x0 > thr_ && left && x1 < x0) { add(g0, x0); }
        }
        std::sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a.second < b.second; });
        for (const auto &c : cand) {
            if (c.second > 0 && c.second < n - 1) { add(g0 + c.second, c.first); }
        }
        // the last sample waits for its right neighbour
        if (n >= 2) {
            prev2_ = xm;
            havePrev2_ = true;
        } else if (havePrev_) {
            prev2_ = prev1_;
            havePrev2_ = true;
        }
        prev1_ = xl;
        havePrev_ = true;
        lastIdx_ = g0 + n - 1;
        settle(lastIdx_, out);
    }

    // The end of the signal: the last sample by the end rule, then every
    // remaining peak.
    void flush(std::vector<long long> &out)
    {
        if (havePrev_) {
            const bool ok = havePrev2_ ? prev1_ > prev2_ : true;  // a lone sample is kept
            if (prev1_ > thr_ && ok) { add(lastIdx_, prev1_); }
            havePrev_ = false;
        }
        settle(std::numeric_limits<long long>::max(), out);
    }

private:
    struct Cand {
        long long idx;
        T val;
        int8_t state;  // 0 unsettled, 1 kept (settled), 2 pruned (settled)
        bool emitted;
    };

    void add(long long idx, T val) { live_.push_back(Cand{idx, val, 0, false}); }

    // future: the smallest index a later candidate can have
    void settle(long long future, std::vector<long long> &out)
    {
        if (md_ <= 0) {  // no pruning: every candidate is a peak
            for (auto &c : live_) { out.push_back(c.idx); }
            live_.clear();
            return;
        }
        const long long md = md_;
        std::vector<size_t> order(live_.size());
        for (size_t i = 0; i < order.size(); ++i) { order[i] = i; }
        // live_ is in index order: a stable sort by value puts ties in index order
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return live_[a].val > live_[b].val; });
        std::set<long long> kept, open;
        auto near = [md](const std::set<long long> &s, long long i) {
            const auto it = s.lower_bound(i - md);
            return it != s.end() && *it <= i + md;
        };
        const long long horizon = future == std::numeric_limits<long long>::max() ? future : future - md;
        for (size_t o : order) {
            Cand &c = live_[o];
            if (c.state != 0) {
                if (c.state == 1) { kept.insert(c.idx); }
                continue;
            }
            const bool keep = !near(kept, c.idx);
            if (keep) { kept.insert(c.idx); }
            if (c.idx >= horizon || near(open, c.idx)) {
                open.insert(c.idx);
            } else {
                c.state = keep ? 1 : 2;
            }
        }
        const long long firstOpen = open.empty() ? std::numeric_limits<long long>::max() : *open.begin();
        for (auto &c : live_) {
            if (c.state == 1 && !c.emitted && c.idx < firstOpen) {
                out.push_back(c.idx);
                c.emitted = true;
            }
        }
        // keep settled candidates only while an open or future one can lie within minDist
        const long long edge = std::min(firstOpen, future);
        const long long keepFrom = edge == std::numeric_limits<long long>::max() ? edge : edge - md;
        live_.erase(std::remove_if(live_.begin(), live_.end(),
                                   [&](const Cand &c) {
                                       return c.state != 0 && (c.emitted || c.state == 2) && c.idx < keepFrom;
                                   }),
                    live_.end());
    }

    T thr_;
    int md_;
    bool havePrev_ = false, havePrev2_ = false;
    T prev1_ = T(0), prev2_ = T(0);
    long long lastIdx_ = -1;  // global index of the last sample seen
    std::vector<Cand> live_;  // in index order
};

}  // namespace mfstream

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
#include <cstring>
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

    // Checkpoint: the whole pending state as bytes, and its restore.
    // Values are stored as double (exact for float).  Layout: magic, version,
    // sizeof(T), minDist, threshold, the two held-back samples and their
    // flags, the last index seen, then every live candidate (index, value,
    // fate, emitted).
    void save(std::vector<unsigned char> &out) const
    {
        out.clear();
        put<uint32_t>(out, kMagic);
        put<uint32_t>(out, kVersion);
        put<uint32_t>(out, static_cast<uint32_t>(sizeof(T)));
        put<int32_t>(out, md_);
        put<double>(out, static_cast<double>(thr_));
        put<double>(out, static_cast<double>(prev1_));
        put<double>(out, static_cast<double>(prev2_));
        put<int64_t>(out, lastIdx_);
        put<uint32_t>(out, (havePrev_ ? 1u : 0u) | (havePrev2_ ? 2u : 0u));
        put<uint32_t>(out, 0);
        put<int64_t>(out, static_cast<int64_t>(live_.size()));
        for (const auto &c : live_) {
            put<int64_t>(out, c.idx);
            put<double>(out, static_cast<double>(c.val));
            put<int32_t>(out, c.state);
            put<int32_t>(out, c.emitted ? 1 : 0);
        }
    }

    // Restores a state written by save() for the same T; false (and the
    // state unchanged) when the bytes are not such a state.
    bool load(const unsigned char *p, size_t n)
    {
        size_t at = 0;
        uint32_t magic = 0, version = 0, tsize = 0, flags = 0, pad = 0;
        int32_t md = 0;
        double thr = 0, p1 = 0, p2 = 0;
        int64_t last = 0, count = 0;
        if (!get(p, n, at, magic) || !get(p, n, at, version) || !get(p, n, at, tsize) || !get(p, n, at, md) ||
            !get(p, n, at, thr) || !get(p, n, at, p1) || !get(p, n, at, p2) || !get(p, n, at, last) ||
            !get(p, n, at, flags) || !get(p, n, at, pad) || !get(p, n, at, count)) {
            return false;
        }
        if (magic != kMagic || version != kVersion || tsize != sizeof(T) || md < 0 || flags > 3 || count < 0 ||
            !(thr >= 0 && thr <= 1)) {
            return false;
        }
        constexpr size_t kEntry = 8 + 8 + 4 + 4;
        if (n - at != static_cast<size_t>(count) * kEntry) { return false; }
        std::vector<Cand> live;
        live.reserve(static_cast<size_t>(count));
        for (int64_t i = 0; i < count; ++i) {
            int64_t idx = 0;
            double v = 0;
            int32_t st = 0, em = 0;
            get(p, n, at, idx);
            get(p, n, at, v);
            get(p, n, at, st);
            get(p, n, at, em);
            if (st < 0 || st > 2 || em < 0 || em > 1 || idx > last || (!live.empty() && idx <= live.back().idx)) {
                return false;
            }
            live.push_back(Cand{idx, static_cast<T>(v), static_cast<int8_t>(st), em != 0});
        }
        thr_ = static_cast<T>(thr);
        md_ = md;
        prev1_ = static_cast<T>(p1);
        prev2_ = static_cast<T>(p2);
        lastIdx_ = last;
        havePrev_ = (flags & 1u) != 0;
        havePrev2_ = (flags & 2u) != 0;
        live_ = std::move(live);
        return true;
    }

private:
    static constexpr uint32_t kMagic = 0x5053464Du;  // "MFSP"
    static constexpr uint32_t kVersion = 1;

    template <typename V>
    static void put(std::vector<unsigned char> &out, V v)
    {
        const size_t at = out.size();
        out.resize(at + sizeof(V));
        std::memcpy(out.data() + at, &v, sizeof(V));
    }

    template <typename V>
    static bool get(const unsigned char *p, size_t n, size_t &at, V &v)
    {
        if (n < sizeof(V) || at > n - sizeof(V)) { return false; }
        std::memcpy(&v, p + at, sizeof(V));
        at += sizeof(V);
        return true;
    }

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

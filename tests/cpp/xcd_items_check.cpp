// CPU check of mfh::appendXcdItems (mflib_amd/csrc/mf_host.hpp), the order
// of a launch's work items over the 8 XCDs (workgroup i runs on XCD i mod
// 8).  For generated sets of lanes it checks that every block of every lane
// appears exactly once with its rowsAt count (empty items only pad the
// bins of a many-lane launch), that a launch of many lanes keeps all items
// of an unsplit lane on one XCD, that the XCDs get about the same work,
// and that equal lanes stay on one XCD each.  Prints "ok" or the first
// failure and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "mf_host.hpp"

static int fail(const char *what, int seed)
{
    std::printf("FAIL seed %d: %s\n", seed, what);
    return 1;
}

int main()
{
    unsigned state = 12345u;
    auto rnd = [&state](int n) {
        state = state * 1664525u + 1013904223u;
        return static_cast<int>((state >> 8) % static_cast<unsigned>(n));
    };
    for (int seed = 0; seed < 300; ++seed) {
        const int nL = seed % 3 == 0 ? 1 + rnd(3) : (seed % 3 == 1 ? 4 + rnd(28) : 32 + rnd(200));
        std::vector<int> laneIds(nL);
        std::vector<long long> blocks(nL);
        std::vector<int> rows(nL);
        for (int i = 0; i < nL; ++i) {
            laneIds[i] = 1000 + 3 * i;
            blocks[i] = seed % 5 == 0 ? 1 + rnd(4) : 1 + rnd(4000);
            rows[i] = 1 + rnd(100);
        }
        // rows with outputs in block k: the lane's rows, thinning out towards its end
        auto rowsAt = [&](int i, long long k) {
            return static_cast<int>(rows[i] - (rows[i] - 1) * k / blocks[i]);
        };
        for (int withRows = 0; withRows < 2; ++withRows) {
            std::vector<mfk::Item> items;
            if (withRows) {
                mfh::appendXcdItems(laneIds, blocks, items, rowsAt);
            } else {
                mfh::appendXcdItems(laneIds, blocks, items);
            }
            long long total = 0, totalWork = 0;
            for (int i = 0; i < nL; ++i) {
                total += blocks[i];
                for (long long k = 0; k < blocks[i]; ++k) { totalWork += 6 + rowsAt(i, k); }
            }
            std::map<int, int> laneOf;
            for (int i = 0; i < nL; ++i) { laneOf[laneIds[i]] = i; }
            std::set<std::pair<int, int>> seen;
            std::map<int, std::set<int>> xcds;  // lane -> XCDs its items run on
            double xcdWork[8] = {};
            long long padding = 0;
            for (size_t p = 0; p < items.size(); ++p) {
                const mfk::Item &it = items[p];
                if (it.nRows == 0) {  // an empty item: only to pad the bins of a many-lane launch
                    if (!withRows || nL < 32) { return fail("empty item", seed); }
                    ++padding;
                    continue;
                }
                const auto li = laneOf.find(it.seg);
                if (li == laneOf.end()) { return fail("unknown lane", seed); }
                if (it.block < 0 || it.block >= blocks[li->second]) { return fail("block out of range", seed); }
                if (!seen.insert({it.seg, it.block}).second) { return fail("block twice", seed); }
                const int want = withRows ? rowsAt(li->second, it.block) : -1;
                if (it.nRows != want) { return fail("nRows", seed); }
                xcds[li->second].insert(static_cast<int>(p % 8));
                xcdWork[p % 8] += 6.0 + std::max(it.nRows, 0);
            }
            if (static_cast<long long>(items.size()) - padding != total) { return fail("item count", seed); }
            if (padding > 0 && items.size() % 8 != 0) { return fail("padded bins unequal", seed); }
            if (withRows && nL >= 32 && padding > 0) {
                // padded work bins: every lane under a quarter of a bin's work
                // stays on one XCD, and the XCDs' work is within a quarter of the
                // mean
                double sum = 0, mx = 0;
                for (double v : xcdWork) {
                    sum += v;
                    mx = std::max(mx, v);
                }
                if (mx > 1.25 * sum / 8) { return fail("XCD work unbalanced (many lanes)", seed); }
                for (int i = 0; i < nL; ++i) {
                    long long w = 0;
                    for (long long k = 0; k < blocks[i]; ++k) { w += 6 + rowsAt(i, k); }
                    if (w <= totalWork / 32 && xcds[i].size() > 1) { return fail("padded lane spread", seed); }
                }
            }
            if (nL >= 32 && padding == 0) {
                // balanced by blocks: a lane under half an XCD's share is never cut,
                // and its items stay on one XCD unless they fall in the tail after
                // the shortest bin ran out
                const long long cap = std::max<long long>(1, total / 16);
                long long maxB = 0;
                for (long long b : blocks) { maxB = std::max(maxB, b); }
                int spread = 0;
                for (int i = 0; i < nL; ++i) {
                    if (blocks[i] <= cap && xcds[i].size() > 1) { ++spread; }
                }
                // the bins differ by less than the largest piece, so only lanes
                // in that tail may have moved
                if (maxB <= cap && spread > nL / 4) { return fail("lanes spread over XCDs", seed); }
            }
            if (withRows && nL < 32 && total >= 512) {
                // few lanes: the work each XCD receives (a block weighs 6 + its
                // rows) stays within a quarter of the mean
                double w[8] = {};
                for (size_t p = 0; p < items.size(); ++p) { w[p % 8] += 6.0 + items[p].nRows; }
                double sum = 0, mx = 0;
                for (double v : w) {
                    sum += v;
                    mx = std::max(mx, v);
                }
                if (mx > 1.25 * sum / 8) { return fail("XCD work unbalanced", seed); }
            }
        }
    }
    // 24 equal lanes of 16 rows (one day on 24 channels, 16 templates):
    // the bins are balanced whole, so every lane stays on one XCD
    {
        std::vector<int> laneIds(24);
        std::vector<long long> blocks(24, 5625);
        for (int i = 0; i < 24; ++i) { laneIds[i] = i; }
        std::vector<mfk::Item> items;
        mfh::appendXcdItems(laneIds, blocks, items, [](int, long long) { return 16; });
        std::map<int, std::set<int>> xcds;
        for (size_t p = 0; p < items.size(); ++p) { xcds[items[p].seg].insert(static_cast<int>(p % 8)); }
        for (const auto &kv : xcds) {
            if (kv.second.size() != 1) { return fail("an equal lane spread over XCDs", -1); }
        }
    }
    std::printf("ok\n");
    return 0;
}

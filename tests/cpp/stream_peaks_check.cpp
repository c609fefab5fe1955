// CPU driver of mfstream::StreamPeaks (mflib_amd/csrc/stream_peaks.hpp):
// feeds a signal in the given segments, each with the candidates a scan of
// the segment alone finds (threshold and local maximum with the end rules
// applied at the segment's own ends, as the device scan does), then flushes.
// Input file: int32 precision (0 float, 1 double), minDist, nSeg, n, ckpt;
// double threshold; int32 segment lengths[nSeg]; n samples.  With ckpt > 0
// the state goes through save() into a new object's load() after every
// ckpt-th segment (a checkpoint and resume).  Output file: int64
// count, then the peaks in the order the pushes returned them, then int64
// pushes-with-output-out-of-order (0 when every push's peaks follow the last).
#include <cstdio>
#include <cstdint>
#include <vector>

#include "stream_peaks.hpp"

template <typename T>
static std::vector<std::pair<T, int>> scan(const T *x, int n, T thr)
{
    std::vector<std::pair<T, int>> c;
    for (int i = 0; i < n; ++i) {
        if (!(x[i] > thr)) { continue; }
        bool ok;
        if (n == 1) {
            ok = true;
        } else if (i == 0) {
            ok = x[1] < x[0];
        } else if (i == n - 1) {
            ok = x[n - 1] > x[n - 2];
        } else {
            ok = x[i] >= x[i - 1] && x[i + 1] < x[i];
        }
        if (ok) { c.push_back({x[i], i}); }
    }
    return c;
}

template <typename T>
static int run(FILE *f, FILE *g, int md, int nSeg, int n, int ckpt, double thr)
{
    std::vector<int> seg(nSeg);
    std::vector<T> x(n);
    if (fread(seg.data(), sizeof(int), nSeg, f) != static_cast<size_t>(nSeg)) { return 2; }
    if (fread(x.data(), sizeof(T), n, f) != static_cast<size_t>(n)) { return 2; }
    mfstream::StreamPeaks<T> sp(static_cast<T>(thr), md);
    std::vector<long long> out;
    long long g0 = 0, disorder = 0;
    for (int s = 0; s < nSeg; ++s) {
        const int m = seg[s];
        if (m < 1 || g0 + m > n) { return 3; }
        const T *p = x.data() + g0;
        const size_t before = out.size();
        sp.push(g0, m, scan<T>(p, m, static_cast<T>(thr)), p[0], p[m > 1 ? 1 : 0], p[m > 1 ? m - 2 : 0], p[m - 1],
                out);
        if (before > 0 && out.size() > before && out[before] <= out[before - 1]) { ++disorder; }
        g0 += m;
        if (ckpt > 0 && (s + 1) % ckpt == 0) {
            std::vector<unsigned char> blob;
            sp.save(blob);
            mfstream::StreamPeaks<T> resumed(T(0), 0);
            if (!resumed.load(blob.data(), blob.size())) { return 4; }
            std::vector<unsigned char> again;
            resumed.save(again);
            if (again != blob) { return 5; }
            // a truncated or altered state is refused and leaves the object as it was
            if (resumed.load(blob.data(), blob.size() - 1)) { return 6; }
            std::vector<unsigned char> bad(blob);
            bad[0] ^= 1u;
            if (resumed.load(bad.data(), bad.size())) { return 6; }
            sp = resumed;
        }
    }
    if (g0 != n) { return 3; }
    const size_t before = out.size();
    sp.flush(out);
    if (before > 0 && out.size() > before && out[before] <= out[before - 1]) { ++disorder; }
    const long long cnt = static_cast<long long>(out.size());
    fwrite(&cnt, sizeof(cnt), 1, g);
    fwrite(out.data(), sizeof(long long), out.size(), g);
    fwrite(&disorder, sizeof(disorder), 1, g);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 3) { return 1; }
    FILE *f = fopen(argv[1], "rb");
    FILE *g = fopen(argv[2], "wb");
    if (!f || !g) { return 1; }
    int hdr[5];
    double thr;
    if (fread(hdr, sizeof(int), 5, f) != 5 || fread(&thr, sizeof(double), 1, f) != 1) { return 2; }
    const int rc = hdr[0] == 0 ? run<float>(f, g, hdr[1], hdr[2], hdr[3], hdr[4], thr)
                               : run<double>(f, g, hdr[1], hdr[2], hdr[3], hdr[4], thr);
    fclose(f);
    fclose(g);
    return rc;
}

/*
 * CPU oracle for the matched-filter detection path (see mf_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py are the only users.  Nothing in the product
 * links or loads this file's library.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_window_length        reference KATs (testing/matchedFilterParameters.cpp:19-31,82-89;
 *                            python/unit_tests.py:159) + brute-force minimisation
 *   orc_naive_pearson_f64    closed forms (rho=1 on a copied window, affine
 *                            invariance, sign flip) + numpy corrcoef brute force
 *   orc_pearson_oracle_*     agrees with the naive restatement and numpy corrcoef
 *   orc_rfft / orc_irfft     numpy.fft.rfft / irfft
 *   orc_mf_apply_*           reference tolerances vs the naive Pearson:
 *                            1e-13 / 2e-6 (testing/matchedFilter.cpp:127,237),
 *                            1e-12 uneven single channel (testing/singleChannel.cpp:154-159),
 *                            1e-11 shift-and-stack inputs (testing/matchedFilter.cpp:425-433)
 *   orc_shift_and_stack_*    stack maximum 1 +/- 1e-3 at 8 s +/- dt/2
 *                            (testing/matchedFilter.cpp:437-443) + closed form
 *   orc_peak_finder_*        the 13- and 21-index KATs (testing/peakFinder.cpp:62-143)
 *                            + brute-force local-max/greedy on tiny inputs
 */
#include "mf_oracle.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ---------------------------------------------------------------------- */
/* Block sizing: src/matchedFilterParameters.cpp:18-104, 371-385           */
/* ---------------------------------------------------------------------- */
int nextPow2(int n)
{
    if (n < 0) { return -1; }
    if (n <= 1) { return 1; }
    std::uint64_t p = 1;
    while (p < static_cast<std::uint64_t>(n)) { p <<= 1; }
    if (p > static_cast<std::uint64_t>(std::numeric_limits<int>::max())) { return -1; }
    return static_cast<int>(p);
}

/* The 21-entry cost table is the reference's (Matlab flop counts for
 * power-of-two transforms); it must match for the sizing KATs to hold. */
const double kFftCost[21] = {
    18, 59, 138, 303, 660, 1441, 3150, 6875, 14952, 32373, 69762,
    149647, 319644, 680105, 1441974, 3047619, 6422736, 13500637,
    28311786, 59244791, 59244791 * 2.09};

int windowLength(int nb, int nx, int *nfft, int *L)
{
    if (nb < 1 || nx <= nb) { return -1; }
    int first = -1;
    for (int i = 0; i < 21; ++i) {
        long long len = 2LL << i;
        if (len > nb - 1) { first = i; break; }
    }
    if (first < 0) {
        int p = nextPow2(nx);
        if (p < 0) { return -1; }
        *nfft = p;
        *L = p - (nb - 1);
        return 0;
    }
    double best = std::numeric_limits<double>::max();
    int bn = 0, bl = 0;
    for (int i = first; i < 21; ++i) {
        long long len = 2LL << i;
        long long lt = len - (nb - 1);
        double obj = std::ceil(nx / static_cast<double>(lt)) * kFftCost[i];
        if (obj < best) { best = obj; bn = static_cast<int>(len); bl = static_cast<int>(lt); }
    }
    *nfft = bn;
    *L = bl;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Compensated summation used by the FP64 oracle of record                 */
/* ---------------------------------------------------------------------- */
struct Neumaier {
    double s = 0, c = 0;
    void add(double v)
    {
        double t = s + v;
        if (std::fabs(s) >= std::fabs(v)) { c += (s - t) + v; }
        else { c += (v - t) + s; }
        s = t;
    }
    double value() const { return s + c; }
};

/* ---------------------------------------------------------------------- */
/* This file's real FFT: radix-2 complex FFT of N/2 points + packing.       */
/* Twiddles come from the C library's cos/sin in FP64, rounded to T.        */
/* ---------------------------------------------------------------------- */
template <typename T>
struct RealFFT {
    using C = std::complex<T>;
    int N = 0, M = 0;
    std::vector<C> tw;   /* exp(-2 pi i k / M), k < M/2            */
    std::vector<C> twr;  /* exp(-pi i k / M),   k <= M/2           */
    std::vector<int> rev;
    explicit RealFFT(int n) : N(n), M(n / 2)
    {
        tw.resize(std::max(1, M / 2));
        for (int k = 0; k < M / 2; ++k) {
            double a = -2.0 * M_PI * k / M;
            tw[k] = C(static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a)));
        }
        twr.resize(M / 2 + 1);
        for (int k = 0; k <= M / 2; ++k) {
            double a = -M_PI * k / M;
            twr[k] = C(static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a)));
        }
        rev.resize(M);
        int bits = 0;
        while ((1 << bits) < M) { ++bits; }
        for (int i = 0; i < M; ++i) {
            int r = 0;
            for (int b = 0; b < bits; ++b) { if (i & (1 << b)) { r |= 1 << (bits - 1 - b); } }
            rev[i] = r;
        }
    }
    void cfft(C *a, bool inverse) const
    {
        for (int i = 0; i < M; ++i) { if (i < rev[i]) { std::swap(a[i], a[rev[i]]); } }
        for (int len = 2; len <= M; len <<= 1) {
            int half = len / 2, step = M / len;
            for (int i = 0; i < M; i += len) {
                for (int j = 0; j < half; ++j) {
                    C w = tw[j * step];
                    if (inverse) { w = std::conj(w); }
                    C u = a[i + j];
                    C v = a[i + j + half] * w;
                    a[i + j] = u + v;
                    a[i + j + half] = u - v;
                }
            }
        }
    }
    /* X[0..M] = DFT_N(x) */
    void forward(const T *x, C *X, std::vector<C> &z) const
    {
        z.resize(M);
        for (int j = 0; j < M; ++j) { z[j] = C(x[2 * j], x[2 * j + 1]); }
        cfft(z.data(), false);
        const T half = static_cast<T>(0.5);
        for (int k = 0; k <= M / 2; ++k) {
            C zk = z[k % M];
            C zm = std::conj(z[(M - k) % M]);
            C E = zk + zm;
            C O = (zk - zm) * C(0, -1);
            C WO = twr[k] * O;
            X[k] = (E + WO) * half;
            X[M - k] = std::conj(E - WO) * half;
        }
    }
    /* x = N * IDFT_N(X)  (unnormalized, the FFTW c2r convention) */
    void inverse(const C *X, T *x, std::vector<C> &z) const
    {
        z.resize(M);
        for (int k = 0; k <= M / 2; ++k) {
            C S = X[k] + std::conj(X[M - k]);
            C D = X[k] - std::conj(X[M - k]);
            C P = std::conj(twr[k]) * D;
            z[k % M] = S + C(0, 1) * P;
            if (k > 0) { z[M - k] = std::conj(S) + C(0, 1) * std::conj(P); }
        }
        cfft(z.data(), true);
        for (int j = 0; j < M; ++j) { x[2 * j] = z[j].real(); x[2 * j + 1] = z[j].imag(); }
    }
};

/* ---------------------------------------------------------------------- */
/* Template preparation: src/matchedFilter.cpp:168-190                     */
/* ---------------------------------------------------------------------- */
template <typename T>
void demeanNormalizeReverse(int nb, const T *b, T *bOut)
{
    double mean = 0;
    for (int i = 0; i < nb; ++i) { mean = mean + b[i]; }
    mean = static_cast<T>(mean / static_cast<double>(nb));
    T norm = 0;
    for (int i = 0; i < nb; ++i) {
        bOut[i] = static_cast<T>(b[nb - 1 - i] - mean);
        norm = norm + bOut[i] * bOut[i];
    }
    norm = std::sqrt(norm);
    if (norm > 0) { norm = static_cast<T>(1.0 / static_cast<double>(norm)); }
    for (int i = 0; i < nb; ++i) { bOut[i] = bOut[i] * norm; }
}

/* ---------------------------------------------------------------------- */
/* Sliding normalization: src/matchedFilter.cpp:94-166                      */
/* (the reference's omp-simd initial reduction is summed sequentially)      */
/* ---------------------------------------------------------------------- */
template <typename T>
void normalizeSignal(int n, int lent, const T *y, T *yNum, T tol, T tooBig)
{
    const T zero = 0;
    std::vector<T> s(lent + 16), s2(lent + 16);
    const T scalNum = static_cast<T>(std::sqrt(static_cast<double>(lent)));
    const int scalDen = lent;
    for (int i = 0; i < n - lent + 1; i = i + lent) {
        const T *yp = &y[i];
        T *yn = &yNum[i];
        T sum = 0, sum2 = 0;
        for (int j = 0; j < lent; ++j) {
            sum = sum + yp[j];
            sum2 = sum2 + yp[j] * yp[j];
        }
        s[0] = sum;
        s2[0] = sum2;
        for (int j = 1; j < lent; ++j) {
            T d = yp[lent - 1 + j] - yp[j - 1];
            T p = yp[lent - 1 + j] + yp[j - 1];
            s[j] = s[j - 1] + d;
            s2[j] = s2[j - 1] + p * d;
        }
        for (int j = 0; j < lent; ++j) {
            T arg = std::max(zero, scalDen * s2[j] - s[j] * s[j]);
            T den = std::sqrt(arg);
            T num = scalNum * yn[j];
            yn[j] = num / std::max(tol, den);
            if (std::abs(yn[j]) > tooBig) { yn[j] = zero; }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* MatchedFilter<T>::initialize + apply: src/matchedFilter.cpp:506-821,     */
/* 862-921 (dead trace), 971-1316 (block loop + normalization)              */
/* ---------------------------------------------------------------------- */
template <typename T>
int mfApply(int mode, int nT, const int *lens, const double *const *templates,
            int n, const T *const *signals, int floatQuirk, int nThreads,
            T *out, int *nfftOut, int *lOut, T tol)
{
    if (nT < 1 || n < 1 || lens == nullptr || templates == nullptr ||
        signals == nullptr || out == nullptr) { return -1; }
    if (mode != 0 && mode != 1) { return -1; }
    int nb = 0;
    for (int it = 0; it < nT; ++it) {
        if (lens[it] < 1 || templates[it] == nullptr) { return -1; }
        nb = std::max(nb, lens[it]);
    }
    const bool multi = (mode == 1);
    const int nx = n + nb - 1;
    int nfft = 0, L = 0;
    if (windowLength(nb, nx, &nfft, &L) != 0) { return -1; }
    if (nfftOut) { *nfftOut = nfft; }
    if (lOut) { *lOut = L; }
    if (n - nb + 1 < 1) { return -1; }
    const int nSpec = nfft / 2 + 1;
    RealFFT<T> fft(nfft);
    using C = std::complex<T>;
    std::vector<C> work;
    /* Template spectra B[it] = r2c(reverse(demean(t))/||.||) */
    std::vector<C> B(static_cast<size_t>(nT) * nSpec);
    std::vector<T> seg(nfft);
    std::vector<T> tData(nb);
    for (int it = 0; it < nT; ++it) {
        std::fill(seg.begin(), seg.end(), T(0));
        std::fill(tData.begin(), tData.end(), T(0));
        for (int i = 0; i < lens[it]; ++i) { tData[i] = static_cast<T>(templates[it][i]); }
        int len = lens[it];
        if (floatQuirk && sizeof(T) == sizeof(float)) { len = nb; }
        demeanNormalizeReverse<T>(len, tData.data(), seg.data());
        fft.forward(seg.data(), &B[static_cast<size_t>(it) * nSpec], work);
    }
    /* Input rows, zero padded beyond n (the reference callocs them) */
    const int nRows = multi ? nT : 1;
    const size_t ldm = static_cast<size_t>(n) + nb + 16;
    std::vector<T> x(static_cast<size_t>(nRows) * ldm, T(0));
    std::vector<char> dead(nT, 0);
    for (int r = 0; r < nRows; ++r) {
        const T *s = signals[r];
        if (s == nullptr) { return -1; }
        std::copy(s, s + n, &x[r * ldm]);
        auto mm = std::minmax_element(s, s + n);
        if (std::abs(*mm.first) == std::abs(*mm.second)) { dead[r] = 1; }
    }
    std::vector<T> filt(static_cast<size_t>(nT) * ldm, T(0));
    std::vector<C> X(nSpec), Xs(nSpec);
    std::vector<T> segOut(nfft);
    const T xnorm = static_cast<T>(1 / static_cast<double>(nfft));
    for (int istart = 0; istart < nx; istart = istart + L) {
        const int iend = std::min(nx, istart + L);
        const int ncopy = iend - istart;
        const int yend = std::min(nx, istart + nfft);
        const int nupdate = yend - istart;
        if (!multi) {
            std::fill(seg.begin(), seg.end(), T(0));
            std::copy(&x[istart], &x[istart] + ncopy, seg.begin());
            fft.forward(seg.data(), Xs.data(), work);
        }
        for (int it = 0; it < nT; ++it) {
            const C *b = &B[static_cast<size_t>(it) * nSpec];
            if (multi) {
                std::fill(seg.begin(), seg.end(), T(0));
                const T *src = &x[it * ldm + istart];
                std::copy(src, src + ncopy, seg.begin());
                fft.forward(seg.data(), X.data(), work);
                for (int w = 0; w < nSpec; ++w) { X[w] = b[w] * X[w]; }
            } else {
                for (int w = 0; w < nSpec; ++w) { X[w] = b[w] * Xs[w]; }
            }
            fft.inverse(X.data(), segOut.data(), work);
            T *ydst = &filt[it * ldm + istart];
            for (int i = 0; i < nupdate; ++i) { ydst[i] = ydst[i] + segOut[i] * xnorm; }
        }
    }
#ifdef _OPENMP
    int nth = nThreads > 0 ? nThreads : omp_get_max_threads();
#pragma omp parallel for num_threads(nth) schedule(dynamic)
#endif
    for (int it = 0; it < nT; ++it) {
        if (dead[it]) { continue; }
        const T *y = &x[(multi ? it : 0) * ldm];
        T *yNum = &filt[it * ldm + lens[it] - 1];
        normalizeSignal<T>(n, lens[it], y, yNum, tol, static_cast<T>(1.1));
    }
    (void)nThreads;
    const int nOut = n - nb + 1;
    for (int it = 0; it < nT; ++it) {
        const T *src = &filt[it * ldm + lens[it] - 1];
        std::copy(src, src + nOut, out + static_cast<size_t>(it) * nOut);
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* shiftAndStack: src/matchedFilter.cpp:1318-1364                           */
/* ---------------------------------------------------------------------- */
template <typename T>
int shiftAndStack(int nT, const T *rho, long long ld, int nOut, const int *shifts,
                  const double *weights, int absval, T *out, int *outLen)
{
    if (nT < 1 || rho == nullptr || shifts == nullptr || weights == nullptr ||
        out == nullptr) { return -1; }
    int smax = shifts[0];
    T xnorm = 0;
    for (int it = 0; it < nT; ++it) {
        if (shifts[it] < 0) { return -1; }
        smax = std::max(smax, shifts[it]);
        xnorm = xnorm + static_cast<T>(weights[it]);
    }
    if (xnorm > 0) { xnorm = 1 / xnorm; }
    const int len = nOut - smax;
    if (len < 0) { return -1; }
    for (int k = 0; k < len; ++k) { out[k] = 0; }
    for (int it = 0; it < nT; ++it) {
        const T w = xnorm * static_cast<T>(weights[it]);
        const T *p = rho + it * ld + shifts[it];
        if (absval) {
            for (int k = 0; k < len; ++k) { out[k] = out[k] + w * std::abs(p[k]); }
        } else {
            for (int k = 0; k < len; ++k) { out[k] = out[k] + w * p[k]; }
        }
    }
    *outLen = len;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Detector reduce + clamp: reduceAllCorrelations and cleanReducedTemplate */
/* (src/singleChannel/detector.cpp:190-250).  det[i] = max over templates  */
/* of rho_t[i] (or |rho_t[i]|), the first template reaching it in id[i]     */
/* (strict >), then det clamped to [-1, 1].  The workspace is calloc'd by  */
/* the caller (detector.cpp:460-463), so id starts at 0 in both branches.  */
/* ---------------------------------------------------------------------- */
template <typename T>
int reduceCorrelations(int nT, const T *rho, long long ld, int len, int absval, T *det, int *id)
{
    if (nT < 1 || len < 0 || rho == nullptr || det == nullptr || id == nullptr) { return -1; }
    for (int i = 0; i < len; ++i) {
        det[i] = absval ? std::abs(rho[i]) : rho[i];
        id[i] = 0;
    }
    for (int it = 1; it < nT; ++it) {
        const T *p = rho + it * ld;
        for (int i = 0; i < len; ++i) {
            const T v = absval ? std::abs(p[i]) : p[i];
            if (v > det[i]) {
                det[i] = v;
                id[i] = it;
            }
        }
    }
    const T one = 1;
    for (int i = 0; i < len; ++i) { det[i] = std::max(-one, std::min(one, det[i])); }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Per-detection products of Detector::detect, FP64 (the reference's        */
/* double instantiation): src/singleChannel/detector.cpp:38-152, 537-664;   */
/* src/singleChannel/relativeMagnitude.cpp:35-58, 76-98, 273-388.           */
/* ---------------------------------------------------------------------- */

/* quadraticRefinement (detector.cpp:38-89). */
double quadraticShift(int n, const double *xc, int p, double dt, int absval)
{
    if (!(p > 0 && p < n - 1)) { return 0.0; }
    double y1 = xc[p - 1], y2 = xc[p], y3 = xc[p + 1];
    if (absval) {
        y1 = std::abs(y1);
        y2 = std::abs(y2);
        y3 = std::abs(y3);
    }
    if (y2 < y1 || y2 < y3) {
        if (y2 < y1 && y2 < y3) {
            y1 = -y1;
            y2 = -y2;
            y3 = -y3;
        } else {
            return 0.0;
        }
    }
    const double eps100 = 100 * std::numeric_limits<double>::epsilon();
    y1 = std::min(y2 - eps100, y1);
    y3 = std::min(y2 - eps100, y3);
    const double num = dt * (y1 - y3);
    const double den = 2 * (y1 - 2 * y2 + y3);
    return num / den;
}

/* sinc / lanczos (detector.cpp:91-109). */
double sincRef(double x)
{
    double r = 1.0;
    if (std::abs(x) > std::numeric_limits<double>::epsilon() * 100) {
        const double pix = M_PI * x;
        r = std::sin(pix) / pix;
    }
    return r;
}

double lanczosRef(double x, int a)
{
    double r = 0.0;
    if (static_cast<int>(std::abs(x)) <= a) { r = sincRef(x) * sincRef(x / a); }
    return r;
}

/* The interpolant of lanczosRefinement (detector.cpp:120-139) at grid
 * points idx = 1 .. nRefine-1 (profile[idx-1]), and the shift it picks
 * (140-148): the first grid point of the largest value, strict '>' from
 * -DBL_MAX; 0 when no point beats that. */
double lanczosShift(int n, const double *xc, int p, double dt, int absval, int a, int nRefine, double *profile)
{
    const double dtWork = 1 / static_cast<double>(nRefine);
    double optShift = 0;
    double yMax = -std::numeric_limits<double>::max();
    for (int idx = 1; idx < nRefine; ++idx) {
        const double shift = -0.5 + dtWork * idx;
        const double x = p + shift;
        double y = 0;
        for (int k = -a; k <= a; ++k) {
            const int i = static_cast<int>(std::floor(x)) - k + 1;
            const double xi = x - static_cast<double>(i);
            double v = 0;
            if (i >= 0 && i < n) { v = xc[i]; }
            y = y + v * lanczosRef(xi, a);
        }
        if (absval) { y = std::abs(y); }
        if (profile) { profile[idx - 1] = y; }
        if (y > yMax) {
            optShift = dt * shift;
            yMax = y;
        }
    }
    return optShift;
}

/* RelativeMagnitude: demeanNorm2 of the template (76-98, 273-299),
 * demeanCopyComputeNorm2 of the window (50-74, 318-339; a zero norm throws
 * there, caught by detect, so no amplitude), Gibbons & Ringdal (35-47) and
 * Schaff & Richards (49-58).  Returns 1 when the factors were computed. */
int amplitudeFactors(int Lt, const double *t, const double *y, double *gr, double *sr)
{
    const double mt = std::accumulate(t, t + Lt, 0.0) / static_cast<double>(Lt);
    const double my = std::accumulate(y, y + Lt, 0.0) / static_cast<double>(Lt);
    std::vector<double> x(Lt), d(Lt);
    double xx = 0, yy = 0;
    for (int i = 0; i < Lt; ++i) {
        x[i] = t[i] - mt;
        xx = xx + x[i] * x[i];
        d[i] = y[i] - my;
        yy = yy + d[i] * d[i];
    }
    if (yy == 0) { return 0; }
    const double xy = std::inner_product(x.begin(), x.end(), d.begin(), 0.0);
    *gr = std::abs(xy) / xx;
    *sr = std::sqrt(yy / xx);
    return 1;
}

/* ---------------------------------------------------------------------- */
/* PeakFinder<T>::apply: src/peakFinder.cpp:468-586                         */
/* ---------------------------------------------------------------------- */
template <typename T>
int peakFinder(int n, const T *x, T thr, int minDist, int *idx, int *nPeaks)
{
    if (n < 1 || x == nullptr || idx == nullptr || nPeaks == nullptr) { return -1; }
    *nPeaks = 0;
    if (*std::max_element(x, x + n) < thr) { return 0; }
    std::vector<std::pair<T, int>> maxes;
    if (n == 1) {
        /* The reference reads x[1] here (out of bounds); a lone sample
         * above threshold is taken as a peak. */
        if (x[0] > thr) { maxes.emplace_back(x[0], 0); }
    } else {
        if (x[1] < x[0] && x[0] > thr) { maxes.emplace_back(x[0], 0); }
        for (int i = 1; i < n - 1; ++i) {
            if (x[i] >= x[i - 1] && x[i + 1] < x[i] && x[i] > thr) {
                maxes.emplace_back(x[i], i);
            }
        }
        if (x[n - 1] > x[n - 2] && x[n - 1] > thr) { maxes.emplace_back(x[n - 1], n - 1); }
    }
    if (maxes.empty()) { return 0; }
    if (minDist > 0) {
        std::stable_sort(maxes.begin(), maxes.end(),
                         [](const std::pair<T, int> &a, const std::pair<T, int> &b) {
                             return a.first > b.first;
                         });
        int np = 1;
        idx[0] = maxes[0].second;
        for (const auto &p : maxes) {
            int dmin = std::abs(idx[0] - p.second);
            for (int i = 1; i < np; ++i) { dmin = std::min(dmin, std::abs(idx[i] - p.second)); }
            if (dmin > minDist) { idx[np++] = p.second; }
        }
        std::sort(idx, idx + np);
        *nPeaks = np;
    } else {
        for (size_t i = 0; i < maxes.size(); ++i) { idx[i] = maxes[i].second; }
        *nPeaks = static_cast<int>(maxes.size());
    }
    return 0;
}

/* Two-pass Pearson at one lag.  Plain sums: testing/utilities.hpp:44-66. */
double pearsonPlain(int nb, const double *bt, double Eb, const double *x)
{
    double xm = 0;
    for (int j = 0; j < nb; ++j) { xm += x[j]; }
    xm /= nb;
    double ex = 0, dot = 0;
    for (int j = 0; j < nb; ++j) {
        double d = x[j] - xm;
        ex += d * d;
        dot += bt[j] * d;
    }
    ex = std::max(std::numeric_limits<double>::epsilon(), std::sqrt(ex));
    return dot / (Eb * ex);
}

template <typename X>
double pearsonComp(int nb, const double *bt, double Eb, const X *x)
{
    Neumaier m;
    for (int j = 0; j < nb; ++j) { m.add(static_cast<double>(x[j])); }
    const double xm = m.value() / nb;
    Neumaier ex, dot;
    for (int j = 0; j < nb; ++j) {
        double d = static_cast<double>(x[j]) - xm;
        ex.add(d * d);
        dot.add(bt[j] * d);
    }
    double e = std::max(std::numeric_limits<double>::epsilon(), std::sqrt(ex.value()));
    return dot.value() / (Eb * e);
}

template <typename X>
int pearsonOracle(int nb, const double *b, int nx, const X *x, int nLags,
                  const long long *lags, double *xc)
{
    if (nb < 1 || nx < nb || b == nullptr || x == nullptr || xc == nullptr) { return -1; }
    Neumaier bm;
    for (int j = 0; j < nb; ++j) { bm.add(b[j]); }
    const double mean = bm.value() / nb;
    std::vector<double> bt(nb);
    Neumaier eb;
    for (int j = 0; j < nb; ++j) { bt[j] = b[j] - mean; eb.add(bt[j] * bt[j]); }
    const double Eb = std::sqrt(eb.value());
    const long long nOut = static_cast<long long>(nx) - nb + 1;
    if (lags == nullptr) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (long long i = 0; i < nOut; ++i) { xc[i] = pearsonComp(nb, bt.data(), Eb, x + i); }
        return 0;
    }
    for (int q = 0; q < nLags; ++q) {
        if (lags[q] < 0 || lags[q] >= nOut) { return -1; }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int q = 0; q < nLags; ++q) { xc[q] = pearsonComp(nb, bt.data(), Eb, x + lags[q]); }
    return 0;
}

// 1 / ||w - mean(w)|| of the window w = x[i .. i + nb) (the Pearson scale
// sqrt(nb) / sqrt(nb S2 - S1^2) of src/matchedFilter.cpp:94-166, whose
// bracket equals nb * sum (w - mean)^2), two-pass with compensated sums;
// 0 for a constant window.
template <typename X>
double windowScale(int nb, const X *w)
{
    Neumaier sm;
    for (int j = 0; j < nb; ++j) { sm.add(static_cast<double>(w[j])); }
    const double mean = sm.value() / nb;
    Neumaier e2;
    for (int j = 0; j < nb; ++j) {
        const double d = static_cast<double>(w[j]) - mean;
        e2.add(d * d);
    }
    const double e = e2.value();
    return e > 0.0 ? 1.0 / std::sqrt(e) : 0.0;
}

}  // namespace

extern "C" {

int orc_window_scale_f32x(int nb, int nx, const float *x, int nLags, const long long *lags, double *out)
{
    if (nb < 1 || nx < nb || x == nullptr || out == nullptr) { return -1; }
    const long long nOut = static_cast<long long>(nx) - nb + 1;
    if (lags == nullptr) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (long long i = 0; i < nOut; ++i) { out[i] = windowScale<float>(nb, x + i); }
        return 0;
    }
    for (int q = 0; q < nLags; ++q) {
        if (lags[q] < 0 || lags[q] >= nOut) { return -1; }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int q = 0; q < nLags; ++q) { out[q] = windowScale<float>(nb, x + lags[q]); }
    return 0;
}

int orc_next_pow2(int n) { return nextPow2(n); }

int orc_window_length(int nb, int nx, int *nfft, int *blockLength)
{
    if (nfft == nullptr || blockLength == nullptr) { return -1; }
    return windowLength(nb, nx, nfft, blockLength);
}

int orc_demean_normalize_reverse_f64(int nb, const double *b, double *out)
{
    if (nb < 1 || b == nullptr || out == nullptr) { return -1; }
    demeanNormalizeReverse<double>(nb, b, out);
    return 0;
}

int orc_demean_normalize_reverse_f32(int nb, const float *b, float *out)
{
    if (nb < 1 || b == nullptr || out == nullptr) { return -1; }
    demeanNormalizeReverse<float>(nb, b, out);
    return 0;
}

int orc_naive_pearson_f64(int nb, const double *b, int nx, const double *x, double *xc)
{
    if (nb < 1 || nx < nb || b == nullptr || x == nullptr || xc == nullptr) { return -1; }
    double mean = std::accumulate(b, b + nb, 0.0) / nb;
    std::vector<double> bt(nb);
    double eb = 0;
    for (int j = 0; j < nb; ++j) { bt[j] = b[j] - mean; eb += bt[j] * bt[j]; }
    eb = std::sqrt(eb);
    for (int i = 0; i < nx - nb + 1; ++i) { xc[i] = pearsonPlain(nb, bt.data(), eb, x + i); }
    return 0;
}

int orc_pearson_oracle_f64(int nb, const double *b, int nx, const double *x,
                           int nLags, const long long *lags, double *xc)
{
    return pearsonOracle<double>(nb, b, nx, x, nLags, lags, xc);
}

int orc_pearson_oracle_f32x(int nb, const double *b, int nx, const float *x,
                            int nLags, const long long *lags, double *xc)
{
    return pearsonOracle<float>(nb, b, nx, x, nLags, lags, xc);
}

int orc_mf_apply_f64(int mode, int nT, const int *lens, const double *const *templates,
                     int n, const double *const *signals, int nThreads, double *out,
                     int *nfftOut, int *lOut)
{
    return mfApply<double>(mode, nT, lens, templates, n, signals, 0, nThreads, out,
                           nfftOut, lOut, 1.e-12);
}

int orc_mf_apply_f32(int mode, int nT, const int *lens, const double *const *templates,
                     int n, const float *const *signals, int floatQuirk, int nThreads,
                     float *out, int *nfftOut, int *lOut)
{
    return mfApply<float>(mode, nT, lens, templates, n, signals, floatQuirk, nThreads,
                          out, nfftOut, lOut, 1.e-6f);
}

int orc_shift_and_stack_f64(int nT, const double *rho, long long ld, int nOut,
                            const int *shifts, const double *weights, int absval,
                            double *out, int *outLen)
{
    return shiftAndStack<double>(nT, rho, ld, nOut, shifts, weights, absval, out, outLen);
}

int orc_reduce_correlations_f32(int nT, const float *rho, long long ld, int len, int absval, float *det,
                                int *id)
{
    return reduceCorrelations<float>(nT, rho, ld, len, absval, det, id);
}

int orc_reduce_correlations_f64(int nT, const double *rho, long long ld, int len, int absval, double *det,
                                int *id)
{
    return reduceCorrelations<double>(nT, rho, ld, len, absval, det, id);
}

int orc_quadratic_shift(int n, const double *xc, int peak, double dt, int absval, double *shift)
{
    if (n < 1 || xc == nullptr || shift == nullptr || peak < 0 || peak >= n) { return -1; }
    *shift = quadraticShift(n, xc, peak, dt, absval);
    return 0;
}

int orc_lanczos_shift(int n, const double *xc, int peak, double dt, int absval, int alpha, int nRefine,
                      double *shift, double *profile)
{
    if (n < 1 || xc == nullptr || shift == nullptr || peak < 0 || peak >= n || alpha < 1 || nRefine < 1) {
        return -1;
    }
    *shift = lanczosShift(n, xc, peak, dt, absval, alpha, nRefine, profile);
    return 0;
}

int orc_amplitude_factors(int Lt, const double *tmpl, const double *window, double *alphaGR, double *alphaSR,
                          int *computed)
{
    if (Lt < 1 || tmpl == nullptr || window == nullptr || alphaGR == nullptr || alphaSR == nullptr ||
        computed == nullptr) {
        return -1;
    }
    *alphaGR = 0;
    *alphaSR = 0;
    *computed = amplitudeFactors(Lt, tmpl, window, alphaGR, alphaSR);
    return 0;
}

int orc_shift_and_stack_f32(int nT, const float *rho, long long ld, int nOut,
                            const int *shifts, const double *weights, int absval,
                            float *out, int *outLen)
{
    return shiftAndStack<float>(nT, rho, ld, nOut, shifts, weights, absval, out, outLen);
}

int orc_peak_finder_f64(int n, const double *x, double threshold, int minDist,
                        int *idx, int *nPeaks)
{
    return peakFinder<double>(n, x, threshold, minDist, idx, nPeaks);
}

int orc_peak_finder_f32(int n, const float *x, float threshold, int minDist,
                        int *idx, int *nPeaks)
{
    return peakFinder<float>(n, x, threshold, minDist, idx, nPeaks);
}

int orc_rfft_f64(int N, const double *in, double *outInterleaved)
{
    if (N < 2 || (N & (N - 1)) != 0 || in == nullptr || outInterleaved == nullptr) { return -1; }
    RealFFT<double> f(N);
    std::vector<std::complex<double>> X(N / 2 + 1), work;
    f.forward(in, X.data(), work);
    for (int k = 0; k <= N / 2; ++k) { outInterleaved[2 * k] = X[k].real(); outInterleaved[2 * k + 1] = X[k].imag(); }
    return 0;
}

int orc_irfft_f64(int N, const double *inInterleaved, double *out)
{
    if (N < 2 || (N & (N - 1)) != 0 || inInterleaved == nullptr || out == nullptr) { return -1; }
    RealFFT<double> f(N);
    std::vector<std::complex<double>> X(N / 2 + 1), work;
    for (int k = 0; k <= N / 2; ++k) { X[k] = std::complex<double>(inInterleaved[2 * k], inInterleaved[2 * k + 1]); }
    f.inverse(X.data(), out, work);
    return 0;
}

}  /* extern "C" */

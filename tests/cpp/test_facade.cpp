// C++ API tests of the facade (MFLib::MatchedFilter, SingleChannel::
// MatchedFilter, PeakFinder, parameter classes) on the GPU engine.  The
// cases restate the reference's own gtest cases:
//   testing/matchedFilter.cpp:58-170  basicTestDouble   (rho within 1e-13)
//   testing/matchedFilter.cpp:172-273 basicTestFloat    (rho within 2e-6)
//   testing/matchedFilter.cpp:275-445 shiftAndStack     (stack max 1 +- 1e-3 at the origin)
//   testing/singleChannel.cpp:104-160 single channel    (rho within 1e-12)
//   testing/peakFinder.cpp:16-120     PeakFinder KATs   (13 and 21 indices)
//   testing/detector.cpp:100-283      Detector          (12 detections: times, templates,
//                                                        interpolated times, scaling factors)
//   testing/relativeMagnitude.cpp:20-124 RelativeMagnitude (log10 factors within 1e-12)
// Expected correlations are the two-pass FP64 Pearson definition
// (testing/utilities.hpp:33-69).  Prints one PASS/FAIL line per case and
// exits non-zero on any failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mflib/matchedFilter.hpp"
#include "mflib/matchedFilterParameters.hpp"
#include "mflib/singleChannel/matchedFilter.hpp"
#include "mflib/singleChannel/matchedFilterParameters.hpp"
#include "mflib/networkStationPhase.hpp"
#include "mflib/singleChannel/detection.hpp"
#include "mflib/singleChannel/detectionTimeInterpolationParameters.hpp"
#include "mflib/singleChannel/detector.hpp"
#include "mflib/singleChannel/detectorParameters.hpp"
#include "mflib/singleChannel/relativeMagnitude.hpp"
#include "mflib/waveformTemplate.hpp"
#include "private/peakFinder.hpp"

namespace
{
int failures = 0;

void expect(bool ok, const std::string &what)
{
    if (!ok) { throw std::runtime_error(what); }
}

void runCase(const char *name, const std::function<void()> &fn)
{
    try
    {
        fn();
        std::printf("PASS %s\n", name);
    }
    catch (const std::exception &e)
    {
        ++failures;
        std::printf("FAIL %s: %s\n", name, e.what());
    }
    std::fflush(stdout);
}

template<class E>
bool throwsType(const std::function<void()> &fn)
{
    try
    {
        fn();
    }
    catch (const E &)
    {
        return true;
    }
    catch (...)
    {
        return false;
    }
    return false;
}

// Two-pass Pearson correlation of template t against every window of x.
std::vector<double> pearson(const std::vector<double> &t, const std::vector<double> &x)
{
    const size_t nb = t.size();
    double tm = 0;
    for (double v : t) { tm += v; }
    tm /= static_cast<double>(nb);
    double tn = 0;
    for (double v : t) { tn += (v - tm) * (v - tm); }
    tn = std::sqrt(tn);
    std::vector<double> rho(x.size() - nb + 1, 0.0);
    for (size_t i = 0; i < rho.size(); ++i)
    {
        double xm = 0;
        for (size_t j = 0; j < nb; ++j) { xm += x[i + j]; }
        xm /= static_cast<double>(nb);
        double num = 0, xn = 0;
        for (size_t j = 0; j < nb; ++j)
        {
            num += (t[j] - tm) * (x[i + j] - xm);
            xn += (x[i + j] - xm) * (x[i + j] - xm);
        }
        const double den = tn * std::sqrt(xn);
        rho[i] = den > 0 ? num / den : 0.0;
    }
    return rho;
}

template<class T>
double maxDiff(const std::vector<T> &a, const std::vector<double> &b, size_t n)
{
    double e = 0;
    for (size_t i = 0; i < n; ++i) { e = std::max(e, std::abs(static_cast<double>(a[i]) - b[i])); }
    return e;
}

MFLib::WaveformTemplate makeTemplate(const std::vector<double> &x, double fs)
{
    MFLib::WaveformTemplate w;
    w.setSamplingRate(fs);
    w.setSignal(static_cast<int>(x.size()), x.data());
    return w;
}

// testing/matchedFilter.cpp:62-95: damped sine/cosine signals, 2 s
// sine/cosine templates.
struct BasicInputs
{
    std::vector<double> x, x2, b, b2;
    double fs = 0;
};
BasicInputs basicInputs()
{
    BasicInputs in;
    const int n = 1000;
    const double tmax = 8.0, dt = tmax / (n - 1), f0 = 4.0 / tmax, amp = 2.0;
    in.fs = 1.0 / dt;
    for (int i = 0; i < n; ++i)
    {
        const double t = i * dt;
        in.x.push_back(amp * std::exp(-0.1 * t) * std::sin(2 * M_PI * f0 * t));
        in.x2.push_back(amp * std::exp(-0.15 * t) * std::cos(2 * M_PI * f0 * t));
    }
    const int nb = static_cast<int>(2.0 / dt + 0.5) + 1;
    for (int i = 0; i < nb; ++i)
    {
        const double t = i * dt;
        in.b.push_back(std::sin(2 * M_PI * f0 * t));
        in.b2.push_back(std::cos(2 * M_PI * f0 * t));
    }
    return in;
}

template<class T>
void basicTest(double tol)
{
    const auto in = basicInputs();
    const int n = static_cast<int>(in.x.size());
    std::vector<T> x(in.x.begin(), in.x.end()), x2(in.x2.begin(), in.x2.end());
    std::vector<double> xd(x.begin(), x.end()), x2d(x2.begin(), x2.end());

    // one template
    MFLib::MatchedFilterParameters options;
    options.addTemplate(makeTemplate(in.b, in.fs));
    options.setSignalSize(n);
    expect(options.getNumberOfTemplates() == 1, "number of templates");
    MFLib::MatchedFilter<T> mf;
    expect(!mf.isInitialized(), "initialized too early");
    mf.initialize(options);
    expect(mf.isInitialized(), "not initialized");
    mf.setSignal(0, n, x.data());
    const T *sp = mf.getSignalPointer(0);
    for (int i = 0; i < n; ++i) { expect(sp[i] == x[i], "getSignalPointer"); }
    expect(!mf.haveMatchedFilteredSignals(), "results before apply");
    mf.apply();
    expect(mf.haveMatchedFilteredSignals(), "no results after apply");
    auto ref = pearson(in.b, xd);
    auto r = mf.getMatchedFilteredSignal(0);
    expect(static_cast<int>(r.size()) == mf.getFilteredSignalLength(), "filtered length");
    expect(r.size() == ref.size(), "filtered length vs n - nb + 1");
    expect(maxDiff(r, ref, r.size()) < tol, "rho (1 template)");

    // four templates, alternating signals
    options.clearTemplates();
    options.addTemplate(makeTemplate(in.b, in.fs));
    options.addTemplate(makeTemplate(in.b2, in.fs));
    options.addTemplate(makeTemplate(in.b2, in.fs));
    options.addTemplate(makeTemplate(in.b, in.fs));
    expect(options.getNumberOfTemplates() == 4, "number of templates (4)");
    mf.initialize(options);
    mf.setSignal(0, n, x.data());
    mf.setSignal(1, n, x2.data());
    mf.setSignal(2, n, x.data());
    mf.setSignal(3, n, x2.data());
    expect(mf.getNumberOfTemplates() == 4, "getNumberOfTemplates");
    mf.apply();
    const std::vector<double> *tp[4] = {&in.b, &in.b2, &in.b2, &in.b};
    const std::vector<double> *sg[4] = {&xd, &x2d, &xd, &x2d};
    for (int it = 0; it < 4; ++it)
    {
        auto rr = pearson(*tp[it], *sg[it]);
        auto got = mf.getMatchedFilteredSignal(it);
        expect(maxDiff(got, rr, got.size()) < tol, "rho (4 templates), row " + std::to_string(it));
        const T *p = mf.getMatchedFilterSignalPointer(it);
        for (size_t i = 0; i < got.size(); ++i) { expect(p[i] == got[i], "pointer and copy differ"); }
    }
}

// testing/matchedFilter.cpp:275-445: three channels, templates offset by
// their onsets, events at origin + travel time; the weighted stack of
// |rho| peaks at 1 at the origin sample.
void shiftAndStackTest()
{
    const double tmax = 120.0, df = 100.0, dt = 1 / df;
    const int n = static_cast<int>(tmax / dt + 0.5) + 1;
    const double tdur = 4.0;
    const int tsize = static_cast<int>(tdur / dt + 0.5) + 1;
    const double toff[3] = {1.0, 2.0, 1.5};
    const double tt[3] = {20.0, 40.0, 80.0};
    const double origin = 8.0;
    const double amps[3][2] = {{2.0, 1.0}, {1.5, 1.5}, {1.2, 2.5}};
    std::vector<double> sig[3], tpl[3];
    for (int k = 0; k < 3; ++k)
    {
        sig[k].assign(n, 0.0);
        for (int i = 0; i < n; ++i) { sig[k][i] = (i % 2 == 0) ? -0.001 : 0.001; }
        const int nextra = static_cast<int>(toff[k] / dt + 0.5);
        tpl[k].assign(tsize + nextra, 0.0);
        const int id = static_cast<int>((origin + tt[k]) / dt + 0.5);
        for (int i = 0; i < tsize; ++i)
        {
            const double t = i * dt;
            double v = amps[k][0] * std::exp(-amps[k][1] * t) * std::sin(2 * M_PI * t / tdur);
            if (i == 0 || i == tsize - 1) { v = 0.0; }
            tpl[k][i + nextra] = v;
            sig[k][id + i] = v;
        }
    }
    MFLib::MatchedFilterParameters options;
    for (int k = 0; k < 3; ++k)
    {
        auto w = makeTemplate(tpl[k], df);
        w.setPhaseOnsetTime(toff[k]);
        w.setTravelTime(tt[k]);
        options.addTemplate(w);
    }
    options.setSignalSize(n);
    MFLib::MatchedFilter<double> mf;
    mf.initialize(options);
    for (int k = 0; k < 3; ++k) { mf.setSignal(k, n, sig[k].data()); }
    expect(throwsType<std::runtime_error>([&] { mf.shiftAndStack(); }), "stack before apply must throw");
    mf.apply();
    for (int k = 0; k < 3; ++k)
    {
        auto ref = pearson(tpl[k], sig[k]);
        auto got = mf.getMatchedFilteredSignal(k);
        expect(maxDiff(got, ref, got.size()) <= 1e-11, "rho of channel " + std::to_string(k));
    }
    auto stack = mf.shiftAndStack();
    // closed form from the rows: mean of |rho_k[i + shift_k]|
    int shifts[3], smax = 0;
    for (int k = 0; k < 3; ++k)
    {
        shifts[k] = static_cast<int>((tt[k] - toff[k]) * df + 0.5);
        smax = std::max(smax, shifts[k]);
    }
    const int nf = mf.getFilteredSignalLength();
    expect(static_cast<int>(stack.size()) == nf - smax, "stack length");
    std::vector<double> rows[3];
    for (int k = 0; k < 3; ++k) { rows[k] = mf.getMatchedFilteredSignal(k); }
    double err = 0;
    for (size_t i = 0; i < stack.size(); ++i)
    {
        double s = 0;
        for (int k = 0; k < 3; ++k) { s += std::abs(rows[k][i + shifts[k]]) / 3.0; }
        err = std::max(err, std::abs(s - stack[i]));
    }
    expect(err < 1e-14, "stack vs closed form");
    const auto it = std::max_element(stack.begin(), stack.end());
    const double tori = static_cast<double>(it - stack.begin()) * dt;
    expect(std::abs(*it - 1.0) < 1e-3, "stack maximum " + std::to_string(*it));
    expect(std::abs(tori - origin) < dt / 2 + 1e-12, "origin time " + std::to_string(tori));

    // a template without travel time makes the instance unstackable
    MFLib::MatchedFilterParameters noShift;
    noShift.addTemplate(makeTemplate(tpl[0], df));
    noShift.setSignalSize(n);
    MFLib::MatchedFilter<double> mf2;
    mf2.initialize(noShift);
    mf2.setSignal(0, n, sig[0].data());
    mf2.apply();
    expect(throwsType<std::runtime_error>([&] { mf2.shiftAndStack(); }), "stack without shifts must throw");
}

// testing/singleChannel.cpp:104-160: templates cut from a U(2,3) signal,
// uneven lengths, one signal for all templates.
void singleChannelTest()
{
    std::mt19937_64 gen(0x5C0DE);
    std::uniform_real_distribution<double> u(2.0, 3.0);
    const int n = 4000;
    std::vector<double> wave(n);
    for (auto &v : wave) { v = u(gen); }
    const int starts[3] = {25, 1124, 3601}, lens[3] = {130, 140, 125};
    MFLib::SingleChannel::MatchedFilterParameters parms;
    expect(!parms.isValid(), "empty parameters valid");
    std::vector<double> tp[3];
    for (int k = 0; k < 3; ++k)
    {
        tp[k].assign(wave.begin() + starts[k], wave.begin() + starts[k] + lens[k]);
        parms.addTemplate(makeTemplate(tp[k], 100.0));
    }
    parms.setSignalSize(n);
    expect(parms.getMaxTemplateLength() == 140, "max template length");
    expect(parms.getDetectionMode() == MFLib::MatchedFilterDetectionMode::SINGLE_CHANNEL, "mode");
    MFLib::SingleChannel::MatchedFilter<double> mf;
    mf.initialize(parms);
    mf.setSignal(n, wave.data());
    mf.apply();
    const int nf = mf.getFilteredSignalLength();
    expect(nf == n - 140 + 1, "filtered length uses the longest template");
    for (int k = 0; k < 3; ++k)
    {
        auto ref = pearson(tp[k], wave);
        auto got = mf.getMatchedFilteredSignal(k);
        expect(maxDiff(got, ref, got.size()) < 1e-12, "rho of template " + std::to_string(k));
        expect(std::abs(got[starts[k]] - 1.0) < 1e-12, "rho at the cut is not 1");
    }
}

// testing/peakFinder.cpp:146-171 generator and 16-120 KATs.
std::vector<double> gaussianSine(double tdur, double dt, double sigma, double f)
{
    const double t0 = -tdur / 2, t1 = tdur / 2;
    const int npts = static_cast<int>((t1 - t0) / dt + 0.5) + 1;
    std::vector<double> r(npts);
    double amax = 0;
    for (int i = 0; i < npts; ++i)
    {
        const double t = t0 + i * dt;
        const double e = (t / sigma) * (t / sigma);
        r[i] = std::exp(-(e * e)) * std::cos(2 * M_PI * (t - t0) / tdur * f);
        amax = std::max(amax, std::abs(r[i]));
    }
    for (auto &v : r) { v /= amax; }
    r.front() = 0;
    r.back() = 0;
    return r;
}

void peakFinderTest()
{
    const int n = 4000;
    const double dt = 0.01;
    const auto tpl = gaussianSine(4.0, dt, 4.0 * 0.3, 12.0);
    std::vector<double> sig(n, 0.0);
    const int ih = static_cast<int>(tpl.size()) / 2;
    std::copy(tpl.begin() + ih, tpl.end(), sig.begin());
    for (double frac : {0.25, 0.5, 0.75})
    {
        const int i1 = static_cast<int>(n * frac) - ih;
        std::copy(tpl.begin(), tpl.end(), sig.begin() + i1);
    }
    std::copy(tpl.begin() + 1, tpl.begin() + 1 + ih, sig.begin() + (n - ih));
    const int minDist = static_cast<int>(0.5 / dt + 1) + 1;
    std::vector<int> ref13 = {0, 67, 933, 1000, 1067, 1933, 2000, 2067, 2933, 3000, 3067, n - 1 - 67, n - 1};
    std::vector<int> ref21 = {0, 33, 67};
    for (int c : {1000, 2000, 3000})
    {
        for (int d : {-67, -33, 0, 33, 67}) { ref21.push_back(c + d); }
    }
    for (int d : {67, 33, 0}) { ref21.push_back(n - 1 - d); }

    MFLib::PeakFinder<double> pf;
    expect(throwsType<std::invalid_argument>([&] { pf.setThreshold(1.5); }), "threshold range");
    pf.setThreshold(0.8);
    pf.setMinimumPeakDistance(minDist);
    pf.setSignal(n, sig.data());
    pf.apply();
    expect(pf.getNumberOfPeaks() == static_cast<int>(ref13.size()),
           "13-peak KAT count " + std::to_string(pf.getNumberOfPeaks()));
    std::vector<int> got(pf.getNumberOfPeaks());
    int *gp = got.data();
    pf.getPeakIndices(static_cast<int>(got.size()), &gp);
    expect(got == ref13, "13-peak KAT indices");
    pf.setMinimumPeakDistance(0);
    pf.apply();
    const int *p = pf.getPeakIndicesPointer();
    expect(pf.getNumberOfPeaks() == static_cast<int>(ref21.size()), "21-peak KAT count");
    expect(std::vector<int>(p, p + pf.getNumberOfPeaks()) == ref21, "21-peak KAT indices");

    MFLib::PeakFinder<float> pff;
    std::vector<float> sf(sig.begin(), sig.end());
    pff.setThreshold(0.8f);
    pff.setMinimumPeakDistance(minDist);
    pff.setSignal(n, sf.data());
    pff.apply();
    const int *pf13 = pff.getPeakIndicesPointer();
    expect(std::vector<int>(pf13, pf13 + pff.getNumberOfPeaks()) == ref13, "13-peak KAT (float)");
}

// testing/detector.cpp:100-283 (createTemplateSignal 285-301): five decaying
// sines; templates 0, 1, 2, 4 written three times each, phase-shifted and
// scaled by 1/4 (even) or 1/2 (odd), into 120 s of uniform noise.
std::vector<double> decayingSine(int npts, double amp, double dt, double f, double decay, double shift)
{
    std::vector<double> r(npts);
    for (int i = 0; i < npts; ++i)
    {
        const double t = i * dt;
        r[i] = amp * std::exp(-t * decay) * std::sin(2 * M_PI * f * (t - shift));
    }
    return r;
}

template<class T>
void detectorTest(MFLib::MatchedFilteredSignalDetectorPolicy policy, bool lanczos)
{
    const double df = 100, dt = 1 / df;
    const int minSpacing = static_cast<int>(4.0 / 2 / dt);
    const int nSignal = static_cast<int>(120.0 / dt + 0.5) + 1;
    const int nTemplate = static_cast<int>(4.0 / dt + 0.5) + 1;
    const double amp[5] = {1.0, 2.0, 0.5, 0.75, 3.0}, freq[5] = {10.0, 5.0, 7.5, 12.5, 2.5};
    const double decay[5] = {2.8, 2.9, 2.6, 2.2, 2.3}, phi[5] = {0.001, 0.002, 0.004, -0.003, -0.002};
    const std::vector<std::vector<double>> inserts = {{4, 28, 90}, {8, 24, 100}, {33, 66, 86}, {}, {16, 44, 106}};
    std::mt19937_64 gen(0xDE7EC7);
    std::uniform_real_distribution<double> noise(-0.05, 0.05);
    std::vector<double> signal(nSignal);
    for (auto &v : signal) { v = noise(gen); }
    MFLib::NetworkStationPhase nsp;
    nsp.setNetwork("UU");
    nsp.setStation("CCUT");
    nsp.setPhase("P");
    struct Expected { double onset, shifted, scale; int it; };
    std::vector<Expected> expected;
    MFLib::SingleChannel::MatchedFilterParameters parms;
    for (int i = 0; i < 5; ++i)
    {
        MFLib::WaveformTemplate wt = makeTemplate(decayingSine(nTemplate, amp[i], dt, freq[i], decay[i], 0), df);
        wt.setIdentifier({nsp, static_cast<uint64_t>(i)});
        parms.addTemplate(wt);
        const double scale = 1.0 / (i % 2 == 0 ? 4.0 : 2.0);
        const auto w = decayingSine(nTemplate, scale * amp[i], dt, freq[i], decay[i], phi[i]);
        for (double toff : inserts[i])
        {
            const int k = static_cast<int>(toff / dt + 0.5);
            std::copy(w.begin(), w.end(), signal.begin() + k);
            expected.push_back({toff, toff + phi[i], scale, i});
        }
    }
    std::sort(expected.begin(), expected.end(), [](const Expected &a, const Expected &b) { return a.onset < b.onset; });
    parms.setSignalSize(nSignal);
    MFLib::SingleChannel::MatchedFilter<T> mf;
    mf.initialize(parms);
    std::vector<T> sig(signal.begin(), signal.end());
    mf.setSignal(nSignal, sig.data());
    mf.apply();

    MFLib::SingleChannel::DetectionTimeInterpolationParameters interp;
    if (lanczos) { interp.enableLanczos(); }
    expect(interp.getLanczosAlpha() == 51 && interp.getLanczosNumberOfInterpolationPoints() == 100,
           "interpolation defaults");
    MFLib::SingleChannel::DetectorParameters dp;
    expect(dp.getDetectionThreshold() == 0.7 && dp.wantAmplitudeScalingFactor() && !dp.wantDetectedWaveform(),
           "detector defaults");
    dp.setDetectionTimeInterpolationParameters(interp);
    dp.setMinimumDetectionSpacing(minSpacing);
    dp.setMatchedFilteredSignalDetectorPolicy(policy);
    dp.enableSaveDetectedWaveform();
    dp.enableSaveAmplitudeScalingFactor();
    MFLib::SingleChannel::Detector<T> detector;
    expect(throwsType<std::invalid_argument>([&] { detector.detect(mf); }), "detect before initialize");
    detector.initialize(dp);
    detector.detect(mf);
    const int ndet = detector.getNumberOfDetections();
    if (policy == MFLib::MatchedFilteredSignalDetectorPolicy::REDUCED)
    {
        expect(ndet == static_cast<int>(expected.size()), "detections: " + std::to_string(ndet));
    }
    // one peak list per template also finds the other templates' events
    // (cross-correlations above 0.7); the written ones are among them
    int matched = 0;
    double lastTime = -1;
    for (int i = 0; i < ndet; ++i)
    {
        const auto &d = detector[i];
        expect(d.getDetectionTime() >= lastTime, "detections sorted by time");
        lastTime = d.getDetectionTime();
        for (const auto &e : expected)
        {
            if (std::abs(d.getDetectionTime() - e.onset) > 1e-10 ||
                d.getTemplateIdentifier().second != static_cast<uint64_t>(e.it))
            {
                continue;
            }
            ++matched;
            expect(d.getTemplateIdentifier().first == nsp, "template identifier");
            const double tol = lanczos ? dt / 2 : dt;
            expect(std::abs(e.shifted - d.getInterpolatedDetectionTime()) < tol,
                   "interpolated time " + std::to_string(d.getInterpolatedDetectionTime()) + " vs " +
                       std::to_string(e.shifted));
            const double a1 = d.getAmplitudeScalingFactor(MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006);
            const double a2 = d.getAmplitudeScalingFactor(MFLib::RelativeMagnitudeType::SCHAFF_RICHARDS_2014);
            expect(std::abs(a1 - e.scale) < 1e-2 && std::abs(a2 - e.scale) < 1e-2,
                   "scaling factors " + std::to_string(a1) + ", " + std::to_string(a2));
            expect(std::abs(d.getMagnitudePerturbation(MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006) -
                            std::log10(a1)) < 1e-6, "magnitude perturbation");
            expect(d.getCorrelationCoefficient() > 0.7 && d.getCorrelationCoefficient() <= 1, "coefficient");
            expect(d.getPolarity() == MFLib::Polarity::UNKNOWN, "unknown polarity stays unknown");
            expect(d.getDetectedSignalLength() == nTemplate, "detected waveform length");
            expect(d.getDetectedSignalPointer()[0] == sig[static_cast<size_t>(d.getDetectionTime() / dt + 0.5)],
                   "detected waveform samples");
            expect(!d.havePhaseOnsetTime() && !d.haveTravelTime(), "no onset or travel time on these templates");
        }
    }
    expect(matched == static_cast<int>(expected.size()), "matched " + std::to_string(matched) + " of 12");
    expect(throwsType<std::invalid_argument>([&] { (void)detector[ndet]; }), "index past the end");
}

// testing/relativeMagnitude.cpp:20-124: factors of random windows against a
// random template, checked through log10 against the test's own
// least-squares and standard-deviation-ratio formulas (get_dmag_1/2).
void relativeMagnitudeTest()
{
    const int n = 400;
    std::mt19937_64 gen(0x4E1A7);
    auto uniform = [&](double lo, double hi) {
        std::uniform_real_distribution<double> u(lo, hi);
        std::vector<double> v(n);
        for (auto &x : v) { x = u(gen); }
        return v;
    };
    auto demean = [](std::vector<double> v) {
        double m = 0;
        for (double x : v) { m += x; }
        m /= static_cast<double>(v.size());
        for (auto &x : v) { x -= m; }
        return v;
    };
    const auto t = uniform(-2, 1);
    MFLib::WaveformTemplate wt;
    wt.setSignal(n, t.data());
    MFLib::SingleChannel::RelativeMagnitude<double> mag;
    expect(throwsType<std::runtime_error>([&] { mag.getDetectedSignalLength(); }), "length before initialize");
    mag.initialize(wt);
    expect(mag.isInitialized() && mag.getDetectedSignalLength() == n && !mag.haveDetectedSignal(), "state");
    expect(throwsType<std::runtime_error>([&] {
               mag.computeAmplitudeScalingFactor(MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006);
           }), "factor before a window");
    const auto x = demean(t);
    for (auto range : {std::pair<double, double>(-3, 8), std::pair<double, double>(-7, 1)})
    {
        const auto sig = uniform(range.first, range.second);
        mag.setDetectedSignal(n, sig.data());
        expect(mag.haveDetectedSignal(), "window set");
        const auto y = demean(sig);
        double xy = 0, xx = 0, yy = 0;
        for (int i = 0; i < n; ++i)
        {
            xy += x[i] * y[i];
            xx += x[i] * x[i];
            yy += y[i] * y[i];
        }
        const double d1 = std::log10(std::abs(xy) / xx);
        const double d2 = std::log10(std::sqrt(yy / n) / std::sqrt(xx / n));
        MFLib::SingleChannel::RelativeMagnitude<double> copy(mag);
        const double m1 = copy.computeMagnitudePerturbation(MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006);
        const double m2 = copy.computeMagnitudePerturbation(MFLib::RelativeMagnitudeType::SCHAFF_RICHARDS_2014);
        expect(std::abs(m1 - d1) < 1e-12 && std::abs(m2 - d2) < 1e-12,
               "magnitudes " + std::to_string(m1 - d1) + ", " + std::to_string(m2 - d2));
    }
    std::vector<double> flat(n, 2.5);
    expect(throwsType<std::invalid_argument>([&] { mag.setDetectedSignal(n, flat.data()); }), "constant window");
    expect(!mag.haveDetectedSignal(), "a rejected window clears the state");
    expect(throwsType<std::invalid_argument>([&] { mag.setDetectedSignal(n - 1, t.data()); }), "window length");
    expect(throwsType<std::invalid_argument>([&] {
               MFLib::SingleChannel::convertAmplitudeScalingFactorToMagnitudePerturbation(0.0);
           }), "non-positive factor");
    MFLib::SingleChannel::RelativeMagnitude<float> magf;
    magf.initialize(wt);
    std::vector<float> sf(t.begin(), t.end());
    for (auto &v : sf) { v = 3.0f * v + 1.0f; }
    magf.setDetectedSignal(n, sf.data());
    expect(std::abs(magf.computeAmplitudeScalingFactor(MFLib::RelativeMagnitudeType::SCHAFF_RICHARDS_2014) - 3.0f) <
               1e-5f, "float instance, scaled copy");
}

void errorConventionTest()
{
    const auto in = basicInputs();
    const int n = static_cast<int>(in.x.size());
    MFLib::MatchedFilter<double> mf;
    expect(throwsType<std::runtime_error>([&] { mf.getNumberOfTemplates(); }), "query before initialize");
    expect(throwsType<std::runtime_error>([&] { mf.apply(); }), "apply before initialize");
    MFLib::MatchedFilterParameters bad;
    expect(throwsType<std::invalid_argument>([&] { mf.initialize(bad); }), "invalid parameters");
    MFLib::MatchedFilterParameters options;
    options.addTemplate(makeTemplate(in.b, in.fs));
    options.setSignalSize(n);
    mf.initialize(options);
    expect(throwsType<std::invalid_argument>([&] { mf.setSignal(0, n - 1, in.x.data()); }), "wrong signal size");
    expect(throwsType<std::invalid_argument>([&] { mf.setSignal(1, n, in.x.data()); }), "template index");
    expect(throwsType<std::invalid_argument>([&] { mf.getMatchedFilteredSignal(-1); }), "row index");
    MFLib::WaveformTemplate w;
    expect(throwsType<std::invalid_argument>([&] { w.setSamplingRate(0); }), "sampling rate");
    expect(throwsType<std::invalid_argument>([&] { w.setShiftAndStackWeight(1.5); }), "weight");
    expect(throwsType<std::runtime_error>([&] { w.getSignalLength(); }), "no signal");
    MFLib::WaveformTemplate w2 = makeTemplate(in.b, 50.0);
    MFLib::MatchedFilterParameters mixed;
    mixed.addTemplate(makeTemplate(in.b, in.fs));
    expect(throwsType<std::invalid_argument>([&] { mixed.addTemplate(w2); }), "sampling-rate mismatch");
    auto kat = MFLib::computeOptimalFFTAndBlockLength(4500, 75000);
    expect(kat.first == 32768 && kat.second == 28269, "window length KAT");
    expect(throwsType<std::invalid_argument>([&] { MFLib::computeOptimalFFTAndBlockLength(100, 100); }),
           "nx <= nb");
}
}  // namespace

int main()
{
    runCase("basicTestDouble", [] { basicTest<double>(1e-13); });
    runCase("basicTestFloat", [] { basicTest<float>(2e-6); });
    runCase("shiftAndStack", shiftAndStackTest);
    runCase("singleChannel", singleChannelTest);
    runCase("peakFinder", peakFinderTest);
    runCase("detectorReducedLanczos", [] {
        detectorTest<double>(MFLib::MatchedFilteredSignalDetectorPolicy::REDUCED, true);
    });
    runCase("detectorReducedQuadraticFloat", [] {
        detectorTest<float>(MFLib::MatchedFilteredSignalDetectorPolicy::REDUCED, false);
    });
    runCase("detectorSingleLanczos", [] {
        detectorTest<double>(MFLib::MatchedFilteredSignalDetectorPolicy::SINGLE, true);
    });
    runCase("relativeMagnitude", relativeMagnitudeTest);
    runCase("errorConventions", errorConventionTest);
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}

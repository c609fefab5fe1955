// SingleChannel::Detector<T>, DetectorParameters and
// DetectionTimeInterpolationParameters
// (include/mflib/singleChannel/detector*.hpp,
// detectionTimeInterpolationParameters.hpp) over the GPU engine.
//
// Detector::detect follows src/singleChannel/detector.cpp:420-691 in the
// reference.  The per-sample and per-peak arithmetic runs on the engine
// that holds the filter's correlations:
//  - REDUCED policy: mf_detect_reduced (reduceAllCorrelations + clamp +
//    PeakFinder, and the template of every peak);
//  - SINGLE policy: mf_detect_templates, every template's peak list in one
//    batch (setCorrelations + PeakFinder per template);
//  - mf_detection_products for all peaks at once: clamped rho, the refined
//    time shift and the amplitude scaling factors.
// The host assembles the Detection objects in the reference's order (peak
// order; for SINGLE, template by template, then sorted by detection time --
// a stable sort here, where the reference's std::sort leaves the order of
// equal times unspecified).  Setter checks apply as in the reference: a
// negative interpolated time throws from detect().
#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>
#include "abi.hpp"
#include "mflib/matchedFilter.hpp"
#include "mflib/networkStationPhase.hpp"
#include "mflib/singleChannel/detectionTimeInterpolationParameters.hpp"
#include "mflib/singleChannel/detector.hpp"
#include "mflib/singleChannel/detectorParameters.hpp"
#include "mflib/singleChannel/matchedFilter.hpp"
#include "mflib/waveformTemplate.hpp"

using namespace MFLib::SingleChannel;

// ---- DetectionTimeInterpolationParameters
// (src/singleChannel/detectionTimeInterpolationParameters.cpp) -----------
class DetectionTimeInterpolationParameters::InterpParametersImpl
{
public:
    int alpha = 51;
    int points = 100;
    MFLib::DetectionTimeInterpolationType type = MFLib::DetectionTimeInterpolationType::QUADRATIC;
};

DetectionTimeInterpolationParameters::DetectionTimeInterpolationParameters()
    : pImpl(std::make_unique<InterpParametersImpl>()) {}
DetectionTimeInterpolationParameters::DetectionTimeInterpolationParameters(
    const DetectionTimeInterpolationParameters &p)
    : pImpl(std::make_unique<InterpParametersImpl>(*p.pImpl)) {}
DetectionTimeInterpolationParameters::DetectionTimeInterpolationParameters(
    DetectionTimeInterpolationParameters &&p) noexcept
    : pImpl(std::move(p.pImpl)) {}
DetectionTimeInterpolationParameters::~DetectionTimeInterpolationParameters() = default;
DetectionTimeInterpolationParameters &DetectionTimeInterpolationParameters::operator=(
    const DetectionTimeInterpolationParameters &p)
{
    if (&p != this) { pImpl = std::make_unique<InterpParametersImpl>(*p.pImpl); }
    return *this;
}
DetectionTimeInterpolationParameters &DetectionTimeInterpolationParameters::operator=(
    DetectionTimeInterpolationParameters &&p) noexcept
{
    if (&p != this) { pImpl = std::move(p.pImpl); }
    return *this;
}
void DetectionTimeInterpolationParameters::clear() noexcept { *pImpl = InterpParametersImpl(); }
void DetectionTimeInterpolationParameters::enableLanczos() noexcept
{
    pImpl->type = MFLib::DetectionTimeInterpolationType::LANCZOS;
}
void DetectionTimeInterpolationParameters::setLanczosAlpha(int alpha)
{
    if (alpha < 1) { throw std::invalid_argument("alpha must be positive"); }
    pImpl->alpha = alpha;
}
int DetectionTimeInterpolationParameters::getLanczosAlpha() const noexcept { return pImpl->alpha; }
void DetectionTimeInterpolationParameters::setLanczosNumberOfInterpolationPoints(int npts)
{
    if (npts < 1) { throw std::invalid_argument("Number of interpolation points must be positive"); }
    pImpl->points = npts;
}
int DetectionTimeInterpolationParameters::getLanczosNumberOfInterpolationPoints() const noexcept
{
    return pImpl->points;
}
void DetectionTimeInterpolationParameters::enableQuadratic() noexcept
{
    pImpl->type = MFLib::DetectionTimeInterpolationType::QUADRATIC;
}
void DetectionTimeInterpolationParameters::disable() noexcept
{
    pImpl->type = MFLib::DetectionTimeInterpolationType::NONE;
}
MFLib::DetectionTimeInterpolationType DetectionTimeInterpolationParameters::getType() const noexcept
{
    return pImpl->type;
}

// ---- DetectorParameters (src/singleChannel/detectorParameters.cpp) -------
class DetectorParameters::DetectorParametersImpl
{
public:
    // clear() restores these but keeps the interpolation settings, as the
    // reference's clear() does (detectorParameters.cpp:68-78)
    void reset()
    {
        threshold = 0.7;
        maxPolicy = MFLib::MaximumMatchedFilterPolicy::MAXIMUM;
        minSpacing = 0;
        policy = MFLib::MatchedFilteredSignalDetectorPolicy::SINGLE;
        wantWaveform = false;
        wantAmplitude = true;
    }
    mutable DetectionTimeInterpolationParameters interp;
    double threshold = 0.7;
    MFLib::MaximumMatchedFilterPolicy maxPolicy = MFLib::MaximumMatchedFilterPolicy::MAXIMUM;
    int minSpacing = 0;
    MFLib::MatchedFilteredSignalDetectorPolicy policy = MFLib::MatchedFilteredSignalDetectorPolicy::SINGLE;
    bool wantWaveform = false;
    bool wantAmplitude = true;
};

DetectorParameters::DetectorParameters() : pImpl(std::make_unique<DetectorParametersImpl>()) {}
DetectorParameters::DetectorParameters(const DetectorParameters &p)
    : pImpl(std::make_unique<DetectorParametersImpl>(*p.pImpl)) {}
DetectorParameters::DetectorParameters(DetectorParameters &&p) noexcept : pImpl(std::move(p.pImpl)) {}
DetectorParameters &DetectorParameters::operator=(const DetectorParameters &p)
{
    if (&p != this) { pImpl = std::make_unique<DetectorParametersImpl>(*p.pImpl); }
    return *this;
}
DetectorParameters &DetectorParameters::operator=(DetectorParameters &&p) noexcept
{
    if (&p != this) { pImpl = std::move(p.pImpl); }
    return *this;
}
DetectorParameters::~DetectorParameters() = default;
void DetectorParameters::clear() noexcept { pImpl->reset(); }
void DetectorParameters::setMaximaPolicy(MFLib::MaximumMatchedFilterPolicy policy) noexcept
{
    pImpl->maxPolicy = policy;
}
MFLib::MaximumMatchedFilterPolicy DetectorParameters::getMaximaPolicy() const noexcept { return pImpl->maxPolicy; }
void DetectorParameters::setMinimumDetectionSpacing(int nSamples) noexcept { pImpl->minSpacing = nSamples; }
int DetectorParameters::getMinimumDetectionSpacing() const noexcept { return pImpl->minSpacing; }
void DetectorParameters::setDetectionThreshold(double tol)
{
    if (tol <= 0 || tol > 1)
    {
        throw std::invalid_argument("Threshold = " + std::to_string(tol) + " must be in range (0,1]");
    }
    pImpl->threshold = tol;
}
double DetectorParameters::getDetectionThreshold() const noexcept { return pImpl->threshold; }
void DetectorParameters::setMatchedFilteredSignalDetectorPolicy(
    MFLib::MatchedFilteredSignalDetectorPolicy policy) noexcept
{
    pImpl->policy = policy;
}
MFLib::MatchedFilteredSignalDetectorPolicy DetectorParameters::getMatchedFilteredSignalDetectorPolicy() const noexcept
{
    return pImpl->policy;
}
void DetectorParameters::setDetectionTimeInterpolationParameters(
    const DetectionTimeInterpolationParameters &interpParameters) noexcept
{
    pImpl->interp = interpParameters;
}
DetectionTimeInterpolationParameters &DetectorParameters::getDetectionTimeInterpolationParameters() const noexcept
{
    return pImpl->interp;
}
void DetectorParameters::enableSaveDetectedWaveform() noexcept { pImpl->wantWaveform = true; }
void DetectorParameters::disableSaveDetectedWaveform() noexcept { pImpl->wantWaveform = false; }
bool DetectorParameters::wantDetectedWaveform() const noexcept { return pImpl->wantWaveform; }
void DetectorParameters::enableSaveAmplitudeScalingFactor() noexcept { pImpl->wantAmplitude = true; }
void DetectorParameters::disableSaveAmplitudeScalingFactor() noexcept { pImpl->wantAmplitude = false; }
bool DetectorParameters::wantAmplitudeScalingFactor() const noexcept { return pImpl->wantAmplitude; }

// ---- Detector<T> -----------------------------------------------------------
template<class T>
class Detector<T>::DetectorImpl
{
public:
    std::vector<Detection<T>> detections;
    DetectorParameters parameters;
    bool useAbsoluteValue = false;
    bool initialized = false;
};

template<class T>
Detector<T>::Detector() : pImpl(std::make_unique<DetectorImpl>()) {}
template<class T>
Detector<T>::~Detector() = default;

template<class T>
void Detector<T>::clear() noexcept
{
    pImpl->detections.clear();
    pImpl->parameters.clear();
    pImpl->useAbsoluteValue = false;
    pImpl->initialized = false;
}

template<class T>
void Detector<T>::initialize(const DetectorParameters &parameters)
{
    clear();
    pImpl->parameters = parameters;
    pImpl->useAbsoluteValue = parameters.getMaximaPolicy() != MFLib::MaximumMatchedFilterPolicy::MAXIMUM;
    pImpl->initialized = true;
}

template<class T>
int Detector<T>::getNumberOfDetections() const noexcept
{
    return static_cast<int>(pImpl->detections.size());
}

template<class T>
bool Detector<T>::isInitialized() const noexcept { return pImpl->initialized; }

template<class T>
const Detection<T> &Detector<T>::operator[](size_t i) const
{
    const size_t n = pImpl->detections.size();
    if (i >= n)
    {
        throw std::invalid_argument(i == 0 ? std::string("No detections")
                                           : "i = " + std::to_string(i) + " cannot exceed " + std::to_string(n));
    }
    return pImpl->detections[i];
}

template<class T>
Detection<T> Detector<T>::getDetection(int id) const
{
    if (id < 0) { throw std::invalid_argument("id = " + std::to_string(id) + " must be non-negative"); }
    return (*this)[static_cast<size_t>(id)];
}

template<class T>
void Detector<T>::detect(const MatchedFilter<T> &mf)
{
    auto &d = *pImpl;
    d.detections.clear();
    if (!d.initialized) { throw std::invalid_argument("Detector class not initialized"); }
    if (!mf.haveMatchedFilteredSignals()) { throw std::invalid_argument("Matched filtered signals not computed"); }
    const int nt = mf.getNumberOfTemplates();
    if (nt < 1) { throw std::invalid_argument("No templates on matched filter class"); }
    const int detectionLength = mf.getFilteredSignalLength();
    if (detectionLength < 1) { throw std::invalid_argument("Matched filtered signal length is 0"); }
    mf_handle *h = MFLib::Facade::Access::handle(MFLib::Facade::Access::inner(mf));
    const auto &parms = d.parameters;
    // the reference clamps the threshold into [0, 1] (detector.cpp:398-402)
    const double threshold = std::min(1.0, std::max(0.0, parms.getDetectionThreshold()));
    const int minSpacing = parms.getMinimumDetectionSpacing();
    const int absval = d.useAbsoluteValue ? 1 : 0;

    // peaks and their templates
    std::vector<int> peaks(std::max(detectionLength, 1));
    std::vector<int> tids;
    int np = 0;
    if (parms.getMatchedFilteredSignalDetectorPolicy() == MFLib::MatchedFilteredSignalDetectorPolicy::REDUCED)
    {
        tids.resize(peaks.size());
        MFLib::Facade::check(mf_detect_reduced(h, 0, absval, threshold, minSpacing, peaks.data(), tids.data(),
                                               static_cast<int>(peaks.size()), &np),
                             "detect");
        peaks.resize(np);
        tids.resize(np);
    }
    else
    {
        // every template's peak list in one batch.  Room for 4 M peaks first;
        // only a pathological threshold needs the worst case (nOut/2 + 1 per
        // row), and then the batch runs again with that room.
        const long long worst = static_cast<long long>(nt) * (detectionLength / 2 + 1) + 1;
        std::vector<int> all(static_cast<size_t>(std::min<long long>(worst, 1LL << 22))), offsets(nt + 1);
        int rc = mf_detect_templates(h, absval, threshold, minSpacing, all.data(), static_cast<int>(all.size()),
                                     offsets.data());
        if (rc == MF_ERR_INVALID_ARGUMENT && static_cast<long long>(all.size()) < worst)
        {
            all.assign(static_cast<size_t>(std::min<long long>(worst, 0x7fffffffLL)), 0);
            rc = mf_detect_templates(h, absval, threshold, minSpacing, all.data(), static_cast<int>(all.size()),
                                     offsets.data());
        }
        MFLib::Facade::check(rc, "detect");
        np = offsets[nt];
        peaks.assign(all.begin(), all.begin() + np);
        tids.resize(np);
        for (int it = 0; it < nt; ++it)
        {
            std::fill(tids.begin() + offsets[it], tids.begin() + offsets[it + 1], it);
        }
    }
    if (np < 1) { return; }

    // per-template facts the products and the Detection objects need
    std::vector<MFLib::WaveformTemplate> templates(nt);
    std::vector<double> dts(nt);
    for (int it = 0; it < nt; ++it)
    {
        templates[it] = mf.getWaveformTemplate(it);
        if (!templates[it].haveIdentifier())
        {
            templates[it].setIdentifier(std::pair<MFLib::NetworkStationPhase, uint64_t>(
                MFLib::NetworkStationPhase(), static_cast<uint64_t>(it)));
        }
        dts[it] = 1 / templates[it].getSamplingRate();
    }
    const auto &interp = parms.getDetectionTimeInterpolationParameters();
    std::vector<double> rho(np), shift(np), alphaGR(np), alphaSR(np);
    std::vector<int> flags(np);
    MFLib::Facade::check(mf_detection_products(h, np, peaks.data(), tids.data(), dts.data(), absval,
                                               static_cast<int>(interp.getType()), interp.getLanczosAlpha(),
                                               interp.getLanczosNumberOfInterpolationPoints(), rho.data(),
                                               shift.data(), alphaGR.data(), alphaSR.data(), flags.data()),
                         "detect");

    const bool wantWaveform = parms.wantDetectedWaveform();
    const bool wantAmplitude = parms.wantAmplitudeScalingFactor();
    const int signalLength = mf.getSignalLength();
    const T *signal = wantWaveform ? mf.getSignalPointer() : nullptr;
    d.detections.reserve(np);
    for (int k = 0; k < np; ++k)
    {
        const int p = peaks[k];
        const int it = tids[k];
        const auto &wt = templates[it];
        Detection<T> det;
        det.setTemplateIdentifier(wt.getIdentifier());
        det.setCorrelationCoefficient(rho[k]);
        int templateLength = wt.getSignalLength();
        bool computeAmplitude = wantAmplitude;
        if (p + templateLength > signalLength)
        {
            std::fprintf(stderr, "Cannot compute amplitude for detection %d\n", k + 1);
            templateLength = signalLength - p;
            computeAmplitude = false;
        }
        if (wantWaveform) { det.setDetectedSignal(templateLength, signal + p); }
        const double detectionTime = p * dts[it];
        det.setDetectionTime(detectionTime);
        const double interpolatedTime = detectionTime + shift[k];
        det.setInterpolatedDetectionTime(interpolatedTime);
        if (wt.havePhaseOnsetTime())
        {
            det.setPhaseOnsetTime(detectionTime + wt.getPhaseOnsetTime());
            det.setInterpolatedPhaseOnsetTime(interpolatedTime + wt.getPhaseOnsetTime());
        }
        // the sign of rho at the peak flips (or keeps) the template's polarity
        const int sign = rho[k] < 0 ? -1 : (rho[k] == 0 ? 0 : 1);
        det.setPolarity(static_cast<MFLib::Polarity>(sign * static_cast<int>(wt.getPolarity())));
        if (wt.haveTravelTime()) { det.setTravelTime(wt.getTravelTime()); }
        if (computeAmplitude)
        {
            if (!(flags[k] & MF_PRODUCT_AMPLITUDE))
            {
                std::fprintf(stderr, "Magnitude calculation failed: Detected signal cannot be constant\n");
            }
            else
            {
                try
                {
                    det.setAmplitudeScalingFactor(static_cast<T>(alphaGR[k]),
                                                  MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006);
                    det.setAmplitudeScalingFactor(static_cast<T>(alphaSR[k]),
                                                  MFLib::RelativeMagnitudeType::SCHAFF_RICHARDS_2014);
                }
                catch (const std::exception &e)
                {
                    std::fprintf(stderr, "Magnitude calculation failed: %s\n", e.what());
                }
            }
        }
        d.detections.push_back(std::move(det));
    }
    if (parms.getMatchedFilteredSignalDetectorPolicy() == MFLib::MatchedFilteredSignalDetectorPolicy::SINGLE)
    {
        std::stable_sort(d.detections.begin(), d.detections.end(), [](const Detection<T> &a, const Detection<T> &b) {
            return a.getDetectionTime() < b.getDetectionTime();
        });
    }
}

template class MFLib::SingleChannel::Detector<double>;
template class MFLib::SingleChannel::Detector<float>;

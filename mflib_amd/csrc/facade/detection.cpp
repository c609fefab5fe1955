// SingleChannel::Detection<T> (include/mflib/singleChannel/detection.hpp):
// a value object.  Setter checks and getter exceptions follow
// src/singleChannel/detection.cpp:124-474 in the reference; one
// difference: haveCorrelationCoefficient() reports whether the value was
// set, where the reference returns the value itself converted to bool
// (so a coefficient of exactly 0 reads as unset there).
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>
#include "mflib/networkStationPhase.hpp"
#include "mflib/singleChannel/detection.hpp"

using namespace MFLib::SingleChannel;

namespace
{
int typeIndex(MFLib::RelativeMagnitudeType type)
{
    return type == MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006 ? 0 : 1;
}

template<class T>
struct Scaling
{
    T alpha = 0;
    T dmag = 0;  // log10(alpha): the magnitude perturbation
    bool set = false;
};

// A time with its "have" flag; negative times are rejected.
struct Time
{
    double value = 0;
    bool set = false;
    void assign(double t, const char *what)
    {
        set = false;
        if (t < 0) { throw std::invalid_argument(std::string(what) + " must be positive"); }
        value = t;
        set = true;
    }
    double get(const char *what) const
    {
        if (!set) { throw std::invalid_argument(std::string(what) + " not yet set"); }
        return value;
    }
};
}  // namespace

template<class T>
class Detection<T>::DetectionImpl
{
public:
    std::pair<MFLib::NetworkStationPhase, uint64_t> id;
    bool haveId = false;
    double rho = 0;
    bool haveRho = false;
    std::vector<T> waveform;
    bool haveWaveform = false;
    Time detection, interpolatedDetection, onset, interpolatedOnset;
    double travel = 0;
    bool haveTravel = false;
    std::array<Scaling<T>, 2> scaling;
    MFLib::Polarity polarity = MFLib::Polarity::UNKNOWN;
};

template<class T>
Detection<T>::Detection() : pImpl(std::make_unique<DetectionImpl>()) {}
template<class T>
Detection<T>::Detection(const Detection &d) : pImpl(std::make_unique<DetectionImpl>(*d.pImpl)) {}
template<class T>
Detection<T>::Detection(Detection &&d) noexcept : pImpl(std::move(d.pImpl)) {}
template<class T>
Detection<T> &Detection<T>::operator=(const Detection &d)
{
    if (&d != this) { pImpl = std::make_unique<DetectionImpl>(*d.pImpl); }
    return *this;
}
template<class T>
Detection<T> &Detection<T>::operator=(Detection &&d) noexcept
{
    if (&d != this) { pImpl = std::move(d.pImpl); }
    return *this;
}
template<class T>
Detection<T>::~Detection() = default;

template<class T>
void Detection<T>::clear() noexcept
{
    *pImpl = DetectionImpl();
}

template<class T>
void Detection<T>::setTemplateIdentifier(const std::pair<NetworkStationPhase, uint64_t> &id) noexcept
{
    pImpl->id = id;
    pImpl->haveId = true;
}
template<class T>
std::pair<MFLib::NetworkStationPhase, uint64_t> Detection<T>::getTemplateIdentifier() const
{
    if (!pImpl->haveId) { throw std::runtime_error("Template ID not set"); }
    return pImpl->id;
}
template<class T>
bool Detection<T>::haveTemplateIdentifier() const noexcept { return pImpl->haveId; }

template<class T>
void Detection<T>::setCorrelationCoefficient(double value)
{
    pImpl->haveRho = false;
    if (value < -1 || value > 1)
    {
        throw std::invalid_argument("Correlation coefficient = " + std::to_string(value) +
                                    " must be in range [-1,1]");
    }
    pImpl->rho = value;
    pImpl->haveRho = true;
}
template<class T>
double Detection<T>::getCorrelationCoefficient() const
{
    if (!pImpl->haveRho) { throw std::runtime_error("Similarity score not yet set"); }
    return pImpl->rho;
}
template<class T>
bool Detection<T>::haveCorrelationCoefficient() const noexcept { return pImpl->haveRho; }

template<class T>
void Detection<T>::setDetectedSignal(int n, const T x[])
{
    pImpl->haveWaveform = false;
    if (n < 1) { throw std::invalid_argument("n must be positive"); }
    if (x == nullptr) { throw std::invalid_argument("x is NULL"); }
    pImpl->waveform.assign(x, x + n);
    pImpl->haveWaveform = true;
}
template<class T>
void Detection<T>::getDetectedSignal(int n, T *x[]) const
{
    const int npts = getDetectedSignalLength();
    if (n < npts)
    {
        throw std::invalid_argument("Signal length = " + std::to_string(n) + " must be at least = " +
                                    std::to_string(npts));
    }
    if (x == nullptr || *x == nullptr) { throw std::invalid_argument("x is NULL"); }
    std::copy(pImpl->waveform.begin(), pImpl->waveform.end(), *x);
}
template<class T>
const T *Detection<T>::getDetectedSignalPointer() const
{
    if (!pImpl->haveWaveform) { throw std::runtime_error("Detected signal not yet set"); }
    return pImpl->waveform.data();
}
template<class T>
int Detection<T>::getDetectedSignalLength() const
{
    if (!pImpl->haveWaveform) { throw std::runtime_error("Detected signal not yet set"); }
    return static_cast<int>(pImpl->waveform.size());
}
template<class T>
bool Detection<T>::haveDetectedSignal() const noexcept { return pImpl->haveWaveform; }

template<class T>
void Detection<T>::setDetectionTime(double t) { pImpl->detection.assign(t, "Detection onset time"); }
template<class T>
double Detection<T>::getDetectionTime() const { return pImpl->detection.get("Detection time"); }
template<class T>
bool Detection<T>::haveDetectionTime() const noexcept { return pImpl->detection.set; }
template<class T>
void Detection<T>::setInterpolatedDetectionTime(double t)
{
    pImpl->interpolatedDetection.assign(t, "Detection onset time");
}
template<class T>
double Detection<T>::getInterpolatedDetectionTime() const
{
    return pImpl->interpolatedDetection.get("Detection time");
}
template<class T>
bool Detection<T>::haveInterpolatedDetectionTime() const noexcept { return pImpl->interpolatedDetection.set; }
template<class T>
void Detection<T>::setPhaseOnsetTime(double t) { pImpl->onset.assign(t, "Phase onset time"); }
template<class T>
double Detection<T>::getPhaseOnsetTime() const { return pImpl->onset.get("Phase onset time"); }
template<class T>
bool Detection<T>::havePhaseOnsetTime() const noexcept { return pImpl->onset.set; }
template<class T>
void Detection<T>::setInterpolatedPhaseOnsetTime(double t) { pImpl->interpolatedOnset.assign(t, "Phase onset time"); }
template<class T>
double Detection<T>::getInterpolatedPhaseOnsetTime() const { return pImpl->interpolatedOnset.get("Phase onset time"); }
template<class T>
bool Detection<T>::haveInterpolatedPhaseOnsetTime() const noexcept { return pImpl->interpolatedOnset.set; }

template<class T>
void Detection<T>::setTravelTime(double travelTime)
{
    if (travelTime < 0) { throw std::invalid_argument("Travel time cannot be negative"); }
    pImpl->travel = travelTime;
    pImpl->haveTravel = true;
}
template<class T>
double Detection<T>::getTravelTime() const
{
    if (!pImpl->haveTravel) { throw std::runtime_error("Travel time not yet set"); }
    return pImpl->travel;
}
template<class T>
bool Detection<T>::haveTravelTime() const noexcept { return pImpl->haveTravel; }

template<class T>
void Detection<T>::setAmplitudeScalingFactor(T value, MFLib::RelativeMagnitudeType type)
{
    auto &s = pImpl->scaling[typeIndex(type)];
    s.alpha = 1;
    s.dmag = 0;
    s.set = false;
    if (value <= 0) { throw std::invalid_argument("Scaling factor = " + std::to_string(value) + " must be positive"); }
    s.alpha = value;
    s.dmag = std::log10(value);
    s.set = true;
}
template<class T>
T Detection<T>::getAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const
{
    const auto &s = pImpl->scaling[typeIndex(type)];
    if (!s.set) { throw std::runtime_error("Scaling factor not yet set for this type"); }
    return s.alpha;
}
template<class T>
T Detection<T>::getMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const
{
    const auto &s = pImpl->scaling[typeIndex(type)];
    if (!s.set) { throw std::runtime_error("Mag pert not yet set for this type"); }
    return s.dmag;
}
template<class T>
bool Detection<T>::haveAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const noexcept
{
    return pImpl->scaling[typeIndex(type)].set;
}
template<class T>
bool Detection<T>::haveMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const noexcept
{
    return pImpl->scaling[typeIndex(type)].set;
}

template<class T>
void Detection<T>::setPolarity(MFLib::Polarity polarity) noexcept { pImpl->polarity = polarity; }
template<class T>
MFLib::Polarity Detection<T>::getPolarity() const noexcept { return pImpl->polarity; }

template class MFLib::SingleChannel::Detection<double>;
template class MFLib::SingleChannel::Detection<float>;

// WaveformTemplate (include/mflib/waveformTemplate.hpp).  Validation and
// defaults follow src/waveformTemplate.cpp:140-420 in the reference:
//  - the signal is kept in FP64 (196-236);
//  - setting the signal or the sampling rate forgets the onset time (157, 206);
//  - the onset must not exceed (npts-1)*samplingRate (287-303; the
//    reference multiplies by the rate, not the period, and this library
//    accepts the same inputs);
//  - a negative travel time throws std::runtime_error (309-315).
#include <stdexcept>
#include <string>
#include <vector>
#include "mflib/networkStationPhase.hpp"
#include "mflib/waveformTemplate.hpp"

using namespace MFLib;

class WaveformTemplate::WaveformTemplateImpl
{
public:
    std::vector<double> signal;
    double samplingRate = 0;
    double weight = 1;
    double onset = -1;
    double travel = -1;
    double magnitude = 0;
    std::pair<NetworkStationPhase, uint64_t> identifier{NetworkStationPhase(), 0};
    Polarity polarity = Polarity::UNKNOWN;
    bool haveIdentifier = false;
    bool haveMagnitude = false;
};

WaveformTemplate::WaveformTemplate() : pImpl(std::make_unique<WaveformTemplateImpl>()) {}
WaveformTemplate::WaveformTemplate(const WaveformTemplate &t) :
    pImpl(std::make_unique<WaveformTemplateImpl>(*t.pImpl)) {}
WaveformTemplate::WaveformTemplate(WaveformTemplate &&t) noexcept : pImpl(std::move(t.pImpl))
{
    t.pImpl = std::make_unique<WaveformTemplateImpl>();
}
WaveformTemplate& WaveformTemplate::operator=(const WaveformTemplate &t)
{
    if (&t != this) { *pImpl = *t.pImpl; }
    return *this;
}
WaveformTemplate& WaveformTemplate::operator=(WaveformTemplate &&t) noexcept
{
    if (&t != this) { std::swap(pImpl, t.pImpl); }
    return *this;
}
WaveformTemplate::~WaveformTemplate() = default;
void WaveformTemplate::clear() noexcept { *pImpl = WaveformTemplateImpl(); }

template<class U>
static void storeSignal(std::vector<double> &dst, int npts, const U *x)
{
    if (npts < 1) { throw std::invalid_argument("npts must be positive"); }
    if (x == nullptr) { throw std::invalid_argument("x is NULL"); }
    dst.assign(x, x + npts);
}
void WaveformTemplate::setSignal(int npts, const double x[])
{
    storeSignal(pImpl->signal, npts, x);
    pImpl->onset = -1;
}
void WaveformTemplate::setSignal(int npts, const float x[])
{
    storeSignal(pImpl->signal, npts, x);
    pImpl->onset = -1;
}

template<class U>
static void loadSignal(const std::vector<double> &src, int maxx, U *x)
{
    if (maxx < static_cast<int>(src.size()))
    {
        throw std::invalid_argument("maxx = " + std::to_string(maxx) + " must be at least " +
                                    std::to_string(src.size()));
    }
    if (x == nullptr) { throw std::invalid_argument("x is NULL"); }
    for (size_t i = 0; i < src.size(); ++i) { x[i] = static_cast<U>(src[i]); }
}
void WaveformTemplate::getSignal(int maxx, double *x[]) const
{
    getSignalLength();
    loadSignal(pImpl->signal, maxx, x ? *x : nullptr);
}
void WaveformTemplate::getSignal(int maxx, float *x[]) const
{
    getSignalLength();
    loadSignal(pImpl->signal, maxx, x ? *x : nullptr);
}
int WaveformTemplate::getSignalLength() const
{
    if (!haveSignal()) { throw std::runtime_error("Signal not set"); }
    return static_cast<int>(pImpl->signal.size());
}
bool WaveformTemplate::haveSignal() const noexcept { return !pImpl->signal.empty(); }

void WaveformTemplate::setSamplingRate(double fs)
{
    if (!(fs > 0)) { throw std::invalid_argument("Sampling rate = " + std::to_string(fs) + " must be positive"); }
    pImpl->onset = -1;
    pImpl->samplingRate = fs;
}
double WaveformTemplate::getSamplingRate() const
{
    if (!haveSamplingRate()) { throw std::runtime_error("Sampling rate not set"); }
    return pImpl->samplingRate;
}
bool WaveformTemplate::haveSamplingRate() const noexcept { return pImpl->samplingRate > 0; }

void WaveformTemplate::setShiftAndStackWeight(double w)
{
    if (!(w >= 0 && w <= 1)) { throw std::invalid_argument("weight = " + std::to_string(w) + " must be in range [0,1]"); }
    pImpl->weight = w;
}
double WaveformTemplate::getShiftAndStackWeight() const noexcept { return pImpl->weight; }

void WaveformTemplate::setPhaseOnsetTime(double t)
{
    if (t < 0) { throw std::invalid_argument("onsetTime = " + std::to_string(t) + " must be positive"); }
    const double fs = getSamplingRate();
    const int npts = getSignalLength();
    const double tmax = static_cast<double>(npts - 1) * fs;
    if (t > tmax)
    {
        throw std::invalid_argument("onsetTime = " + std::to_string(t) + " cannot exceed maxOnsetTime = " +
                                    std::to_string(tmax));
    }
    pImpl->onset = t;
}
double WaveformTemplate::getPhaseOnsetTime() const
{
    if (!havePhaseOnsetTime()) { throw std::runtime_error("Onset time not set"); }
    return pImpl->onset;
}
bool WaveformTemplate::havePhaseOnsetTime() const noexcept { return pImpl->onset >= 0; }

void WaveformTemplate::setTravelTime(double t)
{
    if (t < 0) { throw std::runtime_error("Travel time cannot be negative"); }
    pImpl->travel = t;
}
double WaveformTemplate::getTravelTime() const
{
    if (!haveTravelTime()) { throw std::runtime_error("Travel time not yet set"); }
    return pImpl->travel;
}
bool WaveformTemplate::haveTravelTime() const noexcept { return pImpl->travel >= 0; }

void WaveformTemplate::setIdentifier(const std::pair<NetworkStationPhase, uint64_t> &id) noexcept
{
    pImpl->identifier = id;
    pImpl->haveIdentifier = true;
}
std::pair<NetworkStationPhase, uint64_t> WaveformTemplate::getIdentifier() const
{
    if (!haveIdentifier()) { throw std::runtime_error("Waveform identifier not set"); }
    return pImpl->identifier;
}
bool WaveformTemplate::haveIdentifier() const noexcept { return pImpl->haveIdentifier; }

void WaveformTemplate::setMagnitude(double m) noexcept
{
    pImpl->magnitude = m;
    pImpl->haveMagnitude = true;
}
double WaveformTemplate::getMagnitude() const
{
    if (!haveMagnitude()) { throw std::runtime_error("Magnitude not set"); }
    return pImpl->magnitude;
}
bool WaveformTemplate::haveMagnitude() const noexcept { return pImpl->haveMagnitude; }
void WaveformTemplate::setPolarity(Polarity p) noexcept { pImpl->polarity = p; }
Polarity WaveformTemplate::getPolarity() const noexcept { return pImpl->polarity; }

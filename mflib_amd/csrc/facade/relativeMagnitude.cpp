// SingleChannel::RelativeMagnitude<T>
// (include/mflib/singleChannel/relativeMagnitude.hpp) over the C ABI's
// mf_amplitude_factors: the template's FP64 samples stay here, and each
// detected window is scored on the GPU when it is set.  Checks and
// exceptions follow src/singleChannel/relativeMagnitude.cpp:273-388 in the
// reference.  Both precisions score in FP64; the reference's float
// instantiation subtracts the float sum instead of the mean from the window
// (110-125), which this does not reproduce.
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>
#include "abi.hpp"
#include "mflib/singleChannel/relativeMagnitude.hpp"
#include "mflib/waveformTemplate.hpp"

using namespace MFLib::SingleChannel;

template<class T>
class RelativeMagnitude<T>::RelativeMagnitudeImpl
{
public:
    std::vector<double> tmpl;
    double alphaGR = 1;
    double alphaSR = 1;
    bool haveSignal = false;
    bool initialized = false;
};

template<class T>
RelativeMagnitude<T>::RelativeMagnitude() : pImpl(std::make_unique<RelativeMagnitudeImpl>()) {}
template<class T>
RelativeMagnitude<T>::RelativeMagnitude(const RelativeMagnitude &m)
    : pImpl(std::make_unique<RelativeMagnitudeImpl>(*m.pImpl)) {}
template<class T>
RelativeMagnitude<T>::RelativeMagnitude(RelativeMagnitude &&m) noexcept : pImpl(std::move(m.pImpl)) {}
template<class T>
RelativeMagnitude<T> &RelativeMagnitude<T>::operator=(const RelativeMagnitude &m)
{
    if (&m != this) { pImpl = std::make_unique<RelativeMagnitudeImpl>(*m.pImpl); }
    return *this;
}
template<class T>
RelativeMagnitude<T> &RelativeMagnitude<T>::operator=(RelativeMagnitude &&m) noexcept
{
    if (&m != this) { pImpl = std::move(m.pImpl); }
    return *this;
}
template<class T>
RelativeMagnitude<T>::~RelativeMagnitude() = default;

template<class T>
void RelativeMagnitude<T>::clear() noexcept
{
    if (pImpl) { *pImpl = RelativeMagnitudeImpl(); }
    else { pImpl = std::make_unique<RelativeMagnitudeImpl>(); }
}

template<class T>
void RelativeMagnitude<T>::initialize(const MFLib::WaveformTemplate &wt)
{
    clear();
    const int n = wt.getSignalLength();
    if (n < 1)
    {
        throw std::invalid_argument("Template signal length = " + std::to_string(n) + " must be positive");
    }
    pImpl->tmpl.resize(n);
    double *p = pImpl->tmpl.data();
    wt.getSignal(n, &p);
    pImpl->initialized = true;
}

template<class T>
int RelativeMagnitude<T>::getDetectedSignalLength() const
{
    if (!isInitialized()) { throw std::runtime_error("Class not initialized"); }
    return static_cast<int>(pImpl->tmpl.size());
}

template<class T>
bool RelativeMagnitude<T>::isInitialized() const noexcept { return pImpl && pImpl->initialized; }

template<class T>
void RelativeMagnitude<T>::setDetectedSignal(int n, const T y[])
{
    auto &d = *pImpl;
    d.haveSignal = false;
    d.alphaGR = 1;
    d.alphaSR = 1;
    const int nref = getDetectedSignalLength();
    if (n != nref)
    {
        throw std::invalid_argument("Signal length = " + std::to_string(n) + " must = " + std::to_string(nref));
    }
    if (y == nullptr) { throw std::invalid_argument("Signal is NULL"); }
    int computed = 0;
    MFLib::Facade::check(mf_amplitude_factors(MFLib::Facade::precisionOf<T>(), n, d.tmpl.data(), 1, y, n, &d.alphaGR,
                                              &d.alphaSR, &computed),
                         "setDetectedSignal");
    if (!computed) { throw std::invalid_argument("Detected signal cannot be constant"); }
    d.haveSignal = true;
}

template<class T>
bool RelativeMagnitude<T>::haveDetectedSignal() const noexcept { return pImpl && pImpl->haveSignal; }

template<class T>
T RelativeMagnitude<T>::computeAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const
{
    if (!isInitialized()) { throw std::runtime_error("Class is not initialized"); }
    if (!pImpl->haveSignal) { throw std::runtime_error("Detected waveform not set"); }
    return static_cast<T>(type == MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006 ? pImpl->alphaGR
                                                                                    : pImpl->alphaSR);
}

template<class T>
T RelativeMagnitude<T>::computeMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const
{
    return convertAmplitudeScalingFactorToMagnitudePerturbation<T>(computeAmplitudeScalingFactor(type));
}

template<class T>
T MFLib::SingleChannel::convertAmplitudeScalingFactorToMagnitudePerturbation(T alpha)
{
    if (alpha <= 0) { throw std::invalid_argument("Scaling factor = " + std::to_string(alpha) + " must be positive"); }
    return std::log10(alpha);
}

template class MFLib::SingleChannel::RelativeMagnitude<double>;
template class MFLib::SingleChannel::RelativeMagnitude<float>;
template double MFLib::SingleChannel::convertAmplitudeScalingFactorToMagnitudePerturbation<double>(double);
template float MFLib::SingleChannel::convertAmplitudeScalingFactorToMagnitudePerturbation<float>(float);

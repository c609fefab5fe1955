#ifndef MFLIB_SINGLECHANNEL_RELATIVEMAGNITUDE_HPP
#define MFLIB_SINGLECHANNEL_RELATIVEMAGNITUDE_HPP
#include <memory>
#include "mflib/enums.hpp"
namespace MFLib
{
class WaveformTemplate;
namespace SingleChannel
{
/// Amplitude of a detected window relative to its template
/// (include/mflib/singleChannel/relativeMagnitude.hpp in the reference):
/// the least-squares scaling factor of Gibbons & Ringdal (2006),
/// |<x, y>| / <x, x>, and the norm ratio of Schaff & Richards (2014),
/// sqrt(<y, y> / <x, x>), of the demeaned template x and window y.  The
/// factors are computed on the GPU in FP64 when the window is set.
template<class T = double>
class RelativeMagnitude
{
public:
    RelativeMagnitude();
    RelativeMagnitude(const RelativeMagnitude &magnitude);
    RelativeMagnitude(RelativeMagnitude &&magnitude) noexcept;
    RelativeMagnitude& operator=(const RelativeMagnitude &magnitude);
    RelativeMagnitude& operator=(RelativeMagnitude &&magnitude) noexcept;
    ~RelativeMagnitude();
    void clear() noexcept;
    /// @throws std::invalid_argument if the template has no samples.
    void initialize(const WaveformTemplate &wt);
    /// The template length, which every detected window must have.
    int getDetectedSignalLength() const;
    bool isInitialized() const noexcept;
    /// @throws std::invalid_argument if n differs from the template length,
    ///         y is NULL or the window is constant.
    void setDetectedSignal(int n, const T y[]);
    bool haveDetectedSignal() const noexcept;
    /// @throws std::runtime_error before initialize() or setDetectedSignal().
    T computeAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const;
    /// log10 of the scaling factor.
    T computeMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const;
private:
    class RelativeMagnitudeImpl;
    mutable std::unique_ptr<RelativeMagnitudeImpl> pImpl;
};
/// log10(alpha); @throws std::invalid_argument unless alpha > 0.
template<class T>
T convertAmplitudeScalingFactorToMagnitudePerturbation(T alpha);
}
}
#endif

#ifndef MFLIB_SINGLECHANNEL_DETECTION_HPP
#define MFLIB_SINGLECHANNEL_DETECTION_HPP
#include <cstdint>
#include <memory>
#include <utility>
#include "mflib/enums.hpp"
namespace MFLib
{
class NetworkStationPhase;
namespace SingleChannel
{
/// One detection of the single-channel detector
/// (include/mflib/singleChannel/detection.hpp in the reference): template
/// identifier, correlation coefficient, integer and interpolated detection
/// and phase-onset times, travel time, polarity, amplitude scaling factors
/// (and their log10, the magnitude perturbations) and, optionally, the
/// detected waveform.  Getters of unset values throw.
template<class T = double>
class Detection
{
public:
    Detection();
    Detection(const Detection &detection);
    Detection(Detection &&detection) noexcept;
    Detection& operator=(const Detection &detection);
    Detection& operator=(Detection &&detection) noexcept;
    ~Detection();
    void clear() noexcept;
    void setTemplateIdentifier(const std::pair<NetworkStationPhase, uint64_t> &id) noexcept;
    std::pair<NetworkStationPhase, uint64_t> getTemplateIdentifier() const;
    bool haveTemplateIdentifier() const noexcept;
    /// @throws std::invalid_argument unless -1 <= value <= 1.
    void setCorrelationCoefficient(double value);
    double getCorrelationCoefficient() const;
    bool haveCorrelationCoefficient() const noexcept;
    void setDetectedSignal(int n, const T x[]);
    void getDetectedSignal(int n, T *x[]) const;
    const T *getDetectedSignalPointer() const;
    int getDetectedSignalLength() const;
    bool haveDetectedSignal() const noexcept;
    /// Times in seconds from the signal's first sample; negative throws.
    void setDetectionTime(double time);
    double getDetectionTime() const;
    bool haveDetectionTime() const noexcept;
    void setInterpolatedDetectionTime(double time);
    double getInterpolatedDetectionTime() const;
    bool haveInterpolatedDetectionTime() const noexcept;
    void setTravelTime(double travelTime);
    double getTravelTime() const;
    bool haveTravelTime() const noexcept;
    void setPhaseOnsetTime(double onsetTime);
    double getPhaseOnsetTime() const;
    bool havePhaseOnsetTime() const noexcept;
    void setInterpolatedPhaseOnsetTime(double onsetTime);
    double getInterpolatedPhaseOnsetTime() const;
    bool haveInterpolatedPhaseOnsetTime() const noexcept;
    /// @throws std::invalid_argument unless value > 0.
    void setAmplitudeScalingFactor(T value, MFLib::RelativeMagnitudeType type);
    T getAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const;
    T getMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const;
    bool haveAmplitudeScalingFactor(MFLib::RelativeMagnitudeType type) const noexcept;
    bool haveMagnitudePerturbation(MFLib::RelativeMagnitudeType type) const noexcept;
    void setPolarity(MFLib::Polarity polarity) noexcept;
    MFLib::Polarity getPolarity() const noexcept;
private:
    class DetectionImpl;
    std::unique_ptr<DetectionImpl> pImpl;
};
}
}
#endif

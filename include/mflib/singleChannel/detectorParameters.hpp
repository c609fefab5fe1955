#ifndef MFLIB_SINGLECHANNEL_DETECTORPARAMETERS_HPP
#define MFLIB_SINGLECHANNEL_DETECTORPARAMETERS_HPP
#include <memory>
#include "mflib/enums.hpp"
#include "mflib/singleChannel/detectionTimeInterpolationParameters.hpp"
namespace MFLib
{
namespace SingleChannel
{
/// Settings of the single-channel detector
/// (include/mflib/singleChannel/detectorParameters.hpp in the reference).
/// Defaults: threshold 0.7, maxima of rho (not |rho|), one peak list per
/// template (SINGLE), no minimum spacing, amplitudes on, waveforms off.
class DetectorParameters
{
public:
    DetectorParameters();
    DetectorParameters(const DetectorParameters &parameters);
    DetectorParameters(DetectorParameters &&parameters) noexcept;
    DetectorParameters& operator=(const DetectorParameters &parms);
    DetectorParameters& operator=(DetectorParameters &&parms) noexcept;
    ~DetectorParameters();
    void clear() noexcept;
    void setMaximaPolicy(MaximumMatchedFilterPolicy policy) noexcept;
    MaximumMatchedFilterPolicy getMaximaPolicy() const noexcept;
    void setMinimumDetectionSpacing(int nSamples) noexcept;
    int getMinimumDetectionSpacing() const noexcept;
    /// @throws std::invalid_argument unless 0 < tol <= 1.
    void setDetectionThreshold(double tol);
    double getDetectionThreshold() const noexcept;
    void setMatchedFilteredSignalDetectorPolicy(MFLib::MatchedFilteredSignalDetectorPolicy policy) noexcept;
    MFLib::MatchedFilteredSignalDetectorPolicy getMatchedFilteredSignalDetectorPolicy() const noexcept;
    void setDetectionTimeInterpolationParameters(const DetectionTimeInterpolationParameters &interpParameters) noexcept;
    DetectionTimeInterpolationParameters& getDetectionTimeInterpolationParameters() const noexcept;
    void enableSaveDetectedWaveform() noexcept;
    void disableSaveDetectedWaveform() noexcept;
    bool wantDetectedWaveform() const noexcept;
    void enableSaveAmplitudeScalingFactor() noexcept;
    void disableSaveAmplitudeScalingFactor() noexcept;
    bool wantAmplitudeScalingFactor() const noexcept;
private:
    class DetectorParametersImpl;
    std::unique_ptr<DetectorParametersImpl> pImpl;
};
}
}
#endif

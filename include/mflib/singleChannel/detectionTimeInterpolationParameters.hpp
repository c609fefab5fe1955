#ifndef MFLIB_SINGLECHANNEL_DETECTIONTIMEINTERPOLATIONPARAMETERS_HPP
#define MFLIB_SINGLECHANNEL_DETECTIONTIMEINTERPOLATIONPARAMETERS_HPP
#include <memory>
#include "mflib/enums.hpp"
namespace MFLib
{
namespace SingleChannel
{
/// How a detection's time is refined below one sample
/// (include/mflib/singleChannel/detectionTimeInterpolationParameters.hpp in
/// the reference): a parabola through three lags (the default), or a
/// Lanczos-windowed sinc interpolant of window half-width alpha (default
/// 51) searched on a grid of npts (default 100) shifts in (-1/2, 1/2).
class DetectionTimeInterpolationParameters
{
public:
    DetectionTimeInterpolationParameters();
    DetectionTimeInterpolationParameters(const DetectionTimeInterpolationParameters &parameters);
    DetectionTimeInterpolationParameters(DetectionTimeInterpolationParameters &&parameters) noexcept;
    ~DetectionTimeInterpolationParameters();
    void clear() noexcept;
    DetectionTimeInterpolationParameters& operator=(const DetectionTimeInterpolationParameters &parameters);
    DetectionTimeInterpolationParameters& operator=(DetectionTimeInterpolationParameters &&parameters) noexcept;
    void enableLanczos() noexcept;
    /// @throws std::invalid_argument if alpha < 1.
    void setLanczosAlpha(int alpha);
    int getLanczosAlpha() const noexcept;
    /// @throws std::invalid_argument if npts < 1.
    void setLanczosNumberOfInterpolationPoints(int npts);
    int getLanczosNumberOfInterpolationPoints() const noexcept;
    void enableQuadratic() noexcept;
    void disable() noexcept;
    DetectionTimeInterpolationType getType() const noexcept;
private:
    class InterpParametersImpl;
    std::unique_ptr<InterpParametersImpl> pImpl;
};
}
}
#endif

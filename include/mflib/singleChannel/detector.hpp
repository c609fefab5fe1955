#ifndef MFLIB_SINGLECHANNEL_DETECTOR_HPP
#define MFLIB_SINGLECHANNEL_DETECTOR_HPP
#include <cstddef>
#include <memory>
#include <vector>
#include "mflib/singleChannel/detection.hpp"
namespace MFLib
{
namespace SingleChannel
{
template<class T> class MatchedFilter;
class DetectorParameters;
/// Detections from a single-channel matched filter
/// (include/mflib/singleChannel/detector.hpp in the reference).  detect()
/// runs on the GPU engine that holds the filter's correlations: the
/// reduce over templates (or one peak list per template), PeakFinder, and
/// one batch of per-detection products (clamped rho, refined time,
/// amplitude scaling factors); the Detection objects are assembled on the
/// host in the reference's order.
template<class T = double>
class Detector
{
public:
    Detector();
    ~Detector();
    void clear() noexcept;
    void initialize(const MFLib::SingleChannel::DetectorParameters &parameters);
    /// @throws std::invalid_argument if not initialized, the filter has no
    ///         correlations, no templates or a zero filtered length.
    void detect(const MFLib::SingleChannel::MatchedFilter<T> &mf);
    int getNumberOfDetections() const noexcept;
    bool isInitialized() const noexcept;
    const MFLib::SingleChannel::Detection<T>& operator[](size_t i) const;
    MFLib::SingleChannel::Detection<T> getDetection(int id) const;
private:
    class DetectorImpl;
    std::unique_ptr<DetectorImpl> pImpl;
};
}
}
#endif

#ifndef MFLIB_PEAKFINDER_HPP
#define MFLIB_PEAKFINDER_HPP
#include <memory>
namespace MFLib
{
/// Thresholded local maxima with greedy minimum-distance pruning
/// (include/private/peakFinder.hpp in the reference).  apply() runs the
/// threshold and local-maximum pass on the GPU.
template<class T = double>
class PeakFinder
{
public:
    PeakFinder();
    PeakFinder(const PeakFinder &peaks);
    PeakFinder(PeakFinder &&peaks) noexcept;
    PeakFinder& operator=(const PeakFinder &peaks);
    PeakFinder& operator=(PeakFinder &&peaks) noexcept;
    ~PeakFinder();
    void clear() noexcept;
    /// Peaks closer than this many samples to a larger kept peak are
    /// dropped; 0 or less keeps every candidate.
    void setMinimumPeakDistance(int minDistance) noexcept;
    int getMinimumPeakDistance() const noexcept;
    /// @throws std::invalid_argument unless threshold is in [0,1].
    void setThreshold(T threshhold);
    T getThreshold() const noexcept;
    void setSignal(const int n, const T x[]);
    void apply();
    int getNumberOfPeaks() const noexcept;
    void getPeakIndices(int maxSpace, int *peaksIndices[]) const;
    const int *getPeakIndicesPointer() const;
private:
    class PeakFinderImpl;
    std::unique_ptr<PeakFinderImpl> pImpl;
};
}
#endif

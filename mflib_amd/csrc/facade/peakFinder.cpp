// PeakFinder<T> (include/private/peakFinder.hpp).  Setter validation
// follows src/peakFinder.cpp:216-320 in the reference; apply() calls the
// GPU peak pass through the C ABI (mf_peaks), which keeps equal values in
// ascending index order.
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>
#include "abi.hpp"
#include "private/peakFinder.hpp"

using namespace MFLib;

namespace
{
constexpr double kDefaultThreshold = 0.7;  // src/peakFinder.cpp:24
}

template<class T>
class PeakFinder<T>::PeakFinderImpl
{
public:
    std::vector<T> x;
    std::vector<int> peaks;
    T threshold = static_cast<T>(kDefaultThreshold);
    int minDistance = 0;
    int nPeaks = 0;
};

template<class T>
PeakFinder<T>::PeakFinder() : pImpl(std::make_unique<PeakFinderImpl>()) {}
template<class T>
PeakFinder<T>::PeakFinder(const PeakFinder &o) : pImpl(std::make_unique<PeakFinderImpl>(*o.pImpl)) {}
template<class T>
PeakFinder<T>::PeakFinder(PeakFinder &&o) noexcept : pImpl(std::move(o.pImpl))
{
    o.pImpl = std::make_unique<PeakFinderImpl>();
}
template<class T>
PeakFinder<T>& PeakFinder<T>::operator=(const PeakFinder &o)
{
    if (&o != this) { *pImpl = *o.pImpl; }
    return *this;
}
template<class T>
PeakFinder<T>& PeakFinder<T>::operator=(PeakFinder &&o) noexcept
{
    if (&o != this) { std::swap(pImpl, o.pImpl); }
    return *this;
}
template<class T>
PeakFinder<T>::~PeakFinder() = default;
template<class T>
void PeakFinder<T>::clear() noexcept { *pImpl = PeakFinderImpl(); }

template<class T>
void PeakFinder<T>::setMinimumPeakDistance(int minDistance) noexcept
{
    pImpl->nPeaks = 0;
    pImpl->minDistance = minDistance;
}
template<class T>
int PeakFinder<T>::getMinimumPeakDistance() const noexcept { return pImpl->minDistance; }

template<class T>
void PeakFinder<T>::setThreshold(T threshold)
{
    pImpl->nPeaks = 0;
    if (!(threshold >= 0 && threshold <= 1))
    {
        throw std::invalid_argument("Threshold = " + std::to_string(threshold) + " must be in the range [0,1]");
    }
    pImpl->threshold = threshold;
}
template<class T>
T PeakFinder<T>::getThreshold() const noexcept { return pImpl->threshold; }

template<class T>
void PeakFinder<T>::setSignal(const int n, const T x[])
{
    pImpl->nPeaks = 0;
    if (n < 1) { throw std::invalid_argument("No points in x"); }
    if (x == nullptr) { throw std::invalid_argument("x is NULL"); }
    pImpl->x.assign(x, x + n);
}

template<class T>
void PeakFinder<T>::apply()
{
    auto &d = *pImpl;
    d.nPeaks = 0;
    const int n = static_cast<int>(d.x.size());
    if (n < 1) { return; }
    d.peaks.assign(n, 0);
    int count = 0;
    Facade::check(mf_peaks(d.x.data(), n, Facade::precisionOf<T>(), static_cast<double>(d.threshold),
                           d.minDistance, d.peaks.data(), n, &count), "PeakFinder::apply");
    d.nPeaks = count;
}

template<class T>
int PeakFinder<T>::getNumberOfPeaks() const noexcept { return pImpl->nPeaks; }

template<class T>
void PeakFinder<T>::getPeakIndices(int maxSpace, int *peaksIndices[]) const
{
    const int np = pImpl->nPeaks;
    if (maxSpace < np)
    {
        throw std::invalid_argument("maxx = " + std::to_string(maxSpace) + " must be at least = " +
                                    std::to_string(np));
    }
    if (peaksIndices == nullptr || *peaksIndices == nullptr) { throw std::invalid_argument("peaks is NULL"); }
    std::copy(pImpl->peaks.begin(), pImpl->peaks.begin() + np, *peaksIndices);
}

template<class T>
const int *PeakFinder<T>::getPeakIndicesPointer() const
{
    return pImpl->nPeaks < 1 ? nullptr : pImpl->peaks.data();
}

template class MFLib::PeakFinder<double>;
template class MFLib::PeakFinder<float>;

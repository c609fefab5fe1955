#ifndef MFLIB_SINGLECHANNEL_MATCHEDFILTER_HPP
#define MFLIB_SINGLECHANNEL_MATCHEDFILTER_HPP
#include <complex>
#include <memory>
#include <vector>
#include "mflib/waveformTemplate.hpp"
namespace MFLib
{
namespace Facade
{
struct Access;
}
namespace SingleChannel
{
class MatchedFilterParameters;
/// Many templates against one signal
/// (include/mflib/singleChannel/matchedFilter.hpp in the reference).  One
/// forward FFT per signal block serves every template.
template<class T = double>
class MatchedFilter
{
public:
    MatchedFilter();
    ~MatchedFilter();
    void clear() noexcept;
    void initialize(const MFLib::SingleChannel::MatchedFilterParameters &parameters);
    MFLib::WaveformTemplate getWaveformTemplate(int it) const;
    void setSignal(int nSamples, const T signal[]);
    const T* getSignalPointer() const;
    void zeroSignal();
    int getNumberOfTemplates() const;
    int getSignalLength() const;
    int getFilteredSignalLength() const;
    void apply();
    std::vector<T> getMatchedFilteredSignal(int it) const;
    const T* getMatchedFilterSignalPointer(int it) const;
    bool haveMatchedFilteredSignals() const noexcept;
    bool isInitialized() const noexcept;
private:
    friend struct MFLib::Facade::Access;  // the engine handle, for Detector
    MatchedFilter(const MatchedFilter &mf) = delete;
    MatchedFilter& operator=(const MatchedFilter &mf) = delete;
    class MatchedFilterImpl;
    std::unique_ptr<MatchedFilterImpl> pImpl;
};
}
}
#endif

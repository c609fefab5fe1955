#ifndef MFLIB_MATCHEDFILTER_HPP
#define MFLIB_MATCHEDFILTER_HPP
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
class MatchedFilterParameters;
/// Pearson-correlation matched filter of templates against signals
/// (include/mflib/matchedFilter.hpp in the reference), computed on the
/// GPU.  Results live in device memory; the host views returned by
/// getMatchedFilterSignalPointer are filled on first access after apply()
/// and stay valid until the next apply(), initialize() or clear().
template<class T = double>
class MatchedFilter
{
public:
    MatchedFilter();
    ~MatchedFilter();
    void clear() noexcept;
    void initialize(const MatchedFilterParameters &parameters);
    MFLib::WaveformTemplate getWaveformTemplate(int it) const;
    void setSignal(int it, int nSamples, const T signal[]);
    const T* getSignalPointer(int it) const;
    void zeroSignal(int it);
    int getNumberOfTemplates() const;
    int getSignalLength() const;
    int getFilteredSignalLength() const;
    void apply();
    std::vector<T> shiftAndStack();
    std::vector<T> getMatchedFilteredSignal(int it) const;
    const T* getMatchedFilterSignalPointer(int it) const;
    bool haveMatchedFilteredSignals() const noexcept;
    bool isInitialized() const noexcept;
    std::vector<std::complex<T>> getSpectraOfTemplate(int it) const;
private:
    friend struct MFLib::Facade::Access;  // the engine handle, for Detector
    MatchedFilter& operator=(const MatchedFilter &mf) = delete;
    MatchedFilter(const MatchedFilter &mf) = delete;
    class MatchedFilterImpl;
    std::unique_ptr<MatchedFilterImpl> pImpl;
};
}
#endif

// SingleChannel::MatchedFilter<T>: MFLib::MatchedFilter<T> in
// SINGLE_CHANNEL mode with the signal index fixed at 0
// (src/singleChannel/matchedFilter.cpp:40-141 in the reference).
#include "abi.hpp"
#include "mflib/matchedFilter.hpp"
#include "mflib/matchedFilterParameters.hpp"
#include "mflib/singleChannel/matchedFilter.hpp"
#include "mflib/singleChannel/matchedFilterParameters.hpp"

using namespace MFLib::SingleChannel;

template<class T>
class MatchedFilter<T>::MatchedFilterImpl
{
public:
    MFLib::MatchedFilter<T> mf;
};

template<class T>
MatchedFilter<T>::MatchedFilter() : pImpl(std::make_unique<MatchedFilterImpl>()) {}
template<class T>
MatchedFilter<T>::~MatchedFilter() = default;
template<class T>
void MatchedFilter<T>::clear() noexcept { pImpl->mf.clear(); }
template<class T>
void MatchedFilter<T>::initialize(const MatchedFilterParameters &parameters)
{
    pImpl->mf.initialize(parameters.getParameters());
}
template<class T>
MFLib::WaveformTemplate MatchedFilter<T>::getWaveformTemplate(int it) const { return pImpl->mf.getWaveformTemplate(it); }
template<class T>
void MatchedFilter<T>::setSignal(int nSamples, const T signal[]) { pImpl->mf.setSignal(0, nSamples, signal); }
template<class T>
const T* MatchedFilter<T>::getSignalPointer() const { return pImpl->mf.getSignalPointer(0); }
template<class T>
void MatchedFilter<T>::zeroSignal() { pImpl->mf.zeroSignal(0); }
template<class T>
int MatchedFilter<T>::getNumberOfTemplates() const { return pImpl->mf.getNumberOfTemplates(); }
template<class T>
int MatchedFilter<T>::getSignalLength() const { return pImpl->mf.getSignalLength(); }
template<class T>
int MatchedFilter<T>::getFilteredSignalLength() const { return pImpl->mf.getFilteredSignalLength(); }
template<class T>
void MatchedFilter<T>::apply() { pImpl->mf.apply(); }
template<class T>
std::vector<T> MatchedFilter<T>::getMatchedFilteredSignal(int it) const { return pImpl->mf.getMatchedFilteredSignal(it); }
template<class T>
const T* MatchedFilter<T>::getMatchedFilterSignalPointer(int it) const
{
    return pImpl->mf.getMatchedFilterSignalPointer(it);
}
template<class T>
bool MatchedFilter<T>::haveMatchedFilteredSignals() const noexcept { return pImpl->mf.haveMatchedFilteredSignals(); }
template<class T>
bool MatchedFilter<T>::isInitialized() const noexcept { return pImpl->mf.isInitialized(); }

template class MFLib::SingleChannel::MatchedFilter<double>;
template class MFLib::SingleChannel::MatchedFilter<float>;

template<class T>
const MFLib::MatchedFilter<T> &MFLib::Facade::Access::inner(const MFLib::SingleChannel::MatchedFilter<T> &mf)
{
    return mf.pImpl->mf;
}
template const MFLib::MatchedFilter<float> &MFLib::Facade::Access::inner<float>(
    const MFLib::SingleChannel::MatchedFilter<float> &);
template const MFLib::MatchedFilter<double> &MFLib::Facade::Access::inner<double>(
    const MFLib::SingleChannel::MatchedFilter<double> &);

template<class T>
MFLib::MatchedFilter<T> &MFLib::Facade::Access::inner(MFLib::SingleChannel::MatchedFilter<T> &mf)
{
    return mf.pImpl->mf;
}
template MFLib::MatchedFilter<float> &MFLib::Facade::Access::inner<float>(MFLib::SingleChannel::MatchedFilter<float> &);
template MFLib::MatchedFilter<double> &MFLib::Facade::Access::inner<double>(
    MFLib::SingleChannel::MatchedFilter<double> &);

// MatchedFilter<T> (include/mflib/matchedFilter.hpp) on the GPU engine.
// The pImpl holds one C-ABI handle plus host-side mirrors:
//  - the caller's signals, for getSignalPointer (the reference returns its
//    own copy, src/matchedFilter.cpp:923-938);
//  - correlation rows copied from HBM on first access after apply(), for
//    getMatchedFilterSignalPointer (940-969).
// Index and state checks follow src/matchedFilter.cpp:823-1402.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include "abi.hpp"
#include "mflib/matchedFilter.hpp"
#include "mflib/matchedFilterParameters.hpp"

using namespace MFLib;

template<class T>
class MatchedFilter<T>::MatchedFilterImpl
{
public:
    ~MatchedFilterImpl() { release(); }
    void release() noexcept
    {
        if (handle) { mf_destroy(handle); }
        handle = nullptr;
        signals.clear();
        signalOnDevice.clear();
        rho.clear();
        rhoReady.clear();
        nTemplates = 0;
        nSamples = 0;
        nFiltered = 0;
        singleChannel = false;
        applied = false;
    }
    void requireInitialized() const
    {
        if (!handle) { throw std::runtime_error("Class is not initialized"); }
    }
    void checkIndex(int it) const
    {
        requireInitialized();
        if (it < 0 || it >= nTemplates)
        {
            throw std::invalid_argument("it = " + std::to_string(it) + " must be in the range [0," +
                                        std::to_string(nTemplates - 1) + "]");
        }
    }

    mf_handle *handle = nullptr;
    MatchedFilterParameters parameters;
    std::vector<std::vector<T>> signals;   // per template row, host copy
    std::vector<uint8_t> signalOnDevice;   // row set from device memory: host copy fetched on demand
    std::vector<std::vector<T>> rho;       // per template row, host mirror
    std::vector<uint8_t> rhoReady;
    int nTemplates = 0;
    int nSamples = 0;
    int nFiltered = 0;
    bool singleChannel = false;
    bool applied = false;
};

template<class T>
MatchedFilter<T>::MatchedFilter() : pImpl(std::make_unique<MatchedFilterImpl>()) {}

template<class T>
MatchedFilter<T>::~MatchedFilter() = default;

template<class T>
void MatchedFilter<T>::clear() noexcept
{
    pImpl->release();
    pImpl->parameters.clear();
}

template<class T>
void MatchedFilter<T>::initialize(const MatchedFilterParameters &parms)
{
    clear();
    if (!parms.isValid()) { throw std::invalid_argument("Parameters are not valid"); }
    auto &d = *pImpl;
    const int nt = parms.getNumberOfTemplates();
    const int n = parms.getSignalSize();
    const bool single = parms.getDetectionMode() == MatchedFilterDetectionMode::SINGLE_CHANNEL;

    // Template samples and the shift-and-stack table
    // (src/matchedFilter.cpp:560-593): stacking needs onset and travel
    // time on every template; shift = int((travel - onset)*fs + 0.5).
    std::vector<WaveformTemplate> tmpl;
    tmpl.reserve(nt);
    std::vector<int> lengths(nt);
    std::vector<const double *> samples(nt);
    std::vector<std::vector<double>> buffers(nt);
    std::vector<int> shifts(nt, 0);
    std::vector<double> weights(nt, 1.0);
    bool canStack = true;
    for (int it = 0; it < nt; ++it)
    {
        tmpl.push_back(parms.getTemplate(it));
        const auto &t = tmpl.back();
        lengths[it] = t.getSignalLength();
        buffers[it].resize(lengths[it]);
        double *p = buffers[it].data();
        t.getSignal(lengths[it], &p);
        samples[it] = p;
        if (canStack && t.havePhaseOnsetTime() && t.haveTravelTime())
        {
            const double tdiff = t.getTravelTime() - t.getPhaseOnsetTime();
            shifts[it] = static_cast<int>(tdiff * t.getSamplingRate() + 0.5);
            weights[it] = t.getShiftAndStackWeight();
        }
        else
        {
            canStack = false;
        }
    }

    mf_handle *h = nullptr;
    Facade::check(mf_create(&h, Facade::precisionOf<T>(),
                            single ? MF_MODE_SINGLE_CHANNEL : MF_MODE_MULTI_CHANNEL), "initialize");
    d.handle = h;
    try
    {
        Facade::check(mf_set_signal_size(h, n), "initialize");
        Facade::check(mf_set_templates(h, nt, lengths.data(), samples.data(),
                                       canStack ? shifts.data() : nullptr, weights.data()), "initialize");
        Facade::check(mf_set_stack_absolute_values(h, parms.getStackAbsoluteValues() ? 1 : 0), "initialize");
        Facade::check(mf_initialize(h), "initialize");
        Facade::check(mf_get_filtered_length(h, 0, &d.nFiltered), "initialize");
    }
    catch (...)
    {
        d.release();
        throw;
    }
    d.parameters = parms;
    d.nTemplates = nt;
    d.nSamples = n;
    d.singleChannel = single;
    d.signals.assign(nt, std::vector<T>());  // rows filled on first use
    d.signalOnDevice.assign(nt, 0);
    d.rho.assign(nt, std::vector<T>());
    d.rhoReady.assign(nt, 0);
}

template<class T>
WaveformTemplate MatchedFilter<T>::getWaveformTemplate(int it) const
{
    pImpl->checkIndex(it);
    return pImpl->parameters.getTemplate(it);
}

template<class T>
void MatchedFilter<T>::setSignal(int it, int nSamples, const T signal[])
{
    auto &d = *pImpl;
    d.checkIndex(it);
    if (nSamples != d.nSamples)
    {
        throw std::invalid_argument("Number of samples = " + std::to_string(nSamples) + " must equal " +
                                    std::to_string(d.nSamples));
    }
    if (signal == nullptr) { throw std::invalid_argument("signal is NULL"); }
    // In single-channel mode every template reads signal 0: the reference
    // keeps a copy per row but transforms row 0 only (1176-1285).
    if (!d.singleChannel || it == 0)
    {
        Facade::check(mf_set_signal(d.handle, it, nSamples, signal), "setSignal");
    }
    d.signals[it].assign(signal, signal + nSamples);
    d.signalOnDevice[it] = 0;
}

template<class T>
const T* MatchedFilter<T>::getSignalPointer(int it) const
{
    pImpl->checkIndex(it);
    auto &d = *pImpl;
    auto &row = d.signals[it];
    if (d.signalOnDevice[it])
    {
        row.assign(d.nSamples, T(0));
        Facade::check(mf_get_signal(d.handle, d.singleChannel ? 0 : it, row.data(), d.nSamples), "getSignalPointer");
        d.signalOnDevice[it] = 0;
    }
    if (row.empty()) { row.assign(d.nSamples, T(0)); }
    return row.data();
}

template<class T>
void MatchedFilter<T>::zeroSignal(int it)
{
    auto &d = *pImpl;
    d.checkIndex(it);
    if (!d.singleChannel || it == 0) { Facade::check(mf_zero_signal(d.handle, it), "zeroSignal"); }
    d.signals[it].assign(d.nSamples, T(0));
    d.signalOnDevice[it] = 0;
}

template<class T>
int MatchedFilter<T>::getNumberOfTemplates() const
{
    pImpl->requireInitialized();
    return pImpl->nTemplates;
}

template<class T>
int MatchedFilter<T>::getSignalLength() const
{
    pImpl->requireInitialized();
    return pImpl->nSamples;
}

template<class T>
int MatchedFilter<T>::getFilteredSignalLength() const
{
    pImpl->requireInitialized();
    return pImpl->nFiltered;
}

template<class T>
void MatchedFilter<T>::apply()
{
    auto &d = *pImpl;
    d.requireInitialized();
    d.applied = false;
    std::fill(d.rhoReady.begin(), d.rhoReady.end(), 0);
    Facade::check(mf_apply(d.handle), "apply");
    d.applied = true;
}

template<class T>
std::vector<T> MatchedFilter<T>::shiftAndStack()
{
    auto &d = *pImpl;
    d.requireInitialized();
    if (!d.applied) { throw std::runtime_error("Matched filtered signals not yet computed"); }
    int len = 0;
    std::vector<T> out(std::max(d.nFiltered, 1));
    Facade::check(mf_shift_and_stack(d.handle, 0, out.data(), static_cast<int>(out.size()), &len),
                  "shiftAndStack");
    out.resize(len);
    return out;
}

template<class T>
const T* MatchedFilter<T>::getMatchedFilterSignalPointer(int it) const
{
    auto &d = *pImpl;
    d.checkIndex(it);
    auto &row = d.rho[it];
    if (!d.rhoReady[it])
    {
        row.assign(d.nFiltered, T(0));
        // before apply() the reference's rows are zero (calloc'd, 661-670)
        if (d.applied)
        {
            Facade::check(mf_get_rho(d.handle, it, row.data(), d.nFiltered), "getMatchedFilterSignalPointer");
        }
        d.rhoReady[it] = d.applied ? 1 : 0;
    }
    return row.data();
}

template<class T>
std::vector<T> MatchedFilter<T>::getMatchedFilteredSignal(int it) const
{
    const T *p = getMatchedFilterSignalPointer(it);
    return std::vector<T>(p, p + pImpl->nFiltered);
}

template<class T>
bool MatchedFilter<T>::haveMatchedFilteredSignals() const noexcept
{
    return pImpl->handle != nullptr && pImpl->applied;
}

template<class T>
bool MatchedFilter<T>::isInitialized() const noexcept
{
    return pImpl->handle != nullptr;
}

template<class T>
std::vector<std::complex<T>> MatchedFilter<T>::getSpectraOfTemplate(int it) const
{
    auto &d = *pImpl;
    d.checkIndex(it);
    // size query first: the exported length is the engine's, not always
    // the row's plan (time-domain rows have no FFT plan)
    int len = 0;
    const int rc = mf_template_spectrum(d.handle, it, nullptr, 0, &len);
    if (len < 1) { Facade::check(rc, "getSpectraOfTemplate"); }
    std::vector<std::complex<T>> out(len);
    Facade::check(mf_template_spectrum(d.handle, it, out.data(), static_cast<int>(out.size()), &len),
                  "getSpectraOfTemplate");
    out.resize(len);
    return out;
}

template class MFLib::MatchedFilter<double>;
template class MFLib::MatchedFilter<float>;

template<class T>
mf_handle *MFLib::Facade::Access::handle(const MFLib::MatchedFilter<T> &mf)
{
    return mf.pImpl->handle;
}
template mf_handle *MFLib::Facade::Access::handle<float>(const MFLib::MatchedFilter<float> &);
template mf_handle *MFLib::Facade::Access::handle<double>(const MFLib::MatchedFilter<double> &);

template<class T>
void MFLib::Facade::Access::setSignalDevice(MFLib::MatchedFilter<T> &mf, int it, int nSamples, const T *dx)
{
    auto &d = *mf.pImpl;
    d.checkIndex(it);
    if (nSamples != d.nSamples)
    {
        throw std::invalid_argument("Number of samples = " + std::to_string(nSamples) + " must equal " +
                                    std::to_string(d.nSamples));
    }
    if (dx == nullptr) { throw std::invalid_argument("signal is NULL"); }
    if (!d.singleChannel || it == 0)
    {
        Facade::check(mf_set_signal_device(d.handle, d.singleChannel ? 0 : it, nSamples, dx), "setSignal");
    }
    d.signals[it].clear();
    d.signalOnDevice[it] = 1;
}

template<class T>
const T *MFLib::Facade::Access::rhoDevice(const MFLib::MatchedFilter<T> &mf, int it)
{
    auto &d = *mf.pImpl;
    d.checkIndex(it);
    if (!d.applied) { throw std::runtime_error("Matched filtered signals not yet computed"); }
    const void *p = nullptr;
    Facade::check(mf_get_rho_device(d.handle, it, &p), "getMatchedFilterSignalPointer");
    Facade::check(mf_synchronize(d.handle), "getMatchedFilterSignalPointer");
    return static_cast<const T *>(p);
}

template<class T>
int MFLib::Facade::Access::device(const MFLib::MatchedFilter<T> &mf)
{
    mf.pImpl->requireInitialized();
    int dev = 0;
    Facade::check(mf_get_device(mf.pImpl->handle, &dev), "device");
    return dev;
}

template void MFLib::Facade::Access::setSignalDevice<float>(MFLib::MatchedFilter<float> &, int, int, const float *);
template void MFLib::Facade::Access::setSignalDevice<double>(MFLib::MatchedFilter<double> &, int, int,
                                                             const double *);
template const float *MFLib::Facade::Access::rhoDevice<float>(const MFLib::MatchedFilter<float> &, int);
template const double *MFLib::Facade::Access::rhoDevice<double>(const MFLib::MatchedFilter<double> &, int);
template int MFLib::Facade::Access::device<float>(const MFLib::MatchedFilter<float> &);
template int MFLib::Facade::Access::device<double>(const MFLib::MatchedFilter<double> &);

// MatchedFilterParameters (include/mflib/matchedFilterParameters.hpp).
// Validation and defaults follow src/matchedFilterParameters.cpp:105-385:
//  - adding a template or setting the signal size forgets a chosen FFT
//    length (199-203, 318-321);
//  - sampling rates must agree within 1e-4 (209-216);
//  - the default FFT/block length comes from the window-length rule with
//    nx = signal size (250-283).
// Divergence: isValid() returns false instead of letting getSignalSize()
// throw out of a noexcept function (363-369).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>
#include "abi.hpp"
#include "mflib/matchedFilterParameters.hpp"

using namespace MFLib;

class MatchedFilterParameters::MatchedFilterParametersImpl
{
public:
    std::vector<WaveformTemplate> templates;
    MatchedFilterImplementation implementation = MatchedFilterImplementation::DIRECT;
    MatchedFilterDetectionMode mode = MatchedFilterDetectionMode::MULTI_CHANNEL;
    double samplingRate = -1;
    int signalSize = 0;
    int fftLength = 0;
    int blockLength = 0;
    bool stackAbs = true;
};

MatchedFilterParameters::MatchedFilterParameters() : pImpl(std::make_unique<MatchedFilterParametersImpl>()) {}
MatchedFilterParameters::MatchedFilterParameters(const MatchedFilterParameters &p) :
    pImpl(std::make_unique<MatchedFilterParametersImpl>(*p.pImpl)) {}
MatchedFilterParameters::MatchedFilterParameters(MatchedFilterParameters &&p) noexcept : pImpl(std::move(p.pImpl))
{
    p.pImpl = std::make_unique<MatchedFilterParametersImpl>();
}
MatchedFilterParameters& MatchedFilterParameters::operator=(const MatchedFilterParameters &p)
{
    if (&p != this) { *pImpl = *p.pImpl; }
    return *this;
}
MatchedFilterParameters& MatchedFilterParameters::operator=(MatchedFilterParameters &&p) noexcept
{
    if (&p != this) { std::swap(pImpl, p.pImpl); }
    return *this;
}
MatchedFilterParameters::~MatchedFilterParameters() = default;

void MatchedFilterParameters::clear() noexcept
{
    // the detection mode survives clear(), as in the reference (174-183)
    const auto mode = pImpl->mode;
    *pImpl = MatchedFilterParametersImpl();
    pImpl->mode = mode;
}

void MatchedFilterParameters::addTemplate(const WaveformTemplate &t)
{
    pImpl->fftLength = 0;
    pImpl->blockLength = 0;
    if (!t.haveSignal()) { throw std::invalid_argument("Waveform not yet set on template"); }
    const double fs = t.getSamplingRate();
    if (pImpl->samplingRate < 0) { pImpl->samplingRate = fs; }
    if (std::abs(fs - pImpl->samplingRate) > 1.e-4)
    {
        throw std::invalid_argument("Template sampling rate = " + std::to_string(fs) + " must equal " +
                                    std::to_string(pImpl->samplingRate));
    }
    pImpl->templates.push_back(t);
}

void MatchedFilterParameters::clearTemplates() noexcept
{
    pImpl->samplingRate = -1;
    pImpl->templates.clear();
}

WaveformTemplate MatchedFilterParameters::getTemplate(int it) const
{
    const int nt = getNumberOfTemplates();
    if (it < 0 || it >= nt)
    {
        throw std::invalid_argument("Template " + std::to_string(it) + " must be in range [0," +
                                    std::to_string(nt) + "]");
    }
    return pImpl->templates[it];
}

int MatchedFilterParameters::getNumberOfTemplates() const noexcept
{
    return static_cast<int>(pImpl->templates.size());
}

int MatchedFilterParameters::getMaxTemplateLength() const
{
    if (pImpl->templates.empty()) { throw std::runtime_error("No templates set"); }
    int len = 0;
    for (const auto &t : pImpl->templates) { len = std::max(len, t.getSignalLength()); }
    return len;
}

void MatchedFilterParameters::setSignalSize(int nx)
{
    pImpl->fftLength = 0;
    pImpl->blockLength = 0;
    if (nx < 1) { throw std::invalid_argument("Signal size must be positive"); }
    pImpl->signalSize = nx;
}

int MatchedFilterParameters::getSignalSize() const
{
    if (pImpl->signalSize < 1) { throw std::runtime_error("Signal size not set"); }
    return pImpl->signalSize;
}

void MatchedFilterParameters::setFFTLength(int fftLength)
{
    const int nb = getMaxTemplateLength();
    if (fftLength < nb)
    {
        throw std::invalid_argument("fftLength = " + std::to_string(fftLength) + " must be at least " +
                                    std::to_string(nb));
    }
    pImpl->fftLength = fftLength;
    pImpl->blockLength = fftLength - nb + 1;
}

static void defaultLengths(const MatchedFilterParameters &p, int &nfft, int &L)
{
    const int nb = p.getMaxTemplateLength();
    const int nx = p.getSignalSize();
    Facade::check(mf_compute_optimal_fft_and_block_length(nb, nx, &nfft, &L), "getFFTLength");
}

int MatchedFilterParameters::getFFTLength() const
{
    getMaxTemplateLength();
    if (pImpl->fftLength < 1) { defaultLengths(*this, pImpl->fftLength, pImpl->blockLength); }
    return pImpl->fftLength;
}

int MatchedFilterParameters::getBlockLength() const
{
    getMaxTemplateLength();
    if (pImpl->fftLength < 1) { defaultLengths(*this, pImpl->fftLength, pImpl->blockLength); }
    return pImpl->blockLength;
}

void MatchedFilterParameters::setDetectionMode(const MatchedFilterDetectionMode mode) noexcept { pImpl->mode = mode; }
MatchedFilterDetectionMode MatchedFilterParameters::getDetectionMode() const noexcept { return pImpl->mode; }
void MatchedFilterParameters::setMatchedFilterImplementation(const MatchedFilterImplementation i) noexcept
{
    pImpl->implementation = i;
}
MatchedFilterImplementation MatchedFilterParameters::getMatchedFilterImplementation() const noexcept
{
    return pImpl->implementation;
}
void MatchedFilterParameters::setStackAbsoluteValues(bool labs) noexcept { pImpl->stackAbs = labs; }
bool MatchedFilterParameters::getStackAbsoluteValues() const noexcept { return pImpl->stackAbs; }

bool MatchedFilterParameters::isValid() const noexcept
{
    return !pImpl->templates.empty() && pImpl->signalSize > 0;
}

std::pair<int, int> MFLib::computeOptimalFFTAndBlockLength(int nb, int nx)
{
    int nfft = 0, L = 0;
    Facade::check(mf_compute_optimal_fft_and_block_length(nb, nx, &nfft, &L), "computeOptimalFFTAndBlockLength");
    return {nfft, L};
}

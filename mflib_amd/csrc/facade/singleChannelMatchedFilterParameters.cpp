// SingleChannel::MatchedFilterParameters: MFLib's parameters with the
// detection mode held at SINGLE_CHANNEL
// (src/singleChannel/matchedFilterParameters.cpp:20-159 in the reference).
#include "mflib/singleChannel/matchedFilterParameters.hpp"

using namespace MFLib::SingleChannel;

class MatchedFilterParameters::MatchedFilterParametersImpl
{
public:
    MatchedFilterParametersImpl() { p.setDetectionMode(MFLib::MatchedFilterDetectionMode::SINGLE_CHANNEL); }
    MFLib::MatchedFilterParameters p;
};

MatchedFilterParameters::MatchedFilterParameters() : pImpl(std::make_unique<MatchedFilterParametersImpl>()) {}
MatchedFilterParameters::MatchedFilterParameters(const MatchedFilterParameters &o) :
    pImpl(std::make_unique<MatchedFilterParametersImpl>(*o.pImpl)) {}
MatchedFilterParameters::MatchedFilterParameters(MatchedFilterParameters &&o) noexcept : pImpl(std::move(o.pImpl))
{
    o.pImpl = std::make_unique<MatchedFilterParametersImpl>();
}
MatchedFilterParameters& MatchedFilterParameters::operator=(const MatchedFilterParameters &o)
{
    if (&o != this) { *pImpl = *o.pImpl; }
    return *this;
}
MatchedFilterParameters& MatchedFilterParameters::operator=(MatchedFilterParameters &&o) noexcept
{
    if (&o != this) { std::swap(pImpl, o.pImpl); }
    return *this;
}
MatchedFilterParameters::~MatchedFilterParameters() = default;
void MatchedFilterParameters::clear() noexcept
{
    pImpl->p.clear();
    pImpl->p.setDetectionMode(MFLib::MatchedFilterDetectionMode::SINGLE_CHANNEL);
}
void MatchedFilterParameters::addTemplate(const MFLib::WaveformTemplate &t) { pImpl->p.addTemplate(t); }
void MatchedFilterParameters::clearTemplates() noexcept { pImpl->p.clearTemplates(); }
MFLib::WaveformTemplate MatchedFilterParameters::getTemplate(int it) const { return pImpl->p.getTemplate(it); }
int MatchedFilterParameters::getNumberOfTemplates() const noexcept { return pImpl->p.getNumberOfTemplates(); }
int MatchedFilterParameters::getMaxTemplateLength() const { return pImpl->p.getMaxTemplateLength(); }
void MatchedFilterParameters::setSignalSize(int nx) { pImpl->p.setSignalSize(nx); }
int MatchedFilterParameters::getSignalSize() const { return pImpl->p.getSignalSize(); }
void MatchedFilterParameters::setFFTLength(int n) { pImpl->p.setFFTLength(n); }
int MatchedFilterParameters::getFFTLength() const { return pImpl->p.getFFTLength(); }
int MatchedFilterParameters::getBlockLength() const { return pImpl->p.getBlockLength(); }
MFLib::MatchedFilterDetectionMode MatchedFilterParameters::getDetectionMode() const noexcept
{
    return pImpl->p.getDetectionMode();
}
void MatchedFilterParameters::setMatchedFilterImplementation(const MFLib::MatchedFilterImplementation i) noexcept
{
    pImpl->p.setMatchedFilterImplementation(i);
}
MFLib::MatchedFilterImplementation MatchedFilterParameters::getMatchedFilterImplementation() const noexcept
{
    return pImpl->p.getMatchedFilterImplementation();
}
bool MatchedFilterParameters::isValid() const noexcept { return pImpl->p.isValid(); }
MFLib::MatchedFilterParameters MatchedFilterParameters::getParameters() const { return pImpl->p; }

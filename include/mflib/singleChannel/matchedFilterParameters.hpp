#ifndef MFLIB_SINGLECHANNEL_FILTERPARAMETERS_HPP
#define MFLIB_SINGLECHANNEL_FILTERPARAMETERS_HPP
#include <memory>
#include <vector>
#include "mflib/matchedFilterParameters.hpp"
#include "mflib/waveformTemplate.hpp"
#include "mflib/enums.hpp"
namespace MFLib
{
namespace SingleChannel
{
/// Parameters of a single-channel matched filter: MFLib's parameters with
/// the detection mode fixed to SINGLE_CHANNEL
/// (include/mflib/singleChannel/matchedFilterParameters.hpp in the reference).
class MatchedFilterParameters
{
public:
    MatchedFilterParameters();
    MatchedFilterParameters(const MatchedFilterParameters &parameters);
    MatchedFilterParameters(MatchedFilterParameters &&parameters) noexcept;
    MatchedFilterParameters& operator=(const MatchedFilterParameters &parameters);
    MatchedFilterParameters& operator=(MatchedFilterParameters &&parameters) noexcept;
    ~MatchedFilterParameters();
    void clear() noexcept;
    void addTemplate(const MFLib::WaveformTemplate &waveformTemplate);
    void clearTemplates() noexcept;
    WaveformTemplate getTemplate(int it) const;
    int getNumberOfTemplates() const noexcept;
    int getMaxTemplateLength() const;
    void setSignalSize(int nx);
    int getSignalSize() const;
    void setFFTLength(int fftLength);
    int getFFTLength() const;
    int getBlockLength() const;
    MFLib::MatchedFilterDetectionMode getDetectionMode() const noexcept;
    void setMatchedFilterImplementation(const MatchedFilterImplementation implementation) noexcept;
    MatchedFilterImplementation getMatchedFilterImplementation() const noexcept;
    bool isValid() const noexcept;
    MFLib::MatchedFilterParameters getParameters() const;
private:
    class MatchedFilterParametersImpl;
    std::unique_ptr<MatchedFilterParametersImpl> pImpl;
};
}
}
#endif

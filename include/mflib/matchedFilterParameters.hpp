#ifndef MFLIB_MATCHEDFILTERPARAMETERS_HPP
#define MFLIB_MATCHEDFILTERPARAMETERS_HPP
#include <memory>
#include <utility>
#include <vector>
#include "mflib/waveformTemplate.hpp"
#include "mflib/enums.hpp"
namespace MFLib
{
class WaveformTemplate;
/// Templates, signal size and options of a (multi-channel) matched filter
/// (include/mflib/matchedFilterParameters.hpp in the reference).
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
    /// @throws std::invalid_argument if the template has no signal or its
    ///         sampling rate differs from the first template's by > 1e-4;
    ///         std::runtime_error if it has no sampling rate.
    void addTemplate(const WaveformTemplate &waveformTemplate);
    void clearTemplates() noexcept;
    WaveformTemplate getTemplate(int it) const;
    int getNumberOfTemplates() const noexcept;
    int getMaxTemplateLength() const;
    void setSignalSize(int nx);
    int getSignalSize() const;
    /// FFT length hint.  As in the reference, MatchedFilter::initialize
    /// chooses its own FFT length; this value is reported back only.
    void setFFTLength(int fftLength);
    int getFFTLength() const;
    int getBlockLength() const;
    void setDetectionMode(const MatchedFilterDetectionMode mode) noexcept;
    MatchedFilterDetectionMode getDetectionMode() const noexcept;
    void setMatchedFilterImplementation(const MatchedFilterImplementation implementation) noexcept;
    MatchedFilterImplementation getMatchedFilterImplementation() const noexcept;
    /// Stack |rho| (true, the default) or rho in shiftAndStack.
    void setStackAbsoluteValues(bool labs) noexcept;
    bool getStackAbsoluteValues() const noexcept;
    /// True when there is at least one template and a signal size.
    bool isValid() const noexcept;
private:
    class MatchedFilterParametersImpl;
    std::unique_ptr<MatchedFilterParametersImpl> pImpl;
};
/// Overlap-add FFT length and block length for filter length nb and
/// signal length nx.  @throws std::invalid_argument if nb < 1 or nx <= nb.
std::pair<int, int> computeOptimalFFTAndBlockLength(int nb, int nx);
}
#endif

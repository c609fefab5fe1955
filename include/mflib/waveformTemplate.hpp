#ifndef MFLIB_WAVEFORMTEMPLATE_HPP
#define MFLIB_WAVEFORMTEMPLATE_HPP
#include <cstdint>
#include <memory>
#include <utility>
#include "mflib/enums.hpp"
namespace MFLib
{
class NetworkStationPhase;
/// One template waveform and the metadata the matched filter reads
/// (include/mflib/waveformTemplate.hpp in the reference).  The signal is
/// stored in FP64 whatever precision it is given in.
class WaveformTemplate
{
public:
    WaveformTemplate();
    WaveformTemplate(const WaveformTemplate &tplate);
    WaveformTemplate(WaveformTemplate &&tplate) noexcept;
    WaveformTemplate& operator=(const WaveformTemplate &tplate);
    WaveformTemplate& operator=(WaveformTemplate &&tplate) noexcept;
    ~WaveformTemplate();
    void clear() noexcept;
    /// @throws std::invalid_argument if npts < 1 or x is NULL.
    void setSignal(int npts, const double x[]);
    void setSignal(int npts, const float x[]);
    /// @throws std::runtime_error if no signal; std::invalid_argument if
    ///         maxx < getSignalLength() or *x is NULL.
    void getSignal(int maxx, double *x[]) const;
    void getSignal(int maxx, float *x[]) const;
    int getSignalLength() const;
    bool haveSignal() const noexcept;
    /// @throws std::invalid_argument if samplingRate <= 0.
    void setSamplingRate(double samplingRate);
    double getSamplingRate() const;
    bool haveSamplingRate() const noexcept;
    /// @throws std::invalid_argument unless weight is in [0,1].
    void setShiftAndStackWeight(double weight);
    double getShiftAndStackWeight() const noexcept;
    /// Onset time of the phase in the template, in seconds.
    void setPhaseOnsetTime(double onsetTime);
    double getPhaseOnsetTime() const;
    bool havePhaseOnsetTime() const noexcept;
    /// Travel time of the phase from the source, in seconds.
    void setTravelTime(double travelTime);
    double getTravelTime() const;
    bool haveTravelTime() const noexcept;
    void setIdentifier(const std::pair<NetworkStationPhase, uint64_t> &id) noexcept;
    std::pair<NetworkStationPhase, uint64_t> getIdentifier() const;
    bool haveIdentifier() const noexcept;
    void setMagnitude(double magnitude) noexcept;
    double getMagnitude() const;
    bool haveMagnitude() const noexcept;
    void setPolarity(MFLib::Polarity polarity) noexcept;
    MFLib::Polarity getPolarity() const noexcept;
private:
    class WaveformTemplateImpl;
    std::unique_ptr<WaveformTemplateImpl> pImpl;
};
}
#endif

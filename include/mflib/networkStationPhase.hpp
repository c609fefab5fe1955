#ifndef MFLIB_NETWORKSTATIONPHASE_HPP
#define MFLIB_NETWORKSTATIONPHASE_HPP
#include <cstddef>
#include <memory>
#include <string>
namespace MFLib
{
/// Network, station, channel, location code and phase of a template
/// (include/mflib/networkStationPhase.hpp in the reference).
class NetworkStationPhase
{
public:
    NetworkStationPhase();
    NetworkStationPhase(const NetworkStationPhase &nsp);
    NetworkStationPhase(NetworkStationPhase &&nsp) noexcept;
    NetworkStationPhase& operator=(const NetworkStationPhase &nsp);
    NetworkStationPhase& operator=(NetworkStationPhase &&nsp) noexcept;
    ~NetworkStationPhase();
    void clear() noexcept;
    void setNetwork(const std::string &network) noexcept;
    std::string getNetwork() const noexcept;
    void setStation(const std::string &station) noexcept;
    std::string getStation() const noexcept;
    void setChannel(const std::string &channel) noexcept;
    std::string getChannel() const noexcept;
    void setLocationCode(const std::string &locationCode) noexcept;
    std::string getLocationCode() const noexcept;
    void setPhase(const std::string &phase) noexcept;
    std::string getPhase() const noexcept;
    /// Hash of the five fields.
    size_t getHash() const noexcept;
private:
    class NetworkStationPhaseImpl;
    std::unique_ptr<NetworkStationPhaseImpl> pImpl;
};
bool operator==(const NetworkStationPhase &lhs, const NetworkStationPhase &rhs);
bool operator!=(const NetworkStationPhase &lhs, const NetworkStationPhase &rhs);
}
#endif

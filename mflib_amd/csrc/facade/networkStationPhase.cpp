// NetworkStationPhase (include/mflib/networkStationPhase.hpp; reference
// behaviour in src/networkStationPhase.cpp).  Equality compares the five
// fields exactly.
#include <functional>
#include <string>
#include "mflib/networkStationPhase.hpp"

using namespace MFLib;

class NetworkStationPhase::NetworkStationPhaseImpl
{
public:
    std::string network, station, channel, location, phase;
};

NetworkStationPhase::NetworkStationPhase() : pImpl(std::make_unique<NetworkStationPhaseImpl>()) {}
NetworkStationPhase::NetworkStationPhase(const NetworkStationPhase &nsp) :
    pImpl(std::make_unique<NetworkStationPhaseImpl>(*nsp.pImpl)) {}
NetworkStationPhase::NetworkStationPhase(NetworkStationPhase &&nsp) noexcept :
    pImpl(std::move(nsp.pImpl))
{
    nsp.pImpl = std::make_unique<NetworkStationPhaseImpl>();
}
NetworkStationPhase& NetworkStationPhase::operator=(const NetworkStationPhase &nsp)
{
    if (&nsp != this) { *pImpl = *nsp.pImpl; }
    return *this;
}
NetworkStationPhase& NetworkStationPhase::operator=(NetworkStationPhase &&nsp) noexcept
{
    if (&nsp != this) { std::swap(pImpl, nsp.pImpl); }
    return *this;
}
NetworkStationPhase::~NetworkStationPhase() = default;
void NetworkStationPhase::clear() noexcept { *pImpl = NetworkStationPhaseImpl(); }
void NetworkStationPhase::setNetwork(const std::string &s) noexcept { pImpl->network = s; }
std::string NetworkStationPhase::getNetwork() const noexcept { return pImpl->network; }
void NetworkStationPhase::setStation(const std::string &s) noexcept { pImpl->station = s; }
std::string NetworkStationPhase::getStation() const noexcept { return pImpl->station; }
void NetworkStationPhase::setChannel(const std::string &s) noexcept { pImpl->channel = s; }
std::string NetworkStationPhase::getChannel() const noexcept { return pImpl->channel; }
void NetworkStationPhase::setLocationCode(const std::string &s) noexcept { pImpl->location = s; }
std::string NetworkStationPhase::getLocationCode() const noexcept { return pImpl->location; }
void NetworkStationPhase::setPhase(const std::string &s) noexcept { pImpl->phase = s; }
std::string NetworkStationPhase::getPhase() const noexcept { return pImpl->phase; }

size_t NetworkStationPhase::getHash() const noexcept
{
    // fields joined with a separator that cannot occur in SEED codes
    const std::string key = pImpl->network + '\x1f' + pImpl->station + '\x1f' + pImpl->channel +
                            '\x1f' + pImpl->location + '\x1f' + pImpl->phase;
    return std::hash<std::string>{}(key);
}

bool MFLib::operator==(const NetworkStationPhase &a, const NetworkStationPhase &b)
{
    return a.getNetwork() == b.getNetwork() && a.getStation() == b.getStation() &&
           a.getChannel() == b.getChannel() && a.getLocationCode() == b.getLocationCode() &&
           a.getPhase() == b.getPhase();
}
bool MFLib::operator!=(const NetworkStationPhase &a, const NetworkStationPhase &b) { return !(a == b); }

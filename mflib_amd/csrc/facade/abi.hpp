// Shared by the C++ facade: turns C-ABI status codes into the exception
// types MFLib's API throws (src/matchedFilter.cpp:513-516, 829-834,
// 876-881, 977-980, 1322-1329 in the reference).
#ifndef MFLIB_FACADE_ABI_HPP
#define MFLIB_FACADE_ABI_HPP
#include <new>
#include <stdexcept>
#include <string>
#include "mflib_gpu.h"

namespace MFLib
{
template<class T> class MatchedFilter;
namespace SingleChannel
{
template<class T> class MatchedFilter;
}
}

namespace MFLib
{
namespace Facade
{
inline void check(int rc, const char *where)
{
    if (rc == MF_OK) { return; }
    std::string msg = std::string(where) + ": " + mf_last_error();
    switch (rc)
    {
    case MF_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case MF_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
    default: throw std::runtime_error(msg);
    }
}
// The engine handle behind the reference-compatible classes (friend of
// MFLib::MatchedFilter<T> and SingleChannel::MatchedFilter<T>).
struct Access
{
    template<class T> static mf_handle *handle(const MFLib::MatchedFilter<T> &mf);
    template<class T> static const MFLib::MatchedFilter<T> &inner(const MFLib::SingleChannel::MatchedFilter<T> &mf);
    template<class T> static MFLib::MatchedFilter<T> &inner(MFLib::SingleChannel::MatchedFilter<T> &mf);
    // Device-memory exchange (no reference counterpart), used by pymflib's
    // DLPack methods: the signal of row it from device memory on the
    // handle's device (the host copy behind getSignalPointer is fetched on
    // demand), and the device address of correlation row it after apply(),
    // its work complete; valid until the next apply(), initialize() or clear().
    template<class T> static void setSignalDevice(MFLib::MatchedFilter<T> &mf, int it, int nSamples, const T *dx);
    template<class T> static const T *rhoDevice(const MFLib::MatchedFilter<T> &mf, int it);
    template<class T> static int device(const MFLib::MatchedFilter<T> &mf);
};
template<class T> constexpr int precisionOf();
template<> constexpr int precisionOf<float>() { return MF_PRECISION_FLOAT; }
template<> constexpr int precisionOf<double>() { return MF_PRECISION_DOUBLE; }
}
}
#endif

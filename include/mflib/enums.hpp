#ifndef MFLIB_ENUMS_HPP
#define MFLIB_ENUMS_HPP 1
// Enumerations of MFLib's public API (include/mflib/enums.hpp in the
// reference).  Names and values are kept so that user code compiles
// unchanged against this library.
namespace MFLib
{
/// How the correlations are computed.  The GPU engine always uses FFT
/// block convolution and ignores this setting, as the reference's
/// MatchedFilter does.
enum class MatchedFilterImplementation
{
    DIRECT,
    OVERLAP_ADD,
    AUTO
};
/// SINGLE_CHANNEL: every template is applied to one signal.
/// MULTI_CHANNEL: template i is applied to signal i.
enum class MatchedFilterDetectionMode
{
    SINGLE_CHANNEL,
    MULTI_CHANNEL
};
enum class RelativeMagnitudeType
{
    GIBBONS_RINGDAL_2006 = 0,
    SCHAFF_RICHARDS_2014
};
enum class MaximumMatchedFilterPolicy
{
    MAXIMUM = 0,
    ABSOLUTE_MAXIMUM
};
enum class MatchedFilteredSignalDetectorPolicy
{
    REDUCED = 0,
    SINGLE
};
enum class DetectionTimeInterpolationType
{
    LANCZOS = 0,
    QUADRATIC,
    NONE
};
enum class AssociatorLikelihoodFunction
{
    GAUSSIAN = 0,
    EXPONENTIAL
};
/// First-motion polarity of a phase onset.
enum class Polarity : int
{
    UNKNOWN = 0,
    COMPRESSION = 1,
    DILATATION = -1
};
}
#endif

#ifndef MFLIB_VERSION_HPP
#define MFLIB_VERSION_HPP
// Version of the GPU engine behind MFLib's API (the reference generates
// its counterpart from include/mflib/version.hpp.in).
#define MFLIB_VERSION_MAJOR 0
#define MFLIB_VERSION_MINOR 2
#define MFLIB_VERSION_PATCH 0
#define MFLIB_VERSION "0.2.0"
#endif

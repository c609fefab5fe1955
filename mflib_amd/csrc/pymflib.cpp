// pymflib: the Python module of MFLib's detection path on the GPU engine.
// Class, method and enum names follow the reference's bindings
// (python/pymflib.cpp, python/waveformTemplate.cpp,
// python/networkStationPhase.cpp, python/multiChannelMatchedFilter*.cpp,
// python/singleChannel/matchedFilter*.cpp), so scripts written against
// them run unchanged.  Like the reference, the matched-filter classes are
// the FP64 instantiations and inputs are cast to float64.
//
// Additions with no reference counterpart:
//  - MultiChannel.MatchedFilter.shift_and_stack;
//  - get_signal_length / get_filtered_signal_length;
//  - MultiChannel.MatchedFilterParameters.set/get_stack_absolute_values;
//  - PeakFinder;
//  - SingleChannel.DetectionTimeInterpolationParameters and
//    DetectorParameters.set/get_detection_time_interpolation_parameters
//    (the reference's bindings leave the interpolation at its default);
//  - apply() and detect() release the GIL while the GPU works;
//  - device arrays through DLPack: set_signal also takes a float64 array
//    on the engine's GPU (a torch tensor, say), copied HBM to HBM with no
//    host round trip, and get_matched_filtered_signal_device returns a
//    zero-copy DLPack view of a correlation row (torch.from_dlpack).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mflib/enums.hpp"
#include "mflib/matchedFilter.hpp"
#include "mflib/matchedFilterParameters.hpp"
#include "mflib/networkStationPhase.hpp"
#include "mflib/singleChannel/matchedFilter.hpp"
#include "mflib/singleChannel/matchedFilterParameters.hpp"
#include "mflib/singleChannel/detection.hpp"
#include "mflib/singleChannel/detectionTimeInterpolationParameters.hpp"
#include "mflib/singleChannel/detector.hpp"
#include "mflib/singleChannel/detectorParameters.hpp"
#include "mflib/singleChannel/relativeMagnitude.hpp"
#include "mflib/version.hpp"
#include "mflib/waveformTemplate.hpp"
#include "private/peakFinder.hpp"
#include "facade/abi.hpp"

namespace py = pybind11;
using DArray = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace
{

py::array_t<double> toArray(const std::vector<double> &v)
{
    py::array_t<double> out(static_cast<py::ssize_t>(v.size()));
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

py::array_t<double> toArray(const double *p, int n)
{
    py::array_t<double> out(n);
    std::copy(p, p + n, out.mutable_data());
    return out;
}

int arrayLength(const DArray &x)
{
    if (x.ndim() != 1) { throw std::invalid_argument("signal must be one-dimensional"); }
    if (x.size() > 0x7fffffff) { throw std::invalid_argument("signal is too long"); }
    return static_cast<int>(x.size());
}

// ---- DLPack -------------------------------------------------------------
// The array-exchange structures of dlpack.h (ABI 0.x, as torch, numpy and
// cupy use them): a "dltensor" capsule holds a DLManagedTensor.
namespace dl
{
struct Device
{
    int32_t type;
    int32_t id;
};
struct DataType
{
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct Tensor
{
    void *data;
    Device device;
    int32_t ndim;
    DataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byteOffset;
};
struct Managed
{
    Tensor tensor;
    void *ctx;
    void (*deleter)(Managed *);
};
constexpr int32_t kCPU = 1;
constexpr int32_t kROCm = 10;
constexpr uint8_t kFloat = 2;
}  // namespace dl

// A correlation row in HBM, exported through __dlpack__.  `owner` keeps the
// matched filter (and so its device buffers) alive while any view exists.
struct DeviceRow
{
    py::object owner;
    const double *data = nullptr;
    int64_t n = 0;
    int device = 0;
};

struct RowCtx
{
    py::object owner;
    int64_t shape[1];
    int64_t strides[1];
};

void rowDeleter(dl::Managed *m)
{
    // consumers may call this without the GIL
    PyGILState_STATE g = PyGILState_Ensure();
    delete static_cast<RowCtx *>(m->ctx);
    delete m;
    PyGILState_Release(g);
}

void capsuleDestructor(PyObject *cap)
{
    // a capsule still named "dltensor" was never consumed: free it here
    if (PyCapsule_IsValid(cap, "dltensor"))
    {
        auto *m = static_cast<dl::Managed *>(PyCapsule_GetPointer(cap, "dltensor"));
        if (m && m->deleter) { m->deleter(m); }
    }
}

py::object exportRow(const DeviceRow &r)
{
    auto *ctx = new RowCtx{r.owner, {r.n}, {1}};
    auto *m = new dl::Managed{};
    m->tensor.data = const_cast<double *>(r.data);
    m->tensor.device = {dl::kROCm, r.device};
    m->tensor.ndim = 1;
    m->tensor.dtype = {dl::kFloat, 64, 1};
    m->tensor.shape = ctx->shape;
    m->tensor.strides = ctx->strides;
    m->tensor.byteOffset = 0;
    m->ctx = ctx;
    m->deleter = rowDeleter;
    PyObject *cap = PyCapsule_New(m, "dltensor", capsuleDestructor);
    if (!cap)
    {
        rowDeleter(m);
        throw py::error_already_set();
    }
    return py::reinterpret_steal<py::object>(cap);
}

void bindDeviceRow(py::module_ &m)
{
    py::class_<DeviceRow> c(m, "DeviceRow");
    c.doc() = "A correlation row in GPU memory (float64), exported zero-copy through DLPack.  The next "
              "apply() overwrites it in place; initialize() or clear() on its filter invalidates it.";
    c.def("__dlpack__", [](const DeviceRow &r, py::object, py::object, py::object, py::object) {
        return exportRow(r);
    }, py::arg("stream") = py::none(), py::arg("max_version") = py::none(), py::arg("dl_device") = py::none(),
       py::arg("copy") = py::none());
    c.def("__dlpack_device__", [](const DeviceRow &r) { return py::make_tuple(dl::kROCm, r.device); });
    c.def("__len__", [](const DeviceRow &r) { return r.n; });
    c.def_property_readonly("device", [](const DeviceRow &r) { return r.device; });
}

// A float64 signal on the GPU given through DLPack: (device pointer, n).
// Returns false when x is not a GPU array (the caller treats it as a host
// array).  The producer's pending work is ordered before the engine's by
// passing stream 0 (the default stream, with which the engine's blocking
// stream synchronizes).
bool deviceSignal(const py::object &x, int device, py::object &capsule, const double *&ptr, int &n)
{
    if (!py::hasattr(x, "__dlpack_device__") || !py::hasattr(x, "__dlpack__")) { return false; }
    auto dev = x.attr("__dlpack_device__")().cast<py::tuple>();
    if (dev[0].cast<int>() != dl::kROCm) { return false; }
    if (dev[1].cast<int>() != device)
    {
        throw std::invalid_argument("signal is on GPU " + std::to_string(dev[1].cast<int>()) +
                                    ", the filter on GPU " + std::to_string(device));
    }
    capsule = x.attr("__dlpack__")(py::arg("stream") = 0);
    auto *m = static_cast<dl::Managed *>(PyCapsule_GetPointer(capsule.ptr(), "dltensor"));
    if (!m) { throw py::error_already_set(); }
    const auto &t = m->tensor;
    if (t.ndim != 1) { throw std::invalid_argument("signal must be one-dimensional"); }
    if (t.dtype.code != dl::kFloat || t.dtype.bits != 64 || t.dtype.lanes != 1)
    {
        throw std::invalid_argument("a GPU signal must be float64");
    }
    if (t.strides && t.shape[0] > 1 && t.strides[0] != 1) { throw std::invalid_argument("signal must be contiguous"); }
    if (t.shape[0] > 0x7fffffff) { throw std::invalid_argument("signal is too long"); }
    n = static_cast<int>(t.shape[0]);
    ptr = reinterpret_cast<const double *>(static_cast<const char *>(t.data) + t.byteOffset);
    return true;
}

// setSignal from a host array or a GPU array (DLPack) for the FP64 filter
// behind either Python class.  The capsule (and the producer's buffer)
// lives until the copy into HBM has completed.
void setSignalAny(MFLib::MatchedFilter<double> &mf, int it, const py::object &x)
{
    if (!mf.isInitialized()) { throw std::runtime_error("Class is not initialized"); }
    py::object capsule;
    const double *ptr = nullptr;
    int n = 0;
    if (deviceSignal(x, MFLib::Facade::Access::device(mf), capsule, ptr, n))
    {
        MFLib::Facade::Access::setSignalDevice(mf, it, n, ptr);
        return;
    }
    DArray a = x.cast<DArray>();
    mf.setSignal(it, arrayLength(a), a.data());
}

// ---- NetworkStationPhase (python/networkStationPhase.cpp) ---------------
void bindNetworkStationPhase(py::module_ &m)
{
    using N = MFLib::NetworkStationPhase;
    py::class_<N> c(m, "NetworkStationPhase");
    c.def(py::init<>());
    c.doc() = "Network, station, channel, location code and phase of a waveform template.";
    c.def("set_network", &N::setNetwork);
    c.def("get_network", &N::getNetwork);
    c.def("set_station", &N::setStation);
    c.def("get_station", &N::getStation);
    c.def("set_channel", &N::setChannel);
    c.def("get_channel", &N::getChannel);
    c.def("set_location_code", &N::setLocationCode);
    c.def("get_location_code", &N::getLocationCode);
    c.def("set_phase", &N::setPhase);
    c.def("get_phase", &N::getPhase);
    c.def("clear", &N::clear);
    c.def("__eq__", [](const N &a, const N &b) { return a == b; });
    c.def("__hash__", &N::getHash);
    c.def(py::pickle(
        [](const N &p) {
            return py::make_tuple(p.getNetwork(), p.getStation(), p.getChannel(), p.getLocationCode(),
                                  p.getPhase());
        },
        [](const py::tuple &t) {
            if (t.size() != 5) { throw std::runtime_error("invalid NetworkStationPhase state"); }
            N p;
            p.setNetwork(t[0].cast<std::string>());
            p.setStation(t[1].cast<std::string>());
            p.setChannel(t[2].cast<std::string>());
            p.setLocationCode(t[3].cast<std::string>());
            p.setPhase(t[4].cast<std::string>());
            return p;
        }));
}

// ---- WaveformTemplate (python/waveformTemplate.cpp) ---------------------
std::vector<double> templateSignal(const MFLib::WaveformTemplate &w)
{
    const int n = w.getSignalLength();
    std::vector<double> v(n);
    double *p = v.data();
    w.getSignal(n, &p);
    return v;
}

void bindWaveformTemplate(py::module_ &m)
{
    using W = MFLib::WaveformTemplate;
    py::class_<W> c(m, "WaveformTemplate");
    c.def(py::init<>());
    c.doc() = "A template waveform with its sampling rate, onset time, travel time and stack weight.";
    c.def("set_signal", [](W &w, DArray x) { w.setSignal(arrayLength(x), x.data()); });
    c.def("get_signal", [](const W &w) { return toArray(templateSignal(w)); });
    c.def("set_sampling_rate", &W::setSamplingRate);
    c.def("get_sampling_rate", &W::getSamplingRate);
    c.def("set_shift_and_stack_weight", &W::setShiftAndStackWeight);
    c.def("get_shift_and_stack_weight", &W::getShiftAndStackWeight);
    c.def("set_phase_onset_time_in_signal", &W::setPhaseOnsetTime);
    c.def("get_phase_onset_time_in_signal", &W::getPhaseOnsetTime);
    c.def("set_phase_travel_time", [](W &w, double t) {
        // the reference's binding rejects this as a ValueError
        if (t < 0) { throw std::invalid_argument("travel time = " + std::to_string(t) + " must be positive"); }
        w.setTravelTime(t);
    });
    c.def("get_phase_travel_time", &W::getTravelTime);
    c.def("set_magnitude", &W::setMagnitude);
    c.def("get_magnitude", &W::getMagnitude);
    c.def("set_polarity", &W::setPolarity);
    c.def("get_polarity", &W::getPolarity);
    c.def("set_identifier", [](W &w, const py::sequence &id) {
        if (py::len(id) != 2) { throw std::invalid_argument("identifier must be [NetworkStationPhase, int]"); }
        w.setIdentifier({id[0].cast<MFLib::NetworkStationPhase>(), id[1].cast<uint64_t>()});
    });
    c.def("get_identifier", [](const W &w) {
        auto id = w.getIdentifier();
        return py::make_tuple(id.first, id.second);
    });
    c.def("clear", &W::clear);
    c.def(py::pickle(
        [](const W &w) {
            py::object sig = w.haveSignal() ? py::object(toArray(templateSignal(w))) : py::object(py::none());
            py::object fs = w.haveSamplingRate() ? py::object(py::float_(w.getSamplingRate())) : py::none();
            py::object on = w.havePhaseOnsetTime() ? py::object(py::float_(w.getPhaseOnsetTime())) : py::none();
            py::object tt = w.haveTravelTime() ? py::object(py::float_(w.getTravelTime())) : py::none();
            py::object mag = w.haveMagnitude() ? py::object(py::float_(w.getMagnitude())) : py::none();
            py::object id = py::none();
            if (w.haveIdentifier())
            {
                auto p = w.getIdentifier();
                id = py::make_tuple(p.first, p.second);
            }
            return py::make_tuple(sig, fs, w.getShiftAndStackWeight(), on, tt, mag,
                                  static_cast<int>(w.getPolarity()), id);
        },
        [](const py::tuple &t) {
            if (t.size() != 8) { throw std::runtime_error("invalid WaveformTemplate state"); }
            W w;
            if (!t[0].is_none())
            {
                DArray x = t[0].cast<DArray>();
                w.setSignal(arrayLength(x), x.data());
            }
            if (!t[1].is_none()) { w.setSamplingRate(t[1].cast<double>()); }
            w.setShiftAndStackWeight(t[2].cast<double>());
            if (!t[3].is_none()) { w.setPhaseOnsetTime(t[3].cast<double>()); }
            if (!t[4].is_none()) { w.setTravelTime(t[4].cast<double>()); }
            if (!t[5].is_none()) { w.setMagnitude(t[5].cast<double>()); }
            w.setPolarity(static_cast<MFLib::Polarity>(t[6].cast<int>()));
            if (!t[7].is_none())
            {
                py::tuple id = t[7].cast<py::tuple>();
                w.setIdentifier({id[0].cast<MFLib::NetworkStationPhase>(), id[1].cast<uint64_t>()});
            }
            return w;
        }));
}

// ---- parameters (python/multiChannelMatchedFilterParameters.cpp,
//      python/singleChannel/matchedFilterParameters.cpp) -----------------
template<class P>
void bindParameterCommon(py::class_<P> &c)
{
    c.def(py::init<>());
    c.def("add_template", &P::addTemplate);
    c.def("get_template", &P::getTemplate);
    c.def("get_signal_size", &P::getSignalSize);
    c.def("set_signal_size", &P::setSignalSize);
    c.def("get_number_of_templates", &P::getNumberOfTemplates);
    c.def("get_fft_length", &P::getFFTLength);
    c.def("set_fft_length", &P::setFFTLength);
    c.def("get_block_length", &P::getBlockLength);
    c.def("clear_templates", &P::clearTemplates);
    c.def("clear", &P::clear);
}

// ---- matched filters (python/multiChannelMatchedFilter.cpp,
//      python/singleChannel/matchedFilter.cpp) ---------------------------
void bindMultiChannel(py::module_ &mc)
{
    using P = MFLib::MatchedFilterParameters;
    py::class_<P> p(mc, "MatchedFilterParameters");
    p.doc() = "Templates and signal size of a multi-channel matched filter (template i matches signal i).";
    bindParameterCommon(p);
    // the reference binds block_length to getFFTLength; this returns the block length
    p.def("block_length", &P::getBlockLength);
    p.def("set_stack_absolute_values", &P::setStackAbsoluteValues);
    p.def("get_stack_absolute_values", &P::getStackAbsoluteValues);

    using M = MFLib::MatchedFilter<double>;
    py::class_<M> c(mc, "MatchedFilter");
    c.def(py::init<>());
    c.doc() = "Applies matched filtering to multiple channels where each channel has a specific template.";
    c.def("initialize", [](M &m, const P &parms) {
        if (!parms.isValid()) { throw std::invalid_argument("Parameters are invalid"); }
        m.initialize(parms);
    });
    c.def("set_signal", [](M &m, int it, const py::object &x) { setSignalAny(m, it, x); });
    c.def("zero_signal", &M::zeroSignal);
    c.def("get_number_of_templates", &M::getNumberOfTemplates);
    c.def("get_signal_length", &M::getSignalLength);
    c.def("get_filtered_signal_length", &M::getFilteredSignalLength);
    c.def("apply", &M::apply, py::call_guard<py::gil_scoped_release>());
    c.def("have_matched_filtered_signals", &M::haveMatchedFilteredSignals);
    c.def("get_matched_filtered_signal", [](const M &m, int it) {
        return toArray(m.getMatchedFilterSignalPointer(it), m.getFilteredSignalLength());
    });
    c.def("shift_and_stack", [](M &m) { return toArray(m.shiftAndStack()); });
    c.def("get_matched_filtered_signal_device", [](py::object self, int it) {
        const M &m = self.cast<const M &>();
        return DeviceRow{self, MFLib::Facade::Access::rhoDevice(m, it), m.getFilteredSignalLength(),
                         MFLib::Facade::Access::device(m)};
    });
}

void bindSingleChannel(py::module_ &sc)
{
    using P = MFLib::SingleChannel::MatchedFilterParameters;
    py::class_<P> p(sc, "MatchedFilterParameters");
    p.doc() = "Templates and signal size of a single-channel matched filter (every template matches one signal).";
    bindParameterCommon(p);

    using M = MFLib::SingleChannel::MatchedFilter<double>;
    py::class_<M> c(sc, "MatchedFilter");
    c.def(py::init<>());
    c.doc() = "Applies many templates to one signal.";
    c.def("initialize", [](M &m, const P &parms) {
        if (!parms.isValid()) { throw std::invalid_argument("Parameters are invalid"); }
        m.initialize(parms);
    });
    c.def("set_signal", [](M &m, const py::object &x) {
        if (!m.isInitialized()) { throw std::runtime_error("Class is not initialized"); }
        setSignalAny(MFLib::Facade::Access::inner(m), 0, x);
    });
    c.def("zero_signal", &M::zeroSignal);
    c.def("get_number_of_templates", &M::getNumberOfTemplates);
    c.def("get_signal_length", &M::getSignalLength);
    c.def("get_filtered_signal_length", &M::getFilteredSignalLength);
    c.def("apply", &M::apply, py::call_guard<py::gil_scoped_release>());
    c.def("have_matched_filtered_signals", &M::haveMatchedFilteredSignals);
    c.def("get_matched_filtered_signal", [](const M &m, int it) {
        return toArray(m.getMatchedFilterSignalPointer(it), m.getFilteredSignalLength());
    });
    c.def("get_matched_filtered_signal_device", [](py::object self, int it) {
        const auto &inner = MFLib::Facade::Access::inner(self.cast<const M &>());
        return DeviceRow{self, MFLib::Facade::Access::rhoDevice(inner, it), inner.getFilteredSignalLength(),
                         MFLib::Facade::Access::device(inner)};
    });
}

// ---- detector (python/singleChannel/detector.cpp, detectorParameters.cpp,
//      detection.cpp) ------------------------------------------------------
void bindDetector(py::module_ &sc)
{
    using I = MFLib::SingleChannel::DetectionTimeInterpolationParameters;
    py::class_<I> ip(sc, "DetectionTimeInterpolationParameters");
    ip.def(py::init<>());
    ip.doc() = "Sub-sample refinement of detection times: quadratic (default), Lanczos or none.";
    ip.def("clear", &I::clear);
    ip.def("enable_lanczos", &I::enableLanczos);
    ip.def("enable_quadratic", &I::enableQuadratic);
    ip.def("disable", &I::disable);
    ip.def("set_lanczos_alpha", &I::setLanczosAlpha);
    ip.def("get_lanczos_alpha", &I::getLanczosAlpha);
    ip.def("set_lanczos_number_of_interpolation_points", &I::setLanczosNumberOfInterpolationPoints);
    ip.def("get_lanczos_number_of_interpolation_points", &I::getLanczosNumberOfInterpolationPoints);
    ip.def("get_type", &I::getType);

    using P = MFLib::SingleChannel::DetectorParameters;
    py::class_<P> p(sc, "DetectorParameters");
    p.def(py::init<>());
    p.doc() = "Threshold, spacing, maxima and reduction policies, and products of the single-channel detector.";
    p.def("clear", &P::clear);
    p.def("get_maxima_policy", &P::getMaximaPolicy);
    p.def("set_maxima_policy", &P::setMaximaPolicy);
    p.def("get_matched_filtered_signal_detector_policy", &P::getMatchedFilteredSignalDetectorPolicy);
    p.def("set_matched_filtered_signal_detector_policy", &P::setMatchedFilteredSignalDetectorPolicy);
    p.def("get_minimum_detection_spacing", &P::getMinimumDetectionSpacing);
    p.def("set_minimum_detection_spacing", &P::setMinimumDetectionSpacing);
    p.def("get_detection_threshold", &P::getDetectionThreshold);
    p.def("set_detection_threshold", &P::setDetectionThreshold);
    p.def("enable_save_detected_waveform", &P::enableSaveDetectedWaveform);
    p.def("disable_save_detected_waveform", &P::disableSaveDetectedWaveform);
    p.def("want_detected_waveform", &P::wantDetectedWaveform);
    p.def("enable_save_amplitude_scaling_factor", &P::enableSaveAmplitudeScalingFactor);
    p.def("disable_save_amplitude_scaling_factor", &P::disableSaveAmplitudeScalingFactor);
    p.def("want_amplitude_scaling_factor", &P::wantAmplitudeScalingFactor);
    p.def("set_detection_time_interpolation_parameters", &P::setDetectionTimeInterpolationParameters);
    p.def("get_detection_time_interpolation_parameters",
          [](const P &parms) { return I(parms.getDetectionTimeInterpolationParameters()); });

    using D = MFLib::SingleChannel::Detection<double>;
    using R = MFLib::RelativeMagnitudeType;
    py::class_<D> d(sc, "Detection");
    d.def(py::init<>());
    d.doc() = "One detection: template, correlation, times, polarity, amplitude scaling factors, waveform.";
    d.def("clear", &D::clear);
    d.def("set_template_identifier", [](D &x, const py::sequence &id) {
        if (py::len(id) != 2) { throw std::invalid_argument("identifier must be [NetworkStationPhase, int]"); }
        x.setTemplateIdentifier({id[0].cast<MFLib::NetworkStationPhase>(), id[1].cast<uint64_t>()});
    });
    d.def("get_template_identifier", [](const D &x) {
        auto id = x.getTemplateIdentifier();
        return py::make_tuple(id.first, id.second);
    });
    d.def("have_template_identifier", &D::haveTemplateIdentifier);
    d.def("set_correlation_coefficient", &D::setCorrelationCoefficient);
    d.def("get_correlation_coefficient", &D::getCorrelationCoefficient);
    d.def("have_correlation_coefficient", &D::haveCorrelationCoefficient);
    d.def("set_detected_signal", [](D &x, DArray a) { x.setDetectedSignal(arrayLength(a), a.data()); });
    d.def("get_detected_signal", [](const D &x) {
        return toArray(x.getDetectedSignalPointer(), x.getDetectedSignalLength());
    });
    d.def("have_detected_signal", &D::haveDetectedSignal);
    d.def("set_detection_time", &D::setDetectionTime);
    d.def("get_detection_time", &D::getDetectionTime);
    d.def("have_detection_time", &D::haveDetectionTime);
    d.def("set_interpolated_detection_time", &D::setInterpolatedDetectionTime);
    d.def("get_interpolated_detection_time", &D::getInterpolatedDetectionTime);
    d.def("have_interpolated_detection_time", &D::haveInterpolatedDetectionTime);
    d.def("set_phase_onset_time", &D::setPhaseOnsetTime);
    d.def("get_phase_onset_time", &D::getPhaseOnsetTime);
    d.def("have_phase_onset_time", &D::havePhaseOnsetTime);
    d.def("set_interpolated_phase_onset_time", &D::setInterpolatedPhaseOnsetTime);
    d.def("get_interpolated_phase_onset_time", &D::getInterpolatedPhaseOnsetTime);
    d.def("have_interpolated_phase_onset_time", &D::haveInterpolatedPhaseOnsetTime);
    d.def("set_travel_time", &D::setTravelTime);
    d.def("get_travel_time", &D::getTravelTime);
    d.def("have_travel_time", &D::haveTravelTime);
    d.def("set_polarity", &D::setPolarity);
    d.def("get_polarity", &D::getPolarity);
    d.def("set_amplitude_scaling_factor", &D::setAmplitudeScalingFactor, py::arg("value"),
          py::arg("type") = R::GIBBONS_RINGDAL_2006);
    d.def("get_amplitude_scaling_factor", &D::getAmplitudeScalingFactor, py::arg("type") = R::GIBBONS_RINGDAL_2006);
    d.def("have_amplitude_scaling_factor", &D::haveAmplitudeScalingFactor,
          py::arg("type") = R::GIBBONS_RINGDAL_2006);
    d.def("get_magnitude_perturbation", &D::getMagnitudePerturbation, py::arg("type") = R::GIBBONS_RINGDAL_2006);
    // pickling (python/singleChannel/detection.cpp:497-660 pickles the same
    // fields; the detected waveform is kept here too)
    d.def(py::pickle(
        [](const D &x) {
            auto opt = [](bool have, double v) { return have ? py::object(py::float_(v)) : py::object(py::none()); };
            py::object id = py::none();
            if (x.haveTemplateIdentifier())
            {
                auto t = x.getTemplateIdentifier();
                id = py::make_tuple(t.first, t.second);
            }
            const bool gr = x.haveAmplitudeScalingFactor(R::GIBBONS_RINGDAL_2006);
            const bool sr = x.haveAmplitudeScalingFactor(R::SCHAFF_RICHARDS_2014);
            py::object wave = x.haveDetectedSignal()
                                  ? py::object(toArray(x.getDetectedSignalPointer(), x.getDetectedSignalLength()))
                                  : py::object(py::none());
            return py::make_tuple(
                opt(x.haveCorrelationCoefficient(), x.haveCorrelationCoefficient() ? x.getCorrelationCoefficient() : 0),
                id, opt(x.haveDetectionTime(), x.haveDetectionTime() ? x.getDetectionTime() : 0),
                opt(x.haveInterpolatedDetectionTime(),
                    x.haveInterpolatedDetectionTime() ? x.getInterpolatedDetectionTime() : 0),
                opt(x.havePhaseOnsetTime(), x.havePhaseOnsetTime() ? x.getPhaseOnsetTime() : 0),
                opt(x.haveInterpolatedPhaseOnsetTime(),
                    x.haveInterpolatedPhaseOnsetTime() ? x.getInterpolatedPhaseOnsetTime() : 0),
                opt(x.haveTravelTime(), x.haveTravelTime() ? x.getTravelTime() : 0),
                opt(gr, gr ? x.getAmplitudeScalingFactor(R::GIBBONS_RINGDAL_2006) : 0),
                opt(sr, sr ? x.getAmplitudeScalingFactor(R::SCHAFF_RICHARDS_2014) : 0),
                static_cast<int>(x.getPolarity()), wave);
        },
        [](const py::tuple &t) {
            if (t.size() != 11) { throw std::runtime_error("invalid Detection state"); }
            D x;
            if (!t[0].is_none()) { x.setCorrelationCoefficient(t[0].cast<double>()); }
            if (!t[1].is_none())
            {
                py::tuple id = t[1].cast<py::tuple>();
                x.setTemplateIdentifier({id[0].cast<MFLib::NetworkStationPhase>(), id[1].cast<uint64_t>()});
            }
            if (!t[2].is_none()) { x.setDetectionTime(t[2].cast<double>()); }
            if (!t[3].is_none()) { x.setInterpolatedDetectionTime(t[3].cast<double>()); }
            if (!t[4].is_none()) { x.setPhaseOnsetTime(t[4].cast<double>()); }
            if (!t[5].is_none()) { x.setInterpolatedPhaseOnsetTime(t[5].cast<double>()); }
            if (!t[6].is_none()) { x.setTravelTime(t[6].cast<double>()); }
            if (!t[7].is_none()) { x.setAmplitudeScalingFactor(t[7].cast<double>(), R::GIBBONS_RINGDAL_2006); }
            if (!t[8].is_none()) { x.setAmplitudeScalingFactor(t[8].cast<double>(), R::SCHAFF_RICHARDS_2014); }
            x.setPolarity(static_cast<MFLib::Polarity>(t[9].cast<int>()));
            if (!t[10].is_none())
            {
                DArray a = t[10].cast<DArray>();
                x.setDetectedSignal(arrayLength(a), a.data());
            }
            return x;
        }));

    using RM = MFLib::SingleChannel::RelativeMagnitude<double>;
    py::class_<RM> rm(sc, "RelativeMagnitude");
    rm.def(py::init<>());
    rm.doc() = "Amplitude of a detected window relative to its template (scored on the GPU, FP64).";
    rm.def("initialize", &RM::initialize);
    rm.def("is_initialized", &RM::isInitialized);
    rm.def("get_detected_signal_length", &RM::getDetectedSignalLength);
    rm.def("set_detected_signal", [](RM &m, DArray y) { m.setDetectedSignal(arrayLength(y), y.data()); });
    rm.def("have_detected_signal", &RM::haveDetectedSignal);
    rm.def("compute_amplitude_scaling_factor", &RM::computeAmplitudeScalingFactor,
           py::arg("type") = R::GIBBONS_RINGDAL_2006);
    rm.def("compute_magnitude_perturbation", &RM::computeMagnitudePerturbation,
           py::arg("type") = R::GIBBONS_RINGDAL_2006);

    using X = MFLib::SingleChannel::Detector<double>;
    py::class_<X> x(sc, "Detector");
    x.def(py::init<>());
    x.doc() = "Detections from a single-channel matched filter, computed on the GPU engine.";
    x.def("initialize", &X::initialize);
    x.def("detect", &X::detect, py::call_guard<py::gil_scoped_release>());
    x.def("get_number_of_detections", &X::getNumberOfDetections);
    x.def("get_detection", &X::getDetection);
    x.def("__len__", &X::getNumberOfDetections);
    x.def("__getitem__", [](const X &det, int i) {
        if (i < 0 || i >= det.getNumberOfDetections()) { throw py::index_error("detection index out of range"); }
        return det.getDetection(i);
    });
}

void bindPeakFinder(py::module_ &m)
{
    using F = MFLib::PeakFinder<double>;
    py::class_<F> c(m, "PeakFinder");
    c.def(py::init<>());
    c.doc() = "Thresholded local maxima with greedy minimum-distance pruning.";
    c.def("set_threshold", &F::setThreshold);
    c.def("get_threshold", &F::getThreshold);
    c.def("set_minimum_peak_distance", &F::setMinimumPeakDistance);
    c.def("get_minimum_peak_distance", &F::getMinimumPeakDistance);
    c.def("set_signal", [](F &f, DArray x) { f.setSignal(arrayLength(x), x.data()); });
    c.def("apply", &F::apply, py::call_guard<py::gil_scoped_release>());
    c.def("get_number_of_peaks", &F::getNumberOfPeaks);
    c.def("get_peak_indices", [](const F &f) {
        const int n = f.getNumberOfPeaks();
        py::array_t<int> out(n);
        if (n > 0)
        {
            int *p = out.mutable_data();
            f.getPeakIndices(n, &p);
        }
        return out;
    });
    c.def("clear", &F::clear);
}

}  // namespace

PYBIND11_MODULE(pymflib, m)
{
    m.attr("__version__") = MFLIB_VERSION;
    bindDeviceRow(m);
    bindWaveformTemplate(m);
    bindNetworkStationPhase(m);
    py::enum_<MFLib::RelativeMagnitudeType>(m, "RelativeMagnitudeType")
        .value("gibbons_ringdal_2006", MFLib::RelativeMagnitudeType::GIBBONS_RINGDAL_2006)
        .value("schaff_richards_2014", MFLib::RelativeMagnitudeType::SCHAFF_RICHARDS_2014);
    // the reference spells this enum "MaxmimumMatchedFilterPolicy"; both names are kept
    auto maxPolicy = py::enum_<MFLib::MaximumMatchedFilterPolicy>(m, "MaxmimumMatchedFilterPolicy")
        .value("maximum", MFLib::MaximumMatchedFilterPolicy::MAXIMUM)
        .value("absolute_maximum", MFLib::MaximumMatchedFilterPolicy::ABSOLUTE_MAXIMUM);
    m.attr("MaximumMatchedFilterPolicy") = maxPolicy;
    py::enum_<MFLib::MatchedFilteredSignalDetectorPolicy>(m, "MatchedFilteredSignalDetectorPolicy")
        .value("reduce", MFLib::MatchedFilteredSignalDetectorPolicy::REDUCED)
        .value("single", MFLib::MatchedFilteredSignalDetectorPolicy::SINGLE);
    py::enum_<MFLib::Polarity>(m, "Polarity")
        .value("unknown", MFLib::Polarity::UNKNOWN)
        .value("compression", MFLib::Polarity::COMPRESSION)
        .value("dilatation", MFLib::Polarity::DILATATION);
    py::enum_<MFLib::DetectionTimeInterpolationType>(m, "DetectionTimeInterpolationType")
        .value("quadratic", MFLib::DetectionTimeInterpolationType::QUADRATIC)
        .value("lanczos", MFLib::DetectionTimeInterpolationType::LANCZOS)
        .value("none", MFLib::DetectionTimeInterpolationType::NONE);
    py::module_ sc = m.def_submodule("SingleChannel");
    bindSingleChannel(sc);
    bindDetector(sc);
    py::module_ mc = m.def_submodule("MultiChannel");
    bindMultiChannel(mc);
    bindPeakFinder(m);
}

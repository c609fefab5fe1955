/*
 * mflib_gpu.h -- C ABI of the MI355X matched-filter engine.
 *
 * This is the drop-in boundary: plain pointers, sizes and int status codes,
 * no C++ or GPU-runtime types.  The C++ classes in include/mflib/ (the
 * reference's public API) hold one mf_handle each; pymflib sits on those.
 * Every entry point below names the reference interface it replaces.
 *
 * Status codes: MF_OK (0) or a negative MF_ERR_*.  After a failure,
 * mf_last_error() returns a thread-local message.  The C++ facade maps
 * MF_ERR_INVALID_ARGUMENT to std::invalid_argument and MF_ERR_BAD_STATE
 * to std::runtime_error, as the reference throws them
 * (src/matchedFilter.cpp:513-516, 829-834, 876-881, 977-980, 1322-1329).
 *
 * Threading: a handle is not thread-safe (neither is the reference's
 * MatchedFilter); distinct handles are independent.  Each handle owns one
 * HIP stream.  mf_apply() only enqueues work; every call that returns
 * results to the host waits for it.
 */
#ifndef MFLIB_GPU_H
#define MFLIB_GPU_H
#ifdef __cplusplus
extern "C" {
#endif

#define MF_OK 0
#define MF_ERR_INVALID_ARGUMENT (-1)
#define MF_ERR_BAD_STATE (-2)
#define MF_ERR_OUT_OF_MEMORY (-3)
#define MF_ERR_DEVICE (-4)
#define MF_ERR_UNSUPPORTED (-5)
#define MF_ERR_CANCELLED (-6)

/* Arithmetic type of signals and correlations: MatchedFilter<float> /
 * MatchedFilter<double> (src/matchedFilter.cpp:1434-1435). */
#define MF_PRECISION_FLOAT 0
#define MF_PRECISION_DOUBLE 1

/* MatchedFilterDetectionMode (include/mflib/enums.hpp:14-21), plus the
 * batched layout in which several templates share one channel. */
#define MF_MODE_SINGLE_CHANNEL 0
#define MF_MODE_MULTI_CHANNEL 1
#define MF_MODE_BATCHED 2

typedef struct mf_handle mf_handle;

/* Thread-local message for the last failed call on this thread. */
const char *mf_last_error(void);

/* Library version (include/mflib/version.hpp.in). */
int mf_version(int *major, int *minor, int *patch);

/* ---- instance lifecycle: MatchedFilter<T>::MatchedFilter / ~MatchedFilter
 *      / clear (include/mflib/matchedFilter.hpp:27-41) ------------------- */
int mf_create(mf_handle **h, int precision, int mode);
int mf_destroy(mf_handle *h);
/* MatchedFilter<T>::clear(): drops the plan, templates and device memory. */
int mf_clear(mf_handle *h);

/* ---- parameters: MatchedFilterParameters (include/mflib/matchedFilterParameters.hpp)
 * setSignalSize (src/matchedFilterParameters.cpp:318-327). */
int mf_set_signal_size(mf_handle *h, int n);
/* addTemplate (src/matchedFilterParameters.cpp:199-219) for all templates
 * at once, as initialize() reads them (src/matchedFilter.cpp:560-593).
 * samples[i] holds lengths[i] FP64 samples (WaveformTemplate stores FP64).
 * shifts/weights: the moveout nshift = int((travel-onset)*fs + 0.5) and
 * the shift-and-stack weight of each template (src/matchedFilter.cpp:577-592);
 * pass NULL shifts when any template lacks onset or travel time
 * (shiftAndStack then fails as the reference's does, 1326-1329). */
int mf_set_templates(mf_handle *h, int nTemplates, const int *lengths,
                     const double *const *samples, const int *shifts,
                     const double *weights);
/* setStackAbsoluteValues (src/matchedFilterParameters.cpp:338-348). */
int mf_set_stack_absolute_values(mf_handle *h, int absval);
/* No reference counterpart (MF_MODE_BATCHED only): channel lengths and
 * the channel / stacking group of every template row. */
int mf_set_batch_layout(mf_handle *h, int nChannels, const int *channelLengths,
                        const int *templateChannel, const int *templateGroup,
                        int nGroups);
/* FFT length: 0 lets the autotuner choose (the default); a power of two
 * forces it.  The reference's MatchedFilter ignores setFFTLength and uses
 * computeOptimalFFTAndBlockLength (src/matchedFilter.cpp:519-525). */
int mf_set_fft_length(mf_handle *h, int nfft);
/* No reference counterpart (tuning): correlation kernel variant, 0 = best
 * available (default), 1 = block-wide LDS FFTs, 2 = wave-level register
 * FFTs, 3 / 4 = the short-template time-domain path on MFMA with FP16 /
 * BF16 hi/lo operands (FP32 instances; rows of at most 256 taps, longer
 * rows stay on the FFT path).  A variant that does not exist for a block
 * size falls back to 0.  mf_get_plan reports nfft 0 for time-domain rows. */
int mf_set_kernel_variant(mf_handle *h, int variant);

/* No reference counterpart: with on != 0, mf_apply captures its
 * correlation launches (one per FFT size) into a HIP graph on first use and
 * replays that graph on every later apply, until the next mf_initialize. */
int mf_set_graph_launch(mf_handle *h, int on);

/* No reference counterpart (streaming state for continuous data): with
 * on != 0, set before mf_initialize, every channel keeps the last
 * carry = (longest template - 1) + (largest moveout shift) samples in HBM
 * ahead of its newest chunk, and mf_set_signal{,_device} push a chunk of
 * the signal size (the channel length in MF_MODE_BATCHED).  Rows then hold
 * chunk + carry - nb + 1 lags and stacks chunk + carry - nb + 1 - smax
 * (nb, smax: the group's longest template and largest shift); the LAST
 * `chunk` entries of every row and stack are the lags new with this push
 * (windows whose last sample lies in the chunk), so the tails of
 * consecutive pushes tile the lag axis with no gap or overlap.  Before the
 * first push the carry is zeros; mf_zero_signal restarts a channel.  The
 * dead-trace rule applies to carry + chunk.  mf_get_signal returns the
 * newest chunk. */
int mf_set_streaming(mf_handle *h, int on);
int mf_get_stream_carry(const mf_handle *h, int *carry);
/* No reference counterpart: one streaming push of every channel at once
 * (dx[c] holds channel c's chunk of n samples), waiting once for all the
 * copies instead of once per channel.  When every dx[c] is device memory of
 * the handle's device, two kernels move all channels; otherwise (host
 * memory, another device) each chunk goes through the copy engine. */
int mf_push_chunks_device(mf_handle *h, int nChannels, int n, const void *const *dx);
/* No reference counterpart (checkpoint / resume of a stream): the carry
 * samples of a channel that its next push continues from (*len = carry;
 * a size query when out is NULL), and the restore of such a state into a
 * streaming handle of the same templates and chunk length.  A handle
 * restored from a snapshot gives the next pushes' rows and stacks
 * bit-identical to the handle the snapshot came from.  A restore ends the
 * stream-peak state of every group (mf_set_stream_peaks_state restores
 * it afterwards). */
int mf_get_stream_state(mf_handle *h, int channel, void *out, int cap, int *len);
int mf_set_stream_state(mf_handle *h, int channel, const void *state, int n);
/* No reference counterpart (PeakFinder over a stream, the per-push form of
 * PeakFinder<T>::apply, src/peakFinder.cpp:468-586, on group `group`'s
 * stacks): call once after every apply of a streaming handle, with the same
 * threshold and minDist each time.  It returns in peaks[0..*nPeaks) the
 * stream lags that became final peaks with this push, in increasing order
 * across calls; with last != 0 the push ends the signal (its last lag is
 * judged by the end rule) and every remaining peak is returned, and the
 * next call starts a new stream.  Over a whole stream the returned lags are
 * exactly PeakFinder's list of the whole signal's stack.  Lag K of a
 * stream is the stack lag whose windows start K samples after the first
 * sample of the stream's first push; earlier lags (windows reaching into
 * the carry that push found) are not part of the stream.  A peak is final
 * once no later sample can change it, so it may be returned a few pushes
 * after its own.  When cap is too small the call fails with *nPeaks set and
 * the state unchanged (call again with room).  Needs one chunk length on
 * every channel.  The pending candidates are host state of the handle:
 * mf_get_stream_peaks_state saves them; mf_set_stream_state (a channel's
 * carry) ends every group's stream, so restore the carries first. */
int mf_stream_peaks(mf_handle *h, int group, double threshold, int minDist, int last, long long *peaks, int cap,
                    int *nPeaks);
/* No reference counterpart (checkpoint of mf_stream_peaks, the streamed
 * PeakFinder<T>::apply, src/peakFinder.cpp:468-586): group `group`'s
 * stream-peak state as opaque bytes (*len bytes; a size query when out is
 * NULL or cap too small, which fails with MF_ERR_INVALID_ARGUMENT and *len
 * set), and its restore into a streaming handle of the same layout and
 * precision whose channel carries were restored with mf_set_stream_state.
 * The restored group's next mf_stream_peaks call must follow the handle's
 * next apply; it then continues the saved stream exactly.  Bytes that are
 * not such a state fail with MF_ERR_INVALID_ARGUMENT and change nothing. */
int mf_get_stream_peaks_state(mf_handle *h, int group, void *out, long long cap, long long *len);
int mf_set_stream_peaks_state(mf_handle *h, int group, const void *state, long long n);
/* No reference counterpart: drops group `group`'s stream-peak state
 * (group -1: every group) whatever its push count, for instance after a
 * missed mf_stream_peaks call, which otherwise fails with MF_ERR_BAD_STATE;
 * the group's next call starts a new stream. */
int mf_stream_peaks_reset(mf_handle *h, int group);
/* No reference counterpart: mf_stream_peaks for every group of the handle
 * in one call (one stack launch, one candidate scan, one wait), as
 * mf_detect is to PeakFinder: group g's peaks are
 * peaks[offsets[g] .. offsets[g + 1]), offsets holding nGroups + 1 entries.
 * When cap is too small the call fails with offsets filled and no group's
 * state changed.  A group's stream may be driven by either call, one per
 * apply. */
int mf_stream_peaks_all(mf_handle *h, double threshold, int minDist, int last, long long *peaks, int cap,
                        int *offsets);

/* MatchedFilter<T>::initialize (src/matchedFilter.cpp:506-821): template
 * demean/normalize (FP64), plans, template spectra, device buffers. */
int mf_initialize(mf_handle *h);
int mf_is_initialized(const mf_handle *h, int *yes);
/* getNumberOfTemplates / getSignalLength / getFilteredSignalLength
 * (src/matchedFilter.cpp:1366-1402).  For MF_MODE_BATCHED the lengths are
 * those of channel 0 / group 0.  In streaming mode the signal length is the
 * chunk length. */
int mf_get_number_of_templates(const mf_handle *h, int *n);
int mf_get_signal_length(const mf_handle *h, int *n);
int mf_get_filtered_length(const mf_handle *h, int row, int *n);
/* The block plan of the channel that template row `row` reads. */
int mf_get_plan(const mf_handle *h, int row, int *nfft, int *blockLength);

/* setSignal(it, n, x) (src/matchedFilter.cpp:862-921): copies n host
 * samples of the handle's precision into channel `channel` and applies
 * the dead-trace rule |min| == |max| (894).  In single-channel mode the
 * channel is 0 (src/singleChannel/matchedFilter.cpp:61-67). */
int mf_set_signal(mf_handle *h, int channel, int n, const void *x);
/* No reference counterpart: the same from device memory (already in HBM). */
int mf_set_signal_device(mf_handle *h, int channel, int n, const void *dx);
/* zeroSignal(it) (src/matchedFilter.cpp:823-845). */
int mf_zero_signal(mf_handle *h, int channel);
/* getSignalPointer (923-938) reads the caller-side copy the C++ facade
 * keeps; this reads the device copy back. */
int mf_get_signal(mf_handle *h, int channel, void *out, int cap);

/* apply() (src/matchedFilter.cpp:971-1316): correlation + normalization
 * of every template row.  Enqueues work on the handle's stream. */
int mf_apply(mf_handle *h);
int mf_synchronize(mf_handle *h);
/* No reference counterpart: the HIP device the handle's buffers live on (the
 * device current at mf_create), for callers exchanging device pointers. */
int mf_get_device(const mf_handle *h, int *device);
int mf_have_matched_filtered_signals(const mf_handle *h, int *yes);

/* getMatchedFilteredSignal(it) (940-949): copies the row's
 * getFilteredSignalLength() correlation samples to host memory. */
int mf_get_rho(mf_handle *h, int row, void *out, int cap);
/* No reference counterpart: device pointer of row `row` (valid until the
 * next mf_initialize / mf_clear / mf_destroy).  It does not wait: the row
 * holds the last mf_apply's values once mf_synchronize has returned (or any
 * call that returns results to the host); work on it from another stream
 * (mf_peaks_device, say) must come after that. */
int mf_get_rho_device(mf_handle *h, int row, const void **dptr);

/* shiftAndStack() (src/matchedFilter.cpp:1318-1364) of stacking group
 * `group` (0 for single/multi-channel).  *len receives the stack length
 * n - nb + 1 - max(shift); out must hold that many values. */
int mf_shift_and_stack(mf_handle *h, int group, void *out, int cap, int *len);
/* No reference counterpart: stack into device memory (returns when the
 * stack is complete). */
int mf_shift_and_stack_device(mf_handle *h, int group, void *dout, int cap, int *len);

/* No reference counterpart: the batched detection step.  shiftAndStack of
 * every stacking group, then PeakFinder(threshold, minDist) on every stack
 * (src/matchedFilter.cpp:1318-1364, src/peakFinder.cpp:468-586), all on
 * the handle's stream with one launch per stage.  The peaks of group g are
 * idx[offsets[g] .. offsets[g+1]) in ascending order; offsets holds
 * nGroups+1 entries.  The stacks stay on the device (mf_get_stack_device). */
int mf_detect(mf_handle *h, double threshold, int minDist, int *idx, int cap, int *offsets);
int mf_get_stack_device(mf_handle *h, int group, const void **dptr, int *len);
int mf_get_number_of_groups(const mf_handle *h, int *n);

/* getSpectraOfTemplate(it) (src/matchedFilter.cpp:1411-1431).  Writes
 * nfft/2+1 complex values (re,im interleaved, handle precision) of the
 * spectrum of the demeaned, unit-norm, time-reversed template, zero padded
 * to the row's FFT length, in the reference's convention. */
int mf_template_spectrum(mf_handle *h, int row, void *outInterleaved, int cap, int *len);

/* PeakFinder<T>::apply (src/peakFinder.cpp:468-586) on n samples of the
 * given precision: threshold and local-maximum pass on the GPU, then the
 * greedy minimum-distance pruning.  Ties in value are kept in ascending
 * index order.  idx must hold cap entries; *nPeaks receives the count. */
int mf_peaks(const void *x, int n, int precision, double threshold, int minDist,
             int *idx, int cap, int *nPeaks);
/* No reference counterpart: the same on a device-resident signal.  It runs
 * on the peak finder's own blocking stream, which is ordered after work on
 * the default stream but not after a handle's stream: call mf_synchronize
 * on the handle first when dx is one of its rows. */
int mf_peaks_device(const void *dx, int n, int precision, double threshold,
                    int minDist, int *idx, int cap, int *nPeaks);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code between HIP graph captures, tight serving loops, harnesses that
 * forbid hipMalloc).  mf_peaks_device with its candidate counter and list
 * in a library-held device array (24 MiB on every GPU it runs on, kept
 * until the process ends) instead of buffers it allocates: no hipMalloc,
 * hipFree or pinned host allocation.  n <= 4,194,304 samples (larger n
 * fails with MF_ERR_INVALID_ARGUMENT).  Runs on the null stream, so it is
 * ordered after work on the default stream (synchronize a handle's stream
 * first when dx is one of its rows); blocks until done.  All calls on one
 * GPU share the array; the library serializes them with a process-wide
 * lock, since overlapping calls on one GPU would race on it. */
int mf_peaks_device_noalloc(const void *dx, int n, int precision, double threshold,
                            int minDist, int *idx, int cap, int *nPeaks);

/* No reference counterpart (precision study of normalizeSignal,
 * src/matchedFilter.cpp:94-166): the Pearson scale
 * sqrt(Lt) / sqrt(Lt S2 - S1^2) of each of the n - Lt + 1 windows of a
 * device-resident FP32 signal dx, from prefix scans of the samples built
 * with wavefront shuffle scans: FP64 scans (scanPrecision 0) or
 * compensated FP32 (float-float) scans (scanPrecision 1).  Windows whose
 * bracket is not positive get 0.  dScale (device) receives n - Lt + 1
 * floats; *ms (may be NULL) the kernel time.  2 <= Lt <= 2048.  Blocks
 * until done. */
int mf_window_scale_device(const void *dx, long long n, int Lt, int scanPrecision, void *dScale, double *ms);

/* Detector reduce + clamp: reduceAllCorrelations and cleanReducedTemplate
 * (src/singleChannel/detector.cpp:190-250), the step of Detector::detect
 * (detector.cpp:464-477) that feeds PeakFinder with the REDUCED policy.
 * Over the rows of `group` (group 0 holds every row in single- and
 * multi-channel mode), det[k] = max over rows of rho[k] (of |rho[k]| when
 * absoluteValue), clamped to [-1, 1], and id[k] = the template (row) index
 * of the first row reaching that maximum.  *len receives the group's
 * filtered length (also when det or id is NULL, as a size query that
 * returns MF_ERR_INVALID_ARGUMENT); det and id must hold that many
 * entries (cap). */
int mf_reduce_group(mf_handle *h, int group, int absoluteValue, void *det, int *id, int cap, int *len);
/* No reference counterpart: the same into device buffers. */
int mf_reduce_group_device(mf_handle *h, int group, int absoluteValue, void *dDet, int *dId, int cap, int *len);
/* Detector::detect's REDUCED policy up to the peak list (detector.cpp:464-477,
 * 542-552): the reduce above, PeakFinder<T>::apply on det (threshold,
 * minDist) and, per peak, the template index id[peak].  idx and
 * templateIds hold cap entries; *nPeaks the count. */
int mf_detect_reduced(mf_handle *h, int group, int absoluteValue, double threshold, int minDist, int *idx,
                      int *templateIds, int cap, int *nPeaks);

/* Detector::detect's SINGLE policy for template it (detector.cpp:478-491):
 * setCorrelations (252-284: rho or |rho| of that template, clamped to
 * [-1, 1]) and PeakFinder<T>::apply (threshold, minDist).  idx holds cap
 * entries; *nPeaks the count. */
int mf_detect_template(mf_handle *h, int it, int absoluteValue, double threshold, int minDist, int *idx, int cap,
                       int *nPeaks);
/* The same for every template at once (detector.cpp:478-491, the loop over
 * it): the peaks of template it are idx[offsets[it] .. offsets[it+1]) in
 * ascending order; offsets holds nTemplates + 1 entries, idx cap.  Rows are
 * scanned in chunks with one readback each. */
int mf_detect_templates(mf_handle *h, int absoluteValue, double threshold, int minDist, int *idx, int cap,
                        int *offsets);

/* Time interpolation kinds (values of MFLib::DetectionTimeInterpolationType). */
#define MF_INTERP_LANCZOS 0
#define MF_INTERP_QUADRATIC 1
#define MF_INTERP_NONE 2
/* Bits of mf_detection_products' flags. */
#define MF_PRODUCT_AMPLITUDE 1 /* alphaGR / alphaSR were computed */
#define MF_PRODUCT_TRUNCATED 2 /* peak + template length runs past the signal */
/* The per-detection products of SingleChannel::Detector::detect
 * (src/singleChannel/detector.cpp:537-664), for nPeaks detections, each a
 * lag peakIdx[k] on template row templateIds[k] of the last apply():
 *  - rho[k]: rho at the peak clamped to [-1, 1] (574-575); its sign gives
 *    the detection's polarity (623-629);
 *  - shift[k]: interpolated minus integer detection time, seconds
 *    (601-614): quadraticRefinement (38-89) or lanczosRefinement (111-152)
 *    with lanczosAlpha and lanczosPoints (DetectionTimeInterpolationParameters);
 *    0 for MF_INTERP_NONE;
 *  - alphaGR[k], alphaSR[k]: RelativeMagnitude's amplitude scaling factors
 *    (src/singleChannel/relativeMagnitude.cpp:35-58,362-388) of the window
 *    signal[peak .. peak + Lt) against the template, set with
 *    MF_PRODUCT_AMPLITUDE in flags[k] unless the window runs past the
 *    signal (MF_PRODUCT_TRUNCATED, detector.cpp:578-585) or is constant
 *    (relativeMagnitude.cpp:330-338).
 * samplingPeriods[it] is template it's 1/fs (NULL: 1, shifts in samples).
 * absoluteValue follows the detector's maxima policy.  Arithmetic is FP64
 * for both precisions.  Any output may be NULL.  Lanczos work is bounded:
 * (2 lanczosAlpha + 1) * (lanczosPoints - 1) <= 2^22 terms per detection
 * (MF_ERR_UNSUPPORTED beyond; the defaults 51 / 100 give 10,197). */
int mf_detection_products(mf_handle *h, int nPeaks, const int *peakIdx, const int *templateIds,
                          const double *samplingPeriods, int absoluteValue, int interpolation, int lanczosAlpha,
                          int lanczosPoints, double *rho, double *shift, double *alphaGR, double *alphaSR,
                          int *flags);
/* RelativeMagnitude<T> (src/singleChannel/relativeMagnitude.cpp:35-58,
 * 273-388) for a batch of detected windows against one template, on the
 * GPU: window k is windows[k*ld .. k*ld + Lt) (host samples of
 * `precision`), tmpl the template's Lt FP64 samples.  alphaGR[k] =
 * |<x, y>| / <x, x> (Gibbons & Ringdal 2006) and alphaSR[k] =
 * sqrt(<y, y> / <x, x>) (Schaff & Richards 2014) of the demeaned template x
 * and window y, in FP64; computed[k] = 0 for a constant window, which the
 * reference's setDetectedSignal rejects.  Outputs may be NULL. */
int mf_amplitude_factors(int precision, int Lt, const double *tmpl, int nWindows, const void *windows,
                         long long ld, double *alphaGR, double *alphaSR, int *computed);

/* computeOptimalFFTAndBlockLength(nb, nx)
 * (include/mflib/matchedFilterParameters.hpp:208, src/matchedFilterParameters.cpp:371-385). */
int mf_compute_optimal_fft_and_block_length(int nb, int nx, int *nfft, int *blockLength);

/* No reference counterpart (autotuner): times every FFT size and
 * correlation kernel variant on the running device (16 synthetic lanes x
 * 128 blocks with 1 row, 8 rows of one template length, 8 rows of distinct
 * lengths) and installs the fitted per-block costs for this process; FFT
 * plans (mf_initialize, the request server) then choose the FFT size per
 * lane and the variant per launch from the measured costs.  table
 * receives up to cap rows of (logM, variant, a, bShared, bSolo), in
 * microseconds per block; *n the number of rows.  Takes about a second. */
int mf_autotune(int precision, double *table, int cap, int *n);

/* No reference counterpart: per-apply kernel timings (ms) measured with
 * HIP events on the handle's stream: correlation kernels and the stack
 * kernel of the last mf_shift_and_stack* (for mf_detect, the one launch
 * that stacks and finds the candidates). */
int mf_get_timings(mf_handle *h, double *corrMs, double *stackMs);

/* No reference counterpart: window statistics of the correlation launches
 * since mf_initialize, summed over applies.  *windows: correlation outputs
 * computed (every row's length per apply).  *flagged: windows whose fast
 * value was replaced by an exact FP64 recomputation, counted on the
 * device: the fast path's error bound exceeded the tolerance, or the
 * window's variance lost too many digits.  *full: the flagged windows
 * recomputed in full (two-pass Pearson value: the cancelled variances, and
 * every flagged window of an FP64 handle); the others had their numerator
 * recomputed with the window norm kept.  *solo: windows of rows that
 * computed their window scale on their own wave instead of sharing a
 * block-wide one (rows whose template length fewer than four rows of the
 * same block share).  Any pointer may be NULL.  Waits for the handle's
 * stream. */
int mf_get_window_stats(mf_handle *h, long long *windows, long long *flagged, long long *full,
                        long long *solo);

/* ---- request server: no reference counterpart ---------------------------
 * Serving form of the path: a template library resident in HBM (FP64
 * templates and their spectra at every supported FFT size) and a request
 * loop with continuous batching.  A request is a window of nChannels
 * channels matched against nTemplates library templates: every template is
 * correlated with every channel, as one reference MatchedFilter instance
 * per (request, template) would do (src/matchedFilter.cpp:971-1316).  Row
 * (template i, channel c) has n_c - nb + 1 outputs, nb being the longest
 * template of the request (src/matchedFilter.cpp:1395-1402).
 *
 * A scheduler thread admits pending requests in arrival order while their
 * inputs and outputs fit in the HBM budget, launches every admitted request
 * as one batch (at most maxBatch requests) on one of maxInflight streams,
 * and keeps going while earlier batches still run, so that requests that
 * arrive meanwhile overlap on the device.  A request's results depend only
 * on that request: the same alone and in any mix.  maxBatch = 1 with
 * maxInflight = 1 is the one-request-at-a-time form. */
typedef struct mf_server mf_server;
int mf_server_create(mf_server **s, int precision, double budgetBytes, int maxBatch, int maxInflight);
/* Waits for launched batches, cancels pending requests, frees everything. */
int mf_server_destroy(mf_server *s);
/* Adds n templates (FP64 samples) to the library; their ids are
 * *firstId .. *firstId + n - 1.  Demean, unit norm and spectra are
 * computed on the GPU once (src/matchedFilter.cpp:168-190, 594-649). */
int mf_server_add_templates(mf_server *s, int n, const int *lengths, const double *const *samples,
                            int *firstId);
/* Queues a request.  channels[c] points to lengths[c] samples of the
 * server precision in device memory (onDevice != 0) or host memory; the
 * memory must stay valid until the request completes.  Device channels must
 * already hold their samples: the server reads them from its own streams,
 * which are not ordered after the caller's, so finish (synchronize) the
 * work that writes them before submitting.  Errors in the
 * request (bad ids, channels shorter than a template, a request larger
 * than the whole budget) are reported here. */
int mf_server_submit(mf_server *s, int nChannels, const int *lengths, const void *const *channels,
                     int onDevice, int nTemplates, const int *templateIds, long long *requestId);
/* Blocks until the request completes; *status = MF_OK or the failure code. */
int mf_server_wait(mf_server *s, long long requestId, int *status);
int mf_server_poll(mf_server *s, long long requestId, int *done);
/* Correlation row (template index i of the request, channel c): copy to
 * host, or its device address (valid until mf_server_release). */
int mf_server_get_rho(mf_server *s, long long requestId, int templateIndex, int channel, void *out,
                      int cap, int *len);
int mf_server_get_rho_device(mf_server *s, long long requestId, int templateIndex, int channel,
                             const void **dptr, int *len);
/* Milliseconds from submission to launch and to completion. */
int mf_server_latency(mf_server *s, long long requestId, double *queuedMs, double *totalMs);
/* Frees the request's HBM; waiting requests may then be admitted. */
int mf_server_release(mf_server *s, long long requestId);
/* Cancels a request that has not been launched yet: *cancelled = 1, its
 * waiters return with status MF_ERR_CANCELLED, and requests queued behind
 * it may be admitted.  A request already on the GPU (or done) runs to
 * completion: *cancelled = 0.  Either way release it afterwards. */
int mf_server_cancel(mf_server *s, long long requestId, int *cancelled);

/* Window statistics of the server's correlation launches since
 * mf_server_create, as mf_get_window_stats gives them for a handle: windows
 * of the launched batches, windows recomputed exactly (counted on the
 * device; final for a request once mf_server_wait has returned for it),
 * those recomputed in full, and windows of solo rows.  Any pointer may be
 * NULL. */
int mf_server_window_stats(mf_server *s, long long *windows, long long *flagged, long long *full,
                           long long *solo);
/* Batches launched, requests completed, and correlation samples completed. */
int mf_server_stats(mf_server *s, long long *batches, long long *requests, double *corrSamples);

/* No reference counterpart (tests of failure handling): the next n device
 * allocations the library makes, in any handle or server of this process,
 * fail with MF_ERR_OUT_OF_MEMORY as a failed hipMalloc would; n = 0
 * disarms.  *remaining (may be NULL) receives the count still armed
 * before this call. */
int mf_debug_fail_allocations(int n, int *remaining);

#ifdef __cplusplus
}
#endif
#endif

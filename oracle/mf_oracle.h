/*
 * CPU oracle for the matched-filter detection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 * The product (libmflib_gpu.so / libmflib.so / pymflib) never links,
 * loads or calls it.
 *
 * Two kinds of functions live here:
 *   1. Definitions (FP64 two-pass Pearson, block sizing, PeakFinder),
 *      which pin what the answer must be.
 *   2. A reference-faithful restatement of MFLib::MatchedFilter<T>::apply
 *      (same nfft/L table, overlap-add order, chunk-restart normalization,
 *      tol/tooBig/dead-trace rules) built on this file's own real FFT,
 *      because MKL/FFTW are not available.  It is also the CPU baseline.
 *
 * Every function cites the reference file:line it restates.
 * All functions return 0 on success and a negative value on bad input.
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* src/matchedFilterParameters.cpp:47-104 (computeWindowLength) and
 * 371-385 (computeOptimalFFTAndBlockLength argument checks). */
int orc_window_length(int nb, int nx, int *nfft, int *blockLength);
/* src/matchedFilterParameters.cpp:18-41 */
int orc_next_pow2(int n);

/* src/matchedFilter.cpp:168-190 (demeanNormalizeAndReverseTemplate). */
int orc_demean_normalize_reverse_f64(int nb, const double *b, double *out);
int orc_demean_normalize_reverse_f32(int nb, const float *b, float *out);

/* testing/utilities.hpp:33-69 (naivePearsonCorrelation), FP64, plain sums.
 * xc has nx-nb+1 entries. */
int orc_naive_pearson_f64(int nb, const double *b, int nx, const double *x,
                          double *xc);
/* The FP64 oracle of record: the same two-pass definition with
 * Neumaier-compensated sums.  Evaluated at the nLags lags in lags[]
 * (or at every lag 0..nx-nb when lags == NULL, nLags ignored). */
int orc_pearson_oracle_f64(int nb, const double *b, int nx, const double *x,
                           int nLags, const long long *lags, double *xc);
/* Same on float32 signal samples (promoted exactly to FP64). */
int orc_pearson_oracle_f32x(int nb, const double *b, int nx, const float *x,
                            int nLags, const long long *lags, double *xc);

/* The Pearson window scale 1 / ||w - mean(w)|| (= sqrt(nb) /
 * sqrt(nb S2 - S1^2) of normalizeSignal, src/matchedFilter.cpp:94-166) of
 * the windows x[i .. i + nb) of a float32 signal, FP64 two-pass with
 * compensated sums; 0 for a constant window.  At the nLags lags in lags[]
 * (or every lag 0..nx-nb when lags == NULL). */
int orc_window_scale_f32x(int nb, int nx, const float *x, int nLags, const long long *lags,
                          double *out);

/* Reference-faithful restatement of MatchedFilter<T>::initialize + apply
 * (src/matchedFilter.cpp:506-821, 971-1316).
 *   mode       0 = SINGLE_CHANNEL, 1 = MULTI_CHANNEL (enums.hpp:14-21)
 *   nT         number of templates
 *   lens[nT], templates[nT]  template samples (always double, as
 *              WaveformTemplate stores them, src/waveformTemplate.cpp:196-233)
 *   n          signal length
 *   signals    nT signals (multi-channel) or 1 signal (single-channel)
 *   floatQuirk (f32 only) 1 reproduces the float initialize() demeaning
 *              over the padded max length (src/matchedFilter.cpp:725-726);
 *              0 uses the true template length as the double path does.
 *   nThreads   OpenMP threads for the normalization loop (reference
 *              src/matchedFilter.cpp:1119); <=0 means the runtime default.
 *   out        nT rows of (n - nbMax + 1) samples, row-major.
 * Also writes the chosen nfft / L if the pointers are non-NULL. */
int orc_mf_apply_f64(int mode, int nT, const int *lens,
                     const double *const *templates, int n,
                     const double *const *signals, int nThreads,
                     double *out, int *nfftOut, int *lOut);
int orc_mf_apply_f32(int mode, int nT, const int *lens,
                     const double *const *templates, int n,
                     const float *const *signals, int floatQuirk,
                     int nThreads, float *out, int *nfftOut, int *lOut);

/* src/matchedFilter.cpp:1318-1364 (shiftAndStack).  rho holds nT rows of
 * nOut samples (row stride ld).  shifts/weights as recorded at
 * src/matchedFilter.cpp:577-592.  out must hold nOut - max(shift) values;
 * the produced length is written to *outLen. */
/* Detector reduce + clamp (src/singleChannel/detector.cpp:190-250): over
 * the nT rows of rho (row stride ld), det[i] = max of rho_t[i] (|rho_t[i]|
 * when absval), id[i] = the first t reaching it, then det clamped to
 * [-1, 1]; i < len. */
int orc_reduce_correlations_f32(int nT, const float *rho, long long ld, int len, int absval,
                                float *det, int *id);
int orc_reduce_correlations_f64(int nT, const double *rho, long long ld, int len, int absval,
                                double *det, int *id);

/* Per-detection products of SingleChannel::Detector::detect, FP64.
 * quadraticRefinement: src/singleChannel/detector.cpp:38-89. */
int orc_quadratic_shift(int n, const double *xc, int peak, double dt, int absval, double *shift);
/* lanczosRefinement: detector.cpp:91-152; profile (NULL or nRefine-1
 * entries) receives the interpolant at every grid point. */
int orc_lanczos_shift(int n, const double *xc, int peak, double dt, int absval, int alpha, int nRefine,
                      double *shift, double *profile);
/* RelativeMagnitude's scaling factors (src/singleChannel/relativeMagnitude.cpp:
 * 35-58, 76-98, 273-388) of window[0..Lt) against tmpl[0..Lt); *computed = 0
 * when the window is constant (the reference throws and detect skips it). */
int orc_amplitude_factors(int Lt, const double *tmpl, const double *window, double *alphaGR, double *alphaSR,
                          int *computed);

int orc_shift_and_stack_f64(int nT, const double *rho, long long ld,
                            int nOut, const int *shifts,
                            const double *weights, int absval,
                            double *out, int *outLen);
int orc_shift_and_stack_f32(int nT, const float *rho, long long ld,
                            int nOut, const int *shifts,
                            const double *weights, int absval,
                            float *out, int *outLen);

/* src/peakFinder.cpp:468-586 (PeakFinder<T>::apply).  Ties in the value
 * sort are broken by ascending index (the reference's std::sort leaves
 * them unspecified).  idx must hold n entries. */
int orc_peak_finder_f64(int n, const double *x, double threshold,
                        int minDist, int *idx, int *nPeaks);
int orc_peak_finder_f32(int n, const float *x, float threshold,
                        int minDist, int *idx, int *nPeaks);

/* This file's own real FFT (half-length complex packing, radix-2).
 * Forward: out[0..N/2] complex interleaved (re,im), unnormalized.
 * Used by the restatement; exported so tests can pin it against
 * numpy.fft. */
int orc_rfft_f64(int N, const double *in, double *outInterleaved);
int orc_irfft_f64(int N, const double *inInterleaved, double *out);

#ifdef __cplusplus
}
#endif
#endif

// Host-visible descriptors and launchers of the matched-filter kernels.
// No device code here: the engine (mf_engine.hip) fills these plain
// structs, copies them to device memory and calls the launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace mfk {

constexpr int kThreads = 256;        // block size of every kernel here
constexpr int kStackChunk = 2048;   // outputs per block of the stack kernel
constexpr int kDeadChunk = 16384;   // samples per block of the segment dead-trace pass
constexpr int kMinLogM = 9;          // smallest complex FFT size (M = 512)
constexpr int kMaxLogMFloat = 12;    // M = 4096 (nfft 8192) in FP32
constexpr int kMaxLogMDouble = 11;   // M = 2048 (nfft 4096) in FP64

// One (channel, FFT plan) lane of the correlation work.
struct Seg {
    const void *x;      // channel samples (handle precision), length n
    long long n;        // channel length
    int L;              // correlation outputs per block (nfft - nb + 1)
    int nOut;           // outputs needed on this lane (max over its rows)
    int rowBegin;       // rows [rowBegin, rowEnd) of the row table, by
    int rowEnd;         //   decreasing nOut, then increasing template length
    int dead;           // dead trace: every row of the lane is zero
    float hmax;         // max |template spectrum| over the lane's rows
    unsigned long long *flagged;  // running counts of windows recomputed
                                  //   exactly: [0] all, [1] those whose
                                  //   variance cancelled (null: not counted)
};

// corr2_kernel keeps a block-shared window-scale buffer (rows of a block
// in groups of four or more of equal template length share one scale) when
// a block's N = 2 M outputs of element size tsize fit in 8 KB of LDS; rows of
// smaller groups, and every row without the buffer, compute their scale on
// their own wave ("solo" rows).
constexpr bool corr2_shared_scale(int tsize, int M) { return tsize * M <= 4096; }
constexpr int kCorr2SharedGroup = 4;

// One template row.
struct Row {
    const void *H;        // conj(spectrum)/(2 nfft), nfft/2+1 complex
    const double *tmpl;   // demeaned unit-norm template (FP64), Lt samples
    void *out;            // correlation output row
    int nOut;             // outputs of this row
    int Lt;               // template length
    const void *x;        // the row's channel samples
};

// Queue of flagged windows for the exact recomputation kernel: 64-bit
// entries, output index << 32 | row | cancelled << 31 (row: index into the
// launch's row table).  The correlation launches append while *count stays
// within cap (the rest are recomputed in place); *count is cleared before
// them.  A wave whose reservation straddles cap fills the slots it leaves
// below cap with kFixEmpty, which the recomputation skips.  A null
// `entries` means no queue: every flagged window is recomputed in place.
constexpr unsigned long long kFixEmpty = ~0ull;
struct FixQueue {
    unsigned long long *count;
    int *entries;  // 8 bytes per entry
    unsigned long long cap;
    // row blocks whose flagged windows did not fit the queue: (row, first
    // output, outputs) per record, recomputed whole; sized by the host for
    // every row block of the launches (two records per row block at most),
    // so it cannot overflow.  Only the FFT wave-kernel queue has them.
    unsigned long long *overCount;
    int *over;  // 4 ints per record
};

// One work item: block `block` of lane `seg`, for the first nRows rows of
// the lane (a lane's rows are ordered by decreasing nOut, so the rows with
// outputs in a block are a prefix of them); nRows < 0: every row.
struct Item {
    int seg;
    int block;
    int nRows;
    int pad;
};

// Template preparation job (one per row).
struct TmplJob {
    const double *t;  // raw template samples
    int Lt;
    double *tOut;     // demeaned unit-norm template (FP64)
    void *H;          // spectrum row (handle precision)
    float *hmax;      // max |spectrum| of the demeaned unit-norm template
};

struct StackRow {
    const void *rho;  // correlation row
    int shift;        // moveout in samples
    int pad;
    double weight;    // normalised weight (already rounded to T)
};

// One stacking group: rows[0..nRows) in template order, stack length and
// destination; peak candidates of the group go to candIdx/candVal with
// capacity candCap and the counter *count.
struct StackGroup {
    const StackRow *rows;
    int nRows;
    int candCap;
    long long len;
    void *out;
    int *count;
    int *candIdx;
    void *candVal;
};

// ---- time-domain correlation of short templates on MFMA -----------------
constexpr int kTdFp16 = 3;              // kernel variant: FP16 operands split in two parts
constexpr int kTdBf16 = 4;              // kernel variant: BF16 operands split in three parts
constexpr int kTdOutputsPerItem = 512;  // outputs per work item
constexpr int kTdMaxTaps = 256;         // longest template of the path

// 16-bit parts each operand is split into (hi, lo for FP16; hi, mid, lo for
// BF16, whose 8-bit significand needs three parts for the FP32 contract).
constexpr int td_parts(int kind) { return kind == kTdBf16 ? 3 : 2; }
// bytes of one row's split taps
constexpr size_t td_frag_bytes(int kind, int KT) { return static_cast<size_t>(td_parts(kind)) * 2 * KT; }

// Tolerance of the fast value of a time-domain window, and the relative
// error eps of the split MFMA numerator: the dropped products and split
// residuals (3 u^2 for the FP16 pair, u = 2^-11; 3 u^3 for the BF16
// triple, u = 2^-8), FP32 accumulation over KT / 16 k-steps and the FP32
// epilogue.  A window's fast value is kept when eps sqrt(Lt S2) / den is
// within the tolerance.
constexpr double kTdRhoTol = 2.5e-6;
inline double td_eps(int kind, int KT)
{
    const double u = kind == kTdBf16 ? 1.0 / 256.0 : 1.0 / 2048.0;
    const double split = kind == kTdBf16 ? 3.0 * u * u * u : 3.0 * u * u;
    return split + (4.0 * std::sqrt(KT / 16.0) + 4.0) * 5.9604644775390625e-08;
}

// One template row of a time-domain lane.  The epilogue's per-row
// constants: with S1, S2 the window's sums of d = x - a and d^2 and
// den2 = Lt S2 - S1^2, a window is recomputed in full when
// den2 <= kc S2 + 1e-12 Lt E (E the block's sum of d^2: the variance lost
// its digits) and its numerator is recomputed when den2 < kn S2 (the
// error bound exceeds the tolerance).
struct TdRow {
    const void *frag;     // split taps: KT taps of each part in turn (16-bit)
    const double *tmpl;   // demeaned unit-norm template (FP64), Lt samples
    void *out;            // correlation output row (FP32)
    long long nOut;       // outputs of this row
    int Lt;               // template length
    float sqrtLt;         // sqrt(Lt)
    float kc;             // 2e-9 Lt^2
    float kn;             // (eps / tol)^2 Lt
    const float *x;       // the row's channel samples
    void *pad;
};

// One work item: outputs [seg0, seg0 + kTdOutputsPerItem) of lane `seg`
// for rows [rowBegin, rowBegin + nRows) (nRows <= 128) of the row table.
struct TdItem {
    int seg;
    int rowBegin;
    long long seg0;
    int nRows;
    int pad;
};

// Template split job (one per row).
struct TdPrepJob {
    const double *tmpl;
    int Lt;
    int pad;
    void *frag;
};

bool td_supported(int nb);
int td_taps(int nb);  // padded tap count KT (64, 128 or 256)
hipError_t launch_td_prep(int kind, const TdPrepJob *dJobs, int nJobs, int KT, hipStream_t stream);
hipError_t launch_td(int kind, int KT, const TdItem *dItems, int nItems, const Seg *dSegs, const TdRow *dRows,
                     FixQueue queue, hipStream_t stream);
// exact FP32 recomputation of the queued windows (after the correlation
// launches that filled the queue), for either row table
hipError_t launch_fix(const Row *dRows, FixQueue queue, int blocks, hipStream_t stream);  // and its overflow records
hipError_t launch_fix(const TdRow *dRows, FixQueue queue, int blocks, hipStream_t stream);

// Correlation + normalisation: one launch per FFT size.  Variant 1 runs
// block-wide LDS Stockham FFTs; variant 2 runs wave-level register FFTs
// (one wave per template row).
bool corr_variant_available(int precision, int logM, int variant);
hipError_t launch_corr(int precision, int logM, int variant, const Item *dItems, int nItems,
                       const Seg *dSegs, const Row *dRows, const void *dTwLo,
                       const void *dTwHi, FixQueue queue, hipStream_t stream);
// Template demean/normalise + spectra (FP64 arithmetic): one block per job.
hipError_t launch_tmpl(int precision, int logM, const TmplJob *dJobs, int nJobs,
                       const void *dTwLo64, const void *dTwHi64, hipStream_t stream);
// Shift and stack of nGroups groups (grid.y = group); maxLen bounds the
// grid.
hipError_t launch_stack_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                               int absval, hipStream_t stream);
// The same launch also finding the threshold/local-maximum candidates of
// every stack (into each group's count/candIdx/candVal, counts cleared by
// the caller): the stacks are written and tested in one pass.
hipError_t launch_stack_peak_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                                    int absval, double threshold, hipStream_t stream);
// Candidates of every group's stack (group.out), read back from memory.
hipError_t launch_peak_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                              double threshold, hipStream_t stream);
// The same over template rows as the Detector's per-template policy sees
// them: clamp(rho) or clamp(|rho|) to [-1, 1] (group.out = the row).
hipError_t launch_peak_rows(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                            double threshold, int absval, hipStream_t stream);
// Detector reduce + clamp over nRows device rows (pointer table on the
// device): det[k] = clamp(max_r rho_r[k] (or |rho_r[k]|), -1, 1), id[k] = the
// first r reaching the maximum; k < len.
hipError_t launch_reduce(int precision, const void *const *dRowPtrs, int nRows, long long len, int absval,
                         void *dDet, int *dId, hipStream_t stream);
// Every group's peak candidates packed group by group: dPackedIdx[0..nG)
// = the counts, then the indices from dPackedIdx[nG]; values from
// dPackedVal[0].  Capacity: nG + sum of the groups' candCap.
hipError_t launch_pack_candidates(int precision, const StackGroup *dGroups, int nGroups, int *dPackedIdx,
                                  void *dPackedVal, hipStream_t stream);
// Streaming push of up to kPushMaxChannels channels at once (the table
// travels as a kernel argument): buffer c, carry + chunk words long, keeps
// its last carryWords words in front of the chunkWords words of src c.
// dScratch holds nChannels * carryWords words.
constexpr int kPushMaxChannels = 64;
struct PushTable {
    void *buf[kPushMaxChannels];
    const void *src[kPushMaxChannels];
};
hipError_t launch_push(const PushTable &tab, int nChannels, long long chunkWords, long long carryWords,
                       unsigned *dScratch, hipStream_t stream);
// dst[i] = value, i < n: clears counters with a kernel node instead of a
// memset node (safe inside a captured graph).
hipError_t launch_fill_int(int *dDst, long long n, int value, hipStream_t stream);
// out[i] = src[idx[i]], i < n (all device pointers).
hipError_t launch_gather_int(const int *dSrc, const int *dIdx, int n, int *dOut, hipStream_t stream);
// Threshold + local-maximum candidates (unordered, deterministic set).
hipError_t launch_peaks(int precision, const void *dx, long long n, double threshold,
                        int *dCount, int *dIdx, void *dVal, int cap, hipStream_t stream);
// Dead-trace flag of every segment in a device-resident table, computed
// on the device: seg.dead = (|min x| == |max x|) over the segment's whole
// channel.  dPart holds 2 * nSegs * ceil(maxN / kDeadChunk) doubles.
hipError_t launch_seg_dead(int precision, Seg *dSegs, int nSegs, long long maxN, double *dPart,
                           hipStream_t stream);
// Per-block min and max of a signal (dead-trace test on device input).
hipError_t launch_minmax(int precision, const void *dx, long long n, double *dPartMin,
                         double *dPartMax, int nBlocks, hipStream_t stream);

// Window scales sqrt(Lt)/sqrt(Lt S2 - S1^2) of the n - Lt + 1 windows of a
// FP32 signal from prefix scans: FP64 scans (kScanF64) or compensated
// FP32 (float-float) scans (kScanF32c).  2 <= Lt <= kScanMaxLt.
constexpr int kScanF64 = 0;
constexpr int kScanF32c = 1;
constexpr int kScanMaxLt = 2048;
hipError_t launch_scan_scale(int mode, const float *dx, long long n, int Lt, float *dScale, hipStream_t stream);

// Per-detection products (mf_products.hip): one wavefront per detection.
// Interpolation kinds use the values of MFLib::DetectionTimeInterpolationType.
constexpr int kInterpLanczos = 0;
constexpr int kInterpQuadratic = 1;
constexpr int kInterpNone = 2;
// Lanczos work per detection, (2 alpha + 1) taps at (points - 1) grid
// shifts, is bounded so one wavefront's loop stays within milliseconds
// (the default 51 / 100 is 10,197 terms).
constexpr long long kLanczosMaxTerms = 1LL << 22;
constexpr int kProductAmplitude = 1;  // scaling factors computed
constexpr int kProductTruncated = 2;  // peak + Lt runs past the signal: no amplitude
struct ProductJob {
    const void *rho;     // the template's correlation row (T), nOut lags
    const void *sig;     // the template's channel (T), sigLen samples
    const double *tmpl;  // the template's samples as given (FP64), Lt
    long long sigLen;
    double dt;           // sampling period of the template, s
    int nOut, peak, Lt, pad;
};
struct ProductOut {
    double rho, shift, alphaGR, alphaSR;
    int flags, pad;
};
hipError_t launch_products(int precision, const ProductJob *dJobs, int nJobs, int absval, int interp,
                           int lanczosAlpha, int lanczosPoints, ProductOut *dOut, hipStream_t stream);
// RelativeMagnitude's factors of nWindows windows (window k at
// dWindows + k * ld elements, Lt samples) against one template: alphaGR,
// alphaSR and kProductAmplitude in dOut[k] (0 flags: constant window).
hipError_t launch_amplitude(int precision, const double *dTmpl, int Lt, const void *dWindows, long long ld,
                            int nWindows, ProductOut *dOut, hipStream_t stream);

}  // namespace mfk

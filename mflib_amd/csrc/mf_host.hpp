// Host-side helpers shared by the engine (mf_engine.hip) and the request
// server (mf_server.hip): error plumbing of the C ABI, device and pinned
// buffers, FFT planning and the XCD-aware ordering of work items.
#ifndef MF_HOST_HPP
#define MF_HOST_HPP

#include "mflib_gpu.h"
#include "mf_kernels.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace mfh {


inline thread_local std::string g_lastError;

struct MfError {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw MfError{code, msg}; }

#define MF_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            fail(e_ == hipErrorOutOfMemory ? MF_ERR_OUT_OF_MEMORY : MF_ERR_DEVICE,          \
                 std::string(#expr) + " failed: " + hipGetErrorString(e_));                 \
        }                                                                                   \
    } while (0)

template <class F>
int guard(F &&f)
{
    try {
        f();
        return MF_OK;
    } catch (const MfError &e) {
        g_lastError = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_lastError = "out of host memory";
        return MF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_lastError = e.what();
        return MF_ERR_DEVICE;
    } catch (...) {
        g_lastError = "unknown failure";
        return MF_ERR_DEVICE;
    }
}

// RAII device allocation.
// Allocation-failure injection for tests of the failure paths
// (mf_debug_fail_allocations): while positive, each device allocation of
// the library takes one and fails as out of memory.
inline std::atomic<int> &injectedAllocFailures()
{
    static std::atomic<int> n{0};
    return n;
}

inline bool takeInjectedAllocFailure()
{
    auto &n = injectedAllocFailures();
    int v = n.load();
    while (v > 0) {
        if (n.compare_exchange_weak(v, v - 1)) { return true; }
    }
    return false;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) { (void)hipFree(p); }
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t nb)
    {
        release();
        if (nb == 0) { return; }
        if (takeInjectedAllocFailure()) {
            fail(MF_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(nb) + " bytes failed (injected failure)");
        }
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) {
            p = nullptr;
            fail(e == hipErrorOutOfMemory ? MF_ERR_OUT_OF_MEMORY : MF_ERR_DEVICE,
                 "hipMalloc of " + std::to_string(nb) + " bytes failed: " + hipGetErrorString(e));
        }
        bytes = nb;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

inline size_t roundUp(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Same table and rule as computeWindowLength
// (src/matchedFilterParameters.cpp:47-104, 371-385).
inline const double kFftCost[21] = {18,       59,       138,      303,      660,     1441,    3150,
                             6875,     14952,    32373,    69762,    149647,  319644,  680105,
                             1441974,  3047619,  6422736,  13500637, 28311786, 59244791,
                             59244791 * 2.09};

inline int referenceWindowLength(int nb, int nx, int *nfft, int *L)
{
    if (nb < 1) { fail(MF_ERR_INVALID_ARGUMENT, "Filter length must be positive"); }
    if (nx <= nb) {
        fail(MF_ERR_INVALID_ARGUMENT, "Signal length = " + std::to_string(nx) +
                                          " must exceed filter length = " + std::to_string(nb));
    }
    int first = -1;
    for (int i = 0; i < 21; ++i) {
        if ((2LL << i) > nb - 1) { first = i; break; }
    }
    if (first < 0) {
        long long p = 1;
        while (p < nx) { p <<= 1; }
        if (p > 0x7fffffffLL) { fail(MF_ERR_INVALID_ARGUMENT, "Overflow error"); }
        *nfft = static_cast<int>(p);
        *L = static_cast<int>(p) - (nb - 1);
        return 0;
    }
    double best = 1e308;
    for (int i = first; i < 21; ++i) {
        long long len = 2LL << i;
        long long lt = len - (nb - 1);
        double obj = std::ceil(nx / static_cast<double>(lt)) * kFftCost[i];
        if (obj < best) {
            best = obj;
            *nfft = static_cast<int>(len);
            *L = static_cast<int>(lt);
        }
    }
    return 0;
}

// Two-level twiddle tables for W_{2M}: lo[i] = W^i (i < 64),
// hi[j] = W^(64 j) (j < 2M/64), FP64 from cos/sin, rounded to T.
template <typename T>
std::vector<T> twiddleTable(int logM, bool high)
{
    const long long N = 2LL << logM;
    const int count = high ? static_cast<int>(N / 64) : 64;
    std::vector<T> v(2 * count);
    for (int i = 0; i < count; ++i) {
        const long long e = high ? 64LL * i : i;
        const double ang = -2.0 * M_PI * static_cast<double>(e) / static_cast<double>(N);
        v[2 * i] = static_cast<T>(std::cos(ang));
        v[2 * i + 1] = static_cast<T>(std::sin(ang));
    }
    return v;
}

struct TwTables {
    DevBuf lo, hi, lo64, hi64;
};

// Uploads the FP64 twiddle tables of FFT size 2^(logM+1) and, for FP32
// work, their FP32 roundings.
inline void makeTwiddles(TwTables &t, int logM, bool wantFloat)
{
    auto lo64 = twiddleTable<double>(logM, false);
    auto hi64 = twiddleTable<double>(logM, true);
    t.lo64.alloc(lo64.size() * sizeof(double));
    t.hi64.alloc(hi64.size() * sizeof(double));
    MF_HIP(hipMemcpy(t.lo64.p, lo64.data(), t.lo64.bytes, hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(t.hi64.p, hi64.data(), t.hi64.bytes, hipMemcpyHostToDevice));
    if (wantFloat) {
        auto lo = twiddleTable<float>(logM, false);
        auto hi = twiddleTable<float>(logM, true);
        t.lo.alloc(lo.size() * sizeof(float));
        t.hi.alloc(hi.size() * sizeof(float));
        MF_HIP(hipMemcpy(t.lo.p, lo.data(), t.lo.bytes, hipMemcpyHostToDevice));
        MF_HIP(hipMemcpy(t.hi.p, hi.data(), t.hi.bytes, hipMemcpyHostToDevice));
    }
}

// First device present, or a failure: the engine has no CPU fallback.
inline void requireDevice()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fail(MF_ERR_DEVICE, "no HIP device is available (the engine has no CPU fallback)");
    }
}


// RAII pinned host allocation (hipHostMalloc).
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void release()
    {
        if (p) { (void)hipHostFree(p); }
        p = nullptr;
        bytes = 0;
    }
    void reserve(size_t nb)
    {
        if (nb <= bytes) { return; }
        release();
        hipError_t e = hipHostMalloc(&p, nb, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            fail(MF_ERR_OUT_OF_MEMORY, "hipHostMalloc of " + std::to_string(nb) + " bytes failed: " +
                                           hipGetErrorString(e));
        }
        bytes = nb;
    }
};

inline int maxLogM(int precision)
{
    return precision == MF_PRECISION_FLOAT ? mfk::kMaxLogMFloat : mfk::kMaxLogMDouble;
}

// Row mix of a lane: rows whose template length is shared by at least
// four rows of the lane share one normalisation scale; the others are
// solo rows that each compute their own.
struct RowMix {
    int shared = 0;
    int solo = 0;
    int rows() const { return shared + solo; }
};

inline RowMix rowMix(std::vector<int> lengths)
{
    std::sort(lengths.begin(), lengths.end());
    RowMix m;
    for (size_t i = 0; i < lengths.size();) {
        size_t j = i;
        while (j < lengths.size() && lengths[j] == lengths[i]) { ++j; }
        if (j - i >= 4) { m.shared += static_cast<int>(j - i); } else { m.solo += static_cast<int>(j - i); }
        i = j;
    }
    return m;
}

// Per-block cost of the correlation kernels: a + bShared * shared rows +
// bSolo * solo rows (microseconds per block for a full GPU), per FFT size
// and kernel variant.  The static table comes from MI355X measurements;
// mf_autotune replaces it with a measurement on the running device.
struct CostTable {
    bool measured = false;
    double a[mfk::kMaxLogMFloat + 1][3] = {};
    double bShared[mfk::kMaxLogMFloat + 1][3] = {};
    double bSolo[mfk::kMaxLogMFloat + 1][3] = {};
};

inline std::mutex &costMutex()
{
    static std::mutex m;
    return m;
}

// The FP32 default is mf_autotune's measurement on an MI355X (microseconds
// per block for the whole GPU; variant 1 block-wide, 2 wave-level); FP64
// starts from the static model below.
inline CostTable defaultCostTable(int precision)
{
    CostTable t;
#ifdef MF_STATIC_FORMULA
    return t;
#endif
    if (precision != MF_PRECISION_FLOAT) { return t; }
    struct E { int logM, v; double a, bS, bO; };
    static const E kMi355x[] = {
        {9, 1, 0.00949, 0.00331, 0.00551},  {9, 2, 0.01247, 0.00135, 0.00237},
        {10, 1, 0.01468, 0.00583, 0.00882}, {10, 2, 0.02817, 0.00176, 0.00415},
        {11, 1, 0.03286, 0.01595, 0.02210}, {11, 2, 0.05440, 0.00756, 0.00747},
        {12, 1, 0.08531, 0.05636, 0.07194},
    };
    for (const E &e : kMi355x) {
        t.a[e.logM][e.v] = e.a;
        t.bShared[e.logM][e.v] = e.bS;
        t.bSolo[e.logM][e.v] = e.bO;
    }
    t.measured = true;
    return t;
}

inline CostTable &costTable(int precision)
{
    static CostTable t[2] = {defaultCostTable(MF_PRECISION_FLOAT), defaultCostTable(MF_PRECISION_DOUBLE)};
    return t[precision == MF_PRECISION_FLOAT ? 0 : 1];
}

// Static per-block cost: forward FFT plus, per row, product + inverse FFT +
// epilogue, proportional to M (log2 M + 1.5) and weighted per block size by
// the measured efficiency of the default variant there (16 rows of 512
// samples, relative to M = 1024); solo rows cost 1.6x a shared row
// (measured: 16 distinct lengths ran at 216 against 339 G samples/s).
inline double staticBlockCost(int logM, RowMix mix, int precision)
{
    const double M = static_cast<double>(1LL << logM);
    static const double occF[] = {0.97, 1.00, 1.97, 3.30};  // logM 9..12, FP32
    static const double occD[] = {1.00, 1.60, 2.40, 9.99};
    const double occ = (precision == MF_PRECISION_FLOAT ? occF : occD)[logM - mfk::kMinLogM];
    return (1.0 + 1.15 * mix.shared + 1.84 * mix.solo) * M * (logM + 1.5) * occ;
}

inline bool variantAvailable(int precision, int logM, int variant)
{
    return mfk::corr_variant_available(precision, logM, variant);
}

// Default correlation kernel variant for an FFT size: the wave-level kernel
// up to nfft 4096 in FP32 and 1024 in FP64 (measured faster there: at nfft
// 4096 with 50 rows per lane it runs 1.2-1.3x the block-wide kernel), the
// block-wide kernel for the larger blocks.
inline int staticVariant(int precision, int logM)
{
    const bool waveFaster = precision == MF_PRECISION_FLOAT ? logM <= 11 : logM <= 9;
    return waveFaster && variantAvailable(precision, logM, 2) ? 2 : 1;
}

// The rows of a lane as the cost model sees them: each row's template
// length and number of outputs.  A row works only in the blocks holding its
// outputs, so a lane's cost is its blocks' fixed part plus, per row, the
// blocks of that row.
struct LaneRows {
    std::vector<int> Lt;
    std::vector<long long> nOut;
    long long maxOut() const
    {
        long long m = 0;
        for (long long v : nOut) { m = std::max(m, v); }
        return m;
    }
    int nb() const
    {
        int m = 0;
        for (int v : Lt) { m = std::max(m, v); }
        return m;
    }
    RowMix mix() const { return rowMix(Lt); }
};

// Row blocks of a lane at block length L, split into rows that share a
// normalisation scale and solo rows (as rowMix splits them).
inline void laneRowBlocks(const LaneRows &rs, long long L, double &shared, double &solo)
{
    std::vector<int> order(rs.Lt.size());
    for (size_t i = 0; i < order.size(); ++i) { order[i] = static_cast<int>(i); }
    std::sort(order.begin(), order.end(), [&](int a, int b) { return rs.Lt[a] < rs.Lt[b]; });
    shared = solo = 0;
    for (size_t i = 0; i < order.size();) {
        size_t j = i;
        while (j < order.size() && rs.Lt[order[j]] == rs.Lt[order[i]]) { ++j; }
        double rb = 0;
        for (size_t k = i; k < j; ++k) {
            rb += std::ceil(static_cast<double>(rs.nOut[order[k]]) / static_cast<double>(L));
        }
        (j - i >= 4 ? shared : solo) += rb;
        i = j;
    }
}

// Windows of the work items [first, first + count) of one corr2_kernel
// launch whose row computes its window scale on its own wave, grouped as
// the kernel groups a block's rows (runs of equal template length in the
// row table; runs of kCorr2SharedGroup or more share one scale where the
// kernel keeps the shared buffer).
inline long long soloWindows(const mfk::Item *items, int first, int count, const mfk::Seg *segs,
                             const mfk::Row *rows, bool sharedBuffer)
{
    long long solo = 0;
    for (int k = first; k < first + count; ++k) {
        const mfk::Item &it = items[k];
        const mfk::Seg &sg = segs[it.seg];
        const long long base = static_cast<long long>(it.block) * sg.L;
        const int end = it.nRows < 0 ? sg.rowEnd : std::min(sg.rowEnd, sg.rowBegin + it.nRows);
        for (int g0 = sg.rowBegin; g0 < end;) {
            int g1 = g0 + 1;
            while (g1 < end && rows[g1].Lt == rows[g0].Lt) { ++g1; }
            if (!(sharedBuffer && g1 - g0 >= mfk::kCorr2SharedGroup)) {
                for (int r = g0; r < g1; ++r) {
                    solo += std::max(0ll, std::min(static_cast<long long>(sg.L), rows[r].nOut - base));
                }
            }
            g0 = g1;
        }
    }
    return solo;
}

// Risk margin of larger FFT blocks: each doubling of the block over the
// smallest size adds kBlockRisk to the predicted cost.  Two effects the
// per-block table does not see grow with the block: windows recomputed
// exactly (the FFT noise bound of a block holding a loud stretch flags
// every quiet window of that block, so the flagged share of a row grows
// with the block length), and the per-output scale work of rows whose
// templates are much shorter than the calibration's N/4.  On the
// precision-stress workload (gaps at other levels) nfft 1024 measured 2.1x
// faster than 2048 and 3.1x faster than 4096, while clean data ran within
// 20 % at all three.
constexpr double kBlockRisk = 0.5;
inline double blockRisk(int logM) { return 1.0 + kBlockRisk * (logM - mfk::kMinLogM); }

// Extra cost of a lane whose template spectra do not stay in its XCD's L2
// (4 MB): each row block then streams its row's spectrum, (M + 1) complex
// values, from the MALL or HBM, priced at 4 TB/s for the whole GPU (half the
// HBM peak: the reads compete with the rho writes).  The autotuner's
// synthetic lanes (8 rows) never reach this, so the term is added on top.
inline double spectrumPenalty(int precision, int logM, long long rows, double rowBlocks)
{
    constexpr double kL2Budget = 3.0 * (1 << 20);
    const double specBytes = static_cast<double>((1LL << logM) + 1) * 2.0 * (precision == MF_PRECISION_FLOAT ? 4 : 8);
    if (static_cast<double>(rows) * specBytes <= kL2Budget) { return 0.0; }
    return rowBlocks * specBytes / 4.0e12 * 1.0e6;  // microseconds
}

// Predicted cost of a lane for an FFT size and a variant (0 = the default
// variant), or -1 if the plan is impossible: blocks x the per-block part
// plus row blocks x the per-row part.
inline double laneCost(int precision, int logM, int variant, const LaneRows &rs)
{
    const long long N = 2LL << logM;
    const long long L = N - rs.nb() + 1;
    if (L < 2) { return -1; }  // even block starts need L >= 2
    const double blocks = std::ceil(static_cast<double>(rs.maxOut()) / static_cast<double>(L));
    double shared = 0, solo = 0;
    laneRowBlocks(rs, L, shared, solo);
    const double penalty = spectrumPenalty(precision, logM, static_cast<long long>(rs.Lt.size()), shared + solo);
    std::lock_guard<std::mutex> lk(costMutex());
    const CostTable &t = costTable(precision);
    if (t.measured) {
        int v = variant;
        if (v == 0) {
            double best = -1;
            for (int cand = 1; cand <= 2; ++cand) {
                if (t.a[logM][cand] <= 0) { continue; }
                const double c = blocks * t.a[logM][cand] + t.bShared[logM][cand] * shared + t.bSolo[logM][cand] * solo;
                if (best < 0 || c < best) { best = c; v = cand; }
            }
            if (best >= 0) { return best * blockRisk(logM) + penalty; }
        } else if (t.a[logM][v] > 0) {
            return (blocks * t.a[logM][v] + t.bShared[logM][v] * shared + t.bSolo[logM][v] * solo) * blockRisk(logM) +
                   penalty;
        }
    }
    // static model: per-block part and per-row parts at the same scale
    const double unit = staticBlockCost(logM, RowMix{0, 0}, precision);
    return unit * (blocks + 1.15 * shared + 1.84 * solo) * blockRisk(logM) + penalty;
}

// Correlation kernel variant of one lane: a requested variant if it exists
// for the size; otherwise, with a measured cost table, the variant of least
// predicted per-block cost for the lane's rows; otherwise the default.  It
// depends on the lane alone, so a lane computes the same numbers whatever
// it is launched with.
inline int laneVariant(int precision, int logM, int requested, RowMix mix)
{
    if (requested != 0 && variantAvailable(precision, logM, requested)) { return requested; }
    std::lock_guard<std::mutex> lk(costMutex());
    const CostTable &t = costTable(precision);
    if (t.measured) {
        int bestV = 0;
        double best = -1;
        for (int v = 1; v <= 2; ++v) {
            if (!variantAvailable(precision, logM, v) || t.a[logM][v] <= 0) { continue; }
            const double c = t.a[logM][v] + t.bShared[logM][v] * mix.shared + t.bSolo[logM][v] * mix.solo;
            if (best < 0 || c < best) { best = c; bestV = v; }
        }
        if (bestV != 0) { return bestV; }
    }
    return staticVariant(precision, logM);
}

// FFT size (as logM, nfft = 2^(logM+1)) of a lane whose longest template
// has nb samples and which produces nOut outputs for the rows in `mix`.  A
// positive userNfft forces the size.
inline int chooseLogM(int precision, int userNfft, const LaneRows &rs)
{
    const int nb = rs.nb();
    const int hiLog = maxLogM(precision);
    if (userNfft > 0) {
        int lg = 0;
        while ((1 << (lg + 1)) < userNfft) { ++lg; }
        if ((2 << lg) != userNfft || lg < mfk::kMinLogM || lg > hiLog) {
            fail(MF_ERR_UNSUPPORTED, "FFT length " + std::to_string(userNfft) +
                                         " is not a supported power of two in [" +
                                         std::to_string(2 << mfk::kMinLogM) + ", " +
                                         std::to_string(2 << hiLog) + "]");
        }
        if ((2LL << lg) - nb + 1 < 2) {
            fail(MF_ERR_INVALID_ARGUMENT, "FFT length " + std::to_string(userNfft) +
                                              " is shorter than the template length " + std::to_string(nb));
        }
        return lg;
    }
    int best = -1;
    double bestCost = 0;
    for (int lg = mfk::kMinLogM; lg <= hiLog; ++lg) {
        const long long N = 2LL << lg;
        if (N - nb + 1 < N / 4 && lg < hiLog) { continue; }  // keep >= 25% useful outputs
        const double c = laneCost(precision, lg, 0, rs);
        if (c < 0) { continue; }
        if (best < 0 || c < bestCost) { best = lg; bestCost = c; }
    }
    if (best < 0) {
        fail(MF_ERR_UNSUPPORTED, "template length " + std::to_string(nb) +
                                     " exceeds the largest supported FFT block (" + std::to_string(2 << hiLog) +
                                     " samples)");
    }
    return best;
}

// Block length of a lane: nfft - nb + 1, made even so that every block
// start stays 8-byte aligned for the paired stores of the epilogue.
inline int blockLength(int logM, int nb)
{
    int L = static_cast<int>((2LL << logM) - nb + 1);
    if (L > 1 && (L & 1)) { L -= 1; }
    return L;
}

// Length buckets of one channel's rows.  Rows (sorted by template length)
// are split into contiguous buckets; each bucket becomes a lane with its own
// FFT size and block length, so short templates are not forced onto the
// long templates' short blocks, while a shared bucket saves the forward FFT
// and block loads.  planBuckets finds the split of least predicted cost.
struct Bucket {
    int begin = 0, end = 0;  // [begin, end) of the sorted rows
    int nb = 0;              // longest template of the bucket
    int logM = 0;
    RowMix mix;
    double cost = 0;
};

inline LaneRows laneRows(const std::vector<int> &sortedLt, const std::vector<long long> &nOut, int begin, int end)
{
    LaneRows rs;
    rs.Lt.assign(sortedLt.begin() + begin, sortedLt.begin() + end);
    rs.nOut.assign(nOut.begin() + begin, nOut.begin() + end);
    return rs;
}

inline Bucket makeBucket(int precision, int userNfft, const std::vector<int> &sortedLt,
                         const std::vector<long long> &nOut, int begin, int end)
{
    Bucket b;
    b.begin = begin;
    b.end = end;
    b.nb = sortedLt[end - 1];
    const LaneRows rs = laneRows(sortedLt, nOut, begin, end);
    b.mix = rs.mix();
    b.logM = chooseLogM(precision, userNfft, rs);
    b.cost = laneCost(precision, b.logM, 0, rs);
    return b;
}

// sortedLt: the channel's template lengths in increasing order; nOut[i]:
// the outputs of row i of that order.
inline std::vector<Bucket> planBuckets(int precision, int userNfft, const std::vector<int> &sortedLt,
                                       const std::vector<long long> &nOut)
{
    const int n = static_cast<int>(sortedLt.size());
    if (userNfft > 0 || n <= 1) { return {makeBucket(precision, userNfft, sortedLt, nOut, 0, n)}; }
    // Optimal split into contiguous buckets by dynamic programming over the
    // boundaries between runs of equal length (a run stays in one bucket, so
    // its rows keep sharing a scale): best[J] = min over I < J and FFT size
    // of best[I] + cost(runs I..J-1), where a bucket's cost is its blocks
    // times the per-block part plus its rows' blocks times the per-row
    // parts, at the block length its longest template allows.
    std::vector<int> bnd{0};
    for (int i = 1; i < n; ++i) {
        if (sortedLt[i] != sortedLt[i - 1]) { bnd.push_back(i); }
    }
    bnd.push_back(n);
    const int R = static_cast<int>(bnd.size()) - 1;
    std::vector<char> shared(n, 0);
    for (int k = 0; k < R; ++k) {
        if (bnd[k + 1] - bnd[k] >= 4) { std::fill(shared.begin() + bnd[k], shared.begin() + bnd[k + 1], 1); }
    }
    CostTable t;
    {
        std::lock_guard<std::mutex> lk(costMutex());
        t = costTable(precision);
    }
    const int hiLog = maxLogM(precision);
    // per-block and per-row-block parts of an FFT size, best variant for
    // the given mix of row blocks
    auto cost = [&](int lg, double blocks, double sS, double sO, long long rows) {
        double best = -1;
        if (t.measured) {
            for (int v = 1; v <= 2; ++v) {
                if (t.a[lg][v] <= 0 || !variantAvailable(precision, lg, v)) { continue; }
                const double c = blocks * t.a[lg][v] + sS * t.bShared[lg][v] + sO * t.bSolo[lg][v];
                if (best < 0 || c < best) { best = c; }
            }
        }
        if (best < 0) {
            const double unit = staticBlockCost(lg, RowMix{0, 0}, precision);
            best = unit * (blocks + 1.15 * sS + 1.84 * sO);
        }
        return best * blockRisk(lg) + spectrumPenalty(precision, lg, rows, sS + sO);
    };
    std::vector<double> best(R + 1, -1);
    std::vector<int> fromRun(R + 1, -1), fromLog(R + 1, -1);
    best[0] = 0;
    for (int J = 1; J <= R; ++J) {
        const int j = bnd[J];
        const int nb = sortedLt[j - 1];
        for (int lg = mfk::kMinLogM; lg <= hiLog; ++lg) {
            const long long N = 2LL << lg;
            if (N - nb + 1 < 2) { continue; }
            if (N - nb + 1 < N / 4 && lg < hiLog) { continue; }  // keep >= 25% useful outputs
            const double L = static_cast<double>(blockLength(lg, nb));
            double sS = 0, sO = 0;
            long long maxOut = 0;
            for (int I = J - 1; I >= 0; --I) {
                for (int r = bnd[I]; r < bnd[I + 1]; ++r) {
                    const double rb = std::ceil(static_cast<double>(nOut[r]) / L);
                    (shared[r] ? sS : sO) += rb;
                    maxOut = std::max(maxOut, nOut[r]);
                }
                if (best[I] < 0) { continue; }
                const double c = best[I] + cost(lg, std::ceil(static_cast<double>(maxOut) / L), sS, sO, j - bnd[I]);
                if (best[J] < 0 || c < best[J]) {
                    best[J] = c;
                    fromRun[J] = I;
                    fromLog[J] = lg;
                }
            }
        }
        if (best[J] < 0) {
            fail(MF_ERR_UNSUPPORTED, "template length " + std::to_string(nb) +
                                         " exceeds the largest supported FFT block (" + std::to_string(2 << hiLog) +
                                         " samples)");
        }
    }
    std::vector<Bucket> bk;
    for (int J = R; J > 0; J = fromRun[J]) {
        Bucket b;
        b.begin = bnd[fromRun[J]];
        b.end = bnd[J];
        b.nb = sortedLt[b.end - 1];
        const LaneRows rs = laneRows(sortedLt, nOut, b.begin, b.end);
        b.mix = rs.mix();
        b.logM = fromLog[J];
        b.cost = laneCost(precision, b.logM, 0, rs);
        bk.push_back(b);
    }
    std::reverse(bk.begin(), bk.end());
    return bk;
}

// Work items of a set of lanes sharing one FFT size, in XCD-aware order.
// Workgroup i runs on XCD i mod 8, so the items are dealt from 8 bins in
// turn: while every bin has items, item i comes from bin i mod 8.
//  - Lanes are first kept whole in one bin each, so their template spectra
//    stay in one L2: lanes go to the bin with the fewest blocks, largest
//    first, and the bins hold equal numbers of items.  A lane of more than
//    half a bin's share is cut into block ranges first.
//  - A launch of many lanes (at least 4 per XCD) whose bins then differ in
//    work by more than 1 % is binned by work instead: whole lanes, or
//    pieces of at most a quarter of a bin's work, to the bin with the least
//    work, heaviest first; then the shorter bins are padded with empty
//    items (nRows 0, which the correlation kernels return from at once), so
//    all bins hold equal numbers of items and every item stays on its
//    lane's XCD.
//  - A launch of few lanes (under 4 per XCD) whose bins then differ in work
//    by more than 5 % is dealt by work instead.  A block weighs kBlockRows
//    row units for its own part (load, statistics, forward transform) plus
//    one per row it computes (rowsAt).  The blocks, heaviest first, go to
//    the bins in snake order (0..7, 7..0, ...), so the bins hold equal
//    numbers of items (to one) and about equal work.  Each XCD then reads
//    the spectra of every lane, which few lanes afford.
// rowsAt(i, k), when given, is the number of leading rows of lane
// laneIds[i] with outputs in its block k (the item's nRows); otherwise -1
// (every row).
inline void appendXcdItems(const std::vector<int> &laneIds, const std::vector<long long> &laneBlocks,
                           std::vector<mfk::Item> &items,
                           const std::function<int(int, long long)> &rowsAt = nullptr)
{
    constexpr int kXcd = 8;
    constexpr int kBlockRows = 6;  // measured per-block part over per-row part of a block: about 6
    const int nL = static_cast<int>(laneIds.size());
    std::vector<std::vector<mfk::Item>> bins(kXcd);
    struct Piece {
        int lane;
        long long k0, k1;
    };
    long long total = 0;
    for (long long nb : laneBlocks) { total += nb; }
    const long long cap = std::max<long long>(1, total / (2 * kXcd));
    std::vector<Piece> pieces;
    for (int i = 0; i < nL; ++i) {
        for (long long k0 = 0; k0 < laneBlocks[i]; k0 += cap) {
            pieces.push_back({i, k0, std::min(laneBlocks[i], k0 + cap)});
        }
    }
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const Piece &x, const Piece &y) { return x.k1 - x.k0 > y.k1 - y.k0; });
    std::vector<long long> load(kXcd, 0), work(kXcd, 0);
    for (const Piece &p : pieces) {
        const int b = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        for (long long k = p.k0; k < p.k1; ++k) {
            const int r = rowsAt ? rowsAt(p.lane, k) : -1;
            bins[b].push_back({laneIds[p.lane], static_cast<int>(k), r, 0});
            work[b] += kBlockRows + std::max(r, 0);
        }
        load[b] += p.k1 - p.k0;
    }
    long long workSum = 0, workMax = 0;
    for (long long w : work) {
        workSum += w;
        workMax = std::max(workMax, w);
    }
    if (rowsAt != nullptr && nL >= 4 * kXcd && workMax * kXcd > 1.01 * workSum) {
        for (auto &bin : bins) { bin.clear(); }
        struct WPiece {
            int lane;
            long long k0, k1;
            long long w;
        };
        const long long wcap = std::max<long long>(1, workSum / (4 * kXcd));
        std::vector<WPiece> wp;
        for (int i = 0; i < nL; ++i) {
            WPiece p{i, 0, 0, 0};
            for (long long k = 0; k < laneBlocks[i]; ++k) {
                const long long w = kBlockRows + std::max(rowsAt(i, k), 0);
                if (p.k1 > p.k0 && p.w + w > wcap) {
                    wp.push_back(p);
                    p = {i, k, k, 0};
                }
                p.k1 = k + 1;
                p.w += w;
            }
            if (p.k1 > p.k0) { wp.push_back(p); }
        }
        std::stable_sort(wp.begin(), wp.end(), [](const WPiece &x, const WPiece &y) { return x.w > y.w; });
        std::vector<long long> wl(kXcd, 0);
        for (const WPiece &p : wp) {
            const int b = static_cast<int>(std::min_element(wl.begin(), wl.end()) - wl.begin());
            for (long long k = p.k0; k < p.k1; ++k) {
                bins[b].push_back({laneIds[p.lane], static_cast<int>(k), rowsAt(p.lane, k), 0});
            }
            wl[b] += p.w;
        }
        size_t most = 0;
        for (const auto &bin : bins) { most = std::max(most, bin.size()); }
        for (auto &bin : bins) {
            while (bin.size() < most) { bin.push_back({laneIds[0], 0, 0, 0}); }
        }
    } else if (rowsAt != nullptr && nL < 4 * kXcd && workMax * kXcd > 1.05 * workSum) {
        std::vector<std::pair<int, mfk::Item>> all;
        for (auto &bin : bins) {
            for (const mfk::Item &it : bin) { all.push_back({kBlockRows + it.nRows, it}); }
            bin.clear();
        }
        std::stable_sort(all.begin(), all.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
        for (size_t q = 0; q < all.size(); ++q) {
            const int c = static_cast<int>(q % kXcd);
            bins[(q / kXcd) % 2 == 0 ? c : kXcd - 1 - c].push_back(all[q].second);
        }
    }
    std::vector<size_t> pos(kXcd, 0);
    bool any = true;
    while (any) {
        any = false;
        for (int b = 0; b < kXcd; ++b) {
            if (pos[b] < bins[b].size()) {
                items.push_back(bins[b][pos[b]++]);
                any = true;
            }
        }
    }
}

}  // namespace mfh

#endif

// Matched-filter kernels for gfx950 (MI355X, CDNA4, wave64).
//
// corr_kernel  -- the hot path.  One block = one FFT block of one channel
//   lane; it serves every template row on that channel:
//     1. load nfft samples (coalesced), FP64 block anchor (window mean),
//        FP64 16-sample tile sums of (x-a) and (x-a)^2 via wave shuffles;
//     2. forward real FFT of the anchored block (LDS Stockham), unpacked
//        into X' = 2 X (kept in LDS for all rows);
//     3. per template length: FP64 sliding window sums -> per-output
//        scale sqrt(Lt)/sqrt(Lt*S2 - S1^2), with ill-conditioned windows
//        flagged;
//     4. per row: spectral product with the resident template spectrum
//        (read from L2), inverse packing, inverse FFT, normalisation and
//        the tooBig rule in the epilogue, one coalesced store of rho;
//        flagged windows are recomputed by an exact FP64 two-pass Pearson
//        from the LDS copy of the samples.
//   Blocks are overlap-save: block b covers outputs [bL, bL+L) and reads
//   samples [bL, bL+nfft), so every output is complete inside one block
//   and no cross-block accumulation is needed.  Subtracting the anchor is
//   exact because the template is zero-mean.
//
// tmpl_kernel  -- template demean/normalise and spectrum, FP64.
// stack_kernel -- moveout-shifted weighted stack of one group.
// peak_kernel  -- threshold + local-maximum candidates.
// minmax_kernel -- per-block min/max for the dead-trace test.
#include "mf_device.hpp"
#include "mf_kernels.hpp"
#include <cfloat>
#include <type_traits>

using namespace mfdev;

namespace mfk {

namespace {

constexpr int NT = kThreads;

// Error model of the FFT numerator: |du| <= kNoiseK * eps * sqrt(log2 nfft)
// * max|T^| * rms(block).  A window is recomputed exactly when that bound,
// divided by the window's own norm, could exceed the precision's rho
// tolerance.  Measured against the FP64 oracle with the recomputation
// switched off (band-limited noise with injected events 2-8x the noise RMS
// times sqrt(Lt), templates of 200-2000 samples, nfft 2048/4096), the worst
// error was 1.47x the bound and 4.5e-7 absolute.  The FP32 tolerance of
// 2.5e-6 therefore keeps unflagged windows below ~4e-6, inside the 1e-5
// contract; FP64 correlation keeps 1e-12.
template <typename T>
struct Eps;
template <>
struct Eps<float> {
    static constexpr double value = 5.9604644775390625e-08;  // 2^-24
    static constexpr double rhoTol = 2.5e-6;
};
template <>
struct Eps<double> {
    static constexpr double value = 1.1102230246251565e-16;  // 2^-53
    static constexpr double rhoTol = 1.0e-12;
};
constexpr double kNoiseK = 4.0;

// Scale buffer encoding of a window: f > 0, rho = c f; -f (FP32 correlation:
// a window whose FFT numerator may be off by the noise bound while its FP64
// window sums hold): rho = (exact FP64 numerator) f; -infinity (the window
// sums lost digits, or any flag in FP64 correlation): the whole Pearson
// value is recomputed exactly.
template <typename T>
__device__ __forceinline__ T scale_exact()
{
    return -static_cast<T>(INFINITY);
}

// Exclusive prefix sums tp[q] = sum of ts[q'] for q' < q (q = 0..NTILE) of
// the 16-sample tile sums, computed by wave 0 with a shuffle scan (fixed
// order, deterministic).  The caller synchronises before reading tp.
template <int NTILE>
__device__ __forceinline__ void tile_prefix(const double *ts1, const double *ts2, double *tp1, double *tp2,
                                            int tid)
{
    static_assert(NTILE >= 64 && NTILE % 64 == 0, "tile count");
    constexpr int PER = NTILE / 64;
    if (tid >= 64) { return; }
    double l1[PER], l2[PER];
    double s1 = 0, s2 = 0;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        l1[p] = s1;
        l2[p] = s2;
        s1 += ts1[tid * PER + p];
        s2 += ts2[tid * PER + p];
    }
    double x1 = s1, x2 = s2;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y1 = __shfl_up(x1, o, 64);
        const double y2 = __shfl_up(x2, o, 64);
        if (tid >= o) {
            x1 += y1;
            x2 += y2;
        }
    }
    const double e1 = x1 - s1, e2 = x2 - s2;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        tp1[tid * PER + p] = e1 + l1[p];
        tp2[tid * PER + p] = e2 + l2[p];
    }
    if (tid == 63) {
        tp1[NTILE] = x1;
        tp2[NTILE] = x2;
    }
}

// Sums of d = x - a and d^2 over the window [lo, lo + Lt) of the block's
// samples (TILE-sample tiles): partial tiles sample by sample, whole tiles
// from the tile prefix sums (FP32 correlation) or tile by tile (FP64
// correlation, whose tighter tolerance keeps the summation order of the
// tiles).
template <typename T, int TILE>
__device__ __forceinline__ void window_sums(const T *raw, const double *ts1, const double *ts2, const double *tp1,
                                            const double *tp2, double a, int lo, int Lt, double &S1, double &S2)
{
    const int hi = lo + Lt;
    const int q0 = (lo + TILE - 1) / TILE, q1 = hi / TILE;
    S1 = 0;
    S2 = 0;
    if (q0 < q1) {
        for (int j = lo; j < TILE * q0; ++j) {
            const double d = static_cast<double>(raw[j]) - a;
            S1 += d;
            S2 += d * d;
        }
        if constexpr (sizeof(T) == 4) {
            S1 += tp1[q1] - tp1[q0];
            S2 += tp2[q1] - tp2[q0];
        } else {
            for (int q = q0; q < q1; ++q) {
                S1 += ts1[q];
                S2 += ts2[q];
            }
        }
        for (int j = TILE * q1; j < hi; ++j) {
            const double d = static_cast<double>(raw[j]) - a;
            S1 += d;
            S2 += d * d;
        }
    } else {
        for (int j = lo; j < hi; ++j) {
            const double d = static_cast<double>(raw[j]) - a;
            S1 += d;
            S2 += d * d;
        }
    }
}

// Per-output scale sqrt(Lt)/||x - mean|| for outputs [0, nValid) of this
// block, FP64 sliding sums restarted every C outputs; flagged outputs get
// -1.  Threads t = 0..nthr-1 take part, each a contiguous run of C (odd)
// outputs.  The FP32 scale is rsqrt(den^2) in FP32 plus one Newton step
// (relative error ~2^-23, far inside the FP32 contract); the FP64 scale is
// sqrt/division in FP64.  Returns this thread's number of flagged outputs.
// Outputs [first, first + nValid) of the block (sc[i - first] for output i).
// NANFLAG: flagged outputs get NaN instead (the epilogue then writes NaN
// for them, which the recomputation launch looks for), and the return value
// packs the flagged count with the cancelled count << 16.
template <typename T, int TILE, int UNROLL = 1, bool NANFLAG = false>
__device__ __forceinline__ int compute_scale(const T *raw, const double *ts1, const double *ts2,
                                             const double *tp1, const double *tp2, T *sc, double a, int nValid,
                                             int Lt, double noise, int t, int nthr, int first = 0)
{
    const double sqrtLt = sqrt(static_cast<double>(Lt));
    const double denThr = sqrtLt * noise;
    const double denThr2 = denThr * denThr;
    // odd run length: thread t's accesses start t*C apart, and an odd stride
    // spreads the threads of a wave over all 32 LDS banks
    const int C = ((nValid + nthr - 1) / nthr) | 1;
    const double canc = 2.0e-9 * (Lt + C) * Lt;
    const int i0 = first + t * C;
    const int i1 = min(i0 + C, first + nValid);
    int nflag = 0;
    if (i0 >= i1) { return 0; }
    sc -= first;
    double S1, S2;
    window_sums<T, TILE>(raw, ts1, ts2, tp1, tp2, a, i0, Lt, S1, S2);
    double S2max = S2;
    const float sqrtLtF = static_cast<float>(sqrtLt);
    // UNROLL > 1 (a wave's own long runs) issues the sample reads of
    // several steps together ahead of the serial window-sum chain
#pragma unroll UNROLL
    for (int i = i0; i < i1; ++i) {
        const double den2 = Lt * S2 - S1 * S1;
        const bool canc1 = !(den2 > canc * S2max);
        const bool noisy = !(den2 > denThr2);
        nflag += (canc1 || noisy) ? 1 : 0;
        T f;
        if constexpr (sizeof(T) == 4) {
            const float d2 = static_cast<float>(den2);
            float y = __builtin_amdgcn_rsqf(d2);
            y = y * (1.5f - 0.5f * d2 * y * y);
            f = sqrtLtF * y;
        } else {
            f = static_cast<T>(sqrtLt / sqrt(den2));
        }
        // FP64 correlation recomputes every flagged window whole: its 1e-12
        // tolerance leaves no room for the window sums' own rounding
        if constexpr (NANFLAG) {
            sc[i] = (canc1 || noisy) ? __builtin_nanf("") : f;
            nflag += canc1 ? (1 << 16) : 0;
        } else {
            sc[i] = (canc1 || (sizeof(T) == 8 && noisy)) ? scale_exact<T>() : (noisy ? -f : f);
        }
        if (i + 1 < i1) {
            const double dn = static_cast<double>(raw[i + Lt]) - a;
            const double dd = static_cast<double>(raw[i]) - a;
            S1 += dn - dd;
            S2 += dn * dn - dd * dd;
            S2max = fmax(S2max, S2);
        }
    }
    return nflag;
}

// The FP32 scales of one row's outputs [first, first + nValid) computed by
// one wave, for rows that have no shared scale.  Lane t runs NCH
// interleaved chains (independent work in every step), chain k over the
// outputs [first + (64 k + t) Cc, ... + Cc) with Cc odd: FP64 window sums
// from the tile prefix at the chain start, then one sliding update per
// output, S1 += xn - xo and S2 += (xn - xo)(xn + xo - 2a).  A window is
// flagged (scale -1, recomputed exactly later) when den^2 = Lt S2 - S1^2 is
// not above max(canc S2span, denThr^2): S2span, the energy of the tiles the
// chain's windows cover, bounds every window sum of the chain, so the
// cancellation test of compute_scale holds with one compare per output.
// Returns this lane's number of flagged outputs.
template <int TILE, int NCH, bool NANFLAG = false>
__device__ __forceinline__ int wave_row_scale(const float *raw, const double *tp1, const double *tp2, float *sc,
                                              double a, int nValid, int Lt, double noise, int lane, int first)
{
    const double sqrtLt = sqrt(static_cast<double>(Lt));
    const double denThr2 = sqrtLt * noise * sqrtLt * noise;
    const float sqrtLtF = static_cast<float>(sqrtLt);
    const int Cc = ((nValid + 64 * NCH - 1) / (64 * NCH)) | 1;
    const double canc = 2.0e-9 * (Lt + Cc) * Lt;
    const int end = first + nValid;
    const double a2 = a + a;
    double S1[NCH], S2[NCH], thr[NCH], cancThr[NCH];
    int i0[NCH];
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
        i0[k] = first + (64 * k + lane) * Cc;
        S1[k] = S2[k] = 0;
        thr[k] = denThr2;
        cancThr[k] = 0;
        if (i0[k] < end) {
            window_sums<float, TILE>(raw, nullptr, nullptr, tp1, tp2, a, i0[k], Lt, S1[k], S2[k]);
            const int last = min(i0[k] + Cc, end) - 1;
            const int qa = i0[k] / TILE, qb = (last + Lt + TILE - 1) / TILE;
            cancThr[k] = canc * (tp2[qb] - tp2[qa]);
            thr[k] = fmax(cancThr[k], denThr2);
        }
    }
    sc -= first;
    int nflag = 0;
    for (int j = 0; j < Cc; ++j) {
#pragma unroll
        for (int k = 0; k < NCH; ++k) {
            const int i = i0[k] + j;
            if (i < end) {
                const double den2 = Lt * S2[k] - S1[k] * S1[k];
                const bool flag = !(den2 > thr[k]);
                nflag += flag ? 1 : 0;
                const float d2 = static_cast<float>(den2);
                const float y = __builtin_amdgcn_rsqf(d2);
                const float h = fmaf(-0.5f * d2, y * y, 1.5f);
                const float f = (sqrtLtF * y) * h;
                // noise flag: the scale holds, the numerator is recomputed
                const bool canc1 = !(den2 > cancThr[k]);
                if constexpr (NANFLAG) {
                    sc[i] = flag ? __builtin_nanf("") : f;
                    nflag += canc1 ? (1 << 16) : 0;
                } else {
                    sc[i] = canc1 ? scale_exact<float>() : (flag ? -f : f);
                }
                const double xn = static_cast<double>(raw[i + Lt]);
                const double xo = static_cast<double>(raw[i]);
                const double e = xn - xo;
                S1[k] += e;
                S2[k] = fma(e, (xn + xo) - a2, S2[k]);
            }
        }
    }
    return nflag;
}

// Flagged windows per 64 outputs from which a wave recomputes them one per
// lane (serial FP64 loops, no cross-lane reductions) instead of one per
// wave (strided sums and wave reductions, latency-bound when repeated).
constexpr int kDenseFlags = 16;

// Exact FP64 two-pass Pearson of the window at wi, by one lane.
template <typename T>
__device__ __forceinline__ double exact_window(const T *raw, int wi, gptr<const double> tmpl, int Lt)
{
    double sm = 0;
    for (int mm = 0; mm < Lt; ++mm) { sm += static_cast<double>(raw[wi + mm]); }
    const double mean = sm / Lt;
    double e2 = 0, dn = 0;
    for (int mm = 0; mm < Lt; ++mm) {
        const double d = static_cast<double>(raw[wi + mm]) - mean;
        e2 += d * d;
        dn += tmpl[mm] * d;
    }
    double rho = dn / fmax(DBL_EPSILON, sqrt(e2));
    if (fabs(rho) > 1.1) { rho = 0; }
    return rho;
}

template <typename T, int M>
__global__ __launch_bounds__(NT) void corr_kernel(const Item *__restrict__ items,
                                                  const Seg *__restrict__ segs,
                                                  const Row *__restrict__ rows,
                                                  const Cx<T> *__restrict__ twloG,
                                                  const Cx<T> *__restrict__ twhiG)
{
    constexpr int N = 2 * M;
    constexpr int NTILE = N / 16;
    constexpr int SPT = N / NT;
    constexpr int LOGN = Log2<N>::value;
    static_assert(N % NT == 0, "block size");
    __shared__ T raw[N];
    __shared__ Cx<T> Xs[M + 1];
    __shared__ Cx<T> Ws[M];
    __shared__ T sc[N];
    __shared__ double ts1[NTILE];
    __shared__ double ts2[NTILE];
    __shared__ double tp1[NTILE + 1];
    __shared__ double tp2[NTILE + 1];
    __shared__ Cx<T> twlo[64];
    __shared__ Cx<T> twhi[N / 64];
    __shared__ double red[NT / 64];

    const Item item = items[blockIdx.x];
    if (item.nRows == 0) { return; }  // an empty item that keeps the XCD order
    const Seg sg = segs[item.seg];
    const int tid = threadIdx.x;
    const long long base = static_cast<long long>(item.block) * sg.L;
    const int nValidSeg = static_cast<int>(min(static_cast<long long>(sg.L), sg.nOut - base));
    // the rows with outputs in this block (a prefix of the lane's rows)
    const int rowEnd = item.nRows < 0 ? sg.rowEnd : min(sg.rowEnd, sg.rowBegin + item.nRows);

    if (sg.dead) {
        for (int r = sg.rowBegin; r < rowEnd; ++r) {
            const Row row = rows[r];
            const int nv = static_cast<int>(min(static_cast<long long>(sg.L), row.nOut - base));
            T *out = static_cast<T *>(row.out) + base;
            for (int i = tid; i < nv; i += NT) { out[i] = T(0); }
        }
        return;
    }

    for (int i = tid; i < 64; i += NT) { twlo[i] = twloG[i]; }
    for (int i = tid; i < N / 64; i += NT) { twhi[i] = twhiG[i]; }
    const Twiddle<T> tw{twlo, twhi};

    // 1. load the block, anchor, tile sums, energy
    const T *x = static_cast<const T *>(sg.x);
    const int nSamp = static_cast<int>(min(static_cast<long long>(N), sg.n - base));
    T v[SPT];
    double s = 0;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int j = tid + q * NT;
        const long long g = base + j;
        const T val = (j < nSamp) ? x[g] : T(0);
        v[q] = val;
        raw[j] = val;
        s += static_cast<double>(val);
    }
    const double a = block_sum<NT>(s, red) / nSamp;
    const T aT = static_cast<T>(a);
    double e = 0;
    T *wf = reinterpret_cast<T *>(Ws);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int j = tid + q * NT;
        const bool valid = j < nSamp;
        const double d = valid ? static_cast<double>(v[q]) - a : 0.0;
        double t1 = d, t2 = d * d;
        e += t2;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            t1 += __shfl_xor(t1, o, 64);
            t2 += __shfl_xor(t2, o, 64);
        }
        if ((tid & 15) == 0) {
            ts1[j >> 4] = t1;
            ts2[j >> 4] = t2;
        }
        wf[j] = valid ? static_cast<T>(v[q] - aT) : T(0);
    }
    const double energy = block_sum<NT>(e, red);  // also publishes raw/ts/Ws
    tile_prefix<NTILE>(ts1, ts2, tp1, tp2, tid);   // read after the FFT's barriers

    // 2. forward FFT of the packed block, unpack to X' = 2X
    fft_lds<T, M, NT, false>(Ws, tw);
    for (int k = tid; k <= M / 2; k += NT) {
        const Cx<T> zk = Ws[k];
        const Cx<T> zm = cconj(Ws[(M - k) & (M - 1)]);
        const Cx<T> E = cadd(zk, zm);
        const Cx<T> O = cmulmi(csub(zk, zm));
        const Cx<T> WO = cmul(tw(k), O);
        Xs[k] = cadd(E, WO);
        Xs[M - k] = cconj(csub(E, WO));
    }
    __syncthreads();

    const double rmsBlk = sqrt(energy / nSamp);
    const double noise = kNoiseK / Eps<T>::rhoTol * Eps<T>::value * sqrt(static_cast<double>(LOGN)) *
                         static_cast<double>(sg.hmax) * rmsBlk;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar row loads and branches
    const int lane = tid & 63;
    int curLt = -1;
    for (int r = sg.rowBegin; r < rowEnd; ++r) {
        const Row row = rows[r];
        const int nvRow = static_cast<int>(min(static_cast<long long>(sg.L), row.nOut - base));
        if (nvRow <= 0) { continue; }
        const int Lt = row.Lt;
        if (Lt != curLt) {
            // 3. per-output scale for this template length
            curLt = Lt;
            compute_scale<T, 16>(raw, ts1, ts2, tp1, tp2, sc, a, nValidSeg, Lt, noise, tid, NT);
            __syncthreads();
        }
        // 4a. spectral product + inverse packing into Ws
        const Cx<T> *H = static_cast<const Cx<T> *>(row.H);
        for (int k = tid; k <= M / 2; k += NT) {
            const Cx<T> yk = cmul(Xs[k], H[k]);
            const Cx<T> ym = cmul(Xs[M - k], H[M - k]);
            const Cx<T> S = cadd(yk, cconj(ym));
            const Cx<T> D = csub(yk, cconj(ym));
            const Cx<T> P = cmul(cconj(tw(k)), D);
            Ws[k] = cadd(S, cmuli(P));
            if (k > 0) { Ws[M - k] = cadd(cconj(S), cmuli(cconj(P))); }
        }
        __syncthreads();
        // 4b. inverse FFT: Ws holds the correlation, interleaved = natural order
        fft_lds<T, M, NT, true>(Ws, tw);
        // 4c. normalisation epilogue
        T *out = static_cast<T *>(row.out) + base;
        const T *c = reinterpret_cast<const T *>(Ws);
        for (int i = tid; i < nvRow; i += NT) {
            const T f = sc[i];
            if (f >= T(0)) {
                T rho = c[i] * f;
                if (fabs(rho) > T(1.1)) { rho = T(0); }
                out[i] = rho;
            }
        }
        // 4d. exact FP64 two-pass Pearson for flagged windows
        {
            constexpr int NW = NT / 64;
            const int slice = (nvRow + NW - 1) / NW;
            const int s0 = wave * slice;
            const int s1 = min(s0 + slice, nvRow);
            unsigned long long nFlag = 0, nCanc = 0;
            for (int w0 = s0; w0 < s1; w0 += 64) {
                const int i = w0 + lane;
                const bool fl = (i < s1) && (sc[i] < T(0));
                unsigned long long m = __ballot(fl);
                nFlag += __popcll(m);
                nCanc += __popcll(__ballot(fl && isinf(sc[i])));
                if (__popcll(m) >= kDenseFlags) {
                    if (fl) { out[i] = static_cast<T>(exact_window<T>(raw, i, to_global(row.tmpl), Lt)); }
                    continue;
                }
                while (m != 0ull) {
                    const int b = __ffsll(static_cast<long long>(m)) - 1;
                    m &= m - 1ull;
                    const int wi = w0 + b;
                    double sm = 0;
                    for (int mm = lane; mm < Lt; mm += 64) { sm += static_cast<double>(raw[wi + mm]); }
                    sm = wave_sum(sm);
                    const double mean = sm / Lt;
                    double e2 = 0, dn = 0;
                    for (int mm = lane; mm < Lt; mm += 64) {
                        const double d = static_cast<double>(raw[wi + mm]) - mean;
                        e2 += d * d;
                        dn += row.tmpl[mm] * d;
                    }
                    e2 = wave_sum(e2);
                    dn = wave_sum(dn);
                    double rho = dn / fmax(DBL_EPSILON, sqrt(e2));
                    if (fabs(rho) > 1.1) { rho = 0; }
                    if (lane == 0) { out[wi] = static_cast<T>(rho); }
                }
            }
            if (sg.flagged != nullptr && nFlag != 0ull && lane == 0) {
                atomicAdd(sg.flagged, nFlag);
                if (nCanc != 0ull) { atomicAdd(sg.flagged + 1, nCanc); }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// corr2_kernel: same contract as corr_kernel, built from wave-level FFTs.
// Wave 0 computes the forward FFT while waves 1-3 build the normalisation
// scale; then every wave takes whole template rows (product with the
// template spectrum straight into pass-1 registers, three register passes,
// epilogue straight from pass-3 registers), so the row loop has no block
// barrier.  Rows are processed in groups of equal template length; the
// scale of a group is shared by its rows.
// ---------------------------------------------------------------------------

// In-place exact recomputation of the flagged outputs of one chunk of 64
// (mask m, output w0 + lane for this lane) of a row, sc[i] < 0 flagging
// output i (see scale_exact): a noise flag (-f) keeps the window's scale f
// and recomputes the numerator exactly in FP64 from the anchored samples,
// sum_m t[m] (x[i+m] - a) (the template has zero mean, so the window mean
// drops out of the numerator); -infinity recomputes the two-pass Pearson
// value.  Many flags: one lane per window; few: the wave on each window.
template <typename T>
__device__ __forceinline__ void fixup_chunk(const T *raw, const T *sc, gptr<const double> tmpl, int Lt, gptr<T> out,
                                            int w0, int nv, int lane, double a, unsigned long long m)
{
    const int i = w0 + lane;
    const T s = i < nv ? sc[i] : T(1);
    const bool fl = s < T(0);
    if (__popcll(m) >= kDenseFlags) {
        // many flagged windows (a gap, a loud stretch): one lane per window
        if (fl) {
            double rho;
            if (isinf(s)) {
                rho = exact_window<T>(raw, i, tmpl, Lt);
            } else {
                double dn = 0;
                for (int mm = 0; mm < Lt; ++mm) { dn = fma(tmpl[mm], static_cast<double>(raw[i + mm]) - a, dn); }
                rho = dn * static_cast<double>(-s);
                if (fabs(rho) > 1.1) { rho = 0; }
            }
            out[i] = static_cast<T>(rho);
        }
        return;
    }
    while (m != 0ull) {
        const int b = __ffsll(static_cast<long long>(m)) - 1;
        m &= m - 1ull;
        const int wi = w0 + b;
        const T sw = sc[wi];
        double rho;
        if (isinf(sw)) {
            double sm = 0;
            for (int mm = lane; mm < Lt; mm += 64) { sm += static_cast<double>(raw[wi + mm]); }
            sm = wave_sum(sm);
            const double mean = sm / Lt;
            double e2 = 0, dn = 0;
            for (int mm = lane; mm < Lt; mm += 64) {
                const double d = static_cast<double>(raw[wi + mm]) - mean;
                e2 += d * d;
                dn += tmpl[mm] * d;
            }
            e2 = wave_sum(e2);
            dn = wave_sum(dn);
            rho = dn / fmax(DBL_EPSILON, sqrt(e2));
        } else {
            double dn = 0;
            for (int mm = lane; mm < Lt; mm += 64) { dn = fma(tmpl[mm], static_cast<double>(raw[wi + mm]) - a, dn); }
            dn = wave_sum(dn);
            rho = dn * static_cast<double>(-sw);
        }
        if (fabs(rho) > 1.1) { rho = 0; }
        if (lane == 0) { out[wi] = static_cast<T>(rho); }
    }
}

// Flagged outputs [first, nv) of one row (one wave), sc[i - first] < 0
// flagging output i (see scale_exact), recomputed in place chunk by chunk;
// their count goes to the window statistics.
template <typename T>
__device__ __forceinline__ void fixup_row_wave(const T *raw, const T *sc, gptr<const double> tmpl, int Lt,
                                               gptr<T> out, int nv, int lane, double a, int first,
                                               unsigned long long *flagged)
{
    sc -= first;
    unsigned nFlag = 0, nCanc = 0;
    for (int w0 = first; w0 < nv; w0 += 64) {
        const int i = w0 + lane;
        const T s = i < nv ? sc[i] : T(1);
        const bool fl = s < T(0);
        const unsigned long long m = __ballot(fl);
        nFlag += __popcll(m);
        nCanc += __popcll(__ballot(fl && isinf(s)));
        if (m != 0ull) { fixup_chunk<T>(raw, sc, tmpl, Lt, out, w0, nv, lane, a, m); }
    }
    if (flagged != nullptr && nFlag != 0u && lane == 0) {
        atomicAdd(flagged, static_cast<unsigned long long>(nFlag));
        if (nCanc != 0u) { atomicAdd(flagged + 1, static_cast<unsigned long long>(nCanc)); }
    }
}

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

// Shape of corr2_kernel<T, M>:
//  - SHARED: a block-wide scale buffer for groups of rows of one template
//    length (FP32 up to M = 1024, FP64 up to M = 512);
//  - HALF (the larger blocks): every row computes its own scale, the
//    per-wave exchange buffers are half size, the template spectrum of the
//    next row is not prefetched into registers and the pass-3 twiddles come
//    from the table.  That fits two blocks per CU in LDS and registers
//    where the full layout fitted one;
//  - TILE: samples per FP64 prefix-sum tile (FP32: 4, or 8 at M = 2048; FP64:
//    16, whose windows sum whole tiles);
//  - BLOCKS per CU that LDS and registers allow; the register budget is
//    capped to match.
template <typename T, int M>
struct Corr2Shape {
    static constexpr bool SHARED = corr2_shared_scale(static_cast<int>(sizeof(T)), M);
    static constexpr bool HALF = !SHARED;
    static constexpr int TILE = sizeof(T) == 4 ? (M >= 2048 ? 8 : 4) : 16;
    static constexpr int BLOCKS = sizeof(T) * M <= 2048 ? 3 : 2;
};

template <typename T, int M, bool QUEUE>
__global__ __launch_bounds__(NT, (Corr2Shape<T, M>::BLOCKS)) void corr2_kernel(const Item *__restrict__ items,
                                                   const Seg *__restrict__ segs,
                                                   const Row *__restrict__ rows,
                                                   const Cx<T> *__restrict__ twloG,
                                                   const Cx<T> *__restrict__ twhiG, FixQueue queue)
{
    using Shape = Corr2Shape<T, M>;
    constexpr int N = 2 * M;
    constexpr int TILE = Shape::TILE;
    constexpr bool SHARED = Shape::SHARED;
    constexpr bool HALF = Shape::HALF;
    constexpr int NTILE = N / TILE;
    constexpr int SPT = N / NT;
    constexpr int LOGN = Log2<N>::value;
    constexpr int Q = M / 64;
    // (k, M-k) pair steps per lane for k < M/2; the self-paired bin k = M/2
    // has a closed form (its twiddle is exactly -i) and is done apart
    constexpr int KP = M / 2 / 64;
    constexpr int NWAVE = NT / 64;
    constexpr int XH = XHalf<M>::size;
    constexpr int XW = HALF ? XH : XBuf<M>::size;  // exchange buffer of one wave
    using P = WavePlan<M>;
    static_assert(N % NT == 0 && NWAVE == 4, "block shape");
    static_assert(2 * XH == XBuf<M>::size, "two half buffers make one full buffer");
    __shared__ T raw[N];
    __shared__ Cx<T> Xs[M + 1];
    __shared__ __align__(16) T sc[SHARED ? N : 1];
    // one exchange buffer per wave (half size with HALF; then the first two
    // together are the full buffer of the forward transform and of the
    // packed input)
    __shared__ Cx<T> xb[NWAVE * XW];
    // FP64 tile prefix sums; FP64 correlation also keeps the tile sums (FP32
    // builds the prefix in place over the sums)
    __shared__ double tp1[NTILE + 1];
    __shared__ double tp2[NTILE + 1];
    __shared__ double tsd1[sizeof(T) == 8 ? NTILE : 1];
    __shared__ double tsd2[sizeof(T) == 8 ? NTILE : 1];
    __shared__ Cx<T> twlo[64];
    __shared__ Cx<T> twhi[N / 64];
    __shared__ double red[NWAVE];
    double *ts1 = sizeof(T) == 8 ? tsd1 : tp1;
    double *ts2 = sizeof(T) == 8 ? tsd2 : tp2;

    const Item item = items[blockIdx.x];
    if (item.nRows == 0) { return; }  // an empty item that keeps the XCD order
    const Seg sg = segs[item.seg];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar row loads and branches
    const int lane = tid & 63;
    const long long base = static_cast<long long>(item.block) * sg.L;
    const int nValidSeg = static_cast<int>(min(static_cast<long long>(sg.L), sg.nOut - base));
    // the rows with outputs in this block (a prefix of the lane's rows)
    const int rowEnd = item.nRows < 0 ? sg.rowEnd : min(sg.rowEnd, sg.rowBegin + item.nRows);

    if (sg.dead) {
        for (int r = sg.rowBegin; r < rowEnd; ++r) {
            const Row row = rows[r];
            const int nv = static_cast<int>(min(static_cast<long long>(sg.L), row.nOut - base));
            gptr<T> out = to_global(static_cast<T *>(row.out)) + base;
            for (int i = tid; i < nv; i += NT) { out[i] = T(0); }
        }
        return;
    }
    for (int i = tid; i < 64; i += NT) { twlo[i] = twloG[i]; }
    for (int i = tid; i < N / 64; i += NT) { twhi[i] = twhiG[i]; }
    const Twiddle<T> tw{twlo, twhi};

    // A. load, anchor, tile sums, energy; packed anchored input -> xb (full)
    const gptr<const T> x = to_global(static_cast<const T *>(sg.x));
    const int nSamp = static_cast<int>(min(static_cast<long long>(N), sg.n - base));
    T v0[SPT];
    double s = 0;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int j = tid + q * NT;
        const T val = (j < nSamp) ? x[base + j] : T(0);
        v0[q] = val;
        raw[j] = val;
        s += static_cast<double>(val);
    }
    const double a = uniform(block_sum<NT>(s, red) / nSamp);  // scalar registers: used by every row
    const T aT = static_cast<T>(a);
    double e = 0;
    T *xin = reinterpret_cast<T *>(xb);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int j = tid + q * NT;
        const bool valid = j < nSamp;
        const double d = valid ? static_cast<double>(v0[q]) - a : 0.0;
        double t1 = d, t2 = d * d;
        e += t2;
#pragma unroll
        for (int o = TILE / 2; o > 0; o >>= 1) {
            t1 += __shfl_xor(t1, o, 64);
            t2 += __shfl_xor(t2, o, 64);
        }
        if ((tid & (TILE - 1)) == 0) {
            ts1[j / TILE] = t1;
            ts2[j / TILE] = t2;
        }
        xin[2 * xpad(j >> 1) + (j & 1)] = valid ? static_cast<T>(v0[q] - aT) : T(0);
    }
    const double energy = block_sum<NT>(e, red);
    tile_prefix<NTILE>(ts1, ts2, tp1, tp2, tid);  // in place for FP32 (each lane reads its tiles first)
    __syncthreads();
    const double rmsBlk = sqrt(energy / nSamp);
    const double noise = uniform(kNoiseK / Eps<T>::rhoTol * Eps<T>::value * sqrt(static_cast<double>(LOGN)) *
                                 static_cast<double>(sg.hmax) * rmsBlk);
    WaveTwiddles<T, M, !HALF> wt;
    wt.init(lane, tw);
    Cx<T> v[Q];

    // B. forward FFT + unpack into Xs (wave 0, full buffer)
    auto forwardFft = [&]() {
        Cx<T> *buf = xb;
        wave_load<T, M, P::R1>(v, buf, lane);
        wave_lds_sync();
        wave_fft_regs<T, M, false>(v, buf, lane, wt);
        // natural-order store of pass 3 (point j + r M/R3, j = lane + 64 b)
        constexpr int B3 = M / (P::R3 * 64);
        Cx<T> *po = buf + xpad(lane);
#pragma unroll
        for (int b = 0; b < B3; ++b) {
#pragma unroll
            for (int r = 0; r < P::R3; ++r) { po[cpad(64 * b + r * (M / P::R3))] = v[b * P::R3 + r]; }
        }
        wave_lds_sync();
        // unpack pairs (k, M-k), k = lane + 64 t < M/2
        const Cx<T> *pm = buf + xpad(M - lane);
        const Cx<T> wlo = twlo[lane];
#pragma unroll
        for (int t = 0; t < KP; ++t) {
            const int k = lane + 64 * t;
            const Cx<T> zk = po[68 * t];
            const Cx<T> zm = cconj(k == 0 ? po[0] : pm[-68 * t]);
            const Cx<T> E = cadd(zk, zm);
            const Cx<T> O = cmulmi(csub(zk, zm));
            const Cx<T> WO = cmul(cmul(wlo, twhi[t]), O);
            Xs[k] = cadd(E, WO);
            Xs[M - k] = cconj(csub(E, WO));
        }
        // k = M/2: E = 2 Re z, O = 2 Im z, twiddle -i, so X' = 2 conj(z)
        if (lane == 0) {
            const Cx<T> z = buf[xpad(M / 2)];
            Xs[M / 2] = Cx<T>{z.x + z.x, -(z.y + z.y)};
        }
        wave_lds_sync();
    };

    // C. one template row on this wave, in two halves.  startRow: spectral
    // product + inverse packing and the three register passes (the result
    // stays in v).  finishRow: the row's scale when it has no shared one,
    // then the normalisation epilogue straight from the pass-3 registers.
    // A wave's row sequence runs start(r0), then for each next row r:
    // finish(previous), start(r), and finally finish(last).  start(r)
    // loads the spectrum of the row after r once its product is done, so
    // the load is in flight through r's transform and epilogue; the
    // epilogue issues the same (range-checked) stores on every path, so
    // the wait for that spectrum never waits for those stores.
    Cx<T> hk[KP], hm[KP];  // template spectrum pairs (H[k], H[M-k]), k < M/2
    Cx<T> hmid;            // H[M/2]
    auto loadH = [&](const void *Hp) {
        const gptr<const Cx<T>> H = to_global(static_cast<const Cx<T> *>(Hp));
#pragma unroll
        for (int t = 0; t < KP; ++t) {
            const int k = lane + 64 * t;
            hk[t] = load_cx(H, k);
            hm[t] = load_cx(H, M - k);
        }
        hmid = load_cx(H, M / 2);
    };
    auto rowValid = [&](const Row &row) {
        return static_cast<int>(max(0ll, min(static_cast<long long>(sg.L), row.nOut - base)));
    };
    constexpr bool PREFETCH = !HALF;
    auto startRow = [&](const Row &row, const void *nextH) {
        if (rowValid(row) == 0) {
            if (PREFETCH) { loadH(nextH); }
            return;
        }
        if (!PREFETCH) { loadH(row.H); }
        Cx<T> *buf = xb + wave * XW;
        if constexpr (!HALF) {
            // (k, M-k) pair per lane step (k = lane + 64 t; tw(k) = lo[lane] hi[t])
            const Cx<T> wlo = twlo[lane];
            Cx<T> *pk = buf + xpad(lane);
            Cx<T> *pm = buf + xpad(M - lane);
#pragma unroll
            for (int t = 0; t < KP; ++t) {
                const int k = lane + 64 * t;
                const Cx<T> yk = cmul(Xs[k], hk[t]);
                const Cx<T> ym = cmul(Xs[M - k], hm[t]);
                const Cx<T> S = cadd(yk, cconj(ym));
                const Cx<T> D = csub(yk, cconj(ym));
                const Cx<T> Pp = cmul(cconj(cmul(wlo, twhi[t])), D);
                pk[68 * t] = cadd(S, cmuli(Pp));
                if (k > 0) { pm[-68 * t] = cadd(cconj(S), cmuli(cconj(Pp))); }
            }
            // k = M/2 (y = X H): S = 2 Re y, D = 2i Im y, twiddle -i, so Z = 2 conj(y)
            if (lane == 0) {
                const Cx<T> y = cmul(Xs[M / 2], hmid);
                buf[xpad(M / 2)] = Cx<T>{y.x + y.x, -(y.y + y.y)};
            }
            // the next row's spectrum is in flight through this row's
            // transform and epilogue
            loadH(nextH);
            wave_lds_sync();
            wave_load<T, M, P::R1>(v, buf, lane);
            wave_lds_sync();
            wave_fft_regs<T, M, true>(v, buf, lane, wt);
            return;
        }
        // (k, M-k) pair per lane step (k = lane + 64 t; tw(k) = lo[lane] hi[t]).
        // Points k < M/2 are the first half of the pass-1 input, M - k and
        // M/2 the second: the first half goes out and comes back in while
        // the second waits in registers.
        const Cx<T> wlo = twlo[lane];
        Cx<T> *pk = buf + xpad(lane);
        Cx<T> *pm = buf - cpad(M / 2) + xpad(M - lane);
        Cx<T> zm[KP];
#pragma unroll
        for (int t = 0; t < KP; ++t) {
            const int k = lane + 64 * t;
            const Cx<T> yk = cmul(Xs[k], hk[t]);
            const Cx<T> ym = cmul(Xs[M - k], hm[t]);
            const Cx<T> S = cadd(yk, cconj(ym));
            const Cx<T> D = csub(yk, cconj(ym));
            const Cx<T> Pp = cmul(cconj(cmul(wlo, twhi[t])), D);
            pk[68 * t] = cadd(S, cmuli(Pp));
            zm[t] = cadd(cconj(S), cmuli(cconj(Pp)));
        }
        // k = M/2 (y = X H): S = 2 Re y, D = 2i Im y, twiddle -i, so Z = 2 conj(y)
        const Cx<T> ymid = cmul(Xs[M / 2], hmid);
        wave_lds_sync();
        wave_load_half<T, M, P::R1, 0>(v, buf, lane);
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < KP; ++t) {
            if (lane + 64 * t > 0) { pm[-68 * t] = zm[t]; }
        }
        if (lane == 0) { buf[0] = Cx<T>{ymid.x + ymid.x, -(ymid.y + ymid.y)}; }  // point M/2 of half 1
        wave_lds_sync();
        wave_load_half<T, M, P::R1, 1>(v, buf, lane);
        wave_lds_sync();
        wave_fft_regs_half<T, M, true>(v, buf, lane, wt);
    };
    // rho = c * scale for the pass-3 outputs in half H of the block (the
    // lane's points j + r M/R3 with r in [H R3/2, (H+1) R3/2): outputs
    // [H N/2, (H+1) N/2)), paired stores; scl[i - first] is output i's
    // scale.  Flagged windows (scale -1) are overwritten by the exact
    // fix-up; for the others |rho| <= 1 + O(eps), so the reference's tooBig
    // rule (|rho| > 1.1 -> 0) cannot fire and is not evaluated.  The scale
    // reads stay inside the scale buffer for every position.  The store
    // resource ends at the last whole pair of valid outputs, so invalid
    // pairs are dropped; an odd last output is stored apart.
    using V2 = typename Vec2<T>::type;
    constexpr int B3 = M / (P::R3 * 64);
    auto epilogueHalf = [&](auto hTag, const T *scl, int first, T *outRow, int nvRow) {
        constexpr int H = decltype(hTag)::value;
        constexpr int R0 = H * (P::R3 / 2), R1e = (H + 1) * (P::R3 / 2);
        constexpr int NQH = B3 * (P::R3 / 2);
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(outRow, static_cast<int>(sizeof(T)) * (nvRow & ~1));
        const T *sp = scl - first + 2 * lane;
        V2 f[NQH];
#pragma unroll
        for (int b = 0; b < B3; ++b) {
#pragma unroll
            for (int r = R0; r < R1e; ++r) {
                f[b * (P::R3 / 2) + r - R0] = *reinterpret_cast<const V2 *>(sp + 2 * (64 * b + r * (M / P::R3)));
            }
        }
#pragma unroll
        for (int b = 0; b < B3; ++b) {
#pragma unroll
            for (int r = R0; r < R1e; ++r) {
                const int q = b * P::R3 + r;
                const V2 fq = f[b * (P::R3 / 2) + r - R0];
                const int i0 = 2 * lane + 2 * (64 * b + r * (M / P::R3));
                buffer_store_pair(v[q].x * fq.x, v[q].y * fq.y, rs, static_cast<int>(sizeof(T)) * i0);
            }
        }
        if (nvRow & 1) {
            gptr<T> outG = to_global(outRow);
#pragma unroll
            for (int b = 0; b < B3; ++b) {
#pragma unroll
                for (int r = R0; r < R1e; ++r) {
                    const int q = b * P::R3 + r;
                    const int i0 = 2 * lane + 2 * (64 * b + r * (M / P::R3));
                    if (i0 == nvRow - 1) { outG[i0] = v[q].x * f[b * (P::R3 / 2) + r - R0].x; }
                }
            }
        }
    };
    using Half0 = std::integral_constant<int, 0>;
    using Half1 = std::integral_constant<int, 1>;
    // sharedScale is the row's block-shared scale, or null: the wave then
    // computes the row's scale itself, half a block at a time, into its own
    // exchange buffer, which is free once the inverse FFT is in registers.
    // QUEUE: a row block with flagged outputs (written as NaN by the
    // epilogue) is recorded once for the recomputation launch, and the
    // window statistics are added (flag counts packed as flagged + cancelled
    // << 16; a shared scale's counts cover the block's longest row)
    auto recordRow = [&](int rowIdx, int first, int n) {
        if (lane == 0) {
            const unsigned long long k = atomicAdd(queue.overCount, 1ull);
            gptr<int> rec = to_global(queue.over) + 4 * k;
            rec[0] = rowIdx;
            rec[1] = static_cast<int>(base + first);
            rec[2] = n;
            rec[3] = 0;
        }
    };
    auto countFlags = [&](unsigned long long nf, unsigned long long nc) {
        if (sg.flagged != nullptr) {
            atomicAdd(sg.flagged, nf);
            if (nc != 0ull) { atomicAdd(sg.flagged + 1, nc); }
        }
    };
    auto finishRow = [&](const Row &row, int rowIdx, const T *sharedScale, bool anyFlagShared, double flagSum) {
        const int nvRow = rowValid(row);
        T *outRow = static_cast<T *>(row.out) + base;
        if (sharedScale != nullptr) {
            epilogueHalf(Half0{}, sharedScale, 0, outRow, nvRow);
            epilogueHalf(Half1{}, sharedScale, 0, outRow, nvRow);
            if (anyFlagShared && nvRow > 0) {
                if constexpr (QUEUE) {
                    recordRow(rowIdx, 0, nvRow);
                    if (lane == 0) {
                        const unsigned long long packed = static_cast<unsigned long long>(flagSum);
                        countFlags(packed & 0xffffull, packed >> 16);
                    }
                } else {
                    fixup_row_wave<T>(raw, sharedScale, to_global(row.tmpl), row.Lt, to_global(outRow), nvRow, lane,
                                      a, 0, sg.flagged);
                }
            }
            return;
        }
        T *own = reinterpret_cast<T *>(xb + wave * XW);
        const gptr<const double> tmplG = to_global(row.tmpl);
        // a row's own scale: interleaved chains in FP32, compute_scale in FP64
        auto ownScale = [&](int n, int first) {
            if constexpr (sizeof(T) == 4) {
                return wave_row_scale<TILE, 2, QUEUE>(raw, tp1, tp2, own, a, n, row.Lt, noise, lane, first);
            } else {
                return compute_scale<T, TILE, 4>(raw, ts1, ts2, tp1, tp2, own, a, n, row.Lt, noise, lane, 64, first);
            }
        };
        // after the epilogue of outputs [first, end) with this lane's flag
        // count nf: record / count (QUEUE) or recompute in place
        auto flagged = [&](int nf, bool anyFlag, int first, int end) {
            if constexpr (QUEUE) {
                if (anyFlag) {
                    // the wave's packed counts (each lane's low part is at
                    // most a few hundred, so the sum stays packed), then one
                    // atomic per counter: per-lane atomics on one address
                    // serialize
                    int sum = nf;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) { sum += __shfl_xor(sum, o, 64); }
                    if (lane == 0) {
                        countFlags(static_cast<unsigned long long>(sum & 0xffff), static_cast<unsigned long long>(sum >> 16));
                    }
                    recordRow(rowIdx, first, end - first);
                }
            } else {
                if (anyFlag) { fixup_row_wave<T>(raw, own, tmplG, row.Lt, to_global(outRow), end, lane, a, first, sg.flagged); }
            }
        };
        if constexpr (!HALF) {
            // the whole row's scale at once (the full buffer holds N values)
            int nf = 0;
            if (nvRow > 0) { nf = ownScale(nvRow, 0); }
            const bool anyFlag = __ballot(nf > 0) != 0ull;
            wave_lds_sync();
            epilogueHalf(Half0{}, own, 0, outRow, nvRow);
            epilogueHalf(Half1{}, own, 0, outRow, nvRow);
            flagged(nf, anyFlag, 0, nvRow);
            wave_lds_sync();  // the next row reuses buf
            return;
        }
        // half 0: outputs [0, N/2)
        {
            const int n0 = min(nvRow, N / 2);
            int nf = 0;
            if (n0 > 0) { nf = ownScale(n0, 0); }
            const bool anyFlag = __ballot(nf > 0) != 0ull;
            wave_lds_sync();
            epilogueHalf(Half0{}, own, 0, outRow, nvRow);
            flagged(nf, anyFlag, 0, n0);
            wave_lds_sync();  // the second half reuses the buffer
        }
        // half 1: outputs [N/2, N)
        {
            const int n1 = nvRow - N / 2;
            int nf = 0;
            if (n1 > 0) { nf = ownScale(n1, N / 2); }
            const bool anyFlag = __ballot(nf > 0) != 0ull;
            wave_lds_sync();
            epilogueHalf(Half1{}, own, N / 2, outRow, nvRow);
            flagged(nf, anyFlag, N / 2, nvRow);
            wave_lds_sync();  // the next row reuses the buffer
        }
    };

    // D. rows in groups of equal template length.  With a shared scale
    // buffer, a group of NWAVE rows or more shares one scale computed by
    // the block; rows of smaller groups (every row without the buffer) are
    // "solo": each wave takes every NWAVE-th of them and computes its row's
    // scale itself, so no wave waits for a group it has no row in.
    bool fftDone = false;
    if constexpr (SHARED) {
        for (int g0 = sg.rowBegin; g0 < rowEnd;) {
            const int Lt = rows[g0].Lt;
            int g1 = g0 + 1;
            while (g1 < rowEnd && rows[g1].Lt == Lt) { ++g1; }
            if (g1 - g0 >= kCorr2SharedGroup) {
                int myFlags = 0;
                if (!fftDone) {
                    // wave 0 transforms while waves 1..3 build the scale
                    if (wave == 0) {
                        forwardFft();
                    } else {
                        myFlags = compute_scale<T, TILE, 1, QUEUE>(raw, ts1, ts2, tp1, tp2, sc, a, nValidSeg, Lt, noise,
                                                         tid - 64, NT - 64);
                    }
                    fftDone = true;
                } else {
                    myFlags = compute_scale<T, TILE, 1, QUEUE>(raw, ts1, ts2, tp1, tp2, sc, a, nValidSeg, Lt, noise, tid,
                                                               NT);
                }
                const double flagSum = uniform(block_sum<NT>(static_cast<double>(myFlags), red));  // scalar registers
                const bool anyFlag = flagSum > 0.0;
                int r = g0 + wave;
                if (r < g1) {
                    if (PREFETCH) { loadH(rows[r].H); }
                    startRow(rows[r], rows[r + NWAVE < g1 ? r + NWAVE : r].H);
                    for (int rn = r + NWAVE; rn < g1; rn += NWAVE) {
                        finishRow(rows[r], r, sc, anyFlag, flagSum);
                        startRow(rows[rn], rows[rn + NWAVE < g1 ? rn + NWAVE : rn].H);
                        r = rn;
                    }
                    finishRow(rows[r], r, sc, anyFlag, flagSum);
                }
                __syncthreads();  // sc is rebuilt for the next group
            }
            g0 = g1;
        }
    }
    if (!fftDone) {
        if (wave == 0) { forwardFft(); }
        __syncthreads();
    }
    // Solo rows of this wave in order: the k-th solo row of the segment
    // belongs to wave k mod NWAVE.  nextSolo walks the groups (scalar
    // work) and returns the wave's next solo row, or -1.
    int sg0 = sg.rowBegin, sg1 = sg.rowBegin, sr = sg.rowBegin, sk = 0;
    auto nextSolo = [&]() -> int {
        while (true) {
            if (sr >= sg1) {
                sg0 = sg1;
                if (sg0 >= rowEnd) { return -1; }
                const int Lt = rows[sg0].Lt;
                sg1 = sg0 + 1;
                while (sg1 < rowEnd && rows[sg1].Lt == Lt) { ++sg1; }
                sr = (SHARED && sg1 - sg0 >= kCorr2SharedGroup) ? sg1 : sg0;
                continue;
            }
            const int r = sr++;
            if (((sk++) & (NWAVE - 1)) == wave) { return r; }
        }
    };
    int cur = nextSolo();
    if (cur >= 0) {
        if (PREFETCH) { loadH(rows[cur].H); }
        int nx = nextSolo();
        startRow(rows[cur], rows[nx >= 0 ? nx : cur].H);
        while (nx >= 0) {
            const int nn = nextSolo();
            finishRow(rows[cur], cur, nullptr, false, 0.0);
            startRow(rows[nx], rows[nn >= 0 ? nn : nx].H);
            cur = nx;
            nx = nn;
        }
        finishRow(rows[cur], cur, nullptr, false, 0.0);
    }
}

// Template demean / unit-norm (FP64) and spectrum conj(T^)/(2 nfft).
template <int M>
__global__ __launch_bounds__(NT) void tmpl_kernel(const TmplJob *__restrict__ jobs, int precision,
                                                  const Cx<double> *__restrict__ twloG,
                                                  const Cx<double> *__restrict__ twhiG)
{
    constexpr int N = 2 * M;
    __shared__ Cx<double> Z[M];
    __shared__ Cx<double> twlo[64];
    __shared__ Cx<double> twhi[N / 64];
    __shared__ double red[NT / 64];
    const TmplJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    for (int i = tid; i < 64; i += NT) { twlo[i] = twloG[i]; }
    for (int i = tid; i < N / 64; i += NT) { twhi[i] = twhiG[i]; }
    const Twiddle<double> tw{twlo, twhi};
    double s = 0;
    for (int m = tid; m < job.Lt; m += NT) { s += job.t[m]; }
    const double mean = block_sum<NT>(s, red) / job.Lt;
    double e = 0;
    for (int m = tid; m < job.Lt; m += NT) {
        const double d = job.t[m] - mean;
        e += d * d;
    }
    const double norm = sqrt(block_sum<NT>(e, red));
    const double scale = norm > 0 ? 1.0 / norm : 0.0;
    double *zf = reinterpret_cast<double *>(Z);
    for (int m = tid; m < N; m += NT) {
        double tv = 0;
        if (m < job.Lt) {
            tv = (job.t[m] - mean) * scale;
            job.tOut[m] = tv;
        }
        zf[m] = tv;
    }
    __syncthreads();
    fft_lds<double, M, NT, false>(Z, tw);
    const double hs = 1.0 / (2.0 * N);
    double hmax = 0;
    for (int k = tid; k <= M / 2; k += NT) {
        const Cx<double> zk = Z[k];
        const Cx<double> zm = cconj(Z[(M - k) & (M - 1)]);
        const Cx<double> E = cadd(zk, zm);
        const Cx<double> O = cmulmi(csub(zk, zm));
        const Cx<double> WO = cmul(tw(k), O);
        const Cx<double> tk = cadd(E, WO);             // 2 T^[k]
        const Cx<double> tm = cconj(csub(E, WO));      // 2 T^[M-k]
        hmax = fmax(hmax, 0.5 * sqrt(tk.x * tk.x + tk.y * tk.y));
        hmax = fmax(hmax, 0.5 * sqrt(tm.x * tm.x + tm.y * tm.y));
        // conj(T^)/(2N) = conj(2 T^) / (4N)
        const Cx<double> hk{tk.x * hs * 0.5, -tk.y * hs * 0.5};
        const Cx<double> hm{tm.x * hs * 0.5, -tm.y * hs * 0.5};
        if (precision == 0) {
            Cx<float> *H = static_cast<Cx<float> *>(job.H);
            H[k] = Cx<float>{static_cast<float>(hk.x), static_cast<float>(hk.y)};
            H[M - k] = Cx<float>{static_cast<float>(hm.x), static_cast<float>(hm.y)};
        } else {
            Cx<double> *H = static_cast<Cx<double> *>(job.H);
            H[k] = hk;
            H[M - k] = hm;
        }
    }
    hmax = block_max<NT>(hmax, red);
    if (tid == 0) { *job.hmax = static_cast<float>(hmax); }
}

// out[k] = sum_r w_r * |rho_r[k + s_r]| in row order, no contraction,
// as the reference's shiftAndStack loop (src/matchedFilter.cpp:1341-1361).
template <typename T>
__global__ __launch_bounds__(NT) void stack_kernel(const StackGroup *__restrict__ groups, int absval)
{
    // hipcc contracts a*b+c into an FMA by default; the stack keeps the
    // separately rounded product and sum of the loop it restates.
#pragma clang fp contract(off)
    const StackGroup grp = groups[blockIdx.y];
    const StackRow *__restrict__ srows = grp.rows;
    const int nRows = grp.nRows;
    const long long len = grp.len;
    T *__restrict__ out = static_cast<T *>(grp.out);
    // Each block owns one contiguous chunk of kStackChunk outputs, so the
    // cache lines a block reads are shared only with the chunks beside it
    // (neighbouring blocks sit on other XCDs, whose L2s would each fetch a
    // shared line).  In a pass a thread owns U outputs NT apart, so every
    // load instruction of a wave reads 64 consecutive samples of a row.
    // The row loop is unrolled (8 rows: 64 loads in flight per thread;
    // measured best against 4, 12, 16 and 24) so several rows' loads are in
    // flight while the per-output sums keep their sequential order.
    constexpr int U = 8;
    static_assert(kStackChunk % (NT * U) == 0, "stack chunk");
    const long long chunk0 = static_cast<long long>(blockIdx.x) * kStackChunk;
    if (chunk0 >= len) { return; }
    gptr<T> o = to_global(out);
    for (int pass = 0; pass < kStackChunk / (NT * U); ++pass) {
        const long long k0 = chunk0 + static_cast<long long>(pass) * NT * U + threadIdx.x;
        if (k0 >= len) { break; }
        T acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) { acc[u] = T(0); }
#pragma unroll 8
        for (int r = 0; r < nRows; ++r) {
            const StackRow sr = srows[r];
            const gptr<const T> p = to_global(static_cast<const T *>(sr.rho)) + sr.shift;
            const T w = static_cast<T>(sr.weight);
            T v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const long long k = k0 + u * NT;
                v[u] = k < len ? p[k] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const T a = absval ? fabs(v[u]) : v[u];
                acc[u] = acc[u] + w * a;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long k = k0 + u * NT;
            if (k < len) { o[k] = acc[u]; }
        }
    }
}

// Detector reduce + clamp (src/singleChannel/detector.cpp:190-250): per lag,
// the maximum over the rows of rho (or of |rho|), the first row reaching
// it (strict >, rows in order) and the clamp to [-1, 1].  Same chunking as
// the stack: each block owns kStackChunk lags, a thread U lags NT apart,
// eight rows' loads in flight.
template <typename T>
__global__ __launch_bounds__(NT) void reduce_kernel(const void *const *__restrict__ rowPtrs, int nRows,
                                                    long long len, int absval, T *__restrict__ det,
                                                    int *__restrict__ id)
{
    constexpr int U = 8;
    const long long chunk0 = static_cast<long long>(blockIdx.x) * kStackChunk;
    if (chunk0 >= len) { return; }
    for (int pass = 0; pass < kStackChunk / (NT * U); ++pass) {
        const long long k0 = chunk0 + static_cast<long long>(pass) * NT * U + threadIdx.x;
        if (k0 >= len) { break; }
        T best[U];
        int who[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            best[u] = T(0);
            who[u] = 0;
        }
#pragma unroll 8
        for (int r = 0; r < nRows; ++r) {
            const gptr<const T> p = to_global(static_cast<const T *>(rowPtrs[r]));
            T v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const long long k = k0 + u * NT;
                v[u] = k < len ? p[k] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const T a = absval ? fabs(v[u]) : v[u];
                if (r == 0 || a > best[u]) {
                    best[u] = a;
                    who[u] = r;
                }
            }
        }
        const gptr<T> dg = to_global(det);
        const gptr<int> ig = to_global(id);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long k = k0 + u * NT;
            if (k < len) {
                dg[k] = fmax(T(-1), fmin(T(1), best[u]));
                ig[k] = who[u];
            }
        }
    }
}

// How peak_scan reads its signal: as stored, or as the Detector's
// setCorrelations (src/singleChannel/detector.cpp:252-284) sees a row:
// clamped to [-1, 1] (NaN -> 1, as std::max(-1, std::min(1, x))), of rho
// or of |rho|.
constexpr int kPeakVectors = 4;  // 16-byte vectors per thread per pass
constexpr int kPeakRaw = 0;
constexpr int kPeakClamp = 1;
constexpr int kPeakClampAbs = 2;
template <int XF, typename T>
__device__ __forceinline__ T peak_xf(T v)
{
    if constexpr (XF == kPeakClampAbs) { v = fabs(v); }
    if constexpr (XF != kPeakRaw) { v = fmax(T(-1), fmin(T(1), v)); }
    return v;
}

template <int XF, typename T>
__device__ __forceinline__ T peak_value(gptr<const T> x, long long i)
{
    return peak_xf<XF>(x[i]);
}

// The candidate rule of src/peakFinder.cpp:500-521 for sample i of n, from
// its value and its neighbours': strictly above the threshold and a local
// maximum (a plateau keeps its last sample); an end point compares with its
// one neighbour.
template <typename T>
__device__ __forceinline__ bool peak_rule(T xi, T left, T right, long long i, long long n, T thr)
{
    if (!(xi > thr)) { return false; }
    if (n == 1) { return true; }
    if (i == 0) { return right < xi; }
    if (i == n - 1) { return xi > left; }
    return (xi >= left) && (right < xi);
}

// Candidates of x[0, n).  Each sample is read once: the body, from the
// first 16-byte boundary on, as 16-byte vectors, consecutive lanes holding
// consecutive vectors, so a vector's outer neighbours come from the lanes
// beside it (the first and last lane of a wave, and the last vector, load
// theirs, only when a sample of theirs is above the threshold).  The
// samples before the boundary and after the last whole vector (fewer than
// two vectors' worth) are tested by block 0 one per thread.  Candidates are
// counted per lane, scanned across the wave and appended with one atomic
// per wave; their order in idx/val is arbitrary.
template <typename T, int XF = kPeakRaw>
__device__ __forceinline__ void peak_scan(gptr<const T> x, long long n, T thr, int *count, gptr<int> idx,
                                          gptr<T> val, int cap)
{
    constexpr int VEC = 16 / static_cast<int>(sizeof(T));
    constexpr int U = kPeakVectors;
    static_assert(U * VEC <= 32, "candidate bits");
    using V = T __attribute__((ext_vector_type(VEC)));
    const int lane = threadIdx.x & 63;
    const unsigned mis = static_cast<unsigned>(reinterpret_cast<unsigned long long>(x) & 15ull);
    long long head = static_cast<long long>(((16u - mis) & 15u) / sizeof(T));
    if (head > n) { head = n; }
    const long long nVec = (n - head) / VEC;
    const long long tail0 = head + nVec * VEC;
    const long long nEdge = head + (n - tail0);
    if (blockIdx.x == 0 && threadIdx.x < nEdge) {
        const long long i = threadIdx.x < head ? threadIdx.x : tail0 + (threadIdx.x - head);
        const T xi = peak_value<XF>(x, i);
        if (xi > thr) {
            const T left = i > 0 ? peak_value<XF>(x, i - 1) : T(0);
            const T right = i + 1 < n ? peak_value<XF>(x, i + 1) : T(0);
            if (peak_rule(xi, left, right, i, n, thr)) {
                const int pos = atomicAdd(count, 1);
                if (pos < cap) {
                    idx[pos] = static_cast<int>(i);
                    val[pos] = xi;
                }
            }
        }
    }
    const gptr<const V> xv = reinterpret_cast<gptr<const V>>(x + head);
    const long long span = static_cast<long long>(NT) * U;
    for (long long base = static_cast<long long>(blockIdx.x) * span; base < nVec;
         base += static_cast<long long>(gridDim.x) * span) {
        V v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long k = base + u * NT + threadIdx.x;
            v[u] = k < nVec ? xv[k] : V(T(0));
        }
        unsigned bits = 0u;
        T s[U][VEC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long k = base + u * NT + threadIdx.x;
            bool any = false;
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                s[u][e] = peak_xf<XF>(v[u][e]);
                any = any || (s[u][e] > thr);
            }
            T left = __shfl_up(s[u][VEC - 1], 1, 64);
            T right = __shfl_down(s[u][0], 1, 64);
            if (k < nVec && any) {
                const long long i0 = head + k * VEC;
                if (lane == 0 && i0 > 0) { left = peak_value<XF>(x, i0 - 1); }
                if ((lane == 63 || k + 1 >= nVec) && i0 + VEC < n) { right = peak_value<XF>(x, i0 + VEC); }
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const T l = e > 0 ? s[u][e - 1] : left;
                    const T r = e + 1 < VEC ? s[u][e + 1] : right;
                    if (peak_rule(s[u][e], l, r, i0 + e, n, thr)) { bits |= 1u << (u * VEC + e); }
                }
            }
        }
        if (__ballot(bits != 0u) != 0ull) {
            const int c = __popc(bits);
            int incl = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o, 64);
                if (lane >= o) { incl += t; }
            }
            int basePos = 0;
            if (lane == 63) { basePos = atomicAdd(count, incl); }
            int pos = __shfl(basePos, 63, 64) + incl - c;
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    if ((bits >> (u * VEC + e)) & 1u) {
                        if (pos < cap) {
                            idx[pos] = static_cast<int>(head + (base + u * NT + threadIdx.x) * VEC + e);
                            val[pos] = s[u][e];
                        }
                        ++pos;
                    }
                }
            }
        }
    }
}

// Workgroups of a peak scan over at most maxLen samples of tsize bytes:
// one pass of kPeakVectors vectors per thread, at least one block (block 0
// also tests the unaligned ends), at most `cap` blocks (grid-stride beyond).
inline int peak_grid(long long maxLen, int tsize, int cap)
{
    const long long vecs = maxLen / (16 / tsize);
    const long long want = (vecs + NT * kPeakVectors - 1) / (NT * kPeakVectors);
    return static_cast<int>(want < 1 ? 1 : (want < cap ? want : cap));
}

// Stack and candidates of every group in one launch (the detect path).  As
// stack_kernel, each block owns kStackChunk outputs (one pass of U = 8 per
// thread) and writes them; the block's sums also go to LDS, where each
// output above the threshold finds its neighbours.  An output at the
// block's edge recomputes its outside neighbour from the rows: the same
// sum, in the same order, as the block that owns it.  Candidates as in
// peak_scan (the rule of src/peakFinder.cpp:500-521, one atomic per wave).
template <typename T>
__global__ __launch_bounds__(NT) void stack_peak_kernel(const StackGroup *__restrict__ groups, int absval, T thr)
{
#pragma clang fp contract(off)
    constexpr int U = 8;
    static_assert(kStackChunk == NT * U, "one pass per block");
    __shared__ T sv[kStackChunk];
    const StackGroup grp = groups[blockIdx.y];
    const StackRow *__restrict__ srows = grp.rows;
    const int nRows = grp.nRows;
    const long long len = grp.len;
    const long long chunk0 = static_cast<long long>(blockIdx.x) * kStackChunk;
    if (chunk0 >= len) { return; }  // uniform per block
    const long long k0 = chunk0 + threadIdx.x;
    T acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { acc[u] = T(0); }
#pragma unroll 8
    for (int r = 0; r < nRows; ++r) {
        const StackRow sr = srows[r];
        const gptr<const T> p = to_global(static_cast<const T *>(sr.rho)) + sr.shift;
        const T w = static_cast<T>(sr.weight);
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long k = k0 + u * NT;
            v[u] = k < len ? p[k] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T a = absval ? fabs(v[u]) : v[u];
            acc[u] = acc[u] + w * a;
        }
    }
    gptr<T> o = to_global(static_cast<T *>(grp.out));
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const long long k = k0 + u * NT;
        if (k < len) { o[k] = acc[u]; }
        sv[u * NT + threadIdx.x] = acc[u];
    }
    __syncthreads();
    // the stack at an output outside the block (0 <= k < len)
    const auto stackAt = [&](long long k) {
        T sum = T(0);
        for (int r = 0; r < nRows; ++r) {
            const StackRow sr = srows[r];
            const T x = to_global(static_cast<const T *>(sr.rho))[sr.shift + k];
            sum = sum + static_cast<T>(sr.weight) * (absval ? fabs(x) : x);
        }
        return sum;
    };
    unsigned bits = 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int j = u * NT + threadIdx.x;
        const long long k = chunk0 + j;
        const T xi = acc[u];
        if (k < len && xi > thr) {
            const T left = j > 0 ? sv[j - 1] : (k > 0 ? stackAt(k - 1) : T(0));
            const T right = k + 1 < len ? (j + 1 < kStackChunk ? sv[j + 1] : stackAt(k + 1)) : T(0);
            if (peak_rule(xi, left, right, k, len, thr)) { bits |= 1u << u; }
        }
    }
    const int lane = threadIdx.x & 63;
    if (__ballot(bits != 0u) != 0ull) {
        const int c = __popc(bits);
        int incl = c;
#pragma unroll
        for (int o2 = 1; o2 < 64; o2 <<= 1) {
            const int t = __shfl_up(incl, o2, 64);
            if (lane >= o2) { incl += t; }
        }
        int basePos = 0;
        if (lane == 63) { basePos = atomicAdd(grp.count, incl); }
        int pos = __shfl(basePos, 63, 64) + incl - c;
        const gptr<int> idx = to_global(grp.candIdx);
        const gptr<T> val = to_global(static_cast<T *>(grp.candVal));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if ((bits >> u) & 1u) {
                if (pos < grp.candCap) {
                    idx[pos] = static_cast<int>(chunk0 + u * NT + threadIdx.x);
                    val[pos] = acc[u];
                }
                ++pos;
            }
        }
    }
}

// Streaming push of every channel, in 32-bit words, two launches: first
// each channel's last `carry` words (the buffer's tail) go to its slot of
// the scratch area, then each buffer is rewritten as carry + new chunk.
// Two launches because the ranges overlap when the chunk is shorter than
// the carry.  grid.y = channel.
__global__ __launch_bounds__(NT) void push_save_kernel(PushTable tab, long long chunkWords, long long carryWords,
                                                       unsigned *__restrict__ scratch)
{
    const int c = blockIdx.y;
    const unsigned *__restrict__ buf = static_cast<const unsigned *>(tab.buf[c]);
    unsigned *__restrict__ dst = scratch + static_cast<long long>(c) * carryWords;
    for (long long i = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; i < carryWords;
         i += static_cast<long long>(gridDim.x) * NT) {
        dst[i] = buf[chunkWords + i];
    }
}

__global__ __launch_bounds__(NT) void push_write_kernel(PushTable tab, long long chunkWords, long long carryWords,
                                                        const unsigned *__restrict__ scratch)
{
    const int c = blockIdx.y;
    unsigned *__restrict__ buf = static_cast<unsigned *>(tab.buf[c]);
    const unsigned *__restrict__ src = static_cast<const unsigned *>(tab.src[c]);
    const unsigned *__restrict__ saved = scratch + static_cast<long long>(c) * carryWords;
    const long long total = carryWords + chunkWords;
    for (long long j = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * NT) {
        buf[j] = j < carryWords ? saved[j] : src[j - carryWords];
    }
}

__global__ __launch_bounds__(NT) void fill_int_kernel(int *__restrict__ dst, long long n, int value)
{
    for (long long i = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * NT) {
        dst[i] = value;
    }
}

// out[i] = src[idx[i]]
__global__ __launch_bounds__(NT) void gather_int_kernel(const int *__restrict__ src, const int *__restrict__ idx,
                                                        int n, int *__restrict__ out)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < n) { out[i] = src[idx[i]]; }
}

template <typename T>
__global__ __launch_bounds__(NT) void peak_kernel(const T *__restrict__ x, long long n, T thr,
                                                  int *__restrict__ count, int *__restrict__ idx,
                                                  T *__restrict__ val, int cap)
{
    peak_scan<T>(to_global(x), n, thr, count, to_global(idx), to_global(val), cap);
}

// The candidates of every group's stack (grid.y = group); with XF, of
// clamped rows (the Detector's per-template policy).
template <typename T, int XF = kPeakRaw>
__global__ __launch_bounds__(NT) void peak_groups_kernel(const StackGroup *__restrict__ groups, T thr)
{
    const StackGroup g = groups[blockIdx.y];
    peak_scan<T, XF>(to_global(static_cast<const T *>(g.out)), g.len, thr, g.count, to_global(g.candIdx),
                     to_global(static_cast<T *>(g.candVal)), g.candCap);
}

// Packs every group's candidates into one array, group by group in order
// (block g: offset = the counts of groups before g, each capped at its
// buffer's capacity so the packed writes stay in bounds even when a count
// overflowed), and copies the raw counts in front: packed[0..nG) = counts,
// then the indices; values alike.  The host then reads one contiguous
// prefix (and rejects an overflowed count).
template <typename T>
__global__ __launch_bounds__(NT) void pack_candidates_kernel(const StackGroup *__restrict__ groups, int nG,
                                                             int *__restrict__ packedIdx, T *__restrict__ packedVal)
{
    __shared__ long long red[NT / 64];
    const int g = blockIdx.x;
    long long off = 0;
    for (int q = threadIdx.x; q < g; q += NT) { off += min(*groups[q].count, groups[q].candCap); }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { off += __shfl_xor(off, o, 64); }
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = off; }
    __syncthreads();
    off = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) { off += red[w]; }
    const StackGroup gr = groups[g];
    const int cnt = min(*gr.count, gr.candCap);
    if (threadIdx.x == 0) { packedIdx[g] = *gr.count; }
    const gptr<const int> si = to_global(gr.candIdx);
    const gptr<const T> sv = to_global(static_cast<const T *>(gr.candVal));
    for (int i = threadIdx.x; i < cnt; i += NT) {
        packedIdx[nG + off + i] = si[i];
        packedVal[off + i] = sv[i];
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void minmax_kernel(const T *__restrict__ x, long long n,
                                                    double *__restrict__ pmin,
                                                    double *__restrict__ pmax)
{
    __shared__ double red[NT / 64];
    const long long stride = static_cast<long long>(gridDim.x) * NT;
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (long long i = static_cast<long long>(blockIdx.x) * NT + threadIdx.x; i < n; i += stride) {
        const double v = static_cast<double>(x[i]);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    const double bmax = block_max<NT>(hi, red);
    const double bmin = -block_max<NT>(-lo, red);
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = bmin;
        pmax[blockIdx.x] = bmax;
    }
}

// Dead-trace test of src/matchedFilter.cpp:894 for a device-resident
// segment table.  Pass 1: min and max of each kDeadChunk-sample chunk of
// each segment's channel (grid.y = segment); pass 2: one thread per
// segment folds its chunks and writes seg.dead.
template <typename T>
__global__ __launch_bounds__(NT) void seg_minmax_kernel(const Seg *__restrict__ segs, double *__restrict__ part,
                                                        int chunksPerSeg)
{
    __shared__ double red[NT / 64];
    const Seg sg = segs[blockIdx.y];
    const long long i0 = static_cast<long long>(blockIdx.x) * kDeadChunk;
    double lo = DBL_MAX, hi = -DBL_MAX;
    if (i0 < sg.n) {
        const gptr<const T> x = to_global(static_cast<const T *>(sg.x));
        const long long i1 = min(sg.n, i0 + kDeadChunk);
        for (long long i = i0 + threadIdx.x; i < i1; i += NT) {
            const double v = static_cast<double>(x[i]);
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
    }
    const double bmax = block_max<NT>(hi, red);
    const double bmin = -block_max<NT>(-lo, red);
    if (threadIdx.x == 0) {
        const size_t k = (static_cast<size_t>(blockIdx.y) * chunksPerSeg + blockIdx.x) * 2;
        part[k] = bmin;
        part[k + 1] = bmax;
    }
}

__global__ __launch_bounds__(NT) void seg_dead_kernel(Seg *__restrict__ segs, int nSegs,
                                                      const double *__restrict__ part, int chunksPerSeg)
{
    const int s = blockIdx.x * NT + threadIdx.x;
    if (s >= nSegs) { return; }
    const long long n = segs[s].n;
    const int nch = static_cast<int>((n + kDeadChunk - 1) / kDeadChunk);
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (int c = 0; c < nch; ++c) {
        const size_t k = (static_cast<size_t>(s) * chunksPerSeg + c) * 2;
        lo = fmin(lo, part[k]);
        hi = fmax(hi, part[k + 1]);
    }
    segs[s].dead = (n < 1 || fabs(lo) == fabs(hi)) ? 1 : 0;
}

template <typename T, int M>
hipError_t corr_launch(const Item *dItems, int nItems, const Seg *dSegs, const Row *dRows,
                       const void *lo, const void *hi, hipStream_t stream)
{
    hipLaunchKernelGGL((corr_kernel<T, M>), dim3(nItems), dim3(NT), 0, stream, dItems, dSegs,
                       dRows, static_cast<const Cx<T> *>(lo), static_cast<const Cx<T> *>(hi));
    return hipGetLastError();
}

template <typename T, int M>
hipError_t corr2_launch(const Item *dItems, int nItems, const Seg *dSegs, const Row *dRows,
                        const void *lo, const void *hi, FixQueue queue, hipStream_t stream)
{
    if (sizeof(T) == 4 && queue.over != nullptr) {
        hipLaunchKernelGGL((corr2_kernel<T, M, sizeof(T) == 4>), dim3(nItems), dim3(NT), 0, stream, dItems, dSegs, dRows,
                           static_cast<const Cx<T> *>(lo), static_cast<const Cx<T> *>(hi), queue);
    } else {
        hipLaunchKernelGGL((corr2_kernel<T, M, false>), dim3(nItems), dim3(NT), 0, stream, dItems, dSegs, dRows,
                           static_cast<const Cx<T> *>(lo), static_cast<const Cx<T> *>(hi), FixQueue{});
    }
    return hipGetLastError();
}

// Exact FP64 value of each queued FP32 window, one window per thread: one
// pass over d = x - x0 (x0 the window's first sample, so every d is exact
// and a cancelled variance keeps its digits; a constant window gives 0):
// rho = sum t d / sqrt(sum d^2 - (sum d)^2 / Lt), the template's zero mean
// making sum t d the centred numerator; |rho| > 1.1 gives 0 (the
// reference's tooBig rule).  Rows of either table (FFT rows, time-domain
// rows): their channel samples, FP64 template, output row and length.
// Sums of one window over d = x[m] - shift, m = 0..Lt-1 in order (s1, s2)
// and against the FP64 taps t (dn).  The taps are read two per 16-byte
// load and, from global memory (GLOBAL), the samples four per load; the
// addresses are only 8- and 4-byte aligned, which gfx950 serves for global
// loads.  The sums keep the order of a tap-by-tap loop.
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
template <bool GLOBAL, typename XPtr>
__device__ __forceinline__ void window_tap_sums(XPtr x, gptr<const double> t, int Lt, double shift, double &s1,
                                                double &s2, double &dn)
{
    s1 = 0;
    s2 = 0;
    dn = 0;
    int m = 0;
#pragma unroll 2
    for (; m + 4 <= Lt; m += 4) {
        const f64x2u ta = *reinterpret_cast<const __attribute__((address_space(1))) f64x2u *>(t + m);
        const f64x2u tb = *reinterpret_cast<const __attribute__((address_space(1))) f64x2u *>(t + m + 2);
        float xv[4];
        if constexpr (GLOBAL) {
            const f32x4u v = *reinterpret_cast<const __attribute__((address_space(1))) f32x4u *>(x + m);
#pragma unroll
            for (int e = 0; e < 4; ++e) { xv[e] = v[e]; }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) { xv[e] = x[m + e]; }
        }
        const double tv[4] = {ta[0], ta[1], tb[0], tb[1]};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double d = static_cast<double>(xv[e]) - shift;
            s1 += d;
            s2 = fma(d, d, s2);
            dn = fma(tv[e], d, dn);
        }
    }
    for (; m < Lt; ++m) {
        const double d = static_cast<double>(x[m]) - shift;
        s1 += d;
        s2 = fma(d, d, s2);
        dn = fma(t[m], d, dn);
    }
}

template <typename RowT>
__global__ __launch_bounds__(NT) void fix_kernel(const RowT *__restrict__ rows, FixQueue queue)
{
    const unsigned long long n = min(*queue.count, queue.cap);
    const gptr<const unsigned long long> e = to_global(reinterpret_cast<const unsigned long long *>(queue.entries));
    for (unsigned long long k = blockIdx.x * NT + threadIdx.x; k < n; k += static_cast<unsigned long long>(gridDim.x) * NT) {
        const unsigned long long v = e[k];
        if (v == kFixEmpty) { continue; }  // a slot left by a reservation that straddled cap
        const RowT row = rows[static_cast<unsigned>(v) & 0x7fffffffu];
        const int pos = static_cast<int>(v >> 32);
        const gptr<const float> x = to_global(static_cast<const float *>(row.x)) + pos;
        const gptr<const double> t = to_global(row.tmpl);
        const int Lt = row.Lt;
        double s1, s2, dn;
        window_tap_sums<true>(x, t, Lt, static_cast<double>(x[0]), s1, s2, dn);
        double rho = dn / fmax(DBL_EPSILON, sqrt(fmax(0.0, s2 - s1 * s1 / Lt)));
        if (fabs(rho) > 1.1) { rho = 0; }
        to_global(static_cast<float *>(row.out))[pos] = static_cast<float>(rho);
    }
}

// The recorded row blocks of a queue-mode wave-kernel launch: each NaN
// output (a flagged window) recomputed exactly, as fix_kernel does; one
// record per workgroup.  The record's samples are staged in LDS with their
// equal-neighbour bits, so an exactly constant window (a gap) is 0 at once
// (zero numerator over the floor of the norm); records longer than the
// staging buffer read their samples from global memory.
constexpr int kFixSpan = 6144;  // samples of a record staged in LDS: outputs + Lt - 1
__global__ __launch_bounds__(NT) void fix_blocks_kernel(const Row *__restrict__ rows, FixQueue queue)
{
    __shared__ float xs[kFixSpan];
    __shared__ unsigned eqw[kFixSpan / 32];
    const unsigned long long n = *queue.overCount;
    const int tid = threadIdx.x;
    for (unsigned long long k = blockIdx.x; k < n; k += gridDim.x) {
        const gptr<const int> rec = to_global(static_cast<const int *>(queue.over)) + 4 * k;
        const Row row = rows[rec[0]];
        const int p0 = rec[1], cnt = rec[2];
        const gptr<const double> t = to_global(row.tmpl);
        const int Lt = row.Lt;
        const gptr<const float> xg = to_global(static_cast<const float *>(row.x)) + p0;
        gptr<float> out = to_global(static_cast<float *>(row.out)) + p0;
        const int span = cnt + Lt - 1;
        if (span <= kFixSpan) {  // uniform over the workgroup
            __syncthreads();     // the previous record's reads are done
            for (int j = tid; j < span; j += NT) { xs[j] = xg[j]; }
            __syncthreads();
            for (int w = tid; w < (span + 31) / 32; w += NT) {
                unsigned bits = 0;
                for (int b = 0; b < 32; ++b) {
                    const int j = 32 * w + b;
                    if (j + 1 < span && xs[j] == xs[j + 1]) { bits |= 1u << b; }
                }
                eqw[w] = bits;
            }
            __syncthreads();
            for (int q = tid; q < cnt; q += NT) {
                if (!isnan(out[q])) { continue; }
                if (constant_window(eqw, q, Lt)) {
                    out[q] = 0.0f;
                    continue;
                }
                double s1, s2, dn;
                window_tap_sums<false>(xs + q, t, Lt, static_cast<double>(xs[q]), s1, s2, dn);
                double rho = dn / fmax(DBL_EPSILON, sqrt(fmax(0.0, s2 - s1 * s1 / Lt)));
                if (fabs(rho) > 1.1) { rho = 0; }
                out[q] = static_cast<float>(rho);
            }
            continue;
        }
        for (int q = tid; q < cnt; q += NT) {
            if (!isnan(out[q])) { continue; }
            const gptr<const float> x = xg + q;
            double s1, s2, dn;
            window_tap_sums<true>(x, t, Lt, static_cast<double>(x[0]), s1, s2, dn);
            double rho = dn / fmax(DBL_EPSILON, sqrt(fmax(0.0, s2 - s1 * s1 / Lt)));
            if (fabs(rho) > 1.1) { rho = 0; }
            out[q] = static_cast<float>(rho);
        }
    }
}

template <int M>
hipError_t tmpl_launch(int precision, const TmplJob *dJobs, int nJobs, const void *lo,
                       const void *hi, hipStream_t stream)
{
    hipLaunchKernelGGL((tmpl_kernel<M>), dim3(nJobs), dim3(NT), 0, stream, dJobs, precision,
                       static_cast<const Cx<double> *>(lo), static_cast<const Cx<double> *>(hi));
    return hipGetLastError();
}

}  // namespace

bool corr_variant_available(int precision, int logM, int variant)
{
    if (variant == 2) { return precision == 0 ? (logM >= 9 && logM <= 11) : (logM >= 9 && logM <= 10); }
    if (variant == 1) { return precision == 0 ? (logM >= 9 && logM <= 12) : (logM >= 9 && logM <= 11); }
    return false;
}

hipError_t launch_corr(int precision, int logM, int variant, const Item *dItems, int nItems,
                       const Seg *dSegs, const Row *dRows, const void *dTwLo, const void *dTwHi,
                       FixQueue queue, hipStream_t stream)
{
    if (nItems <= 0) { return hipSuccess; }
    if (variant == 2) {
        if (precision == 0) {
            switch (logM) {
            case 9: return corr2_launch<float, 512>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, queue, stream);
            case 10: return corr2_launch<float, 1024>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, queue, stream);
            case 11: return corr2_launch<float, 2048>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, queue, stream);
            default: return hipErrorInvalidValue;
            }
        }
        // FP64 rows are recomputed in place (their tolerance wants the
        // block's own anchored samples)
        switch (logM) {
        case 9: return corr2_launch<double, 512>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, FixQueue{}, stream);
        case 10: return corr2_launch<double, 1024>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, FixQueue{}, stream);
        default: return hipErrorInvalidValue;
        }
    }
    if (precision == 0) {
        switch (logM) {
        case 9: return corr_launch<float, 512>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
        case 10: return corr_launch<float, 1024>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
        case 11: return corr_launch<float, 2048>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
        case 12: return corr_launch<float, 4096>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (logM) {
    case 9: return corr_launch<double, 512>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
    case 10: return corr_launch<double, 1024>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
    case 11: return corr_launch<double, 2048>(dItems, nItems, dSegs, dRows, dTwLo, dTwHi, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fix(const Row *dRows, FixQueue queue, int blocks, hipStream_t stream)
{
    if (blocks <= 0 || queue.over == nullptr) { return hipSuccess; }
    hipLaunchKernelGGL(fix_blocks_kernel, dim3(blocks), dim3(NT), 0, stream, dRows, queue);
    return hipGetLastError();
}

hipError_t launch_fix(const TdRow *dRows, FixQueue queue, int blocks, hipStream_t stream)
{
    if (blocks <= 0 || queue.entries == nullptr) { return hipSuccess; }
    hipLaunchKernelGGL(fix_kernel<TdRow>, dim3(blocks), dim3(NT), 0, stream, dRows, queue);
    return hipGetLastError();
}

hipError_t launch_tmpl(int precision, int logM, const TmplJob *dJobs, int nJobs,
                       const void *dTwLo64, const void *dTwHi64, hipStream_t stream)
{
    if (nJobs <= 0) { return hipSuccess; }
    switch (logM) {
    case 9: return tmpl_launch<512>(precision, dJobs, nJobs, dTwLo64, dTwHi64, stream);
    case 10: return tmpl_launch<1024>(precision, dJobs, nJobs, dTwLo64, dTwHi64, stream);
    case 11: return tmpl_launch<2048>(precision, dJobs, nJobs, dTwLo64, dTwHi64, stream);
    case 12: return tmpl_launch<4096>(precision, dJobs, nJobs, dTwLo64, dTwHi64, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stack_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                               int absval, hipStream_t stream)
{
    if (maxLen <= 0 || nGroups <= 0) { return hipSuccess; }
    const long long chunks = (maxLen + kStackChunk - 1) / kStackChunk;
    if (chunks > 0x7fffffffLL) { return hipErrorInvalidValue; }
    const int grid = static_cast<int>(chunks);
    if (precision == 0) {
        hipLaunchKernelGGL(stack_kernel<float>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups, absval);
    } else {
        hipLaunchKernelGGL(stack_kernel<double>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups, absval);
    }
    return hipGetLastError();
}

hipError_t launch_reduce(int precision, const void *const *dRowPtrs, int nRows, long long len, int absval,
                         void *dDet, int *dId, hipStream_t stream)
{
    if (len <= 0 || nRows <= 0) { return hipSuccess; }
    const long long chunks = (len + kStackChunk - 1) / kStackChunk;
    if (chunks > 0x7fffffffLL) { return hipErrorInvalidValue; }
    const int grid = static_cast<int>(chunks);
    if (precision == 0) {
        hipLaunchKernelGGL(reduce_kernel<float>, dim3(grid), dim3(NT), 0, stream, dRowPtrs, nRows, len, absval,
                           static_cast<float *>(dDet), dId);
    } else {
        hipLaunchKernelGGL(reduce_kernel<double>, dim3(grid), dim3(NT), 0, stream, dRowPtrs, nRows, len, absval,
                           static_cast<double *>(dDet), dId);
    }
    return hipGetLastError();
}

hipError_t launch_pack_candidates(int precision, const StackGroup *dGroups, int nGroups, int *dPackedIdx,
                                  void *dPackedVal, hipStream_t stream)
{
    if (nGroups <= 0) { return hipSuccess; }
    if (precision == 0) {
        hipLaunchKernelGGL(pack_candidates_kernel<float>, dim3(nGroups), dim3(NT), 0, stream, dGroups, nGroups,
                           dPackedIdx, static_cast<float *>(dPackedVal));
    } else {
        hipLaunchKernelGGL(pack_candidates_kernel<double>, dim3(nGroups), dim3(NT), 0, stream, dGroups, nGroups,
                           dPackedIdx, static_cast<double *>(dPackedVal));
    }
    return hipGetLastError();
}

hipError_t launch_push(const PushTable &tab, int nChannels, long long chunkWords, long long carryWords,
                       unsigned *dScratch, hipStream_t stream)
{
    if (nChannels <= 0 || nChannels > kPushMaxChannels || chunkWords <= 0 || carryWords < 0) {
        return hipErrorInvalidValue;
    }
    // about four words per thread per launch, at most 1024 blocks per channel
    auto blocks = [](long long words) {
        const long long b = (words + NT * 4 - 1) / (NT * 4);
        return static_cast<unsigned>(b < 1 ? 1 : (b > 1024 ? 1024 : b));
    };
    if (carryWords > 0) {
        hipLaunchKernelGGL(push_save_kernel, dim3(blocks(carryWords), nChannels), dim3(NT), 0, stream, tab,
                           chunkWords, carryWords, dScratch);
    }
    hipLaunchKernelGGL(push_write_kernel, dim3(blocks(carryWords + chunkWords), nChannels), dim3(NT), 0, stream, tab,
                       chunkWords, carryWords, dScratch);
    return hipGetLastError();
}

hipError_t launch_gather_int(const int *dSrc, const int *dIdx, int n, int *dOut, hipStream_t stream)
{
    if (n <= 0) { return hipSuccess; }
    hipLaunchKernelGGL(gather_int_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, stream, dSrc, dIdx, n, dOut);
    return hipGetLastError();
}

hipError_t launch_stack_peak_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                                    int absval, double threshold, hipStream_t stream)
{
    if (maxLen <= 0 || nGroups <= 0) { return hipSuccess; }
    const long long chunks = (maxLen + kStackChunk - 1) / kStackChunk;
    if (chunks > 0x7fffffffLL || nGroups > 65535) { return hipErrorInvalidValue; }
    const int grid = static_cast<int>(chunks);
    if (precision == 0) {
        hipLaunchKernelGGL(stack_peak_kernel<float>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups, absval,
                           static_cast<float>(threshold));
    } else {
        hipLaunchKernelGGL(stack_peak_kernel<double>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups, absval,
                           threshold);
    }
    return hipGetLastError();
}

hipError_t launch_peak_groups(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                              double threshold, hipStream_t stream)
{
    if (maxLen <= 0 || nGroups <= 0) { return hipSuccess; }
    const int grid = peak_grid(maxLen, precision == 0 ? 4 : 8, 8192);
    if (precision == 0) {
        hipLaunchKernelGGL(peak_groups_kernel<float>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups,
                           static_cast<float>(threshold));
    } else {
        hipLaunchKernelGGL(peak_groups_kernel<double>, dim3(grid, nGroups), dim3(NT), 0, stream, dGroups,
                           threshold);
    }
    return hipGetLastError();
}

hipError_t launch_peak_rows(int precision, const StackGroup *dGroups, int nGroups, long long maxLen,
                            double threshold, int absval, hipStream_t stream)
{
    if (maxLen <= 0 || nGroups <= 0) { return hipSuccess; }
    if (nGroups > 65535) { return hipErrorInvalidValue; }
    const int grid = peak_grid(maxLen, precision == 0 ? 4 : 8, 8192);
    if (precision == 0) {
        if (absval) {
            hipLaunchKernelGGL((peak_groups_kernel<float, kPeakClampAbs>), dim3(grid, nGroups), dim3(NT), 0, stream,
                               dGroups, static_cast<float>(threshold));
        } else {
            hipLaunchKernelGGL((peak_groups_kernel<float, kPeakClamp>), dim3(grid, nGroups), dim3(NT), 0, stream,
                               dGroups, static_cast<float>(threshold));
        }
    } else {
        if (absval) {
            hipLaunchKernelGGL((peak_groups_kernel<double, kPeakClampAbs>), dim3(grid, nGroups), dim3(NT), 0,
                               stream, dGroups, threshold);
        } else {
            hipLaunchKernelGGL((peak_groups_kernel<double, kPeakClamp>), dim3(grid, nGroups), dim3(NT), 0, stream,
                               dGroups, threshold);
        }
    }
    return hipGetLastError();
}

hipError_t launch_seg_dead(int precision, Seg *dSegs, int nSegs, long long maxN, double *dPart,
                           hipStream_t stream)
{
    if (nSegs <= 0) { return hipSuccess; }
    const long long chunks = maxN < 1 ? 1 : (maxN + kDeadChunk - 1) / kDeadChunk;
    if (chunks > 65535 || nSegs > 65535) { return hipErrorInvalidValue; }
    const int cps = static_cast<int>(chunks);
    if (precision == 0) {
        hipLaunchKernelGGL(seg_minmax_kernel<float>, dim3(cps, nSegs), dim3(NT), 0, stream, dSegs, dPart, cps);
    } else {
        hipLaunchKernelGGL(seg_minmax_kernel<double>, dim3(cps, nSegs), dim3(NT), 0, stream, dSegs, dPart, cps);
    }
    hipLaunchKernelGGL(seg_dead_kernel, dim3((nSegs + NT - 1) / NT), dim3(NT), 0, stream, dSegs, nSegs,
                       static_cast<const double *>(dPart), cps);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *dDst, long long n, int value, hipStream_t stream)
{
    if (n <= 0) { return hipSuccess; }
    const long long want = (n + NT - 1) / NT;
    const int grid = static_cast<int>(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid), dim3(NT), 0, stream, dDst, n, value);
    return hipGetLastError();
}

hipError_t launch_peaks(int precision, const void *dx, long long n, double threshold, int *dCount,
                        int *dIdx, void *dVal, int cap, hipStream_t stream)
{
    if (n <= 0) { return hipSuccess; }
    const int grid = peak_grid(n, precision == 0 ? 4 : 8, 65536);
    if (precision == 0) {
        hipLaunchKernelGGL(peak_kernel<float>, dim3(grid), dim3(NT), 0, stream,
                           static_cast<const float *>(dx), n, static_cast<float>(threshold), dCount,
                           dIdx, static_cast<float *>(dVal), cap);
    } else {
        hipLaunchKernelGGL(peak_kernel<double>, dim3(grid), dim3(NT), 0, stream,
                           static_cast<const double *>(dx), n, threshold, dCount, dIdx,
                           static_cast<double *>(dVal), cap);
    }
    return hipGetLastError();
}

hipError_t launch_minmax(int precision, const void *dx, long long n, double *dPartMin,
                         double *dPartMax, int nBlocks, hipStream_t stream)
{
    if (precision == 0) {
        hipLaunchKernelGGL(minmax_kernel<float>, dim3(nBlocks), dim3(NT), 0, stream,
                           static_cast<const float *>(dx), n, dPartMin, dPartMax);
    } else {
        hipLaunchKernelGGL(minmax_kernel<double>, dim3(nBlocks), dim3(NT), 0, stream,
                           static_cast<const double *>(dx), n, dPartMin, dPartMax);
    }
    return hipGetLastError();
}

}  // namespace mfk

// Time-domain correlation of short templates on MFMA (the optional
// short-template path; no FFT).
//
// For one lane (a channel and its rows, every template at most KT taps)
// the numerators form a Hankel-by-template GEMM
//     D[j][i] = sum_k t~_j[k] * x~[i + k],
// t~_j the demeaned unit-norm template j (zero past its length) and
// x~ = (x - a) / s the block's samples with the block anchor a removed and
// scaled by their RMS s.  Both operands are split into two 16-bit parts
// (hi = round(v), lo = round(v - hi)); the three products hi*hi, hi*lo and
// lo*hi are accumulated in FP32 by v_mfma_f32_32x32x16 (hi*hi in its own
// accumulator, the two cross terms in a second one).  rho = s * D * scale
// with the window normalisation from FP64 prefix sums of the block.
//
// Work item: IS = 512 consecutive outputs of the lane and up to 128 rows.
// Each of the 4 waves owns 32 rows: their split taps stay in registers
// (weight-stationary, KT/2 VGPRs) while the wave walks the 16 output tiles
// of 32; the Hankel operand is read from LDS, where entry p holds the 8
// samples x~[p .. p+7] (one 16-byte read per split and k-step).
//
// Accuracy: per window |d rho| <= eps_m * sqrt(S2) / den with
// eps_m = 3 * 2^-2u (split, u = 11 mantissa bits for FP16, 8 for BF16)
// + 4 sqrt(KT/16) 2^-24 (FP32 accumulation of the hi*hi sums), S2 the
// window's sum of (x - a)^2 and den its centred norm.  Windows whose bound
// exceeds the FP32 tolerance, or whose FP64 variance lost too many digits,
// are recomputed exactly (FP64 two-pass Pearson), as in the FFT path.
#include "mf_device.hpp"
#include "mf_kernels.hpp"
#include <cfloat>

using namespace mfdev;

namespace mfk {

namespace {

constexpr int NT = kThreads;
constexpr int IS = kTdOutputsPerItem;  // outputs per work item
constexpr int NTILE_I = IS / 32;       // output tiles of 32 per item

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

template <typename MT>
struct TdTraits;
template <>
struct TdTraits<_Float16> {
    using vec8 = f16x8;
    static constexpr int mantissa = 11;
    __device__ static f32x16 mfma(vec8 a, vec8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct TdTraits<__bf16> {
    using vec8 = bf16x8;
    static constexpr int mantissa = 8;
    __device__ static f32x16 mfma(vec8 a, vec8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

constexpr double kTdRhoTol = 2.5e-6;

// Exact FP64 two-pass Pearson of one window (single lane).
__device__ double exact_pearson(const float *raw, int i, gptr<const double> t, int Lt)
{
    double s = 0;
    for (int m = 0; m < Lt; ++m) { s += static_cast<double>(raw[i + m]); }
    const double mean = s / Lt;
    double e2 = 0, dn = 0;
    for (int m = 0; m < Lt; ++m) {
        const double d = static_cast<double>(raw[i + m]) - mean;
        e2 += d * d;
        dn += t[m] * d;
    }
    double rho = dn / fmax(DBL_EPSILON, sqrt(e2));
    if (fabs(rho) > 1.1) { rho = 0; }
    return rho;
}

template <typename MT, int KT>
__global__ __launch_bounds__(NT, 2) void tdcorr_kernel(const TdItem *__restrict__ items,
                                                       const Seg *__restrict__ segs,
                                                       const TdRow *__restrict__ rows)
{
    using Tr = TdTraits<MT>;
    using V8 = typename Tr::vec8;
    constexpr int NS = KT / 16;        // k-steps of 16
    constexpr int SPAN = IS + KT;      // samples one item reads
    constexpr int SPT = (SPAN + NT - 1) / NT;
    __shared__ __align__(16) MT xh8[SPAN][8];  // entry p: x~[p .. p+7], hi part
    __shared__ __align__(16) MT xl8[SPAN][8];  //                        lo part
    __shared__ float raw[SPAN];                // samples (FP32), for exact windows
    __shared__ double P1[SPAN + 1];            // prefix sums of d = x - a
    __shared__ double P2[SPAN + 1];            //   and of d^2
    __shared__ double red[NT / 64];
    __shared__ double wsum1[NT / 64], wsum2[NT / 64];

    const TdItem item = items[blockIdx.x];
    const Seg sg = segs[item.seg];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar row loads and branches
    const int lane = tid & 63;
    const int half = lane >> 5;
    const int col = lane & 31;
    const long long base = item.seg0;

    if (sg.dead) {
        for (int r = 0; r < item.nRows; ++r) {
            const TdRow row = rows[item.rowBegin + r];
            const int nv = static_cast<int>(min(static_cast<long long>(IS), row.nOut - base));
            gptr<float> out = to_global(static_cast<float *>(row.out)) + base;
            for (int i = tid; i < nv; i += NT) { out[i] = 0.0f; }
        }
        return;
    }

    // 1. samples, anchor and scale
    const gptr<const float> x = to_global(static_cast<const float *>(sg.x));
    const int nSamp = static_cast<int>(min(static_cast<long long>(SPAN), sg.n - base));
    float v[SPT];
    double s = 0;
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;  // contiguous run per thread (for the scan)
        const float val = (p < nSamp) ? x[base + p] : 0.0f;
        v[q] = val;
        if (p < SPAN) { raw[p] = val; }
        s += static_cast<double>(val);
    }
    const double a = block_sum<NT>(s, red) / nSamp;
    double e = 0, l1 = 0, l2 = 0;
    double d1[SPT], d2[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        const double d = (p < nSamp) ? static_cast<double>(v[q]) - a : 0.0;
        d1[q] = l1;  // exclusive local prefix
        d2[q] = l2;
        l1 += d;
        l2 += d * d;
        e += d * d;
    }
    const double energy = block_sum<NT>(e, red);
    const double sc = energy > 0 ? sqrt(energy / nSamp) : 1.0;
    const float inv = static_cast<float>(1.0 / sc);
    // 2. split samples into the 8-wide Hankel entries: sample p goes to
    //    entry p - q, slot q (q = 0..7)
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        if (p < SPAN) {
            const float xs = (p < nSamp) ? static_cast<float>((static_cast<double>(v[q]) - a) * inv) : 0.0f;
            const MT hi = static_cast<MT>(xs);
            const MT lo = static_cast<MT>(xs - static_cast<float>(hi));
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (p - k >= 0) {
                    xh8[p - k][k] = hi;
                    xl8[p - k][k] = lo;
                }
            }
        }
    }
    // entries past the span read zeros in their upper slots
    for (int p = SPAN - 7 + tid; p < SPAN; p += NT) {
        for (int k = SPAN - p; k < 8; ++k) {
            xh8[p][k] = static_cast<MT>(0.0f);
            xl8[p][k] = static_cast<MT>(0.0f);
        }
    }
    // 3. FP64 prefix sums P[p] = sum_{q < p} d_q (thread runs, then waves)
    double x1 = l1, x2 = l2;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y1 = __shfl_up(x1, o, 64);
        const double y2 = __shfl_up(x2, o, 64);
        if (lane >= o) {
            x1 += y1;
            x2 += y2;
        }
    }
    if (lane == 63) {
        wsum1[wave] = x1;
        wsum2[wave] = x2;
    }
    __syncthreads();
    double w1 = 0, w2 = 0;
    for (int w = 0; w < wave; ++w) {
        w1 += wsum1[w];
        w2 += wsum2[w];
    }
    const double e1 = w1 + x1 - l1, e2 = w2 + x2 - l2;  // exclusive prefix of this thread's run
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int p = tid * SPT + q;
        if (p < SPAN) {
            P1[p] = e1 + d1[q];
            P2[p] = e2 + d2[q];
        }
    }
    if (tid == NT - 1) {
        P1[SPAN] = w1 + x1;
        P2[SPAN] = w2 + x2;
    }

    // 4. this wave's 32 rows: split taps in registers (A operand: row =
    //    lane % 32, taps 16 s + 8 half + 0..7)
    const int rowBase = item.rowBegin + 32 * wave;
    const int nMine = min(32, item.nRows - 32 * wave);
    V8 ah[NS], al[NS];
    {
        const bool has = col < nMine;
        const MT *f = has ? static_cast<const MT *>(rows[rowBase + col].frag) : nullptr;
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            if (has) {
                const gptr<const i32x4> fh = to_global(reinterpret_cast<const i32x4 *>(f));
                ah[st] = __builtin_bit_cast(V8, fh[2 * st + half]);
                al[st] = __builtin_bit_cast(V8, fh[2 * NS + 2 * st + half]);
            } else {
                ah[st] = __builtin_bit_cast(V8, i32x4{0, 0, 0, 0});
                al[st] = __builtin_bit_cast(V8, i32x4{0, 0, 0, 0});
            }
        }
    }
    __syncthreads();  // Hankel entries, prefix sums and raw samples ready
    if (nMine <= 0) { return; }

    constexpr double epsSplit = 3.0 * (1.0 / (1u << Tr::mantissa)) * (1.0 / (1u << Tr::mantissa));
    const double epsM = epsSplit + 4.0 * sqrt(static_cast<double>(NS)) * 5.9604644775390625e-08;
    const double tol2 = kTdRhoTol * kTdRhoTol;
    const double eps2 = epsM * epsM;

    // 5. output tiles: D[j][i], lane holds column i = col and rows
    //    j = 8 (r / 4) + 4 half + r % 4
    for (int tile = 0; tile < NTILE_I; ++tile) {
        const int i0 = tile * 32;
        f32x16 acc0 = {}, acc1 = {};
        const i32x4 *bh = reinterpret_cast<const i32x4 *>(&xh8[i0 + col + 8 * half][0]);
        const i32x4 *bl = reinterpret_cast<const i32x4 *>(&xl8[i0 + col + 8 * half][0]);
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const V8 b_hi = __builtin_bit_cast(V8, bh[16 * st]);  // entry + 16 st
            const V8 b_lo = __builtin_bit_cast(V8, bl[16 * st]);
            acc0 = Tr::mfma(ah[st], b_hi, acc0);
            acc1 = Tr::mfma(ah[st], b_lo, acc1);
            acc1 = Tr::mfma(al[st], b_hi, acc1);
        }
        // epilogue: normalise, flag, store
        const int i = i0 + col;
        const long long gi = base + i;
        const double p1i = P1[i], p2i = P2[i];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j = 8 * (r >> 2) + 4 * half + (r & 3);
            if (j >= nMine) { continue; }
            const TdRow row = rows[rowBase + j];
            if (gi >= row.nOut || i >= IS) { continue; }
            const int Lt = row.Lt;
            const double S1 = P1[i + Lt] - p1i;
            const double S2 = P2[i + Lt] - p2i;
            const double den2 = Lt * S2 - S1 * S1;
            const double num = static_cast<double>(acc0[r]) + static_cast<double>(acc1[r]);
            const bool canc = !(den2 > 2.0e-9 * static_cast<double>(Lt) * Lt * S2 + 1.0e-12 * Lt * P2[SPAN]);
            const bool noisy = eps2 * S2 * Lt > tol2 * den2;
            float rho;
            if (!canc && !noisy) {
                const float d2 = static_cast<float>(den2);
                float y = __builtin_amdgcn_rsqf(d2);
                y = y * (1.5f - 0.5f * d2 * y * y);
                rho = static_cast<float>(num * sc) * (sqrtf(static_cast<float>(Lt)) * y);
            } else {
                rho = static_cast<float>(exact_pearson(raw, i, to_global(row.tmpl), Lt));
            }
            to_global(static_cast<float *>(row.out))[gi] = rho;
        }
    }
}

// Split templates: frag[row] = [hi taps 0..KT) , [lo taps 0..KT), zero past Lt.
template <typename MT>
__global__ __launch_bounds__(NT) void td_prep_kernel(const TdPrepJob *__restrict__ jobs, int KT)
{
    const TdPrepJob job = jobs[blockIdx.x];
    MT *f = static_cast<MT *>(job.frag);
    for (int k = threadIdx.x; k < KT; k += NT) {
        const float t = k < job.Lt ? static_cast<float>(job.tmpl[k]) : 0.0f;
        const MT hi = static_cast<MT>(t);
        const double rest = (k < job.Lt ? job.tmpl[k] : 0.0) - static_cast<double>(static_cast<float>(hi));
        f[k] = hi;
        f[KT + k] = static_cast<MT>(static_cast<float>(rest));
    }
}

}  // namespace

bool td_supported(int nb) { return nb >= 1 && nb <= kTdMaxTaps; }

int td_taps(int nb) { return nb <= 64 ? 64 : (nb <= 128 ? 128 : 256); }

hipError_t launch_td_prep(int kind, const TdPrepJob *dJobs, int nJobs, int KT, hipStream_t stream)
{
    if (nJobs <= 0) { return hipSuccess; }
    if (kind == kTdFp16) {
        hipLaunchKernelGGL(td_prep_kernel<_Float16>, dim3(nJobs), dim3(NT), 0, stream, dJobs, KT);
    } else {
        hipLaunchKernelGGL(td_prep_kernel<__bf16>, dim3(nJobs), dim3(NT), 0, stream, dJobs, KT);
    }
    return hipGetLastError();
}

hipError_t launch_td(int kind, int KT, const TdItem *dItems, int nItems, const Seg *dSegs, const TdRow *dRows,
                     hipStream_t stream)
{
    if (nItems <= 0) { return hipSuccess; }
#define MF_TD_LAUNCH(MT, K) hipLaunchKernelGGL((tdcorr_kernel<MT, K>), dim3(nItems), dim3(NT), 0, stream, dItems, dSegs, dRows)
    if (kind == kTdFp16) {
        switch (KT) {
        case 64: MF_TD_LAUNCH(_Float16, 64); break;
        case 128: MF_TD_LAUNCH(_Float16, 128); break;
        case 256: MF_TD_LAUNCH(_Float16, 256); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (KT) {
        case 64: MF_TD_LAUNCH(__bf16, 64); break;
        case 128: MF_TD_LAUNCH(__bf16, 128); break;
        case 256: MF_TD_LAUNCH(__bf16, 256); break;
        default: return hipErrorInvalidValue;
        }
    }
#undef MF_TD_LAUNCH
    return hipGetLastError();
}

}  // namespace mfk

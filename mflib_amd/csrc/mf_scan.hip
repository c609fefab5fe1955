// Window scales from prefix scans (gfx950): the standalone form of
// normalizeSignal (src/matchedFilter.cpp:94-166) for precision studies.
//
// For a FP32 signal x and a template length Lt, every window
// w_i = x[i .. i + Lt) gets the Pearson scale
//     sqrt(Lt) / sqrt(Lt * S2 - S1^2),   S1 = sum w_i,  S2 = sum w_i^2,
// (0 when the bracket is not positive).  The sums come from prefix scans of
// the anchored samples d = x - a (a = the first sample of the block's span),
// S = P[i + Lt] - P[i], built with wavefront shuffle scans:
//   - FP64 scans (kScanF64): d and d^2 in FP64;
//   - compensated FP32 scans (kScanF32c): d = x - a as an exact pair
//     (TwoSum), d^2 as an exact pair (TwoProduct by FMA) plus the cross
//     term, and every addition of the scans and of the final bracket in
//     float-float arithmetic (Dekker/Knuth TwoSum and FastTwoSum).
// One block covers kScanChunk outputs; the span of samples it scans is
// kScanChunk + Lt - 1 <= kScanSpan.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "mf_device.hpp"
#include "mf_kernels.hpp"

using namespace mfdev;

namespace mfk {

namespace {

constexpr int NT = kThreads;
constexpr int kScanSpan = kScanChunk + kScanMaxLt - 1;
constexpr int kScanGroups = (kScanSpan + 63) / 64;  // 64-sample groups in a span

// ---- float-float arithmetic --------------------------------------------
// No contraction in this file: fusing a product into a later sum (say the
// rounded square of fast_two_sum's first operand) double-counts the
// rounding error the pair already carries.  The one FMA the pairs need
// (two_prod's exact error term) is written out.
#pragma clang fp contract(off)

struct FF {
    float h, l;
};
__device__ __forceinline__ FF two_sum(float a, float b)
{
    const float s = a + b;
    const float v = s - a;
    const float e = (a - (s - v)) + (b - v);
    return {s, e};
}
__device__ __forceinline__ FF fast_two_sum(float a, float b)  // |a| >= |b|
{
    const float s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ FF ff_add(FF a, FF b)
{
    FF s = two_sum(a.h, b.h);
    FF t = two_sum(a.l, b.l);
    s.l += t.h;
    s = fast_two_sum(s.h, s.l);
    s.l += t.l;
    return fast_two_sum(s.h, s.l);
}
__device__ __forceinline__ FF ff_neg(FF a) { return {-a.h, -a.l}; }
__device__ __forceinline__ FF two_prod(float a, float b)
{
    const float p = a * b;
    return {p, __builtin_fmaf(a, b, -p)};
}
__device__ __forceinline__ FF ff_mul(FF a, FF b)
{
    FF p = two_prod(a.h, b.h);
    p.l += a.h * b.l + a.l * b.h;
    return fast_two_sum(p.h, p.l);
}

// ---- scan element types ---------------------------------------------------
template <int MODE>
struct ScanT;
template <>
struct ScanT<kScanF64> {
    using V = double;
    __device__ static V zero() { return 0.0; }
    __device__ static V add(V a, V b) { return a + b; }
    __device__ static V sub(V a, V b) { return a - b; }
    __device__ static V shfl_up(V a, int o) { return __shfl_up(a, o, 64); }
    __device__ static V shfl(V a, int l) { return __shfl(a, l, 64); }
    __device__ static void terms(float x, float a, V &d, V &d2)
    {
        d = static_cast<double>(x) - static_cast<double>(a);  // exact
        d2 = d * d;
    }
    // sqrt(Lt) / sqrt(Lt S2 - S1^2), or 0
    __device__ static float scale(V s1, V s2, int Lt)
    {
        const double den2 = Lt * s2 - s1 * s1;
        return den2 > 0.0 ? static_cast<float>(sqrt(static_cast<double>(Lt) / den2)) : 0.0f;
    }
};
template <>
struct ScanT<kScanF32c> {
    using V = FF;
    __device__ static V zero() { return {0.0f, 0.0f}; }
    __device__ static V add(V a, V b) { return ff_add(a, b); }
    __device__ static V sub(V a, V b) { return ff_add(a, ff_neg(b)); }
    __device__ static V shfl_up(V a, int o) { return {__shfl_up(a.h, o, 64), __shfl_up(a.l, o, 64)}; }
    __device__ static V shfl(V a, int l) { return {__shfl(a.h, l, 64), __shfl(a.l, l, 64)}; }
    __device__ static void terms(float x, float a, V &d, V &d2)
    {
        d = two_sum(x, -a);                 // x - a exactly
        d2 = two_prod(d.h, d.h);            // d.h^2 exactly
        d2.l += 2.0f * d.h * d.l;           // cross term (d.l^2 is below the pair's precision)
        d2 = fast_two_sum(d2.h, d2.l);
    }
    __device__ static float scale(V s1, V s2, int Lt)
    {
        const FF lt = {static_cast<float>(Lt), 0.0f};
        const FF den2 = ff_add(ff_mul(lt, s2), ff_neg(ff_mul(s1, s1)));
        const float d = den2.h + den2.l;
        return d > 0.0f ? sqrtf(static_cast<float>(Lt) / d) : 0.0f;
    }
};

// Inclusive scan over the 64 lanes of a wave (Hillis-Steele, fixed order).
template <int MODE>
__device__ __forceinline__ typename ScanT<MODE>::V wave_scan(typename ScanT<MODE>::V v, int lane)
{
    using S = ScanT<MODE>;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const typename S::V u = S::shfl_up(v, o);
        if (lane >= o) { v = S::add(u, v); }
    }
    return v;
}

template <int MODE>
__global__ __launch_bounds__(NT) void scan_scale_kernel(const float *__restrict__ x, long long n, int Lt,
                                                        float *__restrict__ scale)
{
    using S = ScanT<MODE>;
    using V = typename S::V;
    // P1/P2[j] = sum of the first j anchored samples of the span
    __shared__ V P1[kScanSpan + 1];
    __shared__ V P2[kScanSpan + 1];
    __shared__ V g1[kScanGroups];
    __shared__ V g2[kScanGroups];

    const long long nOut = n - Lt + 1;
    const long long c0 = static_cast<long long>(blockIdx.x) * kScanChunk;
    const int nOutBlk = static_cast<int>(min(static_cast<long long>(kScanChunk), nOut - c0));
    const int span = nOutBlk + Lt - 1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const gptr<const float> xg = to_global(x) + c0;
    const float a = xg[0];

    // 1. per 64-sample group: wave scan of (d, d^2), group totals
    for (int g = wave; g * 64 < span; g += NT / 64) {
        const int j = g * 64 + lane;
        V d = S::zero(), d2 = S::zero();
        if (j < span) { S::terms(xg[j], a, d, d2); }
        d = wave_scan<MODE>(d, lane);
        d2 = wave_scan<MODE>(d2, lane);
        if (j < span) {
            P1[j + 1] = d;
            P2[j + 1] = d2;
        }
        if (lane == 63) {
            g1[g] = d;
            g2[g] = d2;
        }
    }
    if (tid == 0) {
        P1[0] = S::zero();
        P2[0] = S::zero();
    }
    __syncthreads();
    // 2. exclusive scan of the group totals (wave 0, one group per lane)
    static_assert(kScanGroups <= 64, "group count");
    if (wave == 0) {
        const int nG = (span + 63) / 64;
        const V t1 = lane < nG ? g1[lane] : S::zero();
        const V t2 = lane < nG ? g2[lane] : S::zero();
        const V i1 = wave_scan<MODE>(t1, lane), i2 = wave_scan<MODE>(t2, lane);
        const V e1 = S::shfl_up(i1, 1), e2 = S::shfl_up(i2, 1);
        if (lane < nG) {
            g1[lane] = lane > 0 ? e1 : S::zero();
            g2[lane] = lane > 0 ? e2 : S::zero();
        }
    }
    __syncthreads();
    // 3. add the group offsets
    for (int j = tid; j < span; j += NT) {
        const int g = j >> 6;
        P1[j + 1] = S::add(g1[g], P1[j + 1]);
        P2[j + 1] = S::add(g2[g], P2[j + 1]);
    }
    __syncthreads();
    // 4. window sums and scales
    const gptr<float> out = to_global(scale) + c0;
    for (int i = tid; i < nOutBlk; i += NT) {
        const V s1 = S::sub(P1[i + Lt], P1[i]);
        const V s2 = S::sub(P2[i + Lt], P2[i]);
        out[i] = S::scale(s1, s2, Lt);
    }
}

}  // namespace

hipError_t launch_scan_scale(int mode, const float *dx, long long n, int Lt, float *dScale, hipStream_t stream)
{
    if (Lt < 2 || Lt > kScanMaxLt || n < Lt) { return hipErrorInvalidValue; }
    const long long nOut = n - Lt + 1;
    const long long grid = (nOut + kScanChunk - 1) / kScanChunk;
    if (grid > 0x7fffffffLL) { return hipErrorInvalidValue; }
    if (mode == kScanF64) {
        hipLaunchKernelGGL(scan_scale_kernel<kScanF64>, dim3(static_cast<unsigned>(grid)), dim3(NT), 0, stream, dx,
                           n, Lt, dScale);
    } else if (mode == kScanF32c) {
        hipLaunchKernelGGL(scan_scale_kernel<kScanF32c>, dim3(static_cast<unsigned>(grid)), dim3(NT), 0, stream,
                           dx, n, Lt, dScale);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mfk

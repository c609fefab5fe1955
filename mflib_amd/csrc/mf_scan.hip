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
// A block covers a chunk of outputs whose span of samples (chunk + Lt - 1)
// fits one pass of K (8 or 16) consecutive samples per thread: a sequential prefix
// in registers per thread, a wave shuffle scan of the thread totals and the
// wave totals in LDS, then the prefix table in LDS for the window sums.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "mf_device.hpp"
#include "mf_kernels.hpp"

using namespace mfdev;

namespace mfk {

namespace {

constexpr int NT = kThreads;

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

// sqrt(Lt) / sqrt(den2) from the bracket rounded to FP32: hardware rsqrt
// plus one Newton step (relative error ~2^-23, the output's precision), 0
// when the bracket is not positive
__device__ __forceinline__ float scale_of(float den2, int Lt)
{
    if (!(den2 > 0.0f)) { return 0.0f; }
    float y = __builtin_amdgcn_rsqf(den2);
    y = y * (1.5f - 0.5f * den2 * y * y);
    return sqrtf(static_cast<float>(Lt)) * y;
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
        return scale_of(static_cast<float>(Lt * s2 - s1 * s1), Lt);
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
        return scale_of(den2.h + den2.l, Lt);
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

// Outputs per block for template lengths up to MAXLT: the block scans one
// span of NT * K samples (K = 8 for MAXLT 256, so 4 blocks fit a CU; 16
// for MAXLT 2048), so chunk + MAXLT - 1 <= NT * K; the chunk is a multiple
// of 4 so every block's first sample is 16-byte aligned when x is.
template <int MAXLT>
struct ScanChunk {
    static constexpr int K = MAXLT <= 256 ? 8 : 16;  // consecutive samples per thread
    static constexpr int value = ((NT * K - MAXLT + 1) / 4) * 4;
};

// padded LDS index: one spare slot per 16 entries, so the 16 consecutive
// entries a thread writes sit in distinct banks across a quarter-wave
__device__ __forceinline__ int spad(int i) { return i + (i >> 4); }

template <int MODE, int MAXLT, bool VEC>
__global__ __launch_bounds__(NT) void scan_scale_kernel(const float *__restrict__ x, long long n, int Lt,
                                                        float *__restrict__ scale)
{
    using S = ScanT<MODE>;
    using V = typename S::V;
    constexpr int K = ScanChunk<MAXLT>::K;
    constexpr int SPAN = NT * K;
    constexpr int CH = ScanChunk<MAXLT>::value;
    static_assert(CH + MAXLT - 1 <= SPAN, "span");
    // P1/P2[spad(j)] = sum of the first j anchored samples of the span
    __shared__ V P1[SPAN + SPAN / 16 + 1];
    __shared__ V P2[SPAN + SPAN / 16 + 1];
    __shared__ V w1[NT / 64];
    __shared__ V w2[NT / 64];

    const long long nOut = n - Lt + 1;
    const long long c0 = static_cast<long long>(blockIdx.x) * CH;
    const int nOutBlk = static_cast<int>(min(static_cast<long long>(CH), nOut - c0));
    const int span = nOutBlk + Lt - 1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const gptr<const float> xg = to_global(x) + c0;
    const float a = xg[0];

    // 1. the thread's K consecutive samples
    const int j0 = K * tid;
    float xv[K];
    if (VEC && j0 + K <= span) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) f4 *p = (const __attribute__((address_space(1))) f4 *)(xg + j0);
#pragma unroll
        for (int m = 0; m < K / 4; ++m) {
            const f4 v = p[m];
            xv[4 * m] = v.x;
            xv[4 * m + 1] = v.y;
            xv[4 * m + 2] = v.z;
            xv[4 * m + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) { xv[k] = (j0 + k < span) ? xg[j0 + k] : 0.0f; }
    }
    // 2. sequential inclusive prefix over the thread's samples
    V p1[K], p2[K];
    V r1 = S::zero(), r2 = S::zero();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        V d = S::zero(), d2 = S::zero();
        if (j0 + k < span) { S::terms(xv[k], a, d, d2); }
        r1 = S::add(r1, d);
        r2 = S::add(r2, d2);
        p1[k] = r1;
        p2[k] = r2;
    }
    // 3. exclusive offset of the thread: wave shuffle scan of the thread
    // totals, then the totals of the waves before this one
    const V i1 = wave_scan<MODE>(r1, lane), i2 = wave_scan<MODE>(r2, lane);
    const V e1 = S::shfl_up(i1, 1), e2 = S::shfl_up(i2, 1);
    if (lane == 63) {
        w1[wave] = i1;
        w2[wave] = i2;
    }
    __syncthreads();
    V o1 = S::zero(), o2 = S::zero();
    for (int w = 0; w < wave; ++w) {
        o1 = S::add(o1, w1[w]);
        o2 = S::add(o2, w2[w]);
    }
    if (lane > 0) {
        o1 = S::add(o1, e1);
        o2 = S::add(o2, e2);
    }
    // 4. the block's prefix table
#pragma unroll
    for (int k = 0; k < K; ++k) {
        P1[spad(j0 + k + 1)] = S::add(o1, p1[k]);
        P2[spad(j0 + k + 1)] = S::add(o2, p2[k]);
    }
    if (tid == 0) {
        P1[0] = S::zero();
        P2[0] = S::zero();
    }
    __syncthreads();
    // 5. window sums and scales (coalesced stores)
    const gptr<float> out = to_global(scale) + c0;
    for (int i = tid; i < nOutBlk; i += NT) {
        const V s1 = S::sub(P1[spad(i + Lt)], P1[spad(i)]);
        const V s2 = S::sub(P2[spad(i + Lt)], P2[spad(i)]);
        out[i] = S::scale(s1, s2, Lt);
    }
}

template <int MODE, int MAXLT>
hipError_t scan_launch(const float *dx, long long n, int Lt, float *dScale, hipStream_t stream)
{
    constexpr int CH = ScanChunk<MAXLT>::value;
    const long long grid = (n - Lt + 1 + CH - 1) / CH;
    if (grid > 0x7fffffffLL) { return hipErrorInvalidValue; }
    const dim3 g(static_cast<unsigned>(grid));
    if ((reinterpret_cast<uintptr_t>(dx) & 15u) == 0) {
        hipLaunchKernelGGL((scan_scale_kernel<MODE, MAXLT, true>), g, dim3(NT), 0, stream, dx, n, Lt, dScale);
    } else {
        hipLaunchKernelGGL((scan_scale_kernel<MODE, MAXLT, false>), g, dim3(NT), 0, stream, dx, n, Lt, dScale);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_scan_scale(int mode, const float *dx, long long n, int Lt, float *dScale, hipStream_t stream)
{
    if (Lt < 2 || Lt > kScanMaxLt || n < Lt) { return hipErrorInvalidValue; }
    const bool small = Lt <= 256;
    if (mode == kScanF64) {
        return small ? scan_launch<kScanF64, 256>(dx, n, Lt, dScale, stream)
                     : scan_launch<kScanF64, kScanMaxLt>(dx, n, Lt, dScale, stream);
    }
    if (mode == kScanF32c) {
        return small ? scan_launch<kScanF32c, 256>(dx, n, Lt, dScale, stream)
                     : scan_launch<kScanF32c, kScanMaxLt>(dx, n, Lt, dScale, stream);
    }
    return hipErrorInvalidValue;
}

}  // namespace mfk

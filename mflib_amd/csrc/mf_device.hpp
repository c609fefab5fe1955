// Device building blocks for the matched-filter kernels (gfx950 / CDNA4).
//
// FFTs are Stockham auto-sort transforms executed in LDS: each pass reads
// a thread's butterflies into registers, synchronises, applies twiddles and
// a radix-4 (or a final radix-2) DFT and writes back in place.  Real
// transforms of N = 2M samples use one M-point complex transform plus the
// even/odd packing identities; the correlation kernel folds the unpacking,
// the product with the template spectrum and the inverse packing into one
// LDS pass.
//
// Twiddles W_{2M}^e = exp(-2 pi i e / 2M) come from a two-level table
// (64 low entries x 2M/64 high entries, computed in FP64 on the host with
// the C library's cos/sin and rounded to T), so one complex product gives
// any twiddle to about one ulp.
#pragma once
#include <hip/hip_runtime.h>

#include "mf_codelets.hpp"

namespace mfdev {

template <typename T>
using Cx = mfcl::C2<T>;

// Pointers read from descriptor tables are generic; accessing global
// memory through them would compile to FLAT instructions, which also count
// against the LDS wait counter.  These casts make them global.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p)
{
    return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<const T> to_global(const T *p)
{
    return (gptr<const T>)p;
}
// Complex element i of a global array (component loads merge into one
// 2-wide access).
template <typename T>
__device__ __forceinline__ Cx<T> load_cx(gptr<const Cx<T>> p, int i)
{
    const gptr<const T> s = (gptr<const T>)p;
    return Cx<T>{s[2 * i], s[2 * i + 1]};
}

// Raw buffer resource over `bytes` bytes from p (stride 0; dword 3 selects
// 32-bit data, as for every gfx9 raw buffer access).  Stores through it at
// byte offsets >= bytes are dropped by the hardware range check, so a
// store needs no branch of its own and every path issues the same
// instructions.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(void *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, 0x00020000);
}
__device__ __forceinline__ void buffer_store_pair(float x, float y, __amdgpu_buffer_rsrc_t r, int byteOff)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 d = {__float_as_uint(x), __float_as_uint(y)};
    __builtin_amdgcn_raw_buffer_store_b64(d, r, byteOff, 0, 0);
}
__device__ __forceinline__ void buffer_store_pair(double x, double y, __amdgpu_buffer_rsrc_t r, int byteOff)
{
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const unsigned long long a = static_cast<unsigned long long>(__double_as_longlong(x));
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(y));
    const u4 d = {static_cast<unsigned>(a), static_cast<unsigned>(a >> 32), static_cast<unsigned>(b),
                  static_cast<unsigned>(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(d, r, byteOff, 0, 0);
}

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ Cx<T> cconj(Cx<T> a) { return {a.x, -a.y}; }
// i * a
template <typename T>
__device__ __forceinline__ Cx<T> cmuli(Cx<T> a) { return {-a.y, a.x}; }
// -i * a
template <typename T>
__device__ __forceinline__ Cx<T> cmulmi(Cx<T> a) { return {a.y, -a.x}; }

// Two-level twiddle table resident in LDS.
template <typename T>
struct Twiddle {
    const Cx<T> *lo;  // W_{2M}^i,       i < 64
    const Cx<T> *hi;  // W_{2M}^(64 j),  j < 2M/64
    __device__ __forceinline__ Cx<T> operator()(int e) const { return cmul(lo[e & 63], hi[e >> 6]); }
};

template <int R, bool INV, typename T>
__device__ __forceinline__ void dft(Cx<T> *v)
{
    if constexpr (R == 2) {
        Cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {
        static_assert(R == 4, "radix");
        Cx<T> t0 = cadd(v[0], v[2]);
        Cx<T> t1 = csub(v[0], v[2]);
        Cx<T> t2 = cadd(v[1], v[3]);
        Cx<T> d = csub(v[1], v[3]);
        Cx<T> t3 = INV ? cmuli(d) : cmulmi(d);
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, t3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, t3);
    }
}

// One Stockham pass of radix R over M points with input stride Ns.
// Every thread of the block must call it (it synchronises twice).
template <typename T, int M, int R, int NT, bool INV>
__device__ __forceinline__ void stockham_pass(Cx<T> *buf, int Ns, Twiddle<T> tw)
{
    constexpr int NB = M / R;
    constexpr int BPT = (NB + NT - 1) / NT;
    const int t = threadIdx.x;
    Cx<T> v[BPT][R];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int j = t + b * NT;
        if (NB % NT == 0 || j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) { v[b][r] = buf[j + r * NB]; }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int j = t + b * NT;
        if (NB % NT == 0 || j < NB) {
            const int k = j & (Ns - 1);
            // W_{Ns R}^{r k} = W_{2M}^{r * 2k M / (Ns R)}
            const int step = 2 * (M / (Ns * R)) * k;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                Cx<T> w = tw(r * step);
                if (INV) { w.y = -w.y; }
                v[b][r] = cmul(v[b][r], w);
            }
            dft<R, INV>(v[b]);
            const int idxD = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) { buf[idxD + r * Ns] = v[b][r]; }
        }
    }
    __syncthreads();
}

template <int M>
struct Log2 {
    static constexpr int value = 1 + Log2<M / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

// Full M-point complex FFT in LDS (natural order in and out, unnormalised).
template <typename T, int M, int NT, bool INV>
__device__ __forceinline__ void fft_lds(Cx<T> *buf, Twiddle<T> tw)
{
    constexpr int LOG = Log2<M>::value;
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < LOG / 2; ++p) {
        stockham_pass<T, M, 4, NT, INV>(buf, Ns, tw);
        Ns *= 4;
    }
    if constexpr ((LOG & 1) != 0) { stockham_pass<T, M, 2, NT, INV>(buf, Ns, tw); }
}

// A value every lane holds the same copy of, moved to scalar registers
// (block-wide results read back from LDS land in vector registers).
__device__ __forceinline__ double uniform(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// ---- block reductions -------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { v += __shfl_xor(v, o, 64); }
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { v = fmax(v, __shfl_xor(v, o, 64)); }
    return v;
}

// Sum over the block; `scratch` holds NT/64 doubles.  All threads get the
// result.  Deterministic: fixed shuffle tree, then waves in order.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *scratch)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { scratch[w] = v; }
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) { s += scratch[i]; }
    __syncthreads();
    return s;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double *scratch)
{
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { scratch[w] = v; }
    __syncthreads();
    double s = scratch[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) { s = fmax(s, scratch[i]); }
    __syncthreads();
    return s;
}

// ---- wave-level FFTs ----------------------------------------------------
// One wave (64 lanes) transforms M = 64 Q points held as Q complex values
// per lane.  Three register passes (radices R1, R2, R3, R1 R2 R3 = M) with
// two exchanges through a wave-private, padded LDS buffer; no block
// barrier is involved.  Element r of butterfly b of a pass lives in
// v[b R + r]; butterfly j = lane + 64 b reads points j + r M/R.
template <int M>
struct WavePlan;
template <>
struct WavePlan<512> {
    static constexpr int R1 = 8, R2 = 8, R3 = 8;
};
template <>
struct WavePlan<1024> {
    static constexpr int R1 = 16, R2 = 16, R3 = 4;
};
template <>
struct WavePlan<2048> {
    static constexpr int R1 = 16, R2 = 16, R3 = 8;
};

// Padded exchange index: one spare slot every 16 points breaks the
// power-of-two strides of the Stockham stores across LDS banks.
__device__ __forceinline__ int xpad(int i) { return i + (i >> 4); }
template <int M>
struct XBuf {
    static constexpr int size = M + M / 16;
};

// Orders a wave's LDS stores before its later loads (LDS operations of one
// wave complete in issue order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Exchange addresses split into a per-lane part and a compile-time part:
// xpad(lanePart + constPart) == xpad(lanePart) + xpad(constPart) for every
// access below (one of the two parts is a multiple of 16 or their low four
// bits never carry), so each access is one base register plus an
// immediate offset.
constexpr int cpad(int i) { return i + (i >> 4); }

template <typename T, int M, int R>
__device__ __forceinline__ void wave_load(Cx<T> *v, const Cx<T> *buf, int lane)
{
    constexpr int B = M / (R * 64);
    static_assert((M / R) % 16 == 0, "load offsets must be 16-aligned");
    const Cx<T> *p = buf + xpad(lane);
#pragma unroll
    for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int r = 0; r < R; ++r) { v[b * R + r] = p[cpad(64 * b + r * (M / R))]; }
    }
}

// Inter-pass twiddles of one lane.  They depend on the lane only (not on
// the data or the template row), so a block computes them once and every
// transform reuses them (conjugated for inverse transforms).
//   pass 2: butterfly j has k = j mod R1 = lane mod R1 (64 is a multiple
//           of R1), the same for all of the lane's butterflies;
//   pass 3: k = j mod R1 R2, which differs per butterfly when R1 R2 > 64.
// W3REG: pass-3 twiddle rows resident in registers (computed once per
// block); otherwise looked up from the two-level table per transform, which
// saves registers where they limit occupancy.
template <typename T, int M, bool W3REG = true>
struct WaveTwiddles {
    using P = WavePlan<M>;
    static constexpr int NS2 = P::R1;
    static constexpr int NS3 = P::R1 * P::R2;
    static constexpr int B3 = M / (P::R3 * 64);
    static constexpr int W3 = W3REG ? B3 : 1;
    Cx<T> w2[P::R2 - 1];  // pass 2, resident in registers
    Cx<T> w3[W3][P::R3 - 1];
    Twiddle<T> tab;       // pass 3, looked up per transform
    __device__ __forceinline__ void init(int lane, const Twiddle<T> &tw)
    {
        tab = tw;
        const int k2 = lane & (NS2 - 1);
        const int s2 = (2 * M / (NS2 * P::R2)) * k2;
#pragma unroll
        for (int r = 1; r < P::R2; ++r) { w2[r - 1] = tw(r * s2); }
        if constexpr (W3REG) {
#pragma unroll
            for (int b = 0; b < B3; ++b) { lookup3(lane, b, w3[b]); }
        }
    }
    __device__ __forceinline__ void lookup3(int lane, int b, Cx<T> *w) const
    {
        const int k3 = (lane + 64 * b) & (NS3 - 1);
        const int s3 = (2 * M / (NS3 * P::R3)) * k3;
#pragma unroll
        for (int r = 1; r < P::R3; ++r) { w[r - 1] = tab(r * s3); }
    }
    // pass-3 twiddle row of butterfly b of this lane
    __device__ __forceinline__ void pass3(int lane, int b, Cx<T> *w) const
    {
        if constexpr (W3REG) {
#pragma unroll
            for (int r = 1; r < P::R3; ++r) { w[r - 1] = w3[b][r - 1]; }
        } else {
            lookup3(lane, b, w);
        }
    }
};

// Butterflies of one pass with per-butterfly twiddle rows w[b][r-1]
// (wStride 0 when all butterflies share one row).
template <typename T, int M, int R, bool INV, bool TWIDDLE>
__device__ __forceinline__ void wave_butterflies(Cx<T> *v, const Cx<T> *w, int wStride)
{
    constexpr int B = M / (R * 64);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        if constexpr (TWIDDLE) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                Cx<T> t = w[b * wStride + r - 1];
                if (INV) { t.y = -t.y; }
                v[b * R + r] = cmul(v[b * R + r], t);
            }
        }
        mfcl::dftR<R, INV>(&v[b * R]);
    }
}

template <typename T, int M, int R, int NS>
__device__ __forceinline__ void wave_store(const Cx<T> *v, Cx<T> *buf, int lane)
{
    constexpr int B = M / (R * 64);
    static_assert(NS <= 64, "butterfly phase must depend on the lane only");
    // j = lane + 64 b, k = j mod NS = lane mod NS:
    // dst = (lane - k) R + k  +  64 b R + r NS
    const int k = lane & (NS - 1);
    Cx<T> *p = buf + xpad((lane - k) * R + k);
#pragma unroll
    for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int r = 0; r < R; ++r) { p[cpad(64 * b * R + r * NS)] = v[b * R + r]; }
    }
}

// Passes 1..3 on data already in registers as pass-1 input.  On return the
// lane holds output points j + r M/R3 (j = lane + 64 b) in v[b R3 + r].
template <typename T, int M, bool INV, typename W>
__device__ __forceinline__ void wave_fft_regs(Cx<T> *v, Cx<T> *buf, int lane, const W &tw)
{
    using P = WavePlan<M>;
    wave_butterflies<T, M, P::R1, INV, false>(v, nullptr, 0);
    wave_store<T, M, P::R1, 1>(v, buf, lane);
    wave_lds_sync();
    wave_load<T, M, P::R2>(v, buf, lane);
    wave_butterflies<T, M, P::R2, INV, true>(v, tw.w2, 0);
    wave_lds_sync();
    wave_store<T, M, P::R2, P::R1>(v, buf, lane);
    wave_lds_sync();
    wave_load<T, M, P::R3>(v, buf, lane);
#pragma unroll
    for (int b = 0; b < W::B3; ++b) {
        Cx<T> w3[P::R3 - 1];
        tw.pass3(lane, b, w3);
#pragma unroll
        for (int r = 1; r < P::R3; ++r) {
            Cx<T> t = w3[r - 1];
            if (INV) { t.y = -t.y; }
            v[b * P::R3 + r] = cmul(v[b * P::R3 + r], t);
        }
        mfcl::dftR<P::R3, INV>(&v[b * P::R3]);
    }
    wave_lds_sync();
}

// ---- half-size exchanges ------------------------------------------------
// The exchanges of a row transform go through a wave buffer of half the
// padded size (XHalf<M>::size), in two phases: first every value bound for
// points [0, M/2), which the lanes then load, then those bound for
// [M/2, M).  A half buffer holds point i + H M/2 at xpad(i), so the half-H
// base is buf - cpad(H M/2) and every access keeps its full-size index
// expression.  Halving the buffer is what lets more blocks share a CU's LDS.
template <int M>
struct XHalf {
    static constexpr int size = M / 2 + M / 32;
};

// Stores of a pass's outputs (radix R, butterfly phase NS) bound for half
// H.  With two or more butterflies per lane the half follows the butterfly
// index b; with one, the lane: lanes 0-31 write half 0 and 32-63 half 1.
template <typename T, int M, int R, int NS, int H>
__device__ __forceinline__ void wave_store_half(const Cx<T> *v, Cx<T> *buf, int lane)
{
    constexpr int B = M / (R * 64);
    static_assert(NS <= 64, "butterfly phase must depend on the lane only");
    const int k = lane & (NS - 1);
    Cx<T> *p = buf - cpad(H * (M / 2)) + xpad((lane - k) * R + k);
    if constexpr (B >= 2) {
#pragma unroll
        for (int b = H * (B / 2); b < (H + 1) * (B / 2); ++b) {
#pragma unroll
            for (int r = 0; r < R; ++r) { p[cpad(64 * b * R + r * NS)] = v[b * R + r]; }
        }
    } else {
        if ((lane >= 32) == (H == 1)) {
#pragma unroll
            for (int r = 0; r < R; ++r) { p[cpad(r * NS)] = v[r]; }
        }
    }
}

// Loads of the next pass's inputs (radix R) that lie in half H: input r of
// butterfly b is point lane + 64 b + r M/R, in half 0 exactly when r < R/2.
template <typename T, int M, int R, int H>
__device__ __forceinline__ void wave_load_half(Cx<T> *v, const Cx<T> *buf, int lane)
{
    constexpr int B = M / (R * 64);
    static_assert((M / R) % 16 == 0, "load offsets must be 16-aligned");
    const Cx<T> *p = buf - cpad(H * (M / 2)) + xpad(lane);
#pragma unroll
    for (int b = 0; b < B; ++b) {
#pragma unroll
        for (int r = H * (R / 2); r < (H + 1) * (R / 2); ++r) { v[b * R + r] = p[cpad(64 * b + r * (M / R))]; }
    }
}

// Pass outputs in v (radix RS, phase NS) to the next pass's inputs (radix
// RL) in v, through the half buffer.
template <typename T, int M, int RS, int NS, int RL>
__device__ __forceinline__ void wave_exchange_half(Cx<T> *v, Cx<T> *buf, int lane)
{
    constexpr int Q = M / 64;
    Cx<T> u[Q];
    wave_store_half<T, M, RS, NS, 0>(v, buf, lane);
    wave_lds_sync();
    wave_load_half<T, M, RL, 0>(u, buf, lane);
    wave_lds_sync();
    wave_store_half<T, M, RS, NS, 1>(v, buf, lane);
    wave_lds_sync();
    wave_load_half<T, M, RL, 1>(u, buf, lane);
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < Q; ++i) { v[i] = u[i]; }
}

// wave_fft_regs through a half buffer.
template <typename T, int M, bool INV, typename W>
__device__ __forceinline__ void wave_fft_regs_half(Cx<T> *v, Cx<T> *buf, int lane, const W &tw)
{
    using P = WavePlan<M>;
    wave_butterflies<T, M, P::R1, INV, false>(v, nullptr, 0);
    wave_exchange_half<T, M, P::R1, 1, P::R2>(v, buf, lane);
    wave_butterflies<T, M, P::R2, INV, true>(v, tw.w2, 0);
    wave_exchange_half<T, M, P::R2, P::R1, P::R3>(v, buf, lane);
#pragma unroll
    for (int b = 0; b < W::B3; ++b) {
        Cx<T> w3[P::R3 - 1];
        tw.pass3(lane, b, w3);
#pragma unroll
        for (int r = 1; r < P::R3; ++r) {
            Cx<T> t = w3[r - 1];
            if (INV) { t.y = -t.y; }
            v[b * P::R3 + r] = cmul(v[b * P::R3 + r], t);
        }
        mfcl::dftR<P::R3, INV>(&v[b * P::R3]);
    }
    wave_lds_sync();
}

// True when the Lt samples of the window at i are all equal (its exact
// value is then 0: zero numerator over the floor of the norm), from the
// equal-neighbour bits.
__device__ __forceinline__ bool constant_window(const unsigned *eqw, int i, int Lt)
{
    const int end = i + Lt - 1;  // bits [i, end)
    for (int p = i; p < end;) {
        const int b = p & 31;
        const int take = min(32 - b, end - p);
        const unsigned mask = (take == 32 ? ~0u : ((1u << take) - 1u)) << b;
        if ((eqw[p >> 5] & mask) != mask) { return false; }
        p += take;
    }
    return true;
}

}  // namespace mfdev

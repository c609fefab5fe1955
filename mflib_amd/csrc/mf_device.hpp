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

namespace mfdev {

template <typename T>
struct Cx {
    T x, y;
};

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

}  // namespace mfdev

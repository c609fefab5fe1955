// Register DFT codelets (radix 4, 8, 16) and the Stockham pass index rule.
//
// Host+device code so the codelets can be unit-tested on the CPU.  The
// constants are cos/sin of multiples of pi/8 (the 16th roots of unity).
#pragma once
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define MF_HD __host__ __device__ __forceinline__
#else
#define MF_HD inline
#endif

namespace mfcl {

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
MF_HD C2<T> add(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
MF_HD C2<T> sub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
MF_HD C2<T> mul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
MF_HD C2<T> conj(C2<T> a) { return {a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
MF_HD C2<T> rotq(C2<T> a)
{
    if constexpr (INV) { return {-a.y, a.x}; }
    return {a.y, -a.x};
}

// cos(k pi / 8) for k = 0..4
constexpr double kC0 = 1.0;
constexpr double kC1 = 0.92387953251128675613;  // cos(pi/8)
constexpr double kC2 = 0.70710678118654752440;  // cos(pi/4)
constexpr double kC3 = 0.38268343236508977173;  // cos(3pi/8)

// W16^e = exp(-+ 2 pi i e / 16) for e in [0, 16) (sign by INV)
template <bool INV, typename T>
MF_HD C2<T> w16(int e)
{
    // cos/sin of 2 pi e/16 from the first-quadrant constants
    const double cs[5] = {kC0, kC1, kC2, kC3, 0.0};
    const int q = e & 3;          // position inside the quadrant (units of pi/8)
    const int quad = (e >> 2) & 3;
    double c = cs[q], s = cs[4 - q];
    double cr, sr;
    switch (quad) {
    case 0: cr = c; sr = s; break;
    case 1: cr = -s; sr = c; break;
    case 2: cr = -c; sr = -s; break;
    default: cr = s; sr = -c; break;
    }
    // forward twiddle exp(-i theta) = (cos, -sin)
    return {static_cast<T>(cr), static_cast<T>(INV ? sr : -sr)};
}

template <bool INV, typename T>
MF_HD void dft4(C2<T> &a0, C2<T> &a1, C2<T> &a2, C2<T> &a3)
{
    const C2<T> t0 = add(a0, a2), t1 = sub(a0, a2);
    const C2<T> t2 = add(a1, a3), t3 = rotq<INV>(sub(a1, a3));
    a0 = add(t0, t2);
    a2 = sub(t0, t2);
    a1 = add(t1, t3);
    a3 = sub(t1, t3);
}

// 8-point DFT in place (natural order in and out).
template <bool INV, typename T>
MF_HD void dft8(C2<T> *v)
{
    C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const T h = static_cast<T>(kC2);
    // W8^1, W8^2, W8^3
    C2<T> p1, p2, p3;
    if constexpr (INV) {
        p1 = {h * (o1.x - o1.y), h * (o1.x + o1.y)};
        p2 = {-o2.y, o2.x};
        p3 = {-h * (o3.x + o3.y), h * (o3.x - o3.y)};
    } else {
        p1 = {h * (o1.x + o1.y), h * (o1.y - o1.x)};
        p2 = {o2.y, -o2.x};
        p3 = {h * (o3.y - o3.x), -h * (o3.x + o3.y)};
    }
    v[0] = add(e0, o0);
    v[4] = sub(e0, o0);
    v[1] = add(e1, p1);
    v[5] = sub(e1, p1);
    v[2] = add(e2, p2);
    v[6] = sub(e2, p2);
    v[3] = add(e3, p3);
    v[7] = sub(e3, p3);
}

// 16-point DFT in place: 4 x 4 decomposition (n = 4 n1 + n2, k = k1 + 4 k2).
template <bool INV, typename T>
MF_HD void dft16(C2<T> *v)
{
    C2<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        a[n2][0] = v[n2];
        a[n2][1] = v[4 + n2];
        a[n2][2] = v[8 + n2];
        a[n2][3] = v[12 + n2];
        dft4<INV>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1) { a[n2][k1] = mul(a[n2][k1], w16<INV, T>(n2 * k1)); }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C2<T> b0 = a[0][k1], b1 = a[1][k1], b2 = a[2][k1], b3 = a[3][k1];
        dft4<INV>(b0, b1, b2, b3);
        v[k1] = b0;
        v[k1 + 4] = b1;
        v[k1 + 8] = b2;
        v[k1 + 12] = b3;
    }
}

template <int R, bool INV, typename T>
MF_HD void dftR(C2<T> *v)
{
    if constexpr (R == 2) {
        const C2<T> a = v[0], b = v[1];
        v[0] = add(a, b);
        v[1] = sub(a, b);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        dft8<INV>(v);
    } else {
        static_assert(R == 16, "radix");
        dft16<INV>(v);
    }
}

// Stockham pass of radix R, stride Ns over M points: butterfly j reads
// in[j + r M/R], multiplies element r by W_{Ns R}^{r (j mod Ns)}, applies
// the R-point DFT and writes out[(j - j mod Ns) R + j mod Ns + r Ns].
MF_HD int stockhamDst(int j, int Ns, int R, int r)
{
    const int k = j & (Ns - 1);
    return (j - k) * R + k + r * Ns;
}

}  // namespace mfcl

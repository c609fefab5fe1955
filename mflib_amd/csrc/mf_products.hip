// Per-detection products (gfx950): what SingleChannel::Detector::detect
// derives for every peak once PeakFinder has picked it
// (src/singleChannel/detector.cpp:537-664):
//   - the correlation coefficient at the peak, clamped to [-1, 1] (574-575);
//   - the sub-sample time shift, by a parabola through three lags
//     (quadraticRefinement, 38-89) or by Lanczos-windowed sinc
//     interpolation on a grid of shifts (lanczosRefinement, 91-152);
//   - the amplitude scaling factors of the detected window against the
//     template, Gibbons & Ringdal (2006) and Schaff & Richards (2014)
//     (src/singleChannel/relativeMagnitude.cpp:35-58, 273-388), when the
//     window fits the signal (detector.cpp:578-585) and is not constant
//     (relativeMagnitude.cpp:330-338).
// One wavefront per detection.  Everything is FP64 for both signal
// precisions, as the reference's double instantiation computes it; the
// float instantiation's own arithmetic is not reproduced (its Lanczos grid
// position optIndex + shift is rounded to FP32, which drops the fraction
// for lags beyond 2^23, and its detected-window "mean" is the float sum,
// relativeMagnitude.cpp:110-125).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "mf_device.hpp"
#include "mf_kernels.hpp"

using namespace mfdev;

namespace mfk {

namespace {

// The refinements are compared bit for bit with a CPU restatement:
// no contraction of products into sums in this file.
#pragma clang fp contract(off)

constexpr int kWave = 64;

// sinc of the Lanczos kernel sinc(x) sinc(x / a) (detector.cpp:91-109):
// sinc(0) = 1 below |x| <= 100 eps; the window keeps every x with
// int(|x|) <= a.
__device__ __forceinline__ double sinc(double x)
{
    double r = 1.0;
    if (fabs(x) > DBL_EPSILON * 100) {
        const double pix = M_PI * x;
        r = sin(pix) / pix;
    }
    return r;
}

// std::min's selection rule, (b < a) ? b : a, which keeps `a` when either is NaN
__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }

// quadraticRefinement (detector.cpp:38-89): the vertex of the parabola
// through lags p-1, p, p+1 (of |rho| when absval), 0 at either end or when
// lag p is neither the maximum nor the minimum of the three.
template <typename T>
__device__ double quadratic_shift(gptr<const T> xc, int n, int p, double dt, int absval)
{
    if (!(p > 0 && p < n - 1)) { return 0.0; }
    double y1 = static_cast<double>(xc[p - 1]);
    double y2 = static_cast<double>(xc[p]);
    double y3 = static_cast<double>(xc[p + 1]);
    if (absval) {
        y1 = fabs(y1);
        y2 = fabs(y2);
        y3 = fabs(y3);
    }
    if (y2 < y1 || y2 < y3) {
        if (y2 < y1 && y2 < y3) {
            y1 = -y1;
            y2 = -y2;
            y3 = -y3;
        } else {
            return 0.0;
        }
    }
    y1 = std_min(y2 - 100 * DBL_EPSILON, y1);
    y3 = std_min(y2 - 100 * DBL_EPSILON, y3);
    const double num = dt * (y1 - y3);
    const double den = 2 * (y1 - 2 * y2 + y3);
    return num / den;
}

// lanczosRefinement (detector.cpp:111-152) over one wavefront: lane l
// evaluates the interpolant at grid points idx = l + 1, l + 65, ...
// (shift = -0.5 + idx / nRefine), summing its 2a + 1 taps in the
// reference's order; the result is the first grid point of the largest
// value, as the sequential strict '>' scan from -DBL_MAX picks it.
template <typename T>
__device__ double lanczos_shift(gptr<const T> xc, int n, int p, double dt, int absval, int a, int nRefine)
{
    const int lane = threadIdx.x;
    const double dtWork = 1.0 / static_cast<double>(nRefine);
    double best = 0.0;
    int bestIdx = 0;  // 0: none (grid indices start at 1)
    for (int idx = lane + 1; idx < nRefine; idx += kWave) {
        const double shift = -0.5 + dtWork * idx;
        const double x = p + shift;
        const int fl = static_cast<int>(floor(x));
        // The taps sit at xi = frac + k - 1 (frac = x - floor(x), exact), so
        // sin(pi xi) = (-1)^(k-1) sin(pi frac): one sine per grid point for
        // the first sinc factor instead of one per tap.  The window's own
        // factor sinc(xi / a) keeps its sine per tap.
        const double sFrac = sin(M_PI * (x - static_cast<double>(fl)));
        double y = 0.0;
        for (int k = -a; k <= a; ++k) {
            const long long i = static_cast<long long>(fl) - k + 1;  // no int overflow near 2^31 lags
            const double xi = x - static_cast<double>(i);
            double v = 0.0;
            if (i >= 0 && i < n) { v = static_cast<double>(xc[i]); }
            double w = 0.0;
            if (static_cast<int>(fabs(xi)) <= a) {
                const double s1 = ((k - 1) & 1) ? -sFrac : sFrac;
                const double sinc1 = fabs(xi) > DBL_EPSILON * 100 ? s1 / (M_PI * xi) : 1.0;
                w = sinc1 * sinc(xi / a);
            }
            y = y + v * w;
        }
        if (absval) { y = fabs(y); }
        // a lane's own points are visited in increasing idx: strict '>' keeps the first
        if (y > -DBL_MAX && (bestIdx == 0 || y > best)) {
            best = y;
            bestIdx = idx;
        }
    }
    // across lanes: larger value wins, equal values go to the smaller index
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, kWave);
        const int oi = __shfl_xor(bestIdx, o, kWave);
        bool take = false;
        if (oi != 0) {
            if (bestIdx == 0) {
                take = true;
            } else if (ob > best) {
                take = true;
            } else if (ob == best && oi < bestIdx) {
                take = true;
            }
        }
        if (take) {
            best = ob;
            bestIdx = oi;
        }
    }
    if (bestIdx == 0) { return 0.0; }
    const double shift = -0.5 + dtWork * bestIdx;
    return dt * shift;
}

// RelativeMagnitude's factors of y[0..Lt) against t[0..Lt) over one
// wavefront, two passes: the means, then the demeaned sums of squares and
// products.  Returns false for a constant window (no factors).
template <typename T>
__device__ bool amplitude_factors(gptr<const double> t, gptr<const T> y, int Lt, double &gr, double &sr)
{
    const int lane = threadIdx.x;
    double st = 0.0, sy = 0.0;
    for (int i = lane; i < Lt; i += kWave) {
        st += t[i];
        sy += static_cast<double>(y[i]);
    }
    const double mt = wave_sum(st) / Lt;
    const double my = wave_sum(sy) / Lt;
    double xx = 0.0, yy = 0.0, xy = 0.0;
    for (int i = lane; i < Lt; i += kWave) {
        const double dx = t[i] - mt;
        const double dy = static_cast<double>(y[i]) - my;
        xx += dx * dx;
        yy += dy * dy;
        xy += dx * dy;
    }
    xx = wave_sum(xx);
    yy = wave_sum(yy);
    xy = wave_sum(xy);
    if (yy == 0.0) { return false; }  // a constant window throws in the reference
    gr = fabs(xy) / xx;
    sr = sqrt(yy / xx);
    return true;
}

template <typename T>
__global__ __launch_bounds__(kWave) void products_kernel(const ProductJob *__restrict__ jobs, int absval,
                                                        int interp, int lanczosAlpha, int lanczosPoints,
                                                        ProductOut *__restrict__ out)
{
    const ProductJob job = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    const gptr<const T> xc = to_global(static_cast<const T *>(job.rho));
    const int n = job.nOut;
    const int p = job.peak;

    double shift = 0.0;
    if (interp == kInterpQuadratic) {
        shift = quadratic_shift<T>(xc, n, p, job.dt, absval);
    } else if (interp == kInterpLanczos) {
        shift = lanczos_shift<T>(xc, n, p, job.dt, absval, lanczosAlpha, lanczosPoints);
    }

    // amplitude scaling factors: two passes over the template and the window
    int flags = 0;
    double gr = 0.0, sr = 0.0;
    const int Lt = job.Lt;
    if (static_cast<long long>(p) + Lt > job.sigLen) {
        flags |= kProductTruncated;
    } else if (amplitude_factors<T>(to_global(job.tmpl), to_global(static_cast<const T *>(job.sig)) + p, Lt, gr,
                                    sr)) {
        flags |= kProductAmplitude;
    }

    if (lane == 0) {
        const double r = static_cast<double>(xc[p]);
        ProductOut o;
        o.rho = fmax(-1.0, fmin(1.0, r));  // NaN -> 1, as std::max(-1, std::min(1, NaN))
        o.shift = shift;
        o.alphaGR = gr;
        o.alphaSR = sr;
        o.flags = flags;
        o.pad = 0;
        out[blockIdx.x] = o;
    }
}

// One wavefront per window: window k is windows[k * ld .. k * ld + Lt).
template <typename T>
__global__ __launch_bounds__(kWave) void amplitude_kernel(const double *__restrict__ tmpl, int Lt,
                                                         const T *__restrict__ windows, long long ld,
                                                         ProductOut *__restrict__ out)
{
    double gr = 0.0, sr = 0.0;
    const bool ok = amplitude_factors<T>(to_global(tmpl), to_global(windows) + blockIdx.x * ld, Lt, gr, sr);
    if (threadIdx.x == 0) {
        ProductOut o{};
        o.alphaGR = gr;
        o.alphaSR = sr;
        o.flags = ok ? kProductAmplitude : 0;
        out[blockIdx.x] = o;
    }
}

}  // namespace

hipError_t launch_amplitude(int precision, const double *dTmpl, int Lt, const void *dWindows, long long ld,
                            int nWindows, ProductOut *dOut, hipStream_t stream)
{
    if (nWindows <= 0) { return hipSuccess; }
    if (precision == 0) {
        hipLaunchKernelGGL(amplitude_kernel<float>, dim3(nWindows), dim3(kWave), 0, stream, dTmpl, Lt,
                           static_cast<const float *>(dWindows), ld, dOut);
    } else {
        hipLaunchKernelGGL(amplitude_kernel<double>, dim3(nWindows), dim3(kWave), 0, stream, dTmpl, Lt,
                           static_cast<const double *>(dWindows), ld, dOut);
    }
    return hipGetLastError();
}

hipError_t launch_products(int precision, const ProductJob *dJobs, int nJobs, int absval, int interp,
                           int lanczosAlpha, int lanczosPoints, ProductOut *dOut, hipStream_t stream)
{
    if (nJobs <= 0) { return hipSuccess; }
    if (precision == 0) {
        hipLaunchKernelGGL(products_kernel<float>, dim3(nJobs), dim3(kWave), 0, stream, dJobs, absval, interp,
                           lanczosAlpha, lanczosPoints, dOut);
    } else {
        hipLaunchKernelGGL(products_kernel<double>, dim3(nJobs), dim3(kWave), 0, stream, dJobs, absval, interp,
                           lanczosAlpha, lanczosPoints, dOut);
    }
    return hipGetLastError();
}

}  // namespace mfk

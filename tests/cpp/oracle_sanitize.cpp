// Host-side memory and undefined-behaviour check of the oracle: built with
// -fsanitize=address,undefined and run by tests/test_sanitizers.py.  Drives
// every restated routine on small inputs, edge cases included (uneven
// template lengths, single-sample and end-point peaks, truncated Lanczos
// windows, constant windows), and cross-checks a few results so a silent
// miscompile would also fail.  Exits non-zero on any failure; the
// sanitizers abort on their own findings.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "mf_oracle.h"

namespace
{
int failures = 0;
void expect(bool ok, const char *what)
{
    if (!ok)
    {
        ++failures;
        std::printf("FAIL %s\n", what);
    }
}
}  // namespace

int main()
{
    std::mt19937_64 gen(7);
    std::normal_distribution<double> g(0.0, 1.0);
    const int n = 3000;
    std::vector<double> x(n);
    for (auto &v : x) { v = g(gen) + 2.0; }
    for (int i = 1200; i < 1500; ++i) { x[i] = 2.5; }  // constant stretch
    const int lens[3] = {120, 97, 120};
    std::vector<std::vector<double>> t(3);
    std::vector<const double *> tp(3);
    for (int k = 0; k < 3; ++k)
    {
        t[k].resize(lens[k]);
        for (auto &v : t[k]) { v = g(gen); }
        tp[k] = t[k].data();
    }
    // sizing
    int nfft = 0, L = 0;
    expect(orc_window_length(4500, 75000, &nfft, &L) == 0 && nfft == 32768 && L == 28269, "window length KAT");
    expect(orc_window_length(100, 100, &nfft, &L) != 0, "nx <= nb rejected");
    // restated apply, single channel, FP64, and the two-pass oracle beside it
    const int nb = 120, nout = n - nb + 1;
    std::vector<double> rho(3 * static_cast<size_t>(nout));
    const double *sig[1] = {x.data()};
    expect(orc_mf_apply_f64(0, 3, lens, tp.data(), n, sig, 2, rho.data(), &nfft, &L) == 0, "apply");
    std::vector<double> ref(n - lens[1] + 1);  // every lag of the shorter template
    expect(orc_pearson_oracle_f64(lens[1], t[1].data(), n, x.data(), 0, nullptr, ref.data()) == 0, "pearson");
    // away from the constant stretch: inside it the reference's sliding sums
    // leave a rounding-sized denominator (its rho is noise over noise there)
    double worst = 0;
    for (int i = 0; i < nout; ++i)
    {
        if (i + lens[1] > 1200 && i < 1500) { continue; }
        worst = std::fmax(worst, std::fabs(rho[nout + i] - ref[i]));
    }
    expect(worst < 1e-11, "apply vs two-pass Pearson");
    // float apply with the padded-length quirk on and off
    std::vector<float> xf(x.begin(), x.end());
    const float *sigf[1] = {xf.data()};
    std::vector<float> rhof(3 * static_cast<size_t>(nout));
    expect(orc_mf_apply_f32(0, 3, lens, tp.data(), n, sigf, 0, 1, rhof.data(), &nfft, &L) == 0, "apply f32");
    expect(orc_mf_apply_f32(0, 3, lens, tp.data(), n, sigf, 1, 1, rhof.data(), &nfft, &L) == 0, "apply f32 quirk");
    // window scales at a few lags, constant windows included
    std::vector<long long> lags = {0, 1, 1250, 1300, nout - 1};
    std::vector<double> sc(lags.size());
    expect(orc_window_scale_f32x(nb, n, xf.data(), static_cast<int>(lags.size()), lags.data(), sc.data()) == 0,
           "window scale");
    expect(sc[3] == 0.0, "constant window scale is 0");
    // stack, reduce, peaks
    const int shifts[3] = {0, 5, 11};
    const double weights[3] = {1.0, 0.5, 0.25};
    std::vector<double> st(nout);
    int stLen = 0;
    expect(orc_shift_and_stack_f64(3, rho.data(), nout, nout, shifts, weights, 1, st.data(), &stLen) == 0 &&
               stLen == nout - 11,
           "stack");
    std::vector<double> det(nout);
    std::vector<int> id(nout);
    expect(orc_reduce_correlations_f64(3, rho.data(), nout, nout, 1, det.data(), id.data()) == 0, "reduce");
    std::vector<int> pk(nout);
    int np = 0;
    expect(orc_peak_finder_f64(nout, det.data(), 0.0, 0, pk.data(), &np) == 0 && np > 0, "peaks");
    expect(orc_peak_finder_f64(nout, det.data(), 0.3, 40, pk.data(), &np) == 0, "peaks min distance");
    double one = 0.9;
    expect(orc_peak_finder_f64(1, &one, 0.5, 0, pk.data(), &np) == 0 && np == 1, "single sample");
    // per-detection products at the row ends (truncated Lanczos windows)
    double shift = 0;
    std::vector<double> prof(99);
    for (int p : {0, 1, nout / 2, nout - 2, nout - 1})
    {
        expect(orc_quadratic_shift(nout, det.data(), p, 0.01, 1, &shift) == 0, "quadratic");
        expect(orc_lanczos_shift(nout, det.data(), p, 0.01, 1, 51, 100, &shift, prof.data()) == 0, "lanczos");
        expect(std::fabs(shift) < 0.01, "lanczos shift within half a sample");
    }
    double gr = 0, sr = 0;
    int ok = 0;
    expect(orc_amplitude_factors(120, t[0].data(), x.data() + 10, &gr, &sr, &ok) == 0 && ok == 1, "factors");
    expect(orc_amplitude_factors(120, t[0].data(), x.data() + 1250, &gr, &sr, &ok) == 0 && ok == 0,
           "constant window: no factors");
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}

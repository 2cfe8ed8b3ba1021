// Wakefield energy loss per slice (host, FP64): the single-particle wakes of
// Wake::singleWakeResistive / singleWakeGeometric / singleWakeRoughness
// (src/Main/Wake.cpp:136-434) and the current convolution of
// Collective::initWake/update (src/Core/Collective.cpp:36-105, 132-196).
// The quadratures keep the reference's grids and step order so the losses
// agree with its values; the device applies the result per step.
#include "g4mi_wake.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>

#include "g4mi.h"

namespace g4mi {

namespace {
constexpr double kVacimpW = 376.73;        // src/Main/GenMain.cpp:63
constexpr double kCeW = 4.8032045e-11;     // src/Main/GenMain.cpp:68
typedef std::complex<double> cd;

// Roughness kernel on the segment [q1, q2] with N points, trapezoid end
// weights folded in (Wake::KernelRoughness, :136-150)
std::vector<cd> rough_kernel(int n, cd q1, cd q2, double rr) {
    std::vector<cd> k(n);
    const cd dq = (q2 - q1) / (double)(n - 1);
    const cd i(0, 1);
    for (int j = 0; j < n; j++) {
        const cd q = q1 + (double)j * dq;
        const cd sq = (std::sqrt(2. * q + 1.) - i * std::sqrt(2. * q - 1.)) * q / std::sqrt(4. * q * q - 1.);
        k[j] = (sq + 1.) / (1. - i * rr * q * sq) / (1. + i * rr * q);
    }
    k[0] *= 0.5;
    k[n - 1] *= 0.5;
    return k;
}

// Re of dq sum_j exp(-i q_j tau) K_j (TrapIntegrateRoughness, :154-167)
double rough_integral(const std::vector<cd> &k, cd q1, cd q2, double tau) {
    const int n = (int)k.size();
    const cd dq = (q2 - q1) / (double)(n - 1);
    const cd i(0, 1);
    cd v(0, 0);
    for (int j = 0; j < n; j++) {
        const cd q = q1 + (double)j * dq;
        v += std::exp(-i * q * tau) * k[j];
    }
    v *= dq;
    return v.real();
}
}  // namespace

void wake_roughness(int ns, double ds, double radius, double hrough, double lrough, double *w) {
    for (int i = 0; i < ns; i++) w[i] = 0;
    if (hrough <= 0) return;
    const double pi = 2. * asin(1);
    const double rr = pi * pi * pi / lrough / lrough / lrough * hrough * hrough * radius;
    const cd q1(0, 0), q2(0, 2e-3), q3(1, 2e-3), q4(1, 0), q5(100, 0);
    const int n = 128;
    const std::vector<cd> k1 = rough_kernel(n, q1, q2, rr), k2 = rough_kernel(8 * n, q2, q3, rr),
                          k3 = rough_kernel(n, q3, q4, rr), k4 = rough_kernel(8 * n, q4, q5, rr);
    const double coef = rr / pi * 4 / radius / radius * 1.6e-19 / 4 / pi / 8.854e-12;
    for (int i = 0; i < ns; i++) {
        const double tau = 2 * pi * ds * i / lrough;
        double r = rough_integral(k1, q1, q2, tau);
        r += rough_integral(k2, q2, q3, tau);
        r += rough_integral(k3, q3, q4, tau);
        r += rough_integral(k4, q4, q5, tau);
        w[i] = coef * r;
    }
}

void wake_geometric(int ns, double ds, double radius, double gap, double lgap, bool roundpipe, double *w) {
    for (int i = 0; i < ns; i++) w[i] = 0;
    if (gap <= 0) return;
    const double pi = 2. * asin(1.);
    double coef = -kVacimpW * kCeW / (pi * pi * radius * lgap) * 2 * sqrt(0.5 * gap);
    if (!roundpipe) coef *= 0.956;
    for (int i = 0; i < ns; i++) w[i] = coef * sqrt(ds * i);
}

void wake_resistive(int ns, double ds, double radius, double conductivity, double relaxation, bool roundpipe,
                    double *w) {
    for (int i = 0; i < ns; i++) w[i] = 0.0;
    if (conductivity <= 0.0) return;
    const double c_light = 299792458.0, z0 = kVacimpW, a = radius, pi = 2.0 * std::asin(1.0);
    const double s0 = std::pow(2.0 * a * a / (z0 * conductivity), 1.0 / 3.0);
    const unsigned nk = 1000;                 // k in [0, 100 / s0], nk intervals
    const double kmax = 100.0 / s0, dk = kmax / (double)nk;
    const double xmax = 15.0;                 // flat chamber: x-integral on [0, 15], nq points
    const int nq = 20000;
    const double pref = z0 / (2.0 * pi * a);
    // AC conductivity and Leontovich surface impedance (Bane & Stupakov,
    // SLAC-PUB-10707, Eq. 1)
    auto zeta = [&](double k) -> cd {
        const cd sig = conductivity / (cd(1.0, 0.0) - cd(0.0, 1.0) * k * relaxation);
        return cd(1.0, -1.0) * std::sqrt(cd(k, 0.0) / (cd(2.0, 0.0) * sig * cd(z0, 0.0)));
    };
    auto sinhc = [](double x) {
        if (std::abs(x) < 1e-5) {
            const double x2 = x * x;
            return 1.0 + x2 * (1.0 / 6.0 + x2 * (1.0 / 120.0));
        }
        return std::sinh(x) / x;
    };
    auto z_round = [&](double k) -> cd {
        if (k == 0.0) return cd(0.0, 0.0);
        const cd den = cd(1.0, 0.0) / zeta(k) - cd(0.0, 0.5 * k * a);
        if (!std::isfinite(den.real()) || !std::isfinite(den.imag())) return cd(0.0, 0.0);
        return cd(pref, 0.0) / den;
    };
    auto z_flat_at = [&](double k, double x) -> cd {
        const double ch = std::cosh(x);
        const cd den = cd(ch, 0.0) * (cd(ch, 0.0) / zeta(k) + cd(0.0, -k * a * sinhc(x)));
        if (!std::isfinite(den.real()) || !std::isfinite(den.imag())) return cd(0.0, 0.0);
        return cd(pref, 0.0) / den;
    };
    auto z_flat = [&](double k) -> cd {
        if (k == 0.0) return cd(0.0, 0.0);
        const double dx = xmax / (double)(nq - 1);
        cd acc = z_flat_at(k, 0.0) * 0.5;
        for (int j = 1; j < nq - 1; ++j) acc += z_flat_at(k, dx * (double)j);
        acc += z_flat_at(k, xmax) * 0.5;
        return acc * dx;
    };
    std::vector<double> ks(nk + 1), rez(nk + 1);
    for (unsigned j = 0; j <= nk; ++j) {
        ks[j] = dk * (double)j;
        rez[j] = (roundpipe ? z_round(ks[j]) : z_flat(ks[j])).real();
    }
    rez.front() *= 0.5;
    rez.back() *= 0.5;
    const double cpref = (2.0 * c_light / pi) * dk;  // one-sided cosine transform
    for (int i = 0; i < ns; ++i) {
        const double s = i * ds;
        double acc = 0.0;
        for (unsigned j = 0; j <= nk; ++j) acc += rez[j] * std::cos(ks[j] * s);
        w[i] = acc * cpref;
    }
    const double e_charge = 1.602176634e-19;
    for (int i = 0; i < ns; ++i) w[i] *= -e_charge;
}

void wake_loss(int ncur, int sample_hi, double ds, const double *cur, const double *ext, const double *wres,
               const double *wgeo, const double *wrou, double *eloss) {
    const int ns = ncur * sample_hi;  // slices at the highest resolution
    // Collective::initWake: self-loading theorem on the first bin
    std::vector<double> res(wres, wres + ns), geo(wgeo, wgeo + ns), rou(wrou, wrou + ns);
    geo[0] *= 0.5;
    res[0] *= 0.5;
    rou[0] *= 0.5;
    const double dscur = ds * ns / ncur;
    // Collective::update: current profile at high resolution (electrons per
    // bin) and its derivative, then the loss of each simulation slice
    std::vector<double> c(ncur + 1, 0.0), current(ns), dcurrent(ns);
    for (int k = 0; k < ncur; k++) c[k] = cur[k];
    c[ncur] = 0;
    for (int is = 0; is < ns; is++) {
        const double s = ds * (double)is;
        const unsigned idx = (unsigned)(int)floor(s / dscur);
        const double wei = 1 - (s - idx * dscur) / dscur;
        current[is] = wei * c[idx] + (1 - wei) * c[idx + 1];
        current[is] *= ds / kCeW;
        dcurrent[is] = -(c[idx + 1] - c[idx]) * ds / kCeW / dscur;
    }
    int sample = (int)(dscur / ds);
    if (sample < 1) sample = 1;
    for (int ic = 0; ic < ncur; ic++) {
        double wi = 0;
        const int is0 = ic * sample;
        for (int j = 0; j < sample; j++) {
            const int is = is0 + j;
            for (int i = 0; i < ns - is; i++) {
                wi += current[is + i] * (res[i] + rou[i]);
                wi += dcurrent[is + i] * geo[i];
            }
        }
        wi /= (double)sample;
        eloss[ic] = ext[ic] + wi;  // Collective::apply: eloss = wakeext + wakeint
    }
}

}  // namespace g4mi

namespace g4mi {
void set_error(const std::string &msg);
}

extern "C" int g4mi_wake_loss(int nslice, int sample, double ds, const double *current, const double *loss_ext,
                              double radius, int roundpipe, double conductivity, double relaxation, double gap,
                              double lgap, double hrough, double lrough, double *eloss, double *wres,
                              double *wgeo, double *wrou) {
    using namespace g4mi;
    if (nslice < 1 || sample < 1 || !(ds > 0) || !current || !eloss || !(radius > 0) || !(lgap > 0) ||
        !(lrough > 0)) {
        set_error("g4mi_wake_loss: invalid argument");
        return G4MI_E_ARG;
    }
    const long long ns = (long long)nslice * sample;
    if (ns > (1LL << 24)) {
        set_error("g4mi_wake_loss: too many slices");
        return G4MI_E_ARG;
    }
    std::vector<double> r(ns), g(ns), o(ns), ext(nslice, 0.0);
    if (loss_ext) ext.assign(loss_ext, loss_ext + nslice);
    wake_resistive((int)ns, ds, radius, conductivity, relaxation, roundpipe != 0, r.data());
    wake_geometric((int)ns, ds, radius, gap, lgap, roundpipe != 0, g.data());
    wake_roughness((int)ns, ds, radius, hrough, lrough, o.data());
    wake_loss(nslice, sample, ds, current, ext.data(), r.data(), g.data(), o.data(), eloss);
    if (wres) std::copy(r.begin(), r.end(), wres);
    if (wgeo) std::copy(g.begin(), g.end(), wgeo);
    if (wrou) std::copy(o.begin(), o.end(), wrou);
    return G4MI_OK;
}

// Host-side numerics and input generation of the g4mi library.
//
// Compiled with g++ -O2 -ffp-contract=off: the per-step scalars and the
// generated particles/fields are bit-identical to the reference's CPU build
// on the same inputs (same libm, same std::complex, same operation order).
#include "g4mi_host.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>

namespace g4mi {

namespace {
constexpr double kVacimpH = 376.73;
constexpr double kEevH = 510998.95069;

// Rational approximations of J0 and J1 on |x|<8 and the asymptotic form
// beyond (Numerical Recipes bessj0/bessj1, the method of
// src/Util/BesselJ.cpp:16-95).  The coefficient tables are the method's.
struct Poly6 {
    double c[6];
    double eval(double y) const {
        return c[0] + y * (c[1] + y * (c[2] + y * (c[3] + y * (c[4] + y * c[5]))));
    }
};
struct Poly5 {
    double c[5];
    double eval(double y) const { return c[0] + y * (c[1] + y * (c[2] + y * (c[3] + y * c[4]))); }
};
struct BesselTable {
    Poly6 r, s;
    Poly5 p, q;
    double phase;
    bool odd;
};
const BesselTable kJ0 = {{{57568490574.e0, -13362590354.e0, 651619640.7e0, -11214424.18e0,
                           77392.33017e0, -184.9052456e0}},
                         {{57568490411.e0, 1029532985.e0, 9494680.718e0, 59272.64853e0,
                           267.8532712e0, 1.e0}},
                         {{1.e0, -.1098628627e-2, .2734510407e-4, -.2073370639e-5, .2093887211e-6}},
                         {{-.1562499995e-1, .1430488765e-3, -.6911147651e-5, .7621095161e-6,
                           -.934945152e-7}},
                         0.785398164,
                         false};
const BesselTable kJ1 = {{{72362614232.e0, -7895059235.e0, 242396853.1e0, -2972611.439e0,
                           15704.48260e0, -30.16036606e0}},
                         {{144725228442.e0, 2300535178.e0, 18583304.74e0, 99447.43394e0,
                           376.9991397e0, 1.e0}},
                         {{1.e0, .183105e-2, -.3516396496e-4, .2457520174e-5, -.240337019e-6}},
                         {{.04687499995e0, -.2002690873e-3, .8449199096e-5, -.88228987e-6,
                           .105787412e-6}},
                         2.356194491,
                         true};
double bessel_low(const BesselTable &t, double x) {
    if (fabs(x) < 8) {
        double y = x * x;
        double num = t.r.eval(y);
        return t.odd ? x * num / t.s.eval(y) : num / t.s.eval(y);
    }
    double ax = fabs(x);
    double z = 8 / ax;
    double y = z * z;
    double xx = ax - t.phase;
    return sqrt(0.636619772 / ax) * (cos(xx) * t.p.eval(y) - z * sin(xx) * t.q.eval(y));
}
}  // namespace

double bessel_j(int n, double x) {
    if (n == 0) return bessel_low(kJ0, x);
    if (n == 1) return bessel_low(kJ1, x);
    if (x == 0) return 0;
    const double ax = fabs(x);
    double res;
    if (ax > (double)n) {  // upward recurrence
        const double tox = 2 / ax;
        double jm = bessel_low(kJ0, ax), j = bessel_low(kJ1, ax);
        for (int k = 1; k < n; k++) {
            const double jp = k * tox * j - jm;
            jm = j;
            j = jp;
        }
        res = j;
    } else {  // Miller's downward recurrence with renormalisation
        const double tox = 2. / ax;
        const int m = 2 * (n + floor(sqrt((double)(40 * n))) / 2);
        res = 0;
        bool even = false;
        double sum = 0, jp = 0, j = 1;
        for (int k = m; k > 0; k--) {
            const double jm = k * tox * j - jp;
            jp = j;
            j = jm;
            if (fabs(j) > 1e10) {
                j *= 1e-10;
                jp *= 1e-10;
                res *= 1e-10;
                sum *= 1e-10;
            }
            if (even) sum += j;
            even = !even;
            if (k == n) res = jp;
        }
        sum = 2 * sum - j;
        res = res / sum;
    }
    if ((x < 0) && ((n % 2) == 1)) res = -res;
    return res;
}

double coupling(const g4mi_step_params &st, int h) {
    if (st.helical != 0) return (h == 1) ? st.aw : 0;
    if (h % 2 != 1) return 0;
    double xi = st.aw * st.aw;
    xi = 0.5 * xi / (1 + xi) * (double)h;
    const int h0 = (h - 1) / 2;
    return st.aw * (bessel_j(h0, xi) - bessel_j(h0 + 1, xi)) * pow(-1., h0);
}

double adi_factors(int ngrid, double delz, double dgrid, double xks, double *c, double *cbet,
                   double *cwet) {
    typedef std::complex<double> cd;
    const double rtmp = 0.25 * delz / (xks * dgrid * dgrid);
    // Tridiagonal (1 - i rtmp delta^2): sub/super diagonal -i*rtmp, diagonal
    // 1 + 2 i rtmp, with the first row's lower and last row's upper entry 0.
    std::vector<cd> up(ngrid), mid(ngrid), lo(ngrid), bet(ngrid), wet(ngrid);
    for (int i = 0; i < ngrid; i++) {
        up[i] = cd(0, -(i == ngrid - 1 ? 0 : rtmp));
        mid[i] = cd(1, -(-2 * rtmp));
        lo[i] = cd(0, -(i == 0 ? 0 : rtmp));
    }
    bet[0] = 1. / mid[0];
    wet[0] = 0.;
    for (int i = 1; i < ngrid; i++) {
        wet[i] = up[i - 1] * bet[i - 1];
        bet[i] = 1. / (mid[i] - lo[i] * wet[i]);
    }
    for (int i = 0; i < ngrid; i++) {
        c[2 * i] = lo[i].real();
        c[2 * i + 1] = lo[i].imag();
        cbet[2 * i] = bet[i].real();
        cbet[2 * i + 1] = bet[i].imag();
        cwet[2 * i] = wet[i].real();
        cwet[2 * i + 1] = wet[i].imag();
    }
    return rtmp;
}

double adi_band(int ngrid, const double *c, const double *cbet, const double *cwet, int M,
                double *band) {
    typedef std::complex<double> cd;
    const int n = ngrid, W = 2 * M + 1;
    std::vector<cd> C(n), B(n), Wt(n), col(n);
    for (int i = 0; i < n; i++) {
        C[i] = cd(c[2 * i], c[2 * i + 1]);
        B[i] = cd(cbet[2 * i], cbet[2 * i + 1]);
        Wt[i] = cd(cwet[2 * i], cwet[2 * i + 1]);
    }
    std::vector<cd> inv((size_t)n * n);
    for (int j = 0; j < n; j++) {  // column j of A^-1: Thomas on e_j
        for (int i = 0; i < n; i++) {
            const cd r = (i == j) ? cd(1, 0) : cd(0, 0);
            col[i] = (i == 0) ? r * B[0] : (r - C[i] * col[i - 1]) * B[i];
        }
        for (int i = n - 2; i >= 0; i--) col[i] -= Wt[i + 1] * col[i + 1];
        for (int i = 0; i < n; i++) inv[(size_t)i * n + j] = col[i];
    }
    double worst = 0;
    for (int k = 0; k < n; k++) {
        const double diag = std::abs(inv[(size_t)k * n + k]);
        for (int j = 0; j < n; j++) {
            const int m = j - k + M;
            if (m >= 0 && m < W) continue;
            worst = std::max(worst, std::abs(inv[(size_t)k * n + j]) / diag);
        }
        for (int m = 0; m < W; m++) {
            const int j = k + m - M;
            const cd v = (j >= 0 && j < n) ? inv[(size_t)k * n + j] : cd(0, 0);
            band[2 * ((size_t)k * W + m)] = v.real();
            band[2 * ((size_t)k * W + m) + 1] = v.imag();
        }
    }
    return worst;
}

namespace {
// FFT index k of grid point i: (i + (n+1)/2) mod n (FieldSolverFFT.cpp:137-138),
// inverted; returns the point's offset dx = i - (n-1)/2 from the centre.
double fft_offset(int n, int k) {
    const int i = ((k - (n + 1) / 2) % n + n) % n;
    return (double)i - 0.5 * (double)(n - 1);
}
}  // namespace

namespace {
// m[d] = (1/n) sum_k p_k exp(2 pi i k d / n), p_k = exp(i K2_1(k) delz), in
// long double with the DFT exponents reduced mod n exactly
std::vector<std::complex<long double>> propagator_kernel(int n, double dgrid, double xks, double delz) {
    typedef std::complex<long double> cl;
    const double dk = 4. * asin(1.) / ((double)n * dgrid);
    std::vector<cl> p(n), md(n);
    for (int k = 0; k < n; k++) {
        const double dx = fft_offset(n, k);
        const double ph = -(dx * dx) * dk * dk / 2. / xks * delz;
        p[k] = cl(cos(ph), sin(ph));
    }
    const long double twopi = 8.0L * atanl(1.0L);
    for (int d = 0; d < n; d++) {
        cl acc(0, 0);
        for (int k = 0; k < n; k++) {
            const long double a = twopi * (long double)(((long long)k * d) % n) / (long double)n;
            acc += p[k] * cl(cosl(a), sinl(a));
        }
        md[d] = acc / (long double)n;
    }
    return md;
}
}  // namespace

void fft_propagator(int n, double dgrid, double xks, double delz, double *m) {
    typedef std::complex<long double> cl;
    const std::vector<cl> md = propagator_kernel(n, dgrid, xks, delz);
    for (int j = 0; j < n; j++)
        for (int l = 0; l < n; l++) {
            const cl v = md[((j - l) % n + n) % n];
            m[2 * ((size_t)j * n + l)] = (double)v.real();
            m[2 * ((size_t)j * n + l) + 1] = (double)v.imag();
        }
}

int fft_conv_length(int n) {
    int L = 128;
    while (L < 2 * n - 1) L *= 2;
    return (n >= 2 && L <= 512) ? L : 0;
}

int fft_conv_tables(int n, double dgrid, double xks, double delz, double *hk, double *tw) {
    typedef std::complex<long double> cl;
    const int L = fft_conv_length(n);
    if (!L) return 0;
    const std::vector<cl> md = propagator_kernel(n, dgrid, xks, delz);
    const long double twopi = 8.0L * atanl(1.0L);
    std::vector<cl> w(L);
    for (int k = 0; k < L; k++) {
        const long double a = twopi * (long double)k / (long double)L;
        w[k] = cl(cosl(a), -sinl(a));
        tw[2 * k] = (double)w[k].real();
        tw[2 * k + 1] = (double)w[k].imag();
    }
    for (int k = 0; k < L; k++) {  // (1/L) sum over |d| < n of m[d mod n] exp(-2 pi i k d / L)
        cl acc(0, 0);
        for (int d = -(n - 1); d < n; d++) {
            const int dl = (d % L + L) % L;
            acc += md[(d % n + n) % n] * w[(int)(((long long)k * dl) % L)];
        }
        acc /= (long double)L;
        hk[2 * k] = (double)acc.real();
        hk[2 * k + 1] = (double)acc.imag();
    }
    return L;
}

int fft_chirp_tables(int n, int sign, double *hk, double *chirp) {
    // Bluestein: sum_j x_j exp(sign 2 pi i j k / n) = w_k sum_j (x_j w_j) conj(w_(k-j)),
    // w_m = exp(sign i pi m^2 / n), as jk = (j^2 + k^2 - (k - j)^2) / 2; the
    // exponent m^2 is reduced mod 2n exactly
    typedef std::complex<long double> cl;
    const int L = fft_conv_length(n);
    if (!L) return 0;
    const long double pi = 4.0L * atanl(1.0L);
    auto w = [&](long long m) {
        const long long r = (m * m) % (2LL * n);
        const long double a = (long double)sign * pi * (long double)r / (long double)n;
        return cl(cosl(a), sinl(a));
    };
    for (int j = 0; j < n; j++) {
        const cl c = w(j);
        chirp[2 * j] = (double)c.real();
        chirp[2 * j + 1] = (double)c.imag();
    }
    const long double twopi = 8.0L * atanl(1.0L);
    std::vector<cl> tw(L);
    for (int k = 0; k < L; k++) {
        const long double a = twopi * (long double)k / (long double)L;
        tw[k] = cl(cosl(a), -sinl(a));
    }
    for (int k = 0; k < L; k++) {  // (1/L) DFT_L of conj(w_d) placed at d mod L, |d| < n
        cl acc(0, 0);
        for (int d = -(n - 1); d < n; d++) {
            const int dl = (d % L + L) % L;
            acc += std::conj(w(d < 0 ? -d : d)) * tw[(int)(((long long)k * dl) % L)];
        }
        acc /= (long double)L;
        hk[2 * k] = (double)acc.real();
        hk[2 * k + 1] = (double)acc.imag();
    }
    return L;
}

void dft_matrix(int n, int sign, double *w) {
    const long double twopi = 8.0L * atanl(1.0L);
    for (int j = 0; j < n; j++)
        for (int l = 0; l < n; l++) {
            const long double a = twopi * (long double)(((long long)j * l) % n) / (long double)n;
            w[2 * ((size_t)j * n + l)] = (double)cosl(a);
            w[2 * ((size_t)j * n + l) + 1] = (double)(sign * sinl(a));
        }
}

void fft_sigmoid(int n, double xc, double yc, double sig, double *s) {
    const double shift = -0.5 * (double)(n - 1);
    for (int iy = 0; iy < n; iy++) {
        const double y = ((double)iy + shift) / (double)n / yc;
        for (int ix = 0; ix < n; ix++) {
            const double x = ((double)ix + shift) / (double)n / xc;
            const int ii = ((iy + (n + 1) / 2) % n) * n + (ix + (n + 1) / 2) % n;
            const double r = (sqrt(x * x + y * y) - 1) / sig;
            s[ii] = 1. / (1 + exp(r));
        }
    }
}

namespace {
// t = e * m (left-multiplication), as TrackBeam::matmul.
void left_mul(double m[16], const double e[16]) {
    double t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = 0;
            for (int k = 0; k < 4; k++) acc += e[4 * i + k] * m[4 * k + j];
            t[4 * i + j] = acc;
        }
    std::memcpy(m, t, sizeof(t));
}
void unit4(double m[16]) {
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1 : 0;
}
}  // namespace

void chicane_matrix(double angle, double lb, double ld, double lt, double m[16]) {
    double d1[16], d2[16], d3[16], bp[16], bn[16], ep[16], en[16];
    unit4(m);
    for (double *a : {d1, d2, d3, bp, bn, ep, en}) unit4(a);
    d1[1] = ld / cos(angle);
    d1[11] = ld / cos(angle);
    d2[1] = lt - 4 * lb - 2 * ld;
    d2[11] = lt - 4 * lb - 2 * ld;
    d3[1] = -lt;
    d3[11] = -lt;
    const double R = lb / sin(angle);
    const double Lpath = R * angle;
    bp[11] = Lpath;
    bp[0] = cos(angle);
    bp[1] = R * sin(angle);
    bp[4] = -sin(angle) / R;
    bp[5] = cos(angle);
    bn[11] = Lpath;
    bn[0] = cos(-angle);
    bn[1] = R * sin(-angle) * -1;
    bn[4] = -sin(-angle) / R * -1;
    bn[5] = cos(-angle);
    const double efoc = tan(angle) / R;
    ep[4] = efoc;
    ep[14] = -efoc;
    en[4] = -efoc * -1;
    en[14] = efoc * -1;
    for (const double *e : {bp, ep, d1, en, bn, d2, bn, en, d1, ep, bp, d3}) left_mul(m, e);
}

double chicane_r56(double angle, double lb, double ld, double lambda0, double gamma0) {
    double r56 = (4 * lb / sin(angle) * (1 - angle / tan(angle)) + 2 * ld * tan(angle) / cos(angle)) * angle;
    return r56 * 4 * asin(1) / lambda0 / gamma0;
}

// ---------------------------------------------------------------------------
// Random and quasi-random sequences
// ---------------------------------------------------------------------------
namespace {
constexpr int kIm1 = 2147483563, kIm2 = 2147483399, kIa1 = 40014, kIa2 = 40692;
constexpr int kIq1 = 53668, kIq2 = 52774, kIr1 = 12211, kIr2 = 3791;
inline int lcg(int s, int a, int q, int r, int m) {  // Schrage's method
    const int k = s / q;
    s = a * (s - k * q) - k * r;
    return s < 0 ? s + m : s;
}
}  // namespace

Ran2::Ran2(unsigned int seed) {
    s1_ = (1u < seed) ? (int)seed : 1;
    s2_ = s1_;
    for (int i = 39; i >= 0; i--) {
        s1_ = lcg(s1_, kIa1, kIq1, kIr1, kIm1);
        if (i < 32) iv_[i] = s1_;
    }
    iy_ = iv_[0];
}

void Ran2::save(int *state) const {
    for (int i = 0; i < 32; i++) state[i] = iv_[i];
    state[32] = iy_;
    state[33] = s1_;
    state[34] = s2_;
}

void Ran2::load(const int *state) {
    for (int i = 0; i < 32; i++) iv_[i] = state[i];
    iy_ = state[32];
    s1_ = state[33];
    s2_ = state[34];
}

double Ran2::next() {
    s1_ = lcg(s1_, kIa1, kIq1, kIr1, kIm1);
    s2_ = lcg(s2_, kIa2, kIq2, kIr2, kIm2);
    const int j = iy_ / 67108862;
    iy_ = iv_[j] - s2_;
    iv_[j] = s1_;
    if (iy_ < 1) iy_ += kIm1 - 1;
    const double am = 1. / 2147483563;
    const double v = am * iy_;
    return v < 1. - 1.2e-40 ? v : 1. - 1.2e-40;
}

namespace {
// Radical inverse of successive integers in a prime base (Halton
// sequence), counter pre-incremented as src/Util/Hammerslay.cpp:41-49.
class Halton {
  public:
    Halton(double base, unsigned int start) : b_(base), n_(start) {}
    double next() {
        unsigned int q = ++n_;
        double v = 0, scale = 1.0;
        do {
            scale = scale / b_;
            const unsigned int d = q / (unsigned int)b_;
            v += (q - b_ * d) * scale;
            q = d;
        } while (q > 0);
        return v;
    }

  private:
    double b_;
    unsigned int n_;
};

// Normalise a[] to zero mean and unit rms and remove its correlation from
// b[] (QuietLoading.cpp:110-156).
void decorrelate(double *a, double *b, int n, double norm) {
    double sa = 0, saa = 0, sb = 0, sab = 0;
    for (int i = 0; i < n; i++) {
        sa += a[i];
        saa += a[i] * a[i];
        sb += b[i];
        sab += a[i] * b[i];
    }
    sa *= norm;
    sb *= norm;
    saa = sqrt(fabs(saa * norm - sa * sa));
    if (saa > 0) saa = 1. / saa;
    sab = (sab * norm - sa * sb) * saa * saa;
    for (int i = 0; i < n; i++) {
        b[i] = b[i] - sb;
        b[i] -= sab * a[i];
        a[i] = (a[i] - sa) * saa;
    }
}
}  // namespace

double inverse_erfc(double y) {
    // T. Ooura's inverse complementary error function: a rational first
    // guess refined by one high-order correction (the method of
    // src/Util/Inverfc.cpp:16-64); the coefficients are Ooura's.
    const double z = (y > 1) ? 2 - y : y;
    const double w = 0.916461398268964 - log(z);
    double u = sqrt(w);
    double s = (log(u) + 0.488826640273108) / w;
    double t = 1 / (u + 0.231729200323405);
    double x = u * (1 - s * (s * 0.124610454613712 + 0.5)) -
               ((((-0.0728846765585675 * t + 0.269999308670029) * t + 0.150689047360223) * t +
                 0.116065025341614) * t + 0.499999303439796) * t;
    t = 3.97886080735226 / (x + 3.97886080735226);
    u = t - 0.5;
    static const double a[11] = {0.00112648096188977922, 1.05739299623423047e-4,
                                 -0.00351287146129100025, -7.71708358954120939e-4,
                                 0.00685649426074558612, 0.00339721910367775861,
                                 -0.011274916933250487, -0.0118598117047771104,
                                 0.0142961988697898018, 0.0346494207789099922,
                                 0.00220995927012179067};
    static const double b[12] = {-0.0743424357241784861, -0.105872177941595488,
                                 0.0147297938331485121, 0.316847638520135944,
                                 0.713657635868730364, 1.05375024970847138,
                                 1.21448730779995237, 1.16374581931560831,
                                 0.956464974744799006, 0.686265948274097816,
                                 0.434397492331430115, 0.244044510593190935};
    s = a[0];
    for (int k = 1; k < 11; k++) s = s * u + a[k];
    for (int k = 0; k < 12; k++) s = s * u + b[k];
    s = s * t - z * exp(x * x - 0.120782237635245222);
    x = x + s * (x * s + 1);
    return (y > 1) ? -x : x;
}

namespace {
// QuietLoading::loadQuiet after the raw draws (:95-260): exact moments,
// x-px decorrelation, Twiss scaling, beamlet mirroring, phase scale,
// bunching and energy modulation.
void quiet_finish(double *g, double *th, double *x, double *y, double *px, double *py, const SliceBeam &sl,
                  int npart, int nbins, double theta0);
}  // namespace

unsigned int quiet_start(int islice) {
    Ran2 r((unsigned int)islice);
    return (unsigned int)(int)round(r.next() * 1e9);
}

void quiet_load(double *g, double *th, double *x, double *y, double *px, double *py,
                const SliceBeam &sl, int npart, int nbins, double theta0, int islice) {
    const unsigned int start = quiet_start(islice);
    Halton h[6] = {Halton(2, start), Halton(3, start), Halton(5, start),
                   Halton(7, start), Halton(11, start), Halton(13, start)};
    const int mpart = npart / nbins;
    const double dtheta = 1. / (double)nbins;
    for (int i = 0; i < mpart; i++) {
        th[i] = h[0].next() * dtheta;
        g[i] = inverse_erfc(2 * h[1].next());
        x[i] = inverse_erfc(2 * h[2].next());
        y[i] = inverse_erfc(2 * h[3].next());
        px[i] = inverse_erfc(2 * h[4].next());
        py[i] = inverse_erfc(2 * h[5].next());
    }
    quiet_finish(g, th, x, y, px, py, sl, npart, nbins, theta0);
}

unsigned int one4one_seed(int seed, int rank) {
    Ran2 r((unsigned int)seed);
    double v = 0;
    for (int i = 0; i <= rank; i++) v = r.next();
    return (unsigned int)(int)round(v * 1e9);
}

void quiet_load_one4one(Ran2 &stream, double *g, double *th, double *x, double *y, double *px, double *py,
                        const SliceBeam &sl, int npart, double theta0) {
    // all six coordinates from the one stream, particle by particle
    for (int i = 0; i < npart; i++) {
        th[i] = stream.next();
        g[i] = inverse_erfc(2 * stream.next());
        x[i] = inverse_erfc(2 * stream.next());
        y[i] = inverse_erfc(2 * stream.next());
        px[i] = inverse_erfc(2 * stream.next());
        py[i] = inverse_erfc(2 * stream.next());
    }
    quiet_finish(g, th, x, y, px, py, sl, npart, 1, theta0);
}

namespace {
void quiet_finish(double *g, double *th, double *x, double *y, double *px, double *py, const SliceBeam &sl,
                  int npart, int nbins, double theta0) {
    const int mpart = npart / nbins;
    const double dtheta = 1. / (double)nbins;
    const double norm = mpart > 0 ? 1. / (double)mpart : 0;
    {  // energy to exact mean and rms
        double m1 = 0, m2 = 0;
        for (int i = 0; i < mpart; i++) {
            m1 += g[i];
            m2 += g[i] * g[i];
        }
        m1 *= norm;
        m2 = sqrt(fabs(m2 * norm - m1 * m1));
        if (m2 > 0) m2 = 1 / m2;
        m2 *= sl.delgam;
        for (int i = 0; i < mpart; i++) g[i] = (g[i] - m1) * m2 + sl.gamma;
    }
    decorrelate(x, px, mpart, norm);
    decorrelate(y, py, mpart, norm);
    {  // momenta to unit rms
        double a1 = 0, a2 = 0, b1 = 0, b2 = 0;
        for (int i = 0; i < mpart; i++) {
            a1 += px[i];
            a2 += px[i] * px[i];
            b1 += py[i];
            b2 += py[i] * py[i];
        }
        a1 *= norm;
        b1 *= norm;
        a2 = sqrt(fabs(a2 * norm - a1 * a1));
        if (a2 > 0) a2 = 1. / a2;
        b2 = sqrt(fabs(b2 * norm - b1 * b1));
        if (b2 > 0) b2 = 1. / b2;
        for (int i = 0; i < mpart; i++) {
            px[i] = (px[i] - a1) * a2;
            py[i] = (py[i] - b1) * b2;
        }
    }
    const double sigx = sqrt(sl.ex * sl.betax / sl.gamma);
    const double sigy = sqrt(sl.ey * sl.betay / sl.gamma);
    const double sigpx = sqrt(sl.ex / sl.betax / sl.gamma);
    const double sigpy = sqrt(sl.ey / sl.betay / sl.gamma);
    const double corx = -sl.alphax / sl.betax;
    const double cory = -sl.alphay / sl.betay;
    for (int i = 0; i < mpart; i++) {
        x[i] *= sigx;
        y[i] *= sigy;
        px[i] = sigpx * px[i] + corx * x[i];
        py[i] = sigpy * py[i] + cory * y[i];
        px[i] *= g[i];
        py[i] *= g[i];
        x[i] += sl.xcen;
        y[i] += sl.ycen;
        px[i] += sl.pxcen;
        py[i] += sl.pycen;
    }
    for (int i = mpart - 1; i >= 0; i--) {  // beamlets of nbins mirrored phases
        const int d = nbins * i;
        for (int j = nbins - 1; j >= 0; j--) {
            g[d + j] = g[i];
            x[d + j] = x[i];
            y[d + j] = y[i];
            px[d + j] = px[i];
            py[d + j] = py[i];
            th[d + j] = th[i] + j * dtheta;
        }
    }
    for (int i = 0; i < npart; i++) th[i] *= theta0;
    const double pi = 2 * asin(1.);
    double bl = sl.bunch;
    if (bl != 0) {  // invert sin(x)/x for the requested bunching
        double lo = 0, hi = 0.9999, mid = 0;
        while ((hi - lo) > 1e-6) {
            mid = 0.5 * (lo + hi);
            const double probe = sin((1 - mid) * pi) / ((1 - mid) * pi);
            if (bl < probe) hi = mid; else lo = mid;
        }
        bl = mid;
    }
    if ((bl != 0) || (sl.emod != 0)) {
        for (int i = 0; i < npart; i++) {
            g[i] -= sl.emod * sin(th[i] - sl.emodphase);
            if (sl.bunch > 0.1)
                th[i] += -bl * th[i] + bl * pi + sl.bunchphase + pi;
            else
                th[i] -= 2 * sl.bunch * sin(th[i] - sl.bunchphase);
        }
    }
}
}  // namespace

unsigned int shot_noise_seed(int seed, int rank) {
    Ran2 r((unsigned int)seed);
    double v = 0;
    for (int i = 0; i <= rank; i++) v = r.next();
    return (unsigned int)(int)round(v * 1e9);
}

long long shot_noise_draws(Ran2 &stream, int npart, int nbins, double ne, std::vector<double> &out) {
    if (ne <= 0) return 0;
    const int mpart = npart / nbins;
    double nbl = ne / (double)mpart;
    if (nbl < 1) nbl = 1;
    const double twopi = 4 * asin(1.);
    const long long nd = (long long)((nbins - 1) / 2) * mpart;
    for (int ih = 0; ih < (nbins - 1) / 2; ih++) {
        for (int i1 = 0; i1 < mpart; i1++) {
            const double phi = 4 * asin(1) * stream.next();
            double an = sqrt(-log(stream.next()) / nbl) * 2 / (double)(ih + 1);
            if (an > twopi) an = fmod(an, twopi);
            out.push_back(phi);
            out.push_back(an);
        }
    }
    return nd;
}

void shot_noise(Ran2 &stream, double *th, int npart, int nbins, double ne) {
    std::vector<double> d;
    d.reserve(2 * (size_t)((nbins - 1) / 2) * (size_t)(npart / nbins));
    if (shot_noise_draws(stream, npart, nbins, ne, d) == 0) return;
    const int mpart = npart / nbins;
    std::vector<double> dth(npart, 0.0);
    for (int ih = 0; ih < (nbins - 1) / 2; ih++) {
        for (int i1 = 0; i1 < mpart; i1++) {
            const double phi = d[2 * ((size_t)ih * mpart + i1)], an = d[2 * ((size_t)ih * mpart + i1) + 1];
            for (int i2 = 0; i2 < nbins; i2++) {
                const int k = i1 * nbins + i2;
                dth[k] -= an * sin(th[k] * (ih + 1) + phi);
            }
        }
    }
    for (int i = 0; i < npart; i++) th[i] += dth[i];
}

void gauss_field(double *field, int ngrid, double gridmax, double lambda, double power,
                 double w0, double z0in, double phase, double xcen, double ycen, double xangle,
                 double yangle, int harm) {
    typedef std::complex<double> cd;
    const double k = 4. * asin(1) / lambda * (double)harm;
    const double z0 = -z0in;
    const double zr = w0 * w0 * k * 0.5;
    const double f0 = sqrt(k * zr / (zr * zr + z0 * z0));
    const cd qz(-z0, zr);
    const cd coef = cd(0, 1) * 0.5 * k / qz;
    const double unit = sqrt(power * kVacimpH) * k / kEevH;
    const double norm = f0 / sqrt(2. * asin(1.) * 1 * 1 * (1 << 0));
    const cd zscale = unit * norm * cd(cos(phase), sin(phase));
    const double xmid = gridmax + xcen, ymid = gridmax + ycen;
    const double dxy = 2. * gridmax / (ngrid - 1.);
    const double kx = k * xangle, ky = k * yangle;
    for (int iy = 0; iy < ngrid; iy++) {
        const double yy = iy * dxy - ymid;
        const double y2 = yy * yy, phiy = ky * yy;
        for (int ix = 0; ix < ngrid; ix++) {
            const double xx = ix * dxy - xmid;
            const double r2 = y2 + xx * xx;
            const cd v = zscale * exp(-coef * r2 + cd(0, phiy + kx * xx)) * 1.0 * 1.0;
            field[2 * (iy * ngrid + ix)] = v.real();
            field[2 * (iy * ngrid + ix) + 1] = v.imag();
        }
    }
}

// ---------------------------------------------------------------------------
// Lattice parser and unroll
// ---------------------------------------------------------------------------
namespace {
std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}
std::vector<std::string> split(const std::string &s, char sep) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (ch == sep) {
            out.push_back(trim(cur));
            cur.clear();
        } else {
            cur += ch;
        }
    }
    out.push_back(trim(cur));
    return out;
}
enum Kind { kUnd = 0, kDrift = 1, kQuad = 2, kCorr = 3, kChic = 4, kMark = 5, kPhase = 6 };
struct Element {
    Kind kind;
    double z = 0, l = 0;
    // undulator
    double aw = 0, lambdau = 0, nwig = 0, kx = 0, ky = 1, ax = 0, ay = 0, gradx = 0, grady = 0;
    bool helical = false;
    // quad / corrector / chicane / marker / phase shifter
    double k1 = 0, dx = 0, dy = 0, cx = 0, cy = 0, delay = 0, lb = 0, ld = 0, phi = 0;
    int action = 0;
};
struct Def {
    std::string label, type, args;
};
bool to_bool(const std::string &v) {
    std::string t = trim(v);
    return t == "1" || t == "true" || t == "t";
}
bool parse_element(const Def &d, double z, Element &e, std::string &err) {
    const std::string ty = d.type;
    if (ty == "undu") e.kind = kUnd;
    else if (ty == "drif") e.kind = kDrift;
    else if (ty == "quad") e.kind = kQuad;
    else if (ty == "corr") e.kind = kCorr;
    else if (ty == "chic") e.kind = kChic;
    else if (ty == "mark") e.kind = kMark;
    else if (ty == "phas") e.kind = kPhase;
    else {
        err = "unknown element type " + ty;
        return false;
    }
    e.z = z;
    bool haskxy = false;
    for (const std::string &kv : split(d.args, ',')) {
        if (kv.empty()) continue;
        size_t p = kv.find('=');
        if (p == std::string::npos) continue;
        const std::string k = trim(kv.substr(0, p));
        const std::string v = trim(kv.substr(p + 1));
        if (!v.empty() && v[0] == '@') {
            err = "sequence references are not supported: " + kv;
            return false;
        }
        const double f = atof(v.c_str());
        const int tag = atoi(v.c_str()) != 0 ? 1 : 0;
        if (k == "l") e.l = f;
        else if (k == "lambdau") e.lambdau = f;
        else if (k == "aw") e.aw = f;
        else if (k == "nwig") e.nwig = f;
        else if (k == "kx") { e.kx = f; haskxy = true; }
        else if (k == "ky") { e.ky = f; haskxy = true; }
        else if (k == "ax") e.ax = f;
        else if (k == "ay") e.ay = f;
        else if (k == "gradx") e.gradx = f;
        else if (k == "grady") e.grady = f;
        else if (k == "helical") e.helical = to_bool(v);
        else if (k == "k1") e.k1 = f;
        else if (k == "dx") e.dx = f;
        else if (k == "dy") e.dy = f;
        else if (k == "cx") e.cx = f;
        else if (k == "cy") e.cy = f;
        else if (k == "delay") e.delay = f;
        else if (k == "lb") e.lb = f;
        else if (k == "ld") e.ld = f;
        else if (k == "phi") e.phi = f;
        else if (k == "dumpfield") e.action |= 1 * tag;
        else if (k == "dumpbeam") e.action |= 2 * tag;
        else if (k == "sort") e.action |= 4 * tag;
        else if (k == "stop") e.action |= 8 * tag;
    }
    if (e.kind == kUnd) {  // LatticeParser::parseID defaults (:243-331)
        e.l = e.lambdau * e.nwig;
        if (!haskxy) {
            e.kx = e.helical ? 0.5 : 0;
            e.ky = e.helical ? 0.5 : 1;
        }
    }
    if (e.kind == kMark) e.l = 0;
    return true;
}
}  // namespace

LatticeResult parse_and_unroll(const std::string &text, const std::string &beamline_in,
                               double delz, double lambda0, double gamma0) {
    LatticeResult res;
    // coarse parsing: strip comments, join, lower-case, split at ';'
    std::string all;
    {
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            all += " " + line;
        }
    }
    for (char &ch : all) ch = (char)tolower((unsigned char)ch);
    std::vector<Def> defs;
    size_t p;
    while ((p = all.find(';')) != std::string::npos) {
        std::string st = trim(all.substr(0, p));
        all.erase(0, p + 1);
        size_t c = st.find(':'), e = st.find('='), b = st.find('{'), q = st.find('}');
        if (c == std::string::npos || e == std::string::npos || b == std::string::npos ||
            q == std::string::npos) {
            res.error = "invalid lattice statement: " + st;
            return res;
        }
        Def d;
        d.label = trim(st.substr(0, c));
        d.type = trim(st.substr(c + 1, e - c - 1)).substr(0, 4);
        d.args = trim(st.substr(b + 1, q - b - 1));
        for (const Def &o : defs)
            if (o.label == d.label) {
                res.error = "label redefined: " + d.label;
                return res;
            }
        defs.push_back(d);
    }
    auto find = [&](const std::string &l) {
        for (int i = (int)defs.size() - 1; i >= 0; i--)
            if (defs[i].label == l) return i;
        return -1;
    };
    // unroll the beamline (LatticeParser::unroll :612-656)
    std::string bl = beamline_in;
    for (char &ch : bl) ch = (char)tolower((unsigned char)ch);
    struct Seq {
        int def;
        int zref;
        double zoff;
    };
    std::vector<Seq> seq;
    int refele = -1;
    std::string err;
    std::function<bool(int, int)> unroll = [&](int idx, int depth) -> bool {
        if (depth < 0) {
            err = "too many nested lines";
            return false;
        }
        const int save = refele;
        for (std::string item : split(defs[idx].args, ',')) {
            if (item.empty()) continue;
            int count = 1;
            size_t star = item.find('*');
            if (star != std::string::npos) {
                count = atoi(trim(item.substr(0, star)).c_str());
                item = trim(item.substr(star + 1));
            }
            bool reset = false;
            double off = 0;
            size_t at = item.find('@');
            if (at != std::string::npos) {
                off = atof(trim(item.substr(at + 1)).c_str());
                item = trim(item.substr(0, at));
                reset = true;
            }
            const int ix = find(item);
            if (ix < 0) {
                err = "undefined element in beamline: " + item;
                return false;
            }
            for (int j = 0; j < count; j++) {
                if (defs[ix].type == "line") {
                    if (!unroll(ix, depth - 1)) return false;
                } else {
                    seq.push_back({ix, reset ? save : refele, reset ? off : 0.0});
                    refele++;
                }
            }
        }
        return true;
    };
    const int li = find(bl);
    if (li < 0 || defs[li].type != "line") {
        res.error = "lattice does not contain the beamline " + bl;
        return res;
    }
    if (!unroll(li, 19)) {
        res.error = err;
        return res;
    }
    std::vector<Element> lat;
    for (size_t i = 0; i < seq.size(); i++) {
        const double z = seq[i].zref < 0 ? 0 + seq[i].zoff
                                         : lat[seq[i].zref].z + lat[seq[i].zref].l + seq[i].zoff;
        Element e;
        if (!parse_element(defs[seq[i].def], z, e, res.error)) return res;
        lat.push_back(e);
    }
    // layout points and unroll into integration steps (Lattice.cpp:340-516)
    std::map<double, int> layout;
    for (const Element &e : lat) {
        layout[e.z] = 1;
        layout[e.z + e.l] = 1;
    }
    auto locate = [&](double z0, double z1, Kind k) {
        const double zm = 0.5 * (z0 + z1);
        for (size_t i = 0; i < lat.size(); i++)
            if (zm > lat[i].z && zm < lat[i].z + lat[i].l && lat[i].kind == k) return (int)i;
        return -1;
    };
    std::vector<g4mi_step_params> &v = res.steps;
    for (auto it = layout.begin(); std::next(it) != layout.end(); ++it) {
        const double z0 = it->first, z1 = std::next(it)->first;
        double dz = z1 - z0;
        const int u = locate(z0, z1, kUnd);
        if (u < 0) {
            g4mi_step_params s{};
            s.dz = z1 - z0;
            v.push_back(s);
            continue;
        }
        int nz = (int)round(dz / delz);
        if (nz == 0) nz = 1;
        dz = dz / (double)nz;
        const Element &e = lat[u];
        const double ku = 4. * asin(1) / e.lambdau;
        for (int iz = 0; iz < nz; iz++) {
            g4mi_step_params s{};
            s.dz = dz;
            s.aw = e.aw;
            s.ku = ku;
            s.kx = ku * ku * e.kx;
            s.ky = ku * ku * e.ky;
            s.gradx = e.gradx;
            s.grady = e.grady;
            s.ax = e.ax;
            s.ay = e.ay;
            s.helical = e.helical ? 1 : 0;
            v.push_back(s);
        }
    }
    const int nz = (int)v.size();
    if (nz == 0) {
        res.error = "empty lattice";
        return res;
    }
    for (int i = 1; i < nz; i++) v[i].z = v[i - 1].z + v[i - 1].dz;
    std::vector<double> delay(nz, 0), lb(nz, 0), ld(nz, 0), lt(nz, 0), ps(nz, 0);
    int lastChicane = -1;
    for (int i = 0; i < nz; i++) {
        const double z0 = v[i].z, z1 = z0 + v[i].dz;
        const bool inUnd = v[i].aw > 0;
        bool inQuad = false;
        int k = locate(z0, z1, kQuad);
        if (k >= 0) {
            v[i].qf = lat[k].k1;
            v[i].qx = lat[k].dx;
            v[i].qy = lat[k].dy;
            inQuad = true;
        }
        k = locate(z0, z1, kChic);
        if (k >= 0 && k != lastChicane && !inUnd && !inQuad) {
            delay[i] = lat[k].delay;
            lb[i] = lat[k].lb;
            ld[i] = lat[k].ld;
            lt[i] = lat[k].l;
            lastChicane = k;
        }
        k = locate(z0, z1, kCorr);
        if (k >= 0) {
            v[i].cx += lat[k].cx;
            v[i].cy += lat[k].cy;
        }
        k = locate(z0, z1, kPhase);
        if (k >= 0) ps[i] = lat[k].phi;
        for (const Element &e : lat) {
            const double d = e.z - z0;
            if (d * d < 1e-6 && e.kind == kMark) {
                v[i].marker = e.action;
                break;
            }
        }
    }
    // slippage per step (Lattice::calcSlippage :153-197)
    double Lz = 0;
    for (int i = 0; i < nz; i++) {
        if (v[i].aw > 0) {
            const double lr = 2 * gamma0 * gamma0 * lambda0 / (1 + v[i].aw * v[i].aw);
            v[i].slip = v[i].dz / lr;
            v[i].phaseshift = 0;
            if ((Lz > 0) && (i > 0)) {
                v[i - 1].slip += floor(Lz / 2 / gamma0 / gamma0 / lambda0) + 1;
                Lz = 0;
            }
        } else {
            Lz += v[i].dz;
            v[i].slip = floor(delay[i] / lambda0);
            v[i].phaseshift = -ps[i];
        }
    }
    v[nz - 1].slip += floor(Lz / 2 / gamma0 / gamma0 / lambda0) + 1;
    for (int i = 0; i < nz; i++) {  // chicane delay -> bend angle
        v[i].chic_angle = delay[i];
        v[i].chic_lb = lb[i];
        v[i].chic_ld = ld[i];
        v[i].chic_lt = lt[i];
        if (delay[i] != 0) {
            const double target = fabs(delay[i]);
            double tmin = 0, tmax = asin(1) - 0.001, th = 0.5 * (tmax + tmin);
            for (;;) {
                th = 0.5 * (tmax + tmin);
                const double d = 4 * lb[i] * (th / sin(th) - 1) + 2 * ld[i] * (1 / cos(th) - 1);
                if (d > target) tmax = th; else tmin = th;
                if (fabs(target - d) < 1e-15) break;
            }
            v[i].chic_angle = th;
        }
    }
    return res;
}

}  // namespace g4mi

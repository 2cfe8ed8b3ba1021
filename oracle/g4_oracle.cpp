/*
 * FP64 CPU oracle for the GENESIS 1.3 v4 undulator-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Each function
 * restates one reference routine (cited file:line) in the reference's own
 * loop and operation order, so that on the same inputs it reproduces the
 * reference's FP64 results.  Parity status of each function is documented
 * next to it ("pinned by ..."); the tests in tests/test_oracle_*.py hold the
 * checks.  Built with g++ -O2 and no -ffast-math / no FMA contraction so the
 * arithmetic is the plain IEEE sequence the reference's default build uses.
 */
#include "g4_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using std::complex;
using std::vector;

/* Physical constants as defined in src/Main/GenMain.cpp:63-68. */
static const double k_vacimp = 376.73;
static const double k_eev = 510998.95069;
static const double k_ce = 4.8032045e-11;

/* ======================================================================
 * Leaf numerics
 * ==================================================================== */

/* Numerical Recipes ran2 (L'Ecuyer + Bays-Durham shuffle).
 * Follows src/Util/RandomU.cpp:17-35 (seeding, seed clamped to >=1) and
 * :46-73 (draw).  Pinned bit-exactly against the compiled reference
 * RandomU.cpp (oracle/_ref). */
struct Ran2 {
    int iv[32], iy, s1, s2;
    explicit Ran2(unsigned int istart) {
        s1 = (1 < (int)istart) ? (int)istart : 1;
        if (istart > 0x7fffffffu) s1 = (int)istart;  /* int conversion, as reference */
        s2 = s1;
        for (int i = 32 + 7; i >= 0; i--) {
            int k = s1 / 53668;
            s1 = 40014 * (s1 - k * 53668) - k * 12211;
            if (s1 < 0) s1 += 2147483563;
            if (i < 32) iv[i] = s1;
        }
        iy = iv[0];
    }
    double next() {
        const double am = 1. / 2147483563;
        const double rnmx = 1. - 1.2e-40;
        int k = s1 / 53668;
        s1 = 40014 * (s1 - k * 53668) - k * 12211;
        if (s1 < 0) s1 = s1 + 2147483563;
        k = s2 / 52774;
        s2 = 40692 * (s2 - k * 52774) - k * 3791;
        if (s2 < 0) s2 = s2 + 2147483399;
        int j = iy / 67108862;
        iy = iv[j] - s2;
        iv[j] = s1;
        if (iy < 1) iy = iy + 2147483562;
        return (am * iy < rnmx) ? am * iy : rnmx;
    }
};

extern "C" void orc_ran2(unsigned int seed, int n, double *out) {
    Ran2 r(seed);
    for (int i = 0; i < n; i++) out[i] = r.next();
}

/* Radical-inverse (Hammersley/Halton) sequence in a prime base.
 * Follows src/Util/Hammerslay.cpp:15-49: the counter is pre-incremented,
 * digits are peeled with unsigned integer division.  Pinned bit-exactly
 * against oracle/_ref and by the base-2 closed form 1/2,1/4,3/4,... */
static const double k_primes[26] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41,
                                    43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101};
struct Halton {
    double base;
    unsigned int idx;
    explicit Halton(int b) : base(k_primes[b]), idx(0) {}
    double next() {
        double xs = 0, xsi = 1.0;
        unsigned int i2 = ++idx, i1;
        do {
            xsi = xsi / base;
            i1 = i2 / (unsigned int)base;
            xs += (i2 - base * i1) * xsi;
            i2 = i1;
        } while (i2 > 0);
        return xs;
    }
};

extern "C" void orc_hammersley(int base_index, unsigned int start, int n, double *out) {
    Halton h(base_index);
    h.idx = start;
    for (int i = 0; i < n; i++) out[i] = h.next();
}

/* Inverse complementary error function, T. Ooura's published rational +
 * one Newton/Halley correction (the method src/Util/Inverfc.cpp:16-64
 * cites).  Coefficients are Ooura's; they are the algorithm's table.
 * Pinned bit-exactly against oracle/_ref and to ~1e-15 by inverting
 * std::erfc. */
extern "C" double orc_inverfc(double y) {
    double z = y;
    if (y > 1) z = 2 - y;
    double w = 0.916461398268964 - log(z);
    double u = sqrt(w);
    double s = (log(u) + 0.488826640273108) / w;
    double t = 1 / (u + 0.231729200323405);
    double x = u * (1 - s * (s * 0.124610454613712 + 0.5)) -
               ((((-0.0728846765585675 * t + 0.269999308670029) * t +
                  0.150689047360223) * t + 0.116065025341614) * t +
                0.499999303439796) * t;
    t = 3.97886080735226 / (x + 3.97886080735226);
    u = t - 0.5;
    s = (((((((((0.00112648096188977922 * u + 1.05739299623423047e-4) * u -
                0.00351287146129100025) * u - 7.71708358954120939e-4) * u +
              0.00685649426074558612) * u + 0.00339721910367775861) * u -
            0.011274916933250487) * u - 0.0118598117047771104) * u +
          0.0142961988697898018) * u + 0.0346494207789099922) * u +
        0.00220995927012179067;
    s = ((((((((((((s * u - 0.0743424357241784861) * u - 0.105872177941595488) * u +
                  0.0147297938331485121) * u + 0.316847638520135944) * u +
                0.713657635868730364) * u + 1.05375024970847138) * u +
              1.21448730779995237) * u + 1.16374581931560831) * u +
            0.956464974744799006) * u + 0.686265948274097816) * u +
          0.434397492331430115) * u + 0.244044510593190935) * t -
        z * exp(x * x - 0.120782237635245222);
    x = x + s * (x * s + 1);
    if (y > 1) x = -x;
    return x;
}

/* Bessel J0/J1 rational approximations (Numerical Recipes bessj0/bessj1)
 * and J_n by upward recurrence or Miller's downward recurrence, as
 * src/Util/BesselJ.cpp:16-159.  The NR coefficients are the published
 * table of that method.  Pinned bit-exactly against oracle/_ref and to
 * the approximation's ~1e-8 by std::cyl_bessel_j. */
static double bj0(double x) {
    if (fabs(x) < 8) {
        double y = x * x;
        return (57568490574.e0 + y * (-13362590354.e0 + y * (651619640.7e0 +
                y * (-11214424.18e0 + y * (77392.33017e0 + y * (-184.9052456e0)))))) /
               (57568490411.e0 + y * (1029532985.e0 + y * (9494680.718e0 +
                y * (59272.64853e0 + y * (267.8532712e0 + y * 1.e0)))));
    }
    double ax = fabs(x), z = 8 / ax, y = z * z, xx = ax - 0.785398164;
    return sqrt(0.636619772 / ax) *
           (cos(xx) * (1.e0 + y * (-.1098628627e-2 + y * (.2734510407e-4 +
                       y * (-.2073370639e-5 + y * .2093887211e-6)))) -
            z * sin(xx) * (-.1562499995e-1 + y * (.1430488765e-3 +
                       y * (-.6911147651e-5 + y * (.7621095161e-6 + y * -.934945152e-7)))));
}
static double bj1(double x) {
    if (fabs(x) < 8) {
        double y = x * x;
        return x * (72362614232.e0 + y * (-7895059235.e0 + y * (242396853.1e0 +
                    y * (-2972611.439e0 + y * (15704.48260e0 + y * (-30.16036606e0)))))) /
               (144725228442.e0 + y * (2300535178.e0 + y * (18583304.74e0 +
                y * (99447.43394e0 + y * (376.9991397e0 + y * 1.e0)))));
    }
    double ax = fabs(x), z = 8 / ax, y = z * z, xx = ax - 2.356194491;
    return sqrt(0.636619772 / ax) *
           (cos(xx) * (1.e0 + y * (.183105e-2 + y * (-.3516396496e-4 +
                       y * (.2457520174e-5 + y * (-.240337019e-6))))) -
            z * sin(xx) * (.04687499995e0 + y * (-.2002690873e-3 +
                       y * (.8449199096e-5 + y * (-.88228987e-6 + y * .105787412e-6)))));
}
extern "C" double orc_besselj(int n, double x) {
    if (n == 0) return bj0(x);
    if (n == 1) return bj1(x);
    if (x == 0) return 0;
    double ax = fabs(x), bessj;
    if (ax > (double)n) {
        double tox = 2 / ax, bjm = bj0(ax), bj = bj1(ax);
        for (int j = 1; j < n; j++) {
            double bjp = j * tox * bj - bjm;
            bjm = bj;
            bj = bjp;
        }
        bessj = bj;
    } else {
        double tox = 2. / ax;
        int m = 2 * (n + floor(sqrt((double)(40 * n))) / 2);
        bessj = 0;
        int jsum = 0;
        double sum = 0, bjp = 0, bj = 1;
        for (int j = m; j > 0; j--) {
            double bjm = j * tox * bj - bjp;
            bjp = bj;
            bj = bjm;
            if (fabs(bj) > 1e10) {
                bj *= 1e-10;
                bjp *= 1e-10;
                bessj *= 1e-10;
                sum *= 1e-10;
            }
            if (jsum != 0) sum += bj;
            jsum = 1 - jsum;
            if (j == n) bessj = bjp;
        }
        sum = 2 * sum - bj;
        bessj = bessj / sum;
    }
    if ((x < 0) && ((n % 2) == 1)) bessj = -bessj;
    return bessj;
}

/* ======================================================================
 * Undulator helpers (src/Core/Undulator.cpp)
 * ==================================================================== */

/* Coupling factor, Undulator.cpp:141-167 (helical: aw for h=1; planar odd
 * h: aw (J_{(h-1)/2}(xi) - J_{(h+1)/2}(xi)) (-1)^{(h-1)/2}; even h: 0). */
extern "C" double orc_fc(const orc_step *st, int h) {
    double coup = st->aw;
    if (st->helical != 0) return (h == 1) ? coup : 0;
    double xi = st->aw * st->aw;
    xi = 0.5 * xi / (1 + xi) * (double)h;
    if ((h % 2) == 1) {
        int h0 = (h - 1) / 2, h1 = h0 + 1;
        return coup * (orc_besselj(h0, xi) - orc_besselj(h1, xi)) * pow(-1., h0);
    }
    return 0;
}
/* Undulator.cpp:178-182 */
extern "C" double orc_faw(const orc_step *st, double x, double y) {
    double dx = x - st->ax, dy = y - st->ay;
    return (1 + 0.5 * (st->kx * dx * dx + st->ky * dy * dy) + st->gradx * dx + st->grady * dy);
}
/* Undulator.cpp:170-174 */
extern "C" double orc_faw2(const orc_step *st, double x, double y) {
    double dx = x - st->ax, dy = y - st->ay;
    return (1 + st->kx * dx * dx + st->ky * dy * dy + 2 * (st->gradx * dx + st->grady * dy));
}

/* ======================================================================
 * Loaders
 * ==================================================================== */

/* Quiet loading of one slice, src/Loading/QuietLoading.cpp:50-260 for the
 * non-one4one case (Hammersley bases 2,3,5,7,11,13 for theta, gamma, x, y,
 * px, py; init at :15-46 with base = {0..5}).  The six sequences restart at
 * round(1e9 * first ran2 draw of seed islice) (:55-65).  Arrays must hold
 * npart entries.  Pinned bit-exactly against oracle/_ref. */
static void quiet_normalise(double *g, double *th, double *x, double *y, double *px, double *py,
                            const orc_slice_params *sl, int npart, int nbins, double theta0);

extern "C" void orc_quiet_load(double *g, double *th, double *x, double *y, double *px,
                               double *py, const orc_slice_params *sl, int npart, int nbins,
                               double theta0, int islice) {
    Ran2 seed((unsigned int)islice);
    int iseed = (int)round(seed.next() * 1e9);
    Halton st(0), sg(1), sx(2), sy(3), spx(4), spy(5);
    st.idx = sg.idx = sx.idx = sy.idx = spx.idx = spy.idx = (unsigned int)iseed;
    int mpart = npart / nbins;
    double dtheta = 1. / (double)nbins;
    for (int i = 0; i < mpart; i++) {
        th[i] = st.next() * dtheta;
        g[i] = orc_inverfc(2 * sg.next());
        x[i] = orc_inverfc(2 * sx.next());
        y[i] = orc_inverfc(2 * sy.next());
        px[i] = orc_inverfc(2 * spx.next());
        py[i] = orc_inverfc(2 * spy.next());
    }
    quiet_normalise(g, th, x, y, px, py, sl, npart, nbins, theta0);
}

/* one4one loading, QuietLoading::init(true, {seed, rank}) (:15-40) and
 * loadQuiet (:50-260) for every slice in order: the six sequences are one
 * ran2 stream (RandomU::set does nothing), seeded with round(1e9 u), u the
 * (rank+1)-th draw of ran2(seed); nbins = 1; theta0 = 2 pi sample
 * (LoadBeam.cpp:129-132). */
extern "C" void orc_quiet_load_one4one(int seed, int rank, int nslice, const long long *off,
                                       const orc_slice_params *sl, double theta0, double *g, double *th,
                                       double *x, double *y, double *px, double *py) {
    Ran2 rseed((unsigned int)seed);
    double val = 0;
    for (int i = 0; i <= rank; i++) val = rseed.next();
    Ran2 st((unsigned int)(int)round(val * 1e9));
    for (int j = 0; j < nslice; j++) {
        long long a = off[j];
        int n = (int)(off[j + 1] - a);
        for (int i = 0; i < n; i++) {
            th[a + i] = st.next();
            g[a + i] = orc_inverfc(2 * st.next());
            x[a + i] = orc_inverfc(2 * st.next());
            y[a + i] = orc_inverfc(2 * st.next());
            px[a + i] = orc_inverfc(2 * st.next());
            py[a + i] = orc_inverfc(2 * st.next());
        }
        quiet_normalise(g + a, th + a, x + a, y + a, px + a, py + a, &sl[j], n, 1, theta0);
    }
}

static void quiet_normalise(double *g, double *th, double *x, double *y, double *px, double *py,
                            const orc_slice_params *sl, int npart, int nbins, double theta0) {
    int mpart = npart / nbins;
    double dtheta = 1. / (double)nbins;
    double z = 0, zz = 0, norm = 0;
    if (mpart > 0) norm = 1. / (double)mpart;
    for (int i = 0; i < mpart; i++) {
        z += g[i];
        zz += g[i] * g[i];
    }
    z *= norm;
    zz = sqrt(fabs(zz * norm - z * z));
    if (zz > 0) zz = 1 / zz;
    zz *= sl->delgam;
    for (int i = 0; i < mpart; i++) g[i] = (g[i] - z) * zz + sl->gamma;

    /* x/px then y/py decorrelation (:110-156) */
    double *cz[2] = {x, y}, *cp[2] = {px, py};
    for (int c = 0; c < 2; c++) {
        double *a = cz[c], *b = cp[c];
        z = 0; zz = 0;
        double p = 0, zp = 0;
        for (int i = 0; i < mpart; i++) {
            z += a[i];
            zz += a[i] * a[i];
            p += b[i];
            zp += a[i] * b[i];
        }
        z *= norm;
        p *= norm;
        zz = sqrt(fabs(zz * norm - z * z));
        if (zz > 0) zz = 1. / zz;
        zp = (zp * norm - z * p) * zz * zz;
        for (int i = 0; i < mpart; i++) {
            b[i] = b[i] - p;
            b[i] -= zp * a[i];
            a[i] = (a[i] - z) * zz;
        }
    }
    /* momenta to unit rms (:159-180) */
    z = 0; zz = 0;
    double p = 0, pp = 0;
    for (int i = 0; i < mpart; i++) {
        z += px[i];
        zz += px[i] * px[i];
        p += py[i];
        pp += py[i] * py[i];
    }
    z *= norm;
    p *= norm;
    zz = sqrt(fabs(zz * norm - z * z));
    if (zz > 0) zz = 1. / zz;
    pp = sqrt(fabs(pp * norm - p * p));
    if (pp > 0) pp = 1. / pp;
    for (int i = 0; i < mpart; i++) {
        px[i] = (px[i] - z) * zz;
        py[i] = (py[i] - p) * pp;
    }
    /* physical scaling (:185-203) */
    double sigx = sqrt(sl->ex * sl->betax / sl->gamma);
    double sigy = sqrt(sl->ey * sl->betay / sl->gamma);
    double sigpx = sqrt(sl->ex / sl->betax / sl->gamma);
    double sigpy = sqrt(sl->ey / sl->betay / sl->gamma);
    double corx = -sl->alphax / sl->betax;
    double cory = -sl->alphay / sl->betay;
    for (int i = 0; i < mpart; i++) {
        x[i] *= sigx;
        y[i] *= sigy;
        px[i] = sigpx * px[i] + corx * x[i];
        py[i] = sigpy * py[i] + cory * y[i];
        px[i] *= g[i];
        py[i] *= g[i];
        x[i] += sl->xcen;
        y[i] += sl->ycen;
        px[i] += sl->pxcen;
        py[i] += sl->pycen;
    }
    /* beamlet mirroring (:207-218), backwards so it works in place */
    for (int i = mpart; i > 0; i--) {
        int i1 = i - 1, i2 = nbins * i1;
        for (int j = 0; j < nbins; j++) {
            g[i2 + j] = g[i1];
            x[i2 + j] = x[i1];
            y[i2 + j] = y[i1];
            px[i2 + j] = px[i1];
            py[i2 + j] = py[i1];
            th[i2 + j] = th[i1] + j * dtheta;
        }
    }
    for (int i = 0; i < npart; i++) th[i] *= theta0;
    /* bunching / energy modulation (:227-259) */
    double pi = 2 * asin(1.);
    double blocal = sl->bunch;
    if (blocal != 0) {
        double btar = blocal, blow = 0, bupp = 0.9999, bmid = 0;
        while ((bupp - blow) > 1e-6) {
            bmid = 0.5 * (blow + bupp);
            double bprobe = sin((1 - bmid) * pi) / ((1 - bmid) * pi);
            if (btar < bprobe) bupp = bmid; else blow = bmid;
        }
        blocal = bmid;
    }
    if ((blocal != 0) || (sl->emod != 0)) {
        if (sl->bunch > 0.1) {
            for (int i = 0; i < npart; i++) {
                g[i] -= sl->emod * sin(th[i] - sl->emodphase);
                th[i] += -blocal * th[i] + blocal * pi + sl->bunchphase + pi;
            }
        } else {
            for (int i = 0; i < npart; i++) {
                g[i] -= sl->emod * sin(th[i] - sl->emodphase);
                th[i] -= 2 * sl->bunch * sin(th[i] - sl->bunchphase);
            }
        }
    }
}

/* Shot noise, src/Loading/ShotNoise.cpp:16-68.  One ran2 stream seeded by
 * round(1e9 * the (rank+1)-th draw of ran2(seed)), consumed slice after
 * slice (two draws per beamlet per harmonic).  Pinned bit-exactly against
 * oracle/_ref. */
struct ShotNoiseStream {
    Ran2 *r = nullptr;
};
extern "C" void *orc_shotnoise_new(int seed, int rank) {
    Ran2 rs((unsigned int)seed);
    double val = 0;
    for (int i = 0; i <= rank; i++) val = rs.next();
    val *= 1e9;
    int locseed = (int)round(val);
    ShotNoiseStream *s = new ShotNoiseStream;
    s->r = new Ran2((unsigned int)locseed);
    return s;
}
extern "C" void orc_shotnoise_free(void *h) {
    ShotNoiseStream *s = (ShotNoiseStream *)h;
    delete s->r;
    delete s;
}
extern "C" void orc_shotnoise(void *h, double *th, int npart, int nbins, double ne) {
    ShotNoiseStream *s = (ShotNoiseStream *)h;
    if (ne <= 0) return;
    int mpart = npart / nbins;
    double nbl = ne / (double)mpart;
    if (nbl < 1) nbl = 1;
    vector<double> work(npart, 0.0);
    double twopi = 4 * asin(1.);
    for (int ih = 0; ih < (nbins - 1) / 2; ih++) {
        for (int i1 = 0; i1 < mpart; i1++) {
            double phi = 4 * asin(1) * s->r->next();
            double an = sqrt(-log(s->r->next()) / nbl) * 2 / (double)(ih + 1);
            if (an > twopi) an = fmod(an, twopi);
            for (int i2 = 0; i2 < nbins; i2++) {
                int idx = i1 * nbins + i2;
                work[idx] -= an * sin(th[idx] * (ih + 1) + phi);
            }
        }
    }
    for (int i = 0; i < npart; i++) th[i] += work[i];
}

/* Fundamental Gaussian seed field, src/Util/GaussHermite.cpp:20-63 with
 * nx=ny=0, harm=1, no tilt.  'gridmax' is the &field dgrid (half width).
 * Pinned bit-exactly against oracle/_ref and by the power normalisation
 * sum |E|^2 (dgrid eev/ks)^2/vacimp = P. */
extern "C" void orc_load_gauss(double *field, int ngrid, double gridmax, double lambda,
                               double power, double w0, double z0in, double phase,
                               double xcen, double ycen) {
    double k = 4. * asin(1) / lambda * (double)1;
    double z0 = -z0in;
    double zr = w0 * w0 * k * 0.5;
    double f0 = sqrt(k * zr / (zr * zr + z0 * z0));
    complex<double> qz(-z0, zr);
    complex<double> coef = complex<double>(0, 1) * 0.5 * k / qz;
    double unit = sqrt(power * k_vacimp) * k / k_eev;
    double norm = f0 / sqrt(2. * asin(1.) * 1 * 1 * (1 << 0));
    complex<double> zscale = unit * norm * complex<double>(cos(phase), sin(phase));
    double xmid = gridmax + xcen, ymid = gridmax + ycen;
    double dxy = 2. * gridmax / (ngrid - 1.);
    double kx = k * 0, ky = k * 0;
    for (int iy = 0; iy < ngrid; iy++) {
        double yy = iy * dxy - ymid;
        double Hy = 1;
        double y2 = yy * yy;
        double phiy = ky * yy;
        for (int ix = 0; ix < ngrid; ix++) {
            double xx = ix * dxy - xmid;
            double Hx = 1;
            double r2 = y2 + xx * xx;
            complex<double> phi(0, phiy + kx * xx);
            complex<double> v = zscale * exp(-coef * r2 + phi) * Hx * Hy;
            field[2 * (iy * ngrid + ix)] = v.real();
            field[2 * (iy * ngrid + ix) + 1] = v.imag();
        }
    }
}

/* ======================================================================
 * Lattice unroll (src/Lattice/Lattice.cpp:53-149, 153-197, 340-516)
 * ==================================================================== */
extern "C" int orc_unroll(int nele, const int *type, const double *zz, const double *ll,
                          const double *par, double delz, double lambda, double gamma,
                          orc_step *out, int max_steps) {
    std::map<double, int> layout;                 /* Lattice.cpp:38-44 */
    for (int i = 0; i < nele; i++) {
        layout[zz[i]] = 1;
        layout[zz[i] + ll[i]] = 1;
    }
    auto find = [&](double z0, double z1, int t) {   /* findElement :201-212 */
        double zmid = 0.5 * (z0 + z1);
        for (int i = 0; i < nele; i++) {
            double a = zz[i], b = a + ll[i];
            if ((zmid > a) && (zmid < b) && (type[i] == t)) return i;
        }
        return -1;
    };
    vector<orc_step> v;
    auto it = layout.begin();
    auto iend = layout.end();
    iend--;
    while (it != iend) {                            /* unrollLattice :366-421 */
        double z0 = it->first;
        it++;
        double z1 = it->first;
        double dz = z1 - z0;
        int iele = find(z0, z1, 0);
        orc_step s;
        memset(&s, 0, sizeof(s));
        if (iele == -1) {
            s.dz = z1 - z0;
            v.push_back(s);
        } else {
            const double *p = par + 8 * iele;
            int nz = (int)round(dz / delz);
            if (nz == 0) nz = 1;
            dz = dz / (double)nz;
            for (int iz = 0; iz < nz; iz++) {
                orc_step t;
                memset(&t, 0, sizeof(t));
                t.dz = dz;
                t.aw = p[0];
                double ku = 4. * asin(1) / p[1];
                t.ku = ku;
                t.kx = ku * ku * p[3];
                t.ky = ku * ku * p[4];
                t.ax = p[5];
                t.ay = p[6];
                t.helical = (int)p[7];
                v.push_back(t);
            }
        }
    }
    int nz = (int)v.size();
    if (nz + 1 > max_steps) return -1;
    v[0].z = 0;                                      /* :426-431 */
    for (int i = 1; i < nz; i++) v[i].z = v[i - 1].z + v[i - 1].dz;
    vector<double> delay(nz, 0), lb(nz, 0), ld(nz, 0), lt(nz, 0), ps(nz, 0);
    int lastChicane = -1;
    for (int i = 0; i < nz; i++) {                  /* :462-514 */
        double z0 = v[i].z, z1 = z0 + v[i].dz;
        bool inUnd = (v[i].aw > 0), inQuad = false;
        int e = find(z0, z1, 2);
        if (e != -1) {
            v[i].qf = par[8 * e + 0];
            v[i].qx = par[8 * e + 1];
            v[i].qy = par[8 * e + 2];
            inQuad = true;
        }
        e = find(z0, z1, 4);
        if (e != -1 && e != lastChicane && !inUnd && !inQuad) {
            delay[i] = par[8 * e + 0];
            lb[i] = par[8 * e + 1];
            ld[i] = par[8 * e + 2];
            lt[i] = ll[e];
            lastChicane = e;
        }
        e = find(z0, z1, 3);
        if (e != -1) {
            v[i].cx += par[8 * e + 0];
            v[i].cy += par[8 * e + 1];
        }
        e = find(z0, z1, 6);
        if (e != -1) ps[i] = par[8 * e + 0];
        for (int k = 0; k < nele; k++) {            /* findMarker :214-224 */
            double d = zz[k] - z0;
            if (d * d < 1e-6 && type[k] == 5) {
                v[i].marker = (int)par[8 * k];
                break;
            }
        }
    }
    /* calcSlippage :153-197 */
    double Lz = 0, tmp;
    for (int i = 0; i < nz; i++) {
        if (v[i].aw > 0) {
            tmp = 2 * gamma * gamma * lambda / (1 + v[i].aw * v[i].aw);
            v[i].slip = v[i].dz / tmp;
            v[i].phaseshift = 0;
            if ((Lz > 0) && (i > 0)) {
                tmp = Lz / 2 / gamma / gamma / lambda;
                v[i - 1].slip += floor(tmp) + 1;
                Lz = 0;
            }
        } else {
            Lz += v[i].dz;
            tmp = delay[i] / lambda;
            v[i].slip = floor(tmp);
            v[i].phaseshift = -ps[i];
        }
    }
    tmp = Lz / 2 / gamma / gamma / lambda;
    v[nz - 1].slip += floor(tmp) + 1;
    /* chicane delay -> bend angle by bisection, generateLattice :110-128 */
    for (int i = 0; i < nz; i++) {
        v[i].chic_angle = delay[i];
        v[i].chic_lb = lb[i];
        v[i].chic_ld = ld[i];
        v[i].chic_lt = lt[i];
        if (v[i].chic_angle != 0) {
            double dl = fabs(v[i].chic_angle), tmin = 0, tmax = asin(1) - 0.001;
            bool conv = false;
            double th = 0.5 * (tmax + tmin), d;
            while (!conv) {
                th = 0.5 * (tmax + tmin);
                d = 4 * v[i].chic_lb * (th / sin(th) - 1) + 2 * v[i].chic_ld * (1 / cos(th) - 1);
                if (d > dl) tmax = th; else tmin = th;
                if (fabs(dl - d) < 1e-15) conv = true;
            }
            v[i].chic_angle = th;
        }
    }
    for (int i = 0; i < nz; i++) out[i] = v[i];
    return nz;
}

/* ======================================================================
 * Transverse optics, src/Core/TrackBeam.cpp:8-116 (+ corrector :118-128,
 * chicane :131-258).
 * ==================================================================== */
static void mat44(double m[][4], double e[][4]) {   /* TrackBeam.cpp:241-258 */
    double t[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            t[i][j] = 0;
            for (int k = 0; k < 4; k++) t[i][j] += e[i][k] * m[k][j];
        }
    memcpy(m, t, sizeof(t));
}
static void chicane_matrix(double angle, double lb, double ld, double lt, double m[4][4]) {
    double d1[4][4], d2[4][4], d3[4][4], bp[4][4], bn[4][4], ep[4][4], en[4][4];
    double *all[8] = {&m[0][0], &d1[0][0], &d2[0][0], &d3[0][0], &bp[0][0], &bn[0][0], &ep[0][0], &en[0][0]};
    for (int a = 0; a < 8; a++)
        for (int i = 0; i < 16; i++) all[a][i] = (i % 5 == 0) ? 1 : 0;
    d1[0][1] = ld / cos(angle);
    d1[2][3] = ld / cos(angle);
    d2[0][1] = lt - 4 * lb - 2 * ld;
    d2[2][3] = lt - 4 * lb - 2 * ld;
    d3[0][1] = -lt;
    d3[2][3] = -lt;
    double R = lb / sin(angle), Lpath = R * angle;
    bp[2][3] = Lpath;
    bp[0][0] = cos(angle);
    bp[0][1] = R * sin(angle);
    bp[1][0] = -sin(angle) / R;
    bp[1][1] = cos(angle);
    bn[2][3] = Lpath;
    bn[0][0] = cos(-angle);
    bn[0][1] = R * sin(-angle) * -1;
    bn[1][0] = -sin(-angle) / R * -1;
    bn[1][1] = cos(-angle);
    double efoc = tan(angle) / R;
    ep[1][0] = efoc;
    ep[3][2] = -efoc;
    en[1][0] = -efoc * -1;
    en[3][2] = efoc * -1;
    mat44(m, bp); mat44(m, ep); mat44(m, d1); mat44(m, en); mat44(m, bn); mat44(m, d2);
    mat44(m, bn); mat44(m, en); mat44(m, d1); mat44(m, ep); mat44(m, bp); mat44(m, d3);
}

static inline void map_line(int kind, double delz, double q, double *x, double *px,
                            double gammaz, double d) {
    if (kind == 0) {                                   /* applyDrift :84-88 */
        *x += (*px) * delz / gammaz;
    } else if (kind > 0) {                             /* applyFQuad :91-102 */
        double foc = sqrt(q / gammaz), omg = foc * delz;
        double a1 = cos(omg), a2 = sin(omg) / foc, a3 = -a2 * foc * foc;
        double xtmp = *x - d;
        *x = a1 * xtmp + a2 * (*px) / gammaz + d;
        *px = a3 * xtmp * gammaz + a1 * (*px);
    } else {                                           /* applyDQuad :105-116 */
        double foc = sqrt(-q / gammaz), omg = foc * delz;
        double a1 = cosh(omg), a2 = sinh(omg) / foc, a3 = a2 * foc * foc;
        double xtmp = *x - d;
        *x = a1 * xtmp + a2 * (*px) / gammaz + d;
        *px = a3 * xtmp * gammaz + a1 * (*px);
    }
}

extern "C" void orc_track_transverse(orc_state *s, const orc_step *st, double gamma0,
                                     double delz, int last) {
    bool inund = st->aw > 0;
    double aw = inund ? st->aw : 0, dax = inund ? st->ax : 0, day = inund ? st->ay : 0;
    double kx = inund ? st->kx : 0, ky = inund ? st->ky : 0;
    double qf = st->qf, dqx = (qf == 0) ? 0 : st->qx, dqy = (qf == 0) ? 0 : st->qy;
    double cx = st->cx, cy = st->cy;
    double angle = st->chic_angle;
    double betpar0 = sqrt(1 - (1 + aw * aw) / gamma0 / gamma0);
    double qquad = qf * gamma0;
    double qnatx = kx * aw * aw / gamma0 / betpar0;
    double qnaty = ky * aw * aw / gamma0 / betpar0;
    double qx = qquad + qnatx, qy = -qquad + qnaty;
    double xoff = qquad * dqx + qnatx * dax;
    double yoff = -qquad * dqy + qnaty * day;
    long long n0 = s->off[0], n1 = s->off[s->nslice];
    if (last) {
        if ((cx != 0) || (cy != 0)) {
            double ccx = cx * gamma0, ccy = cy * gamma0;
            for (long long i = n0; i < n1; i++) {
                s->px[i] += ccx;
                s->py[i] += ccy;
            }
        }
    } else if (angle != 0) {
        double m[4][4];
        chicane_matrix(angle, st->chic_angle == 0 ? 0 : st->chic_lb,
                       st->chic_ld, st->chic_lt, m);
        for (long long i = n0; i < n1; i++) {
            double gammaz = sqrt(s->gamma[i] * s->gamma[i] - 1 - s->px[i] * s->px[i] - s->py[i] * s->py[i]);
            double tmp = s->x[i];
            s->x[i] = m[0][0] * tmp + m[0][1] * s->px[i] / gammaz;
            s->px[i] = m[1][0] * tmp * gammaz + m[1][1] * s->px[i];
            tmp = s->y[i];
            s->y[i] = m[2][2] * tmp + m[2][3] * s->py[i] / gammaz;
            s->py[i] = m[3][2] * tmp * gammaz + m[3][3] * s->py[i];
        }
    }
    int kxk = 0, kyk = 0;
    if (qx != 0) { xoff = xoff / qx; kxk = (qx > 0) ? 1 : -1; }
    if (qy != 0) { yoff = yoff / qy; kyk = (qy > 0) ? 1 : -1; }
    for (long long i = n0; i < n1; i++) {
        double gammaz = sqrt(s->gamma[i] * s->gamma[i] - 1 - aw * aw - s->px[i] * s->px[i] - s->py[i] * s->py[i]);
        map_line(kxk, delz, qx, &s->x[i], &s->px[i], gammaz, xoff);
        map_line(kyk, delz, qy, &s->y[i], &s->py[i], gammaz, yoff);
    }
}

/* R56 phase shift of a chicane, TrackBeam.cpp:260-277 */
static void apply_r56(orc_state *s, const orc_step *st, double gamma0) {
    double angle = st->chic_angle;
    if (angle == 0) return;
    double lb = st->chic_lb, ld = st->chic_ld;
    double R56 = (4 * lb / sin(angle) * (1 - angle / tan(angle)) + 2 * ld * tan(angle) / cos(angle)) * angle;
    R56 = R56 * 4 * asin(1) / s->reflength / gamma0;
    for (long long i = s->off[0]; i < s->off[s->nslice]; i++)
        s->theta[i] += R56 * (s->gamma[i] - gamma0);
}

/* ======================================================================
 * Longitudinal space charge, src/Core/EFieldSolver.cpp
 * ==================================================================== */

/* Beam::getSize, src/Core/Beam.cpp:42-67 */
static double slice_size(const orc_state *s, int is) {
    if (s->current[is] <= 0) return 1;
    double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
    int count = 0;
    for (long long i = s->off[is]; i < s->off[is + 1]; i++) {
        x1 += s->x[i];
        x2 += s->x[i] * s->x[i];
        y1 += s->y[i];
        y2 += s->y[i] * s->y[i];
        count++;
    }
    if (count == 0) return 1.;
    x1 /= (double)count;
    x2 /= (double)count;
    y1 /= (double)count;
    y2 /= (double)count;
    return sqrt(std::abs(x2 - x1 * x1)) * sqrt(std::abs(y2 - y1 * y1));
}

/* EFieldSolver::longRange, EFieldSolver.cpp:56-117 (one rank). */
extern "C" void orc_lsc_long(orc_state *s, double gamma0, double aw) {
    int ns = s->nslice;
    for (int i = 0; i < ns; i++) s->longESC[i] = 0;
    if (!s->longrange) return;
    double gamma = gamma0 / sqrt(1 + aw * aw);
    vector<double> fc(ns), fs(ns);
    for (int i = 0; i < ns; i++) {
        fc[i] = s->current[i];
        fs[i] = slice_size(s, i);
        if (fs[i] <= 0) fs[i] = 1;
    }
    double sl = s->slicelength;
    double scl = sl / 2. / asin(1) / 299792458.0 / 2 / 8.85e-12;
#pragma omp parallel for num_threads(s->ndiag_threads) schedule(static)
    for (int i = 0; i < ns; i++) {
        double EFld = 0;
        for (int j = 0; j < ns; j++) {
            double ds = (double)(j - i) * sl * gamma;
            double coef = 1 - sqrt(ds * ds / (ds * ds + fs[j]));
            double sgn = (j > i) ? -1. : ((j < i) ? 1. : 0.);
            EFld += sgn * coef * scl * fc[j] / fs[j];
        }
        s->longESC[i] = EFld;
    }
}

/* Short-range solver for one slice: EFieldSolver::shortRange :197-258
 * (with analyseBeam :122-174 split off so the rmax growth stays in slice
 * order), constructLaplaceOperator :176-190, tridiag :260-272. */
struct ScSlice {
    double xcen, ycen, rbound;
};
static ScSlice sc_analyse(const orc_state *s, int is) {
    ScSlice a{0, 0, 0};
    long long n0 = s->off[is], n = s->off[is + 1] - n0;
    if (n < 1) return a;
    for (long long i = 0; i < n; i++) {
        a.xcen += s->x[n0 + i];
        a.ycen += s->y[n0 + i];
    }
    a.xcen /= (double)n;
    a.ycen /= (double)n;
    double rb = 0;
    for (long long i = 0; i < n; i++) {
        double tx = s->x[n0 + i] - a.xcen, ty = s->y[n0 + i] - a.ycen;
        double r2 = tx * tx + ty * ty;
        if (r2 > rb) rb = r2;
    }
    a.rbound = sqrt(rb);
    return a;
}
static void sc_slice(orc_state *s, int is, const ScSlice &a, double rmax, double gz2,
                     double *ez) {
    long long n0 = s->off[is];
    int npart = (int)(s->off[is + 1] - n0);
    for (int i = 0; i < npart; i++) ez[i] = 0;
    s->sc_efield[is] = 0;
    if (!((s->sc_nz > 0) & (s->sc_ngrid > 2))) return;
    if (npart < 1) return;
    int ng = s->sc_ngrid;
    double dr = rmax / (double)(ng - 1);
    vector<complex<double>> cwork(npart), csrc(ng), clow(ng), cmid(ng), cupp(ng), celm(ng), gam(ng);
    vector<int> idxr(npart);
    for (int i = 0; i < npart; i++) {
        double tx = s->x[n0 + i] - a.xcen, ty = s->y[n0 + i] - a.ycen;
        double radi = sqrt(tx * tx + ty * ty);
        cwork[i] = complex<double>(tx / radi, ty / radi);
        idxr[i] = (int)floor(radi / dr);
    }
    double pi = 2 * asin(1);
    vector<double> vol(ng), ldig(ng + 1), rlog(ng), lmid(ng);
    vol[0] = pi * dr * dr;
    rlog[0] = 0.5;
    ldig[0] = 0;
    for (int j = 1; j < ng; j++) {
        vol[j] = pi * dr * dr * (2 * j + 1);
        ldig[j] = 2 * pi * j;
        rlog[j] = log((double)(j + 1) / (double)j);
    }
    ldig[ng] = 0;
    double coef = -gz2 / s->sc_ks / s->sc_ks;
    double econst = k_vacimp / k_eev * s->current[is] / s->sc_ks / (double)npart;
    const double *th = s->theta + n0;
    for (int m = -s->sc_nphi; m <= s->sc_nphi; m++) {
        for (int i = 0; i < ng; i++) lmid[i] = -ldig[i] - ldig[i + 1] - 2. * pi * m * m * rlog[i];
        lmid[ng - 1] -= 2 * pi * ng;
        for (int l = 1; l <= s->sc_nz; l++) {
            for (int i = 0; i < ng; i++) csrc[i] = complex<double>(0, 0);
            for (int i = 0; i < npart; i++)
                csrc[idxr[i]] += std::pow(cwork[i], -m) * complex<double>(cos(l * th[i]), -sin(l * th[i]));
            for (int i = 0; i < ng; i++) {
                csrc[i] *= complex<double>(0, econst / l / vol[i]);
                clow[i] = complex<double>(coef * ldig[i] / l / l / vol[i], 0);
                cmid[i] = complex<double>(1 + coef * lmid[i] / l / l / vol[i], 0);
                cupp[i] = complex<double>(coef * ldig[i + 1] / l / l / vol[i], 0);
            }
            complex<double> bet = cmid[0];
            celm[0] = csrc[0] / bet;
            for (int i = 1; i < ng; i++) {
                gam[i] = cupp[i - 1] / bet;
                bet = cmid[i] - clow[i] * gam[i];
                celm[i] = (csrc[i] - clow[i] * celm[i - 1]) / bet;
            }
            for (int i = ng - 2; i > -1; i--) celm[i] = celm[i] - gam[i + 1] * celm[i + 1];
            if (l == 1) s->sc_efield[is] = std::abs(celm[0]);
            for (int i = 0; i < npart; i++) {
                complex<double> ct = std::pow(cwork[i], m) * complex<double>(cos(l * th[i]), sin(l * th[i]));
                ez[i] += 2 * std::real(ct * celm[idxr[i]]);
            }
        }
    }
}

/* ======================================================================
 * Longitudinal push, src/Core/BeamSolver.cpp:13-163 (fundamental only).
 * ==================================================================== */
struct PushCtx {
    double xks, xku, btpar, ez;
    complex<double> rpart;
    int nh;
    complex<double> rph[8];
    double hh[8];
    double k2gg, k2pp, k3gg, k3pp;
};
static inline void ode(PushCtx &c, double tgam, double tthet) {   /* :144-163 */
    double ztemp1 = -2. / c.xks;
    complex<double> ctmp = 0;
    ctmp += c.rpart * complex<double>(cos(1.0 * tthet), -sin(1.0 * tthet));
    for (int k = 0; k < c.nh; k++) ctmp += c.rph[k] * complex<double>(cos(c.hh[k] * tthet), -sin(c.hh[k] * tthet));
    double btper0 = c.btpar + ztemp1 * ctmp.real();
    double btpar0 = sqrt(1. - btper0 / (tgam * tgam));
    c.k2pp += c.xks * (1. - 1. / btpar0) + c.xku;
    c.k2gg += ctmp.imag() / btpar0 / tgam - c.ez;
}
static inline void rk4(PushCtx &c, double delz, double &gamma, double &theta) {  /* :89-141 */
    c.k2gg = 0;
    c.k2pp = 0;
    ode(c, gamma, theta);
    double stpz = 0.5 * delz;
    gamma += stpz * c.k2gg;
    theta += stpz * c.k2pp;
    c.k3gg = c.k2gg;
    c.k3pp = c.k2pp;
    c.k2gg = 0;
    c.k2pp = 0;
    ode(c, gamma, theta);
    gamma += stpz * (c.k2gg - c.k3gg);
    theta += stpz * (c.k2pp - c.k3pp);
    c.k3gg /= 6;
    c.k3pp /= 6;
    c.k2gg *= -0.5;
    c.k2pp *= -0.5;
    ode(c, gamma, theta);
    stpz = delz;
    gamma += stpz * c.k2gg;
    theta += stpz * c.k2pp;
    c.k3gg -= c.k2gg;
    c.k3pp -= c.k2pp;
    c.k2gg *= 2;
    c.k2pp *= 2;
    ode(c, gamma, theta);
    gamma += stpz * (c.k3gg + c.k2gg / 6.0);
    theta += stpz * (c.k3pp + c.k2pp / 6.0);
}

extern "C" void orc_advance(orc_state *s, const orc_step *st, double gamma0, double delz) {
    double xks = s->xks;                         /* fundamental field present */
    double rtmp = orc_fc(st, 1) / s->xks;
    double xku = st->ku;
    if (xku == 0) xku = xks * 0.5 / gamma0 / gamma0;
    double aw = st->aw, autophase = st->phaseshift;
    orc_lsc_long(s, gamma0, aw);
    double gammaz2 = gamma0 * gamma0 / (1 + aw * aw);
    int ns = s->nslice, ng = s->ngrid;
    /* analyseBeam and the rmax growth in slice order (EFieldSolver.cpp:159-163) */
    bool hassc = (s->sc_nz > 0) & (s->sc_ngrid > 2);
    vector<ScSlice> an(ns);
    vector<double> rmaxs(ns, s->sc_rmax);
    if (hassc) {
        for (int is = 0; is < ns; is++) {
            if (s->off[is + 1] - s->off[is] < 1) { rmaxs[is] = s->sc_rmax; continue; }
            an[is] = sc_analyse(s, is);
            if (an[is].rbound > s->sc_rmax) s->sc_rmax = an[is].rbound * 1.5;
            rmaxs[is] = s->sc_rmax;
        }
    }
    long long maxn = 0;
    for (int is = 0; is < ns; is++) maxn = std::max(maxn, s->off[is + 1] - s->off[is]);
#pragma omp parallel num_threads(s->ndiag_threads)
    {
        vector<double> ezbuf(maxn > 0 ? maxn : 1);
#pragma omp for schedule(dynamic, 1)
        for (int is = 0; is < ns; is++) {
            double eloss = -s->longESC[is] / k_eev;
            sc_slice(s, is, an[is], rmaxs[is], gammaz2, ezbuf.data());
            int islice = (is + s->first) % ns;
            const double *f = s->field + 2LL * islice * ng * ng;
            PushCtx c;
            c.xks = xks;
            c.xku = xku;
            c.nh = s->nh < 8 ? s->nh : 8;
            for (long long ip = s->off[is]; ip < s->off[is + 1]; ip++) {
                double gamma = s->gamma[ip];
                double theta = s->theta[ip] + autophase;
                double x = s->x[ip], y = s->y[ip], px = s->px[ip], py = s->py[ip];
                double awloc = orc_faw(st, x, y);
                c.btpar = 1 + px * px + py * py + aw * aw * awloc * awloc;
                c.ez = ezbuf[ip - s->off[is]] + eloss;
                /* Field::getLLGridpoint, Field.cpp:160-174 */
                if ((x > -s->gridmax) && (x < s->gridmax) && (y > -s->gridmax) && (y < s->gridmax)) {
                    double wx = (x + s->gridmax) / s->dgrid;
                    double wy = (y + s->gridmax) / s->dgrid;
                    int ix = (int)floor(wx), iy = (int)floor(wy);
                    wx = 1 + floor(wx) - wx;
                    wy = 1 + floor(wy) - wy;
                    int idx = ix + iy * ng;
                    complex<double> cpart;
                    complex<double> f0(f[2 * idx], f[2 * idx + 1]);
                    cpart = f0 * wx * wy;
                    idx++;
                    complex<double> f1(f[2 * idx], f[2 * idx + 1]);
                    cpart += f1 * (1 - wx) * wy;
                    idx += ng - 1;
                    complex<double> f2(f[2 * idx], f[2 * idx + 1]);
                    cpart += f2 * wx * (1 - wy);
                    idx++;
                    complex<double> f3(f[2 * idx], f[2 * idx + 1]);
                    cpart += f3 * (1 - wx) * (1 - wy);
                    c.rpart = rtmp * awloc * std::conj(cpart);
                } else {
                    c.rpart = 0;
                }
                for (int k = 0; k < c.nh; k++) {  /* the harmonic fields, :63-79 */
                    const orc_hfield &h = s->hf[k];
                    c.hh[k] = (double)h.harm;
                    c.rph[k] = 0;
                    if ((x > -h.gridmax) && (x < h.gridmax) && (y > -h.gridmax) && (y < h.gridmax)) {
                        double wx = (x + h.gridmax) / h.dgrid;
                        double wy = (y + h.gridmax) / h.dgrid;
                        int ix = (int)floor(wx), iy = (int)floor(wy);
                        wx = 1 + floor(wx) - wx;
                        wy = 1 + floor(wy) - wy;
                        int idx = ix + iy * h.ngrid;
                        const double *fh = h.field + 2LL * islice * h.ngrid * h.ngrid;
                        complex<double> cp = complex<double>(fh[2 * idx], fh[2 * idx + 1]) * wx * wy;
                        idx++;
                        cp += complex<double>(fh[2 * idx], fh[2 * idx + 1]) * (1 - wx) * wy;
                        idx += h.ngrid - 1;
                        cp += complex<double>(fh[2 * idx], fh[2 * idx + 1]) * wx * (1 - wy);
                        idx++;
                        cp += complex<double>(fh[2 * idx], fh[2 * idx + 1]) * (1 - wx) * (1 - wy);
                        c.rph[k] = (orc_fc(st, h.harm) / h.xks) * awloc * std::conj(cp);
                    }
                }
                rk4(c, delz, gamma, theta);
                s->gamma[ip] = gamma;
                s->theta[ip] = theta;
            }
        }
    }
}

/* ======================================================================
 * Field: source deposition + ADI, src/Core/FieldSolverADI.cpp
 * ==================================================================== */

/* FieldSolverADI::init, :125-182.  c/cbet/cwet as interleaved complex
 * arrays [ngrid][2]; cstep is (0, rtmp). */
extern "C" void orc_adi_factors(int ngrid, double delz, double dgrid, double xks, double *c,
                                double *cbet, double *cwet, double *cstep_im) {
    double rtmp = 0.25 * delz / (xks * dgrid * dgrid);
    *cstep_im = rtmp;
    vector<double> mupp(ngrid), mmid(ngrid), mlow(ngrid);
    vector<complex<double>> w1(ngrid), w2(ngrid), cc(ngrid), cb(ngrid), cw(ngrid);
    mupp[0] = rtmp;
    mmid[0] = -2 * rtmp;
    mlow[0] = 0;
    for (int i = 1; i < ngrid - 1; i++) {
        mupp[i] = rtmp;
        mmid[i] = -2 * rtmp;
        mlow[i] = rtmp;
    }
    mupp[ngrid - 1] = 0;
    mmid[ngrid - 1] = -2 * rtmp;
    mlow[ngrid - 1] = rtmp;
    for (int i = 0; i < ngrid; i++) {
        w1[i] = complex<double>(0, -mupp[i]);
        w2[i] = complex<double>(1, -mmid[i]);
        cc[i] = complex<double>(0, -mlow[i]);
    }
    cb[0] = 1. / w2[0];
    cw[0] = 0.;
    for (int i = 1; i < ngrid; i++) {
        cw[i] = w1[i - 1] * cb[i - 1];
        cb[i] = 1. / (w2[i] - cc[i] * cw[i]);
    }
    for (int i = 0; i < ngrid; i++) {
        c[2 * i] = cc[i].real();
        c[2 * i + 1] = cc[i].imag();
        cbet[2 * i] = cb[i].real();
        cbet[2 * i + 1] = cb[i].imag();
        cwet[2 * i] = cw[i].real();
        cwet[2 * i + 1] = cw[i].imag();
    }
}

/* Source term of beam slice ii, FieldSolverADI.cpp:13-51 (harm=1).
 * src holds ngrid^2 complex values and is cleared here. */
/* the source of beam slice ii for the field at harmonic `harm` on a grid
 * (ng, gridmax, dgrid) with wavenumber xks (FieldSolverADI.cpp:13-51) */
static void deposit_h(const orc_state *s, const orc_step *st, int ii, double delz, int harm, int ng,
                      double gridmax, double dgrid, double xks, double *src);

extern "C" void orc_deposit_slice(const orc_state *s, const orc_step *st, int ii, double delz,
                                  double *src) {
    deposit_h(s, st, ii, delz, 1, s->ngrid, s->gridmax, s->dgrid, s->xks, src);
}

static void deposit_h(const orc_state *s, const orc_step *st, int ii, double delz, int harm, int ng,
                      double gridmax, double dgrid, double xks, double *src) {
    for (int ig = 0; ig < 2 * ng * ng; ig++) src[ig] = 0;
    complex<double> *cs = reinterpret_cast<complex<double> *>(src);
    if (!(st->aw > 0) || (harm % 2 != 1)) return;  /* inUndulator, odd harmonic */
    long long n0 = s->off[ii], n = s->off[ii + 1] - n0;
    double scl = orc_fc(st, harm) * k_vacimp * s->current[ii] * xks * delz;
    scl /= 4 * k_eev * (double)n * dgrid * dgrid;
    for (long long k = 0; k < n; k++) {
        double x = s->x[n0 + k], y = s->y[n0 + k];
        double theta = (double)harm * s->theta[n0 + k];
        double gamma = s->gamma[n0 + k];
        if ((x > -gridmax) && (x < gridmax) && (y > -gridmax) && (y < gridmax)) {
            double wx = (x + gridmax) / dgrid;
            double wy = (y + gridmax) / dgrid;
            int ix = (int)floor(wx), iy = (int)floor(wy);
            wx = 1 + floor(wx) - wx;
            wy = 1 + floor(wy) - wy;
            int idx = ix + iy * ng;
            double part = sqrt(orc_faw2(st, x, y)) * scl / gamma;
            complex<double> cpart = complex<double>(sin(theta), cos(theta)) * part;
            double weight = wx * wy;
            cs[idx] += weight * cpart;
            weight = (1 - wx) * wy;
            idx++;
            cs[idx] += weight * cpart;
            weight = wx * (1 - wy);
            idx += ng - 1;
            cs[idx] += weight * cpart;
            weight = (1 - wx) * (1 - wy);
            idx++;
            cs[idx] += weight * cpart;
        }
    }
}

/* FieldSolverADI::ADI + tridagx + tridagy, :60-122.  u is one field slice
 * (ngrid^2 complex, x fastest), r a work array of the same size. */
extern "C" void orc_adi_slice(int ng, const double *cd, const double *cbd, const double *cwd,
                              double cstep_im, const double *srcd, double *ud, double *rd) {
    typedef complex<double> C;
    const C *c = reinterpret_cast<const C *>(cd), *cbet = reinterpret_cast<const C *>(cbd),
            *cwet = reinterpret_cast<const C *>(cwd), *S = reinterpret_cast<const C *>(srcd);
    C *u = reinterpret_cast<C *>(ud), *r = reinterpret_cast<C *>(rd);
    C cstep(0, cstep_im);
    int idx;
    for (idx = 0; idx < ng; idx++) r[idx] = S[idx] + u[idx] + cstep * (u[idx + ng] - 2.0 * u[idx]);
    for (idx = ng; idx < ng * (ng - 1); idx++)
        r[idx] = S[idx] + u[idx] + cstep * (u[idx + ng] - 2.0 * u[idx] + u[idx - ng]);
    for (idx = ng * (ng - 1); idx < ng * ng; idx++)
        r[idx] = S[idx] + u[idx] + cstep * (u[idx - ng] - 2.0 * u[idx]);
    for (int i = 0; i < ng * ng; i += ng) {                 /* tridagx */
        u[i] = r[i] * cbet[0];
        for (int k = 1; k < ng; k++) u[k + i] = (r[k + i] - c[k] * u[k + i - 1]) * cbet[k];
        for (int k = ng - 2; k >= 0; k--) u[k + i] -= cwet[k + 1] * u[k + i + 1];
    }
    for (int ix = 0; ix < ng * ng; ix += ng) {
        idx = ix;
        r[idx] = S[idx] + u[idx] + cstep * (u[idx + 1] - 2.0 * u[idx]);
        for (idx = ix + 1; idx < ix + ng - 1; idx++)
            r[idx] = S[idx] + u[idx] + cstep * (u[idx + 1] - 2.0 * u[idx] + u[idx - 1]);
        idx = ix + ng - 1;
        r[idx] = S[idx] + u[idx] + cstep * (u[idx - 1] - 2.0 * u[idx]);
    }
    for (int i = 0; i < ng; i++) u[i] = r[i] * cbet[0];      /* tridagy */
    for (int k = 1; k < ng; k++) {
        int n = k * ng;
        for (int i = 0; i < ng; i++) u[n + i] = (r[n + i] - c[k] * u[n + i - ng]) * cbet[k];
    }
    for (int k = ng - 2; k >= 0; k--) {
        int n = k * ng;
        for (int i = 0; i < ng; i++) u[n + i] -= cwet[k + 1] * u[n + i + ng];
    }
}

/* 2-D discrete Fourier transform of one ngrid x ngrid slice in place, the
 * transform FFTW's fftw_plan_dft_2d computes (FieldSolverFFT.cpp:155-156):
 * X[ky][kx] = sum_{y,x} x[y][x] exp(sign 2 pi i (ky y + kx x)/n), unscaled,
 * sign -1 forward and +1 backward.  Evaluated as row transforms followed by
 * column transforms, each a direct sum with the exponent reduced to
 * (j k) mod n exactly, so every twiddle is one of n table values. */
extern "C" void orc_dft2(int n, double *data, int sign) {
    typedef complex<double> C;
    C *d = reinterpret_cast<C *>(data);
    vector<C> tw(n), line(n), out(n);
    const double twopi = 4. * asin(1.);
    for (int m = 0; m < n; m++) tw[m] = C(cos(twopi * m / n), sign * sin(twopi * m / n));
    for (int pass = 0; pass < 2; pass++) {
        const long long step = pass == 0 ? 1 : n, stride = pass == 0 ? n : 1;
        for (int l = 0; l < n; l++) {
            for (int j = 0; j < n; j++) line[j] = d[l * stride + j * step];
            for (int k = 0; k < n; k++) {
                C acc(0, 0);
                for (int j = 0; j < n; j++) acc += line[j] * tw[(long long)j * k % n];
                out[k] = acc;
            }
            for (int k = 0; k < n; k++) d[l * stride + k * step] = out[k];
        }
    }
}

/* FieldSolverFFT::init tables (FieldSolverFFT.cpp:127-152): K2 (imaginary
 * part) and the source filter's sigmoid, both at FFT index
 * ((i + (n+1)/2) mod n) of grid point i. */
extern "C" void orc_fft_tables(int ng, double dgrid, double xks, double xc, double yc, double sig,
                               double *k2im, double *sigmoid) {
    const double dk = 4. * asin(1.) / ((double)ng * dgrid);
    const double shift = -0.5 * (double)(ng - 1);
    for (int iy = 0; iy < ng; iy++) {
        double dy = (double)iy + shift;
        double y = dy / (double)ng / yc;
        for (int ix = 0; ix < ng; ix++) {
            double dx = (double)ix + shift;
            double x = dx / (double)ng / xc;
            int iiy = (iy + (ng + 1) / 2) % ng;
            int iix = (ix + (ng + 1) / 2) % ng;
            int ii = iiy * ng + iix;
            k2im[ii] = -(dx * dx + dy * dy) * dk * dk / 2. / xks;
            double r = (sqrt(x * x + y * y) - 1) / sig;
            sigmoid[ii] = 1. / (1 + exp(r));
        }
    }
}

/* FieldSolverFFT::FFT on one slice (FieldSolverFFT.cpp:62-119): propagate
 * the field by exp(K2 delz) in Fourier space and add twice the source,
 * in real space, or filtered by the sigmoid in Fourier space. */
extern "C" void orc_fft_slice(int ng, const double *k2im, const double *sigmoid, int filter, double delz,
                      const double *srcd, double *ud, double *work) {
    typedef complex<double> C;
    C *u = reinterpret_cast<C *>(ud), *w = reinterpret_cast<C *>(work);
    const C *S = reinterpret_cast<const C *>(srcd);
    const int nn = ng * ng;
    orc_dft2(ng, ud, -1);
    if (filter) {
        for (int i = 0; i < nn; i++) w[i] = S[i];
        orc_dft2(ng, work, -1);
        for (int i = 0; i < nn; i++) w[i] = w[i] * sigmoid[i];
        for (int i = 0; i < nn; i++) u[i] = u[i] * exp(C(0, k2im[i]) * delz) + 2. * w[i];
    } else {
        for (int i = 0; i < nn; i++) u[i] = u[i] * exp(C(0, k2im[i]) * delz);
    }
    orc_dft2(ng, ud, +1);
    const double norm = 1. / (double)nn;
    if (filter)
        for (int i = 0; i < nn; i++) u[i] = u[i] * norm;
    else
        for (int i = 0; i < nn; i++) u[i] = u[i] * norm + 2. * S[i];
}

/* Field::track -> FieldSolverFFT::init/advance (FieldSolverFFT.cpp:9-60):
 * the same source as the ADI solver, then fft_slice on slice
 * (ii + first) mod Ns. */
static void field_advance_fft(orc_state *s, const orc_step *st, double delz) {
    int ng = s->ngrid, ns = s->nslice;
    vector<double> k2(ng * ng), sg(ng * ng);
    const int filter = s->fft_filter && s->fft_sig > 0 && s->fft_xcut > 0 && s->fft_ycut > 0;
    orc_fft_tables(ng, s->dgrid, s->xks, filter ? s->fft_xcut : 1, filter ? s->fft_ycut : 1,
                   filter ? s->fft_sig : 1, k2.data(), sg.data());
#pragma omp parallel num_threads(s->ndiag_threads)
    {
        vector<double> src(2 * ng * ng), w(2 * ng * ng);
#pragma omp for schedule(dynamic, 1)
        for (int ii = 0; ii < ns; ii++) {
            orc_deposit_slice(s, st, ii, delz, src.data());
            int i = (ii + s->first) % ns;
            orc_fft_slice(ng, k2.data(), sg.data(), filter, delz, src.data(), s->field + 2LL * i * ng * ng,
                      w.data());
        }
    }
}

/* Field::track -> FieldSolverADI::init/advance, Field.cpp:177-185. */
static void field_advance_harmonics(orc_state *s, const orc_step *st, double delz);

extern "C" void orc_field_advance(orc_state *s, const orc_step *st, double delz) {
    field_advance_harmonics(s, st, delz);
    if (s->fft_solver) {
        field_advance_fft(s, st, delz);
        return;
    }
    int ng = s->ngrid, ns = s->nslice;
    vector<double> c(2 * ng), cb(2 * ng), cw(2 * ng);
    double cim;
    orc_adi_factors(ng, delz, s->dgrid, s->xks, c.data(), cb.data(), cw.data(), &cim);
#pragma omp parallel num_threads(s->ndiag_threads)
    {
        vector<double> src(2 * ng * ng), r(2 * ng * ng);
#pragma omp for schedule(dynamic, 1)
        for (int ii = 0; ii < ns; ii++) {
            orc_deposit_slice(s, st, ii, delz, src.data());
            int i = (ii + s->first) % ns;
            orc_adi_slice(ng, c.data(), cb.data(), cw.data(), cim, src.data(),
                          s->field + 2LL * i * ng * ng, r.data());
        }
    }
}

/* DiagField's power (Diagnostic.cpp:763-780, 868) of harmonic field k in
 * beam slice order: sum |E|^2 (dgrid eev / ks)^2 / vacimp. */
extern "C" void orc_diag_hpower(const orc_state *s, int k, double *power) {
    const orc_hfield &h = s->hf[k];
    const int ns = s->nslice, ng = h.ngrid;
    const double scl = h.dgrid * k_eev / h.xks;
    for (int is0 = 0; is0 < ns; is0++) {  /* storage slice is0 -> output (ns + is0 - first) % ns */
        int is = (ns + is0 - s->first) % ns;
        const double *f = h.field + 2LL * is0 * ng * ng;
        double p = 0;
        for (int i = 0; i < ng * ng; i++) p += f[2 * i] * f[2 * i] + f[2 * i + 1] * f[2 * i + 1];
        power[is] = p * scl * scl / k_vacimp;
    }
}

/* Field::track of the harmonic fields (ADI solver): each has its own
 * factors (its grid and wavenumber) and source. */
static void field_advance_harmonics(orc_state *s, const orc_step *st, double delz) {
    for (int k = 0; k < s->nh; k++) {
        const orc_hfield &h = s->hf[k];
        int ng = h.ngrid, ns = s->nslice;
        vector<double> c(2 * ng), cb(2 * ng), cw(2 * ng);
        double cim;
        orc_adi_factors(ng, delz, h.dgrid, h.xks, c.data(), cb.data(), cw.data(), &cim);
#pragma omp parallel num_threads(s->ndiag_threads)
        {
            vector<double> src(2 * ng * ng), r(2 * ng * ng);
#pragma omp for schedule(dynamic, 1)
            for (int ii = 0; ii < ns; ii++) {
                deposit_h(s, st, ii, delz, h.harm, ng, h.gridmax, h.dgrid, h.xks, src.data());
                int i = (ii + s->first) % ns;
                orc_adi_slice(ng, c.data(), cb.data(), cw.data(), cim, src.data(), h.field + 2LL * i * ng * ng,
                              r.data());
            }
        }
    }
}

/* Control::applySlippage, src/Core/Control.cpp:158-276, one rank. */
extern "C" void orc_slippage(orc_state *s, double slip) {
    if (!s->timerun) return;
    s->accuslip += slip;
    int ns = s->nslice, ng = s->ngrid;
    int direction = 1;
    while (fabs(s->accuslip) > (s->sample * 0.8)) {
        if (s->accuslip < 0) direction = -1;
        s->accuslip -= s->sample * direction;
        int last = (s->first + ns - 1) % ns;
        if (direction < 0) last = (last + 1) % ns;
        if (!s->periodic) {
            /* one rank: rank 0 is also the last rank */
            if (direction > 0 || direction < 0) {
                double *f = s->field + 2LL * last * ng * ng;
                for (int i = 0; i < 2 * ng * ng; i++) f[i] = 0;
                for (int k = 0; k < s->nh; k++) {  /* every field slips alike */
                    const int nh = s->hf[k].ngrid;
                    double *fh = s->hf[k].field + 2LL * last * nh * nh;
                    for (int i = 0; i < 2 * nh * nh; i++) fh[i] = 0;
                }
            }
        }
        s->first = last;
        if (direction < 0) s->first = (last + 1) % ns;
    }
}

struct Incoh {
    Ran2 *r;
    int loss, spread;
};

extern "C" void *orc_incoherent_new(int seed, int rank, int do_loss, int do_spread) {
    Ran2 rseed((unsigned int)seed);
    double val = 0;
    for (int i = 0; i <= rank; i++) val = rseed.next();
    val *= 1e9;
    int locseed = (int)round(val);
    Incoh *h = new Incoh;
    h->r = new Ran2((unsigned int)locseed);
    h->loss = do_loss;
    h->spread = do_spread;
    return h;
}

extern "C" void orc_incoherent_free(void *hv) {
    Incoh *h = (Incoh *)hv;
    if (!h) return;
    delete h->r;
    delete h;
}

/* Incoherent::apply, src/Core/Incoherent.cpp:35-82: one uniform draw per
 * beamlet (every nbins-th particle), in slice and particle order. */
extern "C" void orc_incoherent_apply(orc_state *s, const orc_step *st, double gamma0, double delz) {
    Incoh *h = (Incoh *)s->incoh;
    if (!h) return;
    if (!(st->aw > 0)) return;
    if ((!h->loss) && (!h->spread)) return;
    double gam0 = gamma0;
    double awz = st->aw;
    double xkw0 = st->ku;
    double dgamsig = 1.015e-27 * xkw0 * xkw0 * awz * awz;
    if (st->helical)
        dgamsig *= 1.42 * awz + 1. / (1. + 1.5 * awz + 0.95 * awz * awz);
    else
        dgamsig *= 1.697 * awz + 1. / (1. + 1.88 * awz + 0.8 * awz * awz);
    if (!h->spread) dgamsig = 0;
    dgamsig = sqrt(dgamsig * gam0 * gam0 * gam0 * gam0 * xkw0 * delz) * sqrt(3.);
    double dgamavg = xkw0 * gam0 * awz;
    if (!h->loss) dgamavg = 0;
    dgamavg = dgamavg * dgamavg * 1.88e-15 * delz;
    int nbins = s->nbins > 0 ? s->nbins : 1;
    double dg = 0;
    for (int is = 0; is < s->nslice; is++) {
        long long n0 = s->off[is], npart = s->off[is + 1] - n0;
        for (long long ip = 0; ip < npart; ip++) {
            if ((ip % nbins) == 0) dg = -dgamavg + dgamsig * (2 * h->r->next() - 1);
            s->gamma[n0 + ip] += dg;
        }
    }
}

/* Collective::apply, src/Core/Collective.cpp:107-128 (the wake potential
 * is already folded into s->wake). */
static void wake_apply(orc_state *s, double delz) {
    if (!s->wake) return;
    for (int ic = 0; ic < s->nslice; ic++) {
        s->eloss[ic] = s->wake[ic];
        double dg = s->eloss[ic] * delz / k_eev;
        for (long long ip = s->off[ic]; ip < s->off[ic + 1]; ip++) s->gamma[ip] += dg;
    }
}

/* Beam::track, src/Core/Beam.cpp:97-112. */
extern "C" void orc_beam_track(orc_state *s, const orc_step *st, double gamma0, double delz) {
    orc_track_transverse(s, st, gamma0, delz * 0.5, 0);
    orc_advance(s, st, gamma0, delz);
    orc_incoherent_apply(s, st, gamma0, delz);
    wake_apply(s, delz);
    apply_r56(s, st, gamma0);
    orc_track_transverse(s, st, gamma0, delz * 0.5, 1);
}

/* One step of Gencore::run (Gencore.cpp:104-147): beam, field, slippage. */
extern "C" void orc_step_all(orc_state *s, const orc_step *st, double gamma0) {
    double delz = st->dz;
    orc_beam_track(s, st, gamma0, delz);
    orc_field_advance(s, st, delz);
    orc_slippage(s, st->slip);
}

/* Beam::sort for one4one beams on one rank (src/Core/Beam.cpp:81-94;
 * Sorting::globalSort/fillPushVectors :190-262, 367-428 with s0 = 0,
 * keep window [0, slen ns]; localSort :74-145): particles outside the
 * window are dropped, every other particle moves to slice a + floor(theta /
 * slen) with theta -= slen atar, by swap-with-last and push_back in the
 * reference's order.  The beam is given and returned as per-slice arrays:
 * in: off[ns+1] and the six coordinate arrays (capacity cap); out: the same
 * arrays rewritten, off updated, current[ns] = np ce / slicelength.
 * Returns the new particle count, or -1 if cap is too small. */
extern "C" long long orc_sort_one4one(int ns, long long *off, long long cap, double slen, double slicelength,
                                      double *g, double *th, double *x, double *y, double *px, double *py,
                                      double *current) {
    struct P { double th, g, x, y, px, py; };
    vector<vector<P>> rec(ns);
    for (int a = 0; a < ns; a++)
        for (long long i = off[a]; i < off[a + 1]; i++) rec[a].push_back({th[i], g[i], x[i], y[i], px[i], py[i]});
    /* fillPushVectors' deletion pass (one rank: nothing is sent) */
    double keepmin = 0, keepmax = slen * ns;
    for (int i = 0; i < ns; i++) {
        size_t j = 0;
        while (j < rec[i].size()) {
            double s = slen * i + rec[i][j].th;
            if ((s < keepmin) || (s > keepmax)) {
                rec[i][j] = rec[i].back();
                rec[i].pop_back();
            } else {
                j++;
            }
        }
    }
    /* localSort */
    double invslen = 1. / slen;
    for (int a = 0; a < ns; a++) {
        size_t b = 0;
        while (b < rec[a].size()) {
            double theta = rec[a][b].th;
            int atar = (int)floor(theta * invslen);
            if (atar != 0) {
                P p = rec[a][b];
                p.th = rec[a][b].th - slen * (atar);
                /* a particle exactly at the window's upper edge (s = slen ns)
                 * would index one past the last slice, and one at s = 0 whose
                 * floor(theta / slen) rounds below -a one before the first
                 * (the reference's at() throws there); both are dropped */
                if (a + atar >= 0 && a + atar < ns) rec[a + atar].push_back(p);
                rec[a][b] = rec[a].back();
                rec[a].pop_back();
            } else {
                b++;
            }
        }
    }
    long long n = 0;
    for (int a = 0; a < ns; a++) n += (long long)rec[a].size();
    if (n > cap) return -1;
    n = 0;
    for (int a = 0; a < ns; a++) {
        off[a] = n;
        for (const P &p : rec[a]) {
            th[n] = p.th; g[n] = p.g; x[n] = p.x; y[n] = p.y; px[n] = p.px; py[n] = p.py;
            n++;
        }
        current[a] = (double)rec[a].size() * k_ce / slicelength;
    }
    off[ns] = n;
    return n;
}

/* Localsort target index, src/Util/Sorting.cpp:96 (atar=floor(theta/slen)). */
extern "C" int orc_sort_index(const double *theta, int n, double slen, int *atar) {
    double invslen = 1. / slen;
    int moved = 0;
    for (int i = 0; i < n; i++) {
        atar[i] = (int)floor(theta[i] * invslen);
        if (atar[i] != 0) moved++;
    }
    return moved;
}

/* ======================================================================
 * Diagnostics: DiagBeam::getValues (Diagnostic.cpp:373-608) and
 * DiagField::getValues (:718-996) for the fundamental.
 * ==================================================================== */
extern "C" void orc_diag(const orc_state *s, orc_diag_out *d) {
    int ns = s->nslice, ng = s->ngrid;
    for (int is = 0; is < ns; is++) {
        double g1 = 0, g2 = 0, x1 = 0, x2 = 0, y1 = 0, y2 = 0;
        complex<double> b = 0;
        long long n0 = s->off[is], n1 = s->off[is + 1];
        for (long long i = n0; i < n1; i++) {
            x1 += s->x[i];
            x2 += s->x[i] * s->x[i];
            y1 += s->y[i];
            y2 += s->y[i] * s->y[i];
            g1 += s->gamma[i];
            g2 += s->gamma[i] * s->gamma[i];
            b += complex<double>(cos((0 + 1) * s->theta[i]), sin((0 + 1) * s->theta[i]));
        }
        double norm = 1.;
        if (n1 > n0) norm = 1. / (double)(n1 - n0);
        x1 *= norm; x2 *= norm; y1 *= norm; y2 *= norm; g1 *= norm; g2 *= norm;
        b *= norm;
        if (d->energy) d->energy[is] = g1;
        if (d->energyspread) d->energyspread[is] = sqrt(fabs(g2 - g1 * g1));
        if (d->xposition) d->xposition[is] = x1;
        if (d->xsize) d->xsize[is] = sqrt(fabs(x2 - x1 * x1));
        if (d->yposition) d->yposition[is] = y1;
        if (d->ysize) d->ysize[is] = sqrt(fabs(y2 - y1 * y1));
        if (d->bunching) d->bunching[is] = std::abs(b);
        if (d->bunchingphase) d->bunchingphase[is] = atan2(b.imag(), b.real());
    }
    double ks = 4. * asin(1) / s->xlambda;
    double scl = s->dgrid * k_eev / ks;
    double shift = -0.5 * (double)(ng - 1);
    double g_pow = 0;
    for (int is0 = 0; is0 < ns; is0++) {
        int is = (ns + is0 - s->first) % ns;
        const double *f = s->field + 2LL * is0 * ng * ng;
        double power = 0, x1 = 0, x2 = 0, y1 = 0, y2 = 0;
        complex<double> ff(0, 0);
        for (int iy = 0; iy < ng; iy++) {
            double dy = (double)iy + shift;
            for (int ix = 0; ix < ng; ix++) {
                double dx = (double)ix + shift;
                int i = iy * ng + ix;
                complex<double> loc(f[2 * i], f[2 * i + 1]);
                double wei = loc.real() * loc.real() + loc.imag() * loc.imag();
                ff += loc;
                power += wei;
                x1 += dx * wei;
                x2 += dx * dx * wei;
                y1 += dy * wei;
                y2 += dy * dy * wei;
            }
        }
        g_pow += power;
        if (power > 0) { x1 /= power; x2 /= power; y1 /= power; y2 /= power; }
        x2 = sqrt(fabs(x2 - x1 * x1));
        y2 = sqrt(fabs(y2 - y1 * y1));
        int ic = (ng / 2) * ng + (ng / 2);
        double inten = f[2 * ic] * f[2 * ic] + f[2 * ic + 1] * f[2 * ic + 1];
        double farfield = ff.real() * ff.real() + ff.imag() * ff.imag();
        farfield *= s->dgrid * s->dgrid;
        power *= scl * scl / k_vacimp;
        x1 *= s->dgrid; x2 *= s->dgrid; y1 *= s->dgrid; y2 *= s->dgrid;
        inten *= k_eev * k_eev / ks / ks / k_vacimp;
        if (d->power) d->power[is] = power;
        if (d->f_xposition) d->f_xposition[is] = x1;
        if (d->f_xsize) d->f_xsize[is] = x2;
        if (d->f_yposition) d->f_yposition[is] = y1;
        if (d->f_ysize) d->f_ysize[is] = y2;
        if (d->nf_intensity) d->nf_intensity[is] = inten;
        if (d->ff_intensity) d->ff_intensity[is] = farfield;
    }
    if (d->global_energy) {
        g_pow *= scl * scl / k_vacimp * s->slicelength / 299792458.0;   /* :965 */
        d->global_energy[0] = g_pow;
    }
}

extern "C" void orc_run(orc_state *s, const orc_step *steps, int nsteps, double gamma0,
                        orc_diag_out *rec) {
    orc_diag(s, &rec[0]);
    for (int i = 0; i < nsteps; i++) {
        if (steps[i].marker & 8) break;
        orc_step_all(s, &steps[i], gamma0);
        orc_diag(s, &rec[i + 1]);
    }
}

/*
 * extern "C" harness around the reference's own loader-chain classes
 * (RandomU, Hammerslay, Inverfc, BesselJ, QuietLoading, ShotNoise,
 * GaussHermite), compiled from their sources under the reference tree by
 * build_ref.sh.  TEST INFRASTRUCTURE ONLY: used by tests to pin the oracle
 * restatement bit-for-bit.  Nothing here is copied from the reference; it
 * only calls the reference classes through their public headers.
 */
#include <complex>
#include <cstring>
#include <vector>

#include "BesselJ.h"
#include "GaussHermite.h"
#include "Hammerslay.h"
#include "Inverfc.h"
#include "QuietLoading.h"
#include "RandomU.h"
#include "ShotNoise.h"

/* The two program-wide constants GaussHermite.cpp reads; the reference
 * defines them in its main program file (src/Main/GenMain.cpp:63-68), which
 * is not part of this leaf build. */
extern const double vacimp = 376.73;
extern const double eev = 510998.95069;

extern "C" {

void ref_ran2(unsigned int seed, int n, double *out) {
    RandomU r(seed);
    for (int i = 0; i < n; i++) out[i] = r.getElement();
}

void ref_hammersley(int base_index, unsigned int start, int n, double *out) {
    Hammerslay h(base_index);
    h.set(start);
    for (int i = 0; i < n; i++) out[i] = h.getElement();
}

double ref_inverfc(double y) {
    Inverfc f;
    return f.value(y);
}

double ref_besselj(int n, double x) {
    BesselJ b;
    return b.value(n, x);
}

void ref_quiet_load(double *g, double *th, double *x, double *y, double *px, double *py,
                    const double *sp17, int npart, int nbins, double theta0, int islice) {
    BeamSlice s;
    std::memcpy(&s, sp17, sizeof(BeamSlice));
    std::vector<Particle> p(npart);
    QuietLoading ql;
    int base[6] = {0, 1, 2, 3, 4, 5};
    ql.init(false, base);
    ql.loadQuiet(p.data(), &s, npart, nbins, theta0, islice);
    for (int i = 0; i < npart; i++) {
        g[i] = p[i].gamma; th[i] = p[i].theta; x[i] = p[i].x;
        y[i] = p[i].y; px[i] = p[i].px; py[i] = p[i].py;
    }
}

/* one4one: QuietLoading::init(true, {seed, rank}) once, then loadQuiet
 * for every slice in order on the same object (one stream). */
void ref_quiet_load_one4one(int seed, int rank, int nslice, const long long *off, const double *sp17,
                            double theta0, double *g, double *th, double *x, double *y, double *px,
                            double *py) {
    QuietLoading ql;
    int base[2] = {seed, rank};
    ql.init(true, base);
    for (int j = 0; j < nslice; j++) {
        BeamSlice s;
        std::memcpy(&s, sp17 + 17 * j, sizeof(BeamSlice));
        const long long a = off[j];
        const int n = (int)(off[j + 1] - a);
        std::vector<Particle> p(n > 0 ? n : 1);
        ql.loadQuiet(p.data(), &s, n, 1, theta0, j);
        for (int i = 0; i < n; i++) {
            g[a + i] = p[i].gamma; th[a + i] = p[i].theta; x[a + i] = p[i].x;
            y[a + i] = p[i].y; px[a + i] = p[i].px; py[a + i] = p[i].py;
        }
    }
}

void *ref_shotnoise_new(int seed, int rank) {
    ShotNoise *s = new ShotNoise;
    s->init(seed, rank);
    return s;
}
void ref_shotnoise_free(void *h) { delete (ShotNoise *)h; }
void ref_shotnoise(void *h, double *th, int npart, int nbins, double ne) {
    std::vector<Particle> p(npart);
    for (int i = 0; i < npart; i++) { std::memset(&p[i], 0, sizeof(Particle)); p[i].theta = th[i]; }
    ((ShotNoise *)h)->applyShotNoise(p.data(), npart, nbins, ne);
    for (int i = 0; i < npart; i++) th[i] = p[i].theta;
}

void ref_load_gauss(double *field, int ngrid, double gridmax, double lambda, double power,
                    double w0, double z0, double phase, double xcen, double ycen) {
    FieldSlice s;
    s.lambda = lambda; s.power = power; s.z0 = z0; s.w0 = w0; s.phase = phase;
    s.xcen = xcen; s.ycen = ycen; s.xangle = 0; s.yangle = 0; s.nx = 0; s.ny = 0; s.harm = 1;
    std::vector<std::complex<double>> f(ngrid * ngrid);
    GaussHermite gh;
    gh.loadGauss(f.data(), &s, gridmax, ngrid);
    std::memcpy(field, f.data(), sizeof(double) * 2 * ngrid * ngrid);
}

}

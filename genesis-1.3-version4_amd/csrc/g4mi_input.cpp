// C-ABI of the input generators (lattice, beam, field) -- host only.
#include <cstring>
#include <cmath>
#include <string>

#include "g4mi_host.h"

namespace g4mi {
void set_error(const std::string &msg);
}
using namespace g4mi;

extern "C" {

int g4mi_parse_lattice(const char *text, const char *line, double delz, double lambda0,
                       double gamma0, g4mi_step_params *out, int max_steps, int *nsteps) {
    if (!text || !line || delz <= 0 || lambda0 <= 0 || gamma0 <= 0) {
        set_error("g4mi_parse_lattice: invalid argument");
        return G4MI_E_ARG;
    }
    LatticeResult r = parse_and_unroll(text, line, delz, lambda0, gamma0);
    if (!r.error.empty()) {
        set_error(r.error);
        return G4MI_E_PARSE;
    }
    if (nsteps) *nsteps = (int)r.steps.size();
    for (int i = 0; out && i < (int)r.steps.size() && i < max_steps; i++) out[i] = r.steps[i];
    return G4MI_OK;
}

int g4mi_load_beam(int nslice, const long long *off, const g4mi_slice_beam *slices, int nbins,
                   int shotnoise, int seed, int islice0, double lambda0, double sample,
                   int timerun, double *gamma, double *theta, double *x, double *y, double *px,
                   double *py) {
    if (nslice < 1 || !off || !slices || nbins < 1 || !gamma || !theta || !x || !y || !px || !py) {
        set_error("g4mi_load_beam: invalid argument");
        return G4MI_E_ARG;
    }
    for (int j = 0; j < nslice; j++)
        if ((off[j + 1] - off[j]) % nbins != 0) {
            set_error("g4mi_load_beam: NPART is not a multiple of NBINS");
            return G4MI_E_ARG;
        }
    const double theta0 = 4. * asin(1.);  // LoadBeam.cpp:127
    // quiet loading: each slice from its own reseeded sequence, so slices
    // load independently (threads over slices)
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < nslice; j++) {
        const long long a = off[j];
        const int n = (int)(off[j + 1] - a);
        SliceBeam sb;
        std::memcpy(&sb, &slices[j], sizeof(sb));
        quiet_load(gamma + a, theta + a, x + a, y + a, px + a, py + a, sb, n, nbins, theta0,
                   j + islice0);
    }
    // shot noise: one sequential stream across the slices, in slice order
    if (shotnoise && timerun) {
        Ran2 stream(shot_noise_seed(seed, 0));
        for (int j = 0; j < nslice; j++) {
            const long long a = off[j];
            const int n = (int)(off[j + 1] - a);
            const double ne = slices[j].current * lambda0 * sample / kCeH();
            shot_noise(stream, theta + a, n, nbins, ne);
        }
    }
    return G4MI_OK;
}

int g4mi_one4one_counts(int nslice, const g4mi_slice_beam *slices, double lambda0, double sample,
                        long long *counts) {
    if (nslice < 1 || !slices || !counts || !(lambda0 > 0) || !(sample > 0)) {
        set_error("g4mi_one4one_counts: invalid argument");
        return G4MI_E_ARG;
    }
    for (int j = 0; j < nslice; j++) {  // LoadBeam.cpp:172, 191
        const double ne = slices[j].current * lambda0 * sample / kCeH();
        counts[j] = (long long)(int)round(ne);
        if (counts[j] < 0) {
            set_error("g4mi_one4one_counts: negative current");
            return G4MI_E_ARG;
        }
    }
    return G4MI_OK;
}

int g4mi_load_beam_one4one(int nslice, const long long *off, const g4mi_slice_beam *slices, int seed,
                           double sample, double *gamma, double *theta, double *x, double *y, double *px,
                           double *py) {
    if (nslice < 1 || !off || !slices || !(sample > 0) || !gamma || !theta || !x || !y || !px || !py) {
        set_error("g4mi_load_beam_one4one: invalid argument");
        return G4MI_E_ARG;
    }
    const double theta0 = 4. * asin(1.) * sample;  // LoadBeam.cpp:129-132
    Ran2 stream(one4one_seed(seed, 0));            // one stream through all slices, in order
    for (int j = 0; j < nslice; j++) {
        const long long a = off[j];
        SliceBeam sb;
        std::memcpy(&sb, &slices[j], sizeof(sb));
        quiet_load_one4one(stream, gamma + a, theta + a, x + a, y + a, px + a, py + a, sb,
                           (int)(off[j + 1] - a), theta0);
    }
    return G4MI_OK;
}

int g4mi_load_gauss(double *field, int ngrid, double gridmax, double lambda, double power,
                    double waist, double waist_pos, double phase) {
    if (!field || ngrid < 2 || gridmax <= 0 || lambda <= 0 || waist <= 0) {
        set_error("g4mi_load_gauss: invalid argument");
        return G4MI_E_ARG;
    }
    gauss_field(field, ngrid, gridmax, lambda, power, waist, waist_pos, phase, 0.0, 0.0);
    return G4MI_OK;
}

}  // extern "C"

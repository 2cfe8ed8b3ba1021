// Input-deck driver: a GENESIS 1.3 v4 namelist file and its lattice file ->
// generated beam and field -> the undulator steps on the device ->
// <rootname>.out.h5 in the reference's output layout.
//
// Covered elements (src/Main/GenMain.cpp:80-494 dispatch order, element by
// element in file order): &setup, &lattice (zmatch), &time, &field, &beam,
// &efield, &wake, &sponrad, &track, &write, &importbeam, &importfield and &stop.  Any other element, or a key an element does
// not know, stops the run with an error, as the reference does for unknown
// keys.  Profile references ("@name") are not supported.
#include <sys/stat.h>
#include <unistd.h>
#include <pwd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "g4mi_h5.h"
#include "g4mi_host.h"

namespace g4mi {
void set_error(const std::string &msg);

namespace {

constexpr double kVacimpD = 376.73;         // vacuum impedance [Ohm] (src/Main/GenMain.cpp:63)
constexpr double kEevD = 510998.95069;      // electron rest energy [eV] (src/Main/GenMain.cpp:67)

// ---------------------------------------------------------------------------
// namelist parsing (Parser::parse, src/Main/Parser.cpp:40-92; fillMap :94-124)
// ---------------------------------------------------------------------------
std::string trim_st(const std::string &s) {  // spaces and tabs only
    const size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
    return (a == std::string::npos || b == std::string::npos) ? std::string() : s.substr(a, b - a + 1);
}

bool to_bool(std::string v) {  // StringProcessing::atob
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return v == "1" || v == "true" || v == "t";
}

struct Namelist {
    std::string name;                         // lower case, with the leading '&'
    std::map<std::string, std::string> kv;    // first occurrence of a key wins
};

bool parse_namelists(const std::string &text, std::vector<Namelist> &out, std::string &err) {
    std::istringstream in(text);
    std::string line, list;
    bool open = false;
    Namelist cur;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        line = trim_st(line);
        if (line.empty() || line[0] == '#') continue;
        if (line == "&end" || line == "&END" || line == "&End") {
            if (!open) {
                err = "termination string outside element definition in input file";
                return false;
            }
            // split into key = value pairs
            size_t pos;
            while ((pos = list.find(';')) != std::string::npos) {
                const std::string item = list.substr(0, pos);
                list.erase(0, pos + 1);
                if (item.empty()) continue;
                const size_t eq = item.find('=');
                if (eq == std::string::npos) {
                    err = "invalid format " + item + " in input file";
                    return false;
                }
                const std::string k = trim_st(item.substr(0, eq)), v = trim_st(item.substr(eq + 1));
                cur.kv.emplace(k, v);
            }
            out.push_back(cur);
            cur = Namelist();
            list.clear();
            open = false;
            continue;
        }
        if (line[0] == '&') {
            if (open) {
                err = "nested elements in main input file";
                return false;
            }
            open = true;
            cur.name = line;
            std::transform(cur.name.begin(), cur.name.end(), cur.name.begin(),
                           [](unsigned char c) { return (char)std::tolower(c); });
            continue;
        }
        if (open) {
            list += line;
            list += ";";
        }
    }
    if (open) {
        err = "reached end of input file while parsing namelist '" + cur.name + "'";
        return false;
    }
    return true;
}

// typed key extraction; every key taken is erased so the leftovers are unknown keys
struct Keys {
    Namelist &nl;
    std::string err;
    bool has(const char *k) const { return nl.kv.count(k) > 0; }
    std::string take(const char *k) {
        std::string v = nl.kv.at(k);
        nl.kv.erase(k);
        return v;
    }
    void str(const char *k, std::string &v) {
        if (has(k)) v = take(k);
    }
    void dbl(const char *k, double &v) {
        if (!has(k)) return;
        const std::string s = take(k);
        if (s.find('@') != std::string::npos) {
            if (err.empty()) err = "profile reference '" + s + "' for " + k + " in " + nl.name + " is not supported";
            return;
        }
        v = atof(s.c_str());
    }
    void integer(const char *k, int &v) {
        if (has(k)) v = atoi(take(k).c_str());
    }
    void boolean(const char *k, bool &v) {
        if (has(k)) v = to_bool(take(k));
    }
    bool finish() {
        if (!err.empty()) return false;
        if (!nl.kv.empty()) {
            err = "unknown elements in " + nl.name + ":";
            for (auto &e : nl.kv) err += " " + e.first;
            return false;
        }
        return true;
    }
};

// ---------------------------------------------------------------------------
// deck state
// ---------------------------------------------------------------------------
struct SetupNL {  // Setup::Setup defaults (src/Main/Setup.cpp:7-44)
    std::string rootname, outputdir, lattice, beamline;
    double gamma0 = 5800 / 0.511, lambda0 = 1e-10, delz = 0.015;
    int seed = 123456789, npart = 8192, nbins = 4;
    bool one4one = false, shotnoise = true, beam_global_stat = false, field_global_stat = false;
    bool excl_spatial = false, excl_fft = false, excl_intensity = false, excl_energy = false,
         excl_aux = false, excl_current = true, excl_twiss = true, excl_field_dump = false;
    bool defined = false;
};

struct TimeNL {  // Time::Time / Time::init (src/Main/GenTime.cpp)
    bool dotime = false;
    double s0 = 0, slen = 0;
    int sample = 1;
    int nslice = 1;
    double ds = 0;
};

struct EFieldNL {
    double rmax = 0;
    int ngrid = 100, nz = 0, nphi = 0;
    bool longrange = false;
};

struct WakeNL {  // Wake::Wake defaults (src/Main/Wake.cpp:5-20)
    bool set = false;
    double loss = 0, radius = 2.5e-3, conductivity = 0, relaxation = 0, gap = 0, lgap = 1, hrough = 0,
           lrough = 1;
    bool roundpipe = true;
};

struct SponNL {  // SponRad (src/Main/SponRad.cpp)
    bool set = false;
    int seed = 1234;
    bool loss = false, spread = false;
};

struct Deck {
    std::string input_file, input_text, lattice_text, lattice_path;
    SetupNL setup;
    TimeNL time;
    EFieldNL efield;
    WakeNL wake;
    SponNL spon;
    bool matched = false;
    double mbetax = 0, malphax = 0, mbetay = 0, malphay = 0, mphix = 0, mphiy = 0, zmatch = 0;
    // generated beam (host copy of the loaded state) and field
    std::vector<long long> off;
    std::vector<double> gamma, theta, x, y, px, py, current;
    std::vector<double> field;  // [ns][ng^2][2]
    struct HarmField {
        int harm = 3, ngrid = 0;
        double gridmax = 0;
        std::vector<double> field;  // [ns][ng^2][2], storage order of the fundamental's ring
    };
    std::vector<HarmField> hfields;  // &field harm > 1, in definition order
    int ngrid = 0;
    double gridmax = 0, flambda = 0;
    bool have_beam = false, have_field = false, spon_applied = false;
    // ring offset of the host fields: window slice w is storage slot
    // (w + field_first) % nslice (the device's `first` when the host copy
    // was taken)
    int field_first = 0;
    // device context (created by a &track, released when an element changes
    // the beam or field layout; the host copies are then the state)
    g4mi_ctx *ctx = nullptr;
    int precision = G4MI_FP64;
    bool periodic = false;
    // Setup::runcount: &alter_setup without rootname adds .Run<n+1>
    int runcount = 0;
    // incoherent stream carried over a released context
    bool incoh_saved = false;
    int incoh_state[35] = {0};
    std::string summary;  // JSON of what was set up (check mode)
};

int fail(const std::string &msg, int code = G4MI_E_PARSE) {
    set_error("g4mi_run_deck: " + msg);
    return code;
}

bool read_file(const std::string &path, std::string &out) {
    std::ifstream f(path.c_str(), std::ios::in | std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

std::string dir_of(const std::string &path) {
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

// ---------------------------------------------------------------------------
// Lattice::match (src/Lattice/Lattice.cpp:240-288) with Optics
// (src/Lattice/Optics.cpp): the lattice unrolled at element resolution (a
// step length of 20 m), the transfer matrices of every piece up to z0
// multiplied in order, and the periodic Twiss functions of the product.
// ---------------------------------------------------------------------------
struct Mat2 {
    double a = 1, b = 0, c = 0, d = 1;
};
Mat2 piece(double k1, double L) {  // Optics::getQuad / getDrift
    Mat2 m;
    if (k1 == 0) {
        m.b = L;
        return m;
    }
    const double sk = sqrt(fabs(k1)), w = sk * L;
    if (k1 > 0) {
        m.a = cos(w);
        m.b = sin(w) / sk;
        m.c = -sin(w) * sk;
        m.d = cos(w);
    } else {
        m.a = cosh(w);
        m.b = sinh(w) / sk;
        m.c = sinh(w) * sk;
        m.d = cosh(w);
    }
    return m;
}
Mat2 left_mul(const Mat2 &m, const Mat2 &t) {  // m * t
    Mat2 r;
    r.a = m.a * t.a + m.b * t.c;
    r.b = m.a * t.b + m.b * t.d;
    r.c = m.c * t.a + m.d * t.c;
    r.d = m.c * t.b + m.d * t.d;
    return r;
}

bool match_lattice(Deck &dk, std::string &err) {
    const SetupNL &st = dk.setup;
    LatticeResult lr = parse_and_unroll(dk.lattice_text, st.beamline, 20.0, st.lambda0, st.gamma0);
    if (!lr.error.empty()) {
        err = lr.error;
        return false;
    }
    const auto &s = lr.steps;
    const double z0 = dk.zmatch, g = st.gamma0;
    Mat2 mx, my;
    size_t i = 0;
    while (i < s.size() && s[i].z < z0) {
        const double aw = s[i].aw;
        const double gref2 = g * g - 1 - aw * aw;
        const double qf = s[i].qf * g * g / gref2;
        double qx = aw * aw / gref2;
        const double qy = qx * s[i].ky;
        qx *= s[i].kx;
        double dz = s[i].dz;
        if (s[i].z + s[i].dz > z0) dz = z0 - s[i].z;
        mx = left_mul(piece((qf + qx) == 0 ? 0.0 : qf + qx, dz), mx);
        my = left_mul(piece((-qf + qy) == 0 ? 0.0 : -qf + qy, dz), my);
        i++;
    }
    if (i == s.size() && s.size() > 0 && s.back().z + s.back().dz < z0) {
        dk.matched = false;  // matching position beyond the lattice: ignored
        return true;
    }
    double arg = 2 - mx.a * mx.a - 2 * mx.b * mx.c - mx.d * mx.d;
    if (arg <= 0) {
        dk.matched = false;
        return true;
    }
    dk.mbetax = 2 * mx.b / sqrt(arg);
    dk.malphax = (mx.a - mx.d) / sqrt(arg);
    dk.mphix = acos(0.5 * (mx.a + mx.d)) * 90 / asin(1);
    arg = 2 - my.a * my.a - 2 * my.b * my.c - my.d * my.d;
    if (arg <= 0) {
        dk.matched = false;
        return true;
    }
    dk.mbetay = 2 * my.b / sqrt(arg);
    dk.malphay = (my.a - my.d) / sqrt(arg);
    dk.mphiy = acos(0.5 * (my.a + my.d)) * 90. / asin(1.);
    dk.matched = true;
    return true;
}

// ---------------------------------------------------------------------------
// elements
// ---------------------------------------------------------------------------
bool do_setup(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    SetupNL &s = dk.setup;
    k.str("rootname", s.rootname);
    k.str("outputdir", s.outputdir);
    k.str("lattice", s.lattice);
    k.str("beamline", s.beamline);
    k.dbl("gamma0", s.gamma0);
    k.dbl("lambda0", s.lambda0);
    k.dbl("delz", s.delz);
    k.integer("seed", s.seed);
    k.boolean("one4one", s.one4one);
    k.integer("npart", s.npart);
    k.integer("nbins", s.nbins);
    k.boolean("shotnoise", s.shotnoise);
    k.boolean("beam_global_stat", s.beam_global_stat);
    k.boolean("field_global_stat", s.field_global_stat);
    k.boolean("exclude_spatial_output", s.excl_spatial);
    k.boolean("exclude_fft_output", s.excl_fft);
    k.boolean("exclude_intensity_output", s.excl_intensity);
    k.boolean("exclude_energy_output", s.excl_energy);
    k.boolean("exclude_aux_output", s.excl_aux);
    k.boolean("exclude_current_output", s.excl_current);
    k.boolean("exclude_twiss_output", s.excl_twiss);
    k.boolean("exclude_field_dump", s.excl_field_dump);
    bool dummy = false;
    k.boolean("write_meta_file", dummy);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    std::string path = s.lattice;
    if (!read_file(path, dk.lattice_text)) {
        path = dir_of(dk.input_file) + "/" + s.lattice;
        if (s.lattice.empty() || s.lattice[0] == '/' || !read_file(path, dk.lattice_text)) {
            err = "cannot open lattice file " + s.lattice;
            return false;
        }
    }
    dk.lattice_path = path;
    LatticeResult lr = parse_and_unroll(dk.lattice_text, s.beamline, s.delz, s.lambda0, s.gamma0);
    if (!lr.error.empty()) {
        err = "lattice: " + lr.error;
        return false;
    }
    s.defined = true;
    return true;
}

bool do_lattice(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    k.dbl("zmatch", dk.zmatch);
    if (!k.finish()) {
        err = k.err + " (only zmatch is supported)";
        return false;
    }
    if (dk.zmatch > 0) return match_lattice(dk, err);
    return true;
}

bool do_time(Deck &dk, Namelist &nl, std::string &err) {
    if (dk.have_beam || dk.have_field) return true;  // ignored once beam or field exist (GenMain.cpp)
    Keys k{nl, ""};
    TimeNL &t = dk.time;
    t.dotime = true;
    k.dbl("s0", t.s0);
    k.dbl("slen", t.slen);
    k.integer("sample", t.sample);
    k.boolean("time", t.dotime);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (t.sample < 1) {
        err = "sample rate must be at least 1";
        return false;
    }
    // Time::finishInit (GenTime.cpp) on one rank
    t.ds = dk.setup.lambda0 * t.sample;
    t.nslice = std::max(1, (int)std::round(t.slen / t.ds));
    t.slen = t.ds * t.nslice;
    return true;
}

int time_nslice(const Deck &dk) { return dk.time.dotime || dk.time.ds > 0 ? dk.time.nslice : 1; }
double time_ds(const Deck &dk) { return dk.time.ds > 0 ? dk.time.ds : dk.setup.lambda0 * dk.time.sample; }

bool release_ctx(Deck &dk, std::string &err);
void unroll_ring(std::vector<double> &f, size_t cells, int first);

bool do_field(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    double lambda = dk.setup.lambda0, power = 0, phase = 0, z0 = 0, w0 = 100e-6;
    double xcen = 0, ycen = 0, xangle = 0, yangle = 0, dgrid = 1e-3;
    int ngrid = 151, harm = 1, nx = 0, ny = 0;
    bool add = false;
    k.dbl("lambda", lambda);
    k.dbl("power", power);
    k.dbl("phase", phase);
    k.dbl("waist_pos", z0);
    k.dbl("waist_size", w0);
    k.dbl("xcenter", xcen);
    k.dbl("ycenter", ycen);
    k.dbl("xangle", xangle);
    k.dbl("yangle", yangle);
    k.dbl("dgrid", dgrid);
    k.integer("ngrid", ngrid);
    k.integer("harm", harm);
    k.integer("nx", nx);
    k.integer("ny", ny);
    k.boolean("accumulate", add);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (harm < 1 || nx != 0 || ny != 0) {
        err = "&field: only TEM00 modes (nx = ny = 0) at harm >= 1 are supported";
        return false;
    }
    // a field defined after a &track replaces or adds to the tracked one
    if (!release_ctx(dk, err)) return false;
    Deck::HarmField *hfp = nullptr;  // LoadField.cpp:99-116: the record of this harmonic
    if (harm > 1 && fabs(lambda - dk.setup.lambda0) > 1e-12 * dk.setup.lambda0) {
        err = "&field: harmonic fields at a wavelength other than lambda0 are not supported by this build";
        return false;
    }
    if (harm > 1) {
        if (!dk.have_field) {
            err = "&field: define the fundamental (harm = 1) before its harmonics";
            return false;
        }
        for (auto &e : dk.hfields)
            if (e.harm == harm) hfp = &e;
        if (!hfp) {
            if (dk.hfields.size() >= 4) {
                err = "&field: at most 4 harmonic fields";
                return false;
            }
            dk.hfields.push_back(Deck::HarmField());
            hfp = &dk.hfields.back();
            hfp->harm = harm;
            add = false;
        } else if (add) {
            dgrid = hfp->gridmax;
            ngrid = hfp->ngrid;
        }
    } else if (add && dk.have_field) {
        dgrid = dk.gridmax;
        ngrid = dk.ngrid;
    } else {
        add = false;
    }
    if (ngrid < 3 || dgrid <= 0) {
        err = "&field: invalid grid";
        return false;
    }
    // Field::init resets the ring offset of the field it (re)defines
    // (src/Core/Field.cpp:66): an accumulated field takes its storage order
    // as the new window order, every other field keeps its window order.
    if (dk.field_first != 0) {
        if (dk.have_field && !(add && !hfp)) unroll_ring(dk.field, (size_t)dk.ngrid * dk.ngrid, dk.field_first);
        for (auto &e : dk.hfields)
            if (!(add && hfp == &e)) unroll_ring(e.field, (size_t)e.ngrid * e.ngrid, dk.field_first);
        dk.field_first = 0;
    }
    const int ns = time_nslice(dk);
    const double ds = time_ds(dk);
    const double l0 = dk.setup.lambda0;
    const double dlam = -(lambda - l0) / l0 / l0 * 4 * asin(1);  // LoadField.cpp
    const size_t cells = (size_t)ngrid * ngrid;
    std::vector<double> &target = hfp ? hfp->field : dk.field;
    if (!add) target.assign((size_t)ns * cells * 2, 0.0);
    std::vector<double> slice(cells * 2);
    for (int i = 0; i < ns; i++) {
        const double s = dk.time.s0 + i * ds;
        gauss_field(slice.data(), ngrid, dgrid, lambda, power, w0, z0, fmod(phase + s * dlam, 4 * asin(1.)), xcen,
                    ycen, xangle, yangle, harm);
        double *dst = target.data() + (size_t)i * cells * 2;
        if (add)
            for (size_t q = 0; q < cells * 2; q++) dst[q] += slice[q];
        else
            std::memcpy(dst, slice.data(), cells * 2 * sizeof(double));
    }
    if (hfp) {
        hfp->ngrid = ngrid;
        hfp->gridmax = dgrid;
        return true;
    }
    dk.ngrid = ngrid;
    dk.gridmax = dgrid;
    dk.flambda = lambda;
    dk.have_field = true;
    return true;
}

bool do_beam(Deck &dk, Namelist &nl, std::string &err) {
    if (dk.have_beam) {
        err = "cannot generate beam, because beam is already defined";
        return false;
    }
    Keys k{nl, ""};
    g4mi_slice_beam sb{};
    sb.gamma = dk.setup.gamma0;
    sb.delgam = 0;
    sb.current = 1000;
    sb.ex = sb.ey = 0.3e-6;
    sb.betax = sb.betay = 15;
    if (dk.matched) {
        sb.betax = dk.mbetax;
        sb.alphax = dk.malphax;
        sb.betay = dk.mbetay;
        sb.alphay = dk.malphay;
    }
    k.dbl("gamma", sb.gamma);
    k.dbl("delgam", sb.delgam);
    k.dbl("current", sb.current);
    k.dbl("ex", sb.ex);
    k.dbl("ey", sb.ey);
    k.dbl("betax", sb.betax);
    k.dbl("betay", sb.betay);
    k.dbl("alphax", sb.alphax);
    k.dbl("alphay", sb.alphay);
    k.dbl("xcenter", sb.xcen);
    k.dbl("ycenter", sb.ycen);
    k.dbl("pxcenter", sb.pxcen);
    k.dbl("pycenter", sb.pycen);
    k.dbl("bunch", sb.bunch);
    k.dbl("bunchphase", sb.bunchphase);
    k.dbl("emod", sb.emod);
    k.dbl("emodphase", sb.emodphase);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    const SetupNL &st = dk.setup;
    const int nbins = st.one4one ? 1 : st.nbins;  // LoadBeam.cpp:129-132
    if (st.npart % nbins != 0) {
        err = "NPART is not a multiple of NBINS";
        return false;
    }
    const int ns = time_nslice(dk);
    std::vector<g4mi_slice_beam> slices(ns, sb);
    dk.off.assign(ns + 1, 0);
    if (st.one4one) {  // one electron per macroparticle
        std::vector<long long> cnt(ns);
        if (g4mi_one4one_counts(ns, slices.data(), st.lambda0, (double)dk.time.sample, cnt.data()) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
        for (int i = 0; i < ns; i++) dk.off[i + 1] = dk.off[i] + cnt[i];
    } else {
        for (int i = 0; i < ns; i++) dk.off[i + 1] = dk.off[i] + st.npart;
    }
    const size_t n = (size_t)dk.off[ns];
    for (auto *v : {&dk.gamma, &dk.theta, &dk.x, &dk.y, &dk.px, &dk.py}) v->assign(n, 0.0);
    const int rc = st.one4one
                       ? g4mi_load_beam_one4one(ns, dk.off.data(), slices.data(), st.seed, (double)dk.time.sample,
                                                dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(),
                                                dk.px.data(), dk.py.data())
                       : g4mi_load_beam(ns, dk.off.data(), slices.data(), st.nbins, st.shotnoise ? 1 : 0, st.seed,
                                        0, st.lambda0, (double)dk.time.sample, dk.time.dotime ? 1 : 0,
                                        dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(), dk.px.data(),
                                        dk.py.data());
    if (rc != G4MI_OK) {
        err = "beam loading failed";
        return false;
    }
    dk.current.assign(ns, sb.current);
    dk.have_beam = true;
    return true;
}

// Time window taken from an imported file when no &time element set one
// (ReadBeamHDF5/ReadFieldHDF5::readGlobal: Time::setSampleRate rounds to
// an integer, then Time::init/finishInit).
bool time_from_file(Deck &dk, double s0, double slicelen, double reflen, int count, bool dotime,
                    std::string &err) {
    TimeNL &t = dk.time;
    if (t.ds <= 0 && count > 1) {
        t.sample = (int)std::round(slicelen / reflen);
        t.s0 = s0;
        t.slen = slicelen * count;
        t.dotime = dotime;
        if (t.sample < 1) {
            err = "sample rate must be at least 1";
            return false;
        }
        t.ds = dk.setup.lambda0 * t.sample;
        t.nslice = std::max(1, (int)std::round(t.slen / t.ds));
        t.slen = t.ds * t.nslice;
    }
    const double sample = (double)t.sample;
    if (fabs(slicelen - dk.setup.lambda0 * sample) / slicelen > 1e-6) {
        err = "mismatch in sample rate of run and of input file";
        return false;
    }
    return true;
}

bool read_scalar(const std::string &file, const char *name, double &v, std::string &err) {
    std::vector<double> buf;
    long long dims[2];
    if (h5_read_doubles(file, name, buf, dims, err) < 1) {
        if (err.empty()) err = std::string("dataset ") + name + " is empty";
        err = file + ": " + err;
        return false;
    }
    v = buf[0];
    return true;
}

// &importbeam (ImportBeam::init, src/Loading/ImportBeam.cpp:25-85, with
// ReadBeamHDF5::readGlobal/readSlice, src/IO/readBeamHDF5.cpp:23-185): the
// particles of every slice of the time window from a .par.h5 file; a
// window position outside the file's slices gets the file's first slice
// with zero current and zero transverse coordinates, a position between
// its slices stays empty.
bool do_importbeam(Deck &dk, Namelist &nl, std::string &err) {
    if (dk.have_beam) {
        err = "cannot import beam, because beam is already defined";
        return false;
    }
    Keys k{nl, ""};
    std::string file;
    bool dotime = true;
    k.str("file", file);
    k.boolean("time", dotime);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    double s0f = 0, reflen = 0, slicelen = 0, countd = 0, nbinsd = 0, o4o = 0;
    if (!read_scalar(file, "refposition", s0f, err) || !read_scalar(file, "slicelength", reflen, err) ||
        !read_scalar(file, "slicespacing", slicelen, err) || !read_scalar(file, "slicecount", countd, err) ||
        !read_scalar(file, "beamletsize", nbinsd, err) || !read_scalar(file, "one4one", o4o, err))
        return false;
    const SetupNL &st = dk.setup;
    const int count = (int)countd;
    if (fabs((st.lambda0 - reflen) / st.lambda0) > 1e-6) {
        err = "mismatch between reference length of run and of input file";
        return false;
    }
    if ((int)nbinsd != st.nbins && !(st.one4one && (int)nbinsd == 1)) {
        err = "mismatch between beamlet size (NBINS) of run and of input file";
        return false;
    }
    if (((st.one4one) && (o4o == 0)) || ((!st.one4one) && (o4o != 0))) {
        err = "mismatch between ONE4ONE flag of run and of input file";
        return false;
    }
    if (!time_from_file(dk, s0f, slicelen, reflen, count, dotime, err)) return false;
    const int ns = time_nslice(dk);
    const double ds = time_ds(dk);
    dk.off.assign(ns + 1, 0);
    dk.current.assign(ns, 0.0);
    std::vector<std::vector<double>> parts((size_t)ns * 6);
    static const char *names[6] = {"gamma", "theta", "x", "y", "px", "py"};
    for (int i = 0; i < ns; i++) {
        const double s = dk.time.s0 + i * ds;
        const double rslice = (s - s0f) / slicelen;
        long long n = 0;
        if (fabs(rslice - std::round(rslice)) <= 1e-3) {
            int islice = (int)std::round(rslice) + 1;
            double wei = 1.;
            if (islice < 1 || islice > count) {
                wei = 0;
                islice = 1;
            }
            char g[32];
            snprintf(g, sizeof(g), "slice%6.6d/", islice);
            double cur = 0;
            if (!read_scalar(file, (std::string(g) + "current").c_str(), cur, err)) return false;
            dk.current[i] = cur * wei;
            for (int q = 0; q < 6; q++) {
                long long dims[2];
                std::vector<double> &v = parts[(size_t)i * 6 + q];
                if (h5_read_doubles(file, std::string(g) + names[q], v, dims, err) < 0) {
                    err = file + ": " + err;
                    return false;
                }
                if (q >= 2)
                    for (double &x : v) x *= wei;
                if (q == 0) n = (long long)v.size();
                if ((long long)v.size() != n) {
                    err = file + ": slice " + std::to_string(islice) + " has datasets of different sizes";
                    return false;
                }
            }
        }
        dk.off[i + 1] = dk.off[i] + n;
    }
    const size_t ntot = (size_t)dk.off[ns];
    std::vector<double> *dst[6] = {&dk.gamma, &dk.theta, &dk.x, &dk.y, &dk.px, &dk.py};
    for (int q = 0; q < 6; q++) {
        dst[q]->assign(ntot, 0.0);
        for (int i = 0; i < ns; i++) {
            const std::vector<double> &v = parts[(size_t)i * 6 + q];
            std::copy(v.begin(), v.end(), dst[q]->begin() + dk.off[i]);
        }
    }
    dk.have_beam = true;
    return true;
}

// &importfield (ImportField::init, src/Loading/ImportField.cpp:23-91, with
// ReadFieldHDF5::readGlobal/readSlice, src/IO/readFieldHDF5.cpp:23-120):
// the grid and the slices of a .fld.h5 file, scaled back from sqrt(W) per
// cell by dgrid eev / ks / sqrt(vacimp) and by `attenuation`, the file's
// start shifted by `offset` rounded to whole slices.
bool do_importfield(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    std::string file;
    bool dotime = true;
    int harm = 1;
    double attenuation = 1.0, offset = 0.0;
    k.str("file", file);
    k.boolean("time", dotime);
    k.integer("harmonic", harm);
    k.dbl("attenuation", attenuation);
    k.dbl("offset", offset);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (harm < 1) {
        err = "&importfield: harmonic must be at least 1";
        return false;
    }
    bool exists = harm == 1 && dk.have_field;
    for (const auto &e : dk.hfields) exists |= e.harm == harm;
    if (exists) {  // ImportField.cpp:73-91
        err = "cannot import field, because field is already defined";
        return false;
    }
    if (harm > 1 && !dk.have_field) {
        err = "&importfield: define or import the fundamental (harmonic = 1) before its harmonics";
        return false;
    }
    if (harm > 1 && dk.hfields.size() >= 4) {
        err = "&importfield: at most 4 harmonic fields";
        return false;
    }
    // a field imported after a &track: the tracked fields keep their window
    // order, the new one starts with ring offset 0 (Field::init)
    if (!release_ctx(dk, err)) return false;
    if (dk.field_first != 0) {
        unroll_ring(dk.field, (size_t)dk.ngrid * dk.ngrid, dk.field_first);
        for (auto &e : dk.hfields) unroll_ring(e.field, (size_t)e.ngrid * e.ngrid, dk.field_first);
        dk.field_first = 0;
    }
    double s0f = 0, dgrid = 0, reflen = 0, slicelen = 0, countd = 0, ngd = 0;
    if (!read_scalar(file, "refposition", s0f, err) || !read_scalar(file, "gridsize", dgrid, err) ||
        !read_scalar(file, "wavelength", reflen, err) || !read_scalar(file, "slicespacing", slicelen, err) ||
        !read_scalar(file, "slicecount", countd, err) || !read_scalar(file, "gridpoints", ngd, err))
        return false;
    const int count = (int)countd, ngrid = (int)ngd;
    if (ngrid < 3) {
        err = file + ": invalid grid";
        return false;
    }
    s0f += slicelen * std::round(offset / slicelen);
    const double ks = 4. * asin(1) / reflen;
    const double scl = 1. / (dgrid * kEevD / ks / sqrt(kVacimpD)) * attenuation;
    const double gridmax = 0.5 * dgrid * (double)(ngrid - 1);
    if (fabs((dk.setup.lambda0 - reflen * (double)harm) / dk.setup.lambda0) > 1e-6) {
        err = "mismatch between reference length of run and of input file";
        return false;
    }
    if (!time_from_file(dk, s0f, slicelen, reflen, count, dotime, err)) return false;
    const int ns = time_nslice(dk);
    const double ds = time_ds(dk);
    const size_t cells = (size_t)ngrid * ngrid;
    std::vector<double> &target = harm > 1 ? (dk.hfields.emplace_back(), dk.hfields.back().field) : dk.field;
    target.assign((size_t)ns * cells * 2, 0.0);
    for (int i = 0; i < ns; i++) {
        const double s = dk.time.s0 + i * ds;
        const double rslice = (s - s0f) / slicelen;
        if (fabs(rslice - std::round(rslice)) > 1e-3) continue;
        const int islice = (int)std::round(rslice) + 1;
        if (islice < 1 || islice > count) continue;
        char g[32];
        snprintf(g, sizeof(g), "slice%6.6d/", islice);
        std::vector<double> re, im;
        long long dims[2];
        if (h5_read_doubles(file, std::string(g) + "field-real", re, dims, err) < 0 ||
            h5_read_doubles(file, std::string(g) + "field-imag", im, dims, err) < 0) {
            err = file + ": " + err;
            return false;
        }
        if (re.size() != cells || im.size() != cells) {
            err = file + ": slice " + std::to_string(islice) + " does not hold gridpoints^2 values";
            return false;
        }
        double *f = target.data() + (size_t)i * cells * 2;
        for (size_t j = 0; j < cells; j++) {
            f[2 * j] = scl * re[j];
            f[2 * j + 1] = scl * im[j];
        }
    }
    if (harm > 1) {
        Deck::HarmField &h = dk.hfields.back();
        h.harm = harm;
        h.ngrid = ngrid;
        h.gridmax = gridmax;
        return true;
    }
    dk.ngrid = ngrid;
    dk.gridmax = gridmax;
    dk.flambda = reflen;
    dk.have_field = true;
    return true;
}

// &wake (Wake::init, src/Main/Wake.cpp:48-133): the wake parameters; the
// loss per slice is computed when a &track starts, from the beam's current
// (Collective::update runs at the first step, src/Core/Collective.cpp:110-113).
bool do_wake(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    WakeNL &w = dk.wake;
    w = WakeNL();
    std::string loss, material, output;
    bool transient = false;
    double ztrans = 0;
    k.str("loss", loss);
    k.dbl("radius", w.radius);
    k.dbl("conductivity", w.conductivity);
    k.dbl("relaxation", w.relaxation);
    k.boolean("roundpipe", w.roundpipe);
    k.str("material", material);
    k.dbl("gap", w.gap);
    k.dbl("lgap", w.lgap);
    k.dbl("hrough", w.hrough);
    k.dbl("lrough", w.lrough);
    k.boolean("transient", transient);
    k.dbl("ztrans", ztrans);
    k.str("output", output);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (!loss.empty()) {
        if (loss[0] == '@') {
            err = "&wake: profile references (loss = @...) are not supported by this build";
            return false;
        }
        char *end = nullptr;
        w.loss = strtod(loss.c_str(), &end);
        if (end == loss.c_str()) {
            err = "&wake: loss is not a number";
            return false;
        }
    }
    if (!output.empty()) {
        err = "&wake: output of the wake potentials is not supported by this build";
        return false;
    }
    if (material == "CU" || material == "Cu" || material == "cu") {
        w.conductivity = 5.813e7;
        w.relaxation = 8.1e-6;
    }
    if (material == "AL" || material == "Al" || material == "al") {
        w.conductivity = 3.571e7;
        w.relaxation = 2.4e-6;
    }
    if (!(w.radius > 0) || !(w.lgap > 0) || !(w.lrough > 0)) {
        err = "&wake: radius, lgap and lrough must be positive";
        return false;
    }
    w.set = true;  // transient/ztrans do not change the loss (Collective.cpp:162-191)
    return true;
}

// &sponrad (SponRad::init, src/Main/SponRad.cpp:20-37)
bool do_sponrad(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    SponNL &sp = dk.spon;
    k.integer("seed", sp.seed);
    k.boolean("doLoss", sp.loss);
    k.boolean("doSpread", sp.spread);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    sp.set = true;
    dk.spon_applied = false;  // a new stream from this seed
    dk.incoh_saved = false;
    if (dk.ctx && g4mi_set_incoherent(dk.ctx, sp.seed, dk.setup.one4one ? 1 : dk.setup.nbins, sp.loss ? 1 : 0,
                                      sp.spread ? 1 : 0) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    return true;
}

bool do_efield(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    EFieldNL &e = dk.efield;
    k.dbl("rmax", e.rmax);
    k.integer("ngrid", e.ngrid);
    k.integer("nz", e.nz);
    k.integer("nphi", e.nphi);
    k.boolean("longrange", e.longrange);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    return true;
}

// per-slice twiss moments of the current beam state (DiagBeam::getValues,
// src/Core/Diagnostic.cpp:415-545, the values stored once at iz = 0)
void twiss_once(const Deck &dk, std::vector<double> *out /* ex, ey, bx, by, ax, ay */) {
    const int ns = (int)dk.off.size() - 1;
    for (int q = 0; q < 6; q++) out[q].assign(ns, 0.0);
    for (int is = 0; is < ns; is++) {
        double x1 = 0, x2 = 0, y1 = 0, y2 = 0, px1 = 0, px2 = 0, py1 = 0, py2 = 0, xpx = 0, ypy = 0, g1 = 0;
        for (long long i = dk.off[is]; i < dk.off[is + 1]; i++) {
            x1 += dk.x[i];
            x2 += dk.x[i] * dk.x[i];
            y1 += dk.y[i];
            y2 += dk.y[i] * dk.y[i];
            g1 += dk.gamma[i];
            px1 += dk.px[i];
            py1 += dk.py[i];
            px2 += dk.px[i] * dk.px[i];
            py2 += dk.py[i] * dk.py[i];
            xpx += dk.x[i] * dk.px[i];
            ypy += dk.y[i] * dk.py[i];
        }
        const long long n = dk.off[is + 1] - dk.off[is];
        const double norm = n > 0 ? 1. / (double)n : 1.;
        x1 *= norm, x2 *= norm, y1 *= norm, y2 *= norm, px1 *= norm, px2 *= norm, py1 *= norm, py2 *= norm;
        g1 *= norm, xpx *= norm, ypy *= norm;
        const double ex = sqrt(fabs((x2 - x1 * x1) * (px2 - px1 * px1) - (xpx - x1 * px1) * (xpx - x1 * px1)));
        const double ey = sqrt(fabs((y2 - y1 * y1) * (py2 - py1 * py1) - (ypy - y1 * py1) * (ypy - y1 * py1)));
        out[0][is] = ex;
        out[1][is] = ey;
        out[2][is] = (x2 - x1 * x1) / ex * g1;
        out[3][is] = (y2 - y1 * y1) / ey * g1;
        out[4][is] = -(xpx - x1 * px1) / ex;
        out[5][is] = -(ypy - y1 * py1) / ey;
    }
}

std::string env_or(const char *k, const char *dflt) {
    const char *v = getenv(k);
    return v ? std::string(v) : std::string(dflt);
}

// ---------------------------------------------------------------------------
// dumps
// ---------------------------------------------------------------------------
void write_version(H5File &f, bool &ok) {
    ok &= f.write_single("Meta/Version/Major", " ", {4.0});
    ok &= f.write_single("Meta/Version/Minor", " ", {6.0});
    ok &= f.write_single("Meta/Version/Revision", " ", {0.0});
    ok &= f.write_single("Meta/Version/Beta", " ", {0.0});
    ok &= f.write_string("Meta/Version/Build_Info", "g4mi: undulator step on MI355X (HIP), output layout of GENESIS 1.3 v4");
}

// the state on the host: download from the device when a run has started
bool host_state(Deck &dk, int *first, std::string &err) {
    *first = dk.field_first;
    if (!dk.ctx) return true;
    // one4one sorting moves particles between slices on the device
    if (g4mi_get_slices(dk.ctx, dk.off.data(), dk.current.data()) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    const size_t n = (size_t)dk.off.back();
    for (auto *v : {&dk.gamma, &dk.theta, &dk.x, &dk.y, &dk.px, &dk.py}) v->resize(n);
    if (g4mi_beam_download(dk.ctx, dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(), dk.px.data(),
                           dk.py.data()) != G4MI_OK ||
        g4mi_field_download(dk.ctx, dk.field.data(), first) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    for (size_t k = 0; k < dk.hfields.size(); k++)
        if (g4mi_harmonic_download(dk.ctx, (int)k, dk.hfields[k].field.data()) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
    return true;
}

// a ring-stored field ([nslice][cells] complex, window slice w in storage
// slot (w + first) % nslice) in window order
void unroll_ring(std::vector<double> &f, size_t cells, int first) {
    const size_t per = cells * 2, ns = per ? f.size() / per : 0;
    if (first == 0 || ns < 2) return;
    std::vector<double> t(f.size());
    for (size_t w = 0; w < ns; w++)
        std::memcpy(t.data() + w * per, f.data() + ((w + (size_t)first) % ns) * per, per * sizeof(double));
    f.swap(t);
}

// Take the state off the device: beam, fields, ring offset and the
// incoherent stream go to the host copies and the context is destroyed;
// the next &track builds a new one from them.
bool release_ctx(Deck &dk, std::string &err) {
    if (!dk.ctx) return true;
    int first = 0;
    if (!host_state(dk, &first, err)) return false;
    dk.field_first = first;
    if (dk.spon.set && dk.spon_applied && (dk.spon.loss || dk.spon.spread)) {
        if (g4mi_incoherent_state(dk.ctx, dk.incoh_state, 0) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
        dk.incoh_saved = true;
    }
    g4mi_destroy(dk.ctx);
    dk.ctx = nullptr;
    return true;
}

// Beam::sort (src/Core/Beam.cpp:81-94) of the host beam on the device, in a
// context of its own (the field grid is a placeholder); offsets, currents
// and coordinates come back to the host copies.
bool device_sort(Deck &dk, int device, std::string &err) {
    const int ns = (int)dk.off.size() - 1;
    g4mi_ctx *c = nullptr;
    long long kept = 0;
    bool ok = g4mi_create(&c, device, ns, dk.off.data(), 3, 1.0, dk.setup.lambda0, (double)dk.time.sample,
                          dk.time.dotime ? 1 : 0, 0, G4MI_FP64) == G4MI_OK &&
              g4mi_beam_upload(c, dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(), dk.px.data(),
                               dk.py.data(), dk.current.data()) == G4MI_OK &&
              g4mi_set_sorting(c, 1) == G4MI_OK && g4mi_sort(c, &kept) == G4MI_OK &&
              g4mi_get_slices(c, dk.off.data(), dk.current.data()) == G4MI_OK;
    if (ok) {
        const size_t n = (size_t)dk.off.back();
        for (auto *v : {&dk.gamma, &dk.theta, &dk.x, &dk.y, &dk.px, &dk.py}) v->resize(n);
        ok = g4mi_beam_download(c, dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(), dk.px.data(),
                                dk.py.data()) == G4MI_OK;
    }
    if (!ok) err = std::string("device: ") + g4mi_last_error();
    if (c) g4mi_destroy(c);
    return ok;
}

// <basename>.fld.h5 (WriteFieldHDF5::writeMain, src/IO/writeFieldHDF5.cpp:42-171):
// globals, one group per slice in window order with the field scaled to
// sqrt(W) per cell, and the intensity projections int_xy, int_xz, int_yz.
bool write_field_dump_of(const Deck &dk, const std::vector<double> &field, int ngrid, double gridmax,
                         double xlambda, double slicelen, const std::string &basename, int first,
                         std::string &err);

// every field: the fundamental as <basename>.fld.h5, harmonic h as
// <basename>.h<h>.fld.h5 (WriteFieldHDF5::write, writeFieldHDF5.cpp:13-39)
bool write_field_dump(const Deck &dk, const std::string &basename, int first, std::string &err) {
    if (!write_field_dump_of(dk, dk.field, dk.ngrid, dk.gridmax, dk.flambda, dk.flambda * dk.time.sample,
                             basename, first, err))
        return false;
    for (const auto &h : dk.hfields)  // Field::init: xlambda = lambda / harm
        if (!write_field_dump_of(dk, h.field, h.ngrid, h.gridmax, dk.setup.lambda0 / h.harm,
                                 dk.setup.lambda0 * dk.time.sample, basename + ".h" + std::to_string(h.harm), first,
                                 err))
            return false;
    return true;
}

bool write_field_dump_of(const Deck &dk, const std::vector<double> &field, int ngrid, double gridmax,
                         double xlambda, double slicelen, const std::string &basename, int first,
                         std::string &err) {
    const std::string fname = basename + ".fld.h5";
    H5File f(fname);
    if (!f.ok()) {
        err = "unable to write " + fname + ": " + f.error();
        return false;
    }
    bool ok = true;
    write_version(f, ok);
    const int ng = ngrid;
    const size_t cells = (size_t)ng * ng;
    const int ns = (int)(field.size() / (2 * cells));  // the field's own slices (a beam may not exist yet)
    const double dgrid = 2 * gridmax / (ng - 1);
    ok &= f.write_single("wavelength", "m", {xlambda});
    ok &= f.write_single("slicespacing", "m", {slicelen});
    ok &= f.write_single("refposition", "m", {dk.time.s0});
    ok &= f.write_single("gridsize", "m", {dgrid});
    ok &= f.write_int("gridpoints", {ng});
    ok &= f.write_int("slicecount", {ns});
    const double ks = 4. * asin(1) / xlambda;
    const double scl = dgrid * kEevD / ks / sqrt(kVacimpD);
    std::vector<double> re(cells), im(cells), ixy(cells, 0.0), ixz((size_t)ng * ns, 0.0), iyz((size_t)ng * ns, 0.0);
    char name[32];
    for (int i = 0; i < ns; i++) {
        const double *fs = field.data() + (size_t)((i + first) % ns) * cells * 2;
        for (size_t j = 0; j < cells; j++) {
            re[j] = scl * fs[2 * j];
            im[j] = scl * fs[2 * j + 1];
        }
        if (!dk.setup.excl_field_dump) {
            snprintf(name, sizeof(name), "slice%6.6d", i + 1);
            ok &= f.write_single(std::string(name) + "/field-real", " ", re);
            ok &= f.write_single(std::string(name) + "/field-imag", " ", im);
        }
        for (int iy = 0; iy < ng; iy++)
            for (int ix = 0; ix < ng; ix++) {
                const size_t j = (size_t)iy * ng + ix;
                const double w = fs[2 * j] * fs[2 * j] + fs[2 * j + 1] * fs[2 * j + 1];
                ixy[j] += w;
                ixz[(size_t)ns * ix + i] += w;
                iyz[(size_t)ns * iy + i] += w;
            }
    }
    ok &= f.write_single("int_xy", " ", ixy);
    ok &= f.write_buffer("int_xz", " ", ixz, (size_t)ng, (size_t)ns);
    ok &= f.write_buffer("int_yz", " ", iyz, (size_t)ng, (size_t)ns);
    f.close();
    if (!ok) err = "writing " + fname + " failed";
    return ok;
}

// <basename>.par.h5 (WriteBeamHDF5::write, src/IO/writeBeamHDF5.cpp:14-148):
// globals and one group per slice with the current and every stride-th
// particle's coordinates.
bool write_beam_dump(const Deck &dk, const std::string &basename, int stride, std::string &err) {
    const std::string fname = basename + ".par.h5";
    H5File f(fname);
    if (!f.ok()) {
        err = "unable to write " + fname + ": " + f.error();
        return false;
    }
    bool ok = true;
    write_version(f, ok);
    const int ns = (int)dk.off.size() - 1;
    ok &= f.write_single("slicelength", "m", {dk.setup.lambda0});
    ok &= f.write_single("slicespacing", "m", {dk.setup.lambda0 * dk.time.sample});
    ok &= f.write_single("refposition", "m", {dk.time.s0});
    ok &= f.write_int("beamletsize", {dk.setup.one4one ? 1 : dk.setup.nbins});
    ok &= f.write_int("slicecount", {ns});
    ok &= f.write_int("one4one", {dk.setup.one4one ? 1 : 0});
    if (stride < 1) stride = 1;
    char name[32];
    std::vector<double> w;
    for (int i = 0; i < ns; i++) {
        snprintf(name, sizeof(name), "slice%6.6d", i + 1);
        const std::string g(name);
        ok &= f.write_single(g + "/current", "A", {dk.current[i]});
        const long long a = dk.off[i], b = dk.off[i + 1];
        auto put = [&](const char *n, const char *unit, const std::vector<double> &v) {
            w.clear();
            for (long long k = a; k < b; k += stride) w.push_back(v[(size_t)k]);
            if (!w.empty()) ok &= f.write_single(g + "/" + n, unit, w);
        };
        put("gamma", " ", dk.gamma);
        put("theta", "rad", dk.theta);
        put("x", "m", dk.x);
        put("y", "m", dk.y);
        put("px", "rad", dk.px);
        put("py", "rad", dk.py);
    }
    f.close();
    if (!ok) err = "writing " + fname + " failed";
    return ok;
}

std::string file_name(const SetupNL &st, std::string name) {  // Setup::RootName_to_FileName
    size_t p;
    while ((p = name.find('@')) != std::string::npos) name.replace(p, 1, st.rootname);
    return st.outputdir.empty() ? name : st.outputdir + "/" + name;
}

// &write (src/Main/Dump.cpp:19-62)
bool do_write(Deck &dk, Namelist &nl, std::string &err) {
    Keys k{nl, ""};
    std::string fieldname, beamname;
    int stride = 1;
    k.str("field", fieldname);
    k.str("beam", beamname);
    k.integer("stride", stride);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    int first = 0;
    if (!host_state(dk, &first, err)) return false;
    if (!fieldname.empty()) {
        if (!dk.have_field) {
            err = "&write: no field defined";
            return false;
        }
        if (!write_field_dump(dk, file_name(dk.setup, fieldname), first, err)) return false;
    }
    if (!beamname.empty()) {
        if (!dk.have_beam) {
            err = "&write: no beam defined";
            return false;
        }
        if (!write_beam_dump(dk, file_name(dk.setup, beamname), stride < 1 ? 1 : stride, err)) return false;
    }
    return true;
}

struct DumpList {
    std::vector<int> field_step, beam_step;
    std::vector<std::string> field_file, beam_file;
};

// Writes <rootname>.out.h5: /Meta, /Global, /Lattice, /Beam, /Field
// (Diagnostic::writeToOutputFile, src/Core/Diagnostic.cpp:135-245;
// Output::writeMeta :67-160; HDF5Base writers).
bool write_output(const Deck &dk, const std::string &fname, const std::vector<g4mi_step_params> &steps,
                  const std::vector<double> &zout, const std::map<std::string, std::vector<double>> &rec,
                  size_t nrec, const DumpList &dl, std::string &err) {
    H5File f(fname);
    if (!f.ok()) {
        err = "unable to open output file " + fname + ": " + f.error();
        return false;
    }
    const SetupNL &st = dk.setup;
    const int ns = (int)dk.off.size() - 1;
    bool ok = true;
    // Meta
    write_version(f, ok);
    {
        time_t now = time(nullptr);
        char buf[64];
        ctime_r(&now, buf);
        ok &= f.write_string("Meta/TimeStamp", buf);
    }
    ok &= f.write_string("Meta/HOST", env_or("HOST", "Undefined"));
    {
        struct passwd *pw = getpwuid(getuid());
        ok &= f.write_string("Meta/User", pw ? pw->pw_name : "username lookup failed");
        char cwd[4096];
        ok &= f.write_string("Meta/cwd", getcwd(cwd, sizeof(cwd)) ? cwd : "getcwd call failed");
    }
    ok &= f.write_string("Meta/InputFile", dk.input_text);
    ok &= f.write_string("Meta/LatticeFile", dk.lattice_text);
    // Output::reportDumps (src/IO/Output.cpp:200-247)
    ok &= f.write_int("Meta/Fielddumps/ndumps", {(int)dl.field_step.size()});
    if (!dl.field_step.empty()) ok &= f.write_int("Meta/Fielddumps/intstep", dl.field_step);
    for (size_t k = 0; k < dl.field_file.size(); k++)
        ok &= f.write_string("Meta/Fielddumps/filename" + std::to_string(k + 1), dl.field_file[k]);
    ok &= f.write_int("Meta/Beamdumps/ndumps", {(int)dl.beam_step.size()});
    if (!dl.beam_step.empty()) ok &= f.write_int("Meta/Beamdumps/intstep", dl.beam_step);
    for (size_t k = 0; k < dl.beam_file.size(); k++)
        ok &= f.write_string("Meta/Beamdumps/filename" + std::to_string(k + 1), dl.beam_file[k]);
    ok &= f.write_single("Meta/mpisize", " ", {1.0});
    // Lattice: per-step arrays and the output positions
    auto col = [&](double g4mi_step_params::*m) {
        std::vector<double> v(steps.size());
        for (size_t i = 0; i < steps.size(); i++) v[i] = steps[i].*m;
        return v;
    };
    ok &= f.write_single("Lattice/zplot", "m", zout);
    ok &= f.write_single("Lattice/z", "m", col(&g4mi_step_params::z));
    ok &= f.write_single("Lattice/dz", "m", col(&g4mi_step_params::dz));
    ok &= f.write_single("Lattice/aw", " ", col(&g4mi_step_params::aw));
    ok &= f.write_single("Lattice/ax", "m", col(&g4mi_step_params::ax));
    ok &= f.write_single("Lattice/ay", "m", col(&g4mi_step_params::ay));
    ok &= f.write_single("Lattice/ku", "m^{-1}", col(&g4mi_step_params::ku));
    ok &= f.write_single("Lattice/kx", " ", col(&g4mi_step_params::kx));
    ok &= f.write_single("Lattice/ky", " ", col(&g4mi_step_params::ky));
    ok &= f.write_single("Lattice/qf", "m^{-2}", col(&g4mi_step_params::qf));
    ok &= f.write_single("Lattice/qx", "m", col(&g4mi_step_params::qx));
    ok &= f.write_single("Lattice/qy", "m", col(&g4mi_step_params::qy));
    ok &= f.write_single("Lattice/cx", "rad", col(&g4mi_step_params::cx));
    ok &= f.write_single("Lattice/cy", "rad", col(&g4mi_step_params::cy));
    ok &= f.write_single("Lattice/gradx", "m^{-1}", col(&g4mi_step_params::gradx));
    ok &= f.write_single("Lattice/grady", "m^{-1}", col(&g4mi_step_params::grady));
    ok &= f.write_single("Lattice/slippage", " ", col(&g4mi_step_params::slip));
    ok &= f.write_single("Lattice/phaseshift", "rad", col(&g4mi_step_params::phaseshift));
    ok &= f.write_single("Lattice/chic_angle", "degree", col(&g4mi_step_params::chic_angle));
    ok &= f.write_single("Lattice/chic_lb", "m", col(&g4mi_step_params::chic_lb));
    ok &= f.write_single("Lattice/chic_ld", "m", col(&g4mi_step_params::chic_ld));
    ok &= f.write_single("Lattice/chic_lt", "m", col(&g4mi_step_params::chic_lt));
    // Global
    const double slicelength = st.lambda0 * dk.time.sample;
    ok &= f.write_single("Global/gamma0", "mc^2", {st.gamma0});
    ok &= f.write_single("Global/lambdaref", "m", {st.lambda0});
    ok &= f.write_single("Global/sample", " ", {slicelength / st.lambda0});
    ok &= f.write_single("Global/slen", "m", {slicelength * ns});
    ok &= f.write_single("Global/one4one", " ", {dk.setup.one4one ? 1.0 : 0.0});
    ok &= f.write_single("Global/time", " ", {dk.time.dotime ? 1.0 : 0.0});
    ok &= f.write_single("Global/scan", " ", {0.0});
    {
        std::vector<double> s(ns), fr(ns);
        const double e0 = 1239.842e-9 / st.lambda0;
        double df = e0 * st.lambda0 / slicelength / (double)ns;
        if (ns == 1) df = 0;
        for (int i = 0; i < ns; i++) {
            s[i] = (double)i * slicelength;
            fr[i] = e0 + (double)i * df - 0.5 * df * (double)ns;
        }
        ok &= f.write_single("Global/s", "m", s);
        ok &= f.write_single("Global/frequency", "eV", fr);
    }
    // Beam and Field: [nrec][ns] records; once-values [1][ns]
    auto put = [&](const char *path, const char *unit, const char *key) {
        auto it = rec.find(key);
        if (it != rec.end()) ok &= f.write_buffer(path, unit, it->second, nrec, (size_t)ns);
    };
    if (!st.excl_energy) {
        put("Beam/energy", "mc^2", "energy");
        put("Beam/energyspread", "mc^2", "energyspread");
    }
    if (!st.excl_spatial) {
        put("Beam/xposition", "m", "xposition");
        put("Beam/xsize", "m", "xsize");
        put("Beam/yposition", "m", "yposition");
        put("Beam/ysize", "m", "ysize");
        put("Beam/pxposition", "rad", "pxposition");
        put("Beam/pyposition", "rad", "pyposition");
    }
    put("Beam/bunching", " ", "bunching");
    put("Beam/bunchingphase", "rad", "bunchingphase");
    for (const auto &kv : rec) {  // bunching<h>, bunchingphase<h> (&track bunchharm)
        const std::string &key = kv.first;
        const bool phase = key.rfind("bunchingphase", 0) == 0;
        const size_t n = phase ? 13 : 8;  // length of the name before the harmonic number
        if ((phase || key.rfind("bunching", 0) == 0) && key.size() > n && isdigit((unsigned char)key[n]))
            put(("Beam/" + key).c_str(), phase ? "rad" : " ", key.c_str());
    }
    if (!st.excl_aux) {
        put("Beam/efield", "eV/m", "efield");
        put("Beam/wakefield", "eV/m", "wakefield");
        put("Beam/LSCfield", "eV/m", "lscfield");
        put("Beam/SSCfield", "eV/m", "sscfield");
        put("Beam/xmin", "m", "xmin");
        put("Beam/xmax", "m", "xmax");
        put("Beam/pxmin", "rad", "pxmin");
        put("Beam/pxmax", "rad", "pxmax");
        put("Beam/ymin", "m", "ymin");
        put("Beam/ymax", "m", "ymax");
        put("Beam/pymin", "rad", "pymin");
        put("Beam/pymax", "rad", "pymax");
        put("Beam/emin", "mc^2", "emin");
        put("Beam/emax", "mc^2", "emax");
    }
    // global statistics: only when asked for and when there is more than one
    // slice (Track.cpp: a steady-state run has none)
    auto put1 = [&](const char *path, const char *unit, const char *key) {
        auto it = rec.find(key);
        if (it != rec.end() && nrec > 0) ok &= f.write_single(path, unit, it->second);
    };
    if (st.beam_global_stat && ns > 1) {
        if (!st.excl_energy) {
            put1("Beam/Global/energy", "mc^2", "g_energy");
            put1("Beam/Global/energyspread", "mc^2", "g_energyspread");
        }
        if (!st.excl_spatial) {
            put1("Beam/Global/xposition", "m", "g_xposition");
            put1("Beam/Global/xsize", "m", "g_xsize");
            put1("Beam/Global/yposition", "m", "g_yposition");
            put1("Beam/Global/ysize", "m", "g_ysize");
        }
    }
    ok &= f.write_buffer("Beam/current", "A", dk.current, 1, (size_t)ns);
    {
        std::vector<double> tw[6];
        twiss_once(dk, tw);
        const char *names[6] = {"Beam/emitx", "Beam/emity", "Beam/betax", "Beam/betay", "Beam/alphax", "Beam/alphay"};
        const char *units[6] = {"m", "m", "m", "m", "rad", "rad"};
        for (int q = 0; q < 6; q++) ok &= f.write_buffer(names[q], units[q], tw[q], 1, (size_t)ns);
    }
    put("Field/power", "W", "power");
    if (!st.excl_spatial) {
        put("Field/xposition", "m", "f_xposition");
        put("Field/xsize", "m", "f_xsize");
        put("Field/yposition", "m", "f_yposition");
        put("Field/ysize", "m", "f_ysize");
    }
    if (!st.excl_intensity) {
        put("Field/intensity-nearfield", "W/m^2", "nf_intensity");
        put("Field/phase-nearfield", "rad", "nf_phase");
        put("Field/intensity-farfield", "W/rad^2", "ff_intensity");
        put("Field/phase-farfield", "rad", "ff_phase");
    }
    if (!st.excl_fft) {
        put("Field/xdivergence", "rad", "xdivergence");
        put("Field/ydivergence", "rad", "ydivergence");
        put("Field/xpointing", "rad", "xpointing");
        put("Field/ypointing", "rad", "ypointing");
    }
    if (st.field_global_stat && ns > 1) {
        put1("Field/Global/energy", "J", "global_energy");
        if (!st.excl_fft) {
            put1("Field/Global/xdivergence", "rad", "g_xdivergence");
            put1("Field/Global/ydivergence", "rad", "g_ydivergence");
            put1("Field/Global/xpointing", "rad", "g_xpointing");
            put1("Field/Global/ypointing", "rad", "g_ypointing");
        }
        if (!st.excl_spatial) {
            put1("Field/Global/xposition", "m", "g_f_xposition");
            put1("Field/Global/xsize", "m", "g_f_xsize");
            put1("Field/Global/yposition", "m", "g_f_yposition");
            put1("Field/Global/ysize", "m", "g_f_ysize");
        }
        if (!st.excl_intensity) {
            put1("Field/Global/intensity-nearfield", "W/m^2", "g_nf_intensity");
            put1("Field/Global/intensity-farfield", "W/rad^2", "g_ff_intensity");
        }
    }
    const double dgrid = 2 * dk.gridmax / (dk.ngrid - 1);
    ok &= f.write_single("Field/gridspacing", "m", {dgrid});
    ok &= f.write_single("Field/dgrid", "m", {dk.gridmax});
    ok &= f.write_single("Field/ngrid", " ", {(double)dk.ngrid});
    // the harmonic fields as groups Field<h> (Diagnostic::writeToOutputFile, Diagnostic.cpp:219-225)
    for (size_t k = 0; k < dk.hfields.size(); k++) {
        const auto &h = dk.hfields[k];
        const std::string g = "Field" + std::to_string(h.harm) + "/", pre = "H" + std::to_string(k) + ":";
        auto hput = [&](const char *path, const char *unit, const char *key) {
            auto it = rec.find(pre + key);
            if (it != rec.end()) ok &= f.write_buffer(g + path, unit, it->second, nrec, (size_t)ns);
        };
        auto hput1 = [&](const char *path, const char *unit, const char *key) {
            auto it = rec.find(pre + key);
            if (it != rec.end() && nrec > 0) ok &= f.write_single(g + path, unit, it->second);
        };
        hput("power", "W", "power");
        if (!st.excl_spatial) {
            hput("xposition", "m", "f_xposition");
            hput("xsize", "m", "f_xsize");
            hput("yposition", "m", "f_yposition");
            hput("ysize", "m", "f_ysize");
        }
        if (!st.excl_intensity) {
            hput("intensity-nearfield", "W/m^2", "nf_intensity");
            hput("intensity-farfield", "W/rad^2", "ff_intensity");
            hput("phase-farfield", "rad", "ff_phase");
        }
        if (st.field_global_stat && ns > 1) {
            hput1("Global/energy", "J", "global_energy");
            if (!st.excl_spatial) {
                hput1("Global/xposition", "m", "g_f_xposition");
                hput1("Global/xsize", "m", "g_f_xsize");
                hput1("Global/yposition", "m", "g_f_yposition");
                hput1("Global/ysize", "m", "g_f_ysize");
            }
            if (!st.excl_intensity) {
                hput1("Global/intensity-nearfield", "W/m^2", "g_nf_intensity");
                hput1("Global/intensity-farfield", "W/rad^2", "g_ff_intensity");
            }
        }
        const double hd = 2 * h.gridmax / (h.ngrid - 1);
        ok &= f.write_single(g + "gridspacing", "m", {hd});
        ok &= f.write_single(g + "dgrid", "m", {h.gridmax});
        ok &= f.write_single(g + "ngrid", " ", {(double)h.ngrid});
    }
    f.close();
    if (!ok) err = "writing " + fname + " failed";
    return ok;
}

const char *kRecKeys[] = {"power", "bunching", "bunchingphase", "energy", "energyspread", "xposition",
                          "yposition", "xsize", "ysize", "f_xposition", "f_yposition", "f_xsize",
                          "f_ysize", "nf_intensity", "ff_intensity", "pxposition", "pyposition",
                          "nf_phase", "ff_phase", "efield", "wakefield", "lscfield", "sscfield", "xmin",
                          "xmax", "pxmin", "pxmax", "ymin", "ymax", "pymin", "pymax", "emin", "emax",
                          "xdivergence", "ydivergence", "xpointing", "ypointing"};
// one value per record
const char *kRecScalars[] = {"global_energy", "g_energy", "g_energyspread", "g_xposition", "g_xsize",
                             "g_yposition", "g_ysize", "g_f_xposition", "g_f_xsize", "g_f_yposition",
                             "g_f_ysize", "g_nf_intensity", "g_ff_intensity", "g_xdivergence",
                             "g_ydivergence", "g_xpointing", "g_ypointing"};

bool do_track(Deck &dk, Namelist &nl, int device, bool quiet, bool dry, std::string &err) {
    Keys k{nl, ""};
    double zstop = 1e9, s0 = dk.time.s0, slen = dk.time.slen;
    int output_step = 1, fdump = 0, bdump = 0, sort_step = 0, bunchharm = 1;
    bool dump_ue = false, fft = false, exclharm = false, periodic = false, dbg1 = false, dbg2 = false,
         filt = false;
    double xc = 1, yc = 1, sig = 1;
    std::string prec = "fp64";
    k.dbl("zstop", zstop);
    k.dbl("s0", s0);
    k.dbl("slen", slen);
    k.integer("output_step", output_step);
    k.integer("field_dump_step", fdump);
    k.boolean("field_dump_at_undexit", dump_ue);
    k.integer("beam_dump_step", bdump);
    k.integer("sort_step", sort_step);
    k.integer("bunchharm", bunchharm);
    k.boolean("fft_fieldsolver", fft);
    k.boolean("exclusive_harmonics", exclharm);
    k.boolean("periodic", periodic);
    k.boolean("dbg_report_lattice", dbg1);
    k.boolean("dbg_suppress_outfile", dbg2);
    k.dbl("xcut", xc);
    k.dbl("ycut", yc);
    k.dbl("sigmoid", sig);
    k.boolean("source_filter", filt);
    k.str("precision", prec);  // this build's extension: fp64 (default) or fp32
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (fft && prec == "fp32") {
        err = "&track: fft_fieldsolver runs in the FP64 form only (precision = fp64)";
        return false;
    }
    if (bunchharm > 1 && prec == "fp32") {
        err = "&track: bunching harmonics (bunchharm > 1) run in the FP64 form only (precision = fp64)";
        return false;
    }
    if (prec != "fp64" && prec != "fp32") {
        err = "&track: precision must be fp64 or fp32";
        return false;
    }
    if (!dk.setup.defined || !dk.have_beam || !dk.have_field) {
        err = "&track needs &setup, &beam and &field first";
        return false;
    }
    if (output_step < 1) output_step = 1;
    const SetupNL &st = dk.setup;
    LatticeResult lr = parse_and_unroll(dk.lattice_text, st.beamline, st.delz, st.lambda0, st.gamma0);
    if (!lr.error.empty()) {
        err = "lattice: " + lr.error;
        return false;
    }
    std::vector<g4mi_step_params> steps = lr.steps;
    // Undulator::updateMarker (src/Core/Undulator.cpp:33-64) and markUndExits:
    // dump bits at positions 0..nsteps (before step i; nsteps = after the last)
    const size_t nst = steps.size();
    std::vector<int> dflag(nst + 1, 0);
    for (size_t i = 0; i <= nst; i++) {
        if (i < nst) {
            if (sort_step > 0 && i % (size_t)sort_step == 0) steps[i].marker |= 4;
            if (steps[i].z > zstop) steps[i].marker |= 8;
            dflag[i] = steps[i].marker & 3;
        }
        if (fdump > 0 && i % (size_t)fdump == 0) dflag[i] |= 1;
        if (bdump > 0 && i % (size_t)bdump == 0) dflag[i] |= 2;
        if (dump_ue && i > 0 && i < nst && steps[i - 1].aw != 0 && steps[i].aw == 0) dflag[i] |= 1;
    }
    // steps that run (up to the first stop marker) and the output records
    // (Undulator::updateOutput: the initial state, then after step i (1-based)
    // when i %% output_step == 0 and step i carries no stop marker)
    size_t nrun = 0;
    while (nrun < steps.size() && !(steps[nrun].marker & 8)) nrun++;
    std::vector<size_t> outs;
    for (size_t i = 1; i <= nrun; i++)
        if (i % (size_t)output_step == 0) outs.push_back(i);
    const int ns = (int)dk.off.size() - 1;
    const size_t nrec = outs.size() + 1;
    std::vector<double> zout(nrec, 0.0);
    for (size_t r = 0; r < outs.size(); r++) zout[r + 1] = steps[outs[r] - 1].z + steps[outs[r] - 1].dz;
    char tmp[512];
    snprintf(tmp, sizeof(tmp),
             "{\"nslice\": %d, \"npart\": %lld, \"ngrid\": %d, \"gridmax\": %.17g, \"steps\": %zu, \"run_steps\": "
             "%zu, \"records\": %zu, \"matched\": %s, \"betax\": %.17g, \"alphax\": %.17g, \"phix\": %.17g, "
             "\"betay\": %.17g, \"alphay\": %.17g, \"phiy\": %.17g, \"time\": %s}",
             ns, (long long)dk.off[ns], dk.ngrid, dk.gridmax, steps.size(), nrun, nrec, dk.matched ? "true" : "false",
             dk.mbetax, dk.malphax, dk.mphix, dk.mbetay, dk.malphay, dk.mphiy, dk.time.dotime ? "true" : "false");
    dk.summary = tmp;
    std::string root = st.rootname;
    if (!st.outputdir.empty()) root = st.outputdir + "/" + root;
    if (dk.runcount > 0) root += ".Run" + std::to_string(dk.runcount + 1);
    const std::string fname = root + ".out.h5";

    std::map<std::string, std::vector<double>> rec;
    DumpList dl;
    if (dry) {
        // setup check: the output file without device records
        if (st.rootname.empty()) return true;
        return write_output(dk, fname, steps, zout, rec, 0, dl, err);
    }
    // device context: created by the first &track, reused by later ones
    const int precision = prec == "fp32" ? G4MI_FP32 : G4MI_FP64;
    const bool created = !dk.ctx;
    if (!dk.ctx) {
        dk.precision = precision;
        dk.periodic = periodic;
        int rc = g4mi_create(&dk.ctx, device, ns, dk.off.data(), dk.ngrid, dk.gridmax, st.lambda0,
                             (double)dk.time.sample, dk.time.dotime ? 1 : 0, periodic ? 1 : 0, precision);
        if (rc != G4MI_OK) {
            err = std::string("device context: ") + g4mi_last_error();
            return false;
        }
        const EFieldNL &e = dk.efield;
        if ((rc = g4mi_set_efield(dk.ctx, e.longrange ? 1 : 0, e.nz, e.nphi, e.ngrid, e.rmax)) != G4MI_OK ||
            (rc = g4mi_beam_upload(dk.ctx, dk.gamma.data(), dk.theta.data(), dk.x.data(), dk.y.data(),
                                   dk.px.data(), dk.py.data(), dk.current.data())) != G4MI_OK ||
            (rc = g4mi_field_upload(dk.ctx, dk.field.data(), dk.field_first)) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
        for (const auto &h : dk.hfields) {  // Field records at harmonics (LoadField with harm)
            int idx = -1;
            if (g4mi_add_harmonic(dk.ctx, h.harm, h.ngrid, h.gridmax, &idx) != G4MI_OK ||
                g4mi_harmonic_upload(dk.ctx, idx, h.field.data()) != G4MI_OK) {
                err = g4mi_last_error();
                return false;
            }
        }
    } else {
        if (precision != dk.precision || periodic != dk.periodic) {
            err = "&track: precision and periodic cannot change between &track elements";
            return false;
        }
        // the once-values of the output describe the beam at the start of
        // this run (slices as the last sort left them)
        int first = 0;
        if (!host_state(dk, &first, err)) return false;
    }
    // the incoherent stream is seeded once (it runs on across &track
    // elements); the wake loss follows the current beam
    if (!dry && dk.spon.set && (!dk.spon_applied || created)) {
        if (g4mi_set_incoherent(dk.ctx, dk.spon.seed, dk.setup.one4one ? 1 : dk.setup.nbins,
                                dk.spon.loss ? 1 : 0, dk.spon.spread ? 1 : 0) != G4MI_OK ||
            (dk.spon_applied && dk.incoh_saved && g4mi_incoherent_state(dk.ctx, dk.incoh_state, 1) != G4MI_OK)) {
            err = g4mi_last_error();
            return false;
        }
        dk.spon_applied = true;
        dk.incoh_saved = false;
    }
    // Beam::initSorting: one4one beams are sorted at the sort markers
    if (g4mi_set_sorting(dk.ctx, dk.setup.one4one ? 1 : 0) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    if (dk.wake.set && dk.setup.one4one) {
        err = "&wake with one4one beams is not supported by this build (the loss would follow every sort)";
        return false;
    }
    if (dk.wake.set) {
        const WakeNL &w = dk.wake;
        std::vector<double> eloss(ns), ext(ns, w.loss);
        if (g4mi_wake_loss(ns, dk.time.sample, dk.setup.lambda0, dk.current.data(), ext.data(), w.radius,
                           w.roundpipe ? 1 : 0, w.conductivity, w.relaxation, w.gap, w.lgap, w.hrough, w.lrough,
                           eloss.data(), nullptr, nullptr, nullptr) != G4MI_OK ||
            g4mi_set_wake(dk.ctx, eloss.data()) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
    }
    // Track::init -> Field::initSolver per &track (src/Main/Track.cpp:107-112)
    if (g4mi_set_field_solver(dk.ctx, fft ? 1 : 0, filt ? 1 : 0, xc, yc, sig) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    for (const char *key : kRecKeys) rec[key].assign(nrec * (size_t)ns, 0.0);
    // DiagBeam (Diagnostic.cpp:486-500): bunching<h> for h = 2..bunchharm, or
    // only h = bunchharm with exclusive_harmonics
    std::vector<int> bharm;
    for (int h = 2; h <= bunchharm; h++)
        if (!exclharm || h == bunchharm) bharm.push_back(h);
    for (int h : bharm) {
        rec["bunching" + std::to_string(h)].assign(nrec * (size_t)ns, 0.0);
        rec["bunchingphase" + std::to_string(h)].assign(nrec * (size_t)ns, 0.0);
    }
    std::vector<double> hbun, hpha;
    for (const char *key : kRecScalars) rec[key].assign(nrec, 0.0);
    for (size_t k = 0; k < dk.hfields.size(); k++) {
        const std::string pre = "H" + std::to_string(k) + ":";
        for (const char *key : {"power", "f_xposition", "f_xsize", "f_yposition", "f_ysize", "nf_intensity",
                                "ff_intensity", "ff_phase"})
            rec[pre + key].assign(nrec * (size_t)ns, 0.0);
        for (const char *key : {"global_energy", "g_f_xposition", "g_f_xsize", "g_f_yposition", "g_f_ysize",
                                "g_nf_intensity", "g_ff_intensity"})
            rec[pre + key].assign(nrec, 0.0);
    }
    auto make_diag = [&](size_t r) {
        g4mi_diag_out d{};
        auto ps = [&](const char *k) { return rec[k].data() + r * ns; };
        auto p1 = [&](const char *k) { return rec[k].data() + r; };
        d.pxposition = ps("pxposition");
        d.pyposition = ps("pyposition");
        d.nf_phase = ps("nf_phase");
        d.ff_phase = ps("ff_phase");
        if (!st.excl_aux) {
            d.efield = ps("efield");
            d.wakefield = ps("wakefield");
            d.lscfield = ps("lscfield");
            d.sscfield = ps("sscfield");
            d.xmin = ps("xmin");
            d.xmax = ps("xmax");
            d.pxmin = ps("pxmin");
            d.pxmax = ps("pxmax");
            d.ymin = ps("ymin");
            d.ymax = ps("ymax");
            d.pymin = ps("pymin");
            d.pymax = ps("pymax");
            d.emin = ps("emin");
            d.emax = ps("emax");
        }
        d.g_energy = p1("g_energy");
        d.g_energyspread = p1("g_energyspread");
        d.g_xposition = p1("g_xposition");
        d.g_xsize = p1("g_xsize");
        d.g_yposition = p1("g_yposition");
        d.g_ysize = p1("g_ysize");
        d.g_f_xposition = p1("g_f_xposition");
        d.g_f_xsize = p1("g_f_xsize");
        d.g_f_yposition = p1("g_f_yposition");
        d.g_f_ysize = p1("g_f_ysize");
        d.g_nf_intensity = p1("g_nf_intensity");
        d.g_ff_intensity = p1("g_ff_intensity");
        if (!st.excl_fft) {  // DiagField's spectral block (filter "fft")
            d.xdivergence = ps("xdivergence");
            d.ydivergence = ps("ydivergence");
            d.xpointing = ps("xpointing");
            d.ypointing = ps("ypointing");
            d.g_xdivergence = p1("g_xdivergence");
            d.g_ydivergence = p1("g_ydivergence");
            d.g_xpointing = p1("g_xpointing");
            d.g_ypointing = p1("g_ypointing");
        }
        d.power = rec["power"].data() + r * ns;
        d.bunching = rec["bunching"].data() + r * ns;
        d.bunchingphase = rec["bunchingphase"].data() + r * ns;
        d.energy = rec["energy"].data() + r * ns;
        d.energyspread = rec["energyspread"].data() + r * ns;
        d.xposition = rec["xposition"].data() + r * ns;
        d.yposition = rec["yposition"].data() + r * ns;
        d.xsize = rec["xsize"].data() + r * ns;
        d.ysize = rec["ysize"].data() + r * ns;
        d.f_xposition = rec["f_xposition"].data() + r * ns;
        d.f_yposition = rec["f_yposition"].data() + r * ns;
        d.f_xsize = rec["f_xsize"].data() + r * ns;
        d.f_ysize = rec["f_ysize"].data() + r * ns;
        d.nf_intensity = rec["nf_intensity"].data() + r * ns;
        d.ff_intensity = rec["ff_intensity"].data() + r * ns;
        d.global_energy = rec["global_energy"].data() + r;
        return d;
    };
    auto record = [&](size_t r) {
        g4mi_diag_out d = make_diag(r);
        int rc = g4mi_diag(dk.ctx, &d);
        if (rc == G4MI_OK && !bharm.empty()) {
            hbun.resize((size_t)(bunchharm - 1) * ns);
            hpha.resize(hbun.size());
            rc = g4mi_diag_bunching(dk.ctx, bunchharm, hbun.data(), hpha.data());
            for (int h : bharm) {
                std::copy(hbun.begin() + (size_t)(h - 2) * ns, hbun.begin() + (size_t)(h - 1) * ns,
                          rec["bunching" + std::to_string(h)].begin() + r * ns);
                std::copy(hpha.begin() + (size_t)(h - 2) * ns, hpha.begin() + (size_t)(h - 1) * ns,
                          rec["bunchingphase" + std::to_string(h)].begin() + r * ns);
            }
        }
        for (size_t k = 0; rc == G4MI_OK && k < dk.hfields.size(); k++) {  // DiagField of every harmonic
            const std::string pre = "H" + std::to_string(k) + ":";
            g4mi_diag_out dh{};
            auto hs = [&](const char *key) { return rec[pre + key].data() + r * ns; };
            auto h1 = [&](const char *key) { return rec[pre + key].data() + r; };
            dh.power = hs("power");
            dh.f_xposition = hs("f_xposition");
            dh.f_xsize = hs("f_xsize");
            dh.f_yposition = hs("f_yposition");
            dh.f_ysize = hs("f_ysize");
            dh.nf_intensity = hs("nf_intensity");
            dh.ff_intensity = hs("ff_intensity");
            dh.ff_phase = hs("ff_phase");
            dh.global_energy = h1("global_energy");
            dh.g_f_xposition = h1("g_f_xposition");
            dh.g_f_xsize = h1("g_f_xsize");
            dh.g_f_yposition = h1("g_f_yposition");
            dh.g_f_ysize = h1("g_f_ysize");
            dh.g_nf_intensity = h1("g_nf_intensity");
            dh.g_ff_intensity = h1("g_ff_intensity");
            rc = g4mi_diag_harmonic(dk.ctx, (int)k, &dh);
        }
        return rc;
    };
    if (record(0) != G4MI_OK) {
        err = g4mi_last_error();
        return false;
    }
    if (!quiet) printf("Running %zu steps on the device (%d slices, %lld macroparticles)...\n", nrun, ns,
                       (long long)dk.off[ns]);
    // Control::applyMarker's dumps at position p (before step p, file
    // <root>.<p>), the output records after step p; positions 0..nrun
    auto dump_at = [&](size_t p) {
        if (!dflag[p]) return true;
        int first = 0;
        if (!host_state(dk, &first, err)) return false;
        const std::string base = root + "." + std::to_string(p);
        if (dflag[p] & 1) {
            if (!quiet) printf("Writing field distribution to file: %s.fld.h5 ...\n", base.c_str());
            if (!write_field_dump(dk, base, first, err)) return false;
            dl.field_step.push_back((int)p);
            dl.field_file.push_back(base + ".fld.h5");
        }
        if (dflag[p] & 2) {
            if (!quiet) printf("Writing particle distribution to file: %s.par.h5 ...\n", base.c_str());
            if (!write_beam_dump(dk, base, 1, err)) return false;
            dl.beam_step.push_back((int)p);
            dl.beam_file.push_back(base + ".par.h5");
        }
        return true;
    };
    // FP64 runs whose records need nothing beyond g4mi_diag's blocks hand a
    // stretch of steps that each end in a record to one g4mi_run call, so the
    // records are deferred on the device instead of taken one by one
    const bool batch = precision == G4MI_FP64 && bharm.empty() && dk.hfields.empty() && !dk.setup.one4one;
    std::vector<g4mi_diag_out> dbatch;
    size_t done = 0, r = 0;
    for (size_t p = 0; p <= nrun; p++) {
        const bool out_here = r < outs.size() && outs[r] == p;
        if (batch && out_here && p == done + 1 && !dflag[p]) {
            // the longest stretch p, p+1, ... of record steps without dumps
            size_t q = p;
            dbatch.clear();
            while (q <= nrun && r + dbatch.size() < outs.size() && outs[r + dbatch.size()] == q && !dflag[q]) {
                dbatch.push_back(make_diag(r + dbatch.size() + 1));
                q++;
            }
            const int n = (int)dbatch.size();
            if (g4mi_run(dk.ctx, steps.data() + done, n, st.gamma0, dbatch.data()) != n) {
                err = g4mi_last_error();
                return false;
            }
            done += (size_t)n;
            r += (size_t)n;
            p = done;
            continue;
        }
        if (!out_here && !dflag[p]) continue;
        if (p > done) {
            if (g4mi_run(dk.ctx, steps.data() + done, (int)(p - done), st.gamma0, nullptr) < 0) {
                err = g4mi_last_error();
                return false;
            }
            done = p;
        }
        if (out_here) {
            if (record(++r) != G4MI_OK) {
                err = g4mi_last_error();
                return false;
            }
        }
        if (!dump_at(p)) return false;
    }
    if (done < nrun && g4mi_run(dk.ctx, steps.data() + done, (int)(nrun - done), st.gamma0, nullptr) < 0) {
        err = g4mi_last_error();
        return false;
    }
    if (st.rootname.empty()) {
        err = "no rootname in &setup";
        return false;
    }
    if (!quiet) printf("Writing output file %s...\n", fname.c_str());
    return write_output(dk, fname, steps, zout, rec, nrec, dl, err);
}

// Time::finishInit (src/Main/GenTime.cpp:64-86) after a change of the
// reference length or sample rate, on one rank
void time_refresh(Deck &dk) {
    TimeNL &t = dk.time;
    if (!(t.ds > 0)) return;  // no &time: the window is not initialised
    t.ds = dk.setup.lambda0 * t.sample;
    t.nslice = std::max(1, (int)std::round(t.slen / t.ds));
    t.slen = t.ds * t.nslice;
}

// AlterSetup.cpp:136-170 with Beam::subharmonicConversion (Beam.cpp:243-299)
// and Field::subharmonicConversion (Field.cpp:188-258): the reference
// length grows by h, theta shrinks by h; resampling (one4one) merges h
// slices into one (all particles into slice 0 at theta + i dtheta, then
// sorted); harmonic fields are deleted and the fundamental becomes
// harmonic h of the new reference, its slices merged in groups of h with
// the group's mean power.
bool subharmonic_conversion(Deck &dk, int h, bool resample, int device, bool dry, std::string &err) {
    TimeNL &t = dk.time;
    const double lam_old = dk.setup.lambda0;
    if (!resample && t.dotime && t.sample % h != 0) {
        err = "subharmonic conversion without resampling needs a sample rate that is a multiple of " +
              std::to_string(h);
        return false;
    }
    dk.setup.lambda0 = lam_old * h;
    if (!resample && t.dotime) t.sample = t.sample / h;
    time_refresh(dk);
    const double hd = (double)h;
    if (dk.have_beam) {
        const int ns = (int)dk.off.size() - 1;
        if (resample && !dk.setup.one4one) {
            err = "Cannot convert beam distribution to lower harmonic (resampling needs one4one)";
            return false;
        }
        if (resample && ns % h != 0) {
            err = "Cannot convert beam distribution to lower harmonic (" + std::to_string(ns) +
                  " slices cannot be merged in groups of " + std::to_string(h) + ")";
            return false;
        }
        for (double &v : dk.theta) v /= hd;
        if (resample) {
            // the beam's slicelength and reflength after the conversion
            const double slicelength = lam_old * (double)t.sample * hd, reflength = lam_old * hd;
            const double dtheta = 4. * asin(1) * slicelength / reflength / hd;
            for (int i = 1; i < ns; i++)
                for (long long j = dk.off[i]; j < dk.off[i + 1]; j++) dk.theta[j] += i * dtheta;
            const long long n = dk.off[ns];
            dk.off.assign((size_t)ns / h + 1, n);
            dk.off[0] = 0;
            dk.current.assign((size_t)ns / h, 0.0);
            if (!dry && !device_sort(dk, device, err)) return false;
        }
    }
    dk.hfields.clear();  // AlterSetup.cpp:151-156
    if (dk.have_field) {
        Deck::HarmField f;
        f.harm = h;
        f.ngrid = dk.ngrid;
        f.gridmax = dk.gridmax;
        f.field.swap(dk.field);
        const size_t cells = (size_t)f.ngrid * f.ngrid, per = 2 * cells;
        const int nsize = (int)(f.field.size() / per);
        if (resample && h > 1) {
            if (nsize % h != 0) {
                err = "Cannot convert field distribution to lower harmonic";
                return false;
            }
            std::vector<double> pw(nsize, 0.0);
            for (int i = 0; i < nsize; i++)
                for (size_t k = 0; k < cells; k++) {
                    const double re = f.field[i * per + 2 * k], im = f.field[i * per + 2 * k + 1];
                    pw[i] += re * re + im * im;
                }
            const int first = dk.field_first;
            for (int i = 0; i < nsize; i += h) {  // merge each group into its first slice
                const int is0 = (i + first) % nsize;
                for (int j = 1; j < h; j++) {
                    const int is1 = (i + j + first) % nsize;
                    for (size_t q = 0; q < per; q++) f.field[is0 * per + q] += f.field[is1 * per + q];
                    pw[is0] += pw[is1];
                }
            }
            const double scl = 1. / hd;
            const int di = first % h;
            for (int i = 0; i < nsize / h; i++) {
                if (i != i * h + di)
                    std::memcpy(f.field.data() + i * per, f.field.data() + (size_t)(i * h + di) * per,
                                per * sizeof(double));
                pw[i] = pw[i * h + di] * scl;
            }
            dk.field_first = (first - di) / h;
            f.field.resize((size_t)(nsize / h) * per);
            for (int i = 0; i < nsize / h; i++) {
                double pl = 0;
                for (size_t k = 0; k < cells; k++) {
                    const double re = f.field[i * per + 2 * k], im = f.field[i * per + 2 * k + 1];
                    pl += re * re + im * im;
                }
                if (pw[i] > 0) {
                    const double a = sqrt(pw[i] / pl);
                    for (size_t q = 0; q < per; q++) f.field[i * per + q] *= a;
                }
            }
        }
        dk.hfields.push_back(std::move(f));
        dk.field.clear();
        dk.have_field = false;  // a fundamental of the new reference needs a &field
    }
    if (resample) dk.wake.set = false;  // Beam::subharmonicConversion clears the wakes
    return true;
}

// AlterSetup.cpp:172-209 with Beam::harmonicConversion (Beam.cpp:187-241)
// and Field::harmonicConversion (Field.cpp:261-288): the reference length
// shrinks by h and theta grows by h; resampling (one4one) moves slice i to
// slice h i and sorts; the field at harmonic h becomes the fundamental
// (each slice copied into h slices when resampling) and every other field
// is deleted.
bool harmonic_conversion(Deck &dk, int h, bool resample, int device, bool dry, std::string &err) {
    TimeNL &t = dk.time;
    const double hd = (double)h;
    dk.setup.lambda0 = dk.setup.lambda0 / hd;
    if (!resample && t.dotime) t.sample = t.sample * h;
    time_refresh(dk);
    if (dk.have_beam) {
        if (resample && !dk.setup.one4one) {
            err = "Cannot convert beam distribution to higher harmonic (resampling needs one4one)";
            return false;
        }
        for (double &v : dk.theta) v *= hd;
        if (resample) {
            const int ns = (int)dk.off.size() - 1;
            std::vector<long long> off((size_t)ns * h + 1);
            for (int s = 0; s <= ns * h; s++) off[s] = dk.off[(s + h - 1) / h];
            dk.off.swap(off);
            dk.current.assign((size_t)ns * h, 0.0);
            if (!dry && !device_sort(dk, device, err)) return false;
        }
    }
    Deck::HarmField *match = nullptr;
    for (auto &e : dk.hfields)
        if (e.harm == h) match = &e;
    if (match) {
        dk.field.swap(match->field);
        dk.ngrid = match->ngrid;
        dk.gridmax = match->gridmax;
        dk.flambda = dk.setup.lambda0;  // Field::init: lambda0 / harm of the old reference
        dk.have_field = true;
        if (resample) {
            const size_t per = 2 * (size_t)dk.ngrid * dk.ngrid;
            const int nloc = (int)(dk.field.size() / per);
            std::vector<double> f((size_t)nloc * h * per);
            for (int i = nloc; i > 0; i--)
                for (int j = 0; j < h; j++)
                    std::memcpy(f.data() + (size_t)(i * h - 1 - j) * per, dk.field.data() + (size_t)(i - 1) * per,
                                per * sizeof(double));
            dk.field.swap(f);
            dk.field_first *= h;
        }
    } else {
        dk.field.clear();
        dk.have_field = false;
        dk.field_first = 0;
    }
    dk.hfields.clear();
    if (resample) dk.wake.set = false;  // Beam::harmonicConversion clears the wakes
    return true;
}

// &alter_setup (AlterSetup::init, src/Main/AlterSetup.cpp:37-220): a new
// beamline, step length or rootname (else .Run<n> on the output names), then
// the subharmonic and harmonic conversions of beam and fields.
bool do_alter_setup(Deck &dk, Namelist &nl, int device, bool dry, std::string &err) {
    Keys k{nl, ""};
    std::string rootname, beamline;
    double delz = dk.setup.delz;
    int harmonic = 1, subharmonic = 1;
    bool resample = false, disable = false;
    k.str("rootname", rootname);
    k.str("beamline", beamline);
    k.dbl("delz", delz);
    k.integer("subharmonic", subharmonic);
    k.integer("harmonic", harmonic);
    k.boolean("resample", resample);
    k.boolean("disable", disable);
    if (!k.finish()) {
        err = k.err;
        return false;
    }
    if (harmonic < 1 || subharmonic < 1) {
        err = "harmonic and subharmonic must be at least 1";
        return false;
    }
    if (disable && harmonic > 1) {
        err = "disable = true (keeping non-matching harmonics as disabled fields) is not supported by this build";
        return false;
    }
    dk.setup.delz = delz;
    if (!dk.time.dotime) resample = false;  // AlterSetup.cpp:113
    if (!beamline.empty()) {
        LatticeResult lr = parse_and_unroll(dk.lattice_text, beamline, delz, dk.setup.lambda0, dk.setup.gamma0);
        if (!lr.error.empty()) {
            err = "lattice: " + lr.error;
            return false;
        }
        dk.setup.beamline = beamline;
    }
    if (!rootname.empty()) {
        dk.setup.rootname = rootname;
        dk.runcount = 0;
    } else {
        dk.runcount++;
    }
    if (harmonic == 1 && subharmonic == 1) return true;
    if (!release_ctx(dk, err)) return false;
    if (subharmonic > 1 && !subharmonic_conversion(dk, subharmonic, resample, device, dry, err)) return false;
    if (harmonic > 1 && !harmonic_conversion(dk, harmonic, resample, device, dry, err)) return false;
    return true;
}

// &sort (GenMain.cpp:428-431): Beam::sort, which only moves one4one beams
bool do_sort(Deck &dk, int device, bool dry, std::string &err) {
    if (!dk.have_beam || !dk.setup.one4one || dry) return true;
    if (dk.ctx) {
        long long kept = 0;
        if (g4mi_sort(dk.ctx, &kept) != G4MI_OK) {
            err = g4mi_last_error();
            return false;
        }
        return true;
    }
    return device_sort(dk, device, err);
}

int run_deck(const char *input_file, int device, int quiet, bool dry, std::string *summary) {
    if (!input_file) return fail("no input file", G4MI_E_ARG);
    Deck dk;
    dk.input_file = input_file;
    if (!read_file(dk.input_file, dk.input_text)) return fail("cannot open main input file " + dk.input_file);
    std::vector<Namelist> nls;
    std::string err;
    if (!parse_namelists(dk.input_text, nls, err)) return fail(err);
    int rc = G4MI_OK;
    for (Namelist &nl : nls) {
        bool ok = true;
        const std::string &e = nl.name;
        if (e == "&stop") break;
        if (e == "&setup") ok = do_setup(dk, nl, err);
        else if (e == "&lattice") ok = dk.setup.defined ? do_lattice(dk, nl, err)
                                                      : (err = "&lattice before &setup", false);
        else if (e == "&time") ok = do_time(dk, nl, err);
        else if (e == "&field") ok = do_field(dk, nl, err);
        else if (e == "&beam") ok = do_beam(dk, nl, err);
        else if (e == "&efield") ok = do_efield(dk, nl, err);
        else if (e == "&track") ok = do_track(dk, nl, device, quiet != 0, dry, err);
        else if (e == "&write") ok = do_write(dk, nl, err);
        else if (e == "&importbeam") ok = do_importbeam(dk, nl, err);
        else if (e == "&importfield") ok = do_importfield(dk, nl, err);
        else if (e == "&wake") ok = do_wake(dk, nl, err);
        else if (e == "&sponrad") ok = do_sponrad(dk, nl, err);
        else if (e == "&alter_setup") ok = dk.setup.defined ? do_alter_setup(dk, nl, device, dry, err)
                                                           : (err = "&alter_setup before &setup", false);
        else if (e == "&sort") ok = do_sort(dk, device, dry, err);
        else {
            err = "element " + e + " is not supported by this build";
            ok = false;
        }
        if (!ok) {
            rc = fail(e + ": " + err, err.rfind("device", 0) == 0 ? G4MI_E_HIP : G4MI_E_PARSE);
            break;
        }
    }
    if (summary) *summary = dk.summary;
    if (dk.ctx) g4mi_destroy(dk.ctx);
    return rc;
}

}  // namespace
}  // namespace g4mi

using namespace g4mi;

extern "C" int g4mi_run_deck(const char *input_file, int device, int quiet) {
    return run_deck(input_file, device, quiet, false, nullptr);
}

extern "C" int g4mi_check_deck(const char *input_file, char *json, int len) {
    std::string s;
    const int rc = run_deck(input_file, 0, 1, true, &s);
    if (json && len > 0) snprintf(json, (size_t)len, "%s", s.c_str());
    return rc;
}

extern "C" long long g4mi_h5_read(const char *file, const char *dataset, double *out, long long cap,
                                  long long *dims) {
    if (!file || !dataset) {
        set_error("g4mi_h5_read: invalid argument");
        return G4MI_E_ARG;
    }
    std::vector<double> v;
    long long d[2] = {0, 0};
    std::string err;
    const long long n = h5_read_doubles(file, dataset, v, d, err);
    if (n < 0) {
        set_error("g4mi_h5_read: " + err);
        return G4MI_E_PARSE;
    }
    if (dims) {
        dims[0] = d[0];
        dims[1] = d[1];
    }
    if (out)
        for (long long i = 0; i < std::min(n, cap); i++) out[i] = v[(size_t)i];
    return n;
}

// The step through the C++ host classes (include/g4mi.hpp), against the same
// steps through one g4mi_run call: a small time-dependent SASE window
// (6 slices x 512 macroparticles, shot noise, long-range and short-range
// space charge, 63x63 field with a 100 kW seed, a planar undulator module
// pair with a FODO drift).  Prints one JSON line; exit status 0 when both
// paths give the same power and bunching records to 1e-13.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "g4mi.hpp"

namespace {

const double kGamma0 = 11357.8, kAw = 1.7, kLambdaU = 0.015;
const double kLambda0 = kLambdaU * (1 + kAw * kAw) / (2 * kGamma0 * kGamma0);
const char *kLattice =
    "UND: UNDULATOR = {lambdau=0.015, nwig=30, aw=1.7};\n"
    "D: DRIFT = {l=0.21};\n"
    "QF: QUADRUPOLE = {l=0.08, k1=2.0};\n"
    "FEL: LINE = {UND, D, QF, D, UND};\n";

struct Inputs {
    std::vector<long long> off;
    std::vector<double> g, th, x, y, px, py, cur, field;
    std::vector<g4mi_step_params> steps;
};

Inputs make_inputs(int ns, int npart, int ngrid, double gridmax) {
    Inputs in;
    in.off.assign(ns + 1, 0);
    for (int i = 0; i < ns; i++) in.off[i + 1] = in.off[i] + npart;
    const size_t n = (size_t)in.off[ns];
    for (auto *v : {&in.g, &in.th, &in.x, &in.y, &in.px, &in.py}) v->assign(n, 0.0);
    std::vector<g4mi_slice_beam> sb(ns);
    for (auto &s : sb) {
        s = g4mi_slice_beam{};
        s.gamma = kGamma0;
        s.delgam = 1e-4 * kGamma0;
        s.current = 3000;
        s.ex = s.ey = 0.4e-6;
        s.betax = s.betay = 15;
    }
    in.cur.assign(ns, 3000.0);
    g4mi::check(g4mi_load_beam(ns, in.off.data(), sb.data(), 4, 1, 7, 0, kLambda0, 1.0, 1, in.g.data(),
                               in.th.data(), in.x.data(), in.y.data(), in.px.data(), in.py.data()),
                "g4mi_load_beam");
    const size_t cells = (size_t)ngrid * ngrid;
    std::vector<double> f(cells * 2);
    g4mi::check(g4mi_load_gauss(f.data(), ngrid, gridmax, kLambda0, 1e5, 30e-6, 0.0, 0.0), "g4mi_load_gauss");
    in.field.resize((size_t)ns * cells * 2);
    for (int i = 0; i < ns; i++) std::copy(f.begin(), f.end(), in.field.begin() + (long)i * cells * 2);
    int nst = 0;
    g4mi::check(g4mi_parse_lattice(kLattice, "FEL", 0.015, kLambda0, kGamma0, nullptr, 0, &nst), "lattice");
    in.steps.resize(nst);
    g4mi::check(g4mi_parse_lattice(kLattice, "FEL", 0.015, kLambda0, kGamma0, in.steps.data(), nst, &nst),
                "lattice");
    return in;
}

void setup(g4mi::Context &ctx, const Inputs &in) {
    ctx.initEField(0.0, 50, 1, 0, true);
    ctx.uploadBeam(in.g.data(), in.th.data(), in.x.data(), in.y.data(), in.px.data(), in.py.data(), in.cur.data());
    ctx.uploadField(in.field.data(), 0);
}

}  // namespace

int main(int argc, char **argv) {
    const int device = argc > 1 ? std::stoi(argv[1]) : 0;
    const int ns = 6, npart = 512, ngrid = 63;
    const double gridmax = 0.3e-3;
    try {
        const Inputs in = make_inputs(ns, npart, ngrid, gridmax);
        const int nst = (int)in.steps.size();
        // path 1: the reference's classes, step by step
        std::vector<double> p1((size_t)nst * ns), b1((size_t)nst * ns);
        int done1 = 0;
        {
            g4mi::Context ctx(device, in.off, ngrid, gridmax, kLambda0, 1.0, true, false);
            setup(ctx, in);
            std::vector<g4mi_diag_out> rec(nst);
            done1 = g4mi::gencore_run(ctx, in.steps, kGamma0, 1, [&](size_t i) {
                rec[i] = g4mi_diag_out{};
                rec[i].power = p1.data() + i * ns;
                rec[i].bunching = b1.data() + i * ns;
                return &rec[i];
            });
        }
        // path 2: one g4mi_run call with a record after every step
        std::vector<double> p2((size_t)nst * ns), b2((size_t)nst * ns);
        int done2 = 0;
        {
            g4mi::Context ctx(device, in.off, ngrid, gridmax, kLambda0, 1.0, true, false);
            setup(ctx, in);
            std::vector<g4mi_diag_out> rec(nst);
            for (int i = 0; i < nst; i++) {
                rec[i] = g4mi_diag_out{};
                rec[i].power = p2.data() + (size_t)i * ns;
                rec[i].bunching = b2.data() + (size_t)i * ns;
            }
            done2 = g4mi::check(g4mi_run(ctx.handle(), in.steps.data(), nst, kGamma0, rec.data()), "g4mi_run");
        }
        double ep = 0, eb = 0, pmax = 0;
        for (size_t k = 0; k < (size_t)done1 * ns; k++) pmax = std::fmax(pmax, std::fabs(p2[k]));
        for (size_t k = 0; k < (size_t)done1 * ns; k++) {
            ep = std::fmax(ep, std::fabs(p1[k] - p2[k]) / std::fmax(std::fabs(p2[k]), 1e-12 * pmax));
            eb = std::fmax(eb, std::fabs(b1[k] - b2[k]) / std::fmax(std::fabs(b2[k]), 1e-12));
        }
        const bool same = done1 == done2 && done1 == nst && ep <= 1e-13 && eb <= 1e-13;
        double psum = 0;  // window power of the last record
        for (int k = 0; k < ns; k++) psum += p2[(size_t)(done2 - 1) * ns + k];
        printf("{\"steps\": %d, \"steps_run\": %d, \"max_rel_diff_power\": %.3e, \"max_rel_diff_bunching\": %.3e, "
               "\"p_end_window\": %.10e, \"same\": %s}\n",
               nst, done1, ep, eb, psum, same ? "true" : "false");
        return same ? 0 : 1;
    } catch (const std::exception &e) {
        printf("{\"error\": \"%s\"}\n", e.what());
        return 2;
    }
}

/* C++ host classes over the g4mi C-ABI with the call surface of the
 * reference's step: a FieldSolver with init/advance
 * (include/FieldSolver.h:20-26), Beam::track (src/Core/Beam.cpp:97-112),
 * Control::applySlippage (src/Core/Control.cpp:158-276), Diagnostic::calc
 * (src/Core/Diagnostic.cpp:262-277) and the Gencore::run loop body
 * (src/Core/Gencore.cpp:102-160) that drives them in the same order.
 *
 * Header-only, C++17.  Errors throw g4mi::Error with the library's message
 * (the reference's methods are void and have no error path; a binding
 * that prefers its abort-on-error behaviour catches and aborts).  The
 * per-class calls need the FP64 form: the mixed form advances whole steps
 * only (Gencore::run, or g4mi_run). */
#ifndef G4MI_HPP
#define G4MI_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "g4mi.h"

namespace g4mi {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

inline int check(int rc, const char *what) {
    if (rc < 0) throw Error(std::string(what) + ": " + g4mi_last_error());
    return rc;
}

/* The device state of one time window (Beam + Field + Control of the
 * reference): particles, fields, ring offset, space-charge settings. */
class Context {
  public:
    Context(int device, const std::vector<long long> &off, int ngrid, double gridmax, double lambda0,
            double sample, bool timerun, bool periodic, int precision = G4MI_FP64)
        : ns_((int)off.size() - 1), ngrid_(ngrid), gridmax_(gridmax), precision_(precision) {
        check(g4mi_create(&h_, device, ns_, off.data(), ngrid, gridmax, lambda0, sample, timerun ? 1 : 0,
                          periodic ? 1 : 0, precision),
              "g4mi_create");
    }
    ~Context() {
        if (h_) g4mi_destroy(h_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    /* EFieldSolver::init via Beam::initEField (src/Core/EFieldSolver.cpp:27-49) */
    void initEField(double rmax, int ngrid_r, int nz, int nphi, bool longrange) {
        check(g4mi_set_efield(h_, longrange ? 1 : 0, nz, nphi, ngrid_r, rmax), "g4mi_set_efield");
    }
    void uploadBeam(const double *gamma, const double *theta, const double *x, const double *y, const double *px,
                    const double *py, const double *current) {
        check(g4mi_beam_upload(h_, gamma, theta, x, y, px, py, current), "g4mi_beam_upload");
    }
    void downloadBeam(double *gamma, double *theta, double *x, double *y, double *px, double *py) {
        check(g4mi_beam_download(h_, gamma, theta, x, y, px, py), "g4mi_beam_download");
    }
    void uploadField(const double *field, int first) { check(g4mi_field_upload(h_, field, first), "g4mi_field_upload"); }
    int downloadField(double *field) {
        int first = 0;
        check(g4mi_field_download(h_, field, &first), "g4mi_field_download");
        return first;
    }
    void sync() { check(g4mi_sync(h_), "g4mi_sync"); }
    /* Field::initSolver (src/Core/Field.cpp:72-83): the ADI or the spectral
     * solver (FP64 form), with FieldSolverFFT::initSourceFilter's settings */
    void initSolver(bool isFFT, bool filter = false, double xc = 1, double yc = 1, double sig = 1) {
        check(g4mi_set_field_solver(h_, isFFT ? 1 : 0, filter ? 1 : 0, xc, yc, sig), "g4mi_set_field_solver");
    }

    g4mi_ctx *handle() const { return h_; }
    int nslice() const { return ns_; }
    int ngrid() const { return ngrid_; }
    double gridmax() const { return gridmax_; }
    int precision() const { return precision_; }

  private:
    g4mi_ctx *h_ = nullptr;
    int ns_, ngrid_;
    double gridmax_;
    int precision_;
};

/* FieldSolver (include/FieldSolver.h:20-26) on the device: deposit + ADI
 * (FieldSolverADI) or the spectral step (FieldSolverFFT, after
 * Context::initSolver(true)) for every slice of the context. */
class FieldSolverMI {
  public:
    explicit FieldSolverMI(Context &ctx) : ctx_(ctx) {}
    /* FieldSolverADI::init (:125-182) happens inside advance whenever delz
     * changes; here only the grid is checked against the context's. */
    void init(double /*delz*/, double dgrid, double /*xks*/, unsigned int ngrid) {
        if ((int)ngrid != ctx_.ngrid()) throw Error("FieldSolverMI::init: grid size differs from the context's");
        const double d = 2 * ctx_.gridmax() / (ctx_.ngrid() - 1);
        if (dgrid > 0 && (dgrid > d * (1 + 1e-12) || dgrid < d * (1 - 1e-12)))
            throw Error("FieldSolverMI::init: grid spacing differs from the context's");
    }
    /* FieldSolver::advance(delz, field, beam, und): und's current step */
    void advance(double delz, const g4mi_step_params &und) {
        check(g4mi_field_track(ctx_.handle(), &und, delz), "FieldSolverMI::advance");
    }
    void initSourceFilter(double, double, double, bool) {}

  private:
    Context &ctx_;
};

/* Beam::track (src/Core/Beam.cpp:97-112) through BeamSolver: half
 * transverse step, space charge, RK4, R56, half transverse step. */
class BeamSolverMI {
  public:
    explicit BeamSolverMI(Context &ctx) : ctx_(ctx) {}
    void track(double delz, const g4mi_step_params &und, double gammaref) {
        check(g4mi_beam_track(ctx_.handle(), &und, gammaref, delz), "BeamSolverMI::track");
    }

  private:
    Context &ctx_;
};

/* Control::applySlippage (src/Core/Control.cpp:158-276) */
class ControlMI {
  public:
    explicit ControlMI(Context &ctx) : ctx_(ctx) {}
    void applySlippage(double slippage) { check(g4mi_slippage(ctx_.handle(), slippage), "ControlMI::applySlippage"); }

  private:
    Context &ctx_;
};

/* Diagnostic::calc: one record of the per-slice outputs (any pointer of
 * `out` may be null). */
class DiagnosticMI {
  public:
    explicit DiagnosticMI(Context &ctx) : ctx_(ctx) {}
    void calc(g4mi_diag_out &out) { check(g4mi_diag(ctx_.handle(), &out), "DiagnosticMI::calc"); }

  private:
    Context &ctx_;
};

/* Gencore::run (src/Core/Gencore.cpp:102-160) over an unrolled lattice:
 * for every step until a stop marker, Beam::track, Field::track,
 * slippage, and a diagnostics record on output steps (`record` receives
 * the step's index, and returns the output record to fill or nullptr).
 * Returns the number of steps done.  FP64 form: the per-class calls;
 * mixed form: whole steps. */
template <class RecordFn>
int gencore_run(Context &ctx, const std::vector<g4mi_step_params> &steps, double gammaref, int output_step,
                RecordFn record) {
    BeamSolverMI beam(ctx);
    FieldSolverMI field(ctx);
    ControlMI control(ctx);
    DiagnosticMI diag(ctx);
    if (output_step < 1) output_step = 1;
    int n = 0;
    for (size_t i = 0; i < steps.size(); i++) {
        const g4mi_step_params &und = steps[i];
        if (und.marker & 8) break;  // Undulator::advance stops here
        if (ctx.precision() == G4MI_FP64) {
            beam.track(und.dz, und, gammaref);
            field.init(und.dz, 2 * ctx.gridmax() / (ctx.ngrid() - 1), 0.0, (unsigned)ctx.ngrid());
            field.advance(und.dz, und);
            control.applySlippage(und.slip);
        } else {
            check(g4mi_step(ctx.handle(), &und, gammaref), "g4mi_step");
        }
        n++;
        if (n % output_step == 0) {
            g4mi_diag_out *o = record(i);
            if (o) diag.calc(*o);
        }
    }
    return n;
}

}  // namespace g4mi

#endif

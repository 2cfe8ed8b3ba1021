// Shared host/device definitions of the g4mi library.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "g4mi.h"

namespace g4mi {

// Physical constants of the reference, src/Main/GenMain.cpp:63-68.
constexpr double kVacimp = 376.73;
constexpr double kEev = 510998.95069;
constexpr double kCe = 4.8032045e-11;
constexpr double kClight = 299792458.0;

void set_error(const std::string &msg);

#define G4_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            g4mi::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return G4MI_E_HIP;                                                    \
        }                                                                         \
    } while (0)

#define G4_TRY(x)                     \
    do {                              \
        int r_ = (x);                 \
        if (r_ != G4MI_OK) return r_; \
    } while (0)

// Parameters of the transverse half-step (TrackBeam::track,
// src/Core/TrackBeam.cpp:8-81), prepared on the host once per call.
struct TransParams {
    double aw, delz, qx, qy, xoff, yoff;
    double ccx, ccy;               // corrector kicks (already times gamma0)
    double m00, m01, m10, m11, m22, m23, m32, m33;  // chicane matrix
    int kindx, kindy;              // 0 drift, +1 focusing, -1 defocusing
    int corrector, chicane;
};

// Parameters of BeamSolver::advance for the fundamental.
// A field at a harmonic of the fundamental as the push gathers it
// (BeamSolver::advance :23-32, 63-79): its grid, rtmp = fc(h) / xks_h and h.
struct HarmGather {
    const double2 *field;
    double gridmax, dgrid, rtmp, harm;
    int ngrid;
};
constexpr int kMaxHarm = 4;  // harmonic fields besides the fundamental

struct PushParams {
    double delz, xks, xku, rtmp, aw, autophase;
    double ax, ay, kx, ky, gradx, grady;   // faw(x,y)
    double gridmax, dgrid;
    int ngrid, first, nslice, use_sc;
    int nh;                                // harmonic fields, 0..kMaxHarm
    HarmGather hg[kMaxHarm];
};

// Parameters of the source deposition (FieldSolverADI::advance :13-51).
struct DepositParams {
    double fcoup, xks, delz, dgrid, gridmax;
    double ax, ay, kx, ky, gradx, grady;
    double harm;                           // phase factor h of sin(h theta), cos(h theta)
    int ngrid, enabled;
};

// Device state and host mirror of one context.
struct Ctx;

}  // namespace g4mi

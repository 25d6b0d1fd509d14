// Internal definitions shared by the host-side sources of libe1039or.
#pragma once
#include "../../include/e1039_or.h"

#include <cstdint>
#include <vector>

// Per-plane geometry fields, in the member order of the reference's gPlane
// (reconstruction_classes.cuh:947-977).  nelem is held as a float.
enum PlaneField {
    F_Z, F_NELEM, F_CELLWIDTH, F_SPACING, F_XOFFSET, F_SCALEX, F_X0, F_X1, F_X2, F_COSTHETA, F_SCALEY, F_Y0, F_Y1,
    F_Y2, F_SINTHETA, F_RESOLUTION, F_P1X, F_P1Y, F_P1Z, F_DELTAPX, F_DELTAPY, F_DELTAPZ, F_DP1X, F_DP1Y, F_DP1Z,
    F_SLOPE_MAX, F_INTER_MAX, F_COUNT
};

struct or_geometry {
    float f[F_COUNT][OR_N_DETECTORS];          // plane table
    float wire[OR_N_DETECTORS][OR_MAX_ELEM];   // wire position per (detector, element), fp32
    float deltaW[OR_N_DETECTORS][9];           // alignment offsets read from the file
};

// R-T calibration: per detector the knots and the cubic of each knot interval,
// value(t) = y + dt (b + dt (c + dt d)) with dt = t - x[k] (n - 1 intervals).
struct or_calibration {
    struct Det {
        int n = 0;                          // knots (0: not calibrated)
        float tmin = 0.f, tmax = 0.f;
        std::vector<double> x, y, b, c, d;  // x: n knots; y..d: n - 1 intervals
    };
    Det det[OR_N_DETECTORS];
};

// R-T calibration tables: parse, and the cubic spline of each detector.
//
// The reference reads calibration_*.txt into ROOT TSpline3 objects with option
// "b1e1" (online_reconstruction.cu:220-242): an interpolating cubic spline with
// the first derivative prescribed (0) at both ends.  Here the same spline is
// built from its second derivatives M at the knots, solved from the
// tridiagonal system of C2 continuity with the two end-slope rows, and turned
// into the Taylor coefficients of each knot interval.  All in double.
#include "or_internal.h"

#include <cmath>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

namespace {

bool calibrated_id(int d) { return (d >= 1 && d <= 30) || (d >= 47 && d <= 54); }

// Clamped cubic spline through (x[i], y[i]), zero slope at both ends.
// Rows (h_i = x[i+1] - x[i], s_i = (y[i+1] - y[i]) / h_i):
//   2 h_0 M_0 + h_0 M_1                                 = 6 (s_0 - 0)
//   h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (s_i - s_{i-1})
//   h_{n-2} M_{n-2} + 2 h_{n-2} M_{n-1}                 = 6 (0 - s_{n-2})
// solved by forward elimination and back substitution (diagonally dominant).
void build_spline(or_calibration::Det& D, const double* x, const double* y, int n) {
    D.n = n;
    D.x.assign(x, x + n);
    std::vector<double> h(n - 1), s(n - 1), diag(n), up(n), rhs(n), M(n);
    for (int i = 0; i + 1 < n; ++i) {
        h[i] = x[i + 1] - x[i];
        s[i] = (y[i + 1] - y[i]) / h[i];
    }
    for (int i = 0; i < n; ++i) {
        const double hl = i > 0 ? h[i - 1] : 0.0, hr = i + 1 < n ? h[i] : 0.0;
        const double sl = i > 0 ? s[i - 1] : 0.0, sr = i + 1 < n ? s[i] : 0.0;
        diag[i] = 2.0 * (hl + hr);
        up[i] = hr;          // coefficient of M_{i+1}
        rhs[i] = 6.0 * (sr - sl);
    }
    // forward elimination: row i -= (h_{i-1} / diag'_{i-1}) row i-1
    for (int i = 1; i < n; ++i) {
        const double f = h[i - 1] / diag[i - 1];
        diag[i] -= f * up[i - 1];
        rhs[i] -= f * rhs[i - 1];
    }
    M[n - 1] = rhs[n - 1] / diag[n - 1];
    for (int i = n - 2; i >= 0; --i) M[i] = (rhs[i] - up[i] * M[i + 1]) / diag[i];
    D.y.resize(n - 1); D.b.resize(n - 1); D.c.resize(n - 1); D.d.resize(n - 1);
    for (int i = 0; i + 1 < n; ++i) {
        D.y[i] = y[i];
        D.b[i] = s[i] - h[i] * (2.0 * M[i] + M[i + 1]) / 6.0;
        D.c[i] = 0.5 * M[i];
        D.d[i] = (M[i + 1] - M[i]) / (6.0 * h[i]);
    }
}

}  // namespace

extern "C" {

or_status or_calibration_create(int n_det, const int32_t* det_ids, const int32_t* n_knots, const float* tmin,
                                const float* tmax, const double* t, const double* r, or_calibration** out) {
    if (!out || n_det < 0 || (n_det > 0 && (!det_ids || !n_knots || !tmin || !tmax))) return OR_ERR_INVALID_ARG;
    *out = nullptr;
    int64_t total = 0;
    for (int i = 0; i < n_det; ++i) {
        const int d = det_ids[i], n = n_knots[i];
        if (d < 1 || d >= OR_N_DETECTORS || n < 0 || n == 1) return OR_ERR_INVALID_ARG;
        if (n > OR_RT_MAX_KNOTS) return OR_ERR_CAPACITY;
        if (n > 0 && (!t || !r)) return OR_ERR_INVALID_ARG;
        for (int k = 0; k < n; ++k) {
            if (!std::isfinite(t[total + k]) || !std::isfinite(r[total + k])) return OR_ERR_INVALID_ARG;
            if (k > 0 && !(t[total + k] > t[total + k - 1])) return OR_ERR_INVALID_ARG;
        }
        total += n;
    }
    auto* c = new (std::nothrow) or_calibration;
    if (!c) return OR_ERR_NOMEM;
    total = 0;
    for (int i = 0; i < n_det; ++i) {
        const int d = det_ids[i], n = n_knots[i];
        or_calibration::Det& D = c->det[d];
        D = or_calibration::Det();
        // hodoscope entries are read and never applied (online_reconstruction.cu:389)
        if (calibrated_id(d) && n >= 2) {
            D.tmin = tmin[i];
            D.tmax = tmax[i];
            build_spline(D, t + total, r + total, n);
        }
        total += n;
    }
    *out = c;
    return OR_OK;
}

or_status or_calibration_load(const char* path, or_calibration** out) {
    if (!path || !out) return OR_ERR_INVALID_ARG;
    *out = nullptr;
    std::ifstream in(path);
    if (!in) return OR_ERR_IO;
    std::vector<int32_t> ids, nk;
    std::vector<float> lo, hi;
    std::vector<double> tt, rr;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream head(line);
        int d = 0, n = 0;
        double t0 = 0.0, t1 = 0.0;
        std::string name;
        if (!(head >> d)) continue;          // blank line
        if (!(head >> n >> t0 >> t1)) return OR_ERR_PARSE;
        if (d < 1 || d >= OR_N_DETECTORS || n < 0) return OR_ERR_PARSE;
        if (n > OR_RT_MAX_KNOTS) return OR_ERR_CAPACITY;
        for (int k = 0; k < n; ++k) {
            if (!std::getline(in, line)) return OR_ERR_PARSE;
            std::istringstream row(line);
            int ib = 0;
            double tk = 0.0, rk = 0.0;
            if (!(row >> ib >> tk >> rk)) return OR_ERR_PARSE;
            tt.push_back(tk);
            rr.push_back(rk);
        }
        ids.push_back(d);
        nk.push_back(n);
        lo.push_back((float)t0);   // stored as float (online_reconstruction.cu:121, 232-233)
        hi.push_back((float)t1);
    }
    return or_calibration_create((int)ids.size(), ids.data(), nk.data(), lo.data(), hi.data(), tt.data(), rr.data(),
                                 out);
}

void or_calibration_free(or_calibration* c) { delete c; }

}  // extern "C"

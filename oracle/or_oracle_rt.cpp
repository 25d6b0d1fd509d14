/*
 * or_oracle_rt.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's R-T calibration
 * (OnlineReconstruction/online_reconstruction.cu:220-242 builds one ROOT
 * TSpline3 per detector with option "b1e1"; :387-397 applies it per hit).
 *
 * TSpline3 belongs to ROOT (the reference links it through root-config; no
 * version is pinned and ROOT is absent here).  Its published algorithm is
 * restated: BuildCoeff is de Boor's CUBSPL ("A Practical Guide to Splines",
 * 1978): a tridiagonal system for the slopes s_i at the knots, the "b1"/"e1"
 * rows prescribing s_0 = valbeg and s_{n-1} = valend (both 0 here), forward
 * elimination, back substitution, then the Taylor coefficients of each
 * interval from the values and slopes at its ends.  Eval finds the interval by
 * binary search (x_k < x <= x_k+1; the first interval for x <= x_0) and uses
 * interval n-2 for x >= x_{n-1}, evaluating y + dx (b + dx (c + dx d)).
 */
#include <cstdint>
#include <vector>

namespace {

// de Boor CUBSPL with both end slopes prescribed; b, c, d of intervals 0..n-2
void cubspl_clamped(int n, const double* x, const double* y, double valbeg, double valend, double* B, double* C,
                    double* D) {
    // work arrays in de Boor's roles: c2 = slopes/right-hand side, c3 = h or the
    // super-diagonal, c4 = divided differences or the diagonal
    std::vector<double> c2(n), c3(n), c4(n);
    const int l = n - 1;
    for (int m = 1; m < n; ++m) {
        c3[m] = x[m] - x[m - 1];
        c4[m] = (y[m] - y[m - 1]) / c3[m];
    }
    // "b1": slope given at the left end
    c4[0] = 1.0;
    c3[0] = 0.0;
    c2[0] = valbeg;
    // interior rows, eliminating s_{m-1}
    for (int m = 1; m < l; ++m) {
        const double g = -c3[m + 1] / c4[m - 1];
        c2[m] = g * c2[m - 1] + 3.0 * (c3[m] * c4[m + 1] + c3[m + 1] * c4[m]);
        c4[m] = g * c3[m - 1] + 2.0 * (c3[m] + c3[m + 1]);
    }
    // "e1": slope given at the right end, then back substitution
    c2[l] = valend;
    for (int j = l - 1; j >= 0; --j) c2[j] = (c2[j] - c3[j] * c2[j + 1]) / c4[j];
    for (int i = 1; i < n; ++i) {
        const double dtau = c3[i];
        const double divdf1 = (y[i] - y[i - 1]) / dtau;
        const double divdf3 = c2[i - 1] + c2[i] - 2.0 * divdf1;
        B[i - 1] = c2[i - 1];
        C[i - 1] = (divdf1 - c2[i - 1] - divdf3) / dtau;
        D[i - 1] = divdf3 / dtau / dtau;
    }
}

double spline_eval(int n, const double* x, const double* y, const double* B, const double* C, const double* D,
                   double t) {
    int klow = 0, khig = n - 1;
    if (t <= x[0]) {
        klow = 0;
    } else if (t >= x[n - 1]) {
        klow = n - 1;
    } else {
        while (khig - klow > 1) {
            const int khalf = (klow + khig) / 2;
            if (t > x[khalf]) klow = khalf; else khig = khalf;
        }
    }
    if (klow >= n - 1 && n > 1) klow = n - 2;
    const double dx = t - x[klow];
    return y[klow] + dx * (B[klow] + dx * (C[klow] + dx * D[klow]));
}

}  // namespace

extern "C" {

// Spline coefficients of n knots (n >= 2): b, c, d of the n-1 intervals.
void orc_rt_spline(int n, const double* x, const double* y, double* b, double* c, double* d) {
    cubspl_clamped(n, x, y, 0.0, 0.0, b, c, d);
}

// Spline values at m points.
void orc_rt_eval(int n, const double* x, const double* y, int64_t m, const double* t, double* out) {
    std::vector<double> b(n - 1), c(n - 1), d(n - 1);
    cubspl_clamped(n, x, y, 0.0, 0.0, b.data(), c.data(), d.data());
    for (int64_t i = 0; i < m; ++i) out[i] = spline_eval(n, x, y, b.data(), c.data(), d.data(), t[i]);
}

// Drift of each hit (online_reconstruction.cu:387-397).  nk[55] knots per
// detector (0: no spline), knots of detector 1, 2, ... concatenated in t, r;
// tmin/tmax/cellwidth per detector.  Only detectors 1-30 and 47-54 with knots
// are calibrated; others keep their input drift.
void orc_rt_calibrate(const int32_t* nk, const float* tmin, const float* tmax, const double* t, const double* r,
                      const float* cellwidth, int64_t n_hits, const int16_t* det, const float* tdc,
                      const float* drift, float* out) {
    std::vector<int64_t> off(56, 0);
    for (int dd = 0; dd < 55; ++dd) off[dd + 1] = off[dd] + nk[dd];
    std::vector<std::vector<double>> B(55), C(55), D(55);
    for (int dd = 0; dd < 55; ++dd) {
        if (nk[dd] < 2) continue;
        B[dd].resize(nk[dd] - 1); C[dd].resize(nk[dd] - 1); D[dd].resize(nk[dd] - 1);
        cubspl_clamped(nk[dd], t + off[dd], r + off[dd], 0.0, 0.0, B[dd].data(), C[dd].data(), D[dd].data());
    }
    for (int64_t i = 0; i < n_hits; ++i) {
        const int dd = det[i];
        float v = drift[i];
        const bool cal = ((dd >= 1 && dd <= 30) || (dd >= 47 && dd <= 54)) && nk[dd] >= 2;
        if (cal) {
            if (tdc[i] < tmin[dd]) {
                v = (float)(cellwidth[dd] * 0.5);
            } else if (tdc[i] > tmax[dd]) {
                v = 0;
            } else {
                v = (float)spline_eval(nk[dd], t + off[dd], r + off[dd], B[dd].data(), C[dd].data(), D[dd].data(),
                                       (double)tdc[i]);
            }
        }
        out[i] = v;
    }
}

}  // extern "C"

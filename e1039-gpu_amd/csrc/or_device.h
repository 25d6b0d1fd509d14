// Device-side math of the chain: wire geometry, straight-line fits, left/right
// resolution, hodoscope windows, residuals.  Every expression keeps the
// operation order of the reference code it cites; the library is built with
// -ffp-contract=off so each float operation rounds once.
#pragma once
#include <hip/hip_runtime.h>
#include "or_internal.h"

namespace ordev {

constexpr int ND = OR_N_DETECTORS;

// plane table (reconstruction_classes.cuh:947-977), uploaded per session
extern __constant__ float c_geo[F_COUNT][ND];

__device__ __forceinline__ float G(int f, int d) { return c_geo[f][d]; }

// reconstruction_constants.h:29-34
constexpr float TX_MAX = 0.15f, TY_MAX = 0.1f, X0_MAX = 150.f, Y0_MAX = 50.f;
constexpr float INVP_MAX = 0.2f, INVP_MIN = 0.01f;
// reconstruction_constants.h:75-96 (E1039)
constexpr float Z_DUMP = 42.f, Z_TARGET = -300.f;
constexpr float SAG_T_C = 1.85f, SAG_T_W = 0.25f, SAG_D_C = 1.5f, SAG_D_W = 0.3f;
constexpr float PT_KICK_KMAG = -0.3819216f, FMAGSTR = -1.054f, Z_KMAG_BEND = 1064.26f;

__device__ __forceinline__ float spacing_plane(int sid) {
    // globalconsts::spacingplane (reconstruction_constants.h:64)
    constexpr float t[28] = {0.f, 0.40f, 0.40f, 0.40f, 0.40f, 0.40f, 0.40f, 1.3f, 1.3f, 1.3f, 1.2f, 1.2f, 1.2f, 1.2f,
                             1.2f, 1.2f, 4.0f, 4.0f, 7.0f, 7.0f, 8.0f, 12.0f, 12.0f, 10.0f, 3.0f, 3.0f, 3.0f, 3.0f};
    return t[sid];
}
// globalconsts::detsuperid (reconstruction_constants.h:65)
__device__ __forceinline__ int super_id(int st, int proj) {
    constexpr short t[7][3] = {{2, 1, 3}, {5, 6, 4}, {8, 9, 7}, {11, 12, 10}, {14, 15, 13}, {25, 24, -1}, {26, 27, -1}};
    return t[st][proj];
}
// lrpossibility (reconstruction_constants.h:122): hypothesis h -> (s_i, s_j)
__device__ __forceinline__ short lr_si(int h) { return h < 2 ? 1 : -1; }
__device__ __forceinline__ short lr_sj(int h) { return (h & 1) ? -1 : 1; }

// ---- wire end points (reconstruction_helper.cuh:8-36) ---------------------
__device__ __forceinline__ float x_bep(int d, int e) { return G(F_P1X, d) + G(F_DP1X, d) * (float)(e - 1); }
__device__ __forceinline__ float y_bep(int d, int e) { return G(F_P1Y, d) + G(F_DP1Y, d) * (float)(e - 1); }
__device__ __forceinline__ float z_bep(int d, int e) { return G(F_P1Z, d) + G(F_DP1Z, d) * (float)(e - 1); }

// ---- y of a UV hit from the x track (reconstruction_helper.cuh:75-100) ----
__device__ __forceinline__ bool y_uvhit(int d, int e, float drift, short hs, float x0, float tx, float& y, float& ey) {
    const float p1x = x_bep(d, e);
    const float p2x = p1x + G(F_DELTAPX, d);
    const float xt = x0 + G(F_Z, d) * tx;
    if (xt - drift > p1x && xt - drift > p2x) return false;
    if (xt + drift < p1x && xt + drift < p2x) return false;
    const float p1y = y_bep(d, e);
    float yy = p1y + ((xt - p1x) * G(F_DELTAPY, d)) / G(F_DELTAPX, d);
    if (hs != 0) yy += (copysignf(1.f, G(F_SINTHETA, d)) * 2.f) * (drift - 0.5f * (float)hs);
    yy = fmaxf(yy, p1y);
    yy = fminf(yy, p1y + G(F_DELTAPY, d));
    y = yy;
    ey = G(F_RESOLUTION, d) * fabsf(G(F_DELTAPY, d) / G(F_DELTAPX, d));
    return true;
}

// ---- x, z of a wire at the track's y (reconstruction_helper.cuh:140-158) --
__device__ __forceinline__ void xz_fy(int d, int e, float drift, short hs, float y0, float ty, float& x, float& z) {
    const float p1x = x_bep(d, e), p1y = y_bep(d, e), p1z = z_bep(d, e);
    const float dpy = G(F_DELTAPY, d), dpz = G(F_DELTAPZ, d);
    float zz = p1z - (p1y * dpz) / dpy;
    const float yt = y0 + zz * ty;
    float xx = p1x + ((yt - p1y) * G(F_DELTAPX, d)) / dpy;
    zz = p1z + ((yt - p1y) * dpz) / dpy;
    if (hs != 0) xx += (float)hs * drift;
    x = xx;
    z = zz;
}

// ---- 2x2 inverse with pivoting (trackfittingtools.cuh:3-82) ---------------
__device__ __forceinline__ void matinv2(float A0, float A1, float A2, float A3, float* Ai) {
    float a00 = A0, a10 = A1, a01 = A2, a11 = A3, tmp;
    int p0 = 0, p1 = 1;
    if (fabsf(a10) > fabsf(a00)) {
        tmp = a00; a00 = a10; a10 = tmp;
        tmp = a01; a01 = a11; a11 = tmp;
        p0 = 1; p1 = 0;
    }
    tmp = 1.0f / a00;
    a00 = tmp;
    a01 = tmp * a01;
    tmp = a10;
    a10 = (-tmp) * a00;
    a11 = -(tmp * a01) + a11;
    tmp = 1.0f / a11;
    a10 = tmp * a10;
    a11 = tmp;
    tmp = a01;
    a00 = -(tmp * a10) + a00;
    a01 = (-tmp) * a11;
    // Ainv[row*2 + col]; columns follow the permutation
    if (p0 == 0) { Ai[0] = a00; Ai[2] = a10; Ai[1] = a01; Ai[3] = a11; }
    else { Ai[1] = a00; Ai[3] = a10; Ai[0] = a01; Ai[2] = a11; }
}

// ---- weighted straight-line fit (trackanalyticalminimizer.cuh:32-74) ------
// Normal equations with the reference weighting w = sigma (not 1/sigma^2).
struct Fit { float p0, p1, e0, e1; };

__device__ __forceinline__ Fit fit_finish(float A0, float A1, float A3, float B0, float B1) {
    float Ai[4];
    matinv2(A0, A1, A1, A3, Ai);
    Fit f;
    f.e0 = sqrtf(fabsf(Ai[0]));
    f.e1 = sqrtf(fabsf(Ai[3]));
    float q = 0.0f;
    q += Ai[0] * B0;
    q += Ai[1] * B1;
    f.p0 = q;
    q = 0.0f;
    q += Ai[2] * B0;
    q += Ai[3] * B1;
    f.p1 = q;
    return f;
}

template <int N>
__device__ __forceinline__ Fit fit2d(int n, const float* x, const float* z, const float* w) {
    float A0 = 0.f, A1 = 0.f, A3 = 0.f, B0 = 0.f, B1 = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i < n) {
            B0 += w[i] * x[i];
            B1 += (w[i] * x[i]) * z[i];
            A0 += w[i];
            A1 += w[i] * z[i];
            A3 += (w[i] * z[i]) * z[i];
        }
    }
    return fit_finish(A0, A1, A3, B0, B1);
}

// ---- signed distance to the wire line (reconstruction_helper.cuh:1149-1204)
__device__ __forceinline__ float dca_line(float p1x, float p1y, float p1z, float dpx, float dpy, float dpz, float x0,
                                          float y0, float tx, float ty) {
    const float den2 = (dpy * dpy) * (1.f + tx * tx) + (dpx * dpx) * (1.f + ty * ty) -
                       2.f * (((ty * dpx) * dpz + (ty * dpy) * dpz) + ((tx * ty) * dpx) * dpy);
    return ((ty * dpz - dpy) * (p1x - x0) + (dpx - tx * dpz) * (p1y - y0) + p1z * (tx * dpy - ty * dpx)) / sqrtf(den2);
}
__device__ __forceinline__ float dca_wire(int d, int e, float x0, float y0, float tx, float ty) {
    return dca_line(x_bep(d, e), y_bep(d, e), z_bep(d, e), G(F_DELTAPX, d), G(F_DELTAPY, d), G(F_DELTAPZ, d), x0, y0,
                    tx, ty);
}

__device__ __forceinline__ float position(float pos, float drift, short s) { return pos + (float)s * drift; }

// ---- left/right of x hit pairs (reconstruction_helper.cuh:808-872) ---------
// A partner index at or past n counts as not adjacent.
// x-only decision for one adjacent pair: hypothesis index or -1
__device__ __forceinline__ int lr_xpair_decide(float x0, float tx, float ex0, float etx, int di, int dj, float posi,
                                               float posj, float dri, float drj, float thr) {
    int imin = -1;
    float pmin = 1.e6f;
    const float zi = G(F_Z, di), zj = G(F_Z, dj);
    for (int h = 0; h < 4; ++h) {
        const float pi = position(posi, dri, lr_si(h));
        const float sl = (pi - position(posj, drj, lr_sj(h))) / (zi - zj);
        const float it = pi - sl * zi;
        const float pull = sqrtf(((tx - sl) * (tx - sl)) / etx / etx + ((x0 - it) * (x0 - it)) / ex0 / ex0);
        if (pull < pmin) { imin = h; pmin = pull; }
    }
    return (imin > 0 && pmin < thr) ? imin : -1;
}

template <int N>
__device__ __forceinline__ void lr_xhits(float x0, float tx, float ex0, float etx, int n, const short* det,
                                         const float* pos, const float* drift, short* sign, float thr) {
    for (int k = 0; k < n; k += 2) {
        const int i = k, j = k + 1;
        if (j >= n || abs(det[i] - det[j]) != 1) { k--; continue; }
        if (sign[i] * sign[j] != 0) continue;
        const int h = lr_xpair_decide(x0, tx, ex0, etx, det[i], det[j], pos[i], pos[j], drift[i], drift[j], thr);
        if (h >= 0) { sign[i] = lr_si(h); sign[j] = lr_sj(h); }
    }
}

// reconstruction_helper.cuh:875-892
__device__ __forceinline__ void lr_single_xhits(float x0, float tx, int n, const short* det, const float* pos,
                                                short* sign) {
    for (int k = 0; k < n; ++k) {
        if (sign[k]) continue;
        const float pe = G(F_Z, det[k]) * tx + x0;
        sign[k] = pe > pos[k] ? 1 : -1;
    }
}

// ---- left/right of new hits (reconstruction_helper.cuh:714-804) ------------
// Left/right decision for one adjacent pair (i, j) of a candidate: the
// hypothesis index h (signs lr_si(h), lr_sj(h)) or -1 when the pair keeps its
// signs.  Depends only on the pair's hits and the track; lr_newhits applies it.
__device__ __forceinline__ int lr_pair_decide(float x0, float tx, float y0, float ty, float ex0, float etx, float ey0,
                                              float ety, int di, int dj, float posi, float posj, float dri, float drj,
                                              float thr) {
    const float ct = G(F_COSTHETA, di), st = G(F_SINTHETA, di);
    const float sexp = ct * tx + st * ty;
    const float esl = fabsf(ct * etx) + fabsf(st * ety);
    const float iexp = ct * x0 + st * y0;
    const float ein = fabsf(ct * ex0) + fabsf(st * ey0);
    const float zi = G(F_Z, di), zj = G(F_Z, dj);
    const float smax = G(F_SLOPE_MAX, di), imax = G(F_INTER_MAX, di);
    int imin = -1;
    float pmin = 1.e6f;
    for (int h = 0; h < 4; ++h) {
        const float pi = position(posi, dri, lr_si(h));
        const float sl = (pi - position(posj, drj, lr_sj(h))) / (zi - zj);
        const float it = pi - sl * zi;
        if (fabsf(sl) > smax || fabsf(it) > imax) continue;
        const float pull = sqrtf(((sexp - sl) * (sexp - sl)) / esl / esl + ((iexp - it) * (iexp - it)) / ein / ein);
        if (pull < pmin) { imin = h; pmin = pull; }
    }
    return (imin > 0 && pmin < thr) ? imin : -1;
}

// Pairing walk of the new-hits left/right step (reconstruction_helper.cuh:714-782):
// steps of two, back by one when the detectors are not adjacent.
__device__ __forceinline__ void lr_newhits(float x0, float tx, float y0, float ty, float ex0, float etx, float ey0,
                                           float ety, int n, const short* det, const float* pos, const float* drift,
                                           short* sign, float thr) {
    for (int k = 0; k < n; k += 2) {
        const int i = k, j = k + 1;
        if (j >= n || abs(det[i] - det[j]) != 1) { k--; continue; }
        if (sign[i] * sign[j] != 0) continue;
        const int h = lr_pair_decide(x0, tx, y0, ty, ex0, etx, ey0, ety, det[i], det[j], pos[i], pos[j], drift[i],
                                     drift[j], thr);
        if (h >= 0) { sign[i] = lr_si(h); sign[j] = lr_sj(h); }
    }
}

// Index i < n3 that the walk above pairs with an index j >= n3 (the first
// hit after the station-3 hits), or -1.
__device__ __forceinline__ int lr_cross_index(int n, int n3, const short* det) {
    for (int k = 0; k < n; k += 2) {
        const int i = k, j = k + 1;
        if (j >= n || abs(det[i] - det[j]) != 1) { k--; continue; }
        if (i < n3 && j >= n3) return i;
        if (i >= n3) return -1;
    }
    return -1;
}

__device__ __forceinline__ void lr_single_newhits(float x0, float tx, float y0, float ty, int n, const short* det,
                                                  const float* pos, short* sign) {
    for (int k = 0; k < n; ++k) {
        if (sign[k]) continue;
        const int d = det[k];
        if (d < 0) continue;
        const float pe = (G(F_Z, d) * tx + x0) * G(F_COSTHETA, d) + (G(F_Z, d) * ty + y0) * G(F_SINTHETA, d);
        sign[k] = pe > pos[k] ? 1 : -1;
    }
}

// ---- left/right on hits kept in fixed slots --------------------------------
// A candidate's hits sit in NS fixed slots in the reference's list order; bit s
// of pres says slot s holds a hit.  In every slot layout used here, detectors
// one apart only occur in neighbouring slots, so the pairing walk over the
// present hits (steps of two, back by one at a non-adjacent neighbour) pairs
// slot s with s+1 exactly when both are present, adjacent, and s was not taken
// as the second hit of the pair before.  All indices below are compile-time
// constants after unrolling, which keeps the slot arrays in registers.
template <int NS>
__device__ __forceinline__ void lr_slots_new(float x0, float tx, float y0, float ty, float ex0, float etx, float ey0,
                                             float ety, unsigned pres, const int (&det)[NS], const float (&pos)[NS],
                                             const float (&drift)[NS], int (&sign)[NS], float thr) {
    bool used = false;
#pragma unroll
    for (int s = 0; s + 1 < NS; ++s) {
        const bool start = !used && ((pres >> s) & 1u) && ((pres >> (s + 1)) & 1u) && abs(det[s] - det[s + 1]) == 1;
        if (start && sign[s] * sign[s + 1] == 0) {
            const int h = lr_pair_decide(x0, tx, y0, ty, ex0, etx, ey0, ety, det[s], det[s + 1], pos[s], pos[s + 1],
                                         drift[s], drift[s + 1], thr);
            if (h >= 0) { sign[s] = lr_si(h); sign[s + 1] = lr_sj(h); }
        }
        used = start;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (((pres >> s) & 1u) && sign[s] == 0) {
            const int d = det[s];
            const float pe = (G(F_Z, d) * tx + x0) * G(F_COSTHETA, d) + (G(F_Z, d) * ty + y0) * G(F_SINTHETA, d);
            sign[s] = pe > pos[s] ? 1 : -1;
        }
    }
}

// ---- hodoscope windows (reconstruction_helper.cuh:1002-1144) ---------------
// fudge factors and X cell widths, reconstruction_constants.h:73-74
__device__ __forceinline__ float hodo_fudge(int st) { return st == 0 ? 0.25f : st == 1 ? 0.2f : st == 2 ? 0.15f : 0.0f; }
__device__ __forceinline__ float hodo_xcell(int st) {
    return st == 0 ? 7.32f : st == 1 ? 13.0f : st == 2 ? 14.5875f : 19.6475f;
}

}  // namespace ordev

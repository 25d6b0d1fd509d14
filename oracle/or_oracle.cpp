/*
 * or_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's per-event online reconstruction chain
 * (efuchey/e1039-GPU, OnlineReconstruction/).  It is the checker for the HIP
 * product and the CPU baseline timed beside it.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Each function cites the reference file:line it restates.  Arithmetic keeps
 * the reference's operation order; build with -ffp-contract=off so that every
 * float operation rounds once, as the HIP build (also contraction-off) does.
 *
 * Where the reference is not deterministic the oracle fixes one reading:
 *  (R1) XZ tracklets are listed in the order (thread t, local bin i, combination
 *       i_x) of reconstruction_kernels.cuh:440-465 and handed to the 32 storage
 *       slots by the reference allocator (reconstruction_kernels.cuh:675-807)
 *       serialized in that order: the k-th D3p tracklet goes to slot 2*(k%16),
 *       the k-th D3m one to slot 2*(k%16)+1.
 *  (R2) The cross-slot cleaning pass (reconstruction_helper.cuh:1846-1909) runs
 *       slot 0..31 one after another.
 *  (R3) The pairwise left/right loops read index n+1 when n is the last hit
 *       (reconstruction_helper.cuh:723-727, 825-829); the oracle treats a
 *       partner at index >= nhits as not adjacent.
 *  (R4) Array bounds of the reference (200/60/120 staged hits per plane, 120
 *       reduced hits, 192-hit clusters, 100 hits per pair-builder plane, 100
 *       pairs per window list) become event-reject status bits.
 *  (R5) Station multiplicity sums the 8 rows the reducer writes
 *       (reconstruction_kernels.cuh:128-133 sums 32 rows, 24 never written).
 */
#include "../include/e1039_or.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <mutex>
#include <vector>
#include <omp.h>

namespace orc {

constexpr int ND = OR_N_DETECTORS;
constexpr int NP = OR_TRACK_PARAMS;

/* ---- constants: reconstruction_constants.h:29-162 -------------------------- */
constexpr float TX_MAX = 0.15f, TY_MAX = 0.1f, X0_MAX = 150.f, Y0_MAX = 50.f;
constexpr float INVP_MAX = 0.2f, INVP_MIN = 0.01f;
static const float kSpacingPlane[28] = {0., 0.40, 0.40, 0.40, 0.40, 0.40, 0.40, 1.3, 1.3, 1.3, 1.2, 1.2, 1.2, 1.2,
                                        1.2, 1.2, 4.0, 4.0, 7.0, 7.0, 8.0, 12.0, 12.0, 10.0, 3.0, 3.0, 3.0, 3.0};
static const short kSuperId[7][3] = {{2, 1, 3}, {5, 6, 4}, {8, 9, 7}, {11, 12, 10}, {14, 15, 13}, {25, 24, -1}, {26, 27, -1}};
static const short kPlaneType[15] = {1, 0, 2, 1, 0, 2, 2, 0, 1, 2, 0, 1, 2, 0, 1};
static const short kHodoX[4][2] = {{31, 32}, {37, 38}, {39, 40}, {45, 46}};
static const float kHodoFudge[4] = {0.25f, 0.2f, 0.15f, 0.0f};
static const float kHodoXCell[4] = {7.32f, 13.0f, 14.5875f, 19.6475f};
constexpr float Z_DUMP = 42.f, Z_TARGET = -300.f;
constexpr float SAG_T_C = 1.85f, SAG_T_W = 0.25f, SAG_D_C = 1.5f, SAG_D_W = 0.3f;
constexpr float PT_KICK_KMAG = -0.3819216f, FMAGSTR = -1.054f, Z_KMAG_BEND = 1064.26f;
static const short kLR[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
static const short kXBinLo[3][28] = {
    {1, 5, 9, 13, 17, 21, 25, 29, 33, 37, 41, 45, 49, 53, 57, 61, 65, 69, 73, 77, 81, 85, 89, 93, 97, 101, 105, 109},
    {1, 5, 9, 13, 17, 21, 25, 29, 33, 37, 41, 45, 49, 53, 57, 61, 65, 69, 73, 77, 81, 85, 89, 93, 97, 101, 105, 109},
    {1, 5, 9, 13, 17, 21, 25, 29, 33, 37, 41, 45, 49, 53, 57, 61, 65, 69, 73, 77, 81, 85, 89, 93, 97, 101, 105, 109}};
static const short kXBinHi[3][28] = {
    {4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84, 88, 92, 96, 100, 104, 108, 112},
    {4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84, 88, 92, 96, 100, 104, 108, 116},
    {4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84, 88, 92, 96, 100, 104, 108, 116}};
/* eR plane sets, reconstruction_constants.h:60 (E1039) and reconstruction_kernels.cuh:25-55 */
static const short kEffChambers[24] = {1, 2, 3, 4, 5, 6, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30};

/* ---- geometry: gPlane (reconstruction_classes.cuh:947-977) ---------------- */
struct Geo {
    float z[ND], nelem[ND], cellwidth[ND], spacing[ND], xoffset[ND], scalex[ND], x0[ND], x1[ND], x2[ND],
        costheta[ND], scaley[ND], y0[ND], y1[ND], y2[ND], sintheta[ND], resolution[ND], p1x[ND], p1y[ND], p1z[ND],
        deltapx[ND], deltapy[ND], deltapz[ND], dp1x[ND], dp1y[ND], dp1z[ND], slope_max[ND], inter_max[ND];
    float wire[ND][OR_MAX_ELEM];
};
static_assert(sizeof(float) * 27 * ND == offsetof(Geo, wire), "plane block layout");

/* online_reconstruction.cu:115-218 */
static int load_geometry(const char* path, Geo& g, float* deltaW_out /* 55x9, may be null */) {
    std::ifstream in(path);
    if (!in) return OR_ERR_IO;
    std::memset(&g, 0, sizeof g);
    float dW[ND][9];
    std::memset(dW, 0, sizeof dW);
    std::string line;
    /* declared outside the loop like the reference: a failed extraction keeps
       the previous line's value (online_reconstruction.cu:124-126) */
    int ipl = 0, nelem = 0;
    float z = 0, cw = 0, sp = 0, xo = 0, sx = 0, x0 = 0, x1 = 0, x2 = 0, ct = 0, sy = 0, y0 = 0, y1 = 0, y2 = 0, st = 0,
          res = 0, p1x = 0, p1y = 0, p1z = 0, dpx = 0, dpy = 0, dpz = 0, d1x = 0, d1y = 0, d1z = 0, w = 0;
    while (std::getline(in, line)) {
        if (line[0] == '#') continue;
        std::istringstream iss(line);
        iss >> ipl >> z >> nelem >> cw >> sp >> xo >> sx >> x0 >> x1 >> x2 >> ct >> sy >> y0 >> y1 >> y2 >> st >> res >>
            p1x >> p1y >> p1z >> dpx >> dpy >> dpz >> d1x >> d1y >> d1z;
        if (ipl < 0 || ipl >= ND) return OR_ERR_PARSE;
        g.z[ipl] = z; g.nelem[ipl] = (float)nelem; g.cellwidth[ipl] = cw; g.spacing[ipl] = sp; g.xoffset[ipl] = xo;
        g.scalex[ipl] = sx; g.x0[ipl] = x0; g.x1[ipl] = x1; g.x2[ipl] = x2; g.costheta[ipl] = ct; g.scaley[ipl] = sy;
        g.y0[ipl] = y0; g.y1[ipl] = y1; g.y2[ipl] = y2; g.sintheta[ipl] = st;
        g.resolution[ipl] = res; /* USE_DET_RESOL (online_reconstruction.cu:150-151) */
        g.p1x[ipl] = p1x; g.p1y[ipl] = p1y; g.p1z[ipl] = p1z; g.deltapx[ipl] = dpx; g.deltapy[ipl] = dpy;
        g.deltapz[ipl] = dpz; g.dp1x[ipl] = d1x; g.dp1y[ipl] = d1y; g.dp1z[ipl] = d1z;
        if (ipl > 46) { for (int k = 0; k < 9; ++k) { iss >> w; dW[ipl][k] = w; } }
        else { iss >> w; dW[ipl][0] = w; }
        g.slope_max[ipl] = ct * TX_MAX + st * TY_MAX;
        g.inter_max[ipl] = ct * X0_MAX + st * Y0_MAX;
        if (ipl % 2 == 0 && ipl > 1) { /* online_reconstruction.cu:179-186 */
            double dslope = (g.resolution[ipl] + g.resolution[ipl - 1]) / (g.z[ipl] - g.z[ipl - 1]);
            double dinter = dslope * g.z[ipl];
            g.slope_max[ipl] += dslope; g.inter_max[ipl] += dinter;
            g.slope_max[ipl - 1] += dslope; g.inter_max[ipl - 1] += dinter;
        }
    }
    /* wire table, online_reconstruction.cu:191-218 */
    for (int i = 1; i <= 30; ++i)
        for (int j = 1; j <= (int)g.nelem[i] && j < OR_MAX_ELEM; ++j) {
            double pos = (j - (g.nelem[i] + 1.) / 2.) * g.spacing[i] + g.xoffset[i] + g.x0[i] * g.costheta[i] +
                         g.y0[i] * g.sintheta[i] + dW[i][0];
            g.wire[i][j] = (float)pos;
        }
    for (int i = 31; i <= 46; ++i)
        for (int j = 1; j <= (int)g.nelem[i] && j < OR_MAX_ELEM; ++j) {
            double pos = g.x0[i] * g.costheta[i] + g.y0[i] * g.sintheta[i] + g.xoffset[i] +
                         (j - ((int)g.nelem[i] + 1) / 2.) * g.spacing[i] + dW[i][0];
            g.wire[i][j] = (float)pos;
        }
    for (int i = 47; i <= 54; ++i)
        for (int j = 1; j <= (int)g.nelem[i] && j < OR_MAX_ELEM; ++j) {
            int module = 8 - int((j - 1) / 8);
            double pos = g.x0[i] * g.costheta[i] + g.y0[i] * g.sintheta[i] + g.xoffset[i] +
                         (j - ((int)g.nelem[i] + 1) / 2.) * g.spacing[i] + dW[i][module];
            g.wire[i][j] = (float)pos;
        }
    if (deltaW_out) std::memcpy(deltaW_out, dW, sizeof dW);
    return OR_OK;
}

/* ---- wire end points, reconstruction_helper.cuh:8-36 ----------------------- */
static inline float x_bep(const Geo& g, int d, int e) { return g.p1x[d] + g.dp1x[d] * (e - 1); }
static inline float y_bep(const Geo& g, int d, int e) { return g.p1y[d] + g.dp1y[d] * (e - 1); }
static inline float z_bep(const Geo& g, int d, int e) { return g.p1z[d] + g.dp1z[d] * (e - 1); }

/* reconstruction_helper.cuh:75-100 */
static bool y_uvhit(const Geo& g, int d, int e, float drift, short hs, float x0, float tx, float& y, float& ey) {
    float p1x = x_bep(g, d, e);
    float p2x = p1x + g.deltapx[d];
    float xt = x0 + g.z[d] * tx;
    if (xt - drift > p1x && xt - drift > p2x) return false;
    if (xt + drift < p1x && xt + drift < p2x) return false;
    float p1y = y_bep(g, d, e);
    y = p1y + ((xt - p1x) * g.deltapy[d]) / g.deltapx[d];
    if (hs != 0) y += (std::copysign(1.f, g.sintheta[d]) * 2.f) * (drift - 0.5f * hs);
    y = std::fmax(y, p1y);
    y = std::fmin(y, p1y + g.deltapy[d]);
    ey = g.resolution[d] * std::fabs(g.deltapy[d] / g.deltapx[d]);
    return true;
}

/* reconstruction_helper.cuh:140-158 */
static void xz_fy(const Geo& g, int d, int e, float drift, short hs, float y0, float ty, float& x, float& z) {
    float p1x = x_bep(g, d, e), p1y = y_bep(g, d, e), p1z = z_bep(g, d, e);
    float dpy = g.deltapy[d], dpz = g.deltapz[d];
    z = p1z - (p1y * dpz) / dpy;
    float yt = y0 + z * ty;
    x = p1x + ((yt - p1y) * g.deltapx[d]) / dpy;
    z = p1z + ((yt - p1y) * dpz) / dpy;
    if (hs != 0) x += hs * drift;
}

/* ---- fits: trackanalyticalminimizer.cuh:32-119, trackfittingtools.cuh:3-82 */
static void matinv2(const float* A, float* Ai) {
    float a00 = A[0], a10 = A[1], a01 = A[2], a11 = A[3], tmp;
    int p0 = 0, p1 = 1;
    if (std::fabs(a10) > std::fabs(a00)) {
        tmp = a00; a00 = a10; a10 = tmp;
        tmp = a01; a01 = a11; a11 = tmp;
        p0 = 1; p1 = 0;
    }
    tmp = 1.0f / a00;
    int c0 = p0;
    a00 = tmp;
    a01 = tmp * a01;
    tmp = a10;
    a10 = (-tmp) * a00;
    a11 = -(tmp * a01) + a11;
    tmp = 1.0f / a11;
    int c1 = p1;
    a10 = tmp * a10;
    a11 = tmp;
    tmp = a01;
    a00 = -(tmp * a10) + a00;
    a01 = (-tmp) * a11;
    Ai[0 * 2 + c0] = a00; Ai[1 * 2 + c0] = a10; Ai[0 * 2 + c1] = a01; Ai[1 * 2 + c1] = a11;
}

/* out: par0, par1, err0, err1 */
static void fit2d(int n, const float* x, const float* z, const float* w, float* out) {
    float A[4] = {0, 0, 0, 0}, B[2] = {0, 0}, Ai[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        B[0] += w[i] * x[i];
        B[1] += (w[i] * x[i]) * z[i];
        A[0] += w[i];
        A[1] += w[i] * z[i];
        A[2] += w[i] * z[i];
        A[3] += (w[i] * z[i]) * z[i];
    }
    matinv2(A, Ai);
    for (int j = 0; j < 2; ++j) {
        out[2 + j] = std::sqrt(std::fabs(Ai[j * 2 + j]));
        float p = 0.0f;
        p += Ai[j * 2 + 0] * B[0];
        p += Ai[j * 2 + 1] * B[1];
        out[j] = p;
    }
}

static void fit2d_drift(int n, const float* x, const float* drift, const short* sign, const float* z, const float* w,
                        float* out) {
    float xc[18];
    for (int i = 0; i < n; ++i) xc[i] = x[i] + drift[i] * sign[i];
    fit2d(n, xc, z, w, out);
}

/* ---- left/right: reconstruction_helper.cuh:714-892 -------------------------- */
static inline float position(float pos, float drift, short s) { return pos + s * drift; }

static void lr_xhits(float x0, float tx, float ex0, float etx, int n, const short* det, const float* pos,
                     const float* drift, short* sign, const Geo& g, float thr) {
    for (int k = 0; k < n; k += 2) {
        int i = k, j = k + 1;
        if (j >= n || std::abs(det[i] - det[j]) != 1) { k--; continue; } /* (R3) */
        if (sign[i] * sign[j] != 0) continue;
        int imin = -1;
        float pmin = 1.e6f;
        for (int h = 0; h < 4; ++h) {
            float sl = (position(pos[i], drift[i], kLR[h][0]) - position(pos[j], drift[j], kLR[h][1])) /
                       (g.z[det[i]] - g.z[det[j]]);
            float it = position(pos[i], drift[i], kLR[h][0]) - sl * g.z[det[i]];
            float pull = std::sqrt(((tx - sl) * (tx - sl)) / etx / etx + ((x0 - it) * (x0 - it)) / ex0 / ex0);
            if (pull < pmin) { imin = h; pmin = pull; }
        }
        if (imin > 0 && pmin < thr) { sign[i] = kLR[imin][0]; sign[j] = kLR[imin][1]; }
    }
}

static void lr_single_xhits(float x0, float tx, int n, const short* det, const float* pos, short* sign, const Geo& g) {
    for (int k = 0; k < n; ++k) {
        if (sign[k]) continue;
        float pe = g.z[det[k]] * tx + x0;
        sign[k] = pe > pos[k] ? 1 : -1;
    }
}

static void lr_newhits(float x0, float tx, float y0, float ty, float ex0, float etx, float ey0, float ety, int n,
                       const short* det, const float* pos, const float* drift, short* sign, const Geo& g, float thr) {
    for (int k = 0; k < n; k += 2) {
        int i = k, j = k + 1;
        if (j >= n || std::abs(det[i] - det[j]) != 1) { k--; continue; } /* (R3) */
        int di = det[i];
        float sexp = g.costheta[di] * tx + g.sintheta[di] * ty;
        float esl = std::fabs(g.costheta[di] * etx) + std::fabs(g.sintheta[di] * ety);
        float iexp = g.costheta[di] * x0 + g.sintheta[di] * y0;
        float ein = std::fabs(g.costheta[di] * ex0) + std::fabs(g.sintheta[di] * ey0);
        if (sign[i] * sign[j] != 0) continue;
        int imin = -1;
        float pmin = 1.e6f;
        for (int h = 0; h < 4; ++h) {
            float sl = (position(pos[i], drift[i], kLR[h][0]) - position(pos[j], drift[j], kLR[h][1])) /
                       (g.z[di] - g.z[det[j]]);
            float it = position(pos[i], drift[i], kLR[h][0]) - sl * g.z[di];
            if (std::fabs(sl) > g.slope_max[di] || std::fabs(it) > g.inter_max[di]) continue;
            float pull = std::sqrt(((sexp - sl) * (sexp - sl)) / esl / esl + ((iexp - it) * (iexp - it)) / ein / ein);
            if (pull < pmin) { imin = h; pmin = pull; }
        }
        if (imin > 0 && pmin < thr) { sign[i] = kLR[imin][0]; sign[j] = kLR[imin][1]; }
    }
}

static void lr_single_newhits(float x0, float tx, float y0, float ty, int n, const short* det, const float* pos,
                              short* sign, const Geo& g) {
    for (int k = 0; k < n; ++k) {
        if (sign[k]) continue;
        int d = det[k];
        if (d < 0) continue;
        float pe = (g.z[d] * tx + x0) * g.costheta[d] + (g.z[d] * ty + y0) * g.sintheta[d];
        sign[k] = pe > pos[k] ? 1 : -1;
    }
}

/* ---- hodoscope windows: reconstruction_helper.cuh:1002-1144 ---------------- */
static bool hodo_xz(int stid, int d, int n, const float* hpos, float x0, float tx, float ex0, float etx, const Geo& g) {
    if (n == 0) return false;
    if (d > 40) stid = 3;
    float xh = g.z[d] * tx + x0;
    float f = kHodoFudge[stid], cw = kHodoXCell[stid];
    float ex = 3.f * (std::fabs(g.z[d] * etx) + ex0);
    for (int i = 0; i < n; ++i) {
        float xmin = hpos[i] - cw * 0.5f, xmax = hpos[i] + cw * 0.5f;
        xmin -= (xmax - xmin) * f;
        xmax += (xmax - xmin) * f;
        ex += (xmax - xmin) * 0.15f;
        xmin -= ex;
        xmax += ex;
        if (xmin <= xh && xh <= xmax) return true;
    }
    return false;
}

static bool hodo_tkl(int stid, int d, int n, const float* hpos, float x0, float y0, float tx, float ty, float ex0,
                     float ey0, float etx, float ety, const Geo& g) {
    if (n == 0) return false;
    if (d > 40) stid = 3;
    float f = kHodoFudge[stid], cw = g.cellwidth[d];
    float xh = g.z[d] * tx + x0, yh = g.z[d] * ty + y0;
    float ex = 3.f * (std::fabs(g.z[d] * etx) + ex0);
    float ey = 3.f * (std::fabs(g.z[d] * ety) + ey0);
    for (int i = 0; i < n; ++i) {
        float xmin = hpos[i] - cw * 0.5f, xmax = hpos[i] + cw * 0.5f;
        float ymin = g.y1[d], ymax = g.y2[d];
        xmin -= (xmax - xmin) * f;
        xmax += (xmax - xmin) * f;
        ymin -= (ymax - ymin) * f;
        ymax += (ymax - ymin) * f;
        ex += (xmax - xmin) * 0.15f;
        xmin -= ex; xmax += ex; ymin -= ey; ymax += ey;
        if (xmin <= xh && xh <= xmax && ymin <= yh && yh <= ymax) return true;
    }
    return false;
}

/* ---- residuals: reconstruction_helper.cuh:1149-1204 ------------------------- */
static inline float dca_line(float p1x, float p1y, float p1z, float dpx, float dpy, float dpz, float x0, float y0,
                             float tx, float ty) {
    float den2 = (dpy * dpy) * (1 + tx * tx) + (dpx * dpx) * (1 + ty * ty) -
                 2 * (((ty * dpx) * dpz + (ty * dpy) * dpz) + ((tx * ty) * dpx) * dpy);
    return ((ty * dpz - dpy) * (p1x - x0) + (dpx - tx * dpz) * (p1y - y0) + p1z * (tx * dpy - ty * dpx)) /
           std::sqrt(den2);
}

/* ---- reducer: reconstruction_helper.cuh:168-361 ----------------------------
 * Returns kept count; keep[i] = 1 kept, -1 dropped.  *cl_over set when a
 * chamber cluster outgrows 192 entries (R4). */
static int reduce_plane(int det, int n, const float* chan, const float* pos, const float* tdc, const float* flag,
                        const float* drift, signed char* keep, bool* cl_over) {
    int cl[400];
    int ncl = 0;
    for (int i = 0; i < n; ++i) {
        keep[i] = 1;
        if ((int(flag[i]) & 2) == 0) { keep[i] = -1; continue; }
        if (det > 30) continue; /* declustering is for chambers only (reconstruction_helper.cuh:234) */
        if (ncl == 0) { cl[0] = i; ncl = 1; continue; }
        bool adj = std::fabs(chan[i] - chan[cl[ncl - 1]]) <= 1;
        if (adj && i != n - 1) {
            if (ncl >= 192) *cl_over = true;
            cl[ncl++] = i;
            continue;
        }
        if (i == n - 1 && adj) {
            if (ncl >= 192) *cl_over = true;
            cl[ncl++] = i;
        }
        if (ncl == 2) {
            int a = cl[0], b = cl[1];
            float wmax = 0.9 * 0.5 * (pos[b] - pos[a]);
            float wmin = 4.0 / 9.0 * wmax;
            if ((drift[a] > wmax && drift[b] > wmin) || (drift[a] > wmin && drift[b] > wmax)) {
                if (drift[a] > drift[b]) keep[a] = -1;
                else keep[b] = -1;
            } else {
                float dt = tdc[a] - tdc[b];
                if (((dt >= 0.0 && dt < 8.0) || (dt <= 0.0 && dt > -8.0)) && det >= 19 && det <= 24) {
                    keep[a] = -1;
                    keep[b] = -1;
                }
            }
        }
        if (ncl >= 3) {
            float dtm = 0.0f;
            for (int c = 1; c < ncl; ++c) {
                float d = tdc[cl[c]] - tdc[cl[c - 1]];
                dtm += (d > 0.0 ? d : (tdc[cl[c - 1]] - tdc[cl[c]]));
            }
            dtm = dtm / (ncl - 1);
            if (dtm < 10.0) { for (int c = 0; c < ncl; ++c) keep[cl[c]] = -1; }
            else { for (int c = 1; c < ncl; ++c) keep[cl[c]] = -1; }
        }
        ncl = 0;
    }
    bool has[256];
    std::memset(has, 0, sizeof has);
    int nred = 0;
    for (int i = 0; i < n; ++i) {
        if (keep[i] > 0) {
            int c = (short)chan[i];
            if (has[c]) keep[i] = -1;
            else { has[c] = true; ++nred; }
        }
    }
    return nred;
}

/* ---- per-event state -------------------------------------------------------- */
struct Plane {           /* reduced hits of one plane, staging order */
    int n = 0;
    float chan[400], pos[400], tdc[400], drift[400];
    int orig[400];
};

struct Pair { int a, b; };

struct Tkl {             /* one tracklet: the reference 148-float record + bookkeeping */
    float rec[NP];
    int orig[OR_MAX_HITS_PER_TRACK];
    int t, st3, canon, slot, pos_in_slot;
};

struct Ctx {
    const Geo* g;
    const or_params* p;
    int64_t ev = 0;
    Plane pl[ND];
    std::vector<Tkl> tk;
    uint32_t status;
};

/* make_hitpairs_in_station_bins, reconstruction_helper.cuh:368-460, over all 28 bins */
static void pairs_binned(const Plane& h1, const Plane& h2, int stid, int sbin, int cap, Pair* out /*28*cap*/,
                         int* cnt) {
    int sid = kSuperId[stid][0];
    std::vector<char> f1(h1.n, 0), f2(h2.n, 0);
    for (int b = 0; b < 28; ++b) cnt[b] = 0;
    auto put = [&](float c, int a, int bb) {
        for (int b = 0; b < 28; ++b)
            if (kXBinLo[sbin][b] <= c && c <= kXBinHi[sbin][b]) {
                if (cnt[b] < cap) out[b * cap + cnt[b]] = {a, bb};
                cnt[b]++;
            }
    };
    for (int i = 0; i < h1.n; ++i)
        for (int j = 0; j < h2.n; ++j) {
            if (std::fabs(h1.pos[i] - h2.pos[j]) > kSpacingPlane[sid]) continue;
            put(h1.chan[i], i, j);
            f1[i] = 1;
            f2[j] = 1;
        }
    for (int i = 0; i < h1.n; ++i)
        if (!f1[i]) put(h1.chan[i], i, -1);
    for (int j = 0; j < h2.n; ++j)
        if (!f2[j]) put(h2.chan[j], -1, j);
}

/* make_hitpairs_in_station, reconstruction_helper.cuh:464-549; returns -1 on overflow (R4) */
static int pairs_window(const Ctx& C, int stid, int proj, float xmin, float xmax, Pair* out, int cap) {
    const Geo& g = *C.g;
    int sid = kSuperId[stid][proj];
    int d1 = sid * 2, d2 = sid * 2 - 1;
    const Plane &h1 = C.pl[d1], &h2 = C.pl[d2];
    const int lim = C.p->max_pair_plane_hits; /* hitidx1/2[100] hold the in-window hits */
    int i1[400], i2[400], n1 = 0, n2 = 0;
    for (int i = 0; i < h1.n; ++i) {
        float a, b;
        int e = (int)h1.chan[i];
        if (g.deltapx[d1] > 0) { a = x_bep(g, d1, e); b = x_bep(g, d1, e) + g.deltapx[d1]; }
        else { a = x_bep(g, d1, e) + g.deltapx[d1]; b = x_bep(g, d1, e); }
        if (a <= xmax && b >= xmin) i1[n1++] = i;
    }
    for (int i = 0; i < h2.n; ++i) {
        float a, b;
        int e = (int)h2.chan[i];
        if (g.deltapx[d2] > 0) { a = x_bep(g, d2, e); b = x_bep(g, d2, e) + g.deltapx[d2]; }
        else { a = x_bep(g, d2, e) + g.deltapx[d2]; b = x_bep(g, d2, e); }
        if (a <= xmax && b >= xmin) i2[n2++] = i;
    }
    if (n1 > lim || n2 > lim) return -1;
    char f1[400] = {0}, f2[400] = {0};
    int np = 0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            if (std::fabs(h1.pos[i1[i]] - h2.pos[i2[j]]) > kSpacingPlane[sid]) continue;
            if (np >= cap) return -1;
            out[np++] = {i1[i], i2[j]};
            f1[i] = f2[j] = 1;
        }
    for (int i = 0; i < n1; ++i)
        if (!f1[i]) { if (np >= cap) return -1; out[np++] = {i1[i], -1}; }
    for (int j = 0; j < n2; ++j)
        if (!f2[j]) { if (np >= cap) return -1; out[np++] = {-1, i2[j]}; }
    return np;
}

/* ---- XZ: reconstruction_kernels.cuh:182-855 ------------------------------- */
struct XHit { short det; int idx; };

static void run_xz(Ctx& C) {
    const Geo& g = *C.g;
    const or_params& P = *C.p;
    const int cap = P.max_pairs_per_bin;
    static const int kDets[3][2] = {{16, 15}, {22, 21}, {28, 27}};
    for (int s = 0; s < 3; ++s)
        for (int k = 0; k < 2; ++k)
            if (C.pl[kDets[s][k]].n > P.max_pair_plane_hits) { C.status |= OR_EV_PAIR_OVERFLOW; return; }
    std::vector<Pair> pr[3];
    int cnt[3][28];
    for (int s = 0; s < 3; ++s) {
        pr[s].resize(28 * cap);
        pairs_binned(C.pl[kDets[s][0]], C.pl[kDets[s][1]], 2 + s, s, cap, pr[s].data(), cnt[s]);
        for (int b = 0; b < 28; ++b)
            if (cnt[s][b] > cap) C.status |= OR_EV_XZ_BIN_OVERFLOW;
    }
    if (C.status) return;
    /* prop X planes (detsuperid[5][0]=25, [6][0]=26) and hodo X planes */
    const int P1a = 50, P1b = 49, P2a = 52, P2b = 51;
    int canon = 0;
    for (int t = 0; t < 32; ++t) {
        const int st3 = 3 + t % 2;
        const int s3 = st3 - 2; /* index into pr/cnt: 1 = D3p, 2 = D3m */
        for (int i = 0; i < 49; ++i) {
            int b2 = (t / 8) * 7 + i % 7;
            int b3 = ((t % 8) / 2) * 7 + i / 7;
            int nx2 = cnt[0][b2], nx3 = cnt[s3][b3];
            if (nx2 == 0 || nx3 == 0) continue;
            for (int ix = 0; ix < nx2 * nx3; ++ix) {
                int ix2 = ix % nx2, ix3 = ix / nx2;
                Pair q2 = pr[0][b2 * cap + ix2], q3 = pr[s3][b3 * cap + ix3];
                short det[4]; float X[4], Z[4], ex[4], drift[4], tdc[4]; short sign[4], el[4]; int orig[4];
                int n = 0;
                auto add = [&](int d, int idx) {
                    const Plane& h = C.pl[d];
                    det[n] = (short)d; ex[n] = g.resolution[d]; el[n] = (short)h.chan[idx]; tdc[n] = h.tdc[idx];
                    drift[n] = h.drift[idx]; sign[n] = 0; orig[n] = h.orig[idx];
                    xz_fy(g, d, el[n], drift[n], sign[n], 0, 0, X[n], Z[n]);
                    ++n;
                };
                if (q2.a >= 0) add(16, q2.a);
                if (q2.b >= 0) add(15, q2.b);
                int n2 = n;
                if (n2 == 0) continue;
                if (q3.a >= 0) add(kDets[s3][0], q3.a);
                if (q3.b >= 0) add(kDets[s3][1], q3.b);
                if (n - n2 == 0) continue;
                float f[4];
                fit2d(n, X, Z, ex, f);
                float x0 = f[0], tx = f[1], ex0 = f[2], etx = f[3];
                if (std::fabs(x0) > X0_MAX + 2 * ex0 || std::fabs(tx) > TX_MAX + 2 * etx) continue;
                lr_xhits(x0, tx, ex0, etx, n, det, X, drift, sign, g, 150.f);
                lr_single_xhits(x0, tx, n, det, X, sign, g);
                fit2d_drift(n, X, drift, sign, Z, ex, f);
                x0 = f[0]; tx = f[1]; ex0 = f[2]; etx = f[3];
                if (std::fabs(x0) > X0_MAX || std::fabs(tx) > TX_MAX) continue;
                int nprop = 0;
                bool next = true;
                for (int m = 0; m < C.pl[P1a].n; ++m)
                    if (std::fabs(C.pl[P1a].pos[m] - (tx * g.z[P1a] + x0)) < 5.08f) { nprop++; next = false; break; }
                if (next)
                    for (int m = 0; m < C.pl[P1b].n; ++m)
                        if (std::fabs(C.pl[P1b].pos[m] - (tx * g.z[P1b] + x0)) < 5.08f) { nprop++; break; }
                next = true;
                for (int m = 0; m < C.pl[P2a].n; ++m)
                    if (std::fabs(C.pl[P2a].pos[m] - (tx * g.z[P2a] + x0)) < 7.62f) { nprop++; next = false; break; }
                if (next)
                    for (int m = 0; m < C.pl[P2b].n; ++m)
                        if (std::fabs(C.pl[P2b].pos[m] - (tx * g.z[P2b] + x0)) < 7.62f) { nprop++; break; }
                if (nprop < 2) continue;
                bool ok = true;
                for (int hs = 1; hs <= 3 && ok; ++hs) {
                    int da = kHodoX[hs][0], db = kHodoX[hs][1];
                    bool m = hodo_xz(hs, da, C.pl[da].n, C.pl[da].pos, x0, tx, ex0, etx, g);
                    if (!m) m = hodo_xz(hs, db, C.pl[db].n, C.pl[db].pos, x0, tx, ex0, etx, g);
                    ok = m;
                }
                if (!ok) continue;
                Tkl T;
                std::memset(&T, 0, sizeof T);
                T.t = t; T.st3 = st3; T.canon = canon++;
                T.rec[1] = (float)t;
                T.rec[2] = (float)n;
                T.rec[5] = tx; T.rec[7] = x0; T.rec[10] = etx; T.rec[12] = ex0;
                for (int h = 0; h < n; ++h) {
                    T.rec[22 + h] = det[h]; T.rec[40 + h] = el[h]; T.rec[58 + h] = X[h]; T.rec[76 + h] = drift[h];
                    T.rec[94 + h] = sign[h]; T.rec[112 + h] = tdc[h]; T.rec[130 + h] = 0.0f;
                    T.orig[h] = orig[h];
                }
                C.tk.push_back(T);
            }
        }
    }
    /* slot allocation (R1) and capacity flags (reconstruction_kernels.cuh:719,824,835) */
    int kp = 0, km = 0;
    for (auto& T : C.tk) {
        int& k = (T.st3 == 3) ? kp : km;
        T.slot = 2 * (k % 16) + (T.st3 == 3 ? 0 : 1);
        T.pos_in_slot = k / 16;
        ++k;
    }
    if (kp > 16 * P.max_tracklets_per_slot || km > 16 * P.max_tracklets_per_slot ||
        (int)C.tk.size() >= P.max_tracklets_per_event)
        C.status |= OR_EV_TRACKLET_OVERFLOW;
    std::stable_sort(C.tk.begin(), C.tk.end(), [](const Tkl& a, const Tkl& b) {
        return a.slot != b.slot ? a.slot < b.slot : a.pos_in_slot < b.pos_in_slot;
    });
}

/* ---- YZ: reconstruction_kernels.cuh:864-1832, one tracklet ---------------- */
/* ---- check of the YZ sign-state model the device relies on ------------------
 * Within one (u3, v3) unit the device derives each candidate's starting
 * station-3 signs instead of carrying them: zero for the unit's first
 * first-box survivor; afterwards the signs that survivor's left/right left,
 * with the last station-3 sign replaced by the latest earlier survivor whose
 * station-3/station-2 cross pair took a left/right decision.  When enabled,
 * run_yz_one compares that prediction with the carried state before and after
 * every left/right step.  Counters: checks, mismatches, cross-pair cases. */
/* ---- candidate trace (test infrastructure): every YZ and station-1 candidate
 * that reaches its chi2, in evaluation order, each tracklet's list sizes and
 * its winner.  kind 0 YZ candidate (iu3, iv3, iu2, iv2), 1 YZ winner, 4 YZ
 * list sizes (nu3, nv3, nu2, nv2); 2 station-1 candidate (ix, iu, iv), 3 its
 * winner, 5 station-1 list sizes (nx1, nu1, nv1). */
struct TraceRow { int kind; int64_t ev; int canon; int i[4]; float chi2; };
static std::atomic<int> g_trace_on{0};
static std::mutex g_trace_mu;
static std::vector<TraceRow> g_trace;
static void trace(int kind, const Ctx& C, int canon, int a, int b, int c, int d, float chi2) {
    if (!g_trace_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    g_trace.push_back({kind, C.ev, canon, {a, b, c, d}, chi2});
}

static std::atomic<int> g_yzcheck{0};
static std::atomic<long long> g_yzcnt[3];

static int yzcheck_pair(float x0, float tx, float y0, float ty, float ex0, float etx, float ey0, float ety, int i,
                        int j, const short* det, const float* pos, const float* drift, const Geo& g, float thr) {
    int di = det[i];
    float sexp = g.costheta[di] * tx + g.sintheta[di] * ty;
    float esl = std::fabs(g.costheta[di] * etx) + std::fabs(g.sintheta[di] * ety);
    float iexp = g.costheta[di] * x0 + g.sintheta[di] * y0;
    float ein = std::fabs(g.costheta[di] * ex0) + std::fabs(g.sintheta[di] * ey0);
    int imin = -1;
    float pmin = 1.e6f;
    for (int h = 0; h < 4; ++h) {
        float sl = (position(pos[i], drift[i], kLR[h][0]) - position(pos[j], drift[j], kLR[h][1])) /
                   (g.z[di] - g.z[det[j]]);
        float it = position(pos[i], drift[i], kLR[h][0]) - sl * g.z[di];
        if (std::fabs(sl) > g.slope_max[di] || std::fabs(it) > g.inter_max[di]) continue;
        float pull = std::sqrt(((sexp - sl) * (sexp - sl)) / esl / esl + ((iexp - it) * (iexp - it)) / ein / ein);
        if (pull < pmin) { imin = h; pmin = pull; }
    }
    return (imin > 0 && pmin < thr) ? imin : -1;
}

static int yzcheck_cross(const short* det, int n, int n3) {  /* i < n3 paired with j >= n3 by the walk, or -1 */
    for (int k = 0; k < n; k += 2) {
        int i = k, j = k + 1;
        if (j >= n || std::abs(det[i] - det[j]) != 1) { k--; continue; }
        if (i < n3 && j >= n3) return i;
    }
    return -1;
}

/* Unit filter (test aid, no reference counterpart): the YZ search of a tracklet
 * is restricted to the (u3, v3) units listed for it, as a pruned search leaves
 * them; a tracklet with no listed unit searches all of them.  Keys, sorted:
 * event << 36 | tracklet_id << 20 | unit, where unit = ku3 + kv3 * nku3 counts
 * only the U3 and V3 window-list entries with an accepted hit (the enumeration
 * a pruned search numbers its units by). */
static const int64_t* g_unit_keys = nullptr;
static int64_t g_unit_nkeys = 0;
static int64_t unit_key(int64_t ev, int canon, int unit) { return (ev << 36) | ((int64_t)canon << 20) | unit; }
static bool unit_listed(int64_t key) { return std::binary_search(g_unit_keys, g_unit_keys + g_unit_nkeys, key); }

static bool run_yz_one(Ctx& C, Tkl& T) {
    const Geo& g = *C.g;
    float* R = T.rec;
    const float x0 = R[7], tx = R[5], ex0 = R[12], etx = R[10];
    const int nx = (int)R[2];
    const int st3 = T.st3;
    R[0] = 3;
    /* views: 0 u2, 1 v2, 2 u3, 3 v3 (reconstruction_kernels.cuh:1252-1283) */
    const int vst[4] = {2, 2, st3, st3}, vpj[4] = {1, 2, 1, 2};
    const int wcap = std::max(C.p->max_pairs_window, 1);
    std::vector<Pair> lists[4];
    for (auto& l : lists) l.resize(wcap);
    int nl[4];
    for (int v = 0; v < 4; ++v) {
        int sid = kSuperId[vst[v]][vpj[v]];
        int da = sid * 2, db = sid * 2 - 1;
        float za = g.z[da], zb = g.z[db];
        float xmin = std::fmin((tx - etx) * za, (tx - etx) * zb) + x0 - ex0 - g.spacing[da];
        float xmax = std::fmax((tx + etx) * za, (tx + etx) * zb) + x0 + ex0 + g.spacing[da];
        nl[v] = pairs_window(C, vst[v], vpj[v], xmin, xmax, lists[v].data(), C.p->max_pairs_window);
        if (nl[v] < 0) { C.status |= OR_EV_PAIR_OVERFLOW; return false; }
    }
    const int nu2 = nl[0], nv2 = nl[1], nu3 = nl[2], nv3 = nl[3];
    trace(4, C, T.canon, nu3, nv3, nu2, nv2, 0.f);
    /* enumeration bound of the device (32-bit candidate indices): more than
     * 2^27 (u3, v3, u2, v2) list combinations for one tracklet is an overflow */
    if ((long long)nu2 * nv2 * nu3 * nv3 > (1ll << 27)) { C.status |= OR_EV_PAIR_OVERFLOW; return false; }
    if (nu2 == 0 || nu3 == 0) return true;
    if (nv2 == 0 || nv3 == 0) return true;
    float dd[12], res[12], p1x[12], p1y[12], p1z[12], dpx[12], dpy[12], dpz[12], rsd[12];
    short hs[12];
    for (int m = 0; m < nx; ++m) {
        int d = (int)R[22 + m], e = (int)R[40 + m];
        dd[m] = R[76 + m]; hs[m] = (short)R[94 + m]; res[m] = g.resolution[d];
        p1x[m] = x_bep(g, d, e); p1y[m] = y_bep(g, d, e); p1z[m] = z_bep(g, d, e);
        dpx[m] = g.deltapx[d]; dpy[m] = g.deltapy[d]; dpz[m] = g.deltapz[d];
    }
    const int vdet[4][2] = {{kSuperId[2][1] * 2, kSuperId[2][1] * 2 - 1},
                            {kSuperId[2][2] * 2, kSuperId[2][2] * 2 - 1},
                            {kSuperId[st3][1] * 2, kSuperId[st3][1] * 2 - 1},
                            {kSuperId[st3][2] * 2, kSuperId[st3][2] * 2 - 1}};
    short det[8], el[8], sign[8];
    float Y[8], eY[8], Z[8], pos[8], drift[8], tdc[8];
    int orig[8];
    int nuv = 0;
    auto add = [&](int d, int idx) {
        const Plane& h = C.pl[d];
        det[nuv] = (short)d; Z[nuv] = g.z[d]; el[nuv] = (short)h.chan[idx]; tdc[nuv] = h.tdc[idx]; sign[nuv] = 0;
        drift[nuv] = h.drift[idx]; pos[nuv] = h.pos[idx];
        float y, ey;
        if (y_uvhit(g, d, el[nuv], drift[nuv], 0, x0, tx, y, ey)) {
            int k = nx + nuv;
            res[k] = g.resolution[d]; dd[k] = drift[nuv]; hs[k] = sign[nuv];
            p1x[k] = x_bep(g, d, el[nuv]); p1y[k] = y_bep(g, d, el[nuv]); p1z[k] = z_bep(g, d, el[nuv]);
            dpx[k] = g.deltapx[d]; dpy[k] = g.deltapy[d]; dpz[k] = g.deltapz[d];
            Y[nuv] = y; eY[nuv] = ey; orig[nuv] = h.orig[idx];
            nuv++;
        }
    };
    float chi2min = 10000.1f;
    bool found = false;
    float best[NP];
    int best_orig[8];
    int win_i[4] = {-1, -1, -1, -1};
    const int P1a = 48, P1b = 47, P2a = 54, P2b = 53;
    /* unit filter: index of each U3 / V3 entry among those with an accepted hit */
    bool filtered = false;
    std::vector<int> ku3, kv3;
    int nku3 = 0;
    if (g_unit_keys) {
        const int64_t k0 = unit_key(C.ev, T.canon, 0);
        const int64_t* it = std::lower_bound(g_unit_keys, g_unit_keys + g_unit_nkeys, k0);
        filtered = it != g_unit_keys + g_unit_nkeys && (*it >> 20) == (k0 >> 20);
        if (filtered) {
            auto compact = [&](int v, int n, std::vector<int>& k) {
                k.assign(n, -1);
                int c = 0;
                for (int i = 0; i < n; ++i) {
                    bool acc = false;
                    float y, ey;
                    for (int ab = 0; ab < 2; ++ab) {
                        const int idx = ab ? lists[v][i].b : lists[v][i].a;
                        if (idx < 0) continue;
                        const Plane& h = C.pl[vdet[v][ab]];
                        acc |= y_uvhit(g, vdet[v][ab], (int)h.chan[idx], h.drift[idx], 0, x0, tx, y, ey);
                    }
                    if (acc) k[i] = c++;
                }
                return c;
            };
            nku3 = compact(2, nu3, ku3);
            compact(3, nv3, kv3);
        }
    }
    for (int iuv3 = 0; iuv3 < nu3 * nv3; ++iuv3) {
        int iu3 = iuv3 % nu3, iv3 = iuv3 / nu3;
        if (filtered && ku3[iu3] >= 0 && kv3[iv3] >= 0 &&
            !unit_listed(unit_key(C.ev, T.canon, ku3[iu3] + kv3[iv3] * nku3)))
            continue;
        nuv = 0;
        if (lists[2][iu3].a >= 0) add(vdet[2][0], lists[2][iu3].a);
        if (lists[2][iu3].b >= 0) add(vdet[2][1], lists[2][iu3].b);
        int nu3h = nuv;
        if (nu3h == 0) continue;
        if (lists[3][iv3].a >= 0) add(vdet[3][0], lists[3][iv3].a);
        if (lists[3][iv3].b >= 0) add(vdet[3][1], lists[3][iv3].b);
        int nv3h = nuv - nu3h;
        if (nv3h == 0) continue;
        const int n3 = nu3h + nv3h;
        int yz_seen = 0;
        short yz_seed[4] = {0, 0, 0, 0}, yz_last = 0;
        for (int iuv2 = 0; iuv2 < nu2 * nv2; ++iuv2) {
            int iu2 = iuv2 % nu2, iv2 = iuv2 / nu2;
            nuv = nu3h + nv3h;
            if (lists[0][iu2].a >= 0) add(vdet[0][0], lists[0][iu2].a);
            if (lists[0][iu2].b >= 0) add(vdet[0][1], lists[0][iu2].b);
            int nu2h = nuv - nu3h - nv3h;
            if (nu2h == 0) continue;
            if (lists[1][iv2].a >= 0) add(vdet[1][0], lists[1][iv2].a);
            if (lists[1][iv2].b >= 0) add(vdet[1][1], lists[1][iv2].b);
            int nv2h = nuv - nu3h - nv3h - nu2h;
            if (nv2h == 0) continue;
            float f[4];
            fit2d(nuv, Y, Z, eY, f);
            float y0 = f[0], ty = f[1], ey0 = f[2], ety = f[3];
            if (std::fabs(y0) > Y0_MAX + 2 * ey0 || std::fabs(ty) > TY_MAX + 2 * ety) continue;
            const bool yzc = g_yzcheck.load(std::memory_order_relaxed) != 0;
            int yz_ov = -1;
            if (yzc) {
                bool ok = true;
                for (int q = 0; q < n3; ++q)
                    ok &= sign[q] == (yz_seen == 0 ? 0 : (q == n3 - 1 ? yz_last : yz_seed[q]));
                const int ci = yzcheck_cross(det, nuv, n3);
                if (ci >= 0) {
                    yz_ov = yzcheck_pair(x0, tx, y0, ty, ex0, etx, ey0, ety, ci, ci + 1, det, pos, drift, g, 150.f);
                    g_yzcnt[2]++;
                }
                g_yzcnt[0]++;
                if (!ok) g_yzcnt[1]++;
            }
            lr_newhits(x0, tx, y0, ty, ex0, etx, ey0, ety, nuv, det, pos, drift, sign, g, 150.f);
            lr_single_newhits(x0, tx, y0, ty, nuv, det, pos, sign, g);
            if (yzc) {
                if (yz_seen == 0) {
                    for (int q = 0; q < n3; ++q) yz_seed[q] = sign[q];
                    yz_last = sign[n3 - 1];
                } else {
                    if (yz_ov >= 0) yz_last = kLR[yz_ov][0];
                    bool ok = true;
                    for (int q = 0; q < n3; ++q) ok &= sign[q] == (q == n3 - 1 ? yz_last : yz_seed[q]);
                    g_yzcnt[0]++;
                    if (!ok) g_yzcnt[1]++;
                }
                yz_seen++;
            }
            for (int l = 0; l < nuv; ++l) hs[nx + l] = sign[l];
            fit2d(nuv, Y, Z, eY, f); /* identical refit (Q13, reconstruction_kernels.cuh:1628) */
            y0 = f[0]; ty = f[1]; ey0 = f[2]; ety = f[3];
            if (std::fabs(y0) > Y0_MAX || std::fabs(ty) > TY_MAX) continue;
            int nprop = 0;
            bool next = true;
            for (int m = 0; m < C.pl[P1a].n; ++m)
                if (std::fabs(C.pl[P1a].pos[m] - (ty * g.z[P1a] + y0)) < 5.08f) { nprop++; next = false; break; }
            if (next)
                for (int m = 0; m < C.pl[P1b].n; ++m)
                    if (std::fabs(C.pl[P1b].pos[m] - (ty * g.z[P1b] + y0)) < 5.08f) { nprop++; break; }
            next = true;
            for (int m = 0; m < C.pl[P2a].n; ++m)
                if (std::fabs(C.pl[P2a].pos[m] - (ty * g.z[P2a] + y0)) < 10.16f) { nprop++; next = false; break; }
            if (next)
                for (int m = 0; m < C.pl[P2b].n; ++m)
                    if (std::fabs(C.pl[P2b].pos[m] - (ty * g.z[P2b] + y0)) < 10.16f) { nprop++; break; }
            if (nprop < 2) continue;
            bool ok = true;
            for (int st = 1; st <= 3 && ok; ++st) {
                int da = kHodoX[st][0], db = kHodoX[st][1];
                bool m = hodo_tkl(st, da, C.pl[da].n, C.pl[da].pos, x0, y0, tx, ty, ex0, ey0, etx, ety, g);
                if (!m) m = hodo_tkl(st, db, C.pl[db].n, C.pl[db].pos, x0, y0, tx, ty, ex0, ey0, etx, ety, g);
                ok = m;
            }
            if (!ok) continue;
            float chi2 = 0;
            for (int k = 0; k < nx + nuv; ++k) {
                float dca = dca_line(p1x[k], p1y[k], p1z[k], dpx[k], dpy[k], dpz[k], x0, y0, tx, ty);
                rsd[k] = dd[k] * hs[k] - dca;
                chi2 += ((rsd[k] * rsd[k]) / res[k]) / res[k];
            }
            trace(0, C, T.canon, iu3, iv3, iu2, iv2, chi2);
            if (chi2 < chi2min) {
                chi2min = chi2;
                found = true;
                win_i[0] = iu3; win_i[1] = iv3; win_i[2] = iu2; win_i[3] = iv2;
                best[2] = nuv; best[3] = chi2; best[6] = ty; best[8] = y0; best[11] = ety; best[13] = ey0;
                for (int h = 0; h < nuv; ++h) {
                    best[16 + h] = det[h]; best[34 + h] = el[h]; best[52 + h] = pos[h]; best[70 + h] = drift[h];
                    best[88 + h] = sign[h]; best[106 + h] = tdc[h];
                    if (h < nx) best[124 + h] = rsd[h];
                    best[128 + h] = rsd[nx + h];
                    best_orig[h] = orig[h];
                }
            }
        }
    }
    if (found) {
        trace(1, C, T.canon, win_i[0], win_i[1], win_i[2], win_i[3], chi2min);
        int nb = (int)best[2];
        R[0] = 5;
        R[2] = (float)(nx + nb);
        R[3] = best[3]; R[6] = best[6]; R[8] = best[8]; R[11] = best[11]; R[13] = best[13];
        for (int h = 0; h < nb; ++h) {
            R[22 + nx + h] = best[16 + h]; R[40 + nx + h] = best[34 + h]; R[58 + nx + h] = best[52 + h];
            R[76 + nx + h] = best[70 + h]; R[94 + nx + h] = best[88 + h]; R[112 + nx + h] = best[106 + h];
            if (h < nx) R[130 + h] = best[124 + h];
            R[130 + nx + h] = best[128 + h];
            T.orig[nx + h] = best_orig[h];
        }
    }
    return true;
}

/* ---- back-partial cleaning: reconstruction_helper.cuh:1789-1909 ------------ */
static int shared_hits(const float* a, int na, const float* b, int nb) {
    int c = 0;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j)
            if (b[22 + j] == a[22 + i] && b[40 + j] == a[40 + i]) { c++; break; }
    return c;
}

static void run_cleaning(Ctx& C) {
    /* slot ranges in the slot-sorted list */
    int beg[33];
    {
        int k = 0;
        for (int s = 0; s < 32; ++s) {
            beg[s] = k;
            while (k < (int)C.tk.size() && C.tk[k].slot == s) ++k;
        }
        beg[32] = (int)C.tk.size();
    }
    /* within a slot */
    for (int s = 0; s < 32; ++s) {
        for (int i = beg[s]; i < beg[s + 1]; ++i) {
            float* A = C.tk[i].rec;
            if (A[0] < 5) continue;
            int ni = (short)A[2];
            int common = 0;
            for (int j = i + 1; j < beg[s + 1]; ++j) {
                float* B = C.tk[j].rec;
                if (B[0] < 5) continue;
                int nj = (short)B[2];
                common += shared_hits(A, ni, B, nj);
                if (A[3] / (ni - 4) < B[3] / (nj - 4)) { if (common > ni * 0.3333f) B[0] = 3; }
                else { if (common > nj * 0.3333f) A[0] = 3; }
            }
        }
    }
    /* across slots (R2) */
    for (int s = 0; s < 32; ++s)
        for (int o = 0; o < 32; ++o) {
            if (o == s) continue;
            for (int i = beg[s]; i < beg[s + 1]; ++i) {
                float* A = C.tk[i].rec;
                if (A[0] < 5) continue;
                int ni = (short)A[2];
                int common = 0;
                for (int j = beg[o]; j < beg[o + 1]; ++j) {
                    float* B = C.tk[j].rec;
                    if (B[0] < 5) continue;
                    int nj = (short)B[2];
                    common += shared_hits(A, ni, B, nj);
                    if (A[3] / (ni - 4) < B[3] / (nj - 4)) { if (common > ni * 0.3333f) B[0] = 3; }
                    else { if (common > nj * 0.3333f) A[0] = 3; }
                }
            }
        }
}

/* ---- global: reconstruction_kernels.cuh:1841-2534, one track --------------- */
/* SagittaRatioInStation1, reconstruction_helper.cuh:556-613: expected position
 * and half-width of the station-1 window of each view (0 X, 1 U, 2 V) from the
 * track's projections at station 2 and at its last hit, combined for the
 * target and the dump hypotheses */
static void sagitta_windows(const Geo& g, float x0, float tx, float y0, float ty, int dlast, float* pexp, float* win) {
    float z3 = g.z[dlast], x3 = x0 + tx * z3, y3 = y0 + ty * z3;
    for (int i = 0; i < 3; ++i) {
        int d = 2 * i + 2;
        int idx = kPlaneType[d / 2 - 1];
        int d2 = kSuperId[2][idx] * 2;
        float ps3 = x3 * g.costheta[d2] + y3 * g.sintheta[d2];
        float z1 = g.z[d], z2 = g.z[d2];
        float x2 = x0 + tx * z2, y2 = y0 + ty * z2;
        float ps2 = x2 * g.costheta[d2] + y2 * g.sintheta[d2];
        float s2t = ps2 - (ps3 * (z2 - Z_TARGET)) / (z3 - Z_TARGET);
        float s2d = ps2 - (ps3 * (z2 - Z_DUMP)) / (z3 - Z_DUMP);
        float pet = SAG_T_C * s2t + (ps3 * (z1 - Z_TARGET)) / (z3 - Z_TARGET);
        float ped = SAG_D_C * s2d + (ps3 * (z1 - Z_DUMP)) / (z3 - Z_DUMP);
        float wt = std::fabs(s2t * SAG_T_W), wd = std::fabs(s2d * SAG_D_W);
        float pmn = std::fmin(pet - wt, ped - wd), pmx = std::fmax(pet + wt, ped + wd);
        pexp[idx] = 0.5 * (pmx + pmn);
        win[idx] = 0.5 * (pmx - pmn);
    }
}

static bool run_global_one(Ctx& C, Tkl& T) {
    const Geo& g = *C.g;
    float* R = T.rec;
    const int n23 = (int)R[2];
    if (R[0] < 5) return true;
    const float x0 = R[7], tx = R[5], ex0 = R[12], etx = R[10];
    float y0 = R[8], ty = R[6], ey0 = R[13], ety = R[11];
    /* last hit in z: reconstruction_classes.cuh:407-422 */
    int dlast = 1;
    float zmax = -10000.f;
    for (int k = 0; k < n23; ++k) {
        short d = (short)R[22 + k];
        if (d > 0 && g.z[d] > zmax) { dlast = d; zmax = g.z[d]; }
    }
    /* SagittaRatioInStation1, reconstruction_helper.cuh:556-613 */
    float pexp[3], win[3];
    sagitta_windows(g, x0, tx, y0, ty, dlast, pexp, win);
    const int wcap = std::max(C.p->max_pairs_window, 1);
    std::vector<Pair> lxv(wcap), luv(wcap), lvv(wcap);
    Pair *lx = lxv.data(), *lu = luv.data(), *lv = lvv.data();
    int nx1 = pairs_window(C, 0, 0, pexp[0] - win[0] * 1.0f, pexp[0] + win[0] * 1.0f, lx, C.p->max_pairs_window);
    int nu1 = pairs_window(C, 0, 1, pexp[1] - win[1] * 1.0f, pexp[1] + win[1] * 1.0f, lu, C.p->max_pairs_window);
    int nv1 = pairs_window(C, 0, 2, pexp[2] - win[2] * 1.0f, pexp[2] + win[2] * 1.0f, lv, C.p->max_pairs_window);
    if (nx1 < 0 || nu1 < 0 || nv1 < 0) { C.status |= OR_EV_PAIR_OVERFLOW; return false; }
    if (nx1 == 0 || nu1 == 0 || nv1 == 0) return true;
    trace(5, C, T.canon, nx1, nu1, nv1, 0, 0.f);
    /* y of the station-2/3 UV hits with their signs (reconstruction_kernels.cuh:2086-2102) */
    short d23[12], c23[12], s23[12];
    float dr23[12], res23[12], r23[12];
    float Y[18], eY[18], Zy[18];
    int ny = 0;
    for (int k = 0; k < n23; ++k) {
        d23[k] = (short)R[22 + k]; c23[k] = (short)R[40 + k]; dr23[k] = R[76 + k]; s23[k] = (short)R[94 + k];
        res23[k] = g.resolution[d23[k]];
        if (g.costheta[d23[k]] > 0.99) continue;
        float y, ey;
        if (y_uvhit(g, (int)R[22 + k], (int)R[40 + k], R[76 + k], (short)R[94 + k], x0, tx, y, ey)) {
            Y[ny] = y; eY[ny] = ey; Zy[ny] = g.z[d23[k]]; ny++;
        }
    }
    const int dX = 4, dXp = 3, dU = 2, dUp = 1, dV = 6, dVp = 5;
    short det[18], el[18], sign[18];
    float pos[18], drift[18], tdc[18], X[3], eX[3], Z[3];
    float res[18], p1x[18], p1y[18], p1z[18], dpx[18], dpy[18], dpz[18], rsd[18];
    int orig[18];
    float chi2min = 10000.1f;
    bool found = false;
    float best[NP];
    int best_orig[6];
    int gwin[3] = {-1, -1, -1};
    for (int ix = 0; ix < nx1; ++ix) {
        int nx = 0;
        auto addx = [&](int d, int idx) {
            const Plane& h = C.pl[d];
            det[nx] = (short)d; pos[nx] = h.pos[idx]; eX[nx] = g.resolution[d]; el[nx] = (short)h.chan[idx];
            tdc[nx] = h.tdc[idx]; sign[nx] = 0; drift[nx] = h.drift[idx]; orig[nx] = h.orig[idx];
            xz_fy(g, d, el[nx], drift[nx], sign[nx], y0, ty, X[nx], Z[nx]);
            res[nx] = g.resolution[d];
            p1x[nx] = x_bep(g, d, el[nx]); p1y[nx] = y_bep(g, d, el[nx]); p1z[nx] = z_bep(g, d, el[nx]);
            dpx[nx] = g.deltapx[d]; dpy[nx] = g.deltapy[d]; dpz[nx] = g.deltapz[d];
            ++nx;
        };
        if (lx[ix].a >= 0) addx(dX, lx[ix].a);
        if (lx[ix].b >= 0) addx(dXp, lx[ix].b);
        if (nx == 0) continue;
        X[nx] = x0 + tx * Z_KMAG_BEND;
        eX[nx] = ex0 + etx * Z_KMAG_BEND;
        Z[nx] = Z_KMAG_BEND;
        float f[4];
        fit2d(nx + 1, X, Z, eX, f);
        float x01 = f[0], tx1 = f[1], ex01 = f[0], etx1 = f[1];
        lr_xhits(x01, tx1, ex01, etx1, nx, det, X, drift, sign, g, 150.f);
        lr_single_xhits(x01, tx1, nx, det, X, sign, g);
        for (int l = 0; l < nx; ++l) X[l] += sign[l] * drift[l];
        fit2d(nx + 1, X, Z, eX, f);
        x01 = f[0]; tx1 = f[1]; ex01 = f[0]; etx1 = f[1];
        /* calculate_charge (E1039), reconstruction_helper.cuh:684-691 */
        short q = (double)(-0.0033f) * std::copysign(1.0, (double)FMAGSTR) * (double)x0 < (double)tx ? 1 : -1;
        float invp = (tx1 - tx) * q / PT_KICK_KMAG;
        float einvp = (etx - etx) / PT_KICK_KMAG;
        if (invp > INVP_MAX + einvp || invp < INVP_MIN - einvp) continue;
        for (int iuv = 0; iuv < nu1 * nv1; ++iuv) {
            int iu = iuv % nu1, iv = iuv / nu1;
            int nuv = 0;
            auto adduv = [&](int d, int idx, float zpl) {
                const Plane& h = C.pl[d];
                int k = nx + nuv;
                det[k] = (short)d; el[k] = (short)h.chan[idx]; tdc[k] = h.tdc[idx]; sign[k] = 0;
                drift[k] = h.drift[idx]; pos[k] = h.pos[idx];
                float y, ey;
                if (y_uvhit(g, d, el[k], drift[k], 0, x01, tx1, y, ey)) {
                    res[k] = g.resolution[d];
                    p1x[k] = x_bep(g, d, el[k]); p1y[k] = y_bep(g, d, el[k]); p1z[k] = z_bep(g, d, el[k]);
                    dpx[k] = g.deltapx[d]; dpy[k] = g.deltapy[d]; dpz[k] = g.deltapz[d];
                    Y[ny + nuv] = y; eY[ny + nuv] = ey; Zy[ny + nuv] = zpl; orig[k] = h.orig[idx];
                    nuv++;
                }
            };
            if (lu[iu].a >= 0) adduv(dU, lu[iu].a, g.z[dU]);
            if (lu[iu].b >= 0) adduv(dUp, lu[iu].b, g.z[dUp]);
            int nuh = nuv;
            if (nuh == 0) continue;
            if (lv[iv].a >= 0) adduv(dV, lv[iv].a, g.z[dV]);
            if (lv[iv].b >= 0) adduv(dVp, lv[iv].b, g.z[dVp]);
            if (nuv - nuh == 0) continue;
            fit2d(ny + nuv, Y, Zy, eY, f);
            y0 = f[0]; ty = f[1]; ey0 = f[2]; ety = f[3];
            if (std::fabs(y0) > Y0_MAX + 2 * ey0 || std::fabs(ty) > TY_MAX + 2 * ety) continue;
            lr_newhits(x01, tx1, y0, ty, ex01, etx1, ey0, ety, nx + nuv, det, pos, drift, sign, g, 150.f);
            lr_single_newhits(x01, tx1, y0, ty, nx + nuv, det, pos, sign, g);
            fit2d(ny + nuv, Y, Zy, eY, f);
            y0 = f[0]; ty = f[1]; ey0 = f[2]; ety = f[3];
            if (std::fabs(y0) > Y0_MAX || std::fabs(ty) > TY_MAX) continue;
            /* H1 match is computed but not enforced without HODO_Y_MATCH (reconstruction_kernels.cuh:2424-2440) */
            float chi2 = 0;
            for (int m = 0; m < n23; ++m) {
                int d = d23[m], e = c23[m];
                float dca = dca_line(x_bep(g, d, e), y_bep(g, d, e), z_bep(g, d, e), g.deltapx[d], g.deltapy[d],
                                     g.deltapz[d], x0, y0, tx, ty);
                r23[m] = dr23[m] * s23[m] - dca;
                chi2 += (r23[m] * r23[m]) / (res23[m] * res23[m]);
            }
            float c1 = 0;
            for (int k = 0; k < nx + nuv; ++k) {
                float dca = dca_line(p1x[k], p1y[k], p1z[k], dpx[k], dpy[k], dpz[k], x01, y0, tx1, ty);
                rsd[k] = drift[k] * sign[k] - dca;
                c1 += ((rsd[k] * rsd[k]) / res[k]) / res[k];
            }
            chi2 += c1;
            trace(2, C, T.canon, ix, iu, iv, 0, chi2);
            if (chi2 > 10000.f) continue;
            if (chi2 < chi2min) {
                chi2min = chi2;
                found = true;
                gwin[0] = ix; gwin[1] = iu; gwin[2] = iv;
                best[2] = nx + nuv; best[3] = chi2; best[6] = ty; best[8] = y0; best[9] = invp; best[11] = ety;
                best[13] = ey0; best[14] = einvp; best[15] = q;
                for (int m = 0; m < n23; ++m) best[124 + m] = r23[m];
                for (int m = 0; m < nx + nuv; ++m) {
                    best[16 + m] = det[m]; best[34 + m] = el[m]; best[52 + m] = pos[m]; best[70 + m] = drift[m];
                    best[88 + m] = sign[m]; best[106 + m] = tdc[m]; best[124 + n23 + m] = rsd[m];
                    best_orig[m] = orig[m];
                }
            }
        }
    }
    if (found) {
        trace(3, C, T.canon, gwin[0], gwin[1], gwin[2], 0, chi2min);
        int n1 = (int)best[2];
        R[0] = 6;
        R[2] = (float)(n23 + n1);
        R[3] = best[3]; R[6] = best[6]; R[8] = best[8]; R[9] = best[9]; R[11] = best[11]; R[13] = best[13];
        R[14] = best[14]; R[15] = best[15];
        for (int m = 0; m < n23; ++m) R[130 + m] = best[124 + m];
        for (int m = 0; m < n1; ++m) {
            R[22 + n23 + m] = best[16 + m]; R[40 + n23 + m] = best[34 + m]; R[58 + n23 + m] = best[52 + m];
            R[76 + n23 + m] = best[70 + m]; R[94 + n23 + m] = best[88 + m]; R[112 + n23 + m] = best[106 + m];
            R[130 + n23 + m] = best[124 + n23 + m];
            T.orig[n23 + m] = best_orig[m];
        }
    }
    return true;
}

/* ---- one event: staging (online_reconstruction.cu:506-640) + eR + chain ---- */
/* ---- vertexing: gKernel_Vertexing, reconstruction_kernels.cuh:2547-2810 -------
 * Swims a global track from its most upstream hit through FMAG (100 steps with
 * pT kicks and energy loss outside the beam hole) and the target region (100
 * straight steps), then takes the trajectory point closest to the beam line in
 * x-y (first minimum, points whose x momentum points the wrong way skipped).
 * Constants: reconstruction_constants.h:75-121 (E1039 set).  The reference mixes
 * float and double; every expression keeps its types (sqrtf on a double
 * argument rounds the argument to float first).  The dump-plane point the
 * reference also computes is not part of its output and is not computed. */
constexpr double V_FMAG_LENGTH = 502.92, V_FMAG_HOLE_LENGTH = 27.94, V_FMAG_HOLE_RADIUS = 1.27;
constexpr int V_NSTEPS_FMAG = 100, V_NSTEPS_TARGET = 100;
constexpr double V_STEP_FMAG = 2.514600, V_STEP_TARGET = 5.0, V_PTKICK_UNIT = -0.006096568;
constexpr double V_DEDX0 = 0.014282073, V_DEDX1 = 7.1869681e-05, V_DEDX2 = -1.4279150e-06, V_DEDX3 = 1.5853655e-08,
                 V_DEDX4 = -6.0741470e-11;
constexpr float V_X_BEAM = 0.f, V_Y_BEAM = 0.f;

static double dedx_poly(double p) {
    return V_DEDX0 + V_DEDX1 * p + V_DEDX2 * p * p + V_DEDX3 * p * p * p + V_DEDX4 * p * p * p * p;
}

/* R: the track's 148-float record; writes stationID 7 and fields 16-21 */
static void vertex_track(const Geo& g, float* R) {
    const float x0 = R[7], y0 = R[8], tx = R[5], ty = R[6], invP = R[9], charge = R[15];
    const int nh = (int)R[2];
    /* get_firsthitdetid, reconstruction_classes.cuh:517-531 */
    short det_first = 30;
    float zmin = 10000.f;
    for (int i = 0; i < nh; ++i) {
        const short d = (short)R[22 + i];
        if (d > 0 && g.z[d] < zmin) { det_first = d; zmin = g.z[d]; }
    }
    /* calculate_x0_tx_st1, reconstruction_helper.cuh:622-626 (charge as short) */
    const short q = (short)charge;
    const float tx_st1 = tx + PT_KICK_KMAG * invP * q;
    const float x0_st1 = tx * Z_KMAG_BEND + x0 - tx_st1 * Z_KMAG_BEND;
    const double z0 = g.z[det_first];
    const double s0 = charge * invP * sqrtf((1.f + tx_st1 * tx_st1) / (1.f + tx_st1 * tx_st1 + ty * ty));
    const double s1 = tx_st1, s2 = ty;
    const double s3 = (float)z0 * tx_st1 + x0_st1;   /* x_trk / y_trk take float z */
    const double s4 = (float)z0 * ty + y0;
    double px = s3 + s1 * (V_FMAG_LENGTH - z0), py = s4 + s2 * (V_FMAG_LENGTH - z0), pz = V_FMAG_LENGTH;
    const double pz0 = 1.f / (std::fabs(s0) * sqrtf((float)(1.f + s1 * s1 + s2 * s2)));
    double mx = pz0 * s1, my = pz0 * s2, mz = pz0;
    double dmin = 1.e18;
    bool found = false;
    double v[6] = {0, 0, 0, 0, 0, 0};
    auto consider = [&]() {
        if (FMAGSTR * charge * mx < 0.) return;
        const double d2 = (px - V_X_BEAM) * (px - V_X_BEAM) + (py - V_Y_BEAM) * (py - V_Y_BEAM);
        if (d2 < dmin) { dmin = d2; found = true; v[0] = px; v[1] = py; v[2] = pz; v[3] = mx; v[4] = my; v[5] = mz; }
    };
    consider();
    for (int step = 1; step <= V_NSTEPS_FMAG; ++step) {
        const double tx_i = mx / mz;
        const double tx_f = tx_i + 2.f * charge * V_PTKICK_UNIT * V_STEP_FMAG / std::sqrt(mx * mx + mz * mz);
        const double t1x = tx_i * V_STEP_FMAG, t1y = ty * V_STEP_FMAG, t1z = V_STEP_FMAG;
        const double bx = px - t1x, by = py - t1y, bz = pz - t1z;
        const double ptot_i = std::sqrt(mx * mx + my * my + mz * mz);
        double ptot_b = ptot_i;
        if (bz > V_FMAG_HOLE_LENGTH || bx * bx + by * by > V_FMAG_HOLE_RADIUS)
            ptot_b += dedx_poly(ptot_i) * sqrtf((float)(t1x * t1x + t1y * t1y + t1z * t1z));
        const double t2x = tx_f * V_STEP_FMAG, t2y = ty * V_STEP_FMAG, t2z = V_STEP_FMAG;
        px = bx - t2x; py = by - t2y; pz = bz - t2z;
        double ptot_f = ptot_b;
        if (pz > V_FMAG_HOLE_LENGTH || px * px + py * py > V_FMAG_HOLE_RADIUS)
            ptot_f += dedx_poly(ptot_b) * sqrtf((float)(t2x * t2x + t2y * t2y + t2z * t2z));
        const double pz_f = ptot_f / std::sqrt(1.f + tx_f * tx_f + ty * ty);
        mx = pz_f * tx_f; my = pz_f * ty; mz = pz_f;
        consider();
    }
    for (int step = 1; step <= V_NSTEPS_TARGET; ++step) {
        const double tx_i = mx / mz;
        px -= tx_i * V_STEP_TARGET; py -= ty * V_STEP_TARGET; pz -= V_STEP_TARGET;
        consider();
    }
    if (!found) { v[0] = px; v[1] = py; v[2] = pz; v[3] = mx; v[4] = my; v[5] = mz; }
    R[0] = 7;
    for (int k = 0; k < 6; ++k) R[16 + k] = (float)v[k];
}

/* ---- prop-tube segment matching: gKernel_PropSegmentMatching,
 * reconstruction_helper.cuh:1916-2257.  Every tracklet of the event (any
 * stationID) needs an x segment and a y segment in the prop tubes that agree
 * with it; a tracklet without one gets stationID 5.  A segment is a hit of the
 * first station (P1, either plane) within win1 of the track and a hit of the
 * second (P2, either plane) within win2, whose line has |slope| <= smax,
 * |slope - track slope| <= 0.03, |intercept| <= imax and lies within 3 cm of the
 * track at z = 2028.19.  The reference's loop nesting only decides the order
 * in which the combinations are tried; the outcome is whether any passes. */
static bool prop_segment(Ctx& C, int p1a, int p1b, int p2a, int p2b, float slope, float icpt, float win1, float win2,
                         float smax, float imax) {
    const Geo& g = *C.g;
    const float cut = 0.03f;
    const int first[2] = {p1a, p1b}, second[2] = {p2a, p2b};
    for (int f = 0; f < 2; ++f) {
        const Plane& A = C.pl[first[f]];
        const float za = g.z[first[f]];
        for (int n = 0; n < A.n; ++n) {
            const float pos0 = A.pos[n];
            if (!(std::fabs(pos0 - (slope * za + icpt)) < win1)) continue;
            for (int s2 = 0; s2 < 2; ++s2) {
                const Plane& B = C.pl[second[s2]];
                const float zb = g.z[second[s2]];
                for (int m = 0; m < B.n; ++m) {
                    const float pos1 = B.pos[m];
                    if (std::fabs(pos1 - (slope * zb + icpt)) > win2) continue;
                    const float a = (pos1 - pos0) / (zb - za);
                    if (std::fabs(a) > smax) continue;
                    if (std::fabs(a - slope) > cut) continue;
                    const float b = pos0 - a * za;
                    if (std::fabs(b) > imax) continue;
                    if (std::fabs((slope * 2028.19f + icpt) - (a * 2028.19f + b)) > 3.0f) continue;
                    return true;
                }
            }
        }
    }
    return false;
}

static void run_prop_matching(Ctx& C) {
    for (auto& T : C.tk) {
        float* R = T.rec;
        const float x0 = R[7], tx = R[5], y0 = R[8], ty = R[6];
        if (!prop_segment(C, 50, 49, 52, 51, tx, x0, 5.08f, 7.62f, TX_MAX, X0_MAX) ||
            !prop_segment(C, 48, 47, 54, 53, ty, y0, 5.08f, 10.16f, TY_MAX, Y0_MAX))
            R[0] = 5;
    }
}

/* ---- global track cleaning: gKernel_GlobalTrackCleaning (reconstruction_helper.cuh:2265-2301)
 * per slot, then _crossthreads (:2303-2354) slot against slot in slot order.
 * chi2/ndf = chisq / (nhits - 5); merge threshold 0.015 on the signed relative
 * 1/p difference (no absolute value), same-sign charges only.  In the within-slot
 * pass the else branch compares invP(i) with itself (difference 0, always below
 * the threshold), as the reference does; chi2/ndf > 250 demotes i there too. */
static void run_global_cleaning(Ctx& C) {
    auto ndf = [](const float* R) { return R[3] / (R[2] - 5.f); };
    const float thr = 0.015f;
    /* slot ranges in the slot-sorted list */
    int sstart[33];
    std::vector<int> order(C.tk.size());
    {
        int k = 0;
        for (int s = 0; s < 32; ++s) {
            sstart[s] = k;
            while (k < (int)C.tk.size() && C.tk[k].slot == s) ++k;
        }
        sstart[32] = (int)C.tk.size();
        for (size_t i = 0; i < C.tk.size(); ++i) order[i] = (int)i;
    }
    for (int s = 0; s < 32; ++s)
        for (int ii = sstart[s]; ii < sstart[s + 1]; ++ii) {
            float* A = C.tk[order[ii]].rec;
            if (A[0] < 6) continue;
            if (ndf(A) > 250.f) { A[0] = 5; continue; }
            for (int jj = ii + 1; jj < sstart[s + 1]; ++jj) {
                float* B = C.tk[order[jj]].rec;
                if (B[0] < 6) continue;
                if (ndf(A) < ndf(B)) {
                    if ((A[9] - B[9]) / A[9] < thr && A[15] * B[15] > 0) B[0] = 5;
                } else {
                    if ((A[9] - A[9]) / B[9] < thr && A[15] * B[15] > 0) A[0] = 5;
                }
            }
        }
    for (int s = 0; s < 32; ++s)
        for (int o = 0; o < 32; ++o) {
            if (o == s) continue;
            for (int ii = sstart[s]; ii < sstart[s + 1]; ++ii) {
                float* A = C.tk[order[ii]].rec;
                if (A[0] < 6) continue;
                for (int jj = sstart[o]; jj < sstart[o + 1]; ++jj) {
                    float* B = C.tk[order[jj]].rec;
                    if (B[0] < 6) continue;
                    if (ndf(A) < ndf(B)) {
                        if ((A[9] - B[9]) / A[9] < thr && A[15] * B[15] > 0) B[0] = 5;
                    } else {
                        if ((A[9] - B[9]) / B[9] < thr && A[15] * B[15] > 0) A[0] = 5;
                    }
                }
            }
        }
}

static void run_event(const Geo& g, const or_params& P, const or_hits& in, int64_t e, const or_result& out,
                      std::vector<uint8_t>& keep_bytes, Ctx& C) {
    const int64_t h0 = in.event_offsets[e], h1 = in.event_offsets[e + 1];
    const int nh = (int)(h1 - h0);
    C.g = &g; C.p = &P; C.status = 0; C.tk.clear(); C.ev = e;
    for (int d = 0; d < ND; ++d) C.pl[d].n = 0;
    uint8_t* keep = keep_bytes.data() + h0;
    for (int h = 0; h < nh; ++h) keep[h] = 0;
    if (nh > P.max_hits_per_event) C.status |= OR_EV_EVENT_TOO_LARGE;
    /* raw grouping by detector, input order within a detector */
    struct Raw { std::vector<float> chan, pos, tdc, flag, drift; std::vector<int> orig; };
    static thread_local Raw raw[ND];
    for (int d = 0; d < ND; ++d) { raw[d].chan.clear(); raw[d].pos.clear(); raw[d].tdc.clear(); raw[d].flag.clear(); raw[d].drift.clear(); raw[d].orig.clear(); }
    if (!C.status) {
        for (int h = 0; h < nh; ++h) {
            int d = in.det[h0 + h], el = in.elem[h0 + h];
            if (d > 54) continue; /* planes beyond 54 are skipped (online_reconstruction.cu:506-510) */
            if (d < 1 || el < 1 || el > (int)g.nelem[d] || el >= OR_MAX_ELEM) { C.status |= OR_EV_BAD_HIT; break; }
            Raw& r = raw[d];
            r.chan.push_back((float)el);
            r.pos.push_back(g.wire[d][el]);
            r.tdc.push_back(in.tdc[h0 + h]);
            r.flag.push_back(in.flag ? (float)in.flag[h0 + h] : 2.0f);
            r.drift.push_back(std::fabs(in.drift[h0 + h]));
            r.orig.push_back(h);
        }
    }
    if (!C.status)
        for (int d = 1; d < ND; ++d) {
            int cap = d <= 30 ? P.max_hits_chamber_plane : (d <= 46 ? P.max_hits_hodo_plane : P.max_hits_prop_plane);
            if ((int)raw[d].chan.size() > cap || raw[d].chan.size() > 400) C.status |= OR_EV_PLANE_OVERFLOW;
        }
    int mult[4] = {0, 0, 0, 0};
    if (!C.status) {
        bool cl_over = false;
        signed char kf[400];
        for (int d = 1; d < ND; ++d) {
            Raw& r = raw[d];
            int n = (int)r.chan.size();
            bool reduce = (d >= 31) || (d <= 6) || (d >= 13);
            Plane& p = C.pl[d];
            p.n = 0;
            int nk;
            if (reduce) nk = reduce_plane(d, n, r.chan.data(), r.pos.data(), r.tdc.data(), r.flag.data(), r.drift.data(), kf, &cl_over);
            else { for (int i = 0; i < n; ++i) kf[i] = 1; nk = n; }
            if (reduce && nk > P.max_reduced_per_plane) C.status |= OR_EV_PLANE_OVERFLOW;
            for (int i = 0; i < n; ++i)
                if (kf[i] > 0) {
                    p.chan[p.n] = r.chan[i]; p.pos[p.n] = r.pos[i]; p.tdc[p.n] = r.tdc[i]; p.drift[p.n] = r.drift[i];
                    p.orig[p.n] = r.orig[i]; p.n++;
                }
            if (d >= 1 && d <= 6) mult[0] += nk;
            else if (d >= 13 && d <= 18) mult[1] += nk;
            else if (d >= 19 && d <= 24) mult[2] += nk;
            else if (d >= 25 && d <= 30) mult[3] += nk;
        }
        if (cl_over) C.status |= OR_EV_CLUSTER_OVERFLOW;
        if (!C.status)
            for (int d = 1; d < ND; ++d)
                for (int i = 0; i < C.pl[d].n; ++i) keep[C.pl[d].orig[i]] = 1;
        if (!C.status) {
            if (mult[0] > P.max_d0_mult) C.status |= OR_EV_D0_MULT;
            if (mult[1] > P.max_d2_mult) C.status |= OR_EV_D2_MULT;
            if (mult[2] > P.max_d3_mult) C.status |= OR_EV_D3P_MULT;
            if (mult[3] > P.max_d3_mult) C.status |= OR_EV_D3M_MULT;
        }
    }
    int ntkl = 0;
    if (!C.status) {
        run_xz(C);
        if (!(C.status & (OR_EV_PAIR_OVERFLOW | OR_EV_XZ_BIN_OVERFLOW))) ntkl = (int)C.tk.size();
    }
    if (!C.status)
        for (auto& T : C.tk)
            if (!run_yz_one(C, T)) break;
    if (!C.status) run_cleaning(C);
    if (!C.status)
        for (auto& T : C.tk)
            if (!run_global_one(C, T)) break;
    if (!C.status && P.full_chain) {
        run_prop_matching(C);
        run_global_cleaning(C);
    }
    if (!C.status && (out.vertex || P.full_chain))
        for (auto& T : C.tk)
            if (T.rec[0] >= 6) vertex_track(g, T.rec);
    /* outputs */
    int nt = 0;
    if (!C.status) {
        const float min_stid = P.full_chain ? 7.f : 6.f;
        for (auto& T : C.tk) {
            if (T.rec[0] < min_stid) continue;
            if (nt >= P.max_tracks_out) { C.status |= OR_EV_OUTPUT_OVERFLOW; continue; }
            or_track& o = out.tracks[e * P.max_tracks_out + nt];
            std::memset(&o, 0, sizeof o);
            const float* R = T.rec;
            o.x0 = R[7]; o.tx = R[5]; o.y0 = R[8]; o.ty = R[6]; o.invp = R[9];
            o.err_x0 = R[12]; o.err_tx = R[10]; o.err_y0 = R[13]; o.err_ty = R[11]; o.err_invp = R[14];
            o.chi2 = R[3];
            int nhit = (int)R[2];
            for (int h = 0; h < nhit; ++h) {
                o.hit_index[h] = (uint16_t)T.orig[h];
                if (R[94 + h] > 0) o.sign_plus |= 1u << h;
                if (R[94 + h] < 0) o.sign_minus |= 1u << h;
            }
            o.station_id = (int8_t)R[0]; o.nhits = (uint8_t)nhit; o.charge = (int8_t)R[15];
            o.slot = (uint8_t)T.slot; o.tracklet_id = (uint32_t)T.canon;
            if (out.vertex)
                for (int k = 0; k < 6; ++k) out.vertex[(e * P.max_tracks_out + nt) * 6 + k] = R[16 + k];
            ++nt;
        }
    }
    if (C.status & OR_EV_REJECT_MASK) nt = 0;
    out.status[e] = C.status;
    out.n_tracks[e] = nt;
    if (out.n_tracklets) out.n_tracklets[e] = (C.status & OR_EV_REJECT_MASK) ? 0 : ntkl;
    if (out.records) {
        float* base = out.records + (size_t)e * P.max_tracklets_per_event * NP;
        std::memset(base, 0, sizeof(float) * (size_t)P.max_tracklets_per_event * NP);
        if (!(C.status & OR_EV_REJECT_MASK))
            for (size_t k = 0; k < C.tk.size() && (int)k < P.max_tracklets_per_event; ++k)
                std::memcpy(base + k * NP, C.tk[k].rec, sizeof(float) * NP);
    }
}

}  // namespace orc

/* ============================ C ABI of the oracle ============================ */
extern "C" {

/* prop-tube segment test on explicit planes (first station: planes a, b; second:
 * c, d; positions and plane z), for unit tests of run_prop_matching */
int orc_prop_segment(const float* pa, int na, float za, const float* pb, int nb, float zb, const float* pc, int nc,
                     float zc, const float* pd, int nd, float zd, float slope, float icpt, float win1, float win2,
                     float smax, float imax) {
    orc::Geo g;
    std::memset(&g, 0, sizeof g);
    g.z[50] = za; g.z[49] = zb; g.z[52] = zc; g.z[51] = zd;
    orc::Ctx C;
    C.g = &g;
    auto fill = [&](int d, const float* p, int n) { C.pl[d].n = n; for (int i = 0; i < n; ++i) C.pl[d].pos[i] = p[i]; };
    fill(50, pa, na); fill(49, pb, nb); fill(52, pc, nc); fill(51, pd, nd);
    return orc::prop_segment(C, 50, 49, 52, 51, slope, icpt, win1, win2, smax, imax) ? 1 : 0;
}

/* global track cleaning on explicit tracks (slot-sorted): per track stid, chi2,
 * nhits, invP, charge and slot; stid is updated in place */
void orc_global_cleaning(int n, float* stid, const float* chi2, const float* nhits, const float* invp,
                         const float* charge, const int* slot) {
    orc::Ctx C;
    C.tk.resize(n);
    for (int i = 0; i < n; ++i) {
        std::memset(C.tk[i].rec, 0, sizeof C.tk[i].rec);
        C.tk[i].rec[0] = stid[i]; C.tk[i].rec[3] = chi2[i]; C.tk[i].rec[2] = nhits[i];
        C.tk[i].rec[9] = invp[i]; C.tk[i].rec[15] = charge[i]; C.tk[i].slot = slot[i];
    }
    orc::run_global_cleaning(C);
    for (int i = 0; i < n; ++i) stid[i] = C.tk[i].rec[0];
}

/* YZ unit filter for the following orc_reconstruct calls (keys sorted, see
 * g_unit_keys; NULL turns it off).  The keys must outlive those calls. */
void orc_set_unit_filter(const int64_t* keys, int64_t n) {
    orc::g_unit_keys = n > 0 ? keys : nullptr;
    orc::g_unit_nkeys = n > 0 ? n : 0;
}

void orc_yz_state_check(int enable, long long* counts /* 3, may be NULL */) {
    if (counts)
        for (int i = 0; i < 3; ++i) counts[i] = orc::g_yzcnt[i].load();
    orc::g_yzcheck.store(enable);
    if (enable)
        for (int i = 0; i < 3; ++i) orc::g_yzcnt[i].store(0);
}

struct orc_geometry { orc::Geo g; float deltaW[55 * 9]; };

int orc_geometry_load(const char* path, orc_geometry** out) {
    auto* G = new orc_geometry;
    int st = orc::load_geometry(path, G->g, G->deltaW);
    if (st) { delete G; return st; }
    *out = G;
    return 0;
}
void orc_geometry_free(orc_geometry* G) { delete G; }
/* planes: 27 x 55 in gPlane member order; wire: 55 x OR_MAX_ELEM; deltaW: 55 x 9 */
void orc_geometry_export(const orc_geometry* G, float* planes, float* wire, float* deltaW) {
    if (planes) std::memcpy(planes, &G->g, sizeof(float) * 27 * 55);
    if (wire) std::memcpy(wire, G->g.wire, sizeof G->g.wire);
    if (deltaW) std::memcpy(deltaW, G->deltaW, sizeof G->deltaW);
}

/* Runs the chain on host buffers with nthreads OpenMP threads (0 = default).
 * keep_mask is written as bits (ceil(n_hits/32) words). */
int orc_reconstruct(const orc_geometry* G, const or_params* P, const or_hits* in, or_result* out, int nthreads) {
    const int64_t n = in->n_events;
    std::vector<uint8_t> kb(in->n_hits > 0 ? in->n_hits : 1, 0);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        orc::Ctx* C = new orc::Ctx;
#pragma omp for schedule(dynamic, 1)
        for (int64_t e = 0; e < n; ++e) orc::run_event(G->g, *P, *in, e, *out, kb, *C);
        delete C;
    }
    if (out->keep_mask) {
        int64_t nw = (in->n_hits + 31) / 32;
        for (int64_t w = 0; w < nw; ++w) {
            uint32_t v = 0;
            for (int b = 0; b < 32; ++b) {
                int64_t h = w * 32 + b;
                if (h < in->n_hits && kb[h]) v |= 1u << b;
            }
            out->keep_mask[w] = v;
        }
    }
    return 0;
}

/* ---- single-function entry points used by the pinning tests ------------- */
void orc_fit2d(int n, const float* x, const float* z, const float* w, float* out4) { orc::fit2d(n, x, z, w, out4); }
void orc_fit2d_drift(int n, const float* x, const float* drift, const short* sign, const float* z, const float* w,
                     float* out4) { orc::fit2d_drift(n, x, drift, sign, z, w, out4); }
void orc_matinv2(const float* A, float* Ai) { orc::matinv2(A, Ai); }
int orc_reduce_plane(int det, int n, const float* chan, const float* pos, const float* tdc, const float* flag,
                     const float* drift, signed char* keep) {
    bool over = false;
    int r = orc::reduce_plane(det, n, chan, pos, tdc, flag, drift, keep, &over);
    return over ? -1 : r;
}
float orc_dca(float p1x, float p1y, float p1z, float dpx, float dpy, float dpz, float x0, float y0, float tx, float ty) {
    return orc::dca_line(p1x, p1y, p1z, dpx, dpy, dpz, x0, y0, tx, ty);
}
int orc_y_uvhit(const orc_geometry* G, int d, int e, float drift, int hs, float x0, float tx, float* y, float* ey) {
    return orc::y_uvhit(G->g, d, e, drift, (short)hs, x0, tx, *y, *ey) ? 1 : 0;
}
void orc_xz_fy(const orc_geometry* G, int d, int e, float drift, int hs, float y0, float ty, float* x, float* z) {
    orc::xz_fy(G->g, d, e, drift, (short)hs, y0, ty, *x, *z);
}
int orc_hodo_xz(const orc_geometry* G, int stid, int d, int n, const float* hpos, float x0, float tx, float ex0,
                float etx) { return orc::hodo_xz(stid, d, n, hpos, x0, tx, ex0, etx, G->g) ? 1 : 0; }
/* left/right resolution on explicit hit lists (reconstruction_helper.cuh:714-892) */
void orc_lr_xhits(const orc_geometry* G, float x0, float tx, float ex0, float etx, int n, const short* det,
                  const float* pos, const float* drift, short* sign, float thr) {
    orc::lr_xhits(x0, tx, ex0, etx, n, det, pos, drift, sign, G->g, thr);
}
void orc_lr_newhits(const orc_geometry* G, float x0, float tx, float y0, float ty, float ex0, float etx, float ey0,
                    float ety, int n, const short* det, const float* pos, const float* drift, short* sign, float thr) {
    orc::lr_newhits(x0, tx, y0, ty, ex0, etx, ey0, ety, n, det, pos, drift, sign, G->g, thr);
}
void orc_lr_single_xhits(const orc_geometry* G, float x0, float tx, int n, const short* det, const float* pos,
                         short* sign) {
    orc::lr_single_xhits(x0, tx, n, det, pos, sign, G->g);
}
void orc_lr_single_newhits(const orc_geometry* G, float x0, float tx, float y0, float ty, int n, const short* det,
                           const float* pos, short* sign) {
    orc::lr_single_newhits(x0, tx, y0, ty, n, det, pos, sign, G->g);
}
/* station-1 windows of a track (SagittaRatioInStation1): pexp[3], win[3] by view X, U, V */
void orc_sagitta(const orc_geometry* G, float x0, float tx, float y0, float ty, int dlast, float* pexp, float* win) {
    orc::sagitta_windows(G->g, x0, tx, y0, ty, dlast, pexp, win);
}
/* back-partial cleaning of n explicit tracklets (148-float records, storage
 * slot each), in slot order as the chain runs it; stid_out[i] = stationID of
 * tracklet i afterwards */
void orc_back_cleaning(int n, const float* recs, const int* slot, float* stid_out) {
    orc::Ctx* C = new orc::Ctx;
    C->status = 0;
    C->tk.resize(n);
    for (int i = 0; i < n; ++i) {
        std::memcpy(C->tk[i].rec, recs + (size_t)i * orc::NP, sizeof(float) * orc::NP);
        C->tk[i].slot = slot[i];
        C->tk[i].canon = i;
        C->tk[i].pos_in_slot = i;
    }
    std::stable_sort(C->tk.begin(), C->tk.end(), [](const orc::Tkl& a, const orc::Tkl& b) { return a.slot < b.slot; });
    orc::run_cleaning(*C);
    for (const auto& T : C->tk) stid_out[T.canon] = T.rec[0];
    delete C;
}
/* candidate trace: enable (and clear) / fetch up to cap rows; returns the row count */
void orc_trace(int enable) {
    std::lock_guard<std::mutex> lk(orc::g_trace_mu);
    orc::g_trace.clear();
    orc::g_trace_on.store(enable);
}
long long orc_trace_fetch(long long cap, int* kind, long long* ev, int* canon, int* idx4, float* chi2) {
    std::lock_guard<std::mutex> lk(orc::g_trace_mu);
    const long long n = (long long)orc::g_trace.size();
    for (long long r = 0; r < n && r < cap; ++r) {
        const auto& T = orc::g_trace[r];
        kind[r] = T.kind; ev[r] = T.ev; canon[r] = T.canon; chi2[r] = T.chi2;
        for (int k = 0; k < 4; ++k) idx4[4 * r + k] = T.i[k];
    }
    return n;
}
/* Binned X pair builder on explicit planes: pos/chan arrays; out: 28*cap pairs (a,b), cnt[28] */
void orc_pairs_binned(int stid, int n1, const float* chan1, const float* pos1, int n2, const float* chan2,
                      const float* pos2, int cap, int* out_pairs, int* cnt) {
    orc::Plane* a = new orc::Plane; orc::Plane* b = new orc::Plane;
    a->n = n1; b->n = n2;
    for (int i = 0; i < n1; ++i) { a->chan[i] = chan1[i]; a->pos[i] = pos1[i]; }
    for (int i = 0; i < n2; ++i) { b->chan[i] = chan2[i]; b->pos[i] = pos2[i]; }
    std::vector<orc::Pair> pr(28 * cap);
    orc::pairs_binned(*a, *b, stid, stid - 2, cap, pr.data(), cnt);
    for (int k = 0; k < 28 * cap; ++k) { out_pairs[2 * k] = pr[k].a; out_pairs[2 * k + 1] = pr[k].b; }
    delete a; delete b;
}
}

// Geometry file reader and wire table for libe1039or (host C++).
//
// The text format is the reference's basic_geometry_*.txt: one line per
// detector plane, '#' lines skipped, 26 numeric columns followed by one
// alignment offset (planes 1-46) or nine per-module offsets (prop tubes
// 47-54).  Reading follows online_reconstruction.cu:123-188, including its
// stream semantics: values are extracted with operator>> into variables
// declared outside the line loop, so a token such as "2.749630657183.4666022"
// (basic_geometry_e906.txt:48) yields 2.749630657183 and then .4666022.
#include "or_internal.h"

#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

namespace {

// reconstruction_constants.h:29-32
constexpr float kTxMax = 0.15f, kTyMax = 0.1f, kX0Max = 150.f, kY0Max = 50.f;

// Column order of the file (online_reconstruction.cu:133).
const PlaneField kColumns[] = {F_Z, F_NELEM, F_CELLWIDTH, F_SPACING, F_XOFFSET, F_SCALEX, F_X0, F_X1, F_X2,
                               F_COSTHETA, F_SCALEY, F_Y0, F_Y1, F_Y2, F_SINTHETA, F_RESOLUTION, F_P1X, F_P1Y,
                               F_P1Z, F_DELTAPX, F_DELTAPY, F_DELTAPZ, F_DP1X, F_DP1Y, F_DP1Z};
constexpr int kNCols = sizeof(kColumns) / sizeof(kColumns[0]);

or_status parse(std::istream& in, or_geometry& g) {
    std::memset(&g, 0, sizeof g);
    std::string line;
    int id = 0, nelem = 0;
    float col[kNCols] = {};
    float offset = 0.f;
    while (std::getline(in, line)) {
        if (line[0] == '#') continue;
        std::istringstream ss(line);
        ss >> id;
        for (int c = 0; c < kNCols; ++c) {
            if (kColumns[c] == F_NELEM) ss >> nelem;
            else ss >> col[c];
        }
        if (id < 0 || id >= OR_N_DETECTORS) return OR_ERR_PARSE;
        for (int c = 0; c < kNCols; ++c) g.f[kColumns[c]][id] = kColumns[c] == F_NELEM ? (float)nelem : col[c];
        const int noff = id > 46 ? 9 : 1;
        for (int k = 0; k < noff; ++k) { ss >> offset; g.deltaW[id][k] = offset; }
        // box limits of the left/right pairing (online_reconstruction.cu:177-186)
        const float ct = g.f[F_COSTHETA][id], st = g.f[F_SINTHETA][id];
        g.f[F_SLOPE_MAX][id] = ct * kTxMax + st * kTyMax;
        g.f[F_INTER_MAX][id] = ct * kX0Max + st * kY0Max;
        if (id % 2 == 0 && id > 1) {
            const float dres = g.f[F_RESOLUTION][id] + g.f[F_RESOLUTION][id - 1];
            const float dz = g.f[F_Z][id] - g.f[F_Z][id - 1];
            const double ds = dres / dz;
            const double di = ds * g.f[F_Z][id];
            for (int p = id - 1; p <= id; ++p) {
                g.f[F_SLOPE_MAX][p] = (float)(g.f[F_SLOPE_MAX][p] + ds);
                g.f[F_INTER_MAX][p] = (float)(g.f[F_INTER_MAX][p] + di);
            }
        }
    }
    return OR_OK;
}

// Wire position per element (online_reconstruction.cu:191-218).  The
// association of the sums differs between chambers and the other planes in
// the reference and is kept: for chambers the float products x0*cos and
// y0*sin are added one at a time to the double element offset; for
// hodoscopes and prop tubes the float sum x0*cos + y0*sin + xoffset comes
// first and the double element offset is added to it.
void wire_table(or_geometry& g) {
    for (int d = 1; d <= 54; ++d) {
        const int n = (int)g.f[F_NELEM][d];
        const float x0c = g.f[F_X0][d] * g.f[F_COSTHETA][d];
        const float y0s = g.f[F_Y0][d] * g.f[F_SINTHETA][d];
        for (int e = 1; e <= n && e < OR_MAX_ELEM; ++e) {
            double w;
            if (d <= 30) {
                const double off = (e - (g.f[F_NELEM][d] + 1.) / 2.) * g.f[F_SPACING][d];
                w = off + g.f[F_XOFFSET][d];
                w += x0c;
                w += y0s;
                w += g.deltaW[d][0];
            } else {
                const float base = x0c + y0s + g.f[F_XOFFSET][d];
                const double off = (e - (n + 1) / 2.) * g.f[F_SPACING][d];
                const int module = d >= 47 ? 8 - (e - 1) / 8 : 0;
                w = base + off;
                w += g.deltaW[d][module];
            }
            g.wire[d][e] = (float)w;
        }
    }
}

}  // namespace

extern "C" {

or_status or_geometry_load(const char* path, or_geometry** out) {
    if (!path || !out) return OR_ERR_INVALID_ARG;
    std::ifstream in(path);
    if (!in) return OR_ERR_IO;
    auto* g = new (std::nothrow) or_geometry;
    if (!g) return OR_ERR_NOMEM;
    or_status st = parse(in, *g);
    if (st != OR_OK) { delete g; return st; }
    wire_table(*g);
    *out = g;
    return OR_OK;
}

void or_geometry_free(or_geometry* g) { delete g; }

or_status or_geometry_export(const or_geometry* g, float* planes, float* wire) {
    if (!g) return OR_ERR_INVALID_ARG;
    if (planes) std::memcpy(planes, g->f, sizeof g->f);
    if (wire) std::memcpy(wire, g->wire, sizeof g->wire);
    return OR_OK;
}

void or_params_default(or_params* p) {
    if (!p) return;
    p->max_hits_chamber_plane = 200;
    p->max_hits_hodo_plane = 60;
    p->max_hits_prop_plane = 120;
    p->max_reduced_per_plane = 120;
    p->max_pair_plane_hits = 100;
    p->max_pairs_per_bin = 20;
    p->max_pairs_window = 100;
    p->max_tracklets_per_slot = 16;
    p->max_tracklets_per_event = 512;
    p->max_d0_mult = 210;
    p->max_d2_mult = 120;
    p->max_d3_mult = 90;
    p->max_tracks_out = 16;
    p->max_hits_per_event = 1024;
    p->full_chain = 0;
}

const char* or_status_string(or_status st) {
    switch (st) {
        case OR_OK: return "ok";
        case OR_ERR_INVALID_ARG: return "invalid argument";
        case OR_ERR_IO: return "cannot open file";
        case OR_ERR_PARSE: return "geometry line with a detector id outside 0..54";
        case OR_ERR_DEVICE: return "HIP runtime error";
        case OR_ERR_NOMEM: return "out of memory";
        case OR_ERR_NO_DEVICE: return "no GPU device visible";
        case OR_ERR_CAPACITY: return "parameter exceeds the bound of an index encoding";
        case OR_ERR_CONFLICT: return "a live session on this device uses a different geometry";
    }
    return "unknown status";
}
}

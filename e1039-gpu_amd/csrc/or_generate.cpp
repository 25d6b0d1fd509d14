// Seeded synthetic event generator (conventions in include/e1039_or_gen.h).
// Host C++; one counter-based random stream per event so the output is the
// same for any thread count.
#include "../../include/e1039_or_gen.h"
#include "or_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// KMAG kink and charge rule constants (reconstruction_constants.h:85-94).
constexpr double kZKmag = 1064.26;
constexpr double kPtKick = -0.3819216;
// FMAG: pT kick 2.909 GeV x FMAGSTR -1.054 (reconstruction_constants.h:91-92), bend plane and target z
constexpr double kPtKickF = -3.066086;
constexpr double kZFmag = 251.4;
constexpr double kZTarget = -300.0;

struct Rng {  // xoshiro256** seeded by splitmix64 (Blackman & Vigna)
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (next() >> 11) * 0x1.0p-53; }                 // [0,1)
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double gauss() {  // Box-Muller
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    int poisson(double mu) {
        if (mu < 10.0) {  // Knuth multiplication method
            double L = std::exp(-mu), p = 1.0;
            int k = 0;
            do { ++k; p *= uni(); } while (p > L);
            return k - 1;
        }
        // PTRS transformed rejection (W. Hoermann, 1993)
        double smu = std::sqrt(mu), b = 0.931 + 2.53 * smu, a = -0.059 + 0.02483 * b;
        double inva = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
        for (;;) {
            double U = uni() - 0.5, V = uni(), us = 0.5 - std::fabs(U);
            long k = (long)std::floor((2 * a / us + b) * U + mu + 0.43);
            if (us >= 0.07 && V <= vr) return (int)k;
            if (k < 0 || (us < 0.013 && V > us)) continue;
            if (std::log(V) + std::log(inva) - std::log(a / (us * us) + b) <= -mu + k * std::log(mu) - std::lgamma(k + 1.0))
                return (int)k;
        }
    }
};

struct GHit { int16_t det, elem; float tdc, drift; int seq; };

inline bool instrumented(int d) { return (d >= 1 && d <= 6) || (d >= 13 && d <= 54); }

// drift-time window per detector (calibration_e906.txt header lines)
inline void tdc_window(int d, double& tmin, double& tmax) {
    if (d <= 12) { tmin = 1650; tmax = 1750; }
    else if (d <= 18) { tmin = 1450; tmax = 1710; }
    else if (d <= 24) { tmin = 1360; tmax = 1580; }
    else if (d <= 30) { tmin = 1490; tmax = 1700; }
    else if (d <= 46) { tmin = 1000; tmax = 1000; }
    else { tmin = 560; tmax = 1200; }
}

// distance of closest approach between the track line and the wire line
// p1 + t*dp (double precision; same geometry as the residual of
// reconstruction_helper.cuh:1149-1165).
double wire_dca(const or_geometry& g, int d, int e, double x0, double tx, double y0, double ty) {
    double p1x = g.f[F_P1X][d] + g.f[F_DP1X][d] * (e - 1), p1y = g.f[F_P1Y][d] + g.f[F_DP1Y][d] * (e - 1);
    double p1z = g.f[F_P1Z][d] + g.f[F_DP1Z][d] * (e - 1);
    double dpx = g.f[F_DELTAPX][d], dpy = g.f[F_DELTAPY][d], dpz = g.f[F_DELTAPZ][d];
    // track direction (tx, ty, 1), point (x0, y0, 0); n = dir x dp
    double nx = ty * dpz - dpy, ny = dpx - tx * dpz, nz = tx * dpy - ty * dpx;
    double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
    return ((p1x - x0) * nx + (p1y - y0) * ny + p1z * nz) / nn;
}

struct TrackPar { double x0, tx, y0, ty, invp; int q; double x0u, txu; };

int channel_count(const or_geometry& g) {
    int n = 0;
    for (int d = 1; d < OR_N_DETECTORS; ++d)
        if (instrumented(d)) n += (int)g.f[F_NELEM][d];
    return n;
}

void channel_of(const or_geometry& g, int k, int& d, int& e) {
    for (d = 1; d < OR_N_DETECTORS; ++d) {
        if (!instrumented(d)) continue;
        int n = (int)g.f[F_NELEM][d];
        if (k < n) { e = k + 1; return; }
        k -= n;
    }
    d = 54; e = 1;
}

void noise_hit(const or_geometry& g, Rng& r, int d, int e, std::vector<GHit>& out) {
    double tmin, tmax;
    tdc_window(d, tmin, tmax);
    GHit h;
    h.det = (int16_t)d; h.elem = (int16_t)e;
    h.tdc = (float)r.uni(tmin, tmax);
    h.drift = d >= 31 && d <= 46 ? 0.f : (float)r.uni(0.0, 0.5 * g.f[F_CELLWIDTH][d]);
    h.seq = (int)out.size();
    out.push_back(h);
}

bool draw_track(const or_geometry& g, const or_gen_params& P, Rng& r, TrackPar& T) {
    // Muon from the target (z = Z_TARGET, transverse spread 1 cm), bent at the
    // FMAG and KMAG bend planes.  tx is the downstream (station 2/3) slope.
    for (int tries = 0; tries < 10000; ++tries) {
        const double tx = r.uni(-P.tx_max, P.tx_max), ty = r.uni(-P.ty_max, P.ty_max);
        const int q = r.uni() < 0.5 ? 1 : -1;
        const double invp = r.uni(P.invp_min, P.invp_max);
        const double xv = r.gauss(), yv = r.gauss();
        // upstream slope = downstream slope + kick * q / p (reconstruction_helper.cuh:629-633)
        const double txm = tx + kPtKick * invp * q;         // between FMAG and KMAG
        const double txv = txm + kPtKickF * invp * q;       // at the vertex
        const double xf = xv + txv * (kZFmag - kZTarget);   // at the FMAG bend plane
        const double xk = xf + txm * (kZKmag - kZFmag);     // at the KMAG bend plane
        const double x0 = xk - tx * kZKmag, y0 = yv - ty * kZTarget;
        if (std::fabs(x0) >= 150.0 || std::fabs(y0) >= 50.0) continue;
        const int qr = (double)0.0033f * x0 < tx ? 1 : -1;
        if (qr != q) continue;
        auto crosses = [&](int d) {
            double z = g.f[F_Z][d], x = x0 + tx * z, y = y0 + ty * z;
            return x >= g.f[F_X1][d] && x <= g.f[F_X2][d] && y >= g.f[F_Y1][d] && y <= g.f[F_Y2][d];
        };
        if (!crosses(16)) continue;
        if (!crosses(22) && !crosses(28)) continue;
        // upstream-of-KMAG line through the KMAG point with slope txm
        const double x0u = xk - txm * kZKmag;
        const double z1 = g.f[F_Z][4], x1 = x0u + txm * z1, y1 = y0 + ty * z1;
        if (x1 < g.f[F_X1][4] || x1 > g.f[F_X2][4] || y1 < g.f[F_Y1][4] || y1 > g.f[F_Y2][4]) continue;
        T = {x0, tx, y0, ty, invp, q, x0u, txm};
        return true;
    }
    return false;
}

void track_hits(const or_geometry& g, const or_gen_params& P, Rng& r, const TrackPar& T, std::vector<GHit>& out) {
    for (int d = 1; d < OR_N_DETECTORS; ++d) {
        if (!instrumented(d)) continue;
        double z = g.f[F_Z][d];
        bool up = z < kZKmag;
        double x0 = up ? T.x0u : T.x0, tx = up ? T.txu : T.tx;
        double x = x0 + tx * z, y = T.y0 + T.ty * z;
        if (x < g.f[F_X1][d] || x > g.f[F_X2][d] || y < g.f[F_Y1][d] || y > g.f[F_Y2][d]) continue;
        int n = (int)g.f[F_NELEM][d];
        double u = x * g.f[F_COSTHETA][d] + y * g.f[F_SINTHETA][d];
        double sp = g.f[F_SPACING][d];
        int j0 = (int)std::lround((u - g.wire[d][1]) / sp) + 1;
        j0 = std::min(std::max(j0, 1), n);
        bool hodo = d >= 31 && d <= 46;
        int best = -1;
        double bd = 1e30, bdca = 0;
        for (int j = std::max(1, j0 - 2); j <= std::min(n, j0 + 2); ++j) {
            double dist = hodo ? (u - g.wire[d][j]) : wire_dca(g, d, j, x0, tx, T.y0, T.ty);
            if (std::fabs(dist) < bd) { bd = std::fabs(dist); best = j; bdca = dist; }
        }
        if (best < 0) continue;
        if (!hodo && bd > 0.5 * g.f[F_CELLWIDTH][d] + 0.5 * sp) continue;
        GHit h;
        h.det = (int16_t)d; h.elem = (int16_t)best;
        double tmin, tmax;
        tdc_window(d, tmin, tmax);
        if (hodo) { h.drift = 0.f; h.tdc = (float)tmin; }
        else {
            double dr = std::fabs(bdca) + P.drift_sigma * r.gauss();
            double half = 0.5 * g.f[F_CELLWIDTH][d];
            double dclip = std::min(std::max(std::fabs(dr), 0.0), half);
            h.drift = (float)dr;
            h.tdc = (float)(tmax - dclip * (tmax - tmin) / half);
        }
        h.seq = (int)out.size();
        out.push_back(h);
    }
}

void make_event(const or_geometry& g, const or_gen_params& P, int nchan, int64_t ev, std::vector<GHit>& hits,
                int8_t* ntrue, float* truth) {
    Rng r(P.seed, (uint64_t)ev);
    hits.clear();
    int nt = P.n_tracks_min + (int)(r.uni() * (P.n_tracks_max - P.n_tracks_min + 1));
    nt = std::min(std::max(nt, 0), OR_GEN_MAX_TRACKS);
    int made = 0;
    for (int k = 0; k < nt; ++k) {
        TrackPar T;
        if (!draw_track(g, P, r, T)) continue;
        track_hits(g, P, r, T, hits);
        float* t = truth + 6 * made;
        t[0] = (float)T.x0; t[1] = (float)T.tx; t[2] = (float)T.y0; t[3] = (float)T.ty; t[4] = (float)T.invp;
        t[5] = (float)T.q;
        ++made;
    }
    *ntrue = (int8_t)made;
    const int ntrk = (int)hits.size();
    if (P.mult_mode == 0) {
        double p = P.noise_occupancy;
        if (p > 0) {
            double lq = std::log1p(-p);
            long k = -1;
            for (;;) {
                double u = r.uni();
                k += 1 + (long)std::floor(std::log(1.0 - u) / lq);
                if (k >= nchan) break;
                int d, e;
                channel_of(g, (int)k, d, e);
                noise_hit(g, r, d, e, hits);
            }
        }
    } else {
        double mu = P.mult_mean;
        if (P.mult_mode == 2 && r.uni() < P.mult_frac2) mu = P.mult_mean2;
        int ntot = std::min(r.poisson(mu), P.mult_cap);
        int nbase = (int)std::lround(ntot / (1.0 + P.afterpulse_frac));
        for (int k = ntrk; k < nbase; ++k) {
            int d, e;
            channel_of(g, (int)(r.uni() * nchan), d, e);
            noise_hit(g, r, d, e, hits);
        }
    }
    // after-pulses: duplicate right after the parent
    std::vector<GHit> withap;
    withap.reserve(hits.size() * 2);
    for (auto& h : hits) {
        withap.push_back(h);
        if (r.uni() < P.afterpulse_frac) {
            GHit a = h;
            a.tdc = h.tdc + (float)r.uni(10.0, 200.0);
            a.drift = h.det >= 31 && h.det <= 46 ? 0.f : (float)r.uni(0.0, 0.5 * g.f[F_CELLWIDTH][h.det]);
            withap.push_back(a);
        }
    }
    if (P.mult_mode != 0 && (int)withap.size() > P.mult_cap) {
        // keep the track hits, trim from the end (noise and its after-pulses)
        withap.resize(std::max(P.mult_cap, 0));
    }
    for (size_t i = 0; i < withap.size(); ++i) withap[i].seq = (int)i;
    std::stable_sort(withap.begin(), withap.end(), [](const GHit& a, const GHit& b) {
        return a.det != b.det ? a.det < b.det : a.elem < b.elem;
    });
    hits.swap(withap);
}

}  // namespace

extern "C" {

void or_gen_params_default(or_gen_params* p) {
    p->seed = 0xE1039ull;
    p->n_tracks_min = 1;
    p->n_tracks_max = 2;
    p->tx_max = 0.15f;
    p->ty_max = 0.10f;
    p->invp_min = 0.02f;
    p->invp_max = 0.18f;
    p->drift_sigma = 0.04f;
    p->noise_occupancy = 0.02f;
    p->afterpulse_frac = 0.05f;
    p->mult_mode = 0;
    p->mult_mean = 300.f;
    p->mult_mean2 = 1500.f;
    p->mult_frac2 = 0.2f;
    p->mult_cap = 1000;
}

or_status or_generate(const or_geometry* g, const or_gen_params* p, int64_t n_events, int nthreads,
                      or_gen_events* out) {
    if (!g || !p || !out || n_events < 0) return OR_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof *out);
    const int nchan = channel_count(*g);
    int nth = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nth < 1) nth = 1;
    if (nth > 64) nth = 64;
    // pass 1: generate into per-chunk buffers
    const int64_t chunk = 4096;
    const int64_t nchunks = (n_events + chunk - 1) / chunk;
    std::vector<std::vector<GHit>> buf(nchunks);
    std::vector<std::vector<int32_t>> cnt(nchunks);
    out->n_events = n_events;
    out->n_true = (int8_t*)std::malloc(n_events > 0 ? n_events : 1);
    out->truth = (float*)std::calloc((size_t)(n_events > 0 ? n_events : 1) * OR_GEN_MAX_TRACKS * 6, sizeof(float));
    out->event_offsets = (int64_t*)std::malloc(sizeof(int64_t) * (n_events + 1));
    if (!out->n_true || !out->truth || !out->event_offsets) { or_gen_free(out); return OR_ERR_NOMEM; }
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        std::vector<GHit> h;
        for (;;) {
            int64_t c = next.fetch_add(1);
            if (c >= nchunks) break;
            int64_t e0 = c * chunk, e1 = std::min(n_events, e0 + chunk);
            auto& B = buf[c];
            auto& C = cnt[c];
            C.resize(e1 - e0);
            for (int64_t e = e0; e < e1; ++e) {
                make_event(*g, *p, nchan, e, h, out->n_true + e, out->truth + e * OR_GEN_MAX_TRACKS * 6);
                C[e - e0] = (int32_t)h.size();
                B.insert(B.end(), h.begin(), h.end());
            }
        }
    };
    std::vector<std::thread> th;
    for (int i = 0; i < nth; ++i) th.emplace_back(work);
    for (auto& t : th) t.join();
    int64_t total = 0;
    for (auto& B : buf) total += (int64_t)B.size();
    out->n_hits = total;
    size_t nb = (size_t)(total > 0 ? total : 1);
    out->det = (int16_t*)std::malloc(nb * 2);
    out->elem = (int16_t*)std::malloc(nb * 2);
    out->tdc = (float*)std::malloc(nb * 4);
    out->drift = (float*)std::malloc(nb * 4);
    out->flag = (uint8_t*)std::malloc(nb);
    if (!out->det || !out->elem || !out->tdc || !out->drift || !out->flag) { or_gen_free(out); return OR_ERR_NOMEM; }
    int64_t o = 0, e = 0;
    out->event_offsets[0] = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        for (int32_t k : cnt[c]) { o += k; out->event_offsets[++e] = o; }
    }
    // pass 2: copy out (parallel over chunks)
    std::vector<int64_t> cbase(nchunks + 1, 0);
    for (int64_t c = 0; c < nchunks; ++c) cbase[c + 1] = cbase[c] + (int64_t)buf[c].size();
    next = 0;
    auto copy = [&]() {
        for (;;) {
            int64_t c = next.fetch_add(1);
            if (c >= nchunks) break;
            int64_t b = cbase[c];
            for (size_t i = 0; i < buf[c].size(); ++i) {
                const GHit& h = buf[c][i];
                out->det[b + i] = h.det; out->elem[b + i] = h.elem; out->tdc[b + i] = h.tdc;
                out->drift[b + i] = h.drift; out->flag[b + i] = 2;
            }
            std::vector<GHit>().swap(buf[c]);
        }
    };
    th.clear();
    for (int i = 0; i < nth; ++i) th.emplace_back(copy);
    for (auto& t : th) t.join();
    return OR_OK;
}

void or_gen_free(or_gen_events* ev) {
    if (!ev) return;
    std::free(ev->det); std::free(ev->elem); std::free(ev->tdc); std::free(ev->drift); std::free(ev->flag);
    std::free(ev->event_offsets); std::free(ev->n_true); std::free(ev->truth);
    std::memset(ev, 0, sizeof *ev);
}
}

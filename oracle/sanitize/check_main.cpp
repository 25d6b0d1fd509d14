// TEST INFRASTRUCTURE ONLY: driver of the oracle under the compiler's
// sanitizers.  Generates events with the product's host generator, runs the
// oracle chain (full chain with vertexing) and the oracle dimuons on them with
// one and with several OpenMP threads, and requires identical results.  Built
// with AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer
// by the Makefile beside it; run by tests/test_sanitizers.py.
#include "../../include/e1039_or.h"
#include "../../include/e1039_or_gen.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {
struct orc_geometry;
int orc_geometry_load(const char* path, orc_geometry** out);
void orc_geometry_free(orc_geometry* G);
void orc_geometry_export(const orc_geometry* G, float* planes, float* wire, float* deltaW);
int orc_reconstruct(const orc_geometry* G, const or_params* P, const or_hits* in, or_result* out, int nthreads);
void orc_dimuons(int64_t n_events, int max_out, const int32_t* n_tracks, const uint32_t* status, const or_track* tracks,
                 const float* vertex, const int16_t* det, const int64_t* offsets, const float* z_array,
                 const int32_t* trigger, int32_t* n_dim, float* dim, float* hist, int64_t* lost_out);
}

namespace {

struct Out {
    std::vector<uint32_t> status, keep;
    std::vector<int32_t> ntr, ntkl;
    std::vector<or_track> tracks;
    std::vector<float> rec, vtx;
    std::vector<int32_t> ndim;
    std::vector<float> dim, hist;
};

Out run(const orc_geometry* G, const or_params& P, const or_hits& in, int threads, const float* z) {
    Out o;
    const int64_t ne = in.n_events;
    o.status.assign(ne, 0);
    o.keep.assign((in.n_hits + 31) / 32 + 1, 0);
    o.ntr.assign(ne, 0);
    o.ntkl.assign(ne, 0);
    o.tracks.assign(ne * P.max_tracks_out, or_track{});
    o.rec.assign(ne * (int64_t)P.max_tracklets_per_event * OR_TRACK_PARAMS, 0.f);
    o.vtx.assign(ne * (int64_t)P.max_tracks_out * 6, 0.f);
    or_result r;
    std::memset(&r, 0, sizeof r);
    r.status = o.status.data(); r.keep_mask = o.keep.data(); r.n_tracks = o.ntr.data(); r.tracks = o.tracks.data();
    r.n_tracklets = o.ntkl.data(); r.records = o.rec.data(); r.vertex = o.vtx.data();
    if (orc_reconstruct(G, &P, &in, &r, threads) != 0) {
        std::fprintf(stderr, "orc_reconstruct failed\n");
        std::exit(2);
    }
    o.ndim.assign(ne, 0);
    o.dim.assign(ne * (int64_t)OR_DIMUONS_MAX * OR_DIMUON_FIELDS, 0.f);
    o.hist.assign((size_t)OR_HIST_TRIGGERS * OR_HIST_VARS * OR_HIST_BINS, 0.f);
    int64_t lost = 0;
    orc_dimuons(ne, P.max_tracks_out, o.ntr.data(), o.status.data(), o.tracks.data(), o.vtx.data(), in.det,
                in.event_offsets, z, nullptr, o.ndim.data(), o.dim.data(), o.hist.data(), &lost);
    return o;
}

template <typename T>
bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

int check_case(const char* name, const or_geometry* pg, const orc_geometry* G, const or_params& P, int64_t n,
               float mean, int cap, int threads, const float* z) {
    or_gen_params gp;
    or_gen_params_default(&gp);
    gp.seed = 0x5A17 + (uint64_t)n;
    gp.mult_mode = 1;
    gp.mult_mean = mean;
    gp.mult_cap = cap;
    gp.n_tracks_min = 1;
    gp.n_tracks_max = 3;
    or_gen_events ev;
    if (or_generate(pg, &gp, n, 1, &ev) != OR_OK) {
        std::fprintf(stderr, "%s: or_generate failed\n", name);
        return 2;
    }
    or_hits in;
    std::memset(&in, 0, sizeof in);
    in.n_events = ev.n_events; in.n_hits = ev.n_hits; in.det = ev.det; in.elem = ev.elem; in.tdc = ev.tdc;
    in.drift = ev.drift; in.flag = ev.flag; in.event_offsets = ev.event_offsets;
    const Out a = run(G, P, in, 1, z), b = run(G, P, in, threads, z);
    int64_t tracks = 0, dims = 0;
    for (int32_t t : a.ntr) tracks += t;
    for (int32_t t : a.ndim) dims += t;
    const bool ok = same(a.status, b.status) && same(a.keep, b.keep) && same(a.ntr, b.ntr) && same(a.ntkl, b.ntkl) &&
                    std::memcmp(a.tracks.data(), b.tracks.data(), a.tracks.size() * sizeof(or_track)) == 0 &&
                    same(a.rec, b.rec) && same(a.vtx, b.vtx) && same(a.ndim, b.ndim) && same(a.dim, b.dim) &&
                    same(a.hist, b.hist);
    std::printf("%s: %lld events, %lld hits, %lld tracks, %lld dimuons, 1 vs %d threads %s\n", name,
                (long long)ev.n_events, (long long)ev.n_hits, (long long)tracks, (long long)dims, threads,
                ok ? "identical" : "DIFFER");
    or_gen_free(&ev);
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s geometry_file [threads [events events_high]]\n", argv[0]);
        return 2;
    }
    const int threads = argc > 2 ? std::atoi(argv[2]) : 4;
    const int64_t n1 = argc > 3 ? std::atoll(argv[3]) : 300, n2 = argc > 4 ? std::atoll(argv[4]) : 12;
    or_geometry* pg = nullptr;
    orc_geometry* G = nullptr;
    if (or_geometry_load(argv[1], &pg) != OR_OK || orc_geometry_load(argv[1], &G) != 0) {
        std::fprintf(stderr, "geometry load failed\n");
        return 2;
    }
    std::vector<float> planes(27 * OR_N_DETECTORS), wire(OR_N_DETECTORS * OR_MAX_ELEM), dw(OR_N_DETECTORS * 9);
    orc_geometry_export(G, planes.data(), wire.data(), dw.data());
    const float* z = planes.data();   // field 0: plane z
    or_params P;
    or_params_default(&P);
    P.full_chain = 1;
    P.max_hits_per_event = 2048;
    int rc = check_case("reference caps, Poisson(300)", pg, G, P, n1, 300.f, 1000, threads, z);
    // lifted caps at high occupancy (the runtime-sized vectors of the oracle)
    or_params H = P;
    H.max_hits_chamber_plane = H.max_hits_hodo_plane = H.max_hits_prop_plane = 400;
    H.max_reduced_per_plane = H.max_pair_plane_hits = 254;
    H.max_pairs_per_bin = H.max_pairs_window = 255;
    H.max_tracklets_per_slot = 16384;
    H.max_tracklets_per_event = 2048;
    H.max_d0_mult = H.max_d2_mult = H.max_d3_mult = 1 << 30;
    H.max_tracks_out = 64;
    H.max_hits_per_event = 8192;
    rc |= check_case("lifted caps, Poisson(1200)", pg, G, H, n2, 1200.f, 4000, threads, z);
    orc_geometry_free(G);
    or_geometry_free(pg);
    return rc;
}

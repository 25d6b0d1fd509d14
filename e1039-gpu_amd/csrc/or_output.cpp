// Columnar event output with the branches of the reference's OROutput tree
// (OROutput.cxx:24-102), filled from the library's outputs the way
// online_reconstruction.cu:1063-1222 fills the tree from the device arrays.
// One array per branch in an uncompressed .npz archive (numpy's format: a zip
// of .npy files), written when the writer is closed.
#include "or_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

// Station multiplicity cuts of the event reducer: the reference's
// HasTooManyHits (reconstruction_kernels.cuh:135-138); such events are not written.
constexpr uint32_t MULT_BITS = OR_EV_D0_MULT | OR_EV_D2_MULT | OR_EV_D3P_MULT | OR_EV_D3M_MULT;

struct Branch {
    std::string name;
    char kind;                    // 'i': int32, 'u': uint32, 'f': float32
    std::vector<uint32_t> v;      // raw 32-bit values
};

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320) as the zip format uses it
uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

// .npy version 1.0 header of a 1-D little-endian array, padded to 64 bytes
std::string npy_header(char kind, size_t n) {
    const char* descr = kind == 'i' ? "<i4" : (kind == 'u' ? "<u4" : "<f4");
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" +
                       std::to_string(n) + ",), }";
    size_t total = 10 + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict.append(pad, ' ');
    dict.push_back('\n');
    std::string h("\x93NUMPY\x01\x00", 8);
    const uint16_t len = (uint16_t)dict.size();
    h.push_back((char)(len & 0xFF));
    h.push_back((char)(len >> 8));
    return h + dict;
}

void put16(std::string& s, uint16_t v) {
    s.push_back((char)(v & 0xFF));
    s.push_back((char)(v >> 8));
}
void put32(std::string& s, uint32_t v) {
    for (int k = 0; k < 4; ++k) s.push_back((char)((v >> (8 * k)) & 0xFF));
}

}  // namespace

struct or_output {
    std::string path;
    std::vector<Branch> br;
    int index(const char* name) const {
        for (size_t i = 0; i < br.size(); ++i)
            if (br[i].name == name) return (int)i;
        return -1;
    }
    void add(const char* name, char kind) { br.push_back({name, kind, {}}); }
    void pi(int b, int32_t x) { uint32_t u; std::memcpy(&u, &x, 4); br[b].v.push_back(u); }
    void pu(int b, uint32_t x) { br[b].v.push_back(x); }
    void pf(int b, float x) { uint32_t u; std::memcpy(&u, &x, 4); br[b].v.push_back(u); }
};

namespace {

// Branches in the order of OROutput.cxx:28-101, then the extra status branch.
const char* const TRACK_F[] = {"track.chi2", "track.x0", "track.y0", "track.tx", "track.ty", "track.invp",
                               "track.err_x0", "track.err_y0", "track.err_tx", "track.err_ty", "track.err_invp"};
// record fields of the above (online_reconstruction.cu:1147-1157)
const int TRACK_F_FIELD[] = {3, 7, 8, 5, 6, 9, 12, 13, 10, 11, 14};
const char* const TRACK_V[] = {"track.vx", "track.vy", "track.vz", "track.px", "track.py", "track.pz"};
const char* const DIM_NAMES[] = {
    "dimuon.mass", "dimuon.pT", "dimuon.xF", "dimuon.x1", "dimuon.x2", "dimuon.costheta", "dimuon.phi",
    "dimuon.mass_single", "dimuon.chi2_single", "dimuon.vx", "dimuon.vy", "dimuon.vz", "dimuon.px", "dimuon.py",
    "dimuon.pz", "dimuon.pos_vx", "dimuon.pos_vy", "dimuon.pos_vz", "dimuon.neg_vx", "dimuon.neg_vy",
    "dimuon.neg_vz", "dimuon.pos_e", "dimuon.pos_px", "dimuon.pos_py", "dimuon.pos_pz", "dimuon.neg_e",
    "dimuon.neg_px", "dimuon.neg_py", "dimuon.neg_pz", "dimuon.pos_single_e", "dimuon.pos_single_px",
    "dimuon.pos_single_py", "dimuon.pos_single_pz", "dimuon.neg_single_e", "dimuon.neg_single_px",
    "dimuon.neg_single_py", "dimuon.neg_single_pz", "dimuon.chi2_vtx", "dimuon.chi2_kf"};
// record field of each dimuon branch (online_reconstruction.cu:1184-1219); -1: px, py, pz
// are the sums of the two muons' components (fields 18 + 22, 19 + 23, 20 + 24)
const int DIM_FIELD[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, -1, -1, -1, 12, 13, 14, 15, 16,
                         17, 21, 18, 19, 20, 25, 22, 23, 24, 29, 26, 27, 28, 33, 30, 31, 32, 34, 35};
static_assert(sizeof(DIM_NAMES) / sizeof(DIM_NAMES[0]) == sizeof(DIM_FIELD) / sizeof(DIM_FIELD[0]),
              "dimuon branch table");

void declare(or_output* w) {
    w->add("evtid", 'i');
    w->add("nhits", 'i');
    w->add("hit.detid", 'i');
    w->add("hit.chan", 'i');
    w->add("hit.pos", 'f');
    w->add("hit.tdc", 'f');
    w->add("hit.drift", 'f');
    w->add("ntracks", 'i');
    w->add("track.stid", 'i');
    w->add("track.nhits", 'i');
    for (const char* n : TRACK_F) w->add(n, 'f');
    w->add("track.charge", 'i');
    for (const char* n : TRACK_V) w->add(n, 'f');
    w->add("track.hitdetid", 'i');
    w->add("track.hitchan", 'i');
    w->add("track.hitpos", 'f');
    w->add("track.hitdrift", 'f');
    w->add("track.hitsign", 'i');
    w->add("track.hittdc", 'f');
    w->add("track.hitresid", 'f');
    w->add("track.invp_tgt", 'f');
    w->add("ndimuons", 'i');
    for (const char* n : DIM_NAMES) w->add(n, 'f');
    w->add("status", 'u');
}

}  // namespace

extern "C" {

or_status or_output_open(const char* path, or_output** out) {
    if (!path || !out) return OR_ERR_INVALID_ARG;
    std::FILE* f = std::fopen(path, "wb");   // fail early on an unwritable path
    if (!f) return OR_ERR_IO;
    std::fclose(f);
    auto* w = new (std::nothrow) or_output;
    if (!w) return OR_ERR_NOMEM;
    w->path = path;
    declare(w);
    *out = w;
    return OR_OK;
}

or_status or_output_fill(or_output* w, const or_geometry* g, const or_hits* in, const or_result* chain,
                         int32_t max_tracklets_per_event, const or_dimuon_result* dim, int32_t min_station_id) {
    if (!w || !g || !in || !chain) return OR_ERR_INVALID_ARG;
    if (in->on_device || chain->on_device) return OR_ERR_INVALID_ARG;
    if (in->n_events < 0 || in->n_hits < 0 || max_tracklets_per_event < 1) return OR_ERR_INVALID_ARG;
    if (in->n_events > 0 && (!in->event_offsets || !chain->status || !chain->keep_mask || !chain->records ||
                             !chain->n_tracklets))
        return OR_ERR_INVALID_ARG;
    if (in->n_hits > 0 && (!in->det || !in->elem || !in->tdc || !in->drift)) return OR_ERR_INVALID_ARG;
    if (dim && (!dim->n_dimuons || !dim->dimuons)) return OR_ERR_INVALID_ARG;
    try {
        const int b_evtid = w->index("evtid"), b_nhits = w->index("nhits"), b_hdet = w->index("hit.detid");
        const int b_ntr = w->index("ntracks"), b_stid = w->index("track.stid"), b_tnh = w->index("track.nhits");
        const int b_tf = w->index(TRACK_F[0]), b_charge = w->index("track.charge"), b_tv = w->index(TRACK_V[0]);
        const int b_th = w->index("track.hitdetid"), b_ndim = w->index("ndimuons"), b_dim = w->index(DIM_NAMES[0]);
        const int b_status = w->index("status");
        std::vector<int64_t> byplane[OR_N_DETECTORS];
        for (int64_t e = 0; e < in->n_events; ++e) {
            const int64_t h0 = in->event_offsets[e], h1 = in->event_offsets[e + 1];
            const uint32_t st = chain->status[e];
            // the reference skips events without hits and those its reducer flags
            // (online_reconstruction.cu:1066-1072)
            if (h1 <= h0 || (st & MULT_BITS)) continue;
            w->pi(b_evtid, in->event_id ? in->event_id[e] : (int32_t)e);
            // hits the reducer kept, detector by detector, in staging order
            for (auto& v : byplane) v.clear();
            for (int64_t h = h0; h < h1; ++h) {
                const int d = in->det[h];
                if (d < 1 || d >= OR_N_DETECTORS) continue;
                if ((chain->keep_mask[h >> 5] >> (h & 31)) & 1u) byplane[d].push_back(h);
            }
            int32_t nh = 0;
            for (int d = 1; d < OR_N_DETECTORS; ++d) {
                for (int64_t h : byplane[d]) {
                    const int el = in->elem[h];
                    w->pi(b_hdet, d);
                    w->pi(b_hdet + 1, el);
                    w->pf(b_hdet + 2, (el >= 0 && el < OR_MAX_ELEM) ? g->wire[d][el] : 0.f);
                    w->pf(b_hdet + 3, in->tdc[h]);
                    w->pf(b_hdet + 4, std::fabs(in->drift[h]));
                    ++nh;
                }
            }
            w->pi(b_nhits, nh);
            // tracks: tracklet records of stationID >= min_station_id, slot order
            // (online_reconstruction.cu:1123-1170)
            const int nrec = std::min(chain->n_tracklets[e], max_tracklets_per_event);
            const float* rec = chain->records + e * (int64_t)max_tracklets_per_event * OR_TRACK_PARAMS;
            int32_t ntr = 0;
            for (int k = 0; k < nrec; ++k) {
                const float* r = rec + (int64_t)k * OR_TRACK_PARAMS;
                if ((int)r[0] < min_station_id) continue;
                ++ntr;
                const int nht = (int)r[2];
                w->pi(b_stid, (int)r[0]);
                w->pi(b_tnh, nht);
                for (int q = 0; q < 11; ++q) w->pf(b_tf + q, r[TRACK_F_FIELD[q]]);
                w->pi(b_charge, (int)r[15]);
                for (int q = 0; q < 6; ++q) w->pf(b_tv + q, r[16 + q]);
                for (int l = 0; l < nht && l < OR_MAX_HITS_PER_TRACK; ++l) {
                    w->pi(b_th, (int)r[22 + l]);
                    w->pi(b_th + 1, (int)r[40 + l]);
                    w->pf(b_th + 2, r[58 + l]);
                    w->pf(b_th + 3, r[76 + l]);
                    w->pi(b_th + 4, (int)r[94 + l]);
                    w->pf(b_th + 5, r[112 + l]);
                    w->pf(b_th + 6, r[130 + l]);
                }
            }
            w->pi(b_ntr, ntr);
            // dimuons (online_reconstruction.cu:1178-1220)
            int32_t nd = 0;
            if (dim) {
                nd = std::min(dim->n_dimuons[e], (int32_t)OR_DIMUONS_MAX);
                const float* drow = dim->dimuons + e * (int64_t)OR_DIMUONS_MAX * OR_DIMUON_FIELDS;
                for (int k = 0; k < nd; ++k) {
                    const float* r = drow + (int64_t)k * OR_DIMUON_FIELDS;
                    for (int q = 0; q < (int)(sizeof(DIM_FIELD) / sizeof(DIM_FIELD[0])); ++q) {
                        const int f = DIM_FIELD[q];
                        const float v = f >= 0 ? r[f] : r[18 + (q - 12)] + r[22 + (q - 12)];
                        w->pf(b_dim + q, v);
                    }
                }
            }
            w->pi(b_ndim, nd);
            w->pu(b_status, st);
        }
    } catch (const std::bad_alloc&) {
        return OR_ERR_NOMEM;
    }
    return OR_OK;
}

or_status or_output_close(or_output* w) {
    if (!w) return OR_ERR_INVALID_ARG;
    or_status st = OR_OK;
    std::FILE* f = std::fopen(w->path.c_str(), "wb");
    if (!f) {
        delete w;
        return OR_ERR_IO;
    }
    std::string central;
    uint64_t offset = 0;
    bool ok = true;
    for (const Branch& b : w->br) {
        const std::string name = b.name + ".npy";
        const std::string head = npy_header(b.kind, b.v.size());
        const uint64_t bytes = (uint64_t)head.size() + 4ull * b.v.size();
        if (bytes >= 0xFFFFFFFFull || offset >= 0xFFFFFFFFull) {   // no zip64 records: 4 GiB per member and offset
            st = OR_ERR_CAPACITY;
            ok = false;
            break;
        }
        uint32_t crc = crc32_update(0u, reinterpret_cast<const unsigned char*>(head.data()), head.size());
        crc = crc32_update(crc, reinterpret_cast<const unsigned char*>(b.v.data()), 4 * b.v.size());
        std::string local;
        put32(local, 0x04034b50u);   // local file header
        put16(local, 20);            // version needed
        put16(local, 0);             // flags
        put16(local, 0);             // stored
        put16(local, 0);             // time
        put16(local, 0x21);          // date: 1980-01-01
        put32(local, crc);
        put32(local, (uint32_t)bytes);
        put32(local, (uint32_t)bytes);
        put16(local, (uint16_t)name.size());
        put16(local, 0);
        local += name;
        ok = std::fwrite(local.data(), 1, local.size(), f) == local.size() &&
             std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
             (b.v.empty() || std::fwrite(b.v.data(), 4, b.v.size(), f) == b.v.size());
        if (!ok) break;
        put32(central, 0x02014b50u);   // central directory entry
        put16(central, 20);            // version made by
        put16(central, 20);
        put16(central, 0);
        put16(central, 0);
        put16(central, 0);
        put16(central, 0x21);
        put32(central, crc);
        put32(central, (uint32_t)bytes);
        put32(central, (uint32_t)bytes);
        put16(central, (uint16_t)name.size());
        put16(central, 0);    // extra
        put16(central, 0);    // comment
        put16(central, 0);    // disk
        put16(central, 0);    // internal attributes
        put32(central, 0);    // external attributes
        put32(central, (uint32_t)offset);
        central += name;
        offset += local.size() + bytes;
    }
    if (ok && offset + central.size() >= 0xFFFFFFFFull) {
        st = OR_ERR_CAPACITY;
        ok = false;
    }
    if (ok) {
        std::string end;
        put32(end, 0x06054b50u);   // end of central directory
        put16(end, 0);
        put16(end, 0);
        put16(end, (uint16_t)w->br.size());
        put16(end, (uint16_t)w->br.size());
        put32(end, (uint32_t)central.size());
        put32(end, (uint32_t)offset);
        put16(end, 0);
        ok = std::fwrite(central.data(), 1, central.size(), f) == central.size() &&
             std::fwrite(end.data(), 1, end.size(), f) == end.size();
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok && st == OR_OK) st = OR_ERR_IO;
    delete w;
    return st;
}

}  // extern "C"

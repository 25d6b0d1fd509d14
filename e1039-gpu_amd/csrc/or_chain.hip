// The reconstruction chain on CDNA4: one 64-lane wave per event, persistent
// grid, per-event hits staged in LDS, per-wave tracklet scratch in HBM.
//
// Stages per event (reference kernels in brackets):
//   A  load + stable grouping by detector           [host staging, online_reconstruction.cu:506-640]
//   B  per-plane reduction, one lane per plane      [gkernel_eR, reconstruction_kernels.cuh:8-172]
//   C  binned X pair lists (st2, D3p, D3m)          [make_hitpairs_in_station_bins, reconstruction_helper.cuh:368-460]
//   D  XZ combinations, one lane per combination    [gKernel_XZ_tracking, reconstruction_kernels.cuh:182-855]
//   E  YZ: one lane per (tracklet, u3.v3 pair)      [gKernel_YZ_tracking, reconstruction_kernels.cuh:864-1832]
//   F  back-partial cleaning                        [reconstruction_helper.cuh:1789-1909]
//   G  station-1 extension, one lane per track      [gKernel_Global_tracking, reconstruction_kernels.cuh:1841-2534]
//   H  outputs: status, keep bits, track records
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "or_device.h"
#include "or_prune.h"

namespace ordev {
__constant__ float c_geo[F_COUNT][ND];
// per-plane constants of the final track fit (hi/lo float pairs, see final_fit)
constexpr int FITP_M = 0, FITP_V = 20, FITP_H = 28, FITP_W = 36, FITP_N = 37;
constexpr double FIT_ZREF = 1000.0;
__constant__ float c_fitp[ND][FITP_N];
__constant__ float c_fit_cond = OR_FIT_COND_FP32;   // FP32 / FP64 switch of the final fit

// R-T calibration on the device: per detector n knots x[off .. off+n) and the
// cubic (y, b, c, d) of each knot interval at cf[off + k]; n = 0: not applied.
// x0, xl: first and last knot; scale = (n - 1) / (xl - x0) for the interval guess.
struct RtHead {
    int32_t n, off;
    float tmin, tmax;
    double x0, xl, scale;
};
struct RtDev {
    RtHead h[ND];
    const double* x;
    const double4* cf;
    int32_t nk;   // knots of all detectors
};

// Drift of one hit (online_reconstruction.cu:389-397): tdc < tmin -> cellwidth/2,
// tdc > tmax -> 0, else the spline in FP64.  Knot interval k: the largest with
// t_k < t (TSpline3's binary search), the first interval below the first knot
// and the last one above the last knot.  The interval is guessed from the
// mean knot spacing and corrected against the knots, which gives the same k.
__device__ __forceinline__ float rt_drift(const RtHead& H, const double* x, const double4* cf, int d, float tdc,
                                          float drift) {
    const int n = H.n;
    if (n <= 0) return drift;
    if (tdc < H.tmin) return c_geo[F_CELLWIDTH][d] * 0.5f;
    if (tdc > H.tmax) return 0.f;
    const double t = (double)tdc;
    const double* X = x + H.off;
    int k = 0;
    if (t >= H.xl) {
        k = n - 2;
    } else if (t > H.x0) {
        k = min(max((int)((t - H.x0) * H.scale), 0), n - 2);
        while (k > 0 && !(X[k] < t)) --k;
        while (k < n - 2 && X[k + 1] < t) ++k;
    }
    const double4 c = cf[H.off + k];
    const double dt = t - X[k];
    return (float)(c.x + dt * (c.y + dt * (c.z + dt * c.w)));
}

constexpr int RT_BLOCK = 256, RT_UNROLL = 8;
// knots staged in LDS by rt_calibrate_kernel<true> (tables with more use global memory)
constexpr int RT_LDS_KNOTS = 8192;

// RT_UNROLL hits per lane in flight; the per-detector heads in LDS, and with
// XLDS the knots too (the interval search then reads LDS, and one global load
// per hit, its interval's cubic, remains)
template <bool XLDS>
__global__ __launch_bounds__(RT_BLOCK) void rt_calibrate_kernel(const RtDev* R, const int16_t* det, const float* tdc,
                                                                const float* drift, int64_t n, float* out) {
    __shared__ RtHead H[ND];
    extern __shared__ double rt_x[];
    for (int i = threadIdx.x; i < ND; i += RT_BLOCK) H[i] = R->h[i];
    const double* X = R->x;
    const double4* CF = R->cf;
    if (XLDS) {
        const int nk = R->nk;
        for (int i = threadIdx.x; i < nk; i += RT_BLOCK) rt_x[i] = X[i];
    }
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * RT_BLOCK * RT_UNROLL; base < n;
         base += (int64_t)gridDim.x * RT_BLOCK * RT_UNROLL) {
        int d[RT_UNROLL];
        float t[RT_UNROLL], dr[RT_UNROLL];
#pragma unroll
        for (int j = 0; j < RT_UNROLL; ++j) {
            const int64_t i = base + j * RT_BLOCK + threadIdx.x;
            d[j] = 0; t[j] = 0.f; dr[j] = 0.f;
            if (i < n) { d[j] = det[i]; t[j] = tdc[i]; dr[j] = drift[i]; }
        }
#pragma unroll
        for (int j = 0; j < RT_UNROLL; ++j) {
            const int64_t i = base + j * RT_BLOCK + threadIdx.x;
            if (i < n)
                out[i] = (d[j] >= 1 && d[j] < ND) ? rt_drift(H[d[j]], XLDS ? rt_x : X, CF, d[j], t[j], dr[j]) : dr[j];
        }
    }
}
}

using namespace ordev;

namespace {

constexpr int WAVE = 64;
// Bounds set by index encodings, not by buffers (buffers are sized from the
// session's or_params): in-plane hit indices are 8 bits (0xFF = none), the
// (u, v) and (u3, v3) enumeration indices of a list pair are 16 bits, slot
// order entries are 16 bits.
constexpr int PAIRS_MAX = 255;        // max_pairs_per_bin, max_pairs_window
constexpr int TKL_MAX = 16384;        // max_tracklets_per_event
constexpr int PLANE_CAP = 400;
constexpr int YZ_CHUNK = 16;  // tracklets whose window lists are built together
static_assert((YZ_CHUNK & (YZ_CHUNK - 1)) == 0, "YZ_CHUNK: a power of two (binary searches over a chunk)");
constexpr int YZ_LIST_MIN = 4096;     // survivor list capacity per wave at session creation (longer chunks stream)
constexpr int YZ_BATCH_DEFAULT = 1024;   // survivor batch of the YZ passes (fastest at 4 waves per event)
constexpr int YZ_LIST_MAX = 1 << 20;
// a tracklet whose four window lists span more than this many (u3, v3, u2, v2)
// candidates is an enumeration overflow (keeps chunk sums within 32 bits)
constexpr int64_t YZ_ENUM_MAX = 1ll << 27;
constexpr int NCLASS_MAX = 5;    // event tile classes: 512, 1024, 2048, 4096, 8192 hits

// Sub-stage cycle counters of the profiling build (-DOR_STAGE_TIMING); no-ops
// otherwise.  Indices: YZ window lists, YZ pass 1, 2, 3 (its left/right and
// chi2 part), YZ winners, YZ unit path, global window lists, global pass A,
// global pass B and write, YZ unit scoring, then counts (see OR_COUNT), pass 3's
// sign-free tests (cycles) and the survivors passing them (count).
constexpr int N_SUB_TIMERS = 18;
#ifdef OR_STAGE_TIMING
__device__ unsigned long long g_sub_cycles[N_SUB_TIMERS];
#define OR_SUB_START unsigned long long s_0 = clock64(), s_1;
#define OR_SUB(k)                                                              \
    do {                                                                       \
        s_1 = clock64();                                                       \
        if ((threadIdx.x & 63) == 0) atomicAdd(&g_sub_cycles[k], s_1 - s_0);          \
        s_0 = s_1;                                                             \
    } while (0)
// counts instead of cycles in slots 10..14 (YZ candidates, units scored,
// units kept, candidates of kept units, first-cut survivors)
#define OR_COUNT(k, v)                                                         \
    do {                                                                       \
        if ((threadIdx.x & 63) == 0) atomicAdd(&g_sub_cycles[k], (unsigned long long)(v)); \
    } while (0)
#else
#define OR_SUB_START
#define OR_SUB(k) do { } while (0)
#define OR_COUNT(k, v) do { } while (0)
#endif  // YZ candidates passing the first box cut, per chunk
constexpr uint16_t NOHIT = 0xFFFF;

// Tracklet as kept in scratch between stages.
struct __align__(16) Tk {
    float x0, tx, ex0, etx;
    float y0, ty, ey0, ety;
    float chi2, invp, einvp, pad;
    float vtx[6];           // vertexing: vx, vy, vz, px, py, pz (stationID 7)
    float xpos[4];          // x at y = 0 of the XZ hits (record 'pos' of those hits)
    float resid[18];
    uint16_t hraw[18];      // LDS raw index of each hit
    uint8_t det[18];
    int8_t sign[18];
    uint32_t canon;         // position in the canonical XZ list
    int8_t stid;            // 0 after XZ, 3 rejected/demoted, 5 YZ, 6 global
    uint8_t nx, nuv, n1;
    uint8_t t, st3, slot;
    int8_t charge;
    uint16_t pos_in_slot;
    uint8_t pad2[2];
};
static_assert(sizeof(Tk) == 256, "tracklet record size");

constexpr int CLEAN_WORDS = 64;   // lane-parallel cleaning of up to 64 x 64 live tracklets

// Waves per event: a tile class runs with 1, 2 or 4 waves per workgroup (one
// event per workgroup).  The YZ and station-1 stages hand their tracklet
// chunks to all of them; the other stages run on the first wave.
constexpr int WPE_MAX = 8;
// waves per event a tile class is built for at most: the 512-hit tile one, the
// 1,024- and 2,048-hit tiles four, the 4,096- and 8,192-hit tiles eight (their
// LDS admits two and one workgroups per CU, so eight waves keep four and two
// per SIMD)
__host__ __device__ constexpr int lds_waves(int maxh) { return maxh <= 512 ? 1 : (maxh <= 2048 ? 4 : WPE_MAX); }
__host__ __device__ constexpr int lds_waves_log(int maxh) { return maxh <= 512 ? 0 : (maxh <= 2048 ? 2 : 3); }

// YZ state of one wave's chunk: window list lengths, candidate or unit offsets
// per tracklet, best unit score (pruning) and best candidate key per tracklet.
struct YzWave {
    int nl[YZ_CHUNK][4];      // window list lengths (read as int4)
    int uoff[YZ_CHUNK + 1];   // prefix of (u3, v3) units or of candidates per tracklet of the chunk
    int unoff[YZ_CHUNK + 1];  // prefix of (u3, v3) units per tracklet of the chunk (scoring)
    int umax[YZ_CHUNK];       // pruning: best unit score of each tracklet of the chunk (order-preserving int)
    unsigned long long best[YZ_CHUNK];
    uint32_t hx[YZ_CHUNK];    // hodoscope x-match mask of each tracklet of the chunk (yz_cuts_hx)
    // x hits of each tracklet of the chunk by fixed slot D2 X, X', D3 X, X' (raw
    // index, NOHIT: none) and their signs (yz_chi2_st)
    uint16_t xh[YZ_CHUNK][4];
    int8_t xs[YZ_CHUNK][4];
    // 1/d rounded to double of the enumeration radices (see qdiv): u2 list
    // length, (u2, v2) pairs per unit, u3 list length
    double rx[YZ_CHUNK], rin[YZ_CHUNK], rz[YZ_CHUNK];
};

// floor(a / d) for 0 <= a < 2^27 and 1 <= d < 2^16, from r = 1/d rounded to
// double.  a*r + 2^-20 (one rounding, fused) lies within 2^-25 of
// a/d + 2^-20: above floor(a/d), and below floor(a/d) + 1 because the fraction
// of a/d is 0 or at most 1 - 1/d < 1 - 2^-16.  The YZ enumeration indices stay
// below 2^27 (the enumeration bound) and the radices below 255^2.
__device__ __forceinline__ int qdiv(int a, double r) { return (int)__builtin_fma((double)a, r, 0x1p-20); }
__device__ __forceinline__ double qrcp(int d) { return 1.0 / (double)(d > 0 ? d : 1); }

// The event tile.  Hits are staged grouped by detector; after the reducer the
// kept ones are compacted in place, so plane d's reduced hits are
// [rstart[d], rstart[d] + rcnt[d]) of elem/pos/drift/orig.  Sized so that the
// 512-hit tile stays within 8 KB (20 waves per CU).
// Prop-tube Y planes of the YZ stage's tests (P1Y 48, 47; P2Y 54, 53): their
// reduced hit positions sorted, per event, when a plane holds at most
// psort_cap hits (else the plane is scanned hit by hit)
__host__ __device__ constexpr int psort_cap(int maxh) { return maxh <= 512 ? 32 : 128; }
__device__ __forceinline__ constexpr int psort_plane(int p) { return p == 0 ? 48 : p == 1 ? 47 : p == 2 ? 54 : 53; }

template <int MAXH>
struct Lds {
    int16_t elem[MAXH];
    float pos[MAXH];
    float drift[MAXH];
    uint16_t orig[MAXH];      // event-local input index
    int8_t keep[MAXH];        // staged: 1 in time, -1 out of time; after reduction: 1 kept, -1 dropped
    uint32_t kbits[MAXH / 32];  // reducer keep bit by event-local input index
    int pcnt[56];
    int16_t pstart[56], cursor[56], rcnt[56], rstart[56];
    uint16_t xcnt[3][28];
    float psort[4][psort_cap(MAXH)];   // sorted positions of plane psort_plane(p)
    int pscnt[4];                      // their count, -1: more than psort_cap hits
    union alignas(16) {
        YzWave yw[lds_waves(MAXH)];   // YZ: one per wave; station 1 reads yw[0].nl as 64 list lengths
        uint16_t live[WAVE];      // cleaning: tracklets at stationID >= 5, slot order
        struct {
            unsigned long long alive[CLEAN_WORDS];   // cleaning of more than 64: alive bits by live position
            int lstart[33];                          // first live position of each slot
        } cl;
    };
    int sstart[33];
    uint32_t status;
    int ntkl;                 // tracklets after XZ (read by every wave of the event)
    int yz_next;              // next YZ chunk to hand out
    int gl_next;              // next block of tracklets of the station-1 stage
    int next_ev;              // work queue: the event the group takes next
};

// Per-wave scratch in HBM: one block per resident wave, regions at offsets the
// host computes from the session's caps (ScrLayout); the workgroup keeps the
// addresses of its first wave's regions in LDS (g_cfg), from which each wave
// finds its own, so the stage functions take no pointer arguments for them.
enum ScrRegion {
    R_TK,     // Tk[maxtkl]: the event's tracklets
    R_CELL,   // uint32[1569]: XZ cell offsets
    R_LIST,   // uint16[64][win_cap]: window pair lists (YZ: tracklet x view; station 1: lane)
    R_YV,     // float4[64][win_cap]: per YZ list entry (w y of hit a, w y of hit b, flag byte, 0); station-1 records
    R_YY,     // float2[64][win_cap] (+ 16): y of YZ list entries' hits a, b (station 2: the entry's mean y, 0)
    R_YF,     // uint8[64][win_cap]: accepted-hit flags of YZ list entries | first-fit pattern term << 2
    R_XP,     // uint16[3][28][bin_cap]: binned X pairs
    R_SO,     // uint16[maxtkl]: tracklet index by slot order
    R_PL,     // uint32[list_cap]: YZ first-box survivors of a batch; station-1 pass-A list
    R_PB,     // uchar4[list_cap]: survivor sign state
    R_KL,     // uint32[unit_cap]: kept YZ units of a chunk
    R_KS,     // uint32[unit_cap]: their first candidate in the kept enumeration
    R_YC,     // uint4[3 * list_cap]: survivor candidate cache; station-1 pass-A tables
    R_QL,     // uint32[list_cap]: survivors passing the sign-free tests
    R_BR,     // uint4[YZ_CHUNK][4]: best candidate of each tracklet of a chunk so far
    R_CL,     // uint4[3 * maxtkl]: cleaning, hit per detector 13-30 of each live tracklet
    R_COUNT
};
struct ScrLayout {
    int64_t off[R_COUNT];
    int64_t stride;
    int32_t maxtkl, bin_cap, win_cap, list_cap, unit_cap;
};

struct Args {
    const int16_t* det;
    const int16_t* elem;
    const float* tdc;
    const float* drift;
    const uint8_t* flag;
    const int64_t* off;
    int64_t n_events;
    const int32_t* ev_list;     // events of this launch's size class (NULL: all events in order)
    const int32_t* ev_count;    // device count of ev_list
    int32_t* ev_head;           // work queue head (events handed out beyond the first gridDim.x), zeroed per launch
    const float* wire;          // 55 x OR_MAX_ELEM
    const RtDev* rt;            // R-T calibration applied at load (NULL: input drift)
    or_params p;
    uint32_t* status;
    uint32_t* keep_mask;
    int32_t* n_tracks;
    or_track* tracks;
    int32_t* n_tracklets;
    float* records;
    float* vertex;              // optional: swim global tracks to the beam line
    float* fit;                 // optional: final straight-line fit of each output track
    char* scratch;
    ScrLayout lay;              // scratch layout of this launch's tile class
    int yz_list_cap;            // <= lay.list_cap: survivor batch size of the YZ passes
    const PrunerDev* pruner;    // learned YZ unit pruning (NULL: every unit)
    float* unit_dump;           // optional per-unit feature rows (training aid)
    int64_t dump_cap;
    unsigned long long* dump_count;
};

// regions of the event, shared by its waves; the others are per wave
__host__ __device__ constexpr bool scr_shared(int r) {
    return r == R_TK || r == R_CELL || r == R_XP || r == R_SO || r == R_CL;
}

// region addresses of the workgroup's first wave and the caps (written at
// kernel start, one record per workgroup whatever its waves): wave w's own
// regions lie w scratch blocks (stride bytes each) further
struct WaveCfg {
    char* rp[R_COUNT];
    int64_t stride;
    int32_t maxtkl, bin_cap, win_cap, list_cap, unit_cap;
};
__shared__ WaveCfg g_cfg;

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
// wave of the workgroup (wave-uniform, in a scalar register)
__device__ __forceinline__ int wave_idx() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x) >> 6; }
__device__ __forceinline__ const WaveCfg& cfg() { return g_cfg; }
// address of the wave's scratch region R
template <int R>
__device__ __forceinline__ char* region() {
    return scr_shared(R) ? g_cfg.rp[R] : g_cfg.rp[R] + (int64_t)wave_idx() * g_cfg.stride;
}
// A value the whole wave holds alike, as the wave's scalar: arguments of the
// non-inlined stage functions are taken as per-lane by the compiler, so their
// uniform ones (event and chunk bounds, counts, cursors, shared pointers) are
// marked at entry; loops and index arithmetic on them then run on the scalar unit.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint32_t lo = uni((uint32_t)(uint64_t)v), hi = uni((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
    return reinterpret_cast<T*>((uintptr_t)uni((int64_t)(uintptr_t)p));
}
__device__ __forceinline__ int rank_in(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
// inclusive wave prefix sum
__device__ __forceinline__ int wave_scan(int v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(v, o, WAVE);
        if (l >= o) v += t;
    }
    return v;
}
// Orders the lanes' LDS and memory accesses of one wave without a workgroup
// barrier: writes made by any lane before it are seen by every lane after it.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// float <-> int with the same order (for integer atomic max of float values)
__device__ __forceinline__ int float_order(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float float_unorder(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// ---------------------------------------------------------------------------
// Hit access helpers on the staged event
// The event tile lives in a namespace-scope LDS variable so that non-inlined
// stage functions keep LDS addressing (ds_* instructions) without a pointer.
template <int MAXH>
__shared__ Lds<MAXH> g_lds;

// the calling wave's YZ state
template <int MAXH>
__device__ __forceinline__ YzWave& yzw() { return g_lds<MAXH>.yw[lds_waves(MAXH) == 1 ? 0 : wave_idx()]; }

template <int MAXH, int WPE>
struct Ev {
    struct Ref {
        __device__ __forceinline__ Lds<MAXH>* operator->() const { return &g_lds<MAXH>; }
    };
    static constexpr Ref Lp{};
    __device__ __forceinline__ int nred(int d) const { return g_lds<MAXH>.rcnt[d]; }
    __device__ __forceinline__ int raw(int d, int k) const { return g_lds<MAXH>.rstart[d] + k; }
};

// XZ bin of an element (WCHitsBins_X, reconstruction_constants.h:155-162):
// 4-wire bins from 1, the last D3 bin spans 109-116.
__device__ __forceinline__ int xbin(int s, int c) {
    if (c < 1) return -1;
    int b = (c - 1) >> 2;
    if (b > 27) b = 27;
    const int lo = 1 + 4 * b;
    const int hi = (b == 27 && s > 0) ? 116 : 4 + 4 * b;
    return (lo <= c && c <= hi) ? b : -1;
}

// hodoscope X match for an XZ candidate (reconstruction_helper.cuh:1090-1144)
template <int MAXH, int WPE>
__device__ bool hodo_xz(const Ev<MAXH, WPE>& E, int st, int d, float x0, float tx, float ex0, float etx) {
    const int n = E.nred(d);
    if (n == 0) return false;
    if (d > 40) st = 3;
    const float xh = G(F_Z, d) * tx + x0;
    const float f = hodo_fudge(st), cw = hodo_xcell(st);
    float ex = 3.f * (fabsf(G(F_Z, d) * etx) + ex0);
    // hits in order (the window widening accumulates over them); positions
    // read four at a time so one LDS wait covers four hits
    const int r0 = E.raw(d, 0);
    for (int i0 = 0; i0 < n; i0 += 4) {
        float pv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) pv[k] = E.Lp->pos[r0 + min(i0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k >= n) break;
            const float p = pv[k];
            float xmin = p - cw * 0.5f, xmax = p + cw * 0.5f;
            xmin -= (xmax - xmin) * f;
            xmax += (xmax - xmin) * f;
            ex += (xmax - xmin) * 0.15f;
            xmin -= ex;
            xmax += ex;
            if (xmin <= xh && xh <= xmax) return true;
        }
    }
    return false;
}

// hodoscope match for a tracklet (reconstruction_helper.cuh:1002-1087)
template <int MAXH, int WPE>
__device__ bool hodo_tkl(const Ev<MAXH, WPE>& E, int st, int d, float x0, float y0, float tx, float ty, float ex0, float ey0,
                         float etx, float ety) {
    const int n = E.nred(d);
    if (n == 0) return false;
    if (d > 40) st = 3;
    const float f = hodo_fudge(st), cw = G(F_CELLWIDTH, d);
    const float xh = G(F_Z, d) * tx + x0, yh = G(F_Z, d) * ty + y0;
    float ex = 3.f * (fabsf(G(F_Z, d) * etx) + ex0);
    const float ey = 3.f * (fabsf(G(F_Z, d) * ety) + ey0);
    const int r0 = E.raw(d, 0);
    for (int i0 = 0; i0 < n; i0 += 4) {
        float pv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) pv[k] = E.Lp->pos[r0 + min(i0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k >= n) break;
            const float p = pv[k];
            float xmin = p - cw * 0.5f, xmax = p + cw * 0.5f;
            float ymin = G(F_Y1, d), ymax = G(F_Y2, d);
            xmin -= (xmax - xmin) * f;
            xmax += (xmax - xmin) * f;
            ymin -= (ymax - ymin) * f;
            ymax += (ymax - ymin) * f;
            ex += (xmax - xmin) * 0.15f;
            xmin -= ex;
            xmax += ex;
            ymin -= ey;
            ymax += ey;
            if (xmin <= xh && xh <= xmax && ymin <= yh && yh <= ymax) return true;
        }
    }
    return false;
}

// The x half of hodo_tkl: whether some hit of plane d matches the tracklet's
// x line (x0, tx and their errors) with the window widening accumulated over
// the hits in order.  hodo_tkl's y window is the same for every hit of a plane
// (plane ends widened by the fudge factor and the y error), so hodo_tkl is
// (this) && hodo_y; in the YZ stage the x line is the tracklet's and fixed, so
// this half is computed once per tracklet and plane.
template <int MAXH, int WPE>
__device__ bool hodo_x(const Ev<MAXH, WPE>& E, int st, int d, float x0, float tx, float ex0, float etx) {
    const int n = E.nred(d);
    if (n == 0) return false;
    if (d > 40) st = 3;
    const float f = hodo_fudge(st), cw = G(F_CELLWIDTH, d);
    const float xh = G(F_Z, d) * tx + x0;
    float ex = 3.f * (fabsf(G(F_Z, d) * etx) + ex0);
    const int r0 = E.raw(d, 0);
    for (int i0 = 0; i0 < n; i0 += 4) {
        float pv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) pv[k] = E.Lp->pos[r0 + min(i0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k >= n) break;
            const float p = pv[k];
            float xmin = p - cw * 0.5f, xmax = p + cw * 0.5f;
            xmin -= (xmax - xmin) * f;
            xmax += (xmax - xmin) * f;
            ex += (xmax - xmin) * 0.15f;
            xmin -= ex;
            xmax += ex;
            if (xmin <= xh && xh <= xmax) return true;
        }
    }
    return false;
}

// the y half of hodo_tkl (the window every hit of plane d shares)
__device__ __forceinline__ bool hodo_y(int st, int d, float y0, float ty, float ey0, float ety) {
    if (d > 40) st = 3;
    const float f = hodo_fudge(st);
    const float yh = G(F_Z, d) * ty + y0;
    const float ey = 3.f * (fabsf(G(F_Z, d) * ety) + ey0);
    float ymin = G(F_Y1, d), ymax = G(F_Y2, d);
    ymin -= (ymax - ymin) * f;
    ymax += (ymax - ymin) * f;
    ymin -= ey;
    ymax += ey;
    return ymin <= yh && yh <= ymax;
}

// hodoscope planes of the YZ stage's matches, in the order they are tested:
// bit b of a tracklet's x-match mask is plane yz_hodo(b) (37, 38, 39, 40, 45, 46)
__device__ __forceinline__ constexpr int yz_hodo(int b) { return b < 4 ? 37 + b : 41 + b; }

// a prop-tube hit within win of the expected coordinate, plane a then b; hit
// positions read four at a time so one LDS wait covers four hits
template <int MAXH, int WPE>
__device__ __forceinline__ bool prop_plane_hit(const Ev<MAXH, WPE>& E, int d, float expect, float win) {
    const int n = E.nred(d), r0 = E.raw(d, 0);
    for (int m0 = 0; m0 < n; m0 += 4) {
        float pv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) pv[k] = E.Lp->pos[r0 + min(m0 + k, n - 1)];
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) hit = hit || (m0 + k < n && fabsf(pv[k] - expect) < win);
        if (hit) return true;
    }
    return false;
}
template <int MAXH, int WPE>
__device__ __forceinline__ int prop_hit(const Ev<MAXH, WPE>& E, int da, int db, float slope, float icpt, float win) {
    if (prop_plane_hit(E, da, slope * G(F_Z, da) + icpt, win)) return 1;
    return prop_plane_hit(E, db, slope * G(F_Z, db) + icpt, win) ? 1 : 0;
}

// prop_plane_hit over the sorted positions s[0, n) of a plane (n wave-uniform).
// fl(p - e) is non-decreasing in p, so the smallest |fl(p - e)| over the plane
// is that of the last position below e or of the first at or above it: the
// test of those two is the test of every hit.  NaN or infinite e fails both,
// as it fails every hit.
__device__ __forceinline__ bool prop_sorted_hit(const float* s, int n, float e, float win) {
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (s[lo + half] < e) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const float below = s[max(lo - 1, 0)], above = s[min(lo, max(n - 1, 0))];
    return (lo > 0 && fabsf(below - e) < win) || (lo < n && fabsf(above - e) < win);
}

// min(dmin, |p - e|) over the sorted positions s[0, n) of a plane: by the same
// monotonicity, the smallest |fl(p - e)| is that of one of e's two neighbours
__device__ __forceinline__ float prop_sorted_nearest(const float* s, int n, float e, float dmin) {
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (s[lo + half] < e) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    const float below = s[max(lo - 1, 0)], above = s[min(lo, max(n - 1, 0))];
    if (lo > 0) dmin = fminf(dmin, fabsf(below - e));
    if (lo < n) dmin = fminf(dmin, fabsf(above - e));
    return dmin;
}

// prop_hit for the planes psort_plane(2 pp) then psort_plane(2 pp + 1)
template <int MAXH, int WPE>
__device__ __forceinline__ int prop_hit_y(const Ev<MAXH, WPE>& E, int pp, float slope, float icpt, float win) {
    const Lds<MAXH>& L = g_lds<MAXH>;
    const int da = psort_plane(2 * pp), db = psort_plane(2 * pp + 1);
    const int na = L.pscnt[2 * pp], nb = L.pscnt[2 * pp + 1];
    const float ea = slope * G(F_Z, da) + icpt;
    if (na >= 0 ? prop_sorted_hit(L.psort[2 * pp], na, ea, win) : prop_plane_hit(E, da, ea, win)) return 1;
    const float eb = slope * G(F_Z, db) + icpt;
    return (nb >= 0 ? prop_sorted_hit(L.psort[2 * pp + 1], nb, eb, win) : prop_plane_hit(E, db, eb, win)) ? 1 : 0;
}

// ---- window pair list (make_hitpairs_in_station, reconstruction_helper.cuh:464-549)
// Returns the number of pairs, or -1 when a cap is exceeded.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) int pairs_window(const Ev<MAXH, WPE>& E, int st, int proj, float xmin, float xmax,
                                                      uint16_t* out, int cap_plane, int cap_pairs) {
    const int sid = super_id(st, proj);
    const int d1 = sid * 2, d2 = sid * 2 - 1;
    const int n1 = E.nred(d1), n2 = E.nred(d2);
    const float dx1 = G(F_DELTAPX, d1), dx2 = G(F_DELTAPX, d2);
    // wire span inside [xmin, xmax]; the first 64 hits of a plane keep the answer in a bit mask
    auto inwin = [&](int d, float dx, int i) {
        const int e = E.Lp->elem[E.raw(d, i)];
        const float b = x_bep(d, e), t = x_bep(d, e) + dx;
        const float lo = dx > 0 ? b : t, hi = dx > 0 ? t : b;
        return lo <= xmax && hi >= xmin;
    };
    unsigned long long w1 = 0, w2 = 0;
    int c1 = 0, c2 = 0;
    for (int i = 0; i < n1; ++i)
        if (inwin(d1, dx1, i)) { ++c1; if (i < 64) w1 |= 1ull << i; }
    for (int i = 0; i < n2; ++i)
        if (inwin(d2, dx2, i)) { ++c2; if (i < 64) w2 |= 1ull << i; }
    if (c1 > cap_plane || c2 > cap_plane) return -1;
    const float sp = spacing_plane(sid);
    int np = 0;
    if (n1 <= 64 && n2 <= 64) {
        // only the in-window hits, in index order; the paired hits are marked
        // while the pairs are listed
        unsigned long long p1 = 0, p2 = 0;
        for (unsigned long long m1 = w1; m1; m1 &= m1 - 1) {
            const int i = __ffsll((long long)m1) - 1;
            const float pi = E.Lp->pos[E.raw(d1, i)];
            for (unsigned long long m2 = w2; m2; m2 &= m2 - 1) {
                const int j = __ffsll((long long)m2) - 1;
                if (fabsf(pi - E.Lp->pos[E.raw(d2, j)]) > sp) continue;
                p1 |= 1ull << i;
                p2 |= 1ull << j;
                if (np >= cap_pairs) return -1;
                out[np++] = (uint16_t)(i | (j << 8));
            }
        }
        for (unsigned long long m = w1 & ~p1; m; m &= m - 1) {
            if (np >= cap_pairs) return -1;
            out[np++] = (uint16_t)((__ffsll((long long)m) - 1) | (0xFF << 8));
        }
        for (unsigned long long m = w2 & ~p2; m; m &= m - 1) {
            if (np >= cap_pairs) return -1;
            out[np++] = (uint16_t)(0xFF | ((__ffsll((long long)m) - 1) << 8));
        }
        return np;
    }
    auto in1 = [&](int i) { return i < 64 ? ((w1 >> i) & 1ull) != 0 : inwin(d1, dx1, i); };
    auto in2 = [&](int i) { return i < 64 ? ((w2 >> i) & 1ull) != 0 : inwin(d2, dx2, i); };
    for (int i = 0; i < n1; ++i) {
        if (!in1(i)) continue;
        const float pi = E.Lp->pos[E.raw(d1, i)];
        for (int j = 0; j < n2; ++j) {
            if (!in2(j)) continue;
            if (fabsf(pi - E.Lp->pos[E.raw(d2, j)]) > sp) continue;
            if (np >= cap_pairs) return -1;
            out[np++] = (uint16_t)(i | (j << 8));
        }
    }
    // hits in no pair: X (a) then X' (b)
    for (int i = 0; i < n1; ++i) {
        if (!in1(i)) continue;
        const float pi = E.Lp->pos[E.raw(d1, i)];
        bool paired = false;
        for (int j = 0; j < n2 && !paired; ++j) paired = in2(j) && !(fabsf(pi - E.Lp->pos[E.raw(d2, j)]) > sp);
        if (paired) continue;
        if (np >= cap_pairs) return -1;
        out[np++] = (uint16_t)(i | (0xFF << 8));
    }
    for (int j = 0; j < n2; ++j) {
        if (!in2(j)) continue;
        const float pj = E.Lp->pos[E.raw(d2, j)];
        bool paired = false;
        for (int i = 0; i < n1 && !paired; ++i) paired = in1(i) && !(fabsf(E.Lp->pos[E.raw(d1, i)] - pj) > sp);
        if (paired) continue;
        if (np >= cap_pairs) return -1;
        out[np++] = (uint16_t)(0xFF | (j << 8));
    }
    return np;
}

__device__ __forceinline__ int pa(uint16_t q) { int a = q & 0xFF; return a == 0xFF ? -1 : a; }
__device__ __forceinline__ int pb(uint16_t q) { int b = q >> 8; return b == 0xFF ? -1 : b; }

// ---------------------------------------------------------------------------
// XZ candidate (reconstruction_kernels.cuh:465-672)
template <int MAXH, int WPE>
__device__ __forceinline__ bool xz_candidate(const Ev<MAXH, WPE>& E, int s3, uint16_t q2, uint16_t q3, Tk* out) {
    // hits in fixed slots D2 X, X', D3 X, X' (the reference's list order); see lr_slots_new
    const int d3x = s3 == 1 ? 22 : 28;
    int det[4] = {16, 15, d3x, d3x - 1};
    const int idx[4] = {pa(q2), pb(q2), pa(q3), pb(q3)};
    const unsigned pres = (idx[0] >= 0 ? 1u : 0u) | (idx[1] >= 0 ? 2u : 0u) | (idx[2] >= 0 ? 4u : 0u) |
                          (idx[3] >= 0 ? 8u : 0u);
    if (!(pres & 3u) || !(pres & 12u)) return false;
    int sign[4], raw[4];
    float X[4], Z[4], ex[4], drift[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        sign[s] = 0;
        raw[s] = 0;
        X[s] = Z[s] = drift[s] = 0.f;
        ex[s] = G(F_RESOLUTION, det[s]);
        if ((pres >> s) & 1u) {
            const int r = E.raw(det[s], idx[s]);
            raw[s] = r;
            drift[s] = E.Lp->drift[r];
            xz_fy(det[s], E.Lp->elem[r], drift[s], 0, 0.f, 0.f, X[s], Z[s]);
        }
    }
    auto fit = [&](const float (&x)[4]) {
        float A0 = 0.f, A1 = 0.f, A3 = 0.f, B0 = 0.f, B1 = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (!((pres >> s) & 1u)) continue;
            B0 += ex[s] * x[s];
            B1 += (ex[s] * x[s]) * Z[s];
            A0 += ex[s];
            A1 += ex[s] * Z[s];
            A3 += (ex[s] * Z[s]) * Z[s];
        }
        return fit_finish(A0, A1, A3, B0, B1);
    };
    Fit f = fit(X);
    if (fabsf(f.p0) > X0_MAX + 2.f * f.e0 || fabsf(f.p1) > TX_MAX + 2.f * f.e1) return false;
    // left/right of x pairs (x-only variant), then single hits
    bool used = false;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const bool start = !used && ((pres >> s) & 1u) && ((pres >> (s + 1)) & 1u) && abs(det[s] - det[s + 1]) == 1;
        if (start) {
            const int h = lr_xpair_decide(f.p0, f.p1, f.e0, f.e1, det[s], det[s + 1], X[s], X[s + 1], drift[s],
                                          drift[s + 1], 150.f);
            if (h >= 0) { sign[s] = lr_si(h); sign[s + 1] = lr_sj(h); }
        }
        used = start;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (((pres >> s) & 1u) && sign[s] == 0) sign[s] = G(F_Z, det[s]) * f.p1 + f.p0 > X[s] ? 1 : -1;
    float xc[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) xc[s] = X[s] + drift[s] * (float)sign[s];
    f = fit(xc);
    const float x0 = f.p0, tx = f.p1, ex0 = f.e0, etx = f.e1;
    if (fabsf(x0) > X0_MAX || fabsf(tx) > TX_MAX) return false;
    // prop X: P1X (50 then 49) within 5.08, P2X (52 then 51) within 7.62
    const int nprop = prop_hit(E, 50, 49, tx, x0, 5.08f) + prop_hit(E, 52, 51, tx, x0, 7.62f);
    if (nprop < 2) return false;
    if (!hodo_xz(E, 1, 37, x0, tx, ex0, etx) && !hodo_xz(E, 1, 38, x0, tx, ex0, etx)) return false;
    if (!hodo_xz(E, 2, 39, x0, tx, ex0, etx) && !hodo_xz(E, 2, 40, x0, tx, ex0, etx)) return false;
    if (!hodo_xz(E, 3, 45, x0, tx, ex0, etx) && !hodo_xz(E, 3, 46, x0, tx, ex0, etx)) return false;
    if (out) {
        Tk& T = *out;
        T.x0 = x0; T.tx = tx; T.ex0 = ex0; T.etx = etx;
        T.y0 = T.ty = T.ey0 = T.ety = 0.f;
        T.chi2 = T.invp = T.einvp = 0.f;
        T.nuv = 0; T.n1 = 0;
        T.stid = 0; T.charge = 0;
        int n = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (!((pres >> s) & 1u)) continue;
            T.xpos[n] = X[s];
            T.hraw[n] = (uint16_t)raw[s];
            T.det[n] = (uint8_t)det[s];
            T.sign[n] = (int8_t)sign[s];
            ++n;
        }
        T.nx = (uint8_t)n;
        for (int h = 0; h < 18; ++h) {
            T.resid[h] = 0.f;
            if (h >= n) {
                if (h < 4) T.xpos[h] = 0.f;
                T.hraw[h] = NOHIT;
                T.det[h] = 0;
                T.sign[h] = 0;
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// YZ candidates (reconstruction_kernels.cuh:1320-1800).
//  * The y of every UV hit depends only on the hit and the tracklet's (x0, tx)
//    (calculate_y_uvhit with sign 0), so it is computed once per window-list
//    entry and read here.
//  * The first fit runs over hits in the order u3, u3', v3, v3', u2, u2', v2,
//    v2' (the same sequence of float additions as the reference).
//  * The refit after left/right (reconstruction_kernels.cuh:1628) has the same
//    inputs as the first fit and is not repeated.
// Loads from the wave's scratch (global memory) issued as global, not flat,
// operations: flat ones also count against the LDS counter that shuffles and
// LDS reads wait on.
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ float4 gld4(const float4* p) {
    const __attribute__((address_space(1))) float* f = (const __attribute__((address_space(1))) float*)p;
    return make_float4(f[0], f[1], f[2], f[3]);
}
__device__ __forceinline__ float2 gld2(const float2* p) {
    const __attribute__((address_space(1))) float* f = (const __attribute__((address_space(1))) float*)p;
    return make_float2(f[0], f[1]);
}
__device__ __forceinline__ uint4 gld4(const uint4* p) {
    const __attribute__((address_space(1))) uint32_t* f = (const __attribute__((address_space(1))) uint32_t*)p;
    return make_uint4(f[0], f[1], f[2], f[3]);
}

struct YzLists {
    const uint16_t* l[4];   // pair entries (a | b << 8), views u2, v2, u3, v3
    const float4* y[4];     // (w y of hit a, w y of hit b, flag byte bits, 0)
    const uint8_t* f[4];    // bit 0: hit a present and accepted, bit 1: hit b; bits 2-7: pattern term (yz_pattern)
    int n[4];
};

// Weight of a UV hit of plane d in the YZ fits: the y error calculate_y_uvhit
// gives it (reconstruction_helper.cuh), resolution x |dy/dx| of the wire, a
// constant of the plane.
__device__ __forceinline__ float yz_w(int d) { return G(F_RESOLUTION, d) * fabsf(G(F_DELTAPY, d) / G(F_DELTAPX, d)); }

// fit sums of one list entry's accepted hits (planes da, db), in order a then b
__device__ __forceinline__ int yz_sum(const YzLists& W, int v, int k, int da, int db, float& A0, float& A1, float& A3,
                                      float& B0, float& B1) {
    const uint8_t fl = gld(W.f[v] + k);
    if (!fl) return 0;
    const float4 yv = gld4(W.y[v] + k);
    int c = 0;
    if (fl & 1) {
        const float w = yz_w(da), z = G(F_Z, da);
        B0 += yv.x; B1 += yv.x * z; A0 += w; A1 += w * z; A3 += (w * z) * z; ++c;
    }
    if (fl & 2) {
        const float w = yz_w(db), z = G(F_Z, db);
        B0 += yv.y; B1 += yv.y * z; A0 += w; A1 += w * z; A3 += (w * z) * z; ++c;
    }
    return c;
}

// Hits of one candidate in fixed slots u3, u3', v3, v3', u2, u2', v2, v2' (the
// reference's list order with absent hits left out); bit s of pres = slot s
// holds a hit.  Only compile-time slot indices are used, so this stays in
// registers.
struct YzSlots {
    unsigned pres;
    int det[8], sign[8], raw[8];
    float pos[8], drift[8];
};

__device__ __forceinline__ void yz_dets(YzSlots& H, int st3) {
    const int su3 = super_id(st3, 1) * 2, sv3 = super_id(st3, 2) * 2;
    const int su2 = super_id(2, 1) * 2, sv2 = super_id(2, 2) * 2;
    H.det[0] = su3; H.det[1] = su3 - 1; H.det[2] = sv3; H.det[3] = sv3 - 1;
    H.det[4] = su2; H.det[5] = su2 - 1; H.det[6] = sv2; H.det[7] = sv2 - 1;
}

// accepted hits of entry k of view v into slots S0 (hit a) and S0 + 1 (hit b), signs 0
template <int S0, int MAXH, int WPE>
__device__ __forceinline__ void yz_fill(const Ev<MAXH, WPE>& E, const YzLists& W, int v, int k, YzSlots& H) {
    const uint16_t q = gld(W.l[v] + k);
    const unsigned fl = gld(W.f[v] + k);
    H.pres &= ~(3u << S0);
    H.sign[S0] = 0;
    H.sign[S0 + 1] = 0;
    if (fl & 1u) {
        const int r = E.raw(H.det[S0], pa(q));
        H.pres |= 1u << S0;
        H.raw[S0] = r; H.pos[S0] = E.Lp->pos[r]; H.drift[S0] = E.Lp->drift[r];
    }
    if (fl & 2u) {
        const int r = E.raw(H.det[S0 + 1], pb(q));
        H.pres |= 2u << S0;
        H.raw[S0 + 1] = r; H.pos[S0 + 1] = E.Lp->pos[r]; H.drift[S0 + 1] = E.Lp->drift[r];
    }
}

// slots where the pairing walk starts a pair (see lr_slots_new)
template <int NS>
__device__ __forceinline__ unsigned lr_slot_starts(unsigned pres, const int (&det)[NS]) {
    unsigned m = 0;
    bool used = false;
#pragma unroll
    for (int s = 0; s + 1 < NS; ++s) {
        const bool start = !used && ((pres >> s) & 1u) && ((pres >> (s + 1)) & 1u) && abs(det[s] - det[s + 1]) == 1;
        if (start) m |= 1u << s;
        used = start;
    }
    return m;
}

// The tests of the tail that need no left/right sign: second box cut, prop-tube
// Y and hodoscope matches, on the x line and the first fit alone.
template <int MAXH, int WPE>
__device__ __forceinline__ bool yz_cuts(const Ev<MAXH, WPE>& E, const Tk* Tp, const Fit& f) {
    const float x0 = Tp->x0, tx = Tp->tx, ex0 = Tp->ex0, etx = Tp->etx;
    const float y0 = f.p0, ty = f.p1, ey0 = f.e0, ety = f.e1;
    if (fabsf(y0) > Y0_MAX || fabsf(ty) > TY_MAX) return false;
    // prop Y: P1Y (48 then 47) within 5.08, P2Y (54 then 53) within 10.16
    const int nprop = prop_hit(E, 48, 47, ty, y0, 5.08f) + prop_hit(E, 54, 53, ty, y0, 10.16f);
    if (nprop < 2) return false;
    if (!hodo_tkl(E, 1, 37, x0, y0, tx, ty, ex0, ey0, etx, ety) &&
        !hodo_tkl(E, 1, 38, x0, y0, tx, ty, ex0, ey0, etx, ety))
        return false;
    if (!hodo_tkl(E, 2, 39, x0, y0, tx, ty, ex0, ey0, etx, ety) &&
        !hodo_tkl(E, 2, 40, x0, y0, tx, ty, ex0, ey0, etx, ety))
        return false;
    if (!hodo_tkl(E, 3, 45, x0, y0, tx, ty, ex0, ey0, etx, ety) &&
        !hodo_tkl(E, 3, 46, x0, y0, tx, ty, ex0, ey0, etx, ety))
        return false;
    return true;
}

// yz_cuts with the hodoscope x halves given as the tracklet's mask (bit b:
// plane yz_hodo(b) matches the x line): the same decision
template <int MAXH, int WPE>
__device__ __forceinline__ bool yz_cuts_hx(const Ev<MAXH, WPE>& E, unsigned hx, const Fit& f) {
    const float y0 = f.p0, ty = f.p1, ey0 = f.e0, ety = f.e1;
    if (fabsf(y0) > Y0_MAX || fabsf(ty) > TY_MAX) return false;
    const int nprop = prop_hit_y(E, 0, ty, y0, 5.08f) + prop_hit_y(E, 1, ty, y0, 10.16f);
    if (nprop < 2) return false;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const bool a = ((hx >> (2 * s)) & 1u) && hodo_y(s + 1, yz_hodo(2 * s), y0, ty, ey0, ety);
        if (!a && !(((hx >> (2 * s + 1)) & 1u) && hodo_y(s + 1, yz_hodo(2 * s + 1), y0, ty, ey0, ety))) return false;
    }
    return true;
}

// chi2 over the x hits (stored signs) and the UV hits (signs in H); with Tw set
// the candidate is written into the tracklet
template <int MAXH, int WPE>
__device__ __forceinline__ void yz_chi2(const Ev<MAXH, WPE>& E, const Tk* Tp, const YzSlots& H, const Fit& f,
                                        float& chi2_out, Tk* Tw) {
    const float x0 = Tp->x0, tx = Tp->tx;
    const float y0 = f.p0, ty = f.p1, ey0 = f.e0, ety = f.e1;
    float chi2 = 0.f;
    const int nx = Tp->nx;
    for (int k = 0; k < nx; ++k) {
        const int d = Tp->det[k], r = Tp->hraw[k];
        const float res = G(F_RESOLUTION, d);
        const float rs = E.Lp->drift[r] * (float)Tp->sign[k] - dca_wire(d, E.Lp->elem[r], x0, y0, tx, ty);
        chi2 += ((rs * rs) / res) / res;
        if (Tw) Tw->resid[k] = rs;
    }
    int m = nx;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if (!((H.pres >> s) & 1u)) continue;
        const int d = H.det[s];
        const float res = G(F_RESOLUTION, d);
        const float rs = H.drift[s] * (float)H.sign[s] - dca_wire(d, E.Lp->elem[H.raw[s]], x0, y0, tx, ty);
        chi2 += ((rs * rs) / res) / res;
        if (Tw) {
            Tw->resid[m] = rs;
            Tw->hraw[m] = (uint16_t)H.raw[s];
            Tw->det[m] = (uint8_t)d;
            Tw->sign[m] = (int8_t)H.sign[s];
        }
        ++m;
    }
    if (Tw) {
        Tw->stid = 5;
        Tw->nuv = (uint8_t)(m - nx);
        Tw->chi2 = chi2; Tw->y0 = y0; Tw->ty = ty; Tw->ey0 = ey0; Tw->ety = ety;
    }
    chi2_out = chi2;
}

// The rest of a candidate after the first box cut: left/right on the UV hits
// (station-3 signs in H as they stand), second box cut, prop-tube Y and
// hodoscope matches, chi2 over the x hits (stored signs) and the UV hits
// (reconstruction_helper.cuh:1182-1204).  With Tw set the candidate is written
// into the tracklet.
template <int MAXH, int WPE>
__device__ __forceinline__ bool yz_tail(const Ev<MAXH, WPE>& E, const Tk* Tp, YzSlots& H, const Fit& f, float& chi2_out,
                                        Tk* Tw) {
    lr_slots_new<8>(Tp->x0, Tp->tx, f.p0, f.p1, Tp->ex0, Tp->etx, f.e0, f.e1, H.pres, H.det, H.pos, H.drift, H.sign,
                    150.f);
    if (!yz_cuts(E, Tp, f)) return false;
    yz_chi2(E, Tp, H, f, chi2_out, Tw);
    return true;
}

// YZ unit: one tracklet, one (u3, v3) pair, the (u2, v2) pairs in reference
// order with the station-3 signs carried from candidate to candidate, as the
// reference does.  search (stop < 0): the unit's best (chi2, i_uv2).  replay
// (stop >= 0): writes the candidate at i_uv2 == stop into the tracklet.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) bool yz_unit(const Ev<MAXH, WPE>& E, Tk* Tp, const YzLists& W, int iuv3, int stop,
                                                  float& best_chi2, int& best_i) {
    const int st3 = Tp->st3;
    const int su2 = super_id(2, 1) * 2, sv2 = super_id(2, 2) * 2;
    const int su3 = super_id(st3, 1) * 2, sv3 = super_id(st3, 2) * 2;
    const int nu2 = W.n[0], nv2 = W.n[1], nu3 = W.n[2];
    const int iu3 = iuv3 % nu3, iv3 = iuv3 / nu3;
    float sA0 = 0.f, sA1 = 0.f, sA3 = 0.f, sB0 = 0.f, sB1 = 0.f;
    if (yz_sum(W, 2, iu3, su3, su3 - 1, sA0, sA1, sA3, sB0, sB1) == 0) return false;
    if (yz_sum(W, 3, iv3, sv3, sv3 - 1, sA0, sA1, sA3, sB0, sB1) == 0) return false;
    YzSlots H;
    yz_dets(H, st3);
    H.pres = 0;
    yz_fill<0>(E, W, 2, iu3, H);
    yz_fill<2>(E, W, 3, iv3, H);
    const int last = stop >= 0 ? stop : nu2 * nv2 - 1;
    bool found = false;
    for (int iuv2 = 0; iuv2 <= last; ++iuv2) {
        const int iu2 = iuv2 % nu2, iv2 = iuv2 / nu2;
        float A0 = sA0, A1 = sA1, A3 = sA3, B0 = sB0, B1 = sB1;
        if (yz_sum(W, 0, iu2, su2, su2 - 1, A0, A1, A3, B0, B1) == 0) continue;
        if (yz_sum(W, 1, iv2, sv2, sv2 - 1, A0, A1, A3, B0, B1) == 0) continue;
        const Fit f = fit_finish(A0, A1, A3, B0, B1);
        if (fabsf(f.p0) > Y0_MAX + 2.f * f.e0 || fabsf(f.p1) > TY_MAX + 2.f * f.e1) continue;
        yz_fill<4>(E, W, 0, iu2, H);
        yz_fill<6>(E, W, 1, iv2, H);
        float chi2;
        const bool write = stop >= 0 && iuv2 == stop;
        if (!yz_tail(E, Tp, H, f, chi2, write ? Tp : nullptr)) continue;
        if (write) return true;
        if (chi2 < best_chi2) { best_chi2 = chi2; best_i = iuv2; found = true; }
    }
    return found;
}

// ---------------------------------------------------------------------------
// Stages.  Each is a separate (non-inlined) function so that no single
// function carries the live state of every stage; all lanes of the wave call
// each stage, and stage-level branches depend only on wave-uniform values.
#define STAGE __device__ __attribute__((noinline))

// Accessors of the wave's scratch regions (addresses from g_cfg, region<R>).
struct Scr {
    __device__ Tk* tks() const { return reinterpret_cast<Tk*>(region<R_TK>()); }
    __device__ uint32_t* celloff() const { return reinterpret_cast<uint32_t*>(region<R_CELL>()); }
    __device__ uint16_t* wl() const { return reinterpret_cast<uint16_t*>(region<R_LIST>()); }
    // per YZ list entry: w y of hit a and of hit b (w the plane's weight), its
    // flag byte (as wf holds it) in the bits of z, 0; the y values themselves in wyy
    __device__ float4* wy() const { return reinterpret_cast<float4*>(region<R_YV>()); }
    // (station 2: x = the entry's y, the mean of its accepted hits)
    __device__ float2* wyy() const { return reinterpret_cast<float2*>(region<R_YY>()); }
    __device__ uint8_t* wf() const { return reinterpret_cast<uint8_t*>(region<R_YF>()); }
    // [3][28][bin_cap] X pairs (a | b << 8), 0xFF = none
    __device__ uint16_t* xpairs() const { return reinterpret_cast<uint16_t*>(region<R_XP>()); }
    // tracklet index by slot order
    __device__ uint16_t* slot_order() const { return reinterpret_cast<uint16_t*>(region<R_SO>()); }
    // YZ candidates passing the first box cut, in reference order
    __device__ uint32_t* plist() const { return reinterpret_cast<uint32_t*>(region<R_PL>()); }
    // per candidate: x first-of-unit (bit 0), passed the sign-free tests (bit 1),
    // station 3 is D3m (bit 2); y station-3 signs after its left/right (first
    // of unit), z sign written to the last station-3 hit (0: none), w station-3
    // signs it starts from
    __device__ uchar4* pst() const { return reinterpret_cast<uchar4*>(region<R_PB>()); }
    // kept YZ units of a chunk (j << 16 | u3.v3 index), in order, and their first
    // candidate in the kept enumeration
    __device__ uint32_t* klist() const { return reinterpret_cast<uint32_t*>(region<R_KL>()); }
    __device__ uint32_t* kstart() const { return reinterpret_cast<uint32_t*>(region<R_KS>()); }
    // per box-cut survivor of a batch: its hits (raw LDS indices), which slots
    // hold one, and its first fit, as pass 2 built them (3 x 16 bytes)
    __device__ uint4* ycache() const { return reinterpret_cast<uint4*>(region<R_YC>()); }
    // survivors of a batch that pass the sign-free tests, in order
    __device__ uint32_t* qlist() const { return reinterpret_cast<uint32_t*>(region<R_QL>()); }
    // per tracklet of the chunk: the cache entry and sign state of its best candidate so far
    __device__ uint4* brec() const { return reinterpret_cast<uint4*>(region<R_BR>()); }
    // cleaning: per live tracklet, the raw hit of each of detectors 13-30 (uint16, 0xFFFF none)
    __device__ uint4* clean_hits() const { return reinterpret_cast<uint4*>(region<R_CL>()); }
    __device__ int win() const { return cfg().win_cap; }
};

// Learned unit pruning and unit rows for the YZ stage, by value from the
// launch arguments.
struct Prn {
    const PrunerDev* prune;   // NULL: no pruning
    float* dump;              // NULL: no unit rows
    int64_t dump_cap;
    unsigned long long* dump_count;
    int64_t ev;               // event index (unit rows)
};

// A + B: load, group by detector, reduce.  Returns true when the keep bits are
// valid (staging and reduction caps passed); station cuts set status bits.
template <int MAXH, int WPE>
STAGE bool stage_load_reduce(const Args& a, int64_t h0, int n) {
    h0 = uni(h0);
    n = uni(n);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    const or_params& P = a.p;
    bool staged = false;
    if (n > MAXH || n > P.max_hits_per_event) {
        if (lane == 0) L.status = OR_EV_EVENT_TOO_LARGE;
    } else {
        // ---- A: count per detector, validate ------------------------------
        for (int h = lane; h < n; h += WAVE) {
            const int d = a.det[h0 + h], el = a.elem[h0 + h];
            if (d > 54) continue;
            if (d < 1 || el < 1 || el > (int)G(F_NELEM, d) || el >= OR_MAX_ELEM) atomicOr(&L.status, (uint32_t)OR_EV_BAD_HIT);
            else atomicAdd(&L.pcnt[d], 1);
        }
        wsync();
        const bool badhit = L.status != 0;
        {
            const int d = lane;
            int c = (d >= 1 && d <= 54) ? L.pcnt[d] : 0;
            if (!badhit && d >= 1 && d <= 54) {
                const int cap = d <= 30 ? P.max_hits_chamber_plane : (d <= 46 ? P.max_hits_hodo_plane : P.max_hits_prop_plane);
                if (c > cap || c > PLANE_CAP) atomicOr(&L.status, (uint32_t)OR_EV_PLANE_OVERFLOW);
            }
            const int inc = wave_scan(c);
            if (lane < 56) { L.pstart[lane] = (int16_t)(inc - c); L.cursor[lane] = (int16_t)(inc - c); }
        }
        wsync();
        if (L.status == 0) {
            // stable scatter by detector
            for (int base = 0; base < n; base += WAVE) {
                const int h = base + lane;
                int d = 0, el = 0;
                bool v = h < n;
                if (v) { d = a.det[h0 + h]; el = a.elem[h0 + h]; v = d >= 1 && d <= 54; }
                unsigned long long todo = ballot(v);
                int dest = 0;
                while (todo) {
                    const int leader = __ffsll((long long)todo) - 1;
                    const int dl = __shfl(d, leader, WAVE);
                    const unsigned long long m = ballot(v && d == dl);
                    const int cur = L.cursor[dl];
                    if (v && d == dl) dest = cur + rank_in(m);
                    wsync();
                    if (lane == leader) L.cursor[dl] = (int16_t)(cur + __popcll(m));
                    wsync();
                    todo &= ~m;
                }
                if (v) {
                    L.elem[dest] = (int16_t)el;
                    L.pos[dest] = a.wire[d * OR_MAX_ELEM + el];
                    const float dr = a.drift[h0 + h];
                    L.drift[dest] = fabsf(a.rt ? rt_drift(a.rt->h[d], a.rt->x, a.rt->cf, d, a.tdc[h0 + h], dr) : dr);
                    L.orig[dest] = (uint16_t)h;
                    const int fl = a.flag ? (int)a.flag[h0 + h] : 2;
                    L.keep[dest] = (fl & 2) ? 1 : -1;
                }
            }
            for (int w = lane; w < MAXH / 32; w += WAVE) L.kbits[w] = 0;
            wsync();
            // ---- B: reduction, one lane per plane (reconstruction_helper.cuh:168-361)
            int nred = 0;
            bool clover = false;
            const int d = lane;
            if (d >= 1 && d <= 54) {
                const int s = L.pstart[d], np = L.pcnt[d];
                const bool reduce = d <= 6 || d >= 13;
                if (!reduce) {
                    for (int i = 0; i < np; ++i) L.keep[s + i] = 1;
                    nred = np;
                } else {
                    int ncl = 0, first = 0, lastm = 0;
                    for (int i = 0; i < np; ++i) {
                        const int r = s + i;
                        if (L.keep[r] < 0) continue;  // out of time
                        if (d > 30) continue;
                        if (ncl == 0) { first = i; lastm = i; ncl = 1; continue; }
                        const bool adj = fabsf((float)L.elem[r] - (float)L.elem[s + lastm]) <= 1.f;
                        if (adj && i != np - 1) { if (ncl >= 192) clover = true; ++ncl; lastm = i; continue; }
                        if (i == np - 1 && adj) { if (ncl >= 192) clover = true; ++ncl; lastm = i; }
                        if (ncl == 2) {
                            const int ra = s + first, rb = s + lastm;
                            const float wmax = (float)(0.9 * 0.5 * (double)(L.pos[rb] - L.pos[ra]));
                            const float wmin = (float)(4.0 / 9.0 * (double)wmax);
                            const float da = L.drift[ra], db = L.drift[rb];
                            if ((da > wmax && db > wmin) || (da > wmin && db > wmax)) {
                                if (da > db) L.keep[ra] = -1;
                                else L.keep[rb] = -1;
                            } else {
                                const float dt = a.tdc[h0 + L.orig[ra]] - a.tdc[h0 + L.orig[rb]];
                                if (((dt >= 0.0f && dt < 8.0f) || (dt <= 0.0f && dt > -8.0f)) && d >= 19 && d <= 24) {
                                    L.keep[ra] = -1;
                                    L.keep[rb] = -1;
                                }
                            }
                        }
                        if (ncl >= 3) {
                            float dtm = 0.0f;
                            int prev = -1;
                            for (int c = first; c <= lastm; ++c) {
                                if (L.keep[s + c] < 0) continue;  // not a member (out of time)
                                if (prev >= 0) {
                                    const float tc = a.tdc[h0 + L.orig[s + c]], tp = a.tdc[h0 + L.orig[s + prev]];
                                    const float dd = tc - tp;
                                    dtm += dd > 0.0f ? dd : (tp - tc);
                                }
                                prev = c;
                            }
                            dtm = dtm / (float)(ncl - 1);
                            const bool all = dtm < 10.0f;
                            for (int c = first; c <= lastm; ++c) {
                                if (L.keep[s + c] < 0) continue;
                                if (all || c != first) L.keep[s + c] = -1;
                            }
                        }
                        ncl = 0;
                    }
                    // first good hit per channel wins (reconstruction_helper.cuh:331-349)
                    for (int i = 0; i < np; ++i) {
                        const int r = s + i;
                        if (L.keep[r] <= 0) continue;
                        bool dup = false;
                        for (int j = 0; j < i; ++j)
                            if (L.keep[s + j] > 0 && L.elem[s + j] == L.elem[r]) { dup = true; break; }
                        if (dup) L.keep[r] = -1;
                        else ++nred;
                    }
                }
                if (reduce && nred > P.max_reduced_per_plane) atomicOr(&L.status, (uint32_t)OR_EV_PLANE_OVERFLOW);
                if (clover) atomicOr(&L.status, (uint32_t)OR_EV_CLUSTER_OVERFLOW);
                L.rcnt[d] = (int16_t)nred;
            } else if (lane < 56) {
                L.rcnt[lane] = 0;
            }
            {
                const int c = (d >= 1 && d <= 54) ? nred : 0;
                const int inc = wave_scan(c);
                if (lane < 56) L.rstart[lane] = (int16_t)(inc - c);
                // station multiplicities (reconstruction_kernels.cuh:127-139), 8 written rows summed
                const int m0 = wave_sum((d >= 1 && d <= 6) ? nred : 0);
                const int m2 = wave_sum((d >= 13 && d <= 18) ? nred : 0);
                const int m3 = wave_sum((d >= 19 && d <= 24) ? nred : 0);
                const int m4 = wave_sum((d >= 25 && d <= 30) ? nred : 0);
                wsync();
                if (lane == 0 && L.status == 0) {
                    uint32_t st = 0;
                    if (m0 > P.max_d0_mult) st |= OR_EV_D0_MULT;
                    if (m2 > P.max_d2_mult) st |= OR_EV_D2_MULT;
                    if (m3 > P.max_d3_mult) st |= OR_EV_D3P_MULT;
                    if (m4 > P.max_d3_mult) st |= OR_EV_D3M_MULT;
                    L.status = st;
                    staged = true;
                }
            }
            wsync();
            // compact the kept hits in place (planes stay in detector order, so the
            // k-th kept hit of plane d lands at rstart[d] + k); destinations never
            // pass their sources, so chunks taken in order read before any write
            // can reach them.  Keep bits by input index on the way.
            const int ntot = L.pstart[55];
            int base = 0;
            for (int h0s = 0; h0s < ntot; h0s += WAVE) {
                const int h = h0s + lane;
                const bool kept = h < ntot && L.keep[h] > 0;
                int16_t el = 0;
                float ps = 0.f, dr = 0.f;
                uint16_t og = 0;
                if (kept) { el = L.elem[h]; ps = L.pos[h]; dr = L.drift[h]; og = L.orig[h]; }
                const unsigned long long m = ballot(kept);
                wsync();
                if (kept) {
                    const int t = base + rank_in(m);
                    L.elem[t] = el; L.pos[t] = ps; L.drift[t] = dr; L.orig[t] = og;
                    atomicOr(&L.kbits[og >> 5], 1u << (og & 31));
                }
                base += __popcll(m);
                wsync();
            }
        }
    }
    return __shfl(staged ? 1 : 0, 0, WAVE) != 0;
}

template <int MAXH, int WPE>
STAGE void stage_keep_bits(const Args& a, int64_t h0, int n) {
    h0 = uni(h0);
    n = uni(n);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    const int64_t w0 = h0 >> 5, w1 = (h0 + n - 1) >> 5;
        // global word w holds input hits [32w, 32w + 32); event-local bit i is input hit h0 + i
        const int sh = (int)(h0 & 31);
        for (int64_t w = w0 + lane; n > 0 && w <= w1; w += WAVE) {
            const int64_t lo = w * 32 - h0;        // event-local index of bit 0 of word w
            const int k = (int)((lo + sh) >> 5);   // lo + sh = 32 (w - w0): bits b >= sh come from local word k
            uint32_t bits;
            if (sh == 0) {
                bits = L.kbits[k];
            } else {
                const uint32_t hi = k < MAXH / 32 ? L.kbits[k] : 0u;
                const uint32_t lw = k >= 1 ? L.kbits[k - 1] : 0u;
                bits = (lw >> (32 - sh)) | (hi << sh);
            }
            // only hits of this event: local indices [0, n)
            const int64_t a0 = lo < 0 ? -lo : 0, a1 = (lo + 32 > n) ? n - lo : 32;
            const uint32_t msk = (a1 >= 32 ? 0xFFFFFFFFu : ((1u << a1) - 1u)) & ~(a0 >= 32 ? 0xFFFFFFFFu : ((1u << a0) - 1u));
            bits &= msk;
            if (bits) atomicOr(&a.keep_mask[w], bits);
        }
    }

// C: binned X pairs for station 2, D3p, D3m
template <int MAXH, int WPE>
STAGE void stage_xpairs(or_params P, uint16_t* xpairs) {
    Lds<MAXH>& L = g_lds<MAXH>;
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    // hitflag1/2[100] bound (reconstruction_kernels.cuh:225-226) on all six planes first
        const int pc = (lane < 6) ? L.rcnt[lane < 2 ? 16 - lane : (lane < 4 ? 24 - lane : 32 - lane)] : 0;
        const bool over = ballot(lane < 6 && pc > P.max_pair_plane_hits) != 0ull;
        if (over) {
            if (lane == 0) L.status |= OR_EV_PAIR_OVERFLOW;
        } else if (lane < 3) {
            const int s = lane;
            const int d1 = s == 0 ? 16 : (s == 1 ? 22 : 28), d2 = d1 - 1;
            const int n1 = L.rcnt[d1], n2 = L.rcnt[d2];
            for (int b = 0; b < 28; ++b) L.xcnt[s][b] = 0;
            {
                const float sp = spacing_plane(super_id(2 + s, 0));
                const int cap = cfg().bin_cap;
                auto put = [&](int c, int ia, int ib) {
                    const int b = xbin(s, c);
                    if (b < 0) return;
                    const int k = L.xcnt[s][b];
                    if (k < cap) xpairs[(s * 28 + b) * cap + k] = (uint16_t)((ia & 0xFF) | ((ib & 0xFF) << 8));
                    L.xcnt[s][b] = k + 1;
                };
                auto close = [&](int i, int j) { return !(fabsf(L.pos[E.raw(d1, i)] - L.pos[E.raw(d2, j)]) > sp); };
                // hits that form a pair: bit masks for the first 64 of a plane, recomputed beyond
                unsigned long long f1 = 0, f2 = 0;
                for (int i = 0; i < n1; ++i) {
                    const int ri = E.raw(d1, i);
                    for (int j = 0; j < n2; ++j) {
                        if (!close(i, j)) continue;
                        put(L.elem[ri], i, j);
                        if (i < 64) f1 |= 1ull << i;
                        if (j < 64) f2 |= 1ull << j;
                    }
                }
                for (int i = 0; i < n1; ++i) {
                    bool paired = i < 64 ? ((f1 >> i) & 1ull) != 0 : false;
                    if (i >= 64)
                        for (int j = 0; j < n2 && !paired; ++j) paired = close(i, j);
                    if (!paired) put(L.elem[E.raw(d1, i)], i, 0xFF);
                }
                for (int j = 0; j < n2; ++j) {
                    bool paired = j < 64 ? ((f2 >> j) & 1ull) != 0 : false;
                    if (j >= 64)
                        for (int i = 0; i < n1 && !paired; ++i) paired = close(i, j);
                    if (!paired) put(L.elem[E.raw(d2, j)], 0xFF, j);
                }
                for (int b = 0; b < 28; ++b)
                    if (L.xcnt[s][b] > P.max_pairs_per_bin) atomicOr(&L.status, (uint32_t)OR_EV_XZ_BIN_OVERFLOW);
            }
        }
        wsync();
    }

// D: XZ combinations in canonical order, slot order.  Returns the tracklet count.
template <int MAXH, int WPE>
STAGE int stage_xz(or_params P, Scr S) {
    Lds<MAXH>& L = g_lds<MAXH>;
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    uint32_t* celloff = S.celloff();
    int ntkl = 0;
    uint32_t run = 0;
        for (int c0 = 0; c0 < 1568; c0 += WAVE) {
            const int c = c0 + lane;
            int cnt = 0;
            if (c < 1568) {
                const int t = c / 49, i = c % 49;
                const int s3 = 1 + (t & 1);
                const int b2 = (t / 8) * 7 + i % 7, b3 = ((t % 8) / 2) * 7 + i / 7;
                cnt = L.xcnt[0][b2] * L.xcnt[s3][b3];
            }
            const int inc = wave_scan(cnt);
            if (c < 1568) celloff[c + 1] = run + (uint32_t)inc;
            if (c == 0) celloff[0] = 0;
            run += (uint32_t)__shfl(inc, WAVE - 1, WAVE);
        }
        wsync();
        const uint32_t total = run;
        int np = 0, nm = 0;
        int cb = 0;   // the cell of combination g0 (the same for every lane)
        for (uint32_t g0 = 0; g0 < total; g0 += WAVE) {
            const uint32_t g = g0 + lane;
            bool acc = false;
            int par = 0, t = 0, s3 = 0;
            uint16_t q2 = 0, q3 = 0;
            // lane t holds the end of cell cb + t; the cells ending at or before a
            // combination are a prefix of the lanes (cell offsets never decrease)
            const uint32_t ce = cb + 1 + lane <= 1568 ? celloff[cb + 1 + lane] : 0xFFFFFFFFu;
            const bool inwin = (uint32_t)__builtin_amdgcn_readlane((int)ce, WAVE - 1) > g0 + WAVE - 1;
            int tu = 0;
            if (inwin) {
#pragma unroll
                for (int step = WAVE / 2; step > 0; step >>= 1)
                    if ((uint32_t)__shfl((int)ce, tu + step - 1, WAVE) <= g) tu += step;
            }
            if (g < total) {
                int c = cb + tu;
                if (!inwin) {
                    int lo = 0, hi = 1568;  // largest c with celloff[c] <= g
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (celloff[mid] <= g) lo = mid;
                        else hi = mid;
                    }
                    c = lo;
                }
                t = c / 49;
                const int i = c % 49;
                s3 = 1 + (t & 1);
                const int b2 = (t / 8) * 7 + i % 7, b3 = ((t % 8) / 2) * 7 + i / 7;
                const int nx2 = L.xcnt[0][b2];
                const int ix = (int)(g - celloff[c]);
                const int ix2 = ix % nx2, ix3 = ix / nx2;
                par = t & 1;
                q2 = S.xpairs()[b2 * cfg().bin_cap + ix2];
                q3 = S.xpairs()[(s3 * 28 + b3) * cfg().bin_cap + ix3];
                acc = xz_candidate(E, s3, q2, q3, nullptr);
            }
            const unsigned long long m = ballot(acc);
            const unsigned long long mp = ballot(acc && par == 0);
            const unsigned long long mm = ballot(acc && par == 1);
            if (acc) {
                // accepted: evaluate again writing straight into the tracklet slot
                const int k = ntkl + rank_in(m);
                const int kp = par == 0 ? np + rank_in(mp) : nm + rank_in(mm);
                if (k < cfg().maxtkl) {
                    Tk& T = tks[k];
                    xz_candidate(E, s3, q2, q3, &T);
                    T.t = (uint8_t)t;
                    T.st3 = (uint8_t)(3 + par);
                    T.canon = (uint32_t)k;
                    T.slot = (uint8_t)(2 * (kp % 16) + par);
                    T.pos_in_slot = (uint16_t)(kp / 16);
                }
            }
            ntkl += __popcll(m);
            np += __popcll(mp);
            nm += __popcll(mm);
            // the cell of the next wave's first combination
            const unsigned long long adv = ballot(ce <= g0 + WAVE);
            if (adv != ~0ull) {
                cb += __popcll(adv);
            } else {
                int lo = cb, hi = 1568;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (celloff[mid] <= g0 + WAVE) lo = mid;
                    else hi = mid;
                }
                cb = lo;
            }
        }
        if (lane == 0) {
            if (np > 16 * P.max_tracklets_per_slot || nm > 16 * P.max_tracklets_per_slot ||
                ntkl >= P.max_tracklets_per_event || ntkl > cfg().maxtkl)
                atomicOr(&L.status, (uint32_t)OR_EV_TRACKLET_OVERFLOW);
        }
        // slot order: slot s = 2q + p holds items kp = q, q+16, ...
        {
            const int p = lane & 1, q = lane >> 1;
            const int npar = p ? nm : np;
            const int cs = (lane < 32 && npar > q) ? (npar - q + 15) / 16 : 0;
            const int inc = wave_scan(cs);
            if (lane < 32) L.sstart[lane] = inc - cs;
            if (lane == 32) L.sstart[32] = ntkl < cfg().maxtkl ? ntkl : cfg().maxtkl;
        }
        wsync();
        if (L.status == 0)
            for (int k = lane; k < ntkl; k += WAVE) S.slot_order()[L.sstart[tks[k].slot] + tks[k].pos_in_slot] = (uint16_t)k;
        wsync();
        return ntkl;
}

// E: YZ extension (reconstruction_kernels.cuh:864-1832), chunks of YZ_CHUNK
// tracklets.
//
// Reference order per tracklet: (u3, v3) pairs outer with u3 fastest, (u2, v2)
// pairs inner with u2 fastest; the minimum chi2 with strict < wins.  Within one
// (u3, v3) unit the station-3 hit signs persist from candidate to candidate:
// the first candidate that passes the first box cut sets all of them through
// left/right; afterwards only the last station-3 hit can change, when the
// pairing walk pairs it with the first station-2 hit (D3p' 19 with D2 U 18)
// and that pair's left/right decision is taken.  Every other input of a
// candidate (its fit, which uses y at sign 0, and the pair decisions) depends
// on the candidate alone.  So the sign state each candidate starts from is a
// last-writer scan over the unit's box-cut survivors, and all candidates are
// evaluated in parallel (yz_chunk_par).  yz_chunk_seq walks units one lane
// each, in reference order, for chunks with more survivors than the list holds.

// Sorted prop-tube Y positions of the event (rank sort, ties in plane order),
// for the YZ stage's tests; one wave.
template <int MAXH, int WPE>
STAGE void stage_prop_sort() {
    Lds<MAXH>& L = g_lds<MAXH>;
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    for (int p = 0; p < 4; ++p) {
        const int d = psort_plane(p), n = E.nred(d), r0 = E.raw(d, 0);
        if (n > psort_cap(MAXH)) {
            if (lane == 0) L.pscnt[p] = -1;
            continue;
        }
        for (int i0 = 0; i0 < n; i0 += WAVE) {
            const int i = i0 + lane;
            if (i < n) {
                const float v = L.pos[r0 + i];
                int rank = 0;
                for (int k = 0; k < n; ++k) {
                    const float u = L.pos[r0 + k];
                    rank += (u < v || (u == v && k < i)) ? 1 : 0;
                }
                L.psort[p][rank] = v;
            }
        }
        if (lane == 0) L.pscnt[p] = n;
    }
    wsync();
}

template <int MAXH, int WPE>
__device__ __forceinline__ void yz_lists_of(const Scr& S, int j, YzLists& W) {
    for (int v = 0; v < 4; ++v) {
        W.l[v] = S.wl() + (j * 4 + v) * cfg().win_cap;
        W.y[v] = S.wy() + (j * 4 + v) * cfg().win_cap;
        W.f[v] = S.wf() + (j * 4 + v) * cfg().win_cap;
        W.n[v] = yzw<MAXH>().nl[j][v];
    }
}

struct YzItem { int j, local, iu3, iv3, iu2, iv2; };

// candidate index within the chunk -> tracklet and list entries
template <int MAXH, int WPE>
__device__ __forceinline__ YzItem yz_decode(int i) {
    const YzWave& Y = yzw<MAXH>();
    YzItem it;
    // the last tracklet whose candidates start at or before i (empty ones start
    // where the next does, so the one holding i is the last)
    int j = 0;
#pragma unroll
    for (int st = YZ_CHUNK / 2; st > 0; st >>= 1)
        if (Y.uoff[j + st] <= i) j += st;
    it.j = j;
    it.local = i - Y.uoff[j];
    const int n0 = Y.nl[j][0], nin = n0 * Y.nl[j][1];
    const int iuv3 = qdiv(it.local, Y.rin[j]), iuv2 = it.local - iuv3 * nin;
    it.iv2 = qdiv(iuv2, Y.rx[j]); it.iu2 = iuv2 - it.iv2 * n0;
    it.iv3 = qdiv(iuv3, Y.rz[j]); it.iu3 = iuv3 - it.iv3 * Y.nl[j][2];
    return it;
}

// first fit of a candidate over u3, u3', v3, v3', u2, u2', v2, v2' (the
// reference's float order); false when a view has no accepted hit
// YZ first-fit table.  A hit's weight and z are constants of its plane, so the
// normal matrix (A0, A1, A3) of a YZ candidate, its inverse and the parameter
// errors depend only on which hits of its four list entries are present (per
// entry: a, b or both) and on station 3 being D3p or D3m: 2 x 81 patterns.
// Entry (m, pattern): Ai[0..3], then e0, e1, computed by the fit's own
// arithmetic (sums in the order u3, v3, u2, v2, hit a before b; matinv2;
// square roots) once per device when the geometry is set; a candidate then
// needs only its B sums.  pattern = 27 (fu3 - 1) + 9 (fv3 - 1) + 3 (fu2 - 1) + fv2 - 1.
__device__ float4 g_yz_atab[2 * 81 * 2];

// The pattern index from list flag bytes, which carry each entry's term in bits 2-7
__device__ __forceinline__ int yz_pattern_b(unsigned bu2, unsigned bv2, unsigned bu3, unsigned bv3) {
    return (int)((bu2 >> 2) + (bv2 >> 2) + (bu3 >> 2) + (bv3 >> 2));
}

__global__ __launch_bounds__(256) void yz_atab_kernel() {
    const int t = threadIdx.x;
    if (t >= 2 * 81) return;
    const int m = t / 81, pt = t % 81;
    const unsigned fl[4] = {(unsigned)(pt / 3 % 3 + 1), (unsigned)(pt % 3 + 1), (unsigned)(pt / 27 + 1),
                            (unsigned)(pt / 9 % 3 + 1)};   // u2, v2, u3, v3
    float A0 = 0.f, A1 = 0.f, A3 = 0.f;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int v = o < 2 ? o + 2 : o - 2;
        const int sid = v < 2 ? super_id(2, v + 1) : super_id(3 + m, v - 1);
        const int da = sid * 2, db = sid * 2 - 1;
        if (fl[v] & 1) {
            const float w = yz_w(da), z = G(F_Z, da);
            A0 += w; A1 += w * z; A3 += (w * z) * z;
        }
        if (fl[v] & 2) {
            const float w = yz_w(db), z = G(F_Z, db);
            A0 += w; A1 += w * z; A3 += (w * z) * z;
        }
    }
    float Ai[4];
    matinv2(A0, A1, A1, A3, Ai);
    g_yz_atab[2 * t] = make_float4(Ai[0], Ai[1], Ai[2], Ai[3]);
    g_yz_atab[2 * t + 1] = make_float4(sqrtf(fabsf(Ai[0])), sqrtf(fabsf(Ai[3])), 0.f, 0.f);
}

// fit_finish's parameters from the table entry and the B sums
__device__ __forceinline__ Fit yz_fit_tab(int m, int pattern, float B0, float B1) {
    const float4 ai = g_yz_atab[2 * (m * 81 + pattern)], ee = g_yz_atab[2 * (m * 81 + pattern) + 1];
    Fit f;
    f.e0 = ee.x;
    f.e1 = ee.y;
    float q = 0.0f;
    q += ai.x * B0;
    q += ai.y * B1;
    f.p0 = q;
    q = 0.0f;
    q += ai.z * B0;
    q += ai.w * B1;
    f.p1 = q;
    return f;
}

// B sums of an entry's present hits, a then b: w y, and (w y) z with z the
// plane's (za for hit a, zb for hit b), the products the fit sums.  An entry
// holds 0 for a hit that is absent or not accepted, and its terms are added
// too: x + (+-0) = x for every x but -0, and the sums never are -0 (they start
// at +0, and a round-to-nearest sum is -0 only when both terms are), so the
// sums are those of the present hits alone, bit for bit.
__device__ __forceinline__ void yz_bsum(const float4& t, float za, float zb, float& B0, float& B1) {
    B0 += t.x; B1 += t.x * za;
    B0 += t.y; B1 += t.y * zb;
}

// z of the planes of YZ view v (u2, v2, u3, v3) with station 3 st3: hit a, hit b
__device__ __forceinline__ float2 yz_zab(int v, int st3) {
    const int sid = v < 2 ? super_id(2, v + 1) : super_id(st3, v - 1);
    return make_float2(G(F_Z, sid * 2), G(F_Z, sid * 2 - 1));
}

__device__ __forceinline__ bool yz_fit(const YzLists& W, int st3, const YzItem& it, Fit& f) {
    const float4 tu3 = gld4(W.y[2] + it.iu3), tv3 = gld4(W.y[3] + it.iv3);
    const float4 tu2 = gld4(W.y[0] + it.iu2), tv2 = gld4(W.y[1] + it.iv2);
    const unsigned fu3 = __float_as_uint(tu3.z), fv3 = __float_as_uint(tv3.z);
    const unsigned fu2 = __float_as_uint(tu2.z), fv2 = __float_as_uint(tv2.z);
    if (!fu3 || !fv3 || !fu2 || !fv2) return false;
    float B0 = 0.f, B1 = 0.f;
    const float2 z0 = yz_zab(0, st3), z1 = yz_zab(1, st3), z2 = yz_zab(2, st3), z3 = yz_zab(3, st3);
    yz_bsum(tu3, z2.x, z2.y, B0, B1);
    yz_bsum(tv3, z3.x, z3.y, B0, B1);
    yz_bsum(tu2, z0.x, z0.y, B0, B1);
    yz_bsum(tv2, z1.x, z1.y, B0, B1);
    f = yz_fit_tab(st3 == 4 ? 1 : 0, yz_pattern_b(fu2, fv2, fu3, fv3), B0, B1);
    return true;
}

__device__ __forceinline__ bool yz_box1(const Fit& f) {
    return !(fabsf(f.p0) > Y0_MAX + 2.f * f.e0 || fabsf(f.p1) > TY_MAX + 2.f * f.e1);
}

// station-3 signs (slots 0-3) packed two bits each, 3 = -1
__device__ __forceinline__ uint8_t yz_pack3(const YzSlots& H) {
    unsigned c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) c |= (unsigned)(H.sign[q] & 3) << (2 * q);
    return (uint8_t)c;
}
__device__ __forceinline__ void yz_unpack3(YzSlots& H, unsigned c) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int v = (c >> (2 * q)) & 3;
        H.sign[q] = v == 3 ? -1 : v;
    }
}

__device__ __forceinline__ int wave_max_scan(int v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(v, o, WAVE);
        if (l >= o) v = max(v, t);
    }
    return v;
}

// candidate slots and first fit of candidate i of the chunk
template <int MAXH, int WPE>
__device__ __forceinline__ void yz_candidate(const Ev<MAXH, WPE>& E, const Scr& S, const Tk* Tp, const YzItem& it,
                                             YzSlots& H, Fit& f) {
    YzLists W;
    yz_lists_of<MAXH, WPE>(S, it.j, W);
    yz_fit(W, Tp->st3, it, f);
    yz_dets(H, Tp->st3);
    H.pres = 0;
    yz_fill<0>(E, W, 2, it.iu3, H);
    yz_fill<2>(E, W, 3, it.iv3, H);
    yz_fill<4>(E, W, 0, it.iu2, H);
    yz_fill<6>(E, W, 1, it.iv2, H);
}

template <int ST3, int MAXH, int WPE>
__device__ __forceinline__ float yz_chi2_st(const Ev<MAXH, WPE>& E, const Tk* Tp, int j, const YzSlots& H,
                                            const Fit& f);

// yz_candidate for a tracklet whose station 3 is ST3 (plane constants of the
// candidate's slots known at compile time)
template <int ST3, int MAXH, int WPE>
__device__ __forceinline__ void yz_candidate_st(const Ev<MAXH, WPE>& E, const Scr& S, const YzItem& it, YzSlots& H,
                                                Fit& f) {
    YzLists W;
    yz_lists_of<MAXH, WPE>(S, it.j, W);
    yz_fit(W, ST3, it, f);
    yz_dets(H, ST3);
    H.pres = 0;
    yz_fill<0>(E, W, 2, it.iu3, H);
    yz_fill<2>(E, W, 3, it.iv3, H);
    yz_fill<4>(E, W, 0, it.iu2, H);
    yz_fill<6>(E, W, 1, it.iv2, H);
}

// The representative candidate of a unit (see unit_features), evaluated as the
// chain evaluates a candidate (yz_candidate, first box cut, left/right, the
// sign-free tests, chi2) for a tracklet whose station 3 is ST3: the candidate's
// planes are then compile-time constants, so their constants come through
// scalar loads.  rep_pass: 0 fails the first box cut, 1 a sign-free test, 2 passes.
template <int ST3, int MAXH, int WPE>
__device__ __forceinline__ void yz_rep_eval(const Ev<MAXH, WPE>& E, const Scr& S, const Tk& T, const YzItem& it,
                                            float& chi, float& rep_y0, int& rep_pass) {
    YzSlots H;
    Fit fr;
    yz_candidate_st<ST3>(E, S, it, H, fr);   // the unit's entries and the nearest ones all hold an accepted hit
    rep_y0 = fr.p0;
    if (yz_box1(fr)) {
        rep_pass = 1;
        lr_slots_new<8>(T.x0, T.tx, fr.p0, fr.p1, T.ex0, T.etx, fr.e0, fr.e1, H.pres, H.det, H.pos, H.drift, H.sign,
                        150.f);
        if (yz_cuts_hx(E, yzw<MAXH>().hx[it.j], fr)) rep_pass = 2;
        chi = yz_chi2_st<ST3>(E, &T, it.j, H, fr);
    }
}

// Features of YZ unit (tracklet j, u3.v3 index iuv3) for the pruner.  The
// window lists hold accepted entries only; a view's y is the mean of its
// entry's accepted hits.  With y3 the mean y of the unit's station-3 hits at
// the mean z3 of its U and V planes, the line through (z3, y3) and the beam
// axis at z = 0 predicts the y of any later hit; the box cut |y0| <= 50 cm at
// z = 0 keeps every candidate's line near it.  Features: 10 tx, y3/50, the
// signed and absolute U - V difference at station 3, the a - b difference
// within the U3 and the V3 entry, U3 and V3 hit counts - 1, per station-2
// view the distance of the nearest list entry to the predicted y, then for the
// representative candidate (the unit's station-3 entries with the nearest U2
// and V2 entries, evaluated as every candidate is) log(1 + chi2)/4 and how far
// it gets (0: fails the first box cut, 1: fails a sign-free test, 2: passes),
// the nearest U2 minus nearest V2 y, the distance of the nearest P1 Y and P2 Y
// prop-tube hit to the predicted y, and the candidate's |y0|/50 (distances in
// units of 10 cm; every feature clamped at 10).
template <int MAXH, int WPE>
__device__ __forceinline__ UnitFeatures unit_features(const Scr& S, const Tk& T, int j, int iuv3, const int4 nl) {
    const Ev<MAXH, WPE> E;
    const int win = S.win();
    const int iv3 = qdiv(iuv3, yzw<MAXH>().rz[j]), iu3 = iuv3 - iv3 * nl.z;
    const int qu = (j * 4 + 2) * win + iu3, qv = (j * 4 + 3) * win + iv3;
    const float2 yu = gld2(S.wyy() + qu), yw = gld2(S.wyy() + qv);
    const unsigned fu = gld(S.wf() + qu) & 3u, fv = gld(S.wf() + qv) & 3u;
    const float ya = (fu & 1u) ? yu.x : yu.y, yb = (fu & 2u) ? yu.y : yu.x;
    const float yc = (fv & 1u) ? yw.x : yw.y, yd = (fv & 2u) ? yw.y : yw.x;
    const float yum = 0.5f * (ya + yb), yvm = 0.5f * (yc + yd);
    const float y3 = 0.5f * (yum + yvm);
    const int st3 = T.st3;
    const float z3 = 0.5f * (G(F_Z, super_id(st3, 1) * 2) + G(F_Z, super_id(st3, 2) * 2));
    const float k3 = y3 / z3;   // predicted y per cm of z
    // per station-2 view the list entry nearest to the predicted y (the first in
    // list order at the least distance, if below 100); both lists are read in
    // blocks of 16 entries, one round of loads per block (the lists' region is
    // padded, and entries past a list's end are not compared)
    float d2[2], yb2[2];
    int k2[2];
    {
        const float yp[2] = {k3 * G(F_Z, super_id(2, 1) * 2), k3 * G(F_Z, super_id(2, 2) * 2)};
        const int n[2] = {nl.x, nl.y};
        const float* src[2] = {reinterpret_cast<const float*>(S.wyy() + (j * 4) * win),
                               reinterpret_cast<const float*>(S.wyy() + (j * 4 + 1) * win)};
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            d2[v] = 100.f;
            yb2[v] = yp[v];
            k2[v] = 0;
        }
        const int nmax = max(n[0], n[1]);
        for (int k0 = 0; k0 < nmax; k0 += 16) {
            float ye[2][16];
#pragma unroll
            for (int v = 0; v < 2; ++v)
#pragma unroll
                for (int t = 0; t < 16; ++t) ye[v][t] = gld(src[v] + 2 * (k0 + t));   // .x of the entry's float2
#pragma unroll
            for (int v = 0; v < 2; ++v)
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const float d = fabsf(ye[v][t] - yp[v]);
                    if (k0 + t < n[v] && d < d2[v]) {
                        d2[v] = d;
                        yb2[v] = ye[v][t];
                        k2[v] = k0 + t;
                    }
                }
        }
    }
    // representative candidate: the unit's station-3 entries with the nearest U2
    // and V2 entries, evaluated as the chain evaluates a candidate: first fit,
    // first box cut, left/right, the sign-free tests (second box cut, prop-tube Y,
    // hodoscopes) and chi2
    float chi = 1.0e6f, rep_y0 = 0.f;
    int rep_pass = 0;
    if (nl.x > 0 && nl.y > 0) {
        YzItem it;
        it.j = j; it.local = 0; it.iu3 = iu3; it.iv3 = iv3; it.iu2 = k2[0]; it.iv2 = k2[1];
        if (st3 == 3) yz_rep_eval<3>(E, S, T, it, chi, rep_y0, rep_pass);
        else yz_rep_eval<4>(E, S, T, it, chi, rep_y0, rep_pass);
    }
    float dp[2];
    const Lds<MAXH>& L = g_lds<MAXH>;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int da = psort_plane(2 * p), db = psort_plane(2 * p + 1);   // 48, 47; 54, 53
        const float ypa = k3 * G(F_Z, da), ypb = k3 * G(F_Z, db);
        float dmin = 100.f;
        if (L.pscnt[2 * p] >= 0) {
            dmin = prop_sorted_nearest(L.psort[2 * p], L.pscnt[2 * p], ypa, dmin);
        } else {
#pragma unroll 4
            for (int i = 0; i < E.nred(da); ++i) dmin = fminf(dmin, fabsf(E.Lp->pos[E.raw(da, i)] - ypa));
        }
        if (L.pscnt[2 * p + 1] >= 0) {
            dmin = prop_sorted_nearest(L.psort[2 * p + 1], L.pscnt[2 * p + 1], ypb, dmin);
        } else {
#pragma unroll 4
            for (int i = 0; i < E.nred(db); ++i) dmin = fminf(dmin, fabsf(E.Lp->pos[E.raw(db, i)] - ypb));
        }
        dp[p] = dmin;
    }
    UnitFeatures f;
    f.v[0] = T.tx * 10.f;
    f.v[1] = y3 * 0.02f;
    f.v[2] = (yum - yvm) * 0.1f;
    f.v[3] = fabsf(yum - yvm) * 0.1f;
    f.v[4] = (ya - yb) * 0.1f;
    f.v[5] = (yc - yd) * 0.1f;
    f.v[6] = (float)(__popc(fu) - 1);
    f.v[7] = (float)(__popc(fv) - 1);
    f.v[8] = fminf(d2[0] * 0.1f, 10.f);
    f.v[9] = fminf(d2[1] * 0.1f, 10.f);
    f.v[10] = fminf(log1pf(fmaxf(chi, 0.f)) * 0.25f, 10.f);
    f.v[11] = (float)rep_pass;
    f.v[12] = fminf(fmaxf((yb2[0] - yb2[1]) * 0.1f, -10.f), 10.f);
    f.v[13] = fminf(dp[0] * 0.1f, 10.f);
    f.v[14] = fminf(dp[1] * 0.1f, 10.f);
    f.v[15] = fminf(fabsf(rep_y0) * 0.02f, 10.f);
    return f;
}

// first fit and box cut of YZ candidate (tracklet j of the chunk, list entries
// e = u2, v2, u3, v3; m: station 3 is D3m): B sums in the reference order u3,
// v3, u2, v2, a then b, the rest from the first-fit table
// the chunk's tracklet holding (u3, v3) unit u of the scoring enumeration
template <int MAXH>
__device__ __forceinline__ int unit_tracklet(int u) {
    const YzWave& Y = yzw<MAXH>();
    int j = 0;
#pragma unroll
    for (int t = 1; t < YZ_CHUNK; ++t) j += u >= Y.unoff[t] ? 1 : 0;
    return j;
}

// features of unit u of the chunk starting at tracklet kb (zeros when u >= U)
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) UnitFeatures unit_features_at(Scr S, int kb, int u, int U) {
    kb = uni(kb);
    U = uni(U);
    UnitFeatures f;
#pragma unroll
    for (int q = 0; q < PRUNE_NF; ++q) f.v[q] = 0.f;
    if (u < U) {
        const YzWave& Y = yzw<MAXH>();
        const int j = unit_tracklet<MAXH>(u);
        f = unit_features<MAXH, WPE>(S, S.tks()[kb + j], j, u - Y.unoff[j], *reinterpret_cast<const int4*>(Y.nl[j]));
    }
    return f;
}

// plane z of the YZ views: [u2, v2, u3, v3][station 3 is D3m][hit a, b]
__device__ __forceinline__ void yz_plane_z(float (&zv)[4][2][2]) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const float2 z = yz_zab(v, 3 + m);
            zv[v][m][0] = z.x;
            zv[v][m][1] = z.y;
        }
}

// Pass 1 splits it in two: the loads of a candidate (YzFirst) are issued one
// wave iteration ahead of its evaluation.
struct YzFirst {
    float4 t[4];         // the four list entries (flag bytes in .z)
    unsigned m;
    int ustart, index;   // first candidate of its unit, its own index (-1: no candidate)
};
__device__ __forceinline__ void yz_first_load(const Scr& S, int j, const int (&e)[4], YzFirst& c) {
#pragma unroll
    for (int v = 0; v < 4; ++v) c.t[v] = gld4(S.wy() + (j * 4 + v) * cfg().win_cap + e[v]);
}
// zv: plane z of the views by station 3 (yz_plane_z)
__device__ __forceinline__ bool yz_first_eval(const YzFirst& c, const float (&zv)[4][2][2]) {
    const unsigned fl[4] = {__float_as_uint(c.t[0].z), __float_as_uint(c.t[1].z), __float_as_uint(c.t[2].z),
                            __float_as_uint(c.t[3].z)};
    if (c.index < 0 || !fl[0] || !fl[1] || !fl[2] || !fl[3]) return false;
    const int m = (int)c.m;
    float B0 = 0.f, B1 = 0.f;
    yz_bsum(c.t[2], m ? zv[2][1][0] : zv[2][0][0], m ? zv[2][1][1] : zv[2][0][1], B0, B1);
    yz_bsum(c.t[3], m ? zv[3][1][0] : zv[3][0][0], m ? zv[3][1][1] : zv[3][0][1], B0, B1);
    yz_bsum(c.t[0], zv[0][0][0], zv[0][0][1], B0, B1);
    yz_bsum(c.t[1], zv[1][0][0], zv[1][0][1], B0, B1);
    return yz_box1(yz_fit_tab(m, yz_pattern_b(fl[0], fl[1], fl[2], fl[3]), B0, B1));
}

// Pass 1 streams in batches.  The survivors of the first box cut go into a
// list of cap entries; when the list fills, the batch ends before the first
// (u3, v3) unit that does not fit whole, passes 2 and 3 run on it, and the next
// batch starts at that unit.  The sign state of the YZ passes never crosses a
// unit, so batches cut at unit boundaries give every candidate the state it
// has in one pass over all of them.  A unit with more survivors than the list
// holds sends the chunk to the unit-by-unit path.
//
// Learned pruning and unit rows: every (u3, v3) unit of the chunk is scored
// (and its row written) first, the kept units are listed in order with the
// start of their candidates in the kept enumeration, and pass 1 walks the
// candidates of the kept units only.  Survivors keep their index in the full
// enumeration, so the later passes see the plain chain's candidates with the
// pruned units taken out.
struct UnitScore {
    int nk;          // kept units; -1: more units than the list holds (plain pass 1 runs)
    uint32_t run;    // candidates of the kept units
    int64_t dbase;   // first unit row of the chunk, -1: none written
};

template <int MAXH, int WPE>
__device__ __attribute__((noinline)) UnitScore yz_score_units(Scr S, Prn R, int kb, int ntkl) {
    kb = uni(kb);
    ntkl = uni(ntkl);
    R.prune = uni(R.prune);
    R.dump = uni(R.dump);
    R.dump_cap = uni(R.dump_cap);
    R.dump_count = uni(R.dump_count);
    R.ev = uni(R.ev);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    const int lane = lane_id();
    Tk* tks = S.tks();
    UnitScore res = {-1, 0u, -1};
    OR_SUB_START
    int uo[YZ_CHUNK + 1];
    {
        int units = 0;
        if (lane < YZ_CHUNK && kb + lane < ntkl) {
            const int* nl = Y.nl[lane];
            if (nl[0] > 0 && nl[1] > 0 && nl[2] > 0 && nl[3] > 0) units = nl[2] * nl[3];
        }
        const int inc = wave_scan(units);
        uo[0] = 0;
#pragma unroll
        for (int t = 1; t <= YZ_CHUNK; ++t) uo[t] = __builtin_amdgcn_readfirstlane(__shfl(inc, t - 1, WAVE));
        if (lane < YZ_CHUNK) Y.unoff[lane + 1] = inc;
        if (lane == 0) Y.unoff[0] = 0;
    }
    const int U = uo[YZ_CHUNK];
    if (U > cfg().unit_cap) return res;
    if (R.dump) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(R.dump_count, (unsigned long long)U);
        const unsigned long long b0 = (unsigned)__shfl((int)(uint32_t)b, 0, WAVE);
        const unsigned long long b1 = (unsigned)__shfl((int)(uint32_t)(b >> 32), 0, WAVE);
        const unsigned long long bb = b0 | (b1 << 32);
        if (bb + (unsigned long long)U <= (unsigned long long)R.dump_cap) res.dbase = (int64_t)bb;
    }
    // score every unit (scores held in the candidate-cache region, which pass 2
    // fills only later), each tracklet's best score in LDS
    const bool rel = R.prune && R.prune->margin < INFINITY;
    float* sc = reinterpret_cast<float*>(S.ycache());
    if (lane < YZ_CHUNK) Y.umax[lane] = INT_MIN;
    wsync();
    for (int u0 = 0; u0 < U; u0 += WAVE) {
        const int u = u0 + lane;
        const bool valid = u < U;
        const UnitFeatures f = unit_features_at<MAXH, WPE>(S, kb, u, U);
        if (R.prune) {
            const float score = prune_score(R.prune, f);
            if (valid) {
                sc[u] = score;
                if (rel) atomicMax(&Y.umax[unit_tracklet<MAXH>(u)], float_order(score));
            }
        }
        if (res.dbase >= 0 && valid) {
            const int j = unit_tracklet<MAXH>(u), iuv3 = u - Y.unoff[j];
            float* row = R.dump + (res.dbase + u) * OR_UNIT_ROW;
            row[0] = (float)R.ev;
            row[1] = (float)tks[kb + j].canon;
            row[2] = (float)iuv3;
            row[3] = 0.f;
#pragma unroll
            for (int q = 0; q < PRUNE_NF; ++q) row[4 + q] = f.v[q];
        }
    }
    wsync();
    // keep a unit when it reaches the threshold and lies within the margin of its
    // tracklet's best; the kept ones, in order, with the start of their candidates
    // in the kept enumeration
    int nk = 0;
    uint32_t run = 0;
    for (int u0 = 0; u0 < U; u0 += WAVE) {
        const int u = u0 + lane;
        const bool valid = u < U;
        int j = 0;
#pragma unroll
        for (int t = 1; t < YZ_CHUNK; ++t) j += u >= uo[t] ? 1 : 0;
        int ujs = uo[0];
#pragma unroll
        for (int t = 1; t < YZ_CHUNK; ++t) ujs = j == t ? uo[t] : ujs;
        const int iuv3 = u - ujs;
        int nin = 0;
        bool keep = valid;
        if (valid) {
            const int4 nl = *reinterpret_cast<const int4*>(Y.nl[j]);
            nin = nl.x * nl.y;
            if (R.prune) {
                const float score = sc[u];
                keep = score >= R.prune->thr && (!rel || score >= float_unorder(Y.umax[j]) - R.prune->margin);
            }
        }
        const unsigned long long km = ballot(keep);
        const int c = keep ? nin : 0;
        const int inc = wave_scan(c);
        if (keep) {
            const int r = nk + rank_in(km);
            S.klist()[r] = ((uint32_t)j << 16) | (uint32_t)iuv3;
            S.kstart()[r] = run + (uint32_t)(inc - c);
            if (res.dbase >= 0) R.dump[(res.dbase + u) * OR_UNIT_ROW + 3] = 0.5f;   // searched
        }
        nk += __popcll(km);
        run += (uint32_t)__shfl(inc, WAVE - 1, WAVE);
    }
    wsync();
    OR_SUB(9);
    OR_COUNT(11, U);
    OR_COUNT(12, nk);
    OR_COUNT(13, run);
    res.nk = nk;
    res.run = run;
    return res;
}

// number of entries of the sorted list p[0, n) below v (every lane alike)
__device__ __forceinline__ int lower_bound_u32(const uint32_t* p, int n, uint32_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gld(p + mid) < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One batch of pass 1 over the kept units' candidates from kept-enumeration
// index cursor (the start of kept unit lo0).  Returns the survivors listed; on
// return cursor and lo0 name the next batch's first unit (cursor == run: done).
// -1: the unit at cursor alone has more survivors than the list holds.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) int yz_pass1_kept(Scr S, int kb, int ntkl, int cap, int nk, uint32_t run,
                                                       uint32_t& cursor, int& lo_unit) {
    kb = uni(kb);
    ntkl = uni(ntkl);
    cap = uni(cap);
    nk = uni(nk);
    run = uni(run);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    const int lane = lane_id();
    Tk* tks = S.tks();
    const unsigned d3m = (unsigned)ballot(lane < YZ_CHUNK && kb + lane < ntkl && tks[kb + lane].st3 == 4);
    float zv[4][2][2];
    yz_plane_z(zv);
    const uint32_t start = uni(cursor);
    int np = 0;
    int lo0 = uni(lo_unit);   // last kept unit starting at or before i0 (the same for every lane)
    // The kept units' starts and codes as two windows of 64 in registers: lane t
    // of (ast, akl) holds unit wlo + t, of (bst, bkl) unit wlo + 64 + t (start
    // 0xFFFFFFFF past the last unit).  The second is loaded one window ahead, so
    // the common iteration reads its two units by lane index, without a load.
    int wlo = lo0;
    auto wload = [&](int base, uint32_t& st, uint32_t& kl) {
        const bool in = base + lane < nk;
        st = in ? S.kstart()[base + lane] : 0xFFFFFFFFu;
        kl = in ? S.klist()[base + lane] : 0u;
    };
    uint32_t ast, akl, bst, bkl;
    wload(wlo, ast, akl);
    wload(wlo + WAVE, bst, bkl);
    // start and code of unit wlo + t, t < 128 wave-uniform
    auto wst = [&](int t) -> uint32_t {
        return t < WAVE ? (uint32_t)__builtin_amdgcn_readlane((int)ast, t)
                        : (uint32_t)__builtin_amdgcn_readlane((int)bst, t - WAVE);
    };
    auto wkl = [&](int t) -> uint32_t {
        return t < WAVE ? (uint32_t)__builtin_amdgcn_readlane((int)akl, t)
                        : (uint32_t)__builtin_amdgcn_readlane((int)bkl, t - WAVE);
    };
    // candidate i0 + lane of the kept enumeration: its kept unit (ku), that
    // unit's start in the kept enumeration (ks), and its loads; advances lo0
    auto load = [&](uint32_t i0, YzFirst& c, uint32_t& ks, int& ku) {
        const uint32_t i = i0 + (uint32_t)lane;
        while (lo0 - wlo >= WAVE) {
            ast = bst;
            akl = bkl;
            wlo += WAVE;
            wload(wlo + WAVE, bst, bkl);
        }
        const int s0 = lo0 - wlo;
        // every kept unit has at least one candidate, so the wave's 64 candidates
        // lie in units lo0 .. lo0 + 63, whose starts increase
        const uint32_t st1 = wst(s0 + 1), st2 = wst(s0 + 2);
        int tu, tn;
        uint32_t st_t, code;
        if (st2 > i0 + WAVE - 1) {
            // the wave's candidates lie in units lo0 and lo0 + 1 (the common case);
            // the units starting at or before i0 + 64 are lo0, and lo0 + 1, + 2 as
            // their starts tell
            tu = i >= st1 ? 1 : 0;
            st_t = tu ? st1 : wst(s0);
            code = tu ? wkl(s0 + 1) : wkl(s0);
            tn = (st1 <= i0 + WAVE ? 1 : 0) + (st2 <= i0 + WAVE ? 1 : 0);
        } else {
            // lane t takes the start and code of unit lo0 + t, and each lane
            // finds the last start at or before its candidate by halving
            const int src = s0 + lane;
            const uint32_t sa = __shfl(ast, src & (WAVE - 1), WAVE), sb = __shfl(bst, src & (WAVE - 1), WAVE);
            const uint32_t ka = __shfl(akl, src & (WAVE - 1), WAVE), kbb = __shfl(bkl, src & (WAVE - 1), WAVE);
            const uint32_t st = src < WAVE ? sa : sb, kl = src < WAVE ? ka : kbb;
            // the units starting at or before i0 + 64 are a prefix of them
            tn = __popcll(ballot(st <= i0 + WAVE)) - 1;
            tu = 0;
#pragma unroll
            for (int step = WAVE / 2; step > 0; step >>= 1) {
                const uint32_t sv = __shfl(st, tu + step, WAVE);
                if (sv <= i) tu += step;
            }
            st_t = __shfl(st, tu, WAVE);
            code = __shfl(kl, tu, WAVE);
        }
        ks = st_t;
        ku = lo0 + tu;
        // the next window starts at the last unit starting at or before i0 + 64;
        // a unit past this window can start there only when the window's last
        // unit does (then unit lo0 started at i0)
        lo0 += tn;
        if (tn == WAVE - 1)
            while (lo0 + 1 < nk && S.kstart()[lo0 + 1] <= i0 + WAVE) ++lo0;
        c.index = -1;
        c.ustart = 0;
        if (i < run) {
            const int j = (int)(code >> 16), iuv3 = (int)(code & 0xFFFFu);
            const int inner = (int)(i - st_t);
            const int4 nl = *reinterpret_cast<const int4*>(Y.nl[j]);
            const int ojs = Y.uoff[j];
            const int v2 = qdiv(inner, Y.rx[j]), v3 = qdiv(iuv3, Y.rz[j]);
            const int e[4] = {inner - __umul24(v2, nl.x), v2, iuv3 - __umul24(v3, nl.z), v3};
            c.ustart = ojs + __umul24(iuv3, __umul24(nl.x, nl.y));
            c.index = c.ustart + inner;
            c.m = (d3m >> j) & 1;
            yz_first_load(S, j, e, c);
        }
    };
    auto take = [&](uint32_t i0, const YzFirst& cur, uint32_t st_t, int ku) -> int {
        const bool pass = yz_first_eval(cur, zv);
        const int iorig = cur.index, ustart = cur.ustart;
        const unsigned long long m = ballot(pass);
        const int c = __popcll(m);
        if (np + c > cap) {
            // cut before the unit of the first survivor that does not fit
            const unsigned long long over = ballot(pass && rank_in(m) == cap - np);
            const int lo_l = __ffsll((long long)over) - 1;
            const uint32_t ucut_k = __shfl(st_t, lo_l, WAVE);
            const int ucut_o = __shfl(ustart, lo_l, WAVE), ku_cut = __shfl(ku, lo_l, WAVE);
            if (ucut_k == start) return -1;
            if (pass && iorig < ucut_o) S.plist()[np + rank_in(m)] = (uint32_t)iorig;
            const int nb = __popcll(ballot(pass && iorig < ucut_o));
            wsync();   // the list writes above are read by other lanes below
            // survivors listed before this iteration that belong to the cut unit go too
            const int keep = lower_bound_u32(S.plist(), np + nb, (uint32_t)ucut_o);
            cursor = ucut_k;
            lo_unit = ku_cut;
            return keep;
        }
        if (pass) S.plist()[np + rank_in(m)] = (uint32_t)iorig;
        np += c;
        return INT_MIN;
    };
    // two candidate buffers (see yz_pass1_plain)
    YzFirst ca, cb;
    uint32_t ksa = 0, ksb = 0;
    int kua = 0, kub = 0;
    load(start, ca, ksa, kua);
    for (uint32_t i0 = start; i0 < run; i0 += 2 * WAVE) {
        if (i0 + WAVE < run) load(i0 + WAVE, cb, ksb, kub);
        int r = take(i0, ca, ksa, kua);
        if (r != INT_MIN) return r;
        if (i0 + WAVE >= run) break;
        if (i0 + 2 * WAVE < run) load(i0 + 2 * WAVE, ca, ksa, kua);
        r = take(i0 + WAVE, cb, ksb, kub);
        if (r != INT_MIN) return r;
    }
    cursor = run;
    lo_unit = nk;
    return np;
}

// One batch of plain pass 1 from candidate cursor (a unit start): first fit and
// box cut of every candidate, survivors in order.  Chunk constants go to scalar
// registers (candidate offsets, which tracklets are D3m, the plane z values of
// both station-3 choices) and a candidate's four list entries are loaded
// together.  Returns as yz_pass1_kept.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) int yz_pass1_plain(Scr S, int kb, int ntkl, int cap, int I, int& cursor) {
    kb = uni(kb);
    ntkl = uni(ntkl);
    cap = uni(cap);
    I = uni(I);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    const int lane = lane_id();
    Tk* tks = S.tks();
    int off[YZ_CHUNK + 1];
#pragma unroll
    for (int t = 0; t <= YZ_CHUNK; ++t) off[t] = __builtin_amdgcn_readfirstlane(Y.uoff[t]);
    const unsigned d3m = (unsigned)ballot(lane < YZ_CHUNK && kb + lane < ntkl && tks[kb + lane].st3 == 4);
    float zv[4][2][2];
    yz_plane_z(zv);
    const int start = uni(cursor);
    int np = 0;
    // candidate i0 + lane: tracklet, list entries and unit, and its loads
    auto fill = [&](int i, int j, YzFirst& c) {
        const int4 nl = *reinterpret_cast<const int4*>(Y.nl[j]);
        const int ojs = Y.uoff[j];
        // enumeration indices stay below 2^27 and radices below 2^16: 24-bit products are exact
        const int local = i - ojs, nin = __umul24(nl.x, nl.y);
        const int iuv3 = qdiv(local, Y.rin[j]), iuv2 = local - __umul24(iuv3, nin);
        const int v2 = qdiv(iuv2, Y.rx[j]), v3 = qdiv(iuv3, Y.rz[j]);
        const int e[4] = {iuv2 - __umul24(v2, nl.x), v2, iuv3 - __umul24(v3, nl.z), v3};  // u2, v2, u3, v3
        c.index = i;
        c.ustart = ojs + __umul24(iuv3, nin);
        c.m = (d3m >> j) & 1;
        yz_first_load(S, j, e, c);
    };
    // the tracklet of the wave's first candidate (i0 only grows, and i0 < I =
    // uoff[YZ_CHUNK]); when every candidate of the wave lies in it, the
    // tracklet index is wave-uniform
    int jcur = 0;
    auto load = [&](int i0, YzFirst& c) {
        const int i = i0 + lane;
        c.index = -1;
        c.ustart = 0;
        int end0 = uni(Y.uoff[jcur + 1]);
        while (i0 >= end0) end0 = uni(Y.uoff[++jcur + 1]);
        const int j0 = jcur;
        if (i0 + WAVE <= end0 || I <= end0) {
            if (i < I) fill(i, j0, c);
        } else if (i < I) {
            int j = 0;
#pragma unroll
            for (int t = 1; t < YZ_CHUNK; ++t) j += i >= off[t] ? 1 : 0;
            fill(i, j, c);
        }
    };
    // one wave iteration: evaluate, list the survivors; a batch that fills
    // ends before the first unit that does not fit (returns its survivor count,
    // or -1 when that unit is the batch's first); INT_MIN: go on
    auto take = [&](int i0, const YzFirst& cur) -> int {
        const int i = i0 + lane;
        const bool pass = yz_first_eval(cur, zv);
        const int ustart = cur.ustart;
        const unsigned long long m = ballot(pass);
        const int c = __popcll(m);
        if (np + c > cap) {
            const unsigned long long over = ballot(pass && rank_in(m) == cap - np);
            const int ucut = __shfl(ustart, __ffsll((long long)over) - 1, WAVE);
            if (ucut == start) return -1;
            if (pass && i < ucut) S.plist()[np + rank_in(m)] = (uint32_t)i;
            const int nb = __popcll(ballot(pass && i < ucut));
            wsync();   // the list writes above are read by other lanes below
            const int keep = lower_bound_u32(S.plist(), np + nb, (uint32_t)ucut);
            cursor = ucut;
            return keep;
        }
        if (pass) S.plist()[np + rank_in(m)] = (uint32_t)i;
        np += c;
        return INT_MIN;
    };
    // two candidate buffers: the loads of the next wave iteration are in flight
    // while the current one is evaluated
    YzFirst ca, cb;
    load(start, ca);
    for (int i0 = start; i0 < I; i0 += 2 * WAVE) {
        if (i0 + WAVE < I) load(i0 + WAVE, cb);
        int r = take(i0, ca);
        if (r != INT_MIN) return r;
        if (i0 + WAVE >= I) break;
        if (i0 + 2 * WAVE < I) load(i0 + 2 * WAVE, ca);
        r = take(i0 + WAVE, cb);
        if (r != INT_MIN) return r;
    }
    cursor = I;
    return np;
}

// Candidate cache of pass 2 -> pass 3: raw LDS index of each slot's hit
// (16 bits each), the slot presence bits, the first fit, and the candidate's
// tracklet in the chunk and index among the tracklet's candidates (so that the
// later passes need not decode its index again).
__device__ __forceinline__ void yz_cache_put(uint4* c, const YzSlots& H, const Fit& f, const YzItem& it) {
    uint32_t r[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint32_t lo = ((H.pres >> (2 * t)) & 1u) ? (uint32_t)H.raw[2 * t] : 0u;
        const uint32_t hi = ((H.pres >> (2 * t + 1)) & 1u) ? (uint32_t)H.raw[2 * t + 1] : 0u;
        r[t] = lo | (hi << 16);
    }
    c[0] = make_uint4(r[0], r[1], r[2], r[3]);
    c[1] = make_uint4(H.pres, __float_as_uint(f.p0), __float_as_uint(f.p1), __float_as_uint(f.e0));
    c[2] = make_uint4(__float_as_uint(f.e1), (uint32_t)it.j, (uint32_t)it.local, 0u);
}

// The candidate exactly as yz_candidate builds it, from the cache: slot
// detectors of the station-3 choice, hits of the present slots from LDS, signs 0.
template <int MAXH, int WPE>
__device__ __forceinline__ void yz_cache_get(const Ev<MAXH, WPE>& E, const uint4* c, int st3, YzSlots& H, Fit& f) {
    const uint4 a = c[0], b = c[1], d = c[2];
    const uint32_t r[4] = {a.x, a.y, a.z, a.w};
    yz_dets(H, st3);
    H.pres = b.x;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        H.sign[s] = 0;
        if ((H.pres >> s) & 1u) {
            const int raw = (int)((r[s >> 1] >> (16 * (s & 1))) & 0xFFFFu);
            H.raw[s] = raw;
            H.pos[s] = E.Lp->pos[raw];
            H.drift[s] = E.Lp->drift[raw];
        }
    }
    f.p0 = __uint_as_float(b.y);
    f.p1 = __uint_as_float(b.z);
    f.e0 = __uint_as_float(b.w);
    f.e1 = __uint_as_float(d.x);
}

// One survivor of pass 2 (below) whose tracklet's station 3 is ST3: its
// candidate cache entry, the tests that need no sign (second box cut, prop-tube
// Y, hodoscopes: they depend on the first fit and the tracklet only) and, unless
// it is the first of its unit, the v3' / u2 pair decision.  Returns its state.
template <int ST3, int MAXH, int WPE>
__device__ __forceinline__ uchar4 yz_pass2_one(const Ev<MAXH, WPE>& E, const Scr& S, const Tk* Tp, const YzItem& it,
                                               int k, bool first) {
    YzSlots H;
    Fit f;
    yz_candidate_st<ST3>(E, S, it, H, f);
    yz_cache_put(S.ycache() + 3 * k, H, f, it);
    const bool go = yz_cuts_hx(E, yzw<MAXH>().hx[it.j], f);
    uchar4 ps = make_uchar4((first ? 1 : 0) | (go ? 2 : 0) | (ST3 == 4 ? 4 : 0), 0, 0, 0);
    if (!first && ((lr_slot_starts<8>(H.pres, H.det) >> 3) & 1u)) {
        const int h = lr_pair_decide(Tp->x0, Tp->tx, f.p0, f.p1, Tp->ex0, Tp->etx, f.e0, f.e1, H.det[3], H.det[4],
                                     H.pos[3], H.pos[4], H.drift[3], H.drift[4], 150.f);
        if (h >= 0) ps.z = (uint8_t)(lr_si(h) & 3);
    }
    return ps;
}

// Pass 2 over the survivors k < np of a batch: per survivor, first-of-unit
// flag, the candidate cache entry, the sign-free tests and the v3' / u2 pair
// decision.  Returns the number of first survivors, listed in the pass-3 list
// (free until then) for yz_pass2_firsts.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) int yz_pass2(Scr S, int kb, int np) {
    kb = uni(kb);
    np = uni(np);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    int nf = 0;
    int carry = -1;   // unit of the previous wave iteration's last survivor
    for (int k0 = 0; k0 < np; k0 += WAVE) {
        const int k = k0 + lane;
        YzItem it;
        int unit = -1;   // (tracklet, u3, v3) of the survivor; list indices stay below 256
        if (k < np) {
            it = yz_decode<MAXH, WPE>((int)S.plist()[k]);
            unit = (it.j << 16) | (it.iu3 << 8) | it.iv3;
        }
        // first of its unit: survivor k - 1 (the lane before, or the previous
        // iteration's last) lies in another unit
        const int before = __shfl_up(unit, 1, WAVE);
        const bool first = k < np && (k == 0 || unit != (lane == 0 ? carry : before));
        carry = __shfl(unit, WAVE - 1, WAVE);
        if (k < np) {
            const Tk* Tp = &tks[kb + it.j];
            // per lane by its tracklet's station 3 (a wave's survivors mostly share one tracklet)
            S.pst()[k] = Tp->st3 == 3 ? yz_pass2_one<3>(E, S, Tp, it, k, first) : yz_pass2_one<4>(E, S, Tp, it, k, first);
        }
        const unsigned long long fm = ballot(first);
        if (first) S.qlist()[nf + rank_in(fm)] = (uint32_t)k;
        nf += __popcll(fm);
    }
    return nf;
}

// The first survivors of the units (listed by yz_pass2) run the full
// left/right from zero signs; their station-3 signs seed the unit.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void yz_pass2_firsts(Scr S, int kb, int nf) {
    kb = uni(kb);
    nf = uni(nf);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    for (int q0 = 0; q0 < nf; q0 += WAVE) {
        const int q = q0 + lane;
        if (q < nf) {
            const int k = (int)S.qlist()[q];
            const int j = (int)gld(&S.ycache()[3 * k + 2].y);   // the survivor's tracklet (yz_cache_put)
            const Tk* Tp = &tks[kb + j];
            // per lane by its tracklet's station 3, as in pass 2
            auto lr = [&](auto st3c) {
                YzSlots H;
                Fit f;
                yz_cache_get(E, S.ycache() + 3 * k, decltype(st3c)::value, H, f);
                lr_slots_new<8>(Tp->x0, Tp->tx, f.p0, f.p1, Tp->ex0, Tp->etx, f.e0, f.e1, H.pres, H.det, H.pos,
                                H.drift, H.sign, 150.f);
                uchar4 ps = S.pst()[k];
                ps.y = yz_pack3(H);
                ps.z = (uint8_t)(H.sign[3] & 3);
                S.pst()[k] = ps;
            };
            if (Tp->st3 == 3) lr(std::integral_constant<int, 3>{});
            else lr(std::integral_constant<int, 4>{});
        }
    }
}

// Pass 3 scan: each survivor's starting signs (the unit's seed, and the last
// writer of the v3' sign before it); the survivors that passed the sign-free
// tests are listed.  Returns their number.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) int2 yz_pass3_scan(Scr S, int np) {
    np = uni(np);
    const int lane = lane_id();
    // pass 3: starting signs by scan (the unit's seed, and the last writer of
    // the v3' sign before this candidate); the candidates that passed the
    // sign-free tests in pass 2 are listed and run the rest (left/right, chi2):
    // those of D3p tracklets from the front of the list, those of D3m ones from
    // its back, so that each part has one set of planes
    int carry_f = -1, carry_w = -1, nq = 0, nq4 = 0;
    const int cap = cfg().list_cap;
    for (int k0 = 0; k0 < np; k0 += WAVE) {
        const int k = k0 + lane;
        const bool valid = k < np;
        uchar4 ps = make_uchar4(0, 0, 0, 0);
        if (valid) ps = S.pst()[k];
        const bool first = (ps.x & 1u) != 0;
        int fi = wave_max_scan(valid && first ? k : -1);
        fi = max(fi, carry_f);
        const int wincl = wave_max_scan(valid && (first || ps.z) ? k : -1);
        int wexcl = __shfl_up(wincl, 1, WAVE);
        if (lane == 0) wexcl = -1;
        wexcl = max(wexcl, carry_w);
        carry_f = __shfl(fi, WAVE - 1, WAVE);
        carry_w = max(carry_w, __shfl(wincl, WAVE - 1, WAVE));
        const bool go = valid && (ps.x & 2u);
        if (go && !first) {
            // seed of the unit with the v3' sign (slot 3) of the last writer; the
            // first survivor of the unit is a writer, so wexcl >= fi
            ps.w = (uint8_t)((S.pst()[fi].y & 0x3Fu) | ((unsigned)S.pst()[wexcl].z << 6));
            S.pst()[k] = ps;
        }
        const bool d3m = (ps.x & 4u) != 0;
        const unsigned long long m = ballot(go && !d3m), m4 = ballot(go && d3m);
        if (go && !d3m) S.qlist()[nq + rank_in(m)] = (uint32_t)k;
        if (go && d3m) S.qlist()[cap - 1 - (nq4 + rank_in(m4))] = (uint32_t)k;
        nq += __popcll(m);
        nq4 += __popcll(m4);
    }
    return make_int2(nq, nq4);
}

// chi2 of a YZ candidate of a tracklet whose station 3 is ST3 (yz_chi2 without
// the write): the planes of the x hits (D2 X, X', D3 X, X', absent ones left
// out, the order xz_candidate stores them in) and of the UV slots are then
// wave-uniform, so their constants come through scalar loads; the sum runs in
// the same order
template <int ST3, int MAXH, int WPE>
__device__ __forceinline__ float yz_chi2_st(const Ev<MAXH, WPE>& E, const Tk* Tp, int j, const YzSlots& H,
                                            const Fit& f) {
    const YzWave& Y = yzw<MAXH>();
    const float x0 = Tp->x0, tx = Tp->tx;
    const float y0 = f.p0, ty = f.p1;
    const int d3x = ST3 == 3 ? 22 : 28;   // as xz_candidate sets it
    const int xp[4] = {16, 15, d3x, d3x - 1};
    float chi2 = 0.f;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int d = xp[s], r = Y.xh[j][s];
        if (r != NOHIT) {
            const float res = G(F_RESOLUTION, d);
            const float rs = E.Lp->drift[r] * (float)Y.xs[j][s] - dca_wire(d, E.Lp->elem[r], x0, y0, tx, ty);
            chi2 += ((rs * rs) / res) / res;
        }
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if (!((H.pres >> s) & 1u)) continue;
        const int d = H.det[s];
        const float res = G(F_RESOLUTION, d);
        const float rs = H.drift[s] * (float)H.sign[s] - dca_wire(d, E.Lp->elem[H.raw[s]], x0, y0, tx, ty);
        chi2 += ((rs * rs) / res) / res;
    }
    return chi2;
}

// Pass 3 over the listed survivors: left/right from the starting signs, chi2,
// and the tracklet's best key (64-bit LDS atomicMin of chi2 bits and index);
// the D3p part of the list (nq from its front), then the D3m part (nq4 from its back)
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void yz_pass3(Scr S, int kb, int nq, int nq4) {
    kb = uni(kb);
    nq = uni(nq);
    nq4 = uni(nq4);
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    const int cap = cfg().list_cap;
    auto eval = [&](auto st3c, int k) {
        constexpr int ST3 = decltype(st3c)::value;
        const uchar4 ps = S.pst()[k];
        // the survivor's tracklet and index among the tracklet's candidates (yz_cache_put)
        const uint4 c2 = gld4(S.ycache() + 3 * k + 2);
        const int j = (int)c2.y;
        const uint32_t local = c2.z;
        const Tk* Tp = &tks[kb + j];
        YzSlots H;
        Fit f;
        yz_cache_get(E, S.ycache() + 3 * k, ST3, H, f);
        if (!(ps.x & 1u)) yz_unpack3(H, ps.w);
        lr_slots_new<8>(Tp->x0, Tp->tx, f.p0, f.p1, Tp->ex0, Tp->etx, f.e0, f.e1, H.pres, H.det, H.pos,
                        H.drift, H.sign, 150.f);
        const float chi2 = yz_chi2_st<ST3>(E, Tp, j, H, f);
        // the reference starts its minimum at 10000.1 (reconstruction_kernels.cuh:1240)
        if (chi2 < 10000.1f) {
            const unsigned long long key = ((unsigned long long)__float_as_uint(chi2) << 32) | (unsigned long long)local;
            atomicMin(&Y.best[j], key);
        }
    };
    for (int q0 = 0; q0 < nq; q0 += WAVE) {
        const int q = q0 + lane;
        if (q < nq) eval(std::integral_constant<int, 3>{}, (int)S.qlist()[q]);
    }
    for (int q0 = 0; q0 < nq4; q0 += WAVE) {
        const int q = q0 + lane;
        if (q < nq4) eval(std::integral_constant<int, 4>{}, (int)S.qlist()[cap - 1 - q]);
    }
    wsync();
}

// Parallel form over batches of first-box survivors (see yz_pass1_kept).
// Returns false (nothing written) when one unit alone overflows the list; the
// chunk then takes the unit-by-unit path.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) bool yz_chunk_par(Scr S, Prn R, int kb, int ntkl, int cap) {
    kb = uni(kb);
    ntkl = uni(ntkl);
    cap = uni(cap);
    R.prune = uni(R.prune);
    R.dump = uni(R.dump);
    R.dump_cap = uni(R.dump_cap);
    R.dump_count = uni(R.dump_count);
    R.ev = uni(R.ev);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    {
        const int j = lane, k = kb + j;
        int items = 0;
        if (j < YZ_CHUNK && k < ntkl) {
            const int* nl = Y.nl[j];
            // at most YZ_ENUM_MAX each (stage_yz), so the sum stays below 2^31
            if (nl[0] > 0 && nl[1] > 0 && nl[2] > 0 && nl[3] > 0) items = nl[0] * nl[1] * nl[2] * nl[3];
        }
        const int inc = wave_scan(items);
        if (j < YZ_CHUNK) {
            Y.best[j] = ~0ull;
            Y.uoff[j + 1] = inc;
            const int* nl = Y.nl[j];
            Y.rx[j] = qrcp(nl[0]);
            Y.rin[j] = qrcp(nl[0] * nl[1]);
            Y.rz[j] = qrcp(nl[2]);
        }
        if (j == 0) Y.uoff[0] = 0;
    }
    wsync();
    const int I = Y.uoff[YZ_CHUNK];
    OR_COUNT(10, I);
    UnitScore us = {-1, 0u, -1};
    if (R.prune || R.dump) us = yz_score_units<MAXH, WPE>(S, R, kb, ntkl);
    const bool kept = us.nk >= 0;
    const int64_t dbase = us.dbase;
    uint32_t cursor = 0;
    int icursor = 0, lo_unit = 0;
    unsigned long long saved = ~0ull;   // lane j < 16: best key of tracklet j already in its record
    while (kept ? cursor < us.run : icursor < I) {
        OR_SUB_START
        const int np = kept ? yz_pass1_kept<MAXH, WPE>(S, kb, ntkl, cap, us.nk, us.run, cursor, lo_unit)
                            : yz_pass1_plain<MAXH, WPE>(S, kb, ntkl, cap, I, icursor);
        if (np < 0) {
            // the chunk goes the unit-by-unit way: its unit rows carry no winner flags
            if (dbase >= 0) {
                int units = 0;
                if (lane < YZ_CHUNK && kb + lane < ntkl) {
                    const int* nl = Y.nl[lane];
                    if (nl[0] > 0 && nl[1] > 0 && nl[2] > 0 && nl[3] > 0) units = nl[2] * nl[3];
                }
                const int U = wave_sum(units);
                for (int u = lane; u < U; u += WAVE) R.dump[(dbase + u) * OR_UNIT_ROW + 3] = -1.f;
            }
            return false;
        }
        wsync();
        OR_SUB(1);
        OR_COUNT(14, np);
        // pass 2: per survivor, first-of-unit flag; the first of a unit runs the
        // full left/right from zero signs (its station-3 result seeds the unit),
        // later ones only the decision of a v3' / u2 pair (the one pair that can
        // change a station-3 sign once all are set).  A batch starts at a unit
        // boundary, so its first survivor is the first of its unit.  The first
        // survivors run their left/right in a pass of their own, so that the
        // other lanes of a wave do not wait through it.
        const int nf = yz_pass2<MAXH, WPE>(S, kb, np);
        wsync();
        yz_pass2_firsts<MAXH, WPE>(S, kb, nf);
        wsync();
        OR_SUB(2);
        // pass 3: starting signs by scan; the candidates that passed the
        // sign-free tests in pass 2 run the rest (left/right, chi2)
        const int2 nq = yz_pass3_scan<MAXH, WPE>(S, np);
        wsync();
        OR_SUB(15);
        OR_COUNT(16, nq.x + nq.y);
        yz_pass3<MAXH, WPE>(S, kb, nq.x, nq.y);
        wsync();
        // a tracklet whose best improved in this batch keeps that candidate's
        // cache entry and sign state, since the next batch reuses the lists
        if (lane < YZ_CHUNK && kb + lane < ntkl) {
            const unsigned long long key = Y.best[lane];
            if (key != saved) {
                const uint32_t iw = (uint32_t)Y.uoff[lane] + (uint32_t)key;
                const int k = min(lower_bound_u32(S.plist(), np, iw), np - 1);   // present: set in this batch
                uint4* br = S.brec() + 4 * lane;
                const uint4* c = S.ycache() + 3 * k;
                br[0] = c[0]; br[1] = c[1]; br[2] = c[2];
                const uchar4 ps = S.pst()[k];
                br[3] = make_uint4((uint32_t)ps.x | ((uint32_t)ps.y << 8) | ((uint32_t)ps.z << 16) | ((uint32_t)ps.w << 24),
                                   0u, 0u, 0u);
                saved = key;
            }
        }
        wsync();
        OR_SUB(3);
    }
    OR_SUB_START
    // winners: replay the candidate with its starting signs and write it
    int uo_lane = 0;   // first unit row of tracklet j = lane
    if (dbase >= 0) {
        int units = 0;
        if (lane < YZ_CHUNK && kb + lane < ntkl) {
            const int* nl = Y.nl[lane];
            if (nl[0] > 0 && nl[1] > 0 && nl[2] > 0 && nl[3] > 0) units = nl[2] * nl[3];
        }
        uo_lane = wave_scan(units) - units;
    }
    if (lane < YZ_CHUNK && kb + lane < ntkl) {
        const int j = lane;
        Tk* Tp = &tks[kb + j];
        const unsigned long long key = Y.best[j];
        if (key != ~0ull) {
            const uint4* br = S.brec() + 4 * j;
            const uint32_t psw = br[3].x;
            YzSlots H;
            Fit f;
            yz_cache_get(E, br, Tp->st3, H, f);
            if (!(psw & 1u)) yz_unpack3(H, psw >> 24);
            float chi2;
            yz_tail(E, Tp, H, f, chi2, Tp);
            const int local = (int)(uint32_t)key;
            if (dbase >= 0) R.dump[(dbase + uo_lane + local / (Y.nl[j][0] * Y.nl[j][1])) * OR_UNIT_ROW + 3] = 1.f;
        } else {
            Tp->stid = 3;
        }
    }
    OR_SUB(4);
    return true;
}

// Unit loop: one lane per (tracklet, u3.v3) unit, candidates in reference order.
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void yz_chunk_seq(Scr S, int kb, int ntkl) {
    kb = uni(kb);
    ntkl = uni(ntkl);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    uint16_t* wl = S.wl();
    float4* wy = S.wy();
    uint8_t* wf = S.wf();
    {
        const int j = lane, k = kb + j;
        int u = 0;
        if (j < YZ_CHUNK && k < ntkl) {
            const int* nl = Y.nl[j];
            if (nl[0] > 0 && nl[1] > 0 && nl[2] > 0 && nl[3] > 0) u = nl[2] * nl[3];
        }
        const int inc = wave_scan(u);
        if (j < YZ_CHUNK) {
            Y.best[j] = ~0ull;
            Y.uoff[j + 1] = inc;
        }
        if (j == 0) Y.uoff[0] = 0;
    }
    wsync();
    const int U = Y.uoff[YZ_CHUNK];
    for (int u0 = 0; u0 < U; u0 += WAVE) {
        const int u = u0 + lane;
        if (u < U) {
            int j = 0;
            while (Y.uoff[j + 1] <= u) ++j;
            const int iuv3 = u - Y.uoff[j];
            YzLists W;
            for (int v = 0; v < 4; ++v) {
                W.l[v] = wl + (j * 4 + v) * cfg().win_cap; W.y[v] = wy + (j * 4 + v) * cfg().win_cap;
                W.f[v] = wf + (j * 4 + v) * cfg().win_cap; W.n[v] = Y.nl[j][v];
            }
            float bc = 10000.1f;
            int bi = -1;
            if (yz_unit(E, &tks[kb + j], W, iuv3, -1, bc, bi)) {
                const unsigned long long key =
                    ((unsigned long long)__float_as_uint(bc) << 32) |
                    (unsigned long long)(uint32_t)(iuv3 * (W.n[0] * W.n[1]) + bi);
                atomicMin(&Y.best[j], key);
            }
        }
    }
    wsync();
    if (lane < YZ_CHUNK && kb + lane < ntkl) {
        const int j = lane;
        Tk* Tp = &tks[kb + j];
        const unsigned long long key = Y.best[j];
        if (key != ~0ull) {
            YzLists W;
            for (int v = 0; v < 4; ++v) {
                W.l[v] = wl + (j * 4 + v) * cfg().win_cap; W.y[v] = wy + (j * 4 + v) * cfg().win_cap;
                W.f[v] = wf + (j * 4 + v) * cfg().win_cap; W.n[v] = Y.nl[j][v];
            }
            const int flat = (int)(uint32_t)key, nin = W.n[0] * W.n[1];
            float bc = 0.f;
            int bi = 0;
            yz_unit(E, Tp, W, flat / nin, flat % nin, bc, bi);
        } else {
            Tp->stid = 3;
        }
    }
}

template <int MAXH, int WPE>
STAGE void stage_yz(or_params P, Scr S, Prn R, int ntkl, int list_cap) {
    ntkl = uni(ntkl);
    list_cap = uni(list_cap);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    uint16_t* wl = S.wl();
    float4* wy = S.wy();
    float2* wyy = S.wyy();
    uint8_t* wf = S.wf();
    const int capw = cfg().win_cap;
    // chunks of YZ_CHUNK tracklets, handed to the event's waves as they come free
    for (;;) {
        int kb = 0;
        if (lane == 0) kb = atomicAdd(&L.yz_next, YZ_CHUNK);
        kb = __builtin_amdgcn_readfirstlane(__shfl(kb, 0, WAVE));
        if (kb >= ntkl) break;
        OR_SUB_START
        // window lists: lane = (tracklet j, view v)
        {
            const int j = lane >> 2, v = lane & 3;
            const int k = kb + j;
            unsigned long long raw = 1;   // list length before the accepted-hit compaction
            if (k < ntkl) {
                const Tk& T = tks[k];
                const int st = v < 2 ? 2 : T.st3, pj = (v & 1) ? 2 : 1;
                const int sid = super_id(st, pj);
                const int da = sid * 2, db = sid * 2 - 1;
                const float za = G(F_Z, da), zb = G(F_Z, db);
                const float xmin = fminf((T.tx - T.etx) * za, (T.tx - T.etx) * zb) + T.x0 - T.ex0 - G(F_SPACING, da);
                const float xmax = fmaxf((T.tx + T.etx) * za, (T.tx + T.etx) * zb) + T.x0 + T.ex0 + G(F_SPACING, da);
                uint16_t* lst = wl + (j * 4 + v) * cfg().win_cap;
                const int nl = pairs_window(E, st, pj, xmin, xmax, lst, P.max_pair_plane_hits, capw);
                if (nl < 0) atomicOr(&L.status, (uint32_t)OR_EV_PAIR_OVERFLOW);
                raw = nl > 0 ? (unsigned long long)nl : 0ull;
                // y of each listed hit from the tracklet's x line (calculate_y_uvhit,
                // sign 0).  An entry with no accepted hit ends every candidate that
                // contains it before its fit (reconstruction_kernels.cuh:1390, 1457,
                // 1533, 1600), so the list keeps only entries with an accepted hit,
                // in order: the candidates enumerated over the kept lists are the
                // reference's fitted candidates as an order-preserving subsequence.
                int nacc = 0;
                for (int q = 0; q < nl; ++q) {
                    float4 yv = make_float4(0.f, 0.f, 0.f, 0.f);
                    uint8_t fl = 0;
                    const int ia = pa(lst[q]), ib = pb(lst[q]);
                    if (ia >= 0) {
                        const int r = E.raw(da, ia);
                        if (y_uvhit(da, L.elem[r], L.drift[r], 0, T.x0, T.tx, yv.x, yv.y)) fl |= 1;
                    }
                    if (ib >= 0) {
                        const int r = E.raw(db, ib);
                        if (y_uvhit(db, L.elem[r], L.drift[r], 0, T.x0, T.tx, yv.z, yv.w)) fl |= 2;
                    }
                    if (fl) {
                        const int o = (j * 4 + v) * cfg().win_cap + nacc;
                        lst[nacc] = lst[q];
                        // fit terms w y and (w y) z of each hit, as the fit sums them
                        const uint8_t fb = (uint8_t)(fl | ((fl - 1u) * (v == 0 ? 3u : v == 1 ? 1u : v == 2 ? 27u : 9u) << 2));
                        // (0 for a hit that is absent or not accepted, so that the B
                        // sums can add every term: see yz_bsum)
                        wy[o] = make_float4((fl & 1u) ? yv.y * yv.x : 0.f, (fl & 2u) ? yv.w * yv.z : 0.f,
                                            __uint_as_float(fb), 0.f);
                        // station 3: the y of hits a and b; station 2: the entry's y
                        // (the mean of its accepted hits) for the pruner's features
                        wyy[o] = v >= 2 ? make_float2(yv.x, yv.z)
                                        : make_float2((fl & 1u) ? ((fl & 2u) ? 0.5f * (yv.x + yv.z) : yv.x) : yv.z, 0.f);
                        // the entry's term of the first-fit pattern index (yz_pattern)
                        wf[o] = fb;
                        ++nacc;
                    }
                }
                Y.nl[j][v] = nacc;
            }
            // hodoscope x halves of the tracklet's matches: lane (j, v < 3) takes
            // station v + 1's two planes
            unsigned hb = 0;
            if (k < ntkl && v < 3) {
                const Tk& T = tks[k];
                hb = (hodo_x(E, v + 1, yz_hodo(2 * v), T.x0, T.tx, T.ex0, T.etx) ? 1u : 0u) |
                     (hodo_x(E, v + 1, yz_hodo(2 * v + 1), T.x0, T.tx, T.ex0, T.etx) ? 2u : 0u);
                hb <<= 2 * v;
            }
            hb |= (unsigned)__shfl_xor((int)hb, 1, WAVE);
            hb |= (unsigned)__shfl_xor((int)hb, 2, WAVE);
            if (v == 0) Y.hx[j] = hb;
            // x hit of slot v (D2 X, X', D3 X, X'), as xz_candidate stores them in order
            if (k < ntkl) {
                const Tk& T = tks[k];
                const int d3x = T.st3 == 3 ? 22 : 28;
                const int dv = v == 0 ? 16 : v == 1 ? 15 : v == 2 ? d3x : d3x - 1;
                uint16_t xr = NOHIT;
                int8_t xsg = 0;
                for (int h = 0; h < T.nx; ++h)
                    if (T.det[h] == dv) { xr = T.hraw[h]; xsg = T.sign[h]; }
                Y.xh[j][v] = xr;
                Y.xs[j][v] = xsg;
            }
            // enumeration bound: u2 x v2 x u3 x v3 list entries of one tracklet
            raw *= (unsigned long long)__shfl_xor((long long)raw, 1, WAVE);
            raw *= (unsigned long long)__shfl_xor((long long)raw, 2, WAVE);
            if (raw > (unsigned long long)YZ_ENUM_MAX) atomicOr(&L.status, (uint32_t)OR_EV_PAIR_OVERFLOW);
        }
        wsync();
        OR_SUB(0);
        if (__atomic_load_n(&L.status, __ATOMIC_RELAXED)) break;   // set by any wave: the event is rejected
        if (!yz_chunk_par<MAXH, WPE>(S, R, kb, ntkl, list_cap)) {
            wsync();
            OR_SUB_START
            yz_chunk_seq<MAXH, WPE>(S, kb, ntkl);
            OR_SUB(5);
        }
        wsync();
    }
}

// F: back-partial cleaning (reconstruction_helper.cuh:1789-1909), slot order:
// the within-slot pass for slots 0..31, then the cross pass slot against slot.
//
// Lane form: the tracklets at stationID >= 5 are gathered in slot order, one
// per lane (up to 64), each holding its hits as one raw index per detector
// 13-30 (a tracklet has at most one hit per plane).  The reference's loops run
// in their order over i; its inner loop over j is done by all lanes at once:
// j's alive at the start of the i-step stay alive until their own comparison,
// so nhits_common at j is a prefix sum over them, and the step's decisions
// (demote j, or demote i) are independent and applied together.  More than 64
// such tracklets take the one-lane form.
__device__ __forceinline__ int clean_shared(const int (&ha)[18], const int (&hb)[18]) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < 18; ++t) c += (ha[t] >= 0 && ha[t] == hb[t]) ? 1 : 0;
    return c;
}

// More than 64 live tracklets: each is held in scratch as 9 words of hits (the
// raw hit of detectors 13-30, 16 bits each, 0xFFFF none), then tracklet | slot
// << 16 | nhits << 24 and chi2/(nhits - 4); alive bits by live position in LDS.
// An i-step runs over its j range in chunks of 64 with the shared-hit prefix
// sum carried from chunk to chunk.
struct CleanEnt {
    uint32_t h[9];
    uint32_t kn;
    float ratio;
};
__device__ __forceinline__ CleanEnt clean_load(const uint4* ent, int p) {
    const uint4 a = gld4(ent + 3 * p), b = gld4(ent + 3 * p + 1), c = gld4(ent + 3 * p + 2);
    CleanEnt e;
    e.h[0] = a.x; e.h[1] = a.y; e.h[2] = a.z; e.h[3] = a.w;
    e.h[4] = b.x; e.h[5] = b.y; e.h[6] = b.z; e.h[7] = b.w;
    e.h[8] = c.x; e.kn = c.y; e.ratio = __uint_as_float(c.z);
    return e;
}
__device__ __forceinline__ int clean_shared9(const uint32_t (&a)[9], const uint32_t (&b)[9]) {
    int c = 0;
#pragma unroll
    for (int w = 0; w < 9; ++w) {
        const uint32_t x = a[w] ^ b[w];
        c += ((x & 0xFFFFu) == 0u && (a[w] & 0xFFFFu) != 0xFFFFu) ? 1 : 0;
        c += ((x >> 16) == 0u && (a[w] >> 16) != 0xFFFFu) ? 1 : 0;
    }
    return c;
}

template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void clean_many(Tk* tks, const uint16_t* slot_order, int ntkl, int nlive) {
    tks = uni(tks);
    slot_order = uni(slot_order);
    ntkl = uni(ntkl);
    nlive = uni(nlive);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    uint4* ent = Scr().clean_hits();
    if (lane < 33) L.cl.lstart[lane] = 0;
    for (int w = lane; w < CLEAN_WORDS; w += WAVE) {
        const int b0 = w * 64;
        L.cl.alive[w] = nlive >= b0 + 64 ? ~0ull : (nlive > b0 ? ((1ull << (nlive - b0)) - 1ull) : 0ull);
    }
    wsync();
    // entries in slot order, live tracklets per slot
    int np = 0;
    for (int j0 = 0; j0 < ntkl; j0 += WAVE) {
        const int j = j0 + lane;
        const int k = j < ntkl ? slot_order[j] : 0;
        const bool live = j < ntkl && tks[k].stid >= 5;
        const unsigned long long m = ballot(live);
        if (live) {
            const Tk& T = tks[k];
            const int n = T.nx + T.nuv;
            int hm[18];
#pragma unroll
            for (int t = 0; t < 18; ++t) hm[t] = 0xFFFF;
            for (int h = 0; h < 12; ++h) {
                if (h >= n) break;
                const int d = T.det[h], r = T.hraw[h];
#pragma unroll
                for (int t = 0; t < 18; ++t)
                    if (d == 13 + t) hm[t] = r;
            }
            const int p = np + rank_in(m);
            const float ratio = T.chi2 / (float)(n - 4);
            uint4* e = ent + 3 * p;
            e[0] = make_uint4(hm[0] | (hm[1] << 16), hm[2] | (hm[3] << 16), hm[4] | (hm[5] << 16), hm[6] | (hm[7] << 16));
            e[1] = make_uint4(hm[8] | (hm[9] << 16), hm[10] | (hm[11] << 16), hm[12] | (hm[13] << 16),
                              hm[14] | (hm[15] << 16));
            e[2] = make_uint4(hm[16] | (hm[17] << 16), (uint32_t)k | ((uint32_t)T.slot << 16) | ((uint32_t)n << 24),
                              __float_as_uint(ratio), 0u);
            atomicAdd(&L.cl.lstart[T.slot + 1], 1);
        }
        np += __popcll(m);
    }
    wsync();
    {
        const int c = lane < 33 ? L.cl.lstart[lane] : 0;
        const int inc = wave_scan(c);
        wsync();
        if (lane < 33) L.cl.lstart[lane] = inc;
    }
    wsync();
    // one i-step: A at live position pa against the alive positions of [jb, je)
    auto step = [&](int pa, int jb, int je) {
        const CleanEnt A = clean_load(ent, pa);
        const int na = (int)(A.kn >> 24);
        int carry = 0;
        bool di = false;
        for (int c0 = jb; c0 < je; c0 += WAVE) {
            const int j = c0 + lane;
            const bool part = j < je && ((L.cl.alive[j >> 6] >> (j & 63)) & 1ull);
            int c = 0;
            CleanEnt B;
            if (part) {
                B = clean_load(ent, j);
                c = clean_shared9(A.h, B.h);
            }
            const int common = carry + wave_scan(c);
            bool dj = false, dl = false;
            if (part) {
                if (A.ratio < B.ratio) dj = (float)common > (float)na * 0.3333f;
                else dl = (float)common > (float)(int)(B.kn >> 24) * 0.3333f;
            }
            wsync();
            if (dj) atomicAnd(&L.cl.alive[j >> 6], ~(1ull << (j & 63)));
            di = di || ballot(dl) != 0ull;
            carry = __shfl(common, WAVE - 1, WAVE);
        }
        if (di && lane == 0) atomicAnd(&L.cl.alive[pa >> 6], ~(1ull << (pa & 63)));
        wsync();
    };
    // within each slot (reconstruction_helper.cuh:1789-1843)
    for (int pa = 0; pa < nlive; ++pa) {
        if (!((L.cl.alive[pa >> 6] >> (pa & 63)) & 1ull)) continue;
        const int sa = (int)((gld4(ent + 3 * pa + 2).y >> 16) & 0xFFu);
        step(pa, pa + 1, L.cl.lstart[sa + 1]);
    }
    // slot against slot (reconstruction_helper.cuh:1846-1909): for s, for o != s,
    // for i in s, for j in o.  A step (i, o) reads and changes only the alive
    // bits of i and of o's tracklets, so steps of different i and different o
    // commute: any order that keeps the steps of one i in o order and those of
    // one o in i order gives the same result.  Here, for s, for i in s, all o at
    // once: the j of every other slot in one pass (positions are in slot order),
    // shared hits summed per slot by a segmented scan, and the first slot whose
    // step demotes i is i's last (the j demotions of the slots after it are not
    // applied, as those steps do not run for a dead i).
    auto cross = [&](int pa, int sb, int se) {
        const CleanEnt A = clean_load(ent, pa);
        const int na = (int)(A.kn >> 24);
        int run = 0, base_carry = 0, mlim = 32;
        for (int c0 = 0; c0 < nlive; c0 += WAVE) {
            const int j = c0 + lane;
            const bool in = j < nlive && (j < sb || j >= se);
            const bool al = in && ((L.cl.alive[j >> 6] >> (j & 63)) & 1ull);
            int c = 0, o = 32;
            CleanEnt B;
            if (in) {
                B = clean_load(ent, j);
                o = (int)((B.kn >> 16) & 0xFFu);
            }
            // past the slot that demoted i: nothing is left
            if (mlim < 32 && ballot(in && o <= mlim) == 0ull) break;
            if (al) c = clean_shared9(A.h, B.h);
            const int incl = run + wave_scan(c);
            const bool head = in && j == L.cl.lstart[o];
            const int base = max(base_carry, wave_max_scan(head ? incl - c : -1));
            const int common = incl - base;
            bool dj = false, dl = false;
            if (al) {
                if (A.ratio < B.ratio) dj = (float)common > (float)na * 0.3333f;
                else dl = (float)common > (float)(int)(B.kn >> 24) * 0.3333f;
            }
            const unsigned long long dlm = ballot(dl && o <= mlim);
            if (dlm) mlim = min(mlim, __shfl(o, __ffsll((long long)dlm) - 1, WAVE));
            wsync();
            if (dj && o <= mlim) atomicAnd(&L.cl.alive[j >> 6], ~(1ull << (j & 63)));
            run = __shfl(incl, WAVE - 1, WAVE);
            base_carry = __shfl(base, WAVE - 1, WAVE);
        }
        if (mlim < 32 && lane == 0) atomicAnd(&L.cl.alive[pa >> 6], ~(1ull << (pa & 63)));
        wsync();
    };
    for (int sidx = 0; sidx < 32; ++sidx) {
        const int sb = L.cl.lstart[sidx], se = L.cl.lstart[sidx + 1];
        if (sb == se || se - sb == nlive) continue;
        for (int pa = sb; pa < se; ++pa) {
            if (!((L.cl.alive[pa >> 6] >> (pa & 63)) & 1ull)) continue;
            cross(pa, sb, se);
        }
    }
    for (int p = lane; p < nlive; p += WAVE)
        if (!((L.cl.alive[p >> 6] >> (p & 63)) & 1ull)) tks[gld4(ent + 3 * p + 2).y & 0xFFFFu].stid = 3;
    wsync();
}

template <int MAXH, int WPE>
STAGE void stage_clean(Tk* tks, const uint16_t* slot_order, int ntkl) {
    tks = uni(tks);
    slot_order = uni(slot_order);
    ntkl = uni(ntkl);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    // gather in slot order
    int nlive = 0;
    for (int j0 = 0; j0 < ntkl; j0 += WAVE) {
        const int j = j0 + lane;
        const int k = j < ntkl ? slot_order[j] : 0;
        const bool live = j < ntkl && tks[k].stid >= 5;
        const unsigned long long m = ballot(live);
        if (live && nlive + rank_in(m) < WAVE) L.live[nlive + rank_in(m)] = (uint16_t)k;
        nlive += __popcll(m);
    }
    wsync();
    if (nlive == 0) return;
    if (nlive <= WAVE) {
        const bool mine = lane < nlive;
        const int k = mine ? L.live[lane] : 0;
        int slot = -1, n = 0;
        float ratio = 0.f;
        int hm[18];
#pragma unroll
        for (int t = 0; t < 18; ++t) hm[t] = -1;
        if (mine) {
            const Tk& T = tks[k];
            slot = T.slot;
            n = T.nx + T.nuv;
            ratio = T.chi2 / (float)(n - 4);
            for (int h = 0; h < 12; ++h) {
                if (h >= n) break;
                const int d = T.det[h], r = T.hraw[h];
#pragma unroll
                for (int t = 0; t < 18; ++t)
                    if (d == 13 + t) hm[t] = r;
            }
        }
        unsigned long long alive = nlive == WAVE ? ~0ull : ((1ull << nlive) - 1ull);
        // one i-step: A at position pa against the lanes flagged in part
        auto step = [&](int pa, bool part) {
            int ha[18];
#pragma unroll
            for (int t = 0; t < 18; ++t) ha[t] = __shfl(hm[t], pa, WAVE);
            const int na = __shfl(n, pa, WAVE);
            const float ra = __shfl(ratio, pa, WAVE);
            const int c = part ? clean_shared(ha, hm) : 0;
            const int common = wave_scan(c);
            bool dem_j = false, dem_i = false;
            if (part) {
                if (ra < ratio) dem_j = (float)common > (float)na * 0.3333f;
                else dem_i = (float)common > (float)n * 0.3333f;
            }
            alive &= ~ballot(dem_j);
            if (ballot(dem_i)) alive &= ~(1ull << pa);
        };
        // within each slot
        for (int pa = 0; pa < nlive; ++pa) {
            if (!((alive >> pa) & 1ull)) continue;
            const int sa = __shfl(slot, pa, WAVE);
            step(pa, mine && lane > pa && slot == sa && ((alive >> lane) & 1ull));
        }
        // slot against slot: for s, for o != s, for i in s, for j in o
        unsigned slots = 0;
        {
            const unsigned long long m = ballot(mine);
            (void)m;
            unsigned v = mine ? (1u << slot) : 0u;
#pragma unroll
            for (int o = WAVE / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, WAVE);
            slots = v;
        }
        for (unsigned ss = slots; ss; ss &= ss - 1) {
            const int sidx = __ffs(ss) - 1;
            const unsigned long long in_s = ballot(mine && slot == sidx);
            for (unsigned oo = slots & ~(1u << sidx); oo; oo &= oo - 1) {
                const int oidx = __ffs(oo) - 1;
                for (unsigned long long pm = in_s; pm; pm &= pm - 1) {
                    const int pa = __ffsll((long long)pm) - 1;
                    if (!((alive >> pa) & 1ull)) continue;
                    step(pa, mine && slot == oidx && ((alive >> lane) & 1ull));
                }
            }
        }
        if (mine && !((alive >> lane) & 1ull)) tks[k].stid = 3;
        wsync();
        return;
    }
    if (nlive <= WAVE * CLEAN_WORDS) {
        clean_many<MAXH, WPE>(tks, slot_order, ntkl, nlive);
        return;
    }
    if (lane == 0) {
        auto shared_hits = [&](const Tk& A, int na, const Tk& B, int nb) {
            int c = 0;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j)
                    if (B.det[j] == A.det[i] && B.hraw[j] == A.hraw[i]) { c++; break; }
            return c;
        };
        auto pass = [&](int ib, int ie, int jb_self, int jb, int je) {
            for (int ii = ib; ii < ie; ++ii) {
                Tk& A = tks[slot_order[ii]];
                if (A.stid < 5) continue;
                const int ni = A.nx + A.nuv;
                int common = 0;
                for (int jj = (jb_self ? ii + 1 : jb); jj < je; ++jj) {
                    Tk& B = tks[slot_order[jj]];
                    if (B.stid < 5) continue;
                    const int nj = B.nx + B.nuv;
                    common += shared_hits(A, ni, B, nj);
                    if (A.chi2 / (float)(ni - 4) < B.chi2 / (float)(nj - 4)) {
                        if ((float)common > (float)ni * 0.3333f) B.stid = 3;
                    } else {
                        if ((float)common > (float)nj * 0.3333f) A.stid = 3;
                    }
                }
            }
        };
        for (int s = 0; s < 32; ++s) pass(L.sstart[s], L.sstart[s + 1], 1, 0, L.sstart[s + 1]);
        for (int s = 0; s < 32; ++s)
            for (int o = 0; o < 32; ++o)
                if (o != s) pass(L.sstart[s], L.sstart[s + 1], 0, L.sstart[o], L.sstart[o + 1]);
    }
    wsync();
}

// G: station-1 extension (reconstruction_kernels.cuh:1841-2534), one track at a
// time.
//  * The three station-1 window lists (X, U, V) of a batch of tracks are built
//    together, one lane per (track, view).
//  * Per x pair the x part (x at the current y0/ty, KMAG-point fit, left/right,
//    refit, 1/p window) is computed by every lane alike.
//  * Hits sit in fixed slots X (4), X' (3), U (2), U' (1), V (6), V' (5), the
//    reference's list order; see lr_slots_new.
//  * Two passes.  Pass A walks the x pairs in order with the lanes spread over
//    the (u, v) pairs: y sums, fit and first box cut only, listing the
//    survivors.  y0/ty after an x pair are those of the last (u, v) pair that
//    reached the y fit (the refit after left/right repeats the same fit); they
//    feed the next x pair's x-at-y, so pass A is sequential over x pairs.
//    Pass B evaluates the listed survivors, packed one per lane: left/right,
//    second box cut, residual chi2.  (A full list is evaluated before pass A
//    goes on.)
//  * Survivors are independent except for one case: a lone X' hit next to a U
//    hit forms a left/right pair whose sign writes carry to later (u, v) pairs
//    of the same x pair.  That write does not depend on the current sign, so
//    each lane records whether it writes (sentinel 2 in the sign slot) and a
//    scan over the list hands each lane the sign left by the last writer before
//    it within its x pair, else the x pair's starting sign.
//  * The best candidate is the minimum of (chi2, flat index ix*ncomb + iuv),
//    which is the reference's strict-< first minimum; its residuals are
//    recomputed from its parameters when the track is written.
struct GBest {
    float chi2, y0, ty, ey0, ety, invp, einvp, x01, tx1;
    int q;
    unsigned pres;
    int raw[6], sign[6];
};

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(v, o, WAVE);
        v = t < v ? t : v;
    }
    return v;
}

// window list of view v (0 X, 1 U, 2 V) of station 1 from the sagitta ratio
// (reconstruction_helper.cuh:556-613); -1 on overflow
template <int MAXH, int WPE>
__device__ __forceinline__ int global_window(const Ev<MAXH, WPE>& E, const Tk* T, int v, uint16_t* out, int plane_cap,
                                             int capw) {
    const int n23 = T->nx + T->nuv;
    const float x0 = T->x0, tx = T->tx, y0 = T->y0, ty = T->ty;
    int dlast = 1;
    float zmax = -10000.f;
    for (int m = 0; m < n23; ++m) {
        const int d = T->det[m];
        if (d > 0 && G(F_Z, d) > zmax) { dlast = d; zmax = G(F_Z, d); }
    }
    const int i = v == 0 ? 1 : (v == 1 ? 0 : 2);   // planetype[d/2-1] of d = 2i+2 is v
    const float z3 = G(F_Z, dlast), x3 = x0 + tx * z3, y3 = y0 + ty * z3;
    const int d = 2 * i + 2;
    const int d2 = super_id(2, v) * 2;
    const float ct = G(F_COSTHETA, d2), st = G(F_SINTHETA, d2);
    const float ps3 = x3 * ct + y3 * st;
    const float z1 = G(F_Z, d), z2 = G(F_Z, d2);
    const float x2 = x0 + tx * z2, y2 = y0 + ty * z2;
    const float ps2 = x2 * ct + y2 * st;
    const float s2t = ps2 - (ps3 * (z2 - Z_TARGET)) / (z3 - Z_TARGET);
    const float s2d = ps2 - (ps3 * (z2 - Z_DUMP)) / (z3 - Z_DUMP);
    const float pet = SAG_T_C * s2t + (ps3 * (z1 - Z_TARGET)) / (z3 - Z_TARGET);
    const float ped = SAG_D_C * s2d + (ps3 * (z1 - Z_DUMP)) / (z3 - Z_DUMP);
    const float wt = fabsf(s2t * SAG_T_W), wd = fabsf(s2d * SAG_D_W);
    const float pmn = fminf(pet - wt, ped - wd), pmx = fmaxf(pet + wt, ped + wd);
    const float pexp = (float)(0.5 * (double)(pmx + pmn));
    const float win = (float)(0.5 * (double)(pmx - pmn));
    return pairs_window(E, 0, v, pexp - win * 1.0f, pexp + win * 1.0f, out, plane_cap, capw);
}

// x part of station-1 X entry qx at the current y0/ty (every lane alike);
// false when the entry has no hit or fails the 1/p window
struct GX {
    float x01, tx1, invp;
    int q;
    unsigned xpres;
    int xsign[2], xraw[2];
    float xdr[2], xpos[2];
};

template <int MAXH, int WPE>
__device__ __forceinline__ bool global_xpart(const Ev<MAXH, WPE>& E, uint16_t qx, float x0, float tx, float ex0, float etx,
                                             float y0, float ty, GX& X) {
    constexpr int dx[2] = {4, 3};
    const int xi[2] = {pa(qx), pb(qx)};
    X.xpres = 0;
    float Xp[2] = {0.f, 0.f}, Zx[2] = {0.f, 0.f};
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        X.xraw[w] = 0; X.xsign[w] = 0; X.xdr[w] = 0.f; X.xpos[w] = 0.f;
        if (xi[w] < 0) continue;
        const int r = E.raw(dx[w], xi[w]);
        X.xpres |= 1u << w;
        X.xraw[w] = r; X.xdr[w] = E.Lp->drift[r]; X.xpos[w] = E.Lp->pos[r];
        xz_fy(dx[w], E.Lp->elem[r], X.xdr[w], 0, y0, ty, Xp[w], Zx[w]);
    }
    if (X.xpres == 0) return false;
    const float xk = x0 + tx * Z_KMAG_BEND, exk = ex0 + etx * Z_KMAG_BEND;
    auto xfit = [&](const float (&xv)[2]) {
        float A0 = 0.f, A1 = 0.f, A3 = 0.f, B0 = 0.f, B1 = 0.f;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            if (!((X.xpres >> w) & 1u)) continue;
            const float e = G(F_RESOLUTION, dx[w]);
            B0 += e * xv[w]; B1 += (e * xv[w]) * Zx[w]; A0 += e; A1 += e * Zx[w]; A3 += (e * Zx[w]) * Zx[w];
        }
        B0 += exk * xk; B1 += (exk * xk) * Z_KMAG_BEND; A0 += exk; A1 += exk * Z_KMAG_BEND;
        A3 += (exk * Z_KMAG_BEND) * Z_KMAG_BEND;
        return fit_finish(A0, A1, A3, B0, B1);
    };
    Fit f = xfit(Xp);
    if (X.xpres == 3u) {   // X / X' pair, x-only decision with the fit's parameters as errors
        const int h = lr_xpair_decide(f.p0, f.p1, f.p0, f.p1, 4, 3, Xp[0], Xp[1], X.xdr[0], X.xdr[1], 150.f);
        if (h >= 0) { X.xsign[0] = lr_si(h); X.xsign[1] = lr_sj(h); }
    }
#pragma unroll
    for (int w = 0; w < 2; ++w)
        if (((X.xpres >> w) & 1u) && X.xsign[w] == 0) X.xsign[w] = G(F_Z, dx[w]) * f.p1 + f.p0 > Xp[w] ? 1 : -1;
    float Xc[2];
#pragma unroll
    for (int w = 0; w < 2; ++w) Xc[w] = Xp[w] + (float)X.xsign[w] * X.xdr[w];
    f = xfit(Xc);
    X.x01 = f.p0; X.tx1 = f.p1;  // parameters used as errors (reconstruction_kernels.cuh:2194-2195)
    X.q = (double)(-0.0033f) * copysign(1.0, (double)FMAGSTR) * (double)x0 < (double)tx ? 1 : -1;
    X.invp = ((X.tx1 - tx) * (float)X.q) / PT_KICK_KMAG;
    const float einvp = (etx - etx) / PT_KICK_KMAG;
    return !(X.invp > INVP_MAX + einvp || X.invp < INVP_MIN - einvp);
}

// one accepted UV hit into the y-fit sums
__device__ __forceinline__ void y_acc(float& A0, float& A1, float& A3, float& B0, float& B1, float y, float ey,
                                      float z) {
    B0 += ey * y; B1 += (ey * y) * z; A0 += ey; A1 += ey * z; A3 += (ey * z) * z;
}

// y fit of (u, v) pair iuv of an x entry: the station-2/3 sums As, then the
// accepted U hits (a then b) and, when there is one, the accepted V hits.
// False when the pair does not reach the fit (no accepted U or V hit).  With
// SLOTS the hits go into slots 2-5.
template <bool SLOTS, int MAXH, int WPE>
__device__ __forceinline__ bool global_yfit(const Ev<MAXH, WPE>& E, const uint16_t* lu, const uint16_t* lv, int nu1, int iuv,
                                            const GX& X, const float (&As)[5], unsigned& pres, int (&raw)[6],
                                            float (&pos)[6], float (&drift)[6], Fit& g) {
    constexpr int det6[6] = {4, 3, 2, 1, 6, 5};
    float A0 = As[0], A1 = As[1], A3 = As[2], B0 = As[3], B1 = As[4];
    const int iu = iuv % nu1, iv = iuv / nu1;
    const uint16_t qu = lu[iu], qv = lv[iv];
    const int ui[4] = {pa(qu), pb(qu), pa(qv), pb(qv)};
    int nuh = 0, nvh = 0;
#pragma unroll
    for (int s = 2; s < 6; ++s) {
        const int idx = ui[s - 2];
        if (idx < 0 || (s >= 4 && nuh == 0)) continue;
        const int d = det6[s];
        const int r = E.raw(d, idx);
        float y, ey;
        if (y_uvhit(d, E.Lp->elem[r], E.Lp->drift[r], 0, X.x01, X.tx1, y, ey)) {
            if (SLOTS) {
                pres |= 1u << s;
                raw[s] = r; pos[s] = E.Lp->pos[r]; drift[s] = E.Lp->drift[r];
            }
            y_acc(A0, A1, A3, B0, B1, y, ey, G(F_Z, d));
            if (s < 4) ++nuh;
            else ++nvh;
        }
    }
    if (nuh == 0 || nvh == 0) return false;
    g = fit_finish(A0, A1, A3, B0, B1);
    return true;
}

// Pass-A tables of one x entry: per U entry the sums As with its accepted hits
// added (a then b) and their count; per V entry its accepted hits (y, ey) and
// which of a, b they are.  A (u, v) pair then adds the V hits to the U row,
// the same operations in the same order as global_yfit.
struct GTab {
    float4* u;      // [win_cap][2]: A0, A1, A3, B0 | B1, U hits
    float4* v;      // [win_cap]: y a, ey a, y b, ey b
    uint32_t* vf;   // [win_cap]: bit 0 hit a accepted, bit 1 hit b
};

template <int MAXH, int WPE>
__device__ __forceinline__ void global_tables(const Ev<MAXH, WPE>& E, const GTab& T, const uint16_t* lu,
                                              const uint16_t* lv, int nu1, int nv1, const GX& X,
                                              const float (&As)[5]) {
    constexpr int det6[6] = {4, 3, 2, 1, 6, 5};
    const int lane = lane_id();
    for (int iu = lane; iu < nu1; iu += WAVE) {
        float A0 = As[0], A1 = As[1], A3 = As[2], B0 = As[3], B1 = As[4];
        const uint16_t q = lu[iu];
        const int ui[2] = {pa(q), pb(q)};
        int nuh = 0;
#pragma unroll
        for (int s = 2; s < 4; ++s) {
            const int idx = ui[s - 2];
            if (idx < 0) continue;
            const int d = det6[s];
            const int r = E.raw(d, idx);
            float y, ey;
            if (y_uvhit(d, E.Lp->elem[r], E.Lp->drift[r], 0, X.x01, X.tx1, y, ey)) {
                y_acc(A0, A1, A3, B0, B1, y, ey, G(F_Z, d));
                ++nuh;
            }
        }
        T.u[2 * iu] = make_float4(A0, A1, A3, B0);
        T.u[2 * iu + 1] = make_float4(B1, __int_as_float(nuh), 0.f, 0.f);
    }
    for (int iv = lane; iv < nv1; iv += WAVE) {
        const uint16_t q = lv[iv];
        const int ui[2] = {pa(q), pb(q)};
        float yv[2] = {0.f, 0.f}, eyv[2] = {0.f, 0.f};
        uint32_t fl = 0;
#pragma unroll
        for (int s = 4; s < 6; ++s) {
            const int idx = ui[s - 4];
            if (idx < 0) continue;
            const int d = det6[s];
            const int r = E.raw(d, idx);
            if (y_uvhit(d, E.Lp->elem[r], E.Lp->drift[r], 0, X.x01, X.tx1, yv[s - 4], eyv[s - 4])) fl |= 1u << (s - 4);
        }
        T.v[iv] = make_float4(yv[0], eyv[0], yv[1], eyv[1]);
        T.vf[iv] = fl;
    }
}

// global_yfit<false> of pair iuv from the tables
__device__ __forceinline__ bool global_yfit_tab(const GTab& T, int nu1, int iuv, Fit& g) {
    constexpr int det6[6] = {4, 3, 2, 1, 6, 5};
    const int iu = iuv % nu1, iv = iuv / nu1;
    const float4 a = gld4(T.u + 2 * iu), b = gld4(T.u + 2 * iu + 1);
    if (__float_as_int(b.y) == 0) return false;
    float A0 = a.x, A1 = a.y, A3 = a.z, B0 = a.w, B1 = b.x;
    const uint32_t fl = gld(T.vf + iv);
    if (fl == 0) return false;
    const float4 v = gld4(T.v + iv);
    if (fl & 1u) y_acc(A0, A1, A3, B0, B1, v.x, v.y, G(F_Z, det6[4]));
    if (fl & 2u) y_acc(A0, A1, A3, B0, B1, v.z, v.w, G(F_Z, det6[5]));
    g = fit_finish(A0, A1, A3, B0, B1);
    return true;
}

constexpr int GLOBAL_LIST_CAP = 2048;   // pass-A survivors held before pass B runs (<= list_cap)

// One track: pass A over its x pairs, pass B over the listed survivors, then
// the best candidate is written.  Uses the x-part records (S.wy() + 8), the
// survivor list (S.plist()), the best-candidate record (S.wy()) and the
// pass-A tables (S.ycache()).
template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void global_track(Scr S, Tk* T, const uint16_t* lx, const uint16_t* lu,
                                                       const uint16_t* lv, int nx1, int nu1, int nv1) {
    T = uni(T);
    lx = uni(lx);
    lu = uni(lu);
    lv = uni(lv);
    nx1 = uni(nx1);
    nu1 = uni(nu1);
    nv1 = uni(nv1);
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    constexpr int det6[6] = {4, 3, 2, 1, 6, 5};
    GBest* best = reinterpret_cast<GBest*>(S.wy());
    GX* xrec = reinterpret_cast<GX*>(S.wy() + 8);
    uint32_t* list = S.plist();
    GTab tab;
    tab.u = reinterpret_cast<float4*>(S.ycache());
    tab.v = tab.u + 2 * cfg().win_cap;
    tab.vf = reinterpret_cast<uint32_t*>(tab.v + cfg().win_cap);
    const int n23 = T->nx + T->nuv;
    const float x0 = T->x0, tx = T->tx, ex0 = T->ex0, etx = T->etx;
    const float einvp = (etx - etx) / PT_KICK_KMAG;
    float y0 = T->y0, ty = T->ty;
    OR_SUB_START
    // y sums of the station-2/3 UV hits with their stored signs (reconstruction_kernels.cuh:2086-2102)
    float As[5] = {0.f, 0.f, 0.f, 0.f, 0.f};   // A0, A1, A3, B0, B1
    for (int m = 0; m < n23; ++m) {
        const int d = __builtin_amdgcn_readfirstlane(T->det[m]);
        if ((double)G(F_COSTHETA, d) > 0.99) continue;
        const int h = __builtin_amdgcn_readfirstlane(T->hraw[m]);
        float y, ey;
        if (y_uvhit(d, E.Lp->elem[h], E.Lp->drift[h], T->sign[m], x0, tx, y, ey)) {
            const float z = G(F_Z, d);
            As[3] += ey * y; As[4] += (ey * y) * z; As[0] += ey; As[1] += ey * z; As[2] += (ey * z) * z;
        }
    }
    const int ncomb = nu1 * nv1;
    unsigned long long best_key = ~0ull;
    int carry_ix = -1, carry_sign = 0;   // last lone-X' sign writer evaluated so far
    int ix = 0, base = 0;
    bool have_x = false;
    GX X;
    while (ix < nx1) {
        // ---- pass A: x pairs in order until done or the list is full
        int np = 0;
        while (ix < nx1) {
            if (!have_x) {
                if (!global_xpart(E, lx[ix], x0, tx, ex0, etx, y0, ty, X)) { ++ix; continue; }
                have_x = true;
                base = 0;
                if (lane == 0) xrec[ix] = X;
                global_tables(E, tab, lu, lv, nu1, nv1, X, As);
                wsync();
            }
            if (np + WAVE > GLOBAL_LIST_CAP) break;
            const int iuv = base + lane;
            bool reached = false, pass = false;
            Fit g = {0.f, 0.f, 0.f, 0.f};
            if (iuv < ncomb) {
                reached = global_yfit_tab(tab, nu1, iuv, g);
                pass = reached && !(fabsf(g.p0) > Y0_MAX + 2.f * g.e0 || fabsf(g.p1) > TY_MAX + 2.f * g.e1);
            }
            const unsigned long long pm = ballot(pass);
            if (pass) list[np + rank_in(pm)] = ((uint32_t)ix << 16) | (uint32_t)iuv;
            np += __popcll(pm);
            // y0/ty for the next x pair: the last (u, v) pair that reached the y fit
            const unsigned long long rm = ballot(reached);
            if (rm) {
                const int src = 63 - __clzll((long long)rm);
                y0 = __shfl(g.p0, src, WAVE);
                ty = __shfl(g.p1, src, WAVE);
            }
            base += WAVE;
            if (base >= ncomb) { have_x = false; ++ix; }
        }
        wsync();
        OR_SUB(7);
        // ---- pass B: the listed survivors, one per lane, in list order
        for (int k0 = 0; k0 < np; k0 += WAVE) {
            const int k = k0 + lane;
            int kix = -1, sign[6] = {0, 0, 0, 0, 0, 0}, raw[6] = {0, 0, 0, 0, 0, 0};
            float pos[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, drift[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            bool cand = false, wrote = false, cross = false;
            float fy0 = 0.f, fty = 0.f, fey0 = 0.f, fety = 0.f, chi23 = 0.f;
            unsigned pres = 0;
            GX Y;
            if (k < np) {
                const uint32_t code = list[k];
                kix = (int)(code >> 16);
                const int iuv = (int)(code & 0xFFFFu);
                Y = xrec[kix];
                cross = Y.xpres == 2u;   // lone X'
                pres = Y.xpres;
                sign[0] = Y.xsign[0]; raw[0] = Y.xraw[0]; pos[0] = Y.xpos[0]; drift[0] = Y.xdr[0];
                sign[1] = cross ? 2 : Y.xsign[1]; raw[1] = Y.xraw[1]; pos[1] = Y.xpos[1]; drift[1] = Y.xdr[1];
                Fit g;
                global_yfit<true>(E, lu, lv, nu1, iuv, Y, As, pres, raw, pos, drift, g);
                fy0 = g.p0; fty = g.p1; fey0 = g.e0; fety = g.e1;
                int dets[6] = {4, 3, 2, 1, 6, 5};
                lr_slots_new<6>(Y.x01, Y.tx1, fy0, fty, Y.x01, Y.tx1, fey0, fety, pres, dets, pos, drift, sign, 150.f);
                if (cross) wrote = sign[1] != 2;
                if (!(fabsf(fy0) > Y0_MAX || fabsf(fty) > TY_MAX)) {
                    // station-2/3 residuals with (x0, tx) of the tracklet and the new y line;
                    // the track is the wave's own, so its hits are wave-uniform and the
                    // geometry of their planes comes through scalar loads
                    for (int m = 0; m < n23; ++m) {
                        const int d = __builtin_amdgcn_readfirstlane(T->det[m]);
                        const int h = __builtin_amdgcn_readfirstlane(T->hraw[m]);
                        const float res = G(F_RESOLUTION, d);
                        const float dca = dca_wire(d, E.Lp->elem[h], x0, fy0, tx, fty);
                        const float r23 = E.Lp->drift[h] * (float)T->sign[m] - dca;
                        chi23 += (r23 * r23) / (res * res);
                    }
                    cand = true;
                }
            }
            // sign of a lone X' hit: the last writer at or before this lane in the
            // same x pair, else the x pair's starting sign
            const unsigned long long wm = ballot(wrote);
            {
                const unsigned long long le = wm & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
                const int src = le ? 63 - __clzll((long long)le) : 0;
                const int sv = __shfl(sign[1], src, WAVE), six = __shfl(kix, src, WAVE);
                if (cross) {
                    if (le) sign[1] = six == kix ? sv : Y.xsign[1];
                    else sign[1] = carry_ix == kix ? carry_sign : Y.xsign[1];
                }
                if (wm) {
                    const int last = 63 - __clzll((long long)wm);
                    carry_ix = __shfl(kix, last, WAVE);
                    carry_sign = __shfl(sign[1], last, WAVE);
                }
            }
            unsigned long long key = ~0ull;
            if (cand) {
                float c1 = 0.f;
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    if (!((pres >> s) & 1u)) continue;
                    const int d = det6[s];
                    const float res = G(F_RESOLUTION, d);
                    const float r1 =
                        drift[s] * (float)sign[s] - dca_wire(d, E.Lp->elem[raw[s]], Y.x01, fy0, Y.tx1, fty);
                    c1 += ((r1 * r1) / res) / res;
                }
                const float chi2 = chi23 + c1;
                if (!(chi2 > 10000.f) && chi2 < 10000.1f)
                    key = ((unsigned long long)__float_as_uint(chi2) << 32) |
                          (uint32_t)(kix * ncomb + (int)(list[k] & 0xFFFFu));
            }
            const unsigned long long cmin = wave_min_u64(key);
            if (cmin < best_key) {
                if (key == cmin) {
                    GBest& B = *best;
                    B.chi2 = __uint_as_float((uint32_t)(key >> 32));
                    B.y0 = fy0; B.ty = fty; B.ey0 = fey0; B.ety = fety; B.invp = Y.invp; B.einvp = einvp;
                    B.x01 = Y.x01; B.tx1 = Y.tx1; B.q = Y.q; B.pres = pres;
#pragma unroll
                    for (int s = 0; s < 6; ++s) { B.raw[s] = raw[s]; B.sign[s] = sign[s]; }
                }
                best_key = cmin;
            }
        }
        wsync();
        OR_SUB(8);
    }
    if (best_key != ~0ull && lane == 0) {
        const GBest& B = *best;
        T->stid = 6;
        T->chi2 = B.chi2; T->y0 = B.y0; T->ty = B.ty; T->ey0 = B.ey0; T->ety = B.ety;
        T->invp = B.invp; T->einvp = B.einvp; T->charge = (int8_t)B.q;
        for (int m = 0; m < n23; ++m) {
            const int d = T->det[m], r = T->hraw[m];
            T->resid[m] = E.Lp->drift[r] * (float)T->sign[m] - dca_wire(d, E.Lp->elem[r], x0, B.y0, tx, B.ty);
        }
        int n1 = 0;
        for (int s = 0; s < 6; ++s) {
            if (!((B.pres >> s) & 1u)) continue;
            const int d = det6[s], r = B.raw[s];
            T->resid[n23 + n1] = E.Lp->drift[r] * (float)B.sign[s] - dca_wire(d, E.Lp->elem[r], B.x01, B.y0, B.tx1, B.ty);
            T->hraw[n23 + n1] = (uint16_t)r;
            T->det[n23 + n1] = (uint8_t)d;
            T->sign[n23 + n1] = (int8_t)B.sign[s];
            ++n1;
        }
        T->n1 = (uint8_t)n1;
    }
    wsync();
}

// The window lists depend only on the tracklet, and one overflow rejects the
// event, so the lists of up to GLOBAL_BATCH tracklets (three each) are built
// together, lane (tracklet, view); the tracklets are then extended one after
// the other, each independent of the others.
constexpr int GLOBAL_BATCH = WAVE / 3;
// tracklets handed to a wave at a time: small blocks keep the event's waves
// evenly busy (only the tracklets at stationID >= 5 of a block take work)
constexpr int GLOBAL_BLOCK = 16;
static_assert(GLOBAL_BLOCK <= GLOBAL_BATCH, "a block's window lists are built in one batch");

template <int MAXH, int WPE>
STAGE void stage_global(or_params P, Scr S, int ntkl) {
    ntkl = uni(ntkl);
    Lds<MAXH>& L = g_lds<MAXH>;
    YzWave& Y = yzw<MAXH>();
    Ev<MAXH, WPE> E;
    const int lane = lane_id();
    Tk* tks = S.tks();
    int* nlf = &Y.nl[0][0];
    const int capw = cfg().win_cap;
    // blocks of GLOBAL_BLOCK tracklets, handed to the event's waves as they come free
    for (;;) {
        int k0 = 0;
        if (lane == 0) k0 = atomicAdd(&L.gl_next, GLOBAL_BLOCK);
        k0 = __builtin_amdgcn_readfirstlane(__shfl(k0, 0, WAVE));
        if (k0 >= ntkl) break;
        unsigned long long rest = ballot(lane < GLOBAL_BLOCK && k0 + lane < ntkl && tks[k0 + lane].stid >= 5);
        while (rest) {
            OR_SUB_START
            {
                const int b = lane / 3, v = lane - 3 * b;
                unsigned long long mm = rest;
                for (int t = 0; t < b && mm; ++t) mm &= mm - 1;
                if (b < GLOBAL_BATCH && mm) {
                    const int k = k0 + __ffsll((long long)mm) - 1;
                    const int nl = global_window(E, &tks[k], v, S.wl() + lane * cfg().win_cap, P.max_pair_plane_hits, capw);
                    if (nl < 0) atomicOr(&L.status, (uint32_t)OR_EV_PAIR_OVERFLOW);
                    nlf[lane] = nl;
                }
            }
            wsync();
            OR_SUB(6);
            if (__atomic_load_n(&L.status, __ATOMIC_RELAXED)) return;   // set by any wave: the event is rejected
            for (int b = 0; b < GLOBAL_BATCH && rest; ++b) {
                const int k = k0 + __ffsll((long long)rest) - 1;
                rest &= rest - 1;
                const int nx1 = nlf[3 * b], nu1 = nlf[3 * b + 1], nv1 = nlf[3 * b + 2];
                if (nx1 == 0 || nu1 == 0 || nv1 == 0) continue;
                global_track<MAXH, WPE>(S, &tks[k], S.wl() + (3 * b) * cfg().win_cap, S.wl() + (3 * b + 1) * cfg().win_cap,
                                   S.wl() + (3 * b + 2) * cfg().win_cap, nx1, nu1, nv1);
            }
        }
    }
}

// G': vertexing (gKernel_Vertexing, reconstruction_kernels.cuh:2547-2810), one lane
// per global track.  The track is swum upstream from its most upstream hit
// through FMAG (100 steps: pT kick, energy loss outside the beam hole) and the
// target region (100 straight steps) in double precision; the trajectory point
// closest to the beam line in x-y (first minimum, points with the x momentum on
// the wrong side skipped) is the vertex.  The minimum is taken on the fly in the
// reference's point order, so no trajectory arrays are kept.  Float and double
// are mixed exactly as in the reference (sqrtf of a double argument rounds it
// to float first).  Constants: reconstruction_constants.h:75-121 (E1039).
constexpr double V_FMAG_LENGTH = 502.92, V_FMAG_HOLE_LENGTH = 27.94, V_FMAG_HOLE_RADIUS = 1.27;
constexpr int V_NSTEPS_FMAG = 100, V_NSTEPS_TARGET = 100;
constexpr double V_STEP_FMAG = 2.514600, V_STEP_TARGET = 5.0, V_PTKICK_UNIT = -0.006096568;
constexpr double V_DEDX0 = 0.014282073, V_DEDX1 = 7.1869681e-05, V_DEDX2 = -1.4279150e-06,
                 V_DEDX3 = 1.5853655e-08, V_DEDX4 = -6.0741470e-11;
constexpr float V_X_BEAM = 0.f, V_Y_BEAM = 0.f;

__device__ __forceinline__ double dedx_poly(double p) {
    return V_DEDX0 + V_DEDX1 * p + V_DEDX2 * p * p + V_DEDX3 * p * p * p + V_DEDX4 * p * p * p * p;
}

__device__ void vertex_track(Tk& T) {
    const float x0 = T.x0, y0 = T.y0, tx = T.tx, ty = T.ty, invP = T.invp, charge = (float)T.charge;
    const int nh = T.nx + T.nuv + T.n1;
    // get_firsthitdetid (reconstruction_classes.cuh:517-531)
    int det_first = 30;
    float zmin = 10000.f;
    for (int i = 0; i < nh; ++i) {
        const int d = T.det[i];
        if (d > 0 && G(F_Z, d) < zmin) { det_first = d; zmin = G(F_Z, d); }
    }
    // calculate_x0_tx_st1 (reconstruction_helper.cuh:622-626), charge as short
    const short q = (short)charge;
    const float tx_st1 = tx + PT_KICK_KMAG * invP * q;
    const float x0_st1 = tx * Z_KMAG_BEND + x0 - tx_st1 * Z_KMAG_BEND;
    const double z0 = G(F_Z, det_first);
    const double s0 = charge * invP * sqrtf((1.f + tx_st1 * tx_st1) / (1.f + tx_st1 * tx_st1 + ty * ty));
    const double s1 = tx_st1, s2 = ty;
    const double s3 = (float)z0 * tx_st1 + x0_st1;   // x_trk / y_trk take float z
    const double s4 = (float)z0 * ty + y0;
    double px = s3 + s1 * (V_FMAG_LENGTH - z0), py = s4 + s2 * (V_FMAG_LENGTH - z0), pz = V_FMAG_LENGTH;
    const double pz0 = 1.f / (fabs(s0) * sqrtf((float)(1.f + s1 * s1 + s2 * s2)));
    double mx = pz0 * s1, my = pz0 * s2, mz = pz0;
    double dmin = 1.e18;
    bool found = false;
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    auto consider = [&]() {
        if (FMAGSTR * charge * mx < 0.) return;
        const double d2 = (px - V_X_BEAM) * (px - V_X_BEAM) + (py - V_Y_BEAM) * (py - V_Y_BEAM);
        if (d2 < dmin) { dmin = d2; found = true; v0 = px; v1 = py; v2 = pz; v3 = mx; v4 = my; v5 = mz; }
    };
    consider();
    for (int step = 1; step <= V_NSTEPS_FMAG; ++step) {
        const double tx_i = mx / mz;
        const double tx_f = tx_i + 2.f * charge * V_PTKICK_UNIT * V_STEP_FMAG / sqrt(mx * mx + mz * mz);
        const double t1x = tx_i * V_STEP_FMAG, t1y = ty * V_STEP_FMAG, t1z = V_STEP_FMAG;
        const double bx = px - t1x, by = py - t1y, bz = pz - t1z;
        const double ptot_i = sqrt(mx * mx + my * my + mz * mz);
        double ptot_b = ptot_i;
        if (bz > V_FMAG_HOLE_LENGTH || bx * bx + by * by > V_FMAG_HOLE_RADIUS)
            ptot_b += dedx_poly(ptot_i) * sqrtf((float)(t1x * t1x + t1y * t1y + t1z * t1z));
        const double t2x = tx_f * V_STEP_FMAG, t2y = ty * V_STEP_FMAG, t2z = V_STEP_FMAG;
        px = bx - t2x; py = by - t2y; pz = bz - t2z;
        double ptot_f = ptot_b;
        if (pz > V_FMAG_HOLE_LENGTH || px * px + py * py > V_FMAG_HOLE_RADIUS)
            ptot_f += dedx_poly(ptot_b) * sqrtf((float)(t2x * t2x + t2y * t2y + t2z * t2z));
        const double pz_f = ptot_f / sqrt(1.f + tx_f * tx_f + ty * ty);
        mx = pz_f * tx_f; my = pz_f * ty; mz = pz_f;
        consider();
    }
    for (int step = 1; step <= V_NSTEPS_TARGET; ++step) {
        const double tx_i = mx / mz;
        px -= tx_i * V_STEP_TARGET; py -= ty * V_STEP_TARGET; pz -= V_STEP_TARGET;
        consider();
    }
    if (!found) { v0 = px; v1 = py; v2 = pz; v3 = mx; v4 = my; v5 = mz; }
    T.stid = 7;
    T.vtx[0] = (float)v0; T.vtx[1] = (float)v1; T.vtx[2] = (float)v2;
    T.vtx[3] = (float)v3; T.vtx[4] = (float)v4; T.vtx[5] = (float)v5;
}

// G'': prop-tube segment matching (reconstruction_helper.cuh:1916-2257) and global
// track cleaning (:2265-2354), the post-global steps of the full chain.
//  * Segment test: a hit of P1 (either plane) within win1 of the track and a hit
//    of P2 (either plane) within win2 whose line passes the slope, intercept and
//    z = 2028.19 cuts; any passing combination will do (the reference's loop
//    nesting only orders the tries).  Every tracklet is tested, one per lane;
//    one without an x or a y segment gets stationID 5.
//  * Cleaning: slot by slot, then slot against slot in slot order (the
//    reference runs the cross-slot pass concurrently), one lane; chi2/ndf =
//    chisq / (nhits - 5), merge threshold 0.015 on the signed relative 1/p
//    difference, same-sign charges; the within-slot else branch compares 1/p(i)
//    with itself, as the reference does.
template <int MAXH, int WPE>
__device__ bool prop_segment(int p1a, int p1b, int p2a, int p2b, float slope, float icpt, float win1, float win2,
                             float smax, float imax) {
    Ev<MAXH, WPE> E;
    const float cut = 0.03f;
    const int first[2] = {p1a, p1b}, second[2] = {p2a, p2b};
    for (int f = 0; f < 2; ++f) {
        const int da = first[f];
        const float za = G(F_Z, da);
        for (int n = 0; n < E.nred(da); ++n) {
            const float pos0 = E.Lp->pos[E.raw(da, n)];
            if (!(fabsf(pos0 - (slope * za + icpt)) < win1)) continue;
            for (int s2 = 0; s2 < 2; ++s2) {
                const int db = second[s2];
                const float zb = G(F_Z, db);
                for (int m = 0; m < E.nred(db); ++m) {
                    const float pos1 = E.Lp->pos[E.raw(db, m)];
                    if (fabsf(pos1 - (slope * zb + icpt)) > win2) continue;
                    const float a = (pos1 - pos0) / (zb - za);
                    if (fabsf(a) > smax) continue;
                    if (fabsf(a - slope) > cut) continue;
                    const float b = pos0 - a * za;
                    if (fabsf(b) > imax) continue;
                    if (fabsf((slope * 2028.19f + icpt) - (a * 2028.19f + b)) > 3.0f) continue;
                    return true;
                }
            }
        }
    }
    return false;
}

template <int MAXH, int WPE>
STAGE void stage_postglobal(Tk* tks, const uint16_t* slot_order, int ntkl) {
    tks = uni(tks);
    slot_order = uni(slot_order);
    ntkl = uni(ntkl);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    for (int k = lane; k < ntkl; k += WAVE) {
        Tk& T = tks[k];
        if (!prop_segment<MAXH, WPE>(50, 49, 52, 51, T.tx, T.x0, 5.08f, 7.62f, TX_MAX, X0_MAX) ||
            !prop_segment<MAXH, WPE>(48, 47, 54, 53, T.ty, T.y0, 5.08f, 10.16f, TY_MAX, Y0_MAX))
            T.stid = 5;
    }
    wsync();
    if (lane == 0) {
        const float thr = 0.015f;
        auto ndf = [](const Tk& T) { return T.chi2 / ((float)(T.nx + T.nuv + T.n1) - 5.f); };
        for (int s = 0; s < 32; ++s)
            for (int ii = L.sstart[s]; ii < L.sstart[s + 1]; ++ii) {
                Tk& A = tks[slot_order[ii]];
                if (A.stid < 6) continue;
                if (ndf(A) > 250.f) { A.stid = 5; continue; }
                for (int jj = ii + 1; jj < L.sstart[s + 1]; ++jj) {
                    Tk& B = tks[slot_order[jj]];
                    if (B.stid < 6) continue;
                    if (ndf(A) < ndf(B)) {
                        if ((A.invp - B.invp) / A.invp < thr && (float)A.charge * (float)B.charge > 0) B.stid = 5;
                    } else {
                        if ((A.invp - A.invp) / B.invp < thr && (float)A.charge * (float)B.charge > 0) A.stid = 5;
                    }
                }
            }
        for (int s = 0; s < 32; ++s)
            for (int o = 0; o < 32; ++o) {
                if (o == s) continue;
                for (int ii = L.sstart[s]; ii < L.sstart[s + 1]; ++ii) {
                    Tk& A = tks[slot_order[ii]];
                    if (A.stid < 6) continue;
                    for (int jj = L.sstart[o]; jj < L.sstart[o + 1]; ++jj) {
                        Tk& B = tks[slot_order[jj]];
                        if (B.stid < 6) continue;
                        if (ndf(A) < ndf(B)) {
                            if ((A.invp - B.invp) / A.invp < thr && (float)A.charge * (float)B.charge > 0) B.stid = 5;
                        } else {
                            if ((A.invp - B.invp) / B.invp < thr && (float)A.charge * (float)B.charge > 0) A.stid = 5;
                        }
                    }
                }
            }
    }
    wsync();
}

template <int MAXH, int WPE>
STAGE void stage_vertex(Tk* tks, int ntkl) {
    tks = uni(tks);
    ntkl = uni(ntkl);
    for (int k = lane_id(); k < ntkl; k += WAVE)
        if (tks[k].stid == 6) vertex_track(tks[k]);
    wsync();
}

// H: outputs
// ---- dimuons and display histograms (gKernel_DimuonBuilding, reconstruction_kernels.cuh:2820-3090;
// FindDimuonClosestApproach, reconstruction_helper.cuh:1269-1476; gKernel_fill_display_histograms,
// reconstruction_kernels.cuh:3100-3195) ----
// One lane per event over the chain's output tracks (stationID >= 7, slot order).  Pairs are taken in the
// reference's order (slot a <= slot b, then the tracks of a, then those of b) because momentum values the
// swim leaves in the pair's 4-vectors carry over to the next partner of the same first track.  The swim
// of the two tracks runs in lock step, keeping the closest-approach step on the fly instead of 603-float
// arrays; every statement rounds where the reference's float arrays and double constants make it round.
constexpr float D_MMU = 0.105658f, D_MP = 0.938272f, D_EBEAM = 120.f, D_PBEAM = 119.996331801f;
constexpr float D_Z_DUMP = 42.f, D_CHI2DOF_MAX = 250.f;
constexpr int D_NS = V_NSTEPS_FMAG + V_NSTEPS_TARGET;

struct DimSwim {
    float px, py, pz, mx, my, mz;
};

__device__ __forceinline__ float dim_e(const float* p) { return sqrtf(D_MMU * D_MMU + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// swim start at the FMAG end (reconstruction_kernels.cuh:2953-2977)
__device__ DimSwim dim_start(const or_track& R, int det_first) {
    const short q = (short)R.charge;
    const float charge = (float)R.charge, invP = R.invp, ty = R.ty;
    const float tx_st1 = R.tx + PT_KICK_KMAG * invP * q;
    const float x0_st1 = R.tx * Z_KMAG_BEND + R.x0 - tx_st1 * Z_KMAG_BEND;
    const float z_0 = G(F_Z, det_first);
    const float s0 = charge * invP * sqrtf((1.f + tx_st1 * tx_st1) / (1.f + tx_st1 * tx_st1 + ty * ty));
    const float s3 = z_0 * tx_st1 + x0_st1, s4 = z_0 * ty + R.y0;
    DimSwim w;
    w.px = (float)(s3 + tx_st1 * (V_FMAG_LENGTH - z_0));
    w.py = (float)(s4 + ty * (V_FMAG_LENGTH - z_0));
    w.pz = (float)V_FMAG_LENGTH;
    const float pz_0 = 1.f / (fabsf(s0) * sqrtf(1.f + tx_st1 * tx_st1 + ty * ty));
    w.mx = pz_0 * tx_st1; w.my = pz_0 * ty; w.mz = pz_0;
    return w;
}

// one FMAG step; lin = linear coefficient of the first energy-loss term
// (DEDX_UNIT_2 for the second track of a pair, as the reference has it)
__device__ __forceinline__ void dim_fmag_step(DimSwim& w, float charge, float ty, double lin) {
    const float tx_i = w.mx / w.mz;
    const float tx_f = tx_i + 2.f * charge * V_PTKICK_UNIT * V_STEP_FMAG / sqrtf(w.mx * w.mx + w.mz * w.mz);
    const float t1x = (float)(tx_i * V_STEP_FMAG), t1y = (float)(ty * V_STEP_FMAG), t1z = (float)V_STEP_FMAG;
    const float bx = w.px - t1x, by = w.py - t1y, bz = w.pz - t1z;
    const float ptot_i = sqrtf(w.mx * w.mx + w.my * w.my + w.mz * w.mz);
    float ptot_b = ptot_i;
    if (bz > V_FMAG_HOLE_LENGTH || bx * bx + by * by > V_FMAG_HOLE_RADIUS)
        ptot_b += (V_DEDX0 + lin * ptot_i + V_DEDX2 * ptot_i * ptot_i + V_DEDX3 * ptot_i * ptot_i * ptot_i +
                   V_DEDX4 * ptot_i * ptot_i * ptot_i * ptot_i) * sqrtf(t1x * t1x + t1y * t1y + t1z * t1z);
    const float t2x = (float)(tx_f * V_STEP_FMAG), t2y = (float)(ty * V_STEP_FMAG), t2z = (float)V_STEP_FMAG;
    w.px = bx - t2x; w.py = by - t2y; w.pz = bz - t2z;
    float ptot_f = ptot_b;
    if (w.pz > V_FMAG_HOLE_LENGTH || w.px * w.px + w.py * w.py > V_FMAG_HOLE_RADIUS)
        ptot_f += (V_DEDX0 + V_DEDX1 * ptot_b + V_DEDX2 * ptot_b * ptot_b + V_DEDX3 * ptot_b * ptot_b * ptot_b +
                   V_DEDX4 * ptot_b * ptot_b * ptot_b * ptot_b) * sqrtf(t2x * t2x + t2y * t2y + t2z * t2z);
    const float pz_f = ptot_f / sqrtf(1.f + tx_f * tx_f + ty * ty);
    w.mx = pz_f * tx_f; w.my = pz_f * ty; w.mz = pz_f;
}

__device__ __forceinline__ void dim_target_step(DimSwim& w, float ty) {
    const float tx_i = w.mx / w.mz;
    w.px = w.px - (float)(tx_i * V_STEP_TARGET);
    w.py = w.py - (float)(ty * V_STEP_TARGET);
    w.pz = w.pz - (float)V_STEP_TARGET;
}

// FindDimuonClosestApproach: v, returns the first track's z at the closest step
// (Z_DUMP when no step qualifies); writes p_pos / p_neg xyz at that step.
__device__ float dim_closest(float* v, float* p_pos, float* p_neg, float c1, float c2, float ty1, float ty2, DimSwim a,
                             DimSwim b) {
    float dmin = 1.e18f;
    int best = -1;
    DimSwim ba = a, bb = b;
    float z200a = 0.f, z200b = 0.f;    // flat element NSTEPS_FMAG + NSTEPS_TARGET: z of step 66
    auto consider = [&](int j) {
        if (j == D_NS / 3) { z200a = a.pz; z200b = b.pz; }
        if (FMAGSTR * c1 * a.mx < 0. && FMAGSTR * c2 * b.mx < 0.) return;
        const float d2 = (a.px - b.px) * (a.px - b.px) + (a.py - b.py) * (a.py - b.py);
        if (d2 < dmin) { dmin = d2; best = j; ba = a; bb = b; }
    };
    static_assert(D_NS % 3 == 2, "flat element 200 is a z coordinate");
    consider(0);
    int step = 1;
    for (; step <= V_NSTEPS_FMAG; ++step) {
        dim_fmag_step(a, c1, ty1, V_DEDX1);
        dim_fmag_step(b, c2, ty2, V_DEDX2);
        consider(step);
    }
    for (; step <= D_NS; ++step) {
        dim_target_step(a, ty1);
        dim_target_step(b, ty2);
        consider(step);
    }
    if (best < 0) {
        const float m = (float)((z200a + z200b) * 0.5);
        v[0] = m; v[1] = m; v[2] = m;
        return D_Z_DUMP;
    }
    const DimSwim& P = c1 > 0 ? ba : bb;
    const DimSwim& N = c1 > 0 ? bb : ba;
    p_pos[0] = P.mx; p_pos[1] = P.my; p_pos[2] = P.mz;
    p_neg[0] = N.mx; p_neg[1] = N.my; p_neg[2] = N.mz;
    v[0] = (float)((ba.px + bb.px) * 0.5);
    v[1] = (float)((ba.py + bb.py) * 0.5);
    v[2] = (float)((ba.pz + bb.pz) * 0.5);
    return ba.pz;
}

// display histograms: 11 triggers x 18 variables, bins of (online_reconstruction.cu:670-726) with
// reconstruction_constants.h:225-263; only bins of the variables' ranges are kept, packed per trigger
constexpr int H_NTRG = 11, H_NVARS = 18, H_NVARS_TRK = 6;
__constant__ short h_nb[H_NVARS] = {100, 100, 100, 100, 100, 200, 50, 50, 50, 50, 50, 50, 100, 100, 100, 100, 100, 200};
__constant__ float h_min[H_NVARS] = {-20.0f, -20.0f, 0.0f, -20.0f, -20.0f, -500.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
                                     -3.141592654f, -20.0f, -20.0f, 0.0f, -20.0f, -20.0f, -500.0f};
__constant__ float h_max[H_NVARS] = {20.0f, 20.0f, 120.0f, 20.0f, 20.0f, 500.0f, 10.0f, 1.0f, 1.0f, 0.5f, 5.0f,
                                     3.141592654f, 20.0f, 20.0f, 120.0f, 20.0f, 20.0f, 500.0f};
__constant__ short h_off[H_NVARS + 1] = {0, 100, 200, 300, 400, 500, 700, 750, 800, 850, 900, 950, 1000, 1100,
                                         1200, 1300, 1400, 1500, 1700};
constexpr int H_PACKED = 1700;     // packed bins per trigger

// Every bin j with x_j - hw <= v < x_j + hw (x_j = min + hw (2j + 1), hw = (max - min) / 2 / nbins, in
// float as the reference computes them) gets one count.  Only bins within two of the estimate can
// qualify, so those are tested with the reference's predicate.
__device__ __forceinline__ void hist_fill(uint32_t* lh, int trg, int k, float v) {
    if (!(fabsf(v) < INFINITY)) return;
    const int nb = h_nb[k];
    const float hw = (h_max[k] - h_min[k]) * 0.5f / nb;
    float lf = (v - h_min[k]) / (2.f * hw);
    lf = fminf(fmaxf(lf, -8.f), (float)nb + 8.f);
    const int l0 = (int)floorf(lf);
    for (int j = max(0, l0 - 2); j <= min(nb - 1, l0 + 2); ++j) {
        const float x = h_min[k] + hw * (2.f * j + 1.f);
        if (x - hw <= v && v < x + hw) atomicAdd(&lh[trg * H_PACKED + h_off[k] + j], 1u);
    }
}

struct DimArgs {
    int64_t n_events;
    int max_out;
    const int32_t* n_tracks;
    const uint32_t* status;
    const or_track* tracks;
    const float* vertex;
    const int16_t* det;
    const int64_t* off;
    const int32_t* trigger;
    int32_t* n_dim;
    float* dim;
    float* hist;                 // OR_HIST_* layout, += counts (NULL: no histograms)
    unsigned long long* lost;
};

constexpr int DIM_BLOCK = 256, DIM_MAX_TRACKS = 64;

__global__ __launch_bounds__(DIM_BLOCK) void dimuon_kernel(DimArgs a) {
    __shared__ uint32_t lh[H_NTRG * H_PACKED];
    if (a.hist)
        for (int i = threadIdx.x; i < H_NTRG * H_PACKED; i += DIM_BLOCK) lh[i] = 0;
    __syncthreads();
    const float p_beam[4] = {0, 0, D_PBEAM, D_EBEAM};
    const float p_target[4] = {0, 0, 0, D_MP};
    const float p_cms[4] = {0, 0, D_PBEAM, D_EBEAM + D_MP};
    const float s = p_cms[3] * p_cms[3] - (p_cms[0] * p_cms[0] + p_cms[1] * p_cms[1] + p_cms[2] * p_cms[2]);
    const float sqrt_s = sqrtf(s);
    const float bz = (p_cms[2] / p_cms[3]) * -1;
    const float t_cms = p_target[3] * p_cms[3] - (p_target[0] * p_cms[0] + p_target[1] * p_cms[1] + p_target[2] * p_cms[2]);
    const float b_cms = p_beam[3] * p_cms[3] - (p_beam[0] * p_cms[0] + p_beam[1] * p_cms[1] + p_beam[2] * p_cms[2]);
    for (int64_t e = (int64_t)blockIdx.x * DIM_BLOCK + threadIdx.x; e < a.n_events; e += (int64_t)gridDim.x * DIM_BLOCK) {
        int ndim = 0;
        const int nt = (a.status[e] & OR_EV_REJECT_MASK) ? 0 : min(min(a.n_tracks[e], a.max_out), DIM_MAX_TRACKS);
        const or_track* TR = a.tracks + e * a.max_out;
        const float* VX = a.vertex + e * a.max_out * 6;
        // the event's tracks with stationID >= 7 (slot order), first-hit detectors (get_firsthitdetid,
        // reconstruction_classes.cuh:517-532), slot membership as a bit mask
        uint64_t use = 0;
        unsigned slots = 0;
        int dfirst[DIM_MAX_TRACKS];
        for (int t = 0; t < nt; ++t) {
            const or_track& R = TR[t];
            if (R.station_id < 7) continue;
            use |= 1ull << t;
            slots |= 1u << (R.slot & 31);
            int dmin = 30;
            float zmin = 10000.f;
            const int nh = min((int)R.nhits, OR_MAX_HITS_PER_TRACK);
            for (int h = 0; h < nh; ++h) {
                const int d = a.det[a.off[e] + R.hit_index[h]];
                if (d > 0 && d < ND && G(F_Z, d) < zmin) { dmin = d; zmin = G(F_Z, d); }
            }
            dfirst[t] = dmin;
        }
        float p_pos[4] = {0, 0, 0, 0}, p_neg[4] = {0, 0, 0, 0}, pos_v[3] = {0, 0, 0}, neg_v[3] = {0, 0, 0};
        float* out = a.dim + e * OR_DIMUONS_MAX * OR_DIMUON_FIELDS;
        for (unsigned sa = slots; sa; sa &= sa - 1) {
            const int ia = __ffs(sa) - 1;
            for (unsigned sb = slots & ~((1u << ia) - 1u); sb; sb &= sb - 1) {
                const int ib = __ffs(sb) - 1;
                for (int i = 0; i < nt; ++i) {
                    if (!((use >> i) & 1) || TR[i].slot != ia) continue;
                    const or_track& A = TR[i];
                    const float ca = (float)A.charge;
                    if (ca > 0) {
                        p_pos[0] = VX[i * 6 + 3]; p_pos[1] = VX[i * 6 + 4]; p_pos[2] = VX[i * 6 + 5]; p_pos[3] = dim_e(p_pos);
                        pos_v[0] = VX[i * 6]; pos_v[1] = VX[i * 6 + 1]; pos_v[2] = VX[i * 6 + 2];
                    } else {
                        p_neg[0] = VX[i * 6 + 3]; p_neg[1] = VX[i * 6 + 4]; p_neg[2] = VX[i * 6 + 5]; p_neg[3] = dim_e(p_neg);
                        neg_v[0] = VX[i * 6]; neg_v[1] = VX[i * 6 + 1]; neg_v[2] = VX[i * 6 + 2];
                    }
                    for (int j = 0; j < nt; ++j) {
                        if (!((use >> j) & 1) || TR[j].slot != ib) continue;
                        const or_track& B = TR[j];
                        const float cb = (float)B.charge;
                        if (cb * ca > 0) continue;
                        if (cb > 0) {
                            p_pos[0] = VX[j * 6 + 3]; p_pos[1] = VX[j * 6 + 4]; p_pos[2] = VX[j * 6 + 5]; p_pos[3] = dim_e(p_pos);
                            pos_v[0] = VX[j * 6]; pos_v[1] = VX[j * 6 + 1]; pos_v[2] = VX[j * 6 + 2];
                        } else {
                            p_neg[0] = VX[j * 6 + 3]; p_neg[1] = VX[j * 6 + 4]; p_neg[2] = VX[j * 6 + 5]; p_neg[3] = dim_e(p_neg);
                            neg_v[0] = VX[j * 6]; neg_v[1] = VX[j * 6 + 1]; neg_v[2] = VX[j * 6 + 2];
                        }
                        float ps[4], ns[4], v1[3], v2[3], dv[3], sum[4];
                        for (int k = 0; k < 4; ++k) { ps[k] = p_pos[k]; ns[k] = p_neg[k]; }
                        for (int k = 0; k < 3; ++k) v1[k] = (pos_v[k] + neg_v[k]) * 0.5f;
                        const float vz2 = dim_closest(v2, p_pos, p_neg, ca, cb, A.ty, B.ty, dim_start(A, dfirst[i]),
                                                      dim_start(B, dfirst[j]));
                        p_pos[3] = dim_e(p_pos);
                        p_neg[3] = dim_e(p_neg);
                        if (fabsf(vz2 - v2[2]) < V_STEP_FMAG) {
                            for (int k = 0; k < 4; ++k) sum[k] = p_pos[k] + p_neg[k];
                            for (int k = 0; k < 3; ++k) dv[k] = v2[k];
                        } else {
                            for (int k = 0; k < 4; ++k) sum[k] = ps[k] + ns[k];
                            for (int k = 0; k < 3; ++k) dv[k] = v1[k];
                        }
                        const float m = sqrtf(sum[3] * sum[3] - (sum[0] * sum[0] + sum[1] * sum[1] + sum[2] * sum[2]));
                        const float pT = sqrtf(sum[0] * sum[0] + sum[1] * sum[1]);
                        const float x1 = (p_target[3] * sum[3] - (p_target[0] * sum[0] + p_target[1] * sum[1] + p_target[2] * sum[2])) / t_cms;
                        const float x2 = (p_beam[3] * sum[3] - (p_beam[0] * sum[0] + p_beam[1] * sum[1] + p_beam[2] * sum[2])) / b_cms;
                        {   // Boost (reconstruction_helper.cuh:1244-1258) by -cms velocity (0, 0, bz)
                            const float bx = 0.0f * -1, by = 0.0f * -1;
                            const float b2 = bx * bx + by * by + bz * bz;
                            const float gamma = 1.0f / sqrtf(1.0f - b2);
                            const float bp = sum[0] * bx + sum[1] * by + sum[2] * bz;
                            const float gamma2 = b2 > 0.0f ? (gamma - 1.0f) / b2 : 0.0f;
                            sum[0] = sum[0] + gamma2 * bp * bx + gamma * bx * sum[3];
                            sum[1] = sum[1] + gamma2 * bp * by + gamma * by * sum[3];
                            sum[2] = sum[2] + gamma2 * bp * bz + gamma * bz * sum[3];
                            sum[3] = gamma * (sum[3] + bp);
                        }
                        const float xF = 2.0f * sum[2] / (sqrt_s * (1.f - m * m / s));
                        const float costheta = 2 * (p_neg[3] * p_pos[2] - p_neg[2] * p_pos[3]) / (m * sqrtf(m * m + pT * pT));
                        const float phi = (float)atan2((double)(2 * sqrtf(m * m + pT * pT) * (p_neg[0] * p_pos[1] - p_neg[1] * p_pos[0])),
                                                       (double)(m * (p_pos[0] * p_pos[0] - p_neg[0] * p_neg[0] + p_pos[1] * p_pos[1] -
                                                                     p_neg[1] * p_neg[1])));
                        if (FMAGSTR * (p_pos[0] - p_neg[0]) < 0.) continue;
                        if (m < 0 || m > 10.f) continue;
                        if (fabsf(xF) > 1.f) continue;
                        if (x1 < 0.f || x1 > 1.f) continue;
                        if (x2 < 0.f || x2 > 1.f) continue;
                        if (fabsf(costheta) > 1.f) continue;
                        if (p_pos[2] + p_neg[2] > 120.f || p_pos[2] + p_neg[2] < 30.f) continue;
                        if (fabsf(p_pos[0] + p_neg[0]) > 3.f) continue;
                        if (dv[0] * dv[0] + dv[1] * dv[1] > 225.f) continue;
                        if (ndim >= OR_DIMUONS_MAX) {
                            if (a.lost) atomicAdd(a.lost, 1ull);
                            continue;
                        }
                        float* o = out + ndim * OR_DIMUON_FIELDS;
                        o[0] = m; o[1] = pT; o[2] = xF; o[3] = x1; o[4] = x2; o[5] = costheta; o[6] = phi;
                        o[7] = 0.f; o[8] = 0.f;
                        for (int k = 0; k < 3; ++k) { o[9 + k] = dv[k]; o[12 + k] = pos_v[k]; o[15 + k] = neg_v[k]; }
                        for (int k = 0; k < 4; ++k) { o[18 + k] = p_pos[k]; o[22 + k] = p_neg[k]; o[26 + k] = ps[k]; o[30 + k] = ns[k]; }
                        o[34] = 0.f; o[35] = 0.f;
                        ++ndim;
                    }
                }
            }
        }
        a.n_dim[e] = ndim;
        if (!a.hist) continue;
        // trig_array[n] = TriggerBits & n (not a bit test, reconstruction_kernels.cuh:3108); the last is "all"
        const int tb = a.trigger ? a.trigger[e] : 0;
        for (int t = 0; t < nt; ++t) {
            if (!((use >> t) & 1)) continue;
            const or_track& R = TR[t];
            if (R.chi2 / (float)((int)R.nhits - 5) > D_CHI2DOF_MAX) continue;
            const float vars[H_NVARS_TRK] = {VX[t * 6 + 3], VX[t * 6 + 4], VX[t * 6 + 5], VX[t * 6], VX[t * 6 + 1], VX[t * 6 + 2]};
            for (int trg = 0; trg < H_NTRG; ++trg)
                if ((tb & trg) || trg == H_NTRG - 1)
                    for (int k = 0; k < H_NVARS_TRK; ++k) hist_fill(lh, trg, k, vars[k]);
        }
        for (int i = 0; i < ndim; ++i) {
            const float* o = out + i * OR_DIMUON_FIELDS;
            const float vars[H_NVARS - H_NVARS_TRK] = {o[0], o[2], o[3], o[4], o[1], o[6], o[18] + o[22], o[19] + o[23],
                                                       o[20] + o[24], o[9], o[10], o[11]};
            for (int trg = 0; trg < H_NTRG; ++trg)
                if ((tb & trg) || trg == H_NTRG - 1)
                    for (int k = H_NVARS_TRK; k < H_NVARS; ++k) hist_fill(lh, trg, k, vars[k - H_NVARS_TRK]);
        }
    }
    if (!a.hist) return;
    __syncthreads();
    for (int i = threadIdx.x; i < H_NTRG * H_PACKED; i += DIM_BLOCK) {
        const uint32_t c = lh[i];
        if (!c) continue;
        const int trg = i / H_PACKED, r = i % H_PACKED;
        int k = 0;
        while (r >= h_off[k + 1]) ++k;
        atomicAdd(&a.hist[(trg * H_NVARS + k) * OR_HIST_BINS + (r - h_off[k])], (float)c);
    }
}

// ---- final straight-line fit of an output track (no reference counterpart) ----
// u = pos + sign * drift per hit, model u = cos (x0 + tx z) + sin (y0 + ty z),
// weight 1/resolution^2, z taken from FIT_ZREF.  The normal matrix and
// right-hand side are built from per-plane constants computed on the host in
// double and held as hi/lo float pairs (w h_i h_j, w h_i, h_i with
// h = (cos, cos z, sin, sin z)); the device sums them with Neumaier compensation
// and forms the u terms with error-free products (fmaf), so the FP32 sums carry
// the normal equations to about double precision.  The matrix is equilibrated
// and Cholesky-factorised in FP32; one refinement step against the compensated
// residual follows.  When (max/min pivot)^2 exceeds OR_FIT_COND_FP32, or a
// pivot is not positive, the solve and chi2 run in FP64 from the same sums.

// Neumaier compensated sum
struct CSum {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float t = s + x;
        c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ float f() const { return s + c; }
    __device__ __forceinline__ double d() const { return (double)s + (double)c; }
};

// a + b = s + e exactly, a * b = p + e exactly
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(float a, float b, float& p, float& e) {
    p = a * b;
    e = fmaf(a, b, -p);
}

// In place: a (symmetric, lower part used) = L L^T.  cond = (max/min diagonal
// of L)^2; false when a pivot is not positive.
template <typename T>
__device__ __forceinline__ bool chol4(T (&a)[4][4], T& cond) {
    T lmin = 0, lmax = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T v = a[k][k];
#pragma unroll
        for (int j = 0; j < k; ++j) v -= a[k][j] * a[k][j];
        if (!(v > T(0))) return false;
        const T l = sqrt(v);
        a[k][k] = l;
        lmin = k == 0 ? l : (l < lmin ? l : lmin);
        lmax = k == 0 ? l : (l > lmax ? l : lmax);
#pragma unroll
        for (int i = k + 1; i < 4; ++i) {
            T w = a[i][k];
#pragma unroll
            for (int j = 0; j < k; ++j) w -= a[i][j] * a[k][j];
            a[i][k] = w / l;
        }
    }
    cond = (lmax / lmin) * (lmax / lmin);
    return true;
}
// b <- (L L^T)^-1 b
template <typename T>
__device__ __forceinline__ void chol4_solve(const T (&l)[4][4], T (&b)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T v = b[i];
#pragma unroll
        for (int j = 0; j < i; ++j) v -= l[i][j] * b[j];
        b[i] = v / l[i][i];
    }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        T v = b[i];
#pragma unroll
        for (int j = i + 1; j < 4; ++j) v -= l[j][i] * b[j];
        b[i] = v / l[i][i];
    }
}

// upper-triangle index of (i, j), i <= j
__device__ __forceinline__ constexpr int fit_ut(int i, int j) { return i * 4 - i * (i - 1) / 2 + (j - i); }

template <int MAXH, int WPE>
__device__ __attribute__((noinline)) void final_fit(const Tk& T, float* out) {
    const Lds<MAXH>& L = g_lds<MAXH>;
    const int nh = T.nx + T.nuv + T.n1;
    CSum A[10], B[4];
    for (int h = 0; h < nh; ++h) {
        const int d = T.det[h], r = T.hraw[h];
        float uh, ul;
        two_sum(L.pos[r], (float)T.sign[h] * L.drift[r], uh, ul);
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            A[k].add(c_fitp[d][FITP_M + 2 * k]);
            A[k].add(c_fitp[d][FITP_M + 2 * k + 1]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float vh = c_fitp[d][FITP_V + 2 * i], vl = c_fitp[d][FITP_V + 2 * i + 1];
            float pp, pe;
            two_prod(vh, uh, pp, pe);
            B[i].add(pp); B[i].add(pe); B[i].add(vh * ul); B[i].add(vl * uh);
        }
    }
    // a fit that cannot be solved (fewer than 4 hits, or the FP64 factorisation
    // fails) keeps NaN parameters and chi2 and an infinite condition estimate
    float res8[OR_FIT_FIELDS] = {NAN, NAN, NAN, NAN, NAN, (float)(nh - 4), INFINITY, 0.f};
    if (nh >= 4) {
        float sc[4], l[4][4], cond = INFINITY;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float aii = A[fit_ut(i, i)].f();
            sc[i] = aii > 0.f ? 1.f / sqrtf(aii) : 0.f;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) l[i][j] = (A[i <= j ? fit_ut(i, j) : fit_ut(j, i)].f() * sc[i]) * sc[j];
        const bool ok = chol4(l, cond);
        if (ok && cond <= c_fit_cond) {
            // solve, then one refinement step against r = b - A p in compensated FP32
            float p[4], plo[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int i = 0; i < 4; ++i) p[i] = B[i].f() * sc[i];
            chol4_solve(l, p);
#pragma unroll
            for (int i = 0; i < 4; ++i) p[i] *= sc[i];
            float rr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                CSum acc;
                acc.add(B[i].s); acc.add(B[i].c);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const CSum& aij = A[i <= j ? fit_ut(i, j) : fit_ut(j, i)];
                    float pp, pe;
                    two_prod(aij.s, p[j], pp, pe);
                    acc.add(-pp); acc.add(-pe); acc.add(-(aij.c * p[j]));
                }
                rr[i] = acc.f() * sc[i];
            }
            chol4_solve(l, rr);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float sp, se;
                two_sum(p[i], rr[i] * sc[i], sp, se);
                p[i] = sp; plo[i] = se;
            }
            // to z = 0: x0 = x_ref - tx FIT_ZREF
            const float zr = (float)FIT_ZREF;
            auto ext = [&](float a, float alo, float t, float tlo) {
                float pp, pe, ss, se;
                two_prod(t, zr, pp, pe);
                two_sum(a, -pp, ss, se);
                return ss + ((se - pe) + (alo - tlo * zr));
            };
            res8[0] = ext(p[0], plo[0], p[1], plo[1]); res8[1] = p[1] + plo[1];
            res8[2] = ext(p[2], plo[2], p[3], plo[3]); res8[3] = p[3] + plo[3];
            // chi2 from double-float residuals u - h.p
            CSum c2;
            for (int h = 0; h < nh; ++h) {
                const int d = T.det[h], r = T.hraw[h];
                float uh, ul;
                two_sum(L.pos[r], (float)T.sign[h] * L.drift[r], uh, ul);
                CSum acc;
                acc.add(uh); acc.add(ul);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float hh = c_fitp[d][FITP_H + 2 * i], hl = c_fitp[d][FITP_H + 2 * i + 1];
                    float pp, pe;
                    two_prod(hh, p[i], pp, pe);
                    acc.add(-pp); acc.add(-pe); acc.add(-(hh * plo[i])); acc.add(-(hl * p[i]));
                }
                const float rs = acc.f();
                c2.add((c_fitp[d][FITP_W] * rs) * rs);
            }
            res8[4] = c2.f(); res8[6] = cond;
        } else {
            double ad[4][4], pd[4], condd = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                pd[i] = B[i].d();
#pragma unroll
                for (int j = 0; j < 4; ++j) ad[i][j] = A[i <= j ? fit_ut(i, j) : fit_ut(j, i)].d();
            }
            double scd[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) scd[i] = ad[i][i] > 0.0 ? 1.0 / sqrt(ad[i][i]) : 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
#pragma unroll
                for (int j = 0; j < 4; ++j) ad[i][j] = (ad[i][j] * scd[i]) * scd[j];
                pd[i] *= scd[i];
            }
            if (chol4(ad, condd)) {
                chol4_solve(ad, pd);
#pragma unroll
                for (int i = 0; i < 4; ++i) pd[i] *= scd[i];
                double c2 = 0.0;
                for (int h = 0; h < nh; ++h) {
                    const int d = T.det[h], r = T.hraw[h];
                    double m = 0.0;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        m += ((double)c_fitp[d][FITP_H + 2 * i] + (double)c_fitp[d][FITP_H + 2 * i + 1]) * pd[i];
                    const double rs = ((double)L.pos[r] + (double)T.sign[h] * (double)L.drift[r]) - m;
                    c2 += (double)c_fitp[d][FITP_W] * rs * rs;
                }
                res8[0] = (float)(pd[0] - pd[1] * FIT_ZREF); res8[1] = (float)pd[1];
                res8[2] = (float)(pd[2] - pd[3] * FIT_ZREF); res8[3] = (float)pd[3];
                res8[4] = (float)c2;
            }
            // FP32 factorisation succeeded above the bound: its estimate; failed: the FP64 one,
            // and infinity when that fails too
            res8[6] = ok ? cond : (std::isnan(res8[0]) ? INFINITY : (float)condd);
            res8[7] = 1.f;
        }
    }
#pragma unroll
    for (int q = 0; q < OR_FIT_FIELDS; ++q) out[q] = res8[q];
}

template <int MAXH, int WPE>
STAGE void stage_out(const Args& a, Scr S, int64_t e, int ntkl) {
    e = uni(e);
    ntkl = uni(ntkl);
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    const or_params& P = a.p;
    Tk* tks = S.tks();
    const uint32_t st = L.status;
    const bool rejected = (st & OR_EV_REJECT_MASK) != 0;
    int nout = 0;
    if (!rejected) {
        for (int j0 = 0; j0 < ntkl; j0 += WAVE) {
            const int j = j0 + lane;
            const Tk* Tp = j < ntkl ? &tks[S.slot_order()[j]] : nullptr;
            const bool is6 = Tp && Tp->stid >= (P.full_chain ? 7 : 6);
            const unsigned long long m = ballot(is6);
            if (is6) {
                const int o = nout + rank_in(m);
                if (o < P.max_tracks_out) {
                    const Tk& T = *Tp;
                    or_track R;
                    R.x0 = T.x0; R.tx = T.tx; R.y0 = T.y0; R.ty = T.ty; R.invp = T.invp;
                    R.err_x0 = T.ex0; R.err_tx = T.etx; R.err_y0 = T.ey0; R.err_ty = T.ety; R.err_invp = T.einvp;
                    R.chi2 = T.chi2;
                    const int nh = T.nx + T.nuv + T.n1;
                    uint32_t sp = 0, sm = 0;
                    for (int h = 0; h < OR_MAX_HITS_PER_TRACK; ++h) {
                        R.hit_index[h] = h < nh ? L.orig[T.hraw[h]] : 0;
                        if (h < nh && T.sign[h] > 0) sp |= 1u << h;
                        if (h < nh && T.sign[h] < 0) sm |= 1u << h;
                    }
                    R.sign_plus = sp; R.sign_minus = sm;
                    R.station_id = T.stid; R.nhits = (uint8_t)nh; R.charge = T.charge; R.slot = T.slot;
                    R.tracklet_id = T.canon;
                    a.tracks[e * P.max_tracks_out + o] = R;
                    if (a.vertex)
                        for (int c = 0; c < 6; ++c) a.vertex[(e * P.max_tracks_out + o) * 6 + c] = T.vtx[c];
                    if (a.fit) final_fit<MAXH, WPE>(T, a.fit + (e * P.max_tracks_out + o) * OR_FIT_FIELDS);
                }
            }
            nout += __popcll(m);
        }
    }
    if (lane == 0) {
        uint32_t so = st;
        if (!rejected && nout > P.max_tracks_out) so |= OR_EV_OUTPUT_OVERFLOW;
        a.status[e] = so;
        a.n_tracks[e] = rejected ? 0 : (nout < P.max_tracks_out ? nout : P.max_tracks_out);
        if (a.n_tracklets) a.n_tracklets[e] = rejected ? 0 : ntkl;
    }
    if (a.records) {
        float* base = a.records + (int64_t)e * P.max_tracklets_per_event * OR_TRACK_PARAMS;
        const int nrec = rejected ? 0 : (ntkl < P.max_tracklets_per_event ? ntkl : P.max_tracklets_per_event);
        for (int64_t w = lane; w < (int64_t)P.max_tracklets_per_event * OR_TRACK_PARAMS; w += WAVE) {
            const int j = (int)(w / OR_TRACK_PARAMS), f = (int)(w % OR_TRACK_PARAMS);
            float v = 0.f;
            if (j < nrec) {
                const Tk& T = tks[S.slot_order()[j]];
                const int nh = T.nx + T.nuv + T.n1;
                switch (f) {
                    case 0: v = (float)(T.stid == 0 ? 3 : T.stid); break;
                    case 1: v = (float)T.t; break;
                    case 2: v = (float)nh; break;
                    case 3: v = T.chi2; break;
                    case 5: v = T.tx; break;
                    case 6: v = T.ty; break;
                    case 7: v = T.x0; break;
                    case 8: v = T.y0; break;
                    case 9: v = T.invp; break;
                    case 10: v = T.etx; break;
                    case 11: v = T.ety; break;
                    case 12: v = T.ex0; break;
                    case 13: v = T.ey0; break;
                    case 14: v = T.einvp; break;
                    case 15: v = (float)T.charge; break;
                    case 16: case 17: case 18: case 19: case 20: case 21: v = T.stid == 7 ? T.vtx[f - 16] : 0.f; break;
                    default:
                        if (f >= 22 && f < 130 + 18) {
                            const int blk = (f - 22) / 18, h = (f - 22) % 18;
                            if (h < nh) {
                                const int r = T.hraw[h];
                                switch (blk) {
                                    case 0: v = (float)T.det[h]; break;
                                    case 1: v = (float)L.elem[r]; break;
                                    case 2: v = h < T.nx ? T.xpos[h] : L.pos[r]; break;
                                    case 3: v = L.drift[r]; break;
                                    case 4: v = (float)T.sign[h]; break;
                                    case 5: v = a.tdc[a.off[e] + L.orig[r]]; break;
                                    case 6: v = T.resid[h]; break;
                                }
                            }
                        }
                }
            }
            base[w] = v;
        }
    }
}

// Per-stage cycle counters, compiled in only for the profiling build
// (-DOR_STAGE_TIMING, libe1039or_timing.so); they add no work to the stages.
#ifdef OR_STAGE_TIMING
constexpr int N_STAGE_TIMERS = 7;
__device__ unsigned long long g_stage_cycles[NCLASS_MAX][N_STAGE_TIMERS];
#define OR_TIMER_START unsigned long long t_0 = clock64(), t_1;
#define OR_TIMER(k) t_1 = clock64(); tacc[k] += t_1 - t_0; t_0 = t_1;
#else
#define OR_TIMER_START
#define OR_TIMER(k)
#endif

#ifndef OR_WAVES_PER_SIMD_512
#define OR_WAVES_PER_SIMD_512 4
#endif
// Waves per SIMD a tile class is compiled for (the register budget): as many
// as the workgroups that fit the CU's 160 KB of LDS bring, at most 4.
template <int MAXH, int WPE>
constexpr int chain_waves_per_simd() {
    if (MAXH <= 512) return OR_WAVES_PER_SIMD_512;
    constexpr int lds = (int)sizeof(Lds<MAXH>) + (int)sizeof(WaveCfg);
    constexpr int w = (160 * 1024 / lds) * WPE / 4;
    return w < 1 ? 1 : (w > 4 ? 4 : w);
}

// The chain: one event per workgroup of WPE waves, persistent over a work queue.
// YZ (by chunks of tracklets) and station 1 (by blocks of tracklets) run on
// all waves; the other stages on the first.
template <int MAXH, int WPE>
__global__ __launch_bounds__(WAVE * WPE, (chain_waves_per_simd<MAXH, WPE>())) void chain_kernel(Args a) {
    static_assert(WPE >= 1 && WPE <= lds_waves(MAXH), "waves per event of this tile class");
    Lds<MAXH>& L = g_lds<MAXH>;
    const int lane = lane_id();
    const int w = WPE == 1 ? 0 : wave_idx();
    {
        // the regions of the first wave's block; each wave's own lie its index blocks further (region<R>)
        const int64_t blk = (int64_t)blockIdx.x * WPE;
        if (w == 0 && lane < R_COUNT) g_cfg.rp[lane] = a.scratch + blk * a.lay.stride + a.lay.off[lane];
        if (w == 0 && lane == 0) {
            g_cfg.stride = a.lay.stride;
            g_cfg.maxtkl = a.lay.maxtkl;
            g_cfg.bin_cap = a.lay.bin_cap;
            g_cfg.win_cap = a.lay.win_cap;
            g_cfg.list_cap = a.lay.list_cap;
            g_cfg.unit_cap = a.lay.unit_cap;
        }
    }
    __syncthreads();
    Scr S;
    const int64_t ne = a.ev_list ? (int64_t)*a.ev_count : a.n_events;
#ifdef OR_STAGE_TIMING
    unsigned long long tacc[N_STAGE_TIMERS] = {0, 0, 0, 0, 0, 0, 0};
#endif
    // persistent work queue: the first event of each workgroup is its block
    // index, later ones come from an atomic head, so groups that drew cheap
    // events take more
    for (int64_t i = blockIdx.x; i < ne;) {
        const int64_t e = a.ev_list ? (int64_t)a.ev_list[i] : i;
        OR_TIMER_START
        if (w == 0) {
            const int64_t h0 = a.off[e];
            const int n = (int)(a.off[e + 1] - h0);
            if (lane == 0) L.status = 0;
            if (lane < 56) L.pcnt[lane] = 0;
            wsync();
            const bool staged = stage_load_reduce<MAXH, WPE>(a, h0, n);
            if (staged) stage_keep_bits<MAXH, WPE>(a, h0, n);
            wsync();
            OR_TIMER(0)
            int nt = 0;
            if (L.status == 0) stage_xpairs<MAXH, WPE>(a.p, S.xpairs());
            OR_TIMER(1)
            if (L.status == 0) nt = stage_xz<MAXH, WPE>(a.p, S);
            if (L.status == 0 && nt > 0) stage_prop_sort<MAXH, WPE>();
            OR_TIMER(2)
            if (lane == 0) {
                L.ntkl = nt;
                L.yz_next = 0;
            }
        }
        __syncthreads();
        const int ntkl = L.ntkl;
        if (L.status == 0) {
            const Prn R = {a.pruner, a.unit_dump, a.dump_cap, a.dump_count, e};
            stage_yz<MAXH, WPE>(a.p, S, R, ntkl, a.yz_list_cap);
        }
        __syncthreads();
        OR_TIMER(3)
        if (w == 0) {
            if (L.status == 0) stage_clean<MAXH, WPE>(S.tks(), S.slot_order(), ntkl);
            if (lane == 0) L.gl_next = 0;
        }
        __syncthreads();
        OR_TIMER(4)
        if (L.status == 0) stage_global<MAXH, WPE>(a.p, S, ntkl);
        __syncthreads();
        if (w == 0) {
            if (L.status == 0 && a.p.full_chain) stage_postglobal<MAXH, WPE>(S.tks(), S.slot_order(), ntkl);
            if (L.status == 0 && (a.vertex || a.p.full_chain)) stage_vertex<MAXH, WPE>(S.tks(), ntkl);
            OR_TIMER(5)
            stage_out<MAXH, WPE>(a, S, e, ntkl);
            wsync();
            OR_TIMER(6)
            if (a.ev_head && lane == 0) L.next_ev = atomicAdd(a.ev_head, 1);
        }
        __syncthreads();
        i = a.ev_head ? (int64_t)gridDim.x + L.next_ev : i + gridDim.x;
    }
#ifdef OR_STAGE_TIMING
    constexpr int cls = MAXH == 512 ? 0 : MAXH == 1024 ? 1 : MAXH == 2048 ? 2 : MAXH == 4096 ? 3 : 4;
    if (w == 0 && lane == 0)
        for (int k = 0; k < N_STAGE_TIMERS; ++k) atomicAdd(&g_stage_cycles[cls][k], tacc[k]);
#endif
}

// Size classes: an event goes to the smallest LDS tile that holds its hits
// (the largest class takes the rest and flags them as too large).  The order
// inside a class list depends on scheduling; outputs are indexed by event, so
// results do not.
constexpr int NCLASS = NCLASS_MAX;
constexpr int MAX_EVENT_HITS = 8192;  // largest tile: 8192 hits x 17 B + tables, within the 160 KB of LDS
__host__ __device__ constexpr int class_maxh(int c) { return 512 << c; }

__global__ void classify_events(const int64_t* off, int64_t n, int nclass, int32_t* lists, int64_t stride,
                                int32_t* counts) {
    // one atomic per wave and class (wave-aggregated), slots by lane rank
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += step) {
        const int64_t e = base + lane;
        int c = -1;
        if (e < n) {
            const int64_t h = off[e + 1] - off[e];
            c = 0;
            while (c < nclass - 1 && h > class_maxh(c)) ++c;
        }
        for (int cc = 0; cc < nclass; ++cc) {
            const unsigned long long m = ballot(c == cc);
            if (!m) continue;
            const int leader = __ffsll((long long)m) - 1;
            int k0 = 0;
            if (lane == leader) k0 = atomicAdd(&counts[cc], __popcll(m));
            k0 = __shfl(k0, leader, WAVE);
            if (c == cc) lists[cc * stride + k0 + rank_in(m)] = (int32_t)e;
        }
    }
}

// keep bits are OR-ed in; clear them first (a fill kernel, not a memset)
__global__ void zero_u32(uint32_t* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// one wave per 64 feature rows; with pre1/pre2 also the hidden layers' pre-activations
__global__ void prune_score_kernel(const PrunerDev* W, const float* feats, int64_t n, float* scores, float* pre1,
                                   float* pre2) {
    const int64_t r0 = (int64_t)blockIdx.x * WAVE;
    const int64_t r = r0 + threadIdx.x;
    UnitFeatures f;
#pragma unroll
    for (int q = 0; q < PRUNE_NF; ++q) f.v[q] = r < n ? feats[r * PRUNE_NF + q] : 0.f;
    const float sc = pre1 ? prune_eval<true>(W, f, pre1, pre2, r0, n) : prune_score(W, f);
    if (r < n) scores[r] = sc;
}

}  // namespace

// ============================================================================
// Host side: session and C ABI
// ============================================================================
struct or_session {
    int device = 0;
    bool geo_registered = false;   // counted in g_dev_geo[device].users
    or_params p;
    float* d_wire = nullptr;
    char* d_scratch = nullptr;
    ScrLayout lay;               // per-wave scratch layout, from the caps and the YZ batch capacity
    int grid = 0;                 // scratch blocks (waves) held: the most any class and waves-per-event setting uses
    int maxh = 1024;
    int nclass = 2;
    int cus = 0;                  // compute units of the device
    int occ[NCLASS][4] = {};      // resident workgroups per CU by class and waves per event (1, 2, 4, 8)
    int wpe_log = 3;              // waves per event at most 1 << wpe_log (each class up to its own bound)
    int32_t* d_lists = nullptr;   // NCLASS x list_cap event indices
    int32_t* d_counts = nullptr;  // NCLASS class counts, then NCLASS work queue heads
    int64_t list_cap = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    mutable float last_ms = 0.f;
    mutable bool timing_pending = false;   // ev0/ev1 recorded, elapsed time not read yet
    int yz_list_cap = YZ_BATCH_DEFAULT;   // survivor batch size of the YZ passes (<= lay.list_cap)
    int grid_pct = 100;          // persistent grid, percent of the occupancy limit
    bool use_graph = false;      // replay the launch sequence as a captured hipGraph
    hipGraphExec_t graph_exec = nullptr;
    PrunerDev* d_pruner = nullptr;   // packed pruner weights (first or_session_set_pruner)
    bool prune_on = false;
    float prune_margin = INFINITY;   // per-tracklet score margin (or_session_set_pruner_margin)
    float* dump = nullptr;           // unit rows (training aid)
    int64_t dump_cap = 0;
    unsigned long long* dump_count = nullptr;
    RtDev* d_rt = nullptr;           // R-T calibration table (or_session_set_calibration)
    int64_t rt_nk = 0;               // its knots
    size_t d_rt_bytes = 0;
    bool rt_on = false;
    struct GraphKey {
        const void* p[19];
        int64_t n_events, n_hits, dump_cap;
        hipStream_t stream;
        int yz_list_cap, grid_pct, wpe_log;
        bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof o) == 0; }
    } graph_key;
    // host staging (host-pointer calls)
    void* d_in = nullptr;
    size_t d_in_bytes = 0;
    void* d_out = nullptr;
    size_t d_out_bytes = 0;
};

// Scratch layout of one wave from the session's caps: every region sized by
// the cap that bounds it.
static ScrLayout make_layout(const or_params& p, int list_cap) {
    ScrLayout L;
    std::memset(&L, 0, sizeof L);
    const int64_t maxtkl = p.max_tracklets_per_event, win = p.max_pairs_window, bin = p.max_pairs_per_bin;
    const int64_t unit_cap = std::max<int64_t>(list_cap, 16384);
    int64_t o = 0;
    auto put = [&](int r, int64_t bytes) {
        L.off[r] = o;
        o = (o + std::max<int64_t>(bytes, 1) + 255) / 256 * 256;
    };
    put(R_TK, (int64_t)sizeof(Tk) * maxtkl);
    put(R_CELL, 1569 * 4);
    put(R_LIST, 64 * win * 2);
    put(R_YV, std::max<int64_t>(64 * win * 16, 8 * 16 + win * (int64_t)sizeof(GX)));   // station 1: best + x records
    put(R_YY, (64 * win + 16) * 8);   // 16 entries of padding: the unit features read lists in blocks of 16
    put(R_YF, 64 * win);
    put(R_XP, 3 * 28 * bin * 2);
    put(R_SO, maxtkl * 2);
    put(R_PL, (int64_t)std::max(list_cap, GLOBAL_LIST_CAP) * 4);
    put(R_PB, (int64_t)list_cap * 4);
    put(R_KL, unit_cap * 4);
    put(R_KS, unit_cap * 4);
    put(R_YC, std::max<int64_t>((int64_t)list_cap * 48, win * 52));                     // station 1: pass-A tables
    put(R_QL, (int64_t)list_cap * 4);
    put(R_BR, YZ_CHUNK * 64);
    put(R_CL, std::min<int64_t>(maxtkl, (int64_t)WAVE * CLEAN_WORDS) * 48);
    L.stride = (o + 4095) / 4096 * 4096;
    L.maxtkl = (int32_t)maxtkl;
    L.bin_cap = (int32_t)bin;
    L.win_cap = (int32_t)win;
    L.list_cap = list_cap;
    L.unit_cap = (int32_t)unit_cap;
    return L;
}

// The plane table and the final-fit constants sit in constant memory, one copy
// per device: a session may only be created on a device whose live sessions use
// the same geometry (the tables are uploaded when no session is live).
// Scratch of or_reconstruct_noalloc: one fixed array per device, holding the
// class counters, the class event lists (OR_NOALLOC_MAX_EVENTS per class) and
// as many waves' scratch blocks as fit after them.  Host code uses the device
// address looked up once per device at session creation (g_noalloc_ptr), never
// the array's name.
constexpr int64_t NOALLOC_BYTES = (int64_t)OR_NOALLOC_SCRATCH_MB << 20;
constexpr int64_t NOALLOC_LISTS = 256;
constexpr int64_t NOALLOC_WAVES = ((NOALLOC_LISTS + (int64_t)NCLASS_MAX * OR_NOALLOC_MAX_EVENTS * 4 + 4095) / 4096) * 4096;
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {
char* g_noalloc_ptr[64];   // device address of scratch per device (set under g_geo_mu)
struct DevGeo {
    bool set = false;
    int users = 0;
    float f[F_COUNT][OR_N_DETECTORS];
};
std::mutex g_geo_mu;
DevGeo g_dev_geo[64];
}  // namespace

#define OR_HIP(x)                                                          \
    do {                                                                   \
        hipError_t _e = (x);                                               \
        if (_e != hipSuccess) {                                            \
            std::fprintf(stderr, "libe1039or: %s failed: %s\n", #x, hipGetErrorString(_e)); \
            return OR_ERR_DEVICE;                                          \
        }                                                                  \
    } while (0)

namespace {

template <int MAXH, int WPE>
int occupancy_blocks() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, chain_kernel<MAXH, WPE>, WAVE * WPE, 0) != hipSuccess) nb = 1;
    return nb < 1 ? 1 : nb;
}

// resident workgroups per CU of a tile class at 1, 2, 4, 8 waves per event (0: not built)
template <int MAXH>
void occupancy_of(int (&o)[4]) {
    o[0] = occupancy_blocks<MAXH, 1>();
    o[1] = o[2] = o[3] = 0;
    if constexpr (lds_waves(MAXH) > 1) {
        o[1] = occupancy_blocks<MAXH, 2>();
        o[2] = occupancy_blocks<MAXH, 4>();
    }
    if constexpr (lds_waves(MAXH) > 4) o[3] = occupancy_blocks<MAXH, 8>();
}

template <int MAXH, int WPE>
hipError_t launch_chain(const Args& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((chain_kernel<MAXH, WPE>), dim3(grid), dim3(WAVE * WPE), 0, s, a);
    return hipGetLastError();
}

// k: waves per event 1 << k
template <int MAXH>
hipError_t launch_wpe(int k, const Args& a, int grid, hipStream_t s) {
    if constexpr (lds_waves(MAXH) == 1) {
        return launch_chain<MAXH, 1>(a, grid, s);
    } else {
        switch (k) {
            case 0: return launch_chain<MAXH, 1>(a, grid, s);
            case 1: return launch_chain<MAXH, 2>(a, grid, s);
            case 2: return launch_chain<MAXH, 4>(a, grid, s);
            default:
                if constexpr (lds_waves(MAXH) > 4) return launch_chain<MAXH, 8>(a, grid, s);
                else return launch_chain<MAXH, 4>(a, grid, s);
        }
    }
}

hipError_t launch_class(int c, int k, const Args& a, int grid, hipStream_t s) {
    switch (c) {
        case 0: return launch_wpe<512>(k, a, grid, s);
        case 1: return launch_wpe<1024>(k, a, grid, s);
        case 2: return launch_wpe<2048>(k, a, grid, s);
        case 3: return launch_wpe<4096>(k, a, grid, s);
        default: return launch_wpe<8192>(k, a, grid, s);
    }
}

}  // namespace

extern "C" {

or_status or_session_create(const or_geometry* g, const or_params* p, int device, or_session** out) {
    if (!g || !out) return OR_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return OR_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return OR_ERR_INVALID_ARG;
    or_params prm;
    if (p) prm = *p;
    else or_params_default(&prm);
    if (prm.max_pairs_per_bin < 1 || prm.max_pairs_window < 1 || prm.max_tracklets_per_event < 1 ||
        prm.max_tracklets_per_slot < 1 || prm.max_hits_per_event < 1 || prm.max_tracks_out < 0)
        return OR_ERR_INVALID_ARG;
    if (prm.max_pairs_per_bin > PAIRS_MAX || prm.max_pairs_window > PAIRS_MAX ||
        prm.max_tracklets_per_event > TKL_MAX || prm.max_hits_per_event > MAX_EVENT_HITS ||
        prm.max_hits_chamber_plane > PLANE_CAP || prm.max_hits_prop_plane > PLANE_CAP ||
        prm.max_hits_hodo_plane > PLANE_CAP || prm.max_pair_plane_hits > 254 || prm.max_reduced_per_plane > 254)
        return OR_ERR_CAPACITY;
    auto* s = new (std::nothrow) or_session;
    if (!s) return OR_ERR_NOMEM;
    s->device = device;
    s->p = prm;
    s->nclass = 1;
    while (s->nclass < NCLASS && class_maxh(s->nclass - 1) < prm.max_hits_per_event) ++s->nclass;
    s->maxh = class_maxh(s->nclass - 1);
    OR_HIP(hipSetDevice(device));
    OR_HIP(hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking));
    s->stream = s->own;
    OR_HIP(hipEventCreate(&s->ev0));
    OR_HIP(hipEventCreate(&s->ev1));
    std::unique_lock<std::mutex> geo_lock(g_geo_mu);
    DevGeo& dg = g_dev_geo[device & 63];
    if (dg.users > 0 && std::memcmp(dg.f, g->f, sizeof g->f) != 0) {
        (void)hipEventDestroy(s->ev0);
        (void)hipEventDestroy(s->ev1);
        (void)hipStreamDestroy(s->own);
        delete s;
        return OR_ERR_CONFLICT;
    }
    if (dg.users == 0) {
    OR_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ordev::c_geo), g->f, sizeof g->f));
    {
        // final-fit plane constants in double, split into hi/lo floats
        static_assert(ordev::FITP_N == 37, "final-fit constant layout");
        float fp[OR_N_DETECTORS][ordev::FITP_N] = {};
        auto put = [](float* dst, double x) {
            const float hi = (float)x;
            dst[0] = hi;
            dst[1] = (float)(x - (double)hi);
        };
        for (int d = 0; d < OR_N_DETECTORS; ++d) {
            const double c = g->f[F_COSTHETA][d], sn = g->f[F_SINTHETA][d];
            const double z = (double)g->f[F_Z][d] - ordev::FIT_ZREF, r = g->f[F_RESOLUTION][d];
            const double w = r > 0.0 ? 1.0 / (r * r) : 0.0;
            const double hv[4] = {c, c * z, sn, sn * z};
            int k = 0;
            for (int i = 0; i < 4; ++i)
                for (int j = i; j < 4; ++j, ++k) put(&fp[d][ordev::FITP_M + 2 * k], w * hv[i] * hv[j]);
            for (int i = 0; i < 4; ++i) {
                put(&fp[d][ordev::FITP_V + 2 * i], w * hv[i]);
                put(&fp[d][ordev::FITP_H + 2 * i], hv[i]);
            }
            fp[d][ordev::FITP_W] = (float)w;
        }
        OR_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ordev::c_fitp), fp, sizeof fp));
    }
    // the YZ first-fit table of this geometry (null stream, before any session's work)
    yz_atab_kernel<<<1, 256>>>();
    OR_HIP(hipGetLastError());
    OR_HIP(hipDeviceSynchronize());
    std::memcpy(dg.f, g->f, sizeof g->f);
    dg.set = true;
    }
    ++dg.users;
    s->geo_registered = true;
    if (!g_noalloc_ptr[device & 63]) {
        void* ptr = nullptr;
        const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (se != hipSuccess || !ptr) {
            std::fprintf(stderr, "libe1039or: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(se));
            return OR_ERR_DEVICE;
        }
        g_noalloc_ptr[device & 63] = static_cast<char*>(ptr);
    }
    geo_lock.unlock();
    OR_HIP(hipMalloc(&s->d_wire, sizeof g->wire));
    OR_HIP(hipMemcpy(s->d_wire, g->wire, sizeof g->wire, hipMemcpyHostToDevice));
    hipDeviceProp_t prop;
    OR_HIP(hipGetDeviceProperties(&prop, device));
    occupancy_of<512>(s->occ[0]);
    occupancy_of<1024>(s->occ[1]);
    occupancy_of<2048>(s->occ[2]);
    occupancy_of<4096>(s->occ[3]);
    occupancy_of<8192>(s->occ[4]);
    s->cus = prop.multiProcessorCount;
    s->grid = 0;
    for (int c = 0; c < s->nclass; ++c)
        for (int k = 0; k < 4; ++k) s->grid = std::max(s->grid, s->cus * s->occ[c][k] << k);
    s->lay = make_layout(prm, YZ_LIST_MIN);
    OR_HIP(hipMalloc(&s->d_counts, sizeof(int32_t) * 2 * NCLASS));
    OR_HIP(hipMalloc(&s->d_scratch, (size_t)s->lay.stride * s->grid));
    *out = s;
    return OR_OK;
}

or_status or_session_set_stream(or_session* s, void* stream) {
    if (!s) return OR_ERR_INVALID_ARG;
    s->stream = (hipStream_t)stream;   // NULL: the null stream, ordered with the default-stream work of the caller
    return OR_OK;
}

or_status or_session_sync(or_session* s) {
    if (!s) return OR_ERR_INVALID_ARG;
    OR_HIP(hipStreamSynchronize(s->stream));
    return OR_OK;
}

float or_session_last_kernel_ms(const or_session* s) {
    if (!s) return 0.f;
    if (s->timing_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(s->ev1) == hipSuccess && hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess)
            s->last_ms = ms;
        s->timing_pending = false;
    }
    return s->last_ms;
}

or_status or_session_set_tuning(or_session* s, or_tuning key, int64_t value) {
    if (!s) return OR_ERR_INVALID_ARG;
    switch (key) {
        case OR_TUNE_YZ_LIST_CAP:
            if (value < 1 || value > YZ_LIST_MAX) return OR_ERR_INVALID_ARG;
            if (value > s->lay.list_cap) {
                // a larger batch than the scratch holds: re-size the scratch
                const ScrLayout nl = make_layout(s->p, (int)value);
                OR_HIP(hipSetDevice(s->device));
                OR_HIP(hipStreamSynchronize(s->stream));
                if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
                OR_HIP(hipFree(s->d_scratch));
                s->d_scratch = nullptr;
                OR_HIP(hipMalloc(&s->d_scratch, (size_t)nl.stride * s->grid));
                s->lay = nl;
            }
            s->yz_list_cap = (int)value;
            return OR_OK;
        case OR_TUNE_GRID_PERCENT:
            if (value < 1 || value > 100) return OR_ERR_INVALID_ARG;
            s->grid_pct = (int)value;
            return OR_OK;
        case OR_TUNE_GRAPH:
            if (value != 0 && value != 1) return OR_ERR_INVALID_ARG;
            s->use_graph = value != 0;
            return OR_OK;
        case OR_TUNE_WAVES_PER_EVENT:
            if (value != 1 && value != 2 && value != 4 && value != 8) return OR_ERR_INVALID_ARG;
            s->wpe_log = value == 1 ? 0 : (value == 2 ? 1 : (value == 4 ? 2 : 3));
            return OR_OK;
    }
    return OR_ERR_INVALID_ARG;
}

or_status or_session_get_tuning(const or_session* s, or_tuning key, int64_t* value) {
    if (!s || !value) return OR_ERR_INVALID_ARG;
    switch (key) {
        case OR_TUNE_YZ_LIST_CAP: *value = s->yz_list_cap; return OR_OK;
        case OR_TUNE_GRID_PERCENT: *value = s->grid_pct; return OR_OK;
        case OR_TUNE_GRAPH: *value = s->use_graph ? 1 : 0; return OR_OK;
        case OR_TUNE_WAVES_PER_EVENT: *value = (int64_t)1 << s->wpe_log; return OR_OK;
    }
    return OR_ERR_INVALID_ARG;
}

// Device buffers of one call's launch sequence: per-wave scratch (waves_max
// blocks of the session's stride; a workgroup of W waves takes W consecutive ones), the class event lists (NCLASS x
// list_stride) and the class counters (2 x NCLASS).
struct ChainBufs {
    char* scratch;
    int64_t waves_max;
    int32_t* lists;
    int64_t list_stride;
    int32_t* counts;
};

static or_status issue_chain_with(or_session* s, const or_hits* in, const or_result* out, const ChainBufs& B) {
    Args a;
    a.det = in->det; a.elem = in->elem; a.tdc = in->tdc; a.drift = in->drift; a.flag = in->flag;
    a.off = in->event_offsets; a.n_events = in->n_events;
    a.ev_list = nullptr; a.ev_count = nullptr; a.ev_head = nullptr;
    a.wire = s->d_wire;
    a.p = s->p;
    a.status = out->status; a.keep_mask = out->keep_mask; a.n_tracks = out->n_tracks; a.tracks = out->tracks;
    a.n_tracklets = out->n_tracklets; a.records = out->records; a.vertex = out->vertex; a.fit = out->fit;
    a.scratch = B.scratch;
    a.lay = s->lay;
    a.yz_list_cap = std::min(s->yz_list_cap, s->lay.list_cap);
    a.pruner = s->prune_on ? s->d_pruner : nullptr;
    a.rt = s->rt_on ? s->d_rt : nullptr;
    a.unit_dump = s->dump;
    a.dump_cap = s->dump_cap;
    a.dump_count = s->dump_count;
    const int64_t nw = (in->n_hits + 31) / 32;
    if (nw > 0) {
        hipLaunchKernelGGL(zero_u32, dim3(1024), dim3(256), 0, s->stream, out->keep_mask, nw);
        OR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<uint32_t*>(B.counts),
                       (int64_t)(2 * NCLASS));
    OR_HIP(hipGetLastError());
    hipLaunchKernelGGL(classify_events, dim3(std::min<int64_t>(1024, (in->n_events + 255) / 256)), dim3(256), 0,
                       s->stream, in->event_offsets, in->n_events, s->nclass, B.lists, B.list_stride, B.counts);
    OR_HIP(hipGetLastError());
    for (int c = 0; c < s->nclass; ++c) {
        a.ev_list = B.lists + c * B.list_stride;
        a.ev_count = B.counts + c;
        a.ev_head = B.counts + NCLASS + c;
        int k = std::min(s->wpe_log, lds_waves_log(class_maxh(c)));
        while (k > 0 && ((int64_t)1 << k) > B.waves_max) --k;   // a small fixed scratch: fewer waves per event
        const int64_t g = std::max<int64_t>(1, (int64_t)s->cus * s->occ[c][k] * s->grid_pct / 100);
        const int grid = (int)std::min<int64_t>(std::min<int64_t>(g, in->n_events), B.waves_max >> k);
        OR_HIP(launch_class(c, k, a, grid, s->stream));
    }
    return OR_OK;
}

// The launch sequence of one call: keep-mask fill, class counters fill,
// classification, one persistent launch per tile class.  No memset nodes, so
// it can be captured into a graph.
static or_status issue_chain(or_session* s, const or_hits* in, const or_result* out) {
    const ChainBufs B = {s->d_scratch, s->grid, s->d_lists, s->list_cap, s->d_counts};
    return issue_chain_with(s, in, out, B);
}

static or_status run_device(or_session* s, const or_hits* in, const or_result* out, bool timed) {
    if (in->n_events == 0) {
        const int64_t nw = (in->n_hits + 31) / 32;
        if (nw > 0) {
            hipLaunchKernelGGL(zero_u32, dim3(1024), dim3(256), 0, s->stream, out->keep_mask, nw);
            OR_HIP(hipGetLastError());
        }
        s->last_ms = 0.f;
        s->timing_pending = false;
        return OR_OK;
    }
    if (s->list_cap < in->n_events) {
        if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
        if (s->d_lists) (void)hipFree(s->d_lists);
        s->d_lists = nullptr;
        OR_HIP(hipMalloc(&s->d_lists, sizeof(int32_t) * NCLASS * (size_t)in->n_events));
        s->list_cap = in->n_events;
    }
    if (!s->use_graph) {
        if (timed) OR_HIP(hipEventRecord(s->ev0, s->stream));
        const or_status st = issue_chain(s, in, out);
        if (st != OR_OK) return st;
    } else {
        or_session::GraphKey key;
        std::memset(&key, 0, sizeof key);
        const void* ptrs[19] = {in->det, in->elem, in->tdc, in->drift, in->flag, in->event_offsets, out->status,
                                out->keep_mask, out->n_tracks, out->tracks, out->n_tracklets, out->records,
                                out->vertex, out->fit, s->d_lists, s->prune_on ? s->d_pruner : nullptr, s->dump,
                                s->dump_count, s->rt_on ? s->d_rt : nullptr};
        std::memcpy(key.p, ptrs, sizeof ptrs);
        key.n_events = in->n_events; key.n_hits = in->n_hits; key.dump_cap = s->dump_cap; key.stream = s->stream;
        key.yz_list_cap = s->yz_list_cap; key.grid_pct = s->grid_pct; key.wpe_log = s->wpe_log;
        if (!s->graph_exec || !(key == s->graph_key)) {
            if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
            // captured on the session's own stream (the null stream cannot capture),
            // launched on the caller's stream
            hipGraph_t g = nullptr;
            hipStream_t launch = s->stream;
            const hipError_t be = hipStreamBeginCapture(s->own, hipStreamCaptureModeThreadLocal);
            OR_HIP(be);
            s->stream = s->own;
            const or_status st = issue_chain(s, in, out);
            const hipError_t ce = hipStreamEndCapture(s->own, &g);
            s->stream = launch;
            if (st != OR_OK) { if (g) (void)hipGraphDestroy(g); return st; }
            OR_HIP(ce);
            const hipError_t ie = hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            OR_HIP(ie);
            std::memcpy(&s->graph_key, &key, sizeof key);
        }
        if (timed) OR_HIP(hipEventRecord(s->ev0, s->stream));
        OR_HIP(hipGraphLaunch(s->graph_exec, s->stream));
    }
    if (timed) {
        OR_HIP(hipEventRecord(s->ev1, s->stream));
        s->timing_pending = true;
    }
    return OR_OK;
}

or_status or_reconstruct(or_session* s, const or_hits* in, or_result* out) {
    if (!s || !in || !out || in->n_events < 0 || in->n_hits < 0) return OR_ERR_INVALID_ARG;
    if (!in->det || !in->elem || !in->tdc || !in->drift || !in->event_offsets) return OR_ERR_INVALID_ARG;
    if (!out->status || !out->keep_mask || !out->n_tracks || (!out->tracks && s->p.max_tracks_out > 0))
        return OR_ERR_INVALID_ARG;
    OR_HIP(hipSetDevice(s->device));
    if (in->on_device && out->on_device) {
        // asynchronous; or_session_last_kernel_ms reads the timing events later
        return run_device(s, in, out, true);
    }
    if (in->on_device != out->on_device) return OR_ERR_INVALID_ARG;
    // host buffers: stage through device memory
    const int64_t ne = in->n_events, nh = in->n_hits;
    const int64_t nw = (nh + 31) / 32;
    const size_t in_bytes = (size_t)nh * (2 + 2 + 4 + 4 + 1) + (size_t)(ne + 1) * 8 + 256 * 8;
    const size_t tk_bytes = (size_t)ne * s->p.max_tracks_out * sizeof(or_track);
    const size_t rec_bytes = out->records ? (size_t)ne * s->p.max_tracklets_per_event * OR_TRACK_PARAMS * 4 : 0;
    const size_t vtx_bytes = out->vertex ? (size_t)ne * s->p.max_tracks_out * 6 * 4 : 0;
    const size_t fit_bytes = out->fit ? (size_t)ne * s->p.max_tracks_out * OR_FIT_FIELDS * 4 : 0;
    const size_t out_bytes = (size_t)ne * 4 * 3 + (size_t)nw * 4 + tk_bytes + rec_bytes + vtx_bytes + fit_bytes + 256 * 8;
    if (s->d_in_bytes < in_bytes) {
        if (s->d_in) (void)hipFree(s->d_in);
        s->d_in = nullptr;
        OR_HIP(hipMalloc(&s->d_in, in_bytes));
        s->d_in_bytes = in_bytes;
    }
    if (s->d_out_bytes < out_bytes) {
        if (s->d_out) (void)hipFree(s->d_out);
        s->d_out = nullptr;
        OR_HIP(hipMalloc(&s->d_out, out_bytes));
        s->d_out_bytes = out_bytes;
    }
    auto carve = [](char*& p, size_t bytes) { char* r = p; p += (bytes + 255) / 256 * 256; return r; };
    char* pi = (char*)s->d_in;
    or_hits din = *in;
    din.det = (int16_t*)carve(pi, nh * 2);
    din.elem = (int16_t*)carve(pi, nh * 2);
    din.tdc = (float*)carve(pi, nh * 4);
    din.drift = (float*)carve(pi, nh * 4);
    din.flag = in->flag ? (uint8_t*)carve(pi, nh) : nullptr;
    din.event_offsets = (int64_t*)carve(pi, (ne + 1) * 8);
    din.on_device = 1;
    hipStream_t st = s->stream;
    OR_HIP(hipMemcpyAsync((void*)din.det, in->det, nh * 2, hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync((void*)din.elem, in->elem, nh * 2, hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync((void*)din.tdc, in->tdc, nh * 4, hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync((void*)din.drift, in->drift, nh * 4, hipMemcpyHostToDevice, st));
    if (in->flag) OR_HIP(hipMemcpyAsync((void*)din.flag, in->flag, nh, hipMemcpyHostToDevice, st));
    OR_HIP(hipMemcpyAsync((void*)din.event_offsets, in->event_offsets, (ne + 1) * 8, hipMemcpyHostToDevice, st));
    char* po = (char*)s->d_out;
    or_result dout;
    dout.status = (uint32_t*)carve(po, ne * 4);
    dout.n_tracks = (int32_t*)carve(po, ne * 4);
    dout.n_tracklets = out->n_tracklets ? (int32_t*)carve(po, ne * 4) : nullptr;
    dout.keep_mask = (uint32_t*)carve(po, nw * 4);
    dout.tracks = (or_track*)carve(po, tk_bytes);
    dout.records = out->records ? (float*)carve(po, rec_bytes) : nullptr;
    dout.vertex = out->vertex ? (float*)carve(po, vtx_bytes) : nullptr;
    dout.fit = out->fit ? (float*)carve(po, fit_bytes) : nullptr;
    dout.on_device = 1;
    or_status rs = run_device(s, &din, &dout, true);
    if (rs != OR_OK) return rs;
    OR_HIP(hipMemcpyAsync(out->status, dout.status, ne * 4, hipMemcpyDeviceToHost, st));
    OR_HIP(hipMemcpyAsync(out->n_tracks, dout.n_tracks, ne * 4, hipMemcpyDeviceToHost, st));
    if (out->n_tracklets) OR_HIP(hipMemcpyAsync(out->n_tracklets, dout.n_tracklets, ne * 4, hipMemcpyDeviceToHost, st));
    if (nw) OR_HIP(hipMemcpyAsync(out->keep_mask, dout.keep_mask, nw * 4, hipMemcpyDeviceToHost, st));
    if (tk_bytes) OR_HIP(hipMemcpyAsync(out->tracks, dout.tracks, tk_bytes, hipMemcpyDeviceToHost, st));
    if (rec_bytes) OR_HIP(hipMemcpyAsync(out->records, dout.records, rec_bytes, hipMemcpyDeviceToHost, st));
    if (vtx_bytes) OR_HIP(hipMemcpyAsync(out->vertex, dout.vertex, vtx_bytes, hipMemcpyDeviceToHost, st));
    if (fit_bytes) OR_HIP(hipMemcpyAsync(out->fit, dout.fit, fit_bytes, hipMemcpyDeviceToHost, st));
    OR_HIP(hipStreamSynchronize(st));
    return OR_OK;
}

or_status or_reconstruct_noalloc(or_session* s, const or_hits* in, or_result* out) {
    if (!s || !in || !out || in->n_events < 0 || in->n_hits < 0) return OR_ERR_INVALID_ARG;
    if (!in->on_device || !out->on_device) return OR_ERR_INVALID_ARG;
    if (!in->det || !in->elem || !in->tdc || !in->drift || !in->event_offsets) return OR_ERR_INVALID_ARG;
    if (!out->status || !out->keep_mask || !out->n_tracks || (!out->tracks && s->p.max_tracks_out > 0))
        return OR_ERR_INVALID_ARG;
    if (in->n_events > OR_NOALLOC_MAX_EVENTS) return OR_ERR_CAPACITY;
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_geo_mu);
        base = g_noalloc_ptr[s->device & 63];
    }
    if (!base) return OR_ERR_DEVICE;
    const int64_t waves = (NOALLOC_BYTES - NOALLOC_WAVES) / s->lay.stride;
    if (waves < 1) return OR_ERR_CAPACITY;   // the session's per-wave scratch does not fit
    OR_HIP(hipSetDevice(s->device));
    if (in->n_events == 0) {
        const int64_t nw = (in->n_hits + 31) / 32;
        if (nw > 0) {
            hipLaunchKernelGGL(zero_u32, dim3(1024), dim3(256), 0, s->stream, out->keep_mask, nw);
            OR_HIP(hipGetLastError());
        }
        s->last_ms = 0.f;
        s->timing_pending = false;
        return OR_OK;
    }
    const ChainBufs B = {base + NOALLOC_WAVES, waves, reinterpret_cast<int32_t*>(base + NOALLOC_LISTS),
                         (int64_t)OR_NOALLOC_MAX_EVENTS, reinterpret_cast<int32_t*>(base)};
    OR_HIP(hipEventRecord(s->ev0, s->stream));
    const or_status st = issue_chain_with(s, in, out, B);
    if (st != OR_OK) return st;
    OR_HIP(hipEventRecord(s->ev1, s->stream));
    s->timing_pending = true;
    return OR_OK;
}

or_status or_session_set_pruner(or_session* s, const float* weights, int64_t n_weights, float threshold) {
    if (!s) return OR_ERR_INVALID_ARG;
    if (!weights) {
        s->prune_on = false;
        return OR_OK;
    }
    if (n_weights != PRUNE_NW || std::isnan(threshold)) return OR_ERR_INVALID_ARG;
    for (int64_t i = 0; i < n_weights; ++i)
        if (!std::isfinite(weights[i])) return OR_ERR_INVALID_ARG;
    PrunerDev h;
    prune_pack(weights, threshold, s->prune_margin, &h);
    OR_HIP(hipSetDevice(s->device));
    if (!s->d_pruner) OR_HIP(hipMalloc(&s->d_pruner, sizeof(PrunerDev)));
    OR_HIP(hipStreamSynchronize(s->stream));
    OR_HIP(hipMemcpy(s->d_pruner, &h, sizeof h, hipMemcpyHostToDevice));
    s->prune_on = true;
    return OR_OK;
}

or_status or_session_set_pruner_margin(or_session* s, float margin) {
    if (!s || std::isnan(margin) || margin < 0.f) return OR_ERR_INVALID_ARG;
    s->prune_margin = margin;
    if (s->d_pruner) {
        OR_HIP(hipSetDevice(s->device));
        OR_HIP(hipStreamSynchronize(s->stream));
        OR_HIP(hipMemcpy(reinterpret_cast<char*>(s->d_pruner) + offsetof(PrunerDev, margin), &margin, sizeof margin,
                         hipMemcpyHostToDevice));
    }
    return OR_OK;
}

or_status or_session_set_calibration(or_session* s, const or_calibration* c) {
    if (!s) return OR_ERR_INVALID_ARG;
    if (!c) {
        s->rt_on = false;
        return OR_OK;
    }
    // image: RtDev, then the knots (double), then the interval cubics (double4)
    int64_t nk = 0;
    for (int d = 0; d < OR_N_DETECTORS; ++d) nk += c->det[d].n;
    const size_t head = (sizeof(RtDev) + 255) / 256 * 256;
    const size_t xb = (size_t)std::max<int64_t>(nk, 1) * sizeof(double);
    const size_t bytes = head + (xb + 31) / 32 * 32 + (size_t)std::max<int64_t>(nk, 1) * sizeof(double4);
    std::vector<char> img(bytes, 0);
    OR_HIP(hipSetDevice(s->device));
    if (s->d_rt_bytes < bytes) {
        OR_HIP(hipStreamSynchronize(s->stream));
        if (s->d_rt) OR_HIP(hipFree(s->d_rt));
        s->d_rt = nullptr;
        s->d_rt_bytes = 0;
        s->rt_on = false;
        OR_HIP(hipMalloc(&s->d_rt, bytes));
        s->d_rt_bytes = bytes;
    }
    char* base = reinterpret_cast<char*>(s->d_rt);
    RtDev* h = reinterpret_cast<RtDev*>(img.data());
    double* hx = reinterpret_cast<double*>(img.data() + head);
    double4* hc = reinterpret_cast<double4*>(img.data() + head + (xb + 31) / 32 * 32);
    h->x = reinterpret_cast<const double*>(base + head);
    h->cf = reinterpret_cast<const double4*>(base + head + (xb + 31) / 32 * 32);
    h->nk = (int32_t)nk;
    int64_t o = 0;
    for (int d = 0; d < OR_N_DETECTORS; ++d) {
        const or_calibration::Det& D = c->det[d];
        RtHead& hd = h->h[d];
        hd.n = D.n; hd.off = (int32_t)o; hd.tmin = D.tmin; hd.tmax = D.tmax;
        hd.x0 = D.n ? D.x[0] : 0.0;
        hd.xl = D.n ? D.x[D.n - 1] : 0.0;
        hd.scale = D.n > 1 ? (double)(D.n - 1) / (D.x[D.n - 1] - D.x[0]) : 0.0;
        for (int k = 0; k < D.n; ++k) {
            hx[o + k] = D.x[k];
            hc[o + k] = k + 1 < D.n ? double4{D.y[k], D.b[k], D.c[k], D.d[k]} : double4{0.0, 0.0, 0.0, 0.0};
        }
        o += D.n;
    }
    OR_HIP(hipStreamSynchronize(s->stream));
    OR_HIP(hipMemcpy(s->d_rt, img.data(), bytes, hipMemcpyHostToDevice));
    s->rt_nk = nk;
    s->rt_on = true;
    return OR_OK;
}

or_status or_rt_calibrate(or_session* s, const or_hits* in, float* drift_out) {
    if (!s || !in || !s->rt_on || in->n_hits < 0) return OR_ERR_INVALID_ARG;
    if (in->n_hits == 0) return OR_OK;
    if (!in->det || !in->tdc || !in->drift || !drift_out) return OR_ERR_INVALID_ARG;
    OR_HIP(hipSetDevice(s->device));
    const int64_t n = in->n_hits;
    int cus = 0;
    OR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    const int64_t per_block = (int64_t)RT_BLOCK * RT_UNROLL;
    // knots in LDS when they fit: a few blocks per CU, so that few copies of the table are staged
    const bool xlds = s->rt_nk <= RT_LDS_KNOTS;
    const size_t lds = xlds ? (size_t)std::max<int64_t>(s->rt_nk, 1) * sizeof(double) : 0;
    const unsigned grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((xlds ? 4LL : 16LL) * std::max(1, cus), (n + per_block - 1) / per_block));
    auto launch = [&](const int16_t* d, const float* t, const float* r, float* o) {
        if (xlds)
            hipLaunchKernelGGL(rt_calibrate_kernel<true>, dim3(grid), dim3(RT_BLOCK), lds, s->stream, s->d_rt, d, t, r, n, o);
        else
            hipLaunchKernelGGL(rt_calibrate_kernel<false>, dim3(grid), dim3(RT_BLOCK), 0, s->stream, s->d_rt, d, t, r, n, o);
        return hipGetLastError();
    };
    if (in->on_device) {
        OR_HIP(launch(in->det, in->tdc, in->drift, drift_out));
        return OR_OK;
    }
    char* buf = nullptr;
    const size_t b2 = ((size_t)n * 2 + 255) / 256 * 256, b4 = ((size_t)n * 4 + 255) / 256 * 256;
    OR_HIP(hipMalloc(&buf, b2 + 3 * b4));
    int16_t* dd = reinterpret_cast<int16_t*>(buf);
    float* dt = reinterpret_cast<float*>(buf + b2);
    float* dr = reinterpret_cast<float*>(buf + b2 + b4);
    float* dout = reinterpret_cast<float*>(buf + b2 + 2 * b4);
    hipError_t e = hipMemcpyAsync(dd, in->det, (size_t)n * 2, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, in->tdc, (size_t)n * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, in->drift, (size_t)n * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = launch(dd, dt, dr, dout);
    if (e == hipSuccess) e = hipMemcpyAsync(drift_out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    const hipError_t ef = hipFree(buf);
    OR_HIP(e);
    OR_HIP(ef);
    return OR_OK;
}

or_status or_dimuons(or_session* s, const or_hits* in, const or_result* chain, const int32_t* trigger_bits,
                     or_dimuon_result* out) {
    if (!s || !in || !chain || !out || !in->on_device || !chain->on_device || in->n_events < 0) return OR_ERR_INVALID_ARG;
    if (!in->det || !in->event_offsets || !chain->status || !chain->n_tracks || !chain->tracks || !chain->vertex ||
        !out->n_dimuons || !out->dimuons)
        return OR_ERR_INVALID_ARG;
    if (s->p.max_tracks_out > DIM_MAX_TRACKS) return OR_ERR_CAPACITY;
    if (in->n_events == 0) return OR_OK;
    OR_HIP(hipSetDevice(s->device));
    DimArgs a;
    a.n_events = in->n_events;
    a.max_out = s->p.max_tracks_out;
    a.n_tracks = chain->n_tracks;
    a.status = chain->status;
    a.tracks = chain->tracks;
    a.vertex = chain->vertex;
    a.det = in->det;
    a.off = in->event_offsets;
    a.trigger = trigger_bits;
    a.n_dim = out->n_dimuons;
    a.dim = out->dimuons;
    a.hist = out->histograms;
    a.lost = out->lost;
    int dev_cus = 0;
    OR_HIP(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, s->device));
    // one lane per event; with histograms, per-block LDS counts flushed once per block, so the grid is
    // kept to a few blocks per CU
    const int64_t want = (in->n_events + DIM_BLOCK - 1) / DIM_BLOCK;
    const int64_t cap = out->histograms ? 4LL * std::max(1, dev_cus) : 65535;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, cap));
    hipLaunchKernelGGL(dimuon_kernel, dim3(grid), dim3(DIM_BLOCK), 0, s->stream, a);
    OR_HIP(hipGetLastError());
    return OR_OK;
}

or_status or_session_set_fit_cond_bound(or_session* s, float bound) {
    if (!s || !(bound >= 0.f)) return OR_ERR_INVALID_ARG;
    OR_HIP(hipSetDevice(s->device));
    OR_HIP(hipStreamSynchronize(s->stream));
    OR_HIP(hipMemcpyToSymbol(HIP_SYMBOL(ordev::c_fit_cond), &bound, sizeof bound));
    return OR_OK;
}

or_status or_pruner_score(or_session* s, const float* features, int64_t n, float* scores) {
    if (!s || n < 0 || (n > 0 && (!features || !scores)) || !s->prune_on) return OR_ERR_INVALID_ARG;
    if (n == 0) return OR_OK;
    OR_HIP(hipSetDevice(s->device));
    hipLaunchKernelGGL(prune_score_kernel, dim3((unsigned)((n + WAVE - 1) / WAVE)), dim3(WAVE), 0, s->stream,
                       s->d_pruner, features, n, scores, nullptr, nullptr);
    OR_HIP(hipGetLastError());
    return OR_OK;
}

or_status or_pruner_layers(or_session* s, const float* features, int64_t n, float* scores, float* pre1, float* pre2) {
    if (!s || n < 0 || (n > 0 && (!features || !scores || !pre1 || !pre2)) || !s->prune_on) return OR_ERR_INVALID_ARG;
    if (n == 0) return OR_OK;
    OR_HIP(hipSetDevice(s->device));
    hipLaunchKernelGGL(prune_score_kernel, dim3((unsigned)((n + WAVE - 1) / WAVE)), dim3(WAVE), 0, s->stream,
                       s->d_pruner, features, n, scores, pre1, pre2);
    OR_HIP(hipGetLastError());
    return OR_OK;
}

or_status or_session_set_unit_dump(or_session* s, float* rows, int64_t cap, unsigned long long* count) {
    if (!s) return OR_ERR_INVALID_ARG;
    if (!rows) {
        s->dump = nullptr; s->dump_cap = 0; s->dump_count = nullptr;
        return OR_OK;
    }
    if (!count || cap < 0) return OR_ERR_INVALID_ARG;
    s->dump = rows; s->dump_cap = cap; s->dump_count = count;
    return OR_OK;
}

or_status or_session_autotune(or_session* s, const or_hits* in, or_result* out, int reps, float* best_ms) {
    if (!s || !in || !out || !in->on_device || !out->on_device || reps < 1) return OR_ERR_INVALID_ARG;
    const int wpe_logs[] = {0, 1, 2, 3};
    const int list_caps[] = {s->lay.list_cap, s->lay.list_cap / 2, s->lay.list_cap / 4, s->lay.list_cap / 8};
    static const int grid_pcts[] = {75, 50};
    int best_w = s->wpe_log, best_cap = s->yz_list_cap, best_pct = s->grid_pct;
    float best = -1.f;
    auto time_it = [&](float& ms) -> or_status {
        or_status st = or_reconstruct(s, in, out);   // warm-up at this setting
        if (st != OR_OK) return st;
        float tot = 0.f;
        for (int r = 0; r < reps; ++r) {
            st = or_reconstruct(s, in, out);
            if (st != OR_OK) return st;
            tot += or_session_last_kernel_ms(s);
        }
        ms = tot / (float)reps;
        return OR_OK;
    };
    // waves per event x YZ survivor batch on the full grid, then smaller grids at the best pair
    s->grid_pct = 100;
    for (int wl : wpe_logs) {
        for (int cap : list_caps) {
            s->wpe_log = wl;
            s->yz_list_cap = cap;
            float ms = 0.f;
            const or_status st = time_it(ms);
            if (st != OR_OK) return st;
            if (best < 0.f || ms < best) { best = ms; best_w = wl; best_cap = cap; best_pct = 100; }
        }
    }
    s->wpe_log = best_w;
    s->yz_list_cap = best_cap;
    for (int pct : grid_pcts) {
        s->grid_pct = pct;
        float ms = 0.f;
        const or_status st = time_it(ms);
        if (st != OR_OK) return st;
        if (ms < best) { best = ms; best_pct = pct; }
    }
    s->grid_pct = best_pct;
    if (best_ms) *best_ms = best;
    return OR_OK;
}

void or_session_destroy(or_session* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->geo_registered) {
        std::lock_guard<std::mutex> lk(g_geo_mu);
        --g_dev_geo[s->device & 63].users;
    }
    if (s->d_wire) (void)hipFree(s->d_wire);
    if (s->d_scratch) (void)hipFree(s->d_scratch);
    if (s->d_lists) (void)hipFree(s->d_lists);
    if (s->d_counts) (void)hipFree(s->d_counts);
    if (s->d_pruner) (void)hipFree(s->d_pruner);
    if (s->d_rt) (void)hipFree(s->d_rt);
    if (s->d_in) (void)hipFree(s->d_in);
    if (s->d_out) (void)hipFree(s->d_out);
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->own) (void)hipStreamDestroy(s->own);
    delete s;
}
}

/* Cycles per stage summed over all waves since the last reset, [class][stage]
 * with stages load+reduce, X pairs, XZ, YZ, cleaning, global, outputs.  Only
 * the profiling build (-DOR_STAGE_TIMING) counts; the product build returns
 * OR_ERR_INVALID_ARG. */
extern "C" or_status or_debug_sub_cycles(unsigned long long* out, int reset) {
    if (!out) return OR_ERR_INVALID_ARG;
#ifdef OR_STAGE_TIMING
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_sub_cycles), sizeof(g_sub_cycles)) != hipSuccess) return OR_ERR_DEVICE;
    if (reset) {
        static const unsigned long long zero[N_SUB_TIMERS] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_sub_cycles), zero, sizeof(zero)) != hipSuccess) return OR_ERR_DEVICE;
    }
    return OR_OK;
#else
    (void)reset;
    for (int i = 0; i < N_SUB_TIMERS; ++i) out[i] = 0;
    return OR_ERR_INVALID_ARG;
#endif
}

extern "C" or_status or_debug_stage_cycles(unsigned long long* out, int reset) {
    if (!out) return OR_ERR_INVALID_ARG;
#ifdef OR_STAGE_TIMING
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_stage_cycles), sizeof(g_stage_cycles)) != hipSuccess) return OR_ERR_DEVICE;
    if (reset) {
        static const unsigned long long zero[NCLASS_MAX][N_STAGE_TIMERS] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_stage_cycles), zero, sizeof(zero)) != hipSuccess) return OR_ERR_DEVICE;
    }
    return OR_OK;
#else
    (void)reset;
    for (int i = 0; i < 35; ++i) out[i] = 0;
    return OR_ERR_INVALID_ARG;
#endif
}

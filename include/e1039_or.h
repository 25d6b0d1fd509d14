/*
 * e1039_or.h -- C ABI of the MI355X online-reconstruction chain.
 *
 * The chain: event reduction -> station-2/3 XZ tracklets -> YZ extension ->
 * back-partial cleaning -> station-1 extension (global tracks).
 *
 * The reference (OnlineReconstruction/online_reconstruction.cu) has no library
 * API: main() stages hits into device structs and launches kernels in sequence
 * (online_reconstruction.cu:776-894).  The entry points below replace that
 * sequence; each cites the reference code it stands in for.  All functions
 * have C linkage, take plain pointers and sizes, and never throw.
 *
 * Threading: one or_session per host thread (or per stream).  Sessions are
 * independent.  Outputs are in a canonical order independent of scheduling.
 */
#ifndef E1039_OR_H
#define E1039_OR_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
typedef enum {
    OR_OK = 0,
    OR_ERR_INVALID_ARG = 1,
    OR_ERR_IO = 2,            /* a file cannot be opened, read or written  */
    OR_ERR_PARSE = 3,         /* geometry line with a detector id out of range */
    OR_ERR_DEVICE = 4,        /* HIP runtime error                          */
    OR_ERR_NOMEM = 5,
    OR_ERR_NO_DEVICE = 6,     /* no GPU visible                             */
    OR_ERR_CAPACITY = 7,      /* a parameter exceeds the bound of an index encoding (see or_params) */
    OR_ERR_CONFLICT = 8       /* a live session on the device uses a different geometry */
} or_status;

/* ---- per-event status bits (the reference's HasTooManyHits, split by cause)
 * Any bit set means the event is rejected and produces no tracks, as
 * HasTooManyHits does in the reference (online_reconstruction.cu:1067-1072). */
enum {
    OR_EV_D0_MULT        = 1u << 0,  /* station-1 multiplicity > cap (reconstruction_kernels.cuh:135) */
    OR_EV_D2_MULT        = 1u << 1,  /* (reconstruction_kernels.cuh:136) */
    OR_EV_D3P_MULT       = 1u << 2,  /* (reconstruction_kernels.cuh:137) */
    OR_EV_D3M_MULT       = 1u << 3,  /* (reconstruction_kernels.cuh:138) */
    OR_EV_XZ_BIN_OVERFLOW= 1u << 4,  /* > 20 X pairs in one bin (reconstruction_kernels.cuh:359-375) */
    OR_EV_TRACKLET_OVERFLOW = 1u << 5, /* tracklet slots exhausted (reconstruction_kernels.cuh:719,824,835) */
    OR_EV_PLANE_OVERFLOW = 1u << 6,  /* more hits in a plane than the staging arrays hold (reconstruction_constants.h:173-175) */
    OR_EV_PAIR_OVERFLOW  = 1u << 7,  /* > max_pair_plane_hits in a pair-builder plane, > max_pairs_window pairs in a window list
                                        (reconstruction_kernels.cuh:225,896), or the four YZ window lists of one tracklet
                                        spanning more than 2^27 (u3, v3, u2, v2) combinations */
    OR_EV_CLUSTER_OVERFLOW = 1u << 8,/* chamber cluster longer than 192 hits (reconstruction_helper.cuh:237) */
    OR_EV_EVENT_TOO_LARGE= 1u << 9,  /* event has more hits than the session's per-event bound */
    OR_EV_BAD_HIT        = 1u << 10, /* detectorID < 1 or elementID outside [1, nelem] */
    OR_EV_OUTPUT_OVERFLOW= 1u << 11  /* more global tracks than max_tracks_out (tracks beyond it dropped, event kept) */
};
#define OR_EV_REJECT_MASK 0x7FFu

/* ---- constants of the reference data model ----------------------------- */
#define OR_N_DETECTORS 55          /* ids 0..54, 0 unused (reconstruction_constants.h:13) */
#define OR_MAX_ELEM 404            /* wire table width (online_reconstruction.cu:191) */
#define OR_TRACK_PARAMS 148        /* float record size (reconstruction_constants.h:171) */
#define OR_MAX_HITS_PER_TRACK 18   /* (reconstruction_constants.h:182) */

typedef struct or_geometry or_geometry;   /* opaque */
typedef struct or_session or_session;     /* opaque */

/* Runtime parameters: caps and cuts of the reference as values, so that
 * high-occupancy workloads can lift them.  or_params_default() fills the
 * reference values.  The session sizes its device scratch from them (tracklet
 * storage from max_tracklets_per_event, X pair bins from max_pairs_per_bin,
 * window lists from max_pairs_window).  Bounds come from index encodings only:
 * max_pairs_per_bin and max_pairs_window <= 255, max_tracklets_per_event <=
 * 16384, per-plane hit caps <= 400, max_pair_plane_hits and
 * max_reduced_per_plane <= 254, max_hits_per_event <= 8192; beyond them
 * or_session_create returns OR_ERR_CAPACITY. */
typedef struct {
    int32_t max_hits_chamber_plane;   /* 200  datasizes::NMaxHitsChambers (reconstruction_constants.h:173) */
    int32_t max_hits_hodo_plane;      /* 60   (reconstruction_constants.h:174) */
    int32_t max_hits_prop_plane;      /* 120  (reconstruction_constants.h:175) */
    int32_t max_reduced_per_plane;    /* 120  hitarraycopy[600]/5 (reconstruction_kernels.cuh:16) */
    int32_t max_pair_plane_hits;      /* 100  hitflag[100] (reconstruction_kernels.cuh:225) */
    int32_t max_pairs_per_bin;        /* 20   MaxHitsProj_X (reconstruction_constants.h:154) */
    int32_t max_pairs_window;         /* 100  hitpairs_*[100] (reconstruction_kernels.cuh:896) */
    int32_t max_tracklets_per_slot;   /* 16   TrackSizeMax/THREADS_PER_BLOCK */
    int32_t max_tracklets_per_event;  /* 512  TrackSizeMax (reconstruction_constants.h:181) */
    int32_t max_d0_mult;              /* 210  (reconstruction_constants.h:188) */
    int32_t max_d2_mult;              /* 120  (reconstruction_constants.h:189) */
    int32_t max_d3_mult;              /* 90   (reconstruction_constants.h:190) */
    int32_t max_tracks_out;           /* output records per event (no reference counterpart) */
    int32_t max_hits_per_event;       /* per-event bound the device tiles are sized for, <= 8192; larger events get OR_EV_EVENT_TOO_LARGE */
    int32_t full_chain;               /* 0 (default): stop after global tracking, output stationID >= 6 tracks.
                                         1: then prop-tube segment matching (reconstruction_helper.cuh:1916-2257),
                                         global track cleaning (:2265-2354, slot by slot, then slot against slot in
                                         slot order) and vertexing; output stationID >= 7 tracks, as the reference's
                                         unpack (online_reconstruction.cu:1133-1136) */
} or_params;

/* Input hits, structure of arrays, events contiguous.  Field meaning follows
 * the reference's Hit class (LoadInput.h:10-30): detectorID and elementID are
 * int16, tdcTime and driftDistance fp32.  driftDistance is used as |d|
 * (online_reconstruction.cu:600).  flag: bit 1 set = in time
 * (online_reconstruction.cu:599 stores 1<<is_in_time); NULL = all in time.
 * Within an event the hit order is the reference's staging order: hits are
 * grouped by detector keeping their relative order (online_reconstruction.cu:536-640). */
typedef struct {
    int64_t n_events;
    int64_t n_hits;
    const int16_t* det;
    const int16_t* elem;
    const float* tdc;
    const float* drift;
    const uint8_t* flag;            /* may be NULL */
    const int64_t* event_offsets;   /* n_events+1 */
    int32_t on_device;              /* 1: the pointers above are device pointers */
    const int32_t* event_id;        /* optional, n_events: the events' IDs (LoadInput.h EventID), carried to
                                       the outputs that name events (or_output_fill, host memory); the chain
                                       does not read it.  NULL: the event index */
} or_hits;

/* Compact output track, 96 bytes.  Fields map to the reference record
 * (reconstruction_classes.cuh:245-440) and OROutput branches (OROutput.cxx:34-60). */
typedef struct {
    float x0, tx, y0, ty, invp;
    float err_x0, err_tx, err_y0, err_ty, err_invp;
    float chi2;
    uint16_t hit_index[OR_MAX_HITS_PER_TRACK]; /* event-local input index of each hit */
    uint32_t sign_plus;      /* bit k: hit k has sign +1 */
    uint32_t sign_minus;     /* bit k: hit k has sign -1 */
    int8_t station_id;       /* 6 = global track */
    uint8_t nhits;
    int8_t charge;
    uint8_t slot;            /* storage slot (reference thread array) */
    uint32_t tracklet_id;    /* position in the canonical XZ tracklet list */
} or_track;

/* Output buffers, owned by the caller.  Optional ones may be NULL. */
typedef struct {
    uint32_t* status;        /* n_events: OR_EV_* bits */
    uint32_t* keep_mask;     /* ceil(n_hits/32) words: bit h = input hit h survives the reducer */
    int32_t* n_tracks;       /* n_events: global tracks written */
    or_track* tracks;        /* n_events * max_tracks_out, slot order; only slots [0, n_tracks[e]) of event e are written */
    int32_t* n_tracklets;    /* optional, n_events: XZ tracklets found */
    float* records;          /* optional, n_events * max_tracklets_per_event * 148 floats:
                                every tracklet in slot order, reference record layout */
    float* vertex;           /* optional, n_events * max_tracks_out * 6 floats (vx, vy, vz, px, py, pz):
                                when set, global tracks are swum to the beam line (stationID 7,
                                gKernel_Vertexing, reconstruction_kernels.cuh:2547-2810) and records
                                carry the same values in fields 16-21 */
    float* fit;              /* optional, n_events * max_tracks_out * 8 floats per output track: final
                                straight-line fit over all its hits (x0, tx, y0, ty, chi2, ndf, condition
                                estimate, 1 when the 4x4 solve ran in FP64); see OR_FIT_* below.
                                No reference counterpart (the reference's 4x4 inverse,
                                trackfittingtools.cuh:84-349, is never called) */
    int32_t on_device;       /* 1: device pointers */
} or_result;

/* Final fit of an output track: measurement u = wire position + sign * drift of
 * each hit, model u = cos(theta) (x0 + tx z) + sin(theta) (y0 + ty z) of its
 * plane, weight 1/resolution^2, least squares over the 4 parameters.  The
 * per-plane terms of the normal equations (w h h^T, w h and h, with
 * h = (cos, cos z', sin, sin z') and z' = z - 1000 cm) are computed once per
 * session in double and held as pairs of floats; per hit they are summed in
 * FP32 with compensation (Neumaier), and the u terms are formed with
 * error-free products.  The equilibrated 4x4 system is factorised (Cholesky)
 * and solved in FP32 with one refinement step against the compensated
 * residual, then extrapolated to z = 0 with error-free products; chi2 sums
 * the residuals in double-float.  When the condition estimate (max/min
 * Cholesky pivot)^2 exceeds the bound (OR_FIT_COND_FP32 unless
 * or_session_set_fit_cond_bound changed it) or a pivot is not positive, the
 * solve and chi2 run in FP64 from the same sums.  Fields per track: */
#define OR_FIT_FIELDS 8          /* x0, tx, y0, ty, chi2, ndf (hits - 4), condition estimate, FP64 flag;
                                    a track whose fit cannot be solved (fewer than 4 hits, or the FP64
                                    factorisation fails) gets NaN x0..chi2 and an infinite condition estimate */
#define OR_FIT_COND_FP32 1.0e4f

/* ---- entry points ---------------------------------------------------------- */

/* Fills p with the reference caps and cuts. */
void or_params_default(or_params* p);

/* Parses a basic_geometry_*.txt file and builds the wire table.
 * Replaces online_reconstruction.cu:115-218 (geometry parse into gPlane,
 * slope_max/inter_max, wire_position[55][400]). */
or_status or_geometry_load(const char* path, or_geometry** out);
void or_geometry_free(or_geometry* g);

/* Copies the 27 per-plane geometry arrays (reconstruction_classes.cuh:947-977,
 * 27 x 55 floats, nelem as float) and the wire table (55 x OR_MAX_ELEM floats)
 * into caller buffers; either pointer may be NULL. */
or_status or_geometry_export(const or_geometry* g, float* planes_27x55, float* wire_table);

/* Creates a device session: uploads geometry, sizes the per-wave scratch from
 * the caps in p.  Replaces the cudaMalloc/cudaMemcpy block
 * (online_reconstruction.cu:652-757).  The plane table lives in constant
 * memory, one copy per device: while sessions are live on a device, a session
 * with a different geometry is refused (OR_ERR_CONFLICT). */
or_status or_session_create(const or_geometry* g, const or_params* p, int device, or_session** out);

/* Runs the chain on a batch of events.  Replaces the kernel sequence
 * gkernel_eR -> gKernel_XZ_tracking -> gKernel_YZ_tracking ->
 * gKernel_BackTrackCleaning(_crossthreads) -> gKernel_Global_tracking
 * (online_reconstruction.cu:776-894) and the unpack of reduced hits and
 * tracks (online_reconstruction.cu:1063-1166).  With on_device inputs and
 * outputs the call is asynchronous on the session stream; otherwise it stages
 * through device buffers and returns when results are on the host. */
or_status or_reconstruct(or_session* s, const or_hits* in, or_result* out);

/* or_reconstruct for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference.  Same results as or_reconstruct; device
 * pointers only (in->on_device and out->on_device), asynchronous on the
 * session stream, never captured into the session's graph.  Calls no
 * hipMalloc, hipFree or allocating routine: its per-call device memory (class
 * counters, class event lists, per-wave scratch) is one fixed array of
 * OR_NOALLOC_SCRATCH_MB MiB per device, held until the process ends, so fewer
 * waves run at once than in or_reconstruct.  Limits: at most
 * OR_NOALLOC_MAX_EVENTS events per call, and the session's per-wave scratch
 * (set by its caps) must fit the array; otherwise OR_ERR_CAPACITY.  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (other streams, other threads, other sessions) race and must be ordered by
 * the caller. */
#define OR_NOALLOC_SCRATCH_MB 128
#define OR_NOALLOC_MAX_EVENTS 262144
or_status or_reconstruct_noalloc(or_session* s, const or_hits* in, or_result* out);

/* Waits for the session stream. */
or_status or_session_sync(or_session* s);

/* Sets the HIP stream (hipStream_t) work is issued on; NULL = the null
 * (default) stream.  Until the first call the session uses a stream of its own. */
or_status or_session_set_stream(or_session* s, void* stream);

/* Device time in ms of the last or_reconstruct's chain kernels (all size
 * classes), measured with HIP events on the stream they ran on.  Waits for
 * that call's kernels to finish. */
float or_session_last_kernel_ms(const or_session* s);

/* Tuning knobs of a session; results do not depend on them.  No reference
 * counterpart (the reference hard-codes its launch shapes). */
typedef enum {
    /* YZ candidates surviving the first box cut that are evaluated together
     * (one batch), 1..2^20 (default 1024; capacity 4096 at creation).  A chunk of 16 tracklets with more
     * survivors runs in several batches cut at (u3, v3) unit boundaries; a
     * chunk in which one unit alone has more is evaluated unit by unit.
     * A value above the session's current batch capacity re-sizes its scratch
     * (waits for the session stream). */
    OR_TUNE_YZ_LIST_CAP = 1,
    /* persistent grid size in percent of the occupancy limit, 1..100 (default 100) */
    OR_TUNE_GRID_PERCENT = 2,
    /* 1: capture the launch sequence of a call into a hipGraph and replay it
     * while the buffers, sizes and stream stay the same; 0 (default): launch */
    OR_TUNE_GRAPH = 3,
    /* waves per event at most, 1, 2, 4 or 8 (default 8): each tile class runs
     * min(value, its bound) waves per event, the bound being 1 for the 512-hit
     * class, 4 for the 1,024- and 2,048-hit classes and 8 for the 4,096- and
     * 8,192-hit classes.  The waves of an event share its LDS tile and split its
     * YZ and station-1 work by tracklets; the other stages run on one wave.
     * Results do not depend on it. */
    OR_TUNE_WAVES_PER_EVENT = 4
} or_tuning;
or_status or_session_set_tuning(or_session* s, or_tuning key, int64_t value);

/* Reads a tuning knob of the session into *value. */
or_status or_session_get_tuning(const or_session* s, or_tuning key, int64_t* value);

/* Times the device-resident batch (in, out) at each combination of waves per
 * event (1, 2, 4) and YZ survivor batch (the session's batch capacity, a half,
 * a quarter and an eighth of it), on the full persistent grid, then the best pair at 75
 * and 50 % of the occupancy limit; reps timed calls after one warm-up each.
 * Keeps the fastest setting in the session and returns its mean chain time in
 * *best_ms.  Outputs are those of the last call.  No reference counterpart
 * (the reference launches one block per event, online_reconstruction.cu:776). */
or_status or_session_autotune(or_session* s, const or_hits* in, or_result* out, int reps, float* best_ms);

/* ---- R-T calibration: drift distance from the TDC time ---------------------
 * Per detector a cubic spline R(T) through the calibration knots with zero
 * slope at both ends (ROOT TSpline3 option "b1e1" as online_reconstruction.cu:240
 * builds it), and the rule of online_reconstruction.cu:389-397 for chambers
 * (1-30) and prop tubes (47-54): tdc < tmin -> cellwidth/2, tdc > tmax -> 0,
 * else R(tdc), evaluated in FP64 and rounded to float.  Hodoscopes (31-46) and
 * detectors without knots keep the input drift.  Interval choice as TSpline3's:
 * knot interval k with t_k < tdc <= t_k+1; below the first knot the first
 * interval's cubic, above the last knot the last interval's cubic (the shipped
 * prop-tube calibration has knots 562.5..1197.5 inside tmin, tmax = 560, 1200). */
typedef struct or_calibration or_calibration;
#define OR_RT_MAX_KNOTS 512

/* Parses a calibration_*.txt file: per detector a line "detectorID nBin tmin
 * tmax name" followed by nBin lines "iBin T R".  Replaces
 * online_reconstruction.cu:220-242. */
or_status or_calibration_load(const char* path, or_calibration** out);

/* Builds a calibration from tables: n_det detectors, det_ids[i] in 1..54,
 * n_knots[i] in {0} or 2..OR_RT_MAX_KNOTS knots taken in order from t and r
 * (strictly increasing t per detector), tmin/tmax per detector.  Entries for
 * hodoscopes are accepted and not applied (as in the reference). */
or_status or_calibration_create(int n_det, const int32_t* det_ids, const int32_t* n_knots, const float* tmin,
                                const float* tmax, const double* t, const double* r, or_calibration** out);
void or_calibration_free(or_calibration* c);

/* Uploads c to the session's device; from the next or_reconstruct the chain
 * derives the drift of calibrated detectors from in->tdc inside its event load
 * (the input drift is read for the others).  c == NULL turns it off.  The
 * session keeps its own copy; c may be freed afterwards.  Waits for the
 * session stream. */
or_status or_session_set_calibration(or_session* s, const or_calibration* c);

/* Writes the drift of every input hit to drift_out (n_hits floats): the R-T
 * rule for calibrated detectors, the input drift for the others, without the
 * absolute value the chain takes.  Needs a calibration set on the session.
 * With in->on_device, device pointers and asynchronous on the session stream;
 * otherwise host pointers, returning when drift_out is written.  Replaces the
 * per-hit drift_distance computation of online_reconstruction.cu:387-397. */
or_status or_rt_calibrate(or_session* s, const or_hits* in, float* drift_out);

/* ---- dimuons and display histograms -----------------------------------------
 * Replaces gKernel_DimuonBuilding (reconstruction_kernels.cuh:2820-3090, with
 * FindDimuonClosestApproach, reconstruction_helper.cuh:1269-1476) and
 * gKernel_fill_display_histograms (reconstruction_kernels.cuh:3100-3195), run
 * on the outputs of a full-chain or_reconstruct (or_params.full_chain = 1,
 * or_result.vertex set): opposite-charge pairs of the event's stationID >= 7
 * tracks, taken in the reference's order (slot a <= slot b, tracks of a, tracks
 * of b; a pair inside one slot appears in both orders), swum back through FMAG
 * together to their closest approach, kinematics and the reference's cuts.
 * Record of OR_DIMUON_FIELDS floats (reconstruction_classes.cuh:879-934):
 * mass, pT, xF, x1, x2, costheta, phi, mass_single, chi2_single, vertex xyz,
 * positive-track vertex xyz, negative-track vertex xyz, p+ (px, py, pz, E),
 * p- (px, py, pz, E), single p+ (4), single p- (4), chi2_vtx, chi2_kf.  The
 * reference never computes mass_single, chi2_single, chi2_vtx, chi2_kf; they
 * are 0 here.  phi is atan2 in FP64 rounded to float (the reference's atan2f is
 * within 2 ulp of it).  At most OR_DIMUONS_MAX per event (the reference writes
 * a 17th past its array); further ones are counted in *lost.
 * Histograms: OR_HIST_TRIGGERS x OR_HIST_VARS x OR_HIST_BINS floats, index
 * (trigger * OR_HIST_VARS + variable) * OR_HIST_BINS + bin, the reference's
 * bins (online_reconstruction.cu:670-726); counts are added to what the buffer
 * holds.  Trigger n of an event counts when (trigger_bits & n) != 0, the last
 * always (the reference's test, reconstruction_kernels.cuh:3108).  Variables:
 * track px, py, pz, vx, vy, vz (stationID >= 7, chi2/(nhits-5) <= 250), dimuon
 * mass, xF, x1, x2, pT, phi, px, py, pz, vx, vy, vz. */
#define OR_DIMUON_FIELDS 36
#define OR_DIMUONS_MAX 16
#define OR_HIST_TRIGGERS 11
#define OR_HIST_VARS 18
#define OR_HIST_BINS 256

typedef struct {
    int32_t* n_dimuons;          /* n_events */
    float* dimuons;              /* n_events * OR_DIMUONS_MAX * OR_DIMUON_FIELDS; rows [0, n_dimuons[e]) written */
    float* histograms;           /* optional: OR_HIST_TRIGGERS * OR_HIST_VARS * OR_HIST_BINS, counts added */
    unsigned long long* lost;    /* optional, one counter: dimuons beyond OR_DIMUONS_MAX, added */
} or_dimuon_result;

/* Device pointers only (in->on_device and chain->on_device), asynchronous on
 * the session stream; the session must be the one that produced chain (its
 * max_tracks_out, at most 64, else OR_ERR_CAPACITY).  trigger_bits: n_events
 * int32, or NULL (all 0). */
or_status or_dimuons(or_session* s, const or_hits* in, const or_result* chain, const int32_t* trigger_bits,
                     or_dimuon_result* out);

/* ---- columnar event output (the reference's OROutput tree) ----------------
 * One row per written event and one array per branch, with the branch names
 * and types of OROutput.cxx:24-102 (Int_t and vector<int> as int32,
 * vector<float> as float32).  A vector branch holds the events' values
 * concatenated; its length per event is the event's counter branch (nhits for
 * hit.*, ntracks for track.* but track.hit*, whose lengths are track.nhits,
 * ndimuons for dimuon.*).  Filled as online_reconstruction.cu:1063-1222 fills
 * the tree: events without hits or failing the station multiplicity cuts
 * (OR_EV_D0_MULT .. OR_EV_D3M_MULT, the reference's HasTooManyHits) are not
 * written; hit.* are the hits the reducer kept (keep_mask), detector by
 * detector 1..54 in staging order, chan = elementID, pos = the wire position,
 * tdc and |drift| as given; track.* come from the tracklet records (slot
 * order) whose stationID is at least min_station_id (the reference writes 7
 * and above, the full chain's tracks); dimuon.* from the or_dimuon_result rows,
 * dimuon.px/py/pz being the sums of the two muons' momenta as the reference
 * fills them.  track.invp_tgt stays empty (the reference fills it only when
 * built with TEST_MOMENTUM).  One branch beyond the reference's, status
 * (uint32), holds each written event's OR_EV_* bits.
 * The archive is an uncompressed .npz (numpy.load reads it) written by
 * or_output_close; each array below 4 GiB. */
typedef struct or_output or_output;   /* opaque */
or_status or_output_open(const char* path, or_output** out);
/* Appends the events of one batch, all host pointers: in (event_id optional),
 * chain (status, keep_mask, n_tracklets and records required) from a session
 * with the given max_tracklets_per_event, dim optional (NULL: no dimuons). */
or_status or_output_fill(or_output* w, const or_geometry* g, const or_hits* in, const or_result* chain,
                         int32_t max_tracklets_per_event, const or_dimuon_result* dim, int32_t min_station_id);
/* Writes the archive and frees the writer (also when it fails). */
or_status or_output_close(or_output* w);

/* ---- learned pruning of YZ station-3 units (no reference counterpart) -----
 * A YZ unit is one (u3, v3) window-list pair of a tracklet: the reference
 * evaluates every (u2, v2) pair under every unit (the i_uv3 loop,
 * reconstruction_kernels.cuh:1320-1800).  With a pruner set, each unit is scored by a 16-64-64-1 ReLU
 * perceptron on the matrix cores and units scoring below the threshold are
 * skipped; everything else is the plain chain, so each track either comes out
 * identical to the plain chain's or (when its unit was skipped) differs or is
 * missing.  The 16 features are those of or_session_set_unit_dump.
 * weights: OR_PRUNE_WEIGHTS floats, row-major (out x in): w1[64][16], b1[64],
 * w2[64][64], b2[64], w3[64], b3; score = w3 . relu(w2 relu(w1 f + b1) + b2) + b3
 * with bf16 weights and activations and fp32 sums.  weights == NULL turns
 * pruning off.  Call between reconstructions (it waits for the session stream). */
#define OR_PRUNE_WEIGHTS 5313
or_status or_session_set_pruner(or_session* s, const float* weights, int64_t n_weights, float threshold);

/* Relative pruning: a unit is kept only when its score is also within margin
 * of the best score among the units of its tracklet (the units compete for
 * the tracklet's one YZ winner).  INFINITY (the default) turns it off; the
 * margin is kept across or_session_set_pruner calls.  Waits for the session
 * stream.  No reference counterpart. */
or_status or_session_set_pruner_margin(or_session* s, float margin);

/* Scores n rows of 16 unit features with the session's pruner (device
 * pointers; asynchronous on the session stream). */
or_status or_pruner_score(or_session* s, const float* features, int64_t n, float* scores);

/* As or_pruner_score, and also writes both hidden layers' pre-activations
 * (bias added, before ReLU) of each row: pre1 and pre2 hold n x 64 floats,
 * row-major (device pointers).  A check of the matrix-core arithmetic; no
 * reference counterpart. */
or_status or_pruner_layers(or_session* s, const float* features, int64_t n, float* scores, float* pre1, float* pre2);

/* Training aid: during device-resident calls, one row of OR_UNIT_ROW floats
 * per scored unit is appended to rows (device memory, cap rows): event index,
 * tracklet_id, u3.v3 index (over the window lists' entries with an accepted
 * hit, u3 fastest), unit flag (1: the tracklet's YZ winner is in this unit,
 * 0.5: searched, not the winner, 0: pruned, -1: unknown, the chunk went unit
 * by unit and every unit of it was searched), then the 16 features.  A unit's station-3 y is
 * the mean of the y of its accepted U3 and V3 hits (sign 0); the line through
 * it and the beam axis at z = 0 predicts later hits (distances in 10 cm):
 * 10 tx; y3/50; (yU3 - yV3)/10 and |yU3 - yV3|/10; (yU3a - yU3b)/10 and
 * (yV3a - yV3b)/10; U3 and V3 hits - 1; the distance of the nearest U2 and of
 * the nearest V2 list entry to the prediction; for the representative
 * candidate (the unit with those nearest U2 and V2 entries, evaluated as the
 * chain evaluates every candidate) log(1 + chi2)/4 and the stage it reaches
 * (0: fails the first box cut, 1: a sign-free test, 2: passes); nearest U2 y
 * minus nearest V2 y; the distance of the nearest P1 Y and P2 Y prop-tube hit
 * to the prediction; the representative's |y0|/50.  Each is clamped at 10.
 * Rows of one tracklet are contiguous.  *count (device uint64, zeroed by the
 * caller) counts the rows wanted; rows beyond cap are not written.  Chunks
 * whose units overflow the session's unit list are neither pruned nor
 * dumped.  rows == NULL turns it off. */
#define OR_UNIT_ROW 20
or_status or_session_set_unit_dump(or_session* s, float* rows, int64_t cap, unsigned long long* count);

/* Profiling: device cycles per chain stage summed over all waves since the last
 * reset, out[class * 7 + stage] for the 5 tile classes (512 ... 8192 hits) and
 * the stages load+reduce, X pairs, XZ, YZ, cleaning, global, outputs.  Counts
 * only in the profiling build libe1039or_timing.so; libe1039or.so returns
 * OR_ERR_INVALID_ARG.  No reference counterpart. */
or_status or_debug_stage_cycles(unsigned long long* out /* 35 */, int reset);

/* Profiling: cycles of sub-stages summed over waves (wave-level timers, all
 * tile classes): YZ window lists, YZ passes 1-3, YZ winners, YZ unit path,
 * station-1 window lists, station-1 candidate loop, track write; 16 slots.
 * Same build rule as or_debug_stage_cycles.  No reference counterpart. */
or_status or_debug_sub_cycles(unsigned long long* out /* 18 */, int reset);

/* Sets the condition bound above which the final fit solves in FP64 (0 sends
 * every track to the FP64 path).  The bound is held per device and shared by
 * every session on the session's device; it applies to launches made after the
 * call.  No reference counterpart. */
or_status or_session_set_fit_cond_bound(or_session* s, float bound);

void or_session_destroy(or_session* s);
const char* or_status_string(or_status st);

#ifdef __cplusplus
}
#endif
#endif

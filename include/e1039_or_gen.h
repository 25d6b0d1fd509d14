/*
 * e1039_or_gen.h -- seeded synthetic event generator.
 *
 * No counterpart in the reference: the reference reads hits from ROOT files
 * (online_reconstruction.cu:246-642), which are absent.  This generator
 * produces hits of the same fields (LoadInput.h:10-30) from muon tracks
 * that are straight between the FMAG and KMAG bend planes, on the same
 * geometry, so the chain can be run and timed on seeded inputs.
 *
 * Conventions (all fixed by the seed; results do not depend on thread count):
 *  - each track: a muon from the target (z = -300, x/y ~ N(0, 1 cm)) with
 *    downstream slopes tx ~ U(+-tx_max), ty ~ U(+-ty_max), charge q = +-1 and
 *    1/p ~ U(invp_min, invp_max); it is bent at the FMAG bend plane (z 251.4,
 *    pT kick -3.066086 GeV) and at the KMAG bend plane (z 1064.26, pT kick
 *    -0.3819216 GeV), slope change = kick * q / p; tracks are redrawn until
 *    |x0| < 150, |y0| < 50 (downstream line at z = 0), q agrees with the
 *    charge rule of reconstruction_helper.cuh:684-691 and the track crosses
 *    the station-1 X plane (detector 4), the station-2 X plane (16) and a
 *    station-3 X plane (22 or 28);
 *  - one hit per crossed plane (detectors 1-6, 13-54; 7-12 are not
 *    instrumented in E1039 mode): element = wire of smallest distance,
 *    driftDistance = |distance| + N(0, drift_sigma) for chambers and prop
 *    tubes, 0 for hodoscopes; tdcTime = tmax - drift/v inside the detector's
 *    calibration window;
 *  - noise: mult_mode 0 = each channel hit with probability noise_occupancy;
 *    mult_mode 1 = total hits ~ Poisson(mult_mean) capped at mult_cap;
 *    mult_mode 2 = Poisson(mult_mean) with probability 1-mult_frac2, else
 *    Poisson(mult_mean2), capped; noise spread uniformly over channels;
 *  - after-pulses: each hit duplicated with probability afterpulse_frac,
 *    tdc += U(10,200), placed right after its parent;
 *  - within an event hits are sorted by (detectorID, elementID), flag = 2.
 */
#ifndef E1039_OR_GEN_H
#define E1039_OR_GEN_H
#include "e1039_or.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t seed;
    int32_t n_tracks_min, n_tracks_max;
    float tx_max, ty_max;
    float invp_min, invp_max;
    float drift_sigma;
    float noise_occupancy;
    float afterpulse_frac;
    int32_t mult_mode;
    float mult_mean, mult_mean2, mult_frac2;
    int32_t mult_cap;
} or_gen_params;

#define OR_GEN_MAX_TRACKS 4

typedef struct {
    int64_t n_events, n_hits;
    int16_t* det;
    int16_t* elem;
    float* tdc;
    float* drift;
    uint8_t* flag;
    int64_t* event_offsets;       /* n_events + 1 */
    int8_t* n_true;               /* n_events: tracks generated */
    float* truth;                 /* n_events x OR_GEN_MAX_TRACKS x 6: x0, tx, y0, ty, invp, charge */
} or_gen_events;

void or_gen_params_default(or_gen_params* p);
/* Generates n_events events with nthreads host threads (0 = hardware). */
or_status or_generate(const or_geometry* g, const or_gen_params* p, int64_t n_events, int nthreads,
                      or_gen_events* out);
void or_gen_free(or_gen_events* ev);

#ifdef __cplusplus
}
#endif
#endif

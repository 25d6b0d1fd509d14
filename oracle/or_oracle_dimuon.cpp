/*
 * or_oracle_dimuon.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's dimuon building and display histograms:
 *   gKernel_DimuonBuilding          reconstruction_kernels.cuh:2820-3090
 *   FindDimuonClosestApproach       reconstruction_helper.cuh:1269-1476
 *   4-vector helpers                reconstruction_helper.cuh:1211-1266
 *   gKernel_fill_display_histograms reconstruction_kernels.cuh:3100-3195
 *   histogram bin set-up            online_reconstruction.cu:670-726
 * The reference keeps the swim in float arrays of 3*(NSTEPS+1) elements and
 * mixes double constants into float expressions; this restatement keeps both,
 * so each statement rounds where the reference's does.  Input: the global
 * tracks of the chain with stationID >= 7 (vertexed), in slot order.
 *
 * Fixed readings of the reference:
 *  (D1) momentum and vertex values the swim writes into p_pos / p_neg stay in
 *       place for the next partner of the same first track, as in the
 *       reference (they are only set from the track lists when a track is
 *       taken);
 *  (D2) mass_single, chi2_single, chi2_vtx and chi2_kf are never computed by
 *       the reference (uninitialised); they are written as 0;
 *  (D3) the reference writes a 17th dimuon past the end of the event's 16
 *       slots (ndim > DimuonSizeMax test); here at most 16 are kept and the
 *       rest are counted as lost;
 *  (D4) phi = atan2 evaluated in double and rounded to float (the reference
 *       calls CUDA's atan2f, documented to 2 ulp).
 */
#include "../include/e1039_or.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

// reconstruction_constants.h:88-121 (the E906 set), :183-184, :194, :210-263
constexpr float FMAGSTR = -1.054f, PT_KICK_KMAG = -0.3819216f, Z_KMAG_BEND = 1064.26f, Z_DUMP = 42.f;
constexpr double PTKICK_UNIT = -0.006096568, FMAG_LENGTH = 502.92, FMAG_HOLE_LENGTH = 27.94,
                 FMAG_HOLE_RADIUS = 1.27, STEP_FMAG = 2.514600, STEP_TARGET = 5.0;
constexpr int NSTEPS_FMAG = 100, NSTEPS_TARGET = 100, NS = NSTEPS_FMAG + NSTEPS_TARGET;
constexpr double DEDX_UNIT_0 = 0.014282073, DEDX_UNIT_1 = 7.1869681e-05, DEDX_UNIT_2 = -1.4279150e-06,
                 DEDX_UNIT_3 = 1.5853655e-08, DEDX_UNIT_4 = -6.0741470e-11;
constexpr float Ebeam = 120.f, Mp = 0.938272f, Pbeam = 119.996331801f, Mmu = 0.105658f;
constexpr float chi2dofmax = 250.f;
constexpr int NDIM_PARAM = 36, DIM_MAX = 16, NTRG = 11, NVARS = 18, NVARS_TRK = 6, NBINS = 256;

float E(float m, const float* p) { return sqrtf(m * m + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }
float M2(const float* p) { return p[3] * p[3] - (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }
float Mass(const float* p) { return sqrtf(p[3] * p[3] - (p[0] * p[0] + p[1] * p[1] + p[2] * p[2])); }
float Perp(const float* p) { return sqrtf(p[0] * p[0] + p[1] * p[1]); }
float mult4vec(const float* a, const float* b) { return a[3] * b[3] - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]); }
void Boost(float* p, float bx, float by, float bz) {
    float b2 = bx * bx + by * by + bz * bz;
    float gamma = 1.0f / sqrtf(1.0f - b2);
    float bp = p[0] * bx + p[1] * by + p[2] * bz;
    float gamma2 = b2 > 0.0f ? (gamma - 1.0f) / b2 : 0.0;
    p[0] = p[0] + gamma2 * bp * bx + gamma * bx * p[3];
    p[1] = p[1] + gamma2 * bp * by + gamma * by * p[3];
    p[2] = p[2] + gamma2 * bp * bz + gamma * bz * p[3];
    p[3] = gamma * (p[3] + bp);
}

// FindDimuonClosestApproach; `quirk2` selects DEDX_UNIT_2 as the linear
// coefficient of the first energy-loss term (track 2 in the reference)
void fmag_step(float* pos, float* mom, int step, float charge, float ty, bool quirk2) {
    const int ix = step * 3, iy = ix + 1, iz = ix + 2, jx = ix - 3, jy = ix - 2, jz = ix - 1;
    float tx_i = mom[jx] / mom[jz];
    float tx_f = tx_i + 2.f * charge * PTKICK_UNIT * STEP_FMAG / std::sqrt(mom[jx] * mom[jx] + mom[jz] * mom[jz]);
    float traj1[3] = {(float)(tx_i * STEP_FMAG), (float)(ty * STEP_FMAG), (float)STEP_FMAG};
    float pos_b[3] = {pos[jx] - traj1[0], pos[jy] - traj1[1], pos[jz] - traj1[2]};
    float ptot_i = sqrtf(mom[jx] * mom[jx] + mom[jy] * mom[jy] + mom[jz] * mom[jz]);
    float ptot_b = ptot_i;
    if (pos_b[2] > FMAG_HOLE_LENGTH || pos_b[0] * pos_b[0] + pos_b[1] * pos_b[1] > FMAG_HOLE_RADIUS) {
        const double lin = quirk2 ? DEDX_UNIT_2 : DEDX_UNIT_1;
        ptot_b += (DEDX_UNIT_0 + lin * ptot_i + DEDX_UNIT_2 * ptot_i * ptot_i + DEDX_UNIT_3 * ptot_i * ptot_i * ptot_i +
                   DEDX_UNIT_4 * ptot_i * ptot_i * ptot_i * ptot_i) *
                  sqrtf(traj1[0] * traj1[0] + traj1[1] * traj1[1] + traj1[2] * traj1[2]);
    }
    float traj2[3] = {(float)(tx_f * STEP_FMAG), (float)(ty * STEP_FMAG), (float)STEP_FMAG};
    pos[ix] = pos_b[0] - traj2[0];
    pos[iy] = pos_b[1] - traj2[1];
    pos[iz] = pos_b[2] - traj2[2];
    float ptot_f = ptot_b;
    if (pos[iz] > FMAG_HOLE_LENGTH || pos[ix] * pos[ix] + pos[iy] * pos[iy] > FMAG_HOLE_RADIUS) {
        ptot_f += (DEDX_UNIT_0 + DEDX_UNIT_1 * ptot_b + DEDX_UNIT_2 * ptot_b * ptot_b +
                   DEDX_UNIT_3 * ptot_b * ptot_b * ptot_b + DEDX_UNIT_4 * ptot_b * ptot_b * ptot_b * ptot_b) *
                  sqrtf(traj2[0] * traj2[0] + traj2[1] * traj2[1] + traj2[2] * traj2[2]);
    }
    float pz_f = ptot_f / sqrtf(1.f + tx_f * tx_f + ty * ty);
    mom[ix] = pz_f * tx_f;
    mom[iy] = pz_f * ty;
    mom[iz] = pz_f;
}

void target_step(float* pos, float* mom, int step, float ty) {
    const int ix = step * 3, iy = ix + 1, iz = ix + 2, jx = ix - 3, jy = ix - 2, jz = ix - 1;
    float tx_i = mom[jx] / mom[jz];
    float traj1[3] = {(float)(tx_i * STEP_TARGET), (float)(ty * STEP_TARGET), (float)STEP_TARGET};
    mom[ix] = mom[jx];
    mom[iy] = mom[jy];
    mom[iz] = mom[jz];
    pos[ix] = pos[jx] - traj1[0];
    pos[iy] = pos[jy] - traj1[1];
    pos[iz] = pos[jz] - traj1[2];
}

float closest_approach(float* v, float* p_pos, float* p_neg, float charge_1, float charge_2, float ty_1, float ty_2,
                       float* pos1, float* mom1, float* pos2, float* mom2) {
    int step = 1;
    for (; step <= NSTEPS_FMAG; ++step) {
        fmag_step(pos1, mom1, step, charge_1, ty_1, false);
        fmag_step(pos2, mom2, step, charge_2, ty_2, true);
    }
    for (; step <= NS; ++step) {
        target_step(pos1, mom1, step, ty_1);
        target_step(pos2, mom2, step, ty_2);
    }
    float dca2_min = 1.e18;
    step = -1;
    for (int j = 0; j <= NS; j++) {
        const int ix = j * 3, iy = ix + 1;
        if (FMAGSTR * charge_1 * mom1[ix] < 0. && FMAGSTR * charge_2 * mom2[ix] < 0.) continue;
        float dca2 = (pos1[ix] - pos2[ix]) * (pos1[ix] - pos2[ix]) + (pos1[iy] - pos2[iy]) * (pos1[iy] - pos2[iy]);
        if (dca2 < dca2_min) {
            dca2_min = dca2;
            step = j;
        }
    }
    if (step == -1) {
        // element NSTEPS_FMAG + NSTEPS_TARGET of the flat arrays, for all three coordinates
        v[0] = (pos1[NS] + pos2[NS]) * 0.5;
        v[1] = (pos1[NS] + pos2[NS]) * 0.5;
        v[2] = (pos1[NS] + pos2[NS]) * 0.5;
        return Z_DUMP;
    }
    const int ix = step * 3, iy = ix + 1, iz = ix + 2;
    float* a = charge_1 > 0 ? mom1 : mom2;
    float* b = charge_1 > 0 ? mom2 : mom1;
    p_pos[0] = a[ix]; p_pos[1] = a[iy]; p_pos[2] = a[iz];
    p_neg[0] = b[ix]; p_neg[1] = b[iy]; p_neg[2] = b[iz];
    v[0] = (pos1[ix] + pos2[ix]) * 0.5;
    v[1] = (pos1[iy] + pos2[iy]) * 0.5;
    v[2] = (pos1[iz] + pos2[iz]) * 0.5;
    return pos1[iz];
}

struct Trk {
    float x0, tx, y0, ty, invp, charge, chisq;
    int nhits, det_first;
    float vtx[6];
};

// starting state of the swim at the FMAG end (reconstruction_kernels.cuh:2953-2977)
void start_state(const Trk& T, const float* z_array, float* pos, float* mom) {
    const short q = (short)T.charge;
    const float tx_st1 = T.tx + PT_KICK_KMAG * T.invp * q;                       // calculate_x0_tx_st1
    const float x0_st1 = T.tx * Z_KMAG_BEND + T.x0 - tx_st1 * Z_KMAG_BEND;
    const float z_0 = z_array[T.det_first];
    float s[5];
    s[0] = T.charge * T.invp * sqrtf((1.f + tx_st1 * tx_st1) / (1.f + tx_st1 * tx_st1 + T.ty * T.ty));
    s[1] = tx_st1;
    s[2] = T.ty;
    s[3] = z_0 * tx_st1 + x0_st1;                                                // x_trk
    s[4] = z_0 * T.ty + T.y0;                                                    // y_trk
    pos[0] = s[3] + s[1] * (FMAG_LENGTH - z_0);
    pos[1] = s[4] + s[2] * (FMAG_LENGTH - z_0);
    pos[2] = FMAG_LENGTH;
    float pz_0 = 1.f / (std::fabs(s[0]) * sqrtf(1.f + s[1] * s[1] + s[2] * s[2]));
    mom[0] = pz_0 * s[1];
    mom[1] = pz_0 * s[2];
    mom[2] = pz_0;
}

// gKernel_DimuonBuilding for one event; tracks in slot order with their slots
int build_event(const std::vector<Trk>& trk, const std::vector<int>& slot, const float* z_array, float* out,
                int* lost) {
    const float p_beam[4] = {0, 0, Pbeam, Ebeam};
    const float p_target[4] = {0, 0, 0, Mp};
    const float p_cms[4] = {0, 0, Pbeam, Ebeam + Mp};
    const float s = M2(p_cms);
    const float sqrt_s = sqrtf(s);
    const float cms_bv[3] = {0.0, 0.0, p_cms[2] / p_cms[3]};
    float p_pos[4] = {0, 0, 0, 0}, p_neg[4] = {0, 0, 0, 0}, p_pos_single[4], p_neg_single[4];
    float v_1[3], v_2[3], dim_v[3], pos_v[3] = {0, 0, 0}, neg_v[3] = {0, 0, 0}, p_sum[4];
    static thread_local std::vector<float> pa1(3 * (NS + 1)), ma1(3 * (NS + 1)), pa2(3 * (NS + 1)), ma2(3 * (NS + 1));
    int ndim = 0;
    const int n = (int)trk.size();
    for (int it = 0; it < 32; ++it) {
        for (int jt = it; jt < 32; ++jt) {
            for (int i = 0; i < n; ++i) {
                if (slot[i] != it) continue;
                const Trk& A = trk[i];
                if (A.charge > 0) {
                    p_pos[0] = A.vtx[3]; p_pos[1] = A.vtx[4]; p_pos[2] = A.vtx[5]; p_pos[3] = E(Mmu, p_pos);
                    pos_v[0] = A.vtx[0]; pos_v[1] = A.vtx[1]; pos_v[2] = A.vtx[2];
                } else {
                    p_neg[0] = A.vtx[3]; p_neg[1] = A.vtx[4]; p_neg[2] = A.vtx[5]; p_neg[3] = E(Mmu, p_neg);
                    neg_v[0] = A.vtx[0]; neg_v[1] = A.vtx[1]; neg_v[2] = A.vtx[2];
                }
                for (int j = 0; j < n; ++j) {
                    if (slot[j] != jt) continue;
                    const Trk& B = trk[j];
                    if (B.charge * A.charge > 0) continue;
                    if (B.charge > 0) {
                        p_pos[0] = B.vtx[3]; p_pos[1] = B.vtx[4]; p_pos[2] = B.vtx[5]; p_pos[3] = E(Mmu, p_pos);
                        pos_v[0] = B.vtx[0]; pos_v[1] = B.vtx[1]; pos_v[2] = B.vtx[2];
                    } else {
                        p_neg[0] = B.vtx[3]; p_neg[1] = B.vtx[4]; p_neg[2] = B.vtx[5]; p_neg[3] = E(Mmu, p_neg);
                        neg_v[0] = B.vtx[0]; neg_v[1] = B.vtx[1]; neg_v[2] = B.vtx[2];
                    }
                    for (int k = 0; k < 4; ++k) { p_pos_single[k] = p_pos[k]; p_neg_single[k] = p_neg[k]; }
                    for (int k = 0; k < 3; ++k) v_1[k] = pos_v[k] + neg_v[k];
                    for (int k = 0; k < 3; ++k) v_1[k] *= 0.5f;
                    start_state(A, z_array, pa1.data(), ma1.data());
                    start_state(B, z_array, pa2.data(), ma2.data());
                    const float vz_2 = closest_approach(v_2, p_pos, p_neg, A.charge, B.charge, A.ty, B.ty, pa1.data(),
                                                        ma1.data(), pa2.data(), ma2.data());
                    p_pos[3] = E(Mmu, p_pos);
                    p_neg[3] = E(Mmu, p_neg);
                    if (std::fabs(vz_2 - v_2[2]) < STEP_FMAG) {
                        for (int k = 0; k < 4; ++k) p_sum[k] = p_pos[k] + p_neg[k];
                        for (int k = 0; k < 3; ++k) dim_v[k] = v_2[k];
                    } else {
                        for (int k = 0; k < 4; ++k) p_sum[k] = p_pos_single[k] + p_neg_single[k];
                        for (int k = 0; k < 3; ++k) dim_v[k] = v_1[k];
                    }
                    const float m = Mass(p_sum);
                    const float pT = Perp(p_sum);
                    const float x1 = mult4vec(p_target, p_sum) / mult4vec(p_target, p_cms);
                    const float x2 = mult4vec(p_beam, p_sum) / mult4vec(p_beam, p_cms);
                    Boost(p_sum, cms_bv[0] * -1, cms_bv[1] * -1, cms_bv[2] * -1);
                    const float xF = 2.0f * p_sum[2] / (sqrt_s * (1.f - m * m / s));
                    const float costheta = 2 * (p_neg[3] * p_pos[2] - p_neg[2] * p_pos[3]) / (m * sqrtf(m * m + pT * pT));
                    const float phi = (float)std::atan2(
                        (double)(2 * sqrtf(m * m + pT * pT) * (p_neg[0] * p_pos[1] - p_neg[1] * p_pos[0])),
                        (double)(m * (p_pos[0] * p_pos[0] - p_neg[0] * p_neg[0] + p_pos[1] * p_pos[1] -
                                      p_neg[1] * p_neg[1])));
                    if (FMAGSTR * (p_pos[0] - p_neg[0]) < 0.) continue;
                    if (m < 0 || m > 10.f) continue;
                    if (std::fabs(xF) > 1.f) continue;
                    if (x1 < 0.f || x1 > 1.f) continue;
                    if (x2 < 0.f || x2 > 1.f) continue;
                    if (std::fabs(costheta) > 1.f) continue;
                    if (p_pos[2] + p_neg[2] > 120.f || p_pos[2] + p_neg[2] < 30.f) continue;
                    if (std::fabs(p_pos[0] + p_neg[0]) > 3.f) continue;
                    if (dim_v[0] * dim_v[0] + dim_v[1] * dim_v[1] > 225.f) continue;
                    if (ndim >= DIM_MAX) {
                        ++*lost;
                        continue;
                    }
                    float* o = out + ndim * NDIM_PARAM;
                    o[0] = m; o[1] = pT; o[2] = xF; o[3] = x1; o[4] = x2; o[5] = costheta; o[6] = phi;
                    o[7] = 0.f; o[8] = 0.f;
                    for (int k = 0; k < 3; ++k) { o[9 + k] = dim_v[k]; o[12 + k] = pos_v[k]; o[15 + k] = neg_v[k]; }
                    for (int k = 0; k < 4; ++k) {
                        o[18 + k] = p_pos[k]; o[22 + k] = p_neg[k]; o[26 + k] = p_pos_single[k]; o[30 + k] = p_neg_single[k];
                    }
                    o[34] = 0.f; o[35] = 0.f;
                    ++ndim;
                }
            }
        }
    }
    return ndim;
}

}  // namespace

extern "C" {

// Dimuons and display histograms of a chain result (full chain, vertex set).
// tracks: n_events x max_out or_track; vertex: n_events x max_out x 6; det and
// offsets of the input hits; z_array: plane z (55); trigger: per event, may be
// NULL (all 0).  Outputs: n_dim[n_events], dim[n_events][16][36], hist[11*18*256]
// (accumulated), *lost = dimuons beyond 16 in an event.
void orc_dimuons(int64_t n_events, int max_out, const int32_t* n_tracks, const uint32_t* status, const or_track* tracks,
                 const float* vertex, const int16_t* det, const int64_t* offsets, const float* z_array,
                 const int32_t* trigger, int32_t* n_dim, float* dim, float* hist, int64_t* lost_out) {
    // bins: online_reconstruction.cu:670-726 with reconstruction_constants.h:216-263
    const short nb[NVARS] = {100, 100, 100, 100, 100, 200, 50, 50, 50, 50, 50, 50, 100, 100, 100, 100, 100, 200};
    const float vmin[NVARS] = {-20.0f, -20.0f, 0.0f, -20.0f, -20.0f, -500.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
                               -3.141592654f, -20.0f, -20.0f, 0.0f, -20.0f, -20.0f, -500.0f};
    const float vmax[NVARS] = {20.0f, 20.0f, 120.0f, 20.0f, 20.0f, 500.0f, 10.0f, 1.0f, 1.0f, 0.5f, 5.0f,
                               3.141592654f, 20.0f, 20.0f, 120.0f, 20.0f, 20.0f, 500.0f};
    std::vector<float> hw(NVARS), xpts(NVARS * NBINS);
    for (int k = 0; k < NVARS; ++k) hw[k] = (vmax[k] - vmin[k]) * 0.5f / nb[k];
    for (int k = 0; k < NVARS; ++k)
        for (int l = 0; l < nb[k]; ++l) xpts[k * NBINS + l] = vmin[k] + hw[k] * (2.f * l + 1.f);
    auto fill = [&](int ntrg, int k, float v) {
        for (int j = 0; j < nb[k]; ++j) {
            const float x = xpts[k * NBINS + j];
            if (x - hw[k] <= v && v < x + hw[k]) hist[(ntrg * NVARS + k) * NBINS + j] += 1.f;
        }
    };
    int64_t lost = 0;
    for (int64_t e = 0; e < n_events; ++e) {
        n_dim[e] = 0;
        if (status[e] & OR_EV_REJECT_MASK) continue;
        std::vector<Trk> trk;
        std::vector<int> slot;
        for (int t = 0; t < n_tracks[e]; ++t) {
            const or_track& R = tracks[e * max_out + t];
            if (R.station_id < 7) continue;
            Trk T;
            T.x0 = R.x0; T.tx = R.tx; T.y0 = R.y0; T.ty = R.ty; T.invp = R.invp; T.charge = (float)R.charge;
            T.chisq = R.chi2; T.nhits = R.nhits;
            // get_firsthitdetid (reconstruction_classes.cuh:517-532)
            short dmin = 30;
            float zmin = 10000.f;
            for (int h = 0; h < R.nhits; ++h) {
                const short d = det[offsets[e] + R.hit_index[h]];
                if (d > 0 && z_array[d] < zmin) { dmin = d; zmin = z_array[d]; }
            }
            T.det_first = dmin;
            for (int k = 0; k < 6; ++k) T.vtx[k] = vertex[(e * max_out + t) * 6 + k];
            trk.push_back(T);
            slot.push_back(R.slot);
        }
        int lost_e = 0;
        n_dim[e] = build_event(trk, slot, z_array, dim + e * DIM_MAX * NDIM_PARAM, &lost_e);
        lost += lost_e;
        if (!hist) continue;
        bool trig[NTRG];
        const int tb = trigger ? trigger[e] : 0;
        for (int ntrg = 0; ntrg < NTRG; ++ntrg) trig[ntrg] = tb & ntrg;      // not a bit test (kernels.cuh:3108)
        trig[NTRG - 1] = true;
        for (size_t i = 0; i < trk.size(); ++i) {
            const Trk& T = trk[i];
            if (T.chisq / (T.nhits - 5) > chi2dofmax) continue;
            const float vars[NVARS_TRK] = {T.vtx[3], T.vtx[4], T.vtx[5], T.vtx[0], T.vtx[1], T.vtx[2]};
            for (int ntrg = 0; ntrg < NTRG; ++ntrg)
                if (trig[ntrg])
                    for (int k = 0; k < NVARS_TRK; ++k) fill(ntrg, k, vars[k]);
        }
        for (int i = 0; i < n_dim[e]; ++i) {
            const float* o = dim + (e * DIM_MAX + i) * NDIM_PARAM;
            const float vars[NVARS - NVARS_TRK] = {o[0], o[2], o[3], o[4], o[1], o[6], o[18] + o[22], o[19] + o[23],
                                                   o[20] + o[24], o[9], o[10], o[11]};
            for (int ntrg = 0; ntrg < NTRG; ++ntrg)
                if (trig[ntrg])
                    for (int k = NVARS_TRK; k < NVARS; ++k) fill(ntrg, k, vars[k - NVARS_TRK]);
        }
    }
    if (lost_out) *lost_out = lost;
}

// The swim of build_event for n pairs on their own (start state, then
// closest_approach): a, b hold n rows of (x0, tx, y0, ty, invp, charge, z of the
// track's most upstream hit) for the first and the second track; out gets n rows
// of (v[3], p_pos[3], p_neg[3], z returned).  p_pos and p_neg start at 0 here.
void orc_dimuon_swim(int n, const float* a, const float* b, float* out) {
    std::vector<float> pa1(3 * (NS + 1)), ma1(3 * (NS + 1)), pa2(3 * (NS + 1)), ma2(3 * (NS + 1));
    auto trk = [](const float* r, float* z) {
        Trk T = {};
        T.x0 = r[0]; T.tx = r[1]; T.y0 = r[2]; T.ty = r[3]; T.invp = r[4]; T.charge = r[5];
        z[0] = r[6];
        T.det_first = 0;
        return T;
    };
    for (int i = 0; i < n; ++i) {
        float za[1], zb[1];
        const Trk A = trk(a + 7 * i, za), B = trk(b + 7 * i, zb);
        start_state(A, za, pa1.data(), ma1.data());
        start_state(B, zb, pa2.data(), ma2.data());
        float* o = out + 10 * i;
        float v[3], pp[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
        o[9] = closest_approach(v, pp, pn, A.charge, B.charge, A.ty, B.ty, pa1.data(), ma1.data(), pa2.data(),
                                ma2.data());
        for (int k = 0; k < 3; ++k) { o[k] = v[k]; o[3 + k] = pp[k]; o[6 + k] = pn[k]; }
    }
}

}  // extern "C"

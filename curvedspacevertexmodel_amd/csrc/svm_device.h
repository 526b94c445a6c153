// svm_device.h -- device-side data layout of a batch of spheres.
//
// Layout in HBM (one context = one batch of independent spheres):
//   * each sphere s owns cell rows [co, co+nc) and vertex rows [vo, vo+nv);
//     co and vo are multiples of TILE so a 256-row tile never spans two
//     spheres and every per-sphere reduction has a fixed chunk shape;
//   * positions, directors, velocities, forces: structure-of-arrays FP64
//     (x[], y[], z[]); positions and directors are double-buffered (ping-pong)
//     so a vertex pass can read neighbour positions while writing new ones;
//   * cell rings: int32 slot-major, ring[k*Cs + c] (k < ring capacity), so a
//     thread-per-cell walk reads one coalesced row per ring slot;
//   * per vertex: int4 vertex->cell list (vcn), int4 vertex->vertex list (vn),
//     both in the reference's slot order, plus `vln`, which packs for each
//     cell slot s the vn slots of the previous and next vertex of that cell's
//     ring (bits 4s..4s+1 and 4s+2..4s+3).  vln changes only at T1 flips.
//   * all vertex and cell indices stored in these arrays are global rows.
#pragma once
#include <stdint.h>

#define SVM_TILE 256
#define SVM_HEAP_CAP 4096
#define SVM_LIST_CAP 4096

namespace svm {

struct T1Log;
struct T1GroupLog;

struct SysParams {
    // sizes and offsets
    int32_t nc, nv, co, vo;
    int32_t ctile0, nctiles, vtile0, nvtiles;
    // model
    double R, kr, thr;
    double thr_cos;   // t1_cos_cut(thr, R): edges with a smaller cosine are not short
    uint64_t seed;
    // Brownian / active / Vicsek dynamics
    double dt, mu, T, v0, tau;
    double eta;       // Vicsek noise strength
    uint64_t step;
    // FIRE parameters
    double f_dt0, f_alpha_start, f_dt_max, f_dt_inc, f_dt_dec, f_alpha_dec, f_cutoff, f_alpha_min, f_dt_min;
    int32_t f_nmin, f_max_iter, f_stop_rms, f_pad;
    // FIRE state
    double f_dt, f_alpha, f_force_max, f_power;
    int32_t f_iter, f_nsince, f_frozen, f_active;
    // Adam (energyMinimizerAdam): parameters and state; f_frozen is the
    // minimiser's freeze flag for either minimiser
    double a_beta1, a_beta2, a_eps, a_dt, a_cutoff, a_beta1t, a_beta2t, a_force_max;
    int32_t a_iter, a_max_iter;
    // deferred FIRE velocity mixing
    double mix_alpha, mix_scaling;
    int32_t mix_pending, mix_zero;
    // reductions and diagnostics
    int32_t t1_flag, err;
    int32_t t1_overlap, t1_overflow, t1_head, t1_gen;   // dry-run bookkeeping
    int32_t t1_bskips, pad3;
    double energy_pot, energy_kin;
    long long n_t1, n_fan, n_nan, n_t1_skipped, n_t1_sequential, n_t1_group;
    long long n_seq_reason[4];   // sequential pass because: list overflow, group too large, group out of room, -
    long long n_cand_flip, n_cand_block;   // candidates seen (flippable, blocked)
    long long n_grp_members, n_grp_attempts, n_grp_visits, max_grp_visits;   // group dry runs
    long long n_cl_visits, max_cl_visits;                                      // single-candidate dry runs
};

struct Dev {
    int32_t nsys, Cs, Vs, cap;
    int32_t xcd;                  // 1: tiles dealt to XCDs in contiguous ranges (xcd_tile)
    // The lane a launch covers: a contiguous run of spheres [s0, s0+ns) with
    // vertex tiles [vt0, vt0+vtn) and cell tiles [ct0, ct0+ctn).  Launches of
    // the dynamics step may cover one lane of the batch each, on a stream of
    // their own; every other launch covers the whole batch.
    int32_t s0, ns, vt0, vtn, ct0, ctn;
    SysParams* sys;
    const int32_t* ctile_sys;   // cell tile -> sphere
    const int32_t* vtile_sys;   // vertex tile -> sphere
    // cells
    int32_t* cnn;
    int32_t* ring;
    double *A, *P, *cx, *cy, *cz, *a0, *p0;
    double* ctrig;                // cell centres' cosT, sinT, cosP, sinP: four planes of Cs
    // vertices
    double *px[2], *py[2], *pz[2];
    double *nx[2], *ny[2], *nz[2];
    double *vx, *vy, *vz;
    double *am, *av;              // Adam moment estimates: three planes of Vs each (null until Adam is set up)
    int32_t* cvlist;              // T1 commit: per-sphere region of vertices to rebuild when the LDS list overflows
    uint32_t* cvseen;             // bitmap over all vertices for that list (all zero between commits)
    double *dref, *dpart, *dout;
    double *gen, *cf;             // Voronoi model: generator points and their forces (3 planes of Cs each; lazy)  // dynamical analysis: reference cell positions (3 planes of Cs), tile partials, results
    double *fx, *fy, *fz;
    int4* vcn;
    int4* vn;
    int32_t* vln;
    // work buffers
    unsigned long long* t1mask;   // one bit per vertex row
    unsigned long long* t1fmask;  // one bit per vertex row: the flippable candidates
    int32_t* t1_vidx;             // per vertex row: its slot in its lane's flippable list (this step's candidates)
    int32_t* cand;                // ordered T1 candidates, per-sphere ranges [vo, vo+nv)
    double* cpart;                // per cell tile partial sums
    double* vpart;                // 3 per vertex tile partial sums
    double* disp;                 // AoS displacement staging (3 per vertex row)
    // T1 dry runs
    int32_t* t1_count;            // candidates appended this step (all spheres of the lane)
    int32_t* t1_list;             // candidate vertex rows
    int32_t* t1_flagged;          // per candidate: generation in which its run overlapped
    int32_t* t1_logged;           // per candidate: generation in which its run wrote a log
    int32_t* t1_lead;             // per candidate: slot of the candidate whose dry run covers it
    double t1_near;               // component distance of the dry runs, in cell sizes (0: a run per candidate)
    int32_t* t1_total;            // flippable-list length of the current step (copied before the reset)
    int32_t* t1_btotal;           // blocked-list length of the current step (copied before the reset)
    int32_t* t1_bcount;           // blocked candidates appended this step
    int32_t* t1_blist;            // blocked candidate vertex rows
    int32_t* t1_bflag;            // per blocked candidate: generation in which it was entangled
    int32_t* t1_bskip;            // per blocked candidate: its no-fourth-cell skips
    struct T1Log* t1_log;         // per-candidate result logs
    struct T1GroupLog* t1_group;  // per-sphere group-run log
    unsigned long long* vown;     // ownership stamps, one per vertex row
    unsigned long long* cown;     // ownership stamps, one per cell row
};

}  // namespace svm

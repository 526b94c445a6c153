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

struct SysParams {
    // sizes and offsets
    int32_t nc, nv, co, vo;
    int32_t ctile0, nctiles, vtile0, nvtiles;
    // model
    double R, kr, thr;
    uint64_t seed;
    // Brownian / active dynamics
    double dt, mu, T, v0, tau;
    uint64_t step;
    // FIRE parameters
    double f_dt0, f_alpha_start, f_dt_max, f_dt_inc, f_dt_dec, f_alpha_dec, f_cutoff, f_alpha_min, f_dt_min;
    int32_t f_nmin, f_max_iter, f_stop_rms, f_pad;
    // FIRE state
    double f_dt, f_alpha, f_force_max, f_power;
    int32_t f_iter, f_nsince, f_frozen, f_active;
    // deferred FIRE velocity mixing
    double mix_alpha, mix_scaling;
    int32_t mix_pending, mix_zero;
    // reductions and diagnostics
    int32_t counter_geo, counter_vtx;
    int32_t t1_flag, err;
    int32_t fan_flag, pad2;
    int32_t t1_ncand, t1_done, t1_overlap, t1_head;   // dry-run bookkeeping
    double energy_pot, energy_kin;
    long long n_t1, n_fan, n_nan, n_t1_skipped, n_t1_sequential;
};

struct Dev {
    int32_t nsys, Cs, Vs, cap;
    SysParams* sys;
    const int32_t* ctile_sys;   // cell tile -> sphere
    const int32_t* vtile_sys;   // vertex tile -> sphere
    // cells
    int32_t* cnn;
    int32_t* ring;
    double *A, *P, *cx, *cy, *cz, *a0, *p0;
    // vertices
    double *px[2], *py[2], *pz[2];
    double *nx[2], *ny[2], *nz[2];
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    int4* vcn;
    int4* vn;
    int32_t* vln;
    // work buffers
    unsigned long long* t1mask;   // one bit per vertex row
    int32_t* cand;                // ordered T1 candidates, per-sphere ranges [vo, vo+nv)
    double* cpart;                // per cell tile partial sums
    double* vpart;                // 3 per vertex tile partial sums
    double* disp;                 // AoS displacement staging (3 per vertex row)
    // T1 dry runs
    int32_t* t1_count;            // candidates appended this step (all spheres)
    uint32_t* t1_gen;             // step generation of the ownership stamps
    int32_t* t1_list;             // candidate vertex rows
    struct T1Log* t1_log;         // per-candidate result logs
    unsigned long long* vown;     // ownership stamps, one per vertex row
    unsigned long long* cown;     // ownership stamps, one per cell row
};

}  // namespace svm

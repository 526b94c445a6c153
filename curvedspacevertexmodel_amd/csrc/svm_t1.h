// svm_t1.h -- T1 topology maintenance on the device, in exactly the order of
// the reference's sequential CPU loop (testAndPerformT1TransitionsCPU,
// src/model/sphericalVertexModel.cpp:439-604): vertex1 ascending, slots 0..2,
// every test seeing all earlier flips.
//
// Launches per step:
//   k_t1_detect   thread per vertex: short-edge test on the moved positions
//                 (bit mask + compact lists of flippable and of blocked
//                 candidates).  The lists and their counters belong to the
//                 launch's lane of spheres.
//   k_t1_cluster  one wavefront per candidate vertex u: a dry run of the
//                 sequential loop restricted to u -- u's visit, then the
//                 visits of every later vertex its flips disturbed, in index
//                 order -- on a private LDS copy of the data it touches,
//                 stamping each touched vertex and cell, and logging what it
//                 would write.  When no two candidates of a sphere touched a
//                 common item, their visits commute and any order gives the
//                 sequential result.
//   k_t1_group    block per sphere.  First the sphere's blocked candidates
//                 (every short edge triangle-protected or without a fourth
//                 cell), which do nothing unless a run above touched what
//                 they read; those that were are flagged.  Then, where
//                 candidates are flagged, one dry run over all of them
//                 together, in index order.
//   k_t1_resolve  block per sphere: applies the logs of a sphere whose
//                 candidates did not overlap; otherwise (or when a buffer was
//                 too small) runs the sequential loop itself over the
//                 candidates and disturbed vertices.
// No launch needs a device-scope fence: each kernel consumes what the
// previous one wrote.
// A vertex that is neither a candidate nor disturbed by an earlier flip has
// an unchanged neighbourhood since detection, so its visit is a no-op and
// skipping it is exact.
#pragma once

namespace svm {

#define SVM_T1_MAXCL 16384    // flippable candidates handled by dry runs per step
#define SVM_T1_MAXBL 65536    // blocked candidates checked per step
#define SVM_T1_WV 128         // vertices cached per dry run
#define SVM_T1_WC 64          // cells cached per dry run
#define SVM_T1_RC 32          // ring entries per cached cell
#define SVM_T1_LV 32          // modified vertices logged per dry run
#define SVM_T1_LC 16          // modified cells logged per dry run
#define SVM_T1_WAVES 4        // dry-run wavefronts per block

struct T1Log {
    int next, nvert, ncell, pad;
    long long flips, skipped;
    int vid[SVM_T1_LV];
    int vflag[SVM_T1_LV];
    double vpos[SVM_T1_LV][3];
    int vvn[SVM_T1_LV][3];
    int vvcn[SVM_T1_LV][3];
    int cid[SVM_T1_LC];
    int ccnt[SVM_T1_LC];
    int cring[SVM_T1_LC][SVM_T1_RC];
};

__device__ inline int& ringref(const Dev& d, int c, int k) { return d.ring[(size_t)k * d.Cs + c]; }
__device__ inline int& vcnref(const Dev& d, int v, int k) { return ((int*)&d.vcn[v])[k]; }
__device__ inline int& vnref(const Dev& d, int v, int k) { return ((int*)&d.vn[v])[k]; }

#define SVM_T1_GV 1024        // vertices cached / logged by a group dry run
#define SVM_T1_GC 256         // cells cached / logged by a group dry run
#define SVM_T1_GMAX 64        // overlapping candidates one group dry run takes
#define SVM_T1_GABS 32        // foreign runs a group attempt can absorb

struct T1GroupLog {
    int nvert, ncell, ok, pad;
    long long flips, skipped;
    int vid[SVM_T1_GV];
    int vflag[SVM_T1_GV];
    double vpos[SVM_T1_GV][3];
    int vvn[SVM_T1_GV][3];
    int vvcn[SVM_T1_GV][3];
    int cid[SVM_T1_GC];
    int ccnt[SVM_T1_GC];
    int cring[SVM_T1_GC][SVM_T1_RC];
};

// Min-heap of local vertex indices (the disturbed vertices still to visit).
struct Heap {
    int* a;
    int n;
    __device__ __attribute__((always_inline)) void push(int x) {
        int i = n++;
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (a[p] <= x) break;
            a[i] = a[p];
            i = p;
        }
        a[i] = x;
    }
    __device__ __attribute__((always_inline)) int pop() {
        int top = a[0];
        int x = a[--n];
        int i = 0;
        while (true) {
            int l = 2 * i + 1;
            if (l >= n) break;
            int r = l + 1;
            int m = (r < n && a[r] < a[l]) ? r : l;
            if (a[m] >= x) break;
            a[i] = a[m];
            i = m;
        }
        if (n > 0) a[i] = x;
        return top;
    }
};

// ------------------------------------------------------------ accessors ---
// Direct access to the topology and positions in HBM.
struct GlobalAcc {
    static constexpr bool kCached = false;
    const Dev* d;
    int pcur;
    __device__ __attribute__((always_inline)) void pos(int v, double* p) { p[0] = d->px[pcur][v]; p[1] = d->py[pcur][v]; p[2] = d->pz[pcur][v]; }
    __device__ __attribute__((always_inline)) void set_pos(int v, const double* p) {
        d->px[pcur][v] = p[0]; d->py[pcur][v] = p[1]; d->pz[pcur][v] = p[2];
    }
    __device__ __attribute__((always_inline)) int vn(int v, int k) { return vnref(*d, v, k); }
    __device__ __attribute__((always_inline)) void set_vn(int v, int k, int x) { vnref(*d, v, k) = x; }
    __device__ __attribute__((always_inline)) int vcn(int v, int k) { return vcnref(*d, v, k); }
    __device__ __attribute__((always_inline)) void set_vcn(int v, int k, int x) { vcnref(*d, v, k) = x; }
    __device__ __attribute__((always_inline)) int cnn(int c) { return d->cnn[c]; }
    __device__ __attribute__((always_inline)) void set_cnn(int c, int x) { d->cnn[c] = x; }
    __device__ __attribute__((always_inline)) int ring(int c, int k) { return ringref(*d, c, k); }
    __device__ __attribute__((always_inline)) void set_ring(int c, int k, int x) { ringref(*d, c, k) = x; }
    __device__ __attribute__((always_inline)) bool failed() { return false; }
};

// Private LDS copy of the vertices and cells one dry run reads and writes
// (open addressing on the global row; WV vertex and WC cell slots).
//   * prefetch(): all lanes of the run's wavefront load the neighbourhood of
//     the seed vertices in parallel -- each seed, its three neighbours, the
//     cells of those, and every vertex of those cells' rings -- so the walk,
//     which lane 0 does alone, mostly hits LDS.  What it misses it loads on
//     first touch.
//   * Every slot the walk uses is marked touched.  Ownership is settled after
//     the walk, by all lanes in one pass over the touched slots: single-
//     candidate runs stamp them (stamp_touched), group runs check them
//     (check_touched).  Whether two runs touched a common item does not
//     depend on when during the walks they claimed it.
struct WaveSync {
    __device__ __attribute__((always_inline)) void operator()() const {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
};

template <int WV, int WC>
struct CacheAcc {
    static constexpr bool kCached = true;
    const Dev* d;
    int pcur;
    bool fail;
    int* vkey;        // [WV] global row, or -1
    int* vflag;       // [WV] 1: position written, 2: lists written
    int* vuse;        // [WV] 1: loaded, 2: touched by the walk
    double* vp;       // [WV][3]
    int* vvn;         // [WV][3]
    int* vvcn;        // [WV][3]
    int* ckey;        // [WC]
    int* cflag;       // [WC] 1: written
    int* cuse;        // [WC] 1: loaded, 2: touched, 4: ring longer than a slot holds
    int* ccnt;        // [WC]
    int* cring;       // [WC][RC]
    int* vfill;       // occupied vertex slots
    int* cfill;       // occupied cell slots
    // the walk's last two lookups of each table (key, slot), in registers
    int mvk0 = -1, mvh0 = 0, mvk1 = -1, mvh1 = 0;
    int mck0 = -1, mch0 = 0, mck1 = -1, mch1 = 0;

    static constexpr int log2i(int x) { return x <= 1 ? 0 : 1 + log2i(x >> 1); }
    __device__ __attribute__((always_inline)) static unsigned vhash(int v) { return ((unsigned)v * 2654435761u) >> (32 - log2i(WV)); }
    __device__ __attribute__((always_inline)) static unsigned chash(int c) { return ((unsigned)c * 2654435761u) >> (32 - log2i(WC)); }

    // all lanes: empty tables
    __device__ __attribute__((always_inline)) void reset(int lane, int nl) {
        for (int h = lane; h < WV; h += nl) { vkey[h] = -1; vuse[h] = 0; }
        for (int h = lane; h < WC; h += nl) { ckey[h] = -1; cuse[h] = 0; }
        if (lane == 0) { *vfill = 0; *cfill = 0; }
    }
    __device__ __attribute__((always_inline)) void load_v(int h, int v) {
        vp[3 * h] = d->px[pcur][v];
        vp[3 * h + 1] = d->py[pcur][v];
        vp[3 * h + 2] = d->pz[pcur][v];
        int4 a = d->vn[v], b = d->vcn[v];
        vvn[3 * h] = a.x; vvn[3 * h + 1] = a.y; vvn[3 * h + 2] = a.z;
        vvcn[3 * h] = b.x; vvcn[3 * h + 1] = b.y; vvcn[3 * h + 2] = b.z;
        vflag[h] = 0;
    }
    // sets cuse[h] to 1 (loaded) or 4 (too long)
    __device__ __attribute__((always_inline)) void load_c(int h, int c) {
        int n = d->cnn[c];
        cflag[h] = 0;
        if (n > SVM_T1_RC - 1) { ccnt[h] = 0; cuse[h] = 4; return; }
        ccnt[h] = n;
        for (int k = 0; k < n; ++k) cring[h * SVM_T1_RC + k] = ringref(*d, c, k);
        cuse[h] = 1;
    }

    // ---- the walk (lane 0 only)
    // Every key in a table is loaded (prefetch loads what it inserts), so a
    // hit only marks the slot touched.
    __device__ __attribute__((always_inline)) int vslot(int v) {
        if (v == mvk0) return mvh0;
        if (v == mvk1) return mvh1;
        unsigned h = vhash(v);
        int found = -1;
        for (int probe = 0; probe < WV; ++probe, h = (h + 1) & (WV - 1)) {
            const int k = vkey[h];
            if (k == v) {
                vuse[h] = 3;
                found = (int)h;
                break;
            }
            if (k < 0) {
                vkey[h] = v;
                ++*vfill;
                load_v(h, v);
                vuse[h] = 3;
                found = (int)h;
                break;
            }
        }
        if (found < 0) { fail = true; return 0; }
        mvk1 = mvk0; mvh1 = mvh0;
        mvk0 = v; mvh0 = found;
        return found;
    }
    __device__ __attribute__((always_inline)) int cslot(int c) {
        if (c == mck0) return mch0;
        if (c == mck1) return mch1;
        unsigned h = chash(c);
        int found = -1;
        for (int probe = 0; probe < WC; ++probe, h = (h + 1) & (WC - 1)) {
            const int k = ckey[h];
            if (k == c) {
                found = (int)h;
                break;
            }
            if (k < 0) {
                ckey[h] = c;
                ++*cfill;
                load_c(h, c);
                found = (int)h;
                break;
            }
        }
        if (found < 0 || (cuse[found] & 4)) { fail = true; return 0; }
        cuse[found] = 3;
        mck1 = mck0; mch1 = mch0;
        mck0 = c; mch0 = found;
        return found;
    }
    // forget the memo (after restore(), the slots are untouched again)
    __device__ __attribute__((always_inline)) void forget() { mvk0 = mvk1 = mck0 = mck1 = -1; }
    __device__ __attribute__((always_inline)) void pos(int v, double* p) {
        if (v != mvk0 && v != mvk1) {
            // the first probe's key and position in one round of loads
            const unsigned h = vhash(v);
            const int k = vkey[h];
            const double a0 = vp[3 * h], a1 = vp[3 * h + 1], a2 = vp[3 * h + 2];
            if (k == v) {
                vuse[h] = 3;
                mvk1 = mvk0; mvh1 = mvh0;
                mvk0 = v; mvh0 = (int)h;
                p[0] = a0; p[1] = a1; p[2] = a2;
                return;
            }
        }
        int h = vslot(v);
        p[0] = vp[3 * h]; p[1] = vp[3 * h + 1]; p[2] = vp[3 * h + 2];
    }
    // Slots of N vertices, and of N cells with their counts and first ring
    // entries, each in one round of LDS loads: the first probe of every key is
    // read together; a key found there is marked touched (as vslot / cslot
    // would), any other goes through vslot / cslot, in argument order.  A
    // lookup only ever fills an empty slot, so a key found in the round is
    // where the sequential lookups would have found it.
    template <int N>
    __device__ __attribute__((always_inline)) void vh_n(const int (&keys)[N], int (&hs)[N]) {
        unsigned h[N];
        int k[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            h[i] = vhash(keys[i]);
            k[i] = vkey[h[i]];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (k[i] == keys[i]) {
                vuse[h[i]] = 3;
                hs[i] = (int)h[i];
            } else {
                hs[i] = vslot(keys[i]);
            }
        }
    }
    template <int N, int RR>
    __device__ __attribute__((always_inline)) void ch_ring_n(const int (&keys)[N], int (&hs)[N], int (&ns)[N],
                                                             int (&e)[N][RR]) {
        unsigned h[N];
        int k[N], u[N], c[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            h[i] = chash(keys[i]);
            k[i] = ckey[h[i]];
            u[i] = cuse[h[i]];
            c[i] = ccnt[h[i]];
#pragma unroll
            for (int j = 0; j < RR; ++j) e[i][j] = cring[h[i] * SVM_T1_RC + j];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (k[i] == keys[i] && !(u[i] & 4)) {
                cuse[h[i]] = 3;
                hs[i] = (int)h[i];
                ns[i] = c[i];
            } else {
                hs[i] = cslot(keys[i]);
                ns[i] = ccnt[hs[i]];
#pragma unroll
                for (int j = 0; j < RR; ++j) e[i][j] = cring[hs[i] * SVM_T1_RC + j];
            }
        }
    }
    __device__ __attribute__((always_inline)) void set_pos(int v, const double* p) {
        int h = vslot(v);
        vp[3 * h] = p[0]; vp[3 * h + 1] = p[1]; vp[3 * h + 2] = p[2];
        vflag[h] |= 1;
    }
    __device__ __attribute__((always_inline)) int vn(int v, int k) { return vvn[3 * vslot(v) + k]; }
    __device__ __attribute__((always_inline)) void set_vn(int v, int k, int x) { int h = vslot(v); vvn[3 * h + k] = x; vflag[h] |= 2; }
    __device__ __attribute__((always_inline)) int vcn(int v, int k) { return vvcn[3 * vslot(v) + k]; }
    __device__ __attribute__((always_inline)) void set_vcn(int v, int k, int x) { int h = vslot(v); vvcn[3 * h + k] = x; vflag[h] |= 2; }
    __device__ __attribute__((always_inline)) int cnn(int c) { return ccnt[cslot(c)]; }
    __device__ __attribute__((always_inline)) void set_cnn(int c, int x) {
        int h = cslot(c);
        if (x > SVM_T1_RC - 1) fail = true;
        else { ccnt[h] = x; cflag[h] = 1; }
    }
    __device__ __attribute__((always_inline)) int ring(int c, int k) { return cring[cslot(c) * SVM_T1_RC + (k & (SVM_T1_RC - 1))]; }
    __device__ __attribute__((always_inline)) void set_ring(int c, int k, int x) {
        int h = cslot(c);
        cring[h * SVM_T1_RC + (k & (SVM_T1_RC - 1))] = x;
        cflag[h] = 1;
    }
    __device__ __attribute__((always_inline)) bool failed() { return fail; }
    // slot handles: one lookup (which marks the slot touched), then direct access
    __device__ __attribute__((always_inline)) int vh(int v) { return vslot(v); }
    __device__ __attribute__((always_inline)) int ch(int c) { return cslot(c); }
    __device__ __attribute__((always_inline)) double* vpos_ptr(int h) { return vp + 3 * h; }
    __device__ __attribute__((always_inline)) int* vn_ptr(int h) { return vvn + 3 * h; }
    __device__ __attribute__((always_inline)) int* vcn_ptr(int h) { return vvcn + 3 * h; }
    __device__ __attribute__((always_inline)) int* ring_ptr(int h) { return cring + h * SVM_T1_RC; }
    __device__ __attribute__((always_inline)) int cnn_h(int h) { return ccnt[h]; }
    __device__ __attribute__((always_inline)) void mark_v(int h, int f) { vflag[h] |= f; }
    __device__ __attribute__((always_inline)) void set_cnn_h(int h, int x) {
        if (x > SVM_T1_RC - 1) fail = true;
        else { ccnt[h] = x; cflag[h] = 1; }
    }

    // ---- cooperative parts (every lane of the wavefront)
    // Insert `key` from many lanes at once.  New entries are limited to
    // `limit` occupied slots in all, so that prefetching leaves the walk room.
    template <int W>
    __device__ __attribute__((always_inline)) static void par_insert(int* keys, unsigned h, int key, int* fill, int limit) {
        for (int probe = 0; probe < W; ++probe, h = (h + 1) & (W - 1)) {
            int k = __hip_atomic_load(keys + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (k == key) return;
            if (k < 0) {
                if (atomicAdd(fill, 1) >= limit) { atomicSub(fill, 1); return; }
                int old = atomicCAS(keys + h, -1, key);
                if (old < 0) return;
                atomicSub(fill, 1);
                if (old == key) return;
            }
        }
    }
    // Load the neighbourhood of the seed vertices (local indices seeds[i] +
    // vo), within the budgets vlimit / climit of occupied slots.
    template <class Sync>
    __device__ __attribute__((always_inline)) void prefetch(const int* seeds, int n, int vo, int lane, int nl, int vlimit, int climit, Sync sync) {
        for (int i = lane; i < n; i += nl) {
            const int u = vo + seeds[i];
            par_insert<WV>(vkey, vhash(u), u, vfill, vlimit);
            int4 nb = d->vn[u];
            par_insert<WV>(vkey, vhash(nb.x), nb.x, vfill, vlimit);
            par_insert<WV>(vkey, vhash(nb.y), nb.y, vfill, vlimit);
            par_insert<WV>(vkey, vhash(nb.z), nb.z, vfill, vlimit);
        }
        sync();
        for (int h = lane; h < WV; h += nl) {
            const int v = vkey[h];
            if (v < 0 || (vuse[h] & 1)) continue;
            load_v(h, v);
            vuse[h] = 1;
            for (int k = 0; k < 3; ++k) {
                const int c = vvcn[3 * h + k];
                par_insert<WC>(ckey, chash(c), c, cfill, climit);
            }
        }
        sync();
        for (int h = lane; h < WC; h += nl) {
            const int c = ckey[h];
            if (c < 0 || cuse[h]) continue;
            load_c(h, c);
            for (int k = 0; k < ccnt[h]; ++k) {
                const int x = cring[h * SVM_T1_RC + k];
                par_insert<WV>(vkey, vhash(x), x, vfill, vlimit);
            }
        }
        sync();
        for (int h = lane; h < WV; h += nl) {
            const int v = vkey[h];
            if (v < 0 || (vuse[h] & 1)) continue;
            load_v(h, v);
            vuse[h] = 1;
        }
        sync();
    }
    // Single-candidate runs: stamp every touched item with (generation, id)
    // by an atomic max; meeting another run's stamp of the same generation
    // flags both runs as overlapping.
    __device__ __attribute__((always_inline)) void stamp_touched(unsigned long long stamp, int32_t* flagged, int32_t* overlap, int lane, int nl) {
        const unsigned gen = (unsigned)(stamp >> 32);
        for (int h = lane; h < WV + WC; h += nl) {
            unsigned long long* own;
            if (h < WV) {
                if (vkey[h] < 0 || !(vuse[h] & 2)) continue;
                own = d->vown + vkey[h];
            } else {
                const int j = h - WV;
                if (ckey[j] < 0 || !(cuse[j] & 2)) continue;
                own = d->cown + ckey[j];
            }
            unsigned long long old = atomicMax(own, stamp);
            if ((unsigned)(old >> 32) == gen && old != stamp) {
                *overlap = 1;
                flagged[(int)(stamp & 0xffffffffu) - 1] = (int)gen;
                flagged[(int)(old & 0xffffffffu) - 1] = (int)gen;
            }
        }
    }
    // Group runs: list (up to amax, count in *nabs) the runs that touched what
    // the group touched and whose results would be committed (not flagged).
    __device__ __attribute__((always_inline)) void check_touched(int gen, const int32_t* flagged, int* absorb, int* nabs, int amax, int lane,
                                  int nl) {
        for (int h = lane; h < WV + WC; h += nl) {
            const unsigned long long* own;
            if (h < WV) {
                if (vkey[h] < 0 || !(vuse[h] & 2)) continue;
                own = d->vown + vkey[h];
            } else {
                const int j = h - WV;
                if (ckey[j] < 0 || !(cuse[j] & 2)) continue;
                own = d->cown + ckey[j];
            }
            unsigned long long st = *own;
            if ((int)(st >> 32) != gen) continue;
            const int o = (int)(st & 0xffffffffu) - 1;
            if (flagged[o] == gen) continue;
            int k = atomicAdd(nabs, 1);
            if (k < amax) absorb[k] = o;
        }
    }
    // Undo the walk's writes and touches (the loaded copies stay).
    __device__ __attribute__((always_inline)) void restore(int lane, int nl) {
        forget();
        for (int h = lane; h < WV; h += nl) {
            if (vkey[h] < 0) continue;
            if (vflag[h]) load_v(h, vkey[h]);
            vuse[h] &= 1;
        }
        for (int h = lane; h < WC; h += nl) {
            if (ckey[h] < 0) continue;
            if (cflag[h]) load_c(h, ckey[h]);
            else cuse[h] &= ~2;
        }
    }
    // Copy the modified items into a log (all lanes; order of entries is
    // free since the items are distinct).  False if they do not fit.
    template <int LV, int LC, class Sync>
    __device__ __attribute__((always_inline)) bool write_log(int* lvid, int* lvflag, double (*lvpos)[3], int (*lvvn)[3], int (*lvvcn)[3], int* lcid,
                              int* lccnt, int (*lcring)[SVM_T1_RC], int* nvert, int* ncell, int lane, int nl,
                              Sync sync) {
        if (lane == 0) { *nvert = 0; *ncell = 0; }
        sync();
        for (int h = lane; h < WV; h += nl) {
            if (vkey[h] < 0 || !vflag[h]) continue;
            const int i = atomicAdd(nvert, 1);
            if (i >= LV) continue;
            lvid[i] = vkey[h];
            lvflag[i] = vflag[h];
            for (int k = 0; k < 3; ++k) {
                lvpos[i][k] = vp[3 * h + k];
                lvvn[i][k] = vvn[3 * h + k];
                lvvcn[i][k] = vvcn[3 * h + k];
            }
        }
        for (int h = lane; h < WC; h += nl) {
            if (ckey[h] < 0 || !cflag[h]) continue;
            const int i = atomicAdd(ncell, 1);
            if (i >= LC) continue;
            lcid[i] = ckey[h];
            lccnt[i] = ccnt[h];
            for (int k = 0; k < ccnt[h]; ++k) lcring[i][k] = cring[h * SVM_T1_RC + k];
        }
        sync();
        return *nvert <= LV && *ncell <= LC;
    }
};

// -------------------------------------------------------- T1 logic --------
// getCellVertexSetForT1 (src/model/sphericalVertexModel.cpp:611-731).
// Returns false when the four cells cannot all be determined (vertex1 and
// vertex2 share all three cells, so there is no cell l); the flip is then
// skipped, where the reference would read an uninitialised cell set.
template <class A>
__device__ __attribute__((always_inline)) bool cell_vertex_set(A& a, int vertex1, int vertex2, int* cs, int* vs) {
    cs[0] = cs[1] = cs[2] = cs[3] = -1;
    int cell1 = a.vcn(vertex1, 0), cell2 = a.vcn(vertex1, 1), cell3 = a.vcn(vertex1, 2);
    for (int ff = 0; ff < 3; ++ff) {
        int ctest = a.vcn(vertex2, ff);
        if (ctest != cell1 && ctest != cell2 && ctest != cell3) cs[3] = ctest;
    }
    int cells[3] = {cell1, cell2, cell3};
    int vlast, vcur, vnext = -1, cneigh;
    for (int t = 0; t < 3; ++t) {
        int cell = cells[t];
        cneigh = a.cnn(cell);
        vlast = a.ring(cell, cneigh - 2);
        vcur = a.ring(cell, cneigh - 1);
        for (int cn = 0; cn < cneigh; ++cn) {
            vnext = a.ring(cell, cn);
            if (vcur == vertex1) break;
            vlast = vcur;
            vcur = vnext;
        }
        if (vlast == vertex2) cs[0] = cell;
        else if (vnext == vertex2) cs[2] = cell;
        else cs[1] = cell;
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return false;
    cneigh = a.cnn(cs[1]);
    vlast = a.ring(cs[1], cneigh - 2);
    vcur = a.ring(cs[1], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = a.ring(cs[1], cn);
        if (vcur == vertex1) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[0] = vlast;
    vs[1] = vnext;
    cneigh = a.cnn(cs[3]);
    vlast = a.ring(cs[3], cneigh - 2);
    vcur = a.ring(cs[3], cneigh - 1);
    for (int cn = 0; cn < cneigh; ++cn) {
        vnext = a.ring(cs[3], cn);
        if (vcur == vertex2) break;
        vlast = vcur;
        vcur = vnext;
    }
    vs[3] = vlast;
    vs[2] = vnext;
    return true;
}

// The new positions of the flipped edge's vertices v1, v2 (in place): the
// edge turned by a quarter turn about its normalised midpoint, its vertices
// then put back on the sphere (sphericalVertexModel.cpp:506-527); vC is vertex
// c, whose side fixes the sense of the turn.
__device__ __attribute__((always_inline)) inline void flip_positions(double* v1, double* v2, const double* vC,
                                                                     double R) {
    double mid[3] = {0.5 * (v1[0] + v2[0]), 0.5 * (v1[1] + v2[1]), 0.5 * (v1[2] + v2[2])};
    double mm = 0.0; mm += mid[0] * mid[0]; mm += mid[1] * mid[1]; mm += mid[2] * mid[2];
    double im = 1.0 / svm::sqrt_rn(mm);
    mid[0] = im * mid[0]; mid[1] = im * mid[1]; mid[2] = im * mid[2];
    double determinant = vC[0] * (v1[1] * v2[2] - v1[2] * v2[1])
                       + vC[1] * (v1[2] * v2[0] - v1[0] * v2[2])
                       + vC[2] * (v1[0] * v2[1] - v1[1] * v2[0]);
    // the turn by -0.5 * (determinant > 0 ? 1 : -1) * pi: cos and sin of -+pi/2
    const double s = determinant > 0 ? -1.0 : 1.0;
    rodrigues_cs(v1, mid, kCosQuarterTurn, s);
    rodrigues_cs(v2, mid, kCosQuarterTurn, s);
    double diff[3] = {0.5 * (v1[0] - v2[0]), 0.5 * (v1[1] - v2[1]), 0.5 * (v1[2] - v2[2])};
    for (int k = 0; k < 3; ++k) { v1[k] = v1[k] + diff[k]; v2[k] = v2[k] - diff[k]; }
    double n1 = 0.0; n1 += v1[0] * v1[0]; n1 += v1[1] * v1[1]; n1 += v1[2] * v1[2];
    double n2 = 0.0; n2 += v2[0] * v2[0]; n2 += v2[1] * v2[1]; n2 += v2[2] * v2[2];
    double i1 = 1.0 / svm::sqrt_rn(n1), i2 = 1.0 / svm::sqrt_rn(n2);
    for (int k = 0; k < 3; ++k) { v1[k] = R * (i1 * v1[k]); v2[k] = R * (i2 * v2[k]); }
}

// One T1 flip of edge (vertex1, vertex2): the body of
// testAndPerformT1TransitionsCPU (src/model/sphericalVertexModel.cpp:496-598).
// v1 is the caller's copy of vertex1's position.  Returns 0, or
// SVM_ERR_RING_OVERFLOW (nothing modified) when a cell would outgrow `cap`.
template <class A>
__device__ __attribute__((always_inline)) int t1_flip(A& a, int vertex1, int vertex2, const int* cs, const int* vs, double* v1, double R,
                       int cap, int* tmp) {
    if (a.cnn(cs[1]) + 1 > cap || a.cnn(cs[3]) + 1 > cap) return 4;
    double v2[3], vC[3];
    a.pos(vertex2, v2);
    a.pos(vs[2], vC);
    flip_positions(v1, v2, vC, R);
    a.set_pos(vertex1, v1);
    a.set_pos(vertex2, v2);
    for (int vert = 0; vert < 3; ++vert) {
        if (a.vcn(vertex1, vert) == cs[2]) a.set_vcn(vertex1, vert, cs[3]);
        if (a.vcn(vertex2, vert) == cs[0]) a.set_vcn(vertex2, vert, cs[1]);
        if (a.vn(vs[1], vert) == vertex1) a.set_vn(vs[1], vert, vertex2);
        if (a.vn(vs[2], vert) == vertex2) a.set_vn(vs[2], vert, vertex1);
        if (a.vn(vertex1, vert) == vs[1]) a.set_vn(vertex1, vert, vs[2]);
        if (a.vn(vertex2, vert) == vs[2]) a.set_vn(vertex2, vert, vs[1]);
    }
    // cell i loses vertex2
    int cneigh = a.cnn(cs[0]);
    int cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (a.ring(cs[0], cc) == vertex2) cidx += 1;
        a.set_ring(cs[0], cc, a.ring(cs[0], cidx));
        cidx += 1;
    }
    a.set_cnn(cs[0], cneigh - 1);
    // cell j gains vertex2 between vertex1 and b
    cneigh = a.cnn(cs[1]);
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = a.ring(cs[1], cc);
        tmp[cidx++] = e;
        if (e == vertex1) tmp[cidx++] = vertex2;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) a.set_ring(cs[1], cc, tmp[cc]);
    a.set_cnn(cs[1], cneigh + 1);
    // cell k loses vertex1
    cneigh = a.cnn(cs[2]);
    cidx = 0;
    for (int cc = 0; cc < cneigh - 1; ++cc) {
        if (a.ring(cs[2], cc) == vertex1) cidx += 1;
        a.set_ring(cs[2], cc, a.ring(cs[2], cidx));
        cidx += 1;
    }
    a.set_cnn(cs[2], cneigh - 1);
    // cell l gains vertex1 between vertex2 and a
    cneigh = a.cnn(cs[3]);
    cidx = 0;
    for (int cc = 0; cc < cneigh; ++cc) {
        int e = a.ring(cs[3], cc);
        tmp[cidx++] = e;
        if (e == vertex2) tmp[cidx++] = vertex1;
    }
    for (int cc = 0; cc < cneigh + 1; ++cc) a.set_ring(cs[3], cc, tmp[cc]);
    a.set_cnn(cs[3], cneigh + 1);
    return 0;
}

// ---- cached fast paths (CacheAcc only) -----------------------------------
// The walk runs on one lane, so every dependent LDS access is a full round
// trip.  These versions of cell_vertex_set and t1_flip look each vertex and
// cell up once, read what they need in one batch into registers (rings of up
// to SVM_T1_RR entries) and only then write.  They return -1, having written
// nothing, where that batched form is not provably the same as the in-place
// sequence (repeated vertices or cells, longer rings, a vertex not exactly
// once in a ring); the caller then runs the generic code.
#define SVM_T1_RR 12

// (vlast, vnext) around `target` in the ring r[0..n) (e: its first entries in
// registers), 2 <= n <= SVM_T1_RR, as the loop of getCellVertexSetForT1 finds
// them: the first match among r[n-1], r[0], ..., r[n-2]; without a match
// (r[n-2], r[n-1]).  The search uses constant register indices (a dynamic
// index would put e in scratch); the two results are read from r.
__device__ __attribute__((always_inline)) inline void ring_nbrs(const int* r, const int (&e)[SVM_T1_RR], int n,
                                                                int target, int& vlast, int& vnext) {
    int qlast = -1, qfirst = -1;
#pragma unroll
    for (int j = 0; j < SVM_T1_RR; ++j) {
        const bool m = j < n && e[j] == target;
        qlast = (m && j == n - 1) ? j : qlast;
        qfirst = (m && j < n - 1 && qfirst < 0) ? j : qfirst;
    }
    const int q = qlast >= 0 ? qlast : qfirst;
    const int il = q < 0 ? n - 2 : (q == 0 ? n - 1 : q - 1);
    const int in = q < 0 ? n - 1 : (q == n - 1 ? 0 : q + 1);
    vlast = r[il];
    vnext = r[in];
}

__device__ __attribute__((always_inline)) inline void ring_regs(const int* r, int (&e)[SVM_T1_RR]) {
#pragma unroll
    for (int j = 0; j < SVM_T1_RR; ++j) e[j] = r[j];      // a slot holds SVM_T1_RC >= SVM_T1_RR entries
}

// Position of the only occurrence of x in e[0..n), or -1 (none or several).
__device__ __attribute__((always_inline)) inline int ring_only(const int (&e)[SVM_T1_RR], int n, int x) {
    int p = -1, cnt = 0;
#pragma unroll
    for (int j = 0; j < SVM_T1_RR; ++j)
        if (j < n && e[j] == x) { p = j; ++cnt; }
    return cnt == 1 ? p : -1;
}

// What the cached cell-set search has read, handed to the cached flip that
// follows it (nothing is written in between): the cell set's slots, ring
// lengths and first ring entries, in the order cs[0..3], and the slots of
// vertex1 and vertex2.
struct FlipRegs {
    int hc[4], n[4];
    int e[4][SVM_T1_RR];
    int h1, h2;
};

template <class A>
__device__ __attribute__((always_inline)) int cell_vertex_set_cached(A& a, int vertex1, int vertex2, int* cs,
                                                                     int* vs, FlipRegs& fr) {
    int vh2[2];
    a.template vh_n<2>({vertex1, vertex2}, vh2);
    const int h1 = vh2[0], h2 = vh2[1];
    const int* c1 = a.vcn_ptr(h1);
    const int* c2 = a.vcn_ptr(h2);
    const int cells[3] = {c1[0], c1[1], c1[2]};
    const int ct[3] = {c2[0], c2[1], c2[2]};
    cs[0] = cs[1] = cs[2] = cs[3] = -1;
#pragma unroll
    for (int ff = 0; ff < 3; ++ff)
        if (ct[ff] != cells[0] && ct[ff] != cells[1] && ct[ff] != cells[2]) cs[3] = ct[ff];
    // the three cells of vertex1 and the fourth cell, with their rings, in one round
    int hc[4], n[4];
    int et[4][SVM_T1_RR];
    a.template ch_ring_n<4, SVM_T1_RR>({cells[0], cells[1], cells[2], cs[3] >= 0 ? cs[3] : cells[0]}, hc, n, et);
    if (n[0] < 2 || n[0] > SVM_T1_RR || n[1] < 2 || n[1] > SVM_T1_RR || n[2] < 2 || n[2] > SVM_T1_RR) return -1;
    int nl[3], nn[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int* r = a.ring_ptr(hc[t]);
        ring_nbrs(r, et[t], n[t], vertex1, nl[t], nn[t]);
        if (nl[t] == vertex2) cs[0] = cells[t];
        else if (nn[t] == vertex2) cs[2] = cells[t];
        else cs[1] = cells[t];
    }
    if (cs[0] < 0 || cs[1] < 0 || cs[2] < 0 || cs[3] < 0) return 0;
    // cs[1] is one of the three cells: its walk around vertex1 is done
    const int t1 = cells[2] == cs[1] ? 2 : (cells[1] == cs[1] ? 1 : 0);
    const int h3 = hc[3];
    const int n3 = n[3];
    if (n3 < 2 || n3 > SVM_T1_RR) return -1;
#pragma unroll
    for (int k = 0; k < SVM_T1_RR; ++k) fr.e[3][k] = et[3][k];
    vs[0] = t1 == 2 ? nl[2] : (t1 == 1 ? nl[1] : nl[0]);
    vs[1] = t1 == 2 ? nn[2] : (t1 == 1 ? nn[1] : nn[0]);
    ring_nbrs(a.ring_ptr(h3), fr.e[3], n3, vertex2, vs[3], vs[2]);
    // the hand-off, in cell-set order (a repeated cell has the same slot and ring)
    const int t0 = cells[0] == cs[0] ? 0 : (cells[1] == cs[0] ? 1 : 2);
    const int t2 = cells[0] == cs[2] ? 0 : (cells[1] == cs[2] ? 1 : 2);
    const int tt[3] = {t0, t1, t2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int t = tt[q];
        fr.hc[q] = t == 0 ? hc[0] : (t == 1 ? hc[1] : hc[2]);
        fr.n[q] = t == 0 ? n[0] : (t == 1 ? n[1] : n[2]);
#pragma unroll
        for (int k = 0; k < SVM_T1_RR; ++k) fr.e[q][k] = t == 0 ? et[0][k] : (t == 1 ? et[1][k] : et[2][k]);
    }
    fr.hc[3] = h3;
    fr.n[3] = n3;
    fr.h1 = h1;
    fr.h2 = h2;
    return 1;
}

// The flip with what cell_vertex_set_cached read (fr): only vertices b and c
// (vs[1], vs[2]) are looked up here.
template <class A>
__device__ __attribute__((always_inline)) int t1_flip_cached(A& a, int vertex1, int vertex2, const int* cs,
                                                             const int* vs, double* v1, double R, int cap,
                                                             const FlipRegs& fr) {
    const int hc1 = fr.hc[1], hc3 = fr.hc[3];
    const int n1 = fr.n[1], n3 = fr.n[3];
    if (n1 + 1 > cap || n3 + 1 > cap) return 4;
    int vbc[2];
    a.template vh_n<2>({vs[2], vs[1]}, vbc);
    const int h2 = fr.h2, hC = vbc[0], h1 = fr.h1, hb = vbc[1];
    const int hc0 = fr.hc[0], hc2 = fr.hc[2];
    const int n0 = fr.n[0], n2 = fr.n[2];
    if (h1 == h2 || h1 == hb || h1 == hC || h2 == hb || h2 == hC || hb == hC || hc0 == hc1 || hc0 == hc2 ||
        hc0 == hc3 || hc1 == hc2 || hc1 == hc3 || hc2 == hc3)
        return -1;
    if (n0 < 1 || n0 > SVM_T1_RR || n2 < 1 || n2 > SVM_T1_RR || n1 + 1 > SVM_T1_RR || n3 + 1 > SVM_T1_RR) return -1;
    const int(&e0)[SVM_T1_RR] = fr.e[0];
    const int(&e1)[SVM_T1_RR] = fr.e[1];
    const int(&e2)[SVM_T1_RR] = fr.e[2];
    const int(&e3)[SVM_T1_RR] = fr.e[3];
    const int p0 = ring_only(e0, n0, vertex2), p1 = ring_only(e1, n1, vertex1);
    const int p2 = ring_only(e2, n2, vertex1), p3 = ring_only(e3, n3, vertex2);
    if (p0 < 0 || p1 < 0 || p2 < 0 || p3 < 0) return -1;
    int vc1[3], vc2[3], vnb[3], vnc[3], vn1[3], vn2[3];
    {
        const int* q1 = a.vcn_ptr(h1); const int* q2 = a.vcn_ptr(h2);
        const int* qb = a.vn_ptr(hb); const int* qc = a.vn_ptr(hC);
        const int* r1 = a.vn_ptr(h1); const int* r2 = a.vn_ptr(h2);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vc1[k] = q1[k]; vc2[k] = q2[k]; vnb[k] = qb[k]; vnc[k] = qc[k]; vn1[k] = r1[k]; vn2[k] = r2[k];
        }
    }
    double v2[3], vC[3];
    {
        const double* q = a.vpos_ptr(h2);
        const double* r = a.vpos_ptr(hC);
        v2[0] = q[0]; v2[1] = q[1]; v2[2] = q[2];
        vC[0] = r[0]; vC[1] = r[1]; vC[2] = r[2];
    }
    flip_positions(v1, v2, vC, R);
    {
        double* q = a.vpos_ptr(h1);
        double* r = a.vpos_ptr(h2);
        q[0] = v1[0]; q[1] = v1[1]; q[2] = v1[2];
        r[0] = v2[0]; r[1] = v2[1]; r[2] = v2[2];
        a.mark_v(h1, 1);
        a.mark_v(h2, 1);
    }
#pragma unroll
    for (int vert = 0; vert < 3; ++vert) {
        if (vc1[vert] == cs[2]) { a.vcn_ptr(h1)[vert] = cs[3]; a.mark_v(h1, 2); }
        if (vc2[vert] == cs[0]) { a.vcn_ptr(h2)[vert] = cs[1]; a.mark_v(h2, 2); }
        if (vnb[vert] == vertex1) { a.vn_ptr(hb)[vert] = vertex2; a.mark_v(hb, 2); }
        if (vnc[vert] == vertex2) { a.vn_ptr(hC)[vert] = vertex1; a.mark_v(hC, 2); }
        if (vn1[vert] == vs[1]) { a.vn_ptr(h1)[vert] = vs[2]; a.mark_v(h1, 2); }
        if (vn2[vert] == vs[2]) { a.vn_ptr(h2)[vert] = vs[1]; a.mark_v(h2, 2); }
    }
    // cell i loses vertex2, cell j gains vertex2 after vertex1, cell k loses
    // vertex1, cell l gains vertex1 after vertex2
    int* r0 = a.ring_ptr(hc0);
    int* r1 = a.ring_ptr(hc1);
    int* r2 = a.ring_ptr(hc2);
    int* r3 = a.ring_ptr(hc3);
#pragma unroll
    for (int j = 0; j + 1 < SVM_T1_RR; ++j) {
        if (j >= p0 && j < n0 - 1) r0[j] = e0[j + 1];
        if (j >= p2 && j < n2 - 1) r2[j] = e2[j + 1];
    }
#pragma unroll
    for (int j = 1; j < SVM_T1_RR; ++j) {
        if (j == p1 + 1) r1[j] = vertex2;
        else if (j > p1 + 1 && j <= n1) r1[j] = e1[j - 1];
        if (j == p3 + 1) r3[j] = vertex1;
        else if (j > p3 + 1 && j <= n3) r3[j] = e3[j - 1];
    }
    a.set_cnn_h(hc0, n0 - 1);
    a.set_cnn_h(hc1, n1 + 1);
    a.set_cnn_h(hc2, n2 - 1);
    a.set_cnn_h(hc3, n3 + 1);
    return 0;
}

struct VisitStats {
    long long flips, skipped;
};

// The sequential loop's visit of vertex u (row gu = vo + u): slots 0..2, each
// edge (u, w) with u < w shorter than thr flipped unless protected.  `push`
// receives every later vertex (local index > u) a flip disturbed; `on_flip`
// receives each flip's (vertex1, vertex2, cell set).  Returns 0 or an error.
template <class A, class Push, class OnFlip>
__device__ __attribute__((always_inline)) int visit_vertex(A& a, int gu, int u, int vo, int nv, int co, int nc, double R, double thr, double cos_cut, int cap,
                            int* tmp, VisitStats& st, Push push, OnFlip on_flip) {
    double v1[3];
    a.pos(gu, v1);
    double n1 = pnorm(v1);       // changes only with v1, in a flip
    // u's neighbour list, read again after each flip (a flip rewrites it)
    int nbl[3] = {a.vn(gu, 0), a.vn(gu, 1), a.vn(gu, 2)};
    for (int vv = 0; vv < 3; ++vv) {
        int vertex2 = nbl[vv];
        if (gu < vertex2) {
            double q[3];
            a.pos(vertex2, q);
            if (short_edge(v1, q, n1, pnorm(q), R, thr, cos_cut)) {
                int cs[4], vs[4] = {0, 0, 0, 0};
                int got = -1;
                FlipRegs fr;
                bool have = false;
                if constexpr (A::kCached) {
                    got = cell_vertex_set_cached(a, gu, vertex2, cs, vs, fr);
                    have = got > 0;
                }
                if (got < 0) got = cell_vertex_set(a, gu, vertex2, cs, vs) ? 1 : 0;
                if (!got) {
                    ++st.skipped;
                    continue;
                }
                bool inside = true;
                for (int k = 0; k < 4; ++k) {
                    inside &= cs[k] >= co && cs[k] < co + nc;
                    inside &= vs[k] >= vo && vs[k] < vo + nv;
                }
                if (!inside) return 5;
                if (have ? (fr.n[0] == 3 || fr.n[2] == 3) : (a.cnn(cs[0]) == 3 || a.cnn(cs[2]) == 3)) continue;
                int r = -1;
                if constexpr (A::kCached) {
                    if (have) r = t1_flip_cached(a, gu, vertex2, cs, vs, v1, R, cap, fr);
                }
                if (r < 0) r = t1_flip(a, gu, vertex2, cs, vs, v1, R, cap, tmp);
                if (r) return r;
                n1 = pnorm(v1);
                nbl[0] = a.vn(gu, 0); nbl[1] = a.vn(gu, 1); nbl[2] = a.vn(gu, 2);
                ++st.flips;
                on_flip(gu, vertex2, cs);
                int dirty[6] = {u, vertex2 - vo, vs[0] - vo, vs[1] - vo, vs[2] - vo, vs[3] - vo};
                for (int k = 0; k < 6; ++k)
                    if (dirty[k] > u && dirty[k] < nv) push(dirty[k]);
            }
        }
        if (a.failed()) return 0;
    }
    return 0;
}

// Rebuild vln for vertex x from the current rings; returns false if a ring
// neighbour is not in x's vertex-neighbour list.  The loads are issued in
// three dependent rounds (vcn/vn, the three counts, eight ring entries of each
// cell), not one ring entry at a time.
__device__ inline bool build_vln(const Dev& d, int x) {
    const int4 cl = d.vcn[x];
    const int4 nb = d.vn[x];
    const int cells[3] = {cl.x, cl.y, cl.z};
    int n3[3];
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) n3[sl] = d.cnn[cells[sl]];
    int e[3][8];
#pragma unroll
    for (int sl = 0; sl < 3; ++sl)
#pragma unroll
        for (int j = 0; j < 8; ++j) e[sl][j] = j < n3[sl] ? ringref(d, cells[sl], j) : -1;
    int out = 0;
    bool ok = true;
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        const int c = cells[sl], n = n3[sl];
        int k = -1;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k < 0 && e[sl][j] == x) k = j;
        for (int j = 8; k < 0 && j < n; ++j)
            if (ringref(d, c, j) == x) k = j;
        if (k < 0) { ok = false; continue; }
        const int kp = k == 0 ? n - 1 : k - 1, kn = k + 1 == n ? 0 : k + 1;
        int prev = -1, next = -1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            prev = j == kp ? e[sl][j] : prev;
            next = j == kn ? e[sl][j] : next;
        }
        if (kp >= 8) prev = ringref(d, c, kp);
        if (kn >= 8) next = ringref(d, c, kn);
        const int ls = prev == nb.x ? 0 : (prev == nb.y ? 1 : (prev == nb.z ? 2 : -1));
        const int ns = next == nb.x ? 0 : (next == nb.y ? 1 : (next == nb.z ? 2 : -1));
        if (ls < 0 || ns < 0) { ok = false; continue; }
        out |= (ls << (4 * sl)) | (ns << (4 * sl + 2));
    }
    d.vln[x] = out;
    return ok;
}

__global__ __launch_bounds__(SVM_TILE) void k_build_vln(Dev d, int only) {
    const int tile = blockIdx.x;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if (only >= 0 && s != only) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    if (v - S->vo >= S->nv) return;
    if (!build_vln(d, v)) atomicMax(&S->err, 5);
}

// ------------------------------------------------------------- detect -----
// Short-edge test of testAndPerformT1TransitionsCPU
// (src/model/sphericalVertexModel.cpp:469-481): vertex v is a candidate when
// it has an edge (v, w), v < w, shorter than the threshold.  Candidates go to
// the bit mask (for the sequential path) and to the compact list (for the dry
// runs).  The active updater's director step is in the vertex pass; the T1
// commit redoes it for every vertex a flip moved.
// skip: 1 = spheres a minimiser has stopped (f_frozen), 2 = spheres whose
// minimiser has converged; 0 = none.
__global__ __launch_bounds__(SVM_TILE) void k_t1_detect(Dev d, int pcur, int skip) {
    const int tile = lane_vtile(d);
    if (tile < 0) return;
    const int s = d.vtile_sys[tile];
    SysParams* S = d.sys + s;
    if ((skip == 1 && S->f_frozen) || (skip == 2 && minimiser_converged(S))) return;
    const int v = tile * SVM_TILE + threadIdx.x;
    const int local = v - S->vo;
    const double* X = d.px[pcur];
    const double* Y = d.py[pcur];
    const double* Z = d.pz[pcur];
    bool flag = false, flippable = false;
    int skips = 0;
    if (local < S->nv) {
        double p[3] = {X[v], Y[v], Z[v]};
        int4 nb = d.vn[v];
        int w3[3] = {nb.x, nb.y, nb.z};
        // every neighbour's position in one round of loads (rows of this
        // sphere, in bounds whichever edges are tested), before any test
        double q3[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            q3[k][0] = X[w3[k]];
            q3[k][1] = Y[w3[k]];
            q3[k][2] = Z[w3[k]];
        }
        double n1 = pnorm(p);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int w = w3[k];
            if (v < w) {
                const double* q = q3[k];
                if (short_edge(p, q, n1, pnorm(q), S->R, S->thr, S->thr_cos)) {
                    flag = true;
                    if (!flippable) {
                        // would the sequential visit flip this edge, given the current state?
                        GlobalAcc a{&d, pcur};
                        int cs[4], vs[4];
                        if (!cell_vertex_set(a, v, w, cs, vs)) ++skips;
                        else if (!(d.cnn[cs[0]] == 3 || d.cnn[cs[2]] == 3)) flippable = true;
                    }
                }
            }
        }
    }
    unsigned long long m = __ballot(flag);
    unsigned long long mf = __ballot(flag && flippable);
    unsigned long long mb = m & ~mf;
    const int lane = threadIdx.x & 63;
    int basef = 0, baseb = 0;
    if (lane == 0) {
        d.t1mask[v >> 6] = m;
        d.t1fmask[v >> 6] = mf;
        if (m) atomicOr(&S->t1_flag, 1);
        if (mf) {
            int cnt = __popcll(mf);
            basef = atomicAdd(d.t1_count, cnt);
            atomicAdd((unsigned long long*)&S->n_cand_flip, (unsigned long long)cnt);
            if (basef + cnt > SVM_T1_MAXCL) atomicOr(&S->t1_overflow, 1);
        }
        if (mb) {
            int cnt = __popcll(mb);
            baseb = atomicAdd(d.t1_bcount, cnt);
            atomicAdd((unsigned long long*)&S->n_cand_block, (unsigned long long)cnt);
            if (baseb + cnt > SVM_T1_MAXBL) atomicOr(&S->t1_overflow, 2);
        }
    }
    basef = __shfl(basef, 0, 64);
    baseb = __shfl(baseb, 0, 64);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (flag && flippable) {
        int pos = basef + __popcll(mf & below);
        if (pos < SVM_T1_MAXCL) {
            // list slots are reused across steps and spheres: clear their marks
            d.t1_list[pos] = v;
            d.t1_flagged[pos] = 0;
            d.t1_logged[pos] = 0;
            d.t1_lead[pos] = pos;
            d.t1_vidx[v] = pos;
        }
    } else if (flag) {
        int pos = baseb + __popcll(mb & below);
        if (pos < SVM_T1_MAXBL) {
            d.t1_blist[pos] = v;
            d.t1_bskip[pos] = skips;
            d.t1_bflag[pos] = 0;
        }
    }
}

// A blocked candidate (no short edge of it can flip in the current state)
// has a visit without writes, whose outcome stands unless a flippable
// candidate's run touched what the visit reads: the vertex, its later
// neighbours, its cells and the fourth cell of each short edge.  Run by the
// group kernel's block of the candidate's sphere, a thread per blocked
// candidate, before that block decides whether the sphere needs a group run:
// an entangled candidate is flagged, with the runs it met, for the group run;
// otherwise its no-fourth-cell skips are tallied for the commit.  The runs a
// candidate can meet are its own sphere's, so one block sees all it needs.
__device__ __attribute__((always_inline)) inline void t1_check_sphere(const Dev& d, SysParams* S, int vo, int nv,
                                                                     int t, int nt, int* hit_any, int blocked) {
    const int total = min(blocked, SVM_T1_MAXBL);
    const unsigned gen = (unsigned)S->t1_gen;
    for (int i = t; i < total; i += nt) {
        const int v = d.t1_blist[i];
        if (v < vo || v >= vo + nv) continue;
        bool hit = false;
        auto look = [&](const unsigned long long* own) {
            unsigned long long st = *own;
            if ((unsigned)(st >> 32) == gen) {
                hit = true;
                d.t1_flagged[(int)(st & 0xffffffffu) - 1] = (int)gen;
            }
        };
        look(d.vown + v);
        int4 cl = d.vcn[v];
        look(d.cown + cl.x);
        look(d.cown + cl.y);
        look(d.cown + cl.z);
        int4 nb = d.vn[v];
        int w3[3] = {nb.x, nb.y, nb.z};
        for (int k = 0; k < 3; ++k) {
            int w = w3[k];
            if (v < w) {
                look(d.vown + w);
                int4 cw = d.vcn[w];
                int c3[3] = {cw.x, cw.y, cw.z};
                for (int j = 0; j < 3; ++j)
                    if (c3[j] != cl.x && c3[j] != cl.y && c3[j] != cl.z) look(d.cown + c3[j]);
            }
        }
        if (hit) {
            d.t1_bflag[i] = (int)gen;
            *hit_any = 1;
        } else if (d.t1_bskip[i]) {
            atomicAdd(&S->t1_bskips, d.t1_bskip[i]);   // counted if the dry runs are committed
        }
    }
}

// ------------------------------------------------------------ dry runs ----
// Write one log's items to HBM (whole block).
// One item of a dry-run log written to the topology, with every load issued
// before any store (a store to the topology could alias the next load as far
// as the compiler knows, which would otherwise serialise the items on memory
// latency).  Item i < nvert is vertex i; nvert + j is cell j.  Vertices of a
// written cell go to `verts` (once each, through the LDS bitmap `seen` over
// the sphere's local vertex ids when it is given) and moved vertices to
// `moved`, for the rebuild that follows.
template <int RC>
__device__ __attribute__((always_inline)) inline void apply_item(
    const Dev& d, int pcur, int item, int nvert, const int* vid, const int* vflag, const double (*vpos)[3],
    const int (*vvn)[3], const int (*vvcn)[3], const int* cid, const int* ccnt, const int (*cring)[RC], int dir_mode,
    int vo, unsigned* seen, int seen_words, int* verts, int* nverts, int* moved, int* nmoved, int cap) {
    if (item < nvert) {
        const int v = vid[item], fl = vflag[item];
        const double p0 = vpos[item][0], p1 = vpos[item][1], p2 = vpos[item][2];
        const int n0 = vvn[item][0], n1 = vvn[item][1], n2 = vvn[item][2];
        const int c0 = vvcn[item][0], c1 = vvcn[item][1], c2 = vvcn[item][2];
        if (fl & 1) {
            d.px[pcur][v] = p0;
            d.py[pcur][v] = p1;
            d.pz[pcur][v] = p2;
            if (dir_mode) {
                const int slot = atomicAdd(nmoved, 1);
                if (slot < cap) moved[slot] = v;
            }
        }
        if (fl & 2) {
            vnref(d, v, 0) = n0; vnref(d, v, 1) = n1; vnref(d, v, 2) = n2;
            vcnref(d, v, 0) = c0; vcnref(d, v, 1) = c1; vcnref(d, v, 2) = c2;
        }
        return;
    }
    const int i = item - nvert;
    const int c = cid[i], cnt = ccnt[i];
    int e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = cring[i][k];
    auto add = [&](int x) {
        const unsigned lx = (unsigned)(x - vo);
        if (seen && (lx >> 5) < (unsigned)seen_words) {
            const unsigned bit = 1u << (lx & 31);
            if (atomicOr(&seen[lx >> 5], bit) & bit) return;
        }
        const int slot = atomicAdd(nverts, 1);
        if (slot < cap) verts[slot] = x;
    };
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < cnt) { ringref(d, c, k) = e[k]; add(e[k]); }
    for (int k = 8; k < cnt; ++k) {
        const int x = cring[i][k];
        ringref(d, c, k) = x;
        add(x);
    }
    d.cnn[c] = cnt;
}

// Commit the dry runs of one sphere (whole block): the logs of every
// candidate run that completed and was not flagged as overlapping, plus the
// group run's log when there was one.  These logs write disjoint vertices and
// cells, so their items are applied in parallel, one item per thread; then
// vln of every vertex of a written cell and, for the active updater, the
// director of every moved vertex are rebuilt one vertex per thread.  Every
// write is a pure function of the logs, so the result does not depend on the
// order.
#define SVM_COMMIT_VERTS 4096
#define SVM_COMMIT_LOGS 1024
#define SVM_COMMIT_SEEN 4096      // bitmap words: spheres of up to 131072 vertices are deduplicated
template <int NT>
__device__ inline void commit_sphere(const Dev& d, SysParams* S, int pcur, int nin, int nout, int dir_mode,
                                     int use_group) {
    __shared__ int s_verts[SVM_COMMIT_VERTS];
    __shared__ int s_moved[SVM_COMMIT_VERTS];
    __shared__ int s_logs[SVM_COMMIT_LOGS];
    __shared__ int s_lstart[SVM_COMMIT_LOGS + 1];    // prefix sums of items per log
    __shared__ unsigned s_seen[SVM_COMMIT_SEEN];
    __shared__ int s_part[NT];
    __shared__ int s_nverts, s_nmoved, s_nlogs, s_gcount;
    __shared__ unsigned long long s_flips, s_skipped;
    const int t = threadIdx.x, nt = blockDim.x;
    const int gen = S->t1_gen;
    const int vo = S->vo, nv = S->nv;
    const T1GroupLog& G = d.t1_group[S - d.sys];
    const int seen_words = (nv + 31) >> 5;
    const bool dedup = seen_words <= SVM_COMMIT_SEEN;
    if (t == 0) {
        s_nverts = 0; s_nmoved = 0; s_nlogs = 0;
        s_flips = use_group ? G.flips : 0;
        s_skipped = (use_group ? G.skipped : 0) + S->t1_bskips;
    }
    if (dedup)
        for (int w = t; w < seen_words; w += nt) s_seen[w] = 0u;
    __syncthreads();
    // the flippable list's length as the dry runs saw it (the live counter is
    // reset by this kernel)
    const int total = min(*d.t1_total, SVM_T1_MAXCL);
    for (int j = t; j < total; j += nt) {
        const int row = d.t1_list[j];
        if (!(row >= vo && row < vo + nv && d.t1_logged[j] == gen && d.t1_flagged[j] != gen)) continue;
        const int slot = atomicAdd(&s_nlogs, 1);
        if (slot < SVM_COMMIT_LOGS) s_logs[slot] = j;
        const T1Log& L = d.t1_log[j];
        atomicAdd(&s_flips, (unsigned long long)L.flips);
        atomicAdd(&s_skipped, (unsigned long long)L.skipped);
    }
    __syncthreads();
    const int nlogs = s_nlogs;
    unsigned* seen = dedup ? s_seen : nullptr;
    if (nlogs <= SVM_COMMIT_LOGS) {
        // items per log, then an exclusive scan: each thread owns a run of
        // SVM_COMMIT_LOGS / NT logs, and the run totals are scanned in LDS
        constexpr int PER = SVM_COMMIT_LOGS / NT;
        static_assert(PER * NT == SVM_COMMIT_LOGS, "the block size divides the log list");
        int cnt[PER], run = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int li = t * PER + q;
            cnt[q] = 0;
            if (li < nlogs) {
                const T1Log& L = d.t1_log[s_logs[li]];
                cnt[q] = L.nvert + L.ncell;
            }
            run += cnt[q];
        }
        s_part[t] = run;
        __syncthreads();
        for (int off = 1; off < NT; off <<= 1) {
            const int add = t >= off ? s_part[t - off] : 0;
            __syncthreads();
            s_part[t] += add;
            __syncthreads();
        }
        int acc = s_part[t] - run;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int li = t * PER + q;
            if (li < nlogs) s_lstart[li] = acc;
            acc += cnt[q];
        }
        if (t == nt - 1) s_lstart[nlogs] = s_part[t];
        __syncthreads();
        const int nitems = s_lstart[nlogs];
        int li = 0;
        for (int x = t; x < nitems; x += nt) {
            while (s_lstart[li + 1] <= x) ++li;
            const T1Log& L = d.t1_log[s_logs[li]];
            apply_item<SVM_T1_RC>(d, pcur, x - s_lstart[li], L.nvert, L.vid, L.vflag, L.vpos, L.vvn, L.vvcn, L.cid,
                                  L.ccnt, L.cring, dir_mode, vo, seen, seen_words, s_verts, &s_nverts, s_moved,
                                  &s_nmoved, SVM_COMMIT_VERTS);
        }
    } else {
        // more committed logs than the list holds (not seen in practice):
        // one log per thread, and the whole sphere is rebuilt below
        for (int j = t; j < total; j += nt) {
            const int row = d.t1_list[j];
            if (!(row >= vo && row < vo + nv && d.t1_logged[j] == gen && d.t1_flagged[j] != gen)) continue;
            const T1Log& L = d.t1_log[j];
            for (int x = 0; x < L.nvert + L.ncell; ++x)
                apply_item<SVM_T1_RC>(d, pcur, x, L.nvert, L.vid, L.vflag, L.vpos, L.vvn, L.vvcn, L.cid, L.ccnt,
                                      L.cring, dir_mode, vo, nullptr, 0, s_verts, &s_nverts, s_moved, &s_nmoved, 0);
        }
    }
    if (use_group) {
        const int ng = G.nvert, nitems = ng + G.ncell;
        for (int x = t; x < nitems; x += nt)
            apply_item<SVM_T1_RC>(d, pcur, x, ng, G.vid, G.vflag, G.vpos, G.vvn, G.vvcn, G.cid, G.ccnt, G.cring,
                                  dir_mode, vo, seen, seen_words, s_verts, &s_nverts, s_moved, &s_nmoved,
                                  SVM_COMMIT_VERTS);
    }
    __threadfence_block();
    __syncthreads();
    const bool whole = nlogs > SVM_COMMIT_LOGS;
    // every item of the committed logs and of the group log, one per thread
    // (for the paths where the LDS lists overflowed)
    auto for_items = [&](auto&& fn) {
        const int nitems = s_lstart[nlogs];
        int li = 0;
        for (int x = t; x < nitems; x += nt) {
            while (s_lstart[li + 1] <= x) ++li;
            const T1Log& L = d.t1_log[s_logs[li]];
            fn(x - s_lstart[li], L.nvert, L.vid, L.vflag, L.ccnt, &L.cring[0][0]);
        }
        if (use_group) {
            const int ng = G.nvert;
            for (int x = t; x < ng + G.ncell; x += nt) fn(x, ng, G.vid, G.vflag, G.ccnt, &G.cring[0][0]);
        }
    };
    if (whole) {
        for (int x = vo + t; x < vo + nv; x += nt)
            if (!build_vln(d, x)) atomicMax(&S->err, 5);
    } else if (s_nverts <= SVM_COMMIT_VERTS) {
        for (int i = t; i < s_nverts; i += nt)
            if (!build_vln(d, s_verts[i])) atomicMax(&S->err, 5);
    } else {
        // more than the LDS list holds: the distinct vertices of the written
        // cells go to this sphere's region of a global list, deduplicated by
        // a global bitmap that is cleared again afterwards
        if (t == 0) s_gcount = 0;
        __syncthreads();
        for_items([&](int item, int nvert, const int*, const int*, const int* ccnt, const int* cring) {
            if (item < nvert) return;
            const int i = item - nvert, cnt = ccnt[i];
            for (int k = 0; k < cnt; ++k) {
                const int x = cring[i * SVM_T1_RC + k];
                if ((unsigned)(x - vo) >= (unsigned)nv) continue;
                const unsigned bit = 1u << (x & 31);
                if (atomicOr(&d.cvseen[x >> 5], bit) & bit) continue;
                d.cvlist[vo + atomicAdd(&s_gcount, 1)] = x;   // distinct vertices of the sphere: fewer than nv
            }
        });
        __threadfence_block();
        __syncthreads();
        const int gcount = s_gcount;
        for (int i = t; i < gcount; i += nt)
            if (!build_vln(d, d.cvlist[vo + i])) atomicMax(&S->err, 5);
        for (int i = t; i < gcount; i += nt) {
            const int x = d.cvlist[vo + i];
            atomicAnd(&d.cvseen[x >> 5], ~(1u << (x & 31)));
        }
    }
    if (dir_mode) {
        auto redo = [&](int x) {
            double p[3] = {d.px[pcur][x], d.py[pcur][x], d.pz[pcur][x]};
            double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
            director_update(S, x - vo, p, nI, nO);
            d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
        };
        if (whole) {
            for (int i = t; i < nv; i += nt) redo(vo + i);
        } else if (s_nmoved <= SVM_COMMIT_VERTS) {
            for (int i = t; i < s_nmoved; i += nt) redo(s_moved[i]);
        } else {
            // moved vertices are distinct across the logs: straight from the items
            for_items([&](int item, int nvert, const int* vid, const int* vflag, const int*, const int*) {
                if (item < nvert && (vflag[item] & 1)) redo(vid[item]);
            });
        }
    }
    if (t == 0) {
        S->n_t1 += (long long)s_flips;
        S->n_t1_skipped += (long long)s_skipped;
    }
    __syncthreads();
}

// Predicted components.  Two flippable candidates of a sphere closer than
// d.t1_near cell sizes (sqrt of the mean cell area; SVM_T1_NEAR unless set
// otherwise) are taken to interact,
// and the transitive closure of that relation splits the sphere's candidates
// into components.  One dry run covers a whole component: its members and
// the vertices their flips disturbed, in index order -- the sequential loop
// restricted to the component, as a group run is.  A prediction is not
// trusted for exactness: components whose runs still touch a common item are
// flagged by the stamps and redone by the group run, as single runs were.
// Every wavefront of a component computes the same closure from the same
// inputs (the flippable mask and the moved positions, which no dry run
// writes; the distance test is symmetric bit for bit), and the member with
// the lowest list slot is its leader: it runs, stamps and logs under its own
// slot, and every member records that slot in t1_lead.  Spheres with more
// than 64 flippable candidates, or more than 64 * 64 * SVM_T1_NEAR_MAXW
// vertex rows, keep one run per candidate, as does a distance of 0.
#ifndef SVM_T1_NEAR
#define SVM_T1_NEAR 3.0
#endif
#define SVM_T1_NEAR_MAXW 16

__device__ __attribute__((always_inline)) inline unsigned long long wave_or64(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
    return x;
}
__device__ __attribute__((always_inline)) inline int wave_min32(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
    return x;
}

// The component of candidate `row` (slot idx) in its sphere: members' local
// indices in ascending order into seeds[0..n), and the leader's slot.  All
// 64 lanes of the wavefront call it with the same arguments.
template <class Sync>
__device__ __attribute__((always_inline)) inline int t1_component(const Dev& d, const SysParams* S, int pcur, int row,
                                                                  int idx, int lane, int* seeds, int* crow,
                                                                  double* cpos, int* lead, Sync sync) {
    const int vo = S->vo, nv = S->nv;
    if (lane == 0) seeds[0] = row - vo;
    *lead = idx;
    if (!(d.t1_near > 0.0) || nv > 64 * 64 * SVM_T1_NEAR_MAXW) { sync(); return 1; }
    // the sphere's flippable candidates, in row order: lane L reads mask words
    // [L * MAXW, (L + 1) * MAXW) of the sphere (vo is a multiple of 64)
    const int w0 = vo >> 6, nw = (nv + 63) >> 6;
    unsigned long long words[SVM_T1_NEAR_MAXW];
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < SVM_T1_NEAR_MAXW; ++k) {
        const int wi = lane * SVM_T1_NEAR_MAXW + k;
        words[k] = wi < nw ? d.t1fmask[w0 + wi] : 0ull;
        cnt += __popcll(words[k]);
    }
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    const int ncand = __shfl(inc, 63, 64);
    if (ncand < 2 || ncand > 64) { sync(); return 1; }
    int at = inc - cnt;
#pragma unroll
    for (int k = 0; k < SVM_T1_NEAR_MAXW; ++k) {
        unsigned long long m = words[k];
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1ull;
            crow[at++] = ((w0 + lane * SVM_T1_NEAR_MAXW + k) << 6) + b;
        }
    }
    sync();
    const bool act = lane < ncand;
    const int myrow = act ? crow[lane] : -1;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    int myidx = 0x7fffffff;
    if (act) {
        p0 = d.px[pcur][myrow];
        p1 = d.py[pcur][myrow];
        p2 = d.pz[pcur][myrow];
        myidx = d.t1_vidx[myrow];
        cpos[3 * lane] = p0;
        cpos[3 * lane + 1] = p1;
        cpos[3 * lane + 2] = p2;
    }
    sync();
    const double cell2 = 4.0 * SVM_PI * S->R * S->R / (double)S->nc;
    const double near2 = (d.t1_near * d.t1_near) * cell2;
    unsigned long long adj = 0ull;
    if (act)
        for (int k = 0; k < ncand; ++k) {
            const double e0 = cpos[3 * k] - p0, e1 = cpos[3 * k + 1] - p1, e2 = cpos[3 * k + 2] - p2;
            if (e0 * e0 + e1 * e1 + e2 * e2 < near2) adj |= 1ull << k;
        }
    unsigned long long comp = __ballot(myrow == row);
    if (!comp) { sync(); return 1; }    // not in the mask: a run of its own
    while (true) {
        const unsigned long long nx = wave_or64(((comp >> lane) & 1ull) ? adj : 0ull) | comp;
        if (nx == comp) break;
        comp = nx;
    }
    const bool in = act && ((comp >> lane) & 1ull);
    *lead = wave_min32(in ? myidx : 0x7fffffff);
    if (in) seeds[__popcll(comp & ((1ull << lane) - 1ull))] = myrow - vo;
    sync();
    return __popcll(comp);
}

__global__ __launch_bounds__(64 * SVM_T1_WAVES) void k_t1_cluster(Dev d, int pcur) {
    __shared__ int s_vkey[SVM_T1_WAVES][SVM_T1_WV], s_vflag[SVM_T1_WAVES][SVM_T1_WV], s_vuse[SVM_T1_WAVES][SVM_T1_WV];
    __shared__ double s_vp[SVM_T1_WAVES][SVM_T1_WV * 3];
    __shared__ int s_vvn[SVM_T1_WAVES][SVM_T1_WV * 3], s_vvcn[SVM_T1_WAVES][SVM_T1_WV * 3];
    __shared__ int s_ckey[SVM_T1_WAVES][SVM_T1_WC], s_cflag[SVM_T1_WAVES][SVM_T1_WC];
    __shared__ int s_cuse[SVM_T1_WAVES][SVM_T1_WC], s_ccnt[SVM_T1_WAVES][SVM_T1_WC];
    __shared__ int s_cring[SVM_T1_WAVES][SVM_T1_WC * SVM_T1_RC];
    __shared__ int s_heap[SVM_T1_WAVES][64];
    __shared__ int s_tmp[SVM_T1_WAVES][SVM_T1_RC + 8];
    __shared__ int s_misc[SVM_T1_WAVES][8];   // vfill, cfill, -, outcome, nvert, ncell
    __shared__ int s_seed[SVM_T1_WAVES][64];
    __shared__ int s_crow[SVM_T1_WAVES][64];
    __shared__ double s_cpos[SVM_T1_WAVES][64 * 3];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const WaveSync sync;
    const int nwaves = gridDim.x * SVM_T1_WAVES;
    const int total = min(*d.t1_count, SVM_T1_MAXCL);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *d.t1_total = *d.t1_count;
        *d.t1_btotal = *d.t1_bcount;
    }
    int* misc = s_misc[w];
    for (int idx = blockIdx.x * SVM_T1_WAVES + w; idx < total; idx += nwaves) {
        const int row = d.t1_list[idx];
        SysParams* S = d.sys + d.vtile_sys[row / SVM_TILE];
        if (S->t1_overflow) continue;   // the sphere takes the sequential pass
        const unsigned gen = (unsigned)S->t1_gen;
        const int vo = S->vo, nv = S->nv;
        int lead = idx;
        const int nmem = t1_component(d, S, pcur, row, idx, lane, s_seed[w], s_crow[w], s_cpos[w], &lead, sync);
        if (lane == 0) d.t1_lead[idx] = lead;
        if (lead != idx) continue;      // the leader's run covers this candidate
        CacheAcc<SVM_T1_WV, SVM_T1_WC> a{&d, pcur, false, s_vkey[w], s_vflag[w], s_vuse[w], s_vp[w], s_vvn[w],
                                         s_vvcn[w], s_ckey[w], s_cflag[w], s_cuse[w], s_ccnt[w], s_cring[w],
                                         misc + 0, misc + 1};
        a.reset(lane, 64);
        sync();
        const int* seeds = s_seed[w];
        a.prefetch(seeds, nmem, vo, lane, 64, SVM_T1_WV * 5 / 8, SVM_T1_WC * 5 / 8, sync);
        T1Log& L = d.t1_log[idx];
        if (lane == 0) {
            Heap heap{s_heap[w], 0};
            bool heap_full = false;
            VisitStats st{0, 0};
            int err = 0, last = -1, nvis = 0, mi = 0;
            // the members and the later vertices their flips disturbed, in index order
            while (true) {
                const int cu = mi < nmem ? seeds[mi] : 0x7fffffff;
                const int hu = heap.n > 0 ? heap.a[0] : 0x7fffffff;
                if (cu == 0x7fffffff && hu == 0x7fffffff) break;
                int u;
                if (hu < cu) { u = hu; heap.pop(); }
                else { u = cu; ++mi; }
                if (u <= last) continue;
                last = u;
                ++nvis;
                err = visit_vertex(a, vo + u, u, vo, nv, S->co, S->nc, S->R, S->thr, S->thr_cos, d.cap, s_tmp[w], st,
                                   [&](int x) {
                                       if (heap.n < 64) heap.push(x);
                                       else heap_full = true;
                                   },
                                   [&](int, int, const int*) {});
                if (err || a.failed() || heap_full) break;
            }
            const bool ok = !(err || a.failed() || heap_full);
            misc[3] = !ok ? 0 : ((st.flips || st.skipped) ? 2 : 1);
            atomicAdd((unsigned long long*)&S->n_cl_visits, (unsigned long long)nvis);
            atomicMax((unsigned long long*)&S->max_cl_visits, (unsigned long long)nvis);
            L.flips = st.flips;
            L.skipped = st.skipped;
        }
        sync();
        const int outcome = misc[3];
        a.stamp_touched(((unsigned long long)gen << 32) | (unsigned)(idx + 1), d.t1_flagged, &S->t1_overlap, lane,
                        64);
        bool ok = outcome != 0;
        if (outcome == 2)
            ok = a.template write_log<SVM_T1_LV, SVM_T1_LC>(L.vid, L.vflag, L.vpos, L.vvn, L.vvcn, L.cid, L.ccnt,
                                                            L.cring, misc + 4, misc + 5, lane, 64, sync);
        if (lane == 0) {
            if (!ok) {
                // this component's run could not be completed here: the group run redoes it
                d.t1_flagged[idx] = (int)gen;
                S->t1_overlap = 1;
            } else if (outcome == 2) {
                L.nvert = misc[4];
                L.ncell = misc[5];
                d.t1_logged[idx] = (int)gen;
            }
        }
        sync();
    }
}

struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};

#define SVM_T1_GT 256         // threads of a group block

// Second round, block per sphere whose single-candidate runs
// overlapped: one dry run of the sequential loop over all the flagged
// candidates together (in index order, with their disturbed vertices), in a
// larger cache.  Touching an item of a run that was not flagged means that
// run interacts: it is absorbed into the group and the group retries (the
// cache keeps what it loaded).  Out of room, or too many retries, and the
// sphere takes the sequential pass.
__global__ __launch_bounds__(SVM_T1_GT) void k_t1_group(Dev d, int pcur) {
    __shared__ int s_vkey[SVM_T1_GV], s_vflag[SVM_T1_GV], s_vuse[SVM_T1_GV];
    __shared__ double s_vp[SVM_T1_GV * 3];
    __shared__ int s_vvn[SVM_T1_GV * 3], s_vvcn[SVM_T1_GV * 3];
    __shared__ int s_ckey[SVM_T1_GC], s_cflag[SVM_T1_GC], s_cuse[SVM_T1_GC], s_ccnt[SVM_T1_GC];
    __shared__ int s_cring[SVM_T1_GC * SVM_T1_RC];
    __shared__ int s_heap[SVM_T1_GV];
    __shared__ int s_tmp[SVM_T1_RC + 8];
    __shared__ int s_mem[SVM_T1_GMAX];
    __shared__ int s_absorb[SVM_T1_GABS];
    __shared__ int s_nmem, s_nabs, s_vfill, s_cfill, s_status, s_nlv, s_nlc;
    __shared__ long long s_flips, s_skipped;
    __shared__ int s_hit;
    const int s = d.s0 + (int)blockIdx.x;
    SysParams* S = d.sys + s;
    const int lane = threadIdx.x;     // the walk runs on thread 0; the rest is shared by the block
    const BlockSync sync;
    T1GroupLog& G = d.t1_group[s];
    const int vo = S->vo, nv = S->nv;
    if (lane == 0) s_hit = 0;
    __syncthreads();
    t1_check_sphere(d, S, vo, nv, lane, SVM_T1_GT, &s_hit, *d.t1_bcount);
    __syncthreads();
    if (s_hit && lane == 0) S->t1_overlap = 1;
    __syncthreads();
    if (!S->t1_overlap || S->t1_overflow) {
        if (lane == 0) {
            G.ok = 0;
            if (S->t1_overflow & 1) S->n_seq_reason[0] += 1;
            if (S->t1_overflow & 2) S->n_seq_reason[3] += 1;
        }
        return;
    }
    const int gen = S->t1_gen;
    const int total = min(*d.t1_count, SVM_T1_MAXCL);
    if (lane == 0) s_nmem = 0;
    __syncthreads();
    for (int i = lane; i < total; i += SVM_T1_GT) {
        int row = d.t1_list[i];
        if (row >= vo && row < vo + nv && d.t1_flagged[d.t1_lead[i]] == gen) {
            int k = atomicAdd(&s_nmem, 1);
            if (k < SVM_T1_GMAX) s_mem[k] = row - vo;
        }
    }
    const int btotal = min(*d.t1_bcount, SVM_T1_MAXBL);
    for (int i = lane; i < btotal; i += SVM_T1_GT) {
        int row = d.t1_blist[i];
        if (row >= vo && row < vo + nv && d.t1_bflag[i] == gen) {
            int k = atomicAdd(&s_nmem, 1);
            if (k < SVM_T1_GMAX) s_mem[k] = row - vo;
        }
    }
    __syncthreads();
    int nmem = s_nmem;
    if (nmem > SVM_T1_GMAX) {
        if (lane == 0) { G.ok = 0; S->n_seq_reason[1] += 1; }
        return;
    }
    CacheAcc<SVM_T1_GV, SVM_T1_GC> a{&d, pcur, false, s_vkey, s_vflag, s_vuse, s_vp, s_vvn, s_vvcn,
                                     s_ckey, s_cflag, s_cuse, s_ccnt, s_cring, &s_vfill, &s_cfill};
    a.reset(lane, SVM_T1_GT);
    if (lane == 0) G.ok = 0;
    __syncthreads();
    int seeded = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        a.prefetch(s_mem + seeded, nmem - seeded, vo, lane, SVM_T1_GT, SVM_T1_GV * 5 / 8, SVM_T1_GC * 5 / 8, sync);
        seeded = nmem;
        if (lane == 0) {
            for (int i = 1; i < nmem; ++i) {          // insertion sort
                int x = s_mem[i], j = i - 1;
                while (j >= 0 && s_mem[j] > x) { s_mem[j + 1] = s_mem[j]; --j; }
                s_mem[j + 1] = x;
            }
            a.fail = false;
            Heap heap;
            heap.a = s_heap;
            heap.n = 0;
            bool heap_full = false;
            VisitStats st{0, 0};
            int mi = 0, last = -1, err = 0, nvis = 0;
            while (true) {
                int cu = mi < nmem ? s_mem[mi] : 0x7fffffff;
                int hu = heap.n > 0 ? heap.a[0] : 0x7fffffff;
                if (cu == 0x7fffffff && hu == 0x7fffffff) break;
                int u;
                if (hu < cu) { u = hu; heap.pop(); }
                else { u = cu; ++mi; }
                if (u <= last) continue;
                last = u;
                ++nvis;
                err = visit_vertex(a, vo + u, u, vo, nv, S->co, S->nc, S->R, S->thr, S->thr_cos, d.cap, s_tmp, st,
                                   [&](int x) {
                                       if (heap.n < SVM_T1_GV) heap.push(x);
                                       else heap_full = true;
                                   },
                                   [&](int, int, const int*) {});
                if (err || a.failed() || heap_full) break;
            }
            s_status = (err || a.failed() || heap_full) ? 1 : 0;
            S->n_grp_attempts += 1;
            S->n_grp_visits += nvis;
            if (attempt == 0) S->n_grp_members += nmem;
            if (nvis > S->max_grp_visits) S->max_grp_visits = nvis;
            s_flips = st.flips;
            s_skipped = st.skipped;
            s_nabs = 0;
        }
        __syncthreads();
        if (s_status) {
            if (lane == 0) S->n_seq_reason[2] += 1;   // out of room (or a ring overflow)
            return;
        }
        a.check_touched(gen, d.t1_flagged, s_absorb, &s_nabs, SVM_T1_GABS, lane, SVM_T1_GT);
        __syncthreads();
        const int nabs = s_nabs;
        if (nabs == 0) {
            const bool fit = a.template write_log<SVM_T1_GV, SVM_T1_GC>(G.vid, G.vflag, G.vpos, G.vvn, G.vvcn, G.cid,
                                                                        G.ccnt, G.cring, &s_nlv, &s_nlc, lane, SVM_T1_GT,
                                                                        sync);
            if (lane == 0 && fit) {
                G.nvert = s_nlv;
                G.ncell = s_nlc;
                G.flips = s_flips;
                G.skipped = s_skipped;
                G.ok = 1;
            }
            return;
        }
        if (nabs > SVM_T1_GABS) {
            if (lane == 0) S->n_seq_reason[2] += 1;
            return;
        }
        // absorb the runs met: flag them, then add every candidate each covers
        if (lane == 0) {
            int nnew = 0;
            for (int k = 0; k < nabs; ++k) {
                const int o = s_absorb[k];
                if (d.t1_flagged[o] == gen) continue;
                d.t1_flagged[o] = gen;
                s_absorb[nnew++] = o;
            }
            s_nabs = nnew;
        }
        __syncthreads();
        const int nnew = s_nabs;
        for (int i = lane; i < total; i += SVM_T1_GT) {
            const int ld = d.t1_lead[i];
            bool hit = false;
            for (int k = 0; k < nnew; ++k) hit |= s_absorb[k] == ld;
            if (!hit) continue;
            const int m = atomicAdd(&s_nmem, 1);
            if (m < SVM_T1_GMAX) s_mem[m] = d.t1_list[i] - vo;
        }
        __syncthreads();
        if (s_nmem > SVM_T1_GMAX) {
            if (lane == 0) S->n_seq_reason[1] += 1;
            return;
        }
        nmem = s_nmem;
        a.restore(lane, SVM_T1_GT);
        __syncthreads();
    }
    if (lane == 0) S->n_seq_reason[2] += 1;
}

// ------------------------------------------------------ sequential pass ----
// Closes the T1 step of each sphere (block per sphere): commits the dry-run
// logs of spheres whose candidates did not overlap; otherwise applies the
// flips in the order of the reference's sequential loop, over the candidates
// of the bit mask plus every later vertex a flip disturbed (a min-heap).  If
// the heap fills up, every remaining vertex is visited in order (still
// exact).
// How a sphere's T1 step ends, from the state the detection and dry runs
// left: the sequential pass (run), a commit of the dry-run logs (commit,
// with or without the group log), or nothing.  Read-only, so the large-sphere
// commit and k_t1_resolve reach the same decision.
struct T1Ending {
    int run, commit, group;
};
__device__ inline T1Ending t1_ending(const Dev& d, const SysParams* S, int s, int fire_mode, int force_sequential,
                                    int no_group) {
    const bool frozen = fire_mode && S->f_frozen;
    const int any = S->t1_flag;
    const int group_ok = !force_sequential && !no_group && S->t1_overlap && !S->t1_overflow && d.t1_group[s].ok;
    T1Ending e;
    e.run = !frozen && any && (force_sequential || S->t1_overflow || (S->t1_overlap && !group_ok));
    e.commit = !frozen && any && !e.run;
    e.group = group_ok;
    return e;
}

// Spheres of at least SVM_COMMIT_LARGE vertices commit with a 1024-thread
// block (their logs and the vertex lists to rebuild grow with the sphere),
// before k_t1_resolve, which then only closes their step.
#define SVM_COMMIT_LARGE 65536
__global__ __launch_bounds__(1024) void k_t1_commit_large(Dev d, int pcur, int nin, int nout, int dir_mode,
                                                          int fire_mode, int force_sequential) {
    __shared__ int sh_commit, sh_group;
    const int s = d.s0 + (int)blockIdx.x;
    SysParams* S = d.sys + s;
    if (S->nv < SVM_COMMIT_LARGE) return;
    if (threadIdx.x == 0) {
        const T1Ending e = t1_ending(d, S, s, fire_mode, force_sequential, 0);
        sh_commit = e.commit;
        sh_group = e.group;
    }
    __syncthreads();
    if (!sh_commit) return;
    commit_sphere<1024>(d, S, pcur, nin, nout, dir_mode, sh_group);
    if (threadIdx.x == 0 && sh_group) S->n_t1_group += 1;
}

// check_blocked = 1 closes the step without k_t1_group (which then is not
// launched; spheres of SVM_COMMIT_LARGE vertices or more need it): the block
// of each sphere first checks the sphere's blocked candidates against the
// dry runs' stamps, as k_t1_group does, over the blocked list's length that
// k_t1_cluster copied (the live counter is reset below), and a sphere whose
// runs overlap takes the sequential pass instead of a group run.  Both give
// the sequential order, so the results are the same.
__global__ __launch_bounds__(SVM_TILE) void k_t1_resolve(Dev d, int pcur, int nin, int nout, int dir_mode,
                                                         int advance_step, int fire_mode, int force_sequential,
                                                         int check_blocked) {
    __shared__ int heapbuf[SVM_HEAP_CAP];
    __shared__ int moved[SVM_LIST_CAP];
    __shared__ int touched[SVM_LIST_CAP];
    __shared__ int scan[SVM_TILE];
    __shared__ int tmp[72];
    __shared__ int sh_run, sh_commit, sh_group, sh_nmoved, sh_ntouched, sh_hit;
    const int s = d.s0 + (int)blockIdx.x;
    SysParams* S = d.sys + s;
    const int t = threadIdx.x;
    if (check_blocked && !force_sequential) {
        if (t == 0) sh_hit = 0;
        __syncthreads();
        t1_check_sphere(d, S, S->vo, S->nv, t, SVM_TILE, &sh_hit, *d.t1_btotal);
        __syncthreads();
        if (t == 0) {
            if (sh_hit) S->t1_overlap = 1;
            if (S->t1_overflow & 1) S->n_seq_reason[0] += 1;
            if (S->t1_overflow & 2) S->n_seq_reason[3] += 1;
        }
        __syncthreads();
    }
    if (t == 0) {
        const bool frozen = fire_mode && S->f_frozen;
        const T1Ending e = t1_ending(d, S, s, fire_mode, force_sequential, check_blocked);
        sh_run = e.run;
        sh_commit = e.commit && S->nv < SVM_COMMIT_LARGE;   // large spheres committed by k_t1_commit_large
        sh_group = e.group;
        if (!frozen) {
            S->t1_flag = 0;
            if (fire_mode) S->mix_pending = 0;
        }
        sh_nmoved = 0;
        sh_ntouched = 0;
        if (blockIdx.x == 0) {
            *d.t1_count = 0;   // the lane's candidate lists of this step are used up
            *d.t1_bcount = 0;
        }
    }
    __syncthreads();
    if (!sh_run) {
        if (sh_commit) {
            commit_sphere<SVM_TILE>(d, S, pcur, nin, nout, dir_mode, sh_group);
            if (t == 0 && sh_group) S->n_t1_group += 1;
        }
        if (t == 0) {
            S->t1_overlap = 0;
            S->t1_overflow = 0;
            S->t1_head = -1;
            S->t1_bskips = 0;
            S->t1_gen += 1;
            if (advance_step) S->step += 1;
        }
        return;
    }
    const int vo = S->vo, nv = S->nv;
    const int nw = (nv + 63) >> 6;
    const int wbase = vo >> 6;
    int ncand = 0;
    for (int base = 0; base < nw; base += SVM_TILE) {
        unsigned long long w = (base + t < nw) ? d.t1mask[wbase + base + t] : 0ull;
        int cnt = __popcll(w);
        scan[t] = cnt;
        __syncthreads();
        for (int off = 1; off < SVM_TILE; off <<= 1) {
            int add = t >= off ? scan[t - off] : 0;
            __syncthreads();
            scan[t] += add;
            __syncthreads();
        }
        int pos = ncand + scan[t] - cnt;
        int total = scan[SVM_TILE - 1];
        while (w) {
            int b = __ffsll((long long)w) - 1;
            d.cand[vo + pos++] = (base + t) * 64 + b;
            w &= w - 1;
        }
        __syncthreads();
        ncand += total;
    }
    __threadfence_block();
    __syncthreads();
    if (t == 0) {
        S->n_t1_sequential += 1;
        GlobalAcc a{&d, pcur};
        Heap heap{heapbuf, 0};
        int ci = 0, last = -1, nm = 0, nt = 0, err = 0;
        bool full_scan = false;
        VisitStats st{0, 0};
        auto push = [&](int x) {
            if (full_scan) return;
            if (heap.n < SVM_HEAP_CAP) heap.push(x);
            else full_scan = true;
        };
        auto on_flip = [&](int v1, int v2, const int* cs) {
            if (nm + 2 <= SVM_LIST_CAP) { moved[nm++] = v1; moved[nm++] = v2; }
            else nm = SVM_LIST_CAP + 1;
            for (int k = 0; k < 4; ++k) {
                if (nt < SVM_LIST_CAP) touched[nt++] = cs[k];
                else nt = SVM_LIST_CAP + 1;
            }
        };
        while (true) {
            int u;
            if (!full_scan) {
                int cu = ci < ncand ? d.cand[vo + ci] : 0x7fffffff;
                int hu = heap.n > 0 ? heap.a[0] : 0x7fffffff;
                if (cu == 0x7fffffff && hu == 0x7fffffff) break;
                if (hu < cu) { u = hu; heap.pop(); }
                else { u = cu; ++ci; }
                if (u <= last) continue;
            } else {
                u = last + 1;
                if (u >= nv) break;
            }
            last = u;
            err = visit_vertex(a, vo + u, u, vo, nv, S->co, S->nc, S->R, S->thr, S->thr_cos, d.cap, tmp, st, push, on_flip);
            if (err) break;
        }
        S->n_t1 += st.flips;
        S->n_t1_skipped += st.skipped;
        sh_nmoved = nm;
        sh_ntouched = nt;
        if (err) atomicMax(&S->err, err);
    }
    __threadfence_block();
    __syncthreads();
    const int ntouch = sh_ntouched;
    if (ntouch > SVM_LIST_CAP) {
        for (int x = vo + t; x < vo + nv; x += SVM_TILE)
            if (!build_vln(d, x)) atomicMax(&S->err, 5);
    } else {
        for (int i = t; i < ntouch; i += SVM_TILE) {
            int c = touched[i];
            int n = d.cnn[c];
            for (int k = 0; k < n; ++k)
                if (!build_vln(d, ringref(d, c, k))) atomicMax(&S->err, 5);
        }
    }
    if (dir_mode) {
        const int nmv = sh_nmoved;
        const double* X = d.px[pcur];
        const double* Y = d.py[pcur];
        const double* Z = d.pz[pcur];
        const int hi = nmv > SVM_LIST_CAP ? nv : nmv;
        for (int i = t; i < hi; i += SVM_TILE) {
            int x = nmv > SVM_LIST_CAP ? vo + i : moved[i];
            double p[3] = {X[x], Y[x], Z[x]};
            double nI[3] = {d.nx[nin][x], d.ny[nin][x], d.nz[nin][x]}, nO[3];
            director_update(S, x - vo, p, nI, nO);
            d.nx[nout][x] = nO[0]; d.ny[nout][x] = nO[1]; d.nz[nout][x] = nO[2];
        }
    }
    __syncthreads();
    if (t == 0) {
        S->t1_overlap = 0;
        S->t1_overflow = 0;
        S->t1_head = -1;
        S->t1_bskips = 0;
        S->t1_gen += 1;
        if (advance_step) S->step += 1;
    }
}

}  // namespace svm

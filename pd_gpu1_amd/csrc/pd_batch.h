// pd_batch.h -- device view of a batch of independent 2D PMB specimens.
#pragma once
#include "pd_internal.h"

struct BatchSpec {
    int n, n_int, n_top, n_bot;  // particles, integrated particles, strip counts
    int64_t p_off;               // first particle in the concatenated arrays
    int64_t rp_off;              // first row_ptr entry (each specimen stores n+1)
    int64_t b_off;               // first bond; a multiple of 32 (fail words align)
    int64_t nb;                  // bonds of this specimen
};

// Constants of the bond coefficients recomputed in the kernel (bond_coef).
struct BatchCoef {
    double delta, dx, vol, bc;
};

struct BatchView {
    int nspec;
    const BatchSpec *spec;
    const int32_t *order;  // queue ticket -> specimen, largest bond count first
    const double *x, *y;
    double *ux, *uy, *fx, *fy, *vx, *vy, *vhx, *vhy, *vhox, *vhoy, *fox, *foy;
    double *dupx, *dupy;  // previous increment of u (warm start), then the start state
    double2 *pos;         // per particle (x + ux, y + uy): the one 16-byte gather of a bond
    int warm;             // 1: start each increment from u + previous increment
    int32_t *retired;     // non-null: skip specimens marked non-zero, mark those that end unconverged
                          // (1) or separated (2)
    int nqueue;           // queue tickets: order[0 .. nqueue-1]
    int resume;           // 1: the queued specimens continue an increment paused at iteration it0[s]
    const int32_t *it0;   // resume: the iteration each specimen paused at
    int pause_below;      // > 0: once the queue is empty and at most this many specimens are still
                          // running, they pause at their next iteration (0: never)
    int *active;          // specimens of this launch not yet ended or paused
    int32_t *paused;      // [nspec] 1 for a specimen this launch paused
    double sep;           // > 0: a specimen whose strip reaction falls below sep * react[s] has separated
    double *react;        // [nspec] strip reaction |F| at the specimen's last converged increment (0: none)
    const int32_t *row_ptr, *col;
    const double2 *kd;    // per bond {kappa, |xi|}
    const double2 *xy;    // per particle reference position (x, y)
    int recompute;        // 1: kappa and |xi| recomputed from xy (bit-equal to kd, checked at set-up)
    BatchCoef coef;
    uint32_t *fail;
    double mass, dt, s0;
};

// Several workgroups per specimen (k_batch_adr): workgroups per group, and
// per group an arrival counter, the specimen slot and the parts' damping sums.
// A part owns block/parts lanes of the damping tree (a power of two, at least
// kMinPartLanes); the first launch of an increment uses at most 16 parts,
// the tail launches up to kMaxParts.
constexpr int kMaxParts = 64;
constexpr int kMaxFirstParts = 16;
constexpr int kMinPartLanes = 16;
struct BatchGroups {
    int parts;      // workgroups per specimen: a power of two <= min(kMaxParts, block/kMinPartLanes)
    unsigned *bar;  // [groups] arrival counters, zero at launch
    int *spec;      // [groups] the group's current specimen
    double *sum;    // [groups][parts][4] subtree roots of the damping tree
    int *flag;      // [groups] part 0's pause decision, read by every part after the group barrier
    int *abort;     // set when a group barrier times out
    uint64_t spin_ticks;  // longest wait at a group barrier, in wall-clock ticks
    int xcd;        // 1: the grid is a multiple of 8 x parts and each group's workgroups share an XCD
};

hipError_t launch_pack_bond_pairs(hipStream_t st, int64_t nb, const double *kappa, const double *idist,
                                  double2 *kd);
hipError_t launch_batch_coef_check(hipStream_t st, int n, const int32_t *row_ptr, const int32_t *col,
                                   const double2 *kd, const double2 *xy, const BatchCoef &cg, unsigned *bad);
// Launches with at most this many workgroups per specimen use the dense
// build of the kernel (two 1024-thread workgroups per CU) when it is resident
// more often than the default one.
constexpr int kDenseParts = 4;
int batch_resident_per_cu(int block, int group, bool recompute, bool dense);
hipError_t launch_batch_adr(hipStream_t st, int grid, int block, int group, bool dense, const BatchView &B,
                            double du, double tol, int max_iters, unsigned *queue, int32_t *iters_out,
                            double *resid_out, const BatchGroups &gr);

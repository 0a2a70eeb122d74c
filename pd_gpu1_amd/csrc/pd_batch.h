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

struct BatchView {
    int nspec;
    const BatchSpec *spec;
    const int32_t *order;  // queue ticket -> specimen, largest bond count first
    const double *x, *y;
    double *ux, *uy, *fx, *fy, *vx, *vy, *vhx, *vhy, *vhox, *vhoy, *fox, *foy;
    double *dupx, *dupy;  // previous increment of u (warm start), then the start state
    double2 *pos;         // per particle (x + ux, y + uy): the one 16-byte gather of a bond
    int warm;             // 1: start each increment from u + previous increment
    int32_t *retired;     // non-null: skip specimens marked 1, mark those that end unconverged
    const int32_t *row_ptr, *col;
    const double2 *kd;    // per bond {kappa, |xi|}
    uint32_t *fail;
    double mass, dt, s0;
};

// Several workgroups per specimen (k_batch_adr): workgroups per group, and
// per group an arrival counter, the specimen slot and the parts' damping sums.
constexpr int kMaxParts = 16;
struct BatchGroups {
    int parts;      // workgroups per specimen: a power of two <= min(16, block/64)
    unsigned *bar;  // [groups] arrival counters, zero at launch
    int *spec;      // [groups] the group's current specimen
    double *sum;    // [groups][parts][4] subtree roots of the damping tree
    int *abort;     // set when a group barrier times out
    uint64_t spin_ticks;  // longest wait at a group barrier, in wall-clock ticks
};

hipError_t launch_pack_bond_pairs(hipStream_t st, int64_t nb, const double *kappa, const double *idist,
                                  double2 *kd);
int batch_resident_per_cu(int block, int group);
hipError_t launch_batch_adr(hipStream_t st, int grid, int block, int group, const BatchView &B, double du,
                            double tol, int max_iters, unsigned *queue, int32_t *iters_out, double *resid_out,
                            const BatchGroups &gr);

/*
 * pd_gpu.h -- C-ABI of the MI355X peridynamic time-step engine (gfx950).
 *
 * Drop-in boundary for the hot path of XiaoMmingW/PD_Gpu1: the stages its
 * solvers call (Coord -> Neighbor -> Surf_Corr -> Initial -> Force ->
 * Integrate/Reduce -> Save).  Every entry point below names the reference
 * function or kernel it replaces (file:line in the reference tree).
 *
 * Conventions
 *  - All functions are extern "C", take an opaque pd_handle* and return an int
 *    status (PD_OK = 0).  pd_last_error() returns the message of the last
 *    failure on that handle.  The reference's CHECK() printed and exit(1)ed
 *    (Error.cuh:9-23); nothing here exits the process.
 *  - The handle owns all device memory and one HIP stream.  Host buffers are
 *    caller-owned: inputs are copied in, outputs are written to caller
 *    buffers of the documented size, no pointer is retained.
 *  - A handle is single-threaded; distinct handles are independent.  There are
 *    no globals and no __constant__ state (the reference keeps one problem per
 *    process in `dp`/`p`, main.cu:6-9).
 *  - Per-particle vector fields are SoA doubles laid out [dim][N].
 *  - Bonds are CSR: row_ptr[N+1] and col[Nb] (int32), each row ascending (the
 *    slot order of the reference's CPU neighbour search, Neighbor.cu:44-58).
 *  - Fail state is a bit mask: bit (k & 31) of word k >> 5 is 1 while directed
 *    bond k is intact (the reference's int fail[] slot, Initial.cu:6-14).
 */
#ifndef PD_GPU_H
#define PD_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PD_OK 0
#define PD_E_INVALID 1     /* bad argument or call order */
#define PD_E_HIP 2         /* HIP runtime error */
#define PD_E_CAPACITY 3    /* neighbour count above max_neighbors (MN) or size limit */
#define PD_E_NONFINITE 4   /* NaN/Inf met in a reduction */
#define PD_E_NOT_CONVERGED 5

/* models (Dselect::B_S, Base.cuh:40) */
#define PD_MODEL_PMB 1
#define PD_MODEL_LPS 2

/* arithmetic of the force kernels */
#define PD_PREC_FP64 0     /* FP64 everywhere, bonds summed in CSR order */
#define PD_PREC_MIXED 1    /* FP32 bond math with compensated sums; FP64 state,
                              FP64 re-evaluation of near-critical bonds */

/* fields for pd_get_field / pd_set_field */
#define PD_FIELD_X0 0      /* reference coordinates [dim][N] */
#define PD_FIELD_U 1       /* displacement [dim][N] */
#define PD_FIELD_V 2       /* velocity [dim][N] */
#define PD_FIELD_A 3       /* acceleration [dim][N] */
#define PD_FIELD_F 4       /* pair-force density [dim][N] (FP32 in mixed mode, returned widened) */
#define PD_FIELD_B 5       /* body force [dim][N] */
#define PD_FIELD_VH 6      /* ADR half-step velocity [dim][N] */
#define PD_FIELD_VHO 7     /* ADR previous half-step velocity [dim][N] */
#define PD_FIELD_FO 8      /* ADR previous force [dim][N] */
#define PD_FIELD_DMG 9     /* damage [N] */
#define PD_FIELD_THETA 10  /* LPS dilatation [N] */
#define PD_FIELD_M 11      /* LPS weighted volume [N] */
#define PD_FIELD_IDIST 20  /* bond reference length xi [Nb] */
#define PD_FIELD_FAC 21    /* partial-volume factor [Nb] */
#define PD_FIELD_SCR 22    /* surface-correction factor [Nb] */
#define PD_FIELD_W 23      /* LPS influence weight [Nb] */
#define PD_FIELD_E 24      /* LPS extension of the last theta pass [Nb] (FP64 mode) */

/* init modes for pd_init_state */
#define PD_INIT_DYNAMIC 0  /* base_integrate_initial_* (Initial.cu:24-55,82-112) */
#define PD_INIT_STATIC 1   /* static_initial_bond_* (Initial.cu:57-75,114-132) */

/* Model parameters: the device_parameter the reference copies to __constant__
 * dp (Base.cuh:46-68) plus the switches its code hard-codes. */
typedef struct pd_params {
    int32_t dim;               /* 2 or 3 (Dselect::Dim) */
    int32_t model;             /* PD_MODEL_PMB or PD_MODEL_LPS */
    int32_t precision;         /* PD_PREC_FP64 or PD_PREC_MIXED */
    int32_t max_neighbors;     /* MN: 32 in 2D, 128 in 3D (Base.cuh:27-29) */
    int32_t theta_skip_broken; /* 1: dilatation skips broken bonds (Force_State.cu:345) */
    int32_t reduce_lanes;      /* power-of-two lane count T of the fixed reduction tree */
    int32_t n_top, n_bot;      /* strip particles stored after N_int (Coord.cu:47-66) */
    double dx, delta, vol, bc, mass, K, G;
    double emod, pratio, dens, s0, dt;
    double break_y_min;        /* 3D breaks need y_i, y_j >= this (Force_Bond.cu:117: -0.04) */
    double strip_velocity;     /* strip load speed (Solve_Plate_Crack_Grow.cu:128: 20 m/s) */
    double near_window;        /* mixed mode: every bond with |s-s0| <= near_window*s0, and any whose
                                  FP32 stretch could err across s0, is re-evaluated in FP64 */
} pd_params;

/* Cylinder impactor (Cylinder_Parameter, Base.cuh:227-236). */
typedef struct pd_cyl {
    double rad, mass, height, v, x, y, z;
} pd_cyl;

typedef struct pd_handle pd_handle;

/* ---- parameters: BaseModel_Parameter::cal_parameter (Base.cuh:114-183) ----
 * Derives delta, nx/ny/nz, bc, vol, mass, K, G exactly as the reference does.
 * delta_factor is 3.015 in the reference (Base.cuh:116).  ref_k_bug = 1 keeps
 * the 3D LPS K = E/3/(1-2G) of Base.cuh:175; 0 uses K = E/(3(1-2 nu)).
 * Writes lattice counts to nxyz[3] and the particle count nx*ny*nz to *n. */
int pd_params_init(pd_params *p, int dim, int model, double dt, double dx, double thick,
                   double length, double height, double width, double emod,
                   double pratio, double dens, double s0, double delta_factor,
                   int ref_k_bug, int *nxyz, int *n);

/* ---- coordinate generators (host, bit-identical to the reference) -------- */
int pd_gen_plate(int nx, int ny, double dx, double *x, double *y);          /* Coord.cu:19-31 */
int pd_gen_plate_strips(int nx, int ny, double dx, double *x, double *y);   /* Coord.cu:47-66 */
/* Coord.cu:77-129.  Returns the particle count; x/y/z may be NULL to count. */
int pd_gen_kw(int nx, int ny, int nz, double dx, double rad, double *x, double *y, double *z);
/* Jittered cuboid lattice, SplitMix64 stream (no reference counterpart). */
int pd_gen_block_jitter(int nx, int ny, int nz, double dx, double jitter, uint64_t seed,
                        double *x, double *y, double *z);

/* ---- memory: Base_Allocate / Base_Free (Memory.cu:139-147,276-284) ------- */
int pd_create(const pd_params *p, int device, pd_handle **out);
int pd_destroy(pd_handle *h);
const char *pd_last_error(pd_handle *h);
int pd_get_params(pd_handle *h, pd_params *out);
int pd_set_params(pd_handle *h, const pd_params *p); /* scalars only (dt, s0, ...) */

/* ---- coordinates: H2D copy of Solve_Kalthoff_Winkler.cu:381-383 ---------- */
int pd_set_positions(pd_handle *h, const double *x, const double *y, const double *z,
                     int n_int, int n);

/* ---- neighbours: kernel_find_neighbor_2D/3D (Neighbor.cu:7-40) ----------- */
int pd_build_neighbors(pd_handle *h);            /* cell list -> sorted CSR */
int pd_num_bonds(pd_handle *h, int64_t *nb);
int pd_get_csr(pd_handle *h, int32_t *row_ptr, int32_t *col);

/* ---- volume and surface correction (Surf_Corr.cu:7-35, 235-310) --------- */
int pd_vol_corr(pd_handle *h);
int pd_surface_correct(pd_handle *h);            /* sets scr; scr = 1 until called */

/* ---- initial state (Initial.cu:6-75) ------------------------------------- */
int pd_init_state(pd_handle *h, int mode);

/* ---- cracks: kernel_set_crack_2D (Base_Function.cu:39-81),
 *      kernel_set_crack (Solve_Kalthoff_Winkler.cu:40-63) ------------------ */
int pd_set_crack_2d(pd_handle *h, double clength, double loc_x, double loc_y, double theta);
int pd_set_notches_kw(pd_handle *h, double rad);

/* ---- LPS weights: kernel_initial_weight_2D/3D (Force_State.cu:8-54) ------ */
int pd_lps_weights(pd_handle *h);

/* ---- force: kernel_particle_force_2D / kernel_bond_force_3D
 *      (Force_Bond.cu:52-141); kernel_cal_theta_* + kernel_state_force_*
 *      (Force_State.cu:56-306).  PMB rows i < N; LPS theta i < N, force
 *      i < N_int.  Bonds with s > s0 are marked broken after their force is
 *      added (Force_Bond.cu:32). ------------------------------------------- */
int pd_force(pd_handle *h);

/* ---- explicit integration: kernel_integrate_{2D,3D}_{1,2}
 *      (Integrate.cu:174-249) and kernel_quick_load
 *      (Solve_Plate_Crack_Grow.cu:8-19, applied inside half 1) ------------- */
int pd_verlet_half1(pd_handle *h);
int pd_verlet_half2(pd_handle *h);
/* One full explicit step as the reference time loops order it
 * (Solve_Plate_Crack_Grow.cu:126-149, Solve_Kalthoff_Winkler.cu:407-430),
 * repeated nsteps times from a captured HIP graph. */
int pd_run_verlet(pd_handle *h, int nsteps);

/* ---- ADR: kernel_cnn + cal_cn + sum + kernel_integrate
 *      (Integrate.cu:7-129, Reduce.cu:11-65) ------------------------------- */
int pd_set_body_force_tensile(pd_handle *h, int nx, int ny); /* Solve_Plate_Tensile.cu:7-17 */
/* One ADR iteration: force, fused damping/residual reduction on the device,
 * update.  c_out / resid_out may be NULL (then no host synchronisation). */
int pd_adr_iterate(pd_handle *h, int ct, double *c_out, double *resid_out);
/* Iterate from ct0 until the residual <= tol or max_iters; device-side stop. */
int pd_adr_solve(pd_handle *h, int ct0, int max_iters, double tol,
                 int *iters_out, double *resid_out);
/* Strip displacement increment for displacement-controlled loading. */
int pd_strip_displace(pd_handle *h, double du_top, double du_bot);
/* Reduce.cu:52-65 sum() test hook: fixed-tree sum of field component comp. */
int pd_reduce_sum(pd_handle *h, int field, int comp, double *out);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the fixed-tree sum of pd_reduce_sum over n doubles at the
 * device pointer `values`, with `lanes` lanes (a power of two, as
 * pd_params.reduce_lanes; equal lanes give the bits pd_reduce_sum gives),
 * written to the device double `out`, enqueued on `stream` (a hipStream_t,
 * NULL = the null stream) on GPU `device`; nothing waits for it.  It
 * allocates nothing: the block partials live in a 32 KiB device array the
 * library keeps on each GPU until the process ends, so lanes above
 * 4096 * 256 return PD_E_CAPACITY, and calls that overlap on one GPU (on two
 * streams at once) race for that array.  PD_E_INVALID: null pointers, n < 0,
 * lanes not a power of two >= 64, or a bad device. */
int pd_reduce_sum_noalloc(int device, const double *values, int64_t n, int lanes, double *out, void *stream);

/* ---- damage and fields (Base_Function.cu:149-217) ------------------------ */
int pd_damage(pd_handle *h);
int pd_get_field(pd_handle *h, int field, double *out);
int pd_set_field(pd_handle *h, int field, const double *in);
int pd_get_fail_bits(pd_handle *h, uint32_t *out);   /* (Nb+31)/32 words */
int pd_set_fail_bits(pd_handle *h, const uint32_t *in);

/* ---- KW cylinder and contact (Solve_Kalthoff_Winkler.cu:9-37,89-265) ---- */
int pd_cylinder_init(pd_handle *h, const pd_cyl *c);
int pd_cylinder_get(pd_handle *h, double *state12); /* vel[3] disp[3] acc[3] force[3] */
int pd_contact_step(pd_handle *h);  /* cylinder half 1 is in pd_verlet_half1 */

/* ---- output (Save.cu:6-140) --------------------------------------------- */
#define PD_LAYOUT_PLATE 0
#define PD_LAYOUT_KW 1
int pd_save_text(pd_handle *h, const char *path, int layout, int nx, int ny, int nz, double rad);
int pd_checkpoint_write(pd_handle *h, const char *path, int64_t step);
int pd_checkpoint_read(pd_handle *h, const char *path, int64_t *step);

/* ---- batched quasi-static crack growth (no reference counterpart; built
 *      from the ADR path of solve_plate_tensile, Solve_Plate_Tensile.cu:106-131,
 *      the strip layout of Coord.cu:47-66 and the crack of
 *      Base_Function.cu:83-123) ---------------------------------------------
 * Independent square 2D PMB plates with 3 strip rows above and below and a
 * centred pre-crack, relaxed by ADR in one persistent launch: one workgroup
 * per specimen, specimens handed out by a device work queue.  All specimens
 * share the material, dx, dt and mass of *p. */
typedef struct pd_specimen {
    int32_t nx, ny;            /* plate lattice (strips add 3*nx above and below) */
    double crack_len;          /* centred crack length (m); <= 0: none */
    double crack_angle;        /* crack angle (rad) */
} pd_specimen;
typedef struct pd_batch pd_batch;
int pd_batch_create(const pd_params *p, const pd_specimen *specs, int nspec, int device, pd_batch **out);
int pd_batch_destroy(pd_batch *b);
const char *pd_batch_last_error(pd_batch *b);
int pd_batch_info(pd_batch *b, int64_t *n_particles, int64_t *n_bonds, int *nspec);
int pd_batch_spec_info(pd_batch *b, int spec, int *n, int *n_int, int64_t *nb);
/* One load increment: top strips += du, bottom strips -= du, then ADR from
 * rest (first iteration undamped, Integrate.cu:65-76) until the residual
 * ||F||(free) / ||F||(strips) <= tol or max_iters updates.  block = threads
 * per workgroup (256, 512, 1024), grid = workgroups (0 = fill the device).
 * iters_out/resid_out[nspec] receive the updates done and the final residual. */
int pd_batch_increment(pd_batch *b, double du, double tol, int max_iters, int block, int grid,
                       int32_t *iters_out, double *resid_out, double *ms_out);
/* Force phase of the batched kernel: 1 thread per particle (default) or
 * 4, 8, 16, 32 lanes per particle; every choice adds a row's bond terms in
 * bond order, so results are identical. */
int pd_batch_set_group(pd_batch *b, int group);
/* Workgroups per specimen (no reference counterpart): 1, 2, 4, 8 or 16 (at
 * most block/64), or 0 (default) = the smallest power of two W at which the
 * largest active specimen's work (bonds x sqrt(particles), as its ADR
 * iterations grow with its side) over W workgroups is no more than all
 * active work over the workgroups the device holds (the largest specimen
 * then no longer sets the wall time alone).  With 1024-thread workgroups and
 * W <= 4 an increment's first launch runs a build of the kernel that holds
 * two workgroups per CU (and W is chosen for that capacity); with a larger W
 * and tail launches on (pd_batch_set_tail) the first launch takes W/2, as
 * the tail launches widen the slowest specimens later.  Results are the
 * same for every W.  The workgroups of a specimen
 * meet at barriers in global memory and are all resident (the grid is
 * capped); the damping sums keep the order of one workgroup, so results are
 * identical for every choice.  Other kernels running on the device at the
 * same time (another batch's increment on another stream) can keep a group's
 * workgroups from all being resident: the barrier timeout then ends the
 * increment with an error, so run such increments one at a time per device
 * or with parts = 1.  pd_batch_get_parts reports the value the last
 * increment ran with. */
int pd_batch_set_parts(pd_batch *b, int parts);
int pd_batch_get_parts(pd_batch *b, int *parts_used);
/* Longest wait (seconds, default 20) of a workgroup at a group barrier; a
 * longer wait abandons the increment: every workgroup leaves the kernel,
 * every specimen's state (displacements, velocities, forces, fail bits,
 * retired flags) is put back to what it was before the increment, and
 * pd_batch_increment returns PD_E_HIP with the reason in
 * pd_batch_last_error.  The batch stays usable: the same increment can be
 * run again, e.g. with parts = 1 (no reference counterpart). */
int pd_batch_set_barrier_timeout(pd_batch *b, double seconds);
/* The automatic workgroups per specimen from the specimens' bonds nb[],
 * particles n[] and retired flags (nullable), for a device that holds
 * `capacity` workgroups and at most wmax per specimen (no reference
 * counterpart; host arithmetic only, no device is touched). */
int pd_batch_choose_parts(const int64_t *nb, const int32_t *n, const int32_t *retired, int nspec, int capacity,
                          int wmax);
/* Warm start of the batched increments (no reference counterpart): 0 = plain
 * ADR from the previous solution with the strips moved (default), 1 = the
 * free particles also move by the previous increment, which is exact while
 * the response is linear; ADR still converges every specimen to tol. */
int pd_batch_set_warm(pd_batch *b, int mode);
/* Bond coefficients of the batched force phase (no reference counterpart):
 * -1 (default) = recompute each bond's {kappa, |xi|} from the reference
 * positions with the set-up's own expressions (Surf_Corr.cu:312-337 and
 * kappa = vol*fac*bc*scr) when pd_batch_create found that recomputation
 * equal to the stored values bit for bit (no HBM stream of coefficients),
 * 0 = always stream them, 1 = recompute (PD_E_INVALID when they differ).
 * Results are identical either way.  pd_batch_coef_info: *recomputed = 1
 * when the next increment recomputes them. */
int pd_batch_set_coef_mode(pd_batch *b, int mode);
int pd_batch_coef_info(pd_batch *b, int *recomputed);
/* on = 1: a specimen whose increment ends without reaching tol (it has
 * separated, or needs more than max_iters) is retired; later increments skip
 * it and report 0 iterations (no reference counterpart). */
int pd_batch_set_retire(pd_batch *b, int on);
/* Tail launches (no reference counterpart).  The slowest specimens of an
 * increment set its wall time and would run on a few workgroups while the
 * rest of the device idles: once every specimen has been handed out and at
 * most half of the device's workgroups still run, the running specimens
 * pause at their next iteration and continue in a new launch with at least
 * twice the workgroups each, up to max_parts = 16, 32 (default) or 64 (and
 * at least 16 lanes of the damping tree per workgroup), repeatedly.  A
 * paused specimen resumes at the iteration it paused at, so results are
 * identical to one launch.  max_parts = 0: one persistent launch per
 * increment; off when pd_batch_increment is given a grid.
 * pd_batch_get_tail_parts: workgroups per specimen of the last increment's
 * last tail launch (0: there was none). */
int pd_batch_set_tail(pd_batch *b, int max_parts);
int pd_batch_get_tail_parts(pd_batch *b, int *parts);
/* Separation (no reference counterpart; the reference runs a fixed number of
 * ADR iterations per load step, Solve_Plate_Tensile.cu:106): 0 < ratio < 1
 * ends a specimen's increment when the reaction of its strips falls below
 * ratio x the reaction at its last converged increment -- the plate has
 * fractured through and carries no load, and its relative residual (free
 * forces over the strip reaction) cannot reach tol.  With retire on such a
 * specimen is retired with flag 2.  0 = off (default).
 * pd_batch_get_retired: per specimen 0 = active, 1 = retired unconverged,
 * 2 = retired separated. */
int pd_batch_set_separation(pd_batch *b, double ratio);
int pd_batch_get_retired(pd_batch *b, int32_t *flags);
int pd_batch_get_field(pd_batch *b, int spec, int field, double *out);       /* [2][n] */
int pd_batch_set_field(pd_batch *b, int spec, int field, const double *in);
int pd_batch_get_csr(pd_batch *b, int spec, int32_t *row_ptr, int32_t *col, uint32_t *fail_words);

/* ---- learned ADR warm start (no reference counterpart) --------------------
 * A message-passing network over the PD bond graph: 3 layers, hidden 64,
 * SiLU, BF16 MFMA operands with FP32 accumulation, trained by its own
 * forward/backward kernels with Adam.  Node features [N][16] (the caller's
 * choice; pd_gpu1_amd/warmstart.py: position, previous increment, damage,
 * strip flags, 1, and the node's coordinates in the frame of the pre-crack);
 * edge features [E][3]: reference bond vector / delta (2), intact flag.
 * Output [N][2]: the predicted displacement increment (scaled as the caller
 * scaled the targets). */
typedef struct pd_mpnn pd_mpnn;
int pd_mpnn_create(int device, uint64_t seed, pd_mpnn **out);
int pd_mpnn_destroy(pd_mpnn *m);
const char *pd_mpnn_last_error(pd_mpnn *m);
int pd_mpnn_num_params(pd_mpnn *m, int64_t *n);
int pd_mpnn_get_params(pd_mpnn *m, float *out);
int pd_mpnn_set_params(pd_mpnn *m, const float *in);   /* also resets Adam moments */
/* Symmetric CSR (the bond list; rows ascending). */
int pd_mpnn_set_graph(pd_mpnn *m, int n_nodes, const int32_t *row_ptr, int64_t n_edges, const int32_t *col);
int pd_mpnn_set_features(pd_mpnn *m, const float *node_feat, const float *edge_feat, const float *target,
                         const float *weight);
int pd_mpnn_forward(pd_mpnn *m, float *y_out);
int pd_mpnn_gradient(pd_mpnn *m, float *grad_out, float *loss_out);
int pd_mpnn_train_step(pd_mpnn *m, float lr, float *loss_out);

/* ---- timing and launch shape (no reference counterpart: the reference times
 *      whole runs with gettimeofday, Base_Function.cu:7-12) ---------------- */
int pd_last_force_ms(pd_handle *h, double *ms);       /* events around the last pd_force */
/* Average HIP-event time per pass over reps evaluations: ms_out[0] = PMB kernel
 * or LPS dilatation pass, ms_out[1] = LPS force pass. */
int pd_time_force_passes(pd_handle *h, int reps, double *ms_out);
/* Lanes per particle of the mixed kernels: 4, 8, 16, 32, 64, or 0 for the
 * built-in choice from the mean row length. */
int pd_set_group(pd_handle *h, int group);
/* The same for one pass: pass 0 = PMB or LPS dilatation, 1 = LPS force. */
int pd_set_pass_group(pd_handle *h, int pass, int group);
/* HBM reference: a plain device copy of `bytes` (multiple of 16) repeated
 * reps times; *gbps = 2*bytes*reps/time.  Used to calibrate counters. */
int pd_stream_copy(int device, int64_t bytes, int reps, double *gbps);
/* Work order of the mixed kernels: particles grouped in cubic tiles of edge
 * tile_cells*dx (0 = index order, -1 = default: 5 in 3D and 16 in 2D with
 * LDS-staged tiles, 16 and 64 without).  Results do not depend on it. */
int pd_set_tile(pd_handle *h, int tile_cells);
/* LDS-staged tiles for the mixed kernels: each tile's neighbour records are
 * copied to LDS once per pass and bonds read them through 16-bit local
 * indices.  mode -1 = where every tile fits (default), 0 = never (per-bond
 * gathers from L2), 1 = required (PD_E_CAPACITY otherwise).  Same results. */
int pd_set_lds(pd_handle *h, int mode);
/* Lanes per row of the FP64 PMB kernel: 1 (thread per row), 4, 8, 16, 32, or
 * -1 = 8 in 2D and 16 in 3D (default).  Terms are added in bond order, so the
 * forces and breaks are identical for every choice. */
int pd_set_fp64_group(pd_handle *h, int group);
/* Bond coefficients of the mixed kernels (|xi| and kappa or omega = w fac vol):
 * -1 = recomputed from the bond's reference length when every stored value
 * equals that recomputation to FP32 accuracy, which is checked once per
 * set-up (default; the coefficient stream is then not read), 0 = always
 * streamed, 1 = recomputed (PD_E_INVALID when they differ). */
int pd_set_coef_mode(pd_handle *h, int mode);
/* *recomputed = 1 when the mixed kernels recompute the coefficients. */
int pd_coef_info(pd_handle *h, int *recomputed);
/* Threads per workgroup of the LDS tile kernels: 256, 512 (default) or 1024;
 * for both passes, or for one (pass 0 = PMB or LPS dilatation, 1 = LPS force). */
int pd_set_lds_block(pd_handle *h, int threads);
int pd_set_lds_pass_block(pd_handle *h, int pass, int threads);
/* Whether the LDS tiles are in use, their number and the largest staged set. */
int pd_lds_info(pd_handle *h, int *active, int *ntiles, int *max_staged);
int pd_sync(pd_handle *h);

#ifdef __cplusplus
}
#endif
#endif

/*
 * pd_oracle.h -- FP64 CPU oracle for the peridynamic time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C11 + OpenMP and
 * in double precision, the reference's CPU path for the hot loop of
 * XiaoMmingW/PD_Gpu1 (the `device==1` twins).  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it, and only as the checker or as
 * the reported CPU baseline.  The product library (pd_gpu1_amd) never links,
 * loads or calls anything here.
 *
 * Conventions shared by every function:
 *   - arrays are SoA doubles, one pointer per component (y/z may be NULL when
 *     dim < 3);
 *   - bonds are CSR: row_ptr[N+1] (int64) and col[Nb] (int32), each row sorted
 *     ascending, which is exactly the slot order the reference's CPU neighbour
 *     search produces (Neighbor.cu:44-58);
 *   - fail[] holds one byte per directed bond, 1 = intact, 0 = broken
 *     (Initial.cu:77-80 semantics);
 *   - every floating-point expression keeps the operand order of the reference
 *     line it cites, and the library is compiled with -ffp-contract=off so no
 *     multiply-add is fused.
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Derived model constants, Base.cuh:114-183 (BaseModel_Parameter::cal_parameter). */
typedef struct orc_params {
    int dim;          /* 2 or 3 */
    int bs;           /* 1 = bond-based PMB, 2 = state-based LPS (Dselect::B_S) */
    int nx, ny, nz;   /* lattice counts */
    int N, N_int;     /* all particles, integrated particles */
    int MN;           /* neighbour capacity per particle */
    double dx, delta, radij, sedload, vol, bc, mass, thick, area;
    double emod, pratio, dens, K, G, s0, dt;
} orc_params;

void orc_cal_parameter(orc_params *p, int dim, int bs, double dt, double dx,
                       double thick, double length, double height, double width,
                       double emod, double pratio, double dens, double s0,
                       double delta_factor, int ref_k_bug);

/* Coordinates: Coord.cu:19-31, Coord.cu:47-66, Coord.cu:77-129. */
void orc_coord_plate(int nx, int ny, double dx, double *x, double *y);
void orc_coord_plate_strips(int nx, int ny, double dx, double *x, double *y);
int  orc_coord_kw(int nx, int ny, int nz, double dx, double rad,
                  double *x, double *y, double *z);
/* Jittered lattice block (no reference counterpart; SplitMix64 seeded). */
void orc_coord_block_jitter(int nx, int ny, int nz, double dx, double jitter,
                            uint64_t seed, double *x, double *y, double *z);
uint64_t orc_splitmix64(uint64_t *state);

/* Neighbours.  Brute force is the literal O(N^2) restatement of
 * Neighbor.cu:42-80 into a dense NN/NL (stride MN); returns -1 on overflow. */
int  orc_neighbors_brute(int dim, int N, const double *x, const double *y,
                         const double *z, double delta, int MN, int *NN, int *NL);
/* Cell-list search producing the same rows as CSR.  Pass 1 (col == NULL)
 * fills row_ptr and returns the max row length; pass 2 fills col. */
int  orc_neighbors_cell(int dim, int N, const double *x, const double *y,
                        const double *z, double delta, int64_t *row_ptr, int *col);

/* Volume correction, Surf_Corr.cu:312-337. */
void orc_vol_corr(int dim, int N, const int64_t *row_ptr, const int *col,
                  const double *x, const double *y, const double *z,
                  double delta, double dx, double *idist, double *fac);

/* Surface correction, Surf_Corr.cu:339-524.  scr_mode 0 = trig-free identity
 * (cos^2 = dx^2/(dx^2+dy^2), cos^2 phi = (|dz|/xi)^2), 1 = the literal
 * atan/acos/cos/sin expression.  fncst is [dim][N] scratch (may be NULL). */
void orc_surface_correct(int dim, int N, const int64_t *row_ptr, const int *col,
                         const double *x, const double *y, const double *z,
                         const double *idist, const double *fac,
                         double bc, double vol, double sedload, int scr_mode,
                         double *fncst, double *scr);

/* Cracks: Base_Function.cu:83-123 and Solve_Kalthoff_Winkler.cu:65-87. */
void orc_set_crack_2d(int N_int, const int64_t *row_ptr, const int *col,
                      const double *x, const double *y, double clength,
                      double loc_x, double loc_y, double theta, uint8_t *fail);
void orc_set_crack_kw(int N_int, const int64_t *row_ptr, const int *col,
                      const double *x, const double *y, double rad, uint8_t *fail);

/* PMB force, Force_Bond.cu:159-202.  Rows i < n_rows are evaluated. */
void orc_pmb_force(int dim, int n_rows, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *scr,
                   double vol, double bc, double s0, double break_y_min,
                   uint8_t *fail, double *const *f);

/* LPS weights and m, Force_State.cu:309-321. */
void orc_lps_weights(int N, const int64_t *row_ptr, const double *idist,
                     const double *fac, double delta, double vol, double *w, double *m);
/* LPS dilatation: 2D follows Force_State.cu:339-357, 3D follows the kernel
 * arithmetic at Force_State.cu:218-224.  theta_skip_broken = 1 is the CPU rule. */
void orc_lps_theta(int dim, int N, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *w,
                   const double *m, const uint8_t *fail, double vol, double pratio,
                   int theta_skip_broken, double *e, double *theta);
/* LPS force: 2D Force_State.cu:359-392, 3D Force_State.cu:263-281. */
void orc_lps_force(int dim, int N_int, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *w,
                   const double *m, const double *theta, const double *e,
                   double vol, double pratio, double K, double G, double s0,
                   double break_y_min, uint8_t *fail, double *const *f);

/* Velocity Verlet, Integrate.cu:447-481, and the strip load of
 * Solve_Plate_Crack_Grow.cu:20-27 (n_top/n_bot strip particles after N_int). */
void orc_verlet_half1(int dim, int N_int, double dt, double *const *v,
                      const double *const *a, double *const *u);
void orc_quick_load(int N_int, int n_top, int n_bot, double load, double dt, double *uy);
void orc_verlet_half2(int dim, int N_int, double dt, double dens, double *const *v,
                      double *const *a, const double *const *f, const double *const *b);

/* Canonical reduction order used by the GPU: T (power of two) stride-T
 * sequential partial sums, then a pairwise tree that combines partials at
 * offsets 1, 2, 4, ... (partial[t] += partial[t+off] for t % (2*off) == 0). */
double orc_sum_tree(const double *v, int64_t n, int T);

/* ADR terms of kernel_cnn (Integrate.cu:7-34) per particle. */
void orc_adr_terms(int dim, int N_int, double mass, double dt,
                   const double *const *u, const double *const *f,
                   const double *const *fold, const double *const *vho,
                   double *cn, double *disp);
/* Damping coefficient from the two sums, Integrate.cu:117-128. */
double orc_adr_c(double cn1, double cn2);
/* ADR update, Integrate.cu:54-106. */
void orc_adr_integrate(int dim, int N_int, int ct, double c, double mass, double dt,
                       const double *const *f, const double *const *b,
                       double *const *vh, double *const *v, double *const *vho,
                       double *const *u, double *const *fold);
/* Residual terms: rnum_i = |F+b|^2 for i < N_int; rden_i = |b|^2 for i < N_int
 * and |F|^2 for N_int <= i < N. */
void orc_resid_terms(int dim, int N_int, int N, const double *const *f,
                     const double *const *b, double *rnum, double *rden);

/* Damage, Base_Function.cu:220-238. */
void orc_damage(int N_int, const int64_t *row_ptr, const uint8_t *fail,
                const double *fac, double vol, double *dmg);

/* Kalthoff-Winkler rigid cylinder, Solve_Kalthoff_Winkler.cu:24-37,101-110,
 * 177-239,255-265.  cyl[0..6] = rad, mass, height, v, x, y, z.  The reaction
 * per particle is stored in react[3][N] and summed with orc_sum_tree. */
void orc_cylinder_init(const double *cyl, double *cvel, double *cdisp,
                       double *cacc, double *cforce);
void orc_cylinder_half1(double dt, double *cvel, double *cdisp, const double *cacc);
void orc_contact(int N, const double *dmg, const double *const *x,
                 double *const *u, double *const *v, const double *cyl,
                 const double *cdisp, double dt, double dens, double vol,
                 double *const *react);
void orc_cylinder_half2(double dt, double mass, double *cvel, double *cacc, double *cforce);

#ifdef __cplusplus
}
#endif
/* Threads of the OpenMP loops (the CPU baseline times 1 thread and all). */
void orc_set_threads(int n);
int orc_get_threads(void);

#endif

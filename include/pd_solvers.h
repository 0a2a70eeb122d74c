/*
 * pd_solvers.h -- the reference's solver entry points over the pd_gpu.h C-ABI.
 *
 * Each mirrors one solver of XiaoMmingW/PD_Gpu1 (scenario constants, stage
 * order, disp.txt output) and runs it on GPU `gpu`.  They return a pd_gpu.h
 * status instead of calling exit(1).  checkpoint_path, when not NULL, receives
 * the final state in the binary format of pd_checkpoint_write (no reference
 * counterpart: disp.txt keeps six significant digits).
 */
#ifndef PD_SOLVERS_H
#define PD_SOLVERS_H
#ifdef __cplusplus
extern "C" {
#endif

/* Dselect, Base.cuh:35-43: device (1 CPU, 2 GPU), Dim, B_S (1 bond, 2 state),
 * Mode (1 quasi-static, 2 dynamic).  Solvers overwrite Dim/B_S/Mode as the
 * reference does (e.g. Solve_Plate_Tensile.cu:34-36). */
typedef struct Dselect {
    int device;
    int Dim;
    int B_S;
    int Mode;
} Dselect;

/* solve_plate_tensile(Dselect&, real dx) -- Solve_Plate_Tensile.cuh:17.
 * dx <= 1e-8 selects the reference default 0.01 m. */
int solve_plate_tensile(Dselect *d, double dx, int gpu, const char *out_path, const char *checkpoint_path);
/* solve_plate_crack_grow(Dselect&, real dx) -- Solve_Plate_Crack_Grow.cuh:18. */
int solve_plate_crack_grow(Dselect *d, double dx, int gpu, const char *out_path, const char *checkpoint_path);
/* solve_kalthoff_winkler(Dselect&, real dx) -- Solve_Kalthoff_Winkler.cuh:18.
 * nt < 0 runs 1350 steps (the shipped nt = 0 runs none). */
int solve_kalthoff_winkler(Dselect *d, double dx, int gpu, const char *out_path, int nt,
                           const char *checkpoint_path);

#ifdef __cplusplus
}
#endif
#endif

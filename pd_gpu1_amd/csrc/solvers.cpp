// solvers.cpp -- the reference's three solver entry points, re-expressed over
// the C-ABI of include/pd_gpu.h, and its positional command line.
//
// solve_plate_tensile    Solve_Plate_Tensile.cu:30-138    2D PMB, ADR
// solve_plate_crack_grow Solve_Plate_Crack_Grow.cu:30-171 2D LPS, Verlet
// solve_kalthoff_winkler Solve_Kalthoff_Winkler.cu:268-441 3D PMB, Verlet + cylinder
// main                   main.cu:11-50  `pd_parallel [device] [mode] [dx]`
//
// Each solver keeps the reference's hard-coded scenario constants and stage
// order, runs it on one GPU and writes disp.txt like Save.cu.  Differences:
// dx is parsed as a float (main.cu:23 uses atoi); the GPU ordinal is an
// argument instead of cudaSetDevice(1) (main.cu:16); the Kalthoff-Winkler loop
// runs 1350 steps instead of the shipped nt = 0 (Solve_Kalthoff_Winkler.cu:280);
// the CPU path (device 1) is not part of this engine and is reported as an
// error; failures return a status instead of exit(1).  checkpoint_path
// (nullable) receives the final state in the binary pd_checkpoint_write
// format, for comparisons beyond disp.txt's six significant digits.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/pd_gpu.h"
#include "solvers.h"

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Run {
    pd_handle *h = nullptr;
    int rc = PD_OK;
    ~Run()
    {
        if (h) pd_destroy(h);
    }
    bool ok(int r)
    {
        if (r != PD_OK && rc == PD_OK) {
            rc = r;
            std::fprintf(stderr, "pd error %d: %s\n", r, h ? pd_last_error(h) : "");
        }
        return rc == PD_OK;
    }
};

int check_device(const Dselect &d)
{
    if (d.device != 2) {
        std::fprintf(stderr, "device %d: only the GPU path (device 2) is provided by this engine\n", d.device);
        return PD_E_INVALID;
    }
    return PD_OK;
}

} // namespace

extern "C" int solve_plate_tensile(Dselect *d, double dx, int gpu, const char *out_path, const char *checkpoint_path)
{
    // Solve_Plate_Tensile.cu:30-138
    int rc = check_device(*d);
    if (rc) return rc;
    d->Dim = 2;
    d->B_S = 1;
    d->Mode = 1;
    const double f_dx = dx > 1.0e-8 ? dx : 0.01;
    pd_params p;
    int nxyz[3], n = 0;
    rc = pd_params_init(&p, 2, PD_MODEL_PMB, 1.0, f_dx, 0.01, 1.0, 0.5, 0.0, 200.e9, 1. / 3., 7850, 1.0, 3.015, 0,
                        nxyz, &n);
    if (rc) return rc;
    const int nx = nxyz[0], ny = nxyz[1], nt = 1000;
    std::printf("nx %d ny %d\n", nx, ny);
    std::vector<double> x(n), y(n);
    pd_gen_plate(nx, ny, f_dx, x.data(), y.data());
    Run r;
    if (!r.ok(pd_create(&p, gpu, &r.h))) return r.rc;
    double t0 = now_s();
    r.ok(pd_set_positions(r.h, x.data(), y.data(), nullptr, n, n)) && r.ok(pd_build_neighbors(r.h)) &&
        r.ok(pd_surface_correct(r.h)) && r.ok(pd_init_state(r.h, PD_INIT_STATIC)) &&
        r.ok(pd_set_body_force_tensile(r.h, nx, ny));
    for (int i = 0; i < nt && r.rc == PD_OK; i++) r.ok(pd_adr_iterate(r.h, i, nullptr, nullptr));
    r.ok(pd_damage(r.h)) && r.ok(pd_sync(r.h));
    if (r.rc) return r.rc;
    std::printf("time %f\n", (now_s() - t0) * 1e6);
    if (checkpoint_path) r.ok(pd_checkpoint_write(r.h, checkpoint_path, nt));
    r.ok(pd_save_text(r.h, out_path ? out_path : "disp.txt", PD_LAYOUT_PLATE, nx, ny, 1, 0.0));
    return r.rc;
}

extern "C" int solve_plate_crack_grow(Dselect *d, double dx, int gpu, const char *out_path,
                                      const char *checkpoint_path)
{
    // Solve_Plate_Crack_Grow.cu:30-171
    int rc = check_device(*d);
    if (rc) return rc;
    d->Dim = 2;
    d->B_S = 2;
    d->Mode = 2;
    const double f_dx = dx > 1.0e-8 ? dx : 0.0001;
    pd_params p;
    int nxyz[3], n = 0;
    rc = pd_params_init(&p, 2, PD_MODEL_LPS, 1.3367e-8, f_dx, 0.0001, 0.05, 0.05, 0.0, 192.e9, 0.3, 7850, 0.04472,
                        3.015, 0, nxyz, &n);
    if (rc) return rc;
    const int nx = nxyz[0], ny = nxyz[1], nt = 1250;
    const int n_int = n;
    n += 6 * nx;                                          // :42
    p.n_top = p.n_bot = 3 * nx;
    p.strip_velocity = 20;                                // kernel_quick_load(20, ...) :128
    std::printf("nx %d ny %d\n", nx, ny);
    std::vector<double> x(n), y(n);
    pd_gen_plate_strips(nx, ny, p.dx, x.data(), y.data());
    Run r;
    if (!r.ok(pd_create(&p, gpu, &r.h))) return r.rc;
    double t0 = now_s();
    r.ok(pd_set_positions(r.h, x.data(), y.data(), nullptr, n_int, n)) && r.ok(pd_build_neighbors(r.h)) &&
        r.ok(pd_init_state(r.h, PD_INIT_DYNAMIC)) && r.ok(pd_set_crack_2d(r.h, 0.01, 0.0, 0.0, 0.0)) &&
        r.ok(pd_lps_weights(r.h));
    std::printf("N %d\n", n);
    r.ok(pd_run_verlet(r.h, nt)) && r.ok(pd_damage(r.h)) && r.ok(pd_sync(r.h));
    if (r.rc) return r.rc;
    std::printf("time %f\n", (now_s() - t0) * 1e6);
    if (checkpoint_path) r.ok(pd_checkpoint_write(r.h, checkpoint_path, nt));
    r.ok(pd_save_text(r.h, out_path ? out_path : "disp.txt", PD_LAYOUT_PLATE, nx, ny, 1, 0.0));
    return r.rc;
}

extern "C" int solve_kalthoff_winkler(Dselect *d, double dx, int gpu, const char *out_path, int nt,
                                      const char *checkpoint_path)
{
    // Solve_Kalthoff_Winkler.cu:268-441
    int rc = check_device(*d);
    if (rc) return rc;
    d->Dim = 3;
    d->B_S = 1;
    d->Mode = 2;
    const double f_dx = dx > 1.0e-8 ? dx : 0.001;
    pd_params p;
    int nxyz[3], n = 0;
    rc = pd_params_init(&p, 3, PD_MODEL_PMB, 8.7e-8, f_dx, 0.001, 0.2, 0.1, 0.009, 191.0e9, 0.25, 8000, 0.01, 3.015,
                        0, nxyz, &n);
    if (rc) return rc;
    const int nx = nxyz[0] + 1, ny = nxyz[1] + 1, nz = nxyz[2];  // :282-284
    pd_cyl cyl;
    cyl.rad = 0.025;
    cyl.mass = 1.57;
    cyl.v = -32.0;
    cyl.height = 0.05;
    cyl.x = 0.0;
    cyl.y = 0.1 / 2.0 + cyl.height / 2.0 + 0.1 * f_dx;    // :290
    cyl.z = 0.0;
    p.break_y_min = -0.04;                                // Force_Bond.cu:117
    const int N = pd_gen_kw(nx, ny, nz, f_dx, cyl.rad, nullptr, nullptr, nullptr);
    std::vector<double> x(N), y(N), z(N);
    pd_gen_kw(nx, ny, nz, f_dx, cyl.rad, x.data(), y.data(), z.data());
    std::printf("N %d\n", N);
    std::printf("nx %d ny %d nz %d\n", nx, ny, nz);
    if (nt < 0) nt = 1350;
    Run r;
    if (!r.ok(pd_create(&p, gpu, &r.h))) return r.rc;
    double t0 = now_s();
    r.ok(pd_set_positions(r.h, x.data(), y.data(), z.data(), N, N)) && r.ok(pd_build_neighbors(r.h)) &&
        r.ok(pd_surface_correct(r.h)) && r.ok(pd_init_state(r.h, PD_INIT_DYNAMIC)) &&
        r.ok(pd_set_notches_kw(r.h, cyl.rad)) && r.ok(pd_cylinder_init(r.h, &cyl));
    r.ok(pd_run_verlet(r.h, nt)) && r.ok(pd_damage(r.h)) && r.ok(pd_sync(r.h));
    if (r.rc) return r.rc;
    std::printf("time %f\n", (now_s() - t0) * 1e6);
    if (checkpoint_path) r.ok(pd_checkpoint_write(r.h, checkpoint_path, nt));
    r.ok(pd_save_text(r.h, out_path ? out_path : "disp.txt", PD_LAYOUT_KW, nx, ny, nz, cyl.rad));
    return r.rc;
}

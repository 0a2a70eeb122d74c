// pd_api.hip -- host side of the C-ABI declared in include/pd_gpu.h.
//
// Owns the handle, device memory, the stream and the launch order of the
// stages.  Parameter derivation and coordinate generation are host code that
// follows the reference's own host functions (Base.cuh:114-183,
// Coord.cu:19-129) so that the device sees bit-identical inputs.
#include "pd_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

namespace {

const double kPi = std::acos(-1.0);  // Base.cuh:31

inline double cube(double v) { return v * v * v; }  // Base.cuh:21

template <class T>
hipError_t dalloc(T **p, size_t count)
{
    if (*p) return hipSuccess;
    return hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int fail_words(const pd_handle *h) { return (int)((h->Nb + 31) / 32); }

int check_handle(pd_handle *h)
{
    return h ? PD_OK : PD_E_INVALID;
}

#define REQUIRE(cond, msg)              \
    do {                                \
        if (!(cond)) {                  \
            h->err = msg;               \
            return PD_E_INVALID;        \
        }                               \
    } while (0)

#define LAUNCH(expr) PD_TRY_HIP(h, (expr))

void free_lds(pd_handle *h)
{
    dfree(h->lds.tile_start);
    dfree(h->lds.stage_ptr);
    dfree(h->lds.stage_idx);
    dfree(h->lds.lidx);
    h->lds.ntiles = h->lds.ntiles_int = h->lds.max_staged = 0;
    h->lds_ok = false;
    h->lidx_fail_ok = false;
}

void free_bonds(pd_handle *h)
{
    dfree(h->col);
    dfree(h->fail);
    dfree(h->idist);
    dfree(h->fac);
    dfree(h->scr);
    dfree(h->kappa);
    dfree(h->w);
    dfree(h->e);
    dfree(h->bnd);
    free_lds(h);
    h->Nb = 0;
    h->coef_dirty = true;
    h->have_scr = false;
    h->have_weights = false;
}

void free_particles(pd_handle *h)
{
    free_lds(h);
    D3 *sets[] = {&h->x, &h->u, &h->v, &h->a, &h->b, &h->f, &h->vh, &h->vho, &h->fo};
    for (D3 *s : sets) {
        if (s->p[0]) (void)hipFree(s->p[0]);  // one allocation per set
        s->p[0] = s->p[1] = s->p[2] = nullptr;
    }
    if (h->f32.p[0]) (void)hipFree(h->f32.p[0]);
    h->f32.p[0] = h->f32.p[1] = h->f32.p[2] = nullptr;
    dfree(h->m);
    dfree(h->theta);
    dfree(h->dmg);
    dfree(h->minv32);
    dfree(h->theta32);
    dfree(h->rec);
    dfree(h->terms);
    h->rec_u_valid = false;
    dfree(h->order);
    dfree(h->row_ptr);
}

hipError_t alloc_set(D3 &s, int N, hipStream_t st)
{
    double *p = nullptr;
    hipError_t e = hipMalloc((void **)&p, sizeof(double) * 3 * (size_t)(N > 0 ? N : 1));
    if (e != hipSuccess) return e;
    s.p[0] = p;
    s.p[1] = p + N;
    s.p[2] = p + 2 * (size_t)N;
    return hipMemsetAsync(p, 0, sizeof(double) * 3 * (size_t)N, st);
}

// Staged neighbour lists of the LDS tile kernels for the current order
// (build_tile_order with want_tiles).  PD_E_CAPACITY when a tile's staged
// set does not fit the builder or one workgroup's LDS.
int build_lds_tiles(pd_handle *h)
{
    LdsTiles &T = h->lds;
    hipStream_t s = h->stream;
    if (T.ntiles <= 0 || !h->col) return PD_E_CAPACITY;
    std::vector<int32_t> cnt((size_t)T.ntiles);
    int32_t *dcnt = nullptr;
    LAUNCH(hipMalloc(&dcnt, sizeof(int32_t) * T.ntiles));
    hipError_t e = launch_stage_tiles(s, 0, T.ntiles, T.tile_start, h->order, h->row_ptr, h->col, dcnt, nullptr,
                                      nullptr, nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), dcnt, sizeof(int32_t) * T.ntiles, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dcnt);
    LAUNCH(e);
    int lds_bytes = 0;
    LAUNCH(hipDeviceGetAttribute(&lds_bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    std::vector<int64_t> ptr((size_t)T.ntiles + 1, 0);
    int mx = 0;
    for (int t = 0; t < T.ntiles; ++t) {
        if (cnt[t] < 0) return PD_E_CAPACITY;
        mx = std::max(mx, (int)cnt[t]);
        ptr[t + 1] = ptr[t] + cnt[t];
    }
    if ((int64_t)mx * lds_record_bytes() > lds_bytes || mx > stage_cap()) return PD_E_CAPACITY;
    LAUNCH(hipMalloc(&T.stage_ptr, sizeof(int64_t) * (T.ntiles + 1)));
    LAUNCH(hipMalloc(&T.stage_idx, sizeof(int32_t) * (ptr[T.ntiles] > 0 ? ptr[T.ntiles] : 1)));
    // an even count: breaks clear bit 15 with 32-bit atomics on word k/2
    LAUNCH(hipMalloc(&T.lidx, sizeof(uint16_t) * (size_t)((h->Nb + 2) & ~(int64_t)1)));
    LAUNCH(hipMemcpyAsync(T.stage_ptr, ptr.data(), sizeof(int64_t) * (T.ntiles + 1), hipMemcpyHostToDevice, s));
    LAUNCH(launch_stage_tiles(s, 1, T.ntiles, T.tile_start, h->order, h->row_ptr, h->col, nullptr, T.stage_ptr,
                              T.stage_idx, T.lidx));
    LAUNCH(hipStreamSynchronize(s));
    T.max_staged = mx > 0 ? mx : 1;  // invalid bond slots read record 0
    T.block = h->lds_block;
    T.block2 = h->lds_block2;
    h->lds_ok = true;
    h->lidx_fail_ok = false;
    return PD_OK;
}

// 1/q of the mixed records (1 when there are none).
double rec_iq(const pd_handle *h) { return h->rec_q > 0.0 ? 1.0 / h->rec_q : 1.0; }

// Mixed mode particle records (kRecW words, pd_internal.h).  The quantum q
// is the power of two 2^(e-29) with delta < 2^e, so a bond's coordinate
// difference stays below 2^29 q and its integer difference below 2^31 in
// magnitude; the integers rint(x0 / q) are exact in FP64 for |x0| < 2^52 q.
// lmin, the shortest reference bond, bounds the break band (mix_window).
int setup_records(pd_handle *h)
{
    const pd_params &P = h->P;
    hipStream_t s = h->stream;
    int e2 = 0;
    (void)std::frexp(P.delta, &e2);
    h->rec_q = std::ldexp(1.0, e2 - 29);
    REQUIRE(h->xabs < std::ldexp(h->rec_q, 52), "coordinates too large against the horizon for the mixed records");
    unsigned long long *dmin = nullptr, bits = ~0ull;
    LAUNCH(hipMalloc(&dmin, sizeof(bits)));
    hipError_t e = hipMemcpyAsync(dmin, &bits, sizeof(bits), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_min_bond(s, P.dim, h->N, h->row_ptr, h->col, cd3(h->x), dmin);
    if (e == hipSuccess) e = hipMemcpyAsync(&bits, dmin, sizeof(bits), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dmin);
    LAUNCH(e);
    double m2 = INFINITY;
    if (bits != ~0ull) std::memcpy(&m2, &bits, sizeof(m2));
    h->lmin = std::sqrt(m2);
    REQUIRE(h->Nb == 0 || h->lmin > 8.0 * h->rec_q, "coincident particles: a bond of zero length");
    LAUNCH(dalloc(&h->rec, (size_t)h->N * kRecW));
    LAUNCH(launch_pack_x0(s, P.dim, h->N, cd3(h->x), h->rec_q, h->rec));
    LAUNCH(launch_pack_u(s, nullptr, P.dim, h->N, cd3(h->u), h->rec_q, h->rec));
    h->rec_u_valid = true;
    return PD_OK;
}

int ensure_coef(pd_handle *h)
{
    if (!h->coef_dirty) return PD_OK;
    REQUIRE(h->col, "pd_build_neighbors first");
    hipStream_t s = h->stream;
    const pd_params &P = h->P;
    if (P.precision == PD_PREC_MIXED) {
        int rc = setup_records(h);
        if (rc) return rc;
        // Tiled processing order.  LDS-staged tiles: 5 dx cubes in 3D (about
        // 900 staged records, 70 KB of LDS: two 512-thread workgroups per
        // CU), 16 dx squares in 2D.  Gathering kernels: 16 dx tiles in 3D (a
        // tile plus its horizon halo is ~0.7 MB of records, inside one XCD's
        // 4 MB L2), 64 dx in 2D.
        free_lds(h);
        dfree(h->order);
        const bool want_lds = h->lds_mode != 0;
        int tc = h->tile_cells;
        if (tc < 0) tc = want_lds ? (P.dim == 3 ? 5 : 16) : (P.dim == 3 ? 16 : 64);
        if (tc > 0) {
            int rc = build_tile_order(h, tc * P.dx, want_lds);
            if (rc) return rc;
            if (want_lds) {
                rc = build_lds_tiles(h);
                if (rc == PD_E_CAPACITY && h->lds_mode == 1) {
                    h->err = "a tile's staged neighbour set does not fit the LDS of one workgroup";
                    return rc;
                }
                if (rc && rc != PD_E_CAPACITY) return rc;
                if (rc == PD_E_CAPACITY) {
                    // fall back to the gathering kernels and their tile size
                    free_lds(h);
                    dfree(h->order);
                    if (h->tile_cells < 0) tc = P.dim == 3 ? 16 : 64;
                    rc = build_tile_order(h, tc * P.dx, false);
                    if (rc) return rc;
                }
            }
        } else if (h->lds_mode == 1) {
            h->err = "LDS tiles need a tiled order (tile_cells > 0)";
            return PD_E_INVALID;
        }
    }
    if (P.model == PD_MODEL_PMB) {
        LAUNCH(dalloc(&h->kappa, h->Nb));
        LAUNCH(dalloc(&h->bnd, h->Nb));
        LAUNCH(launch_pmb_coef(s, h->Nb, h->fac, h->scr, P.vol, P.bc, h->idist, rec_iq(h), h->kappa, h->bnd));
    } else {
        if (!h->have_weights) {
            h->err = "LPS force needs pd_lps_weights first";
            return PD_E_INVALID;
        }
        if (P.precision == PD_PREC_MIXED) {
            LAUNCH(dalloc(&h->bnd, h->Nb));
            LAUNCH(dalloc(&h->minv32, h->N));
            LAUNCH(dalloc(&h->theta32, h->N));
            LAUNCH(launch_lps_coef(s, h->Nb, h->w, h->fac, P.vol, h->idist, rec_iq(h), h->bnd, h->N, h->m,
                                   h->minv32));
        } else {
            LAUNCH(dalloc(&h->e, h->Nb));
        }
    }
    h->geom_ok = false;
    if (P.precision == PD_PREC_MIXED && h->coef_mode != 0 && h->Nb > 0) {
        // in units of the record quantum q (the kernels measure lengths in q)
        const double q = h->rec_q;
        GeomCoef gc{};
        gc.hi_dx = (float)((P.delta + P.dx / 2.0) / P.dx);
        gc.inv_dx = (float)(q / P.dx);
        gc.scale = (float)(P.model == PD_MODEL_PMB ? P.vol * P.bc : P.vol);
        gc.k2 = P.model == PD_MODEL_LPS ? (float)(-1.4426950408889634 * q * q / (P.delta * P.delta)) : 0.f;
        int *dok = nullptr, ok = 0;
        LAUNCH(hipMalloc(&dok, sizeof(int)));
        const int one = 1;
        hipError_t e = hipMemcpyAsync(dok, &one, sizeof(int), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = launch_geom_check(s, P.dim, h->N, h->row_ptr, h->col, h->rec, h->bnd, gc, dok);
        if (e == hipSuccess) e = hipMemcpyAsync(&ok, dok, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(dok);
        LAUNCH(e);
        if (!ok && h->coef_mode == 1) {
            h->err = "stored bond coefficients differ from their recomputation from |xi|";
            return PD_E_INVALID;
        }
        h->geom_ok = ok != 0;
        h->gc = gc;
    }
    h->coef_dirty = false;
    return PD_OK;
}

// Launch shape of the mixed kernels: lanes per particle from the mean row
// length (the autotuner in pd_gpu1_amd/autotune.py can override it).
int pick_group(const pd_handle *h, int pass = 0)
{
    if (pass == 1 && h->group2) return h->group2;
    if (h->group) return h->group;
    double avg = h->N ? (double)h->Nb / h->N : 0.0;
    if (avg > 96) return 64;
    if (avg > 40) return 32;
    if (avg > 12) return 16;
    if (avg > 6) return 8;
    return 4;
}

// Lanes per row of the LDS kernels: the autotuned or set group, else 8
// (short per-row walks balance the rows of a tile across the workgroup).
int lds_group(const pd_handle *h, int pass = 0)
{
    if (pass == 1 && h->group2) return h->group2;
    if (h->group) return h->group;
    return 8;
}

const GeomCoef *gcp(const pd_handle *h) { return h->geom_ok ? &h->gc : nullptr; }

// Constants of the mixed LPS passes.  With t = w (A theta xi + B e) / m, the
// pair force (t_i + t_j) vol fac n of Force_State.cu:111-117 (2D: A =
// (2nu-1)/(nu-1)(2K-32G/9), B = 8G) and Force_State.cu:274-279 (3D: 3K theta
// w xi/m + 15G/m w (e - theta xi/3), i.e. A = 3K-5G, B = 15G).  The kernels
// measure lengths in units of the record quantum q, so A and B carry a
// factor q and the dilatation constant a factor q^2 (exact scalings).  The
// dilatation pass stores A theta/m and B/m in each particle record for the
// force pass, and with recomputed coefficients folds vol into the three.
void lps_mixed_consts(const pd_handle *h, float *ctheta, float *A, float *B)
{
    const pd_params &P = h->P;
    const double q = h->rec_q;
    if (P.dim == 2) {
        double c2 = (2.0 * P.pratio - 1.0) / (P.pratio - 1.0);
        *ctheta = (float)(2.0 * c2 * q * q);
        *A = (float)(c2 * (2.0 * P.K - 32.0 * P.G / 9.0) * q);
        *B = (float)(8.0 * P.G * q);
    } else {
        *ctheta = (float)(3.0 * q * q);
        *A = (float)((3.0 * P.K - 5.0 * P.G) * q);
        *B = (float)(15.0 * P.G * q);
    }
}

// Lanes per row of the FP64 PMB kernel: set explicitly, or 8 lanes for the
// short 2D rows (85 us per ADR iteration on a 315 x 315 plate, 99 us with 4)
// and 16 for 3D rows; the bond-order sum makes every choice give the same
// bits.
int fp64_group(const pd_handle *h)
{
    if (h->fp64_group > 0) return h->fp64_group;
    return h->P.dim == 3 ? 16 : 8;
}

// Mixed mode: bring the u half of the particle records up to date when some
// call changed u without refreshing them (the integrators and the contact
// kernel refresh it themselves).  Runs before a graph is captured or
// replayed, so captured sequences never contain the pack pass.
hipError_t sync_rec_u(pd_handle *h)
{
    if (h->P.precision != PD_PREC_MIXED || !h->rec || h->rec_u_valid) return hipSuccess;
    hipError_t e = launch_pack_u(h->stream, nullptr, h->P.dim, h->N, cd3(h->u), h->rec_q, h->rec);
    if (e == hipSuccess) h->rec_u_valid = true;
    return e;
}

// LDS tiles: copy the fail bits into bit 15 of the local index words when
// some call changed them other than through the LDS kernels' own breaks.
// Like sync_rec_u it runs before a graph is captured or replayed.
hipError_t sync_lidx_fail(pd_handle *h)
{
    if (!h->lds_ok || h->lidx_fail_ok) return hipSuccess;
    hipError_t e = launch_lidx_sync(h->stream, h->Nb, h->fail, h->lds.lidx);
    if (e == hipSuccess) h->lidx_fail_ok = true;
    return e;
}

// Every change of the fail words outside the LDS kernels goes through here.
void fail_changed(pd_handle *h) { h->lidx_fail_ok = false; }

RecU rec_or_null(const pd_handle *h)
{
    return RecU{h->P.precision == PD_PREC_MIXED ? h->rec : nullptr, rec_iq(h)};
}

// Band and FP64 state of the mixed kernels' break decision (mixed_bond.h).
MixCtx mix_ctx(const pd_handle *h)
{
    const pd_params &P = h->P;
    MixCtx mc{};
    mc.s0f = (float)P.s0;
    mix_window(P.s0, P.near_window, h->rec_q, h->lmin, &mc);
    mc.k_u = (float)((double)mc.k_u * h->rec_q);  // M in units of q (exact)
    mc.x = cd3(h->x);
    mc.u = cd3(h->u);
    return mc;
}

// Force evaluation (PMB, or LPS theta pass + force pass) on the handle's
// stream.  stop may be null.
int force_launch(pd_handle *h, const int *stop)
{
    int rc = ensure_coef(h);
    if (rc) return rc;
    const pd_params &P = h->P;
    hipStream_t s = h->stream;
    const double ymin = P.dim == 3 ? P.break_y_min : -INFINITY;
    if (P.model == PD_MODEL_PMB) {
        // all rows i < N (Force_Bond.cu:169)
        if (P.precision == PD_PREC_FP64) {
            fail_changed(h);
            LAUNCH(launch_pmb_fp64(s, stop, P.dim, h->N, h->row_ptr, h->col, h->kappa, h->idist, h->fail,
                                   cd3(h->x), cd3(h->u), h->f, P.s0, ymin, fp64_group(h)));
        } else {
            LAUNCH(sync_rec_u(h));
            LAUNCH(sync_lidx_fail(h));
            if (!h->lds_ok) fail_changed(h);
            if (h->lds_ok)
                LAUNCH(launch_pmb_lds(s, stop, h->lds, P.dim, lds_group(h), h->order, h->row_ptr, h->bnd, h->fail,
                                      h->rec, h->f32, P.s0, mix_ctx(h), ymin, gcp(h)));
            else
                LAUNCH(launch_pmb_mixed(s, stop, h->order, P.dim, pick_group(h), h->N, h->row_ptr, h->col, h->bnd,
                                        h->fail, h->rec, h->f32, P.s0, mix_ctx(h), ymin, gcp(h)));
        }
    } else {
        // theta for i < N (Force_State.cu:339), force for i < N_int (:359)
        if (P.precision == PD_PREC_FP64) {
            fail_changed(h);
            LAUNCH(launch_lps_theta_fp64(s, stop, P.dim, h->N, h->row_ptr, h->col, cd3(h->x), cd3(h->u), h->idist,
                                         h->fac, h->w, h->m, h->fail, P.vol, P.pratio, P.theta_skip_broken, h->e,
                                         h->theta));
            LAUNCH(launch_lps_force_fp64(s, stop, P.dim, h->N_int, h->row_ptr, h->col, cd3(h->x), cd3(h->u),
                                         h->idist, h->fac, h->w, h->m, h->theta, h->e, P.vol, P.pratio, P.K, P.G,
                                         P.s0, ymin, h->fail, h->f));
        } else {
            const int g = pick_group(h);
            float ctheta, A, B;
            lps_mixed_consts(h, &ctheta, &A, &B);
            LAUNCH(sync_rec_u(h));
            LAUNCH(sync_lidx_fail(h));
            if (!h->lds_ok) fail_changed(h);
            const MixCtx mc = mix_ctx(h);
            if (h->lds_ok) {
                LAUNCH(launch_lps_theta_lds(s, stop, h->lds, P.dim, lds_group(h), h->order, h->row_ptr, h->bnd,
                                            h->rec, h->minv32, ctheta, A, B, P.theta_skip_broken, h->theta32,
                                            gcp(h)));
                LAUNCH(launch_lps_force_lds(s, stop, h->lds, P.dim, lds_group(h, 1), h->order, h->row_ptr, h->bnd,
                                            h->fail, h->rec, P.s0, mc, ymin, h->f32, gcp(h)));
            } else {
                LAUNCH(launch_lps_theta_mixed(s, stop, h->order, P.dim, g, h->N, h->row_ptr, h->col, h->bnd,
                                              h->fail, h->rec, h->minv32, ctheta, A, B, P.theta_skip_broken,
                                              h->theta32, gcp(h)));
                LAUNCH(launch_lps_force_mixed(s, stop, h->order, P.dim, pick_group(h, 1), h->N_int, h->row_ptr,
                                              h->col, h->bnd, h->fail, h->rec, P.s0, mc, ymin, h->f32, gcp(h)));
            }
        }
    }
    return PD_OK;
}

bool use32(const pd_handle *h) { return h->P.precision == PD_PREC_MIXED; }

// The body force for the Verlet passes: null while it is zero everywhere.
CD3 body(const pd_handle *h) { return h->b_zero ? CD3{} : cd3(h->b); }

int verlet1_launch(pd_handle *h)
{
    const pd_params &P = h->P;
    LAUNCH(launch_verlet1(h->stream, P.dim, h->N_int, P.n_top, P.n_bot, P.dt, P.strip_velocity, h->v, cd3(h->a),
                          h->u, h->have_cyl ? h->S->cyl : nullptr, rec_or_null(h)));
    return PD_OK;
}

int verlet2_launch(pd_handle *h)
{
    const pd_params &P = h->P;
    LAUNCH(launch_verlet2(h->stream, P.dim, h->N_int, P.dt, P.dens, h->v, h->a, cd3(h->f), cf3(h->f32), use32(h),
                          body(h)));
    return PD_OK;
}

int contact_launch(pd_handle *h)
{
    const pd_params &P = h->P;
    LAUNCH(launch_contact(h->stream, h->red_blocks, h->red_threads, h->N_int, h->dmg, cd3(h->x), h->u, h->v,
                          h->cyl, P.dt, P.dens, P.vol, h->partials, h->S, rec_or_null(h), h->terms));
    return PD_OK;
}

// Force of one step, half 2 of it and half 1 of the next (no cylinder).
int fused_step_launch(pd_handle *h)
{
    int rc = force_launch(h, nullptr);
    if (rc) return rc;
    const pd_params &P = h->P;
    LAUNCH(launch_verlet21(h->stream, P.dim, h->N_int, P.n_top, P.n_bot, P.dt, P.dens, P.strip_velocity, h->v,
                           cd3(h->f), cf3(h->f32), use32(h), body(h), h->u, rec_or_null(h)));
    return PD_OK;
}

int step_launch(pd_handle *h)
{
    int rc = verlet1_launch(h);
    if (!rc) rc = force_launch(h, nullptr);
    if (!rc && h->have_cyl) rc = contact_launch(h);
    if (!rc) rc = verlet2_launch(h);
    return rc;
}

int adr_iter_launch(pd_handle *h, double tol)
{
    const pd_params &P = h->P;
    int rc = force_launch(h, &h->S->done);
    if (rc) return rc;
    LAUNCH(launch_adr_reduce(h->stream, h->red_blocks, h->red_threads, P.dim, h->N_int, h->N, P.mass, P.dt,
                             cd3(h->u), cd3(h->f), cf3(h->f32), use32(h), cd3(h->fo), cd3(h->vho), cd3(h->b),
                             h->partials, h->S, tol, h->terms));
    LAUNCH(launch_adr_update(h->stream, P.dim, h->N_int, P.mass, P.dt, cd3(h->f), cf3(h->f32), use32(h),
                             cd3(h->b), h->vh, h->v, h->vho, h->u, h->fo, h->S, rec_or_null(h)));
    return PD_OK;
}

int read_scalars(pd_handle *h, DevScalars *out)
{
    LAUNCH(hipMemcpyAsync(out, h->S, sizeof(DevScalars), hipMemcpyDeviceToHost, h->stream));
    LAUNCH(hipStreamSynchronize(h->stream));
    return PD_OK;
}

// Destroys the cached step graph.  Launches of it may still be queued on the
// handle's stream (pd_run_verlet's odd fused step drops it right after
// replaying it; an error return can leave replays queued), so the stream is
// drained first: the executable graph is never destroyed while in flight.
void drop_graph(pd_handle *h)
{
    if (h->step_graph) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->step_graph);
    }
    h->step_graph = nullptr;
    h->step_graph_kind = -1;
}

// Capture `reps` repetitions of a launch sequence into the handle's graph.
template <class F>
int capture(pd_handle *h, int kind, int reps, F &&fn)
{
    if (h->step_graph && h->step_graph_kind == kind) return PD_OK;
    drop_graph(h);
    hipGraph_t g = nullptr;
    LAUNCH(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = PD_OK;
    for (int r = 0; r < reps && rc == PD_OK; ++r) rc = fn();
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        h->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return PD_E_HIP;
    }
    e = hipGraphInstantiate(&h->step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
        h->step_graph = nullptr;
        return PD_E_HIP;
    }
    h->step_graph_kind = kind;
    return PD_OK;
}

uint64_t splitmix64(uint64_t *state)
{
    // SplitMix64 (Steele, Lea, Flood 2014)
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

} // namespace

extern "C" {

// ------------------------------------------------------------- parameters
int pd_params_init(pd_params *p, int dim, int model, double dt, double dx, double thick, double length,
                   double height, double width, double emod, double pratio, double dens, double s0,
                   double delta_factor, int ref_k_bug, int *nxyz, int *n)
{
    // BaseModel_Parameter::cal_parameter, Base.cuh:114-183
    if (!p || (dim != 2 && dim != 3) || (model != PD_MODEL_PMB && model != PD_MODEL_LPS) || !(dx > 0.0))
        return PD_E_INVALID;
    std::memset(p, 0, sizeof(*p));
    p->dim = dim;
    p->model = model;
    p->precision = PD_PREC_FP64;
    p->theta_skip_broken = 1;
    p->reduce_lanes = 4096;
    p->dx = dx;
    p->emod = emod;
    p->pratio = pratio;
    p->dens = dens;
    p->s0 = s0;
    p->dt = dt;
    p->delta = delta_factor * dx;                        // :116
    p->break_y_min = -INFINITY;
    p->strip_velocity = 0.0;
    p->near_window = 1.0e-4;
    int nx = (int)((length - dx / 2.0) / dx + 1);        // :119
    int ny = (int)((height - dx / 2.0) / dx + 1);
    int nz = dim == 3 ? (int)((width - dx / 2.0) / dx + 1) : 1;
    p->max_neighbors = dim == 2 ? 32 : 128;              // Base.cuh:28-29
    if (model == PD_MODEL_PMB) {
        if (dim == 2) {
            p->vol = dx * dx * thick;                                                   // :134
            p->bc = 6.0 * emod / (kPi * thick * cube(p->delta) * (1.0 - pratio));      // :135
        } else {
            p->vol = dx * dx * dx;                                                      // :142
            p->bc = 6.0 * emod / (kPi * p->delta * cube(p->delta) * (1.0 - 2.0 * pratio)); // :143
        }
        p->mass = 0.25 * dt * dt * (kPi * p->delta * p->delta * thick) * (p->bc) / dx * 5.0; // :150
    } else {
        p->G = 0.5 * emod / (1 + pratio);                                              // :154
        if (dim == 2) {
            double t = dx;                                                              // :166
            p->vol = dx * dx * t;
            p->K = 0.5 * emod / (1 - pratio);                                           // :168
        } else {
            p->K = ref_k_bug ? emod / 3.0 / (1 - 2.0 * p->G) : emod / (3.0 * (1.0 - 2.0 * pratio)); // :175
            p->vol = dx * dx * dx;
        }
    }
    if (nxyz) {
        nxyz[0] = nx;
        nxyz[1] = ny;
        nxyz[2] = nz;
    }
    if (n) *n = nx * ny * nz;
    return PD_OK;
}

// ------------------------------------------------------------ coordinates
int pd_gen_plate(int nx, int ny, double dx, double *x, double *y)
{
    // Coord.cu:19-31
    if (!x || !y || nx <= 0 || ny <= 0) return PD_E_INVALID;
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (-ny / 2.0 + 0.5 + i) * dx;
        }
    return PD_OK;
}

int pd_gen_plate_strips(int nx, int ny, double dx, double *x, double *y)
{
    // Coord.cu:47-66
    if (!x || !y || nx <= 0 || ny <= 0) return PD_E_INVALID;
    for (int i = 0; i < ny + 3; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (-ny / 2.0 + 0.5 + i) * dx;
        }
    for (int i = ny + 3; i < ny + 6; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (ny / 2.0 + 3 - 0.5 - i) * dx;
        }
    return PD_OK;
}

int pd_gen_kw(int nx, int ny, int nz, double dx, double rad, double *x, double *y, double *z)
{
    // Coord.cu:77-129 (returns the count, not count-1)
    int num = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double tx = (-(nx - 1) / 2.0 + i) * dx;
                double ty = (-(ny - 1) / 2.0 + j) * dx;
                double tz = (-nz / 2.0 + 0.5 + k) * dx;
                bool keep;
                if (tx > (-rad - 1.1 * dx) && tx < (-rad + 0.1 * dx))
                    keep = ty < -1.0e-10;
                else if (tx > (rad - 0.1 * dx) && tx < (rad + 1.1 * dx))
                    keep = ty < -1.0e-10;
                else
                    keep = true;
                if (keep) {
                    if (x) {
                        x[num] = tx;
                        y[num] = ty;
                        z[num] = tz;
                    }
                    num++;
                }
            }
    return num;
}

int pd_gen_block_jitter(int nx, int ny, int nz, double dx, double jitter, uint64_t seed, double *x, double *y,
                        double *z)
{
    if (!x || !y || !z) return PD_E_INVALID;
    uint64_t st = seed;
    int64_t n = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++, n++) {
                double b[3] = {(-nx / 2.0 + 0.5 + i) * dx, (-ny / 2.0 + 0.5 + j) * dx, (-nz / 2.0 + 0.5 + k) * dx};
                double o[3];
                for (int d = 0; d < 3; d++) {
                    double u53 = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
                    o[d] = b[d] + (2.0 * u53 - 1.0) * jitter * dx;
                }
                x[n] = o[0];
                y[n] = o[1];
                z[n] = o[2];
            }
    return PD_OK;
}

// ------------------------------------------------------------------ handle
int pd_create(const pd_params *p, int device, pd_handle **out)
{
    if (!p || !out) return PD_E_INVALID;
    *out = nullptr;
    if ((p->dim != 2 && p->dim != 3) || (p->model != PD_MODEL_PMB && p->model != PD_MODEL_LPS) ||
        (p->precision != PD_PREC_FP64 && p->precision != PD_PREC_MIXED) || p->max_neighbors <= 0 ||
        p->max_neighbors > 256 || !(p->delta > 0.0))
        return PD_E_INVALID;
    int T = p->reduce_lanes;
    if (T < 64 || (T & (T - 1)) != 0 || T > (1 << 20)) return PD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PD_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return PD_E_HIP;
    pd_handle *h = new pd_handle();
    h->P = *p;
    h->device = device;
    h->red_threads = T >= 256 ? 256 : T;
    h->red_blocks = T / h->red_threads;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipMalloc((void **)&h->S, sizeof(DevScalars)) != hipSuccess ||
        hipMalloc((void **)&h->partials, sizeof(double) * 8 * h->red_blocks) != hipSuccess ||
        fill_on(h->stream, h->S, 0, sizeof(DevScalars)) != hipSuccess) {
        pd_destroy(h);
        return PD_E_HIP;
    }
    *out = h;
    return PD_OK;
}

int pd_destroy(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_bonds(h);
    free_particles(h);
    for (int d = 0; d < 3; d++) dfree(h->react[d]);
    dfree(h->partials);
    dfree(h->S);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PD_OK;
}

const char *pd_last_error(pd_handle *h)
{
    return h ? h->err.c_str() : "null handle";
}

int pd_get_params(pd_handle *h, pd_params *out)
{
    if (!h || !out) return PD_E_INVALID;
    *out = h->P;
    return PD_OK;
}

int pd_set_params(pd_handle *h, const pd_params *p)
{
    if (!h || !p) return PD_E_INVALID;
    REQUIRE(p->dim == h->P.dim && p->model == h->P.model && p->reduce_lanes == h->P.reduce_lanes,
            "dim, model and reduce_lanes are fixed at pd_create");
    bool coef = p->vol != h->P.vol || p->bc != h->P.bc || p->precision != h->P.precision;
    h->P = *p;
    if (coef) h->coef_dirty = true;
    h->rec_u_valid = false;  // FP64 mode moves u without refreshing the records
    drop_graph(h);
    return PD_OK;
}

int pd_set_positions(pd_handle *h, const double *x, const double *y, const double *z, int n_int, int n)
{
    if (!h) return PD_E_INVALID;
    h->rec_u_valid = false;
    REQUIRE(x && y && (h->P.dim == 2 || z), "null coordinate array");
    REQUIRE(n > 0 && n_int > 0 && n_int <= n, "need 0 < n_int <= n");
    REQUIRE(n_int + h->P.n_top + h->P.n_bot <= n, "strip counts exceed n - n_int");
    LAUNCH(hipSetDevice(h->device));
    LAUNCH(hipStreamSynchronize(h->stream));
    drop_graph(h);
    free_bonds(h);
    free_particles(h);
    for (int d = 0; d < 3; d++) dfree(h->react[d]);
    h->N = n;
    h->N_int = n_int;
    h->xabs = 0.0;
    h->b_zero = true;  // every field starts at zero
    D3 *sets[] = {&h->x, &h->u, &h->v, &h->a, &h->b, &h->f, &h->vh, &h->vho, &h->fo};
    for (D3 *s : sets) LAUNCH(alloc_set(*s, n, h->stream));
    {
        float *p = nullptr;
        LAUNCH(hipMalloc((void **)&p, sizeof(float) * 3 * (size_t)n));
        LAUNCH(fill_on(h->stream, p, 0, sizeof(float) * 3 * (size_t)n));
        h->f32.p[0] = p;
        h->f32.p[1] = p + n;
        h->f32.p[2] = p + 2 * (size_t)n;
    }
    LAUNCH(dalloc(&h->dmg, n));
    LAUNCH(fill_on(h->stream, h->dmg, 0, sizeof(double) * n));
    LAUNCH(dalloc(&h->terms, 4 * (size_t)n));
    LAUNCH(dalloc(&h->m, n));
    LAUNCH(dalloc(&h->theta, n));
    LAUNCH(fill_on(h->stream, h->theta, 0, sizeof(double) * n));
    LAUNCH(dalloc(&h->row_ptr, (size_t)n + 1));
    const double *src[3] = {x, y, z};
    for (int d = 0; d < h->P.dim; d++) {
        LAUNCH(copy_on(h->stream, h->x.p[d], src[d], sizeof(double) * n, hipMemcpyHostToDevice));
        double lo = src[d][0], hi = src[d][0];
        for (int i = 1; i < n; i++) {
            lo = src[d][i] < lo ? src[d][i] : lo;
            hi = src[d][i] > hi ? src[d][i] : hi;
        }
        REQUIRE(std::isfinite(lo) && std::isfinite(hi), "non-finite coordinate");
        h->xabs = std::max(h->xabs, std::max(std::fabs(lo), std::fabs(hi)));
        h->lo[d] = lo;
        double cs = h->P.delta * (1.0 + 1e-9);
        h->ncell[d] = (int)std::floor((hi - lo) / cs) + 1;
    }
    for (int d = h->P.dim; d < 3; d++) {
        h->lo[d] = 0.0;
        h->ncell[d] = 1;
    }
    LAUNCH(hipStreamSynchronize(h->stream));
    return PD_OK;
}

// -------------------------------------------------------------- neighbours
int pd_build_neighbors(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->N > 0, "pd_set_positions first");
    LAUNCH(hipSetDevice(h->device));
    drop_graph(h);
    free_bonds(h);
    int rc = build_neighbors_impl(h);
    if (rc) return rc;
    const int64_t nb = h->Nb;
    LAUNCH(dalloc(&h->fail, (size_t)fail_words(h)));
    fail_changed(h);
    LAUNCH(launch_fill_u32(h->stream, h->fail, fail_words(h), 0xFFFFFFFFu));
    LAUNCH(dalloc(&h->idist, nb));
    LAUNCH(dalloc(&h->fac, nb));
    LAUNCH(dalloc(&h->scr, nb));
    LAUNCH(launch_fill_f64(h->stream, h->scr, nb, 1.0));
    // idist/fac are always needed by every later stage (Surf_Corr.cu:7-35)
    LAUNCH(launch_vol_corr(h->stream, h->P.dim, h->N, h->row_ptr, h->col, cd3(h->x), h->P.delta, h->P.dx,
                           h->idist, h->fac));
    h->coef_dirty = true;
    LAUNCH(hipStreamSynchronize(h->stream));
    return PD_OK;
}

int pd_num_bonds(pd_handle *h, int64_t *nb)
{
    if (!h || !nb) return PD_E_INVALID;
    *nb = h->Nb;
    return PD_OK;
}

int pd_get_csr(pd_handle *h, int32_t *row_ptr, int32_t *col)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    LAUNCH(hipStreamSynchronize(h->stream));
    if (row_ptr) LAUNCH(copy_on(h->stream, row_ptr, h->row_ptr, sizeof(int32_t) * (h->N + 1), hipMemcpyDeviceToHost));
    if (col) LAUNCH(copy_on(h->stream, col, h->col, sizeof(int32_t) * h->Nb, hipMemcpyDeviceToHost));
    return PD_OK;
}

// -------------------------------------------------------- setup physics
int pd_vol_corr(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    LAUNCH(launch_vol_corr(h->stream, h->P.dim, h->N, h->row_ptr, h->col, cd3(h->x), h->P.delta, h->P.dx, h->idist,
                           h->fac));
    h->coef_dirty = true;
    return PD_OK;
}

int pd_surface_correct(pd_handle *h)
{
    // surface_correct_cpu, Surf_Corr.cu:465-524 (sedload Surf_Corr.cu:478,501)
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    const pd_params &P = h->P;
    double sedload = P.dim == 2 ? 0.5 * P.emod / (1.0 - P.pratio * P.pratio) * 1.0e-6 : 0.6 * P.emod * 1.0e-6;
    double *fn = nullptr;
    LAUNCH(hipMalloc((void **)&fn, sizeof(double) * 3 * (size_t)h->N));
    hipError_t e = hipSuccess;
    for (int d = 0; d < P.dim && e == hipSuccess; d++)
        e = launch_surface_probe(h->stream, P.dim, h->N, h->row_ptr, h->col, cd3(h->x), d, h->idist, h->fac, P.bc,
                                 P.vol, sedload, fn + (size_t)d * h->N);
    if (e == hipSuccess) e = launch_surface_scr(h->stream, P.dim, h->N, h->row_ptr, h->col, cd3(h->x), h->idist, fn, h->scr);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(fn);
    LAUNCH(e);
    h->have_scr = true;
    h->coef_dirty = true;
    return PD_OK;
}

int pd_init_state(pd_handle *h, int mode)
{
    // base_integrate_initial_* / static_initial_bond_* (Initial.cu:6-132):
    // every bond intact, u, v, a, b, F zero; static mode also zeroes vhold
    // and Fold.  dmg = 0 (the reference leaves it unset for KW; see
    // Solve_Kalthoff_Winkler.cu:122,187).
    if (!h) return PD_E_INVALID;
    h->rec_u_valid = false;
    REQUIRE(h->col, "pd_build_neighbors first");
    hipStream_t s = h->stream;
    fail_changed(h);
    LAUNCH(launch_fill_u32(s, h->fail, fail_words(h), 0xFFFFFFFFu));
    D3 *sets[] = {&h->u, &h->v, &h->a, &h->b, &h->f, &h->vh, &h->vho, &h->fo};
    for (D3 *st : sets) LAUNCH(launch_fill_f64(s, st->p[0], 3 * (int64_t)h->N, 0.0));
    h->b_zero = true;
    LAUNCH(launch_fill_f64(s, h->dmg, h->N, 0.0));
    LAUNCH(hipMemsetAsync(h->f32.p[0], 0, sizeof(float) * 3 * (size_t)h->N, s));
    (void)mode;
    LAUNCH(hipStreamSynchronize(s));
    return PD_OK;
}

int pd_set_crack_2d(pd_handle *h, double clength, double loc_x, double loc_y, double theta)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col && h->P.dim == 2, "2D problem with bonds required");
    // host scalars of Base_Function.cu:95-110
    const double t = std::tan(theta);
    const double sq1 = std::sqrt(1 + std::tan(theta) * std::tan(theta));
    const double half_s = clength * std::sin(theta) / 2.0;
    const int vertical = std::fabs(theta - kPi / 2.0) < 0.0001;
    fail_changed(h);
    LAUNCH(launch_crack_2d(h->stream, h->N_int, h->row_ptr, h->col, cd3(h->x), clength, loc_x, loc_y, vertical, t,
                           sq1, half_s, h->fail));
    return PD_OK;
}

int pd_set_notches_kw(pd_handle *h, double rad)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    fail_changed(h);
    LAUNCH(launch_crack_kw(h->stream, h->N_int, h->row_ptr, h->col, cd3(h->x), rad, h->fail));
    return PD_OK;
}

// Influence weights w = exp(-xi^2/delta^2) (Force_State.cu:317) on the host:
// the C library's exp is the one the FP64 CPU algorithm calls, so w, and with
// it every FP64 LPS field, is bit-identical to that algorithm's.  The bond
// lengths go down and the weights up in chunks, each chunk split over host
// threads.
static hipError_t host_lps_weights(hipStream_t s, int64_t nb, const double *idist, double delta, double *w)
{
    const int64_t chunk = int64_t(1) << 25;
    std::vector<double> buf((size_t)std::min<int64_t>(nb, chunk));
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
    const double d2 = delta * delta;
    for (int64_t k0 = 0; k0 < nb; k0 += chunk) {
        const int64_t n = std::min(chunk, nb - k0);
        hipError_t e = hipMemcpyAsync(buf.data(), idist + k0, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        const unsigned use = n < (int64_t)1 << 16 ? 1 : nt;
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < use; ++t)
            pool.emplace_back([&, t] {
                const int64_t a = n * t / use, b = n * (t + 1) / use;
                for (int64_t k = a; k < b; ++k) buf[k] = std::exp(-(buf[k] * buf[k]) / d2);
            });
        for (auto &th : pool) th.join();
        e = hipMemcpyAsync(w + k0, buf.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int pd_lps_weights(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col && h->P.model == PD_MODEL_LPS, "LPS problem with bonds required");
    LAUNCH(dalloc(&h->w, h->Nb));
    LAUNCH(host_lps_weights(h->stream, h->Nb, h->idist, h->P.delta, h->w));
    LAUNCH(launch_lps_weights(h->stream, h->N, h->row_ptr, h->idist, h->fac, h->P.delta, h->P.vol, 1, h->w,
                              h->m));
    h->have_weights = true;
    h->coef_dirty = true;
    return PD_OK;
}

// ------------------------------------------------------------------ force
int pd_force(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    LAUNCH(hipEventRecord(h->ev0, h->stream));
    int rc = force_launch(h, nullptr);
    if (rc) return rc;
    LAUNCH(hipEventRecord(h->ev1, h->stream));
    return PD_OK;
}

int pd_time_force_passes(pd_handle *h, int reps, double *ms_out)
{
    // Average device time of each force pass over reps launches, timed with
    // HIP events on the handle's stream: ms_out[0] = PMB kernel or LPS
    // dilatation pass, ms_out[1] = LPS force pass (0 for PMB).
    if (!h || !ms_out || reps <= 0) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    int rc = ensure_coef(h);
    if (rc) return rc;
    HipEvents<3> events;
    LAUNCH(events.create());
    hipEvent_t *e = events.e;
    const pd_params &P = h->P;
    const double ymin = P.dim == 3 ? P.break_y_min : -INFINITY;
    double acc[2] = {0.0, 0.0};
    LAUNCH(sync_lidx_fail(h));
    if (!h->lds_ok || P.precision == PD_PREC_FP64) fail_changed(h);
    const MixCtx mc = P.precision == PD_PREC_MIXED ? mix_ctx(h) : MixCtx{};
    for (int r = 0; r < reps && rc == PD_OK; ++r) {
        if (P.precision == PD_PREC_MIXED)
            LAUNCH(launch_pack_u(h->stream, nullptr, P.dim, h->N, cd3(h->u), h->rec_q, h->rec));
        LAUNCH(hipEventRecord(e[0], h->stream));
        if (P.model == PD_MODEL_PMB && P.precision == PD_PREC_MIXED) {
            if (h->lds_ok)
                LAUNCH(launch_pmb_lds(h->stream, nullptr, h->lds, P.dim, lds_group(h), h->order, h->row_ptr, h->bnd,
                                      h->fail, h->rec, h->f32, P.s0, mc, ymin, gcp(h)));
            else
                LAUNCH(launch_pmb_mixed(h->stream, nullptr, h->order, P.dim, pick_group(h), h->N, h->row_ptr, h->col,
                                        h->bnd, h->fail, h->rec, h->f32, P.s0, mc, ymin, gcp(h)));
            LAUNCH(hipEventRecord(e[1], h->stream));
            LAUNCH(hipEventRecord(e[2], h->stream));
        } else if (P.model == PD_MODEL_PMB) {
            rc = force_launch(h, nullptr);
            LAUNCH(hipEventRecord(e[1], h->stream));
            LAUNCH(hipEventRecord(e[2], h->stream));
        } else if (P.precision == PD_PREC_FP64) {
            LAUNCH(launch_lps_theta_fp64(h->stream, nullptr, P.dim, h->N, h->row_ptr, h->col, cd3(h->x), cd3(h->u),
                                         h->idist, h->fac, h->w, h->m, h->fail, P.vol, P.pratio,
                                         P.theta_skip_broken, h->e, h->theta));
            LAUNCH(hipEventRecord(e[1], h->stream));
            LAUNCH(launch_lps_force_fp64(h->stream, nullptr, P.dim, h->N_int, h->row_ptr, h->col, cd3(h->x),
                                         cd3(h->u), h->idist, h->fac, h->w, h->m, h->theta, h->e, P.vol, P.pratio,
                                         P.K, P.G, P.s0, ymin, h->fail, h->f));
            LAUNCH(hipEventRecord(e[2], h->stream));
        } else {
            const int g = pick_group(h);
            float ctheta, A, B;
            lps_mixed_consts(h, &ctheta, &A, &B);
            if (h->lds_ok)
                LAUNCH(launch_lps_theta_lds(h->stream, nullptr, h->lds, P.dim, lds_group(h), h->order, h->row_ptr,
                                            h->bnd, h->rec, h->minv32, ctheta, A, B, P.theta_skip_broken,
                                            h->theta32, gcp(h)));
            else
                LAUNCH(launch_lps_theta_mixed(h->stream, nullptr, h->order, P.dim, g, h->N, h->row_ptr, h->col,
                                              h->bnd, h->fail, h->rec, h->minv32, ctheta, A, B,
                                              P.theta_skip_broken, h->theta32, gcp(h)));
            LAUNCH(hipEventRecord(e[1], h->stream));
            if (h->lds_ok)
                LAUNCH(launch_lps_force_lds(h->stream, nullptr, h->lds, P.dim, lds_group(h, 1), h->order, h->row_ptr,
                                            h->bnd, h->fail, h->rec, P.s0, mc, ymin, h->f32, gcp(h)));
            else
                LAUNCH(launch_lps_force_mixed(h->stream, nullptr, h->order, P.dim, pick_group(h, 1), h->N_int,
                                              h->row_ptr, h->col, h->bnd, h->fail, h->rec, P.s0, mc, ymin, h->f32,
                                              gcp(h)));
            LAUNCH(hipEventRecord(e[2], h->stream));
        }
        LAUNCH(hipEventSynchronize(e[2]));
        float t0 = 0.f, t1 = 0.f;
        LAUNCH(hipEventElapsedTime(&t0, e[0], e[1]));
        LAUNCH(hipEventElapsedTime(&t1, e[1], e[2]));
        acc[0] += t0;
        acc[1] += t1;
    }
    ms_out[0] = acc[0] / reps;
    ms_out[1] = acc[1] / reps;
    return rc;
}

int pd_set_group(pd_handle *h, int group)
{
    // Lanes per particle of the mixed kernels (0 = heuristic); the autotuner's knob.
    if (!h) return PD_E_INVALID;
    REQUIRE(group == 0 || group == 4 || group == 8 || group == 16 || group == 32 || group == 64,
            "group must be 0, 4, 8, 16, 32 or 64");
    h->group = group;
    h->group2 = group;
    drop_graph(h);
    return PD_OK;
}

int pd_stream_copy(int device, int64_t bytes, int reps, double *gbps)
{
    // Device-to-device copy of `bytes` with a plain 16-byte-per-lane kernel:
    // the reference point for HBM counters and achievable bandwidth.
    if (bytes <= 0 || (bytes % 16) != 0 || reps <= 0 || !gbps) return PD_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return PD_E_HIP;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return PD_E_HIP;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return PD_E_HIP;
    }
    HipEvents<2> ev;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ev.create() == hipSuccess ? PD_OK : PD_E_HIP;
    e0 = ev.e[0];
    e1 = ev.e[1];
    if (!rc && hipMemset(a, 1, bytes) != hipSuccess) rc = PD_E_HIP;
    if (!rc && launch_copy16(0, a, b, bytes) != hipSuccess) rc = PD_E_HIP;  // warm
    if (!rc) {
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps && !rc; r++)
            if (launch_copy16(0, a, b, bytes) != hipSuccess) rc = PD_E_HIP;
        (void)hipEventRecord(e1, 0);
        float ms = 0.f;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = PD_E_HIP;
        if (!rc) *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int pd_set_pass_group(pd_handle *h, int pass, int group)
{
    // Lanes per particle of one pass of the mixed kernels (pass 0: PMB or
    // LPS dilatation, pass 1: LPS force); pd_set_group sets both.
    if (!h) return PD_E_INVALID;
    REQUIRE(pass == 0 || pass == 1, "pass must be 0 or 1");
    REQUIRE(group == 0 || group == 4 || group == 8 || group == 16 || group == 32 || group == 64,
            "group must be 0, 4, 8, 16, 32 or 64");
    if (pass == 0)
        h->group = group;
    else
        h->group2 = group;
    drop_graph(h);
    return PD_OK;
}

int pd_set_tile(pd_handle *h, int tile_cells)
{
    // Edge (in units of dx) of the tiles that order the throughput kernels'
    // work; 0 processes particles in index order, -1 restores the default.
    if (!h) return PD_E_INVALID;
    REQUIRE(tile_cells >= -1 && tile_cells <= 4096, "tile_cells out of range");
    h->tile_cells = tile_cells;
    h->coef_dirty = true;
    drop_graph(h);
    return PD_OK;
}

int pd_set_lds(pd_handle *h, int mode)
{
    // LDS-staged tiles for the mixed kernels: -1 where they fit (default),
    // 0 never (gathering kernels), 1 always (error when a tile does not fit).
    if (!h) return PD_E_INVALID;
    REQUIRE(mode >= -1 && mode <= 1, "mode must be -1, 0 or 1");
    h->lds_mode = mode;
    h->coef_dirty = true;
    drop_graph(h);
    return PD_OK;
}

int pd_set_fp64_group(pd_handle *h, int group)
{
    // Lanes per row of the FP64 PMB kernel: 1 (thread per row), 4, 8, 16,
    // 32, or -1 for the choice by dimension.  Results are identical.
    if (!h) return PD_E_INVALID;
    REQUIRE(group == -1 || group == 1 || group == 4 || group == 8 || group == 16 || group == 32,
            "group must be -1, 1, 4, 8, 16 or 32");
    h->fp64_group = group;
    drop_graph(h);
    return PD_OK;
}

int pd_set_coef_mode(pd_handle *h, int mode)
{
    // Bond coefficients of the mixed kernels: -1 recompute from |xi| when the
    // stored ones match that (default), 0 always stream them, 1 recompute
    // (PD_E_INVALID when they do not match).
    if (!h) return PD_E_INVALID;
    REQUIRE(mode >= -1 && mode <= 1, "mode must be -1, 0 or 1");
    h->coef_mode = mode;
    h->coef_dirty = true;
    drop_graph(h);
    return PD_OK;
}

int pd_coef_info(pd_handle *h, int *recomputed)
{
    if (!h || !recomputed) return PD_E_INVALID;
    int rc = ensure_coef(h);
    if (rc) return rc;
    *recomputed = h->geom_ok ? 1 : 0;
    return PD_OK;
}

int pd_set_lds_pass_block(pd_handle *h, int pass, int threads)
{
    // Workgroup size of one pass of the LDS tile kernels (one workgroup per
    // tile): pass 0 = PMB or LPS dilatation, 1 = LPS force.
    if (!h) return PD_E_INVALID;
    REQUIRE(pass == 0 || pass == 1, "pass must be 0 or 1");
    REQUIRE(threads == 256 || threads == 512 || threads == 1024, "threads must be 256, 512 or 1024");
    if (pass == 0)
        h->lds_block = h->lds.block = threads;
    else
        h->lds_block2 = h->lds.block2 = threads;
    drop_graph(h);
    return PD_OK;
}

int pd_set_lds_block(pd_handle *h, int threads)
{
    int rc = pd_set_lds_pass_block(h, 0, threads);
    return rc ? rc : pd_set_lds_pass_block(h, 1, threads);
}

int pd_lds_info(pd_handle *h, int *active, int *ntiles, int *max_staged)
{
    if (!h) return PD_E_INVALID;
    int rc = ensure_coef(h);
    if (rc) return rc;
    if (active) *active = h->lds_ok ? 1 : 0;
    if (ntiles) *ntiles = h->lds.ntiles;
    if (max_staged) *max_staged = h->lds.max_staged;
    return PD_OK;
}

int pd_last_force_ms(pd_handle *h, double *ms)
{
    if (!h || !ms) return PD_E_INVALID;
    LAUNCH(hipEventSynchronize(h->ev1));
    float t = 0.f;
    LAUNCH(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return PD_OK;
}

int pd_sync(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    LAUNCH(hipStreamSynchronize(h->stream));
    return PD_OK;
}

// -------------------------------------------------------------- explicit
int pd_verlet_half1(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->N > 0, "pd_set_positions first");
    return verlet1_launch(h);
}

int pd_verlet_half2(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->N > 0, "pd_set_positions first");
    return verlet2_launch(h);
}

int pd_run_verlet(pd_handle *h, int nsteps)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col && nsteps >= 0, "bonds and nsteps >= 0 required");
    if (nsteps == 0) return PD_OK;
    int rc = ensure_coef(h);
    if (rc) return rc;
    LAUNCH(sync_rec_u(h));
    LAUNCH(sync_lidx_fail(h));
    // Long runs replay graphs of kGraphSteps steps (one launch per graph
    // instead of per step); the remainder is launched directly.
    constexpr int kGraphSteps = 8;
    if (h->have_cyl || nsteps == 1) {
        const int kind = 1 + (h->have_cyl ? 1 : 0);
        if (nsteps >= kGraphSteps) {
            rc = capture(h, 20 + kind, kGraphSteps, [&]() { return step_launch(h); });
            if (rc) return rc;
            for (int s = 0; s < nsteps / kGraphSteps; ++s) LAUNCH(hipGraphLaunch(h->step_graph, h->stream));
            for (int s = 0; s < nsteps % kGraphSteps && !rc; ++s) rc = step_launch(h);
            if (rc) return rc;
        } else {
            rc = capture(h, kind, 1, [&]() { return step_launch(h); });
            if (rc) return rc;
            for (int s = 0; s < nsteps; ++s) LAUNCH(hipGraphLaunch(h->step_graph, h->stream));
        }
    } else {
        // half 1, then (force, half 2 + half 1 of the next step) from a
        // graph, and the last step's force and half 2
        rc = verlet1_launch(h);
        if (rc) return rc;
        const int mid = nsteps - 1;
        if (mid >= kGraphSteps) {
            rc = capture(h, 24, kGraphSteps, [&]() { return fused_step_launch(h); });
            if (rc) return rc;
            for (int s = 0; s < mid / kGraphSteps; ++s) LAUNCH(hipGraphLaunch(h->step_graph, h->stream));
            for (int s = 0; s < mid % kGraphSteps && !rc; ++s) rc = fused_step_launch(h);
            if (rc) return rc;
        } else {
            rc = capture(h, 4, 1, [&]() { return fused_step_launch(h); });
            if (rc) return rc;
            for (int s = 0; s < mid; ++s) LAUNCH(hipGraphLaunch(h->step_graph, h->stream));
        }
        rc = force_launch(h, nullptr);
        if (!rc) rc = verlet2_launch(h);
        if (rc) return rc;
    }
    LAUNCH(hipStreamSynchronize(h->stream));
    return PD_OK;
}

// -------------------------------------------------------------------- ADR
int pd_set_body_force_tensile(pd_handle *h, int nx, int ny)
{
    // load_plate_tensile_cpu, Solve_Plate_Tensile.cu:19-26
    if (!h) return PD_E_INVALID;
    REQUIRE(nx > 0 && ny > 0 && (int64_t)nx * ny <= h->N, "nx*ny must fit in N");
    std::vector<double> bx((size_t)h->N, 0.0);
    LAUNCH(copy_on(h->stream, bx.data(), h->b.p[0], sizeof(double) * h->N, hipMemcpyDeviceToHost));
    const double E = h->P.emod, dx = h->P.dx;
    for (int i = 0; i < ny; i++) {
        bx[(size_t)i * nx] = -0.001 * E / dx;
        bx[(size_t)i * nx + nx - 1] = 0.001 * E / dx;
    }
    LAUNCH(copy_on(h->stream, h->b.p[0], bx.data(), sizeof(double) * h->N, hipMemcpyHostToDevice));
    if (h->b_zero) {
        h->b_zero = false;
        drop_graph(h);  // captured Verlet steps skip a zero body force
    }
    return PD_OK;
}

int pd_adr_iterate(pd_handle *h, int ct, double *c_out, double *resid_out)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    LAUNCH(launch_set_ct(h->stream, h->S, ct, 1));
    int rc = adr_iter_launch(h, -1.0);
    if (rc) return rc;
    if (c_out || resid_out) {
        DevScalars S;
        rc = read_scalars(h, &S);
        if (rc) return rc;
        if (c_out) *c_out = S.c;
        if (resid_out) *resid_out = S.resid;
        if (S.nonfinite) {
            h->err = "non-finite value in the ADR reduction";
            return PD_E_NONFINITE;
        }
    }
    return PD_OK;
}

int pd_adr_solve(pd_handle *h, int ct0, int max_iters, double tol, int *iters_out, double *resid_out)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col && max_iters >= 0 && tol >= 0.0, "bonds, max_iters >= 0 and tol >= 0 required");
    int rc = ensure_coef(h);
    if (rc) return rc;
    LAUNCH(sync_rec_u(h));
    LAUNCH(sync_lidx_fail(h));
    LAUNCH(launch_set_ct(h->stream, h->S, ct0, 1));
    const int chunk = 32;
    int launched = 0;
    DevScalars S{};
    // tol is a kernel argument of the captured chunk: recapture when it changes.
    if (h->step_graph_kind == 10 && h->graph_tol != tol) drop_graph(h);
    h->graph_tol = tol;
    while (launched < max_iters) {
        if (max_iters - launched >= chunk) {
            rc = capture(h, 10, chunk, [&]() { return adr_iter_launch(h, tol); });
            if (rc) return rc;
            LAUNCH(hipGraphLaunch(h->step_graph, h->stream));
            launched += chunk;
        } else {
            rc = adr_iter_launch(h, tol);
            if (rc) return rc;
            launched += 1;
        }
        rc = read_scalars(h, &S);
        if (rc) return rc;
        if (S.done || S.nonfinite) break;
    }
    if (launched == 0) rc = read_scalars(h, &S);
    // force evaluations = ct_counter - ct0; the converged one applies no update
    int evals = S.ct_counter - ct0;
    if (iters_out) *iters_out = S.done ? evals - 1 : evals;
    if (resid_out) *resid_out = S.resid;
    if (S.nonfinite) {
        h->err = "non-finite value in the ADR reduction";
        return PD_E_NONFINITE;
    }
    return S.done ? PD_OK : PD_E_NOT_CONVERGED;
}

int pd_strip_displace(pd_handle *h, double du_top, double du_bot)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->N > 0, "pd_set_positions first");
    h->rec_u_valid = false;
    LAUNCH(launch_strip_displace(h->stream, h->N_int, h->P.n_top, h->P.n_bot, du_top, du_bot, h->u.p[1]));
    return PD_OK;
}

int pd_reduce_sum(pd_handle *h, int field, int comp, double *out)
{
    if (!h || !out) return PD_E_INVALID;
    REQUIRE(h->N > 0 && comp >= 0 && comp < h->P.dim, "bad component");
    D3 *map[] = {&h->x, &h->u, &h->v, &h->a, &h->f, &h->b, &h->vh, &h->vho, &h->fo};
    REQUIRE(field >= 0 && field <= PD_FIELD_FO && !(field == PD_FIELD_F && use32(h)), "unsupported field");
    LAUNCH(launch_field_sum(h->stream, h->red_blocks, h->red_threads, map[field]->p[comp], h->N, h->partials, h->S));
    DevScalars S;
    int rc = read_scalars(h, &S);
    if (rc) return rc;
    *out = S.sums[7];
    return PD_OK;
}

// --------------------------------------------------------- damage / fields
int pd_damage(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->col, "pd_build_neighbors first");
    LAUNCH(launch_damage(h->stream, h->N_int, h->row_ptr, h->fail, h->fac, h->P.vol, h->dmg));
    return PD_OK;
}

static int field_ptr(pd_handle *h, int field, double **ptr, size_t *count, bool *vec)
{
    *vec = false;
    switch (field) {
    case PD_FIELD_X0: *ptr = h->x.p[0]; *vec = true; break;
    case PD_FIELD_U: *ptr = h->u.p[0]; *vec = true; break;
    case PD_FIELD_V: *ptr = h->v.p[0]; *vec = true; break;
    case PD_FIELD_A: *ptr = h->a.p[0]; *vec = true; break;
    case PD_FIELD_F: *ptr = h->f.p[0]; *vec = true; break;
    case PD_FIELD_B: *ptr = h->b.p[0]; *vec = true; break;
    case PD_FIELD_VH: *ptr = h->vh.p[0]; *vec = true; break;
    case PD_FIELD_VHO: *ptr = h->vho.p[0]; *vec = true; break;
    case PD_FIELD_FO: *ptr = h->fo.p[0]; *vec = true; break;
    case PD_FIELD_DMG: *ptr = h->dmg; *count = h->N; return PD_OK;
    case PD_FIELD_THETA: *ptr = h->theta; *count = h->N; return PD_OK;
    case PD_FIELD_M: *ptr = h->m; *count = h->N; return PD_OK;
    case PD_FIELD_IDIST: *ptr = h->idist; *count = h->Nb; return PD_OK;
    case PD_FIELD_FAC: *ptr = h->fac; *count = h->Nb; return PD_OK;
    case PD_FIELD_SCR: *ptr = h->scr; *count = h->Nb; return PD_OK;
    case PD_FIELD_W: *ptr = h->w; *count = h->Nb; return PD_OK;
    case PD_FIELD_E: *ptr = h->e; *count = h->Nb; return PD_OK;
    default: return PD_E_INVALID;
    }
    *count = (size_t)h->N;
    return PD_OK;
}

int pd_get_field(pd_handle *h, int field, double *out)
{
    if (!h || !out) return PD_E_INVALID;
    double *p = nullptr;
    size_t n = 0;
    bool vec = false;
    REQUIRE(field_ptr(h, field, &p, &n, &vec) == PD_OK, "unknown field");
    REQUIRE(p != nullptr, "field not allocated");
    LAUNCH(hipStreamSynchronize(h->stream));
    const int D = h->P.dim;
    if (vec && field == PD_FIELD_F && use32(h)) {
        std::vector<float> tmp((size_t)D * h->N);
        for (int d = 0; d < D; d++)
            LAUNCH(copy_on(h->stream, tmp.data() + (size_t)d * h->N, h->f32.p[d], sizeof(float) * h->N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < tmp.size(); i++) out[i] = tmp[i];
        return PD_OK;
    }
    if (vec) {
        for (int d = 0; d < D; d++)
            LAUNCH(copy_on(h->stream, out + (size_t)d * h->N, p + (size_t)d * h->N, sizeof(double) * h->N,
                             hipMemcpyDeviceToHost));
        return PD_OK;
    }
    if (field == PD_FIELD_THETA && use32(h) && h->theta32) {
        std::vector<float> tmp((size_t)h->N);
        LAUNCH(copy_on(h->stream, tmp.data(), h->theta32, sizeof(float) * h->N, hipMemcpyDeviceToHost));
        for (int i = 0; i < h->N; i++) out[i] = tmp[i];
        return PD_OK;
    }
    LAUNCH(copy_on(h->stream, out, p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_set_field(pd_handle *h, int field, const double *in)
{
    if (!h || !in) return PD_E_INVALID;
    double *p = nullptr;
    size_t n = 0;
    bool vec = false;
    REQUIRE(field_ptr(h, field, &p, &n, &vec) == PD_OK, "unknown field");
    REQUIRE(p != nullptr, "field not allocated");
    REQUIRE(field != PD_FIELD_X0, "reference coordinates are set by pd_set_positions");
    if (field == PD_FIELD_U) h->rec_u_valid = false;
    if (field == PD_FIELD_B && h->b_zero) {
        h->b_zero = false;
        drop_graph(h);  // captured Verlet steps skip a zero body force
    }
    LAUNCH(hipStreamSynchronize(h->stream));
    if (vec) {
        for (int d = 0; d < h->P.dim; d++)
            LAUNCH(copy_on(h->stream, p + (size_t)d * h->N, in + (size_t)d * h->N, sizeof(double) * h->N,
                             hipMemcpyHostToDevice));
        if (field == PD_FIELD_F && use32(h)) {
            std::vector<float> tmp((size_t)h->N);
            for (int d = 0; d < h->P.dim; d++) {
                for (int i = 0; i < h->N; i++) tmp[i] = (float)in[(size_t)d * h->N + i];
                LAUNCH(copy_on(h->stream, h->f32.p[d], tmp.data(), sizeof(float) * h->N, hipMemcpyHostToDevice));
            }
        }
    } else {
        LAUNCH(copy_on(h->stream, p, in, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (field == PD_FIELD_FAC || field == PD_FIELD_SCR || field == PD_FIELD_IDIST || field == PD_FIELD_W ||
        field == PD_FIELD_M)
        h->coef_dirty = true;
    return PD_OK;
}

int pd_get_fail_bits(pd_handle *h, uint32_t *out)
{
    if (!h || !out) return PD_E_INVALID;
    REQUIRE(h->fail, "pd_build_neighbors first");
    LAUNCH(hipStreamSynchronize(h->stream));
    LAUNCH(copy_on(h->stream, out, h->fail, sizeof(uint32_t) * fail_words(h), hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_set_fail_bits(pd_handle *h, const uint32_t *in)
{
    if (!h || !in) return PD_E_INVALID;
    REQUIRE(h->fail, "pd_build_neighbors first");
    LAUNCH(hipStreamSynchronize(h->stream));
    fail_changed(h);
    LAUNCH(copy_on(h->stream, h->fail, in, sizeof(uint32_t) * fail_words(h), hipMemcpyHostToDevice));
    return PD_OK;
}

// -------------------------------------------------------------- cylinder
int pd_cylinder_init(pd_handle *h, const pd_cyl *c)
{
    // kernel_initial_cylinder / initial_cylinder_cpu, Solve_Kalthoff_Winkler.cu:9-37
    if (!h || !c) return PD_E_INVALID;
    REQUIRE(h->P.dim == 3 && h->N > 0, "3D problem required");
    h->cyl = *c;
    h->have_cyl = true;
    drop_graph(h);
    double st[12] = {0};
    st[1] = c->v;
    LAUNCH(copy_on(h->stream, h->S->cyl, st, sizeof(st), hipMemcpyHostToDevice));
    return PD_OK;
}

int pd_cylinder_get(pd_handle *h, double *state12)
{
    if (!h || !state12) return PD_E_INVALID;
    DevScalars S;
    int rc = read_scalars(h, &S);
    if (rc) return rc;
    std::memcpy(state12, S.cyl, sizeof(S.cyl));
    return PD_OK;
}

int pd_contact_step(pd_handle *h)
{
    if (!h) return PD_E_INVALID;
    REQUIRE(h->have_cyl, "pd_cylinder_init first");
    return contact_launch(h);
}

// ---------------------------------------------------------------- output
int pd_save_text(pd_handle *h, const char *path, int layout, int nx, int ny, int nz, double rad)
{
    // save_disp_cpu / save_kw_cpu, Save.cu:40-56,104-140: default ostream
    // formatting (6 significant digits), "ux uy [uz] dmg" per row.
    if (!h || !path) return PD_E_INVALID;
    const int D = h->P.dim, N = h->N, Ni = h->N_int;
    std::vector<double> u((size_t)D * N), dmg((size_t)N);
    int rc = pd_get_field(h, PD_FIELD_U, u.data());
    if (!rc) rc = pd_get_field(h, PD_FIELD_DMG, dmg.data());
    if (rc) return rc;
    FILE *fp = std::fopen(path, "w");
    REQUIRE(fp, "cannot open output file");
    auto row = [&](size_t i) {
        if (D == 2)
            std::fprintf(fp, "%g %g %g\n", u[i], u[N + i], dmg[i]);
        else
            std::fprintf(fp, "%g %g %g %g\n", u[i], u[N + i], u[2 * (size_t)N + i], dmg[i]);
    };
    if (layout == PD_LAYOUT_KW) {
        // z-layer 0 on the full nx x ny grid, NaN rows where the notches
        // removed particles (Save.cu:111-139)
        (void)nz;
        size_t num = 0;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double tx = (-(nx - 1) / 2.0 + i) * h->P.dx;
                double ty = (-(ny - 1) / 2.0 + j) * h->P.dx;
                bool flag = (tx > (-rad - 1.1 * h->P.dx) && tx < (-rad + 0.1 * h->P.dx)) |
                            (tx > (rad - 0.1 * h->P.dx) && tx < (rad + 1.1 * h->P.dx));
                if (flag && ty >= -1.0e-10) {
                    std::fprintf(fp, "nan nan 0 nan\n");
                } else {
                    if (num < (size_t)Ni) row(num);
                    num++;
                }
            }
    } else {
        for (int i = 0; i < Ni; i++) row((size_t)i);
    }
    std::fclose(fp);
    return PD_OK;
}

static const char kMagic[8] = {'P', 'D', 'M', 'I', '3', '5', '5', 'X'};

int pd_checkpoint_write(pd_handle *h, const char *path, int64_t step)
{
    // Binary checkpoint: header, u v a vh vho fo [dim][N], fail words, cylinder.
    if (!h || !path) return PD_E_INVALID;
    REQUIRE(h->fail, "pd_build_neighbors first");
    const int D = h->P.dim, N = h->N;
    FILE *fp = std::fopen(path, "wb");
    REQUIRE(fp, "cannot open checkpoint file");
    int64_t hdr[5] = {step, D, N, h->N_int, h->Nb};
    bool ok = std::fwrite(kMagic, 1, 8, fp) == 8 && std::fwrite(hdr, sizeof(hdr), 1, fp) == 1;
    std::vector<double> buf((size_t)D * N);
    const int fields[] = {PD_FIELD_U, PD_FIELD_V, PD_FIELD_A, PD_FIELD_VH, PD_FIELD_VHO, PD_FIELD_FO, PD_FIELD_F};
    for (int f : fields) {
        if (!ok) break;
        int rc = pd_get_field(h, f, buf.data());
        if (rc) {
            std::fclose(fp);
            return rc;
        }
        ok = std::fwrite(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
    }
    std::vector<uint32_t> fw((size_t)fail_words(h));
    if (ok && pd_get_fail_bits(h, fw.data()) == PD_OK)
        ok = std::fwrite(fw.data(), sizeof(uint32_t), fw.size(), fp) == fw.size();
    double cyl[12] = {0};
    if (ok && h->have_cyl) pd_cylinder_get(h, cyl);
    if (ok) ok = std::fwrite(cyl, sizeof(cyl), 1, fp) == 1;
    std::fclose(fp);
    REQUIRE(ok, "checkpoint write failed");
    return PD_OK;
}

int pd_checkpoint_read(pd_handle *h, const char *path, int64_t *step)
{
    if (!h || !path) return PD_E_INVALID;
    REQUIRE(h->fail, "pd_build_neighbors first");
    h->rec_u_valid = false;
    const int D = h->P.dim, N = h->N;
    FILE *fp = std::fopen(path, "rb");
    REQUIRE(fp, "cannot open checkpoint file");
    char mg[8];
    int64_t hdr[5];
    bool ok = std::fread(mg, 1, 8, fp) == 8 && std::memcmp(mg, kMagic, 8) == 0 &&
              std::fread(hdr, sizeof(hdr), 1, fp) == 1 && hdr[1] == D && hdr[2] == N && hdr[3] == h->N_int &&
              hdr[4] == h->Nb;
    std::vector<double> buf((size_t)D * N);
    const int fields[] = {PD_FIELD_U, PD_FIELD_V, PD_FIELD_A, PD_FIELD_VH, PD_FIELD_VHO, PD_FIELD_FO, PD_FIELD_F};
    for (int f : fields) {
        if (!ok) break;
        ok = std::fread(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
        if (ok && pd_set_field(h, f, buf.data()) != PD_OK) ok = false;
    }
    std::vector<uint32_t> fw((size_t)fail_words(h));
    if (ok) ok = std::fread(fw.data(), sizeof(uint32_t), fw.size(), fp) == fw.size();
    if (ok) ok = pd_set_fail_bits(h, fw.data()) == PD_OK;
    double cyl[12];
    if (ok) ok = std::fread(cyl, sizeof(cyl), 1, fp) == 1;
    std::fclose(fp);
    REQUIRE(ok, "checkpoint read failed or does not match this problem");
    if (h->have_cyl) LAUNCH(copy_on(h->stream, h->S->cyl, cyl, sizeof(cyl), hipMemcpyHostToDevice));
    if (step) *step = hdr[0];
    return PD_OK;
}

} // extern "C"

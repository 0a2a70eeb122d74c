// pd_internal.h -- handle layout and launcher prototypes shared by the HIP
// translation units of libpd_gpu1_amd.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "../../include/pd_gpu.h"

// Three component pointers passed by value to kernels (SoA [dim][N]).
struct D3 { double *p[3]; };
struct CD3 { const double *p[3]; };
struct F3 { float *p[3]; };
struct CF3 { const float *p[3]; };

// Stream-ordered copy and fill, waited for before they return.  Handles keep
// all of their work on their own (non-blocking) stream and never use the
// legacy stream: HIP refuses a legacy-stream operation while any stream is
// capturing a graph, so this is what lets calls on different handles run
// from different host threads while one of them captures its step graph.
static inline hipError_t copy_on(hipStream_t s, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
static inline hipError_t fill_on(hipStream_t s, void *dst, int value, size_t bytes)
{
    hipError_t e = hipMemsetAsync(dst, value, bytes, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

static inline CD3 cd3(const D3 &a) { return CD3{{a.p[0], a.p[1], a.p[2]}}; }
static inline CF3 cf3(const F3 &a) { return CF3{{a.p[0], a.p[1], a.p[2]}}; }

// Timing events that are destroyed on every return path.
template <int N>
struct HipEvents {
    hipEvent_t e[N] = {};
    hipError_t create()
    {
        for (auto &ev : e) {
            const hipError_t r = hipEventCreate(&ev);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    ~HipEvents()
    {
        for (auto ev : e)
            if (ev) (void)hipEventDestroy(ev);
    }
};

// Device scalars of one handle (one allocation, written by kernels).
struct DevScalars {
    double c;          // ADR damping coefficient of the last iteration
    double resid;      // relative residual of the state before the last update
    double sums[8];    // cn1, cn2, rnum, rden, react x/y/z, spare
    double cyl[12];    // cylinder vel[3], disp[3], acc[3], force[3]
    int nonfinite;     // set when a reduction meets NaN/Inf
    int ct_counter;    // ADR iteration index handed to the next update
    int ct_use;        // ADR iteration index of the current update
    int done;          // ADR stop flag (device-side termination)
    int iters;         // ADR iterations executed since the flag was cleared
    int capacity;      // neighbour-capacity overflow flag
    unsigned pad[4];
};

// Bond coefficients recomputed from |xi| in the mixed kernels (see geom_coef
// in mixed_bond.h): coef = scale * fac(|xi|) [* 2^(k2 |xi|^2)].
struct GeomCoef {
    float hi_dx, inv_dx;   // fac = clamp(hi/dx - |xi|/dx, 0, 1), hi = delta + dx/2
    float scale;           // vol (LPS omega = w fac vol) or vol*bc (PMB kappa, scr = 1), folded
                           // into each pass's constants rather than applied per bond
    float k2;              // LPS: -log2(e)/delta^2, so w = 2^(k2 |xi|^2); PMB: 0
};

// Mixed mode particle record, 8 words (32 bytes) per particle:
//   words 0..2  reference position, quantized: I = rint(x0 / q) mod 2^32
//   word  3     A theta/m (float), written by the LPS dilatation pass
//   words 4..6  displacement u rounded to FP32, in units of q: RN(u) / q
//   word  7     B/m (float), likewise (A, B of the LPS force, see lps_mixed_consts)
// q is a power of two small against the horizon, so the FP32 bond vector
// (float)(int32)(I_c - I_i), in units of q, differs from the exact one by at
// most one unit per component and is the same whichever kernel or tile forms
// it.  The mixed kernels work in units of q throughout (power-of-two scalings
// change no rounding).
constexpr int kRecW = 8;

// Records to refresh with a new u (rec null: none) and 1/q.
struct RecU {
    uint32_t *rec;
    double iq;
};

// What the mixed kernels need besides their bond streams: the quantum, the
// critical stretch, the half-width of the band in which a bond's break is
// decided by its FP64 stretch (mix_window below; M bounds the displacement
// components of the bond's two ends) and the FP64 state that decision reads.
struct MixCtx {
    float s0f;         // critical stretch, FP32
    float wfloor;      // floor of the outer band, max(near_window s0, 3e-6)
    float wn;          // near_window s0: the least inner band
    float wpos, k_u;   // what the records add: wpos + k_u M (k_u per unit q of M)
    CD3 x, u;          // FP64 reference positions and displacements
};

// Half-width of the band around s0 in which the FP64 stretch decides a break.
// Its floor is at least near_window*s0 and at least 3e-6, which bounds the
// FP32 evaluation of the stretch formula: with r = |eta|/|xi| and the inputs
// and the three products of eta.(2 xi + eta) each rounded once, the error is
// below 2^-24 r (12 + 6 r), and |eta| <= |xi + eta| + |xi| gives r <= 2 + s,
// so about 1.5e-6 near s0.  The records perturb the inputs before that
// rounding: the quantized positions move xi by at most d = sqrt(3) q, the FP32
// displacements move eta by at most e = sqrt(3) 2^-23 (1 + 2^-23) M, M the
// larger displacement component of the two ends.  As |xi'+eta'| is within
// d + e of |xi+eta| and |xi'| within d of |xi|, the stretch of the perturbed
// bond is within ((2 + s) d + e) / (|xi| - d) of s.  mix_window gives that,
// with s <= s0 + 0.01 and |xi| >= lmin, the shortest reference bond, as
// wpos + k_u * M.  Outer band: wfloor + wpos + k_u * M, the same for every
// bond of a row (or every bond), a cheap test on the FP32 stretch; outside it
// the FP32 comparison equals the FP64 one for every bond, rotated or not.
// Inner band, for the bonds inside the outer one: the evaluation bound with
// the bond's own r, 2^-24 r (32 + 16 r) (the 12 + 6 r above with a margin for
// the approximate reciprocals and for r itself in FP32), but at least
// near_window * s0, plus wpos + k_u * M.  Only bonds inside the inner band
// read the FP64 state; every break decision equals the FP64 one.
inline float stretch_window(double s0, double near_window)
{
    const double w = near_window * s0;
    return (float)(w > 3.0e-6 ? w : 3.0e-6);
}

inline void mix_window(double s0, double near_window, double q, double lmin, MixCtx *mc)
{
    const double d = 1.7320508075688772 * q;
    const double den = lmin - d;
    const double slack = 1.0 + 1.0e-6;  // covers the FP32 rounding of the band sums
    mc->wfloor = (float)(slack * (double)stretch_window(s0, near_window));
    mc->wn = (float)(slack * near_window * s0);
    mc->wpos = (float)(slack * (2.01 + s0) * d / den);
    mc->k_u = (float)(slack * 1.7320508075688772 * 0x1p-23 * (1.0 + 0x1p-23) / den);
}

// Per-tile staged neighbour lists of the LDS kernels (k_force_lds.hip).
struct LdsTiles {
    int ntiles = 0, ntiles_int = 0;  // tiles; tiles holding integrated rows (they come first)
    int max_staged = 0;              // largest staged set (records of LDS per workgroup)
    int block = 512;                 // threads per workgroup (PMB, LPS dilatation)
    int block2 = 512;                // threads per workgroup of the LPS force pass
    int32_t *tile_start = nullptr;   // [ntiles + 1] row ranges of order[]
    int64_t *stage_ptr = nullptr;    // [ntiles + 1] ranges of stage_idx
    int32_t *stage_idx = nullptr;    // staged particle indices, ascending within a tile
    uint16_t *lidx = nullptr;        // per bond: position of col[k] in its tile's staged list (bits 0-14)
                                     // and a copy of its fail bit (bit 15)
};

struct pd_handle {
    pd_params P{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int N = 0, N_int = 0;
    int64_t Nb = 0;

    // particles (FP64 SoA)
    D3 x{}, u{}, v{}, a{}, b{}, f{}, vh{}, vho{}, fo{};
    F3 f32{};                 // mixed-mode force
    double *m = nullptr, *theta = nullptr, *dmg = nullptr;
    float *minv32 = nullptr, *theta32 = nullptr;  // mixed LPS: 1/m and dilatation
    uint32_t *rec = nullptr;      // mixed mode: 32-byte particle records (see MixCtx)
    bool rec_u_valid = false;     // the u words of every record equal u rounded to FP32
    double rec_q = 0.0;           // quantum of the reference positions in the records
    double lmin = 0.0;            // shortest reference bond length
    double xabs = 0.0;            // largest |coordinate| of the reference positions
    bool b_zero = true;           // the body force b is zero everywhere (the Verlet passes skip it)
    double *terms = nullptr;      // [4][N] per-particle terms of the ADR and contact reductions
    int32_t *order = nullptr;     // mixed mode: tiled processing order (null = index order)
    int tile_cells = -1;          // tile edge in units of dx (0 = index order, -1 = default)
    int lds_mode = -1;            // LDS-staged tiles: -1 where they fit (default), 0 off, 1 required
    bool lds_ok = false;          // the LDS tile lists are built and in use
    bool lidx_fail_ok = false;    // bit 15 of every lidx word equals the bond's fail bit
    int lds_block = 512;          // threads per workgroup of the LDS kernels (pass 0)
    int lds_block2 = 512;         // ... of the LPS force pass
    int fp64_group = -1;          // lanes per row of the FP64 PMB kernel (1 = thread per row, -1 = by row length)
    LdsTiles lds;
    double *react[3] = {nullptr, nullptr, nullptr};

    // bonds (CSR)
    int32_t *row_ptr = nullptr;
    int32_t *col = nullptr;
    uint32_t *fail = nullptr;
    double *idist = nullptr, *fac = nullptr, *scr = nullptr, *kappa = nullptr;
    double *w = nullptr, *e = nullptr;
    float2 *bnd = nullptr;    // mixed mode: per bond {|xi| / q, kappa or omega} in FP32
    int coef_mode = -1;       // mixed mode: -1 recompute from |xi| where that matches bnd, 0 stream bnd, 1 required
    bool geom_ok = false;     // the kernels recompute the bond coefficients (gc) instead of reading bnd
    GeomCoef gc{};
    bool coef_dirty = true;   // kappa / bnd need recomputation
    bool have_scr = false;
    bool have_weights = false;

    // neighbour-search scratch
    double lo[3] = {0, 0, 0};
    int ncell[3] = {1, 1, 1};

    // reductions
    int red_blocks = 16, red_threads = 256;
    double *partials = nullptr;     // [red_blocks * 8]
    DevScalars *S = nullptr;

    // host mirror of the cylinder parameters
    pd_cyl cyl{};
    bool have_cyl = false;

    // force launch shape (autotuned group size)
    int group = 0;   // pass 0 (PMB or LPS dilatation); 0 = built-in choice
    int group2 = 0;  // pass 1 (LPS force)

    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_force_ms = 0.0;

    // captured step graph
    hipGraphExec_t step_graph = nullptr;
    int step_graph_kind = -1;
    double graph_tol = -2.0;
};

#define PD_TRY_HIP(h, expr)                                                       \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(_e);         \
            return PD_E_HIP;                                                      \
        }                                                                         \
    } while (0)

// ---------------------------------------------------------------- launchers
// Each returns hipSuccess or the launch error.  Defined next to the kernels.

// k_neighbor.hip
int build_neighbors_impl(pd_handle *h);
int build_tile_order(pd_handle *h, double tile_len, bool want_tiles);
hipError_t launch_cell_ids(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                           const int *ncell, uint32_t *cell_of, uint32_t *idx);
hipError_t launch_cell_bounds(hipStream_t s, int N, const uint32_t *sorted_cell, int ncells,
                              int32_t *cell_start);
hipError_t launch_neigh_count(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                              const int *ncell, const int32_t *cell_start, const uint32_t *items,
                              double d2, int32_t *cnt);
hipError_t launch_neigh_fill(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                             const int *ncell, const int32_t *cell_start, const uint32_t *items,
                             double d2, const int32_t *row_ptr, int32_t *col);
hipError_t launch_row_sort(hipStream_t s, int N, const int32_t *row_ptr, int32_t *col, int max_row);

// k_setup.hip
hipError_t launch_vol_corr(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col, CD3 x,
                           double delta, double dx, double *idist, double *fac);
hipError_t launch_surface_probe(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                                CD3 x, int axis, const double *idist, const double *fac, double bc,
                                double vol, double sedload, double *fncst);
hipError_t launch_surface_scr(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                              CD3 x, const double *idist, const double *fncst, double *scr);
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, int64_t n, uint32_t v);
hipError_t launch_fill_f64(hipStream_t s, double *p, int64_t n, double v);
hipError_t launch_crack_2d(hipStream_t s, int N_int, const int32_t *rp, const int32_t *col, CD3 x,
                           double clength, double loc_x, double loc_y, int vertical, double t,
                           double sq1, double half_s, uint32_t *fail);
hipError_t launch_crack_kw(hipStream_t s, int N_int, const int32_t *rp, const int32_t *col, CD3 x,
                           double rad, uint32_t *fail);
hipError_t launch_lps_weights(hipStream_t s, int N, const int32_t *rp, const double *idist,
                              const double *fac, double delta, double vol, int w_given, double *w, double *m);
hipError_t launch_pmb_coef(hipStream_t s, int64_t nb, const double *fac, const double *scr, double vol,
                           double bc, const double *idist, double iq, double *kappa, float2 *bnd);
hipError_t launch_lps_coef(hipStream_t s, int64_t nb, const double *w, const double *fac, double vol,
                           const double *idist, double iq, float2 *bnd, int N, const double *m, float *minv32);
hipError_t launch_copy16(hipStream_t s, const void *a, void *b, int64_t bytes);
hipError_t launch_damage(hipStream_t s, int N_int, const int32_t *rp, const uint32_t *fail,
                         const double *fac, double vol, double *dmg);

// k_force.hip
hipError_t launch_pmb_fp64(hipStream_t s, const int *stop, int dim, int n_rows, const int32_t *rp, const int32_t *col,
                           const double *kappa, const double *idist, uint32_t *fail, CD3 x, CD3 u,
                           D3 f, double s0, double ymin, int group);
hipError_t launch_lps_theta_fp64(hipStream_t s, const int *stop, int dim, int N, const int32_t *rp, const int32_t *col,
                                 CD3 x, CD3 u, const double *idist, const double *fac, const double *w,
                                 const double *m, const uint32_t *fail, double vol, double pratio,
                                 int skip_broken, double *e, double *theta);
hipError_t launch_lps_force_fp64(hipStream_t s, const int *stop, int dim, int N_int, const int32_t *rp, const int32_t *col,
                                 CD3 x, CD3 u, const double *idist, const double *fac, const double *w,
                                 const double *m, const double *theta, const double *e, double vol,
                                 double pratio, double K, double G, double s0, double ymin,
                                 uint32_t *fail, D3 f);

// k_force_mixed.hip (particle records: kRecW words per particle, see MixCtx)
hipError_t launch_pack_x0(hipStream_t s, int dim, int N, CD3 x, double q, uint32_t *rec);
hipError_t launch_pack_u(hipStream_t s, const int *stop, int dim, int N, CD3 u, double q, uint32_t *rec);
hipError_t launch_min_bond(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col, CD3 x,
                           unsigned long long *out);
hipError_t launch_geom_check(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                             const uint32_t *rec, const float2 *bnd, GeomCoef gc, int *ok);
hipError_t launch_pmb_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int n_rows,
                            const int32_t *rp, const int32_t *col, const float2 *bnd, uint32_t *fail,
                            const uint32_t *rec, F3 f, double s0, const MixCtx &mc, double ymin, const GeomCoef *gc);
hipError_t launch_lps_theta_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int N,
                                  const int32_t *rp, const int32_t *col, const float2 *bnd, const uint32_t *fail,
                                  uint32_t *rec, const float *minv32, float ctheta, float A, float B,
                                  int skip_broken, float *theta32, const GeomCoef *gc);
hipError_t launch_lps_force_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int N_int,
                                  const int32_t *rp, const int32_t *col, const float2 *bnd, uint32_t *fail,
                                  const uint32_t *rec, double s0, const MixCtx &mc, double ymin, F3 f,
                                  const GeomCoef *gc);
// k_force_lds.hip
int stage_cap();
int lds_record_bytes();
hipError_t launch_lidx_sync(hipStream_t s, int64_t nb, const uint32_t *fail, uint16_t *lidx);
hipError_t launch_stage_tiles(hipStream_t s, int phase, int ntiles, const int32_t *tile_start, const int32_t *order,
                              const int32_t *rp, const int32_t *col, int32_t *count, const int64_t *stage_ptr,
                              int32_t *stage_idx, uint16_t *lidx);
hipError_t launch_pmb_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group, const int32_t *order,
                          const int32_t *rp, const float2 *bnd, uint32_t *fail, const uint32_t *rec, F3 f, double s0,
                          const MixCtx &mc, double ymin, const GeomCoef *gc);
hipError_t launch_lps_theta_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group,
                                const int32_t *order, const int32_t *rp, const float2 *bnd, uint32_t *rec,
                                const float *minv32, float ctheta, float A, float B, int skip_broken, float *theta32,
                                const GeomCoef *gc);
hipError_t launch_lps_force_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group,
                                const int32_t *order, const int32_t *rp, const float2 *bnd, uint32_t *fail,
                                const uint32_t *rec, double s0, const MixCtx &mc, double ymin, F3 f,
                                const GeomCoef *gc);

// k_integrate.hip
hipError_t launch_verlet1(hipStream_t s, int dim, int N_int, int n_top, int n_bot, double dt, double load,
                          D3 v, CD3 a, D3 u, double *cyl_state, RecU rec);
hipError_t launch_verlet21(hipStream_t s, int dim, int N_int, int n_top, int n_bot, double dt, double dens,
                           double load, D3 v, CD3 f, CF3 f32, int use32, CD3 b, D3 u, RecU rec);
hipError_t launch_verlet2(hipStream_t s, int dim, int N_int, double dt, double dens, D3 v, D3 a, CD3 f,
                          CF3 f32, int use32, CD3 b);
hipError_t launch_adr_reduce(hipStream_t s, int blocks, int threads, int dim, int N_int, int N,
                             double mass, double dt, CD3 u, CD3 f, CF3 f32, int use32, CD3 fo, CD3 vho,
                             CD3 b, double *partials, DevScalars *S, double tol, double *terms);
hipError_t launch_adr_update(hipStream_t s, int dim, int N_int, double mass, double dt, CD3 f,
                             CF3 f32, int use32, CD3 b, D3 vh, D3 v, D3 vho, D3 u, D3 fo,
                             DevScalars *S, RecU rec);
hipError_t launch_contact(hipStream_t s, int blocks, int threads, int N, const double *dmg, CD3 x, D3 u,
                          D3 v, pd_cyl cyl, double dt, double dens, double vol, double *partials,
                          DevScalars *S, RecU rec, double *terms);
hipError_t launch_set_ct(hipStream_t s, DevScalars *S, int ct, int clear_done);
hipError_t launch_field_sum(hipStream_t s, int blocks, int threads, const double *v, int64_t n,
                            double *partials, DevScalars *S);
hipError_t launch_field_sum_to(hipStream_t s, int blocks, int threads, const double *v, int64_t n, double *partials,
                               double *out);
hipError_t launch_strip_displace(hipStream_t s, int N_int, int n_top, int n_bot, double dtop,
                                 double dbot, double *uy);
hipError_t launch_f32_to_f64(hipStream_t s, const float *a, double *b, int64_t n);

// k_force_mixed.hip -- mixed-precision PMB and LPS force kernels (throughput
// path) and the particle-record packing they gather from.
//
// Layout.  Neighbour state is gathered from a 32-byte record per particle
// (pd_internal.h): the reference position quantized to a power-of-two quantum
// q (three 32-bit integers), theta/m and 1/m of the LPS dilatation pass, and
// the displacement rounded to FP32.  One gather reads one 32-byte sector
// instead of six scattered SoA words.  The u words are rewritten by the
// integrators that update u (or by the pack kernel when u changed elsewhere);
// the FP64 SoA arrays stay the state of record, and the FP64 re-evaluation
// of a bond near the critical stretch reads them.
//
// Work split.  A group of G lanes (G in 4..64, a divisor of the wave64) owns
// one particle; lanes stride over its CSR row two bonds at a time.  The bond
// stream (col, {|xi|, coefficient}, fail word) of the next pair is fetched
// while the neighbour records of the current pair are gathered, and broken
// or absent bonds are masked arithmetically instead of branched around.  Blocks are renumbered so that
// each of the 8 XCDs (which receive workgroups round-robin) walks one
// contiguous eighth of the particle range, keeping neighbour records in that
// XCD's L2.
//
// Arithmetic.  The stretch is formed without cancellation from
// eta = u_j - u_i (FP32 displacements) and xi = q (I_j - I_i) (an exact
// integer difference rounded once to FP32):
//     s = eta.(2 xi + eta) / (|xi| (|xi + eta| + |xi|)),
// equal to (|xi+eta| - |xi|)/|xi| of Force_Bond.cu:189.  Bond terms are FP32;
// per-lane sums accumulate them in FP64 (exact for each term) and the group
// sum reduces the FP64 partials over xor shuffles.  A bond whose FP32 stretch
// lies within the band of mix_window (mixed_bond.h) around s0 is
// re-evaluated with the FP64 parity expression (Force_Bond.cu:185-189,
// Surf_Corr.cu:324-326) before its break is decided, so breaks follow the
// FP64 stretch of the current state.
#include "mixed_bond.h"

namespace {

// Row walk shared by the three kernels: lanes take bonds k, k+G of the row,
// then k+2G, k+3G, ...; the slots of the next pair are fetched while the
// neighbour records of the current pair are in flight.
#define PD_ROW_WALK(NEED_INTACT, BODY)                                                           \
    {                                                                                            \
        const int32_t kb = rp[i], ke = rp[i + 1];                                                \
        int32_t k = kb + lane;                                                                   \
        Slot sa = fetch_slot<GC>(k, ke, i, col, bnd, fail, NEED_INTACT);                             \
        Slot sb = fetch_slot<GC>(k + G, ke, i, col, bnd, fail, NEED_INTACT);                         \
        while (k < ke) {                                                                         \
            const RecQ ra = load_rec(rec, sa.c), rb = load_rec(rec, sb.c);                       \
            const Slot na = fetch_slot<GC>(k + 2 * G, ke, i, col, bnd, fail, NEED_INTACT);           \
            const Slot nb = fetch_slot<GC>(k + 3 * G, ke, i, col, bnd, fail, NEED_INTACT);           \
            BODY(ra, sa, k);                                                                     \
            BODY(rb, sb, k + G);                                                                 \
            sa = na;                                                                             \
            sb = nb;                                                                             \
            k += 2 * G;                                                                          \
        }                                                                                        \
    }

template <int DIM, int G, bool GC>
__global__ __launch_bounds__(256) void k_pmb_mixed(const int *__restrict__ stop, const int32_t *__restrict__ order, int n_rows,
                                                   const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                   const float2 *__restrict__ bnd, uint32_t *fail,
                                                   const uint32_t *__restrict__ rec, F3 f, double s0, MixCtx mc,
                                                   double ymin, GeomCoef gc)
{
    if (stop && *stop) return;
    const int gid = (xcd_block() * 256 + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (gid >= n_rows) return;  // the whole group leaves together
    const int i = order ? order[gid] : gid;
    const Rec ri = load_rec(rec, i).r;
    const float mi = umag(ri);
    double ax = 0.0, ay = 0.0, az = 0.0;
#define PMB_BODY(R, S, K)                                                                        \
    {                                                                                            \
        const BondGeo<DIM> g = geo<DIM>(ri, R.r);                                          \
        float L0 = S.L0, cf = S.coef, rL0 = 0.f;                                                 \
        if (GC) geom_coef<false>(g.xi2, gc, L0, cf, rL0);                                        \
        const float s = g.num * frcp(L0 * g.len_sum(L0));                                        \
        const float t = masked(cf * s * g.rL, S.mask);                                           \
        ax += (double)(t * g.dx);                                                                \
        ay += (double)(t * g.dy);                                                                \
        if (DIM == 3) az += (double)(t * g.dz);                                                  \
        check_break<DIM, false, false>(i, S.c, nullptr, 0, s, S, K, bond_window(mc, mi, umag(R.r)), ri, R.r, \
                                       g.xi2, mc, s0, ymin, fail);                                                    \
    }
    PD_ROW_WALK(true, PMB_BODY)
#undef PMB_BODY
    const double fs = GC ? (double)gc.scale : 1.0;  // recomputed coefficients leave out vol*bc
    const float sx = group_sum<G>(ax, fs), sy = group_sum<G>(ay, fs);
    const float sz = DIM == 3 ? group_sum<G>(az, fs) : 0.f;
    if (lane == 0) {
        f.p[0][i] = sx;
        f.p[1][i] = sy;
        if (DIM == 3) f.p[2][i] = sz;
    }
}

template <int DIM, int G, bool GC>
__global__ __launch_bounds__(256) void k_lps_theta_mixed(const int *__restrict__ stop, const int32_t *__restrict__ order, int N,
                                                         const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const float2 *__restrict__ bnd,
                                                         const uint32_t *__restrict__ fail, uint32_t *rec,
                                                         const float *__restrict__ minv32, float ctheta, float A,
                                                         float B, int skip_broken, float *theta32, GeomCoef gc)
{
    if (stop && *stop) return;
    const int gid = (xcd_block() * 256 + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (gid >= N) return;
    const int i = order ? order[gid] : gid;
    const Rec ri = load_rec(rec, i).r;
    double acc = 0.0;
    const bool need_intact = skip_broken != 0;
#define THETA_BODY(R, S, K)                                                                      \
    {                                                                                            \
        const BondGeo<DIM> g = geo<DIM>(ri, R.r);                                             \
        float L0 = S.L0, cf = S.coef, rL0 = 0.f;                                                 \
        if (GC) geom_coef<true>(g.xi2, gc, L0, cf, rL0);                                         \
        const float e = g.num * frcp(g.len_sum(L0)); /* |xi+eta| - |xi| */                       \
        acc += (double)masked(cf * L0 * e, S.mask);                                              \
    }
    PD_ROW_WALK(need_intact, THETA_BODY)
#undef THETA_BODY
    const float th = group_sum<G>(acc);
    if (lane == 0) {
        // recomputed coefficients leave out vol: the constants carry it
        const float cth = GC ? ctheta * gc.scale : ctheta, As = GC ? A * gc.scale : A, Bs = GC ? B * gc.scale : B;
        const float r = minv32[i];
        const float theta = cth * r * th;
        theta32[i] = theta;
        rec[(size_t)i * kRecW + 3] = __float_as_uint(As * (theta * r));
        rec[(size_t)i * kRecW + 7] = __float_as_uint(Bs * r);
    }
}

template <int DIM, int G, bool GC>
__global__ __launch_bounds__(256) void k_lps_force_mixed(const int *__restrict__ stop, const int32_t *__restrict__ order, int N_int,
                                                         const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const float2 *__restrict__ bnd, uint32_t *fail,
                                                         const uint32_t *__restrict__ rec, double s0, MixCtx mc,
                                                         double ymin, F3 f, GeomCoef gc)
{
    if (stop && *stop) return;
    const int gid = (xcd_block() * 256 + (int)threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (gid >= N_int) return;
    const int i = order ? order[gid] : gid;
    const RecQ qi = load_rec(rec, i);
    const Rec &ri = qi.r;
    const float mi = umag(ri);
    double ax = 0.0, ay = 0.0, az = 0.0;
    // (t_i + t_j) vol fac = omega [xi (Aq_i + Aq_j) + e (Br_i + Br_j)], Aq = A theta/m and
    // Br = B/m from the records
#define LPSF_BODY(R, S, K)                                                                       \
    {                                                                                            \
        const BondGeo<DIM> g = geo<DIM>(ri, R.r);                                          \
        float L0 = S.L0, cf = S.coef, rL0 = 0.f;                                                 \
        if (GC) geom_coef<true>(g.xi2, gc, L0, cf, rL0);                                         \
        const float e = g.num * frcp(g.len_sum(L0));                                             \
        const float s = e * (GC ? rL0 : frcp(L0));                                               \
        const float tt = masked(cf * __builtin_fmaf(L0, qi.q + R.q, e * (qi.m + R.m)) * g.rL, S.mask); \
        ax += (double)(tt * g.dx);                                                               \
        ay += (double)(tt * g.dy);                                                               \
        if (DIM == 3) az += (double)(tt * g.dz);                                                 \
        check_break<DIM, true, false>(i, S.c, nullptr, 0, s, S, K, bond_window(mc, mi, umag(R.r)), ri, R.r, \
                                      g.xi2, mc, s0, ymin, fail);                                                     \
    }
    PD_ROW_WALK(true, LPSF_BODY)
#undef LPSF_BODY
    const float sx = group_sum<G>(ax), sy = group_sum<G>(ay);
    const float sz = DIM == 3 ? group_sum<G>(az) : 0.f;
    if (lane == 0) {
        f.p[0][i] = sx;
        f.p[1][i] = sy;
        if (DIM == 3) f.p[2][i] = sz;
    }
}

// Records: the quantized reference position (words 0..2; words 3 and 7
// cleared) ...
__global__ void k_pack_x0(int dim, int N, CD3 x, double iq, uint32_t *rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint4 w;
    w.x = (uint32_t)(unsigned long long)__double2ll_rn(x.p[0][i] * iq);
    w.y = (uint32_t)(unsigned long long)__double2ll_rn(x.p[1][i] * iq);
    w.z = dim == 3 ? (uint32_t)(unsigned long long)__double2ll_rn(x.p[2][i] * iq) : 0u;
    w.w = 0u;
    reinterpret_cast<uint4 *>(rec + (size_t)i * kRecW)[0] = w;
    rec[(size_t)i * kRecW + 7] = 0u;
}

// ... and the displacement rounded to FP32, in units of q (words 4..6).
__global__ void k_pack_u(const int *__restrict__ stop, int dim, int N, CD3 u, double iq, uint32_t *rec)
{
    if (stop && *stop) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint32_t *r = rec + (size_t)i * kRecW + 4;
    *reinterpret_cast<float2 *>(r) = make_float2((float)(u.p[0][i] * iq), (float)(u.p[1][i] * iq));
    r[2] = __float_as_uint(dim == 3 ? (float)(u.p[2][i] * iq) : 0.f);
}

// Shortest reference bond: the minimum of |xi|^2 over every bond (FP64), as
// the bits of a non-negative double (which order like the unsigned integers).
__global__ void k_min_bond(int dim, int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col, CD3 x,
                           unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = INFINITY;
    if (i < N)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int c = col[k];
            const double a0 = x.p[0][c] - x.p[0][i], a1 = x.p[1][c] - x.p[1][i];
            double q2 = a0 * a0 + a1 * a1;
            if (dim == 3) {
                const double a2 = x.p[2][c] - x.p[2][i];
                q2 = q2 + a2 * a2;
            }
            m = fmin(m, q2);
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0 && m < INFINITY) atomicMin(out, (unsigned long long)__double_as_longlong(m));
}

// Does every stored FP32 bond coefficient equal its recomputation from |xi|
// (geom_coef, |xi|^2 formed from the records as geo() forms it; lengths in
// units of the record quantum)?  *ok is
// cleared at the first bond that does not (user-set coefficients, surface
// correction, a missing volume correction).
__global__ void k_geom_check(int dim, int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                             const uint32_t *__restrict__ rec, const float2 *__restrict__ bnd, GeomCoef gc, int *ok)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Rec ri = load_rec(rec, i).r;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const Rec rc = load_rec(rec, col[k]).r;
        const float ex = qdiff(ri.ix, rc.ix), ey = qdiff(ri.iy, rc.iy);
        float xi2 = __builtin_fmaf(ey, ey, ex * ex);
        if (dim == 3) {
            const float ez = qdiff(ri.iz, rc.iz);
            xi2 = __builtin_fmaf(ez, ez, xi2);
        }
        float L0, cf, rL0;
        if (gc.k2 != 0.f)
            geom_coef<true>(xi2, gc, L0, cf, rL0);
        else
            geom_coef<false>(xi2, gc, L0, cf, rL0);
        const float2 b = bnd[k];
        cf *= gc.scale;
        if (!(fabsf(L0 - b.x) <= 1e-6f * b.x && fabsf(cf - b.y) <= 1e-5f * fabsf(b.y) + 1e-30f)) {
            atomicAnd(ok, 0);
            return;
        }
    }
}

inline int grid8(int64_t threads)
{
    int64_t b = (threads + 255) / 256;
    return b > 0 ? (int)((b + 7) / 8 * 8) : 8;  // the kernels bounds-check their rows
}

template <int DIM, int G, bool GC>
hipError_t pmb_g(hipStream_t s, const int *stop, const int32_t *order, int n, const int32_t *rp, const int32_t *col,
                 const float2 *bnd, uint32_t *fail, const uint32_t *rec, F3 f, double s0, const MixCtx &mc,
                 double ymin, GeomCoef gc)
{
    k_pmb_mixed<DIM, G, GC><<<grid8((int64_t)n * G), 256, 0, s>>>(stop, order, n, rp, col, bnd, fail, rec, f, s0, mc,
                                                                  ymin, gc);
    return hipGetLastError();
}

template <int DIM, int G, bool GC>
hipError_t theta_g(hipStream_t s, const int *stop, const int32_t *order, int n, const int32_t *rp, const int32_t *col,
                   const float2 *bnd, const uint32_t *fail, uint32_t *rec, const float *minv32, float ctheta,
                   float A, float B, int skip, float *theta32, GeomCoef gc)
{
    k_lps_theta_mixed<DIM, G, GC><<<grid8((int64_t)n * G), 256, 0, s>>>(stop, order, n, rp, col, bnd, fail, rec,
                                                                        minv32, ctheta, A, B, skip, theta32, gc);
    return hipGetLastError();
}

template <int DIM, int G, bool GC>
hipError_t lpsf_g(hipStream_t s, const int *stop, const int32_t *order, int n, const int32_t *rp, const int32_t *col,
                  const float2 *bnd, uint32_t *fail, const uint32_t *rec, double s0, const MixCtx &mc,
                  double ymin, F3 f, GeomCoef gc)
{
    k_lps_force_mixed<DIM, G, GC><<<grid8((int64_t)n * G), 256, 0, s>>>(stop, order, n, rp, col, bnd, fail, rec, s0,
                                                                        mc, ymin, f, gc);
    return hipGetLastError();
}

#define PD_GSWITCH(FN, DIM, G, GC, GCV, ...)              \
    switch (G) {                                          \
    case 4: return FN<DIM, 4, GC>(__VA_ARGS__, GCV);      \
    case 8: return FN<DIM, 8, GC>(__VA_ARGS__, GCV);      \
    case 16: return FN<DIM, 16, GC>(__VA_ARGS__, GCV);    \
    case 32: return FN<DIM, 32, GC>(__VA_ARGS__, GCV);    \
    case 64: return FN<DIM, 64, GC>(__VA_ARGS__, GCV);    \
    default: return hipErrorInvalidValue;                 \
    }

// gc != nullptr: recompute the bond coefficients from |xi| (no coefficient stream)
#define PD_DISPATCH_G(FN, DIM, G, GCP, ...)                                  \
    if (GCP) {                                                               \
        PD_GSWITCH(FN, DIM, G, true, *(GCP), __VA_ARGS__)                    \
    } else {                                                                 \
        PD_GSWITCH(FN, DIM, G, false, GeomCoef{}, __VA_ARGS__)               \
    }

} // namespace

hipError_t launch_geom_check(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                             const uint32_t *rec, const float2 *bnd, GeomCoef gc, int *ok)
{
    if (N <= 0) return hipSuccess;
    k_geom_check<<<(N + 255) / 256, 256, 0, s>>>(dim, N, rp, col, rec, bnd, gc, ok);
    return hipGetLastError();
}

hipError_t launch_pack_x0(hipStream_t s, int dim, int N, CD3 x, double q, uint32_t *rec)
{
    if (N <= 0) return hipSuccess;
    k_pack_x0<<<(N + 255) / 256, 256, 0, s>>>(dim, N, x, 1.0 / q, rec);
    return hipGetLastError();
}

hipError_t launch_pack_u(hipStream_t s, const int *stop, int dim, int N, CD3 u, double q, uint32_t *rec)
{
    if (N <= 0) return hipSuccess;
    k_pack_u<<<(N + 255) / 256, 256, 0, s>>>(stop, dim, N, u, 1.0 / q, rec);
    return hipGetLastError();
}

hipError_t launch_min_bond(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col, CD3 x,
                           unsigned long long *out)
{
    if (N <= 0) return hipSuccess;
    k_min_bond<<<(N + 255) / 256, 256, 0, s>>>(dim, N, rp, col, x, out);
    return hipGetLastError();
}

hipError_t launch_pmb_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int n_rows,
                            const int32_t *rp, const int32_t *col, const float2 *bnd, uint32_t *fail,
                            const uint32_t *rec, F3 f, double s0, const MixCtx &mc, double ymin, const GeomCoef *gc)
{
    if (n_rows <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(pmb_g, 3, group, gc, s, stop, order, n_rows, rp, col, bnd, fail, rec, f, s0, mc, ymin)
    } else {
        PD_DISPATCH_G(pmb_g, 2, group, gc, s, stop, order, n_rows, rp, col, bnd, fail, rec, f, s0, mc, ymin)
    }
}

hipError_t launch_lps_theta_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int N,
                                  const int32_t *rp, const int32_t *col, const float2 *bnd, const uint32_t *fail,
                                  uint32_t *rec, const float *minv32, float ctheta, float A, float B,
                                  int skip_broken, float *theta32, const GeomCoef *gc)
{
    if (N <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(theta_g, 3, group, gc, s, stop, order, N, rp, col, bnd, fail, rec, minv32, ctheta, A, B,
                      skip_broken, theta32)
    } else {
        PD_DISPATCH_G(theta_g, 2, group, gc, s, stop, order, N, rp, col, bnd, fail, rec, minv32, ctheta, A, B,
                      skip_broken, theta32)
    }
}

hipError_t launch_lps_force_mixed(hipStream_t s, const int *stop, const int32_t *order, int dim, int group, int N_int,
                                  const int32_t *rp, const int32_t *col, const float2 *bnd, uint32_t *fail,
                                  const uint32_t *rec, double s0, const MixCtx &mc, double ymin, F3 f,
                                  const GeomCoef *gc)
{
    if (N_int <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(lpsf_g, 3, group, gc, s, stop, order, N_int, rp, col, bnd, fail, rec, s0, mc, ymin, f)
    } else {
        PD_DISPATCH_G(lpsf_g, 2, group, gc, s, stop, order, N_int, rp, col, bnd, fail, rec, s0, mc, ymin, f)
    }
}

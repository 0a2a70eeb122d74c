// k_force.hip -- bond-based PMB and state-based LPS force kernels.
//
// Two families:
//  * FP64 parity kernels (thread per particle).  Each thread walks its CSR
//    row in order and accumulates in the operand order of the reference's
//    CPU twins, so forces and break decisions are bit-identical to the FP64
//    restatement: PMB Force_Bond.cu:159-202, LPS 2D Force_State.cu:339-392,
//    LPS 3D the kernel arithmetic of Force_State.cu:218-224,263-281.
//  * Mixed-precision kernels (G lanes per particle, G in {4..64}).  Lanes
//    stride over the row so bond streams (col, xi, coefficient, fail bit) are
//    coalesced; state stays FP64 and is gathered, the bond stretch is formed
//    without cancellation from eta = u_j - u_i,
//        s = eta.(2 xi + eta) / (|xi| (|xi + eta| + |xi|)),
//    forces are summed per lane with Neumaier compensation and combined across
//    the group with error-free TwoSum steps over xor shuffles (wave64 DPP /
//    ds_swizzle), and any bond whose FP32 stretch lies within
//    near_window*s0 of s0 is re-evaluated in FP64 with the parity expression
//    before its break is decided.
#include "pd_internal.h"

namespace {

__device__ __forceinline__ bool bond_intact(const uint32_t *fail, int64_t k)
{
    return (fail[k >> 5] >> (k & 31)) & 1u;
}

__device__ __forceinline__ void bond_break(uint32_t *fail, int64_t k)
{
    atomicAnd(&fail[k >> 5], ~(1u << (k & 31)));
}

// ------------------------------------------------------------ FP64 parity

template <int DIM>
__global__ __launch_bounds__(256) void k_pmb_fp64(const int *__restrict__ stop, int n_rows, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col,
                                                  const double *__restrict__ kappa,
                                                  const double *__restrict__ idist, uint32_t *fail,
                                                  CD3 x, CD3 u, D3 f, double s0, double ymin)
{
    if (stop && *stop) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t ke = rp[i + 1];
    for (int32_t k = rp[i]; k < ke; ++k) {
        if (!bond_intact(fail, k)) continue;
        const int c = col[k];
        // (x_c + u_c) - x_i - u_i, Force_Bond.cu:185-187
        double d0 = x.p[0][c] + u.p[0][c] - x0 - u0;
        double d1 = x.p[1][c] + u.p[1][c] - x1 - u1;
        double r2 = d0 * d0 + d1 * d1;
        double d2 = 0.0;
        if (DIM == 3) {
            d2 = x.p[2][c] + u.p[2][c] - x2 - u2;
            r2 = r2 + d2 * d2;
        }
        double nl = sqrt(r2);
        double id = idist[k];
        double s = (nl - id) / id;
        double ks = kappa[k] * s;
        fx += ks * d0 / nl;
        fy += ks * d1 / nl;
        if (DIM == 3) {
            fz += ks * d2 / nl;
            if (s > s0 && x1 >= ymin && x.p[1][c] >= ymin) bond_break(fail, k);
        } else {
            if (s > s0) bond_break(fail, k);
        }
    }
    f.p[0][i] = fx;
    f.p[1][i] = fy;
    if (DIM == 3) f.p[2][i] = fz;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_lps_theta_fp64(const int *__restrict__ stop, int N, const int32_t *__restrict__ rp,
                                                        const int32_t *__restrict__ col, CD3 x, CD3 u,
                                                        const double *__restrict__ idist,
                                                        const double *__restrict__ fac,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ m,
                                                        const uint32_t *__restrict__ fail, double vol,
                                                        double ctheta, int skip_broken, double *e,
                                                        double *theta)
{
    if (stop && *stop) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    const double a = ctheta / m[i];
    double th = 0.0;
    const int32_t ke = rp[i + 1];
    for (int32_t k = rp[i]; k < ke; ++k) {
        if (skip_broken && !bond_intact(fail, k)) continue;
        const int c = col[k];
        // x_c - x_i + u_c - u_i, Force_State.cu:349-350 / 221-223
        double d0 = x.p[0][c] - x0 + u.p[0][c] - u0;
        double d1 = x.p[1][c] - x1 + u.p[1][c] - u1;
        double r2 = d0 * d0 + d1 * d1;
        if (DIM == 3) {
            double d2 = x.p[2][c] - x2 + u.p[2][c] - u2;
            r2 = r2 + d2 * d2;
        }
        double ek = sqrt(r2) - idist[k];
        e[k] = ek;
        th += a * w[k] * idist[k] * ek * fac[k] * vol;
    }
    theta[i] = th;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_lps_force_fp64(const int *__restrict__ stop, int N_int, const int32_t *__restrict__ rp,
                                                        const int32_t *__restrict__ col, CD3 x, CD3 u,
                                                        const double *__restrict__ idist,
                                                        const double *__restrict__ fac,
                                                        const double *__restrict__ w,
                                                        const double *__restrict__ m,
                                                        const double *__restrict__ theta,
                                                        const double *__restrict__ e, double vol,
                                                        double A, double B, double s0, double ymin,
                                                        uint32_t *fail, D3 f)
{
    if (stop && *stop) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    const double thi = theta[i], mi = m[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t ke = rp[i + 1];
    for (int32_t k = rp[i]; k < ke; ++k) {
        if (!bond_intact(fail, k)) continue;
        const int c = col[k];
        const double id = idist[k], wk = w[k], ek = e[k], thc = theta[c], mc = m[c];
        double d0 = x.p[0][c] + u.p[0][c] - x0 - u0;
        double d1 = x.p[1][c] + u.p[1][c] - x1 - u1;
        if (DIM == 2) {
            // Force_State.cu:372-386
            double nl = sqrt(d0 * d0 + d1 * d1);
            double s = (nl - id) / id;
            double nxv = d0 / nl, nyv = d1 / nl;
            double ti = A * thi * wk * id / mi + B / mi * wk * ek;
            double tj = A * thc * wk * id / mc + B / mc * wk * ek;
            fx += (ti * vol + tj * vol) * nxv * fac[k];
            fy += (ti * vol + tj * vol) * nyv * fac[k];
            if (s > s0) bond_break(fail, k);
        } else {
            // Force_State.cu:270-280
            double dz_len = x.p[2][c] - x2 + u.p[2][c] - u2;
            double nl = sqrt(d0 * d0 + d1 * d1 + dz_len * dz_len);
            double s = (nl - id) / id;
            double ti = A * thi * wk * id / mi + B / mi * wk * (ek - thi * id / 3.0);
            double tj = A * thc * wk * id / mc + B / mc * wk * (ek - thc * id / 3.0);
            double d2 = x.p[2][c] + u.p[2][c] - x2 - u2;
            double tt = (ti * vol + tj * vol) * fac[k];
            fx += tt * d0 / nl;
            fy += tt * d1 / nl;
            fz += tt * d2 / nl;
            if (s > s0 && x1 >= ymin && x.p[1][c] >= ymin) bond_break(fail, k);
        }
    }
    f.p[0][i] = fx;
    f.p[1][i] = fy;
    if (DIM == 3) f.p[2][i] = fz;
}

// ------------------------------------------------------- mixed precision

struct Comp {
    float s, c;
};

__device__ __forceinline__ void neu_add(Comp &a, float v)
{
    // Neumaier's improved Kahan summation.
    float t = a.s + v;
    if (fabsf(a.s) >= fabsf(v))
        a.c += (a.s - t) + v;
    else
        a.c += (v - t) + a.s;
    a.s = t;
}

template <int G>
__device__ __forceinline__ float group_sum(Comp a)
{
    // Error-free TwoSum combination over xor shuffles inside the G-lane group.
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        float os = __shfl_xor(a.s, off, 64);
        float oc = __shfl_xor(a.c, off, 64);
        float t = a.s + os;
        float bb = t - a.s;
        float err = (a.s - (t - bb)) + (os - bb);
        a.c = a.c + oc + err;
        a.s = t;
    }
    return a.s + a.c;
}

template <int DIM, int G>
__global__ __launch_bounds__(256) void k_pmb_mixed(const int *__restrict__ stop, int n_rows, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col,
                                                   const float *__restrict__ xi32,
                                                   const float *__restrict__ kap32, uint32_t *fail,
                                                   CD3 x, CD3 u, F3 f, double s0, float s0f, float win,
                                                   double ymin)
{
    if (stop && *stop) return;
    const int gid = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    const int lane = threadIdx.x & (G - 1);
    if (gid >= n_rows) return;  // whole group leaves together
    const int i = gid;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    Comp ax{0.f, 0.f}, ay{0.f, 0.f}, az{0.f, 0.f};
    const int32_t kb = rp[i], ke = rp[i + 1];
    for (int32_t k = kb + lane; k < ke; k += G) {
        if (!bond_intact(fail, k)) continue;
        const int c = col[k];
        const double xc0 = x.p[0][c], xc1 = x.p[1][c], xc2 = DIM == 3 ? x.p[2][c] : 0.0;
        const double uc0 = u.p[0][c], uc1 = u.p[1][c], uc2 = DIM == 3 ? u.p[2][c] : 0.0;
        const float ex = (float)(xc0 - x0), ey = (float)(xc1 - x1), ez = (float)(xc2 - x2);
        const float hx = (float)(uc0 - u0), hy = (float)(uc1 - u1), hz = (float)(uc2 - u2);
        const float L0 = xi32[k];
        const float dx = ex + hx, dy = ey + hy, dz = ez + hz;
        float L2 = dx * dx + dy * dy;
        float num = hx * (2.f * ex + hx) + hy * (2.f * ey + hy);
        if (DIM == 3) {
            L2 += dz * dz;
            num += hz * (2.f * ez + hz);
        }
        const float L = sqrtf(L2);
        const float s = num / (L0 * (L + L0));
        bool brk;
        if (fabsf(s - s0f) <= win) {
            // FP64 re-evaluation, Force_Bond.cu:185-189 and Surf_Corr.cu:324-326
            double d0 = xc0 + uc0 - x0 - u0, d1 = xc1 + uc1 - x1 - u1;
            double a0 = xc0 - x0, a1 = xc1 - x1;
            double r2 = d0 * d0 + d1 * d1, q2 = a0 * a0 + a1 * a1;
            if (DIM == 3) {
                double d2 = xc2 + uc2 - x2 - u2, a2 = xc2 - x2;
                r2 = r2 + d2 * d2;
                q2 = q2 + a2 * a2;
            }
            double id = sqrt(q2);
            brk = (sqrt(r2) - id) / id > s0;
        } else {
            brk = s > s0f;
        }
        if (DIM == 3) brk = brk && x1 >= ymin && xc1 >= ymin;
        const float t = kap32[k] * s / L;
        neu_add(ax, t * dx);
        neu_add(ay, t * dy);
        if (DIM == 3) neu_add(az, t * dz);
        if (brk) bond_break(fail, k);
    }
    float sx = group_sum<G>(ax), sy = group_sum<G>(ay);
    float sz = DIM == 3 ? group_sum<G>(az) : 0.f;
    if (lane == 0) {
        f.p[0][i] = sx;
        f.p[1][i] = sy;
        if (DIM == 3) f.p[2][i] = sz;
    }
}

template <int DIM, int G>
__global__ __launch_bounds__(256) void k_lps_theta_mixed(const int *__restrict__ stop, int N, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const float *__restrict__ xi32,
                                                         const float *__restrict__ om32,
                                                         const uint32_t *__restrict__ fail, CD3 x, CD3 u,
                                                         const float *__restrict__ minv32, float ctheta,
                                                         int skip_broken, float *theta32)
{
    if (stop && *stop) return;
    const int gid = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    const int lane = threadIdx.x & (G - 1);
    if (gid >= N) return;
    const int i = gid;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    Comp acc{0.f, 0.f};
    const int32_t kb = rp[i], ke = rp[i + 1];
    for (int32_t k = kb + lane; k < ke; k += G) {
        if (skip_broken && !bond_intact(fail, k)) continue;
        const int c = col[k];
        const float ex = (float)(x.p[0][c] - x0), ey = (float)(x.p[1][c] - x1);
        const float hx = (float)(u.p[0][c] - u0), hy = (float)(u.p[1][c] - u1);
        float ez = 0.f, hz = 0.f;
        if (DIM == 3) {
            ez = (float)(x.p[2][c] - x2);
            hz = (float)(u.p[2][c] - u2);
        }
        const float L0 = xi32[k];
        const float dx = ex + hx, dy = ey + hy, dz = ez + hz;
        float L2 = dx * dx + dy * dy, num = hx * (2.f * ex + hx) + hy * (2.f * ey + hy);
        if (DIM == 3) {
            L2 += dz * dz;
            num += hz * (2.f * ez + hz);
        }
        const float ek = num / (sqrtf(L2) + L0);  // |xi+eta| - |xi|
        neu_add(acc, om32[k] * L0 * ek);
    }
    float th = group_sum<G>(acc);
    if (lane == 0) theta32[i] = ctheta * minv32[i] * th;
}

template <int DIM, int G>
__global__ __launch_bounds__(256) void k_lps_force_mixed(const int *__restrict__ stop, int N_int, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ col,
                                                         const float *__restrict__ xi32,
                                                         const float *__restrict__ om32, uint32_t *fail,
                                                         CD3 x, CD3 u, const float *__restrict__ minv32,
                                                         const float *__restrict__ theta32, float A,
                                                         float B, double s0, float s0f, float win,
                                                         double ymin, F3 f)
{
    if (stop && *stop) return;
    const int gid = (int)((blockIdx.x * 256u + threadIdx.x) / G);
    const int lane = threadIdx.x & (G - 1);
    if (gid >= N_int) return;
    const int i = gid;
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    const float ri = minv32[i], qi = theta32[i] * ri;
    Comp ax{0.f, 0.f}, ay{0.f, 0.f}, az{0.f, 0.f};
    const int32_t kb = rp[i], ke = rp[i + 1];
    for (int32_t k = kb + lane; k < ke; k += G) {
        if (!bond_intact(fail, k)) continue;
        const int c = col[k];
        const double xc0 = x.p[0][c], xc1 = x.p[1][c], xc2 = DIM == 3 ? x.p[2][c] : 0.0;
        const double uc0 = u.p[0][c], uc1 = u.p[1][c], uc2 = DIM == 3 ? u.p[2][c] : 0.0;
        const float ex = (float)(xc0 - x0), ey = (float)(xc1 - x1), ez = (float)(xc2 - x2);
        const float hx = (float)(uc0 - u0), hy = (float)(uc1 - u1), hz = (float)(uc2 - u2);
        const float L0 = xi32[k];
        const float dx = ex + hx, dy = ey + hy, dz = ez + hz;
        float L2 = dx * dx + dy * dy, num = hx * (2.f * ex + hx) + hy * (2.f * ey + hy);
        if (DIM == 3) {
            L2 += dz * dz;
            num += hz * (2.f * ez + hz);
        }
        const float L = sqrtf(L2);
        const float ek = num / (L + L0);
        const float s = ek / L0;
        const float rc = minv32[c];
        const float qc = theta32[c] * rc;
        // (t_i + t_j) vol fac = omega [A xi (q_i + q_j) + B e (r_i + r_j)]
        const float tt = om32[k] * (A * L0 * (qi + qc) + B * ek * (ri + rc)) / L;
        neu_add(ax, tt * dx);
        neu_add(ay, tt * dy);
        if (DIM == 3) neu_add(az, tt * dz);
        bool brk;
        if (fabsf(s - s0f) <= win) {
            // FP64 re-evaluation with the parity expressions
            double d0 = xc0 + uc0 - x0 - u0, d1 = xc1 + uc1 - x1 - u1;
            double a0 = xc0 - x0, a1 = xc1 - x1;
            double r2 = d0 * d0 + d1 * d1, q2 = a0 * a0 + a1 * a1;
            if (DIM == 3) {
                double d2 = xc2 - x2 + uc2 - u2, a2 = xc2 - x2;
                r2 = r2 + d2 * d2;
                q2 = q2 + a2 * a2;
            }
            double id = sqrt(q2);
            brk = (sqrt(r2) - id) / id > s0;
        } else {
            brk = s > s0f;
        }
        if (DIM == 3) brk = brk && x1 >= ymin && xc1 >= ymin;
        if (brk) bond_break(fail, k);
    }
    float sx = group_sum<G>(ax), sy = group_sum<G>(ay);
    float sz = DIM == 3 ? group_sum<G>(az) : 0.f;
    if (lane == 0) {
        f.p[0][i] = sx;
        f.p[1][i] = sy;
        if (DIM == 3) f.p[2][i] = sz;
    }
}

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

template <int DIM, int G>
hipError_t pmb_mixed_g(hipStream_t s, const int *stop, int n_rows, const int32_t *rp, const int32_t *col, const float *xi32,
                       const float *kap32, uint32_t *fail, CD3 x, CD3 u, F3 f, double s0, double window,
                       double ymin)
{
    int64_t threads = (int64_t)n_rows * G;
    k_pmb_mixed<DIM, G><<<nblk(threads, 256), 256, 0, s>>>(stop, n_rows, rp, col, xi32, kap32, fail, x, u, f, s0,
                                                           (float)s0, (float)(window * s0), ymin);
    return hipGetLastError();
}

template <int DIM, int G>
hipError_t theta_mixed_g(hipStream_t s, const int *stop, int N, const int32_t *rp, const int32_t *col, const float *xi32,
                         const float *om32, const uint32_t *fail, CD3 x, CD3 u, const float *m32,
                         float ctheta, int skip, float *theta32)
{
    int64_t threads = (int64_t)N * G;
    k_lps_theta_mixed<DIM, G><<<nblk(threads, 256), 256, 0, s>>>(stop, N, rp, col, xi32, om32, fail, x, u, m32,
                                                                 ctheta, skip, theta32);
    return hipGetLastError();
}

template <int DIM, int G>
hipError_t lps_force_mixed_g(hipStream_t s, const int *stop, int N_int, const int32_t *rp, const int32_t *col,
                             const float *xi32, const float *om32, uint32_t *fail, CD3 x, CD3 u,
                             const float *m32, const float *theta32, float A, float B, double s0,
                             double window, double ymin, F3 f)
{
    int64_t threads = (int64_t)N_int * G;
    k_lps_force_mixed<DIM, G><<<nblk(threads, 256), 256, 0, s>>>(stop, N_int, rp, col, xi32, om32, fail, x, u, m32,
                                                                 theta32, A, B, s0, (float)s0,
                                                                 (float)(window * s0), ymin, f);
    return hipGetLastError();
}

#define PD_DISPATCH_G(FN, DIM, G, ...)                       \
    switch (G) {                                             \
    case 4: return FN<DIM, 4>(__VA_ARGS__);                  \
    case 8: return FN<DIM, 8>(__VA_ARGS__);                  \
    case 16: return FN<DIM, 16>(__VA_ARGS__);                \
    case 32: return FN<DIM, 32>(__VA_ARGS__);                \
    case 64: return FN<DIM, 64>(__VA_ARGS__);                \
    default: return hipErrorInvalidValue;                    \
    }

} // namespace

hipError_t launch_pmb_fp64(hipStream_t s, const int *stop, int dim, int n_rows, const int32_t *rp, const int32_t *col,
                           const double *kappa, const double *idist, uint32_t *fail, CD3 x, CD3 u,
                           D3 f, double s0, double ymin)
{
    if (n_rows <= 0) return hipSuccess;
    if (dim == 3)
        k_pmb_fp64<3><<<nblk(n_rows, 256), 256, 0, s>>>(stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
    else
        k_pmb_fp64<2><<<nblk(n_rows, 256), 256, 0, s>>>(stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
    return hipGetLastError();
}

hipError_t launch_pmb_mixed(hipStream_t s, const int *stop, int dim, int group, int n_rows, const int32_t *rp,
                            const int32_t *col, const float *xi32, const float *kap32, uint32_t *fail,
                            CD3 x, CD3 u, F3 f, double s0, double window, double ymin)
{
    if (n_rows <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(pmb_mixed_g, 3, group, s, stop, n_rows, rp, col, xi32, kap32, fail, x, u, f, s0, window, ymin)
    } else {
        PD_DISPATCH_G(pmb_mixed_g, 2, group, s, stop, n_rows, rp, col, xi32, kap32, fail, x, u, f, s0, window, ymin)
    }
}

hipError_t launch_lps_theta_fp64(hipStream_t s, const int *stop, int dim, int N, const int32_t *rp, const int32_t *col,
                                 CD3 x, CD3 u, const double *idist, const double *fac, const double *w,
                                 const double *m, const uint32_t *fail, double vol, double pratio,
                                 int skip_broken, double *e, double *theta)
{
    if (N <= 0) return hipSuccess;
    // 2.0*(2nu-1)/(nu-1) (Force_State.cu:351) or 3.0 (Force_State.cu:224)
    const double ctheta = dim == 2 ? 2.0 * (2.0 * pratio - 1.0) / (pratio - 1.0) : 3.0;
    if (dim == 3)
        k_lps_theta_fp64<3><<<nblk(N, 256), 256, 0, s>>>(stop, N, rp, col, x, u, idist, fac, w, m, fail, vol, ctheta,
                                                         skip_broken, e, theta);
    else
        k_lps_theta_fp64<2><<<nblk(N, 256), 256, 0, s>>>(stop, N, rp, col, x, u, idist, fac, w, m, fail, vol, ctheta,
                                                         skip_broken, e, theta);
    return hipGetLastError();
}

hipError_t launch_lps_force_fp64(hipStream_t s, const int *stop, int dim, int N_int, const int32_t *rp, const int32_t *col,
                                 CD3 x, CD3 u, const double *idist, const double *fac, const double *w,
                                 const double *m, const double *theta, const double *e, double vol,
                                 double pratio, double K, double G, double s0, double ymin,
                                 uint32_t *fail, D3 f)
{
    if (N_int <= 0) return hipSuccess;
    if (dim == 3) {
        // 3K and 15G, Force_State.cu:274
        k_lps_force_fp64<3><<<nblk(N_int, 256), 256, 0, s>>>(stop, N_int, rp, col, x, u, idist, fac, w, m, theta, e,
                                                             vol, 3.0 * K, 15.0 * G, s0, ymin, fail, f);
    } else {
        // (2nu-1)/(nu-1)*(2K-32G/9) and 8G, Force_State.cu:378-379
        double A = (2.0 * pratio - 1.0) / (pratio - 1.0) * (2.0 * K - 32.0 * G / 9.0);
        k_lps_force_fp64<2><<<nblk(N_int, 256), 256, 0, s>>>(stop, N_int, rp, col, x, u, idist, fac, w, m, theta, e,
                                                             vol, A, 8.0 * G, s0, ymin, fail, f);
    }
    return hipGetLastError();
}

hipError_t launch_lps_theta_mixed(hipStream_t s, const int *stop, int dim, int group, int N, const int32_t *rp,
                                  const int32_t *col, const float *xi32, const float *om32,
                                  const uint32_t *fail, CD3 x, CD3 u, const float *m32, float ctheta,
                                  int skip_broken, float *theta32)
{
    if (N <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(theta_mixed_g, 3, group, s, stop, N, rp, col, xi32, om32, fail, x, u, m32, ctheta, skip_broken,
                      theta32)
    } else {
        PD_DISPATCH_G(theta_mixed_g, 2, group, s, stop, N, rp, col, xi32, om32, fail, x, u, m32, ctheta, skip_broken,
                      theta32)
    }
}

hipError_t launch_lps_force_mixed(hipStream_t s, const int *stop, int dim, int group, int N_int, const int32_t *rp,
                                  const int32_t *col, const float *xi32, const float *om32, uint32_t *fail,
                                  CD3 x, CD3 u, const float *m32, const float *theta32, float A, float B,
                                  double s0, double window, double ymin, F3 f)
{
    if (N_int <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_G(lps_force_mixed_g, 3, group, s, stop, N_int, rp, col, xi32, om32, fail, x, u, m32, theta32, A, B,
                      s0, window, ymin, f)
    } else {
        PD_DISPATCH_G(lps_force_mixed_g, 2, group, s, stop, N_int, rp, col, xi32, om32, fail, x, u, m32, theta32, A, B,
                      s0, window, ymin, f)
    }
}

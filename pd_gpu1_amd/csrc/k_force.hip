// k_force.hip -- FP64 parity force kernels (bond-based PMB, state-based LPS).
//
// Thread per particle.  Each thread walks its CSR row in order and
// accumulates in the operand order of the reference's CPU twins, so forces
// and break decisions are bit-identical to the FP64 restatement: PMB
// Force_Bond.cu:159-202, LPS 2D Force_State.cu:339-392, LPS 3D the kernel
// arithmetic of Force_State.cu:218-224,263-281.  The mixed-precision
// throughput kernels are in k_force_mixed.hip.
#include "pd_internal.h"
#include "lanes.h"

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

// The same force with G lanes per row: the lanes load G consecutive bonds
// (coalesced col / kappa / idist), compute their terms with the arithmetic
// above, and the group adds the terms in bond order through shuffles, so the
// sums are the sequential loop's bit for bit.
template <int DIM, int G>
__global__ __launch_bounds__(256) void k_pmb_fp64_group(const int *__restrict__ stop, int n_rows,
                                                        const int32_t *__restrict__ rp,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ kappa,
                                                        const double *__restrict__ idist, uint32_t *fail, CD3 x,
                                                        CD3 u, D3 f, double s0, double ymin)
{
    if (stop && *stop) return;
    // Workgroups reach the 8 XCDs round-robin.  In 2D they are renumbered so
    // that XCD k takes the k-th eighth of the rows (gridDim.x is a multiple
    // of 8): a band of the plate, whose neighbours then sit in that XCD's L2
    // instead of in all eight (an ADR iteration on a 315 x 315 plate 57 ->
    // 52 us).  The 3D Kalthoff-Winkler run measured 12% slower that way, so
    // 3D keeps the launch order.
    const int b = DIM == 2 ? (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3)
                           : (int)blockIdx.x;
    const int i = (int)(((int64_t)b * blockDim.x + threadIdx.x) / G);
    const int lane = threadIdx.x & (G - 1);
    if (i >= n_rows) return;  // the whole group leaves together
    const double x0 = x.p[0][i], x1 = x.p[1][i], x2 = DIM == 3 ? x.p[2][i] : 0.0;
    const double u0 = u.p[0][i], u1 = u.p[1][i], u2 = DIM == 3 ? u.p[2][i] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int32_t kb = rp[i], ke = rp[i + 1];
    for (int32_t k0 = kb; k0 < ke; k0 += G) {
        const int32_t k = k0 + lane;
        double tx = 0.0, ty = 0.0, tz = 0.0;
        int ok = 0;
        if (k < ke && bond_intact(fail, k)) {
            ok = 1;
            const int c = col[k];
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
            tx = ks * d0 / nl;
            ty = ks * d1 / nl;
            if (DIM == 3) {
                tz = ks * d2 / nl;
                if (s > s0 && x1 >= ymin && x.p[1][c] >= ymin) bond_break(fail, k);
            } else {
                if (s > s0) bond_break(fail, k);
            }
        }
        const int cnt = ke - k0 < G ? ke - k0 : G;
        if constexpr (G == 4 || G == 8) {
            // the terms broadcast over DPP (lanes.h) rather than __shfl
            lanes::Unroll<G>::run([&](auto lc) {
                constexpr int L = decltype(lc)::value;
                const double sx = lanes::bcast<G, L>(tx), sy = lanes::bcast<G, L>(ty);
                const double sz = DIM == 3 ? lanes::bcast<G, L>(tz) : 0.0;
                if (L < cnt && lanes::bcast<G, L>(ok)) {
                    fx += sx;
                    fy += sy;
                    if (DIM == 3) fz += sz;
                }
            });
        } else {
            for (int l = 0; l < cnt; ++l) {
                const double sx = __shfl(tx, l, G), sy = __shfl(ty, l, G);
                const double sz = DIM == 3 ? __shfl(tz, l, G) : 0.0;
                if (__shfl(ok, l, G)) {
                    fx += sx;
                    fy += sy;
                    if (DIM == 3) fz += sz;
                }
            }
        }
    }
    if (lane == 0) {
        f.p[0][i] = fx;
        f.p[1][i] = fy;
        if (DIM == 3) f.p[2][i] = fz;
    }
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

// at least one block: every kernel bounds-checks its index, so an empty
// range (say N_int = 0) is a no-op launch rather than an invalid grid
inline int nblk(int64_t n, int t) { return n > 0 ? (int)((n + t - 1) / t) : 1; }

} // namespace

namespace {
template <int DIM, int G>
void pmb_fp64_g(hipStream_t s, const int *stop, int n_rows, const int32_t *rp, const int32_t *col,
                const double *kappa, const double *idist, uint32_t *fail, CD3 x, CD3 u, D3 f, double s0, double ymin)
{
    const int64_t threads = (int64_t)n_rows * G;
    const unsigned grid = (unsigned)((threads + 255) / 256 + 7) / 8 * 8;  // a multiple of 8 (XCD renumbering)
    k_pmb_fp64_group<DIM, G><<<grid, 256, 0, s>>>(stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
}
} // namespace

hipError_t launch_pmb_fp64(hipStream_t s, const int *stop, int dim, int n_rows, const int32_t *rp, const int32_t *col,
                           const double *kappa, const double *idist, uint32_t *fail, CD3 x, CD3 u,
                           D3 f, double s0, double ymin, int group)
{
    if (n_rows <= 0) return hipSuccess;
    if (group == 4 || group == 8 || group == 16 || group == 32) {
        if (dim == 3) {
            if (group == 4) pmb_fp64_g<3, 4>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 8) pmb_fp64_g<3, 8>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 16) pmb_fp64_g<3, 16>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 32) pmb_fp64_g<3, 32>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
        } else {
            if (group == 4) pmb_fp64_g<2, 4>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 8) pmb_fp64_g<2, 8>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 16) pmb_fp64_g<2, 16>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
            if (group == 32) pmb_fp64_g<2, 32>(s, stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
        }
        return hipGetLastError();
    }
    if (dim == 3)
        k_pmb_fp64<3><<<nblk(n_rows, 256), 256, 0, s>>>(stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
    else
        k_pmb_fp64<2><<<nblk(n_rows, 256), 256, 0, s>>>(stop, n_rows, rp, col, kappa, idist, fail, x, u, f, s0, ymin);
    return hipGetLastError();
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


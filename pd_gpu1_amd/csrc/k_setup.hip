// k_setup.hip -- per-bond setup kernels: volume correction, surface
// correction, cracks, LPS weights, packed bond coefficients and damage.
//
// These run once per problem (damage once per checkpoint).  Each is thread-
// per-particle over its CSR row so that every per-row sum visits bonds in row
// order, the order of the reference's CPU twins; expressions keep the
// reference's operand order and the library is built with -ffp-contract=off.
#include "pd_internal.h"

namespace {

__global__ void k_vol_corr(int dim, int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                           CD3 x, double delta, double dx, double *idist, double *fac)
{
    // Surf_Corr.cu:312-337 (kernel twin Surf_Corr.cu:7-35)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double lo = delta - dx / 2.0, hi = delta + dx / 2.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        int c = col[k];
        double ddx = x.p[0][c] - x.p[0][i], ddy = x.p[1][c] - x.p[1][i];
        double r2 = ddx * ddx + ddy * ddy;
        if (dim == 3) {
            double ddz = x.p[2][c] - x.p[2][i];
            r2 = r2 + ddz * ddz;
        }
        double id = sqrt(r2);
        idist[k] = id;
        if (id <= lo)
            fac[k] = 1.0;
        else if (id <= hi)
            fac[k] = (hi - id) / dx;
        else
            fac[k] = 0.0;
    }
}

__global__ void k_surface_probe(int dim, int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                CD3 x, int axis, const double *__restrict__ idist,
                                const double *__restrict__ fac, double bc, double vol, double sedload,
                                double *fncst)
{
    // Surf_Corr.cu:353-388 with the probe field of Disp (Surf_Corr.cu:339-351):
    // u[axis] = 0.001*x[axis], the other components zero.
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double ui[3] = {0.0, 0.0, 0.0};
    ui[axis] = 0.001 * x.p[axis][i];
    double st = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        int c = col[k];
        double uc[3] = {0.0, 0.0, 0.0};
        uc[axis] = 0.001 * x.p[axis][c];
        double d0 = x.p[0][c] + uc[0] - x.p[0][i] - ui[0];
        double d1 = x.p[1][c] + uc[1] - x.p[1][i] - ui[1];
        double r2 = d0 * d0 + d1 * d1;
        if (dim == 3) {
            double d2 = x.p[2][c] + uc[2] - x.p[2][i] - ui[2];
            r2 = r2 + d2 * d2;
        }
        double nl = sqrt(r2);
        st += 0.25 * bc * (nl - idist[k]) * (nl - idist[k]) / idist[k] * vol * fac[k];
    }
    fncst[i] = sedload / st;
}

__global__ void k_surface_scr(int dim, int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                              CD3 x, const double *__restrict__ idist, const double *__restrict__ fncst,
                              double *scr)
{
    // Surf_Corr.cu:390-463 in trig-free form: with theta = atan(|dy|/|dx|),
    // cos^2 theta = dx^2/(dx^2+dy^2), sin^2 theta = dy^2/(dx^2+dy^2); with
    // phi = acos(|dz|/xi), cos^2 phi = (|dz|/xi)^2.  Axis-aligned branches
    // keep the reference's 1e-10 tests.
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double *fx = fncst, *fy = fncst + N, *fz = fncst + 2 * (size_t)N;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        int c = col[k];
        double adx = fabs(x.p[0][c] - x.p[0][i]), ady = fabs(x.p[1][c] - x.p[1][i]);
        double c2, s2;
        if (dim == 2 || fabs(x.p[2][c] - x.p[2][i]) <= 1.0e-10) {
            if (ady <= 1.0e-10) {
                c2 = 1.0; s2 = 0.0;
            } else if (adx <= 1.0e-10) {
                c2 = 0.0; s2 = 1.0;
            } else {
                double a2 = adx * adx, b2 = ady * ady;
                c2 = a2 / (a2 + b2);
                s2 = b2 / (a2 + b2);
            }
            double scx = (fx[i] + fx[c]) / 2.0;
            double scy = (fy[i] + fy[c]) / 2.0;
            if (dim == 2) {
                scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy)));
            } else {
                double scz = (fz[i] + fz[c]) / 2.0;
                scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy) + 0.0 / (scz * scz)));
            }
        } else if (adx <= 1.0e-10 && ady <= 1.0e-10) {
            scr[k] = (fz[i] + fz[c]) / 2.0;
        } else {
            double a2 = adx * adx, b2 = ady * ady;
            c2 = a2 / (a2 + b2);
            s2 = b2 / (a2 + b2);
            double t = fabs(x.p[2][c] - x.p[2][i]) / idist[k];
            double p2 = t * t;
            double scx = (fx[i] + fx[c]) / 2.0;
            double scy = (fy[i] + fy[c]) / 2.0;
            double scz = (fz[i] + fz[c]) / 2.0;
            scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy) + p2 / (scz * scz)));
        }
    }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

__device__ __forceinline__ void clear_bit(uint32_t *fail, int64_t k)
{
    atomicAnd(&fail[k >> 5], ~(1u << (k & 31)));
}

__global__ void k_crack_2d(int N_int, const int32_t *__restrict__ rp, const int32_t *__restrict__ col, CD3 x,
                           double clength, double loc_x, double loc_y, int vertical, double t,
                           double sq1, double half_s, uint32_t *fail)
{
    // Base_Function.cu:83-123.  tan(theta), sqrt(1+tan^2) and
    // clength*sin(theta)/2 are host-evaluated scalars, so the per-bond work
    // uses only + - * / sqrt.
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    const double *X = x.p[0], *Y = x.p[1];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        int c = col[k];
        if (vertical) {
            if (fabs(Y[c] - loc_y) < half_s && fabs(Y[i] - loc_y) < half_s)
                if ((X[c] - loc_x) * (X[i] - loc_x) < 0.0) clear_bit(fail, k);
        } else {
            double a0 = Y[c] - loc_y, a1 = X[c] - loc_x;
            double dist_c = a0 * a0 + a1 * a1;
            double b0 = Y[i] - loc_y, b1 = X[i] - loc_x;
            double dist_i = b0 * b0 + b1 * b1;
            double vc = (t * X[c] - Y[c] - loc_x * t + loc_y) / sq1;
            double vi = (t * X[i] - Y[i] - loc_x * t + loc_y) / sq1;
            double len_c = sqrt(dist_c - vc * vc);
            double len_i = sqrt(dist_i - vi * vi);
            if (len_c < clength / 2.0 && len_i < clength / 2.0) {
                double bcn = t * X[c] - loc_x * t + loc_y - Y[c];
                double bi = t * X[i] - loc_x * t + loc_y - Y[i];
                if (bcn * bi <= 0) clear_bit(fail, k);
            }
        }
    }
}

__global__ void k_crack_kw(int N_int, const int32_t *__restrict__ rp, const int32_t *__restrict__ col, CD3 x,
                           double rad, uint32_t *fail)
{
    // Solve_Kalthoff_Winkler.cu:65-87
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    const double *X = x.p[0], *Y = x.p[1];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        int c = col[k];
        bool brk = false;
        if ((X[c] > -rad && X[i] < -rad) | (X[i] > -rad && X[c] < -rad))
            if ((Y[c] >= -1.0e-10) | (Y[i] >= -1.0e-10)) brk = true;
        if ((X[c] > rad && X[i] < rad) | (X[i] > rad && X[c] < rad))
            if ((Y[c] >= -1.0e-10) | (Y[i] >= -1.0e-10)) brk = true;
        if (brk) clear_bit(fail, k);
    }
}

__global__ void k_lps_weights(int N, const int32_t *__restrict__ rp, const double *__restrict__ idist,
                              const double *__restrict__ fac, double delta, double vol, int w_given, double *w,
                              double *m)
{
    // Force_State.cu:309-321
    // w_given: w[] already holds the influence weights (computed on the host
    // with the C library's exp, so they equal the FP64 CPU algorithm's bit for
    // bit); only m is formed here, in the same order.
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double mi = 0.0;
    const double d2 = delta * delta;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        double id = idist[k];
        double wk;
        if (w_given) {
            wk = w[k];
        } else {
            wk = exp(-(id * id) / d2);
            w[k] = wk;
        }
        mi += wk * (id * id) * vol * fac[k];
    }
    m[i] = mi;
}

__global__ void k_pmb_coef(int64_t nb, const double *__restrict__ fac, const double *__restrict__ scr,
                           double vol, double bc, const double *__restrict__ idist, double iq, double *kappa,
                           float2 *bnd)
{
    // kappa = vol*fac*bc*scr: the left prefix of Force_Bond.cu:190, so
    // kappa*s*d/nl reproduces the reference product bit for bit.
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    double kap = vol * fac[k] * bc * scr[k];
    kappa[k] = kap;
    bnd[k] = make_float2((float)(idist[k] * iq), (float)kap);  // |xi| in units of the record quantum
}

__global__ void k_lps_coef(int64_t nb, const double *__restrict__ w, const double *__restrict__ fac,
                           double vol, const double *__restrict__ idist, double iq, float2 *bnd)
{
    // omega = w*fac*vol per bond, the bond factor shared by the dilatation
    // and force sums of Force_State.cu:224,274-279.
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    bnd[k] = make_float2((float)(idist[k] * iq), (float)(w[k] * fac[k] * vol));
}

__global__ void k_inv_to_f32(const double *a, float *b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)(1.0 / a[i]);
}

__global__ void k_damage(int N_int, const int32_t *__restrict__ rp, const uint32_t *__restrict__ fail,
                         const double *__restrict__ fac, double vol, double *dmg)
{
    // Base_Function.cu:220-238
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    double p1 = 0.0, p2 = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        double f = (double)((fail[k >> 5] >> (k & 31)) & 1u);
        p1 += f * vol * fac[k];
        p2 += vol * fac[k];
    }
    dmg[i] = 1.0 - p1 / p2;
}

__global__ void k_copy16(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (int64_t)gridDim.x * blockDim.x) b[t] = a[t];
}

// at least one block: every kernel bounds-checks its index, so an empty
// range (say N_int = 0) is a no-op launch rather than an invalid grid
inline int nblk(int64_t n, int t) { return n > 0 ? (int)((n + t - 1) / t) : 1; }

} // namespace

hipError_t launch_copy16(hipStream_t s, const void *a, void *b, int64_t bytes)
{
    int64_t n = bytes / 16;
    k_copy16<<<8192, 256, 0, s>>>((const double2 *)a, (double2 *)b, n);
    return hipGetLastError();
}

hipError_t launch_vol_corr(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col, CD3 x,
                           double delta, double dx, double *idist, double *fac)
{
    k_vol_corr<<<nblk(N, 256), 256, 0, s>>>(dim, N, rp, col, x, delta, dx, idist, fac);
    return hipGetLastError();
}

hipError_t launch_surface_probe(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                                CD3 x, int axis, const double *idist, const double *fac, double bc,
                                double vol, double sedload, double *fncst)
{
    k_surface_probe<<<nblk(N, 256), 256, 0, s>>>(dim, N, rp, col, x, axis, idist, fac, bc, vol, sedload, fncst);
    return hipGetLastError();
}

hipError_t launch_surface_scr(hipStream_t s, int dim, int N, const int32_t *rp, const int32_t *col,
                              CD3 x, const double *idist, const double *fncst, double *scr)
{
    k_surface_scr<<<nblk(N, 256), 256, 0, s>>>(dim, N, rp, col, x, idist, fncst, scr);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, int64_t n, uint32_t v)
{
    if (n <= 0) return hipSuccess;
    int b = nblk(n, 256);
    k_fill_u32<<<b > 65536 ? 65536 : b, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(hipStream_t s, double *p, int64_t n, double v)
{
    if (n <= 0) return hipSuccess;
    int b = nblk(n, 256);
    k_fill_f64<<<b > 65536 ? 65536 : b, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_crack_2d(hipStream_t s, int N_int, const int32_t *rp, const int32_t *col, CD3 x,
                           double clength, double loc_x, double loc_y, int vertical, double t,
                           double sq1, double half_s, uint32_t *fail)
{
    k_crack_2d<<<nblk(N_int, 256), 256, 0, s>>>(N_int, rp, col, x, clength, loc_x, loc_y, vertical, t, sq1,
                                                half_s, fail);
    return hipGetLastError();
}

hipError_t launch_crack_kw(hipStream_t s, int N_int, const int32_t *rp, const int32_t *col, CD3 x,
                           double rad, uint32_t *fail)
{
    k_crack_kw<<<nblk(N_int, 256), 256, 0, s>>>(N_int, rp, col, x, rad, fail);
    return hipGetLastError();
}

hipError_t launch_lps_weights(hipStream_t s, int N, const int32_t *rp, const double *idist,
                              const double *fac, double delta, double vol, int w_given, double *w, double *m)
{
    if (N <= 0) return hipSuccess;
    k_lps_weights<<<nblk(N, 256), 256, 0, s>>>(N, rp, idist, fac, delta, vol, w_given, w, m);
    return hipGetLastError();
}

hipError_t launch_pmb_coef(hipStream_t s, int64_t nb, const double *fac, const double *scr, double vol,
                           double bc, const double *idist, double iq, double *kappa, float2 *bnd)
{
    if (nb <= 0) return hipSuccess;
    k_pmb_coef<<<nblk(nb, 256), 256, 0, s>>>(nb, fac, scr, vol, bc, idist, iq, kappa, bnd);
    return hipGetLastError();
}

hipError_t launch_lps_coef(hipStream_t s, int64_t nb, const double *w, const double *fac, double vol,
                           const double *idist, double iq, float2 *bnd, int N, const double *m, float *m32)
{
    if (nb > 0) k_lps_coef<<<nblk(nb, 256), 256, 0, s>>>(nb, w, fac, vol, idist, iq, bnd);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_inv_to_f32<<<nblk(N, 256), 256, 0, s>>>(m, m32, N);
    return hipGetLastError();
}

hipError_t launch_damage(hipStream_t s, int N_int, const int32_t *rp, const uint32_t *fail,
                         const double *fac, double vol, double *dmg)
{
    k_damage<<<nblk(N_int, 256), 256, 0, s>>>(N_int, rp, fail, fac, vol, dmg);
    return hipGetLastError();
}

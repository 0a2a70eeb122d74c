// k_integrate.hip -- velocity Verlet, adaptive dynamic relaxation (ADR) with
// a fused on-device damping reduction, the Kalthoff-Winkler cylinder contact,
// and the fixed-order reduction they share.
//
// Fixed reduction order (the "canonical tree", lane count T = blocks*threads,
// a power of two): lane t sums elements t, t+T, t+2T, ... in sequence; the T
// partials are then combined pairwise at offsets 1, 2, 4, ... (partial[t] +=
// partial[t+off] for t % (2*off) == 0) -- xor/down shuffles inside a wave64,
// LDS across waves, a one-workgroup finishing launch across blocks.  The result is
// deterministic and independent of scheduling; the FP64 oracle emulates the
// same order, so ADR damping coefficients and trajectories match it bit for
// bit.  This replaces reduce_cp + sum() (Reduce.cu:11-65), which copies each
// sum to the host with a blocking cudaMemcpy, twice per ADR iteration
// (Integrate.cu:112-113).
#include "pd_internal.h"

namespace {

// Mixed-precision force kernels gather u, rounded to FP32 and in units of
// the record quantum, from words 4..6 of a particle record (kRecW words).
// Kernels that move particles refresh those words in the same pass when
// given the record array.
__device__ __forceinline__ void store_rec_u(const RecU &ru, int i, double ux, double uy, double uz)
{
    uint32_t *r = ru.rec + (size_t)i * kRecW + 4;
    *reinterpret_cast<float2 *>(r) = make_float2((float)(ux * ru.iq), (float)(uy * ru.iq));
    r[2] = __float_as_uint((float)(uz * ru.iq));
}


template <int NC>
struct Vec {
    double v[NC];
};

// Combine per-lane partials of one block in canonical order; the block
// partial lands in out[] of thread 0.  B = blockDim.x (power of two, >= 64).
template <int NC>
__device__ __forceinline__ void block_canon(Vec<NC> &p, double (*sh)[16])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double o = __shfl_down(p.v[c], off, 64);
            p.v[c] = p.v[c] + o;
        }
    }
    if (lane == 0)
        for (int c = 0; c < NC; ++c) sh[c][wave] = p.v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int off = 1; off < nwaves; off <<= 1)
            for (int w = 0; w < nwaves; w += 2 * off)
                for (int c = 0; c < NC; ++c) sh[c][w] = sh[c][w] + sh[c][w + off];
        for (int c = 0; c < NC; ++c) p.v[c] = sh[c][0];
    }
}

// Cross-block reduction in two launches: every block writes its canonical
// partial, and a one-workgroup finishing kernel adds the G partials with the
// fixed tree (offsets 1, 2, 4, ... over the blocks).  Measured faster than
// finishing in the last block of the same launch (fences and a ticket):
// an ADR iteration on a 315 x 315 plate 96 -> 62 us.
template <int NC>
__device__ __forceinline__ void block_partial(Vec<NC> &p, double *partials, double (*sh)[16])
{
    block_canon<NC>(p, sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < NC; ++c) partials[blockIdx.x * NC + c] = p.v[c];
}

// Called by every thread of a one-workgroup launch; thread 0 returns true
// with the totals.  The partials are loaded into LDS together (a tree over
// them in global memory would wait for one load per step).
constexpr int kFinishLds = 1024;  // doubles of partials the LDS path holds

template <int NC>
__device__ __forceinline__ bool finish_tree(int G, double *partials, Vec<NC> &tot)
{
    __shared__ double q[kFinishLds];
    const bool in_lds = G * NC <= kFinishLds;
    if (in_lds)
        for (int k = threadIdx.x; k < G * NC; k += blockDim.x) q[k] = partials[k];
    __syncthreads();
    if (threadIdx.x != 0) return false;
    double *v = in_lds ? q : partials;
    for (int off = 1; off < G; off <<= 1)
        for (int b = 0; b < G; b += 2 * off)
            for (int c = 0; c < NC; ++c) v[b * NC + c] = v[b * NC + c] + v[(b + off) * NC + c];
    for (int c = 0; c < NC; ++c) tot.v[c] = v[c];
    return true;
}

template <int DIM>
__global__ void k_verlet1(int N_int, int n_top, int n_bot, double dt, double load, D3 v, CD3 a, D3 u,
                          double *cyl, RecU rec)
{
    // Integrate.cu:447-463 (i < N_int) and Solve_Plate_Crack_Grow.cu:20-27
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cyl && i == 0) {
        // Solve_Kalthoff_Winkler.cu:101-110: vel += dt/2*acc; disp += vel*dt
        for (int d = 0; d < 3; ++d) {
            cyl[d] += (dt / 2.0) * cyl[6 + d];
            cyl[3 + d] += cyl[d] * dt;
        }
    }
    if (i >= N_int + n_top + n_bot) return;
    double ud[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < DIM; ++d) ud[d] = u.p[d][i];
    if (i < N_int) {
        double vd[DIM], ad[DIM];
        for (int d = 0; d < DIM; ++d) {
            vd[d] = v.p[d][i];
            ad[d] = a.p[d][i];
        }
        for (int d = 0; d < DIM; ++d) {
            vd[d] += (dt / 2.0) * ad[d];
            v.p[d][i] = vd[d];
            ud[d] += vd[d] * dt;
            u.p[d][i] = ud[d];
        }
    } else {
        ud[1] = i < N_int + n_top ? ud[1] + load * dt : ud[1] - load * dt;
        u.p[1][i] = ud[1];
    }
    if (rec.rec) store_rec_u(rec, i, ud[0], ud[1], ud[2]);
}

// Half 2 of one step and half 1 of the next in one pass (runs without the
// KW cylinder): the same operations in the same order as k_verlet2 followed
// by k_verlet1, with the acceleration kept in a register.  The acceleration
// of a step that another step follows is never read by the API, so only
// the last step's k_verlet2 stores it.
template <int DIM, bool F32>
__global__ void k_verlet21(int N_int, int n_top, int n_bot, double dt, double dens, double load, D3 v, CD3 f,
                           CF3 f32, CD3 b, D3 u, RecU rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int + n_top + n_bot) return;
    // every load of the particle is issued before the arithmetic
    double ud[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < DIM; ++d) ud[d] = u.p[d][i];
    if (i < N_int) {
        double fd[DIM], bd[DIM], vd[DIM];
        for (int d = 0; d < DIM; ++d) {
            fd[d] = F32 ? (double)f32.p[d][i] : f.p[d][i];
            bd[d] = b.p[0] ? b.p[d][i] : 0.0;  // null: the body force is zero
            vd[d] = v.p[d][i];
        }
        for (int d = 0; d < DIM; ++d) {
            const double ad = (fd[d] + bd[d]) / dens;  // Integrate.cu:465-481
            double w = vd[d] + (dt / 2.0) * ad;
            w = w + (dt / 2.0) * ad;                    // Integrate.cu:447-463
            v.p[d][i] = w;
            ud[d] += w * dt;
            u.p[d][i] = ud[d];
        }
    } else {
        ud[1] = i < N_int + n_top ? ud[1] + load * dt : ud[1] - load * dt;
        u.p[1][i] = ud[1];
    }
    if (rec.rec) store_rec_u(rec, i, ud[0], ud[1], ud[2]);
}

template <int DIM, bool F32>
__global__ void k_verlet2(int N_int, double dt, double dens, D3 v, D3 a, CD3 f, CF3 f32, CD3 b)
{
    // Integrate.cu:465-481
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    double fd[DIM], bd[DIM], vd[DIM];
    for (int d = 0; d < DIM; ++d) {
        fd[d] = F32 ? (double)f32.p[d][i] : f.p[d][i];
        bd[d] = b.p[0] ? b.p[d][i] : 0.0;  // null: the body force is zero
        vd[d] = v.p[d][i];
    }
    for (int d = 0; d < DIM; ++d) {
        const double ad = (fd[d] + bd[d]) / dens;
        a.p[d][i] = ad;
        v.p[d][i] = vd[d] + (dt / 2.0) * ad;
    }
}

__device__ __forceinline__ double adr_c(double cn1, double cn2)
{
    // Integrate.cu:117-128
    double cn;
    if (cn2 != 0.0) {
        if (cn1 / cn2 > 0.0)
            cn = 2.0 * sqrt(cn1 / cn2);
        else
            cn = 0.0;
    } else {
        cn = 0.0;
    }
    if (cn > 2.0) cn = 1.9;
    return cn;
}

// Per-particle terms of the ADR reduction, one thread per particle (the
// whole GPU), into terms[4][N]: the damping Rayleigh-quotient terms of
// kernel_cnn (Integrate.cu:19-31) and the residual terms.  The reduction
// below adds them in the order of the fixed T-lane tree.
__global__ void k_adr_terms(int dim, int N_int, int N, double mass, double dt, CD3 u, CD3 f, CF3 f32, int use32,
                            CD3 fo, CD3 vho, CD3 b, const DevScalars *S, double *terms)
{
    if (S->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double cn = 0.0, disp = 0.0, rn = 0.0, rd = 0.0;
    if (i < N_int) {
        for (int d = 0; d < dim; ++d) {
            double ud = u.p[d][i];
            double fd = use32 ? (double)f32.p[d][i] : f.p[d][i];
            double vd = vho.p[d][i];
            if (d == 0) {
                if (vd != 0.0) cn = -1.0 * ud * ud * (fd / mass - fo.p[d][i] / mass) / (dt * vd);
                disp = ud * ud;
            } else {
                if (vd != 0.0) cn -= 1.0 * ud * ud * (fd / mass - fo.p[d][i] / mass) / (dt * vd);
                disp += ud * ud;
            }
            double r = fd + b.p[d][i];
            rn += r * r;
            rd += b.p[d][i] * b.p[d][i];
        }
    } else {
        for (int d = 0; d < dim; ++d) {
            double fd = use32 ? (double)f32.p[d][i] : f.p[d][i];
            rd += fd * fd;
        }
    }
    terms[i] = cn;
    terms[(size_t)N + i] = disp;
    terms[2 * (size_t)N + i] = rn;
    terms[3 * (size_t)N + i] = rd;
}

// Lane t of the T = gridDim.x * blockDim.x lanes adds terms t, t+T, ... in
// order, then the fixed tree over the lanes: within each block here, across
// the blocks in k_adr_finish, which derives c, the residual and the stop
// flag on the device.
__global__ __launch_bounds__(1024) void k_adr_reduce(int N, const double *__restrict__ terms, double *partials,
                                                     DevScalars *S, double tol)
{
    __shared__ double sh[4][16];
    if (S->done) return;
    const int T = gridDim.x * blockDim.x;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    Vec<4> p{{0.0, 0.0, 0.0, 0.0}};
    int i = t;
    // four strided groups of loads in flight, added in the same order
    for (; i + 3 * T < N; i += 4 * T) {
        double q[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) q[u][c] = terms[(size_t)c * N + i + u * T];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) p.v[c] += q[u][c];
    }
    for (; i < N; i += T) {
        p.v[0] += terms[i];
        p.v[1] += terms[(size_t)N + i];
        p.v[2] += terms[2 * (size_t)N + i];
        p.v[3] += terms[3 * (size_t)N + i];
    }
    block_partial<4>(p, partials, sh);
}

// The damping coefficient, the residual and the stop flag from the G block
// partials of k_adr_reduce (one workgroup; thread 0 writes).
__global__ void k_adr_finish(int G, double *partials, DevScalars *S, double tol)
{
    if (S->done) return;
    Vec<4> tot;
    if (finish_tree<4>(G, partials, tot)) {
        double c = adr_c(tot.v[0], tot.v[1]);
        double rnum = tot.v[2], rden = tot.v[3];
        double r = rden > 0.0 ? sqrt(rnum) / sqrt(rden) : (rnum > 0.0 ? INFINITY : 0.0);
        S->c = c;
        S->resid = r;
        for (int k = 0; k < 4; ++k) S->sums[k] = tot.v[k];
        if (!(isfinite(tot.v[0]) && isfinite(tot.v[1]) && isfinite(rnum) && isfinite(rden))) S->nonfinite = 1;
        S->ct_use = S->ct_counter;
        S->ct_counter = S->ct_counter + 1;
        if (tol >= 0.0 && r <= tol) S->done = 1;
    }
}

__global__ void k_adr_update(int dim, int N_int, double mass, double dt, CD3 f, CF3 f32, int use32, CD3 b,
                             D3 vh, D3 v, D3 vho, D3 u, D3 fo, DevScalars *S, RecU rec)
{
    // kernel_integrate, Integrate.cu:62-105
    if (S->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_int) return;
    const int ct = S->ct_use;
    const double c = S->c;
    for (int d = 0; d < dim; ++d) {
        double fd = use32 ? (double)f32.p[d][i] : f.p[d][i];
        double hv;
        if (ct == 0)
            hv = 1.0 * dt / mass * (fd + b.p[d][i]) / 2.0;
        else
            hv = ((2.0 - c * dt) * vho.p[d][i] + 2.0 * dt / mass * (fd + b.p[d][i])) / (2.0 + c * dt);
        vh.p[d][i] = hv;
        v.p[d][i] = 0.5 * (vho.p[d][i] + hv);
        u.p[d][i] += hv * dt;
        vho.p[d][i] = hv;
        fo.p[d][i] = fd;
    }
    if (rec.rec) store_rec_u(rec, i, u.p[0][i], u.p[1][i], dim == 3 ? u.p[2][i] : 0.0);
}

__global__ void k_set_ct(DevScalars *S, int ct, int clear_done)
{
    S->ct_counter = ct;
    if (clear_done) {
        S->done = 0;
        S->iters = 0;
    }
}

// Contact of every particle with the cylinder, one thread per particle:
// Solve_Kalthoff_Winkler.cu:177-239 moves the particles inside the cylinder
// to its surface and writes their reaction terms to terms[3][N] (zero for the
// others).  The reference sums the reaction by float atomicAdd
// (Solve_Kalthoff_Winkler.cu:168-170); k_contact_sum adds the terms in the
// order of the fixed T-lane tree.
__global__ void k_contact_terms(int N, const double *__restrict__ dmg, CD3 x, D3 u, D3 v, pd_cyl cyl, double dt,
                                double dens, double vol, const DevScalars *S, double *terms, RecU rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double *cd = S->cyl + 3;
    const double cdx = cd[0], cdy = cd[1], cdz = cd[2];
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (dmg[i] < 1.0) {
        double dpx = x.p[0][i] + u.p[0][i] - cyl.x - cdx;
        double dpy = x.p[1][i] + u.p[1][i] - cyl.y - cdy;
        double dpz = x.p[2][i] + u.p[2][i] - cyl.z - cdz;
        double dpd = sqrt(dpx * dpx + dpz * dpz);
        if (dpd <= cyl.rad && fabs(dpy) <= (cyl.height / 2.0)) {
            double rx, ry, rz;
            if (cyl.rad - dpd <= cyl.height / 2.0 - fabs(dpy)) {
                rx = dpx / dpd * cyl.rad;
                ry = dpy;
                rz = dpz / dpd * cyl.rad;
            } else if (dpy >= 0.0) {
                rx = dpx;
                ry = cyl.height / 2.0;
                rz = dpz;
            } else {
                rx = dpx;
                ry = -cyl.height / 2.0;
                rz = dpz;
            }
            double oux = u.p[0][i], ouy = u.p[1][i], ouz = u.p[2][i];
            double ovx = v.p[0][i], ovy = v.p[1][i], ovz = v.p[2][i];
            double nux = cyl.x + cdx + rx - x.p[0][i];
            double nuy = cyl.y + cdy + ry - x.p[1][i];
            double nuz = cyl.z + cdz + rz - x.p[2][i];
            u.p[0][i] = nux;
            u.p[1][i] = nuy;
            u.p[2][i] = nuz;
            if (rec.rec) store_rec_u(rec, i, nux, nuy, nuz);
            double nvx = (nux - oux) / dt, nvy = (nuy - ouy) / dt, nvz = (nuz - ouz) / dt;
            v.p[0][i] = nvx;
            v.p[1][i] = nvy;
            v.p[2][i] = nvz;
            r0 = -dens * (nvx - ovx) / dt * vol;
            r1 = -dens * (nvy - ovy) / dt * vol;
            r2 = -dens * (nvz - ovz) / dt * vol;
        }
    }
    terms[i] = r0;
    terms[(size_t)N + i] = r1;
    terms[2 * (size_t)N + i] = r2;
}

// The reaction on the cylinder in canonical order (block partials here, the
// tree across blocks and the cylinder's second half step in k_contact_finish).
__global__ __launch_bounds__(1024) void k_contact_sum(int N, const double *__restrict__ terms, pd_cyl cyl,
                                                      double dt, double *partials, DevScalars *S)
{
    __shared__ double sh[3][16];
    const int T = gridDim.x * blockDim.x;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    Vec<3> p{{0.0, 0.0, 0.0}};
    int i = t;
    for (; i + 3 * T < N; i += 4 * T) {
        double q[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) q[u][c] = terms[(size_t)c * N + i + u * T];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) p.v[c] += q[u][c];
    }
    for (; i < N; i += T) {
        p.v[0] += terms[i];
        p.v[1] += terms[(size_t)N + i];
        p.v[2] += terms[2 * (size_t)N + i];
    }
    block_partial<3>(p, partials, sh);
}

// The reaction on the cylinder from the G block partials of k_contact_sum,
// then the cylinder's second half step (Solve_Kalthoff_Winkler.cu:255-265).
__global__ void k_contact_finish(int G, double *partials, pd_cyl cyl, double dt, DevScalars *S)
{
    Vec<3> tot;
    if (finish_tree<3>(G, partials, tot)) {
        double *cs = S->cyl;  // vel[0..2] disp[3..5] acc[6..8] force[9..11]
        for (int d = 0; d < 3; ++d) {
            cs[9 + d] = tot.v[d];
            S->sums[4 + d] = tot.v[d];
            cs[6 + d] = cs[9 + d] / cyl.mass;
            cs[d] += dt / 2.0 * cs[6 + d];
            cs[9 + d] = 0.0;
        }
    }
}

__global__ __launch_bounds__(1024) void k_field_sum(const double *__restrict__ v, int64_t n, double *partials,
                                                    DevScalars *S)
{
    __shared__ double sh[1][16];
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Vec<1> p{{0.0}};
    for (int64_t i = t; i < n; i += T) p.v[0] += v[i];
    block_partial<1>(p, partials, sh);
}

__global__ void k_field_finish(int G, double *partials, DevScalars *S)
{
    Vec<1> tot;
    if (finish_tree<1>(G, partials, tot)) S->sums[7] = tot.v[0];
}

// The same finish, writing the total to a caller's device double.
__global__ void k_field_finish_to(int G, double *partials, double *out)
{
    Vec<1> tot;
    if (finish_tree<1>(G, partials, tot)) *out = tot.v[0];
}

__global__ void k_strip_displace(int N_int, int n_top, int n_bot, double dtop, double dbot, double *uy)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_top)
        uy[N_int + k] += dtop;
    else if (k < n_top + n_bot)
        uy[N_int + k] += dbot;
}

__global__ void k_f32_to_f64(const float *a, double *b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// at least one block: every kernel bounds-checks its index, so an empty
// range (say N_int = 0) is a no-op launch rather than an invalid grid
inline int nblk(int64_t n, int t) { return n > 0 ? (int)((n + t - 1) / t) : 1; }

} // namespace

hipError_t launch_verlet1(hipStream_t s, int dim, int N_int, int n_top, int n_bot, double dt, double load,
                          D3 v, CD3 a, D3 u, double *cyl_state, RecU rec)
{
    int n = N_int + n_top + n_bot;
    if (n < 1) n = 1;
    if (dim == 3)
        k_verlet1<3><<<nblk(n, 256), 256, 0, s>>>(N_int, n_top, n_bot, dt, load, v, a, u, cyl_state, rec);
    else
        k_verlet1<2><<<nblk(n, 256), 256, 0, s>>>(N_int, n_top, n_bot, dt, load, v, a, u, cyl_state, rec);
    return hipGetLastError();
}

hipError_t launch_verlet21(hipStream_t s, int dim, int N_int, int n_top, int n_bot, double dt, double dens,
                           double load, D3 v, CD3 f, CF3 f32, int use32, CD3 b, D3 u, RecU rec)
{
    int n = N_int + n_top + n_bot;
    if (n < 1) n = 1;
    const dim3 g(nblk(n, 256)), t(256);
    if (dim == 3) {
        if (use32)
            k_verlet21<3, true><<<g, t, 0, s>>>(N_int, n_top, n_bot, dt, dens, load, v, f, f32, b, u, rec);
        else
            k_verlet21<3, false><<<g, t, 0, s>>>(N_int, n_top, n_bot, dt, dens, load, v, f, f32, b, u, rec);
    } else {
        if (use32)
            k_verlet21<2, true><<<g, t, 0, s>>>(N_int, n_top, n_bot, dt, dens, load, v, f, f32, b, u, rec);
        else
            k_verlet21<2, false><<<g, t, 0, s>>>(N_int, n_top, n_bot, dt, dens, load, v, f, f32, b, u, rec);
    }
    return hipGetLastError();
}

hipError_t launch_verlet2(hipStream_t s, int dim, int N_int, double dt, double dens, D3 v, D3 a, CD3 f,
                          CF3 f32, int use32, CD3 b)
{
    if (N_int <= 0) return hipSuccess;
    const dim3 g(nblk(N_int, 256)), t(256);
    if (dim == 3) {
        if (use32)
            k_verlet2<3, true><<<g, t, 0, s>>>(N_int, dt, dens, v, a, f, f32, b);
        else
            k_verlet2<3, false><<<g, t, 0, s>>>(N_int, dt, dens, v, a, f, f32, b);
    } else {
        if (use32)
            k_verlet2<2, true><<<g, t, 0, s>>>(N_int, dt, dens, v, a, f, f32, b);
        else
            k_verlet2<2, false><<<g, t, 0, s>>>(N_int, dt, dens, v, a, f, f32, b);
    }
    return hipGetLastError();
}

hipError_t launch_adr_reduce(hipStream_t s, int blocks, int threads, int dim, int N_int, int N, double mass,
                             double dt, CD3 u, CD3 f, CF3 f32, int use32, CD3 fo, CD3 vho, CD3 b,
                             double *partials, DevScalars *S, double tol, double *terms)
{
    k_adr_terms<<<nblk(N, 256), 256, 0, s>>>(dim, N_int, N, mass, dt, u, f, f32, use32, fo, vho, b, S, terms);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_adr_reduce<<<blocks, threads, 0, s>>>(N, terms, partials, S, tol);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_adr_finish<<<1, 64, 0, s>>>(blocks, partials, S, tol);
    return hipGetLastError();
}

hipError_t launch_adr_update(hipStream_t s, int dim, int N_int, double mass, double dt, CD3 f, CF3 f32,
                             int use32, CD3 b, D3 vh, D3 v, D3 vho, D3 u, D3 fo, DevScalars *S, RecU rec)
{
    if (N_int <= 0) return hipSuccess;
    k_adr_update<<<nblk(N_int, 256), 256, 0, s>>>(dim, N_int, mass, dt, f, f32, use32, b, vh, v, vho, u, fo, S,
                                                  rec);
    return hipGetLastError();
}

hipError_t launch_set_ct(hipStream_t s, DevScalars *S, int ct, int clear_done)
{
    k_set_ct<<<1, 1, 0, s>>>(S, ct, clear_done);
    return hipGetLastError();
}

hipError_t launch_contact(hipStream_t s, int blocks, int threads, int N, const double *dmg, CD3 x, D3 u, D3 v,
                          pd_cyl cyl, double dt, double dens, double vol, double *partials, DevScalars *S,
                          RecU rec, double *terms)
{
    k_contact_terms<<<nblk(N, 256), 256, 0, s>>>(N, dmg, x, u, v, cyl, dt, dens, vol, S, terms, rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_contact_sum<<<blocks, threads, 0, s>>>(N, terms, cyl, dt, partials, S);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_contact_finish<<<1, 64, 0, s>>>(blocks, partials, cyl, dt, S);
    return hipGetLastError();
}

hipError_t launch_field_sum(hipStream_t s, int blocks, int threads, const double *v, int64_t n,
                            double *partials, DevScalars *S)
{
    k_field_sum<<<blocks, threads, 0, s>>>(v, n, partials, S);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_field_finish<<<1, 64, 0, s>>>(blocks, partials, S);
    return hipGetLastError();
}

hipError_t launch_field_sum_to(hipStream_t s, int blocks, int threads, const double *v, int64_t n, double *partials,
                               double *out)
{
    k_field_sum<<<blocks, threads, 0, s>>>(v, n, partials, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_field_finish_to<<<1, 64, 0, s>>>(blocks, partials, out);
    return hipGetLastError();
}

hipError_t launch_strip_displace(hipStream_t s, int N_int, int n_top, int n_bot, double dtop, double dbot,
                                 double *uy)
{
    int n = n_top + n_bot;
    if (n <= 0) return hipSuccess;
    k_strip_displace<<<nblk(n, 256), 256, 0, s>>>(N_int, n_top, n_bot, dtop, dbot, uy);
    return hipGetLastError();
}

hipError_t launch_f32_to_f64(hipStream_t s, const float *a, double *b, int64_t n)
{
    if (n <= 0) return hipSuccess;
    k_f32_to_f64<<<nblk(n, 256), 256, 0, s>>>(a, b, n);
    return hipGetLastError();
}

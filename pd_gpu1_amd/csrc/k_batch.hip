// k_batch.hip -- batched quasi-static crack growth: many independent 2D PMB
// plates relaxed by adaptive dynamic relaxation (ADR) in one persistent launch.
//
// One workgroup (or a group of W, below) owns one specimen at a time and
// pulls the next specimen index from a device work queue (an atomic counter
// over the specimens sorted by bond count, largest first), so 512 plates of
// 1.6k-40k particles
// load-balance over the 256 CUs without host involvement and the largest
// plates, which set the wall time, start first.  Inside a
// specimen the workgroup runs whole ADR iterations: PMB force over every row
// (Force_Bond.cu:159-202 arithmetic, FP64, rows in CSR order), the damping
// terms of kernel_cnn (Integrate.cu:14-34) and the residual reduced with the
// fixed low-bit-first tree over T = blockDim lanes, the damping coefficient of
// cal_cn (Integrate.cu:117-128), the convergence test, and the update of
// kernel_integrate (Integrate.cu:62-105).  Workgroup barriers separate the
// phases; no grid-wide synchronisation and no host round trip is needed
// (the reference copies both sums to the host every iteration,
// Reduce.cu:61-62, Integrate.cu:112-113).
//
// When there are fewer specimens than resident workgroups, W > 1 workgroups
// (a group, W a power of two) share each specimen: rows and particles are
// split over the group's W*BT threads, the phases are separated by a group
// barrier in global memory (fence, arrival counter, bounded spin, fence), and
// the damping sums keep the single-workgroup order: part p of the group owns
// the T/W consecutive lanes p*T/W ... of the T = BT lane tree, a complete
// subtree, and the W subtree roots are combined by the remaining levels of
// the same tree.  Results are therefore identical for every W.
#include "pd_batch.h"
#include "lanes.h"

namespace {

__device__ __forceinline__ bool intact(const uint32_t *fail, int64_t k)
{
    return (fail[k >> 5] >> (k & 31)) & 1u;
}

template <int BT>
__device__ __forceinline__ void block_canon4(double p[4], double (*sh)[BT / 64], double out[4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < 4; ++c) p[c] = p[c] + __shfl_down(p[c], off, 64);
    if (lane == 0)
        for (int c = 0; c < 4; ++c) sh[c][wave] = p[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int off = 1; off < BT / 64; off <<= 1)
            for (int w = 0; w < BT / 64; w += 2 * off)
                for (int c = 0; c < 4; ++c) sh[c][w] = sh[c][w] + sh[c][w + off];
        for (int c = 0; c < 4; ++c) out[c] = sh[c][0];
    }
}

// Coefficients {kappa, |xi|} of bond k with neighbour c of a row at x0, x1:
// streamed from KD, or (RC) recomputed from the reference positions XY with
// the set-up's own FP64 expressions -- |xi| and the partial-volume factor of
// k_vol_corr (Surf_Corr.cu:312-337), kappa = vol*fac*bc*scr of k_pmb_coef
// with scr = 1 (no surface correction in the plate family) -- which give
// the stored values bit for bit; the batch checks that once per set-up.
// Recomputing takes the 16-byte coefficient stream of every bond off HBM for
// a 16-byte gather of a read-only position that the group's L2 holds.
template <bool RC>
__device__ __forceinline__ double2 bond_coef(int64_t k, int c, const double2 *KD, const double2 *XY, double x0,
                                             double x1, const BatchCoef &cg)
{
    if (!RC) return KD[k];
    const double2 xc = XY[c];
    const double ddx = xc.x - x0, ddy = xc.y - x1;
    const double id = sqrt(ddx * ddx + ddy * ddy);
    const double lo = cg.delta - cg.dx / 2.0, hi = cg.delta + cg.dx / 2.0;
    double fac;
    if (id <= lo)
        fac = 1.0;
    else if (id <= hi)
        fac = (hi - id) / cg.dx;
    else
        fac = 0.0;
    return make_double2(cg.vol * fac * cg.bc, id);  // times scr = 1, which changes no bit
}

// Pair force term of bond k of a row at (x0+u0, x1+u1), Force_Bond.cu:185-196
// operand order, kd = {kappa, |xi|} of the bond; breaks the bond when its
// stretch exceeds s0.  P[c] holds (X[c] + UX[c], Y[c] + UY[c]), the first sum
// of that order, so the neighbour costs one 16-byte load.
__device__ __forceinline__ void bond_term(int64_t k, int c, double2 kd, double x0, double x1, double u0,
                                          double u1, const double2 *P, double s0, uint32_t *fail, double &tx,
                                          double &ty)
{
    const double2 pc = P[c];
    double d0 = pc.x - x0 - u0;
    double d1 = pc.y - x1 - u1;
    double nl = sqrt(d0 * d0 + d1 * d1);
    double id = kd.y;
    double sv = (nl - id) / id;
    double ks = kd.x * sv;
    tx = ks * d0 / nl;
    ty = ks * d1 / nl;
    if (sv > s0) atomicAnd(&fail[k >> 5], ~(1u << (k & 31)));
}

// Workgroup-uniform values read from LDS, moved to scalar registers so that
// the specimen's base pointers and loop bounds stay out of vector registers.
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double uniform_double(double v)
{
    const int2 w = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readfirstlane(w.x), __builtin_amdgcn_readfirstlane(w.y)));
}

// Barrier of the W workgroups of a group.  Every workgroup of the grid is
// resident (the host sizes the grid from the occupancy), so the spin ends
// when the group's other workgroups arrive; a wait longer than
// gr.spin_ticks of the constant-rate wall clock, or another group's abort,
// sets the abort flag and the workgroup leaves the kernel.  Returns false on
// abort.
__device__ __forceinline__ bool group_sync(const BatchGroups &gr, int grp, unsigned &arrivals, int &ok_sh)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        arrivals += (unsigned)gr.parts;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this workgroup's writes before its arrival
        atomicAdd(&gr.bar[grp], 1u);
        const uint64_t t0 = wall_clock64();
        int ok = 1;
        while (__hip_atomic_load(&gr.bar[grp], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < arrivals) {
            if (__hip_atomic_load(gr.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > gr.spin_ticks) {
                atomicExch(gr.abort, 1);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the group's writes before this workgroup's reads
        ok_sh = ok;
    }
    __syncthreads();
    return ok_sh != 0;
}

// The persistent kernel's body; k_batch_adr and k_batch_adr_dense below are
// its two builds.
template <int BT, int G, bool RC>
__device__ __forceinline__ void batch_adr(const BatchView &B, double du, double tol, int max_iters, unsigned *queue,
                                          int32_t *iters_out, double *resid_out, const BatchGroups &gr)
{
    __shared__ double sh[4][BT / 64];
    __shared__ int spec_sh, ok_sh;
    __shared__ double c_sh, r_sh, rd_sh;
    __shared__ int sep_sh, pause_sh;
    __shared__ double q_sh[kMaxParts][4];
    const int tid = threadIdx.x;
    const int W = gr.parts;                       // workgroups per specimen
    // Workgroups reach the 8 XCDs round-robin; with gr.xcd the renumbering
    // b -> (b mod 8)(grid/8) + b/8 gives the W workgroups of a group one XCD,
    // so a specimen's particle state sits in one L2.
    const int b = gr.xcd ? (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3) : (int)blockIdx.x;
    const int grp = b / W, part = b % W;
    const int gt = part * BT + tid, GT = W * BT;  // thread within the group, threads of the group
    const int L = BT / W;                         // damping-tree lanes of this part
    unsigned arrivals = 0;
#define PD_GROUP_SYNC()                                                        \
    do {                                                                       \
        if (W == 1)                                                            \
            __syncthreads();                                                   \
        else if (!group_sync(gr, grp, arrivals, ok_sh))                        \
            return;                                                            \
    } while (0)
    for (;;) {
        if (tid == 0 && part == 0) {
            const int t = (int)atomicAdd(queue, 1u);
            const int q = t < B.nqueue ? B.order[t] : B.nspec;
            if (W == 1)
                spec_sh = q;
            else
                gr.spec[grp] = q;
        }
        PD_GROUP_SYNC();
        if (W > 1 && tid == 0) spec_sh = gr.spec[grp];
        __syncthreads();
        const int s = uniform_int(spec_sh);
        PD_GROUP_SYNC();  // every part has read s before part 0 takes the next ticket
        if (s >= B.nspec) break;
        if (B.retired && B.retired[s]) {
            if (tid == 0 && part == 0) {
                iters_out[s] = 0;
                if (B.paused) B.paused[s] = 0;
                if (B.active) atomicSub(B.active, 1);
            }
            continue;  // uniform: every thread of the group read the same s
        }
        const BatchSpec d = B.spec[s];
        const int N = d.n, Ni = d.n_int;
        const double *X = B.x + d.p_off, *Y = B.y + d.p_off;
        double *UX = B.ux + d.p_off, *UY = B.uy + d.p_off;
        double *FX = B.fx + d.p_off, *FY = B.fy + d.p_off;
        double *VHX = B.vhx + d.p_off, *VHY = B.vhy + d.p_off;
        double *VHOX = B.vhox + d.p_off, *VHOY = B.vhoy + d.p_off;
        double *VX = B.vx + d.p_off, *VY = B.vy + d.p_off;
        double *FOX = B.fox + d.p_off, *FOY = B.foy + d.p_off;
        const int32_t *rp = B.row_ptr + d.rp_off;
        const int32_t *col = B.col + d.b_off;
        const double2 *KD = B.kd + d.b_off;
        const double2 *XY = B.xy + d.p_off;
        uint32_t *fail = B.fail + (d.b_off >> 5);
        const double mass = B.mass, dt = B.dt, s0 = B.s0;
        double *DUPX = B.dupx + d.p_off, *DUPY = B.dupy + d.p_off;
        double2 *P = B.pos + d.p_off;
        if (B.warm && !B.resume) {
            // start from u + the previous increment; keep the start state
            for (int i = gt; i < Ni; i += GT) {
                const double sx = UX[i], sy = UY[i];
                UX[i] = sx + DUPX[i];
                UY[i] = sy + DUPY[i];
                DUPX[i] = sx;
                DUPY[i] = sy;
            }
        }
        // displacement control of the strips (Solve_Plate_Crack_Grow.cu:20-27 layout)
        if (!B.resume)
            for (int k = gt; k < d.n_top + d.n_bot; k += GT) UY[Ni + k] += k < d.n_top ? du : -du;
        PD_GROUP_SYNC();
        for (int i = gt; i < N; i += GT) P[i] = make_double2(X[i] + UX[i], Y[i] + UY[i]);
        PD_GROUP_SYNC();
        // A resumed specimen continues at the iteration it paused at: its
        // state (U, VHO, FO, P) is what the paused launch left, so the
        // iterations are those of one uninterrupted run.  A pause is decided
        // during an iteration and taken at the top of the next, before its
        // force evaluation.
        int it = B.resume ? uniform_int(B.it0[s]) : 0;
        double r = B.resume ? resid_out[s] : 0.0;
        bool paused = false, separated = false;
        int pause_next = 0;
        for (;;) {
            if (pause_next) {
                paused = true;
                break;
            }
            // PMB force, every row (Force_Bond.cu:169).  A part computes the
            // rows of the particles on its own damping-tree lanes (i mod BT
            // in [part*L, part*L + L)), so the damping sums below need only
            // a workgroup barrier; own_row(q) ascends with q.
            const auto own_row = [&](int q) { return part * L + q % L + (q / L) * BT; };
            if (G == 1) {
                for (int q = tid;; q += BT) {
                    const int i = own_row(q);
                    if (i >= N) break;
                    const double x0 = X[i], x1 = Y[i], u0 = UX[i], u1 = UY[i];
                    double fx = 0.0, fy = 0.0;
                    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
                        if (!intact(fail, k)) continue;
                        double tx, ty;
                        const int c = col[k];
                        bond_term(k, c, bond_coef<RC>(k, c, KD, XY, x0, x1, B.coef), x0, x1, u0, u1, P, s0, fail, tx,
                                  ty);
                        fx += tx;
                        fy += ty;
                    }
                    FX[i] = fx;
                    FY[i] = fy;
                }
            } else {
                const int lane = tid & (G - 1);
                for (int q = tid / G;; q += BT / G) {  // the lane group takes the row
                    const int i = own_row(q);
                    if (i >= N) break;
                    const double x0 = X[i], x1 = Y[i], u0 = UX[i], u1 = UY[i];
                    double fx = 0.0, fy = 0.0;
                    const int32_t kb = rp[i], ke = rp[i + 1];
                    // the neighbour index and intact bit of the lane's next
                    // bond load while the current bond computes (a bond's
                    // bit changes only when that bond itself breaks)
                    int32_t k = kb + lane;
                    int c = k < ke ? col[k] : 0;
                    int ok = k < ke && intact(fail, k);
                    for (int32_t k0 = kb; k0 < ke; k0 += G) {
                        const int32_t kn = k + G;
                        const int cn = kn < ke ? col[kn] : 0;
                        const int okn = kn < ke && intact(fail, kn);
                        double tx = 0.0, ty = 0.0;
                        if (ok)
                            bond_term(k, c, bond_coef<RC>(k, c, KD, XY, x0, x1, B.coef), x0, x1, u0, u1, P, s0, fail,
                                      tx, ty);
                        // add the terms in bond order, as the sequential loop does
                        const int cnt = ke - k0 < G ? ke - k0 : G;
                        if constexpr (G == 4 || G == 8) {
                            // broadcasts over DPP (lanes.h): 512 plates x 300
                            // iterations 683 -> 592 ms against __shfl
                            lanes::Unroll<G>::run([&](auto lc) {
                                constexpr int L = decltype(lc)::value;
                                const double sx = lanes::bcast<G, L>(tx), sy = lanes::bcast<G, L>(ty);
                                if (L < cnt && lanes::bcast<G, L>(ok)) {
                                    fx += sx;
                                    fy += sy;
                                }
                            });
                        } else {
#pragma unroll
                            for (int l = 0; l < G; ++l) {
                                const double sx = __shfl(tx, l, G), sy = __shfl(ty, l, G);
                                if (l < cnt && __shfl(ok, l, G)) {
                                    fx += sx;
                                    fy += sy;
                                }
                            }
                        }
                        k = kn;
                        c = cn;
                        ok = okn;
                    }
                    if (lane == 0) {
                        FX[i] = fx;
                        FY[i] = fy;
                    }
                }
            }
            __syncthreads();  // this part's lanes read only rows this workgroup computed
            // damping and residual terms, lane t of the T = BT lane tree owns
            // particles t, t+BT, ...; this part holds lanes part*L ... part*L+L-1
            double p[4] = {0.0, 0.0, 0.0, 0.0};
            for (int i = tid < L ? part * L + tid : N; i < N; i += BT) {
                double cn = 0.0, disp = 0.0, rn = 0.0, rd = 0.0;
                if (i < Ni) {
                    const double ux = UX[i], uy = UY[i], fx = FX[i], fy = FY[i];
                    if (VHOX[i] != 0.0) cn = -1.0 * ux * ux * (fx / mass - FOX[i] / mass) / (dt * VHOX[i]);
                    disp = ux * ux;
                    if (VHOY[i] != 0.0) cn -= 1.0 * uy * uy * (fy / mass - FOY[i] / mass) / (dt * VHOY[i]);
                    disp += uy * uy;
                    rn = fx * fx;
                    rn += fy * fy;
                } else {
                    rd = FX[i] * FX[i];
                    rd += FY[i] * FY[i];
                }
                p[0] += cn;
                p[1] += disp;
                p[2] += rn;
                p[3] += rd;
            }
            double tot[4];
            // lanes from L on add zeros, so with W > 1 the workgroup's tree
            // returns the root of this part's subtree exactly
            block_canon4<BT>(p, sh, tot);
            // pause when every specimen has been handed out and the device
            // runs at most pause_below of them (part 0 decides for the group)
            int pf = 0;
            if (B.pause_below > 0 && tid == 0 && part == 0)
                pf = __hip_atomic_load(queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned)B.nqueue &&
                     __hip_atomic_load(B.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= B.pause_below;
            if (W > 1) {
                if (tid == 0) {
                    for (int c = 0; c < 4; ++c) gr.sum[((size_t)grp * W + part) * 4 + c] = tot[c];
                    if (part == 0) gr.flag[grp] = pf;
                }
                PD_GROUP_SYNC();
                if (tid < W)
                    for (int c = 0; c < 4; ++c) q_sh[tid][c] = gr.sum[((size_t)grp * W + tid) * 4 + c];
                __syncthreads();
                if (tid == 0) {
                    pf = gr.flag[grp];
                    // the remaining levels of the tree, over the W subtree roots
                    for (int off = 1; off < W; off <<= 1)
                        for (int w = 0; w < W; w += 2 * off)
                            for (int c = 0; c < 4; ++c) q_sh[w][c] = q_sh[w][c] + q_sh[w + off][c];
                    for (int c = 0; c < 4; ++c) tot[c] = q_sh[0][c];
                }
            }
            if (tid == 0) {
                double cn1 = tot[0], cn2 = tot[1], cc;
                if (cn2 != 0.0) {
                    if (cn1 / cn2 > 0.0)
                        cc = 2.0 * sqrt(cn1 / cn2);
                    else
                        cc = 0.0;
                } else {
                    cc = 0.0;
                }
                if (cc > 2.0) cc = 1.9;
                c_sh = cc;
                pause_sh = pf;
                r_sh = tot[3] > 0.0 ? sqrt(tot[2]) / sqrt(tot[3]) : (tot[2] > 0.0 ? INFINITY : 0.0);
                rd_sh = sqrt(tot[3]);
                // separation: the strips' reaction has collapsed against the
                // last converged increment's (the plate no longer carries load)
                sep_sh = B.sep > 0.0 && B.react[s] > 0.0 && sqrt(tot[3]) < B.sep * B.react[s];
            }
            __syncthreads();
            r = uniform_double(r_sh);
            if (r <= tol || it >= max_iters || !(r == r)) break;
            if (uniform_int(sep_sh)) {
                separated = true;
                break;
            }
            pause_next = uniform_int(pause_sh);
            const double cc = uniform_double(c_sh);
            for (int i = gt; i < Ni; i += GT) {
                double h0, h1;
                const double fx = FX[i], fy = FY[i];
                if (it == 0) {
                    h0 = 1.0 * dt / mass * (fx + 0.0) / 2.0;
                    h1 = 1.0 * dt / mass * (fy + 0.0) / 2.0;
                } else {
                    h0 = ((2.0 - cc * dt) * VHOX[i] + 2.0 * dt / mass * (fx + 0.0)) / (2.0 + cc * dt);
                    h1 = ((2.0 - cc * dt) * VHOY[i] + 2.0 * dt / mass * (fy + 0.0)) / (2.0 + cc * dt);
                }
                VHX[i] = h0;
                VHY[i] = h1;
                VX[i] = 0.5 * (VHOX[i] + h0);
                VY[i] = 0.5 * (VHOY[i] + h1);
                const double ux1 = UX[i] + h0 * dt, uy1 = UY[i] + h1 * dt;
                UX[i] = ux1;
                UY[i] = uy1;
                P[i] = make_double2(X[i] + ux1, Y[i] + uy1);
                VHOX[i] = h0;
                VHOY[i] = h1;
                FOX[i] = fx;
                FOY[i] = fy;
            }
            PD_GROUP_SYNC();
            ++it;
        }
        if (B.warm && !paused) {
            for (int i = gt; i < Ni; i += GT) {
                DUPX[i] = UX[i] - DUPX[i];
                DUPY[i] = UY[i] - DUPY[i];
            }
        }
        if (tid == 0 && part == 0) {
            iters_out[s] = it;
            resid_out[s] = r;
            if (B.paused) B.paused[s] = paused ? 1 : 0;
            if (B.active) atomicSub(B.active, 1);
            if (!paused) {
                if (B.retired && !(r <= tol)) B.retired[s] = separated ? 2 : 1;
                if (r <= tol && B.react) B.react[s] = rd_sh;
            }
        }
        PD_GROUP_SYNC();
    }
#undef PD_GROUP_SYNC
}

// The default build: up to 128 VGPRs, so one 1024-thread workgroup per CU
// (or four of 256 threads, two of 512).
template <int BT, int G, bool RC>
__global__ __launch_bounds__(BT) void k_batch_adr(BatchView B, double du, double tol, int max_iters,
                                                  unsigned *queue, int32_t *iters_out, double *resid_out,
                                                  BatchGroups gr)
{
    batch_adr<BT, G, RC>(B, du, tol, max_iters, queue, iters_out, resid_out, gr);
}

// Two 1024-thread workgroups per CU: held at 64 VGPRs (the rest spills to
// scratch, mostly outside the bond loop), twice the waves to hide the
// latency of the bond chains when there are enough workgroups to fill them
// (launches with at most kDenseParts workgroups per specimen).
template <int G, bool RC>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void
k_batch_adr_dense(BatchView B, double du, double tol, int max_iters, unsigned *queue, int32_t *iters_out,
                  double *resid_out, BatchGroups gr)
{
    batch_adr<1024, G, RC>(B, du, tol, max_iters, queue, iters_out, resid_out, gr);
}

} // namespace

namespace {
__global__ void k_pack_bond_pairs(int64_t nb, const double *__restrict__ kappa, const double *__restrict__ idist,
                                  double2 *kd)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nb) kd[k] = make_double2(kappa[k], idist[k]);
}

// Per bond of one specimen, thread per row: 1 is added to *bad for every bond
// whose stored {kappa, |xi|} differs from bond_coef's recomputation.
__global__ void k_batch_coef_check(int n, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                   const double2 *__restrict__ kd, const double2 *__restrict__ xy, BatchCoef cg,
                                   unsigned *bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned nbad = 0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const double2 r = bond_coef<true>(k, col[k], kd, xy, xy[i].x, xy[i].y, cg);
        const double2 s = kd[k];
        nbad += __double_as_longlong(r.x) != __double_as_longlong(s.x) ||
                __double_as_longlong(r.y) != __double_as_longlong(s.y);
    }
    if (nbad) atomicAdd(bad, nbad);
}

template <int BT, bool RC>
hipError_t batch_g(hipStream_t st, int grid, int group, bool dense, const BatchView &B, double du, double tol,
                   int max_iters, unsigned *queue, int32_t *iters_out, double *resid_out, const BatchGroups &gr)
{
#define PD_BATCH_CASE(GV)                                                                                    \
    case GV:                                                                                                 \
        if constexpr (BT == 1024) {                                                                          \
            if (dense) {                                                                                     \
                k_batch_adr_dense<GV, RC><<<grid, BT, 0, st>>>(B, du, tol, max_iters, queue, iters_out,      \
                                                               resid_out, gr);                               \
                break;                                                                                       \
            }                                                                                                \
        }                                                                                                    \
        k_batch_adr<BT, GV, RC><<<grid, BT, 0, st>>>(B, du, tol, max_iters, queue, iters_out, resid_out, gr); \
        break;
    switch (group) {
    PD_BATCH_CASE(1) PD_BATCH_CASE(4) PD_BATCH_CASE(8) PD_BATCH_CASE(16) PD_BATCH_CASE(32)
    default: return hipErrorInvalidValue;
    }
#undef PD_BATCH_CASE
    return hipGetLastError();
}

template <bool RC>
hipError_t batch_b(hipStream_t st, int grid, int block, int group, bool dense, const BatchView &B, double du,
                   double tol, int max_iters, unsigned *queue, int32_t *iters_out, double *resid_out,
                   const BatchGroups &gr)
{
    switch (block) {
    case 256: return batch_g<256, RC>(st, grid, group, false, B, du, tol, max_iters, queue, iters_out, resid_out, gr);
    case 512: return batch_g<512, RC>(st, grid, group, false, B, du, tol, max_iters, queue, iters_out, resid_out, gr);
    case 1024: return batch_g<1024, RC>(st, grid, group, dense, B, du, tol, max_iters, queue, iters_out, resid_out, gr);
    default: return hipErrorInvalidValue;
    }
}
} // namespace

hipError_t launch_pack_bond_pairs(hipStream_t st, int64_t nb, const double *kappa, const double *idist,
                                  double2 *kd)
{
    if (nb <= 0) return hipSuccess;
    k_pack_bond_pairs<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(nb, kappa, idist, kd);
    return hipGetLastError();
}

hipError_t launch_batch_coef_check(hipStream_t st, int n, const int32_t *row_ptr, const int32_t *col,
                                   const double2 *kd, const double2 *xy, const BatchCoef &cg, unsigned *bad)
{
    if (n <= 0) return hipSuccess;
    k_batch_coef_check<<<(n + 255) / 256, 256, 0, st>>>(n, row_ptr, col, kd, xy, cg, bad);
    return hipGetLastError();
}

// Workgroups of the batched kernel (block, group, recompute; the dense build
// for block 1024) that fit on one CU at once.
int batch_resident_per_cu(int block, int group, bool recompute, bool dense)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define PD_OCC(BT, GV)                                                                                      \
    if (block == BT && group == GV) {                                                                   \
        const void *k = recompute ? reinterpret_cast<const void *>(k_batch_adr<BT, GV, true>)          \
                                  : reinterpret_cast<const void *>(k_batch_adr<BT, GV, false>);         \
        if (BT == 1024 && dense)                                                                        \
            k = recompute ? reinterpret_cast<const void *>(k_batch_adr_dense<GV, true>)                 \
                          : reinterpret_cast<const void *>(k_batch_adr_dense<GV, false>);               \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, BT, 0);                                 \
    }
    PD_OCC(256, 1) PD_OCC(256, 4) PD_OCC(256, 8) PD_OCC(256, 16) PD_OCC(256, 32)
    PD_OCC(512, 1) PD_OCC(512, 4) PD_OCC(512, 8) PD_OCC(512, 16) PD_OCC(512, 32)
    PD_OCC(1024, 1) PD_OCC(1024, 4) PD_OCC(1024, 8) PD_OCC(1024, 16) PD_OCC(1024, 32)
#undef PD_OCC
    return e == hipSuccess && n > 0 ? n : 1;
}

hipError_t launch_batch_adr(hipStream_t st, int grid, int block, int group, bool dense, const BatchView &B,
                            double du, double tol, int max_iters, unsigned *queue, int32_t *iters_out,
                            double *resid_out, const BatchGroups &gr)
{
    // W must divide the grid and the damping tree (block/64 waves)
    if (gr.parts < 1 || gr.parts > kMaxParts || (gr.parts & (gr.parts - 1)) || block / kMinPartLanes < gr.parts ||
        grid % gr.parts || (gr.parts > 1 && (!gr.bar || !gr.spec || !gr.sum || !gr.abort || !gr.flag)) ||
        (gr.xcd && grid % (8 * gr.parts)) || (B.pause_below > 0 && (!B.active || !B.paused)) ||
        (B.recompute && !B.xy))
        return hipErrorInvalidValue;
    return B.recompute
               ? batch_b<true>(st, grid, block, group, dense, B, du, tol, max_iters, queue, iters_out, resid_out, gr)
               : batch_b<false>(st, grid, block, group, dense, B, du, tol, max_iters, queue, iters_out, resid_out,
                                gr);
}

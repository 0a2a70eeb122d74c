// mixed_bond.h -- device helpers shared by the mixed-precision bond kernels
// (gathering kernels in k_force_mixed.hip, LDS-staged tile kernels in
// k_force_lds.hip): particle records, FP32 bond geometry, the FP64 parity
// stretch, masked bond slots and the break decision.
#pragma once
#include "lanes.h"
#include "pd_internal.h"

namespace {

__device__ __forceinline__ float frsq(float v) { return __builtin_amdgcn_rsqf(v); }
__device__ __forceinline__ float frcp(float v) { return __builtin_amdgcn_rcpf(v); }

// Workgroups are handed to the 8 XCDs round-robin; give XCD k the contiguous
// block range [k*per, (k+1)*per).  gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_block()
{
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
}

// A particle record (pd_internal.h, kRecW words) in registers.
struct Rec {
    int32_t ix, iy, iz;  // quantized reference position (modulo 2^32)
    float ux, uy, uz;    // displacement, FP32
};

struct RecQ {
    Rec r;
    float q, m;  // A theta/m and B/m of the LPS passes
};

__device__ __forceinline__ RecQ rec_of(uint4 a, uint4 b)
{
    return RecQ{Rec{(int32_t)a.x, (int32_t)a.y, (int32_t)a.z, __uint_as_float(b.x), __uint_as_float(b.y),
                    __uint_as_float(b.z)},
                __uint_as_float(a.w), __uint_as_float(b.w)};
}

__device__ __forceinline__ RecQ load_rec(const uint32_t *__restrict__ rec, int c)
{
    const uint4 *p = reinterpret_cast<const uint4 *>(rec + (size_t)c * kRecW);
    return rec_of(p[0], p[1]);
}

// Largest displacement component of a record.
__device__ __forceinline__ float umag(const Rec &r)
{
    return fmaxf(fmaxf(fabsf(r.ux), fabsf(r.uy)), fabsf(r.uz));
}

__device__ __forceinline__ bool bit_of(uint32_t word, int32_t k) { return (word >> (k & 31)) & 1u; }

__device__ __forceinline__ void bond_break(uint32_t *fail, int32_t k)
{
    atomicAnd(&fail[k >> 5], ~(1u << (k & 31)));
}

// Per-lane sums accumulate the FP32 bond terms in FP64: every FP32 term is
// exact in FP64, so the running sum carries 29 guard bits (a compensated sum
// tighter than Neumaier's in FP32).  The group sum reduces those FP64
// partials over xor shuffles and rounds once to FP32.
template <int G>
__device__ __forceinline__ float group_sum(double a, double scale = 1.0)
{
    // Butterfly: every lane adds the same pairs in the same order and ends
    // with the same sum.  Partners within 16 lanes move over DPP.
#pragma unroll
    for (int off = G / 2; off > 8; off >>= 1) a += __shfl_xor(a, off, 64);
    if constexpr (G >= 16) a += lanes::xor_lane<8>(a);
    if constexpr (G >= 8) a += lanes::xor_lane<4>(a);
    if constexpr (G >= 4) a += lanes::xor_lane<2>(a);
    if constexpr (G >= 2) a += lanes::xor_lane<1>(a);
    return (float)(a * scale);
}

// Geometry of one bond in FP32 from the records, in units of the record
// quantum q: xi from the difference of the quantized positions (exact as an
// integer, rounded once to FP32), eta from the FP32 displacements.  Every fused
// multiply-add is written out (the library is built without contraction), so
// every inlined copy of a bond body rounds the same way, whichever kernel,
// tile, lane count or unrolling it sits in.  The FP64 parity stretch below
// keeps the separate roundings of the oracle.
template <int DIM>
struct BondGeo {
    float dx, dy, dz;   // xi + eta
    float num;          // |xi+eta|^2 - |xi|^2 = eta.(2 xi + eta)
    float rL;           // 1/|xi+eta|
    float L2;           // |xi+eta|^2
    float xi2;          // |xi|^2
    // |xi+eta| + |xi| for a reference length L0
    __device__ __forceinline__ float len_sum(float L0) const { return __builtin_fmaf(L2, rL, L0); }
};

// c - i of two quantized coordinates, in units of q: the difference modulo
// 2^32 is the exact difference of the unreduced integers (|xi| / q < 2^31).
__device__ __forceinline__ float qdiff(int32_t i, int32_t c)
{
    return (float)(int32_t)((uint32_t)c - (uint32_t)i);
}

template <int DIM>
__device__ __forceinline__ BondGeo<DIM> geo(const Rec &ri, const Rec &rc)
{
    const float ex = qdiff(ri.ix, rc.ix), ey = qdiff(ri.iy, rc.iy);
    const float hx = rc.ux - ri.ux, hy = rc.uy - ri.uy;
    BondGeo<DIM> g;
    g.dx = ex + hx;
    g.dy = ey + hy;
    float L2 = __builtin_fmaf(g.dy, g.dy, g.dx * g.dx);
    g.num = __builtin_fmaf(hy, __builtin_fmaf(2.f, ey, hy), hx * __builtin_fmaf(2.f, ex, hx));
    g.xi2 = __builtin_fmaf(ey, ey, ex * ex);
    if (DIM == 3) {
        const float ez = qdiff(ri.iz, rc.iz), hz = rc.uz - ri.uz;
        g.dz = ez + hz;
        L2 = __builtin_fmaf(g.dz, g.dz, L2);
        g.num = __builtin_fmaf(hz, __builtin_fmaf(2.f, ez, hz), g.num);
        g.xi2 = __builtin_fmaf(ez, ez, g.xi2);
    } else {
        g.dz = 0.f;
    }
    g.L2 = L2;
    g.rL = frsq(L2);
    return g;
}

// Bond coefficients recomputed from the bond's reference length instead of
// streamed from HBM (used when every stored coefficient equals this
// recomputation to FP32 accuracy, which the engine checks once per set-up):
// the partial-volume factor of Surf_Corr.cu:27-32, and for LPS the influence
// weight exp(-xi^2/delta^2) of Force_State.cu:317.  coef = fac [* w], the
// weight when W (the LPS passes; gc.k2 is then nonzero); the constant factor
// gc.scale of the stored coefficients is carried by each pass's constants.
template <bool W>
__device__ __forceinline__ void geom_coef(float xi2, const GeomCoef &gc, float &L0, float &coef, float &rL0)
{
    rL0 = frsq(xi2);
    L0 = xi2 * rL0;
    // 1 up to delta - dx/2, (hi - |xi|)/dx up to hi = delta + dx/2, 0 beyond
    coef = __builtin_amdgcn_fmed3f(__builtin_fmaf(L0, -gc.inv_dx, gc.hi_dx), 0.f, 1.f);
    if (W) coef *= __builtin_amdgcn_exp2f(gc.k2 * xi2);
}

// FP64 state of one particle, read from the SoA arrays for the FP64 stretch.
struct Rec64 {
    double x0, x1, x2, u0, u1, u2;
};

template <int DIM>
__device__ __forceinline__ Rec64 rec64(const MixCtx &mc, int i)
{
    return Rec64{mc.x.p[0][i], mc.x.p[1][i], DIM == 3 ? mc.x.p[2][i] : 0.0,
                 mc.u.p[0][i], mc.u.p[1][i], DIM == 3 ? mc.u.p[2][i] : 0.0};
}

// FP64 stretch with the parity expressions (PMB order (x_c+u_c)-x_i-u_i for
// every component, LPS-3D order for z when lps3d).
template <int DIM, bool LPS3Z>
__device__ __forceinline__ double stretch64(const Rec64 &ri, const Rec64 &rc)
{
    double d0 = rc.x0 + rc.u0 - ri.x0 - ri.u0, d1 = rc.x1 + rc.u1 - ri.x1 - ri.u1;
    double a0 = rc.x0 - ri.x0, a1 = rc.x1 - ri.x1;
    double r2 = d0 * d0 + d1 * d1, q2 = a0 * a0 + a1 * a1;
    if (DIM == 3) {
        double d2 = LPS3Z ? rc.x2 - ri.x2 + rc.u2 - ri.u2 : rc.x2 + rc.u2 - ri.x2 - ri.u2;
        double a2 = rc.x2 - ri.x2;
        r2 = r2 + d2 * d2;
        q2 = q2 + a2 * a2;
    }
    double id = sqrt(q2);
    return (sqrt(r2) - id) / id;
}

// One bond slot of a lane: neighbour index, |xi| and bond coefficient, and a
// mask that is all ones when the slot holds a bond that acts in this pass.
// Slots past the end of the row point at the particle itself with a zero
// mask, so every lane runs the same branch-free arithmetic.
struct Slot {
    int c;
    float L0, coef;
    uint32_t mask;
};

template <bool GC>
__device__ __forceinline__ Slot fetch_slot(int32_t k, int32_t ke, int self, const int32_t *__restrict__ col,
                                           const float2 *__restrict__ bnd, const uint32_t *__restrict__ fail,
                                           bool need_intact)
{
    const bool valid = k < ke;
    const int32_t kk = valid ? k : 0;
    const int c = col[kk];
    const float2 b = GC ? make_float2(1.f, 0.f) : bnd[kk];
    const uint32_t w = fail[kk >> 5];
    const bool act = valid && (!need_intact || ((w >> (k & 31)) & 1u));
    return Slot{valid ? c : self, valid ? b.x : 1.f, b.y, act ? 0xffffffffu : 0u};
}

__device__ __forceinline__ float masked(float v, uint32_t mask)
{
    return __uint_as_float(__float_as_uint(v) & mask);
}

// Break decision of bond k between row i (record ri) and neighbour cl (a
// particle index, or with MAP a position in the tile's staged list
// map[sb ...]; record rc), s its FP32 stretch, xi2 its |xi|^2, ew = wpos +
// k_u M the records' part of the band (mix_window in pd_internal.h).  lidx,
// when given, is the LDS tile kernels' per-bond local index word whose bit 15
// mirrors the fail bit; it is cleared with the fail bit.  The FP64 state is
// read only for bonds inside the inner band.
template <int DIM, bool LPS3Z, bool MAP>
__device__ __forceinline__ void check_break(int i, int cl, const int32_t *map, int64_t sb, float s, const Slot &sl,
                                            int32_t k, float ew, const Rec &ri, const Rec &rc, float xi2,
                                            const MixCtx &mc, double s0, double ymin, uint32_t *fail,
                                            uint16_t *lidx = nullptr)
{
    const float wo = mc.wfloor + ew;
    if (sl.mask && s >= mc.s0f - wo) {
        bool brk = s > mc.s0f + wo;  // beyond the outer band: broken, no more work
        if (!brk) {
            const float hx = rc.ux - ri.ux, hy = rc.uy - ri.uy, hz = rc.uz - ri.uz;
            const float r = __builtin_sqrtf(__builtin_fmaf(hz, hz, __builtin_fmaf(hy, hy, hx * hx)) * frcp(xi2));
            const float wt = fmaxf(mc.wn, 0x1p-24f * r * __builtin_fmaf(16.f, r, 32.f)) + ew;
            if (s < mc.s0f - wt) return;
            brk = s > mc.s0f + wt;
            if (!brk) brk = stretch64<DIM, LPS3Z>(rec64<DIM>(mc, i), rec64<DIM>(mc, MAP ? map[sb + cl] : cl)) > s0;
        }
        if (DIM == 3 && brk && ymin > -INFINITY) {
            const int c = MAP ? map[sb + cl] : cl;
            brk = mc.x.p[1][i] >= ymin && mc.x.p[1][c] >= ymin;
        }
        if (brk) {
            bond_break(fail, k);
            if (lidx) atomicAnd(reinterpret_cast<uint32_t *>(lidx) + (k >> 1), ~(0x8000u << (16 * (k & 1))));
        }
    }
}

// The records' part of the break band of a bond whose ends have
// displacement components at most mi, mj.
__device__ __forceinline__ float bond_window(const MixCtx &mc, float mi, float mj)
{
    return __builtin_fmaf(mc.k_u, fmaxf(mi, mj), mc.wpos);
}

} // namespace

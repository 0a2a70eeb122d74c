// k_force_lds.hip -- LDS-staged tile variants of the mixed-precision PMB and
// LPS passes, and the builder of the per-tile staged particle lists they use.
//
// Tiles.  The tiled processing order groups the rows into small spatial
// tiles (order[tile_start[t] .. tile_start[t+1])).  The staged set of a tile
// is the sorted set of distinct neighbours of its rows; its particle records
// are copied into LDS once per pass (two consecutive lanes move one 32-byte
// record, so the copy is a coalesced stream of whole records), and every bond
// of the tile then reads its neighbour from LDS through a 16-bit local index
// lidx[k] instead of gathering 32 bytes from the L2 per bond.  Bit 15 of
// lidx[k] mirrors the bond's fail bit (1 = intact), so one 2-byte load per
// bond serves both; the fail words stay the bonds' state, breaks clear both,
// and k_lidx_sync copies the fail words into bit 15 after any other change.
// The bond stream shrinks to lidx (2 B) + {|xi|, coefficient} (8 B, none
// when the coefficients are recomputed).
//
// Arithmetic, bond slots, the group reduction and the break decision are the
// ones of the gathering kernels (mixed_bond.h) with the same bond-to-lane
// assignment, so both paths produce the same bits.  Only the break band
// differs: here one bound per row from the largest displacement of the
// tile's staged records, there one per bond; either band decides every
// break by the FP64 stretch wherever the FP32 one could err.
#include "mixed_bond.h"

#include <climits>

namespace {

constexpr int kHashSlots = 8192;  // per-tile open-addressing table of the builder
constexpr int kStageCap = 4096;   // largest staged set the builder handles (< 2^15)
constexpr uint32_t kIntact = 0x8000u;  // bit 15 of a local index word

// lowbias32 integer hash (C. Wellons, public domain)
__device__ __forceinline__ uint32_t hash32(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352dU;
    v ^= v >> 15;
    v *= 0x846ca68bU;
    v ^= v >> 16;
    return v;
}

// One workgroup per tile.  Phase 0 counts the distinct neighbours (count[t],
// -1 when the tile exceeds kStageCap); phase 1 writes the sorted staged list
// at stage_ptr[t] and the local index of every bond of the tile.
__global__ __launch_bounds__(1024) void k_stage_tiles(int phase, const int32_t *__restrict__ tile_start,
                                                      const int32_t *__restrict__ order,
                                                      const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ col, int32_t *count,
                                                      const int64_t *__restrict__ stage_ptr, int32_t *stage_idx,
                                                      uint16_t *lidx)
{
    __shared__ int32_t table[kHashSlots];
    __shared__ int32_t list[kStageCap];
    __shared__ int n_unique, n_full, n_list;
    const int t = blockIdx.x;
    for (int s = threadIdx.x; s < kHashSlots; s += blockDim.x) table[s] = -1;
    if (threadIdx.x == 0) {
        n_unique = 0;
        n_full = 0;
        n_list = 0;
    }
    __syncthreads();
    const int r0 = tile_start[t], r1 = tile_start[t + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int r = r0 + wave; r < r1; r += nw) {
        const int i = order[r];
        for (int32_t k = rp[i] + lane; k < rp[i + 1]; k += 64) {
            const int c = col[k];
            uint32_t h = hash32((uint32_t)c) & (kHashSlots - 1);
            int probe = 0;
            for (; probe < kHashSlots; ++probe) {
                const int old = atomicCAS(&table[h], -1, c);
                if (old == -1) {
                    atomicAdd(&n_unique, 1);
                    break;
                }
                if (old == c) break;
                h = (h + 1) & (kHashSlots - 1);
            }
            if (probe == kHashSlots) atomicAdd(&n_full, 1);
        }
    }
    __syncthreads();
    const int nu = n_unique;
    const bool fits = n_full == 0 && nu <= kStageCap;
    if (phase == 0) {
        if (threadIdx.x == 0) count[t] = fits ? nu : -1;
        return;
    }
    if (!fits) return;  // phase 1 runs only after every tile fitted
    for (int s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
        const int v = table[s];
        if (v >= 0) list[atomicAdd(&n_list, 1)] = v;
    }
    int p2 = 1;
    while (p2 < nu) p2 <<= 1;
    __syncthreads();
    for (int s = nu + threadIdx.x; s < p2; s += blockDim.x) list[s] = INT_MAX;
    __syncthreads();
    for (int kk = 2; kk <= p2; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int s = threadIdx.x; s < p2; s += blockDim.x) {
                const int q = s ^ j;
                if (q > s) {
                    const bool up = (s & kk) == 0;
                    const int a = list[s], b = list[q];
                    if ((a > b) == up) {
                        list[s] = b;
                        list[q] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int64_t base = stage_ptr[t];
    for (int s = threadIdx.x; s < nu; s += blockDim.x) stage_idx[base + s] = list[s];
    for (int r = r0 + wave; r < r1; r += nw) {
        const int i = order[r];
        for (int32_t k = rp[i] + lane; k < rp[i + 1]; k += 64) {
            const int c = col[k];
            int lo = 0, hi = nu - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (list[mid] < c)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            lidx[k] = (uint16_t)lo;
        }
    }
}

// Copy every bond's fail bit into bit 15 of its local index word.
__global__ void k_lidx_sync(int64_t nb, const uint32_t *__restrict__ fail, uint16_t *lidx)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const uint32_t bit = (fail[k >> 5] >> (k & 31)) & 1u;
    lidx[k] = (uint16_t)((lidx[k] & (kIntact - 1)) | (bit << 15));
}

// Tiles handed to the 8 XCDs in contiguous ranges (gridDim.x is a multiple of 8).
__device__ __forceinline__ int xcd_tile() { return xcd_block(); }

// Copy the tile's staged records into LDS: 2 lanes per 32-byte record, the
// record stride in LDS being the record itself.  With UMAX the largest
// displacement component of the staged records is left in *umax (for the
// break band of the tile's bonds).
template <bool UMAX>
__device__ __forceinline__ void stage_records(uint4 *lds, const uint32_t *__restrict__ rec,
                                              const int32_t *__restrict__ stage_idx, int64_t sb, int ns,
                                              unsigned *umax)
{
    float m = 0.f;
    for (int q = threadIdx.x; q < 2 * ns; q += blockDim.x) {
        const int j = stage_idx[sb + (q >> 1)];
        const uint4 v = reinterpret_cast<const uint4 *>(rec + (size_t)j * kRecW)[q & 1];
        lds[q] = v;
        if (UMAX && (q & 1))
            m = fmaxf(m, fmaxf(fmaxf(fabsf(__uint_as_float(v.x)), fabsf(__uint_as_float(v.y))),
                               fabsf(__uint_as_float(v.z))));
    }
    if (UMAX) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0) atomicMax(umax, __float_as_uint(m));  // non-negative: ordered as integers
    }
    __syncthreads();
}

__device__ __forceinline__ RecQ lds_rec(const uint4 *lds, int l) { return rec_of(lds[2 * l], lds[2 * l + 1]); }

// 2-byte load at a 32-bit byte offset from a uniform base (no 64-bit
// address arithmetic per bond; Nb < 2^31).
__device__ __forceinline__ uint32_t load_lidx(const uint16_t *lidx, uint32_t k)
{
    return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(lidx) + 2u * k);
}

// A slot past the end of the row (k >= ke) reads the row's first bond kb
// instead: its local index is then one of this tile's, so the neighbour
// index needs no select (the mask is zero either way).  kb is in bounds of
// lidx even for an empty row (lidx has Nb + 1 entries or more), whose slots
// are never dereferenced.
template <bool GC>
__device__ __forceinline__ Slot fetch_lslot(int32_t k, int32_t kb, int32_t ke, const uint16_t *lidx,
                                            const float2 *__restrict__ bnd, bool need_intact)
{
    const bool valid = k < ke;
    const uint32_t kk = valid ? (uint32_t)k : (uint32_t)kb;
    const uint32_t v = load_lidx(lidx, kk);
    const float2 b = GC ? make_float2(1.f, 0.f) : bnd[valid ? (uint32_t)k : 0u];
    const bool act = valid && (!need_intact || (v & kIntact));
    return Slot{(int)(v & (kIntact - 1)), valid ? b.x : 1.f, b.y, act ? 0xffffffffu : 0u};
}

// Row walk over LDS records; same lane/bond assignment as the gathering walk.
// The slots are fetched one bond pair ahead of their use.
#define PD_LDS_ROW_WALK(NEED_INTACT, BODY)                                                       \
    {                                                                                            \
        const int32_t kb = rp[i], ke = rp[i + 1];                                                \
        int32_t k = kb + lane;                                                                   \
        Slot sa = fetch_lslot<GC>(k, kb, ke, lidx, bnd, NEED_INTACT);                                \
        Slot sb = fetch_lslot<GC>(k + G, kb, ke, lidx, bnd, NEED_INTACT);                            \
        while (k < ke) {                                                                         \
            const RecQ ra = lds_rec(lds, sa.c), rb = lds_rec(lds, sb.c);                         \
            const Slot na = fetch_lslot<GC>(k + 2 * G, kb, ke, lidx, bnd, NEED_INTACT);              \
            const Slot nb = fetch_lslot<GC>(k + 3 * G, kb, ke, lidx, bnd, NEED_INTACT);              \
            BODY(ra, sa, k);                                                                     \
            BODY(rb, sb, k + G);                                                                 \
            sa = na;                                                                             \
            sb = nb;                                                                             \
            k += 2 * G;                                                                          \
        }                                                                                        \
    }

// sbase: the tile's first entry in stage_idx; ut: largest displacement
// component of its staged records (UMAX).
#define PD_TILE_PROLOGUE(UMAX)                                                                   \
    extern __shared__ uint4 lds[];                                                               \
    __shared__ unsigned umax_sh;                                                                 \
    if (stop && *stop) return;                                                                   \
    const int t = xcd_tile();                                                                    \
    if (t >= ntiles) return;                                                                     \
    if (UMAX && threadIdx.x == 0) umax_sh = 0u;                                                  \
    if (UMAX) __syncthreads();                                                                   \
    const int64_t sbase = stage_ptr[t];                                                          \
    stage_records<UMAX>(lds, rec, stage_idx, sbase, (int)(stage_ptr[t + 1] - sbase), &umax_sh);  \
    const float ut = UMAX ? __uint_as_float(umax_sh) : 0.f;                                      \
    (void)ut;                                                                                    \
    const int lane = threadIdx.x & (G - 1);                                                      \
    const int r1 = tile_start[t + 1];

template <int DIM, int G, bool GC>
__device__ __forceinline__ void pmb_lds_body(const int *__restrict__ stop, int ntiles,
                                             const int32_t *__restrict__ tile_start,
                                             const int32_t *__restrict__ order,
                                             const int64_t *__restrict__ stage_ptr,
                                             const int32_t *__restrict__ stage_idx,
                                             const int32_t *__restrict__ rp, uint16_t *lidx,
                                             const float2 *__restrict__ bnd, uint32_t *fail,
                                             const uint32_t *__restrict__ rec, F3 f, double s0, const MixCtx &mc,
                                             double ymin, GeomCoef gc)
{
    PD_TILE_PROLOGUE(true)
    for (int r = tile_start[t] + (int)threadIdx.x / G; r < r1; r += (int)blockDim.x / G) {
        const int i = order[r];
        const Rec ri = load_rec(rec, i).r;
        const float ew = bond_window(mc, umag(ri), ut);
        double ax = 0.0, ay = 0.0, az = 0.0;
#define PMB_BODY(R, S, K)                                                                        \
    {                                                                                            \
        const BondGeo<DIM> g = geo<DIM>(ri, R.r);                                          \
        float L0 = S.L0, cf = S.coef, rL0 = 0.f;                                                 \
        if (GC) geom_coef<false>(g.xi2, gc, L0, cf, rL0);                                        \
        const float s = g.num * frcp(L0 * g.len_sum(L0));                                        \
        const float tb = masked(cf * s * g.rL, S.mask);                                          \
        ax += (double)(tb * g.dx);                                                               \
        ay += (double)(tb * g.dy);                                                               \
        if (DIM == 3) az += (double)(tb * g.dz);                                                 \
        check_break<DIM, false, true>(i, S.c, stage_idx, sbase, s, S, K, ew, ri, R.r, g.xi2, mc, s0, ymin, fail, \
                                      lidx);                                                     \
    }
        PD_LDS_ROW_WALK(true, PMB_BODY)
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
}

template <int DIM, int G, bool GC>
__global__ __launch_bounds__(1024) void k_lps_theta_lds(const int *__restrict__ stop, int ntiles,
                                                        const int32_t *__restrict__ tile_start,
                                                        const int32_t *__restrict__ order,
                                                        const int64_t *__restrict__ stage_ptr,
                                                        const int32_t *__restrict__ stage_idx,
                                                        const int32_t *__restrict__ rp,
                                                        const uint16_t *__restrict__ lidx,
                                                        const float2 *__restrict__ bnd, uint32_t *rec,
                                                        const float *__restrict__ minv32, float ctheta, float A,
                                                        float B, int skip_broken, float *theta32, GeomCoef gc)
{
    PD_TILE_PROLOGUE(false)
    const bool need_intact = skip_broken != 0;
    // theta/m of this pass is written into the records after every tile has
    // staged them: the position and displacement words are read here, the
    // writes go to words 3 and 7 of rows this tile owns, and the next pass
    // reads them.
    for (int r = tile_start[t] + (int)threadIdx.x / G; r < r1; r += (int)blockDim.x / G) {
        const int i = order[r];
        const Rec ri = load_rec(rec, i).r;
        double acc = 0.0;
#define THETA_BODY(R, S, K)                                                                      \
    {                                                                                            \
        const BondGeo<DIM> g = geo<DIM>(ri, R.r);                                             \
        float L0 = S.L0, cf = S.coef, rL0 = 0.f;                                                 \
        if (GC) geom_coef<true>(g.xi2, gc, L0, cf, rL0);                                         \
        const float e = g.num * frcp(g.len_sum(L0)); /* |xi+eta| - |xi| */                       \
        acc += (double)masked(cf * L0 * e, S.mask);                                              \
    }
        PD_LDS_ROW_WALK(need_intact, THETA_BODY)
#undef THETA_BODY
        const float th = group_sum<G>(acc);
        if (lane == 0) {
            // recomputed coefficients leave out vol: the constants carry it
            const float cth = GC ? ctheta * gc.scale : ctheta, As = GC ? A * gc.scale : A, Bs = GC ? B * gc.scale : B;
            const float rm = minv32[i];
            const float theta = cth * rm * th;
            theta32[i] = theta;
            rec[(size_t)i * kRecW + 3] = __float_as_uint(As * (theta * rm));
            rec[(size_t)i * kRecW + 7] = __float_as_uint(Bs * rm);
        }
    }
}

// amdgpu_waves_per_eu(6): at most 80 VGPRs, so that three 512-thread
// workgroups share a CU (with 64-byte records this measured 1.10 ms instead
// of 1.26 ms per LPS force pass on the 4M-particle block).
template <int DIM, int G, bool GC>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(GC ? 6 : 1))) void k_pmb_lds(
    const int *__restrict__ stop, int ntiles, const int32_t *__restrict__ tile_start,
    const int32_t *__restrict__ order, const int64_t *__restrict__ stage_ptr, const int32_t *__restrict__ stage_idx,
    const int32_t *__restrict__ rp, uint16_t *lidx, const float2 *__restrict__ bnd, uint32_t *fail,
    const uint32_t *__restrict__ rec, F3 f, double s0, MixCtx mc, double ymin, GeomCoef gc)
{
    pmb_lds_body<DIM, G, GC>(stop, ntiles, tile_start, order, stage_ptr, stage_idx, rp, lidx, bnd, fail, rec, f, s0,
                             mc, ymin, gc);
}

template <int DIM, int G, bool GC>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(GC ? 6 : 1))) void k_lps_force_lds(
    const int *__restrict__ stop, int ntiles, const int32_t *__restrict__ tile_start,
    const int32_t *__restrict__ order, const int64_t *__restrict__ stage_ptr, const int32_t *__restrict__ stage_idx,
    const int32_t *__restrict__ rp, uint16_t *lidx, const float2 *__restrict__ bnd, uint32_t *fail,
    const uint32_t *__restrict__ rec, double s0, MixCtx mc, double ymin, F3 f, GeomCoef gc)
{
    PD_TILE_PROLOGUE(true)
    for (int r = tile_start[t] + (int)threadIdx.x / G; r < r1; r += (int)blockDim.x / G) {
        const int i = order[r];
        const RecQ qi = load_rec(rec, i);
        const Rec &ri = qi.r;
        const float ew = bond_window(mc, umag(ri), ut);
        double ax = 0.0, ay = 0.0, az = 0.0;
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
        check_break<DIM, true, true>(i, S.c, stage_idx, sbase, s, S, K, ew, ri, R.r, g.xi2, mc, s0, ymin, fail, \
                                     lidx);                                                      \
    }
        PD_LDS_ROW_WALK(true, LPSF_BODY)
#undef LPSF_BODY
        const float sx = group_sum<G>(ax), sy = group_sum<G>(ay);
        const float sz = DIM == 3 ? group_sum<G>(az) : 0.f;
        if (lane == 0) {
            f.p[0][i] = sx;
            f.p[1][i] = sy;
            if (DIM == 3) f.p[2][i] = sz;
        }
    }
}

inline unsigned tiles_grid(int ntiles) { return (unsigned)((ntiles + 7) / 8 * 8); }

template <class K>
hipError_t lds_attr(K kernel, size_t bytes)
{
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

#define PD_LSWITCH(FN, DIM, G, GC, GCV, ...)              \
    switch (G) {                                          \
    case 4: return FN<DIM, 4, GC>(__VA_ARGS__, GCV);      \
    case 8: return FN<DIM, 8, GC>(__VA_ARGS__, GCV);      \
    case 16: return FN<DIM, 16, GC>(__VA_ARGS__, GCV);    \
    case 32: return FN<DIM, 32, GC>(__VA_ARGS__, GCV);    \
    case 64: return FN<DIM, 64, GC>(__VA_ARGS__, GCV);    \
    default: return hipErrorInvalidValue;                 \
    }

// gc != nullptr: recompute the bond coefficients from |xi| (no coefficient stream)
#define PD_DISPATCH_LDS(FN, DIM, G, GCP, ...)                                \
    if (GCP) {                                                               \
        PD_LSWITCH(FN, DIM, G, true, *(GCP), __VA_ARGS__)                    \
    } else {                                                                 \
        PD_LSWITCH(FN, DIM, G, false, GeomCoef{}, __VA_ARGS__)               \
    }

template <int DIM, int G, bool GC>
hipError_t pmb_lds_g(hipStream_t s, const int *stop, const LdsTiles &T, const int32_t *order, const int32_t *rp,
                     const float2 *bnd, uint32_t *fail, const uint32_t *rec, F3 f, double s0, const MixCtx &mc,
                     double ymin, GeomCoef gc)
{
    const size_t bytes = (size_t)T.max_staged * lds_record_bytes();
    hipError_t e = lds_attr(k_pmb_lds<DIM, G, GC>, bytes);
    if (e != hipSuccess) return e;
    k_pmb_lds<DIM, G, GC><<<tiles_grid(T.ntiles), T.block, bytes, s>>>(stop, T.ntiles, T.tile_start, order,
                                                                      T.stage_ptr, T.stage_idx, rp, T.lidx, bnd,
                                                                      fail, rec, f, s0, mc, ymin, gc);
    return hipGetLastError();
}

template <int DIM, int G, bool GC>
hipError_t theta_lds_g(hipStream_t s, const int *stop, const LdsTiles &T, const int32_t *order, const int32_t *rp,
                       const float2 *bnd, uint32_t *rec, const float *minv32, float ctheta, float A, float B,
                       int skip_broken, float *theta32, GeomCoef gc)
{
    const size_t bytes = (size_t)T.max_staged * lds_record_bytes();
    hipError_t e = lds_attr(k_lps_theta_lds<DIM, G, GC>, bytes);
    if (e != hipSuccess) return e;
    k_lps_theta_lds<DIM, G, GC><<<tiles_grid(T.ntiles), T.block, bytes, s>>>(stop, T.ntiles, T.tile_start, order,
                                                                          T.stage_ptr, T.stage_idx, rp, T.lidx, bnd,
                                                                          rec, minv32, ctheta, A, B, skip_broken,
                                                                          theta32, gc);
    return hipGetLastError();
}

template <int DIM, int G, bool GC>
hipError_t lpsf_lds_g(hipStream_t s, const int *stop, const LdsTiles &T, const int32_t *order, const int32_t *rp,
                      const float2 *bnd, uint32_t *fail, const uint32_t *rec, double s0, const MixCtx &mc,
                      double ymin, F3 f, GeomCoef gc)
{
    const size_t bytes = (size_t)T.max_staged * lds_record_bytes();
    hipError_t e = lds_attr(k_lps_force_lds<DIM, G, GC>, bytes);
    if (e != hipSuccess) return e;
    k_lps_force_lds<DIM, G, GC><<<tiles_grid(T.ntiles_int), T.block2, bytes, s>>>(
        stop, T.ntiles_int, T.tile_start, order, T.stage_ptr, T.stage_idx, rp, T.lidx, bnd, fail, rec, s0, mc, ymin,
        f, gc);
    return hipGetLastError();
}

} // namespace

hipError_t launch_stage_tiles(hipStream_t s, int phase, int ntiles, const int32_t *tile_start, const int32_t *order,
                              const int32_t *rp, const int32_t *col, int32_t *count, const int64_t *stage_ptr,
                              int32_t *stage_idx, uint16_t *lidx)
{
    if (ntiles <= 0) return hipSuccess;
    k_stage_tiles<<<ntiles, 1024, 0, s>>>(phase, tile_start, order, rp, col, count, stage_ptr, stage_idx, lidx);
    return hipGetLastError();
}

hipError_t launch_lidx_sync(hipStream_t s, int64_t nb, const uint32_t *fail, uint16_t *lidx)
{
    if (nb <= 0) return hipSuccess;
    k_lidx_sync<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(nb, fail, lidx);
    return hipGetLastError();
}

int stage_cap() { return kStageCap; }
int lds_record_bytes() { return 4 * kRecW; }

hipError_t launch_pmb_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group, const int32_t *order,
                          const int32_t *rp, const float2 *bnd, uint32_t *fail, const uint32_t *rec, F3 f, double s0,
                          const MixCtx &mc, double ymin, const GeomCoef *gc)
{
    if (T.ntiles <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_LDS(pmb_lds_g, 3, group, gc, s, stop, T, order, rp, bnd, fail, rec, f, s0, mc, ymin)
    } else {
        PD_DISPATCH_LDS(pmb_lds_g, 2, group, gc, s, stop, T, order, rp, bnd, fail, rec, f, s0, mc, ymin)
    }
}

hipError_t launch_lps_theta_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group,
                                const int32_t *order, const int32_t *rp, const float2 *bnd, uint32_t *rec,
                                const float *minv32, float ctheta, float A, float B, int skip_broken, float *theta32,
                                const GeomCoef *gc)
{
    if (T.ntiles <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_LDS(theta_lds_g, 3, group, gc, s, stop, T, order, rp, bnd, rec, minv32, ctheta, A, B,
                        skip_broken, theta32)
    } else {
        PD_DISPATCH_LDS(theta_lds_g, 2, group, gc, s, stop, T, order, rp, bnd, rec, minv32, ctheta, A, B,
                        skip_broken, theta32)
    }
}

hipError_t launch_lps_force_lds(hipStream_t s, const int *stop, const LdsTiles &T, int dim, int group,
                                const int32_t *order, const int32_t *rp, const float2 *bnd, uint32_t *fail,
                                const uint32_t *rec, double s0, const MixCtx &mc, double ymin, F3 f,
                                const GeomCoef *gc)
{
    if (T.ntiles_int <= 0) return hipSuccess;
    if (dim == 3) {
        PD_DISPATCH_LDS(lpsf_lds_g, 3, group, gc, s, stop, T, order, rp, bnd, fail, rec, s0, mc, ymin, f)
    } else {
        PD_DISPATCH_LDS(lpsf_lds_g, 2, group, gc, s, stop, T, order, rp, bnd, fail, rec, s0, mc, ymin, f)
    }
}

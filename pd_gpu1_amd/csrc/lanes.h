// lanes.h -- a lane's value broadcast to its group of 4 or 8 consecutive
// lanes of a wave64 over DPP (data-parallel primitives: a VALU move that
// reads another lane's register), instead of __shfl, which goes through the
// LDS crossbar (ds_bpermute).  Used by the bond-order sums of the FP64 force
// kernels, whose dependent add chain waits on every broadcast, and by the
// butterfly row sums of the mixed kernels.  Every lane of
// the group must be active.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace lanes {

// GFX9 DPP controls: quad_perm [a, b, c, d] = a | b << 2 | c << 4 | d << 6;
// row_shl:n = 0x100 + n, row_shr:n = 0x110 + n (within rows of 16 lanes).
// bank_mask selects which of a row's four banks of four lanes are written.
template <int G, int L>
__device__ __forceinline__ int bcast(int v)
{
    static_assert((G == 4 || G == 8) && L >= 0 && L < G, "groups of 4 or 8 lanes");
    const int q = __builtin_amdgcn_update_dpp(0, v, (L & 3) * 0x55, 0xF, 0xF, false);  // lane L%4 of each quad
    if constexpr (G == 4)
        return q;
    else if constexpr (L < 4)
        return __builtin_amdgcn_update_dpp(q, q, 0x114, 0xF, 0xA, false);  // banks 1, 3 from banks 0, 2
    else
        return __builtin_amdgcn_update_dpp(q, q, 0x104, 0xF, 0x5, false);  // banks 0, 2 from banks 1, 3
}

template <int G, int L>
__device__ __forceinline__ double bcast(double v)
{
    const int2 w = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(bcast<G, L>(w.x), bcast<G, L>(w.y)));
}

// The value of lane ^ M, for M = 1, 2, 4 or 8 (row_ror:n = 0x120 + n; rotating
// a row of 16 by 8 pairs lane i with lane i ^ 8).  Both lanes of each pair
// must be active.
template <int M>
__device__ __forceinline__ int xor_lane(int v)
{
    static_assert(M == 1 || M == 2 || M == 4 || M == 8, "lane ^ 1, 2, 4 or 8");
    if constexpr (M == 1)
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    else if constexpr (M == 2)
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
    else if constexpr (M == 4) {
        const int t = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xF, 0x5, false);  // banks 0, 2 from lane + 4
        return __builtin_amdgcn_update_dpp(t, v, 0x114, 0xF, 0xA, false);         // banks 1, 3 from lane - 4
    } else
        return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
}

template <int M>
__device__ __forceinline__ double xor_lane(double v)
{
    const int2 w = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(xor_lane<M>(w.x), xor_lane<M>(w.y)));
}

// f(std::integral_constant<int, L>{}) for L = 0 .. N-1, in order.
template <int N, int L = 0>
struct Unroll {
    template <class F>
    __device__ __forceinline__ static void run(F &&f)
    {
        if constexpr (L < N) {
            f(std::integral_constant<int, L>{});
            Unroll<N, L + 1>::run(f);
        }
    }
};

} // namespace lanes

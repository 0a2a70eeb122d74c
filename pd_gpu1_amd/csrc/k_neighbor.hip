// k_neighbor.hip -- cell-list neighbour search emitting a per-particle CSR
// bond list with ascending rows.
//
// Replaces the O(N^2) all-pairs kernels kernel_find_neighbor_2D/3D
// (Neighbor.cu:7-40), whose atomicAdd slot append makes the slot order
// nondeterministic.  The predicate is the reference's, evaluated in FP64 with
// the CPU twin's operand order ((dx^2 + dy^2) + dz^2) < delta*delta
// (Neighbor.cu:51,72) and no fused multiply-add, so the bond set is
// bit-identical to the CPU search; rows are then sorted ascending, which is
// the CPU twin's slot order (Neighbor.cu:44-58).
//
// Pipeline: cell id per particle -> radix sort (cell, id) -> cell start table
// -> count pass -> exclusive scan (row_ptr) -> fill pass -> per-row bitonic
// sort in LDS (one wave per row).
#include "pd_internal.h"
#include <climits>

namespace {

__device__ __forceinline__ int cell_coord(double v, double lo, double cs, int n)
{
    int c = (int)floor((v - lo) / cs);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

struct Grid {
    double lo[3];
    double cs;
    int n[3];
};

__global__ void k_cell_ids(int dim, int N, CD3 x, Grid g, uint32_t *cell_of, uint32_t *idx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c0 = cell_coord(x.p[0][i], g.lo[0], g.cs, g.n[0]);
    int c1 = cell_coord(x.p[1][i], g.lo[1], g.cs, g.n[1]);
    int c2 = dim == 3 ? cell_coord(x.p[2][i], g.lo[2], g.cs, g.n[2]) : 0;
    cell_of[i] = (uint32_t)((c2 * g.n[1] + c1) * g.n[0] + c0);
    idx[i] = (uint32_t)i;
}

__global__ void k_cell_bounds(int N, const uint32_t *sorted_cell, int ncells, int32_t *cell_start)
{
    // cell_start[c] = first sorted position whose cell id >= c, for c in [0, ncells].
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    int cur = (int)sorted_cell[t];
    int prev = t == 0 ? -1 : (int)sorted_cell[t - 1];
    for (int c = prev + 1; c <= cur; ++c) cell_start[c] = t;
    if (t == N - 1)
        for (int c = cur + 1; c <= ncells; ++c) cell_start[c] = N;
}

template <int DIM, bool FILL>
__global__ void k_neigh(int N, CD3 x, Grid g, const int32_t *__restrict__ cell_start,
                        const uint32_t *__restrict__ items, double d2, int32_t *cnt,
                        const int32_t *__restrict__ row_ptr, int32_t *__restrict__ col)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double xi = x.p[0][i], yi = x.p[1][i];
    const double zi = DIM == 3 ? x.p[2][i] : 0.0;
    int c0 = cell_coord(xi, g.lo[0], g.cs, g.n[0]);
    int c1 = cell_coord(yi, g.lo[1], g.cs, g.n[1]);
    int c2 = DIM == 3 ? cell_coord(zi, g.lo[2], g.cs, g.n[2]) : 0;
    int n = 0;
    int32_t base = FILL ? row_ptr[i] : 0;
    const int zlo = DIM == 3 ? c2 - 1 : 0, zhi = DIM == 3 ? c2 + 1 : 0;
    for (int cz = zlo; cz <= zhi; ++cz) {
        if (cz < 0 || cz >= g.n[2]) continue;
        for (int cy = c1 - 1; cy <= c1 + 1; ++cy) {
            if (cy < 0 || cy >= g.n[1]) continue;
            int cxl = c0 > 0 ? c0 - 1 : 0, cxh = c0 + 1 < g.n[0] ? c0 + 1 : g.n[0] - 1;
            int rowc = (cz * g.n[1] + cy) * g.n[0];
            int32_t q0 = cell_start[rowc + cxl], q1 = cell_start[rowc + cxh + 1];
            for (int32_t q = q0; q < q1; ++q) {
                int j = (int)items[q];
                if (j == i) continue;
                double ddx = x.p[0][j] - xi;
                double ddy = x.p[1][j] - yi;
                double r2 = ddx * ddx + ddy * ddy;
                if (DIM == 3) {
                    double ddz = x.p[2][j] - zi;
                    r2 = r2 + ddz * ddz;
                }
                if (r2 < d2) {
                    if (FILL) col[base + n] = j;
                    ++n;
                }
            }
        }
    }
    if (!FILL) cnt[i] = n;
}

__global__ void k_tile_keys(int dim, int N, int n_int, CD3 x, Grid g, uint32_t *key, uint32_t *idx)
{
    // tile id of each particle, tiles numbered x-fastest then y then z;
    // particles i >= n_int sort after every integrated particle
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c0 = cell_coord(x.p[0][i], g.lo[0], g.cs, g.n[0]);
    int c1 = cell_coord(x.p[1][i], g.lo[1], g.cs, g.n[1]);
    int c2 = dim == 3 ? cell_coord(x.p[2][i], g.lo[2], g.cs, g.n[2]) : 0;
    uint32_t t = (uint32_t)((c2 * g.n[1] + c1) * g.n[0] + c0);
    key[i] = i < n_int ? t : t + (uint32_t)(g.n[0] * g.n[1] * g.n[2]);
    idx[i] = (uint32_t)i;
}

__global__ void k_tile_flags(int N, const uint32_t *__restrict__ key, uint32_t *flag)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < N) flag[p] = (p == 0 || key[p] != key[p - 1]) ? 1u : 0u;
}

// inc = inclusive prefix of the flags: position p lies in tile inc[p] - 1
__global__ void k_tile_starts(int N, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ inc,
                              int32_t *tile_start)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    if (flag[p]) tile_start[inc[p] - 1] = p;
    if (p == N - 1) tile_start[inc[p]] = N;
}

// One 64-lane wave sorts one row in LDS; four rows per 256-thread block.
template <int NP2>
__global__ __launch_bounds__(256) void k_row_sort(int N, const int32_t *__restrict__ row_ptr,
                                                  int32_t *__restrict__ col)
{
    __shared__ int32_t buf[4][NP2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    int32_t b = 0, len = 0;
    if (row < N) {
        b = row_ptr[row];
        len = row_ptr[row + 1] - b;
    }
    for (int t = lane; t < NP2; t += 64) buf[wave][t] = t < len ? col[b + t] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= NP2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < NP2 / 2; t += 64) {
                int p = 2 * j * (t / j) + (t % j);
                int q = p + j;
                bool up = (p & k) == 0;
                int32_t a = buf[wave][p], c = buf[wave][q];
                if ((a > c) == up) {
                    buf[wave][p] = c;
                    buf[wave][q] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int t = lane; t < len; t += 64) col[b + t] = buf[wave][t];
}

} // namespace

hipError_t launch_cell_ids(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                           const int *ncell, uint32_t *cell_of, uint32_t *idx)
{
    Grid g{{lo[0], lo[1], lo[2]}, cs, {ncell[0], ncell[1], ncell[2]}};
    k_cell_ids<<<(N + 255) / 256, 256, 0, s>>>(dim, N, x, g, cell_of, idx);
    return hipGetLastError();
}

hipError_t launch_cell_bounds(hipStream_t s, int N, const uint32_t *sorted_cell, int ncells,
                              int32_t *cell_start)
{
    k_cell_bounds<<<(N + 255) / 256, 256, 0, s>>>(N, sorted_cell, ncells, cell_start);
    return hipGetLastError();
}

hipError_t launch_neigh_count(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                              const int *ncell, const int32_t *cell_start, const uint32_t *items,
                              double d2, int32_t *cnt)
{
    Grid g{{lo[0], lo[1], lo[2]}, cs, {ncell[0], ncell[1], ncell[2]}};
    if (dim == 3)
        k_neigh<3, false><<<(N + 255) / 256, 256, 0, s>>>(N, x, g, cell_start, items, d2, cnt, nullptr, nullptr);
    else
        k_neigh<2, false><<<(N + 255) / 256, 256, 0, s>>>(N, x, g, cell_start, items, d2, cnt, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_neigh_fill(hipStream_t s, int dim, int N, CD3 x, const double *lo, double cs,
                             const int *ncell, const int32_t *cell_start, const uint32_t *items,
                             double d2, const int32_t *row_ptr, int32_t *col)
{
    Grid g{{lo[0], lo[1], lo[2]}, cs, {ncell[0], ncell[1], ncell[2]}};
    if (dim == 3)
        k_neigh<3, true><<<(N + 255) / 256, 256, 0, s>>>(N, x, g, cell_start, items, d2, nullptr, row_ptr, col);
    else
        k_neigh<2, true><<<(N + 255) / 256, 256, 0, s>>>(N, x, g, cell_start, items, d2, nullptr, row_ptr, col);
    return hipGetLastError();
}

constexpr int kMaxRowSort = 256;  // longest row k_row_sort handles

hipError_t launch_row_sort(hipStream_t s, int N, const int32_t *row_ptr, int32_t *col, int max_row)
{
    int blocks = (N + 3) / 4;
    if (max_row <= 32)
        k_row_sort<32><<<blocks, 256, 0, s>>>(N, row_ptr, col);
    else if (max_row <= 64)
        k_row_sort<64><<<blocks, 256, 0, s>>>(N, row_ptr, col);
    else if (max_row <= 128)
        k_row_sort<128><<<blocks, 256, 0, s>>>(N, row_ptr, col);
    else if (max_row <= 256)
        k_row_sort<256><<<blocks, 256, 0, s>>>(N, row_ptr, col);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ------------------------------------------------------------ host driver
#include <hipcub/hipcub.hpp>
#include <cmath>
#include <vector>

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
    template <class T> T *as() { return static_cast<T *>(p); }
};
} // namespace

#define NB_TRY(expr)                                                              \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            h->err = std::string(#expr) + ": " + hipGetErrorString(_e);           \
            return PD_E_HIP;                                                      \
        }                                                                         \
    } while (0)

int build_neighbors_impl(pd_handle *h)
{
    const int N = h->N, dim = h->P.dim;
    hipStream_t s = h->stream;
    const double cs = h->P.delta * (1.0 + 1e-9);
    const int64_t ncells64 = (int64_t)h->ncell[0] * h->ncell[1] * h->ncell[2];
    if (ncells64 >= (int64_t)1 << 31) {
        h->err = "cell grid too large";
        return PD_E_CAPACITY;
    }
    const int ncells = (int)ncells64;
    CD3 x = cd3(h->x);
    DevBuf cell_of, idx, cell_sorted, items, cell_start, cnt, tmp, maxv;
    NB_TRY(cell_of.alloc(sizeof(uint32_t) * N));
    NB_TRY(idx.alloc(sizeof(uint32_t) * N));
    NB_TRY(cell_sorted.alloc(sizeof(uint32_t) * N));
    NB_TRY(items.alloc(sizeof(uint32_t) * N));
    NB_TRY(cell_start.alloc(sizeof(int32_t) * (ncells + 1)));
    NB_TRY(cnt.alloc(sizeof(int32_t) * (N + 1)));
    NB_TRY(maxv.alloc(sizeof(int32_t)));
    NB_TRY(launch_cell_ids(s, dim, N, x, h->lo, cs, h->ncell, cell_of.as<uint32_t>(), idx.as<uint32_t>()));
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= ncells) ++end_bit;
    size_t tbytes = 0, t2 = 0, t3 = 0;
    NB_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, cell_of.as<uint32_t>(), cell_sorted.as<uint32_t>(),
                                              idx.as<uint32_t>(), items.as<uint32_t>(), N, 0, end_bit, s));
    NB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, cnt.as<int32_t>(), h->row_ptr, N + 1, s));
    NB_TRY(hipcub::DeviceReduce::Max(nullptr, t3, cnt.as<int32_t>(), maxv.as<int32_t>(), N, s));
    size_t need = tbytes > t2 ? tbytes : t2;
    if (t3 > need) need = t3;
    NB_TRY(tmp.alloc(need));
    NB_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tbytes, cell_of.as<uint32_t>(), cell_sorted.as<uint32_t>(),
                                              idx.as<uint32_t>(), items.as<uint32_t>(), N, 0, end_bit, s));
    NB_TRY(launch_cell_bounds(s, N, cell_sorted.as<uint32_t>(), ncells, cell_start.as<int32_t>()));
    const double d2 = h->P.delta * h->P.delta;
    NB_TRY(launch_neigh_count(s, dim, N, x, h->lo, cs, h->ncell, cell_start.as<int32_t>(), items.as<uint32_t>(), d2,
                              cnt.as<int32_t>()));
    NB_TRY(hipMemsetAsync(cnt.as<int32_t>() + N, 0, sizeof(int32_t), s));
    NB_TRY(hipcub::DeviceReduce::Max(tmp.p, t3, cnt.as<int32_t>(), maxv.as<int32_t>(), N, s));
    int32_t max_row = 0;
    NB_TRY(hipMemcpyAsync(&max_row, maxv.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NB_TRY(hipStreamSynchronize(s));
    if (max_row > h->P.max_neighbors) {
        h->err = "neighbour count " + std::to_string(max_row) + " exceeds max_neighbors " +
                 std::to_string(h->P.max_neighbors);
        return PD_E_CAPACITY;
    }
    if (max_row > kMaxRowSort) {
        h->err = "neighbour count " + std::to_string(max_row) + " exceeds the " + std::to_string(kMaxRowSort) +
                 " neighbours per row the CSR row sort holds";
        return PD_E_CAPACITY;
    }
    if ((int64_t)max_row * N >= ((int64_t)1 << 31)) {
        h->err = "bond count exceeds the int32 CSR range";
        return PD_E_CAPACITY;
    }
    NB_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, t2, cnt.as<int32_t>(), h->row_ptr, N + 1, s));
    int32_t nb = 0;
    NB_TRY(hipMemcpyAsync(&nb, h->row_ptr + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NB_TRY(hipStreamSynchronize(s));
    if (h->col) (void)hipFree(h->col);
    h->col = nullptr;
    NB_TRY(hipMalloc(&h->col, sizeof(int32_t) * (nb > 0 ? nb : 1)));
    h->Nb = nb;
    NB_TRY(launch_neigh_fill(s, dim, N, x, h->lo, cs, h->ncell, cell_start.as<int32_t>(), items.as<uint32_t>(), d2,
                             h->row_ptr, h->col));
    NB_TRY(launch_row_sort(s, N, h->row_ptr, h->col, max_row > 0 ? max_row : 1));
    NB_TRY(hipStreamSynchronize(s));
    return PD_OK;
}

int build_tile_order(pd_handle *h, double tile_len, bool want_tiles)
{
    // Processing order of the throughput kernels: particles grouped by cubic
    // tiles of edge tile_len (stable inside a tile), so the particles that run
    // together on one XCD gather their neighbours from one compact region.
    // Integrated particles (i < N_int) come first, so order[0, N_int) is the
    // work list of the passes that only visit integrated rows.  With
    // want_tiles the row range of every non-empty tile is recorded in
    // h->lds.tile_start (tiles of integrated rows first, then the others).
    const int N = h->N, dim = h->P.dim;
    hipStream_t s = h->stream;
    Grid g{{h->lo[0], h->lo[1], h->lo[2]}, tile_len, {1, 1, 1}};
    double hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        hi[d] = h->lo[d] + h->ncell[d] * h->P.delta * (1.0 + 1e-9);
        g.n[d] = (int)std::floor((hi[d] - h->lo[d]) / tile_len) + 1;
    }
    if ((int64_t)g.n[0] * g.n[1] * g.n[2] >= ((int64_t)1 << 30)) {
        h->err = "too many tiles";
        return PD_E_CAPACITY;
    }
    int ntiles = 2 * g.n[0] * g.n[1] * g.n[2];
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= ntiles) ++end_bit;
    DevBuf key, idx, key2, tmp;
    NB_TRY(key.alloc(sizeof(uint32_t) * N));
    NB_TRY(idx.alloc(sizeof(uint32_t) * N));
    NB_TRY(key2.alloc(sizeof(uint32_t) * N));
    if (!h->order) NB_TRY(hipMalloc(&h->order, sizeof(int32_t) * (N > 0 ? N : 1)));
    k_tile_keys<<<(N + 255) / 256, 256, 0, s>>>(dim, N, h->N_int, cd3(h->x), g, key.as<uint32_t>(), idx.as<uint32_t>());
    NB_TRY(hipGetLastError());
    size_t tb = 0;
    NB_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.as<uint32_t>(), key2.as<uint32_t>(), idx.as<uint32_t>(),
                                              reinterpret_cast<uint32_t *>(h->order), N, 0, end_bit, s));
    NB_TRY(tmp.alloc(tb));
    NB_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.as<uint32_t>(), key2.as<uint32_t>(), idx.as<uint32_t>(),
                                              reinterpret_cast<uint32_t *>(h->order), N, 0, end_bit, s));
    if (want_tiles && N > 0) {
        // key (unsorted) is free now: reuse it for the flags, idx for their prefix
        uint32_t *flag = key.as<uint32_t>(), *inc = idx.as<uint32_t>();
        k_tile_flags<<<(N + 255) / 256, 256, 0, s>>>(N, key2.as<uint32_t>(), flag);
        NB_TRY(hipGetLastError());
        size_t sb = 0;
        NB_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, sb, flag, inc, N, s));
        DevBuf tmp2;
        NB_TRY(tmp2.alloc(sb));
        NB_TRY(hipcub::DeviceScan::InclusiveSum(tmp2.p, sb, flag, inc, N, s));
        uint32_t nt = 0, nt_int = 0;
        NB_TRY(hipMemcpyAsync(&nt, inc + (N - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (h->N_int < N) NB_TRY(hipMemcpyAsync(&nt_int, inc + h->N_int, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NB_TRY(hipStreamSynchronize(s));
        if (h->N_int >= N) nt_int = nt + 1;  // every row is integrated
        h->lds.ntiles = (int)nt;
        h->lds.ntiles_int = (int)nt_int - 1;
        NB_TRY(hipMalloc(&h->lds.tile_start, sizeof(int32_t) * (nt + 1)));
        k_tile_starts<<<(N + 255) / 256, 256, 0, s>>>(N, flag, inc, h->lds.tile_start);
        NB_TRY(hipGetLastError());
    }
    NB_TRY(hipStreamSynchronize(s));
    return PD_OK;
}

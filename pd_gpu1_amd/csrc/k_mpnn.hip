// k_mpnn.hip -- kernels of the message-passing warm-start network.
//
// The network (3 message-passing layers, hidden 64, SiLU) predicts the next
// load-step displacement increment of a plate from its PD bond graph.  Dense
// products run on gfx950 matrix cores: v_mfma_f32_16x16x32_bf16 with BF16
// operands and FP32 accumulation.  One wave computes a 16-row x NC-column
// tile; the operand layout used here is: lane l holds A[l%16][8(l/16)..+7]
// and B[8(l/16)..+7][l%16] (k along the 8 packed bf16), and receives
// D[4(l/16)+r][l%16], r = 0..3.  Message aggregation is a gather over the
// CSR rows (one wave per node, one lane per channel) and the backward scatter
// to the source nodes is a gather through the reverse-edge index, so every
// sum has a fixed order and no atomics are used.
#include "pd_mpnn.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p)
{
    return *reinterpret_cast<const bf16x8 *>(p);
}

// C[M x NC] (+)= A[M x K] . BT[NC x K]^T ; A, BT bf16 row-major; C fp32.
template <int K, int NC, bool ACC>
__global__ __launch_bounds__(256) void k_gemm_nt(int M, const __bf16 *__restrict__ A, const __bf16 *__restrict__ BT,
                                                 float *__restrict__ C)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + wave) * 16;
    if (row0 >= M) return;
    const int r = row0 + (lane & 15);
    const int kq = (lane >> 4) * 8;
    f32x4 acc[NC / 16];
#pragma unroll
    for (int t = 0; t < NC / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a;
        if (r < M)
            a = load8(A + (size_t)r * K + k0 + kq);
        else
            for (int t = 0; t < 8; ++t) a[t] = (__bf16)0.0f;
#pragma unroll
        for (int t = 0; t < NC / 16; ++t) {
            bf16x8 b = load8(BT + (size_t)(t * 16 + (lane & 15)) * K + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NC / 16; ++t)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int row = row0 + (lane >> 4) * 4 + rr;
            if (row < M) {
                float *dst = C + (size_t)row * NC + t * 16 + (lane & 15);
                *dst = ACC ? *dst + acc[t][rr] : acc[t][rr];
            }
        }
}

// Partial weight gradient: P[chunk][KA x NC] = A[rows of chunk]^T . B[rows of chunk]
// with A [M x KA] and B [M x NC] bf16 row-major.  Grid (KA/16 * NC/16, chunks).
template <int KA, int NC>
__global__ __launch_bounds__(64) void k_gemm_tn_partial(int M, int chunk_rows, const __bf16 *__restrict__ A,
                                                        const __bf16 *__restrict__ B, float *__restrict__ part)
{
    const int lane = threadIdx.x;
    const int tile = blockIdx.x;
    const int tk = tile / (NC / 16), tn = tile % (NC / 16);
    const int m_begin = blockIdx.y * chunk_rows;
    const int m_end = min(M, m_begin + chunk_rows);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int kq = (lane >> 4) * 8;
    for (int m0 = m_begin; m0 < m_end; m0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int m = m0 + kq + t;
            const bool ok = m < m_end;
            a[t] = ok ? A[(size_t)m * KA + tk * 16 + (lane & 15)] : (__bf16)0.0f;
            b[t] = ok ? B[(size_t)m * NC + tn * 16 + (lane & 15)] : (__bf16)0.0f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    float *dst = part + (size_t)blockIdx.y * KA * NC;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int k = tk * 16 + (lane >> 4) * 4 + rr;
        dst[(size_t)k * NC + tn * 16 + (lane & 15)] = acc[rr];
    }
}

// out[i] = sum over chunks c of part[c][i] (fixed order); optional rows/cols window.
__global__ void k_reduce_chunks(int chunks, int KA, int NC, int k_use, int n_use, const float *__restrict__ part,
                                float *__restrict__ out, int out_ld)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k_use * n_use) return;
    const int k = t / n_use, n = t % n_use;
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * KA * NC + (size_t)k * NC + n];
    out[(size_t)k * out_ld + n] = s;
}

// out[n] = sum_r part[r*stride + off + n], rows in order
__global__ void k_reduce_strided(int rows, int stride, int off, int width, const float *__restrict__ part,
                                 float *__restrict__ out)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= width) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += part[(size_t)r * stride + off + n];
    out[n] = s;
}

// Column sums of X [M x NC] fp32 in fixed order: per-block partials over
// row chunks, then k_reduce_chunks.
__global__ void k_colsum_partial(int M, int NC, int chunk_rows, const float *__restrict__ X, float *__restrict__ part)
{
    const int n = threadIdx.x;
    if (n >= NC) return;
    const int m0 = blockIdx.x * chunk_rows, m1 = min(M, m0 + chunk_rows);
    float s = 0.f;
    for (int m = m0; m < m1; ++m) s += X[(size_t)m * NC + n];
    part[(size_t)blockIdx.x * NC + n] = s;
}

// pre += bias (in place); act = silu(pre) (fp32, and bf16 copy).  If res is
// given: act = res + silu(pre).
__global__ void k_bias_silu(int M, int NC, float *__restrict__ pre, const float *__restrict__ bias,
                            const float *__restrict__ res, float *__restrict__ act, __bf16 *__restrict__ act16)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * NC) return;
    const int n = (int)(t % NC);
    float z = pre[t] + bias[n];
    pre[t] = z;
    float v = silu(z);
    if (res) v += res[t];
    act[t] = v;
    act16[t] = (__bf16)v;
}

__global__ void k_to_bf16(int64_t n, const float *__restrict__ a, __bf16 *__restrict__ b)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (__bf16)a[t];
}

// Message aggregation, one wave per node, lane = channel:
// a_i = dinv_i * sum_{k in row i} silu(P[col_k] + e_k . Q + c)
__global__ __launch_bounds__(256) void k_aggregate(int N, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                   const float *__restrict__ ef, const float *__restrict__ dinv,
                                                   const float *__restrict__ P, const float *__restrict__ Q,
                                                   const float *__restrict__ cvec, float *__restrict__ a,
                                                   __bf16 *__restrict__ a16)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (i >= N) return;
    const float q0 = Q[f], q1 = Q[64 + f], q2 = Q[128 + f], cf = cvec[f];
    float acc = 0.f;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int c = col[k];
        const float pre = P[(size_t)c * 64 + f] + ef[3 * (size_t)k] * q0 + ef[3 * (size_t)k + 1] * q1 +
                          ef[3 * (size_t)k + 2] * q2 + cf;
        acc += silu(pre);
    }
    const float v = acc * dinv[i];
    a[(size_t)i * 64 + f] = v;
    a16[(size_t)i * 64 + f] = (__bf16)v;
}

// Backward of the aggregation for source node j (gather through rev):
// dP_j = sum_{k in row j} da[col_k] dinv[col_k] dsilu(P_j + e_rev(k) . Q + c)
// plus per-wave partial sums of dQ (3 x 64) and dc (64) into part[wave][4][64].
__global__ __launch_bounds__(256) void k_aggregate_bwd(int N, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ col, const int32_t *__restrict__ rev,
                                                       const float *__restrict__ ef, const float *__restrict__ dinv,
                                                       const float *__restrict__ P, const float *__restrict__ Q,
                                                       const float *__restrict__ cvec, const float *__restrict__ da,
                                                       float *__restrict__ dP, __bf16 *__restrict__ dP16,
                                                       float *__restrict__ part)
{
    const int wave_g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    const int f = threadIdx.x & 63;
    const float q0 = Q[f], q1 = Q[64 + f], q2 = Q[128 + f], cf = cvec[f];
    float gq0 = 0.f, gq1 = 0.f, gq2 = 0.f, gc = 0.f;
    for (int j = wave_g; j < N; j += nwaves) {
        const float pj = P[(size_t)j * 64 + f];
        float acc = 0.f;
        for (int32_t k = rp[j]; k < rp[j + 1]; ++k) {
            const int i = col[k];
            const int32_t kr = rev[k];  // edge (i <- j) in row i
            const float e0 = ef[3 * (size_t)kr], e1 = ef[3 * (size_t)kr + 1], e2 = ef[3 * (size_t)kr + 2];
            const float g = da[(size_t)i * 64 + f] * dinv[i] * dsilu(pj + e0 * q0 + e1 * q1 + e2 * q2 + cf);
            acc += g;
            gq0 += e0 * g;
            gq1 += e1 * g;
            gq2 += e2 * g;
            gc += g;
        }
        dP[(size_t)j * 64 + f] = acc;
        dP16[(size_t)j * 64 + f] = (__bf16)acc;
    }
    float *dst = part + (size_t)wave_g * 256;
    dst[f] = gq0;
    dst[64 + f] = gq1;
    dst[128 + f] = gq2;
    dst[192 + f] = gc;
}

// dz = dh * dsilu(z) (fp32 and bf16)
__global__ void k_dsilu_mul(int64_t n, const float *__restrict__ dh, const float *__restrict__ z, float *__restrict__ dz,
                            __bf16 *__restrict__ dz16)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float v = dh[t] * dsilu(z[t]);
    dz[t] = v;
    dz16[t] = (__bf16)v;
}

// Decoder output and loss gradient: y = Y[:, 0:2] + bd; dy = 2 w (y - t) / wsum
// into a zero-padded [N x 32] bf16 matrix; per-node squared errors for the loss.
__global__ void k_loss(int N, const float *__restrict__ Y, const float *__restrict__ bd, const float *__restrict__ tgt,
                       const float *__restrict__ w, float inv_wsum, float *__restrict__ y_out,
                       float *__restrict__ dy, __bf16 *__restrict__ dy16, float *__restrict__ err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float e = 0.f;
    for (int c = 0; c < 32; ++c) {
        float g = 0.f;
        if (c < 2) {
            const float y = Y[(size_t)i * 16 + c] + bd[c];
            if (y_out) y_out[2 * (size_t)i + c] = y;
            if (tgt) {
                const float r = y - tgt[2 * (size_t)i + c];
                e += w[i] * r * r;
                g = 2.f * w[i] * r * inv_wsum;
            }
        }
        if (dy) {
            dy[(size_t)i * 32 + c] = g;
            dy16[(size_t)i * 32 + c] = (__bf16)g;
        }
    }
    if (err) err[i] = e;
}

// Weight preparation: bf16 operand copies of the FP32 master parameters.
__global__ void k_prep_weights(const float *__restrict__ th, MpnnWeights16 w)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const ParamLayout L = ParamLayout::make();
    // W0^T padded: [64][32]
    if (t < 64 * 32) {
        const int n = t / 32, k = t % 32;
        w.W0T[t] = (__bf16)(k < kNodeFeat ? th[L.W0 + k * 64 + n] : 0.f);
    }
    for (int l = 0; l < 3; ++l) {
        if (t < 64 * 64) {
            const int n = t / 64, k = t % 64;
            w.W1T[l][t] = (__bf16)th[L.W1[l] + k * 64 + n];     // [n][k] = W1[k][n]
            w.W1[l][t] = (__bf16)th[L.W1[l] + t];               // [k][n] as stored
            w.W2aT[l][t] = (__bf16)th[L.W2[l] + k * 64 + n];
            w.W2bT[l][t] = (__bf16)th[L.W2[l] + (64 + k) * 64 + n];
            w.W2a[l][t] = (__bf16)th[L.W2[l] + t];
            w.W2b[l][t] = (__bf16)th[L.W2[l] + 64 * 64 + t];
        }
    }
    // Wd^T padded [16][64] (forward) and Wd padded [64][32] (backward)
    if (t < 16 * 64) {
        const int n = t / 64, k = t % 64;
        w.WdT[t] = (__bf16)(n < 2 ? th[L.Wd + k * 2 + n] : 0.f);
    }
    if (t < 64 * 32) {
        const int n = t / 32, k = t % 32;
        w.Wd[t] = (__bf16)(k < 2 ? th[L.Wd + n * 2 + k] : 0.f);
    }
}

__global__ void k_adam(int n, float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    th[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

} // namespace

// ------------------------------------------------------------ launchers
hipError_t mpnn_gemm_nt(hipStream_t s, int K, int NC, bool acc, int M, const __bf16 *A, const __bf16 *BT, float *C)
{
    const unsigned g = nb(M, 64);
    if (M <= 0) return hipSuccess;
#define GEMM_CASE(KK, NN)                                                            \
    if (K == KK && NC == NN) {                                                       \
        if (acc)                                                                     \
            k_gemm_nt<KK, NN, true><<<g, 256, 0, s>>>(M, A, BT, C);                  \
        else                                                                         \
            k_gemm_nt<KK, NN, false><<<g, 256, 0, s>>>(M, A, BT, C);                 \
        return hipGetLastError();                                                    \
    }
    GEMM_CASE(32, 64)
    GEMM_CASE(64, 64)
    GEMM_CASE(64, 16)
#undef GEMM_CASE
    return hipErrorInvalidValue;
}

hipError_t mpnn_gemm_tn(hipStream_t s, int KA, int NC, int M, const __bf16 *A, const __bf16 *B, float *part,
                        int chunks, int chunk_rows, float *out, int k_use, int n_use, int out_ld)
{
    dim3 grid((KA / 16) * (NC / 16), chunks);
#define TN_CASE(KK, NN)                                                                   \
    if (KA == KK && NC == NN) {                                                           \
        k_gemm_tn_partial<KK, NN><<<grid, 64, 0, s>>>(M, chunk_rows, A, B, part);         \
    } else
    TN_CASE(32, 64)
    TN_CASE(64, 64)
    TN_CASE(64, 32)
    return hipErrorInvalidValue;
#undef TN_CASE
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_chunks<<<nb((int64_t)k_use * n_use, 256), 256, 0, s>>>(chunks, KA, NC, k_use, n_use, part, out, out_ld);
    return hipGetLastError();
}

hipError_t mpnn_colsum(hipStream_t s, int M, int NC, const float *X, float *part, int chunks, int chunk_rows, int n_use,
                       float *out)
{
    k_colsum_partial<<<chunks, 64, 0, s>>>(M, NC, chunk_rows, X, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_chunks<<<nb(n_use, 256), 256, 0, s>>>(chunks, 1, NC, 1, n_use, part, out, n_use);
    return hipGetLastError();
}

hipError_t mpnn_bias_silu(hipStream_t s, int M, int NC, float *pre, const float *bias, const float *res, float *act,
                          __bf16 *act16)
{
    k_bias_silu<<<nb((int64_t)M * NC, 256), 256, 0, s>>>(M, NC, pre, bias, res, act, act16);
    return hipGetLastError();
}

hipError_t mpnn_to_bf16(hipStream_t s, int64_t n, const float *a, __bf16 *b)
{
    k_to_bf16<<<nb(n, 256), 256, 0, s>>>(n, a, b);
    return hipGetLastError();
}

hipError_t mpnn_aggregate(hipStream_t s, int N, const int32_t *rp, const int32_t *col, const float *ef,
                          const float *dinv, const float *P, const float *Q, const float *c, float *a, __bf16 *a16)
{
    k_aggregate<<<nb(N, 4), 256, 0, s>>>(N, rp, col, ef, dinv, P, Q, c, a, a16);
    return hipGetLastError();
}

hipError_t mpnn_aggregate_bwd(hipStream_t s, int waves, int N, const int32_t *rp, const int32_t *col,
                              const int32_t *rev, const float *ef, const float *dinv, const float *P, const float *Q,
                              const float *c, const float *da, float *dP, __bf16 *dP16, float *part)
{
    k_aggregate_bwd<<<waves / 4, 256, 0, s>>>(N, rp, col, rev, ef, dinv, P, Q, c, da, dP, dP16, part);
    return hipGetLastError();
}

hipError_t mpnn_reduce_rows(hipStream_t s, int rows, int stride, int off, int width, const float *part, float *out)
{
    k_reduce_strided<<<nb(width, 256), 256, 0, s>>>(rows, stride, off, width, part, out);
    return hipGetLastError();
}

hipError_t mpnn_dsilu_mul(hipStream_t s, int64_t n, const float *dh, const float *z, float *dz, __bf16 *dz16)
{
    k_dsilu_mul<<<nb(n, 256), 256, 0, s>>>(n, dh, z, dz, dz16);
    return hipGetLastError();
}

hipError_t mpnn_loss(hipStream_t s, int N, const float *Y, const float *bd, const float *tgt, const float *w,
                     float inv_wsum, float *y_out, float *dy, __bf16 *dy16, float *err)
{
    k_loss<<<nb(N, 256), 256, 0, s>>>(N, Y, bd, tgt, w, inv_wsum, y_out, dy, dy16, err);
    return hipGetLastError();
}

hipError_t mpnn_prep_weights(hipStream_t s, const float *theta, const MpnnWeights16 &w)
{
    k_prep_weights<<<nb(64 * 64, 256), 256, 0, s>>>(theta, w);
    return hipGetLastError();
}

hipError_t mpnn_adam(hipStream_t s, int n, float *th, const float *g, float *m, float *v, float lr, float b1, float b2,
                     float eps, float c1, float c2)
{
    k_adam<<<nb(n, 256), 256, 0, s>>>(n, th, g, m, v, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

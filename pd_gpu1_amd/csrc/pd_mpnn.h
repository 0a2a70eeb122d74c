// pd_mpnn.h -- shared definitions of the message-passing warm-start network.
#pragma once
#include "pd_internal.h"

constexpr int kNodeFeat = 16;  // node features (the encoder GEMM pads them to K = 32)

// Offsets of the parameter blocks in the flat FP32 vector: W0 [16][64], b0,
// then per layer W1 [64][64], Q [3][64], c [64], W2 [128][64], b2 [64], then
// Wd [64][2], bd [2] (matrices [in][out] row-major).
struct ParamLayout {
    int W0, b0, W1[3], Q[3], c[3], W2[3], b2[3], Wd, bd, total;
    __host__ __device__ static ParamLayout make()
    {
        ParamLayout L{};
        int o = 0;
        L.W0 = o; o += kNodeFeat * 64;
        L.b0 = o; o += 64;
        for (int l = 0; l < 3; ++l) {
            L.W1[l] = o; o += 64 * 64;
            L.Q[l] = o; o += 3 * 64;
            L.c[l] = o; o += 64;
            L.W2[l] = o; o += 128 * 64;
            L.b2[l] = o; o += 64;
        }
        L.Wd = o; o += 64 * 2;
        L.bd = o; o += 2;
        L.total = o;
        return L;
    }
};

// BF16 operand copies of the weights, in the layouts the GEMMs read.
struct MpnnWeights16 {
    __bf16 *W0T;                 // [64][32]   (W0^T, k padded to 32)
    __bf16 *W1T[3], *W1[3];      // [64][64]
    __bf16 *W2aT[3], *W2bT[3];   // [64][64]   (rows 0..63 / 64..127 of W2, transposed)
    __bf16 *W2a[3], *W2b[3];     // [64][64]   (as stored)
    __bf16 *WdT;                 // [16][64]   (Wd^T, rows padded to 16)
    __bf16 *Wd;                  // [64][32]   (Wd, columns padded to 32)
};

hipError_t mpnn_gemm_nt(hipStream_t s, int K, int NC, bool acc, int M, const __bf16 *A, const __bf16 *BT, float *C);
hipError_t mpnn_gemm_tn(hipStream_t s, int KA, int NC, int M, const __bf16 *A, const __bf16 *B, float *part,
                        int chunks, int chunk_rows, float *out, int k_use, int n_use, int out_ld);
hipError_t mpnn_colsum(hipStream_t s, int M, int NC, const float *X, float *part, int chunks, int chunk_rows, int n_use,
                       float *out);
hipError_t mpnn_bias_silu(hipStream_t s, int M, int NC, float *pre, const float *bias, const float *res, float *act,
                          __bf16 *act16);
hipError_t mpnn_to_bf16(hipStream_t s, int64_t n, const float *a, __bf16 *b);
hipError_t mpnn_aggregate(hipStream_t s, int N, const int32_t *rp, const int32_t *col, const float *ef,
                          const float *dinv, const float *P, const float *Q, const float *c, float *a, __bf16 *a16);
hipError_t mpnn_aggregate_bwd(hipStream_t s, int waves, int N, const int32_t *rp, const int32_t *col,
                              const int32_t *rev, const float *ef, const float *dinv, const float *P, const float *Q,
                              const float *c, const float *da, float *dP, __bf16 *dP16, float *part);
hipError_t mpnn_reduce_rows(hipStream_t s, int rows, int stride, int off, int width, const float *part, float *out);
hipError_t mpnn_dsilu_mul(hipStream_t s, int64_t n, const float *dh, const float *z, float *dz, __bf16 *dz16);
hipError_t mpnn_loss(hipStream_t s, int N, const float *Y, const float *bd, const float *tgt, const float *w,
                     float inv_wsum, float *y_out, float *dy, __bf16 *dy16, float *err);
hipError_t mpnn_prep_weights(hipStream_t s, const float *theta, const MpnnWeights16 &w);
hipError_t mpnn_adam(hipStream_t s, int n, float *th, const float *g, float *m, float *v, float lr, float b1, float b2,
                     float eps, float c1, float c2);

// pd_mpnn.hip -- host side of the message-passing warm-start network:
// parameters (FP32 master copy + Adam moments), graph upload, forward pass,
// backward pass and training step on one HIP stream.  The kernels are in
// k_mpnn.hip.  Layer structure (hidden 64, 3 layers, SiLU):
//   h0 = silu(x W0 + b0); per layer P = h W1, a_i = mean_j silu(P_j + e_ij Q + c),
//   h += silu([h, a] W2 + b2); y = h Wd + bd.
#include "pd_mpnn.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

struct pd_mpnn {
    int device = 0;
    hipStream_t s = nullptr;
    std::string err;
    ParamLayout L = ParamLayout::make();
    float *theta = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    int adam_t = 0;
    __bf16 *wbuf = nullptr;
    MpnnWeights16 w{};
    // graph
    int N = 0, capN = 0;
    int64_t E = 0, capE = 0;
    int32_t *rp = nullptr, *col = nullptr, *rev = nullptr;
    float *ef = nullptr, *dinv = nullptr;
    __bf16 *x16 = nullptr;
    float *tgt = nullptr, *wt = nullptr;
    double wsum = 0.0;
    // activations and gradients
    float *u0 = nullptr, *h[4] = {}, *P[3] = {}, *a[3] = {}, *z[3] = {};
    __bf16 *h16[4] = {}, *a16[3] = {};
    float *Y = nullptr, *yout = nullptr, *dy = nullptr, *err_n = nullptr;
    __bf16 *dy16 = nullptr;
    float *dh = nullptr, *dz = nullptr, *da = nullptr, *dP = nullptr;
    __bf16 *dz16 = nullptr, *dP16 = nullptr;
    float *part = nullptr;
    size_t part_cap = 0;
    float *scal = nullptr;  // device scalars: loss
};

namespace {

constexpr int kChunk = 4096;
constexpr int kAggWaves = 1024;

#define M_TRY(expr)                                                               \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            m->err = std::string(#expr) + ": " + hipGetErrorString(_e);           \
            return PD_E_HIP;                                                      \
        }                                                                         \
    } while (0)

template <class T>
void mfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_graph(pd_mpnn *m)
{
    mfree(m->rp); mfree(m->col); mfree(m->rev); mfree(m->ef); mfree(m->dinv);
    mfree(m->x16); mfree(m->tgt); mfree(m->wt);
    mfree(m->u0);
    for (int k = 0; k < 4; ++k) { mfree(m->h[k]); mfree(m->h16[k]); }
    for (int k = 0; k < 3; ++k) { mfree(m->P[k]); mfree(m->a[k]); mfree(m->z[k]); mfree(m->a16[k]); }
    mfree(m->Y); mfree(m->yout); mfree(m->dy); mfree(m->dy16); mfree(m->err_n);
    mfree(m->dh); mfree(m->dz); mfree(m->da); mfree(m->dP); mfree(m->dz16); mfree(m->dP16);
    mfree(m->part);
    m->capN = 0;
    m->capE = 0;
    m->part_cap = 0;
}

int ensure_capacity(pd_mpnn *m, int N, int64_t E)
{
    if (N <= m->capN && E <= m->capE) return PD_OK;
    free_graph(m);
    const int cN = std::max(N, 64);
    const int64_t cE = std::max<int64_t>(E, 64);
    const size_t n64 = (size_t)cN * 64;
    M_TRY(hipMalloc(&m->rp, sizeof(int32_t) * (cN + 1)));
    M_TRY(hipMalloc(&m->col, sizeof(int32_t) * cE));
    M_TRY(hipMalloc(&m->rev, sizeof(int32_t) * cE));
    M_TRY(hipMalloc(&m->ef, sizeof(float) * 3 * cE));
    M_TRY(hipMalloc(&m->dinv, sizeof(float) * cN));
    M_TRY(hipMalloc(&m->x16, sizeof(__bf16) * 32 * (size_t)cN));
    M_TRY(hipMalloc(&m->tgt, sizeof(float) * 2 * (size_t)cN));
    M_TRY(hipMalloc(&m->wt, sizeof(float) * cN));
    M_TRY(hipMalloc(&m->u0, sizeof(float) * n64));
    for (int k = 0; k < 4; ++k) {
        M_TRY(hipMalloc(&m->h[k], sizeof(float) * n64));
        M_TRY(hipMalloc(&m->h16[k], sizeof(__bf16) * n64));
    }
    for (int k = 0; k < 3; ++k) {
        M_TRY(hipMalloc(&m->P[k], sizeof(float) * n64));
        M_TRY(hipMalloc(&m->a[k], sizeof(float) * n64));
        M_TRY(hipMalloc(&m->z[k], sizeof(float) * n64));
        M_TRY(hipMalloc(&m->a16[k], sizeof(__bf16) * n64));
    }
    M_TRY(hipMalloc(&m->Y, sizeof(float) * 16 * (size_t)cN));
    M_TRY(hipMalloc(&m->yout, sizeof(float) * 2 * (size_t)cN));
    M_TRY(hipMalloc(&m->dy, sizeof(float) * 32 * (size_t)cN));
    M_TRY(hipMalloc(&m->dy16, sizeof(__bf16) * 32 * (size_t)cN));
    M_TRY(hipMalloc(&m->err_n, sizeof(float) * cN));
    M_TRY(hipMalloc(&m->dh, sizeof(float) * n64));
    M_TRY(hipMalloc(&m->dz, sizeof(float) * n64));
    M_TRY(hipMalloc(&m->da, sizeof(float) * n64));
    M_TRY(hipMalloc(&m->dP, sizeof(float) * n64));
    M_TRY(hipMalloc(&m->dz16, sizeof(__bf16) * n64));
    M_TRY(hipMalloc(&m->dP16, sizeof(__bf16) * n64));
    const size_t chunks = (size_t)(cN + kChunk - 1) / kChunk;
    size_t need = std::max<size_t>(chunks * 64 * 64, (size_t)kAggWaves * 256);
    need = std::max<size_t>(need, ((size_t)cN + 1023) / 1024 * 64);
    M_TRY(hipMalloc(&m->part, sizeof(float) * need));
    m->part_cap = need;
    m->capN = cN;
    m->capE = cE;
    return PD_OK;
}

uint64_t sm64(uint64_t *st)
{
    // SplitMix64 (Steele, Lea, Flood 2014)
    uint64_t zz = (*st += 0x9E3779B97F4A7C15ULL);
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
    return zz ^ (zz >> 31);
}

int forward(pd_mpnn *m, bool train)
{
    hipStream_t s = m->s;
    const int N = m->N;
    const ParamLayout &L = m->L;
    M_TRY(mpnn_prep_weights(s, m->theta, m->w));
    M_TRY(mpnn_gemm_nt(s, 32, 64, false, N, m->x16, m->w.W0T, m->u0));
    M_TRY(mpnn_bias_silu(s, N, 64, m->u0, m->theta + L.b0, nullptr, m->h[0], m->h16[0]));
    for (int l = 0; l < 3; ++l) {
        M_TRY(mpnn_gemm_nt(s, 64, 64, false, N, m->h16[l], m->w.W1T[l], m->P[l]));
        M_TRY(mpnn_aggregate(s, N, m->rp, m->col, m->ef, m->dinv, m->P[l], m->theta + L.Q[l], m->theta + L.c[l],
                             m->a[l], m->a16[l]));
        M_TRY(mpnn_gemm_nt(s, 64, 64, false, N, m->h16[l], m->w.W2aT[l], m->z[l]));
        M_TRY(mpnn_gemm_nt(s, 64, 64, true, N, m->a16[l], m->w.W2bT[l], m->z[l]));
        M_TRY(mpnn_bias_silu(s, N, 64, m->z[l], m->theta + L.b2[l], m->h[l], m->h[l + 1], m->h16[l + 1]));
    }
    M_TRY(mpnn_gemm_nt(s, 64, 16, false, N, m->h16[3], m->w.WdT, m->Y));
    const float inv_w = m->wsum > 0 ? (float)(1.0 / m->wsum) : 0.f;
    M_TRY(mpnn_loss(s, N, m->Y, m->theta + L.bd, train ? m->tgt : nullptr, m->wt, inv_w, m->yout,
                    train ? m->dy : nullptr, train ? m->dy16 : nullptr, train ? m->err_n : nullptr));
    return PD_OK;
}

int backward(pd_mpnn *m)
{
    hipStream_t s = m->s;
    const int N = m->N;
    const ParamLayout &L = m->L;
    float *g = m->grad;
    const int chunks = (N + kChunk - 1) / kChunk;
    const int cchunks = (N + 1023) / 1024;
    const int64_t n64 = (int64_t)N * 64;
    // decoder
    M_TRY(mpnn_gemm_tn(s, 64, 32, N, m->h16[3], m->dy16, m->part, chunks, kChunk, g + L.Wd, 64, 2, 2));
    M_TRY(mpnn_colsum(s, N, 32, m->dy, m->part, cchunks, 1024, 2, g + L.bd));
    M_TRY(mpnn_gemm_nt(s, 32, 64, false, N, m->dy16, m->w.Wd, m->dh));
    for (int l = 2; l >= 0; --l) {
        M_TRY(mpnn_dsilu_mul(s, n64, m->dh, m->z[l], m->dz, m->dz16));
        M_TRY(mpnn_gemm_tn(s, 64, 64, N, m->h16[l], m->dz16, m->part, chunks, kChunk, g + L.W2[l], 64, 64, 64));
        M_TRY(mpnn_gemm_tn(s, 64, 64, N, m->a16[l], m->dz16, m->part, chunks, kChunk, g + L.W2[l] + 64 * 64, 64, 64,
                           64));
        M_TRY(mpnn_colsum(s, N, 64, m->dz, m->part, cchunks, 1024, 64, g + L.b2[l]));
        M_TRY(mpnn_gemm_nt(s, 64, 64, true, N, m->dz16, m->w.W2a[l], m->dh));
        M_TRY(mpnn_gemm_nt(s, 64, 64, false, N, m->dz16, m->w.W2b[l], m->da));
        M_TRY(mpnn_aggregate_bwd(s, kAggWaves, N, m->rp, m->col, m->rev, m->ef, m->dinv, m->P[l], m->theta + L.Q[l],
                                 m->theta + L.c[l], m->da, m->dP, m->dP16, m->part));
        // per-wave partial rows hold dQ (3 x 64, columns 0..191) and dc (192..255)
        M_TRY(mpnn_reduce_rows(s, kAggWaves, 256, 0, 192, m->part, g + L.Q[l]));
        M_TRY(mpnn_reduce_rows(s, kAggWaves, 256, 192, 64, m->part, g + L.c[l]));
        M_TRY(mpnn_gemm_tn(s, 64, 64, N, m->h16[l], m->dP16, m->part, chunks, kChunk, g + L.W1[l], 64, 64, 64));
        M_TRY(mpnn_gemm_nt(s, 64, 64, true, N, m->dP16, m->w.W1[l], m->dh));
    }
    // encoder: du0 = dh * silu'(u0)
    M_TRY(mpnn_dsilu_mul(s, n64, m->dh, m->u0, m->dz, m->dz16));
    M_TRY(mpnn_gemm_tn(s, 32, 64, N, m->x16, m->dz16, m->part, chunks, kChunk, g + L.W0, kNodeFeat, 64, 64));
    M_TRY(mpnn_colsum(s, N, 64, m->dz, m->part, cchunks, 1024, 64, g + L.b0));
    return PD_OK;
}

int loss_value(pd_mpnn *m, float *loss)
{
    const int cchunks = (m->N + 1023) / 1024;
    M_TRY(mpnn_colsum(m->s, m->N, 1, m->err_n, m->part, cchunks, 1024, 1, m->scal));
    float v = 0.f;
    M_TRY(hipMemcpyAsync(&v, m->scal, sizeof(float), hipMemcpyDeviceToHost, m->s));
    M_TRY(hipStreamSynchronize(m->s));
    *loss = m->wsum > 0 ? (float)(v / m->wsum) : 0.f;
    return PD_OK;
}

} // namespace

extern "C" {

int pd_mpnn_create(int device, uint64_t seed, pd_mpnn **out)
{
    if (!out) return PD_E_INVALID;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return PD_E_HIP;
    pd_mpnn *m = new pd_mpnn();
    m->device = device;
    const int n = m->L.total;
    auto fail = [&]() {
        pd_mpnn_destroy(m);
        return PD_E_HIP;
    };
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) return fail();
    if (hipMalloc(&m->theta, sizeof(float) * n) != hipSuccess || hipMalloc(&m->grad, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&m->am, sizeof(float) * n) != hipSuccess || hipMalloc(&m->av, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&m->scal, sizeof(float) * 8) != hipSuccess)
        return fail();
    // bf16 weight operands
    const size_t wcount = 64 * 32 + 3 * 6 * 64 * 64 + 16 * 64 + 64 * 32;
    if (hipMalloc(&m->wbuf, sizeof(__bf16) * wcount) != hipSuccess) return fail();
    __bf16 *p = m->wbuf;
    m->w.W0T = p; p += 64 * 32;
    for (int l = 0; l < 3; ++l) {
        m->w.W1T[l] = p; p += 4096;
        m->w.W1[l] = p; p += 4096;
        m->w.W2aT[l] = p; p += 4096;
        m->w.W2bT[l] = p; p += 4096;
        m->w.W2a[l] = p; p += 4096;
        m->w.W2b[l] = p; p += 4096;
    }
    m->w.WdT = p; p += 16 * 64;
    m->w.Wd = p;
    // He-uniform weights from a seeded SplitMix64 stream, zero biases
    std::vector<float> th(n, 0.f);
    uint64_t st = seed;
    auto fill = [&](int off, int fan_in, int count) {
        const double bound = std::sqrt(3.0 / fan_in);
        for (int k = 0; k < count; ++k) {
            double u = (double)(sm64(&st) >> 11) * (1.0 / 9007199254740992.0);
            th[off + k] = (float)((2.0 * u - 1.0) * bound);
        }
    };
    const ParamLayout &L = m->L;
    fill(L.W0, kNodeFeat, kNodeFeat * 64);
    for (int l = 0; l < 3; ++l) {
        fill(L.W1[l], 64, 64 * 64);
        fill(L.Q[l], 3, 3 * 64);
        fill(L.W2[l], 128, 128 * 64);
    }
    fill(L.Wd, 64, 64 * 2);
    if (copy_on(m->s, m->theta, th.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        fill_on(m->s, m->am, 0, sizeof(float) * n) != hipSuccess || fill_on(m->s, m->av, 0, sizeof(float) * n) != hipSuccess)
        return fail();
    *out = m;
    return PD_OK;
}

int pd_mpnn_destroy(pd_mpnn *m)
{
    if (!m) return PD_E_INVALID;
    (void)hipSetDevice(m->device);
    if (m->s) (void)hipStreamSynchronize(m->s);
    free_graph(m);
    mfree(m->theta); mfree(m->grad); mfree(m->am); mfree(m->av); mfree(m->wbuf); mfree(m->scal);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
    return PD_OK;
}

const char *pd_mpnn_last_error(pd_mpnn *m) { return m ? m->err.c_str() : "null network"; }

int pd_mpnn_num_params(pd_mpnn *m, int64_t *n)
{
    if (!m || !n) return PD_E_INVALID;
    *n = m->L.total;
    return PD_OK;
}

int pd_mpnn_get_params(pd_mpnn *m, float *out)
{
    if (!m || !out) return PD_E_INVALID;
    M_TRY(hipStreamSynchronize(m->s));
    M_TRY(copy_on(m->s, out, m->theta, sizeof(float) * m->L.total, hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_mpnn_set_params(pd_mpnn *m, const float *in)
{
    if (!m || !in) return PD_E_INVALID;
    M_TRY(hipStreamSynchronize(m->s));
    M_TRY(copy_on(m->s, m->theta, in, sizeof(float) * m->L.total, hipMemcpyHostToDevice));
    M_TRY(fill_on(m->s, m->am, 0, sizeof(float) * m->L.total));
    M_TRY(fill_on(m->s, m->av, 0, sizeof(float) * m->L.total));
    m->adam_t = 0;
    return PD_OK;
}

int pd_mpnn_set_graph(pd_mpnn *m, int n_nodes, const int32_t *row_ptr, int64_t n_edges, const int32_t *col)
{
    // Symmetric CSR graph (j in N(i) <=> i in N(j)), rows ascending -- the PD
    // bond list.  The reverse-edge index turns the backward scatter into a gather.
    if (!m || n_nodes <= 0 || !row_ptr || (n_edges > 0 && !col)) return PD_E_INVALID;
    if (row_ptr[0] != 0 || row_ptr[n_nodes] != n_edges) {
        m->err = "row_ptr does not match n_edges";
        return PD_E_INVALID;
    }
    std::vector<int32_t> rev((size_t)n_edges);
    std::vector<float> dinv(n_nodes);
    for (int i = 0; i < n_nodes; ++i) {
        const int deg = row_ptr[i + 1] - row_ptr[i];
        dinv[i] = deg > 0 ? 1.0f / (float)deg : 0.0f;
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            const int j = col[k];
            if (j < 0 || j >= n_nodes) {
                m->err = "column index out of range";
                return PD_E_INVALID;
            }
            const int32_t *b = col + row_ptr[j], *e = col + row_ptr[j + 1];
            const int32_t *hit = std::lower_bound(b, e, i);
            if (hit == e || *hit != i) {
                m->err = "graph is not symmetric or rows are not sorted";
                return PD_E_INVALID;
            }
            rev[k] = (int32_t)(hit - col);
        }
    }
    M_TRY(hipStreamSynchronize(m->s));  // no kernel may still read what ensure_capacity replaces
    int rc = ensure_capacity(m, n_nodes, n_edges);
    if (rc) return rc;
    m->N = n_nodes;
    m->E = n_edges;
    M_TRY(copy_on(m->s, m->rp, row_ptr, sizeof(int32_t) * (n_nodes + 1), hipMemcpyHostToDevice));
    if (n_edges) {
        M_TRY(copy_on(m->s, m->col, col, sizeof(int32_t) * n_edges, hipMemcpyHostToDevice));
        M_TRY(copy_on(m->s, m->rev, rev.data(), sizeof(int32_t) * n_edges, hipMemcpyHostToDevice));
    }
    M_TRY(copy_on(m->s, m->dinv, dinv.data(), sizeof(float) * n_nodes, hipMemcpyHostToDevice));
    return PD_OK;
}

int pd_mpnn_set_features(pd_mpnn *m, const float *node_feat, const float *edge_feat, const float *target,
                         const float *weight)
{
    // node_feat [N][kNodeFeat], edge_feat [E][3]; target [N][2] and weight [N]
    // may be NULL for inference.
    if (!m || !node_feat || (m->E > 0 && !edge_feat) || m->N <= 0) return PD_E_INVALID;
    M_TRY(hipStreamSynchronize(m->s));
    const int N = m->N;
    std::vector<__bf16> x((size_t)N * 32);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 32; ++c)
            x[(size_t)i * 32 + c] = (__bf16)(c < kNodeFeat ? node_feat[(size_t)i * kNodeFeat + c] : 0.f);
    M_TRY(copy_on(m->s, m->x16, x.data(), sizeof(__bf16) * x.size(), hipMemcpyHostToDevice));
    if (m->E) M_TRY(copy_on(m->s, m->ef, edge_feat, sizeof(float) * 3 * m->E, hipMemcpyHostToDevice));
    if (target && weight) {
        M_TRY(copy_on(m->s, m->tgt, target, sizeof(float) * 2 * (size_t)N, hipMemcpyHostToDevice));
        M_TRY(copy_on(m->s, m->wt, weight, sizeof(float) * N, hipMemcpyHostToDevice));
        double ws = 0.0;
        for (int i = 0; i < N; ++i) ws += weight[i];
        m->wsum = ws;
    } else {
        std::vector<float> ones(N, 1.f);
        M_TRY(copy_on(m->s, m->wt, ones.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        m->wsum = N;
    }
    return PD_OK;
}

int pd_mpnn_forward(pd_mpnn *m, float *y_out)
{
    // y_out [N][2]
    if (!m || !y_out || m->N <= 0) return PD_E_INVALID;
    int rc = forward(m, false);
    if (rc) return rc;
    M_TRY(hipMemcpyAsync(y_out, m->yout, sizeof(float) * 2 * (size_t)m->N, hipMemcpyDeviceToHost, m->s));
    M_TRY(hipStreamSynchronize(m->s));
    return PD_OK;
}

int pd_mpnn_gradient(pd_mpnn *m, float *grad_out, float *loss_out)
{
    // Loss of the current features/targets and its gradient w.r.t. the parameters.
    if (!m || m->N <= 0) return PD_E_INVALID;
    int rc = forward(m, true);
    if (!rc) rc = backward(m);
    if (!rc && loss_out) rc = loss_value(m, loss_out);
    if (rc) return rc;
    if (grad_out) {
        M_TRY(hipMemcpyAsync(grad_out, m->grad, sizeof(float) * m->L.total, hipMemcpyDeviceToHost, m->s));
        M_TRY(hipStreamSynchronize(m->s));
    }
    return PD_OK;
}

int pd_mpnn_train_step(pd_mpnn *m, float lr, float *loss_out)
{
    // One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the current sample.
    if (!m || m->N <= 0 || !(lr > 0.f)) return PD_E_INVALID;
    int rc = forward(m, true);
    if (!rc) rc = backward(m);
    if (rc) return rc;
    m->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)m->adam_t));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)m->adam_t));
    M_TRY(mpnn_adam(m->s, m->L.total, m->theta, m->grad, m->am, m->av, lr, b1, b2, 1e-8f, c1, c2));
    if (loss_out) return loss_value(m, loss_out);
    return PD_OK;
}

} // extern "C"

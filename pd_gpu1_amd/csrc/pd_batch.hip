// pd_batch.hip -- host side of the batched quasi-static ADR entry points.
//
// A batch concatenates independent square 2D PMB plates with boundary strips
// (the Coord.cu:47-66 layout), each with its own centred, angled pre-crack
// (Base_Function.cu:83-123) and local CSR bonds.  Every specimen is set up
// with the single-problem pipeline of pd_api.hip (cell-list CSR, volume
// correction, crack) and copied into the batch arrays; bond ranges start on
// 32-bond boundaries so fail words copy as whole words.
#include "pd_batch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

struct pd_batch {
    pd_params P{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<BatchSpec> spec;
    BatchSpec *d_spec = nullptr;
    int32_t *d_order = nullptr;   // queue order: specimens by bond count, largest first
    int64_t n_total = 0, nb_total = 0, rp_total = 0;
    double *buf = nullptr;        // 16 particle fields of n_total doubles
    double2 *pos = nullptr;       // current positions, refreshed by the kernel
    int32_t *row_ptr = nullptr, *col = nullptr;
    double2 *kd = nullptr;        // per bond {kappa, |xi|}
    double2 *xy = nullptr;        // per particle reference position (x, y)
    int coef_ok = 0;              // every stored {kappa, |xi|} equals its recomputation from xy
    int coef_mode = -1;           // -1: recompute when coef_ok, 0: stream kd, 1: recompute
    uint32_t *fail = nullptr;
    unsigned *queue = nullptr;
    int32_t *d_iters = nullptr;
    double *d_resid = nullptr;
    int32_t *d_retired = nullptr; // per specimen: 1 once an increment ended unconverged, 2 separated (retire on)
    double *d_react = nullptr;    // per specimen: strip reaction at its last converged increment
    int32_t *d_order2 = nullptr;  // queue of the paused specimens a tail launch continues
    int tail = 32;                // widest tail launch in workgroups per specimen (0: none)
    double sep = 0.0;             // separation ratio (0: off)
    int *d_active = nullptr;      // specimens of a launch not yet ended or paused
    int32_t *d_paused = nullptr;  // [nspec] specimens a launch paused
    int retire = 0;
    int last_block = 1024;
    int group = 1;                // lanes per row of the force phase (1 = thread per particle)
    int parts = 0;                // workgroups per specimen (0 = sized from the batch's work)
    int last_parts = 1;           // the value the last increment ran with
    int last_tail_parts = 0;      // workgroups per specimen of its second launch (0: none)
    int groups_cap = 0;           // groups the group buffers hold
    unsigned *g_bar = nullptr;    // [groups_cap] barrier arrival counters
    int *g_spec = nullptr;        // [groups_cap] specimen of each group
    double *g_sum = nullptr;      // [groups_cap][kMaxParts][4] damping subtree roots
    int *g_abort = nullptr;       // a group barrier timed out
    int *g_flag = nullptr;        // [groups_cap] pause decision of each group
    double *snap = nullptr;       // state before a multi-workgroup increment: 14 particle fields,
    double2 *snap_pos = nullptr;  // positions, fail words and retired flags, restored when
    uint32_t *snap_fail = nullptr;  // the increment is abandoned at a barrier timeout
    int32_t *snap_ret = nullptr;
    double *snap_react = nullptr;
    double barrier_timeout = 20.0;  // seconds a workgroup waits at a group barrier
    int warm = 0;                 // 1: increments start from u + the previous increment
    int recompute() const { return coef_mode == 1 || (coef_mode == -1 && coef_ok); }
    BatchView view() const
    {
        BatchView v{};
        v.nspec = (int)spec.size();
        v.spec = d_spec;
        v.order = d_order;
        double *f[16];
        for (int k = 0; k < 16; k++) f[k] = buf + (size_t)k * n_total;
        v.x = f[0]; v.y = f[1]; v.ux = f[2]; v.uy = f[3]; v.fx = f[4]; v.fy = f[5];
        v.vx = f[6]; v.vy = f[7]; v.vhx = f[8]; v.vhy = f[9]; v.vhox = f[10]; v.vhoy = f[11];
        v.fox = f[12]; v.foy = f[13];
        v.dupx = f[14]; v.dupy = f[15];
        v.pos = pos;
        v.warm = warm;
        v.retired = retire ? d_retired : nullptr;
        v.nqueue = (int)spec.size();
        v.resume = 0;
        v.it0 = nullptr;
        v.pause_below = 0;
        v.active = d_active;
        v.paused = d_paused;
        v.sep = sep;
        v.react = d_react;
        v.row_ptr = row_ptr; v.col = col; v.kd = kd; v.fail = fail;
        v.xy = xy;
        v.recompute = recompute();
        v.coef = BatchCoef{P.delta, P.dx, P.vol, P.bc};
        v.mass = P.mass; v.dt = P.dt; v.s0 = P.s0;
        return v;
    }
};

#define B_TRY(expr)                                                               \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            b->err = std::string(#expr) + ": " + hipGetErrorString(_e);           \
            return PD_E_HIP;                                                      \
        }                                                                         \
    } while (0)

static void batch_free(pd_batch *b)
{
    void *ptrs[] = {b->d_spec, b->d_order, b->buf, b->pos, b->row_ptr, b->col, b->kd, b->xy, b->fail, b->queue,
                    b->d_iters, b->d_resid, b->d_retired, b->d_react, b->d_order2, b->snap_react, b->d_active,
                    b->d_paused,
                    b->g_flag, b->g_bar, b->g_spec,
                    b->g_sum, b->g_abort, b->snap, b->snap_pos,
                    b->snap_fail, b->snap_ret};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

extern "C" {

int pd_batch_create(const pd_params *p, const pd_specimen *specs, int nspec, int device, pd_batch **out)
{
    if (!p || !specs || nspec <= 0 || !out || p->dim != 2 || p->model != PD_MODEL_PMB) return PD_E_INVALID;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return PD_E_HIP;
    pd_batch *b = new pd_batch();
    b->P = *p;
    b->P.precision = PD_PREC_FP64;
    b->device = device;
    auto fail_out = [&](int rc) {
        batch_free(b);
        delete b;
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail_out(PD_E_HIP);
    // pass 1: build every specimen with the single-problem pipeline
    std::vector<pd_handle *> hs(nspec, nullptr);
    int rc = PD_OK;
    for (int s = 0; s < nspec && rc == PD_OK; s++) {
        const int nx = specs[s].nx, ny = specs[s].ny;
        if (nx < 4 || ny < 4) {
            rc = PD_E_INVALID;
            break;
        }
        pd_params q = b->P;
        q.n_top = q.n_bot = 3 * nx;
        const int n = nx * (ny + 6);
        std::vector<double> x(n), y(n);
        pd_gen_plate_strips(nx, ny, q.dx, x.data(), y.data());
        rc = pd_create(&q, device, &hs[s]);
        if (!rc) rc = pd_set_positions(hs[s], x.data(), y.data(), nullptr, nx * ny, n);
        if (!rc) rc = pd_build_neighbors(hs[s]);
        if (!rc) rc = pd_init_state(hs[s], PD_INIT_STATIC);
        if (!rc && specs[s].crack_len > 0.0)
            rc = pd_set_crack_2d(hs[s], specs[s].crack_len, 0.0, 0.0, specs[s].crack_angle);
        if (rc && hs[s]) b->err = std::string("specimen ") + std::to_string(s) + ": " + pd_last_error(hs[s]);
        if (!rc) {
            BatchSpec d{};
            d.n = n;
            d.n_int = nx * ny;
            d.n_top = d.n_bot = 3 * nx;
            d.p_off = b->n_total;
            d.rp_off = b->rp_total;
            d.b_off = b->nb_total;
            d.nb = hs[s]->Nb;
            b->n_total += n;
            b->rp_total += n + 1;
            b->nb_total += (d.nb + 31) / 32 * 32;
            b->spec.push_back(d);
        }
    }
    if (!rc && b->nb_total >= ((int64_t)1 << 31)) rc = PD_E_CAPACITY;
    // pass 2: allocate the batch and copy the specimens in
    auto alloc = [&](void **p, size_t bytes) {
        return hipMalloc(p, bytes ? bytes : 16) == hipSuccess ? PD_OK : PD_E_HIP;
    };
    if (!rc) rc = alloc((void **)&b->buf, sizeof(double) * 16 * b->n_total);
    if (!rc) rc = alloc((void **)&b->pos, sizeof(double2) * b->n_total);
    if (!rc) rc = alloc((void **)&b->row_ptr, sizeof(int32_t) * b->rp_total);
    if (!rc) rc = alloc((void **)&b->col, sizeof(int32_t) * b->nb_total);
    if (!rc) rc = alloc((void **)&b->kd, sizeof(double2) * b->nb_total);
    if (!rc) rc = alloc((void **)&b->xy, sizeof(double2) * b->n_total);
    if (!rc) rc = alloc((void **)&b->fail, sizeof(uint32_t) * (b->nb_total / 32 + 1));
    if (!rc) rc = alloc((void **)&b->d_spec, sizeof(BatchSpec) * nspec);
    if (!rc) rc = alloc((void **)&b->d_order, sizeof(int32_t) * nspec);
    if (!rc) rc = alloc((void **)&b->queue, sizeof(unsigned));
    if (!rc) rc = alloc((void **)&b->d_iters, sizeof(int32_t) * nspec);
    if (!rc) rc = alloc((void **)&b->d_resid, sizeof(double) * nspec);
    if (!rc) rc = alloc((void **)&b->d_retired, sizeof(int32_t) * nspec);
    if (!rc && fill_on(b->stream, b->d_retired, 0, sizeof(int32_t) * nspec) != hipSuccess) rc = PD_E_HIP;
    if (!rc) rc = alloc((void **)&b->d_react, sizeof(double) * nspec);
    if (!rc && fill_on(b->stream, b->d_react, 0, sizeof(double) * nspec) != hipSuccess) rc = PD_E_HIP;
    if (!rc) rc = alloc((void **)&b->d_order2, sizeof(int32_t) * nspec);
    if (!rc) rc = alloc((void **)&b->d_active, sizeof(int));
    if (!rc) rc = alloc((void **)&b->d_paused, sizeof(int32_t) * nspec);
    if (!rc && fill_on(b->stream, b->buf, 0, sizeof(double) * 16 * b->n_total) != hipSuccess) rc = PD_E_HIP;
    for (int s = 0; s < nspec && !rc; s++) {
        pd_handle *h = hs[s];
        const BatchSpec &d = b->spec[s];
        // make kappa current (vol*fac*bc*scr, scr = 1: no surface correction
        // in the crack-growth family, Solve_Plate_Crack_Grow.cu:110-117)
        pd_params q;
        pd_get_params(h, &q);
        q.precision = PD_PREC_FP64;
        pd_set_params(h, &q);
        if (pd_force(h) != PD_OK) {  // allocates kappa; the state is discarded below
            rc = PD_E_HIP;
            break;
        }
        (void)hipStreamSynchronize(h->stream);
        std::vector<uint32_t> fw((size_t)((d.nb + 31) / 32));
        if (!fw.empty()) pd_get_fail_bits(h, fw.data());
        hipError_t e = hipSuccess;
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (e == hipSuccess && bytes) e = copy_on(b->stream, dst, src, bytes, hipMemcpyDefault);
        };
        cp(b->buf + 0 * (size_t)b->n_total + d.p_off, h->x.p[0], sizeof(double) * d.n);
        cp(b->buf + 1 * (size_t)b->n_total + d.p_off, h->x.p[1], sizeof(double) * d.n);
        cp(b->row_ptr + d.rp_off, h->row_ptr, sizeof(int32_t) * (d.n + 1));
        cp(b->col + d.b_off, h->col, sizeof(int32_t) * d.nb);
        if (e == hipSuccess) e = launch_pack_bond_pairs(b->stream, d.nb, h->kappa, h->idist, b->kd + d.b_off);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        cp(b->fail + d.b_off / 32, fw.data(), sizeof(uint32_t) * fw.size());
        if (e != hipSuccess) {
            b->err = std::string("batch copy: ") + hipGetErrorString(e);
            rc = PD_E_HIP;
        }
    }
    if (!rc && copy_on(b->stream, b->d_spec, b->spec.data(), sizeof(BatchSpec) * nspec, hipMemcpyHostToDevice) != hipSuccess)
        rc = PD_E_HIP;
    if (!rc) {
        // reference positions as (x, y) pairs, and whether the kernel may
        // recompute every bond's {kappa, |xi|} from them bit for bit
        hipError_t e = launch_pack_bond_pairs(b->stream, b->n_total, b->buf, b->buf + b->n_total, b->xy);
        if (e == hipSuccess) e = fill_on(b->stream, b->queue, 0, sizeof(unsigned));
        const BatchCoef cg{b->P.delta, b->P.dx, b->P.vol, b->P.bc};
        for (int s = 0; s < nspec && e == hipSuccess; s++) {
            const BatchSpec &d = b->spec[s];
            e = launch_batch_coef_check(b->stream, d.n, b->row_ptr + d.rp_off, b->col + d.b_off, b->kd + d.b_off,
                                        b->xy + d.p_off, cg, b->queue);
        }
        unsigned bad = 1;
        if (e == hipSuccess) e = copy_on(b->stream, &bad, b->queue, sizeof(unsigned), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            b->err = std::string("batch coefficients: ") + hipGetErrorString(e);
            rc = PD_E_HIP;
        }
        b->coef_ok = bad == 0;
    }
    if (!rc) {
        // the largest plates set the wall time of an increment: hand them out first
        std::vector<int32_t> ord((size_t)nspec);
        for (int s = 0; s < nspec; ++s) ord[s] = s;
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t c) { return b->spec[a].nb > b->spec[c].nb; });
        if (copy_on(b->stream, b->d_order, ord.data(), sizeof(int32_t) * nspec, hipMemcpyHostToDevice) != hipSuccess)
            rc = PD_E_HIP;
    }
    for (pd_handle *h : hs)
        if (h) pd_destroy(h);
    if (rc) {
        std::string msg = b->err;
        int code = rc;
        batch_free(b);
        delete b;
        (void)msg;
        return code;
    }
    *out = b;
    return PD_OK;
}

int pd_batch_destroy(pd_batch *b)
{
    if (!b) return PD_E_INVALID;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
    return PD_OK;
}

const char *pd_batch_last_error(pd_batch *b) { return b ? b->err.c_str() : "null batch"; }

int pd_batch_info(pd_batch *b, int64_t *n_particles, int64_t *n_bonds, int *nspec)
{
    if (!b) return PD_E_INVALID;
    if (n_particles) *n_particles = b->n_total;
    if (n_bonds) {
        int64_t nb = 0;
        for (const auto &d : b->spec) nb += d.nb;
        *n_bonds = nb;
    }
    if (nspec) *nspec = (int)b->spec.size();
    return PD_OK;
}

// Group buffers for `groups` groups of workgroups (barrier counters,
// specimen slots, damping subtree roots, abort flag).
static int ensure_groups(pd_batch *b, int groups)
{
    if (groups <= b->groups_cap) return PD_OK;
    B_TRY(hipStreamSynchronize(b->stream));
    void *old[] = {b->g_bar, b->g_spec, b->g_sum, b->g_abort, b->g_flag};
    for (void *q : old)
        if (q) (void)hipFree(q);
    b->g_bar = nullptr; b->g_spec = nullptr; b->g_sum = nullptr; b->g_abort = nullptr; b->g_flag = nullptr;
    b->groups_cap = 0;
    B_TRY(hipMalloc((void **)&b->g_bar, sizeof(unsigned) * groups));
    B_TRY(hipMalloc((void **)&b->g_spec, sizeof(int) * groups));
    B_TRY(hipMalloc((void **)&b->g_sum, sizeof(double) * 4 * kMaxParts * groups));
    B_TRY(hipMalloc((void **)&b->g_abort, sizeof(int)));
    B_TRY(hipMalloc((void **)&b->g_flag, sizeof(int) * groups));
    b->groups_cap = groups;
    return PD_OK;
}

// Snapshot (save = true) or restore of the state an increment starts from:
// particle fields 2..15, positions, fail words, retired flags and reactions.
static int batch_snapshot(pd_batch *b, bool save)
{
    const size_t nfield = (size_t)b->n_total * 14, nwords = (size_t)(b->nb_total + 31) / 32;
    const size_t ns = b->spec.size();
    if (!b->snap) {
        B_TRY(hipMalloc((void **)&b->snap, sizeof(double) * nfield));
        B_TRY(hipMalloc((void **)&b->snap_pos, sizeof(double2) * b->n_total));
        B_TRY(hipMalloc((void **)&b->snap_fail, sizeof(uint32_t) * nwords));
        B_TRY(hipMalloc((void **)&b->snap_ret, sizeof(int32_t) * ns));
        B_TRY(hipMalloc((void **)&b->snap_react, sizeof(double) * ns));
    }
    double *u0 = b->buf + (size_t)2 * b->n_total;  // fields 2..15 are contiguous
    struct { void *live, *copy; size_t bytes; } parts[] = {
        {u0, b->snap, sizeof(double) * nfield},
        {b->pos, b->snap_pos, sizeof(double2) * b->n_total},
        {b->fail, b->snap_fail, sizeof(uint32_t) * nwords},
        {b->d_retired, b->snap_ret, sizeof(int32_t) * ns},
        {b->d_react, b->snap_react, sizeof(double) * ns}};
    for (auto &q : parts)
        B_TRY(hipMemcpyAsync(save ? q.copy : q.live, save ? q.live : q.copy, q.bytes, hipMemcpyDeviceToDevice,
                             b->stream));
    if (!save) B_TRY(hipStreamSynchronize(b->stream));
    return PD_OK;
}

int pd_batch_increment(pd_batch *b, double du, double tol, int max_iters, int block, int grid,
                       int32_t *iters_out, double *resid_out, double *ms_out)
{
    // One load increment for every specimen: strips move by +du / -du, then
    // ADR iterates each specimen until its residual <= tol (or max_iters).
    //
    // With the tail launch on (pd_batch_set_tail), the increment runs in up
    // to two persistent launches: the first pauses every specimen that
    // reaches tail x the previous increment's median iteration count, and
    // the second continues only those, with more workgroups each (the
    // slowest specimens otherwise run on a few workgroups while the rest of
    // the device idles).  A paused specimen resumes at the iteration it
    // paused at, so results are identical to one launch.
    if (!b || max_iters < 0 || !(tol >= 0.0)) return PD_E_INVALID;
    if (block != 256 && block != 512 && block != 1024) {
        b->err = "block must be 256, 512 or 1024";
        return PD_E_INVALID;
    }
    B_TRY(hipSetDevice(b->device));
    const int nspec = (int)b->spec.size();
    // every workgroup resident at once (a persistent grid): the capacity
    int dev_cu = 0;
    B_TRY(hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, b->device));
    const bool rc = b->recompute() != 0;
    const int capacity = dev_cu * batch_resident_per_cu(block, b->group, rc, false);
    // The first launch of an increment runs the dense build (two 1024-thread
    // workgroups per CU) when it gives each specimen at most kDenseParts
    // workgroups; tail launches, which continue a few specimens with many
    // workgroups each, run the default one.
    const int capacity_dense = block == 1024 ? dev_cu * batch_resident_per_cu(block, b->group, rc, true) : 0;
    const auto dense_for = [&](int w) { return capacity_dense > capacity && w <= kDenseParts; };
    const auto cap_for = [&](int w) { return dense_for(w) ? capacity_dense : capacity; };
    int W = b->parts;
    const int wmax = std::min(kMaxFirstParts, block / 64);
    const int wtail = std::min(std::min(kMaxParts, b->tail), block / kMinPartLanes);
    if (W == 0) {
        std::vector<int32_t> ret((size_t)nspec, 0);
        if (b->retire)
            B_TRY(copy_on(b->stream, ret.data(), b->d_retired, sizeof(int32_t) * nspec, hipMemcpyDeviceToHost));
        std::vector<int64_t> nb((size_t)nspec);
        std::vector<int32_t> np((size_t)nspec);
        for (int q = 0; q < nspec; ++q) {
            nb[q] = b->spec[q].nb;
            np[q] = b->spec[q].n;
        }
        W = pd_batch_choose_parts(nb.data(), np.data(), ret.data(), nspec, capacity, wmax);
        if (dense_for(W)) {  // the rule again for the capacity of the dense build
            const int Wd = pd_batch_choose_parts(nb.data(), np.data(), ret.data(), nspec, capacity_dense, wmax);
            if (dense_for(Wd)) W = Wd;
        } else if (b->tail && grid <= 0 && W > kDenseParts) {
            // tail launches widen the slowest specimens later, so the first
            // launch can start with half the workgroups per specimen and run
            // twice the specimens at once
            W /= 2;
        }
    } else if (W > wmax) {
        b->err = "workgroups per specimen exceed block/64 (the damping tree's waves)";
        return PD_E_INVALID;
    }
    int groups = grid > 0 ? grid / W : cap_for(W) / W;
    if (W > 1 && groups > cap_for(W) / W) groups = cap_for(W) / W;  // a group's workgroups must all be resident
    if (groups > nspec) groups = nspec;
    if (groups < 1) {
        b->err = "grid holds fewer workgroups than one specimen's group";
        return PD_E_INVALID;
    }
    int clock_khz = 0;  // the wall clock that bounds a barrier wait
    B_TRY(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, b->device));
    const double ticks = b->barrier_timeout * (clock_khz > 0 ? clock_khz : 100000) * 1e3;
    const uint64_t spin = ticks < 1.0 ? 1ull : ticks > 1e18 ? (uint64_t)1e18 : (uint64_t)ticks;
    const bool tail = b->tail && grid <= 0 && W < wtail;
    // With several workgroups per specimen (or further launches) the
    // increment can be abandoned part-way at a barrier timeout (strips moved,
    // some specimens relaxed, others not): keep the state it starts from so
    // that it can be put back.
    const bool snapshot = W > 1 || tail;
    if (snapshot) {
        int rc = batch_snapshot(b, true);
        if (rc) return rc;
    }
    // one persistent launch of `gw` groups of `w` workgroups over the view's queue
    auto launch = [&](const BatchView &v, int w, int gw, bool dense) -> int {
        BatchGroups gr{w, nullptr, nullptr, nullptr, nullptr, nullptr, spin};
        if (w > 1) {
            int rc = ensure_groups(b, gw);
            if (rc) return rc;
            gr = BatchGroups{w, b->g_bar, b->g_spec, b->g_sum, b->g_flag, b->g_abort, spin};
            gr.xcd = (gw * w) % (8 * w) == 0;  // each group's workgroups on one XCD
            B_TRY(hipMemsetAsync(b->g_bar, 0, sizeof(unsigned) * gw, b->stream));
            B_TRY(hipMemsetAsync(b->g_abort, 0, sizeof(int), b->stream));
        }
        B_TRY(hipMemsetAsync(b->queue, 0, sizeof(unsigned), b->stream));
        if (v.pause_below > 0) {
            const int n_active = v.nqueue;
            B_TRY(hipMemcpyAsync(b->d_active, &n_active, sizeof(int), hipMemcpyHostToDevice, b->stream));
            B_TRY(hipMemsetAsync(b->d_paused, 0, sizeof(int32_t) * nspec, b->stream));
        }
        B_TRY(launch_batch_adr(b->stream, gw * w, block, b->group, dense, v, du, tol, max_iters, b->queue,
                               b->d_iters, b->d_resid, gr));
        B_TRY(hipStreamSynchronize(b->stream));  // n_active lives on this frame
        if (w > 1) {
            int aborted = 0;
            B_TRY(copy_on(b->stream, &aborted, b->g_abort, sizeof(int), hipMemcpyDeviceToHost));
            if (aborted) {
                int rc = batch_snapshot(b, false);
                if (rc) return rc;
                b->err = "a workgroup group barrier timed out (the group's workgroups were not all resident); the "
                         "increment was abandoned and every specimen restored to its state before it";
                return PD_E_HIP;
            }
        }
        return PD_OK;
    };
    // Tail launches: once every specimen has been handed out and at most
    // half of the launch's workgroups still run (pause_below specimens of w
    // workgroups), the running specimens pause at their next iteration and
    // continue in a new launch of the default build with as many workgroups
    // each as the device holds for them (at least twice as many, up to
    // wtail; after a dense launch the same number is enough), until w
    // reaches wtail.
    const auto pause_limit = [&](int w, bool dense) {
        return tail && w < wtail ? std::max(1, (dense ? capacity_dense : capacity) / (2 * w)) : 0;
    };
    HipEvents<2> ev;
    B_TRY(ev.create());
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    B_TRY(hipEventRecord(e0, b->stream));
    BatchView v = b->view();
    bool dense = dense_for(W);
    v.pause_below = pause_limit(W, dense);
    int w = W, gw = groups, tail_parts = 0;
    std::vector<int32_t> flags((size_t)nspec);
    for (;;) {
        int rc = launch(v, w, gw, dense);
        if (rc) return rc;
        if (v.pause_below == 0) break;
        B_TRY(copy_on(b->stream, flags.data(), b->d_paused, sizeof(int32_t) * nspec, hipMemcpyDeviceToHost));
        std::vector<int32_t> paused;
        for (int q = 0; q < nspec; ++q)
            if (flags[q]) paused.push_back(q);
        if (paused.empty()) break;
        std::stable_sort(paused.begin(), paused.end(),
                         [&](int32_t a, int32_t c) { return b->spec[a].nb > b->spec[c].nb; });
        int wn = w;
        while (2 * wn <= wtail && (int64_t)paused.size() * 2 * wn <= capacity) wn *= 2;
        B_TRY(copy_on(b->stream, b->d_order2, paused.data(), sizeof(int32_t) * paused.size(),
                      hipMemcpyHostToDevice));
        v.order = b->d_order2;
        v.nqueue = (int)paused.size();
        v.resume = 1;
        v.it0 = b->d_iters;
        // a launch that cannot widen is the last (after a dense one, the
        // default build at the same width may still widen later)
        v.pause_below = wn > w || dense ? pause_limit(wn, false) : 0;
        w = wn;
        dense = false;
        gw = std::min<int>((int)paused.size(), capacity / w);
        tail_parts = w;
    }
    B_TRY(hipEventRecord(e1, b->stream));
    B_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    B_TRY(hipEventElapsedTime(&ms, e0, e1));
    std::vector<int32_t> its((size_t)nspec);
    B_TRY(copy_on(b->stream, its.data(), b->d_iters, sizeof(int32_t) * nspec, hipMemcpyDeviceToHost));
    b->last_block = block;
    b->last_parts = W;
    b->last_tail_parts = tail_parts;
    if (ms_out) *ms_out = ms;
    if (iters_out) std::copy(its.begin(), its.end(), iters_out);
    if (resid_out) B_TRY(copy_on(b->stream, resid_out, b->d_resid, sizeof(double) * nspec, hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_batch_set_tail(pd_batch *b, int max_parts)
{
    // 16, 32 (default) or 64: an increment's slowest specimens finish in
    // further launches with up to max_parts workgroups each once the device
    // is half idle; 0: one persistent launch per increment.  Results are
    // identical.
    if (!b || (max_parts != 0 && max_parts != 16 && max_parts != 32 && max_parts != 64)) return PD_E_INVALID;
    b->tail = max_parts;
    return PD_OK;
}

int pd_batch_set_separation(pd_batch *b, double ratio)
{
    // 0 (default): off.  0 < ratio < 1: a specimen whose strip reaction falls
    // below ratio x its reaction at its last converged increment has
    // fractured through; its increment ends there and, with retire on, it is
    // retired with flag 2.
    if (!b || !(ratio >= 0.0 && ratio < 1.0)) return PD_E_INVALID;
    b->sep = ratio;
    return PD_OK;
}

int pd_batch_get_retired(pd_batch *b, int32_t *flags)
{
    // per specimen: 0 active, 1 retired unconverged, 2 retired separated
    if (!b || !flags) return PD_E_INVALID;
    B_TRY(copy_on(b->stream, flags, b->d_retired, sizeof(int32_t) * b->spec.size(), hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_batch_get_tail_parts(pd_batch *b, int *parts)
{
    // workgroups per specimen of the last increment's second launch (0: none)
    if (!b || !parts) return PD_E_INVALID;
    *parts = b->last_tail_parts;
    return PD_OK;
}

int pd_batch_set_coef_mode(pd_batch *b, int mode)
{
    // -1 (default): the force phase recomputes every bond's {kappa, |xi|}
    // from the reference positions when set-up found the recomputation equal
    // to the stored values bit for bit; 0: always streamed; 1: recomputed
    // (PD_E_INVALID when they differ).  Results are identical either way.
    if (!b || mode < -1 || mode > 1) return PD_E_INVALID;
    if (mode == 1 && !b->coef_ok) {
        b->err = "the stored bond coefficients differ from their recomputation";
        return PD_E_INVALID;
    }
    b->coef_mode = mode;
    return PD_OK;
}

int pd_batch_coef_info(pd_batch *b, int *recomputed)
{
    if (!b || !recomputed) return PD_E_INVALID;
    *recomputed = b->recompute();
    return PD_OK;
}

int pd_batch_set_warm(pd_batch *b, int mode)
{
    // 0: each increment starts from the previous solution with the strips
    // moved (plain ADR); 1: the free particles also move by the previous
    // increment (exact for a linear response, the first increment is plain).
    if (!b || (mode != 0 && mode != 1)) return PD_E_INVALID;
    b->warm = mode;
    return PD_OK;
}

int pd_batch_set_retire(pd_batch *b, int on)
{
    // 1: a specimen whose increment ends without reaching tol (it has
    // separated, or needs more than max_iters) is retired: later increments
    // skip it and report 0 iterations for it.
    if (!b || (on != 0 && on != 1)) return PD_E_INVALID;
    b->retire = on;
    return PD_OK;
}

int pd_batch_set_parts(pd_batch *b, int parts)
{
    // Workgroups per specimen: 1, 2, 4, 8 or 16 (at most block/64), or 0
    // (default) for the smallest power of two at which the largest active
    // specimen's share of the device (its work over W workgroups) is no
    // longer than the whole batch's.  Results are identical for every choice.
    if (!b) return PD_E_INVALID;
    if (parts < 0 || parts > kMaxFirstParts || (parts & (parts - 1))) {
        b->err = "parts must be 0, 1, 2, 4, 8 or 16";
        return PD_E_INVALID;
    }
    b->parts = parts;
    return PD_OK;
}

int pd_batch_choose_parts(const int64_t *nb, const int32_t *n, const int32_t *retired, int nspec, int capacity,
                          int wmax)
{
    // A specimen's work is its bonds times its ADR iterations, which grow
    // with the plate's side (sqrt of its particles).  The largest specimen
    // bounds the wall time from below by its work over W workgroups, the
    // whole batch by all work over the device's workgroups: the smallest
    // power of two W (at most wmax) that brings the first bound down to the
    // second.  Host arithmetic only.
    if (!nb || !n || nspec <= 0 || capacity <= 0 || wmax <= 0) return 1;
    double total = 0.0, largest = 0.0;
    for (int q = 0; q < nspec; ++q)
        if (!(retired && retired[q])) {
            const double work = (double)nb[q] * std::sqrt((double)n[q]);
            total += work;
            largest = std::max(largest, work);
        }
    int W = 1;
    while (2 * W <= wmax && largest * capacity > total * W) W *= 2;
    return W;
}

int pd_batch_set_barrier_timeout(pd_batch *b, double seconds)
{
    // Longest wait of a workgroup at a group barrier before the increment is
    // abandoned with an error (default 20 s).
    if (!b || !(seconds > 0.0)) return PD_E_INVALID;
    b->barrier_timeout = seconds;
    return PD_OK;
}

int pd_batch_get_parts(pd_batch *b, int *parts_used)
{
    if (!b || !parts_used) return PD_E_INVALID;
    *parts_used = b->last_parts;
    return PD_OK;
}

int pd_batch_set_group(pd_batch *b, int group)
{
    // Lanes per row in the force phase: 1 (thread per particle), 4, 8, 16 or
    // 32.  Every choice adds a row's bond terms in bond order.
    if (!b) return PD_E_INVALID;
    if (group != 1 && group != 4 && group != 8 && group != 16 && group != 32) {
        b->err = "group must be 1, 4, 8, 16 or 32";
        return PD_E_INVALID;
    }
    b->group = group;
    return PD_OK;
}

static int batch_field(pd_batch *b, int field, int *slot)
{
    switch (field) {
    case PD_FIELD_X0: *slot = 0; return PD_OK;
    case PD_FIELD_U: *slot = 2; return PD_OK;
    case PD_FIELD_F: *slot = 4; return PD_OK;
    case PD_FIELD_V: *slot = 6; return PD_OK;
    case PD_FIELD_VH: *slot = 8; return PD_OK;
    case PD_FIELD_VHO: *slot = 10; return PD_OK;
    case PD_FIELD_FO: *slot = 12; return PD_OK;
    default: b->err = "unsupported batch field"; return PD_E_INVALID;
    }
}

int pd_batch_get_field(pd_batch *b, int spec, int field, double *out)
{
    // [2][n] of specimen spec
    if (!b || !out || spec < 0 || spec >= (int)b->spec.size()) return PD_E_INVALID;
    int slot = 0;
    int rc = batch_field(b, field, &slot);
    if (rc) return rc;
    const BatchSpec &d = b->spec[spec];
    B_TRY(hipStreamSynchronize(b->stream));
    for (int c = 0; c < 2; c++)
        B_TRY(copy_on(b->stream, out + (size_t)c * d.n, b->buf + (size_t)(slot + c) * b->n_total + d.p_off,
                        sizeof(double) * d.n, hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_batch_set_field(pd_batch *b, int spec, int field, const double *in)
{
    if (!b || !in || spec < 0 || spec >= (int)b->spec.size()) return PD_E_INVALID;
    int slot = 0;
    int rc = batch_field(b, field, &slot);
    if (rc) return rc;
    if (slot == 0) {
        b->err = "reference coordinates are fixed";
        return PD_E_INVALID;
    }
    const BatchSpec &d = b->spec[spec];
    B_TRY(hipStreamSynchronize(b->stream));
    for (int c = 0; c < 2; c++)
        B_TRY(copy_on(b->stream, b->buf + (size_t)(slot + c) * b->n_total + d.p_off, in + (size_t)c * d.n,
                        sizeof(double) * d.n, hipMemcpyHostToDevice));
    return PD_OK;
}

int pd_batch_get_csr(pd_batch *b, int spec, int32_t *row_ptr, int32_t *col, uint32_t *fail_words)
{
    if (!b || spec < 0 || spec >= (int)b->spec.size()) return PD_E_INVALID;
    const BatchSpec &d = b->spec[spec];
    B_TRY(hipStreamSynchronize(b->stream));
    if (row_ptr) B_TRY(copy_on(b->stream, row_ptr, b->row_ptr + d.rp_off, sizeof(int32_t) * (d.n + 1), hipMemcpyDeviceToHost));
    if (col) B_TRY(copy_on(b->stream, col, b->col + d.b_off, sizeof(int32_t) * d.nb, hipMemcpyDeviceToHost));
    if (fail_words)
        B_TRY(copy_on(b->stream, fail_words, b->fail + d.b_off / 32, sizeof(uint32_t) * ((d.nb + 31) / 32),
                        hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_batch_spec_info(pd_batch *b, int spec, int *n, int *n_int, int64_t *nb)
{
    if (!b || spec < 0 || spec >= (int)b->spec.size()) return PD_E_INVALID;
    const BatchSpec &d = b->spec[spec];
    if (n) *n = d.n;
    if (n_int) *n_int = d.n_int;
    if (nb) *nb = d.nb;
    return PD_OK;
}

} // extern "C"

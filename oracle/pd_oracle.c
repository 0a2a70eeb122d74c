/*
 * pd_oracle.c -- FP64 CPU oracle (test infrastructure only; see pd_oracle.h).
 *
 * Each function restates one CPU twin of XiaoMmingW/PD_Gpu1 and cites the
 * reference file:line it follows.  Deliberate, documented departures:
 *   - bonds are CSR instead of the dense N x MN slot array; row order is the
 *     CPU neighbour search's ascending order (Neighbor.cu:44-58), so every
 *     per-row sum visits bonds in the reference's order;
 *   - fail[] covers every row, not only N_int rows (Memory.cu:39 sizes it
 *     N_int*MN although Force_Bond.cu:169-177 reads it for every i < N);
 *   - the ADR sums use the canonical tree of orc_sum_tree instead of the
 *     CPU twin's assignment bug (Integrate.cu:375-376 assigns, the GPU path
 *     Integrate.cu:108-129 sums);
 *   - the 3D LPS bulk modulus uses K = E/(3(1-2 nu)); ref_k_bug = 1 restores
 *     Base.cuh:175 (K = E/3/(1-2G));
 *   - the surface-correction angle factors default to the trig-free identity
 *     (scr_mode 0); scr_mode 1 evaluates the literal atan/acos/cos/sin form.
 */
#include "pd_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static inline double sq(double v) { return v * v; }     /* Base.cuh:20 */
static inline double cube(double v) { return v * v * v; } /* Base.cuh:21 */

/* ------------------------------------------------------------------ params */
void orc_cal_parameter(orc_params *p, int dim, int bs, double dt, double dx,
                       double thick, double length, double height, double width,
                       double emod, double pratio, double dens, double s0,
                       double delta_factor, int ref_k_bug)
{
    /* Base.cuh:114-183 */
    const double pi = acos(-1.0);                        /* Base.cuh:31 */
    memset(p, 0, sizeof(*p));
    p->dim = dim; p->bs = bs; p->dt = dt; p->dx = dx; p->thick = thick;
    p->emod = emod; p->pratio = pratio; p->dens = dens; p->s0 = s0;
    p->delta = delta_factor * dx;                        /* :116 */
    p->radij = dx / 2.0;                                 /* :117 */
    p->sedload = 0.5 * emod / (1.0 - pratio * pratio) * 1.0e-6; /* :118 */
    p->nx = (int)((length - dx / 2.0) / dx + 1);         /* :119 */
    p->nz = 1;
    if (bs == 1) {
        if (dim == 2) {                                  /* :130-136 */
            p->MN = 32;
            p->ny = (int)((height - dx / 2.0) / dx + 1);
            p->N = p->N_int = p->nx * p->ny;
            p->vol = dx * dx * thick;
            p->bc = 6.0 * emod / (pi * thick * cube(p->delta) * (1.0 - pratio));
        } else {                                         /* :137-145 */
            p->MN = 128;
            p->ny = (int)((height - dx / 2.0) / dx + 1);
            p->nz = (int)((width - dx / 2.0) / dx + 1);
            p->N = p->N_int = p->nx * p->ny * p->nz;
            p->vol = dx * dx * dx;
            p->bc = 6.0 * emod / (pi * p->delta * cube(p->delta) * (1.0 - 2.0 * pratio));
        }
        /* :150 */
        p->mass = 0.25 * dt * dt * (pi * p->delta * p->delta * p->thick) * (p->bc) / dx * 5.0;
    } else {
        p->G = 0.5 * emod / (1 + pratio);                /* :154 */
        if (dim == 2) {                                  /* :162-169 */
            p->MN = 32;
            p->ny = (int)((height - dx / 2.0) / dx + 1);
            p->N = p->N_int = p->nx * p->ny;
            p->thick = dx;
            p->vol = dx * dx * p->thick;
            p->K = 0.5 * emod / (1 - pratio);
        } else {                                         /* :170-177 */
            p->MN = 128;
            p->ny = (int)((height - dx / 2.0) / dx + 1);
            p->nz = (int)((width - dx / 2.0) / dx + 1);
            p->N = p->N_int = p->nx * p->ny * p->nz;
            if (ref_k_bug)
                p->K = emod / 3.0 / (1 - 2.0 * p->G);
            else
                p->K = emod / (3.0 * (1.0 - 2.0 * pratio));
            p->vol = dx * dx * dx;
        }
    }
}

/* ------------------------------------------------------------- coordinates */
void orc_coord_plate(int nx, int ny, double dx, double *x, double *y)
{
    /* Coord.cu:19-31 */
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (-ny / 2.0 + 0.5 + i) * dx;
        }
}

void orc_coord_plate_strips(int nx, int ny, double dx, double *x, double *y)
{
    /* Coord.cu:47-66: the plate continued 3 rows upward, then 3 rows below. */
    for (int i = 0; i < ny + 3; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (-ny / 2.0 + 0.5 + i) * dx;
        }
    for (int i = ny + 3; i < ny + 6; i++)
        for (int j = 0; j < nx; j++) {
            x[i * nx + j] = (-nx / 2.0 + 0.5 + j) * dx;
            y[i * nx + j] = (ny / 2.0 + 3 - 0.5 - i) * dx;
        }
}

int orc_coord_kw(int nx, int ny, int nz, double dx, double rad,
                 double *x, double *y, double *z)
{
    /* Coord.cu:77-129; returns the particle count (the reference returns
     * count-1 and its caller adds 1, Solve_Kalthoff_Winkler.cu:296). */
    int num = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double tx = (-(nx - 1) / 2.0 + i) * dx;
                double ty = (-(ny - 1) / 2.0 + j) * dx;
                double tz = (-nz / 2.0 + 0.5 + k) * dx;
                int keep;
                if (tx > (-rad - 1.1 * dx) && tx < (-rad + 0.1 * dx))
                    keep = ty < -1.0e-10;
                else if (tx > (rad - 0.1 * dx) && tx < (rad + 1.1 * dx))
                    keep = ty < -1.0e-10;
                else
                    keep = 1;
                if (keep) {
                    if (x) { x[num] = tx; y[num] = ty; z[num] = tz; }
                    num++;
                }
            }
    return num;
}

uint64_t orc_splitmix64(uint64_t *state)
{
    /* SplitMix64 (Steele, Lea, Flood 2014). */
    uint64_t zz = (*state += 0x9E3779B97F4A7C15ULL);
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ULL;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBULL;
    return zz ^ (zz >> 31);
}

void orc_coord_block_jitter(int nx, int ny, int nz, double dx, double jitter,
                            uint64_t seed, double *x, double *y, double *z)
{
    /* Lattice centred on the origin (the Coord.cu:27-28 rule per axis) plus a
     * uniform jitter in [-jitter*dx, +jitter*dx): one 53-bit uniform per
     * coordinate, particle-major, then x, y, z. */
    uint64_t st = seed;
    int64_t n = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++, n++) {
                double b[3] = {(-nx / 2.0 + 0.5 + i) * dx, (-ny / 2.0 + 0.5 + j) * dx,
                               (-nz / 2.0 + 0.5 + k) * dx};
                double o[3];
                for (int d = 0; d < 3; d++) {
                    double u53 = (double)(orc_splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
                    o[d] = b[d] + (2.0 * u53 - 1.0) * jitter * dx;
                }
                x[n] = o[0]; y[n] = o[1]; z[n] = o[2];
            }
}

/* --------------------------------------------------------------- neighbours */
int orc_neighbors_brute(int dim, int N, const double *x, const double *y,
                        const double *z, double delta, int MN, int *NN, int *NL)
{
    /* Neighbor.cu:42-80 (strict inequality, both directions appended). */
    const double d2 = delta * delta;
    for (int i = 0; i < N; i++) NN[i] = 0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; ++j) {
            double r2 = sq(x[j] - x[i]) + sq(y[j] - y[i]);
            if (dim == 3) r2 = r2 + sq(z[j] - z[i]);
            if (r2 < d2) {
                if (NN[i] >= MN || NN[j] >= MN) return -1;
                NL[i * MN + NN[i]++] = j;
                NL[j * MN + NN[j]++] = i;
            }
        }
    return 0;
}

typedef struct {
    int dim, ncell[3];
    double lo[3], cs;
    int64_t *start; /* [ncells+1] */
    int *items;     /* particle ids grouped by cell, ascending inside a cell */
} cell_grid;

static int cell_coord(const cell_grid *g, int d, double v)
{
    int c = (int)floor((v - g->lo[d]) / g->cs);
    if (c < 0) c = 0;
    if (c >= g->ncell[d]) c = g->ncell[d] - 1;
    return c;
}

static void grid_build(cell_grid *g, int dim, int N, const double *const *xs, double delta)
{
    g->dim = dim;
    g->cs = delta * (1.0 + 1e-9);
    for (int d = 0; d < 3; d++) {
        g->ncell[d] = 1; g->lo[d] = 0.0;
        if (d >= dim) continue;
        double lo = xs[d][0], hi = xs[d][0];
        for (int i = 1; i < N; i++) {
            if (xs[d][i] < lo) lo = xs[d][i];
            if (xs[d][i] > hi) hi = xs[d][i];
        }
        g->lo[d] = lo;
        g->ncell[d] = (int)floor((hi - lo) / g->cs) + 1;
    }
    int64_t nc = (int64_t)g->ncell[0] * g->ncell[1] * g->ncell[2];
    g->start = (int64_t *)calloc(nc + 1, sizeof(int64_t));
    g->items = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int *cid = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int i = 0; i < N; i++) {
        int c[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) c[d] = cell_coord(g, d, xs[d][i]);
        cid[i] = (c[2] * g->ncell[1] + c[1]) * g->ncell[0] + c[0];
        g->start[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; c++) g->start[c + 1] += g->start[c];
    int64_t *fillp = (int64_t *)malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
    memcpy(fillp, g->start, sizeof(int64_t) * nc);
    for (int i = 0; i < N; i++) g->items[fillp[cid[i]]++] = i;
    free(fillp);
    free(cid);
}

static void isort(int *a, int n)
{
    for (int k = 1; k < n; k++) {
        int v = a[k], m = k - 1;
        while (m >= 0 && a[m] > v) { a[m + 1] = a[m]; m--; }
        a[m + 1] = v;
    }
}

int orc_neighbors_cell(int dim, int N, const double *x, const double *y,
                       const double *z, double delta, int64_t *row_ptr, int *col)
{
    /* Same predicate and operand order as Neighbor.cu:51,72:
     * ((dx^2 + dy^2) + dz^2) < delta*delta.  (a-b)^2 == (b-a)^2 exactly, so
     * testing from either end of a pair gives the same answer. */
    const double *xs[3] = {x, y, z};
    const double d2 = delta * delta;
    cell_grid g;
    grid_build(&g, dim, N, xs, delta);
    int maxrow = 0;
    if (col == NULL) row_ptr[0] = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxrow)
    for (int i = 0; i < N; i++) {
        int c[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) c[d] = cell_coord(&g, d, xs[d][i]);
        int cnt = 0;
        int *out = col ? col + row_ptr[i] : NULL;
        int zlo = dim == 3 ? c[2] - 1 : 0, zhi = dim == 3 ? c[2] + 1 : 0;
        for (int cz = zlo; cz <= zhi; cz++) {
            if (cz < 0 || cz >= g.ncell[2]) continue;
            for (int cy = c[1] - 1; cy <= c[1] + 1; cy++) {
                if (cy < 0 || cy >= g.ncell[1]) continue;
                for (int cx = c[0] - 1; cx <= c[0] + 1; cx++) {
                    if (cx < 0 || cx >= g.ncell[0]) continue;
                    int64_t cc = ((int64_t)cz * g.ncell[1] + cy) * g.ncell[0] + cx;
                    for (int64_t q = g.start[cc]; q < g.start[cc + 1]; q++) {
                        int j = g.items[q];
                        if (j == i) continue;
                        double r2 = sq(x[j] - x[i]) + sq(y[j] - y[i]);
                        if (dim == 3) r2 = r2 + sq(z[j] - z[i]);
                        if (r2 < d2) {
                            if (out) out[cnt] = j;
                            cnt++;
                        }
                    }
                }
            }
        }
        if (out) isort(out, cnt);
        else row_ptr[i + 1] = cnt;
        if (cnt > maxrow) maxrow = cnt;
    }
    if (col == NULL)
        for (int i = 0; i < N; i++) row_ptr[i + 1] += row_ptr[i];
    free(g.start);
    free(g.items);
    return maxrow;
}

/* ---------------------------------------------------------- volume / surface */
void orc_vol_corr(int dim, int N, const int64_t *row_ptr, const int *col,
                  const double *x, const double *y, const double *z,
                  double delta, double dx, double *idist, double *fac)
{
    /* Surf_Corr.cu:312-337 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            int c = col[k];
            if (dim == 2)
                idist[k] = sqrt(sq(x[c] - x[i]) + sq(y[c] - y[i]));
            else
                idist[k] = sqrt(sq(x[c] - x[i]) + sq(y[c] - y[i]) + sq(z[c] - z[i]));
            if (idist[k] <= delta - dx / 2.0)
                fac[k] = 1.0;
            else if (idist[k] <= delta + dx / 2.0)
                fac[k] = (delta + dx / 2.0 - idist[k]) / dx;
            else
                fac[k] = 0.0;
        }
}

static void surface_F(int dim, int N, const int64_t *row_ptr, const int *col,
                      const double *const *xs, int axis, const double *idist,
                      const double *fac, double bc, double vol, double sedload,
                      double *fncst)
{
    /* Surf_Corr.cu:339-388 with the probe displacement of Disp (:345):
     * disp[axis] = 0.001*coord[axis], the other components zero. */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double st = 0.0;
        double ui[3] = {0, 0, 0};
        ui[axis] = 0.001 * xs[axis][i];
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            int c = col[k];
            double uc[3] = {0, 0, 0};
            uc[axis] = 0.001 * xs[axis][c];
            double r2 = sq(xs[0][c] + uc[0] - xs[0][i] - ui[0]) +
                        sq(xs[1][c] + uc[1] - xs[1][i] - ui[1]);
            if (dim == 3) r2 = r2 + sq(xs[2][c] + uc[2] - xs[2][i] - ui[2]);
            double nl = sqrt(r2);
            st += 0.25 * bc * (nl - idist[k]) * (nl - idist[k]) / idist[k] * vol * fac[k];
        }
        fncst[i] = sedload / st;
    }
}

void orc_surface_correct(int dim, int N, const int64_t *row_ptr, const int *col,
                         const double *x, const double *y, const double *z,
                         const double *idist, const double *fac,
                         double bc, double vol, double sedload, int scr_mode,
                         double *fncst, double *scr)
{
    /* Surf_Corr.cu:465-524 */
    const double pi = acos(-1.0);
    const double *xs[3] = {x, y, z};
    double *fn = fncst ? fncst : (double *)malloc(sizeof(double) * 3 * (size_t)N);
    for (int d = 0; d < dim; d++)
        surface_F(dim, N, row_ptr, col, xs, d, idist, fac, bc, vol, sedload, fn + (size_t)d * N);
    const double *fx = fn, *fy = fn + N, *fz = fn + 2 * (size_t)N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            int c = col[k];
            double adx = fabs(x[c] - x[i]), ady = fabs(y[c] - y[i]);
            if (dim == 2) {
                /* Surf_Corr.cu:411-426 */
                double c2, s2;
                if (ady <= 1.0e-10) { c2 = 1.0; s2 = 0.0; }
                else if (adx <= 1.0e-10) {
                    if (scr_mode) { double t = 90.0 * pi / 180.0; c2 = cos(t) * cos(t); s2 = sin(t) * sin(t); }
                    else { c2 = 0.0; s2 = 1.0; }
                } else if (scr_mode) {
                    double t = atan(ady / adx);
                    c2 = cos(t) * cos(t); s2 = sin(t) * sin(t);
                } else {
                    double a2 = adx * adx, b2 = ady * ady;
                    c2 = a2 / (a2 + b2); s2 = b2 / (a2 + b2);
                }
                double scx = (fx[i] + fx[c]) / 2.0;
                double scy = (fy[i] + fy[c]) / 2.0;
                scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy)));
            } else {
                /* Surf_Corr.cu:427-459 */
                double adz = fabs(z[c] - z[i]);
                if (adz <= 1.0e-10) {
                    double c2, s2, p2;
                    if (ady <= 1.0e-10) { c2 = 1.0; s2 = 0.0; }
                    else if (adx <= 1.0e-10) {
                        if (scr_mode) { double t = 90.0 * pi / 180.0; c2 = cos(t) * cos(t); s2 = sin(t) * sin(t); }
                        else { c2 = 0.0; s2 = 1.0; }
                    } else if (scr_mode) {
                        double t = atan(ady / adx);
                        c2 = cos(t) * cos(t); s2 = sin(t) * sin(t);
                    } else {
                        double a2 = adx * adx, b2 = ady * ady;
                        c2 = a2 / (a2 + b2); s2 = b2 / (a2 + b2);
                    }
                    if (scr_mode) { double ph = 90.0 * pi / 180.0; p2 = cos(ph) * cos(ph); }
                    else p2 = 0.0;
                    double scx = (fx[i] + fx[c]) / 2.0;
                    double scy = (fy[i] + fy[c]) / 2.0;
                    double scz = (fz[i] + fz[c]) / 2.0;
                    scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy) + p2 / (scz * scz)));
                } else if (adx <= 1.0e-10 && ady <= 1.0e-10) {
                    scr[k] = (fz[i] + fz[c]) / 2.0;
                } else {
                    double c2, s2, p2;
                    if (scr_mode) {
                        double t = atan(ady / adx);
                        double ph = acos(adz / idist[k]);
                        c2 = cos(t) * cos(t); s2 = sin(t) * sin(t); p2 = cos(ph) * cos(ph);
                    } else {
                        double a2 = adx * adx, b2 = ady * ady;
                        c2 = a2 / (a2 + b2); s2 = b2 / (a2 + b2);
                        double t = adz / idist[k];
                        p2 = t * t;
                    }
                    double scx = (fx[i] + fx[c]) / 2.0;
                    double scy = (fy[i] + fy[c]) / 2.0;
                    double scz = (fz[i] + fz[c]) / 2.0;
                    scr[k] = sqrt(1.0 / (c2 / (scx * scx) + s2 / (scy * scy) + p2 / (scz * scz)));
                }
            }
        }
    if (!fncst) free(fn);
}

/* ------------------------------------------------------------------- cracks */
void orc_set_crack_2d(int N_int, const int64_t *row_ptr, const int *col,
                      const double *x, const double *y, double clength,
                      double loc_x, double loc_y, double theta, uint8_t *fail)
{
    /* Base_Function.cu:83-123 */
    const double pi = acos(-1.0);
    const double t = tan(theta);
    const double sq1 = sqrt(1 + tan(theta) * tan(theta));
    const double half_s = clength * sin(theta) / 2.0;
    const int vertical = fabs(theta - pi / 2.0) < 0.0001;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            int c = col[k];
            if (vertical) {
                if (fabs(y[c] - loc_y) < half_s && fabs(y[i] - loc_y) < half_s)
                    if ((x[c] - loc_x) * (x[i] - loc_x) < 0.0) fail[k] = 0;
            } else {
                double dist_c = sq(y[c] - loc_y) + sq(x[c] - loc_x);
                double dist_i = sq(y[i] - loc_y) + sq(x[i] - loc_x);
                double vd_c = sq((t * x[c] - y[c] - loc_x * t + loc_y) / sq1);
                double vd_i = sq((t * x[i] - y[i] - loc_x * t + loc_y) / sq1);
                double len_c = sqrt(dist_c - vd_c);
                double len_i = sqrt(dist_i - vd_i);
                if (len_c < clength / 2.0 && len_i < clength / 2.0) {
                    double bcn = t * x[c] - loc_x * t + loc_y - y[c];
                    double bi = t * x[i] - loc_x * t + loc_y - y[i];
                    if (bcn * bi <= 0) fail[k] = 0;
                }
            }
        }
}

void orc_set_crack_kw(int N_int, const int64_t *row_ptr, const int *col,
                      const double *x, const double *y, double rad, uint8_t *fail)
{
    /* Solve_Kalthoff_Winkler.cu:65-87 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            int c = col[k];
            if ((x[c] > -rad && x[i] < -rad) | (x[i] > -rad && x[c] < -rad))
                if ((y[c] >= -1.0e-10) | (y[i] >= -1.0e-10)) fail[k] = 0;
            if ((x[c] > rad && x[i] < rad) | (x[i] > rad && x[c] < rad))
                if ((y[c] >= -1.0e-10) | (y[i] >= -1.0e-10)) fail[k] = 0;
        }
}

/* ---------------------------------------------------------------- PMB force */
void orc_pmb_force(int dim, int n_rows, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *scr,
                   double vol, double bc, double s0, double break_y_min,
                   uint8_t *fail, double *const *f)
{
    /* Force_Bond.cu:159-202.  2D breaks on s > s0 (:196); 3D also needs
     * y_i, y_j >= break_y_min (:194, hard-coded -0.04 there). */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_rows; i++) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            if (fail[k] != 1) continue;
            int c = col[k];
            double nl;
            if (dim == 2)
                nl = sqrt(sq(x[0][c] + u[0][c] - x[0][i] - u[0][i]) +
                          sq(x[1][c] + u[1][c] - x[1][i] - u[1][i]));
            else
                nl = sqrt(sq(x[0][c] + u[0][c] - x[0][i] - u[0][i]) +
                          sq(x[1][c] + u[1][c] - x[1][i] - u[1][i]) +
                          sq(x[2][c] + u[2][c] - x[2][i] - u[2][i]));
            double s = (nl - idist[k]) / idist[k];
            fx += vol * fac[k] * bc * scr[k] * s * (x[0][c] + u[0][c] - x[0][i] - u[0][i]) / nl;
            fy += vol * fac[k] * bc * scr[k] * s * (x[1][c] + u[1][c] - x[1][i] - u[1][i]) / nl;
            if (dim == 3) {
                fz += vol * fac[k] * bc * scr[k] * s * (x[2][c] + u[2][c] - x[2][i] - u[2][i]) / nl;
                if (s > s0 && x[1][i] >= break_y_min && x[1][c] >= break_y_min) fail[k] = 0;
            } else {
                if (s > s0) fail[k] = 0;
            }
        }
        f[0][i] = fx;
        f[1][i] = fy;
        if (dim == 3) f[2][i] = fz;
    }
}

/* ---------------------------------------------------------------------- LPS */
void orc_lps_weights(int N, const int64_t *row_ptr, const double *idist,
                     const double *fac, double delta, double vol, double *w, double *m)
{
    /* Force_State.cu:309-321 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double mi = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            w[k] = exp(-sq(idist[k]) / sq(delta));
            mi += w[k] * sq(idist[k]) * vol * fac[k];
        }
        m[i] = mi;
    }
}

void orc_lps_theta(int dim, int N, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *w,
                   const double *m, const uint8_t *fail, double vol, double pratio,
                   int theta_skip_broken, double *e, double *theta)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double th = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            if (theta_skip_broken && fail[k] != 1) continue;
            int c = col[k];
            if (dim == 2) {
                /* Force_State.cu:349-351 */
                e[k] = sqrt(sq(x[0][c] - x[0][i] + u[0][c] - u[0][i]) +
                            sq(x[1][c] - x[1][i] + u[1][c] - u[1][i])) - idist[k];
                th += 2.0 * (2.0 * pratio - 1.0) / (pratio - 1.0) / m[i] * w[k] * idist[k] * e[k] * fac[k] * vol;
            } else {
                /* Force_State.cu:221-224 */
                e[k] = sqrt(sq(x[0][c] - x[0][i] + u[0][c] - u[0][i]) +
                            sq(x[1][c] - x[1][i] + u[1][c] - u[1][i]) +
                            sq(x[2][c] - x[2][i] + u[2][c] - u[2][i])) - idist[k];
                th += 3.0 / m[i] * w[k] * idist[k] * e[k] * fac[k] * vol;
            }
        }
        theta[i] = th;
    }
}

void orc_lps_force(int dim, int N_int, const int64_t *row_ptr, const int *col,
                   const double *const *x, const double *const *u,
                   const double *idist, const double *fac, const double *w,
                   const double *m, const double *theta, const double *e,
                   double vol, double pratio, double K, double G, double s0,
                   double break_y_min, uint8_t *fail, double *const *f)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            if (fail[k] != 1) continue;
            int c = col[k];
            if (dim == 2) {
                /* Force_State.cu:370-386 */
                double nl = sqrt(sq(x[0][c] + u[0][c] - x[0][i] - u[0][i]) +
                                 sq(x[1][c] + u[1][c] - x[1][i] - u[1][i]));
                double s = (nl - idist[k]) / idist[k];
                double nxv = (x[0][c] + u[0][c] - x[0][i] - u[0][i]) / nl;
                double nyv = (x[1][c] + u[1][c] - x[1][i] - u[1][i]) / nl;
                double ti = (2.0 * pratio - 1.0) / (pratio - 1.0) * (2.0 * K - 32.0 * G / 9.0) * theta[i] * w[k] * idist[k] / m[i]
                            + 8.0 * G / m[i] * w[k] * e[k];
                double tj = (2.0 * pratio - 1.0) / (pratio - 1.0) * (2.0 * K - 32.0 * G / 9.0) * theta[c] * w[k] * idist[k] / m[c]
                            + 8.0 * G / m[c] * w[k] * e[k];
                fx += (ti * vol + tj * vol) * nxv * fac[k];
                fy += (ti * vol + tj * vol) * nyv * fac[k];
                if (s > s0) fail[k] = 0;
            } else {
                /* Force_State.cu:270-280 */
                double nl = sqrt(sq(x[0][c] + u[0][c] - x[0][i] - u[0][i]) +
                                 sq(x[1][c] + u[1][c] - x[1][i] - u[1][i]) +
                                 sq(x[2][c] - x[2][i] + u[2][c] - u[2][i]));
                double s = (nl - idist[k]) / idist[k];
                double ti = 3.0 * K * theta[i] * w[k] * idist[k] / m[i] + 15.0 * G / m[i] * w[k] * (e[k] - theta[i] * idist[k] / 3.0);
                double tj = 3.0 * K * theta[c] * w[k] * idist[k] / m[c] + 15.0 * G / m[c] * w[k] * (e[k] - theta[c] * idist[k] / 3.0);
                fx += (ti * vol + tj * vol) * fac[k] * (x[0][c] + u[0][c] - x[0][i] - u[0][i]) / nl;
                fy += (ti * vol + tj * vol) * fac[k] * (x[1][c] + u[1][c] - x[1][i] - u[1][i]) / nl;
                fz += (ti * vol + tj * vol) * fac[k] * (x[2][c] + u[2][c] - x[2][i] - u[2][i]) / nl;
                if (s > s0 && x[1][i] >= break_y_min && x[1][c] >= break_y_min) fail[k] = 0;
            }
        }
        f[0][i] = fx;
        f[1][i] = fy;
        if (dim == 3) f[2][i] = fz;
    }
}

/* ------------------------------------------------------------------- Verlet */
void orc_verlet_half1(int dim, int N_int, double dt, double *const *v,
                      const double *const *a, double *const *u)
{
    /* Integrate.cu:447-463 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++)
        for (int d = 0; d < dim; d++) {
            v[d][i] += (dt / 2.0) * a[d][i];
            u[d][i] += v[d][i] * dt;
        }
}

void orc_quick_load(int N_int, int n_top, int n_bot, double load, double dt, double *uy)
{
    /* Solve_Plate_Crack_Grow.cu:20-27 */
    for (int i = 0; i < n_top; i++) uy[N_int + i] += load * dt;
    for (int i = 0; i < n_bot; i++) uy[N_int + n_top + i] -= load * dt;
}

void orc_verlet_half2(int dim, int N_int, double dt, double dens, double *const *v,
                      double *const *a, const double *const *f, const double *const *b)
{
    /* Integrate.cu:465-481 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++)
        for (int d = 0; d < dim; d++) {
            a[d][i] = (f[d][i] + b[d][i]) / dens;
            v[d][i] += (dt / 2.0) * a[d][i];
        }
}

/* ---------------------------------------------------------------- reduction */
double orc_sum_tree(const double *v, int64_t n, int T)
{
    double *p = (double *)calloc((size_t)T, sizeof(double));
    for (int t = 0; t < T; t++) {
        double s = 0.0;
        for (int64_t k = t; k < n; k += T) s += v[k];
        p[t] = s;
    }
    for (int off = 1; off < T; off <<= 1)
        for (int t = 0; t < T; t += 2 * off) p[t] = p[t] + p[t + off];
    double r = p[0];
    free(p);
    return r;
}

/* ---------------------------------------------------------------------- ADR */
void orc_adr_terms(int dim, int N_int, double mass, double dt,
                   const double *const *u, const double *const *f,
                   const double *const *fold, const double *const *vho,
                   double *cn, double *disp)
{
    /* Integrate.cu:14-34 (kernel_cnn) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++) {
        double c = 0.0, dsp;
        if (vho[0][i] != 0.0)
            c = -1.0 * u[0][i] * u[0][i] * (f[0][i] / mass - fold[0][i] / mass) / (dt * vho[0][i]);
        dsp = u[0][i] * u[0][i];
        for (int d = 1; d < dim; d++) {
            if (vho[d][i] != 0.0)
                c -= 1.0 * u[d][i] * u[d][i] * (f[d][i] / mass - fold[d][i] / mass) / (dt * vho[d][i]);
            dsp += u[d][i] * u[d][i];
        }
        cn[i] = c;
        disp[i] = dsp;
    }
}

double orc_adr_c(double cn1, double cn2)
{
    /* Integrate.cu:117-128 */
    double cn;
    if (cn2 != 0.0) {
        if (cn1 / cn2 > 0.0) cn = 2.0 * sqrt(cn1 / cn2);
        else cn = 0.0;
    } else
        cn = 0.0;
    if (cn > 2.0) cn = 1.9;
    return cn;
}

void orc_adr_integrate(int dim, int N_int, int ct, double c, double mass, double dt,
                       const double *const *f, const double *const *b,
                       double *const *vh, double *const *v, double *const *vho,
                       double *const *u, double *const *fold)
{
    /* Integrate.cu:62-105 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++)
        for (int d = 0; d < dim; d++) {
            if (ct == 0)
                vh[d][i] = 1.0 * dt / mass * (f[d][i] + b[d][i]) / 2.0;
            else
                vh[d][i] = ((2.0 - c * dt) * vho[d][i] + 2.0 * dt / mass * (f[d][i] + b[d][i])) / (2.0 + c * dt);
            v[d][i] = 0.5 * (vho[d][i] + vh[d][i]);
            u[d][i] += vh[d][i] * dt;
            vho[d][i] = vh[d][i];
            fold[d][i] = f[d][i];
        }
}

void orc_resid_terms(int dim, int N_int, int N, const double *const *f,
                     const double *const *b, double *rnum, double *rden)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double a = 0.0, bb = 0.0;
        for (int d = 0; d < dim; d++) {
            if (i < N_int) {
                double r = f[d][i] + b[d][i];
                a += r * r;
                bb += b[d][i] * b[d][i];
            } else {
                bb += f[d][i] * f[d][i];
            }
        }
        rnum[i] = a;
        rden[i] = bb;
    }
}

/* ------------------------------------------------------------------- damage */
void orc_damage(int N_int, const int64_t *row_ptr, const uint8_t *fail,
                const double *fac, double vol, double *dmg)
{
    /* Base_Function.cu:220-238 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N_int; i++) {
        double p1 = 0.0, p2 = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            p1 += fail[k] * vol * fac[k];
            p2 += vol * fac[k];
        }
        dmg[i] = 1.0 - p1 / p2;
    }
}

/* ------------------------------------------------------------ KW cylinder */
void orc_cylinder_init(const double *cyl, double *cvel, double *cdisp,
                       double *cacc, double *cforce)
{
    /* Solve_Kalthoff_Winkler.cu:24-37 */
    for (int i = 0; i < 3; i++) {
        cacc[i] = 0.0; cdisp[i] = 0.0; cvel[i] = 0.0; cforce[i] = 0.0;
        if (i == 1) cvel[i] = cyl[3];
    }
}

void orc_cylinder_half1(double dt, double *cvel, double *cdisp, const double *cacc)
{
    /* Solve_Kalthoff_Winkler.cu:101-110 */
    for (int i = 0; i < 3; i++) {
        cvel[i] += (dt / 2.0) * cacc[i];
        cdisp[i] += cvel[i] * dt;
    }
}

void orc_contact(int N, const double *dmg, const double *const *x,
                 double *const *u, double *const *v, const double *cyl,
                 const double *cdisp, double dt, double dens, double vol,
                 double *const *react)
{
    /* Solve_Kalthoff_Winkler.cu:177-239; the reaction contributions are kept
     * per particle and summed by the caller with orc_sum_tree. */
    const double rad = cyl[0], height = cyl[2], cx = cyl[4], cy = cyl[5], cz = cyl[6];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        react[0][i] = 0.0; react[1][i] = 0.0; react[2][i] = 0.0;
        if (!(dmg[i] < 1.0)) continue;
        double dpx = x[0][i] + u[0][i] - cx - cdisp[0];
        double dpy = x[1][i] + u[1][i] - cy - cdisp[1];
        double dpz = x[2][i] + u[2][i] - cz - cdisp[2];
        double dpd = sqrt(dpx * dpx + dpz * dpz);
        if (dpd <= rad && fabs(dpy) <= (height / 2.0)) {
            double rx, ry, rz;
            if (rad - dpd <= height / 2.0 - fabs(dpy)) {
                rx = dpx / dpd * rad; ry = dpy; rz = dpz / dpd * rad;
            } else if (dpy >= 0.0) {
                rx = dpx; ry = height / 2.0; rz = dpz;
            } else {
                rx = dpx; ry = -height / 2.0; rz = dpz;
            }
            double oux = u[0][i], ouy = u[1][i], ouz = u[2][i];
            double ovx = v[0][i], ovy = v[1][i], ovz = v[2][i];
            u[0][i] = cx + cdisp[0] + rx - x[0][i];
            u[1][i] = cy + cdisp[1] + ry - x[1][i];
            u[2][i] = cz + cdisp[2] + rz - x[2][i];
            v[0][i] = (u[0][i] - oux) / dt;
            v[1][i] = (u[1][i] - ouy) / dt;
            v[2][i] = (u[2][i] - ouz) / dt;
            react[0][i] = -dens * (v[0][i] - ovx) / dt * vol;
            react[1][i] = -dens * (v[1][i] - ovy) / dt * vol;
            react[2][i] = -dens * (v[2][i] - ovz) / dt * vol;
        }
    }
}

void orc_cylinder_half2(double dt, double mass, double *cvel, double *cacc, double *cforce)
{
    /* Solve_Kalthoff_Winkler.cu:255-265 */
    for (int i = 0; i < 3; i++) {
        cacc[i] = cforce[i] / mass;
        cvel[i] += dt / 2.0 * cacc[i];
        cforce[i] = 0.0;
    }
}

/* ------------------------------------------------------------- threading */
void orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int orc_get_threads(void) { return omp_get_max_threads(); }

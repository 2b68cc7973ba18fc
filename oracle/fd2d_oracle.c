/*
 * fd2d_oracle.c -- FP64 CPU restatement of fd2d_noise's solvers (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the parity oracle of the HIP engine and the timed CPU baseline. It is loaded only by
 * the tests, the smoke check and the benchmark's CPU-baseline leg, never by the product path.
 *
 * It restates, in the same order of operations, the MATLAB code of
 *   code/run_forward1_green.m, code/run_forward2_correlation.m, code/run_noise_adjoint.m and
 *   code/propagation/{dx_v,dz_v,div_s,init_absbound,define_computational_domain,
 *                     integration_limits}.m
 * Arrays are MATLAB column-major with the first index fastest; loops are 0-based, and every
 * 1-based MATLAB range is written next to its 0-based form.
 *
 * Arithmetic types follow MATLAB: fields and accumulators are double. Where the reference applies
 * dx_v / dz_v to a single array (Born terms, run_forward1_green.m:141-142, run_forward2_correlation.m:
 * 236-237, run_noise_adjoint.m:289-290, and the forward snapshot in the structure kernel,
 * run_noise_adjoint.m:326-327) MATLAB evaluates the stencil in single precision and stores it into
 * a double array; the *_single primitives below do the same.
 *
 * Orders 6 and 8: the reference's dx_v / dz_v / div_s implement only orders 2 and 4 and silently
 * return zeros for any other order (dx_v.m:18-31). Orders 6 and 8 here use the textbook staggered
 * Taylor coefficients, with the zero frame generalised from the order-4 one (a derivative is set
 * where its whole stencil lies inside the array); they are an extension pinned by polynomial
 * exactness in the tests, not a restatement.
 *
 * Pinning (what fixes each function independently of this file; see tests/test_oracle_*.py):
 *   dx_v/dz_v/div_s   polynomial exactness and the zero frames;
 *   absbound          closed form of the Gaussian taper;
 *   green             closed-form 2-D SH line-force response, DFT-of-seismogram identity,
 *                     reciprocity in a heterogeneous medium;
 *   correlation       equivalence with a point-source adjoint run driven by an independently
 *                     synthesised source time function;
 *   adjoint           point/distributed source equivalence, stf_fft DFT identity, source-kernel
 *                     and structure-kernel identities built from receiver traces;
 *   Born terms        finite-difference consistency of the linearised runs (tolerance-based).
 */
#include "fd2d_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
static int g_threads = 1;

void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int oracle_get_threads(void) { return g_threads; }
const char* oracle_last_error(void) { return g_err; }

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static) num_threads(g_threads)")
#else
#define PAR_FOR
#endif

/* ---------------------------------------------------------------------------------------------
 * Stencil coefficients. Orders 2 and 4 are written out literally below, as in dx_v.m:23,29 and
 * div_s.m:24,33. Orders 6 and 8: staggered first-derivative Taylor coefficients
 * c_m for sum_m c_m (f(x+(m-1/2)h) - f(x-(m-1/2)h)) / h.
 * ------------------------------------------------------------------------------------------- */
static const double C6[3] = {75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0};
static const double C8[4] = {1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};

static int order_ok(int order) { return order == 2 || order == 4 || order == 6 || order == 8; }

/* Forward staggered x-derivative, out (nx-1) x nz: dx_v.m:18-31.
 * Order 2: i = 1..nx-1 (1-based). Order 4: i = 2..nx-2, rows 1 and nx-1 stay 0 (dx_v.m:28).
 * Order 2r (r = 3, 4): i = r..nx-r. */
void oracle_dx_v(const double* v, double dx, int nx, int nz, int order, double* out)
{
    const int mx = nx - 1;
    memset(out, 0, sizeof(double) * (size_t)mx * (size_t)nz);
    const int r = order / 2;
    PAR_FOR
    for (int j = 0; j < nz; ++j) {
        const double* c = v + (size_t)nx * j;
        double* o = out + (size_t)mx * j;
        if (order == 2) {
            for (int i = 0; i < nx - 1; ++i) o[i] = (c[i + 1] - c[i]) / dx;
        } else if (order == 4) {
            for (int i = 1; i <= nx - 3; ++i)
                o[i] = 9.0 * (c[i + 1] - c[i]) / (8.0 * dx) - (c[i + 2] - c[i - 1]) / (24.0 * dx);
        } else if (order == 6 || order == 8) {
            const double* cf = order == 6 ? C6 : C8;
            for (int i = r - 1; i <= nx - r - 1; ++i) {
                double s = 0.0;
                for (int m = 1; m <= r; ++m) s += cf[m - 1] * (c[i + m] - c[i - m + 1]);
                o[i] = s / dx;
            }
        }
    }
}

/* Forward staggered z-derivative, out nx x (nz-1): dz_v.m:49-63. */
void oracle_dz_v(const double* v, double dz, int nx, int nz, int order, double* out)
{
    memset(out, 0, sizeof(double) * (size_t)nx * (size_t)(nz - 1));
    const int r = order / 2;
    int jlo, jhi;
    if (order == 2) { jlo = 0; jhi = nz - 2; }
    else { jlo = r - 1; jhi = nz - r - 1; }
    if (!order_ok(order)) return;
    PAR_FOR
    for (int j = jlo; j <= jhi; ++j) {
        double* o = out + (size_t)nx * j;
        if (order == 2) {
            const double* a = v + (size_t)nx * (j + 1);
            const double* b = v + (size_t)nx * j;
            for (int i = 0; i < nx; ++i) o[i] = (a[i] - b[i]) / dz;
        } else if (order == 4) {
            const double* p1 = v + (size_t)nx * (j + 1);
            const double* p0 = v + (size_t)nx * j;
            const double* p2 = v + (size_t)nx * (j + 2);
            const double* m1 = v + (size_t)nx * (j - 1);
            for (int i = 0; i < nx; ++i)
                o[i] = 9.0 * (p1[i] - p0[i]) / (8.0 * dz) - (p2[i] - m1[i]) / (24.0 * dz);
        } else {
            const double* cf = order == 6 ? C6 : C8;
            for (int i = 0; i < nx; ++i) {
                double s = 0.0;
                for (int m = 1; m <= r; ++m)
                    s += cf[m - 1] * (v[i + (size_t)nx * (j + m)] - v[i + (size_t)nx * (j - m + 1)]);
                o[i] = s / dz;
            }
        }
    }
}

/* Single-precision evaluation of dx_v / dz_v on a single array, stored as double (MATLAB's
 * mixed single/double rules: the double operand is converted to single first). */
static void dx_v_single(const float* v, double dx, int nx, int nz, int order, double* out)
{
    const int mx = nx - 1;
    memset(out, 0, sizeof(double) * (size_t)mx * (size_t)nz);
    const int r = order / 2;
    const float fdx = (float)dx, f8dx = (float)(8.0 * dx), f24dx = (float)(24.0 * dx);
    PAR_FOR
    for (int j = 0; j < nz; ++j) {
        const float* c = v + (size_t)nx * j;
        double* o = out + (size_t)mx * j;
        if (order == 2) {
            for (int i = 0; i < nx - 1; ++i) o[i] = (double)((c[i + 1] - c[i]) / fdx);
        } else if (order == 4) {
            for (int i = 1; i <= nx - 3; ++i)
                o[i] = (double)(9.0f * (c[i + 1] - c[i]) / f8dx - (c[i + 2] - c[i - 1]) / f24dx);
        } else if (order == 6 || order == 8) {
            const double* cf = order == 6 ? C6 : C8;
            for (int i = r - 1; i <= nx - r - 1; ++i) {
                float s = 0.0f;
                for (int m = 1; m <= r; ++m) s += (float)cf[m - 1] * (c[i + m] - c[i - m + 1]);
                o[i] = (double)(s / fdx);
            }
        }
    }
}

static void dz_v_single(const float* v, double dz, int nx, int nz, int order, double* out)
{
    memset(out, 0, sizeof(double) * (size_t)nx * (size_t)(nz - 1));
    const int r = order / 2;
    int jlo, jhi;
    if (order == 2) { jlo = 0; jhi = nz - 2; }
    else { jlo = r - 1; jhi = nz - r - 1; }
    if (!order_ok(order)) return;
    const float fdz = (float)dz, f8dz = (float)(8.0 * dz), f24dz = (float)(24.0 * dz);
    PAR_FOR
    for (int j = jlo; j <= jhi; ++j) {
        double* o = out + (size_t)nx * j;
        for (int i = 0; i < nx; ++i) {
            const float* c = v + i;
            const size_t s0 = (size_t)nx;
            if (order == 2) {
                o[i] = (double)((c[s0 * (j + 1)] - c[s0 * j]) / fdz);
            } else if (order == 4) {
                o[i] = (double)(9.0f * (c[s0 * (j + 1)] - c[s0 * j]) / f8dz -
                                (c[s0 * (j + 2)] - c[s0 * (j - 1)]) / f24dz);
            } else {
                const double* cf = order == 6 ? C6 : C8;
                float s = 0.0f;
                for (int m = 1; m <= r; ++m)
                    s += (float)cf[m - 1] * (c[s0 * (j + m)] - c[s0 * (j - m + 1)]);
                o[i] = (double)(s / fdz);
            }
        }
    }
}

/* Backward staggered divergence, out nx x nz: div_s.m:81-103.
 * Order 2: x-part i = 2..nx-1, z-part j = 2..nz-1. Order 4: i = 3..nx-2, j = 3..nz-2.
 * Order 2r: i = r+1..nx-r, j = r+1..nz-r. The z-part is added to the x-part (div_s.m:90,100). */
void oracle_div_s(const double* sxy, const double* szy, double dx, double dz, int nx, int nz,
                  int order, double* out)
{
    const int mx = nx - 1;
    memset(out, 0, sizeof(double) * (size_t)nx * (size_t)nz);
    if (!order_ok(order)) return;
    const int r = order / 2;
    PAR_FOR
    for (int j = 0; j < nz; ++j) {
        double* o = out + (size_t)nx * j;
        const double* s = sxy + (size_t)mx * j;
        if (order == 2) {
            for (int i = 1; i <= nx - 2; ++i) o[i] = (s[i] - s[i - 1]) / dx;
        } else if (order == 4) {
            for (int i = 2; i <= nx - 3; ++i)
                o[i] = 9.0 * (s[i] - s[i - 1]) / (8.0 * dx) - (s[i + 1] - s[i - 2]) / (24.0 * dx);
        } else {
            const double* cf = order == 6 ? C6 : C8;
            for (int i = r; i <= nx - r - 1; ++i) {
                double a = 0.0;
                for (int m = 1; m <= r; ++m) a += cf[m - 1] * (s[i + m - 1] - s[i - m]);
                o[i] = a / dx;
            }
        }
    }
    int jlo = order == 2 ? 1 : r, jhi = order == 2 ? nz - 2 : nz - r - 1;
    if (order == 4) { jlo = 2; jhi = nz - 3; }
    PAR_FOR
    for (int j = jlo; j <= jhi; ++j) {
        double* o = out + (size_t)nx * j;
        for (int i = 0; i < nx; ++i) {
            const double* c = szy + i;
            const size_t sn = (size_t)nx;
            if (order == 2) {
                o[i] = o[i] + (c[sn * j] - c[sn * (j - 1)]) / dz;
            } else if (order == 4) {
                o[i] = o[i] + 9.0 * (c[sn * j] - c[sn * (j - 1)]) / (8.0 * dz) -
                       (c[sn * (j + 1)] - c[sn * (j - 2)]) / (24.0 * dz);
            } else {
                const double* cf = order == 6 ? C6 : C8;
                double a = 0.0;
                for (int m = 1; m <= r; ++m) a += cf[m - 1] * (c[sn * (j + m - 1)] - c[sn * (j - m)]);
                o[i] = o[i] + a / dz;
            }
        }
    }
}

/* Gaussian taper of init_absbound.m:16-30 on the grid of define_computational_domain.m:19-25
 * (X' is nx x nz with X'(i,j) = x(i) = (i-1)dx). */
void oracle_absbound(const fd2d_config* cfg, double* out)
{
    const int nx = cfg->nx, nz = cfg->nz;
    const double dx = cfg->Lx / (nx - 1), dz = cfg->Lz / (nz - 1);
    const double w = cfg->absorb_width, Lx = cfg->Lx, Lz = cfg->Lz;
    for (int j = 0; j < nz; ++j) {
        const double Z = j * dz;
        for (int i = 0; i < nx; ++i) {
            const double X = i * dx;
            double a = 1.0;
            if (cfg->absorb_left == 1)
                a = a * ((double)(X > w) + exp(-(X - w) * (X - w) / ((2 * w) * (2 * w))) * (double)(X <= w));
            if (cfg->absorb_right == 1)
                a = a * ((double)(X < (Lx - w)) +
                         exp(-(X - (Lx - w)) * (X - (Lx - w)) / ((2 * w) * (2 * w))) * (double)(X >= (Lx - w)));
            if (cfg->absorb_bottom == 1)
                a = a * ((double)(Z > w) + exp(-(Z - w) * (Z - w) / ((2 * w) * (2 * w))) * (double)(Z <= w));
            if (cfg->absorb_top == 1)
                a = a * ((double)(Z < (Lz - w)) +
                         exp(-(Z - (Lz - w)) * (Z - (Lz - w)) / ((2 * w) * (2 * w))) * (double)(Z >= (Lz - w)));
            out[i + (size_t)nx * j] = a;
        }
    }
}

/* interpol_matrix of run_forward2_correlation.m:121-165 (identical at run_noise_adjoint.m:175-219)
 * with the band limits of integration_limits.m:2-14. out is K x max(nb,1), column-major. */
int oracle_interpol_matrix(int K, int nb, double* out)
{
    if (nb == 0) {
        for (int k = 0; k < K; ++k) out[k] = 1.0;
        return FD2D_OK;
    }
    if (nb < 2 || nb > K)
        return fail(FD2D_EUNSUPPORTED, "n_basis_fct=%d: the reference's interpolation needs 2 <= n_basis_fct <= n_sample", nb);
    memset(out, 0, sizeof(double) * (size_t)K * (size_t)nb);
    const double step = (double)K / (double)nb;
    for (int k = 1; k <= K; ++k) {
        int ib = -1, lo = 0, hi = 0;
        for (int b = 1; b <= nb; ++b) {
            const int l = (int)floor((b - 1) * step + 1), h = (int)floor(b * step);
            if (k >= l && k <= h) { ib = b; lo = l; hi = h; break; }
        }
        if (ib < 0) return fail(FD2D_EINVAL, "frequency %d lies in no band", k);
        const int ni = hi - lo + 1;
        const int cur = ib;
        const int below = ib == 1 ? 0 : ib - 1;
        const int above = ib == 1 ? 2 : (ib == nb ? 0 : ib + 1);
        const double i = (double)(k - lo + 1), half = ni / 2.0;
#define IM(kk, bb) out[(size_t)((kk) - 1) + (size_t)K * (size_t)((bb) - 1)]
        if (i < half) {
            IM(k, cur) = 1 - (half - i) / ni;
            if (below) IM(k, below) = (half - i) / ni;
        } else if (i == half) {
            IM(k, cur) = 1;
        } else {
            IM(k, cur) = 1 - (i - half) / ni;
            if (above) IM(k, above) = (i - half) / ni;
        }
#undef IM
    }
    return FD2D_OK;
}

/* ------------------------------------------------------------------------------------------- */

static int64_t numel(const fd2d_carray* a)
{
    if (!a || !a->data) return 0;
    return a->dims[0] * a->dims[1] * a->dims[2];
}

static int64_t dim(const fd2d_carray* a, int d) { return numel(a) ? a->dims[d] : (d == 2 ? 1 : 0); }

/* min(find(min(abs(x-p)) == abs(x-p))) over x = 0:h:(n-1)h (run_forward1_green.m:48-49). */
static int nearest_node(double p, double h, int n)
{
    int best = 0;
    double bd = fabs(0.0 - p);
    for (int i = 1; i < n; ++i) {
        const double d = fabs(i * h - p);
        if (d < bd) { bd = d; best = i; }
    }
    return best;
}

static int node_ids(const fd2d_carray* pts, int nx, int nz, double dx, double dz, int** out, int64_t* n)
{
    *n = numel(pts) ? pts->dims[0] : 0;
    *out = NULL;
    if (*n == 0) return FD2D_OK;
    if (pts->dims[1] != 2) return fail(FD2D_EINVAL, "position arrays must be n x 2");
    const double* p = (const double*)pts->data;
    int* id = (int*)malloc(sizeof(int) * 2 * (size_t)*n);
    for (int64_t i = 0; i < *n; ++i) {
        id[2 * i] = nearest_node(p[i], dx, nx);
        id[2 * i + 1] = nearest_node(p[i + *n], dz, nz);
    }
    *out = id;
    return FD2D_OK;
}

typedef struct {
    int nx, nz, order, K, s, f;
    double dx, dz, dt, dw;
    const double* w;
} grid_t;

static int setup(const fd2d_config* cfg, const fd2d_carray* mu, const fd2d_carray* rho, grid_t* g)
{
    if (!cfg) return fail(FD2D_EINVAL, "config is NULL");
    g->nx = cfg->nx; g->nz = cfg->nz; g->order = cfg->order;
    g->K = cfg->n_sample; g->s = cfg->freq_samp; g->f = cfg->fw_nth;
    g->dx = cfg->Lx / (cfg->nx - 1); g->dz = cfg->Lz / (cfg->nz - 1);
    g->dt = cfg->dt; g->dw = cfg->dw; g->w = cfg->w_sample;
    if (g->nx < 3 || g->nz < 3) return fail(FD2D_EINVAL, "grid too small");
    if (!order_ok(g->order)) return fail(FD2D_EUNSUPPORTED, "FD order %d", g->order);
    if (g->s < 1 || g->f < 1 || cfg->nt < 1) return fail(FD2D_EINVAL, "freq_samp, fw_nth, nt must be >= 1");
    if (g->K > 0 && !g->w) return fail(FD2D_EINVAL, "w_sample is NULL");
    const int64_t P = (int64_t)g->nx * g->nz;
    if (numel(mu) != P || numel(rho) != P) return fail(FD2D_EINVAL, "mu/rho must have nx*nz elements");
    return FD2D_OK;
}

/* sigma_xy += dt*mu(1:nx-1,:).*ex ; sigma_zy += dt*mu(:,1:nz-1).*ez (run_forward1_green.m:171-172) */
static void stress_update(const grid_t* g, const double* mu, const double* ex, const double* ez,
                          double* sxy, double* szy)
{
    const int nx = g->nx, nz = g->nz, mx = nx - 1;
    const double dt = g->dt;
    PAR_FOR
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < mx; ++i)
            sxy[i + (size_t)mx * j] = sxy[i + (size_t)mx * j] + dt * mu[i + (size_t)nx * j] * ex[i + (size_t)mx * j];
    PAR_FOR
    for (int j = 0; j < nz - 1; ++j)
        for (int i = 0; i < nx; ++i)
            szy[i + (size_t)nx * j] = szy[i + (size_t)nx * j] + dt * mu[i + (size_t)nx * j] * ez[i + (size_t)nx * j];
}

/* v = v + dt*DS./rho ; v = v.*absbound (run_forward1_green.m:160-164) */
static void velocity_update(const grid_t* g, const double* DS, const double* rho, const double* ab, double* v)
{
    const int64_t P = (int64_t)g->nx * g->nz;
    const double dt = g->dt;
    PAR_FOR
    for (int64_t p = 0; p < P; ++p) {
        double t = v[p] + dt * DS[p] / rho[p];
        v[p] = t * ab[p];
    }
}

/* distribution(x,:) = interpol_matrix * source_distribution(x,:) (run_forward2_correlation.m:168-173) */
static double* make_distribution(const grid_t* g, const double* I, int nb_cols, const double* sd)
{
    const int64_t P = (int64_t)g->nx * g->nz;
    double* dist = (double*)calloc((size_t)P * (size_t)g->K, sizeof(double));
    PAR_FOR
    for (int64_t p = 0; p < P; ++p)
        for (int k = 0; k < g->K; ++k) {
            double a = 0.0;
            for (int b = 0; b < nb_cols; ++b) a += I[k + (size_t)g->K * b] * sd[p + (size_t)P * b];
            dist[p + (size_t)P * k] = a;
        }
    return dist;
}

/* =============================================================================================
 * run_forward1_green (code/run_forward1_green.m)
 * =========================================================================================== */
int oracle_run_forward1_green(const fd2d_config* cfg,
                              const fd2d_carray* mu_a, const fd2d_carray* rho_a,
                              const fd2d_carray* src_a, const fd2d_carray* rec_a, double mode,
                              const fd2d_carray* dmu_a, const fd2d_carray* Gin_a,
                              fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seis)
{
    grid_t g;
    int st = setup(cfg, mu_a, rho_a, &g);
    if (st) return st;
    const int nx = g.nx, nz = g.nz, nt = cfg->nt, K = g.K, mx = nx - 1;
    const int64_t P = (int64_t)nx * nz;
    const double* mu = (const double*)mu_a->data;
    const double* rho = (const double*)rho_a->data;
    const int born = numel(dmu_a) != 0;
    const double* dmu = born ? (const double*)dmu_a->data : NULL;
    const float* G_in = numel(Gin_a) ? (const float*)Gin_a->data : NULL;
    if (born && numel(dmu_a) != P) return fail(FD2D_EINVAL, "dmu must have nx*nz elements");
    if (born && (!G_in || dim(Gin_a, 2) < nt || Gin_a->dims[0] * Gin_a->dims[1] != P))
        return fail(FD2D_EINVAL, "Born mode needs G_in with nt slices of nx*nz");

    int *sid = NULL, *rid = NULL;
    int64_t ns = 0, nr = 0;
    if ((st = node_ids(src_a, nx, nz, g.dx, g.dz, &sid, &ns))) return st;
    if ((st = node_ids(rec_a, nx, nz, g.dx, g.dz, &rid, &nr))) { free(sid); return st; }

    /* run_forward1_green.m:64 */
    const double stf = 1.0e9;

    /* n_fw of :98-103 */
    int64_t n_fw = 0;
    for (int n = nt; n <= 2 * nt - 1; ++n)
        if (n % g.f == 0) ++n_fw;
    float* Gout = (mode != 0 && G_out && G_out->data) ? (float*)G_out->data : NULL;
    double* Gf = (G_fft && G_fft->data) ? (double*)G_fft->data : NULL;
    double* sg = (seis && seis->data) ? (double*)seis->data : NULL;
    if (Gf) memset(Gf, 0, sizeof(double) * 2 * (size_t)P * (size_t)K);
    if (sg) memset(sg, 0, sizeof(double) * (size_t)nr * (size_t)nt);

    double* v = (double*)calloc((size_t)P, sizeof(double));
    double* sxy = (double*)calloc((size_t)mx * nz, sizeof(double));
    double* szy = (double*)calloc((size_t)nx * (nz - 1), sizeof(double));
    double* DS = (double*)malloc(sizeof(double) * (size_t)P);
    double* ex = (double*)malloc(sizeof(double) * (size_t)mx * nz);
    double* ez = (double*)malloc(sizeof(double) * (size_t)nx * (nz - 1));
    double* ab = (double*)malloc(sizeof(double) * (size_t)P);
    double* cre = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    double* cim = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    oracle_absbound(cfg, ab);

    int64_t i_fw_out = 0, i_fw_in = 0;
    const double inv_sqrt_2pi = 1.0 / sqrt(2.0 * M_PI);
    for (int n = 1; n <= nt; ++n) {
        /* :133-137 snapshot of v before the update */
        if (mode != 0 && (n + nt - 1) % g.f == 0) {
            if (Gout) {
                float* o = Gout + (size_t)P * i_fw_out;
                for (int64_t p = 0; p < P; ++p) o[p] = (float)v[p];
            }
            ++i_fw_out;
        }
        /* :140-144 Born source from the stored Green field (single-precision stencil) */
        if (born) {
            const float* Gi = G_in + (size_t)P * i_fw_in;
            dx_v_single(Gi, g.dx, nx, nz, g.order, ex);
            dz_v_single(Gi, g.dz, nx, nz, g.order, ez);
            for (int j = 0; j < nz; ++j)
                for (int i = 0; i < mx; ++i)
                    sxy[i + (size_t)mx * j] = sxy[i + (size_t)mx * j] + g.dt * dmu[i + (size_t)nx * j] * ex[i + (size_t)mx * j];
            for (int j = 0; j < nz - 1; ++j)
                for (int i = 0; i < nx; ++i)
                    szy[i + (size_t)nx * j] = szy[i + (size_t)nx * j] + g.dt * dmu[i + (size_t)nx * j] * ez[i + (size_t)nx * j];
            ++i_fw_in;
        }
        /* :148 */
        oracle_div_s(sxy, szy, g.dx, g.dz, nx, nz, g.order, DS);
        /* :152-156 point force, suppressed in Born mode */
        if (!born)
            for (int64_t i = 0; i < ns; ++i) {
                const size_t p = (size_t)sid[2 * i] + (size_t)nx * sid[2 * i + 1];
                DS[p] = DS[p] + stf;
            }
        /* :160-164 */
        velocity_update(&g, DS, rho, ab, v);
        /* :168-172 */
        oracle_dx_v(v, g.dx, nx, nz, g.order, ex);
        oracle_dz_v(v, g.dz, nx, nz, g.order, ez);
        stress_update(&g, mu, ex, ez, sxy, szy);
        /* :176-183 on-the-fly DFT; fft_coeff of :81-90 with t(n-nt+1) = (n-1)dt at g = n+nt-1 */
        if ((n + nt - 1) % g.s == 0 && Gf) {
            const double t = (n - 1) * g.dt;
            for (int k = 0; k < K; ++k) {
                const double ph = -g.w[k] * t;
                cre[k] = inv_sqrt_2pi * cos(ph) * g.dt;
                cim[k] = inv_sqrt_2pi * sin(ph) * g.dt;
            }
            PAR_FOR
            for (int64_t p = 0; p < P; ++p) {
                const double vp = v[p];
                for (int k = 0; k < K; ++k) {
                    double* a = Gf + 2 * ((size_t)p + (size_t)P * k);
                    a[0] = a[0] + vp * cre[k];
                    a[1] = a[1] + vp * cim[k];
                }
            }
        }
        /* :187-189 */
        if (sg)
            for (int64_t ir = 0; ir < nr; ++ir)
                sg[ir + (size_t)nr * (n - 1)] = v[(size_t)rid[2 * ir] + (size_t)nx * rid[2 * ir + 1]];
    }
    /* :196-198 */
    if (Gf)
        for (int64_t q = 0; q < P * K; ++q) Gf[2 * q + 1] = -Gf[2 * q + 1];
    (void)n_fw;

    free(v); free(sxy); free(szy); free(DS); free(ex); free(ez); free(ab); free(cre); free(cim);
    free(sid); free(rid);
    return FD2D_OK;
}

/* =============================================================================================
 * run_forward2_correlation (code/run_forward2_correlation.m)
 * =========================================================================================== */
int oracle_run_forward2_correlation(const fd2d_config* cfg,
                                    const fd2d_carray* mu_a, const fd2d_carray* rho_a,
                                    const fd2d_carray* G_a, const fd2d_carray* spec_a,
                                    const fd2d_carray* sd_a, const fd2d_carray* rec_a,
                                    double mode, const fd2d_carray* dmu_a, const fd2d_carray* Cin_a,
                                    fd2d_marray* seis, fd2d_marray* C_out)
{
    grid_t g;
    int st = setup(cfg, mu_a, rho_a, &g);
    if (st) return st;
    const int nx = g.nx, nz = g.nz, K = g.K, mx = nx - 1, nb = cfg->n_basis_fct;
    const int nt0 = cfg->nt;
    const int N = 2 * nt0 - 1; /* :68-69 */
    const int64_t P = (int64_t)nx * nz;
    const double* mu = (const double*)mu_a->data;
    const double* rho = (const double*)rho_a->data;
    const double* G = (const double*)G_a->data;
    if (numel(G_a) != P * K) return fail(FD2D_EINVAL, "G_fft must be nx x nz x n_sample");
    const int64_t n_src = dim(spec_a, 1);
    if (dim(spec_a, 0) != K || n_src < 1) return fail(FD2D_EINVAL, "spectrum must be n_sample x n_noise_sources");
    const double* spec = (const double*)spec_a->data;
    const int nb_cols = nb == 0 ? 1 : nb;
    if (nb == 0 && n_src != 1)
        return fail(FD2D_EINVAL, "n_basis_fct == 0 needs one noise source (run_forward2_correlation.m:171)");
    if (numel(sd_a) != P * (nb == 0 ? n_src : nb)) return fail(FD2D_EINVAL, "source_dist size");
    const int born = numel(dmu_a) != 0;
    const double* dmu = born ? (const double*)dmu_a->data : NULL;
    const float* C_in = numel(Cin_a) ? (const float*)Cin_a->data : NULL;
    if (born && (numel(dmu_a) != P || !C_in || dim(Cin_a, 2) < N))
        return fail(FD2D_EINVAL, "Born mode needs dmu (nx*nz) and C_in with 2nt-1 slices");

    int* rid = NULL;
    int64_t nr = 0;
    if ((st = node_ids(rec_a, nx, nz, g.dx, g.dz, &rid, &nr))) return st;

    double* I = (double*)malloc(sizeof(double) * (size_t)K * nb_cols);
    if ((st = oracle_interpol_matrix(K, nb, I))) { free(rid); free(I); return st; }
    double* dist = make_distribution(&g, I, nb_cols, (const double*)sd_a->data);
    /* spectrum summed over noise sources: sum_ns real(S(:,:,ns)) = real(sum_ns S) (:201-211) */
    double* spec_sum = (double*)calloc((size_t)K, sizeof(double));
    for (int k = 0; k < K; ++k)
        for (int64_t ns = 0; ns < n_src; ++ns) spec_sum[k] += spec[k + (size_t)K * ns];

    float* Cout = (mode != 0 && C_out && C_out->data) ? (float*)C_out->data : NULL;
    double* sg = (seis && seis->data) ? (double*)seis->data : NULL;
    if (sg) memset(sg, 0, sizeof(double) * (size_t)nr * (size_t)N);

    double* v = (double*)calloc((size_t)P, sizeof(double));
    double* u = (double*)calloc((size_t)P, sizeof(double));
    double* sxy = (double*)calloc((size_t)mx * nz, sizeof(double));
    double* szy = (double*)calloc((size_t)nx * (nz - 1), sizeof(double));
    double* DS = (double*)malloc(sizeof(double) * (size_t)P);
    double* ex = (double*)malloc(sizeof(double) * (size_t)mx * nz);
    double* ez = (double*)malloc(sizeof(double) * (size_t)nx * (nz - 1));
    double* bx = (double*)malloc(sizeof(double) * (size_t)mx * nz);
    double* bz = (double*)malloc(sizeof(double) * (size_t)nx * (nz - 1));
    float* dC = born ? (float*)malloc(sizeof(float) * (size_t)P) : NULL;
    double* ab = (double*)malloc(sizeof(double) * (size_t)P);
    double* cre = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    double* cim = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    oracle_absbound(cfg, ab);
    const double inv_sqrt_2pi = 1.0 / sqrt(2.0 * M_PI);

    int64_t i_fw_out = 0;
    for (int n = 1; n <= N; ++n) {
        /* :182-186 snapshot of u before the update */
        if (mode != 0 && n % g.f == 0) {
            if (Cout) {
                float* o = Cout + (size_t)P * i_fw_out;
                for (int64_t p = 0; p < P; ++p) o[p] = (float)u[p];
            }
            ++i_fw_out;
        }
        /* :190 */
        oracle_div_s(sxy, szy, g.dx, g.dz, nx, nz, g.order, DS);
        /* :195-216 correlation source; ifft_coeff of :73-85 with t(n) = (n-nt)dt */
        if (n % g.s == 0) {
            const double t = (double)(n - nt0) * g.dt;
            for (int k = 0; k < K; ++k) {
                const double ph = g.w[k] * t;
                cre[k] = inv_sqrt_2pi * cos(ph) * g.dw;
                cim[k] = inv_sqrt_2pi * sin(ph) * g.dw;
            }
            PAR_FOR
            for (int64_t p = 0; p < P; ++p) {
                double Sre = 0.0;
                for (int k = 0; k < K; ++k) {
                    const double a = spec_sum[k] * dist[p + (size_t)P * k];
                    const double* Gk = G + 2 * ((size_t)p + (size_t)P * k);
                    const double zr = a * Gk[0], zi = a * Gk[1];
                    Sre += zr * cre[k] - zi * cim[k];
                }
                DS[p] = DS[p] + Sre;
            }
        }
        /* :220-224 */
        velocity_update(&g, DS, rho, ab, v);
        /* :228-238 */
        oracle_dx_v(v, g.dx, nx, nz, g.order, ex);
        oracle_dz_v(v, g.dz, nx, nz, g.order, ez);
        stress_update(&g, mu, ex, ez, sxy, szy);
        if (born && n != N) {
            const float* c0 = C_in + (size_t)P * (n - 1);
            const float* c1 = C_in + (size_t)P * n;
            for (int64_t p = 0; p < P; ++p) dC[p] = c0[p] - c1[p];
            dx_v_single(dC, g.dx, nx, nz, g.order, bx);
            dz_v_single(dC, g.dz, nx, nz, g.order, bz);
            for (int j = 0; j < nz; ++j)
                for (int i = 0; i < mx; ++i)
                    sxy[i + (size_t)mx * j] -= dmu[i + (size_t)nx * j] * bx[i + (size_t)mx * j];
            for (int j = 0; j < nz - 1; ++j)
                for (int i = 0; i < nx; ++i)
                    szy[i + (size_t)nx * j] -= dmu[i + (size_t)nx * j] * bz[i + (size_t)nx * j];
        }
        /* :242 */
        PAR_FOR
        for (int64_t p = 0; p < P; ++p) u[p] = u[p] + v[p] * g.dt;
        /* :246-248 */
        if (sg)
            for (int64_t ir = 0; ir < nr; ++ir)
                sg[ir + (size_t)nr * (n - 1)] = u[(size_t)rid[2 * ir] + (size_t)nx * rid[2 * ir + 1]];
    }

    free(v); free(u); free(sxy); free(szy); free(DS); free(ex); free(ez); free(bx); free(bz);
    free(dC); free(ab); free(cre); free(cim); free(rid); free(I); free(dist); free(spec_sum);
    return FD2D_OK;
}

/* =============================================================================================
 * run_noise_adjoint (code/run_noise_adjoint.m)
 * =========================================================================================== */
int oracle_run_noise_adjoint(const fd2d_config* cfg,
                             const fd2d_carray* mu_a, const fd2d_carray* rho_a,
                             const fd2d_carray* ufwd_a, const fd2d_carray* stf_a,
                             const fd2d_carray* src_a, const fd2d_carray* rec_a,
                             const fd2d_carray* spec_a, const fd2d_carray* sd_a,
                             const fd2d_carray* G_a, double mode,
                             const fd2d_carray* dmu_a, const fd2d_carray* uin_a,
                             fd2d_marray* K_out, fd2d_marray* stf_fft, fd2d_marray* u_out,
                             fd2d_marray* seis)
{
    grid_t g;
    int st = setup(cfg, mu_a, rho_a, &g);
    if (st) return st;
    const int nx = g.nx, nz = g.nz, K = g.K, mx = nx - 1, nb = cfg->n_basis_fct;
    const int n_zero = cfg->nt;
    const int N = 2 * n_zero - 1; /* :47-49 */
    const int64_t P = (int64_t)nx * nz;
    const double* mu = (const double*)mu_a->data;
    const double* rho = (const double*)rho_a->data;
    const int imode = (int)mode;
    if ((double)imode != mode || !(imode == 0 || imode == 1 || (imode >= 10 && imode <= 13)))
        return fail(FD2D_EINVAL, "mode must be 0, 1, 10, 11, 12 or 13");
    const int nb_cols = nb == 0 ? 1 : nb;

    /* adjoint source type (:249, :255): 2-D stf -> point sources, 3-D -> distributed */
    const int64_t n_stf = numel(stf_a);
    const int stf_point = n_stf != 0 && stf_a->dims[2] == 1;
    const int stf_dist = n_stf != 0 && stf_a->dims[2] != 1;
    const double* stf = n_stf ? (const double*)stf_a->data : NULL;
    const int fft_mode = imode == 1 || imode == 12 || imode == 13;
    if (stf_dist && fft_mode)
        return fail(FD2D_EUNSUPPORTED, "distributed stf with mode %d advances the DFT counter twice (run_noise_adjoint.m:265,343)", imode);
    if (stf_dist && (stf_a->dims[0] * stf_a->dims[1] != P || stf_a->dims[2] != K))
        return fail(FD2D_EINVAL, "distributed stf must be nx x nz x n_sample");

    /* :165-167 half-length run for gradient modes with a distributed source */
    const int nsteps = ((imode == 0 || imode == 1) && stf_dist) ? n_zero : N;

    int *sid = NULL, *rid = NULL;
    int64_t ns_adj = 0, nr = 0;
    if ((st = node_ids(src_a, nx, nz, g.dx, g.dz, &sid, &ns_adj))) return st;
    if ((st = node_ids(rec_a, nx, nz, g.dx, g.dz, &rid, &nr))) { free(sid); return st; }
    if (stf_point && (stf_a->dims[0] < ns_adj || stf_a->dims[1] < nsteps)) {
        free(sid); free(rid);
        return fail(FD2D_EINVAL, "point stf must be n_src x (2nt-1)");
    }

    /* :111-123 source distribution */
    const int have_sd = numel(sd_a) != 0;
    const int64_t n_src = have_sd ? dim(spec_a, 1) : 0;
    const double* spec = numel(spec_a) ? (const double*)spec_a->data : NULL;
    double* I = (double*)malloc(sizeof(double) * (size_t)K * nb_cols);
    if ((st = oracle_interpol_matrix(K, nb, I))) { free(sid); free(rid); free(I); return st; }
    double* dist = NULL;
    double* spec_sum = (double*)calloc((size_t)(K > 0 ? K : 1), sizeof(double));
    if (have_sd) {
        if (dim(spec_a, 0) != K) { free(sid); free(rid); free(I); free(spec_sum); return fail(FD2D_EINVAL, "spectrum must have n_sample rows"); }
        if (nb == 0 && n_src != 1) { free(sid); free(rid); free(I); free(spec_sum); return fail(FD2D_EINVAL, "n_basis_fct == 0 needs one noise source"); }
        if (numel(sd_a) != P * (nb == 0 ? n_src : nb)) { free(sid); free(rid); free(I); free(spec_sum); return fail(FD2D_EINVAL, "source_dist size"); }
        dist = make_distribution(&g, I, nb_cols, (const double*)sd_a->data);
        for (int k = 0; k < K; ++k)
            for (int64_t q = 0; q < n_src; ++q) spec_sum[k] += spec[k + (size_t)K * q];
    }

    /* :302 source kernel gate */
    const int src_kernel = numel(G_a) != 0 && dim(G_a, 2) == K && G_a->dims[0] * G_a->dims[1] == P;
    const double* G = src_kernel ? (const double*)G_a->data : NULL;
    if (src_kernel && (!spec || dim(spec_a, 0) != K || dim(spec_a, 1) != 1)) {
        free(sid); free(rid); free(I); free(spec_sum); free(dist);
        return fail(FD2D_EINVAL, "the source kernel needs an n_sample x 1 spectrum (run_noise_adjoint.m:312)");
    }
    /* :324 structure kernel gate */
    const int64_t n_ufwd = numel(ufwd_a) ? dim(ufwd_a, 2) : 0;
    const float* ufwd = n_ufwd ? (const float*)ufwd_a->data : NULL;
    if (n_ufwd && ufwd_a->dims[0] * ufwd_a->dims[1] != P) {
        free(sid); free(rid); free(I); free(spec_sum); free(dist);
        return fail(FD2D_EINVAL, "u_fwd slices must have nx*nz elements");
    }
    /* :285 Born gate */
    const int64_t n_uin = numel(uin_a) ? dim(uin_a, 2) : 0;
    const int born = numel(dmu_a) != 0 && n_uin == N;
    const double* dmu = born ? (const double*)dmu_a->data : NULL;
    const float* u_in = born ? (const float*)uin_a->data : NULL;

    double* Kd = (K_out && K_out->data) ? (double*)K_out->data : NULL;
    double* sfft = (fft_mode && stf_fft && stf_fft->data) ? (double*)stf_fft->data : NULL;
    float* uo = ((imode == 11 || imode == 13) && u_out && u_out->data) ? (float*)u_out->data : NULL;
    double* sg = (seis && seis->data) ? (double*)seis->data : NULL;
    if (sg) memset(sg, 0, sizeof(double) * (size_t)nr * (size_t)N);
    if (sfft) memset(sfft, 0, sizeof(double) * 2 * (size_t)P * (size_t)K);

    double* v = (double*)calloc((size_t)P, sizeof(double));
    double* u = (double*)calloc((size_t)P, sizeof(double));
    double* sxy = (double*)calloc((size_t)mx * nz, sizeof(double));
    double* szy = (double*)calloc((size_t)nx * (nz - 1), sizeof(double));
    double* DS = (double*)malloc(sizeof(double) * (size_t)P);
    double* ex = (double*)malloc(sizeof(double) * (size_t)mx * nz);
    double* ez = (double*)malloc(sizeof(double) * (size_t)nx * (nz - 1));
    double* bx = (double*)malloc(sizeof(double) * (size_t)mx * nz);
    double* bz = (double*)malloc(sizeof(double) * (size_t)nx * (nz - 1));
    float* dU = born ? (float*)malloc(sizeof(float) * (size_t)P) : NULL;
    double* ab = (double*)malloc(sizeof(double) * (size_t)P);
    double* Kmu = (double*)calloc((size_t)P, sizeof(double));
    double* Ks = (double*)calloc((size_t)P * nb_cols, sizeof(double));
    double* cre = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    double* cim = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    oracle_absbound(cfg, ab);
    const double inv_sqrt_2pi = 1.0 / sqrt(2.0 * M_PI);

    int64_t i_ftc = 0, i_fw_in = 0, i_fw_out = 0;
    for (int n = 1; n <= nsteps; ++n) {
        const double t = (double)(n - n_zero) * g.dt;
        /* :237-241 */
        if ((imode == 11 || imode == 13) && n % g.f == 0) {
            if (uo) {
                float* o = uo + (size_t)P * i_fw_out;
                for (int64_t p = 0; p < P; ++p) o[p] = (float)u[p];
            }
            ++i_fw_out;
        }
        /* :245 */
        oracle_div_s(sxy, szy, g.dx, g.dz, nx, nz, g.order, DS);
        /* :249-268 */
        if (stf_point) {
            const int64_t d0 = stf_a->dims[0];
            for (int64_t i = 0; i < ns_adj; ++i) {
                const size_t p = (size_t)sid[2 * i] + (size_t)nx * sid[2 * i + 1];
                DS[p] = DS[p] + stf[2 * ((size_t)i + (size_t)d0 * (n - 1))];
            }
        } else if (stf_dist && n % g.s == 0) {
            for (int k = 0; k < K; ++k) {
                const double ph = g.w[k] * t;
                cre[k] = inv_sqrt_2pi * cos(ph) * g.dw;
                cim[k] = inv_sqrt_2pi * sin(ph) * g.dw;
            }
            ++i_ftc;
            if (have_sd) {
                PAR_FOR
                for (int64_t p = 0; p < P; ++p) {
                    double Tre = 0.0;
                    for (int k = 0; k < K; ++k) {
                        const double a = spec_sum[k] * dist[p + (size_t)P * k];
                        const double* sk = stf + 2 * ((size_t)p + (size_t)P * k);
                        const double zr = a * sk[0], zi = -(a * sk[1]);
                        Tre += zr * cre[k] - zi * cim[k];
                    }
                    DS[p] = DS[p] + Tre;
                }
            }
        }
        /* :273-277 */
        velocity_update(&g, DS, rho, ab, v);
        /* :281-291 */
        oracle_dx_v(v, g.dx, nx, nz, g.order, ex);
        oracle_dz_v(v, g.dz, nx, nz, g.order, ez);
        stress_update(&g, mu, ex, ez, sxy, szy);
        if (born && n != N) {
            const float* c0 = u_in + (size_t)P * (n - 1);
            const float* c1 = u_in + (size_t)P * n;
            for (int64_t p = 0; p < P; ++p) dU[p] = c0[p] - c1[p];
            dx_v_single(dU, g.dx, nx, nz, g.order, bx);
            dz_v_single(dU, g.dz, nx, nz, g.order, bz);
            for (int j = 0; j < nz; ++j)
                for (int i = 0; i < mx; ++i)
                    sxy[i + (size_t)mx * j] -= dmu[i + (size_t)nx * j] * bx[i + (size_t)mx * j];
            for (int j = 0; j < nz - 1; ++j)
                for (int i = 0; i < nx; ++i)
                    szy[i + (size_t)nx * j] -= dmu[i + (size_t)nx * j] * bz[i + (size_t)nx * j];
        }
        /* :295-297 */
        PAR_FOR
        for (int64_t p = 0; p < P; ++p) u[p] = u[p] + v[p] * g.dt;
        oracle_dx_v(u, g.dx, nx, nz, g.order, ex);
        oracle_dz_v(u, g.dz, nx, nz, g.order, ez);
        /* :302-320 source kernel with ifft_coeff2(n,:) = (2pi)^-1/2 exp(i w t(n)) dw */
        if (src_kernel && n % g.s == 0 && Kd) {
            for (int k = 0; k < K; ++k) {
                const double ph = g.w[k] * t;
                cre[k] = inv_sqrt_2pi * cos(ph) * g.dw;
                cim[k] = inv_sqrt_2pi * sin(ph) * g.dw;
            }
            PAR_FOR
            for (int64_t p = 0; p < P; ++p) {
                for (int b = 0; b < nb_cols; ++b) {
                    double Mre = 0.0;
                    for (int k = 0; k < K; ++k) {
                        const double* Gk = G + 2 * ((size_t)p + (size_t)P * k);
                        const double ar = spec[k] * Gk[0], ai = -(spec[k] * Gk[1]);
                        Mre += I[k + (size_t)K * b] * (ar * cre[k] - ai * cim[k]);
                    }
                    Ks[p + (size_t)P * b] = Ks[p + (size_t)P * b] + Mre * u[p];
                }
            }
        }
        /* :324-331 structure kernel, forward snapshots in reverse order, times fw_nth */
        if (n_ufwd && n_ufwd >= i_fw_in + 1 && n % g.f == 0) {
            const float* sn = ufwd + (size_t)P * (size_t)(n_ufwd - 1 - i_fw_in);
            dx_v_single(sn, g.dx, nx, nz, g.order, bx);
            dz_v_single(sn, g.dz, nx, nz, g.order, bz);
            for (int j = 0; j < nz; ++j)
                for (int i = 0; i < mx; ++i)
                    Kmu[i + (size_t)nx * j] = Kmu[i + (size_t)nx * j] - ex[i + (size_t)mx * j] * bx[i + (size_t)mx * j] * g.f;
            for (int j = 0; j < nz - 1; ++j)
                for (int i = 0; i < nx; ++i)
                    Kmu[i + (size_t)nx * j] = Kmu[i + (size_t)nx * j] - ez[i + (size_t)nx * j] * bz[i + (size_t)nx * j] * g.f;
            ++i_fw_in;
        }
        /* :335-346 DFT of the adjoint state with fft_coeff(i_ftc) at t(n) */
        if (fft_mode && n % g.s == 0) {
            if (sfft) {
                for (int k = 0; k < K; ++k) {
                    const double ph = -g.w[k] * t;
                    cre[k] = inv_sqrt_2pi * cos(ph) * g.dt;
                    cim[k] = inv_sqrt_2pi * sin(ph) * g.dt;
                }
                PAR_FOR
                for (int64_t p = 0; p < P; ++p) {
                    const double up = u[p];
                    for (int k = 0; k < K; ++k) {
                        double* a = sfft + 2 * ((size_t)p + (size_t)P * k);
                        a[0] = a[0] + up * cre[k];
                        a[1] = a[1] + up * cim[k];
                    }
                }
            }
            ++i_ftc;
        }
        /* :350-352 */
        if (sg)
            for (int64_t ir = 0; ir < nr; ++ir)
                sg[ir + (size_t)nr * (n - 1)] = u[(size_t)rid[2 * ir] + (size_t)nx * rid[2 * ir + 1]];
    }
    (void)i_ftc;
    /* :359-367 K = cat(3, K_s, K_mu) */
    if (Kd) {
        memcpy(Kd, Ks, sizeof(double) * (size_t)P * nb_cols);
        memcpy(Kd + (size_t)P * nb_cols, Kmu, sizeof(double) * (size_t)P);
    }

    free(v); free(u); free(sxy); free(szy); free(DS); free(ex); free(ez); free(bx); free(bz);
    free(dU); free(ab); free(Kmu); free(Ks); free(cre); free(cim); free(sid); free(rid);
    free(I); free(dist); free(spec_sum);
    return FD2D_OK;
}

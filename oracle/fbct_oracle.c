/*
 * fbct_oracle.c -- FP64 CPU oracle (test infrastructure only; see fbct_oracle.h).
 *
 * A plain restatement of the reference discretisations in double precision.
 * Threads: a small pthread parallel-for over independent work items; every
 * reduction is done in a fixed order so results are deterministic.
 */
#define _GNU_SOURCE
#include "fbct_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* The reference's PI macro (FP/transform.h:1).  Used in float for the
 * parameter decode, exactly as the reference does, and as a double literal
 * for the ramp taps. */
#define REF_PI_F 3.14159265359f
#define REF_PI_D 3.14159265359

/* ------------------------------------------------------------------ threads */

typedef void (*orc_item_fn)(void *ctx, long long item);

typedef struct {
    orc_item_fn fn;
    void *ctx;
    long long n;
    atomic_llong next;
} orc_pool;

static void *orc_worker(void *arg)
{
    orc_pool *p = (orc_pool *)arg;
    for (;;) {
        long long it = atomic_fetch_add(&p->next, 1);
        if (it >= p->n) break;
        p->fn(p->ctx, it);
    }
    return NULL;
}

int orc_default_threads(void)
{
    long hw = sysconf(_SC_NPROCESSORS_ONLN);
    if (hw < 1) hw = 1;
    const char *e = getenv("FBCT_ORACLE_THREADS");
    if (!e) e = getenv("OMP_NUM_THREADS");
    if (e) {
        long v = strtol(e, NULL, 10);
        if (v >= 1 && v < hw) hw = v;
    }
    if (hw > 64) hw = 64;
    return (int)hw;
}

static void orc_parallel_for(long long n, int nthreads, orc_item_fn fn, void *ctx)
{
    if (n <= 0) return;
    if (nthreads <= 0) nthreads = orc_default_threads();
    if (nthreads > n) nthreads = (int)n;
    orc_pool p;
    p.fn = fn;
    p.ctx = ctx;
    p.n = n;
    atomic_init(&p.next, 0);
    if (nthreads <= 1) {
        orc_worker(&p);
        return;
    }
    pthread_t th[64];
    int started = 0;
    for (int t = 0; t < nthreads - 1 && t < 64; ++t) {
        if (pthread_create(&th[t], NULL, orc_worker, &p) == 0) started++;
        else break;
    }
    orc_worker(&p);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------ decode */

int orc_decode_params(const float *p, orc_geom *g)
{
    /* FP/transform.cc:59-76 and FDK/FBP.cc:61-84: float arithmetic throughout. */
    memset(g, 0, sizeof(*g));
    g->D = p[0];
    g->R = p[1];
    g->nViews = (int)p[2];
    g->dRange = p[3];
    g->nCols = (int)p[4];
    g->nX = (int)p[5];
    g->nY = (int)p[6];
    g->rFOV = p[7];
    g->dShift = p[9];
    g->nSamp = (int)p[11];
    g->detBottom = p[13];
    g->dRow = p[14];
    g->dz = p[15];
    if (g->nViews <= 0 || g->nCols <= 0 || g->nX <= 0 || g->nY <= 0) return -1;
    volatile float t;
    t = g->dRange / (float)g->nViews;
    t = t * REF_PI_F;
    g->dAngle = t / 180.0f;
    t = 2.0f * g->rFOV;
    g->dx = t / (float)g->nX;
    t = 2.0f * g->rFOV;
    g->dy = t / (float)g->nY;
    t = -0.5f * (float)g->nX;
    g->xmin = t * g->dx;
    t = -0.5f * (float)g->nY;
    g->ymin = t * g->dy;
    t = p[8] * REF_PI_F;
    g->dCol = t / 180.0f;
    t = p[10] * REF_PI_F;
    g->dLeft = t / 180.0f;
    t = p[12] * REF_PI_F;
    g->startAngle = t / 180.0f;
    return 0;
}

/* ------------------------------------------------------------------ FP ray */

void orc_ray_setup(const orc_geom *g, double beta, int c,
                   int *N, double *u0, double *w0, double *du, double *dw)
{
    /* FP/transform.cu:64-101, in pixel units of dx.  The reference launches it
     * with fSourceToCenterDistance = R/dx and fFOVRadius = rFOV/dx (:239). */
    const double Rp = (double)g->R / (double)g->dx;
    const double r = (double)g->rFOV / (double)g->dx;
    const double gam = (double)g->dLeft + ((double)c + 0.5) * (double)g->dCol;
    const double th = beta + gam;
    const double d = Rp * sin(gam);
    const double disc = r * r - d * d;
    *N = 0;
    *u0 = *w0 = *du = *dw = 0.0;
    /* A ray that misses the FOV circle yields NaN geometry in the reference and
     * int(NaN) == 0 samples; a chord shorter than one step yields 0 samples. */
    if (!(disc > 0.0)) return;
    const double a = sqrt(disc);
    const double nf = floor(16.0 * a); /* floor(L / 0.125) with L = 2a */
    if (nf < 1.0) return;
    const int n = (int)nf;
    const double ct = cos(th), st = sin(th);
    /* intersection1 = Rot_th(-a, -d); intersection2 = Rot_th(+a, -d) */
    const double e1x = -a * ct + d * st;
    const double e1y = -a * st - d * ct;
    *N = n;
    *du = (2.0 * a * ct) / (double)n;
    *dw = (2.0 * a * st) / (double)n;
    /* start = intersection1 + r (FP/transform.cu:97-98); the texture samples
     * texel centres at +0.5, so the interpolation coordinate is start - 0.5. */
    *u0 = e1x + r - 0.5;
    *w0 = e1y + r - 0.5;
}

long long orc_fp_sample_count(const orc_geom *g, const float *angles)
{
    long long s = 0;
    for (int v = 0; v < g->nViews; ++v)
        for (int c = 0; c < g->nCols; ++c) {
            int N;
            double a, b, e, f;
            orc_ray_setup(g, (double)angles[v], c, &N, &a, &b, &e, &f);
            s += N;
        }
    return s;
}

/* Bilinear fetch with clamp-to-edge on a [nY][nX] image; (u, w) are already in
 * texel-centre coordinates (u = x - 0.5). */
static inline double bilerp(const double *img, int nX, int nY, double u, double w)
{
    const double fu = floor(u), fw = floor(w);
    const int i0 = (int)fu, j0 = (int)fw;
    const double a = u - fu, b = w - fw;
    const int ia = clampi(i0, 0, nX - 1), ib = clampi(i0 + 1, 0, nX - 1);
    const int ja = clampi(j0, 0, nY - 1), jb = clampi(j0 + 1, 0, nY - 1);
    const double top = (1.0 - a) * img[(long long)ja * nX + ia] + a * img[(long long)ja * nX + ib];
    const double bot = (1.0 - a) * img[(long long)jb * nX + ia] + a * img[(long long)jb * nX + ib];
    return (1.0 - b) * top + b * bot;
}

typedef struct {
    const orc_geom *g;
    const float *angles;
    const double *in;
    double *out;
    int nS;
    int weighted;
    double zmin;
    int nRows, nZ;
    const double *h;
    double **priv;
    int nthr;
} orc_ctx;

static void fp_item(void *vctx, long long item)
{
    const orc_ctx *x = (const orc_ctx *)vctx;
    const orc_geom *g = x->g;
    const int v = (int)(item % g->nViews);
    const long long s = item / g->nViews;
    const double *img = x->in + s * (long long)g->nX * g->nY;
    double *row = x->out + (s * g->nViews + v) * (long long)g->nCols;
    const double scale = 0.125 * (double)g->dx; /* stepSize * dx, FP/transform.cu:144 */
    for (int c = 0; c < g->nCols; ++c) {
        int N;
        double u0, w0, du, dw;
        orc_ray_setup(g, (double)x->angles[v], c, &N, &u0, &w0, &du, &dw);
        double acc = 0.0;
        for (int k = 0; k < N; ++k)
            acc += bilerp(img, g->nX, g->nY, u0 + (double)k * du, w0 + (double)k * dw);
        row[c] = acc * scale;
    }
}

void orc_fp(const orc_geom *g, const float *angles, const double *img, double *sino,
            int nS, int nthreads)
{
    orc_ctx x = {g, angles, img, sino, nS, 0, 0.0, 0, 0, NULL, NULL, 0};
    orc_parallel_for((long long)nS * g->nViews, nthreads, fp_item, &x);
}

/* ---------------------------------------------------------- FP adjoint */

typedef struct {
    const orc_geom *g;
    const float *angles;
    const double *sino;  /* one slice */
    double **priv;       /* per-thread images */
    int nchunks;
    atomic_int slot;
} fpa_ctx;

static void scatter_bilerp(double *img, int nX, int nY, double u, double w, double val)
{
    const double fu = floor(u), fw = floor(w);
    const int i0 = (int)fu, j0 = (int)fw;
    const double a = u - fu, b = w - fw;
    const int ia = clampi(i0, 0, nX - 1), ib = clampi(i0 + 1, 0, nX - 1);
    const int ja = clampi(j0, 0, nY - 1), jb = clampi(j0 + 1, 0, nY - 1);
    img[(long long)ja * nX + ia] += (1.0 - b) * (1.0 - a) * val;
    img[(long long)ja * nX + ib] += (1.0 - b) * a * val;
    img[(long long)jb * nX + ia] += b * (1.0 - a) * val;
    img[(long long)jb * nX + ib] += b * a * val;
}

static void fpa_item(void *vctx, long long chunk)
{
    fpa_ctx *x = (fpa_ctx *)vctx;
    const orc_geom *g = x->g;
    double *img = x->priv[chunk];
    const double scale = 0.125 * (double)g->dx;
    for (int v = (int)chunk; v < g->nViews; v += x->nchunks) {
        const double *row = x->sino + (long long)v * g->nCols;
        for (int c = 0; c < g->nCols; ++c) {
            int N;
            double u0, w0, du, dw;
            orc_ray_setup(g, (double)x->angles[v], c, &N, &u0, &w0, &du, &dw);
            const double val = row[c] * scale;
            if (val == 0.0) continue;
            for (int k = 0; k < N; ++k)
                scatter_bilerp(img, g->nX, g->nY, u0 + (double)k * du, w0 + (double)k * dw, val);
        }
    }
}

void orc_fp_adjoint(const orc_geom *g, const float *angles, const double *sino,
                    double *img, int nS, int nthreads)
{
    if (nthreads <= 0) nthreads = orc_default_threads();
    const long long npix = (long long)g->nX * g->nY;
    int nchunks = nthreads < g->nViews ? nthreads : g->nViews;
    if (nchunks < 1) nchunks = 1;
    double **priv = (double **)malloc(sizeof(double *) * nchunks);
    for (int t = 0; t < nchunks; ++t) priv[t] = (double *)malloc(sizeof(double) * npix);
    for (int s = 0; s < nS; ++s) {
        for (int t = 0; t < nchunks; ++t) memset(priv[t], 0, sizeof(double) * npix);
        fpa_ctx x;
        x.g = g;
        x.angles = angles;
        x.sino = sino + (long long)s * g->nViews * g->nCols;
        x.priv = priv;
        x.nchunks = nchunks;
        atomic_init(&x.slot, 0);
        orc_parallel_for(nchunks, nthreads, fpa_item, &x);
        double *out = img + (long long)s * npix;
        for (long long p = 0; p < npix; ++p) {
            double acc = 0.0;
            for (int t = 0; t < nchunks; ++t) acc += priv[t][p];
            out[p] = acc;
        }
    }
    for (int t = 0; t < nchunks; ++t) free(priv[t]);
    free(priv);
}

/* ------------------------------------------------------------------ BP */

/* Column coordinate of pixel (x, y) in view beta and its distance weight, following
 * BP/transform.cu:76-91 and FBP/transform.cu:76-85 (sin/cos passed negated, :110). */
static inline int bp_coord(const orc_geom *g, double cb, double sb, double x, double y,
                           double *xi, double *wgt)
{
    const double U = (double)g->R - x * cb - y * sb;   /* projection onto iso-ray */
    const double t = x * sb - y * cb;                  /* distance to centre ray */
    *wgt = 1.0 / (U * U + t * t);
    *xi = (atan(t / U) - (double)g->dLeft) / (double)g->dCol + (double)g->dShift;
    return (*xi > 0.0 && *xi < (double)g->nCols);
}

static void bp_item(void *vctx, long long item)
{
    const orc_ctx *x = (const orc_ctx *)vctx;
    const orc_geom *g = x->g;
    const int j = (int)(item % g->nY);
    const long long s = item / g->nY;
    const double *sino = x->in + s * (long long)g->nViews * g->nCols;
    double *row = x->out + (s * g->nY + j) * (long long)g->nX;
    const double r2 = (double)g->rFOV * (double)g->rFOV;
    const double y = (double)g->ymin + (double)g->dy * ((double)j + 0.5);
    for (int i = 0; i < g->nX; ++i) {
        const double xx = (double)g->xmin + (double)g->dx * ((double)i + 0.5);
        double acc = 0.0;
        if (xx * xx + y * y <= r2) {
            for (int v = 0; v < g->nViews; ++v) {
                const double beta = (double)x->angles[v];
                double xi, w;
                if (!bp_coord(g, cos(beta), sin(beta), xx, y, &xi, &w)) continue;
                const double ct = xi - 0.5; /* texel-centre convention of tex1D */
                const double fc = floor(ct);
                const int c0 = (int)fc;
                const double al = ct - fc;
                const double *P = sino + (long long)v * g->nCols;
                const double val = (1.0 - al) * P[clampi(c0, 0, g->nCols - 1)] +
                                   al * P[clampi(c0 + 1, 0, g->nCols - 1)];
                acc += x->weighted ? val * w : val;
            }
        }
        row[i] = acc;
    }
}

void orc_bp(const orc_geom *g, const float *angles, const double *sino, double *img,
            int nS, int weighted, int nthreads)
{
    orc_ctx x = {g, angles, sino, img, nS, weighted, 0.0, 0, 0, NULL, NULL, 0};
    orc_parallel_for((long long)nS * g->nY, nthreads, bp_item, &x);
}

static void bpa_item(void *vctx, long long item)
{
    const orc_ctx *x = (const orc_ctx *)vctx;
    const orc_geom *g = x->g;
    const int v = (int)(item % g->nViews);
    const long long s = item / g->nViews;
    const double *img = x->in + s * (long long)g->nX * g->nY;
    double *P = x->out + (s * g->nViews + v) * (long long)g->nCols;
    memset(P, 0, sizeof(double) * g->nCols);
    const double r2 = (double)g->rFOV * (double)g->rFOV;
    const double beta = (double)x->angles[v];
    const double cb = cos(beta), sb = sin(beta);
    for (int j = 0; j < g->nY; ++j) {
        const double y = (double)g->ymin + (double)g->dy * ((double)j + 0.5);
        for (int i = 0; i < g->nX; ++i) {
            const double xx = (double)g->xmin + (double)g->dx * ((double)i + 0.5);
            if (!(xx * xx + y * y <= r2)) continue;
            double xi, w;
            if (!bp_coord(g, cb, sb, xx, y, &xi, &w)) continue;
            const double val = img[(long long)j * g->nX + i] * (x->weighted ? w : 1.0);
            const double ct = xi - 0.5;
            const double fc = floor(ct);
            const int c0 = (int)fc;
            const double al = ct - fc;
            P[clampi(c0, 0, g->nCols - 1)] += (1.0 - al) * val;
            P[clampi(c0 + 1, 0, g->nCols - 1)] += al * val;
        }
    }
}

void orc_bp_adjoint(const orc_geom *g, const float *angles, const double *img,
                    double *sino, int nS, int weighted, int nthreads)
{
    orc_ctx x = {g, angles, img, sino, nS, weighted, 0.0, 0, 0, NULL, NULL, 0};
    orc_parallel_for((long long)nS * g->nViews, nthreads, bpa_item, &x);
}

/* ------------------------------------------------------------ ramp filter */

void orc_ramp_taps(const orc_geom *g, double *h)
{
    /* "Ramp kernel using the h'' form in Hsieh's book, Eq 3.51":
     * FBP/transform.cu:153-162, FDK/GPURecon_GPU.cu:190-200. */
    const int nC = g->nCols;
    const double dc = (double)g->dCol;
    const double kappa = (double)g->R * (double)g->dAngle * dc;
    for (int n = -(nC - 1); n <= nC - 1; ++n) {
        double v;
        if (n == 0) v = 0.5 / (4.0 * dc * dc);
        else if (n % 2 != 0) {
            const double s = sin((double)n * dc);
            v = -0.5 / (REF_PI_D * REF_PI_D * s * s);
        } else v = 0.0;
        h[n + nC - 1] = v * kappa;
    }
}

static void filt_item(void *vctx, long long row)
{
    const orc_ctx *x = (const orc_ctx *)vctx;
    const int nC = x->g->nCols;
    const double *p = x->in + row * (long long)nC;
    double *q = x->out + row * (long long)nC;
    const double *h = x->h;
    for (int c = 0; c < nC; ++c) {
        double acc = 0.0;
        for (int m = 0; m < nC; ++m) acc += p[m] * h[c - m + nC - 1];
        q[c] = acc;
    }
}

void orc_ramp_filter(const orc_geom *g, const double *p, double *q, long long nRowsTotal,
                     int nthreads)
{
    double *h = (double *)malloc(sizeof(double) * (2 * g->nCols - 1));
    orc_ramp_taps(g, h);
    orc_ctx x = {g, NULL, p, q, 0, 0, 0.0, 0, 0, h, NULL, 0};
    orc_parallel_for(nRowsTotal, nthreads, filt_item, &x);
    free(h);
}

void orc_fbp(const orc_geom *g, const float *angles, const double *sino, double *img,
             int nS, int nthreads)
{
    const long long n = (long long)nS * g->nViews * g->nCols;
    double *q = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    orc_ramp_filter(g, sino, q, (long long)nS * g->nViews, nthreads);
    orc_bp(g, angles, q, img, nS, 1, nthreads);
    free(q);
}

void orc_fbp_adjoint(const orc_geom *g, const float *angles, const double *img,
                     double *sino, int nS, int nthreads)
{
    const long long n = (long long)nS * g->nViews * g->nCols;
    double *t = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    orc_bp_adjoint(g, angles, img, t, nS, 1, nthreads);
    orc_ramp_filter(g, t, sino, (long long)nS * g->nViews, nthreads);
    free(t);
}

/* ------------------------------------------------------------------ FDK */

static void fdk_item(void *vctx, long long item)
{
    const orc_ctx *x = (const orc_ctx *)vctx;
    const orc_geom *g = x->g;
    const int nRows = x->nRows, nZ = x->nZ;
    const int j = (int)(item % g->nY);
    const long long s = item / g->nY;
    const double *proj = x->in + s * (long long)g->nViews * nRows * g->nCols;
    double *vol = x->out + s * (long long)nZ * g->nY * g->nX;
    const double r2 = (double)g->rFOV * (double)g->rFOV;
    const double y = (double)g->ymin + (double)g->dy * ((double)j + 0.5);
    for (int i = 0; i < g->nX; ++i)
        for (int k = 0; k < nZ; ++k) vol[((long long)k * g->nY + j) * g->nX + i] = 0.0;
    for (int i = 0; i < g->nX; ++i) {
        const double xx = (double)g->xmin + (double)g->dx * ((double)i + 0.5);
        if (!(xx * xx + y * y <= r2)) continue;
        for (int v = 0; v < g->nViews; ++v) {
            const double beta = (double)x->angles[v];
            const double cb = cos(beta), sb = sin(beta);
            double xi, w;
            if (!bp_coord(g, cb, sb, xx, y, &xi, &w)) continue;
            const double U = (double)g->R - xx * cb - y * sb;
            const double ct = xi - 0.5;
            const double fc = floor(ct);
            const int c0 = (int)fc;
            const double al = ct - fc;
            const int ca = clampi(c0, 0, g->nCols - 1), cb2 = clampi(c0 + 1, 0, g->nCols - 1);
            const double *V = proj + (long long)v * nRows * g->nCols;
            for (int k = 0; k < nZ; ++k) {
                /* FDK/GPURecon_GPU.cu:43-45, source z = 0 (:226) */
                const double z = x->zmin + (double)g->dz * ((double)k + 0.5);
                const double rowIdx = (z * (double)g->D / U - (double)g->detBottom) / (double)g->dRow;
                const double rt = rowIdx - 0.5;
                const double fr = floor(rt);
                const int r0 = (int)fr;
                const double be = rt - fr;
                const int ra = clampi(r0, 0, nRows - 1), rb = clampi(r0 + 1, 0, nRows - 1);
                const double top = (1.0 - al) * V[(long long)ra * g->nCols + ca] + al * V[(long long)ra * g->nCols + cb2];
                const double bot = (1.0 - al) * V[(long long)rb * g->nCols + ca] + al * V[(long long)rb * g->nCols + cb2];
                vol[((long long)k * g->nY + j) * g->nX + i] += ((1.0 - be) * top + be * bot) * w;
            }
        }
    }
}

void orc_fdk(const orc_geom *g, const float *angles, const double *proj, double *vol,
             int nS, int nRows, int nZ, double zmin, int nthreads)
{
    const long long n = (long long)nS * g->nViews * nRows * g->nCols;
    double *q = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    /* FDK/GPURecon_GPU.cu:207-215: every (view, row) line is filtered alike. */
    orc_ramp_filter(g, proj, q, (long long)nS * g->nViews * nRows, nthreads);
    orc_ctx x = {g, angles, q, vol, nS, 1, zmin, nRows, nZ, NULL, NULL, 0};
    orc_parallel_for((long long)nS * g->nY, nthreads, fdk_item, &x);
    free(q);
}

/* ------------------------------------------------------------ cone beam */

void orc_cone_taps(const orc_cone *g, double *h)
{
    const int nC = g->nCols;
    const double dc = g->dCol;
    if (g->detector == 0) {
        /* discrete Ram-Lak of the flat detector, times dCol (the convolution integral),
         * dAngle (the view integral), R D (magnification) and 1/2 (every ray twice) */
        const double pi = 3.14159265358979323846;
        const double kappa = 0.5 * dc * g->dAngle * g->R * g->D;
        for (int n = -(nC - 1); n <= nC - 1; ++n) {
            double v;
            if (n == 0) v = 1.0 / (4.0 * dc * dc);
            else if (n % 2 != 0) v = -1.0 / (pi * pi * (double)n * (double)n * dc * dc);
            else v = 0.0;
            h[n + nC - 1] = v * kappa;
        }
    } else {
        /* the reference taps, FDK/GPURecon_GPU.cu:190-200 */
        const double kappa = g->R * g->dAngle * dc;
        for (int n = -(nC - 1); n <= nC - 1; ++n) {
            double v;
            if (n == 0) v = 0.5 / (4.0 * dc * dc);
            else if (n % 2 != 0) {
                const double s = sin((double)n * dc);
                v = -0.5 / (REF_PI_D * REF_PI_D * s * s);
            } else v = 0.0;
            h[n + nC - 1] = v * kappa;
        }
    }
}

void orc_cone_preweight(const orc_cone *g, double *w)
{
    for (int r = 0; r < g->nRows; ++r)
        for (int c = 0; c < g->nCols; ++c) {
            double v = 1.0;
            if (g->detector == 0) {
                const double u = g->colLeft + ((double)c + 0.5 - g->dShift) * g->dCol;
                const double vv = g->rowBottom + ((double)r + 0.5) * g->dRow;
                v = g->D / sqrt(g->D * g->D + u * u + vv * vv);
            }
            w[(long long)r * g->nCols + c] = v;
        }
}

typedef struct {
    const orc_cone *g;
    const float *angles;
    const double *in;
    double *out;
    const double *h, *w, *pk;
} orc_cone_ctx;

static void cone_filt_item(void *vctx, long long line)
{
    const orc_cone_ctx *x = (const orc_cone_ctx *)vctx;
    const int nC = x->g->nCols, r = (int)(line % x->g->nRows);
    const int v = (int)((line / x->g->nRows) % x->g->nViews);
    const double *p = x->in + line * (long long)nC;
    const double *w = x->w + (long long)r * nC;
    const double *pk = x->pk ? x->pk + (long long)v * nC : NULL;
    double *q = x->out + line * (long long)nC;
    for (int c = 0; c < nC; ++c) {
        double acc = 0.0;
        for (int m = 0; m < nC; ++m) acc += p[m] * w[m] * (pk ? pk[m] : 1.0) * x->h[c - m + nC - 1];
        q[c] = acc;
    }
}

/* Fan angle of column c's centre as the cone backprojection maps it. */
static double cone_gamma(const orc_cone *g, double xi)
{
    const double u = g->colLeft + (xi - g->dShift) * g->dCol;
    return g->detector == 0 ? atan(u / g->D) : u;
}

static int cone_parker(const orc_cone *g, const float *angles, double *pk)
{
    double *gam = (double *)malloc(sizeof(double) * g->nCols);
    for (int c = 0; c < g->nCols; ++c) gam[c] = cone_gamma(g, c + 0.5);
    const double delta = fmax(fabs(cone_gamma(g, 0.0)), fabs(cone_gamma(g, (double)g->nCols)));
    const int rc = orc_parker_weights(angles, g->nViews, gam, g->nCols, delta, pk);
    free(gam);
    return rc;
}

int orc_cone_filter(const orc_cone *g, const float *angles, const double *proj, double *q, int nS, int nthreads)
{
    double *pk = NULL;
    if (g->short_scan) {
        pk = (double *)malloc(sizeof(double) * (size_t)g->nViews * g->nCols);
        if (cone_parker(g, angles, pk) != 0) {
            free(pk);
            return -1;
        }
    }
    double *h = (double *)malloc(sizeof(double) * (2 * g->nCols - 1));
    double *w = (double *)malloc(sizeof(double) * (size_t)g->nRows * g->nCols);
    orc_cone_taps(g, h);
    orc_cone_preweight(g, w);
    orc_cone_ctx x = {g, NULL, proj, q, h, w, pk};
    orc_parallel_for((long long)nS * g->nViews * g->nRows, nthreads, cone_filt_item, &x);
    free(h);
    free(w);
    free(pk);
    return 0;
}

static void cone_bp_item(void *vctx, long long item)
{
    const orc_cone_ctx *x = (const orc_cone_ctx *)vctx;
    const orc_cone *g = x->g;
    const int nC = g->nCols, nR = g->nRows;
    const int j = (int)(item % g->nY);
    const long long s = item / g->nY;
    const double *q = x->in + s * (long long)g->nViews * nR * nC;
    double *vol = x->out + s * (long long)g->nZ * g->nY * g->nX;
    const double y = g->ymin + g->dy * ((double)j + 0.5);
    for (int i = 0; i < g->nX; ++i)
        for (int k = 0; k < g->nZ; ++k) vol[((long long)k * g->nY + j) * g->nX + i] = 0.0;
    for (int i = 0; i < g->nX; ++i) {
        const double xx = g->xmin + g->dx * ((double)i + 0.5);
        if (!(xx * xx + y * y <= g->rFOV * g->rFOV)) continue;
        for (int v = 0; v < g->nViews; ++v) {
            const double beta = (double)x->angles[v];
            const double cb = cos(beta), sb = sin(beta);
            const double U = g->R - xx * cb - y * sb;
            const double t = xx * sb - y * cb;
            const double xi = (g->detector == 0 ? (g->D * t / U - g->colLeft) : (atan(t / U) - g->colLeft)) / g->dCol +
                              g->dShift;
            if (!(xi > 0.0 && xi < (double)nC)) continue;
            const double wgt = g->detector == 0 ? 1.0 / (U * U) : 1.0 / (U * U + t * t);
            const double ct = xi - 0.5, fc = floor(ct), al = ct - fc;
            const int ca = clampi((int)fc, 0, nC - 1), cb2 = clampi((int)fc + 1, 0, nC - 1);
            const double *V = q + (long long)v * nR * nC;
            for (int k = 0; k < g->nZ; ++k) {
                const double z = g->zmin + g->dz * ((double)k + 0.5);
                const double rt = (g->D * z / U - g->rowBottom) / g->dRow - 0.5;
                const double fr = floor(rt), be = rt - fr;
                const int ra = clampi((int)fr, 0, nR - 1), rb = clampi((int)fr + 1, 0, nR - 1);
                const double top = (1.0 - al) * V[(long long)ra * nC + ca] + al * V[(long long)ra * nC + cb2];
                const double bot = (1.0 - al) * V[(long long)rb * nC + ca] + al * V[(long long)rb * nC + cb2];
                vol[((long long)k * g->nY + j) * g->nX + i] += ((1.0 - be) * top + be * bot) * wgt;
            }
        }
    }
}

void orc_cone_bp(const orc_cone *g, const float *angles, const double *q, double *vol, int nS, int nthreads)
{
    orc_cone_ctx x = {g, angles, q, vol, NULL, NULL, NULL};
    orc_parallel_for((long long)nS * g->nY, nthreads, cone_bp_item, &x);
}

int orc_cone_fdk(const orc_cone *g, const float *angles, const double *proj, double *vol, int nS, int nthreads)
{
    const long long n = (long long)nS * g->nViews * g->nRows * g->nCols;
    double *q = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (orc_cone_filter(g, angles, proj, q, nS, nthreads) != 0) {
        free(q);
        return -1;
    }
    orc_cone_bp(g, angles, q, vol, nS, nthreads);
    free(q);
    return 0;
}

/* ------------------------------------------------------------ short scan */

int orc_parker_weights(const float *angles, int nViews, const double *gamma, int nCols, double delta, double *w)
{
    const double pi = 3.14159265358979323846;
    const double s = angles[nViews - 1] >= angles[0] ? 1.0 : -1.0;
    const double span = s * ((double)angles[nViews - 1] - (double)angles[0]);
    /* the last view may sit one view spacing short of the end of the range */
    const double step = nViews > 1 ? span / (nViews - 1) : 0.0;
    if (span + step < pi + 2.0 * delta - 1e-9) return -1;
    for (int v = 0; v < nViews; ++v) {
        const double b = s * ((double)angles[v] - (double)angles[0]);
        for (int c = 0; c < nCols; ++c) {
            const double g = s * gamma[c];
            double x;
            if (b <= 2.0 * delta - 2.0 * g) {
                const double t = delta - g > 0.0 ? sin(0.25 * pi * b / (delta - g)) : 1.0;
                x = t * t;
            } else if (b <= pi - 2.0 * g) {
                x = 1.0;
            } else if (b <= pi + 2.0 * delta) {
                const double t = delta + g > 0.0 ? sin(0.25 * pi * (pi + 2.0 * delta - b) / (delta + g)) : 0.0;
                x = t * t;
            } else {
                x = 0.0;
            }
            w[(long long)v * nCols + c] = 2.0 * x;
        }
    }
    return 0;
}

int orc_parker_fan(const orc_geom *g, const float *angles, double *w)
{
    double *gam = (double *)malloc(sizeof(double) * g->nCols);
    for (int c = 0; c < g->nCols; ++c)
        gam[c] = (double)g->dLeft + ((double)c + 0.5 - (double)g->dShift) * (double)g->dCol;
    const double e0 = (double)g->dLeft - (double)g->dShift * (double)g->dCol;
    const double e1 = (double)g->dLeft + ((double)g->nCols - (double)g->dShift) * (double)g->dCol;
    const int rc = orc_parker_weights(angles, g->nViews, gam, g->nCols, fmax(fabs(e0), fabs(e1)), w);
    free(gam);
    return rc;
}

int orc_fbp_short_scan(const orc_geom *g, const float *angles, const double *sino, double *img, int nS,
                       int nthreads)
{
    const long long per = (long long)g->nViews * g->nCols;
    double *w = (double *)malloc(sizeof(double) * per);
    if (orc_parker_fan(g, angles, w) != 0) {
        free(w);
        return -1;
    }
    double *p = (double *)malloc(sizeof(double) * (nS * per > 0 ? nS * per : 1));
    for (long long k = 0; k < nS * per; ++k) p[k] = sino[k] * w[k % per];
    orc_fbp(g, angles, p, img, nS, nthreads);
    free(p);
    free(w);
    return 0;
}

/* FBP with pre-weights on the sinogram before the filter: cos(gamma_c) (the equiangular
 * FBP's cosine weighting, Hsieh eq. 3.51; the reference's FBP omits it,
 * FBP/transform.cu:162) and/or Parker's weights.  gamma_c is the column's fan angle as
 * the backprojection maps it, dLeft + (c + 1/2 - dShift) dCol. */
int orc_fbp_weighted(const orc_geom *g, const float *angles, const double *sino, double *img, int nS, int cosine,
                     int short_scan, int nthreads)
{
    const long long per = (long long)g->nViews * g->nCols;
    double *w = (double *)malloc(sizeof(double) * per);
    if (short_scan) {
        if (orc_parker_fan(g, angles, w) != 0) {
            free(w);
            return -1;
        }
    } else {
        for (long long k = 0; k < per; ++k) w[k] = 1.0;
    }
    if (cosine)
        for (long long k = 0; k < per; ++k) {
            const int c = (int)(k % g->nCols);
            w[k] *= cos((double)g->dLeft + ((double)c + 0.5 - (double)g->dShift) * (double)g->dCol);
        }
    double *p = (double *)malloc(sizeof(double) * (nS * per > 0 ? nS * per : 1));
    for (long long k = 0; k < nS * per; ++k) p[k] = sino[k] * w[k % per];
    orc_fbp(g, angles, p, img, nS, nthreads);
    free(p);
    free(w);
    return 0;
}

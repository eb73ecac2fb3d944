// plan.cpp -- parameter decode, FP64 geometry tables, plan lifecycle, error reporting.
//
// Everything here runs on the host.  Geometry that the kernels consume per ray or per
// view is computed once in double precision and stored at the precision the kernels
// read (FP32 positions, exact integer sample counts, FP64 ramp taps).

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>

#include "fbct_internal.hpp"

namespace {

thread_local std::string g_last_error;

// The reference's PI macro (FP/transform.h:1): float for the attribute decode, as a
// double literal for the ramp taps (FBP/transform.cu:158).
constexpr float kRefPiF = 3.14159265359f;
constexpr double kRefPiD = 3.14159265359;

}  // namespace

namespace fbct {

fbct_status fail(fbct_status s, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

fbct_status ok()
{
    return FBCT_OK;
}

// ---------------------------------------------------------------- decode
// FP/transform.cc:59-76 (identical in BP/FBP transform.cc and FDK/FBP.cc:61-84):
// float arithmetic in the reference's operation order.
static fbct_status decode(const float *p, fbct_fan_geom *g)
{
    std::memset(g, 0, sizeof(*g));
    g->D = p[0];
    g->R = p[1];
    g->nViews = static_cast<int32_t>(p[2]);
    g->dRange = p[3];
    g->nCols = static_cast<int32_t>(p[4]);
    g->nX = static_cast<int32_t>(p[5]);
    g->nY = static_cast<int32_t>(p[6]);
    g->rFOV = p[7];
    g->dShift = p[9];
    g->nSamp = static_cast<int32_t>(p[11]);
    g->detBottom = p[13];
    g->dRow = p[14];
    g->dz = p[15];
    if (g->nViews <= 0 || g->nCols <= 0 || g->nX <= 0 || g->nY <= 0)
        return fail(FBCT_ERR_INVALID_GEOMETRY, "nViews, nCols, nX, nY must be positive (got %d, %d, %d, %d)",
                    g->nViews, g->nCols, g->nX, g->nY);
    volatile float t;
    t = g->dRange / static_cast<float>(g->nViews);
    t = t * kRefPiF;
    g->dAngle = t / 180.0f;
    t = 2.0f * g->rFOV;
    g->dx = t / static_cast<float>(g->nX);
    t = 2.0f * g->rFOV;
    g->dy = t / static_cast<float>(g->nY);
    t = -0.5f * static_cast<float>(g->nX);
    g->xmin = t * g->dx;
    t = -0.5f * static_cast<float>(g->nY);
    g->ymin = t * g->dy;
    t = p[8] * kRefPiF;
    g->dCol = t / 180.0f;
    t = p[10] * kRefPiF;
    g->dLeft = t / 180.0f;
    t = p[12] * kRefPiF;
    g->startAngle = t / 180.0f;
    return FBCT_OK;
}

static fbct_status validate(const fbct_fan_geom &g)
{
    if (!(g.rFOV > 0.0f) || !std::isfinite(g.rFOV))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "FOV radius must be positive (got %g)", (double)g.rFOV);
    if (!(g.R > g.rFOV) || !std::isfinite(g.R))
        return fail(FBCT_ERR_INVALID_GEOMETRY,
                    "source-isocentre distance R=%g must exceed the FOV radius %g", (double)g.R,
                    (double)g.rFOV);
    if (!(g.dCol != 0.0f) || !std::isfinite(g.dCol) || !std::isfinite(g.dLeft) || !std::isfinite(g.dShift))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "column pitch must be finite and non-zero");
    if ((int64_t)g.nX * g.nY > (int64_t(1) << 30) || (int64_t)g.nViews * g.nCols > (int64_t(1) << 30))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "image or sinogram larger than 2^30 elements");
    if (g.nX > 16384 || g.nY > 16384 || g.nCols > 65536)
        return fail(FBCT_ERR_INVALID_GEOMETRY, "image side or detector width too large");
    return FBCT_OK;
}

// Runs fn(v_lo, v_hi) over [0, nV) on up to 16 threads (disjoint view ranges).
template <typename Fn>
static void parallel_views(int nV, Fn &&fn)
{
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = static_cast<int>(std::min<unsigned>(hw, static_cast<unsigned>(std::max(1, nV / 32))));
    if (nt <= 1) {
        fn(0, nV);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(fn, nV * t / nt, nV * (t + 1) / nt);
    for (std::thread &x : th) x.join();
}

// ------------------------------------------------------------ FP ray table
// Per ray (view v, column c), FP/transform.cu:64-101 restated in FP64 in pixel units of
// dx (the reference launches it with R/dx and rFOV/dx, :239).  Sample k of the ray is
// at (u0 + k du, w0 + k dw) in texel-centre coordinates (the +0.5 texel offset of the
// unnormalised linear texture is folded into u0, w0).
struct RayGeo {
    int n;
    double u0, w0, du, dw;
};

static RayGeo ray_geometry(const fbct_fan_geom &g, double beta, int c)
{
    const double Rp = static_cast<double>(g.R) / static_cast<double>(g.dx);
    const double r = static_cast<double>(g.rFOV) / static_cast<double>(g.dx);
    const double gam = static_cast<double>(g.dLeft) + (static_cast<double>(c) + 0.5) * static_cast<double>(g.dCol);
    const double th = beta + gam;
    const double d = Rp * std::sin(gam);
    const double disc = r * r - d * d;
    RayGeo out{0, 0.0, 0.0, 0.0, 0.0};
    // Missing the FOV circle gives NaN geometry and int(NaN) = 0 samples in the
    // reference; a chord under one step gives floor(L/0.125) = 0.
    if (!(disc > 0.0)) return out;
    const double a = std::sqrt(disc);
    const double nf = std::floor(16.0 * a);  // floor(L / 0.125), L = 2a
    if (nf < 1.0) return out;
    const int n = static_cast<int>(nf);
    const double ct = std::cos(th), st = std::sin(th);
    out.n = n;
    out.du = (2.0 * a * ct) / static_cast<double>(n);
    out.dw = (2.0 * a * st) / static_cast<double>(n);
    out.u0 = (-a * ct + d * st) + r - 0.5;  // Rot_th(-a, -d) + r - 1/2
    out.w0 = (-a * st - d * ct) + r - 0.5;
    return out;
}

// Round a float to the nearest-even bfloat16 (raw bits) and back.
uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return static_cast<uint16_t>((u >> 16) | 0x40u);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

double bf16_value(uint16_t b)
{
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float f;
    std::memcpy(&f, &u, sizeof(f));
    return static_cast<double>(f);
}

bool parker_weights(const std::vector<float> &angles, const std::vector<double> &gamma, double delta,
                    std::vector<float> &w)
{
    const double pi = 3.14159265358979323846;
    const int nV = static_cast<int>(angles.size()), nC = static_cast<int>(gamma.size());
    if (nV < 1) return false;
    const double s = angles[nV - 1] >= angles[0] ? 1.0 : -1.0;
    const double span = s * (static_cast<double>(angles[nV - 1]) - angles[0]);
    const double step = nV > 1 ? span / (nV - 1) : 0.0;  // the last view may end one step short
    if (span + step < pi + 2.0 * delta - 1e-9) return false;
    w.assign(static_cast<size_t>(nV) * nC, 0.0f);
    for (int v = 0; v < nV; ++v) {
        const double b = s * (static_cast<double>(angles[v]) - angles[0]);
        for (int c = 0; c < nC; ++c) {
            const double g = s * gamma[c];
            double x;
            if (b <= 2.0 * delta - 2.0 * g) {
                const double t = delta - g > 0.0 ? std::sin(0.25 * pi * b / (delta - g)) : 1.0;
                x = t * t;
            } else if (b <= pi - 2.0 * g) {
                x = 1.0;
            } else if (b <= pi + 2.0 * delta) {
                const double t = delta + g > 0.0 ? std::sin(0.25 * pi * (pi + 2.0 * delta - b) / (delta + g)) : 0.0;
                x = t * t;
            } else {
                x = 0.0;
            }
            w[static_cast<size_t>(v) * nC + c] = static_cast<float>(2.0 * x);
        }
    }
    return true;
}

// Fan angles of the column centres as the BP maps them, and the detector edge bound.
static void fan_gammas(const fbct_fan_geom &g, std::vector<double> &gamma, double &delta)
{
    const double dl = g.dLeft, dc = g.dCol, sh = g.dShift;
    gamma.resize(g.nCols);
    for (int c = 0; c < g.nCols; ++c) gamma[c] = dl + (c + 0.5 - sh) * dc;
    delta = std::max(std::fabs(dl - sh * dc), std::fabs(dl + (g.nCols - sh) * dc));
}

void ramp_tap_table(const std::vector<double> &taps, int nC, int np, int kp, std::vector<uint16_t> &table, int &lg)
{
    lg = ramp_table_len(np, kp);
    table.assign(static_cast<size_t>(3 * kRampSkew) * lg, 0);
    const size_t plane = static_cast<size_t>(kRampSkew) * lg;
    for (int r = 0; r < kRampSkew; ++r)
        for (int u = 0; u < lg; ++u) {
            const int d = u - r + ramp_table_dlo(np);
            if (d < -(nC - 1) || d > nC - 1) continue;
            const double x = taps[d + nC - 1];
            const uint16_t hi = bf16_rne(static_cast<float>(x));
            const double r1 = x - bf16_value(hi);
            const uint16_t lo = bf16_rne(static_cast<float>(r1));
            const double r2 = r1 - bf16_value(lo);
            const uint16_t ll = bf16_rne(static_cast<float>(r2));
            const size_t o = static_cast<size_t>(r) * lg + u;
            table[o] = hi;
            table[plane + o] = lo;
            table[2 * plane + o] = ll;
        }
}

// Fit atan(z) ~ z + z*s*R(s), s = z^2, on |z| <= zmax: R interpolates
// (atan(sqrt(s))/sqrt(s) - 1)/s at nc Chebyshev nodes of [0, zmax^2] (solved in FP64 in
// the monomial basis, well conditioned for zmax < 1), coefficients rounded to FP32.
// Accepted when the relative error of the rounded polynomial, evaluated in FP64 on a
// dense grid of (0, zmax], is at most 1e-8.
bool fit_atan(double zmax, int nc, float *coef)
{
    const double smax = zmax * zmax;
    double a[kBpAtanMax][kBpAtanMax + 1];
    for (int i = 0; i < nc; ++i) {
        const double si = smax * 0.5 * (1.0 + std::cos(3.14159265358979323846 * (i + 0.5) / nc));
        const double zi = std::sqrt(si);
        double pw = 1.0;
        for (int k = 0; k < nc; ++k, pw *= si) a[i][k] = pw;
        a[i][nc] = (std::atan(zi) / zi - 1.0) / si;
    }
    for (int c = 0; c < nc; ++c) {  // Gaussian elimination, partial pivoting
        int piv = c;
        for (int r = c + 1; r < nc; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return false;
        for (int k = 0; k <= nc; ++k) std::swap(a[c][k], a[piv][k]);
        for (int r = 0; r < nc; ++r) {
            if (r == c) continue;
            const double f = a[r][c] / a[c][c];
            for (int k = c; k <= nc; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int k = 0; k < nc; ++k) coef[k] = static_cast<float>(a[k][nc] / a[k][k]);
    constexpr int kGrid = 4096;
    for (int i = 1; i <= kGrid; ++i) {
        const double z = zmax * i / kGrid, sz = z * z;
        double r = coef[nc - 1];
        for (int k = nc - 2; k >= 0; --k) r = r * sz + coef[k];
        const double at = std::atan(z);
        if (!(std::fabs(z + z * sz * r - at) <= 1e-8 * at)) return false;
    }
    return true;
}

static void build_tables(fbct_plan &p)
{
    const fbct_fan_geom &g = p.g;
    const int nV = g.nViews, nC = g.nCols;
    const double r = static_cast<double>(g.rFOV) / static_cast<double>(g.dx);

    p.ray_n.assign(static_cast<size_t>(nV) * nC, 0);
    p.ray_geo.assign(static_cast<size_t>(nV) * nC * 4, 0.0f);
    p.ray_inv.assign(static_cast<size_t>(nV) * nC * 2, 0.0f);
    p.view_cs.assign(static_cast<size_t>(nV) * 2, 0.0f);
    p.view_csd.assign(static_cast<size_t>(nV) * 2, 0.0);
    int64_t samples = 0;
    for (int v = 0; v < nV; ++v) {
        const double beta = static_cast<double>(p.angles[v]);
        p.view_csd[2 * v] = std::cos(beta);
        p.view_csd[2 * v + 1] = std::sin(beta);
        p.view_cs[2 * v] = static_cast<float>(p.view_csd[2 * v]);
        p.view_cs[2 * v + 1] = static_cast<float>(p.view_csd[2 * v + 1]);
        for (int c = 0; c < nC; ++c) {
            const RayGeo rg = ray_geometry(g, beta, c);
            const size_t k = static_cast<size_t>(v) * nC + c;
            p.ray_n[k] = rg.n;
            p.ray_geo[4 * k + 0] = static_cast<float>(rg.u0);
            p.ray_geo[4 * k + 1] = static_cast<float>(rg.w0);
            p.ray_geo[4 * k + 2] = static_cast<float>(rg.du);
            p.ray_geo[4 * k + 3] = static_cast<float>(rg.dw);
            // reciprocal of the FP32 step the kernels use; a zero step gets a large finite value
            const float du = p.ray_geo[4 * k + 2], dw = p.ray_geo[4 * k + 3];
            p.ray_inv[2 * k + 0] = du != 0.0f ? static_cast<float>(1.0 / static_cast<double>(du)) : 1e30f;
            p.ray_inv[2 * k + 1] = dw != 0.0f ? static_cast<float>(1.0 / static_cast<double>(dw)) : 1e30f;
            samples += rg.n;
        }
    }
    p.fp_samples = samples;

    // FP bundles: rays [b*rays, (b+1)*rays) of a view march slab by slab along the
    // axis their central ray is closest to; slab sl holds the samples whose major
    // coordinate lies in [sl*slab - 1, (sl+1)*slab - 1) and stages taps [alo, alo+slab]
    // along it.  Per (bundle, slab) the tile's minor rows [lo, hi] bound the taps of the
    // bundle's samples in the slab: each ray's segment in the slab, in FP64 on the FP32
    // ray parameters the kernel uses, widened by one row on each side (which covers the
    // kernel's FP32 sample positions).  Per bundle: the first and last slab with samples
    // (a slab outside is empty for every ray, also in FP32: the range is widened by 0.01
    // pixel).  cap is the largest row count of a tile.
    Derived &d = p.d;
    d.fp_scale = static_cast<float>(0.125 * static_cast<double>(g.dx));
    d.two_r = static_cast<float>(2.0 * r);
    for (int cfg = 0; cfg < kFpNumCfg; ++cfg) {
        const int rays = kFpCfgRays[cfg];
        FpBundleCfg &bc = d.fpb[cfg];
        const int slab = kFpCfgSlab[cfg];
        bc.rays = rays;
        bc.slab = slab;
        bc.nslab = static_cast<int>(std::floor((2.0 * r + 1.0) / slab)) + 1;
        bc.nb = (nC + rays - 1) / rays;
        const int nb = bc.nb, nslab = bc.nslab;
        std::vector<uint8_t> &major = p.fp_major[cfg];
        major.assign(static_cast<size_t>(nV) * nb, 1);
        std::vector<int16_t> &slabs = p.fp_slabs[cfg];
        std::vector<int16_t> &rows = p.fp_rows[cfg];
        slabs.assign(static_cast<size_t>(nV) * nb * 2, 0);
        rows.assign(static_cast<size_t>(nV) * nb * nslab * 2, 0);
        std::vector<int> cap_of(nV, 0);
        auto views = [&](int v_lo, int v_hi) {
            std::vector<double> lo(nslab), hi(nslab);
            for (int v = v_lo; v < v_hi; ++v) {
                const double beta = static_cast<double>(p.angles[v]);
                for (int b = 0; b < nb; ++b) {
                    const int c0 = b * rays;
                    const int c1 = std::min(nC, c0 + rays) - 1;
                    const double gmid = static_cast<double>(g.dLeft) + (0.5 * (c0 + c1) + 0.5) * static_cast<double>(g.dCol);
                    const double thm = beta + gmid;
                    const bool xmajor = std::fabs(std::cos(thm)) >= std::fabs(std::sin(thm));
                    const size_t vb = static_cast<size_t>(v) * nb + b;
                    major[vb] = xmajor ? 1 : 0;
                    std::fill(lo.begin(), lo.end(), 1e300);
                    std::fill(hi.begin(), hi.end(), -1e300);
                    int s_first = nslab, s_last = -1;
                    for (int c = c0; c <= c1; ++c) {
                        const size_t k = static_cast<size_t>(v) * nC + c;
                        const int n = p.ray_n[k];
                        if (n <= 0) continue;
                        const double u0 = p.ray_geo[4 * k + 0], w0 = p.ray_geo[4 * k + 1];
                        const double du = p.ray_geo[4 * k + 2], dw = p.ray_geo[4 * k + 3];
                        const double a0 = xmajor ? u0 : w0, da = xmajor ? du : dw;
                        const double m0 = xmajor ? w0 : u0, dm = xmajor ? dw : du;
                        const double a1 = a0 + (n - 1) * da, m1 = m0 + (n - 1) * dm;
                        const double amin = std::min(a0, a1) - 0.01, amax = std::max(a0, a1) + 0.01;
                        // minor coordinate at major coordinate e (a zero major step: the whole ray)
                        auto m_at = [&](double e, bool upper) {
                            if (da == 0.0) return upper ? std::max(m0, m1) : std::min(m0, m1);
                            return m0 + (e - a0) * (dm / da);
                        };
                        const int sl0 = std::max(0, static_cast<int>(std::floor((amin + 1.0) / slab)));
                        const int sl1 = std::min(nslab - 1, static_cast<int>(std::floor((amax + 1.0) / slab)));
                        for (int sl = sl0; sl <= sl1; ++sl) {
                            const double e0 = std::max(amin, sl * static_cast<double>(slab) - 1.0);
                            const double e1 = std::min(amax, (sl + 1) * static_cast<double>(slab) - 1.0);
                            if (e0 > e1) continue;
                            const double mA = m_at(e0, false), mB = m_at(e1, true);
                            lo[sl] = std::min(lo[sl], std::min(mA, mB));
                            hi[sl] = std::max(hi[sl], std::max(mA, mB));
                            s_first = std::min(s_first, sl);
                            s_last = std::max(s_last, sl);
                        }
                    }
                    slabs[2 * vb + 0] = static_cast<int16_t>(s_first);
                    slabs[2 * vb + 1] = static_cast<int16_t>(s_last);
                    for (int sl = 0; sl < nslab; ++sl) {
                        int16_t *e = &rows[2 * (vb * nslab + sl)];
                        if (hi[sl] < lo[sl]) {
                            e[0] = 1;  // empty: lo > hi
                            e[1] = 0;
                            continue;
                        }
                        const int rlo = static_cast<int>(std::floor(lo[sl])) - 1;
                        const int rhi = static_cast<int>(std::floor(hi[sl])) + 2;
                        e[0] = static_cast<int16_t>(rlo);
                        e[1] = static_cast<int16_t>(rhi);
                        cap_of[v] = std::max(cap_of[v], rhi - rlo + 1);
                    }
                }
            }
        };
        parallel_views(nV, views);
        int cap = 1;
        for (int c : cap_of) cap = std::max(cap, c);
        bc.cap = cap;
        // items by decreasing largest tile (ties: item order)
        const size_t nitems = static_cast<size_t>(nV) * nb;
        std::vector<int16_t> item_rows(nitems, 0);
        for (size_t vb = 0; vb < nitems; ++vb)
            for (int sl = 0; sl < nslab; ++sl) {
                const int16_t *e = &rows[2 * (vb * nslab + sl)];
                if (e[0] <= e[1]) item_rows[vb] = std::max<int16_t>(item_rows[vb], static_cast<int16_t>(e[1] - e[0] + 1));
            }
        std::vector<int32_t> &order = p.fp_order[cfg];
        order.resize(nitems);
        for (size_t vb = 0; vb < nitems; ++vb) order[vb] = static_cast<int32_t>(vb);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t x, int32_t y) { return item_rows[x] > item_rows[y]; });
        std::vector<int16_t> &orows = p.fp_order_rows[cfg];
        orows.resize(nitems);
        for (size_t k = 0; k < nitems; ++k) orows[k] = item_rows[order[k]];
    }

    // BP: xi_raw = (atan(t/U) - dLeft)/dCol + dShift  (BP/transform.cu:85).
    d.bp_xi_scale = static_cast<float>(1.0 / static_cast<double>(g.dCol));
    d.bp_xi_offset = static_cast<float>(static_cast<double>(g.dShift) -
                                        static_cast<double>(g.dLeft) / static_cast<double>(g.dCol));
    d.r2 = g.rFOV * g.rFOV;

    // BP^T gathers per (view, column): gamma_c of every column in FP64 and the FOV disk
    // as pixel ranges per row and per column (FP64 membership test of the centres).
    p.col_csd.resize(static_cast<size_t>(2) * nC);
    for (int c = 0; c < nC; ++c) {
        const double gam = static_cast<double>(g.dLeft) +
                           (static_cast<double>(c) + 0.5 - static_cast<double>(g.dShift)) * static_cast<double>(g.dCol);
        p.col_csd[2 * c] = std::cos(gam);
        p.col_csd[2 * c + 1] = std::sin(gam);
    }
    {
        // a line meets the disk in one interval: start from the FP64 chord and settle each
        // end with the membership test itself
        const double rf2 = static_cast<double>(g.rFOV) * static_cast<double>(g.rFOV);
        auto ranges = [rf2](int n_lines, double lmin, double dl, int n, double pmin, double dp,
                            std::vector<int32_t> &out) {
            out.assign(static_cast<size_t>(2) * n_lines, 0);
            for (int k = 0; k < n_lines; ++k) {
                const double y = lmin + dl * (k + 0.5);
                auto in = [&](int i) {
                    const double x = pmin + dp * (i + 0.5);
                    return x * x + y * y <= rf2;
                };
                const double h = rf2 - y * y;
                int lo = n, hi = -1;
                if (h >= 0.0) {
                    const double e = std::sqrt(h);
                    int a = static_cast<int>(std::floor((-e - pmin) / dp - 0.5));
                    int b = static_cast<int>(std::ceil((e - pmin) / dp - 0.5));
                    a = std::min(std::max(a, 0), n - 1);
                    b = std::min(std::max(b, 0), n - 1);
                    while (a > 0 && in(a - 1)) --a;
                    while (a <= b && !in(a)) ++a;
                    while (b < n - 1 && in(b + 1)) ++b;
                    while (b >= a && !in(b)) --b;
                    if (a <= b) lo = a, hi = b;
                }
                out[2 * k] = lo;
                out[2 * k + 1] = hi;
            }
        };
        ranges(g.nY, g.ymin, g.dy, g.nX, g.xmin, g.dx, p.fov_row);
        ranges(g.nX, g.xmin, g.dx, g.nY, g.ymin, g.dy, p.fov_col);
    }
    d.bpa_tan = static_cast<float>(std::tan(std::fabs(static_cast<double>(g.dCol))));
    // FOV pixels see the source at |t/U| <= rFOV/sqrt(R^2 - rFOV^2); the margin covers
    // the FP32 pixel centres on the circle.
    {
        const double rf = static_cast<double>(g.rFOV), R = static_cast<double>(g.R);
        const double zmax = 1.001 * rf / std::sqrt(R * R - rf * rf);
        d.bp_atan_nc = 0;
        std::fill(d.bp_atan_c, d.bp_atan_c + kBpAtanMax, 0.0f);
        for (int nc : {5, 9}) {
            if (fit_atan(zmax, nc, d.bp_atan_c)) {
                d.bp_atan_nc = nc;
                break;
            }
        }
        if (d.bp_atan_nc == 0) std::fill(d.bp_atan_c, d.bp_atan_c + kBpAtanMax, 0.0f);
    }

    // Ramp taps, FBP/transform.cu:153-162 (Hsieh eq. 3.51, scaled by R*dAngle*dCol).
    p.ramp.assign(static_cast<size_t>(2 * nC - 1), 0.0);
    const double dc = static_cast<double>(g.dCol);
    const double kappa = static_cast<double>(g.R) * static_cast<double>(g.dAngle) * dc;
    for (int n = -(nC - 1); n <= nC - 1; ++n) {
        double h;
        if (n == 0) {
            h = 0.5 / (4.0 * dc * dc);
        } else if (n % 2 != 0) {
            const double s = std::sin(static_cast<double>(n) * dc);
            h = -0.5 / (kRefPiD * kRefPiD * s * s);
        } else {
            h = 0.0;
        }
        p.ramp[n + nC - 1] = h * kappa;
    }

    // MFMA filter operand: the taps split into three BF16 planes hi + lo + ll (each the
    // BF16 rounding of what the previous planes left), in skewed copies from which the
    // filter builds H[m][c] = h[c - m] tile by tile; Np and Kp pad the matrix to the tiles.
    d.ramp_np = (nC + kRampTileN - 1) / kRampTileN * kRampTileN;
    d.ramp_kp = (nC + 31) / 32 * 32;
    ramp_tap_table(p.ramp, nC, d.ramp_np, d.ramp_kp, p.ramp_planes, d.ramp_lg);

    // BP in-range test 0 < xi < nCols <=> g in (dLeft - dShift dCol, dLeft + (nCols - dShift) dCol)
    // (either order when dCol < 0); near the edges it is decided on tangents in FP64.
    {
        const double g0 = static_cast<double>(g.dLeft) - static_cast<double>(g.dShift) * dc;
        const double g1 = static_cast<double>(g.dLeft) + (nC - static_cast<double>(g.dShift)) * dc;
        const double gmin = std::min(g0, g1), gmax = std::max(g0, g1);
        const double lim = 1.5707963267948966 - 1e-6;
        d.bp_edge_exact = (gmin > -lim && gmax < lim) ? 1 : 0;
        d.bp_tan_min = d.bp_edge_exact ? std::tan(gmin) : 0.0;
        d.bp_tan_max = d.bp_edge_exact ? std::tan(gmax) : 0.0;
    }

    // FP adjoint (gather): ray c passes through a point at column coordinate
    // (g - dLeft)/dCol - 1/2 = c, the FP ignoring dShift.
    d.fpa_r = static_cast<float>(r);
    d.fpa_c_scale = static_cast<float>(1.0 / dc);
    d.fpa_c_off = static_cast<float>(-static_cast<double>(g.dLeft) / dc - 0.5);
    d.fpa_far_x = static_cast<float>(2.0 * r - g.nX + 0.5);
    d.fpa_far_y = static_cast<float>(2.0 * r - g.nY + 0.5);
}

template <typename T>
static fbct_status upload(T **dst, const void *src, size_t bytes)
{
    if (bytes == 0) bytes = sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), bytes);
    if (e != hipSuccess) return fail(FBCT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    if (src) {
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    }
    return FBCT_OK;
}

static void free_tables(DeviceTables &t)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    if (t.device >= 0) (void)hipSetDevice(t.device);
    (void)hipFree(t.ray_n);
    (void)hipFree(t.ray_geo);
    for (int cfg = 0; cfg < kFpNumCfg; ++cfg) {
        (void)hipFree(t.fp_major[cfg]);
        (void)hipFree(t.fp_slabs[cfg]);
        (void)hipFree(t.fp_rows[cfg]);
        (void)hipFree(t.fp_order[cfg]);
    }
    (void)hipFree(t.ray_inv);
    (void)hipFree(t.view_cs);
    (void)hipFree(t.view_csd);
    (void)hipFree(t.col_csd);
    (void)hipFree(t.fov_row);
    (void)hipFree(t.fov_col);
    (void)hipFree(t.ramp);
    (void)hipFree(t.ramp_planes);
    (void)hipFree(t.parker);
    (void)hipFree(t.cosine);
    if (cur >= 0) (void)hipSetDevice(cur);
    t = DeviceTables{};
}

fbct_status device_parker(fbct_plan *plan, const DeviceTables *t, const float **out)
{
    std::lock_guard<std::mutex> lk(plan->mu);
    DeviceTables *dt = nullptr;
    for (DeviceTables &d : plan->dev)
        if (&d == t) dt = &d;
    if (!dt) return fail(FBCT_ERR_DEVICE, "short scan: tables not resident on this device");
    if (!dt->parker) {
        std::vector<double> gamma;
        double delta;
        fan_gammas(plan->g, gamma, delta);
        std::vector<float> w;
        if (!parker_weights(plan->angles, gamma, delta, w))
            return fail(FBCT_ERR_INVALID_GEOMETRY, "short scan: the views do not cover 180 degrees plus the fan angle");
        fbct_status s = upload_vec(&dt->parker, w);
        if (s != FBCT_OK) return s;
    }
    *out = dt->parker;
    return FBCT_OK;
}

fbct_status device_cosine(fbct_plan *plan, const DeviceTables *t, const float **out)
{
    std::lock_guard<std::mutex> lk(plan->mu);
    DeviceTables *dt = nullptr;
    for (DeviceTables &d : plan->dev)
        if (&d == t) dt = &d;
    if (!dt) return fail(FBCT_ERR_DEVICE, "cosine weights: tables not resident on this device");
    if (!dt->cosine) {
        std::vector<double> gamma;
        double delta;
        fan_gammas(plan->g, gamma, delta);
        std::vector<float> w(gamma.size());
        for (size_t c = 0; c < gamma.size(); ++c) w[c] = static_cast<float>(std::cos(gamma[c]));
        fbct_status s = upload_vec(&dt->cosine, w);
        if (s != FBCT_OK) return s;
    }
    *out = dt->cosine;
    return FBCT_OK;
}

fbct_status resident_tables(fbct_plan *plan, const DeviceTables **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(plan->mu);
    for (const DeviceTables &t : plan->dev)
        if (t.device == dev) {
            *out = &t;
            return FBCT_OK;
        }
    return fail(FBCT_ERR_INVALID_ARGUMENT, "plan tables are not resident on device %d: call fbct_plan_upload first",
                dev);
}

fbct_status device_tables(fbct_plan *plan, const DeviceTables **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(plan->mu);
    for (const DeviceTables &t : plan->dev)
        if (t.device == dev) {
            *out = &t;
            return FBCT_OK;
        }
    // Plans are tiny; keep the vector from reallocating so returned pointers stay valid.
    if (plan->dev.capacity() == 0) plan->dev.reserve(16);
    if (plan->dev.size() == plan->dev.capacity())
        return fail(FBCT_ERR_UNSUPPORTED, "plan used on more than %zu devices", plan->dev.capacity());
    DeviceTables t;
    t.device = dev;
    fbct_status s;
    if ((s = upload(&t.ray_n, plan->ray_n.data(), plan->ray_n.size() * sizeof(int32_t))) != FBCT_OK ||
        (s = upload(&t.ray_geo, plan->ray_geo.data(), plan->ray_geo.size() * sizeof(float))) != FBCT_OK ||
        (s = upload(&t.fp_major[0], plan->fp_major[0].data(), plan->fp_major[0].size())) != FBCT_OK ||
        (s = upload(&t.fp_major[1], plan->fp_major[1].data(), plan->fp_major[1].size())) != FBCT_OK ||
        (s = upload(&t.fp_slabs[0], plan->fp_slabs[0].data(), plan->fp_slabs[0].size() * sizeof(int16_t))) != FBCT_OK ||
        (s = upload(&t.fp_slabs[1], plan->fp_slabs[1].data(), plan->fp_slabs[1].size() * sizeof(int16_t))) != FBCT_OK ||
        (s = upload(&t.fp_rows[0], plan->fp_rows[0].data(), plan->fp_rows[0].size() * sizeof(int16_t))) != FBCT_OK ||
        (s = upload(&t.fp_rows[1], plan->fp_rows[1].data(), plan->fp_rows[1].size() * sizeof(int16_t))) != FBCT_OK ||
        (s = upload(&t.fp_order[0], plan->fp_order[0].data(), plan->fp_order[0].size() * sizeof(int32_t))) != FBCT_OK ||
        (s = upload(&t.fp_order[1], plan->fp_order[1].data(), plan->fp_order[1].size() * sizeof(int32_t))) != FBCT_OK ||
        (s = upload(&t.ray_inv, plan->ray_inv.data(), plan->ray_inv.size() * sizeof(float))) != FBCT_OK ||
        (s = upload(&t.view_cs, plan->view_cs.data(), plan->view_cs.size() * sizeof(float))) != FBCT_OK ||
        (s = upload(&t.view_csd, plan->view_csd.data(), plan->view_csd.size() * sizeof(double))) != FBCT_OK ||
        (s = upload(&t.col_csd, plan->col_csd.data(), plan->col_csd.size() * sizeof(double))) != FBCT_OK ||
        (s = upload(&t.fov_row, plan->fov_row.data(), plan->fov_row.size() * sizeof(int32_t))) != FBCT_OK ||
        (s = upload(&t.fov_col, plan->fov_col.data(), plan->fov_col.size() * sizeof(int32_t))) != FBCT_OK ||
        (s = upload(&t.ramp, plan->ramp.data(), plan->ramp.size() * sizeof(double))) != FBCT_OK ||
        (s = upload(&t.ramp_planes, plan->ramp_planes.data(), plan->ramp_planes.size() * sizeof(uint16_t))) != FBCT_OK) {
        std::string msg = g_last_error;
        free_tables(t);
        g_last_error = msg;
        return s;
    }
    plan->dev.push_back(t);
    *out = &plan->dev.back();
    return FBCT_OK;
}

}  // namespace fbct

// ===================================================================== C ABI

extern "C" {

int fbct_version(void)
{
    return FBCT_VERSION;
}

const char *fbct_status_string(fbct_status s)
{
    switch (s) {
    case FBCT_OK: return "ok";
    case FBCT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case FBCT_ERR_INVALID_GEOMETRY: return "invalid geometry";
    case FBCT_ERR_SHAPE_MISMATCH: return "shape mismatch";
    case FBCT_ERR_UNSUPPORTED: return "unsupported";
    case FBCT_ERR_OUT_OF_MEMORY: return "out of memory";
    case FBCT_ERR_DEVICE: return "device error";
    case FBCT_ERR_WORKSPACE: return "workspace missing or too small";
    }
    return "unknown status";
}

const char *fbct_last_error(void)
{
    return g_last_error.c_str();
}

fbct_status fbct_geom_from_params(const float *parameter, int n_parameter, fbct_fan_geom *out)
{
    if (!parameter || !out) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null pointer");
    if (n_parameter < 16)
        return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "parameter needs 16 values, got %d", n_parameter);
    fbct_status s = fbct::decode(parameter, out);
    if (s != FBCT_OK) return s;
    return fbct::validate(*out);
}

fbct_status fbct_plan_create(const float *parameter, int n_parameter, const float *angles, int n_angles,
                             fbct_plan **out)
{
    if (!out) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null plan output pointer");
    *out = nullptr;
    fbct_fan_geom g;
    fbct_status s = fbct_geom_from_params(parameter, n_parameter, &g);
    if (s != FBCT_OK) return s;
    if (!angles) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null angle list");
    if (n_angles != g.nViews)
        return fbct::fail(FBCT_ERR_SHAPE_MISMATCH, "angle list has %d entries but parameter[2] gives %d views",
                          n_angles, g.nViews);
    for (int v = 0; v < n_angles; ++v)
        if (!std::isfinite(angles[v])) return fbct::fail(FBCT_ERR_INVALID_GEOMETRY, "angle %d is not finite", v);
    fbct_plan *p = new (std::nothrow) fbct_plan();
    if (!p) return fbct::fail(FBCT_ERR_OUT_OF_MEMORY, "plan allocation");
    try {
        p->g = g;
        p->angles.assign(angles, angles + n_angles);
        fbct::build_tables(*p);
    } catch (const std::bad_alloc &) {
        delete p;
        return fbct::fail(FBCT_ERR_OUT_OF_MEMORY, "plan tables");
    }
    *out = p;
    return FBCT_OK;
}

fbct_status fbct_plan_destroy(fbct_plan *plan)
{
    if (!plan) return FBCT_OK;
    for (fbct::DeviceTables &t : plan->dev) fbct::free_tables(t);
    delete plan;
    return FBCT_OK;
}

fbct_status fbct_plan_geom(const fbct_plan *plan, fbct_fan_geom *out)
{
    if (!plan || !out) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null pointer");
    *out = plan->g;
    return FBCT_OK;
}

int64_t fbct_plan_fp_samples(const fbct_plan *plan)
{
    return plan ? plan->fp_samples : -1;
}

fbct_status fbct_plan_ray_table(const fbct_plan *plan, int32_t *n_out, float *geo_out)
{
    if (!plan) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null plan");
    if (n_out) std::memcpy(n_out, plan->ray_n.data(), plan->ray_n.size() * sizeof(int32_t));
    if (geo_out) std::memcpy(geo_out, plan->ray_geo.data(), plan->ray_geo.size() * sizeof(float));
    return FBCT_OK;
}

fbct_status fbct_plan_fp_bundles(const fbct_plan *plan, int cfg, int32_t *info_out, uint8_t *major_out,
                                 int16_t *slabs_out, int16_t *rows_out)
{
    if (!plan) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null plan");
    if (cfg < 0 || cfg >= fbct::kFpNumCfg) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "bundle config %d", cfg);
    const fbct::FpBundleCfg &bc = plan->d.fpb[cfg];
    if (info_out) {
        info_out[0] = bc.rays;
        info_out[1] = bc.slab;
        info_out[2] = bc.nslab;
        info_out[3] = bc.nb;
        info_out[4] = bc.cap;
    }
    if (major_out) std::memcpy(major_out, plan->fp_major[cfg].data(), plan->fp_major[cfg].size());
    if (slabs_out) std::memcpy(slabs_out, plan->fp_slabs[cfg].data(), plan->fp_slabs[cfg].size() * sizeof(int16_t));
    if (rows_out) std::memcpy(rows_out, plan->fp_rows[cfg].data(), plan->fp_rows[cfg].size() * sizeof(int16_t));
    return FBCT_OK;
}

fbct_status fbct_plan_ramp_taps(const fbct_plan *plan, double *taps_out)
{
    if (!plan || !taps_out) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null pointer");
    std::memcpy(taps_out, plan->ramp.data(), plan->ramp.size() * sizeof(double));
    return FBCT_OK;
}

fbct_status fbct_plan_parker_weights(const fbct_plan *plan, float *w_out)
{
    if (!plan || !w_out) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null pointer");
    std::vector<double> gamma;
    double delta;
    fbct::fan_gammas(plan->g, gamma, delta);
    std::vector<float> w;
    if (!fbct::parker_weights(plan->angles, gamma, delta, w))
        return fbct::fail(FBCT_ERR_INVALID_GEOMETRY, "short scan: the views do not cover 180 degrees plus the fan angle");
    std::memcpy(w_out, w.data(), w.size() * sizeof(float));
    return FBCT_OK;
}

fbct_status fbct_plan_upload(fbct_plan *plan)
{
    if (!plan) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "null plan");
    const fbct::DeviceTables *t = nullptr;
    return fbct::device_tables(plan, &t);
}

}  // extern "C"

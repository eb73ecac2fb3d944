// cone.cpp -- cone-beam FDK plans: geometry validation, FP64 tables, C ABI.
//
// Reference: BP_One_Rot_3D_GPU (FDK/GPURecon_GPU.cu:163-268) filters every (view, row)
// line on the host with IPP (:184-217) and backprojects view by view with one kernel
// launch, one pageable host->device copy and an implicit sync each (:245-250).  Here a
// plan holds the per-view trigonometry (FP64, with FP32 copies), the flat-detector
// pre-weights and the Toeplitz planes of the row filter; an execute is one MFMA filter
// launch and one backprojection launch over all views, rows and slices.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr double kPi = 3.14159265358979323846;
// The reference's PI (FP/transform.h:1) as used in the FDK taps.
constexpr double kRefPiD = 3.14159265359;

bool finite_pos(float v)
{
    return std::isfinite(v) && v > 0.0f;
}

fbct_status validate(const fbct_cone_geom &g)
{
    if (!finite_pos(g.R) || !finite_pos(g.D))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: source distances must be positive (R %g, D %g)", (double)g.R,
                    (double)g.D);
    if (!finite_pos(g.rFOV) || !(g.R > g.rFOV))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: need 0 < rFOV < R (rFOV %g, R %g)", (double)g.rFOV, (double)g.R);
    if (g.detector != FBCT_DETECTOR_FLAT && g.detector != FBCT_DETECTOR_CURVED)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "cone: unknown detector type %d", g.detector);
    if (g.nViews < 1 || g.nCols < 1 || g.nRows < 1 || g.nX < 1 || g.nY < 1 || g.nZ < 1)
        return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: counts must be positive");
    if (g.nCols > 16384 || g.nRows > 16384 || g.nX > 16384 || g.nY > 16384 || g.nZ > 16384 || g.nViews > 1 << 20)
        return fail(FBCT_ERR_UNSUPPORTED, "cone: size beyond the supported limits");
    if (!std::isfinite(g.dCol) || g.dCol == 0.0f || !std::isfinite(g.dRow) || g.dRow == 0.0f)
        return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: detector pitches must be finite and non-zero");
    if (!finite_pos(g.dx) || !finite_pos(g.dy) || !finite_pos(g.dz))
        return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: voxel sizes must be positive");
    for (float v : {g.dAngle, g.colLeft, g.rowBottom, g.dShift, g.xmin, g.ymin, g.zmin})
        if (!std::isfinite(v)) return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: non-finite geometry value");
    if (g.short_scan != 0 && g.short_scan != 1)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "cone: short_scan must be 0 or 1 (got %d)", g.short_scan);
    if (g.detector == FBCT_DETECTOR_CURVED) {
        const double a0 = (double)g.colLeft - (double)g.dShift * g.dCol;
        const double a1 = (double)g.colLeft + ((double)g.nCols - g.dShift) * g.dCol;
        if (std::max(std::fabs(a0), std::fabs(a1)) >= kPi / 2)
            return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: curved detector fan must stay within +-90 degrees");
    }
    return FBCT_OK;
}

// Fan angle of detector coordinate xi (columns) as the backprojection maps it.
double cone_gamma(const fbct_cone_geom &g, double xi)
{
    const double u = (double)g.colLeft + (xi - (double)g.dShift) * (double)g.dCol;
    return g.detector == FBCT_DETECTOR_FLAT ? std::atan(u / (double)g.D) : u;
}

bool build_parker(fbct_cone_plan &p)
{
    std::vector<double> gamma(p.g.nCols);
    for (int c = 0; c < p.g.nCols; ++c) gamma[c] = cone_gamma(p.g, c + 0.5);
    const double delta = std::max(std::fabs(cone_gamma(p.g, 0.0)), std::fabs(cone_gamma(p.g, (double)p.g.nCols)));
    return parker_weights(p.angles, gamma, delta, p.parker);
}

void build(fbct_cone_plan &p)
{
    const fbct_cone_geom &g = p.g;
    const int nV = g.nViews, nC = g.nCols, nR = g.nRows;
    p.view_cs.resize(2 * (size_t)nV);
    p.view_csd.resize(2 * (size_t)nV);
    for (int v = 0; v < nV; ++v) {
        const double b = (double)p.angles[v];
        p.view_csd[2 * v] = std::cos(b);
        p.view_csd[2 * v + 1] = std::sin(b);
        p.view_cs[2 * v] = (float)p.view_csd[2 * v];
        p.view_cs[2 * v + 1] = (float)p.view_csd[2 * v + 1];
    }
    const double R = g.R, D = g.D, dc = g.dCol, dr = g.dRow;
    const bool flat = g.detector == FBCT_DETECTOR_FLAT;
    // filter taps with their constant scale
    p.taps.assign(2 * (size_t)nC - 1, 0.0);
    for (int n = -(nC - 1); n <= nC - 1; ++n) {
        double h = 0.0;
        if (flat) {
            if (n == 0) h = 1.0 / (4.0 * dc * dc);
            else if (n % 2 != 0) h = -1.0 / (kPi * kPi * (double)n * (double)n * dc * dc);
            h *= 0.5 * dc * (double)g.dAngle * R * D;
        } else {
            // FDK/GPURecon_GPU.cu:190-200 (Hsieh eq. 3.51), scaled by R dAngle dCol (:202)
            if (n == 0) h = 0.5 / (4.0 * dc * dc);
            else if (n % 2 != 0) {
                const double s = std::sin((double)n * dc);
                h = -0.5 / (kRefPiD * kRefPiD * s * s);
            }
            h *= R * (double)g.dAngle * dc;
        }
        p.taps[n + nC - 1] = h;
    }
    ConeDerived &d = p.d;
    d.ramp_np = (nC + kRampTileN - 1) / kRampTileN * kRampTileN;
    d.ramp_kp = (nC + 31) / 32 * 32;
    ramp_tap_table(p.taps, nC, d.ramp_np, d.ramp_kp, p.ramp_planes, d.ramp_lg);
    if (flat) {
        p.preweight.resize((size_t)nR * nC);
        for (int r = 0; r < nR; ++r)
            for (int c = 0; c < nC; ++c) {
                const double u = (double)g.colLeft + ((double)c + 0.5 - g.dShift) * dc;
                const double w = (double)g.rowBottom + ((double)r + 0.5) * dr;
                p.preweight[(size_t)r * nC + c] = (float)(D / std::sqrt(D * D + u * u + w * w));
            }
    }
    d.xi_scale = (float)(flat ? D / dc : 1.0 / dc);
    d.xi_offset = (float)((double)g.dShift - (double)g.colLeft / dc);
    d.zr_scale = (float)(D / dr);
    d.zr_offset = (float)(-(double)g.rowBottom / dr);
    d.r2 = g.rFOV * g.rFOV;
    // 0 < xi < nCols  <=>  t/U strictly between the detector edge values (either order)
    const double e0 = (double)g.colLeft - (double)g.dShift * dc;
    const double e1 = (double)g.colLeft + ((double)nC - g.dShift) * dc;
    double t0 = flat ? e0 / D : std::tan(e0), t1 = flat ? e1 / D : std::tan(e1);
    d.edge_exact = 1;
    d.tan_min = std::min(t0, t1);
    d.tan_max = std::max(t0, t1);
    d.atan_nc = 0;
    std::fill(d.atan_c, d.atan_c + kBpAtanMax, 0.0f);
    if (!flat) {
        const double rf = g.rFOV;
        const double zmax = 1.001 * rf / std::sqrt(R * R - rf * rf);
        for (int nc : {5, 9})
            if (fit_atan(zmax, nc, d.atan_c)) {
                d.atan_nc = nc;
                break;
            }
        if (d.atan_nc == 0) std::fill(d.atan_c, d.atan_c + kBpAtanMax, 0.0f);
    }
}

void free_cone_tables(ConeDeviceTables &t)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    if (t.device >= 0) (void)hipSetDevice(t.device);
    (void)hipFree(t.view_cs);
    (void)hipFree(t.view_csd);
    (void)hipFree(t.preweight);
    (void)hipFree(t.ramp_planes);
    (void)hipFree(t.parker);
    if (cur >= 0) (void)hipSetDevice(cur);
    t = ConeDeviceTables{};
}

fbct_status cone_tables(fbct_cone_plan *p, const ConeDeviceTables **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(p->mu);
    for (const ConeDeviceTables &t : p->dev)
        if (t.device == dev) {
            *out = &t;
            return FBCT_OK;
        }
    ConeDeviceTables t;
    t.device = dev;
    std::vector<float2> cs(p->g.nViews);
    std::vector<double2> csd(p->g.nViews);
    for (int v = 0; v < p->g.nViews; ++v) {
        cs[v] = make_float2(p->view_cs[2 * v], p->view_cs[2 * v + 1]);
        csd[v] = make_double2(p->view_csd[2 * v], p->view_csd[2 * v + 1]);
    }
    fbct_status s = upload_vec(&t.view_cs, cs);
    if (s == FBCT_OK) s = upload_vec(&t.view_csd, csd);
    if (s == FBCT_OK && !p->preweight.empty()) s = upload_vec(&t.preweight, p->preweight);
    if (s == FBCT_OK) s = upload_vec(&t.ramp_planes, p->ramp_planes);
    if (s == FBCT_OK && !p->parker.empty()) s = upload_vec(&t.parker, p->parker);
    if (s != FBCT_OK) {
        free_cone_tables(t);
        return s;
    }
    if (p->dev.capacity() == 0) p->dev.reserve(16);
    if (p->dev.size() == p->dev.capacity()) {
        free_cone_tables(t);
        return fail(FBCT_ERR_UNSUPPORTED, "cone plan used on more than %zu devices", p->dev.capacity());
    }
    p->dev.push_back(t);
    *out = &p->dev.back();
    return FBCT_OK;
}

size_t proj_elems(const fbct_cone_plan *p, int n)
{
    return (size_t)n * p->g.nViews * p->g.nRows * p->g.nCols;
}

fbct_status check_exec(const fbct_cone_plan *p, const void *in, const void *out, int n, const char *what)
{
    if (!p) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null plan", what);
    if (n < 0) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: negative volume count", what);
    if (n > 0 && (!in || !out)) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null data pointer", what);
    if ((long long)n * p->g.nZ > 65535LL * 16384)
        return fail(FBCT_ERR_UNSUPPORTED, "%s: too many volumes per call", what);
    return FBCT_OK;
}

fbct_status check_dtype(int d, const char *what)
{
    if (d != FBCT_F32 && d != FBCT_BF16) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: unknown storage type %d", what, d);
    return FBCT_OK;
}

size_t proj_bytes(const fbct_cone_plan *p, int n, int dtype)
{
    return (size_t)n * p->g.nViews * p->g.nRows * p->g.nCols * (dtype == FBCT_BF16 ? 2 : 4);
}

size_t vol_bytes(const fbct_cone_plan *p, int n, int dtype)
{
    return (size_t)n * p->g.nZ * p->g.nY * p->g.nX * (dtype == FBCT_BF16 ? 2 : 4);
}

fbct_status run_filter(fbct_cone_plan *p, const ConeDeviceTables &t, const void *proj, fbct_dtype ti, float *q, int n,
                       hipStream_t st)
{
    RampOperand H{t.ramp_planes, p->g.nCols, p->d.ramp_np, p->d.ramp_kp, p->d.ramp_lg};
    const int64_t rows = (int64_t)n * p->g.nViews * p->g.nRows;
    RowWeights rw;
    rw.w = t.preweight;  // [nRows][nCols] cosine weights (flat detector)
    rw.wrows = p->g.nRows;
    rw.w2 = t.parker;    // [nViews][nCols] Parker weights (short scan)
    rw.w2div = p->g.nRows;
    rw.w2rows = p->g.nViews;
    return launch_ramp_gemm(H, proj, ti, q, rows, rw, st);
}

}  // namespace
}  // namespace fbct

extern "C" {

fbct_status fbct_cone_geom_from_params(const float *parameter, int n_parameter, int n_rows, int n_z,
                                       fbct_cone_geom *out)
{
    using fbct::fail;
    if (!parameter || !out) return fail(FBCT_ERR_INVALID_ARGUMENT, "null pointer");
    if (n_rows < 1 || n_z < 1) return fail(FBCT_ERR_INVALID_GEOMETRY, "row and slice counts must be positive");
    fbct_fan_geom f;
    fbct_status s = fbct_geom_from_params(parameter, n_parameter, &f);
    if (s != FBCT_OK) return s;
    std::memset(out, 0, sizeof(*out));
    out->R = f.R;
    out->D = f.D;
    out->nViews = f.nViews;
    out->dAngle = f.dAngle;
    out->detector = FBCT_DETECTOR_CURVED;
    out->nCols = f.nCols;
    out->nRows = n_rows;
    out->dCol = f.dCol;
    out->dRow = f.dRow;
    out->colLeft = f.dLeft;
    out->rowBottom = f.detBottom;
    out->dShift = f.dShift;
    out->nX = f.nX;
    out->nY = f.nY;
    out->nZ = n_z;
    out->dx = f.dx;
    out->dy = f.dy;
    out->dz = f.dz;
    out->xmin = f.xmin;
    out->ymin = f.ymin;
    out->zmin = -0.5f * f.dz * (float)n_z;
    out->rFOV = f.rFOV;
    return fbct::validate(*out);
}

fbct_status fbct_cone_plan_create(const fbct_cone_geom *geom, const float *angles, int n_angles,
                                  fbct_cone_plan **out)
{
    using fbct::fail;
    if (!out) return fail(FBCT_ERR_INVALID_ARGUMENT, "null plan output pointer");
    *out = nullptr;
    if (!geom || !angles) return fail(FBCT_ERR_INVALID_ARGUMENT, "null geometry or angle list");
    fbct_status s = fbct::validate(*geom);
    if (s != FBCT_OK) return s;
    if (n_angles != geom->nViews)
        return fail(FBCT_ERR_SHAPE_MISMATCH, "angle list has %d entries but the geometry has %d views", n_angles,
                    geom->nViews);
    for (int v = 0; v < n_angles; ++v)
        if (!std::isfinite(angles[v])) return fail(FBCT_ERR_INVALID_GEOMETRY, "angle %d is not finite", v);
    fbct_cone_plan *p = new (std::nothrow) fbct_cone_plan();
    if (!p) return fail(FBCT_ERR_OUT_OF_MEMORY, "cone plan allocation");
    try {
        p->g = *geom;
        p->angles.assign(angles, angles + n_angles);
        fbct::build(*p);
        if (geom->short_scan && !fbct::build_parker(*p)) {
            delete p;
            return fail(FBCT_ERR_INVALID_GEOMETRY, "cone: short scan views do not cover 180 degrees plus the fan angle");
        }
    } catch (const std::bad_alloc &) {
        delete p;
        return fail(FBCT_ERR_OUT_OF_MEMORY, "cone plan tables");
    }
    *out = p;
    return FBCT_OK;
}

fbct_status fbct_cone_plan_destroy(fbct_cone_plan *plan)
{
    if (!plan) return FBCT_OK;
    for (fbct::ConeDeviceTables &t : plan->dev) fbct::free_cone_tables(t);
    delete plan;
    return FBCT_OK;
}

size_t fbct_cone_workspace_size(const fbct_cone_plan *plan, int n)
{
    if (!plan || n <= 0) return 0;
    return fbct::proj_elems(plan, n) * sizeof(float);
}

fbct_status fbct_cone_filter(fbct_cone_plan *plan, const void *proj, int proj_dtype, float *filtered, int n,
                             void *stream)
{
    fbct_status s = fbct::check_exec(plan, proj, filtered, n, "fbct_cone_filter");
    if (s != FBCT_OK || n == 0) return s;
    if ((s = fbct::check_dtype(proj_dtype, "fbct_cone_filter")) != FBCT_OK) return s;
    if ((s = fbct::check_disjoint(proj, fbct::proj_bytes(plan, n, proj_dtype), filtered,
                                  fbct::proj_bytes(plan, n, FBCT_F32), "fbct_cone_filter", "proj", "filtered")) != FBCT_OK)
        return s;
    const fbct::ConeDeviceTables *t = nullptr;
    if ((s = fbct::cone_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::run_filter(plan, *t, proj, static_cast<fbct_dtype>(proj_dtype), filtered, n,
                            static_cast<hipStream_t>(stream));
}

fbct_status fbct_cone_backproject(fbct_cone_plan *plan, const float *filtered, void *vol, int vol_dtype, int n,
                                  void *stream)
{
    fbct_status s = fbct::check_exec(plan, filtered, vol, n, "fbct_cone_backproject");
    if (s != FBCT_OK || n == 0) return s;
    if ((s = fbct::check_dtype(vol_dtype, "fbct_cone_backproject")) != FBCT_OK) return s;
    if ((s = fbct::check_disjoint(filtered, fbct::proj_bytes(plan, n, FBCT_F32), vol, fbct::vol_bytes(plan, n, vol_dtype),
                                  "fbct_cone_backproject", "filtered", "vol")) != FBCT_OK)
        return s;
    const fbct::ConeDeviceTables *t = nullptr;
    if ((s = fbct::cone_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_cone_bp(*plan, *t, filtered, vol, static_cast<fbct_dtype>(vol_dtype), n,
                                static_cast<hipStream_t>(stream));
}

fbct_status fbct_cone_fdk(fbct_cone_plan *plan, const void *proj, int proj_dtype, void *vol, int vol_dtype, int n,
                          void *workspace, size_t workspace_bytes, void *stream)
{
    fbct_status s = fbct::check_exec(plan, proj, vol, n, "fbct_cone_fdk");
    if (s != FBCT_OK || n == 0) return s;
    if ((s = fbct::check_dtype(proj_dtype, "fbct_cone_fdk")) != FBCT_OK ||
        (s = fbct::check_dtype(vol_dtype, "fbct_cone_fdk")) != FBCT_OK)
        return s;
    const size_t need = fbct_cone_workspace_size(plan, n);
    if (!workspace || workspace_bytes < need)
        return fbct::fail(FBCT_ERR_WORKSPACE, "fbct_cone_fdk: workspace of %zu bytes needed, got %zu", need,
                          workspace ? workspace_bytes : (size_t)0);
    if ((s = fbct::check_disjoint(proj, fbct::proj_bytes(plan, n, proj_dtype), vol, fbct::vol_bytes(plan, n, vol_dtype),
                                  "fbct_cone_fdk", "proj", "vol")) != FBCT_OK ||
        (s = fbct::check_disjoint(workspace, need, proj, fbct::proj_bytes(plan, n, proj_dtype), "fbct_cone_fdk",
                                  "workspace", "proj")) != FBCT_OK ||
        (s = fbct::check_disjoint(workspace, need, vol, fbct::vol_bytes(plan, n, vol_dtype), "fbct_cone_fdk",
                                  "workspace", "vol")) != FBCT_OK)
        return s;
    const fbct::ConeDeviceTables *t = nullptr;
    if ((s = fbct::cone_tables(plan, &t)) != FBCT_OK) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    float *q = static_cast<float *>(workspace);
    if ((s = fbct::run_filter(plan, *t, proj, static_cast<fbct_dtype>(proj_dtype), q, n, st)) != FBCT_OK)
        return s;
    return fbct::launch_cone_bp(*plan, *t, q, vol, static_cast<fbct_dtype>(vol_dtype), n, st);
}

}  // extern "C"

// capi.cpp -- execute entry points of the C ABI: argument validation and dispatch.
//
// The reference op wrappers (FP/transform.cc:93-117 and siblings) trusted their
// inputs (DCHECKs compile out, :97,105) and allocated / freed device scratch on
// every call (FP/transform.cu:198-217).  Here each call validates its arguments,
// allocates nothing, and enqueues its kernels on the caller's stream.

#include <hip/hip_runtime_api.h>

#include "fbct_internal.hpp"

namespace {

using fbct::fail;

fbct_status check_common(fbct_plan *plan, const void *in, const void *out, int64_t n, const char *what)
{
    if (!plan) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null plan", what);
    if (n < 0) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: negative slice count %lld", what, (long long)n);
    if (n > 0 && (!in || !out)) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: null data pointer", what);
    if (n > 65535) return fail(FBCT_ERR_UNSUPPORTED, "%s: at most 65535 slices per call", what);
    return FBCT_OK;
}

fbct_status check_dtype(int d, const char *what)
{
    if (d != FBCT_F32 && d != FBCT_BF16) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: unknown storage type %d", what, d);
    return FBCT_OK;
}

fbct_status filter_rows(const fbct_plan &p, const fbct::DeviceTables &t, const float *in, float *out, int64_t rows,
                        fbct_filter_mode mode, hipStream_t st)
{
    return mode == FBCT_FILTER_FP64 ? fbct::launch_ramp_filter(p, t, in, out, rows, st)
                                    : fbct::launch_ramp_filter_mfma(p, t, in, out, rows, st);
}

size_t sino_bytes(const fbct_plan *p, int64_t n, int dtype = FBCT_F32)
{
    return static_cast<size_t>(n) * p->g.nViews * p->g.nCols * (dtype == FBCT_BF16 ? 2 : 4);
}

size_t image_bytes(const fbct_plan *p, int64_t n, int dtype = FBCT_F32)
{
    return static_cast<size_t>(n) * p->g.nX * p->g.nY * (dtype == FBCT_BF16 ? 2 : 4);
}

// The image and sinogram buffers of one call must not overlap.
fbct_status check_io(const fbct_plan *p, const void *image, int image_dtype, const void *sino, int sino_dtype,
                     int64_t n, const char *what)
{
    return fbct::check_disjoint(image, image_bytes(p, n, image_dtype), sino, sino_bytes(p, n, sino_dtype), what,
                                "image", "sinogram");
}

// The FBP workspace (the filtered sinograms) must overlap neither the sinogram nor the image.
fbct_status check_workspace(const fbct_plan *p, const void *image, const void *sino, const void *ws, int64_t n,
                            const char *what)
{
    fbct_status s = fbct::check_disjoint(ws, sino_bytes(p, n), sino, sino_bytes(p, n), what, "workspace", "sinogram");
    if (s != FBCT_OK) return s;
    return fbct::check_disjoint(ws, sino_bytes(p, n), image, image_bytes(p, n), what, "workspace", "image");
}

}  // namespace

extern "C" {

fbct_status fbct_fp(fbct_plan *plan, const float *image, float *sino, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, image, sino, n_slices, "fbct_fp");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fp")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_fp(*plan, *t, image, FBCT_F32, sino, FBCT_F32, n_slices, static_cast<hipStream_t>(stream));
}

fbct_status fbct_fp_residual(fbct_plan *plan, const void *image, int image_dtype, const float *target,
                             void *residual, int residual_dtype, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, image, residual, n_slices, "fbct_fp_residual");
    if (s != FBCT_OK || n_slices == 0) return s;
    if (!target) return fbct::fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_fp_residual: null target");
    if ((s = check_dtype(image_dtype, "fbct_fp_residual")) != FBCT_OK ||
        (s = check_dtype(residual_dtype, "fbct_fp_residual")) != FBCT_OK)
        return s;
    if ((s = check_io(plan, image, image_dtype, residual, residual_dtype, n_slices, "fbct_fp_residual")) != FBCT_OK ||
        (s = fbct::check_disjoint(target, sino_bytes(plan, n_slices), residual, sino_bytes(plan, n_slices, residual_dtype),
                                  "fbct_fp_residual", "target", "residual")) != FBCT_OK)
        return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_fp(*plan, *t, image, static_cast<fbct_dtype>(image_dtype), residual,
                           static_cast<fbct_dtype>(residual_dtype), n_slices, static_cast<hipStream_t>(stream), target);
}

fbct_status fbct_fp_ex(fbct_plan *plan, const void *image, int image_dtype, void *sino, int sino_dtype,
                       int n_slices, void *stream)
{
    fbct_status s = check_common(plan, image, sino, n_slices, "fbct_fp_ex");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_dtype(image_dtype, "fbct_fp_ex")) != FBCT_OK || (s = check_dtype(sino_dtype, "fbct_fp_ex")) != FBCT_OK)
        return s;
    if ((s = check_io(plan, image, image_dtype, sino, sino_dtype, n_slices, "fbct_fp_ex")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_fp(*plan, *t, image, static_cast<fbct_dtype>(image_dtype), sino,
                           static_cast<fbct_dtype>(sino_dtype), n_slices, static_cast<hipStream_t>(stream));
}

fbct_status fbct_fp_adjoint(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, sino, image, n_slices, "fbct_fp_adjoint");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fp_adjoint")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_fp_adjoint(*plan, *t, sino, FBCT_F32, image, FBCT_F32, n_slices,
                                   static_cast<hipStream_t>(stream));
}

fbct_status fbct_fp_adjoint_ex(fbct_plan *plan, const void *sino, int sino_dtype, void *image, int image_dtype,
                               int n_slices, void *stream)
{
    fbct_status s = check_common(plan, sino, image, n_slices, "fbct_fp_adjoint_ex");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_dtype(sino_dtype, "fbct_fp_adjoint_ex")) != FBCT_OK ||
        (s = check_dtype(image_dtype, "fbct_fp_adjoint_ex")) != FBCT_OK)
        return s;
    if ((s = check_io(plan, image, image_dtype, sino, sino_dtype, n_slices, "fbct_fp_adjoint_ex")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_fp_adjoint(*plan, *t, sino, static_cast<fbct_dtype>(sino_dtype), image,
                                   static_cast<fbct_dtype>(image_dtype), n_slices, static_cast<hipStream_t>(stream));
}

fbct_status fbct_bp(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, sino, image, n_slices, "fbct_bp");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_bp")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_bp(*plan, *t, sino, image, n_slices, false, static_cast<hipStream_t>(stream));
}

fbct_status fbct_bp_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, image, sino, n_slices, "fbct_bp_adjoint");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_bp_adjoint")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_bp_adjoint(*plan, *t, image, sino, n_slices, false, static_cast<hipStream_t>(stream));
}

fbct_status fbct_fbp_backproject(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream)
{
    fbct_status s = check_common(plan, sino, image, n_slices, "fbct_fbp_backproject");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fbp_backproject")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_bp(*plan, *t, sino, image, n_slices, true, static_cast<hipStream_t>(stream));
}

fbct_status fbct_fbp_backproject_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices,
                                         void *stream)
{
    fbct_status s = check_common(plan, image, sino, n_slices, "fbct_fbp_backproject_adjoint");
    if (s != FBCT_OK || n_slices == 0) return s;
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fbp_backproject_adjoint")) != FBCT_OK) return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return fbct::launch_bp_adjoint(*plan, *t, image, sino, n_slices, true, static_cast<hipStream_t>(stream));
}

fbct_status fbct_ramp_filter_ex(fbct_plan *plan, const float *in, float *out, int64_t n_rows, int mode,
                                void *stream)
{
    if (!plan) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_ramp_filter: null plan");
    if (n_rows < 0) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_ramp_filter: negative row count");
    if (mode != FBCT_FILTER_MFMA && mode != FBCT_FILTER_FP64)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_ramp_filter: unknown filter mode %d", mode);
    if (n_rows == 0) return FBCT_OK;
    if (!in || !out) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_ramp_filter: null data pointer");
    const size_t bytes = static_cast<size_t>(n_rows) * plan->g.nCols * sizeof(float);
    const char *a = reinterpret_cast<const char *>(in), *b = reinterpret_cast<const char *>(out);
    if (a < b + bytes && b < a + bytes)
        return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_ramp_filter: in and out must not overlap");
    const fbct::DeviceTables *t = nullptr;
    fbct_status s;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    return filter_rows(*plan, *t, in, out, n_rows, static_cast<fbct_filter_mode>(mode),
                       static_cast<hipStream_t>(stream));
}

fbct_status fbct_ramp_filter(fbct_plan *plan, const float *in, float *out, int64_t n_rows, void *stream)
{
    return fbct_ramp_filter_ex(plan, in, out, n_rows, FBCT_FILTER_MFMA, stream);
}

size_t fbct_fbp_workspace_size(const fbct_plan *plan, int n_slices)
{
    if (!plan || n_slices <= 0) return 0;
    return sino_bytes(plan, n_slices);
}

fbct_status fbct_fbp_ex(fbct_plan *plan, const float *sino, float *image, int n_slices, void *workspace,
                        size_t workspace_bytes, unsigned flags, void *stream)
{
    fbct_status s = check_common(plan, sino, image, n_slices, "fbct_fbp_ex");
    if (s != FBCT_OK || n_slices == 0) return s;
    if (flags & ~(unsigned)(FBCT_FBP_COSINE | FBCT_FBP_SHORT_SCAN))
        return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_fbp_ex: unknown flags 0x%x", flags);
    if (!workspace || workspace_bytes < sino_bytes(plan, n_slices))
        return fail(FBCT_ERR_WORKSPACE, "fbct_fbp_ex: workspace of %zu bytes needed, got %zu",
                    sino_bytes(plan, n_slices), workspace ? workspace_bytes : (size_t)0);
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fbp_ex")) != FBCT_OK ||
        (s = check_workspace(plan, image, sino, workspace, n_slices, "fbct_fbp_ex")) != FBCT_OK)
        return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    float *q = static_cast<float *>(workspace);
    const int64_t rows = (int64_t)n_slices * plan->g.nViews;
    if (flags == 0) {
        if ((s = filter_rows(*plan, *t, sino, q, rows, FBCT_FILTER_MFMA, st)) != FBCT_OK) return s;
    } else {
        // pre-weights multiply the rows as the MFMA filter stages them (row r = slice * nViews + view)
        fbct::RowWeights rw;
        if (flags & FBCT_FBP_COSINE) {
            if ((s = fbct::device_cosine(plan, t, &rw.w)) != FBCT_OK) return s;
            rw.wrows = 1;
        }
        if (flags & FBCT_FBP_SHORT_SCAN) {
            if ((s = fbct::device_parker(plan, t, &rw.w2)) != FBCT_OK) return s;
            rw.w2div = 1;
            rw.w2rows = plan->g.nViews;
        }
        fbct::RampOperand H{t->ramp_planes, plan->g.nCols, plan->d.ramp_np, plan->d.ramp_kp, plan->d.ramp_lg};
        if ((s = fbct::launch_ramp_gemm(H, sino, FBCT_F32, q, rows, rw, st)) != FBCT_OK) return s;
    }
    return fbct::launch_bp(*plan, *t, q, image, n_slices, true, st);
}

fbct_status fbct_fbp(fbct_plan *plan, const float *sino, float *image, int n_slices, void *workspace,
                     size_t workspace_bytes, void *stream)
{
    return fbct_fbp_ex(plan, sino, image, n_slices, workspace, workspace_bytes, 0u, stream);
}

fbct_status fbct_fbp_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices, void *workspace,
                             size_t workspace_bytes, void *stream)
{
    fbct_status s = check_common(plan, image, sino, n_slices, "fbct_fbp_adjoint");
    if (s != FBCT_OK || n_slices == 0) return s;
    if (!workspace || workspace_bytes < sino_bytes(plan, n_slices))
        return fail(FBCT_ERR_WORKSPACE, "fbct_fbp_adjoint: workspace of %zu bytes needed, got %zu",
                    sino_bytes(plan, n_slices), workspace ? workspace_bytes : (size_t)0);
    if ((s = check_io(plan, image, FBCT_F32, sino, FBCT_F32, n_slices, "fbct_fbp_adjoint")) != FBCT_OK ||
        (s = check_workspace(plan, image, sino, workspace, n_slices, "fbct_fbp_adjoint")) != FBCT_OK)
        return s;
    const fbct::DeviceTables *t = nullptr;
    if ((s = fbct::device_tables(plan, &t)) != FBCT_OK) return s;
    hipStream_t st = static_cast<hipStream_t>(stream);
    float *q = static_cast<float *>(workspace);
    if ((s = fbct::launch_bp_adjoint(*plan, *t, image, q, n_slices, true, st)) != FBCT_OK) return s;
    // The centre-cropped convolution matrix H[c][m] = h[c-m] is symmetric (h[-n] = h[n]).
    return filter_rows(*plan, *t, q, sino, (int64_t)n_slices * plan->g.nViews, FBCT_FILTER_MFMA, st);
}

fbct_status fbct_fbp_short_scan(fbct_plan *plan, const float *sino, float *image, int n_slices, void *workspace,
                                size_t workspace_bytes, void *stream)
{
    return fbct_fbp_ex(plan, sino, image, n_slices, workspace, workspace_bytes, FBCT_FBP_SHORT_SCAN, stream);
}

fbct_status fbct_fdk(fbct_plan *plan, const float *proj, float *vol, int n_slices, void *workspace,
                     size_t workspace_bytes, void *stream)
{
    // nZ = nRows = 1 (FDK/FBP.cc:85,98): the z-loop of FDK/GPURecon_GPU.cu:41-47 runs
    // once at z = zmin + dz/2 = 0 and the single-row texture clamps the row index,
    // so the result is the FBP of each slice with the same dShift and filter scale.
    return fbct_fbp(plan, proj, vol, n_slices, workspace, workspace_bytes, stream);
}

}  // extern "C"

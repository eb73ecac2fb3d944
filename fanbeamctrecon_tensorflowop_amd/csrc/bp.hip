// bp.hip -- pixel-driven fan-beam backprojection (plain and 1/L^2-weighted) and its
// transpose (gfx950).
//
// Reference: BackProjOneView_fan_curve_kernel, BP/transform.cu:67-95 (unweighted) and
// FBP/transform.cu:61-89 (times 1/(U^2+t^2)), launched once per view with a memcpy
// and a device sync per view (BP/transform.cu:110-118).
//
// Here one launch covers every view and slice: a thread owns a pixel, loops over all
// views with the per-view (cos, sin) read as uniform scalars, accumulates in a
// register and stores once.  The detector row is read with two gathers per update;
// the 1-D texture's clamp-to-edge linear filter is reproduced with exact FP32 weights.

#include <hip/hip_runtime.h>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

struct BpArgs {
    const float *sino;
    float *img;
    const float2 *cs;
    long long img_stride, sino_stride;
    int nX, nY, nViews, nCols;
    float R, xmin, ymin, dx, dy, r2;
    float xi_scale, xi_offset;
};

constexpr int kBpTx = 16, kBpTy = 16;

// Column coordinate of the pixel in view (cb, sb) and, optionally, its 1/L^2 weight.
// xi_raw = (atan(t/U) - dLeft)/dCol + dShift, with U = R - x cos b - y sin b and
// t = x sin b - y cos b (BP/transform.cu:82-85 with the negated sin/cos of :116).
template <bool kWeighted>
__device__ __forceinline__ float bp_sample(const float *P, int nCols, float R, float x, float y, float cb,
                                           float sb, float xi_scale, float xi_offset)
{
    const float U = fmaf(-y, sb, fmaf(-x, cb, R));
    const float t = fmaf(x, sb, -y * cb);
    const float xi = fmaf(atanf(t * __builtin_amdgcn_rcpf(U)), xi_scale, xi_offset);
    if (!(xi > 0.0f && xi < (float)nCols)) return 0.0f;
    // texel-centre convention: interpolate at xi - 0.5; clamp to [0, nCols-1]
    const float ct = fminf(fmaxf(xi - 0.5f, 0.0f), (float)(nCols - 1));
    const float fc = fminf(floorf(ct), (float)max(nCols - 2, 0));
    const int c0 = (int)fc;
    const float al = ct - fc;
    const float p0 = P[c0];
    const float p1 = P[min(c0 + 1, nCols - 1)];
    const float val = fmaf(al, p1 - p0, p0);
    if (kWeighted) return val * __builtin_amdgcn_rcpf(fmaf(U, U, t * t));
    return val;
}

template <bool kWeighted>
__global__ __launch_bounds__(kBpTx *kBpTy) void bp_kernel(BpArgs A)
{
    const int i = blockIdx.x * kBpTx + threadIdx.x % kBpTx;
    const int j = blockIdx.y * kBpTy + threadIdx.x / kBpTx;
    const int sl = blockIdx.z;
    if (i >= A.nX || j >= A.nY) return;
    const float x = fmaf(A.dx, (float)i + 0.5f, A.xmin);
    const float y = fmaf(A.dy, (float)j + 0.5f, A.ymin);
    float acc = 0.0f;
    if (x * x + y * y <= A.r2) {
        const float *sino = A.sino + (size_t)sl * A.sino_stride;
        for (int v = 0; v < A.nViews; ++v) {
            const float2 cs = A.cs[v];
            acc += bp_sample<kWeighted>(sino + (size_t)v * A.nCols, A.nCols, A.R, x, y, cs.x, cs.y, A.xi_scale,
                                        A.xi_offset);
        }
    }
    A.img[(size_t)sl * A.img_stride + (size_t)j * A.nX + i] = acc;
}

// Transpose: a workgroup owns one detector row (view v, slice) in LDS, visits every
// pixel of the FOV and scatters the two linear-interpolation weights (clamped taps
// folded into the edge columns) with ds_add_f32; the row is stored once.
template <bool kWeighted>
__global__ __launch_bounds__(256) void bp_adjoint_kernel(BpArgs A)
{
    extern __shared__ float row[];
    const int v = blockIdx.x, sl = blockIdx.z;
    for (int c = threadIdx.x; c < A.nCols; c += 256) row[c] = 0.0f;
    __syncthreads();
    const float2 cs = A.cs[v];
    const float *img = A.img + (size_t)sl * A.img_stride;
    const int npix = A.nX * A.nY;
    for (int p = threadIdx.x; p < npix; p += 256) {
        const int i = p % A.nX, j = p / A.nX;
        const float x = fmaf(A.dx, (float)i + 0.5f, A.xmin);
        const float y = fmaf(A.dy, (float)j + 0.5f, A.ymin);
        if (!(x * x + y * y <= A.r2)) continue;
        const float val0 = img[p];
        if (val0 == 0.0f) continue;
        const float U = fmaf(-y, cs.y, fmaf(-x, cs.x, A.R));
        const float t = fmaf(x, cs.y, -y * cs.x);
        const float xi = fmaf(atanf(t * __builtin_amdgcn_rcpf(U)), A.xi_scale, A.xi_offset);
        if (!(xi > 0.0f && xi < (float)A.nCols)) continue;
        const float ct = fminf(fmaxf(xi - 0.5f, 0.0f), (float)(A.nCols - 1));
        const float fc = fminf(floorf(ct), (float)max(A.nCols - 2, 0));
        const int c0 = (int)fc;
        const float al = ct - fc;
        const float val = kWeighted ? val0 * __builtin_amdgcn_rcpf(fmaf(U, U, t * t)) : val0;
        const float w1 = val * al;
        atomicAdd(&row[c0], val - w1);
        atomicAdd(&row[min(c0 + 1, A.nCols - 1)], w1);
    }
    __syncthreads();
    float *out = const_cast<float *>(A.sino) + (size_t)sl * A.sino_stride + (size_t)v * A.nCols;
    for (int c = threadIdx.x; c < A.nCols; c += 256) out[c] = row[c];
}

BpArgs make_args(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img)
{
    BpArgs A;
    A.sino = sino;
    A.img = img;
    A.cs = t.view_cs;
    A.img_stride = (long long)p.g.nX * p.g.nY;
    A.sino_stride = (long long)p.g.nViews * p.g.nCols;
    A.nX = p.g.nX;
    A.nY = p.g.nY;
    A.nViews = p.g.nViews;
    A.nCols = p.g.nCols;
    A.R = p.g.R;
    A.xmin = p.g.xmin;
    A.ymin = p.g.ymin;
    A.dx = p.g.dx;
    A.dy = p.g.dy;
    A.r2 = p.d.r2;
    A.xi_scale = p.d.bp_xi_scale;
    A.xi_offset = p.d.bp_xi_offset;
    return A;
}

}  // namespace

fbct_status launch_bp(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img, int n,
                      bool weighted, hipStream_t stream)
{
    BpArgs A = make_args(p, t, sino, img);
    dim3 grid((p.g.nX + kBpTx - 1) / kBpTx, (p.g.nY + kBpTy - 1) / kBpTy, n);
    if (weighted)
        hipLaunchKernelGGL(bp_kernel<true>, grid, dim3(kBpTx * kBpTy), 0, stream, A);
    else
        hipLaunchKernelGGL(bp_kernel<false>, grid, dim3(kBpTx * kBpTy), 0, stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "bp_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

fbct_status launch_bp_adjoint(const fbct_plan &p, const DeviceTables &t, const float *img, float *sino, int n,
                              bool weighted, hipStream_t stream)
{
    // bp_adjoint_kernel reads the image through A.img and writes the sinogram through A.sino.
    BpArgs A = make_args(p, t, sino, const_cast<float *>(img));
    const size_t lds = (size_t)p.g.nCols * sizeof(float);
    if (lds > 64 * 1024) return fail(FBCT_ERR_UNSUPPORTED, "fbct_bp_adjoint: detector row too wide for LDS");
    dim3 grid(p.g.nViews, 1, n);
    if (weighted)
        hipLaunchKernelGGL(bp_adjoint_kernel<true>, grid, dim3(256), lds, stream, A);
    else
        hipLaunchKernelGGL(bp_adjoint_kernel<false>, grid, dim3(256), lds, stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "bp_adjoint_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

// fp.hip -- ray-driven fan-beam forward projection and its matched adjoint (gfx950).
//
// Reference: GenerateRayTraceMap_FB_ReconImage_OneView_kernel (FP/transform.cu:64-101)
// builds per-ray start/step/N; RayTrace_ReconImage_OneView_kernel (:103-146) sums
// N bilinear texture fetches per ray at 0.125-pixel steps and scales by 0.125*dx.
//
// Design here:
//   * per-ray (u0, w0, du, dw, N) come from the plan's FP64 table;
//   * a workgroup owns a bundle of kFpRays adjacent rays of one view and marches them
//     slab by slab along the axis the bundle is closest to;  for every slab it stages
//     the exact bounding box of the taps its samples touch into LDS (clamped to the
//     image edge, which reproduces the clamp-to-edge texture address mode) and
//     interpolates from LDS with exact FP32 bilinear weights;
//   * kFpLanesPerRay lanes share a ray, each keeps its own FP32 partial sum across all
//     slabs; one cross-lane reduction per ray at the end and one coalesced store.
// A sample belongs to the slab containing its major coordinate, decided from the same
// FP32 position the interpolation uses, so every sample is counted exactly once.

#include <hip/hip_runtime.h>

#include <climits>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kPitch = kFpSlab + 1;  // LDS row: taps [alo, alo + kFpSlab]

struct FpArgs {
    const float *img;
    float *sino;
    const float4 *ray_geo;
    const int32_t *ray_n;
    const uint8_t *major;
    long long img_stride;
    long long sino_stride;
    int nX, nY, nViews, nCols, nb;
    int nslab, cap;
    float scale;
};

// Smallest k in [0, N] with P(k) true, where P(k) = (a0 + k da >= t) for da >= 0 and
// (a0 + k da < t) for da < 0, evaluated with the same fmaf the samples use (monotone
// in k, so P is false...false true...true).
__device__ __forceinline__ int slab_boundary(float a0, float da, int N, float t)
{
    if (N <= 0) return 0;
    const bool asc = da >= 0.0f;
    float est;
    if (da == 0.0f) {
        est = (a0 >= t) ? 0.0f : (float)N;
    } else {
        const float q = (t - a0) / da;
        est = asc ? ceilf(q) : floorf(q) + 1.0f;
    }
    est = fminf(fmaxf(est, 0.0f), (float)N);
    int k = (int)est;
    auto pred = [&](int kk) {
        const float a = fmaf((float)kk, da, a0);
        return asc ? (a >= t) : (a < t);
    };
    while (k > 0 && pred(k - 1)) --k;
    while (k < N && !pred(k)) ++k;
    return k;
}

__device__ __forceinline__ int wave_min(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

__global__ __launch_bounds__(kFpThreads) void fp_kernel(FpArgs A, int *err)
{
    extern __shared__ float tile[];
    __shared__ int red[2][2][kFpThreads / 64];

    const int tid = threadIdx.x;
    const int lane = tid % kFpLanesPerRay;
    const int b = blockIdx.x, v = blockIdx.y, sl = blockIdx.z;
    const int c = b * kFpRays + tid / kFpLanesPerRay;
    const bool xmajor = A.major[(size_t)v * A.nb + b] != 0;

    float a0 = 0.0f, m0 = 0.0f, da = 0.0f, dm = 0.0f;
    int N = 0;
    if (c < A.nCols) {
        const size_t r = (size_t)v * A.nCols + c;
        const float4 g = A.ray_geo[r];
        N = A.ray_n[r];
        a0 = xmajor ? g.x : g.y;
        m0 = xmajor ? g.y : g.x;
        da = xmajor ? g.z : g.w;
        dm = xmajor ? g.w : g.z;
    }
    const bool asc = da >= 0.0f;
    const float *img = A.img + (size_t)sl * A.img_stride;
    const int na = xmajor ? A.nX : A.nY;  // image extent along the major axis
    const int nm = xmajor ? A.nY : A.nX;

    float acc = 0.0f;
    int kprev = slab_boundary(a0, da, N, -1.0f);
    const int wave = tid / 64;

    for (int s = 0; s < A.nslab; ++s) {
        const int alo = s * kFpSlab - 1;
        const int knext = slab_boundary(a0, da, N, (float)(alo + kFpSlab));
        const int ka = asc ? kprev : knext;
        const int kb = asc ? knext : kprev;
        kprev = knext;

        int mn = INT_MAX, mx = INT_MIN;
        if (ka < kb) {
            const float m1 = fmaf((float)ka, dm, m0);
            const float m2 = fmaf((float)(kb - 1), dm, m0);
            mn = (int)floorf(fminf(m1, m2));
            mx = (int)floorf(fmaxf(m1, m2)) + 1;
        }
        mn = wave_min(mn);
        mx = wave_max(mx);
        if ((tid & 63) == 0) {
            red[s & 1][0][wave] = mn;
            red[s & 1][1][wave] = mx;
        }
        __syncthreads();  // reductions visible; previous slab's reads of `tile` done
        int mlo = red[s & 1][0][0], mhi = red[s & 1][1][0];
#pragma unroll
        for (int w = 1; w < kFpThreads / 64; ++w) {
            mlo = min(mlo, red[s & 1][0][w]);
            mhi = max(mhi, red[s & 1][1][w]);
        }
        if (mlo > mhi) continue;  // no samples of this bundle in the slab (uniform)
        const int M = mhi - mlo + 1;
        if (M > A.cap) {  // cannot happen for tables built by the plan; never read past LDS
            if (tid == 0) atomicOr(err, 1);
            continue;
        }

        // Stage taps [alo, alo+kFpSlab] x [mlo, mhi] with clamp-to-edge addressing.
        // Global loads run along the image row (i) in both orientations.
        const int F = xmajor ? kPitch : M;  // extent of the fast (row-contiguous) index
        const int total = F * (xmajor ? M : kPitch);
        const float invF = 1.0f / (float)F;
        for (int idx = tid; idx < total; idx += kFpThreads) {
            const int sidx = (int)(((float)idx + 0.5f) * invF);
            const int fidx = idx - sidx * F;
            int aa, mm;
            if (xmajor) {
                aa = fidx;
                mm = sidx;
            } else {
                mm = fidx;
                aa = sidx;
            }
            const int ga = min(max(alo + aa, 0), na - 1);
            const int gm = min(max(mlo + mm, 0), nm - 1);
            const int gi = xmajor ? ga : gm;
            const int gj = xmajor ? gm : ga;
            tile[mm * kPitch + aa] = img[(size_t)gj * A.nX + gi];
        }
        __syncthreads();  // tile ready

        for (int k = ka + lane; k < kb; k += kFpLanesPerRay) {
            const float kf = (float)k;
            const float a = fmaf(kf, da, a0);
            const float m = fmaf(kf, dm, m0);
            const float fa = floorf(a), fm = floorf(m);
            const float ta = a - fa, tm = m - fm;
            const int o = ((int)fm - mlo) * kPitch + ((int)fa - alo);
            const float v00 = tile[o], v01 = tile[o + 1];
            const float v10 = tile[o + kPitch], v11 = tile[o + kPitch + 1];
            const float top = fmaf(ta, v01 - v00, v00);
            const float bot = fmaf(ta, v11 - v10, v10);
            acc += fmaf(tm, bot - top, top);
        }
    }

#pragma unroll
    for (int off = kFpLanesPerRay / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0 && c < A.nCols) A.sino[(size_t)sl * A.sino_stride + (size_t)v * A.nCols + c] = acc * A.scale;
}

__device__ int g_fp_err = 0;

}  // namespace

fbct_status launch_fp(const fbct_plan &p, const DeviceTables &t, const float *img, float *sino, int n,
                      hipStream_t stream)
{
    FpArgs A;
    A.img = img;
    A.sino = sino;
    A.ray_geo = t.ray_geo;
    A.ray_n = t.ray_n;
    A.major = t.fp_major;
    A.img_stride = (long long)p.g.nX * p.g.nY;
    A.sino_stride = (long long)p.g.nViews * p.g.nCols;
    A.nX = p.g.nX;
    A.nY = p.g.nY;
    A.nViews = p.g.nViews;
    A.nCols = p.g.nCols;
    A.nb = p.d.fp_nbundles;
    A.nslab = p.d.fp_nslab;
    A.cap = p.d.fp_cap;
    A.scale = p.d.fp_scale;
    const size_t lds = (size_t)kPitch * p.d.fp_cap * sizeof(float);
    if (lds > 160 * 1024 - 1024)
        return fail(FBCT_ERR_UNSUPPORTED, "fbct_fp: slab tile of %zu bytes exceeds LDS", lds);
    int *err = nullptr;
    hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&err), HIP_SYMBOL(g_fp_err));
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipGetSymbolAddress: %s", hipGetErrorString(e));
    dim3 grid(p.d.fp_nbundles, p.g.nViews, n);
    hipLaunchKernelGGL(fp_kernel, grid, dim3(kFpThreads), lds, stream, A, err);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

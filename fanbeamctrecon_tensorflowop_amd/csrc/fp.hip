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

// ------------------------------------------------------------------ adjoint
//
// Exact transpose of fp_kernel: every FP sample (v, c, k) adds sino[v][c]*0.125*dx
// times its four bilinear weights into the four (clamped) taps it reads.  Output
// stationary: a workgroup owns a kFpaTile^2 block of output pixels held in LDS with a
// one-tap halo, visits a group of views, and for each view the candidate rays of the
// plan table.  A ray's samples that touch the tile are those with
// floor(u) in [ti0-1, ti0+T-1] and floor(w) in [tj0-1, tj0+T-1], found with the same
// FP32 positions and boundary search as the forward kernel.  Lanes accumulate into LDS
// (ds_add_f32); the tile is flushed once per workgroup.  Halo taps belong to the
// neighbouring tile, which visits the same samples, except taps -1 and n (the
// clamp-to-edge taps), which are folded into the border pixels.

constexpr int kFpaLanesPerRay = 8;
constexpr int kFpaRaysPerPass = kFpThreads / kFpaLanesPerRay;
constexpr int kFpaSide = kFpaTile + 2;

struct FpaArgs {
    const float *sino;
    float *img;
    const float4 *ray_geo;
    const int32_t *ray_n;
    const int32_t *cols;
    long long img_stride;
    long long sino_stride;
    int nX, nY, nViews, nCols, tiles_x, views_per_group;
    float scale;
};

// Half-open sample range [k0, k1) of a ray whose coordinate x_k = x0 + k dx lies in [lo, hi).
__device__ __forceinline__ void axis_range(float x0, float dx, int N, float lo, float hi, int &k0, int &k1)
{
    const int bl = slab_boundary(x0, dx, N, lo);
    const int bh = slab_boundary(x0, dx, N, hi);
    if (dx >= 0.0f) {
        k0 = bl;
        k1 = bh;
    } else {
        k0 = bh;
        k1 = bl;
    }
}

__global__ __launch_bounds__(kFpThreads) void fp_adjoint_kernel(FpaArgs A)
{
    __shared__ float acc[kFpaSide * kFpaSide];
    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    const int tx = tile % A.tiles_x, ty = tile / A.tiles_x;
    const int ti0 = tx * kFpaTile, tj0 = ty * kFpaTile;
    const int tw = min(kFpaTile, A.nX - ti0), th = min(kFpaTile, A.nY - tj0);
    const int sl = blockIdx.z;
    const int v0 = blockIdx.y * A.views_per_group;
    const int v1 = min(A.nViews, v0 + A.views_per_group);

    for (int i = tid; i < kFpaSide * kFpaSide; i += kFpThreads) acc[i] = 0.0f;
    __syncthreads();

    const float *sino = A.sino + (size_t)sl * A.sino_stride;
    const float ulo = (float)(ti0 - 1), uhi = (float)(ti0 + tw);
    const float wlo = (float)(tj0 - 1), whi = (float)(tj0 + th);
    const int g = tid % kFpaLanesPerRay;

    for (int v = v0; v < v1; ++v) {
        const int *cr = A.cols + ((size_t)tile * A.nViews + v) * 2;
        const int clo = cr[0], chi = cr[1];
        for (int cbase = clo; cbase <= chi; cbase += kFpaRaysPerPass) {
            const int c = cbase + tid / kFpaLanesPerRay;
            if (c > chi) continue;
            const size_t r = (size_t)v * A.nCols + c;
            const int N = A.ray_n[r];
            if (N <= 0) continue;
            const float val = sino[(size_t)v * A.nCols + c] * A.scale;
            if (val == 0.0f) continue;
            const float4 gg = A.ray_geo[r];
            int ku0, ku1, kw0, kw1;
            axis_range(gg.x, gg.z, N, ulo, uhi, ku0, ku1);
            axis_range(gg.y, gg.w, N, wlo, whi, kw0, kw1);
            const int ka = max(ku0, kw0), kb = min(ku1, kw1);
            if (ka >= kb) continue;
            // contiguous segment per lane keeps the 8 lanes of a ray on different taps
            const int len = (kb - ka + kFpaLanesPerRay - 1) / kFpaLanesPerRay;
            const int k0 = ka + g * len, k1 = min(kb, k0 + len);
            for (int k = k0; k < k1; ++k) {
                const float kf = (float)k;
                const float u = fmaf(kf, gg.z, gg.x);
                const float w = fmaf(kf, gg.w, gg.y);
                const float fu = floorf(u), fw = floorf(w);
                const float tu = u - fu, tw2 = w - fw;
                const int o = ((int)fw - (tj0 - 1)) * kFpaSide + ((int)fu - (ti0 - 1));
                const float p1 = val * tw2, p0 = val - p1;
                const float w01 = p0 * tu, w11 = p1 * tu;
                atomicAdd(&acc[o], p0 - w01);
                atomicAdd(&acc[o + 1], w01);
                atomicAdd(&acc[o + kFpaSide], p1 - w11);
                atomicAdd(&acc[o + kFpaSide + 1], w11);
            }
        }
    }
    __syncthreads();

    // Flush: interior taps, plus the clamp-to-edge taps -1 and n into the border pixels.
    float *img = A.img + (size_t)sl * A.img_stride;
    for (int idx = tid; idx < kFpaSide * kFpaSide; idx += kFpThreads) {
        const int lx = idx % kFpaSide, ly = idx / kFpaSide;
        const int gi = ti0 - 1 + lx, gj = tj0 - 1 + ly;
        const bool okx = (gi >= ti0 && gi < ti0 + tw) || gi == -1 || gi == A.nX;
        const bool oky = (gj >= tj0 && gj < tj0 + th) || gj == -1 || gj == A.nY;
        if (!okx || !oky) continue;
        const float a = acc[idx];
        if (a == 0.0f) continue;
        const int pi = min(max(gi, 0), A.nX - 1), pj = min(max(gj, 0), A.nY - 1);
        atomicAdd(&img[(size_t)pj * A.nX + pi], a);
    }
}

__global__ void zero_kernel(float *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0f;
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

fbct_status launch_fp_adjoint(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img, int n,
                              hipStream_t stream)
{
    const long long npix = (long long)p.g.nX * p.g.nY;
    const long long total = npix * n;
    const int zb = (int)std::min<long long>(4096, (total + 255) / 256);
    hipLaunchKernelGGL(zero_kernel, dim3(zb), dim3(256), 0, stream, img, total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "zero_kernel launch: %s", hipGetErrorString(e));

    FpaArgs A;
    A.sino = sino;
    A.img = img;
    A.ray_geo = t.ray_geo;
    A.ray_n = t.ray_n;
    A.cols = t.fpa_cols;
    A.img_stride = npix;
    A.sino_stride = (long long)p.g.nViews * p.g.nCols;
    A.nX = p.g.nX;
    A.nY = p.g.nY;
    A.nViews = p.g.nViews;
    A.nCols = p.g.nCols;
    A.tiles_x = p.d.fpa_tiles_x;
    A.scale = p.d.fp_scale;
    const int ntiles = p.d.fpa_tiles_x * p.d.fpa_tiles_y;
    // enough workgroups to fill 256 CUs several times over
    long long want = 4096 / std::max<long long>(1, (long long)ntiles * n);
    int groups = (int)std::max<long long>(1, std::min<long long>(want, p.g.nViews));
    A.views_per_group = (p.g.nViews + groups - 1) / groups;
    groups = (p.g.nViews + A.views_per_group - 1) / A.views_per_group;
    dim3 grid(ntiles, groups, n);
    hipLaunchKernelGGL(fp_adjoint_kernel, grid, dim3(kFpThreads), 0, stream, A);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_adjoint_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

// fp_adjoint.hip -- matched adjoint of the ray-driven FP, gathered per pixel (gfx950).
//
// The FP (fp.hip; reference FP/transform.cu:103-146) forms
//     p[v][c] = 0.125 dx * sum_k  sum_taps  w_tap(u_k, w_k) * img[clamp(tap)]
// with u_k = fma(k, du, u0), w_k = fma(k, dw, w0) in FP32.  Its transpose is
//     img[j][i] = 0.125 dx * sum_v sum_c p[v][c] * W_c(i, j),
//     W_c(i, j) = sum_k  T_i(u_k) * T_j(w_k),
// where T_i(u) = max(0, 1 - |clamp(u, 0, nX-1) - i|) is the total bilinear weight the
// clamped taps of a sample give pixel i (the clamp folds the taps -1 and nX, and any
// sample beyond the image edge, into the border pixels).
//
// Each thread owns one output pixel and accumulates over all views, so every pixel is
// written exactly once and no atomics are needed.  Per view, the candidate rays are
// the columns whose lines pass within the support of the pixel (radius rho around its
// centre, rho = sqrt(2) inside the image), found from the pixel's fan angle; per
// candidate ray the sample range whose positions fall in the support is computed in
// closed form and the samples are evaluated at exactly the FP32 positions the forward
// kernel interpolates at.  A group of S <= 16 slices that share the geometry is
// processed together: the per-(pixel, view, ray) weight W_c is computed once and applied
// to S sinograms.  Sums over views accumulate in FP64.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "dtype.hpp"
#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kAdjThreads = 256;
constexpr int kAdjTx = 16, kAdjTy = kAdjThreads / kAdjTx;  // tile of the regular launches
constexpr float kBig = 1e30f;

struct FpaArgs {
    const void *sino;  // TI[n][nViews][nCols]
    void *img;         // TO[n][nY][nX]
    const float4 *ray_geo;
    const int32_t *ray_n;
    const float2 *ray_inv;  // (1/du, 1/dw), 1e30 for a zero step
    const float2 *cs;       // (cos beta, sin beta)
    long long img_stride, sino_stride;
    int nX, nY, nViews, nCols, nSlices;
    float scale;            // 0.125 dx
    float r;                // rFOV/dx: FP texel coordinate of the image centre is r - 1/2
    float dx, R;
    float c_scale, c_off;   // column coordinate c_f = atan(t/U) * c_scale + c_off (ray c at c_f = c)
    float c_abs;            // |1/dCol|: radians -> columns
    float far_x, far_y;     // how far samples reach beyond the last pixel centre (texels)
};

// Per-pixel accumulation over all views.  kBorder: the workgroup's tile touches the
// image edge, so the clamp window and the unbounded outer support of border pixels
// apply; interior tiles skip them.
template <int S, bool kBorder, typename TI, typename TO>
__device__ __forceinline__ void adjoint_pixel(const FpaArgs &A, int i, int j, int s0)
{
    const float fi = (float)i, fj = (float)j;
    // pixel centre in mm, in the FP's frame (dx on both axes, centre at texel r - 1/2)
    const float x = (fi + 0.5f - A.r) * A.dx;
    const float y = (fj + 0.5f - A.r) * A.dx;
    float xlo_c = -kBig, xhi_c = kBig, ylo_c = -kBig, yhi_c = kBig;
    float xlo_b = -1.0f, xhi_b = 1.0f, ylo_b = -1.0f, yhi_b = 1.0f;
    float rho = 1.41421356f;
    if (kBorder) {
        // clamp window of the sample coordinate relative to this pixel (border plateaus)
        // and support of T_i in (u - i), unbounded on the outer side of border pixels
        if (i == 0) xlo_c = 0.0f, xlo_b = -kBig;
        if (i == A.nX - 1) xhi_c = 0.0f, xhi_b = kBig;
        if (j == 0) ylo_c = 0.0f, ylo_b = -kBig;
        if (j == A.nY - 1) yhi_c = 0.0f, yhi_b = kBig;
        const float ex = (i == A.nX - 1) ? fmaxf(1.0f, A.far_x) : 1.0f;
        const float ey = (j == A.nY - 1) ? fmaxf(1.0f, A.far_y) : 1.0f;
        rho = sqrtf(ex * ex + ey * ey);
    }

    const TI *sino[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        sino[s] = static_cast<const TI *>(A.sino) + (size_t)min(s0 + s, A.nSlices - 1) * A.sino_stride;

    double acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0;

    for (int v = 0; v < A.nViews; ++v) {
        const float2 cs = A.cs[v];
        const float U = fmaf(-y, cs.y, fmaf(-x, cs.x, A.R));
        const float t = fmaf(x, cs.y, -y * cs.x);
        const float rU = __builtin_amdgcn_rcpf(U);
        const float cf = fmaf(atanf(t * rU), A.c_scale, A.c_off);
        // Rays within distance rho (texels) of the centre subtend at most asin(rho dx / L)
        // from the pixel's fan angle; L >= U and asin(z) <= z (1 + z^2) for z <= 0.9.
        const float z = rho * A.dx * rU;
        const float hw = z < 0.9f ? fmaf(z * fmaf(z, z, 1.0f), A.c_abs, 1e-3f) : 1e9f;
        const int clo = (int)fmaxf(ceilf(cf - hw), 0.0f);
        const int chi = (int)fminf(floorf(cf + hw), (float)(A.nCols - 1));
        float vsum[S];
#pragma unroll
        for (int s = 0; s < S; ++s) vsum[s] = 0.0f;
        const size_t rbase = (size_t)v * A.nCols;
        for (int c = clo; c <= chi; ++c) {
            // The S sinogram values are loaded first, unconditionally, so their latency
            // overlaps the weight computation (a ray that misses the support, or has no
            // samples, ends with wsum = 0; sinograms are expected to be finite).
            const int N = A.ray_n[rbase + c];
            const float4 g = A.ray_geo[rbase + c];
            const float2 inv = A.ray_inv[rbase + c];
            float sv[S];
#pragma unroll
            for (int s = 0; s < S; ++s) sv[s] = load_f(sino[s], rbase + c);
            const float X0 = g.x - fi, Y0 = g.y - fj;
            const float a1 = (xlo_b - X0) * inv.x, a2 = (xhi_b - X0) * inv.x;
            const float b1 = (ylo_b - Y0) * inv.y, b2 = (yhi_b - Y0) * inv.y;
            const float lo = fmaxf(fmaxf(fminf(a1, a2), fminf(b1, b2)), 0.0f);
            const float hi = fminf(fminf(fmaxf(a1, a2), fmaxf(b1, b2)), (float)(N - 1));
            float wsum = 0.0f;
            const float kend = floorf(hi);
            for (float kf = ceilf(lo); kf <= kend; kf += 1.0f) {
                const float X = fmaf(kf, g.z, X0);
                const float Y = fmaf(kf, g.w, Y0);
                if (kBorder) {
                    const float wx = 1.0f - fabsf(fminf(fmaxf(X, xlo_c), xhi_c));
                    const float wy = 1.0f - fabsf(fminf(fmaxf(Y, ylo_c), yhi_c));
                    wsum = fmaf(fmaxf(wx, 0.0f), fmaxf(wy, 0.0f), wsum);
                } else {
                    // inside the support |X|, |Y| < 1 up to rounding at its two ends
                    wsum = fmaf(1.0f - fabsf(X), 1.0f - fabsf(Y), wsum);
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) vsum[s] = fmaf(wsum, sv[s], vsum[s]);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] += (double)vsum[s];
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (s0 + s < A.nSlices)
            store_f(static_cast<TO *>(A.img), (size_t)(s0 + s) * A.img_stride + (size_t)j * A.nX + i,
                    (float)(acc[s] * (double)A.scale));
}

// Pixel tile (tx, ty) of TX x (256/TX) pixels for slices [s0, s0+S).
template <int S, int TX, typename TI, typename TO>
__device__ __forceinline__ void adjoint_tile(const FpaArgs &A, int tx, int ty, int s0)
{
    constexpr int TY = kAdjThreads / TX;
    const int i = tx * TX + threadIdx.x % TX;
    const int j = ty * TY + threadIdx.x / TX;
    if (i >= A.nX || j >= A.nY) return;
    const bool border_tile = tx == 0 || ty == 0 || (tx + 1) * TX >= A.nX || (ty + 1) * TY >= A.nY;
    if (border_tile)
        adjoint_pixel<S, true, TI, TO>(A, i, j, s0);
    else
        adjoint_pixel<S, false, TI, TO>(A, i, j, s0);
}

template <int S, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_kernel(FpaArgs A)
{
    adjoint_tile<S, kAdjTx, TI, TO>(A, blockIdx.x, blockIdx.y, blockIdx.z * S);
}

// Uneven batch: slice s has its own geometry (desc[s]); work item = (slice, pixel tile),
// items of slice s are [prefix[s], prefix[s+1]).  A fixed grid strides over all items.
template <int TX, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_batch_kernel(const FpaSliceDesc *__restrict__ desc,
                                                                       const int *__restrict__ prefix, int n,
                                                                       int total, const TI *sino, TO *img)
{
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = n - 1;  // last slice with prefix[s] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const FpaSliceDesc &d = desc[lo];
        FpaArgs A;
        A.sino = sino + d.sino_off;
        A.img = img + d.img_off;
        A.ray_geo = d.ray_geo;
        A.ray_n = d.ray_n;
        A.ray_inv = d.ray_inv;
        A.cs = d.cs;
        A.img_stride = 0;
        A.sino_stride = 0;
        A.nX = d.nX;
        A.nY = d.nY;
        A.nViews = d.nViews;
        A.nCols = d.nCols;
        A.nSlices = 1;
        A.scale = d.scale;
        A.r = d.r;
        A.dx = d.dx;
        A.R = d.R;
        A.c_scale = d.c_scale;
        A.c_off = d.c_off;
        A.c_abs = d.c_abs;
        A.far_x = d.far_x;
        A.far_y = d.far_y;
        const int local = item - prefix[lo];
        const int ntx = (A.nX + TX - 1) / TX;
        adjoint_tile<1, TX, TI, TO>(A, local % ntx, local / ntx, 0);
    }
}

}  // namespace

template <int S>
static fbct_status launch_fpa_t(fbct_dtype ti, fbct_dtype to, dim3 grid, hipStream_t stream, const FpaArgs &A)
{
    const dim3 block(kAdjThreads);
    if (ti == FBCT_F32 && to == FBCT_F32)
        hipLaunchKernelGGL((fp_adjoint_kernel<S, float, float>), grid, block, 0, stream, A);
    else if (ti == FBCT_BF16 && to == FBCT_F32)
        hipLaunchKernelGGL((fp_adjoint_kernel<S, bf16_t, float>), grid, block, 0, stream, A);
    else if (ti == FBCT_F32 && to == FBCT_BF16)
        hipLaunchKernelGGL((fp_adjoint_kernel<S, float, bf16_t>), grid, block, 0, stream, A);
    else if (ti == FBCT_BF16 && to == FBCT_BF16)
        hipLaunchKernelGGL((fp_adjoint_kernel<S, bf16_t, bf16_t>), grid, block, 0, stream, A);
    else
        return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_fp_adjoint: unknown storage type");
    return FBCT_OK;
}

fbct_status launch_fp_adjoint(const fbct_plan &p, const DeviceTables &t, const void *sino, fbct_dtype ti,
                              void *img, fbct_dtype to, int n, hipStream_t stream)
{
    FpaArgs A;
    A.sino = sino;
    A.img = img;
    A.ray_geo = t.ray_geo;
    A.ray_n = t.ray_n;
    A.ray_inv = t.ray_inv;
    A.cs = t.view_cs;
    A.img_stride = (long long)p.g.nX * p.g.nY;
    A.sino_stride = (long long)p.g.nViews * p.g.nCols;
    A.nX = p.g.nX;
    A.nY = p.g.nY;
    A.nViews = p.g.nViews;
    A.nCols = p.g.nCols;
    A.nSlices = n;
    A.scale = p.d.fp_scale;
    A.r = p.d.fpa_r;
    A.dx = p.g.dx;
    A.R = p.g.R;
    A.c_scale = p.d.fpa_c_scale;
    A.c_off = p.d.fpa_c_off;
    A.c_abs = fabsf(p.d.fpa_c_scale);
    A.far_x = p.d.fpa_far_x;
    A.far_y = p.d.fpa_far_y;
    const unsigned gx = (p.g.nX + kAdjTx - 1) / kAdjTx, gy = (p.g.nY + kAdjTy - 1) / kAdjTy;
    // Slices that share the geometry are processed together when there are enough
    // workgroups left to fill the device.
    const long long tiles = (long long)gx * gy;
    fbct_status st;
    if (n >= 16 && tiles * ((n + 15) / 16) >= 1024)
        st = launch_fpa_t<16>(ti, to, dim3(gx, gy, (n + 15) / 16), stream, A);
    else if (n >= 8 && tiles * ((n + 7) / 8) >= 1024)
        st = launch_fpa_t<8>(ti, to, dim3(gx, gy, (n + 7) / 8), stream, A);
    else if (n >= 4 && tiles * ((n + 3) / 4) >= 1024)
        st = launch_fpa_t<4>(ti, to, dim3(gx, gy, (n + 3) / 4), stream, A);
    else if (n >= 2 && tiles * ((n + 1) / 2) >= 1024)
        st = launch_fpa_t<2>(ti, to, dim3(gx, gy, (n + 1) / 2), stream, A);
    else
        st = launch_fpa_t<1>(ti, to, dim3(gx, gy, n), stream, A);
    if (st != FBCT_OK) return st;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_adjoint_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

template <int TX, typename TI, typename TO>
static void launch_fpa_batch_t(dim3 grid, hipStream_t stream, const BatchDevice &bd, const int *prefix, int n,
                               int total, const void *sino, void *img)
{
    hipLaunchKernelGGL((fp_adjoint_batch_kernel<TX, TI, TO>), grid, dim3(kAdjThreads), 0, stream, bd.fpa_desc,
                       prefix, n, total, static_cast<const TI *>(sino), static_cast<TO *>(img));
}

template <int TX>
static void launch_fpa_batch_x(fbct_dtype ti, fbct_dtype to, dim3 grid, hipStream_t stream, const BatchDevice &bd,
                               const int *prefix, int n, int total, const void *sino, void *img)
{
    if (ti == FBCT_F32 && to == FBCT_F32) launch_fpa_batch_t<TX, float, float>(grid, stream, bd, prefix, n, total, sino, img);
    else if (ti == FBCT_BF16 && to == FBCT_F32) launch_fpa_batch_t<TX, bf16_t, float>(grid, stream, bd, prefix, n, total, sino, img);
    else if (ti == FBCT_F32 && to == FBCT_BF16) launch_fpa_batch_t<TX, float, bf16_t>(grid, stream, bd, prefix, n, total, sino, img);
    else launch_fpa_batch_t<TX, bf16_t, bf16_t>(grid, stream, bd, prefix, n, total, sino, img);
}

fbct_status launch_fp_adjoint_batch(const fbct_batch &b, const BatchDevice &bd, const void *sino, fbct_dtype ti,
                                    void *img, fbct_dtype to, hipStream_t stream)
{
    const int n = (int)b.plans.size();
    int k = 0;
    while (k < kAdjNumTiles && kAdjTileX[k] != b.shape.adj_tile_x) ++k;
    if (k == kAdjNumTiles) return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_batch_fp_adjoint: tile width %d", b.shape.adj_tile_x);
    const int total = b.fpa_prefix[k][n];
    if (total == 0) return FBCT_OK;
    const long long wgs = std::min<long long>(total, (long long)b.shape.adj_wgs_per_cu * bd.num_cu);
    const dim3 grid((unsigned)wgs);
    const int *prefix = bd.fpa_prefix[k];
    switch (b.shape.adj_tile_x) {
    case 8: launch_fpa_batch_x<8>(ti, to, grid, stream, bd, prefix, n, total, sino, img); break;
    case 32: launch_fpa_batch_x<32>(ti, to, grid, stream, bd, prefix, n, total, sino, img); break;
    case 64: launch_fpa_batch_x<64>(ti, to, grid, stream, bd, prefix, n, total, sino, img); break;
    default: launch_fpa_batch_x<16>(ti, to, grid, stream, bd, prefix, n, total, sino, img); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_adjoint_batch_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

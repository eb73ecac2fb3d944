// fp_adjoint.hip -- matched adjoint of the ray-driven FP (gfx950): cell-driven where
// every sample lies in the image's texel cells [-1, n-1], else gathered per pixel.
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
// Both forms write every pixel exactly once, without atomics.  The cell-driven kernels
// (section "cell-driven variant" below) sum each ray's samples per texel cell in closed
// form.  The per-pixel gather: each thread owns one output pixel and accumulates over
// all views.  Per view, the candidate rays are
// the columns whose lines pass within the support of the pixel (radius rho around its
// centre, rho = sqrt(2) inside the image), found from the pixel's fan angle; per
// candidate ray the sample range whose positions fall in the support is computed in
// closed form and the samples are evaluated at exactly the FP32 positions the forward
// kernel interpolates at.  A group of S <= 16 slices that share the geometry is
// processed together: the per-(pixel, view, ray) weight W_c is computed once and applied
// to S sinograms.  Sums over blocks of 8 views accumulate in FP32, across blocks in FP64.

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

// Sample window of one pixel.  kBorder: the pixel's tile touches the image edge, so the
// clamp window and the unbounded outer support of border pixels apply; interior tiles
// skip them.
struct PixelWin {
    float xlo_c, xhi_c, ylo_c, yhi_c;  // clamp window of the sample coordinate (border plateaus)
    float xlo_b, xhi_b, ylo_b, yhi_b;  // support of T_i in (u - i), unbounded outside border pixels
    float rho;                          // support radius (texels)
};

template <bool kBorder>
__device__ __forceinline__ PixelWin pixel_window(const FpaArgs &A, int i, int j)
{
    PixelWin w{-kBig, kBig, -kBig, kBig, -1.0f, 1.0f, -1.0f, 1.0f, 1.41421356f};
    if (kBorder) {
        if (i == 0) w.xlo_c = 0.0f, w.xlo_b = -kBig;
        if (i == A.nX - 1) w.xhi_c = 0.0f, w.xhi_b = kBig;
        if (j == 0) w.ylo_c = 0.0f, w.ylo_b = -kBig;
        if (j == A.nY - 1) w.yhi_c = 0.0f, w.yhi_b = kBig;
        const float ex = (i == A.nX - 1) ? fmaxf(1.0f, A.far_x) : 1.0f;
        const float ey = (j == A.nY - 1) ? fmaxf(1.0f, A.far_y) : 1.0f;
        w.rho = sqrtf(ex * ex + ey * ey);
    }
    return w;
}

// atan(q) for |q| <= 1/2 by the odd polynomial q P(q^2) of degree 7 (coefficients a
// weighted least-squares fit to atan on [0, 1/2]); its error, FP32 evaluation included,
// is below 5.5e-7 rad on [-1/2, 1/2] (checked on 2e5 points against FP64 atan), and
// kCandAtanErr bounds it.
constexpr float kCandAtanErr = 2.0e-6f;
__device__ __forceinline__ float cand_atan(float q)
{
    if (fabsf(q) > 0.5f) return atanf(q);
    const float s = q * q;
    return q * fmaf(fmaf(fmaf(-9.581676484e-02f, s, 1.920663981e-01f), s, -3.328139539e-01f), s,
                    9.999906794e-01f);
}

// Candidate rays of a pixel in one view: columns [clo, chi] whose lines pass within rho
// (texels) of the pixel centre (x, y) (mm).  Such rays subtend at most asin(rho dx / L)
// from the pixel's fan angle; L >= U and asin(z) <= z (1 + z^2) for z <= 0.9.  The fan
// angle comes from cand_atan and the range is widened by its error bound, so it holds
// every column the exact angle admits (an extra candidate adds zero samples).
__device__ __forceinline__ void candidate_rays(const FpaArgs &A, float2 cs, float x, float y, float rho, int &clo,
                                               int &chi)
{
    const float U = fmaf(-y, cs.y, fmaf(-x, cs.x, A.R));
    const float t = fmaf(x, cs.y, -y * cs.x);
    const float rU = __builtin_amdgcn_rcpf(U);
    const float cf = fmaf(cand_atan(t * rU), A.c_scale, A.c_off);
    const float z = rho * A.dx * rU;
    const float hw = z < 0.9f ? fmaf(z * fmaf(z, z, 1.0f), A.c_abs, fmaf(kCandAtanErr, A.c_abs, 1e-3f)) : 1e9f;
    clo = (int)fmaxf(ceilf(cf - hw), 0.0f);
    chi = (int)fminf(floorf(cf + hw), (float)(A.nCols - 1));
}

// Candidate rays of the cells of one 4 x 4 block in one view, from a linear model of the
// column coordinate about the block centre (xc, yc) (mm): cf(p) = c_scale atan(t/U) +
// c_off is c_scale times the polar angle of p about the source, whose Hessian has norm
// 1/r^2 (r the distance to the source).  Over the block's cell centres, within h = 1.5 dx
// of the centre along each axis, the linear model is therefore off by at most
// |c_scale| h^2 / Rmin^2, Rmin the least r over the block.  A ray at fan-angle offset g
// from a cell centre passes r sin|g| from it, so the rays that meet the cell lie within
// asin(rs dx / Rmin) <= z (1 + z^2), z = rs dx / Rmin, of its fan angle (rs the cell's
// support radius, below); the half-width adds the model's error and the atan bound to
// that, so every cell's range holds every column whose ray meets the cell.  One lane
// evaluates a view's model; the 16 lanes of the block read it with shuffles.
struct BlockCand {
    float cf, gx, gy, hw;  // centre column, d cf / dx, d cf / dy (per mm), half-width
};

__device__ __forceinline__ BlockCand block_candidates(const FpaArgs &A, float2 cs, float xc, float yc, float rho)
{
    const float U = fmaf(-yc, cs.y, fmaf(-xc, cs.x, A.R));
    const float t = fmaf(xc, cs.y, -yc * cs.x);
    const float rU = __builtin_amdgcn_rcpf(U);
    BlockCand b;
    b.cf = fmaf(cand_atan(t * rU), A.c_scale, A.c_off);
    // d atan(t/U) = (U dt - t dU) / (U^2 + t^2) with dt = (sin b, -cos b) . dp and
    // dU = -(cos b, sin b) . dp
    const float k = A.c_scale * __builtin_amdgcn_rcpf(fmaf(U, U, t * t));
    b.gx = k * fmaf(U, cs.y, t * cs.x);
    b.gy = k * fmaf(t, cs.y, -U * cs.x);
    const float h = 1.5f * A.dx;
    const float Umin = U - (fabsf(cs.x) + fabsf(cs.y)) * h;
    // every point of the block is at least Rmin from the source: its U is at least Umin,
    // and it lies within 1.415 h of the block centre, at sqrt(U^2 + t^2) from the source
    const float q = fmaf(U, U, t * t);
    const float rL = rsqrtf(q);
    const float rRm = __builtin_amdgcn_rcpf(fmaxf(Umin, fmaf(q, rL, -1.415f * h)));
    auto half_width = [&](float rs) {  // for a support of radius rs (texels) about each cell centre
        const float z = rs * A.dx * rRm;
        return (Umin > 0.0f && z < 0.9f)
                   ? fmaf(z * fmaf(z, z, 1.0f), A.c_abs,
                          fmaf(kCandAtanErr, A.c_abs, fmaf(A.c_abs * h * h, rRm * rRm, 1e-3f)))
                   : 1e9f;
    };
    // A line of unit direction d meets a square cell iff it passes within (|d_x| + |d_y|)/2
    // texels of its centre, which is below rho = 1/sqrt(2) except along the diagonals.  The
    // direction from the source to the block centre is (t sin b - U cos b, -U sin b - t cos b)
    // / sqrt(U^2 + t^2); a candidate ray's direction differs from it by at most 1.415 h / Rmin
    // (its cell's place in the block) plus hw / |c_scale| (its column, within the disk's
    // half-width hw), and (|d_x| + |d_y|)/2 changes by at most 0.7072 per radian; 0.004 texel
    // covers the FP32 sample positions (within 3e-4 texel of the line).  Only rays without
    // samples in the cell are dropped.
    const float hw_disk = half_width(rho);
    const float ax = fabsf(fmaf(t, cs.y, -U * cs.x)), ay = fabsf(fmaf(U, cs.y, t * cs.x));
    const float spread = fmaf(1.415f * h, rRm, hw_disk * __builtin_amdgcn_rcpf(A.c_abs));
    const float rs = fminf(rho, fmaf(0.5f * (ax + ay), rL, fmaf(0.7072f, spread, 0.004f)));
    b.hw = half_width(rs);
    return b;
}

// The cell's candidate range from its block's model (dxl, dyl: cell centre minus block
// centre, mm).
__device__ __forceinline__ void cell_candidates(const FpaArgs &A, const BlockCand &b, float dxl, float dyl, int &clo,
                                                int &chi)
{
    const float cf = fmaf(b.gx, dxl, fmaf(b.gy, dyl, b.cf));
    clo = (int)fmaxf(ceilf(cf - b.hw), 0.0f);
    chi = (int)fminf(floorf(cf + b.hw), (float)(A.nCols - 1));
}

// The model of view j of a group of 16 (lane (lane & ~15) | j holds it): every lane of
// the wave takes part.
__device__ __forceinline__ BlockCand block_model_of(const BlockCand &mine, int j)
{
    const int src = ((int)threadIdx.x & 48) | j;
    return BlockCand{__shfl(mine.cf, src), __shfl(mine.gx, src), __shfl(mine.gy, src), __shfl(mine.hw, src)};
}

// W_c(i, j): total bilinear weight pixel (fi, fj) receives from the N samples of ray
// (u0, w0, du, dw) = g, evaluated at the FP32 sample positions; inv = (1/du, 1/dw).
template <bool kBorder>
__device__ __forceinline__ float ray_weight(float4 g, float2 inv, int N, float fi, float fj, const PixelWin &w)
{
    const float X0 = g.x - fi, Y0 = g.y - fj;
    const float a1 = (w.xlo_b - X0) * inv.x, a2 = (w.xhi_b - X0) * inv.x;
    const float b1 = (w.ylo_b - Y0) * inv.y, b2 = (w.yhi_b - Y0) * inv.y;
    const float lo = fmaxf(fmaxf(fminf(a1, a2), fminf(b1, b2)), 0.0f);
    const float hi = fminf(fminf(fmaxf(a1, a2), fmaxf(b1, b2)), (float)(N - 1));
    float wsum = 0.0f;
    const float kend = floorf(hi);
    auto sample = [&](float kf) {
        const float X = fmaf(kf, g.z, X0);
        const float Y = fmaf(kf, g.w, Y0);
        if (kBorder) {
            const float wx = 1.0f - fabsf(fminf(fmaxf(X, w.xlo_c), w.xhi_c));
            const float wy = 1.0f - fabsf(fminf(fmaxf(Y, w.ylo_c), w.yhi_c));
            wsum = fmaf(fmaxf(wx, 0.0f), fmaxf(wy, 0.0f), wsum);
        } else {
            // inside the support |X|, |Y| < 1 up to rounding at its two ends
            wsum = fmaf(1.0f - fabsf(X), 1.0f - fabsf(Y), wsum);
        }
    };
    float kf = ceilf(lo);
    for (; kf + 1.0f <= kend; kf += 2.0f) {  // two samples per iteration
        sample(kf);
        sample(kf + 1.0f);
    }
    if (kf <= kend) sample(kf);
    return wsum;
}

// FP32 sums over blocks of kViewBlock views, FP64 across blocks.
constexpr int kViewBlock = 8;

// Per-pixel accumulation over all views, every ray's tables and S sinogram values read
// from global memory.
template <int S, bool kBorder, typename TI, typename TO>
__device__ __forceinline__ void adjoint_pixel(const FpaArgs &A, int i, int j, int s0)
{
    const float fi = (float)i, fj = (float)j;
    // pixel centre in mm, in the FP's frame (dx on both axes, centre at texel r - 1/2)
    const float x = (fi + 0.5f - A.r) * A.dx;
    const float y = (fj + 0.5f - A.r) * A.dx;
    const PixelWin pw = pixel_window<kBorder>(A, i, j);

    const TI *sino[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        sino[s] = static_cast<const TI *>(A.sino) + (size_t)min(s0 + s, A.nSlices - 1) * A.sino_stride;

    double acc[S];
    float vsum[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0, vsum[s] = 0.0f;

    for (int v = 0; v < A.nViews; ++v) {
        int clo, chi;
        candidate_rays(A, A.cs[v], x, y, pw.rho, clo, chi);
        const size_t rbase = (size_t)v * A.nCols;
        for (int c = clo; c <= chi; ++c) {
            // The S sinogram values are loaded first, unconditionally, so their latency
            // overlaps the weight computation (a ray that misses the support, or has no
            // samples, ends with wsum = 0; sinograms are expected to be finite).
            const int N = A.ray_n[rbase + c];
            const float4 g = A.ray_geo[rbase + c];
            const float2 inv = A.ray_inv[rbase + c];
            // one 32-bit byte offset serves all S sinograms (a slice's sinogram is < 4 GiB)
            const unsigned boff = (unsigned)(rbase + c) * (unsigned)sizeof(TI);
            float sv[S];
#pragma unroll
            for (int s = 0; s < S; ++s) sv[s] = load_f_at(sino[s], boff);
            const float wsum = ray_weight<kBorder>(g, inv, N, fi, fj, pw);
#pragma unroll
            for (int s = 0; s < S; ++s) vsum[s] = fmaf(wsum, sv[s], vsum[s]);
        }
        if (v % kViewBlock == kViewBlock - 1 || v == A.nViews - 1) {
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] += (double)vsum[s], vsum[s] = 0.0f;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (s0 + s < A.nSlices)
            store_f(static_cast<TO *>(A.img), (size_t)(s0 + s) * A.img_stride + (size_t)j * A.nX + i,
                    (float)(acc[s] * (double)A.scale));
}

// ---------------------------------------------------------------- staged variant
// For S >= 4 slices the per-ray loads dominate: S sinogram values and the ray's tables,
// gathered per pixel.  The staged variant copies, per batch of kStageVB views, the 64
// sinogram columns around the tile's candidate rays into LDS once (coalesced, slices
// interleaved per column, converted to FP32) together with the columns' ray tables; a
// pixel then reads a ray's S values with S/4 16-byte LDS reads.  The window of a view
// is derived from the four corner pixels of the tile (t/U, and so the fan angle, is
// extremal at a corner), lane 4m + c evaluating corner c of view m; a candidate column
// outside the window is read from global memory.
constexpr int kStageCols = 64;  // staged columns per view: one per lane
constexpr int kStageVB = 4;     // views per staging batch

template <int S>
struct AdjStage {
    static constexpr int kPitch = S + 4;  // floats per staged column: 16-byte multiple, off the bank period
    float val[kStageVB][kStageCols][kPitch];
    float4 geo[kStageVB][kStageCols];
    float2 inv[kStageVB][kStageCols];
    int n[kStageVB][kStageCols];
};

template <int S, bool kBorder, typename TI, typename TO>
__device__ __forceinline__ void adjoint_staged_tile(const FpaArgs &A, AdjStage<S> &st, int tx, int ty, int s0)
{
    static_assert(S % 4 == 0, "a staged column is read four slices at a time");
    constexpr int VB = kStageVB;
    constexpr int SQ = S / 4;  // slices staged per wave
    const int i = tx * kAdjTx + (int)threadIdx.x % kAdjTx;
    const int j = ty * kAdjTy + (int)threadIdx.x / kAdjTx;
    const bool valid = i < A.nX && j < A.nY;  // no early exit: every thread stages and waits
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const float fi = (float)i, fj = (float)j;
    const float x = (fi + 0.5f - A.r) * A.dx;
    const float y = (fj + 0.5f - A.r) * A.dx;
    const PixelWin pw = pixel_window<kBorder>(A, i, j);
    // corner of this lane for the window: lane 4m + c evaluates corner c
    const int ci = min(tx * kAdjTx + ((lane & 1) ? kAdjTx - 1 : 0), A.nX - 1);
    const int cj = min(ty * kAdjTy + ((lane & 2) ? kAdjTy - 1 : 0), A.nY - 1);
    const float cx = ((float)ci + 0.5f - A.r) * A.dx;
    const float cy = ((float)cj + 0.5f - A.r) * A.dx;
    // this wave stages slices [s0 + wave SQ, s0 + (wave + 1) SQ)
    const TI *sino_w[SQ];
#pragma unroll
    for (int q = 0; q < SQ; ++q)
        sino_w[q] = static_cast<const TI *>(A.sino) + (size_t)min(s0 + wave * SQ + q, A.nSlices - 1) * A.sino_stride;
    const TI *sino[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        sino[s] = static_cast<const TI *>(A.sino) + (size_t)min(s0 + s, A.nSlices - 1) * A.sino_stride;

    double acc[S];
    float vsum[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0, vsum[s] = 0.0f;

    for (int v0 = 0; v0 < A.nViews; v0 += VB) {
        int c_lo[VB];
        {
            int wlo, whi;
            candidate_rays(A, A.cs[min(v0 + (lane >> 2) % VB, A.nViews - 1)], cx, cy, 1.41421356f, wlo, whi);
#pragma unroll
            for (int sh = 1; sh <= 2; sh <<= 1) wlo = min(wlo, __shfl_xor(wlo, sh));
#pragma unroll
            for (int m = 0; m < VB; ++m)
                c_lo[m] = min(max(__builtin_amdgcn_readlane(wlo, 4 * m) - 1, 0), max(A.nCols - kStageCols, 0));
        }
        __syncthreads();  // the previous batch's reads are done
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            const int c = c_lo[m] + lane;
            if (v0 + m >= A.nViews || c >= A.nCols) continue;
            const size_t rb = (size_t)(v0 + m) * A.nCols + c;
            const unsigned boff = (unsigned)rb * (unsigned)sizeof(TI);
            float q4[SQ];
#pragma unroll
            for (int q = 0; q < SQ; ++q) q4[q] = load_f_at(sino_w[q], boff);
            float *dst = &st.val[m][lane][wave * SQ];  // this wave's SQ slots of the column
            if constexpr (SQ % 4 == 0) {
#pragma unroll
                for (int q = 0; q < SQ; q += 4)
                    *reinterpret_cast<float4 *>(dst + q) = make_float4(q4[q], q4[q + 1], q4[q + 2], q4[q + 3]);
            } else if constexpr (SQ == 2) {
                *reinterpret_cast<float2 *>(dst) = make_float2(q4[0], q4[1]);
            } else {
#pragma unroll
                for (int q = 0; q < SQ; ++q) dst[q] = q4[q];
            }
            if (wave == 0) st.geo[m][lane] = A.ray_geo[rb];
            else if (wave == 1) st.inv[m][lane] = A.ray_inv[rb];
            else if (wave == 2) st.n[m][lane] = A.ray_n[rb];
        }
        __syncthreads();
        if (!valid) continue;
#pragma unroll 1
        for (int m = 0; m < VB; ++m) {
            const int v = v0 + m;
            if (v >= A.nViews) break;
            int clo, chi;
            candidate_rays(A, A.cs[v], x, y, pw.rho, clo, chi);
            for (int c = clo; c <= chi; ++c) {
                const int lc = c - c_lo[m];
                int N;
                float4 g;
                float2 inv;
                float sv[S];
                if ((unsigned)lc < (unsigned)kStageCols) {
                    N = st.n[m][lc];
                    g = st.geo[m][lc];
                    inv = st.inv[m][lc];
#pragma unroll
                    for (int s = 0; s < S; s += 4) {
                        const float4 t4 = *reinterpret_cast<const float4 *>(&st.val[m][lc][s]);
                        sv[s] = t4.x, sv[s + 1] = t4.y, sv[s + 2] = t4.z, sv[s + 3] = t4.w;
                    }
                } else {
                    const size_t rb = (size_t)v * A.nCols + c;
                    N = A.ray_n[rb];
                    g = A.ray_geo[rb];
                    inv = A.ray_inv[rb];
                    const unsigned boff = (unsigned)rb * (unsigned)sizeof(TI);
#pragma unroll
                    for (int s = 0; s < S; ++s) sv[s] = load_f_at(sino[s], boff);
                }
                const float wsum = ray_weight<kBorder>(g, inv, N, fi, fj, pw);
#pragma unroll
                for (int s = 0; s < S; ++s) vsum[s] = fmaf(wsum, sv[s], vsum[s]);
            }
            if (v % kViewBlock == kViewBlock - 1 || v == A.nViews - 1) {
#pragma unroll
                for (int s = 0; s < S; ++s) acc[s] += (double)vsum[s], vsum[s] = 0.0f;
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (s0 + s < A.nSlices)
            store_f(static_cast<TO *>(A.img), (size_t)(s0 + s) * A.img_stride + (size_t)j * A.nX + i,
                    (float)(acc[s] * (double)A.scale));
}

template <int S, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_staged_kernel(FpaArgs A)
{
    __shared__ AdjStage<S> st;
    const int tx = blockIdx.x, ty = blockIdx.y;
    const bool border_tile = tx == 0 || ty == 0 || (tx + 1) * kAdjTx >= A.nX || (ty + 1) * kAdjTy >= A.nY;
    if (border_tile)
        adjoint_staged_tile<S, true, TI, TO>(A, st, tx, ty, blockIdx.z * S);
    else
        adjoint_staged_tile<S, false, TI, TO>(A, st, tx, ty, blockIdx.z * S);
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

// ---------------------------------------------------------------- cell-driven variant
// The per-pixel gather evaluates each sample once for every pixel of its 2x2 support.
// Here a thread owns a texel cell (ci, cj), the square between the pixel centres
// ci..ci+1 and cj..cj+1 (samples inside it interpolate between those four pixels).  Per
// view and candidate ray (lines within half a cell diagonal of the cell centre) it finds
// the ray's samples inside the cell: along each axis the samples before boundary b number
// ceil((b - u0) / du) (du > 0; floor + 1 for du < 0), evaluated by the same expression in
// both cells that share b, so the cells partition every ray's samples exactly.  The sums
// over those n samples of 1, the u- and w-fractions and their products (fractions
// ta + j du, tm + j dw) are closed-form, as in the forward kernel's cell walk, and
// accumulate (FP32) times the ray's value per slice; at the end they give the four corner
// weights, which are combined per pixel through LDS, clamp-to-edge folding the outer
// corners of cells -1 and n-1 into the border pixels.  Valid while every sample lies in the cells [-1, n-1] of both axes
// (2 rFOV/dx - n + 1/2 < 1, square images); otherwise the gather kernels run.  Every
// variant below evaluates a slice with the same arithmetic in the same order, so results
// do not depend on how many slices share a launch.
constexpr int kCellTile = kAdjCellTile;  // cells per tile side; a tile yields 15 x 15 pixels
constexpr int kCellOut = kCellTile - 1;
constexpr int kCells = kCellTile * kCellTile;  // one per thread
constexpr int kCombineS = 4;                   // slices combined per LDS pass
constexpr float kCellRho = 0.70710678f;        // half a cell diagonal (texels)

// Closed-form sums over the samples of ray g = (u0, w0, du, dw) (inv = (1/du, 1/dw), N
// samples) that lie in cell (fci, fcj): r = (n, sum of u-fractions, sum of
// w-fractions, sum of their products).  The bilinear weights the cell's corners (0,0)
// (1,0) (0,1) (1,1) receive are r0 - r1 - r2 + r3, r1 - r3, r2 - r3 and r3 (formed once
// per cell from the accumulated sums in cell_combine).
__device__ __forceinline__ void cell_sums(float4 g, float2 inv, float N, float fci, float fcj, float r[4])
{
    // sample range inside the cell, per axis, then intersected with [0, N): the samples
    // that precede boundary b along the ray number ceil((b - u0) / du) in either
    // direction (a sample exactly on b goes to the cell it leaves, where its bilinear
    // weights are the same), so the range is [ceil(min q), ceil(max q)); the upper
    // boundary uses the expression its other cell uses for its lower one
    const float eu = fci - g.x, ew = fcj - g.y;
    const float qu0 = eu * inv.x, qu1 = (fci + 1.0f - g.x) * inv.x;
    const float qw0 = ew * inv.y, qw1 = (fcj + 1.0f - g.y) * inv.y;
    const float lou = ceilf(fminf(qu0, qu1)), hiu = ceilf(fmaxf(qu0, qu1));
    const float low = ceilf(fminf(qw0, qw1)), hiw = ceilf(fmaxf(qw0, qw1));
    const float klo = fminf(fmaxf(fmaxf(lou, low), 0.0f), N);
    const float khi = fminf(fminf(hiu, hiw), N);
    const float n = fmaxf(khi - klo, 0.0f);
    // fractions of the first sample: the cell index is subtracted before the step is
    // added, so the FP32 rounding is that of the distance, not of the texel coordinate
    const float ta = fmaf(klo, g.z, -eu), tm = fmaf(klo, g.w, -ew);
    const float t1 = 0.5f * fmaf(n, n, -n);                    // sum of j < n
    const float t2 = t1 * fmaf(n, 2.0f / 3.0f, -1.0f / 3.0f);  // sum of j^2
    r[0] = n;
    r[1] = fmaf(g.z, t1, n * ta);  // sum of u-fractions
    r[2] = fmaf(g.w, t1, n * tm);  // sum of w-fractions
    r[3] = fmaf(ta, r[2], g.z * fmaf(tm, t1, g.w * t2));  // sum of products: ta r2 + du (tm t1 + dw t2)
}

// True when cell (ci, cj) can hold samples: every sample lies in the FOV disk of radius
// r about texel coordinate (r - 1/2, r - 1/2) (its chord endpoints are on the circle), so
// a cell whose nearest point is farther away (plus 0.01 texel for FP32 positions)
// receives none and its views can be skipped.
__device__ __forceinline__ bool cell_in_fov(const FpaArgs &A, float fci, float fcj)
{
    const float c0 = A.r - 0.5f;
    const float ex = fmaxf(fmaxf(fci - c0, c0 - (fci + 1.0f)), 0.0f);
    const float ey = fmaxf(fmaxf(fcj - c0, c0 - (fcj + 1.0f)), 0.0f);
    const float rr = A.r + 0.01f;
    return fmaf(ex, ex, ey * ey) <= rr * rr;
}

// Adds the S-slice contribution of one ray to the cell's sums.
template <int S>
__device__ __forceinline__ void cell_accumulate(float4 g, float2 inv, float N, const float (&sv)[S],
                                                float fci, float fcj, float (&acc)[4][S])
{
    float r[4];
    cell_sums(g, inv, N, fci, fcj, r);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < S; ++s) acc[q][s] = fmaf(r[q], sv[s], acc[q][s]);
}

// Cell (lx, ly) of this thread in the 16 x 16 tile: wave w takes cell rows 4w..4w+3 and
// each 16 lanes of it a 4 x 4 block, whose candidate columns lie close together (fewer
// LDS bank conflicts on the staged columns than a 16-cell row).
__device__ __forceinline__ void tile_cell(int &lx, int &ly)
{
    static_assert(kCellTile == 16 && kCells == 256, "lane layout assumes 16 x 16 cells, 256 threads");
    const int t = (int)threadIdx.x, lane = t & 63;
    lx = 4 * (lane >> 4) + (lane & 3);
    ly = 4 * (t >> 6) + ((lane >> 2) & 3);
}

// Combines the cell sums acc[4][S] (n, Su, Sw, Suw, each times the ray values) of the
// tile's cells, as corner partials, into its 15 x 15 pixels of
// slices [s0, s0 + S), kCombineS slices per pass through part (4 x kCombineS x kCells
// floats of LDS, free when called; every thread calls it).  Pixel (px, py) collects
// corner dx of cell px - dx (dx = 0, 1), and, on the borders, the clamped outer corner
// of cell -1 (pixel 0) or of cell n-1 (pixel n-1).
template <int S, typename TO>
__device__ __forceinline__ void cell_combine(const FpaArgs &A, float *part, const float (&acc)[4][S], int tx, int ty,
                                             int s0)
{
    int lx, ly;
    tile_cell(lx, ly);
    const int cid = ly * kCellTile + lx;
    const int px = tx * kCellOut + lx, py = ty * kCellOut + ly;
    const bool out = lx < kCellOut && ly < kCellOut && px < A.nX && py < A.nY;
    int xc[3], xd[3], yc[3], yd[3];  // local cell column and corner of each x (y) term
    int nxt = 0, nyt = 0;
    xc[nxt] = lx + 1, xd[nxt++] = 0;
    xc[nxt] = lx, xd[nxt++] = 1;
    if (px == 0) xc[nxt] = lx, xd[nxt++] = 0;
    if (px == A.nX - 1) xc[nxt] = lx + 1, xd[nxt++] = 1;
    yc[nyt] = ly + 1, yd[nyt++] = 0;
    yc[nyt] = ly, yd[nyt++] = 1;
    if (py == 0) yc[nyt] = ly, yd[nyt++] = 0;
    if (py == A.nY - 1) yc[nyt] = ly + 1, yd[nyt++] = 1;
#pragma unroll
    for (int sb = 0; sb < S; sb += kCombineS) {
        constexpr int kC = S < kCombineS ? S : kCombineS;
#pragma unroll
        for (int k = 0; k < kC; ++k) {  // corner weights from the sums (n, Su, Sw, Suw)
            const float sn = acc[0][sb + k], su = acc[1][sb + k], sw = acc[2][sb + k], suw = acc[3][sb + k];
            const float c01 = sw - suw;
            part[(0 * kC + k) * kCells + cid] = (sn - su) - c01;
            part[(1 * kC + k) * kCells + cid] = su - suw;
            part[(2 * kC + k) * kCells + cid] = c01;
            part[(3 * kC + k) * kCells + cid] = suw;
        }
        __syncthreads();
        if (out) {
#pragma unroll
            for (int k = 0; k < kC; ++k) {
                double sum = 0.0;
                for (int a = 0; a < nyt; ++a)
                    for (int b = 0; b < nxt; ++b)
                        sum += (double)part[((yd[a] * 2 + xd[b]) * kC + k) * kCells + yc[a] * kCellTile + xc[b]];
                if (s0 + sb + k < A.nSlices)
                    store_f(static_cast<TO *>(A.img), (size_t)(s0 + sb + k) * A.img_stride + (size_t)py * A.nX + px,
                            (float)(sum * (double)A.scale));
            }
        }
        __syncthreads();  // part is free again
    }
}

// Cell tile (tx, ty): cells tx*15-1 .. tx*15+14 (and likewise in y), output pixels
// tx*15 .. tx*15+14, slices [s0, s0 + S), every ray read from global memory.  Every
// thread of the workgroup calls it; part: 4 x min(S, kCombineS) x kCells floats of LDS.
template <int S, typename TI, typename TO>
__device__ __forceinline__ void cell_tile(const FpaArgs &A, int tx, int ty, int s0, float *part)
{
    int lx, ly;
    tile_cell(lx, ly);
    const int ci = tx * kCellOut - 1 + lx, cj = ty * kCellOut - 1 + ly;  // >= -1
    const float fci = (float)ci, fcj = (float)cj;
    const TI *sino[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        sino[s] = static_cast<const TI *>(A.sino) + (size_t)min(s0 + s, A.nSlices - 1) * A.sino_stride;
    float acc[4][S];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < S; ++s) acc[q][s] = 0.0f;

    const bool active = ci <= A.nX - 1 && cj <= A.nY - 1 && cell_in_fov(A, fci, fcj);
    // block of 4 x 4 cells of this lane's group of 16: centre and this cell's offset (mm);
    // cell ci's centre is texel coordinate ci + 1/2, at (ci + 1 - r) dx
    const float bx = (float)(tx * kCellOut - 1 + (lx & ~3)) + 2.5f - A.r;  // centre, texel units
    const float by = (float)(ty * kCellOut - 1 + (ly & ~3)) + 2.5f - A.r;
    const float dxl = ((float)(lx & 3) - 1.5f) * A.dx, dyl = ((float)(ly & 3) - 1.5f) * A.dx;
    // all lanes run the view loop (the models are shared by shuffles); inactive cells skip
    // the rays, and a wave without an active cell skips it all
    if (__any(active)) {
        for (int vb = 0; vb < A.nViews; vb += 16) {
            // lane j of each group of 16 evaluates the model of view vb + j
            const int vj = min(vb + ((int)threadIdx.x & 15), A.nViews - 1);
            const BlockCand mine = block_candidates(A, A.cs[vj], bx * A.dx, by * A.dx, kCellRho);
            const int nv = min(16, A.nViews - vb);
            for (int j = 0; j < nv; ++j) {
                const BlockCand bm = block_model_of(mine, j);
                if (!active) continue;
                const int v = vb + j;
                int clo, chi;
                cell_candidates(A, bm, dxl, dyl, clo, chi);
                // 32-bit element offsets from the tables' bases (a plan's tables hold fewer than
                // 2^26 rays)
                const unsigned rbase = (unsigned)v * (unsigned)A.nCols;
                for (int c = clo; c <= chi; ++c) {
                    const unsigned rb = rbase + (unsigned)c;
                    const float N = (float)load_at(A.ray_n, rb * 4u);
                    const float4 g = load_at(A.ray_geo, rb * 16u);
                    const float2 inv = load_at(A.ray_inv, rb * 8u);
                    const unsigned boff = rb * (unsigned)sizeof(TI);
                    float sv[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) sv[s] = load_f_at(sino[s], boff);
                    cell_accumulate<S>(g, inv, N, sv, fci, fcj, acc);
                }
            }
        }
    }
    cell_combine<S, TO>(A, part, acc, tx, ty, s0);
}

template <int S, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_cell_kernel(FpaArgs A)
{
    __shared__ float part[4 * (S < kCombineS ? S : kCombineS) * kCells];
    cell_tile<S, TI, TO>(A, blockIdx.x, blockIdx.y, blockIdx.z * S, part);
}

// Staged cell tile for S >= 4 slices: the sinogram columns around the tile's candidate
// rays are staged in LDS per batch of kStageVB views, as in the staged gather (window
// from the tile's four corner cells), each column's ray tables packed in one 32-byte
// record.  Per view the candidate columns run in three ascending stretches: left of the
// window (global memory), inside it (LDS only) and right of it (global memory), so the
// hot middle loop carries no global addressing and no branch.
template <int S, int VB>
struct CellStage {
    // floats per staged column: S + 4 (a 16-byte multiple, off the bank period) for the
    // slice groups, one value for a single slice
    static constexpr int kPitch = S >= 4 ? S + 4 : 1;
    float val[VB][kStageCols][kPitch];
    // ray tables of the staged columns: (u0, w0, du, dw) and (1/du, 1/dw, N, 0), each
    // a 16-byte record so that 16 lanes reading distinct columns hit distinct banks
    float4 geo[VB][kStageCols];
    float4 aux[VB][kStageCols];
};

template <int S, int VB>
union CellShared {
    CellStage<S, VB> stage;
    float part[4 * (S < kCombineS ? S : kCombineS) * kCells];
};

// VB views per staging batch (VB <= 16: lanes 4m + c evaluate the window corners).
// S = 1 (the uneven batch's slices): wave 2 stages the values.
template <int S, int VB, typename TI, typename TO>
__device__ __forceinline__ void cell_staged_tile(const FpaArgs &A, CellShared<S, VB> &sh, int tx, int ty, int s0)
{
    static_assert(S == 1 || S % 4 == 0, "a staged column is read four slices at a time");
    static_assert(VB >= 1 && VB <= 16, "window corners of at most 16 views per batch");
    constexpr int SQ = S >= 4 ? S / 4 : 1;  // slices staged per wave
    CellStage<S, VB> &st = sh.stage;
    int lx, ly;
    tile_cell(lx, ly);
    const int ci = tx * kCellOut - 1 + lx, cj = ty * kCellOut - 1 + ly;  // >= -1
    // no early exit: every thread stages and waits
    const bool active = ci <= A.nX - 1 && cj <= A.nY - 1 && cell_in_fov(A, (float)ci, (float)cj);
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const float fci = (float)ci, fcj = (float)cj;
    const float x = (fci + 1.0f - A.r) * A.dx;
    const float y = (fcj + 1.0f - A.r) * A.dx;
    // corner cell of this lane for the window: lane 4m + c evaluates corner c
    const int wci = (lane & 1) ? min(tx * kCellOut + kCellOut - 1, A.nX - 1) : tx * kCellOut - 1;
    const int wcj = (lane & 2) ? min(ty * kCellOut + kCellOut - 1, A.nY - 1) : ty * kCellOut - 1;
    const float cx = ((float)wci + 1.0f - A.r) * A.dx;
    const float cy = ((float)wcj + 1.0f - A.r) * A.dx;
    // the group's slices are addressed from its first one by 32-bit byte offsets (the
    // launch checks that S sinograms span less than 4 GiB); a ragged last group repeats
    // its last slice
    const TI *gbase = static_cast<const TI *>(A.sino) + (size_t)s0 * A.sino_stride;
    const int last = A.nSlices - 1 - s0;
    const unsigned sbytes = (unsigned)A.sino_stride * (unsigned)sizeof(TI);

    float acc[4][S];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < S; ++s) acc[q][s] = 0.0f;

    // a candidate column outside the staged window (rare): tables and values from global
    // memory, one slice at a time so that the path holds few registers (the scheduling
    // barrier keeps the loads from being hoisted together); the same FMAs as
    // cell_accumulate
    auto global_ray = [&](int v, int c) {
        const size_t rb = (size_t)v * A.nCols + c;
        float r[4];
        const float4 g = A.ray_geo[rb];
        cell_sums(g, A.ray_inv[rb], (float)A.ray_n[rb], fci, fcj, r);
        unsigned off = (unsigned)rb * (unsigned)sizeof(TI);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const float val = load_f_at(gbase, off);
            if (s < last) off += sbytes;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q][s] = fmaf(r[q], val, acc[q][s]);
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    // Staging is software-pipelined: while a batch is evaluated from LDS, the next batch's
    // values and ray records are loaded into registers (each wave SQ slices of every
    // view's columns, and the records of the views m = wave mod 4), and written to LDS
    // between the two barriers that follow.
    constexpr int kRecViews = (VB + 3) / 4;  // views whose records one wave loads
    auto window = [&](int v0, int (&c_lo)[VB]) {
        int wlo, whi;
        candidate_rays(A, A.cs[min(v0 + (lane >> 2) % VB, A.nViews - 1)], cx, cy, kCellRho, wlo, whi);
#pragma unroll
        for (int sh2 = 1; sh2 <= 2; sh2 <<= 1) wlo = min(wlo, __shfl_xor(wlo, sh2));
#pragma unroll
        for (int m = 0; m < VB; ++m)
            c_lo[m] = min(max(__builtin_amdgcn_readlane(wlo, 4 * m) - 1, 0), max(A.nCols - kStageCols, 0));
    };
    TI pv[VB][SQ];                // next batch: this wave's values of lane's column (storage bits)
    float4 pg[kRecViews];         // its ray records: geo
    float2 pi[kRecViews];         //   1 / steps
    int pn[kRecViews];            //   samples
    auto load_batch = [&](int v0, const int (&c_lo)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            const int c = min(c_lo[m] + lane, A.nCols - 1);
            const unsigned boff = (unsigned)((size_t)min(v0 + m, A.nViews - 1) * A.nCols + c) * (unsigned)sizeof(TI);
            if constexpr (S == 1) {
                if (wave == 0) pv[m][0] = load_at(gbase, boff);
            } else {
#pragma unroll
                for (int q = 0; q < SQ; ++q)
                    pv[m][q] = load_at(gbase, (unsigned)min(wave * SQ + q, last) * sbytes + boff);
            }
        }
#pragma unroll
        for (int r = 0; r < kRecViews; ++r) {
            const int m = wave + 4 * r;
            if (m >= VB) break;
            const int c = min(c_lo[m] + lane, A.nCols - 1);
            const size_t rb = (size_t)min(v0 + m, A.nViews - 1) * A.nCols + c;
            pg[r] = A.ray_geo[rb];
            pi[r] = A.ray_inv[rb];
            pn[r] = A.ray_n[rb];
        }
    };
    auto store_batch = [&](int v0, const int (&c_lo)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            if (v0 + m >= A.nViews || c_lo[m] + lane >= A.nCols) continue;
            if constexpr (S == 1) {
                if (wave == 0) st.val[m][lane][0] = widen(pv[m][0]);
            } else {
                float *dst = &st.val[m][lane][wave * SQ];  // this wave's SQ slots of the column
                if constexpr (SQ % 4 == 0) {
#pragma unroll
                    for (int q = 0; q < SQ; q += 4)
                        *reinterpret_cast<float4 *>(dst + q) = make_float4(
                            widen(pv[m][q]), widen(pv[m][q + 1]), widen(pv[m][q + 2]), widen(pv[m][q + 3]));
                } else if constexpr (SQ == 2) {
                    *reinterpret_cast<float2 *>(dst) = make_float2(widen(pv[m][0]), widen(pv[m][1]));
                } else {
#pragma unroll
                    for (int q = 0; q < SQ; ++q) dst[q] = widen(pv[m][q]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kRecViews; ++r) {
            const int m = wave + 4 * r;
            if (m >= VB || v0 + m >= A.nViews || c_lo[m] + lane >= A.nCols) continue;
            st.geo[m][lane] = pg[r];
            st.aux[m][lane] = make_float4(pi[r].x, pi[r].y, (float)pn[r], 0.0f);
        }
    };

    int c_lo[VB], c_nx[VB];
    window(0, c_nx);
    load_batch(0, c_nx);
    for (int v0 = 0; v0 < A.nViews; v0 += VB) {
#pragma unroll
        for (int m = 0; m < VB; ++m) c_lo[m] = c_nx[m];
        __syncthreads();  // the previous batch's reads are done
        store_batch(v0, c_lo);
        __syncthreads();
        if (v0 + VB < A.nViews) {
            window(v0 + VB, c_nx);
            load_batch(v0 + VB, c_nx);
        }
        if (!active) continue;
        // (each cell evaluates its own candidates here: the shared block model of the
        // unstaged tile ran 3.5% slower in this kernel, which holds more registers)
#pragma unroll 1
        for (int m = 0; m < VB; ++m) {
            const int v = v0 + m;
            if (v >= A.nViews) break;
            int clo, chi;
            candidate_rays(A, A.cs[v], x, y, kCellRho, clo, chi);
            const int wlo = c_lo[m], whi = c_lo[m] + kStageCols - 1;  // staged window
            int c = clo;
            for (; c <= min(chi, wlo - 1); ++c) global_ray(v, c);
            for (; c <= min(chi, whi); ++c) {
                const int lc = c - wlo;
                // the ray record is read before the values: LDS reads complete in order,
                // so the sums can run while the values are still in flight
                const float4 ax = st.aux[m][lc];
                const float4 gg = st.geo[m][lc];
                float sv[S];
                if constexpr (S == 1) {
                    sv[0] = st.val[m][lc][0];
                } else {
#pragma unroll
                    for (int s = 0; s < S; s += 4) {
                        const float4 t4 = *reinterpret_cast<const float4 *>(&st.val[m][lc][s]);
                        sv[s] = t4.x, sv[s + 1] = t4.y, sv[s + 2] = t4.z, sv[s + 3] = t4.w;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);  // keep all the reads ahead of the sums
                cell_accumulate<S>(gg, make_float2(ax.x, ax.y), ax.z, sv, fci, fcj, acc);
            }
            for (; c <= chi; ++c) global_ray(v, c);
        }
    }
    __syncthreads();  // staging reads are done: the union turns into corner weights
    cell_combine<S, TO>(A, sh.part, acc, tx, ty, s0);
}

template <int S, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_cell_staged_kernel(FpaArgs A)
{
    __shared__ CellShared<S, kStageVB> sh;
    unsigned tx, ty, g;  // blocked, XCD-aware order (neighbouring tiles share staged columns in L2)
    xcd_blocked_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x, gridDim.y, gridDim.z,
                     tx, ty, g);
    cell_staged_tile<S, kStageVB, TI, TO>(A, sh, (int)tx, (int)ty, (int)g * S);
}

// Uneven batch: slice s has its own geometry (desc[s]); work item = (slice, tile), items
// of slice s are [prefix[s], prefix[s+1]).  A slice whose samples stay in the cells
// [-1, n-1] takes 15 x 15-pixel cell tiles, as a single-slice call does; any other slice
// TX x (256/TX) gather tiles.  A fixed grid strides over all items, the XCDs taking
// them in chunks of 64, and a slice's tiles run in blocked order, so the tiles resident
// on one XCD cover a compact part of one image and share their rays through its L2.  VB > 0: the cell tiles stage their sinogram columns and ray records
// in LDS, VB views per batch (the same arithmetic in the same order as the unstaged
// tile, VB = 0).
template <int TX, int VB, typename TI, typename TO>
__global__ __launch_bounds__(kAdjThreads) void fp_adjoint_batch_kernel(const FpaSliceDesc *__restrict__ desc,
                                                                       const int *__restrict__ prefix, int n,
                                                                       int total, const TI *sino, TO *img)
{
    __shared__ CellShared<1, (VB > 0 ? VB : 1)> sh;
    float *part = sh.part;
    const int first = (int)xcd_chunk(blockIdx.x, gridDim.x);
    for (int item = first; item < total; item += gridDim.x) {
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
        // the slice's tables, as workgroup-uniform global-memory pointers (read from the
        // descriptor they would be generic, per-lane addresses: flat loads)
        A.ray_geo = uniform_global(d.ray_geo);
        A.ray_n = uniform_global(d.ray_n);
        A.ray_inv = uniform_global(d.ray_inv);
        A.cs = uniform_global(d.cs);
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
        const unsigned local = (unsigned)(item - prefix[lo]);
        unsigned tx, ty;
        if (adj_cells_fit(A.far_x, A.far_y)) {  // same test as the single-slice launch
            blocked_tile(local, adj_cell_tiles(A.nX), adj_cell_tiles(A.nY), tx, ty);
            if constexpr (VB > 0) cell_staged_tile<1, VB, TI, TO>(A, sh, (int)tx, (int)ty, 0);
            else cell_tile<1, TI, TO>(A, (int)tx, (int)ty, 0, part);
        } else {
            blocked_tile(local, (A.nX + TX - 1) / TX, (A.nY + 256 / TX - 1) / (256 / TX), tx, ty);
            adjoint_tile<1, TX, TI, TO>(A, (int)tx, (int)ty, 0);
        }
    }
}

}  // namespace

template <int S, typename TI, typename TO>
static void launch_fpa_k(dim3 grid, hipStream_t stream, const FpaArgs &A)
{
    if (adj_cells_fit(A.far_x, A.far_y) && (S < 4 || (long long)S * A.sino_stride * (long long)sizeof(TI) < (1LL << 32))) {
        const dim3 cg(adj_cell_tiles(A.nX), adj_cell_tiles(A.nY), grid.z);
        if constexpr (S >= 4)
            hipLaunchKernelGGL((fp_adjoint_cell_staged_kernel<S, TI, TO>), cg, dim3(kAdjThreads), 0, stream, A);
        else
            hipLaunchKernelGGL((fp_adjoint_cell_kernel<S, TI, TO>), cg, dim3(kAdjThreads), 0, stream, A);
    } else if constexpr (S >= 4) {
        hipLaunchKernelGGL((fp_adjoint_staged_kernel<S, TI, TO>), grid, dim3(kAdjThreads), 0, stream, A);
    } else {
        hipLaunchKernelGGL((fp_adjoint_kernel<S, TI, TO>), grid, dim3(kAdjThreads), 0, stream, A);
    }
}

template <int S>
static fbct_status launch_fpa_t(fbct_dtype ti, fbct_dtype to, dim3 grid, hipStream_t stream, const FpaArgs &A)
{
    if (ti == FBCT_F32 && to == FBCT_F32) launch_fpa_k<S, float, float>(grid, stream, A);
    else if (ti == FBCT_BF16 && to == FBCT_F32) launch_fpa_k<S, bf16_t, float>(grid, stream, A);
    else if (ti == FBCT_F32 && to == FBCT_BF16) launch_fpa_k<S, float, bf16_t>(grid, stream, A);
    else if (ti == FBCT_BF16 && to == FBCT_BF16) launch_fpa_k<S, bf16_t, bf16_t>(grid, stream, A);
    else return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_fp_adjoint: unknown storage type");
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
static void launch_fpa_batch_t(int vb, dim3 grid, hipStream_t stream, const BatchDevice &bd, const int *prefix, int n,
                               int total, const void *sino, void *img)
{
    const TI *si = static_cast<const TI *>(sino);
    TO *io = static_cast<TO *>(img);
    switch (vb) {
    case 4: hipLaunchKernelGGL((fp_adjoint_batch_kernel<TX, 4, TI, TO>), grid, dim3(kAdjThreads), 0, stream, bd.fpa_desc, prefix, n, total, si, io); break;
    case 8: hipLaunchKernelGGL((fp_adjoint_batch_kernel<TX, 8, TI, TO>), grid, dim3(kAdjThreads), 0, stream, bd.fpa_desc, prefix, n, total, si, io); break;
    default: hipLaunchKernelGGL((fp_adjoint_batch_kernel<TX, 0, TI, TO>), grid, dim3(kAdjThreads), 0, stream, bd.fpa_desc, prefix, n, total, si, io); break;
    }
}

template <int TX>
static void launch_fpa_batch_x(fbct_dtype ti, fbct_dtype to, int vb, dim3 grid, hipStream_t stream,
                               const BatchDevice &bd, const int *prefix, int n, int total, const void *sino, void *img)
{
    if (ti == FBCT_F32 && to == FBCT_F32) launch_fpa_batch_t<TX, float, float>(vb, grid, stream, bd, prefix, n, total, sino, img);
    else if (ti == FBCT_BF16 && to == FBCT_F32) launch_fpa_batch_t<TX, bf16_t, float>(vb, grid, stream, bd, prefix, n, total, sino, img);
    else if (ti == FBCT_F32 && to == FBCT_BF16) launch_fpa_batch_t<TX, float, bf16_t>(vb, grid, stream, bd, prefix, n, total, sino, img);
    else launch_fpa_batch_t<TX, bf16_t, bf16_t>(vb, grid, stream, bd, prefix, n, total, sino, img);
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
    const long long wgs = b.shape.adj_wgs_per_cu == 0
                              ? total
                              : std::min<long long>(total, (long long)b.shape.adj_wgs_per_cu * bd.num_cu);
    const dim3 grid((unsigned)wgs);
    const int *prefix = bd.fpa_prefix[k];
    const int vb = b.shape.adj_stage_views;
    switch (b.shape.adj_tile_x) {
    case 8: launch_fpa_batch_x<8>(ti, to, vb, grid, stream, bd, prefix, n, total, sino, img); break;
    case 32: launch_fpa_batch_x<32>(ti, to, vb, grid, stream, bd, prefix, n, total, sino, img); break;
    case 64: launch_fpa_batch_x<64>(ti, to, vb, grid, stream, bd, prefix, n, total, sino, img); break;
    default: launch_fpa_batch_x<16>(ti, to, vb, grid, stream, bd, prefix, n, total, sino, img); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_adjoint_batch_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

// bp.hip -- pixel-driven fan-beam backprojection (plain and 1/L^2-weighted) and its
// transpose (gfx950).
//
// Reference: BackProjOneView_fan_curve_kernel, BP/transform.cu:67-95 (unweighted) and
// FBP/transform.cu:61-89 (times 1/(U^2+t^2)), launched once per view with a memcpy
// and a device sync per view (BP/transform.cu:110-118).
//
// Here one launch covers every view and slice: a thread owns a pixel of S slices, loops
// over all views with the per-view (cos, sin) read as uniform scalars, computes the
// column coordinate, in-range decision, taps and weight once per (pixel, view) and
// applies them to the S slices, accumulates in registers and stores once.  The detector
// row is read with two gathers per update; the 1-D texture's clamp-to-edge linear
// filter is reproduced with exact FP32 weights.
//
// atan(t/U) is evaluated as z + z^3 R(z^2) with a correction polynomial R fitted per
// plan over the range FOV pixels reach, |t/U| <= rFOV/sqrt(R^2 - rFOV^2) (relative error
// below 1e-8 before FP32 evaluation, about one ulp after); plans whose range needs more
// than 9 terms use atanf.
//
// The reference's hard cut-off 0 < xi < nCols is a discontinuity: where a pixel sits
// within 0.01 column of a fan edge the decision is retaken in FP64 (tangent test on
// U, t from FP64 pixel coordinates and FP64 cos/sin), so FP32 rounding of xi cannot
// flip it.

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "dtype.hpp"
#include "fbct_internal.hpp"

namespace fbct {
namespace {

struct BpArgs {
    const float *sino;
    float *img;
    const float2 *cs;
    const double2 *csd;
    const int2 *fov_row;  // per row the FOV pixel range (FP64 membership of the centres)
    long long img_stride, sino_stride;
    int nX, nY, nViews, nCols;
    float R, xmin, ymin, dx, dy;
    float xi_scale, xi_offset;
    int edge_exact;
    double tan_min, tan_max;
    float atan_c[kBpAtanMax];  // correction polynomial R(s), s = z^2, lowest order first
};

constexpr int kBpTx = 16, kBpTy = 16;
constexpr float kEdgeBand = 0.01f;  // columns

struct PixelGeo {
    float x, y;     // FP32 pixel centre (mm), as the reference computes it
    double xd, yd;  // FP64 pixel centre for edge decisions
};

__device__ __forceinline__ PixelGeo pixel_geo(const BpArgs &A, int i, int j)
{
    PixelGeo p;
    p.x = fmaf(A.dx, (float)i + 0.5f, A.xmin);
    p.y = fmaf(A.dy, (float)j + 0.5f, A.ymin);
    p.xd = (double)A.xmin + (double)A.dx * ((double)i + 0.5);
    p.yd = (double)A.ymin + (double)A.dy * ((double)j + 0.5);
    return p;
}

// atan(z) for |z| within the fitted range: z + z*s*R(s), s = z^2 (kNc terms of R);
// kNc = 0 calls atanf.
template <int kNc>
__device__ __forceinline__ float fan_atan(const BpArgs &A, float z)
{
    if constexpr (kNc == 0) {
        return atanf(z);
    } else {
        const float s = z * z;
        float r = A.atan_c[kNc - 1];
#pragma unroll
        for (int k = kNc - 2; k >= 0; --k) r = fmaf(r, s, A.atan_c[k]);
        return fmaf(z * s, r, z);
    }
}

// Column coordinate xi of the pixel in view v (xi_raw of BP/transform.cu:85), its
// in-range flag and the FP32 U, t.
template <int kNc>
__device__ __forceinline__ bool column_of(const BpArgs &A, const PixelGeo &p, int v, float &xi, float &U, float &t)
{
    const float2 cs = A.cs[v];
    U = fmaf(-p.y, cs.y, fmaf(-p.x, cs.x, A.R));  // projection onto the iso-ray
    t = fmaf(p.x, cs.y, -p.y * cs.x);             // distance to the centre ray
    xi = fmaf(fan_atan<kNc>(A, t * __builtin_amdgcn_rcpf(U)), A.xi_scale, A.xi_offset);
    bool in = xi > 0.0f && xi < (float)A.nCols;
    if (A.edge_exact && (fabsf(xi) < kEdgeBand || fabsf(xi - (float)A.nCols) < kEdgeBand)) {
        const double2 c = A.csd[v];
        const double Ud = (double)A.R - p.xd * c.x - p.yd * c.y;
        const double td = p.xd * c.y - p.yd * c.x;
        in = (td - Ud * A.tan_min > 0.0) && (td - Ud * A.tan_max < 0.0);
    }
    return in;
}

// Texel-centre linear interpolation position: interpolate at xi - 1/2, clamped to the
// row; returns the left tap c0 in [0, nCols-2] and the weight of c0+1.
__device__ __forceinline__ int interp_pos(float xi, int nCols, float &al)
{
    const float ct = fminf(fmaxf(xi - 0.5f, 0.0f), (float)(nCols - 1));
    const float fc = fminf(floorf(ct), (float)max(nCols - 2, 0));
    al = ct - fc;
    return (int)fc;
}

// Weights of the taps c0 and c0 + 1 of one view: (1 - al, al), times the fan-beam
// distance weight 1 / (U^2 + t^2) for the weighted (FBP) form, so that a slice's update
// is two FMAs.
template <bool kWeighted>
__device__ __forceinline__ void bp_tap_weights(float al, float U, float t, float &w0, float &w1)
{
    if constexpr (kWeighted) {
        const float w = __builtin_amdgcn_rcpf(fmaf(U, U, t * t));
        w1 = al * w;
        w0 = w - w1;
    } else {
        w1 = al;
        w0 = 1.0f - al;
    }
}

// Slices [slice0 + S*blockIdx.z, +S) of the batch.
template <bool kWeighted, int S, int kNc>
__global__ __launch_bounds__(kBpTx *kBpTy) void bp_kernel(BpArgs A, int slice0)
{
    unsigned t, g;  // XCD-aware (tile, slice group)
    xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gridDim.y, gridDim.z, t, g);
    const int i = (t % gridDim.x) * kBpTx + threadIdx.x % kBpTx;
    const int j = (t / gridDim.x) * kBpTy + threadIdx.x / kBpTx;
    const size_t sl = (size_t)slice0 + (size_t)S * g;
    if (i >= A.nX || j >= A.nY) return;
    const PixelGeo pg = pixel_geo(A, i, j);
    float acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0f;
    const int2 fr = A.fov_row[j];
    if (i >= fr.x && i <= fr.y) {
        const float *sino = A.sino + sl * A.sino_stride;
        const int c1max = A.nCols - 1;
        for (int v = 0; v < A.nViews; ++v) {
            float xi, U, t;
            if (!column_of<kNc>(A, pg, v, xi, U, t)) continue;
            float al;
            const int c0 = interp_pos(xi, A.nCols, al);
            const int c1 = min(c0 + 1, c1max);
            const float *P = sino + (size_t)v * A.nCols;
            float w0, w1;
            bp_tap_weights<kWeighted>(al, U, t, w0, w1);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const float p0 = P[(size_t)s * A.sino_stride + c0];
                const float p1 = P[(size_t)s * A.sino_stride + c1];
                acc[s] = fmaf(w1, p1, fmaf(w0, p0, acc[s]));
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) A.img[(sl + s) * A.img_stride + (size_t)j * A.nX + i] = acc[s];
}

// Staged variant for S = 4, 8, 16 slices: per batch of kBpStageVB views the workgroup
// copies the 64 detector columns around its tile into LDS once (coalesced, the S slices
// of a column contiguous, so a tap of all S slices is S/4 16-byte LDS reads instead of
// S gathers).  The window of a view starts at the smallest column coordinate of the
// tile's four corner pixels (t/U, and so xi, is extremal at a corner; lane 4m + c
// evaluates corner c of view m); a pixel whose taps leave the window reads them from
// global memory.  Results are those of bp_kernel (same arithmetic per update).
constexpr int kBpStageCols = 64;
constexpr int kBpStageVB = 4;

template <int S>
struct BpStage {
    static constexpr int kPitch = S + 4;  // floats per staged column: 16-byte multiple, off the bank period
    float val[kBpStageVB][kBpStageCols][kPitch];
};

template <bool kWeighted, int S, int kNc>
__global__ __launch_bounds__(kBpTx *kBpTy) void bp_staged_kernel(BpArgs A, int slice0)
{
    static_assert(S % 4 == 0, "a staged column is read four slices at a time");
    constexpr int VB = kBpStageVB, SQ = S / 4;
    __shared__ BpStage<S> st;
    unsigned tbx, tby, g;  // blocked, XCD-aware order (neighbouring tiles share staged columns in L2)
    xcd_blocked_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x, gridDim.y, gridDim.z,
                     tbx, tby, g);
    const int tx0 = (int)tbx * kBpTx, ty0 = (int)tby * kBpTy;
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // wave w takes tile rows 4w..4w+3; each 16 lanes of it a 4x4 pixel block, whose
    // detector columns lie close together (fewer LDS bank conflicts than a 16-pixel row)
    static_assert(kBpTx == 16 && kBpTy == 16, "lane layout assumes 16x16 tiles");
    const int i = tx0 + 4 * (lane >> 4) + (lane & 3);
    const int j = ty0 + 4 * wave + ((lane >> 2) & 3);
    const size_t sl = (size_t)slice0 + (size_t)S * g;
    // no early exit: every thread stages and waits
    const bool valid = i < A.nX && j < A.nY;
    const PixelGeo pg = pixel_geo(A, min(i, A.nX - 1), min(j, A.nY - 1));
    const int2 fr = A.fov_row[min(j, A.nY - 1)];
    const bool inside = valid && i >= fr.x && i <= fr.y;
    const PixelGeo pc = pixel_geo(A, min(tx0 + ((lane & 1) ? kBpTx - 1 : 0), A.nX - 1),
                                  min(ty0 + ((lane & 2) ? kBpTy - 1 : 0), A.nY - 1));
    const float *sino = A.sino + sl * A.sino_stride;  // slice group base
    const float *sino_w = sino + (size_t)(wave * SQ) * A.sino_stride;  // this wave stages slices wave*SQ..
    const unsigned sbytes = (unsigned)A.sino_stride * 4u;
    float acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0f;

    // Staging is software-pipelined: while a batch is read from LDS, the next batch's
    // columns are loaded into registers, and written to LDS between the two barriers that
    // follow.
    auto window = [&](int v0, int (&c_lo)[VB]) {
        float xi, U, tt;
        (void)column_of<kNc>(A, pc, min(v0 + (lane >> 2) % VB, A.nViews - 1), xi, U, tt);
        int wlo = (int)floorf(fminf(fmaxf(xi - 0.5f, -1.0e6f), 1.0e6f));
#pragma unroll
        for (int sh = 1; sh <= 2; sh <<= 1) wlo = min(wlo, __shfl_xor(wlo, sh));
#pragma unroll
        for (int m = 0; m < VB; ++m)
            c_lo[m] = min(max(__builtin_amdgcn_readlane(wlo, 4 * m) - 1, 0), max(A.nCols - kBpStageCols, 0));
    };
    float q[VB][SQ];  // next batch: this wave's slices of lane's column, per view
    auto load_batch = [&](int v0, const int (&c_lo)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            // 32-bit byte offsets from the wave's first slice (the launch stages slice
            // groups only when four sinograms span less than 4 GiB)
            const unsigned off =
                ((unsigned)min(v0 + m, A.nViews - 1) * (unsigned)A.nCols + (unsigned)min(c_lo[m] + lane, A.nCols - 1)) *
                4u;
#pragma unroll
            for (int k = 0; k < SQ; ++k) q[m][k] = load_at(sino_w, off + (unsigned)k * sbytes);
        }
    };
    auto store_batch = [&](int v0, const int (&c_lo)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            if (v0 + m >= A.nViews || c_lo[m] + lane >= A.nCols) continue;
            float *dst = &st.val[m][lane][wave * SQ];
            if constexpr (SQ % 4 == 0) {
#pragma unroll
                for (int k = 0; k < SQ; k += 4)
                    *reinterpret_cast<float4 *>(dst + k) = make_float4(q[m][k], q[m][k + 1], q[m][k + 2], q[m][k + 3]);
            } else if constexpr (SQ == 2) {
                *reinterpret_cast<float2 *>(dst) = make_float2(q[m][0], q[m][1]);
            } else {
                dst[0] = q[m][0];
            }
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
        if (!inside) continue;
#pragma unroll 1
        for (int m = 0; m < VB; ++m) {
            const int v = v0 + m;
            if (v >= A.nViews) break;
            float xi, U, tt;
            if (!column_of<kNc>(A, pg, v, xi, U, tt)) continue;
            float al;
            const int c0 = interp_pos(xi, A.nCols, al);
            float w0, w1;
            bp_tap_weights<kWeighted>(al, U, tt, w0, w1);
            const int lc = c0 - c_lo[m];
            float p0[S], p1[S];
            if ((unsigned)lc < (unsigned)(kBpStageCols - 1) && c0 + 1 < A.nCols) {
#pragma unroll
                for (int s = 0; s < S; s += 4) {
                    const float4 a = *reinterpret_cast<const float4 *>(&st.val[m][lc][s]);
                    const float4 b = *reinterpret_cast<const float4 *>(&st.val[m][lc + 1][s]);
                    p0[s] = a.x, p0[s + 1] = a.y, p0[s + 2] = a.z, p0[s + 3] = a.w;
                    p1[s] = b.x, p1[s + 1] = b.y, p1[s + 2] = b.z, p1[s + 3] = b.w;
                }
            } else {
                const float *P = sino + (size_t)v * A.nCols;
                const int c1 = min(c0 + 1, A.nCols - 1);
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    p0[s] = P[(size_t)s * A.sino_stride + c0];
                    p1[s] = P[(size_t)s * A.sino_stride + c1];
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) acc[s] = fmaf(w1, p1[s], fmaf(w0, p0[s], acc[s]));
        }
    }
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < S; ++s) A.img[(sl + s) * A.img_stride + (size_t)j * A.nX + i] = acc[s];
}

// ------------------------------------------------------------------ transpose
// BP^T: sino[v][c] = sum over FOV pixels p of img[p] w_p hat_c(xi_p), the transpose of
// the two-tap interpolation above (clamped taps folded into the edge columns, the
// 0 < xi < nCols cut kept; w_p = 1 or 1/(U^2+t^2)), gathered per (view, column), so every
// sum has a fixed order and no atomics are needed.
//
// Column c's taps come from the pixels whose column coordinate lies within one column of
// its node xi = c + 1/2, i.e. between the node rays of columns c - 1 and c + 1.  A node
// ray (theta = beta + gamma_c) is marched along image lines: rows or columns of pixels.
// On line k the ray crosses the other axis at pixel coordinate m*(k) = a k + b (one FP64
// FMA), and the candidates are the pixels within hw(k) = L*(k) tan|dCol| /
// ((|pd| - |pl| tan|dCol|) h) of m*(k), cut to the line's FOV range.  Each candidate's
// position relative to the node ray comes from its small offset Delta = (m - m*(k)) h
// along the line: perpendicular distance d = Delta pd, distance from the source
// L = L*(k) + Delta pl, and (xi - 1/2) - c = atan(d/L) / dCol = z (1 - z^2/3) / dCol,
// z = d/L (|z| is at most about tan|dCol|, so the next term is below 1e-9 column).  Every
// quantity is small, so the column offset is accurate to a few 1e-7 column, where an FP32
// column coordinate carries half an ulp of nCols (3e-5 column at 888 columns); that
// error, amplified about 70-fold by the ramp filter that follows in FBP^T, is what a
// scatter with the forward kernels' coordinate would leave.  Sums are FP32 over the lines
// of a chunk and FP64 across chunks, for the same reason.
//
// Work: a workgroup takes 64 adjacent columns (one per lane) of 8 adjacent views (2 per
// thread; wave w holds views w, w + 4) and S slices (8 for most groups).  Its rays march a
// common axis (the one its middle ray runs closest to) in chunks of 8 lines; per chunk the
// workgroup stages the pixels its candidates span -- the union of every ray's candidate
// range over the chunk's lines, each bound linear in k and so extreme at a chunk end --
// for its S slices in LDS, and each candidate is one LDS read of S slices.  Each pixel
// staged serves about two columns of each of the 8 views.  A ray more than 60 degrees off
// that axis (wide fans with few columns) is gathered on its own axis from global memory.
// (Measured on 64 clinical slices: 8 slices x 2 views per thread 6.97 ms, 4 x 4 11.3 ms,
// 8 x 4 8.39 ms, 4 x 2 10.3 ms.)
struct BpaArgs {
    const float *img;
    float *sino;
    const double2 *csd;     // [nViews] cos, sin beta
    const double2 *col_csd; // [nCols] cos, sin gamma_c
    const int2 *fov_row, *fov_col;
    long long img_stride, sino_stride;
    int nX, nY, nViews, nCols;
    double R, xmin, ymin, dx, dy;
    float rFOV, tanc, icol;  // rFOV, tan|dCol|, 1/dCol
};

constexpr int kBpaChunk = 8;                 // lines per staged chunk
constexpr int kBpaS = 8, kBpaV = 2;          // slices per workgroup (largest groups), views per thread
constexpr int kBpaWin = 160;                 // staged pixels per line (wider windows read global memory)

struct float8_t {
    float4 a, b;
};
template <int S> struct BpaTap;
template <> struct BpaTap<1> { using T = float; };
template <> struct BpaTap<2> { using T = float2; };
template <> struct BpaTap<4> { using T = float4; };
template <> struct BpaTap<8> { using T = float8_t; };

template <int S>
__device__ __forceinline__ float tap_at(const typename BpaTap<S>::T &t, int s)
{
    if constexpr (S == 1) return t;
    else return reinterpret_cast<const float *>(&t)[s];
}

// Node ray of one (view, column) on a marching axis.
struct BpaRay {
    double a, b;               // m*(k) = a k + b
    float L0, Ls, pd, pl, hwc;
    int klo, khi;              // lines holding candidates (klo > khi: none)
};

__device__ __forceinline__ void bpa_ray(const BpaArgs &A, double cth, double sth, double Rs, double Rc, bool along_y,
                                        BpaRay &r)
{
    double p0, ext;
    float h;
    if (along_y) {  // lines are image rows j; candidates along x
        const double cot = cth / sth, y0 = A.ymin + 0.5 * A.dy, xs0 = cot * y0 + Rs / sth;
        r.a = cot * A.dy / A.dx;
        r.b = (xs0 - A.xmin) / A.dx - 0.5;
        r.L0 = (float)(Rc - cth * xs0 - sth * y0);
        r.Ls = (float)(-A.dy / sth);
        r.pd = (float)sth, r.pl = (float)(-cth), h = (float)A.dx;
        p0 = -Rs * cth, ext = fabs(sth);
    } else {  // lines are image columns i; candidates along y
        const double tn = sth / cth, x0 = A.xmin + 0.5 * A.dx, ys0 = tn * x0 - Rs / cth;
        r.a = tn * A.dx / A.dy;
        r.b = (ys0 - A.ymin) / A.dy - 0.5;
        r.L0 = (float)(Rc - cth * x0 - sth * ys0);
        r.Ls = (float)(-A.dx / cth);
        r.pd = (float)(-cth), r.pl = (float)(-sth), h = (float)A.dy;
        p0 = Rs * sth, ext = fabs(cth);
    }
    const float apd = fabsf(r.pd) - fabsf(r.pl) * A.tanc;
    r.hwc = apd > 0.05f ? A.tanc / (apd * h) * 1.0001f : 1e9f;
    // lines holding candidates: points within delta of the node ray (delta bounds every
    // candidate's |d|) and inside the FOV disk have |s| <= H along the ray from its foot
    // Rs (sin th, -cos th); their coordinate on the marching axis is p0 + s dir +- delta
    const double delta = ((double)A.R + A.rFOV) * A.tanc * 1.01 + 2.0 * fmax(A.dx, A.dy);
    const double dn = fmax(fabs(Rs) - delta, 0.0);
    const double H2 = (double)A.rFOV * A.rFOV - dn * dn;
    if (!(H2 > 0.0)) {
        r.klo = 1, r.khi = 0;
        return;
    }
    const double span = sqrt(H2) * ext + delta * (along_y ? fabs(cth) : fabs(sth));
    const double org = along_y ? A.ymin : A.xmin, step = along_y ? A.dy : A.dx;
    const int n_major = along_y ? A.nY : A.nX;
    r.klo = max((int)floor((p0 - span - org) / step - 0.5) - 1, 0);
    r.khi = min((int)ceil((p0 + span - org) / step - 0.5) + 1, n_major - 1);
}

// Hat of column c in rel = (xi - 1/2) - c: 1 - |rel| on (-1, 1); the edge columns take
// the clamped taps (weight 1 for xi - 1/2 beyond their node) up to the cut 0 < xi < nCols.
struct BpaHat {
    float clo, chi, cut_lo, cut_hi, icol, c3;
};

// Calls fn(m, w) for the candidates m of line k (FOV range fr) of ray r, w the hat weight
// (times 1/L^2 for the weighted form); n_minor bounds the candidate range.  kEdge: the
// column may be the first or the last (clamped taps, cut at half a column); for every
// other column the hat with its cut at |rel| = 1 is max(1 - |rel|, 0), the same value.
template <bool kWeighted, bool kEdge = true, typename Fn>
__device__ __forceinline__ void bpa_line(const BpaRay &r, const BpaHat &hh, float h, int k, int2 fr, int n_minor, Fn &&fn)
{
    const double ms = fma(r.a, (double)k, r.b);
    const double mfl = floor(ms);
    const int mi = (int)mfl;
    const float mf = (float)(ms - mfl);
    const float Lk = fmaf((float)k, r.Ls, r.L0);
    const float hw = fminf(fmaf(Lk, r.hwc, 1.0e-3f), (float)(2 * n_minor + 2));
    const int lo = max(fr.x, mi + (int)ceilf(mf - hw));
    const int hi = min(fr.y, mi + (int)floorf(mf + hw));
    for (int m = lo; m <= hi; ++m) {
        const float dl = ((float)(m - mi) - mf) * h;
        const float d = dl * r.pd;
        const float L = fmaf(dl, r.pl, Lk);
        const float z = d * __builtin_amdgcn_rcpf(L);
        const float rel = z * fmaf(z * z, hh.c3, hh.icol);
        float w;
        if constexpr (kEdge) {
            w = 1.0f - fabsf(fminf(fmaxf(rel, hh.clo), hh.chi));
            w = (rel > hh.cut_lo && rel < hh.cut_hi) ? w : 0.0f;
        } else {
            w = fmaxf(1.0f - fabsf(rel), 0.0f);
        }
        if constexpr (kWeighted) w *= __builtin_amdgcn_rcpf(fmaf(L, L, d * d));
        fn(m, w);
    }
}

__device__ __forceinline__ int wave_min(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    return x;
}

__device__ __forceinline__ int wave_max(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}

// Slices [slice0 + S*g, +S) of the batch, g = the workgroup's slice group.
template <bool kWeighted, int S, int kBpaVpt>
__global__ __launch_bounds__(256) void bp_adjoint_staged_kernel(BpaArgs A, int slice0)
{
    constexpr int kBpaViewsWg = 4 * kBpaVpt;  // views per workgroup
    using TapT = typename BpaTap<S>::T;
    __shared__ TapT tile[kBpaChunk * kBpaWin];
    __shared__ int2 fovc[kBpaChunk];  // FOV ranges of the chunk's lines
    __shared__ int red[2][4];
    unsigned cb, vg, g;  // blocked, XCD-aware order over (column block, view group, slice group)
    xcd_blocked_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x, gridDim.y, gridDim.z,
                     cb, vg, g);
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int c = (int)cb * 64 + lane;
    const bool cvalid = c < A.nCols;
    // the workgroup holds the first or the last column (uniform)
    const bool edge_wg = cb == 0 || (int)cb * 64 + 63 >= A.nCols - 1;
    const int cc = min(c, A.nCols - 1);
    const size_t sl = (size_t)slice0 + (size_t)S * g;
    const float *img = A.img + sl * A.img_stride;

    // the workgroup's marching axis: the one its middle ray runs closest to
    bool along_y;
    {
        const double2 vm = A.csd[min((int)vg * kBpaViewsWg + kBpaViewsWg / 2, A.nViews - 1)];
        const double2 cm = A.col_csd[min((int)cb * 64 + 32, A.nCols - 1)];
        along_y = fabs(vm.y * cm.x + vm.x * cm.y) >= fabs(vm.x * cm.x - vm.y * cm.y);
    }
    const double2 cg = A.col_csd[cc];
    const double Rs = A.R * cg.y, Rc = A.R * cg.x;
    BpaHat hh;
    hh.clo = c == 0 ? 0.0f : -3.0e38f;
    hh.chi = c == A.nCols - 1 ? 0.0f : 3.0e38f;
    hh.cut_lo = c == 0 ? -0.5f : -1.0f;
    hh.cut_hi = c == A.nCols - 1 ? 0.5f : 1.0f;
    hh.icol = A.icol;
    hh.c3 = -A.icol * (1.0f / 3.0f);

    BpaRay r[kBpaVpt];
    bool own[kBpaVpt];  // marched on its own axis (far from the workgroup's)
    int kmin = INT_MAX, kmax = INT_MIN;
#pragma unroll
    for (int j = 0; j < kBpaVpt; ++j) {
        const int v = (int)vg * kBpaViewsWg + wave + 4 * j;
        const double2 vb = A.csd[min(v, A.nViews - 1)];
        const double cth = vb.x * cg.x - vb.y * cg.y, sth = vb.y * cg.x + vb.x * cg.y;
        own[j] = (along_y ? fabs(sth) : fabs(cth)) < 0.5;
        bpa_ray(A, cth, sth, Rs, Rc, own[j] ? fabs(sth) >= fabs(cth) : along_y, r[j]);
        if (!cvalid || v >= A.nViews) r[j].klo = 1, r[j].khi = 0;
        if (!own[j] && r[j].klo <= r[j].khi) kmin = min(kmin, r[j].klo), kmax = max(kmax, r[j].khi);
    }
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    if (lane == 0) red[0][wave] = kmin, red[1][wave] = kmax;
    __syncthreads();
    const int Klo = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
    const int Khi = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
    __syncthreads();  // red is rewritten per chunk

    const int s_major = along_y ? A.nX : 1, s_minor = along_y ? 1 : A.nX;
    const int n_minor = along_y ? A.nX : A.nY;
    const int2 *fov = along_y ? A.fov_row : A.fov_col;
    const float h = along_y ? (float)A.dx : (float)A.dy;
    // pixel off of the group's S slices: 32-bit byte offsets from the group's first image
    // (the launch keeps S images under 4 GiB)
    const unsigned sbytes = (unsigned)A.img_stride * (unsigned)sizeof(float);
    auto load_taps = [&](unsigned off) {
        TapT t;
        float *tp = reinterpret_cast<float *>(&t);
        const unsigned b = off * (unsigned)sizeof(float);
#pragma unroll
        for (int s = 0; s < S; ++s) tp[s] = load_f_at(img, b + (unsigned)s * sbytes);
        return t;
    };
    // the rare global-memory paths (wide windows, rays on their own axis): one slice at a
    // time, so that they hold few registers (the scheduling barrier keeps the loads from
    // being hoisted together); the same FMAs in the same order as the staged path
    auto add_taps = [&](unsigned off, float w, float (&lacc)[S]) {
        const unsigned b = off * (unsigned)sizeof(float);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            lacc[s] = fmaf(w, load_f_at(img, b + (unsigned)s * sbytes), lacc[s]);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    double accd[kBpaVpt][S];
#pragma unroll
    for (int j = 0; j < kBpaVpt; ++j)
#pragma unroll
        for (int s = 0; s < S; ++s) accd[j][s] = 0.0;

    for (int k0 = Klo; k0 <= Khi; k0 += kBpaChunk) {
        const int k1 = min(k0 + kBpaChunk - 1, Khi);
        // this thread's candidate range over the chunk (bounds linear in k: ends suffice)
        int wlo = INT_MAX, whi = INT_MIN;
#pragma unroll
        for (int j = 0; j < kBpaVpt; ++j) {
            const int ka = max(k0, r[j].klo), kb = min(k1, r[j].khi);
            if (own[j] || ka > kb) continue;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int ke = e ? kb : ka;
                const double ms = fma(r[j].a, (double)ke, r[j].b);
                const float hw = fminf(fmaf(fmaf((float)ke, r[j].Ls, r[j].L0), r[j].hwc, 1.0e-3f), (float)(2 * n_minor + 2));
                wlo = min(wlo, (int)floor(ms - (double)hw));
                whi = max(whi, (int)ceil(ms + (double)hw));
            }
        }
        wlo = wave_min(wlo);
        whi = wave_max(whi);
        if (lane == 0) red[0][wave] = wlo, red[1][wave] = whi;
        __syncthreads();  // also: the previous chunk's reads of the tile are done
        const int mlo = max(min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3])) - 1, 0);
        const int mhi_all = min(max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3])) + 1, n_minor - 1);
        // a window wider than the tile (wide fans): this chunk reads global memory
        const bool wide = mhi_all - mlo + 1 > kBpaWin;
        const int nw = wide ? 0 : max(mhi_all - mlo + 1, 0);  // 0: nothing staged
        const int nl = k1 - k0 + 1;
        if (threadIdx.x < (unsigned)nl) fovc[threadIdx.x] = fov[k0 + threadIdx.x];
        if (nw > 0) {
            // rows (lines along x): a wave per line, lanes along it; columns: 8 lanes along
            // the chunk's lines, which are consecutive pixels of an image row
            static_assert(kBpaChunk == 8, "column staging puts the chunk's lines on 8 lanes");
            if (along_y) {
                for (int li = wave; li < nl; li += 4)
                    for (int mm = lane; mm < nw; mm += 64)
                        tile[li * kBpaWin + mm] = load_taps((unsigned)((k0 + li) * A.nX + mlo + mm));
            } else {
                const int li = lane & 7;
                if (li < nl)
                    for (int mm = (int)threadIdx.x >> 3; mm < nw; mm += 32)
                        tile[li * kBpaWin + mm] = load_taps((unsigned)((mlo + mm) * A.nX + k0 + li));
            }
        }
        __syncthreads();  // tile ready
        // lines run uniformly over the chunk (their FOV ranges from LDS); a ray skips the
        // lines outside its own range
        if (!wide) {
            if (nw == 0) continue;  // no candidates in this chunk (uniform)
            auto staged = [&](auto edge) {
                constexpr bool kEdge = decltype(edge)::value;
#pragma unroll
                for (int j = 0; j < kBpaVpt; ++j) {
                    if (own[j]) continue;
                    float lacc[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) lacc[s] = 0.0f;
                    for (int k = k0; k <= k1; ++k) {
                        if (k < r[j].klo || k > r[j].khi) continue;
                        // every candidate lies in the window (it is the union of the
                        // candidate ranges, one pixel wider); the clamp only keeps reads
                        // inside the tile
                        const TapT *row = &tile[(k - k0) * kBpaWin];
                        bpa_line<kWeighted, kEdge>(r[j], hh, h, k, fovc[k - k0], n_minor, [&](int m, float w) {
                            const TapT t = row[min(max(m - mlo, 0), nw - 1)];
#pragma unroll
                            for (int s = 0; s < S; ++s) lacc[s] = fmaf(w, tap_at<S>(t, s), lacc[s]);
                        });
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) accd[j][s] += (double)lacc[s];
                }
            };
            if (edge_wg) staged(std::true_type{});
            else staged(std::false_type{});
        } else {
#pragma unroll
            for (int j = 0; j < kBpaVpt; ++j) {
                if (own[j]) continue;
                float lacc[S];
#pragma unroll
                for (int s = 0; s < S; ++s) lacc[s] = 0.0f;
                for (int k = max(k0, r[j].klo); k <= min(k1, r[j].khi); ++k)
                    bpa_line<kWeighted>(r[j], hh, h, k, fovc[k - k0], n_minor, [&](int m, float w) {
                        add_taps((unsigned)(k * s_major + m * s_minor), w, lacc);
                    });
#pragma unroll
                for (int s = 0; s < S; ++s) accd[j][s] += (double)lacc[s];
            }
        }
    }

    // rays far from the workgroup's axis: their own axis, taps from global memory, FP32
    // sums over 8 lines and FP64 across them
#pragma unroll
    for (int j = 0; j < kBpaVpt; ++j) {
        if (!own[j] || r[j].klo > r[j].khi) continue;
        const int v = (int)vg * kBpaViewsWg + wave + 4 * j;
        const double2 vb = A.csd[v];
        const double sth = vb.y * cg.x + vb.x * cg.y, cth = vb.x * cg.x - vb.y * cg.y;
        const bool ay = fabs(sth) >= fabs(cth);
        const int sj = ay ? A.nX : 1, si = ay ? 1 : A.nX;
        const int nm = ay ? A.nX : A.nY;
        const int2 *fv = ay ? A.fov_row : A.fov_col;
        const float hj = ay ? (float)A.dx : (float)A.dy;
        for (int k0 = r[j].klo; k0 <= r[j].khi; k0 += kBpaChunk) {
            float lacc[S];
#pragma unroll
            for (int s = 0; s < S; ++s) lacc[s] = 0.0f;
            for (int k = k0; k <= min(k0 + kBpaChunk - 1, r[j].khi); ++k)
                bpa_line<kWeighted>(r[j], hh, hj, k, fv[k], nm, [&](int m, float w) {
                    add_taps((unsigned)(k * sj + m * si), w, lacc);
                });
#pragma unroll
            for (int s = 0; s < S; ++s) accd[j][s] += (double)lacc[s];
        }
    }

    if (!cvalid) return;
#pragma unroll
    for (int j = 0; j < kBpaVpt; ++j) {
        const int v = (int)vg * kBpaViewsWg + wave + 4 * j;
        if (v >= A.nViews) continue;
#pragma unroll
        for (int s = 0; s < S; ++s) A.sino[(sl + s) * A.sino_stride + (size_t)v * A.nCols + c] = (float)accd[j][s];
    }
}

BpArgs make_args(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img)
{
    BpArgs A;
    A.sino = sino;
    A.img = img;
    A.cs = t.view_cs;
    A.csd = t.view_csd;
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
    A.fov_row = t.fov_row;
    A.xi_scale = p.d.bp_xi_scale;
    A.xi_offset = p.d.bp_xi_offset;
    A.edge_exact = p.d.bp_edge_exact;
    A.tan_min = p.d.bp_tan_min;
    A.tan_max = p.d.bp_tan_max;
    for (int k = 0; k < kBpAtanMax; ++k) A.atan_c[k] = p.d.bp_atan_c[k];
    return A;
}

template <bool kWeighted, int kNc>
void launch_bp_groups(const BpArgs &A, dim3 tiles, int n, hipStream_t stream)
{
    // slice groups of 16/8/4/2 while they leave enough workgroups to fill the chip, the
    // remainder one slice per group; groups of 4 or more stage their detector columns
    const long long nt = (long long)tiles.x * tiles.y;
    // the staged kernels address a wave's (up to 4) sinograms by 32-bit byte offsets
    const bool stage_ok = (unsigned long long)A.sino_stride * 16ull < (1ull << 32);
    int done = 0;
    for (int S : {16, 8, 4, 2, 1}) {
        const int groups = (n - done) / S;
        if (groups == 0 || (S > 1 && nt * groups < 2048) || (S >= 4 && !stage_ok)) continue;
        const dim3 grid(tiles.x, tiles.y, groups);
        switch (S) {
        case 16: hipLaunchKernelGGL((bp_staged_kernel<kWeighted, 16, kNc>), grid, dim3(kBpTx * kBpTy), 0, stream, A, done); break;
        case 8: hipLaunchKernelGGL((bp_staged_kernel<kWeighted, 8, kNc>), grid, dim3(kBpTx * kBpTy), 0, stream, A, done); break;
        case 4: hipLaunchKernelGGL((bp_staged_kernel<kWeighted, 4, kNc>), grid, dim3(kBpTx * kBpTy), 0, stream, A, done); break;
        case 2: hipLaunchKernelGGL((bp_kernel<kWeighted, 2, kNc>), grid, dim3(kBpTx * kBpTy), 0, stream, A, done); break;
        default: hipLaunchKernelGGL((bp_kernel<kWeighted, 1, kNc>), grid, dim3(kBpTx * kBpTy), 0, stream, A, done); break;
        }
        done += groups * S;
        if (done == n) break;
    }
}

}  // namespace

fbct_status launch_bp(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img, int n,
                      bool weighted, hipStream_t stream)
{
    BpArgs A = make_args(p, t, sino, img);
    const dim3 tiles((p.g.nX + kBpTx - 1) / kBpTx, (p.g.nY + kBpTy - 1) / kBpTy);
    const int nc = p.d.bp_atan_nc;
    if (weighted) {
        if (nc == 9) launch_bp_groups<true, 9>(A, tiles, n, stream);
        else if (nc == 5) launch_bp_groups<true, 5>(A, tiles, n, stream);
        else launch_bp_groups<true, 0>(A, tiles, n, stream);
    } else {
        if (nc == 9) launch_bp_groups<false, 9>(A, tiles, n, stream);
        else if (nc == 5) launch_bp_groups<false, 5>(A, tiles, n, stream);
        else launch_bp_groups<false, 0>(A, tiles, n, stream);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "bp_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

fbct_status launch_bp_adjoint(const fbct_plan &p, const DeviceTables &t, const float *img, float *sino, int n,
                              bool weighted, hipStream_t stream)
{
    BpaArgs A;
    A.img = img;
    A.sino = sino;
    A.csd = t.view_csd;
    A.col_csd = t.col_csd;
    A.fov_row = t.fov_row;
    A.fov_col = t.fov_col;
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
    A.rFOV = p.g.rFOV;
    A.tanc = p.d.bpa_tan;
    A.icol = p.d.bp_xi_scale;
    // groups of kBpaS slices, the remainder in groups of 4, 2 and 1; every group size has
    // the same workgroups (kBpaV views per thread), so a slice's result does not depend
    // on the group it falls in
    static_assert(kBpaS == 8, "group sizes 8, 4, 2, 1");
    // a group's images are addressed by 32-bit byte offsets: its S images stay below 4 GiB
    const unsigned long long img_bytes = (unsigned long long)p.g.nX * p.g.nY * sizeof(float);
    if (img_bytes >= (1ULL << 32))
        return fail(FBCT_ERR_UNSUPPORTED, "fbct_bp_adjoint: an image of %llu bytes exceeds 4 GiB", img_bytes);
    const dim3 g2((p.g.nCols + 63) / 64, (p.g.nViews + 4 * kBpaV - 1) / (4 * kBpaV), 1);
    int done = 0;
    for (int S : {8, 4, 2, 1}) {
        if ((unsigned long long)S * img_bytes >= (1ULL << 32)) continue;
        const int groups = (n - done) / S;
        if (groups == 0) continue;
        const dim3 grid(g2.x, g2.y, groups);
#define FBCT_BPA(SS)                                                                                            \
    if (weighted) hipLaunchKernelGGL((bp_adjoint_staged_kernel<true, SS, kBpaV>), grid, dim3(256), 0, stream, A, done); \
    else hipLaunchKernelGGL((bp_adjoint_staged_kernel<false, SS, kBpaV>), grid, dim3(256), 0, stream, A, done)
        switch (S) {
        case 8: FBCT_BPA(8); break;
        case 4: FBCT_BPA(4); break;
        case 2: FBCT_BPA(2); break;
        default: FBCT_BPA(1); break;
        }
#undef FBCT_BPA
        done += groups * S;
        if (done == n) break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "bp_adjoint_staged_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

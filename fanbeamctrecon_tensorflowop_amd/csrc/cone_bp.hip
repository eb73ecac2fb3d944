// cone_bp.hip -- voxel-driven cone-beam backprojection (gfx950).
//
// Reference: BackProjOneView_CB_EA_kernel (FDK/GPURecon_GPU.cu:15-50): a thread per
// (x, y) column computes U, t, the column index and 1/(U^2+t^2) once per view and
// loops over z with the row index ((z - zs) D / U - detBottom) / dRow, reading a 2-D
// clamp/linear texture; the host launches it once per view (:245-250).
//
// Here one launch covers all views and volumes.  A thread owns an (x, y) column of KZ
// voxels of one volume, accumulates all views in registers and stores once.  Per
// (column, view) the column coordinate, its taps and weight and the in-range decision
// are computed once; per voxel only the row coordinate (one fma), the taps and a
// bilinear blend.  Texture semantics are reproduced with exact FP32 weights.  Near the
// two fan edges the in-range decision is retaken in FP64 on t/U.
//
// Volumes of 16 slices or more use cone_bp_lds_kernel (z-chunks of 32 voxels from 32
// slices on), which stages each workgroup's detector footprint in LDS (the taps of a
// 16 x 16 x 16 voxel tile span about 30 x 25 detector samples per view, read by 4096
// voxels); thinner volumes read the taps from global memory (cone_bp_kernel).

#include <hip/hip_runtime.h>

#include "dtype.hpp"
#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kTx = 16, kTy = 16;
constexpr float kEdgeBand = 0.01f;  // columns

struct ConeArgs {
    const float *q;        // [n][nViews][nRows][nCols] filtered projections
    void *vol;             // TO[n][nZ][nY][nX]
    const float2 *cs;
    const double2 *csd;
    int nX, nY, nZ, nViews, nCols, nRows, nzc;
    float R, xmin, ymin, zmin, dx, dy, dz, r2;
    float xi_scale, xi_offset, zr_scale, zr_offset;  // zr_offset includes the -1/2 texel shift
    double tan_min, tan_max;
    float atan_c[kBpAtanMax];
};

template <int kNc>
__device__ __forceinline__ float fan_atan(const ConeArgs &A, float z)
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

// kFlat: flat detector (xi linear in t/U, weight 1/U^2); else curved (atan, 1/(U^2+t^2)).
// kNc: atan polynomial terms (curved only).  KZ: voxels per thread along z.
template <bool kFlat, int kNc, int KZ, typename TO>
__global__ __launch_bounds__(kTx *kTy) void cone_bp_kernel(ConeArgs A)
{
    unsigned t, gz;  // XCD-aware (tile, z-chunk of a volume): an XCD keeps whole z-chunks
    xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gridDim.y, gridDim.z, t, gz);
    const int i = (t % gridDim.x) * kTx + threadIdx.x % kTx;
    const int j = (t / gridDim.x) * kTy + threadIdx.x / kTx;
    const int zc = gz % A.nzc;
    const size_t vi = gz / A.nzc;
    const int k0 = zc * KZ;
    if (i >= A.nX || j >= A.nY) return;
    const float x = fmaf(A.dx, (float)i + 0.5f, A.xmin);
    const float y = fmaf(A.dy, (float)j + 0.5f, A.ymin);
    float zk[KZ], acc[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        zk[k] = fmaf(A.dz, (float)(k0 + k) + 0.5f, A.zmin);
        acc[k] = 0.0f;
    }
    if (x * x + y * y <= A.r2) {
        const double xd = (double)A.xmin + (double)A.dx * ((double)i + 0.5);
        const double yd = (double)A.ymin + (double)A.dy * ((double)j + 0.5);
        const int nC = A.nCols, nR = A.nRows;
        const float cmax = (float)(nC - 1), rmax = (float)(nR - 1);
        const float cfl = (float)max(nC - 2, 0), rfl = (float)max(nR - 2, 0);
        const size_t view_elems = (size_t)nR * nC;
        const float *qv = A.q + vi * (size_t)A.nViews * view_elems;
        for (int v = 0; v < A.nViews; ++v, qv += view_elems) {
            const float2 cs = A.cs[v];
            const float U = fmaf(-y, cs.y, fmaf(-x, cs.x, A.R));
            const float t = fmaf(x, cs.y, -y * cs.x);
            const float rU = __builtin_amdgcn_rcpf(U);
            const float z = t * rU;
            const float xi = fmaf(kFlat ? z : fan_atan<kNc>(A, z), A.xi_scale, A.xi_offset);
            bool in = xi > 0.0f && xi < (float)nC;
            if (fabsf(xi) < kEdgeBand || fabsf(xi - (float)nC) < kEdgeBand) {
                const double2 c = A.csd[v];
                const double Ud = (double)A.R - xd * c.x - yd * c.y;
                const double td = xd * c.y - yd * c.x;
                in = (td - Ud * A.tan_min > 0.0) && (td - Ud * A.tan_max < 0.0);
            }
            if (!in) continue;
            const float w = kFlat ? rU * rU : __builtin_amdgcn_rcpf(fmaf(U, U, t * t));
            // column taps (clamp-to-edge, texel centres at c + 1/2)
            const float ct = fminf(fmaxf(xi - 0.5f, 0.0f), cmax);
            const float fc = fminf(floorf(ct), cfl);
            const float al = ct - fc;
            const int c0 = (int)fc;
            const int c1 = min(c0 + 1, nC - 1);
            const float a = rU * A.zr_scale;
            // row coordinate rt_k = z_k a + zr_offset is monotone in k: when the first and
            // last voxel of the chunk lie inside [0, nRows-1) no row clamp is needed
            const float rt_a = fmaf(zk[0], a, A.zr_offset), rt_b = fmaf(zk[KZ - 1], a, A.zr_offset);
            const bool inside = fminf(rt_a, rt_b) >= 0.0f && fmaxf(rt_a, rt_b) < rmax;
            auto tap_rows = [&](int r0, int r1, float &top, float &bot) {
                const float p00 = qv[r0 * nC + c0], p01 = qv[r0 * nC + c1];
                const float p10 = qv[r1 * nC + c0], p11 = qv[r1 * nC + c1];
                top = fmaf(al, p01 - p00, p00);
                bot = fmaf(al, p11 - p10, p10);
            };
            if (inside) {
#pragma unroll
                for (int k = 0; k < KZ; ++k) {
                    const float rt = fmaf(zk[k], a, A.zr_offset);
                    const float fr = floorf(rt);
                    const float be = rt - fr;
                    const int r0 = (int)fr;
                    float top, bot;
                    tap_rows(r0, r0 + 1, top, bot);
                    acc[k] = fmaf(w, fmaf(be, bot - top, top), acc[k]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < KZ; ++k) {
                    const float rt = fminf(fmaxf(fmaf(zk[k], a, A.zr_offset), 0.0f), rmax);
                    const float fr = fminf(floorf(rt), rfl);
                    const float be = rt - fr;
                    const int r0 = (int)fr;
                    float top, bot;
                    tap_rows(r0, min(r0 + 1, nR - 1), top, bot);
                    acc[k] = fmaf(w, fmaf(be, bot - top, top), acc[k]);
                }
            }
        }
    }
    TO *out = static_cast<TO *>(A.vol) + vi * (size_t)A.nZ * A.nY * A.nX + (size_t)j * A.nX + i;
#pragma unroll
    for (int k = 0; k < KZ; ++k)
        if (k0 + k < A.nZ) store_f(out, (size_t)(k0 + k) * A.nY * A.nX, acc[k]);
}

// LDS-staged variant for z-chunks of KZ = 16 or 32 voxels (plain layout).  The workgroup's
// detector footprint per view is derived from the four corner columns of its 16 x 16
// tile (t/U, and so the column coordinate, is extremal at a corner; the row coordinate at
// a corner and an end of the z-chunk), kLdsGroup views at a time: lane 4m + c evaluates
// corner c of view m.  Per batch of VB views the workgroup stages each footprint's rows
// once (one detector row per wave instruction, lane = column, indices clamped to the
// detector, which reproduces clamp-to-edge), and every voxel reads its four taps from
// LDS.  A column whose taps leave the staged window (a footprint taller than
// lds_rows<KZ>() or wider than kLdsCols) reads them from global memory instead.
// views per staging batch: with the next batch's rows loaded into registers while the
// current one is evaluated, one view (126 VGPRs, four waves per SIMD) beat two (150, three
// waves): 16.3 against 17.3 ms on the 512 x 512 x 256 FDK, 19.2 ms for two views unpipelined
constexpr int kLdsVB = 1;
constexpr int kLdsGroup = 16;                  // views per footprint evaluation (64 lanes / 4 corners)
constexpr int kLdsCols = 64;                   // staged columns per row: one per lane
constexpr int kLdsPitch = kLdsCols + 1;        // floats between the views of a staged row
// Staged element (view m, row r, column c) sits at r * stride + m * kLdsPitch + c: the
// four taps of a voxel of any view of the batch are two ds_read2 with immediate offsets
// (at most 255 dwords), and the odd row stride keeps adjacent rows off the same banks.
template <int VB>
constexpr int lds_stride()
{
    // odd, and 10 beyond the packed VB * kLdsPitch: 75 for one view, 141 for two (strides
    // 131 to 141 measured for two views: 141 had the fewest bank conflicts between the
    // lanes' rows)
    return ((VB * kLdsPitch) | 1) + 10;
}
// Staged rows per view for z-chunks of KZ voxels: a 16 x 16 x KZ tile's taps span at most
// 42 (KZ = 16) or 71 (KZ = 32) detector rows at the 0.7 mm / 1.0 mm, SID 541, SDD 949
// geometry (magnification up to 2.6 at the FOV edge), margin included.
template <int KZ>
constexpr int lds_rows()
{
    return KZ == 32 ? 72 : 44;
}

__device__ __forceinline__ int floor_to_int(float v)
{
    return (int)floorf(fminf(fmaxf(v, -1.0e6f), 1.0e6f));
}

// floor(v) as an integer in one instruction (v_cvt_flr_i32_f32); v is a finite row
// coordinate well inside the int range.
__device__ __forceinline__ int cvt_floor(float v)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

template <bool kFlat, int kNc, int KZ, int VB, typename TO>
__global__ __launch_bounds__(kTx *kTy) void cone_bp_lds_kernel(ConeArgs A)
{
    constexpr int kLdsRows = lds_rows<KZ>();
    constexpr int RS = lds_stride<VB>();
    static_assert(kLdsGroup % VB == 0, "a batch of views lies inside one footprint group");
    static_assert(RS + (VB - 1) * kLdsPitch + 1 <= 255, "tap offsets fit ds_read2");
    __shared__ float stage[kLdsRows * RS];
    unsigned t, gz;
    xcd_tile(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gridDim.x * gridDim.y, gridDim.z, t, gz);
    const int tx0 = (int)(t % gridDim.x) * kTx, ty0 = (int)(t / gridDim.x) * kTy;
    // each wave takes an 8x8 pixel quarter of the tile: its taps spread over fewer staged
    // rows and columns than a 16x4 strip's (fewer LDS bank conflicts)
    static_assert(kTx == 16 && kTy == 16, "lane layout assumes 16x16 tiles");
    const int i = tx0 + 8 * (((int)threadIdx.x >> 6) & 1) + ((int)threadIdx.x & 7);
    const int j = ty0 + 8 * ((int)threadIdx.x >> 7) + (((int)threadIdx.x >> 3) & 7);
    const int zc = gz % A.nzc;
    const size_t vi = gz / A.nzc;
    const int k0 = zc * KZ;
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const float x = fmaf(A.dx, (float)i + 0.5f, A.xmin);
    const float y = fmaf(A.dy, (float)j + 0.5f, A.ymin);
    // no early exit: every thread takes part in the staging barriers
    const bool active = i < A.nX && j < A.nY && x * x + y * y <= A.r2;
    // Row coordinate of voxel k: rt_k = rt_0 + k (dz a), rt_0 = z_0 a + zr_offset and
    // a = zr_scale / U -- one fma with k as an immediate.
    const float z0 = fmaf(A.dz, (float)k0 + 0.5f, A.zmin);
    const float z1 = fmaf(A.dz, (float)(k0 + KZ - 1) + 0.5f, A.zmin);
    float acc[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) acc[k] = 0.0f;
    const int ci = min(tx0 + ((lane & 1) ? kTx - 1 : 0), A.nX - 1);
    const int cj = min(ty0 + ((lane & 2) ? kTy - 1 : 0), A.nY - 1);
    const float cx = fmaf(A.dx, (float)ci + 0.5f, A.xmin);
    const float cy = fmaf(A.dy, (float)cj + 0.5f, A.ymin);
    const double xd = (double)A.xmin + (double)A.dx * ((double)i + 0.5);
    const double yd = (double)A.ymin + (double)A.dy * ((double)j + 0.5);
    const int nC = A.nCols, nR = A.nRows;
    const float cmax = (float)(nC - 1), rmax = (float)(nR - 1);
    const float cfl = (float)max(nC - 2, 0), rfl = (float)max(nR - 2, 0);
    const size_t view_elems = (size_t)nR * nC;
    const unsigned row_bytes = 4u * (unsigned)nC;
    const float *qvol = A.q + vi * (size_t)A.nViews * view_elems;

    // footprints of views vg + m, m < 16: after the 4-lane reduction lane 4m holds the
    // first staged column, the first staged row and the last staged row (a margin of one
    // sample on each side absorbs rounding between corner and voxel)
    auto footprints = [&](int vg, int &f_c, int &f_r0, int &f_r1) {
        const float2 cs = A.cs[min(vg + (lane >> 2), A.nViews - 1)];
        const float U = fmaf(-cy, cs.y, fmaf(-cx, cs.x, A.R));
        const float tt = fmaf(cx, cs.y, -cy * cs.x);
        const float rU = __builtin_amdgcn_rcpf(U);
        const float zz = tt * rU;
        float ctn = fmaf(kFlat ? zz : fan_atan<kNc>(A, zz), A.xi_scale, A.xi_offset) - 0.5f;
        const float aa = rU * A.zr_scale;
        const float ra = fmaf(z0, aa, A.zr_offset), rb = fmaf(z1, aa, A.zr_offset);
        float rtn = fminf(ra, rb), rtx = fmaxf(ra, rb);
#pragma unroll
        for (int sh = 1; sh <= 2; sh <<= 1) {
            ctn = fminf(ctn, __shfl_xor(ctn, sh));
            rtn = fminf(rtn, __shfl_xor(rtn, sh));
            rtx = fmaxf(rtx, __shfl_xor(rtx, sh));
        }
        f_c = floor_to_int(ctn) - 1;
        f_r0 = floor_to_int(rtn) - 1;
        f_r1 = floor_to_int(rtx) + 2;
    };
    // window of the batch at v0 (inside the group at vg): first column, first row and
    // row count per view (0 past the last view); wave-uniform
    auto batch_info = [&](int v0, int vg, int f_c, int f_r0, int f_r1, int (&c_lo)[VB], int (&r_lo)[VB],
                          int (&nrow)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            const int src = 4 * (v0 - vg + m);
            c_lo[m] = __builtin_amdgcn_readlane(f_c, src);
            r_lo[m] = __builtin_amdgcn_readlane(f_r0, src);
            const int r_hi = __builtin_amdgcn_readlane(f_r1, src);
            nrow[m] = (v0 + m < A.nViews) ? min(max(r_hi - r_lo[m] + 1, 1), kLdsRows) : 0;
        }
    };
    // Staging is software-pipelined: while a batch is read from LDS, the next batch's
    // rows (wave w: rows w, w + 4, ... of each view) are loaded into registers, and
    // written to LDS between the two barriers that follow.
    constexpr int kRw = kLdsRows / 4;  // staged rows per wave and view
    static_assert(kLdsRows % 4 == 0, "rows are dealt to the 4 waves");
    float pv[VB][kRw];
    auto load_batch = [&](int v0, const int (&c_lo)[VB], const int (&r_lo)[VB], const int (&nrow)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            if (nrow[m] == 0) continue;
            // rows by 32-bit byte offsets from the view's base (the launch checks that a
            // view spans less than 4 GiB)
            const float *qv = qvol + (size_t)(v0 + m) * view_elems;
            const unsigned c4 = 4u * (unsigned)min(max(c_lo[m] + lane, 0), nC - 1);
#pragma unroll
            for (int k = 0; k < kRw; ++k) {
                const int rr = wave + 4 * k;
                if (rr >= nrow[m]) break;
                const unsigned r = (unsigned)min(max(r_lo[m] + rr, 0), nR - 1);
                pv[m][k] = load_at(qv, r * row_bytes + c4);
            }
        }
    };
    auto store_batch = [&](const int (&nrow)[VB]) {
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            float *dst = stage + m * kLdsPitch + lane;
#pragma unroll
            for (int k = 0; k < kRw; ++k) {
                const int rr = wave + 4 * k;
                if (rr >= nrow[m]) break;
                dst[rr * RS] = pv[m][k];
            }
        }
    };

    int f_c, f_r0, f_r1;
    int c_lo[VB], r_lo[VB], nrow[VB], c_nx[VB], r_nx[VB], n_nx[VB];
    footprints(0, f_c, f_r0, f_r1);
    batch_info(0, 0, f_c, f_r0, f_r1, c_nx, r_nx, n_nx);
    load_batch(0, c_nx, r_nx, n_nx);
    for (int v0 = 0; v0 < A.nViews; v0 += VB) {
#pragma unroll
        for (int m = 0; m < VB; ++m) c_lo[m] = c_nx[m], r_lo[m] = r_nx[m], nrow[m] = n_nx[m];
        __syncthreads();  // the previous batch's reads are done
        store_batch(nrow);
        __syncthreads();
        const int v1 = v0 + VB;
        if (v1 < A.nViews) {
            if (v1 % kLdsGroup == 0) footprints(v1, f_c, f_r0, f_r1);
            batch_info(v1, v1 - v1 % kLdsGroup, f_c, f_r0, f_r1, c_nx, r_nx, n_nx);
            load_batch(v1, c_nx, r_nx, n_nx);
        }
        if (!active) continue;
#pragma unroll
        for (int m = 0; m < VB; ++m) {
            const int v = v0 + m;
            if (v >= A.nViews) break;
            const float2 cs = A.cs[v];
            const float U = fmaf(-y, cs.y, fmaf(-x, cs.x, A.R));
            const float tv = fmaf(x, cs.y, -y * cs.x);
            const float rU = __builtin_amdgcn_rcpf(U);
            const float z = tv * rU;
            const float xi = fmaf(kFlat ? z : fan_atan<kNc>(A, z), A.xi_scale, A.xi_offset);
            bool in = xi > 0.0f && xi < (float)nC;
            if (fabsf(xi) < kEdgeBand || fabsf(xi - (float)nC) < kEdgeBand) {
                const double2 c = A.csd[v];
                const double Ud = (double)A.R - xd * c.x - yd * c.y;
                const double td = xd * c.y - yd * c.x;
                in = (td - Ud * A.tan_min > 0.0) && (td - Ud * A.tan_max < 0.0);
            }
            if (!in) continue;
            const float w = kFlat ? rU * rU : __builtin_amdgcn_rcpf(fmaf(U, U, tv * tv));
            const float a = rU * A.zr_scale;
            const float step = a * A.dz;
            const float rt0 = fmaf(z0, a, A.zr_offset);
            const float rt1 = fmaf(step, (float)(KZ - 1), rt0);
            const float ctu = xi - 0.5f;
            const int cl = (int)floorf(ctu) - c_lo[m];
            const int rla = (int)floorf(fminf(rt0, rt1)) - r_lo[m];
            const int rlb = (int)floorf(fmaxf(rt0, rt1)) - r_lo[m];
            if ((unsigned)cl <= (unsigned)(kLdsCols - 2) && rla >= 0 && rlb <= nrow[m] - 2) {
                // unclamped coordinates: the staged rows and columns repeat the edges.
                // Rows are blended first, then the columns with weights w (1 - al), w al.
                const float al = __builtin_amdgcn_fractf(ctu);
                const float wr = w * al, wl = w - wr;
                const int base = m * kLdsPitch + cl - r_lo[m] * RS;
#pragma unroll
                for (int k = 0; k < KZ; ++k) {
                    const float rt = fmaf(step, (float)k, rt0);
                    const float be = __builtin_amdgcn_fractf(rt);
                    const float *p = stage + (cvt_floor(rt) * RS + base);
                    const float left = fmaf(be, p[RS] - p[0], p[0]);
                    const float right = fmaf(be, p[RS + 1] - p[1], p[1]);
                    acc[k] = fmaf(wl, left, fmaf(wr, right, acc[k]));
                }
            } else {
                const float *qv = qvol + (size_t)v * view_elems;
                const float ct = fminf(fmaxf(ctu, 0.0f), cmax);
                const float fc = fminf(floorf(ct), cfl);
                const float al = ct - fc;
                const int c0 = (int)fc, c1 = min(c0 + 1, nC - 1);
#pragma unroll
                for (int k = 0; k < KZ; ++k) {
                    const float rt = fminf(fmaxf(fmaf(step, (float)k, rt0), 0.0f), rmax);
                    const float fr = fminf(floorf(rt), rfl);
                    const float be = rt - fr;
                    const int r0 = (int)fr, r1 = min(r0 + 1, nR - 1);
                    const float p00 = qv[r0 * nC + c0], p01 = qv[r0 * nC + c1];
                    const float p10 = qv[r1 * nC + c0], p11 = qv[r1 * nC + c1];
                    const float top = fmaf(al, p01 - p00, p00);
                    const float bot = fmaf(al, p11 - p10, p10);
                    acc[k] = fmaf(w, fmaf(be, bot - top, top), acc[k]);
                }
            }
        }
    }
    if (!(i < A.nX && j < A.nY)) return;
    TO *out = static_cast<TO *>(A.vol) + vi * (size_t)A.nZ * A.nY * A.nX + (size_t)j * A.nX + i;
#pragma unroll
    for (int k = 0; k < KZ; ++k)
        if (k0 + k < A.nZ) store_f(out, (size_t)(k0 + k) * A.nY * A.nX, acc[k]);
}

template <bool kFlat, int kNc, int KZ>
void launch_lds(dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    constexpr int VB = kLdsVB;
    if (to == FBCT_BF16)
        hipLaunchKernelGGL((cone_bp_lds_kernel<kFlat, kNc, KZ, VB, bf16_t>), grid, dim3(kTx * kTy), 0, st, A);
    else
        hipLaunchKernelGGL((cone_bp_lds_kernel<kFlat, kNc, KZ, VB, float>), grid, dim3(kTx * kTy), 0, st, A);
}

template <bool kFlat, int kNc, int KZ>
void launch_kz(dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    if (to == FBCT_BF16)
        hipLaunchKernelGGL((cone_bp_kernel<kFlat, kNc, KZ, bf16_t>), grid, dim3(kTx * kTy), 0, st, A);
    else
        hipLaunchKernelGGL((cone_bp_kernel<kFlat, kNc, KZ, float>), grid, dim3(kTx * kTy), 0, st, A);
}

template <bool kFlat, int kNc>
void launch_nc(int kz, dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    if (kz == 32) launch_lds<kFlat, kNc, 32>(grid, st, A, to);
    else if (kz == 16) launch_lds<kFlat, kNc, 16>(grid, st, A, to);
    else if (kz == 4) launch_kz<kFlat, kNc, 4>(grid, st, A, to);
    else launch_kz<kFlat, kNc, 1>(grid, st, A, to);
}

}  // namespace

fbct_status launch_cone_bp(const fbct_cone_plan &p, const ConeDeviceTables &t, const float *q,
                           void *vol, fbct_dtype to, int n, hipStream_t stream)
{
    const fbct_cone_geom &g = p.g;
    ConeArgs A;
    A.q = q;
    A.vol = vol;
    A.cs = t.view_cs;
    A.csd = t.view_csd;
    A.nX = g.nX;
    A.nY = g.nY;
    A.nZ = g.nZ;
    A.nViews = g.nViews;
    A.nCols = g.nCols;
    A.nRows = g.nRows;
    A.R = g.R;
    A.xmin = g.xmin;
    A.ymin = g.ymin;
    A.zmin = g.zmin;
    A.dx = g.dx;
    A.dy = g.dy;
    A.dz = g.dz;
    A.r2 = p.d.r2;
    A.xi_scale = p.d.xi_scale;
    A.xi_offset = p.d.xi_offset;
    A.zr_scale = p.d.zr_scale;
    A.zr_offset = p.d.zr_offset - 0.5f;
    A.tan_min = p.d.tan_min;
    A.tan_max = p.d.tan_max;
    for (int k = 0; k < kBpAtanMax; ++k) A.atan_c[k] = p.d.atan_c[k];
    const int kz = g.nZ >= 32 ? 32 : (g.nZ >= 16 ? 16 : (g.nZ >= 4 ? 4 : 1));
    A.nzc = (g.nZ + kz - 1) / kz;
    const long long gz = (long long)A.nzc * n;
    if (gz > 65535) return fail(FBCT_ERR_UNSUPPORTED, "cone backprojection: %lld z-groups exceed the grid", gz);
    if ((unsigned long long)g.nRows * (unsigned long long)g.nCols * 4ull >= (1ull << 32))
        return fail(FBCT_ERR_UNSUPPORTED, "cone backprojection: a view of %d x %d exceeds 4 GiB", g.nRows, g.nCols);
    const dim3 grid((g.nX + kTx - 1) / kTx, (g.nY + kTy - 1) / kTy, (unsigned)gz);
    if (g.detector == FBCT_DETECTOR_FLAT) launch_nc<true, 0>(kz, grid, stream, A, to);
    else if (p.d.atan_nc == 9) launch_nc<false, 9>(kz, grid, stream, A, to);
    else if (p.d.atan_nc == 5) launch_nc<false, 5>(kz, grid, stream, A, to);
    else launch_nc<false, 0>(kz, grid, stream, A, to);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "cone_bp_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

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
// are computed once; per voxel only the row coordinate (one fma), the two row taps
// and a bilinear blend.  Texture semantics are reproduced with exact FP32 weights.
// Near the two fan edges the in-range decision is retaken in FP64 on t/U.

#include <hip/hip_runtime.h>

#include "dtype.hpp"
#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kTx = 16, kTy = 16;
constexpr float kEdgeBand = 0.01f;  // columns

struct ConeArgs {
    const float *q;        // [n][nViews][nRows][nCols] filtered projections (float2 pairs: see kPairs)
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
// kPairs: q holds column pairs (q[c], q[min(c+1, nC-1)]) as float2, so the two column
// taps of a row are one 8-byte load.
template <bool kFlat, int kNc, int KZ, bool kPairs, typename TO>
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
        constexpr int kStride = kPairs ? 2 : 1;  // floats per detector sample
        const float *qv = A.q + vi * (size_t)A.nViews * view_elems * kStride;
        for (int v = 0; v < A.nViews; ++v, qv += view_elems * kStride) {
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
                if constexpr (kPairs) {
                    const float2 u = *reinterpret_cast<const float2 *>(qv + 2 * (r0 * nC + c0));
                    const float2 d = *reinterpret_cast<const float2 *>(qv + 2 * (r1 * nC + c0));
                    top = fmaf(al, u.y - u.x, u.x);
                    bot = fmaf(al, d.y - d.x, d.x);
                } else {
                    const float p00 = qv[r0 * nC + c0], p01 = qv[r0 * nC + c1];
                    const float p10 = qv[r1 * nC + c0], p11 = qv[r1 * nC + c1];
                    top = fmaf(al, p01 - p00, p00);
                    bot = fmaf(al, p11 - p10, p10);
                }
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

template <bool kFlat, int kNc, int KZ, bool kPairs>
void launch_kz(dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    if (to == FBCT_BF16)
        hipLaunchKernelGGL((cone_bp_kernel<kFlat, kNc, KZ, kPairs, bf16_t>), grid, dim3(kTx * kTy), 0, st, A);
    else
        hipLaunchKernelGGL((cone_bp_kernel<kFlat, kNc, KZ, kPairs, float>), grid, dim3(kTx * kTy), 0, st, A);
}

template <bool kFlat, int kNc, bool kPairs>
void launch_nc(int kz, dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    if (kz == 16) launch_kz<kFlat, kNc, 16, kPairs>(grid, st, A, to);
    else if (kz == 4) launch_kz<kFlat, kNc, 4, kPairs>(grid, st, A, to);
    else launch_kz<kFlat, kNc, 1, kPairs>(grid, st, A, to);
}

template <bool kPairs>
void launch_det(const fbct_cone_plan &p, int kz, dim3 grid, hipStream_t st, const ConeArgs &A, fbct_dtype to)
{
    if (p.g.detector == FBCT_DETECTOR_FLAT) launch_nc<true, 0, kPairs>(kz, grid, st, A, to);
    else if (p.d.atan_nc == 9) launch_nc<false, 9, kPairs>(kz, grid, st, A, to);
    else if (p.d.atan_nc == 5) launch_nc<false, 5, kPairs>(kz, grid, st, A, to);
    else launch_nc<false, 0, kPairs>(kz, grid, st, A, to);
}

}  // namespace

fbct_status launch_cone_bp(const fbct_cone_plan &p, const ConeDeviceTables &t, const float *q, bool pairs,
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
    const int kz = g.nZ >= 16 ? 16 : (g.nZ >= 4 ? 4 : 1);
    A.nzc = (g.nZ + kz - 1) / kz;
    const long long gz = (long long)A.nzc * n;
    if (gz > 65535) return fail(FBCT_ERR_UNSUPPORTED, "cone backprojection: %lld z-groups exceed the grid", gz);
    const dim3 grid((g.nX + kTx - 1) / kTx, (g.nY + kTy - 1) / kTy, (unsigned)gz);
    if (pairs) launch_det<true>(p, kz, grid, stream, A, to);
    else launch_det<false>(p, kz, grid, stream, A, to);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "cone_bp_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

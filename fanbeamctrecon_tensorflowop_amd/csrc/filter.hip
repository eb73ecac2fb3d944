// filter.hip -- ramp filter of detector rows (gfx950).
//
// Reference: the host IPP filter of FBP/transform.cu:147-176 (and FDK/GPURecon_GPU.cu:
// 184-217): taps h[n], n in [-(nCols-1), nCols-1], h[0] = 1/(8 dCol^2),
// h[odd n] = -1/(2 pi^2 sin^2(n dCol)), h[even n != 0] = 0, scaled by R*dAngle*dCol;
// each row is convolved in full with ippsConv_32f and the centre nCols outputs kept,
// i.e. q[c] = sum_m p[m] h[c-m].
//
// Here: one workgroup per row (rows of all views and slices in one launch).  The row
// is staged in LDS with nCols zeros on both sides so no bounds checks are needed.  The
// even taps vanish, so output c only pairs with inputs of the opposite parity
// (plus the centre tap).  A thread produces kJ outputs of one parity c, c+2, ...,
// c+2(kJ-1); walking the odd taps n upward, the input those outputs need slides by two
// each step, so each step costs one LDS read and kJ FP64 FMAs with the tap a uniform
// (scalar-loaded) double.  Products p*h are exact in FP64 and the sums accumulate in
// FP64, so the only rounding is the final conversion to FP32.

#include <hip/hip_runtime.h>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

constexpr int kJ = 4;
constexpr int kFiltThreads = 256;

__global__ __launch_bounds__(kFiltThreads) void ramp_filter_kernel(const float *in, float *out, const double *h,
                                                                   int nC)
{
    extern __shared__ float pad[];  // [3*nC + 2*kJ]: zeros | row | zeros
    const long long row = blockIdx.x + (long long)blockIdx.y * gridDim.x;
    const float *p = in + row * nC;
    for (int i = threadIdx.x; i < 3 * nC + 2 * kJ; i += kFiltThreads) {
        const int m = i - nC;
        pad[i] = (m >= 0 && m < nC) ? p[m] : 0.0f;
    }
    __syncthreads();
    const double *hc = h + (nC - 1);  // hc[n] = h[n]
    const double h0 = hc[0];
    const int half = (nC + 1) / 2;                // outputs per parity (upper bound)
    const int groups = (half + kJ - 1) / kJ;      // work items per parity
    float *q = out + row * nC;
    // odd taps n in [nlo, nhi]
    const int nlo = -(nC - 1) | 1;  // smallest odd n >= -(nC-1)
    for (int item = threadIdx.x; item < 2 * groups; item += kFiltThreads) {
        const int par = item % 2;
        const int c0 = par + 2 * kJ * (item / 2);
        double acc[kJ];
        float win[kJ];
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            acc[j] = h0 * (double)pad[nC + c0 + 2 * j];
            win[j] = pad[nC + c0 + 2 * j - nlo];
        }
        for (int n = nlo; n <= nC - 1; n += 2) {
            const double hn = hc[n];
#pragma unroll
            for (int j = 0; j < kJ; ++j) acc[j] = fma((double)win[j], hn, acc[j]);
            // next n (n+2): win[j] <- pad[c0+2j-n-2] = win[j-1]; win[0] from LDS
#pragma unroll
            for (int j = kJ - 1; j > 0; --j) win[j] = win[j - 1];
            const int idx = nC + c0 - n - 2;
            win[0] = pad[max(idx, 0)];
        }
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const int c = c0 + 2 * j;
            if (c < nC) q[c] = (float)acc[j];
        }
    }
}

}  // namespace

fbct_status launch_ramp_filter(const fbct_plan &p, const DeviceTables &t, const float *in, float *out,
                               int64_t rows, hipStream_t stream)
{
    const int nC = p.g.nCols;
    const size_t lds = (size_t)(3 * nC + 2 * kJ) * sizeof(float);
    if (lds > 150 * 1024) return fail(FBCT_ERR_UNSUPPORTED, "ramp filter: row of %d columns too wide", nC);
    const int64_t gx = std::min<int64_t>(rows, 65535);
    const int64_t gy = (rows + gx - 1) / gx;
    if (gy > 65535) return fail(FBCT_ERR_UNSUPPORTED, "ramp filter: too many rows");
    if (gx * gy != rows) {
        // launch full grid rows, then the remainder with a row offset
        const int64_t full = (rows / gx) * gx;
        if (full > 0) {
            hipLaunchKernelGGL(ramp_filter_kernel, dim3((unsigned)gx, (unsigned)(full / gx)), dim3(kFiltThreads), lds,
                               stream, in, out, t.ramp, nC);
        }
        const int64_t rest = rows - full;
        hipLaunchKernelGGL(ramp_filter_kernel, dim3((unsigned)rest, 1), dim3(kFiltThreads), lds, stream,
                           in + full * nC, out + full * nC, t.ramp, nC);
    } else {
        hipLaunchKernelGGL(ramp_filter_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(kFiltThreads), lds, stream, in,
                           out, t.ramp, nC);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "ramp_filter_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

}  // namespace fbct

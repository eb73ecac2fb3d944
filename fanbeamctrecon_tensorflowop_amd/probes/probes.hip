// probes.hip -- throughput probes for calibrating hardware counters (measurement only;
// the operators never call these).
//
// Each kernel is bound by one known resource, so a rocprofv3 counter read on it gives
// that resource's rate at full use on this part:
//   fbct_probe_valu: eight independent v_fma_f32 chains per lane, 8 waves per SIMD
//                    (VALU issue bound; 2 flops per lane per instruction);
//   fbct_probe_lds:  16-byte LDS reads of consecutive addresses by consecutive lanes
//                    (no bank conflicts; LDS bandwidth bound).
// Both write one value per thread so the work is not optimised away.  fbct_probe_copy
// (a plain copy of known size) scales the HBM traffic counters; fbct_probe_marker
// separates the operations of a profiled run.

#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void probe_valu_kernel(float *out, int iters, float m, float c)
{
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = (float)(threadIdx.x + k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], m, c);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

constexpr int kLdsVec = 4096;  // float4 entries: 64 KiB

__global__ __launch_bounds__(kThreads) void probe_lds_kernel(float *out, int iters)
{
    __shared__ float4 buf[kLdsVec];
    for (int k = threadIdx.x; k < kLdsVec; k += kThreads) buf[k] = make_float4((float)k, 1.0f, 2.0f, 3.0f);
    __syncthreads();
    float4 acc[4] = {};
    unsigned idx = threadIdx.x;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4 v = buf[(idx + (unsigned)r * kThreads) & (kLdsVec - 1)];
            acc[r].x += v.x;
            acc[r].y += v.y;
            acc[r].z += v.z;
            acc[r].w += v.w;
        }
        idx += 4 * kThreads;
    }
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) s += acc[r].x + acc[r].y + acc[r].z + acc[r].w;
    out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

// Plain copy of n float4 (16 n bytes read, 16 n bytes written): the HBM traffic counters
// read on it give their scale factors.
__global__ __launch_bounds__(kThreads) void probe_copy_kernel(const float4 *src, float4 *dst, long long n)
{
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        dst[i] = src[i];
}

// Delimiter between the dispatches of a profile: its grid size (tag + 1 workgroups of 64
// threads) names the operation that follows; each workgroup writes one value.
__global__ __launch_bounds__(64) void probe_marker_kernel(float *out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = (float)blockIdx.x;
}

}  // namespace

extern "C" {

// src, dst: n float4 each.
int fbct_probe_copy(const void *src, void *dst, long long n, void *stream)
{
    hipLaunchKernelGGL(probe_copy_kernel, dim3(4096), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
    return (int)hipGetLastError();
}

// out: at least tag + 1 floats.
int fbct_probe_marker(float *out, int tag, void *stream)
{
    hipLaunchKernelGGL(probe_marker_kernel, dim3(tag + 1), dim3(64), 0, static_cast<hipStream_t>(stream), out);
    return (int)hipGetLastError();
}


// out: wgs * 256 floats.  Returns a hipError_t code.
int fbct_probe_valu(float *out, int wgs, int iters, void *stream)
{
    hipLaunchKernelGGL(probe_valu_kernel, dim3(wgs), dim3(kThreads), 0, static_cast<hipStream_t>(stream), out, iters,
                       0.999f, 0.001f);
    return (int)hipGetLastError();
}

int fbct_probe_lds(float *out, int wgs, int iters, void *stream)
{
    hipLaunchKernelGGL(probe_lds_kernel, dim3(wgs), dim3(kThreads), 0, static_cast<hipStream_t>(stream), out, iters);
    return (int)hipGetLastError();
}

}  // extern "C"

// filter_mfma.hip -- ramp filter as a Toeplitz GEMM on the gfx950 BF16 matrix cores.
//
// The centre-cropped convolution of FBP/transform.cu:163-174 is the product
//     Q[rows][nC] = P[rows][nC] * H[nC][nC],   H[m][c] = h[c - m],
// with H symmetric (h[-n] = h[n]).  BF16 keeps the FP32 exponent range but only 8
// mantissa bits, so both operands are split three ways, x = x_hi + x_lo + x_ll (each
// a BF16 number), and six products are accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16:
//     hi*hi + hi*lo + lo*hi + lo*lo + hi*ll + ll*hi,
// which carries the same information as an FP32 product (the dropped terms are below
// 2^-24 relative).  H's three planes are precomputed per plan (padded with zeros to
// [Np][Kp]); P is split in registers as it is loaded (optionally multiplied by a
// per-sample weight first, the cone-beam pre-weight; a BF16 input needs no split).
//
// Tiling: a wave computes 16 rows x 64 columns (4 MFMA tiles) over K in steps of 32;
// a workgroup of 4 waves covers 64 rows.  Operand fragments are loaded straight from
// global memory (P rows as FP32, H rows as BF16: H is symmetric, so the B fragment of
// column c is row c of H, contiguous in k).
// Lane layout of v_mfma_f32_16x16x32_bf16: A lane l holds row l%16, k = 8*(l/16)+j;
// B lane l holds column l%16, k = 8*(l/16)+j; D lane l holds rows 4*(l/16)+j, column l%16.

#include <hip/hip_runtime.h>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaK = 32;
constexpr int kWaveCols = 64;  // 4 tiles of 16

struct MfmaArgs {
    const void *in;    // TI[rows][nC]
    float *out;        // [rows][nC]
    const __bf16 *hp;  // [3][Np][Kp] split planes of H
    const float *w;    // optional per-sample weights [wrows][nC], row r uses w[r % wrows]
    const float *w2;   // optional second factor [w2rows][nC], row r uses w2[(r / w2div) % w2rows]
    long long rows;
    int nC, Np, Kp, wrows, w2div, w2rows;
};

__device__ __forceinline__ void split3(float x, __bf16 &hi, __bf16 &lo, __bf16 &ll)
{
    hi = (__bf16)x;
    const float r1 = x - (float)hi;
    lo = (__bf16)r1;
    const float r2 = r1 - (float)lo;
    ll = (__bf16)r2;
}

__device__ __forceinline__ float widen(float x) { return x; }
__device__ __forceinline__ float widen(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// kAligned: every row starts 16-byte aligned, so the 8 k-values of an A fragment load
// as 16 bytes.  TI: input storage (float or BF16 bits).  kW: inputs are multiplied by a
// per-sample weight before the split.  A BF16 input without weights is exactly its own
// hi part, so only the three products with the H planes are formed.
// kPairs: the output is written as column pairs, out2[r][c] = (q[r][c], q[r][min(c+1, nC-1)])
// (float2), the layout the cone backprojection reads its two column taps from.
template <bool kAligned, typename TI, bool kW, bool kPairs>
__global__ __launch_bounds__(256) void ramp_filter_mfma_kernel(MfmaArgs A)
{
    constexpr bool kExactA = sizeof(TI) == 2 && !kW;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long r0 = (long long)blockIdx.x * 64 + wave * 16;
    const int c0 = blockIdx.y * kWaveCols;
    const long long arow = r0 + (lane & 15);
    const bool arow_ok = arow < A.rows;
    const TI *prow = static_cast<const TI *>(A.in) + (arow_ok ? arow : 0) * (long long)A.nC;
    const float *wrow = kW && A.w ? A.w + (arow_ok ? arow % A.wrows : 0) * (long long)A.nC : nullptr;
    const float *w2row = kW && A.w2 ? A.w2 + (arow_ok ? (arow / A.w2div) % A.w2rows : 0) * (long long)A.nC : nullptr;
    const int kl = 8 * (lane >> 4);  // this lane's k offset within a 32-wide step
    const size_t plane = (size_t)A.Np * A.Kp;

    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < A.Kp; k0 += kMfmaK) {
        // A fragment: 8 consecutive k of one row, split into three BF16 planes
        float x[8];
        const int kb = k0 + kl;
        const bool full = arow_ok && kb + 8 <= A.nC;
        if (kAligned && full) {
            if constexpr (sizeof(TI) == 4) {
                const float4 u = *reinterpret_cast<const float4 *>(prow + kb);
                const float4 v = *reinterpret_cast<const float4 *>(prow + kb + 4);
                x[0] = u.x, x[1] = u.y, x[2] = u.z, x[3] = u.w, x[4] = v.x, x[5] = v.y, x[6] = v.z, x[7] = v.w;
            } else {
                const uint4 u = *reinterpret_cast<const uint4 *>(prow + kb);
                const uint32_t q[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    x[2 * j] = __uint_as_float(q[j] << 16);
                    x[2 * j + 1] = __uint_as_float(q[j] & 0xffff0000u);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = (arow_ok && kb + j < A.nC) ? widen(prow[kb + j]) : 0.0f;
        }
        if constexpr (kW) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const bool ok = arow_ok && kb + j < A.nC;
                float wt = ok ? 1.0f : 0.0f;
                if (wrow) wt *= ok ? wrow[kb + j] : 0.0f;
                if (w2row) wt *= ok ? w2row[kb + j] : 0.0f;
                x[j] *= wt;
            }
        }
        bf16x8 ah, al, all;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if constexpr (kExactA) {
                ah[j] = (__bf16)x[j];
            } else {
                __bf16 h, l, m;
                split3(x[j], h, l, m);
                ah[j] = h, al[j] = l, all[j] = m;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int col = c0 + 16 * t + (lane & 15);  // < Np: planes are zero-padded
            const size_t off = (size_t)col * A.Kp + kb;
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(A.hp + off);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(A.hp + plane + off);
            const bf16x8 bll = *reinterpret_cast<const bf16x8 *>(A.hp + 2 * plane + off);
            if constexpr (!kExactA) {
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(all, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bl, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[t], 0, 0, 0);
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bll, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int col = c0 + 16 * t + (lane & 15);
        if (col >= A.nC) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long row = r0 + 4 * (lane >> 4) + j;
            if (row >= A.rows) continue;
            if constexpr (kPairs) {
                float *o = A.out + 2 * (row * A.nC + col);
                o[0] = acc[t][j];
                if (col > 0) o[-1] = acc[t][j];
                if (col == A.nC - 1) o[1] = acc[t][j];
            } else {
                A.out[row * A.nC + col] = acc[t][j];
            }
        }
    }
}

template <typename TI, bool kW, bool kPairs>
void launch_mfma_t(bool aligned, dim3 grid, hipStream_t stream, const MfmaArgs &A)
{
    if (aligned)
        hipLaunchKernelGGL((ramp_filter_mfma_kernel<true, TI, kW, kPairs>), grid, dim3(256), 0, stream, A);
    else
        hipLaunchKernelGGL((ramp_filter_mfma_kernel<false, TI, kW, kPairs>), grid, dim3(256), 0, stream, A);
}

template <typename TI, bool kW>
void launch_mfma_p(bool pairs, bool aligned, dim3 grid, hipStream_t stream, const MfmaArgs &A)
{
    if (pairs) launch_mfma_t<TI, kW, true>(aligned, grid, stream, A);
    else launch_mfma_t<TI, kW, false>(aligned, grid, stream, A);
}

}  // namespace

fbct_status launch_ramp_gemm(const RampOperand &H, const void *in, fbct_dtype ti, float *out, int64_t rows,
                             const RowWeights &rw, hipStream_t stream, bool pairs)
{
    if (rows <= 0) return FBCT_OK;
    if ((rw.w && rw.wrows <= 0) || (rw.w2 && (rw.w2div <= 0 || rw.w2rows <= 0)))
        return fail(FBCT_ERR_INVALID_ARGUMENT, "ramp filter: weight row counts must be positive");
    const float *w = (rw.w || rw.w2) ? (rw.w ? rw.w : rw.w2) : nullptr;  // non-null: weighted kernel
    MfmaArgs A;
    A.in = in;
    A.out = out;
    A.hp = reinterpret_cast<const __bf16 *>(H.planes);
    A.w = rw.w;
    A.w2 = rw.w2;
    A.rows = rows;
    A.nC = H.nC;
    A.Np = H.Np;
    A.Kp = H.Kp;
    A.wrows = rw.w ? rw.wrows : 1;
    A.w2div = rw.w2 ? rw.w2div : 1;
    A.w2rows = rw.w2 ? rw.w2rows : 1;
    const long long gx = (rows + 63) / 64;
    if (gx > 0x7fffffffLL) return fail(FBCT_ERR_UNSUPPORTED, "ramp filter: too many rows");
    dim3 grid((unsigned)gx, (H.nC + kWaveCols - 1) / kWaveCols);
    const size_t esz = ti == FBCT_BF16 ? 2 : 4;
    const bool aligned = reinterpret_cast<uintptr_t>(in) % 16 == 0 && ((size_t)H.nC * esz) % 16 == 0;
    if (ti == FBCT_BF16) {
        if (w) launch_mfma_p<uint16_t, true>(pairs, aligned, grid, stream, A);
        else launch_mfma_p<uint16_t, false>(pairs, aligned, grid, stream, A);
    } else {
        if (w) launch_mfma_p<float, true>(pairs, aligned, grid, stream, A);
        else launch_mfma_p<float, false>(pairs, aligned, grid, stream, A);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "ramp_filter_mfma_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

fbct_status launch_ramp_filter_mfma(const fbct_plan &p, const DeviceTables &t, const float *in, float *out,
                                    int64_t rows, hipStream_t stream)
{
    RampOperand H{t.ramp_planes, p.g.nCols, p.d.ramp_np, p.d.ramp_kp};
    return launch_ramp_gemm(H, in, FBCT_F32, out, rows, RowWeights{}, stream, false);
}

}  // namespace fbct

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
// 2^-24 relative).  The taps' three planes are precomputed per plan; P is split in
// registers as it is loaded (optionally multiplied by a per-sample weight first, the
// cone-beam pre-weight; a BF16 input needs no split).
//
// Tiling: a workgroup of 4 waves computes 128 rows x 128 (or 256) columns; per K step of 32 it
// stages the BF16 planes of its P rows (split once per element) and the window of split
// taps its H columns read (H is Toeplitz and symmetric: the B fragment of column c at k is
// h[k - c .. k - c + 7], read from skewed copies of the taps so that every fragment is one
// aligned 16-byte LDS read) in LDS, and each wave multiplies its part, 64 rows x 64 (or
// 128) columns of 16 x 16 MFMA tiles.
// Lane layout of v_mfma_f32_16x16x32_bf16: A lane l holds row l%16, k = 8*(l/16)+j;
// B lane l holds column l%16, k = 8*(l/16)+j; D lane l holds rows 4*(l/16)+j, column l%16.

#include <hip/hip_runtime.h>

#include "fbct_internal.hpp"

namespace fbct {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


struct MfmaArgs {
    const void *in;    // TI[rows][nC]
    float *out;        // [rows][nC]
    const __bf16 *hp;  // [3][kRampSkew][Lg] skewed split taps (fbct_internal.hpp)
    const float *w;    // optional per-sample weights [wrows][nC], row r uses w[r % wrows]
    const float *w2;   // optional second factor [w2rows][nC], row r uses w2[(r / w2div) % w2rows]
    long long rows;
    int nC, Np, Kp, Lg, wrows, w2div, w2rows;
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

// A workgroup (4 waves) computes 128 rows x BN columns; per K step of 32 it stages the
// BF16 planes of its P rows (split once per element, weights applied first) and the tap
// window of its H columns in LDS, then each wave multiplies its 64 x BN/2 part from LDS.
// Rows of the P planes are padded to 40 BF16 so the 16 rows a fragment read touches fall
// in distinct banks.
constexpr int kBM = 128, kBK = 32, kKS = 40;

// BN output columns per workgroup (128 or 256; each wave takes BN / 2 of them, NT tiles of
// 16).  Staged taps per plane and copy and K step: the fragments of a K step reach offsets
// s in [k0 - c0 - BN + 1, k0 - c0 + 31], plus 7 for the copies' shift and 7 for the
// window's 8-alignment: BN + 48 BF16 per row, 88 or 152 dwords (an odd multiple of 8, so
// the 16 lanes of a fragment read, 8 copies x 2 windows 16 bytes apart, fall in distinct
// bank groups).
template <typename TI, bool kW, int BN>
__global__ __launch_bounds__(256) void ramp_gemm_kernel(MfmaArgs A)
{
    constexpr int kTW = BN + 48, NT = BN / 32;
    static_assert((kTW / 2 / 8) % 2 == 1, "tap rows an odd multiple of 8 dwords");
    constexpr bool kExactA = sizeof(TI) == 2 && !kW;
    constexpr int kPA = kExactA ? 1 : 3;  // A planes in use
    __shared__ __bf16 As[kPA][kBM][kKS];
    __shared__ __bf16 Ts[3][kRampSkew][kTW];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // each XCD takes whole row tiles, their column tiles consecutive (P rows stay in its L2)
    const unsigned lblk = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const long long r0 = (long long)(lblk / gridDim.y) * kBM;
    const int c0 = (int)(lblk % gridDim.y) * BN;
    // staging assignment: thread -> (row or column tid / 2, 16 k-values at kh)
    const int srow = tid >> 1, kh = (tid & 1) * 16;
    const long long arow = r0 + srow;
    const bool arow_ok = arow < A.rows;
    const TI *prow = static_cast<const TI *>(A.in) + (arow_ok ? arow : 0) * (long long)A.nC;
    const float *wrow = kW && A.w ? A.w + (arow_ok ? arow % A.wrows : 0) * (long long)A.nC : nullptr;
    const float *w2row =
        kW && A.w2 ? A.w2 + (arow_ok ? (arow / A.w2div) % A.w2rows : 0) * (long long)A.nC : nullptr;
    const bool vec = ((size_t)A.nC * sizeof(TI)) % 16 == 0 && reinterpret_cast<uintptr_t>(A.in) % 16 == 0;
    // weight rows 16-byte aligned (FP32 rows of nC, bases from the allocator)
    const bool wvec = A.nC % 4 == 0 && reinterpret_cast<uintptr_t>(A.w) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(A.w2) % 16 == 0;
    // B from the tap table: per K step the window of kTW taps (each plane, each skewed
    // copy) from which column c of H reads h[k0 + k - c], k < 32; the window starts at
    // u0 + k0, and each lane's B fragment of tile t sits at a fixed offset in it
    const int dlo = ramp_table_dlo(A.Np);
    const int u0 = (-c0 - (BN - 1) - dlo) & ~7;  // >= 8: c0 <= Np - BN
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    int boff[NT];  // Ts element offset of the fragment h[s .. s + 7], s = k0 + fk - c
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int s0 = fk - (c0 + wc * (BN / 2) + 16 * t + fr);  // s at k0 = 0
        const int r = (-s0) & (kRampSkew - 1);               // copy whose row holds it aligned
        boff[t] = r * kTW + (s0 - dlo + r - u0);             // in [0, kTW - 8], a multiple of 8
    }

    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[i][t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < A.Kp; k0 += kBK) {
        // ---- stage A: 16 consecutive k of one row, weighted and split
        {
            float x[16];
            const int kb = k0 + kh;
            if (vec && arow_ok && kb + 16 <= A.nC) {
                if constexpr (sizeof(TI) == 4) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float4 u = *reinterpret_cast<const float4 *>(prow + kb + 4 * q);
                        x[4 * q] = u.x, x[4 * q + 1] = u.y, x[4 * q + 2] = u.z, x[4 * q + 3] = u.w;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const uint4 u = *reinterpret_cast<const uint4 *>(prow + kb + 8 * q);
                        const uint32_t d[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            x[8 * q + 2 * j] = __uint_as_float(d[j] << 16);
                            x[8 * q + 2 * j + 1] = __uint_as_float(d[j] & 0xffff0000u);
                        }
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < 16; ++j) x[j] = (arow_ok && kb + j < A.nC) ? widen(prow[kb + j]) : 0.0f;
            }
            // per-sample weights (kW): the product is formed where the split reads it, as
            // one expression
            float wt[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) wt[j] = 1.0f;
            if constexpr (kW) {
                if (wvec && arow_ok && kb + 16 <= A.nC) {
                    // 16-byte loads: a wave's scalar weight loads each touch 32 rows' cache
                    // lines, and 16 of them per step bound the kernel by its memory
                    // instruction rate
                    if (wrow) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const float4 u = *reinterpret_cast<const float4 *>(wrow + kb + 4 * q);
                            wt[4 * q] *= u.x, wt[4 * q + 1] *= u.y, wt[4 * q + 2] *= u.z, wt[4 * q + 3] *= u.w;
                        }
                    }
                    if (w2row) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const float4 u = *reinterpret_cast<const float4 *>(w2row + kb + 4 * q);
                            wt[4 * q] *= u.x, wt[4 * q + 1] *= u.y, wt[4 * q + 2] *= u.z, wt[4 * q + 3] *= u.w;
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const bool ok = arow_ok && kb + j < A.nC;
                        float w1 = ok ? 1.0f : 0.0f;
                        if (wrow) w1 *= ok ? wrow[kb + j] : 0.0f;
                        if (w2row) w1 *= ok ? w2row[kb + j] : 0.0f;
                        wt[j] = w1;
                    }
                }
            }
            bf16x8 h[2], l[2], m[2];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if constexpr (kExactA) {
                    h[j >> 3][j & 7] = (__bf16)x[j];
                } else {
                    __bf16 a, b, c;
                    split3(kW ? x[j] * wt[j] : x[j], a, b, c);
                    h[j >> 3][j & 7] = a, l[j >> 3][j & 7] = b, m[j >> 3][j & 7] = c;
                }
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                *reinterpret_cast<bf16x8 *>(&As[0][srow][kh + 8 * q]) = h[q];
                if constexpr (!kExactA) {
                    *reinterpret_cast<bf16x8 *>(&As[1][srow][kh + 8 * q]) = l[q];
                    *reinterpret_cast<bf16x8 *>(&As[2][srow][kh + 8 * q]) = m[q];
                }
            }
        }
        // ---- stage B: the tap window of this K step (3 planes x 8 copies x kTW taps)
        for (int q = tid; q < 3 * kRampSkew * (kTW / 8); q += 256) {
            const int row = q / (kTW / 8), ch = q % (kTW / 8);  // row = plane * 8 + copy
            *reinterpret_cast<bf16x8 *>(&Ts[0][0][0] + row * kTW + 8 * ch) =
                *reinterpret_cast<const bf16x8 *>(A.hp + (size_t)row * A.Lg + (u0 + k0) + 8 * ch);
        }
        __syncthreads();

        // ---- multiply this wave's 64 x 64 quadrant
        bf16x8 ah[4], al[4], all[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = wr * 64 + 16 * i + fr;
            ah[i] = *reinterpret_cast<const bf16x8 *>(&As[0][r][fk]);
            if constexpr (!kExactA) {
                al[i] = *reinterpret_cast<const bf16x8 *>(&As[1][r][fk]);
                all[i] = *reinterpret_cast<const bf16x8 *>(&As[2][r][fk]);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&Ts[0][0][0] + boff[t]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&Ts[1][0][0] + boff[t]);
            const bf16x8 bll = *reinterpret_cast<const bf16x8 *>(&Ts[2][0][0] + boff[t]);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (!kExactA) {
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(all[i], bh, acc[i][t], 0, 0, 0);
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bl, acc[i][t], 0, 0, 0);
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][t], 0, 0, 0);
                }
                acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bll, acc[i][t], 0, 0, 0);
                acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][t], 0, 0, 0);
                acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][t], 0, 0, 0);
            }
        }
        __syncthreads();  // the next K step overwrites the planes
    }

    // ---- store: D lane l holds rows 4 (l / 16) + j, column l % 16 of each 16 x 16 tile
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int col = c0 + wc * (BN / 2) + 16 * t + (lane & 15);
        if (col >= A.nC) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long row = r0 + wr * 64 + 16 * i + 4 * (lane >> 4) + j;
                if (row >= A.rows) continue;
                A.out[row * A.nC + col] = acc[i][t][j];
            }
    }
}

template <typename TI, bool kW>
void launch_gemm_p(int bn, dim3 grid, hipStream_t stream, const MfmaArgs &A)
{
    if (bn == 256) hipLaunchKernelGGL((ramp_gemm_kernel<TI, kW, 256>), grid, dim3(256), 0, stream, A);
    else hipLaunchKernelGGL((ramp_gemm_kernel<TI, kW, 128>), grid, dim3(256), 0, stream, A);
}

}  // namespace

fbct_status launch_ramp_gemm(const RampOperand &H, const void *in, fbct_dtype ti, float *out, int64_t rows,
                             const RowWeights &rw, hipStream_t stream)
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
    A.Lg = H.Lg;
    A.wrows = rw.w ? rw.wrows : 1;
    A.w2div = rw.w2 ? rw.w2div : 1;
    A.w2rows = rw.w2 ? rw.w2rows : 1;
    if (H.Np % kRampTileN != 0 || H.Np < H.nC || H.Kp % kBK != 0 || H.Kp < H.nC || H.Lg != ramp_table_len(H.Np, H.Kp))
        return fail(FBCT_ERR_UNSUPPORTED, "ramp filter: operand planes not padded to the tile");
    // 256-column tiles (the input rows are read by half as many workgroups) unless they
    // pad the columns further than 128-column tiles do
    const int bn = (H.nC + 255) / 256 * 256 == (H.nC + 127) / 128 * 128 ? 256 : 128;
    const long long gx = (rows + kBM - 1) / kBM;
    if (gx > 0x7fffffffLL || gx * ((H.nC + bn - 1) / bn) > 0xffffffffLL)
        return fail(FBCT_ERR_UNSUPPORTED, "ramp filter: too many rows");
    dim3 grid((unsigned)gx, (H.nC + bn - 1) / bn);
    if (ti == FBCT_BF16) {
        if (w) launch_gemm_p<uint16_t, true>(bn, grid, stream, A);
        else launch_gemm_p<uint16_t, false>(bn, grid, stream, A);
    } else {
        if (w) launch_gemm_p<float, true>(bn, grid, stream, A);
        else launch_gemm_p<float, false>(bn, grid, stream, A);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "ramp_gemm_kernel launch: %s", hipGetErrorString(e));
    return FBCT_OK;
}

fbct_status launch_ramp_filter_mfma(const fbct_plan &p, const DeviceTables &t, const float *in, float *out,
                                    int64_t rows, hipStream_t stream)
{
    RampOperand H{t.ramp_planes, p.g.nCols, p.d.ramp_np, p.d.ramp_kp, p.d.ramp_lg};
    return launch_ramp_gemm(H, in, FBCT_F32, out, rows, RowWeights{}, stream);
}

}  // namespace fbct

// fp.hip -- ray-driven fan-beam forward projection (gfx950); its matched adjoint is in
// fp_adjoint.hip.
//
// Reference: GenerateRayTraceMap_FB_ReconImage_OneView_kernel (FP/transform.cu:64-101)
// builds per-ray start/step/N; RayTrace_ReconImage_OneView_kernel (:103-146) sums
// N bilinear texture fetches per ray at 0.125-pixel steps and scales by 0.125*dx.
//
// Design here:
//   * per-ray (u0, w0, du, dw, N) come from the plan's FP64 table;
//   * a workgroup owns a bundle of 256/G adjacent rays of one view and marches them
//     slab by slab along the axis the bundle is closest to;  for every slab it stages
//     the exact bounding box of the taps its samples touch into LDS (clamped to the
//     image edge, which reproduces the clamp-to-edge texture address mode);
//   * G lanes share a ray, each takes the ray's samples in a fixed share of the slab's
//     major coordinate (ending on texel boundaries) and walks it texel cell by texel
//     cell: the bilinear weights of a cell's samples are summed in closed form and its
//     four taps are read once (a ray spends 6-8 samples in a cell at the 1/8-pixel step);
//   * the plan orders the (view, bundle) work items by the size of their largest tile;
//     the items whose tile fits the LDS share of kFpWgsPerCu workgroups per CU run in a
//     launch with that smaller tile, the others in a launch with the full one;
//   * FP32 partial sums per slab, FP64 across slabs; one cross-lane reduction per ray at
//     the end and one coalesced store (optionally minus a target: the FP residual).
// A sample belongs to the slab containing its major coordinate, decided from the same
// FP32 position the interpolation uses, so every sample is counted exactly once.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>

#include "dtype.hpp"
#include "fbct_internal.hpp"

namespace fbct {
namespace {

// Bundle configuration of a lane count G (kFpCfgRays[cfg] = kFpThreads / G) and its LDS
// row pitch: taps [alo, alo + slab] of a slab.
constexpr int fp_cfg(int G) { return kFpThreads / G == kFpCfgRays[0] ? 0 : 1; }


struct FpArgs {
    const void *img;   // TI[n][nY][nX]
    void *sino;        // TO[n][nViews][nCols]
    const float4 *ray_geo;
    const int32_t *ray_n;
    const uint8_t *major;
    const short2 *slabs;   // [nViews*nb] first and last slab with samples
    const short2 *rows;    // [nViews*nb*nslab] tile rows [lo, hi] (lo > hi: empty)
    long long img_stride;
    long long sino_stride;
    int nX, nY, nViews, nCols, nb, nSlices;
    int nslab, cap;
    float scale;
    const float *target;  // optional FP32 [n][nViews][nCols]: store FP - target
};

// S interleaved slices of one tap: a native vector so one LDS read fetches all of them
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <int S> struct Taps;
template <> struct Taps<1> { using T = float; };
template <> struct Taps<2> { using T = f32x2; };
template <> struct Taps<4> { using T = f32x4; };

template <int S>
__device__ __forceinline__ float lane_of(const typename Taps<S>::T &t, int s)
{
    if constexpr (S == 1) return t;
    else return t[s];
}

// Smallest k in [0, N] with P(k) true, where P(k) = (a0 + k da >= t) for da >= 0 and
// (a0 + k da < t) for da < 0, evaluated with the same fmaf the samples use (monotone
// in k, so P is false...false true...true).  nia = -1/da estimates k; the two scans
// make it exact.
__device__ __forceinline__ int slab_boundary(float a0, float da, float nia, int N, float t)
{
    // selects rather than branches: the lanes of a wave differ in direction
    const bool asc = da >= 0.0f;
    const float q = (a0 - t) * nia;  // (t - a0) / da, with nia = -1/da
    float est = asc ? ceilf(q) : floorf(q) + 1.0f;
    est = da == 0.0f ? ((a0 >= t) ? 0.0f : (float)N) : est;
    est = fminf(fmaxf(est, 0.0f), (float)max(N, 0));
    int k = (int)est;
    // P(k) as one comparison: a >= t for da >= 0, a < t otherwise (a is finite)
    auto pred = [&](int kk) { return (fmaf((float)kk, da, a0) < t) != asc; };
    // one step either way in practice (the estimate's FP32 error is far below a sample);
    // neither unrolled nor interleaved, so that the common exit costs one evaluation of
    // the predicate each
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (k > 0 && pred(k - 1)) --k;
#pragma clang loop unroll(disable) interleave(disable) vectorize(disable)
    while (k < N && !pred(k)) ++k;
    return k;
}

// S slices that share the geometry are projected together: the LDS tile interleaves
// them ([minor][major][slice]) so one sample position, its weights and its LDS
// address serve S slices, with S-wide LDS reads.
// Dynamic LDS of the FP kernels: the slab tile at address 0 (so tap offsets need no
// base add).
extern __shared__ float fp_lds[];

// Bundle b of view v for slices [s0, s0+S): every thread of the workgroup calls it with
// the same arguments (it contains barriers).
template <int S, int G, typename TI, typename TO>
__device__ __forceinline__ void fp_bundle(const FpArgs &A, int b, int v, int s0)
{
    using TapT = typename Taps<S>::T;
    constexpr int kSlab = kFpCfgSlab[fp_cfg(G)];
    constexpr int kTaps = kSlab + 1;         // staged taps per row: [alo, alo + kSlab]
    constexpr int kPitch = fp_pitch(fp_cfg(G));  // row stride of the tile, in taps
    float *tile = fp_lds;

    const int tid = threadIdx.x;
    constexpr int kRays = kFpThreads / G;
    // Share-major lanes: a wave holds kRw rays; lane w of it is ray w % kRw, share
    // w / kRw.  The 16 lanes of one LDS pass then read kRw rays at one or two shares,
    // whose rows differ, and with the tile's row pitch of 25 taps (9 mod 16 in 16-byte
    // bank groups) distinct rows fall on distinct bank groups; with the G shares of a ray
    // in one pass instead, shares three taps apart in adjacent rows collide.
    constexpr int kRw = 64 / G;
    const int wl = tid & 63;
    const int lane = wl / kRw;  // the ray's share
    const int c = b * kRays + (tid >> 6) * kRw + wl % kRw;
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
    // cell walk constants: reciprocal steps (a zero step never crosses: a large finite
    // value) and, with the boundary offset o of a cell in the walking direction (1 or 0),
    // the terms of (o - t) / step + 1 = t (-1/step) + (o / step + 1) for a fraction t
    const float ia = da != 0.0f ? 1.0f / da : 1.0e30f;
    const float im = dm != 0.0f ? 1.0f / dm : 1.0e30f;
    const float nia = -ia, nim = -im;
    const float ca = da >= 0.0f ? ia + 1.0f : 1.0f;
    const float cm = dm >= 0.0f ? im + 1.0f : 1.0f;
    const TI *img[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        img[s] = static_cast<const TI *>(A.img) + (size_t)min(s0 + s, A.nSlices - 1) * A.img_stride;
    const int na = xmajor ? A.nX : A.nY;  // image extent along the major axis
    const int nm = xmajor ? A.nY : A.nX;

    // residual mode: the targets are loaded now, so their latency hides behind the march
    float tgt[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        tgt[s] = (A.target && lane == 0 && c < A.nCols)
                     ? A.target[(size_t)min(s0 + s, A.nSlices - 1) * A.sino_stride + (size_t)v * A.nCols + c]
                     : 0.0f;
    // FP32 sums within a slab, FP64 across slabs
    double accd[S];
#pragma unroll
    for (int s = 0; s < S; ++s) accd[s] = 0.0;
    // slabs of the bundle with samples, and per slab the tile's minor rows (plan tables)
    const size_t vb = (size_t)v * A.nb + b;
    const short2 srange = A.slabs[vb];
    const short2 *rows = A.rows + vb * A.nslab;
    // lane l of a ray takes the samples whose major coordinate lies in
    // [alo + l w, alo + (l + 1) w), w = kSlab / G: it finds the end of its share and takes
    // the start from lane l - 1 (lane 0: the previous slab's end), so the G shares
    // partition the slab's samples and a share ends on a texel boundary whenever l w is
    // an integer
    constexpr float kShare = (float)kSlab / (float)G;
    int kcarry = slab_boundary(a0, da, nia, N, (float)(srange.x * kSlab - 1));
    // byte strides of the interleaved tile
    constexpr float kColB = 4.0f * S;
    constexpr float kRowB = 4.0f * S * kPitch;
    // LDS byte address of the tile (the dynamic-LDS base, folded into obase below)
    const float tbase = (float)(unsigned)(size_t)((__attribute__((address_space(3))) float *)(tile));
    using LdsTap = __attribute__((address_space(3))) const TapT;

    for (int sl = srange.x; sl <= srange.y; ++sl) {
        const int alo = sl * kSlab - 1;
        const short2 rr = rows[sl];  // uniform
        const int mlo = rr.x, mhi = rr.y;
        const float tend = lane == G - 1 ? (float)(alo + kSlab) : fmaf((float)(lane + 1), kShare, (float)alo);
        const int kown = slab_boundary(a0, da, nia, N, tend);
        // share l - 1's end (kRw lanes below; share 0 of a ray takes the carry instead),
        // and for the next slab's share 0 the end of share G - 1
        int kstart = __shfl(kown, max(wl - kRw, 0));
        if (lane == 0) kstart = kcarry;
        kcarry = __shfl(kown, wl % kRw + kRw * (G - 1));
        if (mlo > mhi) continue;  // no samples of this bundle in the slab
        const int M = mhi - mlo + 1;
        // M <= A.cap for every item of a launch: the plan (and the batch) check their row
        // tables against the caps the launches use when they are built; kept so that LDS
        // is never addressed past the tile whatever the tables hold
        if (M > A.cap) continue;
        __syncthreads();  // the previous slab's reads of `tile` are done

        // Stage taps [alo, alo+kSlab] x [mlo, mhi] of the S slices with clamp-to-edge
        // addressing.  Global loads run along the image row (i) in both orientations, at
        // 32-bit byte offsets from the slices' bases; each thread keeps one fixed column
        // (x-major) or minor lane (y-major), so the index split is done once.
        auto stage = [&](int aa, int mm, unsigned go) {
            TapT val;
            float *vp = reinterpret_cast<float *>(&val);
#pragma unroll
            for (int s = 0; s < S; ++s) vp[s] = load_f_at(img[s], go * (unsigned)sizeof(TI));
            reinterpret_cast<TapT *>(tile)[mm * kPitch + aa] = val;
        };
        // A tile inside the image (most of them) needs no clamping: its addresses then
        // advance by a constant per pass (uniform choice per slab).
        auto stage_tile = [&](auto clamp) {
            constexpr bool kClamp = decltype(clamp)::value;
            auto ca = [&](int a) { return kClamp ? min(max(a, 0), na - 1) : a; };
            auto cm = [&](int m) { return kClamp ? min(max(m, 0), nm - 1) : m; };
            if (xmajor) {  // a = i: kRowsPerPass rows of kTaps taps per pass of the workgroup
                constexpr int kRowsPerPass = kFpThreads / kTaps;
                if (tid < kRowsPerPass * kTaps) {
                    const int aa = tid % kTaps;
                    const unsigned ga = (unsigned)ca(alo + aa);
#pragma unroll 2  // two taps' loads in flight per thread
                    for (int mm = tid / kTaps; mm < M; mm += kRowsPerPass)
                        stage(aa, mm, (unsigned)cm(mlo + mm) * (unsigned)A.nX + ga);
                }
            } else {  // a = j: 64 minor taps per wave, one major row per wave and pass
                const int wave = tid >> 6;
                for (int m0 = 0; m0 < M; m0 += 64) {
                    const int mm = m0 + (tid & 63);
                    if (mm >= M) continue;
                    const unsigned gm = (unsigned)cm(mlo + mm);
#pragma unroll 2
                    for (int aa = wave; aa < kTaps; aa += kFpThreads / 64)
                        stage(aa, mm, (unsigned)ca(alo + aa) * (unsigned)A.nX + gm);
                }
            }
        };
        if (alo >= 0 && alo + kSlab <= na - 1 && mlo >= 0 && mhi <= nm - 1) stage_tile(std::false_type{});
        else stage_tile(std::true_type{});
        __syncthreads();  // tile ready

        // byte offset of tap (fa, fm) = (fm - mlo) * kRowB + (fa - alo) * kColB, exact in FP32
        const float obase = tbase - ((float)mlo * kRowB + (float)alo * kColB);
        // A lane's share of the slab is [kl, kend) (major-coordinate split above).  All
        // samples inside one texel cell (floor a, floor m) read the same four taps, so a lane
        // walks its share cell by cell: the n samples of a cell have fractions ta + j da and
        // tm + j dm (j < n), their bilinear weights are summed in closed form and the taps
        // are read and weighted once per cell.  A sample on a cell boundary may be counted
        // in either cell: the bilinear interpolant is continuous there.
        const int kl = asc ? kstart : kown;
        const float kend = (float)(asc ? kown : kstart);
        float acc[S];
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] = 0.0f;
        for (float kf = (float)kl; kf < kend;) {
            const float a = fmaf(kf, da, a0);
            const float m = fmaf(kf, dm, m0);
            const float fa = floorf(a), fm = floorf(m);
            const float ta = a - fa, tm = m - fm;
            // samples before the next boundary along each axis: floor(distance / step) + 1
            // with distance (fa + o) - a = o - ta, one FMA each; the three counts share one
            // min.  A count that rounding puts one past a boundary only moves a sample lying
            // on it (within rounding) into this cell, where its bilinear weights are the same.
            const float ja1 = floorf(fmaf(ta, nia, ca));
            const float jm1 = floorf(fmaf(tm, nim, cm));
            const float n = __builtin_amdgcn_fmed3f(fminf(ja1, jm1), -1.0e30f, kend - kf);
            const float t1 = 0.5f * fmaf(n, n, -n);                        // sum of j
            const float t2 = t1 * fmaf(n, 2.0f / 3.0f, -1.0f / 3.0f);      // sum of j^2
            const float sa = fmaf(da, t1, n * ta);                          // sum of a-fractions
            const float sm = fmaf(dm, t1, n * tm);                          // sum of m-fractions
            // sum of products: sum_j (ta + j da)(tm + j dm) = ta sm + da (tm t1 + dm t2)
            const float sam = fmaf(ta, sm, da * fmaf(tm, t1, dm * t2));
            const float w11 = sam, w01 = sa - sam, w10 = sm - sam, w00 = (n - sa) - w10;
            const unsigned o = (unsigned)fmaf(fm, kRowB, fmaf(fa, kColB, obase));
            const TapT t00 = *reinterpret_cast<LdsTap *>((size_t)o);
            const TapT t01 = *reinterpret_cast<LdsTap *>((size_t)(o + (unsigned)kColB));
            const TapT t10 = *reinterpret_cast<LdsTap *>((size_t)(o + (unsigned)kRowB));
            const TapT t11 = *reinterpret_cast<LdsTap *>((size_t)(o + (unsigned)(kRowB + kColB)));
#pragma unroll
            for (int s = 0; s < S; ++s)
                acc[s] = fmaf(w00, lane_of<S>(t00, s),
                              fmaf(w01, lane_of<S>(t01, s),
                                   fmaf(w10, lane_of<S>(t10, s), fmaf(w11, lane_of<S>(t11, s), acc[s]))));
            kf += n;
        }
#pragma unroll
        for (int s = 0; s < S; ++s) accd[s] += (double)acc[s];
    }

#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int off = 32; off >= kRw; off >>= 1) accd[s] += __shfl_xor(accd[s], off);
        if (lane == 0 && c < A.nCols && s0 + s < A.nSlices) {
            const size_t o = (size_t)(s0 + s) * A.sino_stride + (size_t)v * A.nCols + c;
            const float p = (float)(accd[s] * (double)A.scale);
            store_f(static_cast<TO *>(A.sino), o, A.target ? p - tgt[s] : p);
        }
    }
}

// Items order[item0 + t] (t < gridDim.x) of slice group blockIdx.y; item = v * nb + b.
template <int S, int G, typename TI, typename TO>
__global__ __launch_bounds__(kFpThreads) void fp_kernel(FpArgs A, const int32_t *__restrict__ order, int item0)
{
    unsigned t, g;  // XCD-aware (item, slice group)
    xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x, gridDim.y, t, g);
    const int vb = order[item0 + (int)t];
    fp_bundle<S, G, TI, TO>(A, vb % A.nb, vb / A.nb, g * S);
}

// Uneven batch: slice s has its own geometry (desc[s]); work item = (slice, view,
// bundle), items of slice s are [prefix[s], prefix[s+1]).  A launch takes the items
// order[0, count) (a class of the batch's items ordered by tile size, all fitting `cap`
// rows); a fixed grid of workgroups strides over them (order[blockIdx.x + k gridDim.x]).
template <int G, typename TI, typename TO>
__global__ __launch_bounds__(kFpThreads) void fp_batch_kernel(const FpSliceDesc *__restrict__ desc,
                                                              const int *__restrict__ prefix, int n,
                                                              const int32_t *__restrict__ order, int count, int cap,
                                                              const TI *img, TO *sino)
{
    constexpr int cfg = fp_cfg(G);
    for (int t = blockIdx.x; t < count; t += gridDim.x) {
        const int item = order[t];
        int lo = 0, hi = n - 1;  // last slice with prefix[s] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const FpSliceDesc &d = desc[lo];
        FpArgs A;
        A.target = nullptr;
        A.img = img + d.img_off;
        A.sino = sino + d.sino_off;
        // the slice's tables as workgroup-uniform global-memory pointers (read from the
        // descriptor they would be generic, per-lane addresses: flat loads)
        A.ray_geo = uniform_global(d.ray_geo);
        A.ray_n = uniform_global(d.ray_n);
        A.major = uniform_global(d.major[cfg]);
        A.slabs = uniform_global(d.slabs[cfg]);
        A.rows = uniform_global(d.rows[cfg]);
        A.img_stride = 0;
        A.sino_stride = 0;
        A.nX = d.nX;
        A.nY = d.nY;
        A.nViews = d.nViews;
        A.nCols = d.nCols;
        A.nb = d.nb[cfg];
        A.nSlices = 1;
        A.nslab = d.nslab[cfg];
        A.cap = min(d.cap[cfg], cap);  // the launch's tile; the items it is given fit it
        A.scale = d.scale;
        const int local = item - prefix[lo];
        fp_bundle<1, G, TI, TO>(A, local % A.nb, local / A.nb, 0);
        __syncthreads();  // the next item reuses the LDS tile
    }
}

}  // namespace

template <int S, int G, typename TI, typename TO>
static void launch_fp_s(dim3 grid, size_t lds, hipStream_t stream, const FpArgs &A, const int32_t *order, int item0)
{
    hipLaunchKernelGGL((fp_kernel<S, G, TI, TO>), grid, dim3(kFpThreads), lds, stream, A, order, item0);
}

template <int S, int G>
static fbct_status launch_fp_t(fbct_dtype ti, fbct_dtype to, dim3 grid, size_t lds, hipStream_t stream,
                               const FpArgs &A, const int32_t *order, int item0)
{
    if (ti == FBCT_F32 && to == FBCT_F32) launch_fp_s<S, G, float, float>(grid, lds, stream, A, order, item0);
    else if (ti == FBCT_BF16 && to == FBCT_F32) launch_fp_s<S, G, bf16_t, float>(grid, lds, stream, A, order, item0);
    else if (ti == FBCT_F32 && to == FBCT_BF16) launch_fp_s<S, G, float, bf16_t>(grid, lds, stream, A, order, item0);
    else if (ti == FBCT_BF16 && to == FBCT_BF16) launch_fp_s<S, G, bf16_t, bf16_t>(grid, lds, stream, A, order, item0);
    else return fail(FBCT_ERR_INVALID_ARGUMENT, "fbct_fp: unknown storage type");
    return FBCT_OK;
}

fbct_status launch_fp(const fbct_plan &p, const DeviceTables &t, const void *img, fbct_dtype ti, void *sino,
                      fbct_dtype to, int n, hipStream_t stream, const float *target)
{
    FpArgs A;
    A.target = target;
    A.img = img;
    A.sino = sino;
    A.ray_geo = t.ray_geo;
    A.ray_n = t.ray_n;
    A.img_stride = (long long)p.g.nX * p.g.nY;
    A.sino_stride = (long long)p.g.nViews * p.g.nCols;
    A.nX = p.g.nX;
    A.nY = p.g.nY;
    A.nViews = p.g.nViews;
    A.nCols = p.g.nCols;
    A.scale = p.d.fp_scale;
    A.nSlices = n;
    hipError_t e;

    // Slices per workgroup S and bundle configuration: share each sample's position and
    // weights across 4 slices with 32-ray bundles (smaller tile), else 2 or 1 slices with
    // 64-ray bundles, while the grid keeps >= 2048 workgroups and the tile <= 80 KB.
    const size_t kMaxTile = 80 * 1024;
    auto tile_bytes = [&](int cfg, int S) { return (size_t)fp_pitch(cfg) * p.d.fpb[cfg].cap * sizeof(float) * S; };
    auto grid_wgs = [&](int cfg, int S) { return (long long)p.d.fpb[cfg].nb * p.g.nViews * ((n + S - 1) / S); };
    int S = 1, cfg = 0;
    if (n >= 4 && tile_bytes(1, 4) <= kMaxTile && grid_wgs(1, 4) >= 2048) {
        S = 4;
        cfg = 1;
    } else if (n >= 2 && grid_wgs(0, 2) >= 2048) {
        S = 2;
        cfg = tile_bytes(0, 2) <= kMaxTile ? 0 : 1;
    } else if (tile_bytes(0, 1) > kMaxTile) {
        cfg = 1;
    }
    const FpBundleCfg &bc = p.d.fpb[cfg];
    A.major = t.fp_major[cfg];
    A.slabs = t.fp_slabs[cfg];
    A.rows = t.fp_rows[cfg];
    A.nb = bc.nb;
    A.nslab = bc.nslab;
    if (tile_bytes(cfg, S) > 160 * 1024)
        return fail(FBCT_ERR_UNSUPPORTED, "fbct_fp: slab tile of %zu bytes exceeds LDS", tile_bytes(cfg, S));
    // The LDS tile sets how many workgroups a CU holds (up to kFpWgsPerCu, where the
    // registers cap it).  Items whose largest tile fits the rows of that many workgroups
    // get a launch of their own with that smaller tile; the others (ordered first) a
    // launch with the plan's full tile.
    const size_t row_bytes = (size_t)fp_pitch(cfg) * sizeof(float) * S;
    const int small_cap = (int)std::min<size_t>((size_t)bc.cap, (160 * 1024 / kFpWgsPerCu) / row_bytes);
    const std::vector<int16_t> &orows = p.fp_order_rows[cfg];
    const int nitems = (int)orows.size();
    // first item with at most small_cap rows (counts are non-increasing)
    const int split = (int)(std::upper_bound(orows.begin(), orows.end(), (int16_t)small_cap,
                                             [](int16_t cap_v, int16_t r) { return r <= cap_v; }) -
                            orows.begin());
    const unsigned groups = (unsigned)((n + S - 1) / S);
    for (int part = 0; part < 2; ++part) {
        const int item0 = part == 0 ? 0 : split;
        const int count = part == 0 ? split : nitems - split;
        if (count == 0) continue;
        A.cap = part == 0 ? bc.cap : small_cap;
        const size_t lds = (size_t)A.cap * row_bytes;
        const dim3 grid((unsigned)count, groups);
        fbct_status st;
        if (cfg == 0) {
            if (S == 2) st = launch_fp_t<2, 4>(ti, to, grid, lds, stream, A, t.fp_order[cfg], item0);
            else st = launch_fp_t<1, 4>(ti, to, grid, lds, stream, A, t.fp_order[cfg], item0);
        } else {
            if (S == 4) st = launch_fp_t<4, 8>(ti, to, grid, lds, stream, A, t.fp_order[cfg], item0);
            else if (S == 2) st = launch_fp_t<2, 8>(ti, to, grid, lds, stream, A, t.fp_order[cfg], item0);
            else st = launch_fp_t<1, 8>(ti, to, grid, lds, stream, A, t.fp_order[cfg], item0);
        }
        if (st != FBCT_OK) return st;
        e = hipGetLastError();
        if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_kernel launch: %s", hipGetErrorString(e));
    }
    return FBCT_OK;
}

template <int G, typename TI, typename TO>
static void launch_fp_batch_t(dim3 grid, size_t lds, hipStream_t stream, const fbct_batch &b, const BatchDevice &bd,
                              const int32_t *order, int count, int cap, const void *img, void *sino)
{
    constexpr int cfg = fp_cfg(G);
    const int n = (int)b.plans.size();
    hipLaunchKernelGGL((fp_batch_kernel<G, TI, TO>), grid, dim3(kFpThreads), lds, stream, bd.fp_desc,
                       bd.fp_prefix[cfg], n, order, count, cap, static_cast<const TI *>(img),
                       static_cast<TO *>(sino));
}

fbct_status launch_fp_batch(const fbct_batch &b, const BatchDevice &bd, const void *img, fbct_dtype ti, void *sino,
                            fbct_dtype to, hipStream_t stream)
{
    const int n = (int)b.plans.size();
    const int cfg = b.shape.fp_rays == 32 ? 1 : 0;
    const int total = b.fp_prefix[cfg][n];
    if (total == 0) return FBCT_OK;
    hipError_t e;
    const size_t row_bytes = (size_t)fp_pitch(cfg) * sizeof(float);
    if ((size_t)b.fp_cap_max[cfg] * row_bytes > 64 * 1024)
        return fail(FBCT_ERR_UNSUPPORTED, "fbct_batch_fp: slab tile of %zu bytes exceeds 64 KiB",
                    (size_t)b.fp_cap_max[cfg] * row_bytes);
    // items that fit the tile of kFpWgsPerCu workgroups per CU run in a launch of their
    // own (they follow the larger ones in the order); fp_wgs_per_cu = 0: one workgroup
    // per item, else a grid of that many per CU striding over the items
    const int split = b.fp_split[cfg], small_cap = b.fp_small_cap[cfg];
    for (int part = 0; part < 2; ++part) {
        const int item0 = part == 0 ? 0 : split;
        const int count = part == 0 ? split : total - split;
        if (count == 0) continue;
        const int cap = part == 0 ? b.fp_cap_max[cfg] : small_cap;
        const size_t lds = (size_t)cap * row_bytes;
        const long long wgs = b.shape.fp_wgs_per_cu == 0
                                  ? count
                                  : std::min<long long>(count, (long long)b.shape.fp_wgs_per_cu * bd.num_cu);
        const dim3 grid((unsigned)wgs);
        const int32_t *order = bd.fp_order[cfg] + item0;
#define FBCT_FPB(G)                                                                                       \
    do {                                                                                                  \
        if (ti == FBCT_F32 && to == FBCT_F32) launch_fp_batch_t<G, float, float>(grid, lds, stream, b, bd, order, count, cap, img, sino); \
        else if (ti == FBCT_BF16 && to == FBCT_F32) launch_fp_batch_t<G, bf16_t, float>(grid, lds, stream, b, bd, order, count, cap, img, sino); \
        else if (ti == FBCT_F32 && to == FBCT_BF16) launch_fp_batch_t<G, float, bf16_t>(grid, lds, stream, b, bd, order, count, cap, img, sino); \
        else launch_fp_batch_t<G, bf16_t, bf16_t>(grid, lds, stream, b, bd, order, count, cap, img, sino);          \
    } while (0)
        if (cfg == 0) FBCT_FPB(4);
        else FBCT_FPB(8);
#undef FBCT_FPB
        e = hipGetLastError();
        if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "fp_batch_kernel launch: %s", hipGetErrorString(e));
    }
    return FBCT_OK;
}

}  // namespace fbct

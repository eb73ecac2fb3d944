// fbct_internal.hpp -- internal types shared by the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fbct.h"

namespace fbct {

// FP ray-march tiling (see fp.hip).  Bundle configurations: rays per workgroup (lanes
// per ray = kFpThreads / rays) and slab width in pixels along the major axis (a multiple
// of the lanes per ray, so every lane's share of a slab ends on a texel boundary).  The
// 32-ray, 24-pixel configuration serves the 4-slice tiles: its smaller LDS tile keeps
// five to seven workgroups resident per CU.
constexpr int kFpThreads = 256;
// Workgroups per CU the FP kernels can hold by their registers (4-slice tile kernel: 71
// VGPRs, 7 waves per SIMD); the LDS tile is sized for this many where the work allows.
constexpr int kFpWgsPerCu = 7;
constexpr int kFpNumCfg = 2;
constexpr int kFpCfgRays[kFpNumCfg] = {64, 32};
constexpr int kFpCfgSlab[kFpNumCfg] = {32, 24};
// LDS tile row stride in taps: the slab's kFpCfgSlab + 1 taps and padding (bank spread)
constexpr int kFpCfgPad[kFpNumCfg] = {0, 0};
constexpr int fp_pitch(int cfg) { return kFpCfgSlab[cfg] + 1 + kFpCfgPad[cfg]; }

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each with its own
// L2: the workgroup with linear id b runs on XCD b % 8.  xcd_block maps b to a logical
// block id so that every XCD receives a contiguous range of logical blocks.
constexpr unsigned kNumXcd = 8;

__host__ __device__ inline unsigned xcd_block(unsigned b, unsigned total)
{
    const unsigned q = total / kNumXcd, r = total % kNumXcd;
    const unsigned x = b % kNumXcd, k = b / kNumXcd;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Logical block of workgroup b in a grid of `total`: consecutive chunks of kC logical
// blocks each run on one XCD, the chunks dealt round-robin to the XCDs (blocks past the
// last whole round keep their own id).  Fine-grained enough to balance work items of
// unequal cost across the XCDs, coarse enough that one XCD's resident workgroups take
// neighbouring items.
__host__ __device__ inline unsigned xcd_chunk(unsigned b, unsigned total)
{
    constexpr unsigned kC = 64;
    if (b >= total / (kNumXcd * kC) * (kNumXcd * kC)) return b;
    const unsigned x = b % kNumXcd, k = b / kNumXcd;
    return ((k / kC) * kNumXcd + x) * kC + k % kC;
}

// (tile, slice group) of workgroup b in a grid of `tiles` tiles x `groups` slice groups.
// When the groups divide evenly over the XCDs, each XCD takes whole groups (their inputs
// stay in its L2 and every XCD does the same work); otherwise the hardware order is kept.
__host__ __device__ inline void xcd_tile(unsigned b, unsigned tiles, unsigned groups, unsigned &tile,
                                         unsigned &group)
{
    const unsigned l = groups % kNumXcd == 0 ? xcd_block(b, tiles * groups) : b;
    group = l / tiles;
    tile = l % tiles;
}

// Tile t of a gx x gy grid in blocked order: blocks of kB x kB tiles (block rows top to
// bottom, blocks left to right, partial blocks at the right and bottom edges), in a block
// its tiles row by row.
__host__ __device__ inline void blocked_tile(unsigned t, unsigned gx, unsigned gy, unsigned &tx, unsigned &ty)
{
    constexpr unsigned kB = 8;
    const unsigned br = t / (gx * kB);              // block row (full rows hold gx * kB tiles)
    const unsigned h = gy - br * kB < kB ? gy - br * kB : kB;  // its height
    const unsigned rem = t - br * gx * kB;
    const unsigned bc = rem / (kB * h);             // block in the row (full blocks hold kB * h tiles)
    const unsigned w = gx - bc * kB < kB ? gx - bc * kB : kB;  // its width
    const unsigned r = rem - bc * kB * h;
    ty = br * kB + r / w;
    tx = bc * kB + r % w;
}

// (tile x, tile y, slice group) of workgroup b in a grid of gx x gy tiles and `groups`
// slice groups: each XCD takes a contiguous range of the logical order, which runs over
// the slice groups and in each over the tiles in blocked order.  The workgroups resident
// on one XCD at a time then cover a compact square of the image, whose detector
// footprints in every view overlap, so the columns they stage are shared through that
// XCD's L2.
__host__ __device__ inline void xcd_blocked_tile(unsigned b, unsigned gx, unsigned gy, unsigned groups, unsigned &tx,
                                                 unsigned &ty, unsigned &group)
{
    const unsigned tiles = gx * gy;
    const unsigned l = xcd_block(b, tiles * groups);
    group = l / tiles;
    blocked_tile(l % tiles, gx, gy, tx, ty);
}

// BP atan(z) = z + z^3 R(z^2): at most this many coefficients of R.
constexpr int kBpAtanMax = 9;

struct FpBundleCfg {
    int rays;   // rays per bundle
    int slab;   // slab width (pixels)
    int nslab;  // slabs along the major axis
    int nb;     // bundles per view
    int cap;    // LDS rows reserved for the minor extent of a slab
};

// FP^T pixel tile widths of the uneven-batch kernel (tile height = 256 / width).
constexpr int kAdjNumTiles = 4;
constexpr int kAdjTileX[kAdjNumTiles] = {16, 32, 8, 64};

// FP^T cell tiles: kAdjCellTile^2 texel cells per workgroup yield (kAdjCellTile-1)^2
// pixels.  The cell-driven kernels serve a plan whose samples all lie in the cells
// [-1, n-1] of both axes (adj_cells_fit).
constexpr int kAdjCellTile = 16;
constexpr int adj_cell_tiles(int n) { return (n + kAdjCellTile - 2) / (kAdjCellTile - 1); }
constexpr bool adj_cells_fit(float far_x, float far_y) { return far_x < 1.0f && far_y < 1.0f; }

// Per-slice descriptors of an uneven batch (device resident).  Offsets count elements
// of the storage type from the packed buffers' bases.
struct FpSliceDesc {
    const float4 *ray_geo;
    const int32_t *ray_n;
    const uint8_t *major[kFpNumCfg];
    const short2 *slabs[kFpNumCfg];
    const short2 *rows[kFpNumCfg];
    long long img_off, sino_off;
    int nX, nY, nViews, nCols;
    int nb[kFpNumCfg], cap[kFpNumCfg], nslab[kFpNumCfg];
    float scale;
};

struct FpaSliceDesc {
    const float4 *ray_geo;
    const int32_t *ray_n;
    const float2 *ray_inv;
    const float2 *cs;
    long long img_off, sino_off;
    int nX, nY, nViews, nCols;
    float scale, r, dx, R, c_scale, c_off, c_abs, far_x, far_y;
};

// Device-resident copy of a plan's tables on one HIP device.
struct DeviceTables {
    int device = -1;
    int32_t *ray_n = nullptr;      // [nViews*nCols]
    float4 *ray_geo = nullptr;     // [nViews*nCols] (u0, w0, du, dw)
    float2 *ray_inv = nullptr;     // [nViews*nCols] (1/du, 1/dw)
    uint8_t *fp_major[kFpNumCfg] = {};  // per bundle config: [nViews*nb] 1 = x-major
    short2 *fp_slabs[kFpNumCfg] = {};   // per bundle config: [nViews*nb] first and last slab
    short2 *fp_rows[kFpNumCfg] = {};    // per bundle config: [nViews*nb*nslab] tile rows [lo, hi]
    int32_t *fp_order[kFpNumCfg] = {};  // per bundle config: items v*nb+b by decreasing tile rows
    float2 *view_cs = nullptr;     // [nViews] (cos beta, sin beta)
    double2 *view_csd = nullptr;   // [nViews] (cos beta, sin beta) in FP64
    double2 *col_csd = nullptr;    // [nCols] (cos gamma_c, sin gamma_c) in FP64, gamma_c of the BP column map
    int2 *fov_row = nullptr;       // [nY] pixels i in [x, y] of row j lie in the FOV disk (x > y: none)
    int2 *fov_col = nullptr;       // [nX] pixels j in [x, y] of column i lie in the FOV disk
    double *ramp = nullptr;        // [2*nCols-1]
    uint16_t *ramp_planes = nullptr;  // [3][kRampSkew][Lg] BF16 split planes of the skewed ramp taps
    float *parker = nullptr;       // [nViews][nCols] short-scan weights (uploaded on first use)
    float *cosine = nullptr;       // [nCols] FBP cosine pre-weights cos(gamma_c) (uploaded on first use)
};

// Scalars the kernels need, derived once in FP64 on the host.
struct Derived {
    float fp_scale;       // 0.125 * dx
    float two_r;          // 2 * rFOV/dx (FP sample coordinate extent)
    FpBundleCfg fpb[kFpNumCfg];
    float bp_xi_scale;    // 1/dCol
    float bp_xi_offset;   // dShift - dLeft/dCol
    float r2;             // rFOV*rFOV (float, as the reference)
    float bpa_tan;        // tan |dCol|: BP^T candidate half-width per unit distance from the source
    // In-range decision 0 < xi < nCols taken in FP64 near the two edges: with
    // g = atan(t/U), the pixel is inside iff tan_min < t/U < tan_max.
    int bp_edge_exact;    // 1 when the fan edges lie within (-pi/2, pi/2)
    double bp_tan_min, bp_tan_max;
    int bp_atan_nc;              // coefficients of R in use (5 or 9), 0 = atanf
    float bp_atan_c[kBpAtanMax];
    // FP adjoint (gather)
    float fpa_r;          // rFOV/dx
    float fpa_c_scale;    // 1/dCol
    float fpa_c_off;      // -dLeft/dCol - 1/2
    float fpa_far_x, fpa_far_y;  // 2r - n + 1/2: sample reach beyond the last pixel centre
    int ramp_np, ramp_kp, ramp_lg;  // padded columns / depth of the MFMA filter operand, tap table row
};

}  // namespace fbct

struct fbct_plan {
    fbct_fan_geom g{};
    std::vector<float> angles;

    // host tables (FP64 computed, stored at the precision the kernels read)
    std::vector<int32_t> ray_n;
    std::vector<float> ray_geo;     // 4 per ray
    std::vector<float> ray_inv;     // 2 per ray
    std::vector<uint8_t> fp_major[fbct::kFpNumCfg];
    std::vector<int16_t> fp_slabs[fbct::kFpNumCfg];  // 2 per bundle
    std::vector<int16_t> fp_rows[fbct::kFpNumCfg];   // 2 per (bundle, slab)
    // items (v * nb + b) ordered by decreasing largest tile row count, and those counts:
    // the launch gives the items that fit a smaller tile a launch of their own
    std::vector<int32_t> fp_order[fbct::kFpNumCfg];
    std::vector<int16_t> fp_order_rows[fbct::kFpNumCfg];
    std::vector<float> view_cs;     // 2 per view
    std::vector<double> view_csd;   // 2 per view
    std::vector<double> col_csd;    // 2 per column: cos, sin of gamma_c = dLeft + (c + 1/2 - dShift) dCol
    // FOV disk membership of the pixel centres, x^2 + y^2 <= rFOV^2 in FP64 (the reference
    // tests it in FP32, BP/transform.cu:80-81; the two can differ only for centres within
    // rounding of the circle): per row the range of columns [lo, hi], per column the range
    // of rows (lo > hi: empty)
    std::vector<int32_t> fov_row, fov_col;
    std::vector<double> ramp;       // 2*nCols-1
    std::vector<uint16_t> ramp_planes;  // BF16 hi/lo/ll planes of the skewed taps, [3][kRampSkew][Lg]
    int64_t fp_samples = 0;
    fbct::Derived d{};

    std::mutex mu;
    std::vector<fbct::DeviceTables> dev;
};

namespace fbct {

// A batch's work tables on one device.
struct BatchDevice {
    int device = -1;
    int num_cu = 0;
    FpSliceDesc *fp_desc = nullptr;              // [n]
    FpaSliceDesc *fpa_desc = nullptr;            // [n]
    int *fp_prefix[kFpNumCfg] = {};              // [n+1] work items before slice s
    int32_t *fp_order[kFpNumCfg] = {};           // [items] by decreasing tile rows
    int *fpa_prefix[kAdjNumTiles] = {};          // [n+1]
};

}  // namespace fbct

namespace fbct {

// Scalars of a cone plan's backprojection, derived once in FP64 on the host.
struct ConeDerived {
    float xi_scale, xi_offset;  // flat: xi = (t/U) xi_scale + xi_offset; curved: atan(t/U) xi_scale + xi_offset
    float zr_scale, zr_offset;  // zeta = z (1/U) zr_scale + zr_offset
    float r2;                   // rFOV^2 (float)
    int edge_exact;             // 1: the in-range decision near the edges is retaken in FP64 on t/U
    double tan_min, tan_max;    //    0 < xi < nCols  <=>  tan_min < t/U < tan_max
    int atan_nc;                // curved: coefficients of the atan correction polynomial (0 = atanf)
    float atan_c[kBpAtanMax];
    int ramp_np, ramp_kp, ramp_lg;
};

struct ConeDeviceTables {
    int device = -1;
    float2 *view_cs = nullptr;     // [nViews]
    double2 *view_csd = nullptr;   // [nViews]
    float *preweight = nullptr;    // [nRows][nCols] (flat detector)
    uint16_t *ramp_planes = nullptr;  // [3][kRampSkew][Lg]
    float *parker = nullptr;       // [nViews][nCols] short-scan weights (short_scan plans)
};

}  // namespace fbct

struct fbct_cone_plan {
    fbct_cone_geom g{};
    std::vector<float> angles;
    std::vector<float> view_cs;
    std::vector<double> view_csd;
    std::vector<float> preweight;  // empty for the curved detector
    std::vector<double> taps;      // 2*nCols-1, scaled
    std::vector<uint16_t> ramp_planes;
    std::vector<float> parker;     // [nViews][nCols], short_scan plans only
    fbct::ConeDerived d{};
    std::mutex mu;
    std::vector<fbct::ConeDeviceTables> dev;
};

struct fbct_batch {
    std::vector<fbct_plan *> plans;             // per slice
    std::vector<int64_t> img_off, sino_off;      // [n+1] element offsets
    std::vector<int> fp_prefix[fbct::kFpNumCfg];
    std::vector<int> fpa_prefix[fbct::kAdjNumTiles];
    int fp_cap_max[fbct::kFpNumCfg] = {};
    // FP items (prefix[s] + v * nb + b): first the fp_split items whose largest tile needs
    // more than fp_small_cap rows, then the others (which run with that smaller tile), each
    // class in item order so that a slice's items stay together
    std::vector<int32_t> fp_order[fbct::kFpNumCfg];
    int fp_split[fbct::kFpNumCfg] = {};
    int fp_small_cap[fbct::kFpNumCfg] = {};
    int64_t fp_samples = 0, updates = 0;
    fbct_batch_shape shape{};
    std::mutex mu;
    std::vector<fbct::BatchDevice> dev;
};

namespace fbct {

// Thread-local error reporting.
fbct_status fail(fbct_status s, const char *fmt, ...);
fbct_status ok();

// FBCT_ERR_INVALID_ARGUMENT when the byte ranges [a, a + na) and [b, b + nb) share a byte
// (an execute call whose kernels read one buffer while others write the second would
// return wrong data).
inline fbct_status check_disjoint(const void *a, size_t na, const void *b, size_t nb, const char *what,
                                  const char *aname, const char *bname)
{
    const char *x = static_cast<const char *>(a), *y = static_cast<const char *>(b);
    if (na && nb && x < y + nb && y < x + na) return fail(FBCT_ERR_INVALID_ARGUMENT, "%s: %s and %s must not overlap", what, aname, bname);
    return FBCT_OK;
}

// Host helpers shared by the plan kinds (plan.cpp).
uint16_t bf16_rne(float f);
double bf16_value(uint16_t b);
// atan(z) ~ z + z^3 R(z^2) on |z| <= zmax with nc coefficients of R; false if the fit
// misses 1e-8 relative error.
bool fit_atan(double zmax, int nc, float *coef);
// Ramp taps for the MFMA filter.  The matrix H[m][c] = h[c - m] = h[m - c] (the taps are
// symmetric) is Toeplitz: the 8 consecutive entries h[s .. s + 7] that one lane's B
// fragment holds come from a tap vector.  To read them as one aligned 16-byte load for any
// s, the table holds kRampSkew = 8 copies shifted by one tap each:
//     table[pl][r][u] = plane pl (hi, lo, ll) of h[u - r + dlo],   u < lg,
// zero where |u - r + dlo| > nC - 1; h[s .. s + 7] is row r = (-s) mod 8 from u = s - dlo + r
// (a multiple of 8).  dlo = -(np + 8) and lg = np + kp + 64 cover every offset an np x kp
// tiling reads, with the slack of the 8-aligned windows.
constexpr int kRampSkew = 8;
constexpr int kRampTileN = 256;  // np: a multiple of the widest column tile of the filter
constexpr int ramp_table_dlo(int np) { return -(np + 8); }
constexpr int ramp_table_len(int np, int kp) { return np + kp + 64; }
void ramp_tap_table(const std::vector<double> &taps, int nC, int np, int kp, std::vector<uint16_t> &table, int &lg);

// hipMalloc + copy of a host vector (at least one element is allocated).
template <typename T>
fbct_status upload_vec(T **dst, const std::vector<T> &src)
{
    const size_t bytes = (src.empty() ? 1 : src.size()) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), bytes);
    if (e != hipSuccess) return fail(FBCT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    if (!src.empty()) {
        e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(FBCT_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    }
    return FBCT_OK;
}

// Parker short-scan weights w[v][c] (Parker 1982, smooth sin^2 transitions) times 2, for
// view angles `angles` and column fan angles gamma[c] (the angles the backprojection
// assigns), delta = the largest |fan angle| of the detector edges.  False if the views do
// not cover pi + 2 delta.
bool parker_weights(const std::vector<float> &angles, const std::vector<double> &gamma, double delta,
                    std::vector<float> &w);

// Get (uploading if needed) the tables for the current device.
fbct_status device_tables(fbct_plan *plan, const DeviceTables **out);
// The tables on the current device if they are resident (never uploads or allocates).
fbct_status resident_tables(fbct_plan *plan, const DeviceTables **out);
// The plan's Parker weights on the current device (computed and uploaded on first use).
fbct_status device_parker(fbct_plan *plan, const DeviceTables *t, const float **out);
// cos(gamma_c) per detector column (the backprojection's column map), uploaded on first use.
fbct_status device_cosine(fbct_plan *plan, const DeviceTables *t, const float **out);

// Kernel launchers (fp.hip, bp.hip, filter.hip).  All asynchronous on `stream`.
// target (optional, FP32 sinograms): the output is FP(img) - target.
fbct_status launch_fp(const fbct_plan &p, const DeviceTables &t, const void *img, fbct_dtype ti,
                      void *sino, fbct_dtype to, int n, hipStream_t stream, const float *target = nullptr);
fbct_status launch_fp_adjoint(const fbct_plan &p, const DeviceTables &t, const void *sino,
                              fbct_dtype ti, void *img, fbct_dtype to, int n, hipStream_t stream);
fbct_status launch_bp(const fbct_plan &p, const DeviceTables &t, const float *sino, float *img,
                      int n, bool weighted, hipStream_t stream);
fbct_status launch_bp_adjoint(const fbct_plan &p, const DeviceTables &t, const float *img,
                              float *sino, int n, bool weighted, hipStream_t stream);
fbct_status launch_ramp_filter(const fbct_plan &p, const DeviceTables &t, const float *in,
                               float *out, int64_t rows, hipStream_t stream);
fbct_status launch_fp_batch(const fbct_batch &b, const BatchDevice &bd, const void *img, fbct_dtype ti,
                            void *sino, fbct_dtype to, hipStream_t stream);
fbct_status launch_fp_adjoint_batch(const fbct_batch &b, const BatchDevice &bd, const void *sino, fbct_dtype ti,
                                    void *img, fbct_dtype to, hipStream_t stream);
// Ramp operand of the MFMA filter: the skewed tap table [3][kRampSkew][Lg] of H (nC x nC,
// padded to Np x Kp).
struct RampOperand {
    const uint16_t *planes;
    int nC, Np, Kp, Lg;
};
// rows x nC inputs (FP32 or BF16 storage), optionally times w[r % wrows][c], filtered
// into FP32 outputs.
struct RowWeights {
    const float *w = nullptr;   // row r uses w[r % wrows][c]
    int wrows = 1;
    const float *w2 = nullptr;  // and w2[(r / w2div) % w2rows][c]
    int w2div = 1, w2rows = 1;
};
fbct_status launch_ramp_gemm(const RampOperand &H, const void *in, fbct_dtype ti, float *out, int64_t rows,
                             const RowWeights &rw, hipStream_t stream);
// q: filtered projections, FP32 [n][nViews][nRows][nCols].
fbct_status launch_cone_bp(const fbct_cone_plan &p, const ConeDeviceTables &t, const float *q,
                           void *vol, fbct_dtype to, int n, hipStream_t stream);
fbct_status launch_ramp_filter_mfma(const fbct_plan &p, const DeviceTables &t, const float *in,
                                    float *out, int64_t rows, hipStream_t stream);

}  // namespace fbct

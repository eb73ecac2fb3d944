/*
 * fbct.h -- C ABI of libfbct, the MI355X (gfx950) fan-beam CT operator library.
 *
 * Drop-in replacement for the four TF1 custom ops of xtie97/FanBeamCTRecon_TensorflowOp
 * and the host entry points behind them:
 *
 *   reference op (python name)           reference host entry point              here
 *   FPFanCurve  (fp_fan_curve)           project_fan_curve     FP/transform.h:3    fbct_fp
 *   BPFanCurve  (bp_fan_curve)           backproject_fan_curve BP/transform.h:3    fbct_bp
 *   FBPFanCurve (fbp_fan_curve)          backproject_fan_curve FBP/transform.h:3   fbct_fbp
 *   BackProjection (FDK op)              BP_One_Rot_3D_GPU FDK/GPURecon_GPU.h:59   fbct_fdk
 *   attribute decode (op constructors)   para_setup_fan_curve  FP/transform.h:2    fbct_geom_from_params,
 *                                                                                  fbct_plan_create
 *
 * Conventions (identical to the reference memory layouts, FP/transform.cu:164,250 and
 * BP/transform.cu:87):
 *   image     float[n_slices][nY][nX]           (i fastest)
 *   sinogram  float[n_slices][nViews][nCols]    (column fastest)
 * All data pointers are DEVICE pointers (HIP).  `stream` is a hipStream_t passed as
 * void* (NULL = the default stream).  Every execute call is asynchronous and
 * stream-ordered; none synchronises the device or allocates memory, so a call
 * sequence can be captured into a HIP graph once the plan's tables are resident on
 * the device (fbct_plan_upload, or any earlier execute on that device).
 *
 * The buffers an execute call reads and the buffers it writes (workspaces included)
 * must not overlap: kernels read one while other workgroups write the other, so an
 * overlap is refused with FBCT_ERR_INVALID_ARGUMENT before any device work.
 *
 * Errors are returned as fbct_status; nothing ever exits the process (the reference
 * exits through checkCudaErrors, FP/transform.cu:4).  fbct_last_error() gives a
 * thread-local message for the last failing call on the calling thread.
 */
#ifndef FBCT_H
#define FBCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBCT_VERSION 100 /* 1.0.0 */

typedef enum fbct_status {
    FBCT_OK = 0,
    FBCT_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative count, bad enum       */
    FBCT_ERR_INVALID_GEOMETRY = 2, /* parameter vector describes no valid scanner  */
    FBCT_ERR_SHAPE_MISMATCH = 3,   /* slice count / angle count inconsistent       */
    FBCT_ERR_UNSUPPORTED = 4,      /* valid request this build does not implement  */
    FBCT_ERR_OUT_OF_MEMORY = 5,
    FBCT_ERR_DEVICE = 6,           /* a HIP runtime call failed                    */
    FBCT_ERR_WORKSPACE = 7         /* workspace missing or too small               */
} fbct_status;

/* Storage type of images / sinograms in device memory.  Arithmetic is FP32
 * (accumulation along each ray / over views in FP32; ramp filter in FP64). */
typedef enum fbct_dtype {
    FBCT_F32 = 0,
    FBCT_BF16 = 1
} fbct_dtype;

/* Arithmetic of the ramp filter.  MFMA: Toeplitz GEMM on the BF16 matrix cores with
 * both operands split three ways (six products, FP32 accumulation; FP32-level accuracy).
 * FP64: direct convolution on the vector ALUs with FP64 products and sums. */
typedef enum fbct_filter_mode {
    FBCT_FILTER_MFMA = 0,
    FBCT_FILTER_FP64 = 1
} fbct_filter_mode;

/* Geometry decoded from the reference's 16-float `parameter` attribute exactly as
 * the op constructors decode it (FP/transform.cc:59-76, FDK/FBP.cc:61-86):
 * float arithmetic, PI = 3.14159265359f, angles converted from degrees. */
typedef struct fbct_fan_geom {
    float D;          /* [0]  source-detector distance, mm (FDK row map only) */
    float R;          /* [1]  source-isocentre distance, mm                    */
    int32_t nViews;   /* [2]                                                    */
    float dRange;     /* [3]  angular range, deg                                */
    float dAngle;     /*      dRange/nViews*PI/180, rad (filter scale)          */
    int32_t nCols;    /* [4]                                                    */
    int32_t nX, nY;   /* [5],[6]                                                */
    float rFOV;       /* [7]  mm                                                */
    float dx, dy;     /*      2*rFOV/nX, 2*rFOV/nY, mm                          */
    float xmin, ymin; /*      -nX*dx/2, -nY*dy/2, mm                            */
    float dCol;       /* [8]  rad                                               */
    float dShift;     /* [9]  columns                                           */
    float dLeft;      /* [10] rad                                               */
    int32_t nSamp;    /* [11] slices                                            */
    float startAngle; /* [12] rad (parsed, unused -- as in the reference)       */
    float detBottom;  /* [13] mm (FDK)                                          */
    float dRow;       /* [14] mm (FDK)                                          */
    float dz;         /* [15] mm (FDK)                                          */
} fbct_fan_geom;

typedef struct fbct_plan fbct_plan;

/* ---------------------------------------------------------------- host only --
 * These never touch the GPU. */

int fbct_version(void);
const char *fbct_status_string(fbct_status s);
const char *fbct_last_error(void);

/* Decode the 16-float parameter attribute.  Replaces the attribute decoding in the op
 * constructors FPFanCurveOp/BPFanCurveOp/FBPFanCurveOp (FP/transform.cc:50-91,
 * BP/transform.cc:49-90, FBP/transform.cc:49-90) and BackProjectionOp (FDK/FBP.cc:53-123). */
fbct_status fbct_geom_from_params(const float *parameter, int n_parameter, fbct_fan_geom *out);

/* Create an immutable plan from the op attributes `parameter` (16 floats) and `angle`
 * (nViews floats, rad).  Replaces para_setup_fan_curve (FP/transform.cu:20-62,
 * BP/transform.cu:20-64, FBP/transform.cu:23-58), which kept the geometry in
 * process-global state; a plan owns its geometry, so any number may coexist.
 * Per-ray / per-view tables are computed here in FP64 on the host. */
fbct_status fbct_plan_create(const float *parameter, int n_parameter, const float *angles,
                             int n_angles, fbct_plan **out);
fbct_status fbct_plan_destroy(fbct_plan *plan);
fbct_status fbct_plan_geom(const fbct_plan *plan, fbct_fan_geom *out);

/* FP samples of one slice: sum over rays of N = floor(chord/0.125 px). */
int64_t fbct_plan_fp_samples(const fbct_plan *plan);

/* Host copy of the FP ray table: n_out[nViews*nCols] sample counts and
 * geo_out[nViews*nCols*4] = (u0, w0, du, dw) with u0,w0 the texel-centre
 * coordinates of sample 0 (the FP32 values the kernels use).  Either may be NULL. */
fbct_status fbct_plan_ray_table(const fbct_plan *plan, int32_t *n_out, float *geo_out);

/* Host copy of the FP bundle tables of bundle configuration cfg (0: 64 rays per
 * bundle, 1: 32): info_out[5] = (rays, slab, nslab, nb, cap); major_out[nViews*nb]
 * (1 = bundle marches along x); slabs_out[nViews*nb*2] = first and last slab holding
 * samples; rows_out[nViews*nb*nslab*2] = staged tile rows [lo, hi] per slab (lo > hi:
 * no samples).  Any output may be NULL.  No counterpart in the reference (whose FP
 * samples a texture per ray); it lets tests check the tables on the host. */
fbct_status fbct_plan_fp_bundles(const fbct_plan *plan, int cfg, int32_t *info_out, uint8_t *major_out,
                                 int16_t *slabs_out, int16_t *rows_out);

/* Host copy of the ramp taps h[n], n in [-(nCols-1), nCols-1] (2*nCols-1 doubles). */
fbct_status fbct_plan_ramp_taps(const fbct_plan *plan, double *taps_out);

/* Host copy of the short-scan weights w[nViews][nCols] used by fbct_fbp_short_scan:
 * Parker's weights (Parker 1982) times 2.  FBCT_ERR_INVALID_GEOMETRY if the views do not
 * cover 180 degrees plus the fan angle.  No reference implementation: config.py:23-24
 * declares ParkWeighting and FDK/GPURecon_GPU.h:53 comments out ApplyParkerWeighting. */
fbct_status fbct_plan_parker_weights(const fbct_plan *plan, float *w_out);

/* ------------------------------------------------------------------- device --
 * Upload the plan's tables to the current HIP device (done implicitly by the first
 * execute on a device; call it before capturing a HIP graph). */
fbct_status fbct_plan_upload(fbct_plan *plan);

/* Fan-beam forward projection.  image[n][nY][nX] -> sino[n][nViews][nCols].
 * Replaces project_fan_curve (FP/transform.cu:173-269; kernels :64-146). */
fbct_status fbct_fp(fbct_plan *plan, const float *image, float *sino, int n_slices, void *stream);

/* fbct_fp with explicit storage types (fbct_dtype) for the image and the sinogram;
 * BF16 is widened on load and rounded to nearest even on store, arithmetic is FP32. */
fbct_status fbct_fp_ex(fbct_plan *plan, const void *image, int image_dtype, void *sino, int sino_dtype,
                       int n_slices, void *stream);

/* Residual of the forward projection: residual = FP(image) - target, the FP32 difference
 * of the FP32 projection and target stored in residual_dtype (bitwise what fbct_fp_ex to
 * FP32 followed by the subtraction and a conversion would store), in one pass.  target is
 * an FP32 sinogram batch [n][nViews][nCols].  No reference counterpart: it fuses the
 * data-fidelity residual of an iterative reconstruction (the reference callers subtract
 * in the framework after the FP op). */
fbct_status fbct_fp_residual(fbct_plan *plan, const void *image, int image_dtype, const float *target,
                             void *residual, int residual_dtype, int n_slices, void *stream);

/* Exact matched adjoint of fbct_fp (transpose of the same interpolation weights).
 * sino -> image.  No reference counterpart (the reference has no gradient).  The
 * sinogram must be finite: a NaN or infinity spreads to every pixel whose candidate
 * rays include it, not only to the pixels the ray touches. */
fbct_status fbct_fp_adjoint(fbct_plan *plan, const float *sino, float *image, int n_slices,
                            void *stream);

/* fbct_fp_adjoint with explicit storage types for the sinogram and the image. */
fbct_status fbct_fp_adjoint_ex(fbct_plan *plan, const void *sino, int sino_dtype, void *image, int image_dtype,
                               int n_slices, void *stream);

/* Pixel-driven backprojection (unweighted).  sino -> image.
 * Replaces backproject_fan_curve of BP (BP/transform.cu:123-176; kernel :67-95). */
fbct_status fbct_bp(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream);

/* Transpose of fbct_bp.  image -> sino.  No reference counterpart.  One thread per
 * (view, column) gathers the pixels whose column coordinate lies within one column of
 * the column's node, so every sum has a fixed order (bitwise repeatable, no atomics). */
fbct_status fbct_bp_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices,
                            void *stream);

/* Ramp filter of every row: out[r][c] = sum_m in[r][m] h[c-m] (centre-cropped full
 * linear convolution) on the matrix cores (FBCT_FILTER_MFMA).  Replaces the host IPP
 * filter (FBP/transform.cu:147-176, FDK/GPURecon_GPU.cu:184-217).  in and out must
 * not overlap. */
fbct_status fbct_ramp_filter(fbct_plan *plan, const float *in, float *out, int64_t n_rows,
                             void *stream);

/* fbct_ramp_filter with an explicit fbct_filter_mode. */
fbct_status fbct_ramp_filter_ex(fbct_plan *plan, const float *in, float *out, int64_t n_rows, int mode,
                                void *stream);

/* Workspace in bytes needed by fbct_fbp / fbct_fbp_adjoint / fbct_fdk for n_slices. */
size_t fbct_fbp_workspace_size(const fbct_plan *plan, int n_slices);

/* Filtered backprojection: ramp filter + 1/L^2-weighted pixel-driven BP.
 * sino -> image.  Replaces backproject_fan_curve of FBP (FBP/transform.cu:117-201;
 * kernel :61-89).  workspace: device buffer of fbct_fbp_workspace_size bytes. */
fbct_status fbct_fbp(fbct_plan *plan, const float *sino, float *image, int n_slices,
                     void *workspace, size_t workspace_bytes, void *stream);

/* The backprojection step of FBP alone: 1/L^2-weighted pixel-driven BP of an already
 * filtered sinogram (FBP/transform.cu:61-89,182-188).  sino -> image. */
fbct_status fbct_fbp_backproject(fbct_plan *plan, const float *filtered_sino, float *image, int n_slices,
                                 void *stream);

/* Transpose of fbct_fbp_backproject.  image -> sino.  No reference counterpart. */
fbct_status fbct_fbp_backproject_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices,
                                         void *stream);

/* Exact adjoint of fbct_fbp.  image -> sino.  No reference counterpart. */
fbct_status fbct_fbp_adjoint(fbct_plan *plan, const float *image, float *sino, int n_slices,
                             void *workspace, size_t workspace_bytes, void *stream);

/* FBP with optional pre-weights applied to the sinogram before the ramp filter (they
 * multiply the rows as the MFMA filter stages them; one launch):
 *   FBCT_FBP_COSINE: cos(gamma_c) per column, gamma_c = dLeft + (c + 1/2 - dShift) dCol,
 *     the equiangular FBP's cosine weighting, which the reference omits
 *     (FBP/transform.cu:162); the reference's FBP is flags = 0;
 *   FBCT_FBP_SHORT_SCAN: Parker's weights (as fbct_fbp_short_scan).
 * Flags combine.  No reference counterpart beyond flags = 0 (= fbct_fbp). */
enum { FBCT_FBP_COSINE = 1, FBCT_FBP_SHORT_SCAN = 2 };
fbct_status fbct_fbp_ex(fbct_plan *plan, const float *sino, float *image, int n_slices, void *workspace,
                        size_t workspace_bytes, unsigned flags, void *stream);

/* Short-scan FBP: the sinogram is multiplied by fbct_plan_parker_weights before the
 * ramp filter (one MFMA launch), then backprojected as fbct_fbp.  For a scan over
 * 180 degrees plus the fan angle (parameter[3] gives the angular range).  The first call
 * on a device uploads the weights. */
fbct_status fbct_fbp_short_scan(fbct_plan *plan, const float *sino, float *image, int n_slices,
                                void *workspace, size_t workspace_bytes, void *stream);

/* fbct_fbp without a workspace argument, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The filtered
 * sinograms go to a device array of FBCT_FBP_NOALLOC_SCRATCH_BYTES (64 MiB) that the
 * library holds on each GPU it runs on until the process ends; slices are processed in
 * chunks that fit it (16 clinical 888 x 984 slices per chunk), and a plan whose single
 * slice needs more is refused with FBCT_ERR_WORKSPACE.  The plan's tables must already be
 * resident on the device (fbct_plan_upload or an earlier execute call on it), else
 * FBCT_ERR_INVALID_ARGUMENT.  Results equal fbct_fbp's.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU (different streams or threads) race. */
#define FBCT_FBP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
fbct_status fbct_fbp_noalloc(fbct_plan *plan, const float *sino, float *image, int n_slices, void *stream);

/* The reference FDK op ("BackProjection", FDK/FBP.cc:16-155 -> BP_One_Rot_3D_GPU,
 * FDK/GPURecon_GPU.cu:163-268; kernel :15-50).  The reference pins nZ = nRows = 1
 * (FDK/FBP.cc:85,98), which makes it the FBP of each slice; unlike the reference
 * (slice 0 only, accumulator never zeroed, :178-179,236-238) every slice is
 * processed and the output is fully written. */
fbct_status fbct_fdk(fbct_plan *plan, const float *proj, float *vol, int n_slices,
                     void *workspace, size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------ uneven batches --
 * No reference counterpart: the reference projects one slice geometry per op call
 * (FP/transform.cu:173-269 loops its slices with one process-global geometry).  A
 * batch holds slices of different geometries (image size, views, detector bins), each
 * slice with its own plan, and runs FP or FP^T over all of them in one persistent
 * launch: a fixed grid of workgroups strides over the work items (slice, view, ray
 * bundle) for FP and (slice, pixel tile) for FP^T.  The plans must outlive the batch.
 *
 * Data of slice s sits in one packed device buffer per direction:
 *   images   at element offset image_offsets[s], [nY_s][nX_s]
 *   sinograms at element offset sino_offsets[s],  [nViews_s][nCols_s]
 * (fbct_batch_layout; offsets count elements of the storage type). */
typedef struct fbct_batch fbct_batch;

/* Launch shape of a batch's persistent kernels (a new batch: 64 rays, one FP
 * workgroup per item, 16-pixel gather tiles, one FP^T workgroup per item, no staging). */
typedef struct fbct_batch_shape {
    int32_t fp_rays;        /* rays per FP workgroup: 64 or 32                       */
    int32_t fp_wgs_per_cu;  /* FP workgroups per compute unit in the grid, 1..32;
                               0: one workgroup per work item                       */
    int32_t adj_tile_x;     /* FP^T pixel tile width: 8, 16, 32 or 64 (height 256/w), for
                               slices whose FP samples reach beyond the cells [-1, n-1];
                               other slices take fixed 15 x 15-pixel cell tiles      */
    int32_t adj_wgs_per_cu; /* FP^T workgroups per compute unit in the grid, 1..32;
                               0: one workgroup per work item                       */
    int32_t adj_stage_views; /* FP^T cell tiles: views per LDS staging batch of the
                               sinogram columns and ray records, 4 or 8; 0: no
                               staging (rays read from global memory)               */
} fbct_batch_shape;

fbct_status fbct_batch_create(fbct_plan *const *plans, int n_slices, fbct_batch **out);
fbct_status fbct_batch_destroy(fbct_batch *batch);
int fbct_batch_size(const fbct_batch *batch);

/* image_offsets[n_slices+1], sino_offsets[n_slices+1]: element offsets of each slice in
 * the packed buffers; the last entry is the total element count.  Either may be NULL. */
fbct_status fbct_batch_layout(const fbct_batch *batch, int64_t *image_offsets, int64_t *sino_offsets);

/* Total FP samples and voxel-view updates (nX*nY*nViews summed over slices). */
fbct_status fbct_batch_work(const fbct_batch *batch, int64_t *fp_samples, int64_t *updates);

fbct_status fbct_batch_get_shape(const fbct_batch *batch, fbct_batch_shape *shape);
fbct_status fbct_batch_set_shape(fbct_batch *batch, const fbct_batch_shape *shape);

/* Upload every slice plan's tables and the batch's work tables to the current device
 * (done implicitly by the first execute; call it before capturing a HIP graph). */
fbct_status fbct_batch_upload(fbct_batch *batch);

/* FP of every slice: packed images -> packed sinograms (storage types fbct_dtype). */
fbct_status fbct_batch_fp(fbct_batch *batch, const void *images, int image_dtype, void *sinos, int sino_dtype,
                          void *stream);

/* Matched adjoint of fbct_batch_fp: packed sinograms -> packed images. */
fbct_status fbct_batch_fp_adjoint(fbct_batch *batch, const void *sinos, int sino_dtype, void *images,
                                  int image_dtype, void *stream);

/* ---------------------------------------------------------------- cone beam --
 * Circular-orbit cone-beam FDK with nRows detector rows and an nZ-slice volume, on a
 * flat or a curved (equiangular) detector.
 *
 * The reference's cone kernel (BackProjOneView_CB_EA_kernel, FDK/GPURecon_GPU.cu:15-50,
 * driven by BP_One_Rot_3D_GPU :163-268) is equiangular and already loops over z with
 * the row index ((z - zs) D / U - detBottom) / dRow, but its op pins nZ = nRows = 1
 * (FDK/FBP.cc:85,98).  FBCT_DETECTOR_CURVED is that kernel with any nZ and nRows
 * (reference filter, no pre-weight, weight 1/(U^2+t^2));  FBCT_DETECTOR_FLAT is the
 * textbook Feldkamp-Davis-Kress algorithm for a flat panel (no reference counterpart):
 * cosine pre-weight D / sqrt(D^2+u^2+v^2), Ram-Lak row filter h[0] = 1/(4 dCol^2),
 * h[odd n] = -1/(pi^2 n^2 dCol^2) scaled by dCol dAngle R D / 2, weight 1/U^2.
 *
 * For voxel (x, y, z) and view angle b (source at R (cos b, sin b, 0)):
 *   U = R - x cos b - y sin b,  t = x sin b - y cos b,
 *   column xi = (D t/U - colLeft)/dCol + dShift (flat) or (atan(t/U) - colLeft)/dCol + dShift
 *   (curved), row zeta = (D z/U - rowBottom)/dRow, texel centres at c + 1/2, r + 1/2,
 *   bilinear with clamp-to-edge; a view counts only where 0 < xi < nCols; voxels with
 *   x^2 + y^2 > rFOV^2 are 0.
 * Layouts: projections [n][nViews][nRows][nCols], volumes [n][nZ][nY][nX]. */
typedef enum fbct_detector {
    FBCT_DETECTOR_FLAT = 0,
    FBCT_DETECTOR_CURVED = 1
} fbct_detector;

typedef struct fbct_cone_geom {
    float R, D;           /* source-isocentre, source-detector distance (mm)        */
    int32_t nViews;
    float dAngle;         /* view spacing in the filter scale (rad)                  */
    int32_t detector;     /* fbct_detector                                           */
    int32_t nCols, nRows;
    float dCol;           /* column pitch: mm (flat) or rad (curved)                 */
    float dRow;           /* row pitch at the detector (mm)                          */
    float colLeft;        /* left edge of column 0: mm (flat) or rad (curved)        */
    float rowBottom;      /* lower edge of row 0 at the detector (mm)                */
    float dShift;         /* column shift (columns)                                  */
    int32_t nX, nY, nZ;
    float dx, dy, dz;     /* voxel size (mm)                                         */
    float xmin, ymin, zmin; /* volume corner (mm)                                    */
    float rFOV;           /* reconstructed cylinder radius (mm)                      */
    int32_t short_scan;   /* 1: Parker short-scan weights (per view and column, the    */
                          /*    fan angle atan(u/D) for the flat detector) before the  */
                          /*    filter; the views must cover 180 deg plus the fan      */
} fbct_cone_geom;

typedef struct fbct_cone_plan fbct_cone_plan;

/* The reference FDK op's geometry (FDK/FBP.cc:53-123) with nRows rows and nZ slices:
 * curved detector, colLeft = dLeft, rowBottom = detBottom, dRow, dz from parameter[13..15],
 * zmin = -nZ dz / 2, dAngle = parameter[3]/nViews in rad. */
fbct_status fbct_cone_geom_from_params(const float *parameter, int n_parameter, int n_rows, int n_z,
                                       fbct_cone_geom *out);

fbct_status fbct_cone_plan_create(const fbct_cone_geom *geom, const float *angles, int n_angles,
                                  fbct_cone_plan **out);
fbct_status fbct_cone_plan_destroy(fbct_cone_plan *plan);

/* Workspace bytes for fbct_cone_fdk on n volumes: the filtered projections as FP32
 * rows, 4 bytes per detector sample. */
size_t fbct_cone_workspace_size(const fbct_cone_plan *plan, int n);

/* Row filter alone (pre-weight + ramp): proj (FP32 or BF16 storage) -> FP32. */
fbct_status fbct_cone_filter(fbct_cone_plan *plan, const void *proj, int proj_dtype, float *filtered, int n,
                             void *stream);

/* Backprojection of filtered projections: FP32 -> volume (FP32 or BF16 storage). */
fbct_status fbct_cone_backproject(fbct_cone_plan *plan, const float *filtered, void *vol, int vol_dtype, int n,
                                  void *stream);

/* FDK: filter + backprojection.  proj [n][nViews][nRows][nCols] (fbct_dtype) ->
 * vol [n][nZ][nY][nX] (fbct_dtype); workspace of fbct_cone_workspace_size bytes. */
fbct_status fbct_cone_fdk(fbct_cone_plan *plan, const void *proj, int proj_dtype, void *vol, int vol_dtype, int n,
                          void *workspace, size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FBCT_H */

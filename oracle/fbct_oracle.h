/*
 * fbct_oracle.h -- FP64 CPU oracle for the fan-beam FP / BP / FBP / FDK operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness checker and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library (libfbct) never links, loads or calls it.
 *
 * It restates the discretisations of xtie97/FanBeamCTRecon_TensorflowOp in double
 * precision, with exact (unquantised) bilinear / linear interpolation weights and
 * clamp-to-edge addressing with half-texel centres (the texture semantics the
 * reference configures at FP/transform.cu:224-227 and BP/transform.cu:148-150).
 * Each function cites the reference lines it follows.
 *
 * Memory layouts are the reference's (FP/transform.cu:164,250; BP/transform.cu:87):
 *   image     [slice][j < nY][i < nX]        (i fastest)
 *   sinogram  [slice][view < nViews][c < nCols] (c fastest)
 * Inputs and outputs are double; callers widen float32 data exactly.
 */
#ifndef FBCT_ORACLE_H
#define FBCT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry decoded from the reference's 16-float `parameter` attribute.
 * Scalars are float because the reference decodes them in float
 * (FP/transform.cc:59-76, FDK/FBP.cc:61-86); everything downstream is double. */
typedef struct orc_geom {
    float D;        /* [0]  source-to-detector distance (mm)                 */
    float R;        /* [1]  source-to-isocentre distance (mm)                */
    int   nViews;   /* [2]                                                    */
    float dRange;   /* [3]  angular range (deg)                               */
    float dAngle;   /*      dRange/nViews*PI/180 (rad), FBP/FDK filter scale  */
    int   nCols;    /* [4]                                                    */
    int   nX, nY;   /* [5],[6]                                                */
    float rFOV;     /* [7]  FOV radius (mm)                                   */
    float dx, dy;   /*      2 rFOV / nX, 2 rFOV / nY                          */
    float xmin, ymin;
    float dCol;     /* [8]  column pitch, rad                                 */
    float dShift;   /* [9]  column shift (columns), BP/FBP/FDK only           */
    float dLeft;    /* [10] fan angle of the left detector edge (rad)         */
    int   nSamp;    /* [11] slice count                                       */
    float startAngle; /* [12] rad, parsed but unused by the reference          */
    float detBottom;  /* [13] FDK: bottom edge of the detector rows (mm)       */
    float dRow;       /* [14] FDK: row pitch (mm)                              */
    float dz;         /* [15] FDK: slice thickness (mm)                        */
} orc_geom;

/* Decode p[16] exactly as FP/transform.cc:59-76 (float arithmetic, PI=3.14159265359f).
 * Returns 0 on success, -1 if a count is non-positive. */
int orc_decode_params(const float *p, orc_geom *g);

/* Per-ray setup of the ray-driven FP, FP/transform.cu:64-101, in double.
 * For view angle beta (rad) and column c it returns the sample count N and the
 * texture-space position of sample 0 already shifted by the half-texel
 * (u0 = qx - 0.5, w0 = qy - 0.5) plus the per-sample step. */
void orc_ray_setup(const orc_geom *g, double beta, int c,
                   int *N, double *u0, double *w0, double *du, double *dw);

/* Total number of FP samples of one slice (sum of N over all rays). */
long long orc_fp_sample_count(const orc_geom *g, const float *angles);

/* FP: FP/transform.cu:103-146.  img[nS][nY][nX] -> sino[nS][nViews][nCols]. */
void orc_fp(const orc_geom *g, const float *angles, const double *img, double *sino,
            int nS, int nthreads);

/* Exact transpose of orc_fp (same samples, same clamped bilinear weights).
 * sino[nS][nViews][nCols] -> img[nS][nY][nX]. */
void orc_fp_adjoint(const orc_geom *g, const float *angles, const double *sino,
                    double *img, int nS, int nthreads);

/* Pixel-driven BP: BP/transform.cu:67-95 (weighted=0) and the 1/L^2-weighted
 * variant of FBP/transform.cu:61-89 (weighted=1).  sino -> img. */
void orc_bp(const orc_geom *g, const float *angles, const double *sino, double *img,
            int nS, int weighted, int nthreads);

/* Exact transpose of orc_bp.  img -> sino. */
void orc_bp_adjoint(const orc_geom *g, const float *angles, const double *img,
                    double *sino, int nS, int weighted, int nthreads);

/* Ramp taps h[n], n in [-(nCols-1), nCols-1], stored at h[n + nCols - 1]
 * (2*nCols-1 values), already scaled by R*dAngle*dCol: FBP/transform.cu:149-162. */
void orc_ramp_taps(const orc_geom *g, double *h);

/* Centre-cropped full linear convolution of each row with the ramp taps:
 * q[c] = sum_m p[m] h[c-m], FBP/transform.cu:163-174.  nRowsTotal rows of nCols. */
void orc_ramp_filter(const orc_geom *g, const double *p, double *q, long long nRowsTotal,
                     int nthreads);

/* FBP = weighted BP of the ramp-filtered sinogram: FBP/transform.cu:117-201. */
void orc_fbp(const orc_geom *g, const float *angles, const double *sino, double *img,
             int nS, int nthreads);

/* Exact transpose of orc_fbp (the filter matrix is symmetric). img -> sino. */
void orc_fbp_adjoint(const orc_geom *g, const float *angles, const double *img,
                     double *sino, int nS, int nthreads);

/* Cone-beam backprojection of FDK/GPURecon_GPU.cu:15-50 after the row filter of
 * :184-217, generalised to nZ slices and nRows detector rows.
 *   proj [nS][nViews][nRows][nCols] -> vol [nS][nZ][nY][nX]
 * Column index: (atan(t/U) - dLeft)/dCol + dShift (equiangular, as the reference);
 * row index: ((z - 0) * D / U - detBottom) / dRow; z_k = zmin + dz (k + 0.5).
 * Weight 1/(U^2+t^2).  With nZ = nRows = 1 this is the reference op (FDK/FBP.cc:85,98). */
void orc_fdk(const orc_geom *g, const float *angles, const double *proj, double *vol,
             int nS, int nRows, int nZ, double zmin, int nthreads);

/* Threads the oracle will use when nthreads <= 0. */
int orc_default_threads(void);

#ifdef __cplusplus
}
#endif
#endif

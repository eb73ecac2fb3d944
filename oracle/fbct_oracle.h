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

/* ------------------------------------------------------------ cone beam --
 * Circular-orbit cone-beam FDK on a flat or curved (equiangular) detector with nRows
 * rows and an nZ-slice volume.  No reference counterpart for the flat detector (the
 * reference kernel FDK/GPURecon_GPU.cu:15-50 is equiangular and its op pins
 * nZ = nRows = 1); the curved mode is that kernel generalised, as orc_fdk.
 *
 * Geometry (all double; callers widen the float values the GPU plan receives):
 *   source at R (cos b, sin b, 0); for voxel (x, y, z):
 *     U = R - x cos b - y sin b,  t = x sin b - y cos b;
 *   column coordinate  flat:   xi = (D t / U - colLeft) / dCol + dShift   (dCol, colLeft in mm)
 *                      curved: xi = (atan(t / U) - colLeft) / dCol + dShift (rad)
 *   row coordinate     zeta = (D z / U - rowBottom) / dRow                 (mm)
 *   texel centres at xi = c + 1/2, zeta = r + 1/2; bilinear, clamp-to-edge;
 *   a view contributes only where 0 < xi < nCols (as the reference);
 *   voxels with x^2 + y^2 > rFOV^2 are 0.
 * Flat (Feldkamp-Davis-Kress, textbook): each detector sample is pre-weighted by
 *   D / sqrt(D^2 + u^2 + v^2) at its centre (u, v), rows are convolved with the
 *   discrete Ram-Lak kernel h[0] = 1/(4 dCol^2), h[odd n] = -1/(pi^2 n^2 dCol^2),
 *   scaled by dCol * dAngle * R * D / 2, and backprojected with weight 1/U^2.
 * Curved: the reference filter (orc_ramp_taps, no pre-weight) and weight 1/(U^2 + t^2). */
typedef struct orc_cone {
    double R, D;          /* source-isocentre, source-detector distance (mm)    */
    int nViews;
    double dAngle;        /* view spacing of the filter scale (rad)              */
    int detector;         /* 0 flat, 1 curved (equiangular)                      */
    int nCols, nRows;
    double dCol;          /* flat: mm; curved: rad                               */
    double dRow;          /* mm                                                  */
    double colLeft;       /* left edge of column 0: flat mm, curved rad          */
    double rowBottom;     /* lower edge of row 0 (mm)                            */
    double dShift;        /* columns                                             */
    int nX, nY, nZ;
    double dx, dy, dz;
    double xmin, ymin, zmin;
    double rFOV;
    int short_scan;       /* 1: Parker short-scan weights on the projections          */
} orc_cone;

/* Filter taps h[n + nCols - 1], n in [-(nCols-1), nCols-1], with the constant scale. */
void orc_cone_taps(const orc_cone *g, double *h);

/* Pre-weight w[r][c] (1 for the curved detector). */
void orc_cone_preweight(const orc_cone *g, double *w);

/* proj[nS][nViews][nRows][nCols] -> filtered q (same layout); with short_scan the
 * projections are also multiplied by Parker's weights (orc_parker_weights; the fan angle
 * of a column is atan(u/D) for the flat detector).  Returns -1 if the scan is too short. */
int orc_cone_filter(const orc_cone *g, const float *angles, const double *proj, double *q, int nS, int nthreads);

/* Backprojection of filtered q -> vol[nS][nZ][nY][nX]. */
void orc_cone_bp(const orc_cone *g, const float *angles, const double *q, double *vol, int nS,
                 int nthreads);

/* FDK = orc_cone_filter + orc_cone_bp.  Returns -1 if a short scan is too short. */
int orc_cone_fdk(const orc_cone *g, const float *angles, const double *proj, double *vol, int nS,
                 int nthreads);

/* ------------------------------------------------------------ short scan --
 * Parker's weights (Parker 1982, smooth sin^2 transitions) for a scan over
 * [0, pi + 2 delta] (no reference implementation: config.py:23-24 declares
 * ParkWeighting, FDK/GPURecon_GPU.h:53 comments out ApplyParkerWeighting).  With
 * b = s (angles[v] - angles[0]) and g = s gamma[c] (s = sign of the rotation):
 *   w = sin^2(pi/4 b / (delta - g))               for 0 <= b <= 2 delta - 2 g
 *     = 1                                         up to b = pi - 2 g
 *     = sin^2(pi/4 (pi + 2 delta - b)/(delta + g)) up to b = pi + 2 delta, 0 beyond,
 * times 2: the conjugate ray of (b, g) is (b + pi + 2 g, -g) and their weights sum to
 * one, where a full scan counts every ray twice (the factor 1/2 of the filter taps).
 * gamma[c] is the fan angle of column c's centre as the backprojection maps it;
 * delta = max |fan angle| of the detector edges.  Returns -1 if the views do not
 * cover pi + 2 delta. */
int orc_parker_weights(const float *angles, int nViews, const double *gamma, int nCols, double delta, double *w);

/* Parker weights of a fan geometry, w[nViews][nCols]: gamma = dLeft + (c + 1/2 - dShift) dCol. */
int orc_parker_fan(const orc_geom *g, const float *angles, double *w);

/* Short-scan FBP: filter of the Parker-weighted sinogram + 1/L^2-weighted BP.  Returns
 * -1 (and leaves img untouched) if the scan is too short. */
int orc_fbp_short_scan(const orc_geom *g, const float *angles, const double *sino, double *img, int nS,
                       int nthreads);

/* FBP with the cosine pre-weight cos(gamma_c), gamma_c = dLeft + (c + 1/2 - dShift) dCol
 * (cosine != 0), and/or Parker's weights (short_scan != 0) on the sinogram before the
 * filter.  Returns -1 if short_scan is asked for and the scan is too short. */
int orc_fbp_weighted(const orc_geom *g, const float *angles, const double *sino, double *img, int nS, int cosine,
                     int short_scan, int nthreads);

/* Threads the oracle will use when nthreads <= 0. */
int orc_default_threads(void);

#ifdef __cplusplus
}
#endif
#endif

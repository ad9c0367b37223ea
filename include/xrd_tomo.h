/*
 * xrd_tomo.h -- C ABI of the MI355X (gfx950) XRD-tomography forward projector
 * and MLEM reconstruction.
 *
 * Replaces the MATLAB MEX gateway of airidas-korolkovas/XRD_Tomography:
 *   void mexFunction(int nlhs, mxArray *plhs[], int nrhs, mxArray const *prhs[])
 *   (Forward_XRD.cu:76-302), called from main.m:56-64,
 * and the MLEM loop the reference runs in MATLAB on the CPU (main.m:141-176).
 *
 * Conventions
 *  - Every array argument is a caller-owned DEVICE pointer unless the comment
 *    says otherwise.  No hidden copies of the inputs are made; outputs are
 *    fully overwritten (the reference zero-initialises them,
 *    Forward_XRD.cu:143-166, and then writes every element).
 *  - All layouts are "first index fastest" (MATLAB column-major), exactly as
 *    the reference (Forward_XRD.cu:143-166, integrate_diffraction.cu:365,404).
 *  - Every call returns an int status: XRD_OK (0) or a negative XRD_ERR_*.
 *    xrd_last_error_string() describes the last failure of the calling thread.
 *    Nothing aborts.  The reference performs no argument or CUDA error checks.
 *  - Every call is stream-ordered on `stream` (a hipStream_t passed as void*;
 *    NULL = default stream) and asynchronous with respect to the host unless
 *    documented otherwise.
 *  - Workspace lives in an xrd_context.  A context is not re-entrant: calls
 *    that use the same context must not overlap in time (one host thread at a
 *    time, and calls on different streams must be ordered by the caller).
 *    Different contexts are fully independent.
 */
#ifndef XRD_TOMO_H
#define XRD_TOMO_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define XRD_OK                 0
#define XRD_ERR_INVALID_ARG   -1   /* a scalar argument is out of range        */
#define XRD_ERR_NULL_POINTER  -2   /* a required pointer is NULL               */
#define XRD_ERR_DIMS          -3   /* dimensions <= 0, inconsistent or too big */
#define XRD_ERR_ALIGN         -4   /* a pointer is not suitably aligned        */
#define XRD_ERR_HIP           -5   /* a HIP runtime call failed                */
#define XRD_ERR_ALLOC         -6   /* workspace allocation failed              */
#define XRD_ERR_UNSUPPORTED   -7   /* option not supported by this build       */
#define XRD_ERR_NONFINITE     -8   /* a host-side scalar is NaN or Inf         */

/* ---- reference-quirk switches (xrd_discretization.flags) ------------------ */
/* Compton term uses the (a1,a2) of the previous non-air voxel in the
 * reference's loop order (nx outer, ny inner); the first non-air voxel uses 0.
 * Reproduces the read-before-write at integrate_diffraction.cu:181 vs :192-193. */
#define XRD_QUIRK_STALE_COMPTON_VOXEL  0x1u
/* Source-leg and detector-leg line integrals swapped, as passed at
 * integrate_diffraction.cu:259-260.  Only the Compton term (:347) sees it. */
#define XRD_QUIRK_SWAPPED_LEGS         0x2u
/* Both quirks: the reference kernel's exact semantics. */
#define XRD_QUIRKS_REFERENCE (XRD_QUIRK_STALE_COMPTON_VOXEL | XRD_QUIRK_SWAPPED_LEGS)

/* ---- problem descriptors --------------------------------------------------- */

/* Scanner geometry.  All lengths are in voxel units; the voxel grid covers
 * [0,nx] x [0,ny] and voxel (ix,iy) is centred at (ix+0.5, iy+0.5)
 * (integrate_diffraction.cu:192-193).  These fields are compile-time literals
 * in the reference (integrate_diffraction.cu:81-128, Forward_XRD.cu:17-20). */
typedef struct xrd_geometry {
    int32_t nx, ny;          /* voxel grid (dims of `segmentation`, Forward_XRD.cu:117-119) */
    int32_t n_views;         /* SRC: source positions                                  */
    int32_t n_cols;          /* COL: detector columns                                  */
    int32_t n_rows;          /* ROW: detector rows; pixels DET = COL*ROW               */
    int32_t reserved0;
    double center_x, center_y;   /* rotation centre (Lx/2, Ly/2 in the reference)      */
    double r_src;                /* source trajectory radius (Rsrc)                    */
    double r_det;                /* detector-panel distance from the centre (Rdet)     */
    double view_angle0;          /* phi_v = view_angle0 + v*view_angle_step            */
    double view_angle_step;      /* reference: 2*pi/SRC (integrate_diffraction.cu:93)  */
    double col_pitch, row_pitch; /* detector pixel pitches                             */
    double det_z0;               /* height of detector row 0 above the source plane    */
    double target_angle;         /* anode target angle, rad                            */
    double focal_spot;           /* focal-spot size                                    */
    double wedge_angle;          /* fan-beam wedge opening angle, rad (geomphi)        */
    double voxel_size;           /* in-plane voxel side (geomdvox), 1 in the reference */
} xrd_geometry;

/* Energy / momentum-transfer discretisation and physics constants
 * (Forward_XRD.cu:21-53, integrate_diffraction.cu:138-142, wvt_value.cu:8-23). */
typedef struct xrd_discretization {
    int32_t n_energy;        /* NRG: source and detector energy bins                    */
    int32_t energy_halfwidth;/* DNRG: window is 2*DNRG+1 source bins (needs NRG > 2*DNRG, DNRG <= 7) */
    int32_t n_q;             /* NWT: reconstruction q-bins (parabolic axis)             */
    int32_t n_qf;            /* NQF: fine uniform q-bins used for the counts            */
    int32_t n_mat;           /* MAT: materials (segmentation ids 1..MAT, 0 = air)       */
    int32_t lut_nq, lut_na;  /* NQ, NA: Compton S2 table size                           */
    uint32_t flags;          /* XRD_QUIRK_* bits                                        */
    double e_min;            /* left edge of energy bin 0, keV                          */
    double e_width;          /* dE, keV                                                 */
    double q_min, q_max;     /* q-axis range, 1/Angstrom                                */
    double q_dq0;            /* first coarse bin width (dq0)                            */
    double ray_step;         /* DQ: line-integral step, voxel units                     */
    double hbarc;            /* keV*Angstrom                                            */
    double mc2;              /* electron rest energy, keV                               */
    double kn0, kn1, kn2;    /* 1/Klein-Nishina quadratic fit (Forward_XRD.cu:51-53)    */
} xrd_discretization;

/* Fill `g` with the reference's hard-coded scanner (integrate_diffraction.cu:81-128):
 * mm_over_vox = nx/100, centre (nx,nx), Rsrc 150 mm, Rdet 170 mm, 0.5 mm columns,
 * 1 mm rows, z0 10 mm, 30 deg target, 0.5 mm focal spot, 0.5 deg wedge.  Host-only. */
int xrd_geometry_reference(int32_t nx, int32_t ny, int32_t n_views, int32_t n_cols,
                           int32_t n_rows, xrd_geometry *g);

/* Fill `d` with the reference's discretisation for the given sizes
 * (Forward_XRD.cu:21-53; energy axis 8..80 keV in n_energy bins). Host-only. */
int xrd_discretization_reference(int32_t n_energy, int32_t n_q, int32_t n_qf,
                                 int32_t n_mat, xrd_discretization *d);

/* Number of rows of the model matrix / length of the count vectors:
 * n_views * n_cols * n_rows * n_energy.  Host-only. */
int64_t xrd_num_rows(const xrd_geometry *g, const xrd_discretization *d);
/* Number of columns of the model matrix: n_q * n_mat.  Host-only. */
int64_t xrd_num_cols(const xrd_discretization *d);

/* ---- context ---------------------------------------------------------------- */
typedef struct xrd_context xrd_context;

/* Create a context bound to HIP device `device` (makes it current). */
int xrd_context_create(int device, xrd_context **out);
int xrd_context_destroy(xrd_context *ctx);
/* Bytes of device workspace currently held by the context. Host-only. */
int64_t xrd_context_workspace_bytes(const xrd_context *ctx);

/* Message for the last non-OK status returned on the calling thread. */
const char *xrd_last_error_string(void);
/* ABI version: major*10000 + minor*100 + patch. */
int xrd_version(void);

/* ---- forward projector -------------------------------------------------------
 * Drop-in for Forward_XRD (Forward_XRD.cu:76-302): same inputs in the same order
 * (prhs[0..9]) and the same outputs (plhs[0..2]) with the same layouts.
 *
 *  spectrum     prhs[0]  float [NRG x NRG]   spectrum[e_src + NRG*e_det]
 *                        (integrate_diffraction.cu:61)
 *  ifine        prhs[1]  float [NQF x MAT]   ifine[qf + NQF*mat] (integrate_diffraction.cu:52)
 *  segmentation prhs[2]  int32 [nx x ny]     seg[ix + nx*iy], 1-based material, 0 = air;
 *                        ids outside [0,MAT] are treated as air
 *  a1, a2       prhs[3,4] float [nx x ny]    photoelectric / Compton coefficients in
 *                        1/voxel (main.m:60-61)
 *  s2_table     prhs[5]  float [NQ x NA]     s2[nq + NQ*na], values 0..255 (floored and
 *                        clamped to uint8 on load, integrate_diffraction.cu:66-74)
 *  lut_q_min .. lut_a_max  prhs[6..9]        log2 axes limits of the S2 table
 *  out_counts   plhs[0]  float [SRC x COL x ROW x NRG]  XRD photon counts
 *  out_compton  plhs[1]  float [SRC x COL x ROW x NRG]  Compton counts (x255/S2max units)
 *  out_A        plhs[2]  float [NWT x MAT x SRC x COL x ROW x NRG] model matrix; may be
 *                        NULL to skip it.
 * Row index r = src + SRC*(col + COL*(row + ROW*e_det)); column j = nwt + NWT*mat.
 * Requirements: 16-byte aligned pointers; n_energy > 2*energy_halfwidth.
 * Synchronisation: the call waits once on `stream` for the non-air voxel count,
 * which sizes the view batches of its per-(view, pixel, voxel) workspace (at most
 * 4 GiB); the kernels after that are enqueued asynchronously.                    */
int xrd_forward(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                const float *spectrum, const float *ifine, const int32_t *segmentation,
                const float *a1, const float *a2, const float *s2_table,
                double lut_q_min, double lut_q_max, double lut_a_min, double lut_a_max,
                float *out_counts, float *out_compton, float *out_A, void *stream);

/* Slice `slice` of a volume: segmentation, a1 and a2 are [nx x ny x n_slices]
 * (x fastest, then y, then the slice).  The slice's voxels scatter exactly as in
 * xrd_forward.  The attenuation legs are 3-D segments through the whole volume: the
 * source leg from the focal spot at the slice centre (height slice + 0.5 voxels) to
 * the voxel at its wedge height (integrate_diffraction.cu:199-205), the detector leg
 * from the pixel (slice centre + det_z0 + row*row_pitch) to the voxel.  Each is
 * sampled at the attenuator's 2-D sample positions (attenuator.cu:44-69), at the
 * segment's height there, trilinear with zero outside the volume.  No reference
 * counterpart: the reference scans one slice with attenuation constant in z, which
 * is n_slices = 1 here (identical to xrd_forward).  Outputs as xrd_forward, for the
 * one slice.  Limit: (nx+2)(ny+2)(n_slices+2) <= 2^31. */
int xrd_forward_slice(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                      const float *spectrum, const float *ifine, const int32_t *segmentation,
                      const float *a1, const float *a2, int32_t n_slices, int32_t slice,
                      const float *s2_table, double lut_q_min, double lut_q_max,
                      double lut_a_min, double lut_a_max, float *out_counts, float *out_compton,
                      float *out_A, void *stream);

/* ---- projector launch shape ----------------------------------------------------
 * No reference counterpart: the reference's launch shape is fixed
 * (Forward_XRD.cu:244-246, one block per (view, pixel, energy), 256 threads).
 * Later forward calls on ctx (xrd_forward, xrd_forward_slice, xrd_forward_bf16) use
 * voxels_per_chunk voxels per chunk of the accumulation kernel's per-(voxel, energy)
 * table (1..32; default 16; the chunk's FP32 partials add up at most that many
 * terms, so the last bits of the results depend on it) and at most views_per_batch
 * views per batch of (view, pixel, voxel) records (default: as many as fit 4 GiB).
 * 0 restores a default.  Host-only. */
int xrd_forward_set_launch(xrd_context *ctx, int32_t voxels_per_chunk, int32_t views_per_batch);
/* Autotuner: runs xrd_forward (with A) on the first sample_views views of the scan
 * for voxels_per_chunk in {8, 12, 16} and views_per_batch in {2, 8, 32}, timing each
 * with HIP events on `stream`, keeps the fastest on ctx and reports it with its time
 * in ms.  Allocates and frees scratch outputs; synchronises `stream`. */
int xrd_forward_autotune(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                         const float *spectrum, const float *ifine, const int32_t *segmentation,
                         const float *a1, const float *a2, const float *s2_table,
                         double lut_q_min, double lut_q_max, double lut_a_min, double lut_a_max,
                         int32_t sample_views, void *stream, int32_t *out_voxels_per_chunk,
                         int32_t *out_views_per_batch, float *out_ms);

/* ---- BF16 tables --------------------------------------------------------------
 * The reference reads its form factors and its (a1,a2) attenuation texture in FP32
 * (integrate_diffraction.cu:50-53, Forward_XRD.cu:216-237).  xrd_forward_bf16 takes
 * them as bfloat16 (the upper 16 bits of an IEEE binary32, rounded to nearest even)
 * with one FP32 scale per material, decoded on load (value = bf16 * scale, an FP32
 * product) and then used exactly as the FP32 tables of xrd_forward_slice.  A map scale
 * is picked by the voxel's own segmentation id, so a material of constant (a1,a2)
 * decodes exactly when its scale is that value.  All pointers are device pointers,
 * 16-byte aligned; the struct itself is read on the host during the call. */
typedef struct xrd_bf16_tables {
    const uint16_t *ifine;      /* [NQF x MAT]: Ifine(qf,mat) = bf16(ifine[qf + NQF*mat]) * ifine_scale[mat] */
    const float *ifine_scale;   /* [MAT]                                                     */
    const uint16_t *a1, *a2;    /* [nx x ny x n_slices] (x fastest), 1/voxel before the scale */
    const float *mu_scale;      /* [2 x (MAT+1)]: a1(i) = bf16(a1[i]) * mu_scale[id(i)],
                                   a2(i) = bf16(a2[i]) * mu_scale[MAT+1 + id(i)], id(i) the
                                   segmentation id of voxel i (ids outside [0,MAT]: 0, air) */
} xrd_bf16_tables;

/* xrd_forward_slice (n_slices = 1, slice = 0: xrd_forward) with BF16 tables; the
 * segmentation covers the same voxels as the maps.  Outputs and limits as there. */
int xrd_forward_bf16(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                     const float *spectrum, const xrd_bf16_tables *tables,
                     const int32_t *segmentation, int32_t n_slices, int32_t slice,
                     const float *s2_table, double lut_q_min, double lut_q_max,
                     double lut_a_min, double lut_a_max, float *out_counts, float *out_compton,
                     float *out_A, void *stream);

/* ---- ray traversal (the GetIntersection + attenuator sampling, on its own) ----
 * For `n` segments (x1,y1)->(x2,y2) given in `seg` as 4 doubles each, computes
 * in FP64 the entry point into [0,nx]x[0,ny] with GetIntersection's rules
 * (GetIntersection.cu:9-57), and the number of line-integral samples
 * q = 0,1,.. with q*ray_step <= |P2 - entry| (attenuator.cu:44-69).
 *  out_side  int32 [n]      side hit first (0: y=0, 1: x=nx, 2: y=ny, 3: x=0), -1 if none
 *  out_entry double [2n]    entry point (NaN,NaN if none)
 *  out_count int32 [n]      number of samples (0 if none)
 *  out_texels (optional) int32 [2*n*max_samples]: (floor(x-0.5), floor(y-0.5)) of each
 *            sample, -2^31 padding after the last one.  NULL to skip.
 *  out_integrals (optional) float [n_basis*n]: Kahan-compensated FP32 line integrals
 *            of the n_basis maps `maps` [n_basis x nx x ny] with exact bilinear weights
 *            and zero border, times ray_step.  NULL to skip (then maps may be NULL). */
int xrd_trace_rays(xrd_context *ctx, int32_t nx, int32_t ny, double ray_step, int64_t n,
                   const double *seg, int32_t *out_side, double *out_entry,
                   int32_t *out_count, int32_t max_samples, int32_t *out_texels,
                   int32_t n_basis, const float *maps, float *out_integrals,
                   void *stream);

/* ---- differentiable coefficient model and its VJP -----------------------------
 * No reference counterpart as an entry point: the differentiable form of the
 * Rayleigh/XRD accumulation (integrate_diffraction.cu:266-315) with
 *   - per-voxel form factors F_vox(q_n) = sum_k coef[vox*K + k] basis[k*NWT + n]
 *     on the coarse q axis (every voxel with segmentation > 0 scatters), and
 *   - attenuation mu_e(pix) = sum_b beta[b*NRG + e] maps[b*nx*ny + pix]
 *     (the reference is B = 2, maps = (a1,a2), beta = (1/e^3, 1/f2(e)) of
 *     integrate_diffraction.cu:291-296; a per-energy map is B = NRG, beta = I).
 * out_counts has the Nphotons layout (Compton excluded).
 * Limits: K <= 32, B <= 64, NRG <= 64, NWT <= 128, (nx+2)(ny+2) < 2^24. */
int xrd_forward_coeffs(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                       const float *spectrum, const int32_t *segmentation, const float *coef,
                       int32_t n_coef, const float *basis, int32_t n_basis, const float *maps,
                       const float *beta, float *out_counts, void *stream);

/* Vector-Jacobian product of xrd_forward_coeffs for the upstream gradient
 * grad_counts [rows] (e.g. 1 - y/(counts + background) for the Poisson NLL).
 * Line integrals are recomputed, not stored.  grad_coef [nx*ny x K] (FP64,
 * deterministic) and grad_maps [B x nx*ny] (FP64, accumulated with atomics) are
 * overwritten; either may be NULL.  With B > 4 the map gradient's detector legs are
 * scattered in bundles of adjacent pixels from a per-view gamma workspace (views
 * processed in slabs) of at most 16 GiB and at most half the free device memory; it
 * stays in the context (xrd_context_workspace_bytes counts it).  When one view's
 * gamma exceeds that, the allocation fails, or the environment sets
 * XRD_VJP_SCATTER=inplace, they are scattered inside the VJP launch instead (same
 * results up to atomic order). */
int xrd_vjp_coeffs(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                   const float *spectrum, const int32_t *segmentation, const float *coef,
                   int32_t n_coef, const float *basis, int32_t n_basis, const float *maps,
                   const float *beta, const float *grad_counts, double *grad_coef,
                   double *grad_maps, void *stream);

/* CT transmission channel of the simultaneous CT + XRD scanner (README.md:14).
 * No reference counterpart to bind: the reference names it only in the
 * commented-out launches of integrate_src2det / polychromatic_src2det
 * (Forward_XRD.cu:271-276) and does not ship their kernels.  For view v and
 * detector column c the primary ray runs from the source to the column's point on
 * the detector (the XRD pixel positions of row 0); its part inside the grid
 * (Liang-Barsky clipping, FP64) is sampled like attenuator.cu:44-69 from the
 * source-side entry, and
 *   out_t[v + V*(c + COL*ed)] = sum_{e in win(ed)} spectrum[e + NRG*ed]
 *                               * exp(-sum_b beta[b,e] L_b),
 * L_b the line integral of maps[b] [nx*ny].  Limits: B <= 64, NRG <= 64,
 * (nx+2)(ny+2) < 2^24. */
int xrd_forward_ct(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                   const float *spectrum, int32_t n_basis, const float *maps, const float *beta,
                   float *out_t, void *stream);

/* VJP of xrd_forward_ct with respect to the maps for the upstream gradient
 * grad_t [V*COL*NRG]: grad_maps [B x nx*ny] (FP64, overwritten). */
int xrd_vjp_ct(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
               const float *spectrum, int32_t n_basis, const float *maps, const float *beta,
               const float *grad_t, double *grad_maps, void *stream);

/* Poisson negative log-likelihood sum(f - y log f), f = counts + background
 * (background may be NULL), written to *out_loss_host (synchronises the stream);
 * grad_counts (optional) receives 1 - y/f. */
int xrd_poisson_nll(xrd_context *ctx, int64_t n, const float *counts, const float *background,
                    const float *y, float *grad_counts, double *out_loss_host, void *stream);

/* ---- model-matrix operators (explicit A, row-major [rows x cols]) ------------
 * A is the out_A of xrd_forward viewed as rows (src,col,row,e_det) x cols (nwt,mat),
 * as the reference's MLEM views it (main.m:140). a_dtype selects the storage type. */
#define XRD_DTYPE_F32  0
#define XRD_DTYPE_BF16 1
#define XRD_DTYPE_F16  2
/* BF24: FP32 rounded to nearest even at 24 bits (sign, 8-bit exponent, 15-bit
 * mantissa; relative error <= 2^-16), stored as two planes in one buffer of
 * 3 n_rows n_cols bytes: n_rows*n_cols uint16 holding bits 31..16, then
 * n_rows*n_cols uint8 holding bits 15..8.  Needs n_cols % 8 == 0 and a 16-byte
 * aligned buffer.  No reference counterpart (the reference keeps A in FP32). */
#define XRD_DTYPE_BF24 3

/* out[r] = sum_j A[r,j] x[j]            (the A*Irecon of main.m:162).  */
int xrd_project(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows, int32_t n_cols,
                const float *x, float *out, void *stream);
/* out[j] = sum_r A[r,j] w[r]            (the AmatrixCPU'*(...) of main.m:162). FP64 sums. */
int xrd_backproject(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                    int32_t n_cols, const float *w, double *out, void *stream);
/* out[j] = sum_r A[r,j]                 (the sum(AmatrixCPU,1)' of main.m:162). FP64 sums. */
int xrd_sensitivity(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                    int32_t n_cols, double *out, void *stream);

/* Convert n FP32 values to BF16/F16/BF24 storage (round to nearest even). */
int xrd_convert_matrix(xrd_context *ctx, const float *A, int64_t n, int out_dtype, void *out,
                       void *stream);

/* Convert with a per-column power-of-two storage scale (the per-material scale of
 * the reduced-precision path): out[r,j] = round(A[r,j] * col_scale[j]), where
 * col_scale (device, n_cols floats, written here) maps each column's maximum to
 * about half the dtype's largest finite value.  Pass col_scale to
 * xrd_mlem_plan_create to iterate on the scaled matrix. */
int xrd_convert_matrix_scaled(xrd_context *ctx, const float *A, int64_t n_rows, int32_t n_cols,
                              int out_dtype, void *out, float *col_scale, void *stream);

/* ---- MLEM (main.m:141-176) ----------------------------------------------------
 * x <- x .* A'( y ./ (A x + c + 1e-20) ) ./ (sum(A,1)' + 1e-7); non-finite -> 0; x<0 -> 0.
 * The plan keeps pointers to A, y, c and x (caller-owned, must outlive the plan),
 * computes the sensitivity sum(A,1) once (FP64) at creation, and owns the
 * per-workgroup partial buffers of the fused single-pass forward+adjoint kernel. */
typedef struct xrd_mlem_plan xrd_mlem_plan;

#define XRD_MLEM_FUSED    0   /* one pass over A per iteration (row dot, ratio, row axpy) */
#define XRD_MLEM_UNFUSED  1   /* plain: A x, then A' r as two passes over A               */

/* col_scale: NULL, or the per-column storage scale of A (xrd_convert_matrix_scaled). */
int xrd_mlem_plan_create(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                         int32_t n_cols, const float *col_scale, const float *y,
                         const float *c, float *x, int mode, void *stream,
                         xrd_mlem_plan **out);
/* Run n_iter updates of x in place.  use_graph != 0 replays captured hipGraphs of one
 * and of eight iterations (both captured on first use; n_iter / 8 launches of the
 * second, then the rest of the first), bit for bit equal to n_iter eager updates. */
int xrd_mlem_plan_iterate(xrd_mlem_plan *plan, int32_t n_iter, int use_graph, void *stream);
/* Poisson log-likelihood sum_r (y log f - f), f = A x + c + 1e-20, FP64, written to
 * *out_host after synchronising `stream`. */
int xrd_mlem_plan_loglik(xrd_mlem_plan *plan, double *out_host, void *stream);
/* Autotuner: times the fused pass for each launch shape (workgroups per CU in
 * {1,2,3,4,6} x rows per wave in {1,2,4}) on the plan's own A, y, c and x (x is
 * not modified), keeps the fastest and reports it.  Synchronises `stream`. */
int xrd_mlem_plan_autotune(xrd_mlem_plan *plan, int32_t reps, void *stream,
                           int32_t *out_blocks_per_cu, int32_t *out_rows_per_wave,
                           float *out_ms);
/* Average time in ms of the streaming pass over A alone (the fused pass, or for
 * XRD_MLEM_UNFUSED the A x, ratio and A'r kernels), over reps back-to-back passes
 * timed with HIP events on `stream`; x is not modified.  Synchronises `stream`. */
int xrd_mlem_plan_time_pass(xrd_mlem_plan *plan, int32_t reps, void *stream, float *out_ms);
/* Sets the launch shape explicitly (0 = default for either field), e.g. from a cache
 * of earlier autotuning results. */
int xrd_mlem_plan_set_launch(xrd_mlem_plan *plan, int32_t blocks_per_cu,
                             int32_t rows_per_wave);
/* Device pointer to the FP64 sensitivity vector (n_cols) held by the plan. */
const double *xrd_mlem_plan_sensitivity(const xrd_mlem_plan *plan);
int xrd_mlem_plan_destroy(xrd_mlem_plan *plan);

/* ---- MLEM without allocation ----------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * n_iter MLEM updates of x in place with the kernels, arithmetic and default launch
 * shape of xrd_mlem_plan_create + xrd_mlem_plan_iterate (XRD_MLEM_FUSED, no graph):
 * x bit-identical to that plan's.  The sensitivity sum(A,1)' is recomputed on every
 * call (one extra pass over A) and kept, with the per-workgroup FP64 partials, in a
 * fixed device array of XRD_MLEM_STATIC_SCRATCH_BYTES that the library holds on each
 * GPU until the process ends; the call makes no hipMalloc/hipFree.  A problem needs
 * (workgroups + 1) * n_cols * 8 bytes of it, where workgroups = min(2 * CUs,
 * ceil(n_rows / 16)): up to 1,024 columns at any row count on a 256-CU MI355X.
 * Larger problems get XRD_ERR_DIMS.  Every call on one GPU shares that GPU's array,
 * and it keeps the last call's contents: calls that overlap in time on one GPU (other
 * streams or host threads) race.  ctx selects the device; its workspace is not used. */
#define XRD_MLEM_STATIC_SCRATCH_BYTES 4259840
int xrd_mlem_iterate_static(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                            int32_t n_cols, const float *col_scale, const float *y,
                            const float *c, float *x, int32_t n_iter, void *stream);

/* ---- batched MLEM ------------------------------------------------------------
 * MLEM for n_problems independent explicit matrices (phantoms or slices of uneven
 * size and view count) with one persistent launch per pass over all of them.
 * Host arrays of n_problems device pointers/sizes are read at creation; col_scale
 * may be NULL or hold NULL entries; reserved must be 0.  The problems' rows are
 * cut into one contiguous chunk of about equal cost per persistent workgroup, and
 * each problem is streamed with a row shape picked from its column count.
 * Results are deterministic for a given launch shape.  The sensitivities are
 * computed (FP64) at creation.  No reference counterpart: the reference runs one
 * phantom per MATLAB session (main.m). */
typedef struct xrd_mlem_batch xrd_mlem_batch;

int xrd_mlem_batch_create(xrd_context *ctx, int32_t n_problems, int a_dtype,
                          const void *const *A, const int64_t *n_rows, const int32_t *n_cols,
                          const float *const *col_scale, const float *const *y,
                          const float *const *c, float *const *x, int32_t reserved,
                          void *stream, xrd_mlem_batch **out);
/* n_iter updates of every x[k]; use_graph replays a captured iteration. */
int xrd_mlem_batch_iterate(xrd_mlem_batch *batch, int32_t n_iter, int use_graph, void *stream);
/* Launch shape of the persistent pass (0 = default for either field): workgroups
 * per CU (clamped to what fits on a CU at once) and rows per row slot and step.
 * Waits for the device, then re-partitions the rows. */
int xrd_mlem_batch_set_launch(xrd_mlem_batch *batch, int32_t blocks_per_cu,
                              int32_t rows_per_wave);
/* For every row-span group (problems are launched per column span 256, 384, 512,
 * 768, 1024, 1536 or 2048), times that group's streaming launch (reps launches, x untouched) over
 * blocks_per_cu in {1,2,3,4} (clamped to what fits on a CU), rows_per_wave in
 * {1,2,4} and the chunking of its rows (one chunk per workgroup, or guided chunks of
 * 2 or 6 halving rounds taken from a counter; fixed when the environment variable
 * XRD_BATCH_CHUNK_LEVELS was set at creation) and keeps the fastest.  Returns the shape
 * of the group with the largest matrices and the mean time of a whole pass in ms
 * (outputs may be NULL). */
int xrd_mlem_batch_autotune(xrd_mlem_batch *batch, int32_t reps, void *stream,
                            int32_t *out_blocks_per_cu, int32_t *out_rows_per_wave,
                            float *out_ms);
/* The row-span groups of the batch: writes up to max_groups entries of span,
 * persistent grid (workgroups), rows_per_wave and chunks (equal to the grid when each
 * workgroup streams one fixed chunk, else the number of guided chunks); returns the
 * number of groups (negative on error).  Output arrays may be NULL. */
int xrd_mlem_batch_groups(const xrd_mlem_batch *batch, int32_t max_groups, int32_t *out_span,
                          int32_t *out_grid, int32_t *out_rows_per_wave, int32_t *out_chunks);
/* Mean duration in ms of reps streaming passes at the current launch shape,
 * measured with HIP events on stream (x untouched). */
int xrd_mlem_batch_time_pass(xrd_mlem_batch *batch, int32_t reps, void *stream, float *out_ms);
int xrd_mlem_batch_destroy(xrd_mlem_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* XRD_TOMO_H */

/*
 * xrd_oracle.h -- FP64 host restatement of the reference's XRD forward model
 * and MLEM update.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py may load this library, and only as the checker or
 * the timed CPU baseline.  The product (libxrd_tomo.so) never links it.
 *
 * Every function cites the reference file:line it restates.  All arithmetic is
 * IEEE double with FP contraction disabled (-ffp-contract=off), so the ray
 * traversal (entry point, sample count, texel indices) is reproducible bit for
 * bit by any other IEEE double implementation that performs the same
 * operations in the same order.
 */
#ifndef XRD_ORACLE_H
#define XRD_ORACLE_H

#include <stdint.h>
#include "../include/xrd_tomo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GetIntersection (GetIntersection.cu:9-57): first side of [0,nx]x[0,ny], in the
 * order y=0, x=nx, y=ny, x=0, that the segment (x1,y1)->(x2,y2) crosses with
 * |den| > 1e-5 and both parameters in (-1e-5, 1+1e-5).  Returns the side id
 * (0..3) and the point, or -1 and NaN. */
int xo_get_intersection(double x1, double y1, double x2, double y2, int nx, int ny,
                        double *rx, double *ry);

/* Sampling of one attenuator ray (attenuator.cu:19-69): entry point, sample
 * count n (samples q = 0..n-1 with q*step <= |P2-entry|), and, when texels is
 * non-NULL, the texel base indices (floor(x-0.5), floor(y-0.5)) of the first
 * min(n, max_samples) samples.  Returns the side id. */
int xo_trace_ray(double x1, double y1, double x2, double y2, int nx, int ny, double step,
                 double *entry_x, double *entry_y, int32_t *count, int32_t max_samples,
                 int32_t *texels);

/* Line integrals (attenuator.cu:9-109) of nb maps [nb x nx x ny] from the ROI
 * entry of (x1,y1)->(x2,y2) to (x2,y2): step * sum of exact bilinear samples,
 * zero border (texture filter of Forward_XRD.cu:229-233). */
void xo_line_integrals(const double *maps, int nb, int nx, int ny, double step,
                       double x1, double y1, double x2, double y2, double *out);

/* Line integrals of nb volumes [nb x nx x ny x nz] along the 3-D segment
 * (x1,y1,z1)->(x2,y2,z2): the attenuator's samples of the 2-D projection
 * (x1,y1)->(x2,y2) (ROI entry, then q*step while q*step <= |P2 - entry|), each at
 * the height of the segment there (z linear in the 2-D distance to P2), trilinear
 * with zero outside the volume. */
void xo_line_integrals_3d(const double *vols, int nb, int nx, int ny, int nz, double step,
                          double x1, double y1, double z1, double x2, double y2, double z2,
                          double *out);

/* Bilinear value of map [nx x ny] at texel coordinates (x,y), zero border,
 * texel (i,j) centred at (i+0.5, j+0.5) (tex2D semantics, attenuator.cu:58). */
double xo_bilinear(const double *map, int nx, int ny, double x, double y);

/* find_src_nrg_range (find_src_nrg_range.cu:8-25). */
void xo_find_src_nrg_range(int nrg, int n_energy, int halfwidth, int *lo, int *hi);

/* wvt_value (wvt_value.cu:8-23; XRD_builder.m:16-21): centre and width of coarse
 * bin n on the parabolic q axis. */
void xo_wvt_value(int n, int n_q, double q_min, double q_max, double dq0,
                  double *q, double *dq);

/* Compton channel match (integrate_diffraction.cu:323-328). */
double xo_channel_match(double es1, double es2, double eratio, double ed1, double ed2);

/* Per-(view, pixel, voxel) scattering geometry (integrate_diffraction.cu:81-255). */
typedef struct xo_scatter_geom {
    double src_x, src_y, det_x, det_y;  /* 2-D endpoints of the two legs          */
    double cos_theta, sin_half;         /* cos(theta), sin(theta/2)               */
    double anorm2, bnorm2;              /* |a|^2, |b|^2                           */
    double res2;                        /* sS2 + vV2 + dD2                        */
    double geom;                        /* geom_src2vox * geom_vox2det            */
    double rayleigh;                    /* (1 + cos^2 theta)/2                    */
    double vox_z, det_z;                /* heights of the voxel (wedge) and pixel */
} xo_scatter_geom;

void xo_scatter_geometry(const xrd_geometry *g, int view, int col, int row, int ix, int iy,
                         xo_scatter_geom *out);

/* The forward model of slice `slice` of an n_slices-slice volume (seg_vol, a1_vol,
 * a2_vol [nx x ny x n_slices]): the slice's voxels scatter with the arithmetic of
 * xo_forward; the source leg runs from the focal spot at the slice centre
 * (slice + 0.5) to the voxel at its wedge height, the detector leg from the pixel
 * (slice centre + det_z0 + row*row_pitch) to the voxel, both through the volume
 * (xo_line_integrals_3d).  n_slices = 1 is xo_forward (attenuation constant in z,
 * as in the reference). */
int xo_forward_slice(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                     const double *ifine, const int32_t *seg_vol, const double *a1_vol,
                     const double *a2_vol, int n_slices, int slice, const double *s2_table,
                     double lut_q_min, double lut_q_max, double lut_a_min, double lut_a_max,
                     const int32_t *views, int nvl, const int32_t *cols, int ncl,
                     double *out_N, double *out_C, double *out_A, int nthreads);

/* Full forward model (integrate_diffraction.cu:9-409 with the launch of
 * Forward_XRD.cu:244-259) for the views listed in `views` (n_views_list of them;
 * NULL = all views) and the detector columns listed in `cols` (n_cols_list of them;
 * NULL = all columns).  Outputs use the reference layouts with SRC replaced by
 * n_views_list and COL by n_cols_list: row r = vi + nvl*(ci + ncl*(row + ROW*e_det)).
 *   out_N, out_C: [rows]; out_A: [NWT*MAT*rows] or NULL.
 * Inputs are the xrd_forward inputs converted to double.  nthreads <= 0 uses
 * OpenMP's default.  Returns 0 or a negative XRD_ERR_* code. */
int xo_forward(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
               const double *ifine, const int32_t *segmentation, const double *a1,
               const double *a2, const double *s2_table, double lut_q_min, double lut_q_max,
               double lut_a_min, double lut_a_max, const int32_t *views, int n_views_list,
               const int32_t *cols, int n_cols_list, double *out_N, double *out_C,
               double *out_A, int nthreads);

/* Explicit-matrix operators (main.m:162), A row-major [n_rows x n_cols]. */
void xo_project(const double *A, int64_t n_rows, int n_cols, const double *x, double *out,
                int nthreads);
void xo_backproject(const double *A, int64_t n_rows, int n_cols, const double *w,
                    double *out, int nthreads);
void xo_sensitivity(const double *A, int64_t n_rows, int n_cols, double *out, int nthreads);

/* One MLEM update of x in place given the sensitivity s = sum(A,1)' (main.m:162-163,176). */
void xo_mlem_update(const double *A, int64_t n_rows, int n_cols, const double *y,
                    const double *c, const double *s, double *x, int nthreads);
/* n_iter MLEM updates of x in place (main.m:159-176). */
void xo_mlem(const double *A, int64_t n_rows, int n_cols, const double *y, const double *c,
             double *x, int n_iter, int nthreads);
/* Poisson log-likelihood sum_r (y log f - f), f = A x + c + 1e-20. */
double xo_loglik(const double *A, int64_t n_rows, int n_cols, const double *y,
                 const double *c, const double *x, int nthreads);

/* Coefficient model (differentiable form of integrate_diffraction.cu:266-315):
 * per-voxel form factors coef [nx*ny x K] (k fastest) on a basis [K x NWT] of the
 * coarse q axis, attenuation mu_e = sum_b beta[b*NRG + e] maps[b*nx*ny + pix].
 * Non-air voxels (seg > 0) scatter.  out_yhat: counts in the N layout (no Compton). */
int xo_forward_coeffs(const xrd_geometry *g, const xrd_discretization *d,
                      const double *spectrum, const int32_t *seg, const double *coef, int K,
                      const double *basis, int B, const double *maps, const double *beta,
                      double *out_yhat, int nthreads);
/* Vector-Jacobian product of the coefficient model with upstream gradient w [rows]:
 * grad_coef [nx*ny x K] and grad_maps [B x nx*ny] (either may be NULL). */
int xo_vjp_coeffs(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                  const int32_t *seg, const double *coef, int K, const double *basis, int B,
                  const double *maps, const double *beta, const double *w, double *grad_coef,
                  double *grad_maps, int nthreads);

/* Number of OpenMP threads the oracle would use for nthreads <= 0. */
/* CT transmission channel (new; named by the commented-out integrate_src2det and
 * polychromatic_src2det launches, Forward_XRD.cu:271-276).  xo_ct_chord clips the
 * segment (sx,sy)->(dx,dy) to [0,nx]x[0,ny] (Liang-Barsky, FP64): returns 1 and the
 * entry (ex,ey) and exit (xx,xy), or 0 if it misses.  xo_forward_ct: for view v and
 * detector column c, T[v + V (c + COL ed)] = sum_{e in win(ed)} RF[e,ed]
 * exp(-sum_b beta[b,e] L_b) with L_b the line integral of map b over the clipped
 * source-to-column ray (attenuator.cu:44-69 sampling from the entry).  xo_vjp_ct:
 * grad_maps [B x nx*ny] = dT/dmaps transposed applied to w [V*COL*NRG]. */
int xo_ct_chord(double sx, double sy, double dx, double dy, int nx, int ny, double *ex,
                double *ey, double *xx, double *xy);
int xo_forward_ct(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                  int B, const double *maps, const double *beta, double *out_t, int nthreads);
int xo_vjp_ct(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum, int B,
              const double *maps, const double *beta, const double *w, double *grad_maps,
              int nthreads);

int xo_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif

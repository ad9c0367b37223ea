/*
 * xrd_oracle.c -- FP64 host restatement of the reference XRD forward model and
 * MLEM update.  TEST INFRASTRUCTURE ONLY (see xrd_oracle.h): the checker the
 * HIP path is compared against, and the timed CPU baseline.
 *
 * Build: gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off (oracle/Makefile).
 *
 * Restated from airidas-korolkovas/XRD_Tomography:
 *   GetIntersection.cu:9-57, attenuator.cu:9-109, find_src_nrg_range.cu:8-25,
 *   wvt_value.cu:8-23, integrate_diffraction.cu:33-409, Forward_XRD.cu:143-166
 *   and main.m:141-176.
 * Differences from the reference, by design:
 *   - FP64 everywhere, correctly rounded libm instead of __expf/__log2f/rsqrt
 *     intrinsics (reference quirk A10);
 *   - exact bilinear weights instead of the texture unit's 8-bit fractions (A4);
 *   - runtime geometry/discretisation instead of literals (A6);
 *   - the two reference quirks A1 (stale Compton voxel) and A2 (swapped legs)
 *     are reproduced only when the corresponding XRD_QUIRK_* flag is set.
 * Loop order: the line integrals and per-voxel geometry are hoisted out of the
 * detector-energy loop (the reference recomputes them per energy block); the
 * arithmetic of every term is unchanged.
 */
#include "xrd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XO_PI 3.14159265358979323846

int xo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int xo_threads(int nthreads) { return nthreads > 0 ? nthreads : xo_max_threads(); }

/* GetIntersection.cu:9-57 */
int xo_get_intersection(double x1, double y1, double x2, double y2, int nx, int ny,
                        double *rx, double *ry) {
    const double X = (double)nx, Y = (double)ny;
    /* sides (x3,y3)->(x4,y4): y=0, x=NX, y=NY, x=0 (GetIntersection.cu:12-15) */
    const double x3[4] = {0.0, X, X, 0.0};
    const double x4[4] = {X, X, 0.0, 0.0};
    const double y3[4] = {0.0, 0.0, Y, Y};
    const double y4[4] = {0.0, Y, Y, 0.0};
    const double eps = 1e-5;
    const double x2_x1 = x2 - x1;
    const double y2_y1 = y2 - y1;
    for (int i = 0; i < 4; i++) {
        const double x4_x3 = x4[i] - x3[i];
        const double y3_y1 = y3[i] - y1;
        const double y4_y3 = y4[i] - y3[i];
        const double x3_x1 = x3[i] - x1;
        const double den = x4_x3 * y2_y1 - x2_x1 * y4_y3;
        if (den > eps || den < -eps) {
            const double ua = (x4_x3 * y3_y1 - x3_x1 * y4_y3) / den;
            if (ua > -eps && ua < 1.0 + eps) {
                const double ub = (x2_x1 * y3_y1 - x3_x1 * y2_y1) / den;
                if (ub > -eps && ub < 1.0 + eps) {
                    *rx = x1 + x2_x1 * ua;
                    *ry = y1 + y2_y1 * ua;
                    return i;
                }
            }
        }
    }
    *rx = NAN;
    *ry = NAN;
    return -1;
}

/* Sample count of attenuator.cu:44-69: q = 0,1,... while q*step <= L. */
static int32_t xo_sample_count(double L, double step) {
    if (!(L >= 0.0)) return 0; /* NaN entry: qr <= NaN is false (attenuator.cu:52) */
    double f = floor(L / step);
    int32_t n = (int32_t)f + 1;
    while ((double)n * step <= L) n++;
    while (n > 0 && (double)(n - 1) * step > L) n--;
    return n;
}

/* Unit direction from the entry point to the voxel (attenuator.cu:22-31). */
static void xo_ray_dir(double ex, double ey, double x2, double y2, double *ux, double *uy,
                       double *L) {
    const double lx = x2 - ex, ly = y2 - ey;
    const double lr = sqrt(lx * lx + ly * ly);
    *ux = lx / lr;
    *uy = ly / lr;
    *L = lr;
}

int xo_trace_ray(double x1, double y1, double x2, double y2, int nx, int ny, double step,
                 double *entry_x, double *entry_y, int32_t *count, int32_t max_samples,
                 int32_t *texels) {
    double ex, ey;
    const int side = xo_get_intersection(x1, y1, x2, y2, nx, ny, &ex, &ey);
    *entry_x = ex;
    *entry_y = ey;
    if (side < 0) {
        *count = 0;
        if (texels)
            for (int32_t q = 0; q < max_samples; q++) {
                texels[2 * q] = INT32_MIN;
                texels[2 * q + 1] = INT32_MIN;
            }
        return side;
    }
    double ux, uy, L;
    xo_ray_dir(ex, ey, x2, y2, &ux, &uy, &L);
    const int32_t n = xo_sample_count(L, step);
    *count = n;
    if (texels) {
        for (int32_t q = 0; q < max_samples; q++) {
            if (q < n) {
                const double qr = step * (double)q;
                const double px = ex + ux * qr;
                const double py = ey + uy * qr;
                texels[2 * q] = (int32_t)floor(px - 0.5);
                texels[2 * q + 1] = (int32_t)floor(py - 0.5);
            } else {
                texels[2 * q] = INT32_MIN;
                texels[2 * q + 1] = INT32_MIN;
            }
        }
    }
    return side;
}

static inline double xo_texel(const double *map, int nx, int ny, long i, long j) {
    if (i < 0 || j < 0 || i >= nx || j >= ny) return 0.0; /* cudaAddressModeBorder */
    return map[i + (long)nx * j];
}

double xo_bilinear(const double *map, int nx, int ny, double x, double y) {
    const double xt = x - 0.5, yt = y - 0.5;
    const double fi = floor(xt), fj = floor(yt);
    const double ax = xt - fi, ay = yt - fj;
    const long i = (long)fi, j = (long)fj;
    return (1.0 - ax) * (1.0 - ay) * xo_texel(map, nx, ny, i, j) +
           ax * (1.0 - ay) * xo_texel(map, nx, ny, i + 1, j) +
           (1.0 - ax) * ay * xo_texel(map, nx, ny, i, j + 1) +
           ax * ay * xo_texel(map, nx, ny, i + 1, j + 1);
}

/* Samples of attenuator.cu:44-69 from a given entry point (ex,ey) towards (x2,y2). */
static void xo_line_integrals_from(const double *maps, int nb, int nx, int ny, double step,
                                   double ex, double ey, double x2, double y2, double *out) {
    for (int b = 0; b < nb; b++) out[b] = 0.0;
    double ux, uy, L;
    xo_ray_dir(ex, ey, x2, y2, &ux, &uy, &L);
    const int32_t n = xo_sample_count(L, step);
    const long plane = (long)nx * ny;
    for (int32_t q = 0; q < n; q++) {
        const double qr = step * (double)q;
        const double px = ex + ux * qr;
        const double py = ey + uy * qr;
        for (int b = 0; b < nb; b++) out[b] += xo_bilinear(maps + b * plane, nx, ny, px, py);
    }
    for (int b = 0; b < nb; b++) out[b] *= step;
}

/* attenuator.cu:9-109: step * sum over samples of the bilinear (a1,a2,...). */
void xo_line_integrals(const double *maps, int nb, int nx, int ny, double step, double x1,
                       double y1, double x2, double y2, double *out) {
    double ex, ey;
    if (xo_get_intersection(x1, y1, x2, y2, nx, ny, &ex, &ey) < 0) {
        for (int b = 0; b < nb; b++) out[b] = 0.0;
        return;
    }
    xo_line_integrals_from(maps, nb, nx, ny, step, ex, ey, x2, y2, out);
}

/* Trilinear value of volume [nx x ny x nz] at (x,y,z), zero outside, voxel (i,j,k)
 * centred at (i+0.5, j+0.5, k+0.5). */
static double xo_trilinear(const double *vol, int nx, int ny, int nz, double x, double y,
                           double z) {
    const double zt = z - 0.5;
    const double fk = floor(zt);
    const double az = zt - fk;
    const long k = (long)fk;
    const long plane = (long)nx * ny;
    const double v0 = (k >= 0 && k < nz) ? xo_bilinear(vol + k * plane, nx, ny, x, y) : 0.0;
    const double v1 = (k + 1 >= 0 && k + 1 < nz) ? xo_bilinear(vol + (k + 1) * plane, nx, ny, x, y)
                                                 : 0.0;
    return (1.0 - az) * v0 + az * v1;
}

void xo_line_integrals_3d(const double *vols, int nb, int nx, int ny, int nz, double step,
                          double x1, double y1, double z1, double x2, double y2, double z2,
                          double *out) {
    for (int b = 0; b < nb; b++) out[b] = 0.0;
    double ex, ey;
    if (xo_get_intersection(x1, y1, x2, y2, nx, ny, &ex, &ey) < 0) return;
    double ux, uy, L;
    xo_ray_dir(ex, ey, x2, y2, &ux, &uy, &L);
    const int32_t n = xo_sample_count(L, step);
    const double hx = x2 - x1, hy = y2 - y1;
    const double H = sqrt(hx * hx + hy * hy);
    const long vsize = (long)nx * ny * nz;
    for (int32_t q = 0; q < n; q++) {
        const double qr = step * (double)q;
        const double px = ex + ux * qr;
        const double py = ey + uy * qr;
        const double pz = z2 + (z1 - z2) * ((L - qr) / H);
        for (int b = 0; b < nb; b++) out[b] += xo_trilinear(vols + b * vsize, nx, ny, nz, px, py, pz);
    }
    for (int b = 0; b < nb; b++) out[b] *= step;
}

/* find_src_nrg_range.cu:8-25 (the window is shifted, never truncated). */
void xo_find_src_nrg_range(int nrg, int n_energy, int halfwidth, int *lo, int *hi) {
    int mn = nrg - halfwidth;
    if (mn < 0) mn = 0;
    int mx = mn + 2 * halfwidth;
    if (mx >= n_energy) {
        mx = n_energy - 1;
        mn = mx - 2 * halfwidth;
    }
    *lo = mn;
    *hi = mx;
}

/* wvt_value.cu:8-23 */
void xo_wvt_value(int n, int n_q, double q_min, double q_max, double dq0, double *q,
                  double *dq) {
    const double NW = (double)n_q;
    const double c = (q_max - q_min - dq0 * NW) / (NW * NW);
    const double e1 = q_min + (double)n * dq0 + c * (double)n * (double)n;
    const double e2 = q_min + (double)(n + 1) * dq0 + c * (double)(n + 1) * (double)(n + 1);
    *q = 0.5 * (e1 + e2);
    *dq = e2 - e1;
}

/* integrate_diffraction.cu:81-128 (view/pixel) and :192-255 (voxel). */
void xo_scatter_geometry(const xrd_geometry *g, int view, int col, int row, int ix, int iy,
                         xo_scatter_geom *o) {
    const double phi = g->view_angle0 + (double)view * g->view_angle_step;
    const double sp = sin(phi), cp = cos(phi);
    const double sx = g->center_x + g->r_src * sp;
    const double sy = g->center_y - g->r_src * cp;
    const double sz = 0.0;
    const double rdet = ((double)col - (double)g->n_cols / 2.0 + 0.5) * g->col_pitch;
    const double dx = g->center_x - g->r_det * sp + rdet * cp;
    const double dy = g->center_y + g->r_det * cp + rdet * sp;
    const double dz = g->det_z0 + (double)row * g->row_pitch;
    const double ndx = sp, ndy = -cp, ndz = 0.0;
    const double sta = sin(g->target_angle), cta = cos(g->target_angle);
    const double nsx = -sp * sta, nsy = cp * sta, nsz = -cta;
    const double adet = g->col_pitch * g->row_pitch;
    const double dvox = g->voxel_size;
    const double tw = tan(g->wedge_angle), tw2 = tan(g->wedge_angle / 2.0);

    const double vx = (double)ix + 0.5, vy = (double)iy + 0.5;
    const double ax = vx - sx, ay = vy - sy;
    const double proj = (ax * nsx + ay * nsy) / sqrt(nsx * nsx + nsy * nsy);
    const double vz = -proj * tw2;
    const double hvox = proj * tw;
    const double az = vz - sz;
    const double bx = dx - vx, by = dy - vy, bz = dz - vz;
    const double ab = ax * bx + ay * by + az * bz;
    const double an2 = ax * ax + ay * ay + az * az;
    const double bn2 = bx * bx + by * by + bz * bz;
    const double cth = ab * (1.0 / sqrt(an2)) * (1.0 / sqrt(bn2));
    const double sth = sqrt(fabs(1.0 - cth) / 2.0);

    const double Sx = bx - ab * ax / an2, Sy = by - ab * ay / an2, Sz = bz - ab * az / an2;
    const double Dx = -ax + ab * bx / bn2, Dy = -ay + ab * by / bn2, Dz = -az + ab * bz / bn2;
    const double Vx = -(Sx + Dx), Vy = -(Sy + Dy), Vz = -(Sz + Dz);
    const double Dn = Dx * ndx + Dy * ndy + Dz * ndz;
    const double Sn = Sx * nsx + Sy * nsy + Sz * nsz;
    const double dD2 = (Dx * Dx + Dy * Dy + Dz * Dz - Dn * Dn) * adet / 12.0;
    const double vV2 = (dvox * dvox * (Vx * Vx + Vy * Vy) + hvox * hvox * Vz * Vz) / 12.0;
    const double sS2 = (Sx * Sx + Sy * Sy + Sz * Sz - Sn * Sn) * g->focal_spot * g->focal_spot / 12.0;

    const double rdotn = -sp * ax + cp * ay;
    const double cospsi = rdotn * (1.0 / sqrt(ax * ax + ay * ay));
    const double g_sv = cospsi * (1.0 / sqrt(an2)) * tw;
    const double g_vd = -(ndx * bx + ndy * by) * (1.0 / sqrt(bn2 * bn2 * bn2));

    o->src_x = sx;
    o->src_y = sy;
    o->det_x = dx;
    o->det_y = dy;
    o->cos_theta = cth;
    o->sin_half = sth;
    o->anorm2 = an2;
    o->bnorm2 = bn2;
    o->res2 = sS2 + vV2 + dD2;
    o->geom = g_sv * g_vd;
    o->rayleigh = (1.0 + cth * cth) / 2.0;
    o->vox_z = vz;
    o->det_z = dz;
}

/* Centre and variance of the q-resolution Gaussian of source energy Es
 * (integrate_diffraction.cu:284-288): wvt = 2 Es sin(theta/2)/hbarc and
 * sigma^2 = (2 sin(theta/2)/hbarc)^2 dE^2/12
 *         + (Es/hbarc)^2 (sS2 + vV2 + dD2) / (2 sin(theta/2))^2 / |a|^2 / |b|^2. */
static void xo_q_resolution(const xrd_discretization *d, const xo_scatter_geom *sg, double Es,
                            double *wvt, double *sig2) {
    const double s = sg->sin_half;
    const double k1 = 2.0 * s / d->hbarc;
    const double k2 = Es / d->hbarc;
    *wvt = 2.0 * Es / d->hbarc * s;
    *sig2 = k1 * k1 * d->e_width * d->e_width / 12.0 +
            k2 * k2 * sg->res2 / (2.0 * s) / (2.0 * s) / sg->anorm2 / sg->bnorm2;
}

/* floor(t) clamped to [0, n-1]; NaN -> 0 (the clamps of integrate_diffraction.cu:187-189,
 * :337-339; cvt of NaN to int gives 0). */
static int xo_lut_index(double t, int n) {
    const double f = floor(t);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)n) return n - 1;
    return (int)f;
}

static double xo_s2_u8(double v) { /* integrate_diffraction.cu:66-74 */
    const double f = floor(v);
    if (!(f >= 0.0)) return 0.0;
    if (f > 255.0) return 255.0;
    return f;
}

/* 1/KN(e) quadratic (integrate_diffraction.cu:292, Forward_XRD.cu:51-53) */
static inline double xo_f2(const xrd_discretization *d, double e) {
    return d->kn0 + d->kn1 * e + d->kn2 * e * e;
}

/* Compton channel match (integrate_diffraction.cu:323-328): fraction of the
 * source bin [es1,es2], shifted down by eratio, that lands in [ed1,ed2]. */
double xo_channel_match(double es1, double es2, double eratio, double ed1, double ed2) {
    const double Eo1 = es1 / eratio, Eo2 = es2 / eratio;
    const double lo = Eo1 > ed1 ? Eo1 : ed1;
    const double hi = Eo2 < ed2 ? Eo2 : ed2;
    double cm = hi - lo;
    if (!(cm > 0.0)) cm = 0.0;
    return cm / (Eo2 - Eo1);
}

typedef struct {
    int32_t ix, iy, mat;
    double a1c, a2c; /* (a1,a2) used by the Compton term */
} xo_voxel;

/* Volume attenuation for xo_forward_impl: vol = [a1 volume, a2 volume] of nz
 * slices; the scattering slice sits at z_slice (slice centre). */
typedef struct {
    const double *vol;
    int nz;
    double z_slice;
} xo_volume;

static int xo_forward_impl(const xrd_geometry *g, const xrd_discretization *d,
                           const double *spectrum, const double *ifine, const int32_t *seg,
                           const double *a1, const double *a2, const double *s2_table,
                           double lut_q_min, double lut_q_max, double lut_a_min,
                           double lut_a_max, const int32_t *views, int nvl,
                           const int32_t *cols, int ncl, double *out_N, double *out_C,
                           double *out_A, int nthreads, const xo_volume *volm);

int xo_forward(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
               const double *ifine, const int32_t *seg, const double *a1, const double *a2,
               const double *s2_table, double lut_q_min, double lut_q_max, double lut_a_min,
               double lut_a_max, const int32_t *views, int nvl, const int32_t *cols, int ncl,
               double *out_N, double *out_C, double *out_A, int nthreads) {
    return xo_forward_impl(g, d, spectrum, ifine, seg, a1, a2, s2_table, lut_q_min, lut_q_max,
                           lut_a_min, lut_a_max, views, nvl, cols, ncl, out_N, out_C, out_A,
                           nthreads, NULL);
}

int xo_forward_slice(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                     const double *ifine, const int32_t *seg_vol, const double *a1_vol,
                     const double *a2_vol, int n_slices, int slice, const double *s2_table,
                     double lut_q_min, double lut_q_max, double lut_a_min, double lut_a_max,
                     const int32_t *views, int nvl, const int32_t *cols, int ncl,
                     double *out_N, double *out_C, double *out_A, int nthreads) {
    if (n_slices <= 0 || slice < 0 || slice >= n_slices) return XRD_ERR_DIMS;
    const long plane = (long)g->nx * g->ny;
    const int32_t *seg = seg_vol + plane * slice;
    const double *a1 = a1_vol + plane * slice, *a2 = a2_vol + plane * slice;
    if (n_slices == 1)
        return xo_forward_impl(g, d, spectrum, ifine, seg, a1, a2, s2_table, lut_q_min,
                               lut_q_max, lut_a_min, lut_a_max, views, nvl, cols, ncl, out_N,
                               out_C, out_A, nthreads, NULL);
    double *vol = (double *)malloc(sizeof(double) * 2 * plane * n_slices);
    memcpy(vol, a1_vol, sizeof(double) * plane * n_slices);
    memcpy(vol + plane * n_slices, a2_vol, sizeof(double) * plane * n_slices);
    xo_volume vm = {vol, n_slices, (double)slice + 0.5};
    const int rc = xo_forward_impl(g, d, spectrum, ifine, seg, a1, a2, s2_table, lut_q_min,
                                   lut_q_max, lut_a_min, lut_a_max, views, nvl, cols, ncl,
                                   out_N, out_C, out_A, nthreads, &vm);
    free(vol);
    return rc;
}

static int xo_forward_impl(const xrd_geometry *g, const xrd_discretization *d,
                           const double *spectrum, const double *ifine, const int32_t *seg,
                           const double *a1, const double *a2, const double *s2_table,
                           double lut_q_min, double lut_q_max, double lut_a_min,
                           double lut_a_max, const int32_t *views, int nvl,
                           const int32_t *cols, int ncl, double *out_N, double *out_C,
                           double *out_A, int nthreads, const xo_volume *volm) {
    const int NX = g->nx, NY = g->ny, V = g->n_views, COL = g->n_cols, ROW = g->n_rows;
    /* detector pixels computed: the listed columns (all when cols is NULL) of every row */
    const int NCL = cols ? ncl : COL;
    const int DET = NCL * ROW;
    const int E = d->n_energy, DN = d->energy_halfwidth, Q = d->n_q, QF = d->n_qf;
    const int M = d->n_mat, LQ = d->lut_nq, LA = d->lut_na;
    if (NX <= 0 || NY <= 0 || V <= 0 || COL <= 0 || ROW <= 0 || NCL <= 0 || E <= 2 * DN || Q <= 0 ||
        QF <= 0 || M <= 0 || LQ <= 0 || LA <= 0)
        return XRD_ERR_DIMS;
    int32_t *vlist = (int32_t *)malloc(sizeof(int32_t) * (views ? nvl : V));
    if (!views) {
        nvl = V;
        for (int v = 0; v < V; v++) vlist[v] = v;
    } else {
        memcpy(vlist, views, sizeof(int32_t) * nvl);
    }
    const int swap = (d->flags & XRD_QUIRK_SWAPPED_LEGS) != 0;
    const int stale = (d->flags & XRD_QUIRK_STALE_COMPTON_VOXEL) != 0;

    /* Non-air voxels in the reference's loop order: nx outer, ny inner
     * (integrate_diffraction.cu:173-178). */
    xo_voxel *vox = (xo_voxel *)malloc(sizeof(xo_voxel) * (size_t)NX * NY);
    int nvox = 0;
    double prev_a1 = 0.0, prev_a2 = 0.0;
    for (int ix = 0; ix < NX; ix++)
        for (int iy = 0; iy < NY; iy++) {
            const int s = seg[ix + NX * iy];
            if (s <= 0 || s > M) continue;
            xo_voxel *p = &vox[nvox++];
            p->ix = ix;
            p->iy = iy;
            p->mat = s - 1;
            const double c1 = a1[ix + NX * iy], c2 = a2[ix + NX * iy];
            if (stale) {
                p->a1c = prev_a1;
                p->a2c = prev_a2;
            } else {
                p->a1c = c1;
                p->a2c = c2;
            }
            prev_a1 = c1;
            prev_a2 = c2;
        }

    /* Energy and q axes (integrate_diffraction.cu:138-142, :269-275; wvt_value.cu) */
    double *Es1 = (double *)malloc(sizeof(double) * E * 3);
    double *Es2 = Es1 + E, *Esm = Es1 + 2 * E;
    for (int e = 0; e < E; e++) {
        Es1[e] = d->e_min + (double)e * d->e_width;
        Es2[e] = d->e_min + (double)(e + 1) * d->e_width;
        Esm[e] = 0.5 * (Es1[e] + Es2[e]);
    }
    double *qc = (double *)malloc(sizeof(double) * (2 * Q + QF));
    double *dqc = qc + Q, *qf = qc + 2 * Q;
    for (int n = 0; n < Q; n++) xo_wvt_value(n, Q, d->q_min, d->q_max, d->q_dq0, &qc[n], &dqc[n]);
    const double dq_fine = (d->q_max - d->q_min) / (double)QF;
    for (int n = 0; n < QF; n++) qf[n] = d->q_min + dq_fine * (0.5 + (double)n);
    double *s2u = (double *)malloc(sizeof(double) * LQ * LA);
    for (int i = 0; i < LQ * LA; i++) s2u[i] = xo_s2_u8(s2_table[i]);
    const double da = (lut_a_max - lut_a_min) / (double)LA;
    const double dql = (lut_q_max - lut_q_min) / (double)LQ;
    const double rs2pi = 1.0 / sqrt(2.0 * XO_PI);

    /* Source-leg integrals depend only on (view, voxel): hoisted. */
    const long plane = (long)NX * NY;
    double *maps = (double *)malloc(sizeof(double) * 2 * plane);
    memcpy(maps, a1, sizeof(double) * plane);
    memcpy(maps + plane, a2, sizeof(double) * plane);
    double *Ls = (double *)malloc(sizeof(double) * 2 * (size_t)nvl * (nvox > 0 ? nvox : 1));
    const long nrows = (long)nvl * DET * E;
    const int nt = xo_threads(nthreads);

#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (long t = 0; t < (long)nvl * nvox; t++) {
        const int vi = (int)(t / (nvox > 0 ? nvox : 1)), k = (int)(t % (nvox > 0 ? nvox : 1));
        xo_scatter_geom sg;
        xo_scatter_geometry(g, vlist[vi], 0, 0, vox[k].ix, vox[k].iy, &sg);
        if (volm)  /* the focal spot at the slice centre, the voxel at its wedge height */
            xo_line_integrals_3d(volm->vol, 2, NX, NY, volm->nz, d->ray_step, sg.src_x, sg.src_y,
                                 volm->z_slice, vox[k].ix + 0.5, vox[k].iy + 0.5,
                                 volm->z_slice + sg.vox_z, &Ls[2 * t]);
        else
            xo_line_integrals(maps, 2, NX, NY, d->ray_step, sg.src_x, sg.src_y, vox[k].ix + 0.5,
                              vox[k].iy + 0.5, &Ls[2 * t]);
    }

#pragma omp parallel num_threads(nt)
    {
        double *T = (double *)malloc(sizeof(double) * (size_t)E * Q);  /* coarse G*dq   */
        double *TF = (double *)malloc(sizeof(double) * E);              /* fine sum      */
        double *P = (double *)malloc(sizeof(double) * E * 2);           /* prefactors    */
        double *KC = P + E;                                             /* Compton pieces */
        double *accA = out_A ? (double *)malloc(sizeof(double) * (size_t)E * M * Q) : NULL;
        double *accN = (double *)malloc(sizeof(double) * E * 2);
        double *accC = accN + E;
#pragma omp for schedule(dynamic, 1)
        for (long vp = 0; vp < (long)nvl * DET; vp++) {
            const int vi = (int)(vp / DET), det = (int)(vp % DET);
            const int ci = det % NCL, row = det / NCL;
            const int col = cols ? cols[ci] : ci;
            const int view = vlist[vi];
            if (accA) memset(accA, 0, sizeof(double) * (size_t)E * M * Q);
            memset(accN, 0, sizeof(double) * E * 2);
            for (int k = 0; k < nvox; k++) {
                const xo_voxel *vx = &vox[k];
                const int mat = vx->mat;
                xo_scatter_geom sg;
                xo_scatter_geometry(g, view, col, row, vx->ix, vx->iy, &sg);
                double Ld[2];
                if (volm)
                    xo_line_integrals_3d(volm->vol, 2, NX, NY, volm->nz, d->ray_step, sg.det_x,
                                         sg.det_y, volm->z_slice + sg.det_z, vx->ix + 0.5,
                                         vx->iy + 0.5, volm->z_slice + sg.vox_z, Ld);
                else
                    xo_line_integrals(maps, 2, NX, NY, d->ray_step, sg.det_x, sg.det_y,
                                      vx->ix + 0.5, vx->iy + 0.5, Ld);
                const double *Lsrc = &Ls[2 * ((long)vi * nvox + k)];
                /* Leg feeding Pin and the leg feeding the Compton Pout
                 * (integrate_diffraction.cu:259-260, :295-296, :347). */
                const double *Lin = swap ? Ld : Lsrc;
                const double *Lout = swap ? Lsrc : Ld;
                const double s = sg.sin_half, cth = sg.cos_theta;
                const double gp = sg.geom;
                /* Compton LUT material index (integrate_diffraction.cu:186-189) */
                const int na =
                    xo_lut_index((log2(fabs(vx->a1c / vx->a2c)) - lut_a_min) / da, LA);
                for (int e = 0; e < E; e++) {
                    const double Es = Esm[e];
                    const double ein = Es / d->mc2;
                    double wvt, sig2;
                    xo_q_resolution(d, &sg, Es, &wvt, &sig2);
                    const double f1 = ein * ein * ein, f2 = xo_f2(d, ein);
                    const double Pin = exp(-(Lin[0] / f1 + Lin[1] / f2));
                    const double Pout = exp(-(Lout[0] / f1 + Lout[1] / f2));
                    const double norm = rs2pi / sqrt(sig2);
                    P[e] = Pin * Pout * gp * sg.rayleigh;
                    double sf = 0.0;
                    for (int n = 0; n < QF; n++) {
                        const double dd = wvt - qf[n];
                        sf += exp(-dd * dd / (2.0 * sig2)) * norm * dq_fine *
                              ifine[n + (long)QF * mat];
                    }
                    TF[e] = sf;
                    if (accA)
                        for (int n = 0; n < Q; n++) {
                            const double dd = wvt - qc[n];
                            T[(long)e * Q + n] = exp(-dd * dd / (2.0 * sig2)) * norm * dqc[n];
                        }
                    /* Compton (integrate_diffraction.cu:318-356), without channel_match and RF */
                    const double er = 1.0 + ein * (1.0 - cth);
                    const double eout = ein / er;
                    const double w = ein * s;
                    const double l2w = log2(w * sqrt(er + w * w) / er);
                    const int nq = xo_lut_index((l2w - lut_q_min) / dql, LQ);
                    const double g1 = eout * eout * eout, g2 = xo_f2(d, eout);
                    const double Pc = exp(-(Lout[0] / g1 + Lout[1] / g2));
                    const double cf = er * er * (er + 1.0 / er - 1.0 + cth * cth) / 2.0;
                    KC[e] = Pin * Pc * gp * cf * vx->a2c * s2u[nq + LQ * na];
                }
                for (int ed = 0; ed < E; ed++) {
                    int lo, hi;
                    xo_find_src_nrg_range(ed, E, DN, &lo, &hi);
                    const double Ed1 = d->e_min + (double)ed * d->e_width;
                    const double Ed2 = d->e_min + (double)(ed + 1) * d->e_width;
                    for (int e = lo; e <= hi; e++) {
                        const double RF = spectrum[e + (long)E * ed];
                        accN[ed] += RF * P[e] * TF[e];
                        if (accA) {
                            const double pr = RF * P[e];
                            double *dst = accA + ((long)ed * M + mat) * Q;
                            const double *src = T + (long)e * Q;
                            for (int n = 0; n < Q; n++) dst[n] += pr * src[n];
                        }
                        const double ein = Esm[e] / d->mc2;
                        const double er = 1.0 + ein * (1.0 - cth);
                        accC[ed] += xo_channel_match(Es1[e], Es2[e], er, Ed1, Ed2) * RF * KC[e];
                    }
                }
            }
            for (int ed = 0; ed < E; ed++) {
                const long r = vi + (long)nvl * (det + (long)DET * ed);
                out_N[r] = accN[ed];
                out_C[r] = accC[ed];
                if (accA)
                    for (int m = 0; m < M; m++)
                        for (int n = 0; n < Q; n++)
                            out_A[n + (long)Q * (m + (long)M * r)] = accA[((long)ed * M + m) * Q + n];
            }
        }
        free(T);
        free(TF);
        free(P);
        free(accA);
        free(accN);
    }
    (void)nrows;
    free(Ls);
    free(maps);
    free(s2u);
    free(qc);
    free(Es1);
    free(vox);
    free(vlist);
    return XRD_OK;
}

/* ---- explicit-matrix operators and MLEM (main.m:141-176) -------------------- */

void xo_project(const double *A, int64_t n_rows, int n_cols, const double *x, double *out,
                int nthreads) {
    const int nt = xo_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t r = 0; r < n_rows; r++) {
        const double *a = A + r * (int64_t)n_cols;
        double s = 0.0;
        for (int j = 0; j < n_cols; j++) s += a[j] * x[j];
        out[r] = s;
    }
}

/* Column sums of diag(w) A, per-thread partials merged in thread order. */
static void xo_colsum(const double *A, int64_t n_rows, int n_cols, const double *w,
                      double *out, int nthreads) {
    const int nt = xo_threads(nthreads);
    double *part = (double *)calloc((size_t)nt * n_cols, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *p = part + (size_t)tid * n_cols;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < n_rows; r++) {
            const double *a = A + r * (int64_t)n_cols;
            const double wr = w ? w[r] : 1.0;
            for (int j = 0; j < n_cols; j++) p[j] += a[j] * wr;
        }
    }
    for (int j = 0; j < n_cols; j++) {
        double s = 0.0;
        for (int t = 0; t < nt; t++) s += part[(size_t)t * n_cols + j];
        out[j] = s;
    }
    free(part);
}

void xo_backproject(const double *A, int64_t n_rows, int n_cols, const double *w,
                    double *out, int nthreads) {
    xo_colsum(A, n_rows, n_cols, w, out, nthreads);
}

void xo_sensitivity(const double *A, int64_t n_rows, int n_cols, double *out, int nthreads) {
    xo_colsum(A, n_rows, n_cols, NULL, out, nthreads);
}

void xo_mlem_update(const double *A, int64_t n_rows, int n_cols, const double *y,
                    const double *c, const double *s, double *x, int nthreads) {
    double *f = (double *)malloc(sizeof(double) * n_rows);
    double *g = (double *)malloc(sizeof(double) * n_cols);
    const int nt = xo_threads(nthreads);
    xo_project(A, n_rows, n_cols, x, f, nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t r = 0; r < n_rows; r++) f[r] = (y[r] - 0.0 * c[r]) / (f[r] + c[r] + 1e-20);
    xo_backproject(A, n_rows, n_cols, f, g, nthreads);
    for (int j = 0; j < n_cols; j++) {
        double v = x[j] * g[j] / (s[j] + 1e-7);
        if (!isfinite(v)) v = 0.0; /* main.m:163 */
        if (v < 0.0) v = 0.0;      /* main.m:176 */
        x[j] = v;
    }
    free(f);
    free(g);
}

void xo_mlem(const double *A, int64_t n_rows, int n_cols, const double *y, const double *c,
             double *x, int n_iter, int nthreads) {
    double *s = (double *)malloc(sizeof(double) * n_cols);
    xo_sensitivity(A, n_rows, n_cols, s, nthreads);
    for (int it = 0; it < n_iter; it++) xo_mlem_update(A, n_rows, n_cols, y, c, s, x, nthreads);
    free(s);
}

double xo_loglik(const double *A, int64_t n_rows, int n_cols, const double *y,
                 const double *c, const double *x, int nthreads) {
    double *f = (double *)malloc(sizeof(double) * n_rows);
    xo_project(A, n_rows, n_cols, x, f, nthreads);
    double ll = 0.0;
    for (int64_t r = 0; r < n_rows; r++) {
        const double fr = f[r] + c[r] + 1e-20;
        ll += (y[r] > 0.0 ? y[r] * log(fr) : 0.0) - fr;
    }
    free(f);
    return ll;
}

/* ---- coefficient model and its vector-Jacobian product ------------------------
 * The differentiable form of integrate_diffraction.cu:266-315 used by the VJP:
 * each non-air voxel carries its own form factor F_vox(q_n) = sum_k coef[vox,k]
 * basis[k,n] on the coarse axis, and the attenuation is mu_e = sum_b beta[b,e] m_b
 * (the reference is B = 2 with beta = (1/e^3, 1/f2(e)), integrate_diffraction.cu:291-296;
 * a full per-energy map is B = NRG with beta = identity).
 *   yhat[v,p,ed] = sum_vox sum_{e in win(ed)} RF[e,ed] P(e) S(e)
 *   P(e) = exp(-sum_b beta[b,e] (Ls_b + Ld_b)) g R / sqrt(2 pi sigma_e^2)
 *   S(e) = sum_n exp(-(wvt_e - q_n)^2 / (2 sigma_e^2)) dq_n F_vox(q_n)
 * The Compton term is a constant background here. */

/* Adds gamma[b]*step*(bilinear weight) of every sample of the ray from the entry
 * point (ex,ey) towards (x2,y2) to grad[b][pix]. */
static void xo_line_scatter_from(double *grad, int nb, int nx, int ny, double step, double ex,
                                 double ey, double x2, double y2, const double *gamma) {
    double ux, uy, L;
    xo_ray_dir(ex, ey, x2, y2, &ux, &uy, &L);
    const int32_t n = xo_sample_count(L, step);
    const long plane = (long)nx * ny;
    for (int32_t q = 0; q < n; q++) {
        const double qr = step * (double)q;
        const double px = ex + ux * qr, py = ey + uy * qr;
        const double xt = px - 0.5, yt = py - 0.5;
        const double fi = floor(xt), fj = floor(yt);
        const double ax = xt - fi, ay = yt - fj;
        const long i = (long)fi, j = (long)fj;
        const long ci[4] = {i, i + 1, i, i + 1}, cj[4] = {j, j, j + 1, j + 1};
        const double w[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
        for (int c = 0; c < 4; c++) {
            if (ci[c] < 0 || cj[c] < 0 || ci[c] >= nx || cj[c] >= ny) continue;
            const long pix = ci[c] + (long)nx * cj[c];
            for (int b = 0; b < nb; b++) grad[b * plane + pix] += gamma[b] * step * w[c];
        }
    }
}

/* xo_line_scatter_from for the attenuator ray (x1,y1)->(x2,y2). */
static void xo_line_scatter(double *grad, int nb, int nx, int ny, double step, double x1,
                            double y1, double x2, double y2, const double *gamma) {
    double ex, ey;
    if (xo_get_intersection(x1, y1, x2, y2, nx, ny, &ex, &ey) < 0) return;
    xo_line_scatter_from(grad, nb, nx, ny, step, ex, ey, x2, y2, gamma);
}

typedef struct {
    int E, Q, K, B, nvox;
    double *qc, *dqc, *Esm;
    int32_t *vix, *viy;
} xo_cm_setup;

static int xo_cm_prepare(const xrd_geometry *g, const xrd_discretization *d,
                         const int32_t *seg, int K, int B, xo_cm_setup *s) {
    const int NX = g->nx, NY = g->ny, E = d->n_energy, Q = d->n_q;
    if (NX <= 0 || NY <= 0 || E <= 2 * d->energy_halfwidth || Q <= 0 || K <= 0 || B <= 0)
        return XRD_ERR_DIMS;
    s->E = E;
    s->Q = Q;
    s->K = K;
    s->B = B;
    s->qc = (double *)malloc(sizeof(double) * (2 * Q + E));
    s->dqc = s->qc + Q;
    s->Esm = s->qc + 2 * Q;
    for (int n = 0; n < Q; n++) xo_wvt_value(n, Q, d->q_min, d->q_max, d->q_dq0, &s->qc[n], &s->dqc[n]);
    for (int e = 0; e < E; e++) {
        const double e1 = d->e_min + (double)e * d->e_width;
        const double e2 = d->e_min + (double)(e + 1) * d->e_width;
        s->Esm[e] = 0.5 * (e1 + e2);
    }
    s->vix = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)NX * NY);
    s->viy = s->vix + (size_t)NX * NY;
    s->nvox = 0;
    for (int ix = 0; ix < NX; ix++)
        for (int iy = 0; iy < NY; iy++)
            if (seg[ix + NX * iy] > 0) {
                s->vix[s->nvox] = ix;
                s->viy[s->nvox] = iy;
                s->nvox++;
            }
    return XRD_OK;
}

static void xo_cm_free(xo_cm_setup *s) {
    free(s->qc);
    free(s->vix);
}

/* Per-(view, pixel, voxel) pieces: P[e], sigma2[e], wvt[e], line integrals. */
static void xo_cm_triple(const xrd_geometry *g, const xrd_discretization *d,
                         const xo_cm_setup *s, const double *maps, const double *beta, int view,
                         int col, int row, int k, double *P, double *sig2, double *wvt,
                         double *Ls, double *Ld, xo_scatter_geom *sg) {
    const int NX = g->nx, NY = g->ny, B = s->B, E = s->E;
    const int ix = s->vix[k], iy = s->viy[k];
    xo_scatter_geometry(g, view, col, row, ix, iy, sg);
    xo_line_integrals(maps, B, NX, NY, d->ray_step, sg->src_x, sg->src_y, ix + 0.5, iy + 0.5, Ls);
    xo_line_integrals(maps, B, NX, NY, d->ray_step, sg->det_x, sg->det_y, ix + 0.5, iy + 0.5, Ld);
    const double rs2pi = 1.0 / sqrt(2.0 * XO_PI);
    for (int e = 0; e < E; e++) {
        double mu = 0.0;
        for (int b = 0; b < B; b++) mu += beta[(long)b * E + e] * (Ls[b] + Ld[b]);
        const double Es = s->Esm[e];
        xo_q_resolution(d, sg, Es, &wvt[e], &sig2[e]);
        P[e] = exp(-mu) * sg->geom * sg->rayleigh * rs2pi / sqrt(sig2[e]);
    }
}

int xo_forward_coeffs(const xrd_geometry *g, const xrd_discretization *d,
                      const double *spectrum, const int32_t *seg, const double *coef, int K,
                      const double *basis, int B, const double *maps, const double *beta,
                      double *out_yhat, int nthreads) {
    xo_cm_setup s;
    int rc = xo_cm_prepare(g, d, seg, K, B, &s);
    if (rc) return rc;
    const int V = g->n_views, DET = g->n_cols * g->n_rows, E = s.E, Q = s.Q, DN = d->energy_halfwidth;
    const long npix = (long)g->nx * g->ny;
    const int nt = xo_threads(nthreads);
#pragma omp parallel num_threads(nt)
    {
        double *P = (double *)malloc(sizeof(double) * (5 * E + 2 * B + Q));
        double *sig2 = P + E, *wvt = sig2 + E, *S = wvt + E, *acc = S + E;
        double *Ls = acc + E, *Ld = Ls + B, *F = Ld + B;
#pragma omp for schedule(dynamic, 1)
        for (long vp = 0; vp < (long)V * DET; vp++) {
            const int v = (int)(vp % V), det = (int)(vp / V);
            const int col = det % g->n_cols, row = det / g->n_cols;
            for (int e = 0; e < E; e++) acc[e] = 0.0;
            for (int k = 0; k < s.nvox; k++) {
                xo_scatter_geom sg;
                xo_cm_triple(g, d, &s, maps, beta, v, col, row, k, P, sig2, wvt, Ls, Ld, &sg);
                const long pix = s.vix[k] + (long)g->nx * s.viy[k];
                for (int n = 0; n < Q; n++) {
                    double f = 0.0;
                    for (int kk = 0; kk < K; kk++) f += coef[pix * K + kk] * basis[(long)kk * Q + n];
                    F[n] = f;
                }
                for (int e = 0; e < E; e++) {
                    double sv = 0.0;
                    for (int n = 0; n < Q; n++) {
                        const double dd = wvt[e] - s.qc[n];
                        sv += exp(-dd * dd / (2.0 * sig2[e])) * s.dqc[n] * F[n];
                    }
                    acc[e] += P[e] * sv;
                }
            }
            for (int ed = 0; ed < E; ed++) {
                int lo, hi;
                xo_find_src_nrg_range(ed, E, DN, &lo, &hi);
                double y = 0.0;
                for (int e = lo; e <= hi; e++) y += spectrum[e + (long)E * ed] * acc[e];
                out_yhat[v + (long)V * (det + (long)DET * ed)] = y;
            }
        }
        free(P);
    }
    (void)npix;
    xo_cm_free(&s);
    return XRD_OK;
}

int xo_vjp_coeffs(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                  const int32_t *seg, const double *coef, int K, const double *basis, int B,
                  const double *maps, const double *beta, const double *w, double *grad_coef,
                  double *grad_maps, int nthreads) {
    xo_cm_setup s;
    int rc = xo_cm_prepare(g, d, seg, K, B, &s);
    if (rc) return rc;
    const int V = g->n_views, DET = g->n_cols * g->n_rows, E = s.E, Q = s.Q, DN = d->energy_halfwidth;
    const int NX = g->nx, NY = g->ny;
    const long npix = (long)NX * NY;
    const int nt = xo_threads(nthreads);
    double *hpart = (double *)calloc((size_t)nt * s.nvox * Q, sizeof(double));
    double *gpart = grad_maps ? (double *)calloc((size_t)nt * B * npix, sizeof(double)) : NULL;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *H = hpart + (size_t)tid * s.nvox * Q;
        double *GM = gpart ? gpart + (size_t)tid * B * npix : NULL;
        double *P = (double *)malloc(sizeof(double) * (5 * E + 3 * B + Q));
        double *sig2 = P + E, *wvt = sig2 + E, *S = wvt + E, *Wt = S + E;
        double *Ls = Wt + E, *Ld = Ls + B, *gam = Ld + B, *F = gam + B;
#pragma omp for schedule(dynamic, 1)
        for (long vp = 0; vp < (long)V * DET; vp++) {
            const int v = (int)(vp % V), det = (int)(vp / V);
            const int col = det % g->n_cols, row = det / g->n_cols;
            /* transpose of the detector-response window: Wt[e] = sum_ed RF[e,ed] w[ed] */
            for (int e = 0; e < E; e++) Wt[e] = 0.0;
            for (int ed = 0; ed < E; ed++) {
                int lo, hi;
                xo_find_src_nrg_range(ed, E, DN, &lo, &hi);
                const double wr = w[v + (long)V * (det + (long)DET * ed)];
                for (int e = lo; e <= hi; e++) Wt[e] += spectrum[e + (long)E * ed] * wr;
            }
            for (int k = 0; k < s.nvox; k++) {
                xo_scatter_geom sg;
                xo_cm_triple(g, d, &s, maps, beta, v, col, row, k, P, sig2, wvt, Ls, Ld, &sg);
                const long pix = s.vix[k] + (long)NX * s.viy[k];
                for (int n = 0; n < Q; n++) {
                    double f = 0.0;
                    for (int kk = 0; kk < K; kk++) f += coef[pix * K + kk] * basis[(long)kk * Q + n];
                    F[n] = f;
                }
                double *h = H + (size_t)k * Q;
                for (int e = 0; e < E; e++) {
                    double sv = 0.0;
                    const double we = Wt[e] * P[e];
                    for (int n = 0; n < Q; n++) {
                        const double dd = wvt[e] - s.qc[n];
                        const double gq = exp(-dd * dd / (2.0 * sig2[e])) * s.dqc[n];
                        sv += gq * F[n];
                        h[n] += we * gq;
                    }
                    S[e] = sv;
                }
                if (GM) {
                    for (int b = 0; b < B; b++) {
                        double gb = 0.0;
                        for (int e = 0; e < E; e++) gb -= Wt[e] * beta[(long)b * E + e] * P[e] * S[e];
                        gam[b] = gb;
                    }
                    xo_line_scatter(GM, B, NX, NY, d->ray_step, sg.src_x, sg.src_y, s.vix[k] + 0.5,
                                    s.viy[k] + 0.5, gam);
                    xo_line_scatter(GM, B, NX, NY, d->ray_step, sg.det_x, sg.det_y, s.vix[k] + 0.5,
                                    s.viy[k] + 0.5, gam);
                }
            }
        }
        free(P);
    }
    if (grad_coef) {
        memset(grad_coef, 0, sizeof(double) * npix * K);
        for (int k = 0; k < s.nvox; k++) {
            const long pix = s.vix[k] + (long)NX * s.viy[k];
            for (int kk = 0; kk < K; kk++) {
                double acc = 0.0;
                for (int n = 0; n < Q; n++) {
                    double h = 0.0;
                    for (int t = 0; t < nt; t++) h += hpart[((size_t)t * s.nvox + k) * Q + n];
                    acc += basis[(long)kk * Q + n] * h;
                }
                grad_coef[pix * K + kk] = acc;
            }
        }
    }
    if (grad_maps) {
        for (long i = 0; i < (long)B * npix; i++) {
            double acc = 0.0;
            for (int t = 0; t < nt; t++) acc += gpart[(size_t)t * B * npix + i];
            grad_maps[i] = acc;
        }
    }
    free(hpart);
    free(gpart);
    xo_cm_free(&s);
    return XRD_OK;
}

/* ---- CT transmission channel ---------------------------------------------------
 * The transmitted (CT) signal of the simultaneous CT + XRD scanner (README.md:14).
 * The reference only names it, in the commented-out launches of integrate_src2det
 * and polychromatic_src2det (Forward_XRD.cu:271-276); their kernels are not in the
 * repository, so this definition is new.  For view v and detector column c the
 * primary ray runs from the source to the column's point on the detector (the 2-D
 * endpoints of xo_scatter_geometry).  Its part inside [0,nx]x[0,ny] (Liang-Barsky
 * clipping, FP64, no contraction) is sampled like attenuator.cu:44-69 from the
 * source-side entry point:
 *   L_b = step * sum_q bilinear(m_b, entry + u q step),  q*step <= |exit - entry|
 *   T[v + V (c + COL ed)] = sum_{e in win(ed)} RF[e,ed] exp(-sum_b beta[b,e] L_b). */

int xo_ct_chord(double sx, double sy, double dx, double dy, int nx, int ny, double *ex,
                double *ey, double *xx, double *xy) {
    const double ddx = dx - sx, ddy = dy - sy;
    const double p[4] = {-ddx, ddx, -ddy, ddy};
    const double q[4] = {sx - 0.0, (double)nx - sx, sy - 0.0, (double)ny - sy};
    double t0 = 0.0, t1 = 1.0;
    for (int k = 0; k < 4; k++) {
        if (p[k] == 0.0) {
            if (q[k] < 0.0) return 0;
            continue;
        }
        const double r = q[k] / p[k];
        if (p[k] < 0.0) {
            if (r > t0) t0 = r;
        } else {
            if (r < t1) t1 = r;
        }
    }
    if (!(t0 < t1)) return 0;
    *ex = sx + ddx * t0;
    *ey = sy + ddy * t0;
    *xx = sx + ddx * t1;
    *xy = sy + ddy * t1;
    return 1;
}

/* Per-ray pieces: line integrals L[b] and P[e] = exp(-mu_e L). */
static void xo_ct_ray(const xrd_geometry *g, const xrd_discretization *d, int B,
                      const double *maps, const double *beta, int v, int c, double *L, double *P,
                      double *chord, int *hit) {
    const int E = d->n_energy;
    xo_scatter_geom sg;
    xo_scatter_geometry(g, v, c, 0, 0, 0, &sg);
    *hit = xo_ct_chord(sg.src_x, sg.src_y, sg.det_x, sg.det_y, g->nx, g->ny, &chord[0],
                       &chord[1], &chord[2], &chord[3]);
    if (*hit)
        xo_line_integrals_from(maps, B, g->nx, g->ny, d->ray_step, chord[0], chord[1], chord[2],
                               chord[3], L);
    else
        for (int b = 0; b < B; b++) L[b] = 0.0;
    for (int e = 0; e < E; e++) {
        double mu = 0.0;
        for (int b = 0; b < B; b++) mu += beta[(long)b * E + e] * L[b];
        P[e] = exp(-mu);
    }
}

int xo_forward_ct(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum,
                  int B, const double *maps, const double *beta, double *out_t, int nthreads) {
    const int V = g->n_views, COL = g->n_cols, E = d->n_energy, DN = d->energy_halfwidth;
    if (g->nx <= 0 || g->ny <= 0 || V <= 0 || COL <= 0 || B <= 0 || E <= 2 * DN)
        return XRD_ERR_DIMS;
    const int nt = xo_threads(nthreads);
#pragma omp parallel num_threads(nt)
    {
        double *L = (double *)malloc(sizeof(double) * (B + E + 4));
        double *P = L + B, *chord = P + E;
#pragma omp for schedule(dynamic, 4)
        for (long vc = 0; vc < (long)V * COL; vc++) {
            const int v = (int)(vc % V), c = (int)(vc / V);
            int hit;
            xo_ct_ray(g, d, B, maps, beta, v, c, L, P, chord, &hit);
            for (int ed = 0; ed < E; ed++) {
                int lo, hi;
                xo_find_src_nrg_range(ed, E, DN, &lo, &hi);
                double t = 0.0;
                for (int e = lo; e <= hi; e++) t += spectrum[e + (long)E * ed] * P[e];
                out_t[v + (long)V * (c + (long)COL * ed)] = t;
            }
        }
        free(L);
    }
    return XRD_OK;
}

int xo_vjp_ct(const xrd_geometry *g, const xrd_discretization *d, const double *spectrum, int B,
              const double *maps, const double *beta, const double *w, double *grad_maps,
              int nthreads) {
    const int V = g->n_views, COL = g->n_cols, E = d->n_energy, DN = d->energy_halfwidth;
    if (g->nx <= 0 || g->ny <= 0 || V <= 0 || COL <= 0 || B <= 0 || E <= 2 * DN)
        return XRD_ERR_DIMS;
    const long npix = (long)g->nx * g->ny;
    const int nt = xo_threads(nthreads);
    double *gpart = (double *)calloc((size_t)nt * B * npix, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *GM = gpart + (size_t)tid * B * npix;
        double *L = (double *)malloc(sizeof(double) * (2 * B + 2 * E + 4));
        double *P = L + B, *Wt = P + E, *gam = Wt + E, *chord = gam + B;
#pragma omp for schedule(dynamic, 4)
        for (long vc = 0; vc < (long)V * COL; vc++) {
            const int v = (int)(vc % V), c = (int)(vc / V);
            int hit;
            xo_ct_ray(g, d, B, maps, beta, v, c, L, P, chord, &hit);
            if (!hit) continue;
            for (int e = 0; e < E; e++) Wt[e] = 0.0;
            for (int ed = 0; ed < E; ed++) {
                int lo, hi;
                xo_find_src_nrg_range(ed, E, DN, &lo, &hi);
                const double wr = w[v + (long)V * (c + (long)COL * ed)];
                for (int e = lo; e <= hi; e++) Wt[e] += spectrum[e + (long)E * ed] * wr;
            }
            /* dT/dL_b = -sum_e RF P_e beta[b,e] */
            for (int b = 0; b < B; b++) {
                double gb = 0.0;
                for (int e = 0; e < E; e++) gb -= Wt[e] * beta[(long)b * E + e] * P[e];
                gam[b] = gb;
            }
            xo_line_scatter_from(GM, B, g->nx, g->ny, d->ray_step, chord[0], chord[1], chord[2],
                                 chord[3], gam);
        }
        free(L);
    }
    for (long i = 0; i < (long)B * npix; i++) {
        double acc = 0.0;
        for (int t = 0; t < nt; t++) acc += gpart[(size_t)t * B * npix + i];
        grad_maps[i] = acc;
    }
    free(gpart);
    return XRD_OK;
}

// xrd_device.hpp -- device helpers shared by the gfx950 kernels.
//
// Ray traversal (GetIntersection.cu:9-57, attenuator.cu:19-69) is done in
// FP64 with contraction disabled, so its results (entry point, sample count,
// texel indices) are the same bits as any IEEE double restatement that
// performs the same operations in the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xrd_internal.hpp"

namespace xrd {

constexpr int kWave = 64;

// ---- FP64 ray traversal ------------------------------------------------------

// GetIntersection.cu:9-57: first side (y=0, x=nx, y=ny, x=0) crossed by the
// segment (x1,y1)->(x2,y2) with |den| > 1e-5 and parameters in (-eps, 1+eps).
__device__ __forceinline__ int get_intersection(double x1, double y1, double x2, double y2,
                                                int nx, int ny, double &rx, double &ry) {
#pragma clang fp contract(off)
    const double X = (double)nx, Y = (double)ny;
    const double eps = 1e-5;
    const double x2_x1 = x2 - x1;
    const double y2_y1 = y2 - y1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double x3 = (i == 1 || i == 2) ? X : 0.0;
        const double x4 = (i == 0 || i == 1) ? X : 0.0;
        const double y3 = (i >= 2) ? Y : 0.0;
        const double y4 = (i == 1 || i == 2) ? Y : 0.0;
        const double x4_x3 = x4 - x3;
        const double y3_y1 = y3 - y1;
        const double y4_y3 = y4 - y3;
        const double x3_x1 = x3 - x1;
        const double den = x4_x3 * y2_y1 - x2_x1 * y4_y3;
        if (den > eps || den < -eps) {
            const double ua = (x4_x3 * y3_y1 - x3_x1 * y4_y3) / den;
            if (ua > -eps && ua < 1.0 + eps) {
                const double ub = (x2_x1 * y3_y1 - x3_x1 * y2_y1) / den;
                if (ub > -eps && ub < 1.0 + eps) {
                    rx = x1 + x2_x1 * ua;
                    ry = y1 + y2_y1 * ua;
                    return i;
                }
            }
        }
    }
    rx = __longlong_as_double(0x7ff8000000000000ll);
    ry = rx;
    return -1;
}

// One attenuator ray: entry point e, unit direction u, and sample count n
// (samples q = 0..n-1 with q*step <= |P2 - e|; none if there is no entry).
struct Ray {
    double ex, ey, ux, uy;
    int n;
    int side;
};

__device__ __forceinline__ Ray make_ray(double x1, double y1, double x2, double y2, int nx,
                                        int ny, double step) {
#pragma clang fp contract(off)
    Ray r;
    r.side = get_intersection(x1, y1, x2, y2, nx, ny, r.ex, r.ey);
    r.ux = r.uy = 0.0;
    r.n = 0;
    if (r.side < 0) return r;
    const double lx = x2 - r.ex, ly = y2 - r.ey;
    const double L = sqrt(lx * lx + ly * ly);
    r.ux = lx / L;
    r.uy = ly / L;
    if (!(L >= 0.0)) return r;
    int n = (int)floor(L / step) + 1;
    while ((double)n * step <= L) n++;
    while (n > 0 && (double)(n - 1) * step > L) n--;
    r.n = n;
    return r;
}

// Position of sample q (attenuator.cu:49-55), FP64, no contraction.
__device__ __forceinline__ void ray_sample(const Ray &r, double step, int q, double &px,
                                           double &py) {
#pragma clang fp contract(off)
    const double qr = step * (double)q;
    px = r.ex + r.ux * qr;
    py = r.ey + r.uy * qr;
}

// ---- compensated FP32 accumulation --------------------------------------------

struct Kahan {
    float s, c;
    __device__ __forceinline__ Kahan() : s(0.f), c(0.f) {}
    __device__ __forceinline__ void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s - c; }
};

// Exact-weight bilinear sample of an interleaved (a1,a2) map stored with a one
// texel zero border: padded[(i+1) + (nx+2)*(j+1)] holds texel (i,j).  Corner
// indices are clamped into the border independently, which reproduces the
// texture unit's border-zero addressing (Forward_XRD.cu:229-233) for any
// position.  Texel (i,j) is centred at (i+0.5, j+0.5).
__device__ __forceinline__ float2 bilinear2(const float2 *__restrict__ padded, int nx, int ny,
                                            double px, double py) {
    const double xt = px - 0.5, yt = py - 0.5;
    const double fx = floor(xt), fy = floor(yt);
    const float ax = (float)(xt - fx), ay = (float)(yt - fy);
    int i0 = (int)fx, j0 = (int)fy;
    const int pw = nx + 2;
    const int i0c = min(max(i0, -1), nx) + 1, i1c = min(max(i0 + 1, -1), nx) + 1;
    const int j0c = min(max(j0, -1), ny) + 1, j1c = min(max(j0 + 1, -1), ny) + 1;
    const float2 t00 = padded[i0c + pw * j0c];
    const float2 t10 = padded[i1c + pw * j0c];
    const float2 t01 = padded[i0c + pw * j1c];
    const float2 t11 = padded[i1c + pw * j1c];
    const float bx = 1.f - ax, by = 1.f - ay;
    float2 v;
    v.x = by * (bx * t00.x + ax * t10.x) + ay * (bx * t01.x + ax * t11.x);
    v.y = by * (bx * t00.y + ax * t10.y) + ay * (bx * t01.y + ax * t11.y);
    return v;
}

// Trilinear sample of the padded (a1,a2) volume [(nx+2) x (ny+2) x (nz+2)] at
// (px, py, pz): bilinear in the two bracketing slices (each with its zero
// border), blended in z; slices outside [0, nz) are the zero border slices.
__device__ __forceinline__ float2 trilinear2(const float2 *__restrict__ vol, int nx, int ny, int nz,
                                             double px, double py, double pz) {
    const double zt = pz - 0.5;
    const double fz = floor(zt);
    const float az = (float)(zt - fz);
    const int k0 = (int)fz;
    const int k0c = min(max(k0, -1), nz) + 1, k1c = min(max(k0 + 1, -1), nz) + 1;
    const size_t plane = (size_t)(nx + 2) * (ny + 2);
    const float2 v0 = bilinear2(vol + plane * k0c, nx, ny, px, py);
    const float2 v1 = bilinear2(vol + plane * k1c, nx, ny, px, py);
    const float bz = 1.f - az;
    return make_float2(bz * v0.x + az * v1.x, bz * v0.y + az * v1.y);
}

// Height of sample q of a leg traced from (x1,y1,z1) to the voxel (x2,y2,z2): the
// segment's z, linear in the 2-D distance L - q*step left to the voxel (L the
// distance from the ROI entry, H the whole 2-D length).
__device__ __forceinline__ double leg_height(double z1, double z2, double L, double H, double step,
                                             int q) {
#pragma clang fp contract(off)
    const double qr = step * (double)q;
    return z2 + (z1 - z2) * ((L - qr) / H);
}

// ---- wave reductions (wave64) ---------------------------------------------------

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// v + v of the lane that DPP control CTRL selects (a permutation within rows of
// 16 lanes).
template <int CTRL> __device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf,
                                                          false));
}

// Sum over the L lanes (16, 32 or 64) of each row slot; every lane of the slot
// gets the same bits.  Within 16 lanes: swap neighbours (quad_perm 1,0,3,2), swap
// pairs (quad_perm 2,3,0,1), swap quads (row_half_mirror), swap halves
// (row_mirror); each step adds two equal-rank partial sums, so both partners get
// identical results.  Then lanes 16 and 32 apart are added.
template <int L> __device__ __forceinline__ float group_sum(float v) {
    static_assert(L == 16 || L == 32 || L == 64, "row slots are 16, 32 or 64 lanes");
    v = dpp_add<0xB1>(v);
    v = dpp_add<0x4E>(v);
    v = dpp_add<0x141>(v);
    v = dpp_add<0x140>(v);
    if (L >= 32) v += __shfl_xor(v, 16, kWave);
    if (L >= 64) v += __shfl_xor(v, 32, kWave);
    return v;
}

// find_src_nrg_range.cu:8-25 (the window is shifted, never truncated).
__host__ __device__ __forceinline__ void src_window(int nrg, int n_energy, int hw, int &lo,
                                                    int &hi) {
    int mn = nrg - hw;
    if (mn < 0) mn = 0;
    int mx = mn + 2 * hw;
    if (mx >= n_energy) {
        mx = n_energy - 1;
        mn = mx - 2 * hw;
    }
    lo = mn;
    hi = mx;
}

// floor(t) clamped to [0,n-1], NaN -> 0 (integrate_diffraction.cu:187-189, :337-339).
__device__ __forceinline__ int lut_index(double t, int n) {
    const double f = floor(t);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)n) return n - 1;
    return (int)f;
}

// ---- wave transposition ------------------------------------------------------

// v of the lane that DPP control CTRL selects (a permutation within 16 lanes).
template <int CTRL> __device__ __forceinline__ float dpp_get(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sums 16 values per lane over the wave so that lane l ends up with the wave total
// of value (l >> 2) & 15.  Each step pairs lanes that differ in one lane bit; the
// lane with the bit clear keeps the lower half of its values, its partner the upper
// half, and each adds what the other sends: bits 5 and 4 by shuffles, bit 3 by
// row_mirror (lane i <-> 15 - i), bit 2 by row_half_mirror (i <-> 7 - i); bits 1
// and 0 then add the one remaining value (quad_perm swaps).
__device__ __forceinline__ float transpose_sum16(const float *v, int lane) {
    float w[8];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float keep = b5 ? v[8 + k] : v[k], send = b5 ? v[k] : v[8 + k];
        w[k] = keep + __shfl_xor(send, 32, kWave);
    }
    float u[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float keep = b4 ? w[4 + k] : w[k], send = b4 ? w[k] : w[4 + k];
        u[k] = keep + __shfl_xor(send, 16, kWave);
    }
    float z[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float keep = b3 ? u[2 + k] : u[k], send = b3 ? u[k] : u[2 + k];
        z[k] = keep + dpp_get<0x140>(send);
    }
    const float keep = b2 ? z[1] : z[0], send = b2 ? z[0] : z[1];
    float r = keep + dpp_get<0x141>(send);
    r += dpp_get<0x4E>(r);
    r += dpp_get<0xB1>(r);
    return r;
}

// ---- scattering geometry, FP64 (the projector and the coefficient model) ----------

// Per-(view, pixel) scanner geometry (integrate_diffraction.cu:81-128): source and
// pixel positions, detector normal (ndx, ndy, 0) and focal-spot normal (nsx, nsy, nsz).
struct PixelGeom {
    double sp, cp, sx, sy, dx, dy, dz, ndx, ndy, nsx, nsy, nsz;
};

__device__ __forceinline__ PixelGeom pixel_geometry(const ForwardArgs &a, int v, int p) {
#pragma clang fp contract(off)
    const int col = p % a.n_cols, row = p / a.n_cols;
    PixelGeom pg;
    const ViewConst vc = a.views[v];
    pg.sp = vc.sp;
    pg.cp = vc.cp;
    pg.sx = vc.src_x;
    pg.sy = vc.src_y;
    const double rdet = ((double)col - (double)a.n_cols / 2.0 + 0.5) * a.col_pitch;
    pg.dx = a.center_x - a.r_det * vc.sp + rdet * vc.cp;
    pg.dy = a.center_y + a.r_det * vc.cp + rdet * vc.sp;
    pg.dz = a.det_z0 + (double)row * a.row_pitch;
    pg.ndx = vc.sp;
    pg.ndy = -vc.cp;
    pg.nsx = -vc.sp * a.sta;
    pg.nsy = vc.cp * a.sta;
    pg.nsz = -a.cta;
    return pg;
}

// Per-(view, pixel, voxel) scattering geometry (integrate_diffraction.cu:192-255) in
// the FP64 oracle's operation order: cos(theta), sin(theta/2), the two parts of the
// q-resolution variance sigma^2 = As + Bs (E/hbar c)^2, the geometric weight gp and
// gR = gp times the Rayleigh factor (1 + cos^2 theta)/2.
struct TripleGeom {
    double cth, s, As, Bs, gp, gR;
};

__device__ __forceinline__ TripleGeom triple_geometry(const ForwardArgs &a, const PixelGeom &pg,
                                                      int ix, int iy) {
#pragma clang fp contract(off)
    const double vx = (double)ix + 0.5, vy = (double)iy + 0.5;
    const double ax = vx - pg.sx, ay = vy - pg.sy;
    const double proj = (ax * pg.nsx + ay * pg.nsy) / sqrt(pg.nsx * pg.nsx + pg.nsy * pg.nsy);
    const double vz = -proj * a.tan_w2;
    const double hvox = proj * a.tan_w;
    const double az = vz - 0.0;
    const double bx = pg.dx - vx, by = pg.dy - vy, bz = pg.dz - vz;
    const double ab = ax * bx + ay * by + az * bz;
    const double an2 = ax * ax + ay * ay + az * az;
    const double bn2 = bx * bx + by * by + bz * bz;
    const double cth = ab * (1.0 / sqrt(an2)) * (1.0 / sqrt(bn2));
    const double sth = sqrt(fabs(1.0 - cth) / 2.0);
    const double Sx = bx - ab * ax / an2, Sy = by - ab * ay / an2, Sz = bz - ab * az / an2;
    const double Dx = -ax + ab * bx / bn2, Dy = -ay + ab * by / bn2, Dz = -az + ab * bz / bn2;
    const double Vx = -(Sx + Dx), Vy = -(Sy + Dy), Vz = -(Sz + Dz);
    const double Dn = Dx * pg.ndx + Dy * pg.ndy + Dz * 0.0;
    const double Sn = Sx * pg.nsx + Sy * pg.nsy + Sz * pg.nsz;
    const double dD2 = (Dx * Dx + Dy * Dy + Dz * Dz - Dn * Dn) * a.adet / 12.0;
    const double vV2 =
        (a.voxel_size * a.voxel_size * (Vx * Vx + Vy * Vy) + hvox * hvox * Vz * Vz) / 12.0;
    const double sS2 = (Sx * Sx + Sy * Sy + Sz * Sz - Sn * Sn) * a.focal_spot * a.focal_spot / 12.0;
    const double rdotn = -pg.sp * ax + pg.cp * ay;
    const double cospsi = rdotn * (1.0 / sqrt(ax * ax + ay * ay));
    const double g_sv = cospsi * (1.0 / sqrt(an2)) * a.tan_w;
    const double g_vd = -(pg.ndx * bx + pg.ndy * by) * (1.0 / sqrt(bn2 * bn2 * bn2));
    const double res2 = sS2 + vV2 + dD2;
    const double k1 = 2.0 * sth / a.hbarc;
    TripleGeom o;
    o.cth = cth;
    o.s = sth;
    o.As = k1 * k1 * a.e_width * a.e_width / 12.0;
    o.Bs = res2 / (2.0 * sth) / (2.0 * sth) / an2 / bn2;
    o.gp = g_sv * g_vd;
    o.gR = o.gp * ((1.0 + cth * cth) / 2.0);
    return o;
}

}  // namespace xrd

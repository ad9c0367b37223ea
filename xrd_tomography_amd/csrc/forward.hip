// forward.hip -- gfx950 kernels of the XRD forward projector.
//
// Replaces integrate_diffraction (integrate_diffraction.cu:9-409) and its
// helpers GetIntersection, attenuator, find_src_nrg_range, wvt_value and the
// Copy2Vector/texture set-up of Forward_XRD.cu:208-237.  The reference runs
// one 256-thread block per (view, pixel, detector energy) and loops serially
// over all voxels, recomputing both attenuation line integrals for every
// energy block.  Here:
//   * prepare_*      : quantise the Compton table, build a zero-bordered
//                      interleaved (a1,a2) map, and a material-major list of
//                      the non-air voxels in the reference's loop order;
//   * source_legs    : one FP64-traced, Kahan-summed line integral per
//                      (view, voxel) for the source leg (it does not depend on
//                      the pixel or the energy);
//   * forward_kernel : one workgroup per (view, pixel) that produces all NRG
//                      detector-energy rows at once.  Per chunk of voxels it
//                      traces the detector legs (split across the workgroup),
//                      builds a per-(voxel, source energy) table in LDS (FP64
//                      geometry and exponentials), then every lane accumulates
//                      its (q-bin, energy) Gaussian terms in FP32 registers,
//                      flushed per chunk into compensated totals.  The
//                      detector-response window (find_src_nrg_range) is applied
//                      once per material at the end, which removes the
//                      reference's 11x recomputation of every Gaussian term.
#include "xrd_device.hpp"
#include "xrd_internal.hpp"

namespace xrd {

constexpr double kLog2e = 1.4426950408889634074;
constexpr double kRs2pi = 0.39894228040143267794;  // 1/sqrt(2 pi)
constexpr double kBandScale = 9.419280180123797;    // sqrt(128 / log2(e))

// ---------------------------------------------------------------------------
// preparation

__global__ void prepare_maps_kernel(PrepArgs a) {
    const int pw = a.nx + 2, ph = a.ny + 2;
    const int n = pw * ph;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int px = i % pw, py = i / pw;
        float2 v = make_float2(0.f, 0.f);
        if (px >= 1 && px <= a.nx && py >= 1 && py <= a.ny) {
            const int pix = (px - 1) + a.nx * (py - 1);
            v.x = a.a1.at(pix);
            v.y = a.a2.at(pix);
        }
        a.mu_padded[i] = v;
    }
    // integrate_diffraction.cu:66-74: floor, clamp to [0,255], uint8
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.s2_count;
         i += gridDim.x * blockDim.x) {
        const float f = floorf(a.s2_table[i]);
        int v = (f >= 0.f) ? (f > 255.f ? 255 : (int)f) : 0;
        a.s2u8[i] = (uint8_t)v;
    }
}

// Block-wide exclusive scan of 0/1 flags for a 1024-thread block.
__device__ __forceinline__ int block_rank(bool flag, int *wsum, int &total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const unsigned long long m = __ballot(flag);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; i++) {
        const int s = wsum[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + below;
}

// Non-air voxel list.  Pass 1 walks the reference's loop order (nx outer,
// ny inner; integrate_diffraction.cu:173-178) to find, for the stale-voxel
// quirk, the previous non-air voxel of each one.  Passes 2..M+1 emit the
// voxels of each material in that same order.
__global__ __launch_bounds__(1024) void prepare_list_kernel(PrepArgs a) {
    __shared__ int wsum[16];
    const int npix = a.nx * a.ny;
    const bool stale = (a.flags & XRD_QUIRK_STALE_COMPTON_VOXEL) != 0;
    int base = 0;
    for (int c = 0; c < npix; c += blockDim.x) {
        const int ro = c + threadIdx.x;
        bool f = false;
        int pix = 0;
        if (ro < npix) {
            const int ix = ro / a.ny, iy = ro % a.ny;
            pix = ix + a.nx * iy;
            const int s = a.seg[pix];
            f = (s >= 1 && s <= a.n_mat);
        }
        int tot;
        const int r = block_rank(f, wsum, tot);
        if (f) a.ref_list[base + r] = pix;
        base += tot;
    }
    __syncthreads();
    const int nv = base;
    for (int j = threadIdx.x; j < nv; j += blockDim.x) {
        const int pix = a.ref_list[j];
        if (stale) {
            const int prev = j > 0 ? a.ref_list[j - 1] : -1;
            a.cpx1[pix] = prev >= 0 ? a.a1.at(prev) : 0.f;
            a.cpx2[pix] = prev >= 0 ? a.a2.at(prev) : 0.f;
        } else {
            a.cpx1[pix] = a.a1.at(pix);
            a.cpx2[pix] = a.a2.at(pix);
        }
    }
    __syncthreads();
    int off = 0;
    for (int m = 1; m <= a.n_mat; m++) {
        if (threadIdx.x == 0) a.mat_offsets[m - 1] = off;
        for (int c = 0; c < npix; c += blockDim.x) {
            const int ro = c + threadIdx.x;
            bool f = false;
            int pix = 0;
            if (ro < npix) {
                const int ix = ro / a.ny, iy = ro % a.ny;
                pix = ix + a.nx * iy;
                f = (a.seg[pix] == m);
            }
            int tot;
            const int r = block_rank(f, wsum, tot);
            if (f) {
                VoxelEntry v;
                v.pix = pix;
                v.mat = m - 1;
                v.a1c = a.cpx1[pix];
                v.a2c = a.cpx2[pix];
                a.voxels[off + r] = v;
            }
            off += tot;
        }
    }
    if (threadIdx.x == 0) a.mat_offsets[a.n_mat] = off;
}

__global__ void prepare_volume_kernel(MapSrc a1, MapSrc a2, int nx, int ny, int nz,
                                      float2 *mu_vol) {
    const int pw = nx + 2, ph = ny + 2;
    const size_t n = (size_t)pw * ph * (nz + 2);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(i % pw), py = (int)((i / pw) % ph), pz = (int)(i / ((size_t)pw * ph));
        float2 v = make_float2(0.f, 0.f);
        if (px >= 1 && px <= nx && py >= 1 && py <= ny && pz >= 1 && pz <= nz) {
            const size_t vox = (size_t)(px - 1) + (size_t)nx * ((py - 1) + (size_t)ny * (pz - 1));
            v.x = a1.at(vox);
            v.y = a2.at(vox);
        }
        mu_vol[i] = v;
    }
}

hipError_t launch_prepare_volume(const MapSrc &a1, const MapSrc &a2, int nx, int ny, int nz,
                                 float2 *mu_vol, hipStream_t s) {
    hipLaunchKernelGGL(prepare_volume_kernel, dim3(4096), dim3(256), 0, s, a1, a2, nx, ny, nz,
                       mu_vol);
    return hipGetLastError();
}

hipError_t launch_prepare(const PrepArgs &a, hipStream_t s) {
    const int n = (a.nx + 2) * (a.ny + 2);
    int blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(prepare_maps_kernel, dim3(blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(prepare_list_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// source legs: one thread per (view, voxel)

__global__ __launch_bounds__(256) void source_legs_kernel(ForwardArgs a, float2 *src_int) {
    const int npix = a.nx * a.ny;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    const int nv = a.mat_offsets[a.n_mat];
    if (k >= nv) return;
    const VoxelEntry ve = a.voxels[k];
    const int ix = ve.pix % a.nx, iy = ve.pix / a.nx;
    const ViewConst vc = a.views[v];
    const double vx = (double)ix + 0.5, vy = (double)iy + 0.5;
    const Ray r = make_ray(vc.src_x, vc.src_y, vx, vy, a.nx, a.ny, a.ray_step);
    Kahan s1, s2;
    if (a.nz > 1) {
        // 3-D leg: the focal spot at the slice centre, the voxel at its wedge height
        // (integrate_diffraction.cu:199-205)
#pragma clang fp contract(off)
        const double ax = vx - vc.src_x, ay = vy - vc.src_y;
        const double nsx = -vc.sp * a.sta, nsy = vc.cp * a.sta;
        const double proj = (ax * nsx + ay * nsy) / sqrt(nsx * nsx + nsy * nsy);
        const double vz = -proj * a.tan_w2;
        const double lx = vx - r.ex, ly = vy - r.ey;
        const double L = sqrt(lx * lx + ly * ly), H = sqrt(ax * ax + ay * ay);
        for (int q = 0; q < r.n; q++) {
            double px, py;
            ray_sample(r, a.ray_step, q, px, py);
            const double pz = leg_height(a.z_slice, a.z_slice + vz, L, H, a.ray_step, q);
            const float2 t = trilinear2(a.mu_vol, a.nx, a.ny, a.nz, px, py, pz);
            s1.add(t.x);
            s2.add(t.y);
        }
    } else {
        for (int q = 0; q < r.n; q++) {
            double px, py;
            ray_sample(r, a.ray_step, q, px, py);
            const float2 t = bilinear2(a.mu_padded, a.nx, a.ny, px, py);
            s1.add(t.x);
            s2.add(t.y);
        }
    }
    src_int[(size_t)v * npix + k] =
        make_float2((float)((double)s1.value() * a.ray_step), (float)((double)s2.value() * a.ray_step));
}

hipError_t launch_source_integrals(const ForwardArgs &a, float2 *src_int, hipStream_t s) {
    const int npix = a.nx * a.ny;
    dim3 grid((npix + 255) / 256, a.n_views);
    hipLaunchKernelGGL(source_legs_kernel, grid, dim3(256), 0, s, a, src_int);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// forward accumulation

// exp(-x) for finite x >= 0 to FP32 accuracy (about 2e-7 relative): x log2(e) is
// split in FP64 into an integer n and a fraction |f| <= 1/2, 2^f is v_exp_f32 of
// the fraction, and n goes into the exponent.  The results feed FP32 table
// entries, so the FP64 library exponential is not needed.
__device__ __forceinline__ double exp_neg(double x) {
    double y = -x * 1.4426950408889634074;
    y = y < -400.0 ? -400.0 : y;
    const double n = rint(y);
    const float f = (float)(y - n);
    return (double)ldexpf(__builtin_amdgcn_exp2f(f), (int)n);
}

struct VoxShared {
    double cth, s, As, Bs, gR, gp, a2c;
    double lin0, lin1, lout0, lout1;
    int na, pad;
};

// Per-(view, pixel, voxel) scattering geometry of the record (triple_geometry).
__device__ __forceinline__ void voxel_geometry(const ForwardArgs &a, const PixelGeom &pg, int ix,
                                               int iy, VoxShared &o) {
    const TripleGeom tg = triple_geometry(a, pg, ix, iy);
    o.cth = tg.cth;
    o.s = tg.s;
    o.As = tg.As;
    o.Bs = tg.Bs;
    o.gp = tg.gp;
    o.gR = tg.gR;
}

__device__ __forceinline__ double f2_kn(const ForwardArgs &a, double e) {
    return a.kn0 + a.kn1 * e + a.kn2 * e * e;
}

// One thread per (view of the batch, pixel, voxel): the detector leg's line
// integrals (attenuator.cu:9-109; FP64 sample positions, Kahan FP32 sums) and the
// FP64 scattering geometry (integrate_diffraction.cu:192-255), written as one
// record for forward_kernel.  Consecutive threads take consecutive voxels of the
// material-major list, so their rays run side by side through the map.
template <bool VOLUME>
__global__ __launch_bounds__(256) void pixel_voxel_kernel(ForwardArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nvox) return;
    const int p = blockIdx.y, vb = blockIdx.z;
    const PixelGeom pg = pixel_geometry(a, a.v0 + vb, p);
    const VoxelEntry ve = a.voxels[k];
    const int ix = ve.pix % a.nx, iy = ve.pix / a.nx;
    const double vx = (double)ix + 0.5, vy = (double)iy + 0.5;
    const Ray ray = make_ray(pg.dx, pg.dy, vx, vy, a.nx, a.ny, a.ray_step);
    Kahan s1, s2;
    if (VOLUME) {
        // 3-D leg: from the pixel (slice centre + row height) down or up to the voxel
        // at its wedge height
#pragma clang fp contract(off)
        const double ax = vx - pg.sx, ay = vy - pg.sy;
        const double proj = (ax * pg.nsx + ay * pg.nsy) / sqrt(pg.nsx * pg.nsx + pg.nsy * pg.nsy);
        const double vz = -proj * a.tan_w2;
        const double bx = pg.dx - vx, by = pg.dy - vy;
        const double lx = vx - ray.ex, ly = vy - ray.ey;
        const double L = sqrt(lx * lx + ly * ly), H = sqrt(bx * bx + by * by);
        for (int q = 0; q < ray.n; q++) {
            double px, py;
            ray_sample(ray, a.ray_step, q, px, py);
            const double pz = leg_height(a.z_slice + pg.dz, a.z_slice + vz, L, H, a.ray_step, q);
            const float2 tv = trilinear2(a.mu_vol, a.nx, a.ny, a.nz, px, py, pz);
            s1.add(tv.x);
            s2.add(tv.y);
        }
    } else {
        for (int q = 0; q < ray.n; q++) {
            double px, py;
            ray_sample(ray, a.ray_step, q, px, py);
            const float2 tv = bilinear2(a.mu_padded, a.nx, a.ny, px, py);
            s1.add(tv.x);
            s2.add(tv.y);
        }
    }
    VoxShared vs;
    voxel_geometry(a, pg, ix, iy, vs);
    const size_t ri = ((size_t)vb * a.n_det + p) * a.nvox + k;
    a.pv_s[ri] = vs.s;
    a.pv_cth[ri] = vs.cth;
    a.pv_geo[ri] = make_float4((float)vs.As, (float)vs.Bs, (float)vs.gR, (float)vs.gp);
    a.pv_leg[ri] = make_float2((float)((double)s1.value() * a.ray_step),
                               (float)((double)s2.value() * a.ray_step));
}

hipError_t launch_pixel_voxel(const ForwardArgs &a, hipStream_t s) {
    if (a.nvox <= 0) return hipSuccess;
    dim3 grid((a.nvox + 255) / 256, a.n_det, a.nvb);
    if (a.nz > 1)
        hipLaunchKernelGGL(pixel_voxel_kernel<true>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(pixel_voxel_kernel<false>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// Block of G*TQ lanes: lane (tq, h) owns coarse bin tq, fine bins tq*FS..tq*FS+FS-1
// and the source energies e = h + G*i, i < EPT.
// Four waves per SIMD: two 512-lane workgroups share a CU, so one can trace,
// build its tables or wait at a barrier while the other evaluates Gaussians.
template <int TQ, int G, int EPT, int FS, bool WITH_A>
__global__ __launch_bounds__(G * TQ) __attribute__((amdgpu_waves_per_eu(4))) void forward_kernel(
    ForwardArgs a, int kc) {
    constexpr int NT = G * TQ;
    extern __shared__ __align__(16) unsigned char smem[];
    const int t = threadIdx.x;
    const int tq = t % TQ, h = t / TQ;
    const int V = a.n_views, E = a.n_energy, Q = a.n_q, QF = a.n_qf, M = a.n_mat;
    const int vb = blockIdx.x % a.nvb, p = blockIdx.x / a.nvb;  // view within the batch, pixel
    const int v = a.v0 + vb;
    const int npix = a.nx * a.ny;
    const int Gc = NT / E;  // Compton lane groups

    // ---- LDS carve-up
    VoxShared *vox = reinterpret_cast<VoxShared *>(smem);
    unsigned char *r2 = smem + ((kc * sizeof(VoxShared) + 15) & ~size_t(15));
    double2 *teo = reinterpret_cast<double2 *>(r2);                  // [kc][E] Compton E_out bin
    float4 *tw = reinterpret_cast<float4 *>(teo + kc * E);           // [kc][E]
    float *tb = reinterpret_cast<float *>(tw + kc * E);              // [kc][E] band
    float *tk = tb + kc * E;                                         // [kc][E] Compton K / width
    const int nwin = 2 * a.dn + 1;
    float *rfc = tk + kc * E;     // [E][nwin] spectrum window of each detector bin
    float *hst = rfc + E * nwin;  // [E][Q] per-material coarse totals (A only)
    double *red = reinterpret_cast<double *>(r2);                    // final reductions (aliases)

    // ---- per-lane q constants
    float qch = 0.f, qcl = 0.f, dqc = 0.f;
    if (WITH_A && tq < Q) {
        const double q = a.q_coarse[tq], dq = a.q_coarse[Q + tq];
        qch = (float)q;
        qcl = (float)(q - (double)qch);
        dqc = (float)dq;
    } else {
        qch = 1e30f;  // never in band; dqc = 0 keeps the term at zero
    }
    // fine centres q_min + dq_fine*(0.5 + n) (integrate_diffraction.cu:302);
    // q_coarse holds [Q centres][Q widths][q_min].  The lane's FS centres are evenly
    // spaced, so the walk forms one hi/lo difference, at the middle centre (qmh, qml),
    // and reaches centre s by subtracting dfs[s] = (s - FS/2) dq_fine: the extra
    // rounding is 2^-24 of |d| + dq_fine.  Centres past the axis (n >= QF) keep a
    // finite difference and weigh zero (wf).
    constexpr int SM = FS / 2;
    float qmh, qml, dfs[FS];
    float wlo_c = 1e30f, whi_c = -1e30f, wlo_f = 1e30f, whi_f = -1e30f;
    {
        const double qmin_axis = a.q_coarse[2 * Q];
        const double qm = qmin_axis + a.dq_fine * (0.5 + (double)(tq * FS + SM));
        qmh = (float)qm;
        qml = (float)(qm - (double)qmh);
#pragma unroll
        for (int s = 0; s < FS; s++) dfs[s] = (float)((double)(s - SM) * a.dq_fine);
        float lo = 1e30f, hi = -1e30f;
#pragma unroll
        for (int s = 0; s < FS; s++) {
            const int n = tq * FS + s;
            if (n < QF) {
                const float qf = (float)(qmin_axis + a.dq_fine * (0.5 + (double)n));
                lo = fminf(lo, qf);
                hi = fmaxf(hi, qf);
            }
        }
        float clo = (WITH_A && tq < Q) ? qch : 1e30f, chi = (WITH_A && tq < Q) ? qch : -1e30f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, o, kWave));
            hi = fmaxf(hi, __shfl_xor(hi, o, kWave));
            clo = fminf(clo, __shfl_xor(clo, o, kWave));
            chi = fmaxf(chi, __shfl_xor(chi, o, kWave));
        }
        wlo_f = lo;
        whi_f = hi;
        wlo_c = clo;
        whi_c = chi;
    }

    // ---- accumulators
    // per-chunk partials (at most kc terms each); the per-material coarse totals in
    // LDS (hst[e][tq], FP32, one add per chunk); the fine totals per wave and
    // energy slot in LDS (ftw[wave][i], FP64, one add per chunk of the chunk's
    // wave total)
    static_assert(EPT <= 16, "the fine totals use the 16-value transposed wave sum");
    __shared__ double ftw[NT / kWave][16];
    float pc[EPT], pf[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        pc[i] = 0.f;
        pf[i] = 0.f;
    }
    if (t < NT / kWave * 16) (&ftw[0][0])[t] = 0.0;
    if (WITH_A)
        for (int idx = t; idx < E * Q; idx += NT) hst[idx] = 0.f;
    for (int idx = t; idx < E * nwin; idx += NT) {
        const int ed = idx / nwin, k = idx % nwin;
        int wlo, whi;
        src_window(ed, E, a.dn, wlo, whi);
        rfc[idx] = wlo + k <= whi ? a.spectrum[(wlo + k) + (size_t)E * ed] : 0.f;
    }
    double cacc = 0.0;
    const int ced = t % E, cgrp = t / E;
    const bool compton_lane = t < Gc * E;

    const bool swap = (a.flags & XRD_QUIRK_SWAPPED_LEGS) != 0;

    for (int m = 0; m < M; m++) {
        const int lo = a.mat_offsets[m], hi = a.mat_offsets[m + 1];
        float wf[FS];
#pragma unroll
        for (int s = 0; s < FS; s++) {
            const int n = tq * FS + s;
            float iv = 0.f;
            if (n < QF)
                iv = a.ifine ? a.ifine[n + (size_t)QF * m]
                             : bf16_bits_to_float(a.ifine_b16[n + (size_t)QF * m]) * a.ifine_scale[m];
            wf[s] = (n < QF) ? (float)(a.dq_fine * (double)iv) : 0.f;
        }
        for (int k0 = lo; k0 < hi; k0 += kc) {
            const int nk = min(kc, hi - k0);
            // ---- the chunk's (view, pixel, voxel) records (pixel_voxel_kernel)
            if (t < nk) {
                const int k = k0 + t;
                const size_t ri = ((size_t)vb * a.n_det + p) * a.nvox + k;
                const VoxelEntry ve = a.voxels[k];
                const float4 gg = a.pv_geo[ri];
                const float2 ld = a.pv_leg[ri];
                const float2 ls = a.src_int[(size_t)v * npix + k];
                VoxShared vs;
                vs.s = a.pv_s[ri];
                vs.cth = a.pv_cth[ri];
                vs.As = gg.x;
                vs.Bs = gg.y;
                vs.gR = gg.z;
                vs.gp = gg.w;
                vs.a2c = (double)ve.a2c;
                vs.na = lut_index((log2(fabs((double)ve.a1c / (double)ve.a2c)) - a.lut_a_min) /
                                      a.da_lut,
                                  a.lut_na);
                // integrate_diffraction.cu:259-260 passes the detector leg where the
                // source leg belongs; reproduced only under XRD_QUIRK_SWAPPED_LEGS.
                if (swap) {
                    vs.lin0 = ld.x; vs.lin1 = ld.y; vs.lout0 = ls.x; vs.lout1 = ls.y;
                } else {
                    vs.lin0 = ls.x; vs.lin1 = ls.y; vs.lout0 = ld.x; vs.lout1 = ld.y;
                }
                vox[t] = vs;
            }
            __syncthreads();
            // ---- phase 2: per-(voxel, source energy) table, FP64
            for (int idx = t; idx < nk * E; idx += NT) {
                const int j = idx / E, e = idx % E;
                const VoxShared &vs = vox[j];
                const EnergyConst ec = a.energies[e];
                const double arg = (vs.lin0 + vs.lout0) * ec.inv_f1 + (vs.lin1 + vs.lout1) * ec.inv_f2;
                const double pinpout = exp_neg(arg);
                const double sig2 = vs.As + vs.Bs * ec.k2sq;
                const double w = ec.wcoef * vs.s;
                const float wh = (float)w;
                const float wl = (float)(w - (double)wh);
                // from one 1/sigma: c2 = -log2(e)/(2 sigma^2), the prefactor's
                // 1/(sqrt(2 pi) sigma), and the band half-width sqrt(64/-c2)
                const double isg = 1.0 / sqrt(sig2);
                const double c2 = (-0.5 * kLog2e) * (isg * isg);
                const double P = pinpout * vs.gR * kRs2pi * isg;
                tw[j * E + e] = make_float4(wh, wl, (float)c2, (float)P);
                tb[j * E + e] = (float)(kBandScale * sig2 * isg);
                // Compton pieces (integrate_diffraction.cu:318-356) without RF and
                // channel_match; one reciprocal of er serves every division by it
                const double ein = ec.ein;
                const double er = 1.0 + ein * (1.0 - vs.cth);
                const double ier = 1.0 / er;
                const double eout = ein * ier;
                const double wc = ein * vs.s;
                const double l2w = log2(wc * sqrt(er + wc * wc) * ier);
                const int nq = lut_index((l2w - a.lut_q_min) / a.dq_lut, a.lut_nq);
                const double g1 = eout * eout * eout, g2 = f2_kn(a, eout);
                const double pin = exp_neg(vs.lin0 * ec.inv_f1 + vs.lin1 * ec.inv_f2);
                const double pc_ = exp_neg(vs.lout0 / g1 + vs.lout1 / g2);
                const double cf = er * er * (er + ier - 1.0 + vs.cth * vs.cth) / 2.0;
                const double K = pin * pc_ * vs.gp * cf * vs.a2c *
                                 (double)a.s2u8[nq + a.lut_nq * vs.na];
                // Compton-shifted source bin [es1, es2] / er (channel_match.cu): the
                // detector lanes add (overlap with their bin) * K / (bin width)
                const double eo1 = ec.es1 * ier, eo2 = ec.es2 * ier;
                teo[j * E + e] = make_double2(eo1, eo2);
                tk[j * E + e] = (float)(K / (eo2 - eo1));
            }
            __syncthreads();
            // ---- Compton accumulation over the detector-energy rows
            if (compton_lane) {
                int wlo, whi;
                src_window(ced, E, a.dn, wlo, whi);
                const double Ed1 = a.e_min + (double)ced * a.e_width;
                const double Ed2 = a.e_min + (double)(ced + 1) * a.e_width;
                for (int j = cgrp; j < nk; j += Gc) {
                    // The shifted source bins [es1, es2] / er(e) rise strictly with e (the
                    // edges by about a bin width, far above FP64 rounding), so the bins that
                    // overlap [Ed1, Ed2] are those from the first with eo2 > Ed1 up to the
                    // last with eo1 < Ed2.  Both ends are found by a four-step search
                    // (windows hold at most 2 kMaxEnergyHalfwidth + 1 = 15 bins), the same
                    // number of steps on every lane, and the terms are added in the same
                    // ascending order, with the same arithmetic, as the full window.
                    const double2 *te = teo + j * E;
                    int e_lo = wlo, e_end = wlo;
#pragma unroll
                    for (int step = 8; step > 0; step >>= 1) {
                        const int c = e_lo + step - 1;
                        if (c <= whi && !(te[c].y > Ed1)) e_lo += step;
                        const int d = e_end + step - 1;
                        if (d <= whi && te[d].x < Ed2) e_end += step;
                    }
                    for (int e = e_lo; e < e_end; e++) {
                        const double2 eo = te[e];
                        const double lo_e = eo.x > Ed1 ? eo.x : Ed1;
                        const double hi_e = eo.y < Ed2 ? eo.y : Ed2;
                        const double cm = hi_e - lo_e;
                        if (!(cm > 0.0)) continue;
                        cacc += cm * (double)rfc[ced * nwin + (e - wlo)] * (double)tk[j * E + e];
                    }
                }
            }
            // ---- phase 3: Gaussian resolution terms (integrate_diffraction.cu:298-315)
            // For each source energy of the wave, one ballot gives the chunk's voxels
            // whose band reaches the wave's coarse and fine bins; the wave then walks
            // the set bits only (scalar bit scans), adding voxels in ascending order.
            {
                const int lane = t & (kWave - 1);
#pragma unroll
                for (int i = 0; i < EPT; i++) {
                    const int e = h + G * i;
                    if (e >= E) continue;  // wave-uniform
                    bool in_c = false, in_f = false;
                    if (lane < nk) {
                        const float x = tw[lane * E + e].x;
                        const float band = tb[lane * E + e];
                        in_c = WITH_A && x + band >= wlo_c && x - band <= whi_c;
                        in_f = x + band >= wlo_f && x - band <= whi_f;
                    }
                    uint64_t mc = __ballot(in_c);
                    uint64_t mf = __ballot(in_f);
                    // two voxels per step while two are left, both table reads issued
                    // before either is waited on; the sums keep ascending voxel order
                    auto coarse = [&](const float4 &tv) {
                        const float d = (tv.x - qch) + (tv.y - qcl);
                        return __builtin_amdgcn_exp2f(d * d * tv.z);
                    };
                    auto fine = [&](const float4 &tv) {
                        const float dm = (tv.x - qmh) + (tv.y - qml);
                        float acc = 0.f;
#pragma unroll
                        for (int s = 0; s < FS; s++) {
                            const float d = s == SM ? dm : dm - dfs[s];
                            acc = fmaf(wf[s], __builtin_amdgcn_exp2f(d * d * tv.z), acc);
                        }
                        return acc;
                    };
                    while (mc & (mc - 1)) {
                        const int j0 = __ffsll((unsigned long long)mc) - 1;
                        mc &= mc - 1;
                        const int j1 = __ffsll((unsigned long long)mc) - 1;
                        mc &= mc - 1;
                        const float4 t0 = tw[j0 * E + e], t1 = tw[j1 * E + e];
                        const float g0 = coarse(t0), g1 = coarse(t1);
                        pc[i] = fmaf(t0.w, g0, pc[i]);  // times dqc at the flush
                        pc[i] = fmaf(t1.w, g1, pc[i]);
                    }
                    if (mc) {
                        const float4 tv = tw[(__ffsll((unsigned long long)mc) - 1) * E + e];
                        pc[i] = fmaf(tv.w, coarse(tv), pc[i]);
                    }
                    while (mf & (mf - 1)) {
                        const int j0 = __ffsll((unsigned long long)mf) - 1;
                        mf &= mf - 1;
                        const int j1 = __ffsll((unsigned long long)mf) - 1;
                        mf &= mf - 1;
                        const float4 t0 = tw[j0 * E + e], t1 = tw[j1 * E + e];
                        const float a0 = fine(t0), a1 = fine(t1);
                        pf[i] = fmaf(t0.w, a0, pf[i]);
                        pf[i] = fmaf(t1.w, a1, pf[i]);
                    }
                    if (mf) {
                        const float4 tv = tw[(__ffsll((unsigned long long)mf) - 1) * E + e];
                        pf[i] = fmaf(tv.w, fine(tv), pf[i]);
                    }
                }
            }
            {
                const int lane = t & (kWave - 1);
                float pf16[16];
#pragma unroll
                for (int i = 0; i < 16; i++) pf16[i] = i < EPT ? pf[i] : 0.f;
                const float sf = transpose_sum16(pf16, lane);
                if ((lane & 3) == 0) ftw[t >> 6][(lane >> 2) & 15] += (double)sf;
            }
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = h + G * i;
                if (WITH_A && e < E && tq < Q) hst[e * Q + tq] += pc[i] * dqc;
                pc[i] = 0.f;
                pf[i] = 0.f;
            }
            __syncthreads();
        }
        if (WITH_A) {
            // ---- material flush: apply the detector-response window
            // A[nwt, mat, e_det] = sum_{e in window(e_det)} spectrum[e, e_det] * H[e, nwt]
            for (int ed = h; ed < E; ed += G) {
                if (tq < Q) {
                    int wlo, whi;
                    src_window(ed, E, a.dn, wlo, whi);
                    double acc = 0.0;
                    for (int e = wlo; e <= whi; e++)
                        acc += (double)a.spectrum[e + (size_t)E * ed] * (double)hst[e * Q + tq];
                    const size_t rowi = (size_t)v + (size_t)V * ((size_t)p + (size_t)a.n_det * ed);
                    a.out_a[(size_t)tq + (size_t)Q * ((size_t)m + (size_t)M * rowi)] = (float)acc;
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = h + G * i;
                if (e < E && tq < Q) hst[e * Q + tq] = 0.f;
            }
        }
    }

    // ---- counts: add the fine totals of the waves of each energy group
    const int nw = NT / kWave;  // waves; waves [h*TQ/64, (h+1)*TQ/64) hold group h
    __syncthreads();
    double *hf = red + nw * EPT;  // [E]
    double *cr = hf + E;          // [Gc][E]
    if (t < E) {
        const int hh = t % G, i = t / G;
        double sacc = 0.0;
        for (int w = hh * (TQ / kWave); w < (hh + 1) * (TQ / kWave); w++) sacc += ftw[w][i];
        hf[t] = sacc;
    }
    if (compton_lane) cr[cgrp * E + ced] = cacc;
    __syncthreads();
    if (t < E) {
        int wlo, whi;
        src_window(t, E, a.dn, wlo, whi);
        double nacc = 0.0;
        for (int e = wlo; e <= whi; e++) nacc += (double)a.spectrum[e + (size_t)E * t] * hf[e];
        double cs = 0.0;
        for (int g = 0; g < Gc; g++) cs += cr[g * E + t];
        const size_t rowi = (size_t)v + (size_t)V * ((size_t)p + (size_t)a.n_det * t);
        a.out_n[rowi] = (float)nacc;
        a.out_c[rowi] = (float)cs;
    }
}

template <int TQ, int G, int EPT, int FS, bool WITH_A>
static hipError_t launch_fwd_t(const ForwardArgs &a, hipStream_t s) {
    constexpr int NT = G * TQ;
    const int kc = a.kc > 0 ? a.kc : 16;
    const int E = a.n_energy;
    const int Gc = NT / E;
    const size_t vox_bytes = ((kc * sizeof(VoxShared) + 15) & ~size_t(15));
    // teo (16) | tw (16) | tb (4) | tk (4) per (voxel, energy) | rfc | hst
    size_t r2 = (size_t)kc * E * (16 + 16 + 4 + 4) +
                (size_t)E * (2 * a.dn + 1) * 4 + (WITH_A ? (size_t)E * a.n_q * 4 : 0);
    const size_t redb = ((size_t)(NT / kWave) * EPT + E + (size_t)Gc * E) * 8;
    r2 = r2 > redb ? r2 : redb;
    const size_t shmem = vox_bytes + r2;
    auto kern = forward_kernel<TQ, G, EPT, FS, WITH_A>;
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
    const unsigned nblk = (unsigned)a.nvb * (unsigned)a.n_det;
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(NT), shmem, s, a, kc);
    return hipGetLastError();
}

template <bool WITH_A>
static hipError_t dispatch_fwd(const ForwardArgs &a, hipStream_t s) {
    const int E = a.n_energy, Q = a.n_q, QF = a.n_qf;
    if (Q <= 128 && QF <= 384) {
        if (E <= 16) return launch_fwd_t<128, 4, 4, 3, WITH_A>(a, s);
        if (E <= 64) return launch_fwd_t<128, 4, 16, 3, WITH_A>(a, s);
    }
    if (Q <= 256 && QF <= 768 && E <= 64) return launch_fwd_t<256, 4, 16, 3, WITH_A>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_forward(const ForwardArgs &a, hipStream_t s) {
    return a.out_a ? dispatch_fwd<true>(a, s) : dispatch_fwd<false>(a, s);
}

// ---------------------------------------------------------------------------
// ray traversal on its own (for the intersection-list contract)

__device__ __forceinline__ float texel_planar(const float *map, int nx, int ny, int i, int j) {
    return (i < 0 || j < 0 || i >= nx || j >= ny) ? 0.f : map[i + nx * j];
}

__global__ void trace_rays_kernel(int nx, int ny, double step, int64_t n, const double *seg,
                                  int32_t *side, double *entry, int32_t *count,
                                  int32_t max_samples, int32_t *texels, int32_t n_basis,
                                  const float *maps, float *ints) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Ray r = make_ray(seg[4 * i], seg[4 * i + 1], seg[4 * i + 2], seg[4 * i + 3], nx, ny,
                           step);
    side[i] = r.side;
    entry[2 * i] = r.side >= 0 ? r.ex : __longlong_as_double(0x7ff8000000000000ll);
    entry[2 * i + 1] = r.side >= 0 ? r.ey : __longlong_as_double(0x7ff8000000000000ll);
    count[i] = r.n;
    if (texels) {
        for (int q = 0; q < max_samples; q++) {
            int32_t tx = INT32_MIN, ty = INT32_MIN;
            if (q < r.n) {
                double px, py;
                ray_sample(r, step, q, px, py);
                tx = (int32_t)floor(px - 0.5);
                ty = (int32_t)floor(py - 0.5);
            }
            texels[2 * (i * max_samples + q)] = tx;
            texels[2 * (i * max_samples + q) + 1] = ty;
        }
    }
    if (ints) {
        const size_t plane = (size_t)nx * ny;
        for (int b = 0; b < n_basis; b++) {
            const float *mp = maps + b * plane;
            Kahan acc;
            for (int q = 0; q < r.n; q++) {
                double px, py;
                ray_sample(r, step, q, px, py);
                const double xt = px - 0.5, yt = py - 0.5;
                const double fx = floor(xt), fy = floor(yt);
                const float ax = (float)(xt - fx), ay = (float)(yt - fy);
                const int i0 = (int)fx, j0 = (int)fy;
                const float bx = 1.f - ax, by = 1.f - ay;
                const float val = by * (bx * texel_planar(mp, nx, ny, i0, j0) +
                                        ax * texel_planar(mp, nx, ny, i0 + 1, j0)) +
                                  ay * (bx * texel_planar(mp, nx, ny, i0, j0 + 1) +
                                        ax * texel_planar(mp, nx, ny, i0 + 1, j0 + 1));
                acc.add(val);
            }
            ints[(size_t)b * n + i] = (float)((double)acc.value() * step);
        }
    }
}

hipError_t launch_trace_rays(int nx, int ny, double step, int64_t n, const double *seg,
                             int32_t *side, double *entry, int32_t *count, int32_t max_samples,
                             int32_t *texels, int32_t n_basis, const float *maps, float *ints,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 127) / 128);
    hipLaunchKernelGGL(trace_rays_kernel, dim3(blocks), dim3(128), 0, s, nx, ny, step, n, seg,
                       side, entry, count, max_samples, texels, n_basis, maps, ints);
    return hipGetLastError();
}

}  // namespace xrd

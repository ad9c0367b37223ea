// coeff.hip -- differentiable coefficient model: forward counts and the
// vector-Jacobian product with respect to the per-voxel form-factor
// coefficients and the attenuation basis maps.
//
// Model (the differentiable form of integrate_diffraction.cu:266-315; Compton
// is a constant background here):
//   yhat[v,p,ed] = sum_vox sum_{e in win(ed)} RF[e,ed] P(e) S(e)
//   P(e) = exp(-sum_b beta[b,e] (Ls_b + Ld_b)) g R / sqrt(2 pi sigma_e^2)
//   S(e) = sum_n exp(-(wvt_e - q_n)^2 / (2 sigma_e^2)) dq_n F_vox(q_n),
//   F_vox(q_n) = sum_k coef[vox,k] basis[k,n].
// The reference is B = 2 with beta = (1/e^3, 1/f2(e)) (integrate_diffraction.cu:291-296);
// a full per-energy attenuation map is B = NRG with beta = identity.
//
// The VJP recomputes every line integral (no per-ray storage).  Gradients with
// respect to the coefficients are accumulated per voxel in registers
// (deterministic); gradients with respect to the maps are scattered along the
// rays with FP64 atomics (order-dependent only at the 1e-16 level).
#include <algorithm>
#include <climits>
#include <cstdlib>

#include "xrd_device.hpp"
#include "xrd_internal.hpp"

namespace xrd {

namespace {

constexpr double kLog2e = 1.4426950408889634074;
constexpr double kRs2pi = 0.39894228040143267794;
constexpr int TQ = 128, G = 4, NT = TQ * G, EPT = 16, KC = 16;
static_assert(TQ == kCoeffFqCols, "fq_buf rows hold the kernels' q lanes");
constexpr int MAXB = 64, MAXK = 32;
// Row strides in LDS of the per-voxel rows that the matrix-core contractions read with
// one voxel per lane (lane % 16): FP32 leg integrals [KC][LROW] (68 floats put the 16
// voxels' rows in different banks; at 64 every lane of a column hit one bank) and the
// FP64 per-energy rows mu / W P S [KC][XROW].
constexpr int LROW = MAXB + 4, XROW = 68;

struct VoxC {
    double s, As, Bs, gR;
    double sx, sy;   // unused by the forward, kept for the VJP source ray
    int ix, iy, pix, k;
};

using PixG = PixelGeom;

__device__ __forceinline__ PixG pixel_setup(const CoeffArgs &a, int v, int p) {
    return pixel_geometry(a.f, v, p);
}

// integrate_diffraction.cu:192-255 in FP64 (triple_geometry)
__device__ __forceinline__ void geom_c(const ForwardArgs &a, const PixG &pg, int ix, int iy,
                                       VoxC &o) {
    const TripleGeom tg = triple_geometry(a, pg, ix, iy);
    o.s = tg.s;
    o.As = tg.As;
    o.Bs = tg.Bs;
    o.gR = tg.gR;
}

// Bilinear sample of padded planar map b at (px,py): zero border by index clamping.
__device__ __forceinline__ float bil_planar(const float *__restrict__ base, unsigned plane_off,
                                            int nx, int ny, double px, double py) {
    const double xt = px - 0.5, yt = py - 0.5;
    const double fx = floor(xt), fy = floor(yt);
    const float ax = (float)(xt - fx), ay = (float)(yt - fy);
    const int i0 = (int)fx, j0 = (int)fy;
    const int pw = nx + 2;
    const int i0c = min(max(i0, -1), nx) + 1, i1c = min(max(i0 + 1, -1), nx) + 1;
    const int j0c = min(max(j0, -1), ny) + 1, j1c = min(max(j0 + 1, -1), ny) + 1;
    const float bx = 1.f - ax, by = 1.f - ay;
    const float *__restrict__ pm = base;
    return by * (bx * pm[plane_off + (unsigned)(i0c + pw * j0c)] +
                 ax * pm[plane_off + (unsigned)(i1c + pw * j0c)]) +
           ay * (bx * pm[plane_off + (unsigned)(i0c + pw * j1c)] +
                 ax * pm[plane_off + (unsigned)(i1c + pw * j1c)]);
}

// Phase 1 for a chunk: lane t = j + KC*(bc + nbc*r) marches samples q = r (mod sp) of
// the detector ray of voxel j for maps 2bc, 2bc+1; partial sums combined in LDS.
__device__ __forceinline__ void trace_chunk(const CoeffArgs &a, const PixG &pg, const VoxC *vox,
                                            int nk, float2 *part, float *Ld) {
    const int B = a.n_basis, nbc = (B + 1) / 2, sp = NT / (KC * nbc);
    const int t = threadIdx.x;
    const int j = t % KC, bc = (t / KC) % nbc, r = t / (KC * nbc);
    const size_t plane = (size_t)(a.f.nx + 2) * (a.f.ny + 2);
    if (j < nk && r < sp) {
        const Ray ray = make_ray(pg.dx, pg.dy, (double)vox[j].ix + 0.5, (double)vox[j].iy + 0.5,
                                 a.f.nx, a.f.ny, a.f.ray_step);
        const int b0 = 2 * bc, b1 = 2 * bc + 1;
        const unsigned o0 = (unsigned)(b0 * plane), o1 = (unsigned)((b1 < B ? b1 : b0) * plane);
        Kahan k0, k1;
        // four samples' loads in flight before any is used; samples past the end
        // repeat the lane's last one and are not added
        constexpr int U = 4;
        const int nq = ray.n > r ? (ray.n - r + sp - 1) / sp : 0;
        for (int i0 = 0; i0 < nq; i0 += U) {
            float v0[U], v1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                double px, py;
                ray_sample(ray, a.f.ray_step, r + sp * min(i0 + u, nq - 1), px, py);
                v0[u] = bil_planar(a.maps_padded, o0, a.f.nx, a.f.ny, px, py);
                v1[u] = bil_planar(a.maps_padded, o1, a.f.nx, a.f.ny, px, py);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (i0 + u >= nq) break;
                k0.add(v0[u]);
                if (b1 < B) k1.add(v1[u]);
            }
        }
        part[t] = make_float2(k0.value(), k1.value());
    }
    __syncthreads();
    for (int idx = t; idx < nk * B; idx += NT) {
        const int jj = idx % nk, b = idx / nk;
        const int bcc = b >> 1, comp = b & 1;
        double s = 0.0;
        for (int rr = 0; rr < sp; rr++) {
            const float2 pv = part[jj + KC * (bcc + nbc * rr)];
            s += (double)(comp ? pv.y : pv.x);
        }
        Ld[jj * LROW + b] = (float)(s * a.f.ray_step);
    }
}

// Texel-major layout for many maps: maps_tm[(i+1) + (nx+2)(j+1)][b] holds map b at
// texel (i, j), 64 floats per texel (zero for b >= B and on the border).  One wave
// per ray, one lane per map: each texel a ray touches is one coalesced 256-byte load,
// where the planar layout would touch one cache line per map.
constexpr int kTexelMaps = 64;

__device__ __forceinline__ double read_lane(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sample q0 + lane of a ray (the last sample for lanes past the end): the cell
// (i0, j0) of its bilinear stencil and the fractions (ax, ay).
__device__ __forceinline__ void lane_sample(const Ray &ray, double step, int q, int &i0, int &j0,
                                            double &ax, double &ay) {
    double px, py;
    ray_sample(ray, step, q, px, py);
    const double xt = px - 0.5, yt = py - 0.5;
    const double fx = floor(xt), fy = floor(yt);
    ax = xt - fx;
    ay = yt - fy;
    i0 = (int)fx;
    j0 = (int)fy;
}

// Workgroups are dispatched to the 8 XCDs round-robin by id, and each XCD has its
// own L2.  xcd_contiguous renumbers block id b of nb so that XCD x takes a contiguous
// range of logical indices: the workgroups running together on one XCD then work on
// neighbouring items (a bijection on [0, nb)).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_contiguous(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, slot = b / kXcds;
    return x * q + min(x, r) + slot;
}

// One wave's list in LDS, as arrays indexed by entry so that lane-indexed accesses
// are free of bank conflicts and entry-indexed ones are broadcasts.
//   runs: run r of samples in one cell (ci, cj), with per stencil corner c the element
//         offset texel*64 in the unpadded grid (-1 outside the map, or when a later
//         run of the window adds this corner's weight) and the summed bilinear weight;
//   tex:  the kept corners of a window compacted, with the element offset in the
//         padded texel-major maps; at most four per run.
union WaveList {
    struct {
        double w[4][kWave];
        int off[4][kWave];
        int ci[kWave], cj[kWave];
    } runs;
    struct {
        union {
            double w[4 * kWave];   // the scatter walks' weights
            float wf[4 * kWave];   // the trace walk's weights
        };
        int off[4 * kWave];
    } tex;
};

// Orders a wave's LDS accesses around a list update.  A wave's LDS accesses
// complete in issue order, so only the compiler must keep the reads below the
// writes (a memory fence would also wait for every global access in flight).
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Run r of a window, settled, in lane r's registers: its cell and, per stencil
// corner, the element offset texel*64 in the unpadded grid (-1 outside the map or
// when a later run adds this corner's weight) and the total weight.
struct SettledRun {
    int ci, cj;
    int off[4];
    double w[4];
};

// Passes of 64 samples that one window can take: a pass adds at most one run per
// cell it enters, and its 63 steps of length step cross at most 63 step |ux| + 1
// columns and 63 step |uy| + 1 rows, so it adds at most 3 + 63 step (|ux| + |uy|)
// runs; windows hold up to 64 runs.
constexpr int kMaxPasses = 4;
__device__ __forceinline__ int window_passes(const Ray &ray, double step) {
    const double c = 63.0 * step * (fabs(ray.ux) + fabs(ray.uy));
    const int b = 4 + (int)fmin(c, 64.0);
    return max(1, min(kMaxPasses, kWave / b));
}

// The runs of samples q0 .. q0 + 64 passes - 1 of a ray (clamped to its n samples)
// that share a bilinear cell, in ray order; returns their number nr, run r settled in
// lane r's registers.  Each pass takes one sample per lane; each run-end lane forms
// its run's FP64 bilinear weight sums in closed form over the run's samples in the
// pass and writes its cell to the wave's run list, and a run that continues from the
// previous pass adds its sums to that pass's last entry.  A stencil corner is shared
// by at most three consecutive cells of a straight ray (the cell indices are monotone
// along it), so each run then folds in the weights of the two runs before it for the
// corners it shares with them and drops the corners that one of the next two runs
// shares: every texel of the window appears once, with its total weight.  The run
// list is free again on return.  passes <= window_passes(ray) keeps nr <= 64.  FLAT
// issues every LDS read of the settle before the adjacency tests select among them
// (no dependent read behind a branch); the other form reads under the tests and
// holds fewer registers.  Both give the same bits.  MAXP = 1 compiles the one-pass
// window only.
template <bool FLAT, int MAXP>
__device__ __forceinline__ int cell_runs_settled(const Ray &ray, double step, int n, int q0,
                                                 int passes, int nx, int ny, int lane, WaveList &L,
                                                 SettledRun &sr) {
    const double sx = ray.ux * step, sy = ray.uy * step;
    const unsigned long long lower = (1ull << lane) - 1ull;
    int nr = 0, ci_prev = 0, cj_prev = 0;
    for (int p = 0; p < (MAXP == 1 ? 1 : passes); p++) {
        const int qb = q0 + p * kWave;
        if (qb >= n) break;  // uniform
        const int cnt = min(kWave, n - qb);
        int i0, j0;
        double ax, ay;
        lane_sample(ray, step, min(qb + lane, n - 1), i0, j0, ax, ay);
        const bool act = lane < cnt;
        const int pi = __shfl_up(i0, 1, kWave), pj = __shfl_up(j0, 1, kWave);
        const int ni = __shfl_down(i0, 1, kWave), nj = __shfl_down(j0, 1, kWave);
        const bool first = lane == 0 ? (p == 0 || ci_prev != i0 || cj_prev != j0)
                                     : (pi != i0 || pj != j0);
        const bool last = act && (lane == cnt - 1 || ni != i0 || nj != j0);
        const unsigned long long starts = __ballot(first);
        // the first run of a later pass may continue the previous pass's last run
        const bool cont = (starts & 1ull) == 0ull;
        // first lane of this lane's run in the pass: highest start at or below it
        // (none: the continued run, from lane 0)
        const unsigned long long below = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int rs = below ? 63 - __clzll(below) : 0;
        // the run's weight sums in closed form: along the run the fractions fall back
        // from this lane's (ax, ay) by (sx, sy) per sample (positions are linear in the
        // sample index), so with k = 0 .. nn-1 samples back,
        //   sum ax = nn ax - sx S1,  sum ay = nn ay - sy S1,
        //   sum ax ay = nn ax ay - (ax sy + ay sx) S1 + sx sy S2,
        // S1 = sum k, S2 = sum k^2; only the run-end lanes' values are used
        // (S1 and S2 are integers below 2^19, formed exactly in integer arithmetic)
        const int m = lane - rs + 1;
        const double nn = (double)m;
        const double s1 = (double)((m * (m - 1)) >> 1), s2 = (double)((m - 1) * m * (2 * m - 1) / 6);
        const double sax = nn * ax - sx * s1, say = nn * ay - sy * s1;
        const double saxy = nn * ax * ay - (ax * sy + ay * sx) * s1 + sx * sy * s2;
        const double w[4] = {nn - sax - say + saxy, sax - saxy, say - saxy, saxy};
        const unsigned long long ends = __ballot(last);
        if (last) {
            const int r = nr + __popcll(ends & lower) - (cont ? 1 : 0);
            if (cont && r == nr - 1) {
#pragma unroll
                for (int c = 0; c < 4; c++) L.runs.w[c][r] += w[c];
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int ti = i0 + (c & 1), tj = j0 + (c >> 1);
                    const bool in = ti >= 0 && tj >= 0 && ti < nx && tj < ny;
                    L.runs.off[c][r] = in ? (ti + nx * tj) * kTexelMaps : -1;
                    L.runs.w[c][r] = w[c];
                }
                L.runs.ci[r] = i0;
                L.runs.cj[r] = j0;
            }
        }
        nr += __popcll(ends) - (cont ? 1 : 0);
        ci_prev = __builtin_amdgcn_readlane(i0, kWave - 1);
        cj_prev = __builtin_amdgcn_readlane(j0, kWave - 1);
        // a wave's LDS accesses complete in issue order, so only the compiler must
        // keep the list reads below the writes (a memory fence would also wait for
        // every global access in flight)
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
    // lane r < nr settles run r's corners against runs r-2 .. r+2
    sr.ci = sr.cj = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        sr.off[c] = -1;
        sr.w[c] = 0.0;
    }
    if (FLAT && lane < nr) {
        // every read first (indices clamped into the list, so each is in range),
        // then the adjacency selects: no branch waits on an LDS read
        const int ci = L.runs.ci[lane], cj = L.runs.cj[lane];
        sr.ci = ci;
        sr.cj = cj;
        int nci[4], ncj[4];  // runs lane-2, lane-1, lane+1, lane+2
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = s < 2 ? s - 2 : s - 1;
            const int r = min(max(lane + k, 0), kWave - 1);
            const bool ok = lane + k >= 0 && lane + k < nr;
            const int vi = L.runs.ci[r], vj = L.runs.cj[r];
            nci[s] = ok ? vi : INT_MIN / 2;
            ncj[s] = ok ? vj : INT_MIN / 2;
        }
        double own[4], prev[2][4];
        int offs[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            own[c] = L.runs.w[c][lane];
            offs[c] = L.runs.off[c][lane];
            const int ti = ci + (c & 1), tj = cj + (c >> 1);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                // the corner of run lane-2+s that is texel (ti, tj), if it is adjacent
                const int di = ti - nci[s], dj = tj - ncj[s];
                prev[s][c] = L.runs.w[(di & 1) + 2 * (dj & 1)][max(lane + s - 2, 0)];
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int ti = ci + (c & 1), tj = cj + (c >> 1);
            bool later = false;
            double wsum = own[c];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int di = ti - nci[s], dj = tj - ncj[s];
                const bool adj = di >= 0 && di <= 1 && dj >= 0 && dj <= 1;
                if (s >= 2)
                    later = later || adj;
                else
                    wsum = adj ? wsum + prev[s][c] : wsum;
            }
            sr.off[c] = later ? -1 : offs[c];
            sr.w[c] = wsum;
        }
    }
    if (!FLAT && lane < nr) {
        // the same settle with the reads under the adjacency tests: fewer registers
        const int ci = L.runs.ci[lane], cj = L.runs.cj[lane];
        sr.ci = ci;
        sr.cj = cj;
        int nci[5], ncj[5];
#pragma unroll
        for (int k = -2; k <= 2; k++) {
            const bool ok = k != 0 && lane + k >= 0 && lane + k < nr;
            nci[k + 2] = ok ? L.runs.ci[lane + k] : INT_MIN / 2;
            ncj[k + 2] = ok ? L.runs.cj[lane + k] : INT_MIN / 2;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int ti = ci + (c & 1), tj = cj + (c >> 1);
            bool later = false;
            double wsum = L.runs.w[c][lane];
#pragma unroll
            for (int k = -2; k <= 2; k++) {
                if (k == 0) continue;
                const int di = ti - nci[k + 2], dj = tj - ncj[k + 2];
                if (di < 0 || di > 1 || dj < 0 || dj > 1) continue;
                if (k > 0)
                    later = true;
                else
                    wsum += L.runs.w[di + 2 * dj][lane + k];
            }
            sr.off[c] = later ? -1 : L.runs.off[c][lane];
            sr.w[c] = wsum;
        }
    }
    wave_lds_order();
    return nr;
}

// Compacts the kept corners of the settled runs into the wave's texel list, in lane
// order and corner order within a lane, with each texel's weight:
//   kListScatter: the element offset in the unpadded grid, FP64 weight;
//   kListTrace64: the element offset in the padded texel-major maps, FP64 weight;
//   kListTrace32: the byte offset in the padded texel-major maps (below 2^32: the host
//                 keeps those maps under 4 GiB), FP32 weight.
// The list is padded with (0, 0) entries to a whole number of blocks of kTexBlock, so
// a walk reads aligned 128-bit broadcasts; returns the entry count, nblk the padded
// length.  Offset 0 is the zero border texel of the padded maps and a real texel of
// the unpadded grid, so walks that scatter skip the padding.
constexpr int kTexBlock = 16;
static_assert((4 * kWave) % kTexBlock == 0, "the texel list holds whole blocks");
enum { kListScatter, kListTrace64, kListTrace32 };

template <int MODE>
__device__ __forceinline__ int compact_texels(WaveList &L, int lane, const SettledRun &sr, int nx,
                                              int &nblk) {
    const int pw = nx + 2;
    bool kept[4];
    int off[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        // kept corners lie inside the grid, so the padded index is in range
        kept[c] = sr.off[c] >= 0;
        const unsigned pt = (unsigned)((sr.ci + (c & 1) + 1) + pw * (sr.cj + (c >> 1) + 1));
        off[c] = MODE == kListTrace32   ? (int)(pt * (unsigned)(kTexelMaps * sizeof(float)))
                 : MODE == kListTrace64 ? (int)(pt * (unsigned)kTexelMaps)
                                        : sr.off[c];
    }
    const unsigned long long lower = (1ull << lane) - 1ull;
    int pos = 0, nent = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const unsigned long long bal = __ballot(kept[c]);
        pos += __popcll(bal & lower);
        nent += __popcll(bal);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (kept[c]) {
            L.tex.off[pos] = off[c];
            if (MODE == kListTrace32)
                L.tex.wf[pos] = (float)sr.w[c];
            else
                L.tex.w[pos] = sr.w[c];
            pos++;
        }
    }
    nblk = (nent + kTexBlock - 1) / kTexBlock * kTexBlock;
    if (nent + lane < nblk) {
        L.tex.off[nent + lane] = 0;
        if (MODE == kListTrace32)
            L.tex.wf[nent + lane] = 0.f;
        else
            L.tex.w[nent + lane] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    return nent;
}

// Line integral (before the factor step), for lane b's map, of one ray through the
// texel-major maps: sum over texels of m[texel][b] times the texel's summed bilinear
// weight.  Each window lists its texels once (cell_runs_settled, compact_texels)
// and the wave walks the list with sixteen coalesced 256-byte loads in flight.  A
// texel costs one load per window instead of one per sample and corner.
//   F64: FP64 weights and FP64 sums of the products (four partial sums, added in a
//        fixed order);
//   otherwise: FP32 weights, 32-bit byte offsets from the maps' base (one add per
//        load), and each block's sixteen products summed in FP32 (four partial sums
//        of four) before the block is added in FP64, so a block adds at most
//        4 * 2^-24 of the sum of its terms' magnitudes to the error of the sum, and
//        the FP32 products and weights 2^-23 of it.
template <bool FLAT, int MAXP, bool F64>
__device__ __forceinline__ double wave_trace_tm(const float *__restrict__ m, int nx, int ny,
                                                double step, const Ray &ray, int lane,
                                                WaveList &L) {
    const int n = __builtin_amdgcn_readfirstlane(ray.n);
    const char *__restrict__ mb = reinterpret_cast<const char *>(m);
    const unsigned lane4 = (unsigned)lane * (unsigned)sizeof(float);
    const int passes = MAXP == 1 ? 1 : __builtin_amdgcn_readfirstlane(min(MAXP, window_passes(ray, step)));
    double total = 0.0, acc64[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < n; q0 += kWave * passes) {
        SettledRun sr;
        cell_runs_settled<FLAT, MAXP>(ray, step, n, q0, passes, nx, ny, lane, L, sr);
        int nblk;
        compact_texels<F64 ? kListTrace64 : kListTrace32>(L, lane, sr, nx, nblk);
        for (int e0 = 0; e0 < nblk; e0 += kTexBlock) {
            float v[kTexBlock];
#pragma unroll
            for (int k = 0; k < kTexBlock; k++)
                v[k] = F64 ? m[(unsigned)(L.tex.off[e0 + k] + lane)]
                           : *reinterpret_cast<const float *>(mb + ((unsigned)L.tex.off[e0 + k] + lane4));
            __builtin_amdgcn_sched_barrier(0);  // keep every load above the first use
            if (F64) {
#pragma unroll
                for (int k = 0; k < kTexBlock; k++)
                    acc64[k & 3] = fma(L.tex.w[e0 + k], (double)v[k], acc64[k & 3]);
            } else {
                float acc[4];
#pragma unroll
                for (int k = 0; k < 4; k++) acc[k] = L.tex.wf[e0 + k] * v[k];
#pragma unroll
                for (int k = 4; k < kTexBlock; k++) acc[k & 3] = fmaf(L.tex.wf[e0 + k], v[k], acc[k & 3]);
                total += (double)((acc[0] + acc[1]) + (acc[2] + acc[3]));
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
    return F64 ? (acc64[0] + acc64[1]) + (acc64[2] + acc64[3]) : total;
}

// Settle form and passes per window per kernel (cell_runs_settled): the flat form and
// multi-pass windows where registers are to spare; the branch form and one-pass
// windows in the VJP kernel, whose FP64 gradient totals stay live through every trace,
// and in the bundled scatter's fallback for windows too large for its box.
constexpr bool kFlatForward = true, kFlatVjp = false, kFlatBundle = false;
constexpr int kPassesForward = kMaxPasses, kPassesVjp = 1;

// Detector legs of a voxel chunk, one wave per ray, one lane per map.
template <bool FLAT, int MAXP>
__device__ __forceinline__ void trace_chunk_tm(const CoeffArgs &a, const PixG &pg, const VoxC *vox,
                                               int nk, float *Ld, WaveList *lists) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    for (int j = wid; j < nk; j += NT / kWave) {
        const Ray ray = make_ray(pg.dx, pg.dy, (double)vox[j].ix + 0.5, (double)vox[j].iy + 0.5,
                                 a.f.nx, a.f.ny, a.f.ray_step);
        const double sum = wave_trace_tm<FLAT, MAXP, false>(a.maps_tm, a.f.nx, a.f.ny, a.f.ray_step,
                                                            ray, lane, lists[wid]);
        if (lane < a.n_basis) Ld[j * LROW + lane] = (float)(sum * a.f.ray_step);
    }
}

// Attenuation exponents of a voxel chunk's rays on the matrix cores:
//   mu[j * XROW + e] = sum_b beta[b, e] (Ls[j][b] + Ld[j][b]),  j < KC voxels, e < E,
// the basis contraction mu(E) = sum_b beta[b, E] m_b applied to the rays' line
// integrals (Ls, Ld: [KC x LROW] FP32; bet: beta [B x E], b-major).  Waves 0..3 of the
// workgroup each take 16 energies (one 16 x 16 tile of mu) and step over the maps four
// at a time with v_mfma_f64_16x16x4f64: lane l holds A[l % 16][l / 16] (voxel l % 16,
// map b0 + l / 16), B[l / 16][l % 16] (map b0 + l / 16, energy e0 + l % 16) and
// D[l / 16 + 4 r][l % 16], r < 4 (the output layout measured on gfx950).  The sum of
// the two FP32 integrals and its product with an FP32 beta are exact in FP64, so only
// the FP64 accumulation rounds, in the matrix core's fixed order.  Voxels j >= nk and
// maps b >= B enter as zeros.  The other waves return at once; the caller's barrier
// orders mu before its readers.
typedef double f64x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void chunk_mu(const float *Ls, const float *Ld, const float *bet, int B,
                                         int E, int nk, double *mu) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    const int e0 = wid * 16;
    if (wid >= 4 || e0 >= E) return;  // uniform per wave
    const int i = lane & 15, kq = lane >> 4, e = e0 + (lane & 15);
    f64x4v acc = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = 0; b0 < B; b0 += 4) {
        const int b = b0 + kq;
        const double av = (b < B && i < nk) ? (double)Ls[i * LROW + b] + (double)Ld[i * LROW + b] : 0.0;
        const double bv = (b < B && e < E) ? (double)bet[b * E + e] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) mu[(kq + 4 * r) * XROW + e] = acc[r];
}

// Ray weights of a voxel chunk's map gradient on the matrix cores:
//   gamma[j][b] = -sum_e beta[b, e] X[j * XROW + e],  X = W(e) P(e) S_e,
// and gsrc[j][b] += gamma[j][b] (the source legs' sum over the pixels), for j < nk,
// b < B (both [KC x MAXB]).  Waves 0..3 each take 16 maps and step over the energies
// four at a time with v_mfma_f64_16x16x4f64 (the lane layout of chunk_mu, with voxels
// as rows, energies as the inner dimension and maps as columns).  Voxels j >= nk and
// energies e >= E enter as zeros.  The other waves return at once; the caller's barrier
// orders the results before their readers.
__device__ __forceinline__ void chunk_gamma(const double *X, const float *bet, int B, int E, int nk,
                                            double *gam, double *gsrc) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    const int b0 = wid * 16;
    if (wid >= 4 || b0 >= B) return;  // uniform per wave
    const int j = lane & 15, kq = lane >> 4, b = b0 + (lane & 15);
    f64x4v acc = {0.0, 0.0, 0.0, 0.0};
    for (int e0 = 0; e0 < E; e0 += 4) {
        const int e = e0 + kq;
        const double av = (e < E && j < nk) ? X[j * XROW + e] : 0.0;
        const double bv = (e < E && b < B) ? (double)bet[b * E + e] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int jr = kq + 4 * r;
        if (jr < nk && b < B) {
            gam[jr * MAXB + b] = -acc[r];
            gsrc[jr * MAXB + b] -= acc[r];
        }
    }
}

// Per-(voxel, energy) table: hi/lo wvt, c2 = -log2e/(2 sigma^2), prefactor P, band;
// mu: the ray's attenuation exponent at energy e (chunk_mu).
__device__ __forceinline__ void table_entry(const CoeffArgs &a, const VoxC &vs, double mu, int e,
                                            float4 &tw, float &band, double &P) {
    const EnergyConst ec = a.f.energies[e];
    const double sig2 = vs.As + vs.Bs * ec.k2sq;
    const double w = ec.wcoef * vs.s;
    const float wh = (float)w;
    const float wl = (float)(w - (double)wh);
    const double c2 = -kLog2e / (2.0 * sig2);
    P = exp(-mu) * vs.gR * kRs2pi / sqrt(sig2);
    tw = make_float4(wh, wl, (float)c2, (float)P);
    band = (float)sqrt(64.0 / -c2);
}

}  // namespace

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(NT) void coeff_list_kernel(CoeffArgs a) {
    __shared__ int wsum[16];
    const int npix = a.f.nx * a.f.ny;
    int base = 0;
    for (int c = 0; c < npix; c += blockDim.x) {
        const int ro = c + threadIdx.x;
        bool f = false;
        int pix = 0;
        if (ro < npix) {
            const int ix = ro / a.f.ny, iy = ro % a.f.ny;
            pix = ix + a.f.nx * iy;
            f = a.seg[pix] > 0;
        }
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const unsigned long long m = __ballot(f);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int pre = 0, tot = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
            if (i < w) pre += wsum[i];
            tot += wsum[i];
        }
        __syncthreads();
        if (f) a.vlist[base + pre + below] = pix;
        base += tot;
    }
    if (threadIdx.x == 0) a.nvox[0] = base;
}

__global__ void coeff_pad_kernel(CoeffArgs a) {
    const int pw = a.f.nx + 2, ph = a.f.ny + 2;
    const size_t plane = (size_t)pw * ph;
    const size_t n = plane * a.n_basis;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / plane);
        const int q = (int)(i % plane);
        const int px = q % pw, py = q / pw;
        float v = 0.f;
        if (px >= 1 && px <= a.f.nx && py >= 1 && py <= a.f.ny)
            v = a.maps[(size_t)b * a.f.nx * a.f.ny + (px - 1) + (size_t)a.f.nx * (py - 1)];
        a.maps_padded[i] = v;
    }
}

// maps_tm[(px) + (nx+2) py][b]: texel-major copy of the maps with the zero border.
__global__ void coeff_pad_tm_kernel(CoeffArgs a) {
    const int pw = a.f.nx + 2, ph = a.f.ny + 2;
    const size_t n = (size_t)pw * ph * kTexelMaps;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i % kTexelMaps);
        const int q = (int)(i / kTexelMaps);
        const int px = q % pw, py = q / pw;
        float v = 0.f;
        if (b < a.n_basis && px >= 1 && px <= a.f.nx && py >= 1 && py <= a.f.ny)
            v = a.maps[(size_t)b * a.f.nx * a.f.ny + (px - 1) + (size_t)a.f.nx * (py - 1)];
        a.maps_tm[i] = v;
    }
}

// grad_maps[b][pix] = grad_tm[pix][b]
__global__ void coeff_untranspose_kernel(CoeffArgs a) {
    const size_t npix = (size_t)a.f.nx * a.f.ny;
    const size_t n = npix * a.n_basis;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / npix, pix = i % npix;
        a.grad_maps[i] = a.grad_tm[pix * kTexelMaps + b];
    }
}

// Source legs: one thread per (view, voxel), all B maps.
__global__ __launch_bounds__(256) void coeff_src_kernel(CoeffArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (k >= a.nvox[0]) return;
    const int pix = a.vlist[k];
    const int ix = pix % a.f.nx, iy = pix / a.f.nx;
    const ViewConst vc = a.f.views[v];
    const Ray r = make_ray(vc.src_x, vc.src_y, (double)ix + 0.5, (double)iy + 0.5, a.f.nx, a.f.ny,
                           a.f.ray_step);
    const size_t plane = (size_t)(a.f.nx + 2) * (a.f.ny + 2);
    const size_t npix = (size_t)a.f.nx * a.f.ny;
    for (int b0 = 0; b0 < a.n_basis; b0 += 4) {
        Kahan ks[4];
        for (int q = 0; q < r.n; q++) {
            double px, py;
            ray_sample(r, a.f.ray_step, q, px, py);
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (b0 + u < a.n_basis)
                    ks[u].add(bil_planar(a.maps_padded, (unsigned)((b0 + u) * plane), a.f.nx,
                                         a.f.ny, px, py));
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (b0 + u < a.n_basis)
                a.src_ints[((size_t)v * npix + k) * a.n_basis + b0 + u] =
                    (float)((double)ks[u].value() * a.f.ray_step);
    }
}

// Source legs with texel-major maps (B > 4): one wave per (view, voxel) ray, every
// map at once through wave_trace_tm (lane b: map b) instead of one thread walking
// the ray once per four planar maps.
__global__ __launch_bounds__(256) void coeff_src_tm_kernel(CoeffArgs a) {
    __shared__ WaveList lists[4];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int k = blockIdx.x * 4 + wid, v = blockIdx.y;
    if (k >= a.nvox[0]) return;  // uniform per wave
    const int pix = a.vlist[k];
    const int nx = a.f.nx, ny = a.f.ny;
    const ViewConst vc = a.f.views[v];
    const Ray r = make_ray(vc.src_x, vc.src_y, (double)(pix % nx) + 0.5, (double)(pix / nx) + 0.5,
                           nx, ny, a.f.ray_step);
    const double sum =
        wave_trace_tm<kFlatForward, kPassesForward, false>(a.maps_tm, nx, ny, a.f.ray_step, r, lane,
                                                            lists[wid]);
    if (lane < a.n_basis)
        a.src_ints[((size_t)v * nx * ny + k) * a.n_basis + lane] = (float)(sum * a.f.ray_step);
}

// Detector legs of a slab of views for the VJP: one wave per (view of the slab,
// voxel, pixel) ray, the line integrals of all maps (lane b: map b) stored at
// ld_buf[((view - v_base) nv + voxel) n_det + pixel][b].  The same trace as the
// forward kernel's (trace_chunk_tm), so the VJP sees the forward's bits.
constexpr int kTraceWaves = 4;
// Launches over a slab's rays are capped at kRayLaunchBlocks workgroups: a slab of many
// views has more rays than one dispatch can count in work-items (2^32).  The trace is
// issued as several launches, each from item a.item_base on (its body stays a single
// ray per wave and keeps its registers); the scatter kernels stride over their items.
constexpr unsigned kRayLaunchBlocks = 1u << 20;
__host__ __device__ inline unsigned ray_blocks(int64_t items, int per_block) {
    const int64_t b = (items + per_block - 1) / per_block;
    return (unsigned)(b < (int64_t)kRayLaunchBlocks ? b : (int64_t)kRayLaunchBlocks);
}

// One wave traces kWave consecutive rays: lane l first sets up ray base + l (pixel
// geometry and the FP64 intersection with the grid, the same arithmetic as one wave per
// ray, in parallel over the lanes) into LDS, then the wave traces the rays one after
// another.  Each ray's line integrals are the same bits as when a wave computed the ray
// itself.
__global__ __launch_bounds__(kTraceWaves *kWave) void coeff_trace_kernel(CoeffArgs a) {
    __shared__ WaveList lists[kTraceWaves];
    __shared__ Ray rays[kTraceWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int P = a.f.n_det, nv = a.n_vox, nx = a.f.nx;
    const int64_t total = (int64_t)a.v_count * nv * P;
    const int64_t base = a.item_base + ((int64_t)blockIdx.x * kTraceWaves + wid) * kWave;
    if (base >= total) return;  // uniform per wave
    {
        const int64_t item = base + lane;
        Ray r{};
        if (item < total) {
            const int p = (int)(item % P);
            const int k = (int)((item / P) % nv);
            const int vs = (int)(item / ((int64_t)P * nv));
            const PixG pg = pixel_setup(a, a.v_base + vs, p);
            const int pix = a.vlist[k];
            r = make_ray(pg.dx, pg.dy, (double)(pix % nx) + 0.5, (double)(pix / nx) + 0.5, nx,
                         a.f.ny, a.f.ray_step);
        }
        rays[wid][lane] = r;
    }
    wave_lds_order();
    const int cnt = (int)min((int64_t)kWave, total - base);
    for (int i = 0; i < cnt; i++) {
        const Ray ray = rays[wid][i];
        const double sum = wave_trace_tm<kFlatForward, kPassesForward, false>(
            a.maps_tm, nx, a.f.ny, a.f.ray_step, ray, lane, lists[wid]);
        a.ld_buf[(size_t)(base + i) * kTexelMaps + lane] = (float)(sum * a.f.ray_step);
    }
}

// The detector legs of a slab (a.v_base, a.v_count) into a.ld_buf, in launches of at
// most kRayLaunchBlocks workgroups of kTraceWaves waves of kWave rays.
static void launch_trace_slab(CoeffArgs a, hipStream_t s) {
    const int64_t rays = (int64_t)a.v_count * a.n_vox * a.f.n_det;
    const int64_t per_block = (int64_t)kTraceWaves * kWave;
    const int64_t per_launch = (int64_t)kRayLaunchBlocks * per_block;
    for (int64_t i0 = 0; i0 < rays; i0 += per_launch) {
        a.item_base = i0;
        hipLaunchKernelGGL(coeff_trace_kernel, dim3(ray_blocks(rays - i0, (int)per_block)),
                           dim3(kTraceWaves * kWave), 0, s, a);
    }
}

// Forward counts: one workgroup per (view, pixel).
// Four waves per SIMD (two workgroups per CU) at 128 VGPRs: the spills this costs
// are cheaper than the latency that two waves per SIMD leave exposed (fan256, 12 views:
// 0.98 -> 0.69 s).  Two workgroups need the LDS below 80 KB, so the trace's run
// lists share their space with the chunk's table, which is never live at the same time.
template <bool LDBUF>
__global__ __launch_bounds__(NT, 4) void coeff_forward_kernel(CoeffArgs a) {
    __shared__ VoxC vox[KC];
    __shared__ float2 part[NT];
    __shared__ float Ld[KC * LROW];
    __shared__ float Fq[KC * TQ];
    __shared__ double red[(NT / kWave) * EPT + 64];
    __shared__ float bet[MAXB * 64];
    __shared__ float Lsrc[KC * LROW];
    // table rows of kTS entries: the ballot's per-lane reads of one energy across
    // the chunk's voxels then fall into different banks
    constexpr int kTS = 65;
    __shared__ union {
        WaveList cruns[NT / kWave];
        struct {
            float4 tw[KC * kTS];
            float tb[KC * kTS];
            double mu[KC * XROW];
        } tab;
    } ov;
    WaveList *cruns = ov.cruns;
    float4 *tw = ov.tab.tw;
    float *tb = ov.tab.tb;
    double *mu = ov.tab.mu;
    const ForwardArgs &f = a.f;
    const int t = threadIdx.x, tq = t % TQ, h = t / TQ;
    const int V = f.n_views, E = f.n_energy, Q = f.n_q;
    // neighbouring pixels of one view run together on an XCD: their ray fans from a
    // voxel chunk cross nearly the same texels, which stay in that XCD's L2
    const int l = xcd_contiguous(blockIdx.x, gridDim.x);
    const int v = a.v_base + l / f.n_det, p = l % f.n_det;
    const size_t npix = (size_t)f.nx * f.ny;
    const PixG pg = pixel_setup(a, v, p);
    float qch = 1e30f, qcl = 0.f, dqc = 0.f;
    if (tq < Q) {
        const double q = f.q_coarse[tq];
        qch = (float)q;
        qcl = (float)(q - (double)qch);
        dqc = (float)f.q_coarse[Q + tq];
    }
    float lo = tq < Q ? qch : 1e30f, hi = tq < Q ? qch : -1e30f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, kWave));
        hi = fmaxf(hi, __shfl_xor(hi, o, kWave));
    }
    // fine totals per wave and energy slot, FP64 in LDS: each chunk's FP32 partials
    // are summed over the wave (16-value transposed sum) and added once per chunk,
    // so no per-lane total is held in registers through the ray traces
    static_assert(EPT <= 16, "the totals use the 16-value transposed wave sum");
    __shared__ double ftw[NT / kWave][16];
    if (t < NT / kWave * 16) (&ftw[0][0])[t] = 0.0;
    const int nv = a.nvox[0];
    for (int idx = t; idx < a.n_basis * E; idx += NT) bet[idx] = a.beta[idx];
    for (int k0 = 0; k0 < nv; k0 += KC) {
        const int nk = min(KC, nv - k0);
        if (t < nk) {
            VoxC vs;
            vs.k = k0 + t;
            vs.pix = a.vlist[k0 + t];
            vs.ix = vs.pix % f.nx;
            vs.iy = vs.pix / f.nx;
            geom_c(f, pg, vs.ix, vs.iy, vs);
            vox[t] = vs;
        }
        for (int idx = t; idx < nk * a.n_basis; idx += NT) {
            const int j = idx / a.n_basis, b = idx % a.n_basis;
            Lsrc[j * LROW + b] = a.src_ints[((size_t)v * npix + k0 + j) * a.n_basis + b];
        }
        __syncthreads();
        if (LDBUF) {
            for (int idx = t; idx < nk * a.n_basis; idx += NT) {
                const int j = idx / a.n_basis, b = idx % a.n_basis;
                Ld[j * LROW + b] = a.ld_buf[(((size_t)(v - a.v_base) * a.n_vox + (size_t)(k0 + j)) * f.n_det + p) *
                                                kTexelMaps + b];
            }
        } else if (a.maps_tm) {
            trace_chunk_tm<kFlatForward, kPassesForward>(a, pg, vox, nk, Ld, cruns);
        } else {
            trace_chunk(a, pg, vox, nk, part, Ld);
        }
        // F_vox(q_n) of the chunk, formed once per call (coeff_fq_kernel)
        for (int idx = t; idx < nk * TQ; idx += NT) Fq[idx] = a.fq_buf[(size_t)k0 * TQ + idx];
        __syncthreads();
        chunk_mu(Lsrc, Ld, bet, a.n_basis, E, nk, mu);
        __syncthreads();
        for (int idx = t; idx < nk * E; idx += NT) {
            const int j = idx / E, e = idx % E;
            float4 tv;
            float band;
            double P;
            table_entry(a, vox[j], mu[j * XROW + e], e, tv, band, P);
            tw[j * kTS + e] = tv;
            tb[j * kTS + e] = band;
        }
        __syncthreads();
        // per source energy, one ballot gives the chunk's voxels whose band reaches
        // the wave's bins (the range and the energy are wave-uniform); the wave walks
        // those only, in ascending voxel order
        {
            const int ln = t & (kWave - 1);
            float pf[16];
#pragma unroll
            for (int i = 0; i < 16; i++) pf[i] = 0.f;
#pragma unroll
            for (int i = 0; i < EPT; i++) {
                const int e = h + G * i;
                if (e >= E) continue;
                bool in = false;
                if (ln < nk) {
                    const float x = tw[ln * kTS + e].x, band = tb[ln * kTS + e];
                    in = x + band >= lo && x - band <= hi;
                }
                uint64_t m = __ballot(in);
                auto term = [&](int j) {
                    const float4 tv = tw[j * kTS + e];
                    const float wq = dqc * Fq[j * TQ + tq];
                    const float d = (tv.x - qch) + (tv.y - qcl);
                    return make_float2(tv.w * wq, __builtin_amdgcn_exp2f(d * d * tv.z));
                };
                // two voxels per step while two are left, their reads in flight together
                while (m & (m - 1)) {
                    const int j0 = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const int j1 = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const float2 u0 = term(j0), u1 = term(j1);
                    pf[i] = fmaf(u0.x, u0.y, pf[i]);
                    pf[i] = fmaf(u1.x, u1.y, pf[i]);
                }
                if (m) {
                    const float2 u = term(__ffsll((unsigned long long)m) - 1);
                    pf[i] = fmaf(u.x, u.y, pf[i]);
                }
            }
            const float sf = transpose_sum16(pf, ln);
            if ((ln & 3) == 0) ftw[t >> 6][(ln >> 2) & 15] += (double)sf;
        }
        __syncthreads();
    }
    const int wid = t >> 6;
    if (t < NT / kWave * EPT) red[(t / EPT) * EPT + t % EPT] = ftw[t / EPT][t % EPT];
    __syncthreads();
    double *hf = red + (NT / kWave) * EPT;
    if (t < E) {
        const int hh = t % G, i = t / G;
        double s = 0.0;
        for (int w = hh * (TQ / kWave); w < (hh + 1) * (TQ / kWave); w++) s += red[w * EPT + i];
        hf[t] = s;
    }
    __syncthreads();
    if (t < E) {
        int wlo, whi;
        src_window(t, E, f.dn, wlo, whi);
        double s = 0.0;
        for (int e = wlo; e <= whi; e++) s += (double)f.spectrum[e + (size_t)E * t] * hf[e];
        a.out_counts[(size_t)v + (size_t)V * ((size_t)p + (size_t)f.n_det * t)] = (float)s;
    }
}

// Adds gamma[b] * step * (bilinear weight) of every sample q in [q0, q1) of a ray
// to grad[b][texel] (the transpose of the line integral, attenuator.cu:52-58).
// Consecutive samples in the same texel cell are merged first, so each cell the
// ray crosses costs one FP64 atomic add per corner and map; texels outside the
// map (the zero border) receive nothing.
__device__ __forceinline__ void scatter_ray(const CoeffArgs &a, const Ray &ray, const double *gam,
                                            int q0, int q1) {
    const int nx = a.f.nx, ny = a.f.ny, B = a.n_basis;
    const size_t npix = (size_t)nx * ny;
    int ci = INT_MIN, cj = INT_MIN;
    double w00 = 0.0, w10 = 0.0, w01 = 0.0, w11 = 0.0;
    auto flush = [&]() {
        if (ci == INT_MIN) return;
        const int ii[4] = {ci, ci + 1, ci, ci + 1}, jj[4] = {cj, cj, cj + 1, cj + 1};
        const double ww[4] = {w00, w10, w01, w11};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (ii[c] < 0 || jj[c] < 0 || ii[c] >= nx || jj[c] >= ny) continue;
            const size_t pix = (size_t)ii[c] + (size_t)nx * jj[c];
            const double ws = ww[c] * a.f.ray_step;
            for (int b = 0; b < B; b++)
                if (gam[b] != 0.0) unsafeAtomicAdd(a.grad_maps + (size_t)b * npix + pix, gam[b] * ws);
        }
    };
    for (int q = q0; q < q1; q++) {
        double px, py;
        ray_sample(ray, a.f.ray_step, q, px, py);
        const double xt = px - 0.5, yt = py - 0.5;
        const double fx = floor(xt), fy = floor(yt);
        const double ax = xt - fx, ay = yt - fy;
        const int i0 = (int)fx, j0 = (int)fy;
        if (i0 != ci || j0 != cj) {
            flush();
            ci = i0;
            cj = j0;
            w00 = w10 = w01 = w11 = 0.0;
        }
        w00 += (1 - ax) * (1 - ay);
        w10 += ax * (1 - ay);
        w01 += (1 - ax) * ay;
        w11 += ax * ay;
    }
    flush();
}

// Texel-major form of scatter_ray for one wave: for each window of samples,
// cell_runs_settled and compact_texels list the window's texels once each with their
// summed weights, and the wave walks the list: lane b adds gamma[b] times the weight
// to grad_tm[texel][b], one coalesced atomic instruction per texel.
__device__ __forceinline__ void scatter_ray_tm(const CoeffArgs &a, const Ray &ray, const double *gam,
                                               int lane, WaveList &L) {
    const double gs = (lane < a.n_basis ? gam[lane] : 0.0) * a.f.ray_step;
    const int n = __builtin_amdgcn_readfirstlane(ray.n);
    const int passes = 1;
    for (int q0 = 0; q0 < n; q0 += kWave * passes) {
        SettledRun sr;
        cell_runs_settled<kFlatVjp, kPassesVjp>(ray, a.f.ray_step, n, q0, passes, a.f.nx, a.f.ny,
                                                lane, L, sr);
        int nblk;
        const int nent = compact_texels<kListScatter>(L, lane, sr, a.f.nx, nblk);
        if (gs != 0.0) {
            for (int e0 = 0; e0 < nblk; e0 += kTexBlock) {
                int off[kTexBlock];
                double w[kTexBlock];
#pragma unroll
                for (int k = 0; k < kTexBlock; k++) {
                    off[k] = L.tex.off[e0 + k];
                    w[k] = L.tex.w[e0 + k];
                }
#pragma unroll
                for (int k = 0; k < kTexBlock; k++)
                    if (e0 + k < nent) unsafeAtomicAdd(a.grad_tm + off[k] + lane, gs * w[k]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
}

// Detector legs of the map VJP in bundles of kBundle adjacent pixels: one wave per
// (view of the slab, voxel, pixel bundle).  Rays from one voxel to adjacent pixels
// stay within a small fraction of a texel of each other across the grid, so in each
// 64-sample window their texel lists nearly coincide.  The wave lists each ray's
// texels (cell_runs_settled) and its run lanes add their kept corners' weights into an
// LDS box over the window's texels, one weight column per ray; then every occupied
// texel is emitted once, in box order: sum_i w_i gamma_i[b] (FP64, formed for sixteen
// texels at a time by FP64 matrix products) is added to grad_tm[texel][b].  That is one
// atomic per texel, map and bundle instead of one per texel, map and ray.  A window
// whose box would exceed kBoxN texels scatters each ray's list directly.
constexpr int kBundle = 16, kBoxSide = 15, kBoxN = kBoxSide * kBoxSide, kBundleWaves = 2;
static_assert(kBundle % 4 == 0, "the emission products take the rays four at a time");

// A wave's LDS: the box (per texel, one weight per ray of the bundle; a texel counts as
// touched when one of its weights is nonzero, a corner of weight 0 adding nothing), the
// touched texels of a window in box order, and the bundle's rays.  A window too large
// for the box uses the start of the weight array as the run list of cell_runs_settled
// and clears it again, so 16.0 KB per wave leave room for five two-wave workgroups
// per CU.  Sixteen rays per bundle and a 15 x 15 box measured fastest on MI355X (map
// VJP of 12 views of a 256 x 256 phantom: 8 rays and a 17 x 17 box 0.82 s, 16 rays 0.73
// s with that box and 0.71-0.72 s with 15 x 15, 24 rays 0.82 s, 32 rays 0.90 s).
struct alignas(16) BundleBox {
    float w[kBoxN][kBundle];
    int list[kBoxN];
    Ray rays[kBundle];
};
static_assert(sizeof(WaveList) <= sizeof(float) * kBoxN * kBundle,
              "the fallback run list fits the box's weights");

// One 64-sample window of ray `col` of a bundle, added into the box: the runs of
// samples q0 .. q0 + 63 (clamped to the ray's n samples) that share a bilinear cell,
// each run's four corner weight sums formed in closed form by its run-end lane (as in
// cell_runs_settled), then added to box slot (ti - x0) + bw (tj - y0) of every corner
// (ti, tj) inside the grid, in four phases, one per stencil corner.  The cells along a
// straight ray are all distinct, so within a phase no two lanes touch the same slot and
// no settle against the neighbouring runs is needed: a texel that is a corner of up to
// three runs receives their weights in corner order.  The window's box must hold every
// in-grid corner of its runs.
__device__ __forceinline__ void window_runs_to_box(const Ray &ray, double step, int n, int q0,
                                                   int nx, int ny, int lane, int x0, int y0,
                                                   int bw, BundleBox &bx, int col) {
    const double sx = ray.ux * step, sy = ray.uy * step;
    const int cnt = min(kWave, n - q0);
    int i0, j0;
    double ax, ay;
    lane_sample(ray, step, min(q0 + lane, n - 1), i0, j0, ax, ay);
    const int pi = __shfl_up(i0, 1, kWave), pj = __shfl_up(j0, 1, kWave);
    const int ni = __shfl_down(i0, 1, kWave), nj = __shfl_down(j0, 1, kWave);
    const bool first = lane == 0 || pi != i0 || pj != j0;
    const bool last = lane < cnt && (lane == cnt - 1 || ni != i0 || nj != j0);
    const unsigned long long starts = __ballot(first);
    // first lane of this lane's run: the highest run start at or below it
    const unsigned long long below = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int rs = 63 - __clzll(below);  // lane 0 starts a run, so below != 0
    // closed-form weight sums over the run's m samples (cell_runs_settled)
    const int m = lane - rs + 1;
    const double nn = (double)m;
    const double s1 = (double)((m * (m - 1)) >> 1), s2 = (double)((m - 1) * m * (2 * m - 1) / 6);
    const double sax = nn * ax - sx * s1, say = nn * ay - sy * s1;
    const double saxy = nn * ax * ay - (ax * sy + ay * sx) * s1 + sx * sy * s2;
    const double w[4] = {nn - sax - say + saxy, sax - saxy, say - saxy, saxy};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int ti = i0 + (c & 1), tj = j0 + (c >> 1);
        if (last && ti >= 0 && tj >= 0 && ti < nx && tj < ny) {
            const int sl = (ti - x0) + bw * (tj - y0);
            bx.w[sl][col] += (float)w[c];
        }
        wave_lds_order();  // this phase's writes before the next phase's reads
    }
}

// Three waves per SIMD: at most 168 VGPRs.  The rays live in LDS and the per-ray loop
// is not unrolled, so no ray's parameters are held in registers across the window.
__global__ __launch_bounds__(kBundleWaves *kWave, 3) void coeff_scatter_bundle_kernel(CoeffArgs a,
                                                                                     int nb_pix) {
    __shared__ BundleBox boxes[kBundleWaves];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int nv = a.n_vox, P = a.f.n_det, nx = a.f.nx, ny = a.f.ny;
    const double step = a.f.ray_step;
    BundleBox &bx = boxes[wid];
    WaveList &L = *reinterpret_cast<WaveList *>(&bx.w[0][0]);
    for (int sl = lane; sl < kBoxN; sl += kWave) {
#pragma unroll
        for (int i = 0; i < kBundle; i++) bx.w[sl][i] = 0.f;
    }
    wave_lds_order();
    // one bundle per wave and pass; the launch is capped (ray_blocks) and the waves
    // stride over the slab's bundles.  The box is clear again after every window.
    const int64_t total = (int64_t)a.v_count * nv * nb_pix;
    for (int64_t item = (int64_t)blockIdx.x * kBundleWaves + wid; item < total;
         item += (int64_t)gridDim.x * kBundleWaves) {  // uniform per wave
        const int bnd = (int)(item % nb_pix);
        const int k = (int)((item / nb_pix) % nv);
        const int vs = (int)(item / ((int64_t)nb_pix * nv));
        const int v = a.v_base + vs;
        const int pix = a.vlist[k];
        const double vx = (double)(pix % nx) + 0.5, vy = (double)(pix / nx) + 0.5;
        const float *gam = a.gam_buf + ((size_t)vs * nv + (size_t)k) * P * kTexelMaps;
        // the rays (n = 0 for pixels past the detector), written by lane i
        int nmax = 0;
        {
            Ray r{};
            const int p = bnd * kBundle + lane;  // pixel 0 of a bundle always exists
            if (lane < kBundle) {
                const PixG pg = pixel_setup(a, v, p < P ? p : bnd * kBundle);
                r = make_ray(pg.dx, pg.dy, vx, vy, nx, ny, step);
                if (p >= P) r.n = 0;
                bx.rays[lane] = r;
            }
            int nm = lane < kBundle ? r.n : 0;
#pragma unroll
            for (int o = 1; o < kBundle; o <<= 1) nm = max(nm, __shfl_xor(nm, o, kWave));
            nmax = __builtin_amdgcn_readfirstlane(nm);
        }
        wave_lds_order();
        // the B operands of the emission products: gb[s][t] = gamma * step of ray
        // 4 s + lane / 16 for map 16 t + lane % 16
        double gb[kBundle / 4][4];
#pragma unroll
        for (int s = 0; s < kBundle / 4; s++) {
            const int p = bnd * kBundle + 4 * s + (lane >> 4);
#pragma unroll
            for (int t = 0; t < 4; t++)
                gb[s][t] = p < P ? (double)gam[(size_t)p * kTexelMaps + 16 * t + (lane & 15)] * step : 0.0;
        }
        const unsigned long long lower = (1ull << lane) - 1ull;
        for (int q0 = 0; q0 < nmax; q0 += kWave) {
            // the window's cells are bounded by those of each ray's first and last sample in
            // it (cell indices are monotone along a ray); corners add one, and only texels
            // inside the grid are kept.  Lane 2 i takes ray i's first sample in the window
            // and lane 2 i + 1 its last; the 16 lanes' cells are reduced by shuffles.
            int x0 = INT_MAX, x1 = INT_MIN, y0 = INT_MAX, y1 = INT_MIN;
            if (lane < 2 * kBundle) {
                const Ray rr = bx.rays[lane >> 1];
                if (q0 < rr.n) {
                    int ci, cj;
                    double ax, ay;
                    lane_sample(rr, step, (lane & 1) ? min(q0 + kWave - 1, rr.n - 1) : q0, ci, cj, ax, ay);
                    x0 = ci;
                    x1 = ci + 1;
                    y0 = cj;
                    y1 = cj + 1;
                }
            }
#pragma unroll
            for (int o = 1; o < 2 * kBundle; o <<= 1) {
                x0 = min(x0, __shfl_xor(x0, o, kWave));
                x1 = max(x1, __shfl_xor(x1, o, kWave));
                y0 = min(y0, __shfl_xor(y0, o, kWave));
                y1 = max(y1, __shfl_xor(y1, o, kWave));
            }
            x0 = max(__builtin_amdgcn_readfirstlane(x0), 0);
            y0 = max(__builtin_amdgcn_readfirstlane(y0), 0);
            x1 = min(__builtin_amdgcn_readfirstlane(x1), nx - 1);
            y1 = min(__builtin_amdgcn_readfirstlane(y1), ny - 1);
            const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
            const bool boxed = bw > 0 && bh > 0 && bw * bh <= a.box_texels;
#pragma unroll 1
            for (int i = 0; i < kBundle; i++) {
                const Ray ray = bx.rays[i];
                const int n = __builtin_amdgcn_readfirstlane(ray.n);
                if (q0 >= n) continue;  // uniform
                if (boxed) {
                    window_runs_to_box(ray, step, n, q0, nx, ny, lane, x0, y0, bw, bx, i);
                } else {
                    // a window too large for the box: this ray's texel list, one atomic per
                    // texel and map
                    SettledRun sr;
                    cell_runs_settled<kFlatBundle, 1>(ray, step, n, q0, 1, nx, ny, lane, L, sr);
                    int nblk;
                    const int nent = compact_texels<kListScatter>(L, lane, sr, nx, nblk);
                    const int p = bnd * kBundle + i;
                    const double g = (double)gam[(size_t)p * kTexelMaps + lane] * step;
                    if (g != 0.0)
                        for (int e = 0; e < nent; e++)
                            unsafeAtomicAdd(a.grad_tm + L.tex.off[e] + lane, g * L.tex.w[e]);
                }
                wave_lds_order();
            }
            if (!boxed) {
                // the run list used the box's weights: clear them for the next window
                float *wz = &bx.w[0][0];
                for (int q = lane; q < (int)(sizeof(WaveList) / sizeof(float)); q += kWave) wz[q] = 0.f;
                wave_lds_order();
                continue;
            }
            // the touched texels in box order
            int cnt = 0;
            for (int s0 = 0; s0 < bw * bh; s0 += kWave) {
                const int sl = s0 + lane;
                bool o = false;
                if (sl < bw * bh) {
                    const float4 *wp = reinterpret_cast<const float4 *>(bx.w[sl]);
#pragma unroll
                    for (int h = 0; h < kBundle / 4; h++) {
                        const float4 u = wp[h];
                        o = o || u.x != 0.f || u.y != 0.f || u.z != 0.f || u.w != 0.f;
                    }
                }
                const unsigned long long bal = __ballot(o);
                if (o) bx.list[cnt + __popcll(bal & lower)] = sl;
                cnt += __popcll(bal);
            }
            wave_lds_order();
            // sixteen texels at a time on the matrix cores: the (texel x ray) weights times
            // the (ray x map) gamma, v_mfma_f64_16x16x4f64 over four rays per step (lane l:
            // A[texel l % 16][ray 4 s + l / 16], B = gb[s][t], D[texel l / 16 + 4 r][map
            // 16 t + l % 16]); each lane then adds its four texels' values for its map.
            // Element offsets fit 32 bits: the host keeps the texel-major maps (at least
            // nx ny 64 floats) below 4 GiB.
            for (int f0 = 0; f0 < cnt; f0 += 16) {
                double wa[kBundle / 4];
                const int fa = f0 + (lane & 15);
#pragma unroll
                for (int s = 0; s < kBundle / 4; s++)
                    wa[s] = fa < cnt ? (double)bx.w[bx.list[fa]][4 * s + (lane >> 4)] : 0.0;
                unsigned base[4];
                bool live[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int fd = f0 + (lane >> 4) + 4 * r;
                    live[r] = fd < cnt;
                    const int sl = live[r] ? bx.list[fd] : 0;
                    base[r] = (unsigned)((x0 + sl % bw) + nx * (y0 + sl / bw)) * (unsigned)kTexelMaps +
                              (unsigned)(lane & 15);
                }
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    f64x4v acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s = 0; s < kBundle / 4; s++)
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[s], gb[s][t], acc, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (live[r] && acc[r] != 0.0)
                            unsafeAtomicAdd(a.grad_tm + (base[r] + 16u * (unsigned)t), acc[r]);
                }
            }
            wave_lds_order();
            for (int f = lane; f < cnt; f += kWave) {
                const int sl = bx.list[f];
#pragma unroll
                for (int i = 0; i < kBundle; i++) bx.w[sl][i] = 0.f;
            }
            wave_lds_order();
        }
    }
}

// Source legs of a slab's map VJP: one wave per (view of the slab, voxel) ray from the
// source, lane b adding gsrc_buf[(view - v_base) nv + voxel][b] (the sum of gamma over
// the pixels, written by the VJP launch) times the leg's weights to grad_tm.
__global__ __launch_bounds__(256) void coeff_src_scatter_kernel(CoeffArgs a) {
    __shared__ WaveList lists[4];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int64_t total = (int64_t)a.v_count * a.n_vox;
    for (int64_t item = (int64_t)blockIdx.x * 4 + wid; item < total;
         item += (int64_t)gridDim.x * 4) {  // uniform per wave
        const int k = (int)(item % a.n_vox), vs = (int)(item / a.n_vox);
        const int pix = a.vlist[k];
        const ViewConst vc = a.f.views[a.v_base + vs];
        const Ray ray = make_ray(vc.src_x, vc.src_y, (double)(pix % a.f.nx) + 0.5,
                                 (double)(pix / a.f.nx) + 0.5, a.f.nx, a.f.ny, a.f.ray_step);
        scatter_ray_tm(a, ray, a.gsrc_buf + item * kTexelMaps, lane, lists[wid]);
    }
}

// Lane r of sp lanes that share a ray takes a contiguous run of its samples.
__device__ __forceinline__ void lane_run(int n, int r, int sp, int &q0, int &q1) {
    const int per = (n + sp - 1) / sp;
    q0 = min(n, r * per);
    q1 = min(n, q0 + per);
}

// VJP: one workgroup per (chunk of KC voxels, group of views), looping over the
// views of the group and every pixel.  Per pixel it recomputes the detector legs
// and the forward terms of the chunk, and
//   * adds W(e) P(e) G_e(q_n) dq_n to the lane's FP64 totals h_vox(q_n) (the
//     coefficient gradient is basis . h), written per view group to hpart;
//   * when map gradients are wanted, forms S_e = sum_n G_e(q_n) dq_n F_vox(q_n)
//     from the same Gaussians (16-value transposed wave sums), the ray weights
//     gamma_b = -sum_e beta[b,e] W(e) P(e) S_e, and scatters them along the
//     detector legs; the source legs carry the per-view sums of gamma.
// Four waves per SIMD at 128 VGPRs, like coeff_forward_kernel (fan256, 12 views:
// coefficient VJP 1.11 -> 0.85 s): the trace's lists and the table/Gaussian arrays
// share their LDS space (live in different phases of each pixel), and beta is read
// from global memory, where its 16 KB stay in L1.
// MAPS false: the coefficient-only form, compiled without the S partials (100 against
// 188 B/lane of scratch; 12 views 0.303 -> 0.294 s).  MAPS true still tests a.grad_maps
// at run time: with the test constant the map form spilled more and measured 1% slower.
template <bool LDBUF, bool MAPS>
__global__ __launch_bounds__(NT, 4) void coeff_vjp_kernel(CoeffArgs a, int views_per_group) {
    __shared__ VoxC vox[KC];
    __shared__ float Ld[KC * LROW];
    __shared__ float Fq[KC * TQ];
    __shared__ double gam[KC * MAXB];
    __shared__ double gsrc[KC * MAXB];
    __shared__ double wt[64];
    __shared__ float Lsrc[KC * LROW];
    // table rows of kTS entries: the ballots' lane-indexed reads of one energy across
    // the chunk's voxels then fall into different banks
    constexpr int kTS = 65;
    __shared__ union {
        struct {
            WaveList cruns[NT / kWave];
            float2 part[NT];
        } trace;
        struct {
            float4 tw[KC * kTS];
            float tb[KC * kTS];
            double pw[KC * XROW];              // W(e) P(e)
            float sp_[NT / kWave][KC][16];     // per-wave partial S
        } gauss;
    } ov;
    WaveList *cruns = ov.trace.cruns;
    float2 *part = ov.trace.part;
    float4 *tw = ov.gauss.tw;
    float *tb = ov.gauss.tb;
    double *pw = ov.gauss.pw;
    auto &sp_ = ov.gauss.sp_;
    const float *bet = a.beta;
    const ForwardArgs &f = a.f;
    const int t = threadIdx.x, tq = t % TQ, h = t / TQ;
    const int lane = t & (kWave - 1), wid = t >> 6;
    const int V = f.n_views, E = f.n_energy, Q = f.n_q, B = a.n_basis, P = f.n_det;
    const size_t npix = (size_t)f.nx * f.ny;
    const int nv = a.nvox[0];
    const int k0 = blockIdx.x * KC;
    if (k0 >= nv) return;  // uniform per block
    const int nk = min(KC, nv - k0);
    // views [v_base, v_base + v_count) of a slab (v_count 0: every view)
    const int vend = a.v_count > 0 ? a.v_base + a.v_count : V;
    const int va = a.v_base + blockIdx.y * views_per_group, vb = min(vend, va + views_per_group);
    const bool maps = MAPS && a.grad_maps != nullptr;
    float qch = 1e30f, qcl = 0.f, dqc = 0.f;
    if (tq < Q) {
        const double q = f.q_coarse[tq];
        qch = (float)q;
        qcl = (float)(q - (double)qch);
        dqc = (float)f.q_coarse[Q + tq];
    }
    float lo = tq < Q ? qch : 1e30f, hi = tq < Q ? qch : -1e30f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, kWave));
        hi = fmaxf(hi, __shfl_xor(hi, o, kWave));
    }
    // F_vox(q_n) of the chunk, formed once per call (coeff_fq_kernel)
    for (int idx = t; idx < nk * TQ; idx += NT) Fq[idx] = a.fq_buf[(size_t)k0 * TQ + idx];
    double ht[KC];
#pragma unroll
    for (int j = 0; j < KC; j++) ht[j] = 0.0;
    for (int v = va; v < vb; v++) {
        for (int idx = t; idx < KC * MAXB; idx += NT) gsrc[idx] = 0.0;
        for (int idx = t; idx < nk * B; idx += NT) {
            const int j = idx / B, b = idx % B;
            Lsrc[j * LROW + b] = a.src_ints[((size_t)v * npix + k0 + j) * B + b];
        }
        for (int p = 0; p < P; p++) {
            const PixG pg = pixel_setup(a, v, p);
            if (t < nk) {
                VoxC vs;
                vs.k = k0 + t;
                vs.pix = a.vlist[k0 + t];
                vs.ix = vs.pix % f.nx;
                vs.iy = vs.pix / f.nx;
                geom_c(f, pg, vs.ix, vs.iy, vs);
                vox[t] = vs;
            }
            if (t < E) {
                // transpose of the detector-response window: the detector bins whose
                // window holds source bin t lie within 2*DNRG of it (the windows are
                // 2*DNRG+1 bins, shifted but never truncated), visited in ascending order
                double sw = 0.0;
                const int ed0 = max(0, t - 2 * f.dn), ed1 = min(E - 1, t + 2 * f.dn);
                for (int ed = ed0; ed <= ed1; ed++) {
                    int wlo, whi;
                    src_window(ed, E, f.dn, wlo, whi);
                    if (t >= wlo && t <= whi)
                        sw += (double)f.spectrum[t + (size_t)E * ed] *
                              (double)a.w[(size_t)v + (size_t)V * ((size_t)p + (size_t)P * ed)];
                }
                wt[t] = sw;
            }
            __syncthreads();
            if (LDBUF) {
                for (int idx = t; idx < nk * B; idx += NT) {
                    const int j = idx / B, b = idx % B;
                    Ld[j * LROW + b] = a.ld_buf[(((size_t)(v - a.v_base) * a.n_vox + (size_t)(k0 + j)) * P + p) *
                                                    kTexelMaps + b];
                }
            } else if (a.maps_tm) {
                trace_chunk_tm<kFlatVjp, kPassesVjp>(a, pg, vox, nk, Ld, cruns);
            } else {
                trace_chunk(a, pg, vox, nk, part, Ld);
            }
            __syncthreads();
            // the exponents go to pw, which each (voxel, energy) entry then overwrites
            // with its own W(e) P(e)
            chunk_mu(Lsrc, Ld, bet, B, E, nk, pw);
            __syncthreads();
            for (int idx = t; idx < nk * E; idx += NT) {
                const int j = idx / E, e = idx % E;
                float4 tv;
                float band;
                double Pe;
                table_entry(a, vox[j], pw[j * XROW + e], e, tv, band, Pe);
                pw[j * XROW + e] = wt[e] * Pe;
                tv.w = (float)(wt[e] * Pe);
                tw[j * kTS + e] = tv;
                tb[j * kTS + e] = band;
            }
            __syncthreads();
            if (!MAPS) {
                // coefficients only: one ballot per voxel gives its energy slots whose
                // band reaches this wave's bins (lane i: slot i); the wave then adds
                // those terms in ascending slot order, two per step with both table
                // reads issued before either is used
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    if (j >= nk) break;
                    bool in = false;
                    if (lane < EPT && h + G * lane < E) {
                        const int e = h + G * lane;
                        const float x = tw[j * kTS + e].x, band = tb[j * kTS + e];
                        in = x + band >= lo && x - band <= hi;
                    }
                    uint32_t m = (uint32_t)__ballot(in);
                    auto term = [&](int i) {
                        const float4 tv = tw[j * kTS + h + G * i];
                        const float d = (tv.x - qch) + (tv.y - qcl);
                        return make_float2(tv.w, __builtin_amdgcn_exp2f(d * d * tv.z));
                    };
                    float hp = 0.f;
                    while (m & (m - 1)) {
                        const int i0 = __builtin_ctz(m);
                        m &= m - 1;
                        const int i1 = __builtin_ctz(m);
                        m &= m - 1;
                        const float2 u0 = term(i0), u1 = term(i1);
                        hp = fmaf(u0.x, u0.y, hp);
                        hp = fmaf(u1.x, u1.y, hp);
                    }
                    if (m) {
                        const float2 u = term(__builtin_ctz(m));
                        hp = fmaf(u.x, u.y, hp);
                    }
                    ht[j] += (double)(hp * dqc);
                }
            } else {
                // per energy slot, one ballot gives the chunk's voxels whose band reaches
                // this wave's bins (the range and the energy are wave-uniform); the voxel
                // loop then reads only those entries
                uint32_t inb[16];
    #pragma unroll
                for (int i = 0; i < 16; i++) {
                    const int e = h + G * i;
                    bool in = false;
                    if (i < EPT && e < E && lane < nk) {
                        const float x = tw[lane * kTS + e].x, band = tb[lane * kTS + e];
                        in = x + band >= lo && x - band <= hi;
                    }
                    inb[i] = (uint32_t)__ballot(in);
                }
                // h_vox(q_n) += W(e) P(e) G_e(q_n) dq_n; S_e partials over this wave's bins
    #pragma unroll
                for (int j = 0; j < KC; j++) {
                    if (j >= nk) break;
                    const float fq = Fq[j * TQ + (tq < Q ? tq : 0)] * dqc;
                    float hp = 0.f;
                    float sv[16];
                        // two slots' entries read before either is tested (every slot's
                    // index lies inside the table row, G * 16 <= kTS; a slot past E
                    // reads an unused entry and its bit is clear)
                    static_assert(G * 16 <= kTS, "slot reads stay in the row");
#pragma unroll
                    for (int i = 0; i < 16; i += 2) {
                        const float4 tp[2] = {tw[j * kTS + h + G * i], tw[j * kTS + h + G * (i + 1)]};
#pragma unroll
                        for (int u = 0; u < 2; u++) {
                            sv[i + u] = 0.f;
                            if ((inb[i + u] >> j) & 1u) {
                                const float4 tv = tp[u];
                                const float d = (tv.x - qch) + (tv.y - qcl);
                                const float g = __builtin_amdgcn_exp2f(d * d * tv.z);
                                hp = fmaf(tv.w, g, hp);
                                sv[i + u] = g * fq;
                            }
                        }
                    }
                    ht[j] += (double)(hp * dqc);
                    if (maps) {
                        const float st = transpose_sum16(sv, lane);
                        if ((lane & 3) == 0) sp_[wid][j][(lane >> 2) & 15] = st;
                    }
                }
            }
            __syncthreads();
            if (maps) {
                // gamma_b = -sum_e beta[b,e] W(e) P(e) S_e: first W(e) P(e) S_e into pw
                // (S_e: the partials of the waves of energy e's group h, waves
                // h*(TQ/64) .. h*(TQ/64) + TQ/64 - 1, added in wave order), then the
                // contraction over the energies on the matrix cores
                for (int idx = t; idx < nk * E; idx += NT) {
                    const int j = idx / E, e = idx % E;
                    const int hh = e % G, i = e / G;
                    double se = 0.0;
                    for (int w = hh * (TQ / kWave); w < (hh + 1) * (TQ / kWave); w++)
                        se += (double)sp_[w][j][i];
                    pw[j * XROW + e] *= se;
                }
                __syncthreads();
                chunk_gamma(pw, bet, B, E, nk, gam, gsrc);
                __syncthreads();
                // scatter gamma along the detector legs, or hand it to the bundled
                // scatter: gam_buf[view in slab][voxel][pixel][b]
                if (LDBUF) {
                    for (int idx = t; idx < nk * kTexelMaps; idx += NT) {
                        const int j = idx / kTexelMaps, b = idx % kTexelMaps;
                        a.gam_buf[(((size_t)(v - a.v_base) * a.n_vox + (size_t)(k0 + j)) * P + p) *
                                      kTexelMaps + b] = b < B ? (float)gam[j * MAXB + b] : 0.f;
                    }
                } else if (a.grad_tm) {
                    for (int j = wid; j < nk; j += NT / kWave) {
                        const Ray ray = make_ray(pg.dx, pg.dy, (double)vox[j].ix + 0.5,
                                                 (double)vox[j].iy + 0.5, f.nx, f.ny, f.ray_step);
                        scatter_ray_tm(a, ray, gam + j * MAXB, lane, cruns[wid]);
                    }
                } else {
                    const int j = t % KC, r = t / KC;
                    if (j < nk) {
                        const Ray ray = make_ray(pg.dx, pg.dy, (double)vox[j].ix + 0.5,
                                                 (double)vox[j].iy + 0.5, f.nx, f.ny, f.ray_step);
                        int q0, q1;
                        lane_run(ray.n, r, NT / KC, q0, q1);
                        scatter_ray(a, ray, gam + j * MAXB, q0, q1);
                    }
                }
            }
            __syncthreads();
        }
        // source legs of view v carry the sum of gamma over the pixels; the slab form
        // hands the sums to coeff_src_scatter_kernel, so this launch holds no ray walk
        // (and none of its registers)
        if (maps && LDBUF) {
            for (int idx = t; idx < nk * kTexelMaps; idx += NT) {
                const int j = idx / kTexelMaps, b = idx % kTexelMaps;
                a.gsrc_buf[((size_t)(v - a.v_base) * a.n_vox + (size_t)(k0 + j)) * kTexelMaps + b] =
                    b < B ? gsrc[j * MAXB + b] : 0.0;
            }
        } else if (maps) {
            const ViewConst vc = f.views[v];
            if (a.grad_tm) {
                for (int j = wid; j < nk; j += NT / kWave) {
                    const int pix = a.vlist[k0 + j];
                    const Ray ray = make_ray(vc.src_x, vc.src_y, (double)(pix % f.nx) + 0.5,
                                             (double)(pix / f.nx) + 0.5, f.nx, f.ny, f.ray_step);
                    scatter_ray_tm(a, ray, gsrc + j * MAXB, lane, cruns[wid]);
                }
            } else {
                const int j = t % KC, r = t / KC;
                if (j < nk) {
                    const int pix = a.vlist[k0 + j];
                    const Ray ray = make_ray(vc.src_x, vc.src_y, (double)(pix % f.nx) + 0.5,
                                             (double)(pix / f.nx) + 0.5, f.nx, f.ny, f.ray_step);
                    int q0, q1;
                    lane_run(ray.n, r, NT / KC, q0, q1);
                    scatter_ray(a, ray, gsrc + j * MAXB, q0, q1);
                }
            }
        }
        __syncthreads();
    }
    // this view group's h: add the G energy groups of each bin in order (the shared
    // space of the trace and the table is free now and holds KC*TQ doubles), then
    // store; coeff_grad_kernel adds the view groups and contracts with the basis
    static_assert(sizeof(ov) >= sizeof(double) * KC * TQ, "LDS reuse");
    double *hred = reinterpret_cast<double *>(&ov);
    for (int hh = 0; hh < G; hh++) {
        if (h == hh) {
#pragma unroll
            for (int j = 0; j < KC; j++) {
                if (j < nk) {
                    double *slot = &hred[j * TQ + tq];
                    *slot = (hh == 0 ? 0.0 : *slot) + ht[j];
                }
            }
        }
        __syncthreads();
    }
    if (a.grad_coef)
        for (int idx = t; idx < nk * TQ; idx += NT)
            a.hpart[((size_t)blockIdx.y * a.nvox_cap + k0) * TQ + idx] =
                (a.hpart_add ? a.hpart[((size_t)blockIdx.y * a.nvox_cap + k0) * TQ + idx] : 0.0) +
                hred[idx];
}

// One workgroup of TQ lanes per voxel: h(q_n) = sum of the view groups' partials in
// group order, then grad_coef[vox, k] = sum_n basis[k, n] h(q_n).
__global__ __launch_bounds__(TQ) void coeff_grad_kernel(CoeffArgs a, int n_groups) {
    __shared__ double hv[TQ];
    const int k = blockIdx.x, n = threadIdx.x;
    if (k >= a.nvox[0]) return;  // uniform per block
    const int Q = a.f.n_q;
    double s = 0.0;
    for (int g = 0; g < n_groups; g++) s += a.hpart[((size_t)g * a.nvox_cap + k) * TQ + n];
    hv[n] = n < Q ? s : 0.0;
    __syncthreads();
    if (n < a.n_coef) {
        double r = 0.0;
        for (int i = 0; i < Q; i++) r += (double)a.basis[n * Q + i] * hv[i];
        a.grad_coef[(size_t)a.vlist[k] * a.n_coef + n] = r;
    }
}

// ---- CT transmission channel ---------------------------------------------------
// The transmitted signal of the simultaneous CT + XRD scanner (README.md:14).  The
// reference only names it, in the commented-out launches of integrate_src2det and
// polychromatic_src2det (Forward_XRD.cu:271-276), whose kernels it does not ship:
// for view v and detector column c the primary ray runs from the source to the
// column's point on the detector; its part inside the grid (Liang-Barsky clipping,
// FP64, no contraction) is sampled like attenuator.cu:44-69 from the source-side
// entry, and
//   T[v + V (c + COL ed)] = sum_{e in win(ed)} RF[e,ed] exp(-sum_b beta[b,e] L_b).

// Clips (sx,sy)->(dx,dy) to [0,nx]x[0,ny] and samples the chord from its entry.
__device__ __forceinline__ bool ct_chord(double sx, double sy, double dx, double dy, int nx, int ny,
                                         double step, Ray &r) {
#pragma clang fp contract(off)
    const double ddx = dx - sx, ddy = dy - sy;
    const double p[4] = {-ddx, ddx, -ddy, ddy};
    const double q[4] = {sx - 0.0, (double)nx - sx, sy - 0.0, (double)ny - sy};
    double t0 = 0.0, t1 = 1.0;
    bool miss = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (p[k] == 0.0) {
            if (q[k] < 0.0) miss = true;
            continue;
        }
        const double rr = q[k] / p[k];
        if (p[k] < 0.0) {
            if (rr > t0) t0 = rr;
        } else {
            if (rr < t1) t1 = rr;
        }
    }
    r.n = 0;
    r.side = -1;
    r.ex = r.ey = r.ux = r.uy = 0.0;
    if (miss || !(t0 < t1)) return false;
    r.ex = sx + ddx * t0;
    r.ey = sy + ddy * t0;
    const double xx = sx + ddx * t1, xy = sy + ddy * t1;
    const double lx = xx - r.ex, ly = xy - r.ey;
    const double L = sqrt(lx * lx + ly * ly);
    r.ux = lx / L;
    r.uy = ly / L;
    if (!(L >= 0.0)) return false;
    int n = (int)floor(L / step) + 1;
    while ((double)n * step <= L) n++;
    while (n > 0 && (double)(n - 1) * step > L) n--;
    r.n = n;
    r.side = 0;
    return true;
}

// Per ray (one wave): lane b's line integral L_b, then lane e's exp(-mu_e).
__device__ __forceinline__ double ct_ray_p(const CoeffArgs &a, int v, int c, int lane, Ray &ray,
                                           bool &hit, WaveList &runs) {
    const int E = a.f.n_energy, B = a.n_basis;
    const PixG pg = pixel_setup(a, v, c);
    hit = ct_chord(pg.sx, pg.sy, pg.dx, pg.dy, a.f.nx, a.f.ny, a.f.ray_step, ray);
    double Lb = 0.0;
    if (hit) Lb = wave_trace_tm<kFlatForward, kPassesForward, true>(a.maps_tm, a.f.nx, a.f.ny, a.f.ray_step, ray, lane, runs) *
                  a.f.ray_step;
    const int e = min(lane, E - 1);
    double mu = 0.0;
    for (int b = 0; b < B; b++) mu += (double)a.beta[b * E + e] * read_lane(Lb, b);
    return exp(-mu);
}

// One wave per (view, column) ray; 4 rays per workgroup.
__global__ __launch_bounds__(256) void ct_forward_kernel(CoeffArgs a) {
    __shared__ WaveList cruns[4];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int V = a.f.n_views, COL = a.f.n_cols, E = a.f.n_energy;
    const int ray_id = blockIdx.x * 4 + wid;
    if (ray_id >= V * COL) return;  // uniform per wave
    const int v = ray_id % V, c = ray_id / V;
    Ray ray;
    bool hit;
    const double P = ct_ray_p(a, v, c, lane, ray, hit, cruns[wid]);
    // lane ed: sum over its source-energy window (all windows have 2 dn + 1 bins)
    const int ed = min(lane, E - 1);
    int lo, hi;
    src_window(ed, E, a.f.dn, lo, hi);
    double t = 0.0;
    for (int k = 0; k <= 2 * a.f.dn; k++) {
        const int e = lo + k;
        const double pe = __shfl(P, e, kWave);
        t += (double)a.f.spectrum[e + (size_t)E * ed] * pe;
    }
    if (lane < E) a.out_counts[(size_t)v + (size_t)V * ((size_t)c + (size_t)COL * lane)] = (float)t;
}

// VJP of <w, T> with respect to the maps: per ray gamma_b = -sum_e beta[b,e] P_e
// sum_{ed: e in win(ed)} RF[e,ed] w[v,c,ed], scattered along the chord.
__global__ __launch_bounds__(256) void ct_vjp_kernel(CoeffArgs a) {
    __shared__ double gam[4][kWave];
    __shared__ WaveList cruns[4];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x >> 6;
    const int V = a.f.n_views, COL = a.f.n_cols, E = a.f.n_energy, B = a.n_basis;
    const int ray_id = blockIdx.x * 4 + wid;
    if (ray_id >= V * COL) return;  // uniform per wave
    const int v = ray_id % V, c = ray_id / V;
    Ray ray;
    bool hit;
    const double P = ct_ray_p(a, v, c, lane, ray, hit, cruns[wid]);
    if (!hit) return;
    double wt = 0.0;
    if (lane < E)
        for (int ed = 0; ed < E; ed++) {
            int lo, hi;
            src_window(ed, E, a.f.dn, lo, hi);
            if (lane >= lo && lane <= hi)
                wt += (double)a.f.spectrum[lane + (size_t)E * ed] *
                      (double)a.w[(size_t)v + (size_t)V * ((size_t)c + (size_t)COL * ed)];
        }
    const double ge = lane < E ? wt * P : 0.0;
    const int b = min(lane, B - 1);
    double gb = 0.0;
    for (int e = 0; e < E; e++) gb -= (double)a.beta[b * E + e] * read_lane(ge, e);
    gam[wid][lane] = lane < B ? gb : 0.0;
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
    scatter_ray_tm(a, ray, gam[wid], lane, cruns[wid]);
}

__global__ void zero_f64_kernel(double *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

hipError_t launch_zero_f64(double *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(zero_f64_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Poisson NLL sum(f - y log f), f = counts + background, and its gradient 1 - y/f.
__global__ void poisson_nll_kernel(int64_t n, const float *cnt, const float *bg, const float *y,
                                   float *grad, double *partial) {
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double f = (double)cnt[i] + (bg ? (double)bg[i] : 0.0);
        const double yi = (double)y[i];
        acc += f - (yi > 0.0 ? yi * log(f) : 0.0);
        if (grad) grad[i] = (float)(1.0 - yi / f);
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_poisson_nll(int64_t n, const float *cnt, const float *bg, const float *y,
                              float *grad, double *partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(poisson_nll_kernel, dim3(nblocks), dim3(256), 0, s, n, cnt, bg, y, grad,
                       partial);
    return hipGetLastError();
}

// F_vox(q_n) = sum_k coef[vox, k] basis[k, n] of every voxel of the list, once per call
// (one FP32 fma chain over the coefficients in order), for the forward and VJP kernels,
// which would otherwise form it again in every (view, pixel) or view-group workgroup.
__global__ __launch_bounds__(256) void coeff_fq_kernel(CoeffArgs a) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (int64_t)a.nvox_cap * TQ) return;
    const int k = (int)(idx / TQ), n = (int)(idx % TQ);
    if (k >= a.nvox[0]) return;
    const int Q = a.f.n_q;
    float s = 0.f;
    if (n < Q) {
        const float *cf = a.coef + (size_t)a.vlist[k] * a.n_coef;
        for (int kk = 0; kk < a.n_coef; kk++) s = fmaf(cf[kk], a.basis[kk * Q + n], s);
    }
    a.fq_buf[idx] = s;
}

hipError_t launch_coeff_prepare(const CoeffArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(coeff_list_kernel, dim3(1), dim3(NT), 0, s, a);
    hipLaunchKernelGGL(coeff_fq_kernel, dim3((unsigned)(((int64_t)a.nvox_cap * TQ + 255) / 256)),
                       dim3(256), 0, s, a);
    const size_t n = (size_t)(a.f.nx + 2) * (a.f.ny + 2) * a.n_basis;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(coeff_pad_kernel, dim3(blocks), dim3(256), 0, s, a);
    if (a.maps_tm) hipLaunchKernelGGL(coeff_pad_tm_kernel, dim3(8192), dim3(256), 0, s, a);
    if (a.maps_tm) {
        dim3 grid((a.f.nx * a.f.ny + 3) / 4, a.f.n_views);
        hipLaunchKernelGGL(coeff_src_tm_kernel, grid, dim3(256), 0, s, a);
    } else {
        dim3 grid((a.f.nx * a.f.ny + 255) / 256, a.f.n_views);
        hipLaunchKernelGGL(coeff_src_kernel, grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_coeff_forward(const CoeffArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(coeff_forward_kernel<false>, dim3(a.f.n_views * a.f.n_det), dim3(NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_coeff_forward_slabs(CoeffArgs a, int slab_views, hipStream_t s) {
    const int V = a.f.n_views, P = a.f.n_det;
    const int most = std::max(1, std::min(V, slab_views));
    const int n_slabs = (V + most - 1) / most;
    const int slab = (V + n_slabs - 1) / n_slabs;
    for (int v0 = 0; v0 < V; v0 += slab) {
        a.v_base = v0;
        a.v_count = std::min(slab, V - v0);
        launch_trace_slab(a, s);
        hipLaunchKernelGGL(coeff_forward_kernel<true>, dim3(a.v_count * P), dim3(NT), 0, s, a);
    }
    return hipGetLastError();
}

// View groups of the VJP grid: up to 32 groups of consecutive views, fewer when
// the per-group FP64 partials of h would pass 4 GiB.
int coeff_vjp_groups(int n_views, int npix) {
    const size_t per_group = (size_t)npix * TQ * sizeof(double);
    size_t ng = ((size_t)4 << 30) / per_group;
    if (ng > 32) ng = 32;
    if (ng > (size_t)n_views) ng = (size_t)n_views;
    return ng < 1 ? 1 : (int)ng;
}

bool coeff_texel_major(int n_basis) { return n_basis > 4; }

size_t coeff_texel_floats(int nx, int ny) { return (size_t)(nx + 2) * (ny + 2) * kTexelMaps; }

size_t coeff_hpart_doubles(int n_views, int npix) {
    return (size_t)coeff_vjp_groups(n_views, npix) * (size_t)npix * TQ;
}

hipError_t launch_coeff_vjp(const CoeffArgs &a, hipStream_t s) {
    const int ng = coeff_vjp_groups(a.f.n_views, a.nvox_cap);
    const int vpg = (a.f.n_views + ng - 1) / ng;
    const int nb = (a.nvox_cap + KC - 1) / KC;
    if (nb > 0) {
        if (a.grad_maps)
            hipLaunchKernelGGL((coeff_vjp_kernel<false, true>), dim3(nb, ng), dim3(NT), 0, s, a, vpg);
        else
            hipLaunchKernelGGL((coeff_vjp_kernel<false, false>), dim3(nb, ng), dim3(NT), 0, s, a, vpg);
        if (a.grad_coef)
            hipLaunchKernelGGL(coeff_grad_kernel, dim3(a.nvox_cap), dim3(TQ), 0, s, a, ng);
        if (a.grad_maps && a.grad_tm)
            hipLaunchKernelGGL(coeff_untranspose_kernel, dim3(4096), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

size_t coeff_gam_floats_per_view(int n_vox, int n_det) {
    return (size_t)n_vox * (size_t)n_det * kTexelMaps;
}

hipError_t launch_coeff_vjp_slabs(CoeffArgs a, int slab_views, hipStream_t s) {
    const int V = a.f.n_views, P = a.f.n_det;
    const int most = std::max(1, std::min(V, slab_views));
    // equal slabs (the last one no thinner than the rest), each within the buffers
    const int n_slabs = (V + most - 1) / most;
    const int slab = (V + n_slabs - 1) / n_slabs;
    const int ng = coeff_vjp_groups(slab, a.nvox_cap);
    const int nb = (a.nvox_cap + KC - 1) / KC;
    const int nbp = (P + kBundle - 1) / kBundle;
    if (nb <= 0 || a.n_vox <= 0) return hipSuccess;
    // the slab launch hands every leg's gamma on: map gradients need both buffers
    if (a.grad_maps && (!a.gam_buf || !a.gsrc_buf || !a.grad_tm)) return hipErrorInvalidValue;
    // XRD_VJP_BOX_TEXELS=k accumulates only windows of at most k texels in the box, so
    // that the fallback for larger windows runs on a small problem (for testing)
    a.box_texels = kBoxN;
    if (const char *cap = std::getenv("XRD_VJP_BOX_TEXELS")) {
        const long k = std::strtol(cap, nullptr, 10);
        if (k >= 0 && k < kBoxN) a.box_texels = (int)k;
    }
    for (int v0 = 0; v0 < V; v0 += slab) {
        a.v_base = v0;
        a.v_count = std::min(slab, V - v0);
        a.hpart_add = v0 > 0;
        launch_trace_slab(a, s);
        const int ngs = std::min(ng, a.v_count);
        const int vpg = (a.v_count + ngs - 1) / ngs;
        if (a.grad_maps)
            hipLaunchKernelGGL((coeff_vjp_kernel<true, true>), dim3(nb, ngs), dim3(NT), 0, s, a, vpg);
        else
            hipLaunchKernelGGL((coeff_vjp_kernel<true, false>), dim3(nb, ngs), dim3(NT), 0, s, a, vpg);
        if (a.gam_buf) {
            const int64_t items = (int64_t)a.v_count * a.n_vox * nbp;
            if (items > 0)
                hipLaunchKernelGGL(coeff_scatter_bundle_kernel, dim3(ray_blocks(items, kBundleWaves)),
                                   dim3(kBundleWaves * kWave), 0, s, a, nbp);
            const int64_t src_rays = (int64_t)a.v_count * a.n_vox;
            if (src_rays > 0)
                hipLaunchKernelGGL(coeff_src_scatter_kernel, dim3(ray_blocks(src_rays, 4)),
                                   dim3(256), 0, s, a);
        }
    }
    if (a.grad_coef) hipLaunchKernelGGL(coeff_grad_kernel, dim3(a.nvox_cap), dim3(TQ), 0, s, a, ng);
    if (a.grad_maps && a.grad_tm)
        hipLaunchKernelGGL(coeff_untranspose_kernel, dim3(4096), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ct(const CoeffArgs &a, bool vjp, hipStream_t s) {
    const int rays = a.f.n_views * a.f.n_cols;
    const dim3 grid((rays + 3) / 4);
    if (vjp) {
        hipLaunchKernelGGL(ct_vjp_kernel, grid, dim3(256), 0, s, a);
        hipLaunchKernelGGL(coeff_untranspose_kernel, dim3(4096), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(ct_forward_kernel, grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_ct_prepare(const CoeffArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(coeff_pad_tm_kernel, dim3(8192), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace xrd

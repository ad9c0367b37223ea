// xrd_internal.hpp -- argument blocks and launchers shared by the kernel files
// and the C-ABI layer.  Not part of the public interface (include/xrd_tomo.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xrd_tomo.h"

namespace xrd {

// One entry of the non-air voxel list (reference loop order within a material).
struct VoxelEntry {
    int32_t pix;     // ix + nx*iy
    int32_t mat;     // 0-based material
    float a1c, a2c;  // (a1,a2) used by the Compton term
};

// bfloat16 bits (the upper half of an IEEE binary32) as FP32, exactly.
__host__ __device__ inline float bf16_bits_to_float(uint16_t b) {
    union {
        uint32_t u;
        float f;
    } v;
    v.u = (uint32_t)b << 16;
    return v.f;
}

// A per-voxel table (a1 or a2 map, one slice or a volume) as the preparation kernels
// load it: FP32 values, or bfloat16 values times an FP32 scale chosen by the voxel's
// segmentation id (ids outside [0, n_ids) use scale[0], the air scale).
struct MapSrc {
    const float *f32 = nullptr;
    const uint16_t *b16 = nullptr;
    const float *scale = nullptr;
    const int32_t *seg = nullptr;  // segmentation of the same voxels as b16
    int n_ids = 0;
    __device__ __forceinline__ float at(size_t i) const {
        if (f32) return f32[i];
        int id = seg[i];
        if (id < 0 || id >= n_ids) id = 0;
        return bf16_bits_to_float(b16[i]) * scale[id];
    }
};

// Per-energy-bin constants, host-computed in FP64.
struct EnergyConst {
    double es1, es2, esm;  // source-bin edges and centre, keV
    double ein;            // esm / mc2
    double inv_f1, inv_f2; // 1/e^3 and 1/f2(e) at ein
    double wcoef;          // 2*esm/hbarc : wvt = wcoef * sin(theta/2)
    double k2sq;           // (esm/hbarc)^2
};

// Per-view constants, host-computed in FP64 with the host libm.
struct ViewConst {
    double sp, cp;        // sin/cos of the view angle
    double src_x, src_y;  // source position
};

// Largest detector-response half width (DNRG) the forward kernels support: the
// window of 2 DNRG + 1 source bins is unrolled.
constexpr int kMaxEnergyHalfwidth = 7;
static_assert(2 * kMaxEnergyHalfwidth + 1 <= 15, "the Compton window search takes four steps");

struct ForwardArgs {
    // sizes
    int nx, ny, n_views, n_cols, n_rows, n_det;
    int n_energy, dn, n_q, n_qf, n_mat, lut_nq, lut_na;
    unsigned flags;
    // geometry scalars
    double center_x, center_y, r_src, r_det, col_pitch, row_pitch, det_z0;
    double sta, cta;          // sin/cos target angle
    double tan_w, tan_w2;     // tan(wedge), tan(wedge/2)
    double focal_spot, voxel_size, adet;
    // discretisation scalars
    double e_min, e_width, hbarc, mc2, kn0, kn1, kn2, ray_step;
    double dq_fine, lut_q_min, lut_a_min, dq_lut, da_lut;
    // inputs
    const float *spectrum;      // [E x E]
    const float *ifine;         // [QF x M] FP32, or NULL when ifine_b16 is set
    const uint16_t *ifine_b16;  // [QF x M] bfloat16, times ifine_scale[mat]
    const float *ifine_scale;   // [M]
    const uint8_t *s2u8;        // [NQ x NA] quantised Compton table
    const float2 *mu_padded;    // [(nx+2) x (ny+2)] (a1,a2), zero border
    const VoxelEntry *voxels;   // material-major voxel list
    const int *mat_offsets;     // [M+1] list offsets per material
    const float2 *src_int;      // [V x nx*ny] source-leg integrals (a1,a2)
    const ViewConst *views;     // [V]
    const EnergyConst *energies;// [E]
    const double *q_coarse;     // [2*Q]: centres then widths (FP64)
    // outputs
    float *out_n, *out_c, *out_a;
    // per-(view, pixel, voxel) records of the current batch of views [v0, v0 + nvb)
    // (pixel_voxel_kernel), index (vb * n_det + pixel) * nvox + voxel
    int v0, nvb, nvox;
    double *pv_s, *pv_cth;  // sin(theta/2), cos(theta)
    float4 *pv_geo;         // resolution terms As, Bs and geometry weights gR, gp
    float2 *pv_leg;         // detector-leg integrals (a1, a2)
    // volume attenuation (nz > 1): both legs are 3-D segments through the padded
    // [(nx+2) x (ny+2) x (nz+2)] (a1,a2) volume; the scattering slice's centre is at
    // height z_slice.  nz <= 1: the 2-D map, constant in z.
    const float2 *mu_vol;
    int nz;
    double z_slice;
    // voxels per chunk of forward_kernel (its per-(voxel, energy) LDS table); 0 = 16
    int kc;
};

struct PrepArgs {
    int nx, ny, n_mat;
    unsigned flags;
    const int32_t *seg;
    MapSrc a1, a2;          // the slice's maps [nx x ny], decoded on load
    const float *s2_table;
    int s2_count;
    uint8_t *s2u8;
    float2 *mu_padded;
    float *cpx1, *cpx2;     // per-pixel Compton (a1,a2) images
    int *ref_list;          // scratch [nx*ny]
    VoxelEntry *voxels;
    int *mat_offsets;
};

hipError_t launch_prepare(const PrepArgs &a, hipStream_t s);
// (a1, a2) volume [nx x ny x nz] into the padded volume with a zero border.
hipError_t launch_prepare_volume(const MapSrc &a1, const MapSrc &a2, int nx, int ny, int nz,
                                 float2 *mu_vol, hipStream_t s);
hipError_t launch_source_integrals(const ForwardArgs &a, float2 *src_int, hipStream_t s);
hipError_t launch_pixel_voxel(const ForwardArgs &a, hipStream_t s);
// Largest voxel chunk of forward_kernel: its ballots cover one wave of voxels, and
// 32 voxels' tables still fit a workgroup's LDS.
constexpr int kMaxVoxelsPerChunk = 32;
// Bytes of one (view, pixel, voxel) record, and the workspace the records of one
// batch of views may take.
constexpr size_t kPixelVoxelRecordBytes = 2 * sizeof(double) + sizeof(float4) + sizeof(float2);
constexpr size_t kPixelVoxelBudget = size_t(4) << 30;
// Returns hipErrorInvalidValue when no kernel instantiation fits the sizes.
hipError_t launch_forward(const ForwardArgs &a, hipStream_t s);
hipError_t launch_trace_rays(int nx, int ny, double step, int64_t n, const double *seg,
                             int32_t *side, double *entry, int32_t *count, int32_t max_samples,
                             int32_t *texels, int32_t n_basis, const float *maps, float *ints,
                             hipStream_t s);

// ---- differentiable coefficient model ----------------------------------------------
struct CoeffArgs {
    ForwardArgs f;           // geometry, discretisation, host tables, spectrum
    const int32_t *seg;      // [nx*ny]; voxels with seg > 0 scatter
    const float *coef;       // [nx*ny x K] (k fastest)
    int n_coef;              // K <= 32
    const float *basis;      // [K x NWT]
    int n_basis;             // B <= 64
    const float *maps;       // [B x nx*ny]
    const float *beta;       // [B x NRG]
    float *maps_padded;      // [B x (nx+2)(ny+2)] workspace
    float *src_ints;         // [V x nx*ny x B] workspace
    int *vlist;              // [nx*ny] workspace
    int *nvox;               // [1] workspace
    const float *w;          // upstream gradient [rows]
    float *out_counts;       // [rows]
    double *grad_coef;       // [nx*ny x K] or NULL
    double *grad_maps;       // [B x nx*ny] or NULL (zeroed by the caller)
    float *maps_tm;          // [(nx+2)(ny+2) x 64] texel-major maps, or NULL (B <= 4)
    double *grad_tm;         // [nx*ny x 64] map gradient while accumulating, or NULL
    int nvox_cap;            // nx*ny: bound of the voxel list
    double *hpart;           // coeff_hpart_doubles(V, nx*ny) when grad_coef is set
    // map VJP in view slabs: the VJP launch covers views [v_base, v_base + v_count),
    // adds into hpart when hpart_add, and writes the detector legs' gamma to gam_buf
    // ([v_count x nv x n_det x 64] FP32, voxel-list index nv = n_vox) for the
    // bundled scatter instead of scattering them itself (NULL: scatter in place)
    int v_base, v_count, hpart_add, n_vox;
    float *gam_buf;
    // with gam_buf: the source legs' gamma (the sum over the pixels) of the slab's
    // (view, voxel) rays, [v_count x nv x 64] FP64, written by the VJP launch and
    // scattered by coeff_src_scatter_kernel
    double *gsrc_buf;
    // detector-leg line integrals of the slab's rays ([v_count x nv x n_det x 64]
    // FP32, the layout of gam_buf), written by coeff_trace_kernel and read by the VJP
    // launch instead of tracing the legs itself (NULL: the VJP launch traces)
    float *ld_buf;
    // first ray of a coeff_trace_kernel launch within its slab (launches are capped)
    int64_t item_base;
    // largest window box (texels) the bundled scatter accumulates in LDS; larger
    // windows scatter each ray's texel list directly (set by the slab launcher)
    int box_texels;
    // F_vox(q_n) = sum_k coef[vox,k] basis[k,n] per voxel-list index, [nx*ny x 128]
    // (zero for n >= n_q), formed once per call by coeff_fq_kernel for the forward
    float *fq_buf;
};

// Columns of CoeffArgs::fq_buf (the coefficient kernels' q-lane count).
constexpr int kCoeffFqCols = 128;
hipError_t launch_coeff_prepare(const CoeffArgs &a, hipStream_t s);
hipError_t launch_coeff_forward(const CoeffArgs &a, hipStream_t s);
// Forward in view slabs of slab_views views: per slab the detector legs are traced
// into a.ld_buf (coeff_gam_floats_per_view floats per view) and the forward launch
// reads them; a.n_vox set.
hipError_t launch_coeff_forward_slabs(CoeffArgs a, int slab_views, hipStream_t s);
hipError_t launch_coeff_vjp(const CoeffArgs &a, hipStream_t s);
// VJP in view slabs of slab_views views: per slab the detector legs are traced into
// a.ld_buf, the VJP launch reads them (and, with a.gam_buf, writes gamma there for the
// scatter of the detector legs in bundles of adjacent pixels); a.n_vox set, both
// buffers hold coeff_gam_floats_per_view floats per view of the slab.
hipError_t launch_coeff_vjp_slabs(CoeffArgs a, int slab_views, hipStream_t s);
size_t coeff_gam_floats_per_view(int n_vox, int n_det);
size_t coeff_hpart_doubles(int n_views, int npix);
bool coeff_texel_major(int n_basis);
hipError_t launch_ct_prepare(const CoeffArgs &a, hipStream_t s);
hipError_t launch_ct(const CoeffArgs &a, bool vjp, hipStream_t s);
size_t coeff_texel_floats(int nx, int ny);
hipError_t launch_zero_f64(double *p, size_t n, hipStream_t s);
hipError_t launch_poisson_nll(int64_t n, const float *cnt, const float *bg, const float *y,
                              float *grad, double *partial, int nblocks, hipStream_t s);

// ---- explicit-matrix kernels -------------------------------------------------------
struct MatArgs {
    const void *A;
    int dtype;
    int64_t n_rows;
    int n_cols;
    // Per-column storage scale: the stored matrix is A[r,j]*scale[j] (NULL = 1).
    // Kernels divide x by it on load and divide column sums by it on reduction.
    const float *scale = nullptr;
    // Launch shape of the streaming kernels (0 = default): resident workgroups per
    // CU and consecutive rows each wave loads per step.
    int blocks_per_cu = 0;
    int rows_per_wave = 0;
};

// Largest blocks_per_cu the partial buffers must be sized for.
constexpr int kMaxBlocksPerCu = 8;
// Widest matrix row the streaming kernels hold in registers.
constexpr int kMaxMatCols = 2048;
// Row span of the streaming kernels for n_cols columns: the smallest of 256, 512,
// 1024 and 2048 that is >= n_cols (0 if wider).
// Row spans of the streaming kernels (the span fixes the lanes per row and the
// columns per lane): a matrix uses the smallest span that holds its columns.  The
// 384, 768 and 1536 spans keep the lanes of a matrix whose column count falls
// between two powers of two mostly busy.
constexpr int kNumSpans = 7;
__host__ __device__ constexpr int span_of(int i) {
    return i == 0 ? 256 : i == 1 ? 384 : i == 2 ? 512 : i == 3 ? 768 : i == 4 ? 1024
                                                        : i == 5 ? 1536 : 2048;
}
__host__ __device__ constexpr int row_span(int n_cols) {
    for (int i = 0; i < kNumSpans; i++)
        if (n_cols <= span_of(i)) return span_of(i);
    return 0;
}
__host__ __device__ constexpr int span_index(int span) {
    for (int i = 0; i < kNumSpans; i++)
        if (span_of(i) == span) return i;
    return -1;
}

// Per-column maxima of |A| (FP32 matrix) and power-of-two scales that map each
// column's maximum to about half the largest finite value of dtype (BF16/F16).
hipError_t launch_column_scale(const float *A, int64_t n_rows, int n_cols, int dtype,
                               float *partial_max, int nblocks, float *scale, hipStream_t s);
// Convert with per-column scale (scale may be NULL).
hipError_t launch_convert_scaled(const float *A, int64_t n_rows, int n_cols, int dtype,
                                 const float *scale, void *out, hipStream_t s);

int mat_num_blocks(const MatArgs &m);
hipError_t launch_project(const MatArgs &m, const float *x, float *out, hipStream_t s);
// Writes per-block FP64 partial column sums of diag(w) A (w = NULL: ones)
// into partial[nblocks x n_cols].
hipError_t launch_colsum_partials(const MatArgs &m, const float *w, double *partial,
                                  hipStream_t s);
// out[j] = sum_b partial[b, j]  (fixed order)
hipError_t launch_reduce_partials(const double *partial, int nblocks, int n_cols, double *out,
                                  hipStream_t s, const float *scale = nullptr);
// fused MLEM forward+adjoint pass: partial[b, j] = sum_{r in block b} A[r,j] y_r/(A_r x + c_r + 1e-20)
hipError_t launch_mlem_fused(const MatArgs &m, const float *x, const float *y, const float *c,
                             double *partial, hipStream_t s);
// ratio r = y / (f + c + 1e-20) given f = A x
hipError_t launch_ratio(int64_t n, const float *f, const float *y, const float *c, float *r,
                        hipStream_t s);
// x <- x * (sum_b partial[b,j]) / (sens[j] + 1e-7), non-finite -> 0, < 0 -> 0
hipError_t launch_mlem_update(const double *partial, int nblocks, int n_cols,
                              const double *sens, float *x, hipStream_t s,
                              const float *scale = nullptr);
// per-block FP64 partial Poisson log-likelihood
hipError_t launch_loglik_partials(const MatArgs &m, const float *x, const float *y,
                                  const float *c, double *partial, hipStream_t s);
hipError_t launch_convert(const float *A, int64_t n, int dtype, void *out, hipStream_t s);

// ---- batched MLEM ------------------------------------------------------------------
struct BatchProblem {
    const void *A;        // [n_rows x n_cols], dtype of the batch
    int64_t n_rows;
    int n_cols;           // multiple of 4
    int span;             // row span (span_of(i)) >= n_cols
    int seg_begin, seg_end;  // this problem's segments in the global list
    int sens_offset;      // offset of its sensitivity in the batch's sensitivity buffer
    int pad;
    const float *y, *c;
    float *x;
    const float *scale;   // per-column storage scale or NULL
    const double *sens;   // device pointer to its FP64 sensitivity
};

// Rows [r0, r1) of one problem, streamed by one workgroup.
struct BatchSegment {
    int problem;
    int pad;
    int64_t r0, r1;
};

// Problems of one row span, streamed by one persistent launch of `grid`
// workgroups; chunk c is segments wg_seg[off + c] .. wg_seg[off + c + 1]-1.  Static
// (dynamic == 0): n_chunks == grid and workgroup b streams chunk b.  Dynamic: the
// workgroups take the n_chunks chunks in order from the group's ticket counter.
struct BatchGroup {
    int span;
    int grid;
    int wg_seg_offset;
    int rows_per_wave;    // rows per row slot and step (1, 2 or 4)
    int n_chunks;
    int dynamic;
};

struct BatchArgs {
    const BatchProblem *probs;
    const BatchSegment *segs;
    const int *wg_seg;    // the groups' workgroup-to-segment offsets, concatenated
    int n_problems, n_segs, max_cols, stride, dtype;
    int vec8;             // 2-byte matrices: every n_cols % 8 == 0 and A 16-byte aligned
    int n_groups;
    BatchGroup groups[kNumSpans];  // the spans that occur, ascending
    double *partial;      // [n_segs x stride]
    int *tickets;         // [2 x kNumSpans] chunk counters of dynamic groups, zero between launches
};

// LDS doubles of the batch kernels' block reduction (the widest row shape).
constexpr int kBatchRedDoubles = 2048;

// group < 0: every group in turn; else that group only.
hipError_t launch_batch_stream(const BatchArgs &b, bool mlem, hipStream_t s, int group = -1);
hipError_t launch_batch_update(const BatchArgs &b, double *sens_out, bool update, hipStream_t s);
// Workgroups of the batch kernel for row span `span` and rows_per_wave R that fit
// on one CU at once.
hipError_t batch_resident_blocks(const BatchArgs &b, int span, int R, int *blocks_per_cu);

}  // namespace xrd

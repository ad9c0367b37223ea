// mlem.hip -- explicit model-matrix operators and the MLEM update on gfx950.
//
// Replaces the MATLAB loop of main.m:159-176:
//   x <- x .* A'( y ./ (A x + c + 1e-20) ) ./ (sum(A,1)' + 1e-7)
//   x(~isfinite(x)) = 0; x(x<0) = 0
// A is row-major [rows x cols] (the reference's Amatrix viewed as in
// main.m:140).  The iteration is HBM-bound (2 flop per A element read), so the
// fused kernel streams each row of A exactly once: one wave64 owns a row at a
// time, lanes hold 4 consecutive columns per 256-column group, the row dot
// product is a wave reduction, the ratio y/(Ax+c) is formed in FP64 and the
// row is accumulated into per-lane FP64 column sums.  Per-workgroup FP64
// partials are reduced in a fixed order by the update kernel, so every result
// is deterministic.
#include <hip/hip_fp16.h>

#include "xrd_device.hpp"
#include "xrd_internal.hpp"
#include "matrix_stream.hpp"

namespace xrd {

// Workgroup b handles rows [b*rpb, min((b+1)*rpb, n_rows)) with row shape
// RowShape<W, SPAN>; each step a wave takes R row slots' worth of rows.
template <typename T, int SPAN, int R, int MODE>
__global__ __launch_bounds__(kMatThreads) void colsum_kernel(const void *A, int64_t n_rows,
                                                             int n_cols, int64_t rpb,
                                                             const float *x, const float *y,
                                                             const float *c, const float *w,
                                                             const float *scale, double *partial) {
    __shared__ double red[RowShape<Vec<T>::W, SPAN>::S * SPAN];
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    int64_t r1 = r0 + rpb;
    if (r1 > n_rows) r1 = n_rows;
    colsum_block<T, SPAN, R, MODE>(A, n_rows, n_cols, r0, r1, x, scale, y, c, w, red,
                                   partial + (size_t)blockIdx.x * n_cols);
}

// out[r] = A_r . x  (FP32, L lanes per row)
template <typename T, int SPAN, int R>
__global__ __launch_bounds__(kMatThreads) void project_kernel(const void *A, int64_t n_rows,
                                                              int n_cols, int64_t rpb,
                                                              const float *x, const float *scale,
                                                              float *out) {
    constexpr int W = Vec<T>::W;
    typedef RowShape<W, SPAN> Sh;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = lane / Sh::L, li = lane % Sh::L;
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    int64_t r1 = r0 + rpb;
    if (r1 > n_rows) r1 = n_rows;
    float xv[Sh::NG][W];
    bool gv[Sh::NG];
    load_x<W, Sh::L, Sh::NG>(x, scale, n_cols, li, xv, gv);
    for (int64_t base = r0 + (int64_t)wid * R * Sh::S; base < r1;
         base += (int64_t)kMatWaves * R * Sh::S) {
        typename Vec<T>::raw rv[R][Sh::NG];
        float yv[R], cv[R];  // no per-row values in the projection (kColsOnes)
        load_rows<T, Sh::L, Sh::NG, R, kColsOnes>(A, n_rows * n_cols, n_cols, base + slot, r1, li,
                                                  gv, nullptr, nullptr, nullptr, rv, yv, cv);
#pragma unroll
        for (int k = 0; k < R; k++) {
            const float d = row_dot<T, Sh::L, Sh::NG>(rv[k], xv);
            const int64_t r = base + (int64_t)k * Sh::S + slot;
            if (li == 0 && r < r1) out[r] = d;
        }
    }
}

// Sum per-block partials in a fixed order: 16 waves each take every 16th
// block for 64 columns; the 16 wave sums are then added in wave order.
__global__ __launch_bounds__(1024) void reduce_partials_kernel(const double *partial,
                                                               int nblocks, int n_cols,
                                                               double *out, const double *sens,
                                                               float *x, const float *scale) {
    __shared__ double red[16][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (j < n_cols)
        for (int b = wid; b < nblocks; b += 16) s += partial[(size_t)b * n_cols + j];
    red[wid][lane] = s;
    __syncthreads();
    if (wid == 0 && j < n_cols) {
        double t = red[0][lane];
        for (int q = 1; q < 16; q++) t += red[q][lane];
        if (scale) t /= (double)scale[j];
        if (x) {
            // main.m:162-163,176
            double v = (double)x[j] * t / (sens[j] + 1e-7);
            float f = (float)v;
            if (!isfinite(f)) f = 0.f;
            if (f < 0.f) f = 0.f;
            x[j] = f;
        } else {
            out[j] = t;
        }
    }
}

__global__ void ratio_kernel(int64_t n, const float *f, const float *y, const float *c,
                             float *r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double cr = (double)c[i];
        r[i] = (float)(((double)y[i] - 0.0 * cr) / ((double)f[i] + cr + 1e-20));
    }
}

template <typename T, int SPAN, int R>
__global__ __launch_bounds__(kMatThreads) void loglik_kernel(const void *A, int64_t n_rows,
                                                             int n_cols, int64_t rpb,
                                                             const float *x, const float *y,
                                                             const float *c, const float *scale,
                                                             double *partial) {
    constexpr int W = Vec<T>::W;
    typedef RowShape<W, SPAN> Sh;
    __shared__ double red[kMatThreads];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = lane / Sh::L, li = lane % Sh::L;
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    int64_t r1 = r0 + rpb;
    if (r1 > n_rows) r1 = n_rows;
    float xv[Sh::NG][W];
    bool gv[Sh::NG];
    load_x<W, Sh::L, Sh::NG>(x, scale, n_cols, li, xv, gv);
    double ll = 0.0;
    for (int64_t base = r0 + (int64_t)wid * R * Sh::S; base < r1;
         base += (int64_t)kMatWaves * R * Sh::S) {
        typename Vec<T>::raw rv[R][Sh::NG];
        float yv[R], cv[R];
        load_rows<T, Sh::L, Sh::NG, R, kColsMlem>(A, n_rows * n_cols, n_cols, base + slot, r1, li,
                                                  gv, y, c, nullptr, rv, yv, cv);
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t r = base + (int64_t)k * Sh::S + slot;
            const float d = row_dot<T, Sh::L, Sh::NG>(rv[k], xv);
            if (li == 0 && r < r1) {
                const double f = (double)d + (double)cv[k] + 1e-20;
                const double yr = (double)yv[k];
                ll += (yr > 0.0 ? yr * log(f) : 0.0) - f;
            }
        }
    }
    red[threadIdx.x] = ll;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < kMatThreads; q++) s += red[q];
        partial[blockIdx.x] = s;
    }
}

// Stores v as element i of an n-element BF16/F16/BF24 matrix (round to nearest
// even; NaN stays a quiet NaN).
__device__ __forceinline__ void store_element(float v, int64_t i, int64_t n, int dtype, void *out) {
    uint32_t u = __float_as_uint(v);
    const bool nan = (u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu);
    if (dtype == XRD_DTYPE_BF16) {
        uint16_t b;
        if (nan) {
            b = (uint16_t)((u >> 16) | 0x40u);
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);
            b = (uint16_t)(u >> 16);
        }
        reinterpret_cast<uint16_t *>(out)[i] = b;
    } else if (dtype == XRD_DTYPE_BF24) {
        uint32_t t;  // the top 24 bits
        if (nan) {
            t = (u >> 8) | 0x4000u;
        } else {
            u += 0x7fu + ((u >> 8) & 1u);
            t = u >> 8;
        }
        reinterpret_cast<uint16_t *>(out)[i] = (uint16_t)(t >> 8);
        reinterpret_cast<uint8_t *>(out)[2 * n + i] = (uint8_t)(t & 0xffu);
    } else {
        reinterpret_cast<__half *>(out)[i] = __float2half_rn(v);
    }
}

__global__ void convert_kernel(const float *A, int64_t n, int dtype, void *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        store_element(A[i], i, n, dtype, out);
}

__global__ void convert_scaled_kernel(const float *A, int64_t n_rows, int n_cols, int dtype,
                                      const float *scale, void *out) {
    const int64_t n = n_rows * n_cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float v = A[i];
        if (scale) v *= scale[i % n_cols];  // power-of-two scale: exact
        store_element(v, i, n, dtype, out);
    }
}

// partial_max[b, j] = max over the rows of block b of |A[r, j]|
__global__ __launch_bounds__(256) void column_max_kernel(const float *A, int64_t n_rows,
                                                         int n_cols, int64_t rpb,
                                                         float *partial_max) {
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    int64_t r1 = r0 + rpb;
    if (r1 > n_rows) r1 = n_rows;
    for (int j = threadIdx.x; j < n_cols; j += blockDim.x) {
        float m = 0.f;
        for (int64_t r = r0; r < r1; r++) m = fmaxf(m, fabsf(A[r * n_cols + j]));
        partial_max[(size_t)blockIdx.x * n_cols + j] = m;
    }
}

__global__ void column_scale_kernel(const float *partial_max, int nblocks, int n_cols, int dtype,
                                    float *scale) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_cols) return;
    float m = 0.f;
    for (int b = 0; b < nblocks; b++) m = fmaxf(m, partial_max[(size_t)b * n_cols + j]);
    // F16: map the column maximum to about half the largest finite value (2^15),
    // clamped to 2^-100..2^100.  BF16 has the FP32 exponent range: no scaling.
    float sc = 1.f;
    if (dtype == XRD_DTYPE_F16 && m > 0.f)
        sc = exp2f(fminf(fmaxf(floorf(log2f(32768.f / m)), -100.f), 100.f));
    scale[j] = sc;
}

// ---------------------------------------------------------------------------
// launchers

static int g_num_cus = 0;
static int num_cus() {
    if (g_num_cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            n > 0)
            g_num_cus = n;
        else
            g_num_cus = 256;
    }
    return g_num_cus;
}

int mat_num_blocks(const MatArgs &m) {
    // blocks_per_cu resident workgroups per CU (default 2), never more blocks than
    // 16-row groups
    int bpc = m.blocks_per_cu > 0 ? m.blocks_per_cu : 2;
    if (bpc > kMaxBlocksPerCu) bpc = kMaxBlocksPerCu;
    int64_t nb = (int64_t)num_cus() * bpc;
    const int64_t cap = (m.n_rows + 15) / 16;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return (int)nb;
}

static int64_t rows_per_block(const MatArgs &m, int nb) { return (m.n_rows + nb - 1) / nb; }

// Instantiates the row span covering m.n_cols (256, 384, 512, 768, 1024, 1536 or 2048
// columns).
#define XRD_DISPATCH_SPAN(SPANV, ...)                                  \
    do {                                                               \
        switch (row_span(m.n_cols)) {                                  \
            case 256: { constexpr int SPANV = 256; __VA_ARGS__; break; }   \
            case 384: { constexpr int SPANV = 384; __VA_ARGS__; break; }   \
            case 512: { constexpr int SPANV = 512; __VA_ARGS__; break; }   \
            case 768: { constexpr int SPANV = 768; __VA_ARGS__; break; }   \
            case 1024: { constexpr int SPANV = 1024; __VA_ARGS__; break; } \
            case 1536: { constexpr int SPANV = 1536; __VA_ARGS__; break; } \
            case 2048: { constexpr int SPANV = 2048; __VA_ARGS__; break; } \
            default: return hipErrorInvalidValue;                      \
        }                                                              \
    } while (0)

// Returns CALL evaluated with T = the storage tag of the matrix: 2-byte matrices
// use 16-byte loads when every row starts 16-byte aligned.
#define XRD_DISPATCH_DTYPE(m, CALL)                                                  \
    do {                                                                             \
        const bool v8 = (m).n_cols % 8 == 0 && ((uintptr_t)(m).A & 15) == 0;         \
        switch ((m).dtype) {                                                         \
            case XRD_DTYPE_F32: { using T = float; return CALL; }                    \
            case XRD_DTYPE_BF16:                                                     \
                if (v8) { using T = bf16x8; return CALL; }                           \
                else { using T = bf16x4; return CALL; }                              \
            case XRD_DTYPE_F16:                                                      \
                if (v8) { using T = f16x8; return CALL; }                            \
                else { using T = f16x4; return CALL; }                               \
            case XRD_DTYPE_BF24: { using T = bf24x8; return CALL; }                  \
        }                                                                            \
        return hipErrorInvalidValue;                                                 \
    } while (0)

// Default rows per row slot and step: about 4 loads of 16 bytes per lane in
// flight per column group count.
template <int NG> constexpr int default_rows() { return NG >= 8 ? 1 : (NG >= 4 ? 2 : 4); }

template <typename T, int SPAN, int MODE>
static void colsum_launch_r(const MatArgs &m, int nb, int64_t rpb, const float *x,
                            const float *y, const float *c, const float *w, double *partial,
                            hipStream_t s) {
    constexpr int NG = RowShape<Vec<T>::W, SPAN>::NG;
    // rows per row slot and step: the autotuned value for the MLEM pass, else the default
    const int r = (MODE == kColsMlem && m.rows_per_wave > 0) ? m.rows_per_wave : default_rows<NG>();
    if (r >= 4 && NG <= 4) {
        hipLaunchKernelGGL((colsum_kernel<T, SPAN, 4, MODE>), dim3(nb), dim3(kMatThreads), 0, s,
                           m.A, m.n_rows, m.n_cols, rpb, x, y, c, w, m.scale, partial);
    } else if (r >= 2) {
        hipLaunchKernelGGL((colsum_kernel<T, SPAN, 2, MODE>), dim3(nb), dim3(kMatThreads), 0, s,
                           m.A, m.n_rows, m.n_cols, rpb, x, y, c, w, m.scale, partial);
    } else {
        hipLaunchKernelGGL((colsum_kernel<T, SPAN, 1, MODE>), dim3(nb), dim3(kMatThreads), 0, s,
                           m.A, m.n_rows, m.n_cols, rpb, x, y, c, w, m.scale, partial);
    }
}

template <typename T, int MODE>
static hipError_t colsum_t(const MatArgs &m, const float *x, const float *y, const float *c,
                           const float *w, double *partial, hipStream_t s) {
    const int nb = mat_num_blocks(m);
    const int64_t rpb = rows_per_block(m, nb);
    XRD_DISPATCH_SPAN(SPAN, { colsum_launch_r<T, SPAN, MODE>(m, nb, rpb, x, y, c, w, partial, s); });
    return hipGetLastError();
}

template <int MODE>
static hipError_t colsum_dtype(const MatArgs &m, const float *x, const float *y, const float *c,
                               const float *w, double *partial, hipStream_t s) {
    XRD_DISPATCH_DTYPE(m, (colsum_t<T, MODE>(m, x, y, c, w, partial, s)));
}

hipError_t launch_colsum_partials(const MatArgs &m, const float *w, double *partial,
                                  hipStream_t s) {
    if (w) return colsum_dtype<kColsWeighted>(m, nullptr, nullptr, nullptr, w, partial, s);
    return colsum_dtype<kColsOnes>(m, nullptr, nullptr, nullptr, nullptr, partial, s);
}

hipError_t launch_mlem_fused(const MatArgs &m, const float *x, const float *y, const float *c,
                             double *partial, hipStream_t s) {
    return colsum_dtype<kColsMlem>(m, x, y, c, nullptr, partial, s);
}

template <typename T>
static hipError_t project_t(const MatArgs &m, const float *x, float *out, hipStream_t s) {
    const int nb = mat_num_blocks(m);
    const int64_t rpb = rows_per_block(m, nb);
    XRD_DISPATCH_SPAN(SPAN, {
        constexpr int R = default_rows<RowShape<Vec<T>::W, SPAN>::NG>();
        hipLaunchKernelGGL((project_kernel<T, SPAN, R>), dim3(nb), dim3(kMatThreads), 0, s, m.A,
                           m.n_rows, m.n_cols, rpb, x, m.scale, out);
    });
    return hipGetLastError();
}

hipError_t launch_project(const MatArgs &m, const float *x, float *out, hipStream_t s) {
    XRD_DISPATCH_DTYPE(m, (project_t<T>(m, x, out, s)));
}

template <typename T>
static hipError_t loglik_t(const MatArgs &m, const float *x, const float *y, const float *c,
                           double *partial, hipStream_t s) {
    const int nb = mat_num_blocks(m);
    const int64_t rpb = rows_per_block(m, nb);
    XRD_DISPATCH_SPAN(SPAN, {
        hipLaunchKernelGGL((loglik_kernel<T, SPAN, 1>), dim3(nb), dim3(kMatThreads), 0, s, m.A,
                           m.n_rows, m.n_cols, rpb, x, y, c, m.scale, partial);
    });
    return hipGetLastError();
}

hipError_t launch_loglik_partials(const MatArgs &m, const float *x, const float *y,
                                  const float *c, double *partial, hipStream_t s) {
    XRD_DISPATCH_DTYPE(m, (loglik_t<T>(m, x, y, c, partial, s)));
}

hipError_t launch_reduce_partials(const double *partial, int nblocks, int n_cols, double *out,
                                  hipStream_t s, const float *scale) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((n_cols + 63) / 64), dim3(1024), 0, s,
                       partial, nblocks, n_cols, out, nullptr, nullptr, scale);
    return hipGetLastError();
}

hipError_t launch_mlem_update(const double *partial, int nblocks, int n_cols,
                              const double *sens, float *x, hipStream_t s, const float *scale) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((n_cols + 63) / 64), dim3(1024), 0, s,
                       partial, nblocks, n_cols, nullptr, sens, x, scale);
    return hipGetLastError();
}

hipError_t launch_ratio(int64_t n, const float *f, const float *y, const float *c, float *r,
                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ratio_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, f,
                       y, c, r);
    return hipGetLastError();
}

hipError_t launch_column_scale(const float *A, int64_t n_rows, int n_cols, int dtype,
                               float *partial_max, int nblocks, float *scale, hipStream_t s) {
    const int64_t rpb = (n_rows + nblocks - 1) / nblocks;
    hipLaunchKernelGGL(column_max_kernel, dim3(nblocks), dim3(256), 0, s, A, n_rows, n_cols, rpb,
                       partial_max);
    hipLaunchKernelGGL(column_scale_kernel, dim3((n_cols + 255) / 256), dim3(256), 0, s,
                       partial_max, nblocks, n_cols, dtype, scale);
    return hipGetLastError();
}

hipError_t launch_convert_scaled(const float *A, int64_t n_rows, int n_cols, int dtype,
                                 const float *scale, void *out, hipStream_t s) {
    const int64_t n = n_rows * n_cols;
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(convert_scaled_kernel, dim3((unsigned)blocks), dim3(256), 0, s, A, n_rows,
                       n_cols, dtype, scale, out);
    return hipGetLastError();
}

hipError_t launch_convert(const float *A, int64_t n, int dtype, void *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(convert_kernel, dim3((unsigned)blocks), dim3(256), 0, s, A, n, dtype, out);
    return hipGetLastError();
}

}  // namespace xrd

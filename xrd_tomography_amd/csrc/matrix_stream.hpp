// matrix_stream.hpp -- row streaming of an explicit model matrix shared by the
// single-matrix kernels (mlem.hip) and the batched persistent kernel (batch.hip).
//
// Every lane loads W consecutive elements of a row per column group with one
// 16-byte load (FP32: W = 4; BF16/F16: W = 8), or one 8-byte load for 2-byte
// matrices whose column count is not a multiple of 8 (W = 4).  L lanes share a
// row (L = 16, 32 or 64), so a wave works on S = 64/L rows side by side; lane li
// of its row slot holds columns W (li + L g) + i for column groups g < NG.  A row
// shape (L, NG) covers SPAN = L NG W columns; narrow rows use small L so that a
// row's dot product is reduced over fewer lanes and its FP64 ratio is formed for
// S rows at once.
#pragma once

#include <hip/hip_fp16.h>

#include "xrd_device.hpp"

namespace xrd {

constexpr int kMatThreads = 256;   // 4 waves per workgroup
constexpr int kMatWaves = kMatThreads / kWave;

// A is streamed once per pass with no reuse, so its loads are nontemporal: they
// bypass cache retention and measured 3.24 ms instead of 3.62 ms per 23 GB FP32
// pass on MI355X.
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

// The pointers of a batched problem are read from its descriptor in memory, so the
// compiler cannot see that they address global memory and would emit flat loads;
// glob() states it, which gives global loads with a scalar base.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *glob(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// Storage tags of the matrix element vector one lane loads.
struct bf16x4 {};
struct f16x4 {};
struct bf16x8 {};
struct f16x8 {};
struct bf24x8 {};  // 16-byte load from the uint16 plane + 8-byte load from the uint8 plane

__device__ __forceinline__ void unpack_bf16(unsigned int u, float &lo, float &hi) {
    lo = __uint_as_float(u << 16);
    hi = __uint_as_float(u & 0xffff0000u);
}
__device__ __forceinline__ void unpack_f16(unsigned int u, float &lo, float &hi) {
    const __half2 h = *reinterpret_cast<const __half2 *>(&u);
    const float2 f = __half22float2(h);
    lo = f.x;
    hi = f.y;
}

// Vec<T>::load(A, e, n) loads the W elements starting at element offset e of
// the n-element matrix A as a raw register vector; Vec<T>::unpack(raw, a) widens
// it to W floats.  Rows are kept raw while their dot products are reduced, which
// halves the registers a 2-byte row group holds in flight.
template <typename T> struct Vec;
template <> struct Vec<float> {
    static constexpr int W = 4;
    typedef f32x4v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t) {
        return __builtin_nontemporal_load(
            glob(reinterpret_cast<const f32x4v *>(reinterpret_cast<const float *>(A) + e)));
    }
    static __device__ __forceinline__ void unpack(const raw &v, float *a) {
        a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
    }
};
template <> struct Vec<bf16x4> {
    static constexpr int W = 4;
    typedef u32x2v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t) {
        return __builtin_nontemporal_load(
            glob(reinterpret_cast<const u32x2v *>(reinterpret_cast<const uint16_t *>(A) + e)));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_bf16(u.x, a[0], a[1]);
        unpack_bf16(u.y, a[2], a[3]);
    }
};
template <> struct Vec<f16x4> {
    static constexpr int W = 4;
    typedef u32x2v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t) {
        return __builtin_nontemporal_load(
            glob(reinterpret_cast<const u32x2v *>(reinterpret_cast<const uint16_t *>(A) + e)));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_f16(u.x, a[0], a[1]);
        unpack_f16(u.y, a[2], a[3]);
    }
};
template <> struct Vec<bf16x8> {
    static constexpr int W = 8;
    typedef u32x4v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t) {
        return __builtin_nontemporal_load(
            glob(reinterpret_cast<const u32x4v *>(reinterpret_cast<const uint16_t *>(A) + e)));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_bf16(u.x, a[0], a[1]);
        unpack_bf16(u.y, a[2], a[3]);
        unpack_bf16(u.z, a[4], a[5]);
        unpack_bf16(u.w, a[6], a[7]);
    }
};
template <> struct Vec<f16x8> {
    static constexpr int W = 8;
    typedef u32x4v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t) {
        return __builtin_nontemporal_load(
            glob(reinterpret_cast<const u32x4v *>(reinterpret_cast<const uint16_t *>(A) + e)));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_f16(u.x, a[0], a[1]);
        unpack_f16(u.y, a[2], a[3]);
        unpack_f16(u.z, a[4], a[5]);
        unpack_f16(u.w, a[6], a[7]);
    }
};

// BF24: element i of a pair is (h16_i << 16) | (l8_i << 8).
struct bf24raw {
    u32x4v h;
    u32x2v l;
};
template <> struct Vec<bf24x8> {
    static constexpr int W = 8;
    typedef bf24raw raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e, int64_t n) {
        const uint8_t *b = reinterpret_cast<const uint8_t *>(A);
        raw v;
        v.h = __builtin_nontemporal_load(glob(reinterpret_cast<const u32x4v *>(b + 2 * e)));
        v.l = __builtin_nontemporal_load(glob(reinterpret_cast<const u32x2v *>(b + 2 * n + e)));
        return v;
    }
    static __device__ __forceinline__ void unpack(const raw &v, float *a) {
        const unsigned h[4] = {v.h.x, v.h.y, v.h.z, v.h.w};
        const unsigned l[2] = {v.l.x, v.l.y};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned lb = l[q >> 1] >> (16 * (q & 1));  // bytes of elements 2q, 2q+1
            a[2 * q] = __uint_as_float((h[q] << 16) | ((lb << 8) & 0xff00u));
            a[2 * q + 1] = __uint_as_float((h[q] & 0xffff0000u) | (lb & 0xff00u));
        }
    }
};

enum { kColsOnes = 0, kColsWeighted = 1, kColsMlem = 2 };

// Row shape of SPAN columns at W elements per load: L lanes per row, NG column
// groups.  SPAN is a power of two in [256, 2048]; a lane holds 16 columns (32 for
// SPAN 2048), which bounds its FP64 column sums to 32 (64) registers.
template <int W, int SPAN> struct RowShape {
    // a power of two in [16, 64] lanes per row, about SPAN/16
    static constexpr int L = SPAN / 16 >= 64 ? 64 : (SPAN / 16 >= 32 ? 32 : 16);
    static constexpr int NG = SPAN / (L * W);
    static constexpr int S = kWave / L;
    static_assert(L * NG * W == SPAN, "row shape must tile the span");
};

// x (optionally divided by the per-column storage scale) in the lane's columns
// W (li + L g) + i; gv[g] marks column groups inside the matrix.  x = NULL
// leaves zeros.
template <int W, int L, int NG>
__device__ __forceinline__ void load_x(const float *x, const float *scale, int n_cols, int li,
                                       float (*xv)[W], bool *gv) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int j0 = W * (li + L * g);
        gv[g] = j0 < n_cols;
#pragma unroll
        for (int i = 0; i < W; i++) xv[g][i] = 0.f;
        if (x && gv[g]) {
#pragma unroll
            for (int h = 0; h < W; h += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(x + j0 + h);
                xv[g][h] = v.x; xv[g][h + 1] = v.y; xv[g][h + 2] = v.z; xv[g][h + 3] = v.w;
                if (scale) {
                    const float4 sc = *reinterpret_cast<const float4 *>(scale + j0 + h);
                    xv[g][h] /= sc.x; xv[g][h + 1] /= sc.y;
                    xv[g][h + 2] /= sc.z; xv[g][h + 3] /= sc.w;
                }
            }
        }
    }
}

// The lane's part of rows row0 + k S (k < R), raw (rows >= r1 and groups outside
// the matrix read as zero), and each row's per-row values: y and c for the MLEM
// ratio (kColsMlem) or the weight w (kColsWeighted, in yv).  The per-row loads are
// issued right after the row loads, unconditionally from the row index clamped to
// r1 - 1 (the callers use them only for rows below r1), so both are in flight
// together and a step waits for memory once.
template <typename T, int L, int NG, int R, int MODE>
__device__ __forceinline__ void load_rows(const void *A, int64_t n_elem, int n_cols, int64_t row0,
                                          int64_t r1, int li, const bool *gv, const float *y,
                                          const float *c, const float *w,
                                          typename Vec<T>::raw (*rv)[NG], float *yv, float *cv) {
    constexpr int W = Vec<T>::W;
    constexpr int S = kWave / L;
#pragma unroll
    for (int k = 0; k < R; k++)
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const int64_t r = row0 + (int64_t)k * S;
            rv[k][g] = typename Vec<T>::raw{};
            if (r < r1 && gv[g])
                rv[k][g] = Vec<T>::load(A, r * n_cols + W * (li + L * g), n_elem);
        }
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t r = row0 + (int64_t)k * S;
        const int64_t rc = r < r1 ? r : r1 - 1;
        yv[k] = 0.f;
        cv[k] = 0.f;
        if (MODE == kColsMlem) {
            yv[k] = *glob(y + rc);
            cv[k] = *glob(c + rc);
        } else if (MODE == kColsWeighted) {
            yv[k] = *glob(w + rc);
        }
    }
}

// FP32 dot product of the lane's part of one raw row with x, reduced over the
// row's L lanes.
template <typename T, int L, int NG>
__device__ __forceinline__ float row_dot(const typename Vec<T>::raw *rv,
                                         const float (*xv)[Vec<T>::W]) {
    constexpr int W = Vec<T>::W;
    float d = 0.f;
#pragma unroll
    for (int g = 0; g < NG; g++) {
        float a[W];
        Vec<T>::unpack(rv[g], a);
#pragma unroll
        for (int i = 0; i < W; i++) d = fmaf(a[i], xv[g][i], d);
    }
    return group_sum<L>(d);
}

// Streams rows [r0, r1) of A through the calling workgroup: each step every wave
// takes R S consecutive rows (row slot s of the wave holds rows base + k S + s);
// per row the weight is
//   kColsMlem:     y_r / (A_r x + c_r + 1e-20)   (main.m:162)
//   kColsWeighted: w_r
//   kColsOnes:     1
// and the row is accumulated into the lane's FP64 column sums acc.  Every
// product is added to acc in FP64 (the MLEM ratio formed in FP64), except in the
// MLEM pass over BF16, F16 and BF24 matrices (storage error 2^-8, 2^-11, 2^-16):
// there the ratio is formed in FP32 and the step's R products of a column are
// added in FP32 before one FP64 addition, which halves the vector instructions
// per element.
template <typename T, int L, int NG, int R, int MODE>
__device__ __forceinline__ void stream_rows(const void *A, int64_t n_elem, int n_cols, int64_t r0,
                                            int64_t r1,
                                            const float (*xv)[Vec<T>::W], const bool *gv,
                                            const float *y, const float *c, const float *w,
                                            double (*acc)[Vec<T>::W]) {
    constexpr int W = Vec<T>::W;
    constexpr int S = kWave / L;
    constexpr bool kWide = __is_same(T, float) || MODE != kColsMlem;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = lane / L, li = lane % L;
    for (int64_t base = r0 + (int64_t)wid * R * S; base < r1;
         base += (int64_t)kMatWaves * R * S) {
        typename Vec<T>::raw rv[R][NG];
        float yv[R], cv[R];
        load_rows<T, L, NG, R, MODE>(A, n_elem, n_cols, base + slot, r1, li, gv, y, c, w, rv, yv,
                                     cv);
        if constexpr (kWide) {
            double ratio[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int64_t r = base + (int64_t)k * S + slot;
                if (MODE == kColsMlem) {
                    const float d = row_dot<T, L, NG>(rv[k], xv);
                    double rt = 0.0;
                    if (r < r1) {
                        const double cr = (double)cv[k];
                        rt = ((double)yv[k] - 0.0 * cr) / ((double)d + cr + 1e-20);
                    }
                    ratio[k] = rt;
                } else if (MODE == kColsWeighted) {
                    ratio[k] = (r < r1) ? (double)yv[k] : 0.0;
                } else {
                    ratio[k] = (r < r1) ? 1.0 : 0.0;
                }
            }
#pragma unroll
            for (int k = 0; k < R; k++)
#pragma unroll
                for (int g = 0; g < NG; g++) {
                    float a[W];
                    Vec<T>::unpack(rv[k][g], a);
#pragma unroll
                    for (int i = 0; i < W; i++) acc[g][i] = fma((double)a[i], ratio[k], acc[g][i]);
                }
        } else {
            float ratio[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int64_t r = base + (int64_t)k * S + slot;
                if (MODE == kColsMlem) {
                    const float d = row_dot<T, L, NG>(rv[k], xv);
                    float rt = 0.f;
                    if (r < r1) {
                        const float cr = cv[k];
                        rt = (yv[k] - 0.f * cr) / (d + cr + 1e-20f);
                    }
                    ratio[k] = rt;
                } else if (MODE == kColsWeighted) {
                    ratio[k] = (r < r1) ? yv[k] : 0.f;
                } else {
                    ratio[k] = (r < r1) ? 1.f : 0.f;
                }
            }
#pragma unroll
            for (int g = 0; g < NG; g++) {
                float s[W];
#pragma unroll
                for (int i = 0; i < W; i++) s[i] = 0.f;
#pragma unroll
                for (int k = 0; k < R; k++) {
                    float a[W];
                    Vec<T>::unpack(rv[k][g], a);
#pragma unroll
                    for (int i = 0; i < W; i++) s[i] = fmaf(a[i], ratio[k], s[i]);
                }
#pragma unroll
                for (int i = 0; i < W; i++) acc[g][i] += (double)s[i];
            }
        }
    }
}

// Sums the column sums of the workgroup's waves and row slots through LDS and
// writes the n_cols totals to out.  red holds S * SPAN doubles (S = 64 / L): the
// waves add their slots into it in wave order, then the slots are added in slot
// order, so the result does not depend on timing.
template <int L, int NG, int W>
__device__ __forceinline__ void block_column_sums(const double (*acc)[W], double *red, int n_cols,
                                                  double *out) {
    constexpr int span = L * NG * W;
    constexpr int S = kWave / L;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int slot = lane / L, li = lane % L;
    for (int q = 0; q < kMatWaves; q++) {
        if (wid == q) {
#pragma unroll
            for (int g = 0; g < NG; g++)
#pragma unroll
                for (int i = 0; i < W; i++) {
                    double *p = red + slot * span + W * (li + L * g) + i;
                    *p = q == 0 ? acc[g][i] : *p + acc[g][i];
                }
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < n_cols; j += kMatThreads) {
        double s = red[j];
#pragma unroll
        for (int k = 1; k < S; k++) s += red[k * span + j];
        out[j] = s;
    }
}

// Column sums of diag(weight) A over rows [r0, r1) for one workgroup: loads x,
// streams the rows and writes the workgroup's n_cols FP64 totals to out.
template <typename T, int SPAN, int R, int MODE>
__device__ __forceinline__ void colsum_block(const void *A, int64_t n_rows, int n_cols, int64_t r0,
                                             int64_t r1,
                                             const float *x, const float *scale, const float *y,
                                             const float *c, const float *w, double *red,
                                             double *out) {
    constexpr int W = Vec<T>::W;
    typedef RowShape<W, SPAN> Sh;
    const int li = (threadIdx.x & 63) % Sh::L;
    float xv[Sh::NG][W];
    bool gv[Sh::NG];
    load_x<W, Sh::L, Sh::NG>(MODE == kColsMlem ? x : nullptr, scale, n_cols, li, xv, gv);
    double acc[Sh::NG][W];
#pragma unroll
    for (int g = 0; g < Sh::NG; g++)
#pragma unroll
        for (int i = 0; i < W; i++) acc[g][i] = 0.0;
    stream_rows<T, Sh::L, Sh::NG, R, MODE>(A, n_rows * n_cols, n_cols, r0, r1, xv, gv, y, c, w,
                                           acc);
    block_column_sums<Sh::L, Sh::NG, W>(acc, red, n_cols, out);
}

}  // namespace xrd

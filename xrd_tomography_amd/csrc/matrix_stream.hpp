// matrix_stream.hpp -- row streaming of an explicit model matrix shared by the
// single-matrix kernels (mlem.hip) and the batched persistent kernel (batch.hip).
//
// Every lane loads W consecutive elements of a row per column group with one
// 16-byte load (FP32: W = 4; BF16/F16: W = 8), or one 8-byte load for 2-byte
// matrices whose column count is not a multiple of 8 (W = 4).  A column group
// is 64 W columns wide; a wave covers NG groups of a row.  Lane l of group g
// holds columns W (l + 64 g) + i, i < W.
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

// Storage tags of the matrix element vector one lane loads.
struct bf16x4 {};
struct f16x4 {};
struct bf16x8 {};
struct f16x8 {};

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

// Vec<T>::load(A, e) loads the W elements starting at element offset e as a raw
// register vector; Vec<T>::unpack(raw, a) widens it to W floats.  Rows are kept
// raw while their dot products are reduced, which halves the registers a 2-byte
// row group holds in flight.
template <typename T> struct Vec;
template <> struct Vec<float> {
    static constexpr int W = 4;
    typedef f32x4v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e) {
        return __builtin_nontemporal_load(
            reinterpret_cast<const f32x4v *>(reinterpret_cast<const float *>(A) + e));
    }
    static __device__ __forceinline__ void unpack(const raw &v, float *a) {
        a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
    }
};
template <> struct Vec<bf16x4> {
    static constexpr int W = 4;
    typedef u32x2v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e) {
        return __builtin_nontemporal_load(
            reinterpret_cast<const u32x2v *>(reinterpret_cast<const uint16_t *>(A) + e));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_bf16(u.x, a[0], a[1]);
        unpack_bf16(u.y, a[2], a[3]);
    }
};
template <> struct Vec<f16x4> {
    static constexpr int W = 4;
    typedef u32x2v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e) {
        return __builtin_nontemporal_load(
            reinterpret_cast<const u32x2v *>(reinterpret_cast<const uint16_t *>(A) + e));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_f16(u.x, a[0], a[1]);
        unpack_f16(u.y, a[2], a[3]);
    }
};
template <> struct Vec<bf16x8> {
    static constexpr int W = 8;
    typedef u32x4v raw;
    static __device__ __forceinline__ raw load(const void *A, int64_t e) {
        return __builtin_nontemporal_load(
            reinterpret_cast<const u32x4v *>(reinterpret_cast<const uint16_t *>(A) + e));
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
    static __device__ __forceinline__ raw load(const void *A, int64_t e) {
        return __builtin_nontemporal_load(
            reinterpret_cast<const u32x4v *>(reinterpret_cast<const uint16_t *>(A) + e));
    }
    static __device__ __forceinline__ void unpack(const raw &u, float *a) {
        unpack_f16(u.x, a[0], a[1]);
        unpack_f16(u.y, a[2], a[3]);
        unpack_f16(u.z, a[4], a[5]);
        unpack_f16(u.w, a[6], a[7]);
    }
};

enum { kColsOnes = 0, kColsWeighted = 1, kColsMlem = 2 };

// x (optionally divided by the per-column storage scale) in the lane's columns
// W (lane + 64 g) + i; gv[g] marks column groups inside the matrix.  x = NULL
// leaves zeros.
template <int W, int NG>
__device__ __forceinline__ void load_x(const float *x, const float *scale, int n_cols, int lane,
                                       float (*xv)[W], bool *gv) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int j0 = W * (lane + 64 * g);
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

// R consecutive rows starting at `base`, raw (rows >= r1 and groups outside
// the matrix read as zero).
template <typename T, int NG, int R>
__device__ __forceinline__ void load_rows(const void *A, int n_cols, int64_t base, int64_t r1,
                                          int lane, const bool *gv,
                                          typename Vec<T>::raw (*rv)[NG]) {
    constexpr int W = Vec<T>::W;
#pragma unroll
    for (int k = 0; k < R; k++)
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const int64_t r = base + k;
            rv[k][g] = typename Vec<T>::raw{};
            if (r < r1 && gv[g]) rv[k][g] = Vec<T>::load(A, r * n_cols + W * (lane + 64 * g));
        }
}

// FP32 row dot product of one raw row, reduced over the wave.
template <typename T, int NG>
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
    return wave_sum(d);
}

// Streams rows [r0, r1) of A through the calling workgroup: each step the
// kMatWaves waves take R consecutive rows each; per row the weight is
//   kColsMlem:     y_r / (A_r x + c_r + 1e-20)   (main.m:162, FP64 ratio)
//   kColsWeighted: w_r
//   kColsOnes:     1
// and the row is accumulated into the lane's FP64 column sums acc.
template <typename T, int NG, int R, int MODE>
__device__ __forceinline__ void stream_rows(const void *A, int n_cols, int64_t r0, int64_t r1,
                                            const float (*xv)[Vec<T>::W], const bool *gv,
                                            const float *y, const float *c, const float *w,
                                            double (*acc)[Vec<T>::W]) {
    constexpr int W = Vec<T>::W;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t base = r0 + (int64_t)wid * R; base < r1; base += (int64_t)kMatWaves * R) {
        typename Vec<T>::raw rv[R][NG];
        load_rows<T, NG, R>(A, n_cols, base, r1, lane, gv, rv);
        double ratio[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t r = base + k;
            if (MODE == kColsMlem) {
                const float d = row_dot<T, NG>(rv[k], xv);
                double rt = 0.0;
                if (r < r1) {
                    const double cr = (double)c[r];
                    rt = ((double)y[r] - 0.0 * cr) / ((double)d + cr + 1e-20);
                }
                ratio[k] = rt;
            } else if (MODE == kColsWeighted) {
                ratio[k] = (r < r1) ? (double)w[r] : 0.0;
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
    }
}

// Sums the kMatWaves waves' column sums through LDS (red holds kMatWaves x
// NG*64*W doubles) and writes the workgroup's n_cols totals to out.
template <int NG, int W>
__device__ __forceinline__ void block_column_sums(const double (*acc)[W], double *red, int n_cols,
                                                  double *out) {
    constexpr int span = NG * 64 * W;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
        for (int i = 0; i < W; i++) red[wid * span + W * (lane + 64 * g) + i] = acc[g][i];
    __syncthreads();
    for (int j = threadIdx.x; j < n_cols; j += kMatThreads) {
        double s = red[j];
#pragma unroll
        for (int q = 1; q < kMatWaves; q++) s += red[q * span + j];
        out[j] = s;
    }
}

}  // namespace xrd

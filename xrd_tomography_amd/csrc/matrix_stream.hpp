// matrix_stream.hpp -- row streaming of an explicit model matrix shared by the
// single-matrix kernels (mlem.hip) and the batched persistent kernel (batch.hip).
#pragma once

#include <hip/hip_fp16.h>

#include "xrd_device.hpp"

namespace xrd {

constexpr int kMatThreads = 256;   // 4 waves per workgroup
constexpr int kMatWaves = kMatThreads / kWave;

// 4 consecutive elements of A as floats.  A is streamed once per pass with no
// reuse, so the loads are nontemporal: they bypass cache retention and measured
// 3.24 ms instead of 3.62 ms per 23 GB pass on MI355X.
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2v __attribute__((ext_vector_type(2)));

template <typename T> struct Quad;
template <> struct Quad<float> {
    static __device__ __forceinline__ float4 load(const float *p) {
        const f32x4v v = __builtin_nontemporal_load(reinterpret_cast<const f32x4v *>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    }
};
struct bf16x4 { uint32_t lo, hi; };
struct f16x4 { uint32_t lo, hi; };
template <> struct Quad<bf16x4> {
    static __device__ __forceinline__ float4 load(const bf16x4 *p) {
        const u32x2v u = __builtin_nontemporal_load(reinterpret_cast<const u32x2v *>(p));
        return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                           __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
    }
};
template <> struct Quad<f16x4> {
    static __device__ __forceinline__ float4 load(const f16x4 *p) {
        const u32x2v u = __builtin_nontemporal_load(reinterpret_cast<const u32x2v *>(p));
        const unsigned int ux = u.x, uy = u.y;
        const __half2 a = *reinterpret_cast<const __half2 *>(&ux);
        const __half2 b = *reinterpret_cast<const __half2 *>(&uy);
        const float2 fa = __half22float2(a), fb = __half22float2(b);
        return make_float4(fa.x, fa.y, fb.x, fb.y);
    }
};

// element pointer of column group (row, quad index)
template <typename T> __device__ __forceinline__ const T *quad_ptr(const void *A, int64_t r,
                                                                    int n_cols, int qi);
template <> __device__ __forceinline__ const float *quad_ptr<float>(const void *A, int64_t r,
                                                                    int n_cols, int qi) {
    return reinterpret_cast<const float *>(A) + r * n_cols + 4 * qi;
}
template <> __device__ __forceinline__ const bf16x4 *quad_ptr<bf16x4>(const void *A, int64_t r,
                                                                      int n_cols, int qi) {
    return reinterpret_cast<const bf16x4 *>(reinterpret_cast<const uint16_t *>(A) + r * n_cols +
                                            4 * qi);
}
template <> __device__ __forceinline__ const f16x4 *quad_ptr<f16x4>(const void *A, int64_t r,
                                                                    int n_cols, int qi) {
    return reinterpret_cast<const f16x4 *>(reinterpret_cast<const uint16_t *>(A) + r * n_cols +
                                           4 * qi);
}

enum { kColsOnes = 0, kColsWeighted = 1, kColsMlem = 2 };


// x (optionally divided by the per-column storage scale) in the lane's columns
// 4*(lane + 64 g) + i; gv[g] marks column groups inside the matrix.
template <int NG>
__device__ __forceinline__ void load_x(const float *x, const float *scale, int nq, int lane,
                                       float4 *xv, bool *gv) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int qi = lane + 64 * g;
        gv[g] = qi < nq;
        xv[g] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x && gv[g]) {
            xv[g] = *reinterpret_cast<const float4 *>(x + 4 * qi);
            if (scale) {
                const float4 sc = *reinterpret_cast<const float4 *>(scale + 4 * qi);
                xv[g] = make_float4(xv[g].x / sc.x, xv[g].y / sc.y, xv[g].z / sc.z, xv[g].w / sc.w);
            }
        }
    }
}

// Streams rows [r0, r1) of A through the calling wave group: each step the
// kMatWaves waves take R consecutive rows each; per row the weight is
//   kColsMlem:     y_r / (A_r x + c_r + 1e-20)   (main.m:162, FP64 ratio)
//   kColsWeighted: w_r
//   kColsOnes:     1
// and the row is accumulated into the lane's FP64 column sums acc.
template <typename T, int NG, int R, int MODE>
__device__ __forceinline__ void stream_rows(const void *A, int n_cols, int64_t r0, int64_t r1,
                                            const float4 *xv, const bool *gv, const float *y,
                                            const float *c, const float *w,
                                            double (*acc)[4]) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t base = r0 + (int64_t)wid * R; base < r1; base += (int64_t)kMatWaves * R) {
        float4 av[R][NG];
#pragma unroll
        for (int k = 0; k < R; k++)
#pragma unroll
            for (int g = 0; g < NG; g++) {
                const int64_t r = base + k;
                av[k][g] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (r < r1 && gv[g]) av[k][g] = Quad<T>::load(quad_ptr<T>(A, r, n_cols, lane + 64 * g));
            }
        double ratio[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t r = base + k;
            if (MODE == kColsMlem) {
                float d = 0.f;
#pragma unroll
                for (int g = 0; g < NG; g++) {
                    d = fmaf(av[k][g].x, xv[g].x, d);
                    d = fmaf(av[k][g].y, xv[g].y, d);
                    d = fmaf(av[k][g].z, xv[g].z, d);
                    d = fmaf(av[k][g].w, xv[g].w, d);
                }
                d = wave_sum(d);
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
                acc[g][0] = fma((double)av[k][g].x, ratio[k], acc[g][0]);
                acc[g][1] = fma((double)av[k][g].y, ratio[k], acc[g][1]);
                acc[g][2] = fma((double)av[k][g].z, ratio[k], acc[g][2]);
                acc[g][3] = fma((double)av[k][g].w, ratio[k], acc[g][3]);
            }
    }
}

}  // namespace xrd

// batch.hip -- MLEM for many independent problems (phantoms or slices of uneven
// size and view count) in one persistent launch per pass.
//
// Each problem's model matrix is cut into tiles of consecutive rows.  A
// persistent grid walks the global tile list (tile t, t + grid, ...); each tile
// is streamed exactly like the single-matrix fused pass (matrix_stream.hpp) and
// its FP64 column sums are written to a per-tile partial slot.  The update
// kernel then sums, per problem, the partials of that problem's tiles in tile
// order, so results do not depend on which workgroup processed which tile.
#include "matrix_stream.hpp"
#include "xrd_internal.hpp"

namespace xrd {

template <typename T, int NG, int R, int MODE>
__global__ __launch_bounds__(kMatThreads) void batch_stream_kernel(const BatchProblem *probs,
                                                                   const BatchTile *tiles,
                                                                   int n_tiles, int stride,
                                                                   double *partial) {
    constexpr int W = Vec<T>::W;
    __shared__ double red[kMatWaves * NG * 64 * W];
    const int lane = threadIdx.x & 63;
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const BatchTile tl = tiles[t];
        const BatchProblem pb = probs[tl.problem];
        float xv[NG][W];
        bool gv[NG];
        load_x<W, NG>(MODE == kColsMlem ? pb.x : nullptr, pb.scale, pb.n_cols, lane, xv, gv);
        double acc[NG][W];
#pragma unroll
        for (int g = 0; g < NG; g++)
#pragma unroll
            for (int i = 0; i < W; i++) acc[g][i] = 0.0;
        stream_rows<T, NG, R, MODE>(pb.A, pb.n_cols, tl.r0, tl.r1, xv, gv, pb.y, pb.c, nullptr,
                                    acc);
        block_column_sums<NG, W>(acc, red, pb.n_cols, partial + (size_t)t * stride);
        __syncthreads();
    }
}

// grid (ceil(max_cols/64), n_problems), 1024 lanes: for problem blockIdx.y and 64
// columns, sum the problem's tile partials in tile order (16 waves take every 16th
// tile, then the wave sums are added in wave order); then either write the
// de-scaled sensitivity or apply the MLEM update (main.m:162-163,176).
__global__ __launch_bounds__(1024) void batch_update_kernel(const BatchProblem *probs,
                                                            const double *partial, int stride,
                                                            double *sens_out, int update) {
    __shared__ double red[16][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const BatchProblem pb = probs[blockIdx.y];
    const int j = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (j < pb.n_cols)
        for (int t = pb.tile_begin + wid; t < pb.tile_end; t += 16)
            s += partial[(size_t)t * stride + j];
    red[wid][lane] = s;
    __syncthreads();
    if (wid == 0 && j < pb.n_cols) {
        double tsum = red[0][lane];
        for (int q = 1; q < 16; q++) tsum += red[q][lane];
        if (pb.scale) tsum /= (double)pb.scale[j];
        if (update) {
            double v = (double)pb.x[j] * tsum / (pb.sens[j] + 1e-7);
            float f = (float)v;
            if (!isfinite(f)) f = 0.f;
            if (f < 0.f) f = 0.f;
            pb.x[j] = f;
        } else {
            sens_out[pb.sens_offset + j] = tsum;
        }
    }
}

template <typename T, int MODE>
static hipError_t batch_stream_t(const BatchArgs &b, hipStream_t s) {
    constexpr int W = Vec<T>::W;
    const int ngr = (b.max_cols + 64 * W - 1) / (64 * W);
    const int R = b.rows_per_wave > 0 ? b.rows_per_wave : 2;
#define XRD_BATCH_LAUNCH(NGV, RV)                                                              \
    hipLaunchKernelGGL((batch_stream_kernel<T, NGV, RV, MODE>), dim3(b.grid), dim3(kMatThreads), \
                       0, s, b.probs, b.tiles, b.n_tiles, b.stride, b.partial)
#define XRD_BATCH_R(NGV)                          \
    do {                                          \
        if (R >= 4 && NGV <= 4) XRD_BATCH_LAUNCH(NGV, 4); \
        else if (R >= 2) XRD_BATCH_LAUNCH(NGV, 2);        \
        else XRD_BATCH_LAUNCH(NGV, 1);                    \
    } while (0)
    if (ngr <= 1) XRD_BATCH_R(1);
    else if (ngr <= 2) XRD_BATCH_R(2);
    else if (ngr <= 4) XRD_BATCH_R(4);
    else if constexpr (64 * 8 * W <= kMaxMatCols) {
        if (ngr <= 8) XRD_BATCH_R(8);
        else return hipErrorInvalidValue;
    } else return hipErrorInvalidValue;
#undef XRD_BATCH_R
#undef XRD_BATCH_LAUNCH
    return hipGetLastError();
}

template <typename T>
static hipError_t batch_stream_mode(const BatchArgs &b, bool mlem, hipStream_t s) {
    return mlem ? batch_stream_t<T, kColsMlem>(b, s) : batch_stream_t<T, kColsOnes>(b, s);
}

hipError_t launch_batch_stream(const BatchArgs &b, bool mlem, hipStream_t s) {
    switch (b.dtype) {
        case XRD_DTYPE_F32: return batch_stream_mode<float>(b, mlem, s);
        case XRD_DTYPE_BF16:
            return b.vec8 ? batch_stream_mode<bf16x8>(b, mlem, s) : batch_stream_mode<bf16x4>(b, mlem, s);
        case XRD_DTYPE_F16:
            return b.vec8 ? batch_stream_mode<f16x8>(b, mlem, s) : batch_stream_mode<f16x4>(b, mlem, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_update(const BatchArgs &b, double *sens_out, bool update, hipStream_t s) {
    dim3 grid((b.max_cols + 63) / 64, b.n_problems);
    hipLaunchKernelGGL(batch_update_kernel, grid, dim3(1024), 0, s, b.probs, b.partial, b.stride,
                       sens_out, update ? 1 : 0);
    return hipGetLastError();
}

}  // namespace xrd

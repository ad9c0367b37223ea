// batch.hip -- MLEM for many independent problems (phantoms or slices of uneven
// size and view count) with persistent launches over all of them.
//
// Problems are grouped by row span (the column count rounded up to 256, 384, 512,
// 768, 1024, 1536 or 2048, which fixes how many lanes share a row).  For each group the
// problems' rows are laid end to end and cut into contiguous chunks by cost
// (bytes of A plus a per-row charge for the row's reduction and ratio): either one
// chunk per workgroup of equal cost, or chunks of decreasing size that the
// workgroups take from a counter as they finish (guided order: half of the group's
// cost in the first round of chunks, a quarter in the next, ..., so the last
// chunks are short and no workgroup runs long after the others).  A chunk is a
// short list of segments, each a row range of one problem.  One persistent launch
// per group streams every segment and writes its FP64 column sums to the segment's
// own partial slot; a pass is one such launch per span present and one update
// launch.  The update kernel adds, per problem, the partials of its segments in
// segment order, so the result depends only on the launch shape and the chunking,
// never on timing or on which workgroup streamed a chunk.  A kernel per span (instead
// of one kernel branching on the span) keeps each at the registers of its own
// row shape.
#include "matrix_stream.hpp"
#include "xrd_internal.hpp"

namespace xrd {

// Rows per row slot and step for a span: at most 2 for rows whose FP64 column
// sums take 64 registers per lane.
template <typename T, int SPAN, int R> constexpr int batch_rows() {
    typedef RowShape<Vec<T>::W, SPAN> Sh;
    static_assert(Sh::S * SPAN <= kBatchRedDoubles, "block reduction exceeds the LDS buffer");
    return (Sh::NG * Vec<T>::W >= 32 && R > 2) ? 2 : R;
}

// Chunk c of the group is segments wg_seg[c] .. wg_seg[c + 1] - 1.  With ticket ==
// NULL, workgroup b streams chunk b (grid = n_chunks).  Otherwise the workgroups
// take chunks in index order from the counter ticket[0] until it passes n_chunks;
// ticket[1] counts the workgroups that have stopped, and the last one to stop sets
// both back to zero for the next launch.  Every chunk's segments write their own
// partial slots, so which workgroup streams a chunk does not change any result.
template <typename T, int SPAN, int R, int MODE>
__global__ __launch_bounds__(kMatThreads) void batch_stream_kernel(const BatchProblem *probs,
                                                                   const BatchSegment *segs,
                                                                   const int *wg_seg, int stride,
                                                                   double *partial, int n_chunks,
                                                                   int *ticket) {
    __shared__ double red[RowShape<Vec<T>::W, SPAN>::S * SPAN];
    __shared__ int next;
    int chunk = blockIdx.x;
    if (ticket) {
        if (threadIdx.x == 0) next = atomicAdd(ticket, 1);
        __syncthreads();
        chunk = next;
    }
    while (chunk < n_chunks) {
        const int s0 = wg_seg[chunk], s1 = wg_seg[chunk + 1];
        for (int si = s0; si < s1; si++) {
            const BatchSegment sg = segs[si];
            const BatchProblem pb = probs[sg.problem];
            colsum_block<T, SPAN, batch_rows<T, SPAN, R>(), MODE>(
                pb.A, pb.n_rows, pb.n_cols, sg.r0, sg.r1, MODE == kColsMlem ? pb.x : nullptr,
                pb.scale, pb.y, pb.c, nullptr, red, partial + (size_t)si * stride);
            __syncthreads();
        }
        if (!ticket) break;
        __syncthreads();  // every lane has read `next` before it is replaced
        if (threadIdx.x == 0) next = atomicAdd(ticket, 1);
        __syncthreads();
        chunk = next;
    }
    if (ticket && threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(ticket + 1, 1) == (int)gridDim.x - 1) {
            atomicExch(ticket, 0);
            atomicExch(ticket + 1, 0);
        }
    }
}

// grid (ceil(max_cols/64), n_problems), 1024 lanes: for problem blockIdx.y and 64
// columns, sum the problem's segment partials in a fixed order: wave w takes
// segments w, w + 16, w + 32, ... into sixteen running sums (segment w + 16 k into
// sum k mod 16, so sixteen loads are in flight), adds the sixteen in a fixed tree,
// and the wave sums are added in wave order.  Then it either writes the de-scaled
// sensitivity or applies the MLEM update (main.m:162-163,176).
__global__ __launch_bounds__(1024) void batch_update_kernel(const BatchProblem *probs,
                                                            const double *partial, int stride,
                                                            double *sens_out, int update) {
    __shared__ double red[16][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const BatchProblem pb = probs[blockIdx.y];
    const int j = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (j < pb.n_cols) {
        constexpr int U = 16;
        double u[U];
#pragma unroll
        for (int k = 0; k < U; k++) u[k] = 0.0;
        int t = pb.seg_begin + wid;
        for (; t + (U - 1) * 16 < pb.seg_end; t += U * 16)
#pragma unroll
            for (int k = 0; k < U; k++) u[k] += partial[(size_t)(t + 16 * k) * stride + j];
#pragma unroll
        for (int k = 0; k < U; k++)
            if (t + 16 * k < pb.seg_end) u[k] += partial[(size_t)(t + 16 * k) * stride + j];
#pragma unroll
        for (int h = U / 2; h > 0; h /= 2)
#pragma unroll
            for (int k = 0; k < h; k++) u[k] += u[k + h];
        s = u[0];
    }
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

template <typename T, int SPAN, int MODE>
static const void *batch_kernel_r(int R) {
    if (R >= 4) return reinterpret_cast<const void *>(&batch_stream_kernel<T, SPAN, 4, MODE>);
    if (R >= 2) return reinterpret_cast<const void *>(&batch_stream_kernel<T, SPAN, 2, MODE>);
    return reinterpret_cast<const void *>(&batch_stream_kernel<T, SPAN, 1, MODE>);
}

template <typename T, int MODE>
static const void *batch_kernel_span(int span, int R) {
    switch (span) {
        case 256: return batch_kernel_r<T, 256, MODE>(R);
        case 384: return batch_kernel_r<T, 384, MODE>(R);
        case 512: return batch_kernel_r<T, 512, MODE>(R);
        case 768: return batch_kernel_r<T, 768, MODE>(R);
        case 1024: return batch_kernel_r<T, 1024, MODE>(R);
        case 1536: return batch_kernel_r<T, 1536, MODE>(R);
        case 2048: return batch_kernel_r<T, 2048, MODE>(R);
    }
    return nullptr;
}

template <typename T>
static const void *batch_kernel_mode(bool mlem, int span, int R) {
    return mlem ? batch_kernel_span<T, kColsMlem>(span, R) : batch_kernel_span<T, kColsOnes>(span, R);
}

static const void *batch_kernel(const BatchArgs &b, int span, int R, bool mlem) {
    switch (b.dtype) {
        case XRD_DTYPE_F32: return batch_kernel_mode<float>(mlem, span, R);
        case XRD_DTYPE_BF16:
            return b.vec8 ? batch_kernel_mode<bf16x8>(mlem, span, R)
                          : batch_kernel_mode<bf16x4>(mlem, span, R);
        case XRD_DTYPE_F16:
            return b.vec8 ? batch_kernel_mode<f16x8>(mlem, span, R)
                          : batch_kernel_mode<f16x4>(mlem, span, R);
        case XRD_DTYPE_BF24: return batch_kernel_mode<bf24x8>(mlem, span, R);
    }
    return nullptr;
}

hipError_t batch_resident_blocks(const BatchArgs &b, int span, int R, int *blocks_per_cu) {
    const void *k = batch_kernel(b, span, R, true);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, kMatThreads, 0);
}

hipError_t launch_batch_stream(const BatchArgs &b, bool mlem, hipStream_t s, int group) {
    for (int g = 0; g < b.n_groups; g++) {
        if (group >= 0 && g != group) continue;
        const BatchGroup &gr = b.groups[g];
        const void *k = batch_kernel(b, gr.span, gr.rows_per_wave, mlem);
        if (!k) return hipErrorInvalidValue;
        const BatchProblem *probs = b.probs;
        const BatchSegment *segs = b.segs;
        const int *wg_seg = b.wg_seg + gr.wg_seg_offset;
        int stride = b.stride;
        double *partial = b.partial;
        int n_chunks = gr.n_chunks;
        int *ticket = gr.dynamic ? b.tickets + 2 * g : nullptr;
        void *args[] = {&probs, &segs, &wg_seg, &stride, &partial, &n_chunks, &ticket};
        hipError_t e = hipLaunchKernel(k, dim3(gr.grid), dim3(kMatThreads), args, 0, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_batch_update(const BatchArgs &b, double *sens_out, bool update, hipStream_t s) {
    dim3 grid((b.max_cols + 63) / 64, b.n_problems);
    hipLaunchKernelGGL(batch_update_kernel, grid, dim3(1024), 0, s, b.probs, b.partial, b.stride,
                       sens_out, update ? 1 : 0);
    return hipGetLastError();
}

}  // namespace xrd

// xrd_abi.hip -- the C ABI of include/xrd_tomo.h: argument validation,
// workspace management and kernel dispatch.
//
// Mirrors the MEX gateway Forward_XRD (Forward_XRD.cu:76-302) with caller-owned
// device buffers, explicit dimensions, a stream and status codes instead of
// mxGPUArray copies, hidden allocations and no error checking.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/xrd_tomo.h"
#include "xrd_internal.hpp"

using namespace xrd;

namespace {

thread_local char g_err[512] = "no error";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(XRD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                     \
    do {                                        \
        hipError_t e_ = (expr);                 \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

bool aligned(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// ---- device scratch of xrd_mlem_iterate_static -------------------------------------
// A fixed array in every GPU's copy of the code object: the FP64 sensitivity, then the
// per-workgroup FP64 partials of the fused MLEM pass.  In host code `scratch` is only a
// symbol handle; kernels get the device address that hipGetSymbolAddress returns for
// the current device (scratch_ptr).
constexpr size_t kScratchDoubles = XRD_MLEM_STATIC_SCRATCH_BYTES / sizeof(double);
static_assert(XRD_MLEM_STATIC_SCRATCH_BYTES % 256 == 0, "scratch size");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
bool g_scratch_resolved[kMaxDevices] = {};
double *g_scratch_dev[kMaxDevices] = {};
hipError_t g_scratch_status[kMaxDevices] = {};

// With `device` current: looks up that device's address of `scratch` once.
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_resolved[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_status[device] = e;
    g_scratch_dev[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
    g_scratch_resolved[device] = true;
}

// The device address of `device`'s scratch (NULL if the lookup failed or never ran).
double *scratch_ptr(int device, hipError_t *status) {
    if (device < 0 || device >= kMaxDevices) {
        *status = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    *status = g_scratch_resolved[device] ? g_scratch_status[device] : hipErrorNotInitialized;
    return g_scratch_dev[device];
}

}  // namespace

struct xrd_context {
    int device = 0;
    // forward workspace
    DevBuf mu_padded, s2u8, cpx, ref_list, voxels, offsets, src_int, tables, pv, mu_vol;
    // explicit-matrix partials
    DevBuf partial, small;
    // coefficient-model workspace
    DevBuf cm_maps, cm_src, cm_vlist, cm_h, cm_tm, cm_gtm, cm_gam, cm_ld, cm_fq;
    // pinned staging for host-computed per-view / per-energy tables
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // pinned word for small device-to-host reads (the non-air voxel count)
    int *pinned_word = nullptr;
    hipEvent_t staged = nullptr;
    bool staged_pending = false;
    // projector launch shape (xrd_forward_set_launch / xrd_forward_autotune; 0 = default)
    int fwd_kc = 0, fwd_vpb = 0;
};

struct xrd_mlem_plan {
    xrd_context *ctx = nullptr;
    MatArgs m{};
    const float *y = nullptr, *c = nullptr;
    float *x = nullptr;
    int mode = XRD_MLEM_FUSED;
    int nblocks = 0;
    DevBuf sens, partial, f, r, llp;
    // captured iterations: one, and kPlanGraphIters in a row (one graph launch per
    // kPlanGraphIters iterations leaves no launch gap between them)
    hipGraph_t graph = nullptr, graph_multi = nullptr;
    hipGraphExec_t exec = nullptr, exec_multi = nullptr;
    hipStream_t graph_stream = nullptr;
};

constexpr int kPlanGraphIters = 8;

// Drops the captured iterations (they hold the launch shape they were captured with).
static void plan_drop_graphs(xrd_mlem_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->exec_multi) (void)hipGraphExecDestroy(p->exec_multi);
    if (p->graph_multi) (void)hipGraphDestroy(p->graph_multi);
    p->exec = p->exec_multi = nullptr;
    p->graph = p->graph_multi = nullptr;
}

extern "C" {

const char *xrd_last_error_string(void) { return g_err; }

int xrd_version(void) { return 100; }

int xrd_geometry_reference(int32_t nx, int32_t ny, int32_t n_views, int32_t n_cols,
                           int32_t n_rows, xrd_geometry *g) {
    if (!g) return fail(XRD_ERR_NULL_POINTER, "xrd_geometry_reference: g is NULL");
    if (nx <= 0 || ny <= 0 || n_views <= 0 || n_cols <= 0 || n_rows <= 0)
        return fail(XRD_ERR_DIMS, "xrd_geometry_reference: non-positive size");
    // integrate_diffraction.cu:81-128
    const double mmv = (double)nx / 100.0;
    memset(g, 0, sizeof(*g));
    g->nx = nx;
    g->ny = ny;
    g->n_views = n_views;
    g->n_cols = n_cols;
    g->n_rows = n_rows;
    g->center_x = 100.0 * mmv;
    g->center_y = 100.0 * mmv;
    g->r_src = 150.0 * mmv;
    g->r_det = 170.0 * mmv;
    g->view_angle0 = 0.0;
    g->view_angle_step = 2.0 * M_PI / n_views;
    g->col_pitch = 0.5 * mmv;
    g->row_pitch = 1.0 * mmv;
    g->det_z0 = 10.0 * mmv;
    g->target_angle = M_PI / 6.0;
    g->focal_spot = 0.5 * mmv;
    g->wedge_angle = 0.5 / 180.0 * M_PI;
    g->voxel_size = 1.0;
    return XRD_OK;
}

int xrd_discretization_reference(int32_t n_energy, int32_t n_q, int32_t n_qf, int32_t n_mat,
                                 xrd_discretization *d) {
    if (!d) return fail(XRD_ERR_NULL_POINTER, "xrd_discretization_reference: d is NULL");
    if (n_energy <= 10 || n_q <= 0 || n_qf <= 0 || n_mat <= 0)
        return fail(XRD_ERR_DIMS, "xrd_discretization_reference: bad size");
    // Forward_XRD.cu:21-53, integrate_diffraction.cu:138-140
    memset(d, 0, sizeof(*d));
    d->n_energy = n_energy;
    d->energy_halfwidth = 5;
    d->n_q = n_q;
    d->n_qf = n_qf;
    d->n_mat = n_mat;
    d->lut_nq = 64;
    d->lut_na = 64;
    d->flags = 0;
    d->e_min = 8.0;
    d->e_width = 72.0 / n_energy;
    d->q_min = 0.5;
    d->q_max = 6.0;
    d->q_dq0 = 0.01;
    d->ray_step = 0.25;
    d->hbarc = 1.973269804;
    d->mc2 = 510.99895;
    d->kn0 = 0.751691286166726;
    d->kn1 = 1.4599683207703;
    d->kn2 = -0.64597532690936;
    return XRD_OK;
}

int64_t xrd_num_rows(const xrd_geometry *g, const xrd_discretization *d) {
    if (!g || !d) return -1;
    return (int64_t)g->n_views * g->n_cols * g->n_rows * d->n_energy;
}

int64_t xrd_num_cols(const xrd_discretization *d) {
    if (!d) return -1;
    return (int64_t)d->n_q * d->n_mat;
}

int xrd_context_create(int device, xrd_context **out) {
    if (!out) return fail(XRD_ERR_NULL_POINTER, "xrd_context_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
        return fail(XRD_ERR_INVALID_ARG, "xrd_context_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    // set-up of the allocation-free MLEM entry point: this device's scratch address
    resolve_scratch(device);
    xrd_context *c = new xrd_context();
    c->device = device;
    hipError_t e = hipEventCreateWithFlags(&c->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipEventCreate");
    }
    *out = c;
    return XRD_OK;
}

int xrd_context_destroy(xrd_context *c) {
    if (!c) return XRD_OK;
    (void)hipSetDevice(c->device);
    if (c->staged_pending) hipEventSynchronize(c->staged);
    for (DevBuf *b : {&c->mu_padded, &c->s2u8, &c->cpx, &c->ref_list, &c->voxels, &c->offsets,
                      &c->src_int, &c->tables, &c->pv, &c->mu_vol, &c->partial, &c->small, &c->cm_maps,
                      &c->cm_src, &c->cm_vlist, &c->cm_h, &c->cm_tm, &c->cm_gtm, &c->cm_gam,
                      &c->cm_ld, &c->cm_fq})
        b->release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->pinned_word) (void)hipHostFree(c->pinned_word);
    if (c->staged) (void)hipEventDestroy(c->staged);
    delete c;
    return XRD_OK;
}

int64_t xrd_context_workspace_bytes(const xrd_context *c) {
    if (!c) return 0;
    int64_t s = 0;
    for (const DevBuf *b : {&c->mu_padded, &c->s2u8, &c->cpx, &c->ref_list, &c->voxels,
                            &c->offsets, &c->src_int, &c->tables, &c->pv, &c->mu_vol, &c->partial,
                            &c->small, &c->cm_maps, &c->cm_src, &c->cm_vlist, &c->cm_h,
                            &c->cm_tm, &c->cm_gtm, &c->cm_gam, &c->cm_ld, &c->cm_fq})
        s += (int64_t)b->bytes;
    return s;
}

}  // extern "C"

namespace {

int check_geometry(const xrd_geometry *g, const xrd_discretization *d, const char *fn) {
    if (!g || !d) return fail(XRD_ERR_NULL_POINTER, "%s: geometry/discretization is NULL", fn);
    if (g->nx <= 0 || g->ny <= 0 || g->n_views <= 0 || g->n_cols <= 0 || g->n_rows <= 0)
        return fail(XRD_ERR_DIMS, "%s: non-positive geometry size", fn);
    if ((int64_t)g->nx * g->ny > (1ll << 28))
        return fail(XRD_ERR_DIMS, "%s: voxel grid too large", fn);
    if (d->n_energy <= 0 || d->n_q <= 0 || d->n_qf <= 0 || d->n_mat <= 0 || d->lut_nq <= 0 ||
        d->lut_na <= 0 || d->energy_halfwidth < 0)
        return fail(XRD_ERR_DIMS, "%s: non-positive discretization size", fn);
    if (d->energy_halfwidth > kMaxEnergyHalfwidth)
        return fail(XRD_ERR_UNSUPPORTED, "%s: energy_halfwidth > %d", fn, kMaxEnergyHalfwidth);
    if (d->n_energy <= 2 * d->energy_halfwidth)
        return fail(XRD_ERR_DIMS, "%s: n_energy (%d) must exceed 2*energy_halfwidth (%d)", fn,
                    d->n_energy, 2 * d->energy_halfwidth);
    const double vals[] = {g->center_x, g->center_y, g->r_src, g->r_det, g->view_angle0,
                           g->view_angle_step, g->col_pitch, g->row_pitch, g->det_z0,
                           g->target_angle, g->focal_spot, g->wedge_angle, g->voxel_size,
                           d->e_min, d->e_width, d->q_min, d->q_max, d->q_dq0, d->ray_step,
                           d->hbarc, d->mc2, d->kn0, d->kn1, d->kn2};
    for (double v : vals)
        if (!std::isfinite(v)) return fail(XRD_ERR_NONFINITE, "%s: non-finite descriptor field", fn);
    if (!(d->ray_step > 0.0) || !(d->e_width > 0.0) || !(d->q_max > d->q_min) ||
        !(d->hbarc > 0.0) || !(d->mc2 > 0.0))
        return fail(XRD_ERR_INVALID_ARG, "%s: ray_step, e_width, hbarc, mc2 must be > 0 and "
                    "q_max > q_min", fn);
    if (d->e_min <= 0.0) return fail(XRD_ERR_INVALID_ARG, "%s: e_min must be > 0", fn);
    return XRD_OK;
}

hipError_t stage_to_device(xrd_context *c, const void *host, size_t bytes, void *dev,
                           hipStream_t s) {
    if (c->staged_pending) {
        hipError_t e = hipEventSynchronize(c->staged);
        if (e != hipSuccess) return e;
        c->staged_pending = false;
    }
    if (bytes > c->pinned_bytes) {
        if (c->pinned) (void)hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_bytes = 0;
        hipError_t e = hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        c->pinned_bytes = bytes;
    }
    memcpy(c->pinned, host, bytes);
    hipError_t e = hipMemcpyAsync(dev, c->pinned, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    e = hipEventRecord(c->staged, s);
    if (e == hipSuccess) c->staged_pending = true;
    return e;
}

}  // namespace

namespace {

// Host FP64 tables (views, energies, coarse q axis) staged to the context and the
// geometry/discretisation fields of ForwardArgs shared by all forward-type calls.
int setup_forward_common(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                         const float *spectrum, ForwardArgs &fa, hipStream_t s) {
    const int V = g->n_views, E = d->n_energy, Q = d->n_q;
    const size_t vbytes = (size_t)V * sizeof(ViewConst);
    const size_t ebytes = (size_t)E * sizeof(EnergyConst);
    const size_t qbytes = (size_t)(2 * Q + 1) * sizeof(double);
    std::vector<unsigned char> host(vbytes + ebytes + qbytes);
    ViewConst *vc = reinterpret_cast<ViewConst *>(host.data());
    EnergyConst *ec = reinterpret_cast<EnergyConst *>(host.data() + vbytes);
    double *qc = reinterpret_cast<double *>(host.data() + vbytes + ebytes);
    for (int v = 0; v < V; v++) {
        // integrate_diffraction.cu:93-95
        const double phi = g->view_angle0 + (double)v * g->view_angle_step;
        vc[v].sp = sin(phi);
        vc[v].cp = cos(phi);
        vc[v].src_x = g->center_x + g->r_src * vc[v].sp;
        vc[v].src_y = g->center_y - g->r_src * vc[v].cp;
    }
    for (int e = 0; e < E; e++) {
        // integrate_diffraction.cu:269-292
        EnergyConst &k = ec[e];
        k.es1 = d->e_min + (double)e * d->e_width;
        k.es2 = d->e_min + (double)(e + 1) * d->e_width;
        k.esm = 0.5 * (k.es1 + k.es2);
        k.ein = k.esm / d->mc2;
        k.inv_f1 = 1.0 / (k.ein * k.ein * k.ein);
        k.inv_f2 = 1.0 / (d->kn0 + d->kn1 * k.ein + d->kn2 * k.ein * k.ein);
        k.wcoef = 2.0 * k.esm / d->hbarc;
        k.k2sq = (k.esm / d->hbarc) * (k.esm / d->hbarc);
    }
    {
        // wvt_value.cu:8-23
        const double NW = (double)Q;
        const double cc = (d->q_max - d->q_min - d->q_dq0 * NW) / (NW * NW);
        for (int n = 0; n < Q; n++) {
            const double e1 = d->q_min + (double)n * d->q_dq0 + cc * (double)n * (double)n;
            const double e2 = d->q_min + (double)(n + 1) * d->q_dq0 + cc * (double)(n + 1) * (double)(n + 1);
            qc[n] = 0.5 * (e1 + e2);
            qc[Q + n] = e2 - e1;
        }
        qc[2 * Q] = d->q_min;
    }
    HIP_TRY(ctx->tables.ensure(host.size()), "workspace alloc");
    HIP_TRY(stage_to_device(ctx, host.data(), host.size(), ctx->tables.ptr, s), "stage tables");

    fa = ForwardArgs{};
    fa.nx = g->nx;
    fa.ny = g->ny;
    fa.n_views = V;
    fa.n_cols = g->n_cols;
    fa.n_rows = g->n_rows;
    fa.n_det = g->n_cols * g->n_rows;
    fa.n_energy = E;
    fa.dn = d->energy_halfwidth;
    fa.n_q = Q;
    fa.n_qf = d->n_qf;
    fa.n_mat = d->n_mat;
    fa.lut_nq = d->lut_nq;
    fa.lut_na = d->lut_na;
    fa.flags = d->flags;
    fa.center_x = g->center_x;
    fa.center_y = g->center_y;
    fa.r_src = g->r_src;
    fa.r_det = g->r_det;
    fa.col_pitch = g->col_pitch;
    fa.row_pitch = g->row_pitch;
    fa.det_z0 = g->det_z0;
    fa.sta = sin(g->target_angle);
    fa.cta = cos(g->target_angle);
    fa.tan_w = tan(g->wedge_angle);
    fa.tan_w2 = tan(g->wedge_angle / 2.0);
    fa.focal_spot = g->focal_spot;
    fa.voxel_size = g->voxel_size;
    fa.adet = g->col_pitch * g->row_pitch;
    fa.e_min = d->e_min;
    fa.e_width = d->e_width;
    fa.hbarc = d->hbarc;
    fa.mc2 = d->mc2;
    fa.kn0 = d->kn0;
    fa.kn1 = d->kn1;
    fa.kn2 = d->kn2;
    fa.ray_step = d->ray_step;
    fa.dq_fine = (d->q_max - d->q_min) / (double)d->n_qf;
    fa.spectrum = spectrum;
    fa.views = (const ViewConst *)ctx->tables.ptr;
    fa.energies = (const EnergyConst *)((unsigned char *)ctx->tables.ptr + vbytes);
    fa.q_coarse = (const double *)((unsigned char *)ctx->tables.ptr + vbytes + ebytes);
    return XRD_OK;
}

}  // namespace

namespace {

// The form-factor and attenuation tables of a forward call: FP32 (ifine, a1, a2) or
// bfloat16 with per-material FP32 scales (the xrd_bf16_tables fields).
struct TableArgs {
    const float *ifine = nullptr, *a1 = nullptr, *a2 = nullptr;
    const xrd_bf16_tables *b16 = nullptr;
};

// xrd_forward for the slice `slice` of an n_slices-slice volume (segmentation and the
// maps point at the volumes); n_slices = 1 is the 2-D model.
int forward_impl(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                 const float *spectrum, const TableArgs &tab, const int32_t *seg_vol,
                 int n_slices, int slice, const float *s2_table, double lut_q_min,
                 double lut_q_max, double lut_a_min, double lut_a_max, float *out_counts,
                 float *out_compton, float *out_A, void *stream, const char *fn) {
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    const int32_t *segmentation = seg_vol;
    int rc = check_geometry(g, d, fn);
    if (rc) return rc;
    const xrd_bf16_tables *bt = tab.b16;
    const bool have_tables =
        bt ? (bt->ifine && bt->ifine_scale && bt->a1 && bt->a2 && bt->mu_scale)
           : (tab.ifine && tab.a1 && tab.a2);
    if (!spectrum || !have_tables || !segmentation || !s2_table || !out_counts || !out_compton)
        return fail(XRD_ERR_NULL_POINTER, "%s: a required input/output pointer is NULL", fn);
    const void *ptrs[] = {spectrum, segmentation, s2_table, out_counts, out_compton, out_A,
                          tab.ifine, tab.a1, tab.a2, bt ? bt->ifine : nullptr,
                          bt ? bt->ifine_scale : nullptr, bt ? bt->a1 : nullptr,
                          bt ? bt->a2 : nullptr, bt ? bt->mu_scale : nullptr};
    for (const void *p : ptrs)
        if (p && !aligned(p, 16)) return fail(XRD_ERR_ALIGN, "%s: pointers must be 16-byte aligned", fn);
    const double lims[] = {lut_q_min, lut_q_max, lut_a_min, lut_a_max};
    for (double v : lims)
        if (!std::isfinite(v)) return fail(XRD_ERR_NONFINITE, "%s: non-finite LUT limit", fn);
    if (!(lut_q_max > lut_q_min) || !(lut_a_max > lut_a_min))
        return fail(XRD_ERR_INVALID_ARG, "%s: LUT limits must satisfy max > min", fn);
    if ((int64_t)d->n_q * d->n_mat * (int64_t)g->n_views * g->n_cols * g->n_rows * d->n_energy >
        (int64_t)1 << 40)
        return fail(XRD_ERR_DIMS, "%s: model matrix too large", fn);
    if (d->n_energy > 64 || d->n_q > 256 || d->n_qf > 768)
        return fail(XRD_ERR_UNSUPPORTED,
                    "%s: this build supports n_energy <= 64, n_q <= 256, n_qf <= 768", fn);
    if (n_slices <= 0 || slice < 0 || slice >= n_slices)
        return fail(XRD_ERR_DIMS, "%s: need 0 <= slice < n_slices", fn);
    if ((size_t)(g->nx + 2) * (g->ny + 2) * (size_t)(n_slices + 2) > ((size_t)1 << 31))
        return fail(XRD_ERR_DIMS, "%s: volume too large", fn);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");

    const int nx = g->nx, ny = g->ny, V = g->n_views;
    const size_t npix = (size_t)nx * ny;
    const size_t npad = (size_t)(nx + 2) * (ny + 2);
    // the maps as the preparation kernels load them: the whole volume, and the slice
    MapSrc a1v, a2v;
    if (bt) {
        a1v.b16 = bt->a1;
        a2v.b16 = bt->a2;
        a1v.scale = bt->mu_scale;
        a2v.scale = bt->mu_scale + (d->n_mat + 1);
        a1v.seg = a2v.seg = seg_vol;
        a1v.n_ids = a2v.n_ids = d->n_mat + 1;
    } else {
        a1v.f32 = tab.a1;
        a2v.f32 = tab.a2;
    }
    MapSrc a1s = a1v, a2s = a2v;
    const size_t off = npix * slice;
    for (MapSrc *m : {&a1s, &a2s}) {
        if (m->f32) m->f32 += off;
        if (m->b16) m->b16 += off;
        if (m->seg) m->seg += off;
    }
    segmentation += off;
    HIP_TRY(ctx->mu_padded.ensure(npad * sizeof(float2)), "workspace alloc");
    if (n_slices > 1) {
        HIP_TRY(ctx->mu_vol.ensure(npad * (size_t)(n_slices + 2) * sizeof(float2)),
                "workspace alloc");
        HIP_TRY(launch_prepare_volume(a1v, a2v, nx, ny, n_slices, (float2 *)ctx->mu_vol.ptr, s),
                "volume preparation");
    }
    HIP_TRY(ctx->s2u8.ensure((size_t)d->lut_nq * d->lut_na), "workspace alloc");
    HIP_TRY(ctx->cpx.ensure(2 * npix * sizeof(float)), "workspace alloc");
    HIP_TRY(ctx->ref_list.ensure(npix * sizeof(int)), "workspace alloc");
    HIP_TRY(ctx->voxels.ensure(npix * sizeof(VoxelEntry)), "workspace alloc");
    HIP_TRY(ctx->offsets.ensure((size_t)(d->n_mat + 1) * sizeof(int)), "workspace alloc");
    HIP_TRY(ctx->src_int.ensure((size_t)V * npix * sizeof(float2)), "workspace alloc");
    ForwardArgs fa;
    rc = setup_forward_common(ctx, g, d, spectrum, fa, s);
    if (rc) return rc;

    PrepArgs pa{};
    pa.nx = nx;
    pa.ny = ny;
    pa.n_mat = d->n_mat;
    pa.flags = d->flags;
    pa.seg = segmentation;
    pa.a1 = a1s;
    pa.a2 = a2s;
    pa.s2_table = s2_table;
    pa.s2_count = d->lut_nq * d->lut_na;
    pa.s2u8 = (uint8_t *)ctx->s2u8.ptr;
    pa.mu_padded = (float2 *)ctx->mu_padded.ptr;
    pa.cpx1 = (float *)ctx->cpx.ptr;
    pa.cpx2 = (float *)ctx->cpx.ptr + npix;
    pa.ref_list = (int *)ctx->ref_list.ptr;
    pa.voxels = (VoxelEntry *)ctx->voxels.ptr;
    pa.mat_offsets = (int *)ctx->offsets.ptr;
    HIP_TRY(launch_prepare(pa, s), "prepare kernels");

    fa.ifine = bt ? nullptr : tab.ifine;
    fa.ifine_b16 = bt ? bt->ifine : nullptr;
    fa.ifine_scale = bt ? bt->ifine_scale : nullptr;
    fa.lut_q_min = lut_q_min;
    fa.lut_a_min = lut_a_min;
    fa.dq_lut = (lut_q_max - lut_q_min) / (double)d->lut_nq;
    fa.da_lut = (lut_a_max - lut_a_min) / (double)d->lut_na;
    fa.s2u8 = (const uint8_t *)ctx->s2u8.ptr;
    fa.mu_padded = (const float2 *)ctx->mu_padded.ptr;
    fa.voxels = (const VoxelEntry *)ctx->voxels.ptr;
    fa.mat_offsets = (const int *)ctx->offsets.ptr;
    fa.src_int = (const float2 *)ctx->src_int.ptr;
    fa.out_n = out_counts;
    fa.out_c = out_compton;
    fa.out_a = out_A;
    fa.nz = n_slices;
    fa.z_slice = (double)slice + 0.5;
    fa.mu_vol = n_slices > 1 ? (const float2 *)ctx->mu_vol.ptr : nullptr;
    HIP_TRY(launch_source_integrals(fa, (float2 *)ctx->src_int.ptr, s), "source-leg kernel");
    // the non-air voxel count sizes the per-(view, pixel, voxel) records: views are
    // processed in batches whose records fit in kPixelVoxelBudget bytes
    if (!ctx->pinned_word)
        HIP_TRY(hipHostMalloc((void **)&ctx->pinned_word, 64, hipHostMallocDefault),
                "pinned alloc");
    HIP_TRY(hipMemcpyAsync(ctx->pinned_word, (const int *)ctx->offsets.ptr + d->n_mat, sizeof(int),
                           hipMemcpyDeviceToHost, s),
            "voxel count copy");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    const int nvox = *ctx->pinned_word;
    const size_t per_view = (size_t)fa.n_det * (size_t)(nvox > 0 ? nvox : 1) * kPixelVoxelRecordBytes;
    int nvb = (int)std::min<size_t>((size_t)V, std::max<size_t>(1, kPixelVoxelBudget / per_view));
    if (ctx->fwd_vpb > 0) nvb = std::min(nvb, ctx->fwd_vpb);
    fa.kc = ctx->fwd_kc;
    const size_t nrec = (size_t)nvb * fa.n_det * (size_t)(nvox > 0 ? nvox : 1);
    HIP_TRY(ctx->pv.ensure(nrec * kPixelVoxelRecordBytes), "workspace alloc");
    fa.nvox = nvox;
    fa.pv_s = (double *)ctx->pv.ptr;
    fa.pv_cth = fa.pv_s + nrec;
    fa.pv_geo = (float4 *)(fa.pv_cth + nrec);
    fa.pv_leg = (float2 *)(fa.pv_geo + nrec);
    for (int v0 = 0; v0 < V; v0 += nvb) {
        fa.v0 = v0;
        fa.nvb = std::min(nvb, V - v0);
        HIP_TRY(launch_pixel_voxel(fa, s), "pixel-voxel kernel");
        hipError_t e = launch_forward(fa, s);
        if (e == hipErrorInvalidValue)
            return fail(XRD_ERR_UNSUPPORTED, "%s: no kernel instantiation for these sizes", fn);
        HIP_TRY(e, "forward kernel");
    }
    return XRD_OK;
}

}  // namespace

extern "C" int xrd_forward(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                           const float *spectrum, const float *ifine,
                           const int32_t *segmentation, const float *a1, const float *a2,
                           const float *s2_table, double lut_q_min, double lut_q_max,
                           double lut_a_min, double lut_a_max, float *out_counts,
                           float *out_compton, float *out_A, void *stream) {
    TableArgs tab;
    tab.ifine = ifine;
    tab.a1 = a1;
    tab.a2 = a2;
    return forward_impl(ctx, g, d, spectrum, tab, segmentation, 1, 0, s2_table, lut_q_min,
                        lut_q_max, lut_a_min, lut_a_max, out_counts, out_compton, out_A, stream,
                        "xrd_forward");
}

extern "C" int xrd_forward_slice(xrd_context *ctx, const xrd_geometry *g,
                                 const xrd_discretization *d, const float *spectrum,
                                 const float *ifine, const int32_t *segmentation, const float *a1,
                                 const float *a2, int32_t n_slices, int32_t slice,
                                 const float *s2_table, double lut_q_min, double lut_q_max,
                                 double lut_a_min, double lut_a_max, float *out_counts,
                                 float *out_compton, float *out_A, void *stream) {
    TableArgs tab;
    tab.ifine = ifine;
    tab.a1 = a1;
    tab.a2 = a2;
    return forward_impl(ctx, g, d, spectrum, tab, segmentation, n_slices, slice, s2_table,
                        lut_q_min, lut_q_max, lut_a_min, lut_a_max, out_counts, out_compton,
                        out_A, stream, "xrd_forward_slice");
}

extern "C" int xrd_forward_bf16(xrd_context *ctx, const xrd_geometry *g,
                                const xrd_discretization *d, const float *spectrum,
                                const xrd_bf16_tables *tables, const int32_t *segmentation,
                                int32_t n_slices, int32_t slice, const float *s2_table,
                                double lut_q_min, double lut_q_max, double lut_a_min,
                                double lut_a_max, float *out_counts, float *out_compton,
                                float *out_A, void *stream) {
    if (!tables) return fail(XRD_ERR_NULL_POINTER, "xrd_forward_bf16: tables is NULL");
    TableArgs tab;
    tab.b16 = tables;
    return forward_impl(ctx, g, d, spectrum, tab, segmentation, n_slices, slice, s2_table,
                        lut_q_min, lut_q_max, lut_a_min, lut_a_max, out_counts, out_compton,
                        out_A, stream, "xrd_forward_bf16");
}

extern "C" int xrd_forward_set_launch(xrd_context *ctx, int32_t voxels_per_chunk,
                                      int32_t views_per_batch) {
    const char *fn = "xrd_forward_set_launch";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (voxels_per_chunk < 0 || voxels_per_chunk > kMaxVoxelsPerChunk || views_per_batch < 0)
        return fail(XRD_ERR_INVALID_ARG,
                    "%s: voxels_per_chunk in [0,%d], views_per_batch >= 0 (0 = default)", fn,
                    kMaxVoxelsPerChunk);
    ctx->fwd_kc = voxels_per_chunk;
    ctx->fwd_vpb = views_per_batch;
    return XRD_OK;
}

extern "C" int xrd_forward_autotune(xrd_context *ctx, const xrd_geometry *g,
                                    const xrd_discretization *d, const float *spectrum,
                                    const float *ifine, const int32_t *segmentation,
                                    const float *a1, const float *a2, const float *s2_table,
                                    double lut_q_min, double lut_q_max, double lut_a_min,
                                    double lut_a_max, int32_t sample_views, void *stream,
                                    int32_t *out_voxels_per_chunk, int32_t *out_views_per_batch,
                                    float *out_ms) {
    const char *fn = "xrd_forward_autotune";
    if (!ctx || !g || !d) return fail(XRD_ERR_NULL_POINTER, "%s: NULL argument", fn);
    if (sample_views <= 0) return fail(XRD_ERR_INVALID_ARG, "%s: sample_views must be > 0", fn);
    int rc = check_geometry(g, d, fn);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    // the first sample_views views of the scan, into scratch outputs
    xrd_geometry gs = *g;
    gs.n_views = std::min(sample_views, g->n_views);
    const size_t rows = (size_t)gs.n_views * g->n_cols * g->n_rows * d->n_energy;
    DevBuf out;
    HIP_TRY(out.ensure(rows * ((size_t)d->n_q * d->n_mat + 2) * sizeof(float)),
            "autotune output alloc");
    float *n_out = (float *)out.ptr, *c_out = n_out + rows, *a_out = c_out + rows;
    TableArgs tab;
    tab.ifine = ifine;
    tab.a1 = a1;
    tab.a2 = a2;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    const int kcs[] = {8, 12, 16};
    const int vpbs[] = {2, 8, 32};
    const int saved_kc = ctx->fwd_kc, saved_vpb = ctx->fwd_vpb;
    float best = 1e30f;
    int best_kc = 16, best_vpb = 0;
    rc = XRD_OK;
    hipError_t err = hipSuccess;
    // one untimed call first: workspace allocation and code loading
    rc = forward_impl(ctx, &gs, d, spectrum, tab, segmentation, 1, 0, s2_table, lut_q_min,
                      lut_q_max, lut_a_min, lut_a_max, n_out, c_out, a_out, stream, fn);
    for (int kc : kcs) {
        for (int vpb : vpbs) {
            if (rc != XRD_OK || err != hipSuccess) break;
            if (vpb > gs.n_views && vpb != vpbs[0]) continue;
            ctx->fwd_kc = kc;
            ctx->fwd_vpb = vpb;
            err = hipEventRecord(e0, s);
            if (err != hipSuccess) break;
            rc = forward_impl(ctx, &gs, d, spectrum, tab, segmentation, 1, 0, s2_table,
                              lut_q_min, lut_q_max, lut_a_min, lut_a_max, n_out, c_out, a_out,
                              stream, fn);
            if (rc != XRD_OK) break;
            err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err == hipSuccess && ms < best) {
                best = ms;
                best_kc = kc;
                best_vpb = vpb;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    out.release();
    if (rc != XRD_OK || err != hipSuccess) {
        ctx->fwd_kc = saved_kc;
        ctx->fwd_vpb = saved_vpb;
        return rc != XRD_OK ? rc : hip_fail(err, fn);
    }
    ctx->fwd_kc = best_kc;
    ctx->fwd_vpb = best_vpb;
    if (out_voxels_per_chunk) *out_voxels_per_chunk = best_kc;
    if (out_views_per_batch) *out_views_per_batch = best_vpb;
    if (out_ms) *out_ms = best;
    return XRD_OK;
}

extern "C" int xrd_trace_rays(xrd_context *ctx, int32_t nx, int32_t ny, double ray_step,
                              int64_t n, const double *seg, int32_t *out_side,
                              double *out_entry, int32_t *out_count, int32_t max_samples,
                              int32_t *out_texels, int32_t n_basis, const float *maps,
                              float *out_integrals, void *stream) {
    const char *fn = "xrd_trace_rays";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (nx <= 0 || ny <= 0 || n < 0) return fail(XRD_ERR_DIMS, "%s: bad sizes", fn);
    if (!std::isfinite(ray_step) || !(ray_step > 0.0))
        return fail(XRD_ERR_INVALID_ARG, "%s: ray_step must be finite and > 0", fn);
    if (n == 0) return XRD_OK;
    if (!seg || !out_side || !out_entry || !out_count)
        return fail(XRD_ERR_NULL_POINTER, "%s: a required pointer is NULL", fn);
    if (out_texels && max_samples <= 0)
        return fail(XRD_ERR_INVALID_ARG, "%s: max_samples must be > 0 with out_texels", fn);
    if (out_integrals && (n_basis <= 0 || !maps))
        return fail(XRD_ERR_INVALID_ARG, "%s: out_integrals needs n_basis > 0 and maps", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(launch_trace_rays(nx, ny, ray_step, n, seg, out_side, out_entry, out_count,
                              out_texels ? max_samples : 0, out_texels,
                              out_integrals ? n_basis : 0, maps, out_integrals,
                              (hipStream_t)stream),
            "trace kernel");
    return XRD_OK;
}

namespace {

int check_matrix(const void *A, int dtype, int64_t n_rows, int32_t n_cols, const char *fn) {
    if (!A) return fail(XRD_ERR_NULL_POINTER, "%s: A is NULL", fn);
    if (dtype != XRD_DTYPE_F32 && dtype != XRD_DTYPE_BF16 && dtype != XRD_DTYPE_F16 &&
        dtype != XRD_DTYPE_BF24)
        return fail(XRD_ERR_INVALID_ARG, "%s: unknown dtype %d", fn, dtype);
    if (n_rows <= 0 || n_cols <= 0) return fail(XRD_ERR_DIMS, "%s: non-positive size", fn);
    if (n_cols % 4 != 0) return fail(XRD_ERR_DIMS, "%s: n_cols must be a multiple of 4", fn);
    if (dtype == XRD_DTYPE_BF24 && n_cols % 8 != 0)
        return fail(XRD_ERR_DIMS, "%s: BF24 needs n_cols a multiple of 8", fn);
    if (n_cols > kMaxMatCols)
        return fail(XRD_ERR_UNSUPPORTED, "%s: n_cols > %d", fn, kMaxMatCols);
    const int align = dtype == XRD_DTYPE_F32 || dtype == XRD_DTYPE_BF24 ? 16 : 8;
    if (!aligned(A, align)) return fail(XRD_ERR_ALIGN, "%s: A must be %d-byte aligned", fn, align);
    return XRD_OK;
}

}  // namespace

extern "C" int xrd_project(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                           int32_t n_cols, const float *x, float *out, void *stream) {
    const char *fn = "xrd_project";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    int rc = check_matrix(A, a_dtype, n_rows, n_cols, fn);
    if (rc) return rc;
    if (!x || !out) return fail(XRD_ERR_NULL_POINTER, "%s: x/out is NULL", fn);
    if (!aligned(x, 16)) return fail(XRD_ERR_ALIGN, "%s: x must be 16-byte aligned", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    MatArgs m{A, a_dtype, n_rows, n_cols};
    HIP_TRY(launch_project(m, x, out, (hipStream_t)stream), "project kernel");
    return XRD_OK;
}

static int colsum_common(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                         int32_t n_cols, const float *w, double *out, void *stream,
                         const char *fn) {
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    int rc = check_matrix(A, a_dtype, n_rows, n_cols, fn);
    if (rc) return rc;
    if (!out) return fail(XRD_ERR_NULL_POINTER, "%s: out is NULL", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    MatArgs m{A, a_dtype, n_rows, n_cols};
    const int nb = mat_num_blocks(m);
    HIP_TRY(ctx->partial.ensure((size_t)nb * n_cols * sizeof(double)), "workspace alloc");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(launch_colsum_partials(m, w, (double *)ctx->partial.ptr, s), "colsum kernel");
    HIP_TRY(launch_reduce_partials((const double *)ctx->partial.ptr, nb, n_cols, out, s),
            "reduce kernel");
    return XRD_OK;
}

extern "C" int xrd_backproject(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                               int32_t n_cols, const float *w, double *out, void *stream) {
    if (!w) return fail(XRD_ERR_NULL_POINTER, "xrd_backproject: w is NULL");
    return colsum_common(ctx, A, a_dtype, n_rows, n_cols, w, out, stream, "xrd_backproject");
}

extern "C" int xrd_sensitivity(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                               int32_t n_cols, double *out, void *stream) {
    return colsum_common(ctx, A, a_dtype, n_rows, n_cols, nullptr, out, stream,
                         "xrd_sensitivity");
}

extern "C" int xrd_convert_matrix(xrd_context *ctx, const float *A, int64_t n, int out_dtype,
                                  void *out, void *stream) {
    const char *fn = "xrd_convert_matrix";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!A || !out) return fail(XRD_ERR_NULL_POINTER, "%s: A/out is NULL", fn);
    if (n < 0) return fail(XRD_ERR_DIMS, "%s: n < 0", fn);
    if (out_dtype != XRD_DTYPE_BF16 && out_dtype != XRD_DTYPE_F16 && out_dtype != XRD_DTYPE_BF24)
        return fail(XRD_ERR_INVALID_ARG, "%s: out_dtype must be BF16, F16 or BF24", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(launch_convert(A, n, out_dtype, out, (hipStream_t)stream), "convert kernel");
    return XRD_OK;
}

extern "C" int xrd_convert_matrix_scaled(xrd_context *ctx, const float *A, int64_t n_rows,
                                         int32_t n_cols, int out_dtype, void *out,
                                         float *col_scale, void *stream) {
    const char *fn = "xrd_convert_matrix_scaled";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!A || !out || !col_scale) return fail(XRD_ERR_NULL_POINTER, "%s: NULL argument", fn);
    if (n_rows <= 0 || n_cols <= 0) return fail(XRD_ERR_DIMS, "%s: non-positive size", fn);
    if (out_dtype != XRD_DTYPE_BF16 && out_dtype != XRD_DTYPE_F16)
        return fail(XRD_ERR_INVALID_ARG, "%s: out_dtype must be BF16 or F16", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    int nb = 1024;
    if (nb > n_rows) nb = (int)n_rows;
    HIP_TRY(ctx->small.ensure((size_t)nb * n_cols * sizeof(float)), "workspace alloc");
    HIP_TRY(launch_column_scale(A, n_rows, n_cols, out_dtype, (float *)ctx->small.ptr, nb,
                                col_scale, s),
            "column-scale kernels");
    HIP_TRY(launch_convert_scaled(A, n_rows, n_cols, out_dtype, col_scale, out, s),
            "convert kernel");
    return XRD_OK;
}

// ---- MLEM -------------------------------------------------------------------------

namespace {

hipError_t mlem_one_iteration(xrd_mlem_plan *p, hipStream_t s) {
    hipError_t e;
    if (p->mode == XRD_MLEM_FUSED) {
        e = launch_mlem_fused(p->m, p->x, p->y, p->c, (double *)p->partial.ptr, s);
        if (e != hipSuccess) return e;
    } else {
        e = launch_project(p->m, p->x, (float *)p->f.ptr, s);
        if (e != hipSuccess) return e;
        e = launch_ratio(p->m.n_rows, (const float *)p->f.ptr, p->y, p->c, (float *)p->r.ptr, s);
        if (e != hipSuccess) return e;
        e = launch_colsum_partials(p->m, (const float *)p->r.ptr, (double *)p->partial.ptr, s);
        if (e != hipSuccess) return e;
    }
    return launch_mlem_update((const double *)p->partial.ptr, p->nblocks, p->m.n_cols,
                              (const double *)p->sens.ptr, p->x, s, p->m.scale);
}

}  // namespace

extern "C" int xrd_mlem_plan_create(xrd_context *ctx, const void *A, int a_dtype, int64_t n_rows,
                                    int32_t n_cols, const float *col_scale, const float *y,
                                    const float *c, float *x, int mode, void *stream,
                                    xrd_mlem_plan **out) {
    const char *fn = "xrd_mlem_plan_create";
    if (!out) return fail(XRD_ERR_NULL_POINTER, "%s: out is NULL", fn);
    *out = nullptr;
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    int rc = check_matrix(A, a_dtype, n_rows, n_cols, fn);
    if (rc) return rc;
    if (!y || !c || !x) return fail(XRD_ERR_NULL_POINTER, "%s: y/c/x is NULL", fn);
    if (!aligned(x, 16)) return fail(XRD_ERR_ALIGN, "%s: x must be 16-byte aligned", fn);
    if (col_scale && !aligned(col_scale, 16))
        return fail(XRD_ERR_ALIGN, "%s: col_scale must be 16-byte aligned", fn);
    if (mode != XRD_MLEM_FUSED && mode != XRD_MLEM_UNFUSED)
        return fail(XRD_ERR_INVALID_ARG, "%s: unknown mode %d", fn, mode);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    xrd_mlem_plan *p = new xrd_mlem_plan();
    p->ctx = ctx;
    p->m = MatArgs{A, a_dtype, n_rows, n_cols, col_scale};
    p->y = y;
    p->c = c;
    p->x = x;
    p->mode = mode;
    p->nblocks = mat_num_blocks(p->m);
    hipStream_t s = (hipStream_t)stream;
    // partials sized for the largest launch shape the autotuner may pick
    MatArgs mmax = p->m;
    mmax.blocks_per_cu = kMaxBlocksPerCu;
    const int nb_max = mat_num_blocks(mmax);
    hipError_t e = p->sens.ensure((size_t)n_cols * sizeof(double));
    if (e == hipSuccess) e = p->partial.ensure((size_t)nb_max * n_cols * sizeof(double));
    if (e == hipSuccess) e = p->llp.ensure((size_t)nb_max * sizeof(double) + 8);
    if (e == hipSuccess && mode == XRD_MLEM_UNFUSED) {
        e = p->f.ensure((size_t)n_rows * sizeof(float));
        if (e == hipSuccess) e = p->r.ensure((size_t)n_rows * sizeof(float));
    }
    // sensitivity sum(A,1)' once, FP64 (main.m:162)
    if (e == hipSuccess) e = launch_colsum_partials(p->m, nullptr, (double *)p->partial.ptr, s);
    if (e == hipSuccess)
        e = launch_reduce_partials((const double *)p->partial.ptr, p->nblocks, n_cols,
                                   (double *)p->sens.ptr, s, col_scale);
    if (e != hipSuccess) {
        xrd_mlem_plan_destroy(p);
        return hip_fail(e, "xrd_mlem_plan_create");
    }
    *out = p;
    return XRD_OK;
}

extern "C" int xrd_mlem_plan_iterate(xrd_mlem_plan *p, int32_t n_iter, int use_graph,
                                     void *stream) {
    const char *fn = "xrd_mlem_plan_iterate";
    if (!p) return fail(XRD_ERR_NULL_POINTER, "%s: plan is NULL", fn);
    if (n_iter < 0) return fail(XRD_ERR_INVALID_ARG, "%s: n_iter < 0", fn);
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        for (int i = 0; i < n_iter; i++) HIP_TRY(mlem_one_iteration(p, s), "mlem iteration");
        return XRD_OK;
    }
    if (!p->exec) {
        // capture one iteration, and kPlanGraphIters iterations, on a private stream
        hipStream_t cs;
        HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
        hipError_t e = hipSuccess;
        for (int g = 0; g < 2 && e == hipSuccess; g++) {
            const int iters = g == 0 ? 1 : kPlanGraphIters;
            hipGraph_t &graph = g == 0 ? p->graph : p->graph_multi;
            hipGraphExec_t &exec = g == 0 ? p->exec : p->exec_multi;
            e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e2 = hipSuccess;
                for (int i = 0; i < iters && e2 == hipSuccess; i++) e2 = mlem_one_iteration(p, cs);
                e = hipStreamEndCapture(cs, &graph);
                if (e2 != hipSuccess) e = e2;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        }
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) {
            plan_drop_graphs(p);
            return hip_fail(e, "graph capture");
        }
    }
    int left = n_iter;
    for (; left >= kPlanGraphIters; left -= kPlanGraphIters)
        HIP_TRY(hipGraphLaunch(p->exec_multi, s), "hipGraphLaunch");
    for (; left > 0; left--) HIP_TRY(hipGraphLaunch(p->exec, s), "hipGraphLaunch");
    return XRD_OK;
}

extern "C" int xrd_mlem_plan_loglik(xrd_mlem_plan *p, double *out_host, void *stream) {
    if (!p || !out_host) return fail(XRD_ERR_NULL_POINTER, "xrd_mlem_plan_loglik: NULL argument");
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(launch_loglik_partials(p->m, p->x, p->y, p->c, (double *)p->llp.ptr, s), "loglik kernel");
    std::vector<double> h(p->nblocks);
    HIP_TRY(hipMemcpyAsync(h.data(), p->llp.ptr, h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    double t = 0.0;
    for (double v : h) t += v;
    *out_host = t;
    return XRD_OK;
}

extern "C" int xrd_mlem_plan_autotune(xrd_mlem_plan *p, int32_t reps, void *stream,
                                      int32_t *out_blocks_per_cu, int32_t *out_rows_per_wave,
                                      float *out_ms) {
    const char *fn = "xrd_mlem_plan_autotune";
    if (!p) return fail(XRD_ERR_NULL_POINTER, "%s: plan is NULL", fn);
    if (reps <= 0) return fail(XRD_ERR_INVALID_ARG, "%s: reps must be > 0", fn);
    if (p->mode != XRD_MLEM_FUSED) return fail(XRD_ERR_UNSUPPORTED, "%s: fused plans only", fn);
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    const int bpcs[] = {1, 2, 3, 4, 6};
    const int rows[] = {1, 2, 4};
    float best = 1e30f;
    int best_b = 2, best_r = 0;
    hipError_t err = hipSuccess;
    for (int b : bpcs) {
        for (int r : rows) {
            MatArgs m = p->m;
            m.blocks_per_cu = b;
            m.rows_per_wave = r;
            // the fused pass only reads x, y, c and writes the partials: x is untouched
            err = launch_mlem_fused(m, p->x, p->y, p->c, (double *)p->partial.ptr, s);
            if (err == hipSuccess) err = hipEventRecord(e0, s);
            for (int i = 0; i < reps && err == hipSuccess; i++)
                err = launch_mlem_fused(m, p->x, p->y, p->c, (double *)p->partial.ptr, s);
            if (err == hipSuccess) err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err != hipSuccess) break;
            ms /= (float)reps;
            if (ms < best) {
                best = ms;
                best_b = b;
                best_r = r;
            }
        }
        if (err != hipSuccess) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return hip_fail(err, fn);
    p->m.blocks_per_cu = best_b;
    p->m.rows_per_wave = best_r;
    p->nblocks = mat_num_blocks(p->m);
    plan_drop_graphs(p);  // the captured iterations used the old launch shape
    if (out_blocks_per_cu) *out_blocks_per_cu = best_b;
    if (out_rows_per_wave) *out_rows_per_wave = best_r;
    if (out_ms) *out_ms = best;
    return XRD_OK;
}

extern "C" int xrd_mlem_plan_time_pass(xrd_mlem_plan *p, int32_t reps, void *stream,
                                       float *out_ms) {
    const char *fn = "xrd_mlem_plan_time_pass";
    if (!p || !out_ms) return fail(XRD_ERR_NULL_POINTER, "%s: NULL argument", fn);
    if (reps <= 0) return fail(XRD_ERR_INVALID_ARG, "%s: reps must be > 0", fn);
    HIP_TRY(hipSetDevice(p->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    hipError_t err = hipEventRecord(e0, s);
    for (int i = 0; i < reps && err == hipSuccess; i++) {
        if (p->mode == XRD_MLEM_FUSED) {
            err = launch_mlem_fused(p->m, p->x, p->y, p->c, (double *)p->partial.ptr, s);
        } else {
            err = launch_project(p->m, p->x, (float *)p->f.ptr, s);
            if (err == hipSuccess)
                err = launch_ratio(p->m.n_rows, (const float *)p->f.ptr, p->y, p->c,
                                   (float *)p->r.ptr, s);
            if (err == hipSuccess)
                err = launch_colsum_partials(p->m, (const float *)p->r.ptr,
                                             (double *)p->partial.ptr, s);
        }
    }
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) return hip_fail(err, fn);
    *out_ms = ms / (float)reps;
    return XRD_OK;
}

extern "C" int xrd_mlem_plan_set_launch(xrd_mlem_plan *p, int32_t blocks_per_cu,
                                        int32_t rows_per_wave) {
    if (!p) return fail(XRD_ERR_NULL_POINTER, "xrd_mlem_plan_set_launch: plan is NULL");
    if (blocks_per_cu < 0 || blocks_per_cu > kMaxBlocksPerCu || rows_per_wave < 0 ||
        rows_per_wave > 4)
        return fail(XRD_ERR_INVALID_ARG, "xrd_mlem_plan_set_launch: blocks_per_cu in [0,%d], "
                    "rows_per_wave in [0,4] (0 = default)", kMaxBlocksPerCu);
    p->m.blocks_per_cu = blocks_per_cu;
    p->m.rows_per_wave = rows_per_wave;
    p->nblocks = mat_num_blocks(p->m);
    plan_drop_graphs(p);
    return XRD_OK;
}

extern "C" int xrd_mlem_iterate_static(xrd_context *ctx, const void *A, int a_dtype,
                                       int64_t n_rows, int32_t n_cols, const float *col_scale,
                                       const float *y, const float *c, float *x,
                                       int32_t n_iter, void *stream) {
    const char *fn = "xrd_mlem_iterate_static";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    int rc = check_matrix(A, a_dtype, n_rows, n_cols, fn);
    if (rc) return rc;
    if (!y || !c || !x) return fail(XRD_ERR_NULL_POINTER, "%s: y/c/x is NULL", fn);
    if (!aligned(x, 16)) return fail(XRD_ERR_ALIGN, "%s: x must be 16-byte aligned", fn);
    if (col_scale && !aligned(col_scale, 16))
        return fail(XRD_ERR_ALIGN, "%s: col_scale must be 16-byte aligned", fn);
    if (n_iter < 0) return fail(XRD_ERR_INVALID_ARG, "%s: n_iter < 0", fn);
    hipError_t st = hipSuccess;
    double *base = scratch_ptr(ctx->device, &st);
    if (!base)
        return fail(XRD_ERR_HIP, "%s: no device scratch on device %d (%s)", fn, ctx->device,
                    hipGetErrorString(st));
    // the plan's default launch shape; sensitivity first, partials 256-byte aligned after it
    const MatArgs m{A, a_dtype, n_rows, n_cols, col_scale};
    const int nb = mat_num_blocks(m);
    const size_t sens_doubles = ((size_t)n_cols + 31) / 32 * 32;
    const size_t need = sens_doubles + (size_t)nb * n_cols;
    if (need > kScratchDoubles)
        return fail(XRD_ERR_DIMS,
                    "%s: %d workgroups x %d columns need %zu bytes of device scratch, more "
                    "than the %d bytes of XRD_MLEM_STATIC_SCRATCH_BYTES",
                    fn, nb, n_cols, need * sizeof(double), XRD_MLEM_STATIC_SCRATCH_BYTES);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    double *sens = base;
    double *partial = base + sens_doubles;
    hipStream_t s = (hipStream_t)stream;
    // every workgroup writes all n_cols of its partial slot, so nothing is cleared
    HIP_TRY(launch_colsum_partials(m, nullptr, partial, s), "sensitivity kernel");
    HIP_TRY(launch_reduce_partials(partial, nb, n_cols, sens, s, col_scale), "reduce kernel");
    for (int i = 0; i < n_iter; i++) {
        HIP_TRY(launch_mlem_fused(m, x, y, c, partial, s), "mlem pass");
        HIP_TRY(launch_mlem_update(partial, nb, n_cols, sens, x, s, col_scale), "mlem update");
    }
    return XRD_OK;
}

extern "C" const double *xrd_mlem_plan_sensitivity(const xrd_mlem_plan *p) {
    return p ? (const double *)p->sens.ptr : nullptr;
}

extern "C" int xrd_mlem_plan_destroy(xrd_mlem_plan *p) {
    if (!p) return XRD_OK;
    (void)hipSetDevice(p->ctx->device);
    plan_drop_graphs(p);
    p->sens.release();
    p->partial.release();
    p->f.release();
    p->r.release();
    p->llp.release();
    delete p;
    return XRD_OK;
}

// ---- differentiable coefficient model ---------------------------------------------

namespace {

int coeff_setup(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                const float *spectrum, const int32_t *segmentation, const float *coef,
                int32_t n_coef, const float *basis, int32_t n_basis, const float *maps,
                const float *beta, CoeffArgs &ca, hipStream_t s, const char *fn) {
    int rc = check_geometry(g, d, fn);
    if (rc) return rc;
    if (!spectrum || !segmentation || !coef || !basis || !maps || !beta)
        return fail(XRD_ERR_NULL_POINTER, "%s: a required input pointer is NULL", fn);
    if (n_coef <= 0 || n_coef > 32 || n_basis <= 0 || n_basis > 64)
        return fail(XRD_ERR_UNSUPPORTED, "%s: need 1 <= n_coef <= 32 and 1 <= n_basis <= 64", fn);
    if (d->n_energy > 64 || d->n_q > 128)
        return fail(XRD_ERR_UNSUPPORTED, "%s: need n_energy <= 64 and n_q <= 128", fn);
    if ((size_t)(g->nx + 2) * (g->ny + 2) * 64 * sizeof(float) >= ((size_t)1 << 32))
        return fail(XRD_ERR_UNSUPPORTED,
                    "%s: the texel-major maps, (nx+2)(ny+2)*64 floats, must stay below 4 GiB", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t npix = (size_t)g->nx * g->ny;
    const size_t npad = (size_t)(g->nx + 2) * (g->ny + 2);
    HIP_TRY(ctx->cm_maps.ensure(npad * n_basis * sizeof(float)), "workspace alloc");
    HIP_TRY(ctx->cm_src.ensure((size_t)g->n_views * npix * n_basis * sizeof(float)), "workspace alloc");
    HIP_TRY(ctx->cm_vlist.ensure((npix + 4) * sizeof(int)), "workspace alloc");
    HIP_TRY(ctx->cm_fq.ensure(npix * kCoeffFqCols * sizeof(float)), "workspace alloc");
    const bool tm = coeff_texel_major(n_basis);
    if (tm) HIP_TRY(ctx->cm_tm.ensure(coeff_texel_floats(g->nx, g->ny) * sizeof(float)), "workspace alloc");
    ForwardArgs fa;
    rc = setup_forward_common(ctx, g, d, spectrum, fa, s);
    if (rc) return rc;
    ca = CoeffArgs{};
    ca.f = fa;
    ca.seg = segmentation;
    ca.coef = coef;
    ca.n_coef = n_coef;
    ca.basis = basis;
    ca.n_basis = n_basis;
    ca.maps = maps;
    ca.beta = beta;
    ca.maps_padded = (float *)ctx->cm_maps.ptr;
    ca.src_ints = (float *)ctx->cm_src.ptr;
    ca.vlist = (int *)ctx->cm_vlist.ptr;
    ca.nvox = (int *)ctx->cm_vlist.ptr + npix;
    ca.nvox_cap = (int)npix;
    ca.maps_tm = tm ? (float *)ctx->cm_tm.ptr : nullptr;
    ca.fq_buf = (float *)ctx->cm_fq.ptr;
    HIP_TRY(launch_coeff_prepare(ca, s), "coefficient-model preparation");
    return XRD_OK;
}

// Buffers of the view slabs of the coefficient model with texel-major maps: per view
// of a slab, the detector legs' line integrals and, for the map VJP, their gamma
// (FP32, voxels x pixels x 64 each), within kSlabBudget bytes (64 GiB: 80 views of a
// 256x256 phantom with 256 pixels, 37 with gamma; fewer, longer slabs keep the VJP
// launch's grid fuller) and within half of the free device memory (counting the
// buffers the context already holds), so a slab leaves room for the caller.  Sets ca.n_vox (the voxel count, read back: this synchronises the
// stream), ca.ld_buf and ca.gam_buf, and views to the slab length; views = 0 when
// the maps are planar, no voxel scatters, one view does not fit or an allocation
// fails (the launches then trace the legs themselves and the map VJP scatters in
// place: the same results up to the atomic order and the last bits of the traces).
int coeff_slab_buffers(xrd_context *ctx, const xrd_geometry *g, CoeffArgs &ca, bool gamma,
                       hipStream_t s, size_t &views) {
    constexpr size_t kSlabBudget = (size_t)64 << 30;
    views = 0;
    // XRD_VJP_SCATTER=inplace: the launches trace the legs themselves and the map VJP
    // scatters in place (the path of problems whose buffers do not fit; for testing)
    const char *mode = std::getenv("XRD_VJP_SCATTER");
    if (!ca.maps_tm || (mode && std::strcmp(mode, "inplace") == 0)) return XRD_OK;
    int nv = 0;
    HIP_TRY(hipMemcpyAsync(&nv, ca.nvox, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    ca.n_vox = nv;
    const size_t floats_per_view = coeff_gam_floats_per_view(nv, g->n_cols * g->n_rows);
    // the gamma buffer also holds the source legs' gamma: nv x 64 doubles per view
    const size_t gsrc_per_view = (size_t)nv * 64 * sizeof(double);
    const size_t per_view =
        floats_per_view * sizeof(float) * (gamma ? 2 : 1) + (gamma ? gsrc_per_view : 0);
    size_t budget = kSlabBudget;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        budget = std::min(budget, (free_b + ctx->cm_gam.bytes + ctx->cm_ld.bytes) / 2);
    if (nv <= 0 || per_view > budget) return XRD_OK;
    size_t n = std::max<size_t>(1, std::min<size_t>((size_t)g->n_views, budget / per_view));
    // XRD_VJP_SLAB_VIEWS=k caps the views per slab, so that a small problem runs in
    // several slabs (for testing)
    if (const char *cap = std::getenv("XRD_VJP_SLAB_VIEWS")) {
        const long k = std::strtol(cap, nullptr, 10);
        if (k > 0) n = std::min<size_t>(n, (size_t)k);
    }
    if (ctx->cm_ld.ensure(n * floats_per_view * sizeof(float)) != hipSuccess ||
        (gamma && ctx->cm_gam.ensure(n * (floats_per_view * sizeof(float) + gsrc_per_view)) !=
                      hipSuccess)) {
        (void)hipGetLastError();  // the failed allocation is not the caller's error
        return XRD_OK;
    }
    ca.ld_buf = (float *)ctx->cm_ld.ptr;
    ca.gam_buf = gamma ? (float *)ctx->cm_gam.ptr : nullptr;
    // after the slab's detector-leg gamma (n * floats_per_view floats, a multiple of
    // 64 floats, so the doubles stay aligned)
    ca.gsrc_buf = gamma ? reinterpret_cast<double *>(ca.gam_buf + n * floats_per_view) : nullptr;
    views = n;
    return XRD_OK;
}

}  // namespace

extern "C" int xrd_forward_coeffs(xrd_context *ctx, const xrd_geometry *g,
                                  const xrd_discretization *d, const float *spectrum,
                                  const int32_t *segmentation, const float *coef,
                                  int32_t n_coef, const float *basis, int32_t n_basis,
                                  const float *maps, const float *beta, float *out_counts,
                                  void *stream) {
    const char *fn = "xrd_forward_coeffs";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!out_counts) return fail(XRD_ERR_NULL_POINTER, "%s: out_counts is NULL", fn);
    hipStream_t s = (hipStream_t)stream;
    CoeffArgs ca;
    int rc = coeff_setup(ctx, g, d, spectrum, segmentation, coef, n_coef, basis, n_basis, maps,
                         beta, ca, s, fn);
    if (rc) return rc;
    ca.out_counts = out_counts;
    size_t views = 0;
    rc = coeff_slab_buffers(ctx, g, ca, false, s, views);
    if (rc) return rc;
    if (views > 0)
        HIP_TRY(launch_coeff_forward_slabs(ca, (int)views, s), "coefficient forward in view slabs");
    else
        HIP_TRY(launch_coeff_forward(ca, s), "coefficient forward kernel");
    return XRD_OK;
}

extern "C" int xrd_vjp_coeffs(xrd_context *ctx, const xrd_geometry *g,
                              const xrd_discretization *d, const float *spectrum,
                              const int32_t *segmentation, const float *coef, int32_t n_coef,
                              const float *basis, int32_t n_basis, const float *maps,
                              const float *beta, const float *grad_counts, double *grad_coef,
                              double *grad_maps, void *stream) {
    const char *fn = "xrd_vjp_coeffs";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!grad_counts) return fail(XRD_ERR_NULL_POINTER, "%s: grad_counts is NULL", fn);
    if (!grad_coef && !grad_maps)
        return fail(XRD_ERR_NULL_POINTER, "%s: grad_coef and grad_maps are both NULL", fn);
    hipStream_t s = (hipStream_t)stream;
    CoeffArgs ca;
    int rc = coeff_setup(ctx, g, d, spectrum, segmentation, coef, n_coef, basis, n_basis, maps,
                         beta, ca, s, fn);
    if (rc) return rc;
    const size_t npix = (size_t)g->nx * g->ny;
    ca.w = grad_counts;
    ca.grad_coef = grad_coef;
    ca.grad_maps = grad_maps;
    if (grad_coef) {
        HIP_TRY(ctx->cm_h.ensure(coeff_hpart_doubles(g->n_views, (int)npix) * sizeof(double)),
                "workspace alloc");
        ca.hpart = (double *)ctx->cm_h.ptr;
        HIP_TRY(launch_zero_f64(grad_coef, npix * n_coef, s), "zero kernel");
    }
    if (grad_maps) {
        if (ca.maps_tm) {
            // accumulate texel-major, written out by the VJP launch
            HIP_TRY(ctx->cm_gtm.ensure(npix * 64 * sizeof(double)), "workspace alloc");
            ca.grad_tm = (double *)ctx->cm_gtm.ptr;
            HIP_TRY(launch_zero_f64(ca.grad_tm, npix * 64, s), "zero kernel");
        } else {
            // planar maps: the atomics add into a context buffer (ordinary device
            // memory, where FP64 atomics are supported whatever memory the caller
            // passes), copied out at the end
            HIP_TRY(ctx->cm_gtm.ensure(npix * n_basis * sizeof(double)), "workspace alloc");
            ca.grad_maps = (double *)ctx->cm_gtm.ptr;
            HIP_TRY(launch_zero_f64(ca.grad_maps, npix * n_basis, s), "zero kernel");
        }
    }
    size_t views = 0;
    rc = coeff_slab_buffers(ctx, g, ca, grad_maps != nullptr, s, views);
    if (rc) return rc;
    if (views > 0) {
        HIP_TRY(launch_coeff_vjp_slabs(ca, (int)views, s), "coefficient VJP in view slabs");
        return XRD_OK;
    }
    HIP_TRY(launch_coeff_vjp(ca, s), "coefficient VJP kernel");
    if (grad_maps && !ca.maps_tm)
        HIP_TRY(hipMemcpyAsync(grad_maps, ca.grad_maps, npix * n_basis * sizeof(double),
                               hipMemcpyDeviceToDevice, s),
                "gradient copy");
    return XRD_OK;
}

namespace {

int ct_setup(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
             const float *spectrum, int32_t n_basis, const float *maps, const float *beta,
             CoeffArgs &ca, hipStream_t s, const char *fn) {
    int rc = check_geometry(g, d, fn);
    if (rc) return rc;
    if (!spectrum || !maps || !beta)
        return fail(XRD_ERR_NULL_POINTER, "%s: a required input pointer is NULL", fn);
    if (n_basis <= 0 || n_basis > 64)
        return fail(XRD_ERR_UNSUPPORTED, "%s: need 1 <= n_basis <= 64", fn);
    if (d->n_energy > 64) return fail(XRD_ERR_UNSUPPORTED, "%s: need n_energy <= 64", fn);
    if ((size_t)(g->nx + 2) * (g->ny + 2) * 64 * sizeof(float) >= ((size_t)1 << 32))
        return fail(XRD_ERR_UNSUPPORTED,
                    "%s: the texel-major maps, (nx+2)(ny+2)*64 floats, must stay below 4 GiB", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    HIP_TRY(ctx->cm_tm.ensure(coeff_texel_floats(g->nx, g->ny) * sizeof(float)), "workspace alloc");
    ForwardArgs fa;
    rc = setup_forward_common(ctx, g, d, spectrum, fa, s);
    if (rc) return rc;
    ca = CoeffArgs{};
    ca.f = fa;
    ca.n_basis = n_basis;
    ca.maps = maps;
    ca.beta = beta;
    ca.maps_tm = (float *)ctx->cm_tm.ptr;
    ca.nvox_cap = g->nx * g->ny;
    HIP_TRY(launch_ct_prepare(ca, s), "CT map preparation");
    return XRD_OK;
}

}  // namespace

extern "C" int xrd_forward_ct(xrd_context *ctx, const xrd_geometry *g,
                              const xrd_discretization *d, const float *spectrum,
                              int32_t n_basis, const float *maps, const float *beta, float *out_t,
                              void *stream) {
    const char *fn = "xrd_forward_ct";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!out_t) return fail(XRD_ERR_NULL_POINTER, "%s: out_t is NULL", fn);
    hipStream_t s = (hipStream_t)stream;
    CoeffArgs ca;
    int rc = ct_setup(ctx, g, d, spectrum, n_basis, maps, beta, ca, s, fn);
    if (rc) return rc;
    ca.out_counts = out_t;
    HIP_TRY(launch_ct(ca, false, s), "CT forward kernel");
    return XRD_OK;
}

extern "C" int xrd_vjp_ct(xrd_context *ctx, const xrd_geometry *g, const xrd_discretization *d,
                          const float *spectrum, int32_t n_basis, const float *maps,
                          const float *beta, const float *grad_t, double *grad_maps,
                          void *stream) {
    const char *fn = "xrd_vjp_ct";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (!grad_t || !grad_maps)
        return fail(XRD_ERR_NULL_POINTER, "%s: grad_t or grad_maps is NULL", fn);
    hipStream_t s = (hipStream_t)stream;
    CoeffArgs ca;
    int rc = ct_setup(ctx, g, d, spectrum, n_basis, maps, beta, ca, s, fn);
    if (rc) return rc;
    const size_t npix = (size_t)g->nx * g->ny;
    HIP_TRY(ctx->cm_gtm.ensure(npix * 64 * sizeof(double)), "workspace alloc");
    ca.grad_tm = (double *)ctx->cm_gtm.ptr;
    HIP_TRY(launch_zero_f64(ca.grad_tm, npix * 64, s), "zero kernel");
    ca.w = grad_t;
    ca.grad_maps = grad_maps;
    HIP_TRY(launch_ct(ca, true, s), "CT VJP kernel");
    return XRD_OK;
}

extern "C" int xrd_poisson_nll(xrd_context *ctx, int64_t n, const float *counts,
                               const float *background, const float *y, float *grad_counts,
                               double *out_loss_host, void *stream) {
    const char *fn = "xrd_poisson_nll";
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (n <= 0) return fail(XRD_ERR_DIMS, "%s: n must be > 0", fn);
    if (!counts || !y || !out_loss_host)
        return fail(XRD_ERR_NULL_POINTER, "%s: counts/y/out_loss_host is NULL", fn);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    const int nb = 256;
    HIP_TRY(ctx->small.ensure(nb * sizeof(double)), "workspace alloc");
    HIP_TRY(launch_poisson_nll(n, counts, background, y, grad_counts, (double *)ctx->small.ptr,
                               nb, s), "nll kernel");
    std::vector<double> h(nb);
    HIP_TRY(hipMemcpyAsync(h.data(), ctx->small.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, s),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    double t = 0.0;
    for (double v : h) t += v;
    *out_loss_host = t;
    return XRD_OK;
}

// ---- batched MLEM -----------------------------------------------------------------

// Chunking of a group: levels < 0 one chunk per workgroup (static); L >= 0 guided
// chunks taken from a counter, L rounds of halving sizes then a final round.  Six
// rounds measured fastest on MI355X over a whole batch (48 BF24 phantoms, one process:
// static 14.15 ms, L = 4: 13.98, L = 6: 13.91, L = 8: 13.97 ms per pass); groups whose
// launch is short can do better with fewer, and the autotuner picks per group among
// kBatchLevelChoices.  The environment variable XRD_BATCH_CHUNK_LEVELS (-1 .. 8) fixes
// every group's levels for batches made afterwards, the autotuner's choice included.
constexpr int kBatchLevelsDefault = 6;
constexpr int kBatchLevelChoices[] = {-1, 2, 6};

// Launch shape of one row-span group (index span_index(span)).
struct BatchShape {
    int blocks_per_cu = 2;
    int rows_per_wave = 2;
    int levels = kBatchLevelsDefault;
};

struct xrd_mlem_batch {
    xrd_context *ctx = nullptr;
    BatchArgs args{};
    BatchShape shape[kNumSpans];
    int esize = 4;
    std::vector<BatchProblem> host_probs;
    DevBuf probs, segs, wg_seg, partial, sens, tickets;
    bool levels_fixed = false;  // set by XRD_BATCH_CHUNK_LEVELS
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    return n;
}

void batch_drop_graph(xrd_mlem_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Row-cost charge of one row's reduction and FP64 ratio, in bytes of A.
constexpr double kBatchRowCharge = 64.0;

// Cumulative cost fractions of the chunk boundaries for G workgroups: levels < 0
// gives G equal chunks; levels = L >= 0 gives G chunks of 1/(2G) of the cost, then
// G of 1/(4G), ..., L rounds, and a last round of G equal chunks of what is left.
std::vector<double> chunk_bounds(int G, int levels) {
    std::vector<double> f{0.0};
    double left = 1.0;
    for (int l = 0; l < levels; l++) {
        const double part = 0.5 * left;
        for (int w = 0; w < G; w++) f.push_back(f.back() + part / G);
        left -= part;
    }
    for (int w = 0; w < G; w++) f.push_back(f.back() + left / G);
    f.back() = 1.0;
    return f;
}

// Cuts the rows of the problems in `members` (laid end to end) into C = bounds.size()
// - 1 contiguous chunks, chunk w ending at cost fraction bounds[w + 1], at multiples
// of 16 rows inside a problem; appends the segments to segs and the C + 1 chunk
// offsets to wg.
void partition_group(const std::vector<BatchProblem> &probs, const std::vector<int> &members,
                     int esize, const std::vector<double> &bounds, std::vector<BatchSegment> &segs,
                     std::vector<int> &wg) {
    const int G = (int)bounds.size() - 1;
    const int P = (int)members.size();
    std::vector<double> cost(P);
    double total = 0.0;
    for (int q = 0; q < P; q++) {
        const BatchProblem &pr = probs[members[q]];
        cost[q] = (double)pr.n_rows * ((double)pr.n_cols * esize + kBatchRowCharge);
        total += cost[q];
    }
    // boundary w = (member, row) where workgroup w starts
    std::vector<int> bk(G + 1);
    std::vector<int64_t> br(G + 1);
    bk[0] = 0;
    br[0] = 0;
    bk[G] = P;
    br[G] = 0;
    int k = 0;
    double cum = 0.0;
    for (int w = 1; w < G; w++) {
        const double t = total * bounds[w];
        while (k < P && cum + cost[k] <= t) cum += cost[k++];
        if (k == P) {
            bk[w] = P;
            br[w] = 0;
            continue;
        }
        const int64_t n_rows = probs[members[k]].n_rows;
        int64_t r = (int64_t)((t - cum) / (cost[k] / (double)n_rows));
        r -= r % 16;
        if (r > n_rows) r = n_rows;
        if (bk[w - 1] == k && r < br[w - 1]) r = br[w - 1];  // keep boundaries monotone
        bk[w] = k;
        br[w] = r;
    }
    for (int w = 0; w < G; w++) {
        wg.push_back((int)segs.size());
        int k0 = bk[w];
        int64_t r0 = br[w];
        const int k1 = bk[w + 1];
        const int64_t r1 = br[w + 1];
        while (k0 < k1 || (k0 == k1 && k0 < P && r0 < r1)) {
            const int64_t end = k0 < k1 ? probs[members[k0]].n_rows : r1;
            if (end > r0) {
                BatchSegment sg{};
                sg.problem = members[k0];
                sg.r0 = r0;
                sg.r1 = end;
                segs.push_back(sg);
            }
            if (k0 == k1) break;
            k0++;
            r0 = 0;
        }
    }
    wg.push_back((int)segs.size());
}

// Groups the problems by row span, sizes each group's persistent grid
// (blocks_per_cu per CU, never more than fit on a CU at once: a workgroup that
// waits for a free CU would finish its chunk after all others), partitions every
// group and uploads the segment lists with the problem descriptors.
hipError_t batch_set_grid(xrd_mlem_batch *b) {
    std::vector<BatchProblem> &probs = b->host_probs;
    BatchArgs &a = b->args;
    std::vector<BatchSegment> segs;
    std::vector<int> wg;
    a.n_groups = 0;
    const int cus = device_cus();
    for (int si = 0; si < kNumSpans; si++) {
        const int span = span_of(si);
        std::vector<int> members;
        for (int k = 0; k < (int)probs.size(); k++)
            if (probs[k].span == span) members.push_back(k);
        if (members.empty()) continue;
        const BatchShape &sh = b->shape[si];
        int resident = 0;
        hipError_t e = batch_resident_blocks(a, span, sh.rows_per_wave, &resident);
        if (e != hipSuccess) return e;
        int bpc = sh.blocks_per_cu;
        if (resident > 0 && bpc > resident) bpc = resident;
        BatchGroup &gr = a.groups[a.n_groups++];
        gr.span = span;
        gr.rows_per_wave = sh.rows_per_wave;
        gr.grid = cus * bpc;
        gr.wg_seg_offset = (int)wg.size();
        gr.dynamic = sh.levels >= 0 ? 1 : 0;
        const std::vector<double> bounds = chunk_bounds(gr.grid, sh.levels);
        gr.n_chunks = (int)bounds.size() - 1;
        partition_group(probs, members, b->esize, bounds, segs, wg);
    }
    for (auto &p : probs) {
        p.seg_begin = 0;
        p.seg_end = 0;
    }
    for (int i = (int)segs.size() - 1; i >= 0; i--) probs[segs[i].problem].seg_begin = i;
    for (int i = 0; i < (int)segs.size(); i++) probs[segs[i].problem].seg_end = i + 1;
    hipError_t e = b->segs.ensure(segs.size() * sizeof(BatchSegment));
    if (e == hipSuccess) e = b->wg_seg.ensure(wg.size() * sizeof(int));
    if (e == hipSuccess) e = b->partial.ensure(segs.size() * (size_t)a.stride * sizeof(double));
    // descriptors are small: synchronous copies from pageable memory
    if (e == hipSuccess)
        e = hipMemcpy(b->segs.ptr, segs.data(), segs.size() * sizeof(BatchSegment),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->wg_seg.ptr, wg.data(), wg.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->probs.ptr, probs.data(), probs.size() * sizeof(BatchProblem),
                      hipMemcpyHostToDevice);
    a.probs = (const BatchProblem *)b->probs.ptr;
    a.segs = (const BatchSegment *)b->segs.ptr;
    a.wg_seg = (const int *)b->wg_seg.ptr;
    a.tickets = (int *)b->tickets.ptr;
    a.n_segs = (int)segs.size();
    a.partial = (double *)b->partial.ptr;
    return e;
}

hipError_t batch_iteration(xrd_mlem_batch *b, hipStream_t s) {
    hipError_t e = launch_batch_stream(b->args, true, s);
    if (e != hipSuccess) return e;
    return launch_batch_update(b->args, nullptr, true, s);
}

}  // namespace

extern "C" int xrd_mlem_batch_create(xrd_context *ctx, int32_t n_problems, int a_dtype,
                                     const void *const *A, const int64_t *n_rows,
                                     const int32_t *n_cols, const float *const *col_scale,
                                     const float *const *y, const float *const *c,
                                     float *const *x, int32_t reserved, void *stream,
                                     xrd_mlem_batch **out) {
    const char *fn = "xrd_mlem_batch_create";
    if (!out) return fail(XRD_ERR_NULL_POINTER, "%s: out is NULL", fn);
    *out = nullptr;
    if (!ctx) return fail(XRD_ERR_NULL_POINTER, "%s: context is NULL", fn);
    if (n_problems <= 0) return fail(XRD_ERR_DIMS, "%s: n_problems must be > 0", fn);
    if (!A || !n_rows || !n_cols || !y || !c || !x)
        return fail(XRD_ERR_NULL_POINTER, "%s: a required host array is NULL", fn);
    if (reserved != 0) return fail(XRD_ERR_INVALID_ARG, "%s: reserved must be 0", fn);
    std::vector<BatchProblem> probs(n_problems);
    int max_cols = 0;
    int vec8 = 1;
    for (int k = 0; k < n_problems; k++) {
        int rc = check_matrix(A[k], a_dtype, n_rows[k], n_cols[k], fn);
        if (rc) return rc;
        if (n_cols[k] % 8 != 0 || !aligned(A[k], 16)) vec8 = 0;
        if (!y[k] || !c[k] || !x[k]) return fail(XRD_ERR_NULL_POINTER, "%s: y/c/x[%d] is NULL", fn, k);
        if (!aligned(x[k], 16)) return fail(XRD_ERR_ALIGN, "%s: x[%d] misaligned", fn, k);
        if (col_scale && col_scale[k] && !aligned(col_scale[k], 16))
            return fail(XRD_ERR_ALIGN, "%s: col_scale[%d] misaligned", fn, k);
        BatchProblem &p = probs[k];
        p.A = A[k];
        p.n_rows = n_rows[k];
        p.n_cols = n_cols[k];
        p.span = row_span(n_cols[k]);
        p.y = y[k];
        p.c = c[k];
        p.x = x[k];
        p.scale = col_scale ? col_scale[k] : nullptr;
        if (n_cols[k] > max_cols) max_cols = n_cols[k];
    }
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    xrd_mlem_batch *b = new xrd_mlem_batch();
    b->ctx = ctx;
    b->esize = a_dtype == XRD_DTYPE_F32 ? 4 : (a_dtype == XRD_DTYPE_BF24 ? 3 : 2);
    size_t sens_total = 0;
    for (auto &p : probs) {
        p.sens_offset = (int)sens_total;
        sens_total += (size_t)p.n_cols;
    }
    if (const char *lv = std::getenv("XRD_BATCH_CHUNK_LEVELS")) {
        const int v = std::atoi(lv);
        for (auto &sh : b->shape) sh.levels = v < -1 ? -1 : (v > 8 ? 8 : v);
        b->levels_fixed = true;
    }
    hipError_t e = b->sens.ensure(sens_total * sizeof(double));
    if (e == hipSuccess) e = b->probs.ensure(probs.size() * sizeof(BatchProblem));
    // the chunk counters start at zero; every dynamic launch leaves them at zero
    if (e == hipSuccess) e = b->tickets.ensure(2 * kNumSpans * sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(b->tickets.ptr, 0, 2 * kNumSpans * sizeof(int), s);
    if (e == hipSuccess)
        for (auto &p : probs) p.sens = (const double *)b->sens.ptr + p.sens_offset;
    b->host_probs = probs;
    BatchArgs &a = b->args;
    a.n_problems = n_problems;
    a.max_cols = max_cols;
    a.stride = max_cols;
    a.dtype = a_dtype;
    a.vec8 = vec8;
    if (e == hipSuccess) e = batch_set_grid(b);
    // sensitivities sum(A_k,1)' once, FP64
    if (e == hipSuccess) e = launch_batch_stream(a, false, s);
    if (e == hipSuccess) e = launch_batch_update(a, (double *)b->sens.ptr, false, s);
    if (e != hipSuccess) {
        xrd_mlem_batch_destroy(b);
        return hip_fail(e, fn);
    }
    *out = b;
    return XRD_OK;
}

extern "C" int xrd_mlem_batch_iterate(xrd_mlem_batch *b, int32_t n_iter, int use_graph,
                                      void *stream) {
    const char *fn = "xrd_mlem_batch_iterate";
    if (!b) return fail(XRD_ERR_NULL_POINTER, "%s: batch is NULL", fn);
    if (n_iter < 0) return fail(XRD_ERR_INVALID_ARG, "%s: n_iter < 0", fn);
    HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) {
        for (int i = 0; i < n_iter; i++) HIP_TRY(batch_iteration(b, s), "batched iteration");
        return XRD_OK;
    }
    if (!b->exec) {
        hipStream_t cs;
        HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t e2 = batch_iteration(b, cs);
            e = hipStreamEndCapture(cs, &b->graph);
            if (e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
    }
    for (int i = 0; i < n_iter; i++) HIP_TRY(hipGraphLaunch(b->exec, s), "hipGraphLaunch");
    return XRD_OK;
}

extern "C" int xrd_mlem_batch_set_launch(xrd_mlem_batch *b, int32_t blocks_per_cu,
                                         int32_t rows_per_wave) {
    if (!b) return fail(XRD_ERR_NULL_POINTER, "xrd_mlem_batch_set_launch: batch is NULL");
    if (blocks_per_cu < 0 || blocks_per_cu > kMaxBlocksPerCu || rows_per_wave < 0 ||
        rows_per_wave > 4)
        return fail(XRD_ERR_INVALID_ARG, "xrd_mlem_batch_set_launch: blocks_per_cu in [0,%d], "
                    "rows_per_wave in [0,4]", kMaxBlocksPerCu);
    HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
    // descriptors are replaced below: wait for passes that still read them
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    batch_drop_graph(b);
    for (auto &sh : b->shape) {
        sh.blocks_per_cu = blocks_per_cu > 0 ? blocks_per_cu : 2;
        sh.rows_per_wave = rows_per_wave > 0 ? rows_per_wave : 2;
    }
    HIP_TRY(batch_set_grid(b), "batch partition");
    return XRD_OK;
}

namespace {

// Mean duration (ms) of reps streaming passes of the batch at its current shape
// (group >= 0: that group's launch only).
hipError_t batch_time_stream(xrd_mlem_batch *b, int reps, hipStream_t s, float *ms,
                             int group = -1) {
    hipEvent_t e0, e1;
    hipError_t err = hipEventCreate(&e0);
    if (err != hipSuccess) return err;
    err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return err;
    }
    // the streaming pass reads x, y, c and writes the partials only: x is untouched
    err = launch_batch_stream(b->args, true, s, group);
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    for (int i = 0; i < reps && err == hipSuccess; i++)
        err = launch_batch_stream(b->args, true, s, group);
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float t = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
    *ms = t / (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return err;
}

}  // namespace

extern "C" int xrd_mlem_batch_autotune(xrd_mlem_batch *b, int32_t reps, void *stream,
                                       int32_t *out_blocks_per_cu, int32_t *out_rows_per_wave,
                                       float *out_ms) {
    const char *fn = "xrd_mlem_batch_autotune";
    if (!b) return fail(XRD_ERR_NULL_POINTER, "%s: batch is NULL", fn);
    if (reps <= 0) return fail(XRD_ERR_INVALID_ARG, "%s: reps must be > 0", fn);
    HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    batch_drop_graph(b);
    // each row-span group is launched on its own, so each gets its own shape
    const int bpcs[] = {1, 2, 3, 4};
    const int rows[] = {1, 2, 4};
    for (int g = 0; g < b->args.n_groups; g++) {
        const int si = span_index(b->args.groups[g].span);
        BatchShape best_shape = b->shape[si];
        float best = 1e30f;
        const int n_lv = b->levels_fixed ? 1 : (int)(sizeof(kBatchLevelChoices) / sizeof(int));
        for (int li = 0; li < n_lv; li++) {
            if (!b->levels_fixed) b->shape[si].levels = kBatchLevelChoices[li];
            for (int r : rows) {
                for (int bpc : bpcs) {
                    b->shape[si].rows_per_wave = r;
                    b->shape[si].blocks_per_cu = bpc;
                    HIP_TRY(batch_set_grid(b), "batch partition");
                    float ms = 0.f;
                    HIP_TRY(batch_time_stream(b, reps, s, &ms, g), "timed pass");
                    if (ms < best) {
                        best = ms;
                        best_shape = b->shape[si];
                    }
                }
            }
        }
        b->shape[si] = best_shape;
    }
    HIP_TRY(batch_set_grid(b), "batch partition");
    float total = 0.f;
    HIP_TRY(batch_time_stream(b, reps, s, &total), "timed pass");
    // reported: the shape of the group with the most rows-bytes (the first group if
    // there is one only)
    int main_g = 0;
    double most = -1.0;
    for (int g = 0; g < b->args.n_groups; g++) {
        double cost = 0.0;
        for (const auto &p : b->host_probs)
            if (p.span == b->args.groups[g].span) cost += (double)p.n_rows * p.n_cols;
        if (cost > most) {
            most = cost;
            main_g = g;
        }
    }
    const int si = b->args.n_groups > 0 ? span_index(b->args.groups[main_g].span) : 0;
    if (out_blocks_per_cu) *out_blocks_per_cu = b->shape[si].blocks_per_cu;
    if (out_rows_per_wave) *out_rows_per_wave = b->shape[si].rows_per_wave;
    if (out_ms) *out_ms = total;
    return XRD_OK;
}

extern "C" int xrd_mlem_batch_groups(const xrd_mlem_batch *b, int32_t max_groups,
                                     int32_t *out_span, int32_t *out_grid,
                                     int32_t *out_rows_per_wave, int32_t *out_chunks) {
    if (!b) return fail(XRD_ERR_NULL_POINTER, "xrd_mlem_batch_groups: batch is NULL");
    const int n = b->args.n_groups;
    for (int g = 0; g < n && g < max_groups; g++) {
        if (out_span) out_span[g] = b->args.groups[g].span;
        if (out_grid) out_grid[g] = b->args.groups[g].grid;
        if (out_rows_per_wave) out_rows_per_wave[g] = b->args.groups[g].rows_per_wave;
        if (out_chunks) out_chunks[g] = b->args.groups[g].n_chunks;
    }
    return n;
}

extern "C" int xrd_mlem_batch_time_pass(xrd_mlem_batch *b, int32_t reps, void *stream,
                                        float *out_ms) {
    const char *fn = "xrd_mlem_batch_time_pass";
    if (!b || !out_ms) return fail(XRD_ERR_NULL_POINTER, "%s: NULL argument", fn);
    if (reps <= 0) return fail(XRD_ERR_INVALID_ARG, "%s: reps must be > 0", fn);
    HIP_TRY(hipSetDevice(b->ctx->device), "hipSetDevice");
    HIP_TRY(batch_time_stream(b, reps, (hipStream_t)stream, out_ms), "timed pass");
    return XRD_OK;
}

extern "C" int xrd_mlem_batch_destroy(xrd_mlem_batch *b) {
    if (!b) return XRD_OK;
    (void)hipSetDevice(b->ctx->device);
    batch_drop_graph(b);
    b->probs.release();
    b->segs.release();
    b->wg_seg.release();
    b->partial.release();
    b->sens.release();
    b->tickets.release();
    delete b;
    return XRD_OK;
}

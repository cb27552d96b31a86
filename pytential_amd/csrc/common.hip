// Context, error state, device primitives (scan / sort / compaction) and the
// small elementwise kernels at the boundary (source reorder, upsampling).
#include <hipcub/hipcub.hpp>

#include <array>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "qbx_internal.h"

#include <algorithm>

namespace qbx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int side_streams(int device, hipStream_t out[3])
{
    static std::mutex mu;
    static std::map<int, std::array<hipStream_t, 3>> pool;
    std::lock_guard<std::mutex> lock(mu);
    auto it = pool.find(device);
    if (it == pool.end()) {
        std::array<hipStream_t, 3> s{};
        for (hipStream_t &x : s) QBX_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        it = pool.emplace(device, s).first;
    }
    for (int i = 0; i < 3; ++i) out[i] = it->second[i];
    return QBX_OK;
}

int note_capture(qbx_ctx *ctx, hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    QBX_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) ctx->graph_held = true;
    return QBX_OK;
}

int may_replace(qbx_ctx *ctx, hipStream_t st, const char *what)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    QBX_HIP(hipStreamIsCapturing(st, &cs));
    QBX_CHECK(!ctx->graph_held && cs == hipStreamCaptureStatusNone,
              "%s: the context's %s would be reallocated or rebuilt, but a HIP graph "
              "captured through this context holds the current ones; give each captured "
              "operator its own context (qbx_ctx_create), or warm it up before capturing "
              "and call qbx_ctx_release_graphs after destroying its graphs",
              "qbx context", what);
    return QBX_OK;
}

int scratch_reserve(qbx_ctx *ctx, hipStream_t st, size_t bytes, void **p)
{
    QBX_TRY(note_capture(ctx, st));
    if (bytes > ctx->scratch_bytes) {
        QBX_TRY(may_replace(ctx, st, "scratch buffer"));
        if (ctx->scratch) {
            QBX_HIP(hipDeviceSynchronize());
            QBX_HIP(hipFree(ctx->scratch));
            ctx->scratch = nullptr;
        }
        size_t nb = bytes + bytes / 4 + 4096;
        QBX_HIP(hipMalloc(&ctx->scratch, nb));
        ctx->scratch_bytes = nb;
    }
    *p = ctx->scratch;
    return QBX_OK;
}

int exclusive_scan(qbx_ctx *ctx, hipStream_t st, const int *in, int *out, int n, int *total)
{
    if (n == 0) {
        QBX_HIP(hipMemsetAsync(out, 0, sizeof(int), st));
        if (total) *total = 0;
        return QBX_OK;
    }
    size_t tb = 0;
    QBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, st));
    void *tmp;
    QBX_TRY(scratch_reserve(ctx, st, tb, &tmp));
    QBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, st));
    // out[n] = out[n-1] + in[n-1]
    int last_out = 0, last_in = 0;
    QBX_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    int tot = last_out + last_in;
    QBX_HIP(hipMemcpyAsync(out + n, &tot, sizeof(int), hipMemcpyHostToDevice, st));
    QBX_HIP(hipStreamSynchronize(st));
    if (total) *total = tot;
    return QBX_OK;
}

int sort_pairs_u64(qbx_ctx *ctx, hipStream_t st, const unsigned long long *kin,
                   unsigned long long *kout, const int *vin, int *vout, int n, int end_bit)
{
    if (n == 0) return QBX_OK;
    size_t tb = 0;
    QBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, n, 0,
                                                end_bit, st));
    void *tmp;
    QBX_TRY(scratch_reserve(ctx, st, tb, &tmp));
    QBX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, n, 0, end_bit,
                                                st));
    return QBX_OK;
}

int sort_pairs_i32(qbx_ctx *ctx, hipStream_t st, const int *kin, int *kout, const int *vin,
                   int *vout, int n, int end_bit)
{
    if (n == 0) return QBX_OK;
    size_t tb = 0;
    QBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, n, 0,
                                                end_bit, st));
    void *tmp;
    QBX_TRY(scratch_reserve(ctx, st, tb, &tmp));
    QBX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, vin, vout, n, 0, end_bit,
                                                st));
    return QBX_OK;
}

__global__ void k_scatter_flagged(const int *flags, const int *pos, int n, int *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = i;
}

int compact_indices(qbx_ctx *ctx, hipStream_t st, const int *flags, int n, dbuf<int> &out,
                    int *count)
{
    dbuf<int> pos;
    QBX_TRY(pos.alloc(n + 1));
    int tot = 0;
    QBX_TRY(exclusive_scan(ctx, st, flags, pos.p, n, &tot));
    QBX_TRY(out.alloc(tot > 0 ? tot : 1));
    out.n = tot;
    if (n > 0) {
        hipLaunchKernelGGL(k_scatter_flagged, dim3(cdiv(n, 256)), dim3(256), 0, st, flags,
                           pos.p, n, out.p);
        QBX_HIP(hipGetLastError());
    }
    QBX_HIP(hipStreamSynchronize(st));
    *count = tot;
    return QBX_OK;
}

// zero fill by a kernel (a memset node recorded into a captured graph is not
// reliable on this platform, so the stages clear their outputs this way)
__global__ void k_zero(double *p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = i; k < n; k += stride) p[k] = 0.0;
}

__global__ void k_add(double *dst, const double *src, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long k = i; k < n; k += stride) dst[k] += src[k];
}

int add_into(double *dst, const double *src, long ndoubles, hipStream_t st)
{
    if (ndoubles <= 0) return QBX_OK;
    const long blocks = std::min<long>(cdiv(ndoubles, 256), 8192);
    hipLaunchKernelGGL(k_add, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, ndoubles);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int zero_fill(double *p, long ndoubles, hipStream_t st)
{
    if (ndoubles <= 0) return QBX_OK;
    const long blocks = std::min<long>(cdiv(ndoubles, 256), 8192);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, ndoubles);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

__global__ void k_set_int(int *p, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// clears *flag when any imaginary part is nonzero (every writer stores 0)
__global__ void k_imag_any(const double *z, long n, int *flag)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    bool nz = false;
    for (long k = i; k < n; k += stride) nz |= z[2 * k + 1] != 0.0;
    if (__ballot(nz) != 0 && (threadIdx.x & 63) == 0) *flag = 0;
}

int strengths_all_real(const double *z, long n, int *flag, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, flag, 1);
    if (n > 0) {
        const long blocks = std::min<long>(cdiv(n, 256), 4096);
        hipLaunchKernelGGL(k_imag_any, dim3((unsigned)blocks), dim3(256), 0, st, z, n, flag);
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// one thread per (tree source, right-hand side); density [nsources][nrhs]
__global__ void k_reorder_sources(int ns, int nr, int dim, const int *user_source_ids,
                                  const double *density, const double *weights,
                                  const double *normals, double *str, double *nrm)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ns * nr) return;
    const int i = (int)(idx / nr), r = (int)(idx % nr);
    const int u = user_source_ids[i];
    const double w = weights ? weights[u] : 1.0;
    str[2 * idx] = w * density[2 * ((size_t)u * nr + r)];
    str[2 * idx + 1] = w * density[2 * ((size_t)u * nr + r) + 1];
    if (r == 0 && normals && nrm)
        for (int d = 0; d < dim; ++d) nrm[(size_t)d * ns + i] = normals[(size_t)d * ns + u];
}

// one thread per (element, fine node, right-hand side); in [nel*n1][nrhs]
__global__ void k_upsample(int nel, int n1, int nf, int nr, const double *mat, const double *in,
                           double *out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nel * nf * nr) return;
    const long en = idx / nr;
    const int r = (int)(idx % nr);
    const int e = (int)(en / nf), i = (int)(en % nf);
    double re = 0, im = 0;
    for (int j = 0; j < n1; ++j) {
        const double m = mat[i * n1 + j];
        const size_t k = ((size_t)e * n1 + j) * nr + r;
        re += m * in[2 * k];
        im += m * in[2 * k + 1];
    }
    out[2 * idx] = re;
    out[2 * idx + 1] = im;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

const char *qbx_last_error(void) { return g_err; }

int qbx_version(void) { return 1; }

int qbx_ctx_create(qbx_ctx **ctx, int device)
{
    QBX_CHECK(ctx, "qbx_ctx_create: null output");
    QBX_HIP(hipSetDevice(device));
    double *scr = nullptr;
    QBX_TRY(noalloc_scratch_address(device, &scr));
    qbx_ctx *c = new qbx_ctx();
    c->device = device;
    c->noalloc_scratch = scr;
    *ctx = c;
    return QBX_OK;
}

int qbx_ctx_set_tuning(qbx_ctx *ctx, int param, int value)
{
    QBX_CHECK(ctx, "qbx_ctx_set_tuning: null context");
    QBX_CHECK(param >= 0 && param < QBX_TUNE_COUNT, "qbx_ctx_set_tuning: unknown knob %d",
              param);
    static const int lo[QBX_TUNE_COUNT] = {16, 1, 4, 1, 1, 1};
    static const int hi[QBX_TUNE_COUNT] = {1024, 64, 512, 64, 8, 4};
    QBX_CHECK(value >= lo[param] && value <= hi[param] && (value & (value - 1)) == 0 &&
                  (param != QBX_TUNE_DRIVE_STREAMS || value != 2),
              "qbx_ctx_set_tuning: knob %d takes a power of two in [%d, %d] (got %d)", param,
              lo[param], hi[param], value);
    ctx->tune[param] = value;
    return QBX_OK;
}

int qbx_ctx_get_tuning(qbx_ctx *ctx, int param, int *value)
{
    QBX_CHECK(ctx && value, "qbx_ctx_get_tuning: null argument");
    QBX_CHECK(param >= 0 && param < QBX_TUNE_COUNT, "qbx_ctx_get_tuning: unknown knob %d",
              param);
    *value = ctx->tune[param];
    return QBX_OK;
}

int qbx_ctx_release_graphs(qbx_ctx *ctx)
{
    QBX_CHECK(ctx, "qbx_ctx_release_graphs: null context");
    ctx->graph_held = false;
    return QBX_OK;
}

int qbx_ctx_destroy(qbx_ctx *ctx)
{
    if (!ctx) return QBX_OK;
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (hipEvent_t &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
    return QBX_OK;
}

int qbx_reorder_sources(qbx_ctx *ctx, void *stream, const qbx_tree *tree, int nrhs,
                        const double *density, const double *weights,
                        const double *normals, double *strengths_tree, double *normals_tree)
{
    QBX_CHECK(ctx && tree && density && strengths_tree, "qbx_reorder_sources: null argument");
    QBX_CHECK(nrhs >= 1, "qbx_reorder_sources: nrhs must be >= 1");
    int ns = tree->nsources;
    if (ns == 0) return QBX_OK;
    hipLaunchKernelGGL(k_reorder_sources, dim3(cdiv((long)ns * nrhs, 256)), dim3(256), 0,
                       S(stream), ns, nrhs, tree->dim, tree->user_source_ids.p, density, weights,
                       normals, strengths_tree, normals_tree);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_upsample(qbx_ctx *ctx, void *stream, int nelements, int n1, int nf, int nrhs,
                 const double *mat, const double *in, double *out)
{
    QBX_CHECK(ctx && mat && in && out, "qbx_upsample: null argument");
    QBX_CHECK(nelements >= 0 && n1 > 0 && nf > 0 && nrhs >= 1, "qbx_upsample: bad sizes");
    long n = (long)nelements * nf * nrhs;
    if (n == 0) return QBX_OK;
    hipLaunchKernelGGL(k_upsample, dim3(cdiv(n, 256)), dim3(256), 0, S(stream), nelements, n1,
                       nf, nrhs, mat, in, out);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // extern "C"

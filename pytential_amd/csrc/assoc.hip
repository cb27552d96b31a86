// QBX target/center association on the GPU (pytential/qbx/target_assoc.py).
//
// The association is defined by predicates (the reference's area queries only
// accelerate them, target_assoc.py:185-367):
//   marked(t)   <=> some stage-1 node s has |t-s|^2 <= tunnel(s)^2
//   candidate c <=> side(c)*flag(t) >= 0 and |t-c|^2 <= R_c^2,
//                   side(c) = 2 (c mod 2) - 1, R_c = r_c (1 + tol)
//   per side the candidate with the smallest |t-c|^2 / R_c^2 (ties: lowest c);
//   the plus side wins iff found and (minus not found or d2_plus < d2_minus).
// Candidates are enumerated by walking one quadtree over the stage-1 nodes and
// one over the centers, pruning boxes whose distance from t exceeds the
// largest radius stored below them.  Compiled with -ffp-contract=off.
#include "qbx_internal.h"

namespace qbx {

struct RTree {
    int dim, nboxes, n;
    double mn[3], R;
    const int *level, *child, *coords, *starts, *counts_cumul;
    const double *pts;      // tree order SoA
    const double *rad;      // tree order
    const int *user_ids;    // tree -> user
    const double *boxmax;   // per box max radius in subtree
};

__global__ void k_gather_radius(int n, const int *perm, const double *r, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = r[perm[i]];
}

__global__ void k_leaf_max(int nb, const int *starts, const int *counts, const double *rad,
                           unsigned long long *boxmax)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double m = 0.0;
    for (int i = starts[b]; i < starts[b] + counts[b]; ++i) m = fmax(m, rad[i]);
    boxmax[b] = __double_as_longlong(m);
}

__global__ void k_max_up(int lo, int hi, const int *parent, unsigned long long *boxmax)
{
    int b = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (b < hi) atomicMax(&boxmax[parent[b]], boxmax[b]);
}

// squared distance from point p to box b (computed box bounds, slightly grown)
__device__ double box_dist2(const RTree &t, const double *p, int b)
{
    double h = ldexp(t.R, -t.level[b]);
    double s = 0.0;
    for (int d = 0; d < t.dim; ++d) {
        double lo = t.mn[d] + (double)t.coords[(size_t)d * t.nboxes + b] * h;
        double hi = lo + h;
        double g = 0.0;
        if (p[d] < lo) g = lo - p[d];
        else if (p[d] > hi) g = p[d] - hi;
        s += g * g;
    }
    return s;
}

__device__ bool box_may_hit(const RTree &t, const double *p, int b)
{
    double m = t.boxmax[b];
    double lim = m * (1.0 + 1e-9) + 1e-12 * t.R;
    return box_dist2(t, p, b) <= lim * lim;
}

__device__ double pdist2(int dim, const double *p, const double *x, int n, int i)
{
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        double dd = p[d] - x[(size_t)d * n + i];
        s = s + dd * dd;
    }
    return s;
}

constexpr int ASTACK = 256;

__global__ void k_assoc(RTree ts, RTree tc, int ntargets, const double *targets,
                        const int *flags, int *ttc, int *status, int *nfailed, int *err)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntargets) return;
    int dim = ts.dim;
    double p[3];
    for (int d = 0; d < dim; ++d) p[d] = targets[(size_t)d * ntargets + it];
    int stack[ASTACK];
    bool marked = false;
    int nchild = 1 << dim;
    // ---- mark (QBX_TARGET_MARKER)
    if (ts.n > 0) {
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0 && !marked) {
            int b = stack[--sp];
            if (ts.counts_cumul[b] == 0 || !box_may_hit(ts, p, b)) continue;
            bool leaf = true;
            for (int c = 0; c < nchild; ++c) {
                int ch = ts.child[(size_t)c * ts.nboxes + b];
                if (ch >= 0) {
                    leaf = false;
                    if (sp >= ASTACK) { atomicOr(err, 1); return; }
                    stack[sp++] = ch;
                }
            }
            if (!leaf) continue;
            for (int i = ts.starts[b]; i < ts.starts[b] + ts.counts_cumul[b]; ++i) {
                double r = ts.rad[i];
                if (pdist2(dim, p, ts.pts, ts.n, i) <= r * r) { marked = true; break; }
            }
        }
    }
    ttc[it] = -1;
    if (!marked) { status[it] = 0; return; }
    // ---- find centers (QBX_CENTER_FINDER)
    int fl = flags[it];
    int best[2] = {-1, -1};
    double brel[2] = {INFINITY, INFINITY}, bd2[2] = {INFINITY, INFINITY};
    if (tc.n > 0) {
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            int b = stack[--sp];
            if (tc.counts_cumul[b] == 0 || !box_may_hit(tc, p, b)) continue;
            bool leaf = true;
            for (int c = 0; c < nchild; ++c) {
                int ch = tc.child[(size_t)c * tc.nboxes + b];
                if (ch >= 0) {
                    leaf = false;
                    if (sp >= ASTACK) { atomicOr(err, 1); return; }
                    stack[sp++] = ch;
                }
            }
            if (!leaf) continue;
            for (int i = tc.starts[b]; i < tc.starts[b] + tc.counts_cumul[b]; ++i) {
                int cidx = tc.user_ids[i];
                int side = 2 * (cidx % 2) - 1;
                if (side * fl < 0) continue;
                double d2 = pdist2(dim, p, tc.pts, tc.n, i);
                double Rc = tc.rad[i];
                double R2 = Rc * Rc;
                if (!(d2 <= R2)) continue;
                double rel = d2 / R2;
                int k = side > 0 ? 1 : 0;
                if (rel < brel[k] || (rel == brel[k] && cidx < best[k])) {
                    brel[k] = rel;
                    bd2[k] = d2;
                    best[k] = cidx;
                }
            }
        }
    }
    if (best[0] < 0 && best[1] < 0) {
        status[it] = 1;
        atomicAdd(nfailed, 1);
        return;
    }
    status[it] = 2;
    ttc[it] = (best[1] >= 0 && (best[0] < 0 || bd2[1] < bd2[0])) ? best[1] : best[0];
}

__global__ void k_refine(RTree ts, int ntargets, const double *targets, const int *status,
                         int nel, const int *el_starts, int *refine_flags, int *err)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntargets || status[it] != 1) return;
    int dim = ts.dim, nchild = 1 << dim;
    double p[3];
    for (int d = 0; d < dim; ++d) p[d] = targets[(size_t)d * ntargets + it];
    int stack[ASTACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int b = stack[--sp];
        if (ts.counts_cumul[b] == 0 || !box_may_hit(ts, p, b)) continue;
        bool leaf = true;
        for (int c = 0; c < nchild; ++c) {
            int ch = ts.child[(size_t)c * ts.nboxes + b];
            if (ch >= 0) {
                leaf = false;
                if (sp >= ASTACK) { atomicOr(err, 1); return; }
                stack[sp++] = ch;
            }
        }
        if (!leaf) continue;
        for (int i = ts.starts[b]; i < ts.starts[b] + ts.counts_cumul[b]; ++i) {
            double r = ts.rad[i];
            if (pdist2(dim, p, ts.pts, ts.n, i) <= r * r) {
                int s = ts.user_ids[i];
                int lo = 0, hi = nel;
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (el_starts[mid] <= s) lo = mid; else hi = mid;
                }
                refine_flags[lo] = 1;
            }
        }
    }
}

struct RTreeOwned {
    qbx_tree *tree = nullptr;
    dbuf<double> rad;
    dbuf<unsigned long long> boxmax;
    ~RTreeOwned() { delete tree; }
};

static int make_rtree(qbx_ctx *ctx, hipStream_t st, int dim, int n, const double *pts,
                      const double *radii, RTreeOwned &o, RTree &v)
{
    QBX_TRY(qbx_tree_build(ctx, st, dim, n, pts, 0, nullptr, nullptr, nullptr, nullptr, 32, 0.0,
                           0, &o.tree));
    qbx_tree *T = o.tree;
    QBX_TRY(o.rad.alloc(std::max(n, 1)));
    QBX_TRY(o.boxmax.alloc(T->nboxes));
    if (n > 0)
        hipLaunchKernelGGL(k_gather_radius, dim3(cdiv(n, 256)), dim3(256), 0, st, n,
                           T->user_source_ids.p, radii, o.rad.p);
    hipLaunchKernelGGL(k_leaf_max, dim3(cdiv(T->nboxes, 256)), dim3(256), 0, st, T->nboxes,
                       T->box_source_starts.p, T->box_source_counts_nonchild.p, o.rad.p,
                       o.boxmax.p);
    for (int l = T->nlevels - 1; l >= 1; --l) {
        int lo = T->level_start_box_nrs[l], hi = T->level_start_box_nrs[l + 1];
        hipLaunchKernelGGL(k_max_up, dim3(cdiv(hi - lo, 256)), dim3(256), 0, st, lo, hi,
                           T->box_parent_ids.p, o.boxmax.p);
    }
    QBX_HIP(hipGetLastError());
    v.dim = dim;
    v.nboxes = T->nboxes;
    v.n = n;
    for (int d = 0; d < 3; ++d) v.mn[d] = T->root_min[d];
    v.R = T->root_extent;
    v.level = T->box_levels.p;
    v.child = T->box_child_ids.p;
    v.coords = T->box_coords.p;
    v.starts = T->box_source_starts.p;
    v.counts_cumul = T->box_source_counts_cumul.p;
    v.pts = T->sources.p;
    v.rad = o.rad.p;
    v.user_ids = T->user_source_ids.p;
    v.boxmax = reinterpret_cast<const double *>(o.boxmax.p);
    return QBX_OK;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_target_assoc(qbx_ctx *ctx, void *stream, int dim, int nsrc1, const double *src1,
                     const double *tunnel_radii, int ncenters, const double *centers,
                     const double *radii_with_tol, int ntargets, const double *targets,
                     const int *target_flags, int *target_to_center, int *target_status,
                     int *nfailed)
{
    QBX_CHECK(ctx, "qbx_target_assoc: null ctx");
    QBX_CHECK(dim == 2 || dim == 3, "qbx_target_assoc: bad dim");
    QBX_CHECK(nsrc1 >= 0 && ncenters >= 0 && ntargets >= 0, "qbx_target_assoc: bad sizes");
    hipStream_t st = S(stream);
    if (nfailed) *nfailed = 0;
    if (ntargets == 0) return QBX_OK;
    RTreeOwned os, oc;
    RTree vs{}, vc{};
    QBX_TRY(make_rtree(ctx, st, dim, nsrc1, src1, tunnel_radii, os, vs));
    QBX_TRY(make_rtree(ctx, st, dim, ncenters, centers, radii_with_tol, oc, vc));
    dbuf<int> cnt;
    QBX_TRY(cnt.alloc(2));
    QBX_HIP(hipMemsetAsync(cnt.p, 0, 2 * sizeof(int), st));
    hipLaunchKernelGGL(k_assoc, dim3(cdiv(ntargets, 64)), dim3(64), 0, st, vs, vc, ntargets,
                       targets, target_flags, target_to_center, target_status, cnt.p,
                       cnt.p + 1);
    QBX_HIP(hipGetLastError());
    int h[2];
    QBX_HIP(hipMemcpyAsync(h, cnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    QBX_CHECK(h[1] == 0, "qbx_target_assoc: query stack overflow");
    if (nfailed) *nfailed = h[0];
    if (h[0] > 0) {
        set_error("qbx_target_assoc: %d targets were not assigned a QBX center", h[0]);
        return QBX_ERR_TARGET_ASSOC_FAILED;
    }
    return QBX_OK;
}

int qbx_refine_flags_for_failed(qbx_ctx *ctx, void *stream, int dim, int nsrc1,
                                const double *src1, const double *tunnel_radii, int nelements,
                                const int *el_starts, int ntargets, const double *targets,
                                const int *target_status, int *refine_flags)
{
    QBX_CHECK(ctx && refine_flags, "qbx_refine_flags_for_failed: null argument");
    hipStream_t st = S(stream);
    QBX_HIP(hipMemsetAsync(refine_flags, 0, sizeof(int) * std::max(nelements, 0), st));
    if (ntargets == 0 || nsrc1 == 0) return QBX_OK;
    RTreeOwned os;
    RTree vs{};
    QBX_TRY(make_rtree(ctx, st, dim, nsrc1, src1, tunnel_radii, os, vs));
    dbuf<int> err;
    QBX_TRY(err.alloc(1));
    QBX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_refine, dim3(cdiv(ntargets, 64)), dim3(64), 0, st, vs, ntargets, targets,
                       target_status, nelements, el_starts, refine_flags, err.p);
    QBX_HIP(hipGetLastError());
    int h = 0;
    QBX_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    QBX_CHECK(h == 0, "qbx_refine_flags_for_failed: query stack overflow");
    return QBX_OK;
}

}  // extern "C"

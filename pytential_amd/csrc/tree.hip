// Adaptive 2^d-tree build on the GPU (boxtree TreeBuilder semantics as used by
// pytential/qbx/geometry.py:562-618 and pytential/qbx/utils.py:278-448).
//
// Exact spec (the CPU oracle implements the same one, operation by operation;
// this file is compiled with -ffp-contract=off so no FMA changes a rounding):
//  * root box: bounding box of all points, extent R = max side * (1 + 2^-30)
//  * integer coordinates q = floor((x - min) / R * 2^L), L = 28 (2D) / 19 (3D)
//  * a target with radius r stops at the deepest level l at which it fits every
//    box on its path: l2 norm: A = ((1+s) sqrt(d)/2) h, A - r >= 0 and
//    |x - c|^2 <= (A - r)^2; linf: |x_d - c_d| + r <= (1/2 + s) h per axis
//  * level-synchronous refinement: a box splits iff the summed refine weight of
//    the particles in its subtree exceeds max_leaf_refine_weight (and l < L)
//  * boxes numbered breadth first, children in Morton child order; particles
//    in pre-order rank of their owning box, ties by user index
#include "qbx_internal.h"

namespace qbx {

__constant__ double c_sqrt_dim_half[4] = {0.0, 0.5, 0.70710678118654757, 0.86602540378443865};

__global__ void k_bbox(int dim, int ns, const double *src, int nt, const double *tgt,
                       double *partial)
{
    __shared__ double smin[3][256], smax[3][256];
    int tid = threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    long np = (long)ns + nt;
    for (long p = blockIdx.x * (long)blockDim.x + tid; p < np; p += (long)gridDim.x * blockDim.x) {
        for (int d = 0; d < dim; ++d) {
            double v = p < ns ? src[(size_t)d * ns + p] : tgt[(size_t)d * nt + (p - ns)];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) { smin[d][tid] = mn[d]; smax[d][tid] = mx[d]; }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int d = 0; d < 3; ++d) {
                smin[d][tid] = fmin(smin[d][tid], smin[d][tid + s]);
                smax[d][tid] = fmax(smax[d][tid], smax[d][tid + s]);
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int d = 0; d < 3; ++d) {
            partial[blockIdx.x * 6 + d] = smin[d][0];
            partial[blockIdx.x * 6 + 3 + d] = smax[d][0];
        }
}

struct TreeSpec {
    int dim, L, linf;
    double mn[3];
    double R, scale, K, half_plus_s;
};

__device__ bool particle_fits(const TreeSpec &ts, const double *x, double r, const int *q,
                              int level)
{
    double h = ldexp(ts.R, -level);
    double d2 = 0.0;
    bool ok = true;
    for (int d = 0; d < ts.dim; ++d) {
        int I = q[d] >> (ts.L - level);
        double c = ts.mn[d] + ((double)I + 0.5) * h;
        double dd = x[d] - c;
        if (ts.linf) {
            if (fabs(dd) + r > ts.half_plus_s * h) ok = false;
        } else {
            d2 = d2 + dd * dd;
        }
    }
    if (ts.linf) return ok;
    double thr = ts.K * h - r;
    if (thr < 0.0) return false;
    return d2 <= thr * thr;
}

__global__ void k_particle_init(TreeSpec ts, int ns, const double *src, int nt,
                                const double *tgt, const double *tgt_radii, const int *sw,
                                const int *tw, int *q, int *stop, int *w, int *cur, int *owner)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int np = ns + nt;
    if (p >= np) return;
    double x[3];
    int qq[3];
    int qmax = (1 << ts.L) - 1;
    for (int d = 0; d < ts.dim; ++d) {
        x[d] = p < ns ? src[(size_t)d * ns + p] : tgt[(size_t)d * nt + (p - ns)];
        double u = (x[d] - ts.mn[d]) / ts.R * ts.scale;
        double f = floor(u);
        int qi = f < 0 ? 0 : (f > qmax ? qmax : (int)f);
        qq[d] = qi;
        q[(size_t)d * np + p] = qi;
    }
    int st = ts.L;
    if (p < ns) {
        w[p] = sw ? sw[p] : 1;
    } else {
        int it = p - ns;
        w[p] = tw ? tw[it] : 1;
        double r = tgt_radii ? tgt_radii[it] : 0.0;
        if (r > 0.0) {
            int l = 0;
            while (l < ts.L && particle_fits(ts, x, r, qq, l + 1)) ++l;
            st = l;
        }
    }
    stop[p] = st;
    cur[p] = 0;
    owner[p] = -1;
}

__global__ void k_weight(int np, const int *owner, const int *cur, const int *w, int lo,
                         unsigned long long *wt)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || owner[p] >= 0) return;
    atomicAdd(&wt[cur[p] - lo], (unsigned long long)w[p]);
}

__global__ void k_mark_children(int np, int dim, int L, int l, int lo, long long W,
                                const unsigned long long *wt, int *owner, const int *cur,
                                const int *stop, const int *q, int *cflag)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || owner[p] >= 0) return;
    int b = cur[p] - lo;
    bool split = (wt[b] > (unsigned long long)W) && (l < L);
    if (!split || stop[p] == l) {
        owner[p] = cur[p];
        return;
    }
    int c = 0;
    for (int d = 0; d < dim; ++d) c |= ((q[(size_t)d * np + p] >> (L - l - 1)) & 1) << d;
    cflag[(size_t)b * (1 << dim) + c] = 1;
}

__global__ void k_make_children(int nslots, int dim, int lo, int base, const int *cflag,
                                const int *pos, int *level, int *parent, int *coords_bm,
                                int *child_bm, int *childmap, int l)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nslots) return;
    int nchild = 1 << dim;
    if (!cflag[idx]) {
        childmap[idx] = -1;
        return;
    }
    int id = base + pos[idx];
    int b = lo + idx / nchild, c = idx % nchild;
    level[id] = l + 1;
    parent[id] = b;
    for (int d = 0; d < dim; ++d)
        coords_bm[(size_t)id * dim + d] = 2 * coords_bm[(size_t)b * dim + d] + ((c >> d) & 1);
    for (int cc = 0; cc < nchild; ++cc) child_bm[(size_t)id * nchild + cc] = -1;
    child_bm[(size_t)b * nchild + c] = id;
    childmap[idx] = id;
}

__global__ void k_move(int np, int dim, int L, int l, int lo, const int *owner, int *cur,
                       const int *q, const int *childmap)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || owner[p] >= 0) return;
    int c = 0;
    for (int d = 0; d < dim; ++d) c |= ((q[(size_t)d * np + p] >> (L - l - 1)) & 1) << d;
    cur[p] = childmap[(size_t)(cur[p] - lo) * (1 << dim) + c];
}

__device__ unsigned long long morton_of(int dim, const int *I, int level)
{
    unsigned long long k = 0;
    for (int i = 0; i < level; ++i)
        for (int d = 0; d < dim; ++d)
            k |= (unsigned long long)((I[d] >> i) & 1) << (i * dim + d);
    return k;
}

__global__ void k_box_finalize(int nb, int dim, int L, const int *level, const int *coords_bm,
                               const int *child_bm, double mnx, double mny, double mnz,
                               double R, int *coords, int *child_ids, double *centers,
                               unsigned long long *morton, unsigned long long *prekey, int *iota)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int nchild = 1 << dim;
    int I[3] = {0, 0, 0};
    double mn[3] = {mnx, mny, mnz};
    int lv = level[b];
    double h = ldexp(R, -lv);
    for (int d = 0; d < dim; ++d) {
        I[d] = coords_bm[(size_t)b * dim + d];
        coords[(size_t)d * nb + b] = I[d];
        centers[(size_t)d * nb + b] = mn[d] + ((double)I[d] + 0.5) * h;
    }
    for (int c = 0; c < nchild; ++c) child_ids[(size_t)c * nb + b] = child_bm[(size_t)b * nchild + c];
    unsigned long long m = morton_of(dim, I, lv);
    morton[b] = m;
    prekey[b] = ((m << (dim * (L - lv))) << 5) | (unsigned long long)lv;
    iota[b] = b;
}

__global__ void k_scatter_rank(int nb, const int *byrank, int *rank)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nb) rank[byrank[r]] = r;
}

__global__ void k_particle_keys(int n, int off, const int *owner, const int *rank, int *key,
                                int *val, int *nonchild)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = owner[off + i];
    key[i] = rank[b];
    val[i] = i;
    atomicAdd(&nonchild[b], 1);
}

__global__ void k_perm_by_rank(int nb, const int *rank, const int *cnt, int *out)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) out[rank[b]] = cnt[b];
}

__global__ void k_starts_from_rank(int nb, const int *rank, const int *scan, int *starts)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) starts[b] = scan[rank[b]];
}

__global__ void k_cumul_level(int lo, int hi, const int *parent, int *cumul)
{
    int b = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (b < hi) atomicAdd(&cumul[parent[b]], cumul[b]);
}

__global__ void k_gather_points(int n, int dim, const int *perm, const double *in, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u = perm[i];
    for (int d = 0; d < dim; ++d) out[(size_t)d * n + i] = in[(size_t)d * n + u];
}

static int grow(dbuf<int> &b, long need, long keep)
{
    if (need <= b.n) return QBX_OK;
    long nc = need * 2;
    int *np_;
    QBX_HIP(hipMalloc(&np_, sizeof(int) * nc));
    if (b.p && keep > 0) QBX_HIP(hipMemcpy(np_, b.p, sizeof(int) * keep, hipMemcpyDeviceToDevice));
    if (b.p) QBX_HIP(hipFree(b.p));
    b.p = np_;
    b.n = nc;
    return QBX_OK;
}

static int build(qbx_ctx *ctx, hipStream_t st, int dim, int ns, const double *src, int nt,
                 const double *tgt, const double *tgt_radii, const int *sw, const int *tw,
                 int W, double s, int linf, qbx_tree *T)
{
    const int nchild = 1 << dim;
    const int np = ns + nt;
    const int L = dim == 2 ? 28 : 19;
    T->dim = dim;
    T->nsources = ns;
    T->ntargets = nt;
    T->maxl = L;

    // bounding box
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    if (np > 0) {
        int nblk = std::min(cdiv(np, 256), 512);
        dbuf<double> part;
        QBX_TRY(part.alloc((long)nblk * 6));
        hipLaunchKernelGGL(k_bbox, dim3(nblk), dim3(256), 0, st, dim, ns, src, nt, tgt, part.p);
        QBX_HIP(hipGetLastError());
        std::vector<double> h(nblk * 6);
        QBX_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * nblk * 6,
                               hipMemcpyDeviceToHost, st));
        QBX_HIP(hipStreamSynchronize(st));
        for (int d = 0; d < dim; ++d) { mn[d] = INFINITY; mx[d] = -INFINITY; }
        for (int b = 0; b < nblk; ++b)
            for (int d = 0; d < dim; ++d) {
                mn[d] = std::min(mn[d], h[b * 6 + d]);
                mx[d] = std::max(mx[d], h[b * 6 + 3 + d]);
            }
    }
    double ext = 0.0;
    for (int d = 0; d < dim; ++d)
        if (mx[d] - mn[d] > ext) ext = mx[d] - mn[d];
    double R = ext > 0.0 ? ext * (1.0 + 0x1p-30) : 1.0;
    for (int d = 0; d < 3; ++d) T->root_min[d] = d < dim ? mn[d] : 0.0;
    T->root_extent = R;

    TreeSpec ts;
    ts.dim = dim;
    ts.L = L;
    ts.linf = linf;
    for (int d = 0; d < 3; ++d) ts.mn[d] = T->root_min[d];
    ts.R = R;
    ts.scale = ldexp(1.0, L);
    const double sdh[4] = {0.0, 0.5, 0.70710678118654757, 0.86602540378443865};
    ts.K = (1.0 + s) * sdh[dim];
    ts.half_plus_s = 0.5 + s;

    dbuf<int> q, stop, w, cur, owner;
    QBX_TRY(q.alloc((long)dim * std::max(np, 1)));
    QBX_TRY(stop.alloc(std::max(np, 1)));
    QBX_TRY(w.alloc(std::max(np, 1)));
    QBX_TRY(cur.alloc(std::max(np, 1)));
    QBX_TRY(owner.alloc(std::max(np, 1)));
    if (np > 0) {
        hipLaunchKernelGGL(k_particle_init, dim3(cdiv(np, 256)), dim3(256), 0, st, ts, ns, src,
                           nt, tgt, tgt_radii, sw, tw, q.p, stop.p, w.p, cur.p, owner.p);
        QBX_HIP(hipGetLastError());
    }

    // box arrays (box-major during the build)
    dbuf<int> level, parent, coords_bm, child_bm;
    long cap = std::max(1024, np / 4 + 16);
    QBX_TRY(level.alloc(cap));
    QBX_TRY(parent.alloc(cap));
    QBX_TRY(coords_bm.alloc(cap * dim));
    QBX_TRY(child_bm.alloc(cap * nchild));
    {
        int zero = 0, m1 = -1;
        QBX_HIP(hipMemcpyAsync(level.p, &zero, sizeof(int), hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemcpyAsync(parent.p, &m1, sizeof(int), hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemsetAsync(coords_bm.p, 0, sizeof(int) * dim, st));
        QBX_HIP(hipMemsetAsync(child_bm.p, 0xff, sizeof(int) * nchild, st));
        QBX_HIP(hipStreamSynchronize(st));
    }
    std::vector<int> lstart = {0, 1};
    int nboxes = 1;
    dbuf<unsigned long long> wt;
    dbuf<int> cflag, pos, childmap;
    for (int l = 0;; ++l) {
        int lo = lstart[l], hi = lstart[l + 1];
        if (lo == hi) break;
        QBX_CHECK(l < MAX_LEVELS - 1, "qbx_tree_build: too many levels");
        int nb = hi - lo;
        long nslots = (long)nb * nchild;
        if (wt.n < nb) QBX_TRY(wt.alloc(nb));
        if (cflag.n < nslots) {
            QBX_TRY(cflag.alloc(nslots));
            QBX_TRY(pos.alloc(nslots + 1));
            QBX_TRY(childmap.alloc(nslots));
        }
        QBX_HIP(hipMemsetAsync(wt.p, 0, sizeof(unsigned long long) * nb, st));
        QBX_HIP(hipMemsetAsync(cflag.p, 0, sizeof(int) * nslots, st));
        if (np > 0) {
            hipLaunchKernelGGL(k_weight, dim3(cdiv(np, 256)), dim3(256), 0, st, np, owner.p,
                               cur.p, w.p, lo, wt.p);
            hipLaunchKernelGGL(k_mark_children, dim3(cdiv(np, 256)), dim3(256), 0, st, np, dim,
                               L, l, lo, (long long)W, wt.p, owner.p, cur.p, stop.p, q.p,
                               cflag.p);
            QBX_HIP(hipGetLastError());
        }
        int nnew = 0;
        QBX_TRY(exclusive_scan(ctx, st, cflag.p, pos.p, (int)nslots, &nnew));
        if (nboxes + nnew > cap) {
            long ncap = std::max(cap * 2, (long)(nboxes + nnew) + 1024);
            QBX_TRY(grow(level, ncap, nboxes));
            QBX_TRY(grow(parent, ncap, nboxes));
            QBX_TRY(grow(coords_bm, ncap * dim, (long)nboxes * dim));
            QBX_TRY(grow(child_bm, ncap * nchild, (long)nboxes * nchild));
            cap = ncap;
        }
        hipLaunchKernelGGL(k_make_children, dim3(cdiv(nslots, 256)), dim3(256), 0, st,
                           (int)nslots, dim, lo, nboxes, cflag.p, pos.p, level.p, parent.p,
                           coords_bm.p, child_bm.p, childmap.p, l);
        QBX_HIP(hipGetLastError());
        if (np > 0 && nnew > 0) {
            hipLaunchKernelGGL(k_move, dim3(cdiv(np, 256)), dim3(256), 0, st, np, dim, L, l, lo,
                               owner.p, cur.p, q.p, childmap.p);
            QBX_HIP(hipGetLastError());
        }
        nboxes += nnew;
        lstart.push_back(nboxes);
    }
    int nlevels = (int)lstart.size() - 2;
    T->nboxes = nboxes;
    T->nlevels = nlevels;
    T->level_start_box_nrs.assign(lstart.begin(), lstart.begin() + nlevels + 1);

    // finalize box arrays
    QBX_TRY(T->box_levels.alloc(nboxes));
    QBX_HIP(hipMemcpyAsync(T->box_levels.p, level.p, sizeof(int) * nboxes,
                           hipMemcpyDeviceToDevice, st));
    QBX_TRY(T->box_parent_ids.alloc(nboxes));
    QBX_HIP(hipMemcpyAsync(T->box_parent_ids.p, parent.p, sizeof(int) * nboxes,
                           hipMemcpyDeviceToDevice, st));
    QBX_TRY(T->box_coords.alloc((long)dim * nboxes));
    QBX_TRY(T->box_child_ids.alloc((long)nchild * nboxes));
    QBX_TRY(T->box_centers.alloc((long)dim * nboxes));
    QBX_TRY(T->box_morton.alloc(nboxes));
    dbuf<unsigned long long> prekey, prekey_sorted;
    dbuf<int> iota, byrank, rank;
    QBX_TRY(prekey.alloc(nboxes));
    QBX_TRY(prekey_sorted.alloc(nboxes));
    QBX_TRY(iota.alloc(nboxes));
    QBX_TRY(byrank.alloc(nboxes));
    QBX_TRY(rank.alloc(nboxes));
    hipLaunchKernelGGL(k_box_finalize, dim3(cdiv(nboxes, 256)), dim3(256), 0, st, nboxes, dim, L,
                       level.p, coords_bm.p, child_bm.p, T->root_min[0], T->root_min[1],
                       T->root_min[2], R, T->box_coords.p, T->box_child_ids.p,
                       T->box_centers.p, T->box_morton.p, prekey.p, iota.p);
    QBX_HIP(hipGetLastError());
    QBX_TRY(sort_pairs_u64(ctx, st, prekey.p, prekey_sorted.p, iota.p, byrank.p, nboxes, 64));
    hipLaunchKernelGGL(k_scatter_rank, dim3(cdiv(nboxes, 256)), dim3(256), 0, st, nboxes,
                       byrank.p, rank.p);
    QBX_HIP(hipGetLastError());
    QBX_TRY(T->d_level_start_box_nrs.alloc(nlevels + 1));
    QBX_HIP(hipMemcpyAsync(T->d_level_start_box_nrs.p, T->level_start_box_nrs.data(),
                           sizeof(int) * (nlevels + 1), hipMemcpyHostToDevice, st));
    QBX_HIP(hipStreamSynchronize(st));

    int rank_bits = 1;
    while ((1L << rank_bits) < (long)nboxes + 1) ++rank_bits;

    for (int kind = 0; kind < 2; ++kind) {
        int n = kind == 0 ? ns : nt;
        int off = kind == 0 ? 0 : ns;
        dbuf<int> &perm = kind == 0 ? T->user_source_ids : T->sorted_target_ids;
        dbuf<int> &starts = kind == 0 ? T->box_source_starts : T->box_target_starts;
        dbuf<int> &nonchild = kind == 0 ? T->box_source_counts_nonchild : T->box_target_counts_nonchild;
        dbuf<int> &cumul = kind == 0 ? T->box_source_counts_cumul : T->box_target_counts_cumul;
        QBX_TRY(perm.alloc(std::max(n, 1)));
        perm.n = n;
        QBX_TRY(starts.alloc(nboxes));
        QBX_TRY(nonchild.alloc(nboxes));
        QBX_TRY(cumul.alloc(nboxes));
        QBX_HIP(hipMemsetAsync(nonchild.p, 0, sizeof(int) * nboxes, st));
        if (n > 0) {
            dbuf<int> key, val, key2;
            QBX_TRY(key.alloc(n));
            QBX_TRY(val.alloc(n));
            QBX_TRY(key2.alloc(n));
            hipLaunchKernelGGL(k_particle_keys, dim3(cdiv(n, 256)), dim3(256), 0, st, n, off,
                               owner.p, rank.p, key.p, val.p, nonchild.p);
            QBX_HIP(hipGetLastError());
            QBX_TRY(sort_pairs_i32(ctx, st, key.p, key2.p, val.p, perm.p, n, rank_bits));
        }
        dbuf<int> byr, scan;
        QBX_TRY(byr.alloc(nboxes));
        QBX_TRY(scan.alloc(nboxes + 1));
        hipLaunchKernelGGL(k_perm_by_rank, dim3(cdiv(nboxes, 256)), dim3(256), 0, st, nboxes,
                           rank.p, nonchild.p, byr.p);
        QBX_HIP(hipGetLastError());
        QBX_TRY(exclusive_scan(ctx, st, byr.p, scan.p, nboxes, nullptr));
        hipLaunchKernelGGL(k_starts_from_rank, dim3(cdiv(nboxes, 256)), dim3(256), 0, st,
                           nboxes, rank.p, scan.p, starts.p);
        QBX_HIP(hipMemcpyAsync(cumul.p, nonchild.p, sizeof(int) * nboxes,
                               hipMemcpyDeviceToDevice, st));
        for (int l = nlevels - 1; l >= 1; --l) {
            int lo = lstart[l], hi = lstart[l + 1];
            hipLaunchKernelGGL(k_cumul_level, dim3(cdiv(hi - lo, 256)), dim3(256), 0, st, lo, hi,
                               T->box_parent_ids.p, cumul.p);
        }
        QBX_HIP(hipGetLastError());
    }
    QBX_TRY(T->sources.alloc((long)dim * std::max(ns, 1)));
    QBX_TRY(T->targets.alloc((long)dim * std::max(nt, 1)));
    if (ns > 0)
        hipLaunchKernelGGL(k_gather_points, dim3(cdiv(ns, 256)), dim3(256), 0, st, ns, dim,
                           T->user_source_ids.p, src, T->sources.p);
    if (nt > 0)
        hipLaunchKernelGGL(k_gather_points, dim3(cdiv(nt, 256)), dim3(256), 0, st, nt, dim,
                           T->sorted_target_ids.p, tgt, T->targets.p);
    QBX_HIP(hipGetLastError());
    QBX_HIP(hipStreamSynchronize(st));
    return QBX_OK;
}

struct FieldRef {
    const void *p;
    long bytes;
};

static FieldRef tree_field(const qbx_tree *t, const char *f)
{
    std::string s(f);
    int nb = t->nboxes;
    auto I = [&](const dbuf<int> &b, long n) { return FieldRef{b.p, (long)sizeof(int) * n}; };
    if (s == "box_levels") return I(t->box_levels, nb);
    if (s == "box_parent_ids") return I(t->box_parent_ids, nb);
    if (s == "box_child_ids") return I(t->box_child_ids, (long)nb << t->dim);
    if (s == "box_coords") return I(t->box_coords, (long)nb * t->dim);
    if (s == "box_centers") return FieldRef{t->box_centers.p, (long)sizeof(double) * nb * t->dim};
    if (s == "box_source_starts") return I(t->box_source_starts, nb);
    if (s == "box_source_counts_nonchild") return I(t->box_source_counts_nonchild, nb);
    if (s == "box_source_counts_cumul") return I(t->box_source_counts_cumul, nb);
    if (s == "box_target_starts") return I(t->box_target_starts, nb);
    if (s == "box_target_counts_nonchild") return I(t->box_target_counts_nonchild, nb);
    if (s == "box_target_counts_cumul") return I(t->box_target_counts_cumul, nb);
    if (s == "level_start_box_nrs") return I(t->d_level_start_box_nrs, t->nlevels + 1);
    if (s == "user_source_ids") return I(t->user_source_ids, t->nsources);
    if (s == "sorted_target_ids") return I(t->sorted_target_ids, t->ntargets);
    if (s == "sources") return FieldRef{t->sources.p, (long)sizeof(double) * t->dim * t->nsources};
    if (s == "targets") return FieldRef{t->targets.p, (long)sizeof(double) * t->dim * t->ntargets};
    return FieldRef{nullptr, -1};
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_tree_build(qbx_ctx *ctx, void *stream, int dim, int nsources, const double *sources,
                   int ntargets, const double *targets, const double *target_radii,
                   const int *source_weights, const int *target_weights,
                   int max_leaf_refine_weight, double stick_out_factor, int extent_norm_linf,
                   qbx_tree **out)
{
    QBX_CHECK(ctx && out, "qbx_tree_build: null argument");
    QBX_CHECK(dim == 2 || dim == 3, "qbx_tree_build: dim must be 2 or 3 (got %d)", dim);
    QBX_CHECK(nsources >= 0 && ntargets >= 0, "qbx_tree_build: negative sizes");
    QBX_CHECK((nsources == 0 || sources) && (ntargets == 0 || targets),
              "qbx_tree_build: null coordinates");
    QBX_CHECK(max_leaf_refine_weight > 0, "qbx_tree_build: max_leaf_refine_weight must be > 0");
    QBX_CHECK(stick_out_factor >= 0, "qbx_tree_build: negative stick_out_factor");
    qbx_tree *T = new qbx_tree();
    int rc = build(ctx, S(stream), dim, nsources, sources, ntargets, targets, target_radii,
                   source_weights, target_weights, max_leaf_refine_weight, stick_out_factor,
                   extent_norm_linf, T);
    if (rc != QBX_OK) {
        delete T;
        return rc;
    }
    *out = T;
    return QBX_OK;
}

int qbx_tree_destroy(qbx_tree *tree)
{
    delete tree;
    return QBX_OK;
}

int qbx_tree_sizes(const qbx_tree *t, int *sizes5, double *root4)
{
    QBX_CHECK(t, "qbx_tree_sizes: null tree");
    if (sizes5) {
        sizes5[0] = t->dim;
        sizes5[1] = t->nsources;
        sizes5[2] = t->ntargets;
        sizes5[3] = t->nboxes;
        sizes5[4] = t->nlevels;
    }
    if (root4) {
        for (int d = 0; d < 3; ++d) root4[d] = t->root_min[d];
        root4[3] = t->root_extent;
    }
    return QBX_OK;
}

long qbx_tree_field_bytes(const qbx_tree *t, const char *field)
{
    if (!t || !field) return -1;
    return tree_field(t, field).bytes;
}

int qbx_tree_export(const qbx_tree *t, const char *field, void *dst, long nbytes, void *stream)
{
    QBX_CHECK(t && field, "qbx_tree_export: null argument");
    FieldRef r = tree_field(t, field);
    QBX_CHECK(r.bytes >= 0, "qbx_tree_export: unknown field '%s'", field);
    QBX_CHECK(nbytes >= r.bytes, "qbx_tree_export: '%s' needs %ld bytes, got %ld", field,
              r.bytes, nbytes);
    if (r.bytes > 0)
        QBX_HIP(hipMemcpyAsync(dst, r.p, r.bytes, hipMemcpyDeviceToDevice, S(stream)));
    return QBX_OK;
}

}  // extern "C"

// FMM interaction lists on the GPU (boxtree FMMTraversalBuilder +
// merge_close_lists semantics as used by pytential/qbx/geometry.py:622-641).
//
// Integer spec on box index coordinates (n = well_sep_is_n_away, e = 1 when
// targets carry extent):
//   gap(A,B)   l-inf gap between boxes, in units of the smaller box
//   sep(S,X)   S not smaller than X: gap(S,X) >= n
//   sep3(D,B)  D smaller than B: gap(D, B grown by e of its own boxes) >= n
//   list1(B)   source leaves (a) not smaller than B with !sep, (b) below a
//              same-level colleague of B and under no sep3 box, (c) leaves of
//              sep3 boxes holding fewer than min_nsources_cumul sources
//   list2(P)   children of colleagues of parent(P) that are not colleagues of P
//   list3(B)   first sep3 boxes below colleagues of B
//   list4(P)   source leaves S bigger than P: sep(S,P) and !sep(S,parent(P))
// Lists are sorted by box number.  Each list is built by one thread: a count
// pass, an exclusive scan, a fill pass and an in-place insertion sort.
#include "qbx_internal.h"

namespace qbx {

struct TV {
    int dim, nboxes, nlevels;
    const int *level, *parent, *child, *coords, *lstart;
    const unsigned long long *morton;
    const int *src_nonchild, *src_cumul;
};

__device__ unsigned long long tv_morton(int dim, const int *I, int level)
{
    unsigned long long k = 0;
    for (int i = 0; i < level; ++i)
        for (int d = 0; d < dim; ++d)
            k |= (unsigned long long)((I[d] >> i) & 1) << (i * dim + d);
    return k;
}

__device__ int tv_find(const TV &t, int level, const int *I)
{
    if (level >= t.nlevels) return -1;
    for (int d = 0; d < t.dim; ++d)
        if (I[d] < 0 || I[d] >= (1 << level)) return -1;
    unsigned long long key = tv_morton(t.dim, I, level);
    int lo = t.lstart[level], hi = t.lstart[level + 1], end = hi;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (t.morton[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < end && t.morton[lo] == key) ? lo : -1;
}

__device__ bool tv_leaf(const TV &t, int b)
{
    for (int c = 0; c < (1 << t.dim); ++c)
        if (t.child[(size_t)c * t.nboxes + b] >= 0) return false;
    return true;
}

__device__ long long tv_gap(const TV &t, int a, int b, int grow)
{
    int la = t.level[a], lb = t.level[b];
    int m = la > lb ? la : lb;
    long long g = 0;
    for (int d = 0; d < t.dim; ++d) {
        long long sa = 1LL << (m - la), sb = 1LL << (m - lb);
        long long ca = t.coords[(size_t)d * t.nboxes + a], cb = t.coords[(size_t)d * t.nboxes + b];
        long long alo = ca * sa, ahi = (ca + 1) * sa;
        long long blo = (cb - grow) * sb, bhi = (cb + 1 + grow) * sb;
        long long gd = 0;
        if (alo - bhi > gd) gd = alo - bhi;
        if (blo - ahi > gd) gd = blo - ahi;
        if (gd > g) g = gd;
    }
    return g;
}

// colleague k (k in [0, (2n+1)^dim)) of box b, or -1
__device__ int tv_colleague(const TV &t, int b, int n, int k)
{
    int w = 2 * n + 1, I[3];
    for (int d = 0; d < t.dim; ++d) {
        I[d] = t.coords[(size_t)d * t.nboxes + b] + (k % w) - n;
        k /= w;
    }
    return tv_find(t, t.level[b], I);
}

struct ListOut {
    int *cnt;          // count mode: per-item count
    const int *starts; // fill mode
    int *lists;
    int fill;          // 0 count, 1 fill
};

__device__ inline void emit(ListOut &o, int item, int &k, int v)
{
    if (o.fill) o.lists[o.starts[item] + k] = v;
    ++k;
}

constexpr int STACK = 192;

__global__ void k_list13(TV t, int n, int e, int minsrc, int ntb, const int *target_boxes,
                         ListOut o1, ListOut o3, int *err)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntb) return;
    int B = target_boxes[it];
    int k1 = 0, k3 = 0;
    int ncol = 1;
    for (int d = 0; d < t.dim; ++d) ncol *= (2 * n + 1);
    // (a) leaves not smaller than B: colleagues of each ancestor-or-self
    int anc[MAX_LEVELS];
    int na = 0;
    for (int a = B; a >= 0 && na < MAX_LEVELS; a = t.parent[a]) anc[na++] = a;
    for (int j = na - 1; j >= 0; --j) {
        for (int kc = 0; kc < ncol; ++kc) {
            int C = tv_colleague(t, anc[j], n, kc);
            if (C < 0 || !tv_leaf(t, C) || t.src_nonchild[C] == 0) continue;
            if (tv_gap(t, C, B, 0) < n) emit(o1, it, k1, C);
        }
    }
    // (b) below same-level colleagues
    int stack[STACK];
    for (int kc = 0; kc < ncol; ++kc) {
        int C = tv_colleague(t, B, n, kc);
        if (C < 0 || tv_leaf(t, C)) continue;
        int sp = 0;
        for (int ch = (1 << t.dim) - 1; ch >= 0; --ch) {
            int x = t.child[(size_t)ch * t.nboxes + C];
            if (x >= 0) stack[sp++] = x;
        }
        while (sp > 0) {
            int D = stack[--sp];
            if (t.src_cumul[D] == 0) continue;
            bool leaf = tv_leaf(t, D);
            if (tv_gap(t, D, B, e) >= n) {
                if (t.src_cumul[D] < minsrc) {
                    // all source leaves below D go to list 1
                    int sub[STACK];
                    int sq = 0;
                    sub[sq++] = D;
                    while (sq > 0) {
                        int X = sub[--sq];
                        if (t.src_cumul[X] == 0) continue;
                        if (tv_leaf(t, X)) {
                            if (t.src_nonchild[X] > 0) emit(o1, it, k1, X);
                            continue;
                        }
                        for (int ch = (1 << t.dim) - 1; ch >= 0; --ch) {
                            int x = t.child[(size_t)ch * t.nboxes + X];
                            if (x < 0) continue;
                            if (sq >= STACK) { atomicOr(err, 1); return; }
                            sub[sq++] = x;
                        }
                    }
                } else {
                    emit(o3, it, k3, D);
                }
            } else if (leaf) {
                if (t.src_nonchild[D] > 0) emit(o1, it, k1, D);
            } else {
                for (int ch = (1 << t.dim) - 1; ch >= 0; --ch) {
                    int x = t.child[(size_t)ch * t.nboxes + D];
                    if (x < 0) continue;
                    if (sp >= STACK) { atomicOr(err, 1); return; }
                    stack[sp++] = x;
                }
            }
        }
    }
    if (!o1.fill) o1.cnt[it] = k1;
    if (!o3.fill) o3.cnt[it] = k3;
}

__global__ void k_list24(TV t, int n, int nttp, const int *ttp_boxes, ListOut o2, ListOut o4)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nttp) return;
    int P = ttp_boxes[it];
    int k2 = 0, k4 = 0;
    int lp = t.level[P];
    if (lp > 0) {
        int ncol = 1;
        for (int d = 0; d < t.dim; ++d) ncol *= (2 * n + 1);
        int Q = t.parent[P];
        for (int kc = 0; kc < ncol; ++kc) {
            int C = tv_colleague(t, Q, n, kc);
            if (C < 0) continue;
            for (int ch = 0; ch < (1 << t.dim); ++ch) {
                int A = t.child[(size_t)ch * t.nboxes + C];
                if (A < 0 || t.src_cumul[A] == 0) continue;
                long long g = 0;
                for (int d = 0; d < t.dim; ++d) {
                    long long dd = (long long)t.coords[(size_t)d * t.nboxes + A] -
                                   t.coords[(size_t)d * t.nboxes + P];
                    if (dd < 0) dd = -dd;
                    if (dd > g) g = dd;
                }
                if (g > n) emit(o2, it, k2, A);
            }
        }
        int anc[MAX_LEVELS];
        int na = 0;
        for (int a = Q; a >= 0 && na < MAX_LEVELS; a = t.parent[a]) anc[na++] = a;
        for (int j = na - 1; j >= 0; --j) {
            for (int kc = 0; kc < ncol; ++kc) {
                int Sb = tv_colleague(t, anc[j], n, kc);
                if (Sb < 0 || !tv_leaf(t, Sb) || t.src_nonchild[Sb] == 0) continue;
                if (tv_gap(t, Sb, P, 0) >= n && tv_gap(t, Sb, Q, 0) < n) emit(o4, it, k4, Sb);
            }
        }
    }
    if (!o2.fill) o2.cnt[it] = k2;
    if (!o4.fill) o4.cnt[it] = k4;
}

__global__ void k_sort_lists(int nitems, const int *starts, int *lists)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems) return;
    int a = starts[it], b = starts[it + 1];
    for (int i = a + 1; i < b; ++i) {
        int v = lists[i];
        int j = i - 1;
        while (j >= a && lists[j] > v) {
            lists[j + 1] = lists[j];
            --j;
        }
        lists[j + 1] = v;
    }
}

__global__ void k_box_flags(TV t, const int *tgt_nonchild, int *f_src, int *f_srcpar,
                            int *f_tgt, int *f_ttp)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= t.nboxes) return;
    bool leaf = tv_leaf(t, b);
    f_src[b] = (leaf && t.src_nonchild[b] > 0) ? 1 : 0;
    f_srcpar[b] = (!leaf && t.src_cumul[b] > 0) ? 1 : 0;
    int tg = tgt_nonchild[b] > 0 ? 1 : 0;
    f_tgt[b] = tg;
    if (tg)
        for (int a = b; a >= 0; a = t.parent[a]) f_ttp[a] = 1;  // idempotent stores
}

__global__ void k_level_starts(int nlevels, const int *boxes, int n, const int *box_level,
                               int *out)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > nlevels) return;
    if (l == nlevels) { out[l] = n; return; }
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (box_level[boxes[mid]] < l) lo = mid + 1; else hi = mid;
    }
    out[l] = lo;
}

__global__ void k_inverse_map(int n, const int *boxes, int *map)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[boxes[i]] = i;
}

// Coarsest box level, into *out: with starts == nullptr over the n list
// entries (boxes named by entries[i]), otherwise over the boxes[i] whose list
// starts[i] .. starts[i + 1] is nonempty.  Minimum per wavefront, then one
// atomic per wavefront.
__global__ void __launch_bounds__(256) k_min_level(int n, const int *entries, const int *starts,
                                                   const int *boxes, const int *level, int *out)
{
    int m = INT_MAX;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        if (!starts)
            m = min(m, level[entries[i]]);
        else if (starts[i + 1] > starts[i])
            m = min(m, level[boxes[i]]);
    }
    for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m != INT_MAX) atomicMin(out, m);
}

static void min_level(hipStream_t st, int n, const int *entries, const int *starts,
                      const int *boxes, const int *level, int *out)
{
    if (n <= 0) return;
    const int nb = (int)std::min<long>(cdiv(n, 256), 512);
    hipLaunchKernelGGL(k_min_level, dim3(nb), dim3(256), 0, st, n, entries, starts, boxes, level,
                       out);
}

static int level_starts(hipStream_t st, const qbx_tree *T, const dbuf<int> &boxes, int n,
                        dbuf<int> &d_out, std::vector<int> &h_out)
{
    QBX_TRY(d_out.alloc(T->nlevels + 1));
    hipLaunchKernelGGL(k_level_starts, dim3(1), dim3(MAX_LEVELS + 1), 0, st, T->nlevels,
                       boxes.p, n, T->box_levels.p, d_out.p);
    QBX_HIP(hipGetLastError());
    h_out.resize(T->nlevels + 1);
    QBX_HIP(hipMemcpyAsync(h_out.data(), d_out.p, sizeof(int) * (T->nlevels + 1),
                           hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    return QBX_OK;
}

static int build_lists(qbx_ctx *ctx, hipStream_t st, int nitems, dbuf<int> &starts,
                       dbuf<int> &lists, int *total, dbuf<int> &cnt)
{
    QBX_TRY(starts.alloc(nitems + 1));
    QBX_TRY(exclusive_scan(ctx, st, cnt.p, starts.p, nitems, total));
    QBX_TRY(lists.alloc(std::max(*total, 1)));
    lists.n = *total;
    return QBX_OK;
}

static int traverse(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, int n, int ext,
                    int minsrc, qbx_trav *R)
{
    TV t;
    t.dim = T->dim;
    t.nboxes = T->nboxes;
    t.nlevels = T->nlevels;
    t.level = T->box_levels.p;
    t.parent = T->box_parent_ids.p;
    t.child = T->box_child_ids.p;
    t.coords = T->box_coords.p;
    t.lstart = T->d_level_start_box_nrs.p;
    t.morton = T->box_morton.p;
    t.src_nonchild = T->box_source_counts_nonchild.p;
    t.src_cumul = T->box_source_counts_cumul.p;
    int nb = T->nboxes;
    R->nlevels = T->nlevels;
    R->nsep = n;

    dbuf<int> f_src, f_srcpar, f_tgt, f_ttp;
    QBX_TRY(f_src.alloc(nb));
    QBX_TRY(f_srcpar.alloc(nb));
    QBX_TRY(f_tgt.alloc(nb));
    QBX_TRY(f_ttp.alloc(nb));
    QBX_HIP(hipMemsetAsync(f_ttp.p, 0, sizeof(int) * nb, st));
    hipLaunchKernelGGL(k_box_flags, dim3(cdiv(nb, 256)), dim3(256), 0, st, t,
                       T->box_target_counts_nonchild.p, f_src.p, f_srcpar.p, f_tgt.p, f_ttp.p);
    QBX_HIP(hipGetLastError());
    QBX_TRY(compact_indices(ctx, st, f_src.p, nb, R->source_boxes, &R->nsource_boxes));
    QBX_TRY(compact_indices(ctx, st, f_srcpar.p, nb, R->source_parent_boxes,
                            &R->nsource_parent_boxes));
    QBX_TRY(compact_indices(ctx, st, f_tgt.p, nb, R->target_boxes, &R->ntarget_boxes));
    QBX_TRY(compact_indices(ctx, st, f_ttp.p, nb, R->ttp_boxes, &R->nttp_boxes));
    QBX_TRY(level_starts(st, T, R->source_boxes, R->nsource_boxes, R->level_start_source_box_nrs,
                         R->h_level_start_source_box_nrs));
    QBX_TRY(level_starts(st, T, R->source_parent_boxes, R->nsource_parent_boxes,
                         R->level_start_source_parent_box_nrs,
                         R->h_level_start_source_parent_box_nrs));
    QBX_TRY(level_starts(st, T, R->target_boxes, R->ntarget_boxes, R->level_start_target_box_nrs,
                         R->h_level_start_target_box_nrs));
    QBX_TRY(level_starts(st, T, R->ttp_boxes, R->nttp_boxes, R->level_start_ttp_box_nrs,
                         R->h_level_start_ttp_box_nrs));
    QBX_TRY(R->box_to_ttp.alloc(nb));
    QBX_TRY(R->box_to_target_box.alloc(nb));
    QBX_HIP(hipMemsetAsync(R->box_to_ttp.p, 0xff, sizeof(int) * nb, st));
    QBX_HIP(hipMemsetAsync(R->box_to_target_box.p, 0xff, sizeof(int) * nb, st));
    if (R->nttp_boxes)
        hipLaunchKernelGGL(k_inverse_map, dim3(cdiv(R->nttp_boxes, 256)), dim3(256), 0, st,
                           R->nttp_boxes, R->ttp_boxes.p, R->box_to_ttp.p);
    if (R->ntarget_boxes)
        hipLaunchKernelGGL(k_inverse_map, dim3(cdiv(R->ntarget_boxes, 256)), dim3(256), 0, st,
                           R->ntarget_boxes, R->target_boxes.p, R->box_to_target_box.p);
    QBX_HIP(hipGetLastError());

    // err[0]: stack overflow flag; err[1], err[2]: the coarsest levels of
    // qbx_trav::mpole_min_level and local_min_level
    dbuf<int> err;
    QBX_TRY(err.alloc(3));
    const int err_init[3] = {0, T->nlevels, T->nlevels};
    QBX_HIP(hipMemcpyAsync(err.p, err_init, sizeof(err_init), hipMemcpyHostToDevice, st));
    // lists 1 and 3
    {
        int ntb = R->ntarget_boxes;
        dbuf<int> c1, c3;
        QBX_TRY(c1.alloc(std::max(ntb, 1)));
        QBX_TRY(c3.alloc(std::max(ntb, 1)));
        ListOut o1{c1.p, nullptr, nullptr, 0}, o3{c3.p, nullptr, nullptr, 0};
        if (ntb)
            hipLaunchKernelGGL(k_list13, dim3(cdiv(ntb, 64)), dim3(64), 0, st, t, n, ext, minsrc,
                               ntb, R->target_boxes.p, o1, o3, err.p);
        QBX_HIP(hipGetLastError());
        QBX_TRY(build_lists(ctx, st, ntb, R->l1_starts, R->l1_lists, &R->n1, c1));
        QBX_TRY(build_lists(ctx, st, ntb, R->l3_starts, R->l3_lists, &R->n3, c3));
        o1 = ListOut{nullptr, R->l1_starts.p, R->l1_lists.p, 1};
        o3 = ListOut{nullptr, R->l3_starts.p, R->l3_lists.p, 1};
        if (ntb) {
            hipLaunchKernelGGL(k_list13, dim3(cdiv(ntb, 64)), dim3(64), 0, st, t, n, ext, minsrc,
                               ntb, R->target_boxes.p, o1, o3, err.p);
            hipLaunchKernelGGL(k_sort_lists, dim3(cdiv(ntb, 64)), dim3(64), 0, st, ntb,
                               R->l1_starts.p, R->l1_lists.p);
            hipLaunchKernelGGL(k_sort_lists, dim3(cdiv(ntb, 64)), dim3(64), 0, st, ntb,
                               R->l3_starts.p, R->l3_lists.p);
        }
        QBX_HIP(hipGetLastError());
    }
    // lists 2 and 4
    {
        int nt = R->nttp_boxes;
        dbuf<int> c2, c4;
        QBX_TRY(c2.alloc(std::max(nt, 1)));
        QBX_TRY(c4.alloc(std::max(nt, 1)));
        ListOut o2{c2.p, nullptr, nullptr, 0}, o4{c4.p, nullptr, nullptr, 0};
        if (nt)
            hipLaunchKernelGGL(k_list24, dim3(cdiv(nt, 64)), dim3(64), 0, st, t, n, nt,
                               R->ttp_boxes.p, o2, o4);
        QBX_HIP(hipGetLastError());
        QBX_TRY(build_lists(ctx, st, nt, R->l2_starts, R->l2_lists, &R->n2, c2));
        QBX_TRY(build_lists(ctx, st, nt, R->l4_starts, R->l4_lists, &R->n4, c4));
        o2 = ListOut{nullptr, R->l2_starts.p, R->l2_lists.p, 1};
        o4 = ListOut{nullptr, R->l4_starts.p, R->l4_lists.p, 1};
        if (nt) {
            hipLaunchKernelGGL(k_list24, dim3(cdiv(nt, 64)), dim3(64), 0, st, t, n, nt,
                               R->ttp_boxes.p, o2, o4);
            hipLaunchKernelGGL(k_sort_lists, dim3(cdiv(nt, 64)), dim3(64), 0, st, nt,
                               R->l2_starts.p, R->l2_lists.p);
            hipLaunchKernelGGL(k_sort_lists, dim3(cdiv(nt, 64)), dim3(64), 0, st, nt,
                               R->l4_starts.p, R->l4_lists.p);
        }
        QBX_HIP(hipGetLastError());
    }
    min_level(st, R->n2, R->l2_lists.p, nullptr, nullptr, T->box_levels.p, err.p + 1);
    min_level(st, R->n3, R->l3_lists.p, nullptr, nullptr, T->box_levels.p, err.p + 1);
    min_level(st, R->nttp_boxes, nullptr, R->l2_starts.p, R->ttp_boxes.p, T->box_levels.p,
              err.p + 2);
    min_level(st, R->nttp_boxes, nullptr, R->l4_starts.p, R->ttp_boxes.p, T->box_levels.p,
              err.p + 2);
    QBX_HIP(hipGetLastError());
    int herr[3] = {0, 0, 0};
    QBX_HIP(hipMemcpyAsync(herr, err.p, sizeof(herr), hipMemcpyDeviceToHost, st));
    QBX_HIP(hipStreamSynchronize(st));
    QBX_CHECK(herr[0] == 0, "qbx_traverse: traversal stack overflow");
    R->mpole_min_level = herr[1];
    R->local_min_level = herr[2];
    return QBX_OK;
}

struct TFieldRef {
    const void *p;
    long bytes;
};

static TFieldRef trav_field(const qbx_trav *r, const char *f)
{
    std::string s(f);
    auto I = [&](const dbuf<int> &b, long n) { return TFieldRef{b.p, (long)sizeof(int) * n}; };
    int L1 = r->nlevels + 1;
    if (s == "source_boxes") return I(r->source_boxes, r->nsource_boxes);
    if (s == "source_parent_boxes") return I(r->source_parent_boxes, r->nsource_parent_boxes);
    if (s == "target_boxes") return I(r->target_boxes, r->ntarget_boxes);
    if (s == "target_or_target_parent_boxes") return I(r->ttp_boxes, r->nttp_boxes);
    if (s == "level_start_source_box_nrs") return I(r->level_start_source_box_nrs, L1);
    if (s == "level_start_source_parent_box_nrs") return I(r->level_start_source_parent_box_nrs, L1);
    if (s == "level_start_target_box_nrs") return I(r->level_start_target_box_nrs, L1);
    if (s == "level_start_target_or_target_parent_box_nrs") return I(r->level_start_ttp_box_nrs, L1);
    if (s == "neighbor_source_boxes_starts") return I(r->l1_starts, r->ntarget_boxes + 1);
    if (s == "neighbor_source_boxes_lists") return I(r->l1_lists, r->n1);
    if (s == "from_sep_siblings_starts") return I(r->l2_starts, r->nttp_boxes + 1);
    if (s == "from_sep_siblings_lists") return I(r->l2_lists, r->n2);
    if (s == "from_sep_smaller_starts") return I(r->l3_starts, r->ntarget_boxes + 1);
    if (s == "from_sep_smaller_lists") return I(r->l3_lists, r->n3);
    if (s == "from_sep_bigger_starts") return I(r->l4_starts, r->nttp_boxes + 1);
    if (s == "from_sep_bigger_lists") return I(r->l4_lists, r->n4);
    return TFieldRef{nullptr, -1};
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_traverse(qbx_ctx *ctx, void *stream, const qbx_tree *tree, int well_sep_is_n_away,
                 int targets_have_extent, int from_sep_smaller_min_nsources_cumul,
                 qbx_trav **out)
{
    QBX_CHECK(ctx && tree && out, "qbx_traverse: null argument");
    QBX_CHECK(well_sep_is_n_away >= 1 && well_sep_is_n_away <= 3,
              "qbx_traverse: well_sep_is_n_away must be in [1, 3]");
    qbx_trav *R = new qbx_trav();
    int rc = traverse(ctx, S(stream), tree, well_sep_is_n_away, targets_have_extent ? 1 : 0,
                      from_sep_smaller_min_nsources_cumul, R);
    if (rc != QBX_OK) {
        delete R;
        return rc;
    }
    *out = R;
    return QBX_OK;
}

int qbx_trav_destroy(qbx_trav *t)
{
    delete t;
    return QBX_OK;
}

int qbx_trav_sizes(const qbx_trav *r, int *c)
{
    QBX_CHECK(r && c, "qbx_trav_sizes: null argument");
    c[0] = r->nsource_boxes;
    c[1] = r->nsource_parent_boxes;
    c[2] = r->ntarget_boxes;
    c[3] = r->nttp_boxes;
    c[4] = r->n1;
    c[5] = r->n2;
    c[6] = r->n3;
    c[7] = r->n4;
    return QBX_OK;
}

long qbx_trav_field_bytes(const qbx_trav *r, const char *field)
{
    if (!r || !field) return -1;
    return trav_field(r, field).bytes;
}

int qbx_trav_export(const qbx_trav *r, const char *field, void *dst, long nbytes, void *stream)
{
    QBX_CHECK(r && field, "qbx_trav_export: null argument");
    TFieldRef f = trav_field(r, field);
    QBX_CHECK(f.bytes >= 0, "qbx_trav_export: unknown field '%s'", field);
    QBX_CHECK(nbytes >= f.bytes, "qbx_trav_export: '%s' needs %ld bytes", field, f.bytes);
    if (f.bytes > 0)
        QBX_HIP(hipMemcpyAsync(dst, f.p, f.bytes, hipMemcpyDeviceToDevice, S(stream)));
    return QBX_OK;
}

}  // extern "C"

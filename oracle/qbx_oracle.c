/*
 * CPU oracle for the QBX-FMM hot path (test infrastructure; see qbx_oracle.h).
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * -ffp-contract=off matters: the tree build and target association make FP
 * comparisons whose rounding the HIP path reproduces operation by operation.
 */
#define _GNU_SOURCE
#include <complex.h>
#undef I
#define JI _Complex_I
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "qbx_oracle.h"

#define PI 3.14159265358979323846

/* sqrt(dim)/2 as decimal literals shared with the HIP code */
static const double SQRT_DIM_HALF[4] = {0.0, 0.5, 0.70710678118654757, 0.86602540378443865};

static int maxl_for_dim(int dim) { return dim == 2 ? 28 : 19; }

/* ------------------------------------------------------------------ */
/* growable int vector                                                 */
/* ------------------------------------------------------------------ */
typedef struct { int *a; int n, cap; } ivec;
static void iv_push(ivec *v, int x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->a = (int *)realloc(v->a, sizeof(int) * v->cap);
    }
    v->a[v->n++] = x;
}
static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------ */
/* Tree build                                                          */
/*                                                                     */
/* Restates boxtree's TreeBuilder(kind="adaptive") as used by          */
/* pytential/qbx/geometry.py:562-618: particles = sources + targets,   */
/* targets may carry a radius (stick-out, extent_norm l2/linf), a box  */
/* is split while its refine weight exceeds max_leaf_refine_weight.    */
/* Exact spec (shared with the HIP builder):                           */
/*  - root: bbox of all points; extent R = max side * (1 + 2^-30)      */
/*  - integer coords q = floor((x - min) / R * 2^L), L = 28 (2D), 19   */
/*  - stop level of a particle with radius r: deepest level l such     */
/*    that it fits the level-j box for all j <= l; l2 fit: with        */
/*    A = ((1+s) * sqrt(d)/2) * h, A - r >= 0 and |x-c|^2 <= (A-r)^2   */
/*  - level-synchronous split: box weight = sum of weights of the      */
/*    particles in its subtree; split iff weight > W and level < L;    */
/*    particles that cannot descend stay (non-child particles)         */
/*  - boxes numbered breadth-first, children in Morton child order     */
/*  - particle tree order: by pre-order (DFS) rank of the owner box,   */
/*    then by user index (stable)                                      */
/* ------------------------------------------------------------------ */
static int particle_fits(int dim, const double *x, double r, const int *q, int L,
                         int level, const double *mn, double R, double K,
                         int linf, double half_plus_s)
{
    double h = ldexp(R, -level);
    double d2 = 0.0;
    int ok = 1;
    for (int d = 0; d < dim; ++d) {
        int I = q[d] >> (L - level);
        double c = mn[d] + ((double)I + 0.5) * h;
        double dd = x[d] - c;
        if (linf) {
            if (fabs(dd) + r > half_plus_s * h) ok = 0;
        } else {
            d2 = d2 + dd * dd;
        }
    }
    if (linf) return ok;
    double thr = K * h - r;
    if (thr < 0.0) return 0;
    return d2 <= thr * thr;
}

int orc_tree_build(int dim,
                   int nsources, const double *sources,
                   int ntargets, const double *targets, const double *target_radii,
                   const int *source_weights, const int *target_weights,
                   int max_leaf_refine_weight, double stick_out_factor,
                   int extent_norm_linf, orc_tree *out)
{
    memset(out, 0, sizeof(*out));
    const int nchild = 1 << dim;
    const int np = nsources + ntargets;
    const int L = maxl_for_dim(dim);

    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    for (int d = 0; d < dim; ++d) {
        for (int i = 0; i < nsources; ++i) {
            double v = sources[(size_t)d * nsources + i];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
        for (int i = 0; i < ntargets; ++i) {
            double v = targets[(size_t)d * ntargets + i];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
    }
    if (np == 0) for (int d = 0; d < dim; ++d) { mn[d] = 0; mx[d] = 0; }
    double ext = 0.0;
    for (int d = 0; d < dim; ++d) if (mx[d] - mn[d] > ext) ext = mx[d] - mn[d];
    double R = ext > 0.0 ? ext * (1.0 + 0x1p-30) : 1.0;
    double scale = ldexp(1.0, L);
    int qmax = (1 << L) - 1;

    int *q = (int *)malloc(sizeof(int) * (size_t)dim * (np ? np : 1));
    int *stop = (int *)malloc(sizeof(int) * (np ? np : 1));
    int *w = (int *)malloc(sizeof(int) * (np ? np : 1));
    const double K = (1.0 + stick_out_factor) * SQRT_DIM_HALF[dim];
    const double half_plus_s = 0.5 + stick_out_factor;

    for (int p = 0; p < np; ++p) {
        double x[3];
        for (int d = 0; d < dim; ++d)
            x[d] = p < nsources ? sources[(size_t)d * nsources + p]
                                : targets[(size_t)d * ntargets + (p - nsources)];
        int qq[3];
        for (int d = 0; d < dim; ++d) {
            double u = (x[d] - mn[d]) / R * scale;
            double f = floor(u);
            int qi = f < 0 ? 0 : (f > qmax ? qmax : (int)f);
            qq[d] = qi;
            q[(size_t)d * np + p] = qi;
        }
        if (p < nsources) {
            w[p] = source_weights ? source_weights[p] : 1;
            stop[p] = L;
        } else {
            int it = p - nsources;
            w[p] = target_weights ? target_weights[it] : 1;
            double r = target_radii ? target_radii[it] : 0.0;
            if (r > 0.0) {
                int l = 0;
                while (l < L && particle_fits(dim, x, r, qq, L, l + 1, mn, R, K,
                                              extent_norm_linf, half_plus_s))
                    ++l;
                stop[p] = l;
            } else {
                stop[p] = L;
            }
        }
    }

    /* level-synchronous construction */
    int cap = 1024;
    int *blevel = (int *)malloc(sizeof(int) * cap);
    int *bparent = (int *)malloc(sizeof(int) * cap);
    int *bcoord = (int *)malloc(sizeof(int) * cap * dim);
    int *bchild = (int *)malloc(sizeof(int) * cap * nchild);
    ivec lstart = {0};
    int nboxes = 1;
    blevel[0] = 0; bparent[0] = -1;
    for (int d = 0; d < dim; ++d) bcoord[d] = 0;
    for (int c = 0; c < nchild; ++c) bchild[c] = -1;
    iv_push(&lstart, 0); iv_push(&lstart, 1);

    int *cur = (int *)calloc(np ? np : 1, sizeof(int));
    int *owner = (int *)malloc(sizeof(int) * (np ? np : 1));
    for (int p = 0; p < np; ++p) owner[p] = -1;

    for (int l = 0;; ++l) {
        int lo = lstart.a[l], hi = lstart.a[l + 1];
        if (lo == hi) break;
        int nb = hi - lo;
        long long *wt = (long long *)calloc(nb, sizeof(long long));
        char *split = (char *)calloc(nb, 1);
        int *cid = (int *)malloc(sizeof(int) * (size_t)nb * nchild);
        char *cflag = (char *)calloc((size_t)nb * nchild, 1);
        for (int p = 0; p < np; ++p)
            if (owner[p] < 0) wt[cur[p] - lo] += w[p];
        for (int b = 0; b < nb; ++b)
            split[b] = (wt[b] > max_leaf_refine_weight) && (l < L);
        for (int p = 0; p < np; ++p) {
            if (owner[p] >= 0) continue;
            int b = cur[p] - lo;
            if (!split[b] || stop[p] == l) { owner[p] = cur[p]; continue; }
            int c = 0;
            for (int d = 0; d < dim; ++d)
                c |= ((q[(size_t)d * np + p] >> (L - l - 1)) & 1) << d;
            cflag[(size_t)b * nchild + c] = 1;
        }
        for (int b = 0; b < nb; ++b) {
            for (int c = 0; c < nchild; ++c) {
                cid[(size_t)b * nchild + c] = -1;
                if (!cflag[(size_t)b * nchild + c]) continue;
                if (nboxes == cap) {
                    cap *= 2;
                    blevel = (int *)realloc(blevel, sizeof(int) * cap);
                    bparent = (int *)realloc(bparent, sizeof(int) * cap);
                    bcoord = (int *)realloc(bcoord, sizeof(int) * cap * dim);
                    bchild = (int *)realloc(bchild, sizeof(int) * cap * nchild);
                }
                int id = nboxes++;
                blevel[id] = l + 1;
                bparent[id] = lo + b;
                for (int d = 0; d < dim; ++d)
                    bcoord[(size_t)id * dim + d] = 2 * bcoord[(size_t)(lo + b) * dim + d] + ((c >> d) & 1);
                for (int cc = 0; cc < nchild; ++cc) bchild[(size_t)id * nchild + cc] = -1;
                bchild[(size_t)(lo + b) * nchild + c] = id;
                cid[(size_t)b * nchild + c] = id;
            }
        }
        for (int p = 0; p < np; ++p) {
            if (owner[p] >= 0) continue;
            int b = cur[p] - lo;
            int c = 0;
            for (int d = 0; d < dim; ++d)
                c |= ((q[(size_t)d * np + p] >> (L - l - 1)) & 1) << d;
            cur[p] = cid[(size_t)b * nchild + c];
        }
        iv_push(&lstart, nboxes);
        free(wt); free(split); free(cid); free(cflag);
    }
    int nlevels = lstart.n - 2;  /* last level range is empty */

    /* pre-order rank */
    int *rank = (int *)malloc(sizeof(int) * nboxes);
    int *byrank = (int *)malloc(sizeof(int) * nboxes);
    {
        int *stack = (int *)malloc(sizeof(int) * (nboxes + 1));
        int sp = 0, r = 0;
        stack[sp++] = 0;
        while (sp) {
            int b = stack[--sp];
            rank[b] = r; byrank[r] = b; ++r;
            for (int c = nchild - 1; c >= 0; --c) {
                int ch = bchild[(size_t)b * nchild + c];
                if (ch >= 0) stack[sp++] = ch;
            }
        }
        free(stack);
    }

    out->dim = dim; out->nsources = nsources; out->ntargets = ntargets;
    out->nboxes = nboxes; out->nlevels = nlevels; out->maxl = L;
    for (int d = 0; d < dim; ++d) out->root_min[d] = mn[d];
    out->root_extent = R;
    out->box_levels = (int *)malloc(sizeof(int) * nboxes);
    out->box_parent_ids = (int *)malloc(sizeof(int) * nboxes);
    out->box_child_ids = (int *)malloc(sizeof(int) * nboxes * nchild);
    out->box_coords = (int *)malloc(sizeof(int) * nboxes * dim);
    out->box_centers = (double *)malloc(sizeof(double) * nboxes * dim);
    for (int b = 0; b < nboxes; ++b) {
        out->box_levels[b] = blevel[b];
        out->box_parent_ids[b] = bparent[b];
        for (int c = 0; c < nchild; ++c)
            out->box_child_ids[(size_t)c * nboxes + b] = bchild[(size_t)b * nchild + c];
        double h = ldexp(R, -blevel[b]);
        for (int d = 0; d < dim; ++d) {
            int I = bcoord[(size_t)b * dim + d];
            out->box_coords[(size_t)d * nboxes + b] = I;
            out->box_centers[(size_t)d * nboxes + b] = mn[d] + ((double)I + 0.5) * h;
        }
    }
    out->level_start_box_nrs = (int *)malloc(sizeof(int) * (nlevels + 1));
    for (int l = 0; l <= nlevels; ++l) out->level_start_box_nrs[l] = lstart.a[l];

    /* particle order + box ranges, separately for sources and targets */
    for (int kind = 0; kind < 2; ++kind) {
        int n = kind == 0 ? nsources : ntargets;
        int off = kind == 0 ? 0 : nsources;
        int *nonchild = (int *)calloc(nboxes, sizeof(int));
        for (int i = 0; i < n; ++i) nonchild[owner[off + i]]++;
        int *rstart = (int *)malloc(sizeof(int) * (nboxes + 1));
        rstart[0] = 0;
        for (int r = 0; r < nboxes; ++r) rstart[r + 1] = rstart[r] + nonchild[byrank[r]];
        int *fill = (int *)malloc(sizeof(int) * nboxes);
        for (int r = 0; r < nboxes; ++r) fill[r] = rstart[r];
        int *perm = (int *)malloc(sizeof(int) * (n ? n : 1));
        for (int i = 0; i < n; ++i) perm[fill[rank[owner[off + i]]]++] = i;
        int *starts = (int *)malloc(sizeof(int) * nboxes);
        int *cumul = (int *)malloc(sizeof(int) * nboxes);
        for (int b = 0; b < nboxes; ++b) { starts[b] = rstart[rank[b]]; cumul[b] = nonchild[b]; }
        for (int b = nboxes - 1; b > 0; --b) cumul[bparent[b]] += cumul[b];
        if (kind == 0) {
            out->user_source_ids = perm;
            out->box_source_starts = starts;
            out->box_source_counts_nonchild = nonchild;
            out->box_source_counts_cumul = cumul;
        } else {
            out->sorted_target_ids = perm;
            out->box_target_starts = starts;
            out->box_target_counts_nonchild = nonchild;
            out->box_target_counts_cumul = cumul;
        }
        free(rstart); free(fill);
    }

    free(q); free(stop); free(w); free(blevel); free(bparent); free(bcoord); free(bchild);
    free(lstart.a); free(cur); free(owner); free(rank); free(byrank);
    return 0;
}

void orc_tree_free(orc_tree *t)
{
    free(t->box_levels); free(t->box_parent_ids); free(t->box_child_ids);
    free(t->box_coords); free(t->box_centers);
    free(t->box_source_starts); free(t->box_source_counts_nonchild); free(t->box_source_counts_cumul);
    free(t->box_target_starts); free(t->box_target_counts_nonchild); free(t->box_target_counts_cumul);
    free(t->level_start_box_nrs); free(t->user_source_ids); free(t->sorted_target_ids);
    memset(t, 0, sizeof(*t));
}

/* ------------------------------------------------------------------ */
/* Traversal                                                           */
/*                                                                     */
/* Restates boxtree's FMMTraversalBuilder + merge_close_lists as used  */
/* by pytential/qbx/geometry.py:622-641 (well_sep_is_n_away=n), as an  */
/* integer spec on box index coordinates:                              */
/*  gap(A,B): l-inf gap between the boxes in units of the smaller box  */
/*  sep(S,X)  (S not smaller than X): gap(S,X) >= n                    */
/*  sep3(D,B) (D smaller than B):     gap(D, B grown by e boxes) >= n  */
/*            e = 1 when targets carry extent (stick-out), else 0      */
/*  list1(B): leaves with sources, (a) not smaller than B with         */
/*            !sep(S,B), (b) below a same-level colleague of B, not    */
/*            under any sep3 box; plus leaves of sep3 boxes holding    */
/*            fewer than min_nsources_cumul sources (merged "close")   */
/*  list2(P): children of colleagues of parent(P), not colleagues of P */
/*  list3(B): first sep3 boxes below colleagues of B                   */
/*  list4(P): leaves S bigger than P with sep(S,P), !sep(S,parent(P))  */
/*  all lists sorted by box number.                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    const orc_tree *t;
    int dim, nchild;
    uint64_t *morton;   /* per box */
} tctx;

static uint64_t morton_of(int dim, const int *I, int level)
{
    uint64_t k = 0;
    for (int i = 0; i < level; ++i)
        for (int d = 0; d < dim; ++d)
            k |= (uint64_t)((I[d] >> i) & 1) << (i * dim + d);
    return k;
}

static int box_coord(const orc_tree *t, int b, int d) { return t->box_coords[(size_t)d * t->nboxes + b]; }

static int find_box(const tctx *c, int level, const int *I)
{
    const orc_tree *t = c->t;
    if (level >= t->nlevels) return -1;
    for (int d = 0; d < c->dim; ++d)
        if (I[d] < 0 || I[d] >= (1 << level)) return -1;
    uint64_t key = morton_of(c->dim, I, level);
    int lo = t->level_start_box_nrs[level], hi = t->level_start_box_nrs[level + 1];
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (c->morton[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo < t->level_start_box_nrs[level + 1] && c->morton[lo] == key) return lo;
    return -1;
}

/* gap between boxes a and b, extending b by `grow` boxes of its own size */
static long long box_gap(const orc_tree *t, int a, int b, int grow)
{
    int la = t->box_levels[a], lb = t->box_levels[b];
    int m = la > lb ? la : lb;
    long long g = 0;
    for (int d = 0; d < t->dim; ++d) {
        long long sa = 1LL << (m - la), sb = 1LL << (m - lb);
        long long alo = (long long)box_coord(t, a, d) * sa;
        long long ahi = (long long)(box_coord(t, a, d) + 1) * sa;
        long long blo = (long long)(box_coord(t, b, d) - grow) * sb;
        long long bhi = (long long)(box_coord(t, b, d) + 1 + grow) * sb;
        long long gd = 0;
        if (alo - bhi > gd) gd = alo - bhi;
        if (blo - ahi > gd) gd = blo - ahi;
        if (gd > g) g = gd;
    }
    return g;
}

static int is_leaf(const orc_tree *t, int b)
{
    int nchild = 1 << t->dim;
    for (int c = 0; c < nchild; ++c)
        if (t->box_child_ids[(size_t)c * t->nboxes + b] >= 0) return 0;
    return 1;
}

/* enumerate same-level colleagues (|dI|_inf <= n, including b itself) */
static int colleagues(const tctx *c, int b, int n, int *outv)
{
    const orc_tree *t = c->t;
    int dim = c->dim, level = t->box_levels[b];
    int I0[3], I[3], cnt = 0, w = 2 * n + 1;
    int total = 1;
    for (int d = 0; d < dim; ++d) { I0[d] = box_coord(t, b, d); total *= w; }
    for (int k = 0; k < total; ++k) {
        int kk = k;
        for (int d = 0; d < dim; ++d) { I[d] = I0[d] + (kk % w) - n; kk /= w; }
        int f = find_box(c, level, I);
        if (f >= 0) outv[cnt++] = f;
    }
    return cnt;
}

static void push_source_leaves(const orc_tree *t, int b, ivec *v)
{
    int nchild = 1 << t->dim;
    if (t->box_source_counts_cumul[b] == 0) return;
    if (is_leaf(t, b)) {
        if (t->box_source_counts_nonchild[b] > 0) iv_push(v, b);
        return;
    }
    for (int c = 0; c < nchild; ++c) {
        int ch = t->box_child_ids[(size_t)c * t->nboxes + b];
        if (ch >= 0) push_source_leaves(t, ch, v);
    }
}

static void csr_from_lists(ivec *lists, int n, int **starts, int **flat, int *total)
{
    int *s = (int *)malloc(sizeof(int) * (n + 1));
    s[0] = 0;
    for (int i = 0; i < n; ++i) s[i + 1] = s[i] + lists[i].n;
    int *f = (int *)malloc(sizeof(int) * (s[n] ? s[n] : 1));
    for (int i = 0; i < n; ++i) {
        if (lists[i].n) {
            qsort(lists[i].a, lists[i].n, sizeof(int), cmp_int);
            memcpy(f + s[i], lists[i].a, sizeof(int) * lists[i].n);
        }
        free(lists[i].a);
    }
    *starts = s; *flat = f; *total = s[n];
}

static void level_starts(const orc_tree *t, const int *boxes, int n, int **out)
{
    int *ls = (int *)malloc(sizeof(int) * (t->nlevels + 1));
    int j = 0;
    for (int l = 0; l <= t->nlevels; ++l) {
        while (j < n && t->box_levels[boxes[j]] < l) ++j;
        ls[l] = j;
    }
    ls[t->nlevels] = n;
    *out = ls;
}

int orc_traverse(const orc_tree *t, int n, int targets_have_extent,
                 int min_nsources_cumul, orc_trav *out)
{
    memset(out, 0, sizeof(*out));
    tctx c;
    c.t = t; c.dim = t->dim; c.nchild = 1 << t->dim;
    c.morton = (uint64_t *)malloc(sizeof(uint64_t) * t->nboxes);
    for (int b = 0; b < t->nboxes; ++b) {
        int I[3];
        for (int d = 0; d < t->dim; ++d) I[d] = box_coord(t, b, d);
        c.morton[b] = morton_of(t->dim, I, t->box_levels[b]);
    }
    const int e = targets_have_extent ? 1 : 0;
    const int nb = t->nboxes;
    out->nlevels = t->nlevels;

    ivec sb = {0}, spb = {0}, tb = {0}, ttp = {0};
    char *is_ttp = (char *)calloc(nb, 1);
    for (int b = 0; b < nb; ++b) {
        int leaf = is_leaf(t, b);
        if (leaf && t->box_source_counts_nonchild[b] > 0) iv_push(&sb, b);
        if (!leaf && t->box_source_counts_cumul[b] > 0) iv_push(&spb, b);
        if (t->box_target_counts_nonchild[b] > 0) {
            iv_push(&tb, b);
            for (int a = b; a >= 0 && !is_ttp[a]; a = t->box_parent_ids[a]) is_ttp[a] = 1;
        }
    }
    for (int b = 0; b < nb; ++b) if (is_ttp[b]) iv_push(&ttp, b);
    out->nsource_boxes = sb.n; out->source_boxes = sb.a ? sb.a : (int *)malloc(4);
    out->nsource_parent_boxes = spb.n; out->source_parent_boxes = spb.a ? spb.a : (int *)malloc(4);
    out->ntarget_boxes = tb.n; out->target_boxes = tb.a ? tb.a : (int *)malloc(4);
    out->nttp_boxes = ttp.n; out->ttp_boxes = ttp.a ? ttp.a : (int *)malloc(4);
    level_starts(t, out->source_boxes, sb.n, &out->level_start_source_box_nrs);
    level_starts(t, out->source_parent_boxes, spb.n, &out->level_start_source_parent_box_nrs);
    level_starts(t, out->target_boxes, tb.n, &out->level_start_target_box_nrs);
    level_starts(t, out->ttp_boxes, ttp.n, &out->level_start_ttp_box_nrs);

    int maxcol = 1;
    for (int d = 0; d < t->dim; ++d) maxcol *= (2 * n + 1);

    /* lists 1 and 3 (per target box) */
    ivec *l1 = (ivec *)calloc(tb.n ? tb.n : 1, sizeof(ivec));
    ivec *l3 = (ivec *)calloc(tb.n ? tb.n : 1, sizeof(ivec));
#pragma omp parallel
    {
        int *col = (int *)malloc(sizeof(int) * maxcol);
        int *stack = (int *)malloc(sizeof(int) * (size_t)(nb + 1));
#pragma omp for schedule(dynamic, 16)
        for (int it = 0; it < tb.n; ++it) {
            int B = out->target_boxes[it];
            /* (a) leaves not smaller than B */
            int anc[64], na = 0;
            for (int a = B; a >= 0; a = t->box_parent_ids[a]) anc[na++] = a;
            for (int k = na - 1; k >= 0; --k) {
                int P = anc[k];
                int nc = colleagues(&c, P, n, col);
                for (int j = 0; j < nc; ++j) {
                    int C = col[j];
                    if (!is_leaf(t, C) || t->box_source_counts_nonchild[C] == 0) continue;
                    if (box_gap(t, C, B, 0) < n) iv_push(&l1[it], C);
                }
            }
            /* (b) below same-level colleagues */
            int nc = colleagues(&c, B, n, col);
            for (int j = 0; j < nc; ++j) {
                int C = col[j];
                if (is_leaf(t, C)) continue;
                int sp = 0;
                for (int ch = c.nchild - 1; ch >= 0; --ch) {
                    int x = t->box_child_ids[(size_t)ch * nb + C];
                    if (x >= 0) stack[sp++] = x;
                }
                while (sp) {
                    int D = stack[--sp];
                    if (t->box_source_counts_cumul[D] == 0) continue;
                    if (box_gap(t, D, B, e) >= n) {
                        if (t->box_source_counts_cumul[D] < min_nsources_cumul)
                            push_source_leaves(t, D, &l1[it]);
                        else
                            iv_push(&l3[it], D);
                    } else if (is_leaf(t, D)) {
                        if (t->box_source_counts_nonchild[D] > 0) iv_push(&l1[it], D);
                    } else {
                        for (int ch = c.nchild - 1; ch >= 0; --ch) {
                            int x = t->box_child_ids[(size_t)ch * nb + D];
                            if (x >= 0) stack[sp++] = x;
                        }
                    }
                }
            }
        }
        free(col); free(stack);
    }
    csr_from_lists(l1, tb.n, &out->l1_starts, &out->l1_lists, &out->n1);
    csr_from_lists(l3, tb.n, &out->l3_starts, &out->l3_lists, &out->n3);
    free(l1); free(l3);

    /* lists 2 and 4 (per ttp box) */
    ivec *l2 = (ivec *)calloc(ttp.n ? ttp.n : 1, sizeof(ivec));
    ivec *l4 = (ivec *)calloc(ttp.n ? ttp.n : 1, sizeof(ivec));
#pragma omp parallel
    {
        int *col = (int *)malloc(sizeof(int) * maxcol);
#pragma omp for schedule(dynamic, 16)
        for (int it = 0; it < ttp.n; ++it) {
            int P = out->ttp_boxes[it];
            int lp = t->box_levels[P];
            if (lp == 0) continue;
            int Q = t->box_parent_ids[P];
            int nc = colleagues(&c, Q, n, col);
            for (int j = 0; j < nc; ++j) {
                int C = col[j];
                for (int ch = 0; ch < c.nchild; ++ch) {
                    int A = t->box_child_ids[(size_t)ch * nb + C];
                    if (A < 0 || t->box_source_counts_cumul[A] == 0) continue;
                    long long g = 0;
                    for (int d = 0; d < t->dim; ++d) {
                        long long dd = llabs((long long)box_coord(t, A, d) - box_coord(t, P, d));
                        if (dd > g) g = dd;
                    }
                    if (g > n) iv_push(&l2[it], A);
                }
            }
            int anc[64], na = 0;
            for (int a = Q; a >= 0; a = t->box_parent_ids[a]) anc[na++] = a;
            for (int k = na - 1; k >= 0; --k) {
                int nc2 = colleagues(&c, anc[k], n, col);
                for (int j = 0; j < nc2; ++j) {
                    int S = col[j];
                    if (!is_leaf(t, S) || t->box_source_counts_nonchild[S] == 0) continue;
                    if (box_gap(t, S, P, 0) >= n && box_gap(t, S, Q, 0) < n)
                        iv_push(&l4[it], S);
                }
            }
        }
        free(col);
    }
    csr_from_lists(l2, ttp.n, &out->l2_starts, &out->l2_lists, &out->n2);
    csr_from_lists(l4, ttp.n, &out->l4_starts, &out->l4_lists, &out->n4);
    free(l2); free(l4);
    free(is_ttp); free(c.morton);
    return 0;
}

void orc_trav_free(orc_trav *tr)
{
    free(tr->source_boxes); free(tr->level_start_source_box_nrs);
    free(tr->source_parent_boxes); free(tr->level_start_source_parent_box_nrs);
    free(tr->target_boxes); free(tr->level_start_target_box_nrs);
    free(tr->ttp_boxes); free(tr->level_start_ttp_box_nrs);
    free(tr->l1_starts); free(tr->l1_lists); free(tr->l2_starts); free(tr->l2_lists);
    free(tr->l3_starts); free(tr->l3_lists); free(tr->l4_starts); free(tr->l4_lists);
    memset(tr, 0, sizeof(*tr));
}

/* ------------------------------------------------------------------ */
/* Target association                                                  */
/*                                                                     */
/* Restates pytential/qbx/target_assoc.py:185-367, 853-952 as           */
/* predicates (the area queries there only accelerate them):           */
/*  marked(t)  <=> exists stage-1 source s: |t-s|^2 <= tunnel(s)^2      */
/*  candidate c for marked t <=> side(c)*flag(t) >= 0 and               */
/*     |t-c|^2 <= R_c^2, R_c = r_c (1 + tol); side(c) = 2 (c % 2) - 1   */
/*  per side: minimum of |t-c|^2 / R_c^2 (ties -> lowest c);            */
/*  choose plus iff both found and d2_plus < d2_minus, or only plus.    */
/*  status: 0 unmarked, 1 marked without center, 2 center found.        */
/* ------------------------------------------------------------------ */
static double dist2(int dim, const double *a, int na, int ia, const double *b, int nb_, int ib)
{
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        double dd = a[(size_t)d * na + ia] - b[(size_t)d * nb_ + ib];
        s = s + dd * dd;
    }
    return s;
}

typedef struct {
    int dim, n;
    double mn[3], h;
    int nc[3];
    int *cell_start, *items;
} grid_t;

static void grid_build(grid_t *g, int dim, int n, const double *x, double h)
{
    g->dim = dim; g->n = n;
    for (int d = 0; d < 3; ++d) { g->mn[d] = 0; g->nc[d] = 1; }
    double mx[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        g->mn[d] = INFINITY; mx[d] = -INFINITY;
        for (int i = 0; i < n; ++i) {
            double v = x[(size_t)d * n + i];
            if (v < g->mn[d]) g->mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
        if (n == 0) { g->mn[d] = 0; mx[d] = 0; }
    }
    double ext = 0;
    for (int d = 0; d < dim; ++d) if (mx[d] - g->mn[d] > ext) ext = mx[d] - g->mn[d];
    if (!(h > 0)) h = ext > 0 ? ext : 1.0;
    /* cap the cell count */
    long long tot;
    for (;;) {
        tot = 1;
        for (int d = 0; d < dim; ++d) {
            g->nc[d] = (int)((mx[d] - g->mn[d]) / h) + 1;
            tot *= g->nc[d];
        }
        if (tot <= 4LL * n + 64) break;
        h *= 1.5;
    }
    g->h = h;
    g->cell_start = (int *)calloc(tot + 1, sizeof(int));
    g->items = (int *)malloc(sizeof(int) * (n ? n : 1));
    int *cellof = (int *)malloc(sizeof(int) * (n ? n : 1));
    for (int i = 0; i < n; ++i) {
        long long cell = 0, mul = 1;
        for (int d = 0; d < dim; ++d) {
            int ci = (int)((x[(size_t)d * n + i] - g->mn[d]) / h);
            if (ci < 0) ci = 0;
            if (ci >= g->nc[d]) ci = g->nc[d] - 1;
            cell += ci * mul; mul *= g->nc[d];
        }
        cellof[i] = (int)cell;
        g->cell_start[cell + 1]++;
    }
    for (long long c2 = 0; c2 < tot; ++c2) g->cell_start[c2 + 1] += g->cell_start[c2];
    int *fill = (int *)malloc(sizeof(int) * (tot + 1));
    memcpy(fill, g->cell_start, sizeof(int) * (tot + 1));
    for (int i = 0; i < n; ++i) g->items[fill[cellof[i]]++] = i;
    free(fill); free(cellof);
}

static void grid_free(grid_t *g) { free(g->cell_start); free(g->items); }

/* visit all items in cells within one cell of point p (cells are >= radius) */
#define GRID_FOREACH(g, p, body)                                                      \
    do {                                                                              \
        int c0[3] = {0, 0, 0};                                                        \
        for (int d = 0; d < (g)->dim; ++d) {                                          \
            int ci = (int)floor(((p)[d] - (g)->mn[d]) / (g)->h);                      \
            c0[d] = ci;                                                               \
        }                                                                             \
        int span = 1;                                                                 \
        for (int d = 0; d < (g)->dim; ++d) span *= 3;                                 \
        for (int k = 0; k < span; ++k) {                                              \
            int kk = k, ok = 1;                                                       \
            long long cell = 0, mul = 1;                                              \
            for (int d = 0; d < (g)->dim; ++d) {                                      \
                int ci = c0[d] + (kk % 3) - 1; kk /= 3;                               \
                if (ci < 0 || ci >= (g)->nc[d]) { ok = 0; break; }                    \
                cell += ci * mul; mul *= (g)->nc[d];                                  \
            }                                                                         \
            if (!ok) continue;                                                        \
            for (int _q = (g)->cell_start[cell]; _q < (g)->cell_start[cell + 1]; ++_q) { \
                int item = (g)->items[_q];                                            \
                body                                                                  \
            }                                                                         \
        }                                                                             \
    } while (0)

int orc_target_assoc(int dim, int nsrc1, const double *src1, const double *tunnel,
                     int ncenters, const double *centers, const double *rtol,
                     int ntargets, const double *targets, const int *flags,
                     int accelerated, int *ttc, int *status)
{
    grid_t gs, gc;
    double maxt = 0, maxr = 0;
    for (int i = 0; i < nsrc1; ++i) if (tunnel[i] > maxt) maxt = tunnel[i];
    for (int i = 0; i < ncenters; ++i) if (rtol[i] > maxr) maxr = rtol[i];
    if (accelerated) {
        grid_build(&gs, dim, nsrc1, src1, maxt);
        grid_build(&gc, dim, ncenters, centers, maxr);
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int it = 0; it < ntargets; ++it) {
        double p[3];
        for (int d = 0; d < dim; ++d) p[d] = targets[(size_t)d * ntargets + it];
        int marked = 0;
        if (accelerated) {
            GRID_FOREACH(&gs, p, {
                if (!marked) {
                    double d2 = dist2(dim, targets, ntargets, it, src1, nsrc1, item);
                    if (d2 <= tunnel[item] * tunnel[item]) marked = 1;
                }
            });
        } else {
            for (int s = 0; s < nsrc1 && !marked; ++s) {
                double d2 = dist2(dim, targets, ntargets, it, src1, nsrc1, s);
                if (d2 <= tunnel[s] * tunnel[s]) marked = 1;
            }
        }
        ttc[it] = -1;
        if (!marked) { status[it] = 0; continue; }
        int best[2] = {-1, -1};
        double brel[2] = {INFINITY, INFINITY}, bd2[2] = {INFINITY, INFINITY};
        int fl = flags[it];
#define CONSIDER(cidx)                                                        \
        do {                                                                  \
            int side = 2 * ((cidx) % 2) - 1;                                  \
            if (side * fl < 0) break;                                         \
            double d2 = dist2(dim, targets, ntargets, it, centers, ncenters, (cidx)); \
            double R2 = rtol[(cidx)] * rtol[(cidx)];                          \
            if (!(d2 <= R2)) break;                                           \
            double rel = d2 / R2;                                             \
            int k = side > 0 ? 1 : 0;                                         \
            if (rel < brel[k] || (rel == brel[k] && (cidx) < best[k])) {      \
                brel[k] = rel; bd2[k] = d2; best[k] = (cidx);                 \
            }                                                                 \
        } while (0)
        if (accelerated) {
            GRID_FOREACH(&gc, p, { CONSIDER(item); });
        } else {
            for (int cc = 0; cc < ncenters; ++cc) CONSIDER(cc);
        }
#undef CONSIDER
        if (best[0] < 0 && best[1] < 0) { status[it] = 1; continue; }
        status[it] = 2;
        if (best[1] >= 0 && (best[0] < 0 || bd2[1] < bd2[0])) ttc[it] = best[1];
        else ttc[it] = best[0];
    }
    if (accelerated) { grid_free(&gs); grid_free(&gc); }
    return 0;
}

int orc_refine_flags_for_failed(int dim, int nsrc1, const double *src1, const double *tunnel,
                                int nelements, const int *el_starts,
                                int ntargets, const double *targets,
                                const int *status, int *refine_flags)
{
    for (int e = 0; e < nelements; ++e) refine_flags[e] = 0;
    for (int it = 0; it < ntargets; ++it) {
        if (status[it] != 1) continue;
        for (int s = 0; s < nsrc1; ++s) {
            double d2 = dist2(dim, targets, ntargets, it, src1, nsrc1, s);
            if (d2 <= tunnel[s] * tunnel[s]) {
                int lo = 0, hi = nelements;
                while (hi - lo > 1) {
                    int mid = (lo + hi) / 2;
                    if (el_starts[mid] <= s) lo = mid; else hi = mid;
                }
                refine_flags[lo] = 1;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* 2D expansions                                                       */
/*                                                                     */
/* Laplace, G = -log|x-y| / (2 pi) (sumpy LaplaceKernel(2)):           */
/*   complex potential Phi with phi = -Re(Phi)/(2 pi)                   */
/*   multipole: a0 log(z-c) + sum_k a_k (R/(z-c))^k                     */
/*   local:     sum_l b_l ((z-c)/R)^l                                   */
/*   (complex Taylor = the harmonic-polynomial Taylor space of sumpy's  */
/*    LaplaceConformingVolumeTaylor expansions in 2D, same truncation)  */
/* Helmholtz, G = (i/4) H0(k|x-y|) (sumpy HelmholtzKernel(2)), H2D:      */
/*   multipole: sum_n M_n H_n(k rho) e^{i n theta}                      */
/*   local:     sum_n L_n J_n(k rho) e^{i n theta},  |n| <= p           */
/* Double layer = directional source derivative along the normal.       */
/* ------------------------------------------------------------------ */
typedef double complex cplx;

static double binom_tab[130][130];
static int binom_ready = 0;
static void binom_init(void)
{
    if (binom_ready) return;
#pragma omp critical(binom)
    {
        if (!binom_ready) {
            for (int n = 0; n < 130; ++n) {
                binom_tab[n][0] = 1.0;
                for (int k = 1; k <= n; ++k)
                    binom_tab[n][k] = binom_tab[n - 1][k - 1] + (k < n ? binom_tab[n - 1][k] : 0.0);
                for (int k = n + 1; k < 130; ++k) binom_tab[n][k] = 0.0;
            }
            binom_ready = 1;
        }
    }
}
static inline double binom(int n, int k) { return (k < 0 || k > n) ? 0.0 : binom_tab[n][k]; }

static inline cplx cpt(const double *x, int n, int i)
{
    return x[i] + JI * x[(size_t)n + i];
}

void orc_bessel_jy(int nmax, double x, double *j, double *y)
{
    for (int n = 0; n <= nmax; ++n) {
        j[n] = jn(n, x);
        if (y) y[n] = yn(n, x);
    }
}

/* Z_n(k rho) e^{-i n theta} for n in [-m, m] (index n+m); hank selects H */
static void cyl_waves_conj(int hank, double k, cplx v, int m, cplx *out)
{
    double rho = cabs(v);
    double th = carg(v);
    for (int n = 0; n <= m; ++n) {
        double jv = jn(n, k * rho);
        cplx z = hank ? (jv + JI * yn(n, k * rho)) : jv;
        cplx e = cexp(-JI * (double)n * th);
        out[m + n] = z * e;
        double sgn = (n & 1) ? -1.0 : 1.0;
        out[m - n] = sgn * z * conj(e);  /* Z_{-n} e^{+i n th} */
    }
}

/* Z_l(k rho) e^{+i l theta} for l in [-m, m] (index l+m) */
static void cyl_waves(int hank, double k, cplx v, int m, cplx *out)
{
    double rho = cabs(v);
    double th = carg(v);
    for (int n = 0; n <= m; ++n) {
        double jv = jn(n, k * rho);
        cplx z = hank ? (jv + JI * yn(n, k * rho)) : jv;
        cplx e = cexp(JI * (double)n * th);
        out[m + n] = z * e;
        double sgn = (n & 1) ? -1.0 : 1.0;
        out[m - n] = sgn * z * conj(e);
    }
}

static inline int ncoeff_of(int kernel, int p) { return kernel == 0 ? p + 1 : 2 * p + 1; }

/* P2E: add source contributions to a multipole (local=0) or local (local=1) */
static void p2e_one(const orc_kparams *kp, int local, int p, cplx c, double R,
                    cplx s, cplx q, cplx nu, cplx *coef)
{
    int dip = kp->source_kind == 1;
    if (kp->kernel == 0) {
        cplx w = s - c;
        if (local) {
            cplx t = R / w, pw = 1.0;
            if (!dip) {
                coef[0] += q * clog(-w);
                for (int l = 1; l <= p; ++l) { pw *= t; coef[l] += -q / (double)l * pw; }
            } else {
                cplx f = nu * q / w;
                coef[0] += f;
                for (int l = 1; l <= p; ++l) { pw *= t; coef[l] += f * pw; }
            }
        } else {
            cplx t = w / R, pw = 1.0;
            if (!dip) {
                coef[0] += q;
                for (int kk = 1; kk <= p; ++kk) { pw *= t; coef[kk] += -q / (double)kk * pw; }
            } else {
                cplx f = nu * q / R;
                for (int kk = 1; kk <= p; ++kk) { coef[kk] += -f * pw; pw *= t; }
            }
        }
    } else {
        cplx g[2 * 130 + 3];
        int m = p + 1;
        cyl_waves_conj(local, kp->k, s - c, m, g);
        if (!dip) {
            for (int n = -p; n <= p; ++n) coef[n + p] += q * g[n + m];
        } else {
            double hk = 0.5 * kp->k;
            for (int n = -p; n <= p; ++n)
                coef[n + p] += q * hk * (conj(nu) * g[n - 1 + m] - nu * g[n + 1 + m]);
        }
    }
}

/* E2P of the derivative along d (complex dx + i dy) at t.  Laplace: the
 * potential is -Re F / (2 pi) with F analytic, the derivative -Re(F' d) /
 * (2 pi), F' of the local sum a_l ((t - c)/R)^l and of the multipole
 * a_0 log w + sum a_k (R/w)^k differentiated term by term.  Helmholtz: the
 * derivative of Z_n(k rho) e^{in theta} along d is
 * (k/2) (d Z_{n-1} e^{i(n-1) theta} - conj(d) Z_{n+1} e^{i(n+1) theta})
 * (Graf's recurrences), summed over the coefficients. */
static cplx e2p_dir(const orc_kparams *kp, int local, int p, cplx c, double R,
                    const cplx *coef, cplx t, cplx d)
{
    if (kp->kernel == 0) {
        cplx w = t - c, fp = 0;
        if (local) {
            cplx x = w / R, pw = 1.0;
            for (int l = 1; l <= p; ++l) { fp += (double)l * coef[l] * pw / R; pw *= x; }
        } else {
            cplx x = R / w, pw = x;
            fp = coef[0] / w;
            for (int kk = 1; kk <= p; ++kk) { fp -= (double)kk * coef[kk] * pw / w; pw *= x; }
        }
        return -creal(fp * d) / (2 * PI);
    } else {
        cplx g[2 * 131 + 1];
        int m = p + 1;
        cyl_waves(!local, kp->k, t - c, m, g);
        cplx phi = 0;
        for (int n = -p; n <= p; ++n)
            phi += coef[n + p] * 0.5 * kp->k * (d * g[n - 1 + m] - conj(d) * g[n + 1 + m]);
        return 0.25 * JI * phi;
    }
}

/* E2P: evaluate; returns the potential (kernel scaling included) */
static cplx e2p_one(const orc_kparams *kp, int local, int p, cplx c, double R,
                    const cplx *coef, cplx t)
{
    if (kp->kernel == 0) {
        cplx w = t - c, phi = 0;
        if (local) {
            cplx x = w / R;
            for (int l = p; l >= 0; --l) phi = phi * x + coef[l];
        } else {
            cplx x = R / w;
            for (int kk = p; kk >= 1; --kk) phi = (phi + coef[kk]) * x;
            phi += coef[0] * clog(w);
        }
        return -creal(phi) / (2 * PI);
    } else {
        cplx g[2 * 130 + 1];
        cyl_waves(!local, kp->k, t - c, p, g);
        cplx phi = 0;
        for (int n = -p; n <= p; ++n) phi += coef[n + p] * g[n + p];
        return 0.25 * JI * phi;
    }
}

/* E2E: kind 0 = M2M, 1 = M2L, 2 = L2L (adds into tgt) */
static void e2e_one(const orc_kparams *kp, int kind, int p1, int p2,
                    cplx c1, double R1, cplx c2, double R2, const cplx *a, cplx *b)
{
    if (kp->kernel == 0) {
        binom_init();
        if (kind == 0) {
            cplx z0 = c1 - c2;
            cplx u = z0 / R2, r12 = R1 / R2;
            b[0] += a[0];
            cplx upow[260], rpow[260];
            upow[0] = 1; rpow[0] = 1;
            for (int i = 1; i <= p2; ++i) upow[i] = upow[i - 1] * u;
            for (int i = 1; i <= p1; ++i) rpow[i] = rpow[i - 1] * r12;
            for (int l = 1; l <= p2; ++l) {
                cplx acc = -a[0] * upow[l] / (double)l;
                int kmax = l < p1 ? l : p1;
                for (int kk = 1; kk <= kmax; ++kk)
                    acc += a[kk] * rpow[kk] * upow[l - kk] * binom(l - 1, kk - 1);
                b[l] += acc;
            }
        } else if (kind == 1) {
            cplx z0 = c1 - c2;
            cplx s = R1 / z0, t = R2 / z0;
            cplx spow[260], tpow[260];
            spow[0] = 1; tpow[0] = 1;
            for (int i = 1; i <= p1; ++i) spow[i] = -spow[i - 1] * s;  /* (-1)^k (R1/z0)^k */
            for (int i = 1; i <= p2; ++i) tpow[i] = tpow[i - 1] * t;
            cplx acc0 = a[0] * clog(-z0);
            for (int kk = 1; kk <= p1; ++kk) acc0 += a[kk] * spow[kk];
            b[0] += acc0;
            for (int l = 1; l <= p2; ++l) {
                cplx acc = -a[0] / (double)l;
                for (int kk = 1; kk <= p1; ++kk)
                    acc += a[kk] * spow[kk] * binom(l + kk - 1, kk - 1);
                b[l] += acc * tpow[l];
            }
        } else {
            cplx z0 = c2 - c1;
            cplx u = z0 / R1;
            double r21 = R2 / R1;
            cplx upow[260];
            upow[0] = 1;
            for (int i = 1; i <= p1; ++i) upow[i] = upow[i - 1] * u;
            double rp = 1.0;
            for (int j = 0; j <= p2; ++j) {
                cplx acc = 0;
                for (int l = j; l <= p1; ++l) acc += a[l] * binom(l, j) * upow[l - j];
                b[j] += acc * rp;
                rp *= r21;
            }
        }
    } else {
        int m = p1 + p2;
        cplx g[2 * 260 + 1];
        cyl_waves(kind == 1, kp->k, c2 - c1, m, g);
        for (int mm = -p2; mm <= p2; ++mm) {
            cplx acc = 0;
            for (int n = -p1; n <= p1; ++n) acc += a[n + p1] * g[n - mm + m];
            b[mm + p2] += acc;
        }
    }
}

/* P2P of the derivative along d at t (closed forms of the kernels'
 * gradients in t; w = t - s, w^ = w / r):
 *   Laplace S -Re(d / w) / (2 pi), D -Re(nu d / w^2) / (2 pi)
 *   Helmholtz S -(i/4) k (d . w^) H1, D (i/4) k [H1 (d . nu - (d . w^)(nu . w^)) / r
 *     + k (nu . w^)(d . w^) H1'(k r)], H1' = H0 - H1 / x */
static cplx p2p_dir(const orc_kparams *kp, cplx s, cplx q, cplx nu, cplx t, cplx d)
{
    cplx w = t - s;
    double r = cabs(w);
    if (r == 0.0) return 0;
    if (kp->kernel == 0) {
        if (kp->source_kind == 0) return -creal(d / w) * q / (2 * PI);
        return -creal(nu * d / (w * w)) * q / (2 * PI);
    } else {
        double kr = kp->k * r;
        double pd = (creal(w) * creal(d) + cimag(w) * cimag(d)) / r;
        cplx h0 = j0(kr) + JI * y0(kr), h1 = j1(kr) + JI * y1(kr);
        if (kp->source_kind == 0) return -0.25 * JI * q * kp->k * pd * h1;
        double pn = (creal(w) * creal(nu) + cimag(w) * cimag(nu)) / r;
        double dn = creal(d) * creal(nu) + cimag(d) * cimag(nu);
        cplx v = h1 * (dn - pd * pn) / r + kp->k * pn * pd * (h0 - h1 / kr);
        return 0.25 * JI * q * kp->k * v;
    }
}

static cplx p2p_one(const orc_kparams *kp, cplx s, cplx q, cplx nu, cplx t)
{
    cplx w = t - s;
    double r = cabs(w);
    if (r == 0.0) return 0;
    if (kp->kernel == 0) {
        if (kp->source_kind == 0) return -q * log(r) / (2 * PI);
        return creal(nu / w) * q / (2 * PI);
    } else {
        double kr = kp->k * r;
        if (kp->source_kind == 0) return 0.25 * JI * q * (j0(kr) + JI * y0(kr));
        double proj = (creal(w) * creal(nu) + cimag(w) * cimag(nu)) / r;
        return 0.25 * JI * q * kp->k * (j1(kr) + JI * y1(kr)) * proj;
    }
}

/* the potential, or its derivative along d when kp->target_deriv */
static inline cplx e2p_at(const orc_kparams *kp, int local, int p, cplx c, double R,
                          const cplx *coef, cplx t, cplx d)
{
    return kp->target_deriv ? e2p_dir(kp, local, p, c, R, coef, t, d)
                            : e2p_one(kp, local, p, c, R, coef, t);
}

static inline cplx p2p_at(const orc_kparams *kp, cplx s, cplx q, cplx nu, cplx t, cplx d)
{
    return kp->target_deriv ? p2p_dir(kp, s, q, nu, t, d) : p2p_one(kp, s, q, nu, t);
}

static inline cplx dir_at(const orc_kparams *kp, int n, int i)
{
    return kp->target_deriv ? cpt(kp->target_dir, n, i) : 0;
}

/* ------------------------------------------------------------------ */
/* single-operator entry points (for tests)                            */
/* ------------------------------------------------------------------ */
int orc_p2e(const orc_kparams *kp, int local, int order, const double *c, double rscale,
            int ns, const double *src, const double *nrm, const double *str, double *coeffs)
{
    if (kp->kernel == 2) return orc_p2e3(kp, local, order, c, ns, src, nrm, str, coeffs);
    if (kp->kernel == 3) return orc_p2eh3(kp, local, order, c, ns, src, nrm, str, coeffs);
    int nc = ncoeff_of(kp->kernel, order);
    cplx *cf = (cplx *)coeffs;
    for (int i = 0; i < nc; ++i) cf[i] = 0;
    cplx cc = c[0] + JI * c[1];
    for (int s = 0; s < ns; ++s) {
        cplx q = str[2 * s] + JI * str[2 * s + 1];
        cplx nu = nrm ? cpt(nrm, ns, s) : 0;
        p2e_one(kp, local, order, cc, rscale, cpt(src, ns, s), q, nu, cf);
    }
    return 0;
}

int orc_e2p(const orc_kparams *kp, int local, int order, const double *c, double rscale,
            const double *coeffs, int nt, const double *tgt, double *pot)
{
    if (kp->kernel == 2) return orc_e2p3(local, order, c, coeffs, nt, tgt, pot);
    if (kp->kernel == 3) return orc_e2ph3(kp, local, order, c, coeffs, nt, tgt, pot);
    cplx cc = c[0] + JI * c[1];
    for (int t = 0; t < nt; ++t) {
        cplx v = e2p_at(kp, local, order, cc, rscale, (const cplx *)coeffs, cpt(tgt, nt, t),
                        dir_at(kp, nt, t));
        pot[2 * t] = creal(v); pot[2 * t + 1] = cimag(v);
    }
    return 0;
}

int orc_e2e(const orc_kparams *kp, int kind, int p1, int p2,
            const double *c1, double R1, const double *c2, double R2,
            const double *src, double *tgt)
{
    if (kp->kernel == 2) return orc_e2e3(kind, p1, p2, c1, c2, src, tgt);
    if (kp->kernel == 3) return orc_e2eh3(kp, kind, p1, p2, c1, c2, src, tgt);
    int nc = ncoeff_of(kp->kernel, p2);
    cplx *b = (cplx *)tgt;
    for (int i = 0; i < nc; ++i) b[i] = 0;
    e2e_one(kp, kind, p1, p2, c1[0] + JI * c1[1], R1, c2[0] + JI * c2[1], R2,
            (const cplx *)src, b);
    return 0;
}

int orc_p2p(const orc_kparams *kp, int ns, const double *src, const double *nrm,
            const double *str, int nt, const double *tgt, double *pot)
{
    if (kp->kernel == 2) return orc_p2p3(kp, ns, src, nrm, str, nt, tgt, pot);
    if (kp->kernel == 3) return orc_p2ph3(kp, ns, src, nrm, str, nt, tgt, pot);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        cplx acc = 0, tt = cpt(tgt, nt, t);
        for (int s = 0; s < ns; ++s) {
            cplx q = str[2 * s] + JI * str[2 * s + 1];
            cplx nu = nrm ? cpt(nrm, ns, s) : 0;
            acc += p2p_at(kp, cpt(src, ns, s), q, nu, tt, dir_at(kp, nt, t));
        }
        pot[2 * t] = creal(acc); pot[2 * t + 1] = cimag(acc);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* QBX-FMM driver (pytential/qbx/fmm.py:414-594)                       */
/* ------------------------------------------------------------------ */
int orc_qbx_fmm(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                const double *sources, const double *normals, const double *strengths,
                int ncenters, const double *centers, const double *radii,
                const double *targets, const int *uttc, double *result, double *stage_times)
{
    return orc_qbx_fmm_ex(t, tr, kp, sources, normals, strengths, ncenters, centers, radii,
                          targets, uttc, result, stage_times, NULL, NULL, NULL);
}

/* copy per-box coefficients (stored at n + p_level) into pmax-centred rows */
static void dump_coeffs(const orc_kparams *kp, int nrows, int ncf, int pmax,
                        const int *row_order, const cplx *src, double *dst)
{
    for (int b = 0; b < nrows; ++b) {
        int p = row_order[b];
        for (int j = 0; j < ncf; ++j) { dst[2 * ((size_t)b * ncf + j)] = 0; dst[2 * ((size_t)b * ncf + j) + 1] = 0; }
        if (kp->kernel == 0) {
            for (int j = 0; j <= p; ++j) {
                dst[2 * ((size_t)b * ncf + j)] = creal(src[(size_t)b * ncf + j]);
                dst[2 * ((size_t)b * ncf + j) + 1] = cimag(src[(size_t)b * ncf + j]);
            }
        } else {
            for (int n = -p; n <= p; ++n) {
                cplx v = src[(size_t)b * ncf + n + p];
                dst[2 * ((size_t)b * ncf + n + pmax)] = creal(v);
                dst[2 * ((size_t)b * ncf + n + pmax) + 1] = cimag(v);
            }
        }
    }
}

int orc_qbx_fmm_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                   const double *sources, const double *normals, const double *strengths,
                   int ncenters, const double *centers, const double *radii,
                   const double *targets, const int *uttc, double *result, double *stage_times,
                   double *mpole_out, double *local_out, double *qexp_out)
{
    if (t->dim == 3 && kp->kernel == 3)
        return orc_qbx_fmmh3d_ex(t, tr, kp, sources, normals, strengths, ncenters, centers,
                                 radii, targets, uttc, result, stage_times, mpole_out,
                                 local_out, qexp_out);
    if (t->dim == 3)
        return orc_qbx_fmm3d_ex(t, tr, kp, sources, normals, strengths, ncenters, centers,
                                radii, targets, uttc, result, stage_times, mpole_out,
                                local_out, qexp_out);
    if (t->dim != 2 || kp->kernel > 1) return -1;
    if (kp->nthreads > 0) omp_set_num_threads(kp->nthreads);
    binom_init();
    const int ns = t->nsources, nT = t->ntargets, nb = t->nboxes, nl = t->nlevels;
    const int nchild = 4;
    int pmax = 0;
    for (int l = 0; l < nl; ++l) if (kp->level_orders[l] > pmax) pmax = kp->level_orders[l];
    const int ncf = ncoeff_of(kp->kernel, pmax);
    const int pq = kp->qbx_order;
    const int ncq = ncoeff_of(kp->kernel, pq);
    const int dip = kp->source_kind == 1;
    double tm[16] = {0};
    double t0 = omp_get_wtime(), t1;

    /* reorder_sources (fmm.py:436-437) */
    cplx *S = (cplx *)malloc(sizeof(cplx) * (ns ? ns : 1));
    cplx *Q = (cplx *)malloc(sizeof(cplx) * (ns ? ns : 1));
    cplx *NU = (cplx *)malloc(sizeof(cplx) * (ns ? ns : 1));
    for (int i = 0; i < ns; ++i) {
        int u = t->user_source_ids[i];
        S[i] = cpt(sources, ns, u);
        Q[i] = strengths[2 * u] + JI * strengths[2 * u + 1];
        NU[i] = dip ? cpt(normals, ns, u) : 0;
    }
    cplx *T = (cplx *)malloc(sizeof(cplx) * (nT ? nT : 1));
    cplx *DIR = (cplx *)malloc(sizeof(cplx) * (nT ? nT : 1));
    int *tree_ttc = (int *)malloc(sizeof(int) * (nT ? nT : 1));
    for (int i = 0; i < nT; ++i) {
        int u = t->sorted_target_ids[i];
        T[i] = cpt(targets, nT, u);
        DIR[i] = dir_at(kp, nT, u);
        tree_ttc[i] = u < ncenters ? -2 : uttc[u];  /* -2: a center, not evaluated */
    }
    int *tgt_box_of_box = (int *)malloc(sizeof(int) * nb);
    for (int b = 0; b < nb; ++b) tgt_box_of_box[b] = -1;
    for (int i = 0; i < tr->ntarget_boxes; ++i) tgt_box_of_box[tr->target_boxes[i]] = i;
    int *ttp_of_box = (int *)malloc(sizeof(int) * nb);
    for (int b = 0; b < nb; ++b) ttp_of_box[b] = -1;
    for (int i = 0; i < tr->nttp_boxes; ++i) ttp_of_box[tr->ttp_boxes[i]] = i;

    double *rsc = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    for (int l = 0; l < nl; ++l) rsc[l] = ldexp(t->root_extent, -l);
#define BOXC(b) (t->box_centers[b] + JI * t->box_centers[(size_t)nb + (b)])
#define LEV(b) (t->box_levels[b])
#define PL(b) (kp->level_orders[LEV(b)])

    cplx *M = (cplx *)calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *Lc = (cplx *)calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *pot = (cplx *)calloc(nT ? nT : 1, sizeof(cplx));
    t1 = omp_get_wtime(); tm[0] = t1 - t0; t0 = t1;

    /* P2M */
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < tr->nsource_boxes; ++i) {
        int b = tr->source_boxes[i];
        int s0 = t->box_source_starts[b], s1 = s0 + t->box_source_counts_nonchild[b];
        for (int s = s0; s < s1; ++s)
            p2e_one(kp, 0, PL(b), BOXC(b), rsc[LEV(b)], S[s], Q[s], NU[s], M + (size_t)b * ncf);
    }
    t1 = omp_get_wtime(); tm[1] = t1 - t0; t0 = t1;

    /* M2M, upward */
    for (int l = nl - 2; l >= 0; --l) {
        int i0 = tr->level_start_source_parent_box_nrs[l], i1 = tr->level_start_source_parent_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            int P = tr->source_parent_boxes[i];
            for (int c = 0; c < nchild; ++c) {
                int C = t->box_child_ids[(size_t)c * nb + P];
                if (C < 0 || t->box_source_counts_cumul[C] == 0) continue;
                e2e_one(kp, 0, PL(C), PL(P), BOXC(C), rsc[LEV(C)], BOXC(P), rsc[LEV(P)],
                        M + (size_t)C * ncf, M + (size_t)P * ncf);
            }
        }
    }
    t1 = omp_get_wtime(); tm[2] = t1 - t0; t0 = t1;

    /* P2P (list 1) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l1_starts[i]; j < tr->l1_starts[i + 1]; ++j) {
                int Sb = tr->l1_lists[j];
                int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
                for (int s = s0; s < s1; ++s) acc += p2p_at(kp, S[s], Q[s], NU[s], T[ti], DIR[ti]);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[3] = t1 - t0; t0 = t1;

    /* M2L (list 2) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        int P = tr->ttp_boxes[i];
        for (int j = tr->l2_starts[i]; j < tr->l2_starts[i + 1]; ++j) {
            int A = tr->l2_lists[j];
            e2e_one(kp, 1, PL(A), PL(P), BOXC(A), rsc[LEV(A)], BOXC(P), rsc[LEV(P)],
                    M + (size_t)A * ncf, Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[4] = t1 - t0; t0 = t1;

    /* M2P (list 3) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l3_starts[i]; j < tr->l3_starts[i + 1]; ++j) {
                int D = tr->l3_lists[j];
                acc += e2p_at(kp, 0, PL(D), BOXC(D), rsc[LEV(D)], M + (size_t)D * ncf, T[ti],
                              DIR[ti]);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[5] = t1 - t0; t0 = t1;

    /* P2L (list 4) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        int P = tr->ttp_boxes[i];
        for (int j = tr->l4_starts[i]; j < tr->l4_starts[i + 1]; ++j) {
            int Sb = tr->l4_lists[j];
            int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e_one(kp, 1, PL(P), BOXC(P), rsc[LEV(P)], S[s], Q[s], NU[s], Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[6] = t1 - t0; t0 = t1;

    /* L2L, downward */
    for (int l = 1; l < nl; ++l) {
        int i0 = tr->level_start_ttp_box_nrs[l], i1 = tr->level_start_ttp_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            int P = tr->ttp_boxes[i];
            int Pp = t->box_parent_ids[P];
            e2e_one(kp, 2, PL(Pp), PL(P), BOXC(Pp), rsc[LEV(Pp)], BOXC(P), rsc[LEV(P)],
                    Lc + (size_t)Pp * ncf, Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[7] = t1 - t0; t0 = t1;

    /* L2P on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            pot[ti] += e2p_at(kp, 1, PL(B), BOXC(B), rsc[LEV(B)], Lc + (size_t)B * ncf, T[ti],
                              DIR[ti]);
        }
    }
    t1 = omp_get_wtime(); tm[8] = t1 - t0; t0 = t1;

    /* QBX: global centers = centers with >= 1 target */
    char *used = (char *)calloc(ncenters ? ncenters : 1, 1);
    int ntot = nT;
    for (int u = ncenters; u < ntot; ++u) if (uttc[u] >= 0) used[uttc[u]] = 1;
    int *center_tbox = (int *)malloc(sizeof(int) * (ncenters ? ncenters : 1));
    for (int b = 0; b < nb; ++b) {
        int t0i = t->box_target_starts[b], t1i = t0i + t->box_target_counts_nonchild[b];
        for (int ti = t0i; ti < t1i; ++ti) {
            int u = t->sorted_target_ids[ti];
            if (u < ncenters) center_tbox[u] = tgt_box_of_box[b];
        }
    }
    cplx *QX = (cplx *)calloc((size_t)(ncenters ? ncenters : 1) * ncq, sizeof(cplx));
    /* P2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int it = center_tbox[c];
        cplx cc = cpt(centers, ncenters, c);
        for (int j = tr->l1_starts[it]; j < tr->l1_starts[it + 1]; ++j) {
            int Sb = tr->l1_lists[j];
            int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e_one(kp, 1, pq, cc, radii[c], S[s], Q[s], NU[s], QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[9] = t1 - t0; t0 = t1;
    /* M2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int it = center_tbox[c];
        cplx cc = cpt(centers, ncenters, c);
        for (int j = tr->l3_starts[it]; j < tr->l3_starts[it + 1]; ++j) {
            int D = tr->l3_lists[j];
            e2e_one(kp, 1, PL(D), pq, BOXC(D), rsc[LEV(D)], cc, radii[c],
                    M + (size_t)D * ncf, QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[10] = t1 - t0; t0 = t1;
    /* L2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int B = tr->target_boxes[center_tbox[c]];
        cplx cc = cpt(centers, ncenters, c);
        e2e_one(kp, 2, PL(B), pq, BOXC(B), rsc[LEV(B)], cc, radii[c],
                Lc + (size_t)B * ncf, QX + (size_t)c * ncq);
    }
    t1 = omp_get_wtime(); tm[11] = t1 - t0; t0 = t1;
    /* QBXL2P */
#pragma omp parallel for schedule(dynamic, 64)
    for (int ti = 0; ti < nT; ++ti) {
        int c = tree_ttc[ti];
        if (c < 0) continue;
        pot[ti] = e2p_at(kp, 1, pq, cpt(centers, ncenters, c), radii[c], QX + (size_t)c * ncq,
                         T[ti], DIR[ti]);
    }
    t1 = omp_get_wtime(); tm[12] = t1 - t0; t0 = t1;

    if (mpole_out || local_out) {
        int *ro = (int *)malloc(sizeof(int) * (nb ? nb : 1));
        for (int b = 0; b < nb; ++b) ro[b] = PL(b);
        if (mpole_out) dump_coeffs(kp, nb, ncf, pmax, ro, M, mpole_out);
        if (local_out) dump_coeffs(kp, nb, ncf, pmax, ro, Lc, local_out);
        free(ro);
    }
    if (qexp_out) {
        int *ro = (int *)malloc(sizeof(int) * (ncenters ? ncenters : 1));
        for (int c = 0; c < ncenters; ++c) ro[c] = pq;
        dump_coeffs(kp, ncenters, ncq, pq, ro, QX, qexp_out);
        free(ro);
    }
    /* back to user order */
    for (int ti = 0; ti < nT; ++ti) {
        int u = t->sorted_target_ids[ti];
        cplx v = u < ncenters ? 0 : pot[ti];
        result[2 * u] = creal(v); result[2 * u + 1] = cimag(v);
    }
    t1 = omp_get_wtime(); tm[13] = t1 - t0;
    if (stage_times) memcpy(stage_times, tm, sizeof(tm));
#undef BOXC
#undef LEV
#undef PL
    free(S); free(Q); free(NU); free(T); free(DIR); free(tree_ttc); free(tgt_box_of_box); free(ttp_of_box);
    free(rsc); free(M); free(Lc); free(pot); free(used); free(center_tbox); free(QX);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Direct QBX (pytential/qbx/__init__.py:866-1068 with the QBX local    */
/* of the FMM path as the expansion, per the accuracy-oracle role)      */
/* ------------------------------------------------------------------ */
int orc_direct_qbx(int dim, const orc_kparams *kp,
                   int ns, const double *src, const double *nrm, const double *str,
                   int ncenters, const double *centers, const double *radii,
                   int nt, const double *tgt, const int *ttc, double *result)
{
    if (dim == 3) {
        if (kp->nthreads > 0) omp_set_num_threads(kp->nthreads);
        if (kp->kernel == 3)
            return orc_direct_qbxh3d(kp, ns, src, nrm, str, ncenters, centers, nt, tgt, ttc,
                                     result);
        return orc_direct_qbx3d(kp, ns, src, nrm, str, ncenters, centers, nt, tgt, ttc,
                                result);
    }
    if (dim != 2 || kp->kernel > 1) return -1;
    if (kp->nthreads > 0) omp_set_num_threads(kp->nthreads);
    const int pq = kp->qbx_order;
    const int ncq = ncoeff_of(kp->kernel, pq);
    const int dip = kp->source_kind == 1;
#pragma omp parallel for schedule(dynamic, 8)
    for (int it = 0; it < nt; ++it) {
        cplx tt = cpt(tgt, nt, it);
        cplx v = 0;
        int c = ttc[it];
        if (c >= 0) {
            cplx coef[2 * 130 + 1];
            for (int i = 0; i < ncq; ++i) coef[i] = 0;
            cplx cc = cpt(centers, ncenters, c);
            for (int s = 0; s < ns; ++s) {
                cplx q = str[2 * s] + JI * str[2 * s + 1];
                cplx nu = dip ? cpt(nrm, ns, s) : 0;
                p2e_one(kp, 1, pq, cc, radii[c], cpt(src, ns, s), q, nu, coef);
            }
            v = e2p_at(kp, 1, pq, cc, radii[c], coef, tt, dir_at(kp, nt, it));
        } else {
            for (int s = 0; s < ns; ++s) {
                cplx q = str[2 * s] + JI * str[2 * s + 1];
                cplx nu = dip ? cpt(nrm, ns, s) : 0;
                v += p2p_at(kp, cpt(src, ns, s), q, nu, tt, dir_at(kp, nt, it));
            }
        }
        result[2 * it] = creal(v); result[2 * it + 1] = cimag(v);
    }
    return 0;
}

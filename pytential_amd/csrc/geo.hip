// QBX geometry lookups on the GPU (pytential/qbx/geometry.py:643-899):
// center -> target box, global QBX centers, center -> tree targets (CSR,
// targets in ascending tree order), and the per-box lists of targets that need
// no QBX (boxtree FilteredTargetListsInTreeOrder semantics).
#include "qbx_internal.h"

namespace qbx {

__global__ void k_tree_ttc(int nt, const int *sorted_target_ids, const int *uttc, int *tree_ttc,
                           int *user_to_tree)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    int u = sorted_target_ids[i];
    tree_ttc[i] = uttc[u];
    user_to_tree[u] = i;
}

__global__ void k_center_tbox(int ntb, const int *target_boxes, const int *tstarts,
                              const int *tcounts, const int *sorted_target_ids, int ncenters,
                              int *c2tb)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntb) return;
    int b = target_boxes[it];
    for (int i = tstarts[b]; i < tstarts[b] + tcounts[b]; ++i) {
        int u = sorted_target_ids[i];
        if (u < ncenters) c2tb[u] = it;
    }
}

__global__ void k_used_centers(int nt, int ncenters, const int *uttc, int *used)
{
    int u = ncenters + blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nt) return;
    int c = uttc[u];
    if (c >= 0 && c < ncenters) used[c] = 1;
}

__global__ void k_qbx_target_flags(int nt, int ncenters, const int *sorted_target_ids,
                                   const int *tree_ttc, int *fq, int *fn)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    int u = sorted_target_ids[i];
    int c = tree_ttc[i];
    fq[i] = (u >= ncenters && c >= 0) ? 1 : 0;
    fn[i] = (u >= ncenters && c == -1) ? 1 : 0;
}

__global__ void k_keys_of(int n, const int *idx, const int *tree_ttc, int *keys, int *counts)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int c = tree_ttc[idx[j]];
    keys[j] = c;
    atomicAdd(&counts[c], 1);
}

__global__ void k_nqbtl_boxes(int nb, const int *tstarts, const int *tcounts, const int *pos,
                              int *fstarts, int *fcounts)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int s = tstarts[b];
    fstarts[b] = pos[s];
    fcounts[b] = pos[s + tcounts[b]] - pos[s];
}

// per target box: units and lanes per target from its target count (the
// box's non-QBX targets, fcount by box id, or the global QBX centers of the
// box, cstarts by target box index) and its list work: the sources of its
// list boxes (scount given) or the number of list entries (scount null)
__global__ void k_unit_work(int ntb, const int *tboxes, const int *starts, const int *lists,
                            const int *scount, const int *fcount, const int *cstarts, int iter,
                            int *units, int *glanes)
{
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ntb) return;
    const int nf = cstarts ? cstarts[it + 1] - cstarts[it] : fcount[tboxes[it]];
    long work = 0;
    if (scount)
        for (int e = starts[it]; e < starts[it + 1]; ++e) work += scount[lists[e]];
    else
        work = starts[it + 1] - starts[it];
    int g = 1;
    while (g < 64 && 2 * g * nf <= 64) g <<= 1;             // fill the wavefront
    while (g < 64 && (long)g * iter < work) g <<= 1;        // bound a lane's walk
    const int tg = 64 / g;
    units[it] = (nf > 0 && work > 0) ? (nf + tg - 1) / tg : 0;
    glanes[it] = g;
}

int build_unit_plan(hipStream_t st, const qbx_tree *T, const qbx_trav *R, const qbx_geo *G,
                    const int *starts, const int *lists, bool by_sources, bool centers,
                    int iter, unit_plan &plan)
{
    const int ntb = R->ntarget_boxes;
    plan.nunits = 0;
    QBX_TRY(plan.box_g.alloc(std::max(ntb, 1)));
    if (ntb == 0) return QBX_OK;
    dbuf<int> units;
    QBX_TRY(units.alloc(ntb));
    hipLaunchKernelGGL(k_unit_work, dim3(cdiv(ntb, 256)), dim3(256), 0, st, ntb,
                       R->target_boxes.p, starts, lists,
                       by_sources ? T->box_source_counts_nonchild.p : nullptr,
                       G->nqbtl_box_target_counts_nonchild.p,
                       centers ? G->gc_by_tbox_starts.p : nullptr, iter, units.p, plan.box_g.p);
    QBX_HIP(hipGetLastError());
    std::vector<int> hu(ntb), hg(ntb);
    QBX_HIP(hipMemcpyAsync(hu.data(), units.p, sizeof(int) * ntb, hipMemcpyDeviceToHost, st));
    QBX_HIP(hipMemcpyAsync(hg.data(), plan.box_g.p, sizeof(int) * ntb, hipMemcpyDeviceToHost,
                           st));
    QBX_HIP(hipStreamSynchronize(st));
    std::vector<int> ub, ut;
    for (int it = 0; it < ntb; ++it) {
        const int tg = 64 / hg[it];
        for (int u = 0; u < hu[it]; ++u) {
            ub.push_back(it);
            ut.push_back(u * tg);
        }
    }
    plan.nunits = (int)ub.size();
    QBX_TRY(plan.unit_box.alloc(std::max(plan.nunits, 1)));
    QBX_TRY(plan.unit_t0.alloc(std::max(plan.nunits, 1)));
    if (plan.nunits > 0) {
        QBX_HIP(hipMemcpyAsync(plan.unit_box.p, ub.data(), sizeof(int) * plan.nunits,
                               hipMemcpyHostToDevice, st));
        QBX_HIP(hipMemcpyAsync(plan.unit_t0.p, ut.data(), sizeof(int) * plan.nunits,
                               hipMemcpyHostToDevice, st));
    }
    QBX_HIP(hipStreamSynchronize(st));
    return QBX_OK;
}

__global__ void k_gc_keys(int ng, const int *gc, const int *c2tb, int *keys, int *counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ng) return;
    const int it = c2tb[gc[i]];
    keys[i] = it;
    atomicAdd(&counts[it], 1);
}

// global QBX centers grouped by target box (ascending center id within a box)
static int build_gc_by_tbox(qbx_ctx *ctx, hipStream_t st, const qbx_trav *R, qbx_geo *G)
{
    const int ntb = R->ntarget_boxes, ng = G->nglobal;
    QBX_TRY(G->gc_by_tbox_starts.alloc(ntb + 1));
    QBX_TRY(G->gc_by_tbox.alloc(std::max(ng, 1)));
    dbuf<int> keys, keys2, counts;
    QBX_TRY(keys.alloc(std::max(ng, 1)));
    QBX_TRY(keys2.alloc(std::max(ng, 1)));
    QBX_TRY(counts.alloc(std::max(ntb, 1)));
    QBX_HIP(hipMemsetAsync(counts.p, 0, sizeof(int) * std::max(ntb, 1), st));
    if (ng > 0) {
        hipLaunchKernelGGL(k_gc_keys, dim3(cdiv(ng, 256)), dim3(256), 0, st, ng,
                           G->global_qbx_centers.p, G->center_to_target_box.p, keys.p, counts.p);
        QBX_HIP(hipGetLastError());
        int bits = 1;
        while ((1L << bits) < (long)ntb + 1) ++bits;
        QBX_TRY(sort_pairs_i32(ctx, st, keys.p, keys2.p, G->global_qbx_centers.p,
                               G->gc_by_tbox.p, ng, bits));
    }
    if (ntb > 0) QBX_TRY(exclusive_scan(ctx, st, counts.p, G->gc_by_tbox_starts.p, ntb, nullptr));
    QBX_HIP(hipStreamSynchronize(st));
    // groups of up to GC_GROUP consecutive centers of one target box (the
    // rows of the batched P2QBXL contraction)
    std::vector<int> hs(ntb + 1, 0), ub, ut;
    if (ntb > 0)
        QBX_HIP(hipMemcpy(hs.data(), G->gc_by_tbox_starts.p, sizeof(int) * (ntb + 1),
                          hipMemcpyDeviceToHost));
    for (int it = 0; it < ntb; ++it)
        for (int c0 = 0; c0 < hs[it + 1] - hs[it]; c0 += GC_GROUP) {
            ub.push_back(it);
            ut.push_back(c0);
        }
    unit_plan &U = G->gc_group_plan;
    U.nunits = (int)ub.size();
    QBX_TRY(U.unit_box.alloc(std::max(U.nunits, 1)));
    QBX_TRY(U.unit_t0.alloc(std::max(U.nunits, 1)));
    if (U.nunits > 0) {
        QBX_HIP(hipMemcpy(U.unit_box.p, ub.data(), sizeof(int) * U.nunits,
                          hipMemcpyHostToDevice));
        QBX_HIP(hipMemcpy(U.unit_t0.p, ut.data(), sizeof(int) * U.nunits,
                          hipMemcpyHostToDevice));
    }
    // the QBX targets of the centers of each target box, in rows of TSQ_ROWS
    // (the rows of the batched target-specific sum)
    const int nc = G->ncenters;
    std::vector<int> cs(nc + 1, 0), cl(std::max(G->nctt, 1)), gcl(std::max(hs[ntb], 1));
    if (nc > 0)
        QBX_HIP(hipMemcpy(cs.data(), G->ctt_starts.p, sizeof(int) * (nc + 1),
                          hipMemcpyDeviceToHost));
    if (G->nctt > 0)
        QBX_HIP(hipMemcpy(cl.data(), G->ctt_lists.p, sizeof(int) * G->nctt,
                          hipMemcpyDeviceToHost));
    if (hs[ntb] > 0)
        QBX_HIP(hipMemcpy(gcl.data(), G->gc_by_tbox.p, sizeof(int) * hs[ntb],
                          hipMemcpyDeviceToHost));
    std::vector<int> tb, rows;
    for (int it = 0; it < ntb; ++it) {
        int fill = 0;
        for (int g = hs[it]; g < hs[it + 1]; ++g) {
            const int c = gcl[g];
            for (int j = cs[c]; j < cs[c + 1]; ++j) {
                if (fill == 0) {
                    tb.push_back(it);
                    rows.insert(rows.end(), TSQ_ROWS, -1);
                }
                rows[rows.size() - TSQ_ROWS + fill] = cl[j];
                fill = (fill + 1) % TSQ_ROWS;
            }
        }
    }
    unit_plan &V = G->tsq_plan;
    V.nunits = (int)tb.size();
    QBX_TRY(V.unit_box.alloc(std::max(V.nunits, 1)));
    QBX_TRY(G->tsq_rows.alloc(std::max((long)rows.size(), 1L)));
    if (V.nunits > 0) {
        QBX_HIP(hipMemcpy(V.unit_box.p, tb.data(), sizeof(int) * V.nunits,
                          hipMemcpyHostToDevice));
        QBX_HIP(hipMemcpy(G->tsq_rows.p, rows.data(), sizeof(int) * rows.size(),
                          hipMemcpyHostToDevice));
    }
    return QBX_OK;
}

static int build_geo(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                     int nc, const double *centers, const double *radii, const int *uttc,
                     qbx_geo *G)
{
    int nt = T->ntargets, nb = T->nboxes, dim = T->dim;
    QBX_CHECK(nc >= 0 && nc <= nt, "qbx_geo_build: ncenters out of range");
    G->ncenters = nc;
    QBX_TRY(G->centers.alloc((long)dim * std::max(nc, 1)));
    QBX_TRY(G->radii.alloc(std::max(nc, 1)));
    if (nc > 0) {
        QBX_HIP(hipMemcpyAsync(G->centers.p, centers, sizeof(double) * dim * nc,
                               hipMemcpyDeviceToDevice, st));
        QBX_HIP(hipMemcpyAsync(G->radii.p, radii, sizeof(double) * nc, hipMemcpyDeviceToDevice,
                               st));
    }
    QBX_TRY(G->tree_ttc.alloc(std::max(nt, 1)));
    QBX_TRY(G->user_to_tree_target.alloc(std::max(nt, 1)));
    if (nt > 0)
        hipLaunchKernelGGL(k_tree_ttc, dim3(cdiv(nt, 256)), dim3(256), 0, st, nt,
                           T->sorted_target_ids.p, uttc, G->tree_ttc.p,
                           G->user_to_tree_target.p);
    QBX_TRY(G->center_to_target_box.alloc(std::max(nc, 1)));
    QBX_HIP(hipMemsetAsync(G->center_to_target_box.p, 0xff, sizeof(int) * std::max(nc, 1), st));
    if (R->ntarget_boxes > 0)
        hipLaunchKernelGGL(k_center_tbox, dim3(cdiv(R->ntarget_boxes, 64)), dim3(64), 0, st,
                           R->ntarget_boxes, R->target_boxes.p, T->box_target_starts.p,
                           T->box_target_counts_nonchild.p, T->sorted_target_ids.p, nc,
                           G->center_to_target_box.p);
    QBX_HIP(hipGetLastError());

    dbuf<int> used;
    QBX_TRY(used.alloc(std::max(nc, 1)));
    QBX_HIP(hipMemsetAsync(used.p, 0, sizeof(int) * std::max(nc, 1), st));
    if (nt > nc)
        hipLaunchKernelGGL(k_used_centers, dim3(cdiv(nt - nc, 256)), dim3(256), 0, st, nt, nc,
                           uttc, used.p);
    QBX_HIP(hipGetLastError());
    QBX_TRY(compact_indices(ctx, st, used.p, nc, G->global_qbx_centers, &G->nglobal));

    dbuf<int> fq, fn;
    QBX_TRY(fq.alloc(std::max(nt, 1)));
    QBX_TRY(fn.alloc(std::max(nt, 1)));
    if (nt > 0)
        hipLaunchKernelGGL(k_qbx_target_flags, dim3(cdiv(nt, 256)), dim3(256), 0, st, nt, nc,
                           T->sorted_target_ids.p, G->tree_ttc.p, fq.p, fn.p);
    QBX_HIP(hipGetLastError());
    // center -> tree targets
    dbuf<int> filt;
    QBX_TRY(compact_indices(ctx, st, fq.p, nt, filt, &G->nctt));
    QBX_TRY(G->ctt_starts.alloc(nc + 1));
    QBX_TRY(G->ctt_lists.alloc(std::max(G->nctt, 1)));
    {
        dbuf<int> keys, keys2, counts;
        QBX_TRY(keys.alloc(std::max(G->nctt, 1)));
        QBX_TRY(keys2.alloc(std::max(G->nctt, 1)));
        QBX_TRY(counts.alloc(std::max(nc, 1)));
        QBX_HIP(hipMemsetAsync(counts.p, 0, sizeof(int) * std::max(nc, 1), st));
        if (G->nctt > 0) {
            hipLaunchKernelGGL(k_keys_of, dim3(cdiv(G->nctt, 256)), dim3(256), 0, st, G->nctt,
                               filt.p, G->tree_ttc.p, keys.p, counts.p);
            QBX_HIP(hipGetLastError());
            int bits = 1;
            while ((1L << bits) < (long)nc + 1) ++bits;
            QBX_TRY(sort_pairs_i32(ctx, st, keys.p, keys2.p, filt.p, G->ctt_lists.p, G->nctt,
                                   bits));
        }
        QBX_TRY(exclusive_scan(ctx, st, counts.p, G->ctt_starts.p, nc, nullptr));
    }
    // non-QBX filtered target lists
    {
        dbuf<int> pos;
        QBX_TRY(pos.alloc(nt + 1));
        QBX_TRY(exclusive_scan(ctx, st, fn.p, pos.p, nt, &G->nfiltered));
        QBX_TRY(compact_indices(ctx, st, fn.p, nt, G->nqbtl_unfiltered, &G->nfiltered));
        QBX_TRY(G->nqbtl_box_target_starts.alloc(nb));
        QBX_TRY(G->nqbtl_box_target_counts_nonchild.alloc(nb));
        hipLaunchKernelGGL(k_nqbtl_boxes, dim3(cdiv(nb, 256)), dim3(256), 0, st, nb,
                           T->box_target_starts.p, T->box_target_counts_nonchild.p, pos.p,
                           G->nqbtl_box_target_starts.p, G->nqbtl_box_target_counts_nonchild.p);
        QBX_HIP(hipGetLastError());
    }
    G->trav = R;
    QBX_TRY(build_gc_by_tbox(ctx, st, R, G));
    QBX_TRY(build_unit_plan(st, T, R, G, R->l1_starts.p, R->l1_lists.p, true, false,
                            ctx->tune[QBX_TUNE_P2P_UNIT_ITER],
                            G->p2p_plan));
    QBX_TRY(build_unit_plan(st, T, R, G, R->l3_starts.p, R->l3_lists.p, false, false,
                            ctx->tune[QBX_TUNE_M2P_UNIT_ITER], G->m2p_plan));
    QBX_TRY(build_unit_plan(st, T, R, G, R->l1_starts.p, R->l1_lists.p, true, true,
                            ctx->tune[QBX_TUNE_P2QBXL_UNIT_ITER], G->p2q_plan));
    QBX_TRY(build_unit_plan(st, T, R, G, R->l3_starts.p, R->l3_lists.p, false, true,
                            ctx->tune[QBX_TUNE_M2QBXL_UNIT_ITER], G->m2q_plan));
    QBX_HIP(hipStreamSynchronize(st));
    return QBX_OK;
}

struct GFieldRef {
    const void *p;
    long bytes;
};

static GFieldRef geo_field(const qbx_geo *g, const char *f, int nboxes)
{
    std::string s(f);
    auto I = [&](const dbuf<int> &b, long n) { return GFieldRef{b.p, (long)sizeof(int) * n}; };
    if (s == "qbx_center_to_target_box") return I(g->center_to_target_box, g->ncenters);
    if (s == "global_qbx_centers") return I(g->global_qbx_centers, g->nglobal);
    if (s == "center_to_tree_targets_starts") return I(g->ctt_starts, g->ncenters + 1);
    if (s == "center_to_tree_targets_lists") return I(g->ctt_lists, g->nctt);
    if (s == "nqbtl_box_target_starts") return I(g->nqbtl_box_target_starts, nboxes);
    if (s == "nqbtl_box_target_counts_nonchild")
        return I(g->nqbtl_box_target_counts_nonchild, nboxes);
    if (s == "nqbtl_unfiltered_from_filtered_target_indices")
        return I(g->nqbtl_unfiltered, g->nfiltered);
    if (s == "tree_target_to_center") return I(g->tree_ttc, g->tree_ttc.n);
    return GFieldRef{nullptr, -1};
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_geo_build(qbx_ctx *ctx, void *stream, const qbx_tree *tree, const qbx_trav *trav,
                  int ncenters, const double *centers, const double *radii,
                  const int *user_target_to_center, qbx_geo **out)
{
    QBX_CHECK(ctx && tree && trav && out && user_target_to_center, "qbx_geo_build: null argument");
    qbx_geo *G = new qbx_geo();
    int rc = build_geo(ctx, S(stream), tree, trav, ncenters, centers, radii,
                       user_target_to_center, G);
    if (rc != QBX_OK) {
        delete G;
        return rc;
    }
    G->tree_ttc.n = tree->ntargets;
    *out = G;
    return QBX_OK;
}

int qbx_geo_destroy(qbx_geo *g)
{
    delete g;
    return QBX_OK;
}

int qbx_geo_sizes(const qbx_geo *g, int *c)
{
    QBX_CHECK(g && c, "qbx_geo_sizes: null argument");
    c[0] = g->ncenters;
    c[1] = g->nglobal;
    c[2] = g->nctt;
    c[3] = g->nfiltered;
    return QBX_OK;
}

long qbx_geo_field_bytes(const qbx_geo *g, const char *field)
{
    if (!g || !field) return -1;
    return geo_field(g, field, (int)g->nqbtl_box_target_starts.n).bytes;
}

int qbx_geo_export(const qbx_geo *g, const char *field, void *dst, long nbytes, void *stream)
{
    QBX_CHECK(g && field, "qbx_geo_export: null argument");
    GFieldRef f = geo_field(g, field, (int)g->nqbtl_box_target_starts.n);
    QBX_CHECK(f.bytes >= 0, "qbx_geo_export: unknown field '%s'", field);
    QBX_CHECK(nbytes >= f.bytes, "qbx_geo_export: '%s' needs %ld bytes", field, f.bytes);
    if (f.bytes > 0)
        QBX_HIP(hipMemcpyAsync(dst, f.p, f.bytes, hipMemcpyDeviceToDevice, S(stream)));
    return QBX_OK;
}

}  // extern "C"

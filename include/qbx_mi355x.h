/*
 * libqbx_mi355x: MI355X (gfx950) QBX-FMM layer-potential evaluation.
 *
 * Plain C ABI: raw device pointers, sizes and opaque handles; no torch or
 * C++ types cross it.  Every call is ordered on the hipStream_t it is given
 * (passed as void*), returns an int status (0 = ok), and on failure leaves a
 * message in qbx_last_error().  The library never frees caller memory; the
 * tree/traversal/geometry handles own their device arrays until destroyed.
 *
 * Layout conventions (the reference's, pytential/qbx/interactions.py:64-68):
 *   coordinates      SoA double, axis d of point i at x[d*n + i]
 *   indices          int32
 *   complex values   interleaved (re, im) doubles
 *   expansions       box expansions level-compact (qbx_expansion_size);
 *                    QBX expansions [ncenters][ncoeff_qbx] complex, row-major;
 *                    Laplace 2D coefficient k at k (k = 0..p),
 *                    Helmholtz 2D coefficient n at n + p (|n| <= p),
 *                    Laplace 3D coefficient (n, m) at n*n + n + m
 *   right-hand sides several densities share one evaluation (nrhs in the
 *                    kernel description); every per-source, per-coefficient
 *                    and per-target array then carries nrhs values
 *                    innermost: strengths [nsources][nrhs], expansions
 *                    [...][ncoeff][nrhs], potentials [ntargets][nrhs]
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef QBX_MI355X_H
#define QBX_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    QBX_OK = 0,
    QBX_ERR_INVALID = 1,
    QBX_ERR_HIP = 2,
    QBX_ERR_TARGET_ASSOC_FAILED = 3,
    QBX_ERR_UNSUPPORTED = 4
};

enum { QBX_LAPLACE2D = 0, QBX_HELMHOLTZ2D = 1, QBX_LAPLACE3D = 2, QBX_HELMHOLTZ3D = 3 };
enum { QBX_SOURCE_S = 0, QBX_SOURCE_D = 1 };

typedef struct qbx_ctx qbx_ctx;
typedef struct qbx_tree qbx_tree;
typedef struct qbx_trav qbx_trav;
typedef struct qbx_geo qbx_geo;

/* ---- context / errors --------------------------------------------------- */
int qbx_ctx_create(qbx_ctx **ctx, int device);
int qbx_ctx_destroy(qbx_ctx *ctx);
/* A context owns the FMM workspaces, translation tables and scratch that the
 * stage entry points and qbx_drive_fmm use.  Once a HIP stream capture has
 * recorded work that uses them, the context keeps them fixed: a later call
 * that would reallocate or rebuild them fails (QBX_ERR_INVALID) instead of
 * leaving the graph with freed memory or another operator's tables.  Give
 * each captured operator its own context.  qbx_ctx_release_graphs tells the
 * context that no graph captured through it is replayed any more.  (No
 * reference counterpart: the reference has no graph capture.) */
int qbx_ctx_release_graphs(qbx_ctx *ctx);
const char *qbx_last_error(void);
int qbx_version(void);

/* ---- tree build -----------------------------------------------------------
 * Replaces boxtree TreeBuilder.__call__ as invoked by
 * QBXFMMGeometryData.tree() (pytential/qbx/geometry.py:562-618) and by
 * build_tree_with_qbx_metadata (pytential/qbx/utils.py:278-448).
 * sources [dim*nsources], targets [dim*ntargets], target_radii [ntargets] or
 * NULL, weights [n] or NULL (= 1).  extent_norm_linf: 0 = "l2", 1 = "linf". */
int qbx_tree_build(qbx_ctx *ctx, void *stream, int dim,
                   int nsources, const double *sources,
                   int ntargets, const double *targets, const double *target_radii,
                   const int *source_weights, const int *target_weights,
                   int max_leaf_refine_weight, double stick_out_factor,
                   int extent_norm_linf, qbx_tree **out);
int qbx_tree_destroy(qbx_tree *tree);

/* sizes: {dim, nsources, ntargets, nboxes, nlevels}; root: {min_x, min_y, min_z, extent} */
int qbx_tree_sizes(const qbx_tree *tree, int *sizes5, double *root4);
/* Copy one boxtree Tree field (by name, e.g. "box_source_starts",
 * "box_centers", "user_source_ids", "sorted_target_ids", "level_start_box_nrs",
 * "box_child_ids" [2^dim][nboxes], "box_coords" [dim][nboxes], "sources",
 * "targets" (tree order)) into caller device memory `dst` of `nbytes`. */
int qbx_tree_export(const qbx_tree *tree, const char *field, void *dst, long nbytes,
                    void *stream);
long qbx_tree_field_bytes(const qbx_tree *tree, const char *field);

/* ---- traversal ------------------------------------------------------------
 * Replaces boxtree FMMTraversalBuilder + merge_close_lists as invoked by
 * QBXFMMGeometryData.traversal() (pytential/qbx/geometry.py:622-641). */
int qbx_traverse(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                 int well_sep_is_n_away, int targets_have_extent,
                 int from_sep_smaller_min_nsources_cumul, qbx_trav **out);
int qbx_trav_destroy(qbx_trav *trav);
/* counts: {nsource_boxes, nsource_parent_boxes, ntarget_boxes, nttp_boxes,
 *          n_list1, n_list2, n_list3, n_list4} */
int qbx_trav_sizes(const qbx_trav *trav, int *counts8);
/* fields: "source_boxes", "source_parent_boxes", "target_boxes",
 * "target_or_target_parent_boxes", "level_start_*" (same names as boxtree),
 * "neighbor_source_boxes_starts/lists", "from_sep_siblings_starts/lists",
 * "from_sep_smaller_starts/lists", "from_sep_bigger_starts/lists" */
int qbx_trav_export(const qbx_trav *trav, const char *field, void *dst, long nbytes,
                    void *stream);
long qbx_trav_field_bytes(const qbx_trav *trav, const char *field);

/* ---- target association ----------------------------------------------------
 * Replaces associate_targets_to_qbx_centers (pytential/qbx/target_assoc.py:
 * 853-952: mark_targets :534-634, find_centers :637-748).  Inputs on device:
 * stage-1 source nodes + tunnel radii, interleaved centers + radii*(1+tol),
 * targets + side flags.  Outputs (device): target_to_center (-1 = none),
 * target_status (0 unmarked, 1 marked without center, 2 found).
 * *nfailed (host) receives the number of status-1 targets; if it is nonzero
 * the call returns QBX_ERR_TARGET_ASSOC_FAILED (outputs still written). */
int qbx_target_assoc(qbx_ctx *ctx, void *stream, int dim,
                     int nsrc1, const double *src1, const double *tunnel_radii,
                     int ncenters, const double *centers, const double *radii_with_tol,
                     int ntargets, const double *targets, const int *target_flags,
                     int *target_to_center, int *target_status, int *nfailed);
/* Replaces mark_elements_for_refinement (target_assoc.py:751-825):
 * refine_flags[nelements] (device, int) set to 1 for elements owning a source
 * within tunnel distance of a failed (status 1) target. */
int qbx_refine_flags_for_failed(qbx_ctx *ctx, void *stream, int dim,
                                int nsrc1, const double *src1, const double *tunnel_radii,
                                int nelements, const int *element_to_source_starts,
                                int ntargets, const double *targets,
                                const int *target_status, int *refine_flags);

/* ---- QBX geometry lookups ---------------------------------------------------
 * Replaces QBXFMMGeometryData.qbx_center_to_target_box (:643-700),
 * global_qbx_centers (:752-793), center_to_tree_targets (:839-872) and
 * non_qbx_box_target_lists (:874-899).  user_target_to_center [ntargets of
 * the tree] in user order: centers first (-1), then user targets. */
int qbx_geo_build(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                  const qbx_trav *trav, int ncenters, const double *centers,
                  const double *radii, const int *user_target_to_center,
                  qbx_geo **out);
int qbx_geo_destroy(qbx_geo *geo);
/* counts: {ncenters, nglobal_qbx_centers, ncenter_targets, nfiltered_targets} */
int qbx_geo_sizes(const qbx_geo *geo, int *counts4);
/* fields: "qbx_center_to_target_box", "global_qbx_centers",
 * "center_to_tree_targets_starts", "center_to_tree_targets_lists",
 * "nqbtl_box_target_starts", "nqbtl_box_target_counts_nonchild",
 * "nqbtl_unfiltered_from_filtered_target_indices" */
int qbx_geo_export(const qbx_geo *geo, const char *field, void *dst, long nbytes,
                   void *stream);
long qbx_geo_field_bytes(const qbx_geo *geo, const char *field);

/* ---- QBX-FMM wrangler stages ------------------------------------------------
 * kernel description for all stages */
typedef struct {
    int kernel;             /* QBX_LAPLACE2D | QBX_HELMHOLTZ2D | QBX_LAPLACE3D |
                               QBX_HELMHOLTZ3D (real k > 0, spherical-wave expansions) */
    double k;               /* Helmholtz wave number (real) */
    int source_kind;        /* QBX_SOURCE_S | QBX_SOURCE_D */
    int qbx_order;          /* p_q */
    int nlevels;
    const int *level_orders;/* host array [nlevels]: fmm_level_to_order */
    int nrhs;               /* right-hand sides per evaluation (0 is taken as 1) */
    /* Target derivative, sumpy's DirectionalTargetDerivative as pytential's
     * Sp / Dp apply it (symbolic/primitives.py:2303-2465, the target normal
     * as the direction): 0 = the potential; 1 = its derivative along
     * target_dir at each target (2D kernels; the FMM evaluation stages P2P,
     * M2P, L2P, QBXL2P and direct QBX; the expansions are unchanged).
     * target_dir: SoA doubles [2][ntargets] over the tree's targets in tree
     * order (the centers' entries are not read). */
    int target_deriv;
    const double *target_dir;
    /* 3D Laplace in qbx_drive_fmm: 1 = the list-1 part of the QBX targets by
     * the target-specific sum (qbx_eval_target_specific_qbx_locals_nrhs)
     * instead of P2QBXL, as pytential's fmmlib backend does when
     * _use_target_specific_qbx allows it (qbx/fmmlib.py:73-108) */
    int target_specific_qbx;
    /* 3D Laplace, several densities: 1 = the caller states that every
     * strength is real, so the box translations (M2M, M2L, L2L) use the
     * conjugate symmetry of the expansions (half the coefficients, a third
     * of the products) in the stage entry points too; qbx_drive_fmm finds
     * this out on the device by itself.  Not a reference field: pytential's
     * expansions are complex whatever the density. */
    int real_strengths;
} qbx_kernel_desc;

/* Box expansions (multipoles and locals) are level-compact: the boxes of
 * level l are stored contiguously with ncoeff(p_l) complex coefficients each
 * (Laplace p_l + 1, Helmholtz 2 p_l + 1 with coefficient n at n + p_l),
 * level l starting at complex offset level_offsets[l].  Total size in complex
 * entries, or -1 on error: */
long qbx_expansion_size(const qbx_kernel_desc *kd, const qbx_tree *tree);
int qbx_expansion_level_offsets(const qbx_kernel_desc *kd, const qbx_tree *tree,
                                long *level_offsets /* [nlevels + 1] */);
/* complex coefficients per QBX center (Laplace p_q + 1, Helmholtz 2 p_q + 1) */
int qbx_ncoeff_qbx(const qbx_kernel_desc *kd);

/* Source data in tree order.  Replaces QBXExpansionWrangler.reorder_sources
 * (pytential/qbx/fmm.py:198-199) together with get_flat_strengths_from_densities
 * (pytential/qbx/__init__.py:1075-1096): strengths_tree[i][r] =
 * weights[u] * density[u][r], u = user_source_ids[i]; normals gathered likewise
 * (source.py:392-433).  density complex [nsources][nrhs]; weights double
 * [nsources]; normals [dim*nsources] or NULL. Outputs complex / SoA doubles. */
int qbx_reorder_sources(qbx_ctx *ctx, void *stream, const qbx_tree *tree, int nrhs,
                        const double *density, const double *weights,
                        const double *normals, double *strengths_tree,
                        double *normals_tree);

/* Per-panel upsampling of stage-1 nodal values to quad-stage-2 nodes (the
 * meshmode connection behind InterpolationPreprocessor,
 * pytential/symbolic/mappers.py:743-826): out[e*nf + i][r] = sum_j
 * mat[i*n1 + j] * in[e*n1 + j][r], complex in/out with nrhs values innermost,
 * mat double [nf*n1]. */
int qbx_upsample(qbx_ctx *ctx, void *stream, int nelements, int n1, int nf, int nrhs,
                 const double *mat, const double *in, double *out);

/* The stages below mirror the wrangler methods called by drive_fmm
 * (pytential/qbx/fmm.py:436-565).  `accumulate` = 0 overwrites the output.
 *   mpole/local: level-compact, qbx_expansion_size complex entries
 *   qbx_exp:     [ncenters][qbx_ncoeff_qbx] complex
 *   pot:         [ntargets of the tree] complex, tree target order */
int qbx_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                        const qbx_trav *trav, const qbx_kernel_desc *kd,
                        const double *strengths_tree, const double *normals_tree,
                        double *mpole);                         /* P2M, sumpy */
int qbx_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                           const qbx_trav *trav, const qbx_kernel_desc *kd,
                           double *mpole);                      /* M2M */
int qbx_eval_direct(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                    const qbx_trav *trav, const qbx_geo *geo,
                    const qbx_kernel_desc *kd, const double *strengths_tree,
                    const double *normals_tree, double *pot, int accumulate); /* P2P list 1 */
int qbx_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                           const qbx_trav *trav, const qbx_kernel_desc *kd,
                           const double *mpole, double *local); /* M2L list 2 */
int qbx_eval_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                        const qbx_trav *trav, const qbx_geo *geo,
                        const qbx_kernel_desc *kd, const double *mpole,
                        double *pot, int accumulate);           /* M2P list 3 */
int qbx_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                    const qbx_trav *trav, const qbx_kernel_desc *kd,
                    const double *strengths_tree, const double *normals_tree,
                    double *local, int accumulate);             /* P2L list 4 */
int qbx_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                      const qbx_trav *trav, const qbx_kernel_desc *kd,
                      double *local);                           /* L2L */
int qbx_eval_locals(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                    const qbx_trav *trav, const qbx_geo *geo,
                    const qbx_kernel_desc *kd, const double *local,
                    double *pot, int accumulate);               /* L2P */
/* P2QBXL: form_global_qbx_locals (fmm.py:231-265, interactions.py:40-167) */
int qbx_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                               const qbx_trav *trav, const qbx_geo *geo,
                               const qbx_kernel_desc *kd,
                               const double *strengths_tree, const double *normals_tree,
                               double *qbx_exp, int accumulate);
/* M2QBXL: translate_box_multipoles_to_qbx_local (fmm.py:267-312,
 * interactions.py:174-294), all source levels in one call */
int qbx_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream,
                               const qbx_tree *tree, const qbx_trav *trav,
                               const qbx_geo *geo, const qbx_kernel_desc *kd,
                               const double *mpole, double *qbx_exp, int accumulate);
/* L2QBXL: translate_box_local_to_qbx_local (fmm.py:314-356,
 * interactions.py:301-412) */
int qbx_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream,
                               const qbx_tree *tree, const qbx_trav *trav,
                               const qbx_geo *geo, const qbx_kernel_desc *kd,
                               const double *local, double *qbx_exp, int accumulate);
/* QBXL2P: eval_qbx_expansions (fmm.py:358-389, interactions.py:419-532);
 * overwrites pot at every QBX target */
int qbx_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                            const qbx_geo *geo, const qbx_kernel_desc *kd,
                            const double *qbx_exp, double *pot);
/* Reorder (fmm.py:569-588): user-order output of the user targets
 * (tree targets >= ncenters): out[u - ncenters][r] = pot[tree index of u][r]. */
int qbx_finalize_potentials(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                            const qbx_geo *geo, int nrhs, const double *pot,
                            double *out_user);

/* Whole drive_fmm (fmm.py:414-594) in one call: the stages run on `stream`
 * and three side streams owned by ctx, forked from and joined back into
 * `stream` by events (the result is ordered on `stream`; concurrent calls on
 * one ctx race); workspaces owned by ctx.  M2M and L2L skip the coarse levels
 * whose multipoles no stage reads and whose locals are zero, so the
 * potentials are those of the stage-by-stage sequence.  strengths_tree/
 * normals_tree from qbx_reorder_sources; out_user [ntargets - ncenters][nrhs]
 * complex. */
int qbx_drive_fmm(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                  const qbx_trav *trav, const qbx_geo *geo,
                  const qbx_kernel_desc *kd, const double *strengths_tree,
                  const double *normals_tree, double *out_user);

/* qbx_drive_fmm without allocation.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * stage order and results as qbx_drive_fmm (2D Laplace and 2D Helmholtz
 * only; other kernels return QBX_ERR_UNSUPPORTED), with every stage on the
 * caller's stream and its workspace (expansions, potentials, Helmholtz
 * translation tables) in one device array of QBX_NOALLOC_SCRATCH_BYTES per
 * device that the library holds until the process ends; an evaluation that
 * needs more returns QBX_ERR_INVALID.  It calls no hipMalloc or hipFree and
 * creates no stream or event.  All calls on one device share the array and
 * it keeps the last call's contents (each call clears what it needs): calls
 * that overlap on one device race. */
#define QBX_NOALLOC_SCRATCH_BYTES (256L << 20)
int qbx_drive_fmm_noalloc(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                          const qbx_trav *trav, const qbx_geo *geo,
                          const qbx_kernel_desc *kd, const double *strengths_tree,
                          const double *normals_tree, double *out_user);

/* ---- tuning knobs ------------------------------------------------------------
 * No reference counterpart: the work-unit shapes and lane splits that the
 * autotuner (pytential_amd/autotune.py) searches per workload.  Unit plans
 * (lanes per target so that a lane walks about this many list items) take
 * effect at the next qbx_geo_build on the context; the L2QBXL lane split at
 * the next evaluation.  Every setting gives the same results up to the
 * order of the last additions (the L2QBXL split: bit for bit). */
enum {
    QBX_TUNE_P2P_UNIT_ITER = 0,     /* P2P: list-1 sources per lane (16..1024) */
    QBX_TUNE_M2P_UNIT_ITER = 1,     /* M2P: list-3 boxes per lane (1..64) */
    QBX_TUNE_P2QBXL_UNIT_ITER = 2,  /* P2QBXL: list-1 sources per lane (4..512) */
    QBX_TUNE_M2QBXL_UNIT_ITER = 3,  /* M2QBXL: list-3 boxes per lane (1..64) */
    QBX_TUNE_L2QBXL_LANES = 4,      /* 2D Helmholtz one-density L2QBXL: lanes per
                                       center (1, 2, 4, 8) */
    QBX_TUNE_DRIVE_STREAMS = 5,     /* qbx_drive_fmm: 4 = its stage DAG on the caller's
                                       stream and three side streams (default); 1 = every
                                       stage on the caller's stream, so that many contexts
                                       can run side by side on streams of their own (one
                                       graph holding several evaluations) */
    QBX_TUNE_COUNT = 6
};
int qbx_ctx_set_tuning(qbx_ctx *ctx, int param, int value);
int qbx_ctx_get_tuning(qbx_ctx *ctx, int param, int *value);

/* ---- adjoint QBX-FMM (one density) ------------------------------------------
 * No reference counterpart: pytential differentiates only by differencing or
 * through its dense matrix builder (pytential/symbolic/matrix.py:376-522),
 * whose conjugate transpose this applies in O(N).  For the forward map
 * out_user = A strengths_tree of qbx_drive_fmm (2D Laplace or Helmholtz, S or
 * D; 3D Helmholtz S; nrhs 1, no target derivative) it computes
 * str_adj = A^H pot_adj_user
 * through the transposed stages on the same tree, lists and orders (Laplace:
 * the transpose of the real-linear map strengths -> real potentials).
 * pot_adj_user [n_user_targets] complex, str_adj [nsources] complex, tree
 * order (overwritten).  Workspaces are the context's. */
int qbx_drive_fmm_adjoint(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                          const qbx_trav *trav, const qbx_geo *geo,
                          const qbx_kernel_desc *kd, const double *pot_adj_user,
                          const double *normals_tree, double *str_adj);

/* One transposed stage of the above (for stage-by-stage checks): in is the
 * adjoint of the forward stage's output, out that of its input; every stage
 * adds into out except QBXL2P (overwrites the QBX expansions' adjoint) and
 * L2L / M2M (update in in place, out unused).  Arrays as the forward stages
 * (tree-order potentials [ntargets], level-compact expansions, QBX expansions
 * [ncenters][ncoeff_q], tree-order strengths [nsources], all complex). */
enum {
    QBX_ADJ_QBXL2P = 0,  /* potentials -> QBX expansions */
    QBX_ADJ_L2QBXL = 1,  /* QBX expansions -> box locals */
    QBX_ADJ_M2QBXL = 2,  /* QBX expansions -> multipoles */
    QBX_ADJ_P2QBXL = 3,  /* QBX expansions -> strengths */
    QBX_ADJ_L2P = 4,     /* potentials -> box locals */
    QBX_ADJ_M2P = 5,     /* potentials -> multipoles */
    QBX_ADJ_P2P = 6,     /* potentials -> strengths */
    QBX_ADJ_L2L = 7,     /* locals, child -> parent */
    QBX_ADJ_M2L = 8,     /* locals -> multipoles */
    QBX_ADJ_P2L = 9,     /* locals -> strengths */
    QBX_ADJ_M2M = 10,    /* multipoles, parent -> child */
    QBX_ADJ_P2M = 11     /* multipoles -> strengths */
};
int qbx_adjoint_stage(qbx_ctx *ctx, void *stream, const qbx_tree *tree, const qbx_trav *trav,
                      const qbx_geo *geo, const qbx_kernel_desc *kd, int stage,
                      const double *in, double *out, const double *normals_tree);

/* ---- direct QBX (no FMM) ---------------------------------------------------
 * Replaces exec_compute_potential_insn_direct (pytential/qbx/__init__.py:
 * 866-1068) with LayerPotentialOnTargetAndCenterSubset (qbx/direct.py:41-154):
 * the evaluation of fmm_order=False, and the accuracy oracle at full size.
 * Every global QBX center's order-p_q expansion is formed from all sources
 * and evaluated at its targets; every non-QBX target sums all sources.  2D
 * Laplace / Helmholtz (nrhs right-hand sides) and 3D Laplace / Helmholtz (one
 * right-hand side; the QBX targets through the target-specific form, which
 * equals the expansion pipeline); same inputs and output layout as the FMM
 * stages (strengths_tree, normals_tree, pot [ntargets][nrhs] in tree order;
 * finish with qbx_finalize_potentials).  The level orders of kd are unused. */
int qbx_direct_qbx(qbx_ctx *ctx, void *stream, const qbx_tree *tree, const qbx_geo *geo,
                   const qbx_kernel_desc *kd, const double *strengths_tree,
                   const double *normals_tree, double *pot);

/* ---- point-to-point sums over every source ---------------------------------
 * Replaces sumpy's P2P as UnregularizedLayerPotentialSource.
 * exec_compute_potential_insn_direct calls it (pytential/unregularized.py:
 * 154-198: every target sums every source, no tree): pot[t] = sum_s G(t, s)
 * q_s, G the kernel of kd (its source derivative along the normal for
 * QBX_SOURCE_D) in the same units as the FMM stages.  Plain (user) order in
 * and out: sources [dim][nsources], normals [dim][nsources] (double layer
 * only), strengths complex [nsources][nrhs], targets [dim][ntargets], pot
 * complex [ntargets][nrhs] (written, not accumulated).  3D kernels take one
 * right-hand side.  A coincident source and target contribute zero.  kd's
 * qbx_order and level orders are unused (level_orders must still name at
 * least one valid order). */
int qbx_p2p_all(qbx_ctx *ctx, void *stream, const qbx_kernel_desc *kd, int nsources,
                const double *sources, const double *normals, const double *strengths,
                int ntargets, const double *targets, double *pot);

/* ---- geometry of the discretization nodes ----------------------------------
 * Replaces the geometry primitives the QBX source evaluates
 * (pytential/symbolic/primitives.py:1150-1447: parametrization derivative,
 * normal, area element, mapping stretch factor, expansion_radii,
 * expansion_centers, weights_and_area_elements; used by
 * pytential/qbx/geometry.py:466-499 and qbx/__init__.py:1075-1096), from the
 * mesh's nodal coordinates el_nodes [dim][nelements*n1] (element-major) and
 * the element's polynomial interpolant:
 *   interp [m][n1] (NULL: identity, m == n1) -- values at the m output nodes,
 *   deriv [dim-1][m][n1] -- reference derivatives at the output nodes,
 *   ref_weights [m] -- reference quadrature weights (needed for weights),
 *   stretch_map [(dim-1)^2] row-major -- E with stretch = the largest singular
 *     value of J E (curves on [-1, 1]: E = 2; triangles in unit coordinates:
 *     E = 2 x the equilateral-to-unit map),
 *   radius_factor -- expansion radius = factor x stretch (0.5 curves, 0.25
 *     surfaces).
 * Outputs (NULL skips): nodes, normals [dim][nelements*m]; weights
 * (area element x reference weight), radii [nelements*m]; centers
 * [dim][2*nelements*m] interleaved (c-, c+) per node.  Curves: normal =
 * (dy, -dx) / |.|; surfaces: (J_0 x J_1) / |.| (the element's orientation). */
int qbx_element_geometry(qbx_ctx *ctx, void *stream, int dim, int nelements, int n1,
                         const double *el_nodes, int m, const double *interp,
                         const double *deriv, const double *ref_weights,
                         const double *stretch_map, double radius_factor, double *nodes,
                         double *normals, double *weights, double *radii, double *centers);

/* ---- refinement checkers ---------------------------------------------------
 * Replace RefinerWrangler.check_expansion_disks_undisturbed_by_sources and
 * check_sufficient_source_quadrature_resolution (pytential/qbx/refinement.py:
 * 155-281, 363-501), the ball queries of _refine_qbx_stage1 / _stage2
 * (:667-884).  Every ball is tested against every point (the same answer as
 * the reference's area query, without a tree).  Flags are only ever set to 1
 * (clear them first); *found becomes 1 when any flag is set.
 *
 * qbx_check_expansion_disks_undisturbed: condition 1.  nodes [dim][nnodes]
 * (stage 1), element_to_node_starts [nelements+1], centers [dim][2*nnodes]
 * interleaved (c-, c+) per node, danger_radii [2*nnodes] (the expansion radius
 * of the center's node); flags the element of every center with a node at
 * distance <= (1 - tolerance) * danger radius.
 *
 * qbx_check_source_quadrature_resolution: condition 2.  sources
 * [dim][nsources] (stage-2 nodes), element_to_source_starts [nelements+1],
 * danger_radii_by_element [nelements] (0.75 x the element's largest stage-2
 * expansion radius), centers [dim][ncenters] (stage 1); flags the element of
 * every source with a center at distance <= its element's danger radius. */
int qbx_check_expansion_disks_undisturbed(qbx_ctx *ctx, void *stream, int dim, int nelements,
                                          const int *element_to_node_starts, int nnodes,
                                          const double *nodes, const double *centers,
                                          const double *danger_radii, double tolerance,
                                          int *refine_flags, int *found);
int qbx_check_source_quadrature_resolution(qbx_ctx *ctx, void *stream, int dim, int nelements,
                                           const int *element_to_source_starts, int nsources,
                                           const double *sources,
                                           const double *danger_radii_by_element,
                                           int ncenters, const double *centers,
                                           int *refine_flags, int *found);

/* ---- dense layer-potential matrices ----------------------------------------
 * Replaces MatrixBuilder.map_int_g (pytential/symbolic/matrix.py:451-522, the
 * QBX matrix of sumpy's LayerPotentialMatrixGenerator with the quadrature
 * weights and the stage-1 -> quad-stage-2 resampling of map_interpolation,
 * :401-449) and, with centers == NULL, P2PMatrixBuilder (matrix.py:530-600):
 *   mat[t][e*n1 + j] = sum_i K(t, s_{e,i}) weights[e*nf + i] upsample[i][j]
 * K(t, s): centers given -- the order-qbx_order QBX expansion of a unit
 * source at s about target t's center (centers[:, t], radius radii[t]; the
 * center of the side the caller chose), evaluated at t; 3D through the
 * target-specific form, which equals it; centers NULL -- the kernel itself
 * (coincident points zero).  QBX_SOURCE_D: the source-normal derivative.
 * Units as the FMM stages.  sources / normals [dim][nelements*nf]
 * (quad-stage-2 nodes, element-major), weights [nelements*nf] (area element
 * times rule weight), upsample [nf][n1] row-major (NULL: identity, nf == n1,
 * which gives the matrix over the fine nodes), targets / centers [dim]
 * [ntargets], mat complex128 rows of leading dimension ld >= nelements*n1.
 * nf <= 256 and nf*n1 <= 4096.  kd's level orders are unused (2D: they must
 * still name at least one valid order). */
int qbx_lpot_matrix(qbx_ctx *ctx, void *stream, const qbx_kernel_desc *kd, int nelements,
                    int n1, int nf, const double *upsample, const double *sources,
                    const double *normals, const double *weights, int ntargets,
                    const double *targets, const double *centers, const double *radii,
                    long ld, double *mat);

/* ---- target-specific QBX (3D) ---------------------------------------------
 * Replaces eval_target_specific_qbx_locals (pytential/qbx/target_specific/
 * impl.pyx:566-785, called from pytential/qbx/fmmlib.py:587-639): for each
 * (center c, target t) pair of the association and each source s of list 1 of
 * c's target box, the order-p_q expansion about c evaluated directly at t:
 *   QBX_LAPLACE3D   S  q sum_n rho_t^n / rho_s^(n+1) P_n(cos th)
 *   QBX_HELMHOLTZ3D S  i k q sum_n (2n+1) j_n(k rho_t) h_n(k rho_s) P_n(cos th)
 * (fmmlib scaling: 4 pi times the Green's function), op QBX_TSQBX_SP its
 * target gradient, QBX_TSQBX_D its dipole-direction source derivative with
 * strength q.  Tree must be 3D; k = k_re + i k_im; p_q in [0, 12].
 * strengths_tree complex [nsources] (weights * density, tree order);
 * dipvec_tree [3*nsources] (op D only); pot complex [ntargets] (ops S, D) or
 * grad complex [3][ntargets] (op Sp), tree target order.  Entries at QBX
 * targets are overwritten; other entries are left untouched. */
enum { QBX_TSQBX_S = 0, QBX_TSQBX_D = 1, QBX_TSQBX_SP = 2 };
int qbx_eval_target_specific_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *tree,
                                        const qbx_trav *trav, const qbx_geo *geo,
                                        int kernel, int op, double k_re, double k_im,
                                        int qbx_order, const double *strengths_tree,
                                        const double *dipvec_tree, double *pot,
                                        double *grad);
/* The same sum for nrhs densities in one call (3D Laplace, S or D): the
 * kernel value of each (target, source) pair is computed once and applied to
 * every density on the FP64 matrix cores.  strengths_tree [nsources][nrhs]
 * complex, pot [ntargets][nrhs] tree order; accumulate: add into pot at the
 * QBX targets (0: overwrite them).  Output unscaled like the above. */
int qbx_eval_target_specific_qbx_locals_nrhs(qbx_ctx *ctx, void *stream,
                                             const qbx_tree *tree, const qbx_trav *trav,
                                             const qbx_geo *geo, int kernel, int op,
                                             int qbx_order, int nrhs,
                                             const double *strengths_tree,
                                             const double *dipvec_tree, double *pot,
                                             int accumulate);

/* ---- differentiable 3D Helmholtz evaluation at smooth targets --------------
 * Targets far from the surface need no QBX: pytential evaluates them through
 * the point-FMM part of drive_fmm (pytential/qbx/fmm.py:216-225, 436-565) with
 * sumpy's HelmholtzKernel(3), G_k(t, s) = e^{ik|t-s|} / (4 pi |t-s|).  These
 * entry points are the exact sums, with the k-derivative and the adjoint that
 * a torch.autograd.Function backward needs (the reference computes gradients
 * only by differencing or through its dense matrix builder,
 * pytential/symbolic/matrix.py:376-522).
 *   pot[t]     = sum_s G_k(t, s) q_s
 *   dpot_dk[t] = sum_s i e^{ik|t-s|} / (4 pi) q_s        (NULL: skipped)
 *   v[s]       = sum_t conj(G_k(t, s)) u_t               (adjoint)
 * sources/targets SoA double [3][n]; q, pot, dpot_dk, u, v complex; real k.
 * Coincident points contribute zero.  The forward uses the context's scratch
 * for per-chunk partial sums (added in a fixed order). */
int qbx_helmholtz3d_p2p(qbx_ctx *ctx, void *stream, int nsources, const double *sources,
                        const double *strengths, int ntargets, const double *targets, double k,
                        double *pot, double *dpot_dk);
int qbx_helmholtz3d_p2p_adjoint(qbx_ctx *ctx, void *stream, int nsources, const double *sources,
                                int ntargets, const double *targets, double k, const double *u,
                                double *v);
/* The transpose of qbx_upsample: out[e*n1 + j][r] = sum_i mat[i*n1 + j] *
 * in[e*nf + i][r] (the adjoint of the stage-1 -> quad-stage-2 interpolation). */
int qbx_upsample_adjoint(qbx_ctx *ctx, void *stream, int nelements, int n1, int nf, int nrhs,
                         const double *mat, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif

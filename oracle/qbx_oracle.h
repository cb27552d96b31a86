/*
 * CPU oracle for the QBX-FMM hot path.  TEST INFRASTRUCTURE ONLY: it is the
 * checker that the HIP product path is compared against, and the CPU
 * baseline that bench.py times.  The product library never links it.
 *
 * It restates, in plain C (+ OpenMP), the algorithm of
 *   pytential/qbx/fmm.py:414-594        drive_fmm stage order
 *   pytential/qbx/interactions.py       P2QBXL / M2QBXL / L2QBXL / QBXL2P
 *   pytential/qbx/geometry.py:505-899   tree inputs, center/target lookups
 *   pytential/qbx/target_assoc.py       mark_targets / find_centers
 * and of the (un-vendored) boxtree tree build and FMM traversal, whose
 * published semantics (adaptive 2^d-tree with particle stick-out, lists 1-4
 * with well_sep_is_n_away) are restated as an exact integer spec.
 *
 * All arrays are SoA ("sep,C" in loopy terms): coordinate axis d of particle i
 * is at x[d*n + i].  Indices are int32.
 */
#ifndef QBX_ORACLE_H
#define QBX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim, nsources, ntargets, nboxes, nlevels, maxl;
    double root_min[3];
    double root_extent;
    int *box_levels;                 /* [nboxes] */
    int *box_parent_ids;             /* [nboxes], -1 for the root */
    int *box_child_ids;              /* [2^dim * nboxes], child-major, -1 = none */
    int *box_coords;                 /* [dim * nboxes] integer box index at its level */
    double *box_centers;             /* [dim * nboxes] */
    int *box_source_starts, *box_source_counts_nonchild, *box_source_counts_cumul;
    int *box_target_starts, *box_target_counts_nonchild, *box_target_counts_cumul;
    int *level_start_box_nrs;        /* [nlevels + 1] */
    int *user_source_ids;            /* [nsources] tree index -> user index */
    int *sorted_target_ids;          /* [ntargets] tree index -> user index */
} orc_tree;

typedef struct {
    int nlevels;
    int nsource_boxes, nsource_parent_boxes, ntarget_boxes, nttp_boxes;
    int *source_boxes, *level_start_source_box_nrs;
    int *source_parent_boxes, *level_start_source_parent_box_nrs;
    int *target_boxes, *level_start_target_box_nrs;
    int *ttp_boxes, *level_start_ttp_box_nrs;
    int n1, n2, n3, n4;
    int *l1_starts, *l1_lists;       /* by target box index, merged close lists */
    int *l2_starts, *l2_lists;       /* by ttp box index */
    int *l3_starts, *l3_lists;       /* by target box index, all levels */
    int *l4_starts, *l4_lists;       /* by ttp box index */
} orc_trav;

int orc_tree_build(int dim,
                   int nsources, const double *sources,
                   int ntargets, const double *targets, const double *target_radii,
                   const int *source_weights, const int *target_weights,
                   int max_leaf_refine_weight, double stick_out_factor,
                   int extent_norm_linf, orc_tree *out);
void orc_tree_free(orc_tree *t);

int orc_traverse(const orc_tree *t, int well_sep_is_n_away, int targets_have_extent,
                 int from_sep_smaller_min_nsources_cumul, orc_trav *out);
void orc_trav_free(orc_trav *tr);

/* target association: brute force (exact spec) and grid-accelerated */
int orc_target_assoc(int dim, int nsrc1, const double *src1, const double *tunnel_radii,
                     int ncenters, const double *centers, const double *radii_with_tol,
                     int ntargets, const double *targets, const int *target_flags,
                     int accelerated,
                     int *target_to_center, int *target_status);

int orc_refine_flags_for_failed(int dim, int nsrc1, const double *src1,
                                const double *tunnel_radii,
                                int nelements, const int *element_to_source_starts,
                                int ntargets, const double *targets,
                                const int *target_status, int *refine_flags);

/* QBX-FMM and direct QBX (2D Laplace / Helmholtz, S or D source kernel) */
typedef struct {
    int kernel;            /* 0 laplace2d, 1 helmholtz2d, 2 laplace3d, 3 helmholtz3d */
    double k;              /* helmholtz wave number (real) */
    int source_kind;       /* 0 single layer (charge), 1 double layer (dipole) */
    int qbx_order;
    const int *level_orders;  /* [nlevels] */
    int nthreads;          /* 0 = OpenMP default */
    /* 1: evaluate the derivative along target_dir at each target instead of
     * the potential (2D; target_dir SoA [2][n] indexed like the targets array
     * of the call) */
    int target_deriv;
    const double *target_dir;
} orc_kparams;

/* Evaluate the QBX-FMM.  Geometry arrays are in user order; the
 * function reorders into tree order itself (reorder_sources of drive_fmm).
 *   centers/radii: interleaved QBX centers (ncenters = tree targets [0, ncenters))
 *   user_target_to_center: [ntargets_total], -1 no QBX, >=0 center index
 *   strengths: complex [nsources] (user order), normals [dim*nsources] (D only)
 *   result: complex [ntargets_total] in user order (centers get 0)
 * stage_times (optional, [16]) gets per-stage seconds.
 */
int orc_qbx_fmm(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                const double *sources, const double *normals, const double *strengths,
                int ncenters, const double *centers, const double *radii,
                const double *targets, const int *user_target_to_center,
                double *result, double *stage_times);

/* same, also returning multipoles / locals [nboxes][ncoeff(pmax)] and QBX
 * expansions [ncenters][ncoeff(p_q)] (complex, pmax-centred like the GPU) */
int orc_qbx_fmm_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                   const double *sources, const double *normals, const double *strengths,
                   int ncenters, const double *centers, const double *radii,
                   const double *targets, const int *user_target_to_center,
                   double *result, double *stage_times,
                   double *mpole_out, double *local_out, double *qexp_out);

/* Direct QBX oracle (fmm_order=False): every QBX target gets the order-p
 * QBX local of *all* sources at its center; other targets get the full P2P. */
int orc_direct_qbx(int dim, const orc_kparams *kp,
                   int nsources, const double *sources, const double *normals,
                   const double *strengths,
                   int ncenters, const double *centers, const double *radii,
                   int ntargets, const double *targets, const int *target_to_center,
                   double *result);

/* Single-stage entry points used by operator-level tests */
int orc_p2e(const orc_kparams *kp, int local, int order, const double *c, double rscale,
            int nsources, const double *sources, const double *normals,
            const double *strengths, double *coeffs);
int orc_e2p(const orc_kparams *kp, int local, int order, const double *c, double rscale,
            const double *coeffs, int ntargets, const double *targets, double *pot);
int orc_e2e(const orc_kparams *kp, int kind, int src_order, int tgt_order,
            const double *csrc, double rsrc, const double *ctgt, double rtgt,
            const double *src_coeffs, double *tgt_coeffs);
int orc_p2p(const orc_kparams *kp, int nsources, const double *sources,
            const double *normals, const double *strengths,
            int ntargets, const double *targets, double *pot);

/* 3D Laplace (fmm3d_oracle.c): same contracts, coefficient (n, m) at n*n + n + m */
int orc_qbx_fmm3d_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                     const double *sources, const double *normals, const double *strengths,
                     int ncenters, const double *centers, const double *radii,
                     const double *targets, const int *user_target_to_center,
                     double *result, double *stage_times,
                     double *mpole_out, double *local_out, double *qexp_out);
int orc_direct_qbx3d(const orc_kparams *kp, int nsources, const double *sources,
                     const double *normals, const double *strengths, int ncenters,
                     const double *centers, int ntargets, const double *targets,
                     const int *target_to_center, double *result);
int orc_p2e3(const orc_kparams *kp, int local, int order, const double *c, int nsources,
             const double *sources, const double *normals, const double *strengths,
             double *coeffs);
int orc_e2p3(int local, int order, const double *c, const double *coeffs, int ntargets,
             const double *targets, double *pot);
int orc_e2e3(int kind, int src_order, int tgt_order, const double *csrc, const double *ctgt,
             const double *src_coeffs, double *tgt_coeffs);
int orc_p2p3(const orc_kparams *kp, int nsources, const double *sources,
             const double *normals, const double *strengths,
             int ntargets, const double *targets, double *pot);
/* regular / irregular solid harmonics, interleaved complex [(p+1)^2] */
void orc_sh3d(int p, const double *x, double *ups, double *the);

/* 3D Helmholtz (fmmh3d_oracle.c, kernel code 3): same contracts, spherical
 * waves R = j_n Y_n^m (locals), S = h_n Y_n^m (multipoles) */
int orc_qbx_fmmh3d_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                      const double *sources, const double *normals, const double *strengths,
                      int ncenters, const double *centers, const double *radii,
                      const double *targets, const int *user_target_to_center,
                      double *result, double *stage_times,
                      double *mpole_out, double *local_out, double *qexp_out);
int orc_direct_qbxh3d(const orc_kparams *kp, int nsources, const double *sources,
                      const double *normals, const double *strengths, int ncenters,
                      const double *centers, int ntargets, const double *targets,
                      const int *target_to_center, double *result);
int orc_p2eh3(const orc_kparams *kp, int local, int order, const double *c, int nsources,
              const double *sources, const double *normals, const double *strengths,
              double *coeffs);
int orc_e2ph3(const orc_kparams *kp, int local, int order, const double *c,
              const double *coeffs, int ntargets, const double *targets, double *pot);
int orc_e2eh3(const orc_kparams *kp, int kind, int src_order, int tgt_order,
              const double *csrc, const double *ctgt, const double *src_coeffs,
              double *tgt_coeffs);
int orc_p2ph3(const orc_kparams *kp, int nsources, const double *sources,
              const double *normals, const double *strengths,
              int ntargets, const double *targets, double *pot);
/* Gaunt coefficient int Y_n^m conj(Y_nu^mu) conj(Y_l^(m-mu)) */
double orc_gaunt3(int n, int m, int nu, int mu, int l);

/* Bessel helpers (for tests): J_n, Y_n for n = 0..nmax at real x */
void orc_bessel_jy(int nmax, double x, double *j, double *y);

#ifdef __cplusplus
}
#endif
#endif

// Internal definitions shared by the libqbx_mi355x translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qbx_mi355x.h"

namespace qbx {

void set_error(const char *fmt, ...);

#define QBX_HIP(call)                                                            \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::qbx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                             hipGetErrorString(e_));                             \
            return QBX_ERR_HIP;                                                  \
        }                                                                        \
    } while (0)

#define QBX_CHECK(cond, ...)                                                     \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ::qbx::set_error(__VA_ARGS__);                                       \
            return QBX_ERR_INVALID;                                              \
        }                                                                        \
    } while (0)

#define QBX_TRY(call)                                                            \
    do {                                                                         \
        int r_ = (call);                                                         \
        if (r_ != QBX_OK) return r_;                                             \
    } while (0)

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Owned device buffer (freed with the handle that holds it).
template <class T>
struct dbuf {
    T *p = nullptr;
    long n = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(long count) {
        release();
        n = count;
        if (count <= 0) return QBX_OK;
        QBX_HIP(hipMalloc(&p, sizeof(T) * count));
        return QBX_OK;
    }
    long bytes() const { return (long)sizeof(T) * n; }
};

constexpr int MAX_LEVELS = 64;

// Balanced work units for the per-target-box kernels.  A unit is one 64-lane
// wavefront over TG = 64 / G consecutive "targets" of one target box (its
// non-QBX targets for P2P / M2P, its global QBX centers for P2QBXL / M2QBXL),
// G lanes per target splitting the box's interaction list; G is chosen per
// box from its list work (list-1 sources, or list-3 boxes) so that no lane
// walks more than about *_UNIT_ITER items, and so that boxes with few targets
// still fill the wavefront.  A target belongs to exactly one unit, so units
// write their targets' sums without atomics.
constexpr int UNIT_ITER = 128;     // P2P: list-1 sources per lane
constexpr int M2P_UNIT_ITER = 8;   // M2P: list-3 boxes per lane
constexpr int P2Q_UNIT_ITER = 32;  // P2QBXL: list-1 sources per lane
constexpr int M2Q_UNIT_ITER = 2;   // M2QBXL: list-3 boxes per lane
constexpr int GC_GROUP = 4;        // centers per block of the batched 3D P2QBXL
constexpr int TSQ_ROWS = 16;       // QBX targets per block of the batched 3D TSQBX
int tsqbx_nrhs(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
               const qbx_geo *G, int kernel, int op, int qbx_order, int nrhs,
               const double *strengths_tree, const double *dipvec_tree, double scale,
               double *pot, int accumulate);

struct unit_plan {
    int nunits = 0;
    dbuf<int> unit_box;  // [nunits] index into the traversal's target boxes
    dbuf<int> unit_t0;   // [nunits] first non-QBX target of the unit (box-local)
    dbuf<int> box_g;     // [ntarget_boxes] lanes per target (power of two <= 64)
};

}  // namespace qbx

struct qbx_ctx {
    int device = 0;
    // stream-ordered scratch (grown on demand, never shrunk)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // set once a stream capture (HIP graph) has recorded kernels that use this
    // context's scratch, workspaces or tables: from then on those buffers are
    // neither reallocated nor rebuilt (a graph replays the raw pointers and
    // table contents it captured); qbx_ctx_release_graphs clears it
    bool graph_held = false;
    // FMM workspaces for qbx_drive_fmm
    qbx::dbuf<double> ws_mpole, ws_local, ws_local2, ws_qbx, ws_pot;
    // Helmholtz translation tables and the (tree, k, orders) they were built for
    qbx::dbuf<double> ws_tab_m2l, ws_tab_mm;
    std::vector<double> tab_key;
    // set only for the duration of a qbx_drive_fmm_noalloc call: the tables
    // are then built into this memory (inside the device scratch array) on
    // every call instead of into ws_tab_*
    double *tab_m2l_ext = nullptr, *tab_mm_ext = nullptr;
    // this device's allocation-free scratch array (hipGetSymbolAddress at
    // context creation), used only by qbx_drive_fmm_noalloc
    double *noalloc_scratch = nullptr;
    // 3D Laplace unit-offset harmonic tables and the (nsep, pmax) they hold
    qbx::dbuf<double> ws_tab3_m2l, ws_tab3_ch;
    // qbx_drive_fmm (3D Laplace): 1 when every strength of the call is real
    // (the translations then use the conjugate symmetry of the expansions)
    qbx::dbuf<int> ws_allreal;
    // qbx_drive_fmm_adjoint: the adjoints of the multipoles, locals, QBX
    // expansions and tree-order potentials
    qbx::dbuf<double> ws_adj_m, ws_adj_l, ws_adj_q, ws_adj_p;
    // tuning knobs (QBX_TUNE_*): unit-plan list items per lane, L2QBXL lanes
    int tune[QBX_TUNE_COUNT] = {qbx::UNIT_ITER, qbx::M2P_UNIT_ITER, qbx::P2Q_UNIT_ITER,
                                qbx::M2Q_UNIT_ITER, 4, 4};
    std::vector<double> tab3_key;
    // 3D Helmholtz: the Gaunt table (degrees <= h3_pg) and the translation
    // tables of the (k, tree extent, nsep, orders) in h3_key
    qbx::dbuf<double> ws_h3_gaunt, ws_h3_tab;
    // class-grouped one-density M2L: the per-entry products before the
    // per-box sums
    qbx::dbuf<double> ws_m2lg;
    qbx::dbuf<double> ws_h3_m2qb;  // 3D Helmholtz M2QBXL: per-box contraction matrices
    int h3_pg = -1;
    std::vector<double> h3_key;
    // side streams of qbx_drive_fmm (default priority; shared by every
    // context on the device, see side_streams) and its fork/join events
    // (this context's own), set up on its first call
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

struct qbx_tree {
    int dim = 0, nsources = 0, ntargets = 0, nboxes = 0, nlevels = 0, maxl = 0;
    double root_min[3] = {0, 0, 0};
    double root_extent = 1.0;
    std::vector<int> level_start_box_nrs;   // host copy [nlevels + 1]
    qbx::dbuf<int> box_levels, box_parent_ids, box_child_ids, box_coords;
    qbx::dbuf<double> box_centers;
    qbx::dbuf<int> box_source_starts, box_source_counts_nonchild, box_source_counts_cumul;
    qbx::dbuf<int> box_target_starts, box_target_counts_nonchild, box_target_counts_cumul;
    qbx::dbuf<int> d_level_start_box_nrs;
    qbx::dbuf<int> user_source_ids, sorted_target_ids;
    qbx::dbuf<double> sources, targets;      // tree order, SoA
    qbx::dbuf<unsigned long long> box_morton;  // unpadded Morton code at the box's level
};

struct qbx_trav {
    int nlevels = 0;
    int nsep = 2;  // well_sep_is_n_away
    int nsource_boxes = 0, nsource_parent_boxes = 0, ntarget_boxes = 0, nttp_boxes = 0;
    int n1 = 0, n2 = 0, n3 = 0, n4 = 0;
    std::vector<int> h_level_start_source_box_nrs, h_level_start_source_parent_box_nrs,
        h_level_start_target_box_nrs, h_level_start_ttp_box_nrs;
    qbx::dbuf<int> source_boxes, source_parent_boxes, target_boxes, ttp_boxes;
    qbx::dbuf<int> level_start_source_box_nrs, level_start_source_parent_box_nrs,
        level_start_target_box_nrs, level_start_ttp_box_nrs;
    qbx::dbuf<int> l1_starts, l1_lists, l2_starts, l2_lists, l3_starts, l3_lists,
        l4_starts, l4_lists;
    qbx::dbuf<int> box_to_ttp;        // box -> ttp index or -1
    qbx::dbuf<int> box_to_target_box; // box -> target box index or -1
    // coarsest level of a box in list 2 or list 3 (the multipoles anything
    // reads) and of a box with a nonempty list 2 or list 4 (the coarsest
    // nonzero local); nlevels when there is none
    int mpole_min_level = 0, local_min_level = 0;
    // class-grouped one-density M2L (2D Helmholtz; built on first use): the
    // list-2 entries of the grouped levels sorted by (level, offset class),
    // each entry's sorted position (-1 for the other levels), and work items
    // of up to 16 entries of one (level, class): {level, class, first, count},
    // each entry's row offset in the scratch of products
    mutable qbx::dbuf<int> m2lg_e, m2lg_pos, m2lg_yoff, m2lg_items;
    // items of the one-wave kernel (ncl <= 48), then of the four-wave one;
    // the pairs' rows (complex) in all
    mutable int m2lg_nitems = -1, m2lg_nitems2 = 0;
    mutable long m2lg_nrows = 0;
    mutable std::vector<int> m2lg_levels;  // the grouped levels it was built for
};

struct qbx_geo {
    int ncenters = 0, nglobal = 0, nctt = 0, nfiltered = 0;
    qbx::dbuf<double> centers, radii;          // copies, SoA [dim][ncenters]
    qbx::dbuf<int> tree_ttc;                   // [ntargets] tree order
    qbx::dbuf<int> center_to_target_box;       // [ncenters]
    qbx::dbuf<int> global_qbx_centers;         // [nglobal]
    qbx::dbuf<int> ctt_starts, ctt_lists;      // [ncenters + 1], [nctt]
    qbx::dbuf<int> nqbtl_box_target_starts, nqbtl_box_target_counts_nonchild;
    qbx::dbuf<int> nqbtl_unfiltered;           // [nfiltered] tree target indices
    qbx::dbuf<int> user_to_tree_target;        // [ntargets]
    // global QBX centers grouped by target box: CSR over target box indices
    qbx::dbuf<int> gc_by_tbox_starts, gc_by_tbox;
    // work units over the traversal the geometry was built with: P2P (list 1)
    // and M2P (list 3) over non-QBX targets, P2QBXL and M2QBXL over centers
    const qbx_trav *trav = nullptr;
    qbx::unit_plan p2p_plan, m2p_plan, p2q_plan, m2q_plan;
    // groups of qbx::GC_GROUP consecutive global centers of one target box
    qbx::unit_plan gc_group_plan;
    // the QBX targets grouped for the many-density target-specific sum:
    // unit u covers up to qbx::TSQ_ROWS (center, target) entries whose
    // centers lie in one target box (unit_box[u]); tsq_rows[TSQ_ROWS u + j]
    // is the tree target index of row j, or -1
    qbx::unit_plan tsq_plan;
    qbx::dbuf<int> tsq_rows;
};

namespace qbx {
int scratch_reserve(qbx_ctx *ctx, hipStream_t st, size_t bytes, void **p);
// the three side streams of the device's fused drivers, created once per
// device and kept until the process ends: contexts share them, so a process
// with many contexts (one per captured operator) does not spread its
// streams over the hardware queues the main stream uses
int side_streams(int device, hipStream_t out[3]);
// the device address of this device's allocation-free scratch array (looked
// up once per device, with the device current)
int noalloc_scratch_address(int device, double **out);
// marks the context as held by a graph when st is capturing
int note_capture(qbx_ctx *ctx, hipStream_t st);
// QBX_OK when the context's buffers may be replaced (no graph holds them),
// else an error naming `what`; call before every reallocation or rebuild
int may_replace(qbx_ctx *ctx, hipStream_t st, const char *what);
// grows a context workspace to at least n elements (never shrinks it)
template <class T>
int ws_reserve(qbx_ctx *ctx, hipStream_t st, dbuf<T> &b, long n, const char *what)
{
    QBX_TRY(note_capture(ctx, st));
    if (b.n >= n) return QBX_OK;
    QBX_TRY(may_replace(ctx, st, what));
    return b.alloc(n);
}
// exclusive scan of int flags/counts (n items) into out (n+1 items), returns total via host copy
int exclusive_scan(qbx_ctx *ctx, hipStream_t st, const int *in, int *out, int n, int *total);
// stable sort pairs by 64-bit or 32-bit key
int sort_pairs_u64(qbx_ctx *ctx, hipStream_t st, const unsigned long long *kin,
                   unsigned long long *kout, const int *vin, int *vout, int n, int end_bit);
int sort_pairs_i32(qbx_ctx *ctx, hipStream_t st, const int *kin, int *kout, const int *vin,
                   int *vout, int n, int end_bit);
int compact_indices(qbx_ctx *ctx, hipStream_t st, const int *flags, int n, dbuf<int> &out,
                    int *count);
// stream-ordered zero fill of n doubles by a kernel (safe inside graph capture)
int zero_fill(double *p, long ndoubles, hipStream_t st);
// dst[i] += src[i] for n doubles, stream-ordered
int add_into(double *dst, const double *src, long ndoubles, hipStream_t st);
// work units of a target-box list (starts/lists over the traversal's target
// boxes) for the geometry's non-QBX targets, or its global QBX centers
// (centers); work per entry is the entry box's source count (by_sources) or one
int build_unit_plan(hipStream_t st, const qbx_tree *T, const qbx_trav *R, const qbx_geo *G,
                    const int *starts, const int *lists, bool by_sources, bool centers,
                    int iter, unit_plan &plan);

// direct 3D QBX (tsqbx.hip): target-specific sums over all sources at QBX
// targets, direct sums elsewhere; pot complex [ntargets] tree order, scaled
// dense layer-potential matrices (fmm2d.hip, tsqbx.hip): one block of
// LM_THREADS per target row; the upsampling matrix (nf x n1 doubles, at most
// LM_UMAX) is staged in LDS; nf <= LM_THREADS.  Contract: qbx_lpot_matrix.
constexpr int LM_THREADS = 256;
constexpr int LM_UMAX = 4096;
int lpot_matrix_2d(hipStream_t st, const qbx_kernel_desc *d, int nel, int n1, int nf,
                   const double *U, const double *src, const double *nrm, const double *w,
                   int nt, const double *tgt, const double *ctr, const double *rad, long ld,
                   double *mat);
int lpot_matrix_3d(hipStream_t st, const qbx_kernel_desc *d, int nel, int n1, int nf,
                   const double *U, const double *src, const double *nrm, const double *w,
                   int nt, const double *tgt, const double *ctr, const double *rad, long ld,
                   double *mat);

int direct_p2p3d(hipStream_t st, double k, bool dip, int nf, const int *unfilt,
                 const double *tgt, int nt, const double *src, int ns, const double *str,
                 const double *nrm, double *pot);
int direct_qbx3d(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_geo *G, int kernel,
                 int kind, double k, int qbx_order, const double *str, const double *nrm,
                 double *pot);

// 3D Laplace stages (fmm3d.hip), same contracts as the C-ABI entry points
long f3_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T);
int f3_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets);
int f3_ncoeff_qbx(const qbx_kernel_desc *d);
int f3_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                       const qbx_kernel_desc *d, const double *str, const double *nrm,
                       double *mpole);
// M2M into parent levels >= lmin, L2L from parent levels >= lmin (0: all)
int f3_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, double *mpole, int lmin = 0,
                          const int *allreal = nullptr);
int f3_eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, int mode, const double *str,
                    const double *nrm, const double *expn, double *pot, int accumulate);
// allreal: null, or a device flag that is 1 when every strength behind the
// multipoles is real (then the conjugate-symmetric matrix-core kernel runs)
int f3_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, const double *mpole, double *local,
                          const int *allreal = nullptr);
// *flag = 1 if every imaginary part of the n complex values z is zero, else 0
int strengths_all_real(const double *z, long n, int *flag, hipStream_t st);
int f3_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                   const qbx_kernel_desc *d, const double *str, const double *nrm, double *local,
                   int accumulate);
int f3_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                     const qbx_kernel_desc *d, double *local, int lmin = 0,
                     const int *allreal = nullptr);
int f3_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                              const qbx_trav *R, const qbx_geo *G, const qbx_kernel_desc *d,
                              const double *str, const double *nrm, double *qexp, int accumulate);
int f3_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                             const qbx_trav *R, const qbx_geo *G,
                                             const qbx_kernel_desc *d, const double *mpole,
                                             double *qexp, int accumulate);
int f3_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                        const qbx_trav *R, const qbx_geo *G,
                                        const qbx_kernel_desc *d, const double *local,
                                        double *qexp, int accumulate);
int f3_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                           const qbx_kernel_desc *d, const double *qexp, double *pot);

// 3D Helmholtz stages (fmmh3d.hip): the f3_* contracts for QBX_HELMHOLTZ3D
long h3_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T);
int h3_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets);
int h3_ncoeff_qbx(const qbx_kernel_desc *d);
int h3_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                       const qbx_kernel_desc *d, const double *str, const double *nrm,
                       double *mpole);
int h3_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, double *mpole, int lmin);
int h3_eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, int mode, const double *str,
                    const double *nrm, const double *expn, double *pot, int accumulate);
int h3_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, const double *mpole, double *local);
int h3_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                   const qbx_kernel_desc *d, const double *str, const double *nrm, double *local,
                   int accumulate);
int h3_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                     const qbx_kernel_desc *d, double *local, int lmin);
int h3_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                              const qbx_trav *R, const qbx_geo *G, const qbx_kernel_desc *d,
                              const double *str, const double *nrm, double *qexp, int accumulate);
int h3_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                             const qbx_trav *R, const qbx_geo *G,
                                             const qbx_kernel_desc *d, const double *mpole,
                                             double *qexp, int accumulate);
int h3_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                        const qbx_trav *R, const qbx_geo *G,
                                        const qbx_kernel_desc *d, const double *local,
                                        double *qexp, int accumulate);
int h3_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                           const qbx_kernel_desc *d, const double *qexp, double *pot);
// the conjugate-transposed stages and driver (single layer, one density);
// arguments as qbx_adjoint_stage / qbx_drive_fmm_adjoint
int h3_adjoint_stage(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                     const qbx_geo *G, const qbx_kernel_desc *d, int stage, const double *in,
                     double *out);
int h3_drive_fmm_adjoint(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                         const qbx_geo *G, const qbx_kernel_desc *d, const double *pot_adj_user,
                         double *str_adj);

}  // namespace qbx

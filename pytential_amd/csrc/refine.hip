// QBX refinement checkers (pytential/qbx/refinement.py:155-281, driven by
// _refine_qbx_stage1 / _refine_qbx_stage2, :667-884):
//
//   condition 1 (expansion disk undisturbed by sources): a stage-1 center c
//     whose ball of radius (1 - tol) * danger_radius(c) contains a stage-1
//     source flags the element of c;
//   condition 2 (sufficient source quadrature resolution): a stage-2 source
//     whose ball of radius danger_radius(element of the source) contains a
//     center flags the element of the source.
//
// The reference finds the sources (centers) near each ball by an area query
// on a tree; here every ball is tested against every point, tiled through
// LDS: the same flags (a point is near or not), with no tree to build, and at
// the sizes refinement runs on (1e4..1e6 balls) a sweep of a few milliseconds.
// Flags are set to 1 by plain stores (every writer writes the same value).
#include "qbx_internal.h"

namespace qbx {

constexpr int RF_THREADS = 256;

// element containing point i of an element -> point CSR (starts[nel + 1])
__device__ __forceinline__ int element_of(const int *starts, int nel, int i)
{
    int lo = 0, hi = nel;  // starts[lo] <= i < starts[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (starts[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One thread per ball (center or point), the points through LDS.  ball_elem:
// the element flagged for ball b; radius: rad[b] * scale, or, with
// rad_by_element, rad[element of b] * scale.
template <int DIM>
__global__ void __launch_bounds__(RF_THREADS)
k_ball_contains_any(int nballs, const double *balls, const double *rad, int rad_by_element,
                    double scale, const int *el_starts, int nel, int ball_stride_div,
                    int npts, const double *pts, int *flags, int *found)
{
    __shared__ double sp[DIM][RF_THREADS];
    const int b = blockIdx.x * RF_THREADS + threadIdx.x;
    const bool act = b < nballs;
    double c[DIM];
    double r = -1.0;
    int e = 0;
    if (act) {
        for (int a = 0; a < DIM; ++a) c[a] = balls[(size_t)a * nballs + b];
        // balls are centers (two per node: ball_stride_div 2) or points (1)
        e = element_of(el_starts, nel, b / ball_stride_div);
        r = scale * (rad_by_element ? rad[e] : rad[b]);
    } else {
        for (int a = 0; a < DIM; ++a) c[a] = 0.0;
    }
    bool hit = false;
    for (int t0 = 0; t0 < npts; t0 += RF_THREADS) {
        const int n = min(RF_THREADS, npts - t0);
        __syncthreads();
        if (threadIdx.x < n)
            for (int a = 0; a < DIM; ++a) sp[a][threadIdx.x] = pts[(size_t)a * npts + t0 + threadIdx.x];
        __syncthreads();
        if (act && !hit)
            for (int j = 0; j < n; ++j) {
                double d2 = 0.0;
                for (int a = 0; a < DIM; ++a) {
                    const double d = c[a] - sp[a][j];
                    d2 += d * d;
                }
                if (sqrt(d2) <= r) {  // the reference's distance(...) <= radius
                    hit = true;
                    break;
                }
            }
    }
    if (hit) {
        flags[e] = 1;
        *found = 1;
    }
}

static int ball_sweep(hipStream_t st, int dim, int nballs, const double *balls,
                      const double *rad, int rad_by_element, double scale, const int *el_starts,
                      int nel, int stride_div, int npts, const double *pts, int *flags,
                      int *found)
{
    if (nballs == 0 || npts == 0) return QBX_OK;
    const dim3 g(cdiv(nballs, RF_THREADS)), blk(RF_THREADS);
    if (dim == 2)
        hipLaunchKernelGGL(k_ball_contains_any<2>, g, blk, 0, st, nballs, balls, rad,
                           rad_by_element, scale, el_starts, nel, stride_div, npts, pts, flags,
                           found);
    else
        hipLaunchKernelGGL(k_ball_contains_any<3>, g, blk, 0, st, nballs, balls, rad,
                           rad_by_element, scale, el_starts, nel, stride_div, npts, pts, flags,
                           found);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_check_expansion_disks_undisturbed(qbx_ctx *ctx, void *stream, int dim, int nelements,
                                          const int *element_to_node_starts, int nnodes,
                                          const double *nodes, const double *centers,
                                          const double *danger_radii, double tolerance,
                                          int *refine_flags, int *found)
{
    QBX_CHECK(ctx && element_to_node_starts && refine_flags && found,
              "qbx_check_expansion_disks_undisturbed: null argument");
    QBX_CHECK(dim == 2 || dim == 3, "dim must be 2 or 3");
    QBX_CHECK(nelements >= 0 && nnodes >= 0, "negative size");
    QBX_CHECK(nnodes == 0 || (nodes && centers && danger_radii), "null geometry");
    QBX_CHECK(tolerance >= 0.0 && tolerance < 1.0, "tolerance must be in [0, 1)");
    return ball_sweep(S(stream), dim, 2 * nnodes, centers, danger_radii, 0, 1.0 - tolerance,
                      element_to_node_starts, nelements, 2, nnodes, nodes, refine_flags, found);
}

int qbx_check_source_quadrature_resolution(qbx_ctx *ctx, void *stream, int dim, int nelements,
                                           const int *element_to_source_starts, int nsources,
                                           const double *sources,
                                           const double *danger_radii_by_element,
                                           int ncenters, const double *centers,
                                           int *refine_flags, int *found)
{
    QBX_CHECK(ctx && element_to_source_starts && refine_flags && found,
              "qbx_check_source_quadrature_resolution: null argument");
    QBX_CHECK(dim == 2 || dim == 3, "dim must be 2 or 3");
    QBX_CHECK(nelements >= 0 && nsources >= 0 && ncenters >= 0, "negative size");
    QBX_CHECK(nsources == 0 || (sources && danger_radii_by_element), "null sources");
    QBX_CHECK(ncenters == 0 || centers, "null centers");
    return ball_sweep(S(stream), dim, nsources, sources, danger_radii_by_element, 1, 1.0,
                      element_to_source_starts, nelements, 1, ncenters, centers, refine_flags,
                      found);
}

}  // extern "C"

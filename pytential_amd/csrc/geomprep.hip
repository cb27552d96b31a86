// Geometry of the discretization nodes from the mesh's nodal coordinates:
// the device counterpart of pytential's geometry primitives
// (pytential/symbolic/primitives.py:1150-1447 -- parametrization derivative,
// normal, area element, mapping stretch factor, expansion radii and centers,
// weights_and_area_elements) as the QBX source evaluates them
// (pytential/qbx/geometry.py:466-499, pytential/qbx/__init__.py:1075-1096).
//
// Every quantity comes from the element's polynomial interpolant of its
// nodal coordinates, like meshmode's num_reference_derivative:
//   x  = sum_j I[i][j] x_j          (I = identity at the stage-1 nodes, the
//                                    stage-1 -> quad-stage-2 interpolation
//                                    matrix at the fine nodes)
//   J  = sum_j D_k[i][j] x_j        (dim x rd, rd = dim - 1 reference axes)
//   area = |J| (curve) or |J_0 x J_1| (surface);  normal = J^perp / area
//   weight = area * reference weight
//   stretch = largest singular value of J E (E: the reference-to-equilateral
//             scaling of primitives.py:1150-1213, times 2 for the bi-unit side)
//   radius = factor * stretch,  centers = x -/+ radius * normal (interleaved)
// One thread per output node; memory-bound and small (run once per geometry).
#include "qbx_internal.h"

namespace qbx {

__global__ void k_element_geometry(int dim, int nel, int n1, const double *xn, int m,
                                   const double *interp, const double *deriv,
                                   const double *refw, const double *E, double factor,
                                   double *x_out, double *n_out, double *w_out,
                                   double *rad_out, double *ctr_out)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nel * m) return;
    const int e = (int)(g / m), i = (int)(g % m);
    const long N1 = (long)nel * n1, NM = (long)nel * m;
    const int rd = dim - 1;
    double x[3] = {0, 0, 0}, J0[3] = {0, 0, 0}, J1[3] = {0, 0, 0};
    for (int j = 0; j < n1; ++j) {
        const double c = interp ? interp[(long)i * n1 + j] : (i == j ? 1.0 : 0.0);
        const double d0 = deriv[(long)i * n1 + j];
        const double d1 = rd > 1 ? deriv[((long)m + i) * n1 + j] : 0.0;
        for (int a = 0; a < dim; ++a) {
            const double v = xn[a * N1 + (long)e * n1 + j];
            x[a] += c * v;
            J0[a] += d0 * v;
            J1[a] += d1 * v;
        }
    }
    double nv[3] = {0, 0, 0}, area, stretch;
    if (dim == 2) {
        area = sqrt(J0[0] * J0[0] + J0[1] * J0[1]);
        nv[0] = J0[1] / area;
        nv[1] = -J0[0] / area;
        const double s = E[0];
        stretch = fabs(s) * area;
    } else {
        const double cr[3] = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
                              J0[0] * J1[1] - J0[1] * J1[0]};
        area = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int a = 0; a < 3; ++a) nv[a] = cr[a] / area;
        // Je = J E (3 x 2), G = Je^T Je, largest eigenvalue of the 2 x 2 G
        double A0[3], A1[3];
        for (int a = 0; a < 3; ++a) {
            A0[a] = J0[a] * E[0] + J1[a] * E[2];
            A1[a] = J0[a] * E[1] + J1[a] * E[3];
        }
        const double g00 = A0[0] * A0[0] + A0[1] * A0[1] + A0[2] * A0[2];
        const double g11 = A1[0] * A1[0] + A1[1] * A1[1] + A1[2] * A1[2];
        const double g01 = A0[0] * A1[0] + A0[1] * A1[1] + A0[2] * A1[2];
        const double h = 0.5 * (g00 - g11);
        stretch = sqrt(0.5 * (g00 + g11) + sqrt(h * h + g01 * g01));
    }
    const double r = factor * stretch;
    for (int a = 0; a < dim; ++a) {
        x_out[a * NM + g] = x[a];
        n_out[a * NM + g] = nv[a];
        if (ctr_out) {
            ctr_out[a * 2 * NM + 2 * g] = x[a] - r * nv[a];
            ctr_out[a * 2 * NM + 2 * g + 1] = x[a] + r * nv[a];
        }
    }
    if (w_out) w_out[g] = area * refw[i];
    if (rad_out) rad_out[g] = r;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_element_geometry(qbx_ctx *ctx, void *stream, int dim, int nelements, int n1,
                         const double *el_nodes, int m, const double *interp,
                         const double *deriv, const double *ref_weights,
                         const double *stretch_map, double radius_factor, double *nodes,
                         double *normals, double *weights, double *radii, double *centers)
{
    QBX_CHECK(ctx && el_nodes && deriv && stretch_map && nodes && normals,
              "qbx_element_geometry: null argument");
    QBX_CHECK(dim == 2 || dim == 3, "qbx_element_geometry: dim must be 2 or 3");
    QBX_CHECK(nelements >= 0 && n1 >= 1 && m >= 1, "qbx_element_geometry: bad size");
    QBX_CHECK(interp || m == n1, "qbx_element_geometry: no interpolation needs m == n1");
    QBX_CHECK(!weights || ref_weights, "qbx_element_geometry: weights need ref_weights");
    const long n = (long)nelements * m;
    if (n == 0) return QBX_OK;
    hipLaunchKernelGGL(k_element_geometry, dim3(cdiv(n, 256)), dim3(256), 0, S(stream), dim,
                       nelements, n1, el_nodes, m, interp, deriv, ref_weights, stretch_map,
                       radius_factor, nodes, normals, weights, radii, centers);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // extern "C"

// 3D Helmholtz point sums for smooth (far) targets and their adjoints: the
// forward and backward kernels of the differentiable single-layer evaluation
// at targets that need no QBX (pytential evaluates such targets through the
// plain FMM part of the QBX-FMM, qbx/fmm.py:216-225, P2P/M2P/L2P with
// HelmholtzKernel(3) = e^{ikr} / (4 pi r)).
//
//   pot[t]    = sum_s G_k(t, s) q_s,          G_k = e^{ik|t-s|} / (4 pi |t-s|)
//   dpot_dk[t] = sum_s dG_k/dk q_s = sum_s i e^{ik|t-s|} / (4 pi) q_s
//   v[s]      = sum_t conj(G_k(t, s)) u_t     (the adjoint, G^H u)
//
// Forward: few targets (thousands) against many sources (10^5-10^6), so the
// sources are split into chunks across blocks: block (target tile, source
// chunk) sums its chunk for 256 targets through LDS tiles of sources and
// writes a partial; a second kernel adds the partials of each target in chunk
// order (deterministic, no atomics).  Adjoint: one thread per source, the
// targets staged through LDS.  Coincident points contribute zero.
#include "qbx_internal.h"
#include "qbx_math.h"

namespace qbx {

constexpr int H3_THREADS = 256;   // targets (forward) / sources (adjoint) per block
constexpr int H3_TILE = 256;      // staged points per LDS tile
constexpr double INV4PI = 0.07957747154594766788;

__global__ void __launch_bounds__(H3_THREADS)
k_h3_p2p_partial(int ns, const double *src, const double *q, int nt, const double *tgt,
                 double k, int chunk, double *part_pot, double *part_dk)
{
    __shared__ double sx[H3_TILE], sy[H3_TILE], sz[H3_TILE], sqr[H3_TILE], sqi[H3_TILE];
    const int t = blockIdx.x * H3_THREADS + threadIdx.x;
    const int c = blockIdx.y;
    const int s0 = c * chunk, s1 = min(ns, s0 + chunk);
    const bool act = t < nt;
    const double tx = act ? tgt[t] : 0.0, ty = act ? tgt[(size_t)nt + t] : 0.0,
                 tz = act ? tgt[2 * (size_t)nt + t] : 0.0;
    double pr = 0, pi = 0, dr = 0, di = 0;
    for (int b = s0; b < s1; b += H3_TILE) {
        const int nb = min(H3_TILE, s1 - b);
        __syncthreads();
        for (int j = threadIdx.x; j < nb; j += H3_THREADS) {
            const int s = b + j;
            sx[j] = src[s];
            sy[j] = src[(size_t)ns + s];
            sz[j] = src[2 * (size_t)ns + s];
            sqr[j] = q[2 * (size_t)s];
            sqi[j] = q[2 * (size_t)s + 1];
        }
        __syncthreads();
        if (act)
            for (int j = 0; j < nb; ++j) {
                const double dx = tx - sx[j], dy = ty - sy[j], dz = tz - sz[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 == 0.0) continue;
                // 1/r and r from the refined hardware estimate; sin and cos
                // by qbx_sincos (a few ulp, no library call)
                const double ir = qbx_rsqrt(r2);
                const double r = r2 * ir;
                double sn, cs;
                qbx_sincos(k * r, &sn, &cs);
                // e^{ikr} q
                const double er = cs * sqr[j] - sn * sqi[j];
                const double ei = cs * sqi[j] + sn * sqr[j];
                pr += er * ir;
                pi += ei * ir;
                // i e^{ikr} q
                dr -= ei;
                di += er;
            }
    }
    if (act) {
        const size_t o = (size_t)c * nt + t;
        part_pot[2 * o] = pr;
        part_pot[2 * o + 1] = pi;
        if (part_dk) {
            part_dk[2 * o] = dr;
            part_dk[2 * o + 1] = di;
        }
    }
}

__global__ void k_h3_reduce(int nchunks, int nt, const double *part, double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double re = 0, im = 0;
    for (int c = 0; c < nchunks; ++c) {
        re += part[2 * ((size_t)c * nt + t)];
        im += part[2 * ((size_t)c * nt + t) + 1];
    }
    out[2 * (size_t)t] = re * INV4PI;
    out[2 * (size_t)t + 1] = im * INV4PI;
}

__global__ void __launch_bounds__(H3_THREADS)
k_h3_p2p_adjoint(int ns, const double *src, int nt, const double *tgt, double k,
                 const double *u, double *v)
{
    __shared__ double tx_[H3_TILE], ty_[H3_TILE], tz_[H3_TILE], ur[H3_TILE], ui[H3_TILE];
    const int s = blockIdx.x * H3_THREADS + threadIdx.x;
    const bool act = s < ns;
    const double sx = act ? src[s] : 0.0, sy = act ? src[(size_t)ns + s] : 0.0,
                 sz = act ? src[2 * (size_t)ns + s] : 0.0;
    double vr = 0, vi = 0;
    for (int b = 0; b < nt; b += H3_TILE) {
        const int nb = min(H3_TILE, nt - b);
        __syncthreads();
        for (int j = threadIdx.x; j < nb; j += H3_THREADS) {
            const int t = b + j;
            tx_[j] = tgt[t];
            ty_[j] = tgt[(size_t)nt + t];
            tz_[j] = tgt[2 * (size_t)nt + t];
            ur[j] = u[2 * (size_t)t];
            ui[j] = u[2 * (size_t)t + 1];
        }
        __syncthreads();
        if (act)
            for (int j = 0; j < nb; ++j) {
                const double dx = tx_[j] - sx, dy = ty_[j] - sy, dz = tz_[j] - sz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 == 0.0) continue;
                const double ir = qbx_rsqrt(r2);
                const double r = r2 * ir;
                double sn, cs;
                qbx_sincos(k * r, &sn, &cs);
                // e^{-ikr} u / r
                vr += (cs * ur[j] + sn * ui[j]) * ir;
                vi += (cs * ui[j] - sn * ur[j]) * ir;
            }
    }
    if (act) {
        v[2 * (size_t)s] = vr * INV4PI;
        v[2 * (size_t)s + 1] = vi * INV4PI;
    }
}

// out[e, j, r] = sum_i mat[i, j] in[e, i, r]  (the transpose of qbx_upsample)
__global__ void k_upsample_adjoint(int nel, int n1, int nf, int nr, const double *mat,
                                   const double *in, double *out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nel * n1 * nr) return;
    const long ej = idx / nr;
    const int r = (int)(idx % nr);
    const int e = (int)(ej / n1), j = (int)(ej % n1);
    double re = 0, im = 0;
    for (int i = 0; i < nf; ++i) {
        const double m = mat[i * n1 + j];
        const size_t kk = ((size_t)e * nf + i) * nr + r;
        re += m * in[2 * kk];
        im += m * in[2 * kk + 1];
    }
    out[2 * idx] = re;
    out[2 * idx + 1] = im;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_helmholtz3d_p2p(qbx_ctx *ctx, void *stream, int nsources, const double *sources,
                        const double *strengths, int ntargets, const double *targets, double k,
                        double *pot, double *dpot_dk)
{
    QBX_CHECK(ctx && pot, "qbx_helmholtz3d_p2p: null argument");
    QBX_CHECK(nsources >= 0 && ntargets >= 0, "qbx_helmholtz3d_p2p: negative size");
    QBX_CHECK(nsources == 0 || (sources && strengths), "qbx_helmholtz3d_p2p: null sources");
    QBX_CHECK(ntargets == 0 || targets, "qbx_helmholtz3d_p2p: null targets");
    hipStream_t st = S(stream);
    if (ntargets == 0) return QBX_OK;
    if (nsources == 0) {
        QBX_TRY(zero_fill(pot, 2L * ntargets, st));
        if (dpot_dk) QBX_TRY(zero_fill(dpot_dk, 2L * ntargets, st));
        return QBX_OK;
    }
    const int tiles = cdiv(ntargets, H3_THREADS);
    // enough blocks to fill the chip (>= 8 per CU on 256 CUs), chunks of at
    // least one LDS tile
    int nchunks = std::max(1, 2048 / tiles);
    nchunks = std::min(nchunks, cdiv(nsources, H3_TILE));
    const int chunk = cdiv(nsources, nchunks);
    nchunks = cdiv(nsources, chunk);
    const size_t pb = sizeof(double) * 2 * (size_t)nchunks * ntargets;
    void *scr;
    QBX_TRY(scratch_reserve(ctx, st, dpot_dk ? 2 * pb : pb, &scr));
    double *ppot = (double *)scr;
    double *pdk = dpot_dk ? ppot + 2 * (size_t)nchunks * ntargets : nullptr;
    hipLaunchKernelGGL(k_h3_p2p_partial, dim3(tiles, nchunks), dim3(H3_THREADS), 0, st, nsources,
                       sources, strengths, ntargets, targets, k, chunk, ppot, pdk);
    hipLaunchKernelGGL(k_h3_reduce, dim3(cdiv(ntargets, 256)), dim3(256), 0, st, nchunks,
                       ntargets, ppot, pot);
    if (dpot_dk)
        hipLaunchKernelGGL(k_h3_reduce, dim3(cdiv(ntargets, 256)), dim3(256), 0, st, nchunks,
                           ntargets, pdk, dpot_dk);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_helmholtz3d_p2p_adjoint(qbx_ctx *ctx, void *stream, int nsources, const double *sources,
                                int ntargets, const double *targets, double k, const double *u,
                                double *v)
{
    QBX_CHECK(ctx && v, "qbx_helmholtz3d_p2p_adjoint: null argument");
    QBX_CHECK(nsources >= 0 && ntargets >= 0, "qbx_helmholtz3d_p2p_adjoint: negative size");
    QBX_CHECK(nsources == 0 || sources, "qbx_helmholtz3d_p2p_adjoint: null sources");
    QBX_CHECK(ntargets == 0 || (targets && u), "qbx_helmholtz3d_p2p_adjoint: null targets");
    if (nsources == 0) return QBX_OK;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_h3_p2p_adjoint, dim3(cdiv(nsources, H3_THREADS)), dim3(H3_THREADS), 0,
                       st, nsources, sources, ntargets, targets, k, u, v);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_upsample_adjoint(qbx_ctx *ctx, void *stream, int nelements, int n1, int nf, int nrhs,
                         const double *mat, const double *in, double *out)
{
    QBX_CHECK(ctx && mat && in && out, "qbx_upsample_adjoint: null argument");
    QBX_CHECK(nelements >= 0 && n1 > 0 && nf > 0 && nrhs >= 1, "qbx_upsample_adjoint: bad size");
    const long n = (long)nelements * n1 * nrhs;
    if (n == 0) return QBX_OK;
    hipLaunchKernelGGL(k_upsample_adjoint, dim3(cdiv(n, 256)), dim3(256), 0, S(stream),
                       nelements, n1, nf, nrhs, mat, in, out);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // extern "C"

// Target-specific QBX (TSQBX) for 3D Laplace and Helmholtz: the GPU
// counterpart of eval_target_specific_qbx_locals
// (pytential/qbx/target_specific/impl.pyx:566-785, called from
// pytential/qbx/fmmlib.py:587-639).
//
// For every QBX center c, every target t associated with c and every source s
// in list 1 of c's target box:
//   Laplace   S   q   sum_{n<=p} rho_t^n / rho_s^{n+1} P_n(cos th)
//             Sp  grad_t of S,   D  dipstr nu . grad_s of S
//   Helmholtz S   i k q sum_{n<=p} (2n+1) j_n(k rho_t) h_n(k rho_s) P_n(cos th)
//             Sp  grad_t,        D  dipstr nu . grad_s
// (fmmlib convention: 4 pi times the free-space Green's function), with
// rho_t = |t-c|, rho_s = |s-c| and th the angle between t-c and s-c.
//
// Mapping: TSQ_LANES lanes per (center, target) pair split the sources; the
// Legendre and spherical-Hankel sequences run as recurrences per source, the
// target's j_n(k rho_t) (complex k) comes from Miller's backward recurrence
// into registers; lane groups reduce with shuffles.  Output overwrites
// pot / grad at the QBX targets, like the reference.
#include "qbx_internal.h"

namespace qbx {

struct zc {
    double x, y;
};
__device__ __forceinline__ zc operator+(zc a, zc b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ zc operator-(zc a, zc b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ zc operator*(zc a, zc b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ zc operator*(double s, zc a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ zc zinv(zc a)
{
    const double d = 1.0 / (a.x * a.x + a.y * a.y);
    return {a.x * d, -a.y * d};
}
__device__ __forceinline__ double zabs(zc a) { return sqrt(a.x * a.x + a.y * a.y); }
__device__ __forceinline__ zc zsin(zc z)
{
    return {sin(z.x) * cosh(z.y), cos(z.x) * sinh(z.y)};
}
__device__ __forceinline__ zc zcos(zc z)
{
    return {cos(z.x) * cosh(z.y), -sin(z.x) * sinh(z.y)};
}
__device__ __forceinline__ zc zexp_i(zc z)  // e^{i z}
{
    const double e = exp(-z.y);
    return {e * cos(z.x), e * sin(z.x)};
}

constexpr int TSQ_LANES = 16;

// j_n(z), j_n'(z) for n = 0..P (complex z) into registers
template <int P>
__device__ void sph_j(zc z, zc *j, zc *jd)
{
    const double az = zabs(z);
    if (az < 1e-300) {
#pragma unroll
        for (int n = 0; n <= P; ++n) {
            j[n] = {n == 0 ? 1.0 : 0.0, 0.0};
            jd[n] = {n == 1 ? 1.0 / 3.0 : 0.0, 0.0};
        }
        return;
    }
    const double m = fmax((double)(P + 1), az);
    const int M = (int)m + 30 + (int)sqrt(40.0 * m);
    const zc iz = zinv(z);
    zc fp1 = {0, 0}, f = {1e-30, 0};
    zc jj[P + 2];
#pragma unroll
    for (int n = 0; n <= P + 1; ++n) jj[n] = {0, 0};
    for (int n = M; n >= 1; --n) {
#pragma unroll
        for (int q = 0; q <= P + 1; ++q)
            if (q == n) jj[q] = f;
        const zc fm1 = ((2.0 * n + 1.0) * iz) * f - fp1;
        fp1 = f;
        f = fm1;
        if (fabs(f.x) + fabs(f.y) > 1e250) {
            f = 1e-250 * f;
            fp1 = 1e-250 * fp1;
#pragma unroll
            for (int q = 0; q <= P + 1; ++q) jj[q] = 1e-250 * jj[q];
        }
    }
    jj[0] = f;
    const zc j0 = zsin(z) * iz;
    const zc j1 = zsin(z) * iz * iz - zcos(z) * iz;
    const zc sc = zabs(j0) >= zabs(j1) ? j0 * zinv(jj[0]) : j1 * zinv(jj[1]);
#pragma unroll
    for (int n = 0; n <= P + 1; ++n) jj[n] = jj[n] * sc;
#pragma unroll
    for (int n = 0; n <= P; ++n) {
        j[n] = jj[n];
        jd[n] = n == 0 ? zc{-jj[1].x, -jj[1].y} : jj[n - 1] - ((n + 1.0) * iz) * jj[n];
    }
}

// P_n(ct) and P_n'(ct) from the three-term recurrences; n = 0, 1, 2, ... in
// order, (Pm2, Pm1, dPm2, dPm1) the running state initialised to (1, ct, 0, 1)
__device__ __forceinline__ void legendre_step(int n, double ct, double &Pm2, double &Pm1,
                                              double &dPm2, double &dPm1, double &Pn,
                                              double &dPn)
{
    if (n == 0) {
        Pn = 1.0;
        dPn = 0.0;
    } else if (n == 1) {
        Pn = ct;
        dPn = 1.0;
    } else {
        Pn = ((2.0 * n - 1.0) * ct * Pm1 - (n - 1.0) * Pm2) / n;
        dPn = ((2.0 * n - 1.0) * (Pm1 + ct * dPm1) - (n - 1.0) * dPm2) / n;
        Pm2 = Pm1;
        Pm1 = Pn;
        dPm2 = dPm1;
        dPm1 = dPn;
    }
}

// h_n(z) for n = 0, 1, 2, ... in order by the forward recurrence
// h_n = (2n-1)/z h_{n-1} - h_{n-2}; (hm2, hm1) the running state (h0, h1)
__device__ __forceinline__ zc hankel_step(int n, zc iz, zc h0, zc h1, zc &hm2, zc &hm1)
{
    if (n == 0) return h0;
    if (n == 1) return h1;
    const zc hn = ((2.0 * n - 1.0) * iz) * hm1 - hm2;
    hm2 = hm1;
    hm1 = hn;
    return hn;
}

// One source's term of the order-P target-specific expansion about a center,
// unit strength, fmmlib scaling (4 pi G): tc = t - c (|tc| = rt), sc = s - c
// (|sc| = rs), ct = cos of the angle between them; jv = j_n(k rt) (Helmholtz).
//   OP 0 (S):  Laplace sum_n rt^n / rs^{n+1} P_n(ct)
//              Helmholtz i k sum_n (2n+1) j_n(k rt) h_n(k rs) P_n(ct)
//   OP 1 (D):  nu . grad_s of the S term
template <int KERNEL, int OP, int P>
__device__ __forceinline__ zc tsq_pair_sd(zc k, const double *tc, double rt, const zc *jv,
                                          const double *sc, double rs, double ct,
                                          const double *nu)
{
    double Pm2 = 1.0, Pm1 = ct, dPm2 = 0.0, dPm1 = 1.0;
    if (KERNEL == 0) {
        if (OP == 0) {
            double r = 1.0 / rs, acc = 0.0;
#pragma unroll
            for (int n = 0; n <= P; ++n) {
                double Pn, dPn;
                legendre_step(n, ct, Pm2, Pm1, dPm2, dPm1, Pn, dPn);
                acc += r * Pn;
                r *= rt / rs;
            }
            return {acc, 0.0};
        }
        double gv[3] = {0, 0, 0};
        double Rn = 1.0 / (rs * rs);
#pragma unroll
        for (int n = 0; n <= P; ++n) {
            double Pn, dPn;
            legendre_step(n, ct, Pm2, Pm1, dPm2, dPm1, Pn, dPn);
            for (int m = 0; m < 3; ++m)
                gv[m] += Rn * (-(n + 1.0) * (sc[m] / rs) * Pn +
                               (tc[m] / rt - ct * sc[m] / rs) * dPn);
            Rn *= rt / rs;
        }
        return {nu[0] * gv[0] + nu[1] * gv[1] + nu[2] * gv[2], 0.0};
    }
    // spherical Hankel h_n(k rho_s) by forward recurrence
    const zc z = rs * k;
    const zc iz = zinv(z);
    const zc h0 = zexp_i(z) * zinv(zc{-z.y, z.x});  // e^{iz} / (iz)
    const zc h1 = h0 * (iz - zc{0, 1});
    zc hm2 = h0, hm1 = h1;
    zc acc = {0, 0};
    zc av[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
    for (int n = 0; n <= P; ++n) {
        double Pn, dPn;
        legendre_step(n, ct, Pm2, Pm1, dPm2, dPm1, Pn, dPn);
        const zc hprev = n == 1 ? h0 : hm1;
        const zc hn = hankel_step(n, iz, h0, h1, hm2, hm1);
        const double w = 2.0 * n + 1.0;
        if (OP == 0) {
            acc = acc + (w * Pn) * (jv[n] * hn);
        } else {
            // h_n' = h_{n-1} - (n+1)/z h_n,  h_0' = -h_1
            const zc hd = n == 0 ? zc{-h1.x, -h1.y} : hprev - ((n + 1.0) * iz) * hn;
            const zc a = (w * Pn) * (k * hd);
            const zc b = (w * dPn) * hn;
            for (int m = 0; m < 3; ++m)
                av[m] = av[m] + (1.0 / rs) * (jv[n] * (sc[m] * a +
                                               (tc[m] / rt - ct * sc[m] / rs) * b));
        }
    }
    const zc ik = {-k.y, k.x};
    if (OP == 0) return ik * acc;
    return ik * (nu[0] * av[0] + nu[1] * av[1] + nu[2] * av[2]);
}

template <int KERNEL, int OP, int P>
__global__ void __launch_bounds__(256) k_tsqbx(double kre, double kim, int nctt,
                                               const int *ctt_lists, const int *tree_ttc,
                                               const int *c2tb, const double *ctr, int nc,
                                               const double *tgt, int nt, const double *src,
                                               int ns, const int *l1s, const int *l1l,
                                               const int *sstart, const int *scount,
                                               const double *q, const double *dip,
                                               double *pot, double *grad)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int pr = gid / TSQ_LANES, lane = gid % TSQ_LANES;
    const bool active = pr < nctt;
    const zc k = {kre, kim};
    zc res = {0, 0};
    zc g[3] = {{0, 0}, {0, 0}, {0, 0}};
    int ti = 0;
    if (active) {
        ti = ctt_lists[pr];
        const int c = tree_ttc[ti];
        const int tb = c2tb[c];
        double tc[3];
        for (int m = 0; m < 3; ++m) tc[m] = tgt[(size_t)m * nt + ti] - ctr[(size_t)m * nc + c];
        const double rt = sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
        zc jv[P + 1], jdv[P + 1];
        if (KERNEL == 1) sph_j<P>(rt * k, jv, jdv);
        for (int e = l1s[tb]; e < l1s[tb + 1]; ++e) {
            const int sb = l1l[e];
            const int s1 = sstart[sb] + scount[sb];
            for (int s = sstart[sb] + lane; s < s1; s += TSQ_LANES) {
                double sc[3];
                for (int m = 0; m < 3; ++m) sc[m] = src[(size_t)m * ns + s] - ctr[(size_t)m * nc + c];
                const double rs = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]);
                const double ct = (tc[0] * sc[0] + tc[1] * sc[1] + tc[2] * sc[2]) / (rt * rs);
                const zc qs = {q[2 * s], q[2 * s + 1]};
                if (OP != 2) {
                    double nu[3] = {0, 0, 0};
                    if (OP == 1) {
                        nu[0] = dip[s];
                        nu[1] = dip[(size_t)ns + s];
                        nu[2] = dip[2 * (size_t)ns + s];
                    }
                    res = res + tsq_pair_sd<KERNEL, OP, P>(k, tc, rt, jv, sc, rs, ct, nu) * qs;
                } else if (KERNEL == 0) {
                    double Pm2 = 1.0, Pm1 = ct, dPm2 = 0.0, dPm1 = 1.0;
                    double gv[3] = {0, 0, 0};
                    double Rn = 1.0 / (rs * rs);
#pragma unroll
                    for (int n = 0; n <= P; ++n) {
                        double Pn, dPn;
                        legendre_step(n, ct, Pm2, Pm1, dPm2, dPm1, Pn, dPn);
                        if (n >= 1) {
                            for (int m = 0; m < 3; ++m)
                                gv[m] += Rn * (n * (tc[m] / rt) * Pn +
                                               (sc[m] / rs - ct * tc[m] / rt) * dPn);
                            Rn *= rt / rs;
                        }
                    }
                    for (int m = 0; m < 3; ++m) g[m] = g[m] + gv[m] * qs;
                } else {
                    // spherical Hankel h_n(k rho_s) by forward recurrence
                    double Pm2 = 1.0, Pm1 = ct, dPm2 = 0.0, dPm1 = 1.0;
                    const zc z = rs * k;
                    const zc iz = zinv(z);
                    const zc h0 = zexp_i(z) * zinv(zc{-z.y, z.x});  // e^{iz} / (iz)
                    const zc h1 = h0 * (iz - zc{0, 1});
                    zc hm2 = h0, hm1 = h1;
                    zc av[3] = {{0, 0}, {0, 0}, {0, 0}};
#pragma unroll
                    for (int n = 0; n <= P; ++n) {
                        double Pn, dPn;
                        legendre_step(n, ct, Pm2, Pm1, dPm2, dPm1, Pn, dPn);
                        const zc hn = hankel_step(n, iz, h0, h1, hm2, hm1);
                        const double w = 2.0 * n + 1.0;
                        const zc a = (w * Pn) * (k * jdv[n]);
                        const zc b = (w * dPn) * jv[n];
                        for (int m = 0; m < 3; ++m)
                            av[m] = av[m] + (1.0 / rt) * (hn * (tc[m] * a +
                                                   (sc[m] / rs - ct * tc[m] / rt) * b));
                    }
                    const zc ikq = zc{-k.y, k.x} * qs;  // i k q
                    for (int m = 0; m < 3; ++m) g[m] = g[m] + ikq * av[m];
                }
            }
        }
    }
    for (int m = TSQ_LANES / 2; m >= 1; m >>= 1) {
        res.x += __shfl_xor(res.x, m, TSQ_LANES);
        res.y += __shfl_xor(res.y, m, TSQ_LANES);
        for (int d = 0; d < 3; ++d) {
            g[d].x += __shfl_xor(g[d].x, m, TSQ_LANES);
            g[d].y += __shfl_xor(g[d].y, m, TSQ_LANES);
        }
    }
    if (active && lane == 0) {
        if (OP == 2) {
            for (int d = 0; d < 3; ++d) {
                grad[2 * ((size_t)d * nt + ti)] = g[d].x;
                grad[2 * ((size_t)d * nt + ti) + 1] = g[d].y;
            }
        } else {
            pot[2 * (size_t)ti] = res.x;
            pot[2 * (size_t)ti + 1] = res.y;
        }
    }
}

// The target-specific sum for many densities (3D Laplace S and D) on the
// FP64 matrix cores: pot[t][r] (+)= sum_s G(t, s) q[s][r] with G the
// unit-strength term of tsq_pair_sd -- real for Laplace, computed once per
// (target, source) pair for all right-hand sides.  A block per unit of
// tsq_plan: up to TSQ_ROWS QBX targets whose centers share target box it (the
// A rows: lane i evaluates G for row i % 16 and source k0 + i / 16, exactly
// the A operand of v_mfma_f64_16x16x4f64), K = the box's list-1 sources in
// steps of four, shared round-robin by the four waves, columns two tiles of
// 16 right-hand sides (blockIdx.y chunks of 32).  The waves' partial sums
// are added in LDS in wave order and written once.
typedef double tsq_d4 __attribute__((ext_vector_type(4)));
constexpr int TSQM_RB = 32;

template <int OP, int P>
__global__ void __launch_bounds__(256) k3_tsqbx_mr(int nr, const int *ubox, const int *rows,
                                                   const int *tree_ttc, const double *ctr,
                                                   int nc, const double *tgt, int nt,
                                                   const double *src, int ns,
                                                   const int *l1s, const int *l1l,
                                                   const int *sstart, const int *scount,
                                                   const double *q, const double *dip,
                                                   double scale, double *pot, int accumulate)
{
    __shared__ double part[4][TSQ_ROWS][TSQM_RB][2];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rb0 = blockIdx.y * TSQM_RB;
    const int nrb = min(TSQM_RB, nr - rb0);
    // this lane's A row: target ti of center c
    const int ti = rows[TSQ_ROWS * u + col];
    const bool rin = ti >= 0;
    double tc[3] = {0, 0, 0}, cc[3] = {0, 0, 0}, rt = 1.0;
    if (rin) {
        const int c = tree_ttc[ti];
        for (int m = 0; m < 3; ++m) {
            cc[m] = ctr[(size_t)m * nc + c];
            tc[m] = tgt[(size_t)m * nt + ti] - cc[m];
        }
        rt = sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
    }
    const zc k0 = {0, 0};
    tsq_d4 cre[2], cim[2];
    for (int h = 0; h < 2; ++h) cre[h] = cim[h] = tsq_d4{0, 0, 0, 0};
    int step = 0;  // K-step counter over the whole list, dealt to the waves
    for (int e = l1s[it]; e < l1s[it + 1]; ++e) {
        const int sb = l1l[e];
        const int s0 = sstart[sb], n = scount[sb];
        const int nks = (n + 3) / 4;
        // the K-steps of this box that fall to this wave
        int ks = (wv - step) & 3;
        for (; ks < nks; ks += 4) {
            const int k = ks * 4 + kr;
            const bool kin = k < n;
            const int s = s0 + (kin ? k : 0);
            double g = 0.0;
            if (rin && kin) {
                double sc[3], nu[3] = {0, 0, 0};
                for (int m = 0; m < 3; ++m) sc[m] = src[(size_t)m * ns + s] - cc[m];
                const double rs = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]);
                const double ct = (tc[0] * sc[0] + tc[1] * sc[1] + tc[2] * sc[2]) / (rt * rs);
                if (OP == 1)
                    for (int m = 0; m < 3; ++m) nu[m] = dip[(size_t)m * ns + s];
                g = tsq_pair_sd<0, OP, P>(k0, tc, rt, nullptr, sc, rs, ct, nu).x;
            }
            double bre[2] = {0, 0}, bim[2] = {0, 0};
            for (int h = 0; h < 2; ++h) {
                const int r = h * 16 + col;
                if (kin && r < nrb) {
                    bre[h] = q[2 * ((size_t)s * nr + rb0 + r)];
                    bim[h] = q[2 * ((size_t)s * nr + rb0 + r) + 1];
                }
            }
            // real strengths (the usual case) need no imaginary-part products;
            // the test is wave-uniform
            const bool anyim = __ballot(bim[0] != 0.0 || bim[1] != 0.0) != 0;
            for (int h = 0; h < 2; ++h) {
                cre[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(g, bre[h], cre[h], 0, 0, 0);
                if (anyim)
                    cim[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(g, bim[h], cim[h], 0, 0, 0);
            }
        }
        step = (step + nks) & 3;
    }
    // D register j of lane i: row 4 j + i / 16, column i % 16
    for (int h = 0; h < 2; ++h)
        for (int j = 0; j < 4; ++j) {
            part[wv][4 * j + kr][h * 16 + col][0] = cre[h][j];
            part[wv][4 * j + kr][h * 16 + col][1] = cim[h][j];
        }
    __syncthreads();
    for (int e = tid; e < TSQ_ROWS * TSQM_RB; e += 256) {
        const int row = e / TSQM_RB, r = e % TSQM_RB;
        const int tr = rows[TSQ_ROWS * u + row];
        if (tr < 0 || r >= nrb) continue;
        double vr = 0.0, vi = 0.0;
        for (int w = 0; w < 4; ++w) {
            vr += part[w][row][r][0];
            vi += part[w][row][r][1];
        }
        vr *= scale;
        vi *= scale;
        double *dst = pot + 2 * ((size_t)tr * nr + rb0 + r);
        if (accumulate) {
            dst[0] += vr;
            dst[1] += vi;
        } else {
            dst[0] = vr;
            dst[1] = vi;
        }
    }
}

template <int KERNEL, int OP>
static int launch_p(int order, dim3 grid, hipStream_t st, double kre, double kim, int nctt,
                    const int *a0, const int *a1, const int *a2, const double *ctr, int nc,
                    const double *tgt, int nt, const double *src, int ns, const int *l1s,
                    const int *l1l, const int *ss, const int *sc, const double *q,
                    const double *dip, double *pot, double *grad)
{
#define TSQ_CASE(PP)                                                                        \
    case PP:                                                                                \
        hipLaunchKernelGGL((k_tsqbx<KERNEL, OP, PP>), grid, dim3(256), 0, st, kre, kim,     \
                           nctt, a0, a1, a2, ctr, nc, tgt, nt, src, ns, l1s, l1l, ss, sc, q, \
                           dip, pot, grad);                                                 \
        break;
    switch (order) {
        TSQ_CASE(0) TSQ_CASE(1) TSQ_CASE(2) TSQ_CASE(3) TSQ_CASE(4) TSQ_CASE(5) TSQ_CASE(6)
        TSQ_CASE(7) TSQ_CASE(8) TSQ_CASE(9) TSQ_CASE(10) TSQ_CASE(11) TSQ_CASE(12)
    default:
        set_error("TSQBX order must be in [0, 12]");
        return QBX_ERR_INVALID;
    }
#undef TSQ_CASE
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// One block per target row t of a dense layer-potential matrix:
//   mat[t][e n1 + j] = sum_i K(t, s_ei) w_ei U[i][j]
// K the order-P target-specific entry about the row's center (the QBX
// expansion of a unit source evaluated at t) or, for P < 0, the kernel
// itself; both with the 1/(4 pi) of G.  Layout and contraction as in the 2D
// kernel: the fine sources of LM_THREADS / nf elements one per thread, their
// weighted values through LDS into the per-element contraction with U.
template <int KERNEL, int OP, int P>
__global__ void __launch_bounds__(LM_THREADS)
k_lpot_matrix_3d(double kre, int nel, int n1, int nf, const double *U, const double *src,
                 const double *nrm, const double *w, int nt, const double *tgt,
                 const double *ctr, long ld, double *mat)
{
    constexpr int PJ = P >= 0 ? P : 0;
    constexpr double INV4PI_ = 0.07957747154594766788;
    __shared__ double kv[2 * LM_THREADS];
    __shared__ double Us[LM_UMAX];
    const int t = blockIdx.x, tid = threadIdx.x, ns = nel * nf;
    const zc k = {kre, 0.0};
    const double tp[3] = {tgt[t], tgt[(size_t)nt + t], tgt[2 * (size_t)nt + t]};
    double tc[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, rt = 0.0;
    zc jv[PJ + 1], jdv[PJ + 1];
    if (P >= 0) {
        for (int m = 0; m < 3; ++m) {
            cv[m] = ctr[(size_t)m * nt + t];
            tc[m] = tp[m] - cv[m];
        }
        rt = sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
        if (KERNEL == 1) sph_j<PJ>(rt * k, jv, jdv);
    }
    for (int i = tid; i < nf * n1; i += LM_THREADS)
        Us[i] = U ? U[i] : (i / n1 == i % n1 ? 1.0 : 0.0);
    __syncthreads();
    const int epb = LM_THREADS / nf;
    const int le = tid / nf, fi = tid % nf;
    for (int e0 = 0; e0 < nel; e0 += epb) {
        const int e = e0 + le;
        zc v = {0, 0};
        if (le < epb && e < nel) {
            const int s = e * nf + fi;
            const double sp[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
            double nu[3] = {0, 0, 0};
            if (OP == 1)
                for (int m = 0; m < 3; ++m) nu[m] = nrm[(size_t)m * ns + s];
            if constexpr (P < 0) {
                const double d[3] = {tp[0] - sp[0], tp[1] - sp[1], tp[2] - sp[2]};
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 > 0.0) {
                    const double r = sqrt(r2), ir = 1.0 / r;
                    double sn, cs;
                    sincos(kre * r, &sn, &cs);
                    v = {cs * ir, sn * ir};
                    if (OP == 1) {
                        const double proj = (d[0] * nu[0] + d[1] * nu[1] + d[2] * nu[2]) * ir * ir;
                        v = (proj * zc{1.0, -kre * r}) * v;
                    }
                }
            } else {
                double sc[3];
                for (int m = 0; m < 3; ++m) sc[m] = sp[m] - cv[m];
                const double rs = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]);
                const double ct = (tc[0] * sc[0] + tc[1] * sc[1] + tc[2] * sc[2]) / (rt * rs);
                v = tsq_pair_sd<KERNEL, OP, PJ>(k, tc, rt, jv, sc, rs, ct, nu);
            }
            v = (w[s] * INV4PI_) * v;
        }
        kv[2 * tid] = v.x;
        kv[2 * tid + 1] = v.y;
        __syncthreads();
        const int ne = min(epb, nel - e0);
        for (int o = tid; o < ne * n1; o += LM_THREADS) {
            const int le2 = o / n1, j = o % n1;
            const double *kr = kv + 2 * le2 * nf;
            double ax = 0.0, ay = 0.0;
            for (int ii = 0; ii < nf; ++ii) {
                const double u = Us[ii * n1 + j];
                ax += u * kr[2 * ii];
                ay += u * kr[2 * ii + 1];
            }
            const size_t off = 2 * ((size_t)t * ld + (size_t)(e0 + le2) * n1 + j);
            mat[off] = ax;
            mat[off + 1] = ay;
        }
        __syncthreads();
    }
}

template <int KERNEL, int OP>
static int launch_lpot_matrix_3d(int order, bool qbx, hipStream_t st, double k, int nel, int n1,
                                 int nf, const double *U, const double *src, const double *nrm,
                                 const double *w, int nt, const double *tgt, const double *ctr,
                                 long ld, double *mat)
{
    const dim3 grid(nt), blk(LM_THREADS);
#define LM3_ARGS k, nel, n1, nf, U, src, nrm, w, nt, tgt, ctr, ld, mat
    if (!qbx) {
        hipLaunchKernelGGL((k_lpot_matrix_3d<KERNEL, OP, -1>), grid, blk, 0, st, LM3_ARGS);
    } else {
        switch (order) {
#define LM3_CASE(PP)                                                                          \
    case PP:                                                                                  \
        hipLaunchKernelGGL((k_lpot_matrix_3d<KERNEL, OP, PP>), grid, blk, 0, st, LM3_ARGS); \
        break;
            LM3_CASE(0) LM3_CASE(1) LM3_CASE(2) LM3_CASE(3) LM3_CASE(4) LM3_CASE(5)
            LM3_CASE(6) LM3_CASE(7) LM3_CASE(8) LM3_CASE(9) LM3_CASE(10) LM3_CASE(11)
            LM3_CASE(12)
#undef LM3_CASE
        default:
            set_error("qbx_order must be in [0, 12]");
            return QBX_ERR_INVALID;
        }
    }
#undef LM3_ARGS
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// one virtual list-1 box holding every source, for the direct (no FMM) form:
// l1s[i] = i, l1l[i] = 0 for each target box, box 0 = sources [0, ns)
__global__ void k_virtual_lists(int ntb, int ns, int *buf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= ntb) buf[i] = i;
    if (i < ntb) buf[ntb + 1 + i] = 0;
    if (i == 0) {
        buf[2 * ntb + 1] = 0;
        buf[2 * ntb + 2] = ns;
    }
}

__global__ void k_scale_c(long n, double s, double *a)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * n) a[i] *= s;
}

// Direct 3D sums for the targets that need no QBX: one thread per target,
// sources through LDS.  G = e^{ikr}/(4 pi r) (k = 0: Laplace); DIP: its
// derivative along the source normal, n.(t - s) (1 - ikr) e^{ikr}/(4 pi r^3).
constexpr int D3_TILE = 256;

template <bool DIP>
__global__ void __launch_bounds__(256) k3_direct_p2p(double kre, double kim, int nf,
                                                     const int *unfilt, const double *tgt,
                                                     int nt, const double *src, int ns,
                                                     const double *q, const double *nrm,
                                                     double *pot)
{
    __shared__ double sx[D3_TILE], sy[D3_TILE], sz[D3_TILE], sqr[D3_TILE], sqi[D3_TILE];
    __shared__ double snx[DIP ? D3_TILE : 1], sny[DIP ? D3_TILE : 1], snz[DIP ? D3_TILE : 1];
    const int f = blockIdx.x * 256 + threadIdx.x;
    const bool act = f < nf;
    int ti = 0;
    double tx = 0, ty = 0, tz = 0;
    if (act) {
        ti = unfilt ? unfilt[f] : f;
        tx = tgt[ti];
        ty = tgt[(size_t)nt + ti];
        tz = tgt[2 * (size_t)nt + ti];
    }
    const zc ik = {-kim, kre};
    double ar = 0, ai = 0;
    for (int b = 0; b < ns; b += D3_TILE) {
        const int nb = min(D3_TILE, ns - b);
        __syncthreads();
        for (int j = threadIdx.x; j < nb; j += 256) {
            const int s = b + j;
            sx[j] = src[s];
            sy[j] = src[(size_t)ns + s];
            sz[j] = src[2 * (size_t)ns + s];
            sqr[j] = q[2 * (size_t)s];
            sqi[j] = q[2 * (size_t)s + 1];
            if (DIP) {
                snx[j] = nrm[s];
                sny[j] = nrm[(size_t)ns + s];
                snz[j] = nrm[2 * (size_t)ns + s];
            }
        }
        __syncthreads();
        if (act)
            for (int j = 0; j < nb; ++j) {
                const double dx = tx - sx[j], dy = ty - sy[j], dz = tz - sz[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 == 0.0) continue;
                const double r = sqrt(r2), ir = 1.0 / r;
                // e^{ikr} = e^{-kim r} (cos kre r + i sin kre r)
                double sn, cs;
                sincos(kre * r, &sn, &cs);
                const double dec = kim == 0.0 ? 1.0 : exp(-kim * r);
                zc g = {dec * cs * ir, dec * sn * ir};
                if (DIP) {
                    const double proj = (dx * snx[j] + dy * sny[j] + dz * snz[j]) * ir * ir;
                    const zc fac = {1.0 - r * ik.x, -r * ik.y};
                    g = (proj * fac) * g;
                }
                ar += g.x * sqr[j] - g.y * sqi[j];
                ai += g.x * sqi[j] + g.y * sqr[j];
            }
    }
    if (act) {
        constexpr double INV4PI_ = 0.07957747154594766788;
        pot[2 * (size_t)ti] = ar * INV4PI_;
        pot[2 * (size_t)ti + 1] = ai * INV4PI_;
    }
}

// The many-density target-specific sum; scale multiplies the result (1 for
// the public entry point's unscaled form, 1/(4 pi) inside the fused driver).
int tsqbx_nrhs(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
               const qbx_geo *G, int kernel, int op, int qbx_order, int nrhs,
               const double *strengths_tree, const double *dipvec_tree, double scale,
               double *pot, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && strengths_tree && pot, "TSQBX: null argument");
    QBX_CHECK(T->dim == 3, "TSQBX is defined for 3D trees (got dim %d)", T->dim);
    QBX_CHECK(kernel == QBX_LAPLACE3D, "many-density TSQBX: 3D Laplace only (got %d)", kernel);
    QBX_CHECK(op == QBX_TSQBX_S || op == QBX_TSQBX_D, "many-density TSQBX: op S or D (got %d)",
              op);
    QBX_CHECK(op != QBX_TSQBX_D || dipvec_tree, "TSQBX double layer needs dipole vectors");
    QBX_CHECK(qbx_order >= 0 && qbx_order <= 12, "TSQBX order must be in [0, 12]");
    QBX_CHECK(nrhs >= 1, "TSQBX: nrhs must be >= 1");
    QBX_CHECK(G->trav == R, "TSQBX: geometry was built for another traversal");
    const unit_plan &V = G->tsq_plan;
    if (V.nunits == 0) return QBX_OK;
    hipStream_t st = S(stream);
    const dim3 grid(V.nunits, cdiv(nrhs, TSQM_RB));
#define TSQM_CASE(OPV, PP)                                                                  \
    hipLaunchKernelGGL((k3_tsqbx_mr<OPV, PP>), grid, dim3(256), 0, st, nrhs, V.unit_box.p,  \
                       G->tsq_rows.p, G->tree_ttc.p, G->centers.p, G->ncenters,            \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources,               \
                       R->l1_starts.p, R->l1_lists.p, T->box_source_starts.p,              \
                       T->box_source_counts_nonchild.p, strengths_tree, dipvec_tree, scale, \
                       pot, accumulate);                                                         \
    break;
#define TSQM_OP(OPV)                  \
    switch (qbx_order) {              \
    case 0: TSQM_CASE(OPV, 0)         \
    case 1: TSQM_CASE(OPV, 1)         \
    case 2: TSQM_CASE(OPV, 2)         \
    case 3: TSQM_CASE(OPV, 3)         \
    case 4: TSQM_CASE(OPV, 4)         \
    case 5: TSQM_CASE(OPV, 5)         \
    case 6: TSQM_CASE(OPV, 6)         \
    case 7: TSQM_CASE(OPV, 7)         \
    case 8: TSQM_CASE(OPV, 8)         \
    case 9: TSQM_CASE(OPV, 9)         \
    case 10: TSQM_CASE(OPV, 10)       \
    case 11: TSQM_CASE(OPV, 11)       \
    default: TSQM_CASE(OPV, 12)       \
    }
    if (op == QBX_TSQBX_S) {
        TSQM_OP(0)
    } else {
        TSQM_OP(1)
    }
#undef TSQM_OP
#undef TSQM_CASE
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_eval_target_specific_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                        const qbx_trav *R, const qbx_geo *G, int kernel,
                                        int op, double k_re, double k_im, int qbx_order,
                                        const double *strengths_tree, const double *dipvec_tree,
                                        double *pot, double *grad)
{
    QBX_CHECK(ctx && T && R && G && strengths_tree, "TSQBX: null argument");
    QBX_CHECK(T->dim == 3, "TSQBX is defined for 3D trees (got dim %d)", T->dim);
    QBX_CHECK(kernel == QBX_LAPLACE3D || kernel == QBX_HELMHOLTZ3D, "TSQBX: bad kernel %d",
              kernel);
    QBX_CHECK(op >= QBX_TSQBX_S && op <= QBX_TSQBX_SP, "TSQBX: bad op %d", op);
    QBX_CHECK(op != QBX_TSQBX_SP || grad, "TSQBX gradient needs a grad array");
    QBX_CHECK(op == QBX_TSQBX_SP || pot, "TSQBX potential needs a pot array");
    QBX_CHECK(op != QBX_TSQBX_D || dipvec_tree, "TSQBX double layer needs dipole vectors");
    QBX_CHECK(qbx_order >= 0 && qbx_order <= 12, "TSQBX order must be in [0, 12]");
    if (G->nctt == 0) return QBX_OK;
    hipStream_t st = S(stream);
    const dim3 grid(cdiv((long)G->nctt * TSQ_LANES, 256));
    const int hk = kernel == QBX_HELMHOLTZ3D;
#define TSQ_ARGS                                                                          \
    qbx_order, grid, st, k_re, k_im, G->nctt, G->ctt_lists.p, G->tree_ttc.p,             \
        G->center_to_target_box.p, G->centers.p, G->ncenters, T->targets.p, T->ntargets, \
        T->sources.p, T->nsources, R->l1_starts.p, R->l1_lists.p, T->box_source_starts.p, \
        T->box_source_counts_nonchild.p, strengths_tree, dipvec_tree, pot, grad
    if (!hk && op == QBX_TSQBX_S) return launch_p<0, 0>(TSQ_ARGS);
    if (!hk && op == QBX_TSQBX_D) return launch_p<0, 1>(TSQ_ARGS);
    if (!hk && op == QBX_TSQBX_SP) return launch_p<0, 2>(TSQ_ARGS);
    if (hk && op == QBX_TSQBX_S) return launch_p<1, 0>(TSQ_ARGS);
    if (hk && op == QBX_TSQBX_D) return launch_p<1, 1>(TSQ_ARGS);
    return launch_p<1, 2>(TSQ_ARGS);
#undef TSQ_ARGS
}

int qbx_eval_target_specific_qbx_locals_nrhs(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                             const qbx_trav *R, const qbx_geo *G, int kernel,
                                             int op, int qbx_order, int nrhs,
                                             const double *strengths_tree,
                                             const double *dipvec_tree, double *pot,
                                             int accumulate)
{
    return qbx::tsqbx_nrhs(ctx, stream, T, R, G, kernel, op, qbx_order, nrhs, strengths_tree,
                           dipvec_tree, 1.0, pot, accumulate);
}

}  // extern "C"

namespace qbx {

int lpot_matrix_3d(hipStream_t st, const qbx_kernel_desc *d, int nel, int n1, int nf,
                   const double *U, const double *src, const double *nrm, const double *w,
                   int nt, const double *tgt, const double *ctr, const double *rad, long ld,
                   double *mat)
{
    (void)rad;  // the target-specific form needs no expansion radius
    QBX_CHECK(d->source_kind == QBX_SOURCE_S || d->source_kind == QBX_SOURCE_D,
              "unsupported source kind %d", d->source_kind);
    const bool dip = d->source_kind == QBX_SOURCE_D;
    QBX_CHECK(!dip || nrm, "double layer needs normals");
    const double k = d->kernel == QBX_HELMHOLTZ3D ? d->k : 0.0;
    const bool qbx = ctr != nullptr;
#define LM3(KN, OPV) \
    launch_lpot_matrix_3d<KN, OPV>(d->qbx_order, qbx, st, k, nel, n1, nf, U, src, nrm, w, nt, \
                                   tgt, ctr, ld, mat)
    if (d->kernel == QBX_LAPLACE3D) return dip ? LM3(0, 1) : LM3(0, 0);
    return dip ? LM3(1, 1) : LM3(1, 0);
#undef LM3
}

// Direct 3D point sums: pot[ti] = sum_s G(t_ti, s) q_s for the nf targets
// ti = unfilt[f] (unfilt null: ti = f), written, not accumulated.
int direct_p2p3d(hipStream_t st, double k, bool dip, int nf, const int *unfilt,
                 const double *tgt, int nt, const double *src, int ns, const double *str,
                 const double *nrm, double *pot)
{
    if (nf == 0) return QBX_OK;
    if (dip)
        hipLaunchKernelGGL(k3_direct_p2p<true>, dim3(cdiv(nf, 256)), dim3(256), 0, st, k, 0.0, nf,
                           unfilt, tgt, nt, src, ns, str, nrm, pot);
    else
        hipLaunchKernelGGL(k3_direct_p2p<false>, dim3(cdiv(nf, 256)), dim3(256), 0, st, k, 0.0,
                           nf, unfilt, tgt, nt, src, ns, str, nrm, pot);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// Direct 3D QBX: the target-specific sums over every source at QBX targets
// (exactly the order-p_q QBX expansion of all sources, evaluated), direct
// sums at the other targets.  pot complex [ntargets] tree order, scaled
// (1/(4 pi) included).  kind: QBX_SOURCE_S or QBX_SOURCE_D.
int direct_qbx3d(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_geo *G, int kernel,
                 int kind, double k, int qbx_order, const double *str, const double *nrm,
                 double *pot)
{
    QBX_CHECK(T->dim == 3, "direct 3D QBX needs a 3D tree");
    QBX_CHECK(qbx_order >= 0 && qbx_order <= 12, "direct 3D QBX: order must be in [0, 12]");
    QBX_CHECK(kind == QBX_SOURCE_S || nrm, "double layer needs normals");
    QBX_TRY(zero_fill(pot, 2L * T->ntargets, st));
    const double kk = kernel == QBX_HELMHOLTZ3D ? k : 0.0;
    if (G->nctt > 0) {
        // center_to_target_box indexes the target boxes of the geometry's traversal
        QBX_CHECK(G->trav, "direct 3D QBX: geometry without a traversal");
        const int ntb = G->trav->ntarget_boxes;
        void *scr;
        QBX_TRY(scratch_reserve(ctx, st, sizeof(int) * (2L * ntb + 3), &scr));
        int *vl = (int *)scr;
        hipLaunchKernelGGL(k_virtual_lists, dim3(cdiv(ntb + 1, 256)), dim3(256), 0, st, ntb,
                           T->nsources, vl);
        const dim3 grid(cdiv((long)G->nctt * TSQ_LANES, 256));
        const int op = kind == QBX_SOURCE_S ? 0 : 1;
#define DQ3_ARGS                                                                          \
    qbx_order, grid, st, kk, 0.0, G->nctt, G->ctt_lists.p, G->tree_ttc.p,                \
        G->center_to_target_box.p, G->centers.p, G->ncenters, T->targets.p, T->ntargets, \
        T->sources.p, T->nsources, vl, vl + ntb + 1, vl + 2 * ntb + 1, vl + 2 * ntb + 2, \
        str, nrm, pot, nullptr
        int rc;
        if (kernel == QBX_LAPLACE3D)
            rc = op == 0 ? launch_p<0, 0>(DQ3_ARGS) : launch_p<0, 1>(DQ3_ARGS);
        else
            rc = op == 0 ? launch_p<1, 0>(DQ3_ARGS) : launch_p<1, 1>(DQ3_ARGS);
        QBX_TRY(rc);
#undef DQ3_ARGS
        constexpr double INV4PI_ = 0.07957747154594766788;
        hipLaunchKernelGGL(k_scale_c, dim3(cdiv(2L * T->ntargets, 256)), dim3(256), 0, st,
                           (long)T->ntargets, INV4PI_, pot);
    }
    if (G->nfiltered > 0)
        QBX_TRY(direct_p2p3d(st, kk, kind == QBX_SOURCE_D, G->nfiltered, G->nqbtl_unfiltered.p,
                             T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm,
                             pot));
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // namespace qbx

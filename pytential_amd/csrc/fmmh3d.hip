// QBX-FMM wrangler stages for 3D Helmholtz (pytential/qbx/fmm.py drive_fmm and
// pytential/qbx/interactions.py), complex128, G(x, y) = e^{ik|x-y|} /
// (4 pi |x-y|) with real k > 0.
//
// Expansions in spherical waves (coefficient (n, m) at n*n + n + m):
//   Y_n^m     orthonormal spherical harmonics, Condon-Shortley phase,
//             Y_n^{-m} = (-1)^m conj(Y_n^m)
//   R_n^m(x) = j_n(k|x|) Y_n^m(x/|x|),  S_n^m(x) = h_n(k|x|) Y_n^m(x/|x|)
//   multipole about c:  M_n^m = ik sum q (-1)^m R_n^{-m}(s - c),  phi = sum M S(x - c)
//   local about c:      L_n^m = ik sum q (-1)^m S_n^{-m}(s - c),  phi = sum L R(x - c)
//   translations, t = c2 - c1 (M2M and L2L with X = R, M2L / M2QBXL with S):
//     b_{nu,mu} = sum_{n,m} C^X_{nu mu, n m}(t) a_{n,m},
//     C^X_{nu mu, n m}(t) = 4 pi sum_l (-1)^((nu+l-n)/2) G(n,m; nu,mu; l,m-mu) X_l^{m-mu}(t)
//   with the Gaunt coefficients G = int Y_n^m conj(Y_nu^mu) conj(Y_l^(m-mu))
//   (l = |n-nu|, |n-nu|+2, ..., n+nu), computed on the device by
//   Gauss-Legendre quadrature of normalised Legendre functions (exact: the
//   integrands are polynomials of degree n + nu + l).
//   Double layer (q nu . grad_s): the ladder relations
//     d_z F_n^m = k (A_{n-1}^m F_{n-1}^m - A_n^m F_{n+1}^m),
//     (d_x + i d_y) F_n^m = k (a F_{n+1}^{m+1} + b F_{n-1}^{m+1}),
//     (d_x - i d_y) F_n^m = -k (c F_{n+1}^{m-1} + d F_{n-1}^{m-1})
//   (coefficients in nu_grad below), F = R or S.
//
// Storage (HBM): box expansions level-compact like the 3D Laplace stages,
// (p_l+1)^2 complex per box of level l (times nr), scaled per level by
// s_l = min(1, k h_l) (h_l the box size): multipoles hold M_n^m / s^n,
// locals L_n^m s^n, so that j_n and h_n of small arguments stay O(1) at
// every order.  QBX expansions [ncenters][(p_q+1)^2][nr] are unscaled.
//
// Box translations use dense tables, one complex (p_in+1)^2 x (p_out+1)^2
// matrix per level and offset class (8 child octants for M2M and L2L, the
// (4 nsep + 3)^3 integer offsets for M2L), built on the device from the
// Gaunt table and cached in the context per (k, tree extent, orders, nsep).
// M2QBXL and L2QBXL translate to arbitrary centers: the matrix entries are
// summed from the Gaunt table and X_l^la(t) on the fly.
//
// Thread mappings:
//  P2M, P2L, P2QBXL  a block per output expansion (box or QBX center); each
//                    chunk of sources is expanded into R or S waves in LDS
//                    (a thread per (source, m) column), then a thread per
//                    output coefficient sums the chunk
//  M2M, M2L, L2L     a block per output box, a thread per output
//                    coefficient; each entry's input expansion staged in
//                    LDS, the table read row by row (coalesced)
//  P2P, M2P, L2P     a 64-lane block per target box, lanes over targets
//  M2QBXL, L2QBXL    a 64-lane block per QBX center; per input box X(t) in
//                    LDS, lanes over (output coefficient, input slice)
//  QBXL2P            a thread per (center, target) pair
#include "qbx_internal.h"
#include "qbx_math.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace qbx {
namespace h3 {

constexpr int MAXP = 20;    // box expansion orders
constexpr int MAXPQ = 8;    // QBX expansion orders
constexpr int FML = 21;     // levels of a 3D tree
constexpr int NRC_B = 4;    // right-hand-side chunk of the box stages
constexpr int NRC_E = 8;    // right-hand-side chunk of the evaluations
constexpr int NRC_Q = 2;    // right-hand-side chunk of the QBX translations
constexpr double FOURPI = 12.566370614359172954;
constexpr double INV4PI = 0.079577471545947667884;
constexpr long TAB_BYTES_MAX = 32L << 30;  // translation tables per context

__host__ __device__ constexpr int IDX(int n, int m) { return n * n + n + m; }

struct HD {
    int dip, pq, nlevels, ncq, nsep, W, pmax, pg, nr, r0, rc;
    double k;
    int order[FML];
    int ncl[FML];
    int lbox[FML + 1];
    long lcoef[FML + 1];
    double h[FML];  // box size per level
    double s[FML];  // expansion scale per level
    const double *gaunt;  // [(pg+1)^2 (n,m)][(pg+1)^2 (nu,mu)][pg+1]
    const double *tab;    // every translation table (complex)
    long m2m_off[FML], l2l_off[FML], m2l_off[FML];  // complex offsets, -1: none
};

struct TD {
    int nboxes;
    const int *level, *parent, *child, *coords;
    const double *centers;  // [3][nboxes]
    const int *sstart, *scount, *scumul;
};

struct v3 {
    double x, y, z;
};
__device__ __forceinline__ v3 ld3(const double *p, long n, long i)
{
    return {p[i], p[n + i], p[2 * n + i]};
}
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 sub3(v3 a, v3 b) { return sub(a, b); }
__device__ __forceinline__ v3 boxc(const TD &t, int b) { return ld3(t.centers, t.nboxes, b); }

__device__ __forceinline__ long ecoef_off(const HD &hd, int b, int lv)
{
    return 2 * (hd.lcoef[lv] + (long)(b - hd.lbox[lv]) * hd.ncl[lv]) * hd.nr;
}

__device__ __forceinline__ void nm_of(int i, int &n, int &m)
{
    n = (int)sqrt((double)i);
    while (n * n > i) --n;
    while ((n + 1) * (n + 1) <= i) ++n;
    m = i - n * n - n;
}

__device__ __forceinline__ cplx sgn_conj(cplx v, int am)  // (-1)^m conj(v)
{
    return (am & 1) ? cplx{-v.x, v.y} : cplx{v.x, -v.y};
}

// ---- special functions (real argument) ------------------------------------
// j_n(x), n = 0..p (p >= 1 slots written), by Miller's backward recurrence
// from N = max(p, x) + 10 + 4 sqrt(max(p, x)) (j_N / j_p below 1e-10 there,
// so the start's error enters squared), normalised by j_0 or j_1 (the larger)
__device__ void sph_j_seq(int p, double x, double *j)
{
    const int pp = p < 1 ? 1 : p;
    if (x < 1e-300) {
        for (int n = 0; n <= pp; ++n) j[n] = n == 0 ? 1.0 : 0.0;
        return;
    }
    const double mx = fmax((double)pp, x);
    const int N = (int)mx + 10 + (int)(4.0 * sqrt(mx));
    const double ix = qbx_rcp(x);
    double fp1 = 0.0, f = 1e-300;
    for (int n = N; n >= 0; --n) {
        if (n <= pp) j[n] = f;
        const double fm1 = (2.0 * n + 1.0) * ix * f - fp1;
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e200) {
            f *= 1e-200;
            fp1 *= 1e-200;
            for (int q = n; q <= pp; ++q) j[q] *= 1e-200;
        }
    }
    double sn, cs;
    qbx_sincos(x, &sn, &cs);
    const double j0 = sn * ix, j1 = (sn * ix - cs) * ix;
    const double sc = fabs(j0) >= fabs(j1) ? j0 / j[0] : j1 / j[1];
    for (int n = 0; n <= pp; ++n) j[n] *= sc;
}

// radial sequence into rad (complex, n = 0..p, p + 1 >= 2 slots): j_n
// (sing 0, Miller's recurrence; jtmp holds p + 2 doubles) or h_n = j_n + i y_n
// (sing 1, x > 0) by the forward recurrence h_{n+1} = (2n+1)/x h_n - h_{n-1}
// from h_0 = (sin x - i cos x) / x and h_1: the recurrence is stable for the
// dominant y_n, and the error it leaves in the small j_n of x < n stays at
// the rounding of |h_n| (only h_n as a whole enters the singular expansions)
__device__ void radial_seq(int sing, int p, double x, cplx *rad, double *jtmp)
{
    if (!sing) {
        sph_j_seq(p, x, jtmp);
        for (int n = 0; n <= p; ++n) rad[n] = {jtmp[n], 0.0};
        return;
    }
    double sn, cs;
    qbx_sincos(x, &sn, &cs);
    const double ix = qbx_rcp(x);
    cplx hm2 = {sn * ix, -cs * ix};
    cplx hm1 = {(sn * ix - cs) * ix, (-cs * ix - sn) * ix};
    rad[0] = hm2;
    rad[1] = hm1;
    for (int n = 2; n <= p; ++n) {
        const cplx h = ((2.0 * n - 1.0) * ix) * hm1 - hm2;
        rad[n] = h;
        hm2 = hm1;
        hm1 = h;
    }
}

// Column m >= 0 of Y_n^m(d / |d|) for n = m..p: calls f(n, Y_n^m).
// Q_m^m = -sqrt((2m+1)/(2m)) Q_{m-1}^{m-1} (Q_0^0 = 1/sqrt(4 pi)),
// Q_{m+1}^m = sqrt(2m+3) z Q_m^m, Q_n^m = a_n (z Q_{n-1}^m - b_n Q_{n-2}^m),
// a_n = sqrt((4n^2 - 1) / (n^2 - m^2)), b_n = 1 / a_{n-1};
// Y_n^m = Q_n^m ((x + i y) / r)^m.  sqrt(A / B) = A / sqrt(A B) and its
// inverse B / sqrt(A B) from one refined reciprocal square root per step
// (no division, no library square root).
template <class F>
__device__ __forceinline__ void ylm_column(int m, int p, v3 d, F f)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double ir = r2 > 0.0 ? qbx_rsqrt(r2) : 0.0;
    const double z = r2 > 0.0 ? d.z * ir : 1.0;
    const cplx w = {d.x * ir, d.y * ir};
    double qmm = 0.28209479177387814347;  // 1 / sqrt(4 pi)
    cplx wm = {1.0, 0.0};
    for (int i = 1; i <= m; ++i) {
        const double A = 2.0 * i + 1.0, B = 2.0 * i;
        qmm *= -A * qbx_rsqrt(A * B);
        wm = wm * w;
    }
    double a2 = 0.0, a1 = qmm, binv = 0.0;
    f(m, qmm * wm);
    for (int n = m + 1; n <= p; ++n) {
        double q;
        if (n == m + 1) {
            const double rs = qbx_rsqrt(2.0 * m + 3.0);
            q = (2.0 * m + 3.0) * rs * z * qmm;
            binv = rs;
        } else {
            const double A = 4.0 * n * n - 1.0, B = (double)n * n - (double)m * m;
            const double rs = qbx_rsqrt(A * B);
            q = (A * rs) * (z * a1 - binv * a2);
            binv = B * rs;
        }
        f(n, q * wm);
        a2 = a1;
        a1 = q;
    }
}

__device__ __forceinline__ cplx getf(const cplx *F, int p, int n, int m)
{
    return (n < 0 || n > p || m < -n || m > n) ? cplx{0.0, 0.0} : F[IDX(n, m)];
}

// nu . grad F_n^m from the ladder relations; F holds orders <= pf (>= n + 1)
__device__ __forceinline__ cplx nu_grad(const cplx *F, int pf, double k, int n, int m, v3 nu)
{
    const double dn = n;
    const double An = sqrt(((dn + 1.0) * (dn + 1.0) - (double)m * m) /
                           ((2.0 * dn + 1.0) * (2.0 * dn + 3.0)));
    const double An1 = (n >= 1 && abs(m) <= n - 1)
                           ? sqrt((dn * dn - (double)m * m) / ((2.0 * dn - 1.0) * (2.0 * dn + 1.0)))
                           : 0.0;
    const cplx dz = k * (An1 * getf(F, pf, n - 1, m) - An * getf(F, pf, n + 1, m));
    const double a = sqrt((dn + m + 1.0) * (dn + m + 2.0) / ((2.0 * dn + 1.0) * (2.0 * dn + 3.0)));
    const double b = n >= 1 ? sqrt(fmax(0.0, (dn - m) * (dn - m - 1.0)) /
                                   ((2.0 * dn - 1.0) * (2.0 * dn + 1.0)))
                            : 0.0;
    const double c = sqrt((dn - m + 1.0) * (dn - m + 2.0) / ((2.0 * dn + 1.0) * (2.0 * dn + 3.0)));
    const double d = n >= 1 ? sqrt(fmax(0.0, (dn + m) * (dn + m - 1.0)) /
                                   ((2.0 * dn - 1.0) * (2.0 * dn + 1.0)))
                            : 0.0;
    const cplx dp = k * (a * getf(F, pf, n + 1, m + 1) + b * getf(F, pf, n - 1, m + 1));
    const cplx dm = -k * (c * getf(F, pf, n + 1, m - 1) + d * getf(F, pf, n - 1, m - 1));
    const cplx nm = {0.5 * nu.x, -0.5 * nu.y}, np = {0.5 * nu.x, 0.5 * nu.y};
    return nu.z * dz + nm * dp + np * dm;
}

__device__ __forceinline__ double ipow(double s, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= s;
    return r;
}

// ---- Gaunt table ----------------------------------------------------------
// g[(IDX(n,m) ncg + IDX(nu,mu)) (pg+1) + i] = G(n,m; nu,mu; |n-nu| + 2i, m-mu)
// from 2 pi sum_q w_q Pb_n^m(x_q) Pb_nu^mu(x_q) Pb_l^(m-mu)(x_q), Pb the
// normalised associated Legendre functions at the nodes ([nq][(2pg+1)^2]).
__global__ void kh_gaunt(int pg, int nq, const double *wq, const double *pb, double *g)
{
    const int ncg = (pg + 1) * (pg + 1), L = 2 * pg, ncl = (L + 1) * (L + 1);
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)ncg * ncg) return;
    const int j = (int)(e / ncg), i = (int)(e % ncg);
    int n, m, nu, mu;
    nm_of(j, n, m);
    nm_of(i, nu, mu);
    double *out = g + e * (pg + 1);
    const int imax = min(n, nu);
    for (int ii = 0; ii <= pg; ++ii) {
        double acc = 0.0;
        const int l = abs(n - nu) + 2 * ii;
        if (ii <= imax && abs(m - mu) <= l) {
            for (int q = 0; q < nq; ++q) {
                const double *P = pb + (long)q * ncl;
                acc += wq[q] * P[IDX(n, m)] * P[IDX(nu, mu)] * P[IDX(l, m - mu)];
            }
            acc *= 6.283185307179586477;
        }
        out[ii] = acc;
    }
}

// ---- wave tables ----------------------------------------------------------
// X_l^la(t), l <= L, into X (LDS, (L+1)^2 complex) by the block; rad and
// jtmp scratch of L + 3 entries; sing 0: R, 1: S
__device__ void waves_block(int sing, int L, double k, v3 t, cplx *X, cplx *rad, double *jtmp)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        const double r = sqrt(t.x * t.x + t.y * t.y + t.z * t.z);
        radial_seq(sing, L, k * r, rad, jtmp);
    }
    __syncthreads();
    for (int m = tid; m <= L; m += blockDim.x)
        ylm_column(m, L, t, [&](int n, cplx y) {
            X[IDX(n, m)] = rad[n] * y;
            if (m > 0) X[IDX(n, -m)] = rad[n] * sgn_conj(y, m);
        });
    __syncthreads();
}

// Translation tables: tab[cls][j = (n,m) < ncin][i = (nu,mu) < ncout] =
// C^X_{i,j}(t_cls) sout^nu / sin^n (the scales of the stored expansions).
__global__ void __launch_bounds__(256) kh_build_tab(int sing, double k, int pin, int pout,
                                                    double sin_, double sout,
                                                    const double *tv, int pg,
                                                    const double *gaunt, double *tab)
{
    __shared__ double Xs[2 * 41 * 41];
    __shared__ double rads[2 * 44];
    __shared__ double jt[44];
    const int cls = blockIdx.x;
    const v3 t = {tv[3 * cls], tv[3 * cls + 1], tv[3 * cls + 2]};
    const int ncin = (pin + 1) * (pin + 1), ncout = (pout + 1) * (pout + 1);
    double *T = tab + 2L * cls * ncin * ncout;
    const bool zero = t.x == 0.0 && t.y == 0.0 && t.z == 0.0;
    if (zero) {
        for (long e = threadIdx.x; e < (long)ncin * ncout; e += blockDim.x) {
            T[2 * e] = 0.0;
            T[2 * e + 1] = 0.0;
        }
        return;
    }
    cplx *X = reinterpret_cast<cplx *>(Xs);
    waves_block(sing, pin + pout, k, t, X, reinterpret_cast<cplx *>(rads), jt);
    const int ncg = (pg + 1) * (pg + 1);
    for (long e = threadIdx.x; e < (long)ncin * ncout; e += blockDim.x) {
        const int j = (int)(e / ncout), i = (int)(e % ncout);
        int n, m, nu, mu;
        nm_of(j, n, m);
        nm_of(i, nu, mu);
        const double *g = gaunt + ((long)j * ncg + i) * (pg + 1);
        cplx c = {0.0, 0.0};
        const int imax = min(n, nu);
        for (int ii = 0; ii <= imax; ++ii) {
            const int l = abs(n - nu) + 2 * ii;
            if (abs(m - mu) > l) continue;
            const int ex = (nu + l - n) / 2;
            c += ((ex & 1) ? -g[ii] : g[ii]) * X[IDX(l, m - mu)];
        }
        const double sc = FOURPI * ipow(sout, nu) / ipow(sin_, n);
        T[2 * e] = sc * c.x;
        T[2 * e + 1] = sc * c.y;
    }
}

// ---- P2M, P2L, P2QBXL -----------------------------------------------------
// MODE 0: P2M, items = source boxes, a box's own sources, out = multipoles;
// MODE 1: P2L, items = target-or-target-parent boxes, list-4 sources;
// MODE 2: P2QBXL, items = global QBX centers, list-1 sources of the
// center's target box, out = QBX expansions (order pq, unscaled).
// Dynamic LDS: CH sources x [(pf+1)^2 waves, pf+1 radial, rc strengths].
template <int MODE, int NRC>
__global__ void __launch_bounds__(512) kh_p2e(HD hd, TD t, int nitems, const int *items,
                                             const int *starts, const int *lists,
                                             const double *ctrs, int nctr, const int *c2tb,
                                             const double *src, int ns, const double *str,
                                             const double *nrm, int CH, int pfmax,
                                             double *out)
{
    extern __shared__ double sh[];
    __shared__ int gs[64];  // the chunk's source indices (CH <= 32)
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int ncf = (pfmax + 1) * (pfmax + 1);
    cplx *sF = reinterpret_cast<cplx *>(sh);                       // [CH][ncf]
    cplx *srad = sF + (long)CH * ncf;                               // [CH][pfmax+1]
    cplx *sq = srad + (long)CH * (pfmax + 1);                       // [CH][NRC]
    double *snu = reinterpret_cast<double *>(sq + (long)CH * NRC);  // [CH][3]
    double *sd = snu + 3 * CH;                                      // [CH][3]
    double *jt = sd + 3 * CH;                                       // [CH][pfmax+3]
    int B, p, e0, e1, lv = 0;
    v3 c;
    double *o;
    double sc;
    if (MODE == 2) {
        const int ci = items[it];
        const int tb = c2tb[ci];
        c = ld3(ctrs, nctr, ci);
        p = hd.pq;
        e0 = starts[tb];
        e1 = starts[tb + 1];
        B = -1;
        o = out + 2L * ci * hd.ncq * hd.nr;
        sc = 1.0;
    } else {
        B = items[it];
        lv = t.level[B];
        c = boxc(t, B);
        p = hd.order[lv];
        if (MODE == 0) {
            e0 = 0;
            e1 = 1;
        } else {
            e0 = starts[it];
            e1 = starts[it + 1];
        }
        o = out + ecoef_off(hd, B, lv);
        sc = MODE == 0 ? 1.0 / hd.s[lv] : hd.s[lv];  // M / s^n, L s^n
    }
    const int pf = p + (hd.dip ? 1 : 0);
    const int nco = (p + 1) * (p + 1);
    const int sing = MODE == 0 ? 0 : 1;
    const int tid = threadIdx.x;
    int n_o = 0, m_o = 0;
    if (tid < nco) nm_of(tid, n_o, m_o);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0.0, 0.0};
    const double k = hd.k;
    // walk the source ranges in chunks of CH
    int e = e0, s = 0, s1 = 0;
    bool have = false;
    auto next_range = [&]() {
        while (e < e1) {
            const int Sb = MODE == 0 ? B : lists[e];
            ++e;
            const int a = t.sstart[Sb], b = a + t.scount[Sb];
            if (b > a) {
                s = a;
                s1 = b;
                return true;
            }
        }
        return false;
    };
    have = next_range();
    while (have) {
        // gather up to CH sources (uniform across the block)
        __syncthreads();  // the previous chunk consumed
        int cnt = 0;
        while (have && cnt < CH) {
            const int take = min(s1 - s, CH - cnt);
            for (int q = tid; q < take; q += blockDim.x) gs[cnt + q] = s + q;
            cnt += take;
            s += take;
            if (s >= s1) have = next_range();
        }
        __syncthreads();
        for (int q = tid; q < cnt; q += blockDim.x) {
            const int si = gs[q];
            const v3 d = sub(ld3(src, ns, si), c);
            sd[3 * q] = d.x;
            sd[3 * q + 1] = d.y;
            sd[3 * q + 2] = d.z;
            if (hd.dip) {
                const v3 nu = ld3(nrm, ns, si);
                snu[3 * q] = nu.x;
                snu[3 * q + 1] = nu.y;
                snu[3 * q + 2] = nu.z;
            }
            for (int r = 0; r < NRC; ++r) {
                cplx qv = {0.0, 0.0};
                if (r < hd.rc) {
                    const double *pq_ = str + 2 * ((size_t)si * hd.nr + hd.r0 + r);
                    qv = {pq_[0], pq_[1]};
                }
                sq[q * NRC + r] = qv;
            }
            const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            radial_seq(sing, pf, k * rr, srad + (long)q * (pfmax + 1), jt + (long)q * (pfmax + 3));
        }
        __syncthreads();
        for (int w = tid; w < cnt * (pf + 1); w += blockDim.x) {
            const int q = w / (pf + 1), m = w % (pf + 1);
            const v3 d = {sd[3 * q], sd[3 * q + 1], sd[3 * q + 2]};
            cplx *F = sF + (long)q * ncf;
            const cplx *rd = srad + (long)q * (pfmax + 1);
            ylm_column(m, pf, d, [&](int n, cplx y) {
                F[IDX(n, m)] = rd[n] * y;
                if (m > 0) F[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
            });
        }
        __syncthreads();
        if (tid < nco) {
            for (int q = 0; q < cnt; ++q) {
                const cplx *F = sF + (long)q * ncf;
                cplx f;
                if (hd.dip) {
                    const v3 nu = {snu[3 * q], snu[3 * q + 1], snu[3 * q + 2]};
                    f = nu_grad(F, pf, k, n_o, -m_o, nu);
                } else {
                    f = F[IDX(n_o, -m_o)];
                }
                if (m_o & 1) f = -f;
#pragma unroll
                for (int r = 0; r < NRC; ++r) acc[r] += sq[q * NRC + r] * f;
            }
        }
    }
    if (tid < nco) {
        // ik, and the stored scale s^(-n) (multipole) or s^n (local)
        const double scn = ipow(sc, n_o);
        const cplx ik = {0.0, k * scn};
#pragma unroll
        for (int r = 0; r < NRC; ++r)
            if (r < hd.rc) {
                double *dst = o + 2 * ((long)tid * hd.nr + hd.r0 + r);
                const cplx v = ik * acc[r];
                dst[0] += v.x;
                dst[1] += v.y;
            }
    }
}

// S waves of one source about a center, in registers: F[IDX(n, m)] =
// h_n(k r) Y_n^m(d) for n <= PF, d = s - c, r = |d| > 0.  h_n by the forward
// recurrence (as radial_seq: stable for the dominant y_n, the small j_n of
// x < n accurate to the rounding of |h_n|), Y by the column recurrences, all
// unrolled.
template <int PF>
__device__ __forceinline__ void s_waves_reg(double k, v3 d, cplx *F)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double ir = qbx_rsqrt(r2);
    const double r = r2 * ir;
    const double x = k * r, ix = qbx_rcp(x);
    double sn, cs;
    qbx_sincos(x, &sn, &cs);
    double yv[PF + 2], jv[PF + 2];
    yv[0] = -cs * ix;
    yv[1] = (-cs * ix - sn) * ix;
    jv[0] = sn * ix;
    jv[1] = (sn * ix - cs) * ix;
#pragma unroll
    for (int n = 2; n <= PF; ++n) {
        yv[n] = (2.0 * n - 1.0) * ix * yv[n - 1] - yv[n - 2];
        jv[n] = (2.0 * n - 1.0) * ix * jv[n - 1] - jv[n - 2];
    }
    const double scl = 1.0;
    const double z = d.z * ir;
    const cplx w = {d.x * ir, d.y * ir};
    cplx wm = {1.0, 0.0};
    double qmm = 0.28209479177387814347;
#pragma unroll
    for (int m = 0; m <= PF; ++m) {
        if (m > 0) {
            qmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
            wm = wm * w;
        }
        double a2 = 0.0, a1 = qmm;
#pragma unroll
        for (int n = m; n <= PF; ++n) {
            double q;
            if (n == m) {
                q = qmm;
            } else if (n == m + 1) {
                q = sqrt(2.0 * m + 3.0) * z * qmm;
            } else {
                const double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
                const double b = sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) /
                                      (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
                q = a * (z * a1 - b * a2);
            }
            if (n > m) {
                a2 = a1;
                a1 = q;
            }
            const cplx hn = {scl * jv[n], yv[n]};
            const cplx y = q * wm;
            F[IDX(n, m)] = hn * y;
            if (m > 0) F[IDX(n, -m)] = hn * sgn_conj(y, m);
        }
    }
}

// P2QBXL for one density with the QBX expansion in registers (orders <= 5):
// a wavefront per global QBX center, the center's list-1 sources dealt to
// the lanes round-robin across the boxes.  Each lane forms a source's S
// waves -- y_n by the forward recurrence, j_n by Miller's backward
// recurrence from N = max(PF, x) + 10 + 4 sqrt(max(PF, x)) normalised by j_0
// or j_1, Y by the column recurrences, all unrolled -- and adds
// q (-1)^m S_n^{-m}(s - c) (or its nu-gradient) into its registers; the
// lanes' sums are reduced by shuffles and lane 0 adds ik times them.
template <int PQ, int DIP>
__global__ void __launch_bounds__(64) kh_p2qbxl_reg(HD hd, TD t, int ng, const int *gcent,
                                                    const int *c2tb, const int *starts,
                                                    const int *lists, const double *ctrs,
                                                    int nctr, const double *src, int ns,
                                                    const double *str, const double *nrm,
                                                    double *qexp)
{
    constexpr int PF = PQ + DIP;
    constexpr int NC = (PQ + 1) * (PQ + 1);
    constexpr int NF = (PF + 1) * (PF + 1);
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int ci = gcent[g];
    const int tb = c2tb[ci];
    const v3 c = ld3(ctrs, nctr, ci);
    const int lane = threadIdx.x;
    const double k = hd.k;
    cplx acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) acc[i] = {0.0, 0.0};
    int base = 0;
    for (int e = starts[tb]; e < starts[tb + 1]; ++e) {
        const int Sb = lists[e];
        const int s0 = t.sstart[Sb], cnt = t.scount[Sb];
        for (int j = (lane - (base & 63) + 64) & 63; j < cnt; j += 64) {
            const int si = s0 + j;
            const v3 d = sub(ld3(src, ns, si), c);
            cplx F[NF];
            s_waves_reg<PF>(k, d, F);
            const cplx qs = ldc(str + 2 * ((size_t)si * hd.nr + hd.r0));
            v3 nu = {0.0, 0.0, 0.0};
            if (DIP) nu = ld3(nrm, ns, si);
#pragma unroll
            for (int n = 0; n <= PQ; ++n)
#pragma unroll
                for (int m = -n; m <= n; ++m) {
                    cplx fv = DIP ? nu_grad(F, PF, k, n, -m, nu) : F[IDX(n, -m)];
                    if (m & 1) fv = -fv;
                    acc[IDX(n, m)] += qs * fv;
                }
        }
        base += cnt;
    }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            acc[i].x += __shfl_xor(acc[i].x, off);
            acc[i].y += __shfl_xor(acc[i].y, off);
        }
    if (lane == 0) {
        const cplx ik = {0.0, k};
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            double *dst = qexp + 2 * (((size_t)ci * NC + i) * hd.nr + hd.r0);
            const cplx v = ik * acc[i];
            dst[0] += v.x;
            dst[1] += v.y;
        }
    }
}

// ---- M2M, M2L, L2L ---------------------------------------------------------
// OP 0 M2M: boxes = source parents of level lv, entries = their children
//           with sources (level lv + 1), table by child octant
// OP 1 M2L: boxes = target-or-target-parent boxes [i0, i0 + nbox) of level
//           lv, entries = list 2, table by integer offset class
// OP 2 L2L: boxes = target-or-target-parent boxes of level lv >= 1, entry =
//           the parent (level lv - 1), table by the box's octant
// A thread per output coefficient; dynamic LDS: the entry's input
// (ncin x NRC complex).
template <int OP, int NRC>
__global__ void __launch_bounds__(448) kh_trans(HD hd, TD t, int lv, int i0, int nbox,
                                                const int *boxes, const int *starts,
                                                const int *lists, const double *in, double *out)
{
    extern __shared__ double sh[];
    cplx *a = reinterpret_cast<cplx *>(sh);
    const int ib = blockIdx.x;
    if (ib >= nbox) return;
    const int i = i0 + ib;
    const int B = boxes[i];
    const int lvin = OP == 0 ? lv + 1 : OP == 2 ? lv - 1 : lv;
    const int pin = hd.order[lvin], pout = hd.order[lv];
    const int ncin = (pin + 1) * (pin + 1), ncout = (pout + 1) * (pout + 1);
    const int tid = threadIdx.x;
    const int S1 = 2 * hd.W + 1;
    const int nb = t.nboxes;
    const int bx = t.coords[B], by = t.coords[nb + B], bz = t.coords[2 * nb + B];
    cplx acc[NRC];
#pragma unroll
    for (int r = 0; r < NRC; ++r) acc[r] = {0.0, 0.0};
    int ne;
    if (OP == 0) ne = 8;
    else if (OP == 1) ne = starts[i + 1] - starts[i];
    else ne = 1;
    for (int e = 0; e < ne; ++e) {
        int A, cls;
        if (OP == 0) {
            A = t.child[(long)e * nb + B];
            if (A < 0 || t.scumul[A] == 0) continue;  // uniform over the block
            cls = (t.coords[A] & 1) | ((t.coords[nb + A] & 1) << 1) |
                  ((t.coords[2 * nb + A] & 1) << 2);
        } else if (OP == 1) {
            A = lists[starts[i] + e];
            const int dx = bx - t.coords[A] + hd.W, dy = by - t.coords[nb + A] + hd.W,
                      dz = bz - t.coords[2 * nb + A] + hd.W;
            if ((unsigned)dx >= (unsigned)S1 || (unsigned)dy >= (unsigned)S1 ||
                (unsigned)dz >= (unsigned)S1)
                continue;
            cls = (dx * S1 + dy) * S1 + dz;
        } else {
            A = t.parent[B];
            if (A < 0) continue;
            cls = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
        }
        const long toff = OP == 0 ? hd.m2m_off[lv] : OP == 1 ? hd.m2l_off[lv] : hd.l2l_off[lv];
        const double *T = hd.tab + 2 * (toff + (long)cls * ncin * ncout);
        const double *src = in + ecoef_off(hd, A, lvin);
        __syncthreads();
        for (int w = tid; w < ncin * NRC; w += blockDim.x) {
            const int j = w / NRC, r = w % NRC;
            a[w] = r < hd.rc ? ldc(src + 2 * ((long)j * hd.nr + hd.r0 + r)) : cplx{0.0, 0.0};
        }
        __syncthreads();
        if (tid < ncout) {
            for (int j = 0; j < ncin; ++j) {
                const cplx tv = ldc(T + 2 * ((long)j * ncout + tid));
#pragma unroll
                for (int r = 0; r < NRC; ++r) acc[r] += tv * a[j * NRC + r];
            }
        }
    }
    if (tid < ncout) {
        double *o = out + ecoef_off(hd, B, lv);
#pragma unroll
        for (int r = 0; r < NRC; ++r)
            if (r < hd.rc) {
                double *dst = o + 2 * ((long)tid * hd.nr + hd.r0 + r);
                dst[0] += acc[r].x;
                dst[1] += acc[r].y;
            }
    }
}

// ---- P2P, M2P, L2P ---------------------------------------------------------
__device__ __forceinline__ cplx p2p_val(double k, bool dip, v3 s, v3 nu, v3 x)
{
    const v3 d = sub(x, s);
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 == 0.0) return {0.0, 0.0};
    const double ir = qbx_rsqrt(r2), r = r2 * ir;
    double sn, cs;
    qbx_sincos(k * r, &sn, &cs);
    const cplx e = {cs * ir * INV4PI, sn * ir * INV4PI};  // e^{ikr} / (4 pi r)
    if (!dip) return e;
    const double pr = (nu.x * d.x + nu.y * d.y + nu.z * d.z) * ir * ir;
    return (e * cplx{1.0, -k * r}) * pr;
}

// sum over (n, m) of coef * rad[n] Y_n^m(d) (rad already scaled)
template <int NRC>
__device__ void sum_waves(const HD &hd, const double *coef, int p, v3 d, const cplx *rad,
                          cplx *acc)
{
    for (int m = 0; m <= p; ++m)
        ylm_column(m, p, d, [&](int n, cplx y) {
            const cplx fp = rad[n] * y;
            const cplx fm = rad[n] * sgn_conj(y, m);
#pragma unroll
            for (int r = 0; r < NRC; ++r)
                if (r < hd.rc) {
                    acc[r] += ldc(coef + 2 * ((long)IDX(n, m) * hd.nr + hd.r0 + r)) * fp;
                    if (m > 0)
                        acc[r] += ldc(coef + 2 * ((long)IDX(n, -m) * hd.nr + hd.r0 + r)) * fm;
                }
        });
}

template <int MODE, int NRC>
__global__ void __launch_bounds__(64) kh_eval(HD hd, TD t, int ntb, const int *tboxes,
                                              const int *starts, const int *lists,
                                              const int *fstart, const int *fcount,
                                              const int *unfilt, const double *tgt, int nt,
                                              const double *src, int ns, const double *str,
                                              const double *nrm, const double *expn,
                                              double *pot)
{
    __shared__ double rads[64 * 2 * (MAXP + 1)];
    __shared__ double jts[64 * (MAXP + 3)];
    const int itb = blockIdx.x;
    if (itb >= ntb) return;
    const int B = tboxes[itb];
    const int f0 = fstart[B], nf = fcount[B];
    if (nf == 0) return;
    const int lane = threadIdx.x;
    cplx *rad = reinterpret_cast<cplx *>(rads) + lane * (MAXP + 1);
    double *jt = jts + lane * (MAXP + 3);
    const double k = hd.k;
    for (int f = f0 + lane; f < f0 + nf; f += 64) {
        const int ti = unfilt[f];
        const v3 x = ld3(tgt, nt, ti);
        cplx acc[NRC];
#pragma unroll
        for (int r = 0; r < NRC; ++r) acc[r] = {0.0, 0.0};
        if (MODE == 0) {
            for (int e = starts[itb]; e < starts[itb + 1]; ++e) {
                const int Sb = lists[e];
                const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
                for (int s = s0; s < s1; ++s) {
                    const v3 nu = hd.dip ? ld3(nrm, ns, s) : v3{0, 0, 0};
                    const cplx g = p2p_val(k, hd.dip != 0, ld3(src, ns, s), nu, x);
#pragma unroll
                    for (int r = 0; r < NRC; ++r)
                        if (r < hd.rc)
                            acc[r] += g * ldc(str + 2 * ((size_t)s * hd.nr + hd.r0 + r));
                }
            }
        } else {
            const int eb = MODE == 1 ? starts[itb] : 0;
            const int ee = MODE == 1 ? starts[itb + 1] : 1;
            for (int e = eb; e < ee; ++e) {
                const int D = MODE == 1 ? lists[e] : B;
                const int lv = t.level[D];
                const int p = hd.order[lv];
                const v3 d = sub(x, boxc(t, D));
                const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                radial_seq(MODE == 1 ? 1 : 0, p, k * rr, rad, jt);
                // stored M / s^n: h_n s^n; stored L s^n: j_n / s^n
                const double s = MODE == 1 ? hd.s[lv] : 1.0 / hd.s[lv];
                double sp = 1.0;
                for (int n = 0; n <= p; ++n) {
                    rad[n] = sp * rad[n];
                    sp *= s;
                }
                sum_waves<NRC>(hd, expn + ecoef_off(hd, D, lv), p, d, rad, acc);
            }
        }
#pragma unroll
        for (int r = 0; r < NRC; ++r)
            if (r < hd.rc) {
                double *dst = pot + 2 * ((size_t)ti * hd.nr + hd.r0 + r);
                dst[0] += acc[r].x;
                dst[1] += acc[r].y;
            }
    }
}

// ---- M2L grouped by translation class, on the FP64 matrix cores ----------
// One density.  The list-2 pairs of the levels with M2L tables, sorted by
// (level, offset class), in items of 16 pairs of one class: per item the
// complex product Y[16 pairs][nc] = A[16][nc] T_cls[nc][nc] (A the pairs'
// source multipoles M~, T the scaled table, so a row of Y is the pair's L~
// contribution), tiled for v_mfma_f64_16x16x4f64 (lane i holds A[i%16][k0 +
// i/16] and B[k0 + i/16][i%16]; D register j of lane i is D[4j + i/16][i%16];
// the complex product is four real MFMAs, A_im negated for the real part).
// The item's 16 source rows are staged in LDS; the block's four waves take
// every fourth 16-column tile of T, so each item reads its table once.  The
// rows go to per-pair scratch; kh_m2l_cls_sum adds each target box's rows in
// list order (deterministic).
typedef double h3_d4 __attribute__((ext_vector_type(4)));
constexpr int H3G_NP = 16;  // pairs per item

__global__ void __launch_bounds__(256) kh_m2l_cls(HD hd, TD t, int nitems, const int *items,
                                                  const int *pe, const int *l2_lists,
                                                  const int *yoff, const double *mpole,
                                                  double *Y)
{
    extern __shared__ double sh[];
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int lv = items[4 * it], cls = items[4 * it + 1], a0 = items[4 * it + 2],
              cnt = items[4 * it + 3];
    const int nc = hd.ncl[lv];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, col = lane & 15,
              kr = lane >> 4;
    cplx *As = reinterpret_cast<cplx *>(sh);
    for (int w = tid; w < H3G_NP * nc; w += blockDim.x) {
        const int q = w / nc, j = w - q * nc;
        cplx v = {0.0, 0.0};
        if (q < cnt) v = ldc(mpole + ecoef_off(hd, l2_lists[pe[a0 + q]], lv) + 2L * j);
        As[w] = v;
    }
    __syncthreads();
    const double *T = hd.tab + 2 * (hd.m2l_off[lv] + (long)cls * nc * nc);
    const int KS = (nc + 3) / 4, OT = (nc + 15) / 16;
    const int nw = blockDim.x >> 6;
    for (int ot = wv; ot < OT; ot += nw) {
        const int o = ot * 16 + col;
        h3_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + kr;
            double are = 0.0, aim = 0.0, bre = 0.0, bim = 0.0;
            if (k < nc) {
                const cplx a = As[col * nc + k];
                are = a.x;
                aim = a.y;
                if (o < nc) {
                    const cplx b = ldc(T + 2 * ((long)k * nc + o));
                    bre = b.x;
                    bim = b.y;
                }
            }
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = 4 * j + kr;
            if (q < cnt && o < nc)
                stc(Y + 2 * ((long)yoff[a0 + q] + o), cplx{cre[j], cim[j]});
        }
    }
}

// the grouped rows of each target-or-target-parent box's list-2 pairs added
// into its local, in list order (levels without M2L tables are skipped)
__global__ void __launch_bounds__(128) kh_m2l_cls_sum(HD hd, TD t, int nttp, const int *ttp,
                                                      const int *starts, const int *pos,
                                                      const int *yoff, const double *Y,
                                                      double *local)
{
    const int i = blockIdx.x;
    if (i >= nttp) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int B = ttp[i];
    const int lv = t.level[B];
    if (hd.m2l_off[lv] < 0) return;
    const int nc = hd.ncl[lv];
    double *o = local + ecoef_off(hd, B, lv);
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        cplx acc = {0.0, 0.0};
        for (int e = e0; e < e1; ++e) {
            const int j = pos[e];
            if (j >= 0) acc += ldc(Y + 2 * ((long)yoff[j] + c));
        }
        stc(o + 2L * c, ldc(o + 2L * c) + acc);
    }
}

// P2P (MODE 0, list 1) and M2P (MODE 1, list 3) over balanced work units
// (unit_plan: a 64-lane wavefront per unit, the unit's targets take G lanes
// apiece, each lane walks every G-th source / list-3 entry) with the G
// partial sums reduced by shuffles in a fixed order.  Same sums as kh_eval,
// for boxes whose lists are long next to their few targets (far targets in
// coarse leaves).
template <int MODE, int NRC>
__global__ void __launch_bounds__(64) kh_eval_u(HD hd, TD t, const int *ubox, const int *ut0,
                                                const int *box_g, const int *tboxes,
                                                const int *starts, const int *lists,
                                                const int *fstart, const int *fcount,
                                                const int *unfilt, const double *tgt, int nt,
                                                const double *src, int ns, const double *str,
                                                const double *nrm, const double *expn,
                                                double *pot)
{
    __shared__ double rads[MODE == 1 ? 64 * 2 * (MAXP + 1) : 1];
    __shared__ double jts[MODE == 1 ? 64 * (MAXP + 3) : 1];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int B = tboxes[it];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int tl = ut0[u] + tsub;
    const bool active = tl < fcount[B];
    const double k = hd.k;
    cplx acc[NRC];
#pragma unroll
    for (int r = 0; r < NRC; ++r) acc[r] = {0.0, 0.0};
    int ti = 0;
    if (active) {
        ti = unfilt[fstart[B] + tl];
        const v3 x = ld3(tgt, nt, ti);
        if (MODE == 0) {
            for (int e = starts[it]; e < starts[it + 1]; ++e) {
                const int Sb = lists[e];
                const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
                for (int s = s0 + sub; s < s1; s += G) {
                    const v3 nu = hd.dip ? ld3(nrm, ns, s) : v3{0, 0, 0};
                    const cplx g = p2p_val(k, hd.dip != 0, ld3(src, ns, s), nu, x);
#pragma unroll
                    for (int r = 0; r < NRC; ++r)
                        if (r < hd.rc)
                            acc[r] += g * ldc(str + 2 * ((size_t)s * hd.nr + hd.r0 + r));
                }
            }
        } else {
            cplx *rad = reinterpret_cast<cplx *>(rads) + lane * (MAXP + 1);
            double *jt = jts + lane * (MAXP + 3);
            for (int e = starts[it] + sub; e < starts[it + 1]; e += G) {
                const int D = lists[e];
                const int lv = t.level[D];
                const int p = hd.order[lv];
                const v3 d = sub3(x, boxc(t, D));
                const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                radial_seq(1, p, k * rr, rad, jt);
                double sp = 1.0;
                for (int n = 0; n <= p; ++n) {
                    rad[n] = sp * rad[n];
                    sp *= hd.s[lv];
                }
                sum_waves<NRC>(hd, expn + ecoef_off(hd, D, lv), p, d, rad, acc);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NRC; ++r)
        for (int off = G >> 1; off >= 1; off >>= 1) {
            acc[r].x += __shfl_xor(acc[r].x, off);
            acc[r].y += __shfl_xor(acc[r].y, off);
        }
    if (active && sub == 0) {
#pragma unroll
        for (int r = 0; r < NRC; ++r)
            if (r < hd.rc) {
                double *dst = pot + 2 * ((size_t)ti * hd.nr + hd.r0 + r);
                dst[0] += acc[r].x;
                dst[1] += acc[r].y;
            }
    }
}

// ---- M2QBXL, L2QBXL --------------------------------------------------------
// MODE 0: M2QBXL over list 3 of the center's target box (C^S);
// MODE 1: L2QBXL from the target box's own local (C^R).
template <int MODE, int NRC>
__global__ void __launch_bounds__(64) kh_e2qbxl(HD hd, TD t, int ng, const int *gcent,
                                                const int *c2tb, const int *tboxes,
                                                const int *starts, const int *lists,
                                                const double *ctrs, int nctr,
                                                const double *expn, double *qexp)
{
    // LDS sized for this launch's orders (e2qbxl_lds): X (L+1)^2, the box
    // coefficients (p+1)^2 NRC, the (output, slice) sums 96 NRC and the
    // radial sequence, L = max p + p_q -- a few KB, so that many blocks
    // share a CU while one lane computes a radial sequence
    extern __shared__ double sh[];
    const int Lm = hd.pmax + hd.pq;
    cplx *X = reinterpret_cast<cplx *>(sh);
    cplx *a = X + (Lm + 1) * (Lm + 1);
    cplx *sacc = a + (hd.pmax + 1) * (hd.pmax + 1) * NRC;
    double *rads = reinterpret_cast<double *>(sacc + 96 * NRC);  // Lm + 3 complex
    double *jt = rads + 2 * (Lm + 3);                             // Lm + 4
    const int g = blockIdx.x;
    if (g >= ng) return;
    const int ci = gcent[g];
    const int itb = c2tb[ci];
    const v3 c = ld3(ctrs, nctr, ci);
    const int tid = threadIdx.x;
    const int pq = hd.pq, nco = hd.ncq;
    const int parts = nco >= 64 ? 1 : 64 / nco;
    const int ncg = (hd.pg + 1) * (hd.pg + 1);
    for (int w = tid; w < 96 * NRC; w += 64) sacc[w] = {0.0, 0.0};
    const int eb = MODE == 0 ? starts[itb] : 0;
    const int ee = MODE == 0 ? starts[itb + 1] : 1;
    for (int e = eb; e < ee; ++e) {
        const int D = MODE == 0 ? lists[e] : tboxes[itb];
        const int lv = t.level[D];
        const int p = hd.order[lv];
        const int ncin = (p + 1) * (p + 1);
        const v3 tv = sub(c, boxc(t, D));
        __syncthreads();  // previous entry's X and a consumed
        waves_block(MODE == 0 ? 1 : 0, p + pq, hd.k, tv, X, reinterpret_cast<cplx *>(rads), jt);
        // the box's coefficients, unscaled (M = M~ s^n, L = L~ / s^n)
        const double s = MODE == 0 ? hd.s[lv] : 1.0 / hd.s[lv];
        const double *src = expn + ecoef_off(hd, D, lv);
        for (int w = tid; w < ncin * NRC; w += 64) {
            const int j = w / NRC, r = w % NRC;
            int n, m;
            nm_of(j, n, m);
            a[w] = r < hd.rc ? ipow(s, n) * ldc(src + 2 * ((long)j * hd.nr + hd.r0 + r))
                             : cplx{0.0, 0.0};
        }
        __syncthreads();
        for (int w = tid; w < nco * parts; w += 64) {
            const int o = w % nco, part = w / nco;
            int nu, mu;
            nm_of(o, nu, mu);
            cplx acc[NRC];
#pragma unroll
            for (int r = 0; r < NRC; ++r) acc[r] = {0.0, 0.0};
            for (int j = part; j < ncin; j += parts) {
                int n, m;
                nm_of(j, n, m);
                const double *gg = hd.gaunt + ((long)j * ncg + o) * (hd.pg + 1);
                cplx cf = {0.0, 0.0};
                const int imax = min(n, nu);
                for (int ii = 0; ii <= imax; ++ii) {
                    const int l = abs(n - nu) + 2 * ii;
                    if (abs(m - mu) > l) continue;
                    const int ex = (nu + l - n) / 2;
                    cf += ((ex & 1) ? -gg[ii] : gg[ii]) * X[IDX(l, m - mu)];
                }
#pragma unroll
                for (int r = 0; r < NRC; ++r) acc[r] += cf * a[j * NRC + r];
            }
#pragma unroll
            for (int r = 0; r < NRC; ++r) sacc[w * NRC + r] += acc[r];
        }
    }
    __syncthreads();
    for (int w = tid; w < nco * NRC; w += 64) {
        const int o = w / NRC, r = w % NRC;
        if (r >= hd.rc) continue;
        cplx v = {0.0, 0.0};
        for (int part = 0; part < parts; ++part) v += sacc[(part * nco + o) * NRC + r];
        double *dst = qexp + 2 * (((size_t)ci * nco + o) * hd.nr + hd.r0 + r);
        dst[0] += FOURPI * v.x;
        dst[1] += FOURPI * v.y;
    }
}

// shares (blockIdx.y) of the list-1 sources / list-3 boxes per center group
// in M2QBXL, P2QBXL^T and M2QBXL^T
constexpr int ADJ_SHARES = 4;

__device__ __forceinline__ void atomic_addc(double *p, cplx v)
{
    atomicAdd(p, v.x);
    atomicAdd(p + 1, v.y);
}

// ---- M2QBXL for one density, p_q <= 3, as a product on the matrix cores -----
// The M2QBXL sum of kh_e2qbxl regrouped by the waves' index q = (l, m'):
//   Q_c[o] = sum_D sum_q X_q(c - c_D) Bm_D[q][o],
//   Bm_D[q = (l, m')][o = (nu, mu)] = 4 pi sum_{j = (n, m' + mu)}
//       (-1)^((nu + l - n)/2) G(j, o, (l - |n - nu|)/2) s^n M~_D[j]
// (the terms with l = |n - nu| + 2 ii, ii <= min(n, nu)).  Bm depends on the
// box alone (kh_m2qb_mat, once per box and evaluation); the sum over q is,
// per group of centers and list-3 box, a product of the centers' waves
// X[centers][q] with Bm_D[q][16] on v_mfma_f64_16x16x4f64.
__global__ void __launch_bounds__(256) kh_m2qb_mat(HD hd, TD t, int nqmax, const double *mpole,
                                                   double *bm)
{
    __shared__ double as[2 * (MAXP + 1) * (MAXP + 1)];
    const int D = blockIdx.x;
    if (D >= t.nboxes) return;
    if (t.scumul[D] == 0) {  // no multipole: a zero matrix
        double *z = bm + 2L * D * nqmax * 16;
        for (int w = threadIdx.x; w < 2 * nqmax * 16; w += blockDim.x) z[w] = 0.0;
        return;
    }
    const int lv = t.level[D];
    const int p = hd.order[lv];
    const int ncin = (p + 1) * (p + 1);
    const int pq = hd.pq, nco = hd.ncq;
    const int L = p + pq, nq = (L + 1) * (L + 1);
    const int ncg = (hd.pg + 1) * (hd.pg + 1);
    cplx *a = reinterpret_cast<cplx *>(as);
    const double s = hd.s[lv];
    const double *src = mpole + ecoef_off(hd, D, lv);
    for (int j = threadIdx.x; j < ncin; j += blockDim.x) {
        int n, m;
        nm_of(j, n, m);
        a[j] = ipow(s, n) * ldc(src + 2L * j);
    }
    __syncthreads();
    cplx *out = reinterpret_cast<cplx *>(bm) + (long)D * nqmax * 16;
    for (int w = threadIdx.x; w < nqmax * 16; w += blockDim.x) {
        const int q = w >> 4, o = w & 15;
        cplx acc = {0.0, 0.0};
        if (q < nq && o < nco) {
            int l, mq, nu, mu;
            nm_of(q, l, mq);
            nm_of(o, nu, mu);
            const int m = mq + mu;
            const int n0 = max(max(abs(m), l - nu), nu - l), n1 = min(p, l + nu);
            for (int n = n0; n <= n1; ++n) {
                if ((n + nu + l) & 1) continue;
                const int ii = (l - abs(n - nu)) >> 1;
                if (ii > min(n, nu)) continue;
                const int j = IDX(n, m);
                const double g = hd.gaunt[((long)j * ncg + o) * (hd.pg + 1) + ii];
                const int ex = (nu + l - n) / 2;
                acc += ((ex & 1) ? -g : g) * a[j];
            }
            acc = FOURPI * acc;
        }
        out[w] = acc;
    }
}

// a block (four waves) per (group of up to GC_GROUP global centers of one
// target box, share blockIdx.y of its list-3 boxes); per list-3 box D the
// radial sequences (a lane per center) and the Y columns (a lane per
// (center, m)) give the group's waves X[centers][nq] in LDS (rows past the
// group unused), the waves split the K = nq steps of the 16 x 16 product with
// Bm_D read from HBM, their partial tiles are added in LDS in wave order and
// the shares' sums added into the QBX expansions with atomics
__global__ void __launch_bounds__(256) kh_m2qbxl_mm(HD hd, TD t, int nunits, const int *ubox,
                                                    const int *uc0, const int *gcs,
                                                    const int *gcl, const int *l3s,
                                                    const int *l3l, const double *ctrs, int nctr,
                                                    int nqmax, const double *bm, double *qexp)
{
    extern __shared__ double sh[];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int itb = ubox[u];
    const int c0 = gcs[itb] + uc0[u], nct = min(GC_GROUP, gcs[itb + 1] - c0);
    const int e0 = l3s[itb], e1 = l3s[itb + 1];
    if (nct <= 0 || e0 + (int)blockIdx.y >= e1) return;
    const int Lm = hd.pmax + hd.pq;
    const int nco = hd.ncq;
    h3_d4 *red = reinterpret_cast<h3_d4 *>(sh);                    // [3][2][64]
    cplx *X = reinterpret_cast<cplx *>(red + 3 * 2 * 64);           // [GC_GROUP][nqmax]
    cplx *rad = X + GC_GROUP * nqmax;                               // [GC_GROUP][Lm + 3]
    double *jt = reinterpret_cast<double *>(rad + GC_GROUP * (Lm + 3));  // [GC_GROUP][Lm + 4]
    __shared__ double cxs[GC_GROUP * 3];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, col = lane & 15,
              kr = lane >> 4;
    if (tid < GC_GROUP * 3) {
        const int q = tid / 3, dd = tid % 3;
        cxs[tid] = q < nct ? ctrs[(size_t)dd * nctr + gcl[c0 + q]] : 0.0;
    }
    h3_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
    for (int e = e0 + blockIdx.y; e < e1; e += gridDim.y) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int L = hd.order[lv] + hd.pq, nq = (L + 1) * (L + 1);
        const v3 cd = boxc(t, D);
        __syncthreads();  // the previous entry's X consumed, cxs written
        if (tid < nct) {
            const v3 dv = sub(v3{cxs[3 * tid], cxs[3 * tid + 1], cxs[3 * tid + 2]}, cd);
            const double rr = sqrt(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
            radial_seq(1, L, hd.k * rr, rad + tid * (Lm + 3), jt + tid * (Lm + 4));
        }
        __syncthreads();
        for (int w = tid; w < nct * (L + 1); w += blockDim.x) {
            const int q = w / (L + 1), m = w % (L + 1);
            const v3 dv = sub(v3{cxs[3 * q], cxs[3 * q + 1], cxs[3 * q + 2]}, cd);
            const cplx *rd = rad + q * (Lm + 3);
            cplx *Xq = X + (long)q * nqmax;
            ylm_column(m, L, dv, [&](int n, cplx y) {
                Xq[IDX(n, m)] = rd[n] * y;
                if (m > 0) Xq[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
            });
        }
        __syncthreads();
        const double *B = bm + 2L * (long)D * nqmax * 16;
        const int KS = (nq + 3) / 4;
        for (int ks = wv; ks < KS; ks += 4) {
            const int k = ks * 4 + kr;
            double are = 0.0, aim = 0.0, bre = 0.0, bim = 0.0;
            if (k < nq) {
                if (col < nct) {
                    const cplx a = X[(long)col * nqmax + k];
                    are = a.x;
                    aim = a.y;
                }
                const cplx b = ldc(B + 2L * (k * 16 + col));
                bre = b.x;
                bim = b.y;
            }
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
        }
    }
    // the four waves' partial tiles: D register j of lane i is the entry
    // (center 4j + i/16, coefficient i%16)
    if (wv > 0) {
        red[((wv - 1) * 2 + 0) * 64 + lane] = cre;
        red[((wv - 1) * 2 + 1) * 64 + lane] = cim;
    }
    __syncthreads();
    if (wv == 0) {
        for (int h = 0; h < 3; ++h) {
            cre += red[(h * 2 + 0) * 64 + lane];
            cim += red[(h * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = 4 * j + kr;
            if (q < nct && col < nco)
                atomic_addc(qexp + 2 * ((size_t)gcl[c0 + q] * nco + col), cplx{cre[j], cim[j]});
        }
    }
}

// ---- QBXL2P ---------------------------------------------------------------
template <int NRC>
__global__ void kh_qbxl2p(HD hd, int nctt, const int *ctt_lists, const int *tree_ttc,
                          const double *tgt, int nt, const double *ctr, int nc,
                          const double *qexp, double *pot)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nctt) return;
    const int ti = ctt_lists[j];
    const int c = tree_ttc[ti];
    const v3 d = sub(ld3(tgt, nt, ti), ld3(ctr, nc, c));
    cplx rad[MAXPQ + 1];
    double jt[MAXPQ + 3];
    const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    sph_j_seq(hd.pq, hd.k * rr, jt);
    for (int n = 0; n <= hd.pq; ++n) rad[n] = {jt[n], 0.0};
    cplx acc[NRC];
#pragma unroll
    for (int r = 0; r < NRC; ++r) acc[r] = {0.0, 0.0};
    sum_waves<NRC>(hd, qexp + 2 * (size_t)c * hd.ncq * hd.nr, hd.pq, d, rad, acc);
#pragma unroll
    for (int r = 0; r < NRC; ++r)
        if (r < hd.rc) stc(pot + 2 * ((size_t)ti * hd.nr + hd.r0 + r), acc[r]);
}

// ---- adjoint (conjugate transpose) of the point-FMM stages, single layer --
// With the scaled storage of the forward (M~ = M / s^n, L~ = L s^n):
//   P2M^T  q'_s += -ik sum_{n,m} R_n^m(s - c) s^-n M~'[n,m]
//   P2L^T  q'_s += -ik sum_B sum conj(h_n) Y_n^m(s - c_B) s^n L~'_B[n,m]
//   L2P^T  L~'[n,m] += sum_t (-1)^m R_n^{-m}(t - c) s^-n pot'_t
//   M2P^T  M~'_D[n,m] += sum_t conj(h_n) (-1)^m Y_n^{-m}(t - c_D) s^n pot'_t
//   M2M^T, M2L^T, L2L^T  in' += T^H out' with the forward's tables
//   P2P^T  q'_s += sum_t conj(G(t, s)) pot'_t
// Scatters (list 2, 3, 4 from the target side, L2L^T into parents, P2P^T)
// add with atomics; the rest have one writer.

// P2M^T (LOCAL false: the box's own sources, coefficients M~') and P2L^T
// (LOCAL true: the sources of the box's list-4 boxes, coefficients L~'):
// a 64-lane block per box, lanes over the points
template <bool LOCAL>
__global__ void __launch_bounds__(64) kh_adj_p2e(HD hd, TD t, int nitems, const int *items,
                                                 const int *starts, const int *lists,
                                                 const double *src, int ns,
                                                 const double *expn, double *sadj)
{
    __shared__ double rads[64 * 2 * (MAXP + 1)];
    __shared__ double jts[64 * (MAXP + 3)];
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int B = items[it];
    const int lv = t.level[B];
    const int p = hd.order[lv];
    const v3 c = boxc(t, B);
    const int lane = threadIdx.x;
    cplx *rad = reinterpret_cast<cplx *>(rads) + lane * (MAXP + 1);
    double *jt = jts + lane * (MAXP + 3);
    const double *coef = expn + ecoef_off(hd, B, lv);
    const double s = LOCAL ? hd.s[lv] : 1.0 / hd.s[lv];
    const int e0 = LOCAL ? starts[it] : 0, e1 = LOCAL ? starts[it + 1] : 1;
    for (int e = e0; e < e1; ++e) {
        const int Sb = LOCAL ? lists[e] : B;
        const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
        for (int si = s0 + lane; si < s1; si += 64) {
            const v3 d = sub(ld3(src, ns, si), c);
            const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            radial_seq(LOCAL ? 1 : 0, p, hd.k * rr, rad, jt);
            double sp = 1.0;
            for (int n = 0; n <= p; ++n) {
                rad[n] = sp * (LOCAL ? conj(rad[n]) : rad[n]);
                sp *= s;
            }
            cplx acc[1] = {{0.0, 0.0}};
            sum_waves<1>(hd, coef, p, d, rad, acc);
            const cplx v = cplx{0.0, -hd.k} * acc[0];
            if (LOCAL)
                atomic_addc(sadj + 2 * (size_t)si, v);
            else
                stc(sadj + 2 * (size_t)si, ldc(sadj + 2 * (size_t)si) + v);
        }
    }
}

// L2P^T (LOCAL: the box's own local; M2P^T uses kh_adj_m2p, a block per
// list-3 entry): a block per target box, a thread per coefficient of the
// output expansion; the waves about the center at the box's non-QBX targets
// staged in LDS in chunks of 8 targets
template <bool LOCAL>
__global__ void __launch_bounds__(448) kh_adj_e2p(HD hd, TD t, int ntb, const int *tboxes,
                                                  const int *starts, const int *lists,
                                                  const int *fstart, const int *fcount,
                                                  const int *unfilt, const double *tgt,
                                                  int nt, const double *pot, double *xadj)
{
    constexpr int TC = 8;
    __shared__ double Fs[2 * TC * (MAXP + 1) * (MAXP + 1)];
    __shared__ double rads[2 * TC * (MAXP + 1)];
    __shared__ double jts[TC * (MAXP + 3)];
    __shared__ double ds[3 * TC];
    __shared__ double gs[2 * TC];
    const int itb = blockIdx.x;
    if (itb >= ntb) return;
    const int B = tboxes[itb];
    const int f0 = fstart[B], nf = fcount[B];
    if (nf == 0) return;
    const int e0 = LOCAL ? 0 : starts[itb], e1 = LOCAL ? 1 : starts[itb + 1];
    const int tid = threadIdx.x;
    cplx *F = reinterpret_cast<cplx *>(Fs);
    cplx *rad = reinterpret_cast<cplx *>(rads);
    for (int e = e0; e < e1; ++e) {
        const int D = LOCAL ? B : lists[e];
        const int lv = t.level[D];
        const int p = hd.order[lv];
        const int nco = (p + 1) * (p + 1), ncf = (MAXP + 1) * (MAXP + 1);
        const v3 c = boxc(t, D);
        int n_o = 0, m_o = 0;
        if (tid < nco) nm_of(tid, n_o, m_o);
        cplx acc = {0.0, 0.0};
        for (int c0 = 0; c0 < nf; c0 += TC) {
            const int cnt = min(TC, nf - c0);
            __syncthreads();  // the previous chunk consumed
            if (tid < cnt) {
                const int ti = unfilt[f0 + c0 + tid];
                const v3 d = sub(ld3(tgt, nt, ti), c);
                ds[3 * tid] = d.x;
                ds[3 * tid + 1] = d.y;
                ds[3 * tid + 2] = d.z;
                gs[2 * tid] = pot[2 * (size_t)ti];
                gs[2 * tid + 1] = pot[2 * (size_t)ti + 1];
                const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                cplx *rd = rad + tid * (MAXP + 1);
                radial_seq(LOCAL ? 0 : 1, p, hd.k * rr, rd, jts + tid * (MAXP + 3));
                if (!LOCAL)
                    for (int n = 0; n <= p; ++n) rd[n] = conj(rd[n]);
            }
            __syncthreads();
            for (int w = tid; w < cnt * (p + 1); w += blockDim.x) {
                const int q = w / (p + 1), m = w % (p + 1);
                const v3 d = {ds[3 * q], ds[3 * q + 1], ds[3 * q + 2]};
                cplx *Fq = F + (long)q * ncf;
                const cplx *rd = rad + q * (MAXP + 1);
                ylm_column(m, p, d, [&](int n, cplx y) {
                    Fq[IDX(n, m)] = rd[n] * y;
                    if (m > 0) Fq[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
                });
            }
            __syncthreads();
            if (tid < nco)
                for (int q = 0; q < cnt; ++q) {
                    cplx f = F[(long)q * ncf + IDX(n_o, -m_o)];
                    if (m_o & 1) f = -f;
                    acc += cplx{gs[2 * q], gs[2 * q + 1]} * f;
                }
        }
        if (tid < nco) {
            const double sc = ipow(LOCAL ? 1.0 / hd.s[lv] : hd.s[lv], n_o);
            double *o = xadj + ecoef_off(hd, D, lv) + 2L * tid * hd.nr;
            if (LOCAL)
                stc(o, ldc(o) + sc * acc);
            else
                atomic_addc(o, sc * acc);
        }
    }
}

// M2M^T (OP 0: parents of level lv to their children), M2L^T (OP 1, per
// target box: target-or-target-parent boxes of level lv to their list-2
// boxes; the adjoint driver uses the class-grouped kh_m2l_cls_adj), L2L^T
// (OP 2: boxes of level lv >= 1 to their parents): in'_j += sum_i
// conj(T[j][i]) out'_i.  The output box's adjoint staged in LDS; a
// wavefront per table row j at a time, its lanes along the row (coalesced)
// and a shuffle reduction.
template <int OP>
__global__ void __launch_bounds__(256) kh_trans_adj(HD hd, TD t, int lv, int i0, int nbox,
                                                    const int *boxes, const int *starts,
                                                    const int *lists, const double *outa,
                                                    double *ina)
{
    __shared__ double ys[2 * (MAXP + 1) * (MAXP + 1)];
    cplx *y = reinterpret_cast<cplx *>(ys);
    const int ib = blockIdx.x;
    if (ib >= nbox) return;
    const int i = i0 + ib;
    const int B = boxes[i];
    const int lvin = OP == 0 ? lv + 1 : OP == 2 ? lv - 1 : lv;
    const int pin = hd.order[lvin], pout = hd.order[lv];
    const int ncin = (pin + 1) * (pin + 1), ncout = (pout + 1) * (pout + 1);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    const int S1 = 2 * hd.W + 1;
    const int nb = t.nboxes;
    const int bx = t.coords[B], by = t.coords[nb + B], bz = t.coords[2 * nb + B];
    for (int w = tid; w < ncout; w += blockDim.x)
        y[w] = ldc(outa + ecoef_off(hd, B, lv) + 2L * w * hd.nr);
    __syncthreads();
    int ne;
    if (OP == 0) ne = 8;
    else if (OP == 1) ne = starts[i + 1] - starts[i];
    else ne = 1;
    for (int e = 0; e < ne; ++e) {
        int A, cls;
        if (OP == 0) {
            A = t.child[(long)e * nb + B];
            if (A < 0 || t.scumul[A] == 0) continue;
            cls = (t.coords[A] & 1) | ((t.coords[nb + A] & 1) << 1) |
                  ((t.coords[2 * nb + A] & 1) << 2);
        } else if (OP == 1) {
            A = lists[starts[i] + e];
            const int dx = bx - t.coords[A] + hd.W, dy = by - t.coords[nb + A] + hd.W,
                      dz = bz - t.coords[2 * nb + A] + hd.W;
            if ((unsigned)dx >= (unsigned)S1 || (unsigned)dy >= (unsigned)S1 ||
                (unsigned)dz >= (unsigned)S1)
                continue;
            cls = (dx * S1 + dy) * S1 + dz;
        } else {
            A = t.parent[B];
            if (A < 0) continue;
            cls = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
        }
        const long toff = OP == 0 ? hd.m2m_off[lv] : OP == 1 ? hd.m2l_off[lv] : hd.l2l_off[lv];
        const double *T = hd.tab + 2 * (toff + (long)cls * ncin * ncout);
        double *obase = ina + ecoef_off(hd, A, lvin);
        for (int j = wv; j < ncin; j += nw) {
            const double *row = T + 2L * j * ncout;
            cplx acc = {0.0, 0.0};
            for (int q = lane; q < ncout; q += 64) acc += conj(ldc(row + 2 * q)) * y[q];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                acc.x += __shfl_xor(acc.x, off);
                acc.y += __shfl_xor(acc.y, off);
            }
            if (lane == 0) {
                double *o = obase + 2L * j * hd.nr;
                if (OP == 0)
                    stc(o, ldc(o) + acc);
                else
                    atomic_addc(o, acc);
            }
        }
    }
}

// the target box index of global list entry e (starts: CSR over target
// boxes, ntb + 1 entries)
__device__ __forceinline__ int entry_box(const int *starts, int ntb, int e)
{
    int lo = 0, hi = ntb;  // starts[lo] <= e < starts[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (starts[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// P2P^T: a 64-lane block per list-1 entry (target box B, source box Sb),
// lanes over Sb's sources, each summing over B's non-QBX targets; atomics
// into q' (a source is in the list 1 of several target boxes)
__global__ void __launch_bounds__(64) kh_adj_p2p(HD hd, TD t, int ntb, int n1, const int *tboxes,
                                                 const int *starts, const int *lists,
                                                 const int *fstart, const int *fcount,
                                                 const int *unfilt, const double *tgt, int nt,
                                                 const double *src, int ns, const double *pot,
                                                 double *sadj)
{
    __shared__ double tx[3 * 64], tq[2 * 64];
    const int e = blockIdx.x;
    if (e >= n1) return;
    const int itb = entry_box(starts, ntb, e);
    const int B = tboxes[itb];
    const int f0 = fstart[B], nf = fcount[B];
    if (nf == 0) return;
    const int Sb = lists[e];
    const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
    const int lane = threadIdx.x;
    for (int sb = s0; sb < s1; sb += 64) {
        const int si = sb + lane;
        const v3 sp = si < s1 ? ld3(src, ns, si) : v3{0, 0, 0};
        cplx acc = {0.0, 0.0};
        for (int c0 = 0; c0 < nf; c0 += 64) {
            const int cnt = min(64, nf - c0);
            __syncthreads();  // the previous chunk consumed
            if (lane < cnt) {
                const int ti = unfilt[f0 + c0 + lane];
                tx[3 * lane] = tgt[ti];
                tx[3 * lane + 1] = tgt[(size_t)nt + ti];
                tx[3 * lane + 2] = tgt[2 * (size_t)nt + ti];
                tq[2 * lane] = pot[2 * (size_t)ti];
                tq[2 * lane + 1] = pot[2 * (size_t)ti + 1];
            }
            __syncthreads();
            if (si < s1)
                for (int q = 0; q < cnt; ++q) {
                    const cplx g = p2p_val(hd.k, false, sp, v3{0, 0, 0},
                                           v3{tx[3 * q], tx[3 * q + 1], tx[3 * q + 2]});
                    acc += conj(g) * cplx{tq[2 * q], tq[2 * q + 1]};
                }
        }
        if (si < s1) atomic_addc(sadj + 2 * (size_t)si, acc);
    }
}

// M2P^T: a block per list-3 entry (target box B, box D), a thread per
// coefficient of D's multipole; the waves about D's center at B's non-QBX
// targets staged in LDS in chunks of 8; atomics into M~'_D
__global__ void __launch_bounds__(448) kh_adj_m2p(HD hd, TD t, int ntb, int n3, const int *tboxes,
                                                  const int *starts, const int *lists,
                                                  const int *fstart, const int *fcount,
                                                  const int *unfilt, const double *tgt, int nt,
                                                  const double *pot, double *madj)
{
    constexpr int TC = 8;
    __shared__ double Fs[2 * TC * (MAXP + 1) * (MAXP + 1)];
    __shared__ double rads[2 * TC * (MAXP + 1)];
    __shared__ double jts[TC * (MAXP + 3)];
    __shared__ double ds[3 * TC];
    __shared__ double gs[2 * TC];
    const int e = blockIdx.x;
    if (e >= n3) return;
    const int itb = entry_box(starts, ntb, e);
    const int B = tboxes[itb];
    const int f0 = fstart[B], nf = fcount[B];
    if (nf == 0) return;
    const int tid = threadIdx.x;
    cplx *F = reinterpret_cast<cplx *>(Fs);
    cplx *rad = reinterpret_cast<cplx *>(rads);
    const int D = lists[e];
    const int lv = t.level[D];
    const int p = hd.order[lv];
    const int nco = (p + 1) * (p + 1), ncf = (MAXP + 1) * (MAXP + 1);
    const v3 c = boxc(t, D);
    int n_o = 0, m_o = 0;
    if (tid < nco) nm_of(tid, n_o, m_o);
    cplx acc = {0.0, 0.0};
    for (int c0 = 0; c0 < nf; c0 += TC) {
        const int cnt = min(TC, nf - c0);
        __syncthreads();  // the previous chunk consumed
        if (tid < cnt) {
            const int ti = unfilt[f0 + c0 + tid];
            const v3 d = sub(ld3(tgt, nt, ti), c);
            ds[3 * tid] = d.x;
            ds[3 * tid + 1] = d.y;
            ds[3 * tid + 2] = d.z;
            gs[2 * tid] = pot[2 * (size_t)ti];
            gs[2 * tid + 1] = pot[2 * (size_t)ti + 1];
            const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            cplx *rd = rad + tid * (MAXP + 1);
            radial_seq(1, p, hd.k * rr, rd, jts + tid * (MAXP + 3));
            for (int n = 0; n <= p; ++n) rd[n] = conj(rd[n]);
        }
        __syncthreads();
        for (int w = tid; w < cnt * (p + 1); w += blockDim.x) {
            const int q = w / (p + 1), m = w % (p + 1);
            const v3 d = {ds[3 * q], ds[3 * q + 1], ds[3 * q + 2]};
            cplx *Fq = F + (long)q * ncf;
            const cplx *rd = rad + q * (MAXP + 1);
            ylm_column(m, p, d, [&](int n, cplx y) {
                Fq[IDX(n, m)] = rd[n] * y;
                if (m > 0) Fq[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
            });
        }
        __syncthreads();
        if (tid < nco)
            for (int q = 0; q < cnt; ++q) {
                cplx f = F[(long)q * ncf + IDX(n_o, -m_o)];
                if (m_o & 1) f = -f;
                acc += cplx{gs[2 * q], gs[2 * q + 1]} * f;
            }
    }
    if (tid < nco)
        atomic_addc(madj + ecoef_off(hd, D, lv) + 2L * tid * hd.nr, ipow(hd.s[lv], n_o) * acc);
}

// ---- adjoint of the QBX stages (unscaled QBX expansions Lq_c) -------------
//   QBXL2P^T  Lq'_c[n,m] = sum_{t of c} (-1)^m R_n^{-m}(t - c) pot'_t
//   L2QBXL^T  L~'_B[j] += 4 pi s^-n sum_o conj(cf^R(j, o)) Lq'_c[o]
//   M2QBXL^T  M~'_D[j] += 4 pi s^n sum_o conj(cf^S(j, o)) Lq'_c[o]
//   P2QBXL^T  q'_s += -ik sum_c sum_o conj(h_nu) Y_nu^mu(s - c) Lq'_c[o]
// with cf(j, o) the forward's Gaunt sum (kh_e2qbxl).

// QBXL2P^T: a block per center (every center: those without targets get
// zeros), a thread per QBX coefficient, the center's targets staged in LDS
// in chunks of 16
__global__ void __launch_bounds__(128) kh_adj_qbxl2p(HD hd, int nc, const int *cts,
                                                     const int *ctl, const double *tgt, int nt,
                                                     const double *ctr, const double *pot,
                                                     double *qadj)
{
    constexpr int TC = 16;
    constexpr int NCF = (MAXPQ + 1) * (MAXPQ + 1);
    __shared__ double Fs[2 * TC * NCF];
    __shared__ double rads[2 * TC * (MAXPQ + 1)];
    __shared__ double jts[TC * (MAXPQ + 3)];
    __shared__ double ds[3 * TC];
    __shared__ double gs[2 * TC];
    const int c = blockIdx.x;
    if (c >= nc) return;
    const int tid = threadIdx.x;
    const int pq = hd.pq, nco = hd.ncq;
    cplx *F = reinterpret_cast<cplx *>(Fs);
    cplx *rad = reinterpret_cast<cplx *>(rads);
    const v3 cc = ld3(ctr, nc, c);
    int n_o = 0, m_o = 0;
    if (tid < nco) nm_of(tid, n_o, m_o);
    cplx acc = {0.0, 0.0};
    const int t0 = cts[c], ntc = cts[c + 1] - t0;
    for (int c0 = 0; c0 < ntc; c0 += TC) {
        const int cnt = min(TC, ntc - c0);
        __syncthreads();  // the previous chunk consumed
        if (tid < cnt) {
            const int ti = ctl[t0 + c0 + tid];
            const v3 d = sub(ld3(tgt, nt, ti), cc);
            ds[3 * tid] = d.x;
            ds[3 * tid + 1] = d.y;
            ds[3 * tid + 2] = d.z;
            gs[2 * tid] = pot[2 * (size_t)ti];
            gs[2 * tid + 1] = pot[2 * (size_t)ti + 1];
            const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            double *jt = jts + tid * (MAXPQ + 3);
            sph_j_seq(pq, hd.k * rr, jt);
            for (int n = 0; n <= pq; ++n) rad[tid * (MAXPQ + 1) + n] = {jt[n], 0.0};
        }
        __syncthreads();
        for (int w = tid; w < cnt * (pq + 1); w += blockDim.x) {
            const int q = w / (pq + 1), m = w % (pq + 1);
            const v3 d = {ds[3 * q], ds[3 * q + 1], ds[3 * q + 2]};
            cplx *Fq = F + q * NCF;
            const cplx *rd = rad + q * (MAXPQ + 1);
            ylm_column(m, pq, d, [&](int n, cplx y) {
                Fq[IDX(n, m)] = rd[n] * y;
                if (m > 0) Fq[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
            });
        }
        __syncthreads();
        if (tid < nco)
            for (int q = 0; q < cnt; ++q) {
                cplx f = F[q * NCF + IDX(n_o, -m_o)];
                if (m_o & 1) f = -f;
                acc += cplx{gs[2 * q], gs[2 * q + 1]} * f;
            }
    }
    if (tid < nco) stc(qadj + 2 * ((size_t)c * nco + tid), acc);
}

// L2QBXL^T (MODE 1: into the target box's own local) and M2QBXL^T (MODE 0:
// into the list-3 boxes' multipoles): a block per (group of up to GC_GROUP
// centers of one target box, share blockIdx.y of the list-3 entries), a
// thread per box coefficient summing over the group's centers, per center
// the waves of c - c_D in LDS; atomics into the box expansions
template <int MODE>
__global__ void __launch_bounds__(448) kh_adj_e2qbxl(HD hd, TD t, int nunits,
                                                     const int *unit_box, const int *unit_c0,
                                                     const int *tboxes, const int *starts,
                                                     const int *lists, const int *gcs,
                                                     const int *gcl, const double *ctrs,
                                                     int nctr, const double *qadj, double *xadj)
{
    __shared__ double Xs[2 * (MAXP + MAXPQ + 1) * (MAXP + MAXPQ + 1)];
    __shared__ double ys[2 * (MAXPQ + 1) * (MAXPQ + 1)];
    __shared__ double rads[2 * (MAXP + MAXPQ + 3)];
    __shared__ double jt[MAXP + MAXPQ + 4];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int itb = unit_box[u];
    const int g0 = gcs[itb] + unit_c0[u], g1 = min(g0 + GC_GROUP, gcs[itb + 1]);
    cplx *X = reinterpret_cast<cplx *>(Xs);
    cplx *y = reinterpret_cast<cplx *>(ys);
    const int tid = threadIdx.x;
    const int pq = hd.pq, nco = hd.ncq;
    const int ncg = (hd.pg + 1) * (hd.pg + 1);
    const int eb = MODE == 0 ? starts[itb] + (int)blockIdx.y : 0;
    const int ee = MODE == 0 ? starts[itb + 1] : (blockIdx.y == 0 ? 1 : 0);
    const int es = MODE == 0 ? (int)gridDim.y : 1;
    for (int e = eb; e < ee; e += es) {
        const int D = MODE == 0 ? lists[e] : tboxes[itb];
        const int lv = t.level[D];
        const int p = hd.order[lv];
        const int ncin = (p + 1) * (p + 1);
        const v3 cd = boxc(t, D);
        int n = 0, m = 0;
        if (tid < ncin) nm_of(tid, n, m);
        cplx acc = {0.0, 0.0};
        for (int g = g0; g < g1; ++g) {
            const int ci = gcl[g];
            // waves_block synchronises before it writes X: the previous
            // center's X and y are consumed by then
            waves_block(MODE == 0 ? 1 : 0, p + pq, hd.k, sub(ld3(ctrs, nctr, ci), cd), X,
                        reinterpret_cast<cplx *>(rads), jt);
            for (int w = tid; w < nco; w += blockDim.x)
                y[w] = ldc(qadj + 2 * ((size_t)ci * nco + w));
            __syncthreads();
            if (tid < ncin) {
                const double *gj = hd.gaunt + (long)tid * ncg * (hd.pg + 1);
                for (int o = 0; o < nco; ++o) {
                    int nu, mu;
                    nm_of(o, nu, mu);
                    const double *gg = gj + (long)o * (hd.pg + 1);
                    cplx cf = {0.0, 0.0};
                    const int imax = min(n, nu);
                    for (int ii = 0; ii <= imax; ++ii) {
                        const int l = abs(n - nu) + 2 * ii;
                        if (abs(m - mu) > l) continue;
                        const int ex = (nu + l - n) / 2;
                        cf += ((ex & 1) ? -gg[ii] : gg[ii]) * X[IDX(l, m - mu)];
                    }
                    acc += conj(cf) * y[o];
                }
            }
        }
        if (tid < ncin) {
            const double sc = FOURPI * ipow(MODE == 0 ? hd.s[lv] : 1.0 / hd.s[lv], n);
            atomic_addc(xadj + ecoef_off(hd, D, lv) + 2L * tid, sc * acc);
        }
    }
}

// L2QBXL^T / M2QBXL^T with the group's centers summed before the Gaunt
// contraction (p_q <= 3): per list-3 box D (or the own box) the waves of all
// the group's centers in LDS (a lane per center for the radial sequences, a
// lane per (center, m) for the Y columns), Z[q][o] = sum_c conj(X_c[q])
// Lq'_c[o] in LDS, then a thread per box coefficient j: sum_o sum_ii
// (-1)^ex G(j, o, ii) Z[(l, m - mu)][o] -- one contraction per group and box
// instead of one per center.  Dynamic LDS: X [GC_GROUP][nqmax], Z [nqmax][16],
// the radial sequences.
template <int MODE>
__global__ void __launch_bounds__(448) kh_adj_e2qbxl_z(HD hd, TD t, int nunits,
                                                       const int *unit_box, const int *unit_c0,
                                                       const int *tboxes, const int *starts,
                                                       const int *lists, const int *gcs,
                                                       const int *gcl, const double *ctrs,
                                                       int nctr, int nqmax, const double *qadj,
                                                       double *xadj)
{
    extern __shared__ double sh[];
    const int Lm = hd.pmax + hd.pq;
    cplx *X = reinterpret_cast<cplx *>(sh);                          // [GC_GROUP][nqmax]
    cplx *Z = X + GC_GROUP * nqmax;                                   // [nqmax][16]
    cplx *rad = Z + 16 * nqmax;                                       // [GC_GROUP][Lm + 3]
    double *jt = reinterpret_cast<double *>(rad + GC_GROUP * (Lm + 3));  // [GC_GROUP][Lm + 4]
    __shared__ double cxs[GC_GROUP * 3];
    __shared__ cplx ys[GC_GROUP * 16];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int itb = unit_box[u];
    const int c0 = gcs[itb] + unit_c0[u], nct = min(GC_GROUP, gcs[itb + 1] - c0);
    if (nct <= 0) return;
    const int tid = threadIdx.x;
    const int pq = hd.pq, nco = hd.ncq;
    const int ncg = (hd.pg + 1) * (hd.pg + 1);
    if (tid < GC_GROUP * 3) {
        const int q = tid / 3, dd = tid % 3;
        cxs[tid] = q < nct ? ctrs[(size_t)dd * nctr + gcl[c0 + q]] : 0.0;
    }
    if (tid < GC_GROUP * 16) {
        const int q = tid >> 4, o = tid & 15;
        ys[tid] = q < nct && o < nco ? ldc(qadj + 2 * ((size_t)gcl[c0 + q] * nco + o))
                                     : cplx{0.0, 0.0};
    }
    const int eb = MODE == 0 ? starts[itb] + (int)blockIdx.y : 0;
    const int ee = MODE == 0 ? starts[itb + 1] : (blockIdx.y == 0 ? 1 : 0);
    const int es = MODE == 0 ? (int)gridDim.y : 1;
    for (int e = eb; e < ee; e += es) {
        const int D = MODE == 0 ? lists[e] : tboxes[itb];
        const int lv = t.level[D];
        const int p = hd.order[lv];
        const int L = p + pq, nq = (L + 1) * (L + 1);
        const int ncin = (p + 1) * (p + 1);
        const v3 cd = boxc(t, D);
        __syncthreads();  // the previous box's Z consumed; cxs, ys written
        if (tid < nct) {
            const v3 dv = sub(v3{cxs[3 * tid], cxs[3 * tid + 1], cxs[3 * tid + 2]}, cd);
            const double rr = sqrt(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
            radial_seq(MODE == 0 ? 1 : 0, L, hd.k * rr, rad + tid * (Lm + 3),
                       jt + tid * (Lm + 4));
        }
        __syncthreads();
        for (int w = tid; w < nct * (L + 1); w += blockDim.x) {
            const int q = w / (L + 1), m = w % (L + 1);
            const v3 dv = sub(v3{cxs[3 * q], cxs[3 * q + 1], cxs[3 * q + 2]}, cd);
            const cplx *rd = rad + q * (Lm + 3);
            cplx *Xq = X + (long)q * nqmax;
            ylm_column(m, L, dv, [&](int n, cplx y) {
                Xq[IDX(n, m)] = rd[n] * y;
                if (m > 0) Xq[IDX(n, -m)] = rd[n] * sgn_conj(y, m);
            });
        }
        __syncthreads();
        for (int w = tid; w < nq * 16; w += blockDim.x) {
            const int q = w >> 4, o = w & 15;
            cplx z = {0.0, 0.0};
            for (int c = 0; c < nct; ++c) z += conj(X[(long)c * nqmax + q]) * ys[c * 16 + o];
            Z[w] = z;
        }
        __syncthreads();
        if (tid < ncin) {
            int n, m;
            nm_of(tid, n, m);
            const double *gj = hd.gaunt + (long)tid * ncg * (hd.pg + 1);
            cplx acc = {0.0, 0.0};
            for (int o = 0; o < nco; ++o) {
                int nu, mu;
                nm_of(o, nu, mu);
                const double *gg = gj + (long)o * (hd.pg + 1);
                const int imax = min(n, nu);
                for (int ii = 0; ii <= imax; ++ii) {
                    const int l = abs(n - nu) + 2 * ii;
                    if (abs(m - mu) > l) continue;
                    const int ex = (nu + l - n) / 2;
                    acc += ((ex & 1) ? -gg[ii] : gg[ii]) * Z[IDX(l, m - mu) * 16 + o];
                }
            }
            const double sc = FOURPI * ipow(MODE == 0 ? hd.s[lv] : 1.0 / hd.s[lv], n);
            atomic_addc(xadj + ecoef_off(hd, D, lv) + 2L * tid, sc * acc);
        }
    }
}

// P2QBXL^T: a 64-lane block per (group of up to GC_GROUP centers of one
// target box, share blockIdx.y of its list-1 sources), lanes over the
// sources, each summing over the group's centers; one atomic per source.
// PQ >= 0: the S waves in registers (s_waves_reg); PQ = -1: any order, the
// radial sequence in LDS.
template <int PQ>
__global__ void __launch_bounds__(64) kh_adj_p2qbxl(HD hd, TD t, int nunits, const int *unit_box,
                                                    const int *unit_c0, const int *gcs,
                                                    const int *gcl, const double *ctrs,
                                                    int nctr, const int *starts,
                                                    const int *lists, const double *src,
                                                    int ns, const double *qadj, double *sadj)
{
    constexpr int NCR = PQ >= 0 ? (PQ + 1) * (PQ + 1) : 1;
    __shared__ double cq[2 * GC_GROUP * (MAXPQ + 1) * (MAXPQ + 1)];
    __shared__ double cc[3 * GC_GROUP];
    __shared__ double rads[PQ >= 0 ? 1 : 64 * 2 * (MAXPQ + 1)];
    __shared__ double jts[PQ >= 0 ? 1 : 64 * (MAXPQ + 3)];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int itb = unit_box[u];
    const int g0 = gcs[itb] + unit_c0[u], ng = min(GC_GROUP, gcs[itb + 1] - g0);
    const int lane = threadIdx.x;
    const int pq = PQ >= 0 ? PQ : hd.pq, nco = hd.ncq;
    cplx *Lq = reinterpret_cast<cplx *>(cq);
    for (int w = lane; w < ng * nco; w += 64) {
        const int ci = gcl[g0 + w / nco];
        Lq[w] = ldc(qadj + 2 * ((size_t)ci * nco + w % nco));
    }
    if (lane < 3 * ng) {
        const int ci = gcl[g0 + lane / 3];
        cc[lane] = ctrs[(size_t)(lane % 3) * nctr + ci];
    }
    __syncthreads();
    int base = 0;
    const int ns_ = gridDim.y, sh = blockIdx.y;
    for (int e = starts[itb]; e < starts[itb + 1]; ++e) {
        const int Sb = lists[e];
        const int s0 = t.sstart[Sb], cnt = t.scount[Sb];
        // the box's sources, numbered across its list-1 boxes, dealt to the
        // shares round-robin in runs of 64
        for (int j0 = 0; j0 < cnt; j0 += 64) {
            const int run = (base + j0) >> 6;
            if (run % ns_ != sh) continue;
            const int j = j0 + lane;
            if (j >= cnt) continue;
            const int si = s0 + j;
            const v3 sp = ld3(src, ns, si);
            cplx acc = {0.0, 0.0};
            for (int q = 0; q < ng; ++q) {
                const v3 d = sub(sp, v3{cc[3 * q], cc[3 * q + 1], cc[3 * q + 2]});
                const cplx *L = Lq + q * nco;
                if constexpr (PQ >= 0) {
                    cplx F[NCR];
                    s_waves_reg<PQ>(hd.k, d, F);
                    // sum conj(h_n) Y_n^m L = sum conj((-1)^m F[n, -m]) L
#pragma unroll
                    for (int nn = 0; nn <= PQ; ++nn)
#pragma unroll
                        for (int mm = -nn; mm <= nn; ++mm) {
                            cplx f = conj(F[IDX(nn, -mm)]);
                            if (mm & 1) f = -f;
                            acc += f * L[IDX(nn, mm)];
                        }
                } else {
                    cplx *rad = reinterpret_cast<cplx *>(rads) + lane * (MAXPQ + 1);
                    double *jt = jts + lane * (MAXPQ + 3);
                    const double rr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                    radial_seq(1, pq, hd.k * rr, rad, jt);
                    for (int nn = 0; nn <= pq; ++nn) rad[nn] = conj(rad[nn]);
                    cplx a1[1] = {{0.0, 0.0}};
                    sum_waves<1>(hd, reinterpret_cast<const double *>(L), pq, d, rad, a1);
                    acc += a1[0];
                }
            }
            atomic_addc(sadj + 2 * (size_t)si, cplx{0.0, -hd.k} * acc);
        }
        base += cnt;
    }
}

// M2L^T on the same items: Y[16 pairs][nc] = A[16][nc] B with A the pairs'
// target-box local adjoints L~' and B[k][o] = conj(T_cls[o][k]); each row
// added into its source box's multipole adjoint with atomics
__global__ void __launch_bounds__(256) kh_m2l_cls_adj(HD hd, TD t, int nitems, const int *items,
                                                      const int *pe, const int *l2_starts,
                                                      const int *l2_lists, int nttp,
                                                      const int *ttp, const double *ladj,
                                                      double *madj)
{
    extern __shared__ double sh[];
    __shared__ int srcbox[H3G_NP];
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int lv = items[4 * it], cls = items[4 * it + 1], a0 = items[4 * it + 2],
              cnt = items[4 * it + 3];
    const int nc = hd.ncl[lv];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, col = lane & 15,
              kr = lane >> 4;
    cplx *As = reinterpret_cast<cplx *>(sh);
    if (tid < H3G_NP) srcbox[tid] = tid < cnt ? l2_lists[pe[a0 + tid]] : -1;
    for (int w = tid; w < H3G_NP * nc; w += blockDim.x) {
        const int q = w / nc, j = w - q * nc;
        cplx v = {0.0, 0.0};
        if (q < cnt) {
            const int B = ttp[entry_box(l2_starts, nttp, pe[a0 + q])];
            v = ldc(ladj + ecoef_off(hd, B, lv) + 2L * j);
        }
        As[w] = v;
    }
    __syncthreads();
    const double *T = hd.tab + 2 * (hd.m2l_off[lv] + (long)cls * nc * nc);
    const int KS = (nc + 3) / 4, OT = (nc + 15) / 16;
    const int nw = blockDim.x >> 6;
    for (int ot = wv; ot < OT; ot += nw) {
        const int o = ot * 16 + col;
        h3_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + kr;
            double are = 0.0, aim = 0.0, bre = 0.0, bim = 0.0;
            if (k < nc) {
                const cplx a = As[col * nc + k];
                are = a.x;
                aim = a.y;
                if (o < nc) {
                    const cplx b = ldc(T + 2 * ((long)o * nc + k));
                    bre = b.x;
                    bim = -b.y;
                }
            }
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
            cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = 4 * j + kr;
            if (q < cnt && o < nc)
                atomic_addc(madj + ecoef_off(hd, srcbox[q], lv) + 2L * o, cplx{cre[j], cim[j]});
        }
    }
}

// user-order potential adjoints to tree order (centers' entries stay zero)
__global__ void kh_adj_scatter(int nt, int nc, const int *user_to_tree, const double *in,
                               double *out)
{
    const int u = nc + blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nt) return;
    const int ti = user_to_tree[u];
    out[2 * (size_t)ti] = in[2 * (size_t)(u - nc)];
    out[2 * (size_t)ti + 1] = in[2 * (size_t)(u - nc) + 1];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int make_hd(const qbx_kernel_desc *d, const qbx_tree *T, HD *hd, int nsep = 2)
{
    QBX_CHECK(d, "null kernel description");
    QBX_CHECK(d->kernel == QBX_HELMHOLTZ3D, "3D Helmholtz stages: kernel %d", d->kernel);
    QBX_CHECK(!d->target_deriv, "3D Helmholtz stages: target derivatives are not supported");
    QBX_CHECK(d->k > 0.0 && std::isfinite(d->k), "3D Helmholtz stages need a real k > 0 (got %g)",
              d->k);
    QBX_CHECK(d->source_kind == QBX_SOURCE_S || d->source_kind == QBX_SOURCE_D,
              "unsupported source kind %d", d->source_kind);
    QBX_CHECK(d->qbx_order >= 0 && d->qbx_order <= MAXPQ, "3D qbx_order must be in [0, %d]",
              MAXPQ);
    QBX_CHECK(d->level_orders && d->nlevels >= 1, "level_orders missing");
    QBX_CHECK(d->nrhs >= 0, "nrhs must be >= 0 (0 means 1)");
    if (T) {
        QBX_CHECK(T->dim == 3, "3D kernels need a 3D tree");
        QBX_CHECK(T->nlevels <= FML, "tree has too many levels (%d)", T->nlevels);
        QBX_CHECK(d->nlevels >= T->nlevels, "level_orders has %d entries, tree has %d levels",
                  d->nlevels, T->nlevels);
    }
    memset(hd, 0, sizeof(*hd));
    hd->dip = d->source_kind == QBX_SOURCE_D;
    hd->pq = d->qbx_order;
    hd->k = d->k;
    hd->nsep = nsep;
    hd->W = 2 * nsep + 1;
    hd->nr = d->nrhs > 0 ? d->nrhs : 1;
    hd->r0 = 0;
    hd->rc = 1;
    const int nl = T ? T->nlevels : std::min(d->nlevels, FML);
    hd->nlevels = nl;
    long coef = 0;
    int pmax = 0;
    for (int l = 0; l < nl; ++l) {
        const int p = d->level_orders[l];
        QBX_CHECK(p >= 1 && p <= MAXP, "3D fmm order %d at level %d outside [1, %d]", p, l, MAXP);
        hd->order[l] = p;
        pmax = std::max(pmax, p);
        hd->ncl[l] = (p + 1) * (p + 1);
        hd->h[l] = T ? ldexp(T->root_extent, -l) : 1.0;
        hd->s[l] = std::min(1.0, hd->k * hd->h[l]);
        hd->lbox[l] = T ? T->level_start_box_nrs[l] : 0;
        hd->lcoef[l] = coef;
        if (T)
            coef += (long)(T->level_start_box_nrs[l + 1] - T->level_start_box_nrs[l]) * hd->ncl[l];
        hd->m2m_off[l] = hd->l2l_off[l] = hd->m2l_off[l] = -1;
    }
    hd->lbox[nl] = T ? T->nboxes : 0;
    hd->lcoef[nl] = coef;
    hd->pmax = pmax;
    hd->pg = std::max(pmax, hd->pq);
    hd->ncq = (hd->pq + 1) * (hd->pq + 1);
    return QBX_OK;
}

static TD make_td(const qbx_tree *T)
{
    TD t;
    t.nboxes = T->nboxes;
    t.level = T->box_levels.p;
    t.parent = T->box_parent_ids.p;
    t.child = T->box_child_ids.p;
    t.coords = T->box_coords.p;
    t.centers = T->box_centers.p;
    t.sstart = T->box_source_starts.p;
    t.scount = T->box_source_counts_nonchild.p;
    t.scumul = T->box_source_counts_cumul.p;
    return t;
}

static void gauss_legendre(int N, std::vector<double> &x, std::vector<double> &w)
{
    x.resize(N);
    w.resize(N);
    for (int i = 0; i < N; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (N + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int n = 2; n <= N; ++n) {
                const double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
                p0 = p1;
                p1 = p2;
            }
            dp = N * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// normalised associated Legendre functions (Condon-Shortley) at x, all
// |m| <= n <= L
static void pbar_host(int L, double x, double *P)
{
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    double qmm = 1.0 / std::sqrt(4.0 * M_PI), sm = 1.0;
    for (int m = 0; m <= L; ++m) {
        if (m > 0) {
            qmm *= -std::sqrt((2.0 * m + 1.0) / (2.0 * m));
            sm *= s;
        }
        double a2 = 0.0, a1 = qmm;
        P[IDX(m, m)] = qmm * sm;
        for (int n = m + 1; n <= L; ++n) {
            double q;
            if (n == m + 1) {
                q = std::sqrt(2.0 * m + 3.0) * x * qmm;
            } else {
                const double a = std::sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
                const double b = std::sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) /
                                           (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
                q = a * (x * a1 - b * a2);
            }
            P[IDX(n, m)] = q * sm;
            a2 = a1;
            a1 = q;
        }
    }
    for (int n = 1; n <= L; ++n)
        for (int m = 1; m <= n; ++m) P[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * P[IDX(n, m)];
}

// The Gaunt table and the translation tables of this tree, kernel and
// orders, cached in the context.
static int ensure_tables(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, HD *hd,
                         bool need_trans)
{
    QBX_TRY(note_capture(ctx, st));
    // Gaunt table for degrees <= pg
    if (ctx->h3_pg < hd->pg) {
        QBX_TRY(may_replace(ctx, st, "3D Helmholtz Gaunt table"));
        const int pg = std::max(hd->pg, 8), L = 2 * pg, nq = 2 * pg + 2;
        const int ncg = (pg + 1) * (pg + 1), ncl = (L + 1) * (L + 1);
        std::vector<double> xq, wq;
        gauss_legendre(nq, xq, wq);
        std::vector<double> pb((size_t)nq * ncl);
        for (int q = 0; q < nq; ++q) pbar_host(L, xq[q], pb.data() + (size_t)q * ncl);
        dbuf<double> dw, dp;
        QBX_TRY(dw.alloc(nq));
        QBX_TRY(dp.alloc((long)pb.size()));
        QBX_HIP(hipMemcpy(dw.p, wq.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
        QBX_HIP(hipMemcpy(dp.p, pb.data(), sizeof(double) * pb.size(), hipMemcpyHostToDevice));
        QBX_TRY(ctx->ws_h3_gaunt.alloc((long)ncg * ncg * (pg + 1)));
        hipLaunchKernelGGL(kh_gaunt, dim3(cdiv((long)ncg * ncg, 256)), dim3(256), 0, st, pg, nq,
                           dw.p, dp.p, ctx->ws_h3_gaunt.p);
        QBX_HIP(hipGetLastError());
        QBX_HIP(hipStreamSynchronize(st));  // before the staging buffers go
        ctx->h3_pg = pg;
        ctx->h3_key.clear();  // tables index the Gaunt table's layout
    }
    hd->pg = ctx->h3_pg;
    hd->gaunt = ctx->ws_h3_gaunt.p;
    if (!need_trans) return QBX_OK;
    const int nl = hd->nlevels;
    std::vector<double> key = {hd->k, T->root_extent, (double)hd->nsep, (double)nl,
                               (double)hd->pg};
    for (int l = 0; l < nl; ++l) key.push_back(hd->order[l]);
    // table layout: M2M per parent level (8 classes), L2L per child level
    // (8 classes), M2L per level with separated boxes ((2W+1)^3 classes)
    const int S1 = 2 * hd->W + 1;
    const long nm2l = (long)S1 * S1 * S1;
    long off = 0;
    long m2m[FML], l2l[FML], m2l[FML];
    for (int l = 0; l < nl; ++l) {
        m2m[l] = l2l[l] = m2l[l] = -1;
        if (l + 1 < nl) {
            m2m[l] = off;
            off += 8L * hd->ncl[l + 1] * hd->ncl[l];
        }
        if (l >= 1) {
            l2l[l] = off;
            off += 8L * hd->ncl[l - 1] * hd->ncl[l];
        }
        if ((1L << l) > hd->nsep + 1) {
            m2l[l] = off;
            off += nm2l * hd->ncl[l] * hd->ncl[l];
        }
    }
    QBX_CHECK(16L * off <= TAB_BYTES_MAX,
              "3D Helmholtz translation tables would take %ld MiB (orders too high for this "
              "tree depth)", (16L * off) >> 20);
    if (key != ctx->h3_key) {
        QBX_TRY(may_replace(ctx, st, "3D Helmholtz translation tables"));
        QBX_TRY(ctx->ws_h3_tab.alloc(2 * std::max(off, 1L)));
        // class offset vectors t = c2 - c1 per table
        std::vector<double> tv;
        std::vector<long> tvoff(3 * nl, -1);
        for (int l = 0; l < nl; ++l) {
            if (m2m[l] >= 0) {  // child (level l+1) -> parent: t = parent - child
                tvoff[3 * l] = (long)tv.size() / 3;
                const double hc = hd->h[l + 1];
                for (int ci = 0; ci < 8; ++ci) {
                    tv.push_back(-((ci & 1) ? 0.5 : -0.5) * hc);
                    tv.push_back(-((ci & 2) ? 0.5 : -0.5) * hc);
                    tv.push_back(-((ci & 4) ? 0.5 : -0.5) * hc);
                }
            }
            if (l2l[l] >= 0) {  // parent -> child (level l): t = child - parent
                tvoff[3 * l + 1] = (long)tv.size() / 3;
                const double hc = hd->h[l];
                for (int ci = 0; ci < 8; ++ci) {
                    tv.push_back(((ci & 1) ? 0.5 : -0.5) * hc);
                    tv.push_back(((ci & 2) ? 0.5 : -0.5) * hc);
                    tv.push_back(((ci & 4) ? 0.5 : -0.5) * hc);
                }
            }
            if (m2l[l] >= 0) {  // source -> target: t = (target - source) h
                tvoff[3 * l + 2] = (long)tv.size() / 3;
                const double h = hd->h[l];
                for (int dx = -hd->W; dx <= hd->W; ++dx)
                    for (int dy = -hd->W; dy <= hd->W; ++dy)
                        for (int dz = -hd->W; dz <= hd->W; ++dz) {
                            tv.push_back(dx * h);
                            tv.push_back(dy * h);
                            tv.push_back(dz * h);
                        }
            }
        }
        dbuf<double> dtv;
        QBX_TRY(dtv.alloc(std::max<long>(1, (long)tv.size())));
        if (!tv.empty())
            QBX_HIP(hipMemcpy(dtv.p, tv.data(), sizeof(double) * tv.size(),
                              hipMemcpyHostToDevice));
        for (int l = 0; l < nl; ++l) {
            const double k = hd->k;
            if (m2m[l] >= 0)  // M~ = M / s^n: in scale 1/s_child, out 1/s_parent
                hipLaunchKernelGGL(kh_build_tab, dim3(8), dim3(256), 0, st, 0, k, hd->order[l + 1],
                                   hd->order[l], 1.0 / hd->s[l + 1], 1.0 / hd->s[l],
                                   dtv.p + 3 * tvoff[3 * l], hd->pg, hd->gaunt,
                                   ctx->ws_h3_tab.p + 2 * m2m[l]);
            if (l2l[l] >= 0)  // L~ = L s^n: in scale s_parent, out s_child
                hipLaunchKernelGGL(kh_build_tab, dim3(8), dim3(256), 0, st, 0, k, hd->order[l - 1],
                                   hd->order[l], hd->s[l - 1], hd->s[l],
                                   dtv.p + 3 * tvoff[3 * l + 1], hd->pg, hd->gaunt,
                                   ctx->ws_h3_tab.p + 2 * l2l[l]);
            if (m2l[l] >= 0)  // in 1/s, out s
                hipLaunchKernelGGL(kh_build_tab, dim3(nm2l), dim3(256), 0, st, 1, k, hd->order[l],
                                   hd->order[l], 1.0 / hd->s[l], hd->s[l],
                                   dtv.p + 3 * tvoff[3 * l + 2], hd->pg, hd->gaunt,
                                   ctx->ws_h3_tab.p + 2 * m2l[l]);
        }
        QBX_HIP(hipGetLastError());
        QBX_HIP(hipStreamSynchronize(st));  // before the offset vectors go
        ctx->h3_key = key;
    }
    for (int l = 0; l < nl; ++l) {
        hd->m2m_off[l] = m2m[l];
        hd->l2l_off[l] = l2l[l];
        hd->m2l_off[l] = m2l[l];
    }
    hd->tab = ctx->ws_h3_tab.p;
    return QBX_OK;
}

#define FOR_CHUNKS(hd, CH, LAUNCH)                                                       \
    do {                                                                                 \
        if ((hd).nr == 1) {                                                              \
            (hd).r0 = 0;                                                                 \
            (hd).rc = 1;                                                                 \
            LAUNCH(1);                                                                   \
        } else {                                                                         \
            for (int r0_ = 0; r0_ < (hd).nr; r0_ += (CH)) {                             \
                (hd).r0 = r0_;                                                           \
                (hd).rc = std::min((CH), (hd).nr - r0_);                                 \
                LAUNCH(CH);                                                              \
            }                                                                            \
        }                                                                                \
    } while (0)

static inline long exp_doubles(const HD &hd) { return 2L * hd.lcoef[hd.nlevels] * hd.nr; }

// LDS bytes and chunk size of kh_p2e for orders up to pfmax: chunks within
// P2E_LDS_BUDGET KiB, so that several blocks share a CU
constexpr long P2E_LDS_BUDGET = 20;
static void p2e_shape(int pfmax, int nrc, int *CH, size_t *bytes)
{
    const long ncf = (long)(pfmax + 1) * (pfmax + 1);
    const long per = 16 * ncf + 16L * (pfmax + 1) + 16L * nrc + 48 + 8L * (pfmax + 3);
    int ch = (int)std::min(32L, (P2E_LDS_BUDGET * 1024) / per);
    ch = std::max(ch, 1);
    *CH = ch;
    *bytes = (size_t)(per * ch);
}

static int threads_for(int nco)
{
    return std::min(512, std::max(64, 64 * ((nco + 63) / 64)));
}

template <int MODE>
static int launch_p2e(hipStream_t st, HD hd, const TD &td, int nitems, const int *items,
                      const int *starts, const int *lists, const double *ctrs, int nctr,
                      const int *c2tb, const double *src, int ns, const double *str,
                      const double *nrm, int pmax_out, double *out)
{
    if (nitems == 0) return QBX_OK;
    const int pfmax = pmax_out + (hd.dip ? 1 : 0);
    const int nco = (pmax_out + 1) * (pmax_out + 1);
    const int th = threads_for(nco);
#define H3_P2E(N)                                                                          \
    do {                                                                                   \
        int CH;                                                                            \
        size_t by;                                                                         \
        p2e_shape(pfmax, N, &CH, &by);                                                     \
        hipLaunchKernelGGL((kh_p2e<MODE, N>), dim3(nitems), dim3(th), by, st, hd, td,      \
                           nitems, items, starts, lists, ctrs, nctr, c2tb, src, ns, str,   \
                           nrm, CH, pfmax, out);                                           \
    } while (0)
    FOR_CHUNKS(hd, NRC_B, H3_P2E);
#undef H3_P2E
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // namespace h3

using namespace h3;

long h3_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T)
{
    HD hd;
    if (!T || make_hd(d, T, &hd) != QBX_OK) return -1;
    return hd.lcoef[hd.nlevels] * hd.nr;
}

int h3_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    for (int l = 0; l <= hd.nlevels; ++l) offsets[l] = hd.lcoef[l] * hd.nr;
    return QBX_OK;
}

int h3_ncoeff_qbx(const qbx_kernel_desc *d)
{
    HD hd;
    if (make_hd(d, nullptr, &hd) != QBX_OK) return -1;
    return hd.ncq;
}

int h3_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                       const qbx_kernel_desc *d, const double *str, const double *nrm,
                       double *mpole)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    QBX_CHECK(!hd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    QBX_TRY(zero_fill(mpole, exp_doubles(hd), st));
    const TD td = make_td(T);
    return launch_p2e<0>(st, hd, td, R->nsource_boxes, R->source_boxes.p, nullptr, nullptr,
                         nullptr, 0, nullptr, T->sources.p, T->nsources, str, nrm, hd.pmax,
                         mpole);
}

static int h3_trans_level(hipStream_t st, const HD &hd0, const TD &td, int op, int lv, int i0,
                          int nbox, const int *boxes, const int *starts, const int *lists,
                          const double *in, double *out)
{
    if (nbox <= 0) return QBX_OK;
    HD hd = hd0;
    const int lvin = op == 0 ? lv + 1 : op == 2 ? lv - 1 : lv;
    const int ncin = hd.ncl[lvin], ncout = hd.ncl[lv];
    const int th = std::min(448, std::max(64, 64 * ((ncout + 63) / 64)));
#define H3_TR(N)                                                                           \
    do {                                                                                   \
        const size_t by = sizeof(double) * 2 * (size_t)ncin * (N);                         \
        if (op == 0)                                                                       \
            hipLaunchKernelGGL((kh_trans<0, N>), dim3(nbox), dim3(th), by, st, hd, td, lv, \
                               i0, nbox, boxes, starts, lists, in, out);                   \
        else if (op == 1)                                                                  \
            hipLaunchKernelGGL((kh_trans<1, N>), dim3(nbox), dim3(th), by, st, hd, td, lv, \
                               i0, nbox, boxes, starts, lists, in, out);                   \
        else                                                                               \
            hipLaunchKernelGGL((kh_trans<2, N>), dim3(nbox), dim3(th), by, st, hd, td, lv, \
                               i0, nbox, boxes, starts, lists, in, out);                   \
    } while (0)
    FOR_CHUNKS(hd, NRC_B, H3_TR);
#undef H3_TR
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int h3_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, double *mpole, int lmin)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd, R->nsep));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &hd, true));
    const TD td = make_td(T);
    for (int l = T->nlevels - 2; l >= std::max(lmin, 0); --l) {
        const int i0 = R->h_level_start_source_parent_box_nrs[l];
        const int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
        QBX_TRY(h3_trans_level(st, hd, td, 0, l, i0, i1 - i0, R->source_parent_boxes.p, nullptr,
                               nullptr, mpole, mpole));
    }
    return QBX_OK;
}

// the class-grouped list-2 pairs of the levels with M2L tables (kh_m2l_cls),
// cached on the traversal: pe[j] the list entry of sorted pair j, pos[e] its
// position (-1: level without a table), yoff[j] its scratch row, items
// (level, class, first pair, count)
static int h3_m2l_groups(const qbx_tree *T, const qbx_trav *R, const HD &hd)
{
    std::vector<int> lev = {-3};  // the 3D Helmholtz grouping (a 2D one is keyed apart)
    for (int l = 0; l < hd.nlevels; ++l)
        if (hd.m2l_off[l] >= 0) lev.push_back(l * 1000 + hd.ncl[l]);
    if (R->m2lg_nitems >= 0 && R->m2lg_levels == lev) return QBX_OK;
    const int nttp = R->nttp_boxes, n2 = R->n2, nb = T->nboxes;
    std::vector<int> starts(nttp + 1), lists(std::max(n2, 1)), ttp(std::max(nttp, 1)),
        coords(3L * nb);
    QBX_HIP(hipDeviceSynchronize());
    QBX_HIP(hipMemcpy(starts.data(), R->l2_starts.p, sizeof(int) * (nttp + 1),
                      hipMemcpyDeviceToHost));
    if (n2 > 0)
        QBX_HIP(hipMemcpy(lists.data(), R->l2_lists.p, sizeof(int) * n2, hipMemcpyDeviceToHost));
    if (nttp > 0)
        QBX_HIP(hipMemcpy(ttp.data(), R->ttp_boxes.p, sizeof(int) * nttp, hipMemcpyDeviceToHost));
    QBX_HIP(hipMemcpy(coords.data(), T->box_coords.p, sizeof(int) * 3L * nb,
                      hipMemcpyDeviceToHost));
    const int W = hd.W, S1 = 2 * W + 1, S3 = S1 * S1 * S1;
    std::vector<std::pair<long, int>> kv;
    for (int l = 0; l < T->nlevels; ++l) {
        if (hd.m2l_off[l] < 0) continue;
        for (int i = R->h_level_start_ttp_box_nrs[l]; i < R->h_level_start_ttp_box_nrs[l + 1];
             ++i) {
            const int P = ttp[i];
            for (int e = starts[i]; e < starts[i + 1]; ++e) {
                const int A = lists[e];
                const int dx = coords[P] - coords[A], dy = coords[nb + P] - coords[nb + A],
                          dz = coords[2L * nb + P] - coords[2L * nb + A];
                QBX_CHECK(std::abs(dx) <= W && std::abs(dy) <= W && std::abs(dz) <= W,
                          "M2L: list-2 offset (%d, %d, %d) outside the class table", dx, dy, dz);
                kv.push_back({(long)l * S3 + ((dx + W) * S1 + (dy + W)) * S1 + (dz + W), e});
            }
        }
    }
    std::sort(kv.begin(), kv.end());
    const int np = (int)kv.size();
    std::vector<int> pos(std::max(n2, 1), -1), eo(std::max(np, 1)), items;
    for (int j = 0; j < np; ++j) {
        pos[kv[j].second] = j;
        eo[j] = kv[j].second;
    }
    for (int j = 0; j < np;) {
        int j1 = j;
        while (j1 < np && kv[j1].first == kv[j].first) ++j1;
        const int l = (int)(kv[j].first / S3), cls = (int)(kv[j].first % S3);
        for (int a = j; a < j1; a += H3G_NP) {
            items.push_back(l);
            items.push_back(cls);
            items.push_back(a);
            items.push_back(std::min(H3G_NP, j1 - a));
        }
        j = j1;
    }
    std::vector<int> yoff(std::max(np, 1));
    long y = 0;
    for (int j = 0; j < np; ++j) {
        yoff[j] = (int)y;
        y += hd.ncl[kv[j].first / S3];
    }
    QBX_CHECK(y < (1L << 31), "M2L: %ld grouped rows overflow the row offsets", y);
    const int ni = (int)items.size() / 4;
    QBX_TRY(R->m2lg_e.alloc(std::max(np, 1)));
    QBX_TRY(R->m2lg_pos.alloc(std::max(n2, 1)));
    QBX_TRY(R->m2lg_yoff.alloc(std::max(np, 1)));
    QBX_TRY(R->m2lg_items.alloc(std::max(4 * ni, 4)));
    QBX_HIP(hipMemcpy(R->m2lg_e.p, eo.data(), sizeof(int) * eo.size(), hipMemcpyHostToDevice));
    QBX_HIP(hipMemcpy(R->m2lg_pos.p, pos.data(), sizeof(int) * pos.size(),
                      hipMemcpyHostToDevice));
    QBX_HIP(hipMemcpy(R->m2lg_yoff.p, yoff.data(), sizeof(int) * yoff.size(),
                      hipMemcpyHostToDevice));
    if (ni > 0)
        QBX_HIP(hipMemcpy(R->m2lg_items.p, items.data(), sizeof(int) * items.size(),
                          hipMemcpyHostToDevice));
    R->m2lg_nitems = ni;
    R->m2lg_nitems2 = 0;
    R->m2lg_nrows = y;
    R->m2lg_levels = lev;
    return QBX_OK;
}

int h3_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, const double *mpole, double *local)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd, R->nsep));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &hd, true));
    QBX_TRY(zero_fill(local, exp_doubles(hd), st));
    if (R->n2 == 0) return QBX_OK;
    const TD td = make_td(T);
    if (hd.nr == 1) {
        // one density: class-grouped on the matrix cores
        QBX_TRY(h3_m2l_groups(T, R, hd));
        if (R->m2lg_nitems > 0) {
            QBX_TRY(ws_reserve(ctx, st, ctx->ws_m2lg, 2 * std::max(R->m2lg_nrows, 1L),
                               "grouped M2L rows"));
            int ncmax = 1;
            for (int l = 0; l < hd.nlevels; ++l)
                if (hd.m2l_off[l] >= 0) ncmax = std::max(ncmax, hd.ncl[l]);
            const size_t by = sizeof(double) * 2 * H3G_NP * (size_t)ncmax;
            hipLaunchKernelGGL(kh_m2l_cls, dim3(R->m2lg_nitems), dim3(256), by, st, hd, td,
                               R->m2lg_nitems, R->m2lg_items.p, R->m2lg_e.p, R->l2_lists.p,
                               R->m2lg_yoff.p, mpole, ctx->ws_m2lg.p);
            hipLaunchKernelGGL(kh_m2l_cls_sum, dim3(R->nttp_boxes), dim3(128), 0, st, hd, td,
                               R->nttp_boxes, R->ttp_boxes.p, R->l2_starts.p, R->m2lg_pos.p,
                               R->m2lg_yoff.p, ctx->ws_m2lg.p, local);
            QBX_HIP(hipGetLastError());
        }
        return QBX_OK;
    }
    for (int l = 0; l < T->nlevels; ++l) {
        const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        if (i1 == i0 || hd.m2l_off[l] < 0) continue;
        QBX_TRY(h3_trans_level(st, hd, td, 1, l, i0, i1 - i0, R->ttp_boxes.p, R->l2_starts.p,
                               R->l2_lists.p, mpole, local));
    }
    return QBX_OK;
}

int h3_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                     const qbx_kernel_desc *d, double *local, int lmin)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd, R->nsep));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &hd, true));
    const TD td = make_td(T);
    for (int l = std::max(lmin, 0) + 1; l < T->nlevels; ++l) {
        const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        QBX_TRY(h3_trans_level(st, hd, td, 2, l, i0, i1 - i0, R->ttp_boxes.p, nullptr, nullptr,
                               local, local));
    }
    return QBX_OK;
}

int h3_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                   const qbx_kernel_desc *d, const double *str, const double *nrm, double *local,
                   int accumulate)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    QBX_CHECK(!hd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate) QBX_TRY(zero_fill(local, exp_doubles(hd), st));
    if (R->n4 == 0 || R->nttp_boxes == 0) return QBX_OK;
    const TD td = make_td(T);
    return launch_p2e<1>(st, hd, td, R->nttp_boxes, R->ttp_boxes.p, R->l4_starts.p,
                         R->l4_lists.p, nullptr, 0, nullptr, T->sources.p, T->nsources, str, nrm,
                         hd.pmax, local);
}

int h3_eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, int mode, const double *str,
                    const double *nrm, const double *expn, double *pot, int accumulate)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    QBX_CHECK(mode != 0 || !hd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(pot, 2L * (long)T->ntargets * hd.nr, st));
    const int ntb = R->ntarget_boxes;
    if (ntb == 0 || G->nfiltered == 0) return QBX_OK;
    const TD td = make_td(T);
    const int *starts = mode == 1 ? R->l3_starts.p : R->l1_starts.p;
    const int *lists = mode == 1 ? R->l3_lists.p : R->l1_lists.p;
    if (mode != 2) {
        // P2P and M2P over the geometry's balanced work units
        QBX_CHECK(G->trav == R, "P2P / M2P: geometry was built for another traversal");
        const unit_plan &U = mode == 0 ? G->p2p_plan : G->m2p_plan;
        if (U.nunits == 0) return QBX_OK;
#define H3_EU(M, N)                                                                        \
    hipLaunchKernelGGL((kh_eval_u<M, N>), dim3(U.nunits), dim3(64), 0, st, hd, td,         \
                       U.unit_box.p, U.unit_t0.p, U.box_g.p, R->target_boxes.p, starts,   \
                       lists, G->nqbtl_box_target_starts.p,                              \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm, expn, \
                       pot)
#define H3_EU0(N) H3_EU(0, N)
#define H3_EU1(N) H3_EU(1, N)
        if (mode == 0) FOR_CHUNKS(hd, NRC_E, H3_EU0);
        else FOR_CHUNKS(hd, NRC_E, H3_EU1);
#undef H3_EU0
#undef H3_EU1
#undef H3_EU
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define H3_EV(M, N)                                                                        \
    hipLaunchKernelGGL((kh_eval<M, N>), dim3(ntb), dim3(64), 0, st, hd, td, ntb,           \
                       R->target_boxes.p, starts, lists, G->nqbtl_box_target_starts.p,    \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm,    \
                       expn, pot)
#define H3_EV0(N) H3_EV(0, N)
#define H3_EV1(N) H3_EV(1, N)
#define H3_EV2(N) H3_EV(2, N)
    if (mode == 0) FOR_CHUNKS(hd, NRC_E, H3_EV0);
    else if (mode == 1) FOR_CHUNKS(hd, NRC_E, H3_EV1);
    else FOR_CHUNKS(hd, NRC_E, H3_EV2);
#undef H3_EV0
#undef H3_EV1
#undef H3_EV2
#undef H3_EV
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int h3_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                              const qbx_trav *R, const qbx_geo *G, const qbx_kernel_desc *d,
                              const double *str, const double *nrm, double *qexp, int accumulate)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    QBX_CHECK(!hd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, 2L * G->ncenters * hd.ncq * hd.nr, st));
    if (G->nglobal == 0) return QBX_OK;
    QBX_CHECK(G->trav == R, "P2QBXL: geometry was built for another traversal");
    const TD td = make_td(T);
    if (hd.pq <= 5) {
        // the expansion in registers, one density per launch
        const int ng = G->nglobal;
        for (int r0 = 0; r0 < hd.nr; ++r0) {
            hd.r0 = r0;
            hd.rc = 1;
#define H3_PQR(PQV, DV)                                                                    \
    hipLaunchKernelGGL((kh_p2qbxl_reg<PQV, DV>), dim3(ng), dim3(64), 0, st, hd, td, ng,    \
                       G->global_qbx_centers.p, G->center_to_target_box.p, R->l1_starts.p, \
                       R->l1_lists.p, G->centers.p, G->ncenters, T->sources.p, T->nsources, \
                       str, nrm, qexp)
#define H3_PQD(PQV)          \
    if (hd.dip)              \
        H3_PQR(PQV, 1);      \
    else                     \
        H3_PQR(PQV, 0)
            switch (hd.pq) {
            case 0: H3_PQD(0); break;
            case 1: H3_PQD(1); break;
            case 2: H3_PQD(2); break;
            case 3: H3_PQD(3); break;
            case 4: H3_PQD(4); break;
            default: H3_PQD(5); break;
            }
#undef H3_PQD
#undef H3_PQR
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    return launch_p2e<2>(st, hd, td, G->nglobal, G->global_qbx_centers.p, R->l1_starts.p,
                         R->l1_lists.p, G->centers.p, G->ncenters, G->center_to_target_box.p,
                         T->sources.p, T->nsources, str, nrm, hd.pq, qexp);
}

// dynamic LDS bytes of kh_e2qbxl
static size_t e2qbxl_lds(const HD &hd, int nrc)
{
    const long Lm = hd.pmax + hd.pq;
    const long cpl = (Lm + 1) * (Lm + 1) + (long)(hd.pmax + 1) * (hd.pmax + 1) * nrc + 96L * nrc +
                     (Lm + 3);
    return (size_t)(16 * cpl + 8 * (Lm + 4));
}

static int h3_e2qbxl(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                     const qbx_geo *G, const qbx_kernel_desc *d, int mode, const double *expn,
                     double *qexp, int accumulate)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    QBX_TRY(ensure_tables(ctx, st, T, &hd, false));
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, 2L * G->ncenters * hd.ncq * hd.nr, st));
    const int ng = G->nglobal;
    if (ng == 0 || (mode == 0 && R->n3 == 0)) return QBX_OK;
    QBX_CHECK(G->trav == R, "M2QBXL / L2QBXL: geometry was built for another traversal");
    const TD td = make_td(T);
    const int Lm = hd.pmax + hd.pq, nqmax = (Lm + 1) * (Lm + 1);
    const size_t mm_by = 32 * 3 * 2 * 64 + 16 * ((long)GC_GROUP * nqmax + GC_GROUP * (Lm + 3)) +
                         8L * GC_GROUP * (Lm + 4);
    const unit_plan &U = G->gc_group_plan;
    if (mode == 0 && hd.nr == 1 && hd.ncq <= 16 && U.nunits > 0 && mm_by <= 64 * 1024) {
        // one density, p_q <= 3 (and the waves' tile within 96 KiB of LDS):
        // the per-box matrices, then the products
        QBX_TRY(ws_reserve(ctx, st, ctx->ws_h3_m2qb, 2L * T->nboxes * nqmax * 16,
                           "M2QBXL box matrices"));
        hipLaunchKernelGGL(kh_m2qb_mat, dim3(T->nboxes), dim3(256), 0, st, hd, td, nqmax, expn,
                           ctx->ws_h3_m2qb.p);
        const size_t by = mm_by;
        hipLaunchKernelGGL(kh_m2qbxl_mm, dim3(U.nunits, ADJ_SHARES), dim3(256), by, st, hd, td,
                           U.nunits, U.unit_box.p, U.unit_t0.p, G->gc_by_tbox_starts.p,
                           G->gc_by_tbox.p, R->l3_starts.p, R->l3_lists.p, G->centers.p,
                           G->ncenters, nqmax, ctx->ws_h3_m2qb.p, qexp);
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define H3_EQ(M, N)                                                                        \
    hipLaunchKernelGGL((kh_e2qbxl<M, N>), dim3(ng), dim3(64), e2qbxl_lds(hd, N), st, hd, td, ng, \
                       G->global_qbx_centers.p, G->center_to_target_box.p,                 \
                       R->target_boxes.p, R->l3_starts.p, R->l3_lists.p, G->centers.p,     \
                       G->ncenters, expn, qexp)
#define H3_EQ0(N) H3_EQ(0, N)
#define H3_EQ1(N) H3_EQ(1, N)
    if (mode == 0) FOR_CHUNKS(hd, NRC_Q, H3_EQ0);
    else FOR_CHUNKS(hd, NRC_Q, H3_EQ1);
#undef H3_EQ0
#undef H3_EQ1
#undef H3_EQ
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int h3_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                             const qbx_trav *R, const qbx_geo *G,
                                             const qbx_kernel_desc *d, const double *mpole,
                                             double *qexp, int accumulate)
{
    return h3_e2qbxl(ctx, S(stream), T, R, G, d, 0, mpole, qexp, accumulate);
}

int h3_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                        const qbx_trav *R, const qbx_geo *G,
                                        const qbx_kernel_desc *d, const double *local,
                                        double *qexp, int accumulate)
{
    return h3_e2qbxl(ctx, S(stream), T, R, G, d, 1, local, qexp, accumulate);
}

int h3_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                           const qbx_kernel_desc *d, const double *qexp, double *pot)
{
    HD hd;
    QBX_TRY(make_hd(d, T, &hd));
    hipStream_t st = S(stream);
    if (G->nctt > 0) {
#define H3_Q2P(N)                                                                          \
    hipLaunchKernelGGL(kh_qbxl2p<N>, dim3(cdiv(G->nctt, 64)), dim3(64), 0, st, hd, G->nctt, \
                       G->ctt_lists.p, G->tree_ttc.p, T->targets.p, T->ntargets,            \
                       G->centers.p, G->ncenters, qexp, pot)
        FOR_CHUNKS(hd, NRC_E, H3_Q2P);
#undef H3_Q2P
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

static int adjoint_hd(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                      const qbx_kernel_desc *d, HD *hd)
{
    QBX_CHECK(d && d->source_kind == QBX_SOURCE_S,
              "the 3D Helmholtz adjoint QBX-FMM is for the single layer");
    QBX_CHECK(d->nrhs <= 1, "the adjoint QBX-FMM takes one density (nrhs %d)", d->nrhs);
    QBX_TRY(make_hd(d, T, hd, R->nsep));
    QBX_TRY(ensure_tables(ctx, st, T, hd, true));
    hd->r0 = 0;
    hd->rc = 1;
    return QBX_OK;
}

static int adjoint_launch(hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                          const qbx_geo *G, const HD &hd, int stage, const double *in,
                          double *out)
{
    const TD td = make_td(T);
    const int ntb = R->ntarget_boxes, nc = G->ncenters;
    // groups of up to GC_GROUP global centers of one target box
    const unit_plan &U = G->gc_group_plan;
    int ncl_max = 1;
    for (int l = 0; l < hd.nlevels; ++l) ncl_max = std::max(ncl_max, hd.ncl[l]);
    switch (stage) {
    case QBX_ADJ_QBXL2P:
        if (nc > 0)
            hipLaunchKernelGGL(kh_adj_qbxl2p, dim3(nc), dim3(128), 0, st, hd, nc,
                               G->ctt_starts.p, G->ctt_lists.p, T->targets.p, T->ntargets,
                               G->centers.p, in, out);
        break;
    case QBX_ADJ_L2QBXL:
    case QBX_ADJ_M2QBXL:
        if (U.nunits > 0 && (stage == QBX_ADJ_L2QBXL || R->n3 > 0)) {
            const int th = std::min(448, 64 * ((ncl_max + 63) / 64));
            const int ny = stage == QBX_ADJ_L2QBXL ? 1 : ADJ_SHARES;
            const int Lm = hd.pmax + hd.pq, nqmax = (Lm + 1) * (Lm + 1);
            const size_t zby = 16 * ((long)GC_GROUP * nqmax + 16L * nqmax + GC_GROUP * (Lm + 3)) +
                               8L * GC_GROUP * (Lm + 4);
            if (hd.ncq <= 16 && zby <= 96 * 1024) {
                // the group's centers summed before the contraction
                auto kz = stage == QBX_ADJ_L2QBXL ? kh_adj_e2qbxl_z<1> : kh_adj_e2qbxl_z<0>;
                hipLaunchKernelGGL(kz, dim3(U.nunits, ny), dim3(th), zby, st, hd, td, U.nunits,
                                   U.unit_box.p, U.unit_t0.p, R->target_boxes.p, R->l3_starts.p,
                                   R->l3_lists.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,
                                   G->centers.p, nc, nqmax, in, out);
                break;
            }
            auto kern = stage == QBX_ADJ_L2QBXL ? kh_adj_e2qbxl<1> : kh_adj_e2qbxl<0>;
            hipLaunchKernelGGL(kern, dim3(U.nunits, ny), dim3(th), 0, st, hd, td, U.nunits,
                               U.unit_box.p, U.unit_t0.p, R->target_boxes.p, R->l3_starts.p,
                               R->l3_lists.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,
                               G->centers.p, nc, in, out);
        }
        break;
    case QBX_ADJ_P2QBXL:
        if (U.nunits > 0) {
#define H3_AP2Q(PQV)                                                                       \
    hipLaunchKernelGGL(kh_adj_p2qbxl<PQV>, dim3(U.nunits, ADJ_SHARES), dim3(64), 0, st, hd, \
                       td, U.nunits, U.unit_box.p, U.unit_t0.p, G->gc_by_tbox_starts.p,    \
                       G->gc_by_tbox.p, G->centers.p, nc, R->l1_starts.p, R->l1_lists.p,   \
                       T->sources.p, T->nsources, in, out)
            switch (hd.pq) {
            case 0: H3_AP2Q(0); break;
            case 1: H3_AP2Q(1); break;
            case 2: H3_AP2Q(2); break;
            case 3: H3_AP2Q(3); break;
            case 4: H3_AP2Q(4); break;
            case 5: H3_AP2Q(5); break;
            default: H3_AP2Q(-1); break;
            }
#undef H3_AP2Q
        }
        break;
    case QBX_ADJ_L2P:
        if (ntb > 0 && G->nfiltered > 0)
            hipLaunchKernelGGL(kh_adj_e2p<true>, dim3(ntb), dim3(448), 0, st, hd, td, ntb,
                               R->target_boxes.p, nullptr, nullptr, G->nqbtl_box_target_starts.p,
                               G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,
                               T->targets.p, T->ntargets, in, out);
        break;
    case QBX_ADJ_M2P:
        if (R->n3 > 0 && G->nfiltered > 0)
            hipLaunchKernelGGL(kh_adj_m2p, dim3(R->n3), dim3(448), 0, st, hd, td, ntb, R->n3,
                               R->target_boxes.p, R->l3_starts.p, R->l3_lists.p,
                               G->nqbtl_box_target_starts.p,
                               G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,
                               T->targets.p, T->ntargets, in, out);
        break;
    case QBX_ADJ_P2P:
        if (R->n1 > 0 && G->nfiltered > 0)
            hipLaunchKernelGGL(kh_adj_p2p, dim3(R->n1), dim3(64), 0, st, hd, td, ntb, R->n1,
                               R->target_boxes.p, R->l1_starts.p, R->l1_lists.p,
                               G->nqbtl_box_target_starts.p,
                               G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,
                               T->targets.p, T->ntargets, T->sources.p, T->nsources, in, out);
        break;
    case QBX_ADJ_L2L:
        // out ignored: in (the locals' adjoint) is updated in place, finest
        // level first
        for (int l = T->nlevels - 1; l >= 1; --l) {
            const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
            if (i1 > i0)
                hipLaunchKernelGGL(kh_trans_adj<2>, dim3(i1 - i0), dim3(256), 0, st, hd, td, l, i0,
                                   i1 - i0, R->ttp_boxes.p, nullptr, nullptr, in,
                                   const_cast<double *>(in));
        }
        break;
    case QBX_ADJ_M2L:
        if (R->n2 > 0) {
            // class-grouped on the matrix cores (the forward's items)
            QBX_TRY(h3_m2l_groups(T, R, hd));
            if (R->m2lg_nitems > 0) {
                const size_t by = sizeof(double) * 2 * H3G_NP * (size_t)ncl_max;
                hipLaunchKernelGGL(kh_m2l_cls_adj, dim3(R->m2lg_nitems), dim3(256), by, st, hd,
                                   td, R->m2lg_nitems, R->m2lg_items.p, R->m2lg_e.p,
                                   R->l2_starts.p, R->l2_lists.p, R->nttp_boxes,
                                   R->ttp_boxes.p, in, out);
            }
        }
        break;
    case QBX_ADJ_P2L:
        if (R->nttp_boxes > 0 && R->n4 > 0)
            hipLaunchKernelGGL(kh_adj_p2e<true>, dim3(R->nttp_boxes), dim3(64), 0, st, hd, td,
                               R->nttp_boxes, R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p,
                               T->sources.p, T->nsources, in, out);
        break;
    case QBX_ADJ_M2M:
        // out ignored: in (the multipoles' adjoint) is updated in place,
        // coarsest level first
        for (int l = 0; l <= T->nlevels - 2; ++l) {
            const int i0 = R->h_level_start_source_parent_box_nrs[l];
            const int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
            if (i1 > i0)
                hipLaunchKernelGGL(kh_trans_adj<0>, dim3(i1 - i0), dim3(256), 0, st, hd, td, l, i0,
                                   i1 - i0, R->source_parent_boxes.p, nullptr, nullptr, in,
                                   const_cast<double *>(in));
        }
        break;
    case QBX_ADJ_P2M:
        if (R->nsource_boxes > 0)
            hipLaunchKernelGGL(kh_adj_p2e<false>, dim3(R->nsource_boxes), dim3(64), 0, st, hd,
                               td, R->nsource_boxes, R->source_boxes.p, nullptr, nullptr,
                               T->sources.p, T->nsources, in, out);
        break;
    default:
        set_error("qbx_adjoint_stage: unknown stage %d", stage);
        return QBX_ERR_INVALID;
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int h3_adjoint_stage(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                     const qbx_geo *G, const qbx_kernel_desc *d, int stage, const double *in,
                     double *out)
{
    HD hd;
    QBX_TRY(adjoint_hd(ctx, st, T, R, d, &hd));
    return adjoint_launch(st, T, R, G, hd, stage, in, out);
}

int h3_drive_fmm_adjoint(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, const qbx_trav *R,
                         const qbx_geo *G, const qbx_kernel_desc *d, const double *pot_adj_user,
                         double *str_adj)
{
    HD hd;
    QBX_TRY(adjoint_hd(ctx, st, T, R, d, &hd));
    const long nm = 2L * std::max(hd.lcoef[hd.nlevels], 1L);
    const long nq = 2L * std::max(G->ncenters, 1) * hd.ncq;
    const long np = 2L * std::max(T->ntargets, 1);
    QBX_TRY(ws_reserve(ctx, st, ctx->ws_adj_m, nm, "adjoint multipole workspace"));
    QBX_TRY(ws_reserve(ctx, st, ctx->ws_adj_l, nm, "adjoint local workspace"));
    QBX_TRY(ws_reserve(ctx, st, ctx->ws_adj_q, nq, "adjoint QBX workspace"));
    QBX_TRY(ws_reserve(ctx, st, ctx->ws_adj_p, np, "adjoint potential workspace"));
    double *Ma = ctx->ws_adj_m.p, *La = ctx->ws_adj_l.p, *Qa = ctx->ws_adj_q.p,
           *Pa = ctx->ws_adj_p.p;
    QBX_TRY(zero_fill(Ma, nm, st));
    QBX_TRY(zero_fill(La, nm, st));
    QBX_TRY(zero_fill(Pa, np, st));
    QBX_TRY(zero_fill(str_adj, 2L * T->nsources, st));
    const int nc = G->ncenters, nt = T->ntargets;
    if (nt > nc)
        hipLaunchKernelGGL(kh_adj_scatter, dim3(cdiv(nt - nc, 256)), dim3(256), 0, st, nt, nc,
                           G->user_to_tree_target.p, pot_adj_user, Pa);
    QBX_HIP(hipGetLastError());
    auto run = [&](int stage, const double *in, double *out) {
        return adjoint_launch(st, T, R, G, hd, stage, in, out);
    };
    // the forward's stages in reverse: the potentials' adjoint into the QBX
    // and box locals, those down the tree, then into the multipoles and up
    QBX_TRY(run(QBX_ADJ_QBXL2P, Pa, Qa));
    QBX_TRY(run(QBX_ADJ_L2QBXL, Qa, La));
    QBX_TRY(run(QBX_ADJ_L2P, Pa, La));
    QBX_TRY(run(QBX_ADJ_L2L, La, nullptr));
    QBX_TRY(run(QBX_ADJ_M2L, La, Ma));
    QBX_TRY(run(QBX_ADJ_M2QBXL, Qa, Ma));
    QBX_TRY(run(QBX_ADJ_M2P, Pa, Ma));
    QBX_TRY(run(QBX_ADJ_P2L, La, str_adj));
    QBX_TRY(run(QBX_ADJ_P2QBXL, Qa, str_adj));
    QBX_TRY(run(QBX_ADJ_P2P, Pa, str_adj));
    QBX_TRY(run(QBX_ADJ_M2M, Ma, nullptr));
    QBX_TRY(run(QBX_ADJ_P2M, Ma, str_adj));
    return QBX_OK;
}

}  // namespace qbx

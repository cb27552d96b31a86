/*
 * CPU oracle, 3D Helmholtz QBX-FMM (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the QBX-FMM stage sequence of pytential/qbx/fmm.py:414-594 for
 * the 3D Helmholtz kernel G(x, y) = e^{ik|x-y|} / (4 pi |x-y|) (real k > 0)
 * with spherical-wave expansions, written from the mathematics.  The
 * reference reaches 3D Helmholtz through sumpy's expansion classes
 * (qbx/__init__.py:98-112; not present here); spherical waves are the
 * family of its fmmlib backend and of the target-specific path
 * (qbx/target_specific/impl.pyx), with the same coefficient count (p+1)^2.
 *
 *   Y_n^m     orthonormal spherical harmonics with the Condon-Shortley
 *             phase, Y_n^{-m} = (-1)^m conj(Y_n^m); coefficient (n, m) at
 *             n*n + n + m
 *   R_n^m(x) = j_n(k|x|) Y_n^m(x/|x|),  S_n^m(x) = h_n(k|x|) Y_n^m(x/|x|)
 *             (h_n = j_n + i y_n, the outgoing spherical Hankel function)
 *   G(x, y) = ik sum_{n,m} S_n^m(x - c) (-1)^m R_n^{-m}(y - c),  |y-c| < |x-c|
 *   multipole about c:  M_n^m = ik sum q (-1)^m R_n^{-m}(s - c),  phi = sum M S(x - c)
 *   local about c:      L_n^m = ik sum q (-1)^m S_n^{-m}(s - c),  phi = sum L R(x - c)
 *   translations (addition theorems), t = c2 - c1:
 *     M2M, L2L:  b_{nu,mu} = sum_{n,m} C^R_{nu mu, n m}(t) a_{n,m}
 *     M2L:       b_{nu,mu} = sum_{n,m} C^S_{nu mu, n m}(t) a_{n,m}
 *     C^X_{nu mu, n m}(t) = 4 pi sum_l (-1)^((nu + l - n)/2)
 *                           G(n,m; nu,mu; l,m-mu) X_l^{m-mu}(t)
 *     with the Gaunt coefficients G(n,m; nu,mu; l,la) = int Y_n^m
 *     conj(Y_nu^mu) conj(Y_l^la) over the unit sphere (nonzero for
 *     n + nu + l even, |n - nu| <= l <= n + nu), from the plane-wave
 *     expansion e^{ik s.x} = 4 pi sum i^n R_n^m(x) conj(Y_n^m(s)); they are
 *     computed here by Gauss-Legendre quadrature, exact for these
 *     polynomial integrands.
 *   dipole sources (q nu . grad_s) through the ladder relations (F = R, S):
 *     d_z F_n^m = k (A_{n-1}^m F_{n-1}^m - A_n^m F_{n+1}^m),
 *       A_n^m = sqrt(((n+1)^2 - m^2) / ((2n+1)(2n+3)))
 *     (d_x + i d_y) F_n^m = k (a F_{n+1}^{m+1} + b F_{n-1}^{m+1}),
 *       a = sqrt((n+m+1)(n+m+2) / ((2n+1)(2n+3))), b = sqrt((n-m)(n-m-1) / ((2n-1)(2n+1)))
 *     (d_x - i d_y) F_n^m = -k (c F_{n+1}^{m-1} + d F_{n-1}^{m-1}),
 *       c = sqrt((n-m+1)(n-m+2) / ((2n+1)(2n+3))), d = sqrt((n+m)(n+m-1) / ((2n-1)(2n+1)))
 *
 * Potentials are sums of q G (the 1/(4 pi) is part of G).  Expansions are
 * not rescaled (FP64 covers the ranges of the orders and depths tested).
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "qbx_oracle.h"
#undef I
#define JI _Complex_I

typedef double complex cplx;
#define PI_ 3.14159265358979323846
#define FOURPI (4.0 * PI_)
#define IDX(n, m) ((n) * (n) + (n) + (m))
#define PMAXH 24

static inline int nch(int p) { return (p + 1) * (p + 1); }

/* Normalised associated Legendre functions without the sin^m factor:
 * Yn^m(x) = Q_n^m(cos th) ((x + i y)/r)^m for m >= 0, Q_0^0 = 1/sqrt(4 pi),
 * Q_m^m = -sqrt((2m+1)/(2m)) Q_{m-1}^{m-1}, Q_{m+1}^m = sqrt(2m+3) z Q_m^m,
 * Q_n^m = a (z Q_{n-1}^m - b Q_{n-2}^m), a = sqrt((4n^2-1)/(n^2-m^2)),
 * b = sqrt(((n-1)^2-m^2)/(4(n-1)^2-1)). */
static void ylm(int p, const double *x, cplx *Y)
{
    const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double z = x[2] / r;
    const cplx w = (x[0] + JI * x[1]) / r;
    double qmm = 1.0 / sqrt(FOURPI);
    cplx wm = 1.0;
    for (int m = 0; m <= p; ++m) {
        if (m > 0) {
            qmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
            wm *= w;
        }
        double a2 = 0.0, a1 = qmm;
        Y[IDX(m, m)] = qmm * wm;
        for (int n = m + 1; n <= p; ++n) {
            double q;
            if (n == m + 1) {
                q = sqrt(2.0 * m + 3.0) * z * qmm;
            } else {
                const double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
                const double b = sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) /
                                      (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
                q = a * (z * a1 - b * a2);
            }
            Y[IDX(n, m)] = q * wm;
            a2 = a1;
            a1 = q;
        }
    }
    for (int n = 1; n <= p; ++n)
        for (int m = 1; m <= n; ++m) Y[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * conj(Y[IDX(n, m)]);
}

/* j_n(x), n = 0..p, x > 0: Miller's backward recurrence normalised by j_0 or j_1 */
static void sph_jn(int p, double x, double *j)
{
    if (x < 1e-300) {
        for (int n = 0; n <= p; ++n) j[n] = n == 0 ? 1.0 : 0.0;
        return;
    }
    const double mx = fmax((double)p, x);
    const int N = (int)mx + 30 + (int)sqrt(40.0 * mx);
    double fp1 = 0.0, f = 1e-300;
    double buf[PMAXH * 2 + 4];
    for (int n = N; n >= 0; --n) {
        if (n <= p + 1) buf[n] = f;
        const double fm1 = (2.0 * n + 1.0) / x * f - fp1;
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e200) {
            f *= 1e-200;
            fp1 *= 1e-200;
            for (int q = n; q <= p + 1 && q <= N; ++q) buf[q] *= 1e-200;
        }
    }
    const double j0 = sin(x) / x, j1 = sin(x) / (x * x) - cos(x) / x;
    const double sc = fabs(j0) >= fabs(j1) ? j0 / buf[0] : j1 / buf[1];
    for (int n = 0; n <= p; ++n) j[n] = buf[n] * sc;
}

/* h_n(x) = j_n(x) + i y_n(x), y_n by the forward recurrence */
static void sph_hn(int p, double x, cplx *h)
{
    double j[PMAXH * 2 + 4];
    sph_jn(p, x, j);
    double ym2 = -cos(x) / x, ym1 = -cos(x) / (x * x) - sin(x) / x;
    for (int n = 0; n <= p; ++n) {
        double y;
        if (n == 0) y = ym2;
        else if (n == 1) y = ym1;
        else {
            y = (2.0 * n - 1.0) / x * ym1 - ym2;
            ym2 = ym1;
            ym1 = y;
        }
        h[n] = j[n] + JI * y;
    }
}

/* F_n^m(x) for n <= p: R (sing 0) or S (sing 1) */
static void waves(int sing, int p, double k, const double *x, cplx *F)
{
    const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    cplx *Y = F;
    ylm(p, x, Y);
    if (sing) {
        cplx h[PMAXH * 2 + 4];
        sph_hn(p, k * r, h);
        for (int n = 0; n <= p; ++n)
            for (int m = -n; m <= n; ++m) F[IDX(n, m)] = h[n] * Y[IDX(n, m)];
    } else {
        double j[PMAXH * 2 + 4];
        sph_jn(p, k * r, j);
        for (int n = 0; n <= p; ++n)
            for (int m = -n; m <= n; ++m) F[IDX(n, m)] = j[n] * Y[IDX(n, m)];
    }
}

static inline cplx getf(const cplx *F, int p, int n, int m)
{
    return (n < 0 || n > p || m < -n || m > n) ? 0 : F[IDX(n, m)];
}

/* nu . grad F_n^m from the ladder relations; F holds orders <= n + 1 */
static cplx nu_grad(const cplx *F, int pf, double k, int n, int m, const double *nu)
{
    const double An = sqrt(((n + 1.0) * (n + 1.0) - (double)m * m) / ((2.0 * n + 1) * (2.0 * n + 3)));
    const double An1 = n >= 1 && abs(m) <= n - 1
                           ? sqrt(((double)n * n - (double)m * m) / ((2.0 * n - 1) * (2.0 * n + 1)))
                           : 0.0;
    const cplx dz = k * (An1 * getf(F, pf, n - 1, m) - An * getf(F, pf, n + 1, m));
    const double a = sqrt((n + m + 1.0) * (n + m + 2.0) / ((2.0 * n + 1) * (2.0 * n + 3)));
    const double b = n >= 1 ? sqrt(fmax(0.0, (n - m) * (n - m - 1.0)) / ((2.0 * n - 1) * (2.0 * n + 1))) : 0.0;
    const double c = sqrt((n - m + 1.0) * (n - m + 2.0) / ((2.0 * n + 1) * (2.0 * n + 3)));
    const double d = n >= 1 ? sqrt(fmax(0.0, (n + m) * (n + m - 1.0)) / ((2.0 * n - 1) * (2.0 * n + 1))) : 0.0;
    const cplx dp = k * (a * getf(F, pf, n + 1, m + 1) + b * getf(F, pf, n - 1, m + 1));
    const cplx dm = -k * (c * getf(F, pf, n + 1, m - 1) + d * getf(F, pf, n - 1, m - 1));
    return nu[2] * dz + 0.5 * (nu[0] - JI * nu[1]) * dp + 0.5 * (nu[0] + JI * nu[1]) * dm;
}

/* ---- Gaunt coefficients by Gauss-Legendre quadrature */
static void gauss_legendre(int N, double *x, double *w)
{
    for (int i = 0; i < N; ++i) {
        double z = cos(PI_ * (i + 0.75) / (N + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int n = 2; n <= N; ++n) {
                const double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
                p0 = p1;
                p1 = p2;
            }
            if (N == 1) p1 = z, p0 = 1.0;
            dp = N * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* P-bar_n^m(x) = sqrt((2n+1)/(4 pi) (n-m)!/(n+m)!) P_n^m(x) (Condon-Shortley),
 * all 0 <= m <= n <= p, at one point; negative m from (-1)^m symmetry */
static void pbar(int p, double x, double *P)
{
    const double s = sqrt(fmax(0.0, 1.0 - x * x));
    double qmm = 1.0 / sqrt(FOURPI);
    double sm = 1.0;
    for (int m = 0; m <= p; ++m) {
        if (m > 0) {
            qmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m));
            sm *= s;
        }
        double a2 = 0.0, a1 = qmm;
        P[IDX(m, m)] = qmm * sm;
        for (int n = m + 1; n <= p; ++n) {
            double q;
            if (n == m + 1) q = sqrt(2.0 * m + 3.0) * x * qmm;
            else {
                const double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
                const double b = sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) /
                                      (4.0 * (n - 1.0) * (n - 1.0) - 1.0));
                q = a * (x * a1 - b * a2);
            }
            P[IDX(n, m)] = q * sm;
            a2 = a1;
            a1 = q;
        }
    }
    for (int n = 1; n <= p; ++n)
        for (int m = 1; m <= n; ++m) P[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * P[IDX(n, m)];
}

/* gaunt table for n, nu <= pg: g[(IDX(n,m) * nch(pg) + IDX(nu,mu)) * (pg+1) + i]
 * holds G(n,m; nu,mu; l = |n-nu| + 2i, m-mu) */
static double *g_tab = NULL;
static int g_pg = -1;

static int gaunt_ensure(int pg)
{
    int ok = 1;
#pragma omp critical(orc_gaunt)
    {
        if (pg > g_pg) {
            const int nc = nch(pg), L = 2 * pg, NQ = 2 * pg + 2;  /* degree n + nu + l <= 4 pg */
            double *xq = malloc(sizeof(double) * NQ), *wq = malloc(sizeof(double) * NQ);
            double *P = malloc(sizeof(double) * (size_t)NQ * nch(L));
            double *g = calloc((size_t)nc * nc * (pg + 1), sizeof(double));
            if (!xq || !wq || !P || !g) ok = 0;
            else {
                gauss_legendre(NQ, xq, wq);
                for (int q = 0; q < NQ; ++q) pbar(L, xq[q], P + (size_t)q * nch(L));
                for (int n = 0; n <= pg; ++n)
                    for (int m = -n; m <= n; ++m)
                        for (int nu = 0; nu <= pg; ++nu)
                            for (int mu = -nu; mu <= nu; ++mu) {
                                const int la = m - mu;
                                for (int i = 0; i <= (n < nu ? n : nu); ++i) {
                                    const int l = abs(n - nu) + 2 * i;
                                    if (abs(la) > l) continue;
                                    double acc = 0.0;
                                    for (int q = 0; q < NQ; ++q) {
                                        const double *Pq = P + (size_t)q * nch(L);
                                        acc += wq[q] * Pq[IDX(n, m)] * Pq[IDX(nu, mu)] * Pq[IDX(l, la)];
                                    }
                                    g[((size_t)IDX(n, m) * nc + IDX(nu, mu)) * (pg + 1) + i] = 2.0 * PI_ * acc;
                                }
                            }
                free(g_tab);
                g_tab = g;
                g_pg = pg;
            }
            free(xq);
            free(wq);
            free(P);
        }
    }
    return ok;
}

/* exported for tests: G(n,m; nu,mu; l, m-mu) */
double orc_gaunt3(int n, int m, int nu, int mu, int l)
{
    const int pg = n > nu ? n : nu;
    if (pg > PMAXH || l < abs(n - nu) || l > n + nu || ((n + nu + l) & 1) || abs(m - mu) > l ||
        abs(m) > n || abs(mu) > nu)
        return 0.0;
    if (!gaunt_ensure(pg > 8 ? pg : 8)) return NAN;
    return g_tab[((size_t)IDX(n, m) * nch(g_pg) + IDX(nu, mu)) * (g_pg + 1) + (l - abs(n - nu)) / 2];
}

/* b += C^X(t) a: a of order p1 about c1, b of order p2 about c2, t = c2 - c1 */
static void e2e_h(int sing, double k, int p1, int p2, const double *c1, const double *c2,
                  const cplx *a, cplx *b)
{
    const double t[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
    const int pl = p1 + p2, pg = g_pg;
    cplx *X = malloc(sizeof(cplx) * nch(pl));
    waves(sing, pl, k, t, X);
    for (int nu = 0; nu <= p2; ++nu)
        for (int mu = -nu; mu <= nu; ++mu) {
            cplx acc = 0;
            for (int n = 0; n <= p1; ++n)
                for (int m = -n; m <= n; ++m) {
                    const double *g = g_tab + ((size_t)IDX(n, m) * nch(pg) + IDX(nu, mu)) * (pg + 1);
                    cplx c = 0;
                    for (int i = 0; i <= (n < nu ? n : nu); ++i) {
                        const int l = abs(n - nu) + 2 * i;
                        if (abs(m - mu) > l) continue;
                        const int e = (nu + l - n) / 2;
                        c += ((e & 1) ? -1.0 : 1.0) * g[i] * X[IDX(l, m - mu)];
                    }
                    acc += FOURPI * c * a[IDX(n, m)];
                }
            b[IDX(nu, mu)] += acc;
        }
    free(X);
}

/* P2E: one source into a multipole (local 0) or local (local 1) of order p */
static void p2e_h(int local, int dip, double k, int p, const double *c, const double *s,
                  cplx q, const double *nu, cplx *coef)
{
    const double d[3] = {s[0] - c[0], s[1] - c[1], s[2] - c[2]};
    cplx F[(PMAXH + 2) * (PMAXH + 2)];
    const int pf = p + (dip ? 1 : 0);
    waves(local, pf, k, d, F);
    const cplx ikq = JI * k * q;
    for (int n = 0; n <= p; ++n)
        for (int m = -n; m <= n; ++m) {
            const cplx f = dip ? nu_grad(F, pf, k, n, -m, nu) : F[IDX(n, -m)];
            coef[IDX(n, m)] += ikq * ((m & 1) ? -1.0 : 1.0) * f;
        }
}

/* E2P: potential of a multipole (local 0) or local (local 1) at t */
static cplx e2p_h(int local, double k, int p, const double *c, const cplx *coef, const double *t)
{
    const double d[3] = {t[0] - c[0], t[1] - c[1], t[2] - c[2]};
    cplx F[(PMAXH + 1) * (PMAXH + 1)];
    waves(!local, p, k, d, F);
    cplx phi = 0;
    for (int i = 0; i < nch(p); ++i) phi += coef[i] * F[i];
    return phi;
}

static cplx p2p_h(int dip, double k, const double *s, cplx q, const double *nu, const double *t)
{
    const double d[3] = {t[0] - s[0], t[1] - s[1], t[2] - s[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == 0.0) return 0;
    const double r = sqrt(r2);
    const cplx e = cexp(JI * k * r);
    if (!dip) return q * e / (FOURPI * r);
    return q * e * (1.0 - JI * k * r) * (nu[0] * d[0] + nu[1] * d[1] + nu[2] * d[2]) /
           (FOURPI * r2 * r);
}

/* ---- single-operator entry points (dispatched from qbx_oracle.c) */
int orc_p2eh3(const orc_kparams *kp, int local, int order, const double *c, int ns,
              const double *src, const double *nrm, const double *str, double *coeffs)
{
    if (order > PMAXH) return -1;
    cplx *cf = (cplx *)coeffs;
    for (int i = 0; i < nch(order); ++i) cf[i] = 0;
    for (int s = 0; s < ns; ++s) {
        const double x[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
        double nu[3] = {0, 0, 0};
        if (nrm) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
        p2e_h(local, kp->source_kind == 1, kp->k, order, c, x, str[2 * s] + JI * str[2 * s + 1],
              nu, cf);
    }
    return 0;
}

int orc_e2ph3(const orc_kparams *kp, int local, int order, const double *c,
              const double *coeffs, int nt, const double *tgt, double *pot)
{
    if (order > PMAXH) return -1;
    for (int t = 0; t < nt; ++t) {
        const double x[3] = {tgt[t], tgt[(size_t)nt + t], tgt[2 * (size_t)nt + t]};
        const cplx v = e2p_h(local, kp->k, order, c, (const cplx *)coeffs, x);
        pot[2 * t] = creal(v);
        pot[2 * t + 1] = cimag(v);
    }
    return 0;
}

/* kind 0 = M2M, 1 = M2L, 2 = L2L (b overwritten) */
int orc_e2eh3(const orc_kparams *kp, int kind, int p1, int p2, const double *c1,
              const double *c2, const double *src, double *tgt)
{
    if (p1 > PMAXH || p2 > PMAXH || !gaunt_ensure(p1 > p2 ? p1 : p2)) return -1;
    cplx *b = (cplx *)tgt;
    for (int i = 0; i < nch(p2); ++i) b[i] = 0;
    e2e_h(kind == 1, kp->k, p1, p2, c1, c2, (const cplx *)src, b);
    return 0;
}

int orc_p2ph3(const orc_kparams *kp, int ns, const double *src, const double *nrm,
              const double *str, int nt, const double *tgt, double *pot)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        const double x[3] = {tgt[t], tgt[(size_t)nt + t], tgt[2 * (size_t)nt + t]};
        cplx acc = 0;
        for (int s = 0; s < ns; ++s) {
            const double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
            double nu[3] = {0, 0, 0};
            if (nrm) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
            acc += p2p_h(kp->source_kind == 1, kp->k, y, str[2 * s] + JI * str[2 * s + 1], nu, x);
        }
        pot[2 * t] = creal(acc);
        pot[2 * t + 1] = cimag(acc);
    }
    return 0;
}

/* ---- QBX-FMM driver, 3D Helmholtz (pytential/qbx/fmm.py:414-594) */
int orc_qbx_fmmh3d_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                      const double *sources, const double *normals, const double *strengths,
                      int ncenters, const double *centers, const double *radii,
                      const double *targets, const int *uttc, double *result,
                      double *stage_times, double *mpole_out, double *local_out,
                      double *qexp_out)
{
    (void)radii;
    if (t->dim != 3 || kp->kernel != 3 || !(kp->k > 0)) return -1;
    if (kp->nthreads > 0) omp_set_num_threads(kp->nthreads);
    const int ns = t->nsources, nT = t->ntargets, nb = t->nboxes, nl = t->nlevels;
    const double k = kp->k;
    int pmax = 0;
    for (int l = 0; l < nl; ++l) if (kp->level_orders[l] > pmax) pmax = kp->level_orders[l];
    const int pq = kp->qbx_order;
    if (pmax > PMAXH || pq > PMAXH || !gaunt_ensure(pmax > pq ? pmax : pq)) return -1;
    const int ncf = nch(pmax), ncq = nch(pq);
    const int dip = kp->source_kind == 1;
    double tm[16] = {0};
    double t0 = omp_get_wtime(), t1;

    double *S = malloc(sizeof(double) * 3 * (ns ? ns : 1));
    double *NU = calloc(3 * (size_t)(ns ? ns : 1), sizeof(double));
    cplx *Q = malloc(sizeof(cplx) * (ns ? ns : 1));
    for (int i = 0; i < ns; ++i) {
        const int u = t->user_source_ids[i];
        for (int d = 0; d < 3; ++d) {
            S[3 * i + d] = sources[(size_t)d * ns + u];
            if (dip) NU[3 * i + d] = normals[(size_t)d * ns + u];
        }
        Q[i] = strengths[2 * u] + JI * strengths[2 * u + 1];
    }
    double *Tp = malloc(sizeof(double) * 3 * (nT ? nT : 1));
    int *tree_ttc = malloc(sizeof(int) * (nT ? nT : 1));
    for (int i = 0; i < nT; ++i) {
        const int u = t->sorted_target_ids[i];
        for (int d = 0; d < 3; ++d) Tp[3 * i + d] = targets[(size_t)d * nT + u];
        tree_ttc[i] = u < ncenters ? -2 : uttc[u];
    }
    double *C = malloc(sizeof(double) * 3 * (ncenters ? ncenters : 1));
    for (int c = 0; c < ncenters; ++c)
        for (int d = 0; d < 3; ++d) C[3 * c + d] = centers[(size_t)d * ncenters + c];
    double *BC = malloc(sizeof(double) * 3 * nb);
    for (int b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) BC[3 * b + d] = t->box_centers[(size_t)d * nb + b];
    int *tgt_box_of_box = malloc(sizeof(int) * nb);
    for (int b = 0; b < nb; ++b) tgt_box_of_box[b] = -1;
    for (int i = 0; i < tr->ntarget_boxes; ++i) tgt_box_of_box[tr->target_boxes[i]] = i;
#define LEV(b) (t->box_levels[b])
#define PL(b) (kp->level_orders[LEV(b)])

    cplx *M = calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *Lc = calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *pot = calloc(nT ? nT : 1, sizeof(cplx));
    t1 = omp_get_wtime(); tm[0] = t1 - t0; t0 = t1;

    /* P2M */
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < tr->nsource_boxes; ++i) {
        const int b = tr->source_boxes[i];
        const int s0 = t->box_source_starts[b], s1 = s0 + t->box_source_counts_nonchild[b];
        for (int s = s0; s < s1; ++s)
            p2e_h(0, dip, k, PL(b), BC + 3 * b, S + 3 * s, Q[s], NU + 3 * s, M + (size_t)b * ncf);
    }
    t1 = omp_get_wtime(); tm[1] = t1 - t0; t0 = t1;

    /* M2M, upward */
    for (int l = nl - 2; l >= 0; --l) {
        const int i0 = tr->level_start_source_parent_box_nrs[l];
        const int i1 = tr->level_start_source_parent_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            const int P = tr->source_parent_boxes[i];
            for (int c = 0; c < 8; ++c) {
                const int Ch = t->box_child_ids[(size_t)c * nb + P];
                if (Ch < 0 || t->box_source_counts_cumul[Ch] == 0) continue;
                e2e_h(0, k, PL(Ch), PL(P), BC + 3 * Ch, BC + 3 * P, M + (size_t)Ch * ncf,
                      M + (size_t)P * ncf);
            }
        }
    }
    t1 = omp_get_wtime(); tm[2] = t1 - t0; t0 = t1;

    /* P2P (list 1) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        const int B = tr->target_boxes[i];
        const int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l1_starts[i]; j < tr->l1_starts[i + 1]; ++j) {
                const int Sb = tr->l1_lists[j];
                const int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
                for (int s = s0; s < s1; ++s)
                    acc += p2p_h(dip, k, S + 3 * s, Q[s], NU + 3 * s, Tp + 3 * ti);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[3] = t1 - t0; t0 = t1;

    /* M2L (list 2) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        const int P = tr->ttp_boxes[i];
        for (int j = tr->l2_starts[i]; j < tr->l2_starts[i + 1]; ++j) {
            const int A = tr->l2_lists[j];
            e2e_h(1, k, PL(A), PL(P), BC + 3 * A, BC + 3 * P, M + (size_t)A * ncf,
                  Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[4] = t1 - t0; t0 = t1;

    /* M2P (list 3) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        const int B = tr->target_boxes[i];
        const int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l3_starts[i]; j < tr->l3_starts[i + 1]; ++j) {
                const int D = tr->l3_lists[j];
                acc += e2p_h(0, k, PL(D), BC + 3 * D, M + (size_t)D * ncf, Tp + 3 * ti);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[5] = t1 - t0; t0 = t1;

    /* P2L (list 4) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        const int P = tr->ttp_boxes[i];
        for (int j = tr->l4_starts[i]; j < tr->l4_starts[i + 1]; ++j) {
            const int Sb = tr->l4_lists[j];
            const int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e_h(1, dip, k, PL(P), BC + 3 * P, S + 3 * s, Q[s], NU + 3 * s,
                      Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[6] = t1 - t0; t0 = t1;

    /* L2L, downward */
    for (int l = 1; l < nl; ++l) {
        const int i0 = tr->level_start_ttp_box_nrs[l], i1 = tr->level_start_ttp_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            const int P = tr->ttp_boxes[i];
            const int Pp = t->box_parent_ids[P];
            e2e_h(0, k, PL(Pp), PL(P), BC + 3 * Pp, BC + 3 * P, Lc + (size_t)Pp * ncf,
                  Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[7] = t1 - t0; t0 = t1;

    /* L2P on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        const int B = tr->target_boxes[i];
        const int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            pot[ti] += e2p_h(1, k, PL(B), BC + 3 * B, Lc + (size_t)B * ncf, Tp + 3 * ti);
        }
    }
    t1 = omp_get_wtime(); tm[8] = t1 - t0; t0 = t1;

    /* QBX stages: centers with >= 1 target */
    char *used = calloc(ncenters ? ncenters : 1, 1);
    for (int u = ncenters; u < nT; ++u) if (uttc[u] >= 0) used[uttc[u]] = 1;
    int *center_tbox = malloc(sizeof(int) * (ncenters ? ncenters : 1));
    for (int b = 0; b < nb; ++b) {
        const int t0i = t->box_target_starts[b], t1i = t0i + t->box_target_counts_nonchild[b];
        for (int ti = t0i; ti < t1i; ++ti) {
            const int u = t->sorted_target_ids[ti];
            if (u < ncenters) center_tbox[u] = tgt_box_of_box[b];
        }
    }
    cplx *QX = calloc((size_t)(ncenters ? ncenters : 1) * ncq, sizeof(cplx));
    /* P2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        const int it = center_tbox[c];
        for (int j = tr->l1_starts[it]; j < tr->l1_starts[it + 1]; ++j) {
            const int Sb = tr->l1_lists[j];
            const int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e_h(1, dip, k, pq, C + 3 * c, S + 3 * s, Q[s], NU + 3 * s, QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[9] = t1 - t0; t0 = t1;
    /* M2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        const int it = center_tbox[c];
        for (int j = tr->l3_starts[it]; j < tr->l3_starts[it + 1]; ++j) {
            const int D = tr->l3_lists[j];
            e2e_h(1, k, PL(D), pq, BC + 3 * D, C + 3 * c, M + (size_t)D * ncf, QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[10] = t1 - t0; t0 = t1;
    /* L2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        const int B = tr->target_boxes[center_tbox[c]];
        e2e_h(0, k, PL(B), pq, BC + 3 * B, C + 3 * c, Lc + (size_t)B * ncf, QX + (size_t)c * ncq);
    }
    t1 = omp_get_wtime(); tm[11] = t1 - t0; t0 = t1;
    /* QBXL2P */
#pragma omp parallel for schedule(dynamic, 64)
    for (int ti = 0; ti < nT; ++ti) {
        const int c = tree_ttc[ti];
        if (c < 0) continue;
        pot[ti] = e2p_h(1, k, pq, C + 3 * c, QX + (size_t)c * ncq, Tp + 3 * ti);
    }
    t1 = omp_get_wtime(); tm[12] = t1 - t0; t0 = t1;

    if (mpole_out) memcpy(mpole_out, M, sizeof(cplx) * (size_t)nb * ncf);
    if (local_out) memcpy(local_out, Lc, sizeof(cplx) * (size_t)nb * ncf);
    if (qexp_out) memcpy(qexp_out, QX, sizeof(cplx) * (size_t)ncenters * ncq);
    for (int ti = 0; ti < nT; ++ti) {
        const int u = t->sorted_target_ids[ti];
        const cplx v = u < ncenters ? 0 : pot[ti];
        result[2 * u] = creal(v);
        result[2 * u + 1] = cimag(v);
    }
    t1 = omp_get_wtime(); tm[13] = t1 - t0;
    if (stage_times) memcpy(stage_times, tm, sizeof(tm));
#undef LEV
#undef PL
    free(S); free(NU); free(Q); free(Tp); free(tree_ttc); free(C); free(BC);
    free(tgt_box_of_box); free(M); free(Lc); free(pot); free(used); free(center_tbox);
    free(QX);
    return 0;
}

/* Direct QBX, 3D Helmholtz: every QBX target gets the order-p_q local of all
 * sources about its center; other targets the full point-to-point sum. */
int orc_direct_qbxh3d(const orc_kparams *kp, int ns, const double *src, const double *nrm,
                      const double *str, int ncenters, const double *centers, int nt,
                      const double *tgt, const int *ttc, double *result)
{
    if (kp->kernel != 3 || kp->qbx_order > PMAXH || !(kp->k > 0)) return -1;
    const int pq = kp->qbx_order, ncq = nch(pq);
    const int dip = kp->source_kind == 1;
    const double k = kp->k;
#pragma omp parallel for schedule(dynamic, 8)
    for (int it = 0; it < nt; ++it) {
        const double x[3] = {tgt[it], tgt[(size_t)nt + it], tgt[2 * (size_t)nt + it]};
        cplx v = 0;
        const int c = ttc[it];
        if (c >= 0) {
            cplx coef[(PMAXH + 1) * (PMAXH + 1)];
            for (int i = 0; i < ncq; ++i) coef[i] = 0;
            const double cc[3] = {centers[c], centers[(size_t)ncenters + c],
                                  centers[2 * (size_t)ncenters + c]};
            for (int s = 0; s < ns; ++s) {
                const double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
                double nu[3] = {0, 0, 0};
                if (dip) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
                p2e_h(1, dip, k, pq, cc, y, str[2 * s] + JI * str[2 * s + 1], nu, coef);
            }
            v = e2p_h(1, k, pq, cc, coef, x);
        } else {
            for (int s = 0; s < ns; ++s) {
                const double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
                double nu[3] = {0, 0, 0};
                if (dip) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
                v += p2p_h(dip, k, y, str[2 * s] + JI * str[2 * s + 1], nu, x);
            }
        }
        result[2 * it] = creal(v);
        result[2 * it + 1] = cimag(v);
    }
    return 0;
}

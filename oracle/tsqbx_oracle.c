/*
 * CPU oracle for target-specific QBX (TSQBX), 3D Laplace and Helmholtz
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Restates pytential/qbx/target_specific/impl.pyx:566-785
 * (eval_target_specific_qbx_locals) with the kernels of :222-561:
 *   Laplace  S   q   sum_{n<=p} rho_t^n / rho_s^{n+1} P_n(cos th)
 *            Sp  grad_t of the above,  D  dipstr * nu . grad_s of the above
 *   Helmholtz S  i k q sum_{n<=p} (2n+1) j_n(k rho_t) h_n(k rho_s) P_n(cos th)
 *            Sp / D  target / source gradients of the above
 * (unscaled, fmmlib convention: 4 pi times the free-space Green's function),
 * rho_t = |t - c|, rho_s = |s - c|, cos th the angle between t - c and s - c.
 *
 * Spherical Bessel functions of complex argument (the role of jfuns3d_ and
 * h3dall_ in helmholtz_utils.c:70-285, 400-515), written from their
 * definitions:
 *   j_n: Miller's backward recurrence j_{n-1} = (2n+1)/z j_n - j_{n+1},
 *        normalised with j_0 = sin z / z or j_1 = sin z / z^2 - cos z / z;
 *   h_n: h_0 = e^{iz}/(iz), h_1 = h_0 (1/z - i), forward recurrence;
 *   f_n' = f_{n-1} - (n+1)/z f_n, f_0' = -f_1.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#undef I
#define JI _Complex_I

typedef double complex cplx;

void orc_sph_jn_c(int nmax, cplx z, cplx *j, cplx *jd)
{
    if (cabs(z) < 1e-300) {
        for (int n = 0; n <= nmax; ++n) {
            j[n] = n == 0 ? 1.0 : 0.0;
            if (jd) jd[n] = n == 1 ? 1.0 / 3.0 : 0.0;
        }
        return;
    }
    double az = cabs(z);
    int M = (int)(fmax((double)nmax + 1, az)) + 30 + (int)sqrt(40.0 * fmax(nmax + 1, az));
    cplx *f = (cplx *)malloc(sizeof(cplx) * (M + 2));
    cplx zinv = 1.0 / z;
    f[M + 1] = 0;
    f[M] = 1e-30;
    for (int n = M; n >= 1; --n) {
        f[n - 1] = (2.0 * n + 1.0) * zinv * f[n] - f[n + 1];
        if (cabs(f[n - 1]) > 1e250) {
            for (int m = n - 1; m <= M + 1; ++m) f[m] *= 1e-250;
        }
    }
    cplx j0 = csin(z) * zinv;
    cplx j1 = csin(z) * zinv * zinv - ccos(z) * zinv;
    cplx scale = cabs(j0) >= cabs(j1) ? j0 / f[0] : j1 / f[1];
    cplx *jj = (cplx *)malloc(sizeof(cplx) * (nmax + 2));
    for (int n = 0; n <= nmax + 1; ++n) jj[n] = f[n] * scale;
    for (int n = 0; n <= nmax; ++n) {
        j[n] = jj[n];
        if (jd) jd[n] = n == 0 ? -jj[1] : jj[n - 1] - (n + 1.0) * zinv * jj[n];
    }
    free(f);
    free(jj);
}

void orc_sph_hn_c(int nmax, cplx z, cplx *h, cplx *hd)
{
    cplx zinv = 1.0 / z;
    cplx h0 = cexp(JI * z) / (JI * z);
    cplx h1 = h0 * (zinv - JI);
    cplx *hh = (cplx *)malloc(sizeof(cplx) * (nmax + 3));
    hh[0] = h0;
    hh[1] = h1;
    for (int n = 1; n <= nmax; ++n) hh[n + 1] = (2.0 * n + 1.0) * zinv * hh[n] - hh[n - 1];
    for (int n = 0; n <= nmax; ++n) {
        h[n] = hh[n];
        if (hd) hd[n] = n == 0 ? -hh[1] : hh[n - 1] - (n + 1.0) * zinv * hh[n];
    }
    free(hh);
}

/* ctypes-friendly wrappers: out arrays are interleaved complex */
void orc_sph_jn(int nmax, double zr, double zi, double unused, double *j, double *jd)
{
    (void)unused;
    orc_sph_jn_c(nmax, zr + JI * zi, (cplx *)j, (cplx *)jd);
}

void orc_sph_hn(int nmax, double zr, double zi, double unused, double *h, double *hd)
{
    (void)unused;
    orc_sph_hn_c(nmax, zr + JI * zi, (cplx *)h, (cplx *)hd);
}

/* P_n(x), P_n'(x), n = 0..p (fmmlib's legvals recurrence, impl.pyx:132-173) */
static void legendre(double x, int p, double *P, double *dP)
{
    P[0] = 1.0;
    dP[0] = 0.0;
    if (p == 0) return;
    P[1] = x;
    dP[1] = 1.0;
    for (int j = 2; j <= p; ++j) {
        P[j] = ((2.0 * j - 1.0) * x * P[j - 1] - (j - 1.0) * P[j - 2]) / j;
        dP[j] = ((2.0 * j - 1.0) * (P[j - 1] + x * dP[j - 1]) - (j - 1.0) * dP[j - 2]) / j;
    }
}

/*
 * kernel: 0 Laplace, 1 Helmholtz (k complex); op: 0 S (pot), 1 D (pot), 2 Sp (grad)
 * Arrays as in impl.pyx:566-600: sources [3][ns], targets [3][nt],
 * centers [3][nc], CSR in tree numbering; charge complex [ns] (the dipole
 * strength for D), dipvec [3][ns]; pot complex [nt], grad complex [3][nt].
 */
int orc_tsqbx(int kernel, int op, double kre, double kim, int order,
               int ns, const double *sources, int nt, const double *targets,
               int nc, const double *centers,
               int ng, const int *qbx_centers, const int *c2tb,
               const int *ctt_starts, const int *ctt_lists,
               const int *l1_starts, const int *l1_lists,
               const int *box_src_starts, const int *box_src_counts,
               const double *charge, const double *dipvec,
               double *pot, double *grad)
{
    const cplx k = kre + JI * kim;
    if (order < 0 || order > 60) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int ig = 0; ig < ng; ++ig) {
        int c = qbx_centers[ig];
        double cc[3] = {centers[c], centers[(size_t)nc + c], centers[2 * (size_t)nc + c]};
        int tb = c2tb[c];
        double P[64], dP[64];
        cplx jv[64], jd[64], hv[64], hd[64];
        for (int it = ctt_starts[c]; it < ctt_starts[c + 1]; ++it) {
            int tg = ctt_lists[it];
            double tc[3];
            for (int m = 0; m < 3; ++m) tc[m] = targets[(size_t)m * nt + tg] - cc[m];
            double rt = sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
            if (kernel == 1) orc_sph_jn_c(order, k * rt, jv, jd);
            cplx res = 0, g[3] = {0, 0, 0};
            for (int e = l1_starts[tb]; e < l1_starts[tb + 1]; ++e) {
                int sb = l1_lists[e];
                for (int s = box_src_starts[sb]; s < box_src_starts[sb] + box_src_counts[sb]; ++s) {
                    double sc[3];
                    for (int m = 0; m < 3; ++m) sc[m] = sources[(size_t)m * ns + s] - cc[m];
                    double rs = sqrt(sc[0] * sc[0] + sc[1] * sc[1] + sc[2] * sc[2]);
                    double ct = (tc[0] * sc[0] + tc[1] * sc[1] + tc[2] * sc[2]) / (rt * rs);
                    cplx q = charge[2 * s] + JI * charge[2 * s + 1];
                    legendre(ct, order, P, dP);
                    if (kernel == 0) {
                        if (op == 0) {
                            double r = 1.0 / rs, acc = 0;
                            for (int n = 0; n <= order; ++n) { acc += r * P[n]; r *= rt / rs; }
                            res += q * acc;
                        } else if (op == 2) {
                            double gt[3] = {0, 0, 0};
                            double Rn = 1.0 / (rs * rs);
                            for (int n = 1; n <= order; ++n) {
                                for (int m = 0; m < 3; ++m)
                                    gt[m] += Rn * (n * (tc[m] / rt) * P[n] +
                                                   (sc[m] / rs - ct * tc[m] / rt) * dP[n]);
                                Rn *= rt / rs;
                            }
                            for (int m = 0; m < 3; ++m) g[m] += q * gt[m];
                        } else {
                            double gs[3] = {0, 0, 0};
                            double Rn = 1.0 / (rs * rs);
                            for (int n = 0; n <= order; ++n) {
                                for (int m = 0; m < 3; ++m)
                                    gs[m] += Rn * (-(n + 1.0) * (sc[m] / rs) * P[n] +
                                                   (tc[m] / rt - ct * sc[m] / rs) * dP[n]);
                                Rn *= rt / rs;
                            }
                            double nu[3] = {dipvec[s], dipvec[(size_t)ns + s], dipvec[2 * (size_t)ns + s]};
                            res += q * (nu[0] * gs[0] + nu[1] * gs[1] + nu[2] * gs[2]);
                        }
                    } else {
                        orc_sph_hn_c(order, k * rs, hv, op == 1 ? hd : NULL);
                        if (op == 0) {
                            cplx acc = 0;
                            for (int n = 0; n <= order; ++n) acc += (2.0 * n + 1.0) * jv[n] * hv[n] * P[n];
                            res += JI * k * q * acc;
                        } else if (op == 2) {
                            cplx gt[3] = {0, 0, 0};
                            for (int n = 0; n <= order; ++n)
                                for (int m = 0; m < 3; ++m)
                                    gt[m] += (2.0 * n + 1.0) * hv[n] / rt *
                                             (k * jd[n] * P[n] * tc[m] +
                                              (sc[m] / rs - ct * tc[m] / rt) * jv[n] * dP[n]);
                            for (int m = 0; m < 3; ++m) g[m] += JI * k * q * gt[m];
                        } else {
                            cplx gs[3] = {0, 0, 0};
                            for (int n = 0; n <= order; ++n)
                                for (int m = 0; m < 3; ++m)
                                    gs[m] += (2.0 * n + 1.0) * jv[n] / rs *
                                             (k * sc[m] * hd[n] * P[n] +
                                              (tc[m] / rt - ct * sc[m] / rs) * hv[n] * dP[n]);
                            double nu[3] = {dipvec[s], dipvec[(size_t)ns + s], dipvec[2 * (size_t)ns + s]};
                            res += JI * k * q * (gs[0] * nu[0] + gs[1] * nu[1] + gs[2] * nu[2]);
                        }
                    }
                }
            }
            if (op != 2) {
                pot[2 * tg] = creal(res);
                pot[2 * tg + 1] = cimag(res);
            } else {
                for (int m = 0; m < 3; ++m) {
                    grad[2 * ((size_t)m * nt + tg)] = creal(g[m]);
                    grad[2 * ((size_t)m * nt + tg) + 1] = cimag(g[m]);
                }
            }
        }
    }
    return 0;
}

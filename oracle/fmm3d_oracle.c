/*
 * CPU oracle, 3D Laplace QBX-FMM (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the 3D Laplace stages of the QBX-FMM driven by
 * pytential/qbx/fmm.py:414-594 (stage order and data flow as in the 2D
 * oracle in qbx_oracle.c) with solid-harmonic expansions written from the
 * mathematics (the role of sumpy's 3D Laplace expansions; their exact
 * coefficient normalisation cannot be checked here):
 *
 *   Y_n^m(x) = r^n P_n^m(cos th) e^{i m phi} / (n+m)!        (regular)
 *   T_n^m(x) = (n-m)! P_n^m(cos th) e^{i m phi} / r^{n+1}    (irregular)
 *   P_n^m with the Condon-Shortley phase; Y_n^{-m} = (-1)^m conj(Y_n^m),
 *   T_n^{-m} = (-1)^m conj(T_n^m); coefficient (n, m) at n*n + n + m.
 *
 *   1/|x - y| = sum_{n,m} conj(Y_n^m(y)) T_n^m(x),  |y| < |x|
 *   multipole about c:  M_n^m = sum q conj(Y_n^m(s - c)),  phi = sum M T(x - c)
 *   local about c:      L_n^m = sum q T_n^m(s - c),         phi = sum L conj(Y(x - c))
 *   M2M c -> c':  M'_n^m = sum_{k,l} conj(Y_k^l(c - c')) M_{n-k}^{m-l}
 *   M2L c -> c':  L_k^l  = (-1)^k sum_{n,m} M_n^m T_{n+k}^{m+l}(c' - c)
 *   L2L c -> c':  L'_a^b = sum_{j,i} L_{a+j}^{b+i} conj(Y_j^i(c' - c))
 *   dipole sources (q nu . grad_s): gradients from
 *     d_z Y_n^m = Y_{n-1}^m, (d_x + i d_y) Y_n^m = Y_{n-1}^{m+1},
 *     (d_x - i d_y) Y_n^m = -Y_{n-1}^{m-1};
 *     d_z T_n^m = -T_{n+1}^m, (d_x + i d_y) T_n^m = T_{n+1}^{m+1},
 *     (d_x - i d_y) T_n^m = -T_{n+1}^{m-1}.
 *
 * Potentials carry the Green's function 1/(4 pi r) (the target-specific
 * path's output stays unscaled, like impl.pyx).  Expansions are not
 * rescaled by box size (FP64 covers r^{+-p} for the orders and depths used).
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
#define FOURPI (4.0 * 3.14159265358979323846)
#define IDX(n, m) ((n) * (n) + (n) + (m))
#define PMAX3 40

static inline int nc3(int p) { return (p + 1) * (p + 1); }

/* Y_n^m (regular) for n <= p, all m, by the recurrences
 *   Y_m^m = -(x + i y)/(2m) Y_{m-1}^{m-1},
 *   (n-m)(n+m) Y_n^m = (2n-1) z Y_{n-1}^m - r^2 Y_{n-2}^m */
static void ups3(int p, const double *x, cplx *Y)
{
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const cplx w = x[0] + JI * x[1];
    cplx ymm = 1.0;
    for (int m = 0; m <= p; ++m) {
        if (m > 0) ymm = -w / (2.0 * m) * ymm;
        Y[IDX(m, m)] = ymm;
        cplx a = 0, b = ymm;  /* Y_{n-2}, Y_{n-1} */
        for (int n = m + 1; n <= p; ++n) {
            cplx c = ((2.0 * n - 1.0) * x[2] * b - r2 * a) / ((double)(n - m) * (n + m));
            Y[IDX(n, m)] = c;
            a = b;
            b = c;
        }
    }
    for (int n = 0; n <= p; ++n)
        for (int m = 1; m <= n; ++m)
            Y[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * conj(Y[IDX(n, m)]);
}

/* T_n^m (irregular) for n <= p:
 *   T_0^0 = 1/r, T_m^m = -(2m-1) (x + i y)/r^2 T_{m-1}^{m-1},
 *   T_n^m = ((2n-1) z T_{n-1}^m - (n+m-1)(n-m-1) T_{n-2}^m) / r^2 */
static void the3(int p, const double *x, cplx *T)
{
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double ir2 = 1.0 / r2;
    const cplx w = x[0] + JI * x[1];
    cplx tmm = 1.0 / sqrt(r2);
    for (int m = 0; m <= p; ++m) {
        if (m > 0) tmm = -(2.0 * m - 1.0) * w * ir2 * tmm;
        T[IDX(m, m)] = tmm;
        cplx a = 0, b = tmm;
        for (int n = m + 1; n <= p; ++n) {
            cplx c = ((2.0 * n - 1.0) * x[2] * b - (double)(n + m - 1) * (n - m - 1) * a) * ir2;
            T[IDX(n, m)] = c;
            a = b;
            b = c;
        }
    }
    for (int n = 0; n <= p; ++n)
        for (int m = 1; m <= n; ++m)
            T[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * conj(T[IDX(n, m)]);
}

/* exported for tests: interleaved complex [(p+1)^2] each */
void orc_sh3d(int p, const double *x, double *ups, double *the)
{
    if (ups) ups3(p, x, (cplx *)ups);
    if (the) the3(p, x, (cplx *)the);
}

static inline cplx getc(const cplx *A, int p, int n, int m)
{
    return (n < 0 || n > p || m < -n || m > n) ? 0 : A[IDX(n, m)];
}

/* P2E: add one source to a multipole (local = 0) or local (local = 1) */
static void p2e3(int local, int dip, int p, const double *c, const double *s, cplx q,
                 const double *nu, cplx *coef)
{
    double d[3] = {s[0] - c[0], s[1] - c[1], s[2] - c[2]};
    cplx H[(PMAX3 + 2) * (PMAX3 + 2)];
    const int pe = p + (dip && local ? 1 : 0);
    if (local) the3(pe, d, H);
    else ups3(pe, d, H);
    if (!dip) {
        for (int n = 0; n <= p; ++n)
            for (int m = -n; m <= n; ++m)
                coef[IDX(n, m)] += local ? q * H[IDX(n, m)] : q * conj(H[IDX(n, m)]);
        return;
    }
    const cplx nm = nu[0] - JI * nu[1], np = nu[0] + JI * nu[1];
    for (int n = 0; n <= p; ++n)
        for (int m = -n; m <= n; ++m) {
            cplx g;
            if (local)
                g = -nu[2] * getc(H, pe, n + 1, m) +
                    0.5 * (nm * getc(H, pe, n + 1, m + 1) - np * getc(H, pe, n + 1, m - 1));
            else
                g = nu[2] * getc(H, pe, n - 1, m) +
                    0.5 * (nm * getc(H, pe, n - 1, m + 1) - np * getc(H, pe, n - 1, m - 1));
            coef[IDX(n, m)] += local ? q * g : q * conj(g);
        }
}

/* E2P: potential (with 1/(4 pi)) of a multipole (local = 0) or local at t */
static cplx e2p3(int local, int p, const double *c, const cplx *coef, const double *t)
{
    double d[3] = {t[0] - c[0], t[1] - c[1], t[2] - c[2]};
    cplx H[(PMAX3 + 1) * (PMAX3 + 1)];
    cplx phi = 0;
    if (local) {
        ups3(p, d, H);
        for (int i = 0; i < nc3(p); ++i) phi += coef[i] * conj(H[i]);
    } else {
        the3(p, d, H);
        for (int i = 0; i < nc3(p); ++i) phi += coef[i] * H[i];
    }
    return phi / FOURPI;
}

/* E2E: kind 0 = M2M, 1 = M2L, 2 = L2L (adds into b) */
static void e2e3(int kind, int p1, int p2, const double *c1, const double *c2,
                 const cplx *a, cplx *b)
{
    if (kind == 0) {
        double d[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
        cplx Y[(PMAX3 + 1) * (PMAX3 + 1)];
        ups3(p2, d, Y);
        for (int n = 0; n <= p2; ++n)
            for (int m = -n; m <= n; ++m) {
                cplx acc = 0;
                for (int k = 0; k <= n; ++k)
                    for (int l = -k; l <= k; ++l)
                        acc += conj(Y[IDX(k, l)]) * getc(a, p1, n - k, m - l);
                b[IDX(n, m)] += acc;
            }
    } else if (kind == 1) {
        double d[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
        cplx *Tt = (cplx *)malloc(sizeof(cplx) * nc3(p1 + p2));
        the3(p1 + p2, d, Tt);
        for (int k = 0; k <= p2; ++k)
            for (int l = -k; l <= k; ++l) {
                cplx acc = 0;
                for (int n = 0; n <= p1; ++n)
                    for (int m = -n; m <= n; ++m)
                        acc += a[IDX(n, m)] * getc(Tt, p1 + p2, n + k, m + l);
                b[IDX(k, l)] += ((k & 1) ? -1.0 : 1.0) * acc;
            }
        free(Tt);
    } else {
        double d[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
        cplx Y[(PMAX3 + 1) * (PMAX3 + 1)];
        ups3(p1, d, Y);
        for (int aa = 0; aa <= p2; ++aa)
            for (int bb = -aa; bb <= aa; ++bb) {
                cplx acc = 0;
                for (int j = 0; j + aa <= p1; ++j)
                    for (int i = -j; i <= j; ++i)
                        acc += getc(a, p1, aa + j, bb + i) * conj(Y[IDX(j, i)]);
                b[IDX(aa, bb)] += acc;
            }
    }
}

static cplx p2p3(int dip, const double *s, cplx q, const double *nu, const double *t)
{
    double d[3] = {t[0] - s[0], t[1] - s[1], t[2] - s[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == 0.0) return 0;
    double r = sqrt(r2);
    if (!dip) return q / (FOURPI * r);
    return q * (nu[0] * d[0] + nu[1] * d[1] + nu[2] * d[2]) / (FOURPI * r2 * r);
}

/* ---- single-operator entry points (dispatched from qbx_oracle.c) */
int orc_p2e3(const orc_kparams *kp, int local, int order, const double *c, int ns,
             const double *src, const double *nrm, const double *str, double *coeffs)
{
    if (order > PMAX3) return -1;
    cplx *cf = (cplx *)coeffs;
    for (int i = 0; i < nc3(order); ++i) cf[i] = 0;
    for (int s = 0; s < ns; ++s) {
        double x[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
        double nu[3] = {0, 0, 0};
        if (nrm) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
        p2e3(local, kp->source_kind == 1, order, c, x, str[2 * s] + JI * str[2 * s + 1], nu, cf);
    }
    return 0;
}

int orc_e2p3(int local, int order, const double *c, const double *coeffs, int nt,
             const double *tgt, double *pot)
{
    if (order > PMAX3) return -1;
    for (int t = 0; t < nt; ++t) {
        double x[3] = {tgt[t], tgt[(size_t)nt + t], tgt[2 * (size_t)nt + t]};
        cplx v = e2p3(local, order, c, (const cplx *)coeffs, x);
        pot[2 * t] = creal(v);
        pot[2 * t + 1] = cimag(v);
    }
    return 0;
}

int orc_e2e3(int kind, int p1, int p2, const double *c1, const double *c2, const double *src,
             double *tgt)
{
    if (p1 > PMAX3 || p2 > PMAX3) return -1;
    cplx *b = (cplx *)tgt;
    for (int i = 0; i < nc3(p2); ++i) b[i] = 0;
    e2e3(kind, p1, p2, c1, c2, (const cplx *)src, b);
    return 0;
}

int orc_p2p3(const orc_kparams *kp, int ns, const double *src, const double *nrm,
             const double *str, int nt, const double *tgt, double *pot)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        double x[3] = {tgt[t], tgt[(size_t)nt + t], tgt[2 * (size_t)nt + t]};
        cplx acc = 0;
        for (int s = 0; s < ns; ++s) {
            double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
            double nu[3] = {0, 0, 0};
            if (nrm) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
            acc += p2p3(kp->source_kind == 1, y, str[2 * s] + JI * str[2 * s + 1], nu, x);
        }
        pot[2 * t] = creal(acc);
        pot[2 * t + 1] = cimag(acc);
    }
    return 0;
}

/* ---- QBX-FMM driver, 3D (pytential/qbx/fmm.py:414-594) */
int orc_qbx_fmm3d_ex(const orc_tree *t, const orc_trav *tr, const orc_kparams *kp,
                     const double *sources, const double *normals, const double *strengths,
                     int ncenters, const double *centers, const double *radii,
                     const double *targets, const int *uttc, double *result,
                     double *stage_times, double *mpole_out, double *local_out,
                     double *qexp_out)
{
    (void)radii;
    if (t->dim != 3 || kp->kernel != 2) return -1;
    if (kp->nthreads > 0) omp_set_num_threads(kp->nthreads);
    const int ns = t->nsources, nT = t->ntargets, nb = t->nboxes, nl = t->nlevels;
    int pmax = 0;
    for (int l = 0; l < nl; ++l) if (kp->level_orders[l] > pmax) pmax = kp->level_orders[l];
    const int pq = kp->qbx_order;
    if (pmax > PMAX3 || pq > PMAX3) return -1;
    const int ncf = nc3(pmax), ncq = nc3(pq);
    const int dip = kp->source_kind == 1;
    double tm[16] = {0};
    double t0 = omp_get_wtime(), t1;

    /* reorder_sources (fmm.py:436-437) into tree order, AoS for convenience */
    double *S = (double *)malloc(sizeof(double) * 3 * (ns ? ns : 1));
    double *NU = (double *)calloc(3 * (size_t)(ns ? ns : 1), sizeof(double));
    cplx *Q = (cplx *)malloc(sizeof(cplx) * (ns ? ns : 1));
    for (int i = 0; i < ns; ++i) {
        int u = t->user_source_ids[i];
        for (int d = 0; d < 3; ++d) {
            S[3 * i + d] = sources[(size_t)d * ns + u];
            if (dip) NU[3 * i + d] = normals[(size_t)d * ns + u];
        }
        Q[i] = strengths[2 * u] + JI * strengths[2 * u + 1];
    }
    double *Tp = (double *)malloc(sizeof(double) * 3 * (nT ? nT : 1));
    int *tree_ttc = (int *)malloc(sizeof(int) * (nT ? nT : 1));
    for (int i = 0; i < nT; ++i) {
        int u = t->sorted_target_ids[i];
        for (int d = 0; d < 3; ++d) Tp[3 * i + d] = targets[(size_t)d * nT + u];
        tree_ttc[i] = u < ncenters ? -2 : uttc[u];
    }
    double *C = (double *)malloc(sizeof(double) * 3 * (ncenters ? ncenters : 1));
    for (int c = 0; c < ncenters; ++c)
        for (int d = 0; d < 3; ++d) C[3 * c + d] = centers[(size_t)d * ncenters + c];
    double *BC = (double *)malloc(sizeof(double) * 3 * nb);
    for (int b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) BC[3 * b + d] = t->box_centers[(size_t)d * nb + b];
    int *tgt_box_of_box = (int *)malloc(sizeof(int) * nb);
    for (int b = 0; b < nb; ++b) tgt_box_of_box[b] = -1;
    for (int i = 0; i < tr->ntarget_boxes; ++i) tgt_box_of_box[tr->target_boxes[i]] = i;
#define LEV(b) (t->box_levels[b])
#define PL(b) (kp->level_orders[LEV(b)])

    cplx *M = (cplx *)calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *Lc = (cplx *)calloc((size_t)nb * ncf, sizeof(cplx));
    cplx *pot = (cplx *)calloc(nT ? nT : 1, sizeof(cplx));
    t1 = omp_get_wtime(); tm[0] = t1 - t0; t0 = t1;

    /* P2M */
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < tr->nsource_boxes; ++i) {
        int b = tr->source_boxes[i];
        int s0 = t->box_source_starts[b], s1 = s0 + t->box_source_counts_nonchild[b];
        for (int s = s0; s < s1; ++s)
            p2e3(0, dip, PL(b), BC + 3 * b, S + 3 * s, Q[s], NU + 3 * s, M + (size_t)b * ncf);
    }
    t1 = omp_get_wtime(); tm[1] = t1 - t0; t0 = t1;

    /* M2M, upward */
    for (int l = nl - 2; l >= 0; --l) {
        int i0 = tr->level_start_source_parent_box_nrs[l];
        int i1 = tr->level_start_source_parent_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            int P = tr->source_parent_boxes[i];
            for (int c = 0; c < 8; ++c) {
                int Ch = t->box_child_ids[(size_t)c * nb + P];
                if (Ch < 0 || t->box_source_counts_cumul[Ch] == 0) continue;
                e2e3(0, PL(Ch), PL(P), BC + 3 * Ch, BC + 3 * P, M + (size_t)Ch * ncf,
                     M + (size_t)P * ncf);
            }
        }
    }
    t1 = omp_get_wtime(); tm[2] = t1 - t0; t0 = t1;

    /* P2P (list 1) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l1_starts[i]; j < tr->l1_starts[i + 1]; ++j) {
                int Sb = tr->l1_lists[j];
                int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
                for (int s = s0; s < s1; ++s)
                    acc += p2p3(dip, S + 3 * s, Q[s], NU + 3 * s, Tp + 3 * ti);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[3] = t1 - t0; t0 = t1;

    /* M2L (list 2) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        int P = tr->ttp_boxes[i];
        for (int j = tr->l2_starts[i]; j < tr->l2_starts[i + 1]; ++j) {
            int A = tr->l2_lists[j];
            e2e3(1, PL(A), PL(P), BC + 3 * A, BC + 3 * P, M + (size_t)A * ncf,
                 Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[4] = t1 - t0; t0 = t1;

    /* M2P (list 3) on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            cplx acc = 0;
            for (int j = tr->l3_starts[i]; j < tr->l3_starts[i + 1]; ++j) {
                int D = tr->l3_lists[j];
                acc += e2p3(0, PL(D), BC + 3 * D, M + (size_t)D * ncf, Tp + 3 * ti);
            }
            pot[ti] += acc;
        }
    }
    t1 = omp_get_wtime(); tm[5] = t1 - t0; t0 = t1;

    /* P2L (list 4) */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->nttp_boxes; ++i) {
        int P = tr->ttp_boxes[i];
        for (int j = tr->l4_starts[i]; j < tr->l4_starts[i + 1]; ++j) {
            int Sb = tr->l4_lists[j];
            int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e3(1, dip, PL(P), BC + 3 * P, S + 3 * s, Q[s], NU + 3 * s,
                     Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[6] = t1 - t0; t0 = t1;

    /* L2L, downward */
    for (int l = 1; l < nl; ++l) {
        int i0 = tr->level_start_ttp_box_nrs[l], i1 = tr->level_start_ttp_box_nrs[l + 1];
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = i0; i < i1; ++i) {
            int P = tr->ttp_boxes[i];
            int Pp = t->box_parent_ids[P];
            e2e3(2, PL(Pp), PL(P), BC + 3 * Pp, BC + 3 * P, Lc + (size_t)Pp * ncf,
                 Lc + (size_t)P * ncf);
        }
    }
    t1 = omp_get_wtime(); tm[7] = t1 - t0; t0 = t1;

    /* L2P on non-QBX targets */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < tr->ntarget_boxes; ++i) {
        int B = tr->target_boxes[i];
        int t0i = t->box_target_starts[B], t1i = t0i + t->box_target_counts_nonchild[B];
        for (int ti = t0i; ti < t1i; ++ti) {
            if (tree_ttc[ti] != -1) continue;
            pot[ti] += e2p3(1, PL(B), BC + 3 * B, Lc + (size_t)B * ncf, Tp + 3 * ti);
        }
    }
    t1 = omp_get_wtime(); tm[8] = t1 - t0; t0 = t1;

    /* QBX stages: centers with >= 1 target */
    char *used = (char *)calloc(ncenters ? ncenters : 1, 1);
    for (int u = ncenters; u < nT; ++u) if (uttc[u] >= 0) used[uttc[u]] = 1;
    int *center_tbox = (int *)malloc(sizeof(int) * (ncenters ? ncenters : 1));
    for (int b = 0; b < nb; ++b) {
        int t0i = t->box_target_starts[b], t1i = t0i + t->box_target_counts_nonchild[b];
        for (int ti = t0i; ti < t1i; ++ti) {
            int u = t->sorted_target_ids[ti];
            if (u < ncenters) center_tbox[u] = tgt_box_of_box[b];
        }
    }
    cplx *QX = (cplx *)calloc((size_t)(ncenters ? ncenters : 1) * ncq, sizeof(cplx));
    /* P2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int it = center_tbox[c];
        for (int j = tr->l1_starts[it]; j < tr->l1_starts[it + 1]; ++j) {
            int Sb = tr->l1_lists[j];
            int s0 = t->box_source_starts[Sb], s1 = s0 + t->box_source_counts_nonchild[Sb];
            for (int s = s0; s < s1; ++s)
                p2e3(1, dip, pq, C + 3 * c, S + 3 * s, Q[s], NU + 3 * s, QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[9] = t1 - t0; t0 = t1;
    /* M2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int it = center_tbox[c];
        for (int j = tr->l3_starts[it]; j < tr->l3_starts[it + 1]; ++j) {
            int D = tr->l3_lists[j];
            e2e3(1, PL(D), pq, BC + 3 * D, C + 3 * c, M + (size_t)D * ncf, QX + (size_t)c * ncq);
        }
    }
    t1 = omp_get_wtime(); tm[10] = t1 - t0; t0 = t1;
    /* L2QBXL */
#pragma omp parallel for schedule(dynamic, 16)
    for (int c = 0; c < ncenters; ++c) {
        if (!used[c]) continue;
        int B = tr->target_boxes[center_tbox[c]];
        e2e3(2, PL(B), pq, BC + 3 * B, C + 3 * c, Lc + (size_t)B * ncf, QX + (size_t)c * ncq);
    }
    t1 = omp_get_wtime(); tm[11] = t1 - t0; t0 = t1;
    /* QBXL2P */
#pragma omp parallel for schedule(dynamic, 64)
    for (int ti = 0; ti < nT; ++ti) {
        int c = tree_ttc[ti];
        if (c < 0) continue;
        pot[ti] = e2p3(1, pq, C + 3 * c, QX + (size_t)c * ncq, Tp + 3 * ti);
    }
    t1 = omp_get_wtime(); tm[12] = t1 - t0; t0 = t1;

    /* dumps: [rows][ncoeff] complex, coefficient (n, m) at n*n + n + m */
    if (mpole_out) memcpy(mpole_out, M, sizeof(cplx) * (size_t)nb * ncf);
    if (local_out) memcpy(local_out, Lc, sizeof(cplx) * (size_t)nb * ncf);
    if (qexp_out) memcpy(qexp_out, QX, sizeof(cplx) * (size_t)ncenters * ncq);
    for (int ti = 0; ti < nT; ++ti) {
        int u = t->sorted_target_ids[ti];
        cplx v = u < ncenters ? 0 : pot[ti];
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

/* Direct QBX, 3D: every QBX target gets the order-p_q local of all sources
 * about its center; other targets the full point-to-point sum. */
int orc_direct_qbx3d(const orc_kparams *kp, int ns, const double *src, const double *nrm,
                     const double *str, int ncenters, const double *centers, int nt,
                     const double *tgt, const int *ttc, double *result)
{
    if (kp->kernel != 2 || kp->qbx_order > PMAX3) return -1;
    const int pq = kp->qbx_order, ncq = nc3(pq);
    const int dip = kp->source_kind == 1;
#pragma omp parallel for schedule(dynamic, 8)
    for (int it = 0; it < nt; ++it) {
        double x[3] = {tgt[it], tgt[(size_t)nt + it], tgt[2 * (size_t)nt + it]};
        cplx v = 0;
        int c = ttc[it];
        if (c >= 0) {
            cplx coef[(PMAX3 + 1) * (PMAX3 + 1)];
            for (int i = 0; i < ncq; ++i) coef[i] = 0;
            double cc[3] = {centers[c], centers[(size_t)ncenters + c],
                            centers[2 * (size_t)ncenters + c]};
            for (int s = 0; s < ns; ++s) {
                double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
                double nu[3] = {0, 0, 0};
                if (dip) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
                p2e3(1, dip, pq, cc, y, str[2 * s] + JI * str[2 * s + 1], nu, coef);
            }
            v = e2p3(1, pq, cc, coef, x);
        } else {
            for (int s = 0; s < ns; ++s) {
                double y[3] = {src[s], src[(size_t)ns + s], src[2 * (size_t)ns + s]};
                double nu[3] = {0, 0, 0};
                if (dip) { nu[0] = nrm[s]; nu[1] = nrm[(size_t)ns + s]; nu[2] = nrm[2 * (size_t)ns + s]; }
                v += p2p3(dip, y, str[2 * s] + JI * str[2 * s + 1], nu, x);
            }
        }
        result[2 * it] = creal(v);
        result[2 * it + 1] = cimag(v);
    }
    return 0;
}

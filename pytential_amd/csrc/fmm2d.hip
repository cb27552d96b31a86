// QBX-FMM wrangler stages for 2D Laplace and Helmholtz (pytential/qbx/fmm.py
// drive_fmm and pytential/qbx/interactions.py), FP64 / complex128.
//
// Expansions:
//  Laplace   G = -log|x-y|/(2 pi); complex potential Phi, phi = -Re(Phi)/(2 pi)
//            multipole a0 log(z-c) + sum_k a_k (R/(z-c))^k, local sum_l b_l ((z-c)/R)^l
//  Helmholtz G = (i/4) H0(k|x-y|); multipole sum_n M_n H_n(k rho) e^{in theta},
//            local sum_n L_n J_n(k rho) e^{in theta}; translations by Graf's
//            addition theorem (M2M/L2L with J, M2L/M2QBXL with H).
// The double layer is the source-normal derivative of G.
//
// Storage (HBM): box expansions are level-compact.  Level l holds its boxes
// contiguously with p_l + 1 (Laplace) or 2 p_l + 1 (Helmholtz) complex
// coefficients each; coefficient n of a Helmholtz box sits at n + p_l.  QBX
// expansions are [ncenters][ncoeff(p_q)].
//
// Thread mappings:
//  P2M, P2L        one thread per (box, |n|)
//  M2M, L2L, M2L   one thread per (box, output coefficient); Helmholtz
//                  translation vectors come from per-level tables (4 child
//                  offsets for M2M/L2L, the offset classes of list 2 for
//                  M2L), so each output is one dot product
//  P2P, M2P, L2P   one block per target box, one thread per target
//  P2QBXL          16 lanes per QBX center over its list-1 sources, shuffle
//                  reduction, order-p_q expansion in registers
//  M2QBXL, L2QBXL  one thread per QBX center, expansion in registers
//  QBXL2P          one thread per (center, target) pair
#include "qbx_internal.h"
#include "qbx_math.h"

#include <cstring>

namespace qbx {

constexpr int MAXP = 120;
constexpr int BINOM_N = 2 * MAXP + 4;
constexpr int FML = 30;        // levels of a 2D tree (coordinates have 28 bits)

struct KD {
    int kernel, dip, pq, nlevels, ncq;
    int nsep;                  // well_sep_is_n_away of the traversal
    int m2l_w;                 // list-2 offsets satisfy |dI| <= m2l_w = 2 nsep + 1
    double k;
    int order[FML];
    int ncl[FML];
    int lbox[FML + 1];
    long lcoef[FML + 1];
    double rscale[FML];
    const double *tab_m2l;     // Helmholtz M2L table (centred rows)
    const double *tab_mm;      // Helmholtz M2M/L2L table (centred rows)
    long m2l_off[FML];
    long mm_off[FML];
};

struct TD {
    int nboxes;
    const int *level, *parent, *child, *coords;
    const double *centers;  // [2][nboxes]
    const int *sstart, *scount, *scumul;
};

// thread index -> (level, item, j) for launches that cover several levels
struct Flat {
    int nl;
    long tstart[FML + 1];
    int istart[FML];
    int nper[FML];
};

__device__ __forceinline__ bool flat_locate(const Flat &f, long gid, int &lv, int &i, int &j)
{
    if (gid >= f.tstart[f.nl]) return false;
    int l = 0;
    while (gid >= f.tstart[l + 1]) ++l;
    const long r = gid - f.tstart[l];
    lv = l;
    i = f.istart[l] + (int)(r / f.nper[l]);
    j = (int)(r % f.nper[l]);
    return true;
}

__device__ __forceinline__ cplx boxc(const TD &t, int b)
{
    return {t.centers[b], t.centers[(size_t)t.nboxes + b]};
}

__device__ __forceinline__ cplx ldpt(const double *x, int n, int i)
{
    return {x[i], x[(size_t)n + i]};
}

// pointer to coefficient 0 of box b (level lv)
__device__ __forceinline__ long ecoef_off(const KD &kd, int b, int lv)
{
    const long off = kd.lcoef[lv] + (long)(b - kd.lbox[lv]) * kd.ncl[lv];
    return 2 * (off + (kd.kernel == QBX_HELMHOLTZ2D ? kd.order[lv] : 0));
}

constexpr double INV2PI = 0.15915494309189533577;

__device__ __forceinline__ cplx cpow_int(cplx z, int m)
{
    cplx r = {1, 0};
    while (m > 0) {
        if (m & 1) r = r * z;
        z = z * z;
        m >>= 1;
    }
    return r;
}

// -------------------------------------------------------------------------
// pairwise kernel
// -------------------------------------------------------------------------
__device__ __forceinline__ cplx p2p_pair(const KD &kd, cplx s, cplx q, cplx nu, cplx t)
{
    const cplx w = t - s;
    const double r2 = abs2(w);
    if (r2 == 0.0) return {0.0, 0.0};
    if (kd.kernel == QBX_LAPLACE2D) {
        if (!kd.dip) return (-0.5 * log(r2) * INV2PI) * q;
        const cplx v = (nu * q) * cinv(w);
        return {v.x * INV2PI, 0.0};
    }
    const double r = sqrt(r2);
    const double kr = kd.k * r;
    cplx h0, h1;
    if (!kd.dip) {
        hankel01_fast<true, false>(kr, h0, h1);
        const cplx v = q * h0;
        return {-0.25 * v.y, 0.25 * v.x};
    }
    hankel01_fast<false, true>(kr, h0, h1);
    const double proj = (w.x * nu.x + w.y * nu.y) / r;
    const cplx v = (kd.k * proj) * (q * h1);
    return {-0.25 * v.y, 0.25 * v.x};
}

// -------------------------------------------------------------------------
// P2M / P2L
// -------------------------------------------------------------------------
// Helmholtz: coefficients n = m and n = -m (m >= 0) of the expansion about c
// of sources [s0,s1).  g_j = Z_j(k rho) e^{-i j theta} of v = s - c, Z = H for
// locals and J for multipoles, g_{-j} = (-1)^j Z_j e^{+i j theta}.
__device__ void helm_p2e_pm(const KD &kd, int local, int m, cplx c, const double *src, int ns,
                            int s0, int s1, const double *str, const double *nrm, cplx &cp,
                            cplx &cm)
{
    cp = {0, 0};
    cm = {0, 0};
    const int top = m + 1;
    const double sm = (m & 1) ? -1.0 : 1.0;  // (-1)^m
    for (int s = s0; s < s1; ++s) {
        const cplx v = ldpt(src, ns, s) - c;
        const double rho = sqrt(abs2(v));
        const double x = kd.k * rho;
        const cplx q = ldc(str + 2 * s);
        const cplx e1 = rho > 0 ? cplx{v.x / rho, -v.y / rho} : cplx{1, 0};  // e^{-i theta}
        // Z_{|m-1|}, Z_m, Z_{m+1}
        cplx zlo, z0, zp1;
        if (local) {
            cplx a, b;
            hankel01_fast<true, true>(x, a, b);
            const cplx h1v = b;
            const double tox = 2.0 / x;
            cplx hs0 = {0, 0}, hs1 = {0, 0}, hs2 = {0, 0};
            for (int j = 0; j <= top; ++j) {
                if (j == m - 1) hs0 = a;
                if (j == m) hs1 = a;
                if (j == m + 1) hs2 = a;
                const cplx nb = ((j + 1) * tox) * b - a;
                a = b;
                b = nb;
            }
            zlo = m >= 1 ? hs0 : h1v;
            z0 = hs1;
            zp1 = hs2;
        } else {
            double jlo = 0, j0v = 0, jp1 = 0;
            if (x < 1e-300) {
                j0v = m == 0 ? 1.0 : 0.0;
                jlo = m == 1 ? 1.0 : 0.0;
            } else {
                const double norm = miller_j(
                    x, top,
                    [&](int j, double f) {
                        if (j == m - 1 || (m == 0 && j == 1)) jlo = f;
                        if (j == m) j0v = f;
                        if (j == m + 1) jp1 = f;
                    },
                    [&](double sc) {
                        jlo *= sc;
                        j0v *= sc;
                        jp1 *= sc;
                    });
                jlo *= norm;
                j0v *= norm;
                jp1 *= norm;
            }
            zlo = {jlo, 0};
            z0 = {j0v, 0};
            zp1 = {jp1, 0};
        }
        const cplx em = cpow_int(e1, m);  // e^{-i m theta}
        const cplx emp1 = em * e1;        // e^{-i (m+1) theta}
        const cplx g0 = z0 * em;
        if (!kd.dip) {
            cp += q * g0;
            if (m > 0) cm += (sm * q) * (z0 * conj(em));
        } else {
            const cplx nu = ldpt(nrm, ns, s);
            const cplx hq = (0.5 * kd.k) * q;
            // g_{m-1}: for m = 0 it is g_{-1} = -Z_1 e^{+i theta}
            const cplx gm1 = m >= 1 ? zlo * (em * conj(e1)) : -(zlo * conj(e1));
            const cplx gp1 = zp1 * emp1;
            cp += hq * (conj(nu) * gm1 - nu * gp1);
            if (m > 0) {
                // n = -m: g_{-m-1} = (-1)^{m+1} Z_{m+1} e^{+i(m+1)th},
                //         g_{-m+1} = (-1)^{m-1} Z_{m-1} e^{+i(m-1)th}
                const cplx gA = (-sm) * (zp1 * conj(emp1));
                const cplx gB = (-sm) * (zlo * conj(em * conj(e1)));
                cm += hq * (conj(nu) * gA - nu * gB);
            }
        }
    }
}

__device__ cplx lap_p2e_coeff(const KD &kd, int local, int l, cplx c, double R,
                              const double *src, int ns, int s0, int s1, const double *str,
                              const double *nrm)
{
    cplx acc = {0, 0};
    for (int s = s0; s < s1; ++s) {
        const cplx w = ldpt(src, ns, s) - c;
        const cplx q = ldc(str + 2 * s);
        if (local) {
            const cplx pw = cpow_int(R * cinv(w), l);
            if (!kd.dip) {
                if (l == 0) acc += q * clogz(-w);
                else acc += (-1.0 / l) * (q * pw);
            } else {
                const cplx nu = ldpt(nrm, ns, s);
                acc += ((nu * q) * cinv(w)) * pw;
            }
        } else {
            const cplx t = (1.0 / R) * w;
            if (!kd.dip) {
                if (l == 0) acc += q;
                else acc += (-1.0 / l) * (q * cpow_int(t, l));
            } else if (l >= 1) {
                const cplx nu = ldpt(nrm, ns, s);
                acc += -((1.0 / R) * (nu * q)) * cpow_int(t, l - 1);
            }
        }
    }
    return acc;
}

// P2M over source leaf boxes (flat over levels, p+1 threads per box)
__global__ void k_p2m(KD kd, TD t, Flat fl, const int *boxes, const double *src, int ns,
                      const double *str, const double *nrm, double *mpole)
{
    int lv, i, j;
    if (!flat_locate(fl, blockIdx.x * (long)blockDim.x + threadIdx.x, lv, i, j)) return;
    const int b = boxes[i];
    const int s0 = t.sstart[b], s1 = s0 + t.scount[b];
    const cplx c = boxc(t, b);
    double *a = mpole + ecoef_off(kd, b, lv);
    if (kd.kernel == QBX_LAPLACE2D) {
        stc(a + 2 * j, lap_p2e_coeff(kd, 0, j, c, kd.rscale[lv], src, ns, s0, s1, str, nrm));
    } else {
        cplx cp, cm;
        helm_p2e_pm(kd, 0, j, c, src, ns, s0, s1, str, nrm, cp, cm);
        stc(a + 2 * j, cp);
        if (j > 0) stc(a - 2 * j, cm);
    }
}

// P2L from list 4 (flat over ttp levels, p+1 threads per box), accumulates
__global__ void k_p2l(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
                      const double *src, int ns, const double *str, const double *nrm,
                      double *local)
{
    int lv, i, j;
    if (!flat_locate(fl, blockIdx.x * (long)blockDim.x + threadIdx.x, lv, i, j)) return;
    if (starts[i] == starts[i + 1]) return;
    const int P = ttp[i];
    const cplx c = boxc(t, P);
    cplx acc = {0, 0}, accm = {0, 0};
    for (int e = starts[i]; e < starts[i + 1]; ++e) {
        const int Sb = lists[e];
        const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
        if (kd.kernel == QBX_LAPLACE2D) {
            acc += lap_p2e_coeff(kd, 1, j, c, kd.rscale[lv], src, ns, s0, s1, str, nrm);
        } else {
            cplx cp, cm;
            helm_p2e_pm(kd, 1, j, c, src, ns, s0, s1, str, nrm, cp, cm);
            acc += cp;
            accm += cm;
        }
    }
    double *a = local + ecoef_off(kd, P, lv);
    stc(a + 2 * j, ldc(a + 2 * j) + acc);
    if (kd.kernel != QBX_LAPLACE2D && j > 0) stc(a - 2 * j, ldc(a - 2 * j) + accm);
}

// -------------------------------------------------------------------------
// translation tables (Helmholtz)
// -------------------------------------------------------------------------
// M2L row (level l, class (dx,dy)): u_j(d) = H_j(k|d|) e^{i j arg d} with
// d = (dx, dy) h_l, j in [-2 p_l, 2 p_l], stored centred.
__global__ void k_tab_m2l(KD kd, int nl, double *tab)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = kd.m2l_w, ncls = (2 * W + 1) * (2 * W + 1);
    const int l = gid / ncls, cls = gid % ncls;
    if (l >= nl) return;
    const int L = 2 * kd.order[l];
    double *row = tab + 2 * (kd.m2l_off[l] + (long)cls * (2 * L + 1) + L);
    const int dx = cls / (2 * W + 1) - W, dy = cls % (2 * W + 1) - W;
    if (abs(dx) <= kd.nsep && abs(dy) <= kd.nsep) {
        for (int j = -L; j <= L; ++j) stc(row + 2 * j, {0, 0});
        return;
    }
    const double h = kd.rscale[l];
    const cplx d = {dx * h, dy * h};
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    const cplx e1 = {d.x / rho, d.y / rho};
    cplx a, b;
    hankel01_fast<true, true>(x, a, b);
    const double tox = 2.0 / x;
    cplx e = {1, 0};
    for (int j = 0; j <= L; ++j) {
        stc(row + 2 * j, a * e);
        if (j > 0) {
            const cplx w = a * conj(e);
            stc(row - 2 * j, (j & 1) ? -w : w);
        }
        const cplx nb = ((j + 1) * tox) * b - a;
        a = b;
        b = nb;
        e = e * e1;
    }
}

// M2M/L2L row (child level l, child index c): uhat_j(c_parent - c_child),
// j in [-(p_l + p_{l-1}), +], stored centred.
__global__ void k_tab_mm(KD kd, int nl, double *tab)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = 1 + gid / 4, c = gid % 4;
    if (l >= nl) return;
    const int L = kd.order[l] + kd.order[l - 1];
    double *row = tab + 2 * (kd.mm_off[l] + (long)c * (2 * L + 1) + L);
    const double h = kd.rscale[l];
    const cplx d = {(0.5 - (c & 1)) * h, (0.5 - ((c >> 1) & 1)) * h};
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    const cplx e1 = {d.x / rho, d.y / rho};
    for (int j = 0; j <= L; ++j) row[2 * j] = 0.0;
    const double norm = miller_j(
        x, L, [&](int j, double f) { row[2 * j] = f; },
        [&](double s) {
            for (int j = 0; j <= L; ++j) row[2 * j] *= s;
        });
    cplx e = {1, 0};
    for (int j = 0; j <= L; ++j) {
        const double J = row[2 * j] * norm;
        stc(row + 2 * j, J * e);
        if (j > 0) {
            const cplx w = J * conj(e);
            stc(row - 2 * j, (j & 1) ? -w : w);
        }
        e = e * e1;
    }
}

// -------------------------------------------------------------------------
// Laplace translations, one output coefficient
// -------------------------------------------------------------------------
__device__ __forceinline__ double g_binom(const double *B, int n, int k)
{
    return (k < 0 || k > n) ? 0.0 : B[n * BINOM_N + k];
}

// M2M (c1 -> c2, z0 = c1 - c2), output l:
// -a0 u^l / l + sum_{k=1}^{min(l,p1)} a_k r^k u^{l-k} C(l-1,k-1), u = z0/R2, r = R1/R2
__device__ cplx lap_m2m(const double *B, const double *a, int p1, int l, cplx z0, double R1,
                        double R2)
{
    if (l == 0) return ldc(a);
    const cplx u = (1.0 / R2) * z0;
    const double r = R1 / R2;
    const cplx acc = (-1.0 / l) * (ldc(a) * cpow_int(u, l));
    const int kmax = l < p1 ? l : p1;
    cplx hsum = {0, 0};
    double rk = 1.0;
    for (int k = 1; k <= kmax; ++k) {
        rk *= r;
        hsum = hsum * u + (rk * g_binom(B, l - 1, k - 1)) * ldc(a + 2 * k);
    }
    return acc + hsum * cpow_int(u, l - kmax);
}

// M2L (c1 multipole -> c2 local, z0 = c1 - c2), output l
__device__ cplx lap_m2l(const double *B, const double *a, int p1, int l, cplx z0, double R1,
                        double R2)
{
    const cplx iz = cinv(z0);
    const cplx ms = -(R1 * iz), t = R2 * iz;
    if (l == 0) {
        cplx acc = ldc(a) * clogz(-z0);
        cplx pw = {1, 0};
        for (int k = 1; k <= p1; ++k) {
            pw = pw * ms;
            acc += ldc(a + 2 * k) * pw;
        }
        return acc;
    }
    cplx acc = (-1.0 / l) * ldc(a);
    cplx pw = {1, 0};
    for (int k = 1; k <= p1; ++k) {
        pw = pw * ms;
        acc += g_binom(B, l + k - 1, k - 1) * (ldc(a + 2 * k) * pw);
    }
    return acc * cpow_int(t, l);
}

// L2L (c1 -> c2, z0 = c2 - c1), output j
__device__ cplx lap_l2l(const double *B, const double *a, int p1, int j, cplx z0, double R1,
                        double R2)
{
    if (j > p1) return {0, 0};
    const cplx u = (1.0 / R1) * z0;
    cplx acc = {0, 0};
    for (int l = p1; l >= j; --l) acc = acc * u + g_binom(B, l, j) * ldc(a + 2 * l);
    double rp = 1.0;
    const double r21 = R2 / R1;
    for (int i = 0; i < j; ++i) rp *= r21;
    return rp * acc;
}

// -------------------------------------------------------------------------
// M2M (parents at one level), M2L (list 2), L2L (one level)
// -------------------------------------------------------------------------
__global__ void k_m2m(KD kd, TD t, const double *B, int lvp, int i0, int i1, const int *spb,
                      double *mpole)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int ncp = kd.ncl[lvp];
    const int i = i0 + (int)(gid / ncp), j = (int)(gid % ncp);
    if (i >= i1) return;
    const int P = spb[i];
    const int lvc = lvp + 1;
    const int p1 = kd.order[lvc], p2 = kd.order[lvp];
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
    const int m = helm ? j - p2 : j;
    const cplx cp = boxc(t, P);
    cplx acc = {0, 0};
    for (int c = 0; c < 4; ++c) {
        const int C = t.child[(size_t)c * t.nboxes + P];
        if (C < 0 || t.scumul[C] == 0) continue;
        const double *a = mpole + ecoef_off(kd, C, lvc);
        if (!helm) {
            acc += lap_m2m(B, a, p1, m, boxc(t, C) - cp, kd.rscale[lvc], kd.rscale[lvp]);
        } else {
            const int L = p1 + p2;
            const double *row = kd.tab_mm + 2 * (kd.mm_off[lvc] + (long)c * (2 * L + 1) + L);
            cplx s = {0, 0};
            for (int n = -p1; n <= p1; ++n) s += ldc(a + 2 * n) * ldc(row + 2 * (n - m));
            acc += s;
        }
    }
    double *o = mpole + ecoef_off(kd, P, lvp) + 2 * m;
    stc(o, ldc(o) + acc);
}

__global__ void k_m2l(KD kd, TD t, const double *B, Flat fl, const int *ttp, const int *starts,
                      const int *lists, const double *mpole, double *local)
{
    int lv, i, j;
    if (!flat_locate(fl, blockIdx.x * (long)blockDim.x + threadIdx.x, lv, i, j)) return;
    if (starts[i] == starts[i + 1]) return;
    const int P = ttp[i];
    const int p = kd.order[lv];
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
    const int m = helm ? j - p : j;
    const cplx cp = boxc(t, P);
    const int Ix = t.coords[P], Iy = t.coords[(size_t)t.nboxes + P];
    const int L = 2 * p;
    cplx acc = {0, 0};
    for (int e = starts[i]; e < starts[i + 1]; ++e) {
        const int A = lists[e];
        const double *a = mpole + ecoef_off(kd, A, lv);
        if (!helm) {
            acc += lap_m2l(B, a, p, m, boxc(t, A) - cp, kd.rscale[lv], kd.rscale[lv]);
        } else {
            const int dx = Ix - t.coords[A], dy = Iy - t.coords[(size_t)t.nboxes + A];
            const int cls = (dx + kd.m2l_w) * (2 * kd.m2l_w + 1) + (dy + kd.m2l_w);
            const double *row = kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
            cplx s = {0, 0};
            for (int n = -p; n <= p; ++n) s += ldc(a + 2 * n) * ldc(row + 2 * (n - m));
            acc += s;
        }
    }
    double *o = local + ecoef_off(kd, P, lv) + 2 * m;
    stc(o, ldc(o) + acc);
}

__global__ void k_l2l(KD kd, TD t, const double *B, int lv, int i0, int i1, const int *ttp,
                      double *local)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int ncp = kd.ncl[lv];
    const int i = i0 + (int)(gid / ncp), j = (int)(gid % ncp);
    if (i >= i1) return;
    const int P = ttp[i];
    const int Q = t.parent[P];
    if (Q < 0) return;
    const int p2 = kd.order[lv], p1 = kd.order[lv - 1];
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
    const int m = helm ? j - p2 : j;
    const double *a = local + ecoef_off(kd, Q, lv - 1);
    cplx v;
    if (!helm) {
        v = lap_l2l(B, a, p1, m, boxc(t, P) - boxc(t, Q), kd.rscale[lv - 1], kd.rscale[lv]);
    } else {
        // uhat_j(c_P - c_Q) = (-1)^j uhat_j(c_Q - c_P): the M2M row with sign
        const int c = (t.coords[P] & 1) | ((t.coords[(size_t)t.nboxes + P] & 1) << 1);
        const int L = p1 + p2;
        const double *row = kd.tab_mm + 2 * (kd.mm_off[lv] + (long)c * (2 * L + 1) + L);
        v = {0, 0};
        for (int n = -p1; n <= p1; ++n) {
            const cplx term = ldc(a + 2 * n) * ldc(row + 2 * (n - m));
            v += ((n - m) & 1) ? -term : term;
        }
    }
    double *o = local + ecoef_off(kd, P, lv) + 2 * m;
    stc(o, ldc(o) + v);
}

// -------------------------------------------------------------------------
// evaluation at points; a: coefficient 0 (Helmholtz n at a[2n])
// -------------------------------------------------------------------------
__device__ cplx eval_local(const KD &kd, const double *a, int p, cplx c, double R, cplx tp)
{
    if (kd.kernel == QBX_LAPLACE2D) {
        const cplx x = (1.0 / R) * (tp - c);
        cplx phi = {0, 0};
        for (int l = p; l >= 0; --l) phi = phi * x + ldc(a + 2 * l);
        return {-phi.x * INV2PI, 0.0};
    }
    const cplx d = tp - c;
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    cplx phi;
    if (x < 1e-300) {
        phi = ldc(a);
    } else {
        const cplx e1 = {d.x / rho, d.y / rho};
        const cplx em1 = conj(e1);
        cplx e = cpow_int(e1, p);
        cplx acc = {0, 0};
        const double norm = miller_j(
            x, p,
            [&](int n, double f) {
                cplx v = ldc(a + 2 * n) * e;
                if (n > 0) {
                    const cplx w = ldc(a - 2 * n) * conj(e);
                    v += (n & 1) ? -w : w;
                }
                acc += f * v;
                e = e * em1;
            },
            [&](double s) { acc = s * acc; });
        phi = norm * acc;
    }
    return {-0.25 * phi.y, 0.25 * phi.x};
}

__device__ cplx eval_mpole(const KD &kd, const double *a, int p, cplx c, double R, cplx tp)
{
    const cplx w = tp - c;
    if (kd.kernel == QBX_LAPLACE2D) {
        const cplx x = R * cinv(w);
        cplx phi = {0, 0};
        for (int k = p; k >= 1; --k) phi = (phi + ldc(a + 2 * k)) * x;
        phi += ldc(a) * clogz(w);
        return {-phi.x * INV2PI, 0.0};
    }
    const double rho = sqrt(abs2(w));
    const double x = kd.k * rho;
    const cplx e1 = {w.x / rho, w.y / rho};
    cplx e = {1, 0};
    cplx hm1, hm;
    hankel01_fast<true, true>(x, hm1, hm);
    const double tox = 2.0 / x;
    cplx acc = ldc(a) * hm1;
    for (int n = 1; n <= p; ++n) {
        cplx h;
        if (n == 1) {
            h = hm;
        } else {
            h = ((n - 1) * tox) * hm - hm1;
            hm1 = hm;
            hm = h;
        }
        e = e * e1;
        cplx v = ldc(a + 2 * n) * e;
        const cplx wv = ldc(a - 2 * n) * conj(e);
        v += (n & 1) ? -wv : wv;
        acc += h * v;
    }
    return {-0.25 * acc.y, 0.25 * acc.x};
}

// One 64-lane block per target box.  P2P and M2P give each of the box's
// non-QBX targets a group of G lanes (G a power of two, G * ntargets <= 64)
// that splits the list-1 sources / list-3 boxes and is reduced by shuffles,
// so boxes with few targets still keep the wavefront busy.
// MODE 0: P2P (list 1), 1: M2P (list 3), 2: L2P (one lane per target)
template <int MODE>
__global__ void __launch_bounds__(64) k_eval_targets(KD kd, TD t, int ntb, const int *tboxes,
                                                     const int *starts, const int *lists,
                                                     const int *fstart, const int *fcount,
                                                     const int *unfilt, const double *tgt,
                                                     int nt, const double *src, int ns,
                                                     const double *str, const double *nrm,
                                                     const double *expn, double *pot)
{
    const int it = blockIdx.x;
    if (it >= ntb) return;
    const int B = tboxes[it];
    const int f0 = fstart[B], nf = fcount[B];
    if (nf == 0) return;
    const int lane = threadIdx.x;
    if (MODE == 2) {
        const int lv = t.level[B];
        for (int f = f0 + lane; f < f0 + nf; f += 64) {
            const int ti = unfilt[f];
            const cplx v = eval_local(kd, expn + ecoef_off(kd, B, lv), kd.order[lv], boxc(t, B),
                                      kd.rscale[lv], ldpt(tgt, nt, ti));
            double *o = pot + 2 * (size_t)ti;
            stc(o, ldc(o) + v);
        }
        return;
    }
    int G = 64;
    while (G > 1 && G * nf > 64) G >>= 1;
    const int tpr = 64 / G;
    const int tsub = lane / G, sub = lane % G;
    const int e0 = starts[it], e1 = starts[it + 1];
    for (int base = 0; base < nf; base += tpr) {
        const bool active = base + tsub < nf;
        int ti = 0;
        cplx acc = {0, 0};
        if (active) {
            ti = unfilt[f0 + base + tsub];
            const cplx tp = ldpt(tgt, nt, ti);
            if (MODE == 0) {
                for (int e = e0; e < e1; ++e) {
                    const int Sb = lists[e];
                    const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
                    for (int s = s0 + sub; s < s1; s += G) {
                        const cplx nu = kd.dip ? ldpt(nrm, ns, s) : cplx{0, 0};
                        acc += p2p_pair(kd, ldpt(src, ns, s), ldc(str + 2 * s), nu, tp);
                    }
                }
            } else {
                for (int e = e0 + sub; e < e1; e += G) {
                    const int D = lists[e];
                    const int lv = t.level[D];
                    acc += eval_mpole(kd, expn + ecoef_off(kd, D, lv), kd.order[lv], boxc(t, D),
                                      kd.rscale[lv], tp);
                }
            }
        }
        for (int m = G / 2; m >= 1; m >>= 1) {
            acc.x += __shfl_xor(acc.x, m);
            acc.y += __shfl_xor(acc.y, m);
        }
        if (active && sub == 0) {
            double *o = pot + 2 * (size_t)ti;
            stc(o, ldc(o) + acc);
        }
    }
}

// -------------------------------------------------------------------------
// QBX stages
// -------------------------------------------------------------------------
template <int PQ>
struct QAcc {
    cplx v[2 * PQ + 1];
};

// A holds Laplace coefficients l at v[l], Helmholtz n at v[n + PQ]
template <int PQ>
__device__ void qbx_store(const KD &kd, double *qexp, int c, const QAcc<PQ> &A, int accumulate)
{
    const int nq = kd.kernel == QBX_LAPLACE2D ? PQ + 1 : 2 * PQ + 1;
    double *o = qexp + 2 * (size_t)c * kd.ncq;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) {
        if (i < nq) {
            cplx v = A.v[i];
            if (accumulate) v += ldc(o + 2 * i);
            stc(o + 2 * i, v);
        }
    }
}

constexpr int P2Q_LANES = 16;

__device__ __forceinline__ double shfl_xor_d(double v, int m)
{
    return __shfl_xor(v, m, P2Q_LANES);
}

// P2QBXL: P2Q_LANES lanes per global center split the list-1 sources of the
// center's target box; the order-PQ expansion stays in registers and is
// reduced across the lane group with shuffles.  No atan2/sincos: the phase
// e^{-i theta} is conj(w)/|w|.
template <int PQ>
__global__ void __launch_bounds__(256) k_p2qbxl(KD kd, TD t, int ng, const int *gcent,
                                                const int *c2tb, const double *ctr,
                                                const double *rad, int nc, const int *l1s,
                                                const int *l1l, const double *src, int ns,
                                                const double *str, const double *nrm,
                                                double *qexp, int accumulate)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = gid / P2Q_LANES, lane = gid % P2Q_LANES;
    const bool active = g < ng;
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    int c = 0;
    if (active) {
        c = gcent[g];
        const int it = c2tb[c];
        const cplx cc = ldpt(ctr, nc, c);
        const double R = rad[c];
        const int e0 = l1s[it], e1 = l1s[it + 1];
        for (int e = e0; e < e1; ++e) {
            const int Sb = l1l[e];
            const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
            for (int s = s0 + lane; s < s1; s += P2Q_LANES) {
                const cplx w = ldpt(src, ns, s) - cc;
                const cplx q = ldc(str + 2 * s);
                if (kd.kernel == QBX_LAPLACE2D) {
                    const cplx iw = cinv(w);
                    const cplx tt = R * iw;
                    cplx pw = {1, 0};
                    if (!kd.dip) {
                        A.v[0] += q * clogz(-w);
#pragma unroll
                        for (int l = 1; l <= PQ; ++l) {
                            pw = pw * tt;
                            A.v[l] += (-1.0 / l) * (q * pw);
                        }
                    } else {
                        const cplx f = (ldpt(nrm, ns, s) * q) * iw;
                        A.v[0] += f;
#pragma unroll
                        for (int l = 1; l <= PQ; ++l) {
                            pw = pw * tt;
                            A.v[l] += f * pw;
                        }
                    }
                } else {
                    const double r = sqrt(abs2(w));
                    const double x = kd.k * r;
                    const double ir = 1.0 / r;
                    const cplx em = {w.x * ir, -w.y * ir};  // e^{-i theta}
                    cplx H[PQ + 2];
                    hankel01_fast<true, true>(x, H[0], H[1]);
                    const double tox = 2.0 / x;
#pragma unroll
                    for (int m = 2; m <= PQ + 1; ++m)
                        H[m] = ((m - 1) * tox) * H[m - 1] - H[m - 2];
                    // G[m + PQ + 1] = H_m e^{-i m theta}, m in [-(PQ+1), PQ+1]
                    cplx G[2 * PQ + 3];
                    cplx e = {1, 0};
#pragma unroll
                    for (int m = 0; m <= PQ + 1; ++m) {
                        G[PQ + 1 + m] = H[m] * e;
                        const cplx gm = H[m] * conj(e);
                        G[PQ + 1 - m] = (m & 1) ? -gm : gm;
                        e = e * em;
                    }
                    if (!kd.dip) {
#pragma unroll
                        for (int n = -PQ; n <= PQ; ++n) A.v[n + PQ] += q * G[n + PQ + 1];
                    } else {
                        const cplx nu = ldpt(nrm, ns, s);
                        const cplx hq = (0.5 * kd.k) * q;
                        const cplx a1 = hq * conj(nu), a2 = hq * nu;
#pragma unroll
                        for (int n = -PQ; n <= PQ; ++n)
                            A.v[n + PQ] += a1 * G[n + PQ] - a2 * G[n + PQ + 2];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) {
#pragma unroll
        for (int m = P2Q_LANES / 2; m >= 1; m >>= 1) {
            A.v[i].x += shfl_xor_d(A.v[i].x, m);
            A.v[i].y += shfl_xor_d(A.v[i].y, m);
        }
    }
    if (active && lane == 0) qbx_store<PQ>(kd, qexp, c, A, accumulate);
}

constexpr int M2Q_LANES = 4;

// M2QBXL: M2Q_LANES lanes per global center split its list-3 boxes
template <int PQ>
__global__ void __launch_bounds__(256) k_m2qbxl(KD kd, TD t, const double *B, int ng,
                                                const int *gcent, const int *c2tb,
                                                const double *ctr, const double *rad, int nc,
                                                const int *l3s, const int *l3l,
                                                const double *mpole, double *qexp,
                                                int accumulate)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = gid / M2Q_LANES, sub = gid % M2Q_LANES;
    const bool active = g < ng;
    const int c = active ? gcent[g] : 0;
    const int it = active ? c2tb[c] : 0;
    const cplx cc = active ? ldpt(ctr, nc, c) : cplx{0, 0};
    const double R2 = active ? rad[c] : 1.0;
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    const int ebeg = active ? l3s[it] + sub : 0, eend = active ? l3s[it + 1] : 0;
    for (int e = ebeg; e < eend; e += M2Q_LANES) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int p1 = kd.order[lv];
        const double *a = mpole + ecoef_off(kd, D, lv);
        const cplx cd = boxc(t, D);
        if (kd.kernel == QBX_LAPLACE2D) {
            const cplx z0 = cd - cc;
            const cplx iz = cinv(z0);
            const cplx ms = -(kd.rscale[lv] * iz), tt = R2 * iz;
            const cplx a0 = ldc(a);
            cplx Sv[PQ + 1];
            Sv[0] = a0 * clogz(-z0);
#pragma unroll
            for (int l = 1; l <= PQ; ++l) Sv[l] = (-1.0 / l) * a0;
            cplx pw = {1, 0};
            for (int k = 1; k <= p1; ++k) {
                pw = pw * ms;
                const cplx ak = ldc(a + 2 * k) * pw;
                Sv[0] += ak;
#pragma unroll
                for (int l = 1; l <= PQ; ++l) Sv[l] += g_binom(B, l + k - 1, k - 1) * ak;
            }
            cplx tl = {1, 0};
#pragma unroll
            for (int l = 0; l <= PQ; ++l) {
                A.v[l] += Sv[l] * tl;
                tl = tl * tt;
            }
        } else {
            const cplx d = cc - cd;
            const double rho = sqrt(abs2(d));
            const double x = kd.k * rho;
            const cplx e1 = {d.x / rho, d.y / rho};
            cplx ep = {1, 0};
            cplx hm1, hm;
            hankel01_fast<true, true>(x, hm1, hm);
            const double tox = 2.0 / x;
            const int lmax = p1 + PQ;
            cplx h = hm1;
            for (int l = 0; l <= lmax; ++l) {
                if (l == 1) {
                    h = hm;
                } else if (l >= 2) {
                    h = ((l - 1) * tox) * hm - hm1;
                    hm1 = hm;
                    hm = h;
                }
                const cplx he = h * ep;
                cplx hc = h * conj(ep);
                if (l & 1) hc = -hc;
#pragma unroll
                for (int m = -PQ; m <= PQ; ++m) {
                    const int n1 = m + l;
                    if (n1 >= -p1 && n1 <= p1) A.v[m + PQ] += ldc(a + 2 * n1) * he;
                    const int n2 = m - l;
                    if (l > 0 && n2 >= -p1 && n2 <= p1) A.v[m + PQ] += ldc(a + 2 * n2) * hc;
                }
                ep = ep * e1;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) {
#pragma unroll
        for (int m = M2Q_LANES / 2; m >= 1; m >>= 1) {
            A.v[i].x += __shfl_xor(A.v[i].x, m, M2Q_LANES);
            A.v[i].y += __shfl_xor(A.v[i].y, m, M2Q_LANES);
        }
    }
    if (active && sub == 0) qbx_store<PQ>(kd, qexp, c, A, accumulate);
}

template <int PQ>
__global__ void k_l2qbxl(KD kd, TD t, const double *B, int ng, const int *gcent, const int *c2tb,
                         const int *tboxes, const double *ctr, const double *rad, int nc,
                         const double *local, double *qexp, int accumulate)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int c = gcent[g];
    const int Bx = tboxes[c2tb[c]];
    const int lv = t.level[Bx];
    const int p1 = kd.order[lv];
    const cplx cc = ldpt(ctr, nc, c);
    const double R2 = rad[c];
    const double *a = local + ecoef_off(kd, Bx, lv);
    const cplx cb = boxc(t, Bx);
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    if (kd.kernel == QBX_LAPLACE2D) {
        const double R1 = kd.rscale[lv];
        const cplx u = (1.0 / R1) * (cc - cb);
        for (int l = p1; l >= 0; --l) {
            const cplx al = ldc(a + 2 * l);
#pragma unroll
            for (int j = 0; j <= PQ; ++j)
                if (j <= l) A.v[j] = A.v[j] * u + g_binom(B, l, j) * al;
        }
        double rp = 1.0;
        const double r21 = R2 / R1;
#pragma unroll
        for (int j = 0; j <= PQ; ++j) {
            A.v[j] = rp * A.v[j];
            rp *= r21;
        }
    } else {
        const cplx d = cc - cb;
        const double rho = sqrt(abs2(d));
        const double x = kd.k * rho;
        const int lmax = p1 + PQ;
        if (x < 1e-300) {
#pragma unroll
            for (int m = -PQ; m <= PQ; ++m)
                if (abs(m) <= p1) A.v[m + PQ] = ldc(a + 2 * m);
        } else {
            const cplx e1 = {d.x / rho, d.y / rho};
            const cplx em1 = conj(e1);
            cplx e = cpow_int(e1, lmax);
            const double norm = miller_j(
                x, lmax,
                [&](int l, double f) {
                    const cplx fe = f * e;
                    cplx fc = f * conj(e);
                    if (l & 1) fc = -fc;
#pragma unroll
                    for (int m = -PQ; m <= PQ; ++m) {
                        const int n1 = m + l;
                        if (n1 >= -p1 && n1 <= p1) A.v[m + PQ] += ldc(a + 2 * n1) * fe;
                        const int n2 = m - l;
                        if (l > 0 && n2 >= -p1 && n2 <= p1) A.v[m + PQ] += ldc(a + 2 * n2) * fc;
                    }
                    e = e * em1;
                },
                [&](double s) {
#pragma unroll
                    for (int m = 0; m < 2 * PQ + 1; ++m) A.v[m] = s * A.v[m];
                });
#pragma unroll
            for (int m = 0; m < 2 * PQ + 1; ++m) A.v[m] = norm * A.v[m];
        }
    }
    qbx_store<PQ>(kd, qexp, c, A, accumulate);
}

// QBXL2P: one thread per (center, target) entry of center_to_tree_targets
__global__ void k_qbxl2p(KD kd, int nctt, const int *ctt_lists, const int *tree_ttc,
                         const double *tgt, int nt, const double *ctr, const double *rad, int nc,
                         const double *qexp, double *pot)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nctt) return;
    const int ti = ctt_lists[j];
    const int c = tree_ttc[ti];
    const cplx tp = ldpt(tgt, nt, ti);
    const cplx cc = ldpt(ctr, nc, c);
    const double *a =
        qexp + 2 * ((size_t)c * kd.ncq + (kd.kernel == QBX_HELMHOLTZ2D ? kd.pq : 0));
    stc(pot + 2 * (size_t)ti, eval_local(kd, a, kd.pq, cc, rad[c], tp));
}

__global__ void k_finalize(int nt, int nc, const int *user_to_tree, const double *pot,
                           double *out)
{
    const int u = nc + blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nt) return;
    const int ti = user_to_tree[u];
    out[2 * (size_t)(u - nc)] = pot[2 * (size_t)ti];
    out[2 * (size_t)(u - nc) + 1] = pot[2 * (size_t)ti + 1];
}

// -------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------
static double *g_binom_dev[64] = {nullptr};

static int binom_table(const double **out)
{
    int dev = 0;
    QBX_HIP(hipGetDevice(&dev));
    QBX_CHECK(dev >= 0 && dev < 64, "device index out of range");
    if (!g_binom_dev[dev]) {
        std::vector<double> h((size_t)BINOM_N * BINOM_N, 0.0);
        for (int n = 0; n < BINOM_N; ++n) {
            h[(size_t)n * BINOM_N] = 1.0;
            for (int k = 1; k <= n; ++k)
                h[(size_t)n * BINOM_N + k] = h[(size_t)(n - 1) * BINOM_N + k - 1] +
                                             (k < n ? h[(size_t)(n - 1) * BINOM_N + k] : 0.0);
        }
        double *p;
        QBX_HIP(hipMalloc(&p, sizeof(double) * h.size()));
        QBX_HIP(hipMemcpy(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        g_binom_dev[dev] = p;
    }
    *out = g_binom_dev[dev];
    return QBX_OK;
}

static int make_kd(const qbx_kernel_desc *d, const qbx_tree *T, KD *kd, int nsep = 2)
{
    QBX_CHECK(d, "null kernel description");
    QBX_CHECK(d->kernel == QBX_LAPLACE2D || d->kernel == QBX_HELMHOLTZ2D,
              "unsupported kernel %d", d->kernel);
    QBX_CHECK(d->source_kind == QBX_SOURCE_S || d->source_kind == QBX_SOURCE_D,
              "unsupported source kind %d", d->source_kind);
    QBX_CHECK(d->qbx_order >= 0 && d->qbx_order <= 12, "qbx_order must be in [0, 12]");
    QBX_CHECK(d->level_orders && d->nlevels >= 1 && d->nlevels <= FML,
              "level_orders missing or nlevels outside [1, %d]", FML);
    if (T) {
        QBX_CHECK(T->dim == 2, "2D kernels need a 2D tree");
        QBX_CHECK(T->nlevels <= FML, "tree has too many levels (%d)", T->nlevels);
        QBX_CHECK(d->nlevels >= T->nlevels, "level_orders has %d entries, tree has %d levels",
                  d->nlevels, T->nlevels);
    }
    if (d->kernel == QBX_HELMHOLTZ2D) QBX_CHECK(d->k > 0, "Helmholtz needs k > 0");
    memset(kd, 0, sizeof(*kd));
    kd->kernel = d->kernel;
    kd->dip = d->source_kind == QBX_SOURCE_D;
    kd->pq = d->qbx_order;
    kd->k = d->k;
    kd->nsep = nsep;
    kd->m2l_w = 2 * nsep + 1;
    const int ncls = (2 * kd->m2l_w + 1) * (2 * kd->m2l_w + 1);
    const int nl = T ? T->nlevels : d->nlevels;
    kd->nlevels = nl;
    const bool helm = kd->kernel == QBX_HELMHOLTZ2D;
    long coef = 0, m2l = 0, mm = 0;
    for (int l = 0; l < nl; ++l) {
        const int p = d->level_orders[l];
        QBX_CHECK(p >= 1 && p <= MAXP, "fmm order %d at level %d outside [1, %d]", p, l, MAXP);
        kd->order[l] = p;
        kd->ncl[l] = helm ? 2 * p + 1 : p + 1;
        kd->rscale[l] = T ? ldexp(T->root_extent, -l) : 1.0;
        kd->lbox[l] = T ? T->level_start_box_nrs[l] : 0;
        kd->lcoef[l] = coef;
        if (T)
            coef += (long)(T->level_start_box_nrs[l + 1] - T->level_start_box_nrs[l]) * kd->ncl[l];
        kd->m2l_off[l] = m2l;
        m2l += (long)ncls * (4 * p + 1);
        kd->mm_off[l] = mm;
        if (l >= 1) mm += 4L * (2 * (p + d->level_orders[l - 1]) + 1);
    }
    kd->lbox[nl] = T ? T->nboxes : 0;
    kd->lcoef[nl] = coef;
    kd->ncq = helm ? 2 * kd->pq + 1 : kd->pq + 1;
    return QBX_OK;
}

// Helmholtz translation tables, cached in the context per (tree, k, orders)
static int ensure_tables(qbx_ctx *ctx, hipStream_t st, const qbx_tree *T, KD *kd)
{
    if (kd->kernel != QBX_HELMHOLTZ2D) return QBX_OK;
    std::vector<double> key;
    key.push_back((double)(uintptr_t)T);
    key.push_back(kd->nsep);
    key.push_back(T->nboxes);
    key.push_back(T->root_extent);
    key.push_back(kd->k);
    for (int l = 0; l < kd->nlevels; ++l) key.push_back(kd->order[l]);
    long nm2l = 0, nmm = 0;
    const int ncls = (2 * kd->m2l_w + 1) * (2 * kd->m2l_w + 1);
    for (int l = 0; l < kd->nlevels; ++l) nm2l += (long)ncls * (4 * kd->order[l] + 1);
    for (int l = 1; l < kd->nlevels; ++l) nmm += 4L * (2 * (kd->order[l] + kd->order[l - 1]) + 1);
    if (key != ctx->tab_key) {
        if (ctx->ws_tab_m2l.n < 2 * nm2l) QBX_TRY(ctx->ws_tab_m2l.alloc(2 * nm2l));
        if (ctx->ws_tab_mm.n < 2 * std::max(nmm, 1L))
            QBX_TRY(ctx->ws_tab_mm.alloc(2 * std::max(nmm, 1L)));
        const int n1 = kd->nlevels * ncls;
        hipLaunchKernelGGL(k_tab_m2l, dim3(cdiv(n1, 64)), dim3(64), 0, st, *kd, kd->nlevels,
                           ctx->ws_tab_m2l.p);
        const int n2 = 4 * std::max(kd->nlevels - 1, 0);
        if (n2 > 0)
            hipLaunchKernelGGL(k_tab_mm, dim3(cdiv(n2, 64)), dim3(64), 0, st, *kd, kd->nlevels,
                               ctx->ws_tab_mm.p);
        QBX_HIP(hipGetLastError());
        ctx->tab_key = key;
    }
    kd->tab_m2l = ctx->ws_tab_m2l.p;
    kd->tab_mm = ctx->ws_tab_mm.p;
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

// thread layout over a level-sorted box list: p+1 (half) or ncl threads per box
static Flat make_flat(const KD &kd, const std::vector<int> &level_starts, bool half)
{
    Flat f;
    memset(&f, 0, sizeof(f));
    f.nl = kd.nlevels;
    long t = 0;
    for (int l = 0; l < kd.nlevels; ++l) {
        f.tstart[l] = t;
        f.istart[l] = level_starts[l];
        f.nper[l] = half ? kd.order[l] + 1 : kd.ncl[l];
        t += (long)(level_starts[l + 1] - level_starts[l]) * f.nper[l];
    }
    f.tstart[kd.nlevels] = t;
    return f;
}

#define DISPATCH_PQ(pq, KERNEL, ...)                                                    \
    switch (pq) {                                                                        \
    case 0: hipLaunchKernelGGL(KERNEL<0>, __VA_ARGS__); break;                           \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                           \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                           \
    case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                           \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                           \
    case 5: hipLaunchKernelGGL(KERNEL<5>, __VA_ARGS__); break;                           \
    case 6: hipLaunchKernelGGL(KERNEL<6>, __VA_ARGS__); break;                           \
    case 7: hipLaunchKernelGGL(KERNEL<7>, __VA_ARGS__); break;                           \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                           \
    case 9: hipLaunchKernelGGL(KERNEL<9>, __VA_ARGS__); break;                           \
    case 10: hipLaunchKernelGGL(KERNEL<10>, __VA_ARGS__); break;                         \
    case 11: hipLaunchKernelGGL(KERNEL<11>, __VA_ARGS__); break;                         \
    case 12: hipLaunchKernelGGL(KERNEL<12>, __VA_ARGS__); break;                         \
    default: set_error("qbx_order out of range"); return QBX_ERR_INVALID;               \
    }

static inline long exp_bytes(const KD &kd) { return 16L * kd.lcoef[kd.nlevels]; }

}  // namespace qbx

using namespace qbx;

extern "C" {

long qbx_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T)
{
    if (d && d->kernel == QBX_LAPLACE3D) return f3_expansion_size(d, T);
    KD kd;
    if (!T || make_kd(d, T, &kd) != QBX_OK) return -1;
    return kd.lcoef[kd.nlevels];
}

int qbx_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets)
{
    QBX_CHECK(T && offsets, "qbx_expansion_level_offsets: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_expansion_level_offsets(d, T, offsets);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    for (int l = 0; l <= kd.nlevels; ++l) offsets[l] = kd.lcoef[l];
    return QBX_OK;
}

int qbx_ncoeff_qbx(const qbx_kernel_desc *d)
{
    if (d && d->kernel == QBX_LAPLACE3D) return f3_ncoeff_qbx(d);
    KD kd;
    if (make_kd(d, nullptr, &kd) != QBX_OK) return -1;
    return kd.ncq;
}

int qbx_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_kernel_desc *d, const double *str, const double *nrm,
                        double *mpole)
{
    QBX_CHECK(ctx && T && R && str && mpole, "qbx_form_multipoles: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_form_multipoles(ctx, stream, T, R, d, str, nrm, mpole);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    QBX_HIP(hipMemsetAsync(mpole, 0, exp_bytes(kd), st));
    const Flat fl = make_flat(kd, R->h_level_start_source_box_nrs, true);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0)
        hipLaunchKernelGGL(k_p2m, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), fl,
                           R->source_boxes.p, T->sources.p, T->nsources, str, nrm, mpole);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                           const qbx_kernel_desc *d, double *mpole)
{
    QBX_CHECK(ctx && T && R && mpole, "qbx_coarsen_multipoles: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_coarsen_multipoles(ctx, stream, T, R, d, mpole);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    for (int l = T->nlevels - 2; l >= 0; --l) {
        const int i0 = R->h_level_start_source_parent_box_nrs[l];
        const int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr > 0)
            hipLaunchKernelGGL(k_m2m, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B,
                               l, i0, i1, R->source_parent_boxes.p, mpole);
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

static int eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_geo *G, const qbx_kernel_desc *d, int mode,
                        const double *str, const double *nrm, const double *expn, double *pot,
                        int accumulate)
{
    QBX_CHECK(ctx && T && R && G && pot, "stage: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_eval_targets(ctx, stream, T, R, G, d, mode, str, nrm, expn, pot, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(pot, 0, sizeof(double) * 2 * (size_t)T->ntargets, st));
    const int ntb = R->ntarget_boxes;
    if (ntb == 0 || G->nfiltered == 0) return QBX_OK;
    const TD td = make_td(T);
    const int *starts = mode == 1 ? R->l3_starts.p : R->l1_starts.p;
    const int *lists = mode == 1 ? R->l3_lists.p : R->l1_lists.p;
#define EVAL_LAUNCH(M)                                                                     \
    hipLaunchKernelGGL(k_eval_targets<M>, dim3(ntb), dim3(64), 0, st, kd, td, ntb,         \
                       R->target_boxes.p, starts, lists, G->nqbtl_box_target_starts.p,    \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm,    \
                       expn, pot)
    if (mode == 0) EVAL_LAUNCH(0);
    else if (mode == 1) EVAL_LAUNCH(1);
    else EVAL_LAUNCH(2);
#undef EVAL_LAUNCH
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_eval_direct(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                    const double *nrm, double *pot, int accumulate)
{
    QBX_CHECK(str, "qbx_eval_direct: null strengths");
    if (d && d->source_kind == QBX_SOURCE_D) QBX_CHECK(nrm, "double layer needs normals");
    return eval_targets(ctx, stream, T, R, G, d, 0, str, nrm, nullptr, pot, accumulate);
}

int qbx_eval_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_geo *G, const qbx_kernel_desc *d, const double *mpole,
                        double *pot, int accumulate)
{
    QBX_CHECK(mpole, "qbx_eval_multipoles: null multipoles");
    return eval_targets(ctx, stream, T, R, G, d, 1, nullptr, nullptr, mpole, pot, accumulate);
}

int qbx_eval_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, const double *local,
                    double *pot, int accumulate)
{
    QBX_CHECK(local, "qbx_eval_locals: null locals");
    return eval_targets(ctx, stream, T, R, G, d, 2, nullptr, nullptr, local, pot, accumulate);
}

int qbx_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                           const qbx_kernel_desc *d, const double *mpole, double *local)
{
    QBX_CHECK(ctx && T && R && mpole && local, "qbx_multipole_to_local: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_multipole_to_local(ctx, stream, T, R, d, mpole, local);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    QBX_HIP(hipMemsetAsync(local, 0, exp_bytes(kd), st));
    const Flat fl = make_flat(kd, R->h_level_start_ttp_box_nrs, false);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0 && R->n2 > 0)
        hipLaunchKernelGGL(k_m2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B, fl,
                           R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_kernel_desc *d, const double *str, const double *nrm,
                    double *local, int accumulate)
{
    QBX_CHECK(ctx && T && R && str && local, "qbx_form_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_form_locals(ctx, stream, T, R, d, str, nrm, local, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate) QBX_HIP(hipMemsetAsync(local, 0, exp_bytes(kd), st));
    const Flat fl = make_flat(kd, R->h_level_start_ttp_box_nrs, true);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0 && R->n4 > 0)
        hipLaunchKernelGGL(k_p2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), fl,
                           R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p, T->sources.p,
                           T->nsources, str, nrm, local);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                      const qbx_kernel_desc *d, double *local)
{
    QBX_CHECK(ctx && T && R && local, "qbx_refine_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_refine_locals(ctx, stream, T, R, d, local);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    for (int l = 1; l < T->nlevels; ++l) {
        const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr > 0)
            hipLaunchKernelGGL(k_l2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B,
                               l, i0, i1, R->ttp_boxes.p, local);
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                               const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                               const double *nrm, double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && str && qexp, "qbx_form_global_qbx_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_form_global_qbx_locals(ctx, stream, T, R, G, d, str, nrm, qexp, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    const int ng = G->nglobal;
    if (ng == 0) return QBX_OK;
    DISPATCH_PQ(kd.pq, k_p2qbxl, dim3(cdiv((long)ng * P2Q_LANES, 256)), dim3(256), 0, st, kd,
                make_td(T), ng, G->global_qbx_centers.p, G->center_to_target_box.p,
                G->centers.p, G->radii.p, G->ncenters, R->l1_starts.p, R->l1_lists.p,
                T->sources.p, T->nsources, str, nrm, qexp, accumulate);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                              const qbx_trav *R, const qbx_geo *G,
                                              const qbx_kernel_desc *d, const double *mpole,
                                              double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && mpole && qexp, "M2QBXL: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_translate_box_multipoles_to_qbx_local(ctx, stream, T, R, G, d, mpole, qexp,
                                                        accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    const int ng = G->nglobal;
    if (ng == 0 || R->n3 == 0) return QBX_OK;
    DISPATCH_PQ(kd.pq, k_m2qbxl, dim3(cdiv((long)ng * M2Q_LANES, 256)), dim3(256), 0, st, kd,
                make_td(T), B, ng,
                G->global_qbx_centers.p, G->center_to_target_box.p, G->centers.p, G->radii.p,
                G->ncenters, R->l3_starts.p, R->l3_lists.p, mpole, qexp, 1);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                         const qbx_trav *R, const qbx_geo *G,
                                         const qbx_kernel_desc *d, const double *local,
                                         double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && local && qexp, "L2QBXL: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_translate_box_local_to_qbx_local(ctx, stream, T, R, G, d, local, qexp,
                                                   accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    const int ng = G->nglobal;
    if (ng == 0) return QBX_OK;
    DISPATCH_PQ(kd.pq, k_l2qbxl, dim3(cdiv(ng, 64)), dim3(64), 0, st, kd, make_td(T), B, ng,
                G->global_qbx_centers.p, G->center_to_target_box.p, R->target_boxes.p,
                G->centers.p, G->radii.p, G->ncenters, local, qexp, 1);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                            const qbx_kernel_desc *d, const double *qexp, double *pot)
{
    QBX_CHECK(ctx && T && G && qexp && pot, "qbx_eval_qbx_expansions: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_eval_qbx_expansions(ctx, stream, T, G, d, qexp, pot);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (G->nctt > 0)
        hipLaunchKernelGGL(k_qbxl2p, dim3(cdiv(G->nctt, 64)), dim3(64), 0, st, kd, G->nctt,
                           G->ctt_lists.p, G->tree_ttc.p, T->targets.p, T->ntargets,
                           G->centers.p, G->radii.p, G->ncenters, qexp, pot);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_finalize_potentials(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                            const double *pot, double *out)
{
    QBX_CHECK(ctx && T && G && pot && out, "qbx_finalize_potentials: null argument");
    const int nt = T->ntargets, nc = G->ncenters;
    if (nt > nc)
        hipLaunchKernelGGL(k_finalize, dim3(cdiv(nt - nc, 256)), dim3(256), 0, S(stream), nt, nc,
                           G->user_to_tree_target.p, pot, out);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_drive_fmm(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                  const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                  const double *nrm, double *out_user)
{
    QBX_CHECK(ctx && T && R && G && str && out_user, "qbx_drive_fmm: null argument");
    const long nexp = qbx_expansion_size(d, T);
    const int ncq = qbx_ncoeff_qbx(d);
    if (nexp < 0 || ncq < 0) return QBX_ERR_INVALID;
    const long nm = 2L * std::max(nexp, 1L);
    const long nq = 2L * std::max(G->ncenters, 1) * ncq;
    const long np = 2L * std::max(T->ntargets, 1);
    if (ctx->ws_mpole.n < nm) QBX_TRY(ctx->ws_mpole.alloc(nm));
    if (ctx->ws_local.n < nm) QBX_TRY(ctx->ws_local.alloc(nm));
    if (ctx->ws_qbx.n < nq) QBX_TRY(ctx->ws_qbx.alloc(nq));
    if (ctx->ws_pot.n < np) QBX_TRY(ctx->ws_pot.alloc(np));
    double *M = ctx->ws_mpole.p, *L = ctx->ws_local.p, *Q = ctx->ws_qbx.p, *P = ctx->ws_pot.p;
    QBX_TRY(qbx_form_multipoles(ctx, stream, T, R, d, str, nrm, M));
    QBX_TRY(qbx_coarsen_multipoles(ctx, stream, T, R, d, M));
    QBX_TRY(qbx_eval_direct(ctx, stream, T, R, G, d, str, nrm, P, 0));
    QBX_TRY(qbx_multipole_to_local(ctx, stream, T, R, d, M, L));
    QBX_TRY(qbx_eval_multipoles(ctx, stream, T, R, G, d, M, P, 1));
    QBX_TRY(qbx_form_locals(ctx, stream, T, R, d, str, nrm, L, 1));
    QBX_TRY(qbx_refine_locals(ctx, stream, T, R, d, L));
    QBX_TRY(qbx_eval_locals(ctx, stream, T, R, G, d, L, P, 1));
    QBX_TRY(qbx_form_global_qbx_locals(ctx, stream, T, R, G, d, str, nrm, Q, 0));
    QBX_TRY(qbx_translate_box_multipoles_to_qbx_local(ctx, stream, T, R, G, d, M, Q, 1));
    QBX_TRY(qbx_translate_box_local_to_qbx_local(ctx, stream, T, R, G, d, L, Q, 1));
    QBX_TRY(qbx_eval_qbx_expansions(ctx, stream, T, G, d, Q, P));
    QBX_TRY(qbx_finalize_potentials(ctx, stream, T, G, P, out_user));
    return QBX_OK;
}

}  // extern "C"

// QBX-FMM wrangler stages for 2D Laplace and Helmholtz (pytential/qbx/fmm.py
// drive_fmm and pytential/qbx/interactions.py), FP64 / complex128.
//
// Expansions (coefficient layout in qbx_mi355x.h):
//  Laplace   G = -log|x-y|/(2 pi); complex potential Phi, phi = -Re(Phi)/(2 pi)
//            multipole a0 log(z-c) + sum_k a_k (R/(z-c))^k, local sum_l b_l ((z-c)/R)^l
//  Helmholtz G = (i/4) H0(k|x-y|); multipole sum_n M_n H_n(k rho) e^{in theta},
//            local sum_n L_n J_n(k rho) e^{in theta}; translations by Graf's
//            addition theorem (M2M/L2L with J, M2L with H).
// The double layer is the source-normal derivative of G.
//
// Thread mappings (first version; the measured hot spots get specialised):
//  P2M/M2M/M2L/P2L/L2L  one thread per (box, output coefficient)
//  P2P/M2P/L2P          one block per target box, one thread per target
//  P2QBXL/M2QBXL/L2QBXL one thread per global QBX center, the order-p_q
//                       expansion kept in registers (template on p_q)
//  QBXL2P               one thread per (center, target) pair
#include "qbx_internal.h"
#include "qbx_math.h"

namespace qbx {

constexpr int MAXP = 120;
constexpr int BINOM_N = 2 * MAXP + 4;

struct KD {
    int kernel, dip, pq, pmax, ncf, ncq, nlevels;
    double k;
    int order[MAX_LEVELS];
    double rscale[MAX_LEVELS];
};

struct TD {
    int nboxes;
    const int *level, *parent, *child;
    const double *centers;  // [2][nboxes]
    const int *sstart, *scount, *scumul;
    const int *tstart, *tcount;
};

__device__ __forceinline__ cplx boxc(const TD &t, int b)
{
    return {t.centers[b], t.centers[(size_t)t.nboxes + b]};
}

__device__ __forceinline__ cplx ldpt(const double *x, int n, int i)
{
    return {x[i], x[(size_t)n + i]};
}

constexpr double INV2PI = 0.15915494309189533577;

// -------------------------------------------------------------------------
// pairwise kernels
// -------------------------------------------------------------------------
__device__ __forceinline__ cplx p2p_pair(const KD &kd, cplx s, cplx q, cplx nu, cplx t)
{
    cplx w = t - s;
    double r2 = abs2(w);
    if (r2 == 0.0) return {0.0, 0.0};
    if (kd.kernel == QBX_LAPLACE2D) {
        if (!kd.dip) return (-0.5 * log(r2) * INV2PI) * q;
        cplx v = (nu * q) * cinv(w);
        return {v.x * INV2PI, 0.0};
    } else {
        double r = sqrt(r2);
        double kr = kd.k * r;
        if (!kd.dip) {
            cplx h0 = {j0(kr), y0(kr)};
            cplx v = q * h0;
            return {-0.25 * v.y, 0.25 * v.x};
        }
        double proj = (w.x * nu.x + w.y * nu.y) / r;
        cplx h1 = {j1(kr), y1(kr)};
        cplx v = (kd.k * proj) * (q * h1);
        return {-0.25 * v.y, 0.25 * v.x};
    }
}

// -------------------------------------------------------------------------
// P2M / P2L: one coefficient (index n) of the expansion of sources [s0,s1)
// -------------------------------------------------------------------------
// Helmholtz: Z_m(k rho) e^{-i m theta} for m in {n-1, n, n+1} of v = s - c.
// local=1 uses H (forward recurrence), local=0 uses J (Miller).
__device__ cplx helm_p2e_coeff(const KD &kd, int local, int n, cplx c, const double *src,
                               int ns, int s0, int s1, const double *str, const double *nrm)
{
    cplx acc = {0, 0};
    int an = abs(n) + 1;
    for (int s = s0; s < s1; ++s) {
        cplx v = ldpt(src, ns, s) - c;
        double rho = sqrt(abs2(v));
        double x = kd.k * rho;
        double th = atan2(v.y, v.x);
        cplx q = ldc(str + 2 * s);
        // z[m] for m = |n|-1, |n|, |n|+1 (m >= 0) via |m| symmetry
        double zr[3], zi[3];
        int mm[3] = {n - 1, n, n + 1};
        if (local) {
            cplx h0, h1;
            hankel01(x, h0, h1);
            cplx hm1 = h0, hm = h1;  // H_0, H_1
            cplx hv[3];
            // need H_|m| for |m| up to an
            int need[3] = {abs(mm[0]), abs(mm[1]), abs(mm[2])};
            for (int j = 0; j < 3; ++j) hv[j] = need[j] == 0 ? h0 : (need[j] == 1 ? h1 : cplx{0, 0});
            double tox = 2.0 / x;
            for (int m = 1; m < an; ++m) {
                cplx hn = (m * tox) * hm - hm1;
                hm1 = hm;
                hm = hn;
                for (int j = 0; j < 3; ++j)
                    if (need[j] == m + 1) hv[j] = hn;
            }
            for (int j = 0; j < 3; ++j) { zr[j] = hv[j].x; zi[j] = hv[j].y; }
        } else {
            double jv[3] = {0, 0, 0};
            int need[3] = {abs(mm[0]), abs(mm[1]), abs(mm[2])};
            if (x < 1e-300) {
                for (int j = 0; j < 3; ++j) jv[j] = need[j] == 0 ? 1.0 : 0.0;
            } else {
                double norm = miller_j(
                    x, an,
                    [&](int m, double f) {
                        for (int j = 0; j < 3; ++j)
                            if (need[j] == m) jv[j] = f;
                    },
                    [&](double sc) {
                        for (int j = 0; j < 3; ++j) jv[j] *= sc;
                    });
                for (int j = 0; j < 3; ++j) jv[j] *= norm;
            }
            for (int j = 0; j < 3; ++j) { zr[j] = jv[j]; zi[j] = 0.0; }
        }
        // g_m = Z_m e^{-i m th}, Z_{-m} = (-1)^m Z_m
        cplx g[3];
        for (int j = 0; j < 3; ++j) {
            int m = mm[j];
            double sgn = (m < 0 && (m & 1)) ? -1.0 : 1.0;
            double sn, cs;
            sincos(-(double)m * th, &sn, &cs);
            g[j] = (sgn * cplx{zr[j], zi[j]}) * cplx{cs, sn};
        }
        if (!kd.dip) {
            acc += q * g[1];
        } else {
            cplx nu = ldpt(nrm, ns, s);
            acc += (0.5 * kd.k) * (q * (conj(nu) * g[0] - nu * g[2]));
        }
    }
    return acc;
}

__device__ cplx lap_p2e_coeff(const KD &kd, int local, int l, cplx c, double R, const double *src,
                              int ns, int s0, int s1, const double *str, const double *nrm)
{
    cplx acc = {0, 0};
    for (int s = s0; s < s1; ++s) {
        cplx w = ldpt(src, ns, s) - c;
        cplx q = ldc(str + 2 * s);
        if (local) {
            cplx t = R * cinv(w), pw = {1, 0};
            for (int i = 0; i < l; ++i) pw = pw * t;
            if (!kd.dip) {
                if (l == 0) acc += q * clogz(-w);
                else acc += (-1.0 / l) * (q * pw);
            } else {
                cplx nu = ldpt(nrm, ns, s);
                acc += ((nu * q) * cinv(w)) * pw;
            }
        } else {
            cplx t = (1.0 / R) * w, pw = {1, 0};
            if (!kd.dip) {
                for (int i = 0; i < l; ++i) pw = pw * t;
                if (l == 0) acc += q;
                else acc += (-1.0 / l) * (q * pw);
            } else if (l >= 1) {
                for (int i = 0; i < l - 1; ++i) pw = pw * t;
                cplx nu = ldpt(nrm, ns, s);
                acc += -((1.0 / R) * (nu * q)) * pw;
            }
        }
    }
    return acc;
}

__device__ __forceinline__ int coeff_index(const KD &kd, int j)
{
    return kd.kernel == QBX_LAPLACE2D ? j : j + kd.pmax;
}

// threads: (box index i in `boxes`, j in [0, ncf))
__global__ void k_p2m(KD kd, TD t, int nboxes_list, const int *boxes, const double *src, int ns,
                      const double *str, const double *nrm, double *mpole)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i = (int)(gid / kd.ncf), j = (int)(gid % kd.ncf);
    if (i >= nboxes_list) return;
    int b = boxes[i];
    int p = kd.order[t.level[b]];
    int n = kd.kernel == QBX_LAPLACE2D ? j : j - kd.pmax;
    cplx v = {0, 0};
    if ((kd.kernel == QBX_LAPLACE2D && n <= p) || (kd.kernel != QBX_LAPLACE2D && abs(n) <= p)) {
        int s0 = t.sstart[b], s1 = s0 + t.scount[b];
        cplx c = boxc(t, b);
        if (kd.kernel == QBX_LAPLACE2D)
            v = lap_p2e_coeff(kd, 0, n, c, kd.rscale[t.level[b]], src, ns, s0, s1, str, nrm);
        else
            v = helm_p2e_coeff(kd, 0, n, c, src, ns, s0, s1, str, nrm);
    }
    stc(mpole + 2 * ((size_t)b * kd.ncf + j), v);
}

// P2L from list 4: threads (ttp index i, j)
__global__ void k_p2l(KD kd, TD t, int nttp, const int *ttp, const int *starts, const int *lists,
                      const double *src, int ns, const double *str, const double *nrm,
                      double *local)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i = (int)(gid / kd.ncf), j = (int)(gid % kd.ncf);
    if (i >= nttp) return;
    int P = ttp[i];
    int p = kd.order[t.level[P]];
    int n = kd.kernel == QBX_LAPLACE2D ? j : j - kd.pmax;
    if ((kd.kernel == QBX_LAPLACE2D && n > p) || (kd.kernel != QBX_LAPLACE2D && abs(n) > p)) return;
    if (starts[i] == starts[i + 1]) return;
    cplx c = boxc(t, P);
    cplx acc = {0, 0};
    for (int e = starts[i]; e < starts[i + 1]; ++e) {
        int Sb = lists[e];
        int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
        if (kd.kernel == QBX_LAPLACE2D)
            acc += lap_p2e_coeff(kd, 1, n, c, kd.rscale[t.level[P]], src, ns, s0, s1, str, nrm);
        else
            acc += helm_p2e_coeff(kd, 1, n, c, src, ns, s0, s1, str, nrm);
    }
    double *o = local + 2 * ((size_t)P * kd.ncf + j);
    stc(o, ldc(o) + acc);
}

// -------------------------------------------------------------------------
// E2E translations producing one output coefficient
// -------------------------------------------------------------------------
// Helmholtz, regular waves (M2M, L2L, L2QBXL): out_m = sum_n a_n J_{n-m}(k|d|) e^{i(n-m) th}
__device__ cplx helm_translate_J(const KD &kd, const double *a, int p1, int m, cplx d)
{
    double rho = sqrt(abs2(d));
    double x = kd.k * rho;
    int lmax = p1 + abs(m);
    if (x < 1e-300) {
        // J_l(0) = delta_l0
        if (abs(m) <= p1) return ldc(a + 2 * (m + kd.pmax));
        return {0, 0};
    }
    double th = atan2(d.y, d.x);
    double sn, cs, sn1, cs1;
    sincos((double)lmax * th, &sn, &cs);
    sincos(-th, &sn1, &cs1);
    cplx e = {cs, sn}, em1 = {cs1, sn1};  // e^{i l th}, e^{-i th}
    cplx acc = {0, 0};
    double norm = miller_j(
        x, lmax,
        [&](int l, double f) {
            // e currently e^{i l th}
            int n1 = m + l;
            if (n1 >= -p1 && n1 <= p1) acc += f * (ldc(a + 2 * (n1 + kd.pmax)) * e);
            if (l > 0) {
                int n2 = m - l;
                if (n2 >= -p1 && n2 <= p1) {
                    double sg = (l & 1) ? -f : f;
                    acc += sg * (ldc(a + 2 * (n2 + kd.pmax)) * conj(e));
                }
            }
            e = e * em1;
        },
        [&](double s) { acc = s * acc; });
    return norm * acc;
}

// Helmholtz, singular waves (M2L, M2QBXL): out_m = sum_n a_n H_{n-m}(k|d|) e^{i(n-m) th}
__device__ cplx helm_translate_H(const KD &kd, const double *a, int p1, int m, cplx d)
{
    double rho = sqrt(abs2(d));
    double x = kd.k * rho;
    int lmax = p1 + abs(m);
    double th = atan2(d.y, d.x);
    double sn, cs;
    sincos(th, &sn, &cs);
    cplx e1 = {cs, sn}, e = {1, 0};
    cplx hm1, hm;
    hankel01(x, hm1, hm);  // H_0, H_1
    double tox = 2.0 / x;
    cplx acc = {0, 0};
    cplx h = hm1;
    for (int l = 0; l <= lmax; ++l) {
        if (l == 1) h = hm;
        else if (l >= 2) {
            cplx hn = ((l - 1) * tox) * hm - hm1;
            hm1 = hm;
            hm = hn;
            h = hn;
        }
        int n1 = m + l;
        if (n1 >= -p1 && n1 <= p1) acc += h * (ldc(a + 2 * (n1 + kd.pmax)) * e);
        if (l > 0) {
            int n2 = m - l;
            if (n2 >= -p1 && n2 <= p1) {
                cplx v = h * (ldc(a + 2 * (n2 + kd.pmax)) * conj(e));
                acc += (l & 1) ? -v : v;
            }
        }
        e = e * e1;
    }
    return acc;
}

__device__ double g_binom(const double *B, int n, int k)
{
    return (k < 0 || k > n) ? 0.0 : B[n * BINOM_N + k];
}

// Laplace M2M (c1 -> c2, z0 = c1 - c2), one output l:
// -a0 u^l / l + sum_{k=1}^{min(l,p1)} a_k r^k u^{l-k} C(l-1,k-1), u = z0/R2, r = R1/R2
__device__ cplx lap_m2m(const double *B, const double *a, int p1, int l, cplx z0, double R1,
                        double R2)
{
    if (l == 0) return ldc(a);
    cplx u = (1.0 / R2) * z0;
    double r = R1 / R2;
    cplx ul = {1, 0};
    for (int i = 0; i < l; ++i) ul = ul * u;
    cplx acc = (-1.0 / l) * (ldc(a) * ul);
    int kmax = l < p1 ? l : p1;
    cplx hsum = {0, 0};
    double rk = 1.0;
    for (int k = 1; k <= kmax; ++k) {
        rk *= r;
        hsum = hsum * u + (rk * g_binom(B, l - 1, k - 1)) * ldc(a + 2 * k);
    }
    cplx ulk = {1, 0};
    for (int i = 0; i < l - kmax; ++i) ulk = ulk * u;
    return acc + hsum * ulk;
}

// Laplace M2L (c1 multipole -> c2 local), one output l
__device__ cplx lap_m2l(const double *B, const double *a, int p1, int l, cplx z0, double R1,
                        double R2)
{
    cplx iz = cinv(z0);
    cplx s = R1 * iz, t = R2 * iz;
    cplx ms = -s;
    cplx acc = {0, 0};
    if (l == 0) {
        acc = ldc(a) * clogz(-z0);
        cplx pw = {1, 0};
        for (int k = 1; k <= p1; ++k) {
            pw = pw * ms;
            acc += ldc(a + 2 * k) * pw;
        }
        return acc;
    }
    acc = (-1.0 / l) * ldc(a);
    cplx pw = {1, 0};
    for (int k = 1; k <= p1; ++k) {
        pw = pw * ms;
        acc += g_binom(B, l + k - 1, k - 1) * (ldc(a + 2 * k) * pw);
    }
    cplx tl = {1, 0};
    for (int i = 0; i < l; ++i) tl = tl * t;
    return acc * tl;
}

// Laplace L2L (c1 -> c2, z0 = c2 - c1), one output j
__device__ cplx lap_l2l(const double *B, const double *a, int p1, int j, cplx z0, double R1,
                        double R2)
{
    if (j > p1) return {0, 0};
    cplx u = (1.0 / R1) * z0;
    cplx acc = {0, 0};
    for (int l = p1; l >= j; --l) acc = acc * u + g_binom(B, l, j) * ldc(a + 2 * l);
    double rp = 1.0, r21 = R2 / R1;
    for (int i = 0; i < j; ++i) rp *= r21;
    return rp * acc;
}

__device__ __forceinline__ bool coeff_valid(const KD &kd, int j, int p, int *n)
{
    if (kd.kernel == QBX_LAPLACE2D) {
        *n = j;
        return j <= p;
    }
    *n = j - kd.pmax;
    return abs(*n) <= p;
}

// M2M for parents at one level: threads (i over [i0,i1) of source_parent_boxes, j)
__global__ void k_m2m(KD kd, TD t, const double *B, int i0, int i1, const int *spb, double *mpole)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i = i0 + (int)(gid / kd.ncf), j = (int)(gid % kd.ncf);
    if (i >= i1) return;
    int P = spb[i];
    int p2 = kd.order[t.level[P]], n;
    if (!coeff_valid(kd, j, p2, &n)) return;
    cplx cp = boxc(t, P);
    cplx acc = {0, 0};
    for (int c = 0; c < 4; ++c) {
        int C = t.child[(size_t)c * t.nboxes + P];
        if (C < 0 || t.scumul[C] == 0) continue;
        int p1 = kd.order[t.level[C]];
        const double *a = mpole + 2 * (size_t)C * kd.ncf;
        cplx cc = boxc(t, C);
        if (kd.kernel == QBX_LAPLACE2D)
            acc += lap_m2m(B, a, p1, n, cc - cp, kd.rscale[t.level[C]], kd.rscale[t.level[P]]);
        else
            acc += helm_translate_J(kd, a, p1, n, cp - cc);
    }
    double *o = mpole + 2 * ((size_t)P * kd.ncf + j);
    stc(o, ldc(o) + acc);
}

// M2L list 2: threads (ttp index i, j)
__global__ void k_m2l(KD kd, TD t, const double *B, int nttp, const int *ttp, const int *starts,
                      const int *lists, const double *mpole, double *local)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i = (int)(gid / kd.ncf), j = (int)(gid % kd.ncf);
    if (i >= nttp) return;
    if (starts[i] == starts[i + 1]) return;
    int P = ttp[i];
    int p2 = kd.order[t.level[P]], n;
    if (!coeff_valid(kd, j, p2, &n)) return;
    cplx cp = boxc(t, P);
    cplx acc = {0, 0};
    for (int e = starts[i]; e < starts[i + 1]; ++e) {
        int A = lists[e];
        int p1 = kd.order[t.level[A]];
        const double *a = mpole + 2 * (size_t)A * kd.ncf;
        cplx ca = boxc(t, A);
        if (kd.kernel == QBX_LAPLACE2D)
            acc += lap_m2l(B, a, p1, n, ca - cp, kd.rscale[t.level[A]], kd.rscale[t.level[P]]);
        else
            acc += helm_translate_H(kd, a, p1, n, cp - ca);
    }
    double *o = local + 2 * ((size_t)P * kd.ncf + j);
    stc(o, ldc(o) + acc);
}

// L2L for boxes at one level (ttp range [i0,i1)): threads (i, j)
__global__ void k_l2l(KD kd, TD t, const double *B, int i0, int i1, const int *ttp, double *local)
{
    long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int i = i0 + (int)(gid / kd.ncf), j = (int)(gid % kd.ncf);
    if (i >= i1) return;
    int P = ttp[i];
    int Q = t.parent[P];
    if (Q < 0) return;
    int p2 = kd.order[t.level[P]], n;
    if (!coeff_valid(kd, j, p2, &n)) return;
    int p1 = kd.order[t.level[Q]];
    const double *a = local + 2 * (size_t)Q * kd.ncf;
    cplx cq = boxc(t, Q), cp = boxc(t, P);
    cplx v;
    if (kd.kernel == QBX_LAPLACE2D)
        v = lap_l2l(B, a, p1, n, cp - cq, kd.rscale[t.level[Q]], kd.rscale[t.level[P]]);
    else
        v = helm_translate_J(kd, a, p1, n, cp - cq);
    double *o = local + 2 * ((size_t)P * kd.ncf + j);
    stc(o, ldc(o) + v);
}

// -------------------------------------------------------------------------
// evaluation at points
// -------------------------------------------------------------------------
__device__ cplx eval_local(const KD &kd, const double *a, int p, cplx c, double R, cplx tp)
{
    if (kd.kernel == QBX_LAPLACE2D) {
        cplx x = (1.0 / R) * (tp - c);
        cplx phi = {0, 0};
        for (int l = p; l >= 0; --l) phi = phi * x + ldc(a + 2 * l);
        return {-phi.x * INV2PI, 0.0};
    }
    cplx d = tp - c;
    double x = kd.k * sqrt(abs2(d));
    cplx phi;
    if (x < 1e-300) {
        phi = ldc(a + 2 * kd.pmax);
    } else {
        double th = atan2(d.y, d.x);
        double sn, cs, sn1, cs1;
        sincos((double)p * th, &sn, &cs);
        sincos(-th, &sn1, &cs1);
        cplx e = {cs, sn}, em1 = {cs1, sn1};
        cplx acc = {0, 0};
        double norm = miller_j(
            x, p,
            [&](int n, double f) {
                cplx v = ldc(a + 2 * (n + kd.pmax)) * e;
                if (n > 0) {
                    cplx w = ldc(a + 2 * (-n + kd.pmax)) * conj(e);
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
    cplx w = tp - c;
    if (kd.kernel == QBX_LAPLACE2D) {
        cplx x = R * cinv(w);
        cplx phi = {0, 0};
        for (int k = p; k >= 1; --k) phi = (phi + ldc(a + 2 * k)) * x;
        phi += ldc(a) * clogz(w);
        return {-phi.x * INV2PI, 0.0};
    }
    double x = kd.k * sqrt(abs2(w));
    double th = atan2(w.y, w.x);
    double sn, cs;
    sincos(th, &sn, &cs);
    cplx e1 = {cs, sn}, e = {1, 0};
    cplx hm1, hm;
    hankel01(x, hm1, hm);
    double tox = 2.0 / x;
    cplx acc = ldc(a + 2 * kd.pmax) * hm1;
    for (int n = 1; n <= p; ++n) {
        cplx h;
        if (n == 1) h = hm;
        else {
            h = ((n - 1) * tox) * hm - hm1;
            hm1 = hm;
            hm = h;
        }
        e = e * e1;
        cplx v = ldc(a + 2 * (n + kd.pmax)) * e;
        cplx wv = ldc(a + 2 * (-n + kd.pmax)) * conj(e);
        v += (n & 1) ? -wv : wv;
        acc += h * v;
    }
    return {-0.25 * acc.y, 0.25 * acc.x};
}

// block per target box (index it), threads over its non-QBX targets
// mode 0: P2P (list 1), 1: M2P (list 3), 2: L2P
template <int MODE>
__global__ void k_eval_targets(KD kd, TD t, int ntb, const int *tboxes, const int *starts,
                               const int *lists, const int *fstart, const int *fcount,
                               const int *unfilt, const double *tgt, int nt, const double *src,
                               int ns, const double *str, const double *nrm,
                               const double *expn, double *pot)
{
    int it = blockIdx.x;
    if (it >= ntb) return;
    int B = tboxes[it];
    int f0 = fstart[B], nf = fcount[B];
    for (int f = f0 + threadIdx.x; f < f0 + nf; f += blockDim.x) {
        int ti = unfilt[f];
        cplx tp = ldpt(tgt, nt, ti);
        cplx acc = {0, 0};
        if (MODE == 0) {
            for (int e = starts[it]; e < starts[it + 1]; ++e) {
                int Sb = lists[e];
                int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
                for (int s = s0; s < s1; ++s) {
                    cplx nu = kd.dip ? ldpt(nrm, ns, s) : cplx{0, 0};
                    acc += p2p_pair(kd, ldpt(src, ns, s), ldc(str + 2 * s), nu, tp);
                }
            }
        } else if (MODE == 1) {
            for (int e = starts[it]; e < starts[it + 1]; ++e) {
                int D = lists[e];
                int lv = t.level[D];
                acc += eval_mpole(kd, expn + 2 * (size_t)D * kd.ncf, kd.order[lv], boxc(t, D),
                                  kd.rscale[lv], tp);
            }
        } else {
            int lv = t.level[B];
            acc = eval_local(kd, expn + 2 * (size_t)B * kd.ncf, kd.order[lv], boxc(t, B),
                             kd.rscale[lv], tp);
        }
        double *o = pot + 2 * (size_t)ti;
        stc(o, ldc(o) + acc);
    }
}

// -------------------------------------------------------------------------
// QBX stages, one thread per global QBX center, order PQ in registers
// -------------------------------------------------------------------------
template <int PQ>
struct QAcc {
    static constexpr int NL = PQ + 1;        // Laplace
    static constexpr int NH = 2 * PQ + 1;    // Helmholtz
    cplx v[NH];
};

template <int PQ>
__device__ void qbx_store(const KD &kd, double *qexp, int c, const QAcc<PQ> &A, int accumulate)
{
    int nq = kd.kernel == QBX_LAPLACE2D ? PQ + 1 : 2 * PQ + 1;
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

template <int PQ>
__global__ void k_p2qbxl(KD kd, TD t, int ng, const int *gcent, const int *c2tb,
                         const double *ctr, const double *rad, int nc, const int *l1s,
                         const int *l1l, const double *src, int ns, const double *str,
                         const double *nrm, double *qexp, int accumulate)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int c = gcent[g];
    int it = c2tb[c];
    cplx cc = ldpt(ctr, nc, c);
    double R = rad[c];
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    for (int e = l1s[it]; e < l1s[it + 1]; ++e) {
        int Sb = l1l[e];
        int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
        for (int s = s0; s < s1; ++s) {
            cplx w = ldpt(src, ns, s) - cc;
            cplx q = ldc(str + 2 * s);
            if (kd.kernel == QBX_LAPLACE2D) {
                cplx iw = cinv(w);
                cplx tt = R * iw;
                cplx pw = {1, 0};
                if (!kd.dip) {
                    A.v[0] += q * clogz(-w);
#pragma unroll
                    for (int l = 1; l <= PQ; ++l) {
                        pw = pw * tt;
                        A.v[l] += (-1.0 / l) * (q * pw);
                    }
                } else {
                    cplx f = (ldpt(nrm, ns, s) * q) * iw;
                    A.v[0] += f;
#pragma unroll
                    for (int l = 1; l <= PQ; ++l) {
                        pw = pw * tt;
                        A.v[l] += f * pw;
                    }
                }
            } else {
                double x = kd.k * sqrt(abs2(w));
                double th = atan2(w.y, w.x);
                double sn, cs;
                sincos(-th, &sn, &cs);
                cplx em = {cs, sn};  // e^{-i th}
                // H_m for m = 0..PQ+1
                cplx H[PQ + 2];
                hankel01(x, H[0], H[1]);
                double tox = 2.0 / x;
#pragma unroll
                for (int m = 2; m <= PQ + 1; ++m) H[m] = ((m - 1) * tox) * H[m - 1] - H[m - 2];
                // g_m = H_m e^{-i m th}, m in [-(PQ+1), PQ+1], index m + PQ + 1
                cplx G[2 * PQ + 3];
                cplx e = {1, 0};
#pragma unroll
                for (int m = 0; m <= PQ + 1; ++m) {
                    G[PQ + 1 + m] = H[m] * e;
                    cplx gm = H[m] * conj(e);
                    G[PQ + 1 - m] = (m & 1) ? -gm : gm;
                    e = e * em;
                }
                if (!kd.dip) {
#pragma unroll
                    for (int n = -PQ; n <= PQ; ++n) A.v[n + PQ] += q * G[n + PQ + 1];
                } else {
                    cplx nu = ldpt(nrm, ns, s);
                    cplx hq = (0.5 * kd.k) * q;
                    cplx a1 = hq * conj(nu), a2 = hq * nu;
#pragma unroll
                    for (int n = -PQ; n <= PQ; ++n)
                        A.v[n + PQ] += a1 * G[n + PQ] - a2 * G[n + PQ + 2];
                }
            }
        }
    }
    qbx_store<PQ>(kd, qexp, c, A, accumulate);
}

template <int PQ>
__global__ void k_m2qbxl(KD kd, TD t, const double *B, int ng, const int *gcent, const int *c2tb,
                         const double *ctr, const double *rad, int nc, const int *l3s,
                         const int *l3l, const double *mpole, double *qexp, int accumulate)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int c = gcent[g];
    int it = c2tb[c];
    cplx cc = ldpt(ctr, nc, c);
    double R2 = rad[c];
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    for (int e = l3s[it]; e < l3s[it + 1]; ++e) {
        int D = l3l[e];
        int lv = t.level[D];
        int p1 = kd.order[lv];
        const double *a = mpole + 2 * (size_t)D * kd.ncf;
        cplx cd = boxc(t, D);
        if (kd.kernel == QBX_LAPLACE2D) {
            cplx z0 = cd - cc;
            cplx iz = cinv(z0);
            cplx ms = -(kd.rscale[lv] * iz), tt = R2 * iz;
            cplx a0 = ldc(a);
            cplx S[PQ + 1];
            S[0] = a0 * clogz(-z0);
#pragma unroll
            for (int l = 1; l <= PQ; ++l) S[l] = (-1.0 / l) * a0;
            cplx pw = {1, 0};
            for (int k = 1; k <= p1; ++k) {
                pw = pw * ms;
                cplx ak = ldc(a + 2 * k) * pw;
                S[0] += ak;
#pragma unroll
                for (int l = 1; l <= PQ; ++l) S[l] += g_binom(B, l + k - 1, k - 1) * ak;
            }
            cplx tl = {1, 0};
#pragma unroll
            for (int l = 0; l <= PQ; ++l) {
                A.v[l] += S[l] * tl;
                tl = tl * tt;
            }
        } else {
            cplx d = cc - cd;
            double x = kd.k * sqrt(abs2(d));
            double th = atan2(d.y, d.x);
            double sn, cs;
            sincos(th, &sn, &cs);
            cplx e1 = {cs, sn}, ep = {1, 0};
            cplx hm1, hm;
            hankel01(x, hm1, hm);
            double tox = 2.0 / x;
            int lmax = p1 + PQ;
            cplx h = hm1;
            for (int l = 0; l <= lmax; ++l) {
                if (l == 1) h = hm;
                else if (l >= 2) {
                    h = ((l - 1) * tox) * hm - hm1;
                    hm1 = hm;
                    hm = h;
                }
                cplx he = h * ep, hc = h * conj(ep);
                if (l & 1) hc = -hc;
#pragma unroll
                for (int m = -PQ; m <= PQ; ++m) {
                    int n1 = m + l;
                    if (n1 >= -p1 && n1 <= p1) A.v[m + PQ] += ldc(a + 2 * (n1 + kd.pmax)) * he;
                    int n2 = m - l;
                    if (l > 0 && n2 >= -p1 && n2 <= p1)
                        A.v[m + PQ] += ldc(a + 2 * (n2 + kd.pmax)) * hc;
                }
                ep = ep * e1;
            }
        }
    }
    qbx_store<PQ>(kd, qexp, c, A, accumulate);
}

template <int PQ>
__global__ void k_l2qbxl(KD kd, TD t, const double *B, int ng, const int *gcent, const int *c2tb,
                         const int *tboxes, const double *ctr, const double *rad, int nc,
                         const double *local, double *qexp, int accumulate)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    int c = gcent[g];
    int Bx = tboxes[c2tb[c]];
    int lv = t.level[Bx];
    int p1 = kd.order[lv];
    cplx cc = ldpt(ctr, nc, c);
    double R2 = rad[c];
    const double *a = local + 2 * (size_t)Bx * kd.ncf;
    cplx cb = boxc(t, Bx);
    QAcc<PQ> A;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) A.v[i] = {0, 0};
    if (kd.kernel == QBX_LAPLACE2D) {
        double R1 = kd.rscale[lv];
        cplx u = (1.0 / R1) * (cc - cb);
        for (int l = p1; l >= 0; --l) {
            cplx al = ldc(a + 2 * l);
#pragma unroll
            for (int j = 0; j <= PQ; ++j)
                if (j <= l) A.v[j] = A.v[j] * u + g_binom(B, l, j) * al;
        }
        double rp = 1.0, r21 = R2 / R1;
#pragma unroll
        for (int j = 0; j <= PQ; ++j) {
            A.v[j] = rp * A.v[j];
            rp *= r21;
        }
    } else {
        cplx d = cc - cb;
        double x = kd.k * sqrt(abs2(d));
        int lmax = p1 + PQ;
        if (x < 1e-300) {
#pragma unroll
            for (int m = -PQ; m <= PQ; ++m)
                if (abs(m) <= p1) A.v[m + PQ] = ldc(a + 2 * (m + kd.pmax));
        } else {
            double th = atan2(d.y, d.x);
            double sn, cs, sn1, cs1;
            sincos((double)lmax * th, &sn, &cs);
            sincos(-th, &sn1, &cs1);
            cplx e = {cs, sn}, em1 = {cs1, sn1};
            double norm = miller_j(
                x, lmax,
                [&](int l, double f) {
                    cplx fe = f * e, fc = f * conj(e);
                    if (l & 1) fc = -fc;
#pragma unroll
                    for (int m = -PQ; m <= PQ; ++m) {
                        int n1 = m + l;
                        if (n1 >= -p1 && n1 <= p1) A.v[m + PQ] += ldc(a + 2 * (n1 + kd.pmax)) * fe;
                        int n2 = m - l;
                        if (l > 0 && n2 >= -p1 && n2 <= p1)
                            A.v[m + PQ] += ldc(a + 2 * (n2 + kd.pmax)) * fc;
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
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nctt) return;
    int ti = ctt_lists[j];
    int c = tree_ttc[ti];
    cplx tp = ldpt(tgt, nt, ti);
    cplx cc = ldpt(ctr, nc, c);
    const double *a = qexp + 2 * (size_t)c * kd.ncq;
    // QBX expansions use the same storage rule with pmax = pq
    KD k2 = kd;
    k2.pmax = kd.pq;
    cplx v = eval_local(k2, a, kd.pq, cc, rad[c], tp);
    stc(pot + 2 * (size_t)ti, v);
}

__global__ void k_finalize(int nt, int nc, const int *user_to_tree, const double *pot,
                           double *out)
{
    int u = nc + blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nt) return;
    int ti = user_to_tree[u];
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
    if (dev < 0 || dev >= 64) return QBX_ERR_INVALID;
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

static int make_kd(const qbx_kernel_desc *d, const qbx_tree *T, KD *kd)
{
    QBX_CHECK(d, "null kernel description");
    QBX_CHECK(d->kernel == QBX_LAPLACE2D || d->kernel == QBX_HELMHOLTZ2D,
              "unsupported kernel %d", d->kernel);
    QBX_CHECK(d->source_kind == QBX_SOURCE_S || d->source_kind == QBX_SOURCE_D,
              "unsupported source kind %d", d->source_kind);
    QBX_CHECK(d->qbx_order >= 0 && d->qbx_order <= 12, "qbx_order must be in [0, 12]");
    QBX_CHECK(d->level_orders && d->nlevels >= 1 && d->nlevels <= MAX_LEVELS,
              "level_orders missing or nlevels out of range");
    if (T) {
        QBX_CHECK(T->dim == 2, "2D kernels need a 2D tree");
        QBX_CHECK(d->nlevels >= T->nlevels, "level_orders has %d entries, tree has %d levels",
                  d->nlevels, T->nlevels);
    }
    kd->kernel = d->kernel;
    kd->dip = d->source_kind == QBX_SOURCE_D;
    kd->pq = d->qbx_order;
    kd->k = d->k;
    kd->nlevels = d->nlevels;
    int pmax = 0;
    for (int l = 0; l < d->nlevels; ++l) {
        QBX_CHECK(d->level_orders[l] >= 1 && d->level_orders[l] <= MAXP,
                  "fmm order %d at level %d outside [1, %d]", d->level_orders[l], l, MAXP);
        kd->order[l] = d->level_orders[l];
        pmax = std::max(pmax, d->level_orders[l]);
        kd->rscale[l] = T ? ldexp(T->root_extent, -l) : 1.0;
    }
    for (int l = d->nlevels; l < MAX_LEVELS; ++l) {
        kd->order[l] = 0;
        kd->rscale[l] = 1.0;
    }
    if (kd->kernel == QBX_HELMHOLTZ2D) QBX_CHECK(d->k > 0, "Helmholtz needs k > 0");
    kd->pmax = pmax;
    kd->ncf = kd->kernel == QBX_LAPLACE2D ? pmax + 1 : 2 * pmax + 1;
    kd->ncq = kd->kernel == QBX_LAPLACE2D ? kd->pq + 1 : 2 * kd->pq + 1;
    return QBX_OK;
}

static TD make_td(const qbx_tree *T)
{
    TD t;
    t.nboxes = T->nboxes;
    t.level = T->box_levels.p;
    t.parent = T->box_parent_ids.p;
    t.child = T->box_child_ids.p;
    t.centers = T->box_centers.p;
    t.sstart = T->box_source_starts.p;
    t.scount = T->box_source_counts_nonchild.p;
    t.scumul = T->box_source_counts_cumul.p;
    t.tstart = T->box_target_starts.p;
    t.tcount = T->box_target_counts_nonchild.p;
    return t;
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

}  // namespace qbx

using namespace qbx;

extern "C" {

int qbx_ncoeff_fmm(const qbx_kernel_desc *d)
{
    KD kd;
    if (make_kd(d, nullptr, &kd) != QBX_OK) return -1;
    return kd.ncf;
}

int qbx_ncoeff_qbx(const qbx_kernel_desc *d)
{
    KD kd;
    if (make_kd(d, nullptr, &kd) != QBX_OK) return -1;
    return kd.ncq;
}

int qbx_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_kernel_desc *d, const double *str, const double *nrm,
                        double *mpole)
{
    QBX_CHECK(ctx && T && R && str && mpole, "qbx_form_multipoles: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    QBX_HIP(hipMemsetAsync(mpole, 0, sizeof(double) * 2 * (size_t)T->nboxes * kd.ncf, st));
    long nthr = (long)R->nsource_boxes * kd.ncf;
    if (nthr > 0)
        hipLaunchKernelGGL(k_p2m, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T),
                           R->nsource_boxes, R->source_boxes.p, T->sources.p, T->nsources, str,
                           nrm, mpole);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                           const qbx_kernel_desc *d, double *mpole)
{
    QBX_CHECK(ctx && T && R && mpole, "qbx_coarsen_multipoles: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    for (int l = T->nlevels - 2; l >= 0; --l) {
        int i0 = R->h_level_start_source_parent_box_nrs[l];
        int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
        long nthr = (long)(i1 - i0) * kd.ncf;
        if (nthr > 0)
            hipLaunchKernelGGL(k_m2m, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B,
                               i0, i1, R->source_parent_boxes.p, mpole);
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
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(pot, 0, sizeof(double) * 2 * (size_t)T->ntargets, st));
    int ntb = R->ntarget_boxes;
    if (ntb == 0 || G->nfiltered == 0) return QBX_OK;
    TD td = make_td(T);
    if (mode == 0)
        hipLaunchKernelGGL(k_eval_targets<0>, dim3(ntb), dim3(64), 0, st, kd, td, ntb,
                           R->target_boxes.p, R->l1_starts.p, R->l1_lists.p,
                           G->nqbtl_box_target_starts.p, G->nqbtl_box_target_counts_nonchild.p,
                           G->nqbtl_unfiltered.p, T->targets.p, T->ntargets, T->sources.p,
                           T->nsources, str, nrm, expn, pot);
    else if (mode == 1)
        hipLaunchKernelGGL(k_eval_targets<1>, dim3(ntb), dim3(64), 0, st, kd, td, ntb,
                           R->target_boxes.p, R->l3_starts.p, R->l3_lists.p,
                           G->nqbtl_box_target_starts.p, G->nqbtl_box_target_counts_nonchild.p,
                           G->nqbtl_unfiltered.p, T->targets.p, T->ntargets, T->sources.p,
                           T->nsources, str, nrm, expn, pot);
    else
        hipLaunchKernelGGL(k_eval_targets<2>, dim3(ntb), dim3(64), 0, st, kd, td, ntb,
                           R->target_boxes.p, R->l1_starts.p, R->l1_lists.p,
                           G->nqbtl_box_target_starts.p, G->nqbtl_box_target_counts_nonchild.p,
                           G->nqbtl_unfiltered.p, T->targets.p, T->ntargets, T->sources.p,
                           T->nsources, str, nrm, expn, pot);
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
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_HIP(hipMemsetAsync(local, 0, sizeof(double) * 2 * (size_t)T->nboxes * kd.ncf, st));
    long nthr = (long)R->nttp_boxes * kd.ncf;
    if (nthr > 0)
        hipLaunchKernelGGL(k_m2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B,
                           R->nttp_boxes, R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole,
                           local);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_kernel_desc *d, const double *str, const double *nrm,
                    double *local, int accumulate)
{
    QBX_CHECK(ctx && T && R && str && local, "qbx_form_locals: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(local, 0, sizeof(double) * 2 * (size_t)T->nboxes * kd.ncf, st));
    long nthr = (long)R->nttp_boxes * kd.ncf;
    if (nthr > 0 && R->n4 > 0)
        hipLaunchKernelGGL(k_p2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T),
                           R->nttp_boxes, R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p,
                           T->sources.p, T->nsources, str, nrm, local);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                      const qbx_kernel_desc *d, double *local)
{
    QBX_CHECK(ctx && T && R && local, "qbx_refine_locals: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    for (int l = 1; l < T->nlevels; ++l) {
        int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        long nthr = (long)(i1 - i0) * kd.ncf;
        if (nthr > 0)
            hipLaunchKernelGGL(k_l2l, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, make_td(T), B,
                               i0, i1, R->ttp_boxes.p, local);
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                               const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                               const double *nrm, double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && str && qexp, "qbx_form_global_qbx_locals: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    int ng = G->nglobal;
    if (ng == 0) return QBX_OK;
    DISPATCH_PQ(kd.pq, k_p2qbxl, dim3(cdiv(ng, 64)), dim3(64), 0, st, kd, make_td(T), ng,
                G->global_qbx_centers.p, G->center_to_target_box.p, G->centers.p, G->radii.p,
                G->ncenters, R->l1_starts.p, R->l1_lists.p, T->sources.p, T->nsources, str, nrm,
                qexp, accumulate);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                              const qbx_trav *R, const qbx_geo *G,
                                              const qbx_kernel_desc *d, const double *mpole,
                                              double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && mpole && qexp, "M2QBXL: null argument");
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    int ng = G->nglobal;
    if (ng == 0 || R->n3 == 0) return QBX_OK;
    DISPATCH_PQ(kd.pq, k_m2qbxl, dim3(cdiv(ng, 64)), dim3(64), 0, st, kd, make_td(T), B, ng,
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
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_HIP(hipMemsetAsync(qexp, 0, sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq, st));
    int ng = G->nglobal;
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
    int nt = T->ntargets, nc = G->ncenters;
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
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    long nm = 2L * T->nboxes * kd.ncf;
    long nq = 2L * std::max(G->ncenters, 1) * kd.ncq;
    long np = 2L * std::max(T->ntargets, 1);
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

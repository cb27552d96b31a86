// QBX-FMM wrangler stages for 2D Laplace and Helmholtz (pytential/qbx/fmm.py
// drive_fmm and pytential/qbx/interactions.py), FP64 / complex128.
//
// Expansions:
//  Laplace   G = -log|x-y|/(2 pi); complex potential Phi, phi = -Re(Phi)/(2 pi)
//            multipole a0 log(z-c) + sum_k a_k (R/(z-c))^k, local sum_l b_l ((z-c)/R)^l
//            (strengths must be real: the real part of Phi is taken)
//  Helmholtz G = (i/4) H0(k|x-y|); multipole sum_n M_n H_n(k rho) e^{in theta},
//            local sum_n L_n J_n(k rho) e^{in theta}; translations by Graf's
//            addition theorem (M2M/L2L with J, M2L/M2QBXL with H).
// The double layer is the source-normal derivative of G.
//
// Right-hand sides: nr densities share one evaluation.  Every per-source,
// per-coefficient and per-target array carries its nr values innermost, and
// every kernel computes the geometry of an interaction (distances, Bessel and
// Hankel values, translation weights) once and applies it to a chunk of up to
// NRC right-hand sides held in registers; the host launches one chunk
// [r0, r0 + rc) at a time.
//
// Storage (HBM): box expansions are level-compact.  Level l holds its boxes
// contiguously with p_l + 1 (Laplace) or 2 p_l + 1 (Helmholtz) coefficients
// each (times nr); coefficient n of a Helmholtz box sits at n + p_l.  QBX
// expansions are [ncenters][ncoeff(p_q)][nr].
//
// Thread mappings:
//  P2M, P2L        one thread per (box, |n|)
//  M2M, L2L, M2L   one thread per (box, output coefficient); Helmholtz
//                  translation vectors come from per-level tables (4 child
//                  offsets for M2M/L2L, the offset classes of list 2 for
//                  M2L), so each output is one dot product
//  P2P, M2P, L2P   one block per target box, lane groups per target
//  P2QBXL          16 lanes per QBX center over its list-1 sources, shuffle
//                  reduction, order-p_q expansion in registers
//  M2QBXL          4 lanes per QBX center over its list-3 boxes
//  L2QBXL          one thread per QBX center
//  QBXL2P          one thread per (center, target) pair
#include <algorithm>

#include "qbx_internal.h"
#include "qbx_math.h"

#include <cstring>
#include <mutex>
#include <type_traits>

namespace qbx {

constexpr int MAXP = 120;
constexpr int BINOM_N = 2 * MAXP + 4;
constexpr int FML = 30;        // levels of a 2D tree (coordinates have 28 bits)
constexpr int NRC_S = 8;       // RHS chunk of the scalar-accumulator kernels
constexpr int NRC_Q = 4;       // RHS chunk of the QBX-expansion kernels

struct KD {
    int kernel, dip, pq, nlevels, ncq;
    int nsep;                  // well_sep_is_n_away of the traversal
    int m2l_w;                 // list-2 offsets satisfy |dI| <= m2l_w = 2 nsep + 1
    int nr;                    // right-hand sides (stride of the innermost axis)
    int r0, rc;                // the chunk of right-hand sides of this launch
    double k;
    double ik;                 // 1 / k (Helmholtz)
    int order[FML];
    int ncl[FML];
    int lbox[FML + 1];
    long lcoef[FML + 1];
    double rscale[FML];
    const double *tab_m2l;     // Helmholtz M2L table (centred rows)
    const double *tab_mm;      // Helmholtz M2M/L2L table (centred rows)
    long m2l_off[FML];
    long mm_off[FML];
    int tder;                  // evaluate the derivative along tdir at the targets
    int ntgt;                  // stride of tdir (the tree's targets)
    const double *tdir;        // [2][ntgt] tree order
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

// Blocks over the boxes of several levels, each level's range padded to a
// multiple of the 8 XCDs.  The dispatcher deals blocks to the XCDs in turn
// (block b to XCD b % 8), so renumbering within a level gives each XCD one
// contiguous eighth of every level (a bijection whatever the dispatch
// order): neighbouring boxes, which read the same list-2 multipoles, share
// one L2, and every XCD still gets its share of each level's work.
struct FlatX {
    int nl;
    long tstart[FML + 1];  // padded block offset of each level
    int istart[FML];       // first box (index into the level-ordered boxes)
    int count[FML];        // boxes of the level
};

__device__ __forceinline__ bool flatx_locate(const FlatX &f, long b, int &lv, int &i)
{
    if (b >= f.tstart[f.nl]) return false;
    int l = 0;
    while (b >= f.tstart[l + 1]) ++l;
    const long r = b - f.tstart[l];
    const long per = (f.tstart[l + 1] - f.tstart[l]) / 8;
    const long k = (r % 8) * per + r / 8;
    if (k >= f.count[l]) return false;
    lv = l;
    i = f.istart[l] + (int)k;
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

// coefficient 0 (Helmholtz: n = 0, centred), right-hand side 0 of box b;
// coefficient n, right-hand side r at [2 (n nr + r)]
__device__ __forceinline__ long ecoef_off(const KD &kd, int b, int lv)
{
    const long off = kd.lcoef[lv] + (long)(b - kd.lbox[lv]) * kd.ncl[lv];
    return 2 * (off + (kd.kernel == QBX_HELMHOLTZ2D ? kd.order[lv] : 0)) * kd.nr;
}

// strength of source s, right-hand side r
__device__ __forceinline__ cplx ldq(const KD &kd, const double *str, int s, int r)
{
    return ldc(str + 2 * ((size_t)s * kd.nr + r));
}

// coefficient n, right-hand side r relative to an ecoef_off pointer
__device__ __forceinline__ cplx ldcn(const KD &kd, const double *a, int n, int r)
{
    return ldc(a + 2 * ((long)n * kd.nr + r));
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
// pairwise kernel without the strength (potential = value * q): Laplace
// -log(r)/(2 pi) and its dipole Re(nu / w)/(2 pi); Helmholtz (i/4) H0(k r)
// and its dipole (i/4) k (w . nu / r) H1(k r), w = t - s.  The Hankel series
// length is chosen per wavefront; coincident points give zero.
// -------------------------------------------------------------------------
//
// DER: the derivative of that value along the target direction d instead
// (complex d = dx + i dy; w^ = w / r):
//   Laplace S  -Re(d / w) / (2 pi)          Laplace D  -Re(nu d / w^2) / (2 pi)
//   Helmholtz S  -(i/4) k (d . w^) H1(k r)
//   Helmholtz D  (i/4) k [H1(k r) (d . nu - (d . w^)(nu . w^)) / r
//                         + k (nu . w^)(d . w^) (H0(k r) - H1(k r) / (k r))]
// (H1' = H0 - H1 / x).
// KN, DP: the kernel and source kind fixed at compile time (-1: read from kd);
// SC: Hankel series coefficients as scalar operands (off: an opaque
// wave-uniform zero, see hankel01_wave_sc) instead of from the LDS table tab.
template <int KN, int DP, bool DER, bool SC>
__device__ __forceinline__ cplx p2p_value_impl(const KD &kd, cplx s, cplx nu, cplx t,
                                               const double *tab, int off, cplx dir)
{
    const int kern = KN >= 0 ? KN : kd.kernel;
    const bool dip = DP >= 0 ? DP != 0 : kd.dip != 0;
    const cplx w = t - s;
    const double r2 = abs2(w);
    const bool zero = r2 == 0.0;
    const cplx ws = zero ? cplx{1.0, 0.0} : w;
    const double r2s = zero ? 1.0 : r2;
    cplx v;
    if (kern == QBX_LAPLACE2D) {
        if (!DER) {
            v = !dip ? cplx{-0.5 * qbx_log(r2s) * INV2PI, 0.0}
                     : cplx{(nu * cinv(ws)).x * INV2PI, 0.0};
        } else {
            const cplx iw = cinv(ws);
            v = !dip ? cplx{-(dir * iw).x * INV2PI, 0.0}
                     : cplx{-((nu * dir) * (iw * iw)).x * INV2PI, 0.0};
        }
    } else {
        const double ir = qbx_rsqrt(r2s);
        const double r = r2s * ir;
        const double kr = kd.k * r;
        const double ikr = ir * kd.ik;
        cplx h0, h1;
        auto hank = [&](auto w0, auto w1) {
            constexpr bool W0 = decltype(w0)::value, W1 = decltype(w1)::value;
            if (SC)
                hankel01_wave_sc<W0, W1>(kr, ikr, off, h0, h1);
            else
                hankel01_wave<W0, W1>(kr, ikr, tab, h0, h1);
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        if (!DER) {
            if (!dip) {
                hank(T_{}, F_{});
                v = {-0.25 * h0.y, 0.25 * h0.x};
            } else {
                hank(F_{}, T_{});
                const double proj = (ws.x * nu.x + ws.y * nu.y) * ir;
                const cplx hv = (kd.k * proj) * h1;
                v = {-0.25 * hv.y, 0.25 * hv.x};
            }
        } else {
            const double pd = (ws.x * dir.x + ws.y * dir.y) * ir;
            if (!dip) {
                hank(F_{}, T_{});
                const cplx hv = (-kd.k * pd) * h1;
                v = {-0.25 * hv.y, 0.25 * hv.x};
            } else {
                hank(T_{}, T_{});
                const double pn = (ws.x * nu.x + ws.y * nu.y) * ir;
                const double dn = dir.x * nu.x + dir.y * nu.y;
                const cplx hv = kd.k * ((ir * (dn - pd * pn)) * h1 +
                                        (kd.k * pn * pd) * (h0 - ikr * h1));
                v = {-0.25 * hv.y, 0.25 * hv.x};
            }
        }
    }
    return zero ? cplx{0.0, 0.0} : v;
}

template <bool DER = false>
__device__ __forceinline__ cplx p2p_value_wave(const KD &kd, cplx s, cplx nu, cplx t,
                                            const double *tab, cplx dir = {0.0, 0.0})
{
    return p2p_value_impl<-1, -1, DER, false>(kd, s, nu, t, tab, 0, dir);
}

// the target direction of tree target ti (zero when no derivative is taken)
__device__ __forceinline__ cplx ldir(const KD &kd, int ti)
{
    return kd.tder ? cplx{kd.tdir[ti], kd.tdir[(size_t)kd.ntgt + ti]} : cplx{0.0, 0.0};
}

// -------------------------------------------------------------------------
// P2M / P2L: the basis value(s) one source contributes to coefficient(s)
// -------------------------------------------------------------------------
// Helmholtz: coefficients n = m and n = -m (m >= 0) of the expansion about c
// of the source at v + c.  g_j = Z_j(k rho) e^{-i j theta} of v, Z = H for
// locals and J for multipoles, g_{-j} = (-1)^j Z_j e^{+i j theta}.
__device__ void helm_p2e_basis(const KD &kd, int local, int m, cplx v, cplx nu, cplx &bp,
                               cplx &bm)
{
    const int top = m + 1;
    const double sm = (m & 1) ? -1.0 : 1.0;  // (-1)^m
    const double rho = sqrt(abs2(v));
    const double x = kd.k * rho;
    const cplx e1 = rho > 0 ? cplx{v.x / rho, -v.y / rho} : cplx{1, 0};  // e^{-i theta}
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
    if (!kd.dip) {
        bp = z0 * em;
        bm = sm * (z0 * conj(em));
    } else {
        const double hk = 0.5 * kd.k;
        // g_{m-1}: for m = 0 it is g_{-1} = -Z_1 e^{+i theta}
        const cplx gm1 = m >= 1 ? zlo * (em * conj(e1)) : -(zlo * conj(e1));
        const cplx gp1 = zp1 * emp1;
        bp = hk * (conj(nu) * gm1 - nu * gp1);
        // n = -m: g_{-m-1} = (-1)^{m+1} Z_{m+1} e^{+i(m+1)th},
        //         g_{-m+1} = (-1)^{m-1} Z_{m-1} e^{+i(m-1)th}
        const cplx gA = (-sm) * (zp1 * conj(emp1));
        const cplx gB = (-sm) * (zlo * conj(em * conj(e1)));
        bm = hk * (conj(nu) * gA - nu * gB);
    }
}

// Laplace: coefficient l of one source at w + c (R the expansion's rscale)
__device__ __forceinline__ cplx lap_p2e_basis(const KD &kd, int local, int l, cplx w, double R,
                                              cplx nu)
{
    if (local) {
        const cplx pw = cpow_int(R * cinv(w), l);
        if (!kd.dip) return l == 0 ? clogz(-w) : (-1.0 / l) * pw;
        return (nu * cinv(w)) * pw;
    }
    const cplx t = (1.0 / R) * w;
    if (!kd.dip) return l == 0 ? cplx{1, 0} : (-1.0 / l) * cpow_int(t, l);
    if (l == 0) return {0, 0};
    return -((1.0 / R) * nu) * cpow_int(t, l - 1);
}

// contributions of sources [s0, s1) to coefficient j (and -j for Helmholtz)
template <int NRC>
__device__ __forceinline__ void p2e_sources(const KD &kd, int local, int j, cplx c, double R,
                                            const double *src, int ns, int s0, int s1,
                                            const double *str, const double *nrm, cplx *ap,
                                            cplx *am)
{
    for (int s = s0; s < s1; ++s) {
        const cplx w = ldpt(src, ns, s) - c;
        const cplx nu = kd.dip ? ldpt(nrm, ns, s) : cplx{0, 0};
        if (kd.kernel == QBX_LAPLACE2D) {
            const cplx b = lap_p2e_basis(kd, local, j, w, R, nu);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) ap[q] += ldq(kd, str, s, kd.r0 + q) * b;
        } else {
            cplx bp, bm;
            helm_p2e_basis(kd, local, j, w, nu, bp, bm);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) {
                    const cplx qs = ldq(kd, str, s, kd.r0 + q);
                    ap[q] += qs * bp;
                    am[q] += qs * bm;
                }
        }
    }
}

// P2M over source leaf boxes (flat over levels, p+1 threads per box)
template <int NRC>
__global__ void k_p2m(KD kd, TD t, Flat fl, const int *boxes, const double *src, int ns,
                      const double *str, const double *nrm, double *mpole)
{
    int lv, i, j;
    if (!flat_locate(fl, blockIdx.x * (long)blockDim.x + threadIdx.x, lv, i, j)) return;
    const int b = boxes[i];
    const int s0 = t.sstart[b], s1 = s0 + t.scount[b];
    cplx ap[NRC], am[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) ap[q] = am[q] = {0, 0};
    p2e_sources<NRC>(kd, 0, j, boxc(t, b), kd.rscale[lv], src, ns, s0, s1, str, nrm, ap, am);
    double *a = mpole + ecoef_off(kd, b, lv);
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) {
            const int r = kd.r0 + q;
            stc(a + 2 * ((long)j * kd.nr + r), ap[q]);
            if (helm && j > 0) stc(a + 2 * ((long)-j * kd.nr + r), am[q]);
        }
}

// P2L from list 4 (flat over ttp levels, p+1 threads per box), accumulates
template <int NRC>
__global__ void k_p2l(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
                      const double *src, int ns, const double *str, const double *nrm,
                      double *local)
{
    int lv, i, j;
    if (!flat_locate(fl, blockIdx.x * (long)blockDim.x + threadIdx.x, lv, i, j)) return;
    if (starts[i] == starts[i + 1]) return;
    const int P = ttp[i];
    const cplx c = boxc(t, P);
    cplx ap[NRC], am[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) ap[q] = am[q] = {0, 0};
    for (int e = starts[i]; e < starts[i + 1]; ++e) {
        const int Sb = lists[e];
        const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
        p2e_sources<NRC>(kd, 1, j, c, kd.rscale[lv], src, ns, s0, s1, str, nrm, ap, am);
    }
    double *a = local + ecoef_off(kd, P, lv);
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) {
            const int r = kd.r0 + q;
            double *o = a + 2 * ((long)j * kd.nr + r);
            stc(o, ldc(o) + ap[q]);
            if (helm && j > 0) {
                double *o2 = a + 2 * ((long)-j * kd.nr + r);
                stc(o2, ldc(o2) + am[q]);
            }
        }
}

// -------------------------------------------------------------------------
// Helmholtz P2M / P2L through LDS, one block per box.  Sources are taken S
// at a time.  Phase A (one lane per source) runs the source's sequence once:
// J_j by Miller's recurrence (multipoles) or H_j by forward recurrence
// (locals), j = 0..p+1, and stores g+_j = Z_j e^{-ij theta} and
// g-_j = Z_j e^{+ij theta} of v = source - center.  Phase B (one lane per
// coefficient n in [-p, p]) sums q_s b_n(s) over the chunk, with
// g_{-j} = (-1)^j g-_j and b_n = g_n (single layer) or
// (k/2)(conj(nu) g_{n-1} - nu g_{n+1}) (double layer): O(p) work per source
// instead of a recurrence per coefficient.
// -------------------------------------------------------------------------
struct P2EShape {
    int S;  // sources per chunk
    int R;  // row stride of g+- in complex entries (odd, >= p + 2)
    __host__ __device__ int doubles(int nrc) const { return 4 * S * R + 2 * S * nrc + 2 * S + S; }
};

template <int NRC, bool LOCAL>
__global__ void __launch_bounds__(256) k_p2e_h(KD kd, TD t, int lv, int i0, P2EShape sh, const int *boxes,
                        const int *starts, const int *lists, const double *src, int ns,
                        const double *str, const double *nrm, double *expn)
{
    extern __shared__ double lds[];
    const int i = i0 + blockIdx.x;
    const int Bx = boxes[i];
    int e0 = 0, e1 = 1;
    if (LOCAL) {
        e0 = starts[i];
        e1 = starts[i + 1];
        if (e0 == e1) return;
    }
    const int p = kd.order[lv];
    const int S = sh.S, R = sh.R;
    double *Gp = lds, *Gm = Gp + 2 * S * R, *Qs = Gm + 2 * S * R, *Nu = Qs + 2 * S * NRC;
    int *idx = (int *)(Nu + 2 * S);
    const int tid = threadIdx.x, nth = blockDim.x;
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    const cplx c = boxc(t, Bx);
    const bool act = tid < 2 * p + 1;
    const int n = tid - p;
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};

    auto gval = [&](int sl, int j) -> cplx {
        if (j >= 0) return {Gp[2 * (sl * R + j)], Gp[2 * (sl * R + j) + 1]};
        const cplx g = {Gm[2 * (sl * R - j)], Gm[2 * (sl * R - j) + 1]};
        return (j & 1) ? -g : g;
    };

    auto flush = [&](int cnt) {
        __syncthreads();
        if (tid < cnt) {
            const int si = idx[tid];
            double *gp = Gp + 2 * tid * R, *gm = Gm + 2 * tid * R;
            const cplx v = ldpt(src, ns, si) - c;
            const double rho = sqrt(abs2(v));
            const double x = kd.k * rho;
            const cplx e1 = rho > 0 ? cplx{v.x / rho, -v.y / rho} : cplx{1, 0};
            if (LOCAL) {
                cplx a, b;
                hankel01_call(x, a, b);
                const double tox = 2.0 / x;
                for (int j = 0; j <= p + 1; ++j) {
                    gp[2 * j] = a.x;
                    gp[2 * j + 1] = a.y;
                    const cplx nb = ((j + 1) * tox) * b - a;
                    a = b;
                    b = nb;
                }
            } else if (x < 1e-300) {
                for (int j = 0; j <= p + 1; ++j) {
                    gp[2 * j] = j == 0 ? 1.0 : 0.0;
                    gp[2 * j + 1] = 0.0;
                }
            } else {
                const double norm = miller_j(
                    x, p + 1, [&](int j, double f) { gp[2 * j] = f; },
                    [&](double sc) {
                        for (int j = 0; j <= p + 1; ++j) gp[2 * j] *= sc;
                    });
                for (int j = 0; j <= p + 1; ++j) {
                    gp[2 * j] *= norm;
                    gp[2 * j + 1] = 0.0;
                }
            }
            cplx e = {1, 0};
            for (int j = 0; j <= p + 1; ++j) {
                const cplx z = {gp[2 * j], gp[2 * j + 1]};
                const cplx zp = z * e, zm = z * conj(e);
                gp[2 * j] = zp.x;
                gp[2 * j + 1] = zp.y;
                gm[2 * j] = zm.x;
                gm[2 * j + 1] = zm.y;
                e = e * e1;
            }
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                    Qs[2 * (tid * NRC + q)] = pq[0];
                    Qs[2 * (tid * NRC + q) + 1] = pq[1];
                }
            if (dip) {
                Nu[2 * tid] = nrm[si];
                Nu[2 * tid + 1] = nrm[(size_t)ns + si];
            }
        }
        __syncthreads();
        if (act) {
            const double hk = 0.5 * kd.k;
            for (int sl = 0; sl < cnt; ++sl) {
                cplx b;
                if (!dip) {
                    b = gval(sl, n);
                } else {
                    const cplx nu = {Nu[2 * sl], Nu[2 * sl + 1]};
                    b = hk * (conj(nu) * gval(sl, n - 1) - nu * gval(sl, n + 1));
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc)
                        acc[q] += cplx{Qs[2 * (sl * NRC + q)], Qs[2 * (sl * NRC + q) + 1]} * b;
            }
        }
    };

    int fill = 0;
    for (int e = e0; e < e1; ++e) {
        const int Sb = LOCAL ? lists[e] : Bx;
        int s = t.sstart[Sb];
        const int s1 = s + t.scount[Sb];
        while (s < s1) {
            const int take = min(s1 - s, S - fill);
            for (int j = tid; j < take; j += nth) idx[fill + j] = s + j;
            fill += take;
            s += take;
            if (fill == S) {
                flush(S);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
    if (act) {
        double *o = expn + ecoef_off(kd, Bx, lv) + 2L * n * nr;
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc) {
                double *pq = o + 2 * (r0 + q);
                stc(pq, LOCAL ? ldc(pq) + acc[q] : acc[q]);
            }
    }
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
// Laplace translations, one output coefficient, a chunk of right-hand sides
// -------------------------------------------------------------------------
__device__ __forceinline__ double g_binom(const double *B, int n, int k)
{
    return (k < 0 || k > n) ? 0.0 : B[n * BINOM_N + k];
}

// M2M (c1 -> c2, z0 = c1 - c2), output l:
// -a0 u^l / l + sum_{k=1}^{min(l,p1)} a_k r^k u^{l-k} C(l-1,k-1), u = z0/R2, r = R1/R2
template <int NRC>
__device__ void lap_m2m(const KD &kd, const double *B, const double *a, int p1, int l, cplx z0,
                        double R1, double R2, cplx *acc)
{
    if (l == 0) {
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) acc[q] += ldcn(kd, a, 0, kd.r0 + q);
        return;
    }
    const cplx u = (1.0 / R2) * z0;
    const double r = R1 / R2;
    const cplx ul = cpow_int(u, l);
    const int kmax = l < p1 ? l : p1;
    cplx hsum[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) hsum[q] = {0, 0};
    double rk = 1.0;
    for (int k = 1; k <= kmax; ++k) {
        rk *= r;
        const double wk = rk * g_binom(B, l - 1, k - 1);
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) hsum[q] = hsum[q] * u + wk * ldcn(kd, a, k, kd.r0 + q);
    }
    const cplx ut = cpow_int(u, l - kmax);
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc)
            acc[q] += (-1.0 / l) * (ldcn(kd, a, 0, kd.r0 + q) * ul) + hsum[q] * ut;
}

// M2L (c1 multipole -> c2 local, z0 = c1 - c2), output l
template <int NRC>
__device__ void lap_m2l(const KD &kd, const double *B, const double *a, int p1, int l, cplx z0,
                        double R1, double R2, cplx *acc)
{
    const cplx iz = cinv(z0);
    const cplx ms = -(R1 * iz), t = R2 * iz;
    cplx s[NRC];
    if (l == 0) {
        const cplx lg = clogz(-z0);
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            s[q] = q < kd.rc ? ldcn(kd, a, 0, kd.r0 + q) * lg : cplx{0, 0};
        cplx pw = {1, 0};
        for (int k = 1; k <= p1; ++k) {
            pw = pw * ms;
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) s[q] += ldcn(kd, a, k, kd.r0 + q) * pw;
        }
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) acc[q] += s[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        s[q] = q < kd.rc ? (-1.0 / l) * ldcn(kd, a, 0, kd.r0 + q) : cplx{0, 0};
    cplx pw = {1, 0};
    for (int k = 1; k <= p1; ++k) {
        pw = pw * ms;
        const cplx wk = g_binom(B, l + k - 1, k - 1) * pw;
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) s[q] += ldcn(kd, a, k, kd.r0 + q) * wk;
    }
    const cplx tl = cpow_int(t, l);
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) acc[q] += s[q] * tl;
}

// L2L (c1 -> c2, z0 = c2 - c1), output j
template <int NRC>
__device__ void lap_l2l(const KD &kd, const double *B, const double *a, int p1, int j, cplx z0,
                        double R1, double R2, cplx *acc)
{
    if (j > p1) return;
    const cplx u = (1.0 / R1) * z0;
    cplx s[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) s[q] = {0, 0};
    for (int l = p1; l >= j; --l) {
        const double bl = g_binom(B, l, j);
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) s[q] = s[q] * u + bl * ldcn(kd, a, l, kd.r0 + q);
    }
    double rp = 1.0;
    const double r21 = R2 / R1;
    for (int i = 0; i < j; ++i) rp *= r21;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) acc[q] += rp * s[q];
}

// Helmholtz table translation: acc[q] += sum_n a[n] row[n - m] ((-1)^{n-m} if flip).
// One right-hand side: four partial sums over n mod 4, so that four
// coefficient/row load pairs are in flight and the multiply-add chains are
// independent (the loop is latency-bound, not throughput-bound).
template <int NRC>
__device__ __forceinline__ void helm_row_dot(const KD &kd, const double *a, int p1, int m,
                                             const double *row, bool flip, cplx *acc)
{
    if (NRC == 1) {
        const int r = kd.r0;
        cplx s4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        int n = -p1;
        for (; n + 3 <= p1; n += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cplx w = ldc(row + 2 * (n + u - m));
                if (flip && ((n + u - m) & 1)) w = -w;
                s4[u] += ldcn(kd, a, n + u, r) * w;
            }
        }
        for (; n <= p1; ++n) {
            cplx w = ldc(row + 2 * (n - m));
            if (flip && ((n - m) & 1)) w = -w;
            s4[0] += ldcn(kd, a, n, r) * w;
        }
        acc[0] += (s4[0] + s4[1]) + (s4[2] + s4[3]);
        return;
    }
    for (int n = -p1; n <= p1; ++n) {
        cplx w = ldc(row + 2 * (n - m));
        if (flip && ((n - m) & 1)) w = -w;
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) acc[q] += ldcn(kd, a, n, kd.r0 + q) * w;
    }
}

// o: coefficient (or target) pointer for right-hand side 0
template <int NRC>
__device__ __forceinline__ void store_add(const KD &kd, double *o, const cplx *acc)
{
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) {
            double *p = o + 2 * (kd.r0 + q);
            stc(p, ldc(p) + acc[q]);
        }
}

// -------------------------------------------------------------------------
// M2M (parents at one level), M2L (list 2), L2L (one level)
// -------------------------------------------------------------------------
template <int NRC>
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
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int c = 0; c < 4; ++c) {
        const int C = t.child[(size_t)c * t.nboxes + P];
        if (C < 0 || t.scumul[C] == 0) continue;
        const double *a = mpole + ecoef_off(kd, C, lvc);
        if (!helm) {
            lap_m2m<NRC>(kd, B, a, p1, m, boxc(t, C) - cp, kd.rscale[lvc], kd.rscale[lvp], acc);
        } else {
            const int L = p1 + p2;
            const double *row = kd.tab_mm + 2 * (kd.mm_off[lvc] + (long)c * (2 * L + 1) + L);
            helm_row_dot<NRC>(kd, a, p1, m, row, false, acc);
        }
    }
    store_add<NRC>(kd, mpole + ecoef_off(kd, P, lvp) + 2L * m * kd.nr, acc);
}

// M2L with one block per target-or-target-parent box (fl: one entry per box):
// thread t takes output coefficient j = t % ncl and every S-th list-2 entry
// from t / ncl (S = M2L_THREADS / ncl groups), so a box's entries are split
// S ways instead of one thread walking all of them -- the long serial
// chains of the coarse levels (ncl ~ 60-150 outputs, ~27 entries each) were
// the tail of the one-thread-per-coefficient kernel.  The groups' partial
// sums are added in LDS in group order (deterministic), then stored.
constexpr int M2L_THREADS = 256;

template <int NRC>
__global__ void __launch_bounds__(M2L_THREADS)
k_m2l_blk(KD kd, TD t, const double *B, FlatX fl, const int *ttp, const int *starts,
          const int *lists, const double *mpole, double *local)
{
    __shared__ cplx red[M2L_THREADS * NRC];
    int lv, i;
    if (!flatx_locate(fl, blockIdx.x, lv, i)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int ncl = kd.ncl[lv];
    const int S = M2L_THREADS / ncl;
    const int tid = threadIdx.x;
    const int g = tid / ncl, j = tid % ncl;
    const int P = ttp[i];
    const int p = kd.order[lv];
    const bool helm = kd.kernel == QBX_HELMHOLTZ2D;
    const int m = helm ? j - p : j;
    const cplx cp = boxc(t, P);
    const int Ix = t.coords[P], Iy = t.coords[(size_t)t.nboxes + P];
    const int L = 2 * p;
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    if (g < S) {
        for (int e = e0 + g; e < e1; e += S) {
            const int A = lists[e];
            const double *a = mpole + ecoef_off(kd, A, lv);
            if (!helm) {
                lap_m2l<NRC>(kd, B, a, p, m, boxc(t, A) - cp, kd.rscale[lv], kd.rscale[lv],
                             acc);
            } else {
                const int dx = Ix - t.coords[A], dy = Iy - t.coords[(size_t)t.nboxes + A];
                const int cls = (dx + kd.m2l_w) * (2 * kd.m2l_w + 1) + (dy + kd.m2l_w);
                const double *row =
                    kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
                helm_row_dot<NRC>(kd, a, p, m, row, false, acc);
            }
        }
        if (g > 0) {
#pragma unroll
            for (int q = 0; q < NRC; ++q) red[(tid - ncl) * NRC + q] = acc[q];
        }
    }
    __syncthreads();
    if (g == 0) {
        for (int h = 1; h < S; ++h)
#pragma unroll
            for (int q = 0; q < NRC; ++q) acc[q] += red[((h - 1) * ncl + j) * NRC + q];
        store_add<NRC>(kd, local + ecoef_off(kd, P, lv) + 2L * m * kd.nr, acc);
    }
}

// ---- M2L on the FP64 matrix cores (Helmholtz, many right-hand sides) ------
// For a target box P and its list-2 boxes A_e:  L_P[m][r] += sum_e sum_n
// T_e[m][n] M_Ae[n][r], T_e[m][n] = row_e[n - m] (Toeplitz, from the
// translation table), r over up to 16 right-hand sides: per entry a complex
// (ncl x ncl) x (ncl x 16) product, tiled for v_mfma_f64_16x16x4f64 (lane i
// holds A[i%16][k0 + i/16] and B[k0 + i/16][i%16]; D register j of lane i is
// D[4j + i/16][i%16]).  The complex product is four real MFMAs (A_im negated
// for the real part).  One block per box; its four waves take every fourth
// entry, each stages the entry's row window in LDS (zero outside the row) and
// reads the multipole coefficients straight from HBM (16 right-hand sides of
// a coefficient are contiguous); the waves' partial sums are added in LDS in
// wave order, then stored.  Levels with ncl = 2p + 1 <= M2LM_ROWS * 16.
typedef double m2l_d4 __attribute__((ext_vector_type(4)));
constexpr int M2LM_ROWS = 3;                  // row tiles of 16 (ncl <= 48)
constexpr int M2LM_WIN = M2LM_ROWS * 16 + 48; // staged row window (complex)

__global__ void __launch_bounds__(256)
k_m2l_mfma(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
           const double *mpole, double *local)
{
    __shared__ double win[4][2 * M2LM_WIN];
    __shared__ m2l_d4 red[3][M2LM_ROWS][2][64];
    int lv, i, unused;
    if (!flat_locate(fl, blockIdx.x, lv, i, unused)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int p = kd.order[lv];
    const int ncl = kd.ncl[lv];
    const int RT = (ncl + 15) / 16, KS = (ncl + 3) / 4;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int P = ttp[i];
    const int Ix = t.coords[P], Iy = t.coords[(size_t)t.nboxes + P];
    const int L = 2 * p;
    const int rc = kd.rc, r0 = kd.r0;
    const int col = lane & 15, kr = lane >> 4;
    // window index w <-> row offset d = w - (RT * 16 - 1): n - m in [-(16 RT - 1), 4 KS - 1]
    const int dlo = -(RT * 16 - 1), nwin = RT * 16 + KS * 4 - 1;
    m2l_d4 cre[M2LM_ROWS], cim[M2LM_ROWS];
#pragma unroll
    for (int rt = 0; rt < M2LM_ROWS; ++rt) cre[rt] = cim[rt] = m2l_d4{0, 0, 0, 0};
    double *wn = win[wv];
    for (int e = e0 + wv; e < e1; e += 4) {
        const int A = lists[e];
        const int dx = Ix - t.coords[A], dy = Iy - t.coords[(size_t)t.nboxes + A];
        const int cls = (dx + kd.m2l_w) * (2 * kd.m2l_w + 1) + (dy + kd.m2l_w);
        const double *row = kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
        const double *a = mpole + ecoef_off(kd, A, lv);
        for (int w = lane; w < nwin; w += 64) {
            const int d = w + dlo;
            const bool in = d >= -L && d <= L;
            wn[2 * w] = in ? row[2 * d] : 0.0;
            wn[2 * w + 1] = in ? row[2 * d + 1] : 0.0;
        }
        // the window is written and read by this wave only: LDS keeps one
        // wave's accesses in order; the barrier keeps the compiler's too
        __builtin_amdgcn_wave_barrier();
        for (int ks = 0; ks < KS; ++ks) {
            const int n = ks * 4 + kr;  // 0-based coefficient of the source
            double bre = 0.0, bim = 0.0;
            if (n < ncl && col < rc) {
                const cplx b = ldcn(kd, a, n - p, r0 + col);
                bre = b.x;
                bim = b.y;
            }
#pragma unroll
            for (int rt = 0; rt < M2LM_ROWS; ++rt) {
                if (rt < RT) {
                    const int m = rt * 16 + col;  // A row of this lane
                    const int w = n - m - dlo;
                    const double are = wn[2 * w], aim = wn[2 * w + 1];
                    cre[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[rt], 0, 0, 0);
                    cre[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[rt], 0, 0, 0);
                    cim[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[rt], 0, 0, 0);
                    cim[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[rt], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (wv > 0) {
#pragma unroll
        for (int rt = 0; rt < M2LM_ROWS; ++rt) {
            red[wv - 1][rt][0][lane] = cre[rt];
            red[wv - 1][rt][1][lane] = cim[rt];
        }
    }
    __syncthreads();
    if (wv == 0) {
        for (int h = 0; h < 3; ++h)
#pragma unroll
            for (int rt = 0; rt < M2LM_ROWS; ++rt) {
                cre[rt] += red[h][rt][0][lane];
                cim[rt] += red[h][rt][1][lane];
            }
        double *o = local + ecoef_off(kd, P, lv);
#pragma unroll
        for (int rt = 0; rt < M2LM_ROWS; ++rt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = rt * 16 + 4 * j + kr;  // D row of register j
                if (rt < RT && m < ncl && col < rc) {
                    double *q = o + 2 * ((long)(m - p) * kd.nr + r0 + col);
                    q[0] += cre[rt][j];
                    q[1] += cim[rt][j];
                }
            }
    }
}

// One-density M2L grouped by translation class (Helmholtz, levels with
// ncl = 2p + 1 <= 48).  All list-2 pairs of one level and integer offset
// class share one Toeplitz operator T[m][n] = row[n - m], so a work item of up
// to 16 such pairs is one complex (ncl x ncl) x (ncl x 16) product with the
// pairs' source multipoles as the columns, on v_mfma_f64_16x16x4f64 (lane i
// holds A[i%16][k0 + i/16] and B[k0 + i/16][i%16]; D register j of lane i is
// D[4j + i/16][i%16]); a wavefront per item stages the row window in LDS.
// Column c's product goes to the pair's scratch row (at yoff[first + c],
// ncl complex); k_m2l_cls_sum then adds each box's rows in list order, so
// the sums are deterministic.
constexpr int M2LG_NP = 16;  // pairs per work item
constexpr int M2LM_RTW = 4;  // row tiles per wave of the four-wave kernels (ncl <= 256)
constexpr int M2LM_WIN2 = 16 * 4 * M2LM_RTW + 256;

__global__ void __launch_bounds__(64)
k_m2l_cls(KD kd, TD t, int nitems, const int *items, const int *pe, const int *yoff,
          const int *lists, const double *mpole, double *Y)
{
    __shared__ double win[2 * (M2LM_ROWS * 16 + 48)];
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int lv = items[4 * it], cls = items[4 * it + 1];
    const int first = items[4 * it + 2], cnt = items[4 * it + 3];
    const int p = kd.order[lv];
    const int ncl = kd.ncl[lv];
    const int RT = (ncl + 15) / 16, KS = (ncl + 3) / 4;
    const int lane = threadIdx.x;
    const int col = lane & 15, kr = lane >> 4;
    const int L = 2 * p;
    const double *row = kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
    const int dlo = -(RT * 16 - 1), nwin = RT * 16 + KS * 4 - 1;
    for (int w = lane; w < nwin; w += 64) {
        const int d = w + dlo;
        const bool in = d >= -L && d <= L;
        win[2 * w] = in ? row[2 * d] : 0.0;
        win[2 * w + 1] = in ? row[2 * d + 1] : 0.0;
    }
    __syncthreads();
    // this lane's B column: the source multipole of pair `col`
    const double *a = nullptr;
    if (col < cnt) a = mpole + ecoef_off(kd, lists[pe[first + col]], lv);
    m2l_d4 cre[M2LM_ROWS], cim[M2LM_ROWS];
#pragma unroll
    for (int rt = 0; rt < M2LM_ROWS; ++rt) cre[rt] = cim[rt] = m2l_d4{0, 0, 0, 0};
    for (int ks = 0; ks < KS; ++ks) {
        const int n = ks * 4 + kr;
        double bre = 0.0, bim = 0.0;
        if (a && n < ncl) {
            const cplx b = ldcn(kd, a, n - p, kd.r0);
            bre = b.x;
            bim = b.y;
        }
#pragma unroll
        for (int rt = 0; rt < M2LM_ROWS; ++rt) {
            if (rt < RT) {
                const int m = rt * 16 + col;
                const int w = n - m - dlo;
                const double are = win[2 * w], aim = win[2 * w + 1];
                cre[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[rt], 0, 0, 0);
                cre[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[rt], 0, 0, 0);
                cim[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[rt], 0, 0, 0);
                cim[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[rt], 0, 0, 0);
            }
        }
    }
    if (col < cnt) {
        double *y = Y + 2L * yoff[first + col];
#pragma unroll
        for (int rt = 0; rt < M2LM_ROWS; ++rt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = rt * 16 + 4 * j + kr;
                if (rt < RT && m < ncl) {
                    y[2 * m] = cre[rt][j];
                    y[2 * m + 1] = cim[rt][j];
                }
            }
    }
}

// k_m2l_cls for the coarse levels (48 < ncl <= 256): four waves share the
// item's row window (staged once) and take the row tiles w, w + 4, ...
__global__ void __launch_bounds__(256)
k_m2l_cls_rows(KD kd, TD t, int nitems, const int *items, const int *pe, const int *yoff,
               const int *lists, const double *mpole, double *Y)
{
    __shared__ double win[2 * M2LM_WIN2];
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int lv = items[4 * it], cls = items[4 * it + 1];
    const int first = items[4 * it + 2], cnt = items[4 * it + 3];
    const int p = kd.order[lv];
    const int ncl = kd.ncl[lv];
    const int RT = (ncl + 15) / 16, KS = (ncl + 3) / 4;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int L = 2 * p;
    const double *row = kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
    const int dlo = -(RT * 16 - 1), nwin = RT * 16 + KS * 4 - 1;
    for (int w = threadIdx.x; w < nwin; w += 256) {
        const int d = w + dlo;
        const bool in = d >= -L && d <= L;
        win[2 * w] = in ? row[2 * d] : 0.0;
        win[2 * w + 1] = in ? row[2 * d + 1] : 0.0;
    }
    __syncthreads();
    const double *a = nullptr;
    if (col < cnt) a = mpole + ecoef_off(kd, lists[pe[first + col]], lv);
    m2l_d4 cre[M2LM_RTW], cim[M2LM_RTW];
#pragma unroll
    for (int q = 0; q < M2LM_RTW; ++q) cre[q] = cim[q] = m2l_d4{0, 0, 0, 0};
    for (int ks = 0; ks < KS; ++ks) {
        const int n = ks * 4 + kr;
        double bre = 0.0, bim = 0.0;
        if (a && n < ncl) {
            const cplx b = ldcn(kd, a, n - p, kd.r0);
            bre = b.x;
            bim = b.y;
        }
#pragma unroll
        for (int q = 0; q < M2LM_RTW; ++q) {
            const int rt = wv + 4 * q;
            if (rt < RT) {
                const int m = rt * 16 + col;
                const int w = n - m - dlo;
                const double are = win[2 * w], aim = win[2 * w + 1];
                cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[q], 0, 0, 0);
                cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[q], 0, 0, 0);
                cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[q], 0, 0, 0);
                cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[q], 0, 0, 0);
            }
        }
    }
    if (col < cnt) {
        double *y = Y + 2L * yoff[first + col];
#pragma unroll
        for (int q = 0; q < M2LM_RTW; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int rt = wv + 4 * q;
                const int m = rt * 16 + 4 * j + kr;
                if (rt < RT && m < ncl) {
                    y[2 * m] = cre[q][j];
                    y[2 * m + 1] = cim[q][j];
                }
            }
    }
}

// the per-box sums of the grouped rows, in list-2 order: a 64-lane block per
// target-or-target-parent box of a grouped level (levmask bit), lanes over
// the output coefficients
__global__ void __launch_bounds__(64)
k_m2l_cls_sum(KD kd, TD t, int nttp, const int *ttp, const int *starts, const int *pos,
              const int *yoff, unsigned levmask, const double *Y, double *local)
{
    const int i = blockIdx.x;
    if (i >= nttp) return;
    const int P = ttp[i];
    const int lv = t.level[P];
    if (!((levmask >> lv) & 1u)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int p = kd.order[lv], ncl = kd.ncl[lv];
    double *o = local + ecoef_off(kd, P, lv);
    for (int m = threadIdx.x; m < ncl; m += 64) {
        cplx acc = {0.0, 0.0};
        for (int e = e0; e < e1; ++e) acc += ldc(Y + 2 * ((long)yoff[pos[e]] + m));
        double *q = o + 2 * ((long)(m - p) * kd.nr + kd.r0);
        q[0] += acc.x;
        q[1] += acc.y;
    }
}

// The same product for the coarse levels (48 < ncl <= 256): the block walks
// the entries together, stages each entry's row window in LDS once, and wave
// w computes the row tiles w, w + 4, ... (at most M2LM_RTW each) over all
// entries, so no partial sums need combining.
__global__ void __launch_bounds__(256)
k_m2l_mfma_rows(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
                const double *mpole, double *local)
{
    __shared__ double win[2 * M2LM_WIN2];
    int lv, i, unused;
    if (!flat_locate(fl, blockIdx.x, lv, i, unused)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int p = kd.order[lv];
    const int ncl = kd.ncl[lv];
    const int RT = (ncl + 15) / 16, KS = (ncl + 3) / 4;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int P = ttp[i];
    const int Ix = t.coords[P], Iy = t.coords[(size_t)t.nboxes + P];
    const int L = 2 * p;
    const int rc = kd.rc, r0 = kd.r0;
    const int col = lane & 15, kr = lane >> 4;
    const int dlo = -(RT * 16 - 1), nwin = RT * 16 + KS * 4 - 1;
    m2l_d4 cre[M2LM_RTW], cim[M2LM_RTW];
#pragma unroll
    for (int q = 0; q < M2LM_RTW; ++q) cre[q] = cim[q] = m2l_d4{0, 0, 0, 0};
    for (int e = e0; e < e1; ++e) {
        const int A = lists[e];
        const int dx = Ix - t.coords[A], dy = Iy - t.coords[(size_t)t.nboxes + A];
        const int cls = (dx + kd.m2l_w) * (2 * kd.m2l_w + 1) + (dy + kd.m2l_w);
        const double *row = kd.tab_m2l + 2 * (kd.m2l_off[lv] + (long)cls * (2 * L + 1) + L);
        const double *a = mpole + ecoef_off(kd, A, lv);
        __syncthreads();
        for (int w = threadIdx.x; w < nwin; w += 256) {
            const int d = w + dlo;
            const bool in = d >= -L && d <= L;
            win[2 * w] = in ? row[2 * d] : 0.0;
            win[2 * w + 1] = in ? row[2 * d + 1] : 0.0;
        }
        __syncthreads();
        for (int ks = 0; ks < KS; ++ks) {
            const int n = ks * 4 + kr;
            double bre = 0.0, bim = 0.0;
            if (n < ncl && col < rc) {
                const cplx b = ldcn(kd, a, n - p, r0 + col);
                bre = b.x;
                bim = b.y;
            }
#pragma unroll
            for (int q = 0; q < M2LM_RTW; ++q) {
                const int rt = wv + 4 * q;
                if (rt < RT) {
                    const int w = n - (rt * 16 + col) - dlo;
                    const double are = win[2 * w], aim = win[2 * w + 1];
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[q], 0, 0, 0);
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[q], 0, 0, 0);
                }
            }
        }
    }
    double *o = local + ecoef_off(kd, P, lv);
#pragma unroll
    for (int q = 0; q < M2LM_RTW; ++q) {
        const int rt = wv + 4 * q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = rt * 16 + 4 * j + kr;
            if (rt < RT && m < ncl && col < rc) {
                double *dst = o + 2 * ((long)(m - p) * kd.nr + r0 + col);
                dst[0] += cre[q][j];
                dst[1] += cim[q][j];
            }
        }
    }
}

// Helmholtz M2M (MODE 0: parents at level lv from their children) and L2L
// (MODE 1: boxes at level lv from their parent) on the matrix cores, up to 16
// right-hand sides: out[m][r] += sum_n row[n - m] in[n][r] (L2L: times
// (-1)^(n-m)), the rows from the shift table, n over the input order, m over
// the output order.  Same tiling as k_m2l_mfma_rows: the block walks the
// inputs (the nonempty children, or the parent), stages each row window in
// LDS once, wave w computes the row tiles w, w + 4, ...
template <int MODE>
__global__ void __launch_bounds__(256)
k_shift_mfma(KD kd, TD t, int lv, int i0, const int *boxes, double *expn)
{
    __shared__ double win[2 * M2LM_WIN2];
    const int Bx = boxes[i0 + blockIdx.x];
    const int lin = MODE == 0 ? lv + 1 : lv - 1;
    const int pin = kd.order[lin], pout = kd.order[lv];
    const int nin = 2 * pin + 1, nout = 2 * pout + 1;
    const int RT = (nout + 15) / 16, KS = (nin + 3) / 4;
    const int L = pin + pout;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rc = kd.rc, r0 = kd.r0;
    const int col = lane & 15, kr = lane >> 4;
    // window index w <-> d = n - m (0-based n, m): d in [-(16 RT - 1), 4 KS - 1];
    // the table offset of (n, m) is (n - pin) - (m - pout) = d + pout - pin
    const int dlo = -(RT * 16 - 1), nwin = RT * 16 + KS * 4 - 1;
    m2l_d4 cre[M2LM_RTW], cim[M2LM_RTW];
#pragma unroll
    for (int q = 0; q < M2LM_RTW; ++q) cre[q] = cim[q] = m2l_d4{0, 0, 0, 0};
    const int nsrc = MODE == 0 ? 4 : 1;
    for (int c = 0; c < nsrc; ++c) {
        int IN, cls;
        if (MODE == 0) {
            IN = t.child[(size_t)c * t.nboxes + Bx];
            if (IN < 0 || t.scumul[IN] == 0) continue;  // uniform over the block
            cls = c;
        } else {
            IN = t.parent[Bx];
            if (IN < 0) return;
            cls = (t.coords[Bx] & 1) | ((t.coords[(size_t)t.nboxes + Bx] & 1) << 1);
        }
        const double *row =
            kd.tab_mm + 2 * (kd.mm_off[MODE == 0 ? lin : lv] + (long)cls * (2 * L + 1) + L);
        const double *a = expn + ecoef_off(kd, IN, lin);
        __syncthreads();
        for (int w = threadIdx.x; w < nwin; w += 256) {
            const int d = w + dlo + pout - pin;
            const bool in = d >= -L && d <= L;
            double sg = 1.0;
            if (MODE == 1 && (d & 1)) sg = -1.0;
            win[2 * w] = in ? sg * row[2 * d] : 0.0;
            win[2 * w + 1] = in ? sg * row[2 * d + 1] : 0.0;
        }
        __syncthreads();
        for (int ks = 0; ks < KS; ++ks) {
            const int n = ks * 4 + kr;
            double bre = 0.0, bim = 0.0;
            if (n < nin && col < rc) {
                const cplx b = ldcn(kd, a, n - pin, r0 + col);
                bre = b.x;
                bim = b.y;
            }
#pragma unroll
            for (int q = 0; q < M2LM_RTW; ++q) {
                const int rt = wv + 4 * q;
                if (rt < RT) {
                    const int w = n - (rt * 16 + col) - dlo;
                    const double are = win[2 * w], aim = win[2 * w + 1];
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[q], 0, 0, 0);
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[q], 0, 0, 0);
                }
            }
        }
    }
    double *o = expn + ecoef_off(kd, Bx, lv);
#pragma unroll
    for (int q = 0; q < M2LM_RTW; ++q) {
        const int rt = wv + 4 * q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = rt * 16 + 4 * j + kr;
            if (rt < RT && m < nout && col < rc) {
                double *dst = o + 2 * ((long)(m - pout) * kd.nr + r0 + col);
                dst[0] += cre[q][j];
                dst[1] += cim[q][j];
            }
        }
    }
}

template <int NRC>
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
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    if (!helm) {
        lap_l2l<NRC>(kd, B, a, p1, m, boxc(t, P) - boxc(t, Q), kd.rscale[lv - 1], kd.rscale[lv],
                     acc);
    } else {
        // uhat_j(c_P - c_Q) = (-1)^j uhat_j(c_Q - c_P): the M2M row with sign
        const int c = (t.coords[P] & 1) | ((t.coords[(size_t)t.nboxes + P] & 1) << 1);
        const int L = p1 + p2;
        const double *row = kd.tab_mm + 2 * (kd.mm_off[lv] + (long)c * (2 * L + 1) + L);
        helm_row_dot<NRC>(kd, a, p1, m, row, true, acc);
    }
    store_add<NRC>(kd, local + ecoef_off(kd, P, lv) + 2L * m * kd.nr, acc);
}

// -------------------------------------------------------------------------
// Helmholtz M2M / L2L through LDS.  A block owns one output box and its
// first 2 p_out + 1 threads one output coefficient m each.  Source boxes are
// taken one at a time: the coefficient vector (2 p_in + 1 coefficients of
// the chunk's right-hand sides) and the translation row (2 L + 1, centred)
// are staged into LDS with coalesced loads, then
//     out_m += sum_n a_n row[n - m]      (row sign-flipped at odd n - m: L2L)
// reads only LDS (the a_n are broadcast, the rows contiguous across m).  The
// per-level launches at coarse levels have few boxes and long rows, where
// the plain one-thread-per-coefficient kernels wait on serial global loads.
// (M2L takes a block per box with its list-2 entries split across thread
// groups, k_m2l_blk; staging the entries' coefficient vectors in LDS and
// register blocking over outputs both measured slower.)
// -------------------------------------------------------------------------
constexpr int HT_MAX_THREADS = 256;
constexpr long HT_LDS_BUDGET = 32768;  // bytes of LDS per block for staged sources

// doubles of LDS one staged source box takes
__host__ __device__ __forceinline__ int ht_entry_doubles(int nrc, int p_in, int L)
{
    return 2 * nrc * (2 * p_in + 1) + 2 * (2 * L + 1);
}

// sa[n + p_in][q] <- coefficient n, right-hand side r0 + q; ag points at n = -p_in
template <int NRC>
__device__ __forceinline__ void ht_stage_coeffs(const KD &kd, double *sa, const double *ag,
                                                int na, int tid, int nth)
{
    const int rc = kd.rc;
    for (int i = tid; i < na * rc; i += nth) {
        const int n = i / rc, q = i - n * rc;
        const double *g = ag + 2 * ((long)n * kd.nr + kd.r0 + q);
        sa[2 * (n * NRC + q)] = g[0];
        sa[2 * (n * NRC + q) + 1] = g[1];
    }
}

__device__ __forceinline__ void ht_stage_row(double *sr, const double *rg, int nrow, int tid,
                                             int nth)
{
    for (int i = tid; i < 2 * nrow; i += nth) sr[i] = rg[i];
}

// acc[q] += sum_{n=-p1}^{p1} sa[n + p1][q] row[n - m]; row centred.  One
// right-hand side: four independent partial sums (n mod 4) for latency.
template <int NRC>
__device__ __forceinline__ void ht_dot(int rc, const double *sa, int p1, int m,
                                       const double *row, bool flip, cplx *acc)
{
    if (NRC == 1) {
        cplx s4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        int n = -p1;
        for (; n + 3 <= p1; n += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int d = n + u - m;
                cplx w = {row[2 * d], row[2 * d + 1]};
                if (flip && (d & 1)) w = -w;
                s4[u] += cplx{sa[2 * (n + u + p1)], sa[2 * (n + u + p1) + 1]} * w;
            }
        }
        for (; n <= p1; ++n) {
            const int d = n - m;
            cplx w = {row[2 * d], row[2 * d + 1]};
            if (flip && (d & 1)) w = -w;
            s4[0] += cplx{sa[2 * (n + p1)], sa[2 * (n + p1) + 1]} * w;
        }
        acc[0] += (s4[0] + s4[1]) + (s4[2] + s4[3]);
        return;
    }
    for (int n = -p1; n <= p1; ++n) {
        const int d = n - m;
        cplx w = {row[2 * d], row[2 * d + 1]};
        if (flip && (d & 1)) w = -w;
        const double *an = sa + 2 * NRC * (n + p1);
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc) acc[q] += cplx{an[2 * q], an[2 * q + 1]} * w;
    }
}

// M2M: one block per source parent box at level lvp; its children are staged
// NS at a time (all four when they fit), so one round of global loads serves
// several children
// NG thread groups (blockDim >= NG (2 p2 + 1)) split the staged children;
// their partial sums are added in LDS in group order (NG = 1: one group).
template <int NRC>
__global__ void k_m2m_h(KD kd, TD t, int lvp, int i0, int NS, int NG, const int *spb,
                        double *mpole)
{
    extern __shared__ double lds[];
    const int P = spb[i0 + blockIdx.x];
    const int lvc = lvp + 1, p1 = kd.order[lvc], p2 = kd.order[lvp], L = p1 + p2;
    const int na = 2 * p1 + 1, nrow = 2 * L + 1;
    const int per = ht_entry_doubles(NRC, p1, L);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nout = 2 * p2 + 1;
    const int grp = tid / nout;
    const bool act = grp < NG;
    const int m = tid % nout - p2;
    int kids[4], nk = 0;
    for (int c = 0; c < 4; ++c) {
        const int C = t.child[(size_t)c * t.nboxes + P];
        if (C >= 0 && t.scumul[C] > 0) kids[nk++] = c;
    }
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int k0 = 0; k0 < nk; k0 += NS) {
        const int kn = min(NS, nk - k0);
        __syncthreads();
        for (int k = 0; k < kn; ++k) {
            const int c = kids[k0 + k];
            const int C = t.child[(size_t)c * t.nboxes + P];
            double *sa = lds + (long)k * per, *sr = sa + 2 * NRC * na;
            ht_stage_coeffs<NRC>(kd, sa, mpole + ecoef_off(kd, C, lvc) - 2L * p1 * kd.nr, na,
                                 tid, nth);
            ht_stage_row(sr, kd.tab_mm + 2 * (kd.mm_off[lvc] + (long)c * nrow), nrow, tid, nth);
        }
        __syncthreads();
        if (act)
            for (int k = grp; k < kn; k += NG) {
                const double *sa = lds + (long)k * per;
                ht_dot<NRC>(kd.rc, sa, p1, m, sa + 2 * NRC * na + 2 * L, false, acc);
            }
    }
    if (NG > 1) {
        // the groups' partial sums through LDS (after the last chunk's reads)
        __syncthreads();
        cplx *red = reinterpret_cast<cplx *>(lds);
        if (act && grp > 0)
#pragma unroll
            for (int q = 0; q < NRC; ++q) red[((grp - 1) * nout + (m + p2)) * NRC + q] = acc[q];
        __syncthreads();
        if (act && grp == 0)
            for (int g = 1; g < NG; ++g)
#pragma unroll
                for (int q = 0; q < NRC; ++q) acc[q] += red[((g - 1) * nout + (m + p2)) * NRC + q];
    }
    if (act && grp == 0) store_add<NRC>(kd, mpole + ecoef_off(kd, P, lvp) + 2L * m * kd.nr, acc);
}

// L2L: one block per target-or-target-parent box at level lv
template <int NRC>
__global__ void k_l2l_h(KD kd, TD t, int lv, int i0, const int *ttp, double *local)
{
    extern __shared__ double lds[];
    const int P = ttp[i0 + blockIdx.x];
    const int Q = t.parent[P];
    if (Q < 0) return;
    const int p2 = kd.order[lv], p1 = kd.order[lv - 1], L = p1 + p2;
    const int na = 2 * p1 + 1, nrow = 2 * L + 1;
    const int tid = threadIdx.x, nth = blockDim.x;
    double *sa = lds, *sr = lds + 2 * NRC * na;
    // uhat_j(c_P - c_Q) = (-1)^j uhat_j(c_Q - c_P): the M2M row of P's child slot
    const int c = (t.coords[P] & 1) | ((t.coords[(size_t)t.nboxes + P] & 1) << 1);
    ht_stage_coeffs<NRC>(kd, sa, local + ecoef_off(kd, Q, lv - 1) - 2L * p1 * kd.nr, na, tid,
                         nth);
    ht_stage_row(sr, kd.tab_mm + 2 * (kd.mm_off[lv] + (long)c * nrow), nrow, tid, nth);
    __syncthreads();
    if (tid < 2 * p2 + 1) {
        const int m = tid - p2;
        cplx acc[NRC];
#pragma unroll
        for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
        ht_dot<NRC>(kd.rc, sa, p1, m, sr + 2 * L, true, acc);
        store_add<NRC>(kd, local + ecoef_off(kd, P, lv) + 2L * m * kd.nr, acc);
    }
}

// -------------------------------------------------------------------------
// evaluation at points of a chunk of expansions; a: ecoef_off-style pointer
// (coefficient n, right-hand side r at a[2 (n nr + r)]); out[q] receives the
// potential of right-hand side r0 + q
// -------------------------------------------------------------------------
// DER: the derivative along the complex direction d instead.  Laplace: the
// potential is -Re F / (2 pi) with F analytic, so the derivative is
// -Re(F' d) / (2 pi).  Helmholtz: with D+- = d/dx +- i d/dy,
// D+ [Z_n e^{in th}] = -k Z_{n+1} e^{i(n+1) th} and
// D- [Z_n e^{in th}] = k Z_{n-1} e^{i(n-1) th}, so the derivative of
// sum_n a_n Z_n e^{in th} along d is sum_m b_m Z_m e^{im th} with
// b_m = (k/2) (d a_{m+1} - conj(d) a_{m-1}), |m| <= p + 1.
template <int NRC, bool DER = false>
__device__ __forceinline__ cplx dcoef(const KD &kd, const double *a, int p, int m, int r,
                                      cplx dir)
{
    if (!DER) return ldcn(kd, a, m, r);
    const cplx up = (m + 1 >= -p && m + 1 <= p) ? ldcn(kd, a, m + 1, r) : cplx{0, 0};
    const cplx dn = (m - 1 >= -p && m - 1 <= p) ? ldcn(kd, a, m - 1, r) : cplx{0, 0};
    return (0.5 * kd.k) * (dir * up - conj(dir) * dn);
}

template <int NRC, bool DER = false>
__device__ void eval_local_n(const KD &kd, const double *a, int p, cplx c, double R, cplx tp,
                             cplx *out, cplx dir = {0.0, 0.0})
{
    const int r0 = kd.r0, rc = kd.rc;
    if (kd.kernel == QBX_LAPLACE2D) {
        const cplx x = (1.0 / R) * (tp - c);
        cplx phi[NRC];
#pragma unroll
        for (int q = 0; q < NRC; ++q) phi[q] = {0, 0};
        if (!DER) {
            for (int l = p; l >= 0; --l)
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) phi[q] = phi[q] * x + ldcn(kd, a, l, r0 + q);
#pragma unroll
            for (int q = 0; q < NRC; ++q) out[q] = {-phi[q].x * INV2PI, 0.0};
        } else {
            // F' = sum_l l a_l x^(l-1) / R
            for (int l = p; l >= 1; --l)
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) phi[q] = phi[q] * x + (double)l * ldcn(kd, a, l, r0 + q);
            const cplx dr = (1.0 / R) * dir;
#pragma unroll
            for (int q = 0; q < NRC; ++q) out[q] = {-(phi[q] * dr).x * INV2PI, 0.0};
        }
        return;
    }
    const int pt = DER ? p + 1 : p;
    const cplx d = tp - c;
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    double norm = 1.0;
    if (x < 1e-300) {
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc) acc[q] = dcoef<NRC, DER>(kd, a, p, 0, r0 + q, dir);
    } else {
        const cplx e1 = {d.x / rho, d.y / rho};
        const cplx em1 = conj(e1);
        cplx e = cpow_int(e1, pt);
        norm = miller_j(
            x, pt,
            [&](int n, double f) {
                const cplx fe = f * e;
                cplx fc = f * conj(e);
                if (n & 1) fc = -fc;
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) {
                        acc[q] += dcoef<NRC, DER>(kd, a, p, n, r0 + q, dir) * fe;
                        if (n > 0) acc[q] += dcoef<NRC, DER>(kd, a, p, -n, r0 + q, dir) * fc;
                    }
                e = e * em1;
            },
            [&](double s) {
#pragma unroll
                for (int q = 0; q < NRC; ++q) acc[q] = s * acc[q];
            });
    }
#pragma unroll
    for (int q = 0; q < NRC; ++q) {
        const cplx phi = norm * acc[q];
        out[q] = {-0.25 * phi.y, 0.25 * phi.x};
    }
}

// DER: the derivative along d (see eval_local_n): Laplace
// F' = (a_0 - sum_k k a_k (R/w)^k) / w; Helmholtz the b_m of the H_m series.
template <int NRC, bool DER = false>
__device__ void eval_mpole_add(const KD &kd, const double *a, int p, cplx c, double R, cplx tp,
                               cplx *out, cplx dir = {0.0, 0.0})
{
    const int r0 = kd.r0, rc = kd.rc;
    const cplx w = tp - c;
    if (kd.kernel == QBX_LAPLACE2D) {
        const cplx x = R * cinv(w);
        cplx phi[NRC];
#pragma unroll
        for (int q = 0; q < NRC; ++q) phi[q] = {0, 0};
        if (!DER) {
            const cplx lw = clogz(w);
            for (int k = p; k >= 1; --k)
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) phi[q] = (phi[q] + ldcn(kd, a, k, r0 + q)) * x;
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const cplx v = phi[q] + ldcn(kd, a, 0, r0 + q) * lw;
                    out[q] += cplx{-v.x * INV2PI, 0.0};
                }
        } else {
            for (int k = p; k >= 1; --k)
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) phi[q] = (phi[q] + (double)k * ldcn(kd, a, k, r0 + q)) * x;
            const cplx dw = dir * cinv(w);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const cplx v = (ldcn(kd, a, 0, r0 + q) - phi[q]) * dw;
                    out[q] += cplx{-v.x * INV2PI, 0.0};
                }
        }
        return;
    }
    const int pt = DER ? p + 1 : p;
    const double iw = qbx_rsqrt(abs2(w));
    const double rho = abs2(w) * iw;
    const double x = kd.k * rho;
    const cplx e1 = {w.x * iw, w.y * iw};
    cplx e = {1, 0};
    cplx hm1, hm;
    hankel01_call(x, hm1, hm);
    const double tox = 2.0 * iw * kd.ik;
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        acc[q] = q < rc ? dcoef<NRC, DER>(kd, a, p, 0, r0 + q, dir) * hm1 : cplx{0, 0};
    for (int n = 1; n <= pt; ++n) {
        cplx h;
        if (n == 1) {
            h = hm;
        } else {
            h = ((n - 1) * tox) * hm - hm1;
            hm1 = hm;
            hm = h;
        }
        e = e * e1;
        const cplx he = h * e;
        cplx hc = h * conj(e);
        if (n & 1) hc = -hc;
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc)
                acc[q] += dcoef<NRC, DER>(kd, a, p, n, r0 + q, dir) * he +
                          dcoef<NRC, DER>(kd, a, p, -n, r0 + q, dir) * hc;
    }
#pragma unroll
    for (int q = 0; q < NRC; ++q) out[q] += cplx{-0.25 * acc[q].y, 0.25 * acc[q].x};
}

// P2P over balanced work units (one 64-lane wavefront each, see unit_plan):
// the unit's TG targets take G lanes apiece, the box's list-1 sources are
// staged through LDS in chunks of CH (positions, normals, strengths, read
// coalesced), each target's lanes split a chunk, and the G partial sums are
// reduced by shuffles in a fixed order.
template <int NRC>
struct P2PTile {
    static constexpr int CH = NRC == 1 ? 128 : 64;
};

template <int NRC, bool DER, int KN, int DP>
__global__ void __launch_bounds__(64) k_p2p(KD kd, TD t, const int *ubox, const int *ut0,
                                            const int *box_g, const int *tboxes,
                                            const int *starts, const int *lists,
                                            const int *fstart, const int *fcount,
                                            const int *unfilt, const double *tgt, int nt,
                                            const double *src, int ns, const double *str,
                                            const double *nrm, double *pot)
{
    constexpr int CH = P2PTile<NRC>::CH;
    __shared__ double sx[CH], sy[CH], snx[CH], sny[CH];
    __shared__ double sq[CH * 2 * NRC];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int B = tboxes[it];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int tl = ut0[u] + tsub;
    const bool active = tl < fcount[B];
    const bool dip = DP >= 0 ? DP != 0 : kd.dip != 0;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    int ti = 0;
    cplx tp = {0, 0}, dir = {0, 0};
    if (active) {
        ti = unfilt[fstart[B] + tl];
        tp = ldpt(tgt, nt, ti);
        if (DER) dir = ldir(kd, ti);
    }
    hankel_table_fill(htab, lane, 64);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};

    auto flush = [&](int n) {
        __syncthreads();
        if (active) {
            for (int j = sub; j < n; j += G) {
                // an offset the compiler cannot see through keeps the series
                // coefficients as LDS reads inside the loop instead of
                // hoisting all of them into registers
                int toff = 0;
                asm volatile("" : "+s"(toff));
                const cplx nu = dip ? cplx{snx[j], sny[j]} : cplx{0, 0};
                const cplx g = p2p_value_impl<KN, DP, DER, false>(
                    kd, cplx{sx[j], sy[j]}, nu, tp, htab + toff, 0, dir);
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) acc[q] += cplx{sq[2 * (j * NRC + q)], sq[2 * (j * NRC + q) + 1]} * g;
            }
        }
        __syncthreads();
    };

    int fill = 0;
    const int e1 = starts[it + 1];
    for (int e = starts[it]; e < e1; ++e) {
        const int Sb = lists[e];
        int s = t.sstart[Sb];
        const int s1 = s + t.scount[Sb];
        while (s < s1) {
            const int take = min(s1 - s, CH - fill);
            for (int j = lane; j < take; j += 64) {
                const int si = s + j, d = fill + j;
                sx[d] = src[si];
                sy[d] = src[(size_t)ns + si];
                if (dip) {
                    snx[d] = nrm[si];
                    sny[d] = nrm[(size_t)ns + si];
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) {
                        const double *p = str + 2 * ((size_t)si * nr + r0 + q);
                        sq[2 * (d * NRC + q)] = p[0];
                        sq[2 * (d * NRC + q) + 1] = p[1];
                    }
            }
            fill += take;
            s += take;
            if (fill == CH) {
                flush(CH);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        for (int m = G / 2; m >= 1; m >>= 1) {
            acc[q].x += __shfl_xor(acc[q].x, m);
            acc[q].y += __shfl_xor(acc[q].y, m);
        }
    if (active && sub == 0) store_add<NRC>(kd, pot + 2 * (size_t)ti * nr, acc);
}

// M2P over balanced work units (see unit_plan): the unit's TG targets take
// G lanes apiece that split the box's list-3 boxes; shuffle reduction.
template <int NRC, bool DER = false>
__global__ void __launch_bounds__(64) k_m2p(KD kd, TD t, const int *ubox, const int *ut0,
                                            const int *box_g, const int *tboxes,
                                            const int *starts, const int *lists,
                                            const int *fstart, const int *fcount,
                                            const int *unfilt, const double *tgt, int nt,
                                            const double *mpole, double *pot)
{
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int B = tboxes[it];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int tl = ut0[u] + tsub;
    const bool active = tl < fcount[B];
    int ti = 0;
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    if (active) {
        ti = unfilt[fstart[B] + tl];
        const cplx tp = ldpt(tgt, nt, ti);
        const cplx dir = DER ? ldir(kd, ti) : cplx{0, 0};
        for (int e = starts[it] + sub; e < starts[it + 1]; e += G) {
            const int D = lists[e];
            const int lv = t.level[D];
            eval_mpole_add<NRC, DER>(kd, mpole + ecoef_off(kd, D, lv), kd.order[lv], boxc(t, D),
                                     kd.rscale[lv], tp, acc, dir);
        }
    }
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        for (int m = G / 2; m >= 1; m >>= 1) {
            acc[q].x += __shfl_xor(acc[q].x, m);
            acc[q].y += __shfl_xor(acc[q].y, m);
        }
    if (active && sub == 0) store_add<NRC>(kd, pot + 2 * (size_t)ti * kd.nr, acc);
}

// L2P: one 64-lane block per target box, one lane per non-QBX target
template <int MODE, int NRC, bool DER = false>
__global__ void __launch_bounds__(64) k_eval_targets(KD kd, TD t, int ntb, const int *tboxes,
                                                     const int *starts, const int *lists,
                                                     const int *fstart, const int *fcount,
                                                     const int *unfilt, const double *tgt,
                                                     int nt, const double *expn, double *pot)
{
    static_assert(MODE == 2, "k_eval_targets evaluates box locals only");
    const int it = blockIdx.x;
    if (it >= ntb) return;
    const int B = tboxes[it];
    const int f0 = fstart[B], nf = fcount[B];
    const int lv = t.level[B];
    for (int f = f0 + threadIdx.x; f < f0 + nf; f += 64) {
        const int ti = unfilt[f];
        cplx v[NRC];
        eval_local_n<NRC, DER>(kd, expn + ecoef_off(kd, B, lv), kd.order[lv], boxc(t, B),
                               kd.rscale[lv], ldpt(tgt, nt, ti), v,
                               DER ? ldir(kd, ti) : cplx{0, 0});
        store_add<NRC>(kd, pot + 2 * (size_t)ti * kd.nr, v);
    }
}

// -------------------------------------------------------------------------
// QBX stages
// -------------------------------------------------------------------------
template <int PQ, int NRC>
struct QAcc {
    cplx v[NRC][2 * PQ + 1];
};

template <int PQ, int NRC>
__device__ __forceinline__ void qacc_zero(QAcc<PQ, NRC> &A)
{
#pragma unroll
    for (int q = 0; q < NRC; ++q)
#pragma unroll
        for (int i = 0; i < 2 * PQ + 1; ++i) A.v[q][i] = {0, 0};
}

// A holds Laplace coefficients l at v[.][l], Helmholtz n at v[.][n + PQ]
template <int PQ, int NRC>
__device__ void qbx_store(const KD &kd, double *qexp, int c, const QAcc<PQ, NRC> &A,
                          int accumulate)
{
    const int nq = kd.kernel == QBX_LAPLACE2D ? PQ + 1 : 2 * PQ + 1;
    double *o = qexp + 2 * (size_t)c * kd.ncq * kd.nr;
#pragma unroll
    for (int i = 0; i < 2 * PQ + 1; ++i) {
        if (i < nq) {
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) {
                    double *p = o + 2 * ((long)i * kd.nr + kd.r0 + q);
                    cplx v = A.v[q][i];
                    if (accumulate) v += ldc(p);
                    stc(p, v);
                }
        }
    }
}

// P2QBXL contribution of one source (w = source - center, normal nu,
// strengths qs) to the order-PQ expansion A about a center of radius R.  No
// atan2/sincos: the phase e^{-i theta} is conj(w)/|w|.  tab: the Hankel
// series coefficients (hankel01_wave).
// KN, DP: the kernel and source kind fixed at compile time (-1: read from kd)
template <int PQ, int NRC, int KN = -1, int DP = -1>
__device__ __forceinline__ void p2qbxl_pair(const KD &kd, cplx w, cplx nu, const cplx *qs,
                                            double R, const double *tab, QAcc<PQ, NRC> &A)
{
    const int kern = KN >= 0 ? KN : kd.kernel;
    const bool dip = DP >= 0 ? DP != 0 : kd.dip != 0;
    if (kern == QBX_LAPLACE2D) {
        const cplx iw = cinv(w);
        const cplx tt = R * iw;
        cplx pw = {1, 0};
        if (!dip) {
            const cplx lg = clogz(-w);
#pragma unroll
            for (int q = 0; q < NRC; ++q) A.v[q][0] += qs[q] * lg;
#pragma unroll
            for (int l = 1; l <= PQ; ++l) {
                pw = pw * tt;
                const cplx b = (-1.0 / l) * pw;
#pragma unroll
                for (int q = 0; q < NRC; ++q) A.v[q][l] += qs[q] * b;
            }
        } else {
            const cplx f = nu * iw;
#pragma unroll
            for (int q = 0; q < NRC; ++q) A.v[q][0] += qs[q] * f;
#pragma unroll
            for (int l = 1; l <= PQ; ++l) {
                pw = pw * tt;
                const cplx b = f * pw;
#pragma unroll
                for (int q = 0; q < NRC; ++q) A.v[q][l] += qs[q] * b;
            }
        }
        return;
    }
    const double ir = qbx_rsqrt(abs2(w));
    const double r = abs2(w) * ir;
    const double x = kd.k * r;
    const double ix = ir * kd.ik;
    const cplx em = {w.x * ir, -w.y * ir};  // e^{-i theta}
    cplx H[PQ + 2];
    hankel01_wave<true, true>(x, ix, tab, H[0], H[1]);
    const double tox = 2.0 * ix;
#pragma unroll
    for (int m = 2; m <= PQ + 1; ++m) H[m] = ((m - 1) * tox) * H[m - 1] - H[m - 2];
    // Gv[m + PQ + 1] = H_m e^{-i m theta}, m in [-(PQ+1), PQ+1]
    cplx Gv[2 * PQ + 3];
    cplx e = {1, 0};
#pragma unroll
    for (int m = 0; m <= PQ + 1; ++m) {
        Gv[PQ + 1 + m] = H[m] * e;
        const cplx gm = H[m] * conj(e);
        Gv[PQ + 1 - m] = (m & 1) ? -gm : gm;
        e = e * em;
    }
    if (!dip) {
#pragma unroll
        for (int n = -PQ; n <= PQ; ++n)
#pragma unroll
            for (int q = 0; q < NRC; ++q) A.v[q][n + PQ] += qs[q] * Gv[n + PQ + 1];
    } else {
        const double hk = 0.5 * kd.k;
        const cplx a1 = hk * conj(nu), a2 = hk * nu;
#pragma unroll
        for (int n = -PQ; n <= PQ; ++n) {
            const cplx b = a1 * Gv[n + PQ] - a2 * Gv[n + PQ + 2];
#pragma unroll
            for (int q = 0; q < NRC; ++q) A.v[q][n + PQ] += qs[q] * b;
        }
    }
}

// reduce A over the G lanes of a group (consecutive lanes, G a power of two)
template <int PQ, int NRC>
__device__ __forceinline__ void qacc_group_reduce(QAcc<PQ, NRC> &A, int G)
{
#pragma unroll
    for (int q = 0; q < NRC; ++q)
#pragma unroll
        for (int i = 0; i < 2 * PQ + 1; ++i)
            for (int m = G / 2; m >= 1; m >>= 1) {
                A.v[q][i].x += __shfl_xor(A.v[q][i].x, m);
                A.v[q][i].y += __shfl_xor(A.v[q][i].y, m);
            }
}

// P2QBXL over balanced work units (see unit_plan): the unit's TG global
// centers of one target box take G lanes apiece; the box's list-1 sources are
// staged through LDS in chunks (positions, normals, strengths) and each
// center's lanes split a chunk; the expansions are reduced by shuffles.
template <int NRC>
struct P2QTile {
    static constexpr int CH = NRC == 1 ? 256 : 64;
};

template <int PQ, int NRC, int KN = -1, int DP = -1>
__global__ void __launch_bounds__(64) k_p2qbxl(KD kd, TD t, const int *ubox, const int *ut0,
                                               const int *box_g, const int *gcs,
                                               const int *gcl, const double *ctr,
                                               const double *rad, int nc, const int *l1s,
                                               const int *l1l, const double *src, int ns,
                                               const double *str, const double *nrm,
                                               double *qexp, int accumulate)
{
    constexpr int CH = P2QTile<NRC>::CH;
    __shared__ double sx[CH], sy[CH], snx[CH], sny[CH];
    __shared__ double sq[CH * 2 * NRC];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int cl = ut0[u] + tsub;
    const bool active = cl < gcs[it + 1] - gcs[it];
    const bool dip = DP >= 0 ? DP != 0 : kd.dip != 0;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    int c = 0;
    cplx cc = {0, 0};
    double R = 1.0;
    if (active) {
        c = gcl[gcs[it] + cl];
        cc = ldpt(ctr, nc, c);
        R = rad[c];
    }
    hankel_table_fill(htab, lane, 64);
    QAcc<PQ, NRC> A;
    qacc_zero<PQ, NRC>(A);

    auto flush = [&](int n) {
        __syncthreads();
        if (active) {
            for (int j = sub; j < n; j += G) {
                // an offset the compiler cannot see through keeps the series
                // coefficients as LDS reads inside the loop
                int toff = 0;
                asm volatile("" : "+s"(toff));
                cplx qs[NRC];
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    qs[q] = q < rc ? cplx{sq[2 * (j * NRC + q)], sq[2 * (j * NRC + q) + 1]}
                                   : cplx{0, 0};
                const cplx nu = dip ? cplx{snx[j], sny[j]} : cplx{0, 0};
                p2qbxl_pair<PQ, NRC, KN, DP>(kd, cplx{sx[j], sy[j]} - cc, nu, qs, R,
                                             htab + toff, A);
            }
        }
        __syncthreads();
    };

    int fill = 0;
    const int e1 = l1s[it + 1];
    for (int e = l1s[it]; e < e1; ++e) {
        const int Sb = l1l[e];
        int s = t.sstart[Sb];
        const int s1 = s + t.scount[Sb];
        while (s < s1) {
            const int take = min(s1 - s, CH - fill);
            for (int j = lane; j < take; j += 64) {
                const int si = s + j, d = fill + j;
                sx[d] = src[si];
                sy[d] = src[(size_t)ns + si];
                if (dip) {
                    snx[d] = nrm[si];
                    sny[d] = nrm[(size_t)ns + si];
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) {
                        const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                        sq[2 * (d * NRC + q)] = pq[0];
                        sq[2 * (d * NRC + q) + 1] = pq[1];
                    }
            }
            fill += take;
            s += take;
            if (fill == CH) {
                flush(CH);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
    qacc_group_reduce<PQ, NRC>(A, G);
    if (active && sub == 0) qbx_store<PQ, NRC>(kd, qexp, c, A, accumulate);
}

// Direct QBX (no FMM; pytential/qbx/__init__.py:866-1068 with
// qbx/direct.py:41-154): every global QBX center's expansion from ALL
// sources, and every non-QBX target's potential summed over ALL sources.
// The sources pass through LDS in chunks of DQ_CH shared by the block: 4
// centers x 16 lanes (P2QBXL) or 64 targets x 1 lane (P2P) per block.
constexpr int DQ_CH = 256;

template <int PQ, int NRC>
__global__ void __launch_bounds__(64) k_direct_p2qbxl(KD kd, int ng, const int *gcent,
                                                      const double *ctr, const double *rad,
                                                      int nc, const double *src, int ns,
                                                      const double *str, const double *nrm,
                                                      double *qexp)
{
    __shared__ double sx[DQ_CH], sy[DQ_CH], snx[DQ_CH], sny[DQ_CH];
    __shared__ double sq[DQ_CH * 2 * NRC];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    const int lane = threadIdx.x;
    const int g = blockIdx.x * 4 + lane / 16, sub = lane % 16;
    const bool active = g < ng;
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    int c = 0;
    cplx cc = {0, 0};
    double R = 1.0;
    if (active) {
        c = gcent[g];
        cc = ldpt(ctr, nc, c);
        R = rad[c];
    }
    hankel_table_fill(htab, lane, 64);
    QAcc<PQ, NRC> A;
    qacc_zero<PQ, NRC>(A);
    for (int b = 0; b < ns; b += DQ_CH) {
        const int nb = min(DQ_CH, ns - b);
        __syncthreads();
        for (int j = lane; j < nb; j += 64) {
            const int si = b + j;
            sx[j] = src[si];
            sy[j] = src[(size_t)ns + si];
            if (dip) {
                snx[j] = nrm[si];
                sny[j] = nrm[(size_t)ns + si];
            }
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                    sq[2 * (j * NRC + q)] = pq[0];
                    sq[2 * (j * NRC + q) + 1] = pq[1];
                }
        }
        __syncthreads();
        if (active)
            for (int j = sub; j < nb; j += 16) {
                int toff = 0;
                asm volatile("" : "+s"(toff));
                cplx qs[NRC];
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    qs[q] = q < rc ? cplx{sq[2 * (j * NRC + q)], sq[2 * (j * NRC + q) + 1]}
                                   : cplx{0, 0};
                const cplx nu = dip ? cplx{snx[j], sny[j]} : cplx{0, 0};
                p2qbxl_pair<PQ, NRC>(kd, cplx{sx[j], sy[j]} - cc, nu, qs, R, htab + toff, A);
            }
    }
#pragma unroll
    for (int q = 0; q < NRC; ++q)
#pragma unroll
        for (int i = 0; i < 2 * PQ + 1; ++i)
#pragma unroll
            for (int m = 8; m >= 1; m >>= 1) {
                A.v[q][i].x += __shfl_xor(A.v[q][i].x, m, 16);
                A.v[q][i].y += __shfl_xor(A.v[q][i].y, m, 16);
            }
    if (active && sub == 0) qbx_store<PQ, NRC>(kd, qexp, c, A, 0);
}

template <int NRC, bool DER = false>
__global__ void __launch_bounds__(64) k_direct_p2p(KD kd, int nf, const int *unfilt,
                                                   const double *tgt, int nt, const double *src,
                                                   int ns, const double *str, const double *nrm,
                                                   double *pot)
{
    __shared__ double sx[DQ_CH], sy[DQ_CH], snx[DQ_CH], sny[DQ_CH];
    __shared__ double sq[DQ_CH * 2 * NRC];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    const int lane = threadIdx.x;
    const int f = blockIdx.x * 64 + lane;
    const bool active = f < nf;
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    int ti = 0;
    cplx tp = {0, 0}, dir = {0, 0};
    if (active) {
        ti = unfilt ? unfilt[f] : f;
        tp = ldpt(tgt, nt, ti);
        if (DER) dir = ldir(kd, ti);
    }
    hankel_table_fill(htab, lane, 64);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int b = 0; b < ns; b += DQ_CH) {
        const int nb = min(DQ_CH, ns - b);
        __syncthreads();
        for (int j = lane; j < nb; j += 64) {
            const int si = b + j;
            sx[j] = src[si];
            sy[j] = src[(size_t)ns + si];
            if (dip) {
                snx[j] = nrm[si];
                sny[j] = nrm[(size_t)ns + si];
            }
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                    sq[2 * (j * NRC + q)] = pq[0];
                    sq[2 * (j * NRC + q) + 1] = pq[1];
                }
        }
        __syncthreads();
        if (active)
            for (int j = 0; j < nb; ++j) {
                int toff = 0;
                asm volatile("" : "+s"(toff));
                const cplx nu = dip ? cplx{snx[j], sny[j]} : cplx{0, 0};
                const cplx g = p2p_value_wave<DER>(kd, cplx{sx[j], sy[j]}, nu, tp, htab + toff,
                                                   dir);
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) acc[q] += cplx{sq[2 * (j * NRC + q)], sq[2 * (j * NRC + q) + 1]} * g;
            }
    }
    if (active) {
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc) stc(pot + 2 * ((size_t)ti * nr + r0 + q), acc[q]);
    }
}

// ---- dense layer-potential matrices -------------------------------------
// Local-expansion basis at the target of a row: Laplace B_l = ((t - c)/R)^l,
// l = 0..PQ; Helmholtz B_{n+PQ} = J_n(k rho) e^{i n theta}, n = -PQ..PQ, with
// t - c = rho e^{i theta} (the same Miller recurrence as eval_local_n).
template <int PQ>
__device__ void local_basis_2d(const KD &kd, cplx c, double R, cplx tp, cplx *B)
{
    if (kd.kernel == QBX_LAPLACE2D) {
        const cplx x = (1.0 / R) * (tp - c);
        cplx pw = {1, 0};
        for (int l = 0; l <= PQ; ++l) {
            B[l] = pw;
            pw = pw * x;
        }
        return;
    }
    for (int m = 0; m <= 2 * PQ; ++m) B[m] = {0, 0};
    const cplx d = tp - c;
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    if (x < 1e-300) {
        B[PQ] = {1, 0};
        return;
    }
    const cplx e1 = {d.x / rho, d.y / rho};
    const cplx em1 = conj(e1);
    cplx e = cpow_int(e1, PQ);
    const double norm = miller_j(
        x, PQ,
        [&](int n, double f) {
            B[PQ + n] = f * e;
            if (n > 0) {
                cplx fc = f * conj(e);
                if (n & 1) fc = -fc;
                B[PQ - n] = fc;
            }
            e = e * em1;
        },
        [&](double sc) {
            for (int m = 0; m <= 2 * PQ; ++m) B[m] = sc * B[m];
        });
    for (int m = 0; m <= 2 * PQ; ++m) B[m] = norm * B[m];
}

// The order-PQ QBX expansion of a unit source at w = s - c (normal nu) about
// the center, evaluated against the row's basis B: one entry of the QBX
// layer-potential matrix (P2QBXL of one source, then QBXL2P).
template <int PQ>
__device__ __forceinline__ cplx qbx_pair_value(const KD &kd, cplx w, cplx nu, double R,
                                               const double *tab, const cplx *B)
{
    QAcc<PQ, 1> A;
    qacc_zero(A);
    const cplx one[1] = {{1.0, 0.0}};
    p2qbxl_pair<PQ, 1>(kd, w, nu, one, R, tab, A);
    cplx phi = {0, 0};
    if (kd.kernel == QBX_LAPLACE2D) {
#pragma unroll
        for (int l = 0; l <= PQ; ++l) phi += A.v[0][l] * B[l];
        return {-phi.x * INV2PI, 0.0};
    }
#pragma unroll
    for (int n = 0; n <= 2 * PQ; ++n) phi += A.v[0][n] * B[n];
    return {-0.25 * phi.y, 0.25 * phi.x};
}

// One block per target row t:  mat[t][e n1 + j] = sum_i K(t, s_ei) w_ei U[i][j]
// with K the QBX entry about the row's center (PQ >= 0) or the kernel itself
// (PQ < 0).  The fine sources of LM_THREADS / nf consecutive elements take one
// thread each; their weighted values pass through LDS into the per-element
// contraction with U, whose outputs are consecutive in the row (coalesced).
template <int PQ>
__global__ void __launch_bounds__(LM_THREADS)
k_lpot_matrix_2d(KD kd, int nel, int n1, int nf, const double *U, const double *src,
                 const double *nrm, const double *w, int nt, const double *tgt,
                 const double *ctr, const double *rad, long ld, double *mat)
{
    __shared__ double kv[2 * LM_THREADS];
    __shared__ double Us[LM_UMAX];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    __shared__ cplx B[PQ >= 0 ? 2 * PQ + 1 : 1];
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int ns = nel * nf;
    const cplx tp = ldpt(tgt, nt, t);
    cplx cc = {0, 0};
    double R = 1.0;
    if (PQ >= 0) {
        cc = ldpt(ctr, nt, t);
        R = rad[t];
    }
    for (int i = tid; i < nf * n1; i += LM_THREADS)
        Us[i] = U ? U[i] : (i / n1 == i % n1 ? 1.0 : 0.0);
    hankel_table_fill(htab, tid, LM_THREADS);
    if (PQ >= 0 && tid == 0) local_basis_2d<(PQ >= 0 ? PQ : 0)>(kd, cc, R, tp, B);
    __syncthreads();
    const int epb = LM_THREADS / nf;
    const int le = tid / nf, fi = tid % nf;
    for (int e0 = 0; e0 < nel; e0 += epb) {
        const int e = e0 + le;
        cplx v = {0, 0};
        if (le < epb && e < nel) {
            const int s = e * nf + fi;
            const cplx sp = ldpt(src, ns, s);
            const cplx nu = kd.dip ? ldpt(nrm, ns, s) : cplx{0, 0};
            int toff = 0;
            asm volatile("" : "+s"(toff));
            if constexpr (PQ < 0)
                v = p2p_value_wave(kd, sp, nu, tp, htab + toff);
            else
                v = qbx_pair_value<PQ>(kd, sp - cc, nu, R, htab + toff, B);
            v = w[s] * v;
        }
        kv[2 * tid] = v.x;
        kv[2 * tid + 1] = v.y;
        __syncthreads();
        const int ne = min(epb, nel - e0);
        for (int o = tid; o < ne * n1; o += LM_THREADS) {
            const int le2 = o / n1, j = o % n1;
            const double *kr = kv + 2 * le2 * nf;
            cplx acc = {0, 0};
            for (int ii = 0; ii < nf; ++ii) {
                const double u = Us[ii * n1 + j];
                acc.x += u * kr[2 * ii];
                acc.y += u * kr[2 * ii + 1];
            }
            stc(mat + 2 * ((size_t)t * ld + (size_t)(e0 + le2) * n1 + j), acc);
        }
        __syncthreads();
    }
}

// P2QBXL for many right-hand sides (Helmholtz, qbx_order <= 5) on the
// matrix cores.  A block takes a group of up to GC_GROUP centers of one target
// box (gc_group_plan) and walks the box's list-1 sources in chunks of PQM_CH:
// the threads evaluate each (center, source) pair's unit-strength expansion
// once (p2qbxl_pair) into LDS -- rows (center, coefficient), columns sources:
// the A operand -- next to the chunk's strengths for up to 16 right-hand sides
// (the B operand), and waves 0..RT-1 accumulate the complex product of their
// 16-row tile with v_mfma_f64_16x16x4f64.  The expansion work is done once per
// pair instead of once per right-hand-side chunk, and the sums over sources
// become matrix products.  Results are added to qexp (zeroed by the caller).
constexpr int PQM_CH = 48;

template <int PQ>
__global__ void __launch_bounds__(256)
k_p2qbxl_mfma(KD kd, TD t, const int *ubox, const int *ut0, const int *gcs, const int *gcl,
              const double *ctr, const double *rad, int nc, const int *l1s, const int *l1l,
              const double *src, int ns, const double *str, const double *nrm, double *qexp)
{
    constexpr int NQ = 2 * PQ + 1;
    constexpr int ROWS = GC_GROUP * NQ;
    constexpr int RT = (ROWS + 15) / 16;
    __shared__ double Ab[2][RT * 16][PQM_CH];
    __shared__ double Sb[2][PQM_CH][16];
    __shared__ double sx[PQM_CH], sy[PQM_CH], snx[PQM_CH], sny[PQM_CH];
    __shared__ __align__(16) double htab[4 * HANKEL_NSER];
    const int u = blockIdx.x, it = ubox[u], c0 = ut0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    const bool dip = kd.dip;
    // the pair this thread evaluates: center ci, chunk slot j
    const int ci = tid / PQM_CH, jslot = tid % PQM_CH;
    const bool has_c = ci < ncen;
    cplx cc = {0, 0};
    double R = 1.0;
    if (has_c) {
        const int c = gcl[gcs[it] + c0 + ci];
        cc = ldpt(ctr, nc, c);
        R = rad[c];
    }
    hankel_table_fill(htab, tid, 256);
    m2l_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};

    auto flush = [&](int n) {
        __syncthreads();  // sx.. and Sb of the chunk written
        if (ci < GC_GROUP) {
            QAcc<PQ, 1> A;
            qacc_zero<PQ, 1>(A);
            if (has_c && jslot < n) {
                int toff = 0;
                asm volatile("" : "+s"(toff));
                const cplx one[1] = {{1.0, 0.0}};
                const cplx nu = dip ? cplx{snx[jslot], sny[jslot]} : cplx{0, 0};
                p2qbxl_pair<PQ, 1>(kd, cplx{sx[jslot], sy[jslot]} - cc, nu, one, R, htab + toff,
                                   A);
            }
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                Ab[0][ci * NQ + i][jslot] = A.v[0][i].x;
                Ab[1][ci * NQ + i][jslot] = A.v[0][i].y;
            }
        }
        __syncthreads();  // basis complete
        if (wv < RT) {
#pragma unroll 4
            for (int ks = 0; ks < PQM_CH / 4; ++ks) {
                const int k = ks * 4 + kr;
                const double are = Ab[0][wv * 16 + col][k], aim = Ab[1][wv * 16 + col][k];
                const double bre = Sb[0][k][col], bim = Sb[1][k][col];
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
            }
        }
        __syncthreads();  // chunk consumed
    };

    int fill = 0;
    const int e1 = l1s[it + 1];
    for (int e = l1s[it]; e < e1; ++e) {
        const int Sbx = l1l[e];
        int s = t.sstart[Sbx];
        const int s1 = s + t.scount[Sbx];
        while (s < s1) {
            const int take = min(s1 - s, PQM_CH - fill);
            for (int j = tid; j < take; j += 256) {
                const int si = s + j, d = fill + j;
                sx[d] = src[si];
                sy[d] = src[(size_t)ns + si];
                if (dip) {
                    snx[d] = nrm[si];
                    sny[d] = nrm[(size_t)ns + si];
                }
            }
            for (int q = tid; q < take * 16; q += 256) {
                const int j = q >> 4, r = q & 15;
                const int si = s + j, d = fill + j;
                double vr = 0.0, vi = 0.0;
                if (r < rc) {
                    const double *pq = str + 2 * ((size_t)si * nr + r0 + r);
                    vr = pq[0];
                    vi = pq[1];
                }
                Sb[0][d][r] = vr;
                Sb[1][d][r] = vi;
            }
            fill += take;
            s += take;
            if (fill == PQM_CH) {
                flush(PQM_CH);
                fill = 0;
            }
        }
    }
    if (fill > 0) {
        // zero the unused source slots' strengths (their basis is zeroed in flush)
        for (int q = tid; q < (PQM_CH - fill) * 16; q += 256) {
            const int d = fill + (q >> 4), r = q & 15;
            Sb[0][d][r] = 0.0;
            Sb[1][d][r] = 0.0;
        }
        flush(fill);
    }
    if (wv < RT) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = wv * 16 + 4 * j + kr;
            const int cj = row / NQ, i = row % NQ;
            if (cj < ncen && col < rc) {
                const int c = gcl[gcs[it] + c0 + cj];
                double *o = qexp + 2 * (((size_t)c * kd.ncq + i) * nr + r0 + col);
                o[0] += cre[j];
                o[1] += cim[j];
            }
        }
    }
}

// M2QBXL for many right-hand sides (Helmholtz, qbx_order <= 5) on the
// matrix cores: for a group of up to GC_GROUP centers of one target box and
// each box D of the target box's list 3, Q_c[m][r] += sum_n g_c(n - m)
// M_D[n][r] with the Graf sequence g_c(d) = H_d(k rho) e^{i d theta} (d >= 0),
// (-1)^d H_d e^{-i d theta} at -d (rho e^{i theta} = c - c_D; the same terms
// as k_m2qbxl).  One thread per center runs its Hankel recurrence into LDS,
// the rows (center, coefficient) form the A operand, the multipole
// coefficients of up to 16 right-hand sides the B operand.
constexpr int M2QM_G = 2 * (MAXP + 5) + 1;  // g length bound (p <= MAXP, p_q <= 5)

template <int PQ>
__global__ void __launch_bounds__(256)
k_m2qbxl_mfma(KD kd, TD t, const int *ubox, const int *ut0, const int *gcs, const int *gcl,
              const double *ctr, int nc, const int *l3s, const int *l3l, const double *mpole,
              double *qexp)
{
    constexpr int NQ = 2 * PQ + 1;
    constexpr int ROWS = GC_GROUP * NQ;
    constexpr int RT = (ROWS + 15) / 16;
    __shared__ double gs[GC_GROUP][2 * M2QM_G];
    const int u = blockIdx.x, it = ubox[u], c0 = ut0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    // this lane's A row (wave wv < RT)
    const int row = wv * 16 + col;
    const int rci = row / NQ, rm = row % NQ - PQ;
    const bool rvalid = wv < RT && rci < ncen;
    cplx cc = {0, 0};
    if (tid < ncen) cc = ldpt(ctr, nc, gcl[gcs[it] + c0 + tid]);
    m2l_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
    for (int e = l3s[it]; e < l3s[it + 1]; ++e) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int p1 = kd.order[lv];
        const int lmax = p1 + PQ;
        const int ncl = 2 * p1 + 1;
        const double *a = mpole + ecoef_off(kd, D, lv);
        __syncthreads();  // previous g consumed
        if (tid < ncen) {
            double *g = gs[tid];
            const cplx d = cc - boxc(t, D);
            const double rho = sqrt(abs2(d));
            const double x = kd.k * rho;
            const cplx e1 = {d.x / rho, d.y / rho};
            cplx ep = {1, 0};
            cplx hm1, hm;
            hankel01_call(x, hm1, hm);
            const double tox = 2.0 / x;
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
                g[2 * (lmax + l)] = he.x;
                g[2 * (lmax + l) + 1] = he.y;
                if (l > 0) {
                    g[2 * (lmax - l)] = hc.x;
                    g[2 * (lmax - l) + 1] = hc.y;
                }
                ep = ep * e1;
            }
        }
        __syncthreads();  // g ready
        if (wv < RT) {
            const int KS = (ncl + 3) / 4;
            for (int ks = 0; ks < KS; ++ks) {
                const int k = ks * 4 + kr;  // 0-based coefficient n = k - p1
                double bre = 0.0, bim = 0.0;
                if (k < ncl && col < rc) {
                    const cplx b = ldcn(kd, a, k - p1, r0 + col);
                    bre = b.x;
                    bim = b.y;
                }
                double are = 0.0, aim = 0.0;
                if (rvalid && k < ncl) {
                    const int dd = (k - p1) - rm + lmax;  // d = n - m in [0, 2 lmax]
                    are = gs[rci][2 * dd];
                    aim = gs[rci][2 * dd + 1];
                }
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
                cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
            }
        }
    }
    if (wv < RT) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int orow = wv * 16 + 4 * j + kr;
            const int cj = orow / NQ, i = orow % NQ;
            if (cj < ncen && col < rc) {
                const int c = gcl[gcs[it] + c0 + cj];
                double *o = qexp + 2 * (((size_t)c * kd.ncq + i) * nr + r0 + col);
                o[0] += cre[j];
                o[1] += cim[j];
            }
        }
    }
}

// L2QBXL for many right-hand sides (Helmholtz, qbx_order <= 5) on the
// matrix cores: the centers of a group share their target box's local
// expansion L_B; Q_c[m][r] += sum_n g_c(n - m) L_B[n][r] with g_c(d) =
// J_d(k rho) e^{i d theta} (d >= 0), (-1)^d J_d e^{-i d theta} at -d
// (rho e^{i theta} = c - c_B; Miller's recurrence, as k_l2qbxl).  One thread
// per center writes its normalised sequence into LDS (the A operand rows).
template <int PQ>
__global__ void __launch_bounds__(256)
k_l2qbxl_mfma(KD kd, TD t, const int *ubox, const int *ut0, const int *gcs, const int *gcl,
              const int *tboxes, const double *ctr, int nc, const double *local, double *qexp)
{
    constexpr int NQ = 2 * PQ + 1;
    constexpr int ROWS = GC_GROUP * NQ;
    constexpr int RT = (ROWS + 15) / 16;
    __shared__ double gs[GC_GROUP][2 * M2QM_G];
    const int u = blockIdx.x, it = ubox[u], c0 = ut0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    const int Bx = tboxes[it];
    const int lv = t.level[Bx];
    const int p1 = kd.order[lv];
    const int lmax = p1 + PQ;
    const int ncl = 2 * p1 + 1;
    const double *a = local + ecoef_off(kd, Bx, lv);
    if (tid < ncen) {
        double *g = gs[tid];
        for (int w = 0; w <= 2 * lmax; ++w) g[2 * w] = g[2 * w + 1] = 0.0;
        const cplx d = ldpt(ctr, nc, gcl[gcs[it] + c0 + tid]) - boxc(t, Bx);
        const double rho = sqrt(abs2(d));
        const double x = kd.k * rho;
        if (x < 1e-300) {
            g[2 * lmax] = 1.0;
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
                    g[2 * (lmax + l)] = fe.x;
                    g[2 * (lmax + l) + 1] = fe.y;
                    if (l > 0) {
                        g[2 * (lmax - l)] = fc.x;
                        g[2 * (lmax - l) + 1] = fc.y;
                    }
                    e = e * em1;
                },
                [&](double sc) {
                    for (int w = 0; w <= 2 * lmax; ++w) {
                        g[2 * w] *= sc;
                        g[2 * w + 1] *= sc;
                    }
                });
            for (int w = 0; w <= 2 * lmax; ++w) {
                g[2 * w] *= norm;
                g[2 * w + 1] *= norm;
            }
        }
    }
    __syncthreads();
    if (wv >= RT) return;
    const int row = wv * 16 + col;
    const int rci = row / NQ, rm = row % NQ - PQ;
    const bool rvalid = rci < ncen;
    m2l_d4 cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
    const int KS = (ncl + 3) / 4;
    for (int ks = 0; ks < KS; ++ks) {
        const int k = ks * 4 + kr;
        double bre = 0.0, bim = 0.0;
        if (k < ncl && col < rc) {
            const cplx b = ldcn(kd, a, k - p1, r0 + col);
            bre = b.x;
            bim = b.y;
        }
        double are = 0.0, aim = 0.0;
        if (rvalid && k < ncl) {
            const int dd = (k - p1) - rm + lmax;
            are = gs[rci][2 * dd];
            aim = gs[rci][2 * dd + 1];
        }
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int orow = wv * 16 + 4 * j + kr;
        const int cj = orow / NQ, i = orow % NQ;
        if (cj < ncen && col < rc) {
            const int c = gcl[gcs[it] + c0 + cj];
            double *o = qexp + 2 * (((size_t)c * kd.ncq + i) * nr + r0 + col);
            o[0] += cre[j];
            o[1] += cim[j];
        }
    }
}

// M2QBXL over balanced work units (see unit_plan): the unit's TG global
// centers of one target box take G lanes apiece that split the box's list-3
// boxes; shuffle reduction, then the expansions are added.
template <int PQ, int NRC>
__global__ void __launch_bounds__(64) k_m2qbxl(KD kd, TD t, const double *B, const int *ubox,
                                               const int *ut0, const int *box_g,
                                               const int *gcs, const int *gcl,
                                               const double *ctr, const double *rad, int nc,
                                               const int *l3s, const int *l3l,
                                               const double *mpole, double *qexp)
{
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int cl = ut0[u] + tsub;
    const bool active = cl < gcs[it + 1] - gcs[it];
    const int c = active ? gcl[gcs[it] + cl] : 0;
    const cplx cc = active ? ldpt(ctr, nc, c) : cplx{0, 0};
    const double R2 = active ? rad[c] : 1.0;
    QAcc<PQ, NRC> A;
    qacc_zero<PQ, NRC>(A);
    const int ebeg = active ? l3s[it] + sub : 0, eend = active ? l3s[it + 1] : 0;
    for (int e = ebeg; e < eend; e += G) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int p1 = kd.order[lv];
        const double *a = mpole + ecoef_off(kd, D, lv);
        const cplx cd = boxc(t, D);
        if (kd.kernel == QBX_LAPLACE2D) {
            const cplx z0 = cd - cc;
            const cplx iz = cinv(z0);
            const cplx ms = -(kd.rscale[lv] * iz), tt = R2 * iz;
            const cplx lg = clogz(-z0);
            cplx Sv[NRC][PQ + 1];
#pragma unroll
            for (int q = 0; q < NRC; ++q) {
                const cplx a0 = q < kd.rc ? ldcn(kd, a, 0, kd.r0 + q) : cplx{0, 0};
                Sv[q][0] = a0 * lg;
#pragma unroll
                for (int l = 1; l <= PQ; ++l) Sv[q][l] = (-1.0 / l) * a0;
            }
            cplx pw = {1, 0};
            for (int k = 1; k <= p1; ++k) {
                pw = pw * ms;
#pragma unroll
                for (int q = 0; q < NRC; ++q) {
                    if (q >= kd.rc) continue;
                    const cplx ak = ldcn(kd, a, k, kd.r0 + q) * pw;
                    Sv[q][0] += ak;
#pragma unroll
                    for (int l = 1; l <= PQ; ++l) Sv[q][l] += g_binom(B, l + k - 1, k - 1) * ak;
                }
            }
            cplx tl = {1, 0};
#pragma unroll
            for (int l = 0; l <= PQ; ++l) {
#pragma unroll
                for (int q = 0; q < NRC; ++q) A.v[q][l] += Sv[q][l] * tl;
                tl = tl * tt;
            }
        } else {
            const cplx d = cc - cd;
            const double rho = sqrt(abs2(d));
            const double x = kd.k * rho;
            const cplx e1 = {d.x / rho, d.y / rho};
            cplx ep = {1, 0};
            cplx hm1, hm;
            hankel01_call(x, hm1, hm);
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
                    const int n2 = m - l;
#pragma unroll
                    for (int q = 0; q < NRC; ++q) {
                        if (q >= kd.rc) continue;
                        if (n1 >= -p1 && n1 <= p1)
                            A.v[q][m + PQ] += ldcn(kd, a, n1, kd.r0 + q) * he;
                        if (l > 0 && n2 >= -p1 && n2 <= p1)
                            A.v[q][m + PQ] += ldcn(kd, a, n2, kd.r0 + q) * hc;
                    }
                }
                ep = ep * e1;
            }
        }
    }
    qacc_group_reduce<PQ, NRC>(A, G);
    if (active && sub == 0) qbx_store<PQ, NRC>(kd, qexp, c, A, 1);
}

// (64-thread blocks declared: the default bound of 1024 threads capped the
// kernel at 128 VGPRs, and with 4 right-hand sides it spilled 676 bytes per
// lane -- 2.4x slower)
template <int PQ, int NRC>
__global__ void __launch_bounds__(64)
k_l2qbxl(KD kd, TD t, const double *B, int ng, const int *gcent, const int *c2tb,
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
    QAcc<PQ, NRC> A;
    qacc_zero<PQ, NRC>(A);
    if (kd.kernel == QBX_LAPLACE2D) {
        const double R1 = kd.rscale[lv];
        const cplx u = (1.0 / R1) * (cc - cb);
        for (int l = p1; l >= 0; --l) {
#pragma unroll
            for (int q = 0; q < NRC; ++q) {
                if (q >= kd.rc) continue;
                const cplx al = ldcn(kd, a, l, kd.r0 + q);
#pragma unroll
                for (int j = 0; j <= PQ; ++j)
                    if (j <= l) A.v[q][j] = A.v[q][j] * u + g_binom(B, l, j) * al;
            }
        }
        double rp = 1.0;
        const double r21 = R2 / R1;
#pragma unroll
        for (int j = 0; j <= PQ; ++j) {
#pragma unroll
            for (int q = 0; q < NRC; ++q) A.v[q][j] = rp * A.v[q][j];
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
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (abs(m) <= p1 && q < kd.rc) A.v[q][m + PQ] = ldcn(kd, a, m, kd.r0 + q);
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
                        const int n2 = m - l;
#pragma unroll
                        for (int q = 0; q < NRC; ++q) {
                            if (q >= kd.rc) continue;
                            if (n1 >= -p1 && n1 <= p1)
                                A.v[q][m + PQ] += ldcn(kd, a, n1, kd.r0 + q) * fe;
                            if (l > 0 && n2 >= -p1 && n2 <= p1)
                                A.v[q][m + PQ] += ldcn(kd, a, n2, kd.r0 + q) * fc;
                        }
                    }
                    e = e * em1;
                },
                [&](double s) {
#pragma unroll
                    for (int q = 0; q < NRC; ++q)
#pragma unroll
                        for (int m = 0; m < 2 * PQ + 1; ++m) A.v[q][m] = s * A.v[q][m];
                });
#pragma unroll
            for (int q = 0; q < NRC; ++q)
#pragma unroll
                for (int m = 0; m < 2 * PQ + 1; ++m) A.v[q][m] = norm * A.v[q][m];
        }
    }
    qbx_store<PQ, NRC>(kd, qexp, c, A, accumulate);
}

// Helmholtz L2QBXL for one right-hand side with LS lanes per center: lane
// `sub` keeps the outputs m = -PQ + sub + LS i.  Each lane runs the center's
// Miller recurrence itself (a few flops per step) and adds its outputs in the
// order of k_l2qbxl, so the sums are the same; the split gives the launch LS
// times the wavefronts of the one-thread-per-center kernel, which on its own
// (one wavefront per 64 centers, a long serial chain each) was latency-bound.
template <int PQ, int LS>
__global__ void __launch_bounds__(256)
k_l2qbxl_split(KD kd, TD t, int ng, const int *gcent, const int *c2tb, const int *tboxes,
               const double *ctr, int nc, const double *local, double *qexp)
{
    constexpr int MI = (2 * PQ + 1 + LS - 1) / LS;
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int g = (int)(gid / LS), sub = (int)(gid % LS);
    if (g >= ng) return;
    const int c = gcent[g];
    const int Bx = tboxes[c2tb[c]];
    const int lv = t.level[Bx];
    const int p1 = kd.order[lv];
    const cplx cc = ldpt(ctr, nc, c);
    const double *a = local + ecoef_off(kd, Bx, lv);
    const cplx d = cc - boxc(t, Bx);
    const double rho = sqrt(abs2(d));
    const double x = kd.k * rho;
    cplx A[MI];
#pragma unroll
    for (int i = 0; i < MI; ++i) A[i] = {0, 0};
    if (x < 1e-300) {
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            const int m = -PQ + sub + LS * i;
            if (m <= PQ && abs(m) <= p1) A[i] = ldcn(kd, a, m, 0);
        }
    } else {
        const cplx e1 = {d.x / rho, d.y / rho};
        const cplx em1 = conj(e1);
        cplx e = cpow_int(e1, p1 + PQ);
        const double norm = miller_j(
            x, p1 + PQ,
            [&](int l, double f) {
                const cplx fe = f * e;
                cplx fc = f * conj(e);
                if (l & 1) fc = -fc;
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int m = -PQ + sub + LS * i;
                    const int n1 = m + l, n2 = m - l;
                    if (m <= PQ) {
                        if (n1 >= -p1 && n1 <= p1) A[i] += ldcn(kd, a, n1, 0) * fe;
                        if (l > 0 && n2 >= -p1 && n2 <= p1) A[i] += ldcn(kd, a, n2, 0) * fc;
                    }
                }
                e = e * em1;
            },
            [&](double sc) {
#pragma unroll
                for (int i = 0; i < MI; ++i) A[i] = sc * A[i];
            });
#pragma unroll
        for (int i = 0; i < MI; ++i) A[i] = norm * A[i];
    }
    double *o = qexp + 2 * (size_t)c * kd.ncq;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const int m = -PQ + sub + LS * i;
        if (m <= PQ) {
            double *pm = o + 2 * (m + PQ);
            stc(pm, A[i] + ldc(pm));
        }
    }
}


// QBXL2P: one thread per (center, target) entry of center_to_tree_targets
template <int NRC, bool DER = false>
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
    const double *a = qexp + 2 * ((size_t)c * kd.ncq +
                                  (kd.kernel == QBX_HELMHOLTZ2D ? kd.pq : 0)) * kd.nr;
    cplx v[NRC];
    eval_local_n<NRC, DER>(kd, a, kd.pq, cc, rad[c], tp, v, DER ? ldir(kd, ti) : cplx{0, 0});
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) stc(pot + 2 * ((size_t)ti * kd.nr + kd.r0 + q), v[q]);
}

__global__ void k_finalize(int nt, int nc, int nr, const int *user_to_tree, const double *pot,
                           double *out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(nt - nc) * nr) return;
    const int u = nc + (int)(idx / nr), r = (int)(idx % nr);
    const int ti = user_to_tree[u];
    out[2 * idx] = pot[2 * ((size_t)ti * nr + r)];
    out[2 * idx + 1] = pot[2 * ((size_t)ti * nr + r) + 1];
}

// -------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------
// Pascal's triangle to BINOM_N rows, generated at compile time by the
// additions of the recurrence (IEEE round to nearest, as at run time) into an
// initialised device array: no allocation, no copy, nothing to order against
// the streams that read it.  The array has external linkage and is not const:
// kernels reach it only through the pointer the host looks up, and the device
// compiler drops an internal constant that no device code names (the
// runtime's symbol lookup then aborts).
struct BinomTab {
    double v[BINOM_N * BINOM_N];
};

constexpr BinomTab make_binom()
{
    BinomTab t{};
    for (int n = 0; n < BINOM_N; ++n) {
        t.v[n * BINOM_N] = 1.0;
        for (int k = 1; k <= n; ++k)
            t.v[n * BINOM_N + k] = t.v[(n - 1) * BINOM_N + k - 1] +
                                   (k < n ? t.v[(n - 1) * BINOM_N + k] : 0.0);
    }
    return t;
}

__device__ BinomTab qbx_binom_tab = make_binom();
static const double *g_binom_dev[64] = {nullptr};

static int binom_table(const double **out)
{
    static std::mutex mu;
    int dev = 0;
    QBX_HIP(hipGetDevice(&dev));
    QBX_CHECK(dev >= 0 && dev < 64, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!g_binom_dev[dev]) {
        void *p = nullptr;
        QBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(qbx_binom_tab)));
        g_binom_dev[dev] = (const double *)p;
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
    QBX_CHECK(d->nrhs >= 0, "nrhs must be >= 0 (0 means 1)");
    if (T) {
        QBX_CHECK(T->dim == 2, "2D kernels need a 2D tree");
        QBX_CHECK(T->nlevels <= FML, "tree has too many levels (%d)", T->nlevels);
        QBX_CHECK(d->nlevels >= T->nlevels, "level_orders has %d entries, tree has %d levels",
                  d->nlevels, T->nlevels);
    }
    if (d->kernel == QBX_HELMHOLTZ2D) QBX_CHECK(d->k > 0, "Helmholtz needs k > 0");
    QBX_CHECK(d->target_deriv == 0 || d->target_deriv == 1, "target_deriv must be 0 or 1");
    QBX_CHECK(!d->target_deriv || d->target_dir, "target_deriv needs target_dir");
    memset(kd, 0, sizeof(*kd));
    kd->tder = d->target_deriv;
    kd->tdir = d->target_dir;
    kd->ntgt = T ? T->ntargets : 0;
    kd->kernel = d->kernel;
    kd->dip = d->source_kind == QBX_SOURCE_D;
    kd->pq = d->qbx_order;
    kd->k = d->k;
    kd->ik = d->k > 0 ? 1.0 / d->k : 0.0;
    kd->nsep = nsep;
    kd->m2l_w = 2 * nsep + 1;
    kd->nr = d->nrhs > 0 ? d->nrhs : 1;
    kd->r0 = 0;
    kd->rc = 1;
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
    const int ncls_ = (2 * kd->m2l_w + 1) * (2 * kd->m2l_w + 1);
    if (ctx->tab_m2l_ext) {
        // allocation-free call: build into its scratch every time (other
        // calls on the device may have overwritten the array since)
        hipLaunchKernelGGL(k_tab_m2l, dim3(cdiv(kd->nlevels * ncls_, 64)), dim3(64), 0, st, *kd,
                           kd->nlevels, ctx->tab_m2l_ext);
        if (kd->nlevels > 1)
            hipLaunchKernelGGL(k_tab_mm, dim3(cdiv(4 * (kd->nlevels - 1), 64)), dim3(64), 0, st,
                               *kd, kd->nlevels, ctx->tab_mm_ext);
        QBX_HIP(hipGetLastError());
        kd->tab_m2l = ctx->tab_m2l_ext;
        kd->tab_mm = ctx->tab_mm_ext;
        return QBX_OK;
    }
    QBX_TRY(note_capture(ctx, st));
    if (key != ctx->tab_key) {
        // a graph replays the tables it captured: never rebuild under one
        QBX_TRY(may_replace(ctx, st, "Helmholtz translation tables"));
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
// Launch shape of the LDS translation kernels (Helmholtz only): threads
// cover the 2 p_out + 1 outputs and one staged source box fits in 64 KiB.
// False means the plain kernels are used instead.
struct HtLaunch {
    int threads = 0, p_in = 0, L = 0;
};

template <int NRC>
static size_t ht_bytes(const HtLaunch &h)
{
    return 8UL * ht_entry_doubles(NRC, h.p_in, h.L);
}

static bool ht_launch(const KD &kd, int p_out, int p_in, int L, HtLaunch *h)
{
    if (kd.kernel != QBX_HELMHOLTZ2D) return false;
    const int nout = 2 * p_out + 1;
    if (nout > HT_MAX_THREADS) return false;
    if (8L * ht_entry_doubles(NRC_S, p_in, L) > 65536) return false;
    h->threads = 64 * ((nout + 63) / 64);
    h->p_in = p_in;
    h->L = L;
    return true;
}

// Shape of k_p2e_h at order p: S sources per chunk within the LDS budget,
// threads covering max(2 p + 1, S).  False: use the plain kernels.
static bool p2e_launch(const KD &kd, int p, P2EShape *sh, int *threads, size_t *bytes)
{
    if (kd.kernel != QBX_HELMHOLTZ2D || 2 * p + 1 > HT_MAX_THREADS) return false;
    P2EShape h;
    h.R = (p + 2) | 1;
    const long per = 8L * (4 * h.R + 2 * NRC_S + 2 + 1);
    h.S = (int)std::min(64L, HT_LDS_BUDGET / per);
    if (h.S < 4) return false;
    *sh = h;
    *threads = 64 * ((std::max(2 * p + 1, h.S) + 63) / 64);
    *bytes = 8UL * h.doubles(NRC_S);
    return true;
}

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

// Run LAUNCH(N) over the right-hand-side chunks: N = 1 for a single
// right-hand side, CH otherwise; kd.r0 / kd.rc are set per chunk.
#define FOR_RHS_CHUNKS(kd, CH, LAUNCH)                                                  \
    do {                                                                                 \
        if ((kd).nr == 1) {                                                              \
            (kd).r0 = 0;                                                                 \
            (kd).rc = 1;                                                                 \
            LAUNCH(1);                                                                   \
        } else {                                                                         \
            for (int r0_ = 0; r0_ < (kd).nr; r0_ += (CH)) {                             \
                (kd).r0 = r0_;                                                           \
                (kd).rc = std::min((CH), (kd).nr - r0_);                                 \
                LAUNCH(CH);                                                              \
            }                                                                            \
        }                                                                                \
    } while (0)

#define DISPATCH_PQ_NR(pq, NR, KERNEL, ...)                                             \
    switch (pq) {                                                                        \
    case 0: hipLaunchKernelGGL((KERNEL<0, NR>), __VA_ARGS__); break;                     \
    case 1: hipLaunchKernelGGL((KERNEL<1, NR>), __VA_ARGS__); break;                     \
    case 2: hipLaunchKernelGGL((KERNEL<2, NR>), __VA_ARGS__); break;                     \
    case 3: hipLaunchKernelGGL((KERNEL<3, NR>), __VA_ARGS__); break;                     \
    case 4: hipLaunchKernelGGL((KERNEL<4, NR>), __VA_ARGS__); break;                     \
    case 5: hipLaunchKernelGGL((KERNEL<5, NR>), __VA_ARGS__); break;                     \
    case 6: hipLaunchKernelGGL((KERNEL<6, NR>), __VA_ARGS__); break;                     \
    case 7: hipLaunchKernelGGL((KERNEL<7, NR>), __VA_ARGS__); break;                     \
    case 8: hipLaunchKernelGGL((KERNEL<8, NR>), __VA_ARGS__); break;                     \
    case 9: hipLaunchKernelGGL((KERNEL<9, NR>), __VA_ARGS__); break;                     \
    case 10: hipLaunchKernelGGL((KERNEL<10, NR>), __VA_ARGS__); break;                   \
    case 11: hipLaunchKernelGGL((KERNEL<11, NR>), __VA_ARGS__); break;                   \
    case 12: hipLaunchKernelGGL((KERNEL<12, NR>), __VA_ARGS__); break;                   \
    default: set_error("qbx_order out of range"); return QBX_ERR_INVALID;               \
    }

static inline long exp_bytes(const KD &kd) { return 16L * kd.lcoef[kd.nlevels] * kd.nr; }

// Helmholtz P2M (LDS kernel) of the source boxes at level l; the levels are
// independent (each writes its own boxes)
static void p2m_level_h(hipStream_t st, KD &kd, const TD &td, const qbx_tree *T,
                        const qbx_trav *R, int l, const double *str, const double *nrm,
                        double *mpole)
{
    const int i0 = R->h_level_start_source_box_nrs[l];
    const int i1 = R->h_level_start_source_box_nrs[l + 1];
    if (i1 == i0) return;
    P2EShape sh;
    int th;
    size_t by;
    p2e_launch(kd, kd.order[l], &sh, &th, &by);
#define L_P2MH(N)                                                                          \
    hipLaunchKernelGGL((k_p2e_h<N, false>), dim3(i1 - i0), dim3(th), by, st, kd, td, l, i0, sh, \
                       R->source_boxes.p, nullptr, nullptr, T->sources.p, T->nsources, str,  \
                       nrm, mpole)
    FOR_RHS_CHUNKS(kd, NRC_S, L_P2MH);
#undef L_P2MH
}

// P2M with the levels spread over the streams sts[0..nst) (zeroing on
// sts[0] first, every stream joined back into sts[0] at the end): the
// per-level launches are small and independent, so they run side by side
// instead of one after another.  False when the LDS kernel does not apply
// (the caller then runs qbx_form_multipoles).
static int p2m_multistream(qbx_ctx *ctx, hipStream_t *sts, int nst, hipEvent_t *evs,
                           const qbx_tree *T, const qbx_trav *R, const qbx_kernel_desc *d,
                           const double *str, const double *nrm, double *mpole, bool *done)
{
    *done = false;
    if (!d || d->kernel != QBX_HELMHOLTZ2D) return QBX_OK;
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    for (int l = 0; l < T->nlevels; ++l) {
        P2EShape sh;
        int th;
        size_t by;
        if (!p2e_launch(kd, kd.order[l], &sh, &th, &by)) return QBX_OK;
    }
    const TD td = make_td(T);
    QBX_TRY(zero_fill(mpole, (exp_bytes(kd)) / 8, sts[0]));
    QBX_HIP(hipEventRecord(evs[0], sts[0]));
    for (int k = 1; k < nst; ++k) QBX_HIP(hipStreamWaitEvent(sts[k], evs[0], 0));
    // the busiest levels first, each to the stream with the least work so far
    // (boxes x coefficients): a small level queued behind a big one would
    // start only when the near field, issued meanwhile on the other streams,
    // has dispatched all its workgroups
    std::vector<std::pair<long, int>> order;
    for (int l = 0; l < T->nlevels; ++l) {
        const long nb = R->h_level_start_source_box_nrs[l + 1] - R->h_level_start_source_box_nrs[l];
        if (nb > 0) order.push_back({-nb * (2L * kd.order[l] + 1), l});
    }
    std::sort(order.begin(), order.end());
    std::vector<long> load(nst, 0);
    for (size_t k = 0; k < order.size(); ++k) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[j] -= order[k].first;
        p2m_level_h(sts[j], kd, td, T, R, order[k].second, str, nrm, mpole);
    }
    QBX_HIP(hipGetLastError());
    for (int k = 1; k < nst; ++k) {
        QBX_HIP(hipEventRecord(evs[k], sts[k]));
        QBX_HIP(hipStreamWaitEvent(sts[0], evs[k], 0));
    }
    *done = true;
    return QBX_OK;
}

int lpot_matrix_2d(hipStream_t st, const qbx_kernel_desc *d, int nel, int n1, int nf,
                   const double *U, const double *src, const double *nrm, const double *w,
                   int nt, const double *tgt, const double *ctr, const double *rad, long ld,
                   double *mat)
{
    KD kd;
    QBX_TRY(make_kd(d, nullptr, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    QBX_CHECK(!kd.tder, "qbx_lpot_matrix: target derivatives are not supported");
    const dim3 grid(nt), blk(LM_THREADS);
#define LM_ARGS kd, nel, n1, nf, U, src, nrm, w, nt, tgt, ctr, rad, ld, mat
    if (!ctr) {
        hipLaunchKernelGGL(k_lpot_matrix_2d<-1>, grid, blk, 0, st, LM_ARGS);
    } else {
        switch (kd.pq) {
#define LM_CASE(PP) \
    case PP: hipLaunchKernelGGL(k_lpot_matrix_2d<PP>, grid, blk, 0, st, LM_ARGS); break;
            LM_CASE(0) LM_CASE(1) LM_CASE(2) LM_CASE(3) LM_CASE(4) LM_CASE(5) LM_CASE(6)
            LM_CASE(7) LM_CASE(8) LM_CASE(9) LM_CASE(10) LM_CASE(11) LM_CASE(12)
#undef LM_CASE
        default: set_error("qbx_order out of range"); return QBX_ERR_INVALID;
        }
    }
#undef LM_ARGS
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// The class grouping of k_m2l_cls for the levels `lev` of a traversal,
// built on the host at the first call and kept with the traversal (pairs
// sorted by (level, class, entry), so the items and sums are the same on
// every call).
static int m2l_groups(const qbx_tree *T, const qbx_trav *R, const KD &kd,
                      const std::vector<int> &lev)
{
    if (R->m2lg_nitems >= 0 && R->m2lg_levels == lev) return QBX_OK;
    const int nttp = R->nttp_boxes, n2 = R->n2, nb = T->nboxes;
    std::vector<int> starts(nttp + 1), lists(std::max(n2, 1)), ttp(std::max(nttp, 1)),
        coords(2L * nb);
    QBX_HIP(hipDeviceSynchronize());
    QBX_HIP(hipMemcpy(starts.data(), R->l2_starts.p, sizeof(int) * (nttp + 1),
                      hipMemcpyDeviceToHost));
    if (n2 > 0)
        QBX_HIP(hipMemcpy(lists.data(), R->l2_lists.p, sizeof(int) * n2, hipMemcpyDeviceToHost));
    if (nttp > 0)
        QBX_HIP(hipMemcpy(ttp.data(), R->ttp_boxes.p, sizeof(int) * nttp, hipMemcpyDeviceToHost));
    QBX_HIP(hipMemcpy(coords.data(), T->box_coords.p, sizeof(int) * 2L * nb,
                      hipMemcpyDeviceToHost));
    const int W = kd.m2l_w, S1 = 2 * W + 1;
    std::vector<std::pair<long, int>> kv;
    for (int l : lev)
        for (int i = R->h_level_start_ttp_box_nrs[l]; i < R->h_level_start_ttp_box_nrs[l + 1];
             ++i) {
            const int P = ttp[i];
            for (int e = starts[i]; e < starts[i + 1]; ++e) {
                const int A = lists[e];
                const int dx = coords[P] - coords[A], dy = coords[nb + P] - coords[nb + A];
                QBX_CHECK(std::abs(dx) <= W && std::abs(dy) <= W,
                          "M2L: list-2 offset (%d, %d) outside the class table", dx, dy);
                kv.push_back({(long)l * S1 * S1 + (dx + W) * S1 + (dy + W), e});
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
        const int l = (int)(kv[j].first / (S1 * S1)), cls = (int)(kv[j].first % (S1 * S1));
        for (int a = j; a < j1; a += M2LG_NP) {
            items.push_back(l);
            items.push_back(cls);
            items.push_back(a);
            items.push_back(std::min(M2LG_NP, j1 - a));
        }
        j = j1;
    }
    // per-pair row offsets (rows of the pair's level length, in sorted order);
    // items split into the one-wave (ncl <= 48) and four-wave lists
    std::vector<int> yoff(std::max(np, 1));
    long y = 0;
    for (int j = 0; j < np; ++j) {
        yoff[j] = (int)y;
        y += kd.ncl[kv[j].first / (S1 * S1)];
    }
    QBX_CHECK(y < (1L << 31), "M2L: %ld grouped rows overflow the row offsets", y);
    std::vector<int> small, big;
    for (size_t q = 0; q < items.size(); q += 4) {
        std::vector<int> &dst = kd.ncl[items[q]] <= M2LM_ROWS * 16 ? small : big;
        dst.insert(dst.end(), items.begin() + q, items.begin() + q + 4);
    }
    const int ns = (int)small.size() / 4, nbg = (int)big.size() / 4;
    small.insert(small.end(), big.begin(), big.end());
    QBX_TRY(R->m2lg_e.alloc(std::max(np, 1)));
    QBX_TRY(R->m2lg_pos.alloc(std::max(n2, 1)));
    QBX_TRY(R->m2lg_yoff.alloc(std::max(np, 1)));
    QBX_TRY(R->m2lg_items.alloc(std::max(4 * (ns + nbg), 4)));
    QBX_HIP(hipMemcpy(R->m2lg_e.p, eo.data(), sizeof(int) * eo.size(), hipMemcpyHostToDevice));
    QBX_HIP(hipMemcpy(R->m2lg_pos.p, pos.data(), sizeof(int) * pos.size(),
                      hipMemcpyHostToDevice));
    QBX_HIP(hipMemcpy(R->m2lg_yoff.p, yoff.data(), sizeof(int) * yoff.size(),
                      hipMemcpyHostToDevice));
    if (ns + nbg > 0)
        QBX_HIP(hipMemcpy(R->m2lg_items.p, small.data(), sizeof(int) * small.size(),
                          hipMemcpyHostToDevice));
    R->m2lg_nitems = ns;
    R->m2lg_nitems2 = nbg;
    R->m2lg_nrows = y;
    R->m2lg_levels = lev;
    return QBX_OK;
}

}  // namespace qbx

using namespace qbx;

extern "C" {

long qbx_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T)
{
    if (d && d->kernel == QBX_LAPLACE3D) return f3_expansion_size(d, T);
    if (d && d->kernel == QBX_HELMHOLTZ3D) return h3_expansion_size(d, T);
    KD kd;
    if (!T || make_kd(d, T, &kd) != QBX_OK) return -1;
    return kd.lcoef[kd.nlevels] * kd.nr;
}

int qbx_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets)
{
    QBX_CHECK(T && offsets, "qbx_expansion_level_offsets: null argument");
    if (d && d->kernel == QBX_LAPLACE3D) return f3_expansion_level_offsets(d, T, offsets);
    if (d && d->kernel == QBX_HELMHOLTZ3D) return h3_expansion_level_offsets(d, T, offsets);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    for (int l = 0; l <= kd.nlevels; ++l) offsets[l] = kd.lcoef[l] * kd.nr;
    return QBX_OK;
}

int qbx_ncoeff_qbx(const qbx_kernel_desc *d)
{
    if (d && d->kernel == QBX_LAPLACE3D) return f3_ncoeff_qbx(d);
    if (d && d->kernel == QBX_HELMHOLTZ3D) return h3_ncoeff_qbx(d);
    KD kd;
    if (make_kd(d, nullptr, &kd) != QBX_OK) return -1;
    return kd.ncq;
}

int qbx_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_kernel_desc *d, const double *str, const double *nrm,
                        double *mpole)
{
    QBX_CHECK(ctx && T && R && str && mpole, "qbx_form_multipoles: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_form_multipoles(ctx, stream, T, R, d, str, nrm, mpole);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_form_multipoles(ctx, stream, T, R, d, str, nrm, mpole);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    QBX_TRY(zero_fill(mpole, (exp_bytes(kd)) / 8, st));
    if (kd.kernel == QBX_HELMHOLTZ2D) {
        bool ok = true;
        for (int l = 0; l < T->nlevels; ++l) {
            P2EShape sh;
            int th;
            size_t by;
            ok = ok && p2e_launch(kd, kd.order[l], &sh, &th, &by);
        }
        if (ok) {
            const TD td = make_td(T);
            for (int l = 0; l < T->nlevels; ++l) p2m_level_h(st, kd, td, T, R, l, str, nrm, mpole);
            QBX_HIP(hipGetLastError());
            return QBX_OK;
        }
    }
    const Flat fl = make_flat(kd, R->h_level_start_source_box_nrs, true);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0) {
        const TD td = make_td(T);
#define L_P2M(N)                                                                           \
    hipLaunchKernelGGL(k_p2m<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, fl,      \
                       R->source_boxes.p, T->sources.p, T->nsources, str, nrm, mpole)
        FOR_RHS_CHUNKS(kd, NRC_S, L_P2M);
#undef L_P2M
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// M2M into the parent levels lmin .. nlevels - 2 (lmin = 0: every level, as
// the reference's coarsen_multipoles; the fused driver passes the coarsest
// level whose multipoles a later stage reads)
static int coarsen_multipoles_from(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                   const qbx_trav *R, const qbx_kernel_desc *d, double *mpole,
                                   int lmin)
{
    QBX_CHECK(ctx && T && R && mpole, "qbx_coarsen_multipoles: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_coarsen_multipoles(ctx, stream, T, R, d, mpole, lmin);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_coarsen_multipoles(ctx, stream, T, R, d, mpole, lmin);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    const TD td = make_td(T);
    bool mfma = kd.kernel == QBX_HELMHOLTZ2D && kd.nr >= 4;
    for (int l = 0; l < T->nlevels; ++l) mfma = mfma && 2 * kd.order[l] + 1 <= 16 * 4 * M2LM_RTW;
    for (int l = T->nlevels - 2; l >= std::max(lmin, 0); --l) {
        const int i0 = R->h_level_start_source_parent_box_nrs[l];
        const int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr == 0) continue;
        if (mfma) {
            for (int r0 = 0; r0 < kd.nr; r0 += 16) {
                kd.r0 = r0;
                kd.rc = std::min(16, kd.nr - r0);
                hipLaunchKernelGGL(k_shift_mfma<0>, dim3(i1 - i0), dim3(256), 0, st, kd, td, l,
                                   i0, R->source_parent_boxes.p, mpole);
            }
            continue;
        }
        HtLaunch hl;
        if (ht_launch(kd, kd.order[l], kd.order[l + 1], kd.order[l] + kd.order[l + 1], &hl)) {
        // levels with few parents (the coarse, high-order ones): a thread
        // group per staged child, so a block's four children run side by side
#define L_M2MH(N)                                                                          \
    do {                                                                                   \
        const int ns_ = (int)std::max(1L, std::min(4L, 65536L / (long)ht_bytes<N>(hl)));   \
        const int nout_ = 2 * kd.order[l] + 1;                                             \
        int ng_ = 1, th_ = hl.threads;                                                     \
        const size_t red_ = sizeof(cplx) * 3 * nout_ * (N);                                \
        if (i1 - i0 <= 16 && ns_ == 4 && 4 * nout_ <= 1024 &&                              \
            red_ <= (size_t)ns_ * ht_bytes<N>(hl)) {                                       \
            ng_ = 4;                                                                       \
            th_ = 64 * ((4 * nout_ + 63) / 64);                                            \
        }                                                                                  \
        hipLaunchKernelGGL(k_m2m_h<N>, dim3(i1 - i0), dim3(th_), ns_ * ht_bytes<N>(hl), st,  \
                           kd, td, l, i0, ns_, ng_, R->source_parent_boxes.p, mpole);       \
    } while (0)
            FOR_RHS_CHUNKS(kd, NRC_S, L_M2MH);
#undef L_M2MH
            continue;
        }
#define L_M2M(N)                                                                           \
    hipLaunchKernelGGL(k_m2m<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, B, l, i0, \
                       i1, R->source_parent_boxes.p, mpole)
        FOR_RHS_CHUNKS(kd, NRC_S, L_M2M);
#undef L_M2M
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                           const qbx_kernel_desc *d, double *mpole)
{
    return coarsen_multipoles_from(ctx, stream, T, R, d, mpole, 0);
}

static int eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                        const qbx_geo *G, const qbx_kernel_desc *d, int mode,
                        const double *str, const double *nrm, const double *expn, double *pot,
                        int accumulate)
{
    QBX_CHECK(ctx && T && R && G && pot, "stage: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_eval_targets(ctx, stream, T, R, G, d, mode, str, nrm, expn, pot, accumulate);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_eval_targets(ctx, stream, T, R, G, d, mode, str, nrm, expn, pot, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(pot, (sizeof(double) * 2 * (size_t)T->ntargets * kd.nr) / 8, st));
    const int ntb = R->ntarget_boxes;
    if (ntb == 0 || G->nfiltered == 0) return QBX_OK;
    const TD td = make_td(T);
    const int *starts = mode == 1 ? R->l3_starts.p : R->l1_starts.p;
    const int *lists = mode == 1 ? R->l3_lists.p : R->l1_lists.p;
    if (mode == 1) {
        const unit_plan &U = G->m2p_plan;
        QBX_CHECK(G->trav == R, "qbx_eval_multipoles: geometry was built for another traversal");
        if (U.nunits == 0) return QBX_OK;
#define M2P_LAUNCH_D(N, D)                                                                 \
    hipLaunchKernelGGL((k_m2p<N, D>), dim3(U.nunits), dim3(64), 0, st, kd, td,            \
                       U.unit_box.p, U.unit_t0.p, U.box_g.p, R->target_boxes.p, starts,   \
                       lists, G->nqbtl_box_target_starts.p,                              \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, expn, pot)
#define M2P_LAUNCH(N)              \
    if (kd.tder)                   \
        M2P_LAUNCH_D(N, true);     \
    else                           \
        M2P_LAUNCH_D(N, false)
        FOR_RHS_CHUNKS(kd, NRC_S, M2P_LAUNCH);
#undef M2P_LAUNCH
#undef M2P_LAUNCH_D
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    if (mode == 0) {
        const unit_plan &U = G->p2p_plan;
        QBX_CHECK(G->trav == R, "qbx_eval_direct: geometry was built for another traversal");
        if (U.nunits == 0) return QBX_OK;
#define P2P_LAUNCH_K(N, D, KN, DP)                                                         \
    hipLaunchKernelGGL((k_p2p<N, D, KN, DP>), dim3(U.nunits), dim3(64), 0, st, kd, td,    \
                       U.unit_box.p, U.unit_t0.p, U.box_g.p, R->target_boxes.p, starts,   \
                       lists, G->nqbtl_box_target_starts.p,                              \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm, pot)
#define P2P_LAUNCH_D(N, D)                                                                 \
    do {                                                                                   \
        if (kd.kernel == QBX_HELMHOLTZ2D) {                                                \
            if (kd.dip)                                                                    \
                P2P_LAUNCH_K(N, D, QBX_HELMHOLTZ2D, 1);                                    \
            else                                                                           \
                P2P_LAUNCH_K(N, D, QBX_HELMHOLTZ2D, 0);                                    \
        } else {                                                                           \
            if (kd.dip)                                                                    \
                P2P_LAUNCH_K(N, D, QBX_LAPLACE2D, 1);                                      \
            else                                                                           \
                P2P_LAUNCH_K(N, D, QBX_LAPLACE2D, 0);                                      \
        }                                                                                  \
    } while (0)
#define P2P_LAUNCH(N)              \
    if (kd.tder)                   \
        P2P_LAUNCH_D(N, true);     \
    else                           \
        P2P_LAUNCH_D(N, false)
        FOR_RHS_CHUNKS(kd, NRC_S, P2P_LAUNCH);
#undef P2P_LAUNCH
#undef P2P_LAUNCH_D
#undef P2P_LAUNCH_K
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define EVAL_LAUNCH(M, N, D)                                                               \
    hipLaunchKernelGGL((k_eval_targets<M, N, D>), dim3(ntb), dim3(64), 0, st, kd, td, ntb, \
                       R->target_boxes.p, starts, lists, G->nqbtl_box_target_starts.p,    \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, expn, pot)
#define EV2(N)                     \
    if (kd.tder)                   \
        EVAL_LAUNCH(2, N, true);   \
    else                           \
        EVAL_LAUNCH(2, N, false)
    FOR_RHS_CHUNKS(kd, NRC_S, EV2);
#undef EV2
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
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_multipole_to_local(ctx, stream, T, R, d, mpole, local);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_multipole_to_local(ctx, stream, T, R, d, mpole, local);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    QBX_TRY(zero_fill(local, (exp_bytes(kd)) / 8, st));
    if (R->n2 == 0) return QBX_OK;
    const TD td = make_td(T);
    // many right-hand sides (Helmholtz): on the matrix cores, 16 right-hand
    // sides per launch -- levels with 2p + 1 <= 48 splitting the entries over
    // the waves, coarser ones (<= 256) splitting the row tiles
    const bool mfma = kd.kernel == QBX_HELMHOLTZ2D && kd.nr >= 4;
    // one right-hand side (Helmholtz): the levels with 2p + 1 <= 48 grouped by
    // translation class on the matrix cores (not in the allocation-free
    // driver, whose workspace has no room for the per-pair rows)
    const bool grp = kd.kernel == QBX_HELMHOLTZ2D && kd.nr == 1 && !ctx->tab_m2l_ext;
    // path of level l: 0 vector kernel, 1 k_m2l_mfma, 2 k_m2l_mfma_rows,
    // 3 k_m2l_cls, 4 k_m2l_cls_rows
    auto path = [&](int l) {
        if (grp) return kd.ncl[l] <= M2LM_ROWS * 16 ? 3 : kd.ncl[l] <= 16 * 4 * M2LM_RTW ? 4 : 0;
        if (!mfma) return 0;
        if (kd.ncl[l] <= M2LM_ROWS * 16) return 1;
        return kd.ncl[l] <= 16 * 4 * M2LM_RTW ? 2 : 0;
    };
    auto box_flat = [&](int want, long *count) {
        Flat f = make_flat(kd, R->h_level_start_ttp_box_nrs, false);
        long n = 0;
        for (int l = 0; l < f.nl; ++l) {
            f.tstart[l] = n;
            f.nper[l] = 1;
            if (path(l) == want)
                n += R->h_level_start_ttp_box_nrs[l + 1] - R->h_level_start_ttp_box_nrs[l];
        }
        f.tstart[f.nl] = n;
        *count = n;
        return f;
    };
    if (mfma) {
        long nb1, nb2;
        const Flat f1 = box_flat(1, &nb1), f2 = box_flat(2, &nb2);
        for (int r0 = 0; r0 < kd.nr; r0 += 16) {
            kd.r0 = r0;
            kd.rc = std::min(16, kd.nr - r0);
            if (nb2 > 0)
                hipLaunchKernelGGL(k_m2l_mfma_rows, dim3(nb2), dim3(256), 0, st, kd, td, f2,
                                   R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local);
            if (nb1 > 0)
                hipLaunchKernelGGL(k_m2l_mfma, dim3(nb1), dim3(256), 0, st, kd, td, f1,
                                   R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local);
        }
    }
    if (grp) {
        std::vector<int> lev;
        unsigned mask = 0;
        for (int l = 0; l < kd.nlevels; ++l)
            if (path(l) >= 3) {
                lev.push_back(l);
                mask |= 1u << l;
            }
        QBX_TRY(m2l_groups(T, R, kd, lev));
        if (R->m2lg_nitems + R->m2lg_nitems2 > 0) {
            QBX_TRY(ws_reserve(ctx, st, ctx->ws_m2lg, 2L * R->m2lg_nrows,
                               "class-grouped M2L rows"));
            if (R->m2lg_nitems2 > 0)
                hipLaunchKernelGGL(k_m2l_cls_rows, dim3(R->m2lg_nitems2), dim3(256), 0, st, kd,
                                   td, R->m2lg_nitems2, R->m2lg_items.p + 4L * R->m2lg_nitems,
                                   R->m2lg_e.p, R->m2lg_yoff.p, R->l2_lists.p, mpole,
                                   ctx->ws_m2lg.p);
            if (R->m2lg_nitems > 0)
                hipLaunchKernelGGL(k_m2l_cls, dim3(R->m2lg_nitems), dim3(64), 0, st, kd, td,
                                   R->m2lg_nitems, R->m2lg_items.p, R->m2lg_e.p,
                                   R->m2lg_yoff.p, R->l2_lists.p, mpole, ctx->ws_m2lg.p);
            hipLaunchKernelGGL(k_m2l_cls_sum, dim3(R->nttp_boxes), dim3(64), 0, st, kd, td,
                               R->nttp_boxes, R->ttp_boxes.p, R->l2_starts.p, R->m2lg_pos.p,
                               R->m2lg_yoff.p, mask, ctx->ws_m2lg.p, local);
            QBX_HIP(hipGetLastError());
        }
    }
    // one block per target-or-target-parent box
    for (int l = 0; l < kd.nlevels; ++l)
        QBX_CHECK(kd.ncl[l] <= M2L_THREADS, "M2L: %d coefficients exceed one block", kd.ncl[l]);
    FlatX fl;
    memset(&fl, 0, sizeof(fl));
    fl.nl = kd.nlevels;
    long nb = 0;
    for (int l = 0; l < kd.nlevels; ++l) {
        fl.tstart[l] = nb;
        fl.istart[l] = R->h_level_start_ttp_box_nrs[l];
        fl.count[l] = path(l) == 0 ? R->h_level_start_ttp_box_nrs[l + 1] -
                                         R->h_level_start_ttp_box_nrs[l]
                                   : 0;
        nb += (fl.count[l] + 7) / 8 * 8;
    }
    fl.tstart[kd.nlevels] = nb;
    if (nb > 0) {
#define L_M2L(N)                                                                           \
    hipLaunchKernelGGL(k_m2l_blk<N>, dim3(nb), dim3(M2L_THREADS), 0, st, kd, td, B, fl,    \
                       R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local)
        FOR_RHS_CHUNKS(kd, NRC_S, L_M2L);
#undef L_M2L
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_kernel_desc *d, const double *str, const double *nrm,
                    double *local, int accumulate)
{
    QBX_CHECK(ctx && T && R && str && local, "qbx_form_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_form_locals(ctx, stream, T, R, d, str, nrm, local, accumulate);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_form_locals(ctx, stream, T, R, d, str, nrm, local, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate) QBX_TRY(zero_fill(local, (exp_bytes(kd)) / 8, st));
    if (kd.kernel == QBX_HELMHOLTZ2D && R->n4 > 0) {
        bool ok = true;
        for (int l = 0; l < T->nlevels; ++l) {
            P2EShape sh;
            int th;
            size_t by;
            ok = ok && p2e_launch(kd, kd.order[l], &sh, &th, &by);
        }
        if (ok) {
            const TD td = make_td(T);
            for (int l = 0; l < T->nlevels; ++l) {
                const int i0 = R->h_level_start_ttp_box_nrs[l];
                const int i1 = R->h_level_start_ttp_box_nrs[l + 1];
                if (i1 == i0) continue;
                P2EShape sh;
                int th;
                size_t by;
                p2e_launch(kd, kd.order[l], &sh, &th, &by);
#define L_P2LH(N)                                                                          \
    hipLaunchKernelGGL((k_p2e_h<N, true>), dim3(i1 - i0), dim3(th), by, st, kd, td, l, i0, sh, \
                       R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p, T->sources.p,         \
                       T->nsources, str, nrm, local)
                FOR_RHS_CHUNKS(kd, NRC_S, L_P2LH);
#undef L_P2LH
            }
            QBX_HIP(hipGetLastError());
            return QBX_OK;
        }
    }
    const Flat fl = make_flat(kd, R->h_level_start_ttp_box_nrs, true);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0 && R->n4 > 0) {
        const TD td = make_td(T);
#define L_P2L(N)                                                                           \
    hipLaunchKernelGGL(k_p2l<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, fl,      \
                       R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p, T->sources.p,        \
                       T->nsources, str, nrm, local)
        FOR_RHS_CHUNKS(kd, NRC_S, L_P2L);
#undef L_P2L
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// L2L from the parent levels lmin .. nlevels - 2 (lmin = 0: every level, as
// the reference's refine_locals; the fused driver passes the coarsest level
// with a nonzero local, above which the shifts would add zeros)
static int refine_locals_from(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                              const qbx_kernel_desc *d, double *local, int lmin)
{
    QBX_CHECK(ctx && T && R && local, "qbx_refine_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_refine_locals(ctx, stream, T, R, d, local, lmin);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_refine_locals(ctx, stream, T, R, d, local, lmin);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    QBX_TRY(ensure_tables(ctx, st, T, &kd));
    const TD td = make_td(T);
    bool mfma = kd.kernel == QBX_HELMHOLTZ2D && kd.nr >= 4;
    for (int l = 0; l < T->nlevels; ++l) mfma = mfma && 2 * kd.order[l] + 1 <= 16 * 4 * M2LM_RTW;
    for (int l = std::max(lmin, 0) + 1; l < T->nlevels; ++l) {
        const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr == 0) continue;
        if (mfma) {
            for (int r0 = 0; r0 < kd.nr; r0 += 16) {
                kd.r0 = r0;
                kd.rc = std::min(16, kd.nr - r0);
                hipLaunchKernelGGL(k_shift_mfma<1>, dim3(i1 - i0), dim3(256), 0, st, kd, td, l,
                                   i0, R->ttp_boxes.p, local);
            }
            continue;
        }
        HtLaunch hl;
        if (ht_launch(kd, kd.order[l], kd.order[l - 1], kd.order[l] + kd.order[l - 1], &hl)) {
#define L_L2LH(N)                                                                          \
    hipLaunchKernelGGL(k_l2l_h<N>, dim3(i1 - i0), dim3(hl.threads), ht_bytes<N>(hl), st, kd, \
                       td, l, i0, R->ttp_boxes.p, local)
            FOR_RHS_CHUNKS(kd, NRC_S, L_L2LH);
#undef L_L2LH
            continue;
        }
#define L_L2L(N)                                                                           \
    hipLaunchKernelGGL(k_l2l<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, B, l, i0, \
                       i1, R->ttp_boxes.p, local)
        FOR_RHS_CHUNKS(kd, NRC_S, L_L2L);
#undef L_L2L
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                      const qbx_kernel_desc *d, double *local)
{
    return refine_locals_from(ctx, stream, T, R, d, local, 0);
}

int qbx_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                               const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                               const double *nrm, double *qexp, int accumulate)
{
    QBX_CHECK(ctx && T && R && G && str && qexp, "qbx_form_global_qbx_locals: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_form_global_qbx_locals(ctx, stream, T, R, G, d, str, nrm, qexp, accumulate);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_form_global_qbx_locals(ctx, stream, T, R, G, d, str, nrm, qexp, accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, (sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq * kd.nr) / 8, st));
    const unit_plan &U = G->p2q_plan;
    QBX_CHECK(G->trav == R, "P2QBXL: geometry was built for another traversal");
    if (G->nglobal == 0 || U.nunits == 0) return QBX_OK;
    const TD td = make_td(T);
    const unit_plan &V = G->gc_group_plan;
    if (kd.kernel == QBX_HELMHOLTZ2D && kd.nr >= 4 && kd.pq <= 5 && V.nunits > 0) {
        for (int r0 = 0; r0 < kd.nr; r0 += 16) {
            kd.r0 = r0;
            kd.rc = std::min(16, kd.nr - r0);
#define L_P2QM(PQV)                                                                        \
    hipLaunchKernelGGL(k_p2qbxl_mfma<PQV>, dim3(V.nunits), dim3(256), 0, st, kd, td,       \
                       V.unit_box.p, V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,  \
                       G->centers.p, G->radii.p, G->ncenters, R->l1_starts.p,               \
                       R->l1_lists.p, T->sources.p, T->nsources, str, nrm, qexp)
            switch (kd.pq) {
            case 0: L_P2QM(0); break;
            case 1: L_P2QM(1); break;
            case 2: L_P2QM(2); break;
            case 3: L_P2QM(3); break;
            case 4: L_P2QM(4); break;
            default: L_P2QM(5); break;
            }
#undef L_P2QM
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    if (kd.kernel == QBX_HELMHOLTZ2D && kd.nr == 1) {
        // one Helmholtz density: the kernel and source kind fixed at compile time
#define L_P2QH(PQV, DV)                                                                    \
    hipLaunchKernelGGL((k_p2qbxl<PQV, 1, QBX_HELMHOLTZ2D, DV>), dim3(U.nunits), dim3(64), 0, \
                       st, kd, td, U.unit_box.p, U.unit_t0.p, U.box_g.p,                   \
                       G->gc_by_tbox_starts.p, G->gc_by_tbox.p, G->centers.p, G->radii.p,  \
                       G->ncenters, R->l1_starts.p, R->l1_lists.p, T->sources.p,           \
                       T->nsources, str, nrm, qexp, accumulate)
#define L_P2QHD(PQV)          \
    if (kd.dip)               \
        L_P2QH(PQV, 1);       \
    else                      \
        L_P2QH(PQV, 0)
        kd.r0 = 0;
        kd.rc = 1;
        switch (kd.pq) {
        case 0: L_P2QHD(0); break;
        case 1: L_P2QHD(1); break;
        case 2: L_P2QHD(2); break;
        case 3: L_P2QHD(3); break;
        case 4: L_P2QHD(4); break;
        case 5: L_P2QHD(5); break;
        case 6: L_P2QHD(6); break;
        case 7: L_P2QHD(7); break;
        case 8: L_P2QHD(8); break;
        case 9: L_P2QHD(9); break;
        case 10: L_P2QHD(10); break;
        case 11: L_P2QHD(11); break;
        case 12: L_P2QHD(12); break;
        default: set_error("qbx_order out of range"); return QBX_ERR_INVALID;
        }
#undef L_P2QHD
#undef L_P2QH
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define L_P2Q(N)                                                                           \
    DISPATCH_PQ_NR(kd.pq, N, k_p2qbxl, dim3(U.nunits), dim3(64), 0, st, kd, td,            \
                   U.unit_box.p, U.unit_t0.p, U.box_g.p, G->gc_by_tbox_starts.p,            \
                   G->gc_by_tbox.p, G->centers.p, G->radii.p, G->ncenters, R->l1_starts.p,  \
                   R->l1_lists.p, T->sources.p, T->nsources, str, nrm, qexp, accumulate)
    FOR_RHS_CHUNKS(kd, NRC_Q, L_P2Q);
#undef L_P2Q
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
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_translate_box_multipoles_to_qbx_local(ctx, stream, T, R, G, d, mpole, qexp,
                                                        accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, (sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq * kd.nr) / 8, st));
    const unit_plan &U = G->m2q_plan;
    QBX_CHECK(G->trav == R, "M2QBXL: geometry was built for another traversal");
    if (G->nglobal == 0 || R->n3 == 0 || U.nunits == 0) return QBX_OK;
    const TD td = make_td(T);
    const unit_plan &V = G->gc_group_plan;
    // (also for one right-hand side: the Graf sequence built once per (center,
    // box) pair for four centers beats the per-thread form even with 15 of 16
    // columns idle)
    if (kd.kernel == QBX_HELMHOLTZ2D && kd.pq <= 5 && V.nunits > 0) {
        for (int r0 = 0; r0 < kd.nr; r0 += 16) {
            kd.r0 = r0;
            kd.rc = std::min(16, kd.nr - r0);
#define L_M2QM(PQV)                                                                        \
    hipLaunchKernelGGL(k_m2qbxl_mfma<PQV>, dim3(V.nunits), dim3(256), 0, st, kd, td,       \
                       V.unit_box.p, V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,  \
                       G->centers.p, G->ncenters, R->l3_starts.p, R->l3_lists.p, mpole, qexp)
            switch (kd.pq) {
            case 0: L_M2QM(0); break;
            case 1: L_M2QM(1); break;
            case 2: L_M2QM(2); break;
            case 3: L_M2QM(3); break;
            case 4: L_M2QM(4); break;
            default: L_M2QM(5); break;
            }
#undef L_M2QM
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define L_M2Q(N)                                                                           \
    DISPATCH_PQ_NR(kd.pq, N, k_m2qbxl, dim3(U.nunits), dim3(64), 0, st, kd, td, B,         \
                   U.unit_box.p, U.unit_t0.p, U.box_g.p, G->gc_by_tbox_starts.p,            \
                   G->gc_by_tbox.p, G->centers.p, G->radii.p, G->ncenters, R->l3_starts.p,  \
                   R->l3_lists.p, mpole, qexp)
    FOR_RHS_CHUNKS(kd, NRC_Q, L_M2Q);
#undef L_M2Q
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
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_translate_box_local_to_qbx_local(ctx, stream, T, R, G, d, local, qexp,
                                                   accumulate);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    const double *B;
    QBX_TRY(binom_table(&B));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, (sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq * kd.nr) / 8, st));
    const int ng = G->nglobal;
    if (ng == 0) return QBX_OK;
    const TD td = make_td(T);
    const unit_plan &V = G->gc_group_plan;
    if (kd.kernel == QBX_HELMHOLTZ2D && kd.nr >= 4 && kd.pq <= 5 && V.nunits > 0) {
        for (int r0 = 0; r0 < kd.nr; r0 += 16) {
            kd.r0 = r0;
            kd.rc = std::min(16, kd.nr - r0);
#define L_L2QM(PQV)                                                                        \
    hipLaunchKernelGGL(k_l2qbxl_mfma<PQV>, dim3(V.nunits), dim3(256), 0, st, kd, td,       \
                       V.unit_box.p, V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,  \
                       R->target_boxes.p, G->centers.p, G->ncenters, local, qexp)
            switch (kd.pq) {
            case 0: L_L2QM(0); break;
            case 1: L_L2QM(1); break;
            case 2: L_L2QM(2); break;
            case 3: L_L2QM(3); break;
            case 4: L_L2QM(4); break;
            default: L_L2QM(5); break;
            }
#undef L_L2QM
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    if (kd.kernel == QBX_HELMHOLTZ2D && kd.nr == 1 && kd.pq <= 12) {
        // lanes per center: the tuning knob (every choice gives the same sums)
        const int lanes = ctx->tune[QBX_TUNE_L2QBXL_LANES];
#define L_L2QS_LS(PQV, LSV)                                                                \
    hipLaunchKernelGGL((k_l2qbxl_split<PQV, LSV>), dim3(cdiv((long)ng * LSV, 256)),        \
                       dim3(256), 0, st, kd, td, ng, G->global_qbx_centers.p,               \
                       G->center_to_target_box.p, R->target_boxes.p, G->centers.p,          \
                       G->ncenters, local, qexp)
#define L_L2QS(PQV)                                        \
    do {                                                   \
        if (lanes == 1) L_L2QS_LS(PQV, 1);                 \
        else if (lanes == 2) L_L2QS_LS(PQV, 2);            \
        else if (lanes == 8) L_L2QS_LS(PQV, 8);            \
        else L_L2QS_LS(PQV, 4);                            \
    } while (0)
        switch (kd.pq) {
        case 0: L_L2QS(0); break;
        case 1: L_L2QS(1); break;
        case 2: L_L2QS(2); break;
        case 3: L_L2QS(3); break;
        case 4: L_L2QS(4); break;
        case 5: L_L2QS(5); break;
        case 6: L_L2QS(6); break;
        case 7: L_L2QS(7); break;
        case 8: L_L2QS(8); break;
        case 9: L_L2QS(9); break;
        case 10: L_L2QS(10); break;
        case 11: L_L2QS(11); break;
        default: L_L2QS(12); break;
        }
#undef L_L2QS_LS
#undef L_L2QS
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define L_L2Q(N)                                                                           \
    DISPATCH_PQ_NR(kd.pq, N, k_l2qbxl, dim3(cdiv(ng, 64)), dim3(64), 0, st, kd, td, B, ng,  \
                   G->global_qbx_centers.p, G->center_to_target_box.p, R->target_boxes.p,  \
                   G->centers.p, G->radii.p, G->ncenters, local, qexp, 1)
    FOR_RHS_CHUNKS(kd, NRC_Q, L_L2Q);
#undef L_L2Q
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                            const qbx_kernel_desc *d, const double *qexp, double *pot)
{
    QBX_CHECK(ctx && T && G && qexp && pot, "qbx_eval_qbx_expansions: null argument");
    if (d && d->kernel == QBX_LAPLACE3D)
        return f3_eval_qbx_expansions(ctx, stream, T, G, d, qexp, pot);
    if (d && d->kernel == QBX_HELMHOLTZ3D)
        return h3_eval_qbx_expansions(ctx, stream, T, G, d, qexp, pot);
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (G->nctt > 0) {
#define L_Q2P_D(N, D)                                                                      \
    hipLaunchKernelGGL((k_qbxl2p<N, D>), dim3(cdiv(G->nctt, 64)), dim3(64), 0, st, kd,     \
                       G->nctt, G->ctt_lists.p, G->tree_ttc.p, T->targets.p, T->ntargets,  \
                       G->centers.p, G->radii.p, G->ncenters, qexp, pot)
#define L_Q2P(N)                   \
    if (kd.tder)                   \
        L_Q2P_D(N, true);          \
    else                           \
        L_Q2P_D(N, false)
        FOR_RHS_CHUNKS(kd, NRC_S, L_Q2P);
#undef L_Q2P
#undef L_Q2P_D
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_direct_qbx(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                   const qbx_kernel_desc *d, const double *str, const double *nrm, double *pot)
{
    QBX_CHECK(ctx && T && G && d && str && pot, "qbx_direct_qbx: null argument");
    if (d->kernel == QBX_LAPLACE3D || d->kernel == QBX_HELMHOLTZ3D) {
        QBX_CHECK(d->nrhs <= 1, "qbx_direct_qbx: 3D evaluates one density at a time");
        QBX_CHECK(!d->target_deriv, "qbx_direct_qbx: 3D target derivatives are not supported");
        return direct_qbx3d(ctx, S(stream), T, G, d->kernel, d->source_kind, d->k, d->qbx_order,
                            str, nrm, pot);
    }
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    const int nr = kd.nr;
    QBX_TRY(zero_fill(pot, 2L * T->ntargets * nr, st));
    if (G->nfiltered > 0) {
#define L_DP2P_D(N, D)                                                                     \
    hipLaunchKernelGGL((k_direct_p2p<N, D>), dim3(cdiv(G->nfiltered, 64)), dim3(64), 0, st, \
                       kd, G->nfiltered, G->nqbtl_unfiltered.p, T->targets.p, T->ntargets, \
                       T->sources.p, T->nsources, str, nrm, pot)
#define L_DP2P(N)                  \
    if (kd.tder)                   \
        L_DP2P_D(N, true);         \
    else                           \
        L_DP2P_D(N, false)
        FOR_RHS_CHUNKS(kd, NRC_S, L_DP2P);
#undef L_DP2P
#undef L_DP2P_D
    }
    if (G->nglobal > 0) {
        const long nq = 2L * std::max(G->ncenters, 1) * kd.ncq * nr;
        QBX_TRY(ws_reserve(ctx, st, ctx->ws_qbx, nq, "QBX expansion workspace"));
        double *Q = ctx->ws_qbx.p;
        QBX_TRY(zero_fill(Q, nq, st));
#define L_DP2Q(N)                                                                          \
    DISPATCH_PQ_NR(kd.pq, N, k_direct_p2qbxl, dim3(cdiv(G->nglobal, 4)), dim3(64), 0, st,  \
                   kd, G->nglobal, G->global_qbx_centers.p, G->centers.p, G->radii.p,      \
                   G->ncenters, T->sources.p, T->nsources, str, nrm, Q)
        FOR_RHS_CHUNKS(kd, NRC_Q, L_DP2Q);
#undef L_DP2Q
        QBX_TRY(qbx_eval_qbx_expansions(ctx, stream, T, G, d, Q, pot));
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_p2p_all(qbx_ctx *ctx, void *stream, const qbx_kernel_desc *d, int nsources,
                const double *sources, const double *normals, const double *str, int ntargets,
                const double *targets, double *pot)
{
    QBX_CHECK(ctx && d && pot, "qbx_p2p_all: null argument");
    QBX_CHECK(nsources >= 0 && ntargets >= 0, "qbx_p2p_all: negative size");
    QBX_CHECK(nsources == 0 || (sources && str), "qbx_p2p_all: null sources");
    QBX_CHECK(ntargets == 0 || targets, "qbx_p2p_all: null targets");
    hipStream_t st = S(stream);
    if (d->kernel == QBX_LAPLACE3D || d->kernel == QBX_HELMHOLTZ3D) {
        QBX_CHECK(d->nrhs <= 1, "qbx_p2p_all: 3D kernels take one density at a time");
        QBX_CHECK(d->source_kind == QBX_SOURCE_S || d->source_kind == QBX_SOURCE_D,
                  "unsupported source kind %d", d->source_kind);
        QBX_CHECK(d->source_kind == QBX_SOURCE_S || nsources == 0 || normals,
                  "double layer needs normals");
        if (ntargets == 0) return QBX_OK;
        return direct_p2p3d(st, d->kernel == QBX_HELMHOLTZ3D ? d->k : 0.0,
                            d->source_kind == QBX_SOURCE_D, ntargets, nullptr, targets, ntargets,
                            sources, nsources, str, normals, pot);
    }
    KD kd;
    QBX_TRY(make_kd(d, nullptr, &kd));
    QBX_CHECK(!kd.dip || nsources == 0 || normals, "double layer needs normals");
    QBX_CHECK(!kd.tder, "qbx_p2p_all: target derivatives are not supported");
    if (ntargets == 0) return QBX_OK;
    // every target is written (zero when there are no sources)
#define L_AP2P(N)                                                                          \
    hipLaunchKernelGGL(k_direct_p2p<N>, dim3(cdiv(ntargets, 64)), dim3(64), 0, st, kd,     \
                       ntargets, nullptr, targets, ntargets, sources, nsources, str,      \
                       normals, pot)
    FOR_RHS_CHUNKS(kd, NRC_S, L_AP2P);
#undef L_AP2P
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int qbx_lpot_matrix(qbx_ctx *ctx, void *stream, const qbx_kernel_desc *d, int nelements,
                    int n1, int nf, const double *upsample, const double *sources,
                    const double *normals, const double *weights, int ntargets,
                    const double *targets, const double *centers, const double *radii, long ld,
                    double *mat)
{
    QBX_CHECK(ctx && d && mat, "qbx_lpot_matrix: null argument");
    QBX_CHECK(!d->target_deriv, "qbx_lpot_matrix: target derivatives are not supported");
    QBX_CHECK(nelements >= 0 && ntargets >= 0, "qbx_lpot_matrix: negative size");
    QBX_CHECK(n1 >= 1 && nf >= 1 && nf <= LM_THREADS && (long)nf * n1 <= LM_UMAX,
              "qbx_lpot_matrix: element sizes n1=%d nf=%d outside the supported range", n1, nf);
    QBX_CHECK(upsample || nf == n1, "qbx_lpot_matrix: no upsampling matrix needs nf == n1");
    QBX_CHECK(ld >= (long)nelements * n1, "qbx_lpot_matrix: ld below the column count");
    QBX_CHECK(nelements == 0 || (sources && weights), "qbx_lpot_matrix: null sources");
    QBX_CHECK(ntargets == 0 || targets, "qbx_lpot_matrix: null targets");
    QBX_CHECK(!centers || radii, "qbx_lpot_matrix: centers need radii");
    if (ntargets == 0 || nelements == 0) return QBX_OK;
    hipStream_t st = S(stream);
    if (d->kernel == QBX_LAPLACE3D || d->kernel == QBX_HELMHOLTZ3D)
        return lpot_matrix_3d(st, d, nelements, n1, nf, upsample, sources, normals, weights,
                              ntargets, targets, centers, radii, ld, mat);
    return lpot_matrix_2d(st, d, nelements, n1, nf, upsample, sources, normals, weights,
                          ntargets, targets, centers, radii, ld, mat);
}

int qbx_finalize_potentials(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                            int nrhs, const double *pot, double *out)
{
    QBX_CHECK(ctx && T && G && pot && out, "qbx_finalize_potentials: null argument");
    QBX_CHECK(nrhs >= 1, "qbx_finalize_potentials: nrhs must be >= 1");
    const int nt = T->ntargets, nc = G->ncenters;
    if (nt > nc)
        hipLaunchKernelGGL(k_finalize, dim3(cdiv((long)(nt - nc) * nrhs, 256)), dim3(256), 0,
                           S(stream), nt, nc, nrhs, G->user_to_tree_target.p, pot, out);
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // extern "C"

// The allocation-free driver's workspace: one array per device, loaded with
// the code object and kept until the process ends.  Host code never uses the
// name as an address (on the host it is only a handle for the runtime):
// noalloc_scratch_address looks up each device's own copy once, with that
// device current, and the driver passes that address to its kernels.  Every
// qbx_drive_fmm_noalloc call on a device shares the array, so calls that
// overlap on one device race.
static __device__ __align__(256) double scratch[QBX_NOALLOC_SCRATCH_BYTES / 8];

namespace qbx {
int noalloc_scratch_address(int device, double **out)
{
    static std::mutex mu;
    static double *addr[64] = {nullptr};
    QBX_CHECK(device >= 0 && device < 64, "device index %d out of range", device);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        QBX_HIP(hipGetDevice(&cur));
        QBX_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        QBX_HIP(hipSetDevice(cur));
        QBX_HIP(e);
        QBX_CHECK(p, "hipGetSymbolAddress returned a null scratch address");
        addr[device] = (double *)p;
    }
    *out = addr[device];
    return QBX_OK;
}
}  // namespace qbx

// Workspace of one fused evaluation: multipoles M, locals L, the list-4
// locals L4, QBX expansions Q and tree-order potentials P.
struct DriveBufs {
    double *M, *L, *L4, *Q, *P;
    int *allreal = nullptr;  // 3D Laplace: the all-strengths-real flag
};

struct DriveSizes {
    long nexp, nm, nq, np;
};

static int drive_sizes(const qbx_tree *T, const qbx_geo *G, const qbx_kernel_desc *d,
                       DriveSizes *z)
{
    z->nexp = qbx_expansion_size(d, T);
    const int ncq = qbx_ncoeff_qbx(d);
    if (z->nexp < 0 || ncq < 0) return QBX_ERR_INVALID;
    const int nr = d->nrhs > 0 ? d->nrhs : 1;
    z->nm = 2L * std::max(z->nexp, 1L);
    z->nq = 2L * std::max(G->ncenters, 1) * ncq * nr;
    z->np = 2L * std::max(T->ntargets, 1) * nr;
    return QBX_OK;
}

// The stage DAG of qbx_drive_fmm.  multistream: on the caller's stream plus
// the device's three side streams joined by this context's events (below);
// otherwise every stage on the caller's stream in the order of drive_fmm.
// G = 1/(4 pi r) of the 3D potentials, applied to the target-specific sum
constexpr double INV4PI_3D = 0.079577471545947667884;

static int drive_impl(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                      const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                      const double *nrm, double *out_user, const DriveSizes &z,
                      const DriveBufs &b, bool multistream)
{
    const int nr = d->nrhs > 0 ? d->nrhs : 1;
    const long nexp = z.nexp, np = z.np;
    double *M = b.M, *L = b.L, *Q = b.Q, *P = b.P, *L4 = b.L4;
    hipStream_t st = S(stream);
    // 3D Laplace with the target-specific sum: P2QBXL leaves the expansions
    // at zero (M2QBXL and L2QBXL add into them) and the list-1 part of the
    // QBX targets is added after QBXL2P
    const bool tsq = d->target_specific_qbx != 0;
    // 3D Laplace: whether every strength is real (the M2L then forms half
    // the local coefficients and mirrors the rest)
    if (b.allreal) QBX_TRY(strengths_all_real(str, (long)T->nsources * nr, b.allreal, st));
    auto m2l = [&](hipStream_t s) {
        if (d->kernel == QBX_LAPLACE3D)
            return f3_multipole_to_local(ctx, s, T, R, d, M, L, b.allreal);
        return qbx_multipole_to_local(ctx, s, T, R, d, M, L);
    };
    auto m2m = [&](hipStream_t s) {
        if (d->kernel == QBX_LAPLACE3D)
            return f3_coarsen_multipoles(ctx, s, T, R, d, M, R->mpole_min_level, b.allreal);
        if (d->kernel == QBX_HELMHOLTZ3D)
            return h3_coarsen_multipoles(ctx, s, T, R, d, M, R->mpole_min_level);
        return coarsen_multipoles_from(ctx, s, T, R, d, M, R->mpole_min_level);
    };
    auto l2l = [&](hipStream_t s) {
        if (d->kernel == QBX_LAPLACE3D)
            return f3_refine_locals(ctx, s, T, R, d, L, R->local_min_level, b.allreal);
        if (d->kernel == QBX_HELMHOLTZ3D)
            return h3_refine_locals(ctx, s, T, R, d, L, R->local_min_level);
        return refine_locals_from(ctx, s, T, R, d, L, R->local_min_level);
    };
    if (!multistream) {
        QBX_TRY(zero_fill(P, np, st));
        QBX_TRY(qbx_form_multipoles(ctx, st, T, R, d, str, nrm, M));
        QBX_TRY(qbx_eval_direct(ctx, st, T, R, G, d, str, nrm, P, 1));
        if (tsq)
            QBX_TRY(zero_fill(Q, z.nq, st));
        else
            QBX_TRY(qbx_form_global_qbx_locals(ctx, st, T, R, G, d, str, nrm, Q, 0));
        QBX_TRY(qbx_form_locals(ctx, st, T, R, d, str, nrm, L4, 0));
        QBX_TRY(m2m(st));
        QBX_TRY(m2l(st));
        QBX_TRY(qbx_eval_multipoles(ctx, st, T, R, G, d, M, P, 1));
        QBX_TRY(qbx_translate_box_multipoles_to_qbx_local(ctx, st, T, R, G, d, M, Q, 1));
        QBX_TRY(add_into(L, L4, nexp * 2, st));
        QBX_TRY(l2l(st));
        QBX_TRY(qbx_eval_locals(ctx, st, T, R, G, d, L, P, 1));
        QBX_TRY(qbx_translate_box_local_to_qbx_local(ctx, st, T, R, G, d, L, Q, 1));
        QBX_TRY(qbx_eval_qbx_expansions(ctx, st, T, G, d, Q, P));
        if (tsq)
            QBX_TRY(tsqbx_nrhs(ctx, st, T, R, G, d->kernel,
                               d->source_kind == QBX_SOURCE_D ? QBX_TSQBX_D : QBX_TSQBX_S,
                               d->qbx_order, nr, str, nrm, INV4PI_3D, P, 1));
        QBX_TRY(qbx_finalize_potentials(ctx, stream, T, G, nr, P, out_user));
        return QBX_OK;
    }
    if (!ctx->side[0]) {
        QBX_TRY(side_streams(ctx->device, ctx->side));
        for (hipEvent_t &e : ctx->ev) QBX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // The stage DAG on four streams (the stage order of drive_fmm is kept
    // within each stream; every array has one writer stream at a time):
    //   main: zero P, M; P2M levels -(join)-> M2M -> M2L -[eM]-> (wait eP) L += L4
    //         -> L2L -[eL]-> (wait e2) L2QBXL, QBXL2P -> (wait e1) finalize
    //   s3:   P2M levels, P2L into L4 (its own local buffer) -[eP]->
    //   s1:   P2M levels, P2P, wait eM: M2P, wait eL: L2P -[e1]->  (non-QBX
    //         targets of P)
    //   s2:   P2M levels, P2QBXL, wait eM: M2QBXL -[e2]->  (Q)
    // (QBX targets of P are written on main: L2QBXL and QBXL2P follow L2L
    // there with no event hop)
    // (the side streams start after the zero fills; without the multistream
    // P2M they wait for the whole P2M instead)
    // so the latency-bound level-by-level translations overlap the near field;
    // P2M, the head of the chain, is issued before the near field (measured
    // 4% faster than forking first), its independent per-level launches
    // spread over the four streams, and each side stream goes on to its
    // near-field stage when its own P2M levels are done (3% faster than
    // waiting for the whole P2M: the join and fork cost two event hops).  M2P and M2QBXL, which need only
    // M2M, start after M2L: M2L then has the chip to itself and they fill the
    // idle slots of the L2L launches (measured 4% faster than starting them
    // beside M2L).
    hipStream_t s1 = ctx->side[0], s2 = ctx->side[1], s3 = ctx->side[2];
    hipEvent_t e0 = ctx->ev[0], eM = ctx->ev[1], eL = ctx->ev[2], e1 = ctx->ev[3],
               e2 = ctx->ev[4], eP = ctx->ev[5];
    QBX_TRY(zero_fill(P, np, st));
    bool done;
    {
        hipStream_t sts[4] = {st, s1, s2, s3};
        hipEvent_t evs[4] = {e0, e1, e2, eP};
        QBX_TRY(p2m_multistream(ctx, sts, 4, evs, T, R, d, str, nrm, M, &done));
        if (!done) QBX_TRY(qbx_form_multipoles(ctx, st, T, R, d, str, nrm, M));
    }
    // after the multistream P2M each side stream has already waited for the
    // zero fills and runs its near-field stage (which reads no multipole)
    // right after its own P2M levels; only M2M waits for the join
    if (!done) {
        QBX_HIP(hipEventRecord(e0, st));
        QBX_HIP(hipStreamWaitEvent(s1, e0, 0));
        QBX_HIP(hipStreamWaitEvent(s2, e0, 0));
        QBX_HIP(hipStreamWaitEvent(s3, e0, 0));
    }
    QBX_TRY(qbx_eval_direct(ctx, s1, T, R, G, d, str, nrm, P, 1));
    if (tsq)
        QBX_TRY(zero_fill(Q, z.nq, s2));
    else
        QBX_TRY(qbx_form_global_qbx_locals(ctx, s2, T, R, G, d, str, nrm, Q, 0));
    QBX_TRY(qbx_form_locals(ctx, s3, T, R, d, str, nrm, L4, 0));
    QBX_HIP(hipEventRecord(eP, s3));
    QBX_TRY(m2m(st));
    QBX_TRY(m2l(st));
    QBX_HIP(hipEventRecord(eM, st));
    QBX_HIP(hipStreamWaitEvent(s1, eM, 0));
    QBX_HIP(hipStreamWaitEvent(s2, eM, 0));
    QBX_TRY(qbx_eval_multipoles(ctx, s1, T, R, G, d, M, P, 1));
    QBX_TRY(qbx_translate_box_multipoles_to_qbx_local(ctx, s2, T, R, G, d, M, Q, 1));
    QBX_HIP(hipEventRecord(e2, s2));
    QBX_HIP(hipStreamWaitEvent(st, eP, 0));
    QBX_TRY(add_into(L, L4, nexp * 2, st));
    QBX_TRY(l2l(st));
    QBX_HIP(hipEventRecord(eL, st));
    QBX_HIP(hipStreamWaitEvent(s1, eL, 0));
    QBX_TRY(qbx_eval_locals(ctx, s1, T, R, G, d, L, P, 1));
    QBX_HIP(hipEventRecord(e1, s1));
    QBX_HIP(hipStreamWaitEvent(st, e2, 0));
    QBX_TRY(qbx_translate_box_local_to_qbx_local(ctx, st, T, R, G, d, L, Q, 1));
    QBX_TRY(qbx_eval_qbx_expansions(ctx, st, T, G, d, Q, P));
    if (tsq)
        QBX_TRY(tsqbx_nrhs(ctx, st, T, R, G, d->kernel,
                           d->source_kind == QBX_SOURCE_D ? QBX_TSQBX_D : QBX_TSQBX_S,
                           d->qbx_order, nr, str, nrm, INV4PI_3D, P, 1));
    QBX_HIP(hipStreamWaitEvent(st, e1, 0));
    QBX_TRY(qbx_finalize_potentials(ctx, stream, T, G, nr, P, out_user));
    return QBX_OK;
}


extern "C" {

int qbx_drive_fmm(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                  const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                  const double *nrm, double *out_user)
{
    QBX_CHECK(ctx && T && R && G && d && str && out_user, "qbx_drive_fmm: null argument");
    QBX_CHECK(!d->target_specific_qbx || d->kernel == QBX_LAPLACE3D,
              "qbx_drive_fmm: the target-specific sum is for 3D Laplace");
    DriveSizes z;
    QBX_TRY(drive_sizes(T, G, d, &z));
    QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_mpole, z.nm, "multipole workspace"));
    QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_local, z.nm, "local workspace"));
    QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_local2, z.nm, "list-4 local workspace"));
    QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_qbx, z.nq, "QBX expansion workspace"));
    QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_pot, z.np, "potential workspace"));
    DriveBufs b{ctx->ws_mpole.p, ctx->ws_local.p, ctx->ws_local2.p, ctx->ws_qbx.p,
                      ctx->ws_pot.p};
    if (d->kernel == QBX_LAPLACE3D) {
        QBX_TRY(ws_reserve(ctx, S(stream), ctx->ws_allreal, 1, "strength flag"));
        b.allreal = ctx->ws_allreal.p;
    }
    return drive_impl(ctx, stream, T, R, G, d, str, nrm, out_user, z, b,
                      ctx->tune[QBX_TUNE_DRIVE_STREAMS] > 1);
}

int qbx_drive_fmm_noalloc(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_geo *G, const qbx_kernel_desc *d, const double *str,
                          const double *nrm, double *out_user)
{
    QBX_CHECK(ctx && T && R && G && d && str && out_user,
              "qbx_drive_fmm_noalloc: null argument");
    QBX_CHECK(ctx->noalloc_scratch, "qbx_drive_fmm_noalloc: context has no scratch address");
    if (d->kernel != QBX_LAPLACE2D && d->kernel != QBX_HELMHOLTZ2D) {
        set_error("qbx_drive_fmm_noalloc: 2D Laplace and 2D Helmholtz only (got kernel %d)",
                  d->kernel);
        return QBX_ERR_UNSUPPORTED;
    }
    DriveSizes z;
    QBX_TRY(drive_sizes(T, G, d, &z));
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    // Helmholtz translation tables (the sizes ensure_tables builds)
    long ntab_m2l = 0, ntab_mm = 0;
    if (kd.kernel == QBX_HELMHOLTZ2D) {
        const int ncls = (2 * kd.m2l_w + 1) * (2 * kd.m2l_w + 1);
        for (int l = 0; l < kd.nlevels; ++l) ntab_m2l += 2L * ncls * (4 * kd.order[l] + 1);
        for (int l = 1; l < kd.nlevels; ++l)
            ntab_mm += 8L * (2 * (kd.order[l] + kd.order[l - 1]) + 1);
    }
    auto up = [](long n) { return (n + 31) & ~31L; };   // 256-byte aligned pieces
    const long need = 3 * up(z.nm) + up(z.nq) + up(z.np) + up(ntab_m2l) + up(std::max(ntab_mm, 1L));
    QBX_CHECK(need <= QBX_NOALLOC_SCRATCH_BYTES / 8,
              "qbx_drive_fmm_noalloc: this evaluation needs %ld bytes of scratch, more than the "
              "library's %ld (QBX_NOALLOC_SCRATCH_BYTES); use qbx_drive_fmm",
              need * 8, (long)QBX_NOALLOC_SCRATCH_BYTES);
    double *p = ctx->noalloc_scratch;
    DriveBufs b;
    b.M = p; p += up(z.nm);
    b.L = p; p += up(z.nm);
    b.L4 = p; p += up(z.nm);
    b.Q = p; p += up(z.nq);
    b.P = p; p += up(z.np);
    ctx->tab_m2l_ext = p; p += up(ntab_m2l);
    ctx->tab_mm_ext = p;
    const int rc = drive_impl(ctx, stream, T, R, G, d, str, nrm, out_user, z, b, false);
    ctx->tab_m2l_ext = ctx->tab_mm_ext = nullptr;
    return rc;
}

}  // extern "C"

// the adjoint (conjugate transpose) of the stages above
#include "adjoint2d.inc"

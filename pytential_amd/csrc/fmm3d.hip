// QBX-FMM wrangler stages for 3D Laplace (pytential/qbx/fmm.py drive_fmm and
// pytential/qbx/interactions.py), FP64 / complex128.
//
// Expansions in solid harmonics (coefficient (n, m) at n*n + n + m):
//   Y_n^m(x) = r^n P_n^m(cos th) e^{i m phi} / (n+m)!        regular
//   T_n^m(x) = (n-m)! P_n^m(cos th) e^{i m phi} / r^{n+1}    irregular
//   Y_n^{-m} = (-1)^m conj(Y_n^m), T_n^{-m} = (-1)^m conj(T_n^m)
//   1/|x-y| = sum conj(Y_n^m(y)) T_n^m(x) for |y| < |x|
//   multipole  M_n^m = sum q conj(Y_n^m(s-c)),  phi(x) = sum M T(x-c)
//   local      L_n^m = sum q T_n^m(s-c),        phi(x) = sum L conj(Y(x-c))
//   M2M  M'_n^m = sum_{k,l} conj(Y_k^l(c-c')) M_{n-k}^{m-l}
//   M2L  L_k^l  = (-1)^k sum_{n,m} M_n^m T_{n+k}^{m+l}(c'-c)
//   L2L  L'_a^b = sum_{j,i} L_{a+j}^{b+i} conj(Y_j^i(c'-c))
// Double layer: q nu . grad_s through
//   d_z Y_n^m = Y_{n-1}^m, (d_x +- i d_y) Y_n^m = +-Y_{n-1}^{m+-1},
//   d_z T_n^m = -T_{n+1}^m, (d_x +- i d_y) T_n^m = +-T_{n+1}^{m+-1}.
// Potentials carry G = 1/(4 pi r).
//
// Recurrences (no trigonometry, w = x + i y, r^2 = |x|^2):
//   Y_m^m = -(w/2m) Y_{m-1}^{m-1}, (n-m)(n+m) Y_n^m = (2n-1) z Y_{n-1}^m - r^2 Y_{n-2}^m
//   T_m^m = -(2m-1)(w/r^2) T_{m-1}^{m-1},
//   T_n^m = ((2n-1) z T_{n-1}^m - (n+m-1)(n-m-1) T_{n-2}^m) / r^2
//
// Right-hand sides: as in the 2D stages, nr densities share one evaluation,
// arrays carry them innermost, and each kernel applies an interaction's
// geometry (harmonics, translation weights) to a chunk of right-hand sides
// held in registers; the host launches one chunk [r0, r0 + rc) at a time.
//
// Storage (HBM): box expansions level-compact, (p_l+1)^2 complex per box of
// level l (times nr); QBX expansions [ncenters][(p_q+1)^2][nr].
//
// Thread mappings:
//  P2M, P2L        one thread per (box, coefficient)
//  M2M, L2L        one thread per (box, coefficient), child offsets from a
//                  table of Y at the 8 unit child offsets
//  M2L             one 64-lane block per target box; each list-2 multipole is
//                  staged (rescaled by h^-n) in LDS and contracted against
//                  the irregular harmonics of its integer offset class, read
//                  from a table T(delta), |delta_i| <= 2 nsep + 1
//  P2P, M2P, L2P   one block per target box, lane groups per target
//  P2QBXL          16 lanes per QBX center over its list-1 sources, the
//                  order-p_q expansion in registers, shuffle reduction
//  M2QBXL          16 lanes per QBX center: per list-3 box the irregular
//                  harmonics of the offset are built in LDS, then each lane
//                  contracts its output coefficients
//  L2QBXL          one thread per QBX center, expansion in registers
//  QBXL2P          one thread per (center, target) pair
#include "qbx_internal.h"
#include "qbx_math.h"

#include <complex>
#include <cstring>
#include <vector>

namespace qbx {
namespace f3 {

constexpr int MAXP = 20;       // box expansion orders
constexpr int MAXPQ = 8;       // QBX expansion orders (templated)
constexpr int FML = 21;        // levels of a 3D tree (19-bit coordinates)
constexpr int NRC_S = 8;       // RHS chunk of the scalar-accumulator kernels
constexpr int NRC_Q = 2;       // RHS chunk of the QBX-expansion kernels
constexpr double INV4PI = 0.079577471545947667884;

__host__ __device__ constexpr int IDX(int n, int m) { return n * n + n + m; }

// accumulator of v_mfma_f64_16x16x4f64 (4 doubles per lane)
typedef double f3_d4 __attribute__((ext_vector_type(4)));

struct KD {
    int dip, pq, nlevels, ncq, nsep, W, pmax, tabn;   // tabn = 2 pmax + 1
    int nr, r0, rc;            // right-hand sides; the chunk of this launch
    int order[FML];
    int ncl[FML];
    int lbox[FML + 1];
    long lcoef[FML + 1];
    double h[FML];
    const double *tab_m2l;   // [(2W+1)^3][tabn^2] complex: T(delta), unit box size
    const double *tab_ch;    // [8][(pmax+1)^2] complex: Y(bit - 1/2), unit child size
};

struct TD {
    int nboxes;
    const int *level, *parent, *child, *coords;
    const double *centers;  // [3][nboxes]
    const int *sstart, *scount, *scumul;
};

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

struct v3 {
    double x, y, z;
};
__device__ __forceinline__ v3 ld3(const double *p, long n, long i)
{
    return {p[i], p[n + i], p[2 * n + i]};
}
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 boxc(const TD &t, int b) { return ld3(t.centers, t.nboxes, b); }

// coefficient 0, right-hand side 0 of box b; coefficient i, right-hand side
// r at [2 (i nr + r)]
__device__ __forceinline__ long ecoef_off(const KD &kd, int b, int lv)
{
    return 2 * (kd.lcoef[lv] + (long)(b - kd.lbox[lv]) * kd.ncl[lv]) * kd.nr;
}

__device__ __forceinline__ cplx ldq(const KD &kd, const double *str, int s, int r)
{
    return ldc(str + 2 * ((size_t)s * kd.nr + r));
}

__device__ __forceinline__ cplx ldcn(const KD &kd, const double *a, int i, int r)
{
    return ldc(a + 2 * ((long)i * kd.nr + r));
}

__device__ __forceinline__ void nm_of(int i, int &n, int &m)
{
    n = (int)sqrt((double)i);
    while (n * n > i) --n;
    while ((n + 1) * (n + 1) <= i) ++n;
    m = i - n * n - n;
}

__device__ __forceinline__ cplx neg_m(cplx v, int am)  // (-1)^m conj(v)
{
    return (am & 1) ? cplx{-v.x, v.y} : cplx{v.x, -v.y};
}

// single Y_n^m(d), O(n)
__device__ cplx ups_nm(int n, int m, v3 d)
{
    const int am = m < 0 ? -m : m;
    if (am > n || n < 0) return {0, 0};
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const cplx hw = {-0.5 * d.x, -0.5 * d.y};
    cplx b = {1, 0}, a = {0, 0};
    for (int i = 1; i <= am; ++i) b = (1.0 / i) * (b * hw);
    for (int k = am + 1; k <= n; ++k) {
        const cplx c =
            (1.0 / ((double)(k - am) * (k + am))) * (((2.0 * k - 1.0) * d.z) * b - r2 * a);
        a = b;
        b = c;
    }
    return m < 0 ? neg_m(b, am) : b;
}

// single T_n^m(d), O(n)
__device__ cplx the_nm(int n, int m, v3 d)
{
    const int am = m < 0 ? -m : m;
    if (am > n || n < 0) return {0, 0};
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double ir2 = 1.0 / r2;
    const cplx w = {d.x * ir2, d.y * ir2};
    cplx b = {sqrt(ir2), 0}, a = {0, 0};
    for (int i = 1; i <= am; ++i) b = (-(2.0 * i - 1.0)) * (b * w);
    for (int k = am + 1; k <= n; ++k) {
        const cplx c = ir2 * (((2.0 * k - 1.0) * d.z) * b -
                              ((double)(k + am - 1) * (k - am - 1)) * a);
        a = b;
        b = c;
    }
    return m < 0 ? neg_m(b, am) : b;
}

// full table T_n^m(d), n <= P, in registers
template <int P>
__device__ __forceinline__ void the_table(v3 d, cplx *T)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double ir2 = 1.0 / r2;
    const cplx w = {d.x * ir2, d.y * ir2};
    cplx tmm = {sqrt(ir2), 0};
#pragma unroll
    for (int m = 0; m <= P; ++m) {
        if (m > 0) tmm = (-(2.0 * m - 1.0)) * (tmm * w);
        T[IDX(m, m)] = tmm;
        cplx a = {0, 0}, b = tmm;
#pragma unroll
        for (int n = m + 1; n <= P; ++n) {
            const cplx c = ir2 * (((2.0 * n - 1.0) * d.z) * b -
                                  ((double)(n + m - 1) * (n - m - 1)) * a);
            T[IDX(n, m)] = c;
            a = b;
            b = c;
        }
    }
#pragma unroll
    for (int n = 1; n <= P; ++n)
#pragma unroll
        for (int m = 1; m <= n; ++m) T[IDX(n, -m)] = neg_m(T[IDX(n, m)], m);
}

// out[q] += sum_{n<=p,m} coef[n,m][r0+q] T_n^m(d); coef strided by nr,
// streamed column by column with O(1) harmonic state
template <int NRC>
__device__ void sum_irregular_n(const KD &kd, const double *coef, int p, v3 d, cplx *out)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const double ir2 = 1.0 / r2;
    const cplx w = {d.x * ir2, d.y * ir2};
    cplx tmm = {sqrt(ir2), 0};
    for (int m = 0; m <= p; ++m) {
        if (m > 0) tmm = (-(2.0 * m - 1.0)) * (tmm * w);
        cplx a = {0, 0}, b = tmm;
        for (int n = m; n <= p; ++n) {
            if (n > m) {
                const cplx c = ir2 * (((2.0 * n - 1.0) * d.z) * b -
                                      ((double)(n + m - 1) * (n - m - 1)) * a);
                a = b;
                b = c;
            }
            const cplx bn = neg_m(b, m);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) {
                    out[q] += ldcn(kd, coef, IDX(n, m), kd.r0 + q) * b;
                    if (m > 0) out[q] += ldcn(kd, coef, IDX(n, -m), kd.r0 + q) * bn;
                }
        }
    }
}

// out[q] += sum_{n<=p,m} coef[n,m][r0+q] conj(Y_n^m(d)); conj(Y_n^{-m}) = (-1)^m Y_n^m
template <int NRC>
__device__ void sum_regular_conj_n(const KD &kd, const double *coef, int p, v3 d, cplx *out)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const cplx hw = {-0.5 * d.x, -0.5 * d.y};
    cplx ymm = {1, 0};
    for (int m = 0; m <= p; ++m) {
        if (m > 0) ymm = (1.0 / m) * (ymm * hw);
        cplx a = {0, 0}, b = ymm;
        for (int n = m; n <= p; ++n) {
            if (n > m) {
                const cplx c = (1.0 / ((double)(n - m) * (n + m))) *
                               (((2.0 * n - 1.0) * d.z) * b - r2 * a);
                a = b;
                b = c;
            }
            const cplx bc = conj(b);
            const cplx bs = (m & 1) ? -b : b;
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) {
                    out[q] += ldcn(kd, coef, IDX(n, m), kd.r0 + q) * bc;
                    if (m > 0) out[q] += ldcn(kd, coef, IDX(n, -m), kd.r0 + q) * bs;
                }
        }
    }
}

__device__ __forceinline__ double p2p_value(const KD &kd, v3 s, v3 nu, v3 t)
{
    const v3 d = sub(t, s);
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 == 0.0) return 0.0;
    const double ir = 1.0 / sqrt(r2);
    if (!kd.dip) return INV4PI * ir;
    return INV4PI * ir * ir * ir * (nu.x * d.x + nu.y * d.y + nu.z * d.z);
}

// coefficient (n, m) one unit-strength source contributes to a multipole
// (local = 0) or local (1)
__device__ __forceinline__ cplx p2e_basis(const KD &kd, int local, int n, int m, v3 d, v3 nu)
{
    if (!kd.dip) return local ? the_nm(n, m, d) : conj(ups_nm(n, m, d));
    const cplx nm = {nu.x, -nu.y}, np = {nu.x, nu.y};
    cplx g;
    if (local)
        g = (-nu.z) * the_nm(n + 1, m, d) +
            0.5 * (nm * the_nm(n + 1, m + 1, d) - np * the_nm(n + 1, m - 1, d));
    else
        g = nu.z * ups_nm(n - 1, m, d) +
            0.5 * (nm * ups_nm(n - 1, m + 1, d) - np * ups_nm(n - 1, m - 1, d));
    return local ? g : conj(g);
}

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

// P2M over source boxes: thread per (box, coefficient)
template <int NRC>
__global__ void k3_p2m(KD kd, TD t, Flat fl, const int *boxes, const double *src, int ns,
                       const double *str, const double *nrm, double *mpole)
{
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int lv, i, j;
    if (!flat_locate(fl, gid, lv, i, j)) return;
    const int b = boxes[i];
    int n, m;
    nm_of(j, n, m);
    const v3 c = boxc(t, b);
    const int s0 = t.sstart[b], s1 = s0 + t.scount[b];
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int s = s0; s < s1; ++s) {
        const v3 nu = kd.dip ? ld3(nrm, ns, s) : v3{0, 0, 0};
        const cplx bs = p2e_basis(kd, 0, n, m, sub(ld3(src, ns, s), c), nu);
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) acc[q] += ldq(kd, str, s, kd.r0 + q) * bs;
    }
    double *o = mpole + ecoef_off(kd, b, lv) + 2L * j * kd.nr;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) stc(o + 2 * (kd.r0 + q), acc[q]);
}

// P2L over list 4: thread per (ttp box, coefficient), accumulating
template <int NRC>
__global__ void k3_p2l(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
                       const double *src, int ns, const double *str, const double *nrm,
                       double *local)
{
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int lv, i, j;
    if (!flat_locate(fl, gid, lv, i, j)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int b = ttp[i];
    int n, m;
    nm_of(j, n, m);
    const v3 c = boxc(t, b);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int e = e0; e < e1; ++e) {
        const int S = lists[e];
        const int s0 = t.sstart[S], s1 = s0 + t.scount[S];
        for (int s = s0; s < s1; ++s) {
            const v3 nu = kd.dip ? ld3(nrm, ns, s) : v3{0, 0, 0};
            const cplx bs = p2e_basis(kd, 1, n, m, sub(ld3(src, ns, s), c), nu);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) acc[q] += ldq(kd, str, s, kd.r0 + q) * bs;
        }
    }
    store_add<NRC>(kd, local + ecoef_off(kd, b, lv) + 2L * j * kd.nr, acc);
}

// P2M / P2L through LDS, one block per box, sources S at a time.  Phase A
// (one lane per source) computes the source's harmonic table once, m >= 0
// half: Y_n^m(s - c) for n <= p (multipoles) or T_n^m(s - c) for n <= p + 1
// (locals, the dipole needs n + 1), at ht[s][n(n+1)/2 + m]; phase B (one lane
// per coefficient (n, m)) sums q_s b_nm(s) over the chunk, with the negative
// orders from (-1)^m conj and the dipole rules of p2e_basis.  O(p^2) work per
// source instead of O(p^3) from one harmonic per thread and source.
struct P2EShape3 {
    int S;  // sources per chunk
    int R;  // row stride (complex entries, odd) >= (pt+1)(pt+2)/2
    __host__ __device__ int doubles(int nrc) const { return 2 * S * R + 2 * S * nrc + 3 * S + S; }
};

__device__ __forceinline__ int tri(int n, int m) { return n * (n + 1) / 2 + m; }

// the source's table for k3_p2e: h[tri(n, m)], m >= 0, n <= pt, of d = source -
// center -- irregular harmonics (LOCAL) or regular ones (multipoles), by the
// recurrences in m and n
template <bool LOCAL>
__device__ __forceinline__ void p2e3_table(v3 d, int pt, double *h)
{
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (LOCAL) {
        const double ir2 = 1.0 / r2;
        const cplx w = {d.x * ir2, d.y * ir2};
        cplx tmm = {sqrt(ir2), 0};
        for (int mm = 0; mm <= pt; ++mm) {
            if (mm > 0) tmm = (-(2.0 * mm - 1.0)) * (tmm * w);
            cplx a = {0, 0}, b = tmm;
            stc(h + 2 * tri(mm, mm), b);
            for (int nn = mm + 1; nn <= pt; ++nn) {
                const cplx cc = ir2 * (((2.0 * nn - 1.0) * d.z) * b -
                                       ((double)(nn + mm - 1) * (nn - mm - 1)) * a);
                stc(h + 2 * tri(nn, mm), cc);
                a = b;
                b = cc;
            }
        }
    } else {
        const cplx hw = {-0.5 * d.x, -0.5 * d.y};
        cplx ymm = {1, 0};
        for (int mm = 0; mm <= pt; ++mm) {
            if (mm > 0) ymm = (1.0 / mm) * (ymm * hw);
            cplx a = {0, 0}, b = ymm;
            stc(h + 2 * tri(mm, mm), b);
            for (int nn = mm + 1; nn <= pt; ++nn) {
                const cplx cc = (1.0 / ((double)(nn - mm) * (nn + mm))) *
                                (((2.0 * nn - 1.0) * d.z) * b - r2 * a);
                stc(h + 2 * tri(nn, mm), cc);
                a = b;
                b = cc;
            }
        }
    }
}

template <int NRC, bool LOCAL>
__global__ void __launch_bounds__(512) k3_p2e(KD kd, TD t, int lv, int i0, P2EShape3 sh,
                                               const int *boxes, const int *starts,
                                               const int *lists, const double *src, int ns,
                                               const double *str, const double *nrm,
                                               double *expn)
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
    const int pt = LOCAL ? p + 1 : p;
    const int S = sh.S, R = sh.R;
    double *Ht = lds, *Qs = Ht + 2 * S * R, *Nu = Qs + 2 * S * NRC;
    int *idx = (int *)(Nu + 3 * S);
    const int tid = threadIdx.x, nth = blockDim.x;
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    const v3 c = boxc(t, Bx);
    const bool act = tid < (p + 1) * (p + 1);
    int n = 0, m = 0;
    if (act) nm_of(tid, n, m);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};

    // table entry (nn, mm) of staged source sl, any sign of mm, zero outside
    auto hval = [&](int sl, int nn, int mm) -> cplx {
        const int am = mm < 0 ? -mm : mm;
        if (nn < 0 || am > nn) return {0, 0};
        const cplx v = {Ht[2 * (sl * R + tri(nn, am))], Ht[2 * (sl * R + tri(nn, am)) + 1]};
        return mm < 0 ? neg_m(v, am) : v;
    };

    auto flush = [&](int cnt) {
        __syncthreads();
        if (tid < cnt) {
            const int si = idx[tid];
            p2e3_table<LOCAL>(sub(ld3(src, ns, si), c), pt, Ht + 2 * tid * R);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < rc) {
                    const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                    Qs[2 * (tid * NRC + q)] = pq[0];
                    Qs[2 * (tid * NRC + q) + 1] = pq[1];
                }
            if (dip) {
                Nu[3 * tid] = nrm[si];
                Nu[3 * tid + 1] = nrm[(size_t)ns + si];
                Nu[3 * tid + 2] = nrm[2 * (size_t)ns + si];
            }
        }
        __syncthreads();
        if (act)
            for (int sl = 0; sl < cnt; ++sl) {
                cplx b;
                if (!dip) {
                    b = LOCAL ? hval(sl, n, m) : conj(hval(sl, n, m));
                } else {
                    const double nx = Nu[3 * sl], ny = Nu[3 * sl + 1], nz = Nu[3 * sl + 2];
                    const cplx nmn = {nx, -ny}, npl = {nx, ny};
                    if (LOCAL) {
                        b = (-nz) * hval(sl, n + 1, m) +
                            0.5 * (nmn * hval(sl, n + 1, m + 1) - npl * hval(sl, n + 1, m - 1));
                    } else {
                        b = conj(nz * hval(sl, n - 1, m) +
                                 0.5 * (nmn * hval(sl, n - 1, m + 1) -
                                        npl * hval(sl, n - 1, m - 1)));
                    }
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc)
                        acc[q] += cplx{Qs[2 * (sl * NRC + q)], Qs[2 * (sl * NRC + q) + 1]} * b;
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
        double *o = expn + ecoef_off(kd, Bx, lv) + 2L * tid * nr;
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < rc) {
                double *pq = o + 2 * (r0 + q);
                stc(pq, LOCAL ? ldc(pq) + acc[q] : acc[q]);
            }
    }
}

// P2M / P2L for many right-hand sides on the FP64 matrix cores: the sources
// of the box (P2M) or of its list-4 boxes (P2L) are staged P2E3M_S at a time
// with their tables (p2e3_table) and the strengths of up to 16 right-hand
// sides; out[(n,m)][r] (+)= sum_s b_(n,m)(s) q_s[r] with b as in k3_p2e
// (single or double layer) is a complex product: rows the coefficients
// (wave w takes the 16-row tiles w, w + 4, ...), k the staged sources.
constexpr int P2E3M_S = 32;
constexpr int P2E3M_RTW = 4;  // (p + 1)^2 <= 256

template <bool LOCAL>
__global__ void __launch_bounds__(256)
k3_p2e_mfma(KD kd, TD t, int lv, int i0, int R, const int *boxes, const int *starts,
            const int *lists, const double *src, int ns, const double *str, const double *nrm,
            double *expn)
{
    extern __shared__ double lds[];
    double *Ht = lds, *Qs = Ht + 2 * P2E3M_S * R, *Nu = Qs + 2 * P2E3M_S * 16;
    int *idx = (int *)(Nu + 3 * P2E3M_S);
    __shared__ short nof[256], mof[256];
    const int i = i0 + blockIdx.x;
    const int Bx = boxes[i];
    int e0 = 0, e1 = 1;
    if (LOCAL) {
        e0 = starts[i];
        e1 = starts[i + 1];
        if (e0 == e1) return;
    }
    const int p = kd.order[lv];
    const int pt = LOCAL ? p + 1 : p;
    const int nout = (p + 1) * (p + 1);
    const int RT = (nout + 15) / 16;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    const v3 c = boxc(t, Bx);
    {
        int n = (int)sqrt((double)tid);
        while (n * n > tid) --n;
        while ((n + 1) * (n + 1) <= tid) ++n;
        nof[tid] = (short)n;
        mof[tid] = (short)(tid - n * n - n);
    }
    f3_d4 cre[P2E3M_RTW], cim[P2E3M_RTW];
#pragma unroll
    for (int q = 0; q < P2E3M_RTW; ++q) cre[q] = cim[q] = f3_d4{0, 0, 0, 0};

    auto hval = [&](int sl, int nn, int mm) -> cplx {
        const int am = mm < 0 ? -mm : mm;
        if (nn < 0 || am > nn) return {0, 0};
        const cplx v = {Ht[2 * (sl * R + tri(nn, am))], Ht[2 * (sl * R + tri(nn, am)) + 1]};
        return mm < 0 ? neg_m(v, am) : v;
    };

    auto flush = [&](int cnt) {
        __syncthreads();
        if (tid < cnt) {
            const int si = idx[tid];
            p2e3_table<LOCAL>(sub(ld3(src, ns, si), c), pt, Ht + 2 * tid * R);
            if (dip) {
                Nu[3 * tid] = nrm[si];
                Nu[3 * tid + 1] = nrm[(size_t)ns + si];
                Nu[3 * tid + 2] = nrm[2 * (size_t)ns + si];
            }
        }
        int imag = 0;
        for (int e = tid; e < P2E3M_S * 16; e += 256) {
            const int k = e >> 4, r = e & 15;
            double vr = 0.0, vi = 0.0;
            if (k < cnt && r < rc) {
                const double *pq = str + 2 * ((size_t)idx[k] * nr + r0 + r);
                vr = pq[0];
                vi = pq[1];
            }
            imag |= vi != 0.0;
            Qs[2 * e] = vr;
            Qs[2 * e + 1] = vi;
        }
        // real strengths (a real density): the products with the zero
        // imaginary parts are skipped (block-uniform)
        imag = __syncthreads_or(imag);
        for (int ks = 0; ks < P2E3M_S / 4; ++ks) {
            const int sl = ks * 4 + kr;
            const double bre = Qs[2 * (sl * 16 + col)], bim = Qs[2 * (sl * 16 + col) + 1];
#pragma unroll
            for (int q = 0; q < P2E3M_RTW; ++q) {
                const int rt = wv + 4 * q;
                if (rt < RT) {
                    const int row = rt * 16 + col;
                    cplx a = {0, 0};
                    if (row < nout && sl < cnt) {
                        const int n = nof[row], m = mof[row];
                        if (!dip) {
                            a = LOCAL ? hval(sl, n, m) : conj(hval(sl, n, m));
                        } else {
                            const double nx = Nu[3 * sl], ny = Nu[3 * sl + 1], nz = Nu[3 * sl + 2];
                            const cplx nmn = {nx, -ny}, npl = {nx, ny};
                            if (LOCAL)
                                a = (-nz) * hval(sl, n + 1, m) +
                                    0.5 * (nmn * hval(sl, n + 1, m + 1) -
                                           npl * hval(sl, n + 1, m - 1));
                            else
                                a = conj(nz * hval(sl, n - 1, m) +
                                         0.5 * (nmn * hval(sl, n - 1, m + 1) -
                                                npl * hval(sl, n - 1, m - 1)));
                        }
                    }
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bre, cre[q], 0, 0, 0);
                    if (imag) {
                        cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, bim, cre[q], 0, 0, 0);
                        cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bim, cim[q], 0, 0, 0);
                    }
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bre, cim[q], 0, 0, 0);
                }
            }
        }
    };

    int fill = 0;
    for (int e = e0; e < e1; ++e) {
        const int Sb = LOCAL ? lists[e] : Bx;
        int s = t.sstart[Sb];
        const int s1 = s + t.scount[Sb];
        while (s < s1) {
            const int take = min(s1 - s, P2E3M_S - fill);
            for (int j = tid; j < take; j += 256) idx[fill + j] = s + j;
            fill += take;
            s += take;
            if (fill == P2E3M_S) {
                flush(P2E3M_S);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
    double *o = expn + ecoef_off(kd, Bx, lv);
#pragma unroll
    for (int q = 0; q < P2E3M_RTW; ++q) {
        const int rt = wv + 4 * q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = rt * 16 + 4 * j + kr;
            if (rt < RT && row < nout && col < rc) {
                double *pq = o + 2 * ((long)row * nr + r0 + col);
                const cplx v = {cre[q][j], cim[q][j]};
                stc(pq, LOCAL ? ldc(pq) + v : v);
            }
        }
    }
}

// child offset index bits: bit d set = upper half in dimension d
__device__ __forceinline__ const double *child_tab(const KD &kd, int ci)
{
    return kd.tab_ch + 2L * ci * (kd.pmax + 1) * (kd.pmax + 1);
}

// M2M into parents of level lvp: thread per (parent, coefficient)
template <int NRC>
__global__ void k3_m2m(KD kd, TD t, int lvp, int i0, int i1, const int *spb, double *mpole,
                       const int *allreal)
{
    if (allreal && *allreal) return;  // k3_trans_real's case
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ncp = kd.ncl[lvp];
    const long nthr = (long)(i1 - i0) * ncp;
    if (gid >= nthr) return;
    const int P = spb[i0 + (int)(gid / ncp)];
    const int j = (int)(gid % ncp);
    int n, m;
    nm_of(j, n, m);
    const int lvc = lvp + 1;
    const int pc = kd.order[lvc];
    const double hc = kd.h[lvc];
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    for (int ci = 0; ci < 8; ++ci) {
        const int C = t.child[(size_t)ci * t.nboxes + P];
        if (C < 0 || t.scumul[C] == 0) continue;
        const double *a = mpole + ecoef_off(kd, C, lvc);
        const double *U = child_tab(kd, ci);
        // conj(Y_k^l(c_child - c_parent)) = conj(U_k^l) hc^k
        double hk = 1.0;
        for (int k = 0; k <= n; ++k) {
            const int nn = n - k;
            if (nn <= pc) {
                for (int l = -k; l <= k; ++l) {
                    const int mm = m - l;
                    if (mm < -nn || mm > nn) continue;
                    const cplx wgt = hk * conj(ldc(U + 2 * IDX(k, l)));
#pragma unroll
                    for (int q = 0; q < NRC; ++q)
                        if (q < kd.rc) acc[q] += wgt * ldcn(kd, a, IDX(nn, mm), kd.r0 + q);
                }
            }
            hk *= hc;
        }
    }
    double *o = mpole + ecoef_off(kd, P, lvp) + 2L * j * kd.nr;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) stc(o + 2 * (kd.r0 + q), acc[q]);
}

// L2L into ttp boxes of level lv: thread per (box, coefficient), accumulating
template <int NRC>
__global__ void k3_l2l(KD kd, TD t, int lv, int i0, int i1, const int *ttp, double *local,
                       const int *allreal)
{
    if (allreal && *allreal) return;  // k3_trans_real's case
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = kd.ncl[lv];
    const long nthr = (long)(i1 - i0) * nc;
    if (gid >= nthr) return;
    const int Bx = ttp[i0 + (int)(gid / nc)];
    const int j = (int)(gid % nc);
    int a, b;
    nm_of(j, a, b);
    const int Pp = t.parent[Bx];
    const int lvp = lv - 1;
    const int pp = kd.order[lvp];
    const double hc = kd.h[lv];
    // child index of Bx within its parent from coordinates
    int ci = 0;
    ci |= (t.coords[Bx] & 1);
    ci |= (t.coords[(size_t)t.nboxes + Bx] & 1) << 1;
    ci |= (t.coords[2 * (size_t)t.nboxes + Bx] & 1) << 2;
    const double *U = child_tab(kd, ci);
    const double *L = local + ecoef_off(kd, Pp, lvp);
    cplx acc[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
    double hj = 1.0;
    for (int jj = 0; a + jj <= pp; ++jj) {
        for (int i = -jj; i <= jj; ++i) {
            const int mm = b + i;
            if (mm < -(a + jj) || mm > a + jj) continue;
            const cplx wgt = hj * conj(ldc(U + 2 * IDX(jj, i)));
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < kd.rc) acc[q] += wgt * ldcn(kd, L, IDX(a + jj, mm), kd.r0 + q);
        }
        hj *= hc;
    }
    store_add<NRC>(kd, local + ecoef_off(kd, Bx, lv) + 2L * j * kd.nr, acc);
}

constexpr int M2L_THREADS = 64;

// M2L: one 64-lane block per ttp box with list-2 entries; NO output
// coefficients per lane, NRC right-hand sides; LDS holds the rescaled source
// multipole [ncl][NRC]
template <int NO, int NRC>
__global__ void __launch_bounds__(M2L_THREADS) k3_m2l(KD kd, TD t, const int *ttp,
                                                      const int *starts, const int *lists,
                                                      const double *mpole, double *local)
{
    extern __shared__ double Ms[];
    const int i = blockIdx.x;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int Bx = ttp[i];
    const int lv = t.level[Bx];
    const int p = kd.order[lv];
    const int nc = kd.ncl[lv];
    const double h = kd.h[lv];
    const double ih = 1.0 / h;
    const int W = kd.W, S1 = 2 * W + 1;
    const int tn = kd.tabn;
    const int bx = t.coords[Bx], by = t.coords[(size_t)t.nboxes + Bx],
              bz = t.coords[2 * (size_t)t.nboxes + Bx];
    cplx acc[NO][NRC];
#pragma unroll
    for (int r = 0; r < NO; ++r)
#pragma unroll
        for (int q = 0; q < NRC; ++q) acc[r][q] = {0, 0};
    for (int e = e0; e < e1; ++e) {
        const int A = lists[e];
        const int dx = bx - t.coords[A] + W;
        const int dy = by - t.coords[(size_t)t.nboxes + A] + W;
        const int dz = bz - t.coords[2 * (size_t)t.nboxes + A] + W;
        // list 2 offsets satisfy |delta_i| <= 2 nsep + 1 by construction
        if ((unsigned)dx >= (unsigned)S1 || (unsigned)dy >= (unsigned)S1 ||
            (unsigned)dz >= (unsigned)S1)
            continue;
        const double *a = mpole + ecoef_off(kd, A, lv);
        __syncthreads();
        for (int qq = threadIdx.x; qq < nc; qq += M2L_THREADS) {
            int n, m;
            nm_of(qq, n, m);
            double s = 1.0;
            for (int k = 0; k < n; ++k) s *= ih;
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                stc(Ms + 2 * (qq * NRC + q),
                    q < kd.rc ? s * ldcn(kd, a, qq, kd.r0 + q) : cplx{0, 0});
        }
        __syncthreads();
        const double *T = kd.tab_m2l + 2L * ((dx * S1 + dy) * S1 + dz) * tn * tn;
#pragma unroll
        for (int r = 0; r < NO; ++r) {
            const int qo = threadIdx.x + r * M2L_THREADS;
            if (qo < nc) {
                int k, l;
                nm_of(qo, k, l);
                for (int n = 0; n <= p; ++n) {
                    const int N = n + k;
                    for (int m = -n; m <= n; ++m) {
                        const int M = m + l;
                        if (M < -N || M > N) continue;
                        const cplx tv = ldc(T + 2 * IDX(N, M));
#pragma unroll
                        for (int q = 0; q < NRC; ++q)
                            acc[r][q] += ldc(Ms + 2 * (IDX(n, m) * NRC + q)) * tv;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NO; ++r) {
        const int qo = threadIdx.x + r * M2L_THREADS;
        if (qo < nc) {
            int k, l;
            nm_of(qo, k, l);
            double s = (k & 1) ? -ih : ih;  // (-1)^k h^{-k-1}
            for (int kk = 0; kk < k; ++kk) s *= ih;
            cplx v[NRC];
#pragma unroll
            for (int q = 0; q < NRC; ++q) v[q] = s * acc[r][q];
            store_add<NRC>(kd, local + ecoef_off(kd, Bx, lv) + 2L * qo * kd.nr, v);
        }
    }
}

// M2L for many right-hand sides on the FP64 matrix cores: per target box
// out(k,l)[r] = sum_(n,m) T(n+k, m+l) h^-n a(n,m)[r] (the sum of k3_m2l), a
// complex product of the (k,l) x (n,m) operator -- entries gathered from the
// entry's unit-offset table, staged in LDS -- with the multipole coefficients
// of up to 16 right-hand sides, on v_mfma_f64_16x16x4f64 (lane i holds
// A[i%16][k0 + i/16] and B[k0 + i/16][i%16]; D register j of lane i is
// D[4j + i/16][i%16]).  Wave w takes the 16-row tiles w, w + 4, ... over all
// entries; the epilogue scales by (-1)^k h^(-k-1) as k3_m2l.
constexpr int M2L3M_RTW = 4;             // row tiles per wave: (p + 1)^2 <= 256
constexpr int M2L3M_T = 2 * 15 + 1;      // table side for p <= 15

__global__ void __launch_bounds__(256)
k3_m2l_mfma(KD kd, TD t, Flat fl, const int *ttp, const int *starts, const int *lists,
            const double *mpole, double *local, const int *allreal)
{
    __shared__ double Ts[2 * M2L3M_T * M2L3M_T];
    __shared__ short nof[256], mof[256];
    __shared__ double ihp[32];
    if (allreal && *allreal) return;  // k3_trans_real's case
    int lv, i, unused;
    if (!flat_locate(fl, blockIdx.x, lv, i, unused)) return;
    const int e0 = starts[i], e1 = starts[i + 1];
    if (e0 == e1) return;
    const int Bx = ttp[i];
    const int p = kd.order[lv];
    const int nc = kd.ncl[lv];
    const int RT = (nc + 15) / 16, KS = (nc + 3) / 4;
    const int tnl = (2 * p + 1) * (2 * p + 1);  // table entries used: IDX(N, M), N <= 2p
    const double ih = 1.0 / kd.h[lv];
    const int W = kd.W, S1 = 2 * W + 1, tn = kd.tabn;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rc = kd.rc, r0 = kd.r0;
    for (int q = tid; q < 256; q += 256) {
        int n = (int)sqrt((double)q);
        while (n * n > q) --n;
        while ((n + 1) * (n + 1) <= q) ++n;
        nof[q] = (short)n;
        mof[q] = (short)(q - n * n - n);
    }
    if (tid < 32) {
        double s = 1.0;
        for (int k = 0; k < tid; ++k) s *= ih;
        ihp[tid] = s;
    }
    __syncthreads();  // nof, mof, ihp (the epilogue reads them even with no entry)
    const int bx = t.coords[Bx], by = t.coords[(size_t)t.nboxes + Bx],
              bz = t.coords[2 * (size_t)t.nboxes + Bx];
    f3_d4 cre[M2L3M_RTW], cim[M2L3M_RTW];
#pragma unroll
    for (int q = 0; q < M2L3M_RTW; ++q) cre[q] = cim[q] = f3_d4{0, 0, 0, 0};
    for (int e = e0; e < e1; ++e) {
        const int A = lists[e];
        const int dx = bx - t.coords[A] + W;
        const int dy = by - t.coords[(size_t)t.nboxes + A] + W;
        const int dz = bz - t.coords[2 * (size_t)t.nboxes + A] + W;
        if ((unsigned)dx >= (unsigned)S1 || (unsigned)dy >= (unsigned)S1 ||
            (unsigned)dz >= (unsigned)S1)
            continue;  // uniform over the block
        const double *a = mpole + ecoef_off(kd, A, lv);
        const double *T = kd.tab_m2l + 2L * ((dx * S1 + dy) * S1 + dz) * tn * tn;
        __syncthreads();  // previous table consumed (and nof/ihp written)
        for (int q = tid; q < tnl; q += 256) {
            Ts[2 * q] = T[2 * q];
            Ts[2 * q + 1] = T[2 * q + 1];
        }
        __syncthreads();
        for (int ks = 0; ks < KS; ++ks) {
            const int kq = ks * 4 + kr;  // (n, m) of the B row / A column
            double bre = 0.0, bim = 0.0;
            int n = 0, m = 0;
            if (kq < nc) {
                n = nof[kq];
                m = mof[kq];
                if (col < rc) {
                    const cplx b = ldcn(kd, a, kq, r0 + col);
                    bre = ihp[n] * b.x;
                    bim = ihp[n] * b.y;
                }
            }
#pragma unroll
            for (int q = 0; q < M2L3M_RTW; ++q) {
                const int rt = wv + 4 * q;
                if (rt < RT) {
                    const int row = rt * 16 + col;
                    double are = 0.0, aim = 0.0;
                    if (row < nc && kq < nc) {
                        const int N = n + nof[row], M = m + mof[row];
                        if (M >= -N && M <= N) {
                            are = Ts[2 * IDX(N, M)];
                            aim = Ts[2 * IDX(N, M) + 1];
                        }
                    }
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, cre[q], 0, 0, 0);
                    cre[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim, cre[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, cim[q], 0, 0, 0);
                    cim[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, cim[q], 0, 0, 0);
                }
            }
        }
    }
    double *o = local + ecoef_off(kd, Bx, lv);
#pragma unroll
    for (int q = 0; q < M2L3M_RTW; ++q) {
        const int rt = wv + 4 * q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = rt * 16 + 4 * j + kr;
            if (rt < RT && row < nc && col < rc) {
                const int k = nof[row];
                const double sc = ((k & 1) ? -1.0 : 1.0) * ihp[k + 1];
                double *dst = o + 2 * ((long)row * kd.nr + r0 + col);
                dst[0] += sc * cre[q][j];
                dst[1] += sc * cim[q][j];
            }
        }
    }
}

// Box translations for many real densities on the FP64 matrix cores.  Real
// strengths give M_n^{-m} = (-1)^m conj(M_n^m) and L_n^{-m} = (-1)^m
// conj(L_n^m) for every multipole and local, so a translation out = A in
// forms only the output rows i = (k, l), l >= 0, from the input columns
// j = (n, m), m >= 0: out_i = sum_j A1 a_j + A2 conj(a_j) with A1 = A(i, j)
// and A2 = (-1)^m A(i, (n, -m)) (0 for m = 0), a real-linear map
//   [Re out]   [A1r + A2r   A2i - A1i] [Re a]
//   [Im out] = [A1i + A2i   A1r - A2r] [Im a]
// of (P+1)(P+2) rows and columns -- about a third of the real products of
// the complex (P+1)^2 square.  The operator of one entry, read from a table
// staged in LDS:
//   M2L (OPK 0)  A = T_{n+k}^{m+l}(delta), the unit-offset irregular
//                harmonics of the list-2 offset; B scaled h^-n, output
//                (-1)^k h^-(k+1); entries: the list-2 boxes
//   M2M (OPK 1)  A = conj(Y_{k-n}^{l-m}(child offset)) for n <= k (unit
//                child size); B scaled hc^-n, output hc^k; entries: the
//                children, output the parent (stored)
//   L2L (OPK 2)  A = conj(Y_{n-k}^{m-l}(child offset)) for n >= k; B (the
//                parent's locals) scaled hc^n, output hc^-k; entry: the
//                parent, output the box (added)
// (row i = (k, l), column j = (n, m); for M2M and L2L the tiles on the zero
// side of n = k are skipped).  Rows and columns interleave (re, im) of each
// coefficient (row 2i + pr, column 2j + pk).  A block per output box of one
// level and 16 NCT right-hand sides.  Per entry the table and the input
// box's m >= 0 coefficients (scaled, split into re / im rows) are staged in
// LDS; the next entry's are loaded into registers while the current one is
// contracted.  The K-steps of each entry are dealt round-robin to the four
// waves, each wave holding every row tile; the waves' sums are added in LDS
// at the end and the epilogue writes row (k, l) and its mirror (k, -l).
// With allreal given the kernel runs only when *allreal is set (the complex
// kernels take the other case); without, the caller has stated it.
constexpr int TR3_ECH = 256;  // entries whose offsets are staged at once

__host__ __device__ constexpr int tri_n(int j)  // n of triangular index j
{
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= j) ++n;
    return n;
}

template <int OPK, int P, int NCT>
__global__ void __launch_bounds__(256)
k3_trans_real(KD kd, TD t, int lv, int i0, const int *boxes, const int *starts,
              const int *lists, const double *in, double *out, const int *allreal)
{
    constexpr int R = (P + 1) * (P + 2) / 2;      // coefficients with m >= 0
    constexpr int R2 = 2 * R;                     // real rows / columns
    constexpr int RT = (R2 + 15) / 16;            // row tiles
    constexpr int KS = (R2 + 3) / 4;              // K-steps
    // high orders split the row tiles over RG groups of waves (fewer
    // accumulators per wave) and the K-steps over the KG waves of a group
    constexpr int RG = (OPK == 0 && P >= 10) ? 2 : 1, KG = 4 / RG;
    constexpr int RTW = (RT + RG - 1) / RG;       // row tiles per wave
    constexpr int RC = 16 * NCT;                  // right-hand sides per block
    constexpr int TNL = OPK == 0 ? (2 * P + 1) * (2 * P + 1) : (P + 1) * (P + 1);
    constexpr int NT = (2 * TNL + 255) / 256;       // table doubles per thread
    constexpr int NB = (R * RC + 255) / 256;        // (coefficient, rhs) pairs per thread
    // staging area, reused for the reduction at the end
    constexpr int LDSN = (2 * TNL + R2 * RC) > RT * 16 * 17 ? (2 * TNL + R2 * RC)
                                                            : RT * 16 * 17;
    __shared__ double lds[LDSN];
    __shared__ double bsc[P + 1], osc[P + 1];
    __shared__ short jn[R];
    __shared__ int etab[TR3_ECH];
    __shared__ long eoff[TR3_ECH];
    __shared__ int nent;
    double *Ts = lds, *Bs = lds + 2 * TNL;
    if (allreal && *allreal == 0) return;
    const int i = i0 + blockIdx.x;
    const int Bx = boxes[i];
    // the entry range: list 2, the children, or the parent
    int e0 = 0, e1 = 0;
    if (OPK == 0) {
        e0 = starts[i];
        e1 = starts[i + 1];
        if (e0 == e1) return;
    } else if (OPK == 1) {
        e1 = 8;
    } else {
        e1 = 1;
    }
    const int W = kd.W, S1 = 2 * W + 1, tn = kd.tabn;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const int rg = wv / KG, kg = wv % KG;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    // the level of the input boxes and the scalings
    const int lin = OPK == 0 ? lv : (OPK == 1 ? lv + 1 : lv - 1);
    if (tid <= P) {
        const double h = OPK == 0 ? kd.h[lv] : (OPK == 1 ? kd.h[lv + 1] : kd.h[lv]);
        const double ih = 1.0 / h;
        double pw = 1.0, ipw = 1.0;
        for (int k = 0; k < tid; ++k) {
            pw *= h;
            ipw *= ih;
        }
        if (OPK == 0) {
            bsc[tid] = ipw;
            osc[tid] = ((tid & 1) ? -ipw : ipw) * ih;
        } else if (OPK == 1) {
            bsc[tid] = ipw;
            osc[tid] = pw;
        } else {
            bsc[tid] = pw;
            osc[tid] = ipw;
        }
    }
    for (int j = tid; j < R; j += 256) jn[j] = (short)tri_n(j);
    // this lane's A rows: (k, l, pr) of row (rg RTW + rt) 16 + col
    int rowk[RTW], rowl[RTW], rowpr[RTW];
#pragma unroll
    for (int rt = 0; rt < RTW; ++rt) {
        const int rho = (rg * RTW + rt) * 16 + col;
        const int ii = rho >> 1;
        const int k = tri_n(ii);
        rowk[rt] = rho < R2 ? k : -1;
        rowl[rt] = ii - k * (k + 1) / 2;
        rowpr[rt] = rho & 1;
    }
    const int bx = t.coords[Bx], by = t.coords[(size_t)t.nboxes + Bx],
              bz = t.coords[2 * (size_t)t.nboxes + Bx];
    // the staged data of an entry: table doubles and (re, im) of the B pairs;
    // the entries' table and coefficient offsets come from LDS (filled per
    // chunk of TR3_ECH entries), so the loads of the next entry depend on no
    // global read and overlap the current entry's products
    double tr[NT], br[NB][2];
    auto fetch = [&](int ee) -> bool {
        const int tix = etab[ee];
        if (tix < 0) return false;  // uniform over the block
        const double *T = OPK == 0 ? kd.tab_m2l + 2L * tix * tn * tn
                                   : kd.tab_ch + 2L * tix * (kd.pmax + 1) * (kd.pmax + 1);
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const int q = tid + 256 * u;
            // the child table holds Y: stage conj(Y) (flip the imaginary parts)
            tr[u] = q < 2 * TNL ? ((OPK != 0 && (q & 1)) ? -T[q] : T[q]) : 0.0;
        }
        const double *a = in + eoff[ee];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int f = tid + 256 * u;
            const int j = f / RC, r = f % RC;
            br[u][0] = br[u][1] = 0.0;
            if (f < R * RC && r < rc) {
                const int n = jn[j];
                const int m = j - n * (n + 1) / 2;
                const double *z = a + 2 * ((long)(n * n + n + m) * nr + r0 + r);
                br[u][0] = z[0];
                br[u][1] = z[1];
            }
        }
        return true;
    };
    f3_d4 acc[RTW][NCT];
#pragma unroll
    for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
        for (int c = 0; c < NCT; ++c) acc[rt][c] = f3_d4{0, 0, 0, 0};
    int step = 0;  // K-steps dealt so far (rotates the round-robin start)
    for (int cb = e0; cb < e1; cb += TR3_ECH) {
        __syncthreads();  // the previous chunk's entries consumed (bsc, jn written)
        if (OPK == 0) {
            const int cn = min(TR3_ECH, e1 - cb);
            for (int q = tid; q < cn; q += 256) {
                const int A = lists[cb + q];
                const int dx = bx - t.coords[A] + W;
                const int dy = by - t.coords[(size_t)t.nboxes + A] + W;
                const int dz = bz - t.coords[2 * (size_t)t.nboxes + A] + W;
                const bool ok = (unsigned)dx < (unsigned)S1 && (unsigned)dy < (unsigned)S1 &&
                                (unsigned)dz < (unsigned)S1;
                etab[q] = ok ? (dx * S1 + dy) * S1 + dz : -1;
                eoff[q] = ecoef_off(kd, A, lin);
            }
            if (tid == 0) nent = cn;
        } else if (OPK == 1) {
            // the nonempty children, octant ci = child slot ci
            if (tid == 0) {
                int c = 0;
                for (int ci = 0; ci < 8; ++ci) {
                    const int C = t.child[(size_t)ci * t.nboxes + Bx];
                    if (C < 0 || t.scumul[C] == 0) continue;
                    etab[c] = ci;
                    eoff[c] = ecoef_off(kd, C, lin);
                    ++c;
                }
                nent = c;
            }
        } else {
            if (tid == 0) {
                etab[0] = (bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2);
                eoff[0] = ecoef_off(kd, t.parent[Bx], lin);
                nent = 1;
            }
        }
        __syncthreads();
        const int cn = nent;
        bool have = cn > 0 ? fetch(0) : false;
        for (int ee = 0; ee < cn; ++ee) {
            __syncthreads();  // the previous entry's staging consumed
            if (have) {
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    const int q = tid + 256 * u;
                    if (q < 2 * TNL) Ts[q] = tr[u];
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    const int f = tid + 256 * u;
                    if (f < R * RC) {
                        const int j = f / RC, r = f % RC;
                        const double sc = bsc[jn[j]];
                        Bs[(2 * j) * RC + r] = sc * br[u][0];
                        Bs[(2 * j + 1) * RC + r] = sc * br[u][1];
                    }
                }
            }
            const bool cur = have;
            __syncthreads();
            have = ee + 1 < cn ? fetch(ee + 1) : false;
            if (!cur) continue;
            for (int ks = (kg - step + KG * 4) % KG; ks < KS; ks += KG) {
                const int kap = ks * 4 + kr;  // B row / A column
                const int j = kap >> 1, pk = kap & 1;
                const bool kin = kap < R2;
                int n = 0, m = 0;
                double b[NCT];
#pragma unroll
                for (int c = 0; c < NCT; ++c) b[c] = 0.0;
                if (kin) {
                    n = jn[j];
                    m = j - n * (n + 1) / 2;
#pragma unroll
                    for (int c = 0; c < NCT; ++c) b[c] = Bs[kap * RC + c * 16 + col];
                }
                const double sm = (m == 0) ? 0.0 : ((m & 1) ? -1.0 : 1.0);
                // the K-step's column orders (uniform): M2M and L2L skip the
                // row tiles on the zero side of n = k
                const int nlo = tri_n(min(2 * ks, R - 1)), nhi = tri_n(min(2 * ks + 1, R - 1));
                // branch-free gather: every row tile's two table reads are
                // issued together (invalid entries read 0 and are masked),
                // then the matrix products follow
                double c1[RTW], c2[RTW];
                bool v1[RTW], v2[RTW];
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt) {
                    const int k = rowk[rt], l = rowl[rt];
                    const int off = rowpr[rt] == pk ? 0 : 1;
                    int i1, i2;
                    bool ok1, ok2;
                    if (OPK == 0) {
                        const int N = n + k, base = N * N + N;
                        i1 = base + m + l;
                        i2 = base + l - m;
                        ok1 = ok2 = kin && k >= 0;
                    } else if (OPK == 1) {
                        const int K = k - n, base = K * K + K;
                        i1 = base + l - m;
                        i2 = base + l + m;
                        ok1 = kin && k >= 0 && K >= 0 && l - m <= K && m - l <= K;
                        ok2 = kin && k >= 0 && K >= 0 && l + m <= K;
                    } else {
                        const int K = n - k, base = K * K + K;
                        i1 = base + m - l;
                        i2 = base - m - l;
                        ok1 = kin && k >= 0 && K >= 0 && m - l <= K && l - m <= K;
                        ok2 = kin && k >= 0 && K >= 0 && m + l <= K;
                    }
                    v1[rt] = ok1;
                    v2[rt] = ok2;
                    c1[rt] = Ts[ok1 ? 2 * i1 + off : 0];
                    c2[rt] = Ts[ok2 ? 2 * i2 + off : 0];
                }
                const double s2 = pk ? -sm : sm;
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt) {
                    // row orders of the tile
                    const int rtg = rg * RTW + rt;
                    const int rklo = tri_n(min(8 * rtg, R - 1)),
                              rkhi = tri_n(min(8 * rtg + 7, R - 1));
                    const bool live =
                        rtg < RT && (OPK == 0 || (OPK == 1 ? nlo <= rkhi : nhi >= rklo));
                    double v = (v1[rt] ? c1[rt] : 0.0) + s2 * (v2[rt] ? c2[rt] : 0.0);
                    v = (rowpr[rt] == 0 && pk == 1) ? -v : v;
                    if (live) {
#pragma unroll
                        for (int c = 0; c < NCT; ++c)
                            acc[rt][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(v, b[c], acc[rt][c],
                                                                               0, 0, 0);
                    }
                }
            }
            step = (step + KS) % KG;
        }
    }
    // the sums of a row group's waves, one 16-column tile at a time (in the
    // staging area): D register j of lane i is row 4 j + i / 16, column
    // i % 16 of its tile
    double(*red)[17] = reinterpret_cast<double(*)[17]>(lds);
    double *o = out + ecoef_off(kd, Bx, lv);
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        for (int w = 0; w < KG; ++w) {
            __syncthreads();
            if (kg == w) {
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt) {
                    const int rtg = rg * RTW + rt;
                    if (rtg < RT) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            double &d = red[rtg * 16 + 4 * j + kr][col];
                            d = (w == 0 ? 0.0 : d) + acc[rt][c][j];
                        }
                    }
                }
            }
        }
        __syncthreads();
        for (int q = tid; q < R * 16; q += 256) {
            const int ii = q >> 4, cc = q & 15;
            const int r = c * 16 + cc;
            if (r >= rc) continue;
            const int k = jn[ii];
            const int l = ii - k * (k + 1) / 2;
            const double sc = osc[k];
            const double re = sc * red[2 * ii][cc], im = sc * red[2 * ii + 1][cc];
            double *dst = o + 2 * ((long)(k * k + k + l) * nr + r0 + r);
            double *dm = o + 2 * ((long)(k * k + k - l) * nr + r0 + r);
            const double sl = (l & 1) ? -1.0 : 1.0;
            if (OPK == 1) {
                dst[0] = re;
                dst[1] = im;
                if (l > 0) {
                    dm[0] = sl * re;
                    dm[1] = -sl * im;
                }
            } else {
                dst[0] += re;
                dst[1] += im;
                if (l > 0) {
                    dm[0] += sl * re;
                    dm[1] -= sl * im;
                }
            }
        }
    }
}

// P2P (MODE 0, list 1), M2P (1, list 3), L2P (2) for non-QBX targets
template <int MODE, int NRC>
__global__ void __launch_bounds__(64) k3_eval_targets(KD kd, TD t, int ntb, const int *tboxes,
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
        const v3 c = boxc(t, B);
        for (int f = f0 + lane; f < f0 + nf; f += 64) {
            const int ti = unfilt[f];
            cplx v[NRC];
#pragma unroll
            for (int q = 0; q < NRC; ++q) v[q] = {0, 0};
            sum_regular_conj_n<NRC>(kd, expn + ecoef_off(kd, B, lv), kd.order[lv],
                                    sub(ld3(tgt, nt, ti), c), v);
#pragma unroll
            for (int q = 0; q < NRC; ++q) v[q] = INV4PI * v[q];
            store_add<NRC>(kd, pot + 2 * (size_t)ti * kd.nr, v);
        }
        return;
    }
    int G = 64;
    while (G > 1 && G * nf > 64) G >>= 1;
    const int tpr = 64 / G;
    const int tsub = lane / G, sub_ = lane % G;
    const int e0 = starts[it], e1 = starts[it + 1];
    for (int base = 0; base < nf; base += tpr) {
        const bool active = base + tsub < nf;
        int ti = 0;
        cplx acc[NRC];
#pragma unroll
        for (int q = 0; q < NRC; ++q) acc[q] = {0, 0};
        if (active) {
            ti = unfilt[f0 + base + tsub];
            const v3 tp = ld3(tgt, nt, ti);
            if (MODE == 0) {
                for (int e = e0; e < e1; ++e) {
                    const int Sb = lists[e];
                    const int s0 = t.sstart[Sb], s1 = s0 + t.scount[Sb];
                    for (int s = s0 + sub_; s < s1; s += G) {
                        const v3 nu = kd.dip ? ld3(nrm, ns, s) : v3{0, 0, 0};
                        const double g = p2p_value(kd, ld3(src, ns, s), nu, tp);
#pragma unroll
                        for (int q = 0; q < NRC; ++q)
                            if (q < kd.rc) acc[q] += g * ldq(kd, str, s, kd.r0 + q);
                    }
                }
            } else {
                for (int e = e0 + sub_; e < e1; e += G) {
                    const int D = lists[e];
                    const int lv = t.level[D];
                    sum_irregular_n<NRC>(kd, expn + ecoef_off(kd, D, lv), kd.order[lv],
                                         sub(tp, boxc(t, D)), acc);
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q) acc[q] = INV4PI * acc[q];
            }
        }
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            for (int m = G / 2; m >= 1; m >>= 1) {
                acc[q].x += __shfl_xor(acc[q].x, m);
                acc[q].y += __shfl_xor(acc[q].y, m);
            }
        if (active && sub_ == 0) store_add<NRC>(kd, pot + 2 * (size_t)ti * kd.nr, acc);
    }
}

// ---- QBX stages
template <int PQ, int NRC>
struct QAcc {
    cplx v[NRC][(PQ + 1) * (PQ + 1)];
};

template <int PQ, int NRC>
__device__ void qbx_store(const KD &kd, double *qexp, int c, const QAcc<PQ, NRC> &A,
                          int accumulate)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    double *o = qexp + 2 * (size_t)c * NQ * kd.nr;
#pragma unroll
    for (int i = 0; i < NQ; ++i)
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < kd.rc) {
                double *p = o + 2 * ((long)i * kd.nr + kd.r0 + q);
                cplx v = A.v[q][i];
                if (accumulate) v += ldc(p);
                stc(p, v);
            }
}

// P2QBXL over balanced work units (unit_plan over the global QBX centers of
// each target box): the unit's TG centers take G lanes apiece; the box's
// list-1 sources are staged through LDS in chunks (positions, normals,
// strengths, coalesced) and each center's lanes split a chunk; the expansions
// are reduced by shuffles within the G-lane groups.
template <int NRC>
struct P2QTile3 {
    static constexpr int CH = NRC == 1 ? 256 : 64;
};

template <int PQ, int NRC>
__global__ void __launch_bounds__(64) k3_p2qbxl(KD kd, TD t, const int *ubox, const int *ut0,
                                                const int *box_g, const int *gcs,
                                                const int *gcl, const double *ctr, int nc,
                                                const int *l1s, const int *l1l,
                                                const double *src, int ns, const double *str,
                                                const double *nrm, double *qexp, int accumulate)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    constexpr int CH = P2QTile3<NRC>::CH;
    __shared__ double sx[CH], sy[CH], sz[CH], snx[CH], sny[CH], snz[CH];
    __shared__ double sq[CH * 2 * NRC];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int G = box_g[it];
    const int lane = threadIdx.x;
    const int tsub = lane / G, sub = lane & (G - 1);
    const int cl = ut0[u] + tsub;
    const bool active = cl < gcs[it + 1] - gcs[it];
    const bool dip = kd.dip;
    const int rc = kd.rc, r0 = kd.r0, nr = kd.nr;
    int c = 0;
    v3 cc = {0, 0, 0};
    if (active) {
        c = gcl[gcs[it] + cl];
        cc = ld3(ctr, nc, c);
    }
    QAcc<PQ, NRC> A;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
#pragma unroll
        for (int i = 0; i < NQ; ++i) A.v[q][i] = {0, 0};

    auto flush = [&](int n) {
        __syncthreads();
        if (active) {
            for (int j = sub; j < n; j += G) {
                const v3 d = {sx[j] - cc.x, sy[j] - cc.y, sz[j] - cc.z};
                cplx qs[NRC];
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    qs[q] = q < rc ? cplx{sq[2 * (j * NRC + q)], sq[2 * (j * NRC + q) + 1]}
                                   : cplx{0, 0};
                if (!dip) {
                    cplx T[NQ];
                    the_table<PQ>(d, T);
#pragma unroll
                    for (int i = 0; i < NQ; ++i)
#pragma unroll
                        for (int q = 0; q < NRC; ++q) A.v[q][i] += qs[q] * T[i];
                } else {
                    constexpr int P1 = PQ + 1;
                    cplx T[(P1 + 1) * (P1 + 1)];
                    the_table<P1>(d, T);
                    const cplx nmn = {snx[j], -sny[j]}, npl = {snx[j], sny[j]};
                    const double nz = snz[j];
#pragma unroll
                    for (int n = 0; n <= PQ; ++n)
#pragma unroll
                        for (int m = -n; m <= n; ++m) {
                            const cplx g2 = (-nz) * T[IDX(n + 1, m)] +
                                            0.5 * (nmn * T[IDX(n + 1, m + 1)] -
                                                   npl * T[IDX(n + 1, m - 1)]);
#pragma unroll
                            for (int q = 0; q < NRC; ++q) A.v[q][IDX(n, m)] += qs[q] * g2;
                        }
                }
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
                const int si = s + j, dd = fill + j;
                sx[dd] = src[si];
                sy[dd] = src[(size_t)ns + si];
                sz[dd] = src[2 * (size_t)ns + si];
                if (dip) {
                    snx[dd] = nrm[si];
                    sny[dd] = nrm[(size_t)ns + si];
                    snz[dd] = nrm[2 * (size_t)ns + si];
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q)
                    if (q < rc) {
                        const double *pq = str + 2 * ((size_t)si * nr + r0 + q);
                        sq[2 * (dd * NRC + q)] = pq[0];
                        sq[2 * (dd * NRC + q) + 1] = pq[1];
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
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            for (int m = G / 2; m >= 1; m >>= 1) {
                A.v[q][i].x += __shfl_xor(A.v[q][i].x, m);
                A.v[q][i].y += __shfl_xor(A.v[q][i].y, m);
            }
    if (active && sub == 0) qbx_store<PQ, NRC>(kd, qexp, c, A, accumulate);
}

// Batched P2QBXL (many right-hand sides), a complex contraction per target
// box: for one group of up to GC_GROUP global centers of a box and MR_RB
// right-hand sides, the box's list-1 sources are taken SC at a time; phase A
// builds the basis of every (center, source) pair once into LDS,
// A[k][row] with row = (center, coefficient), and stages the chunk's
// strengths B[k][rhs]; phase B accumulates C[row][rhs] += sum_k A[k][row]
// B[k][rhs] on the FP64 matrix cores (v_mfma_f64_16x16x4f64, 16 x 16 output
// tiles, the complex product as four real ones).  The basis is computed once
// for all right-hand sides instead of once per register chunk of them.
constexpr int MR_RB = 32;

template <int PQ>
struct MRShape {
    static constexpr int NQ = (PQ + 1) * (PQ + 1);
    static constexpr int R = GC_GROUP * NQ;
    static constexpr int SC = PQ <= 3 ? 32 : 16;
};

template <int PQ>
__global__ void __launch_bounds__(256) k3_p2qbxl_mr(KD kd, TD t, const int *ubox,
                                                    const int *uc0, const int *gcs,
                                                    const int *gcl, const double *ctr, int nc,
                                                    const int *l1s, const int *l1l,
                                                    const double *src, int ns, const double *str,
                                                    const double *nrm, double *qexp,
                                                    int accumulate)
{
    using SH = MRShape<PQ>;
    constexpr int NQ = SH::NQ, R = SH::R, SC = SH::SC;
    __shared__ double As[SC * R * 2];
    __shared__ double Bs[SC * MR_RB * 2];
    __shared__ int sidx[SC];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int c0 = uc0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int nr = kd.nr;
    const int rb0 = blockIdx.y * MR_RB;
    const int nrb = min(MR_RB, nr - rb0);
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kr = lane >> 4;
    const bool dip = kd.dip;
    // phase B on the matrix cores: output tiles (16 rows x 16 right-hand
    // sides); wave w takes the row tiles w, w + 4, ... for both column tiles
    constexpr int RT = (R + 15) / 16, RTW = (RT + 3) / 4;
    f3_d4 cre[RTW][2], cim[RTW][2];
#pragma unroll
    for (int j = 0; j < RTW; ++j)
#pragma unroll
        for (int h = 0; h < 2; ++h) cre[j][h] = cim[j][h] = f3_d4{0, 0, 0, 0};

    auto flush = [&](int n) {
        __syncthreads();
        for (int pr = tid; pr < GC_GROUP * n; pr += 256) {
            const int ci = pr / n, k = pr - ci * n;
            if (ci >= ncen) continue;
            const int si = sidx[k];
            const v3 cc = ld3(ctr, nc, gcl[gcs[it] + c0 + ci]);
            const v3 d = sub(ld3(src, ns, si), cc);
            double *a = As + 2 * ((size_t)k * R + ci * NQ);
            if (!dip) {
                cplx T[NQ];
                the_table<PQ>(d, T);
#pragma unroll
                for (int i = 0; i < NQ; ++i) stc(a + 2 * i, T[i]);
            } else {
                constexpr int P1 = PQ + 1;
                cplx T[(P1 + 1) * (P1 + 1)];
                the_table<P1>(d, T);
                const v3 nu = ld3(nrm, ns, si);
                const cplx nmn = {nu.x, -nu.y}, npl = {nu.x, nu.y};
#pragma unroll
                for (int nn = 0; nn <= PQ; ++nn)
#pragma unroll
                    for (int m = -nn; m <= nn; ++m)
                        stc(a + 2 * IDX(nn, m),
                            (-nu.z) * T[IDX(nn + 1, m)] +
                                0.5 * (nmn * T[IDX(nn + 1, m + 1)] - npl * T[IDX(nn + 1, m - 1)]));
            }
        }
        int imag = 0;
        for (int e = tid; e < n * nrb; e += 256) {
            const int k = e / nrb, r = e - k * nrb;
            const double *pq = str + 2 * ((size_t)sidx[k] * nr + rb0 + r);
            Bs[2 * (k * MR_RB + r)] = pq[0];
            Bs[2 * (k * MR_RB + r) + 1] = pq[1];
            imag |= pq[1] != 0.0;
        }
        // real strengths (a real density): the products with the zero
        // imaginary parts are skipped (block-uniform)
        imag = __syncthreads_or(imag);
        // C[row][rhs] += sum_k A[k][row] B[k][rhs]: lane holds A[row0 + col][k]
        // and B[k][16 h + col], k = 4 ks + kr; slots k >= n hold stale data
        // and are masked (rows and columns past the valid ones only reach
        // outputs that are not stored)
        for (int ks = 0; ks < (SC + 3) / 4; ++ks) {
            const int k = ks * 4 + kr;
            const bool kin = k < n;
            double bre[2], bim[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                bre[h] = kin ? Bs[2 * (k * MR_RB + 16 * h + col)] : 0.0;
                bim[h] = kin ? Bs[2 * (k * MR_RB + 16 * h + col) + 1] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < RTW; ++j) {
                const int rtile = wv + 4 * j;
                if (rtile < RT) {
                    const int row = rtile * 16 + col;
                    double are = 0.0, aim = 0.0;
                    if (kin && row < R) {
                        are = As[2 * (k * R + row)];
                        aim = As[2 * (k * R + row) + 1];
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        cre[j][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre[h], cre[j][h], 0, 0, 0);
                        if (imag) {
                            cre[j][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim[h], cre[j][h], 0, 0, 0);
                            cim[j][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim[h], cim[j][h], 0, 0, 0);
                        }
                        cim[j][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre[h], cim[j][h], 0, 0, 0);
                    }
                }
            }
        }
    };

    int fill = 0;
    const int e1 = l1s[it + 1];
    for (int e = l1s[it]; e < e1; ++e) {
        const int Sb = l1l[e];
        int s = t.sstart[Sb];
        const int s1 = s + t.scount[Sb];
        while (s < s1) {
            const int take = min(s1 - s, SC - fill);
            for (int j = tid; j < take; j += 256) sidx[fill + j] = s + j;
            fill += take;
            s += take;
            if (fill == SC) {
                flush(SC);
                fill = 0;
            }
        }
    }
    if (fill > 0) flush(fill);
#pragma unroll
    for (int j = 0; j < RTW; ++j) {
        const int rtile = wv + 4 * j;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int row = rtile * 16 + 4 * g + kr;  // D row of register g
                const int r = 16 * h + col;
                if (rtile >= RT || row >= R || r >= nrb) continue;
                const int ci = row / NQ, i = row - ci * NQ;
                if (ci >= ncen) continue;
                const int c = gcl[gcs[it] + c0 + ci];
                const cplx v = {cre[j][h][g], cim[j][h][g]};
                double *pq = qexp + 2 * (((size_t)c * NQ + i) * nr + rb0 + r);
                stc(pq, accumulate ? ldc(pq) + v : v);
            }
    }
}

constexpr int M2Q_LANES = 16;

// A lane group lives inside one wavefront: order its LDS writes before the
// group's reads (and reads before the next overwrite) without a block barrier,
// which the groups' different trip counts would make unsafe.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// M2QBXL: 16 lanes per global center; per list-3 box the irregular harmonics
// T_N^M(c - b), N <= p_b + PQ, are built in LDS (one column m >= 0 per lane),
// then lane j contracts output coefficients j, j + 16, ...
template <int PQ, int NRC>
__global__ void __launch_bounds__(256) k3_m2qbxl(KD kd, TD t, int ng, const int *gcent,
                                                 const int *c2tb, const double *ctr, int nc,
                                                 const int *l3s, const int *l3l,
                                                 const double *mpole, double *qexp)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    constexpr int NO = (NQ + M2Q_LANES - 1) / M2Q_LANES;
    extern __shared__ double lds[];
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = gid / M2Q_LANES, lane = gid % M2Q_LANES;
    const int slot = threadIdx.x / M2Q_LANES;
    const int pn = kd.pmax + PQ;
    const int ntab = (pn + 1) * (pn + 1);
    double *T = lds + 2L * slot * ntab;
    const bool active = g < ng;
    const int c = active ? gcent[g] : 0;
    const int it = active ? c2tb[c] : 0;
    const v3 cc = active ? ld3(ctr, nc, c) : v3{0, 0, 0};
    cplx acc[NO][NRC];
#pragma unroll
    for (int r = 0; r < NO; ++r)
#pragma unroll
        for (int q = 0; q < NRC; ++q) acc[r][q] = {0, 0};
    const int e0 = active ? l3s[it] : 0, e1 = active ? l3s[it + 1] : 0;
    for (int e = e0; e < e1; ++e) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int p1 = kd.order[lv];
        const int N1 = p1 + PQ;
        const v3 d = sub(cc, boxc(t, D));
        const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        const double ir2 = 1.0 / r2;
        const cplx w = {d.x * ir2, d.y * ir2};
        // column m = lane, lane + 16, ...: T_m^m by its product formula, then up in n
        for (int m = lane; m <= N1; m += M2Q_LANES) {
            cplx tmm = {sqrt(ir2), 0};
            for (int i = 1; i <= m; ++i) tmm = (-(2.0 * i - 1.0)) * (tmm * w);
            cplx a = {0, 0}, b = tmm;
            for (int n = m; n <= N1; ++n) {
                if (n > m) {
                    const cplx cn = ir2 * (((2.0 * n - 1.0) * d.z) * b -
                                           ((double)(n + m - 1) * (n - m - 1)) * a);
                    a = b;
                    b = cn;
                }
                stc(T + 2 * IDX(n, m), b);
                if (m > 0) stc(T + 2 * IDX(n, -m), neg_m(b, m));
            }
        }
        wave_sync();
        const double *a = mpole + ecoef_off(kd, D, lv);
#pragma unroll
        for (int r = 0; r < NO; ++r) {
            const int qo = lane + r * M2Q_LANES;
            if (qo < NQ) {
                int k, l;
                nm_of(qo, k, l);
                cplx s[NRC];
#pragma unroll
                for (int q = 0; q < NRC; ++q) s[q] = {0, 0};
                for (int n = 0; n <= p1; ++n) {
                    const int N = n + k;
                    for (int m = -n; m <= n; ++m) {
                        const int M = m + l;
                        if (M < -N || M > N) continue;
                        const cplx tv = ldc(T + 2 * IDX(N, M));
#pragma unroll
                        for (int q = 0; q < NRC; ++q)
                            if (q < kd.rc) s[q] += ldcn(kd, a, IDX(n, m), kd.r0 + q) * tv;
                    }
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q) acc[r][q] += (k & 1) ? -s[q] : s[q];
            }
        }
        wave_sync();
    }
    if (active) {
#pragma unroll
        for (int r = 0; r < NO; ++r) {
            const int qo = lane + r * M2Q_LANES;
            if (qo < NQ) store_add<NRC>(kd, qexp + 2 * ((size_t)c * NQ + qo) * kd.nr, acc[r]);
        }
    }
}

// M2QBXL for many right-hand sides on the FP64 matrix cores: a block per
// group of GC_GROUP global centers of one target box (gc_group_plan), wave w
// takes center w of the group and the 32 right-hand sides of blockIdx.y.
// Per list-3 box D of the target box the wave builds T_N^M(c - b_D),
// N <= p_D + PQ, in its own LDS slice (as k3_m2qbxl, one column m per lane),
// then the complex product out[(k,l)][r] += sum_(n,m) T(n+k, m+l)
// a_D[(n,m)][r] runs as four real v_mfma_f64_16x16x4f64 per K-step of four
// (n,m) (lane i holds A[i%16][k0 + i/16] and B[k0 + i/16][i%16]; D register
// j of lane i is D[4j + i/16][i%16]): rows (k,l) in 16-row tiles, columns
// two tiles of 16 right-hand sides, multipoles read straight from HBM (the
// group's waves read the same box).  The epilogue applies (-1)^k and adds
// into qexp.  Waves only touch their own LDS slice: no block barrier.
constexpr int M2QM_RB = 32;  // right-hand sides per block (blockIdx.y)

template <int PQ>
__global__ void __launch_bounds__(256) k3_m2qbxl_mfma(KD kd, TD t, const int *ubox,
                                                      const int *uc0, const int *gcs,
                                                      const int *gcl, const double *ctr, int nc,
                                                      const int *l3s, const int *l3l,
                                                      const double *mpole, double *qexp)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    constexpr int RT = (NQ + 15) / 16;
    extern __shared__ double lds[];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int c0 = uc0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    if (wv >= ncen) return;  // no barriers below: idle waves leave at once
    const int col = lane & 15, kr = lane >> 4;
    const int nr = kd.nr;
    const int rb0 = blockIdx.y * M2QM_RB;
    const int nrb = min(M2QM_RB, nr - rb0);
    const int c = gcl[gcs[it] + c0 + wv];
    const v3 cc = ld3(ctr, nc, c);
    const int pn = kd.pmax + PQ;
    double *T = lds + 2L * wv * (pn + 1) * (pn + 1);
    // the (k, l) of this lane's A rows, tile by tile
    int rk[RT], rl[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) {
        const int row = q * 16 + col;
        if (row < NQ)
            nm_of(row, rk[q], rl[q]);
        else
            rk[q] = rl[q] = 0;
    }
    f3_d4 cre[RT][2], cim[RT][2];
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) cre[q][h] = cim[q][h] = f3_d4{0, 0, 0, 0};
    const int e0 = l3s[it], e1 = l3s[it + 1];
    for (int e = e0; e < e1; ++e) {
        const int D = l3l[e];
        const int lv = t.level[D];
        const int p1 = kd.order[lv];
        const int N1 = p1 + PQ;
        const int ncf = (p1 + 1) * (p1 + 1);
        const v3 d = sub(cc, boxc(t, D));
        const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        const double ir2 = 1.0 / r2;
        const cplx w = {d.x * ir2, d.y * ir2};
        for (int m = lane; m <= N1; m += 64) {
            cplx tmm = {sqrt(ir2), 0};
            for (int i = 1; i <= m; ++i) tmm = (-(2.0 * i - 1.0)) * (tmm * w);
            cplx a = {0, 0}, b = tmm;
            for (int n = m; n <= N1; ++n) {
                if (n > m) {
                    const cplx cn = ir2 * (((2.0 * n - 1.0) * d.z) * b -
                                           ((double)(n + m - 1) * (n - m - 1)) * a);
                    a = b;
                    b = cn;
                }
                stc(T + 2 * IDX(n, m), b);
                if (m > 0) stc(T + 2 * IDX(n, -m), neg_m(b, m));
            }
        }
        wave_sync();
        const double *am = mpole + ecoef_off(kd, D, lv);
        const int KS = (ncf + 3) / 4;
        for (int ks = 0; ks < KS; ++ks) {
            const int kq = ks * 4 + kr;  // (n, m) of the A column and B row
            int n = 0, m = 0;
            const bool kin = kq < ncf;
            if (kin) nm_of(kq, n, m);
            double bre[2], bim[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int rr = h * 16 + col;
                bre[h] = bim[h] = 0.0;
                if (kin && rr < nrb) {
                    const cplx bv = ldcn(kd, am, kq, rb0 + rr);
                    bre[h] = bv.x;
                    bim[h] = bv.y;
                }
            }
#pragma unroll
            for (int q = 0; q < RT; ++q) {
                const int row = q * 16 + col;
                double are = 0.0, aim = 0.0;
                if (kin && row < NQ) {
                    const int N = n + rk[q], M = m + rl[q];
                    if (M >= -N && M <= N) {
                        are = T[2 * IDX(N, M)];
                        aim = T[2 * IDX(N, M) + 1];
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    cre[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre[h], cre[q][h], 0, 0, 0);
                    cre[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim[h], cre[q][h], 0, 0, 0);
                    cim[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim[h], cim[q][h], 0, 0, 0);
                    cim[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre[h], cim[q][h], 0, 0, 0);
                }
            }
        }
        wave_sync();  // T read before the next box overwrites it
    }
    double *o = qexp + 2 * (size_t)c * NQ * nr;
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = q * 16 + 4 * j + kr;
            if (row < NQ) {
                int k, l;
                nm_of(row, k, l);
                const double sg = (k & 1) ? -1.0 : 1.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rr = h * 16 + col;
                    if (rr < nrb) {
                        double *dst = o + 2 * ((long)row * nr + rb0 + rr);
                        dst[0] += sg * cre[q][h][j];
                        dst[1] += sg * cim[q][h][j];
                    }
                }
            }
        }
}

// L2QBXL for many right-hand sides on the FP64 matrix cores: a block per
// group of GC_GROUP global centers of one target box (they share the box's
// local expansion L), wave w takes center w and the 32 right-hand sides of
// blockIdx.y.  The wave builds Y_j^i(c - b), j <= p (the recurrences of
// k3_l2qbxl, one column i per lane, Y_j^-i = (-1)^i conj(Y_j^i)) in its LDS
// slice; then out[(a,b)][r] += sum_(n,m) conj(Y_(n-a)^(m-b)) L[(n,m)][r]
// (zero unless n >= a and |m - b| <= n - a) as four real
// v_mfma_f64_16x16x4f64 per K-step (lane layout as k3_m2qbxl_mfma).
template <int PQ>
__global__ void __launch_bounds__(256) k3_l2qbxl_mfma(KD kd, TD t, const int *ubox,
                                                      const int *uc0, const int *gcs,
                                                      const int *gcl, const int *tboxes,
                                                      const double *ctr, int nc,
                                                      const double *local, double *qexp)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    constexpr int RT = (NQ + 15) / 16;
    extern __shared__ double lds[];
    const int u = blockIdx.x;
    const int it = ubox[u];
    const int c0 = uc0[u];
    const int ncen = min(GC_GROUP, gcs[it + 1] - gcs[it] - c0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    if (wv >= ncen) return;  // no barriers below: idle waves leave at once
    const int col = lane & 15, kr = lane >> 4;
    const int nr = kd.nr;
    const int rb0 = blockIdx.y * M2QM_RB;
    const int nrb = min(M2QM_RB, nr - rb0);
    const int c = gcl[gcs[it] + c0 + wv];
    const int Bx = tboxes[it];
    const int lv = t.level[Bx];
    const int p1 = kd.order[lv];
    const int ncf = (p1 + 1) * (p1 + 1);
    double *Y = lds + 2L * wv * (kd.pmax + 1) * (kd.pmax + 1);
    const v3 d = sub(ld3(ctr, nc, c), boxc(t, Bx));
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const cplx hw = {-0.5 * d.x, -0.5 * d.y};
    for (int i = lane; i <= p1; i += 64) {
        cplx ymm = {1, 0};
        for (int q = 1; q <= i; ++q) ymm = (1.0 / q) * (ymm * hw);
        cplx ya = {0, 0}, yb = ymm;
        for (int j = i; j <= p1; ++j) {
            if (j > i) {
                const cplx cn = (1.0 / ((double)(j - i) * (j + i))) *
                                (((2.0 * j - 1.0) * d.z) * yb - r2 * ya);
                ya = yb;
                yb = cn;
            }
            stc(Y + 2 * IDX(j, i), yb);
            if (i > 0) stc(Y + 2 * IDX(j, -i), (i & 1) ? -conj(yb) : conj(yb));
        }
    }
    wave_sync();
    int ra[RT], rb[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) {
        const int row = q * 16 + col;
        if (row < NQ)
            nm_of(row, ra[q], rb[q]);
        else
            ra[q] = rb[q] = 0;
    }
    f3_d4 cre[RT][2], cim[RT][2];
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) cre[q][h] = cim[q][h] = f3_d4{0, 0, 0, 0};
    const double *L = local + ecoef_off(kd, Bx, lv);
    const int KS = (ncf + 3) / 4;
    for (int ks = 0; ks < KS; ++ks) {
        const int kq = ks * 4 + kr;
        int n = 0, m = 0;
        const bool kin = kq < ncf;
        if (kin) nm_of(kq, n, m);
        double bre[2], bim[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int rr = h * 16 + col;
            bre[h] = bim[h] = 0.0;
            if (kin && rr < nrb) {
                const cplx bv = ldcn(kd, L, kq, rb0 + rr);
                bre[h] = bv.x;
                bim[h] = bv.y;
            }
        }
#pragma unroll
        for (int q = 0; q < RT; ++q) {
            const int row = q * 16 + col;
            double are = 0.0, aim = 0.0;
            const int j = n - ra[q], i = m - rb[q];
            if (kin && row < NQ && j >= 0 && i >= -j && i <= j) {
                are = Y[2 * IDX(j, i)];
                aim = -Y[2 * IDX(j, i) + 1];  // conj
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                cre[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre[h], cre[q][h], 0, 0, 0);
                cre[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim, bim[h], cre[q][h], 0, 0, 0);
                cim[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim[h], cim[q][h], 0, 0, 0);
                cim[q][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre[h], cim[q][h], 0, 0, 0);
            }
        }
    }
    double *o = qexp + 2 * (size_t)c * NQ * nr;
#pragma unroll
    for (int q = 0; q < RT; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = q * 16 + 4 * j + kr;
            if (row < NQ) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int rr = h * 16 + col;
                    if (rr < nrb) {
                        double *dst = o + 2 * ((long)row * nr + rb0 + rr);
                        dst[0] += cre[q][h][j];
                        dst[1] += cim[q][h][j];
                    }
                }
            }
        }
}

// L2QBXL: one thread per global center
template <int PQ, int NRC>
__global__ void k3_l2qbxl(KD kd, TD t, int ng, const int *gcent, const int *c2tb,
                          const int *tboxes, const double *ctr, int nc, const double *local,
                          double *qexp)
{
    constexpr int NQ = (PQ + 1) * (PQ + 1);
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int c = gcent[g];
    const int Bx = tboxes[c2tb[c]];
    const int lv = t.level[Bx];
    const int p1 = kd.order[lv];
    const v3 d = sub(ld3(ctr, nc, c), boxc(t, Bx));
    const double *L = local + ecoef_off(kd, Bx, lv);
    QAcc<PQ, NRC> A;
#pragma unroll
    for (int q = 0; q < NRC; ++q)
#pragma unroll
        for (int i = 0; i < NQ; ++i) A.v[q][i] = {0, 0};
    const double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    const cplx hw = {-0.5 * d.x, -0.5 * d.y};
    cplx ymm = {1, 0};
    for (int i = 0; i <= p1; ++i) {        // column i >= 0 of Y_j^i, j = i..p1
        if (i > 0) ymm = (1.0 / i) * (ymm * hw);
        cplx ya = {0, 0}, yb = ymm;
        for (int j = i; j <= p1; ++j) {
            if (j > i) {
                const cplx cn = (1.0 / ((double)(j - i) * (j + i))) *
                                (((2.0 * j - 1.0) * d.z) * yb - r2 * ya);
                ya = yb;
                yb = cn;
            }
            const cplx yc = conj(yb);                  // conj(Y_j^i)
            const cplx yn = (i & 1) ? -yb : yb;        // conj(Y_j^{-i}) = (-1)^i Y_j^i
#pragma unroll
            for (int a = 0; a <= PQ; ++a) {
                if (a + j > p1) continue;
#pragma unroll
                for (int b = -a; b <= a; ++b) {
                    const int n = a + j;
                    const int m1 = b + i;
                    const int m2 = b - i;
#pragma unroll
                    for (int q = 0; q < NRC; ++q) {
                        if (q >= kd.rc) continue;
                        if (m1 >= -n && m1 <= n)
                            A.v[q][IDX(a, b)] += ldcn(kd, L, IDX(n, m1), kd.r0 + q) * yc;
                        if (i > 0 && m2 >= -n && m2 <= n)
                            A.v[q][IDX(a, b)] += ldcn(kd, L, IDX(n, m2), kd.r0 + q) * yn;
                    }
                }
            }
        }
    }
    qbx_store<PQ, NRC>(kd, qexp, c, A, 1);
}

// QBXL2P: one thread per (center, target) entry
template <int NRC>
__global__ void k3_qbxl2p(KD kd, int nctt, const int *ctt_lists, const int *tree_ttc,
                          const double *tgt, int nt, const double *ctr, int nc,
                          const double *qexp, double *pot)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nctt) return;
    const int ti = ctt_lists[j];
    const int c = tree_ttc[ti];
    const v3 d = sub(ld3(tgt, nt, ti), ld3(ctr, nc, c));
    cplx v[NRC];
#pragma unroll
    for (int q = 0; q < NRC; ++q) v[q] = {0, 0};
    sum_regular_conj_n<NRC>(kd, qexp + 2 * (size_t)c * kd.ncq * kd.nr, kd.pq, d, v);
#pragma unroll
    for (int q = 0; q < NRC; ++q)
        if (q < kd.rc) stc(pot + 2 * ((size_t)ti * kd.nr + kd.r0 + q), INV4PI * v[q]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using zc = std::complex<double>;

static void host_irregular(int p, const double *x, zc *T)
{
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double ir2 = 1.0 / r2;
    const zc w(x[0] * ir2, x[1] * ir2);
    zc tmm(std::sqrt(ir2), 0.0);
    for (int m = 0; m <= p; ++m) {
        if (m > 0) tmm = -(2.0 * m - 1.0) * w * tmm;
        zc a = 0.0, b = tmm;
        T[IDX(m, m)] = tmm;
        for (int n = m + 1; n <= p; ++n) {
            zc c = ir2 * ((2.0 * n - 1.0) * x[2] * b - (double)(n + m - 1) * (n - m - 1) * a);
            T[IDX(n, m)] = c;
            a = b;
            b = c;
        }
    }
    for (int n = 1; n <= p; ++n)
        for (int m = 1; m <= n; ++m) T[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * std::conj(T[IDX(n, m)]);
}

static void host_regular(int p, const double *x, zc *Y)
{
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const zc hw(-0.5 * x[0], -0.5 * x[1]);
    zc ymm = 1.0;
    for (int m = 0; m <= p; ++m) {
        if (m > 0) ymm = (1.0 / m) * hw * ymm;
        zc a = 0.0, b = ymm;
        Y[IDX(m, m)] = ymm;
        for (int n = m + 1; n <= p; ++n) {
            zc c = ((2.0 * n - 1.0) * x[2] * b - r2 * a) / ((double)(n - m) * (n + m));
            Y[IDX(n, m)] = c;
            a = b;
            b = c;
        }
    }
    for (int n = 1; n <= p; ++n)
        for (int m = 1; m <= n; ++m) Y[IDX(n, -m)] = ((m & 1) ? -1.0 : 1.0) * std::conj(Y[IDX(n, m)]);
}

static int make_kd(const qbx_kernel_desc *d, const qbx_tree *T, KD *kd, int nsep = 2)
{
    QBX_CHECK(d, "null kernel description");
    QBX_CHECK(!d->target_deriv, "3D stages: target derivatives are not supported (the "
                                "target-specific stage gives 3D gradients)");
    QBX_CHECK(d->kernel == QBX_LAPLACE3D, "3D stages support QBX_LAPLACE3D only (got %d)",
              d->kernel);
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
    memset(kd, 0, sizeof(*kd));
    kd->dip = d->source_kind == QBX_SOURCE_D;
    kd->pq = d->qbx_order;
    kd->nsep = nsep;
    kd->W = 2 * nsep + 1;
    kd->nr = d->nrhs > 0 ? d->nrhs : 1;
    kd->r0 = 0;
    kd->rc = 1;
    const int nl = T ? T->nlevels : std::min(d->nlevels, FML);
    kd->nlevels = nl;
    long coef = 0;
    int pmax = 0;
    for (int l = 0; l < nl; ++l) {
        const int p = d->level_orders[l];
        QBX_CHECK(p >= 1 && p <= MAXP, "3D fmm order %d at level %d outside [1, %d]", p, l, MAXP);
        kd->order[l] = p;
        pmax = std::max(pmax, p);
        kd->ncl[l] = (p + 1) * (p + 1);
        kd->h[l] = T ? ldexp(T->root_extent, -l) : 1.0;
        kd->lbox[l] = T ? T->level_start_box_nrs[l] : 0;
        kd->lcoef[l] = coef;
        if (T)
            coef += (long)(T->level_start_box_nrs[l + 1] - T->level_start_box_nrs[l]) * kd->ncl[l];
    }
    kd->lbox[nl] = T ? T->nboxes : 0;
    kd->lcoef[nl] = coef;
    kd->pmax = pmax;
    kd->tabn = 2 * pmax + 1;
    kd->ncq = (kd->pq + 1) * (kd->pq + 1);
    return QBX_OK;
}

// unit-offset tables, cached in the context per (nsep, pmax)
static int ensure_tables(qbx_ctx *ctx, hipStream_t st, KD *kd)
{
    std::vector<double> key = {3.0, (double)kd->nsep, (double)kd->pmax};
    QBX_TRY(note_capture(ctx, st));
    if (key != ctx->tab3_key) {
        // a graph replays the tables it captured: never rebuild under one
        QBX_TRY(may_replace(ctx, st, "3D translation tables"));
        const int W = kd->W, S1 = 2 * W + 1, tn = kd->tabn;
        const long ncls = (long)S1 * S1 * S1;
        std::vector<zc> tab((size_t)ncls * tn * tn, zc(0.0, 0.0));
        for (int ix = -W; ix <= W; ++ix)
            for (int iy = -W; iy <= W; ++iy)
                for (int iz = -W; iz <= W; ++iz) {
                    if (!ix && !iy && !iz) continue;
                    const double x[3] = {(double)ix, (double)iy, (double)iz};
                    const long cl = ((long)(ix + W) * S1 + (iy + W)) * S1 + (iz + W);
                    host_irregular(2 * kd->pmax, x, tab.data() + cl * tn * tn);
                }
        const int nc1 = (kd->pmax + 1) * (kd->pmax + 1);
        std::vector<zc> ch((size_t)8 * nc1);
        for (int ci = 0; ci < 8; ++ci) {
            const double x[3] = {(ci & 1) ? 0.5 : -0.5, (ci & 2) ? 0.5 : -0.5,
                                 (ci & 4) ? 0.5 : -0.5};
            host_regular(kd->pmax, x, ch.data() + (size_t)ci * nc1);
        }
        QBX_TRY(ctx->ws_tab3_m2l.alloc(2L * (long)tab.size()));
        QBX_TRY(ctx->ws_tab3_ch.alloc(2L * (long)ch.size()));
        QBX_HIP(hipMemcpy(ctx->ws_tab3_m2l.p, tab.data(), sizeof(zc) * tab.size(),
                          hipMemcpyHostToDevice));
        QBX_HIP(hipMemcpy(ctx->ws_tab3_ch.p, ch.data(), sizeof(zc) * ch.size(),
                          hipMemcpyHostToDevice));
        ctx->tab3_key = key;
    }
    kd->tab_m2l = ctx->ws_tab3_m2l.p;
    kd->tab_ch = ctx->ws_tab3_ch.p;
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

static Flat make_flat(const KD &kd, const std::vector<int> &level_starts)
{
    Flat f;
    memset(&f, 0, sizeof(f));
    f.nl = kd.nlevels;
    long t = 0;
    for (int l = 0; l < kd.nlevels; ++l) {
        f.tstart[l] = t;
        f.istart[l] = level_starts[l];
        f.nper[l] = kd.ncl[l];
        t += (long)(level_starts[l + 1] - level_starts[l]) * f.nper[l];
    }
    f.tstart[kd.nlevels] = t;
    return f;
}

// Run LAUNCH(N) over the right-hand-side chunks: N = 1 for a single
// right-hand side, CH otherwise; kd.r0 / kd.rc are set per chunk.
#define FOR_RHS_CHUNKS3(kd, CH, LAUNCH)                                                 \
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

#define DISPATCH3_PQ_NR(pq, NR, KERNEL, ...)                                            \
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
    default: set_error("qbx_order out of range"); return QBX_ERR_INVALID;               \
    }

#define DISPATCH3_NO_NR(no, NR, KERNEL, ...)                                            \
    switch (no) {                                                                        \
    case 1: hipLaunchKernelGGL((KERNEL<1, NR>), __VA_ARGS__); break;                     \
    case 2: hipLaunchKernelGGL((KERNEL<2, NR>), __VA_ARGS__); break;                     \
    case 3: hipLaunchKernelGGL((KERNEL<3, NR>), __VA_ARGS__); break;                     \
    case 4: hipLaunchKernelGGL((KERNEL<4, NR>), __VA_ARGS__); break;                     \
    case 5: hipLaunchKernelGGL((KERNEL<5, NR>), __VA_ARGS__); break;                     \
    case 6: hipLaunchKernelGGL((KERNEL<6, NR>), __VA_ARGS__); break;                     \
    case 7: hipLaunchKernelGGL((KERNEL<7, NR>), __VA_ARGS__); break;                     \
    default: set_error("M2L: order out of range"); return QBX_ERR_INVALID;              \
    }

static inline long exp_bytes(const KD &kd) { return 16L * kd.lcoef[kd.nlevels] * kd.nr; }

}  // namespace f3

using namespace f3;

long f3_expansion_size(const qbx_kernel_desc *d, const qbx_tree *T)
{
    KD kd;
    if (!T || make_kd(d, T, &kd) != QBX_OK) return -1;
    return kd.lcoef[kd.nlevels] * kd.nr;
}

int f3_expansion_level_offsets(const qbx_kernel_desc *d, const qbx_tree *T, long *offsets)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    for (int l = 0; l <= kd.nlevels; ++l) offsets[l] = kd.lcoef[l] * kd.nr;
    return QBX_OK;
}

int f3_ncoeff_qbx(const qbx_kernel_desc *d)
{
    KD kd;
    if (make_kd(d, nullptr, &kd) != QBX_OK) return -1;
    return kd.ncq;
}

// Shape of k3_p2e at order p: S sources per chunk in 32 KiB of LDS, threads
// covering max((p+1)^2, S) (at most 512).  False: the plain kernels.
static bool p2e3_launch(int p, bool local, P2EShape3 *sh, int *threads, size_t *bytes)
{
    const int pt = local ? p + 1 : p;
    P2EShape3 h;
    h.R = ((pt + 1) * (pt + 2) / 2) | 1;
    const long per = 8L * (2 * h.R + 2 * NRC_S + 3 + 1);
    h.S = (int)std::min(64L, 32768L / per);
    const int nout = (p + 1) * (p + 1);
    if (h.S < 4 || nout > 512) return false;
    *sh = h;
    *threads = 64 * ((std::max(nout, h.S) + 63) / 64);
    *bytes = 8UL * h.doubles(NRC_S);
    return true;
}

// dynamic LDS of k3_p2e_mfma for table row stride R (complex)
static size_t p2e3m_bytes(int R)
{
    return 8UL * (2 * P2E3M_S * R + 2 * P2E3M_S * 16 + 3 * P2E3M_S) + 4UL * P2E3M_S;
}

static bool p2e3_all_levels(const KD &kd, int nlevels, bool local)
{
    for (int l = 0; l < nlevels; ++l) {
        P2EShape3 sh;
        int th;
        size_t by;
        if (!p2e3_launch(kd.order[l], local, &sh, &th, &by)) return false;
    }
    return true;
}

int f3_form_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                       const qbx_kernel_desc *d, const double *str, const double *nrm,
                       double *mpole)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    QBX_TRY(zero_fill(mpole, (exp_bytes(kd)) / 8, st));
    if (p2e3_all_levels(kd, T->nlevels, false)) {
        const TD td = make_td(T);
        for (int l = 0; l < T->nlevels; ++l) {
            const int i0 = R->h_level_start_source_box_nrs[l];
            const int i1 = R->h_level_start_source_box_nrs[l + 1];
            if (i1 == i0) continue;
            P2EShape3 sh;
            int th;
            size_t by;
            p2e3_launch(kd.order[l], false, &sh, &th, &by);
            if (kd.nr >= 4 && (kd.order[l] + 1) * (kd.order[l] + 1) <= 16 * 4 * P2E3M_RTW) {
                for (int r0 = 0; r0 < kd.nr; r0 += 16) {
                    kd.r0 = r0;
                    kd.rc = std::min(16, kd.nr - r0);
                    hipLaunchKernelGGL(k3_p2e_mfma<false>, dim3(i1 - i0), dim3(256),
                                       p2e3m_bytes(sh.R), st, kd, td, l, i0, sh.R,
                                       R->source_boxes.p, nullptr, nullptr, T->sources.p,
                                       T->nsources, str, nrm, mpole);
                }
                continue;
            }
#define L3_P2MH(N)                                                                         \
    hipLaunchKernelGGL((k3_p2e<N, false>), dim3(i1 - i0), dim3(th), by, st, kd, td, l, i0, sh, \
                       R->source_boxes.p, nullptr, nullptr, T->sources.p, T->nsources, str,  \
                       nrm, mpole)
            FOR_RHS_CHUNKS3(kd, NRC_S, L3_P2MH);
#undef L3_P2MH
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    const Flat fl = make_flat(kd, R->h_level_start_source_box_nrs);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0) {
        const TD td = make_td(T);
#define L3_P2M(N)                                                                          \
    hipLaunchKernelGGL(k3_p2m<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, fl,     \
                       R->source_boxes.p, T->sources.p, T->nsources, str, nrm, mpole)
        FOR_RHS_CHUNKS3(kd, NRC_S, L3_P2M);
#undef L3_P2M
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// Launch k3_trans_real<OPK, p> over nb output boxes of level l (orders p <= 15;
// 32 right-hand sides per block, 16 for M2M and L2L above order 9).
template <int OPK>
static int launch_trans_real(hipStream_t st, KD kd, const TD &td, int l, int p, int i0,
                             int nb, const int *boxes, const int *starts, const int *lists,
                             const double *in, double *out, const int *allreal)
{
    const int rcb = (OPK != 0 && p >= 10) ? 16 : 32;
    for (int r0 = 0; r0 < kd.nr; r0 += rcb) {
        kd.r0 = r0;
        kd.rc = std::min(rcb, kd.nr - r0);
#define TR_CASE(PP)                                                                        \
    case PP:                                                                               \
        hipLaunchKernelGGL((k3_trans_real<OPK, PP, (OPK != 0 && PP >= 10) ? 1 : 2>), dim3(nb), \
                           dim3(256), 0, st, kd, td,                                       \
                           l, i0, boxes, starts, lists, in, out, allreal);                 \
        break;
        switch (p) {
            TR_CASE(1) TR_CASE(2) TR_CASE(3) TR_CASE(4) TR_CASE(5) TR_CASE(6) TR_CASE(7)
            TR_CASE(8) TR_CASE(9) TR_CASE(10) TR_CASE(11) TR_CASE(12) TR_CASE(13)
            TR_CASE(14) TR_CASE(15)
        default: set_error("real-strength translation: order %d out of range", p);
            return QBX_ERR_INVALID;
        }
#undef TR_CASE
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

// Which kernels run a box translation: the real-strength matrix-core kernel
// (stated by the caller, or decided on the device by *flag), the complex one,
// or both, each returning at once in the other's case.
struct RealSel {
    bool real, complex_;
    const int *flag;  // passed to both kernels (null: no device decision)
};

static RealSel real_sel(const qbx_kernel_desc *d, const KD &kd, const int *flag, bool fits)
{
    const bool ok = fits && kd.nr >= 4;
    if (ok && d->real_strengths) return {true, false, nullptr};
    if (ok && flag) return {true, true, flag};
    return {false, true, nullptr};
}

int f3_coarsen_multipoles(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, double *mpole, int lmin, const int *allreal)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    QBX_TRY(ensure_tables(ctx, S(stream), &kd));
    hipStream_t st = S(stream);
    const TD td = make_td(T);
    for (int l = T->nlevels - 2; l >= std::max(lmin, 0); --l) {
        const int i0 = R->h_level_start_source_parent_box_nrs[l];
        const int i1 = R->h_level_start_source_parent_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr > 0) {
            // the real-strength kernel where parent and child orders agree
            const int p = kd.order[l];
            const RealSel rs = real_sel(d, kd, allreal, p <= 15 && kd.order[l + 1] == p);
            if (rs.real)
                QBX_TRY(launch_trans_real<1>(st, kd, td, l, p, i0, i1 - i0,
                                             R->source_parent_boxes.p, nullptr, nullptr, mpole,
                                             mpole, rs.flag));
            if (rs.complex_) {
#define L3_M2M(N)                                                                          \
    hipLaunchKernelGGL(k3_m2m<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, l, i0,  \
                       i1, R->source_parent_boxes.p, mpole, rs.flag)
                FOR_RHS_CHUNKS3(kd, NRC_S, L3_M2M);
#undef L3_M2M
            }
        }
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_eval_targets(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                    const qbx_geo *G, const qbx_kernel_desc *d, int mode, const double *str,
                    const double *nrm, const double *expn, double *pot, int accumulate)
{
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
#define EVAL3_LAUNCH(M, N)                                                                 \
    hipLaunchKernelGGL((k3_eval_targets<M, N>), dim3(ntb), dim3(64), 0, st, kd, td, ntb,   \
                       R->target_boxes.p, starts, lists, G->nqbtl_box_target_starts.p,    \
                       G->nqbtl_box_target_counts_nonchild.p, G->nqbtl_unfiltered.p,      \
                       T->targets.p, T->ntargets, T->sources.p, T->nsources, str, nrm,    \
                       expn, pot)
#define EV30(N) EVAL3_LAUNCH(0, N)
#define EV31(N) EVAL3_LAUNCH(1, N)
#define EV32(N) EVAL3_LAUNCH(2, N)
    if (mode == 0) FOR_RHS_CHUNKS3(kd, NRC_S, EV30);
    else if (mode == 1) FOR_RHS_CHUNKS3(kd, NRC_S, EV31);
    else FOR_RHS_CHUNKS3(kd, NRC_S, EV32);
#undef EV30
#undef EV31
#undef EV32
#undef EVAL3_LAUNCH
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_multipole_to_local(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                          const qbx_kernel_desc *d, const double *mpole, double *local,
                          const int *allreal)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    QBX_TRY(ensure_tables(ctx, S(stream), &kd));
    hipStream_t st = S(stream);
    QBX_TRY(zero_fill(local, (exp_bytes(kd)) / 8, st));
    const RealSel rs = real_sel(d, kd, allreal, kd.pmax <= 15);
    if (R->nttp_boxes > 0 && R->n2 > 0 && kd.nr >= 4 && kd.pmax <= 15) {
        // many right-hand sides: on the matrix cores, 16 per launch (complex)
        const TD td = make_td(T);
        if (rs.complex_) {
            Flat fl;
            fl.nl = kd.nlevels;
            long nb = 0;
            for (int l = 0; l < kd.nlevels; ++l) {
                fl.tstart[l] = nb;
                fl.istart[l] = R->h_level_start_ttp_box_nrs[l];
                fl.nper[l] = 1;
                nb += R->h_level_start_ttp_box_nrs[l + 1] - R->h_level_start_ttp_box_nrs[l];
            }
            fl.tstart[kd.nlevels] = nb;
            for (int r0 = 0; r0 < kd.nr; r0 += 16) {
                kd.r0 = r0;
                kd.rc = std::min(16, kd.nr - r0);
                hipLaunchKernelGGL(k3_m2l_mfma, dim3(nb), dim3(256), 0, st, kd, td, fl,
                                   R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local,
                                   rs.flag);
            }
        }
        // real strengths: the conjugate-symmetric kernel, level by level (its
        // tile counts are fixed by the level's order)
        if (rs.real) {
            for (int l = 0; l < kd.nlevels; ++l) {
                const int i0 = R->h_level_start_ttp_box_nrs[l];
                const int nbl = R->h_level_start_ttp_box_nrs[l + 1] - i0;
                if (nbl == 0) continue;
                QBX_TRY(launch_trans_real<0>(st, kd, td, l, kd.order[l], i0, nbl,
                                             R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p,
                                             mpole, local, rs.flag));
            }
        }
    } else if (R->nttp_boxes > 0 && R->n2 > 0) {
        const TD td = make_td(T);
        const int no = ((kd.pmax + 1) * (kd.pmax + 1) + M2L_THREADS - 1) / M2L_THREADS;
#define L3_M2L(N)                                                                          \
    DISPATCH3_NO_NR(no, N, k3_m2l, dim3(R->nttp_boxes), dim3(M2L_THREADS),                 \
                    sizeof(double) * 2 * (size_t)(kd.pmax + 1) * (kd.pmax + 1) * (N), st,  \
                    kd, td, R->ttp_boxes.p, R->l2_starts.p, R->l2_lists.p, mpole, local)
        FOR_RHS_CHUNKS3(kd, NRC_S, L3_M2L);
#undef L3_M2L
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_form_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                   const qbx_kernel_desc *d, const double *str, const double *nrm, double *local,
                   int accumulate)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    QBX_CHECK(!kd.dip || nrm, "double layer needs normals");
    hipStream_t st = S(stream);
    if (!accumulate) QBX_TRY(zero_fill(local, (exp_bytes(kd)) / 8, st));
    if (R->n4 > 0 && p2e3_all_levels(kd, T->nlevels, true)) {
        const TD td = make_td(T);
        for (int l = 0; l < T->nlevels; ++l) {
            const int i0 = R->h_level_start_ttp_box_nrs[l];
            const int i1 = R->h_level_start_ttp_box_nrs[l + 1];
            if (i1 == i0) continue;
            P2EShape3 sh;
            int th;
            size_t by;
            p2e3_launch(kd.order[l], true, &sh, &th, &by);
            if (kd.nr >= 4 && (kd.order[l] + 1) * (kd.order[l] + 1) <= 16 * 4 * P2E3M_RTW) {
                for (int r0 = 0; r0 < kd.nr; r0 += 16) {
                    kd.r0 = r0;
                    kd.rc = std::min(16, kd.nr - r0);
                    hipLaunchKernelGGL(k3_p2e_mfma<true>, dim3(i1 - i0), dim3(256),
                                       p2e3m_bytes(sh.R), st, kd, td, l, i0, sh.R,
                                       R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p,
                                       T->sources.p, T->nsources, str, nrm, local);
                }
                continue;
            }
#define L3_P2LH(N)                                                                         \
    hipLaunchKernelGGL((k3_p2e<N, true>), dim3(i1 - i0), dim3(th), by, st, kd, td, l, i0, sh, \
                       R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p, T->sources.p,         \
                       T->nsources, str, nrm, local)
            FOR_RHS_CHUNKS3(kd, NRC_S, L3_P2LH);
#undef L3_P2LH
        }
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    const Flat fl = make_flat(kd, R->h_level_start_ttp_box_nrs);
    const long nthr = fl.tstart[fl.nl];
    if (nthr > 0 && R->n4 > 0) {
        const TD td = make_td(T);
#define L3_P2L(N)                                                                          \
    hipLaunchKernelGGL(k3_p2l<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, fl,     \
                       R->ttp_boxes.p, R->l4_starts.p, R->l4_lists.p, T->sources.p,        \
                       T->nsources, str, nrm, local)
        FOR_RHS_CHUNKS3(kd, NRC_S, L3_P2L);
#undef L3_P2L
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_refine_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_trav *R,
                     const qbx_kernel_desc *d, double *local, int lmin, const int *allreal)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd, R->nsep));
    QBX_TRY(ensure_tables(ctx, S(stream), &kd));
    hipStream_t st = S(stream);
    const TD td = make_td(T);
    for (int l = std::max(lmin, 0) + 1; l < T->nlevels; ++l) {
        const int i0 = R->h_level_start_ttp_box_nrs[l], i1 = R->h_level_start_ttp_box_nrs[l + 1];
        const long nthr = (long)(i1 - i0) * kd.ncl[l];
        if (nthr > 0) {
            const int p = kd.order[l];
            const RealSel rs = real_sel(d, kd, allreal, p <= 15 && kd.order[l - 1] == p);
            if (rs.real)
                QBX_TRY(launch_trans_real<2>(st, kd, td, l, p, i0, i1 - i0, R->ttp_boxes.p,
                                             nullptr, nullptr, local, local, rs.flag));
            if (rs.complex_) {
#define L3_L2L(N)                                                                          \
    hipLaunchKernelGGL(k3_l2l<N>, dim3(cdiv(nthr, 128)), dim3(128), 0, st, kd, td, l, i0,  \
                       i1, R->ttp_boxes.p, local, rs.flag)
                FOR_RHS_CHUNKS3(kd, NRC_S, L3_L2L);
#undef L3_L2L
            }
        }
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_form_global_qbx_locals(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                              const qbx_trav *R, const qbx_geo *G, const qbx_kernel_desc *d,
                              const double *str, const double *nrm, double *qexp, int accumulate)
{
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
    if (kd.nr >= 8 && kd.pq <= 5 && G->gc_group_plan.nunits > 0) {
        const unit_plan &V = G->gc_group_plan;
        const dim3 grid(V.nunits, cdiv(kd.nr, MR_RB));
#define L3_MR(PQV)                                                                         \
    hipLaunchKernelGGL(k3_p2qbxl_mr<PQV>, grid, dim3(256), 0, st, kd, td, V.unit_box.p,    \
                       V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p, G->centers.p, \
                       G->ncenters, R->l1_starts.p, R->l1_lists.p, T->sources.p,            \
                       T->nsources, str, nrm, qexp, accumulate)
        switch (kd.pq) {
        case 0: L3_MR(0); break;
        case 1: L3_MR(1); break;
        case 2: L3_MR(2); break;
        case 3: L3_MR(3); break;
        case 4: L3_MR(4); break;
        default: L3_MR(5); break;
        }
#undef L3_MR
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define L3_P2Q(N)                                                                          \
    DISPATCH3_PQ_NR(kd.pq, N, k3_p2qbxl, dim3(U.nunits), dim3(64), 0, st, kd, td,          \
                    U.unit_box.p, U.unit_t0.p, U.box_g.p, G->gc_by_tbox_starts.p,           \
                    G->gc_by_tbox.p, G->centers.p, G->ncenters, R->l1_starts.p,             \
                    R->l1_lists.p, T->sources.p, T->nsources, str, nrm, qexp, accumulate)
    FOR_RHS_CHUNKS3(kd, NRC_Q, L3_P2Q);
#undef L3_P2Q
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_translate_box_multipoles_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                             const qbx_trav *R, const qbx_geo *G,
                                             const qbx_kernel_desc *d, const double *mpole,
                                             double *qexp, int accumulate)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, (sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq * kd.nr) / 8, st));
    const int ng = G->nglobal;
    if (ng == 0 || R->n3 == 0) return QBX_OK;
    const int pn = kd.pmax + kd.pq;
    if (kd.nr >= 4 && kd.pq <= 5 && G->gc_group_plan.nunits > 0) {
        // many right-hand sides: the matrix-core kernel, one harmonic table
        // per wave (center)
        QBX_CHECK(G->trav == R, "M2QBXL: geometry was built for another traversal");
        const unit_plan &V = G->gc_group_plan;
        const size_t lds4 = sizeof(double) * 2 * (size_t)GC_GROUP * (pn + 1) * (pn + 1);
        QBX_CHECK(lds4 <= 64 * 1024, "M2QBXL: orders too high for LDS staging (%d + %d)",
                  kd.pmax, kd.pq);
        const TD td4 = make_td(T);
        const dim3 grid(V.nunits, cdiv(kd.nr, M2QM_RB));
#define L3_M2QM(PQV)                                                                       \
    hipLaunchKernelGGL(k3_m2qbxl_mfma<PQV>, grid, dim3(256), lds4, st, kd, td4, V.unit_box.p, \
                       V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p, G->centers.p, \
                       G->ncenters, R->l3_starts.p, R->l3_lists.p, mpole, qexp)
        switch (kd.pq) {
        case 0: L3_M2QM(0); break;
        case 1: L3_M2QM(1); break;
        case 2: L3_M2QM(2); break;
        case 3: L3_M2QM(3); break;
        case 4: L3_M2QM(4); break;
        default: L3_M2QM(5); break;
        }
#undef L3_M2QM
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
    // one harmonic table per 16-lane center group: 16 groups per block, or
    // 8 / 4 for the orders whose tables would not fit 64 KiB of LDS
    int block = 256;
    size_t lds = sizeof(double) * 2 * (size_t)(block / M2Q_LANES) * (pn + 1) * (pn + 1);
    while (lds > 64 * 1024 && block > 64) {
        block /= 2;
        lds /= 2;
    }
    QBX_CHECK(lds <= 64 * 1024, "M2QBXL: orders too high for LDS staging (%d + %d)", kd.pmax,
              kd.pq);
    const TD td = make_td(T);
#define L3_M2Q(N)                                                                          \
    DISPATCH3_PQ_NR(kd.pq, N, k3_m2qbxl, dim3(cdiv((long)ng * M2Q_LANES, block)), dim3(block), \
                    lds, st, kd, td, ng, G->global_qbx_centers.p, G->center_to_target_box.p, \
                    G->centers.p, G->ncenters, R->l3_starts.p, R->l3_lists.p, mpole, qexp)
    FOR_RHS_CHUNKS3(kd, NRC_Q, L3_M2Q);
#undef L3_M2Q
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_translate_box_local_to_qbx_local(qbx_ctx *ctx, void *stream, const qbx_tree *T,
                                        const qbx_trav *R, const qbx_geo *G,
                                        const qbx_kernel_desc *d, const double *local,
                                        double *qexp, int accumulate)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (!accumulate)
        QBX_TRY(zero_fill(qexp, (sizeof(double) * 2 * (size_t)G->ncenters * kd.ncq * kd.nr) / 8, st));
    const int ng = G->nglobal;
    if (ng == 0) return QBX_OK;
    const TD td = make_td(T);
    if (kd.nr >= 4 && kd.pq <= 5 && G->gc_group_plan.nunits > 0) {
        // many right-hand sides: the matrix-core kernel (a Y table per wave)
        QBX_CHECK(G->trav == R, "L2QBXL: geometry was built for another traversal");
        const unit_plan &V = G->gc_group_plan;
        const size_t lds4 = sizeof(double) * 2 * (size_t)GC_GROUP * (kd.pmax + 1) * (kd.pmax + 1);
        QBX_CHECK(lds4 <= 64 * 1024, "L2QBXL: order too high for LDS staging (%d)", kd.pmax);
        const dim3 grid(V.nunits, cdiv(kd.nr, M2QM_RB));
#define L3_L2QM(PQV)                                                                       \
    hipLaunchKernelGGL(k3_l2qbxl_mfma<PQV>, grid, dim3(256), lds4, st, kd, td, V.unit_box.p, \
                       V.unit_t0.p, G->gc_by_tbox_starts.p, G->gc_by_tbox.p,               \
                       R->target_boxes.p, G->centers.p, G->ncenters, local, qexp)
        switch (kd.pq) {
        case 0: L3_L2QM(0); break;
        case 1: L3_L2QM(1); break;
        case 2: L3_L2QM(2); break;
        case 3: L3_L2QM(3); break;
        case 4: L3_L2QM(4); break;
        default: L3_L2QM(5); break;
        }
#undef L3_L2QM
        QBX_HIP(hipGetLastError());
        return QBX_OK;
    }
#define L3_L2Q(N)                                                                          \
    DISPATCH3_PQ_NR(kd.pq, N, k3_l2qbxl, dim3(cdiv(ng, 64)), dim3(64), 0, st, kd, td, ng,  \
                    G->global_qbx_centers.p, G->center_to_target_box.p, R->target_boxes.p, \
                    G->centers.p, G->ncenters, local, qexp)
    FOR_RHS_CHUNKS3(kd, NRC_Q, L3_L2Q);
#undef L3_L2Q
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

int f3_eval_qbx_expansions(qbx_ctx *ctx, void *stream, const qbx_tree *T, const qbx_geo *G,
                           const qbx_kernel_desc *d, const double *qexp, double *pot)
{
    KD kd;
    QBX_TRY(make_kd(d, T, &kd));
    hipStream_t st = S(stream);
    if (G->nctt > 0) {
#define L3_Q2P(N)                                                                          \
    hipLaunchKernelGGL(k3_qbxl2p<N>, dim3(cdiv(G->nctt, 64)), dim3(64), 0, st, kd,         \
                       G->nctt, G->ctt_lists.p, G->tree_ttc.p, T->targets.p, T->ntargets,  \
                       G->centers.p, G->ncenters, qexp, pot)
        FOR_RHS_CHUNKS3(kd, NRC_S, L3_Q2P);
#undef L3_Q2P
    }
    QBX_HIP(hipGetLastError());
    return QBX_OK;
}

}  // namespace qbx

// Device math for the 2D QBX-FMM: complex helpers and Bessel recurrences.
//
// Bessel functions of real argument x > 0:
//  * J_n: Miller's backward recurrence J_{n-1} = (2n/x) J_n - J_{n+1}, started
//    well above max(n, x), normalised with J_0 + 2 sum_k J_{2k} = 1, with
//    rescaling before overflow.  Consumers receive unnormalised values and the
//    final normalisation factor (they accumulate linearly).
//  * H_n = J_n + i Y_n (first kind Hankel): forward recurrence from
//    H_0, H_1 (j0/j1/y0/y1 of the device math library); accurate relative to
//    |H_n| because Y dominates where the J part loses accuracy.
#pragma once

#include <hip/hip_runtime.h>

namespace qbx {

// ---- cheap elementary functions for the pair loops -------------------------
// The device library's log (double-double internally), sqrt and IEEE
// division are correctly rounded or nearly so and cost 30..90 instructions
// each; the pair loops only need a few ulp.

// 1 / sqrt(x) for positive normal x: the hardware estimate refined by two
// Newton steps (each squares the relative error)
__device__ __forceinline__ double qbx_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-h * y, y, 0.5);
    return __builtin_fma(y, e, y);
}

// 1 / x for normal x: the hardware estimate refined by two Newton steps
__device__ __forceinline__ double qbx_rcp(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// sin and cos of one argument, within 2 ulp of 1 for |x| <= 1.6e6 (checked
// against the host library over [0, 1.6e6]): Cody and Waite's reduction
// (W. J. Cody and W. Waite, Software Manual for the Elementary Functions,
// 1980) by q = rint(x 2/pi) with pi/2 split into three doubles, the first
// two of 33 significant bits so that q P1 and q P2 are exact for |q| < 2^20,
// then the Taylor series of sin and cos on |r| <= pi/4 (through r^17 and
// r^16: truncation below 1e-17); the quadrant from q mod 4.  Larger
// arguments take the device library.
__device__ __forceinline__ void qbx_sincos(double x, double *sn, double *cs)
{
    if (!(__builtin_fabs(x) <= 1.6e6)) {
        sincos(x, sn, cs);
        return;
    }
    constexpr double TWO_OVER_PI = 0.6366197723675814;
    constexpr double P1 = 1.5707963267341256, P2 = 6.077100506303966e-11,
                     P3 = 2.0222662487959506e-21;
    const double q = __builtin_rint(x * TWO_OVER_PI);
    const double r = __builtin_fma(-q, P3, __builtin_fma(-q, P2, __builtin_fma(-q, P1, x)));
    const double z = r * r;
    // 1/n! with alternating signs
    double ps = -8.22063524662433e-18;
    ps = __builtin_fma(ps, z, 2.8114572543455206e-15);
    ps = __builtin_fma(ps, z, -7.647163731819816e-13);
    ps = __builtin_fma(ps, z, 1.6059043836821613e-10);
    ps = __builtin_fma(ps, z, -2.505210838544172e-08);
    ps = __builtin_fma(ps, z, 2.7557319223985893e-06);
    ps = __builtin_fma(ps, z, -0.0001984126984126984);
    ps = __builtin_fma(ps, z, 0.008333333333333333);
    ps = __builtin_fma(ps, z, -0.16666666666666666);
    double pc = -1.5619206968586225e-16;
    pc = __builtin_fma(pc, z, 4.779477332387385e-14);
    pc = __builtin_fma(pc, z, -1.1470745597729725e-11);
    pc = __builtin_fma(pc, z, 2.08767569878681e-09);
    pc = __builtin_fma(pc, z, -2.755731922398589e-07);
    pc = __builtin_fma(pc, z, 2.48015873015873e-05);
    pc = __builtin_fma(pc, z, -0.001388888888888889);
    pc = __builtin_fma(pc, z, 0.041666666666666664);
    pc = __builtin_fma(pc, z, -0.5);
    const double sr = __builtin_fma(r * z, ps, r);
    const double cr = __builtin_fma(z, pc, 1.0);
    const int quad = (int)(long long)q & 3;
    const double a = (quad & 1) ? cr : sr, b = (quad & 1) ? sr : cr;
    *sn = (quad & 2) ? -a : a;
    *cs = ((quad + 1) & 2) ? -b : b;
}

// natural logarithm of a positive normal x, within a few ulp: x = m 2^e with
// m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with s = (m - 1)/(m + 1),
// |s| <= 0.1716, summed as 2 s sum_{j<=11} s^(2j) / (2j + 1) (the first
// dropped term is below 1e-18 relative), e ln 2 in two parts (Cody and
// Waite's split of ln 2: the high part has trailing zeros, so e times it is
// exact)
__device__ __forceinline__ double qbx_log(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);   // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double num = m - 1.0, den = m + 1.0;
    const double r = qbx_rcp(den);
    double sv = num * r;
    sv = __builtin_fma(r, __builtin_fma(-den, sv, num), sv);
    const double s2 = sv * sv;
    double p = 1.0 / 23.0;
    p = __builtin_fma(p, s2, 1.0 / 21.0);
    p = __builtin_fma(p, s2, 1.0 / 19.0);
    p = __builtin_fma(p, s2, 1.0 / 17.0);
    p = __builtin_fma(p, s2, 1.0 / 15.0);
    p = __builtin_fma(p, s2, 1.0 / 13.0);
    p = __builtin_fma(p, s2, 1.0 / 11.0);
    p = __builtin_fma(p, s2, 1.0 / 9.0);
    p = __builtin_fma(p, s2, 1.0 / 7.0);
    p = __builtin_fma(p, s2, 1.0 / 5.0);
    p = __builtin_fma(p, s2, 1.0 / 3.0);
    const double t = (2.0 * sv) * s2 * p;      // 2 s (s^2/3 + s^4/5 + ...)
    constexpr double LN2_HI = 6.93147180369123816490e-01;
    constexpr double LN2_LO = 1.90821492927058770002e-10;
    const double de = (double)e;
    return __builtin_fma(de, LN2_HI, __builtin_fma(de, LN2_LO, t) + 2.0 * sv);
}


struct cplx {
    double x, y;
};

__device__ __forceinline__ cplx mk(double a, double b) { return cplx{a, b}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cplx operator-(cplx a) { return {-a.x, -a.y}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cplx operator*(double s, cplx a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ cplx operator*(cplx a, double s) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ cplx &operator+=(cplx &a, cplx b)
{
    a.x += b.x;
    a.y += b.y;
    return a;
}
__device__ __forceinline__ cplx conj(cplx a) { return {a.x, -a.y}; }
__device__ __forceinline__ double abs2(cplx a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ cplx cinv(cplx a)
{
    double d = qbx_rcp(abs2(a));
    return {a.x * d, -a.y * d};
}
__device__ __forceinline__ cplx cdivz(cplx a, cplx b) { return a * cinv(b); }
__device__ __forceinline__ cplx clogz(cplx a) { return {0.5 * log(abs2(a)), atan2(a.y, a.x)}; }
__device__ __forceinline__ cplx ldc(const double *p) { return {p[0], p[1]}; }
__device__ __forceinline__ void stc(double *p, cplx v)
{
    p[0] = v.x;
    p[1] = v.y;
}

// H_0(x), H_1(x) for x > 0
template <bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_fast(double x, cplx &h0, cplx &h1);

__device__ __forceinline__ void hankel01(double x, cplx &h0, cplx &h1)
{
    hankel01_fast<true, true>(x, h0, h1);
}

__device__ __forceinline__ int miller_start(int nmax, double x)
{
    double m = fmax((double)nmax, x);
    int N = (int)m + 16 + (int)sqrt(40.0 * m + 1.0);
    return N + (N & 1);
}

// Backward recurrence for J_0..J_nmax at x > 0.  consume(n, f) sees
// unnormalised f_n for n = nmax..0 (descending); rescale(s) is called when the
// running values are multiplied by s.  Returns the normalisation factor.
template <class Consume, class Rescale>
__device__ __forceinline__ double miller_j(double x, int nmax, Consume consume, Rescale rescale)
{
    int N = miller_start(nmax, x);
    double fp1 = 0.0, f = 1e-30, sum = 0.0;
    double tox = 2.0 / x;
    for (int n = N; n >= 1; --n) {
        if (n <= nmax) consume(n, f);
        if ((n & 1) == 0) sum += f;
        double fm1 = (double)n * tox * f - fp1;
        fp1 = f;
        f = fm1;
        if (fabs(f) > 1e250) {
            const double s = 1e-250;
            f *= s;
            fp1 *= s;
            sum *= s;
            rescale(s);
        }
    }
    consume(0, f);
    return 1.0 / (f + 2.0 * sum);
}

// ---------------------------------------------------------------------------
// H_0 and H_1 of real argument, evaluated together.  For 0 < x <= 4 the power
// series of Abramowitz & Stegun 9.1.10 / 9.1.11 in t = x^2/4:
//   J0 = sum A_k t^k,  J1 = (x/2) sum B_k t^k,
//   Y0 = (2/pi) ln(x/2) J0 - (1/pi) sum C_k t^k,
//   Y1 = -2/(pi x) + (2/pi) ln(x/2) J1 - (x/(2 pi)) sum D_k t^k,
//   A_k = (-1)^k/(k!)^2,  B_k = (-1)^k/(k!(k+1)!),
//   C_k = (-1)^k 2 psi(k+1)/(k!)^2,  D_k = (-1)^k (psi(k+1)+psi(k+2))/(k!(k+1)!)
// (coefficients generated at 50 digits, psi(k+1) = H_k - gamma).  Relative
// error of H0 and H1 against scipy: 10 terms < 1.1e-15 for x <= 1.5, 14 terms
// < 9e-16 for x <= 3, 18 terms < 2.7e-15 for x <= 4, 22 terms < 6.2e-15 for
// x <= 5, 26 terms < 1.4e-14 for x <= 6 (cancellation grows with x);
// hankel01_fast switches to the device library above 4, hankel01_wave above 6.
// ---------------------------------------------------------------------------
static __constant__ double c_ser_A[26] = {
    1.0, -1.0, 0.25,
    -0.027777777777777776, 0.001736111111111111, -6.944444444444444e-05,
    1.9290123456790124e-06, -3.936759889140842e-08, 6.151187326782565e-10,
    -7.594058428126624e-12, 7.594058428126623e-14, -6.276081345559193e-16,
    4.358389823304995e-18, -2.5789288895295828e-20, 1.3157800456783586e-22,
    -5.8479113141260385e-25, 2.2843403570804838e-27, -7.904291893012054e-30,
    2.4395962632753253e-32, -6.757884385804225e-35, 1.6894710964510564e-37,
    -3.8310002187098785e-40, 7.915289708078262e-43, -1.4962740468957016e-45,
    2.5976979980828152e-48, -4.156316796932504e-51,
};
static __constant__ double c_ser_B[26] = {
    1.0, -0.5, 0.08333333333333333,
    -0.006944444444444444, 0.00034722222222222224, -1.1574074074074073e-05,
    2.755731922398589e-07, -4.920949861426052e-09, 6.834652585313961e-11,
    -7.594058428126623e-13, 6.903689480115112e-15, -5.230067787965994e-17,
    3.352607556388458e-19, -1.842092063949702e-21, 8.771866971189057e-24,
    -3.654944571328774e-26, 1.3437296218120491e-28, -4.391273273895586e-31,
    1.2839980333028028e-33, -3.3789421929021126e-36, 8.045100459290744e-39,
    -1.7413637357772174e-41, 3.4414303078601135e-44, -6.234475195398757e-47,
    1.0390791992331261e-49, -1.5985833834355786e-52,
};
static __constant__ double c_ser_C[26] = {
    -1.1544313298030657, -0.8455686701969343, 0.46139216754923357,
    -0.0697843149128778, 0.0052295752376104185, -0.00023696078728219452,
    7.2252482063984075e-06, -1.5870193042608417e-07, 2.6334973460771295e-09,
    -3.419988145411616e-11, 3.571869313974148e-13, -3.0660686806416945e-15,
    2.2018541919451488e-17, -1.3425481319958094e-19, 7.03770394527961e-22,
    -3.205840570979285e-24, 1.2808357275022893e-26, -4.5249492039395507e-29,
    1.4236958856226736e-31, -4.0148915409660105e-34, 1.0206175962060132e-36,
    -2.350811331348287e-39, 4.929005384429404e-42, -9.447701305456722e-45,
    1.6618734044035933e-47, -2.692247981421209e-50,
};
static __constant__ double c_ser_D[26] = {
    -0.15443132980306573, -0.6727843350984671, 0.18157516696085563,
    -0.019182189839330562, 0.001115359491966528, -4.142247689271143e-05,
    1.071545914091181e-06, -2.045286003593878e-08, 3.002048746589188e-10,
    -3.495928729692882e-12, 3.309914735250272e-14, -2.5986411321011286e-16,
    1.7195232826992565e-18, -9.721207518823618e-21, 4.750281743327667e-23,
    -2.0264937604328578e-25, 7.613370727767116e-28, -2.5382566314880593e-30,
    7.560715083977377e-33, -2.0243404814475158e-35, 4.898393793644304e-38,
    -1.0764658948663905e-40, 2.1580085597860893e-43, -3.962519190587796e-46,
    6.689056785583698e-49, -1.0416283904829096e-51,
};

template <int K, bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_series(double x, cplx &h0, cplx &h1)
{
    const double t = 0.25 * x * x;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        if (WANT0) {
            a = a * t + c_ser_A[k];
            c = c * t + c_ser_C[k];
        }
        if (WANT1) {
            b = b * t + c_ser_B[k];
            d = d * t + c_ser_D[k];
        }
    }
    const double lg = qbx_log(0.5 * x);
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    constexpr double ONE_OVER_PI = 0.31830988618379067154;
    if (WANT0) h0 = {a, TWO_OVER_PI * lg * a - ONE_OVER_PI * c};
    if (WANT1) {
        double j1v = 0.5 * x * b;
        h1 = {j1v, -TWO_OVER_PI / x + TWO_OVER_PI * lg * j1v - 0.5 * x * ONE_OVER_PI * d};
    }
}

template <bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_fast(double x, cplx &h0, cplx &h1)
{
    if (x <= 1.5) {
        hankel01_series<10, WANT0, WANT1>(x, h0, h1);
    } else if (x <= 3.0) {
        hankel01_series<14, WANT0, WANT1>(x, h0, h1);
    } else if (x <= 4.0) {
        hankel01_series<18, WANT0, WANT1>(x, h0, h1);
    } else {
        if (WANT0) h0 = {j0(x), y0(x)};
        if (WANT1) h1 = {j1(x), y1(x)};
    }
}

// H_0, H_1 from the device library, kept out of line so that its
// coefficients are materialised inside the call rather than held in
// registers across a caller's loop; returned by value (in registers: a
// by-reference result would go through scratch memory)
struct H01 {
    cplx h0, h1;
};

template <bool WANT0, bool WANT1>
__device__ __noinline__ H01 hankel01_libm(double x)
{
    H01 r = {{0, 0}, {0, 0}};
    if (WANT0) r.h0 = {j0(x), y0(x)};
    if (WANT1) r.h1 = {j1(x), y1(x)};
    return r;
}

// hankel01_fast out of line: for callers that evaluate it once per outer
// iteration, so that neither the series coefficients nor the device library's
// are held in registers across the caller's loops
__device__ __noinline__ H01 hankel01_call_v(double x)
{
    H01 r;
    hankel01_fast<true, true>(x, r.h0, r.h1);
    return r;
}

__device__ __forceinline__ void hankel01_call(double x, cplx &h0, cplx &h1)
{
    const H01 r = hankel01_call_v(x);
    h0 = r.h0;
    h1 = r.h1;
}

// The series of hankel01_series with its coefficients read from tab, laid
// out [k][A_k, C_k, B_k, D_k] so that each step reads its pair with one
// 16-byte load (typically a copy in LDS, so that a loop around it does not
// pin the coefficients in registers)
constexpr int HANKEL_NSER = 26;

template <int K, bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_series_tab(double x, double ix, const double *tab,
                                                    cplx &h0, cplx &h1)
{
    const double t = 0.25 * x * x;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        if (WANT0) {
            const double2 ac = *reinterpret_cast<const double2 *>(tab + 4 * k);
            a = a * t + ac.x;
            c = c * t + ac.y;
        }
        if (WANT1) {
            const double2 bd = *reinterpret_cast<const double2 *>(tab + 4 * k + 2);
            b = b * t + bd.x;
            d = d * t + bd.y;
        }
    }
    const double lg = qbx_log(0.5 * x);
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    constexpr double ONE_OVER_PI = 0.31830988618379067154;
    if (WANT0) h0 = {a, TWO_OVER_PI * lg * a - ONE_OVER_PI * c};
    if (WANT1) {
        double j1v = 0.5 * x * b;
        h1 = {j1v, -TWO_OVER_PI * ix + TWO_OVER_PI * lg * j1v - 0.5 * x * ONE_OVER_PI * d};
    }
}

// fill tab (4 * HANKEL_NSER doubles, 16-byte aligned) cooperatively by the
// lanes of a block
__device__ __forceinline__ void hankel_table_fill(double *tab, int lane, int nlanes)
{
    for (int i = lane; i < 4 * HANKEL_NSER; i += nlanes) {
        const int k = i / 4, w = i % 4;
        tab[i] = w == 0 ? c_ser_A[k] : w == 1 ? c_ser_C[k] : w == 2 ? c_ser_B[k] : c_ser_D[k];
    }
}

// H_0, H_1 with the series length chosen once per wavefront: the lanes with
// x <= 6 share one series of the length the largest of them needs (one pass
// instead of several divergent ones); lanes beyond 6 take the device library
// in their own pass.  ix = 1 / x (callers that have 1/r at hand save a
// division).
template <bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_wave(double x, double ix, const double *tab, cplx &h0,
                                              cplx &h1)
{
    const bool big = x > 6.0;
    if (__any(big)) {
        if (big) {
            const H01 r = hankel01_libm<WANT0, WANT1>(x);
            h0 = r.h0;
            h1 = r.h1;
        }
    }
    if (__any(!big)) {
        cplx a0 = {0, 0}, a1 = {0, 0};
        if (__any(!big && x > 5.0))
            hankel01_series_tab<26, WANT0, WANT1>(x, ix, tab, a0, a1);
        else if (__any(!big && x > 4.0))
            hankel01_series_tab<22, WANT0, WANT1>(x, ix, tab, a0, a1);
        else if (__any(!big && x > 3.0))
            hankel01_series_tab<18, WANT0, WANT1>(x, ix, tab, a0, a1);
        else if (__any(!big && x > 1.5))
            hankel01_series_tab<14, WANT0, WANT1>(x, ix, tab, a0, a1);
        else
            hankel01_series_tab<10, WANT0, WANT1>(x, ix, tab, a0, a1);
        if (!big) {
            h0 = a0;
            h1 = a1;
        }
    }
}

// The series of hankel01_series with its coefficients read from the
// constant tables through an offset the compiler cannot see through (a
// wave-uniform zero kept in a scalar register): the reads stay scalar loads
// inside the caller's loop, served by the scalar cache, and feed the FMAs as
// scalar operands -- no vector registers held across the loop and no LDS
// traffic.
template <int K, bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_series_sc(double x, double ix, int off, cplx &h0,
                                                   cplx &h1)
{
    const double t = 0.25 * x * x;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        if (WANT0) {
            a = a * t + c_ser_A[k + off];
            c = c * t + c_ser_C[k + off];
        }
        if (WANT1) {
            b = b * t + c_ser_B[k + off];
            d = d * t + c_ser_D[k + off];
        }
    }
    const double lg = qbx_log(0.5 * x);
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    constexpr double ONE_OVER_PI = 0.31830988618379067154;
    if (WANT0) h0 = {a, TWO_OVER_PI * lg * a - ONE_OVER_PI * c};
    if (WANT1) {
        double j1v = 0.5 * x * b;
        h1 = {j1v, -TWO_OVER_PI * ix + TWO_OVER_PI * lg * j1v - 0.5 * x * ONE_OVER_PI * d};
    }
}

// hankel01_wave with the scalar-operand series (off: a wave-uniform zero the
// caller hides from the optimiser once per iteration)
template <bool WANT0, bool WANT1>
__device__ __forceinline__ void hankel01_wave_sc(double x, double ix, int off, cplx &h0,
                                                 cplx &h1)
{
    const bool big = x > 6.0;
    if (__any(big)) {
        if (big) {
            const H01 r = hankel01_libm<WANT0, WANT1>(x);
            h0 = r.h0;
            h1 = r.h1;
        }
    }
    if (__any(!big)) {
        cplx a0 = {0, 0}, a1 = {0, 0};
        if (__any(!big && x > 5.0))
            hankel01_series_sc<26, WANT0, WANT1>(x, ix, off, a0, a1);
        else if (__any(!big && x > 4.0))
            hankel01_series_sc<22, WANT0, WANT1>(x, ix, off, a0, a1);
        else if (__any(!big && x > 3.0))
            hankel01_series_sc<18, WANT0, WANT1>(x, ix, off, a0, a1);
        else if (__any(!big && x > 1.5))
            hankel01_series_sc<14, WANT0, WANT1>(x, ix, off, a0, a1);
        else
            hankel01_series_sc<10, WANT0, WANT1>(x, ix, off, a0, a1);
        if (!big) {
            h0 = a0;
            h1 = a1;
        }
    }
}

}  // namespace qbx

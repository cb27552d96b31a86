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
    double d = 1.0 / abs2(a);
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
__device__ __forceinline__ void hankel01(double x, cplx &h0, cplx &h1)
{
    h0 = {j0(x), y0(x)};
    h1 = {j1(x), y1(x)};
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

}  // namespace qbx

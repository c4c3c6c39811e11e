// Elementary functions in double precision, for float arguments, on the device
// and the host.
//
// The reference's Sin, Cos, Atan, Power, Exp, ArcSine and ArcCosine
// (math.cpp:136-260) wrap the platform's float functions, whose last bit
// differs between C libraries and the GPU.  The renderer and its CPU
// restatement both use the value of the function at the float argument,
// rounded to float: the restatement through the C library's double functions,
// the device through the functions below, which stay within a few units of the
// 53-bit last place.  Rounded to float, the two agree in all but about one case
// in 2^26, so both sides draw the same directions and weights.
//
// Methods: Cody and Waite's additive argument reduction (pi/2 and ln 2 split
// into parts whose products with the quotient are exact); Taylor polynomials on
// the reduced interval, carried past the 53-bit last place; atan by a 17-entry
// table of atan(k/16) and the addition formula; log by the atanh series of
// (m - 1) / (m + 1).  No library routine is called besides sqrt, fma, rint,
// frexp and ldexp, which are exact or correctly rounded.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace az {
namespace dm {

constexpr double kPio2_1 = 1.5707963267948966;        // pi/2 = kPio2_1 + kPio2_2 + kPio2_3 + O(1e-49)
constexpr double kPio2_2 = 6.123233995736766e-17;
constexpr double kPio2_3 = -1.4973849048591698e-33;
constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPiHi = 3.141592653589793, kPiLo = 1.2246467991473532e-16;
constexpr double kLn2Hi = 0.6931467056274414;          // 21 significant bits: k * kLn2Hi is exact
constexpr double kLn2Lo = 4.7493250390316726e-07;      // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.4426950408889634;

// atan(k / 16), k = 0..16, correctly rounded.
constexpr double kAtan16[17] = {
    0.0,                 0.06241880999595735, 0.12435499454676144, 0.18534794999569476, 0.24497866312686414,
    0.3028848683749714,  0.35877067027057225, 0.4124104415973873,  0.4636476090008061,  0.5123894603107377,
    0.5585993153435624,  0.6022873461349642,  0.6435011087932844,  0.6823165548747481,  0.7188299996216245,
    0.7531512809621944,  0.7853981633974483};

// sin(x) and cos(x) for |x| < 2^30.  x - k pi/2 is exact for the first part
// (x and k * kPio2_1 are multiples of 2^-52 and the difference is below 1 in
// magnitude), so the reduced argument carries only the rounding of the last
// two parts.  Taylor terms to r^17 and r^18 on |r| <= pi/4 leave a truncation
// error below 1e-19.
__host__ __device__ inline void sincos_d(double x, double &s, double &c) {
    const double k = rint(x * kTwoOverPi);
    double r = fma(-k, kPio2_1, x);
    r = fma(-k, kPio2_2, r);
    r = fma(-k, kPio2_3, r);
    const double z = r * r;
    double ps = 2.8114572543455206e-15;                // 1/17!
    ps = fma(ps, z, -7.647163731819816e-13);           // -1/15!
    ps = fma(ps, z, 1.6059043836821613e-10);           // 1/13!
    ps = fma(ps, z, -2.505210838544172e-08);           // -1/11!
    ps = fma(ps, z, 2.755731922398589e-06);            // 1/9!
    ps = fma(ps, z, -1.984126984126984e-04);           // -1/7!
    ps = fma(ps, z, 8.333333333333333e-03);            // 1/5!
    ps = fma(ps, z, -1.6666666666666666e-01);          // -1/3!
    const double sr = fma(r * z, ps, r);
    // cos r = 1 - z/2 + z^2 q(z), q(z) = 1/4! - z/6! + ... - z^7/18!
    double pc = -1.5619206968586225e-16;               // -1/18!
    pc = fma(pc, z, 4.779477332387385e-14);            // 1/16!
    pc = fma(pc, z, -1.1470745597729725e-11);          // -1/14!
    pc = fma(pc, z, 2.08767569878681e-09);             // 1/12!
    pc = fma(pc, z, -2.755731922398589e-07);           // -1/10!
    pc = fma(pc, z, 2.48015873015873e-05);             // 1/8!
    pc = fma(pc, z, -1.388888888888889e-03);           // -1/6!
    pc = fma(pc, z, 4.1666666666666664e-02);           // 1/4!
    const double cr = fma(z * z, pc, fma(z, -0.5, 1.0));
    // Quadrant k mod 4, computed in double so a NaN argument stays NaN.
    const double q = k - 4.0 * floor(k * 0.25);
    const bool swap = q == 1.0 || q == 3.0;
    double sv = swap ? cr : sr, cv = swap ? sr : cr;
    if (q >= 2.0) sv = -sv;
    if (q == 1.0 || q == 2.0) cv = -cv;
    s = x == 0.0 ? x : sv;                             // sin(-0) = -0
    c = cv;
}

// exp(xh + xl) for |xl| <= 1e-15 |xh|; 0 below -746 and inf above 710.
// xh - k kLn2Hi is exact (k * kLn2Hi is a multiple of 2^-21 and the difference
// lies within the bits of xh).  Taylor terms to r^13 on |r| <= 0.35 leave a
// truncation error below 5e-18.
__host__ __device__ inline double exp_dd(double xh, double xl) {
    if (!(xh < 710.0)) return xh * 1e300;              // inf, or NaN
    if (xh < -746.0) return 0.0;
    const double k = rint(xh * kInvLn2);
    double r = fma(-k, kLn2Hi, xh);
    r = fma(-k, kLn2Lo, r) + xl;
    double p = 1.6059043836821613e-10;                 // 1/13!
    p = fma(p, r, 2.08767569878681e-09);               // 1/12!
    p = fma(p, r, 2.505210838544172e-08);              // 1/11!
    p = fma(p, r, 2.755731922398589e-07);              // 1/10!
    p = fma(p, r, 2.755731922398589e-06);              // 1/9!
    p = fma(p, r, 2.48015873015873e-05);               // 1/8!
    p = fma(p, r, 1.984126984126984e-04);              // 1/7!
    p = fma(p, r, 1.388888888888889e-03);              // 1/6!
    p = fma(p, r, 8.333333333333333e-03);              // 1/5!
    p = fma(p, r, 4.1666666666666664e-02);             // 1/4!
    p = fma(p, r, 1.6666666666666666e-01);             // 1/3!
    p = fma(p, r, 0.5);
    p = fma(p, r * r, r);                              // r + r^2/2 + ... ; added to 1 last
    return ldexp(1.0 + p, (int)k);
}

// log(x) = hi + lo for finite x > 0 given as a float: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.172.
// m - 1 and m + 1 are exact for a float m, the quotient is carried as s + sl,
// and the series runs to s^23 (truncation below 1e-19 relative).
__host__ __device__ inline void log_dd(double x, double &hi, double &lo) {
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0, den = m + 1.0;
    const double s = f / den;
    const double sl = fma(-s, den, f) / den;
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    // log x = e ln2 + 2s + 2s z p(z) + 2 sl, summed as hi + lo with exact
    // two-term sums (|e ln2| >= |2s| whenever e != 0, |2s| > |2s z p|).
    const double b = 2.0 * s, t = (2.0 * s * z) * p;
    const double h1 = b + t, l1 = (b - h1) + t;
    const double a = (double)e * kLn2Hi;
    hi = a + h1;
    lo = ((a - hi) + h1) + (l1 + 2.0 * sl + (double)e * kLn2Lo);
}

// x^y for finite x > 0 and finite y given as floats.  Out of line: inlined
// into each BRDF evaluation it pushes the shading kernels past their register
// budget (128 VGPRs at 4 waves per SIMD) into scratch spills.
__host__ __device__ inline __attribute__((noinline)) double pow_pos(double x, double y) {
    double lh, ll;
    log_dd(x, lh, ll);
    const double th = y * lh;
    const double tl = fma(y, lh, -th) + y * ll;
    return exp_dd(th, tl);
}

// atan2(y, x) with the C library's signs and zeros: reduce |y|/|x| or |x|/|y|
// to z in [0, 1], then atan z = atan(k/16) + atan(d), d = (z - k/16)/(1 + z k/16),
// |d| <= 1/32, by its series to d^13 (truncation below 2e-24).
__host__ __device__ inline double atan2_d(double y, double x) {
    if (x != x || y != y) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    double r = 0.0;
    if (den != 0.0) {
        const double z = num == den ? 1.0 : num / den;   // both infinite: pi/4
        const double kd = fmin(fmax(rint(z * 16.0), 0.0), 16.0);
        const double c = kd * 0.0625;
        const double d = (z - c) / fma(z, c, 1.0);
        const double w = d * d;
        double p = 1.0 / 13.0;
        p = fma(p, w, -1.0 / 11.0);
        p = fma(p, w, 1.0 / 9.0);
        p = fma(p, w, -1.0 / 7.0);
        p = fma(p, w, 1.0 / 5.0);
        p = fma(p, w, -1.0 / 3.0);
        r = kAtan16[(int)kd] + fma(d * w, p, d);
    }
    if (swap) r = (kPio2_1 - r) + kPio2_2;
    if (__builtin_signbit(x)) r = (kPiHi - r) + kPiLo;
    return copysign(r, y);
}

}  // namespace dm

// The float functions, as the renderer calls them.
__host__ __device__ inline void SinCosF(float v, float &sv, float &cv) {   // Sin and Cos of one angle
    double s, c;
    dm::sincos_d((double)v, s, c);
    sv = (float)s;
    cv = (float)c;
}
__host__ __device__ inline float SinF(float v) {
    float s, c;
    SinCosF(v, s, c);
    return s;
}
__host__ __device__ inline float CosF(float v) {
    float s, c;
    SinCosF(v, s, c);
    return c;
}
__host__ __device__ inline float Atan2F(float y, float x) { return (float)dm::atan2_d((double)y, (double)x); }
__host__ __device__ inline float Exp(float v) {                            // math.cpp:149-156
    const double d = v;
    return d != d ? v : (float)dm::exp_dd(d, 0.0);
}
__host__ __device__ inline float Asin(float v) {                           // math.cpp:221-228
    const double d = v;
    return (float)dm::atan2_d(d, sqrt(fma(-d, d, 1.0)));
}
__host__ __device__ inline float Acos(float v) {                           // math.cpp:239-246
    const double d = v;
    return (float)dm::atan2_d(sqrt(fma(-d, d, 1.0)), d);
}
// x^y with the C library's special cases (zeros, negative bases with integer
// exponents, infinities, NaN).
__host__ __device__ inline float PowF(float x, float y) {
    if (x > 0.0f && x < __builtin_huge_valf() && y - y == 0.0f) return (float)dm::pow_pos((double)x, (double)y);
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    const float ax = fabsf(x);
    const bool yInt = floorf(y) == y;
    const bool yOdd = yInt && fabsf(y) < 16777216.0f && floorf(0.5f * y) != 0.5f * y;
    float r;
    if (ax == 0.0f) r = y > 0.0f ? 0.0f : __builtin_huge_valf();
    else if (ax == __builtin_huge_valf()) r = y > 0.0f ? __builtin_huge_valf() : 0.0f;
    else if (y - y != 0.0f) r = ax == 1.0f ? 1.0f : (ax < 1.0f) == (y > 0.0f) ? 0.0f : __builtin_huge_valf();   // y = +-inf
    else r = (float)dm::pow_pos((double)ax, (double)y);
    if (x < 0.0f || (x == 0.0f && __builtin_signbit(x))) {
        if (!yInt && ax != 0.0f && ax != __builtin_huge_valf() && y - y == 0.0f) return __builtin_nanf("");
        if (yOdd) r = -r;
    }
    return r;
}
// x^N for a small integer N by products in double: at most two roundings
// after the exact square.
template <int N>
__host__ __device__ inline float PowIF(float x) {
    const double d = x, d2 = d * d;
    if constexpr (N == 4) return (float)(d2 * d2);
    else return (float)((d2 * d2) * d);
}

}  // namespace az

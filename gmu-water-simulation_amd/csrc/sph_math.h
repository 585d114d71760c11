// sph_math.h -- device arithmetic of the grid-SPH kernels whose results must be
// the IEEE / libm results the reference's CPU code gets (src/CCPUParticleSimulator.cpp),
// bit for bit. Shared by the kernels and by the exhaustive device checks of them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// t^3 correctly rounded (error-free products via FMA), the value libm's
// pow(t, 3.0) returns (src/CCPUParticleSimulator.cpp:14) for every t = h^2 - r^2
// of an accepted pair (checked on the host against glibc pow for every float r^2
// in [0, h^2], and the device results against the host's). Host and device.
__host__ __device__ __forceinline__ double cube_cr(double t)
{
    const double p = t * t;
    const double e = __builtin_fma(t, t, -p);   // t*t = p + e exactly
    const double q = p * t;
    const double f = __builtin_fma(p, t, -q);   // p*t = q + f exactly
    const double c = __builtin_fma(e, t, f);    // f + e*t
    return q + c;
}

// The C library's pow is not correctly rounded (an error bound of about 0.52
// ulp): where the exact power lies within a few thousandths of an ulp of a
// rounding midpoint it may return the other neighbour. The inputs where it does
// are enumerated at build time over the whole domain of accepted pairs (every
// float r^2 in [0, fl(h^2)], gen_pow_tables.hip) into sorted tables of the float
// bit pattern of r^2, bit 31 set when pow returns the next double up (else the
// next double down). pow_exception looks one input up: +1 / -1 / 0 (not listed).
__device__ __forceinline__ int pow_exception(const uint32_t *tab, uint32_t n, uint32_t r2_bits)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((tab[mid] & 0x7FFFFFFFu) < r2_bits)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < n && (tab[lo] & 0x7FFFFFFFu) == r2_bits) return (tab[lo] >> 31) ? 1 : -1;
    return 0;
}

// The double next to a positive finite x, above (dir = +1) or below (dir = -1).
__device__ __forceinline__ double next_double(double x, int dir)
{
    return __longlong_as_double(__double_as_longlong(x) + (long long) dir);
}

// Whether rounding the double x to float can change when x moves by up to 32 of
// its ulps: only if x lies within 32 ulps of a float rounding boundary, the
// midpoint of two adjacent floats, whose 29 low significand bits are 1 followed
// by 28 zeros in x's binade (a boundary in the neighbouring binade is at least
// 2^-25 x away).
//
// The kernels use it to compute a power cheaply and take the exact path (the
// correctly rounded power, corrected by the pow tables) only when the float they
// keep can depend on the difference. A power p' within k ulps of the library's
// p moves a product c*p by at most 2k ulps before rounding (the relative spacing
// of doubles is within a factor 2), so fl(c*p) by at most 2k + 1, and a sum
// fl(D + fl(c*p)) with D >= 0 by at most 2k + 2 of its own ulps. The cheap cube
// fl(fl(t*t)*t) is within 2 ulps of t^3 and pow(t, 3) within 1.5 (glibc: one
// ulp from the correctly rounded cube at most, enumerated), so k <= 3.5 and the
// sum moves by at most 9 ulps; the square hr*hr is correctly rounded, pow(hr, 2)
// at most one ulp away. Outside the window every choice gives the same float.
__device__ __forceinline__ bool near_float_boundary(double x)
{
    const uint32_t lo = (uint32_t) __double_as_longlong(x);
    return ((lo & 0x1FFFFFFFu) - (0x10000000u - 32u)) < 64u;
}

// The library's pow(t, 3) for t = h^2 - r^2 of the float r^2 with bits r2_bits:
// the correctly rounded cube, moved to its neighbour where the table lists r^2.
__device__ __forceinline__ double pow3_lib(double t, uint32_t r2_bits, const uint32_t *tab, uint32_t n)
{
    double c = cube_cr(t);
    const int d = pow_exception(tab, n, r2_bits);
    return d ? next_double(c, d) : c;
}

// The library's pow(hr, 2) for hr = h - sqrt(r^2) of the float r^2 with bits r2_bits.
__device__ __forceinline__ double pow2_lib(double hr, uint32_t r2_bits, const uint32_t *tab, uint32_t n)
{
    const double q = hr * hr;
    const int d = pow_exception(tab, n, r2_bits);
    return d ? next_double(q, d) : q;
}

// One term of updateDensityPressure (src/CCPUParticleSimulator.cpp:126): the
// float density plus poly6 * pow(hsq - r^2, 3) (Wpoly6, :14) in double, rounded
// to float. The cube is computed as fl(fl(t*t)*t); where the float result could
// depend on its last bits (near_float_boundary), the sum is redone with the
// library's cube (pow3_lib, table `tab3`).
//
// Split in two so that a caller can compute the terms of several pairs before
// adding them in order: density_poly is the part that does not depend on the
// running density, density_accumulate the add and its rounding.
__device__ __forceinline__ double density_poly(double t, double poly6)
{
    return poly6 * ((t * t) * t);
}

__device__ __forceinline__ float density_accumulate(float density, double poly, double t, float r2, double poly6,
                                                    const uint32_t *tab3, uint32_t n3)
{
    double sum = (double) density + poly;
    if (__builtin_expect(near_float_boundary(sum), 0))
        sum = (double) density + poly6 * pow3_lib(t, __float_as_uint(r2), tab3, n3);
    return (float) sum;
}

__device__ __forceinline__ float density_term(float density, float r2, double hsq, double poly6,
                                              const uint32_t *tab3, uint32_t n3)
{
    const double t = hsq - (double) r2;
    return density_accumulate(density, density_poly(t, poly6), t, r2, poly6, tab3, n3);
}

// (float)(spiky * pow(hr, 2)), the factor of WspikyGradient
// (src/CCPUParticleSimulator.cpp:21) that meets a QVector3D: the correctly
// rounded square, and the library's square (pow2_lib, table `tab2`) where the
// float could depend on the difference.
__device__ __forceinline__ float spiky_factor(double hr, float r2, double spiky, const uint32_t *tab2, uint32_t n2)
{
    double sg = spiky * (hr * hr);
    if (__builtin_expect(near_float_boundary(sg), 0)) sg = spiky * pow2_lib(hr, __float_as_uint(r2), tab2, n2);
    return (float) sg;
}

// sqrt(x), correctly rounded, for x the double of a finite non-negative float
// (the r^2 of an accepted pair, src/CCPUParticleSimulator.cpp:20): from the
// hardware estimates g = x * rsq(x) and rsq(x) / 2, two Newton corrections
// g += (x - g^2) * rsq(x) / 2, each residual exact by FMA, and sqrt(0) = 0
// selected. Not correctly rounded for every double, but for every float in
// [0, h^2]: checked against sqrt() for each of them on the device (a third
// fewer FP64 operations than the compiler's Goldschmidt sequence for a double
// sqrt, which this replaced after the same check).
__device__ __forceinline__ double sqrt_with_rsq(double x, double y)
{
    const double hh = y * 0.5;
    double g = x * y;
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, hh, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, hh, g);
    return x == 0.0 ? 0.0 : g;
}

__device__ __forceinline__ double sqrt_of_float(double x)
{
    return sqrt_with_rsq(x, __builtin_amdgcn_rsq(x));
}

// 1/b from an estimate y of it good to about 2^-22 (relative): two
// Newton-Raphson steps. The force kernels pass rsq(r^2) for b = (float) r, which
// is within about 2^-23 of 1/b; a device test checks |1 - b * rb| <= 2^-52 for
// every float r^2 in (0, h^2]. (After one step it is about 2^-46.)
__device__ __forceinline__ double rcp_from(double b, double y)
{
    double e = __builtin_fma(-b, y, 1.0);
    double r = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Float quotients a/b from double products a * rb, rb ~ 1/b to within 2^-52.
// For float a, b whose quotient rounds to a normal float, the exact a/b is never
// within 2^-49 (relative) of a float rounding midpoint: a = b*m for a 25-bit
// midpoint m is impossible, a 24-bit midpoint (the one between the largest
// subnormal and the smallest normal) is hit exactly only when b is a power of
// two, where rb and the product are exact, and otherwise |a - b*m| is at least
// one unit of b*m's last bit. So the rounded product is the IEEE quotient. A
// result that rounds to zero or to a subnormal is redone with the IEEE division.
__device__ __forceinline__ bool zero_or_subnormal(float q)
{
    return __builtin_amdgcn_classf(q, 0x0F0);   // -subnormal, -0, +0, +subnormal
}

__device__ __forceinline__ bool subnormal(float q)
{
    return __builtin_amdgcn_classf(q, 0x090);   // -subnormal, +subnormal
}

// Whether none of three quotients is zero or subnormal: their smallest
// magnitude is at least the smallest normal float (NaNs drop out of the
// minimum, and are neither zero nor subnormal). One minimum and one compare
// before the per-quotient tests, which only run when this fails.
__device__ __forceinline__ bool all_normal3(float q0, float q1, float q2)
{
    return __builtin_fminf(__builtin_fminf(__builtin_fabsf(q0), __builtin_fabsf(q1)), __builtin_fabsf(q2)) >=
           0x1p-126f;
}

// a_c / b for b in (0, 1): |a/b| > |a| >= 2^-149 for a != 0, so a zero result
// means a = 0 (exact); only subnormal results take the IEEE division.
__device__ __forceinline__ void div3_small_divisor(float a0, float a1, float a2, float b, float &q0, float &q1,
                                                  float &q2, double est)
{
    const double rb = rcp_from((double) b, est);
    q0 = (float) ((double) a0 * rb);
    q1 = (float) ((double) a1 * rb);
    q2 = (float) ((double) a2 * rb);
    if (__builtin_expect(!all_normal3(q0, q1, q2), 0) &&
        ((int) subnormal(q0) | (int) subnormal(q1) | (int) subnormal(q2))) {
        q0 = a0 / b;
        q1 = a1 / b;
        q2 = a2 / b;
    }
}

// a_c / b with rb = 1/b rounded to double: zero or subnormal results of a
// non-zero a take the IEEE division (b is re-read through `divisor` only then).
template <typename Divisor>
__device__ __forceinline__ void div3_rcp(float a0, float a1, float a2, double rb, Divisor divisor, float &q0,
                                         float &q1, float &q2)
{
    q0 = (float) ((double) a0 * rb);
    q1 = (float) ((double) a1 * rb);
    q2 = (float) ((double) a2 * rb);
    if (__builtin_expect(all_normal3(q0, q1, q2), 1)) return;
    const bool t0 = zero_or_subnormal(q0) & (a0 != 0.0f);
    const bool t1 = zero_or_subnormal(q1) & (a1 != 0.0f);
    const bool t2 = zero_or_subnormal(q2) & (a2 != 0.0f);
    if (__builtin_expect((int) t0 | (int) t1 | (int) t2, 0)) {
        const float b = divisor();
        q0 = a0 / b;
        q1 = a1 / b;
        q2 = a2 / b;
    }
}

}  // namespace sph

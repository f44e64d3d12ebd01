/*
 * pt_math.h -- deterministic single-precision math shared by the HIP kernels
 * and by the host CPU build of the same arithmetic.
 *
 * Every function here is written so that the host (gcc, -ffp-contract=off,
 * hardware FMA) and the device (hipcc for gfx950, -ffp-contract=off) execute
 * the identical sequence of correctly rounded IEEE-754 operations, so results
 * agree bit for bit.  Nothing here calls libm transcendentals.
 *
 *  - pt_sincosf / pt_sinf / pt_cosf:
 *      Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software Manual
 *      for the Elementary Functions", 1980): k = rint(x * 2/pi) and
 *      r = x - k*(C1 + C2 + C3) in three fused steps, followed by the truncated
 *      Taylor series of sin and cos on [-pi/4, pi/4] (truncation error below
 *      2e-9, i.e. far under half an ulp) evaluated by Horner's rule with fmaf.
 *      The reference calls cos()/sin() on float angles in
 *      screenToExemplarBasis (texSynth_util.h:95-104) and in the scatter
 *      direction (texSynth_kernel.cu:918-919).
 *  - pt_atan2f:
 *      octant reduction to t in [0,1], the Cody & Waite step
 *      atan(t) = pi/6 + atan((sqrt3*t - 1)/(t + sqrt3)) for t > 2 - sqrt3,
 *      and the Taylor series of atan to t^13 on |t| <= 0.268.  The reference
 *      calls atan2f when converting a structure tensor to an angle
 *      (imageIO.h:136-139).
 *  - pt_philox4x32_10:
 *      Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes,
 *      R. O. Dror, D. E. Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
 *      SC'11).  It replaces the host glibc rand() stream of
 *      synthesisFrame.cpp:66-150 so the random pyramid is a pure function of
 *      (seed, frame*level, x, y, draw).
 *
 * This header compiles as C99 (host) and as HIP C++ (device and host).
 */
#ifndef PT_MATH_H
#define PT_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PT_HD __host__ __device__ static inline
#else
#define PT_HD static inline
#endif

/* pi/2 split for Cody-Waite reduction: C1 + C2 + C3 ~= pi/2. */
#define PT_PIO2_C1 1.57079637050628662109375f
#define PT_PIO2_C2 (-4.371138828673793e-08f)
#define PT_PIO2_C3 (-1.7151245100058819e-15f)
#define PT_TWO_OVER_PI 0.636619772367581343f

PT_HD void pt_sincosf(float x, float *s_out, float *c_out)
{
    float k = rintf(x * PT_TWO_OVER_PI);
    float r = fmaf(-k, PT_PIO2_C1, x);
    r = fmaf(-k, PT_PIO2_C2, r);
    r = fmaf(-k, PT_PIO2_C3, r);
    float r2 = r * r;
    /* sin(r) = r + r^3 * (-1/3! + r^2 (1/5! + r^2 (-1/7! + r^2 / 9!))) */
    float ps = fmaf(r2, 2.75573192239858907e-06f, -1.98412698412698413e-04f);
    ps = fmaf(r2, ps, 8.33333333333333333e-03f);
    ps = fmaf(r2, ps, -1.66666666666666667e-01f);
    float sr = fmaf(r * r2, ps, r);
    /* cos(r) = 1 + r^2 * (-1/2! + r^2 (1/4! + r^2 (-1/6! + r^2 (1/8! - r^2/10!)))) */
    float pc = fmaf(r2, -2.75573192239858907e-07f, 2.48015873015873016e-05f);
    pc = fmaf(r2, pc, -1.38888888888888889e-03f);
    pc = fmaf(r2, pc, 4.16666666666666667e-02f);
    pc = fmaf(r2, pc, -0.5f);
    float cr = fmaf(r2, pc, 1.0f);
    int q = ((int)k) & 3;
    float s, c;
    if (q == 0) { s = sr; c = cr; }
    else if (q == 1) { s = cr; c = -sr; }
    else if (q == 2) { s = -sr; c = -cr; }
    else { s = -cr; c = sr; }
    *s_out = s;
    *c_out = c;
}

PT_HD float pt_sinf(float x) { float s, c; pt_sincosf(x, &s, &c); return s; }
PT_HD float pt_cosf(float x) { float s, c; pt_sincosf(x, &s, &c); return c; }

#define PT_SQRT3_F 1.73205080756887729f
#define PT_PI_6_F 0.523598775598298873f
#define PT_PI_2_F 1.57079632679489662f
#define PT_PI_F 3.14159265358979324f

/* atan(t) for t in [0, 1]. */
PT_HD float pt_atan_unit(float t)
{
    float off = 0.0f;
    if (t > 0.267949192431122706f) {
        t = fmaf(t, PT_SQRT3_F, -1.0f) / (t + PT_SQRT3_F);
        off = PT_PI_6_F;
    }
    float t2 = t * t;
    float p = fmaf(t2, 7.69230769230769231e-02f, -9.09090909090909091e-02f);
    p = fmaf(t2, p, 1.11111111111111111e-01f);
    p = fmaf(t2, p, -1.42857142857142857e-01f);
    p = fmaf(t2, p, 2.00000000000000000e-01f);
    p = fmaf(t2, p, -3.33333333333333333e-01f);
    float a = fmaf(t * t2, p, t);
    return off + a;
}

PT_HD float pt_atan2f(float y, float x)
{
    float ax = fabsf(x), ay = fabsf(y);
    float a;
    if (ax == 0.0f && ay == 0.0f) {
        a = signbit(x) ? PT_PI_F : 0.0f;
    } else if (ay > ax) {
        a = PT_PI_2_F - pt_atan_unit(ax / ay);
        if (signbit(x)) a = PT_PI_F - a;
    } else {
        a = pt_atan_unit(ay / ax);
        if (signbit(x)) a = PT_PI_F - a;
    }
    return copysignf(a, y);
}

/*
 * Angle of a 2-D structure tensor (xx, xy, yy), as strucTensorToAngle
 * (imageIO.h:136-139): theta = 0.5*atan2(2xy, xx-yy) + pi/2 evaluated in
 * double and rounded to float, then folded from [pi/2, ...) down by pi.
 */
PT_HD float pt_tensor_to_angle(float xx, float xy, float yy)
{
    float at = pt_atan2f((float)(2.0 * (double)xy), xx - yy);
    float angle = (float)(0.5 * (double)at + 1.57079632679489661923);
    if ((double)angle >= 1.57079632679489661923)
        return (float)((double)angle - 3.14159265358979323846);
    return angle;
}

/* Philox4x32-10. ctr[4], key[2] in; out[4]. */
#define PT_PHILOX_M0 0xD2511F53u
#define PT_PHILOX_M1 0xCD9E8D57u
#define PT_PHILOX_W0 0x9E3779B9u
#define PT_PHILOX_W1 0xBB67AE85u

PT_HD void pt_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PT_PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)PT_PHILOX_M1 * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PT_PHILOX_W0;
        k1 += PT_PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* A 31-bit draw with glibc rand()'s range [0, 2^31-1]. */
PT_HD int32_t pt_rand31(uint32_t word) { return (int32_t)(word >> 1); }

/* rand()/RAND_MAX as evaluated in float by synthesisFrame.cpp:133-134. */
PT_HD float pt_rand_unit(uint32_t word)
{
    return (float)pt_rand31(word) / (float)2147483647;
}

#endif /* PT_MATH_H */

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
 *  - pt_logf:
 *      x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s) with
 *      s = (m-1)/(m+1), |s| <= 0.1716, by the series 2s (1 + s^2/3 + s^4/5 +
 *      ... + s^10/11) (truncation below 5e-11 relative), plus e*ln2 with ln2
 *      split in a 16-bit head and a tail.  The reference calls logf in the
 *      coherence term (texSynth_kernel.cu:396,400).
 *  - pt_acosf:
 *      acos(x) = pi/2 - asin(x) for |x| <= 1/2, 2 asin(sqrt((1-x)/2)) above
 *      and pi - 2 asin(sqrt((1+x)/2)) below, asin(z) on |z| <= 1/2 by its
 *      Taylor series to z^25 (truncation below 4e-10 relative); NaN outside
 *      [-1, 1].  The reference calls acosf in the coherence term
 *      (texSynth_kernel.cu:417).
 *  - pt_coherence_term: the coherence energy of one tap (below).
 *  - pt_log_d / pt_exp_d / pt_pow_d (double): log as pt_logf with the series
 *      to s^25; exp by k = rint(x/ln2), r = x - k ln2 (two-part ln2) and the
 *      Taylor series to r^14 (|r| <= 0.35, truncation below 1e-19), scaled by
 *      2^k; pow(x, y) = exp(y log x) for x > 0.  Used by the colour
 *      conversion below, where the reference calls pow in double and powf.
 *  - pt_rgb_to_lab: RGBtoLAB (texSynth_util.h:220-249) with Clinear and f_Lab
 *      (:166-195), keeping the reference's mix of double and float
 *      arithmetic; powf(c, 3) and powf(c, 2) of its constants are the
 *      correctly rounded float values.
 *  - pt_rgb_to_lab_fast: the same function, bit for bit, with its two powers
 *      evaluated by inverse-root Newton iterations (below); the kernels use
 *      it, the CPU oracle the plain form.
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
/* Color4::toLuminance (types.h:87-89) of the per-channel difference of two
 * colours: 0.2126 r + 0.7152 g + 0.0722 b, the products and sums in double
 * (the constants are double literals), rounded to float. */
PT_HD float pt_luminance_delta(float cr, float cg, float cb, float pr, float pg, float pb)
{
    float r = cr - pr, g = cg - pg, b = cb - pb;
    double y = 0.2126 * (double)r;
    y = y + 0.7152 * (double)g;
    y = y + 0.0722 * (double)b;
    return (float)y;
}

PT_HD float pt_tensor_to_angle(float xx, float xy, float yy)
{
    float at = pt_atan2f((float)(2.0 * (double)xy), xx - yy);
    float angle = (float)(0.5 * (double)at + 1.57079632679489661923);
    if ((double)angle >= 1.57079632679489661923)
        return (float)((double)angle - 3.14159265358979323846);
    return angle;
}

/* ln 2 = PT_LN2_HI + PT_LN2_LO; the head has 16 significant bits, so
 * e * PT_LN2_HI is exact for |e| < 256. */
#define PT_LN2_HI 0.693145751953125f
#define PT_LN2_LO 1.42860682030941723212e-06f
#define PT_SQRT1_2_F 0.707106781186547524f

PT_HD float pt_logf(float x)
{
    if (!(x >= 0.0f)) return NAN; /* negative or NaN */
    if (x == 0.0f) return -INFINITY;
    if (isinf(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { /* subnormal: scale by 2^24 (exact) */
        x *= 16777216.0f;
        e = -24;
    }
    union {
        float f;
        uint32_t u;
    } b;
    b.f = x;
    e += (int)((b.u >> 23) & 0xffu) - 127;
    b.u = (b.u & 0x007fffffu) | 0x3f800000u; /* m in [1, 2) */
    float m = b.f;
    if (m >= 1.41421356237309505f) {
        m *= 0.5f; /* exact */
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f); /* m - 1 is exact */
    float z = s * s;
    float p = fmaf(z, 9.09090909090909091e-02f, 1.11111111111111111e-01f);
    p = fmaf(z, p, 1.42857142857142857e-01f);
    p = fmaf(z, p, 2.00000000000000000e-01f);
    p = fmaf(z, p, 3.33333333333333333e-01f);
    float lm = 2.0f * fmaf(s * z, p, s); /* 2 (s + s^3 p) */
    float ef = (float)e;
    return ef * PT_LN2_HI + fmaf(ef, PT_LN2_LO, lm);
}

/* asin(z) for |z| <= 1/2: z + z^3 * P(z^2), P's coefficients the Taylor
 * series (2k)! / (4^k (k!)^2 (2k+1)), k = 1..12. */
PT_HD float pt_asin_half(float z)
{
    float t = z * z;
    float p = fmaf(t, 6.44721031188964875e-03f, 7.31252587359884545e-03f);
    p = fmaf(t, p, 8.39033580961681506e-03f);
    p = fmaf(t, p, 9.76160952919407840e-03f);
    p = fmaf(t, p, 1.15518008961397051e-02f);
    p = fmaf(t, p, 1.39648437500000007e-02f);
    p = fmaf(t, p, 1.73527644230769239e-02f);
    p = fmaf(t, p, 2.23721590909090919e-02f);
    p = fmaf(t, p, 3.03819444444444441e-02f);
    p = fmaf(t, p, 4.46428571428571438e-02f);
    p = fmaf(t, p, 7.49999999999999972e-02f);
    p = fmaf(t, p, 1.66666666666666657e-01f);
    return fmaf(z * t, p, z);
}

/* pi/2 = PT_PIO2_HI + PT_PIO2_LO */
#define PT_PIO2_HI 1.57079637050628662109375f
#define PT_PIO2_LO (-4.37113882867379e-08f)

PT_HD float pt_acosf(float x)
{
    if (!(x >= -1.0f && x <= 1.0f)) return NAN;
    if (x > 0.5f) {
        float z = (1.0f - x) * 0.5f; /* exact */
        return 2.0f * pt_asin_half(sqrtf(z));
    }
    if (x < -0.5f) {
        float z = (1.0f + x) * 0.5f; /* exact */
        float a = 2.0f * pt_asin_half(sqrtf(z));
        return (2.0f * PT_PIO2_HI - a) + 2.0f * PT_PIO2_LO;
    }
    return (PT_PIO2_HI - pt_asin_half(x)) + PT_PIO2_LO;
}

/*
 * coherenceTerm (texSynth_kernel.cu:377-423) for tap (i, j) != (0, 0) of a
 * candidate centred at (tcx, tcy) whose tap lands at (tx, ty) (after the
 * toroidal wrap), with (cx, cy) the current offset of pixel (x+i, y+j):
 * the squared log of the exemplar/image distance ratio (clamped to
 * [1/ratio_range, ratio_range], normalised by log(ratio_range)^2) and the
 * angle between the two directions (acos clamped to [-1, 1], as the
 * reference's clamp does, then relative to angle_range*pi), combined as
 * 1 - (1 - Eratio)(1 - Eangle).  The reference forms angle_range*pi and the
 * min/division in double (M_PI is a double) and stores the result in a
 * float; the same is done here.  Both sides include this one definition.
 */
/* pt_coherence_term with its launch constants formed by the caller:
 * inv_ratio = 1 / ratio_range, z = log(ratio_range)^2 (pt_logf, squared in
 * float) and lim = angle_range * pi in double */
PT_HD float pt_coherence_term_k(float cx, float cy, float tcx, float tcy, float tx, float ty, int i, int j, int level,
                                float inv_ratio, float ratio_range, float z, double lim)
{
    float dx = cx - tcx, dy = cy - tcy;
    float lvl = (float)(1 << level);
    float dist_ex = (dx * dx + dy * dy) * lvl;
    float dist_im = (float)(i * i + j * j) * lvl;
    float ratio = fmaxf(inv_ratio, fminf(dist_ex / dist_im, ratio_range));
    float lr = pt_logf(ratio);
    float e_ratio = (lr * lr) / z;
    float ux = tx - tcx, uy = ty - tcy;
    float lt = sqrtf(ux * ux + uy * uy);
    ux = (lt > 1e-4f) ? ux * (1.0f / lt) : 1.0f;
    uy = (lt > 1e-4f) ? uy * (1.0f / lt) : 0.0f;
    float lc = sqrtf(dx * dx + dy * dy);
    float vx = (lc > 1e-4f) ? dx * (1.0f / lc) : 1.0f;
    float vy = (lc > 1e-4f) ? dy * (1.0f / lc) : 0.0f;
    float d = ux * vx + uy * vy;
    float angle = fmaxf(-1.0f, fminf(pt_acosf(d), 1.0f));
    float e_angle = (float)(fmin((double)angle, lim) / lim);
    return 1.0f - (1.0f - e_ratio) * (1.0f - e_angle);
}

PT_HD float pt_coherence_term(float cx, float cy, float tcx, float tcy, float tx, float ty, int i, int j, int level,
                              float ratio_range, float angle_range)
{
    float lz = pt_logf(ratio_range);
    return pt_coherence_term_k(cx, cy, tcx, tcy, tx, ty, i, j, level, 1.0f / ratio_range, ratio_range, lz * lz,
                               (double)angle_range * 3.14159265358979323846);
}

/* ln 2 = PT_LN2_HI_D + PT_LN2_LO_D, the head with 32 significant bits */
#define PT_LN2_HI_D 6.93147180369123816490e-01
#define PT_LN2_LO_D 1.90821492927058770002e-10

PT_HD double pt_log_d(double x)
{
    if (!(x >= 0.0)) return NAN;
    if (x == 0.0) return -INFINITY;
    if (isinf(x)) return x;
    int e = 0;
    if (x < 2.2250738585072014e-308) { /* subnormal: scale by 2^54 (exact) */
        x *= 18014398509481984.0;
        e = -54;
    }
    union {
        double d;
        uint64_t u;
    } b;
    b.d = x;
    e += (int)((b.u >> 52) & 0x7ffu) - 1023;
    b.u = (b.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull; /* m in [1, 2) */
    double m = b.d;
    if (m >= 1.4142135623730951) {
        m *= 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 27.0;
    p = fma(z, p, 1.0 / 25.0);
    p = fma(z, p, 1.0 / 23.0);
    p = fma(z, p, 1.0 / 21.0);
    p = fma(z, p, 1.0 / 19.0);
    p = fma(z, p, 1.0 / 17.0);
    p = fma(z, p, 1.0 / 15.0);
    p = fma(z, p, 1.0 / 13.0);
    p = fma(z, p, 1.0 / 11.0);
    p = fma(z, p, 1.0 / 9.0);
    p = fma(z, p, 1.0 / 7.0);
    p = fma(z, p, 1.0 / 5.0);
    p = fma(z, p, 1.0 / 3.0);
    double lm = 2.0 * fma(s * z, p, s);
    double ef = (double)e;
    return ef * PT_LN2_HI_D + fma(ef, PT_LN2_LO_D, lm);
}

PT_HD double pt_exp_d(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double k = rint(x * 1.4426950408889634);
    double r = fma(-k, PT_LN2_HI_D, x);
    r = fma(-k, PT_LN2_LO_D, r);
    double p = 1.0 / 87178291200.0; /* 1/14! */
    p = fma(r, p, 1.0 / 6227020800.0);
    p = fma(r, p, 1.0 / 479001600.0);
    p = fma(r, p, 1.0 / 39916800.0);
    p = fma(r, p, 1.0 / 3628800.0);
    p = fma(r, p, 1.0 / 362880.0);
    p = fma(r, p, 1.0 / 40320.0);
    p = fma(r, p, 1.0 / 5040.0);
    p = fma(r, p, 1.0 / 720.0);
    p = fma(r, p, 1.0 / 120.0);
    p = fma(r, p, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = fma(r, p, 1.0);
    return ldexp(p, (int)k);
}

/* x^y for x > 0 (0 for x == 0 and y > 0) */
PT_HD double pt_pow_d(double x, double y)
{
    if (x == 0.0) return y > 0.0 ? 0.0 : (y == 0.0 ? 1.0 : INFINITY);
    if (!(x > 0.0)) return NAN;
    return pt_exp_d(y * pt_log_d(x));
}

PT_HD float pt_powf(float x, float y) { return (float)pt_pow_d((double)x, (double)y); }

/* Clinear (texSynth_util.h:175-183): (c + a) / (1.0 + a) and the division by
 * 12.92 in double, pow in double */
PT_HD float pt_srgb_to_linear(float c)
{
    const float a = 0.055f;
    if ((double)c > 0.04045) return (float)pt_pow_d((double)(c + a) / (1.0 + (double)a), 2.4);
    return (float)((double)c / 12.92);
}

/* f_Lab (texSynth_util.h:185-192): powf((float)(6/29), 3) and
 * powf((float)(29/6), 2) correctly rounded; the linear branch in double */
#define PT_LAB_T0 8.856453001e-03f
#define PT_LAB_K2 2.336111259e+01f
PT_HD float pt_f_lab(float t)
{
    if (t > PT_LAB_T0) return pt_powf(t, 0.3333333432674408f);
    return (float)(((1.0 / 3.0) * (double)PT_LAB_K2) * (double)t + (4.0 / 29.0));
}

/* RGBtoLAB (texSynth_util.h:220-249): D65 white, L/100, (a+100)/200,
 * (b+100)/200; the XYZ matrix products in double; alpha unchanged */
PT_HD void pt_rgb_to_lab(float r, float g, float b, float *lo, float *ao, float *bo)
{
    const float ref_x = 95.0469970703125f, ref_y = 100.0f, ref_z = 108.88300323486328f;
    float rl = 100 * pt_srgb_to_linear(r);
    float gl = 100 * pt_srgb_to_linear(g);
    float bl = 100 * pt_srgb_to_linear(b);
    float x = (float)(0.4124 * (double)rl + 0.3576 * (double)gl + 0.1805 * (double)bl);
    float y = (float)(0.2126 * (double)rl + 0.7152 * (double)gl + 0.0722 * (double)bl);
    float z = (float)(0.0193 * (double)rl + 0.1192 * (double)gl + 0.9505 * (double)bl);
    float fy = pt_f_lab(y / ref_y);
    *lo = (116.f * fy - 16) / 100.f;
    *ao = (500.f * (pt_f_lab(x / ref_x) - fy) + 100) / 200.f;
    *bo = (200.f * (fy - pt_f_lab(z / ref_z)) + 100) / 200.f;
}

/*
 * Fast forms of the two powers of the colour conversion, equal bit for bit
 * to pt_srgb_to_linear and pt_f_lab.  Inside a fast range each forms the
 * power in double from an inverse root w: an exponent-bits estimate (the
 * float's bit pattern read as a scaled base-2 logarithm), three Newton steps
 * in float, one in double, which leaves the power within about 2^-41
 * relative of its exact value.  The result is rounded to float only when
 * every value within 2^-38 relative of it rounds to the same float, which
 * then is also the rounding of the plain form's double-series value (within
 * 2^-48 of the exact power); otherwise, and outside the fast range, the plain
 * form is evaluated.  The identity is checked over every float input of the
 * fast ranges by tests/lab_fast_check.c.
 */
#define PT_ROOT5_MAGIC 0x4c2c47e6u /* 6/5 * 2^23 * (127 - 0.0450465) */
#define PT_ROOT3_MAGIC 0x54a2fa8eu /* 4/3 * 2^23 * (127 - 0.0450465) */
#define PT_SRGB_K 0x1.c242be51c25dfp-1 /* (1 + (double)0.055f)^(-12/5) */
#define PT_LAB_DP 0x1.5555555555555p-27 /* (double)0.33333334f - 1/3 */
#define PT_FAST_TOL 0x1p-38

PT_HD float pt_bits_to_float(uint32_t u)
{
    union {
        uint32_t u;
        float f;
    } b;
    b.u = u;
    return b.f;
}

PT_HD uint32_t pt_float_to_bits(float f)
{
    union {
        float f;
        uint32_t u;
    } b;
    b.f = f;
    return b.u;
}

/* (v / (1 + (double)0.055f))^2.4 for v = c + 0.055f in (0.0955, 2.06]:
 * (v w)^3 K with w = v^(-1/5); the exponent 12/5 differs from (double)2.4
 * by 9e-17. */
PT_HD double pt_srgb_pow_est(float v)
{
    float w = pt_bits_to_float(PT_ROOT5_MAGIC - pt_float_to_bits(v) / 5u);
    for (int k = 0; k < 3; k++) {
        float w2 = w * w;
        float q = v * (w2 * w2 * w);
        w = w * fmaf(-0.2f, q, 1.2f);
    }
    double V = (double)v, W = (double)w;
    double W2 = W * W;
    W = W * fma(-0.2, V * (W2 * W2 * W), 1.2);
    double p = V * W;
    return p * p * p * PT_SRGB_K;
}

/* t^p for p = (double)0.33333334f and t in (PT_LAB_T0, 4]: t w^2 with
 * w = t^(-1/3), times t^(p - 1/3) = 1 + (p - 1/3) log t (the next term is
 * below 2e-15). */
PT_HD double pt_f_lab_pow_est(float t)
{
    float w = pt_bits_to_float(PT_ROOT3_MAGIC - pt_float_to_bits(t) / 3u);
    for (int k = 0; k < 3; k++) {
        float q = t * (w * w * w);
        w = w * fmaf(-0.333333343f, q, 1.33333337f);
    }
    double T = (double)t, W = (double)w;
    W = W * fma(-1.0 / 3.0, T * (W * W * W), 4.0 / 3.0);
    return (T * (W * W)) * fma(PT_LAB_DP, (double)pt_logf(t), 1.0);
}

/* 1 when every value within PT_FAST_TOL relative of r > 0 rounds to the
 * same float, stored in *f */
PT_HD int pt_round_decided(double r, float *f)
{
    float lo = (float)(r * (1.0 - PT_FAST_TOL)), hi = (float)(r * (1.0 + PT_FAST_TOL));
    *f = lo;
    return lo == hi;
}

PT_HD float pt_srgb_to_linear_fast(float c)
{
    if ((double)c > 0.04045 && c <= 2.0f) {
        float f;
        if (pt_round_decided(pt_srgb_pow_est(c + 0.055f), &f)) return f;
    }
    return pt_srgb_to_linear(c);
}

PT_HD float pt_f_lab_fast(float t)
{
    if (t > PT_LAB_T0 && t <= 4.0f) {
        float f;
        if (pt_round_decided(pt_f_lab_pow_est(t), &f)) return f;
    }
    return pt_f_lab(t);
}

/* pt_rgb_to_lab with the fast powers: the same operations otherwise */
PT_HD void pt_rgb_to_lab_fast(float r, float g, float b, float *lo, float *ao, float *bo)
{
    const float ref_x = 95.0469970703125f, ref_y = 100.0f, ref_z = 108.88300323486328f;
    float rl = 100 * pt_srgb_to_linear_fast(r);
    float gl = 100 * pt_srgb_to_linear_fast(g);
    float bl = 100 * pt_srgb_to_linear_fast(b);
    float x = (float)(0.4124 * (double)rl + 0.3576 * (double)gl + 0.1805 * (double)bl);
    float y = (float)(0.2126 * (double)rl + 0.7152 * (double)gl + 0.0722 * (double)bl);
    float z = (float)(0.0193 * (double)rl + 0.1192 * (double)gl + 0.9505 * (double)bl);
    float fy = pt_f_lab_fast(y / ref_y);
    *lo = (116.f * fy - 16) / 100.f;
    *ao = (500.f * (pt_f_lab_fast(x / ref_x) - fy) + 100) / 200.f;
    *bo = (200.f * (fy - pt_f_lab_fast(z / ref_z)) + 100) / 200.f;
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

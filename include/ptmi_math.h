/*
 * ptmi_math.h — the declared fp32 arithmetic of the transport loop.
 *
 * GLSL ES leaves the precision of sin/cos/exp/log/atan/pow implementation-defined, so
 * the reference program (its generated transport loop, e.g.
 * tests/compiler/__snapshots__/generated-glsl.snapshot.test.ts.snap:8176) has no single
 * bit-level answer for them. This header fixes one: every function below is built only
 * from IEEE-754 binary32 +, -, *, / and sqrt (all correctly rounded on gfx950 and on
 * x86-64 SSE), evaluated in a fixed order with no contraction into FMA. The same
 * source therefore produces the same bits on the CPU (gcc -ffp-contract=off) and on the
 * GPU (hipcc -ffp-contract=off), which is what lets per-pixel radiance be compared
 * bit for bit between the HIP kernels and the CPU restatement.
 *
 * Methods (textbook, no vendor constants):
 *   sin/cos  Cody–Waite reduction by pi/2 (three-part constant, exact k*C1 for
 *            |k| < 2^12), Taylor polynomials to degree 11/12 on [-pi/4, pi/4].
 *   exp      reduction by ln 2 (two-part constant), Taylor to degree 8, exact 2^k scale.
 *   log      m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1), log m = 2 atanh s (odd series).
 *   atan     |x| > 1 → pi/2 - atan(1/x); |x| > tan(pi/12) → pi/6 + atan((x√3-1)/(x+√3));
 *            odd Taylor series to degree 15.
 * Accuracy against correctly rounded results is pinned by tests (a few ulp).
 *
 * Include from C (gcc) or HIP (hipcc). Nothing here allocates or touches global state.
 */
#ifndef PTMI_MATH_H
#define PTMI_MATH_H

#include <stdint.h>
#include <math.h>
#include <string.h>

#if defined(__HIPCC__)
#define PTM_FN __host__ __device__ static inline
#else
#define PTM_FN static inline
#endif

#define PTM_PI      3.14159265359f   /* the GLSL literal PI (glsl/core/math.glsl:6) */
#define PTM_TWO_PI  6.28318530718f   /* the GLSL literal TWO_PI (glsl/core/math.glsl:7) */

PTM_FN uint32_t ptm_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
PTM_FN float ptm_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* 2^k for integer k in [-252, 254], exact (two-step for the subnormal/overflow tails). */
PTM_FN float ptm_ldexp(float x, int k) {
    if (k > 127) { x = x * ptm_u2f(0x7F000000u); k -= 127; if (k > 127) k = 127; }
    if (k < -126) { x = x * ptm_u2f(0x00800000u); k += 126; if (k < -126) k = -126; }
    return x * ptm_u2f((uint32_t)(k + 127) << 23);
}

PTM_FN void ptm_sincos(float x, float* s_out, float* c_out) {
    if (!(fabsf(x) < 8192.0f)) {           /* outside the reduction's exact range (or NaN) */
        if (x != x || fabsf(x) == INFINITY) { *s_out = x - x; *c_out = x - x; return; }
    }
    const float C1 = 1.5703125f;                 /* 0x3fc90000 */
    const float C2 = 4.837512969970703125e-4f;   /* 0x39fda000 */
    const float C3 = 7.549790126404332e-08f;     /* 0x33a22169 */
    float fk = floorf(x * 0.636619746685028076f + 0.5f);
    int k = (int)fk;
    float r = ((x - fk * C1) - fk * C2) - fk * C3;
    float r2 = r * r;
    float ps = ((((-2.50521083854417188e-08f * r2 + 2.75573192239858907e-06f) * r2
                 - 1.98412698412698413e-04f) * r2 + 8.33333333333333333e-03f) * r2
                - 1.66666666666666667e-01f);
    float sn = r + (r * r2) * ps;
    float pc = (((((2.08767569878680990e-09f * r2 - 2.75573192239858907e-07f) * r2
                  + 2.48015873015873016e-05f) * r2 - 1.38888888888888889e-03f) * r2
                 + 4.16666666666666667e-02f));
    float cs = (1.0f - 0.5f * r2) + (r2 * r2) * pc;
    switch (k & 3) {
        case 0: *s_out = sn;  *c_out = cs;  break;
        case 1: *s_out = cs;  *c_out = -sn; break;
        case 2: *s_out = -sn; *c_out = -cs; break;
        default: *s_out = -cs; *c_out = sn; break;
    }
}
PTM_FN float ptm_sin(float x) { float s, c; ptm_sincos(x, &s, &c); return s; }
PTM_FN float ptm_cos(float x) { float s, c; ptm_sincos(x, &s, &c); return c; }
PTM_FN float ptm_tan(float x) { float s, c; ptm_sincos(x, &s, &c); return s / c; }

PTM_FN float ptm_exp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    const float L1 = 0.693145751953125f;         /* 0x3f317200 */
    const float L2 = 1.428606765330187e-06f;     /* 0x35bfbe8e */
    float fk = floorf(x * 1.44269502162933350f + 0.5f);
    int k = (int)fk;
    float r = (x - fk * L1) - fk * L2;
    float p = (((((((2.48015873015873016e-05f * r + 1.98412698412698413e-04f) * r
                    + 1.38888888888888889e-03f) * r + 8.33333333333333333e-03f) * r
                  + 4.16666666666666667e-02f) * r + 1.66666666666666667e-01f) * r
                + 0.5f) * r);
    float e = 1.0f + (r + r * p);
    return ptm_ldexp(e, k);
}

PTM_FN float ptm_log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return (x - x) / (x - x);      /* NaN */
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t b = ptm_f2u(x);
    if (b < 0x00800000u) { x = x * 8388608.0f; b = ptm_f2u(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float m = ptm_u2f((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421353816986084f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = ((((((1.0f / 13.0f) * s2 + (1.0f / 11.0f)) * s2 + (1.0f / 9.0f)) * s2
                 + (1.0f / 7.0f)) * s2 + 0.2f) * s2 + (1.0f / 3.0f));
    float lm = 2.0f * s + (2.0f * s) * (s2 * p);
    float fe = (float)e;
    return fe * 0.693145751953125f + (lm + fe * 1.428606765330187e-06f);
}

/* atan on [0, +inf]; sign handled by callers. */
PTM_FN float ptm_atan_pos(float a) {
    float base = 0.0f;
    int inv = 0;
    if (a > 1.0f) { a = 1.0f / a; inv = 1; }
    if (a > 0.267949193716049194f) {
        a = (a * 1.73205077648162842f - 1.0f) / (a + 1.73205077648162842f);
        base = 0.523598790168762207f;
    }
    float a2 = a * a;
    float p = (((((((-1.0f / 15.0f) * a2 + (1.0f / 13.0f)) * a2 - (1.0f / 11.0f)) * a2
                  + (1.0f / 9.0f)) * a2 - (1.0f / 7.0f)) * a2 + 0.2f) * a2 - (1.0f / 3.0f));
    float r = base + (a + (a * a2) * p);
    return inv ? 1.57079637050628662f - r : r;
}
PTM_FN float ptm_atan(float x) {
    if (x != x) return x;
    float r = ptm_atan_pos(fabsf(x));
    return x < 0.0f ? -r : r;
}
/* GLSL atan(y, x). atan(0, 0) returns 0 (GLSL leaves it undefined). */
PTM_FN float ptm_atan2(float y, float x) {
    if (x != x || y != y) return x + y;
    float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return 0.0f;
    float r;
    if (ay <= ax) r = ptm_atan_pos(ay / ax);
    else r = 1.57079637050628662f - ptm_atan_pos(ax / ay);
    if (x < 0.0f) r = 3.14159274101257324f - r;
    return y < 0.0f ? -r : r;
}
PTM_FN float ptm_acos(float x) {
    return ptm_atan2(sqrtf((1.0f - x) * (1.0f + x)), x);
}
/* The equal-area octahedral environment chart (src/components/env/octahedral/octahedral.glsl;
 * Clarberg 2008, pbrt-v4's ImageInfiniteLight mapping, y up). ptm_oct_dir maps table uv to a
 * direction and applies the inverse of the rotation (c, s); ptm_oct_uv maps a direction, rotated
 * by (c, s), to table uv. */
PTM_FN float ptm_copysign_abs(float m, float s) { return s < 0.0f ? -fabsf(m) : fabsf(m); }
PTM_FN void ptm_oct_dir(float u, float v, float c, float s, float* dx, float* dy, float* dz) {
    float up = 2.0f * u - 1.0f, vp = 2.0f * v - 1.0f;
    float upAbs = fabsf(up), vpAbs = fabsf(vp);
    float sd = 1.0f - (upAbs + vpAbs);
    float d = fabsf(sd);
    float r = 1.0f - d;
    float phi = ((r == 0.0f ? 1.0f : (vpAbs - upAbs) / r) + 1.0f) * PTM_PI * 0.25f;
    float y = ptm_copysign_abs(1.0f - r * r, sd);
    float cosPhi = ptm_copysign_abs(ptm_cos(phi), up);
    float sinPhi = ptm_copysign_abs(ptm_sin(phi), vp);
    float t = 2.0f - r * r;
    float ss = r * sqrtf(t > 0.0f ? t : 0.0f);
    float x = cosPhi * ss, z = sinPhi * ss, ns = -s;           /* env_rotate_cs(d, (c, -s)) */
    *dx = c * x - ns * z; *dy = y; *dz = ns * x + c * z;
}
PTM_FN void ptm_oct_uv(float dx, float dy, float dz, float c, float s, float* uo, float* vo) {
    float il = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    float nx0 = dx * il, ny = dy * il, nz0 = dz * il;
    float nx = c * nx0 - s * nz0, nz = s * nx0 + c * nz0;      /* env_rotate_cs(n, (c, s)) */
    float x = fabsf(nx), y = fabsf(ny), z = fabsf(nz);
    float t = 1.0f - y;
    float r = sqrtf(t > 0.0f ? t : 0.0f);
    float a = x > z ? x : z;
    float b = (a == 0.0f) ? 0.0f : (x < z ? x : z) / a;
    float phi = ptm_atan(b) * (2.0f / PTM_PI);
    if (x < z) phi = 1.0f - phi;
    float vv = phi * r;
    float uu = r - vv;
    if (ny < 0.0f) { float w = uu; uu = vv; vv = w; uu = 1.0f - uu; vv = 1.0f - vv; }
    uu = ptm_copysign_abs(uu, nx);
    vv = ptm_copysign_abs(vv, nz);
    *uo = (uu + 1.0f) * 0.5f; *vo = (vv + 1.0f) * 0.5f;
}
/* asin(x) = atan(x, sqrt(1 - x^2)) on [-1, 1] (exactly +-pi/2 at +-1). */
PTM_FN float ptm_asin(float x) {
    return ptm_atan2(x, sqrtf((1.0f - x) * (1.0f + x)));
}
/* pow(x, y) = exp(y log x) for x > 0; pow(0, y>0) = 0. GLSL leaves negative x undefined;
 * declared here: an integer exponent gives the real power (sign of x^y), any other NaN.
 * Schlick's pow(1 - cos, 5) meets x = -1 ulp when a rounded |dot| exceeds 1. */
PTM_FN float ptm_pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
    if (y == 0.0f) return 1.0f;
    if (x < 0.0f) {
        if (floorf(y) != y) return NAN;
        float r = ptm_exp(y * ptm_log(-x));
        return fmodf(y, 2.0f) != 0.0f ? -r : r;
    }
    return ptm_exp(y * ptm_log(x));
}
PTM_FN float ptm_tanh(float x) {
    float ax = fabsf(x);
    if (ax > 9.0f) return x < 0.0f ? -1.0f : 1.0f;
    float e = ptm_exp(2.0f * ax);
    float r = (e - 1.0f) / (e + 1.0f);
    return x < 0.0f ? -r : r;
}
PTM_FN float ptm_atanh(float x) {
    return 0.5f * ptm_log((1.0f + x) / (1.0f - x));
}
PTM_FN float ptm_cosh(float x) {
    float e = ptm_exp(fabsf(x));
    return 0.5f * (e + 1.0f / e);
}
/* GLSL smoothstep (genType form): t = clamp((x-e0)/(e1-e0), 0, 1); t*t*(3-2t). */
PTM_FN float ptm_smoothstep(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}
/* GLSL fract */
PTM_FN float ptm_fract(float x) { return x - floorf(x); }

#endif /* PTMI_MATH_H */

/*
 * or_math.h -- vector arithmetic of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle is a plain-C restatement of YMRT's
 * host-visible algorithms, used by tests/ and by bench.py's CPU baseline as the
 * checker.  Nothing in the product may include or call it.
 *
 * The helpers follow the glm 0.9.8 formulas the reference calls
 * (external/glm/glm/detail/func_geometric.inl:54-95: dot = (x*x + y*y) + z*z,
 * normalize = v * (1 / sqrt(dot(v, v))), length = sqrt(dot(v, v));
 * func_common.inl:169 mix = x + a * (y - x); type_mat4x4.inl:501-535 mat4*vec4 =
 * (m0*x + m1*y) + (m2*z + m3*w)) and src/MathCommon.h.  Built with
 * -ffp-contract=off so every operation rounds as written.
 */
#ifndef OR_MATH_H
#define OR_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;

#define OR_PI 3.14159265358f
#define OR_TWO_PI 6.28318530716f
#define OR_INV_PI 0.318309886183f
#define OR_FLOAT_EPSILON 1.569209E-07f
#define OR_MIN_COLOR_EPSILON 3.7252903E-9f
#define OR_MAX_COLOR_CLAMP 1E16f
#define OR_SHADOW_RAY_CLAMP 0.9995f
#define OR_FLOAT_MAX 1E36f
#define OR_FLOAT_MIN -1E36f
#define OR_TMAX 1E36f
#define OR_TMIN 1E-7f
#define OR_EMPTY 0xFFFFFFFFu

static inline v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 s3(float s) { v3 r = {s, s, s}; return r; }
static inline v2 mk2(float x, float y) { v2 r = {x, y}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 div3(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 sclf3(float s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline v3 divs3(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b)
{
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float length3(v3 a) { return sqrtf(dot3(a, a)); }
static inline v3 normalize3(v3 a) { return scl3(a, 1.0f / sqrtf(dot3(a, a))); }
static inline v3 abs3(v3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline float fminr(float a, float b) { return b < a ? b : a; } /* glm::min */
static inline float fmaxr(float a, float b) { return a < b ? b : a; } /* glm::max */
static inline v3 min3(v3 a, v3 b) { return mk3(fminr(a.x, b.x), fminr(a.y, b.y), fminr(a.z, b.z)); }
static inline v3 max3(v3 a, v3 b) { return mk3(fmaxr(a.x, b.x), fmaxr(a.y, b.y), fmaxr(a.z, b.z)); }
static inline float clampf(float x, float lo, float hi) { return fminr(fmaxr(x, lo), hi); }
static inline float comp3(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc3(v3 *a, int i, float v) { if (i == 0) a->x = v; else if (i == 1) a->y = v; else a->z = v; }
static inline float maxcomp3(v3 a) { return fmaxr(a.x, fmaxr(a.y, a.z)); }
static inline v3 mix3(v3 x, v3 y, float a) { return add3(x, sclf3(a, sub3(y, x))); }
static inline v3 mixv3(v3 x, v3 y, v3 a) { return add3(x, mul3(a, sub3(y, x))); }
static inline float sqrf(float x) { return x * x; }

/* Transcendentals of float arguments evaluated in double precision and
 * rounded once to float: the correctly rounded result except where the double
 * value lies within one of its own ulps of a float rounding boundary.  CUDA's
 * float functions that the reference calls are specified only to a few ulps;
 * the oracle fixes them to the correctly rounded value, which is also what
 * the GPU library computes, so both sides take the same branches on the last
 * bits of a free-flight distance, a transmittance or a sampled direction.
 * Host-side setup (camera field of view, light cones, transforms) and the
 * indirect clamp's powers keep glibc's float functions, as the library's host
 * code and its clamp keep float functions. */
static inline float or_expf(float x) { return (float)exp((double)x); }
static inline float or_logf(float x) { return (float)log((double)x); }
static inline float or_sinf(float x) { return (float)sin((double)x); }
static inline float or_cosf(float x) { return (float)cos((double)x); }
static inline float or_powf(float x, float y) { return (float)pow((double)x, (double)y); }
static inline float or_acosf(float x) { return (float)acos((double)x); }
static inline float or_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }

/* src/MathCommon.h:103-106 */
static inline float safe_rcp(float f)
{
    return fabsf(f) < 1E-30f ? (f > 0.0f ? 1E30f : -1E30f) : (1.0f / f);
}

/* src/MathCommon.h:151-154 */
static inline float error_gamma(int n)
{
    return ((float)n * OR_FLOAT_EPSILON) / (1.0f - (float)n * OR_FLOAT_EPSILON);
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* src/MathCommon.h:263-297 */
static inline float next_float_up(float f)
{
    if (isinf(f) && f > 0.0f) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = f2u(f);
    if (f >= 0.0f) ++ui; else --ui;
    return u2f(ui);
}
static inline float next_float_down(float f)
{
    if (isinf(f) && f < 0.0f) return f;
    if (f == 0.0f) f = -0.0f;
    uint32_t ui = f2u(f);
    if (f <= 0.0f) ++ui; else --ui;
    return u2f(ui);
}

/* src/MathCommon.h:146-149 (the 0.7152 literal is a double in the reference) */
static inline float luminance3(v3 c)
{
    return (float)((double)(0.2126f * c.x) + 0.7152 * (double)c.y + (double)(0.0722f * c.z));
}

/* column-major 4x4, m[col*4 + row] */
typedef struct { float m[16]; } m4;

/* glm::mat4 * vec4 with w = 1 (position) or 0 (direction):
 * (c0*x + c1*y) + (c2*z + c3*w) per row. */
static inline v3 m4_point(const float *m, v3 p)
{
    v3 r;
    r.x = (m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * 1.0f);
    r.y = (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * 1.0f);
    r.z = (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * 1.0f);
    return r;
}
static inline v3 m4_dir(const float *m, v3 d)
{
    v3 r;
    r.x = (m[0] * d.x + m[4] * d.y) + (m[8] * d.z + m[12] * 0.0f);
    r.y = (m[1] * d.x + m[5] * d.y) + (m[9] * d.z + m[13] * 0.0f);
    r.z = (m[2] * d.x + m[6] * d.y) + (m[10] * d.z + m[14] * 0.0f);
    return r;
}
/* glm::transpose(m) * vec4(v, 0): row i of the result is column i of m dotted
 * with v, summed as (a*x + b*y) + (c*z + d*0). */
static inline v3 m4_transpose_dir(const float *m, v3 d)
{
    v3 r;
    r.x = (m[0] * d.x + m[1] * d.y) + (m[2] * d.z + m[3] * 0.0f);
    r.y = (m[4] * d.x + m[5] * d.y) + (m[6] * d.z + m[7] * 0.0f);
    r.z = (m[8] * d.x + m[9] * d.y) + (m[10] * d.z + m[11] * 0.0f);
    /* m[3], m[7], m[11] are the affine bottom row (0) */
    return r;
}

/* src/MathCommon.h:198-205 TransformVector (left-to-right sums) */
static inline v3 transform_vector(const float *m, v3 v)
{
    return mk3(m[0] * v.x + m[4] * v.y + m[8] * v.z,
               m[1] * v.x + m[5] * v.y + m[9] * v.z,
               m[2] * v.x + m[6] * v.y + m[10] * v.z);
}
/* src/MathCommon.h:207-214 TransposeTransformVector */
static inline v3 transpose_transform_vector(const float *m, v3 v)
{
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z,
               m[4] * v.x + m[5] * v.y + m[6] * v.z,
               m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
/* src/MathCommon.h:216-234 TransformPosition with optional error bound */
static inline v3 transform_position(const float *m, v3 p, v3 *err)
{
    v3 r = mk3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12],
               m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
               m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]);
    if (err) {
        float xs = fabsf(m[0] * p.x) + fabsf(m[4] * p.y) + fabsf(m[8] * p.z) + fabsf(m[12]);
        float ys = fabsf(m[1] * p.x) + fabsf(m[5] * p.y) + fabsf(m[9] * p.z) + fabsf(m[13]);
        float zs = fabsf(m[2] * p.x) + fabsf(m[6] * p.y) + fabsf(m[10] * p.z) + fabsf(m[14]);
        float xe = fabsf(m[0]) * err->x + fabsf(m[4]) * err->y + fabsf(m[8]) * err->z;
        float ye = fabsf(m[1]) * err->x + fabsf(m[5]) * err->y + fabsf(m[9]) * err->z;
        float ze = fabsf(m[2]) * err->x + fabsf(m[6]) * err->y + fabsf(m[10]) * err->z;
        float g3 = error_gamma(3);
        *err = add3(sclf3(g3 + 1.0f, mk3(xe, ye, ze)), sclf3(g3, mk3(xs, ys, zs)));
    }
    return r;
}

/* src/MathCommon.h:113-129 */
static inline void orthogonal_basis(v3 *N, v3 *T, v3 *B)
{
    *N = normalize3(*N);
    if (fabsf(N->z) > fabsf(N->x)) {
        float l = sqrtf(fmaxr(1.f - N->x * N->x, 0.0f));
        float r = 1.f / l;
        *B = mk3(0.f, N->z * r, -N->y * r);
    } else {
        float l = sqrtf(fmaxr(1.f - N->z * N->z, 0.0f));
        float r = 1.f / l;
        *B = mk3(N->y * r, -N->x * r, 0.f);
    }
    *T = cross3(*B, *N);
}

#endif

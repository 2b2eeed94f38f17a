// dev_math.hpp -- float3 arithmetic shared by host and gfx950 device code.
//
// Every helper rounds per operation in source order (the library is built with
// -ffp-contract=off), matching the formulas the reference takes from glm 0.9.8
// (dot = (x*x + y*y) + z*z, normalize = v * (1 / sqrt(dot)), mix = x + a*(y - x),
// mat4*vec4 = (c0*x + c1*y) + (c2*z + c3*w)) and from src/MathCommon.h.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>
#include <type_traits>

#define YM_HD __host__ __device__ __forceinline__

// Calls f(integral_constant<int, I>) for I in [B, E).  Small fixed-size arrays
// whose elements are picked by a run-time index stay in registers only when
// every access uses a compile-time index: an unrolled loop is not enough, as
// the optimiser may first fold "if (i == k) a[i] = v" into a store through a
// computed address, which pins the array in scratch memory.
template <int B, int E, typename F>
YM_HD void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

namespace ym {

// Loads and stores through global-memory instructions.  The device arrays
// reach the pass kernels as pointers read from the frame-slot table in
// memory, so the compiler cannot tell them from LDS or scratch addresses and
// emits flat instructions, which count against both the vector-memory and
// the LDS wait counters: an LDS stack access must then also wait for every
// node load in flight.  Casting the address through an integer to the
// global address space keeps the global form (with a scalar base where the
// index is per lane); the arrays are device allocations, so it is exact.
template <typename T>
YM_HD T gload(const T *p)
{
    T out;
    __builtin_memcpy(&out, (const __attribute__((address_space(1))) T *)(uintptr_t)p, sizeof(T));
    return out;
}
template <typename T>
YM_HD void gstore(T *p, const T &v)
{
    __builtin_memcpy((__attribute__((address_space(1))) T *)(uintptr_t)p, &v, sizeof(T));
}

constexpr float kPi = 3.14159265358f;
constexpr float kTwoPi = 6.28318530716f;
constexpr float kInvPi = 0.318309886183f;
constexpr float kFloatEpsilon = 1.569209E-07f;
constexpr float kMinColorEpsilon = 3.7252903E-9f;
constexpr float kMaxColorClamp = 1E16f;
constexpr float kShadowRayClamp = 0.9995f;
constexpr float kFloatMax = 1E36f;
constexpr float kTMax = 1E36f;
constexpr float kTMin = 1E-7f;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kMaxBoundaryRecord = 6;
constexpr int kMaxRayDepth = 64;
constexpr int kMaxFreePathLength = 2048;

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};

YM_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
YM_HD f3 s3(float s) { return f3{s, s, s}; }
YM_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
YM_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
YM_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
YM_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
YM_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
YM_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
YM_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
YM_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
YM_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
YM_HD f3 cross(f3 a, f3 b)
{
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
YM_HD float length(f3 a) { return sqrtf(dot(a, a)); }
YM_HD f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
YM_HD f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
// glm::min / glm::max semantics
YM_HD float gmin(float a, float b) { return b < a ? b : a; }
YM_HD float gmax(float a, float b) { return a < b ? b : a; }
YM_HD f3 min3(f3 a, f3 b) { return f3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
YM_HD f3 max3(f3 a, f3 b) { return f3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
YM_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
YM_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
YM_HD void setc(f3 &a, int i, float v)
{
    if (i == 0) a.x = v;
    else if (i == 1) a.y = v;
    else a.z = v;
}
YM_HD float maxcomp(f3 a) { return gmax(a.x, gmax(a.y, a.z)); }
YM_HD f3 mix3(f3 x, f3 y, float a) { return x + a * (y - x); }
YM_HD f3 mixv(f3 x, f3 y, f3 a) { return x + a * (y - x); }
YM_HD float sqr(float x) { return x * x; }

// src/MathCommon.h:103-106
YM_HD float safe_rcp(float f) { return fabsf(f) < 1E-30f ? (f > 0.0f ? 1E30f : -1E30f) : (1.0f / f); }

// Transcendentals of float arguments, evaluated in double precision and
// rounded once: the correctly rounded float result except when the double
// result lies within one of its own ulps of a float rounding boundary.  The
// GPU's own float versions are up to a few ulps off, and the last bits of a
// free-flight distance, a transmittance or a sampled direction decide later
// branches; a CPU evaluation of the same form rounds to the same float.
YM_HD float ym_expf(float x) { return (float)exp((double)x); }
YM_HD float ym_logf(float x) { return (float)log((double)x); }
YM_HD float ym_sinf(float x) { return (float)sin((double)x); }
YM_HD float ym_cosf(float x) { return (float)cos((double)x); }
// sin and cos of an angle |x| <= 64 (the sampled azimuths, [0, 2 pi)) in
// double precision with few registers, each rounded once to float: Cody-Waite
// reduction by pi/2 with fdlibm's two-part pi/2 (pio2_1 has 33 significant
// bits, so k * pio2_1 is exact for |k| < 2^20), then the Taylor series of sin
// and cos on |r| <= pi/4 up to r^19 and r^18, whose truncation errors stay
// below 1e-19.  The general double sin/cos carry a reduction for huge
// arguments that costs the shading kernels spilled registers.
YM_HD void ym_sincosf(float x, float &s, float &c)
{
    constexpr double kPio2Hi = 1.57079632673412561417e+00;  // fdlibm pio2_1
    constexpr double kPio2Lo = 6.07710050650619224932e-11;  // fdlibm pio2_1t
    const double xd = (double)x;
    const double k = rint(xd * 0.63661977236758134308);      // 2 / pi
    const double r = (xd - k * kPio2Hi) - k * kPio2Lo;
    const double r2 = r * r;
    double ps = 1.0 / 121645100408832000.0;                  // 1/19!
    ps = ps * -r2 + 1.0 / 355687428096000.0;                 // 1/17!
    ps = ps * -r2 + 1.0 / 1307674368000.0;                   // 1/15!
    ps = ps * -r2 + 1.0 / 6227020800.0;                      // 1/13!
    ps = ps * -r2 + 1.0 / 39916800.0;                        // 1/11!
    ps = ps * -r2 + 1.0 / 362880.0;                          // 1/9!
    ps = ps * -r2 + 1.0 / 5040.0;                            // 1/7!
    ps = ps * -r2 + 1.0 / 120.0;                             // 1/5!
    ps = ps * -r2 + 1.0 / 6.0;                               // 1/3!
    const double sr = r - r * r2 * ps;
    double pc = 1.0 / 6402373705728000.0;                    // 1/18!
    pc = pc * -r2 + 1.0 / 20922789888000.0;                  // 1/16!
    pc = pc * -r2 + 1.0 / 87178291200.0;                     // 1/14!
    pc = pc * -r2 + 1.0 / 479001600.0;                       // 1/12!
    pc = pc * -r2 + 1.0 / 3628800.0;                         // 1/10!
    pc = pc * -r2 + 1.0 / 40320.0;                           // 1/8!
    pc = pc * -r2 + 1.0 / 720.0;                             // 1/6!
    pc = pc * -r2 + 1.0 / 24.0;                              // 1/4!
    pc = pc * -r2 + 0.5;                                     // 1/2!
    const double cr = 1.0 - r2 * pc;
    const int q = (int)k & 3;
    const double sd = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    const double cd = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
    s = (float)sd;
    c = (float)cd;
}
YM_HD float ym_powf(float x, float y) { return (float)pow((double)x, (double)y); }
YM_HD float ym_acosf(float x) { return (float)acos((double)x); }
YM_HD float ym_atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
// src/MathCommon.h:151-154
YM_HD constexpr float error_gamma(int n)
{
    return (float(n) * kFloatEpsilon) / (1.0f - float(n) * kFloatEpsilon);
}

YM_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
YM_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// src/MathCommon.h:263-297
YM_HD float next_float_up(float f)
{
    if (f == INFINITY) return f;
    if (f == -0.0f) f = 0.0f;
    uint32_t ui = f2u(f);
    if (f >= 0.0f) ++ui;
    else --ui;
    return u2f(ui);
}
YM_HD float next_float_down(float f)
{
    if (f == -INFINITY) return f;
    if (f == 0.0f) f = -0.0f;
    uint32_t ui = f2u(f);
    if (f <= 0.0f) ++ui;
    else --ui;
    return u2f(ui);
}

// src/MathCommon.h:146-149; the green weight is a double literal there.
YM_HD float luminance(f3 c)
{
    return (float)((double)(0.2126f * c.x) + 0.7152 * (double)c.y + (double)(0.0722f * c.z));
}

// Affine 3x4 transform stored as three rows (x' = r0 . (x,y,z,1)), evaluated in
// glm's mat4*vec4 order: (m0*x + m1*y) + (m2*z + m3*w).
struct Affine {
    float r[12];
};
YM_HD f3 xf_point(const Affine &m, f3 p)
{
    return f3{(m.r[0] * p.x + m.r[1] * p.y) + (m.r[2] * p.z + m.r[3] * 1.0f),
              (m.r[4] * p.x + m.r[5] * p.y) + (m.r[6] * p.z + m.r[7] * 1.0f),
              (m.r[8] * p.x + m.r[9] * p.y) + (m.r[10] * p.z + m.r[11] * 1.0f)};
}
YM_HD f3 xf_dir(const Affine &m, f3 d)
{
    return f3{(m.r[0] * d.x + m.r[1] * d.y) + (m.r[2] * d.z + m.r[3] * 0.0f),
              (m.r[4] * d.x + m.r[5] * d.y) + (m.r[6] * d.z + m.r[7] * 0.0f),
              (m.r[8] * d.x + m.r[9] * d.y) + (m.r[10] * d.z + m.r[11] * 0.0f)};
}
// glm::transpose(m) * vec4(v, 0); the bottom row (0,0,0,1) times w = 0 adds +0.
YM_HD f3 xf_transpose_dir(const Affine &m, f3 d)
{
    return f3{(m.r[0] * d.x + m.r[4] * d.y) + (m.r[8] * d.z + 0.0f),
              (m.r[1] * d.x + m.r[5] * d.y) + (m.r[9] * d.z + 0.0f),
              (m.r[2] * d.x + m.r[6] * d.y) + (m.r[10] * d.z + 0.0f)};
}
// src/MathCommon.h:198-214 TransformVector / TransposeTransformVector
YM_HD f3 xf_vector(const Affine &m, f3 v)
{
    return f3{m.r[0] * v.x + m.r[1] * v.y + m.r[2] * v.z, m.r[4] * v.x + m.r[5] * v.y + m.r[6] * v.z,
              m.r[8] * v.x + m.r[9] * v.y + m.r[10] * v.z};
}
YM_HD f3 xf_transpose_vector(const Affine &m, f3 v)
{
    return f3{m.r[0] * v.x + m.r[4] * v.y + m.r[8] * v.z, m.r[1] * v.x + m.r[5] * v.y + m.r[9] * v.z,
              m.r[2] * v.x + m.r[6] * v.y + m.r[10] * v.z};
}
// src/MathCommon.h:216-234 TransformPosition with the running error bound
YM_HD f3 xf_position(const Affine &m, f3 p, f3 *err)
{
    f3 r{m.r[0] * p.x + m.r[1] * p.y + m.r[2] * p.z + m.r[3], m.r[4] * p.x + m.r[5] * p.y + m.r[6] * p.z + m.r[7],
         m.r[8] * p.x + m.r[9] * p.y + m.r[10] * p.z + m.r[11]};
    if (err) {
        float xs = fabsf(m.r[0] * p.x) + fabsf(m.r[1] * p.y) + fabsf(m.r[2] * p.z) + fabsf(m.r[3]);
        float ys = fabsf(m.r[4] * p.x) + fabsf(m.r[5] * p.y) + fabsf(m.r[6] * p.z) + fabsf(m.r[7]);
        float zs = fabsf(m.r[8] * p.x) + fabsf(m.r[9] * p.y) + fabsf(m.r[10] * p.z) + fabsf(m.r[11]);
        float xe = fabsf(m.r[0]) * err->x + fabsf(m.r[1]) * err->y + fabsf(m.r[2]) * err->z;
        float ye = fabsf(m.r[4]) * err->x + fabsf(m.r[5]) * err->y + fabsf(m.r[6]) * err->z;
        float ze = fabsf(m.r[8]) * err->x + fabsf(m.r[9]) * err->y + fabsf(m.r[10]) * err->z;
        const float g3 = error_gamma(3);
        *err = (g3 + 1.0f) * f3{xe, ye, ze} + g3 * f3{xs, ys, zs};
    }
    return r;
}

// src/MathCommon.h:113-129
YM_HD void orthogonal_basis(f3 &N, f3 &T, f3 &B)
{
    N = normalize(N);
    if (fabsf(N.z) > fabsf(N.x)) {
        float l = sqrtf(gmax(1.f - N.x * N.x, 0.0f));
        float r = 1.f / l;
        B = f3{0.f, N.z * r, -N.y * r};
    } else {
        float l = sqrtf(gmax(1.f - N.z * N.z, 0.0f));
        float r = 1.f / l;
        B = f3{N.y * r, -N.x * r, 0.f};
    }
    T = cross(B, N);
}

// src/ShadingUtilities.cuh:993-1011
YM_HD f3 offset_ray_origin(f3 p, f3 perr, f3 dir, f3 ng)
{
    float dist = dot(abs3(ng), perr);
    f3 off = dist * ng;
    if (dot(dir, ng) < 0.0f) off = -off;
    f3 po = p + off;
    if (off.x > 0.0f) po.x = next_float_up(po.x);
    else if (off.x < 0.0f) po.x = next_float_down(po.x);
    if (off.y > 0.0f) po.y = next_float_up(po.y);
    else if (off.y < 0.0f) po.y = next_float_down(po.y);
    if (off.z > 0.0f) po.z = next_float_up(po.z);
    else if (off.z < 0.0f) po.z = next_float_down(po.z);
    return po;
}

}  // namespace ym

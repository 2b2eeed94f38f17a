// dev_sampler.hpp -- Halton and PCG samplers, bit-exact with the reference.
//
//   HaltonEnumerator (src/HaltonEnumerator.h:11-115): pixel -> global index
//   RadicalInversion / RadicalInversion2 / RadicalInversion3
//     (src/RandomSampler.cuh:456-507)
//   HaltonSampler::Random1D dimension walk (430-440); pcg4d / PCGSampler (360-404)
//
// Integer digit extraction is exact; the float part is
// fma(float(radical_sum), base_multiply, postfix_zero), the single fused
// operation the reference's nvcc build (-fmad=true) emits for
// `radical_sum * base_multiply + postfix_zero`.  Indices below 2^32 take an
// exact double-precision digit loop (same quotients and sums, far cheaper
// than 64-bit integer division on gfx950).
#pragma once

#include "dev_scene.hpp"

namespace ym {

struct HaltonEnum {
    float x_scale, y_scale;
    int m_x, m_y, exp2, exp3, increment;

    static void ext_euclid(int a, int b, int &x, int &y)
    {
        if (b == 0) {
            x = 1;
            y = 0;
            return;
        }
        int lx, ly;
        ext_euclid(b, a % b, lx, ly);
        x = ly;
        y = lx - (a / b) * ly;
    }
    static HaltonEnum make(int width, int height)
    {
        HaltonEnum e;
        int s2 = 1, s3 = 1;
        e.exp2 = 0;
        e.exp3 = 0;
        while (s2 < width) {
            s2 *= 2;
            ++e.exp2;
        }
        while (s3 < height) {
            s3 *= 3;
            ++e.exp3;
        }
        int mx, my;
        ext_euclid(s3, s2, mx, my);
        int inv2 = mx < 0 ? (mx + s2) : (mx % s2);
        int inv3 = my < 0 ? (my + s3) : (my % s3);
        e.m_x = s3 * inv2;
        e.m_y = s2 * inv3;
        e.increment = s2 * s3;
        e.x_scale = (float)s2;
        e.y_scale = (float)s3;
        return e;
    }
    YM_HD uint64_t index(uint32_t px, uint32_t py, uint32_t sample) const
    {
        uint64_t ix = 0, iy = 0;
        uint32_t a = px, b = py;
        for (int i = 0; i < exp2; ++i) {
            ix = ix * 2u + (a & 1u);
            a >>= 1;
        }
        for (int i = 0; i < exp3; ++i) {
            uint32_t q = b / 3u;
            iy = iy * 3u + (b - 3u * q);
            b = q;
        }
        const uint64_t off = (ix * (uint64_t)m_x + iy * (uint64_t)m_y) % (uint64_t)increment;
        return off + (uint64_t)sample * (uint64_t)increment;
    }
};

template <uint32_t kBase>
YM_HD float radical_inverse_plain(uint64_t index)
{
    constexpr float inv_base = 1.0f / float(kBase);
    uint64_t radical_sum = 0;
    float base_multiply = 1.0f;
    while (index) {
        uint64_t next = index / kBase;
        uint64_t digit = index - next * kBase;
        radical_sum = radical_sum * kBase + digit;
        base_multiply *= inv_base;
        index = next;
    }
    return gmax(0.000001f, gmin((float)radical_sum * base_multiply, 0.999999f));
}

// Permuted radical inverse of dimension dim >= 2 (base = dim-th prime),
// HaltonSampler::RadicalInversion (src/RandomSampler.cuh:442-472).
//
// Indices below 2^32 run the digit loop in double precision, where every
// value is an integer below 2^53 and so exact: the quotient q = n / base is
// floor(fma(n, 1/base, 2^-20)) -- with n < 2^32 and 5 <= base < 2^13 the
// fma lands within 2^-21 of n / base + 2^-20, which lies in (q, q + 1)
// because the fractional part of n / base is 0 or between 1/base and
// 1 - 1/base -- and the sum radical_sum * base + perm[digit] stays below
// base * 2^32 < 2^45.  Converting the exact sum to float rounds it as the
// reference's integer conversion does, so the value is the same; the
// double operations replace a 64-bit multiply-high and two 64-bit
// multiply-adds per digit.
YM_HD float radical_inverse(const DevScene &sc, uint64_t index, uint32_t dim)
{
    const HaltonDim hd = sc.halton_dims[dim];
    const uint32_t base = hd.base;
    const uint32_t *perm = sc.halton + hd.perm_off;
    const float inv_base = 1.0f / (float)base;
    float base_multiply = 1.0f;
    float sum;
    if (index <= 0xFFFFFFFFull) {
        const double base_d = (double)base;
        double radical_sum = 0.0;
        uint32_t idx = (uint32_t)index;
        while (idx) {
            const uint32_t next = (uint32_t)__builtin_fma((double)idx, hd.inv_base, 0x1p-20);
            const uint32_t digit = idx - next * base;
            radical_sum = __builtin_fma(radical_sum, base_d, (double)perm[digit]);
            base_multiply *= inv_base;
            idx = next;
        }
        sum = (float)radical_sum;
    } else {
        uint64_t radical_sum = 0;
        while (index) {
            const uint64_t next = index / base;
            const uint64_t digit = index - next * base;
            radical_sum = radical_sum * base + perm[digit];
            base_multiply *= inv_base;
            index = next;
        }
        sum = (float)radical_sum;
    }
    const float postfix_zero = (float)perm[0] * base_multiply * inv_base / (1.0f - inv_base);
    return gmax(0.000001f, gmin(__builtin_fmaf(sum, base_multiply, postfix_zero), 0.999999f));
}

YM_HD void pcg4d(uint32_t v[4])
{
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}

// Path sampler state: 16 B in HBM between stages.
struct Sampler {
    uint64_t index;    // Halton global index, or (pixel | sample << 32) for PCG
    uint32_t dim;
    uint32_t frame_type;  // frame (PCG) << 1 | type
};

constexpr uint32_t kSamplerHalton = 1;

__device__ __forceinline__ float sampler_next(const DevScene &sc, Sampler &s)
{
    if (s.frame_type & 1u) {
        float r = radical_inverse(sc, s.index, s.dim);
        ++s.dim;
        if (s.dim >= 1000u) s.dim = 2;
        return r;
    }
    uint32_t v[4] = {(uint32_t)s.index, s.dim, (uint32_t)(s.index >> 32), s.frame_type >> 1};
    pcg4d(v);
    s.dim += 1;
    return (float)v[0] / (float)0xFFFFFFFFu;
}

}  // namespace ym

/*
 * or_sampler.c -- oracle restatement of YMRT's samplers.
 *
 * TEST INFRASTRUCTURE ONLY (checker for the HIP sampler; never shipped).
 *
 *   - first 1000 primes and their prefix sums: src/RandomSampler.cuh:14-212
 *     (generated here by trial division; the tests check count, last prime 7919
 *     and the total 3 682 913)
 *   - HaltonEnumerator: src/HaltonEnumerator.h:11-115
 *   - HaltonSampler::RadicalInversion / RadicalInversion2 / RadicalInversion3:
 *     src/RandomSampler.cuh:456-507
 *   - HaltonSampler::Random1D dimension walk and wrap at 1000:
 *     src/RandomSampler.cuh:430-440
 *   - pcg4d / PCGSampler::Random1D: src/RandomSampler.cuh:360-404
 *   - permutation table layout: src/core/SceneModule.cpp:1412-1432.  The
 *     reference seeds std::mt19937 from std::random_device and uses
 *     std::shuffle (implementation-defined).  Here the table is fixed by a seed:
 *     MT19937 (the standard generator; known answer: the 10000th output of the
 *     default seed 5489 is 4123659995) and a Fisher-Yates pass from the top,
 *     j = rng() % (i + 1).
 *
 * Floating point: RadicalInversion's final `radical_sum * base_multiply +
 * postfix_zero` is evaluated as one fused multiply-add, which is what the
 * reference's nvcc build (default -fmad=true) emits for that expression; the
 * pixel offset `ScaleY(y) - py` likewise.  Everything else rounds per
 * operation (-ffp-contract=off).
 */
#include <stdlib.h>
#include <math.h>
#include "oracle.h"

static uint32_t g_primes[OR_PRIME_COUNT];
static uint32_t g_prime_sums[OR_PRIME_COUNT + 1];
static int g_primes_ready = 0;

static void init_primes(void)
{
    if (g_primes_ready) return;
    uint32_t n = 0, c = 2;
    while (n < OR_PRIME_COUNT) {
        int is_p = 1;
        for (uint32_t d = 2; d * d <= c; ++d) {
            if (c % d == 0) { is_p = 0; break; }
        }
        if (is_p) g_primes[n++] = c;
        ++c;
    }
    g_prime_sums[0] = 0;
    for (uint32_t i = 0; i < OR_PRIME_COUNT; ++i) g_prime_sums[i + 1] = g_prime_sums[i] + g_primes[i];
    g_primes_ready = 1;
}

uint32_t or_prime(uint32_t i) { init_primes(); return g_primes[i]; }
uint32_t or_prime_sum(uint32_t i) { init_primes(); return g_prime_sums[i]; }

/* ------------------------------------------------------------------ MT19937 */
void or_mt_seed(or_mt19937 *mt, uint32_t seed)
{
    mt->s[0] = seed;
    for (int i = 1; i < 624; ++i)
        mt->s[i] = 1812433253u * (mt->s[i - 1] ^ (mt->s[i - 1] >> 30)) + (uint32_t)i;
    mt->i = 624;
}

uint32_t or_mt_next(or_mt19937 *mt)
{
    if (mt->i >= 624) {
        for (int k = 0; k < 624; ++k) {
            uint32_t y = (mt->s[k] & 0x80000000u) | (mt->s[(k + 1) % 624] & 0x7fffffffu);
            uint32_t v = mt->s[(k + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            mt->s[k] = v;
        }
        mt->i = 0;
    }
    uint32_t y = mt->s[mt->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

uint32_t or_halton_table_size(void) { init_primes(); return g_prime_sums[OR_PRIME_COUNT]; }

void or_halton_table(uint32_t seed, uint32_t *table)
{
    init_primes();
    or_mt19937 mt;
    or_mt_seed(&mt, seed);
    for (uint32_t i = 0; i < OR_PRIME_COUNT; ++i) {
        uint32_t p = g_primes[i];
        uint32_t *perm = table + g_prime_sums[i];
        for (uint32_t j = 0; j < p; ++j) perm[j] = j;
        for (uint32_t j = p - 1; j > 0; --j) {
            uint32_t k = or_mt_next(&mt) % (j + 1);
            uint32_t t = perm[j]; perm[j] = perm[k]; perm[k] = t;
        }
    }
}

/* --------------------------------------------------------------- enumerator */
static void ext_euclid(int a, int b, int *x, int *y)
{
    if (b == 0) { *x = 1; *y = 0; return; }
    int lx, ly;
    ext_euclid(b, a % b, &lx, &ly);
    *x = ly;
    *y = lx - (a / b) * ly;
}

void or_enum_init(or_enumerator *e, int width, int height)
{
    int scale2 = 1, scale3 = 1;
    e->exp2 = 0; e->exp3 = 0;
    while (scale2 < width) { scale2 *= 2; ++e->exp2; }
    while (scale3 < height) { scale3 *= 3; ++e->exp3; }
    int w = scale2, h = scale3;
    int mx, my;
    ext_euclid(h, w, &mx, &my);
    int inv2 = mx < 0 ? (mx + w) : (mx % w);
    int inv3 = my < 0 ? (my + h) : (my % h);
    e->m_x = h * inv2;
    e->m_y = w * inv3;
    e->increment = w * h;
    e->x_scale = (float)scale2;
    e->y_scale = (float)scale3;
}

static uint64_t digit_reverse(uint64_t index, uint64_t base, uint64_t digits)
{
    uint64_t r = 0;
    for (uint64_t i = 0; i < digits; ++i) {
        r = r * base + (index % base);
        index /= base;
    }
    return r;
}

uint64_t or_enum_index(const or_enumerator *e, int px, int py, int sample)
{
    uint64_t ix = digit_reverse((uint64_t)px, 2, (uint64_t)e->exp2);
    uint64_t iy = digit_reverse((uint64_t)py, 3, (uint64_t)e->exp3);
    uint64_t off = (ix * (uint64_t)e->m_x + iy * (uint64_t)e->m_y) % (uint64_t)e->increment;
    return off + (uint64_t)sample * (uint64_t)e->increment;
}

/* ---------------------------------------------------------- radical inverse */
float or_radical_inverse(uint64_t index, uint32_t dim, const uint32_t *table)
{
    init_primes();
    uint32_t base = g_primes[dim];
    const uint32_t *perm = table + g_prime_sums[dim];
    float inv_base = 1.0f / (float)base;
    uint64_t radical_sum = 0;
    float base_multiply = 1.0f;
    while (index) {
        uint64_t next = index / base;
        uint64_t digit = index - next * base;
        radical_sum = radical_sum * base + perm[digit];
        base_multiply *= inv_base;
        index = next;
    }
    float postfix_zero = (float)perm[0] * base_multiply * inv_base / (1.0f - inv_base);
    float v = fmaf((float)radical_sum, base_multiply, postfix_zero);
    return fmaxr(0.000001f, fminr(v, 0.999999f));
}

static float radical_inverse_plain(uint64_t index, uint32_t base)
{
    const float inv_base = 1.0f / (float)base;
    uint64_t radical_sum = 0;
    float base_multiply = 1.0f;
    while (index) {
        uint64_t next = index / base;
        uint64_t digit = index - next * base;
        radical_sum = radical_sum * base + digit;
        base_multiply *= inv_base;
        index = next;
    }
    return fmaxr(0.000001f, fminr((float)radical_sum * base_multiply, 0.999999f));
}

float or_radical_inverse2(uint64_t index) { return radical_inverse_plain(index, 2u); }
float or_radical_inverse3(uint64_t index) { return radical_inverse_plain(index, 3u); }

/* Value of dimension `dim` of Halton sample `index` as the renderer sees it:
 * dims 0 and 1 are the unpermuted pixel-offset values, dims >= 2 the permuted
 * radical inverses. */
float or_halton_value(uint64_t index, uint32_t dim, const uint32_t *table)
{
    if (dim == 0) return or_radical_inverse2(index);
    if (dim == 1) return or_radical_inverse3(index);
    return or_radical_inverse(index, dim, table);
}

float g_probe_u = 0.5f;

float or_sampler_next(or_sampler *s)
{
    if (s->type < 0) return g_probe_u; /* fixed value, used only by test probes */
    if (s->type == OR_SAMPLER_HALTON) {
        float r = or_radical_inverse(s->index, s->dim, s->table);
        ++s->dim;
        if (s->dim >= OR_PRIME_COUNT) s->dim = 2;
        return r;
    }
    /* PCGSampler::Random1D: pcg4d(pixel, dim, sample, frame).x / 0xFFFFFFFF */
    uint32_t v[4] = {s->pixel, s->dim, s->sample, s->frame};
    or_pcg4d(v);
    s->dim += 1;
    return (float)v[0] / (float)0xFFFFFFFFu;
}

void or_pcg4d(uint32_t v[4])
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

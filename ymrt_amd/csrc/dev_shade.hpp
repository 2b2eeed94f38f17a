// dev_shade.hpp -- device shading library: medium records, image textures,
// the DefaultMtl BSDF, shape lights, Henyey-Greenstein phases and free-flight
// sampling.  Each function restates the reference function cited beside it in
// the same floating-point operation order (contraction off), so the HIP path
// and the CPU oracle compute the same floats wherever the math library does.
#pragma once

#include "../../include/ymrt.h"
#include "dev_sampler.hpp"

namespace ym {

// ---------------------------------------------------------------- RayTransfer
// src/Ray.h:106-260 (records of (instance, ior, priority, volume)).  Every
// access to the six fixed slots uses a compile-time index (static_for), is
// unconditional, and picks by selecting values: conditional loads or stores
// at different slots get merged by the optimiser into one access through a
// computed address, which would pin the records in scratch memory.
YM_HD CameraRecord select_rec(bool c, const CameraRecord &a, const CameraRecord &b)
{
    return CameraRecord{c ? a.prim : b.prim, c ? a.ior : b.ior, c ? a.priority : b.priority, c ? a.vol : b.vol};
}

struct Transfer {
    int max_idx, count;
    CameraRecord rec[kMaxBoundaryRecord];

    YM_HD void clear()
    {
        max_idx = -1;
        count = 0;
    }
    YM_HD CameraRecord get(int idx) const
    {
        CameraRecord r = rec[0];
        static_for<1, kMaxBoundaryRecord>([&](auto I) {
            const CameraRecord c = rec[I()];
            r = select_rec(I() == idx, c, r);
        });
        return r;
    }
    YM_HD int find_max() const
    {
        uint32_t mp = rec[0].priority;
        int mi = 0;
        static_for<1, kMaxBoundaryRecord>([&](auto I) {
            const uint32_t pr = rec[I()].priority;
            const bool take = I() < count && pr >= mp;
            mp = take ? pr : mp;
            mi = take ? (int)I() : mi;
        });
        return count == 0 ? -1 : mi;
    }
    // RayTransfer::FindRecordWithPrim scans down from the top: the highest match wins
    YM_HD int find_prim(uint32_t prim) const
    {
        int r = -1;
        static_for<0, kMaxBoundaryRecord>([&](auto I) {
            const uint32_t pp = rec[I()].prim;
            r = (I() < count && pp == prim) ? (int)I() : r;
        });
        return r;
    }
    YM_HD bool push(uint32_t prim, float ior, uint32_t prio, int vol)
    {
        if (count == kMaxBoundaryRecord) return false;
        const CameraRecord nr{prim, ior, prio, vol};
        static_for<0, kMaxBoundaryRecord>([&](auto I) {
            const CameraRecord c = rec[I()];
            rec[I()] = select_rec(I() == count, nr, c);
        });
        ++count;
        if (count == 1) {
            max_idx = 0;
            return true;
        }
        if (prio >= get(max_idx).priority) max_idx = count - 1;
        return true;
    }
    YM_HD bool pop(int idx)
    {
        if (count == 0) return false;
        if (!(idx >= 0 && idx < count)) return false;
        static_for<0, kMaxBoundaryRecord - 1>([&](auto I) {
            const CameraRecord c = rec[I()], nx = rec[I() + 1];
            rec[I()] = select_rec(I() >= idx && I() + 1 < count, nx, c);
        });
        --count;
        max_idx = find_max();
        return true;
    }
    // GetCurrentVolumeIndices: records of the top priority, from the top down
    YM_HD int volumes(int *vols, uint32_t volume_count) const
    {
        static_for<0, kMaxBoundaryRecord>([&](auto J) { vols[J()] = 0; });
        if (count == 0 || max_idx == -1) return 0;
        const uint32_t mp = get(max_idx).priority;
        int n = 0;
        static_for<0, kMaxBoundaryRecord>([&](auto I) {
            constexpr int i = kMaxBoundaryRecord - 1 - I();
            const CameraRecord c = rec[i];
            const bool take = i < count && c.priority == mp && c.vol >= 0 && (uint32_t)c.vol < volume_count;
            static_for<0, kMaxBoundaryRecord>([&](auto J) {
                const int v = vols[J()];
                vols[J()] = (take && J() == n) ? c.vol : v;
            });
            n += take ? 1 : 0;
        });
        return n;
    }
    YM_HD float ray_ior() const { return count > 0 ? get(max_idx).ior : 1.0f; }
    YM_HD float ex_ior(uint32_t hit_prim) const
    {
        uint32_t mp = 0;
        int mi = -1;
        static_for<0, kMaxBoundaryRecord>([&](auto I) {
            const CameraRecord c = rec[I()];
            const bool take = I() < count && c.prim != hit_prim && c.priority >= mp;
            mp = take ? c.priority : mp;
            mi = take ? (int)I() : mi;
        });
        return mi == -1 ? 1.0f : get(mi).ior;
    }
};

YM_HD float fetch_ior(const DevMaterial &m, uint32_t *prio)
{
    if (m.type == YMRT_MATERIAL_DEFAULT) {
        *prio = m.ior_priority;
        return m.ior_n;
    }
    *prio = 0;
    return 1.0f;
}

constexpr int kNearbyMax = 4;

struct NearbyList {
    int n;
    uint32_t prim[kNearbyMax];
    float t[kNearbyMax];
    int back[kNearbyMax];
};

// BoundaryTransitionBunch (src/ShadingUtilities.cuh:954-990): pushes for the
// front-face hits first, then pops for the back-face hits.
__device__ __forceinline__ void boundary_bunch(const DevScene &sc, Transfer &t, const NearbyList &nb)
{
#pragma unroll
    for (int i = 0; i < kNearbyMax; ++i) {
        if (i >= nb.n || nb.back[i]) continue;
        const DevInstance &in = sc.instances[nb.prim[i]];
        uint32_t prio;
        float ior = fetch_ior(sc.materials[in.material], &prio);
        if (t.find_prim(nb.prim[i]) == -1) t.push(nb.prim[i], ior, prio, in.inner_volume);
    }
#pragma unroll
    for (int i = 0; i < kNearbyMax; ++i) {
        if (i >= nb.n || !nb.back[i]) continue;
        int ri = t.find_prim(nb.prim[i]);
        if (ri != -1) t.pop(ri);
    }
}

// ---------------------------------------------------------------- textures
// EvalTexImage (src/TextureEval.cuh:162-237) over RGBA8 texels decoded through
// the sRGB / linear table (LoadPng8, src/FileIO.cpp:17-45).
// ((x % w) + w) % w, the repeat wrap, for w >= 1: one unsigned remainder
// for x >= 0, none for x = -1 (the left or upper neighbour of texel 0, the
// only negative coordinate a lookup makes), the signed formula otherwise.
__device__ __forceinline__ int wrap_repeat(int x, int w)
{
    if (x >= 0) return (int)((uint32_t)x % (uint32_t)w);
    if (x == -1) return w - 1;
    return ((x % w) + w) % w;
}

// Texel columns (or rows) c and c + 1 after the texture's wrap: clamped to
// [0, w - 1], or the repeat wrap of c and its successor ((c + 1) mod w
// follows from c mod w without a second remainder).
__device__ __forceinline__ void wrap_pair(const DevTexture &t, int c, int w, int &c0, int &c1)
{
    if (t.wrap == YMRT_WRAP_CLAMP) {
        c0 = c < 0 ? 0 : (c > w - 1 ? w - 1 : c);
        c1 = c + 1 < 0 ? 0 : (c + 1 > w - 1 ? w - 1 : c + 1);
    } else {
        c0 = wrap_repeat(c, w);
        c1 = c0 + 1 == w ? 0 : c0 + 1;
    }
}

// The texel at wrapped coordinates (x, y).
__device__ __forceinline__ f3 texel(const DevScene &sc, const DevTexture &t, int x, int y)
{
    const int w = (int)t.width;
    if (t.f32) {
        const float *p = sc.texels_f32 + t.offset + ((uint64_t)y * w + x) * t.channels;
        if (t.channels == 1) return s3(p[0]);
        return f3{p[0], t.channels > 1 ? p[1] : 0.0f, t.channels > 2 ? p[2] : 0.0f};
    }
    const uint32_t word = sc.texels[t.offset + (uint64_t)y * w + x];
    const float *lut = sc.tex_lut + (t.color_space == YMRT_COLOR_SRGB ? 0 : 256);
    if (t.channels == 1) return s3(lut[word & 0xFFu]);
    f3 r = s3(0.0f);
    r.x = lut[word & 0xFFu];
    if (t.channels > 1) r.y = lut[(word >> 8) & 0xFFu];
    if (t.channels > 2) r.z = lut[(word >> 16) & 0xFFu];
    return r;
}

__device__ __forceinline__ f3 image_eval(const DevScene &sc, const DevTexture &t, f2 uv)
{
    float tu = uv.x * t.su + t.ou;
    if (tu < 0.0f) tu = gmax(0.0f, tu + fabsf(floorf(tu)));
    float tv = uv.y * t.sv + t.ov;
    if (tv < 0.0f) tv = gmax(0.0f, tv + fabsf(floorf(tv)));
    const float cx = tu * (float)t.width - 0.5f, cy = tv * (float)t.height - 0.5f;
    const int x0 = (int)floorf(cx), y0 = (int)floorf(cy);
    const float fx = cx - (float)x0, fy = cy - (float)y0;
    int xa, xb, ya, yb;
    wrap_pair(t, x0, (int)t.width, xa, xb);
    wrap_pair(t, y0, (int)t.height, ya, yb);
    f3 r = ((1.0f - fx) * (1.0f - fy)) * texel(sc, t, xa, ya);
    r = r + (fx * (1.0f - fy)) * texel(sc, t, xb, ya);
    r = r + ((1.0f - fx) * fy) * texel(sc, t, xa, yb);
    r = r + (fx * fy) * texel(sc, t, xb, yb);
    return r;
}

// The four bilinear taps image_eval blends at uv: the offset of each tap's
// first channel in the texture's own texel order (row-major, channels
// interleaved) and its weight.
__device__ __forceinline__ void image_taps(const DevTexture &t, f2 uv, uint32_t (&off)[4], float (&w)[4])
{
    float tu = uv.x * t.su + t.ou;
    if (tu < 0.0f) tu = gmax(0.0f, tu + fabsf(floorf(tu)));
    float tv = uv.y * t.sv + t.ov;
    if (tv < 0.0f) tv = gmax(0.0f, tv + fabsf(floorf(tv)));
    const float cx = tu * (float)t.width - 0.5f, cy = tv * (float)t.height - 0.5f;
    const int x0 = (int)floorf(cx), y0 = (int)floorf(cy);
    const float fx = cx - (float)x0, fy = cy - (float)y0;
    w[0] = (1.0f - fx) * (1.0f - fy);
    w[1] = fx * (1.0f - fy);
    w[2] = (1.0f - fx) * fy;
    w[3] = fx * fy;
    int xs[2], ys[2];
    wrap_pair(t, x0, (int)t.width, xs[0], xs[1]);
    wrap_pair(t, y0, (int)t.height, ys[0], ys[1]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        off[k] = ((uint32_t)ys[k >> 1] * t.width + (uint32_t)xs[k & 1]) * t.channels;
}

// Perlin noise (PerlinNoiseEval, src/TextureEval.cuh:55-131): Ken Perlin's
// improved noise over the reference's lattice permutation and its 16
// gradient cases.
__constant__ uint8_t kPerlinPerm[256] = {
#include "perlin_perm.inc"
};

__device__ __forceinline__ int perlin_hash(int x, int y, int z)
{
    return kPerlinPerm[(kPerlinPerm[(kPerlinPerm[x] + y) & 255] + z) & 255];
}

__device__ __forceinline__ float perlin_grad(int h, float x, float y, float z)
{
    switch (h & 15) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    case 11: return -y - z;
    case 12: return y + x;
    case 13: return -x + y;
    case 14: return -y + z;
    default: return -y - z;
    }
}

__device__ __forceinline__ float perlin_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

__device__ __noinline__ float perlin_noise(f3 p)
{
    int x0 = (int)floorf(p.x), y0 = (int)floorf(p.y), z0 = (int)floorf(p.z);
    const float tx = p.x - (float)x0, ty = p.y - (float)y0, tz = p.z - (float)z0;
    const int x1 = (x0 + 1) & 255, y1 = (y0 + 1) & 255, z1 = (z0 + 1) & 255;
    x0 &= 255;
    y0 &= 255;
    z0 &= 255;
    const float v000 = perlin_grad(perlin_hash(x0, y0, z0), tx, ty, tz);
    const float v100 = perlin_grad(perlin_hash(x1, y0, z0), tx - 1, ty, tz);
    const float v010 = perlin_grad(perlin_hash(x0, y1, z0), tx, ty - 1, tz);
    const float v001 = perlin_grad(perlin_hash(x0, y0, z1), tx, ty, tz - 1);
    const float v110 = perlin_grad(perlin_hash(x1, y1, z0), tx - 1, ty - 1, tz);
    const float v011 = perlin_grad(perlin_hash(x0, y1, z1), tx, ty - 1, tz - 1);
    const float v101 = perlin_grad(perlin_hash(x1, y0, z1), tx - 1, ty, tz - 1);
    const float v111 = perlin_grad(perlin_hash(x1, y1, z1), tx - 1, ty - 1, tz - 1);
    const float ix = perlin_fade(tx), iy = perlin_fade(ty), iz = perlin_fade(tz);
    const float v00 = v000 + ix * (v100 - v000);
    const float v01 = v001 + ix * (v101 - v001);
    const float v10 = v010 + ix * (v110 - v010);
    const float v11 = v011 + ix * (v111 - v011);
    const float v0 = v00 + iy * (v10 - v00);
    const float v1 = v01 + iy * (v11 - v01);
    return v0 + iz * (v1 - v0);
}

// PerlinNoiseFBM / PerlinNoiseTurbulence (src/TextureEval.cuh:132-160)
__device__ __forceinline__ float perlin_layers(f3 p, float lacunarity, float gain, int layers, float amplitude,
                                               float offset, bool turbulence)
{
    float r = 0.0f, frequency = 1.0f;
    for (int i = 0; i < layers; ++i) {
        const float n = perlin_noise(frequency * p);
        r += amplitude * (turbulence ? fabsf(n) : n) + offset;
        amplitude *= gain;
        offset *= gain;
        frequency *= lacunarity;
    }
    return r;
}

struct TexCoord {
    f2 st;
    f3 p_world, p_object;
};

// One node of a texture graph given its inputs (EvalTexResult and the EvalTex*
// functions, src/TextureEval.cuh:162-404).
__device__ __forceinline__ f3 texture_node(const DevScene &sc, const DevTexture &t, f3 black, f3 white,
                                           const TexCoord &tc)
{
    const f3 po = tc.p_object;
    switch (t.type) {
    case YMRT_TEXTURE_IMAGE: return image_eval(sc, t, tc.st);
    case YMRT_TEXTURE_CONSTANT_FLOAT: return s3(t.value[0]);
    case YMRT_TEXTURE_CONSTANT_RGB: return f3{t.value[0], t.value[1], t.value[2]};
    case YMRT_TEXTURE_CHECKERBOARD: {
        const float x = ym_sinf(po.x * t.frequency), y = ym_sinf(po.y * t.frequency), z = ym_sinf(po.z * t.frequency);
        return x * y * z < 0.0f ? white : black;
    }
    case YMRT_TEXTURE_NOISE: {
        float n = perlin_noise(po * t.frequency);
        if (t.normalized) n = n * 0.5f + 0.5f;
        return black + n * (white - black);
    }
    case YMRT_TEXTURE_FBM:
    case YMRT_TEXTURE_TURBULENCE: {
        const float r = perlin_layers(t.frequency * po, t.lacunarity, t.gain, t.layers, t.amplitude, t.offset_f,
                                      t.type == YMRT_TEXTURE_TURBULENCE);
        return black + r * (white - black);
    }
    case YMRT_TEXTURE_MARBLE: {
        const float fbm = perlin_layers(t.frequency * po, t.lacunarity, t.gain, t.layers, 1.0f, 0.0f, false);
        float turb = 0.0f;
        turb += (t.axes & 1) ? po.x : 0.0f;
        turb += (t.axes & 2) ? po.y : 0.0f;
        turb += (t.axes & 4) ? po.z : 0.0f;
        const float m = ym_sinf(turb + t.variation * fbm) * 0.5f + 0.5f;
        return black + m * (white - black);
    }
    case YMRT_TEXTURE_WOOD: {
        float w = perlin_noise(po * t.frequency) * t.variation;
        w = w - floorf(w);
        return black + w * (white - black);
    }
    case YMRT_TEXTURE_POLKA_DOT: {
        const f2 st{tc.st.x * t.frequency, tc.st.y * t.frequency};
        const float fs = floorf(st.x), ft = floorf(st.y);
        if (!(perlin_noise(f3{fs + 0.5f, ft + 0.5f, 0.0f}) > 0.0f)) return black;
        auto hash = [](float f, float multi) {
            const float x = f * multi;
            return (x - floorf(x)) * 114.f;
        };
        const float max_offset = 0.5f - t.radius;
        const float ds = st.x - fs, dt = st.y - ft;
        const float sc0 = 0.5f + max_offset * perlin_noise(f3{hash(st.x, 1145.14f), hash(st.y, 9807.26f), 0.0f});
        const float tc0 = 0.5f + max_offset * perlin_noise(f3{hash(st.x, 4444.44f), hash(st.y, 2333.33f), 0.0f});
        return sqr(ds - sc0) + sqr(dt - tc0) < t.radius * t.radius ? white : black;
    }
    default: {  // YMRT_TEXTURE_WAVE
        const float f0 = perlin_layers(po * t.frequency, t.lacunarity, t.gain, t.layers, 1.0f, 0.0f, false);
        const float f1 = perlin_layers(po * t.frequency1, t.lacunarity1, t.gain1, t.layers1, 1.0f, 0.0f, false);
        const float w = fabsf(f0) * f1;
        return black + w * (white - black);
    }
    }
}

// Texture graph evaluation (TextureEval, src/TextureEval.cuh:415-481): the
// graph's post-order program runs on a value stack of kTexStack entries kept
// per lane in LDS ([entry * 3 + channel][lane]); a one-node graph (an image
// or a constant) is evaluated inline without the stack.
template <bool kGraph = true>
struct TexEval {
    const DevScene &sc;
    TexCoord tc;
    float *stk;
    int lane;

    __device__ __forceinline__ void put(int i, f3 v)
    {
        stk[(i * 3 + 0) * 64 + lane] = v.x;
        stk[(i * 3 + 1) * 64 + lane] = v.y;
        stk[(i * 3 + 2) * 64 + lane] = v.z;
    }
    __device__ __forceinline__ f3 get(int i) const
    {
        return f3{stk[(i * 3 + 0) * 64 + lane], stk[(i * 3 + 1) * 64 + lane], stk[(i * 3 + 2) * 64 + lane]};
    }
    __device__ __forceinline__ f3 eval(uint32_t ti, const TexCoord &at)
    {
        const DevTexture &root = sc.textures[ti];
        if (root.type == YMRT_TEXTURE_IMAGE) return image_eval(sc, root, at.st);
        if (root.type == YMRT_TEXTURE_CONSTANT_FLOAT) return s3(root.value[0]);
        if (root.type == YMRT_TEXTURE_CONSTANT_RGB) return f3{root.value[0], root.value[1], root.value[2]};
        if constexpr (kGraph) return graph(root, at);
        return s3(0.0f);  // kGraph is false only for scenes without procedural nodes
    }
    __device__ __noinline__ f3 graph(const DevTexture &root, const TexCoord &at)
    {
        int sp = 0;
        for (uint32_t k = 0; k < root.prog_len; ++k) {
            const uint32_t code = sc.tex_prog[root.prog_start + k];
            f3 v = s3(0.0f);
            if (code != kTexBlack) {
                const DevTexture &t = sc.textures[code];
                f3 black = s3(0.0f), white = s3(0.0f);
                if (t.type >= YMRT_TEXTURE_CHECKERBOARD) {
                    sp -= 2;
                    black = get(sp);
                    white = get(sp + 1);
                }
                v = texture_node(sc, t, black, white, at);
            }
            put(sp++, v);
        }
        return get(sp - 1);
    }
    __device__ __forceinline__ f3 eval(uint32_t ti) { return eval(ti, tc); }
};

// ---------------------------------------------------------------- BSDF lobes
// src/BSDF.cuh:20-375
YM_HD bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }
YM_HD f3 reflect3(f3 wo, f3 n) { return -wo + (2.0f * n) * dot(n, wo); }
YM_HD bool refract3(f3 wo, f3 n, float ni, float nt, f3 &wi)
{
    float ci = gmax(dot(wo, n), 0.0f);
    float ci2 = sqr(ci);
    float si2 = 1.0f - ci2;
    float ct2 = 1.0f - sqr(ni / nt) * si2;
    if (!(ct2 > 0.0f)) return false;
    float ct = sqrtf(ct2);
    wi = normalize(-(ni / nt) * (wo - n * ci) - n * ct);
    return true;
}
YM_HD float fresnel_dd(float eta, float c)
{
    c = gmax(0.0f, c);
    float s2 = 1.f - c * c;
    float t0 = sqrtf(gmax(1.0f - (s2 / (eta * eta)), 0.0f));
    float t1 = eta * t0;
    float t2 = eta * c;
    float rs = (c - t1) / gmax(c + t1, 1E-12f);
    float rp = (t0 - t2) / gmax(t0 + t2, 1E-12f);
    return (float)(0.5 * (double)(rs * rs + rp * rp));
}
YM_HD f3 sqrt3(f3 a) { return f3{sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
YM_HD f3 fresnel_dc(f3 eta, f3 etak, float c)
{
    float c2 = c * c, s2 = 1.0f - c2;
    f3 e2 = eta * eta, k2 = etak * etak;
    f3 t0 = e2 - k2 - s3(s2);
    f3 ab = f3{sqrtf(t0.x * t0.x + 4.0f * e2.x * k2.x), sqrtf(t0.y * t0.y + 4.0f * e2.y * k2.y),
               sqrtf(t0.z * t0.z + 4.0f * e2.z * k2.z)};
    f3 t1 = ab + s3(c2);
    f3 a = f3{sqrtf(0.5f * (ab.x + t0.x)), sqrtf(0.5f * (ab.y + t0.y)), sqrtf(0.5f * (ab.z + t0.z))};
    f3 t2 = (2.0f * a) * c;
    f3 rs = (t1 - t2) / (t1 + t2);
    f3 t3 = c2 * ab + s3(s2 * s2);
    f3 t4 = t2 * s2;
    f3 rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}
YM_HD void edge_tint(f3 refl, f3 tint, f3 &n, f3 &k)
{
    f3 r{gclamp(refl.x, 0.f, 0.99f), gclamp(refl.y, 0.f, 0.99f), gclamp(refl.z, 0.f, 0.99f)};
    f3 rs = sqrt3(r);
    const f3 one = s3(1.0f);
    f3 nmin = (one - r) / (one + r);
    f3 nmax = (one + rs) / (one - rs);
    n = mixv(nmax, nmin, tint);
    f3 np1 = n + one, nm1 = n - one;
    f3 k2 = (np1 * np1 * r - nm1 * nm1) / (one - r);
    k2 = max3(k2, s3(0.0f));
    k = sqrt3(k2);
}
YM_HD float ggx_d(f3 h, float ax, float ay)
{
    float den = kPi * ax * ay * sqr(sqr(h.x / ax) + sqr(h.y / ay) + sqr(h.z));
    return 1.0f / gmax(den, 1E-12f);
}
YM_HD float ggx_lambda(f3 w, float ax, float ay)
{
    if (fabsf(w.z) < 1E-8f) return 0.0f;
    return (-1.0f + sqrtf(1.0f + (sqr(w.x * ax) + sqr(w.y * ay)) / sqr(w.z))) * 0.5f;
}
YM_HD float ggx_g1(f3 w, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(w, ax, ay)); }
YM_HD float ggx_g2(f3 a, f3 b, float ax, float ay)
{
    return 1.0f / (1.0f + ggx_lambda(a, ax, ay) + ggx_lambda(b, ax, ay));
}
// d ln f / d r of an isotropic GGX conductor lobe at (wo, wi) in the shading
// frame, alpha = r^2 (InitBSDFSettings' clamp to (0.001, 0.999) inside: the
// derivative is 0 where it holds).  f = F D G2 / (4 cos_o cos_i) depends on
// alpha through D = 1 / (pi a^2 (s / a^2 + hz^2)^2), s = hx^2 + hy^2 of the
// unit half vector, and G2 = 1 / (1 + Lambda(wo) + Lambda(wi)) with
// Lambda(w) = (sqrt(1 + a^2 tan^2(w)) - 1) / 2; Fresnel does not depend on
// it.  d ln D / da = -2 / a + 4 s / (a^3 q) with q = s / a^2 + hz^2, d Lambda
// / da = a tan^2 / (2 sqrt(1 + a^2 tan^2)), d / dr = 2 r d / da.  Double
// precision: it is summed over many paths.
__device__ __forceinline__ float ggx_rough_dlog(float r, f3 wo, f3 wi)
{
    if (!(r > 0.001f && r < 0.999f) || !(wo.z > 0.0f && wi.z > 0.0f)) return 0.0f;
    const double a = (double)r * r;
    double hx = (double)wo.x + wi.x, hy = (double)wo.y + wi.y, hz = (double)wo.z + wi.z;
    const double hl = sqrt(hx * hx + hy * hy + hz * hz);
    if (hl < 1e-12) return 0.0f;
    hx /= hl;
    hy /= hl;
    hz /= hl;
    const double sh = hx * hx + hy * hy, q = sh / (a * a) + hz * hz;
    const double dlnd = -2.0 / a + 4.0 * sh / (a * a * a * q);
    double lam = 0.0, dlam = 0.0;
    const f3 ws[2] = {wo, wi};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double wz = ws[k].z;
        if (fabs(wz) < 1e-8) continue;
        const double t2 = ((double)ws[k].x * ws[k].x + (double)ws[k].y * ws[k].y) / (wz * wz);
        const double root = sqrt(1.0 + a * a * t2);
        lam += 0.5 * (root - 1.0);
        dlam += 0.5 * a * t2 / root;
    }
    return (float)((dlnd - dlam / (1.0 + lam)) * 2.0 * r);
}

__device__ __forceinline__ f3 ggx_sample_wh(f3 wo, float u0, float u1, float ax, float ay)
{
    f3 wh = normalize(f3{ax * wo.x, ay * wo.y, wo.z});
    float lensq = wh.x * wh.x + wh.y * wh.y;
    f3 T1 = lensq > 0 ? f3{-wh.y, wh.x, 0} * (1.0f / sqrtf(lensq)) : f3{1, 0, 0};
    f3 T2 = cross(wh, T1);
    float r = sqrtf(gmax(u0, 0.0f));
    float phi = 2.0f * kPi * u1;
    float sp, cp;
    ym_sincosf(phi, sp, cp);
    float t1 = r * cp;
    float t2 = r * sp;
    float s = (float)(0.5 * (1.0 + (double)wh.z));
    t2 = (float)((1.0 - (double)s) * sqrt(1.0 - (double)(t1 * t1)) + (double)s * (double)t2);
    f3 Nh = t1 * T1 + t2 * T2 + sqrtf(gmax(0.0f, 1.0f - t1 * t1 - t2 * t2)) * wh;
    return normalize(f3{ax * Nh.x, ay * Nh.y, gmax(0.0f, Nh.z)});
}
__device__ __forceinline__ f3 sample_cosine(float u0, float u1)
{
    float r = sqrtf(u0);
    float phi = u1 * kTwoPi;
    float sp, cp;
    ym_sincosf(phi, sp, cp);
    float x = r * cp, y = r * sp;
    float r2 = x * x + y * y;
    return f3{x, y, r2 < 1.0f ? sqrtf(1.0f - r2) : 0.0f};
}

enum : int { kLobeLambert = 0, kLobeReflect = 1, kLobeTransmit = 2 };

struct Lobe {
    int type;
    int metal;
    f3 albedo;
    f3 nt, kt;   // reflection: eta, k (transmission uses nt.x)
    float ni;
    float ax, ay;
};

YM_HD f3 lobe_eval(const Lobe &l, f3 wo, f3 wi, float &pdf)
{
    if (l.type == kLobeLambert) {
        if (!same_hemi(wo, wi)) {
            pdf = 0.0f;
            return s3(0.0f);
        }
        float c = gmax(wi.z, 0.0f);
        pdf = c * kInvPi;
        return l.albedo * c * kInvPi;
    }
    if (l.type == kLobeReflect) {
        if (!same_hemi(wo, wi) || wo.z < 0.f || wi.z < 0.f) {
            pdf = 0.0f;
            return s3(0.0f);
        }
        f3 wh = wi + wo;
        float len = length(wh);
        if (fabsf(len) < 1E-8f) {
            pdf = 0.0f;
            return s3(0.0f);
        }
        wh = wh / len;
        float D = ggx_d(wh, l.ax, l.ay);
        float G1 = ggx_g1(wo, l.ax, l.ay);
        float G2 = ggx_g2(wo, wi, l.ax, l.ay);
        float ini = safe_rcp(l.ni);
        float c = gmax(dot(wo, wh), 0.0f);
        f3 fr = l.metal ? fresnel_dc(l.nt * ini, l.kt * ini, c) : s3(fresnel_dd(l.nt.x * ini, c));
        pdf = D * G1 * 0.25f / gmax(wo.z, 1E-8f);
        return l.albedo * D * G2 * 0.25f * fr / gmax(wo.z, 1E-8f);
    }
    if (same_hemi(wo, wi)) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    const float nts = l.nt.x;
    f3 wh = -(l.ni * wo + nts * wi);
    if (!same_hemi(wo, wh)) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    float len = length(wh);
    if (fabsf(len) < 1E-8f) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    wh = wh / len;
    float D = ggx_d(wh, l.ax, l.ay);
    float G1 = ggx_g1(wo, l.ax, l.ay);
    float G2 = ggx_g2(wo, wi, l.ax, l.ay);
    float dwh = sqr(nts) * fabsf(dot(wi, wh)) / gmax(sqr(l.ni * dot(wo, wh) + nts * dot(wi, wh)), 1E-12f);
    float ini = safe_rcp(l.ni);
    float c = gmax(dot(wo, wh), 0.0f);
    float fr = fresnel_dd(nts * ini, c);
    pdf = D * G1 * fabsf(dot(wo, wh)) * dwh / gmax(fabsf(wo.z), 1E-8f);
    f3 r = l.albedo * (1.0f - fr);
    r = r * fabsf(dot(wo, wh) / gmax(fabsf(wo.z), 1E-8f));
    r = r * G2;
    r = r * D;
    r = r * dwh;
    r = r * sqr(l.ni / nts);
    return r;
}

__device__ __forceinline__ bool lobe_sample(const Lobe &l, float u0, float u1, f3 wo, f3 &weight, f3 &wi,
                                            float &pdf)
{
    if (l.type == kLobeLambert) {
        wi = sample_cosine(u0, u1);
        float c = gmax(wi.z, 0.0f);
        pdf = c * kInvPi;
        weight = l.albedo;
        return true;
    }
    f3 wh = ggx_sample_wh(wo, u0, u1, l.ax, l.ay);
    if (!same_hemi(wo, wh) || dot(wo, wh) < 0.0f) return false;
    if (l.type == kLobeReflect) {
        wi = reflect3(wo, wh);
        if (!same_hemi(wi, wo)) return false;
        float D = ggx_d(wh, l.ax, l.ay);
        float G1 = ggx_g1(wo, l.ax, l.ay);
        float G2 = ggx_g2(wo, wi, l.ax, l.ay);
        float ini = safe_rcp(l.ni);
        float c = gmax(dot(wo, wh), 0.0f);
        f3 fr = l.metal ? fresnel_dc(l.nt * ini, l.kt * ini, c) : s3(fresnel_dd(l.nt.x * ini, c));
        pdf = D * G1 * 0.25f / gmax(wo.z, 1E-8f);
        weight = l.albedo * G2 * fr / G1;
        return true;
    }
    const float nts = l.nt.x;
    if (!refract3(wo, wh, l.ni, nts, wi)) return false;
    if (same_hemi(wi, wo)) return false;
    float D = ggx_d(wh, l.ax, l.ay);
    float G1 = ggx_g1(wo, l.ax, l.ay);
    float G2 = ggx_g2(wi, wo, l.ax, l.ay);
    float dwh = sqr(nts) * fabsf(dot(wi, wh)) / gmax(sqr(l.ni * dot(wo, wh) + nts * dot(wi, wh)), 1E-12f);
    float ini = safe_rcp(l.ni);
    float c = gmax(dot(wo, wh), 0.0f);
    float fr = fresnel_dd(nts * ini, c);
    pdf = D * G1 * fabsf(dot(wo, wh)) * dwh / gmax(fabsf(wo.z), 1E-8f);
    weight = l.albedo * (1.0f - fr) * sqr(l.ni / nts) * G2 / G1;
    return true;
}

// DefaultMtlBSDF + MaterialBSDF (src/BSDF.cuh:432-807), no normal/bump maps.
// The reference appends up to five lobes to an array (coat, metal, specular,
// transmission, diffuse -- always in that order, absent ones skipped).  Here
// each lobe has a fixed slot and a presence bit; eval/sample walk the slots in
// the same order, so sums and the lobe-selection CDF are the reference's, and
// the per-lobe data (colours plus the shared roughness/IOR terms) stays in
// registers instead of a dynamically indexed array.
enum : int { kSlotCoat = 0, kSlotMetal = 1, kSlotSpec = 2, kSlotTrans = 3, kSlotDiffuse = 4, kSlots = 5 };

struct Bsdf {
    f3 t, b, n;
    uint32_t mask;      // presence bit per slot; 0 for non-default materials
    float w[kSlots];
    f3 col[kSlots];
    float ni, ax, ay, cax, cay, ior_coat, ior_out;
    f3 metal_n, metal_k;

    YM_HD f3 to_local(f3 v) const { return f3{dot(v, t), dot(v, b), dot(v, n)}; }
    YM_HD f3 to_world(f3 v) const { return v.x * t + v.y * b + v.z * n; }
    YM_HD Lobe lobe(int slot) const
    {
        Lobe l;
        l.albedo = col[slot];
        l.ni = ni;
        l.kt = s3(0.0f);
        l.metal = 0;
        l.ax = ax;
        l.ay = ay;
        if (slot == kSlotCoat) {
            l.type = kLobeReflect;
            l.ax = cax;
            l.ay = cay;
            l.nt = s3(ior_coat);
        } else if (slot == kSlotMetal) {
            l.type = kLobeReflect;
            l.nt = metal_n;
            l.kt = metal_k;
            l.metal = 1;
        } else if (slot == kSlotSpec) {
            l.type = kLobeReflect;
            l.nt = s3(ior_out);
        } else if (slot == kSlotTrans) {
            l.type = kLobeTransmit;
            l.nt = s3(ior_out);
        } else {
            l.type = kLobeLambert;
            l.nt = s3(0.0f);
        }
        return l;
    }
};

// Camera::TextureCoordinatesDifferential (src/Camera.h:56-140): (du/dx,
// dv/dx, du/dy, dv/dy) of a world-space hit, from rays through pixel
// neighbours offset by (sx, sy) pixels.
__device__ __forceinline__ float4 camera_uv_differentials(const DevCamera &cam, f3 hp, f3 sn, f3 ng, f3 dpdu,
                                                          f3 dpdv, uint32_t width, uint32_t height, float sx,
                                                          float sy)
{
    f3 look = cam.position - hp;
    const float dist = length(look);
    look = look * safe_rcp(dist);
    const f3 right = normalize(cam.u - look * dot(cam.u, look));
    const f3 up = cross(look, right);
    const float iph = 2.0f * cam.tan_half_fov;
    const float yoff = iph * sy / float(height);
    const float ipw = iph * cam.aspect;
    const float xoff = ipw * sx / float(width);
    auto min_abs = [](float f, float th) { return fabsf(f) < th ? (f > 0.0f ? th : -th) : f; };
    const f3 dxd = normalize(f3{0.0f, 0.0f, -1.0f} + right * xoff);
    const f3 dyd = normalize(f3{0.0f, 0.0f, -1.0f} + up * yoff);
    const f3 lsn{dot(sn, right), dot(sn, up), dot(sn, look)};
    const f3 lp{0.0f, 0.0f, -dist};
    const float tx = dot(lsn, lp) / min_abs(dot(lsn, dxd), 1E-4f);
    f3 dpdx = (dxd * tx - lp) * safe_rcp(sx);
    const float ty = dot(lsn, lp) / min_abs(dot(lsn, dyd), 1E-4f);
    f3 dpdy = (dyd * ty - lp) * safe_rcp(sy);
    dpdx = dpdx.x * right + dpdx.y * up + dpdx.z * look;
    dpdy = dpdy.x * right + dpdy.y * up + dpdy.z * look;
    const f3 an = abs3(ng);
    int d0, d1;
    if (an.z > an.y && an.z > an.x) {
        d0 = 0;
        d1 = 1;
    } else if (an.y > an.x) {
        d0 = 0;
        d1 = 2;
    } else {
        d0 = 1;
        d1 = 2;
    }
    const float m00 = comp(dpdu, d0), m01 = comp(dpdv, d0), m10 = comp(dpdu, d1), m11 = comp(dpdv, d1);
    const float det = m00 * m11 - m10 * m01;
    if (fabsf(det) < 1E-7f) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float idet = 1.0f / det;
    const float dudx = m11 * comp(dpdx, d0) - m01 * comp(dpdx, d1);
    const float dvdx = -m10 * comp(dpdx, d0) + m00 * comp(dpdx, d1);
    const float dudy = m11 * comp(dpdy, d0) - m01 * comp(dpdy, d1);
    const float dvdy = -m10 * comp(dpdy, d0) + m00 * comp(dpdy, d1);
    return make_float4(dudx * idet, dvdx * idet, dudy * idet, dvdy * idet);
}

YM_HD void shading_frame(Bsdf &bs, int hit_back, f3 sn, f3 su)
{
    if (!hit_back) {
        bs.n = sn;
        bs.t = normalize(su);
        bs.b = cross(bs.n, bs.t);
    } else {
        bs.n = -sn;
        bs.t = -normalize(su);
        bs.b = -cross(bs.n, bs.t);
    }
}

// MaterialBSDF::InitShadingSpace (src/BSDF.cuh:694-807): shading frame from
// the interpolated normal and dpdu, perturbed by a normal map or a bump map
// (finite differences of the height over the texture-coordinate footprint
// duv), then flipped to face the incoming ray.  normal_tex / bump_tex are
// kEmpty when the material has none.
template <class TX>
__device__ __forceinline__ void init_shading_space(Bsdf &bs, TX &tx, uint32_t normal_tex, uint32_t bump_tex,
                                                   int hit_back, f3 ray_d, f3 ng, f3 hsn, f3 dpdu, f3 dpdv,
                                                   const Affine &w2o, float4 duv)
{
    f3 sn = hsn;
    if (dot(hsn, ng) < 0.0f) sn = -hsn;
    f3 su = dpdu;
    if (dot(cross(dpdu, dpdv), ng) < 0.0f) su = -dpdu;
    f3 sv = dpdv;
    su = su - sn * dot(su, sn);
    sv = sv - sn * dot(sv, sn);
    if (normal_tex != kEmpty) {
        shading_frame(bs, hit_back, sn, su);
        const f3 tn = normalize(tx.eval(normal_tex) * 2.0f - s3(1.0f));
        bs.n = tn.x * bs.t + tn.y * bs.b + tn.z * bs.n;
        bs.t = normalize(bs.t - bs.n * dot(bs.t, bs.n));
        bs.b = hit_back ? -cross(bs.n, bs.t) : cross(bs.n, bs.t);
    } else if (bump_tex != kEmpty) {
        const f3 osu = xf_vector(w2o, su), osv = xf_vector(w2o, sv);
        float du = gmax((fabsf(duv.x) + fabsf(duv.z)) * 0.5f, 0.0005f);
        if (!(du > 0.0f)) du = 0.0005f;
        TexCoord tu = tx.tc;
        tu.st = f2{tx.tc.st.x + du, tx.tc.st.y};
        tu.p_world = tx.tc.p_world + su * du;
        tu.p_object = tx.tc.p_object + osu * du;
        const float hu = tx.eval(bump_tex, tu).x;
        float dv = gmax((fabsf(duv.y) + fabsf(duv.w)) * 0.5f, 0.0005f);
        if (!(dv > 0.0f)) dv = 0.0005f;
        TexCoord tvv = tx.tc;
        tvv.st = f2{tx.tc.st.x, tx.tc.st.y + dv};
        tvv.p_world = tx.tc.p_world + sv * dv;
        tvv.p_object = tx.tc.p_object + osv * dv;
        const float hv = tx.eval(bump_tex, tvv).x;
        const float h0 = tx.eval(bump_tex).x;
        su = su + (hu - h0) * safe_rcp(du) * sn;
        sv = sv + (hv - h0) * safe_rcp(dv) * sn;
        sn = normalize(cross(su, sv));
        shading_frame(bs, hit_back, sn, su);
    } else {
        shading_frame(bs, hit_back, sn, su);
    }
    if (dot(-ray_d, bs.n) < 0.0f) {
        bs.n = reflect3(bs.n, dot(bs.n, ng) > 0.0f ? ng : -ng);
        bs.t = normalize(bs.t - bs.n * dot(bs.t, bs.n));
        bs.b = hit_back ? -cross(bs.n, bs.t) : cross(bs.n, bs.t);
    }
}

template <class TX>
__device__ __forceinline__ float tex_float(TX &tx, uint32_t tex, float def)
{
    return tex != kEmpty ? tx.eval(tex).x : def;
}
template <class TX>
__device__ __forceinline__ f3 tex_color(TX &tx, uint32_t tex, const float *def)
{
    return tex != kEmpty ? tx.eval(tex) : f3{def[0], def[1], def[2]};
}

// DefaultMtlBSDF::InitBSDFSettings (src/BSDF.cuh:453-618).  kClass names a
// material class whose weights the general code below fixes exactly:
//  kMatDiffuse: zero coat, metal, specular and transmission weights and a
//   coat that cannot overflow; the diffuse lobe is left alone with the albedo
//   as its colour (tint, (1 - metal), (1 - F * 0) and (1 - 0) are all 1).
//  kMatMetal: metalness 1, zero coat weight and a finite coat transmittance;
//   tint is exactly 1, so the metal lobe has colour (1, 1, 1), and every
//   product with (1 - metal) = 0 is a signed zero or NaN, which the "> 0"
//   tests of the specular and transmission lobes reject.  The diffuse lobe
//   stays, weighted as in the general code.
template <uint32_t kClass = kMatGeneral, class TX>
__device__ __forceinline__ void init_default_bsdf(TX &tx, Bsdf &b, const DevMaterial &m, f3 ray_d,
                                                  int hit_back, float ray_ior, float ex_ior)
{
    const float ior_min = 0.0001f;
    constexpr bool kNoCoat = kClass == kMatMetal;
    if constexpr (kClass == kMatDiffuse) {
        const f3 dif = tex_color(tx, m.diffuse_albedo_tex, m.diffuse_albedo);
        b.ni = gmax(ray_ior, ior_min);
        b.ax = b.ay = b.cax = b.cay = 0.0f;
        b.ior_coat = b.ior_out = 1.0f;
        b.metal_n = b.metal_k = s3(0.0f);
#pragma unroll
        for (int i = 0; i < kSlots; ++i) {
            b.w[i] = 0.0f;
            b.col[i] = s3(0.0f);
        }
        b.mask = 1u << kSlotDiffuse;
        b.col[kSlotDiffuse] = dif;
        const float dif_lum = luminance(dif);
        const float wsum = 0.0f + dif_lum;
        b.w[kSlotDiffuse] = wsum > 0.0f ? dif_lum * safe_rcp(wsum) : 1.0f;
        return;
    }
    const float coat_w = kNoCoat ? 0.0f : tex_float(tx, m.coat_weight_tex, m.coat_weight);
    const float ior_in = gmax(ray_ior, ior_min);
    const float ior_coat = gmax(m.coat_ior, ior_min);
    const float ior_out = gmax(hit_back ? ex_ior : m.ior_n, ior_min);
    const float eta_ic = ior_coat * safe_rcp(ior_in);
    const float eta_io = ior_out * safe_rcp(ior_in);
    const float cos_in = gmax(dot(b.n, -ray_d), 0.0f);
    const float fr_ic = !kNoCoat && coat_w > 0.0f ? fresnel_dd(eta_ic, cos_in) : 1.0f;
    const float fr_io = fresnel_dd(eta_io, cos_in);
    const float metal = kNoCoat ? 1.0f : tex_float(tx, m.metalness_tex, m.metalness);
    const float ax = sqr(gclamp(tex_float(tx, m.roughness_x_tex, m.roughness_x), 0.001f, 0.999f));
    const float ay = sqr(gclamp(tex_float(tx, m.roughness_y_tex, m.roughness_y), 0.001f, 0.999f));
    const float spec_w = tex_float(tx, m.specular_weight_tex, m.specular_weight);
    const float trans_w = tex_float(tx, m.transmission_weight_tex, m.transmission_weight);
    const f3 dif = tex_color(tx, m.diffuse_albedo_tex, m.diffuse_albedo);
    const f3 spe = tex_color(tx, m.specular_albedo_tex, m.specular_albedo);
    const float coat_th = tex_float(tx, m.coat_thickness_tex, m.coat_thickness);
    const float cax = sqr(gclamp(tex_float(tx, m.coat_roughness_x_tex, m.coat_roughness_x), 0.001f, 0.999f));
    const float cay = sqr(gclamp(tex_float(tx, m.coat_roughness_y_tex, m.coat_roughness_y), 0.001f, 0.999f));
    const f3 coat_alb = tex_color(tx, m.coat_albedo_tex, m.coat_albedo);
    // coat transmittance; multiplied by the coat weight (0 for kNoCoat, where
    // it is finite, so the product is 0 whatever its value)
    f3 coat_tr = s3(0.0f);
    if constexpr (!kNoCoat) {
        const f3 ca = max3(coat_alb, s3(1E-8f));
        const f3 sigma = (-f3{ym_logf(ca.x), ym_logf(ca.y), ym_logf(ca.z)}) * coat_th;
        const float rc = safe_rcp(cos_in);
        const f3 ns = -sigma;
        coat_tr = f3{ym_expf(ns.x * rc), ym_expf(ns.y * rc), ym_expf(ns.z * rc)};
    }
    b.ni = ior_in;
    b.ax = ax;
    b.ay = ay;
    b.cax = cax;
    b.cay = cay;
    b.ior_coat = ior_coat;
    b.ior_out = ior_out;
    b.metal_n = s3(0.0f);
    b.metal_k = s3(0.0f);
    b.mask = 0;
    float wsum = 0.0f;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
        b.w[i] = 0.0f;
        b.col[i] = s3(0.0f);
    }
    const f3 coat_color = s3(coat_w * fr_ic);
    const float coat_lum = luminance(coat_color);
    if (!kNoCoat && coat_lum > 0.0f) {
        b.mask |= 1u << kSlotCoat;
        b.col[kSlotCoat] = coat_color;
        b.w[kSlotCoat] = coat_lum;
        wsum += coat_lum;
    }
    const f3 lerp = (1.0f - coat_w) * s3(1.0f) + coat_tr * coat_w;
    const f3 tint = (1.0f - coat_w * fr_ic) * lerp;
    const f3 metal_c = tint * metal;
    const float metal_lum = luminance(metal_c);
    if (metal_lum > 0.0f) {
        edge_tint(dif, spe, b.metal_n, b.metal_k);
        b.mask |= 1u << kSlotMetal;
        b.col[kSlotMetal] = metal_c;
        b.w[kSlotMetal] = metal_lum;
        wsum += metal_lum;
    }
    const f3 spec_c = tint * (1.0f - metal) * fr_io * spec_w * spe;
    const float spec_lum = luminance(spec_c);
    if (!kNoCoat && spec_lum > 0.0f) {
        b.mask |= 1u << kSlotSpec;
        b.col[kSlotSpec] = spec_c;
        b.w[kSlotSpec] = spec_lum;
        wsum += spec_lum;
    }
    const f3 trans_c = tint * (1.0f - metal) * (1.0f - fr_io * spec_w) * trans_w * spe;
    const float trans_lum = luminance(trans_c);
    if (!kNoCoat && trans_lum > 0.0f) {
        b.mask |= 1u << kSlotTrans;
        b.col[kSlotTrans] = trans_c;
        b.w[kSlotTrans] = trans_lum;
        wsum += trans_lum;
    }
    const f3 dif_c = tint * (1.0f - metal) * (1.0f - fr_io * spec_w) * (1.0f - trans_w) * dif;
    const float dif_lum = luminance(dif_c);
    b.mask |= 1u << kSlotDiffuse;
    b.col[kSlotDiffuse] = dif_c;
    b.w[kSlotDiffuse] = dif_lum;
    wsum += dif_lum;
    if (wsum > 0.0f) {
        const float iw = safe_rcp(wsum);
#pragma unroll
        for (int i = 0; i < kSlots; ++i)
            if (b.mask & (1u << i)) b.w[i] *= iw;
    } else {
        const float iw = 1.0f / (float)__popc(b.mask);
#pragma unroll
        for (int i = 0; i < kSlots; ++i)
            if (b.mask & (1u << i)) b.w[i] = iw;
    }
}

// DefaultMtlBSDF::EvalWi (src/BSDF.cuh:619-632).  kFixed: the lobe mask
// known at compile time (0: b.mask).
template <uint32_t kFixed = 0>
YM_HD f3 bsdf_eval(const Bsdf &b, f3 wo, f3 wi, float &pdf)
{
    const uint32_t mask = kFixed ? kFixed : b.mask;
    f3 f = s3(0.0f);
    float p = 0.0f;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
        if (!(mask & (1u << i))) continue;
        float lp = 0.0f;
        f = f + lobe_eval(b.lobe(i), wo, wi, lp);
        p += b.w[i] * lp;
    }
    pdf = p;
    return f;
}

// DefaultMtlBSDF::SampleWi (src/BSDF.cuh:633-679); kFixed as for bsdf_eval.
template <uint32_t kFixed = 0>
__device__ __forceinline__ bool bsdf_sample(const Bsdf &b, f3 wo, f3 &weight, f3 &wi, float &pdf, float u0,
                                            float u1, float u2)
{
    const uint32_t mask = kFixed ? kFixed : b.mask;
    if (!mask) return false;
    float cdf = 0.0f;
    int sel = -1, last = -1;
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
        if (!(mask & (1u << i))) continue;
        last = i;
        if (sel < 0) {
            if (u0 < cdf + b.w[i]) sel = i;
            else cdf += b.w[i];
        }
    }
    if (sel == -1) sel = last;
    float lp = 0.0f;
    bool ok = false;
#pragma unroll
    for (int i = 0; i < kSlots; ++i)
        if (i == sel) ok = lobe_sample(b.lobe(i), u1, u2, wo, weight, wi, lp);
    if (!ok) return false;
    float mp = 0.0f;
    f3 mf = s3(0.0f);
#pragma unroll
    for (int i = 0; i < kSlots; ++i)
        if (i == sel) {
            mp = lp * b.w[i];
            mf = weight * lp;
        }
#pragma unroll
    for (int i = 0; i < kSlots; ++i) {
        if (!(mask & (1u << i)) || i == sel) continue;
        float p = 0.0f;
        mf = mf + lobe_eval(b.lobe(i), wo, wi, p);
        mp += b.w[i] * p;
    }
    pdf = mp;
    weight = mf * safe_rcp(mp);
    return true;
}

// ---------------------------------------------------------------- phases
// src/Volume.h:21-48, src/VolumeEval.cuh:470-538
YM_HD float hg_eval(float g, f3 wo, f3 wi, float &pdf)
{
    float c = dot(wo, wi);
    float t = 1.f + g * g + 2.f * g * c;
    pdf = 0.25f * kInvPi * (1.f - g * g) * safe_rcp(t * sqrtf(gmax(t, 0.0f)));
    return pdf;
}
__device__ __forceinline__ void hg_sample(float g, float u0, float u1, f3 wo, f3 &wi, float &pdf)
{
    float c;
    if (fabsf(g) < 1e-4f) {
        c = 1.f - 2.f * u0;
    } else {
        float t = (1.f - g * g) * safe_rcp(1.f - g + 2.f * g * u0);
        c = (1.f + g * g - t * t) / (2.f * g);
    }
    float s = sqrtf(gmax(0.0f, 1.0f - c * c));
    float phi = 2.f * kPi * u1;
    f3 w = wo, u, v;
    orthogonal_basis(w, u, v);
    float sp, cp;
    ym_sincosf(phi, sp, cp);
    wi = u * s * cp + v * s * sp + w * c;
    hg_eval(g, wo, wi, pdf);
}
// SampleMixedPhases (src/VolumeEval.cuh:470-514); loops over the six fixed
// slots with a count guard so vw/gs stay in registers.
__device__ __forceinline__ bool mixed_phase_sample(const float *vw, const float *gs, int n, float u0, float u1,
                                                   f3 wo, float &weight, f3 &wi, float &pdf)
{
    if (n == 0) return false;
    float cdf = 0.0f;
    int sel = -1;
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n || sel >= 0) continue;
        if (u0 < cdf + vw[i]) {
            sel = i;
            u0 = (u0 - cdf) * safe_rcp(vw[i]);
        } else {
            cdf += vw[i];
        }
    }
    if (sel == -1) sel = n - 1;
    float gsel = gs[0], vsel = vw[0];
#pragma unroll
    for (int i = 1; i < kMaxBoundaryRecord; ++i)
        if (i == sel) {
            gsel = gs[i];
            vsel = vw[i];
        }
    float pp = 0.0f;
    hg_sample(gsel, u0, u1, wo, wi, pp);
    const float w1 = 1.0f;
    float mp = pp * vsel;
    float mg = w1 * pp * vsel;
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n || i == sel) continue;
        float p;
        mg += vw[i] * hg_eval(gs[i], wo, wi, p);
        mp += vw[i] * p;
    }
    pdf = mp;
    weight = mg * safe_rcp(mp);
    return true;
}
// EvalMixedPhases (src/VolumeEval.cuh:516-538)
YM_HD float mixed_phase_eval(const float *vw, const float *gs, int n, f3 wo, f3 wi, float &pdf)
{
    if (n == 0) {
        pdf = 0.0f;
        return 0.0f;
    }
    float mg = 0.0f, mp = 0.0f;
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n) continue;
        float p = 0.0f;
        float g = hg_eval(gs[i], wo, wi, p);
        mg += vw[i] * g;
        mp += vw[i] * p;
    }
    pdf = mp;
    return mg;
}

// ---------------------------------------------------------------- media
__device__ __forceinline__ float grid_density(const DevScene &sc, int gi, f3 p)
{
    const DevGrid g = sc.grids[gi];
    if (!(p.x >= 0.0f && p.x < 1.0f && p.y >= 0.0f && p.y < 1.0f && p.z >= 0.0f && p.z < 1.0f)) return 0.0f;
    uint32_t ix = (uint32_t)(p.x * (float)g.nx), iy = (uint32_t)(p.y * (float)g.ny), iz = (uint32_t)(p.z * (float)g.nz);
    if (ix >= g.nx) ix = g.nx - 1;
    if (iy >= g.ny) iy = g.ny - 1;
    if (iz >= g.nz) iz = g.nz - 1;
    return sc.voxels[g.offset + ((uint64_t)iz * g.ny + iy) * g.nx + ix];
}

// The voxel of grid gi that grid_density reads at object point p, or -1
// outside the unit cube.
__device__ __forceinline__ int grid_voxel(const DevScene &sc, int gi, f3 p)
{
    const DevGrid g = sc.grids[gi];
    if (!(p.x >= 0.0f && p.x < 1.0f && p.y >= 0.0f && p.y < 1.0f && p.z >= 0.0f && p.z < 1.0f)) return -1;
    uint32_t ix = (uint32_t)(p.x * (float)g.nx), iy = (uint32_t)(p.y * (float)g.ny), iz = (uint32_t)(p.z * (float)g.nz);
    if (ix >= g.nx) ix = g.nx - 1;
    if (iy >= g.ny) iy = g.ny - 1;
    if (iz >= g.nz) iz = g.nz - 1;
    return (int)((iz * g.ny + iy) * g.nx + ix);
}

// Heterogeneous media: a voxel grid, or a texture graph as the reference's
// density_texture_idx (SampleVolumeScattering's test, src/VolumeEval.cuh:231-238).
__device__ __forceinline__ bool vol_is_het(const DevScene &sc, int v)
{
    const DevVolume &dv = sc.volumes[v];
    return (dv.grid >= 0 && (uint32_t)dv.grid < sc.grid_count) ||
           (dv.tex >= 0 && (uint32_t)dv.tex < sc.texture_count);
}

// Majorant of a heterogeneous volume's density: Texture::GetUpperBound
// (src/Texture.h:1181-1251, computed on the host) or the grid's maximum.
__device__ __forceinline__ float vol_bound(const DevScene &sc, const DevVolume &v)
{
    return v.tex >= 0 ? v.tex_bound : sc.grids[v.grid].max_density;
}

// Density of a heterogeneous volume at a tracking step, max(., 0): the x
// channel of its texture graph at (st = 0, world point, object point) as
// src/VolumeEval.cuh:165-176 evaluates it, or the grid's voxel at the object
// point.  kTex: the scene has texture-density media (the graph evaluator and
// its LDS stack, tex_stk, are compiled in).
template <bool kTex>
__device__ __forceinline__ float vol_density(const DevScene &sc, const DevVolume &v, f3 pw, f3 po, float *tex_stk)
{
    if constexpr (kTex) {
        if (v.tex >= 0) {
            TexEval<true> tx{sc, TexCoord{f2{0.0f, 0.0f}, pw, po}, tex_stk, (int)__lane_id()};
            return gmax(tx.eval((uint32_t)v.tex).x, 0.0f);
        }
    }
    return gmax(grid_density(sc, v.grid, po), 0.0f);
}

__device__ __forceinline__ f3 vsig(const DevVolume &v) { return f3{v.sigma_t[0], v.sigma_t[1], v.sigma_t[2]}; }
__device__ __forceinline__ f3 valb(const DevVolume &v) { return f3{v.albedo[0], v.albedo[1], v.albedo[2]}; }
YM_HD f3 exp3(f3 a) { return f3{ym_expf(a.x), ym_expf(a.y), ym_expf(a.z)}; }

// Random source for the shadow-ray transmittance of grid media: PCG hash of
// (pixel, vertex, sample, frame) and a per-vertex counter.
struct HashRng {
    uint32_t key[4];
    uint32_t ctr;
    __device__ __forceinline__ float next()
    {
        uint32_t v[4] = {key[0], key[1] ^ (ctr << 8), key[2], key[3]};
        ctr++;
        pcg4d(v);
        return (float)(v[0] >> 8) * (1.0f / 16777216.0f);
    }
};

// Gradient events of the delta and ratio trackers (the adjoint's replay
// pass): note() sees each heterogeneous volume's object-space point at a
// tentative collision, then real() or null() gets the collision's kind with
// the summed sigma_t (channel x) there and the majorant.  NoTrackEvents
// compiles to nothing.
struct NoTrackEvents {
    __device__ __forceinline__ void note(int, f3) {}
    __device__ __forceinline__ void real(float, float) {}
    __device__ __forceinline__ void null(float, float) {}
};

// src/VolumeEval.cuh:18-78
__device__ __forceinline__ bool sample_homogeneous(const DevScene &sc, int n, const int *vi, float tmax, f3 &trw,
                                                   float &dist, Sampler &s)
{
    trw = s3(1.0f);
    dist = kFloatMax;
    f3 st = s3(0.0f), ss = s3(0.0f);
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n) continue;
        const DevVolume &v = sc.volumes[vi[i]];
        st = st + vsig(v);
        ss = ss + vsig(v) * valb(v);
    }
    const float csum = st.x + st.y + st.z;
    const f3 cw = csum > 0.0f ? st / csum : s3(1.0f / 3.0f);
    float u0 = sampler_next(sc, s);
    int ch;
    if (u0 < cw.x) {
        ch = 0;
        u0 = u0 * safe_rcp(cw.x);
    } else if (u0 < cw.x + cw.y) {
        ch = 1;
        u0 = (u0 - cw.x) * safe_rcp(cw.y);
    } else {
        ch = 2;
        u0 = (u0 - cw.x - cw.y) * safe_rcp(cw.z);
    }
    const float sc_ = comp(st, ch);
    const float fpl = sc_ > 0.0f ? (-ym_logf(1.0f - u0) / sc_) : kFloatMax;
    const bool scatter = fpl < tmax;
    const float m = gmin(tmax, fpl);
    const f3 tr{ym_expf(-st.x * m), ym_expf(-st.y * m), ym_expf(-st.z * m)};
    const float pdf = scatter ? (cw.x * st.x * tr.x + cw.y * st.y * tr.y + cw.z * st.z * tr.z)
                              : (cw.x * tr.x + cw.y * tr.y + cw.z * tr.z);
    if (!(fabsf(pdf) > 0.0f)) return false;
    trw = scatter ? (tr * ss) / pdf : tr / pdf;
    dist = fpl;
    return scatter;
}

// src/VolumeEval.cuh:80-214 with voxel-grid or texture densities
template <bool kTex, typename Ev = NoTrackEvents>
__device__ __forceinline__ bool sample_heterogeneous(const DevScene &sc, int n, const int *vi, f3 ro_w, f3 rd_w,
                                                     float tmax, f3 &trw, float &dist, Sampler &s, float *tex_stk,
                                                     Ev &&ev = Ev{})
{
    trw = s3(1.0f);
    dist = kFloatMax;
    f3 ro[kMaxBoundaryRecord], rd[kMaxBoundaryRecord];
    f3 sm = s3(0.0f);
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n) continue;
        const DevVolume &v = sc.volumes[vi[i]];
        float ub = 1.0f;
        if (vol_is_het(sc, vi[i])) {
            ro[i] = xf_point(v.w2o, ro_w);
            rd[i] = xf_dir(v.w2o, rd_w);
            ub = vol_bound(sc, v);
        }
        sm = sm + vsig(v) * ub;
    }
    const float csum = sm.x + sm.y + sm.z;
    const f3 cw = csum > 0.0f ? sm / csum : s3(1.0f / 3.0f);
    const float u0 = sampler_next(sc, s);
    const int ch = u0 < cw.x ? 0 : (u0 < cw.x + cw.y ? 1 : 2);
    if (!(comp(sm, ch) > 0.0f)) return false;
    const f3 nf{safe_rcp(sm.x + 1.0f), safe_rcp(sm.y + 1.0f), safe_rcp(sm.z + 1.0f)};
    const f3 smr{safe_rcp(sm.x), safe_rcp(sm.y), safe_rcp(sm.z)};
    f3 tr = s3(1.0f), pdf = s3(1.0f);
    float acc = 0.0f;
    int it = 0;
    for (;;) {
        if (it >= kMaxFreePathLength) break;
        const float fpl = -ym_logf(1.0f - sampler_next(sc, s)) / comp(sm, ch);
        if (acc + fpl > tmax) {
            const float dd = tmax - acc;
            const f3 ptr{ym_expf(-sm.x * dd), ym_expf(-sm.y * dd), ym_expf(-sm.z * dd)};
            tr = tr * (ptr * nf);
            pdf = pdf * (ptr * nf);
            acc += fpl;
            break;
        }
        acc += fpl;
        f3 st = s3(0.0f), ss = s3(0.0f);
#pragma unroll
        for (int i = 0; i < kMaxBoundaryRecord; ++i) {
            if (i >= n) continue;
            const DevVolume &v = sc.volumes[vi[i]];
            float dens = 1.0f;
            if (vol_is_het(sc, vi[i])) {
                dens = vol_density<kTex>(sc, v, ro_w + rd_w * acc, ro[i] + rd[i] * acc, tex_stk);
                ev.note(vi[i], ro[i] + rd[i] * acc);
            }
            st = st + vsig(v) * dens;
            ss = ss + (vsig(v) * valb(v)) * dens;
        }
        f3 rp;
        rp.x = gmin(sm.x > 0.0f ? st.x * smr.x : 0.0f, 1.0f);
        rp.y = gmin(sm.y > 0.0f ? st.y * smr.y : 0.0f, 1.0f);
        rp.z = gmin(sm.z > 0.0f ? st.z * smr.z : 0.0f, 1.0f);
        const f3 fp = s3(1.0f) - rp;
        const f3 ptr{ym_expf(-sm.x * fpl), ym_expf(-sm.y * fpl), ym_expf(-sm.z * fpl)};
        if (sampler_next(sc, s) < comp(rp, ch)) {
            tr = tr * ((ptr * ss) * nf);
            pdf = pdf * (((ptr * sm) * rp) * nf);
            ev.real(st.x, sm.x);
            break;
        }
        ev.null(st.x, sm.x);
        tr = tr * (((ptr * sm) * fp) * nf);
        pdf = pdf * (((ptr * sm) * fp) * nf);
        ++it;
    }
    trw = tr * safe_rcp(dot(cw, pdf));
    dist = acc;
    return (acc > tmax || it >= kMaxFreePathLength) ? false : true;
}

// EvalTr (src/VolumeEval.cuh:253-418): transmittance of one segment inside
// the current media.
// kHet: 0 every medium is homogeneous (only the closed form is compiled),
// 1 voxel-grid media, 2 texture-density media as well (tex_stk: the texture
// graph's LDS stack).
template <int kHet, typename Ev = NoTrackEvents>
__device__ __forceinline__ f3 eval_tr(const DevScene &sc, f3 ro_w, f3 rd_w, const Transfer &t, float tmax,
                                      HashRng &rng, float *tex_stk, Ev &&ev = Ev{})
{
    if (sc.volume_count == 0) return s3(1.0f);
    int vi[kMaxBoundaryRecord];
    const int n = t.volumes(vi, sc.volume_count);
    if (n == 0) return s3(1.0f);
    if constexpr (kHet == 0) {
        f3 st = s3(0.0f);
#pragma unroll
        for (int i = 0; i < kMaxBoundaryRecord; ++i)
            if (i < n) st = st + vsig(sc.volumes[vi[i]]);
        return f3{ym_expf(-st.x * tmax), ym_expf(-st.y * tmax), ym_expf(-st.z * tmax)};
    }
    bool het = false;
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i)
        if (i < n && vol_is_het(sc, vi[i])) het = true;
    if (!het) {
        f3 st = s3(0.0f);
#pragma unroll
        for (int i = 0; i < kMaxBoundaryRecord; ++i)
            if (i < n) st = st + vsig(sc.volumes[vi[i]]);
        return f3{ym_expf(-st.x * tmax), ym_expf(-st.y * tmax), ym_expf(-st.z * tmax)};
    }
    f3 ro[kMaxBoundaryRecord], rd[kMaxBoundaryRecord];
    f3 sm = s3(0.0f);
#pragma unroll
    for (int i = 0; i < kMaxBoundaryRecord; ++i) {
        if (i >= n) continue;
        const DevVolume &v = sc.volumes[vi[i]];
        float ub = 1.0f;
        if (vol_is_het(sc, vi[i])) {
            ro[i] = xf_point(v.w2o, ro_w);
            rd[i] = xf_dir(v.w2o, rd_w);
            ub = vol_bound(sc, v);
        }
        sm = sm + vsig(v) * ub;
    }
    const float csum = sm.x + sm.y + sm.z;
    const f3 cw = csum > 0.0f ? sm / csum : s3(1.0f / 3.0f);
    const float u0 = rng.next();
    const int ch = u0 < cw.x ? 0 : (u0 < cw.x + cw.y ? 1 : 2);
    if (!(comp(sm, ch) > 0.0f)) return s3(1.0f);
    const f3 nf{safe_rcp(sm.x + 1.0f), safe_rcp(sm.y + 1.0f), safe_rcp(sm.z + 1.0f)};
    const f3 smr{safe_rcp(sm.x), safe_rcp(sm.y), safe_rcp(sm.z)};
    f3 tr = s3(1.0f), pdf = s3(1.0f);
    float acc = 0.0f;
    int it = 0;
    for (;;) {
        if (it >= kMaxFreePathLength) break;
        const float fpl = -ym_logf(1.0f - rng.next()) / comp(sm, ch);
        if (acc + fpl > tmax) {
            const float dd = tmax - acc;
            const f3 ptr{ym_expf(-sm.x * dd), ym_expf(-sm.y * dd), ym_expf(-sm.z * dd)};
            tr = tr * (ptr * nf);
            pdf = pdf * (ptr * nf);
            break;
        }
        acc += fpl;
        f3 st = s3(0.0f);
#pragma unroll
        for (int i = 0; i < kMaxBoundaryRecord; ++i) {
            if (i >= n) continue;
            const DevVolume &v = sc.volumes[vi[i]];
            float dens = 1.0f;
            if (vol_is_het(sc, vi[i])) {
                dens = vol_density<kHet == 2>(sc, v, ro_w + rd_w * acc, ro[i] + rd[i] * acc, tex_stk);
                ev.note(vi[i], ro[i] + rd[i] * acc);
            }
            st = st + vsig(v) * dens;
        }
        // ratio tracking's factor 1 - sigma / sigma_m at this step
        ev.null(st.x, sm.x);
        f3 rp;
        rp.x = gmin(sm.x > 0.0f ? st.x * smr.x : 0.0f, 1.0f);
        rp.y = gmin(sm.y > 0.0f ? st.y * smr.y : 0.0f, 1.0f);
        rp.z = gmin(sm.z > 0.0f ? st.z * smr.z : 0.0f, 1.0f);
        const f3 fp = s3(1.0f) - rp;
        const f3 ptr{ym_expf(-sm.x * fpl), ym_expf(-sm.y * fpl), ym_expf(-sm.z * fpl)};
        tr = tr * (((ptr * sm) * fp) * nf);
        pdf = pdf * ((ptr * sm) * nf);
        if (it + 1 > 32) {
            const f3 pt = tr * safe_rcp(dot(cw, pdf));
            const float rr = gmin(gmax(pt.x, gmax(pt.y, pt.z)), 1.0f);
            if (rng.next() < rr) tr = tr / gmax(rr, kMinColorEpsilon);
            else break;
        }
        ++it;
    }
    return tr * safe_rcp(dot(cw, pdf));
}

// ---------------------------------------------------------------- lights
// SampleLightPower (src/Light.h:398-416).  The binary search returns the
// first index whose CDF value exceeds u (or the last light), which on a
// non-decreasing CDF is the number of values <= u: for up to
// kLightCdfInline lights that count is taken over the copy of the CDF in the
// scene record, with no dependent loads.
__device__ __forceinline__ int sample_light_power(const DevScene &sc, float u, float &pdf)
{
    if (sc.light_count <= kLightCdfInline) {
        int idx = 0;
        float c_idx = 0.0f, c_prev = 0.0f;
#pragma unroll
        for (uint32_t i = 0; i < kLightCdfInline; ++i) {
            if (i < sc.light_count && sc.light_cdf_inline[i] <= u) idx = (int)i + 1;
        }
        if (idx > (int)sc.light_count - 1) idx = (int)sc.light_count - 1;
#pragma unroll
        for (uint32_t i = 0; i < kLightCdfInline; ++i) {
            if ((int)i == idx) c_idx = sc.light_cdf_inline[i];
            if ((int)i + 1 == idx) c_prev = sc.light_cdf_inline[i];
        }
        pdf = idx == 0 ? c_idx : c_idx - c_prev;
        return idx;
    }
    int left = 0, right = (int)sc.light_count;
    while (left < right) {
        int mid = (left + right) / 2;
        if (sc.light_cdf[mid] <= u) left = mid + 1;
        else right = mid;
    }
    int idx = left;
    if (idx > (int)sc.light_count - 1) idx = (int)sc.light_count - 1;
    pdf = idx == 0 ? sc.light_cdf[idx] : (sc.light_cdf[idx] - sc.light_cdf[idx - 1]);
    return idx;
}

struct LightTri {
    f3 p0, p1, p2, n0, n1, n2;
    f3 le;
};

__device__ __forceinline__ LightTri load_light(const DevLightRec &l)
{
    LightTri lt;
    lt.p0 = f3{l.p[0], l.p[1], l.p[2]};
    lt.p1 = f3{l.p[3], l.p[4], l.p[5]};
    lt.p2 = f3{l.p[6], l.p[7], l.p[8]};
    lt.n0 = f3{l.n[0], l.n[1], l.n[2]};
    lt.n1 = f3{l.n[3], l.n[4], l.n[5]};
    lt.n2 = f3{l.n[6], l.n[7], l.n[8]};
    lt.le = f3{l.le[0], l.le[1], l.le[2]};
    return lt;
}

YM_HD f3 interp3(f3 a, f3 b, f3 c, f3 w) { return a * w.x + b * w.y + c * w.z; }

// RayIntersectSphere (src/Light.h:42-65): the discriminant is formed in
// double precision as the reference's 4.0 literal makes it.
YM_HD bool ray_sphere_light(f3 o, f3 d, float radius, float &t0, float &t1)
{
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float b = 2.0f * (d.x * o.x + d.y * o.y + d.z * o.z);
    const float c = o.x * o.x + o.y * o.y + o.z * o.z - radius * radius;
    const float disc = (float)((double)(b * b) - 4.0 * (double)a * (double)c);
    if (disc < 0.0f) return false;
    const float sq = sqrtf(disc);
    const float ia = 1.0f / a;
    t0 = (-b - sq) * 0.5f * ia;
    t1 = (-b + sq) * 0.5f * ia;
    if (t1 < 1E-10f) return false;
    return true;
}

// SampleCone (src/SampleUtilities.h:33-39)
YM_HD f3 sample_cone(float u0, float u1, float cos_theta_max)
{
    const float ct = (1.0f - u0) + u0 * cos_theta_max;
    const float st = sqrtf(gmax(0.0f, 1.0f - ct * ct));
    const float phi = u1 * 2.0f * kPi;
    float sp, cp;
    ym_sincosf(phi, sp, cp);
    return f3{cp * st, sp * st, ct};
}


YM_HD f3 affine_translate(const Affine &m) { return f3{m.r[3], m.r[7], m.r[11]}; }

// ShapeLight::SampleLi, sphere branch (src/Light.h:282-317): a direction in
// the cone the sphere subtends, uniform in solid angle.
__device__ __forceinline__ bool sphere_light_sample_li(const DevLightRec &l, f3 pos, float u0, float u1, f3 &wi,
                                                       f3 &lpos, f3 &lerr, f3 &lng, f3 &color, float &pdf)
{
    const f3 opos = xf_position(l.w2o, pos, nullptr);
    const float lpo = length(opos);
    const float radius = l.radius;
    if (lpo < radius) {
        pdf = 0.0f;
        return false;
    }
    const float ctm = sqrtf(gmax(sqr(lpo) - sqr(radius), 0.0f)) / lpo;
    const float cone = 2.0f * kPi * (1.0f - ctm);
    f3 cs = sample_cone(u0, u1, ctm);
    f3 w = -opos, u, v;
    orthogonal_basis(w, u, v);
    cs = cs.x * u + cs.y * v + cs.z * w;
    float t0 = 0.0f, t1 = 0.0f;
    ray_sphere_light(opos, cs, radius, t0, t1);
    f3 ip = opos + cs * t0;
    const float lio = length(ip);
    ip = ip * (radius * safe_rcp(lio));
    f3 ie = error_gamma(5) * abs3(ip);
    lpos = xf_position(l.o2w, ip, &ie);
    lerr = ie;
    lng = normalize(lpos - affine_translate(l.o2w));
    wi = normalize(xf_vector(l.o2w, cs));
    pdf = 1.0f / cone;
    color = f3{l.le[0], l.le[1], l.le[2]} * cone;
    return true;
}

// ShapeLight::EvalLi, sphere branch (src/Light.h:163-199)
__device__ __forceinline__ f3 sphere_light_eval_li(const DevLightRec &l, f3 pos, f3 wi, f3 &lpos, f3 &lerr, f3 &lng,
                                                   float &pdf)
{
    const f3 owi = xf_vector(l.w2o, wi);
    const f3 opos = xf_position(l.w2o, pos, nullptr);
    const float lpo = length(opos);
    const float radius = l.radius;
    float t0 = 0.0f, t1 = 0.0f;
    if (lpo < radius || !ray_sphere_light(opos, owi, radius, t0, t1)) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    const float ctm = sqrtf(gmax(sqr(lpo) - sqr(radius), 0.0f)) / lpo;
    const float cone = 2.0f * kPi * (1.0f - ctm);
    const float ct = dot(normalize(owi), -opos / lpo);
    if (ct < ctm) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    f3 ip = opos + owi * t0;
    const float lio = length(ip);
    ip = ip * (radius * safe_rcp(lio));
    f3 ie = error_gamma(5) * abs3(ip);
    lpos = xf_position(l.o2w, ip, &ie);
    lerr = ie;
    lng = normalize(lpos - affine_translate(l.o2w));
    pdf = 1.0f / cone;
    return f3{l.le[0], l.le[1], l.le[2]};
}

// ShapeLight::SampleLi, mesh branch (src/Light.h:207-281)
__device__ __forceinline__ bool light_sample_li(const DevLightRec &l, f3 pos, float u0, float u1, f3 &wi, f3 &lpos,
                                                f3 &lerr, f3 &lng, f3 &color, float &pdf)
{
    if (l.triangle == kEmpty) return sphere_light_sample_li(l, pos, u0, u1, wi, lpos, lerr, lng, color, pdf);
    const LightTri lt = load_light(l);
    const f3 gn = cross(lt.p1 - lt.p0, lt.p2 - lt.p0);
    const float su = sqrtf(gmax(0.0f, u0));
    const float bu = 1.0f - su, bv = u1 * su;
    const f3 b{bu, bv, 1.0f - bu - bv};
    const float xs = fabsf(lt.p0.x * b.x) + fabsf(lt.p1.x * b.y) + fabsf(lt.p2.x * b.z);
    const float ys = fabsf(lt.p0.y * b.x) + fabsf(lt.p1.y * b.y) + fabsf(lt.p2.y * b.z);
    const float zs = fabsf(lt.p0.z * b.x) + fabsf(lt.p1.z * b.y) + fabsf(lt.p2.z * b.z);
    const f3 sp = interp3(lt.p0, lt.p1, lt.p2, b);
    f3 se = error_gamma(7) * f3{xs, ys, zs};
    const f3 sn = interp3(lt.n0, lt.n1, lt.n2, b);
    const f3 wn = normalize(xf_transpose_dir(l.w2o, sn));
    lpos = xf_position(l.o2w, sp, &se);
    lerr = se;
    lng = normalize(xf_transpose_vector(l.w2o, gn));
    const float dist = length(lpos - pos);
    wi = (lpos - pos) * safe_rcp(dist);
    const float ct = -dot(wn, wi);
    if (ct <= 0.0f) {
        pdf = 0.0f;
        return false;
    }
    pdf = sqr(dist) * safe_rcp(l.area * ct);
    color = lt.le / pdf;
    return true;
}

// RayIntersectTriangle (src/Light.h:14-40) with glm 0.9.8's mat3 inverse
YM_HD bool ray_tri_light(f3 o, f3 d, f3 p0, f3 p1, f3 p2, f3 &bary)
{
    const f3 a = p0 - p2, b = p1 - p2;
    float m[3][3];
    m[0][0] = a.x;
    m[0][1] = a.y;
    m[0][2] = a.z;
    m[1][0] = b.x;
    m[1][1] = b.y;
    m[1][2] = b.z;
    m[2][0] = -d.x;
    m[2][1] = -d.y;
    m[2][2] = -d.z;
    const float ood = 1.0f / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                              m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                              m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
    float iv[3][3];
    iv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * ood;
    iv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * ood;
    iv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * ood;
    iv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * ood;
    iv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * ood;
    iv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * ood;
    iv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * ood;
    iv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * ood;
    iv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * ood;
    const f3 r = o - p2;
    const float u = iv[0][0] * r.x + iv[1][0] * r.y + iv[2][0] * r.z;
    const float v = iv[0][1] * r.x + iv[1][1] * r.y + iv[2][1] * r.z;
    const float t = iv[0][2] * r.x + iv[1][2] * r.y + iv[2][2] * r.z;
    const float w = 1.0f - u - v;
    if (t < kTMin) return false;
    bary = f3{u, v, w};
    return (u > 0.0f && u < 1.0f) && (v > 0.0f && v < 1.0f) && (w > 0.0f && w < 1.0f);
}

// ShapeLight::EvalLi, mesh branch (src/Light.h:87-162)
__device__ __forceinline__ f3 light_eval_li(const DevLightRec &l, f3 pos, f3 wi, f3 &lpos, f3 &lerr, f3 &lng,
                                            float &pdf)
{
    if (l.triangle == kEmpty) return sphere_light_eval_li(l, pos, wi, lpos, lerr, lng, pdf);
    const LightTri lt = load_light(l);
    const f3 gn = cross(lt.p1 - lt.p0, lt.p2 - lt.p0);
    const f3 owi = xf_vector(l.w2o, wi);
    const f3 opos = xf_position(l.w2o, pos, nullptr);
    f3 b = s3(0.0f);
    if (!ray_tri_light(opos, owi, lt.p0, lt.p1, lt.p2, b)) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    const float xs = fabsf(lt.p0.x * b.x) + fabsf(lt.p1.x * b.y) + fabsf(lt.p2.x * b.z);
    const float ys = fabsf(lt.p0.y * b.x) + fabsf(lt.p1.y * b.y) + fabsf(lt.p2.y * b.z);
    const float zs = fabsf(lt.p0.z * b.x) + fabsf(lt.p1.z * b.y) + fabsf(lt.p2.z * b.z);
    const f3 sp = interp3(lt.p0, lt.p1, lt.p2, b);
    f3 se = error_gamma(7) * f3{xs, ys, zs};
    const f3 sn = interp3(lt.n0, lt.n1, lt.n2, b);
    const f3 wn = normalize(xf_transpose_dir(l.w2o, sn));
    const float ct = -dot(wn, wi);
    if (ct <= 0.0f) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    lpos = xf_position(l.o2w, sp, &se);
    lerr = se;
    lng = normalize(xf_transpose_vector(l.w2o, gn));
    const float dist = length(lpos - pos);
    pdf = sqr(dist) * safe_rcp(l.area * ct);
    return lt.le;
}

// DistantLight::SampleLi (src/Light.h:361-385): a direction uniform in the
// cone; the returned radiance is not divided by the pdf (the reference's
// distant light is normalised by its cone, see EvalLi).
YM_HD void distant_sample_li(const DevDistant &dl, float u0, float u1, f3 &wi, f3 &color, float &pdf)
{
    const float cone = 2.0f * kPi * (1.0f - dl.cos_max);
    f3 cs;
    if (cone < 1E-10f) {
        cs = f3{0.0f, 1.0f, 0.0f};
    } else {
        cs = sample_cone(u0, u1, dl.cos_max);
        cs = f3{cs.x, cs.z, cs.y};
    }
    wi = -normalize(xf_vector(dl.o2w, cs));
    pdf = 1.0f / gmax(cone, 1E-10f);
    color = dl.color * dl.intensity;
}

// DistantLight::EvalLi (src/Light.h:344-358): radiance times the cone pdf
YM_HD f3 distant_eval_li(const DevDistant &dl, f3 wi, float &pdf)
{
    const f3 lw = xf_vector(dl.w2o, wi);
    const float ct = -lw.y;
    if (ct < dl.cos_max) {
        pdf = 0.0f;
        return s3(0.0f);
    }
    const float cone = 2.0f * kPi * (1.0f - dl.cos_max);
    pdf = 1.0f / cone;
    return (dl.color * dl.intensity) * pdf;
}

// EvalDistantLight's light pick (src/Shading.cu:109-110)
YM_HD uint32_t pick_distant(uint32_t count, float u)
{
    int k = (int)(u * (float)count);
    k = k > (int)count - 1 ? (int)count - 1 : k;
    return (uint32_t)(k < 0 ? 0 : k);
}

YM_HD float power_heuristic(float main_pdf, float vice_pdf)
{
    return safe_rcp(1.0f + sqr(vice_pdf * safe_rcp(main_pdf)));
}

YM_HD f3 background(f3 d) { return mix3(s3(1.0f), f3{0.05f, 0.2f, 0.85f}, d.y * 0.5f + 0.5f); }

}  // namespace ym

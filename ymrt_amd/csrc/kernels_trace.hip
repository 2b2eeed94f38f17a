// kernels_trace.hip -- ray-query kernels (closest hit, any hit) and the
// sampler probes behind ymrt_trace_* / ymrt_halton_* in include/ymrt.h.
//
// One traversal per lane, 64-lane workgroups (one wave64 each), the traversal
// stack in LDS.  A ray is read as two 16-B loads and its hit written as one
// 16-B store.
#include <algorithm>

#include "dev_sampler.hpp"
#include "dev_traverse.hpp"
#include "internal.hpp"

namespace ym {

constexpr int kTraceBlock = 64;

// Persistent waves over static ranges of the rays: wave w owns rays
// [w * per, (w + 1) * per) and its lanes take the next ray of the range as
// they finish (no atomics: the call owns no work counter), stepping with the
// node-first rule of the path tracer's traversal kernels.
constexpr uint32_t kTraceWaves = 4096;
constexpr int kTraceIdle = 3;

template <bool kAny, bool kCount, bool kSphere, int kStack>
__global__ __launch_bounds__(kTraceBlock) void k_trace(DevScene sc, const float4 *__restrict__ rays, uint32_t n,
                                                        uint32_t per, uint4 *__restrict__ hits,
                                                        uint32_t *__restrict__ occluded, uint2 *__restrict__ visits,
                                                        uint32_t *__restrict__ err)
{
    __shared__ uint32_t stack[kStack * kTraceBlock];
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t cur = blockIdx.x * per;
    const uint32_t end = cur + per < n ? cur + per : n;
    LaneScratch<kTraceBlock> ls{stack, nullptr, nullptr, nullptr, (int)lane, nullptr, 0, 0};
    Traversal<kAny ? kModeSkipBoundary : kModeAll, kAny, false, kCount, kSphere, kStack> tv;
    TraceCounters cnt{0, 0};
    uint32_t i = 0;
    int r = kTraceIdle;
    for (;;) {
        const uint64_t idle = __ballot(r == kTraceIdle);
        if (idle && cur < end) {
            const uint32_t take = min((uint32_t)__popcll(idle), end - cur);
            if (r == kTraceIdle) {
                const uint32_t rank = (uint32_t)__popcll(idle & below);
                if (rank < take) {
                    i = cur + rank;
                    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
                    Query q;
                    q.o = f3{r0.x, r0.y, r0.z};
                    q.tmin = r0.w;
                    q.d = f3{r1.x, r1.y, r1.z};
                    q.tmax = r1.w;
                    q.last_inst = kEmpty;
                    q.last_tri = kEmpty;
                    tv.hit.back = 0;
                    tv.hit.inst = kEmpty;
                    tv.hit.tri = kEmpty;
                    tv.nb_count = 0;
                    cnt = TraceCounters{0, 0};
                    r = tv.init(sc, q);
                }
            }
            cur += take;
        }
        const uint32_t want = cur < end ? 16u : (uint32_t)kTraceBlock;
        for (;;) {
            const uint64_t busy = __ballot(r == kTravContinue);
            if (busy == 0 || (uint32_t)(kTraceBlock - __popcll(busy)) >= want) break;
            wave_step<kAny ? 24 : 16, !kAny>(tv, r, sc, ls, cnt);
        }
        if (r != kTravContinue && r != kTraceIdle) {
            if (r == kTravOverflow) atomicOr(err, 1u);
            const bool hit = r == kTravHit;
            if (kAny) {
                occluded[i] = hit ? 1u : 0u;
            } else {
                hits[i] = hit ? make_uint4(__float_as_uint(tv.hit.t), tv.hit.tri, tv.hit.inst, (uint32_t)tv.hit.back)
                              : make_uint4(__float_as_uint(kTMax), kEmpty, kEmpty, 0u);
            }
            if (kCount) visits[i] = make_uint2(cnt.nodes, cnt.tris);
            r = kTraceIdle;
        }
        if (cur >= end && __ballot(r != kTraceIdle) == 0) break;
    }
}

template <bool kAny, bool kCount, int kStack>
static int launch_one(const ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, uint32_t *occ,
                      uint32_t *visits, hipStream_t s)
{
    const uint32_t waves = std::min<uint32_t>(kTraceWaves, (n + kTraceBlock - 1) / kTraceBlock);
    const uint32_t per = (n + waves - 1) / waves;
    const dim3 grid((n + per - 1) / per);
    if (sc->dev.has_spheres)
        hipLaunchKernelGGL((k_trace<kAny, kCount, true, kStack>), grid, dim3(kTraceBlock), 0, s, sc->dev,
                           reinterpret_cast<const float4 *>(rays), n, per, reinterpret_cast<uint4 *>(hits), occ,
                           reinterpret_cast<uint2 *>(visits), sc->d_error);
    else
        hipLaunchKernelGGL((k_trace<kAny, kCount, false, kStack>), grid, dim3(kTraceBlock), 0, s, sc->dev,
                           reinterpret_cast<const float4 *>(rays), n, per, reinterpret_cast<uint4 *>(hits), occ,
                           reinterpret_cast<uint2 *>(visits), sc->d_error);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

template <bool kAny, bool kCount>
static int launch_stack(const ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, uint32_t *occ,
                        uint32_t *visits, hipStream_t s)
{
    if (sc->max_stack <= 32) return launch_one<kAny, kCount, 32>(sc, rays, n, hits, occ, visits, s);
    if (sc->max_stack <= 48) return launch_one<kAny, kCount, 48>(sc, rays, n, hits, occ, visits, s);
    if (sc->max_stack <= 64) return launch_one<kAny, kCount, 64>(sc, rays, n, hits, occ, visits, s);
    if (sc->max_stack <= 96) return launch_one<kAny, kCount, 96>(sc, rays, n, hits, occ, visits, s);
    set_error("scene BVH too deep for the traversal stack (needs " + std::to_string(sc->max_stack) + " entries)");
    return YMRT_ERR_UNSUPPORTED;
}

int launch_trace(const ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, uint32_t *occluded,
                 uint32_t *visits, int any, hipStream_t stream)
{
    if (n == 0) return YMRT_OK;
    if (any) return launch_stack<true, false>(sc, rays, n, hits, occluded, visits, stream);
    if (visits) return launch_stack<false, true>(sc, rays, n, hits, occluded, visits, stream);
    return launch_stack<false, false>(sc, rays, n, hits, occluded, visits, stream);
}

// ------------------------------------------------------------- sampler probes
__global__ void k_halton_values(DevScene sc, const uint64_t *idx, uint32_t n, uint32_t dim0, uint32_t dims,
                                float *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t index = idx[i];
    for (uint32_t k = 0; k < dims; ++k) {
        const uint32_t d = dim0 + k;
        float v;
        if (d == 0) v = radical_inverse_plain<2>(index);
        else if (d == 1) v = radical_inverse_plain<3>(index);
        else v = radical_inverse(sc, index, d);
        out[(size_t)i * dims + k] = v;
    }
}

int launch_halton_values(const ymrt_scene *sc, const uint64_t *idx, uint32_t n, uint32_t dim0, uint32_t dims,
                         float *out, hipStream_t stream)
{
    YM_REQUIRE(dim0 + dims <= 1000, "Halton dimension out of range (max 1000)");
    if (n == 0 || dims == 0) return YMRT_OK;
    hipLaunchKernelGGL(k_halton_values, dim3((n + 255) / 256), dim3(256), 0, stream, sc->dev, idx, n, dim0, dims,
                       out);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

__global__ void k_halton_enumerate(HaltonEnum e, const uint32_t *px, const uint32_t *py, const uint32_t *s,
                                   uint32_t n, uint64_t *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = e.index(px[i], py[i], s[i]);
}

int launch_halton_enumerate(uint32_t w, uint32_t h, const uint32_t *px, const uint32_t *py, const uint32_t *s,
                            uint32_t n, uint64_t *out, hipStream_t stream)
{
    YM_REQUIRE(w > 0 && h > 0, "image size must be positive");
    if (n == 0) return YMRT_OK;
    HaltonEnum e = HaltonEnum::make((int)w, (int)h);
    hipLaunchKernelGGL(k_halton_enumerate, dim3((n + 255) / 256), dim3(256), 0, stream, e, px, py, s, n, out);
    YM_HIP(hipGetLastError());
    return YMRT_OK;
}

}  // namespace ym

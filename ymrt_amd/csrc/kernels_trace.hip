// kernels_trace.hip -- ray-query kernels (closest hit, any hit) and the
// sampler probes behind ymrt_trace_* / ymrt_halton_* in include/ymrt.h.
//
// One ray per lane, 64-lane workgroups (one wave64 each), the traversal stack
// in LDS.  Rays are read as two 16-B loads and hits written as one 16-B store,
// so a wave moves 2 KB of rays and 1 KB of hits in fully coalesced lines.
#include "dev_sampler.hpp"
#include "dev_traverse.hpp"
#include "internal.hpp"

namespace ym {

constexpr int kTraceBlock = 64;

template <bool kAny, bool kCount, bool kSphere, int kStack>
__global__ __launch_bounds__(kTraceBlock) void k_trace(DevScene sc, const float4 *__restrict__ rays, uint32_t n,
                                                        uint4 *__restrict__ hits, uint32_t *__restrict__ occluded,
                                                        uint2 *__restrict__ visits, uint32_t *__restrict__ err)
{
    __shared__ uint32_t stack[kStack * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    Query q;
    q.o = f3{r0.x, r0.y, r0.z};
    q.tmin = r0.w;
    q.d = f3{r1.x, r1.y, r1.z};
    q.tmax = r1.w;
    q.last_inst = kEmpty;
    q.last_tri = kEmpty;
    LaneScratch<kTraceBlock> ls{stack, nullptr, nullptr, nullptr, (int)threadIdx.x, nullptr, 0, 0};
    Hit h;
    h.back = 0;
    int nbc = 0;
    TraceCounters cnt{0, 0};
    int r;
    if (kAny)
        r = traverse<kModeSkipBoundary, true, false, kCount, kSphere, kStack>(sc, q, h, nbc, ls, cnt);
    else
        r = traverse<kModeAll, false, false, kCount, kSphere, kStack>(sc, q, h, nbc, ls, cnt);
    if (r < 0) {
        atomicOr(err, 1u);
        r = 0;
    }
    if (kAny) {
        occluded[i] = (uint32_t)r;
    } else {
        hits[i] = r ? make_uint4(__float_as_uint(h.t), h.tri, h.inst, (uint32_t)h.back)
                    : make_uint4(__float_as_uint(kTMax), kEmpty, kEmpty, 0u);
    }
    if (kCount) visits[i] = make_uint2(cnt.nodes, cnt.tris);
}

template <bool kAny, bool kCount, int kStack>
static int launch_one(const ymrt_scene *sc, const ymrt_ray *rays, uint32_t n, ymrt_hit *hits, uint32_t *occ,
                      uint32_t *visits, hipStream_t s)
{
    dim3 grid((n + kTraceBlock - 1) / kTraceBlock);
    if (sc->dev.has_spheres)
        hipLaunchKernelGGL((k_trace<kAny, kCount, true, kStack>), grid, dim3(kTraceBlock), 0, s, sc->dev,
                           reinterpret_cast<const float4 *>(rays), n, reinterpret_cast<uint4 *>(hits), occ,
                           reinterpret_cast<uint2 *>(visits), sc->d_error);
    else
    hipLaunchKernelGGL((k_trace<kAny, kCount, false, kStack>), grid, dim3(kTraceBlock), 0, s, sc->dev,
                       reinterpret_cast<const float4 *>(rays), n, reinterpret_cast<uint4 *>(hits), occ,
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
